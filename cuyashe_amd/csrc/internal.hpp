// Access to the library's object internals for host translation units other
// than yashe_hip.hip (which defines the objects).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/yashe_hip.h"

namespace yhi {

// Records msg as the calling thread's yh_last_error() and returns code.
int fail(int code, const std::string& msg);
// HIP device of a context, and its stream.
int ctx_device(const yh_ctx* ctx);
hipStream_t ctx_stream(const yh_ctx* ctx);
// The stream a ring's operations run on.
hipStream_t ring_stream(const yh_ring* ring);
// q of a ring (limbs, least significant first) and its limb count.
const uint64_t* ring_q(const yh_ring* ring, int* limbs);
// The context a ring or key belongs to.
yh_ctx* ring_ctx(const yh_ring* ring);
const yh_ring* evk_ring(const yh_evk* evk);
// References held by a request loop: while any is held the ring or key
// refuses destruction, and the context counts the loop as live.
void hold_ring(yh_ring* ring, int delta);
void hold_evk(const yh_evk* evk, int delta);
void hold_ctx(yh_ctx* ctx, int delta);

}  // namespace yhi
