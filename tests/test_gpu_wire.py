"""Keys and ciphertexts leaving one context and entering another through the
wire format: a key file written by one process-like owner (its own context,
ring and keys) and read by another gives the same public key, the same
resident keyswitch key (bit-identical mul+keyswitch results) and the same
decryptions."""
import numpy as np
import pytest

import cuyashe_amd as ya
from cuyashe_amd.scheme import Yashe, negacyclic_mul_mod

pytestmark = pytest.mark.gpu

Q127 = (1 << 127) - 1


@pytest.mark.parametrize("n,q,t", [(64, Q127, 1024), (4096, Q127, 1024)])
def test_key_file_and_ciphertexts_cross_contexts(gpu_ctx, tmp_path, n, q, t):
    owner = Yashe(gpu_ctx, n, q, t, 32, seed=5).generate_keys()
    path = str(tmp_path / "keys.yhw")
    owner.save_keys(path, secret=True)
    rng = np.random.default_rng(6)
    m1 = rng.integers(0, t, size=(2, n))
    m2 = rng.integers(0, t, size=(2, n))
    c1, c2 = owner.encrypt(m1), owner.encrypt(m2)
    blob1, blob2 = owner.export_ciphertexts(c1, 2), owner.export_ciphertexts(c2, 2)
    want = gpu_ctx.alloc(2 * owner.elt_bytes)
    owner.mul_keyswitch(c1, c2, want, 2)
    want_host = want.download((2, owner.L, n))

    ctx2 = ya.Context(0)
    try:
        other = Yashe(ctx2, n, q, t, 32, seed=99).load_keys(path)
        assert np.array_equal(other.h.download((1, other.L, n)), owner.h.download((1, owner.L, n)))
        d1, d2 = other.import_ciphertexts(blob1), other.import_ciphertexts(blob2)
        got = ctx2.alloc(2 * other.elt_bytes)
        other.mul_keyswitch(d1, d2, got, 2)
        assert np.array_equal(got.download((2, other.L, n)), want_host)
        dec = other.decrypt(got, 2)
        for i in range(2):
            assert np.array_equal(dec[i], negacyclic_mul_mod(m1[i], m2[i], t))
        # a ciphertext encrypted under the loaded public key decrypts with
        # the owner's secret key
        c3 = other.encrypt(m1[:1])
        back = owner.import_ciphertexts(other.export_ciphertexts(c3, 1))
        assert np.array_equal(owner.decrypt(back, 1)[0], m1[0])
        # a blob of another ring is refused before anything is uploaded
        ring2 = ya.Ring(ctx2, n, q, t=t + 1, w_bits=32)
        with pytest.raises(ya.YashError, match="another ring"):
            ring2.import_(ya.WIRE_CIPHERTEXT, blob1)
        ring2.close()
        other.ring.close()
    finally:
        ctx2.close()
    owner.ring.close()
