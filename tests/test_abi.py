"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its host-only parameter planning follows the exactness rules."""
import ctypes
import os
import re

import pytest

import cuyashe_amd as ya
from cuyashe_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "yashe_hip.h")
P = (1 << 64) - (1 << 32) + 1


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(yh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(_lib.exported_symbols())


def test_no_gpu_fails_loudly():
    # Without a device the context cannot be created; there is no CPU fallback.
    try:
        import subprocess
        has_gpu = subprocess.run(["rocminfo"], capture_output=True, timeout=20).returncode == 0 and \
            b"gfx9" in subprocess.run(["rocminfo"], capture_output=True, timeout=20).stdout
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(ya.YashError):
        ya.Context(0)


@pytest.mark.parametrize("n,q,t,wb", [
    (4096, (1 << 127) - 1, 1024, 32),
    (8192, (1 << 255) - 19, 1 << 16, 32),
    (1024, (1 << 127) - 1, 0, 0),
    (16384, (1 << 511) - 187, 1024, 32),
    (8, 77287149995192912462927307869, 35951, 32),
    (2, 1171313591017775093490277364417, 0, 0),
    (16, (1 << 511) - 187, 1024, 32),
])
def test_ring_plan_bounds(n, q, t, wb):
    info, primes, ks_primes = ya.ring_plan(n, q, t, wb, keyswitch=True)
    assert info.R == len(primes) and info.n == n
    # modulus 0 is the Goldilocks prime itself (its convolution is exact mod P
    # for any operands); the small primes follow in descending order, sized
    # for one product of centred residues per convolution in the product set,
    # and in the keyswitch set for l_w products of a centred but unreduced
    # digit (transformed once for every modulus) by a centred key residue
    def exact_prod(p):
        return n * ((p - 1) // 2) ** 2 <= (P - 1) // 2

    def exact_ks(p):
        # centred digits: |d - 2^(wb-1)| <= 2^(wb-1)
        return max(1, info.lw) * n * (1 << (wb - 1)) * ((p - 1) // 2) <= (P - 1) // 2

    def is_prime(p):
        return p > 1 and all(p % d for d in range(2, int(p ** 0.5) + 1))

    for mods, exact in ((primes, exact_prod), (ks_primes, exact_ks)):
        if not mods:
            continue
        assert mods[0] == P
        small = mods[1:]
        assert small == sorted(small, reverse=True) and len(set(small)) == len(small)
        for p in small:
            assert is_prime(p)
            assert exact(p)
            assert p < (1 << 28)
        # the reconstruction epilogues sum y_i * (32-bit word) over the small
        # primes in one 64-bit column
        assert sum(small) < (1 << 32)
        if small and small[0] < (1 << 28) - 64 and sum(small) < (1 << 31):
            # the first small prime is the largest the bound allows
            nxt = small[0] + 2
            while not is_prime(nxt):
                nxt += 2
            assert not exact(nxt) or nxt >= (1 << 28)
    M = 1
    for p in primes:
        M *= p
    assert M >= 4 * n * (q - 1) ** 2
    assert M.bit_length() == info.m_bits
    assert M.bit_length() + 2 <= 64 * info.work_limbs
    M_short = M // primes[-1]
    assert M_short < 4 * n * (q - 1) ** 2  # no more primes than needed
    if wb:
        lw = (q.bit_length() + wb - 1) // wb
        assert info.lw == lw and len(ks_primes) == info.R_ks
        Mk = 1
        for p in ks_primes:
            Mk *= p
        assert Mk >= 4 * lw * n * ((1 << wb) - 1) * (q - 1)
        assert Mk.bit_length() == info.m_ks_bits
        if len(ks_primes) > 1:
            assert Mk // ks_primes[-1] < 4 * lw * n * ((1 << wb) - 1) * (q - 1)


def test_ring_plan_rejects_bad_parameters():
    with pytest.raises(ya.YashError):
        ya.ring_plan(3, (1 << 127) - 1)
    with pytest.raises(ya.YashError):
        ya.ring_plan(32768, (1 << 127) - 1)
    with pytest.raises(ya.YashError):
        ya.ring_plan(64, 1)
    with pytest.raises(ya.YashError):
        ya.ring_plan(64, (1 << 600) + 1)


def test_plain_length_rule_matches_oracle():
    from oracle import cpuref
    L = _lib.lib()
    for da in range(0, 40):
        for db in range(0, 40):
            for mn in (1, 16):
                assert L.yh_poly_mul_plain_length(da, db, mn) == cpuref.ntt_length(da, db, mn), (da, db, mn)
