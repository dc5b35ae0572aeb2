"""The identities behind the negacyclic transforms whose twist is folded into
the first forward pass and whose untwist is folded into the last inverse pass
(cuyashe_amd/csrc/ntt16.hpp, MERGED; tables in ring_params.hpp), checked with
Python integers.

Thread g of a length-N transform holds elements j = g + c N/16 (c < 16) in
those passes. psi = 7^((P-1)/(2N)) and psi^(N/16) = 7^((P-1)/32) = 2^78, so
psi^j = psi^g 2^(78 c): the slot factor is a shift and psi^g is common to the
thread's inputs. A 16-point DFT commutes with a common factor, so the twisted
transform equals the DFT of the slot-shifted inputs with psi^g applied to
every output. The GPU parity tests check the kernels end to end; these tests
pin the algebra on its own, including a planted wrong slot exponent."""
import random

P = (1 << 64) - (1 << 32) + 1


def brev4(c):
    return ((c & 1) << 3) | ((c & 2) << 1) | ((c & 4) >> 1) | ((c & 8) >> 3)


def dft16(x, w16):
    return [sum(x[k] * pow(w16, k * m, P) for k in range(16)) % P for m in range(16)]


def test_psi_slot_factor_is_2_pow_78():
    assert pow(7, (P - 1) // 32, P) == pow(2, 78, P)
    for logn in range(4, 15):
        n = 1 << logn
        psi = pow(7, (P - 1) // (2 * n), P)
        assert pow(psi, n // 16, P) == pow(2, 78, P)
        g, c = random.Random(logn).randrange(n // 16), random.Random(logn + 1).randrange(16)
        assert pow(psi, g + c * (n // 16), P) == pow(psi, g, P) * pow(2, 78 * c, P) % P


def _first_pass(x, n, g, merged, slot_exp=78):
    """Forward first pass of thread g: twisted inputs -> DFT16 -> inter-pass
    twiddles (the unmerged form), or slot-shifted inputs -> DFT16 -> merged
    twiddles w_N^(brev4(c) g) psi^g (output slot c is DFT index brev4(c))."""
    s = n // 16
    w = pow(7, (P - 1) // n, P)
    psi = pow(7, (P - 1) // (2 * n), P)
    w16 = pow(w, s, P)
    if merged:
        xin = [x[c] * pow(2, slot_exp * c, P) % P for c in range(16)]
    else:
        xin = [x[c] * pow(psi, g + c * s, P) % P for c in range(16)]
    X = dft16(xin, w16)
    out = []
    for c in range(16):
        v = X[brev4(c)]
        tw = pow(w, brev4(c) * g, P)
        out.append(v * tw * (pow(psi, g, P) if merged else 1) % P)
    return out


def test_merged_first_pass_equals_twisted_pass():
    rng = random.Random(7)
    for n in (16, 32, 64, 4096):
        for g in sorted({0, min(1, n // 16 - 1), n // 16 - 1, rng.randrange(n // 16)}):
            x = [rng.randrange(P) for _ in range(16)]
            assert _first_pass(x, n, g, merged=True) == _first_pass(x, n, g, merged=False)


def test_planted_wrong_slot_exponent_fails():
    rng = random.Random(8)
    x = [rng.randrange(P) for _ in range(16)]
    assert _first_pass(x, 4096, 5, merged=True, slot_exp=79) != _first_pass(x, 4096, 5, merged=False)


def test_merged_last_pass_equals_untwist():
    """Inverse last pass: inputs times w_N^(-brev4(c) g) psi^-g N^-1, inverse
    DFT16, output slot c times 2^(-78 c) == the unmerged pass followed by
    psi^-j N^-1 on element j = g + c N/16."""
    rng = random.Random(9)
    for n in (16, 32, 4096):
        s = n // 16
        w = pow(7, (P - 1) // n, P)
        wi = pow(w, P - 2, P)
        psi = pow(7, (P - 1) // (2 * n), P)
        psii = pow(psi, P - 2, P)
        ninv = pow(n, P - 2, P)
        wi16 = pow(wi, s, P)
        for g in sorted({0, s - 1, rng.randrange(s)}):
            y = [rng.randrange(P) for _ in range(16)]
            # the inverse pass reads slot c as DIT input brev4(c)-ordered data:
            # twiddle, then inverse DFT over the bit-reversed slots
            t_old = [y[c] * pow(wi, brev4(c) * g, P) % P for c in range(16)]
            z_old = [sum(t_old[c] * pow(wi16, brev4(c) * m, P) for c in range(16)) % P for m in range(16)]
            want = [z_old[m] * pow(psii, g + m * s, P) * ninv % P for m in range(16)]
            t_new = [y[c] * pow(wi, brev4(c) * g, P) * pow(psii, g, P) * ninv % P for c in range(16)]
            z_new = [sum(t_new[c] * pow(wi16, brev4(c) * m, P) for c in range(16)) % P for m in range(16)]
            got = [z_new[m] * pow(2, (192 - (78 * m) % 192) % 192, P) % P for m in range(16)]
            assert got == want
