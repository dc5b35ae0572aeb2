"""Writes tests/golden/reference_kats.json.

Every vector is either data held by the reference's own tests (the expected
arrays of src/test_mul.cu, the constants of src/test_bnt.cu and
src/test_reduce.cu) or a closed form computed here with Python integers
(schoolbook products, explicit folds, exact divisions). The operand streams of
the reference tests come from glibc ``srand(36251); rand()`` exactly as the
reference draws them (Polynomial::random, src/polynomial.h:1369-1381). Nothing
from the reference is imported or executed.

Run: python tests/golden/gen_golden.py
"""
import ctypes
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "reference_kats.json")
P = (1 << 64) - (1 << 32) + 1


class GlibcRand:
    def __init__(self, seed):
        self.libc = ctypes.CDLL("libc.so.6")
        self.libc.rand.restype = ctypes.c_int
        self.libc.srand(seed)

    def __call__(self):
        return int(self.libc.rand())


def schoolbook(a, b):
    out = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def cyclic(a, b, N):
    out = [0] * N
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[(i + j) % N] += x * y
    return out


def negacyclic_mod(a, b, n, q):
    out = [0] * n
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            k = i + j
            if k >= n:
                out[k - n] -= x * y
            else:
                out[k] += x * y
    return [v % q for v in out]


def fold_mod(c, n, q):
    c = list(c) + [0] * (2 * n - len(c))
    return [(c[i] - c[i + n]) % q for i in range(n)]


def dft(x, inverse=False):
    N = len(x)
    w = pow(7, (P - 1) // N, P)
    if inverse:
        w = pow(w, P - 2, P)
    out = [sum(x[j] * pow(w, j * k, P) for j in range(N)) % P for k in range(N)]
    if inverse:
        ninv = pow(N, P - 2, P)
        out = [v * ninv % P for v in out]
    return out


def main():
    kats = {}

    # ---- src/test_mul.cu first test (:17-50): q, degree 8, srand(36251)
    q1 = 1171313591017775093490277364417
    r = GlibcRand(36251)
    a = [r() % q1 for _ in range(9)]   # Polynomial::random(&a, 8)
    b = [r() % q1 for _ in range(9)]   # Polynomial::random(&b, 8)
    expected1 = [52549502886738203, 729378972269732879, 3315814916818960945, 5373031442410099707,
                 5487719755647497370, 6002790757075884044, 7641866224560971944, 8444884424660384993,
                 10787872446794447944, 10671875472706824289, 7566104911804466568, 5469026184047188285,
                 4332423149998847987, 2715876362046054824, 2047760673349179087, 1502584925726431011]
    assert cyclic(a, b, 16) == expected1, "KAT1 is the product mod x^16 - 1"
    kats["test_mul_1"] = {"source": "src/test_mul.cu:17-50", "q": str(q1), "n": 8, "N": 16,
                          "a": [str(v) for v in a], "b": [str(v) for v in b],
                          "expected": [str(v) for v in expected1]}

    # ---- src/test_mul.cu second test (:52-89): literal operands, exact Z[x] product
    a2 = [1304768405, 1912295543, 584306160, 476779113, 2057832156, 1274012458, 1818971124, 1943566066]
    b2 = [13860721405712617825882419437, 911979968073012785910226533, 60134818568832314735148266017,
          30228208549061091796489949751, 42496336825615196107672318132, 2538357487121433195951809389,
          19401757472799351763805908821, 68415359680988770693366249506]
    expected2 = [18085031360681010249051172126355487985, 27695718415243509623513993478082759156,
                 88304891436672768443904444999435867224, 161577734953354918933760345835324042047,
                 177348142234080397165460901588496065374, 150446511525560453344956466169095450596,
                 219533428887126844535620298106253625669, 315527818352347626764175594181751346765,
                 380494624177053395698844189209855453911, 280450476250800780083801312965924383119,
                 211828603909665970332138017649830954350, 252717044998320508082871060193586798902,
                 127386722625350085506673181613116824226, 162154161142687163774421077602378932930,
                 132969771469154360046681933849550863396]
    assert schoolbook(a2, b2) == expected2, "KAT2 is the exact product"
    kats["test_mul_2"] = {"source": "src/test_mul.cu:52-89", "q": str(q1), "n": 8, "N": 16,
                          "a": [str(v) for v in a2], "b": [str(v) for v in b2],
                          "expected": [str(v) for v in expected2]}

    # ---- src/test_bnt.cu:128-190: CRT -> ICRT of X under a 512-bit q, degree 1
    X = 5342748479728798479824982
    q_bnt = int("11728242269442118225223215137823427308012851928365899315927446782605336945293708352036740673805944948355911992121288532925023234127632779222593494093664877")
    kats["test_bnt_crt"] = {"source": "src/test_bnt.cu:128-190", "X": str(X), "q": str(q_bnt), "n": 1}

    # ---- src/test_reduce.cu:139-163: one coefficient mod q
    big = 100841831415836174481434746063997936280
    qr = 77287149995192912462927307869
    kats["test_reduce_modn"] = {"source": "src/test_reduce.cu:139-163", "value": str(big), "q": str(qr),
                                "expected": str(big % qr)}
    # ---- src/test_reduce.cu:52-135: fold mod x^8 + 1 of two literal vectors (q = 655615111)
    qs = 655615111
    v1 = [1] * 9
    v2 = [722, 173, 735, 651, 460, 161, 56, 80, 762, 9, 220, 281, 62, 96, 239]
    kats["test_reduce_fold"] = {"source": "src/test_reduce.cu:52-135", "q": str(qs), "n": 8,
                                "inputs": [v1, v2], "expected": [fold_mod(v1, 8, qs), fold_mod(v2, 8, qs)]}

    # ---- src/test.cu PolySuite (:31-70): n = 8, q = 77287149995192912462927307869,
    # a fresh srand(36251) per test case.
    qp = 77287149995192912462927307869
    n = 8
    r = GlibcRand(36251)
    cases = []
    for _ in range(10):  # multiplyByPolynomial (:472-524)
        a = [r() % qp for _ in range(n)]
        b = [r() % qp for _ in range(n)]
        cases.append({"a": [str(v) for v in a], "b": [str(v) for v in b],
                      "expected": [str(v) for v in negacyclic_mod(a, b, n, qp)]})
    kats["test_multiplyByPolynomial"] = {"source": "src/test.cu:472-524", "q": str(qp), "n": n, "cases": cases}

    r = GlibcRand(36251)  # severalMultiplications (:823-900)
    a = [r() % qp for _ in range(n)]
    b = [r() % qp for _ in range(n)]
    chain = [negacyclic_mod(a, b, n, qp)]
    for _ in range(1, 10):
        chain.append(negacyclic_mod(chain[-1], a, n, qp))
    kats["test_severalMultiplications"] = {"source": "src/test.cu:823-900", "q": str(qp), "n": n,
                                           "a": [str(v) for v in a], "b": [str(v) for v in b],
                                           "chain": [[str(v) for v in c] for c in chain]}

    # simpleMultiplyByPolynomial (:421-470): a_i = i, b_i = 1, exact product
    a = list(range(n))
    b = [1] * n
    kats["test_simpleMultiplyByPolynomial"] = {"source": "src/test.cu:421-470", "q": str(qp), "n": n,
                                               "a": a, "b": b, "expected": schoolbook(a, b)}

    r = GlibcRand(36251)  # phiReduceCPU (:901-933): degree 2n-2 inputs, fold, mod q
    red = []
    for _ in range(10):
        c = [r() % qp for _ in range(2 * n - 1)]
        red.append({"c": [str(v) for v in c], "expected": [str(v) for v in fold_mod(c, n, qp)]})
    kats["test_phiReduce"] = {"source": "src/test.cu:901-969", "q": str(qp), "n": n, "cases": red}

    # ---- NTT definition (ntt_test.py:7-34, src/test.cu:305-329): brute-force DFT
    ntt = []
    for N in (16, 64):
        x = [(k * 0x9E3779B97F4A7C15 + 12345) % P for k in range(N)]
        ntt.append({"N": N, "x": [str(v) for v in x], "X": [str(v) for v in dft(x)],
                    "wN": str(pow(7, (P - 1) // N, P))})
    kats["ntt_dft"] = {"source": "ntt_test.py:7-34, src/cuda_functions.cu:878-918", "cases": ntt}

    # ---- s_mul host branch (src/test.cu:330-346): uint128 % P, plus edge values
    edge = [0, 1, 2, P - 1, P - 2, 1 << 32, (1 << 32) - 1, (1 << 64) - (1 << 32), (1 << 63), P >> 1]
    muls = [[str(x), str(y), str(x * y % P)] for x in edge for y in edge]
    kats["s_mul"] = {"source": "src/test.cu:330-346, src/cuda_functions.cu:351", "cases": muls}

    with open(OUT, "w") as f:
        json.dump(kats, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
