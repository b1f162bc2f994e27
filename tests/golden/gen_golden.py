"""Writes tests/golden/golden.json: small known-answer vectors computed with
exact rational arithmetic (fractions.Fraction), independently of the C oracle.

The reference ships no tests, fixtures or golden vectors (its only check is
the unasserted residual of src/util.rs:55-85), so these vectors are derived
from the arithmetic the reference's code fixes:
  - "ukernel": the 6x8 micro-kernel, an fma chain in p order from zero, then
    alpha*acc, then fma(beta, c, alpha*acc), beta == 0 not reading C
    (ukernels/bli_gemm_haswell_asm_d6x8.c:691, :725-888, :899-911, :939-1193);
  - "goto": the Goto tree's rounding order, one rounding of C per Kc = 192
    block with beta := 1 after the first block (src/composables/part.rs:96-110),
    fused update on full 6x8 tiles and c = tau + beta*c on edge tiles
    (src/kern/kernel_nm.rs:75-95), Mc = 120, Nc = 3000 (l3_packing.rs:51-54);
  - "triple": TripleLoop's x,z,y order with c = a*b + c rounded twice
    (src/composables/triple_loop.rs:8-17);
  - "exact": correctly rounded beta*C0 + alpha*A*B (for the double-double oracle);
  - "ozaki": the library's split-BF16 algorithm (no reference counterpart);
  - "crt": the library's Chinese-remainder split-BF16 algorithm: exact integer
    images of the exponent-aligned rows/columns, their exact product (Python
    integers), one correctly rounded conversion, scaling, alpha, fma with beta.
Every double is stored with float.hex so the vectors are bit exact.

Run: python tests/golden/gen_golden.py
"""
from __future__ import annotations

import json
import math
from fractions import Fraction as F
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "golden.json"


def rn(x: F) -> float:
    """Round a rational to the nearest double (ties to even): int/int true
    division is correctly rounded in CPython."""
    return x.numerator / x.denominator if x.denominator != 1 else float(x.numerator)


def fma(a: float, b: float, c: float) -> float:
    return rn(F(a) * F(b) + F(c))


def mul(a: float, b: float) -> float:
    return rn(F(a) * F(b))


def add(a: float, b: float) -> float:
    return rn(F(a) + F(b))


def hexs(a) -> list:
    return [float(x).hex() for x in np.asarray(a, dtype=np.float64).ravel(order="F")]


def rand_mat(rng, h, w, spread=6):
    """Values with varied exponents so that rounding order matters."""
    return rng.uniform(-1, 1, (h, w)) * np.exp2(rng.integers(-spread, spread + 1, (h, w)))


def ukernel_case(rng, k, alpha, beta):
    a = rand_mat(rng, 6, k)            # a[i][p]
    b = rand_mat(rng, k, 8)            # b[p][j]
    c = rand_mat(rng, 6, 8)
    out = np.zeros((6, 8))
    for i in range(6):
        for j in range(8):
            acc = 0.0
            for p in range(k):
                acc = fma(a[i, p], b[p, j], acc)
            t = mul(alpha, acc)
            out[i, j] = t if beta == 0.0 else fma(beta, c[i, j], t)
    return {"k": k, "alpha": float(alpha).hex(), "beta": float(beta).hex(),
            "a": hexs(a), "b": hexs(b), "c": hexs(c), "out": hexs(out)}


KC, MC, NC, MR, NR = 192, 120, 3000, 6, 8


def goto_entry(A, B, c0, y, x, m, n, k, alpha, beta):
    n0 = (x // NC) * NC
    nc = min(NC, n - n0)
    jr = ((x - n0) // NR) * NR
    m0 = (y // MC) * MC
    mc = min(MC, m - m0)
    ir = ((y - m0) // MR) * MR
    full = (nc - jr >= NR) and (mc - ir >= MR)
    c = c0
    bb = beta
    for k0 in range(0, k, KC):
        acc = 0.0
        for p in range(k0, min(k, k0 + KC)):
            acc = fma(A[y, p], B[p, x], acc)
        s = mul(alpha, acc)
        if full:
            c = s if bb == 0.0 else fma(bb, c, s)
        else:
            c = add(s, mul(bb, c))
        bb = 1.0
    return c


def goto_case(rng, m, n, k, alpha, beta):
    A, B, C = rand_mat(rng, m, k), rand_mat(rng, k, n), rand_mat(rng, m, n)
    out = np.array([[goto_entry(A, B, C[y, x], y, x, m, n, k, alpha, beta) for x in range(n)]
                    for y in range(m)])
    return {"m": m, "n": n, "k": k, "alpha": float(alpha).hex(), "beta": float(beta).hex(),
            "A": hexs(A), "B": hexs(B), "C": hexs(C), "out": hexs(out)}


def triple_case(rng, m, n, k):
    A, B, C = rand_mat(rng, m, k), rand_mat(rng, k, n), rand_mat(rng, m, n)
    out = C.copy()
    for x in range(n):
        for z in range(k):
            for y in range(m):
                out[y, x] = add(mul(A[y, z], B[z, x]), out[y, x])
    return {"m": m, "n": n, "k": k, "A": hexs(A), "B": hexs(B), "C": hexs(C), "out": hexs(out)}


def exact_case(rng, m, n, k, alpha, beta, spread):
    A, B, C = rand_mat(rng, m, k, spread), rand_mat(rng, k, n, spread), rand_mat(rng, m, n, spread)
    out = np.zeros((m, n))
    for y in range(m):
        for x in range(n):
            s = sum(F(A[y, p]) * F(B[p, x]) for p in range(k))
            out[y, x] = rn(F(alpha) * s + F(beta) * F(C[y, x]))
    return {"m": m, "n": n, "k": k, "alpha": float(alpha).hex(), "beta": float(beta).hex(),
            "A": hexs(A), "B": hexs(B), "C": hexs(C), "out": hexs(out)}


def ozaki_chunk(s):
    return min(256, 16 * (1024 // (16 * (s + 2))))


def ozaki_split(v, s):
    mu = max(abs(x) for x in v)
    e = math.frexp(mu)[1] if mu > 0 else 0
    qs = []
    for x in v:
        r = F(x) / F(2) ** e
        q = []
        for t in range(1, s + 1):
            scaled = r * F(2) ** (8 * t)
            qt = round(scaled)  # Python round(Fraction) is round-half-even
            q.append(qt)
            r -= F(qt) / F(2) ** (8 * t)
        qs.append(q)
    return qs, e


def ozaki_case(rng, m, n, k, s, alpha, beta, spread):
    A, B, C = rand_mat(rng, m, k, spread), rand_mat(rng, k, n, spread), rand_mat(rng, m, n, spread)
    rows = [ozaki_split(list(A[i, :]), s) for i in range(m)]
    cols = [ozaki_split(list(B[:, j]), s) for j in range(n)]
    kch = ozaki_chunk(s)
    out = np.zeros((m, n))
    for i in range(m):
        qa, ea = rows[i]
        for j in range(n):
            qb, eb = cols[j]
            acc = 0.0
            for c0 in range(0, k, kch):
                for d in range(s + 1, 1, -1):
                    S = 0
                    for t in range(1, d):
                        u = d - t
                        S += sum(qa[p][t - 1] * qb[p][u - 1] for p in range(c0, min(k, c0 + kch)))
                    assert abs(S) < 2 ** 24
                    acc = rn(F(acc) + F(S) / F(2) ** (8 * d))
            v = rn(F(alpha) * F(acc))
            v = math.ldexp(v, ea + eb)
            out[i, j] = v if beta == 0.0 else fma(beta, C[i, j], v)
    return {"m": m, "n": n, "k": k, "s": s, "alpha": float(alpha).hex(),
            "beta": float(beta).hex(), "A": hexs(A), "B": hexs(B), "C": hexs(C),
            "out": hexs(out)}


CRT_MODULI = [511, 509, 507, 505, 503, 499, 493, 491, 487, 479, 473, 467, 463, 461, 457, 449]


def crt_beta(k, N):
    P = math.prod(CRT_MODULI[:N])
    kb = (k - 1).bit_length() if k > 1 else 0
    room = P.bit_length() - 2 - kb
    b = room // 2 if room >= 0 else -1
    return min(b, (125 - kb) // 2, 53)


def crt_integers(v, beta):
    mu = max((abs(x) for x in v), default=0.0)
    e = math.frexp(mu)[1] if mu > 0 else 0
    return [round(F(x) * F(2) ** (beta - e)) for x in v], e


def crt_case(rng, m, n, k, N, alpha, beta, spread, zero_row=False):
    A, B, C = rand_mat(rng, m, k, spread), rand_mat(rng, k, n, spread), rand_mat(rng, m, n, spread)
    if zero_row:
        A[0, :] = 0.0
    b = crt_beta(k, N)
    rows = [crt_integers(list(A[i, :]), b) for i in range(m)]
    cols = [crt_integers(list(B[:, j]), b) for j in range(n)]
    out = np.zeros((m, n))
    for i in range(m):
        ia, ea = rows[i]
        for j in range(n):
            ib, eb = cols[j]
            c = sum(x * y for x, y in zip(ia, ib))
            assert 2 * abs(c) < math.prod(CRT_MODULI[:N])
            v = math.ldexp(float(c), ea + eb - 2 * b)   # float(int) rounds to nearest even
            v = rn(F(alpha) * F(v))
            out[i, j] = v if beta == 0.0 else fma(beta, C[i, j], v)
    return {"m": m, "n": n, "k": k, "N": N, "alpha": float(alpha).hex(),
            "beta": float(beta).hex(), "A": hexs(A), "B": hexs(B), "C": hexs(C),
            "out": hexs(out)}


def main():
    rng = np.random.default_rng(0x6D6F6D6D73)
    g = {"ukernel": [], "goto": [], "triple": [], "exact": [], "ozaki": [], "crt": []}
    for k, al, be in [(0, 1.0, 1.0), (1, 1.0, 1.0), (4, 1.0, 0.0), (5, -2.0, 0.5), (9, 0.75, 1.0),
                      (17, 1.0, -1.5)]:
        g["ukernel"].append(ukernel_case(rng, k, al, be))
    for m, n, k, al, be in [(7, 9, 5, 1.0, 1.0), (6, 8, 200, 1.0, 1.0), (7, 9, 197, -1.0, 0.5),
                            (5, 3, 385, 2.0, 0.0), (125, 2, 3, 1.0, 1.0)]:
        g["goto"].append(goto_case(rng, m, n, k, al, be))
    g["triple"].append(triple_case(rng, 3, 4, 5))
    for m, n, k, al, be, sp in [(3, 3, 40, 1.0, 1.0, 0), (2, 2, 64, 1.0, 1.0, 30),
                                (3, 2, 17, -0.5, 2.0, 8)]:
        g["exact"].append(exact_case(rng, m, n, k, al, be, sp))
    for m, n, k, s, al, be, sp in [(3, 2, 300, 3, 1.0, 1.0, 0), (2, 3, 130, 7, 1.0, 1.0, 8),
                                   (2, 2, 50, 4, -1.0, 0.0, 16)]:
        g["ozaki"].append(ozaki_case(rng, m, n, k, s, al, be, sp))
    for m, n, k, N, al, be, sp, zr in [(3, 2, 300, 14, 1.0, 1.0, 0, False),
                                       (2, 3, 130, 16, 1.0, 1.0, 16, True),
                                       (3, 3, 64, 8, -0.75, 0.0, 8, False),
                                       (2, 2, 4200, 13, 1.0, 2.0, 4, False),
                                       (2, 2, 0, 14, 1.0, 0.5, 0, False)]:
        g["crt"].append(crt_case(rng, m, n, k, N, al, be, sp, zr))
    OUT.write_text(json.dumps(g))
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
