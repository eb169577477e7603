"""The arithmetic of the matrix-core run products (k_subchunk_prod's byte
form, k_mfma_table / the in-LDS table build, k_mfma_runs' epilogue) restated
step by step in Python integers with the kernels' word sizes, and checked
against big-integer arithmetic: the identities V = P 2^30 (mod N) and
Y = P (mod N), and every bound the kernels rely on (int32 column sums,
V > 0, Y < N (1 + 2^-12) < 2N, signed 64-bit sums without overflow, the top
limb of Y below 2^30). Products of real prime runs, all-ones digits and
extreme table values; moduli across the 3-limb band."""
import random

import pytest
import sympy

KB = 768        # bytes per sub-chunk product in the byte form
COLS = 12       # bytes of c < N < 2^89
M64 = (1 << 64) - 1
LMASK = (1 << 30) - 1


def s64(x):
    """Two's-complement 64-bit value of x (x must not overflow)."""
    assert -(1 << 63) <= x < (1 << 63), "signed 64-bit sum overflows"
    return x


def digits30(P):
    d = []
    while P:
        d.append(P & LMASK)
        P >>= 30
    return d


def balanced_bytes_from_digits(D, nbytes):
    """k_subchunk_prod: byte i from the 30-bit digits, then the carry rule
    (a byte >= 128 becomes byte - 256 with a carry into the next)."""
    out, carry = [], 0
    for i in range(nbytes):
        k, off = divmod(8 * i, 30)
        d0 = D[k] if k < len(D) else 0
        d1 = D[k + 1] if k + 1 < len(D) else 0
        v = (((d0 >> off) | ((d1 << (30 - off)) & 0xFFFFFFFF)) & 255) + carry
        carry = 1 if v >= 128 else 0
        b = v & 255
        out.append(b - 256 if b >= 128 else b)
    return out


def table(N, nbytes):
    """c_i = 2^(8 i + 30) mod N as balanced bytes: bytes of c + 0x80..80 with
    the top bit flipped (k_mfma_runs' build), checked equal to the carry rule
    (k_mfma_table)."""
    M = int("80" * COLS, 16)
    c = pow(2, 30, N)
    rows = []
    for _ in range(nbytes):
        y = c + M
        assert y < 1 << (8 * COLS)
        alt = [(((y >> (8 * k)) & 255) ^ 0x80) for k in range(COLS)]
        row, carry = [], 0
        for k in range(COLS):
            v = ((c >> (8 * k)) & 255) + carry
            carry = 1 if v >= 128 else 0
            row.append(v & 255)
        assert carry == 0 and row == alt
        rows.append([b - 256 if b >= 128 else b for b in row])
        c = c * 256 % N
    return rows


def epilogue(cols, N):
    """mfma_pair: V + m N column by column in signed 64-bit sums, / 2^30."""
    n = [(N >> (30 * j)) & LMASK for j in range(3)]
    ninv = (-pow(n[0], -1, 1 << 30)) % (1 << 30)
    W = N << 17
    Wlo, Whi = W & ((1 << 60) - 1), W >> 60
    sh = [1, 1 << 8, 1 << 16, 1 << 24, 1 << 2, 1 << 10, 1 << 18, 1 << 26, 1 << 4, 1 << 12, 1 << 20, 1 << 28]
    s0 = Wlo
    for k in range(4):
        s0 = s64(s0 + cols[k] * sh[k])
    m = ((s0 & 0xFFFFFFFF) * ninv) & LMASK
    s0 = s64(s0 + m * n[0])
    assert s0 & LMASK == 0
    s1 = s64((s0 >> 30) + m * n[1])
    for k in range(4, 8):
        s1 = s64(s1 + cols[k] * sh[k])
    s2 = s64((s1 >> 30) + Whi + m * n[2])
    for k in range(8, 12):
        s2 = s64(s2 + cols[k] * sh[k])
    top = s2 >> 30
    assert 0 <= top < 1 << 30
    return (s1 & LMASK) | ((s2 & LMASK) << 30) | (top << 60)


def check(P, N, tab):
    D = digits30(P)
    nbytes = min(KB, (30 * len(D) + 7) // 8 + 1)
    a = balanced_bytes_from_digits(D, KB)
    assert sum(x << (8 * i) for i, x in enumerate(a)) == P
    assert all(x == 0 for x in a[nbytes:])
    cols = [sum(a[i] * tab[i][k] for i in range(nbytes)) for k in range(COLS)]
    assert all(abs(c) < 768 * 128 * 128 for c in cols)  # int32 accumulators
    V = sum(c << (8 * k) for k, c in enumerate(cols))
    assert V % N == P * 2**30 % N and V + (N << 17) > 0
    Y = epilogue(cols, N)
    assert Y % N == P % N
    assert Y < N + (N >> 12) < 2 * N


@pytest.mark.parametrize("pbits", [30, 37, 44])
def test_run_products_by_bytes(pbits):
    rng = random.Random(pbits)
    p = sympy.prevprime(int(2 ** 44.49)) if pbits == 44 else sympy.nextprime(rng.getrandbits(pbits) | 1 << (pbits - 1))
    N = p * p
    assert 2 * N < 2**90
    tab = table(N, KB)
    # products of 128 consecutive primes (the shape of a sub-chunk), near the
    # largest targets of the band and small
    for q0 in (2 * 10**12, 10**9, 2**43):
        P, q = 1, q0
        for _ in range(128):
            q = sympy.nextprime(q)
            P *= q
        check(P, N, tab)
    # all-ones digits (every byte at its carry edge) and random long products
    for nd in (1, 60, 187, 200):
        check((1 << (30 * nd)) - 1, N, tab)
    for _ in range(5):
        check(rng.getrandbits(rng.randrange(8, 5600)) | 1, N, tab)


def test_extreme_table_columns():
    """Column sums at their extremes: every byte -128 against table bytes of
    one sign stay inside int32, and V stays positive."""
    p = sympy.prevprime(int(2 ** 44.49))
    N = p * p
    a = [-128] * KB
    tab = table(N, KB)
    for k in range(COLS):
        worst = sum(abs(a[i] * tab[i][k]) for i in range(KB))
        assert worst < 2**31
    V = sum(sum(a[i] * tab[i][k] for i in range(KB)) << (8 * k) for k in range(COLS))
    assert V + (N << 17) > 0


def _in_one_run(q_first, q_last, T, sqrtT):
    """The batch kernels' test that a sub-chunk of sieving primes
    [q_first, q_last] lies in one exponent run of target T above sqrt(T)
    (then k_mfma_runs applies it and k_batch_mul skips it)."""
    return q_first > sqrtT and q_last <= T and q_last <= T // (T // q_first)


def test_whole_chunk_skip_implies_every_sub_chunk_taken():
    """k_batch_mul skips a work item when its chunk's whole integer span
    [qa, qz] passes the one-run test. Brute force over small targets: such a
    span holds one exponent floor(T / q) (the one of the Legendre formula for
    q > sqrt(T)), and every sub-interval [qf, ql] of it passes the per-block
    test, so no sub-chunk is left to the vector units; a span one past the
    run's end fails it."""
    import math
    rng = random.Random(5)
    checked = 0
    for T in [97, 1000, 4096, 12345, 99991] + [rng.randrange(50, 200000) for _ in range(30)]:
        sqrtT = math.isqrt(T)
        for _ in range(300):
            qa = rng.randrange(2, T + 2)
            qz = qa + rng.randrange(0, 2 * T // max(qa, 1) + 3)
            if not _in_one_run(qa, qz, T, sqrtT):
                continue
            checked += 1
            e = T // qa
            assert all(T // q == e and q > sqrtT and q <= T for q in range(qa, qz + 1))
            for _ in range(5):
                qf = rng.randrange(qa, qz + 1)
                ql = rng.randrange(qf, qz + 1)
                assert _in_one_run(qf, ql, T, sqrtT)
            # the run's last q is T // e: one past it starts the next run (or passes T)
            assert not _in_one_run(qa, T // e + 1, T, sqrtT)
    assert checked > 1000
