"""The device arithmetic headers (clw_arith.h, clw_nt64.h) compiled for the
host CPU (libclw_hostcheck.so) and checked against Python big integers,
sympy and the oracle's exhaustive searches."""
import ctypes as C
import os
import random

import pytest
import sympy

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HC = os.path.join(ROOT, "clwilson_amd", "libclw_hostcheck.so")


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(HC):
        from clwilson_amd import build
        build.build_hostcheck()
    L = C.CDLL(HC)
    u64p = C.POINTER(C.c_uint64)
    L.clwh_montmul.argtypes = [C.c_int, u64p, u64p, u64p, u64p]
    L.clwh_undrift.argtypes = [C.c_int, C.c_uint64, u64p, C.c_uint64, u64p]
    L.clwh_montmul_lockstep.argtypes = [C.c_int, u64p, u64p, u64p, u64p]
    L.clwh_unmasked_m.argtypes = [C.c_int]
    u32p = C.POINTER(C.c_uint32)
    L.clwh_montmul_lazy.argtypes = [C.c_int, u32p, u64p, u64p, u32p]
    L.clwh_lazy_norm.argtypes = [C.c_int, u32p, u64p, u32p]
    L.clwh_lazy_params.argtypes = [C.c_int, C.POINTER(C.c_int)]
    L.clwh_redc_long.argtypes = [C.c_int, u32p, C.c_uint32, u64p, u64p]
    L.clwh_neg_inv30.restype = C.c_uint32
    L.clwh_neg_inv30.argtypes = [C.c_uint32]
    L.clwh_legendre.restype = C.c_uint64
    L.clwh_legendre.argtypes = [C.c_uint64, C.c_uint64]
    L.clwh_isqrt.restype = C.c_uint64
    L.clwh_isqrt.argtypes = [C.c_uint64]
    L.clwh_inv_mod.restype = C.c_uint64
    L.clwh_inv_mod.argtypes = [C.c_uint64, C.c_uint64]
    L.clwh_mr_bases_needed.argtypes = [C.c_uint64]
    L.clwh_is_prime_mr.argtypes = [C.c_uint64, C.c_int]
    L.clwh_two_squares_a.restype = C.c_int64
    L.clwh_two_squares_a.argtypes = [C.c_uint64]
    L.clwh_eisenstein_cu.argtypes = [C.c_uint64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    return L


def _arr(v):
    return (C.c_uint64 * 2)(v & (2**64 - 1), v >> 64)


@pytest.mark.parametrize("L", [2, 3, 4, 5])
def test_montmul_lazy_bounds(hc, L):
    """a*b*R^-1 mod N with the result < 2N for a < 2N, b < R/2 (R = 2^(30L))."""
    R = 2 ** (30 * L)
    rng = random.Random(L)
    n_ok = 0
    while n_ok < 1500:
        bits = rng.randint(3, min(15 * L - 1, 62))
        p = sympy.nextprime(rng.getrandbits(bits) | 3)
        N = p * p
        if 2 * N >= R:
            continue
        a = rng.randrange(2 * N)
        b = rng.randrange(min(R // 2, 2**128))
        out = (C.c_uint64 * 2)()
        hc.clwh_montmul(L, _arr(a), _arr(b), _arr(N), out)
        r = out[0] | (out[1] << 64)
        assert r < 2 * N and r % N == a * b * pow(R, -1, N) % N
        # both operands < 2N (mul_full before its conditional subtraction)
        b2 = rng.randrange(2 * N)
        hc.clwh_montmul(L, _arr(a), _arr(b2), _arr(N), out)
        r = out[0] | (out[1] << 64)
        assert r < 3 * N and r % N == a * b2 * pow(R, -1, N) % N
        n_ok += 1


def test_montmul_chain_without_cond_sub(hc):
    """The matrix-core run accumulator (3 limbs) multiplies X by the reduced
    run product Y < N (1 + 2^-12) with montmul alone: for N < 0.4998 R the
    chain stays below 2N (and right mod N), from the largest X and Y up."""
    L, R = 3, 2**90
    rng = random.Random(11)
    for p in (sympy.prevprime(int((0.4998 * R) ** 0.5)), sympy.nextprime(2**44), sympy.nextprime(10**13),
              sympy.nextprime(2**30)):
        N = p * p
        assert N < 0.4998 * R
        ymax = N + (N >> 12)
        x = 2 * N - 1
        for k in range(3000):
            y = ymax - 1 if k % 3 == 0 else rng.randrange(ymax)
            out = (C.c_uint64 * 2)()
            hc.clwh_montmul(L, _arr(x), _arr(y), _arr(N), out)
            r = out[0] | (out[1] << 64)
            assert r < 2 * N and r % N == x * y * pow(R, -1, N) % N
            x = r if k % 5 else 2 * N - 1


def _column_bound(L, U):
    """Largest 64-bit column sum of montmul_lockstep: limbs of a, b, n below
    2^30, quotient digits m_k below 2^32 for k < U (unmasked), else 2^30."""
    A = 2**30 - 1
    M = [2**32 - 1 if k < U else A for k in range(L)]
    carry = worst = 0
    for k in range(2 * L - 1):
        s = carry + sum(A * A for i in range(L) if 0 <= k - i < L)
        s += sum(M[i] * A for i in range(L) if i < k and 0 <= k - i < L)
        if k < L:
            s += M[k] * A
        worst = max(worst, s)
        carry = s >> 30
    return worst


@pytest.mark.parametrize("L", [2, 3, 4, 5])
def test_lockstep_unmasked_digits_fit(hc, L):
    """The unmasked quotient digits keep every column below 2^64, the top
    digit stays masked (so m < R (1 + 2^-26)), and for 3 and 4 limbs one
    more unmasked digit would overflow a column."""
    U = hc.clwh_unmasked_m(L)
    assert 1 <= U <= L - 1
    assert _column_bound(L, U) < 2**64
    if L in (3, 4):
        assert U == L - 1 or _column_bound(L, U + 1) >= 2**64


@pytest.mark.parametrize("L", [2, 3, 4, 5])
def test_montmul_lockstep_bounds(hc, L):
    """The lockstep product (unmasked low digits) is a b R^-1 mod N and stays
    below 2N for a < 2N and b < N/4 (record factors Q <= T^2 < p^2/4), for
    random primes of the band and the largest prime of the band, with
    extreme operands."""
    R = 2 ** (30 * L)
    rng = random.Random(100 + L)
    pmax = {2: 759250124, 3: 24879108095803, 4: 815238614083298888, 5: 2**62}[L]
    primes = [sympy.prevprime(pmax)] + [sympy.nextprime(rng.randrange(pmax // 4, pmax)) for _ in range(40)]
    out = (C.c_uint64 * 2)()
    for p in primes:
        N = p * p
        assert 2 * N < R
        cases = [(2 * N - 1, N // 4 - 1), (2 * N - 1, 1), (0, N // 4 - 1)]
        cases += [(rng.randrange(2 * N), rng.randrange(N // 4)) for _ in range(60)]
        for a, b in cases:
            hc.clwh_montmul_lockstep(L, _arr(a), _arr(b), _arr(N), out)
            r = out[0] | (out[1] << 64)
            assert r < 2 * N and r % N == a * b * pow(R, -1, N) % N, (L, p, a, b)


# Largest p of each batch limb band (2 p^2 < R = 2^(30 L)).
BAND_PMAX = {2: 759250124, 3: 24879108095803, 4: 815238614083298888}


def _lazy_params(hc, L):
    b = C.c_int()
    u = hc.clwh_lazy_params(L, C.byref(b))
    return u, b.value


def _lazy_limbs(v, L, t):
    """Lazy-form limbs of value v: limb 0 takes t extra units of 2^30 (so up
    to 32 bits), the top limb takes what is left."""
    lo = (v % 2**30) + 2**30 * t
    assert lo < 2**32 and lo <= v
    rest = (v - lo) >> 30
    limbs = [lo]
    for i in range(1, L):
        limbs.append(rest if i == L - 1 else rest % 2**30)
        rest >>= 30
    assert limbs[-1] < 2**32
    return limbs


def _limbs_value(x):
    return sum(int(w) << (30 * i) for i, w in enumerate(x))


@pytest.mark.parametrize("L", [2, 3, 4])
def test_lazy_column_bounds(hc, L):
    """Every 64-bit column sum of montmul_lazy fits, limb by limb, for the
    largest modulus of the band: a in lazy form (limb 0 below 2^32, value
    below B N), b < N/4, N's top limb from the band limit, unmasked quotient
    digits below 2^32; and the lazy bound a < B N is preserved:
    B N^2 / (4 R) + m N / R < B N."""
    U, B = _lazy_params(hc, L)
    R = 2 ** (30 * L)
    N = BAND_PMAX[L] ** 2
    assert 2 * N < R
    top = 30 * (L - 1)
    amax = [2**32 - 1] + [2**30 - 1] * (L - 2) + [min(2**32 - 1, (B * N) >> top)]
    bmax = [2**30 - 1] * (L - 1) + [(N // 4) >> top]
    nmax = [2**30 - 1] * (L - 1) + [N >> top]
    mmax = [2**32 - 1 if k < U else 2**30 - 1 for k in range(L)]
    carry = worst = 0
    for k in range(2 * L - 1):
        s = carry + sum(amax[i] * bmax[k - i] for i in range(L) if 0 <= k - i < L)
        s += sum(mmax[i] * nmax[k - i] for i in range(L) if i < k and 0 <= k - i < L)
        if k < L:
            s += mmax[k] * nmax[0]
        worst = max(worst, s)
        carry = 4 * (s >> 32) if k == L else s >> 30
    assert worst < 2**64, (L, worst / 2**64)
    m_total = sum(mk << (30 * k) for k, mk in enumerate(mmax))
    assert B * N * N // 4 + m_total * N < B * N * R
    # the top limb of a result below B N fits 32 bits
    assert (B * N) >> top < 2**32


@pytest.mark.parametrize("L", [2, 3, 4])
def test_montmul_lazy_products_and_norm(hc, L):
    """The lazy product on extreme and random lazy operands is a b R^-1 mod
    N, below B N, with limb 0 below 2^32 and the middle limbs below 2^30;
    chains of products stay exact; lazy_norm gives the strict form (30-bit
    limbs, below 2N) of the same residue."""
    U, B = _lazy_params(hc, L)
    R = 2 ** (30 * L)
    Rinv_cache = {}
    rng = random.Random(500 + L)
    pmax = BAND_PMAX[L]
    primes = [sympy.prevprime(pmax)] + [sympy.nextprime(rng.randrange(pmax // 4, pmax)) for _ in range(25)]
    a_buf, out = (C.c_uint32 * L)(), (C.c_uint32 * L)()
    for p in primes:
        N = p * p
        Rinv = pow(R, -1, N)
        vals = [B * N - 1, B * N - 2**31, 1, 2**32 - 1] + [rng.randrange(B * N) for _ in range(40)]
        for v in vals:
            tmax = min(3, (v >> 30), (2**32 - 1 - v % 2**30) >> 30)
            a = _lazy_limbs(v, L, rng.randint(0, tmax) if tmax > 0 else 0)
            for b in (N // 4 - 1, 1, rng.randrange(N // 4)):
                a_buf[:] = a
                hc.clwh_montmul_lazy(L, a_buf, _arr(b), _arr(N), out)
                r = _limbs_value(out)
                assert r % N == v * b * Rinv % N and r < B * N, (L, p, v, b)
                assert out[0] < 2**32 and all(out[i] < 2**30 for i in range(1, L - 1))
                hc.clwh_lazy_norm(L, out, _arr(N), a_buf)
                s = _limbs_value(a_buf)
                assert s < 2 * N and s % N == r % N and all(w < 2**30 for w in a_buf)
        # a chain of 300 products from the largest lazy value
        x = _lazy_limbs(B * N - 1, L, 0)
        want = (B * N - 1) % N
        for _ in range(300):
            b = rng.choice([N // 4 - 1, rng.randrange(N // 4)])
            a_buf[:] = x
            hc.clwh_montmul_lazy(L, a_buf, _arr(b), _arr(N), out)
            x = list(out)
            want = want * b * Rinv % N
            assert _limbs_value(x) < B * N
        assert _limbs_value(x) % N == want


@pytest.mark.parametrize("L", [2, 3, 4])
def test_redc_long(hc, L):
    """Long Montgomery reduction of a D-digit integer (the product of a run
    of sieving primes): P 2^(-30 D) mod N in strict form (30-bit limbs,
    below 2N), for all-ones digits, products of real primes and random
    digits, D from one group to 270 digits."""
    rng = random.Random(700 + L)
    pmax = BAND_PMAX[L]
    primes = [sympy.prevprime(pmax)] + [sympy.nextprime(rng.randrange(pmax // 4, pmax)) for _ in range(12)]
    out = (C.c_uint64 * 2)()
    for p in primes:
        N = p * p
        G = 2 * L if L % 2 else L  # digits per unrolled group
        for D in sorted({G, 2 * G, 30 - 30 % G, 96 - 96 % G, 270 - 270 % G}):
            cases = [[2**30 - 1] * D, [0] * (D - 1) + [1], [rng.randrange(2**30) for _ in range(D)]]
            # a product of consecutive primes (the run products the kernel reduces)
            q0 = sympy.nextprime(rng.randrange(2, min(p, 2**40)))
            P, q = 1, q0
            while P.bit_length() + q.bit_length() <= 30 * D:
                P *= q
                q = sympy.nextprime(q)
            cases.append([(P >> (30 * i)) & (2**30 - 1) for i in range(D)])
            for digits in cases:
                P = sum(d << (30 * i) for i, d in enumerate(digits))
                arr = (C.c_uint32 * D)(*digits)
                assert hc.clwh_redc_long(L, arr, D, _arr(N), out) == 0
                y = out[0] | (out[1] << 64)
                assert y < 2 * N, (L, p, D)
                assert y % N == P * pow(2, -30 * D, N) % N, (L, p, D)


@pytest.mark.parametrize("L", [2, 3, 4, 5])
def test_undrift(hc, L):
    R = 2 ** (30 * L)
    rng = random.Random(10 + L)
    for _ in range(300):
        bits = rng.randint(3, min(15 * L - 1, 62))
        p = sympy.nextprime(rng.getrandbits(bits) | 3)
        N = p * p
        if 2 * N >= R:
            continue
        a = rng.randrange(2 * N)
        d = rng.choice([0, 1, 2, rng.randrange(1000), rng.randrange(2**40)])
        out = (C.c_uint64 * 2)()
        hc.clwh_undrift(L, p, _arr(a), d, out)
        assert (out[0] | (out[1] << 64)) == a * pow(R, d, N) % N


def test_neg_inv30(hc):
    for n0 in [1, 3, 0x3FFFFFFF, 123456789 | 1]:
        assert (n0 * hc.clwh_neg_inv30(n0)) % 2**30 == 2**30 - 1


def test_miller_rabin_exact(hc):
    for n in range(39, 300000, 2):
        assert hc.clwh_is_prime_mr(n, hc.clwh_mr_bases_needed(n)) == int(sympy.isprime(n)), n
    rng = random.Random(1)
    for _ in range(20000):
        n = rng.getrandbits(rng.randint(20, 64)) | 1
        assert hc.clwh_is_prime_mr(n, hc.clwh_mr_bases_needed(n)) == int(sympy.isprime(n)), n
    # base-2 strong pseudoprimes are rejected by the extra bases
    for n in (2047, 3277, 4033, 4681, 8321, 3215031751, 2152302898747, 3474749660383):
        assert hc.clwh_is_prime_mr(n, hc.clwh_mr_bases_needed(n)) == 0


def test_legendre_isqrt_inverse(hc):
    rng = random.Random(2)
    for _ in range(5000):
        T = rng.randrange(1, 2**61)
        q = int(sympy.prevprime(rng.randrange(3, min(T + 3, 2**40))))
        assert hc.clwh_legendre(q, T) == O.legendre(q, T)[0]
        n = rng.getrandbits(64)
        r = hc.clwh_isqrt(n)
        assert r * r <= n < (r + 1) * (r + 1)
        p = sympy.nextprime(rng.randrange(10**6, 2**62))
        a = rng.randrange(1, p)
        assert a * hc.clwh_inv_mod(a, p) % p == 1


def test_cornacchia_vs_exhaustive(hc):
    """find_a / find_c (exhaustive, find.cl) equal Cornacchia's results."""
    L = O.lib()
    rng = random.Random(4)
    ps = list(sympy.primerange(5, 30000)) + [sympy.nextprime(rng.randrange(10**6, 10**10)) for _ in range(400)]
    for p in ps:
        if p % 4 == 1:
            assert hc.clwh_two_squares_a(p) == L.orc_find_a(p)
        if p % 3 == 1:
            c, u = C.c_int64(), C.c_int64()
            hc.clwh_eisenstein_cu(p, C.byref(c), C.byref(u))
            assert c.value == L.orc_find_c(p)
            umod = 1 if ((p - 1) // 6) % 2 == 0 else 2
            assert abs(u.value) == abs(c.value) and u.value % 3 == umod


def test_cornacchia_large(hc):
    """p up to 2^62: identities checked directly (exhaustive search too slow)."""
    import math
    rng = random.Random(6)
    for _ in range(3000):
        p = sympy.nextprime(rng.randrange(10**10, 2**62))
        if p % 4 == 1:
            a = hc.clwh_two_squares_a(p)
            b2 = p - a * a
            assert a % 4 == 1 and b2 > 0 and math.isqrt(b2) ** 2 == b2
        if p % 3 == 1:
            c, u = C.c_int64(), C.c_int64()
            hc.clwh_eisenstein_cu(p, C.byref(c), C.byref(u))
            d2 = 4 * p - c.value ** 2
            assert c.value % 3 == 1 and d2 % 27 == 0 and math.isqrt(d2 // 27) ** 2 == d2 // 27
