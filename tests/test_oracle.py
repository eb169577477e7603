"""The CPU oracle pinned against mathematics and the reference's own answers.

Every oracle function is checked against something it does not share code
with: Python big integers, sympy, direct factorials, the reference's self-test
constants and the Costa-Gerbicz-Harvey published w_p values.
"""
import math
import random

import numpy as np
import pytest
import sympy

import oracle as O

R64 = 1 << 64


def _val(x, p):
    return x.s0 + p * x.s1


def _digits(v, p):
    return O.U64x2(v % p, v // p)


def _rand_prime(rng, bits):
    while True:
        n = rng.getrandbits(bits) | 1 | (1 << (bits - 1))
        if n > 5 and sympy.isprime(n):
            return n


@pytest.mark.parametrize("bits", [3, 8, 16, 31, 32, 33, 44, 45, 52, 61, 62])
def test_m2p_arithmetic_vs_bigint(bits):
    """common.cl value semantics: mul = xyR^-1, mul_r2 = x*val(y)*R^-1,
    square = x^2 R^-1, get = xR^-1, one = R mod p^2 (R = 2^64)."""
    L = O.lib()
    rng = random.Random(bits)
    for _ in range(150):
        p = _rand_prime(rng, bits) if bits > 3 else rng.choice([5, 7])
        m = p * p
        rinv = pow(R64, -1, m)
        q = L.orc_invert(p)
        assert (p * q) % R64 == 1
        one = L.orc_m2p_one(p)
        assert _val(one, p) == R64 % m and one.s0 < p and one.s1 < p
        x, y = rng.randrange(m), rng.randrange(m)
        X, Y = _digits(x, p), _digits(y, p)
        z = L.orc_m2p_mul(X, Y, p, q)
        assert z.s0 < p and z.s1 < p and _val(z, p) == x * y * rinv % m
        z = L.orc_m2p_square(X, p, q)
        assert _val(z, p) == x * x * rinv % m
        a = rng.randrange(R64)  # any 64-bit value, also >= p^2
        z = L.orc_m2p_mul_r2(a, Y, p, q)
        assert _val(z, p) == a * y * rinv % m
        assert _val(L.orc_m2p_get(X, p, q), p) == x * rinv % m
        assert _val(L.orc_m2p_add(X, Y, p), p) == (x + y) % m
        assert _val(L.orc_m2p_dup(X, p), p) == 2 * x % m


def test_setup_constants():
    L = O.lib()
    import ctypes as C
    for p in [7, 1000003, 4294967311, 20000000000021, (1 << 61) - 1]:
        o = (C.c_uint64 * 6)()
        L.orc_setup_constants(p, o)
        m = p * p
        assert o[2] + p * o[3] == R64 % m
        assert o[4] + p * o[5] == R64 * R64 % m


def _spsp2_def(n):
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(2, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(1, s):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def test_strong_prp_two():
    """getsegprps.cl:101-147 against the definition, all odd n in [117, 2e5)
    and random 64-bit n; the first base-2 strong pseudoprimes pass."""
    L = O.lib()
    for n in range(117, 200000, 2):
        assert L.orc_strong_prp_two(n) == int(_spsp2_def(n)), n
    for n in (2047, 3277, 4033, 4681, 8321):
        assert not sympy.isprime(n) and L.orc_strong_prp_two(n) == 1
    rng = random.Random(5)
    for _ in range(3000):
        n = rng.getrandbits(rng.randint(20, 62)) | 1
        if n > 117:
            assert L.orc_strong_prp_two(n) == int(_spsp2_def(n))


def test_primes_range():
    assert list(O.primes_range(0, 1000)) == list(sympy.primerange(0, 1000))
    assert list(O.primes_range(10**9, 10**9 + 5000)) == list(sympy.primerange(10**9, 10**9 + 5000))
    assert len(O.primes_range(0, 10**6)) == 78498
    assert len(O.primes_range(0, 10**7)) == 664579


def test_legendre_vs_factorisation():
    """getPower (cl_wilson.cpp:1293-1313): exponent of q in T! and curBit."""
    for T in [1, 2, 10, 57, 113, 300]:
        f = sympy.factorint(math.factorial(T)) if T > 1 else {}
        for q in sympy.primerange(2, 320):
            e, cb = O.legendre(q, T)
            assert e == f.get(q, 0)
            if e > 1:
                assert cb == 1 << (e.bit_length() - 2)
    e, _ = O.legendre(3, 10**12)
    assert e == sum(10**12 // 3**i for i in range(1, 30))


def _all_reps(p):
    a = [x for x in range(1, math.isqrt(p) + 1, 2) if math.isqrt(p - x * x) ** 2 == p - x * x]
    c = [x for x in range(1, 2 * math.isqrt(p) + 2) if (4 * p - x * x) > 0 and (4 * p - x * x) % 27 == 0
         and math.isqrt((4 * p - x * x) // 27) ** 2 == (4 * p - x * x) // 27]
    return a, c


def test_find_acu():
    """find.cl:46-117 normalisation: a = 1 (mod 4) with a^2+b^2 = p;
    c = 1 (mod 3) with c^2+27d^2 = 4p; u = umod (mod 3) with u^2+3v^2 = 4p."""
    L = O.lib()
    for p in sympy.primerange(5, 6000):
        a_s, c_s = _all_reps(p)
        if p % 4 == 1:
            a = L.orc_find_a(p)
            assert abs(a) in a_s and a % 4 == 1 and len(a_s) == 1
        if p % 3 == 1:
            c = L.orc_find_c(p)
            assert abs(c) in c_s and c % 3 == 1 and len(c_s) == 1
            umod = 1 if ((p - 1) // 6) % 2 == 0 else 2
            for k in range(3):
                u = L.orc_find_u_nth(p, k)
                v2 = 4 * p - u * u
                assert v2 % 3 == 0 and math.isqrt(v2 // 3) ** 2 == v2 // 3 and u % 3 == umod


def _fact_mod(n, m):
    r = 1
    for k in range(2, n + 1):
        r = r * k % m
    return r


def test_finalize_identities_vs_direct_factorial():
    """processResult's type identities (cl_wilson.cpp:731-846): from T_p! the
    oracle must reproduce (p-1)! + 1 mod p^2 computed directly."""
    for p in list(sympy.primerange(5, 4000)) + [100003, 1000003, 1000033]:
        m = p * p
        t = 0 if p % 3 == 1 else (1 if p % 12 == 5 else 2)
        T = (p - 1) // (6, 4, 2)[t]
        quot, rem = O.finalize(p, t, _fact_mod(T, m))
        full = (_fact_mod(p - 1, m) + 1) % m
        assert (quot, rem) == (full // p, full % p), p


def test_u_choice_invariance():
    """The reference keeps whichever u an atomic race finds first
    (find.cl:92-114); every solution must give the same result."""
    L = O.lib()
    rng = random.Random(3)
    ps = [p for p in sympy.primerange(7, 3000) if p % 3 == 1]
    ps += [q for q in (sympy.nextprime(rng.randrange(10**5, 10**6)) for _ in range(60)) if q % 3 == 1]
    for p in ps:
        V = _fact_mod((p - 1) // 6, p * p)
        c = L.orc_find_c(p)
        outs = {O.finalize(p, 0, V, L.orc_find_u_nth(p, k), c) for k in range(3)}
        assert len(outs) == 1


def test_wilson_range_selftest_case5(golden):
    """cl_wilson.cpp:2076-2089: [200, 564) -> 57 results, checksum 0x80A3,
    pi(maxT) = 30, last (563, 0)."""
    kat = golden("kat_selftest.json")[4]
    arr, st = O.wilson_range(kat["pmin"], kat["pmax"])
    assert st["rc"] == 0
    assert st["resultcount"] == kat["results"]
    assert f"{st['checksum']:016X}" == kat["checksum"]
    assert st["totalcount"] == kat["pi_maxt"]
    assert (st["lastp"], st["lastw"]) == (kat["lastp"], kat["lastw"])


def test_wilson_range_vs_direct():
    """Every prime of a small range against (p-1)! mod p^2 directly."""
    arr, st = O.wilson_range(5, 3000)
    for p, q, r in zip(arr["p"], arr["quot"], arr["rem"]):
        p = int(p)
        full = (_fact_mod(p - 1, p * p) + 1) % (p * p)
        assert (int(q), int(r)) == (full // p, full % p)


def test_direct_factorial_samples():
    L = O.lib()
    import ctypes as C
    for p in [1000003, 1999993]:
        o = (C.c_uint64 * 2)()
        L.orc_factorial_mod_p2_direct(p, p - 1, o)
        assert o[0] == _fact_mod(p - 1, p * p)


@pytest.mark.parametrize("p,w", [(5, 0), (13, 0), (563, 0), (8315831, 3), (10746881, -7), (11892977, -7),
                                 (14296621, 2), (29085907, 9), (39198017, -7), (51802061, 4),
                                 (56151923, -1), (87467099, -2), (93559087, -3)])
def test_cgh_published_values(p, w):
    arr, st = O.wilson_range(p, p + 1)
    assert st["rc"] == 0 and st["resultcount"] == 1
    assert st["lastw"] == w


def test_golden_small_consistent(golden):
    """oracle_small.json was written by this oracle; p < 20000 also direct."""
    for case in golden("oracle_small.json"):
        if case["pmax"] > 200000:
            continue
        arr, st = O.wilson_range(case["pmin"], case["pmax"])
        assert [int(x) for x in arr["quot"]] == case["quot"]
        assert f"{st['checksum']:016X}" == case["checksum"]


def test_direct_fixture_matches_definition_and_plain_method(golden):
    """direct_lt_2e6.json holds every prime p < 2e6 by the definition
    (p-1)! = 2*3*...*(p-1) mod p^2: its samples are recomputed here with
    Python integers (p < 2e4) and the oracle's direct product, and on
    [1e6, 2e6) its quotient hash and checksum equal cfg1.json, which the
    plain-method oracle wrote -- two independent algorithms agree on all
    70,435 primes."""
    d = golden("direct_lt_2e6.json")
    cfg1 = golden("cfg1.json")
    assert [x["tpcount"] for x in d] == [78496, 70435]  # pi(1e6) - 2, pi(2e6) - pi(1e6)
    big = d[1]
    assert (big["quot_sha256"], big["checksum"]) == (cfg1["quot_sha256"], cfg1["checksum"])
    samples = np.array([p for p, _ in d[0]["samples"] + d[1]["samples"]], dtype=np.uint64)
    q, r = O.direct_quotients(samples)
    assert [int(x) for x in q] == [w for _, w in d[0]["samples"] + d[1]["samples"]] and not r.any()
    for p, w in d[0]["samples"][:40]:
        if p < 20000:
            assert (_fact_mod(p - 1, p * p) + 1) % (p * p) == w * p
