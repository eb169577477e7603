/*
 * wilson_oracle.c -- CPU restatement of CLWilson's Wilson-prime algorithm.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * the library built from it. The product (clwilson_amd/) never links it.
 *
 * It restates, in plain C with exact 64/128-bit integer arithmetic, what the
 * reference computes for one range [pmin, pmax):
 *   - the "double-precision" Montgomery arithmetic mod p^2 (kernels/common.cl)
 *   - the base-2 strong PRP test of the segment sieve (kernels/getsegprps.cl)
 *   - Legendre power tables (cl_wilson.cpp getPower)
 *   - test-prime typing and type targets (cl_wilson.cpp:1552-1586)
 *   - the plain method: setup -> mulsmall/mullarge -> reduce -> iterate -> get
 *     (setup.cl, mulsmall.cl, mullarge.cl, reduce.cl, iterate.cl)
 *   - find a/c/u (find.cl) and the final check processResult
 *     (cl_wilson.cpp:691-933), checksum and result lines.
 *
 * Where the reference multiplies base-2 strong PRPs into the products and later
 * divides the composite ones back out with prps.dat (cl_wilson.cpp:709-726),
 * this oracle multiplies exact primes only. The net value is identical and
 * prps.dat is not part of the mount.
 *
 * Pinning (each function has a check in tests/test_oracle.py):
 *   m2p_* arithmetic     vs Python big integers on random operands, all widths
 *   strong_prp_two       vs the known base-2 strong pseudoprimes and sympy
 *   exact prime sieve    vs sympy / pi(x) values
 *   legendre             vs brute-force factorisation of n!
 *   find_a/c/u           vs brute force identities, uniqueness, Cornacchia
 *   wilson_range         vs direct (p-1)! mod p^2 for small p; the reference
 *                        self-test case [200,564); CGH published w_p values
 *   finalize             vs Python big-integer restatement of processResult
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef unsigned __int128 u128;
typedef struct { uint64_t s0, s1; } u64x2;

/* ------------------------------------------------------------------------ */
/* common.cl:17-113  128-bit helpers                                          */
/* ------------------------------------------------------------------------ */

/* common.cl:17-38: r0 + 2^64 r1 = a + b (mod 2^128) */
static inline u64x2 add_wide(u64x2 a, u64x2 b) {
  u64x2 r = {a.s0 + b.s0, a.s1 + b.s1};
  if (r.s0 < a.s0) r.s1 += 1;
  return r;
}
/* common.cl:41-67: full 64x64 -> 128 product */
static inline u64x2 mul_wide(uint64_t a, uint64_t b) {
  u128 t = (u128)a * b;
  u64x2 r = {(uint64_t)t, (uint64_t)(t >> 64)};
  return r;
}
static inline uint64_t mul_hi(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) >> 64); }
/* common.cl:70-73 */
static inline int is_less_than(u64x2 a, u64x2 b) { return (a.s1 < b.s1) || ((a.s1 == b.s1) && (a.s0 < b.s0)); }
/* common.cl:76-81: p * p_inv == 1 (mod 2^64) by Newton iteration */
uint64_t orc_invert(uint64_t p) {
  uint64_t p_inv = 1, prev = 0;
  while (p_inv != prev) { prev = p_inv; p_inv *= 2 - p * p_inv; }
  return p_inv;
}
/* common.cl:84-113 */
static inline uint64_t add_mod(uint64_t x, uint64_t y, uint64_t p) { uint64_t cp = (x >= p - y) ? p : 0; return x + y - cp; }
static inline uint64_t add_mod_c(uint64_t x, uint64_t y, uint64_t p, uint64_t* c) {
  int carry = (x >= p - y); *c = carry ? 1 : 0; return x + y - (carry ? p : 0);
}
static inline uint64_t sub_mod(uint64_t x, uint64_t y, uint64_t p) { uint64_t cp = (x < y) ? p : 0; return x - y + cp; }
static inline uint64_t sub_mod_c(uint64_t x, uint64_t y, uint64_t p, uint64_t* c) {
  int carry = (x < y); *c = carry ? 1 : 0; return x - y + (carry ? p : 0);
}

/* ------------------------------------------------------------------------ */
/* common.cl:115-260  digit-wise Montgomery arithmetic mod p^2, R = 2^64      */
/* A residue (x0, x1) has value x0 + p*x1 with 0 <= x0, x1 < p.              */
/* ------------------------------------------------------------------------ */

/* common.cl:120-130: R mod p^2 */
u64x2 orc_m2p_one(uint64_t p) {
  if ((p >> 32) == 0) {
    const uint64_t p2 = p * p, r_p2 = (0 - p2) % p2;
    u64x2 r = {r_p2 % p, r_p2 / p};
    return r;
  }
  const uint64_t mp = 0 - p;
  u64x2 r = {mp % p, mp / p + 1};
  return r;
}
/* common.cl:133-139 */
u64x2 orc_m2p_dup(u64x2 x, uint64_t p) {
  uint64_t c; uint64_t l = add_mod_c(x.s0, x.s0, p, &c); uint64_t h = add_mod(x.s1 + c, x.s1, p);
  u64x2 r = {l, h}; return r;
}
/* common.cl:142-148 */
u64x2 orc_m2p_add(u64x2 x, u64x2 y, uint64_t p) {
  uint64_t c; uint64_t l = add_mod_c(x.s0, y.s0, p, &c); uint64_t h = add_mod(x.s1 + c, y.s1, p);
  u64x2 r = {l, h}; return r;
}
/* common.cl:151-173 */
u64x2 orc_m2p_square(u64x2 x, uint64_t p, uint64_t q) {
  const u64x2 t = mul_wide(x.s0, x.s0);
  const uint64_t u0 = q * t.s0, t1 = t.s1, v1 = mul_hi(p, u0);
  const u64x2 x01 = mul_wide(x.s0, x.s1);
  const u64x2 u0w = {u0, 0};
  const u64x2 x01u = add_wide(x01, u0w);
  const u64x2 tp = add_wide(x01u, x01);
  const int tp_carry = is_less_than(tp, x01);
  const uint64_t tp_h = tp.s1, tpc = (tp_carry | (tp_h >= p)) ? p : 0;
  const uint64_t up0 = q * tp.s0, t1p = tp_h - tpc, v1p = mul_hi(p, up0);
  uint64_t c;
  const uint64_t z0 = sub_mod_c(t1, v1, p, &c);
  const uint64_t z1 = sub_mod(t1p, v1p + c, p);
  u64x2 r = {z0, z1}; return r;
}
/* common.cl:176-198 */
u64x2 orc_m2p_mul(u64x2 x, u64x2 y, uint64_t p, uint64_t q) {
  const u64x2 t = mul_wide(x.s0, y.s0);
  const uint64_t u0 = q * t.s0, t1 = t.s1, v1 = mul_hi(p, u0);
  const u64x2 u0w = {u0, 0};
  const u64x2 x0y1u = add_wide(mul_wide(x.s0, y.s1), u0w);
  const u64x2 x1y0 = mul_wide(x.s1, y.s0);
  const u64x2 tp = add_wide(x0y1u, x1y0);
  const int tp_carry = is_less_than(tp, x1y0);
  const uint64_t tp_h = tp.s1, tpc = (tp_carry | (tp_h >= p)) ? p : 0;
  const uint64_t up0 = q * tp.s0, t1p = tp_h - tpc, v1p = mul_hi(p, up0);
  uint64_t c;
  const uint64_t z0 = sub_mod_c(t1, v1, p, &c);
  const uint64_t z1 = sub_mod(t1p, v1p + c, p);
  u64x2 r = {z0, z1}; return r;
}
/* common.cl:217-243: x * (y0 + p y1) * R^-1 for a plain 64-bit x */
u64x2 orc_m2p_mul_r2(uint64_t x, u64x2 y, uint64_t p, uint64_t q) {
  const u64x2 t = mul_wide(x, y.s0);
  const uint64_t u0 = q * t.s0, t1 = t.s1, v1 = mul_hi(p, u0);
  const u64x2 u0w = {u0, 0};
  const u64x2 tp = add_wide(mul_wide(x, y.s1), u0w);
  const uint64_t tp_h = tp.s1, tpc = (tp_h >= p) ? p : 0;
  const uint64_t up0 = q * tp.s0, t1p = tp_h - tpc, v1p = mul_hi(p, up0);
  uint64_t c;
  const uint64_t z0 = sub_mod_c(t1, v1, p, &c);
  const uint64_t z1 = sub_mod(t1p, v1p + c, p);
  u64x2 r = {z0, z1}; return r;
}
/* common.cl:246-260: REDC to canonical digits */
u64x2 orc_m2p_get(u64x2 x, uint64_t p, uint64_t q) {
  const uint64_t u0 = q * x.s0, v1 = mul_hi(p, u0);
  const uint64_t tp = x.s1 + u0, up0 = q * tp;
  const uint64_t t1p = (tp < x.s1) ? 1 : 0;
  const uint64_t v1p = mul_hi(p, up0);
  uint64_t c;
  const uint64_t z0 = sub_mod_c(0, v1, p, &c);
  const uint64_t z1 = sub_mod(t1p, v1p + c, p);
  u64x2 r = {z0, z1}; return r;
}

/* setup.cl:20-29: per-prime constants q, one, R^2 mod p^2 */
typedef struct { uint64_t p, q; u64x2 one, r2; } tpconst;
static tpconst make_tpconst(uint64_t p) {
  tpconst c; c.p = p; c.q = orc_invert(p);
  c.one = orc_m2p_one(p);
  u64x2 two = orc_m2p_dup(c.one, p);
  u64x2 r2 = orc_m2p_dup(two, p);
  for (int i = 0; i < 5; ++i) r2 = orc_m2p_square(r2, p, c.q); /* 4^(2^5) = 2^64 */
  c.r2 = r2;
  return c;
}
void orc_setup_constants(uint64_t p, uint64_t out[6]) {
  tpconst c = make_tpconst(p);
  out[0] = c.p; out[1] = c.q; out[2] = c.one.s0; out[3] = c.one.s1; out[4] = c.r2.s0; out[5] = c.r2.s1;
}

/* ------------------------------------------------------------------------ */
/* getsegprps.cl:67-147  base-2 strong probable-prime test                   */
/* ------------------------------------------------------------------------ */
static inline uint64_t m_mul(uint64_t a, uint64_t b, uint64_t p, uint64_t q) {
  u64x2 ab = mul_wide(a, b);
  uint64_t m = ab.s0 * q, mp = mul_hi(m, p), r = ab.s1 - mp;
  return (ab.s1 < mp) ? r + p : r;
}
int orc_strong_prp_two(uint64_t p) {
  const uint64_t q = orc_invert(p);
  const uint64_t one = (0 - p) % p, nmo = p - one;
  const uint64_t two = add_mod(one, one, p);
  const int t = __builtin_ctzll(p - 1);
  const uint64_t exp = p >> t;
  uint64_t curBit = 0x8000000000000000ull;
  curBit >>= (__builtin_clzll(exp) + 1);
  uint64_t a = two;
  while (curBit) {
    a = m_mul(a, a, p, q);
    if (exp & curBit) a = add_mod(a, a, p);
    curBit >>= 1;
  }
  if (a == one || a == nmo) return 1;
  for (int s = 1; s < t; ++s) {
    a = m_mul(a, a, p, q);
    if (a == nmo) return 1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Exact prime generation (the primesieve calls of cl_wilson.cpp:1316-1329,  */
/* 1529): a segmented sieve of Eratosthenes over odd numbers.                */
/* ------------------------------------------------------------------------ */
static uint32_t* g_base = NULL;  /* odd primes up to g_base_lim */
static size_t g_nbase = 0;
static uint64_t g_base_lim = 0;

static void ensure_base(uint64_t lim) {
  if (lim <= g_base_lim) return;
  uint64_t L = lim < 1024 ? 1024 : lim;
  uint8_t* comp = (uint8_t*)calloc(L + 1, 1);
  for (uint64_t i = 2; i * i <= L; ++i)
    if (!comp[i]) for (uint64_t j = i * i; j <= L; j += i) comp[j] = 1;
  size_t n = 0;
  for (uint64_t i = 3; i <= L; i += 2) if (!comp[i]) ++n;
  uint32_t* b = (uint32_t*)malloc(n * sizeof(uint32_t));
  n = 0;
  for (uint64_t i = 3; i <= L; i += 2) if (!comp[i]) b[n++] = (uint32_t)i;
  free(comp);
  free(g_base);
  g_base = b; g_nbase = n; g_base_lim = L;
}

static uint64_t isqrt64(uint64_t n) {
  uint64_t r = (uint64_t)sqrtl((long double)n);
  while (r * r > n) --r;
  while ((r + 1) * (r + 1) <= n) ++r;
  return r;
}

/* Primes in [lo, hi) written to out (at most cap). Returns the total count
 * (which may exceed cap; then only cap are written). */
uint64_t orc_primes_range(uint64_t lo, uint64_t hi, uint64_t* out, uint64_t cap) {
  if (hi <= lo) return 0;
  uint64_t cnt = 0;
  if (lo <= 2 && hi > 2) { if (cnt < cap && out) out[cnt] = 2; ++cnt; }
  uint64_t s = lo < 3 ? 3 : (lo | 1);
  if (s >= hi) return cnt;
  ensure_base(isqrt64(hi) + 1);
  const uint64_t SEG = 1u << 22; /* odd numbers per segment */
  uint8_t* mark = (uint8_t*)malloc(SEG);
  for (uint64_t a = s; a < hi; a += 2 * SEG) {
    uint64_t b = a + 2 * SEG; if (b > hi) b = hi;
    uint64_t n = (b - a + 1) / 2;
    memset(mark, 0, n);
    for (size_t k = 0; k < g_nbase; ++k) {
      uint64_t q = g_base[k];
      if (q * q >= b) break;
      uint64_t st = q * q;
      if (st < a) { st = ((a + q - 1) / q) * q; if ((st & 1) == 0) st += q; }
      for (uint64_t m = st; m < b; m += 2 * q) mark[(m - a) >> 1] = 1;
    }
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t v = a + 2 * i;
      if (v >= b) break;
      if (!mark[i] && v > 1) { if (cnt < cap && out) out[cnt] = v; ++cnt; }
    }
  }
  free(mark);
  return cnt;
}

/* ------------------------------------------------------------------------ */
/* cl_wilson.cpp:1293-1313 getPower: Legendre exponent of q in T! and the    */
/* bit below its top bit (the left-to-right square-and-multiply start).      */
/* ------------------------------------------------------------------------ */
void orc_legendre(uint64_t prime, uint64_t target, uint64_t out[2]) {
  if (prime > target) { out[0] = 0; out[1] = 0; return; }
  uint64_t totalpower = 0, currp = prime, qq = target / currp;
  while (qq) {
    totalpower += qq;
    u128 pp = (u128)currp * prime;
    if (pp > target) break;
    currp = (uint64_t)pp;
    qq = target / currp;
  }
  uint64_t curBit = 0x8000000000000000ull;
  if (totalpower > 1) curBit >>= (__builtin_clzll(totalpower) + 1);
  out[0] = totalpower; out[1] = curBit;
}

/* mulsmall.cl:32-51 / mullarge.cl:38-58: q^e in Montgomery form */
static u64x2 prime_power_mont(uint64_t prime, uint64_t e, uint64_t curBit, const tpconst* c) {
  const u64x2 base = orc_m2p_mul_r2(prime, c->r2, c->p, c->q);
  if (e == 1) return base;
  u64x2 a = base;
  while (curBit) {
    a = orc_m2p_square(a, c->p, c->q);
    if (e & curBit) a = orc_m2p_mul(a, base, c->p, c->q);
    curBit >>= 1;
  }
  return a;
}

/* ------------------------------------------------------------------------ */
/* Exact arithmetic mod m = p^2 < 2^126 for the final check (GMP in the      */
/* reference, cl_wilson.cpp:691-933). Deliberately simple: 256-bit product,  */
/* bitwise reduction.                                                        */
/* ------------------------------------------------------------------------ */
static u128 mulmod128(u128 a, u128 b, u128 m) {
  uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64), b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
  u128 p00 = (u128)a0 * b0, p01 = (u128)a0 * b1, p10 = (u128)a1 * b0, p11 = (u128)a1 * b1;
  /* 256-bit value w3:w2:w1:w0 */
  uint64_t w0 = (uint64_t)p00;
  u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;
  uint64_t w1 = (uint64_t)mid;
  u128 hi = (mid >> 64) + (p01 >> 64) + (p10 >> 64) + p11;
  uint64_t w2 = (uint64_t)hi, w3 = (uint64_t)(hi >> 64);
  uint64_t w[4] = {w0, w1, w2, w3};
  u128 r = 0;
  for (int i = 255; i >= 0; --i) {
    r = (r << 1) | ((w[i >> 6] >> (i & 63)) & 1);
    if (r >= m) r -= m;
  }
  return r;
}
static u128 addmod128(u128 a, u128 b, u128 m) { u128 s = a + b; return s >= m ? s - m : s; }
static u128 submod128(u128 a, u128 b, u128 m) { return a >= b ? a - b : a + (m - b); }
static u128 powmod128(u128 b, u128 e, u128 m) {
  u128 r = 1 % m;
  b %= m;
  while (e) { if (e & 1) r = mulmod128(r, b, m); b = mulmod128(b, b, m); e >>= 1; }
  return r;
}
/* signed 64-bit value reduced mod m */
static u128 smod128(int64_t v, u128 m) {
  if (v >= 0) return (u128)(uint64_t)v % m;
  u128 a = (u128)(uint64_t)(-(v + 1)) + 1; /* |v| without overflow */
  a %= m;
  return a ? m - a : 0;
}
/* modular inverse mod m by extended Euclid on signed 128-bit; 0 if none */
static int invmod128(u128 a, u128 m, u128* out) {
  __int128 t = 0, newt = 1;
  __int128 r = (__int128)m, newr = (__int128)(a % m);
  while (newr != 0) {
    __int128 qt = r / newr, tmp;
    tmp = t - qt * newt; t = newt; newt = tmp;
    tmp = r - qt * newr; r = newr; newr = tmp;
  }
  if (r != 1) return 0;
  if (t < 0) t += (__int128)m;
  *out = (u128)t;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* find.cl:14-117 + cl_wilson.cpp:587-654: a, c, u by exhaustive search      */
/* ------------------------------------------------------------------------ */
/* find.cl:14-43 bitwise integer square root */
uint64_t orc_isqrt_bitwise(uint64_t n) {
  uint64_t x = n, c = 0, d = 1ull << 62;
  while (d > n) d >>= 2;
  while (d != 0) {
    if (x >= c + d) { x -= c + d; c = (c >> 1) + d; }
    else c >>= 1;
    d >>= 2;
  }
  return c;
}
/* finda: a^2 + b^2 = p with a odd, a = 1 (mod 4) after the sign flip */
int64_t orc_find_a(uint64_t p) {
  uint64_t maxa = (uint64_t)sqrtl((long double)p);
  for (uint64_t a = 1; a <= maxa; a += 2) {
    uint64_t b2 = p - a * a, b = orc_isqrt_bitwise(b2);
    if (b * b == b2) return (a % 4 == 3) ? -(int64_t)a : (int64_t)a;
  }
  return 0;
}
/* findc: c^2 + 27 d^2 = 4p, c = 1 (mod 3) after the sign flip */
int64_t orc_find_c(uint64_t p) {
  uint64_t p4 = 4 * p, maxd = (uint64_t)sqrtl(((long double)p4) / 27.0);
  for (uint64_t d = 1; d <= maxd; ++d) {
    uint64_t c2 = p4 - 27 * d * d; int64_t c = (int64_t)orc_isqrt_bitwise(c2);
    if ((uint64_t)(c * c) == c2) return (c % 3 == 2) ? -c : c;
  }
  return 0;
}
/* findu: u^2 + 3 v^2 = 4p, u = umod (mod 3); nth selects among solutions in
 * increasing v (the reference takes whichever lane wins an atomic race). */
int64_t orc_find_u_nth(uint64_t p, int nth) {
  uint64_t p4 = 4 * p, maxv = (uint64_t)sqrtl(((long double)p4) / 3.0);
  uint32_t umod = (((p - 1) / 6) % 2 == 0) ? 1 : 2;
  int k = 0;
  for (uint64_t v = 1; v <= maxv; ++v) {
    uint64_t u2 = p4 - 3 * v * v; int64_t u = (int64_t)orc_isqrt_bitwise(u2);
    if ((uint64_t)(u * u) == u2) {
      int64_t r = (u % 3 == (int64_t)umod) ? u : -u;
      if (k == nth) return r;
      ++k;
    }
  }
  return 0;
}
int64_t orc_find_u(uint64_t p) { return orc_find_u_nth(p, 0); }

/* ------------------------------------------------------------------------ */
/* cl_wilson.cpp:691-870 processResult: from the canonical value of          */
/* T_p! mod p^2 (T_p = (p-1)/6, /4, /2 by type) to (quot, rem) of            */
/* (p-1)! + 1 = rem + p*quot (mod p^2). u_override/ c/a override: 0 = search.*/
/* ------------------------------------------------------------------------ */
void orc_finalize_ex(uint64_t p, uint32_t type, uint64_t res_lo, uint64_t res_hi,
                     int64_t acu1, int64_t acu2, uint64_t out[2]) {
  const u128 mp = p, m = mp * mp;
  u128 residue = (((u128)res_hi << 64) | res_lo) % m;
  u128 a, b;
  if (type == 0) {
    int64_t uu = acu1 ? acu1 : orc_find_u(p);
    int64_t cc = acu2 ? acu2 : orc_find_c(p);
    const u128 mu = smod128(uu, m), mc = smod128(cc, m);
    residue = powmod128(residue, 6, m);                       /* ((p-1)/6)!^6 */
    a = submod128(powmod128(2, p, m), 1, m);                  /* 2^p - 1 */
    b = submod128(0, mulmod128(mulmod128(mu, mu, m), mu, m), m); /* -u^3 */
    a = mulmod128(b, a, m);
    b = mulmod128(mulmod128(mu, mp % m, m), 3, m);            /* 3 p u */
    a = addmod128(a, b, m);
    residue = mulmod128(residue, a, m);
    u128 ci;
    if (!invmod128(mc, m, &ci)) { out[0] = out[1] = ~0ull; return; }
    a = submod128(mulmod128(mp, ci, m), mc, m);               /* p/c - c */
    residue = mulmod128(residue, a, m);
    a = submod128(powmod128(3, p, m), 1, m);                  /* 3^p - 1 */
    u128 i2 = 0; invmod128(2, m, &i2);
    a = mulmod128(a, i2, m);                                  /* (3^p-1)/2 */
    residue = mulmod128(residue, a, m);
  } else if (type == 1) {
    int64_t aa = acu1 ? acu1 : orc_find_a(p);
    const u128 ma = smod128(aa, m);
    residue = powmod128(residue, 4, m);                       /* ((p-1)/4)!^4 */
    a = submod128(mulmod128(powmod128(2, p, m), 3, m), 4 % m, m); /* 3*2^p - 4 */
    residue = mulmod128(residue, a, m);
    a = submod128(mulmod128(mulmod128(ma, ma, m), 2, m), mp % m, m); /* 2a^2 - p */
    residue = mulmod128(residue, a, m);
  } else {
    residue = mulmod128(residue, residue, m);                 /* ((p-1)/2)!^2 */
    a = submod128(1, powmod128(2, p, m), m);                  /* 1 - 2^p */
    residue = mulmod128(residue, a, m);
  }
  residue = addmod128(residue, 1, m);
  out[0] = (uint64_t)(residue / mp); /* quot */
  out[1] = (uint64_t)(residue % mp); /* rem */
}
void orc_finalize(uint64_t p, uint32_t type, uint64_t res_lo, uint64_t res_hi, uint64_t out[2]) {
  orc_finalize_ex(p, type, res_lo, res_hi, 0, 0, out);
}

/* Direct (p-1)! mod p^2 for p < 2^32 (independent of everything above). */
void orc_factorial_mod_p2_direct(uint64_t p, uint64_t n, uint64_t out[2]) {
  const uint64_t m = p * p;
  uint64_t r = 1 % m;
  for (uint64_t k = 2; k <= n; ++k) r = (uint64_t)(((u128)r * k) % m);
  out[0] = r; out[1] = 0;
}

/* Direct Wilson quotients for many primes: (p-1)! mod p^2 by multiplying
 * 2, 3, ..., p-1 one after another (the textbook definition, nothing of the
 * plain method), then quot = ((p-1)!+1)/p mod p and rem = ((p-1)!+1) mod p.
 * Montgomery products mod m = p^2 with R = 2^64 keep it to a few ns per
 * factor: every product of a plain k adds R^-1, removed at the end with
 * R^(p-2) mod m. p < 2^31 (m < 2^62), OpenMP over the primes. */
static uint64_t mont64_inv(uint64_t m) { /* -m^-1 mod 2^64, m odd */
  uint64_t x = m;
  for (int i = 0; i < 6; ++i) x *= 2 - m * x;
  return (uint64_t)0 - x;
}
static inline uint64_t mont64_mul(uint64_t a, uint64_t b, uint64_t m, uint64_t minv) {
  const u128 t = (u128)a * b;
  const uint64_t q = (uint64_t)t * minv;
  const u128 u = (t + (u128)q * m) >> 64;
  return (uint64_t)(u >= m ? u - m : u);
}
void orc_direct_quotients(const uint64_t* ps, uint64_t n, uint64_t* quot, uint64_t* rem, int nthreads) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t p = ps[i], m = p * p, minv = mont64_inv(m);
    uint64_t r = 1;
    for (uint64_t k = 2; k < p; ++k) r = mont64_mul(r, k, m, minv); /* (p-1)! R^-(p-2) */
    /* R^(p-2) mod m by square and multiply with u128 remainders */
    const uint64_t Rm = (uint64_t)(((u128)1 << 64) % m);
    uint64_t fix = 1 % m, b = Rm, e = p - 2;
    while (e) {
      if (e & 1) fix = (uint64_t)(((u128)fix * b) % m);
      b = (uint64_t)(((u128)b * b) % m);
      e >>= 1;
    }
    const uint64_t f = (uint64_t)(((u128)r * fix) % m); /* (p-1)! mod p^2 */
    const uint64_t w = (f + 1) % m;
    quot[i] = w / p;
    rem[i] = w % p;
  }
}

/* ------------------------------------------------------------------------ */
/* The plain method for one range (cl_wilson.cpp:1442-1974).                 */
/* ------------------------------------------------------------------------ */
static uint32_t type_of(uint64_t p, uint64_t* ptarget) {
  /* cl_wilson.cpp:1552-1586 */
  if (p % 3 == 1) { *ptarget = (p - 1) / 6; return 0; }
  if (p % 12 == 5) { *ptarget = (p - 1) / 4; return 1; }
  *ptarget = (p - 1) / 2; return 2; /* p % 12 == 11 */
}

/* canonical u128 value of digits (z0, z1): z0 + p*z1 */
static inline u128 digits_value(u64x2 z, uint64_t p) { return (u128)z.s0 + (u128)p * z.s1; }

/* Per test prime: canonical value of prod_{q prime in [qlo,qhi), q <= T}
 * q^{e_q(T)} mod p^2, computed as mulsmall/mullarge + reduce do
 * (Montgomery digits, left-to-right powers, one running product). */
static u128 window_product_one(const tpconst* c, uint64_t T, const uint64_t* primes, uint64_t np) {
  u64x2 acc = c->one;
  for (uint64_t i = 0; i < np; ++i) {
    const uint64_t q = primes[i];
    if (q > T) break;
    uint64_t pw[2];
    orc_legendre(q, T, pw);
    u64x2 f = prime_power_mont(q, pw[0], pw[1], c);
    acc = orc_m2p_mul(acc, f, c->p, c->q);
  }
  return digits_value(orc_m2p_get(acc, c->p, c->q), c->p);
}

void orc_window_product(const uint64_t* tp, const uint64_t* targets, uint32_t ntp, uint64_t qlo, uint64_t qhi,
                        int nthreads, uint64_t* out_lo, uint64_t* out_hi) {
  uint64_t maxT = 0;
  for (uint32_t i = 0; i < ntp; ++i) if (targets[i] > maxT) maxT = targets[i];
  if (qhi > maxT + 1) qhi = maxT + 1;
  uint64_t np = qhi > qlo ? orc_primes_range(qlo, qhi, NULL, 0) : 0;
  uint64_t* primes = (uint64_t*)malloc((np ? np : 1) * sizeof(uint64_t));
  if (np) orc_primes_range(qlo, qhi, primes, np);
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (uint32_t i = 0; i < ntp; ++i) {
    tpconst c = make_tpconst(tp[i]);
    u128 v = window_product_one(&c, targets[i], primes, np);
    out_lo[i] = (uint64_t)v; out_hi[i] = (uint64_t)(v >> 64);
  }
  free(primes);
}

/* iterate.cl:28-58: prod_{n=a+1}^{b} n, times the canonical residue r, mod p^2 */
static u128 iterate_one(const tpconst* c, uint64_t a, uint64_t b, u128 r) {
  u64x2 acc = c->one;
  if (b > a) {
    u64x2 mn = orc_m2p_mul_r2(a + 1, c->r2, c->p, c->q);
    const u64x2 m1 = orc_m2p_mul_r2(1, c->r2, c->p, c->q);
    for (uint64_t n = a + 1; n <= b; ++n) {
      acc = orc_m2p_mul(mn, acc, c->p, c->q);
      mn = orc_m2p_add(mn, m1, c->p);
    }
  }
  /* continue from r: convert r (canonical) to Montgomery form, multiply */
  const uint64_t p = c->p;
  u64x2 rd = {(uint64_t)(r % p), (uint64_t)(r / p)};
  u64x2 rm = orc_m2p_mul(rd, c->r2, p, c->q); /* r * R */
  acc = orc_m2p_mul(acc, rm, p, c->q);
  return digits_value(orc_m2p_get(acc, p, c->q), p);
}

void orc_iterate(const uint64_t* tp, const uint64_t* from, const uint64_t* to, uint32_t ntp, int nthreads,
                 uint64_t* io_lo, uint64_t* io_hi) {
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (uint32_t i = 0; i < ntp; ++i) {
    tpconst c = make_tpconst(tp[i]);
    u128 r = ((u128)io_hi[i] << 64) | io_lo[i];
    u128 v = iterate_one(&c, from[i], to[i], r);
    io_lo[i] = (uint64_t)v; io_hi[i] = (uint64_t)(v >> 64);
  }
}

typedef struct {
  uint64_t checksum;       /* sd.checksum */
  uint64_t totalcount;     /* pi(maxtarget): st.totalcount - sd.prpsremoved */
  uint64_t typeTarget[3];
  uint64_t maxtarget;
  uint64_t lastp;          /* sd.testResultPrime */
  int64_t lastw;           /* sd.testResultValue */
  uint32_t tpcount, tpcnt[3], resultcount, pad;
} orc_stats;

/* Whole range [pmin, pmax): per test prime (p, type, quot, rem) and stats.
 * out arrays may be NULL; if non-NULL they need room for pi(pmax)-pi(pmin). */
int orc_wilson_range(uint64_t pmin, uint64_t pmax, int nthreads,
                     uint64_t* out_p, uint32_t* out_type, uint64_t* out_quot, uint64_t* out_rem, orc_stats* st) {
  memset(st, 0, sizeof(*st));
  if (pmin < 5 || pmax <= pmin) return -1;
  uint64_t ntp = orc_primes_range(pmin, pmax, NULL, 0);
  if (!ntp) return -2;
  uint64_t* tpl = (uint64_t*)malloc(ntp * sizeof(uint64_t));
  orc_primes_range(pmin, pmax, tpl, ntp);
  uint32_t* ty = (uint32_t*)malloc(ntp * sizeof(uint32_t));
  uint64_t* ptarget = (uint64_t*)malloc(ntp * sizeof(uint64_t));
  uint64_t* ttarget = (uint64_t*)malloc(ntp * sizeof(uint64_t));
  for (uint64_t i = 0; i < ntp; ++i) {
    ty[i] = type_of(tpl[i], &ptarget[i]);
    st->tpcnt[ty[i]]++;
    if (!st->typeTarget[ty[i]]) st->typeTarget[ty[i]] = ptarget[i];
  }
  for (uint64_t i = 0; i < ntp; ++i) ttarget[i] = st->typeTarget[ty[i]];
  st->maxtarget = st->typeTarget[0];
  for (int t = 1; t < 3; ++t) if (st->typeTarget[t] > st->maxtarget) st->maxtarget = st->typeTarget[t];
  st->tpcount = (uint32_t)ntp;
  st->totalcount = st->maxtarget >= 2 ? orc_primes_range(2, st->maxtarget + 1, NULL, 0) : 0;

  uint64_t* lo = (uint64_t*)malloc(ntp * sizeof(uint64_t));
  uint64_t* hi = (uint64_t*)malloc(ntp * sizeof(uint64_t));
  orc_window_product(tpl, ttarget, (uint32_t)ntp, 2, st->maxtarget + 1, nthreads, lo, hi);
  orc_iterate(tpl, ttarget, ptarget, (uint32_t)ntp, nthreads, lo, hi);
  uint64_t* quot = (uint64_t*)malloc(ntp * sizeof(uint64_t));
  uint64_t* rem = (uint64_t*)malloc(ntp * sizeof(uint64_t));
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (uint64_t i = 0; i < ntp; ++i) {
    uint64_t o[2];
    orc_finalize(tpl[i], ty[i], lo[i], hi[i], o);
    quot[i] = o[0]; rem[i] = o[1];
  }
  /* cl_wilson.cpp:866-900: rem check, near-Wilson classification, checksum */
  int bad = 0;
  for (uint64_t i = 0; i < ntp; ++i) {
    const uint64_t p = tpl[i];
    if (rem[i] != 0) bad = 1;
    const uint64_t negquot = p - quot[i];
    const uint64_t smallest = (quot[i] <= negquot) ? quot[i] : negquot;
    if (quot[i] == 0) { st->resultcount++; st->lastp = p; st->lastw = 0; }
    else if (smallest < 1000) {
      int32_t dq = (smallest == quot[i]) ? (int32_t)quot[i] : -((int32_t)negquot);
      st->resultcount++; st->lastp = p; st->lastw = dq;
    }
    st->checksum += p + rem[i] + quot[i];
    if (out_p) out_p[i] = p;
    if (out_type) out_type[i] = ty[i];
    if (out_quot) out_quot[i] = quot[i];
    if (out_rem) out_rem[i] = rem[i];
  }
  free(tpl); free(ty); free(ptarget); free(ttarget); free(lo); free(hi); free(quot); free(rem);
  return bad ? -3 : 0;
}
