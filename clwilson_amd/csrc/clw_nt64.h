// clw_nt64.h -- 64-bit number theory on the device: Montgomery arithmetic mod
// an odd n < 2^64, deterministic Miller-Rabin, Legendre exponents, integer
// square roots, and the two-squares / Eisenstein representations of a prime
// (Cornacchia's algorithm) used by the final Wilson check.
#pragma once
#include "clw_hd.h"

namespace clw {

// n^-1 mod 2^64 for odd n (Newton; each step doubles the correct bits).
CLW_HD uint64_t inv64(uint64_t n) {
  uint64_t x = n;  // 3 bits
  for (int i = 0; i < 5; ++i) x *= 2 - n * x;
  return x;
}

CLW_HD uint64_t umulhi64(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) >> 64); }

// a*b*2^-64 mod n for a, b < n (REDC with the subtraction form).
CLW_HD uint64_t mmul64(uint64_t a, uint64_t b, uint64_t n, uint64_t ninv) {
  const uint64_t lo = a * b, hi = umulhi64(a, b);
  const uint64_t m = lo * ninv;
  const uint64_t mh = umulhi64(m, n);
  const uint64_t r = hi - mh;
  return hi < mh ? r + n : r;
}

CLW_HD uint64_t addmod64(uint64_t a, uint64_t b, uint64_t n) {
  const uint64_t s = a + b;
  return (s < a || s >= n) ? s - n : s;
}
CLW_HD uint64_t submod64(uint64_t a, uint64_t b, uint64_t n) {
  return a >= b ? a - b : a + (n - b);
}

struct M64 {
  uint64_t n, ninv, one;  // one = 2^64 mod n
};
CLW_HD M64 m64_make(uint64_t n) {
  M64 m;
  m.n = n;
  m.ninv = inv64(n);
  m.one = (0 - n) % n;
  return m;
}
// Montgomery form of a small integer s (s < 2^32) by double-and-add of one.
CLW_HD uint64_t m64_small(const M64& m, uint32_t s) {
  uint64_t r = 0, a = m.one;
  while (s) {
    if (s & 1) r = addmod64(r, a, m.n);
    a = addmod64(a, a, m.n);
    s >>= 1;
  }
  return r;
}
// Montgomery form of an arbitrary x < n: x * 2^64 mod n by 64 doublings.
CLW_HD uint64_t m64_to(const M64& m, uint64_t x) {
  uint64_t r = x % m.n;
  for (int i = 0; i < 64; ++i) r = addmod64(r, r, m.n);
  return r;
}
CLW_HD uint64_t m64_from(const M64& m, uint64_t xm) { return mmul64(xm, 1, m.n, m.ninv); }
CLW_HD uint64_t m64_pow(const M64& m, uint64_t bm, uint64_t e) {
  uint64_t r = m.one;
  while (e) {
    if (e & 1) r = mmul64(r, bm, m.n, m.ninv);
    bm = mmul64(bm, bm, m.n, m.ninv);
    e >>= 1;
  }
  return r;
}

// One strong-probable-prime round to base a (a given in Montgomery form).
CLW_HD bool sprp_m(const M64& m, uint64_t am, uint64_t d, int s) {
  const uint64_t nm1 = m.n - m.one;  // Montgomery form of n-1
  uint64_t x = m64_pow(m, am, d);
  if (x == m.one || x == nm1) return true;
  for (int i = 1; i < s; ++i) {
    x = mmul64(x, x, m.n, m.ninv);
    if (x == nm1) return true;
    if (x == m.one) return false;
  }
  return false;
}

// Base-2 strong test with doubling in place of multiplication by the base.
CLW_HD bool sprp2(const M64& m, uint64_t d, int s) {
  const uint64_t nm1 = m.n - m.one;
  const int top = 63 - __builtin_clzll(d);
  uint64_t x = addmod64(m.one, m.one, m.n);  // 2 in Montgomery form
  for (int b = top - 1; b >= 0; --b) {
    x = mmul64(x, x, m.n, m.ninv);
    if ((d >> b) & 1) x = addmod64(x, x, m.n);
  }
  if (x == m.one || x == nm1) return true;
  for (int i = 1; i < s; ++i) {
    x = mmul64(x, x, m.n, m.ninv);
    if (x == nm1) return true;
    if (x == m.one) return false;
  }
  return false;
}

// Number of leading prime bases {2,3,5,...,37} that make Miller-Rabin
// deterministic below the bound (Jaeschke 1993; Feitsma/Galway base-2 data;
// Sorenson & Webster 2015 for the first 12 primes).
CLW_HD int mr_bases_needed(uint64_t n_upper) {
  if (n_upper < 2047ull) return 1;
  if (n_upper < 1373653ull) return 2;
  if (n_upper < 25326001ull) return 3;
  if (n_upper < 3215031751ull) return 4;
  if (n_upper < 2152302898747ull) return 5;
  if (n_upper < 3474749660383ull) return 6;
  if (n_upper < 341550071728321ull) return 7;
  if (n_upper < 3825123056546413051ull) return 9;
  return 12;
}

// The extra bases (after 2) for an odd n > 37 known to pass base 2.
CLW_HD bool mr_extra(const M64& m, uint64_t d, int s, int nbases) {
  const uint32_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (int k = 1; k < nbases; ++k) {
    if (!sprp_m(m, m64_small(m, bases[k]), d, s)) return false;
  }
  return true;
}

// Legendre exponent of prime q in T!: sum_{j>=1} floor(T / q^j).
CLW_HD uint64_t legendre_exp(uint64_t q, uint64_t T) {
  if (q > T) return 0;
  uint64_t x = T / q, e = x;
  while (x >= q) {
    x /= q;
    e += x;
  }
  return e;
}

// floor(sqrt(n)) by the digit-by-digit (binary restoring) method.
CLW_HD uint64_t isqrt64(uint64_t n) {
  uint64_t res = 0, bit = 1ull << 62;
  while (bit > n) bit >>= 2;
  while (bit) {
    if (n >= res + bit) {
      n -= res + bit;
      res = (res >> 1) + bit;
    } else {
      res >>= 1;
    }
    bit >>= 2;
  }
  return res;
}

// Inverse of a mod n (gcd(a, n) = 1, n < 2^63) by the extended Euclidean
// algorithm on signed 64-bit values.
CLW_HD uint64_t inv_mod64(uint64_t a, uint64_t n) {
  int64_t t = 0, nt = 1;
  uint64_t r = n, nr = a % n;
  while (nr) {
    const uint64_t q = r / nr;
    const int64_t tt = t - (int64_t)q * nt;
    t = nt;
    nt = tt;
    const uint64_t rr = r - q * nr;
    r = nr;
    nr = rr;
  }
  return t < 0 ? (uint64_t)(t + (int64_t)n) : (uint64_t)t;
}

// Two squares: p = a^2 + b^2 for a prime p = 1 (mod 4). Returns the odd one
// with the sign chosen so that a = 1 (mod 4), as find.cl:46-65 normalises it.
CLW_HD int64_t two_squares_a(uint64_t p) {
  const M64 m = m64_make(p);
  const uint64_t nm1 = p - m.one;
  uint64_t x = 0;
  for (uint32_t z = 2; z < 1000; ++z) {
    const uint64_t y = m64_pow(m, m64_small(m, z), (p - 1) / 4);
    if (mmul64(y, y, p, m.ninv) == nm1) {
      x = m64_from(m, y);
      break;
    }
  }
  // Cornacchia: Euclid on (p, x) until the remainder drops below sqrt(p).
  uint64_t r0 = p, r1 = x;
  while ((unsigned __int128)r1 * r1 > p) {
    const uint64_t t = r0 % r1;
    r0 = r1;
    r1 = t;
  }
  const uint64_t a1 = r1;
  const uint64_t b1 = isqrt64(p - a1 * a1);
  const uint64_t odd = (a1 & 1) ? a1 : b1;
  return (odd % 4 == 3) ? -(int64_t)odd : (int64_t)odd;
}

// Eisenstein representation 4p = x^2 + 3y^2 for a prime p = 1 (mod 3), via a
// square root of -3 (2w+1 for a primitive cube root of unity w) and
// Cornacchia's algorithm on (2p, s). Returns c with 4p = c^2 + 27d^2 and
// c = 1 (mod 3) (find.cl:68-88), and u = +-c with u = umod (mod 3) where
// umod = 1 if (p-1)/6 is even else 2 (cl_wilson.cpp:636).
CLW_HD void eisenstein_cu(uint64_t p, int64_t* c_out, int64_t* u_out) {
  const M64 m = m64_make(p);
  uint64_t w = 0;
  for (uint32_t g = 2; g < 1000; ++g) {
    const uint64_t y = m64_pow(m, m64_small(m, g), (p - 1) / 3);
    if (y != m.one) {
      w = m64_from(m, y);
      break;
    }
  }
  uint64_t s = addmod64(addmod64(w, w, p), 1, p);  // s^2 = -3 mod p
  if ((s & 1) == 0) s = p - s;                      // odd, so s^2 = -3 (mod 4p) too
  const uint64_t p2 = 2 * p, p4 = 4 * p;
  uint64_t r0 = p2, r1 = s;
  while ((unsigned __int128)r1 * r1 >= p4) {
    const uint64_t t = r0 % r1;
    r0 = r1;
    r1 = t;
  }
  const int64_t x = (int64_t)r1;
  const int64_t y = (int64_t)isqrt64((p4 - (uint64_t)x * (uint64_t)x) / 3);
  // The three solutions up to sign: (x, y), ((x-3y)/2, (x+y)/2), ((x+3y)/2, (x-y)/2).
  int64_t cu = x, cv = y;
  if (y % 3 != 0) {
    const int64_t u2 = (x - 3 * y) / 2, v2 = (x + y) / 2;
    const int64_t u3 = (x + 3 * y) / 2, v3 = (x - y) / 2;
    if (v2 % 3 == 0) { cu = u2; cv = v2; }
    else { cu = u3; cv = v3; }
  }
  (void)cv;
  int64_t c = cu < 0 ? -cu : cu;
  const int64_t umod = (((p - 1) / 6) % 2 == 0) ? 1 : 2;
  *u_out = (c % 3 == umod) ? c : -c;
  *c_out = (c % 3 == 2) ? -c : c;
}

}  // namespace clw
