// clw_arith.h -- exact integer arithmetic mod p^2 for gfx950 (device side).
//
// Residues mod N = p^2 are held as L limbs of 30 bits (little-endian), value
// sum v[i] * 2^(30 i). Multiplication is Montgomery's REDC with R = 2^(30 L),
// computed column by column ("finely integrated product scanning"): every
// 30x30-bit partial product is one v_mad_u64_u32 into a 64-bit column
// accumulator. 30-bit limbs leave 4 bits of headroom, so a column of up to
// 2L <= 16 partial products never overflows and no carry chains are needed.
//
// Bounds used throughout (N odd, 2N < R):
//   mul_plain(a, q): a < 2N, q < R/2          ->  a q R^-1 mod N, result < 2N
//   mul_full(a, b) : a, b < 2N                ->  a b R^-1 mod N, result < 2N
//                    (one conditional subtraction; 4N^2/R + N < 3N)
//
// The reference keeps residues as two base-p digits with a digit-wise
// Montgomery radix 2^64 (kernels/common.cl:115-260). The value computed for a
// test prime is the same integer mod p^2; only the representation differs.
#pragma once
#include "clw_hd.h"

namespace clw {

constexpr uint32_t LB = 30;
constexpr uint32_t LMASK = (1u << LB) - 1;

typedef unsigned __int128 u128;

template <int L>
struct Res {
  uint32_t v[L];
};

CLW_HD uint64_t mad32(uint32_t a, uint32_t b, uint64_t c) {
  return (uint64_t)a * b + c;
}

template <int L>
CLW_HD Res<L> res_from_u128(u128 x) {
  Res<L> r;
#pragma unroll
  for (int i = 0; i < L; ++i) r.v[i] = (uint32_t)(x >> (LB * i)) & LMASK;
  return r;
}

template <int L>
CLW_HD u128 res_to_u128(const Res<L>& a) {
  u128 x = 0;
#pragma unroll
  for (int i = L - 1; i >= 0; --i) x = (x << LB) | a.v[i];
  return x;
}

template <int L>
CLW_HD Res<L> res_small(uint32_t s) {
  Res<L> r;
#pragma unroll
  for (int i = 0; i < L; ++i) r.v[i] = 0;
  r.v[0] = s;
  return r;
}

// Montgomery product a*b*R^-1 mod n (lazy; see the bounds in the header).
template <int L>
CLW_HD Res<L> montmul(const Res<L>& a, const Res<L>& b, const Res<L>& n, uint32_t ninv) {
  uint32_t m[L];
  Res<L> r;
  uint64_t c = 0;
#pragma unroll
  for (int k = 0; k < 2 * L - 1; ++k) {
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const int j = k - i;
      if (j >= 0 && j < L) c = mad32(a.v[i], b.v[j], c);
    }
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const int j = k - i;
      if (i < k && j >= 0 && j < L) c = mad32(m[i], n.v[j], c);
    }
    if (k < L) {
      m[k] = ((uint32_t)c * ninv) & LMASK;
      c = mad32(m[k], n.v[0], c);
      c >>= LB;
    } else {
      r.v[k - L] = (uint32_t)c & LMASK;
      c >>= LB;
    }
  }
  r.v[L - 1] = (uint32_t)c;
  return r;
}

// Keeps the compiler from reassociating a column sum into several partial
// chains (each extra chain costs a 64-bit add); independent products of
// different test primes are interleaved instead.
#if defined(__HIP_DEVICE_COMPILE__)
#define CLW_PIN1(a) asm volatile("" : "+v"(a))
#define CLW_PIN32(a) asm volatile("" : "+v"(a))
#define CLW_PIN_S(a) asm volatile("" : "+s"(a))
#else
#define CLW_PIN1(a) ((void)0)
#define CLW_PIN32(a) ((void)0)
#define CLW_PIN_S(a) ((void)0)
#endif

template <int K>
CLW_HD void pin_all(uint64_t (&c)[K]) {
#pragma unroll
  for (int j = 0; j < K; ++j) CLW_PIN1(c[j]);
}

// Number of low quotient digits m_0..m_{U-1} that montmul_lockstep leaves
// unmasked (32 bits instead of 30): m_k only has to be right mod 2^30 for the
// division by 2^30 to be exact, and each mask is one instruction. A 32-bit
// m_k makes its products m_k n_j < 2^62; U is the largest count for which
// every column sum (L^2 + L products and the carry) still fits in 64 bits:
// L = 3: column 2 holds 3 a b + 2 (unmasked m n) + m_2 n_0 + carry
// < 3 2^60 + 2 2^62 + 2^60 + 2^34 < 2^64.
template <int L>
constexpr int unmasked_m() { return (L == 3 || L == 4) ? 2 : 1; }

// K independent Montgomery products r[j] = a[j] * b * R^-1 mod n[j] with a
// common (wave-uniform) multiplier b, computed column by column in lockstep:
// consecutive instructions belong to different test primes, so the 64-bit
// accumulator chains never stall on each other and need no extra adds.
// Requires a[j] < 2 n[j] and b < n[j] / 4 (every record factor Q <= T^2 <
// p^2 / 4); then r[j] < 2 n[j] although m may exceed R by ~3 2^(30 U):
// (a b + m n) / R < n/4 + n (1 + 2^-28) < 2 n.
template <int L, int K>
CLW_HD void montmul_lockstep(Res<L> (&r)[K], const Res<L> (&a)[K], const Res<L>& b, const Res<L> (&n)[K],
                             const uint32_t (&ninv)[K]) {
  uint32_t m[K][L];
  uint64_t c[K];
#pragma unroll
  for (int j = 0; j < K; ++j) c[j] = 0;
#pragma unroll
  for (int k = 0; k < 2 * L - 1; ++k) {
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const int jj = k - i;
      if (jj >= 0 && jj < L) {
#pragma unroll
        for (int j = 0; j < K; ++j) c[j] = mad32(a[j].v[i], b.v[jj], c[j]);
        pin_all<K>(c);
      }
    }
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const int jj = k - i;
      if (i < k && jj >= 0 && jj < L) {
#pragma unroll
        for (int j = 0; j < K; ++j) c[j] = mad32(m[j][i], n[j].v[jj], c[j]);
        pin_all<K>(c);
      }
    }
    if (k < L) {
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const uint32_t mk = (uint32_t)c[j] * ninv[j];
        m[j][k] = k < unmasked_m<L>() ? mk : (mk & LMASK);
        c[j] = mad32(m[j][k], n[j].v[0], c[j]) >> LB;
      }
      pin_all<K>(c);
    } else {
#pragma unroll
      for (int j = 0; j < K; ++j) {
        r[j].v[k - L] = (uint32_t)c[j] & LMASK;
        c[j] >>= LB;
      }
      pin_all<K>(c);
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j) r[j].v[L - 1] = (uint32_t)c[j];
}

// Lazy form of the batched kernel's run accumulators. montmul_lazy is
// montmul_lockstep with two instructions fewer per 3-limb product:
//  * the top quotient digit is left unmasked too for 2 and 3 limbs (all
//    digits: m < 4 R (1 + 2^-30)), so the accumulators grow to a < 5 n:
//    (a b + m n) / R < 5 n (n/4) / R + 4 n (1 + 2^-30) < 5 n/8 + 4.01 n;
//    4 limbs keep the top two digits masked (m < R (1 + 2^-28): a < 2 n);
//  * the first output limb keeps all 32 bits of its column: the column sum
//    S = lo + 2^32 hi is the limb lo plus a carry of 4 hi into the next
//    column (one MAD instead of a mask and a 64-bit shift).
// A lazy value is sum v[i] 2^(30 i) with v[0] < 2^32 and v[L-1] < 2^32; it
// is only multiplied by record factors (b < n/4, 30-bit limbs) and brought
// back to the strict form (< 2n, 30-bit limbs) by lazy_norm. Every column
// stays below 2^64 for the band limits (2 p^2 < R, n's top limb below 2^29):
// tests/test_arith_host.py computes the bounds limb by limb.
template <int L>
constexpr int lazy_unmasked_m() { return L <= 3 ? L : 2; }
template <int L>
constexpr uint32_t lazy_bound() { return L <= 3 ? 5 : 2; }  // a, r < lazy_bound * n

template <int L, int K>
CLW_HD void montmul_lazy(Res<L> (&r)[K], const Res<L> (&a)[K], const Res<L>& b, const Res<L> (&n)[K],
                         const uint32_t (&ninv)[K]) {
  uint32_t m[K][L];
  uint64_t c[K];
  uint32_t four = 4;
  CLW_PIN_S(four);  // an opaque 4: the carry stays one MAD (4 hi would become a 64-bit shift of {hi, 0})
#pragma unroll
  for (int j = 0; j < K; ++j) c[j] = 0;
#pragma unroll
  for (int k = 0; k < 2 * L - 1; ++k) {
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const int jj = k - i;
      if (jj >= 0 && jj < L) {
#pragma unroll
        for (int j = 0; j < K; ++j) c[j] = mad32(a[j].v[i], b.v[jj], c[j]);
        pin_all<K>(c);
      }
    }
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const int jj = k - i;
      if (i < k && jj >= 0 && jj < L) {
#pragma unroll
        for (int j = 0; j < K; ++j) c[j] = mad32(m[j][i], n[j].v[jj], c[j]);
        pin_all<K>(c);
      }
    }
    if (k < L) {
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const uint32_t mk = (uint32_t)c[j] * ninv[j];
        m[j][k] = k < lazy_unmasked_m<L>() ? mk : (mk & LMASK);
        c[j] = mad32(m[j][k], n[j].v[0], c[j]) >> LB;
      }
      pin_all<K>(c);
    } else if (k == L) {
#pragma unroll
      for (int j = 0; j < K; ++j) {
        r[j].v[0] = (uint32_t)c[j];
        uint32_t hi = (uint32_t)(c[j] >> 32);
        CLW_PIN32(hi);  // else 4 hi is rebuilt as (c >> 30) & ~3: three instructions
        c[j] = mad32(hi, four, 0);
      }
      pin_all<K>(c);
    } else {
#pragma unroll
      for (int j = 0; j < K; ++j) {
        r[j].v[k - L] = (uint32_t)c[j] & LMASK;
        c[j] >>= LB;
      }
      pin_all<K>(c);
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j) r[j].v[L - 1] = (uint32_t)c[j];
}

// Long Montgomery reduction of a D-digit integer P (30-bit digits, D a
// multiple of L, wave-uniform: the product of a run of sieving primes) for K
// moduli in lockstep: y[j] = P 2^(-30 D) mod n[j], y[j] < 2 n[j] (strict
// form). Column i (weight 2^(30 i)) receives p_i, the carry and m_{i-d} n_d
// for d = 1..L-1, then m_i = c * ninv mod 2^32 zeroes it mod 2^30 with
// m_i n_0; per digit: L MADs, one mul_lo, one 64-bit shift and half a 64-bit
// add (digits enter in pairs). The quotient digits of the last L columns are
// masked, so M = sum m_i 2^(30 i) < 2^(30 D) (1 + 2^-28) and
// y = (P + M n) / 2^(30 D) < 1 + n (1 + 2^-28) < 2 n. Column sums stay below
// 2^35 + 2^60 + (L - 1) 2^62 + 2^62 < 2^64 for L <= 3; L = 4 adds one more
// 2^62 term (n's top limb is below 2^29, so that product is below 2^61).
// Digits per unrolled group of redc_long: a multiple of L (the ring of
// quotient digits) and of 2 (digits enter in pairs).
template <int L>
constexpr int redc_span() { return L % 2 ? 2 * L : L; }

// One group of redc_span digits of redc_long (MASKED: the last group, whose
// last L quotient digits are masked). Digits enter two at a time: column i
// gets p_i + p_{i+1} 2^30 with one 64-bit add (p_{i+1} lands in column i + 1
// after the shift; the low 30 bits that fix m_i are unchanged).
template <int L, int K, bool MASKED>
CLW_HD void redc_group(uint64_t (&c)[K], uint32_t (&m)[K][L], const uint32_t* __restrict__ pg, const Res<L> (&n)[K],
                       const uint32_t (&ninv)[K]) {
  constexpr int G = redc_span<L>();
#pragma unroll
  for (int t = 0; t < G; ++t) {
    if (t % 2 == 0) {
      const uint64_t pr = (uint64_t)pg[t] | ((uint64_t)pg[t + 1] << LB);
#pragma unroll
      for (int j = 0; j < K; ++j) c[j] += pr;
      pin_all<K>(c);
    }
#pragma unroll
    for (int d = 1; d < L; ++d) {
#pragma unroll
      for (int j = 0; j < K; ++j) c[j] = mad32(m[j][((t - d) % L + L) % L], n[j].v[d], c[j]);
      pin_all<K>(c);
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const uint32_t mk = (uint32_t)c[j] * ninv[j];
      m[j][t % L] = (MASKED && t >= G - L) ? (mk & LMASK) : mk;
      c[j] = mad32(m[j][t % L], n[j].v[0], c[j]) >> LB;
    }
    pin_all<K>(c);
  }
}

// D: a multiple of redc_span<L>().
template <int L, int K>
CLW_HD void redc_long(Res<L> (&y)[K], const uint32_t* __restrict__ p, uint32_t D, const Res<L> (&n)[K],
                      const uint32_t (&ninv)[K]) {
  constexpr uint32_t G = redc_span<L>();
  uint32_t m[K][L];
  uint64_t c[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    c[j] = 0;
#pragma unroll
    for (int t = 0; t < L; ++t) m[j][t] = 0;
  }
  const uint32_t groups = D / G;  // >= 1
  for (uint32_t g = 0; g + 1 < groups; ++g) redc_group<L, K, false>(c, m, p + g * G, n, ninv);
  redc_group<L, K, true>(c, m, p + (groups - 1) * G, n, ninv);
  // columns D .. D + L - 1: the remaining m_{D-1-s} n_d, then the carry
#pragma unroll
  for (int k = 0; k < L - 1; ++k) {
#pragma unroll
    for (int d = k + 1; d < L; ++d) {
#pragma unroll
      for (int j = 0; j < K; ++j) c[j] = mad32(m[j][(L + k - d) % L], n[j].v[d], c[j]);
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
      y[j].v[k] = (uint32_t)c[j] & LMASK;
      c[j] >>= LB;
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j) y[j].v[L - 1] = (uint32_t)c[j];
}

// Strict form (< 2n, 30-bit limbs) of a lazy value; same residue mod n.
// 32-bit limb arithmetic only (a 128-bit version raises the kernel's
// register peak, which sits in the run flush).
template <int L>
CLW_HD Res<L> lazy_norm(const Res<L>& x, const Res<L>& n) {
  Res<L> y;
  uint32_t carry = 0;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    const uint32_t t = x.v[i] + carry;  // x.v[i] < 2^32 only for i = 0 and the top limb
    y.v[i] = i < L - 1 ? (t & LMASK) : t;
    carry = i == 0 ? (x.v[0] >> LB) : (t >> LB);
  }
  // y < lazy_bound n with 30-bit limbs below the top one: subtract n while
  // y >= 2 n (at most lazy_bound - 2 times)
#pragma unroll
  for (uint32_t s = 2; s < lazy_bound<L>(); ++s) {
    Res<L> d, d2;
    uint32_t b1 = 0, b2 = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) {  // d = y - n, d2 = y - 2n
      const uint32_t s1 = n.v[i] + b1;
      d.v[i] = i < L - 1 ? ((y.v[i] - s1) & LMASK) : (y.v[i] - s1);
      b1 = y.v[i] < s1;
      const uint32_t s2 = n.v[i] + b2;
      d2.v[i] = i < L - 1 ? ((d.v[i] - s2) & LMASK) : (d.v[i] - s2);
      b2 = d.v[i] < s2;
    }
    if (!b1 && !b2) y = d;  // y >= 2 n
  }
  return y;
}

// r - n if r >= n, else r (limb-wise borrow).
template <int L>
CLW_HD Res<L> cond_sub(const Res<L>& r, const Res<L>& n) {
  Res<L> t;
  int32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    int32_t d = (int32_t)r.v[i] - (int32_t)n.v[i] - borrow;
    borrow = d < 0 ? 1 : 0;
    t.v[i] = (uint32_t)d & LMASK;
  }
  return borrow ? r : t;
}

template <int L>
CLW_HD Res<L> mul_plain(const Res<L>& a, const Res<L>& q, const Res<L>& n, uint32_t ninv) {
  return montmul<L>(a, q, n, ninv);
}

template <int L>
CLW_HD Res<L> mul_full(const Res<L>& a, const Res<L>& b, const Res<L>& n, uint32_t ninv) {
  return cond_sub<L>(montmul<L>(a, b, n, ninv), n);
}

// a + b mod n for canonical a, b < n.
template <int L>
CLW_HD Res<L> add_mod(const Res<L>& a, const Res<L>& b, const Res<L>& n) {
  Res<L> s;
  uint32_t carry = 0;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    uint32_t t = a.v[i] + b.v[i] + carry;
    s.v[i] = t & LMASK;
    carry = t >> LB;
  }
  return cond_sub<L>(s, n);
}

// a - b mod n for canonical a, b < n.
template <int L>
CLW_HD Res<L> sub_mod(const Res<L>& a, const Res<L>& b, const Res<L>& n) {
  Res<L> t;
  int32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    int32_t d = (int32_t)a.v[i] - (int32_t)b.v[i] - borrow;
    borrow = d < 0 ? 1 : 0;
    t.v[i] = (uint32_t)d & LMASK;
  }
  if (borrow) {
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      uint32_t s = t.v[i] + n.v[i] + carry;
      t.v[i] = s & LMASK;
      carry = s >> LB;
    }
  }
  return t;
}

// Per-test-prime Montgomery context.
template <int L>
struct Ctx {
  Res<L> n;       // N = p^2
  uint32_t ninv;  // -N^-1 mod 2^30
  Res<L> one;     // R mod N      (Montgomery form of 1)
  Res<L> r2;      // R^2 mod N    (Montgomery form of R)
};

// -n0^-1 mod 2^30 for odd n0 (Newton iteration on 32-bit words).
CLW_HD uint32_t neg_inv30(uint32_t n0) {
  uint32_t inv = n0;  // correct to 3 bits for odd n0
#pragma unroll
  for (int i = 0; i < 4; ++i) inv *= 2u - n0 * inv;
  return (0u - inv) & LMASK;
}

// 2^k mod m for m < 2^126 by doubling (used once per test prime).
CLW_HD u128 pow2_mod(uint32_t k, u128 m) {
  u128 x = 1 % m;
  for (uint32_t i = 0; i < k; ++i) {
    x <<= 1;
    if (x >= m) x -= m;
  }
  return x;
}

// Montgomery form of R^e: returns R^(e+1) mod N (lazy < 2N). e is any 64-bit
// exponent; the loop branches on e's bits (uniform when e is uniform).
template <int L>
CLW_HD Res<L> pow_r(const Ctx<L>& c, uint64_t e) {
  if (e == 0) return c.one;
  Res<L> r = c.r2;
  const int top = 63 - __builtin_clzll(e);
  for (int b = top - 1; b >= 0; --b) {
    r = mul_full<L>(r, r, c.n, c.ninv);
    if ((e >> b) & 1) r = mul_full<L>(r, c.r2, c.n, c.ninv);
  }
  return r;
}

// Montgomery power of a value already in Montgomery form (x R): (x^e) R.
template <int L>
CLW_HD Res<L> mont_pow(const Ctx<L>& c, const Res<L>& xm, uint64_t e) {
  if (e == 0) return c.one;
  Res<L> r = xm;
  const int top = 63 - __builtin_clzll(e);
  for (int b = top - 1; b >= 0; --b) {
    r = mul_full<L>(r, r, c.n, c.ninv);
    if ((e >> b) & 1) r = mul_full<L>(r, xm, c.n, c.ninv);
  }
  return r;
}

// Drift bookkeeping: a value "a with drift d" stands for a * R^d mod N.
// Returns the canonical value (< N) of a * R^d.
template <int L>
CLW_HD Res<L> undrift(const Ctx<L>& c, const Res<L>& a, uint64_t d) {
  const Res<L> w = pow_r<L>(c, d);  // R^(d+1)
  Res<L> r = mul_full<L>(a, w, c.n, c.ninv);
  return cond_sub<L>(r, c.n);
}

}  // namespace clw
