// clw_hostcheck.cpp -- the device arithmetic headers compiled for the host
// CPU, exported for unit tests (tests/test_arith_host.py). This library is a
// test harness for clw_arith.h / clw_nt64.h; the product runs them on the GPU.
#include <string.h>
#include "clw_arith.h"
#include "clw_nt64.h"

using namespace clw;

template <int L>
static void montmul_t(const uint64_t a[2], const uint64_t b[2], const uint64_t n[2], uint64_t out[2]) {
  const u128 A = ((u128)a[1] << 64) | a[0], B = ((u128)b[1] << 64) | b[0], Nv = ((u128)n[1] << 64) | n[0];
  const Res<L> ra = res_from_u128<L>(A), rb = res_from_u128<L>(B), rn = res_from_u128<L>(Nv);
  const uint32_t ninv = neg_inv30(rn.v[0]);
  const Res<L> r = montmul<L>(ra, rb, rn, ninv);
  const u128 R = res_to_u128<L>(r);
  out[0] = (uint64_t)R;
  out[1] = (uint64_t)(R >> 64);
}

template <int L>
static void montmul_lockstep_t(const uint64_t a[2], const uint64_t b[2], const uint64_t n[2], uint64_t out[2]) {
  const u128 A = ((u128)a[1] << 64) | a[0], B = ((u128)b[1] << 64) | b[0], Nv = ((u128)n[1] << 64) | n[0];
  Res<L> ra[1] = {res_from_u128<L>(A)}, rn[1] = {res_from_u128<L>(Nv)}, r[1];
  const uint32_t ninv[1] = {neg_inv30(rn[0].v[0])};
  montmul_lockstep<L, 1>(r, ra, res_from_u128<L>(B), rn, ninv);
  const u128 R = res_to_u128<L>(r[0]);
  out[0] = (uint64_t)R;
  out[1] = (uint64_t)(R >> 64);
}

// Lazy product on raw limbs (the lazy form is not a plain 128-bit value:
// limb 0 may hold 32 bits).
template <int L>
static void montmul_lazy_t(const uint32_t* a, const uint64_t b[2], const uint64_t n[2], uint32_t* out) {
  const u128 B = ((u128)b[1] << 64) | b[0], Nv = ((u128)n[1] << 64) | n[0];
  Res<L> ra[1], rn[1] = {res_from_u128<L>(Nv)}, r[1];
  for (int i = 0; i < L; ++i) ra[0].v[i] = a[i];
  const uint32_t ninv[1] = {neg_inv30(rn[0].v[0])};
  montmul_lazy<L, 1>(r, ra, res_from_u128<L>(B), rn, ninv);
  for (int i = 0; i < L; ++i) out[i] = r[0].v[i];
}

template <int L>
static void redc_long_t(const uint32_t* pd, uint32_t D, const uint64_t n[2], uint64_t out[2]) {
  const u128 Nv = ((u128)n[1] << 64) | n[0];
  Res<L> rn[1] = {res_from_u128<L>(Nv)}, y[1];
  const uint32_t ninv[1] = {neg_inv30(rn[0].v[0])};
  redc_long<L, 1>(y, pd, D, rn, ninv);
  for (int i = 0; i < L; ++i)
    if (y[0].v[i] >> LB) {  // limbs must be 30-bit
      out[0] = out[1] = ~0ull;
      return;
    }
  const u128 R = res_to_u128<L>(y[0]);
  out[0] = (uint64_t)R;
  out[1] = (uint64_t)(R >> 64);
}

template <int L>
static void lazy_norm_t(const uint32_t* x, const uint64_t n[2], uint32_t* out) {
  const u128 Nv = ((u128)n[1] << 64) | n[0];
  Res<L> rx;
  for (int i = 0; i < L; ++i) rx.v[i] = x[i];
  const Res<L> y = lazy_norm<L>(rx, res_from_u128<L>(Nv));
  for (int i = 0; i < L; ++i) out[i] = y.v[i];
}

template <int L>
static Ctx<L> make_ctx(uint64_t p) {
  const u128 N = (u128)p * p;
  Ctx<L> c;
  c.n = res_from_u128<L>(N);
  c.ninv = neg_inv30(c.n.v[0]);
  const u128 one = pow2_mod(30 * L, N);
  u128 r2 = one;
  for (int i = 0; i < 30 * L; ++i) { r2 <<= 1; if (r2 >= N) r2 -= N; }
  c.one = res_from_u128<L>(one);
  c.r2 = res_from_u128<L>(r2);
  return c;
}

template <int L>
static void undrift_t(uint64_t p, const uint64_t a[2], uint64_t d, uint64_t out[2]) {
  const Ctx<L> c = make_ctx<L>(p);
  const u128 A = ((u128)a[1] << 64) | a[0];
  const Res<L> r = undrift<L>(c, res_from_u128<L>(A), d);
  const u128 R = res_to_u128<L>(r);
  out[0] = (uint64_t)R;
  out[1] = (uint64_t)(R >> 64);
}

extern "C" {

int clwh_montmul(int L, const uint64_t* a, const uint64_t* b, const uint64_t* n, uint64_t* out) {
  switch (L) {
    case 2: montmul_t<2>(a, b, n, out); return 0;
    case 3: montmul_t<3>(a, b, n, out); return 0;
    case 4: montmul_t<4>(a, b, n, out); return 0;
    case 5: montmul_t<5>(a, b, n, out); return 0;
  }
  return -1;
}

int clwh_montmul_lockstep(int L, const uint64_t* a, const uint64_t* b, const uint64_t* n, uint64_t* out) {
  switch (L) {
    case 2: montmul_lockstep_t<2>(a, b, n, out); return 0;
    case 3: montmul_lockstep_t<3>(a, b, n, out); return 0;
    case 4: montmul_lockstep_t<4>(a, b, n, out); return 0;
    case 5: montmul_lockstep_t<5>(a, b, n, out); return 0;
  }
  return -1;
}

int clwh_montmul_lazy(int L, const uint32_t* a, const uint64_t* b, const uint64_t* n, uint32_t* out) {
  switch (L) {
    case 2: montmul_lazy_t<2>(a, b, n, out); return 0;
    case 3: montmul_lazy_t<3>(a, b, n, out); return 0;
    case 4: montmul_lazy_t<4>(a, b, n, out); return 0;
  }
  return -1;
}

int clwh_redc_long(int L, const uint32_t* p, uint32_t D, const uint64_t* n, uint64_t* out) {
  if (D == 0 || D % (uint32_t)(L % 2 ? 2 * L : L)) return -1;
  switch (L) {
    case 2: redc_long_t<2>(p, D, n, out); return 0;
    case 3: redc_long_t<3>(p, D, n, out); return 0;
    case 4: redc_long_t<4>(p, D, n, out); return 0;
  }
  return -1;
}

int clwh_lazy_norm(int L, const uint32_t* x, const uint64_t* n, uint32_t* out) {
  switch (L) {
    case 2: lazy_norm_t<2>(x, n, out); return 0;
    case 3: lazy_norm_t<3>(x, n, out); return 0;
    case 4: lazy_norm_t<4>(x, n, out); return 0;
  }
  return -1;
}

// (unmasked quotient digits, bound B of the lazy form a < B n)
int clwh_lazy_params(int L, int* bound) {
  switch (L) {
    case 2: *bound = (int)lazy_bound<2>(); return lazy_unmasked_m<2>();
    case 3: *bound = (int)lazy_bound<3>(); return lazy_unmasked_m<3>();
    case 4: *bound = (int)lazy_bound<4>(); return lazy_unmasked_m<4>();
  }
  return -1;
}

int clwh_unmasked_m(int L) {
  switch (L) {
    case 2: return unmasked_m<2>();
    case 3: return unmasked_m<3>();
    case 4: return unmasked_m<4>();
    case 5: return unmasked_m<5>();
  }
  return -1;
}

int clwh_undrift(int L, uint64_t p, const uint64_t* a, uint64_t d, uint64_t* out) {
  switch (L) {
    case 2: undrift_t<2>(p, a, d, out); return 0;
    case 3: undrift_t<3>(p, a, d, out); return 0;
    case 4: undrift_t<4>(p, a, d, out); return 0;
    case 5: undrift_t<5>(p, a, d, out); return 0;
  }
  return -1;
}

uint32_t clwh_neg_inv30(uint32_t n0) { return neg_inv30(n0); }
uint64_t clwh_inv64(uint64_t n) { return inv64(n); }
uint64_t clwh_legendre(uint64_t q, uint64_t T) { return legendre_exp(q, T); }
uint64_t clwh_isqrt(uint64_t n) { return isqrt64(n); }
uint64_t clwh_inv_mod(uint64_t a, uint64_t n) { return inv_mod64(a, n); }
int clwh_mr_bases_needed(uint64_t n) { return mr_bases_needed(n); }

// Full deterministic primality test as the sieve kernel applies it to a
// candidate n > 37 that survived trial division (base 2, then extra bases).
int clwh_is_prime_mr(uint64_t n, int nbases) {
  if (n < 2) return 0;
  if ((n & 1) == 0) return n == 2;
  const uint32_t small[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (int i = 0; i < 12; ++i) {
    if (n == small[i]) return 1;
    if (n % small[i] == 0) return 0;
  }
  const M64 m = m64_make(n);
  const int s = __builtin_ctzll(n - 1);
  const uint64_t d = (n - 1) >> s;
  if (!sprp2(m, d, s)) return 0;
  return mr_extra(m, d, s, nbases) ? 1 : 0;
}

int64_t clwh_two_squares_a(uint64_t p) { return two_squares_a(p); }
void clwh_eisenstein_cu(uint64_t p, int64_t* c, int64_t* u) { eisenstein_cu(p, c, u); }

}  // extern "C"
