// cl_wilson.cpp -- host orchestration of one Wilson-prime range on the GPU
// (the role of cl_wilson() in the reference, cl_wilson.cpp:1442-1974).
//
// Steps: exact test primes of [pmin, pmax) on the host; typing and type
// targets; per-prime constants on the device; a loop over segments of the
// sieving-prime axis [7, maxtarget] (sieve + products + reduction per type);
// the iterate step to each prime's own target; the batched final check; the
// checksum and result classification on the host.
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <mutex>
#include <vector>

#include "clw_internal.h"
#include "clw_nt64.h"

namespace {

const uint64_t kMaxP = 0xFFFFFFFFFFFFFFFFull / 4;  // cl_wilson.h:9
const uint64_t kMaxWidth = 10000000;               // cl_wilson.cpp:1162-1166
const uint64_t kDefaultSpan = 4370ull * CLW_BLOCK_SPAN;
const uint32_t kSideCap = 1u << 22;
const uint32_t kMaxKRec = 32, kMaxKSide = 8, kMaxKLanes = 8192;
const uint32_t kLanesModeBelow = 8192;              // test primes per type
const uint64_t kLanesSpan = 16ull * kDefaultSpan;   // ~4.3e9 integers per segment
const uint32_t kPlainGroups = 1024;                 // work-groups per plain launch

// Odd primes below `lim` (simple sieve; lim is at most a few million).
std::vector<uint32_t> small_odd_primes(uint32_t lim);

// Odd primes up to at least `lim`, computed once per process and grown on
// demand (the base of every host segmented sieve).
std::shared_ptr<const std::vector<uint32_t>> base_primes(uint32_t lim) {
  static std::mutex mu;
  static std::shared_ptr<const std::vector<uint32_t>> cache;
  static uint32_t cache_lim = 0;
  std::lock_guard<std::mutex> g(mu);
  if (!cache || cache_lim < lim) {
    const uint32_t want = std::max(lim, 1u << 22);
    cache = std::make_shared<const std::vector<uint32_t>>(small_odd_primes(want));
    cache_lim = want;
  }
  return cache;
}

std::vector<uint32_t> small_odd_primes(uint32_t lim) {
  std::vector<uint8_t> comp(lim + 1, 0);
  std::vector<uint32_t> out;
  for (uint64_t i = 3; i <= lim; i += 2) {
    if (comp[i]) continue;
    out.push_back((uint32_t)i);
    for (uint64_t j = i * i; j <= lim; j += 2 * i) comp[j] = 1;
  }
  return out;
}

bool is_prime_det(uint64_t n) {
  if (n < 2) return false;
  static const uint32_t sp[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (uint32_t q : sp) {
    if (n == q) return true;
    if (n % q == 0) return false;
  }
  const clw::M64 m = clw::m64_make(n);
  const int s = __builtin_ctzll(n - 1);
  const uint64_t d = (n - 1) >> s;
  if (!clw::sprp2(m, d, s)) return false;
  return clw::mr_extra(m, d, s, clw::mr_bases_needed(n));
}

int64_t signed_w(uint64_t p, uint64_t quot) {
  const uint64_t neg = p - quot;
  return quot <= neg ? (int64_t)quot : -(int64_t)neg;
}

}  // namespace

// Segmented sieve over [lo, hi) of the odd numbers, then (when it stops
// short of sqrt(hi)) deterministic Miller-Rabin on what is left. Replaces
// primesieve_generate_primes (cl_wilson.cpp:1529). How far to sieve: a base
// prime q costs one 64-bit division (its first multiple in the range) and
// takes out about 1/q of the candidates, and a candidate left costs a
// Miller-Rabin test (about 50 divisions' worth); a narrow range far out
// (1e4 wide at 1e14) is best sieved by the primes below 2^16 and tested, a
// wide one sieved to sqrt(hi) (base primes up to 2^25 are kept: complete
// below 1.1e15). Each base prime's next multiple is carried from window to
// window, so its division is paid once per range.
extern "C" uint64_t clw_host_primes(uint64_t lo, uint64_t hi, uint64_t* out, uint64_t cap) {
  if (hi <= lo) return 0;
  uint64_t cnt = 0;
  auto emit = [&](uint64_t v) {
    if (out && cnt < cap) out[cnt] = v;
    ++cnt;
  };
  if (lo <= 2 && hi > 2) emit(2);
  const uint64_t s = lo < 3 ? 3 : (lo | 1);
  if (s >= hi) return cnt;
  const uint64_t nodd = (hi - s + 1) / 2;  // candidates s, s + 2, ..., below hi
  const uint64_t r = (uint64_t)clw::isqrt64(hi - 1);  // a composite below hi has a prime factor <= r
  const uint64_t top = std::min<uint64_t>(r, 1u << 25);  // the furthest the kept base primes reach
  uint64_t lim = top;
  {
    double best = -1;
    for (uint64_t q = 1u << 16;; q <<= 2) {
      const uint64_t qq = std::min(q, top);
      const double lq = std::log((double)std::max<uint64_t>(qq, 16));
      double cost = (double)qq / lq;                          // divisions
      if (qq < r) cost += 50.0 * (double)nodd * 1.1229 / lq;  // Miller-Rabin on the survivors
      if (best < 0 || cost < best) {
        best = cost;
        lim = qq;
      }
      if (qq == top) break;
    }
  }
  const bool need_mr = lim < r;
  const auto base_ptr = base_primes((uint32_t)lim);
  const std::vector<uint32_t>& base = *base_ptr;
  size_t nb = 0;  // base primes in use: q <= lim
  while (nb < base.size() && base[nb] <= lim) ++nb;
  // next[j]: odd-candidate index of base[j]'s next multiple to strike (the
  // first at or above max(q^2, s))
  std::vector<uint64_t> next(nb);
  for (size_t j = 0; j < nb; ++j) {
    const uint64_t q = base[j];
    uint64_t st = q * q;
    if (st < s) {
      st = (s + q - 1) / q * q;
      if ((st & 1) == 0) st += q;
    }
    next[j] = (st - s) / 2;
  }
  const uint64_t SEG = 1u << 18;  // odd candidates per window (256 KB of flags)
  std::vector<uint8_t> mark(SEG);
  for (uint64_t w0 = 0; w0 < nodd; w0 += SEG) {
    const uint64_t w1 = std::min(w0 + SEG, nodd);
    std::fill(mark.begin(), mark.begin() + (w1 - w0), 0);
    for (size_t j = 0; j < nb; ++j) {
      uint64_t k = next[j];
      if (k >= w1) continue;
      const uint64_t q = base[j];
      for (; k < w1; k += q) mark[k - w0] = 1;
      next[j] = k;
    }
    for (uint64_t k = w0; k < w1; ++k) {
      if (mark[k - w0]) continue;
      const uint64_t v = s + 2 * k;
      if (need_mr && !is_prime_det(v)) continue;
      emit(v);
    }
  }
  return cnt;
}

// cl_wilson.cpp:1552-1586: type 0 if p = 1 (mod 3) with target (p-1)/6,
// type 1 if p = 5 (mod 12) with (p-1)/4, type 2 if p = 11 (mod 12) with
// (p-1)/2; the type target is the target of the type's first prime.
static inline uint32_t prime_type(uint64_t p) { return p % 3 == 1 ? 0 : (p % 12 == 5 ? 1 : 2); }
static inline uint64_t type_divisor(uint32_t t) { return t == 0 ? 6 : (t == 1 ? 4 : 2); }

extern "C" int clw_range_types(uint64_t pmin, uint64_t pmax, uint32_t tpcnt[3], uint64_t targets[3]) {
  if (!tpcnt || !targets || pmin < 5 || pmax <= pmin) return clw_fail(CLW_ERR_ARG, "clw_range_types: bad range");
  std::vector<uint64_t> ps(clw_host_primes(pmin, pmax, nullptr, 0));
  clw_host_primes(pmin, pmax, ps.data(), ps.size());
  for (int t = 0; t < 3; ++t) tpcnt[t] = 0, targets[t] = 0;
  for (uint64_t p : ps) {
    const uint32_t t = prime_type(p);
    if (!tpcnt[t]++) targets[t] = (p - 1) / type_divisor(t);
  }
  return CLW_OK;
}

// cl_wilson.cpp:657-688 (result lines) and 872-898 (classification).
extern "C" int clw_format_result(uint64_t p, uint64_t quot, char* buf, size_t buflen) {
  const uint64_t neg = p - quot, smallest = quot <= neg ? quot : neg;
  if (quot == 0) {
    if (buf) snprintf(buf, buflen, "%llu is a Wilson prime\n", (unsigned long long)p);
    return 1;
  }
  if (smallest < 1000) {
    const int32_t dq = (smallest == quot) ? (int32_t)quot : -((int32_t)neg);
    if (buf) snprintf(buf, buflen, "%llu is a Near-Wilson prime %+d\n", (unsigned long long)p, dq);
    return 1;
  }
  if (buf && buflen) buf[0] = 0;
  return 0;
}

// Result verification against a known-good list (cl_wilson.cpp:902-930,
// 936-1014, 1073-1087).
extern "C" int clw_verify_results(const char* path, uint64_t expected_sum, uint64_t pmin, uint64_t pmax,
                                  const uint64_t* p, const uint64_t* quot, uint64_t n, uint32_t* matched,
                                  uint32_t* listed) {
  if (!path || (n && (!p || !quot))) return clw_fail(CLW_ERR_ARG, "clw_verify_results: null argument");
  if (matched) *matched = 0;
  if (listed) *listed = 0;
  FILE* f = fopen(path, "r");
  if (!f) return clw_fail(CLW_ERR_IO, "Cannot open %s !!!", path);
  std::vector<std::pair<uint64_t, int32_t>> good;  // the entries in [pmin, pmax)
  uint64_t sum = 0;
  char line[256];
  while (fgets(line, sizeof(line), f)) {
    unsigned long long gp = 0;
    int gv = 0;
    if (sscanf(line, "%llu %d", &gp, &gv) != 2) {
      fclose(f);
      return clw_fail(CLW_ERR_IO, "Error reading %s !!!", path);
    }
    sum += (uint64_t)gp + (uint64_t)(int64_t)gv;  // uint64 + int32, as cl_wilson.cpp:954
    if (gp >= pmin && gp < pmax) good.push_back({(uint64_t)gp, (int32_t)gv});
  }
  fclose(f);
  if (sum != expected_sum) return clw_fail(CLW_ERR_CHECK, "Checksum error in %s !!!", path);
  if (listed) *listed = (uint32_t)good.size();
  uint32_t nm = 0;
  const double threshold = 1.0 / 50000.0;
  for (uint64_t k = 0; k < n; ++k) {
    const uint64_t neg = p[k] - quot[k], smallest = quot[k] <= neg ? quot[k] : neg;
    // the ratio rounded to double before the compare, as cl_wilson.cpp:906-907
    const double wpp = (double)((long double)smallest / (long double)p[k]);
    if (!(wpp < threshold)) continue;
    const int32_t dq = smallest == quot[k] ? (int32_t)quot[k] : -(int32_t)neg;
    bool found = false;
    for (auto& e : good) {
      if (e.first != p[k]) continue;
      if (e.second != dq)
        return clw_fail(CLW_ERR_CHECK, "result check failed! p: %llu %d, result file is %d",
                        (unsigned long long)p[k], dq, e.second);
      e.first = 0;  // matched once
      found = true;
      ++nm;
      break;
    }
    if (!found)
      return clw_fail(CLW_ERR_CHECK, "result check failed! p: %llu is not in result file!", (unsigned long long)p[k]);
  }
  if (matched) *matched = nm;
  for (const auto& e : good)
    if (e.first)
      return clw_fail(CLW_ERR_CHECK, "result check failed! p: %llu was not found in results!",
                      (unsigned long long)e.first);
  return CLW_OK;
}

namespace {

#define CK(x)                        \
  do {                               \
    const int rc_ = (x);             \
    if (rc_ != CLW_OK) return rc_;   \
  } while (0)

}  // namespace

// A range being searched: host test primes, the device table, segment
// buffers and partial-product scratch (the state cl_wilson() keeps in
// progData/searchData, cl_wilson.h:27-70).
struct clw_session {
  clw_ctx* ctx = nullptr;
  clw_options opt{};
  uint64_t pmin = 0, pmax = 0;
  std::vector<uint64_t> tp;   // increasing
  std::vector<uint32_t> pos;  // pos[k]: index of tp[k] in the type-grouped table
  uint32_t cnt[3] = {0, 0, 0}, off[3] = {0, 0, 0};
  uint64_t T[3] = {0, 0, 0}, maxT = 0;
  uint32_t stride = 0, target_waves = 0;
  bool lanes_mode = false;
  uint64_t span = 0;
  clw_tp_table tab{};
  clw_seg_buffers seg{};
  uint64_t* part_lo = nullptr;
  uint64_t* part_hi = nullptr;
  // run products of a segment's sub-chunks (lane-per-test-prime mode, 2-4
  // limbs): digits with stride ds per (block, sub-chunk), their counts
  uint32_t* subp = nullptr;
  uint32_t* subn = nullptr;
  uint32_t ds = 0;
  uint64_t* pprod = nullptr;  // device counter: partial products of the run-product kernel
  uint64_t total_primes = 0;  // sieving primes applied so far
  bool iterated = false;
  double gpu_seconds = 0;
  // Events around the launches of one segment: sieve, then per type
  // (records, side, reduce).
  hipEvent_t ev[2 + 3 * 4] = {};
  bool ev_ok = false;
  clw_kernel_stats ks{};

  ~clw_session() {
    if (ev_ok)
      for (hipEvent_t e : ev) (void)hipEventDestroy(e);  // teardown: nothing to report
    clw_tp_table_free(ctx, &tab);
    clw_seg_free(ctx, &seg);
    clw_free(ctx, part_lo);
    clw_free(ctx, part_hi);
    clw_free(ctx, subp);
    clw_free(ctx, subn);
    clw_free(ctx, pprod);
  }
};

extern "C" int clw_session_kernel_stats(const clw_session* S, clw_kernel_stats* out) {
  if (!S || !out) return clw_fail(CLW_ERR_ARG, "clw_session_kernel_stats: null argument");
  *out = S->ks;
  return CLW_OK;
}

extern "C" int clw_session_reset_stats(clw_session* S) {
  if (!S) return clw_fail(CLW_ERR_ARG, "clw_session_reset_stats: null session");
  memset(&S->ks, 0, sizeof(S->ks));
  return CLW_OK;
}

static float ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms : 0.0f;
}

extern "C" int clw_session_create(clw_ctx* ctx, uint64_t pmin, uint64_t pmax, const clw_options* opt,
                                  clw_session** out) {
  if (!ctx || !out) return clw_fail(CLW_ERR_ARG, "clw_session_create: null argument");
  *out = nullptr;
  // cl_wilson.cpp:1146-1182 setupSearch
  if (pmin < 5 || pmax > kMaxP || pmin >= pmax)
    return clw_fail(CLW_ERR_ARG, "range [%llu, %llu) invalid: need 5 <= pmin < pmax <= %llu",
                    (unsigned long long)pmin, (unsigned long long)pmax, (unsigned long long)kMaxP);
  if (pmax - pmin > kMaxWidth)
    return clw_fail(CLW_ERR_ARG, "range width %llu > 10000000", (unsigned long long)(pmax - pmin));
  clw_session* S = new clw_session();
  S->ctx = ctx;
  S->opt = opt ? *opt : clw_options{};
  S->pmin = pmin;
  S->pmax = pmax;
  {
    std::vector<uint64_t> v((pmax - pmin) / 2 + 4);
    uint64_t n = 0;
    const int rc = clw_device_primes(ctx, pmin, pmax, v.data(), v.size(), &n);
    if (rc != CLW_OK) {
      delete S;
      return rc;
    }
    v.resize(n);
    S->tp.swap(v);
  }
  if (S->tp.empty()) {
    delete S;
    return clw_fail(CLW_ERR_ARG, "there are no primes to test in this range");
  }
  const uint32_t ntp = (uint32_t)S->tp.size();
  for (uint64_t p : S->tp) {
    const uint32_t t = prime_type(p);
    if (!S->cnt[t]++) S->T[t] = (p - 1) / type_divisor(t);
  }
  S->off[1] = S->cnt[0];
  S->off[2] = S->cnt[0] + S->cnt[1];
  std::vector<uint64_t> gp(ntp);
  std::vector<uint32_t> gt(ntp);
  S->pos.resize(ntp);
  {
    uint32_t fill[3] = {S->off[0], S->off[1], S->off[2]};
    for (uint32_t k = 0; k < ntp; ++k) {
      const uint32_t t = prime_type(S->tp[k]);
      S->pos[k] = fill[t]++;
      gp[S->pos[k]] = S->tp[k];
      gt[S->pos[k]] = t;
    }
  }
  S->maxT = std::max(S->T[0], std::max(S->T[1], S->T[2]));
  S->stride = std::max(S->cnt[0], std::max(S->cnt[1], S->cnt[2]));
  // Few test primes: a wave per prime (clw_mul_records_lanes) and longer
  // segments; many: a lane per prime (clw_mul_records).
  S->lanes_mode = S->opt.mode == 2 || (S->opt.mode == 0 && S->stride < kLanesModeBelow);
  S->span = S->opt.segment_span
                ? (S->opt.segment_span + CLW_BLOCK_SPAN - 1) / CLW_BLOCK_SPAN * CLW_BLOCK_SPAN
                : (S->lanes_mode ? kLanesSpan : kDefaultSpan);
  S->target_waves = S->opt.target_waves ? S->opt.target_waves : (uint32_t)ctx->prop.multiProcessorCount * 4 * 8;
  int rc = clw_tp_table_alloc(ctx, ntp, (uint32_t)clw_limbs_for(S->tp.back()), &S->tab);
  if (rc == CLW_OK) rc = clw_h2d(ctx, S->tab.p, gp.data(), (size_t)ntp * 8, 1);
  if (rc == CLW_OK) rc = clw_h2d(ctx, S->tab.type, gt.data(), (size_t)ntp * 4, 1);
  if (rc == CLW_OK) rc = clw_setup(ctx, &S->tab, 0);  // setup.cl
  if (rc == CLW_OK) rc = clw_seg_alloc(ctx, (uint32_t)(S->span / CLW_BLOCK_SPAN + 1), kSideCap, &S->seg);
  // Chunk products of type t live at part[c * cnt[t] + i]: at most
  // (kMaxKRec + kMaxKSide) * cnt (lane per prime) or target_waves + 2 cnt.
  const size_t part_entries = std::max<size_t>(
      std::max<size_t>((size_t)(kMaxKRec + kMaxKSide) * S->stride, (size_t)S->target_waves + 2 * (size_t)S->stride),
      kPlainGroups);
  if (rc == CLW_OK) rc = clw_alloc(ctx, part_entries * 8, (void**)&S->part_lo);
  if (rc == CLW_OK) rc = clw_alloc(ctx, part_entries * 8, (void**)&S->part_hi);
  if (rc == CLW_OK && !S->lanes_mode && S->opt.mode != 3 && S->tab.limbs <= 4) {
    // 2 kBatchSubRecs sieving primes up to maxT: at most 2 kBatchSubRecs
    // (bits(maxT) + 1) bits, in 30-bit digits rounded up to 12
    const uint32_t bits = (S->maxT ? 64 - __builtin_clzll(S->maxT) : 1) + 1;
    const uint32_t digits = (2 * kBatchSubRecs * bits + 29) / 30;
    S->ds = std::min<uint32_t>(kBatchSubMaxDigits, std::max<uint32_t>(12, (digits + 11) / 12 * 12));
    const size_t slots = (size_t)S->seg.max_blocks * kBatchSubCap;
    rc = clw_alloc(ctx, slots * S->ds * 4, (void**)&S->subp);
    if (rc == CLW_OK) rc = clw_alloc(ctx, slots * 4, (void**)&S->subn);
    if (rc == CLW_OK) rc = clw_alloc(ctx, 8, (void**)&S->pprod);
    if (rc == CLW_OK) rc = clw_memset(ctx, S->pprod, 0, 8);
  }
  if (rc == CLW_OK) rc = clw_finish(ctx);
  if (rc == CLW_OK) {
    S->ev_ok = true;
    for (hipEvent_t& e : S->ev)
      if (hipEventCreate(&e) != hipSuccess) S->ev_ok = false;
    if (!S->ev_ok) rc = clw_fail(CLW_ERR_HIP, "hipEventCreate failed");
  }
  if (rc != CLW_OK) {
    delete S;
    return rc;
  }
  *out = S;
  return CLW_OK;
}

extern "C" int clw_session_destroy(clw_session* S) {
  delete S;
  return CLW_OK;
}

extern "C" int clw_session_info(const clw_session* S, clw_range_stats* st) {
  if (!S || !st) return clw_fail(CLW_ERR_ARG, "clw_session_info: null argument");
  memset(st, 0, sizeof(*st));
  st->pmin = S->pmin;
  st->pmax = S->pmax;
  st->tpcount = (uint32_t)S->tp.size();
  for (int t = 0; t < 3; ++t) {
    st->tpcnt[t] = S->cnt[t];
    st->typeTarget[t] = S->T[t];
  }
  st->maxtarget = S->maxT;
  st->totalcount = S->total_primes;
  st->limbs = S->tab.limbs;
  st->seconds = S->gpu_seconds;
  return CLW_OK;
}

// Applies every sieving prime q in [qlo, qhi) (clipped to [2, maxtarget]) to
// the residues: the segment loop of cl_wilson.cpp:1856-1929. Windows must not
// overlap between calls; their union over a search is [2, maxtarget].
extern "C" int clw_session_run(clw_session* S, uint64_t qlo, uint64_t qhi) {
  if (!S) return clw_fail(CLW_ERR_ARG, "clw_session_run: null session");
  if (S->iterated) return clw_fail(CLW_ERR_STATE, "clw_session_run after iterate");
  const auto t0 = std::chrono::steady_clock::now();
  clw_ctx* ctx = S->ctx;
  if (qhi > S->maxT + 1) qhi = S->maxT + 1;
  // Primes 2, 3, 5 are outside the mod-30 wheel: their powers go to the side
  // lists of the window that contains them.
  std::vector<uint64_t> small_side[3];
  for (uint32_t t = 0; t < 3; ++t)
    for (uint64_t q : {2ull, 3ull, 5ull})
      if (S->cnt[t] && q >= qlo && q < qhi && q <= S->T[t]) {
        small_side[t].push_back(q);
        small_side[t].push_back(clw::legendre_exp(q, S->T[t]));
      }
  for (uint64_t q : {2ull, 3ull, 5ull}) S->total_primes += (q >= qlo && q < qhi && q <= S->maxT);
  const bool have_small = !small_side[0].empty() || !small_side[1].empty() || !small_side[2].empty();
  uint64_t lo = std::max<uint64_t>(qlo, 7);
  bool first = true;
  auto last_report = std::chrono::steady_clock::now();
  while ((first && have_small) || lo < qhi) {
    const uint64_t hi = std::min<uint64_t>(lo + S->span - (lo % 30), qhi);
    const bool have_range = lo < hi;
    CK(clw_seg_clear(ctx, &S->seg));
    if (first && have_small) {
      uint32_t h_side_cnt[4] = {0, 0, 0, 0};
      for (int t = 0; t < 3; ++t) {
        if (small_side[t].empty()) continue;
        h_side_cnt[t] = (uint32_t)(small_side[t].size() / 2);
        CK(clw_h2d(ctx, S->seg.side[t], small_side[t].data(), small_side[t].size() * 8, 1));
      }
      CK(clw_h2d(ctx, S->seg.side_count, h_side_cnt, 16, 1));
    }
    hipStream_t st = clw_stream(ctx);
    bool ran_sieve = false, ran_t[3] = {false, false, false}, ran_rec[3] = {false, false, false};
    uint32_t nblk_t[3] = {0, 0, 0};
    if (have_range) {
      uint32_t nb = 0;
      CK(clw_hip(hipEventRecord(S->ev[0], st), "hipEventRecord"));
      CK(clw_sieve_segment(ctx, &S->seg, lo, hi, S->T, &nb));
      if (S->subp) CK(clwi_session_subprod(ctx, S->seg.rec, S->seg.nrec, nb, S->subp, S->subn, S->ds, S->seg.flags));
      CK(clw_hip(hipEventRecord(S->ev[1], st), "hipEventRecord"));
      ran_sieve = true;
    }
    for (uint32_t t = 0; t < 3; ++t) {
      const uint32_t n = S->cnt[t];
      if (!n) continue;
      const bool recs = have_range && S->T[t] >= lo;
      if (!recs && !(first && !small_side[t].empty())) continue;
      const uint32_t nblk = recs ? clw_blocks_upto(lo, hi, S->T[t]) : 0;
      hipEvent_t* e4 = S->ev + 2 + 4 * t;
      ran_t[t] = true;
      ran_rec[t] = nblk > 0;
      nblk_t[t] = nblk;
      CK(clw_hip(hipEventRecord(e4[0], st), "hipEventRecord"));
      uint32_t nparts = 0, Ks = 1;
      if (S->opt.mode == 3) {
        // one test prime per launch pair, as the reference's segment loop
        // (cl_wilson.cpp:1895-1923) enqueues mullarge + reduce per prime
        const uint32_t groups = std::max<uint32_t>(1, std::min<uint32_t>(kPlainGroups, nblk));
        if (nblk)
          for (uint32_t i = 0; i < n; ++i) {
            CK(clw_plain_mul(ctx, &S->tab, &S->seg, t, S->off[t] + i, nblk, groups, S->part_lo, S->part_hi));
            CK(clw_plain_reduce(ctx, &S->tab, S->off[t] + i, groups, S->part_lo, S->part_hi));
          }
      } else if (S->lanes_mode) {
        // a wave per test prime, lanes over the records; ~2+ blocks per chunk
        uint32_t K = std::max<uint32_t>(1, (S->target_waves + n - 1) / n);
        K = std::min<uint32_t>(K, std::max<uint32_t>(1, nblk / 2));
        K = std::min<uint32_t>(K, kMaxKLanes);
        if (nblk) {
          CK(clw_mul_records_lanes(ctx, &S->tab, &S->seg, t, S->off[t], n, nblk, K, S->part_lo, S->part_hi, n));
          nparts = K;
        }
      } else {
        const uint32_t per_lane = S->opt.per_lane ? S->opt.per_lane : clw_default_per_lane(S->tab.limbs);
        const uint32_t tiles = (n + 256 * per_lane - 1) / (256 * per_lane);
        uint32_t K = std::max<uint32_t>(1, (S->target_waves + 4 * tiles - 1) / (4 * tiles));
        K = std::min<uint32_t>(K, std::min<uint32_t>(kMaxKRec, std::max<uint32_t>(nblk, 1)));
        Ks = std::min<uint32_t>(kMaxKSide, std::max<uint32_t>(1, K / 4));
        if (nblk) {
          if (S->subp)
            CK(clwi_mul_records_runs(ctx, &S->tab, &S->seg, S->subp, S->subn, S->ds, t, S->off[t], n, nblk, K,
                                     S->part_lo, S->part_hi, n, per_lane, S->pprod));
          else
            CK(clw_mul_records(ctx, &S->tab, &S->seg, t, S->off[t], n, nblk, K, S->part_lo, S->part_hi, n,
                               per_lane));
          nparts = K;
        }
      }
      CK(clw_hip(hipEventRecord(e4[1], st), "hipEventRecord"));
      CK(clw_mul_side(ctx, &S->tab, &S->seg, t, S->off[t], n, Ks, S->part_lo + (size_t)nparts * n,
                      S->part_hi + (size_t)nparts * n, n));
      CK(clw_hip(hipEventRecord(e4[2], st), "hipEventRecord"));
      nparts += Ks;
      CK(clw_reduce(ctx, &S->tab, S->off[t], n, nparts, S->part_lo, S->part_hi, n));  // reduce.cl
      CK(clw_hip(hipEventRecord(e4[3], st), "hipEventRecord"));
    }
    // clearn.cl: add the segment's prime count; check the device flags.
    uint32_t flags[4] = {0, 0, 0, 0};
    uint64_t seg_total[2] = {0, 0};
    CK(clw_d2h(ctx, flags, S->seg.flags, 16, 0));
    CK(clw_d2h(ctx, seg_total, S->seg.total, 16, 1));
    if (flags[0])
      return clw_fail(CLW_ERR_OVERFLOW, "device capacity exceeded in segment [%llu, %llu): flags %u",
                      (unsigned long long)lo, (unsigned long long)hi, flags[0]);
    S->total_primes += seg_total[0];
    if (S->pprod) {
      uint64_t pp = 0;
      CK(clw_d2h(ctx, &pp, S->pprod, 8, 1));
      CK(clw_memset(ctx, S->pprod, 0, 8));
      S->ks.mul_partial_products += pp;
    }
    // Kernel timings and work of this segment.
    S->ks.segments++;
    if (ran_sieve) {
      S->ks.sieve_launches++;
      S->ks.sieve_ms += ev_ms(S->ev[0], S->ev[1]);
      S->ks.sieve_integers += hi - lo;
    }
    uint32_t maxblk = std::max(nblk_t[0], std::max(nblk_t[1], nblk_t[2]));
    std::vector<uint32_t> nrec_h(maxblk), npr_h(maxblk);
    if (maxblk) {
      CK(clw_d2h(ctx, nrec_h.data(), S->seg.nrec, (size_t)maxblk * 4, 0));
      CK(clw_d2h(ctx, npr_h.data(), S->seg.nprimes, (size_t)maxblk * 4, 1));
    }
    for (int t = 0; t < 3; ++t) {
      if (!ran_t[t]) continue;
      hipEvent_t* e4 = S->ev + 2 + 4 * t;
      if (ran_rec[t]) {
        S->ks.mul_launches++;
        S->ks.mul_ms += ev_ms(e4[0], e4[1]);
        uint64_t recs = 0, prs = 0;
        for (uint32_t b = 0; b < nblk_t[t]; ++b) recs += nrec_h[b], prs += npr_h[b];
        S->ks.mul_record_apps += recs * S->cnt[t];
        S->ks.mul_prime_apps += prs * S->cnt[t];
        if (!S->subp)  // record products only (the run-product kernel counts its own on the device)
          S->ks.mul_partial_products += recs * S->cnt[t] * (2 * S->tab.limbs * S->tab.limbs + S->tab.limbs);
      }
      S->ks.side_launches++;
      S->ks.side_ms += ev_ms(e4[1], e4[2]);
      S->ks.reduce_launches++;
      S->ks.reduce_ms += ev_ms(e4[2], e4[3]);
    }
    first = false;
    if (!have_range) break;
    lo = hi;
    if (S->opt.verbose) {
      const auto now = std::chrono::steady_clock::now();
      if (std::chrono::duration<double>(now - last_report).count() > 4.0) {
        last_report = now;
        fprintf(stderr, "%.4f%%\n", 100.0 * (double)lo / (double)(S->maxT + 1));
      }
    }
  }
  S->gpu_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return CLW_OK;
}

// Canonical residues (value mod p^2) in increasing p.
extern "C" int clw_session_residues(clw_session* S, uint64_t* lo, uint64_t* hi, uint64_t cap) {
  if (!S || !lo || !hi) return clw_fail(CLW_ERR_ARG, "clw_session_residues: null argument");
  const uint32_t n = (uint32_t)S->tp.size();
  std::vector<uint64_t> a(n), b(n);
  CK(clw_d2h(S->ctx, a.data(), S->tab.acc_lo, (size_t)n * 8, 0));
  CK(clw_d2h(S->ctx, b.data(), S->tab.acc_hi, (size_t)n * 8, 1));
  for (uint32_t k = 0; k < n && k < cap; ++k) {
    lo[k] = a[S->pos[k]];
    hi[k] = b[S->pos[k]];
  }
  return CLW_OK;
}

extern "C" int clw_session_primes(const clw_session* S, uint64_t* out, uint64_t cap) {
  if (!S || !out) return clw_fail(CLW_ERR_ARG, "clw_session_primes: null argument");
  for (size_t k = 0; k < S->tp.size() && k < cap; ++k) out[k] = S->tp[k];
  return CLW_OK;
}

// iterate.cl (cl_wilson.cpp:1938-1946).
extern "C" int clw_session_iterate(clw_session* S) {
  if (!S) return clw_fail(CLW_ERR_ARG, "clw_session_iterate: null session");
  if (S->iterated) return clw_fail(CLW_ERR_STATE, "clw_session_iterate called twice");
  const auto t0 = std::chrono::steady_clock::now();
  CK(clw_iterate(S->ctx, &S->tab, S->T));
  CK(clw_finish(S->ctx));
  S->iterated = true;
  S->gpu_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return CLW_OK;
}

// Final check (processResult, cl_wilson.cpp:691-900) for every test prime;
// statistics as the reference accumulates them.
extern "C" int clw_session_finalize(clw_session* S, clw_range_stats* stats, uint64_t* out_p, uint64_t* out_quot,
                                    uint64_t* out_rem, uint64_t out_cap) {
  if (!S || !stats) return clw_fail(CLW_ERR_ARG, "clw_session_finalize: null argument");
  if (!S->iterated) return clw_fail(CLW_ERR_STATE, "clw_session_finalize before iterate");
  const auto t0 = std::chrono::steady_clock::now();
  clw_ctx* ctx = S->ctx;
  const uint32_t ntp = (uint32_t)S->tp.size();
  uint64_t *dq = nullptr, *dr = nullptr;
  int rc = clw_alloc(ctx, (size_t)ntp * 8, (void**)&dq);
  if (rc == CLW_OK) rc = clw_alloc(ctx, (size_t)ntp * 8, (void**)&dr);
  std::vector<uint64_t> q(ntp), r(ntp);
  if (rc == CLW_OK) rc = clw_finalize(ctx, &S->tab, dq, dr, nullptr);
  if (rc == CLW_OK) rc = clw_d2h(ctx, q.data(), dq, (size_t)ntp * 8, 0);
  if (rc == CLW_OK) rc = clw_d2h(ctx, r.data(), dr, (size_t)ntp * 8, 1);
  clw_free(ctx, dq);
  clw_free(ctx, dr);
  if (rc != CLW_OK) return rc;
  S->gpu_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  clw_session_info(S, stats);
  int bad = 0;
  uint64_t checksum = 0;
  for (uint32_t k = 0; k < ntp; ++k) {
    const uint64_t p = S->tp[k], quot = q[S->pos[k]], rem = r[S->pos[k]];
    if (rem != 0) bad = 1;  // Wilson's theorem mod p (cl_wilson.cpp:866-870)
    if (clw_format_result(p, quot, nullptr, 0)) {
      stats->resultcount++;
      stats->lastp = p;
      stats->lastw = signed_w(p, quot);
    }
    checksum += p + rem + quot;
    if (k < out_cap) {
      if (out_p) out_p[k] = p;
      if (out_quot) out_quot[k] = quot;
      if (out_rem) out_rem[k] = rem;
    }
  }
  stats->checksum = checksum;
  if (bad)
    return clw_fail(CLW_ERR_CHECK, "Wilson quotient check failed (rem != 0) in [%llu, %llu)",
                    (unsigned long long)S->pmin, (unsigned long long)S->pmax);
  return CLW_OK;
}

// Ranges with at least this many test primes (estimated as width / ln pmin)
// run through a one-range batch in the automatic mode: the batched engine's
// run products are ~2x faster per prime than the session's record products,
// while the session's wave-per-test-prime kernel suits small ranges.
static const double kBatchMinTestPrimes = 65536;

static int wilson_range_batch(clw_ctx* ctx, uint64_t pmin, uint64_t pmax, const clw_options* opt,
                              clw_range_stats* stats, uint64_t* out_p, uint64_t* out_quot, uint64_t* out_rem,
                              uint64_t out_cap) {
  clw_options bo = opt ? *opt : clw_options{};
  bo.mode = 0;
  clw_batch* B = nullptr;
  CK(clw_batch_create(ctx, &pmin, &pmax, 1, &bo, &B));
  const auto t0 = std::chrono::steady_clock::now();
  clw_range_stats info{};
  int rc = clw_batch_info(B, 0, &info);
  if (rc == CLW_OK) rc = clw_batch_run(B, 2, info.maxtarget + 1);
  if (rc == CLW_OK) rc = clw_batch_iterate(B);
  if (rc == CLW_OK) rc = clw_batch_finalize(B);
  if (rc == CLW_OK) rc = clw_batch_results(B, 0, stats, out_p, out_quot, out_rem, out_cap);
  if (rc == CLW_OK) {
    stats->totalcount = clw_batch_applied(B);  // one range: pi(its largest target)
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  const int rd = clw_batch_destroy(B);
  return rc != CLW_OK ? rc : rd;
}

extern "C" int clw_wilson_range(clw_ctx* ctx, uint64_t pmin, uint64_t pmax, const clw_options* opt,
                                clw_range_stats* stats, uint64_t* out_p, uint64_t* out_quot, uint64_t* out_rem,
                                uint64_t out_cap) {
  if (!ctx || !stats) return clw_fail(CLW_ERR_ARG, "clw_wilson_range: null argument");
  memset(stats, 0, sizeof(*stats));
  if ((!opt || opt->mode == 0) && pmin >= 5 && pmin < pmax && pmax - pmin <= 10000000 &&
      clw_limbs_for(pmax - 1) <= 4 && (double)(pmax - pmin) / std::log((double)pmin) >= kBatchMinTestPrimes)
    return wilson_range_batch(ctx, pmin, pmax, opt, stats, out_p, out_quot, out_rem, out_cap);
  clw_session* S = nullptr;
  CK(clw_session_create(ctx, pmin, pmax, opt, &S));
  int rc = clw_session_run(S, 2, S->maxT + 1);
  if (rc == CLW_OK) rc = clw_session_iterate(S);
  if (rc == CLW_OK) rc = clw_session_finalize(S, stats, out_p, out_quot, out_rem, out_cap);
  clw_session_destroy(S);
  return rc;
}

// ---------------------------------------------------------------------------
// Checkpoints (cl_wilson.cpp:110-319)
// ---------------------------------------------------------------------------
// The reference's additive sum (cl_wilson.cpp:116-119) plus a format
// constant: this library stores canonical residues where the reference stores
// Montgomery digits, so a file of one program must never validate in the
// other (the residues would be read in the wrong form).
static const uint64_t kCkptFormatSalt = 0x434C57414D443031ull;  // "CLWAMD01"
static uint64_t ckpt_sum(const clw_work_status* st, const uint64_t* lo, const uint64_t* hi) {
  uint64_t s = kCkptFormatSalt + st->pmin + st->pmax + st->currp + st->trickle + st->totalcount + st->tpcount +
               st->done;
  for (uint32_t i = 0; i < st->tpcount; ++i) s += lo[i] + hi[i];
  return s;
}

extern "C" int clw_ckpt_write(const char* path, clw_work_status* st, const uint64_t* res_lo,
                              const uint64_t* res_hi) {
  static_assert(sizeof(clw_work_status) == 56, "workStatus layout");
  if (!path || !st || (st->tpcount && (!res_lo || !res_hi))) return clw_fail(CLW_ERR_ARG, "clw_ckpt_write: null");
  st->state_sum = ckpt_sum(st, res_lo, res_hi);
  FILE* f = fopen(path, "wb");
  if (!f) return clw_fail(CLW_ERR_IO, "cannot open %s", path);
  bool ok = fwrite(st, sizeof(*st), 1, f) == 1;
  for (uint32_t i = 0; ok && i < st->tpcount; ++i) {
    const uint64_t pair[2] = {res_lo[i], res_hi[i]};
    ok = fwrite(pair, 16, 1, f) == 1;
  }
  if (fclose(f) != 0) ok = false;
  if (!ok) return clw_fail(CLW_ERR_IO, "cannot write checkpoint %s", path);
  return CLW_OK;
}

namespace {
// 0 invalid/missing, 1 valid, 2 valid and done.
int ckpt_read_one(const char* path, uint64_t pmin, uint64_t pmax, uint32_t tpcount, clw_work_status* st,
                  std::vector<uint64_t>& lo, std::vector<uint64_t>& hi) {
  FILE* f = path ? fopen(path, "rb") : nullptr;
  if (!f) return 0;
  int r = 1;
  if (fread(st, sizeof(*st), 1, f) != 1) r = 0;
  if (r && (st->tpcount != tpcount || st->pmin != pmin || st->pmax != pmax)) r = 0;
  if (r) {
    lo.assign(tpcount, 0);
    hi.assign(tpcount, 0);
    for (uint32_t i = 0; r && i < tpcount; ++i) {
      uint64_t pair[2];
      if (fread(pair, 16, 1, f) != 1) r = 0;
      lo[i] = pair[0];
      hi[i] = pair[1];
    }
  }
  fclose(f);
  if (!r) return 0;
  // the sum is checked for finished files too (the reference returns 2
  // before its check, cl_wilson.cpp:192-194): a foreign or damaged file
  // never ends a workunit
  if (ckpt_sum(st, lo.data(), hi.data()) != st->state_sum) return 0;
  return st->done ? 2 : 1;
}
}  // namespace

extern "C" int clw_ckpt_read(const char* path_a, const char* path_b, uint64_t pmin, uint64_t pmax, uint32_t tpcount,
                             clw_work_status* st, uint64_t* res_lo, uint64_t* res_hi, int* which, int* result) {
  if (!st || !result || (tpcount && (!res_lo || !res_hi))) return clw_fail(CLW_ERR_ARG, "clw_ckpt_read: null");
  *result = 0;
  if (which) *which = -1;
  clw_work_status sa{}, sb{};
  std::vector<uint64_t> la, ha, lb, hb;
  const int ra = ckpt_read_one(path_a, pmin, pmax, tpcount, &sa, la, ha);
  const int rb = ckpt_read_one(path_b, pmin, pmax, tpcount, &sb, lb, hb);
  if (ra == 2 || rb == 2) {
    *result = 2;
    *st = ra == 2 ? sa : sb;
    return CLW_OK;
  }
  int pick = -1;
  if (ra == 1 && rb == 1) pick = sa.currp > sb.currp ? 0 : 1;  // cl_wilson.cpp:253-259
  else if (ra == 1) pick = 0;
  else if (rb == 1) pick = 1;
  if (pick < 0) return CLW_OK;
  *st = pick == 0 ? sa : sb;
  const std::vector<uint64_t>& L = pick == 0 ? la : lb;
  const std::vector<uint64_t>& H = pick == 0 ? ha : hb;
  for (uint32_t i = 0; i < tpcount; ++i) {
    res_lo[i] = L[i];
    res_hi[i] = H[i];
  }
  if (which) *which = pick;
  *result = 1;
  return CLW_OK;
}

extern "C" int clw_session_set_residues(clw_session* S, const uint64_t* lo, const uint64_t* hi, uint64_t totalcount) {
  if (!S || !lo || !hi) return clw_fail(CLW_ERR_ARG, "clw_session_set_residues: null argument");
  const uint32_t n = (uint32_t)S->tp.size();
  std::vector<uint64_t> a(n), b(n);
  for (uint32_t k = 0; k < n; ++k) {
    const unsigned __int128 v = ((unsigned __int128)hi[k] << 64) | lo[k];
    if (v >= (unsigned __int128)S->tp[k] * S->tp[k])
      return clw_fail(CLW_ERR_ARG, "residue %u is not reduced mod p^2", k);
    a[S->pos[k]] = lo[k];
    b[S->pos[k]] = hi[k];
  }
  CK(clw_h2d(S->ctx, S->tab.acc_lo, a.data(), (size_t)n * 8, 0));
  CK(clw_h2d(S->ctx, S->tab.acc_hi, b.data(), (size_t)n * 8, 1));
  S->total_primes = totalcount;
  return CLW_OK;
}
