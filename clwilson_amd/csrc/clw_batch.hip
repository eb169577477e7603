// clw_batch.hip -- the batched engine: many ranges searched together over
// one shared stream of sieving primes.
//
// Every range needs the same sieving primes (only up to different targets
// and with different exponents), so a batch sieves each segment once and a
// persistent product kernel per limb band pops (chunk, group, tile) work
// items from a device queue. A group is one (range, type) pair with its own
// factorial target; exponents are computed on the fly from the target. The
// segment step (advance descriptor, sieve, products, reductions) reads its
// parameters from device memory and is captured once into a HIP graph and
// replayed for every further segment.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "clw_internal.h"
#include "clw_nt64.h"

namespace {

struct SegHost {  // mirror of BatchSeg (clw_batch_kernels.inc)
  uint64_t lo, hi, base, span, end;
  uint32_t nblk, nbases;
  uint32_t queue[4];
};
struct GroupHost {  // mirror of BatchGroup
  uint64_t T, sqrtT;
  uint32_t tp_off, tp_n;
  uint64_t part_off;
};

const uint64_t kMaxP = 0xFFFFFFFFFFFFFFFFull / 4;
const uint64_t kMaxWidth = 10000000;
const uint64_t kBatchSpan = 4370ull * CLW_BLOCK_SPAN;

inline uint32_t prime_type(uint64_t p) { return p % 3 == 1 ? 0 : (p % 12 == 5 ? 1 : 2); }
inline uint64_t type_divisor(uint32_t t) { return t == 0 ? 6 : (t == 1 ? 4 : 2); }

struct Band {
  int L = 0;
  uint32_t per_lane = 0;
  clw_tp_table tab{};
  std::vector<uint64_t> p, from;
  std::vector<uint32_t> type, tp_group, item_gt;
  std::vector<GroupHost> groups;
  GroupHost* groups_d = nullptr;
  uint32_t* tp_group_d = nullptr;
  uint32_t* item_gt_d = nullptr;
  uint64_t* from_d = nullptr;
  uint64_t* part_lo = nullptr;
  uint64_t* part_hi = nullptr;
  uint32_t K = 1, items_per_chunk = 0;
  std::vector<uint64_t> quot, rem;
};

struct RangeHost {
  uint64_t pmin, pmax;
  std::vector<uint64_t> tp;   // increasing
  std::vector<uint32_t> pos;  // index of tp[k] in its band table
  int band;
  uint32_t cnt[3] = {0, 0, 0};
  uint64_t T[3] = {0, 0, 0};
};

#define CK(x)                      \
  do {                             \
    const int rc_ = (x);           \
    if (rc_ != CLW_OK) return rc_; \
  } while (0)

int64_t signed_w(uint64_t p, uint64_t quot) {
  const uint64_t neg = p - quot;
  return quot <= neg ? (int64_t)quot : -(int64_t)neg;
}

}  // namespace

struct clw_batch {
  clw_ctx* ctx = nullptr;
  clw_options opt{};
  std::vector<RangeHost> ranges;
  Band bands[6];  // by limb count
  uint64_t maxT = 0, span = 0;
  uint32_t max_blocks = 0;
  uint32_t* rec = nullptr;
  uint32_t* nrec = nullptr;
  uint32_t* flags = nullptr;
  uint64_t* totals = nullptr;  // [0] this segment, [1] accumulated
  SegHost* seg_d = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool iterated = false, finalized = false;
  clw_kernel_stats ks{};
  uint64_t applied = 0;  // sieving primes applied (all windows)

  ~clw_batch() {
    if (graph_exec) hipGraphExecDestroy(graph_exec);
    if (graph) hipGraphDestroy(graph);
    for (Band& b : bands) {
      if (!b.L) continue;
      clw_tp_table_free(ctx, &b.tab);
      void* ptrs[] = {b.groups_d, b.tp_group_d, b.item_gt_d, b.from_d, b.part_lo, b.part_hi};
      for (void* x : ptrs) clw_free(ctx, x);
    }
    void* ptrs[] = {rec, nrec, flags, totals, seg_d};
    for (void* x : ptrs) clw_free(ctx, x);
  }
};

extern "C" int clw_batch_create(clw_ctx* ctx, const uint64_t* pmin, const uint64_t* pmax, uint32_t nranges,
                                const clw_options* opt, clw_batch** out) {
  if (!ctx || !pmin || !pmax || !out || nranges == 0) return clw_fail(CLW_ERR_ARG, "clw_batch_create: bad argument");
  if (clwi_batch_seg_size() != sizeof(SegHost) || clwi_batch_group_size() != sizeof(GroupHost))
    return clw_fail(CLW_ERR_STATE, "batch descriptor layout mismatch");
  *out = nullptr;
  clw_batch* B = new clw_batch();
  B->ctx = ctx;
  B->opt = opt ? *opt : clw_options{};
  B->ranges.resize(nranges);
  for (uint32_t r = 0; r < nranges; ++r) {
    RangeHost& R = B->ranges[r];
    R.pmin = pmin[r];
    R.pmax = pmax[r];
    if (R.pmin < 5 || R.pmax > kMaxP || R.pmin >= R.pmax || R.pmax - R.pmin > kMaxWidth) {
      delete B;
      return clw_fail(CLW_ERR_ARG, "range %u [%llu, %llu) invalid", r, (unsigned long long)R.pmin,
                      (unsigned long long)R.pmax);
    }
    R.tp.resize(clw_host_primes(R.pmin, R.pmax, nullptr, 0));
    clw_host_primes(R.pmin, R.pmax, R.tp.data(), R.tp.size());
    if (R.tp.empty()) {
      delete B;
      return clw_fail(CLW_ERR_ARG, "range %u has no primes", r);
    }
    for (uint64_t p : R.tp) {
      const uint32_t t = prime_type(p);
      if (!R.cnt[t]++) R.T[t] = (p - 1) / type_divisor(t);
    }
    R.band = clw_limbs_for(R.tp.back());
    if (R.band > 4) {
      delete B;
      return clw_fail(CLW_ERR_ARG, "range %u: p above 8.2e17 is served by sessions, not batches", r);
    }
    for (int t = 0; t < 3; ++t) B->maxT = std::max(B->maxT, R.T[t]);
  }
  // Band tables: test primes grouped by (range, type); one group per pair.
  for (uint32_t r = 0; r < nranges; ++r) {
    RangeHost& R = B->ranges[r];
    Band& bd = B->bands[R.band];
    bd.L = R.band;
    R.pos.resize(R.tp.size());
    for (uint32_t t = 0; t < 3; ++t) {
      if (!R.cnt[t]) continue;
      GroupHost g{};
      g.T = R.T[t];
      g.sqrtT = clw::isqrt64(R.T[t]);
      g.tp_off = (uint32_t)bd.p.size();
      g.tp_n = R.cnt[t];
      const uint32_t gi = (uint32_t)bd.groups.size();
      for (size_t k = 0; k < R.tp.size(); ++k) {
        if (prime_type(R.tp[k]) != t) continue;
        R.pos[k] = (uint32_t)bd.p.size();
        bd.p.push_back(R.tp[k]);
        bd.type.push_back(t);
        bd.from.push_back(R.T[t]);
        bd.tp_group.push_back(gi);
      }
      bd.groups.push_back(g);
    }
  }
  // Segment buffers and descriptor.
  B->span = B->opt.segment_span ? (B->opt.segment_span + CLW_BLOCK_SPAN - 1) / CLW_BLOCK_SPAN * CLW_BLOCK_SPAN
                                : kBatchSpan;
  B->max_blocks = (uint32_t)(B->span / CLW_BLOCK_SPAN + 2);  // +1 slot for the primes 2, 3, 5
  int rc = clw_alloc(ctx, (size_t)B->max_blocks * CLW_REC_CAP * CLW_REC_WORDS * 4, (void**)&B->rec);
  if (rc == CLW_OK) rc = clw_alloc(ctx, (size_t)B->max_blocks * 4, (void**)&B->nrec);
  if (rc == CLW_OK) rc = clw_alloc(ctx, 16, (void**)&B->flags);
  if (rc == CLW_OK) rc = clw_alloc(ctx, 16, (void**)&B->totals);
  if (rc == CLW_OK) rc = clw_alloc(ctx, sizeof(SegHost), (void**)&B->seg_d);
  if (rc == CLW_OK) rc = clw_memset(ctx, B->flags, 0, 16);
  if (rc == CLW_OK) rc = clw_memset(ctx, B->totals, 0, 16);
  const uint32_t waves = B->opt.target_waves ? B->opt.target_waves : (uint32_t)ctx->prop.multiProcessorCount * 4 * 8;
  for (Band& bd : B->bands) {
    if (!bd.L || rc != CLW_OK) continue;
    bd.per_lane = B->opt.per_lane ? B->opt.per_lane : clw_default_per_lane((uint32_t)bd.L);
    const uint32_t n = (uint32_t)bd.p.size();
    // items: (group, tile); chunks K so that items * K covers the wave target
    bd.items_per_chunk = 0;
    for (uint32_t gi = 0; gi < bd.groups.size(); ++gi) {
      const uint32_t tiles = (bd.groups[gi].tp_n + 256 * bd.per_lane - 1) / (256 * bd.per_lane);
      for (uint32_t tile = 0; tile < tiles; ++tile) bd.item_gt.push_back((gi << 16) | tile);
      bd.items_per_chunk += tiles;
    }
    if (bd.groups.size() >= 65536 || bd.items_per_chunk == 0) {
      rc = clw_fail(CLW_ERR_ARG, "too many groups in a batch band");
      break;
    }
    bd.K = std::max<uint32_t>(1, std::min<uint32_t>(64, (2 * waves / 4 + bd.items_per_chunk - 1) / bd.items_per_chunk));
    uint64_t off = 0;
    for (GroupHost& g : bd.groups) {
      g.part_off = off;
      off += (uint64_t)bd.K * g.tp_n;
    }
    rc = clw_tp_table_alloc(ctx, n, (uint32_t)bd.L, &bd.tab);
    if (rc == CLW_OK) rc = clw_h2d(ctx, bd.tab.p, bd.p.data(), (size_t)n * 8, 1);
    if (rc == CLW_OK) rc = clw_h2d(ctx, bd.tab.type, bd.type.data(), (size_t)n * 4, 1);
    if (rc == CLW_OK) rc = clw_setup(ctx, &bd.tab, 0);
    if (rc == CLW_OK) rc = clw_alloc(ctx, bd.groups.size() * sizeof(GroupHost), (void**)&bd.groups_d);
    if (rc == CLW_OK) rc = clw_h2d(ctx, bd.groups_d, bd.groups.data(), bd.groups.size() * sizeof(GroupHost), 1);
    if (rc == CLW_OK) rc = clw_alloc(ctx, (size_t)n * 4, (void**)&bd.tp_group_d);
    if (rc == CLW_OK) rc = clw_h2d(ctx, bd.tp_group_d, bd.tp_group.data(), (size_t)n * 4, 1);
    if (rc == CLW_OK) rc = clw_alloc(ctx, bd.item_gt.size() * 4, (void**)&bd.item_gt_d);
    if (rc == CLW_OK) rc = clw_h2d(ctx, bd.item_gt_d, bd.item_gt.data(), bd.item_gt.size() * 4, 1);
    if (rc == CLW_OK) rc = clw_alloc(ctx, (size_t)n * 8, (void**)&bd.from_d);
    if (rc == CLW_OK) rc = clw_h2d(ctx, bd.from_d, bd.from.data(), (size_t)n * 8, 1);
    if (rc == CLW_OK) rc = clw_alloc(ctx, (size_t)off * 8, (void**)&bd.part_lo);
    if (rc == CLW_OK) rc = clw_alloc(ctx, (size_t)off * 8, (void**)&bd.part_hi);
  }
  if (rc == CLW_OK) rc = clw_finish(ctx);
  if (rc != CLW_OK) {
    delete B;
    return rc;
  }
  *out = B;
  return CLW_OK;
}

extern "C" int clw_batch_destroy(clw_batch* B) {
  delete B;
  return CLW_OK;
}

extern "C" uint32_t clw_batch_size(const clw_batch* B) { return B ? (uint32_t)B->ranges.size() : 0; }

extern "C" int clw_batch_info(const clw_batch* B, uint32_t r, clw_range_stats* st) {
  if (!B || !st || r >= B->ranges.size()) return clw_fail(CLW_ERR_ARG, "clw_batch_info: bad argument");
  const RangeHost& R = B->ranges[r];
  memset(st, 0, sizeof(*st));
  st->pmin = R.pmin;
  st->pmax = R.pmax;
  st->tpcount = (uint32_t)R.tp.size();
  for (int t = 0; t < 3; ++t) {
    st->tpcnt[t] = R.cnt[t];
    st->typeTarget[t] = R.T[t];
  }
  st->maxtarget = std::max(R.T[0], std::max(R.T[1], R.T[2]));
  st->limbs = (uint32_t)R.band;
  return CLW_OK;
}

static int batch_step(clw_batch* B, uint32_t extra_block) {
  clw_ctx* ctx = B->ctx;
  CK(clwi_seg_advance(ctx, B->seg_d, B->flags, B->totals + 1, B->totals));
  CK(clwi_batch_sieve(ctx, B->seg_d, B->rec, B->nrec, B->max_blocks - 1, B->flags, B->totals));
  for (int band = 2; band <= 4; ++band) {
    Band& bd = B->bands[band];
    if (!bd.L) continue;
    CK(clwi_batch_mul(ctx, &bd.tab, bd.per_lane, bd.groups_d, bd.item_gt_d, bd.items_per_chunk, bd.K, B->rec,
                      B->nrec, B->seg_d, (uint32_t)(band - 2), extra_block, 0, bd.part_lo, bd.part_hi));
    CK(clwi_batch_reduce(ctx, &bd.tab, bd.tp_group_d, bd.groups_d, bd.K, bd.part_lo, bd.part_hi));
  }
  return CLW_OK;
}

// Applies every sieving prime q in [qlo, qhi) (clipped to [2, max target])
// to every range of the batch. Windows must be disjoint across calls.
extern "C" int clw_batch_run(clw_batch* B, uint64_t qlo, uint64_t qhi) {
  if (!B) return clw_fail(CLW_ERR_ARG, "clw_batch_run: null batch");
  if (B->iterated) return clw_fail(CLW_ERR_STATE, "clw_batch_run after iterate");
  clw_ctx* ctx = B->ctx;
  const auto t0 = std::chrono::steady_clock::now();
  if (qhi > B->maxT + 1) qhi = B->maxT + 1;
  if (qlo < 2) qlo = 2;
  if (qhi <= qlo) return CLW_OK;
  const uint64_t lo = std::max<uint64_t>(qlo, 7);
  // Host replica of k_seg_advance: the segments of [lo, qhi).
  std::vector<std::pair<uint64_t, uint64_t>> segs;
  for (uint64_t a = lo; a < qhi;) {
    uint64_t b = a + B->span - (a % 30);
    if (b > qhi) b = qhi;
    segs.push_back({a, b});
    a = b;
  }
  const bool small = qlo <= 5;
  if (segs.empty() && small) segs.push_back({lo, lo});  // only 2, 3, 5
  SegHost h{};
  h.lo = lo;
  h.hi = lo;  // the first advance yields [lo, ...)
  h.span = B->span;
  h.end = std::max(qhi, lo);
  CK(clw_h2d(ctx, B->seg_d, &h, sizeof(h), 1));
  for (size_t s = 0; s < segs.size(); ++s) {
    // The first segment of a window that holds 2, 3, 5 carries them as an
    // extra record block (slot nblk); its step runs eagerly.
    const bool extra = (s == 0 && small);
    if (extra) {
      const uint32_t nb = segs[0].second > segs[0].first
                              ? (uint32_t)((segs[0].second - segs[0].first / 30 * 30 + CLW_BLOCK_SPAN - 1) /
                                           CLW_BLOCK_SPAN)
                              : 0;
      CK(clwi_small_primes_block(ctx, qlo, qhi, B->rec, B->nrec, nb));
      CK(batch_step(B, 1));
    } else if (B->opt.graph) {
      if (!B->graph_exec) {
        hipStream_t st = clw_stream(ctx);
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return clw_fail_hip(e, "hipStreamBeginCapture");
        const int rc = batch_step(B, 0);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(st, &g);
        if (rc != CLW_OK) return rc;
        if (e != hipSuccess) return clw_fail_hip(e, "hipStreamEndCapture");
        B->graph = g;
        e = hipGraphInstantiate(&B->graph_exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) return clw_fail_hip(e, "hipGraphInstantiate");
      }
      const hipError_t e = hipGraphLaunch(B->graph_exec, clw_stream(ctx));
      if (e != hipSuccess) return clw_fail_hip(e, "hipGraphLaunch");
    } else {
      CK(batch_step(B, 0));
    }
    B->ks.segments++;
    B->ks.sieve_integers += segs[s].second - segs[s].first;
  }
  // Flags and the number of sieving primes applied in this window.
  CK(clwi_seg_advance(ctx, B->seg_d, B->flags, B->totals + 1, B->totals));  // folds the last segment's count
  uint32_t flags[4] = {0, 0, 0, 0};
  uint64_t tot[2] = {0, 0};
  CK(clw_d2h(ctx, flags, B->flags, 16, 0));
  CK(clw_d2h(ctx, tot, B->totals, 16, 1));
  if (flags[0]) return clw_fail(CLW_ERR_OVERFLOW, "device capacity exceeded: flags %u", flags[0]);
  uint64_t smallc = 0;
  for (uint64_t q : {2ull, 3ull, 5ull}) smallc += (q >= qlo && q < qhi);
  B->applied += tot[1] + smallc;
  CK(clw_memset(ctx, B->totals, 0, 16));
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  B->ks.mul_ms += sec * 1e3;  // whole-window wall time of the batched steps
  B->ks.mul_launches += segs.size();
  return CLW_OK;
}

extern "C" int clw_batch_iterate(clw_batch* B) {
  if (!B) return clw_fail(CLW_ERR_ARG, "clw_batch_iterate: null batch");
  if (B->iterated) return clw_fail(CLW_ERR_STATE, "clw_batch_iterate called twice");
  for (Band& bd : B->bands)
    if (bd.L) CK(clw_iterate_from(B->ctx, &bd.tab, bd.from_d));
  CK(clw_finish(B->ctx));
  B->iterated = true;
  return CLW_OK;
}

extern "C" int clw_batch_finalize(clw_batch* B) {
  if (!B) return clw_fail(CLW_ERR_ARG, "clw_batch_finalize: null batch");
  if (!B->iterated) return clw_fail(CLW_ERR_STATE, "clw_batch_finalize before iterate");
  for (Band& bd : B->bands) {
    if (!bd.L) continue;
    const uint32_t n = bd.tab.count;
    uint64_t *dq = nullptr, *dr = nullptr;
    int rc = clw_alloc(B->ctx, (size_t)n * 8, (void**)&dq);
    if (rc == CLW_OK) rc = clw_alloc(B->ctx, (size_t)n * 8, (void**)&dr);
    if (rc == CLW_OK) rc = clw_finalize(B->ctx, &bd.tab, dq, dr, nullptr);
    bd.quot.assign(n, 0);
    bd.rem.assign(n, 0);
    if (rc == CLW_OK) rc = clw_d2h(B->ctx, bd.quot.data(), dq, (size_t)n * 8, 0);
    if (rc == CLW_OK) rc = clw_d2h(B->ctx, bd.rem.data(), dr, (size_t)n * 8, 1);
    clw_free(B->ctx, dq);
    clw_free(B->ctx, dr);
    CK(rc);
  }
  B->finalized = true;
  return CLW_OK;
}

// Per-range results after clw_batch_finalize, in increasing p; stats as
// clw_session_finalize fills them (checksum, results, last result).
extern "C" int clw_batch_results(const clw_batch* B, uint32_t r, clw_range_stats* st, uint64_t* out_p,
                                 uint64_t* out_quot, uint64_t* out_rem, uint64_t cap) {
  if (!B || !st || r >= B->ranges.size()) return clw_fail(CLW_ERR_ARG, "clw_batch_results: bad argument");
  if (!B->finalized) return clw_fail(CLW_ERR_STATE, "clw_batch_results before clw_batch_finalize");
  CK(clw_batch_info(B, r, st));
  const RangeHost& R = B->ranges[r];
  const Band& bd = B->bands[R.band];
  int bad = 0;
  uint64_t checksum = 0;
  for (size_t k = 0; k < R.tp.size(); ++k) {
    const uint64_t p = R.tp[k], quot = bd.quot[R.pos[k]], rem = bd.rem[R.pos[k]];
    if (rem != 0) bad = 1;
    if (clw_format_result(p, quot, nullptr, 0)) {
      st->resultcount++;
      st->lastp = p;
      st->lastw = signed_w(p, quot);
    }
    checksum += p + rem + quot;
    if (k < cap) {
      if (out_p) out_p[k] = p;
      if (out_quot) out_quot[k] = quot;
      if (out_rem) out_rem[k] = rem;
    }
  }
  st->checksum = checksum;
  // sieving primes applied to this range: those up to its largest target
  st->totalcount = 0;
  if (bad) return clw_fail(CLW_ERR_CHECK, "Wilson quotient check failed in batch range %u", r);
  return CLW_OK;
}

extern "C" int clw_batch_residues(clw_batch* B, uint32_t r, uint64_t* lo, uint64_t* hi, uint64_t cap) {
  if (!B || !lo || !hi || r >= B->ranges.size()) return clw_fail(CLW_ERR_ARG, "clw_batch_residues: bad argument");
  const RangeHost& R = B->ranges[r];
  const Band& bd = B->bands[R.band];
  const uint32_t n = bd.tab.count;
  std::vector<uint64_t> a(n), b(n);
  CK(clw_d2h(B->ctx, a.data(), bd.tab.acc_lo, (size_t)n * 8, 0));
  CK(clw_d2h(B->ctx, b.data(), bd.tab.acc_hi, (size_t)n * 8, 1));
  for (size_t k = 0; k < R.tp.size() && k < cap; ++k) {
    lo[k] = a[R.pos[k]];
    hi[k] = b[R.pos[k]];
  }
  return CLW_OK;
}

extern "C" uint64_t clw_batch_applied(const clw_batch* B) { return B ? B->applied : 0; }

extern "C" int clw_batch_kernel_stats(const clw_batch* B, clw_kernel_stats* out) {
  if (!B || !out) return clw_fail(CLW_ERR_ARG, "clw_batch_kernel_stats: null argument");
  *out = B->ks;
  return CLW_OK;
}
