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
#include <atomic>
#include <cmath>
#include <chrono>
#include <thread>
#include <vector>

#include "clw_internal.h"
#include "clw_nt64.h"

namespace {

struct SegHost {  // mirror of BatchSeg (clw_batch_kernels.inc)
  uint64_t lo, hi, base, span, end;
  uint32_t nblk, nbases;
  uint32_t queue[3][8];
};
struct GroupHost {  // mirror of BatchGroup
  uint64_t T, sqrtT;
  uint32_t tp_off, tp_n;
  uint64_t part_off;
};
struct PlanHost {  // mirror of BatchBand
  uint32_t items_per_chunk, K;
  uint32_t mfma, pad;
};
const uint64_t kNoPart = ~0ull;

const uint64_t kMaxP = 0xFFFFFFFFFFFFFFFFull / 4;
const uint64_t kMaxWidth = 10000000;
const uint32_t kMaxSpanBlocks = 4370;  // largest default segment (268 M integers)
const uint32_t kMinSpanBlocks = 16;
const double kAppsPerSegment = 2e12;   // factor applications aimed at per segment
const uint32_t kItemsPerBlock = 32;    // work items per resident block (balance)
const uint32_t kMaxChunks = 256;
const uint64_t kTableBytesPerPrime = 88;  // clw_tp_table arrays per test prime
const size_t kTotalsBytes = 64;            // device counters (8 words)
const size_t kStageBack = 128;             // staging bytes for the counters handed back (flags 16, totals 64)
const uint32_t LB_DIGIT = 30;              // bits per digit of a run product
const uint64_t kMfmaTableCap = 8ull << 30;  // byte tables kept in HBM up to this size
const uint64_t kMfmaMaxP4 = 1ull << 47;   // 4-limb band on the matrix cores below this p
const uint32_t kMfmaMinRows = 1024;        // sub-chunks per segment / 8 for the matrix cores (a table pays off)
const uint32_t kMarks = 5;                 // events per timed segment: start, sieved, matrix-core
                                           // run products, products, reductions

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
  PlanHost* plan_d = nullptr;
  PlanHost plan{};
  uint64_t part_cap = 0;   // chunk-product slots
  bool part_own = true;    // part_lo/hi allocated by the batch (else a caller's scratch)
  uint32_t resident = 0;   // persistent grid of the product kernel
  uint32_t items_per_block = kItemsPerBlock;
  uint64_t active_tp = 0;  // test primes of the groups active in this window
  std::vector<uint64_t> quot, rem;
  // run products on the matrix cores (3-limb band): tiles of kMfmaTile test
  // primes of one group, their byte tables and one partial per (slice, test prime)
  bool mfma = false;
  std::vector<uint32_t> mt_group, mt_first, mt_n, mt_active;
  uint32_t* mt_d = nullptr;  // mt_group | mt_first | mt_n | mt_active | active count
  void* ctab = nullptr;      // byte tables in HBM (allocated by the first window that uses them)
  bool ctab_tried = false;
  bool mfma_lazy = false;   // largest N below 0.4998 R: k_mfma_runs<true>
  uint32_t nslices = kMfmaSlices;  // row slices of k_mfma_runs (partials per test prime)
  uint32_t pm_slices = 0;          // slices pm_lo/pm_hi hold
  uint32_t mfma_min_rows = kMfmaMinRows;  // sub-chunks per slice and segment from which the matrix cores
                                          // take the run products (the autotuner measures it)
  uint64_t* pm_lo = nullptr;
  uint64_t* pm_hi = nullptr;
  uint32_t n_mt() const { return (uint32_t)mt_group.size(); }
};

struct RangeHost {
  uint64_t pmin, pmax;
  std::vector<uint64_t> tp;   // increasing
  std::vector<uint32_t> pos;  // index of tp[k] in its band table
  int band;
  uint32_t cnt[3] = {0, 0, 0};
  uint64_t T[3] = {0, 0, 0};
  uint32_t band_off = 0;      // its test primes are [band_off, band_off + tp.size()) of the band table
  bool finished = false;      // iterated and checked on its own (clw_batch_finish_range)
  std::vector<uint64_t> quot, rem;  // by band-table order, once finished
  // a finish in flight (clw_batch_finish_launch): quotients then remainders
  // on the device, their pinned host copy, and the event after the copy
  bool finishing = false;
  uint64_t* fin_d = nullptr;
  uint64_t* fin_h = nullptr;
  hipEvent_t fin_ev = nullptr;
};

// The table restricted to entries [off, off + n) (same arrays, offset).
clw_tp_table table_view(const clw_tp_table& t, uint32_t off, uint32_t n) {
  clw_tp_table v = t;
  v.p += off;
  v.type += off;
  v.tp_target += off;
  v.n_lo += off;
  v.n_hi += off;
  v.one_lo += off;
  v.one_hi += off;
  v.r2_lo += off;
  v.r2_hi += off;
  v.ninv += off;
  v.acc_lo += off;
  v.acc_hi += off;
  v.count = n;
  return v;
}

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

// Window buffers that batches running their windows on one stream can share
// (clw_workspace_create): record buffer and counts, sieve bitmap, sub-chunk
// digits and bytes, row lists; sized for any batch of the same segment span
// (digits for sieving primes up to 2^62).
struct clw_workspace {
  clw_ctx* ctx = nullptr;
  uint32_t max_blocks = 0;
  uint32_t *rec = nullptr, *nrec = nullptr, *subp = nullptr, *subn = nullptr, *era = nullptr;
  int8_t* subb = nullptr;
  void* subq = nullptr;
  uint32_t *subm = nullptr, *rowlist = nullptr, *rowpos = nullptr, *mrows = nullptr;
  uint64_t bytes = 0;
};

struct clw_batch {
  clw_ctx* ctx = nullptr;
  clw_options opt{};
  std::vector<RangeHost> ranges;
  Band bands[6];  // by limb count
  uint64_t maxT = 0;
  uint32_t max_blocks = 0;  // record-buffer slots (largest segment + 2)
  uint32_t* rec = nullptr;
  uint32_t* nrec = nullptr;
  uint32_t* flags = nullptr;
  uint64_t* totals = nullptr;  // primes: [0] this segment, [1] accumulated; products: [2] records,
                               // [3] primes, [4] 30x30-bit partial products, [5] int8 multiply-adds
                               // of the matrix-core GEMM (algorithmic), [6] (issued)
  uint32_t* subp = nullptr;    // run products: digits of each (block, sub-chunk)
  uint32_t* subn = nullptr;    // their digit counts
  uint32_t ds = 0;             // digit stride (a multiple of 12)
  int8_t* subb = nullptr;      // run products as kMfmaKB signed bytes, in row blocks (for a band with mfma)
  void* subq = nullptr;         // per sub-chunk: first and last prime (uint4)
  uint32_t* subm = nullptr;     // per sub-chunk: primes | bytes << 16
  uint32_t* rowlist = nullptr;  // the segment's sub-chunks, dense (k_mfma_rows)
  uint32_t* rowpos = nullptr;   // and each sub-chunk's place in that list
  uint32_t* era = nullptr;      // wheel bitmap of the sieve of Eratosthenes (segments below 2^50)
  uint32_t* mrows = nullptr;
  // the window buffers above belong to a clw_workspace (clw_batch_set_workspace)
  // and are not the batch's to free
  bool ws_borrowed = false;
  const clw_workspace* ws = nullptr;
  bool any_mfma = false;
  SegHost* seg_d = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool iterated = false, finalized = false;
  clw_kernel_stats ks{};
  uint64_t applied = 0;  // sieving primes applied (all windows)
  // Eager mode: kMarks events per segment (start, after sieve, after the
  // matrix-core run products, after products, after reductions), read once
  // per window.
  std::vector<hipEvent_t> events;
  hipEvent_t done_ev = nullptr;  // recorded at the end of each launched window
  uint32_t ev_used = 0;
  double wall_ms = 0;  // host wall time of clw_batch_run calls
  uint64_t next_q = 2;  // sieving primes [2, next_q) applied to every range
  uint64_t bytes = 0;   // device memory held
  // Pinned staging for a window's uploads (plans, work lists, the segment
  // descriptor) and for the counters it hands back: a copy from pageable
  // memory blocks the host until the device runs it, which, with other
  // streams' kernels holding every CU, can take a whole window of theirs.
  // Bytes [0, kStageBack) take the flags and counters of the last window;
  // uploads go after them, and stage_ev follows the last one.
  uint8_t* stage = nullptr;
  size_t stage_cap = 0, stage_off = 0;
  hipEvent_t stage_ev = nullptr;
  bool back_staged = false;  // the last window's flags and counters are copied into stage
  // creation: the tables' uploads (pinned, released once ready_ev has passed)
  // and the event after them and the setup kernels
  uint8_t* create_stage = nullptr;
  hipEvent_t ready_ev = nullptr;
  struct {
    bool active = false;
    uint64_t qlo = 0, qhi = 0;
    size_t nsegs = 0;
    std::chrono::steady_clock::time_point t0;
  } pending;            // window launched by clw_batch_launch, not yet waited for

  // Waits for the batch's own device work, not for its stream (which other
  // batches may share): every call that enqueues work either waits for it
  // before it returns or ends it with an event (done_ev after a window,
  // fin_ev after a finish), so those events cover all of it.
  // (with_ready false: creation's work, ready_ev, is left to the caller)
  bool quiesce(bool with_ready = true) const {
    bool ok = !done_ev || hipEventSynchronize(done_ev) == hipSuccess;
    ok &= !stage_ev || hipEventSynchronize(stage_ev) == hipSuccess;
    ok &= !with_ready || !ready_ev || hipEventSynchronize(ready_ev) == hipSuccess;
    for (const RangeHost& R : ranges)
      if (R.fin_ev) ok &= hipEventSynchronize(R.fin_ev) == hipSuccess;
    return ok;
  }

  ~clw_batch() {
    // teardown: errors here have nobody to report to. Once the batch's work
    // is done, its blocks go back to the cache of idle device memory instead
    // of hipFree, which would wait for every stream of the device.
    const bool drained = !ctx || quiesce();
    auto give = [&](void* x) {
      if (drained)
        (void)clwi_release(ctx, x);
      else
        (void)clw_free(ctx, x);
    };
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    if (done_ev) (void)hipEventDestroy(done_ev);
    if (stage_ev) (void)hipEventDestroy(stage_ev);
    if (ready_ev) (void)hipEventDestroy(ready_ev);
    if (drained) {  // else a copy may still read them
      clwi_host_release(stage);
      clwi_host_release(create_stage);
    }
    for (RangeHost& R : ranges) {
      if (R.finishing) {
        give(R.fin_d);
        clwi_host_release(R.fin_h);
      }
      if (R.fin_ev) (void)hipEventDestroy(R.fin_ev);
    }
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (Band& b : bands) {
      if (!b.L) continue;
      void* tab[] = {b.tab.p, b.tab.tp_target, b.tab.n_lo, b.tab.n_hi, b.tab.one_lo, b.tab.one_hi, b.tab.r2_lo,
                     b.tab.r2_hi, b.tab.acc_lo, b.tab.acc_hi, b.tab.type, b.tab.ninv};
      for (void* x : tab) give(x);
      void* ptrs[] = {b.groups_d, b.tp_group_d, b.item_gt_d, b.from_d, b.plan_d, b.mt_d, b.pm_lo, b.pm_hi};
      for (void* x : ptrs) give(x);
      if (b.ctab) (void)hipFree(b.ctab);
      if (b.part_own) {
        give(b.part_lo);
        give(b.part_hi);
      }
    }
    void* own[] = {flags, totals, seg_d};
    for (void* x : own) give(x);
    if (!ws_borrowed) {
      void* win[] = {rec, nrec, subp, subn, subb, subq, subm, rowlist, rowpos, mrows, era};
      for (void* x : win) give(x);
    }
  }
};

// bytes of the run products' row buffer: every sub-chunk slot of the record
// buffer as a row, in whole row blocks, and the step k_mfma_runs loads past
// the last one
static size_t subb_bytes(const clw_batch* B) { return mfma_rows_bytes((size_t)B->max_blocks * kBatchSubCap); }

// An upload through the pinned staging area: the bytes are copied there and
// the device copy is enqueued without blocking the host (a blocking copy
// from the source when the area is full or missing). Uploads only append
// until stage_reset, which first waits for the ones before it (normally long
// done: they precede the last window's kernels).
static int stage_h2d(clw_batch* B, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return CLW_OK;
  if (!B->stage || B->stage_off + bytes > B->stage_cap) return clw_h2d(B->ctx, dst, src, bytes, 1);
  uint8_t* const at = B->stage + B->stage_off;
  memcpy(at, src, bytes);
  CK(clw_hip(hipMemcpyAsync(dst, at, bytes, hipMemcpyHostToDevice, clw_stream(B->ctx)), "staged upload"));
  if (!B->stage_ev) CK(clw_hip(hipEventCreateWithFlags(&B->stage_ev, hipEventDisableTiming), "hipEventCreate"));
  CK(clw_hip(hipEventRecord(B->stage_ev, clw_stream(B->ctx)), "hipEventRecord"));
  B->stage_off += (bytes + 15) & ~(size_t)15;
  return CLW_OK;
}

static int stage_reset(clw_batch* B) {
  if (B->stage_ev) CK(clw_hip(hipEventSynchronize(B->stage_ev), "hipEventSynchronize"));
  B->stage_off = kStageBack;
  return CLW_OK;
}

extern "C" int clw_batch_create(clw_ctx* ctx, const uint64_t* pmin, const uint64_t* pmax, uint32_t nranges,
                                const clw_options* opt, clw_batch** out) {
  if (!ctx || !pmin || !pmax || !out || nranges == 0) return clw_fail(CLW_ERR_ARG, "clw_batch_create: bad argument");
  if (clwi_batch_seg_size() != sizeof(SegHost) || clwi_batch_group_size() != sizeof(GroupHost) ||
      clwi_batch_band_size() != sizeof(PlanHost))
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
  }
  if (B->opt.host_primes) {  // test primes on up to 16 host threads
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
      for (uint32_t r; (r = next.fetch_add(1)) < nranges;) {
        RangeHost& R = B->ranges[r];
        std::vector<uint64_t> v(clw_host_primes(R.pmin, R.pmax, nullptr, 0));
        clw_host_primes(R.pmin, R.pmax, v.data(), v.size());
        R.tp.swap(v);
      }
    };
    const uint32_t nt = std::min<uint32_t>(nranges, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (uint32_t k = 1; k < nt; ++k) pool.emplace_back(work);
    work();
    for (std::thread& t : pool) t.join();
  } else {
    for (uint32_t r = 0; r < nranges; ++r) {  // test primes sieved on the device
      RangeHost& R = B->ranges[r];
      std::vector<uint64_t> v((R.pmax - R.pmin) / 2 + 4);
      uint64_t n = 0;
      const int rc = clw_device_primes(ctx, R.pmin, R.pmax, v.data(), v.size(), &n);
      if (rc != CLW_OK) {
        delete B;
        return rc;
      }
      v.resize(n);
      R.tp.swap(v);
    }
  }
  for (uint32_t r = 0; r < nranges; ++r) {
    RangeHost& R = B->ranges[r];
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
    R.band_off = (uint32_t)bd.p.size();
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
  const uint32_t span_blocks =
      B->opt.segment_span ? (uint32_t)std::min<uint64_t>((B->opt.segment_span + CLW_BLOCK_SPAN - 1) / CLW_BLOCK_SPAN,
                                                         kMaxSpanBlocks)
                          : kMaxSpanBlocks;
  // no segment needs more blocks than cover [0, maxT]
  const uint32_t need = (uint32_t)std::min<uint64_t>(kMaxSpanBlocks, B->maxT / CLW_BLOCK_SPAN + 2);
  B->max_blocks = std::max<uint32_t>(kMinSpanBlocks, std::min(span_blocks, need)) + 2;  // + the slot for 2, 3, 5
  auto balloc = [&](size_t sz, void** ptr) {
    B->bytes += sz;
    return clw_alloc(ctx, sz, ptr);
  };
  // window buffers: the batch's own, or (external_workspace) a shared
  // workspace's, given by clw_batch_set_workspace before the first window
  const bool own_win = !B->opt.external_workspace;
  int rc = CLW_OK;
  if (own_win) rc = balloc((size_t)B->max_blocks * CLW_REC_CAP * CLW_REC_WORDS * 4, (void**)&B->rec);
  // every segment ends at most at maxT + 1: below 2^50 the segments are
  // sieved by Eratosthenes (the table of sieving primes built here, before
  // any graph capture), else by trial division and Miller-Rabin
  if (rc == CLW_OK && B->maxT + 1 <= clwi_era_hi_max()) {
    if (own_win) rc = balloc((size_t)B->max_blocks * (CLW_BLOCK_SPAN / 30), (void**)&B->era);
    if (rc == CLW_OK) rc = clwi_era_prepare(ctx, B->maxT + 1);
  }
  if (rc == CLW_OK && own_win) rc = balloc((size_t)B->max_blocks * 4, (void**)&B->nrec);
  if (rc == CLW_OK) rc = balloc(16, (void**)&B->flags);
  if (rc == CLW_OK) rc = balloc(kTotalsBytes, (void**)&B->totals);
  if (rc == CLW_OK) rc = balloc(sizeof(SegHost), (void**)&B->seg_d);
  // run products: 2 kBatchSubRecs sieving primes below maxT + 1 take at most
  // 2 kBatchSubRecs bits(maxT) bits (sized with one bit more per prime)
  {
    const uint32_t bits = (B->maxT ? 64 - __builtin_clzll(B->maxT) : 1) + 1;
    const uint32_t digits = (2 * kBatchSubRecs * bits + LB_DIGIT - 1) / LB_DIGIT;
    B->ds = std::min<uint32_t>(kBatchSubMaxDigits, std::max<uint32_t>(12, (digits + 11) / 12 * 12));
  }
  if (rc == CLW_OK && own_win) rc = balloc((size_t)B->max_blocks * kBatchSubCap * B->ds * 4, (void**)&B->subp);
  if (rc == CLW_OK && own_win) rc = balloc((size_t)B->max_blocks * kBatchSubCap * 4, (void**)&B->subn);
  if (B->opt.run_products > 2) {
    const uint32_t v = B->opt.run_products;
    delete B;
    return clw_fail(CLW_ERR_ARG, "clw_options.run_products %u: 0, 1 or 2", v);
  }
  if (B->opt.mfma_slices && B->opt.mfma_slices != 8 && B->opt.mfma_slices != 16 && B->opt.mfma_slices != 32) {
    const uint32_t v = B->opt.mfma_slices;
    delete B;
    return clw_fail(CLW_ERR_ARG, "clw_options.mfma_slices %u: 0, 8, 16 or 32", v);
  }
  // Matrix-core run products: the 3-limb band, and the 4-limb band while its
  // test primes stay below 2^47 (N < 2^94: the same 12-byte tables, and
  // sub-chunk products of primes below its targets, < 2^46, within kMfmaKB
  // bytes)
  B->any_mfma = false;
  for (int L = 3; L <= 4 && B->opt.run_products != 1; ++L) {
    Band& bd = B->bands[L];
    if (bd.L != L) continue;
    uint64_t pmax = 0;
    for (uint64_t p : bd.p) pmax = std::max(pmax, p);
    if (L == 4 && pmax >= kMfmaMaxP4) continue;
    bd.mfma = true;
    // lazy run accumulator: 2 p^2 < 0.9996 R for every test prime of the band
    // (R = 2^90 at 3 limbs; at 4 limbs N < 2^94 is far below R = 2^120)
    bd.mfma_lazy = L == 4 || 2.0 * (double)pmax * (double)pmax < 0.9996 * std::ldexp(1.0, 90);
    B->any_mfma = true;
  }
  if (B->any_mfma && own_win) {
    if (rc == CLW_OK) rc = balloc(subb_bytes(B), (void**)&B->subb);
    if (rc == CLW_OK) rc = balloc((size_t)B->max_blocks * kBatchSubCap * 16, &B->subq);
    if (rc == CLW_OK) rc = balloc((size_t)B->max_blocks * kBatchSubCap * 4, (void**)&B->subm);
    if (rc == CLW_OK) rc = balloc((size_t)B->max_blocks * kBatchSubCap * 4, (void**)&B->rowlist);
    if (rc == CLW_OK) rc = balloc((size_t)B->max_blocks * kBatchSubCap * 4, (void**)&B->rowpos);
    if (rc == CLW_OK) rc = balloc(4, (void**)&B->mrows);
  }
  if (rc == CLW_OK) rc = clw_memset(ctx, B->flags, 0, 16);
  if (rc == CLW_OK) rc = clw_memset(ctx, B->totals, 0, kTotalsBytes);
  size_t stage_bytes = kStageBack + 2 * (sizeof(SegHost) + 16) + 4096;  // the staging area (two windows' uploads)
  // the tables' uploads, enqueued together after the loop from one pinned
  // block (a copy from pageable memory would block until the device runs it)
  struct Upload {
    void* dst;
    const void* src;
    size_t bytes;
  };
  std::vector<Upload> uploads;
  std::vector<std::vector<uint32_t>> upload_words;  // sources that live only here
  for (Band& bd : B->bands) {
    if (!bd.L || rc != CLW_OK) continue;
    bd.per_lane = B->opt.per_lane ? B->opt.per_lane : clw_batch_default_per_lane((uint32_t)bd.L);
    const uint32_t n = (uint32_t)bd.p.size();
    if (bd.groups.size() >= 65536) {
      rc = clw_fail(CLW_ERR_ARG, "too many ranges in one limb band of a batch (%zu groups)", bd.groups.size());
      break;
    }
    size_t tiles = 0;  // at one test prime per lane (the most items any plan has)
    for (const GroupHost& g : bd.groups) tiles += (g.tp_n + 255) / 256;
    const uint64_t chunks = std::min<uint32_t>(B->opt.max_chunks ? B->opt.max_chunks : kMaxChunks, kMaxChunks);
    bd.part_cap = std::max<uint64_t>(n, std::min<uint64_t>(chunks * n, 1ull << 26));
    rc = clwi_batch_resident(ctx, (uint32_t)bd.L, bd.per_lane, &bd.resident);
    if (rc == CLW_OK && B->opt.target_waves) bd.resident = std::max(1u, B->opt.target_waves / 4);
    if (rc == CLW_OK) rc = clw_tp_table_alloc(ctx, n, (uint32_t)bd.L, &bd.tab);
    B->bytes += (uint64_t)n * kTableBytesPerPrime;
    if (rc == CLW_OK) {
      uploads.push_back({bd.tab.p, bd.p.data(), (size_t)n * 8});
      uploads.push_back({bd.tab.type, bd.type.data(), (size_t)n * 4});
    }
    if (rc == CLW_OK) rc = balloc(bd.groups.size() * sizeof(GroupHost), (void**)&bd.groups_d);
    if (rc == CLW_OK) rc = balloc((size_t)n * 4, (void**)&bd.tp_group_d);
    if (rc == CLW_OK) uploads.push_back({bd.tp_group_d, bd.tp_group.data(), (size_t)n * 4});
    if (rc == CLW_OK) rc = balloc(tiles * 4, (void**)&bd.item_gt_d);
    if (rc == CLW_OK) rc = balloc(sizeof(PlanHost), (void**)&bd.plan_d);
    if (rc == CLW_OK) rc = balloc((size_t)n * 8, (void**)&bd.from_d);
    if (rc == CLW_OK) uploads.push_back({bd.from_d, bd.from.data(), (size_t)n * 8});
    if (bd.mfma && rc == CLW_OK) {
      for (uint32_t gi = 0; gi < bd.groups.size(); ++gi) {
        const GroupHost& g = bd.groups[gi];
        for (uint32_t t0 = 0; t0 < g.tp_n; t0 += kMfmaTile) {
          bd.mt_group.push_back(gi);
          bd.mt_first.push_back(g.tp_off + t0);
          bd.mt_n.push_back(std::min(kMfmaTile, g.tp_n - t0));
        }
      }
      const uint32_t nm = bd.n_mt();
      std::vector<uint32_t> words(4 * (size_t)nm + 1, 0);
      std::copy(bd.mt_group.begin(), bd.mt_group.end(), words.begin());
      std::copy(bd.mt_first.begin(), bd.mt_first.end(), words.begin() + nm);
      std::copy(bd.mt_n.begin(), bd.mt_n.end(), words.begin() + 2 * (size_t)nm);
      rc = balloc(words.size() * 4, (void**)&bd.mt_d);
      if (rc == CLW_OK) {
        upload_words.push_back(std::move(words));
        uploads.push_back({bd.mt_d, upload_words.back().data(), upload_words.back().size() * 4});
      }
      // row slices of k_mfma_runs (16 B of partial per slice and test prime)
      bd.nslices = B->opt.mfma_slices ? B->opt.mfma_slices : kMfmaSlices;
      if (rc == CLW_OK) rc = balloc((size_t)bd.nslices * n * 8, (void**)&bd.pm_lo);
      if (rc == CLW_OK) rc = balloc((size_t)bd.nslices * n * 8, (void**)&bd.pm_hi);
      bd.pm_slices = bd.nslices;
    }
    // a window's uploads: groups, tiles, active matrix-core tiles and their
    // count, the plan twice, the count again
    stage_bytes += 2 * (bd.groups.size() * sizeof(GroupHost) + tiles * 4 + ((size_t)bd.n_mt() + 1) * 4 +
                        3 * sizeof(PlanHost) + 8 * 16);
    if (B->opt.external_scratch) {  // chunk products in a caller's buffer (clw_batch_set_scratch)
      bd.part_cap = 0;
      bd.part_own = false;
    } else {
      if (rc == CLW_OK) rc = balloc((size_t)bd.part_cap * 8, (void**)&bd.part_lo);
      if (rc == CLW_OK) rc = balloc((size_t)bd.part_cap * 8, (void**)&bd.part_hi);
    }
  }
  if (rc == CLW_OK && clwi_host_alloc(stage_bytes, (void**)&B->stage) == CLW_OK) {
    B->stage_cap = stage_bytes;
    B->stage_off = kStageBack;
  } else if (rc == CLW_OK) {  // without it, uploads and counters take blocking copies
    B->stage = nullptr;
  }
  // The uploads through one pinned block, then the Montgomery setup of each
  // band (it reads p and type); ready_ev follows. Nothing here waits for the
  // device: a request loop builds batches on one stream while others run,
  // and clw_batch_bind makes the batch's next stream wait for ready_ev.
  size_t up_total = 0;
  for (const Upload& u : uploads) up_total += (u.bytes + 15) & ~(size_t)15;
  if (rc == CLW_OK && clwi_host_alloc(up_total, (void**)&B->create_stage) != CLW_OK) B->create_stage = nullptr;
  size_t off = 0;
  for (const Upload& u : uploads) {
    if (rc != CLW_OK) break;
    if (B->create_stage) {
      memcpy(B->create_stage + off, u.src, u.bytes);
      rc = clw_hip(hipMemcpyAsync(u.dst, B->create_stage + off, u.bytes, hipMemcpyHostToDevice, clw_stream(ctx)),
                   "table upload");
      off += (u.bytes + 15) & ~(size_t)15;
    } else {
      rc = clw_h2d(ctx, u.dst, u.src, u.bytes, 1);
    }
  }
  for (Band& bd : B->bands)
    if (bd.L && rc == CLW_OK) rc = clw_setup(ctx, &bd.tab, 0);
  if (rc == CLW_OK) rc = clw_hip(hipEventCreateWithFlags(&B->ready_ev, hipEventDisableTiming), "hipEventCreate");
  if (rc == CLW_OK) rc = clw_hip(hipEventRecord(B->ready_ev, clw_stream(ctx)), "hipEventRecord");
  if (rc != CLW_OK) {
    delete B;
    return rc;
  }
  *out = B;
  return CLW_OK;
}

// Moves the batch to another context (stream) of the same device: later
// windows, waits and finishes run there. No window may be in flight.
extern "C" int clw_batch_bind(clw_batch* B, clw_ctx* ctx) {
  if (!B || !ctx) return clw_fail(CLW_ERR_ARG, "clw_batch_bind: null argument");
  if (B->pending.active) return clw_fail(CLW_ERR_STATE, "clw_batch_bind with a window in flight");
  if (ctx->device != B->ctx->device) return clw_fail(CLW_ERR_ARG, "clw_batch_bind: another device");
  if (!B->quiesce(false)) return clw_fail(CLW_ERR_HIP, "clw_batch_bind: waiting for the batch's work failed");
  // creation's uploads and setup: the new stream waits for them on the device
  if (B->ready_ev) CK(clw_hip(hipStreamWaitEvent(clw_stream(ctx), B->ready_ev, 0), "hipStreamWaitEvent"));
  B->ctx = ctx;
  return CLW_OK;
}

extern "C" int clw_batch_destroy(clw_batch* B) {
  delete B;
  return CLW_OK;
}

extern "C" int clw_workspace_create(clw_ctx* ctx, uint64_t segment_span, clw_workspace** out) {
  if (!ctx || !out) return clw_fail(CLW_ERR_ARG, "clw_workspace_create: null argument");
  *out = nullptr;
  clw_workspace* W = new clw_workspace();
  W->ctx = ctx;
  const uint32_t span_blocks =
      segment_span ? (uint32_t)std::min<uint64_t>((segment_span + CLW_BLOCK_SPAN - 1) / CLW_BLOCK_SPAN, kMaxSpanBlocks)
                   : kMaxSpanBlocks;
  W->max_blocks = std::max<uint32_t>(kMinSpanBlocks, span_blocks) + 2;  // as clw_batch_create sizes them
  const size_t slots = (size_t)W->max_blocks * kBatchSubCap;
  auto walloc = [&](size_t sz, void** ptr) {
    W->bytes += sz;
    return clw_alloc(ctx, sz, ptr);
  };
  int rc = walloc((size_t)W->max_blocks * CLW_REC_CAP * CLW_REC_WORDS * 4, (void**)&W->rec);
  if (rc == CLW_OK) rc = walloc((size_t)W->max_blocks * (CLW_BLOCK_SPAN / 30), (void**)&W->era);
  if (rc == CLW_OK) rc = walloc((size_t)W->max_blocks * 4, (void**)&W->nrec);
  if (rc == CLW_OK) rc = walloc(slots * kBatchSubMaxDigits * 4, (void**)&W->subp);
  if (rc == CLW_OK) rc = walloc(slots * 4, (void**)&W->subn);
  if (rc == CLW_OK) rc = walloc(mfma_rows_bytes(slots), (void**)&W->subb);
  if (rc == CLW_OK) rc = walloc(slots * 16, &W->subq);
  if (rc == CLW_OK) rc = walloc(slots * 4, (void**)&W->subm);
  if (rc == CLW_OK) rc = walloc(slots * 4, (void**)&W->rowlist);
  if (rc == CLW_OK) rc = walloc(slots * 4, (void**)&W->rowpos);
  if (rc == CLW_OK) rc = walloc(4, (void**)&W->mrows);
  if (rc != CLW_OK) {
    clw_workspace_destroy(W);
    return rc;
  }
  *out = W;
  return CLW_OK;
}

// The caller has destroyed (or rebound) every batch that used the workspace.
extern "C" int clw_workspace_destroy(clw_workspace* W) {
  if (!W) return CLW_OK;
  const bool drained = hipStreamSynchronize(clw_stream(W->ctx)) == hipSuccess;
  void* ptrs[] = {W->rec, W->era, W->nrec, W->subp, W->subn, W->subb, W->subq, W->subm, W->rowlist, W->rowpos, W->mrows};
  for (void* x : ptrs) (void)(drained ? clwi_release(W->ctx, x) : clw_free(W->ctx, x));
  delete W;
  return CLW_OK;
}

extern "C" uint64_t clw_workspace_bytes(const clw_workspace* W) { return W ? W->bytes : 0; }

// The batch's windows use W's buffers from now on (its own, if it had any,
// go back to the cache once its work is done).
extern "C" int clw_batch_set_workspace(clw_batch* B, clw_workspace* W) {
  if (!B || !W) return clw_fail(CLW_ERR_ARG, "clw_batch_set_workspace: null argument");
  if (B->pending.active) return clw_fail(CLW_ERR_STATE, "clw_batch_set_workspace with a window in flight");
  // one stream orders the windows of every batch in the workspace's buffers
  if (W->ctx != B->ctx)
    return clw_fail(CLW_ERR_ARG, "clw_batch_set_workspace: the workspace belongs to another context (stream)");
  if (W->max_blocks < B->max_blocks)
    return clw_fail(CLW_ERR_ARG, "workspace of %u blocks < the batch's %u", W->max_blocks, B->max_blocks);
  if (!B->ws_borrowed && B->rec) {  // the batch's own buffers
    if (!B->quiesce()) return clw_fail(CLW_ERR_HIP, "clw_batch_set_workspace: waiting for the batch's work failed");
    void* win[] = {B->rec, B->nrec, B->subp, B->subn, B->subb, B->subq, B->subm, B->rowlist, B->rowpos, B->mrows, B->era};
    for (void* x : win) (void)clwi_release(B->ctx, x);
  }
  B->rec = W->rec;
  B->nrec = W->nrec;
  B->subp = W->subp;
  B->subn = W->subn;
  B->era = B->maxT + 1 <= clwi_era_hi_max() ? W->era : nullptr;
  B->subb = B->any_mfma ? W->subb : nullptr;
  B->subq = W->subq;
  B->subm = W->subm;
  B->rowlist = W->rowlist;
  B->rowpos = W->rowpos;
  B->mrows = W->mrows;
  B->ws_borrowed = true;
  B->ws = W;
  if (B->graph_exec) {  // captured with the old buffers
    CK(clw_hip(hipGraphExecDestroy(B->graph_exec), "hipGraphExecDestroy"));
    CK(clw_hip(hipGraphDestroy(B->graph), "hipGraphDestroy"));
    B->graph_exec = nullptr;
    B->graph = nullptr;
  }
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

// Chunk products in a caller's scratch of `cap` slots (lo and hi words),
// shared by batches that run one after another on one stream; split among
// the batch's limb bands by their test primes.
extern "C" int clw_batch_set_scratch(clw_batch* B, uint64_t* part_lo, uint64_t* part_hi, uint64_t cap) {
  if (!B || !part_lo || !part_hi) return clw_fail(CLW_ERR_ARG, "clw_batch_set_scratch: null argument");
  if (B->pending.active) return clw_fail(CLW_ERR_STATE, "clw_batch_set_scratch with a window in flight");
  uint64_t total = 0;
  for (const Band& bd : B->bands) total += bd.L ? bd.tab.count : 0;
  if (cap < total) return clw_fail(CLW_ERR_ARG, "scratch of %llu slots < %llu test primes", (unsigned long long)cap,
                                   (unsigned long long)total);
  uint64_t off = 0;
  for (Band& bd : B->bands) {
    if (!bd.L) continue;
    if (bd.part_own) {
      clw_free(B->ctx, bd.part_lo);
      clw_free(B->ctx, bd.part_hi);
      B->bytes -= 16 * bd.part_cap;
      bd.part_own = false;
    }
    const uint64_t share = (uint64_t)((double)cap * bd.tab.count / total);
    bd.part_cap = std::max<uint64_t>(bd.tab.count, std::min<uint64_t>(share, cap - off));
    bd.part_lo = part_lo + off;
    bd.part_hi = part_hi + off;
    off += bd.part_cap;
  }
  if (off > cap) return clw_fail(CLW_ERR_ARG, "scratch too small for the bands of this batch");
  // kernel arguments changed: capture the segment step again
  if (B->graph_exec) CK(clw_hip(hipGraphExecDestroy(B->graph_exec), "hipGraphExecDestroy"));
  if (B->graph) CK(clw_hip(hipGraphDestroy(B->graph), "hipGraphDestroy"));
  B->graph_exec = nullptr;
  B->graph = nullptr;
  return CLW_OK;
}

extern "C" uint32_t clw_batch_default_per_lane(uint32_t limbs) { return limbs <= 2 ? 4 : limbs == 3 ? 3 : 2; }

// Work plan of a window whose first sieving prime is q0: the groups with
// target >= q0 get items (their tiles), the others none; K chunks per
// segment so that items x K covers kItemsPerBlock per resident block, within
// the chunk-product slots. Returns the test primes active.
static int plan_window(clw_batch* B, uint64_t q0, uint64_t* active_total) {
  uint64_t act = 0;
  for (int band = 2; band <= 4; ++band) {
    Band& bd = B->bands[band];
    if (!bd.L) continue;
    bd.item_gt.clear();
    bd.active_tp = 0;
    for (uint32_t gi = 0; gi < bd.groups.size(); ++gi) {
      const GroupHost& g = bd.groups[gi];
      if (g.T < q0) continue;
      const uint32_t tiles = (g.tp_n + 256 * bd.per_lane - 1) / (256 * bd.per_lane);
      for (uint32_t tile = 0; tile < tiles; ++tile) bd.item_gt.push_back((gi << 16) | tile);
      bd.active_tp += g.tp_n;
    }
    const uint32_t ipc = (uint32_t)bd.item_gt.size();
    uint32_t K = 1;
    if (ipc) {
      const uint64_t want = ((uint64_t)bd.items_per_block * bd.resident + ipc - 1) / ipc;
      const uint64_t cap = std::max<uint64_t>(1, bd.part_cap / bd.active_tp);
      K = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({want, cap, (uint64_t)kMaxChunks}));
    }
    uint64_t off = 0;
    for (GroupHost& g : bd.groups) {
      if (g.T < q0) {
        g.part_off = kNoPart;
      } else {
        g.part_off = off;
        off += (uint64_t)K * g.tp_n;
      }
    }
    bd.plan.items_per_chunk = ipc;
    bd.plan.K = K;
    bd.plan.mfma = 0;  // set by mfma_window once the segment span is known
    CK(stage_h2d(B, bd.groups_d, bd.groups.data(), bd.groups.size() * sizeof(GroupHost)));
    if (ipc) CK(stage_h2d(B, bd.item_gt_d, bd.item_gt.data(), (size_t)ipc * 4));
    if (bd.mfma) {  // the tiles of the active groups, then their count
      const uint32_t nm = bd.n_mt();
      bd.mt_active.clear();
      for (uint32_t mt = 0; mt < nm; ++mt)
        if (bd.groups[bd.mt_group[mt]].T >= q0) bd.mt_active.push_back(mt);
      bd.mt_active.push_back((uint32_t)bd.mt_active.size());
      const size_t na = bd.mt_active.size() - 1;
      if (na) CK(stage_h2d(B, bd.mt_d + 3 * (size_t)nm, bd.mt_active.data(), na * 4));
      CK(stage_h2d(B, bd.mt_d + 4 * (size_t)nm, &bd.mt_active[na], 4));
    }
    CK(stage_h2d(B, bd.plan_d, &bd.plan, sizeof(PlanHost)));
    act += bd.active_tp;
  }
  *active_total = act;
  return CLW_OK;
}

// Run products of the window on the matrix cores or on the vector units: a
// workgroup of k_mfma_runs loads (or builds) a 147 KB table for its 16 test
// primes, which pays off over about a thousand sub-chunks of its slice; short
// segments (large batches at the default segment_apps) keep k_batch_mul's
// reduction. Sets the plan flag both kernels read and the tile count.
// The byte tables in HBM (9.2 KB per test prime) up to kMfmaTableCap, made
// when a window (or a tuning trial) first uses the matrix cores; without them
// (run_products 2, larger, or no memory) k_mfma_runs builds each tile's table
// in LDS.
static int mfma_tables(clw_batch* B, Band& bd) {
  if (bd.ctab_tried) return CLW_OK;
  bd.ctab_tried = true;
  const uint32_t nm = bd.n_mt();
  if (B->opt.run_products == 0 && (uint64_t)nm * kMfmaTileBytes <= kMfmaTableCap &&
      hipMalloc(&bd.ctab, (size_t)nm * kMfmaTileBytes) == hipSuccess) {
    B->bytes += (uint64_t)nm * kMfmaTileBytes;
    CK(clwi_mfma_table(B->ctx, &bd.tab, bd.mt_d + nm, bd.mt_d + 2 * (size_t)nm, nm, bd.ctab));
    if (B->graph_exec) {  // captured without the tables: capture the segment step again
      CK(clw_hip(hipGraphExecDestroy(B->graph_exec), "hipGraphExecDestroy"));
      CK(clw_hip(hipGraphDestroy(B->graph), "hipGraphDestroy"));
      B->graph_exec = nullptr;
      B->graph = nullptr;
    }
  } else {
    bd.ctab = nullptr;
    (void)hipGetLastError();  // a failed allocation is not an error here
  }
  return CLW_OK;
}

// Sets the matrix-core flag (the slice count, or 0 for the vector units) of
// the band's plan and the count of active tiles, both read by the kernels.
static int mfma_plan(clw_batch* B, Band& bd, uint32_t slices) {
  bd.plan.mfma = slices;
  const uint32_t nm = bd.n_mt(), na = slices ? (uint32_t)bd.mt_active.size() - 1 : 0u;
  if (slices) CK(mfma_tables(B, bd));
  CK(stage_h2d(B, bd.plan_d, &bd.plan, sizeof(PlanHost)));
  CK(stage_h2d(B, bd.mt_d + 4 * (size_t)nm, &na, 4));
  return CLW_OK;
}

static int mfma_window(clw_batch* B, uint64_t seg_len, uint64_t qhi) {
  for (Band& bd : B->bands) {
    if (!bd.L || !bd.mfma) continue;
    const double rows = (double)seg_len / std::log((double)std::max<uint64_t>(qhi, 16)) / (2.0 * kBatchSubRecs);
    const bool use = rows / bd.nslices >= bd.mfma_min_rows && bd.mt_active.size() > 1;
    CK(mfma_plan(B, bd, use ? bd.nslices : 0u));
  }
  return CLW_OK;
}

static int mark(clw_batch* B, bool timed) {
  if (!timed) return CLW_OK;
  if (B->ev_used == B->events.size()) {
    hipEvent_t e;
    const hipError_t err = hipEventCreate(&e);
    if (err != hipSuccess) return clw_fail_hip(err, "hipEventCreate");
    B->events.push_back(e);
  }
  const hipError_t err = hipEventRecord(B->events[B->ev_used++], clw_stream(B->ctx));
  return err == hipSuccess ? CLW_OK : clw_fail_hip(err, "hipEventRecord");
}

// One segment: advance the descriptor, sieve, products of every band, then
// fold every band's chunk products. `timed` brackets the phases with events
// (eager launches only: a captured graph holds no event records).
static int batch_step(clw_batch* B, uint32_t extra_block, bool timed) {
  clw_ctx* ctx = B->ctx;
  CK(mark(B, timed));
  CK(clwi_seg_advance(ctx, B->seg_d, B->flags, B->totals + 1, B->totals));
  CK(clwi_batch_sieve(ctx, B->seg_d, B->rec, B->nrec, B->max_blocks - 1, B->flags, B->totals, B->era, B->maxT + 1));
  // the row list first: the run products' bytes go to their rows
  if (B->any_mfma) CK(clwi_mfma_rows(ctx, B->seg_d, extra_block, B->nrec, B->rowlist, B->rowpos, B->mrows));
  CK(clwi_batch_subprod(ctx, B->seg_d, extra_block, B->rec, B->nrec, B->max_blocks, B->subp, B->subn, B->ds,
                        B->any_mfma ? B->subb : nullptr, B->rowpos, B->subq, B->subm, B->flags));
  CK(mark(B, timed));
  for (int band = 2; band <= 4; ++band) {
    Band& bd = B->bands[band];
    if (!bd.L || !bd.mfma) continue;
    const uint32_t nm = bd.n_mt();
    CK(clwi_mfma_runs(ctx, &bd.tab, bd.groups_d, bd.mt_d, bd.mt_d + nm, bd.mt_d + 2 * (size_t)nm,
                      bd.mt_d + 3 * (size_t)nm, bd.mt_d + 4 * (size_t)nm, nm, B->rowlist, B->mrows, B->subq, B->subm,
                      B->subb, subb_bytes(B), bd.ctab, bd.pm_lo, bd.pm_hi,
                      B->totals + 2, bd.mfma_lazy ? 1u : 0u, bd.nslices, B->seg_d));
  }
  CK(mark(B, timed));
  for (int band = 2; band <= 4; ++band) {
    Band& bd = B->bands[band];
    if (!bd.L) continue;
    CK(clwi_batch_mul(ctx, &bd.tab, bd.per_lane, bd.groups_d, bd.item_gt_d, bd.plan_d, B->rec, B->nrec,
                      B->seg_d, (uint32_t)(band - 2), extra_block, B->subp, B->subn, B->ds, bd.resident,
                      bd.part_lo, bd.part_hi, B->totals + 2));
  }
  CK(mark(B, timed));
  for (int band = 2; band <= 4; ++band) {
    Band& bd = B->bands[band];
    if (bd.L)
      CK(clwi_batch_reduce(ctx, &bd.tab, bd.tp_group_d, bd.groups_d, bd.plan_d, bd.part_lo, bd.part_hi,
                           bd.mfma ? bd.pm_lo : nullptr, bd.mfma ? bd.pm_hi : nullptr));
  }
  CK(mark(B, timed));
  return CLW_OK;
}

// Enqueues every sieving prime q in [qlo, qhi) (clipped to [2, max target])
// for every range of the batch on the batch's stream; clw_batch_wait ends
// the window. Windows must be disjoint across calls.
extern "C" int clw_batch_launch(clw_batch* B, uint64_t qlo, uint64_t qhi) {
  if (!B) return clw_fail(CLW_ERR_ARG, "clw_batch_launch: null batch");
  if (B->iterated) return clw_fail(CLW_ERR_STATE, "clw_batch_launch after iterate");
  if (B->pending.active) return clw_fail(CLW_ERR_STATE, "clw_batch_launch: previous window not waited for");
  if (!B->rec) return clw_fail(CLW_ERR_STATE, "batch without window buffers (clw_batch_set_workspace)");
  if (B->ws && B->ws->ctx != B->ctx)
    return clw_fail(CLW_ERR_STATE, "batch bound to another stream than its workspace (clw_batch_set_workspace again)");
  clw_ctx* ctx = B->ctx;
  const auto t0 = std::chrono::steady_clock::now();
  if (qhi > B->maxT + 1) qhi = B->maxT + 1;
  if (qlo < 2) qlo = 2;
  if (qhi <= qlo) return CLW_OK;
  const uint64_t lo = std::max<uint64_t>(qlo, 7);
  for (const Band& bd : B->bands)
    if (bd.L && bd.part_cap < bd.tab.count)
      return clw_fail(CLW_ERR_STATE, "batch without chunk-product slots (clw_batch_set_scratch)");
  CK(stage_reset(B));
  if (B->create_stage && hipEventQuery(B->ready_ev) == hipSuccess) {  // creation's uploads are done
    clwi_host_release(B->create_stage);
    B->create_stage = nullptr;
  }
  uint64_t active = 0;
  CK(plan_window(B, qlo, &active));
  uint64_t span_blocks = B->max_blocks - 2;
  if (active) {
    // about opt.segment_apps (default kAppsPerSegment) factor applications
    // per segment at the prime density ~1/ln q, within the record buffer
    const double apps = B->opt.segment_apps ? (double)B->opt.segment_apps : kAppsPerSegment;
    const double per_block = (double)active * CLW_BLOCK_SPAN / std::log((double)std::max<uint64_t>(qhi, 16));
    span_blocks = std::min<uint64_t>(span_blocks, std::max<uint64_t>(kMinSpanBlocks, (uint64_t)(apps / per_block) + 1));
  }
  const uint64_t span = span_blocks * CLW_BLOCK_SPAN;
  CK(mfma_window(B, std::min<uint64_t>(span, qhi - lo), qhi));
  // Host replica of k_seg_advance: the segments of [lo, qhi).
  std::vector<std::pair<uint64_t, uint64_t>> segs;
  for (uint64_t a = lo; a < qhi;) {
    uint64_t b = a + span - (a % 30);
    if (b > qhi) b = qhi;
    segs.push_back({a, b});
    a = b;
  }
  const bool small = qlo <= 5;
  if (segs.empty() && small) segs.push_back({lo, lo});  // only 2, 3, 5
  SegHost h{};
  h.lo = lo;
  h.hi = lo;  // the first advance yields [lo, ...)
  h.span = span;
  h.end = std::max(qhi, lo);
  CK(stage_h2d(B, B->seg_d, &h, sizeof(h)));
  B->ev_used = 0;
  const bool timed = !B->opt.graph;
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
      CK(batch_step(B, 1, timed));
    } else if (B->opt.graph) {
      if (!B->graph_exec) {
        hipStream_t st = clw_stream(ctx);
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return clw_fail_hip(e, "hipStreamBeginCapture");
        const int rc = batch_step(B, 0, false);
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
      CK(batch_step(B, 0, timed));
    }
    B->ks.segments++;
    B->ks.sieve_integers += segs[s].second - segs[s].first;
  }
  // folds the last segment's prime count into the window total
  CK(clwi_seg_advance(ctx, B->seg_d, B->flags, B->totals + 1, B->totals));
  // flags and counters into pinned memory, read by clw_batch_wait without a
  // copy of its own
  B->back_staged = false;
  if (B->stage) {
    CK(clw_hip(hipMemcpyAsync(B->stage, B->flags, 16, hipMemcpyDeviceToHost, clw_stream(ctx)), "counters copy"));
    CK(clw_hip(hipMemcpyAsync(B->stage + 16, B->totals, kTotalsBytes, hipMemcpyDeviceToHost, clw_stream(ctx)),
               "counters copy"));
    B->back_staged = true;
  }
  B->pending.active = true;
  if (!B->done_ev) {
    const hipError_t e = hipEventCreateWithFlags(&B->done_ev, hipEventDisableTiming);
    if (e != hipSuccess) return clw_fail_hip(e, "hipEventCreate");
  }
  const hipError_t e = hipEventRecord(B->done_ev, clw_stream(ctx));
  if (e != hipSuccess) return clw_fail_hip(e, "hipEventRecord");
  B->pending.qlo = qlo;
  B->pending.qhi = qhi;
  B->pending.nsegs = segs.size();
  B->pending.t0 = t0;
  return CLW_OK;
}

// 1 if no window is in flight or the launched one has finished on the
// device (clw_batch_wait then returns at once), 0 while it runs.
extern "C" int clw_batch_query(const clw_batch* B) {
  if (!B) return clw_fail(CLW_ERR_ARG, "clw_batch_query: null batch");
  if (!B->pending.active || !B->done_ev) return 1;
  const hipError_t e = hipEventQuery(B->done_ev);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return clw_fail_hip(e, "hipEventQuery");
}

// Waits for the launched window; checks the device flags and folds counters
// and event times into the statistics.
extern "C" int clw_batch_wait(clw_batch* B) {
  if (!B) return clw_fail(CLW_ERR_ARG, "clw_batch_wait: null batch");
  if (!B->pending.active) return CLW_OK;
  B->pending.active = false;
  clw_ctx* ctx = B->ctx;
  const uint64_t qlo = B->pending.qlo, qhi = B->pending.qhi;
  uint32_t flags[4] = {0, 0, 0, 0};
  uint64_t tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  static_assert(sizeof(tot) == kTotalsBytes && 16 + kTotalsBytes <= kStageBack, "counters in the staging area");
  if (B->back_staged) {  // copied behind the window
    CK(clw_hip(hipEventSynchronize(B->done_ev), "hipEventSynchronize"));
    memcpy(flags, B->stage, 16);
    memcpy(tot, B->stage + 16, sizeof(tot));
  } else {
    CK(clw_d2h(ctx, flags, B->flags, 16, 0));
    CK(clw_d2h(ctx, tot, B->totals, sizeof(tot), 1));
  }
  if (flags[0]) return clw_fail(CLW_ERR_OVERFLOW, "device capacity exceeded: flags %u", flags[0]);
  uint64_t smallc = 0;
  for (uint64_t q : {2ull, 3ull, 5ull}) smallc += (q >= qlo && q < qhi);
  B->applied += tot[1] + smallc;
  B->ks.mul_record_apps += tot[2];
  B->ks.mul_prime_apps += tot[3];
  B->ks.mul_partial_products += tot[4];
  B->ks.mfma_int8_macs += tot[5];
  B->ks.mfma_int8_issued += tot[6];
  B->ks.mfma_partial_products += tot[7];
  CK(clw_memset(ctx, B->totals, 0, kTotalsBytes));
  for (uint32_t k = 0; k + kMarks <= B->ev_used; k += kMarks) {
    float a = 0, m = 0, b = 0, c = 0;
    CK(clw_hip(hipEventElapsedTime(&a, B->events[k], B->events[k + 1]), "hipEventElapsedTime"));
    CK(clw_hip(hipEventElapsedTime(&m, B->events[k + 1], B->events[k + 2]), "hipEventElapsedTime"));
    CK(clw_hip(hipEventElapsedTime(&b, B->events[k + 1], B->events[k + 3]), "hipEventElapsedTime"));
    CK(clw_hip(hipEventElapsedTime(&c, B->events[k + 3], B->events[k + 4]), "hipEventElapsedTime"));
    B->ks.sieve_ms += a;
    B->ks.mfma_ms += m;
    B->ks.mul_ms += b;
    B->ks.reduce_ms += c;
    B->ks.sieve_launches++;
    B->ks.reduce_launches++;
  }
  for (int band = 2; band <= 4; ++band) B->ks.mul_launches += B->bands[band].L ? B->pending.nsegs : 0;
  B->wall_ms += std::chrono::duration<double>(std::chrono::steady_clock::now() - B->pending.t0).count() * 1e3;
  if (qlo <= B->next_q && qhi > B->next_q) B->next_q = qhi;
  return CLW_OK;
}

extern "C" int clw_batch_run(clw_batch* B, uint64_t qlo, uint64_t qhi) {
  CK(clw_batch_launch(B, qlo, qhi));
  return clw_batch_wait(B);
}

// Sieving primes [2, next_q) have reached every range (windows from 2 on,
// back to back).
extern "C" uint64_t clw_batch_covered(const clw_batch* B) { return B ? B->next_q : 0; }

extern "C" uint64_t clw_batch_device_bytes(const clw_batch* B) { return B ? B->bytes : 0; }

static int range_results(const clw_batch* B, uint32_t r, clw_range_stats* st, uint64_t* out_p, uint64_t* out_quot,
                         uint64_t* out_rem, uint64_t cap);

// One range done early: once the sieving primes up to its largest target
// have been applied, iterate and check its test primes alone (the rest of
// the batch keeps running). finish_launch enqueues that on the batch's
// stream with a copy of the results into pinned host memory and returns;
// finish_query tells whether they are in; finish_range collects them
// (launching first if nothing is in flight) and returns the range's results.
// The device buffer goes back to the idle-block cache, never to hipFree,
// which would wait for every stream of the device.
extern "C" int clw_batch_finish_launch(clw_batch* B, uint32_t r) {
  if (!B || r >= B->ranges.size()) return clw_fail(CLW_ERR_ARG, "clw_batch_finish_launch: bad argument");
  if (B->pending.active) return clw_fail(CLW_ERR_STATE, "clw_batch_finish_launch with a window in flight");
  if (B->iterated) return clw_fail(CLW_ERR_STATE, "clw_batch_finish_launch after clw_batch_iterate");
  RangeHost& R = B->ranges[r];
  if (R.finished || R.finishing) return CLW_OK;
  const uint64_t mt = std::max(R.T[0], std::max(R.T[1], R.T[2]));
  if (B->next_q <= mt)
    return clw_fail(CLW_ERR_STATE, "range %u needs sieving primes up to %llu, covered up to %llu", r,
                    (unsigned long long)mt, (unsigned long long)B->next_q);
  Band& bd = B->bands[R.band];
  const uint32_t n = (uint32_t)R.tp.size();
  const clw_tp_table v = table_view(bd.tab, R.band_off, n);
  clw_ctx* ctx = B->ctx;
  int rc = clw_alloc(ctx, (size_t)2 * n * 8, (void**)&R.fin_d);
  if (rc == CLW_OK) rc = clwi_host_alloc((size_t)2 * n * 8, (void**)&R.fin_h);
  if (rc == CLW_OK) rc = clw_iterate_from(ctx, &v, bd.from_d + R.band_off);
  if (rc == CLW_OK) rc = clw_finalize(ctx, &v, R.fin_d, R.fin_d + n, nullptr);
  if (rc == CLW_OK)
    rc = clw_hip(hipMemcpyAsync(R.fin_h, R.fin_d, (size_t)2 * n * 8, hipMemcpyDeviceToHost, clw_stream(ctx)),
                 "finish copy");
  if (rc == CLW_OK && !R.fin_ev) rc = clw_hip(hipEventCreateWithFlags(&R.fin_ev, hipEventDisableTiming), "hipEventCreate");
  if (rc == CLW_OK) rc = clw_hip(hipEventRecord(R.fin_ev, clw_stream(ctx)), "hipEventRecord");
  if (rc != CLW_OK) {
    // nothing may still use the buffers when they go back
    (void)hipStreamSynchronize(clw_stream(ctx));
    (void)clwi_release(ctx, R.fin_d);
    clwi_host_release(R.fin_h);
    R.fin_d = nullptr;
    R.fin_h = nullptr;
    return rc;
  }
  R.finishing = true;
  return CLW_OK;
}

extern "C" int clw_batch_finish_query(const clw_batch* B, uint32_t r) {
  if (!B || r >= B->ranges.size()) return clw_fail(CLW_ERR_ARG, "clw_batch_finish_query: bad argument");
  const RangeHost& R = B->ranges[r];
  if (R.finished) return 1;
  if (!R.finishing) return clw_fail(CLW_ERR_STATE, "clw_batch_finish_query: range %u not launched", r);
  const hipError_t e = hipEventQuery(R.fin_ev);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return clw_fail_hip(e, "hipEventQuery");
}

// Waits for range r's launched finish and keeps its results on the host.
static int collect_finish(clw_batch* B, RangeHost& R) {
  CK(clw_hip(hipEventSynchronize(R.fin_ev), "hipEventSynchronize"));
  const size_t n = R.tp.size();
  R.quot.assign(R.fin_h, R.fin_h + n);
  R.rem.assign(R.fin_h + n, R.fin_h + 2 * n);
  // the copy, the buffers' last use, is behind the event
  (void)clwi_release(B->ctx, R.fin_d);
  clwi_host_release(R.fin_h);
  R.fin_d = nullptr;
  R.fin_h = nullptr;
  R.finishing = false;
  R.finished = true;
  return CLW_OK;
}

extern "C" int clw_batch_finish_range(clw_batch* B, uint32_t r, clw_range_stats* st, uint64_t* out_p,
                                      uint64_t* out_quot, uint64_t* out_rem, uint64_t cap) {
  if (!B || r >= B->ranges.size()) return clw_fail(CLW_ERR_ARG, "clw_batch_finish_range: bad argument");
  RangeHost& R = B->ranges[r];
  if (!R.finished) {
    if (!R.finishing) CK(clw_batch_finish_launch(B, r));
    CK(collect_finish(B, R));
  }
  return range_results(B, r, st, out_p, out_quot, out_rem, cap);
}

extern "C" int clw_batch_iterate(clw_batch* B) {
  if (!B) return clw_fail(CLW_ERR_ARG, "clw_batch_iterate: null batch");
  if (B->iterated) return clw_fail(CLW_ERR_STATE, "clw_batch_iterate called twice");
  bool any_finished = false;
  for (const RangeHost& R : B->ranges) any_finished |= R.finished || R.finishing;
  if (!any_finished) {
    for (Band& bd : B->bands)
      if (bd.L) CK(clw_iterate_from(B->ctx, &bd.tab, bd.from_d));
  } else {
    for (const RangeHost& R : B->ranges) {
      if (R.finished || R.finishing) continue;
      Band& bd = B->bands[R.band];
      const clw_tp_table v = table_view(bd.tab, R.band_off, (uint32_t)R.tp.size());
      CK(clw_iterate_from(B->ctx, &v, bd.from_d + R.band_off));
    }
  }
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

static int range_results(const clw_batch* B, uint32_t r, clw_range_stats* st, uint64_t* out_p, uint64_t* out_quot,
                         uint64_t* out_rem, uint64_t cap) {
  if (!st) return clw_fail(CLW_ERR_ARG, "range results: null stats");
  CK(clw_batch_info(B, r, st));
  const RangeHost& R = B->ranges[r];
  const Band& bd = B->bands[R.band];
  int bad = 0;
  uint64_t checksum = 0;
  for (size_t k = 0; k < R.tp.size(); ++k) {
    const uint32_t i = R.pos[k];
    const uint64_t p = R.tp[k];
    const uint64_t quot = R.finished ? R.quot[i - R.band_off] : bd.quot[i];
    const uint64_t rem = R.finished ? R.rem[i - R.band_off] : bd.rem[i];
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
  st->totalcount = 0;  // not tracked per range in a batch
  if (bad) return clw_fail(CLW_ERR_CHECK, "Wilson quotient check failed in batch range %u", r);
  return CLW_OK;
}

// Per-range results after clw_batch_finalize (or clw_batch_finish_range), in
// increasing p; stats as clw_session_finalize fills them (checksum, results,
// last result).
extern "C" int clw_batch_results(const clw_batch* B, uint32_t r, clw_range_stats* st, uint64_t* out_p,
                                 uint64_t* out_quot, uint64_t* out_rem, uint64_t cap) {
  if (!B || !st || r >= B->ranges.size()) return clw_fail(CLW_ERR_ARG, "clw_batch_results: bad argument");
  if (!B->finalized && !B->ranges[r].finished)
    return clw_fail(CLW_ERR_STATE, "clw_batch_results before clw_batch_finalize");
  return range_results(B, r, st, out_p, out_quot, out_rem, cap);
}

// Resume with a caller's own checkpoint format (the command line keeps the
// reference's, cl_wilson.cpp:1736-1777): range r's canonical residues in
// increasing p (as clw_batch_residues returns them), holding the sieving
// primes [2, covered), `applied` of them. One-range batches, before any
// window of this batch.
extern "C" int clw_batch_set_residues(clw_batch* B, uint32_t r, const uint64_t* lo, const uint64_t* hi, uint64_t n,
                                      uint64_t covered, uint64_t applied) {
  if (!B || !lo || !hi || r >= B->ranges.size()) return clw_fail(CLW_ERR_ARG, "clw_batch_set_residues: bad argument");
  if (B->ranges.size() != 1) return clw_fail(CLW_ERR_ARG, "clw_batch_set_residues: one-range batches only");
  if (B->pending.active || B->iterated || B->next_q != 2 || B->applied != 0)
    return clw_fail(CLW_ERR_STATE, "clw_batch_set_residues: the batch has run windows already");
  const RangeHost& R = B->ranges[r];
  if (n != R.tp.size()) return clw_fail(CLW_ERR_ARG, "clw_batch_set_residues: %llu residues for %zu test primes",
                                        (unsigned long long)n, R.tp.size());
  for (uint64_t k = 0; k < n; ++k) {
    const unsigned __int128 v = ((unsigned __int128)hi[k] << 64) | lo[k];
    if (v >= (unsigned __int128)R.tp[k] * R.tp[k]) return clw_fail(CLW_ERR_ARG, "residue %llu not below p^2", (unsigned long long)k);
  }
  Band& bd = B->bands[R.band];
  const uint32_t cnt = bd.tab.count;
  std::vector<uint64_t> a(cnt), b(cnt);
  CK(clw_d2h(B->ctx, a.data(), bd.tab.acc_lo, (size_t)cnt * 8, 0));
  CK(clw_d2h(B->ctx, b.data(), bd.tab.acc_hi, (size_t)cnt * 8, 1));
  for (uint64_t k = 0; k < n; ++k) {
    a[R.pos[k]] = lo[k];
    b[R.pos[k]] = hi[k];
  }
  CK(clw_h2d(B->ctx, bd.tab.acc_lo, a.data(), (size_t)cnt * 8, 0));
  CK(clw_h2d(B->ctx, bd.tab.acc_hi, b.data(), (size_t)cnt * 8, 1));
  B->next_q = std::max<uint64_t>(covered, 2);
  B->applied = applied;
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
  out->side_ms = B->wall_ms;  // a batch has no side lists: wall time instead
  return CLW_OK;
}

// ---- autotuner -------------------------------------------------------------
//
// Per limb band, at the caller's magnitude q (the reference sizes its
// launches from a timed probe instead, cl_wilson.cpp:1184-1290):
//  1. for a band with matrix-core run products, the row slices of
//     k_mfma_runs (8 or 16) and the switch between the matrix cores and the
//     vector units: both product paths are timed on trial segments of the
//     length a window at q runs and of a quarter of it; the matrix-core time
//     is a + b r and the vector time c r in the rows r per slice, so the
//     matrix cores pay from r* = a / (c - b) rows per slice on
//     (mfma_min_rows; never if c <= b);
//  2. the test primes per lane of k_batch_mul (1, 2 or 4 Montgomery products
//     in lockstep, i.e. the tile of 256 x KP test primes a work-group
//     carries) x work items per resident block (the chunk count K, i.e. the
//     fan-in of the reduction that folds chunk products), timed on the work
//     step 1 leaves it (where the matrix cores take the run products, only
//     the records across run boundaries).
// Trials write chunk products and partials to scratch and run no reduction,
// so no residue changes. Results can be kept in a text cache keyed by device,
// band and the log2 of the band's active test primes.

namespace {

struct TuneKey {
  char arch[64];
  int cus, L, size_class;
};

struct TuneChoice {
  uint32_t kp = 0, ipb = 0, slices = 0, min_rows = 0;
};

const char kTuneTag[] = "clwtune2";  // cache lines: tag arch cus L size_class kp ipb slices min_rows

bool cache_lookup(const char* path, const TuneKey& k, TuneChoice* c) {
  if (!path) return false;
  FILE* f = fopen(path, "r");
  if (!f) return false;
  char line[256];
  bool hit = false;
  while (fgets(line, sizeof(line), f)) {
    char tag[16], arch[64];
    int cus, L, sc;
    unsigned a, b, sl, mr;
    if (sscanf(line, "%15s %63s %d %d %d %u %u %u %u", tag, arch, &cus, &L, &sc, &a, &b, &sl, &mr) != 9) continue;
    if (!strcmp(tag, kTuneTag) && !strcmp(arch, k.arch) && cus == k.cus && L == k.L && sc == k.size_class) {
      *c = TuneChoice{a, b, sl, mr};
      hit = true;  // the last matching line wins
    }
  }
  fclose(f);
  return hit;
}

void cache_store(const char* path, const TuneKey& k, const TuneChoice& c) {
  if (!path) return;
  FILE* f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s %s %d %d %d %u %u %u %u\n", kTuneTag, k.arch, k.cus, k.L, k.size_class, c.kp, c.ipb, c.slices,
          c.min_rows);
  fclose(f);
}

}  // namespace

// Segment blocks of a window at q with `active` test primes (the length
// clw_batch_launch would choose).
static uint64_t window_blocks(const clw_batch* B, uint64_t q, uint64_t active) {
  const double apps = B->opt.segment_apps ? (double)B->opt.segment_apps : kAppsPerSegment;
  const double per_block = (double)std::max<uint64_t>(active, 1) * CLW_BLOCK_SPAN /
                           std::log((double)std::max<uint64_t>(q, 16));
  return std::min<uint64_t>(B->max_blocks - 2, std::max<uint64_t>(kMinSpanBlocks, (uint64_t)(apps / per_block) + 1));
}

// Trial segment [q, q + span_blocks blocks) (clipped to maxT): sieved, run
// products formed (digits, and for the matrix cores bytes and the dense row
// list). Returns the segment's sub-chunks (rows); *desc receives the device
// segment descriptor as the product kernels find it (work queues empty), so
// that each trial on the segment can start from it.
static int trial_segment(clw_batch* B, uint64_t q, uint64_t span_blocks, uint32_t* rows, SegHost* desc) {
  clw_ctx* ctx = B->ctx;
  SegHost h{};
  h.lo = q;
  h.hi = q;
  h.span = span_blocks * CLW_BLOCK_SPAN;
  h.end = std::min(q + h.span, B->maxT + 1);  // run products are sized for primes up to maxT
  CK(clw_h2d(ctx, B->seg_d, &h, sizeof(h), 1));
  CK(clw_memset(ctx, B->totals, 0, kTotalsBytes));
  CK(clwi_seg_advance(ctx, B->seg_d, B->flags, B->totals + 1, B->totals));
  CK(clwi_batch_sieve(ctx, B->seg_d, B->rec, B->nrec, B->max_blocks - 1, B->flags, B->totals, B->era, B->maxT + 1));
  *rows = 0;
  if (B->any_mfma) CK(clwi_mfma_rows(ctx, B->seg_d, 0, B->nrec, B->rowlist, B->rowpos, B->mrows));
  CK(clwi_batch_subprod(ctx, B->seg_d, 0, B->rec, B->nrec, B->max_blocks, B->subp, B->subn, B->ds,
                        B->any_mfma ? B->subb : nullptr, B->rowpos, B->subq, B->subm, B->flags));
  if (B->any_mfma) CK(clw_d2h(ctx, rows, B->mrows, 4, 1));
  CK(clw_d2h(ctx, desc, B->seg_d, sizeof(SegHost), 1));
  return CLW_OK;
}

// The band's product kernels on the trial segment (descriptor desc), timed
// with events: k_mfma_runs with `slices` row slices (0: none, the run
// products stay on the vector units; with_mfma false: the plan only, so
// k_batch_mul skips what the matrix cores would take, and k_mfma_runs, whose
// time does not depend on k_batch_mul's shape, is not run) and k_batch_mul,
// whose work queues start empty. *ms and the factor applications they made.
static int trial_products(clw_batch* B, Band& bd, const SegHost& desc, uint32_t slices, bool with_mfma, double* ms,
                          double* apps) {
  clw_ctx* ctx = B->ctx;
  CK(clw_h2d(ctx, B->seg_d, &desc, sizeof(SegHost), 0));
  if (bd.mfma) CK(mfma_plan(B, bd, slices));
  CK(clw_memset(ctx, B->totals, 0, kTotalsBytes));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = clw_hip(hipEventCreate(&e0), "hipEventCreate");
  if (rc == CLW_OK) rc = clw_hip(hipEventCreate(&e1), "hipEventCreate");
  if (rc == CLW_OK) rc = clw_hip(hipEventRecord(e0, clw_stream(ctx)), "hipEventRecord");
  if (rc == CLW_OK && slices && with_mfma) {
    const uint32_t nm = bd.n_mt();
    rc = clwi_mfma_runs(ctx, &bd.tab, bd.groups_d, bd.mt_d, bd.mt_d + nm, bd.mt_d + 2 * (size_t)nm,
                        bd.mt_d + 3 * (size_t)nm, bd.mt_d + 4 * (size_t)nm, nm, B->rowlist, B->mrows, B->subq,
                        B->subm, B->subb, subb_bytes(B), bd.ctab, bd.pm_lo,
                        bd.pm_hi, B->totals + 2, bd.mfma_lazy ? 1u : 0u, slices, B->seg_d);
  }
  if (rc == CLW_OK)
    rc = clwi_batch_mul(ctx, &bd.tab, bd.per_lane, bd.groups_d, bd.item_gt_d, bd.plan_d, B->rec, B->nrec, B->seg_d,
                        (uint32_t)(bd.L - 2), 0, B->subp, B->subn, B->ds, bd.resident, bd.part_lo, bd.part_hi,
                        B->totals + 2);
  if (rc == CLW_OK) rc = clw_hip(hipEventRecord(e1, clw_stream(ctx)), "hipEventRecord");
  uint64_t tot[8] = {0};
  if (rc == CLW_OK) rc = clw_d2h(ctx, tot, B->totals, sizeof(tot), 1);
  float t = 0;
  if (rc == CLW_OK) rc = clw_hip(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  CK(rc);
  CK(clw_memset(ctx, B->totals, 0, kTotalsBytes));
  *ms = t;
  *apps = (double)tot[3];
  return CLW_OK;
}

// Partials of k_mfma_runs for `slices` slices (16 B per slice and test prime).
static int ensure_pm(clw_batch* B, Band& bd, uint32_t slices) {
  if (bd.pm_slices >= slices) return CLW_OK;
  const size_t n = bd.p.size();
  uint64_t *lo = nullptr, *hi = nullptr;
  CK(clw_alloc(B->ctx, (size_t)slices * n * 8, (void**)&lo));
  const int rc = clw_alloc(B->ctx, (size_t)slices * n * 8, (void**)&hi);
  if (rc != CLW_OK) {
    clw_free(B->ctx, lo);
    return rc;
  }
  clw_free(B->ctx, bd.pm_lo);
  clw_free(B->ctx, bd.pm_hi);
  B->bytes += 16ull * n * (slices - bd.pm_slices);
  bd.pm_lo = lo;
  bd.pm_hi = hi;
  bd.pm_slices = slices;
  return CLW_OK;
}

// Step 1 of the tuner for a band with matrix-core run products.
static int tune_mfma(clw_batch* B, Band& bd, uint64_t q, uint64_t active, TuneChoice* ch, clw_band_tune* t) {
  std::vector<uint32_t> cand;
  if (B->opt.mfma_slices) {
    cand.push_back(B->opt.mfma_slices);
  } else {
    cand.push_back(8);
    // 16 slices: twice the partials (kept below 1 GiB)
    if (32ull * bd.p.size() * 8 <= (1ull << 30)) cand.push_back(16);
  }
  for (uint32_t sl : cand) CK(ensure_pm(B, bd, sl));
  const uint64_t full = window_blocks(B, q, active);
  const uint64_t part = std::max<uint64_t>(kMinSpanBlocks, full / 4);
  double tm[2][4] = {{0}}, tv[2] = {0, 0};
  uint32_t rows[2] = {0, 0};
  bool warmed = false;
  for (int li = 0; li < 2; ++li) {
    SegHost desc{};
    CK(trial_segment(B, q, li == 0 ? full : part, &rows[li], &desc));
    double ms = 0, apps = 0;
    if (!warmed) {  // code objects, tables, clocks
      for (uint32_t sl : cand) CK(trial_products(B, bd, desc, sl, true, &ms, &apps));
      warmed = true;
    }
    CK(trial_products(B, bd, desc, 0, true, &tv[li], &apps));
    if (t->mfma_trials < 8) {
      t->mfma_trial_slices[t->mfma_trials] = 0;
      t->mfma_trial_rows[t->mfma_trials] = rows[li];
      t->mfma_trial_ms[t->mfma_trials] = tv[li];
    }
    t->mfma_trials++;
    for (size_t ci = 0; ci < cand.size(); ++ci) {
      CK(trial_products(B, bd, desc, cand[ci], true, &tm[li][ci], &apps));
      if (t->mfma_trials < 8) {
        t->mfma_trial_slices[t->mfma_trials] = cand[ci];
        t->mfma_trial_rows[t->mfma_trials] = rows[li];
        t->mfma_trial_ms[t->mfma_trials] = tm[li][ci];
      }
      t->mfma_trials++;
    }
  }
  size_t best = 0;
  for (size_t ci = 1; ci < cand.size(); ++ci)
    if (tm[0][ci] < tm[0][best]) best = ci;
  ch->slices = cand[best];
  // crossover in rows per slice
  const double r0 = (double)rows[0] / ch->slices, r1 = (double)rows[1] / ch->slices;
  double min_rows = 4294967295.0;
  if (r0 > r1 && r1 > 0) {
    const double b = (tm[0][best] - tm[1][best]) / (r0 - r1);
    const double a = std::max(0.0, tm[0][best] - b * r0);
    const double c = tv[0] / r0;
    if (c > b) min_rows = a / (c - b);
  } else if (tm[0][best] < tv[0]) {
    min_rows = r0;  // one length only: the matrix cores from this segment's rows on
  }
  ch->min_rows = (uint32_t)std::min(4294967295.0, std::max(64.0, min_rows));
  return CLW_OK;
}

extern "C" int clw_batch_autotune(clw_batch* B, uint64_t q, const char* cache_path, clw_band_tune* report,
                                  uint32_t report_cap) {
  if (!B) return clw_fail(CLW_ERR_ARG, "clw_batch_autotune: null batch");
  if (B->iterated) return clw_fail(CLW_ERR_STATE, "clw_batch_autotune after iterate");
  if (B->pending.active) return clw_fail(CLW_ERR_STATE, "clw_batch_autotune with a window in flight");
  if (!B->rec) return clw_fail(CLW_ERR_STATE, "batch without window buffers (clw_batch_set_workspace)");
  if (B->ws && B->ws->ctx != B->ctx)
    return clw_fail(CLW_ERR_STATE, "batch bound to another stream than its workspace (clw_batch_set_workspace again)");
  clw_ctx* ctx = B->ctx;
  if (q < 7) q = 7;
  if (q > B->maxT) q = B->maxT;
  uint32_t nrep = 0;
  for (int band = 2; band <= 4; ++band) {
    Band& bd = B->bands[band];
    if (!bd.L) continue;
    uint64_t active = 0;
    for (const GroupHost& g : bd.groups) active += g.T >= q ? g.tp_n : 0;
    clw_band_tune t{};
    t.limbs = (uint32_t)bd.L;
    t.active_test_primes = active;
    TuneKey key{};
    snprintf(key.arch, sizeof(key.arch), "%s", ctx->prop.gcnArchName);
    for (char* c = key.arch; *c; ++c)
      if (*c == ' ') *c = '_';
    key.cus = ctx->prop.multiProcessorCount;
    key.L = bd.L;
    key.size_class = active ? 63 - __builtin_clzll(active) : 0;
    TuneChoice ch;
    if (active == 0) {
      // nothing of this band reaches q: keep its plan
      ch = TuneChoice{bd.per_lane, bd.items_per_block, bd.nslices, bd.mfma_min_rows};
    } else if (cache_lookup(cache_path, key, &ch)) {
      t.from_cache = 1;
      if (bd.mfma) CK(ensure_pm(B, bd, ch.slices));
    } else {
      uint64_t act_all = 0;  // test primes of all bands active at q: they set the window length
      CK(plan_window(B, q, &act_all));  // the groups and tiles active at q
      bd.per_lane = B->opt.per_lane ? B->opt.per_lane : clw_batch_default_per_lane((uint32_t)bd.L);
      bd.items_per_block = kItemsPerBlock;
      CK(clwi_batch_resident(ctx, (uint32_t)bd.L, bd.per_lane, &bd.resident));
      if (bd.mfma) {
        CK(tune_mfma(B, bd, q, act_all, &ch, &t));
      } else {
        ch.slices = bd.nslices;
        ch.min_rows = bd.mfma_min_rows;
      }
      // step 2 on one window-length segment, the run products where step 1 put them
      uint32_t rows = 0;
      SegHost desc{};
      CK(trial_segment(B, q, window_blocks(B, q, act_all), &rows, &desc));
      const uint32_t sl = bd.mfma && rows / ch.slices >= ch.min_rows ? ch.slices : 0u;
      static const uint32_t kps_by_band[5][6] = {{0}, {0}, {1, 2, 3, 4}, {2, 3, 4, 5, 6}, {1, 2, 3, 4, 5}};
      const uint32_t ipbs[3] = {8, 32, 128};
      double best = -1, best_ms = 1e300, ms = 0, apps = 0;  // every candidate on the same segment: by time
      uint64_t act = 0;
      CK(plan_window(B, q, &act));
      CK(trial_products(B, bd, desc, sl, false, &ms, &apps));  // warm-up
      for (uint32_t a : kps_by_band[bd.L]) {
        if (!a || (B->opt.per_lane && a != B->opt.per_lane)) continue;
        for (uint32_t b : ipbs) {
          bd.per_lane = a;
          bd.items_per_block = b;
          CK(clwi_batch_resident(ctx, (uint32_t)bd.L, bd.per_lane, &bd.resident));
          if (B->opt.target_waves) bd.resident = std::max(1u, B->opt.target_waves / 4);
          CK(plan_window(B, q, &act));  // items and K for this tile size
          CK(trial_products(B, bd, desc, sl, false, &ms, &apps));
          const double rate = ms > 0 ? apps / (ms * 1e-3) : 0.0;
          if (t.trials < 16) {
            t.trial_per_lane[t.trials] = a;
            t.trial_items_per_block[t.trials] = b;
            t.trial_apps_per_s[t.trials] = rate;
            t.trial_ms[t.trials] = ms;
          }
          t.trials++;
          if (ms < best_ms) {
            best_ms = ms;
            best = rate;
            ch.kp = a;
            ch.ipb = b;
          }
        }
      }
      t.best_apps_per_s = best;
      cache_store(cache_path, key, ch);
    }
    bd.per_lane = ch.kp;
    bd.items_per_block = ch.ipb;
    if (bd.mfma) {
      bd.nslices = ch.slices;
      bd.mfma_min_rows = ch.min_rows;
    }
    CK(clwi_batch_resident(ctx, (uint32_t)bd.L, bd.per_lane, &bd.resident));
    if (B->opt.target_waves) bd.resident = std::max(1u, B->opt.target_waves / 4);
    t.per_lane = bd.per_lane;
    t.items_per_block = bd.items_per_block;
    t.mfma_slices = bd.mfma ? bd.nslices : 0u;
    t.mfma_min_rows = bd.mfma ? bd.mfma_min_rows : 0u;
    if (report && nrep < report_cap) report[nrep] = t;
    nrep++;
  }
  // the product kernels and plans may have changed: capture the segment step again
  if (B->graph_exec) CK(clw_hip(hipGraphExecDestroy(B->graph_exec), "hipGraphExecDestroy"));
  if (B->graph) CK(clw_hip(hipGraphDestroy(B->graph), "hipGraphDestroy"));
  B->graph_exec = nullptr;
  B->graph = nullptr;
  return CLW_OK;
}

// ---- checkpoint / resume of a batch ----------------------------------------
//
// The reference's A/B checkpoint (cl_wilson.cpp:110-319: a header with the
// range, progress and a checksum, then the residues; two files written in
// turn, the valid one with more progress read back) extended to a batch:
//   header  "CLWBATCH", version, ranges, covered (sieving primes [2, covered)
//           applied to every range), applied count, state_sum
//   ranges  {pmin, pmax, tpcount, finished}
//   data    per range in increasing p: canonical residues (lo, hi), or for a
//           range already finished on its own, its (quot, rem)
// state_sum = the additive sum (mod 2^64) of every other 64-bit word.

namespace {

const char kBatchMagic[8] = {'C', 'L', 'W', 'B', 'A', 'T', 'C', 'H'};
const uint64_t kBatchCkptVersion = 1;

struct CkptHead {
  char magic[8];
  uint64_t version, nranges, covered, applied, state_sum;
};
struct CkptRange {
  uint64_t pmin, pmax, tpcount, finished;
};

uint64_t sum_words(const void* p, size_t bytes) {
  const uint64_t* w = static_cast<const uint64_t*>(p);
  uint64_t s = 0;
  for (size_t i = 0; i < bytes / 8; ++i) s += w[i];
  return s;
}

}  // namespace

extern "C" int clw_batch_ckpt_write(clw_batch* B, const char* path) {
  if (!B || !path) return clw_fail(CLW_ERR_ARG, "clw_batch_ckpt_write: null argument");
  if (B->pending.active) return clw_fail(CLW_ERR_STATE, "clw_batch_ckpt_write with a window in flight");
  if (B->iterated) return clw_fail(CLW_ERR_STATE, "clw_batch_ckpt_write after clw_batch_iterate");
  // a launched finish has iterated its range's accumulators in place: the
  // file takes that range's results, never those values as residues
  for (RangeHost& R : B->ranges)
    if (R.finishing) CK(collect_finish(B, R));
  std::vector<CkptRange> rr(B->ranges.size());
  std::vector<uint64_t> data;
  std::vector<uint64_t> acc_lo[6], acc_hi[6];
  for (int band = 2; band <= 4; ++band) {
    Band& bd = B->bands[band];
    if (!bd.L) continue;
    acc_lo[band].resize(bd.tab.count);
    acc_hi[band].resize(bd.tab.count);
    CK(clw_d2h(B->ctx, acc_lo[band].data(), bd.tab.acc_lo, (size_t)bd.tab.count * 8, 0));
    CK(clw_d2h(B->ctx, acc_hi[band].data(), bd.tab.acc_hi, (size_t)bd.tab.count * 8, 1));
  }
  for (size_t r = 0; r < B->ranges.size(); ++r) {
    const RangeHost& R = B->ranges[r];
    rr[r] = {R.pmin, R.pmax, (uint64_t)R.tp.size(), R.finished ? 1ull : 0ull};
    for (size_t k = 0; k < R.tp.size(); ++k) {
      const uint32_t i = R.pos[k];
      if (R.finished) {
        data.push_back(R.quot[i - R.band_off]);
        data.push_back(R.rem[i - R.band_off]);
      } else {
        data.push_back(acc_lo[R.band][i]);
        data.push_back(acc_hi[R.band][i]);
      }
    }
  }
  CkptHead h{};
  memcpy(h.magic, kBatchMagic, 8);
  h.version = kBatchCkptVersion;
  h.nranges = B->ranges.size();
  h.covered = B->next_q;
  h.applied = B->applied;
  h.state_sum = h.version + h.nranges + h.covered + h.applied + sum_words(rr.data(), rr.size() * sizeof(CkptRange)) +
                sum_words(data.data(), data.size() * 8);
  FILE* f = fopen(path, "wb");
  if (!f) return clw_fail(CLW_ERR_IO, "cannot open %s", path);
  bool ok = fwrite(&h, sizeof(h), 1, f) == 1;
  ok = ok && fwrite(rr.data(), sizeof(CkptRange), rr.size(), f) == rr.size();
  ok = ok && (data.empty() || fwrite(data.data(), 8, data.size(), f) == data.size());
  if (fclose(f) != 0) ok = false;
  if (!ok) return clw_fail(CLW_ERR_IO, "cannot write batch checkpoint %s", path);
  return CLW_OK;
}

// Reads one checkpoint file; 1 if it is valid for this batch (head and data
// filled), 0 if missing, unreadable, for other ranges or failing its sum.
static int read_batch_file(const clw_batch* B, const char* path, CkptHead* h, std::vector<CkptRange>* rr,
                           std::vector<uint64_t>* data) {
  if (!path) return 0;
  FILE* f = fopen(path, "rb");
  if (!f) return 0;
  bool ok = fread(h, sizeof(*h), 1, f) == 1 && !memcmp(h->magic, kBatchMagic, 8) &&
            h->version == kBatchCkptVersion && h->nranges == B->ranges.size();
  uint64_t words = 0;
  if (ok) {
    rr->resize(h->nranges);
    ok = fread(rr->data(), sizeof(CkptRange), rr->size(), f) == rr->size();
    for (size_t r = 0; ok && r < rr->size(); ++r) {
      const RangeHost& R = B->ranges[r];
      const CkptRange& c = (*rr)[r];
      ok = c.pmin == R.pmin && c.pmax == R.pmax && c.tpcount == R.tp.size() && c.finished <= 1;
      words += 2 * c.tpcount;
    }
  }
  if (ok) {
    data->resize(words);
    ok = words == 0 || fread(data->data(), 8, words, f) == words;
  }
  fclose(f);
  if (!ok) return 0;
  const uint64_t s = h->version + h->nranges + h->covered + h->applied +
                     sum_words(rr->data(), rr->size() * sizeof(CkptRange)) + sum_words(data->data(), words * 8);
  return s == h->state_sum ? 1 : 0;
}

extern "C" int clw_batch_ckpt_read(clw_batch* B, const char* path_a, const char* path_b, int* which,
                                   int* result) {
  if (!B || !which || !result) return clw_fail(CLW_ERR_ARG, "clw_batch_ckpt_read: null argument");
  if (B->pending.active || B->iterated) return clw_fail(CLW_ERR_STATE, "clw_batch_ckpt_read: batch in use");
  // a launched finish would later hand back results of the state replaced here
  for (const RangeHost& R : B->ranges)
    if (R.finishing) return clw_fail(CLW_ERR_STATE, "clw_batch_ckpt_read: a range's finish is in flight");
  *which = -1;
  *result = 0;
  CkptHead ha{}, hb{};
  std::vector<CkptRange> ra, rb;
  std::vector<uint64_t> da, db;
  const int va = read_batch_file(B, path_a, &ha, &ra, &da);
  const int vb = read_batch_file(B, path_b, &hb, &rb, &db);
  if (!va && !vb) return CLW_OK;
  const bool use_b = vb && (!va || hb.covered > ha.covered);
  const CkptHead& h = use_b ? hb : ha;
  const std::vector<CkptRange>& rr = use_b ? rb : ra;
  const std::vector<uint64_t>& d = use_b ? db : da;
  std::vector<uint64_t> acc_lo[6], acc_hi[6];
  for (int band = 2; band <= 4; ++band) {
    Band& bd = B->bands[band];
    if (!bd.L) continue;
    acc_lo[band].assign(bd.tab.count, 0);
    acc_hi[band].assign(bd.tab.count, 0);
  }
  size_t w = 0;
  for (size_t r = 0; r < B->ranges.size(); ++r) {
    RangeHost& R = B->ranges[r];
    R.finished = rr[r].finished != 0;
    if (R.finished) {
      R.quot.assign(R.tp.size(), 0);
      R.rem.assign(R.tp.size(), 0);
    }
    for (size_t k = 0; k < R.tp.size(); ++k, w += 2) {
      const uint32_t i = R.pos[k];
      if (R.finished) {
        R.quot[i - R.band_off] = d[w];
        R.rem[i - R.band_off] = d[w + 1];
      } else {
        acc_lo[R.band][i] = d[w];
        acc_hi[R.band][i] = d[w + 1];
      }
    }
  }
  for (int band = 2; band <= 4; ++band) {
    Band& bd = B->bands[band];
    if (!bd.L) continue;
    CK(clw_h2d(B->ctx, bd.tab.acc_lo, acc_lo[band].data(), (size_t)bd.tab.count * 8, 0));
    CK(clw_h2d(B->ctx, bd.tab.acc_hi, acc_hi[band].data(), (size_t)bd.tab.count * 8, 1));
  }
  B->next_q = h.covered;
  B->applied = h.applied;
  *which = use_b ? 1 : 0;
  *result = 1;
  return CLW_OK;
}
