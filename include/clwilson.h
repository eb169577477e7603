/*
 * clwilson.h -- C ABI of the MI355X Wilson-prime engine (libclwilson.so).
 *
 * This replaces the reference's OpenCL layer: simpleCL.h/simpleCL.c (device
 * handles, runtime kernel builds, positional kernel arguments, NDRange
 * launches, buffer I/O, profiling) and the raw cl* calls of main.cpp and
 * cl_wilson.cpp. Kernels are compiled offline for gfx950; every launch is a
 * typed entry point with explicit sizes. Functions return an int status
 * (CLW_OK = 0) and never exit the process; clw_last_error() describes the
 * last failure on the calling thread. No torch or C++ types cross this ABI.
 *
 * Residues are exchanged in canonical form: the integer value V mod p^2,
 * 0 <= V < p^2, split into two uint64 words (lo, hi) in separate arrays.
 */
#ifndef CLWILSON_H
#define CLWILSON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CLW_OK 0
#define CLW_ERR_HIP (-1)      /* HIP runtime error (message in clw_last_error) */
#define CLW_ERR_ARG (-2)      /* invalid argument / range                       */
#define CLW_ERR_OVERFLOW (-3) /* a device capacity was exceeded (sieve, side)   */
#define CLW_ERR_NOMEM (-4)    /* allocation failed                              */
#define CLW_ERR_CHECK (-5)    /* Wilson check failed: (p-1)!+1 not = 0 mod p    */
#define CLW_ERR_STATE (-6)    /* calls out of order / unusable state            */
#define CLW_ERR_IO (-7)       /* file I/O (results, checkpoints)                */

/* Device-side error flags (clw_seg_buffers.flags[0] bits). */
#define CLW_FLAG_SIEVE_OVERFLOW 1u /* more primes in a sieve block than fit     */
#define CLW_FLAG_REC_OVERFLOW 2u   /* more records in a block than rec_cap      */
#define CLW_FLAG_SIDE_OVERFLOW 4u  /* side list longer than side_cap            */

/* ---- device & runtime (simpleCL.h:46-141; main.cpp:192-283) ------------- */

/* Opaque device context: one HIP device, one in-order stream.
 * Replaces sclHard {platform, context, device, queue} (simpleCL.h:46-51). */
typedef struct clw_ctx clw_ctx;

typedef struct {
  char name[256];          /* CL_DEVICE_NAME (main.cpp:247)                   */
  char arch[64];           /* gcnArchName, e.g. gfx950                        */
  int32_t compute_units;   /* CL_DEVICE_MAX_COMPUTE_UNITS (main.cpp:262)      */
  int32_t clock_khz;
  uint64_t global_mem;     /* CL_DEVICE_GLOBAL_MEM_SIZE (main.cpp:272)        */
  uint64_t max_alloc;      /* CL_DEVICE_MAX_MEM_ALLOC_SIZE (main.cpp:267)     */
} clw_device_info;

/* Device selection, context and queue creation: main.cpp:199-241. */
int clw_init(int device_index, clw_ctx** out);
/* sclReleaseClHard (simpleCL.c:579-583). */
int clw_shutdown(clw_ctx* ctx);
/* clGetDeviceInfo queries of main.cpp:243-275. */
int clw_get_device_info(clw_ctx* ctx, clw_device_info* info);
/* Replaces sclPrintErrorFlags + exit (simpleCL.c:30-350): text of the last
 * failure on this thread ("" if none). */
const char* clw_last_error(void);

/* clCreateBuffer call sites (e.g. cl_wilson.cpp:1459-1470, 1597-1612). */
int clw_alloc(clw_ctx* ctx, size_t bytes, void** dptr);
/* sclReleaseMemObject (simpleCL.c:585-597); NULL-safe. */
int clw_free(clw_ctx* ctx, void* dptr);
/* sclWrite / sclWriteNB (simpleCL.c:718-742); blocking=0 is asynchronous on
 * the context stream (the host buffer must stay valid until clw_finish). */
int clw_h2d(clw_ctx* ctx, void* dst, const void* src, size_t bytes, int blocking);
/* sclRead / sclReadNB (simpleCL.c:744-768). */
int clw_d2h(clw_ctx* ctx, void* dst, const void* src, size_t bytes, int blocking);
/* Fills device memory with a byte value (no counterpart: the reference uses
 * clearresult.cl / clearacu kernels for its counters). */
int clw_memset(clw_ctx* ctx, void* dptr, int value, size_t bytes);
/* sclFinish (simpleCL.c:770-785): wait for all queued work. */
int clw_finish(clw_ctx* ctx);

/* Events replace clEnqueueMarker + clGetEventInfo polling (cl_wilson.cpp:357-405)
 * and ProfilesclEnqueueKernel's START/END timestamps (simpleCL.c:501-524). */
typedef struct clw_event clw_event;
int clw_event_create(clw_ctx* ctx, clw_event** ev);
int clw_event_record(clw_ctx* ctx, clw_event* ev);
/* 1 if complete, 0 if pending, <0 on error. */
int clw_event_query(clw_event* ev);
int clw_event_wait(clw_event* ev);
int clw_event_elapsed_ms(clw_event* start, clw_event* stop, float* ms);
int clw_event_destroy(clw_event* ev);

/* ---- per-test-prime table (setup.cl:11-52; cl_wilson.h:11-15) ----------- */

/* Structure-of-arrays table of test primes in device memory. All arrays have
 * `count` entries. Test primes are grouped by type (0, 1, 2); the caller
 * fills p[] and type[]; clw_setup fills the rest. `limbs` selects the
 * Montgomery limb schedule (2..5 limbs of 30 bits, R = 2^(30*limbs)); use
 * clw_limbs_for(pmax). acc_lo/acc_hi hold the running canonical residues. */
typedef struct {
  uint64_t* p;
  uint32_t* type;
  uint64_t* tp_target;   /* (p-1)/6, /4, /2 by type                         */
  uint64_t* n_lo;        /* N = p^2                                          */
  uint64_t* n_hi;
  uint64_t* one_lo;      /* R mod N                                          */
  uint64_t* one_hi;
  uint64_t* r2_lo;       /* R^2 mod N                                        */
  uint64_t* r2_hi;
  uint32_t* ninv;        /* -N^-1 mod 2^30                                   */
  uint64_t* acc_lo;      /* canonical running residue                        */
  uint64_t* acc_hi;
  uint32_t count;
  uint32_t limbs;
} clw_tp_table;

/* Smallest limb schedule whose Montgomery radix covers p_max^2 (2N < R). */
int clw_limbs_for(uint64_t p_max);
/* Allocates every array of the table for `count` test primes. */
int clw_tp_table_alloc(clw_ctx* ctx, uint32_t count, uint32_t limbs, clw_tp_table* t);
int clw_tp_table_free(clw_ctx* ctx, clw_tp_table* t);

/* Kernel setup (setup.cl:11-52): per test prime N, -N^-1, R mod N, R^2 mod N
 * and its own factorial target; residues set to 1 unless resume != 0. */
int clw_setup(clw_ctx* ctx, const clw_tp_table* t, int resume);

/* ---- sieving-prime segments (getsegprps.cl:211-293, cl_wilson.cpp:1293-1398) */

/* Per-segment device buffers. One sieve block covers CLW_BLOCK_SPAN
 * consecutive integers. For every block the sieve writes its primes, paired
 * in increasing order, as 32-byte records {Q limbs[5] (30-bit), e0, e1, e2}:
 * Q = product of the pair (or one prime), e_t = the Legendre exponent of the
 * pair's primes in T_t! for type t (0 = not applied). Prime powers whose
 * exponent does not fit a record go to the per-type side lists as
 * {q, e} uint64 pairs. */
#define CLW_BLOCK_SPAN 61440u
#define CLW_REC_CAP 4096u     /* records per block (8192 primes)              */
#define CLW_REC_WORDS 8u
typedef struct {
  uint32_t* rec;          /* max_blocks * CLW_REC_CAP * 8 words               */
  uint32_t* nrec;         /* records per block                                */
  uint32_t* nprimes;      /* primes per block                                 */
  uint64_t* side[3];      /* side_cap * 2 words per type                      */
  uint32_t* side_count;   /* [3]                                              */
  uint32_t* flags;        /* [1] CLW_FLAG_* bits                              */
  uint64_t* total;        /* [1] primes emitted since the last clear          */
  uint32_t* era;          /* max_blocks * CLW_BLOCK_SPAN / 30 bytes: the wheel
                             bitmap of the segment's sieve of Eratosthenes
                             (segments ending at most at 2^50); NULL: every
                             segment by trial division and Miller-Rabin      */
  uint32_t max_blocks;
  uint32_t side_cap;
} clw_seg_buffers;

int clw_seg_alloc(clw_ctx* ctx, uint32_t max_blocks, uint32_t side_cap, clw_seg_buffers* s);
int clw_seg_free(clw_ctx* ctx, clw_seg_buffers* s);
/* clearresult.cl: zero side counts, flags and the prime total. */
int clw_seg_clear(clw_ctx* ctx, const clw_seg_buffers* s);

/* Segment sieve, replacing getsegprps.cl (GPU 2-PRPs above 2^32) and the CPU
 * primesieve tables below 2^32 (cl_wilson.cpp:1316-1365): exact primes q in
 * [lo, hi) (lo >= 7), paired into records with their Legendre exponents for
 * the type targets T[3] (T_t = 0: type absent). Blocks start at
 * base = floor(lo/30)*30. *nblocks receives the block count. For hi <= 2^50
 * (and s->era allocated) a sieve of Eratosthenes by every prime up to
 * sqrt(hi) marks the composites; above, each block strikes the primes
 * 7..1021 and tests the survivors by deterministic Miller-Rabin. */
int clw_sieve_segment(clw_ctx* ctx, const clw_seg_buffers* s, uint64_t lo, uint64_t hi,
                      const uint64_t targets[3], uint32_t* nblocks);

/* Number of sieve blocks of a segment starting at lo that hold primes <= T. */
uint32_t clw_blocks_upto(uint64_t lo, uint64_t hi, uint64_t T);

/* mulsmall.cl / mullarge.cl: for the test primes [tp_off, tp_off+tp_n) of one
 * type, the product over the records of blocks [0, nblk) split into K
 * contiguous chunks; chunk c writes its canonical product to
 * part_lo/hi[c * part_stride + i] (i relative to tp_off). per_lane test
 * primes share one lane (grid (ceil(tp_n / (256 per_lane)), K)). */
int clw_mul_records(clw_ctx* ctx, const clw_tp_table* t, const clw_seg_buffers* s, uint32_t type,
                    uint32_t tp_off, uint32_t tp_n, uint32_t nblk, uint32_t K,
                    uint64_t* part_lo, uint64_t* part_hi, uint32_t part_stride, uint32_t per_lane);
/* Test primes per lane of clw_mul_records for a limb schedule (4 up to 3
 * limbs, else 2; the product kernel interleaves their Montgomery products).
 * 0 in clw_mul_records / clw_options means this default for sessions. */
uint32_t clw_default_per_lane(uint32_t limbs);
/* The same product with a wave per test prime whose 64 lanes share the
 * records (for ranges with few test primes); grid (ceil(tp_n/4), K). */
int clw_mul_records_lanes(clw_ctx* ctx, const clw_tp_table* t, const clw_seg_buffers* s, uint32_t type,
                          uint32_t tp_off, uint32_t tp_n, uint32_t nblk, uint32_t K,
                          uint64_t* part_lo, uint64_t* part_hi, uint32_t part_stride);
/* The same for the side list of `type` (count read on the device), K chunks. */
int clw_mul_side(clw_ctx* ctx, const clw_tp_table* t, const clw_seg_buffers* s, uint32_t type,
                 uint32_t tp_off, uint32_t tp_n, uint32_t K,
                 uint64_t* part_lo, uint64_t* part_hi, uint32_t part_stride);
/* The reference's launch structure for one test prime ("plain" path):
 * mullarge.cl (groups work-groups over the segment's records, per-record
 * Montgomery conversion and binary powering, LDS tree, one total per group
 * into grp_lo/hi) and reduce.cl (one work-group folds the totals into the
 * residue of test prime tp). */
int clw_plain_mul(clw_ctx* ctx, const clw_tp_table* t, const clw_seg_buffers* s, uint32_t type, uint32_t tp,
                  uint32_t nblk, uint32_t groups, uint64_t* grp_lo, uint64_t* grp_hi);
int clw_plain_reduce(clw_ctx* ctx, const clw_tp_table* t, uint32_t tp, uint32_t groups, const uint64_t* grp_lo,
                     const uint64_t* grp_hi);
/* reduce.cl: acc[tp_off+i] *= prod_{c<nparts} part[c * part_stride + i]. */
int clw_reduce(clw_ctx* ctx, const clw_tp_table* t, uint32_t tp_off, uint32_t tp_n, uint32_t nparts,
               const uint64_t* part_lo, const uint64_t* part_hi, uint32_t part_stride);

/* iterate.cl: acc[i] *= prod_{n = from[type]+1}^{tp_target[i]} n for every
 * test prime (from = the type targets). */
int clw_iterate(clw_ctx* ctx, const clw_tp_table* t, const uint64_t from[3]);
/* The same with a start per test prime: acc[i] *= prod_{n = from_tp[i]+1}^
 * {tp_target[i]} n (from_tp is a device array of count entries). Used when
 * one table holds test primes of several ranges, each with its own targets. */
int clw_iterate_from(clw_ctx* ctx, const clw_tp_table* t, const uint64_t* from_tp);

/* find.cl + processResult (cl_wilson.cpp:542-870): from acc = T_p! mod p^2
 * compute (p-1)! + 1 = rem + p*quot (mod p^2) with the type identities; the
 * a / c / u representations come from Cornacchia's algorithm. acu may be
 * NULL; else acu[2*i] = u or a, acu[2*i+1] = c. */
int clw_finalize(clw_ctx* ctx, const clw_tp_table* t, uint64_t* quot, uint64_t* rem, int64_t* acu);

/* ---- whole range: cl_wilson() (cl_wilson.cpp:1442-1974) ---------------- */

typedef struct {
  uint64_t segment_span;   /* integers per sieve segment (0 = default)        */
  uint32_t target_waves;   /* waves to aim for per product launch (0 = auto)  */
  uint32_t verbose;        /* progress lines on stderr                        */
  uint32_t mode;           /* product kernel: 0 auto (clw_wilson_range: a
                              one-range batch for large ranges), 1 lane per test prime
                              (clw_mul_records), 2 wave per test prime
                              (clw_mul_records_lanes), 3 the reference's
                              launch structure, one test prime per launch
                              (clw_plain_mul + clw_plain_reduce)              */
  uint32_t per_lane;       /* test primes per lane in clw_mul_records (0 auto) */
  uint32_t graph;          /* batches: replay the segment step as a captured
                              HIP graph (0 = launch it eagerly)               */
  uint64_t segment_apps;   /* batches: factor applications aimed at per
                              segment step (0 = 2e12, about a second); the
                              segment never exceeds segment_span (batches:
                              the record buffer, 0 = 268,492,800 integers)   */
  uint32_t host_primes;    /* batches: test primes from the host sieve on
                              worker threads (no device work while building,
                              e.g. beside busy streams) instead of the device */
  uint32_t max_chunks;     /* batches: chunks a segment may be split into
                              (chunk-product slots per test prime; 0 = 256)  */
  uint32_t external_scratch; /* batches: allocate no chunk-product slots; the
                              caller provides them (clw_batch_set_scratch)   */
  uint32_t run_products;   /* batches, 3-limb band: 0 = reduce the run
                              products as an int8 GEMM on the matrix cores
                              (k_mfma_runs; the byte tables, 9.2 KB per test
                              prime, kept in HBM up to 8 GiB, else built per
                              workgroup), 1 = on the vector units
                              (k_batch_mul), 2 = matrix cores with the tables
                              always built per workgroup (no table memory);
                              the same results                                */
  uint32_t mfma_slices;    /* batches, matrix cores: row slices per segment
                              (8, 16 or 32; 0 = 8); each slice's rows are read
                              by one XCD, each slice adds 16 B of partial per
                              test prime; the same results                    */
  uint32_t external_workspace; /* batches: allocate no window buffers (record
                              buffer, sieve bitmap, run products); the caller
                              provides a workspace (clw_batch_set_workspace) */
} clw_options;

typedef struct {
  uint64_t pmin, pmax;
  uint64_t checksum;       /* sum of p + rem + quot (cl_wilson.cpp:900)       */
  uint64_t totalcount;     /* sieving primes applied = pi(maxtarget)          */
  uint64_t typeTarget[3];
  uint64_t maxtarget;
  uint64_t lastp;          /* last reported (near-)Wilson prime               */
  int64_t lastw;           /* its signed Wilson quotient                      */
  uint32_t tpcount, tpcnt[3], resultcount, limbs;
  double seconds;          /* wall time of the GPU phases                     */
} clw_range_stats;

/* Runs the plain method on [pmin, pmax) (5 <= pmin < pmax <= 2^62-1, width
 * <= 10^7). If out_* are non-NULL they receive, in increasing p, the test
 * primes, their Wilson quotients quot = ((p-1)!+1)/p mod p and rem. In the
 * automatic mode a range with about 65536 test primes or more (width / ln
 * pmin) and p below 8.2e17 runs as a one-range batch (same results, the
 * batched engine's faster product path); otherwise, and for modes 1-3, a
 * session. */
int clw_wilson_range(clw_ctx* ctx, uint64_t pmin, uint64_t pmax, const clw_options* opt,
                     clw_range_stats* stats, uint64_t* out_p, uint64_t* out_quot, uint64_t* out_rem,
                     uint64_t out_cap);

/* The same run in phases (the structure of cl_wilson(): setup, segment loop,
 * iterate, getResults). A session owns the device state of one range. */
typedef struct clw_session clw_session;
/* Test primes, typing, device table and setup (cl_wilson.cpp:1442-1843). */
int clw_session_create(clw_ctx* ctx, uint64_t pmin, uint64_t pmax, const clw_options* opt, clw_session** out);
int clw_session_destroy(clw_session* s);
/* Counts and targets of the session (checksum / results fields are 0). */
int clw_session_info(const clw_session* s, clw_range_stats* stats);
/* Applies every sieving prime q in [qlo, qhi), clipped to [2, maxtarget], to
 * the running residues (the segment loop, cl_wilson.cpp:1856-1929). Calls
 * must cover disjoint windows; a full search covers [2, maxtarget]. */
int clw_session_run(clw_session* s, uint64_t qlo, uint64_t qhi);
/* Canonical running residues and the test primes, in increasing p. */
int clw_session_residues(clw_session* s, uint64_t* lo, uint64_t* hi, uint64_t cap);
int clw_session_primes(const clw_session* s, uint64_t* out, uint64_t cap);
/* iterate.cl for every test prime (cl_wilson.cpp:1938-1946). */
int clw_session_iterate(clw_session* s);
/* getResults/processResult (cl_wilson.cpp:1017-1090, 691-900) on the GPU;
 * fills stats (checksum, results, last result) and the per-prime outputs. */
int clw_session_finalize(clw_session* s, clw_range_stats* stats, uint64_t* out_p, uint64_t* out_quot,
                         uint64_t* out_rem, uint64_t out_cap);

/* Per-kernel timing of a session, from HIP events recorded around every
 * launch of clw_session_run (no counterpart in the reference beyond
 * ProfilesclEnqueueKernel). record_apps = full Montgomery products x test
 * primes; prime_apps = sieving primes x test primes applied;
 * partial_products = the 30x30-bit partial products of the product kernels
 * (2L^2 + L per full product, L + 1 per digit of a run product reduced). */
typedef struct {
  uint64_t segments;
  uint64_t sieve_launches, sieve_integers;
  uint64_t mul_launches, mul_record_apps, mul_prime_apps;
  uint64_t side_launches, reduce_launches;
  double sieve_ms, mul_ms, side_ms, reduce_ms;
  uint64_t mul_partial_products;
  uint64_t mfma_int8_macs;  /* batches: int8 multiply-adds of the matrix-core
                               run products (k_mfma_runs): 12 per byte of a
                               sub-chunk product per test prime it is
                               applied to (the GEMM's algorithmic count)      */
  double mfma_ms;           /* batches: k_mfma_runs' share of mul_ms          */
  uint64_t mfma_int8_issued; /* the multiply-adds its MFMA instructions issue
                               (rows padded to groups of 48, bytes to 64)     */
  uint64_t mfma_partial_products; /* k_mfma_runs' own 30x30-bit partial
                               products (one Montgomery digit and one full
                               product per sub-chunk and test prime; not in
                               mul_partial_products)                          */
} clw_kernel_stats;
int clw_session_kernel_stats(const clw_session* s, clw_kernel_stats* out);
int clw_session_reset_stats(clw_session* s);

/* ---- batches: many ranges over one shared sieve ----------------------- */

/* The reference searches one range per process (main.cpp:313-405 ->
 * cl_wilson()); a batch runs the same plain method for up to thousands of
 * ranges at once. Each segment of sieving primes is sieved once and applied
 * to every range whose targets reach it: a persistent product kernel per limb
 * band takes (chunk, range-type group, tile) items from a device work queue
 * and computes each group's Legendre exponents on the fly. Results are those
 * of clw_wilson_range on each range, bit for bit. Ranges must have p below
 * 8.2e17 (at most 4 limbs); wider p are served by sessions. */
typedef struct clw_batch clw_batch;
int clw_batch_create(clw_ctx* ctx, const uint64_t* pmin, const uint64_t* pmax, uint32_t nranges,
                     const clw_options* opt, clw_batch** out);
int clw_batch_destroy(clw_batch* b);
/* Test primes per lane of the batched product kernel for a limb band when
 * neither clw_options.per_lane nor the autotuner sets it: 4 for 2 limbs, 3
 * for 3 limbs, 2 for 4 limbs (the fastest at each band's production
 * magnitude). */
uint32_t clw_batch_default_per_lane(uint32_t limbs);
/* Moves the batch to another context (stream) of the same device; no window
 * may be in flight. Lets one thread build batches while others run. */
int clw_batch_bind(clw_batch* b, clw_ctx* ctx);
/* Chunk products in caller-owned device buffers of cap 64-bit slots each
 * (at least the batch's test primes): batches that run one after another on
 * one stream can share one scratch instead of holding their own. */
int clw_batch_set_scratch(clw_batch* b, uint64_t* part_lo, uint64_t* part_hi, uint64_t cap);
/* Window buffers shared by batches that run their windows one after another
 * on one stream (no counterpart in the reference, which runs one range per
 * process; for a request loop holding many batches): the record buffer, the
 * sieve bitmap and the run products of a segment, sized for any batch of
 * the given segment span (0 = the default). A batch created with
 * external_workspace has none of its own until clw_batch_set_workspace; a
 * batch with its own gives them back. The batch must be bound to the
 * workspace's context (its stream orders every window in the buffers):
 * set_workspace refuses another context, and a batch rebound elsewhere
 * afterwards refuses windows until it is given a workspace of its new
 * context. The workspace outlives its batches. */
typedef struct clw_workspace clw_workspace;
int clw_workspace_create(clw_ctx* ctx, uint64_t segment_span, clw_workspace** out);
int clw_workspace_destroy(clw_workspace* ws);
uint64_t clw_workspace_bytes(const clw_workspace* ws);
int clw_batch_set_workspace(clw_batch* b, clw_workspace* ws);
uint32_t clw_batch_size(const clw_batch* b);
/* Counts and targets of range r (as clw_session_info). */
int clw_batch_info(const clw_batch* b, uint32_t r, clw_range_stats* stats);
/* Applies the sieving primes q in [qlo, qhi), clipped to [2, the largest
 * target of the batch], to every range (disjoint windows across calls). */
int clw_batch_run(clw_batch* b, uint64_t qlo, uint64_t qhi);
/* clw_batch_run in two halves: launch enqueues the window on the batch's
 * stream and returns; wait blocks until it is done and checks the device
 * flags. Batches on different contexts (streams) overlap on the device. */
int clw_batch_launch(clw_batch* b, uint64_t qlo, uint64_t qhi);
int clw_batch_wait(clw_batch* b);
/* 1 when the launched window has finished on the device (or none is in
 * flight), 0 while it runs; never blocks. */
int clw_batch_query(const clw_batch* b);
/* End of the prefix [2, covered) of sieving primes applied to every range. */
uint64_t clw_batch_covered(const clw_batch* b);
/* Device memory held by the batch (tables, buffers, scratch). */
uint64_t clw_batch_device_bytes(const clw_batch* b);
/* Range r done early: once clw_batch_covered() exceeds its largest target,
 * iterates and checks its test primes alone, leaving the others running, and
 * returns its results as clw_batch_results does. */
int clw_batch_finish_range(clw_batch* b, uint32_t r, clw_range_stats* stats, uint64_t* out_p,
                           uint64_t* out_quot, uint64_t* out_rem, uint64_t out_cap);
/* clw_batch_finish_range in two halves, for a request loop that must not
 * block (no counterpart in the reference, which finishes one range per
 * process): finish_launch enqueues range r's iterate and check on the batch's
 * stream, with a copy of its results into pinned host memory, and returns
 * (no window may be in flight; windows may be launched after it);
 * finish_query returns 1 once the results are in, 0 while they are not
 * (never blocks; CLW_ERR_STATE if range r was not launched);
 * clw_batch_finish_range then collects them. */
int clw_batch_finish_launch(clw_batch* b, uint32_t r);
int clw_batch_finish_query(const clw_batch* b, uint32_t r);
/* Sieving primes applied so far (all windows). */
uint64_t clw_batch_applied(const clw_batch* b);
/* Canonical running residues of range r, in increasing p. */
int clw_batch_residues(clw_batch* b, uint32_t r, uint64_t* lo, uint64_t* hi, uint64_t cap);
/* Resume from a caller's checkpoint (the command line's stateA/B.ckp, read
 * by clw_ckpt_read; cl_wilson.cpp:1736-1777 read_state): range r's residues
 * in increasing p, holding the sieving primes [2, covered), `applied` of
 * them. One-range batches, before their first window. */
int clw_batch_set_residues(clw_batch* b, uint32_t r, const uint64_t* lo, const uint64_t* hi, uint64_t n,
                           uint64_t covered, uint64_t applied);
int clw_batch_iterate(clw_batch* b);
int clw_batch_finalize(clw_batch* b);
/* Range r after clw_batch_finalize or clw_batch_finish_range: stats
 * (checksum, results, last result;
 * totalcount is not tracked per range and reads 0) and per-prime outputs. */
int clw_batch_results(const clw_batch* b, uint32_t r, clw_range_stats* stats, uint64_t* out_p,
                      uint64_t* out_quot, uint64_t* out_rem, uint64_t out_cap);
/* Autotuner per limb band, at magnitude q (sieving primes near q), on trial
 * segments (sieve + product kernels into scratch; no residue changes):
 *  1. a band whose run products can go to the matrix cores (3 limbs; 4 limbs
 *     below p = 2^47): the row slices of k_mfma_runs (8 or 16; or
 *     clw_options.mfma_slices) and the switch between the matrix cores and
 *     the vector units. Both paths are timed on a segment of the length a
 *     window at q runs and on a quarter of it; with the matrix-core time
 *     a + b r and the vector time c r in the sub-chunks r per slice, windows
 *     use the matrix cores from mfma_min_rows = a / (c - b) sub-chunks per
 *     slice on (never if c <= b). mfma_trial_* hold the trials (slices 0:
 *     vector units; rows: the segment's sub-chunks).
 *  2. test primes per lane of the vector product kernel (1-4 for 2 limbs,
 *     2-6 for 3, 1-5 for 4 limbs: the Montgomery products a lane runs in
 *     lockstep, which set the instruction schedule and the work-group tile
 *     of 256 x per_lane test primes) x work items per resident block (8, 32,
 *     128: the chunk count, i.e. the fan-in of the chunk-product reduction),
 *     timed on the work step 1 leaves to that kernel (where the matrix cores
 *     take the run products, the records across run boundaries only:
 *     best_apps_per_s and trial_apps_per_s are then that kernel's own
 *     factor applications per second, 0 for a segment with no record
 *     across a run boundary); the fastest on the trial segment is kept
 *     (trial_ms: each candidate's time).
 * With cache_path, a tuned (device, band, size class) is read from /
 * appended to that text file instead of re-timed. report receives one entry
 * per band present (at most report_cap). The reference sizes its launches
 * from a timed probe instead (cl_wilson.cpp:1184-1290) and has fixed launch
 * shapes otherwise (cl_wilson.cpp:1795-1843). */
typedef struct {
  uint32_t limbs, per_lane, items_per_block, trials, from_cache;
  uint64_t active_test_primes;
  double best_apps_per_s;
  uint32_t trial_per_lane[16], trial_items_per_block[16];
  double trial_apps_per_s[16];
  uint32_t mfma_slices;   /* chosen row slices of k_mfma_runs (0: band without matrix-core run products) */
  uint32_t mfma_min_rows; /* sub-chunks per slice and segment from which windows use the matrix cores */
  uint32_t mfma_trials;
  uint32_t mfma_trial_slices[8];
  uint64_t mfma_trial_rows[8];
  double mfma_trial_ms[8];
  double trial_ms[16];    /* step 2: each per-lane trial's time on the trial segment (the fastest is kept) */
} clw_band_tune;
int clw_batch_autotune(clw_batch* b, uint64_t q, const char* cache_path, clw_band_tune* report,
                       uint32_t report_cap);
/* Checkpoint of a batch (the A/B scheme of cl_wilson.cpp:110-319 extended
 * to many ranges): every range's canonical running residues (or, for a range
 * already finished on its own, its results) and the covered prefix of
 * sieving primes, with an additive checksum. Write alternately to two files;
 * read picks the valid file (right ranges, right sum) that covers more.
 * *result = 0 (no valid file: start from scratch) or 1 (resumed; continue
 * with windows from clw_batch_covered(); *which = 0 for a, 1 for b). Write
 * first collects any finish in flight (clw_batch_finish_launch), so the file
 * holds that range's results; read refuses a batch with a finish in flight. */
int clw_batch_ckpt_write(clw_batch* b, const char* path);
int clw_batch_ckpt_read(clw_batch* b, const char* path_a, const char* path_b, int* which, int* result);
/* segments, sieve_integers, record/prime applications (device counters);
 * sieve_ms, mul_ms, reduce_ms from HIP events when the batch launches its
 * steps eagerly (opt.graph = 0; 0 under graph replay); side_ms = host wall
 * time of all clw_batch_run calls. */
int clw_batch_kernel_stats(const clw_batch* b, clw_kernel_stats* out);

/* ---- checkpoint / resume (cl_wilson.cpp:110-319, 1734-1777) ------------- */

/* The reference's checkpoint header (workStatus, cl_wilson.h:22-25), 56 bytes,
 * followed in the file by tpcount (lo, hi) residue pairs, under the
 * reference's file names. The files are not interchangeable with the
 * reference's: this build stores each residue in canonical form (value mod
 * p^2, increasing p) where the reference stores Montgomery digits, so its
 * sum carries a format constant the reference's lacks (either program
 * rejects the other's files and starts over). state_sum = 0x434C57414D443031
 * ("CLWAMD01") + pmin + pmax + currp + trickle + totalcount + tpcount + done
 * + sum of all residue words (mod 2^64); the reference's sum has no constant
 * (cl_wilson.cpp:116-119). */
typedef struct {
  uint64_t pmin, pmax, currp, trickle, state_sum, totalcount;
  uint32_t tpcount, done;
} clw_work_status;

/* write_state (cl_wilson.cpp:110-156): writes header + residues to path,
 * filling st->state_sum. */
int clw_ckpt_write(const char* path, clw_work_status* st, const uint64_t* res_lo, const uint64_t* res_hi);
/* read_state (cl_wilson.cpp:159-288): reads both files, rejects a file that
 * cannot be parsed, belongs to another range/count, or fails its checksum
 * (finished files included, unlike the reference), and picks the one with
 * the larger currp. *result = 0 (no valid file),
 * 1 (resume: st and residues filled, *which = 0 for A / 1 for B),
 * 2 (a valid file says done). res_* need room for tpcount entries. */
int clw_ckpt_read(const char* path_a, const char* path_b, uint64_t pmin, uint64_t pmax, uint32_t tpcount,
                  clw_work_status* st, uint64_t* res_lo, uint64_t* res_hi, int* which, int* result);
/* Uploads canonical residues (increasing p) and the applied-prime count of a
 * checkpoint into a session (the sclWrite of residues + setup(resume=1)). */
int clw_session_set_residues(clw_session* s, const uint64_t* lo, const uint64_t* hi, uint64_t totalcount);

/* Result line of cl_wilson.cpp:657-688 for one test prime into buf; returns
 * 1 if the prime is reported (quot = 0 or min(quot, p-quot) < 1000), else 0. */
int clw_format_result(uint64_t p, uint64_t quot, char* buf, size_t buflen);

/* Result verification (the -r option; cl_wilson.cpp:902-930, 936-1014,
 * 1073-1087, readGoodResultFile + the resultTest branch of processResult):
 * reads a known-good list of "p w_p" lines, checks that the additive sum of
 * every line's p + w_p (mod 2^64) equals expected_sum (the reference's file:
 * 0x6659912234F44428), then requires every tested prime with |w_p/p| <
 * 1/50000 (the ratio rounded to double, as the reference) to be listed with
 * the same w_p, and every listed p in [pmin, pmax) to have been tested and
 * matched. CLW_OK, CLW_ERR_IO (missing or unreadable file) or CLW_ERR_CHECK
 * (checksum or match failure; the reference's message in clw_last_error).
 * matched / listed (may be NULL): primes matched, lines in [pmin, pmax). */
int clw_verify_results(const char* path, uint64_t expected_sum, uint64_t pmin, uint64_t pmax, const uint64_t* p,
                       const uint64_t* quot, uint64_t n, uint32_t* matched, uint32_t* listed);

/* Counts of test primes by type and the type targets of [pmin, pmax)
 * (cl_wilson.cpp:1552-1586), computed on the host. */
int clw_range_types(uint64_t pmin, uint64_t pmax, uint32_t tpcnt[3], uint64_t targets[3]);

/* Exact primes in [lo, hi) on the host (the primesieve_generate_primes call of
 * cl_wilson.cpp:1529). Returns the count; writes at most cap primes. */
uint64_t clw_host_primes(uint64_t lo, uint64_t hi, uint64_t* out, uint64_t cap);
/* The same list generated on the device by the segment sieve (wheel-30 trial
 * division in LDS, then deterministic Miller-Rabin): the test primes of
 * sessions and batches. *count = number of primes (those beyond cap are
 * counted, not written). Replaces primesieve_generate_primes for the test
 * primes (cl_wilson.cpp:1529). */
int clw_device_primes(clw_ctx* ctx, uint64_t lo, uint64_t hi, uint64_t* out, uint64_t cap, uint64_t* count);
/* No counterpart in the reference: clw_device_primes for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc). The same list, with no hipMalloc/hipFree and no host heap use:
 * the count pass, offsets, flags and prime list live in one static device
 * array per GPU (8.4 MB, held until the process ends, looked up by clw_init).
 * Ranges are limited to hi - lo <= 10^7 (the reference's range limit,
 * cl_wilson.cpp:1162-1166); wider ones fail with CLW_ERR_ARG. Calls that
 * overlap on one GPU share that array and race: serialize them. */
int clw_device_primes_noalloc(clw_ctx* ctx, uint64_t lo, uint64_t hi, uint64_t* out, uint64_t cap, uint64_t* count);

#ifdef __cplusplus
}
#endif
#endif /* CLWILSON_H */
