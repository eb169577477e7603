// clw_internal.h -- shared declarations of the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clwilson.h"

#define CLW_SP_COUNT 169u  // odd primes 7..1021 used for trial division

struct clw_ctx {
  int device;
  hipStream_t stream;
  hipDeviceProp_t prop;
};

struct clw_event {
  hipEvent_t ev;
};

// Records an error message for clw_last_error() and returns `code`.
int clw_fail(int code, const char* fmt, ...);
int clw_fail_hip(hipError_t e, const char* what);
// Checks hipGetLastError() after a launch.
int clw_check_launch(const char* kernel);
// CLW_OK, or the HIP error recorded for clw_last_error().
static inline int clw_hip(hipError_t e, const char* what) { return e == hipSuccess ? CLW_OK : clw_fail_hip(e, what); }
static inline hipStream_t clw_stream(clw_ctx* ctx) { return ctx->stream; }
int clw_upload_sieve_tables();
// Looks up the current device's address of the allocation-free entry point's
// scratch array (clw_kernels.hip), once per device; called by clw_init.
int clwi_noalloc_setup(int device);

// Run products of the batched engine: a sub-chunk is kBatchSubRecs pair
// records (128 sieving primes), at most kBatchSubCap per sieve block, each
// multiplied out into at most kBatchSubMaxDigits 30-bit digits (128 primes
// below 2^62: 265 digits, rounded up to a multiple of 12).
constexpr uint32_t kBatchSubRecs = 64;
constexpr uint32_t kBatchSubCap = CLW_REC_CAP / kBatchSubRecs;
constexpr uint32_t kBatchSubMaxDigits = 276;

// Run products on the matrix cores (the 3-limb band, and the 4-limb band while
// its test primes stay below 2^47): each sub-chunk product P
// is also kept as kMfmaKB signed bytes a_i (P = sum a_i 2^(8 i), a_i in
// [-128, 127]); every test prime of the band has a table of the
// kMfmaCols signed bytes of c_i = 2^(8 i + 30) mod p^2 for i < kMfmaKB, so that
// sum_i a_i c_i = P 2^30 (mod p^2) is an int8 GEMM (rows: sub-chunks, columns:
// test primes x byte columns). kMfmaTile test primes share one table tile in
// LDS; the rows of a segment are split into kMfmaSlices slices.
constexpr uint32_t kMfmaTile = 16;
constexpr uint32_t kMfmaCols = 12;   // bytes of c < p^2 < 2^94
constexpr uint32_t kMfmaKB = 768;    // bytes per product: 128 primes below 2^46 take <= 737
constexpr uint32_t kMfmaSlices = 8;      // fewest slices (one per XCD)
constexpr uint32_t kMfmaMaxSlices = 32;  // most: a slice's rows then stay in one XCD's L2
constexpr size_t kMfmaTileBytes = (size_t)kMfmaCols * kMfmaKB * kMfmaTile;  // 147,456
// The bytes of the segment's sub-chunk products sit in the order the matrix
// cores read them: row r (its place in the segment's row list) belongs to
// row block r / 16; a row block holds, for each 64-byte step kb, the 1 KB
// operand of v_mfma_i32_16x16x64_i8 in lane order (lane 16 g + t: bytes
// [64 kb + 16 g, + 16) of row 16 (r / 16) + t). One wave's load of a step is
// then one contiguous KB.
constexpr uint32_t kMfmaRowBlockBytes = 16 * kMfmaKB;  // 12,288
__host__ __device__ constexpr size_t mfma_row_byte(uint32_t r, uint32_t i) {
  return (size_t)(r / 16) * kMfmaRowBlockBytes + (size_t)(i / 64) * 1024 + ((i / 16) % 4 * 16 + r % 16) * 16 + i % 16;
}
// bytes of the row buffer for `rows` rows: whole row blocks and the 1 KB step
// that k_mfma_runs loads past the last one (unused)
constexpr size_t mfma_rows_bytes(size_t rows) { return (rows + 15) / 16 * kMfmaRowBlockBytes + 1024; }

// Batched-engine launchers (clw_kernels.hip) used by clw_batch.hip.
extern "C" {
int clwi_batch_sieve(clw_ctx* ctx, void* seg, uint32_t* rec, uint32_t* nrec, uint32_t max_blocks, uint32_t* flags,
                     uint64_t* total, uint32_t* era, uint64_t era_hi_max);
int clwi_era_prepare(clw_ctx* ctx, uint64_t hi_max);
uint64_t clwi_era_hi_max();
int clwi_batch_resident(clw_ctx* ctx, uint32_t limbs, uint32_t per_lane, uint32_t* out);
int clwi_batch_mul(clw_ctx* ctx, const clw_tp_table* t, uint32_t per_lane, const void* groups,
                   const uint32_t* item_gt, const void* plan, const uint32_t* rec, const uint32_t* nrec, void* seg,
                   uint32_t qslot, uint32_t extra_block, const uint32_t* subp, const uint32_t* subn, uint32_t ds,
                   uint32_t grid, uint64_t* part_lo, uint64_t* part_hi, uint64_t* apps);
int clwi_batch_subprod(clw_ctx* ctx, void* seg, uint32_t extra_block, const uint32_t* rec, const uint32_t* nrec,
                       uint32_t max_blocks, uint32_t* subp, uint32_t* subn, uint32_t ds, int8_t* subb,
                       const uint32_t* rowpos, void* subq, uint32_t* subm, uint32_t* flags);
int clwi_mfma_table(clw_ctx* ctx, const clw_tp_table* t, const uint32_t* mt_first, const uint32_t* mt_n,
                    uint32_t n_mt, void* ctab);
int clwi_mfma_rows(clw_ctx* ctx, void* seg, uint32_t extra_block, const uint32_t* nrec, uint32_t* rowlist,
                   uint32_t* rowpos, uint32_t* mrows);
// Returns a block from clw_alloc that no kernel uses any more (every stream
// that used it has drained) to the device's cache of idle blocks.
int clwi_release(clw_ctx* ctx, void* dptr);
// Pinned host memory from a cache of idle blocks (clw_runtime.hip); a block
// goes back with clwi_host_release once no copy uses it.
int clwi_host_alloc(size_t bytes, void** hptr);
void clwi_host_release(void* hptr);
int clwi_mfma_runs(clw_ctx* ctx, const clw_tp_table* t, const void* groups, const uint32_t* mt_group,
                   const uint32_t* mt_first, const uint32_t* mt_n, const uint32_t* mt_active, const uint32_t* n_active,
                   uint32_t n_mt, const uint32_t* rowlist, const uint32_t* mrows, const void* subq,
                   const uint32_t* subm, const int8_t* subb, uint64_t subb_bytes, const void* ctab, uint64_t* pm_lo,
                   uint64_t* pm_hi, uint64_t* apps, uint32_t lazy, uint32_t nslices, const void* seg);
int clwi_session_subprod(clw_ctx* ctx, const uint32_t* rec, const uint32_t* nrec, uint32_t nblk, uint32_t* subp,
                         uint32_t* subn, uint32_t ds, uint32_t* flags);
int clwi_mul_records_runs(clw_ctx* ctx, const clw_tp_table* t, const clw_seg_buffers* s, const uint32_t* subp,
                          const uint32_t* subn, uint32_t ds, uint32_t type, uint32_t tp_off, uint32_t tp_n,
                          uint32_t nblk, uint32_t K, uint64_t* part_lo, uint64_t* part_hi, uint32_t part_stride,
                          uint32_t per_lane, uint64_t* pprod);
int clwi_batch_reduce(clw_ctx* ctx, const clw_tp_table* t, const uint32_t* tp_group, const void* groups,
                      const void* plan, const uint64_t* part_lo, const uint64_t* part_hi, const uint64_t* pm_lo,
                      const uint64_t* pm_hi);
int clwi_seg_advance(clw_ctx* ctx, void* seg, uint32_t* flags, uint64_t* total_prev, uint64_t* total);
int clwi_small_primes_block(clw_ctx* ctx, uint64_t qlo, uint64_t qhi, uint32_t* rec, uint32_t* nrec, uint32_t blk);
size_t clwi_batch_seg_size();
size_t clwi_batch_group_size();
size_t clwi_batch_band_size();
}
