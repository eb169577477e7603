// clw_runtime.hip -- device context, memory, copies, events and error state
// of the C ABI (include/clwilson.h). Replaces simpleCL.c and the OpenCL
// set-up of main.cpp:192-283. Nothing here exits the process: every failure
// is returned as a status code with a message for clw_last_error().
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <unordered_map>

#include "clw_internal.h"

#define CK_HIP_OR_RETURN(expr, what)                              \
  do {                                                            \
    const hipError_t e_ = (expr);                                 \
    if (e_ != hipSuccess) return clw_fail_hip(e_, what);          \
  } while (0)

static thread_local char g_err[512] = "";

int clw_fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int clw_fail_hip(hipError_t e, const char* what) {
  return clw_fail(CLW_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

int clw_check_launch(const char* kernel) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return clw_fail_hip(e, kernel);
  return CLW_OK;
}

extern "C" const char* clw_last_error(void) { return g_err; }

static std::mutex g_tables_mu;
static bool g_tables_done[64] = {};

extern "C" int clw_init(int device_index, clw_ctx** out) {
  if (!out) return clw_fail(CLW_ERR_ARG, "clw_init: null out");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return clw_fail_hip(e, "hipGetDeviceCount");
  if (n <= 0) return clw_fail(CLW_ERR_HIP, "no HIP device available");
  if (device_index < 0 || device_index >= n || device_index >= 64)
    return clw_fail(CLW_ERR_ARG, "device %d out of range (%d devices)", device_index, n);
  e = hipSetDevice(device_index);
  if (e != hipSuccess) return clw_fail_hip(e, "hipSetDevice");
  clw_ctx* c = new clw_ctx();
  c->device = device_index;
  e = hipGetDeviceProperties(&c->prop, device_index);
  if (e != hipSuccess) {
    delete c;
    return clw_fail_hip(e, "hipGetDeviceProperties");
  }
  // One in-order stream at the default priority (the reference's single
  // in-order profiling queue, main.cpp:235).
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return clw_fail_hip(e, "hipStreamCreate");
  }
  {
    std::lock_guard<std::mutex> lk(g_tables_mu);
    if (!g_tables_done[device_index]) {
      const int rc = clw_upload_sieve_tables();
      if (rc != CLW_OK) {
        (void)hipStreamDestroy(c->stream);  // already failing: report the first error
        delete c;
        return rc;
      }
      g_tables_done[device_index] = true;
    }
  }
  {
    const int rc = clwi_noalloc_setup(device_index);  // with the device current
    if (rc != CLW_OK) {
      (void)hipStreamDestroy(c->stream);
      delete c;
      return rc;
    }
  }
  *out = c;
  return CLW_OK;
}

extern "C" int clw_shutdown(clw_ctx* ctx) {
  if (!ctx) return CLW_OK;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  const hipError_t d = hipStreamDestroy(ctx->stream);
  delete ctx;
  if (e != hipSuccess) return clw_fail_hip(e, "shutdown sync");
  if (d != hipSuccess) return clw_fail_hip(d, "hipStreamDestroy");
  return CLW_OK;
}

extern "C" int clw_get_device_info(clw_ctx* ctx, clw_device_info* info) {
  if (!ctx || !info) return clw_fail(CLW_ERR_ARG, "clw_get_device_info: null argument");
  memset(info, 0, sizeof(*info));
  snprintf(info->name, sizeof(info->name), "%s", ctx->prop.name);
  snprintf(info->arch, sizeof(info->arch), "%s", ctx->prop.gcnArchName);
  info->compute_units = ctx->prop.multiProcessorCount;
  info->clock_khz = ctx->prop.clockRate;
  info->global_mem = ctx->prop.totalGlobalMem;
  info->max_alloc = ctx->prop.totalGlobalMem;
  return CLW_OK;
}

// Idle device blocks kept for reuse, per device. hipFree waits for the whole
// device, so a request loop that retires a batch while other streams run
// would stall them all, and hipMalloc of large blocks is slow: a batch that
// is destroyed hands its blocks here instead (clwi_release: its stream has
// drained, no kernel uses them), and clw_alloc takes a cached block of the
// same size class before it asks HIP. Sizes go up in steps of 1/8 of a power
// of two (at most 12.5 % more than asked); at most kCacheCap bytes are kept,
// and the cache is emptied when HIP runs out of memory.
namespace {
constexpr int kMaxDevices = 64;
constexpr size_t kCacheCap = 16ull << 30;
std::mutex g_cache_mu;
std::unordered_multimap<size_t, void*> g_cache[kMaxDevices];  // size class -> idle block
std::unordered_map<void*, size_t> g_block_size[kMaxDevices];  // every block clw_alloc handed out
size_t g_cached_bytes[kMaxDevices] = {};

size_t size_class(size_t bytes) {
  if (bytes <= 256) return 256;
  const int top = 63 - __builtin_clzll(bytes - 1);  // 2^top < bytes <= 2^(top+1)
  const size_t step = (size_t)1 << (top - 3);
  return (bytes + step - 1) / step * step;
}

void cache_trim(int dev) {  // caller holds g_cache_mu
  for (auto& kv : g_cache[dev]) {
    (void)hipFree(kv.second);
    g_block_size[dev].erase(kv.second);
  }
  g_cache[dev].clear();
  g_cached_bytes[dev] = 0;
}
}  // namespace

extern "C" int clw_alloc(clw_ctx* ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr) return clw_fail(CLW_ERR_ARG, "clw_alloc: null argument");
  *dptr = nullptr;
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return clw_fail(CLW_ERR_ARG, "clw_alloc: device index");
  const size_t cls = size_class(bytes == 0 ? 16 : bytes);
  CK_HIP_OR_RETURN(hipSetDevice(ctx->device), "hipSetDevice");
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = g_cache[ctx->device].find(cls);
  if (it != g_cache[ctx->device].end()) {
    *dptr = it->second;
    g_cache[ctx->device].erase(it);
    g_cached_bytes[ctx->device] -= cls;
    return CLW_OK;
  }
  hipError_t e = hipMalloc(dptr, cls);
  if (e != hipSuccess && g_cached_bytes[ctx->device]) {  // the idle blocks may be in the way
    (void)hipGetLastError();
    cache_trim(ctx->device);
    e = hipMalloc(dptr, cls);
  }
  if (e != hipSuccess) {
    *dptr = nullptr;
    clw_fail_hip(e, "hipMalloc");
    return CLW_ERR_NOMEM;
  }
  g_block_size[ctx->device][*dptr] = cls;
  return CLW_OK;
}

// A block from clw_alloc that no kernel uses any more (the caller has drained
// every stream that used it) goes to the cache; others are freed.
extern "C" int clwi_release(clw_ctx* ctx, void* dptr) {
  if (!dptr) return CLW_OK;
  if (!ctx || ctx->device < 0 || ctx->device >= kMaxDevices) return clw_free(ctx, dptr);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = g_block_size[ctx->device].find(dptr);
    if (it != g_block_size[ctx->device].end() && g_cached_bytes[ctx->device] + it->second <= kCacheCap) {
      g_cache[ctx->device].emplace(it->second, dptr);
      g_cached_bytes[ctx->device] += it->second;
      return CLW_OK;
    }
  }
  return clw_free(ctx, dptr);
}

// Pinned host blocks for copies that must not block the host (a copy from
// or to pageable memory does), kept for reuse by size class: hipHostFree,
// like hipFree, may wait for the device. Blocks come back through
// clwi_host_release once no copy uses them; at most kHostCacheCap bytes are
// kept idle.
namespace {
constexpr size_t kHostCacheCap = 1ull << 30;
std::unordered_multimap<size_t, void*> g_host_cache;  // size class -> idle pinned block
std::unordered_map<void*, size_t> g_host_size;        // every pinned block handed out
size_t g_host_cached = 0;
}  // namespace

extern "C" int clwi_host_alloc(size_t bytes, void** hptr) {
  if (!hptr) return clw_fail(CLW_ERR_ARG, "clwi_host_alloc: null argument");
  const size_t cls = size_class(bytes == 0 ? 16 : bytes);
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = g_host_cache.find(cls);
  if (it != g_host_cache.end()) {
    *hptr = it->second;
    g_host_cache.erase(it);
    g_host_cached -= cls;
    return CLW_OK;
  }
  const hipError_t e = hipHostMalloc(hptr, cls, hipHostMallocDefault);
  if (e != hipSuccess) {
    *hptr = nullptr;
    clw_fail_hip(e, "hipHostMalloc");
    return CLW_ERR_NOMEM;
  }
  g_host_size[*hptr] = cls;
  return CLW_OK;
}

extern "C" void clwi_host_release(void* hptr) {
  if (!hptr) return;
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = g_host_size.find(hptr);
  if (it == g_host_size.end()) return;
  if (g_host_cached + it->second <= kHostCacheCap) {
    g_host_cache.emplace(it->second, hptr);
    g_host_cached += it->second;
    return;
  }
  g_host_size.erase(it);
  (void)hipHostFree(hptr);
}

extern "C" int clw_free(clw_ctx* ctx, void* dptr) {
  if (!dptr) return CLW_OK;
  if (ctx) CK_HIP_OR_RETURN(hipSetDevice(ctx->device), "hipSetDevice");
  if (ctx && ctx->device >= 0 && ctx->device < kMaxDevices) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_block_size[ctx->device].erase(dptr);
  }
  const hipError_t e = hipFree(dptr);
  if (e != hipSuccess) return clw_fail_hip(e, "hipFree");
  return CLW_OK;
}

extern "C" int clw_h2d(clw_ctx* ctx, void* dst, const void* src, size_t bytes, int blocking) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return clw_fail(CLW_ERR_ARG, "clw_h2d: null argument");
  if (bytes == 0) return CLW_OK;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && blocking) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return clw_fail_hip(e, "h2d copy");
  return CLW_OK;
}

extern "C" int clw_d2h(clw_ctx* ctx, void* dst, const void* src, size_t bytes, int blocking) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return clw_fail(CLW_ERR_ARG, "clw_d2h: null argument");
  if (bytes == 0) return CLW_OK;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && blocking) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return clw_fail_hip(e, "d2h copy");
  return CLW_OK;
}

extern "C" int clw_memset(clw_ctx* ctx, void* dptr, int value, size_t bytes) {
  if (!ctx || (!dptr && bytes)) return clw_fail(CLW_ERR_ARG, "clw_memset: null argument");
  if (bytes == 0) return CLW_OK;
  const hipError_t e = hipMemsetAsync(dptr, value, bytes, ctx->stream);
  if (e != hipSuccess) return clw_fail_hip(e, "memset");
  return CLW_OK;
}

extern "C" int clw_finish(clw_ctx* ctx) {
  if (!ctx) return clw_fail(CLW_ERR_ARG, "clw_finish: null ctx");
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return clw_fail_hip(e, "stream synchronize");
  return CLW_OK;
}

extern "C" int clw_event_create(clw_ctx* ctx, clw_event** ev) {
  if (!ctx || !ev) return clw_fail(CLW_ERR_ARG, "clw_event_create: null argument");
  clw_event* x = new clw_event();
  const hipError_t e = hipEventCreate(&x->ev);
  if (e != hipSuccess) {
    delete x;
    return clw_fail_hip(e, "hipEventCreate");
  }
  *ev = x;
  return CLW_OK;
}

extern "C" int clw_event_record(clw_ctx* ctx, clw_event* ev) {
  if (!ctx || !ev) return clw_fail(CLW_ERR_ARG, "clw_event_record: null argument");
  const hipError_t e = hipEventRecord(ev->ev, ctx->stream);
  if (e != hipSuccess) return clw_fail_hip(e, "hipEventRecord");
  return CLW_OK;
}

extern "C" int clw_event_query(clw_event* ev) {
  if (!ev) return clw_fail(CLW_ERR_ARG, "clw_event_query: null event");
  const hipError_t e = hipEventQuery(ev->ev);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return clw_fail_hip(e, "hipEventQuery");
}

extern "C" int clw_event_wait(clw_event* ev) {
  if (!ev) return clw_fail(CLW_ERR_ARG, "clw_event_wait: null event");
  const hipError_t e = hipEventSynchronize(ev->ev);
  if (e != hipSuccess) return clw_fail_hip(e, "hipEventSynchronize");
  return CLW_OK;
}

extern "C" int clw_event_elapsed_ms(clw_event* a, clw_event* b, float* ms) {
  if (!a || !b || !ms) return clw_fail(CLW_ERR_ARG, "clw_event_elapsed_ms: null argument");
  const hipError_t e = hipEventElapsedTime(ms, a->ev, b->ev);
  if (e != hipSuccess) return clw_fail_hip(e, "hipEventElapsedTime");
  return CLW_OK;
}

extern "C" int clw_event_destroy(clw_event* ev) {
  if (!ev) return CLW_OK;
  const hipError_t e = hipEventDestroy(ev->ev);
  delete ev;
  return clw_hip(e, "hipEventDestroy");
}

extern "C" int clw_limbs_for(uint64_t p_max) {
  // 2N < R = 2^(30 L) with N = p^2  <=>  p < 2^(15 L - 1/2).
  const unsigned __int128 n2 = 2 * (unsigned __int128)p_max * p_max;
  for (int L = 2; L <= 5; ++L) {
    const int bits = 30 * L;
    if (bits >= 128 || (n2 >> bits) == 0) return L;
  }
  return 5;
}

extern "C" int clw_tp_table_alloc(clw_ctx* ctx, uint32_t count, uint32_t limbs, clw_tp_table* t) {
  if (!ctx || !t) return clw_fail(CLW_ERR_ARG, "clw_tp_table_alloc: null argument");
  if (limbs < 2 || limbs > 5) return clw_fail(CLW_ERR_ARG, "limbs %u out of range", limbs);
  memset(t, 0, sizeof(*t));
  t->count = count;
  t->limbs = limbs;
  const size_t n = count ? count : 1;
  uint64_t** u64s[] = {&t->p, &t->tp_target, &t->n_lo, &t->n_hi, &t->one_lo, &t->one_hi,
                       &t->r2_lo, &t->r2_hi, &t->acc_lo, &t->acc_hi};
  for (uint64_t** a : u64s) {
    const int rc = clw_alloc(ctx, n * 8, (void**)a);
    if (rc != CLW_OK) {
      clw_tp_table_free(ctx, t);
      return rc;
    }
  }
  uint32_t** u32s[] = {&t->type, &t->ninv};
  for (uint32_t** a : u32s) {
    const int rc = clw_alloc(ctx, n * 4, (void**)a);
    if (rc != CLW_OK) {
      clw_tp_table_free(ctx, t);
      return rc;
    }
  }
  return CLW_OK;
}

extern "C" int clw_tp_table_free(clw_ctx* ctx, clw_tp_table* t) {
  if (!t) return CLW_OK;
  void* ptrs[] = {t->p, t->tp_target, t->n_lo, t->n_hi, t->one_lo, t->one_hi, t->r2_lo,
                  t->r2_hi, t->acc_lo, t->acc_hi, t->type, t->ninv};
  for (void* p : ptrs) clw_free(ctx, p);
  memset(t, 0, sizeof(*t));
  return CLW_OK;
}

extern "C" int clw_seg_alloc(clw_ctx* ctx, uint32_t max_blocks, uint32_t side_cap, clw_seg_buffers* s) {
  if (!ctx || !s || max_blocks == 0 || side_cap == 0) return clw_fail(CLW_ERR_ARG, "clw_seg_alloc: bad argument");
  memset(s, 0, sizeof(*s));
  s->max_blocks = max_blocks;
  s->side_cap = side_cap;
  int rc = clw_alloc(ctx, (size_t)max_blocks * CLW_REC_CAP * CLW_REC_WORDS * 4, (void**)&s->rec);
  if (rc == CLW_OK) rc = clw_alloc(ctx, (size_t)max_blocks * 4, (void**)&s->nrec);
  if (rc == CLW_OK) rc = clw_alloc(ctx, (size_t)max_blocks * 4, (void**)&s->nprimes);
  for (int t = 0; t < 3 && rc == CLW_OK; ++t) rc = clw_alloc(ctx, (size_t)side_cap * 16, (void**)&s->side[t]);
  if (rc == CLW_OK) rc = clw_alloc(ctx, 16, (void**)&s->side_count);
  if (rc == CLW_OK) rc = clw_alloc(ctx, 16, (void**)&s->flags);
  if (rc == CLW_OK) rc = clw_alloc(ctx, 16, (void**)&s->total);
  if (rc == CLW_OK) rc = clw_alloc(ctx, (size_t)max_blocks * (CLW_BLOCK_SPAN / 30), (void**)&s->era);
  if (rc != CLW_OK) {
    clw_seg_free(ctx, s);
    return rc;
  }
  return clw_seg_clear(ctx, s);
}

extern "C" int clw_seg_free(clw_ctx* ctx, clw_seg_buffers* s) {
  if (!s) return CLW_OK;
  void* ptrs[] = {s->rec,  s->nrec,       s->nprimes, s->side[0], s->side[1],
                  s->side[2], s->side_count, s->flags,   s->total,   s->era};
  for (void* p : ptrs) clw_free(ctx, p);
  memset(s, 0, sizeof(*s));
  return CLW_OK;
}

extern "C" int clw_seg_clear(clw_ctx* ctx, const clw_seg_buffers* s) {
  if (!ctx || !s) return clw_fail(CLW_ERR_ARG, "clw_seg_clear: null argument");
  int rc = clw_memset(ctx, s->side_count, 0, 16);
  if (rc == CLW_OK) rc = clw_memset(ctx, s->flags, 0, 16);
  if (rc == CLW_OK) rc = clw_memset(ctx, s->total, 0, 16);
  return rc;
}
