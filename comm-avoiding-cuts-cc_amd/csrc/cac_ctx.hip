// cac_ctx.hip -- context lifecycle, errors, device buffers and kernel timing.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

#include "cac_common.h"
#include "cac_internal.h"

namespace cac {

int fail(cac_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->last_error = msg;
  return code;
}

int hip_fail(cac_ctx* ctx, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory) return fail(ctx, CAC_ENOMEM, m);
  return fail(ctx, CAC_EDEVICE, m);
}

int bind(cac_ctx* ctx) {
  if (!ctx) return CAC_EINVAL;
  CAC_HIP(ctx, hipSetDevice(ctx->device));
  return CAC_OK;
}

int scratch_reserve(cac_ctx* ctx, size_t bytes, void** base) {
  if (bytes > ctx->scratch_bytes) {
    if (ctx->scratch) {
      CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      CAC_HIP(ctx, hipDeviceSynchronize());
      CAC_HIP(ctx, hipFree(ctx->scratch));
      ctx->scratch = nullptr;
      ctx->scratch_bytes = 0;
    }
    size_t want = bytes + bytes / 8;
    hipError_t e = hipMalloc(&ctx->scratch, want);
    if (e != hipSuccess) {
      ctx->scratch = nullptr;
      return hip_fail(ctx, e, "hipMalloc(scratch)");
    }
    ctx->scratch_bytes = want;
  }
  *base = ctx->scratch;
  return CAC_OK;
}

static hipEvent_t take_event(cac_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void prof_begin(cac_ctx* ctx, const char* name, hipStream_t s) {
  if (!ctx->profiling) return;
  cac_prof_rec r{name, take_event(ctx), take_event(ctx)};
  if (r.start) (void)hipEventRecord(r.start, s);
  ctx->prof_open.push_back(ctx->prof.size());
  ctx->prof.push_back(r);
}

void prof_end(cac_ctx* ctx, hipStream_t s) {
  if (!ctx->profiling || ctx->prof_open.empty()) return;
  cac_prof_rec& r = ctx->prof[ctx->prof_open.back()];
  ctx->prof_open.pop_back();
  if (r.stop) (void)hipEventRecord(r.stop, s);
}

// Device-to-device copy as a plain streaming kernel: 16-byte vector loads and
// stores, grid-stride, tail handled bytewise.  It reads and writes exactly
// `bytes` bytes each, which makes it the calibration kernel for HBM byte counters.
__global__ void __launch_bounds__(256) k_copy(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                              size_t n16, unsigned char* __restrict__ dtail,
                                              const unsigned char* __restrict__ stail, unsigned tail) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
  if (blockIdx.x == 0 && threadIdx.x < tail) dtail[threadIdx.x] = stail[threadIdx.x];
}

// Frees every grow-only buffer except `keep` whose id is not pinned.
static void release_unpinned(cac_ctx* ctx, void** keep) {
  (void)hipDeviceSynchronize();
  auto pinned = [&](int id) {
    for (int p : ctx->pins)
      if (p == id) return true;
    return false;
  };
  auto drop = [&](int id, void** b, size_t* c) {
    if (b == keep || !*b || pinned(id)) return;
    (void)hipFree(*b);
    *b = nullptr;
    *c = 0;
  };
  drop(kBufKs0, &ctx->ks_buf[0], &ctx->ks_bytes[0]);
  drop(kBufKs1, &ctx->ks_buf[1], &ctx->ks_bytes[1]);
  drop(kBufKsRetry, &ctx->ks_retry_buf, &ctx->ks_retry_bytes);
  drop(kBufSmp, &ctx->smp_buf, &ctx->smp_bytes);
  drop(kBufLt, &ctx->lt_buf, &ctx->lt_bytes);
  drop(kBufSpDense, &ctx->sp_dense_buf, &ctx->sp_dense_bytes);
  drop(kBufSpTmp, &ctx->sp_tmp, &ctx->sp_tmp_bytes);
  drop(kBufServe, &ctx->serve_buf, &ctx->serve_bytes);
  for (size_t d = 0; d < ctx->sp_bufs.size(); ++d) drop(kBufSpDepth + (int)d, &ctx->sp_bufs[d], &ctx->sp_bytes[d]);
}

int grow_buffer(cac_ctx* ctx, void** buf, size_t* cap, size_t bytes) {
  if (bytes <= *cap && *buf) return CAC_OK;
  if (*buf) CAC_HIP(ctx, hipFree(*buf));
  *buf = nullptr;
  *cap = 0;
  // 1/8 headroom (1/2 with grow_wide), at least 1 MiB, so a later call a
  // little larger than this one does not reallocate (a large hipFree +
  // hipMalloc costs far more than the memory: the free waits for the device)
  const size_t want = std::max<size_t>(bytes + (ctx->grow_wide ? bytes / 2 : bytes / 8), (size_t)1 << 20);
  if (hipMalloc(buf, want) == hipSuccess) {
    *cap = want;
    return CAC_OK;
  }
  (void)hipGetLastError();
  *buf = nullptr;
  if (hipMalloc(buf, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *buf = nullptr;
    release_unpinned(ctx, buf);
    CAC_HIP(ctx, hipMalloc(buf, bytes));
  }
  *cap = bytes;
  return CAC_OK;
}

int device_memory_available(cac_ctx* ctx, size_t* bytes) {
  size_t free_b = 0, total_b = 0;
  CAC_HIP(ctx, hipMemGetInfo(&free_b, &total_b));
  *bytes = free_b + ctx->ks_bytes[0] + ctx->ks_bytes[1] + ctx->smp_bytes + ctx->ks_retry_bytes;
  return CAC_OK;
}

}  // namespace cac

extern "C" {

int cac_ctx_create(int device, cac_ctx** out) {
  if (!out) return CAC_EINVAL;
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return CAC_EDEVICE;
  if (device < 0 || device >= count) return CAC_EINVAL;
  cac_ctx* ctx = new (std::nothrow) cac_ctx();
  if (!ctx) return CAC_ENOMEM;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc(&ctx->pinned, 4096, hipHostMallocDefault) != hipSuccess) {
    delete ctx;
    return CAC_EDEVICE;
  }
  // the static scratch of cac_cc_static_dev: its address on this device, taken
  // now so that entry point never needs to (it retries if this failed)
  (void)cac::cc_static_init(ctx);
  *out = ctx;
  return CAC_OK;
}

int cac_ctx_destroy(cac_ctx* ctx) {
  if (!ctx) return CAC_EINVAL;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& r : ctx->prof) {
    if (r.start) (void)hipEventDestroy(r.start);
    if (r.stop) (void)hipEventDestroy(r.stop);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->cc_graph) (void)hipGraphExecDestroy(ctx->cc_graph);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  for (int b = 0; b < 2; ++b)
    if (ctx->ks_buf[b]) (void)hipFree(ctx->ks_buf[b]);
  if (ctx->smp_buf) (void)hipFree(ctx->smp_buf);
  if (ctx->ks_retry_buf) (void)hipFree(ctx->ks_retry_buf);
  for (void* b : ctx->sp_bufs)
    if (b) (void)hipFree(b);
  if (ctx->sp_dense_buf) (void)hipFree(ctx->sp_dense_buf);
  if (ctx->sp_tmp) (void)hipFree(ctx->sp_tmp);
  if (ctx->lt_buf) (void)hipFree(ctx->lt_buf);
  if (ctx->serve_buf) (void)hipFree(ctx->serve_buf);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CAC_OK;
}

const char* cac_strerror(int code) {
  switch (code) {
    case CAC_OK: return "ok";
    case CAC_EINVAL: return "invalid argument";
    case CAC_ENOMEM: return "out of memory";
    case CAC_EDEVICE: return "device error";
    case CAC_ENOTCONN: return "graph is not connected";
    case CAC_ERANGE: return "weight sum overflows int64";
    case CAC_EIO: return "input file cannot be read";
    default: return "unknown error";
  }
}

const char* cac_last_error(const cac_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

void* cac_ctx_stream(cac_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int cac_ctx_synchronize(cac_ctx* ctx) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CAC_OK;
}

int cac_dev_alloc(cac_ctx* ctx, size_t bytes, void** dptr) {
  if (!dptr) return CAC_EINVAL;
  int rc = cac::bind(ctx);
  if (rc) return rc;
  CAC_HIP(ctx, hipMalloc(dptr, bytes ? bytes : 1));
  return CAC_OK;
}

int cac_dev_free(cac_ctx* ctx, void* dptr) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  CAC_HIP(ctx, hipFree(dptr));
  return CAC_OK;
}

int cac_memcpy_h2d(cac_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!bytes) return CAC_OK;
  if (!dst || !src) return CAC_EINVAL;
  CAC_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CAC_OK;
}

int cac_memcpy_d2h(cac_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!bytes) return CAC_OK;
  if (!dst || !src) return CAC_EINVAL;
  CAC_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  CAC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CAC_OK;
}

int cac_memcpy_d2d(cac_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  if (!bytes) return CAC_OK;
  if (!dst || !src) return CAC_EINVAL;
  if (((uintptr_t)dst | (uintptr_t)src) & 15) return cac::fail(ctx, CAC_EINVAL, "cac_memcpy_d2d: pointers must be 16-byte aligned");
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  const size_t n16 = bytes / 16;
  const unsigned tail = (unsigned)(bytes % 16);
  cac::prof_begin(ctx, "copy", s);
  hipLaunchKernelGGL(cac::k_copy, dim3(cac::grid_for(n16, 256, 8192)), dim3(256), 0, s, (uint4*)dst,
                     (const uint4*)src, n16, (unsigned char*)dst + n16 * 16,
                     (const unsigned char*)src + n16 * 16, tail);
  CAC_LAUNCH_CHECK(ctx, "k_copy");
  cac::prof_end(ctx, s);
  return CAC_OK;
}

// Tuning table: CAC_TUNE_KS_CONTRACT_BLOCK keys 512, 1024, 2048, 4096 (row-size
// class upper bounds), a value 0 (built-in rule) or a power of two from key / 8
// (at least 64) to 1024; CAC_TUNE_SAMPLE_OCCUPANCY keys 16, 20, 24, 32 (log2 of
// the edge-count class bound), a value 0 (built-in: 2), 1 or 2;
// CAC_TUNE_WIDE_SAMPLING key 0, a value 0 (built-in: 2^20 edges) to 63 (never);
// CAC_TUNE_KS_COMPACT_BLOCK keys 512, 1024, 2048, 4096, 8192 (child-size class
// upper bounds), a value 0 (built-in: 512), 256, 512 or 1024.
static int tune_slot(int knob, uint32_t key) {
  if (knob == CAC_TUNE_KS_CONTRACT_BLOCK) {
    switch (key) {
      case 512: return 0;
      case 1024: return 1;
      case 2048: return 2;
      case 4096: return 3;
      default: return -1;
    }
  }
  if (knob == CAC_TUNE_SAMPLE_OCCUPANCY) {
    switch (key) {
      case 16: return 0;
      case 20: return 1;
      case 24: return 2;
      case 32: return 3;
      default: return -1;
    }
  }
  if (knob == CAC_TUNE_WIDE_SAMPLING) return key == 0 ? 0 : -1;
  if (knob == CAC_TUNE_KS_COMPACT_BLOCK) {
    switch (key) {
      case 512: return 0;
      case 1024: return 1;
      case 2048: return 2;
      case 4096: return 3;
      case 8192: return 4;
      default: return -1;
    }
  }
  return -1;
}

int cac_tune_set(cac_ctx* ctx, int knob, uint32_t key, int32_t value) {
  if (!ctx) return CAC_EINVAL;
  const int slot = tune_slot(knob, key);
  if (slot < 0) return cac::fail(ctx, CAC_EINVAL, "unknown tuning knob or key");
  if (knob == CAC_TUNE_SAMPLE_OCCUPANCY) {
    if (value < 0 || value > 2) return cac::fail(ctx, CAC_EINVAL, "sampling workgroups per CU must be 0, 1 or 2");
    ctx->sample_occupancy[slot] = value;
    return CAC_OK;
  }
  if (knob == CAC_TUNE_WIDE_SAMPLING) {
    if (value < 0 || value > 63) return cac::fail(ctx, CAC_EINVAL, "wide sampling threshold must be 0 to 63");
    ctx->wide_log2 = value;
    return CAC_OK;
  }
  if (knob == CAC_TUNE_KS_COMPACT_BLOCK) {
    if (value != 0 && value != 256 && value != 512 && value != 1024)
      return cac::fail(ctx, CAC_EINVAL, "compaction block size must be 0, 256, 512 or 1024");
    ctx->ks_compact_block[slot] = value;
    return CAC_OK;
  }
  const int32_t lo = key / 8 < 64 ? 64 : (int32_t)(key / 8);
  if (value != 0 && (value < lo || value > 1024 || (value & (value - 1))))
    return cac::fail(ctx, CAC_EINVAL, "block size must be 0 or a power of two from key/8 (at least 64) to 1024");
  ctx->ks_contract_block[slot] = value;
  return CAC_OK;
}

int cac_tune_get(cac_ctx* ctx, int knob, uint32_t key, int32_t* value) {
  if (!ctx || !value) return CAC_EINVAL;
  const int slot = tune_slot(knob, key);
  if (slot < 0) return cac::fail(ctx, CAC_EINVAL, "unknown tuning knob or key");
  *value = knob == CAC_TUNE_SAMPLE_OCCUPANCY ? ctx->sample_occupancy[slot]
           : knob == CAC_TUNE_WIDE_SAMPLING  ? ctx->wide_log2
           : knob == CAC_TUNE_KS_COMPACT_BLOCK ? ctx->ks_compact_block[slot]
                                               : ctx->ks_contract_block[slot];
  return CAC_OK;
}

int cac_prof_enable(cac_ctx* ctx, int on) {
  if (!ctx) return CAC_EINVAL;
  ctx->profiling = on != 0;
  return CAC_OK;
}

int cac_prof_reset(cac_ctx* ctx) {
  int rc = cac::bind(ctx);
  if (rc) return rc;
  CAC_HIP(ctx, hipDeviceSynchronize());
  for (auto& r : ctx->prof) {
    if (r.start) ctx->event_pool.push_back(r.start);
    if (r.stop) ctx->event_pool.push_back(r.stop);
  }
  ctx->prof.clear();
  ctx->prof_open.clear();
  return CAC_OK;
}

int cac_prof_count(cac_ctx* ctx) { return ctx ? (int)ctx->prof.size() : CAC_EINVAL; }

int cac_prof_get(cac_ctx* ctx, int i, const char** name, float* ms) {
  if (!ctx || i < 0 || i >= (int)ctx->prof.size()) return CAC_EINVAL;
  int rc = cac::bind(ctx);
  if (rc) return rc;
  cac_prof_rec& r = ctx->prof[(size_t)i];
  if (!r.start || !r.stop) return cac::fail(ctx, CAC_EDEVICE, "event creation failed");
  CAC_HIP(ctx, hipEventSynchronize(r.stop));
  float t = 0.f;
  CAC_HIP(ctx, hipEventElapsedTime(&t, r.start, r.stop));
  if (name) *name = r.name;
  if (ms) *ms = t;
  return CAC_OK;
}

}  // extern "C"
