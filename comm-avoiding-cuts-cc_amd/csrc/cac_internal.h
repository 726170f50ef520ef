// cac_internal.h -- context object, error plumbing and scratch arena shared by
// the entry points of libcac.so.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "cac.h"

struct cac_prof_rec {
  const char* name;
  hipEvent_t start, stop;
};

struct cac_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // Grow-only device scratch arena; entry points carve sub-buffers from it.
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // Small pinned host buffer for scalar results.
  void* pinned = nullptr;
  std::string last_error;
  bool profiling = false;
  std::vector<cac_prof_rec> prof;
  std::vector<size_t> prof_open;  // records begun and not yet ended (ranges nest)
  std::vector<hipEvent_t> event_pool;
  // Two grow-only buffers for the levels of the Karger-Stein recursion (parents
  // in one, children in the other).
  void* ks_buf[2] = {nullptr, nullptr};
  size_t ks_bytes[2] = {0, 0};
  // compaction work-item prefix of the last dense Karger-Stein level: its
  // upload may still be in flight when the level returns, so it lives here
  std::vector<uint32_t> ks_items;
  void* ks_retry_buf = nullptr;  // children redone with a full-size row pool
  size_t ks_retry_bytes = 0;
  // Grow-only buffer for the per-trial scratch of the sampling contraction.
  void* smp_buf = nullptr;
  size_t smp_bytes = 0;
  // Sparse Karger-Stein: one grow-only buffer per recursion depth (the children
  // of a chunk of nodes stay alive while their subtree runs), the scratch of
  // one contraction launch, and the dense matrices handed over to the dense
  // engine.
  std::vector<void*> sp_bufs;
  std::vector<size_t> sp_bytes;
  void* sp_tmp = nullptr;  // contraction scratch of one launch
  size_t sp_tmp_bytes = 0;
  void* sp_dense_buf = nullptr;
  size_t sp_dense_bytes = 0;
  // Square-root-cut trials beyond the dense engine: sampling scratch and the
  // merged contracted edge list of one trial.
  void* lt_buf = nullptr;
  size_t lt_bytes = 0;
  // a served launch group's uploads, checks and results
  void* serve_buf = nullptr;
  size_t serve_bytes = 0;
  // Grow-only buffers whose contents are live (ids below); the others may be
  // released when an allocation fails.
  std::vector<int> pins;
  size_t smp_wave = 0;  // sampling workgroups resident at once (0: not asked yet)
  // Grow-only buffers regrow by half again rather than an eighth: for contexts
  // whose chunks vary in size from call to call (serving workers), where each
  // regrowth's hipFree waits for the whole device.
  bool grow_wide = false;
  // Tuned launch shapes (cac_tune_set): dense Karger-Stein contraction block
  // size per row-size class, keys 512, 1024, 2048, 4096; 0 = built-in rule.
  int32_t ks_contract_block[4] = {0, 0, 0, 0};
  // Sampling rule of the square-root cut's shrink phase (cac_set_sampling_rule).
  int sample_rule = CAC_SAMPLE_LC;
  // Sampling workgroups per CU for launches of more trials than CUs, by edge
  // count class m < 2^16, 2^20, 2^24, larger (cac_tune_set); 0 = built-in (2).
  int32_t sample_occupancy[4] = {0, 0, 0, 0};
  // threads per k_ks_compact workgroup by child-size class (<= 512, 1024, 2048,
  // 4096, larger); 0 = built-in (512)
  int32_t ks_compact_block[5] = {0, 0, 0, 0, 0};
  // log2 of the smallest edge count whose sqrt-cut trials run in the wide
  // (grid-per-trial) form: 0 = built-in (2^20), 63 = never
  int32_t wide_log2 = 0;
  // The captured CSR connected-components step (cac_cc_csr_graph_dev) and the
  // arguments it was captured for: offsets, columns, labels, scratch, stream,
  // n, label mode.
  hipGraphExec_t cc_graph = nullptr;
  const void* cc_graph_ptrs[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  uint64_t cc_graph_args[2] = {0, 0};
};

namespace cac {

// Records a failure message on the context and returns the code.
int fail(cac_ctx* ctx, int code, const std::string& msg);
int hip_fail(cac_ctx* ctx, hipError_t e, const char* what);

// Makes the context's device current for the calling thread.
int bind(cac_ctx* ctx);

// Ensures the scratch arena holds at least 'bytes'; returns its base.
int scratch_reserve(cac_ctx* ctx, size_t bytes, void** base);

// Grow-only device buffer (*buf, *cap): reallocated only when 'bytes' exceeds it.
// If the allocation fails, every grow-only buffer of the context that is not
// pinned is released and the allocation is tried once more.
int grow_buffer(cac_ctx* ctx, void** buf, size_t* cap, size_t bytes);

// Ids of the context's grow-only buffers, for pinning.
enum : int { kBufKs0 = 0, kBufKs1, kBufKsRetry, kBufSmp, kBufLt, kBufSpDense, kBufSpTmp, kBufServe, kBufSpDepth = 16 };

// Keeps a grow-only buffer from being released while its contents are live.
struct BufPin {
  cac_ctx* ctx;
  int id;
  BufPin(cac_ctx* c, int i) : ctx(c), id(i) { c->pins.push_back(i); }
  ~BufPin() {
    for (size_t k = ctx->pins.size(); k-- > 0;)
      if (ctx->pins[k] == id) {
        ctx->pins.erase(ctx->pins.begin() + (long)k);
        break;
      }
  }
  BufPin(const BufPin&) = delete;
  BufPin& operator=(const BufPin&) = delete;
};

// Takes the device address of the static connected-components scratch for the
// context's device (once per device); cac_ctx_create calls it at set-up.
int cc_static_init(cac_ctx* ctx);

// Free device memory plus what the context's grow-only buffers already hold.
int device_memory_available(cac_ctx* ctx, size_t* bytes);

inline hipStream_t pick_stream(cac_ctx* ctx, void* s) {
  return s ? (hipStream_t)s : ctx->stream;
}

// Carves aligned sub-buffers from a base pointer.
struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* b) : base((char*)b) {}
  template <class T>
  T* take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = (T*)(base + off);
    off += count * sizeof(T);
    return p;
  }
};
inline size_t carve_size(size_t bytes) { return (bytes + 255) & ~size_t(255); }

// Kernel timing: brackets one launch (or a nested range of them) with events
// when profiling is enabled; prof_end closes the innermost open range.
void prof_begin(cac_ctx* ctx, const char* name, hipStream_t s);
void prof_end(cac_ctx* ctx, hipStream_t s);

}  // namespace cac

#define CAC_HIP(ctx, call)                                  \
  do {                                                      \
    hipError_t _e = (call);                                 \
    if (_e != hipSuccess) return cac::hip_fail(ctx, _e, #call); \
  } while (0)

#define CAC_LAUNCH_CHECK(ctx, what)                          \
  do {                                                       \
    hipError_t _e = hipGetLastError();                       \
    if (_e != hipSuccess) return cac::hip_fail(ctx, _e, what); \
  } while (0)
