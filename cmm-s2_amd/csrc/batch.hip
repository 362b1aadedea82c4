// Host side of member-batched launches (batch.h): a batch object owns a
// device buffer for the per-member arguments and block prefixes of one step
// and two pinned staging buffers (the one a pending copy reads is waited for
// before it is refilled).  Recording is per host thread.
#include <algorithm>
#include <cstring>
#include <vector>

#include "batch.h"

using namespace cmm;

namespace {
struct Rec {
  const void* fn;
  unsigned blocks, threads;
  size_t lds, size, align;
  const char* what;
  std::vector<unsigned char> args;
};
}  // namespace

struct cmm_batch {
  unsigned char* dev = nullptr;
  unsigned char* host[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  int64_t cap = 0;
  int cur = 0;
  int member = -1;
  std::vector<std::vector<Rec>> recs;
};

namespace {
thread_local cmm_batch* t_active = nullptr;
int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }
}  // namespace

namespace cmm {
bool batch_recording() { return t_active != nullptr; }

int batch_record(const void* fn, unsigned blocks, unsigned threads, size_t lds, const void* args, size_t size,
                 size_t align, const char* what) {
  cmm_batch* b = t_active;
  if (!b || b->member < 0) return fail(CMM_E_STATE, std::string(what) + ": batched launch outside a member");
  Rec r{fn, blocks, threads, lds, size, align, what, {}};
  r.args.assign(static_cast<const unsigned char*>(args), static_cast<const unsigned char*>(args) + size);
  b->recs.back().push_back(std::move(r));
  return CMM_OK;
}
}  // namespace cmm

extern "C" int cmm_batch_create(int64_t bytes, cmm_batch** out) {
  if (bytes < 1024 || !out) return fail(CMM_E_ARG, "cmm_batch_create: bad arguments");
  cmm_batch* b = new cmm_batch();
  b->cap = align_up(bytes, 256);
  hipError_t e = hipMalloc(&b->dev, b->cap);
  for (int k = 0; k < 2 && e == hipSuccess; ++k) {
    e = hipHostMalloc(&b->host[k], b->cap);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    cmm_batch_destroy(b);
    return hip_check(e, "cmm_batch_create");
  }
  *out = b;
  return CMM_OK;
}

extern "C" int cmm_batch_destroy(cmm_batch* b) {
  if (!b) return CMM_OK;
  if (t_active == b) t_active = nullptr;
  for (int k = 0; k < 2; ++k) {
    if (b->ev[k]) {
      (void)hipEventSynchronize(b->ev[k]);
      (void)hipEventDestroy(b->ev[k]);
    }
    (void)hipHostFree(b->host[k]);
  }
  (void)hipFree(b->dev);
  delete b;
  return CMM_OK;
}

extern "C" int cmm_batch_begin(cmm_batch* b) {
  if (!b) return fail(CMM_E_ARG, "cmm_batch_begin: bad arguments");
  if (t_active) return fail(CMM_E_STATE, "cmm_batch_begin: a batch is already recording on this thread");
  b->recs.clear();
  b->member = -1;
  t_active = b;
  return CMM_OK;
}

extern "C" int cmm_batch_member(cmm_batch* b, int i) {
  if (!b || t_active != b) return fail(CMM_E_STATE, "cmm_batch_member: the batch is not recording");
  if (i != (int)b->recs.size()) return fail(CMM_E_ARG, "cmm_batch_member: members must come in order 0, 1, ...");
  b->recs.emplace_back();
  b->member = i;
  return CMM_OK;
}

extern "C" int cmm_batch_cancel(cmm_batch* b) {
  if (!b) return fail(CMM_E_ARG, "cmm_batch_cancel: bad arguments");
  if (t_active == b) t_active = nullptr;
  b->recs.clear();
  b->member = -1;
  return CMM_OK;
}

// Ends recording and issues, for every recorded position j, one grid with
// the blocks of all members' j-th launches.  Every member must have recorded
// the same kernels in the same order with the same block size.
extern "C" int cmm_batch_launch(cmm_batch* b, void* stream, int64_t* launches) {
  if (!b || t_active != b) return fail(CMM_E_STATE, "cmm_batch_launch: the batch is not recording");
  t_active = nullptr;
  const int nm = (int)b->recs.size();
  if (launches) *launches = 0;
  if (nm == 0) return CMM_OK;
  const size_t J = b->recs[0].size();
  for (int i = 1; i < nm; ++i)
    if (b->recs[i].size() != J)
      return fail(CMM_E_STATE, "cmm_batch_launch: members recorded different numbers of launches");
  // layout: per position, the block prefix (nm + 1 ints) and the argument
  // array (nm records of the kernel's argument struct)
  std::vector<int64_t> off_start(J), off_args(J);
  int64_t total = 0;
  for (size_t j = 0; j < J; ++j) {
    const Rec& r0 = b->recs[0][j];
    for (int i = 1; i < nm; ++i) {
      const Rec& r = b->recs[i][j];
      if (r.fn != r0.fn || r.threads != r0.threads || r.size != r0.size)
        return fail(CMM_E_STATE, std::string("cmm_batch_launch: members diverge at ") + r0.what);
    }
    off_start[j] = total;
    total = align_up(total + 4 * (int64_t)(nm + 1), 256);
    off_args[j] = total;
    total = align_up(total + (int64_t)r0.size * nm, 256);
  }
  if (total > b->cap) return fail(CMM_E_RANGE, "cmm_batch_launch: arguments exceed the batch buffer");
  hipError_t e = hipEventSynchronize(b->ev[b->cur]);  // the copy that last read this staging buffer
  if (e != hipSuccess) return hip_check(e, "cmm_batch_launch: staging");
  unsigned char* h = b->host[b->cur];
  std::vector<int64_t> blocks(J);
  for (size_t j = 0; j < J; ++j) {
    int* start = reinterpret_cast<int*>(h + off_start[j]);
    int64_t s = 0;
    size_t lds = 0;
    for (int i = 0; i < nm; ++i) {
      const Rec& r = b->recs[i][j];
      start[i] = (int)s;
      s += r.blocks;
      lds = std::max(lds, r.lds);
      std::memcpy(h + off_args[j] + (int64_t)r.size * i, r.args.data(), r.size);
    }
    if (s > 0x7fffffff) return fail(CMM_E_RANGE, "cmm_batch_launch: too many blocks");
    start[nm] = (int)s;
    blocks[j] = s;
    b->recs[0][j].lds = lds;
  }
  hipStream_t st = as_stream(stream);
  e = hipMemcpyAsync(b->dev, h, total, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipEventRecord(b->ev[b->cur], st);
  if (e != hipSuccess) return hip_check(e, "cmm_batch_launch: upload");
  b->cur ^= 1;
  int64_t n = 0;
  for (size_t j = 0; j < J; ++j) {
    if (blocks[j] == 0) continue;
    const Rec& r0 = b->recs[0][j];
    const void* pa = b->dev + off_args[j];
    const int* ps = reinterpret_cast<const int*>(b->dev + off_start[j]);
    int nmv = nm;
    void* kargs[3] = {&pa, &ps, &nmv};
    e = hipLaunchKernel(r0.fn, dim3((unsigned)blocks[j]), dim3(r0.threads), kargs, r0.lds, st);
    if (e != hipSuccess) return hip_check(e, r0.what);
    ++n;
  }
  if (launches) *launches = n;
  b->recs.clear();
  b->member = -1;
  return CMM_OK;
}

// ---------------------------------------------------------------------------
// Device copy of n doubles (the remap's copy of the current map's jets into
// the submap stack), a kernel of the library so that it can run batched.
// ---------------------------------------------------------------------------
struct KCopyF64 {
  struct Args {
    double* dst;
    const double* src;
    int64_t n;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a, int bid) {
    const int64_t i = (int64_t)bid * blockDim.x + threadIdx.x;
    if (i < a.n) a.dst[i] = a.src[i];
  }
};

__global__ __launch_bounds__(256) void k_copy_f64(KCopyF64::Args a) { KCopyF64::run(a, blockIdx.x); }

extern "C" int cmm_copy_f64(double* dst, const double* src, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && (!dst || !src))) return fail(CMM_E_ARG, "cmm_copy_f64: bad arguments");
  if (n == 0) return CMM_OK;
  return launch<KCopyF64>(k_copy_f64, (unsigned)grid_for(n, 256), 256, 0, as_stream(stream),
                          KCopyF64::Args{dst, src, n}, "cmm_copy_f64");
}
