// Grouped scatter-add of Bernstein-Bezier adjoint records: counting sort by
// target block, then one thread per distinct key sums its group and adds it
// to the target with plain loads and stores.
#include <hipcub/hipcub.hpp>

#include "scatter.h"

namespace cmm {

namespace {
inline int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// hipCUB takes the bin count (key_range + 1) as an int: the largest key
// range whose bin count is still a positive int
constexpr int64_t kMaxKeyRange = (int64_t)0x7fffffff - 1;

// key k has 1 .. split records / more than split records
struct ShortGroup {
  const uint32_t* offs;
  uint32_t split;
  __host__ __device__ bool operator()(const uint32_t& k) const {
    const uint32_t d = offs[k + 1] - offs[k];
    return d != 0u && d <= split;
  }
};
struct LongGroup {
  const uint32_t* offs;
  uint32_t split;
  __host__ __device__ bool operator()(const uint32_t& k) const { return offs[k + 1] - offs[k] > split; }
};

hipError_t list_keys(void* temp, size_t& bytes, const uint32_t* offs, uint32_t split, uint32_t* keys_u,
                     uint32_t* keys_l, uint32_t* meta, int nkeys, hipStream_t st) {
  return hipcub::DevicePartition::If(temp, bytes, hipcub::CountingInputIterator<uint32_t>(0), keys_u, keys_l,
                                     hipcub::DiscardOutputIterator<>(), meta, nkeys, ShortGroup{offs, split},
                                     LongGroup{offs, split}, st);
}

// temporary storage of the scan over the bins and of the key lists (the
// larger of the two)
size_t scan_temp_bytes(int64_t bins) {
  size_t scan = 0, sel = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                  (int)bins, (hipStream_t)0);
  if (e != hipSuccess) return 0;
  e = list_keys(nullptr, sel, nullptr, 1, nullptr, nullptr, nullptr, (int)(bins - 1), (hipStream_t)0);
  if (e != hipSuccess) return 0;
  return std::max(scan, sel);
}
}  // namespace

int64_t scatter_bytes(int64_t n, int64_t key_range) {
  if (n < 1 || n > (int64_t)0x7fffffff || key_range < 1 || key_range > kMaxKeyRange) return 0;
  const int64_t bins = key_range + 1;
  const size_t tb = scan_temp_bytes(bins);
  if (!tb) return 0;
  return 5 * align256(4 * n) + 2 * align256(4 * bins) + 256 + align256(kScatterRecBytes * n) + align256((int64_t)tb);
}

int scatter_layout(void* work, int64_t work_bytes, int64_t n, int64_t key_range, ScatterWork& w) {
  if (key_range < 1 || key_range > kMaxKeyRange)
    return fail(CMM_E_RANGE, "scatter workspace: key range above 2^31 - 2");
  const int64_t need = scatter_bytes(n, key_range);
  if (!need) return fail(CMM_E_RANGE, "scatter workspace: bad record count");
  if (!work || work_bytes < need)
    return fail(CMM_E_ARG, "scatter workspace too small: need " + std::to_string(need) + " bytes");
  char* p = static_cast<char*>(work);
  auto take = [&](int64_t b) {
    char* q = p;
    p += align256(b);
    return q;
  };
  const int64_t bins = key_range + 1;
  w.n = n;
  w.keys = reinterpret_cast<uint32_t*>(take(4 * n));
  w.keys_s = reinterpret_cast<uint32_t*>(take(4 * n));
  w.idx_s = reinterpret_cast<uint32_t*>(take(4 * n));
  w.count = reinterpret_cast<uint32_t*>(take(4 * bins));
  w.offs = reinterpret_cast<uint32_t*>(take(4 * bins));
  w.keys_u = reinterpret_cast<uint32_t*>(take(4 * n));
  w.keys_l = reinterpret_cast<uint32_t*>(take(4 * n));
  w.meta = reinterpret_cast<uint32_t*>(take(8));
  char* recs = take(kScatterRecBytes * n);
  w.rec = reinterpret_cast<ScatterRec*>(recs);
  w.recv = reinterpret_cast<ScatterRecV*>(recs);
  w.empty = (uint32_t)key_range;
  w.temp_bytes = scan_temp_bytes(bins);
  w.temp = take((int64_t)w.temp_bytes);
  return CMM_OK;
}

// zero the per-key counters (the producer counts its records into them)
__global__ __launch_bounds__(256) void k_scatter_init(uint32_t* __restrict__ count, int64_t bins) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < bins) count[i] = 0u;
}

// slot i goes to the next free position of its key's range (the counters
// hold the group sizes on entry and are left at zero); empty slots are not
// placed: positions from offs[empty] on are never read
__global__ __launch_bounds__(256) void k_scatter_place(const uint32_t* __restrict__ keys, int64_t n, uint32_t empty,
                                                       const uint32_t* __restrict__ offs,
                                                       uint32_t* __restrict__ count, uint32_t* __restrict__ keys_s,
                                                       uint32_t* __restrict__ idx_s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  if (k >= empty) return;
  const uint32_t pos = offs[k] + __hip_atomic_fetch_add(count + k, 0xffffffffu, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT) - 1u;
  keys_s[pos] = k;
  idx_s[pos] = (uint32_t)i;
}

// Records per reducing thread: a group (records of one key) is cut into
// chunks that end at multiples of kChunk.  A group inside one chunk is added
// with plain loads / stores (its only writer); the chunks of a longer group
// (e.g. the 2L samples of a pole ring, all in one sub-triangle) add with
// atomics.  Every chunk of a group sees the same group extent, so a target
// block is never written both ways.
constexpr int kChunk = 16;

// add the 6 NC sums of one block into a packed-layout adjoint array
template <int NC>
__device__ inline void add_block(double* base, const double* acc, bool atomic) {
  if (atomic) {
#pragma unroll
    for (int e = 0; e < NC * 6; ++e) unsafeAtomicAdd(base + e, acc[e]);
  } else {
    double2* dst = reinterpret_cast<double2*>(base);
#pragma unroll
    for (int e = 0; e < NC * 3; ++e) {
      double2 v = dst[e];
      v.x += acc[2 * e];
      v.y += acc[2 * e + 1];
      dst[e] = v;
    }
  }
}

template <int NC, int MODE>
__global__ __launch_bounds__(256) void k_scatter_reduce(ScatterWork w, int64_t n, ScatterTargets t) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = w.offs[w.empty];  // records placed
  if (j >= total) return;
  const uint32_t key = w.keys_s[j];
  const bool starts_group = j == 0 || w.keys_s[j - 1] != key;
  if (!starts_group && j % kChunk != 0) return;  // inside another thread's chunk
  constexpr int NO = MODE == 0 ? 1 : 3;
  double acc[NO][NC * 6];
#pragma unroll
  for (int o = 0; o < NO; ++o)
#pragma unroll
    for (int e = 0; e < NC * 6; ++e) acc[o][e] = 0.0;
  const int64_t stop = (j / kChunk + 1) * kChunk;  // next chunk boundary
  int64_t r = j;
  for (; r < total && r < stop && w.keys_s[r] == key; ++r) {
    if (MODE == 2) {
      const ScatterRecV& rv = w.recv[w.idx_s[r]];
      const int cl = rv.cls;
#pragma unroll
      for (int o = 0; o < NO; ++o) {
        if (o < t.nout) {
          const double wt = t.W[cl][o];
#pragma unroll
          for (int e = 0; e < NC * 6; ++e) acc[o][e] += wt * rv.v[e];
        }
      }
      continue;
    }
    const ScatterRec rc = w.rec[w.idx_s[r]];
    const d3 cb = d3{rc.cb[0], rc.cb[1], rc.cb[2]};
    const double a[3] = {rc.a[0], rc.a[1], rc.a[2]};
    double bas[6];
    bb6_basis(cb, bas);
    if (MODE == 0) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[0][6 * c + k] += a[c] * bas[k];
    } else {
      const int cl = rc.cls;
#pragma unroll
      for (int o = 0; o < NO; ++o) {
        if (o < t.nout) {
          const double wt = t.W[cl][o];
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[o][6 * c + k] += wt * a[c] * bas[k];
        }
      }
    }
  }
  const bool ends_group = r >= total || w.keys_s[r] != key;
  const bool atomic = !(starts_group && ends_group);
  if (MODE == 0) {
    const int cl = (int)(key / (uint64_t)t.span);
    const int64_t blk = (int64_t)(key - (uint64_t)cl * t.span);
    add_block<NC>(t.T[cl] + blk * (NC * 6), acc[0], atomic);
  } else {
#pragma unroll
    for (int o = 0; o < NO; ++o)
      if (o < t.nout) add_block<NC>(t.T[o] + (int64_t)key * (NC * 6), acc[o], atomic);
  }
}

int scatter_begin(const ScatterWork& w, hipStream_t st) {
  const int64_t bins = (int64_t)w.empty + 1;
  hipLaunchKernelGGL(k_scatter_init, dim3(grid_for(bins, 256)), dim3(256), 0, st, w.count, bins);
  CMM_LAUNCH_CHECK("scatter: init");
  return CMM_OK;
}

int scatter_group(const ScatterWork& w, int64_t n, hipStream_t st, uint32_t split) {
  if (n < 1) return CMM_OK;
  if (n > w.n) return fail(CMM_E_ARG, "scatter: more records than the workspace holds");
  const int64_t bins = (int64_t)w.empty + 1;
  size_t tb = w.temp_bytes;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(w.temp, tb, w.count, w.offs, (int)bins, st);
  if (e != hipSuccess) return hip_check(e, "scatter: scan");
  if (split) {
    tb = w.temp_bytes;
    e = list_keys(w.temp, tb, w.offs, split, w.keys_u, w.keys_l, w.meta, (int)w.empty, st);
    if (e != hipSuccess) return hip_check(e, "scatter: key lists");
  }
  hipLaunchKernelGGL(k_scatter_place, dim3(grid_for(n, 256)), dim3(256), 0, st, w.keys, n, w.empty, w.offs, w.count,
                     w.keys_s, w.idx_s);
  CMM_LAUNCH_CHECK("scatter: place");
  return CMM_OK;
}

int scatter_reduce(const ScatterWork& w, int64_t n, const ScatterTargets& t, hipStream_t st) {
  if (n < 1) return CMM_OK;
  int rc = scatter_group(w, n, st);
  if (rc) return rc;
  const int grid = grid_for(n, 256);
  if (t.mode == 0) {
    if (t.nc == 3)
      hipLaunchKernelGGL((k_scatter_reduce<3, 0>), dim3(grid), dim3(256), 0, st, w, n, t);
    else
      hipLaunchKernelGGL((k_scatter_reduce<1, 0>), dim3(grid), dim3(256), 0, st, w, n, t);
  } else if (t.mode == 1) {
    if (t.nc == 3)
      hipLaunchKernelGGL((k_scatter_reduce<3, 1>), dim3(grid), dim3(256), 0, st, w, n, t);
    else
      hipLaunchKernelGGL((k_scatter_reduce<1, 1>), dim3(grid), dim3(256), 0, st, w, n, t);
  } else {
    if (t.nc != 3) return fail(CMM_E_ARG, "scatter: value records need 3 components");
    hipLaunchKernelGGL((k_scatter_reduce<3, 2>), dim3(grid), dim3(256), 0, st, w, n, t);
  }
  CMM_LAUNCH_CHECK("scatter: reduce");
  return CMM_OK;
}

}  // namespace cmm

extern "C" int cmm_scatter_workspace_bytes(int64_t records, int64_t key_range, int64_t* bytes) {
  if (!bytes || records < 1 || key_range < 1) return cmm::fail(CMM_E_ARG, "cmm_scatter_workspace_bytes: bad arguments");
  if (records > (int64_t)0x7fffffff || key_range > cmm::kMaxKeyRange)
    return cmm::fail(CMM_E_RANGE, "cmm_scatter_workspace_bytes: record count or key range out of range");
  const int64_t b = cmm::scatter_bytes(records, key_range);
  if (!b) return cmm::fail(CMM_E_HIP, "cmm_scatter_workspace_bytes: hipCUB temporary storage query failed");
  *bytes = b;
  return CMM_OK;
}
