// Sorted scatter-add of Bernstein-Bezier adjoint records: radix sort by
// target block, then one thread per distinct key sums its group and adds it
// to the target with plain loads and stores.
#include <rocprim/device/device_radix_sort.hpp>

#include "scatter.h"

namespace cmm {

namespace {
inline int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// Onesweep radix sort at every size (rocPRIM's default switches to a merge
// sort of ~17 launches below 2^20 items, the record counts here) with 6 key
// bits per pass and 2048 items per block: measured on 655k / 786k pairs of
// 21-22-bit keys, 125 / 134 us against 145 / 150 us for the default (and
// 221 us with 8 bits per pass).
using SortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<256, 12>, rocprim::kernel_config<256, 8>, 6>, 0>;

hipError_t sort_pairs(void* temp, size_t& bytes, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                      uint32_t* vout, int64_t n, int end_bit, hipStream_t st) {
  return rocprim::radix_sort_pairs<SortConfig>(temp, bytes, kin, kout, vin, vout, (unsigned int)n, 0u,
                                               (unsigned int)end_bit, st);
}

size_t sort_temp_bytes(int64_t n) {
  size_t bytes = 0;
  hipError_t e = sort_pairs(nullptr, bytes, nullptr, nullptr, nullptr, nullptr, n, 32, (hipStream_t)0);
  return e == hipSuccess ? bytes : 0;
}
}  // namespace

int64_t scatter_bytes(int64_t n) {
  if (n < 1 || n > (int64_t)0x7fffffff) return 0;
  const size_t tb = sort_temp_bytes(n);
  if (!tb) return 0;
  return 4 * align256(4 * n) + align256(kScatterRecBytes * n) + align256((int64_t)tb);
}

int scatter_layout(void* work, int64_t work_bytes, int64_t n, int64_t key_range, ScatterWork& w) {
  if (key_range < 1 || key_range >= (int64_t)kScatterEmpty)
    return fail(CMM_E_RANGE, "scatter workspace: keys do not fit 32 bits");
  const int64_t need = scatter_bytes(n);
  if (!need) return fail(CMM_E_RANGE, "scatter workspace: bad record count");
  if (!work || work_bytes < need)
    return fail(CMM_E_ARG, "scatter workspace too small: need " + std::to_string(need) + " bytes");
  char* p = static_cast<char*>(work);
  auto take = [&](int64_t b) {
    char* q = p;
    p += align256(b);
    return q;
  };
  w.n = n;
  w.keys = reinterpret_cast<uint32_t*>(take(4 * n));
  w.idx = reinterpret_cast<uint32_t*>(take(4 * n));
  w.keys_s = reinterpret_cast<uint32_t*>(take(4 * n));
  w.idx_s = reinterpret_cast<uint32_t*>(take(4 * n));
  char* recs = take(kScatterRecBytes * n);
  w.rec = reinterpret_cast<ScatterRec*>(recs);
  w.recv = reinterpret_cast<ScatterRecV*>(recs);
  w.empty = (uint32_t)key_range;
  w.end_bit = 1;
  while (w.end_bit < 32 && ((int64_t)1 << w.end_bit) <= key_range) ++w.end_bit;
  w.temp_bytes = sort_temp_bytes(n);
  w.temp = take((int64_t)w.temp_bytes);
  return CMM_OK;
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
  if (j >= n) return;
  const uint32_t key = w.keys_s[j];
  if (key >= w.empty) return;
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
  for (; r < n && r < stop && w.keys_s[r] == key; ++r) {
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
  const bool ends_group = r >= n || w.keys_s[r] != key;
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

int scatter_reduce(const ScatterWork& w, int64_t n, const ScatterTargets& t, hipStream_t st) {
  if (n < 1) return CMM_OK;
  if (n > w.n) return fail(CMM_E_ARG, "scatter: more records than the workspace holds");
  size_t tb = w.temp_bytes;
  hipError_t e = sort_pairs(w.temp, tb, w.keys, w.keys_s, w.idx, w.idx_s, n, w.end_bit, st);
  if (e != hipSuccess) return hip_check(e, "scatter: radix sort");
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

extern "C" int cmm_scatter_workspace_bytes(int64_t records, int64_t* bytes) {
  if (!bytes || records < 1) return cmm::fail(CMM_E_ARG, "cmm_scatter_workspace_bytes: bad arguments");
  const int64_t b = cmm::scatter_bytes(records);
  if (!b) return cmm::fail(CMM_E_RANGE, "cmm_scatter_workspace_bytes: record count out of range");
  *bytes = b;
  return CMM_OK;
}
