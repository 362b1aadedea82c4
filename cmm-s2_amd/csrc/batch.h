// Member-batched launches.
//
// A kernel that can run batched is written as a struct K with
//   struct Args { ... };                       its arguments,
//   static constexpr int kThreads;             its launch bound,
//   static __device__ void run(const Args&, int bid);   its body, with the
//                                              block index as a parameter,
// and a __global__ wrapper that calls K::run(a, blockIdx.x) (the kernel the
// library launches on its own).  While a batch records on the host thread
// (cmm_batch_begin .. cmm_batch_launch), launch<K>() records the launch for
// the current member instead of issuing it; cmm_batch_launch then issues ONE
// grid per recorded position for all members, k_batched<K>, whose block b
// runs block b - start[i] of member i (start[i] <= b < start[i + 1]) with
// member i's arguments.  Every block runs the code and arguments it would
// run in the member's own launch, so batched members give the bits of their
// own runs.
#pragma once

#include "common.h"

namespace cmm {

// member of block b: the largest i with start[i] <= b (members with no
// blocks have start[i] == start[i + 1] and are skipped)
__device__ __forceinline__ int batch_member(const int* __restrict__ start, int nm, int b) {
  int lo = 0, hi = nm;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

// the member's arguments, copied from the constant address space at entry
// (uniform scalar loads, as from the kernel-argument segment of a solo launch)
template <class A>
__device__ __forceinline__ A batch_args(const A* args, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) A CA;
  const CA* pa = (const CA*)args;
  return pa[i];
#else
  return args[i];  // host pass: the kernel body is not compiled for the host
#endif
}

// arguments with a large table indexed at run time (the composition's 64
// level pointers) are read in place instead of copied (a copy does not fit
// the scalar registers): K::kArgsInPlace
template <class K, class = void>
struct ArgsInPlace {
  static constexpr bool value = false;
};
template <class K>
struct ArgsInPlace<K, decltype((void)K::kArgsInPlace)> {
  static constexpr bool value = K::kArgsInPlace;
};

template <class K>
__device__ __forceinline__ void batch_run(const typename K::Args* __restrict__ args, const int* __restrict__ start,
                                          int nm) {
  const int b = (int)blockIdx.x;
  const int i = batch_member(start, nm, b);
  if constexpr (ArgsInPlace<K>::value) K::run(args[i], b - start[i]);
  else K::run(batch_args(args, i), b - start[i]);
}

template <class K>
__global__ __launch_bounds__(K::kThreads) void k_batched(const typename K::Args* __restrict__ args,
                                                         const int* __restrict__ start, int nm) {
  batch_run<K>(args, start, nm);
}

// kernels with a minimum occupancy in their launch bounds
template <class K>
__global__ __launch_bounds__(K::kThreads, K::kMinEU) void k_batched_eu(const typename K::Args* __restrict__ args,
                                                                       const int* __restrict__ start, int nm) {
  batch_run<K>(args, start, nm);
}

template <class K, class = void>
struct BatchedFn {
  static const void* get() { return reinterpret_cast<const void*>(&k_batched<K>); }
};
template <class K>
struct BatchedFn<K, decltype((void)K::kMinEU)> {
  static const void* get() { return reinterpret_cast<const void*>(&k_batched_eu<K>); }
};

// host: is a batch recording on this thread; record one launch
bool batch_recording();
int batch_record(const void* fn, unsigned blocks, unsigned threads, size_t lds, const void* args, size_t size,
                 size_t align, const char* what);

// Issue (or record) kernel K with `blocks` blocks of `threads` threads:
// `solo` is K's own __global__ wrapper.
template <class K>
int launch(void (*solo)(typename K::Args), unsigned blocks, unsigned threads, size_t lds, hipStream_t st,
           const typename K::Args& a, const char* what) {
  if (batch_recording())
    return batch_record(BatchedFn<K>::get(), blocks, threads, lds, &a, sizeof(a), alignof(typename K::Args), what);
  if (blocks == 0) return CMM_OK;
  hipLaunchKernelGGL(solo, dim3(blocks), dim3(threads), lds, st, a);
  CMM_LAUNCH_CHECK(what);
  return CMM_OK;
}

}  // namespace cmm
