// Grouped scatter-add of Bernstein-Bezier adjoint records (internal).
//
// The adjoints of spline evaluation at scattered points add
//   a_c * basis_j(cb)   (c < NC components, j < 6 basis functions)
// into the 6 NC coefficients of the located sub-triangle.  Device-wide FP64
// atomics for that run at ~15 G/s here (they are performed beyond the
// per-XCD L2), so the adjoint kernels write one record per (point, use)
// instead: key = target block, a (3 doubles), cb (3 doubles), class.  A
// counting sort (int32 histogram, one scan, placement by counter) groups
// equal keys and one thread per group adds it with plain stores (one writer
// per target block, no FP64 atomics).
//
// Sequence: scatter_layout, scatter_begin (zero the counters), the producer
// kernel (scatter_put* / scatter_empty: each record counts its key as it is
// written), then scatter_reduce (scan, placement, sums) or scatter_group
// (scan, placement only) and a caller-specific reduction.
#pragma once
#include "common.h"

namespace cmm {

constexpr uint32_t kScatterEmpty = 0xffffffffu;

// one record: 64 bytes, one cache line: the sums are a_c basis_j(cb)
struct alignas(64) ScatterRec {
  double a[3];
  double cb[3];
  int32_t cls;
  int32_t sub;  // sub-triangle of the record's target block
  int32_t pad[2];
};

// a record that carries its 18 sums (NC = 3) directly: partial sums formed
// before the scatter (e.g. over a quad of lanes)
struct alignas(16) ScatterRecV {
  double v[18];
  int32_t cls;
  int32_t sub;
  int32_t pad[2];
};
constexpr int64_t kScatterRecBytes = sizeof(ScatterRecV);  // slot size (both kinds)

struct ScatterWork {
  uint32_t* keys = nullptr;    // (n) written by the producer; `empty` = unused slot
  uint32_t* keys_s = nullptr;  // grouped by key: positions [0, offs[empty]) hold the records
  uint32_t* idx_s = nullptr;   // slot of the record at each grouped position
  uint32_t* count = nullptr;   // (empty + 1) records per key (counted by the producer), then
                               // placement counters, left at zero
  uint32_t* offs = nullptr;    // (empty + 1) first grouped position of each key; offs[empty]
                               // = the number of records
  uint32_t* keys_u = nullptr;  // (n) with key lists: the keys with 1 .. split records, in
                               // increasing order, meta[0] of them
  uint32_t* keys_l = nullptr;  // (n) with key lists: the keys with more than split records,
                               // in increasing order, meta[1] of them
  uint32_t* meta = nullptr;    // (2) list lengths
  ScatterRec* rec = nullptr;   // (n) slots of kScatterRecBytes
  ScatterRecV* recv = nullptr; // the same slots as value records
  void* temp = nullptr;        // scan temporary storage
  size_t temp_bytes = 0;
  int64_t n = 0;
  uint32_t empty = kScatterEmpty;  // key of an unused slot: 1 + the largest key
};

// where a group of records goes
struct ScatterTargets {
  // mode 0: key = class * span + block, target T[class] (classes < 8), one
  //         output per key, weight 1
  // mode 1: key = block, nout targets T[t] with weight W[record class][t]
  // mode 2: as mode 1 with value records (ScatterRecV, NC = 3)
  int mode = 0;
  int nc = 3;
  int nout = 1;
  int64_t span = 0;
  double* T[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double W[4][3] = {{1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0}};
};

// bytes of a workspace for n records with keys < key_range (0 on error)
int64_t scatter_bytes(int64_t n, int64_t key_range);
// carve a caller workspace for n records with keys < key_range; fails if it
// is too small
int scatter_layout(void* work, int64_t work_bytes, int64_t n, int64_t key_range, ScatterWork& w);
// zero the per-key counters (before the producer kernel)
int scatter_begin(const ScatterWork& w, hipStream_t st);
// group the n records by key (after the producer kernel); with split > 0,
// also list the keys with records: keys_u (at most split records) and
// keys_l (more)
int scatter_group(const ScatterWork& w, int64_t n, hipStream_t st, uint32_t split = 0);
// group the n records by key and add every group into its target
int scatter_reduce(const ScatterWork& w, int64_t n, const ScatterTargets& t, hipStream_t st);

__device__ inline void scatter_count(const ScatterWork& w, uint32_t key) {
  __hip_atomic_fetch_add(w.count + key, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void scatter_put(const ScatterWork& w, int64_t slot, uint32_t key, d3 a, d3 cb, int cls,
                                   int sub = 0) {
  w.keys[slot] = key;
  scatter_count(w, key);
  ScatterRec r;
  r.a[0] = a.x;
  r.a[1] = a.y;
  r.a[2] = a.z;
  r.cb[0] = cb.x;
  r.cb[1] = cb.y;
  r.cb[2] = cb.z;
  r.cls = cls;
  r.sub = sub;
  r.pad[0] = r.pad[1] = 0;
  w.rec[slot] = r;
}

__device__ inline void scatter_put_values(const ScatterWork& w, int64_t slot, uint32_t key, const double v[18],
                                          int cls, int sub = 0) {
  w.keys[slot] = key;
  scatter_count(w, key);
  ScatterRecV r;
#pragma unroll
  for (int e = 0; e < 18; ++e) r.v[e] = v[e];
  r.cls = cls;
  r.sub = sub;
  r.pad[0] = r.pad[1] = 0;
  w.recv[slot] = r;
}

__device__ inline void scatter_empty(const ScatterWork& w, int64_t slot) { w.keys[slot] = w.empty; }

}  // namespace cmm
