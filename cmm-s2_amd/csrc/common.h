// Internal helpers shared by the HIP translation units of libcmm_s2_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/cmm_s2.h"

namespace cmm {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);

#define CMM_HIP(call)                                         \
  do {                                                        \
    hipError_t _e = (call);                                   \
    if (_e != hipSuccess) return cmm::hip_check(_e, #call);   \
  } while (0)

#define CMM_LAUNCH_CHECK(what)                                \
  do {                                                        \
    hipError_t _e = hipGetLastError();                        \
    if (_e != hipSuccess) return cmm::hip_check(_e, what);    \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return (int)(g < 1 ? 1 : g);
}

// ---------------------------------------------------------------------------
// small vector helpers (device + host)
// ---------------------------------------------------------------------------
struct d3 {
  double x, y, z;
};

__host__ __device__ inline d3 mk(double x, double y, double z) { return d3{x, y, z}; }
__host__ __device__ inline d3 add(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline d3 sub(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline d3 scl(double s, d3 a) { return d3{s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// utils.cross (utils.py:103-104)
__host__ __device__ inline d3 cross(d3 a, d3 b) {
  return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline double norm3(d3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// utils.det_vec (utils.py:49-56), columns a, b, c
__host__ __device__ inline double det3(d3 a, d3 b, d3 c) {
  return a.x * b.y * c.z + b.x * c.y * a.z + a.y * b.z * c.x -
         (c.x * b.y * a.z + c.y * b.z * a.x + b.x * a.y * c.z);
}

// utils.bary_coords (utils.py:58-67)
__host__ __device__ inline d3 bary(d3 v1, d3 v2, d3 v3, d3 q) {
  double den = det3(v1, v2, v3);
  return d3{det3(q, v2, v3) / den, det3(v1, q, v3) / den, det3(v1, v2, q) / den};
}

// same with one reciprocal (hot paths; differs from bary by rounding)
__host__ __device__ inline d3 bary_fast(d3 v1, d3 v2, d3 v3, d3 q) {
  double r = 1.0 / det3(v1, v2, v3);
  return d3{det3(q, v2, v3) * r, det3(v1, q, v3) * r, det3(v1, v2, q) * r};
}

// utils.div_norm on one point: x / |x|
__host__ __device__ inline d3 unit(d3 a) {
  double n = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return d3{a.x / n, a.y / n, a.z / n};
}

// mesh_functions.sphere_tan_proj (mesh_functions.py:320-329)
__host__ __device__ inline d3 sphere_tan_proj(d3 a, d3 b) {
  double c = dot(b, a);
  double d = sqrt(dot(a, a) - 2 * c * c + dot(b, b) * c * c);
  return d3{(a.x - c * b.x) / d, (a.y - c * b.y) / d, (a.z - c * b.z) / d};
}

// Split sub-triangle index from (argmin, argmax) of the barycentrics with
// first-occurrence tie breaking (spherical_spline.py:691-697).
__host__ __device__ inline int split_index(double b0, double b1, double b2) {
  int mn = 0, mx = 0;
  double vmn = b0, vmx = b0;
  if (b1 < vmn) { vmn = b1; mn = 1; }
  if (b2 < vmn) { vmn = b2; mn = 2; }
  if (b1 > vmx) { vmx = b1; mx = 1; }
  if (b2 > vmx) { vmx = b2; mx = 2; }
  if (mn == mx) { mx = (mn + 1) % 3; }  // all equal: any sub-triangle (the reference yields NaN)
  return 2 * mn + (mx - (mx > mn ? 1 : 0));
}

// Coefficient indices of the 6 split sub-triangles, by split_index order
// (argmin, argmax) = (0,1),(0,2),(1,0),(1,2),(2,0),(2,1): rows of the Cs table
// of spherical_spline.py:660-662.
__host__ __device__ constexpr int cs_table(int s, int j) {
  constexpr int T[6][6] = {{18, 1, 13, 7, 6, 16}, {13, 2, 18, 11, 10, 16}, {18, 14, 0, 17, 5, 4},
                           {14, 18, 2, 17, 10, 9}, {0, 12, 18, 3, 15, 4},  {1, 18, 12, 7, 15, 8}};
  return T[s][j];
}

// row s of the Cs table as six 5-bit fields (entry j at bit 5 j), selected
// without an indexed local array
__host__ __device__ constexpr uint32_t cs_row_bits(int s) {
  uint32_t c = 0;
  for (int j = 0; j < 6; ++j) c |= (uint32_t)cs_table(s, j) << (5 * j);
  return c;
}
__host__ __device__ inline uint32_t cs_code(int s) {
  constexpr uint32_t c0 = cs_row_bits(0), c1 = cs_row_bits(1), c2 = cs_row_bits(2), c3 = cs_row_bits(3),
                     c4 = cs_row_bits(4), c5 = cs_row_bits(5);
  return s == 0 ? c0 : s == 1 ? c1 : s == 2 ? c2 : s == 3 ? c3 : s == 4 ? c4 : c5;
}

// Barycentrics of q in split sub-triangle s from the macro barycentrics b and
// the norms s12 = |v1+v2|, s23, s31, s4 = |v1+v2+v3|.  The sub-triangle
// vertices are edges_ps (spherical_spline.py:664-666) over
// [v1,v2,v3,v4,e1,e2,e3] with e1 = (v1+v2)/s12 ..., v4 = (v1+v2+v3)/s4; the
// formulas solve q = sum c_i P_i exactly as Cramer's rule on those vertices
// does (new_vpoints + bary_coords), by back substitution.
__host__ __device__ inline d3 sub_bary(int s, double b0, double b1, double b2, double s12,
                                       double s23, double s31, double s4) {
  switch (s) {
    case 0: return d3{s4 * b0, b1 - b2, s23 * (b2 - b0)};          // (v4, v2, e2)
    case 1: return d3{s23 * (b1 - b0), b2 - b1, s4 * b0};          // (e2, v3, v4)
    case 2: return d3{s4 * b1, s31 * (b2 - b1), b0 - b2};          // (v4, e3, v1)
    case 3: return d3{s31 * (b0 - b1), s4 * b1, b2 - b0};          // (e3, v4, v3)
    case 4: return d3{b0 - b1, s12 * (b1 - b2), s4 * b2};          // (v1, e1, v4)
    default: return d3{b1 - b0, s4 * b2, s12 * (b0 - b2)};         // (v2, v4, e1)
  }
}

// 6-term quadratic Bernstein-Bezier sum (spherical_spline.py:196-201).
__host__ __device__ inline double bb6(d3 c, const double* p) {
  return (c.x * c.x) * p[0] + (c.y * c.y) * p[1] + (c.z * c.z) * p[2] + 2 * (c.y * c.x) * p[3] +
         2 * (c.z * c.y) * p[4] + 2 * (c.z * c.x) * p[5];
}

// gradient of bb6 with respect to the three barycentrics (grad_HBB,
// spherical_spline.py:700-711)
__host__ __device__ inline d3 bb6_db(d3 c, const double* p) {
  return d3{2 * (p[0] * c.x + p[3] * c.y + p[5] * c.z), 2 * (p[1] * c.y + p[3] * c.x + p[4] * c.z),
            2 * (p[2] * c.z + p[4] * c.y + p[5] * c.x)};
}

// Jacobian of x -> p(x) / |p(x)| restricted to the tangent plane at q:
// J = (I / r - X X^T / r^3) G (I - q q^T), G[c] the ambient gradient of
// component c, X the raw value, r = |X| (the normalisation of
// spherical_spline.py:250-263, with the input-side projection that makes it
// the differential T_q S^2 -> T_X S^2).  Row-major J[c][d].
__host__ __device__ inline void normalized_jacobian(d3 X, const d3 G[3], d3 q, double J[9]) {
  const double r2 = X.x * X.x + X.y * X.y + X.z * X.z, r = sqrt(r2), ir = 1 / r, ir3 = ir / r2;
  const double Xv[3] = {X.x, X.y, X.z}, qv[3] = {q.x, q.y, q.z};
  double A[9];
  for (int d = 0; d < 3; ++d) {
    const double gd[3] = {d == 0 ? G[0].x : (d == 1 ? G[0].y : G[0].z),
                          d == 0 ? G[1].x : (d == 1 ? G[1].y : G[1].z),
                          d == 0 ? G[2].x : (d == 1 ? G[2].y : G[2].z)};
    const double xg = Xv[0] * gd[0] + Xv[1] * gd[1] + Xv[2] * gd[2];
    for (int c = 0; c < 3; ++c) A[3 * c + d] = gd[c] * ir - Xv[c] * xg * ir3;
  }
  for (int c = 0; c < 3; ++c) {
    const double aq = A[3 * c] * qv[0] + A[3 * c + 1] * qv[1] + A[3 * c + 2] * qv[2];
    for (int d = 0; d < 3; ++d) J[3 * c + d] = A[3 * c + d] - aq * qv[d];
  }
}

// 19 Bernstein-Bezier coefficients of one component on one macro triangle
// (spherical_spline.py:300-337) from the vertex values f[i] and vertex
// gradients gr[i] (i = 0..2) with the Powell-Sabin geometry recomputed from
// the vertices (mesh_functions.py:92-126).
struct PSGeom {
  d3 H[9];        // h12,h21,h23,h32,h13,h31,h41,h42,h43
  double G[10][3];// g12,g21,g23,g32,g13,g31,g14,g24,g34,mid
  double E[3][3]; // bary of e1,e2,e3
  double Gi[9];   // 1/G of the divisor entry of coefficients c3..c11 (the
                  // reference also forms 1/g first and multiplies)
};

// (G row, column) of the divisor and H index of coefficients c3..c11
__host__ __device__ inline int ps_div_row(int k) {
  const int r[9] = {0, 6, 4, 2, 7, 1, 5, 8, 3};
  return r[k];
}
__host__ __device__ inline int ps_div_col(int k) {
  const int c[9] = {1, 1, 2, 1, 1, 2, 1, 1, 2};
  return c[k];
}

__host__ __device__ inline void ps_geometry(d3 v1, d3 v2, d3 v3, PSGeom& o) {
  // the reference normalises (v1+v2+v3)/3 and v/2 + w/2
  d3 s3 = add(add(v1, v2), v3);
  d3 v4 = unit(d3{s3.x / 3, s3.y / 3, s3.z / 3});
  d3 e1 = unit(d3{v1.x / 2 + v2.x / 2, v1.y / 2 + v2.y / 2, v1.z / 2 + v2.z / 2});
  d3 e2 = unit(d3{v2.x / 2 + v3.x / 2, v2.y / 2 + v3.y / 2, v2.z / 2 + v3.z / 2});
  d3 e3 = unit(d3{v3.x / 2 + v1.x / 2, v3.y / 2 + v1.y / 2, v3.z / 2 + v1.z / 2});
  o.H[0] = sphere_tan_proj(sub(v2, v1), v1);
  o.H[1] = sphere_tan_proj(sub(v1, v2), v2);
  o.H[2] = sphere_tan_proj(sub(v3, v2), v2);
  o.H[3] = sphere_tan_proj(sub(v2, v3), v3);
  o.H[4] = sphere_tan_proj(sub(v3, v1), v1);
  o.H[5] = sphere_tan_proj(sub(v1, v3), v3);
  o.H[6] = sphere_tan_proj(sub(v4, v1), v1);
  o.H[7] = sphere_tan_proj(sub(v4, v2), v2);
  o.H[8] = sphere_tan_proj(sub(v4, v3), v3);
  d3 g;
  g = bary(v1, e1, v4, o.H[0]); o.G[0][0] = g.x; o.G[0][1] = g.y; o.G[0][2] = g.z;
  g = bary(v2, v4, e1, o.H[1]); o.G[1][0] = g.x; o.G[1][1] = g.y; o.G[1][2] = g.z;
  g = bary(v2, e2, v4, o.H[2]); o.G[2][0] = g.x; o.G[2][1] = g.y; o.G[2][2] = g.z;
  g = bary(v3, v4, e2, o.H[3]); o.G[3][0] = g.x; o.G[3][1] = g.y; o.G[3][2] = g.z;
  g = bary(v1, v4, e3, o.H[4]); o.G[4][0] = g.x; o.G[4][1] = g.y; o.G[4][2] = g.z;
  g = bary(v3, e3, v4, o.H[5]); o.G[5][0] = g.x; o.G[5][1] = g.y; o.G[5][2] = g.z;
  g = bary(v1, v4, e3, sphere_tan_proj(v4, v1)); o.G[6][0] = g.x; o.G[6][1] = g.y; o.G[6][2] = g.z;
  g = bary(v2, v4, e1, sphere_tan_proj(v4, v2)); o.G[7][0] = g.x; o.G[7][1] = g.y; o.G[7][2] = g.z;
  g = bary(v3, v4, e2, sphere_tan_proj(v4, v3)); o.G[8][0] = g.x; o.G[8][1] = g.y; o.G[8][2] = g.z;
  g = bary(v1, v2, v3, v4); o.G[9][0] = g.x; o.G[9][1] = g.y; o.G[9][2] = g.z;
  g = bary(v1, v2, v3, e1); o.E[0][0] = g.x; o.E[0][1] = g.y; o.E[0][2] = g.z;
  g = bary(v1, v2, v3, e2); o.E[1][0] = g.x; o.E[1][1] = g.y; o.E[1][2] = g.z;
  g = bary(v1, v2, v3, e3); o.E[2][0] = g.x; o.E[2][1] = g.y; o.E[2][2] = g.z;
  for (int k = 0; k < 9; ++k) o.Gi[k] = 1 / o.G[ps_div_row(k)][ps_div_col(k)];
}

// PSGeom as 66 doubles (H 27, G 30, E 9) at p[i * stride]
constexpr int kPSGeomLen = 75;
__host__ __device__ inline void ps_pack(const PSGeom& P, double* p, int64_t stride) {
  int k = 0;
  for (int i = 0; i < 9; ++i) { p[(k++) * stride] = P.H[i].x; p[(k++) * stride] = P.H[i].y; p[(k++) * stride] = P.H[i].z; }
  for (int i = 0; i < 10; ++i) for (int j = 0; j < 3; ++j) p[(k++) * stride] = P.G[i][j];
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) p[(k++) * stride] = P.E[i][j];
  for (int i = 0; i < 9; ++i) p[(k++) * stride] = P.Gi[i];
}
__host__ __device__ __forceinline__ void ps_unpack(const double* __restrict__ p, int64_t stride, PSGeom& P) {
  int k = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    double x = p[(k++) * stride], y = p[(k++) * stride], z = p[(k++) * stride];
    P.H[i] = d3{x, y, z};
  }
#pragma unroll
  for (int i = 0; i < 10; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) P.G[i][j] = p[(k++) * stride];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) P.E[i][j] = p[(k++) * stride];
#pragma unroll
  for (int i = 0; i < 9; ++i) P.Gi[i] = p[(k++) * stride];
}

// Write the packed evaluation layout (Ntri, 6, NC, 6) of a block of up to
// 256 triangles through LDS in rounds of 64 triangles so that the global
// stores are contiguous.  cf[c][k] are this thread's coefficients; lds and
// packed are 16-byte aligned.
template <int NC>
__device__ inline void store_packed_block(const double (&cf)[NC][19], int64_t t0, int64_t ntri,
                                          double* __restrict__ packed, double* lds) {
  constexpr int PER = 6 * NC * 6;
  const int tid = threadIdx.x;
  const int64_t t = t0 + tid;
  for (int r = 0; r < 4; ++r) {
    if ((tid >> 6) == r && t < ntri) {
      double* dst = lds + (tid & 63) * PER;
#pragma unroll
      for (int s = 0; s < 6; ++s)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int j = 0; j < 6; ++j) dst[(s * NC + c) * 6 + j] = cf[c][cs_table(s, j)];
    }
    __syncthreads();
    const int64_t tr = t0 + 64 * r;
    if (tr < ntri) {
      // 16-byte non-temporal stores: the block is written once and read by
      // a later kernel, keeping it out of the caches costs nothing
      typedef double dbl2 __attribute__((ext_vector_type(2)));
      const int nvalid = (int)((ntri - tr) < 64 ? (ntri - tr) : 64);
      dbl2* out = reinterpret_cast<dbl2*>(packed + tr * PER);
      const dbl2* src = reinterpret_cast<const dbl2*>(lds);
      for (int e = tid; e < nvalid * (PER / 2); e += blockDim.x) __builtin_nontemporal_store(src[e], out + e);
    }
    __syncthreads();
  }
}

__host__ __device__ inline void ps_coeffs(const PSGeom& P, const double f[3], const d3 gr[3],
                                          double c[19]) {
  c[0] = f[0];
  c[1] = f[1];
  c[2] = f[2];
  // c_{3+k} = (1/g)(h . grad f(v_i) / 2 - g_0 f(v_i)) with g = G[row], h = H[row]
  // and i = k / 3 (spherical_spline.py:314-324)
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int r = ps_div_row(k), i = k / 3;
    c[3 + k] = P.Gi[k] * (dot(P.H[r], gr[i]) / 2 - P.G[r][0] * f[i]);
  }
  c[12] = P.E[0][0] * c[3] + P.E[0][1] * c[8];
  c[13] = P.E[1][1] * c[6] + P.E[1][2] * c[11];
  c[14] = P.E[2][2] * c[9] + P.E[2][0] * c[5];
  c[15] = P.E[0][0] * c[4] + P.E[0][1] * c[7];
  c[16] = P.E[1][1] * c[7] + P.E[1][2] * c[10];
  c[17] = P.E[2][0] * c[4] + P.E[2][2] * c[10];
  c[18] = P.G[9][0] * c[4] + P.G[9][1] * c[7] + P.G[9][2] * c[10];
}

// Transpose of ps_coeffs (a linear map of f, gr for fixed geometry): the
// adjoints cb[19] of the coefficients -> adjoints added into fb[3], gb[3].
__host__ __device__ __forceinline__ void ps_coeffs_adj(const PSGeom& P, const double cb_in[19], double fb[3], d3 gb[3]) {
  double c[19];
  for (int k = 0; k < 19; ++k) c[k] = cb_in[k];
  c[3] += P.E[0][0] * c[12];
  c[8] += P.E[0][1] * c[12];
  c[6] += P.E[1][1] * c[13];
  c[11] += P.E[1][2] * c[13];
  c[9] += P.E[2][2] * c[14];
  c[5] += P.E[2][0] * c[14];
  c[4] += P.E[0][0] * c[15] + P.E[2][0] * c[17] + P.G[9][0] * c[18];
  c[7] += P.E[0][1] * c[15] + P.E[1][1] * c[16] + P.G[9][1] * c[18];
  c[10] += P.E[1][2] * c[16] + P.E[2][2] * c[17] + P.G[9][2] * c[18];
  for (int i = 0; i < 3; ++i) fb[i] += c[i];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int r = ps_div_row(k), i = k / 3;
    const double t = P.Gi[k] * c[3 + k];
    gb[i] = add(gb[i], scl(t / 2, P.H[r]));
    fb[i] -= t * P.G[r][0];
  }
}

// the 6 quadratic Bernstein-Bezier basis values of bb6
__host__ __device__ inline void bb6_basis(d3 c, double b[6]) {
  b[0] = c.x * c.x;
  b[1] = c.y * c.y;
  b[2] = c.z * c.z;
  b[3] = 2 * (c.y * c.x);
  b[4] = 2 * (c.z * c.y);
  b[5] = 2 * (c.z * c.x);
}

// transpose of the (linear) sub-triangle map b -> sub_bary(s, b, ...)
__host__ __device__ inline d3 sub_bary_T(int s, d3 cbar, double s12, double s23, double s31, double s4) {
  const d3 c0 = sub_bary(s, 1.0, 0.0, 0.0, s12, s23, s31, s4);
  const d3 c1 = sub_bary(s, 0.0, 1.0, 0.0, s12, s23, s31, s4);
  const d3 c2 = sub_bary(s, 0.0, 0.0, 1.0, s12, s23, s31, s4);
  return d3{dot(cbar, c0), dot(cbar, c1), dot(cbar, c2)};
}

// adjoint of x -> x / |x| at x (raw vector r): (I - u u^T) ybar / |r|
__host__ __device__ inline d3 unit_adj(d3 r, d3 ybar) {
  const double n = sqrt(r.x * r.x + r.y * r.y + r.z * r.z), in = 1 / n;
  const d3 u = d3{r.x * in, r.y * in, r.z * in};
  const double p = dot(u, ybar);
  return d3{(ybar.x - p * u.x) * in, (ybar.y - p * u.y) * in, (ybar.z - p * u.z) * in};
}

// numpy's floor division of doubles (npy_divmod): NOT floor(a / b).
__host__ __device__ inline double npy_floordiv(double a, double b) {
  double mod = fmod(a, b);
  double div = (a - mod) / b;
  if (mod != 0.0) {
    if ((b < 0) != (mod < 0)) {
      div -= 1.0;
    }
  }
  double fl;
  if (div != 0.0) {
    fl = floor(div);
    if (div - fl > 0.5) fl += 1.0;
  } else {
    fl = copysign(0.0, a / b);
  }
  return fl;
}

// numpy remainder of doubles (sign of the divisor).
__host__ __device__ inline double npy_mod(double a, double b) {
  double mod = fmod(a, b);
  if (mod != 0.0) {
    if ((b < 0) != (mod < 0)) mod += b;
  } else {
    mod = copysign(0.0, b);
  }
  return mod;
}

// strided point access
__host__ __device__ inline d3 load_pt(const double* q, int64_t i, int64_t sp, int64_t sc) {
  const double* p = q + i * sp;
  return d3{p[0], p[sc], p[2 * sc]};
}
__host__ __device__ inline void store_pt(double* q, int64_t i, int64_t sp, int64_t sc, d3 v) {
  double* p = q + i * sp;
  p[0] = v.x;
  p[sc] = v.y;
  p[2 * sc] = v.z;
}

}  // namespace cmm
