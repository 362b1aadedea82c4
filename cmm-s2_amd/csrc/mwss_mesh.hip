// MWSS structured velocity mesh: closed-form location with numpy semantics,
// Powell-Sabin assembly / multi-level evaluation, the backward
// characteristic integrators (fused RK4 over the time-interpolated velocity
// spline), project_stream's grid post-processing and vorticity sampling.
//
// The mesh is implicit: vertex (ring r, column c) = (sin t_r cos p_c,
// sin t_r sin p_c, cos t_r) from the four small sine / cosine tables, so
// location and geometry cost no HBM traffic; only the jets (assembly) and
// the packed coefficients (evaluation, 6 per component per level, one
// 144-byte block) are read from HBM.
#include <cmath>
#include <vector>

#include "batch.h"
#include "common.h"
#include "scatter.h"

using namespace cmm;

struct cmm_mwss_mesh {
  int L = 0, N = 0, M = 0;  // N = 2L longitudes, M = L+1 colatitudes
  int64_t Nu = 0, Tu = 0;
  double dphi = 0, dthe = 0, phi0 = 0, the0 = 0;
  double* theta = nullptr;   // (M)
  double* st = nullptr;      // sin theta (M)
  double* ct = nullptr;      // cos theta (M)
  double* sp = nullptr;      // sin phi (N)
  double* cp = nullptr;      // cos phi (N)
  double* rowgeom = nullptr; // (L, 2, kPSGeomLen) Powell-Sabin geometry of column 0 per (row, half)
};

extern "C" int cmm_mwss_mesh_create(int L, const double* thetas, const double* phis,
                                    const double* sin_t, const double* cos_t, const double* sin_p,
                                    const double* cos_p, cmm_mwss_mesh** out) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_mwss_mesh_create: no batched form");
  if (L < 2 || !thetas || !phis || !sin_t || !cos_t || !sin_p || !cos_p || !out)
    return fail(CMM_E_ARG, "cmm_mwss_mesh_create: bad arguments");
  if (L > 8192) return fail(CMM_E_RANGE, "cmm_mwss_mesh_create: L > 8192");
  cmm_mwss_mesh* m = new cmm_mwss_mesh();
  m->L = L; m->N = 2 * L; m->M = L + 1;
  m->Nu = (int64_t)2 * L * (L - 1) + 2;
  m->Tu = (int64_t)4 * L * (L - 1);
  m->dphi = fabs(phis[1] - phis[0]);
  m->dthe = fabs(thetas[1] - thetas[0]);
  m->phi0 = phis[0];
  m->the0 = thetas[0];
  hipError_t e;
  size_t bm = (size_t)m->M * 8, bn = (size_t)m->N * 8;
  if ((e = hipMalloc(&m->theta, bm)) != hipSuccess || (e = hipMalloc(&m->st, bm)) != hipSuccess ||
      (e = hipMalloc(&m->ct, bm)) != hipSuccess || (e = hipMalloc(&m->sp, bn)) != hipSuccess ||
      (e = hipMalloc(&m->cp, bn)) != hipSuccess) {
    cmm_mwss_mesh_destroy(m);
    return hip_check(e, "cmm_mwss_mesh_create: hipMalloc");
  }
  if ((e = hipMemcpy(m->theta, thetas, bm, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->st, sin_t, bm, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->ct, cos_t, bm, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->sp, sin_p, bn, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->cp, cos_p, bn, hipMemcpyHostToDevice)) != hipSuccess) {
    cmm_mwss_mesh_destroy(m);
    return hip_check(e, "cmm_mwss_mesh_create: hipMemcpy");
  }
  // Powell-Sabin geometry of the column-0 triangle of every (row, half): the
  // mesh is invariant under rotation about z by phi_c, so the assembly
  // rotates the tangent vectors H by phi_c and reuses G and E as they are
  {
    auto V = [&](int r, int c) { return d3{sin_t[r] * cos_p[c], sin_t[r] * sin_p[c], cos_t[r]}; };
    std::vector<double> tab((size_t)L * 2 * kPSGeomLen);
    for (int r = 0; r < L; ++r)
      for (int h = 0; h < 2; ++h) {
        d3 a, b, c;
        if (r == 0) { a = V(0, 0); b = V(1, 0); c = V(1, 1); }
        else if (r == L - 1) { a = V(L - 1, 0); b = V(L, 0); c = V(L - 1, 1); }
        else if (h == 0) { a = V(r, 0); b = V(r + 1, 1); c = V(r, 1); }
        else { a = V(r, 0); b = V(r + 1, 0); c = V(r + 1, 1); }
        PSGeom P;
        ps_geometry(a, b, c, P);
        ps_pack(P, &tab[((size_t)r * 2 + h) * kPSGeomLen], 1);
      }
    if ((e = hipMalloc(&m->rowgeom, tab.size() * 8)) != hipSuccess ||
        (e = hipMemcpy(m->rowgeom, tab.data(), tab.size() * 8, hipMemcpyHostToDevice)) != hipSuccess) {
      cmm_mwss_mesh_destroy(m);
      return hip_check(e, "cmm_mwss_mesh_create: geometry table");
    }
  }
  *out = m;
  return CMM_OK;
}

extern "C" int cmm_mwss_mesh_destroy(cmm_mwss_mesh* m) {
  if (!m) return CMM_OK;
  (void)hipFree(m->theta);
  (void)hipFree(m->st);
  (void)hipFree(m->ct);
  (void)hipFree(m->sp);
  (void)hipFree(m->cp);
  (void)hipFree(m->rowgeom);
  delete m;
  return CMM_OK;
}

// ---------------------------------------------------------------------------
// location (mesh_functions.py:196-282)
// ---------------------------------------------------------------------------
struct MLoc {
  int64_t tri;
  int64_t vi[3];
  d3 v[3];
  d3 b;
};

__device__ inline d3 ring_pt(const cmm_mwss_mesh& M, int r, int c) {
  return d3{M.st[r] * M.cp[c], M.st[r] * M.sp[c], M.ct[r]};
}

// numpy-faithful decisions: no FMA contraction in the tests that pick cells
__device__ __forceinline__ bool mwss_locate(const cmm_mwss_mesh& M, d3 q, MLoc& o) {
#pragma clang fp contract(off)
  const double TWO_PI = 2 * M_PI;
  double phi = npy_mod(atan2(q.y, q.x) + TWO_PI, TWO_PI);
  double theta = atan2(sqrt(q.y * q.y + q.x * q.x), q.z);
  if (!isfinite(phi) || !isfinite(theta)) return false;
  double phin = npy_mod(phi, TWO_PI);
  int N = M.N, Mr = M.M;
  long long plr = (long long)npy_floordiv(phin - M.phi0, M.dphi);
  int pl = (int)(((plr % N) + N) % N);
  int tl = (int)npy_floordiv(theta - M.the0, M.dthe);
  if (tl == Mr - 1) tl = Mr - 2;
  if (tl < 0 || tl > Mr - 2) return false;
  double thc = (theta - M.theta[tl]) / M.dthe;
  int pp1 = (pl + 1) % N;
  bool in0 = (thc <= 0.03) && (tl != 0);
  bool in1 = (thc > 0.85) && (tl != Mr - 2);
  if (in0 || in1) {
    // the reference re-derives the point from its angles here
    // (sphere2cart(phi, theta), mesh_functions.py:240, 249); that differs
    // from q by rounding only, which can change the side of a great circle
    // only for points on it (a location tie)
    const d3 qq = q;
    if (in0) {
      d3 n = cross(ring_pt(M, tl, pl), ring_pt(M, tl, pp1));
      double dd = n.x * qq.x + n.y * qq.y + n.z * qq.z;
      if (dd > 0) tl -= 1;
    } else {
      d3 n = cross(ring_pt(M, tl + 1, pl), ring_pt(M, tl + 1, pp1));
      double dd = n.x * qq.x + n.y * qq.y + n.z * qq.z;
      if (-dd > 0) tl += 1;
    }
  }
  d3 a = ring_pt(M, tl, pl), b = ring_pt(M, tl + 1, pp1);
  d3 n = cross(a, b);
  double dd = n.x * q.x + n.y * q.y + n.z * q.z;
  int s = (-dd > 0) ? 1 : 0;
  const int64_t Nu = M.Nu;
  auto R = [&](int r, int c) -> int64_t { return 1 + (int64_t)(r - 1) * N + c; };
  if (tl == 0) {
    o.tri = pl;
    o.vi[0] = 0; o.vi[1] = R(1, pl); o.vi[2] = R(1, pp1);
    o.v[0] = ring_pt(M, 0, 0); o.v[1] = ring_pt(M, 1, pl); o.v[2] = ring_pt(M, 1, pp1);
  } else if (tl == Mr - 2) {
    o.tri = M.Tu - N + pl;
    o.vi[0] = R(Mr - 2, pl); o.vi[1] = Nu - 1; o.vi[2] = R(Mr - 2, pp1);
    o.v[0] = ring_pt(M, Mr - 2, pl); o.v[1] = ring_pt(M, Mr - 1, 0); o.v[2] = ring_pt(M, Mr - 2, pp1);
  } else {
    o.tri = N + 2 * ((int64_t)(tl - 1) * N + pl) + s;
    if (s == 0) {
      o.vi[0] = R(tl, pl); o.vi[1] = R(tl + 1, pp1); o.vi[2] = R(tl, pp1);
      o.v[0] = ring_pt(M, tl, pl); o.v[1] = ring_pt(M, tl + 1, pp1); o.v[2] = ring_pt(M, tl, pp1);
    } else {
      o.vi[0] = R(tl, pl); o.vi[1] = R(tl + 1, pl); o.vi[2] = R(tl + 1, pp1);
      o.v[0] = ring_pt(M, tl, pl); o.v[1] = ring_pt(M, tl + 1, pl); o.v[2] = ring_pt(M, tl + 1, pp1);
    }
  }
  o.b = bary_fast(o.v[0], o.v[1], o.v[2], q);
  return true;
}

// vertices of triangle t (closed form of grid_assemble, mesh_functions.py:380-412)
__device__ inline void mwss_tri_vertices(const cmm_mwss_mesh& M, int64_t t, int64_t vi[3], d3 v[3]) {
  int N = M.N, Mr = M.M;
  auto R = [&](int r, int c) -> int64_t { return 1 + (int64_t)(r - 1) * N + c; };
  if (t < N) {
    int c = (int)t, c1 = (c + 1) % N;
    vi[0] = 0; vi[1] = R(1, c); vi[2] = R(1, c1);
    v[0] = ring_pt(M, 0, 0); v[1] = ring_pt(M, 1, c); v[2] = ring_pt(M, 1, c1);
  } else if (t >= M.Tu - N) {
    int c = (int)(t - (M.Tu - N)), c1 = (c + 1) % N;
    vi[0] = R(Mr - 2, c); vi[1] = M.Nu - 1; vi[2] = R(Mr - 2, c1);
    v[0] = ring_pt(M, Mr - 2, c); v[1] = ring_pt(M, Mr - 1, 0); v[2] = ring_pt(M, Mr - 2, c1);
  } else {
    int64_t u = t - N;
    int h = (int)(u % 2);
    int64_t k = u / 2;
    int r = 1 + (int)(k / N), c = (int)(k % N), c1 = (c + 1) % N;
    if (h == 0) {
      vi[0] = R(r, c); vi[1] = R(r + 1, c1); vi[2] = R(r, c1);
      v[0] = ring_pt(M, r, c); v[1] = ring_pt(M, r + 1, c1); v[2] = ring_pt(M, r, c1);
    } else {
      vi[0] = R(r, c); vi[1] = R(r + 1, c); vi[2] = R(r + 1, c1);
      v[0] = ring_pt(M, r, c); v[1] = ring_pt(M, r + 1, c); v[2] = ring_pt(M, r + 1, c1);
    }
  }
}

// sub-triangle barycentrics from the macro ones, norms from the vertices
__device__ inline d3 sub_bary_v(int s, d3 b, const d3 v[3]) {
  d3 s3 = add(add(v[0], v[1]), v[2]);
  double s4 = norm3(s3);
  double se;
  if (s <= 1) se = norm3(add(v[1], v[2]));
  else if (s <= 3) se = norm3(add(v[2], v[0]));
  else se = norm3(add(v[0], v[1]));
  return sub_bary(s, b.x, b.y, b.z, se, se, se, s4);
}

struct LevelSet {
  const double* p[3];
  double w[3];
  int nlev;
};

// Velocity levels assembled on a subset of triangles (cmm_mwss_assemble_sparse):
// bit slot[k] of vbits[t] is set where level k's packed blocks of triangle t
// are current; jets[k] (3, 4, Nu) is the level's source; mark (may be NULL)
// collects every triangle an evaluation locates.
struct SparseLv {
  const uint8_t* vbits;
  int slot[3];
  const double* jets[3];
  uint8_t* mark;
  int mark_every;  // 0: mark only the triangles with a missing block
};

// (row, half) of the per-row geometry table and the column whose angle
// rotates it, for MWSS triangle t (the ordering of mwss_tri_vertices)
__device__ inline void mwss_tri_rowcol(const cmm_mwss_mesh& M, int64_t t, int& row, int& half, int& col) {
  const int N = M.N;
  if (t < N) { row = 0; col = (int)t; half = 0; }
  else if (t >= M.Tu - N) { row = M.M - 2; col = (int)(t - (M.Tu - N)); half = 0; }
  else {
    const int64_t u = t - N;
    half = (int)(u % 2);
    const int64_t k = u / 2;
    row = 1 + (int)(k / N);
    col = (int)(k % N);
  }
}

// Coefficient kk < 12 of component cc of triangle t's Powell-Sabin spline
// from the jets J (3, 4, Nu): vertex values and the c_{3+k} formulas of
// ps_coeffs with the column-rotated row geometry g (ps_pack layout: H at 0,
// G at 27, E at 57, 1/G at 66), written exactly as k_mwss_assemble
// evaluates them so that both give the same bits.
__device__ inline double mwss_coeff_direct(const cmm_mwss_mesh& M, const double* __restrict__ J,
                                           const double* __restrict__ g, double c, double sn, int64_t v0,
                                           int64_t v1, int64_t v2, int cc, int kk) {
  const int64_t Nu = M.Nu;
  if (kk < 3) return J[(int64_t)(cc * 4) * Nu + (kk == 0 ? v0 : (kk == 1 ? v1 : v2))];
  const int k9 = kk - 3, i = k9 / 3;
  const int r = (int)((0x385172460ull >> (4 * k9)) & 0xF);  // ps_div_row(k9)
  const int64_t vi = i == 0 ? v0 : (i == 1 ? v1 : v2);
  const d3 h0 = d3{g[3 * r], g[3 * r + 1], g[3 * r + 2]};
  const d3 h = d3{c * h0.x - sn * h0.y, sn * h0.x + c * h0.y, h0.z};
  const double f = J[(int64_t)(cc * 4) * Nu + vi];
  const d3 gr = d3{J[(int64_t)(0 * 4 + 1 + cc) * Nu + vi], J[(int64_t)(1 * 4 + 1 + cc) * Nu + vi],
                   J[(int64_t)(2 * 4 + 1 + cc) * Nu + vi]};
  return g[66 + k9] * (dot(h, gr) / 2 - g[27 + 3 * r] * f);
}

// Coefficients 12..18 from the direct ones x[0..11] (ps_coeffs)
__device__ inline double mwss_coeff_derived(const double* __restrict__ g, const double* x, int kk) {
  switch (kk) {
    case 12: return g[57] * x[3] + g[58] * x[8];
    case 13: return g[61] * x[6] + g[62] * x[11];
    case 14: return g[65] * x[9] + g[63] * x[5];
    case 15: return g[57] * x[4] + g[58] * x[7];
    case 16: return g[61] * x[7] + g[62] * x[10];
    case 17: return g[63] * x[4] + g[65] * x[10];
    default: return g[54] * x[4] + g[55] * x[7] + g[56] * x[10];
  }
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Level K's values for the lanes of this wave whose triangle's block of that
// level is not assembled (have == false): while such a lane remains, the
// wave takes the lowest one's triangle, computes the 57 coefficients of
// level K there together (one or two per lane, through LDS), and every
// lane with that triangle evaluates its sub-triangle from them.  Each round
// serves at least the leading lane, so the loop ends after at most 64
// rounds; it is entered only when a lane misses.  Blocks of at most 256
// threads.
template <int K>
__device__ __forceinline__ void mwss_fill_missing(const cmm_mwss_mesh& M, const SparseLv& sv, int64_t tri, int s,
                                                  d3 cb, bool have, double (&bk)[3]) {
  __shared__ double fb_lds[4][64];
  double* X = fb_lds[threadIdx.x >> 6];
  uint64_t pend = __ballot(!have);
  while (pend) {
    // the leading lane's triangle as a wave-uniform (scalar) value: the
    // geometry, vertex ids and jets addresses below stay in scalar registers
    const int leader = __ffsll((unsigned long long)pend) - 1;
    const int lt = __builtin_amdgcn_readlane((int)tri, leader);
    int row, half, col;
    mwss_tri_rowcol(M, lt, row, half, col);
    int64_t vi[3];
    d3 vv[3];
    mwss_tri_vertices(M, lt, vi, vv);
    const double* g = M.rowgeom + ((int64_t)row * 2 + half) * kPSGeomLen;
    const double c = M.cp[col], sn = M.sp[col];
    const uint64_t act = __ballot(1);
    const int nact = __popcll(act);
    const int rank = __popcll(act & ((1ull << __lane_id()) - 1));
    wave_lds_sync();
    for (int j = rank; j < 36; j += nact)
      X[(j / 12) * 19 + j % 12] = mwss_coeff_direct(M, sv.jets[K], g, c, sn, vi[0], vi[1], vi[2], j / 12, j % 12);
    wave_lds_sync();
    for (int j = rank; j < 21; j += nact) X[(j / 7) * 19 + 12 + j % 7] = mwss_coeff_derived(g, X + (j / 7) * 19, 12 + j % 7);
    wave_lds_sync();
    if (!have && tri == lt) {
      const uint32_t code = cs_code(s);
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        double a[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) a[j] = X[cc * 19 + ((code >> (5 * j)) & 31u)];
        bk[cc] = bb6(cb, a);
      }
      have = true;
    }
    pend = __ballot(!have);
  }
}

// sum_k w_k * PS_k(q) for ncomp components; sv (ncomp 3 only): levels kept
// on a subset of triangles.  Every level's block is loaded together with its
// validity byte (a stale block costs nothing when it is not used), so the
// common case keeps the loads of complete levels in flight at once; a wave
// in which some lane's block is missing sums again, level by level, with the
// missing blocks built from the jets.  Levels are added in the same order
// either way, so both give the bits of complete levels.
__device__ __forceinline__ bool mwss_eval_levels(const cmm_mwss_mesh& M, const LevelSet& ls, int ncomp, d3 q,
                                        double* r, int64_t* tri_out = nullptr, const SparseLv* sv = nullptr) {
  MLoc o;
  if (!mwss_locate(M, q, o)) return false;
  if (tri_out) *tri_out = o.tri;
  int s = split_index(o.b.x, o.b.y, o.b.z);
  d3 cb = sub_bary_v(s, o.b, o.v);
  unsigned have = 7;
  if (sv) {
    const unsigned vb = sv->vbits[o.tri];
    have = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (k >= ls.nlev || ((vb >> sv->slot[k]) & 1u)) have |= 1u << k;
    if (sv->mark && sv->mark_every) sv->mark[o.tri] = 1;
  }
  // compile-time trip counts with guards: runtime-indexed arrays would be
  // placed in scratch memory
#pragma unroll
  for (int c = 0; c < 3; ++c)
    if (c < ncomp) r[c] = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (k < ls.nlev) {
      const double* p = ls.p[k] + (o.tri * 6 + s) * 6 * ncomp;
      const double w = ls.w[k];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if (c < ncomp) r[c] += w * bb6(cb, p + 6 * c);
    }
  }
  if (sv && __ballot(have != 7)) {
    if (sv->mark && have != 7) sv->mark[o.tri] = 1;
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (k < ls.nlev) {
        const bool hk = (have >> k) & 1u;
        double bk[3] = {0, 0, 0};
        if (hk) {
          const double* p = ls.p[k] + (o.tri * 6 + s) * 18;
#pragma unroll
          for (int c = 0; c < 3; ++c) bk[c] = bb6(cb, p + 6 * c);
        }
        if (k == 0) mwss_fill_missing<0>(M, *sv, o.tri, s, cb, hk, bk);
        if (k == 1) mwss_fill_missing<1>(M, *sv, o.tri, s, cb, hk, bk);
        if (k == 2) mwss_fill_missing<2>(M, *sv, o.tri, s, cb, hk, bk);
        const double w = ls.w[k];
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] += w * bk[c];
      }
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// analytic velocity fields (dynamical_fields.py:18-42) and RH-4
// ---------------------------------------------------------------------------
__device__ inline d3 rot_apply(const double R[3][3], d3 x) {
  return d3{R[0][0] * x.x + R[0][1] * x.y + R[0][2] * x.z, R[1][0] * x.x + R[1][1] * x.y + R[1][2] * x.z,
            R[2][0] * x.x + R[2][1] * x.y + R[2][2] * x.z};
}

__device__ inline void matmul3(const double A[3][3], const double B[3][3], double C[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
}

__device__ inline d3 vel_sbr(d3 X, double T, double alpha) {
  double c = -2 * M_PI / T;
  return cross(X, d3{c * sin(alpha), 0.0, c * cos(alpha)});
}

__device__ inline d3 vel_deform_rot(double t, d3 X, double alpha) {
  double tau = 2 * M_PI * t;
  double ct = cos(tau), st = sin(tau), ca = cos(alpha), sa = sin(alpha);
  double Rz[3][3] = {{ct, -st, 0}, {st, ct, 0}, {0, 0, 1}};
  double RzT[3][3] = {{ct, st, 0}, {-st, ct, 0}, {0, 0, 1}};
  double Ry[3][3] = {{ca, 0, -sa}, {0, 1, 0}, {sa, 0, ca}};
  double RyT[3][3] = {{ca, 0, sa}, {0, 1, 0}, {-sa, 0, ca}};
  double A[3][3], B[3][3];
  matmul3(RzT, Ry, A);
  matmul3(RyT, Rz, B);
  d3 Xr = rot_apply(A, X);
  double c1 = 2 * 2 * cos(M_PI * t / 1.0);
  d3 us = cross(Xr, d3{0.0, c1 * Xr.y, 0.0});
  d3 usr = rot_apply(B, us);
  return add(usr, vel_sbr(X, 1.0, 0.0));
}

__device__ inline d3 vel_rh4(double t, d3 X, double w, double K, double Om) {
  const double R = 4;
  double nu = (R * (3 + R) * w - 2 * Om) / ((1 + R) * (2 + R));
  double c = cos(4 * nu * t), s = sin(4 * nu * t);
  double x = X.x, y = X.y, z = X.z, x2 = x * x, y2 = y * y;
  double re4 = x2 * x2 - 6 * x2 * y2 + y2 * y2;
  double im4 = 4 * x2 * x * y - 4 * x * y2 * y;
  double A = c * re4 + s * im4;
  double px = K * z * (c * (4 * x2 * x - 12 * x * y2) + s * (12 * x2 * y - 4 * y2 * y));
  double py = K * z * (c * (-12 * x2 * y + 4 * y2 * y) + s * (4 * x2 * x - 12 * x * y2));
  double pz = -w + K * A;
  return d3{y * pz - z * py, z * px - x * pz, x * py - y * px};
}

// utils.cart2sphere (utils.py:18-23)
__device__ inline void cart2sphere(d3 X, double& phi, double& theta) {
  phi = npy_mod(atan2(X.y, X.x) + 2 * M_PI, 2 * M_PI);
  theta = atan2(sqrt(X.y * X.y + X.x * X.x), X.z);
}

// dynamical_fields.u_deform (dynamical_fields.py:24-30)
__device__ inline d3 vel_deform(double t, d3 X, double T) {
  const double c1 = 2 * 2 * cos(M_PI * t / T);
  return cross(X, d3{0.0, c1 * X.y, 0.0});
}

// dynamical_fields.omega_r (dynamical_fields.py:48-55), rho_0 = 3
__device__ inline double vortex_omega(double theta) {
  if (theta == 0.0) return 0.0;
  const double r = 3.0 * sin(theta), ch = cosh(r);
  return (2 * M_PI) * sqrt(3.0) * 1.5 * (1 / r) * tanh(r) * (1 / (ch * ch));
}

// dynamical_fields.u_static_vortex (dynamical_fields.py:57-66)
__device__ inline d3 vel_static_vortex(d3 X) {
  double phi, theta;
  cart2sphere(X, phi, theta);
  const double w = vortex_omega(theta);
  return d3{-X.y * w, X.x * w, 0.0};
}

// dynamical_fields.u_rotating_vortex (dynamical_fields.py:68-78): the static
// vortex in the frame Rot_x(pi/2) Rot_z(tau)^T, tau = 2 pi t / T, plus the
// solid-body rotation
__device__ inline d3 vel_rotating_vortex(double t, d3 X, double T) {
  const double tau = 2 * M_PI * t / T, ct = cos(tau), st = sin(tau);
  const double c = cos(M_PI / 2), s = sin(M_PI / 2);
  const double Rx[3][3] = {{1, 0, 0}, {0, c, -s}, {0, s, c}};
  const double RzT[3][3] = {{ct, st, 0}, {-st, ct, 0}, {0, 0, 1}};
  const double Rz[3][3] = {{ct, -st, 0}, {st, ct, 0}, {0, 0, 1}};
  const double RxT[3][3] = {{1, 0, 0}, {0, c, s}, {0, -s, c}};
  double A[3][3], Bm[3][3];
  matmul3(Rx, RzT, A);
  matmul3(Rz, RxT, Bm);
  const d3 up = vel_static_vortex(rot_apply(A, X));
  return add(rot_apply(Bm, up), vel_sbr(X, 1.0, 0.0));
}

// dynamical_fields.equator_vortex (dynamical_fields.py:80-85)
__device__ inline d3 vel_equator_vortex(d3 X) {
  const double c = cos(M_PI / 2), s = sin(M_PI / 2);
  const double Rx[3][3] = {{1, 0, 0}, {0, c, -s}, {0, s, c}};
  const double RxT[3][3] = {{1, 0, 0}, {0, c, s}, {0, -s, c}};
  return rot_apply(RxT, vel_static_vortex(rot_apply(Rx, X)));
}

// dynamical_fields.u_div (dynamical_fields.py:99-108)
__device__ inline d3 vel_div(double t, d3 X, double T) {
  double phi, theta;
  cart2sphere(X, phi, theta);
  const double sp2 = sin(phi / 2), st = sin(theta), ct = cos(theta), cpt = cos(M_PI * t / T);
  const double u = -(sp2 * sp2) * sin(2 * theta) * st * cpt;
  const double v = (-0.5) * sin(phi) * (st * st * st) * cpt;
  const double a0 = ct * cos(phi) * v - st * sin(phi) * u;
  const double a1 = ct * sin(phi) * v + st * cos(phi) * u;
  const double a2 = -st * v;
  // utils.tan_proj (utils.py:29-34)
  return d3{(1 - X.x * X.x) * a0 + (-X.x * X.y) * a1 + (-X.x * X.z) * a2,
            (-X.x * X.y) * a0 + (1 - X.y * X.y) * a1 + (-X.y * X.z) * a2,
            (-X.x * X.z) * a0 + (-X.y * X.z) * a1 + (1 - X.z * X.z) * a2};
}

struct VelDev {
  int kind;
  cmm_mwss_mesh mesh;
  LevelSet ls[4];
  double stage_t[4];
  double params[8];
  SparseLv sv;  // sv.vbits == NULL: every level complete
};

// VK: the kernels are instantiated for the analytic fields (kVelAnalytic),
// the spline velocity with complete levels (kVelSpline: the CMM loop spends
// no registers on the analytic fields or on missing blocks) and with levels
// kept on a subset of triangles (kVelSparse)
constexpr int kVelAnalytic = 0, kVelSpline = 1, kVelSparse = 2;

template <int VK>
__device__ __forceinline__ d3 vel_eval(const VelDev& V, int stage, d3 x) {
  if constexpr (VK != kVelAnalytic) {
    double r[3];
    if (!mwss_eval_levels(V.mesh, V.ls[stage], 3, x, r, nullptr, VK == kVelSparse ? &V.sv : nullptr)) {
      double nan = __builtin_nan("");
      return d3{nan, nan, nan};
    }
    return d3{r[0], r[1], r[2]};
  }
  switch (V.kind) {
    case CMM_VEL_SBR: return vel_sbr(x, V.params[0], V.params[1]);
    case CMM_VEL_DEFORM_ROT: return vel_deform_rot(V.stage_t[stage], x, V.params[1]);
    case CMM_VEL_RH4: return vel_rh4(V.stage_t[stage], x, V.params[0], V.params[1], V.params[2]);
    case CMM_VEL_DEFORM: return vel_deform(V.stage_t[stage], x, V.params[0]);
    case CMM_VEL_STATIC_VORTEX: return vel_static_vortex(x);
    case CMM_VEL_ROTATING_VORTEX: return vel_rotating_vortex(V.stage_t[stage], x, V.params[0]);
    case CMM_VEL_EQUATOR_VORTEX: return vel_equator_vortex(x);
    default: return vel_div(V.stage_t[stage], x, V.params[0]);
  }
}

// spline velocity at x with the located triangle (-1: not located), for an
// adjoint that revisits x
__device__ __forceinline__ d3 vel_eval_t(const VelDev& V, int stage, d3 x, int64_t& tri) {
  double r[3];
  tri = -1;
  if (!mwss_eval_levels(V.mesh, V.ls[stage], 3, x, r, &tri)) {
    const double nan = __builtin_nan("");
    tri = -1;
    return d3{nan, nan, nan};
  }
  return d3{r[0], r[1], r[2]};
}

// evolution_functions.div_norm (82-85): x * (1 / |x|)
__device__ inline d3 proj_norm(d3 x) {
  double inv = 1 / sqrt(x.x * x.x + x.y * x.y + x.z * x.z);
  return d3{inv * x.x, inv * x.y, inv * x.z};
}

// one backward step of integrator `integ` for the point xn
template <int VK>
__device__ __forceinline__ d3 advect_point(const VelDev& V, int integ, double dt, d3 xn) {
  auto vel_eval = [&](const VelDev& W, int stage, d3 x) { return ::vel_eval<VK>(W, stage, x); };
  if (integ == CMM_INT_EVAL) return vel_eval(V, 0, xn);
  if (integ == CMM_INT_EULER) {
    d3 v1 = vel_eval(V, 0, xn);
    return proj_norm(sub(xn, scl(dt, v1)));
  }
  if (integ == CMM_INT_IMPROVED_EULER) {
    d3 v1 = vel_eval(V, 0, xn);
    d3 s1 = proj_norm(sub(xn, scl(dt, v1)));
    d3 v2 = vel_eval(V, 1, s1);
    return proj_norm(sub(xn, scl(dt / 2, add(v1, v2))));
  }
  if (integ == CMM_INT_RK2) {
    d3 v1 = vel_eval(V, 0, xn);
    d3 s1 = proj_norm(sub(xn, scl(dt / 2, v1)));
    d3 v2 = vel_eval(V, 1, s1);
    return proj_norm(sub(xn, scl(dt, v2)));
  }
  d3 v1 = vel_eval(V, 0, xn);
  d3 s1 = proj_norm(d3{xn.x - dt * v1.x / 2, xn.y - dt * v1.y / 2, xn.z - dt * v1.z / 2});
  d3 v2 = vel_eval(V, 1, s1);
  d3 s2 = proj_norm(sub(xn, scl(dt / 2, v2)));
  d3 v3 = vel_eval(V, 2, s2);
  d3 s3 = proj_norm(sub(xn, scl(dt, v3)));
  d3 v4 = vel_eval(V, 3, s3);
  d3 acc = d3{v1.x + 2 * v2.x + 2 * v3.x + v4.x, v1.y + 2 * v2.y + 2 * v3.y + v4.y,
              v1.z + 2 * v2.z + 2 * v3.z + v4.z};
  return proj_norm(sub(xn, scl(dt / 6, acc)));
}

template <int VK>
__global__ __launch_bounds__(128) void k_advect(VelDev V, int integ, double dt, const double* __restrict__ x, int64_t m,
                                                double* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const d3 out = advect_point<VK>(V, integ, dt, d3{x[i], x[m + i], x[2 * m + i]});
  y[i] = out.x;
  y[m + i] = out.y;
  y[2 * m + i] = out.z;
}

// vertex feet and stencil feet of one map in one launch: point i is
// (vertex v = i / P, slot k = i % P), slot 0 the vertex itself when P = 5,
// the other slots its four stencil points, so the feet of one vertex sit in
// neighbouring lanes and their velocity gathers share cache lines
template <int P, int VK>
struct KAdvectStencil {
  struct Args {
    VelDev V;
    int integ;
    double dt;
    const double* xv;
    int64_t xv_row;
    const double* xs;
    int64_t xs_row, nv;
    double* yv;
    int64_t yv_row;
    double* ys;
    int64_t ys_row;
  };
  static constexpr int kThreads = 128;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

template <int P, int VK>
__device__ __forceinline__ void KAdvectStencil<P, VK>::run(const Args& a, int bid) {
  const VelDev& V = a.V;
  const int integ = a.integ;
  const double dt = a.dt;
  const double* __restrict__ xv = a.xv;
  const double* __restrict__ xs = a.xs;
  double* __restrict__ yv = a.yv;
  double* __restrict__ ys = a.ys;
  const int64_t xv_row = a.xv_row, xs_row = a.xs_row, nv = a.nv, yv_row = a.yv_row, ys_row = a.ys_row;
  const int64_t i = (int64_t)bid * blockDim.x + threadIdx.x;
  if (i >= P * nv) return;
  const int64_t v = i / P;
  const int k = (int)(i - v * P) - (P - 4);  // stencil index, -1 = vertex
  const double* src = k < 0 ? xv + v : xs + (int64_t)k * nv + v;
  const int64_t row = k < 0 ? xv_row : xs_row;
  const d3 out = advect_point<VK>(V, integ, dt, d3{src[0], src[row], src[2 * row]});
  double* dst = k < 0 ? yv + v : ys + (int64_t)k * nv + v;
  const int64_t orow = k < 0 ? yv_row : ys_row;
  dst[0] = out.x;
  dst[orow] = out.y;
  dst[2 * orow] = out.z;
}

template <int P, int VK>
__global__ __launch_bounds__(128) void k_advect_stencil(typename KAdvectStencil<P, VK>::Args a) {
  KAdvectStencil<P, VK>::run(a, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_mwss_locate(cmm_mwss_mesh M, const double* __restrict__ q, int64_t n, int64_t sp,
                              int64_t sc, int32_t* __restrict__ tri, double* __restrict__ bcc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  MLoc o;
  bool ok = mwss_locate(M, load_pt(q, i, sp, sc), o);
  tri[i] = ok ? (int32_t)o.tri : -1;
  if (bcc) {
    double nan = __builtin_nan("");
    bcc[3 * i] = ok ? o.b.x : nan;
    bcc[3 * i + 1] = ok ? o.b.y : nan;
    bcc[3 * i + 2] = ok ? o.b.z : nan;
  }
}

__global__ __launch_bounds__(256) void k_mwss_eval(cmm_mwss_mesh M, LevelSet ls, int ncomp, const double* __restrict__ q,
                            int64_t n, int64_t sp, int64_t sc, double* __restrict__ out, int64_t osp,
                            int64_t osc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r[3];
  if (!mwss_eval_levels(M, ls, ncomp, load_pt(q, i, sp, sc), r)) {
    for (int c = 0; c < ncomp; ++c) r[c] = __builtin_nan("");
  }
  for (int c = 0; c < ncomp; ++c) out[i * osp + c * osc] = r[c];
}

// tangential gradient of a scalar spline on the MWSS mesh
// (spherical_spline.py:73-111): out (3, n), NaN where not located
__global__ __launch_bounds__(256) void k_mwss_eval_grad(cmm_mwss_mesh M, const double* __restrict__ packed,
                                                        const double* __restrict__ q, int64_t n, int64_t sp,
                                                        int64_t sc, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const d3 x = load_pt(q, i, sp, sc);
  MLoc o;
  if (!mwss_locate(M, x, o)) {
    for (int k = 0; k < 3; ++k) out[(int64_t)k * n + i] = __builtin_nan("");
    return;
  }
  const int s = split_index(o.b.x, o.b.y, o.b.z);
  const d3 cb = sub_bary_v(s, o.b, o.v);
  // macro dual basis w_i = (v_j x v_k) / det: grad of the barycentric b_i
  const double D = det3(o.v[0], o.v[1], o.v[2]);
  const d3 w0 = scl(1 / D, cross(o.v[1], o.v[2])), w1 = scl(1 / D, cross(o.v[2], o.v[0])),
           w2 = scl(1 / D, cross(o.v[0], o.v[1]));
  const d3 gx = sub_bary_v(s, d3{w0.x, w1.x, w2.x}, o.v);
  const d3 gy = sub_bary_v(s, d3{w0.y, w1.y, w2.y}, o.v);
  const d3 gz = sub_bary_v(s, d3{w0.z, w1.z, w2.z}, o.v);
  const double* p = packed + (o.tri * 6 + s) * 6;
  const d3 db = bb6_db(cb, p);
  const d3 g = d3{dot(db, gx), dot(db, gy), dot(db, gz)};
  const double gq = dot(g, x);
  out[i] = g.x - gq * x.x;
  out[n + i] = g.y - gq * x.y;
  out[2 * n + i] = g.z - gq * x.z;
}

// the 19 coefficients per component of MWSS triangle t from the jets
// (NC, 4, Nu): the per-row geometry rotated to the triangle's column
template <int NC>
__device__ __forceinline__ void mwss_tri_coeffs(const cmm_mwss_mesh& M, const double* __restrict__ vals, int64_t t,
                                                double (&cf)[NC][19]) {
  int row, col, half;
  mwss_tri_rowcol(M, t, row, half, col);
  int64_t vi[3];
  d3 v[3];
  mwss_tri_vertices(M, t, vi, v);
  PSGeom P;
  ps_unpack(M.rowgeom + ((int64_t)row * 2 + half) * kPSGeomLen, 1, P);
  const double c = M.cp[col], sn = M.sp[col];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    d3 h = P.H[i];
    P.H[i] = d3{c * h.x - sn * h.y, sn * h.x + c * h.y, h.z};
  }
  const int64_t Nu = M.Nu;
#pragma unroll
  for (int cc = 0; cc < NC; ++cc) {
    double f[3];
    d3 gr[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (NC == 1) {
        f[i] = vals[vi[i]];
        gr[i] = d3{vals[Nu + vi[i]], vals[2 * Nu + vi[i]], vals[3 * Nu + vi[i]]};
      } else {
        f[i] = vals[(cc * 4) * Nu + vi[i]];
        gr[i] = d3{vals[(0 * 4 + 1 + cc) * Nu + vi[i]], vals[(1 * 4 + 1 + cc) * Nu + vi[i]],
                   vals[(2 * 4 + 1 + cc) * Nu + vi[i]]};
      }
    }
    ps_coeffs(P, f, gr, cf[cc]);
  }
}

template <int NC>
__global__ __launch_bounds__(256) void k_mwss_assemble(cmm_mwss_mesh M, const double* __restrict__ vals,
                                                       double* __restrict__ coeffs, double* __restrict__ packed) {
  __shared__ __align__(16) double lds[64 * 6 * NC * 6];
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x;
  const int64_t t = t0 + threadIdx.x;
  double cf[NC][19];
  if (t < M.Tu) {
    mwss_tri_coeffs<NC>(M, vals, t, cf);
    if (coeffs)
#pragma unroll
      for (int cc = 0; cc < NC; ++cc)
        for (int k = 0; k < 19; ++k) coeffs[((int64_t)cc * 19 + k) * M.Tu + t] = cf[cc][k];
  }
  if (packed) store_packed_block<NC>(cf, t0, M.Tu, packed, lds);
}

// Rows of MWSS triangles next to either pole that a sparse velocity level
// always holds: a foot near a pole crosses many of their thin triangles from
// one step to the next, and every missing block costs its wave one round of
// the fallback in mwss_fill_missing.
constexpr int kPolarRows = 4;

// cmm_mwss_assemble_sparse: one wave per chunk of triangles decides which
// to assemble (mark != NULL: the marked ones and the polar rows, bit `slot`
// of vbits set for them and cleared for the others, mark cleared; mark ==
// NULL: those whose bit is clear, bit set), compacts their ids in LDS and
// assembles them a lane each, so the waves that run the 212-register
// assembly are about as many as the selected triangles fill.  The polar rows
// (always selected) come in chunks of 64, the rest in chunks of 256, so no
// wave carries more than one round of the assembly on average.
struct KMwssAssembleSparse {
  struct Args {
    cmm_mwss_mesh M;
    const double* vals;
    double* packed;
    uint8_t* vbits;
    int slot, inherit;
    uint8_t* mark;
    int64_t npolar, polar_chunks;
  };
  static constexpr int kThreads = 64;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

__device__ __forceinline__ void KMwssAssembleSparse::run(const Args& a, int bid) {
  const cmm_mwss_mesh& M = a.M;
  const double* __restrict__ vals = a.vals;
  double* __restrict__ packed = a.packed;
  uint8_t* __restrict__ vbits = a.vbits;
  uint8_t* __restrict__ mark = a.mark;
  const int slot = a.slot, inherit = a.inherit;
  const int64_t npolar = a.npolar, polar_chunks = a.polar_chunks;
  __shared__ int32_t ids[256];
  __shared__ __align__(16) double stage[16 * 108];
  const int lane = threadIdx.x;
  const unsigned bit = 1u << slot;
  const int64_t Tu = M.Tu;
  const int64_t b = bid;
  int64_t c0, c1;
  if (b < polar_chunks) {  // [0, npolar) then [Tu - npolar, Tu), 64 at a time
    const int64_t half = (polar_chunks + 1) / 2;
    c0 = b < half ? 64 * b : Tu - npolar + 64 * (b - half);
    c1 = b < half ? min(c0 + 64, npolar) : min(c0 + 64, Tu);
  } else {
    c0 = npolar + 256 * (b - polar_chunks);
    c1 = min(c0 + 256, Tu - npolar);
  }
  int nsel = 0;
  for (int64_t base = c0; base < c1; base += 64) {
    const int64_t t = base + lane;
    bool sel = false;
    if (t < c1) {
      const unsigned vb = vbits[t];
      if (mark) {
        const bool polar = t < npolar || t >= Tu - npolar;
        const bool mk = mark[t] != 0;
        sel = mk || polar || (inherit >= 0 && ((vb >> inherit) & 1u));
        if (mk) mark[t] = 0;
        const unsigned nb = sel ? (vb | bit) : (vb & ~bit);
        if (nb != vb) vbits[t] = (uint8_t)nb;
      } else {
        sel = (vb & bit) == 0;
        if (sel) vbits[t] = (uint8_t)(vb | bit);
      }
    }
    const uint64_t bl = __ballot(sel);
    if (sel) ids[nsel + __popcll(bl & ((1ull << lane) - 1))] = (int32_t)t;
    nsel += __popcll(bl);
  }
  __syncthreads();
  // 64 selected triangles at a time a lane each, their blocks written 16 at
  // a time through LDS so that every store instruction covers whole blocks
  typedef double dbl2 __attribute__((ext_vector_type(2)));
  for (int i0 = 0; i0 < nsel; i0 += 64) {
    double cf[3][19];
    const bool mine = i0 + lane < nsel;
    if (mine) mwss_tri_coeffs<3>(M, vals, ids[i0 + lane], cf);
    const int cnt = min(64, nsel - i0);
    for (int r0 = 0; r0 < cnt; r0 += 16) {
      if (mine && lane >= r0 && lane < r0 + 16) {
        double* dst = stage + (lane - r0) * 108;
#pragma unroll
        for (int ss = 0; ss < 6; ++ss)
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int j = 0; j < 6; ++j) dst[(ss * 3 + c) * 6 + j] = cf[c][cs_table(ss, j)];
      }
      __syncthreads();
      const int nb = min(16, cnt - r0);
      const dbl2* src = reinterpret_cast<const dbl2*>(stage);
      for (int e = lane; e < nb * 54; e += 64) {
        const int tl = e / 54;
        dbl2* out = reinterpret_cast<dbl2*>(packed + (int64_t)ids[i0 + r0 + tl] * 108);
        __builtin_nontemporal_store(src[e], out + (e - tl * 54));
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(64) void k_mwss_assemble_sparse(KMwssAssembleSparse::Args a) {
  KMwssAssembleSparse::run(a, blockIdx.x);
}

// grid index (ring, column) of mesh vertex v: poles take the phi = 0 sample
__device__ inline void vertex_grid(const cmm_mwss_mesh& M, int64_t v, int& t, int& p) {
  if (v == 0) { t = 0; p = 0; }
  else if (v == M.Nu - 1) { t = M.L; p = 0; }
  else { t = 1 + (int)((v - 1) / M.N); p = (int)((v - 1) % M.N); }
}

// spherical_harmonic_tools.tan_proj (145-152) of cross(g, s) at s
__device__ inline d3 rot_grad(d3 g, d3 s) {
  d3 a = cross(g, s);
  double scl_ = (a.x * s.x + a.y * s.y + a.z * s.z) / sqrt(s.x * s.x + s.y * s.y + s.z * s.z);
  return d3{a.x - scl_ * s.x, a.y - scl_ * s.y, a.z - scl_ * s.z};
}

struct KVelocityJets {
  struct Args {
    cmm_mwss_mesh M;
    const double* F;
    double* vals;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

__device__ __forceinline__ void KVelocityJets::run(const Args& a, int bid) {
  const cmm_mwss_mesh& M = a.M;
  const double* __restrict__ F = a.F;
  double* __restrict__ vals = a.vals;
  int64_t v = (int64_t)bid * blockDim.x + threadIdx.x;
  if (v >= M.Nu) return;
  int t, p;
  vertex_grid(M, v, t, p);
  const int64_t G = (int64_t)(M.L + 1) * M.N;
  int64_t gi = (int64_t)t * M.N + p;
  d3 s = ring_pt(M, t, p);
  const int64_t Nu = M.Nu;
  for (int i = 0; i < 3; ++i) {
    vals[(i * 4) * Nu + v] = F[i * G + gi];
    d3 g = d3{F[(3 + 3 * i) * G + gi], F[(4 + 3 * i) * G + gi], F[(5 + 3 * i) * G + gi]};
    d3 r = rot_grad(g, s);
    vals[(0 * 4 + 1 + i) * Nu + v] = r.x;
    vals[(1 * 4 + 1 + i) * Nu + v] = r.y;
    vals[(2 * 4 + 1 + i) * Nu + v] = r.z;
  }
}

__global__ __launch_bounds__(256) void k_velocity_jets(KVelocityJets::Args a) { KVelocityJets::run(a, blockIdx.x); }

__global__ __launch_bounds__(256) void k_scalar_jets(cmm_mwss_mesh M, const double* __restrict__ F, double* __restrict__ vals) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= M.Nu) return;
  int t, p;
  vertex_grid(M, v, t, p);
  const int64_t G = (int64_t)(M.L + 1) * M.N;
  int64_t gi = (int64_t)t * M.N + p;
  d3 s = ring_pt(M, t, p);
  const int64_t Nu = M.Nu;
  vals[v] = F[gi];
  d3 r = rot_grad(d3{F[G + gi], F[2 * G + gi], F[3 * G + gi]}, s);
  vals[Nu + v] = r.x;
  vals[2 * Nu + v] = r.y;
  vals[3 * Nu + v] = r.z;
}

struct VortArgs {
  int kind;
  double params[4];
  double rot;
  int ring_len;
};

__device__ inline double vort_analytic(int kind, double phi, double theta, const double* prm) {
  const double pi = M_PI;
  switch (kind) {
    case CMM_VORT_ROSSBY: {
      double alpha = 1.0 / 30.0, Om = 2 * pi, t = prm[0];
      double s = sin(theta);
      return 30 * cos(theta) * cos(4 * (phi + 2 * Om * alpha * t)) * (s * s * s * s);
    }
    case CMM_VORT_ZONAL_JET:
    case CMM_VORT_PERTURBED_JET:
    case CMM_VORT_MULTI_JET: {
      auto jet = [&](double ph, double th, bool pert) {
        double beta2 = 144.0;
        double thc = pi / 4 + (pert ? 0.01 * cos(12 * ph) : 0.0);
        double ult = (pi / 2) * exp(-2 * beta2 * (1 - cos(-th + pi / 2 - thc)));
        return (cos(-th + pi / 2) * (2 * beta2 * (cos(thc) * sin(-th + pi / 2) - sin(thc) * cos(-th + pi / 2))) +
                sin(pi / 2 - th)) * ult;
      };
      if (kind == CMM_VORT_ZONAL_JET) return jet(phi, theta, false);
      if (kind == CMM_VORT_PERTURBED_JET) return jet(phi, theta, true);
      return jet(phi, theta, true) + jet(phi, theta - pi / 8, true);
    }
    case CMM_VORT_GAUSSIAN: {
      double x = sin(theta) * cos(phi), y = sin(theta) * sin(phi), z = cos(theta);
      return 4 * pi * exp(-((x - 1) * (x - 1) + y * y + z * z) / (1.0 / 16));
    }
    case CMM_VORT_PERTURBED_ROSSBY: {
      double s = sin(theta);
      return 30 * cos(theta) * cos(4 * phi) * (s * s * s * s) + 0.1 * sin(theta) * cos(30 * theta);
    }
    case CMM_VORT_COSINE_BELL:
    case CMM_VORT_SLOTTED_CYLINDER: {
      // dynamical_fields.py:110-156: two bells / slotted cylinders of radius
      // R = 1/2 centred at (pi +- pi/6, pi/2)
      const double R = 0.5, h0 = 1.0;
      const double pc[2] = {pi + pi / 6, pi - pi / 6}, tc = pi / 2;
      double o[2];
      for (int k = 0; k < 2; ++k) {
        const double arg = sin(tc - pi / 2) * sin(theta - pi / 2) +
                           cos(tc - pi / 2) * cos(theta - pi / 2) * cos(phi - pc[k]);
        const double r = acos(arg);
        if (kind == CMM_VORT_COSINE_BELL) {
          o[k] = r < R ? (h0 / 2) * (1 + cos(r * pi / R)) : 0.0;
        } else {
          const bool in = r <= R, slot = fabs(phi - pc[k]) < R / 6;
          const bool keep = k == 0 ? theta < (-5 * R / 12 + pi / 2) : theta > (5 * R / 12 + pi / 2);
          o[k] = (in && (!slot || keep)) ? 0.9 : 0.0;
        }
      }
      return kind == CMM_VORT_COSINE_BELL ? 0.9 * (o[0] + o[1]) + 0.1 : o[0] + o[1] + 0.1;
    }
    case CMM_VORT_STATIC_VORTEX_TRACER:
    case CMM_VORT_ROTATING_VORTEX_TRACER: {
      // dynamical_fields.py:87-97: 1 - tanh(rho/5 sin(phi - omega_r(theta) t)),
      // rho = 3 sin(theta); the rotating form in the frame Rot_x(pi/2)
      double ph = phi, th = theta;
      if (kind == CMM_VORT_ROTATING_VORTEX_TRACER) {
        const double c = cos(M_PI / 2), s = sin(M_PI / 2);
        const double Rx[3][3] = {{1, 0, 0}, {0, c, -s}, {0, s, c}};
        cart2sphere(rot_apply(Rx, d3{sin(theta) * cos(phi), sin(theta) * sin(phi), cos(theta)}), ph, th);
      }
      const double rho = 3.0 * sin(th);
      return 1 - tanh((rho / 5) * sin(ph - vortex_omega(th) * prm[0]));
    }
    default:
      return __builtin_nan("");
  }
}

// real field sum_l sum_{m >= 0} w_m Re(a_lm Y_lm(theta, phi)), w_0 = 1,
// w_m = 2 (a_l,-m = (-1)^m conj a_lm), coefficients a [m][l] complex (Ls, Ls)
__device__ inline double sh_series(const double* __restrict__ a, int Ls, double phi, double theta) {
  const double x = cos(theta), st = sin(theta);
  double pmm = 1.0 / sqrt(4 * M_PI), sum = 0.0;
  for (int m = 0; m < Ls; ++m) {
    if (m > 0) pmm *= -sqrt((2.0 * m + 1.0) / (2.0 * m)) * st;
    const double cm = cos(m * phi), sm = sin(m * phi);
    const double* am = a + 2 * (int64_t)m * Ls;
    double p0 = pmm, acc = am[2 * m] * cm - am[2 * m + 1] * sm;
    acc *= p0;
    if (m + 1 < Ls) {
      double p1 = sqrt(2.0 * m + 3.0) * x * p0;
      acc += p1 * (am[2 * (m + 1)] * cm - am[2 * (m + 1) + 1] * sm);
      for (int l = m + 2; l < Ls; ++l) {
        const double al = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
        const double bl = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
        const double p2 = al * (x * p1 - bl * p0);
        acc += p2 * (am[2 * l] * cm - am[2 * l + 1] * sm);
        p0 = p1;
        p1 = p2;
      }
    }
    sum += (m == 0 ? 1.0 : 2.0) * acc;
  }
  return sum;
}

struct KSampleVorticity {
  struct Args {
    VortArgs A;
    cmm_mwss_mesh W;
    const double* wpacked;
    const double* X;
    int64_t n, sp, sc;
    const double* cosg;
    double* out;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

__device__ __forceinline__ void KSampleVorticity::run(const Args& a, int bid) {
  const VortArgs& A = a.A;
  const cmm_mwss_mesh& W = a.W;
  const double* __restrict__ wpacked = a.wpacked;
  const double* __restrict__ X = a.X;
  const double* __restrict__ cosg = a.cosg;
  double* __restrict__ out = a.out;
  const int64_t n = a.n, sp = a.sp, sc = a.sc;
  int64_t i = (int64_t)bid * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d3 x = load_pt(X, i, sp, sc);
  double w0, cth;
  if (A.kind == CMM_VORT_SPLINE) {
    // spherical_spline.__call__ (spherical_spline.py:42-70) takes the angles
    // of X (cart2sphere, utils.py:18-23) back to xyz and locates that point:
    // the round trip is X / |X| (to rounding), located here directly, and
    // cos theta(X) = z / |X| (differences of a few ulp, no transcendental
    // functions before the location's own angles)
    const double nn = sqrt(x.x * x.x + x.y * x.y + x.z * x.z);
    const d3 q = d3{x.x / nn, x.y / nn, x.z / nn};
    LevelSet ls;
    ls.p[0] = wpacked; ls.w[0] = 1.0; ls.nlev = 1;
    double r[1];
    if (!mwss_eval_levels(W, ls, 1, q, r)) r[0] = __builtin_nan("");
    w0 = r[0];
    cth = q.z;
  } else {
    // utils.cart2sphere (utils.py:18-23)
    const double phi = npy_mod(atan2(x.y, x.x) + 2 * M_PI, 2 * M_PI);
    const double theta = atan2(sqrt(x.y * x.y + x.x * x.x), x.z);
    w0 = A.kind == CMM_VORT_SH_SERIES ? sh_series(wpacked, (int)A.params[0], phi, theta)
                                      : vort_analytic(A.kind, phi, theta, A.params);
    cth = cos(theta);
  }
  int ring = (int)(i / A.ring_len);
  out[i] = w0 + 2 * A.rot * cth - 2 * A.rot * cosg[ring];
}

__global__ __launch_bounds__(256) void k_sample_vorticity(KSampleVorticity::Args a) { KSampleVorticity::run(a, blockIdx.x); }

// ---------------------------------------------------------------------------
// adjoints (reverse mode).  Kernels marked "accumulate" add into their
// outputs with atomics (the caller clears them); the others overwrite.
// ---------------------------------------------------------------------------
// transpose of rot_grad in g: r = P (g x s), P = I - s s^T / |s| symmetric
__device__ inline d3 rot_grad_T(d3 rb, d3 s) {
  const double c = dot(rb, s) / sqrt(dot(s, s));
  const d3 ab = d3{rb.x - c * s.x, rb.y - c * s.y, rb.z - c * s.z};
  return cross(s, ab);
}

// mesh vertex sampled by grid point (t, p), -1 for the pole samples p > 0
__device__ inline int64_t grid_vertex(const cmm_mwss_mesh& M, int t, int p) {
  if (t == 0) return p == 0 ? 0 : -1;
  if (t == M.L) return p == 0 ? M.Nu - 1 : -1;
  return 1 + (int64_t)(t - 1) * M.N + p;
}

// transpose of k_velocity_jets (NF = 12) / k_scalar_jets (NF = 4): every grid
// point receives from at most one vertex, so one thread per grid point
// overwrites all NF fields
template <int NF>
__global__ __launch_bounds__(256) void k_grid_jets_adj(cmm_mwss_mesh M, const double* __restrict__ vb,
                                                       double* __restrict__ Fb) {
  const int64_t G = (int64_t)(M.L + 1) * M.N;
  const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= G) return;
  const int t = (int)(gi / M.N), p = (int)(gi % M.N);
  const int64_t v = grid_vertex(M, t, p);
  const int64_t Nu = M.Nu;
  if (v < 0) {
    for (int k = 0; k < NF; ++k) Fb[k * G + gi] = 0.0;
    return;
  }
  const d3 s = ring_pt(M, t, p);
  if (NF == 12) {
    for (int i = 0; i < 3; ++i) {
      Fb[i * G + gi] = vb[(i * 4) * Nu + v];
      const d3 rb = d3{vb[(0 * 4 + 1 + i) * Nu + v], vb[(1 * 4 + 1 + i) * Nu + v], vb[(2 * 4 + 1 + i) * Nu + v]};
      const d3 g = rot_grad_T(rb, s);
      Fb[(3 + 3 * i) * G + gi] = g.x;
      Fb[(4 + 3 * i) * G + gi] = g.y;
      Fb[(5 + 3 * i) * G + gi] = g.z;
    }
  } else {
    Fb[gi] = vb[v];
    const d3 g = rot_grad_T(d3{vb[Nu + v], vb[2 * Nu + v], vb[3 * Nu + v]}, s);
    Fb[G + gi] = g.x;
    Fb[2 * G + gi] = g.y;
    Fb[3 * G + gi] = g.z;
  }
}

// transpose of k_mwss_assemble (accumulate into the jets vb): packed
// adjoints -> 19 coefficient adjoints per component (sum over the
// sub-triangles that share a coefficient) -> ps_coeffs^T -> the 3 vertices
template <int NC>
__global__ __launch_bounds__(256) void k_mwss_assemble_adj(cmm_mwss_mesh M, const double* __restrict__ pb,
                                                           double* __restrict__ vb) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M.Tu) return;
  const int N = M.N, Mr = M.M;
  int row, col, half;
  if (t < N) { row = 0; col = (int)t; half = 0; }
  else if (t >= M.Tu - N) { row = Mr - 2; col = (int)(t - (M.Tu - N)); half = 0; }
  else {
    const int64_t u = t - N;
    half = (int)(u % 2);
    const int64_t k = u / 2;
    row = 1 + (int)(k / N);
    col = (int)(k % N);
  }
  int64_t vi[3];
  d3 v[3];
  mwss_tri_vertices(M, t, vi, v);
  PSGeom P;
  ps_unpack(M.rowgeom + ((int64_t)row * 2 + half) * kPSGeomLen, 1, P);
  const double c = M.cp[col], sn = M.sp[col];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const d3 h = P.H[i];
    P.H[i] = d3{c * h.x - sn * h.y, sn * h.x + c * h.y, h.z};
  }
  const int64_t Nu = M.Nu;
  const double* src = pb + t * (36 * NC);
#pragma unroll
  for (int cc = 0; cc < NC; ++cc) {
    double cb[19];
#pragma unroll
    for (int k = 0; k < 19; ++k) cb[k] = 0.0;
#pragma unroll
    for (int ss = 0; ss < 6; ++ss)
#pragma unroll
      for (int j = 0; j < 6; ++j) cb[cs_table(ss, j)] += src[(ss * NC + cc) * 6 + j];
    double fb[3] = {0.0, 0.0, 0.0};
    d3 gb[3] = {d3{0, 0, 0}, d3{0, 0, 0}, d3{0, 0, 0}};
    ps_coeffs_adj(P, cb, fb, gb);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (NC == 1) {
        unsafeAtomicAdd(vb + vi[i], fb[i]);
        unsafeAtomicAdd(vb + Nu + vi[i], gb[i].x);
        unsafeAtomicAdd(vb + 2 * Nu + vi[i], gb[i].y);
        unsafeAtomicAdd(vb + 3 * Nu + vi[i], gb[i].z);
      } else {
        unsafeAtomicAdd(vb + (cc * 4) * Nu + vi[i], fb[i]);
        unsafeAtomicAdd(vb + (0 * 4 + 1 + cc) * Nu + vi[i], gb[i].x);
        unsafeAtomicAdd(vb + (1 * 4 + 1 + cc) * Nu + vi[i], gb[i].y);
        unsafeAtomicAdd(vb + (2 * 4 + 1 + cc) * Nu + vi[i], gb[i].z);
      }
    }
  }
}

// Two-phase transpose of k_mwss_assemble<NC> without atomics.  Phase 1, one
// thread per triangle: the jet adjoints of its corners, corner i row r at
// tmp[(4 NC i + r) Tu + t] (rows in the jets order).  Phase 2, one thread per
// non-pole vertex: its (at most 6) incident (triangle, corner) pairs in closed
// form, added into vb; one block per pole reduces the 2L cap triangles.
template <int NC>
__global__ __launch_bounds__(64 * NC) void k_mwss_assemble_adj_corners(cmm_mwss_mesh M, const double* __restrict__ pb,
                                                                       double* __restrict__ tmp) {
  // 64 triangles per block: their packed blocks (contiguous) are staged
  // through LDS with coalesced 16-byte loads, then thread (component cc,
  // triangle tt) forms that component's corner adjoints
  constexpr int W = 36 * NC;  // doubles per triangle
  __shared__ double sh[64 * (W + 1)];
  const int64_t t0 = (int64_t)blockIdx.x * 64;
  const int64_t Tu = M.Tu;
  const int64_t nvalid = min((int64_t)64, Tu - t0) * W;  // doubles of this block in pb
  const double2* src = reinterpret_cast<const double2*>(pb + t0 * W);
#pragma unroll 4
  for (int q = threadIdx.x; q < 32 * W; q += 64 * NC) {
    const int e = 2 * q, tt = e / W, r = e - W * tt;
    const double2 v = e < nvalid ? src[q] : double2{0.0, 0.0};
    sh[tt * (W + 1) + r] = v.x;
    sh[tt * (W + 1) + r + 1] = v.y;
  }
  __syncthreads();
  const int tt = threadIdx.x & 63, cc = threadIdx.x >> 6;
  const int64_t t = t0 + tt;
  if (t >= Tu) return;
  const int N = M.N, Mr = M.M;
  int row = 0, col = 0, half = 0;
  if (t < N) { row = 0; col = (int)t; half = 0; }
  else if (t >= Tu - N) { row = Mr - 2; col = (int)(t - (Tu - N)); half = 0; }
  else {
    const int64_t u = t - N;
    half = (int)(u % 2);
    const int64_t k = u / 2;
    row = 1 + (int)(k / N);
    col = (int)(k % N);
  }
  PSGeom P;
  ps_unpack(M.rowgeom + ((int64_t)row * 2 + half) * kPSGeomLen, 1, P);
  const double c = M.cp[col], sn = M.sp[col];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const d3 h = P.H[i];
    P.H[i] = d3{c * h.x - sn * h.y, sn * h.x + c * h.y, h.z};
  }
  constexpr int R = 4 * NC;
  double cb[19];
#pragma unroll
  for (int k = 0; k < 19; ++k) cb[k] = 0.0;
#pragma unroll
  for (int ss = 0; ss < 6; ++ss)
#pragma unroll
    for (int j = 0; j < 6; ++j) cb[cs_table(ss, j)] += sh[tt * (W + 1) + (ss * NC + cc) * 6 + j];
  double fb[3] = {0.0, 0.0, 0.0};
  d3 gb[3] = {d3{0, 0, 0}, d3{0, 0, 0}, d3{0, 0, 0}};
  ps_coeffs_adj(P, cb, fb, gb);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (NC == 1) {
      tmp[(R * i + 0) * Tu + t] = fb[i];
      tmp[(R * i + 1) * Tu + t] = gb[i].x;
      tmp[(R * i + 2) * Tu + t] = gb[i].y;
      tmp[(R * i + 3) * Tu + t] = gb[i].z;
    } else {
      tmp[(R * i + cc * 4) * Tu + t] = fb[i];
      tmp[(R * i + 0 * 4 + 1 + cc) * Tu + t] = gb[i].x;
      tmp[(R * i + 1 * 4 + 1 + cc) * Tu + t] = gb[i].y;
      tmp[(R * i + 2 * 4 + 1 + cc) * Tu + t] = gb[i].z;
    }
  }
}

// Corner jets of the transposed assembly, in one of two layouts: the
// per-triangle pass of a packed adjoint writes corner i row r of triangle t
// at p[(R i + r) Tu + t] for every triangle; the sparse velocity adjoint
// (k_mwss_corner_reduce) at p[3 R t + R i + r] for the triangles with
// touched[t] != 0 only (the others hold no adjoint and are never read).
template <int NC>
struct CornerSrc {
  static constexpr int R = 4 * NC;
  const double* p;
  const uint8_t* touched;  // null: the first layout
  int64_t Tu;
  __device__ bool has(int64_t t) const { return !touched || touched[t] != 0; }
  __device__ double at(int64_t t, int i, int r) const {
    return touched ? p[(3 * R) * t + R * i + r] : p[(R * i + r) * Tu + t];
  }
};

// pole p (0 north: corner 0 of triangles 0..N-1; 1 south: corner 1 of the
// last N triangles), reduced by one block of 256 threads
template <int NC>
__device__ inline void mwss_pole_reduce(const cmm_mwss_mesh& M, const CornerSrc<NC>& src, double* __restrict__ vb,
                                        int p) {
  constexpr int R = 4 * NC;
  __shared__ double red[R][4];
  const int N = M.N;
  const int64_t Tu = M.Tu;
  const int64_t t0 = p == 0 ? 0 : Tu - N;
  const int corner = p == 0 ? 0 : 1;
  const int64_t v = p == 0 ? 0 : M.Nu - 1;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double a[R];
#pragma unroll
  for (int q = 0; q < R; ++q) a[q] = 0.0;
  for (int c = threadIdx.x; c < N; c += 256)
    if (src.has(t0 + c))
#pragma unroll
      for (int q = 0; q < R; ++q) a[q] += src.at(t0 + c, corner, q);
#pragma unroll
  for (int q = 0; q < R; ++q) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a[q] += __shfl_xor(a[q], o);
    if (lane == 0) red[q][w] = a[q];
  }
  __syncthreads();
  if (threadIdx.x < R)
    vb[threadIdx.x * M.Nu + v] += ((red[threadIdx.x][0] + red[threadIdx.x][1]) + red[threadIdx.x][2]) +
                                  red[threadIdx.x][3];
}

// blocks 0 .. ceil((Nu - 2) / 256) - 1: the interior vertices; the last two
// blocks: the poles
template <int NC>
__global__ __launch_bounds__(256) void k_mwss_assemble_adj_gather(cmm_mwss_mesh M, CornerSrc<NC> src,
                                                                  double* __restrict__ vb) {
  const int nbi = (int)((M.Nu - 2 + 255) / 256);
  if ((int)blockIdx.x >= nbi) {
    mwss_pole_reduce<NC>(M, src, vb, (int)blockIdx.x - nbi);
    return;
  }
  const int64_t v = 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= M.Nu - 1) return;
  const int N = M.N, L = M.L;
  const int r = 1 + (int)((v - 1) / N), c = (int)((v - 1) % N);
  const int cm = (c + N - 1) % N;
  const int64_t Tu = M.Tu;
  auto mid = [&](int rr, int cc, int h) -> int64_t { return N + 2 * ((int64_t)(rr - 1) * N + cc) + h; };
  int64_t ts[6];
  int cs[6];
  int k = 0;
  if (r == 1) { ts[k] = c; cs[k++] = 1; ts[k] = cm; cs[k++] = 2; }
  if (r == L - 1) { ts[k] = Tu - N + c; cs[k++] = 0; ts[k] = Tu - N + cm; cs[k++] = 2; }
  if (r <= L - 2) {
    ts[k] = mid(r, c, 0); cs[k++] = 0;
    ts[k] = mid(r, c, 1); cs[k++] = 0;
    ts[k] = mid(r, cm, 0); cs[k++] = 2;
  }
  if (r >= 2) {
    ts[k] = mid(r - 1, cm, 0); cs[k++] = 1;
    ts[k] = mid(r - 1, cm, 1); cs[k++] = 2;
    ts[k] = mid(r - 1, c, 1); cs[k++] = 1;
  }
  constexpr int R = 4 * NC;
  double acc[R];
#pragma unroll
  for (int q = 0; q < R; ++q) acc[q] = 0.0;
#pragma unroll
  for (int e = 0; e < 6; ++e) {  // fixed trip count: every load of the vertex in flight together
    if (e < k && src.has(ts[e])) {
#pragma unroll
      for (int q = 0; q < R; ++q) acc[q] += src.at(ts[e], cs[e], q);
    }
  }
  const int64_t Nu = M.Nu;
#pragma unroll
  for (int q = 0; q < R; ++q) vb[q * Nu + v] += acc[q];
}

// (row, column, half) of MWSS triangle t (< 2^31): N cap triangles, 2 N per
// band, N cap triangles
__device__ inline void mwss_tri_cell(const cmm_mwss_mesh& M, int t, int& row, int& col, int& half) {
  const int N = M.N, Tu = (int)M.Tu;
  if (t < N) { row = 0; col = t; half = 0; }
  else if (t >= Tu - N) { row = M.M - 2; col = t - (Tu - N); half = 0; }
  else {
    const int u = t - N;
    half = u & 1;
    const int k = u >> 1;
    row = 1 + k / N;
    col = k - (row - 1) * N;
  }
}

// Sparse transpose of the velocity assembly for the stencil-feet adjoint.
// The records (value records, 18 sums, class = RK stage, sub = sub-triangle)
// are grouped by triangle, with the lists of the triangles that have records
// (at most kCornerSplit of them: one thread per (triangle, level); more: one
// wave).  For triangle t and level o, the stage-weighted sums of its records
// through the transposed sub-triangle packing and ps_coeffs_adj give the jets
// adjoints of the 3 corners, added into corners[o] (layout of CornerSrc with
// touched flags; stored on the first touch).  Each (t, o) has one writer.
constexpr uint32_t kCornerSplit = 8;

struct CornerBars {
  double* c[3];
  uint8_t* touched[3];
  double W[4][3];
  int nout;
};

// level o's arrays and stage weights, selected without indexing the kernel
// argument block (a dynamic index would copy it to scratch)
struct CornerLevel {
  double* c;
  uint8_t* touched;
  double w[4];
  __device__ __forceinline__ double weight(int cls) const {
    return cls == 0 ? w[0] : (cls == 1 ? w[1] : (cls == 2 ? w[2] : w[3]));
  }
};
#define CORNER_LEVEL(out, o, L)                                                                      \
  CornerLevel L;                                                                                     \
  L.c = (o) == 0 ? (out).c[0] : ((o) == 1 ? (out).c[1] : (out).c[2]);                                \
  L.touched = (o) == 0 ? (out).touched[0] : ((o) == 1 ? (out).touched[1] : (out).touched[2]);        \
  for (int c_ = 0; c_ < 4; ++c_)                                                                     \
    L.w[c_] = (o) == 0 ? (out).W[c_][0] : ((o) == 1 ? (out).W[c_][1] : (out).W[c_][2]);

// one record into the coefficient adjoints: every sub-triangle's 18 terms
// with the weight zeroed unless it is the record's (branch-free, constant
// indices: a switch over the sub-triangle is merged by the compiler into an
// indexed access that puts cb in scratch)
__device__ __forceinline__ void corner_accumulate(const ScatterRecV& rv, const CornerLevel& lv, double (&cb)[3][19]) {
  const double wt = lv.weight(rv.cls);
  const int sub = rv.sub;
  double v[18];
#pragma unroll
  for (int e = 0; e < 18; ++e) v[e] = rv.v[e];
#pragma unroll
  for (int S = 0; S < 6; ++S) {
    const double m = sub == S ? wt : 0.0;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int j = 0; j < 6; ++j) cb[cc][cs_table(S, j)] += m * v[6 * cc + j];
  }
}

// the row geometry (unrotated) and column rotation of triangle t
struct CornerGeom {
  PSGeom P;
  double c, sn;
  __device__ __forceinline__ CornerGeom(const cmm_mwss_mesh& M, uint32_t t) {
    int row, col, half;
    mwss_tri_cell(M, (int)t, row, col, half);
    ps_unpack(M.rowgeom + ((int64_t)row * 2 + half) * kPSGeomLen, 1, P);
    c = M.cp[col];
    sn = M.sp[col];
  }
};

// component cc: coefficient adjoints cb -> its 12 corner jets adjoints
// res[12 i + row].  ps_coeffs_adj runs on the unrotated row geometry and the
// gradient adjoints are rotated to the column afterwards (H enters them
// linearly).
__device__ __forceinline__ void corner_jets(const CornerGeom& G, int cc, const double (&cb)[19], double (&res)[36]) {
  double fb[3] = {0.0, 0.0, 0.0};
  d3 gb[3] = {d3{0, 0, 0}, d3{0, 0, 0}, d3{0, 0, 0}};
  ps_coeffs_adj(G.P, cb, fb, gb);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    res[12 * i + cc * 4] = fb[i];                            // f
    res[12 * i + 1 + cc] = G.c * gb[i].x - G.sn * gb[i].y;   // d/dx
    res[12 * i + 5 + cc] = G.sn * gb[i].x + G.c * gb[i].y;   // d/dy
    res[12 * i + 9 + cc] = gb[i].z;                          // d/dz
  }
}

// short groups: a fixed grid strides over the blocks (list chunk k / nout,
// level k % nout: the levels of a chunk run together and share its records
// in L2)
__global__ __launch_bounds__(128) void k_mwss_corner_reduce(cmm_mwss_mesh M, ScatterWork w, CornerBars out) {
  const uint32_t nlist = w.meta[0];
  const uint32_t nblk = (nlist + 127) / 128 * (uint32_t)out.nout;
  for (uint32_t k = blockIdx.x; k < nblk; k += gridDim.x) {
    const int o = (int)(k % out.nout);
    const uint32_t j = (k / out.nout) * 128 + threadIdx.x;
    if (j >= nlist) continue;
    const uint32_t t = w.keys_u[j];
    const uint32_t a = w.offs[t], b = w.offs[t + 1];
    CORNER_LEVEL(out, o, lv)
    double cb[3][19];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int q = 0; q < 19; ++q) cb[cc][q] = 0.0;
    for (uint32_t r = a; r < b; ++r) corner_accumulate(w.recv[w.idx_s[r]], lv, cb);
    const CornerGeom G(M, t);
    double res[36];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) corner_jets(G, cc, cb[cc], res);
    double2* dst = reinterpret_cast<double2*>(lv.c + (int64_t)t * 36);
    if (lv.touched[t] != 0) {
      double2 old[18];
#pragma unroll
      for (int q = 0; q < 18; ++q) old[q] = dst[q];
#pragma unroll
      for (int q = 0; q < 18; ++q) {
        res[2 * q] += old[q].x;
        res[2 * q + 1] += old[q].y;
      }
    }
#pragma unroll
    for (int q = 0; q < 18; ++q) dst[q] = double2{res[2 * q], res[2 * q + 1]};
    lv.touched[t] = 1;
  }
}

// long groups: 16-lane segments (4 per wave) of a fixed grid stride over
// (list entry g / nout, level g % nout); the segment's lanes stride over the
// records, a butterfly within the segment sums them, segment lanes 0..2
// store the 3 components.  Segments past the end take part in the
// butterflies with zeros and store nothing.
__global__ __launch_bounds__(256) void k_mwss_corner_reduce_long(cmm_mwss_mesh M, ScatterWork w, CornerBars out) {
  const int lane = (int)(threadIdx.x & 63), sl = lane & 15;
  const uint32_t ng = w.meta[1] * (uint32_t)out.nout;
  const uint32_t nseg = gridDim.x * (blockDim.x / 16);
  const uint32_t seg0 = (blockIdx.x * blockDim.x + (threadIdx.x & ~63u)) / 16;  // first segment of this wave
  for (uint32_t gw = seg0; gw < ng; gw += nseg) {  // wave-uniform
    const uint32_t g = gw + (uint32_t)(lane >> 4);
    const bool live = g < ng;
    const int o = live ? (int)(g % out.nout) : 0;
    const uint32_t t = live ? w.keys_l[g / out.nout] : 0u;
    const uint32_t a = live ? w.offs[t] : 0u, b = live ? w.offs[t + 1] : 0u;
    CORNER_LEVEL(out, o, lv)
    double cb[3][19];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int q = 0; q < 19; ++q) cb[cc][q] = 0.0;
    for (uint32_t r = a + sl; r < b; r += 16) corner_accumulate(w.recv[w.idx_s[r]], lv, cb);
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        double v = cb[cc][q];
#pragma unroll
        for (int m = 8; m > 0; m >>= 1) v += __shfl_xor(v, m);
        cb[cc][q] = v;
      }
    const bool fresh = live && lv.touched[t] == 0;
    if (live && sl < 3) {
      const CornerGeom G(M, t);
      double res[36];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        if (sl == cc) corner_jets(G, cc, cb[cc], res);
      double* dst = lv.c + (int64_t)t * 36;
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        if (sl == cc)
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const int rr[4] = {12 * i + cc * 4, 12 * i + 1 + cc, 12 * i + 5 + cc, 12 * i + 9 + cc};
#pragma unroll
            for (int q = 0; q < 4; ++q) dst[rr[q]] = fresh ? res[rr[q]] : dst[rr[q]] + res[rr[q]];
          }
    }
    // every lane of the segment has read the flag above
    if (live && sl == 0) lv.touched[t] = 1;
  }
}

// located velocity-spline point: triangle, sub-triangle, sub-triangle
// barycentrics and the pieces of their position gradient
struct VLoc {
  bool ok;
  int64_t tri;
  int s;
  d3 cb;
  d3 w[3];  // macro dual basis: b_i = x . w_i
  double se, s4;
};

__device__ __forceinline__ VLoc vel_locate(const cmm_mwss_mesh& M, d3 x) {
  VLoc r;
  MLoc o;
  r.ok = mwss_locate(M, x, o);
  if (!r.ok) {
    r.tri = -1;
    r.s = 0;
    return r;
  }
  r.tri = o.tri;
  r.s = split_index(o.b.x, o.b.y, o.b.z);
  const d3 s3 = add(add(o.v[0], o.v[1]), o.v[2]);
  r.s4 = norm3(s3);
  if (r.s <= 1) r.se = norm3(add(o.v[1], o.v[2]));
  else if (r.s <= 3) r.se = norm3(add(o.v[2], o.v[0]));
  else r.se = norm3(add(o.v[0], o.v[1]));
  r.cb = sub_bary(r.s, o.b.x, o.b.y, o.b.z, r.se, r.se, r.se, r.s4);
  const double D = det3(o.v[0], o.v[1], o.v[2]);
  r.w[0] = scl(1 / D, cross(o.v[1], o.v[2]));
  r.w[1] = scl(1 / D, cross(o.v[2], o.v[0]));
  r.w[2] = scl(1 / D, cross(o.v[0], o.v[1]));
  return r;
}

// packed adjoints of the velocity levels (atomic form)
struct LevelBars {
  double* p[3];
};

// vel_locate for a point already located in triangle tri: the closed-form
// vertices of tri are those mwss_locate builds, bit for bit, without the
// angles and cell search; the barycentrics then agree with vel_locate's to
// rounding (the compiler may fuse the Cramer products differently in the two
// inlined contexts)
__device__ __forceinline__ VLoc vel_locate_at(const cmm_mwss_mesh& M, d3 x, int64_t tri) {
  VLoc r;
  MLoc o;
  o.tri = tri;
  mwss_tri_vertices(M, tri, o.vi, o.v);
  o.b = bary_fast(o.v[0], o.v[1], o.v[2], x);
  r.ok = true;
  r.tri = o.tri;
  r.s = split_index(o.b.x, o.b.y, o.b.z);
  const d3 s3 = add(add(o.v[0], o.v[1]), o.v[2]);
  r.s4 = norm3(s3);
  if (r.s <= 1) r.se = norm3(add(o.v[1], o.v[2]));
  else if (r.s <= 3) r.se = norm3(add(o.v[2], o.v[0]));
  else r.se = norm3(add(o.v[0], o.v[1]));
  r.cb = sub_bary(r.s, o.b.x, o.b.y, o.b.z, r.se, r.se, r.se, r.s4);
  const double D = det3(o.v[0], o.v[1], o.v[2]);
  r.w[0] = scl(1 / D, cross(o.v[1], o.v[2]));
  r.w[1] = scl(1 / D, cross(o.v[2], o.v[0]));
  r.w[2] = scl(1 / D, cross(o.v[0], o.v[1]));
  return r;
}

// split location of points on the velocity mesh: the full search, or from a
// known triangle (the form the RK-feet adjoint uses for its stage points)
__global__ __launch_bounds__(256) void k_mwss_split_locate(cmm_mwss_mesh M, const double* __restrict__ q, int64_t n,
                                                           int64_t sp, int64_t sc, const int32_t* __restrict__ known,
                                                           int32_t* __restrict__ tri, int32_t* __restrict__ sub,
                                                           double* __restrict__ cb, double* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const d3 x = load_pt(q, i, sp, sc);
  const int64_t k = known ? (int64_t)known[i] : -1;
  const VLoc r = known ? (k >= 0 && k < M.Tu ? vel_locate_at(M, x, k) : VLoc{false, -1, 0}) : vel_locate(M, x);
  tri[i] = r.ok ? (int32_t)r.tri : -1;
  if (sub) sub[i] = r.ok ? r.s : -1;
  const double nan = __builtin_nan("");
  if (cb) {
    cb[3 * i] = r.ok ? r.cb.x : nan;
    cb[3 * i + 1] = r.ok ? r.cb.y : nan;
    cb[3 * i + 2] = r.ok ? r.cb.z : nan;
  }
  if (w) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      w[9 * i + 3 * j] = r.ok ? r.w[j].x : nan;
      w[9 * i + 3 * j + 1] = r.ok ? r.w[j].y : nan;
      w[9 * i + 3 * j + 2] = r.ok ? r.w[j].z : nan;
    }
  }
}

// Adjoint of one velocity-spline evaluation v = sum_k w_k PS_k(x) of stage
// `stage`: adds w_k vbar_c basis_j into the level adjoints lb[k] (atomics,
// summed over the 4 lanes of a quad first when the whole quad sits in one
// sub-triangle: the four stencil feet of a vertex are 1e-5 apart) and
// returns xbar = (dv/dx)^T vbar.  Every lane of the quad must call it.
// Records are keyed by sub-triangle (tri_keys false) or by triangle (true,
// the sub-triangle travels in the record).
__device__ __forceinline__ d3 vel_adj(const VelDev& V, int stage, d3 x, d3 vb, const LevelBars& lb,
                                      const ScatterWork& sw, bool rec, int64_t slot, bool tri_keys,
                                      int64_t tri_known) {
  const VLoc r = tri_known >= 0 ? vel_locate_at(V.mesh, x, tri_known) : vel_locate(V.mesh, x);
  const LevelSet& ls = V.ls[stage];
  const int lane = threadIdx.x & 63;
  const int64_t key = r.ok ? r.tri * 6 + r.s : -1 - (int64_t)lane;
  const int64_t k1 = __shfl_xor(key, 1), k2 = __shfl_xor(key, 2);
  const bool e1 = key == k1, e2 = key == k2;
  // every lane must execute the shuffle (no short-circuit: a lane that skips
  // it leaves its partner reading an inactive lane)
  const int e1p = __shfl_xor((int)e1, 2);
  const bool same = e1 && e2 && e1p;
  double bas[6];
  bb6_basis(r.cb, bas);
  double cv[18];
  const double vv[3] = {vb.x, vb.y, vb.z};
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int j = 0; j < 6; ++j) cv[6 * c + j] = r.ok ? vv[c] * bas[j] : 0.0;
  // quad sums in convergent code (every lane shuffles), then select
#pragma unroll
  for (int e = 0; e < 18; ++e) {
    double t = cv[e] + __shfl_xor(cv[e], 1);
    t += __shfl_xor(t, 2);
    cv[e] = same ? t : cv[e];
  }
  const bool issue = r.ok && (!same || (lane & 3) == 0);
  if (rec) {
    // value record for the sorted scatter-add (class = stage; the stage
    // weights are applied there); slot < 0: a lane past the end
    if (slot >= 0) {
      if (issue)
        scatter_put_values(sw, slot, (uint32_t)(tri_keys ? r.tri : r.tri * 6 + r.s), cv, stage, r.s);
      else
        scatter_empty(sw, slot);
    }
  }
  d3 cbar = d3{0, 0, 0};
  if (r.ok) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (k < ls.nlev) {
        const double w = ls.w[k];
        const double* p = ls.p[k] + (r.tri * 6 + r.s) * 18;
#pragma unroll
        for (int c = 0; c < 3; ++c) cbar = add(cbar, scl(w * vv[c], bb6_db(r.cb, p + 6 * c)));
        if (!rec && issue && w != 0.0) {
          double* dst = lb.p[k] + (r.tri * 6 + r.s) * 18;
#pragma unroll
          for (int e = 0; e < 18; ++e) unsafeAtomicAdd(dst + e, w * cv[e]);
        }
      }
    }
  }
  if (!r.ok) return d3{0, 0, 0};
  const d3 bbar = sub_bary_T(r.s, cbar, r.se, r.se, r.se, r.s4);
  return add(add(scl(bbar.x, r.w[0]), scl(bbar.y, r.w[1])), scl(bbar.z, r.w[2]));
}


// Adjoint of advect_point for the stencil feet (integrators Euler, improved
// Euler, RK4): recompute the stages, then reverse them.  Point i = (vertex
// i / 4, stencil k = i % 4) as k_advect_stencil<4>, so a quad of lanes holds
// the four feet of one vertex.  ybar (3, 4 nv) with row stride yb_row; the
// level adjoints accumulate.
__global__ __launch_bounds__(128) void k_advect_stencil_adj(VelDev V, int integ, double dt,
                                                            const double* __restrict__ xs, int64_t xs_row, int64_t nv,
                                                            const double* __restrict__ yb, int64_t yb_row,
                                                            LevelBars LB, double* __restrict__ xb, int64_t xb_row,
                                                            ScatterWork swk, int rec, int tri_keys) {
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // the last quad may reach past the end: clamp (its lanes then repeat a
  // valid point and add nothing)
  const bool live = i0 < 4 * nv;
  const int64_t i = live ? i0 : 4 * nv - 1;
  const int64_t v = i / 4;
  const int k = (int)(i - v * 4);
  const double* src = xs + (int64_t)k * nv + v;
  const d3 xn = d3{src[0], src[xs_row], src[2 * xs_row]};
  const double* bsrc = yb + (int64_t)k * nv + v;
  d3 ybar = live ? d3{bsrc[0], bsrc[yb_row], bsrc[2 * yb_row]} : d3{0, 0, 0};
  const LevelBars& lb = LB;
  const ScatterWork& sw = swk;
  // record slots 4 i + stage (unused stages marked empty)
  const int64_t s0 = 4 * i0;
  if (rec && live) {
    const int used = integ == CMM_INT_EULER ? 1 : (integ == CMM_INT_IMPROVED_EULER ? 2 : 4);
    for (int st = used; st < 4; ++st) scatter_empty(swk, s0 + st);
  }
  d3 xbar;
  // the stage points are located once: the adjoint of each evaluation
  // revisits its triangle (t0 .. t3)
  int64_t t0, t1, t2, t3;
  if (integ == CMM_INT_EULER) {
    const d3 v1 = vel_eval_t(V, 0, xn, t0);
    const d3 z = sub(xn, scl(dt, v1));
    const d3 zb = unit_adj(z, ybar);
    xbar = add(zb, vel_adj(V, 0, xn, scl(-dt, zb), lb, sw, rec != 0, live ? s0 : -1, tri_keys != 0, t0));
  } else if (integ == CMM_INT_IMPROVED_EULER) {
    const d3 v1 = vel_eval_t(V, 0, xn, t0);
    const d3 z1 = sub(xn, scl(dt, v1));
    const d3 s1 = proj_norm(z1);
    const d3 v2 = vel_eval_t(V, 1, s1, t1);
    const d3 z = sub(xn, scl(dt / 2, add(v1, v2)));
    const d3 zb = unit_adj(z, ybar);
    const d3 v2b = scl(-dt / 2, zb);
    d3 v1b = v2b;
    const d3 s1b = vel_adj(V, 1, s1, v2b, lb, sw, rec != 0, live ? s0 + 1 : -1, tri_keys != 0, t1);
    const d3 z1b = unit_adj(z1, s1b);
    v1b = add(v1b, scl(-dt, z1b));
    xbar = add(add(zb, z1b), vel_adj(V, 0, xn, v1b, lb, sw, rec != 0, live ? s0 : -1, tri_keys != 0, t0));
  } else {
  const d3 v1 = vel_eval_t(V, 0, xn, t0);
  const d3 z1 = d3{xn.x - dt * v1.x / 2, xn.y - dt * v1.y / 2, xn.z - dt * v1.z / 2};
  const d3 s1 = proj_norm(z1);
  const d3 v2 = vel_eval_t(V, 1, s1, t1);
  const d3 z2 = sub(xn, scl(dt / 2, v2));
  const d3 s2 = proj_norm(z2);
  const d3 v3 = vel_eval_t(V, 2, s2, t2);
  const d3 z3 = sub(xn, scl(dt, v3));
  const d3 s3 = proj_norm(z3);
  const d3 v4 = vel_eval_t(V, 3, s3, t3);
  const d3 acc = d3{v1.x + 2 * v2.x + 2 * v3.x + v4.x, v1.y + 2 * v2.y + 2 * v3.y + v4.y,
                    v1.z + 2 * v2.z + 2 * v3.z + v4.z};
  const d3 z4 = sub(xn, scl(dt / 6, acc));
  const d3 ab = scl(-dt / 6, unit_adj(z4, ybar));
  d3 v1b = ab, v2b = scl(2, ab), v3b = scl(2, ab);
  const d3 v4b = ab;
  const d3 s3b = vel_adj(V, 3, s3, v4b, lb, sw, rec != 0, live ? s0 + 3 : -1, tri_keys != 0, t3);
  v3b = add(v3b, scl(-dt, unit_adj(z3, s3b)));
  const d3 s2b = vel_adj(V, 2, s2, v3b, lb, sw, rec != 0, live ? s0 + 2 : -1, tri_keys != 0, t2);
  v2b = add(v2b, scl(-dt / 2, unit_adj(z2, s2b)));
  const d3 s1b = vel_adj(V, 1, s1, v2b, lb, sw, rec != 0, live ? s0 + 1 : -1, tri_keys != 0, t1);
  const d3 z1b = unit_adj(z1, s1b);
  v1b = add(v1b, scl(-dt / 2, z1b));
  // z4 = xn - ..., z3 = xn - ..., z2 = xn - ..., z1 = xn - ... all carry xn
  const d3 z4b = unit_adj(z4, ybar), z3b = unit_adj(z3, s3b), z2b = unit_adj(z2, s2b);
  xbar = add(add(add(z4b, z3b), add(z2b, z1b)), vel_adj(V, 0, xn, v1b, lb, sw, rec != 0, live ? s0 : -1, tri_keys != 0, t0));
  }
  if (xb && live) {
    double* dst = xb + (int64_t)k * nv + v;
    dst[0] = xbar.x;
    dst[xb_row] = xbar.y;
    dst[2 * xb_row] = xbar.z;
  }
}

// Adjoint of k_sample_vorticity for a spline w0: xbar (overwrite, may be
// null) and the adjoint of the w0 coefficients (accumulate, may be null).
// The spline is located at q = x / |x| and cos theta(x) = z / |x|, as in the
// forward sampling.
__global__ __launch_bounds__(256) void k_sample_vorticity_adj(cmm_mwss_mesh W, const double* __restrict__ wpacked,
                                                              const double* __restrict__ X, int64_t n, int64_t sp,
                                                              int64_t sc, const double* __restrict__ wbar, double rot,
                                                              double* __restrict__ xbar, int64_t bsp, int64_t bsc,
                                                              double* __restrict__ pbar, ScatterWork sw, int rec) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const d3 x = load_pt(X, i, sp, sc);
  const double nq = sqrt(x.x * x.x + x.y * x.y + x.z * x.z);
  const d3 q = d3{x.x / nq, x.y / nq, x.z / nq};
  const double wb = wbar[i];
  const VLoc r = vel_locate(W, q);
  d3 gq = d3{0, 0, 0};
  if (rec) {
    // a full wave at one point (the 2L samples of a pole ring) writes one
    // record with the wave's sum: one key is otherwise counted and placed
    // 2L times per ring and call
    const int lane = (int)(threadIdx.x & 63);
    const int64_t key = r.ok ? r.tri * 6 + r.s : -1 - (int64_t)lane;
    const bool full = __ballot(1) == ~0ull;
    const int64_t k0 = __shfl(key, 0);
    const double c0x = __shfl(r.cb.x, 0), c0y = __shfl(r.cb.y, 0), c0z = __shfl(r.cb.z, 0);
    const bool same = full && __all(key == k0 && r.cb.x == c0x && r.cb.y == c0y && r.cb.z == c0z);
    double a = wb;
    if (same) {
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) a += __shfl_xor(a, o);
    }
    if (r.ok && (!same || lane == 0))
      scatter_put(sw, i, (uint32_t)(r.tri * 6 + r.s), d3{a, 0.0, 0.0}, r.cb, 0);
    else
      scatter_empty(sw, i);
  }
  if (pbar && !rec) {
    // atomics; a full wave on one sub-triangle (the 2L samples of a pole
    // ring are one point) is summed first and added once
    double bas[6];
    bb6_basis(r.cb, bas);
    const int64_t key = r.ok ? r.tri * 6 + r.s : -1 - (int64_t)(threadIdx.x & 63);
    const bool full = __ballot(1) == ~0ull;
    const int64_t k0 = __shfl(key, 0);
    const bool uniform = full && __all(key == k0) && r.ok;
    double v[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) v[j] = r.ok ? wb * bas[j] : 0.0;
    if (uniform) {
#pragma unroll
      for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) v[j] += __shfl_xor(v[j], o);
    }
    if (r.ok && (!uniform || (threadIdx.x & 63) == 0)) {
      double* dst = pbar + (r.tri * 6 + r.s) * 6;
      for (int j = 0; j < 6; ++j) unsafeAtomicAdd(dst + j, v[j]);
    }
  }
  if (r.ok) {
    const double* p = wpacked + (r.tri * 6 + r.s) * 6;
    const d3 bb = sub_bary_T(r.s, bb6_db(r.cb, p), r.se, r.se, r.se, r.s4);
    gq = add(add(scl(bb.x, r.w[0]), scl(bb.y, r.w[1])), scl(bb.z, r.w[2]));
  }
  if (!xbar) return;
  const double nx = norm3(x), inx = 1 / nx;
  // d(x / |x|)^T wb gq + wb 2 rot d(z / |x|)
  d3 xb = unit_adj(x, scl(wb, gq));
  const d3 u = scl(inx, x);
  const double cz_ = 2 * rot * wb * inx;
  xb = add(xb, d3{-cz_ * u.z * u.x, -cz_ * u.z * u.y, cz_ * (1 - u.z * u.z)});
  store_pt(xbar, i, bsp, bsc, xb);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int cmm_mwss_locate(const cmm_mwss_mesh* m, const double* q, int64_t n, int64_t sp,
                               int64_t sc, int32_t* tri, double* bcc, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_mwss_locate: no batched form");
  if (!m || n < 0 || (n > 0 && (!q || !tri))) return fail(CMM_E_ARG, "cmm_mwss_locate: bad arguments");
  if (n == 0) return CMM_OK;
  hipLaunchKernelGGL(k_mwss_locate, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), *m, q, n,
                     sp, sc, tri, bcc);
  CMM_LAUNCH_CHECK("cmm_mwss_locate");
  return CMM_OK;
}

extern "C" int cmm_mwss_split_locate(const cmm_mwss_mesh* m, const double* q, int64_t n, int64_t sp, int64_t sc,
                                     const int32_t* known, int32_t* tri, int32_t* sub, double* cb, double* w,
                                     void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_mwss_split_locate: no batched form");
  if (!m || n < 0 || (n > 0 && (!q || !tri))) return fail(CMM_E_ARG, "cmm_mwss_split_locate: bad arguments");
  if (n == 0) return CMM_OK;
  hipLaunchKernelGGL(k_mwss_split_locate, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), *m, q, n, sp,
                     sc, known, tri, sub, cb, w);
  CMM_LAUNCH_CHECK("cmm_mwss_split_locate");
  return CMM_OK;
}

extern "C" int cmm_mwss_assemble(const cmm_mwss_mesh* m, const double* vals, int ncomp,
                                 double* coeffs, double* packed, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_mwss_assemble: no batched form");
  if (!m || !vals || (ncomp != 1 && ncomp != 3)) return fail(CMM_E_ARG, "cmm_mwss_assemble: bad arguments");
  if (!coeffs && !packed) return CMM_OK;
  if (ncomp == 3)
    hipLaunchKernelGGL(k_mwss_assemble<3>, dim3(grid_for(m->Tu, 256)), dim3(256), 0, as_stream(stream), *m,
                       vals, coeffs, packed);
  else
    hipLaunchKernelGGL(k_mwss_assemble<1>, dim3(grid_for(m->Tu, 256)), dim3(256), 0, as_stream(stream), *m,
                       vals, coeffs, packed);
  CMM_LAUNCH_CHECK("cmm_mwss_assemble");
  return CMM_OK;
}

extern "C" int cmm_mwss_assemble_sparse(const cmm_mwss_mesh* m, const double* vals, double* packed, uint8_t* vbits,
                                        int slot, int inherit, uint8_t* mark, void* stream) {
  if (!m || !vals || !packed || !vbits || slot < 0 || slot > 7 || inherit < -1 || inherit > 7 || inherit == slot)
    return fail(CMM_E_ARG, "cmm_mwss_assemble_sparse: bad arguments");
  // the rows of triangles next to either pole: a cap of N, then 2N a row
  const int64_t npolar = std::min<int64_t>(m->N + (int64_t)2 * m->N * (kPolarRows - 1), m->Tu / 2);
  const int64_t pc = 2 * grid_for(npolar, 64);
  const int64_t inner = grid_for(m->Tu - 2 * npolar, 256);
  return launch<KMwssAssembleSparse>(k_mwss_assemble_sparse, (unsigned)(pc + inner), 64, 0, as_stream(stream),
                                     KMwssAssembleSparse::Args{*m, vals, packed, vbits, slot, inherit, mark, npolar, pc},
                                     "cmm_mwss_assemble_sparse");
}

extern "C" int cmm_mwss_eval(const cmm_mwss_mesh* m, const double* const* levels, const double* w,
                             int nlev, int ncomp, const double* q, int64_t n, int64_t sp, int64_t sc,
                             double* out, int64_t osp, int64_t osc, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_mwss_eval: no batched form");
  if (!m || !levels || !w || nlev < 1 || nlev > 3 || (ncomp != 1 && ncomp != 3) || n < 0 ||
      (n > 0 && (!q || !out)))
    return fail(CMM_E_ARG, "cmm_mwss_eval: bad arguments");
  LevelSet ls;
  ls.nlev = nlev;
  for (int k = 0; k < nlev; ++k) {
    if (!levels[k]) return fail(CMM_E_ARG, "cmm_mwss_eval: null level");
    ls.p[k] = levels[k];
    ls.w[k] = w[k];
  }
  if (n == 0) return CMM_OK;
  hipLaunchKernelGGL(k_mwss_eval, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), *m, ls, ncomp,
                     q, n, sp, sc, out, osp, osc);
  CMM_LAUNCH_CHECK("cmm_mwss_eval");
  return CMM_OK;
}

static int velocity_device(const cmm_velocity* vel, const char* who, VelDev& V) {
  V.kind = vel->kind;
  for (int s = 0; s < 4; ++s) V.stage_t[s] = vel->stage_t[s];
  for (int k = 0; k < 8; ++k) V.params[k] = vel->params[k];
  if (vel->kind == CMM_VEL_SPLINE) {
    if (!vel->mesh || vel->nlev < 1 || vel->nlev > 3) return fail(CMM_E_ARG, std::string(who) + ": bad spline velocity");
    V.mesh = *vel->mesh;
    for (int s = 0; s < 4; ++s) {
      V.ls[s].nlev = vel->nlev;
      for (int k = 0; k < 3; ++k) {
        V.ls[s].p[k] = k < vel->nlev ? vel->levels[k] : nullptr;
        V.ls[s].w[k] = vel->stage_w[s][k];
      }
      for (int k = 0; k < vel->nlev; ++k)
        if (!vel->levels[k]) return fail(CMM_E_ARG, std::string(who) + ": null velocity level");
    }
    V.sv = SparseLv{nullptr, {0, 0, 0}, {nullptr, nullptr, nullptr}, nullptr, 1};
    if (vel->vbits) {
      V.sv.vbits = vel->vbits;
      for (int k = 0; k < vel->nlev; ++k) {
        if (!vel->jets[k] || vel->vslot[k] < 0 || vel->vslot[k] > 7)
          return fail(CMM_E_ARG, std::string(who) + ": a sparse velocity needs jets and a bit for every level");
        V.sv.slot[k] = vel->vslot[k];
        V.sv.jets[k] = vel->jets[k];
      }
      for (int k = vel->nlev; k < 3; ++k) {  // never read: point at level 0
        V.sv.slot[k] = V.sv.slot[0];
        V.sv.jets[k] = vel->jets[0];
      }
      V.sv.mark = vel->mark;
      V.sv.mark_every = vel->mark_every;
    }
  } else if (vel->kind < 0 || vel->kind > CMM_VEL_DIV) {
    return fail(CMM_E_ARG, std::string(who) + ": unknown velocity kind");
  }
  return CMM_OK;
}

extern "C" int cmm_advect_points(const cmm_velocity* vel, int integ, double dt, const double* x,
                                 int64_t m, double* y, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_advect_points: no batched form");
  if (!vel || m < 0 || (m > 0 && (!x || !y)) || integ < 0 || integ > 4)
    return fail(CMM_E_ARG, "cmm_advect_points: bad arguments");
  VelDev V;
  int rc = velocity_device(vel, "cmm_advect_points", V);
  if (rc) return rc;
  if (m == 0) return CMM_OK;
  auto kern = V.kind != CMM_VEL_SPLINE ? k_advect<kVelAnalytic>
              : (V.sv.vbits ? k_advect<kVelSparse> : k_advect<kVelSpline>);
  hipLaunchKernelGGL(kern, dim3(grid_for(m, 128)), dim3(128), 0, as_stream(stream), V, integ, dt, x, m, y);
  CMM_LAUNCH_CHECK("cmm_advect_points");
  return CMM_OK;
}

extern "C" int cmm_advect_stencil(const cmm_velocity* vel, int integ, double dt, const double* verts,
                                  int64_t verts_row, const double* st, int64_t st_row, int64_t nv, double* yv,
                                  int64_t yv_row, double* yst, int64_t yst_row, void* stream) {
  if (!vel || nv < 0 || (nv > 0 && (!st || !yst)) || (verts && !yv) || integ < 0 || integ > 4 ||
      st_row < 4 * nv || yst_row < 4 * nv || (verts && (verts_row < nv || yv_row < nv)))
    return fail(CMM_E_ARG, "cmm_advect_stencil: bad arguments");
  VelDev V;
  int rc = velocity_device(vel, "cmm_advect_stencil", V);
  if (rc) return rc;
  if (nv == 0) return CMM_OK;
  const int vk = V.kind != CMM_VEL_SPLINE ? kVelAnalytic : (V.sv.vbits ? kVelSparse : kVelSpline);
  const int P = verts ? 5 : 4;
  const unsigned blocks = (unsigned)grid_for(P * nv, 128);
  hipStream_t st_ = as_stream(stream);
  const char* what = "cmm_advect_stencil";
#define CMM_ADVECT_LAUNCH(PP, VKK)                                                                              \
  return launch<KAdvectStencil<PP, VKK>>(k_advect_stencil<PP, VKK>, blocks, 128, 0, st_,                        \
                                         KAdvectStencil<PP, VKK>::Args{V, integ, dt, verts, verts_row, st, st_row, \
                                                                       nv, yv, yv_row, yst, yst_row},             \
                                         what)
  if (verts) {
    if (vk == kVelAnalytic) CMM_ADVECT_LAUNCH(5, kVelAnalytic);
    if (vk == kVelSpline) CMM_ADVECT_LAUNCH(5, kVelSpline);
    CMM_ADVECT_LAUNCH(5, kVelSparse);
  }
  if (vk == kVelAnalytic) CMM_ADVECT_LAUNCH(4, kVelAnalytic);
  if (vk == kVelSpline) CMM_ADVECT_LAUNCH(4, kVelSpline);
  CMM_ADVECT_LAUNCH(4, kVelSparse);
#undef CMM_ADVECT_LAUNCH
}

extern "C" int cmm_mwss_eval_grad(const cmm_mwss_mesh* m, const double* packed, const double* q, int64_t n,
                                  int64_t sp, int64_t sc, double* out, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_mwss_eval_grad: no batched form");
  if (!m || !packed || n < 0 || (n > 0 && (!q || !out))) return fail(CMM_E_ARG, "cmm_mwss_eval_grad: bad arguments");
  if (n == 0) return CMM_OK;
  hipLaunchKernelGGL(k_mwss_eval_grad, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), *m, packed, q, n,
                     sp, sc, out);
  CMM_LAUNCH_CHECK("cmm_mwss_eval_grad");
  return CMM_OK;
}

extern "C" int cmm_mwss_velocity_jets(const cmm_mwss_mesh* m, const double* fields, double* vals,
                                      void* stream) {
  if (!m || !fields || !vals) return fail(CMM_E_ARG, "cmm_mwss_velocity_jets: bad arguments");
  return launch<KVelocityJets>(k_velocity_jets, (unsigned)grid_for(m->Nu, 256), 256, 0, as_stream(stream),
                               KVelocityJets::Args{*m, fields, vals}, "cmm_mwss_velocity_jets");
}

extern "C" int cmm_mwss_scalar_jets(const cmm_mwss_mesh* m, const double* fields, double* vals,
                                    void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_mwss_scalar_jets: no batched form");
  if (!m || !fields || !vals) return fail(CMM_E_ARG, "cmm_mwss_scalar_jets: bad arguments");
  hipLaunchKernelGGL(k_scalar_jets, dim3(grid_for(m->Nu, 256)), dim3(256), 0, as_stream(stream), *m,
                     fields, vals);
  CMM_LAUNCH_CHECK("cmm_mwss_scalar_jets");
  return CMM_OK;
}

extern "C" int cmm_sample_vorticity(int kind, const cmm_mwss_mesh* wmesh, const double* wpacked,
                                    const double* params, const double* X, int64_t n, int64_t sp,
                                    int64_t sc, const double* cosg, int ring_len, double rot_rate,
                                    double* out, void* stream) {
  if (kind < 0 || kind > CMM_VORT_SH_SERIES || n < 0 || (n > 0 && (!X || !out || !cosg)) || ring_len < 1)
    return fail(CMM_E_ARG, "cmm_sample_vorticity: bad arguments");
  if (kind == CMM_VORT_SPLINE && (!wmesh || !wpacked))
    return fail(CMM_E_ARG, "cmm_sample_vorticity: spline vorticity needs a mesh and coefficients");
  if (kind == CMM_VORT_SH_SERIES && (!wpacked || !params || !(params[0] >= 1 && params[0] <= 1024) ||
                                     params[0] != (double)(int)params[0]))
    return fail(CMM_E_ARG, "cmm_sample_vorticity: a spherical-harmonic series needs its table and 1 <= Ls <= 1024");
  VortArgs A;
  A.kind = kind;
  for (int k = 0; k < 4; ++k) A.params[k] = params ? params[k] : 0.0;
  A.rot = rot_rate;
  A.ring_len = ring_len;
  cmm_mwss_mesh W;
  if (wmesh) W = *wmesh;
  if (n == 0) return CMM_OK;
  return launch<KSampleVorticity>(k_sample_vorticity, (unsigned)grid_for(n, 256), 256, 0, as_stream(stream),
                                  KSampleVorticity::Args{A, W, wpacked, X, n, sp, sc, cosg, out},
                                  "cmm_sample_vorticity");
}

// ---- adjoints -------------------------------------------------------------
extern "C" int cmm_mwss_velocity_jets_adjoint(const cmm_mwss_mesh* m, const double* vals_bar, double* fields_bar,
                                              void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_mwss_velocity_jets_adjoint: no batched form");
  if (!m || !vals_bar || !fields_bar) return fail(CMM_E_ARG, "cmm_mwss_velocity_jets_adjoint: bad arguments");
  const int64_t G = (int64_t)(m->L + 1) * m->N;
  hipLaunchKernelGGL(k_grid_jets_adj<12>, dim3(grid_for(G, 256)), dim3(256), 0, as_stream(stream), *m, vals_bar,
                     fields_bar);
  CMM_LAUNCH_CHECK("cmm_mwss_velocity_jets_adjoint");
  return CMM_OK;
}

extern "C" int cmm_mwss_scalar_jets_adjoint(const cmm_mwss_mesh* m, const double* vals_bar, double* fields_bar,
                                            void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_mwss_scalar_jets_adjoint: no batched form");
  if (!m || !vals_bar || !fields_bar) return fail(CMM_E_ARG, "cmm_mwss_scalar_jets_adjoint: bad arguments");
  const int64_t G = (int64_t)(m->L + 1) * m->N;
  hipLaunchKernelGGL(k_grid_jets_adj<4>, dim3(grid_for(G, 256)), dim3(256), 0, as_stream(stream), *m, vals_bar,
                     fields_bar);
  CMM_LAUNCH_CHECK("cmm_mwss_scalar_jets_adjoint");
  return CMM_OK;
}

extern "C" int cmm_mwss_assemble_adjoint(const cmm_mwss_mesh* m, const double* packed_bar, int ncomp,
                                         double* vals_bar, void* work, int64_t work_bytes, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_mwss_assemble_adjoint: no batched form");
  if (!m || !packed_bar || !vals_bar || (ncomp != 1 && ncomp != 3))
    return fail(CMM_E_ARG, "cmm_mwss_assemble_adjoint: bad arguments");
  if (work) {
    if (work_bytes < 96 * ncomp * m->Tu) return fail(CMM_E_ARG, "cmm_mwss_assemble_adjoint: workspace too small");
    double* tmp = static_cast<double*>(work);
    hipStream_t st = as_stream(stream);
    const int64_t ni = m->Nu - 2;
    if (ncomp == 3) {
      hipLaunchKernelGGL(k_mwss_assemble_adj_corners<3>, dim3(grid_for(m->Tu, 64)), dim3(192), 0, st, *m, packed_bar, tmp);
      hipLaunchKernelGGL(k_mwss_assemble_adj_gather<3>, dim3(grid_for(ni, 256) + 2), dim3(256), 0, st, *m,
                         CornerSrc<3>{tmp, nullptr, m->Tu}, vals_bar);
    } else {
      hipLaunchKernelGGL(k_mwss_assemble_adj_corners<1>, dim3(grid_for(m->Tu, 64)), dim3(64), 0, st, *m, packed_bar, tmp);
      hipLaunchKernelGGL(k_mwss_assemble_adj_gather<1>, dim3(grid_for(ni, 256) + 2), dim3(256), 0, st, *m,
                         CornerSrc<1>{tmp, nullptr, m->Tu}, vals_bar);
    }
    CMM_LAUNCH_CHECK("cmm_mwss_assemble_adjoint");
    return CMM_OK;
  }
  if (ncomp == 3)
    hipLaunchKernelGGL(k_mwss_assemble_adj<3>, dim3(grid_for(m->Tu, 256)), dim3(256), 0, as_stream(stream), *m,
                       packed_bar, vals_bar);
  else
    hipLaunchKernelGGL(k_mwss_assemble_adj<1>, dim3(grid_for(m->Tu, 256)), dim3(256), 0, as_stream(stream), *m,
                       packed_bar, vals_bar);
  CMM_LAUNCH_CHECK("cmm_mwss_assemble_adjoint");
  return CMM_OK;
}

extern "C" int cmm_advect_stencil_adjoint(const cmm_velocity* vel, int integ, double dt, const double* st,
                                          int64_t st_row, int64_t nv, const double* yst_bar, int64_t yb_row,
                                          double* const* level_bars, double* st_bar, int64_t sb_row, void* work,
                                          int64_t work_bytes, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_advect_stencil_adjoint: no batched form");
  if (!vel || !st || !yst_bar || !level_bars || nv < 1 || st_row < 4 * nv || yb_row < 4 * nv ||
      (st_bar && sb_row < 4 * nv) ||
      (integ != CMM_INT_EULER && integ != CMM_INT_IMPROVED_EULER && integ != CMM_INT_RK4))
    return fail(CMM_E_ARG, "cmm_advect_stencil_adjoint: bad arguments");
  if (vel->kind != CMM_VEL_SPLINE) return fail(CMM_E_ARG, "cmm_advect_stencil_adjoint: needs a spline velocity");
  if (vel->vbits)
    return fail(CMM_E_ARG, "cmm_advect_stencil_adjoint: sparse velocity levels (complete them first)");
  VelDev V;
  int rc = velocity_device(vel, "cmm_advect_stencil_adjoint", V);
  if (rc) return rc;
  LevelBars LB;
  for (int k = 0; k < 3; ++k) {
    LB.p[k] = k < vel->nlev ? level_bars[k] : nullptr;
    if (k < vel->nlev && !LB.p[k]) return fail(CMM_E_ARG, "cmm_advect_stencil_adjoint: null level adjoint");
  }
  ScatterWork sw;
  const int64_t nrec = 16 * nv;  // 4 stages per stencil foot
  if (work) {
    rc = scatter_layout(work, work_bytes, nrec, vel->mesh->Tu * 6, sw);
    if (rc) return rc;
  }
  hipStream_t s = as_stream(stream);
  if (work) {
    rc = scatter_begin(sw, s);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_advect_stencil_adj, dim3(grid_for(4 * nv, 128)), dim3(128), 0, s, V, integ, dt, st, st_row, nv,
                     yst_bar, yb_row, LB, st_bar, sb_row, sw, work ? 1 : 0, 0);
  CMM_LAUNCH_CHECK("cmm_advect_stencil_adjoint");
  if (work) {
    ScatterTargets t;
    t.mode = 2;
    t.nc = 3;
    t.nout = vel->nlev;
    for (int k = 0; k < vel->nlev; ++k) {
      t.T[k] = level_bars[k];
      for (int g = 0; g < 4; ++g) t.W[g][k] = vel->stage_w[g][k];
    }
    return scatter_reduce(sw, nrec, t, s);
  }
  return CMM_OK;
}

extern "C" int cmm_advect_stencil_adjoint_corners(const cmm_velocity* vel, int integ, double dt, const double* st,
                                                  int64_t st_row, int64_t nv, const double* yst_bar, int64_t yb_row,
                                                  double* const* level_corners, uint8_t* const* level_touched,
                                                  double* st_bar, int64_t sb_row, void* work, int64_t work_bytes,
                                                  void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_advect_stencil_adjoint_corners: no batched form");
  if (!vel || !st || !yst_bar || !level_corners || !level_touched || !work || nv < 1 || st_row < 4 * nv ||
      yb_row < 4 * nv || (st_bar && sb_row < 4 * nv) ||
      (integ != CMM_INT_EULER && integ != CMM_INT_IMPROVED_EULER && integ != CMM_INT_RK4))
    return fail(CMM_E_ARG, "cmm_advect_stencil_adjoint_corners: bad arguments");
  if (vel->kind != CMM_VEL_SPLINE)
    return fail(CMM_E_ARG, "cmm_advect_stencil_adjoint_corners: needs a spline velocity");
  if (vel->vbits)
    return fail(CMM_E_ARG, "cmm_advect_stencil_adjoint_corners: sparse velocity levels (complete them first)");
  VelDev V;
  int rc = velocity_device(vel, "cmm_advect_stencil_adjoint_corners", V);
  if (rc) return rc;
  CornerBars out;
  out.nout = vel->nlev;
  for (int k = 0; k < 3; ++k) {
    out.c[k] = k < vel->nlev ? level_corners[k] : nullptr;
    out.touched[k] = k < vel->nlev ? level_touched[k] : nullptr;
    if (k < vel->nlev && (!out.c[k] || !out.touched[k]))
      return fail(CMM_E_ARG, "cmm_advect_stencil_adjoint_corners: null level adjoint");
    for (int g = 0; g < 4; ++g) out.W[g][k] = k < vel->nlev ? vel->stage_w[g][k] : 0.0;
  }
  LevelBars LB;
  for (int k = 0; k < 3; ++k) LB.p[k] = nullptr;
  ScatterWork sw;
  const int64_t nrec = 16 * nv;  // 4 stages per stencil foot
  if (vel->mesh->Tu > (int64_t)0x7fffffff) return fail(CMM_E_RANGE, "cmm_advect_stencil_adjoint_corners: mesh too large");
  rc = scatter_layout(work, work_bytes, nrec, vel->mesh->Tu, sw);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  rc = scatter_begin(sw, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_advect_stencil_adj, dim3(grid_for(4 * nv, 128)), dim3(128), 0, s, V, integ, dt, st, st_row, nv,
                     yst_bar, yb_row, LB, st_bar, sb_row, sw, 1, 1);
  CMM_LAUNCH_CHECK("cmm_advect_stencil_adjoint_corners");
  rc = scatter_group(sw, nrec, s, kCornerSplit);
  if (rc) return rc;
  // at most min(records, Tu) short and min(records / (split + 1), Tu) long
  // groups
  // fixed grids that stride over the lists (their lengths are on the
  // device): up to 8 blocks of 128 / 2 of 256 threads per CU
  const int64_t Tu = vel->mesh->Tu;
  const int64_t nshort = grid_for(std::min(nrec, Tu), 128) * vel->nlev;
  hipLaunchKernelGGL(k_mwss_corner_reduce, dim3((unsigned)std::min<int64_t>(nshort, 2048)), dim3(128), 0, s, *vel->mesh,
                     sw, out);
  const int64_t nlong = std::min(nrec / (kCornerSplit + 1), Tu) * vel->nlev;  // 16-lane segments
  if (nlong > 0)
    hipLaunchKernelGGL(k_mwss_corner_reduce_long, dim3((unsigned)std::min<int64_t>(grid_for(nlong, 16), 512)),
                       dim3(256), 0, s, *vel->mesh, sw, out);
  CMM_LAUNCH_CHECK("cmm_advect_stencil_adjoint_corners: reduce");
  return CMM_OK;
}

extern "C" int cmm_mwss_corner_jets_adjoint(const cmm_mwss_mesh* m, const double* corners, const uint8_t* touched,
                                            double* vals_bar, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_mwss_corner_jets_adjoint: no batched form");
  if (!m || !corners || !touched || !vals_bar) return fail(CMM_E_ARG, "cmm_mwss_corner_jets_adjoint: bad arguments");
  const int64_t ni = m->Nu - 2;
  hipLaunchKernelGGL(k_mwss_assemble_adj_gather<3>, dim3(grid_for(ni, 256) + 2), dim3(256), 0, as_stream(stream), *m,
                     CornerSrc<3>{corners, touched, m->Tu}, vals_bar);
  CMM_LAUNCH_CHECK("cmm_mwss_corner_jets_adjoint");
  return CMM_OK;
}

extern "C" int cmm_sample_vorticity_adjoint(const cmm_mwss_mesh* wmesh, const double* wpacked, const double* X,
                                            int64_t n, int64_t sp, int64_t sc, const double* wbar, double rot_rate,
                                            double* xbar, int64_t bsp, int64_t bsc, double* wpacked_bar, void* work,
                                            int64_t work_bytes, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_sample_vorticity_adjoint: no batched form");
  if (!wmesh || !wpacked || n < 0 || (n > 0 && (!X || !wbar)))
    return fail(CMM_E_ARG, "cmm_sample_vorticity_adjoint: bad arguments");
  if (n == 0 || (!xbar && !wpacked_bar)) return CMM_OK;
  ScatterWork sw;
  const bool rec = work && wpacked_bar;
  if (rec) {
    int rc = scatter_layout(work, work_bytes, n, wmesh->Tu * 6, sw);
    if (rc) return rc;
  }
  hipStream_t s = as_stream(stream);
  if (rec) {
    int rc = scatter_begin(sw, s);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_sample_vorticity_adj, dim3(grid_for(n, 256)), dim3(256), 0, s, *wmesh, wpacked, X, n, sp, sc,
                     wbar, rot_rate, xbar, bsp, bsc, wpacked_bar, sw, rec ? 1 : 0);
  CMM_LAUNCH_CHECK("cmm_sample_vorticity_adjoint");
  if (rec) {
    ScatterTargets t;
    t.mode = 1;
    t.nc = 1;
    t.nout = 1;
    t.T[0] = wpacked_bar;
    return scatter_reduce(sw, n, t, s);
  }
  return CMM_OK;
}
