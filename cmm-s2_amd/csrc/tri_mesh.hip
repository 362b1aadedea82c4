// Unstructured spherical triangulations: cube-map bucket location,
// Powell-Sabin assembly / evaluation, fused composition and the stencil ->
// jet projection of the map update.
//
// Device layout of one mesh (all resident in HBM):
//   geom   (Ntri, 16) f64, one 128-byte line per triangle:
//          w0, w1, w2 (dual basis: b_i = q . w_i are the homogeneous
//          barycentrics of utils.bary_coords), s12 = |v1+v2|, s23, s31,
//          s4 = |v1+v2+v3|, 3 pad
//   cells  (6 R R) int4 {count, c0, c1, c2}: candidates of one cube-map cell
//          inline when count <= 3, else c0 = offset into `over`
//   over   overflow candidate lists (int32)
//   verts  (Nv, 3) f64, tris (Ntri, 3) int32
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "batch.h"
#include "common.h"
#include "scatter.h"

namespace cmm {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int hip_check(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return CMM_E_HIP;
}

constexpr int kMaxDepth = 64;

struct LevelPtrs {
  const double* p[kMaxDepth];
};

}  // namespace cmm

using namespace cmm;

struct cmm_tri_mesh {
  int64_t nv = 0, ntri = 0;
  int R = 0;
  int64_t total_cand = 0, max_cand = 0, bytes = 0;
  double* geom = nullptr;
  int4* cells = nullptr;
  int32_t* over = nullptr;
  double* verts = nullptr;
  int32_t* tris = nullptr;
  double* psgeom = nullptr;  // (kPSGeomLen, Ntri) Powell-Sabin geometry, structure of arrays
  int32_t* vstart = nullptr; // (Nv + 1) incidence lists: vertex -> 3 t + corner
  int32_t* vinc = nullptr;   // (3 Ntri)
  int device = 0;
};

extern "C" int cmm_abi_version(void) { return CMM_ABI_VERSION; }
extern "C" const char* cmm_last_error(void) { return cmm::g_err.c_str(); }

// ---------------------------------------------------------------------------
// cube map (host and device agree bit for bit on the cell of a point)
// ---------------------------------------------------------------------------
__host__ __device__ inline int cube_cell(d3 q, int R, bool* ok) {
  double ax = fabs(q.x), ay = fabs(q.y), az = fabs(q.z);
  int k = 0;
  double m = ax;
  if (ay > m) { k = 1; m = ay; }
  if (az > m) { k = 2; m = az; }
  double c[3] = {q.x, q.y, q.z};
  int sg = c[k] >= 0 ? 0 : 1;
  double u = c[(k + 1) % 3] / m, v = c[(k + 2) % 3] / m;
  *ok = (m > 0) && isfinite(u) && isfinite(v);
  if (!*ok) return 0;
  double fu = floor((u + 1.0) * 0.5 * R), fv = floor((v + 1.0) * 0.5 * R);
  int i = (int)fmin(fmax(fu, 0.0), (double)(R - 1));
  int j = (int)fmin(fmax(fv, 0.0), (double)(R - 1));
  return ((2 * k + sg) * R + i) * R + j;
}

namespace {

struct Tri2 {
  double u[3], v[3];
};

// separating-axis overlap test of a 2-D triangle and an axis-aligned box
bool tri_box_overlap(const Tri2& t, double x0, double x1, double y0, double y1) {
  for (int e = 0; e < 3; ++e) {
    int a = e, b = (e + 1) % 3, c = (e + 2) % 3;
    double nx = -(t.v[b] - t.v[a]), ny = t.u[b] - t.u[a];
    double sc = nx * (t.u[c] - t.u[a]) + ny * (t.v[c] - t.v[a]);
    if (sc < 0) { nx = -nx; ny = -ny; }
    // triangle lies on the side n.(p - a) >= 0; box fully on the other side?
    double bx = nx >= 0 ? x1 : x0, by = ny >= 0 ? y1 : y0;
    if (nx * (bx - t.u[a]) + ny * (by - t.v[a]) < 0) return false;
  }
  return true;
}

void raster_piece(const d3 p[3], int32_t tri, int R, std::vector<std::pair<int32_t, int32_t>>& out) {
  const double eps = 1e-9;
  for (int k = 0; k < 3; ++k) {
    for (int sg = 0; sg < 2; ++sg) {
      double s = sg == 0 ? 1.0 : -1.0;
      bool inside = true;
      Tri2 t;
      for (int i = 0; i < 3; ++i) {
        double c[3] = {p[i].x, p[i].y, p[i].z};
        double m = s * c[k];
        if (m < 1e-3) { inside = false; break; }
        t.u[i] = c[(k + 1) % 3] / m;
        t.v[i] = c[(k + 2) % 3] / m;
      }
      if (!inside) continue;
      double umin = std::min({t.u[0], t.u[1], t.u[2]}) - eps;
      double umax = std::max({t.u[0], t.u[1], t.u[2]}) + eps;
      double vmin = std::min({t.v[0], t.v[1], t.v[2]}) - eps;
      double vmax = std::max({t.v[0], t.v[1], t.v[2]}) + eps;
      if (umax < -1 || umin > 1 || vmax < -1 || vmin > 1) continue;
      int i0 = std::max(0, (int)std::floor((umin + 1) * 0.5 * R));
      int i1 = std::min(R - 1, (int)std::floor((umax + 1) * 0.5 * R));
      int j0 = std::max(0, (int)std::floor((vmin + 1) * 0.5 * R));
      int j1 = std::min(R - 1, (int)std::floor((vmax + 1) * 0.5 * R));
      for (int i = i0; i <= i1; ++i) {
        double x0 = -1 + 2.0 * i / R - eps, x1 = -1 + 2.0 * (i + 1) / R + eps;
        for (int j = j0; j <= j1; ++j) {
          double y0 = -1 + 2.0 * j / R - eps, y1 = -1 + 2.0 * (j + 1) / R + eps;
          if (tri_box_overlap(t, x0, x1, y0, y1))
            out.push_back({(int32_t)(((2 * k + sg) * R + i) * R + j), tri});
        }
      }
    }
  }
}

// split a spherical triangle into geodesic pieces of at most ~20 degrees
void raster_tri(d3 a, d3 b, d3 c, int32_t tri, int R, std::vector<std::pair<int32_t, int32_t>>& out,
                int depth = 0) {
  const double cmax = std::cos(20.0 * M_PI / 180.0);
  if (depth < 8 && (dot(a, b) < cmax || dot(b, c) < cmax || dot(c, a) < cmax)) {
    d3 ab = unit(add(a, b)), bc = unit(add(b, c)), ca = unit(add(c, a));
    raster_tri(a, ab, ca, tri, R, out, depth + 1);
    raster_tri(ab, b, bc, tri, R, out, depth + 1);
    raster_tri(ca, bc, c, tri, R, out, depth + 1);
    raster_tri(ab, bc, ca, tri, R, out, depth + 1);
    return;
  }
  d3 p[3] = {a, b, c};
  raster_piece(p, tri, R, out);
}

}  // namespace

extern "C" int cmm_tri_mesh_create(const double* vertices, int64_t nv, const int64_t* simplices,
                                   int64_t ntri, int cells_per_face, cmm_tri_mesh** out) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_tri_mesh_create: no batched form");
  if (!vertices || !simplices || !out || nv < 3 || ntri < 1)
    return fail(CMM_E_ARG, "cmm_tri_mesh_create: bad arguments");
  if (ntri > (int64_t)1 << 30 || nv > (int64_t)1 << 30)
    return fail(CMM_E_RANGE, "cmm_tri_mesh_create: mesh too large");
  for (int64_t t = 0; t < 3 * ntri; ++t)
    if (simplices[t] < 0 || simplices[t] >= nv)
      return fail(CMM_E_ARG, "cmm_tri_mesh_create: simplex index out of range");
  int R = cells_per_face;
  // default: about 36 cells per triangle (fewer cells straddle an edge, so
  // the waves of a composition run fewer candidate tests: 0.853 -> 0.801 ms
  // for the depth-5 composition of 8.4 M points on ico-7, whose 23.6 M cells
  // take 378 MB; 144 cells per triangle measured slower, 0.911 ms)
  if (R <= 0) R = 2 * std::max(2, (int)std::ceil(std::sqrt((double)ntri * 3.0)));
  if ((int64_t)6 * R * R > ((int64_t)1 << 31) - 1) return fail(CMM_E_RANGE, "cmm_tri_mesh_create: too many cells");

  // geometry records
  std::vector<double> geom((size_t)ntri * 16, 0.0);
  auto V = [&](int64_t i) { return d3{vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]}; };
  std::vector<int32_t> tris((size_t)ntri * 3);
  std::vector<std::pair<int32_t, int32_t>> pairs;
  pairs.reserve((size_t)ntri * 8);
  std::vector<double> psg((size_t)ntri * kPSGeomLen);
  for (int64_t t = 0; t < ntri; ++t) {
    d3 v1 = V(simplices[3 * t]), v2 = V(simplices[3 * t + 1]), v3 = V(simplices[3 * t + 2]);
    for (int i = 0; i < 3; ++i) tris[3 * t + i] = (int32_t)simplices[3 * t + i];
    double D = det3(v1, v2, v3);
    if (!(D > 0)) return fail(CMM_E_ARG, "cmm_tri_mesh_create: triangle not counter-clockwise");
    d3 w0 = cross(v2, v3), w1 = cross(v3, v1), w2 = cross(v1, v2);
    double* g = &geom[16 * t];
    g[0] = w0.x / D; g[1] = w0.y / D; g[2] = w0.z / D;
    g[3] = w1.x / D; g[4] = w1.y / D; g[5] = w1.z / D;
    g[6] = w2.x / D; g[7] = w2.y / D; g[8] = w2.z / D;
    g[10] = norm3(add(v1, v2));
    g[11] = norm3(add(v2, v3));
    g[12] = norm3(add(v3, v1));
    g[13] = norm3(add(add(v1, v2), v3));
    raster_tri(v1, v2, v3, (int32_t)t, R, pairs);
    PSGeom P;
    ps_geometry(v1, v2, v3, P);
    ps_pack(P, &psg[t], ntri);
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  int64_t ncell = (int64_t)6 * R * R;
  std::vector<int4> cells((size_t)ncell, int4{0, -1, -1, -1});
  std::vector<int32_t> over;
  int64_t maxc = 0;
  for (size_t a = 0; a < pairs.size();) {
    size_t b = a;
    while (b < pairs.size() && pairs[b].first == pairs[a].first) ++b;
    int64_t n = (int64_t)(b - a);
    maxc = std::max(maxc, n);
    int4& c = cells[pairs[a].first];
    c.x = (int)n;
    if (n <= 3) {
      int ids[3] = {-1, -1, -1};
      for (int64_t i = 0; i < n; ++i) ids[i] = pairs[a + i].second;
      c.y = ids[0]; c.z = ids[1]; c.w = ids[2];
    } else {
      c.y = (int)over.size();
      for (size_t i = a; i < b; ++i) over.push_back(pairs[i].second);
    }
    a = b;
  }
  for (int64_t c = 0; c < ncell; ++c)
    if (cells[c].x == 0) return fail(CMM_E_ARG, "cmm_tri_mesh_create: triangulation does not cover the sphere");
  // Order the candidates of every cell by how much of the cell they cover
  // (containment of 5 sample points: centre and the four quarter points), so
  // the first test usually succeeds for every lane of a wave.  Only the
  // order changes, which matters only for points on a shared edge (ties).
  {
    const double su[5] = {0.5, 0.25, 0.75, 0.25, 0.75}, sv[5] = {0.5, 0.25, 0.25, 0.75, 0.75};
    std::vector<int32_t> ids;
    std::vector<std::pair<int, int32_t>> score;
    for (int64_t cidx = 0; cidx < ncell; ++cidx) {
      int4& c = cells[cidx];
      if (c.x < 2) continue;
      const int n = c.x;
      ids.clear();
      if (n <= 3) {
        const int inl[3] = {c.y, c.z, c.w};
        for (int i = 0; i < n; ++i) ids.push_back(inl[i]);
      } else {
        for (int i = 0; i < n; ++i) ids.push_back(over[c.y + i]);
      }
      const int face = (int)(cidx / ((int64_t)R * R)), k = face / 2;
      const double sgn = (face % 2) ? -1.0 : 1.0;
      const int ci = (int)((cidx / R) % R), cj = (int)(cidx % R);
      score.clear();
      for (int32_t t : ids) {
        const double* g = &geom[16 * (size_t)t];
        int in = 0;
        for (int sidx = 0; sidx < 5; ++sidx) {
          double pt[3];
          pt[k] = sgn;
          pt[(k + 1) % 3] = -1.0 + 2.0 * (ci + su[sidx]) / R;
          pt[(k + 2) % 3] = -1.0 + 2.0 * (cj + sv[sidx]) / R;
          double b0 = g[0] * pt[0] + g[1] * pt[1] + g[2] * pt[2];
          double b1 = g[3] * pt[0] + g[4] * pt[1] + g[5] * pt[2];
          double b2 = g[6] * pt[0] + g[7] * pt[1] + g[8] * pt[2];
          in += (b0 >= 0 && b1 >= 0 && b2 >= 0) ? 1 : 0;
        }
        score.push_back({-in, t});
      }
      std::stable_sort(score.begin(), score.end(),
                       [](const std::pair<int, int32_t>& a, const std::pair<int, int32_t>& b) { return a.first < b.first; });
      if (n <= 3) {
        int32_t o[3] = {-1, -1, -1};
        for (int i = 0; i < n; ++i) o[i] = score[i].second;
        c.y = o[0]; c.z = o[1]; c.w = o[2];
      } else {
        for (int i = 0; i < n; ++i) over[c.y + i] = score[i].second;
      }
    }
  }
  if (over.empty()) over.push_back(0);
  // vertex -> (triangle, corner) incidence (transposed assembly)
  std::vector<int32_t> vstart((size_t)nv + 1, 0), vinc((size_t)ntri * 3);
  for (int64_t k = 0; k < ntri * 3; ++k) ++vstart[(size_t)tris[k] + 1];
  for (int64_t v = 0; v < nv; ++v) vstart[v + 1] += vstart[v];
  {
    std::vector<int32_t> fill(vstart.begin(), vstart.end() - 1);
    for (int64_t k = 0; k < ntri * 3; ++k) vinc[fill[tris[k]]++] = (int32_t)k;
  }

  cmm_tri_mesh* m = new cmm_tri_mesh();
  m->nv = nv; m->ntri = ntri; m->R = R;
  m->total_cand = (int64_t)pairs.size();
  m->max_cand = maxc;
  (void)hipGetDevice(&m->device);
  size_t bg = geom.size() * 8, bc = cells.size() * sizeof(int4), bo = over.size() * 4,
         bv = (size_t)nv * 24, bt = tris.size() * 4, bp = psg.size() * 8;
  m->bytes = (int64_t)(bg + bc + bo + bv + bt + bp);
  hipError_t e;
  if ((e = hipMalloc(&m->geom, bg)) != hipSuccess || (e = hipMalloc(&m->cells, bc)) != hipSuccess ||
      (e = hipMalloc(&m->over, bo)) != hipSuccess || (e = hipMalloc(&m->verts, bv)) != hipSuccess ||
      (e = hipMalloc(&m->tris, bt)) != hipSuccess || (e = hipMalloc(&m->psgeom, bp)) != hipSuccess ||
      (e = hipMalloc(&m->vstart, vstart.size() * 4)) != hipSuccess ||
      (e = hipMalloc(&m->vinc, vinc.size() * 4)) != hipSuccess) {
    cmm_tri_mesh_destroy(m);
    return hip_check(e, "cmm_tri_mesh_create: hipMalloc");
  }
  if ((e = hipMemcpy(m->geom, geom.data(), bg, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->cells, cells.data(), bc, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->over, over.data(), bo, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->verts, vertices, bv, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->tris, tris.data(), bt, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->psgeom, psg.data(), bp, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->vstart, vstart.data(), vstart.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->vinc, vinc.data(), vinc.size() * 4, hipMemcpyHostToDevice)) != hipSuccess) {
    cmm_tri_mesh_destroy(m);
    return hip_check(e, "cmm_tri_mesh_create: hipMemcpy");
  }
  *out = m;
  return CMM_OK;
}

extern "C" int cmm_tri_mesh_destroy(cmm_tri_mesh* m) {
  if (!m) return CMM_OK;
  (void)hipFree(m->geom);
  (void)hipFree(m->cells);
  (void)hipFree(m->over);
  (void)hipFree(m->verts);
  (void)hipFree(m->tris);
  (void)hipFree(m->psgeom);
  (void)hipFree(m->vstart);
  (void)hipFree(m->vinc);
  delete m;
  return CMM_OK;
}

extern "C" int cmm_tri_mesh_stats(const cmm_tri_mesh* m, int64_t* out) {
  if (!m || !out) return fail(CMM_E_ARG, "cmm_tri_mesh_stats: bad arguments");
  out[0] = m->nv; out[1] = m->ntri; out[2] = m->R;
  out[3] = m->total_cand; out[4] = m->max_cand; out[5] = m->bytes;
  return CMM_OK;
}

// ---------------------------------------------------------------------------
// device location
// ---------------------------------------------------------------------------
struct Located {
  int tri;
  double b0, b1, b2;
  double s12, s23, s31, s4;
  d3 w0, w1, w2;
};

// geometry line of a triangle (16 doubles): w0 w1 w2 (dual basis, doubles
// 0..8) and s12 s23 s31 s4 (doubles 10..13).  A candidate test reads the
// first 72 bytes (5 loads); the winner's norms are read once (2 loads).
__device__ inline void load_geom_test(const double* __restrict__ geom, int t, Located& L) {
  const double* g = geom + (int64_t)t * 16;
  const double2* g2 = reinterpret_cast<const double2*>(g);
  double2 a = g2[0], b = g2[1], c = g2[2], d = g2[3];
  L.w0 = d3{a.x, a.y, b.x};
  L.w1 = d3{b.y, c.x, c.y};
  L.w2 = d3{d.x, d.y, g[8]};
}
__device__ inline void load_geom_rest(const double* __restrict__ geom, Located& L) {
  const double2* g2 = reinterpret_cast<const double2*>(geom + (int64_t)L.tri * 16);
  const double2 e = g2[5], f = g2[6];
  L.s12 = e.x;
  L.s23 = e.y;
  L.s31 = f.x;
  L.s4 = f.y;
}

// containing triangle: first candidate with all barycentrics >= 0, else the
// candidate whose smallest barycentric is largest
__device__ inline bool tri_locate(const cmm_tri_mesh& M, const double* __restrict__ geom,
                                  const int4* __restrict__ cells, const int32_t* __restrict__ over,
                                  d3 q, Located& L) {
  bool ok;
  int cell = cube_cell(q, M.R, &ok);
  if (!ok) { L.tri = -1; return false; }
  int4 c = cells[cell];
  int n = c.x;
  int best = -1;
  double bestm = -INFINITY;
  Located cur;
  for (int i = 0; i < n; ++i) {
    int t = (n <= 3) ? (i == 0 ? c.y : (i == 1 ? c.z : c.w)) : over[c.y + i];
    load_geom_test(geom, t, cur);
    double b0 = dot(q, cur.w0), b1 = dot(q, cur.w1), b2 = dot(q, cur.w2);
    double mn = fmin(b0, fmin(b1, b2));
    if (mn > bestm) {
      bestm = mn;
      best = t;
      L = cur;
      L.b0 = b0; L.b1 = b1; L.b2 = b2;
    }
    if (mn >= 0) break;
  }
  L.tri = best;
  if (best < 0) return false;
  load_geom_rest(geom, L);
  return true;
}

// evaluation of coefficient block `blk` (= 6 triangle + sub-triangle) at
// sub-triangle barycentrics c (raw polynomial)
__device__ inline d3 eval_block3(const double* __restrict__ packed, int64_t blk, d3 c) {
  const double2* p = reinterpret_cast<const double2*>(packed + blk * 18);
  double buf[18];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    double2 v = p[i];
    buf[2 * i] = v.x;
    buf[2 * i + 1] = v.y;
  }
  return d3{bb6(c, buf), bb6(c, buf + 6), bb6(c, buf + 12)};
}

// evaluation of one packed level at a located point (raw polynomial)
__device__ inline d3 eval_level3(const double* __restrict__ packed, const Located& L, d3 bq,
                                 int s) {
  const d3 c = sub_bary(s, bq.x, bq.y, bq.z, L.s12, L.s23, L.s31, L.s4);
  return eval_block3(packed, (int64_t)L.tri * 6 + s, c);
}

// location of points for evaluation: coefficient block 6 t + s of the
// containing triangle t and split sub-triangle s, and the sub-triangle
// barycentrics (the values eval_level3 works with); -1 / NaN if not located
__global__ __launch_bounds__(256) void k_tri_split_locate(cmm_tri_mesh M, const double* __restrict__ q, int64_t n,
                                                          int64_t sp, int64_t sc, int32_t* __restrict__ blk,
                                                          double* __restrict__ cb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Located L;
  if (!tri_locate(M, M.geom, M.cells, M.over, load_pt(q, i, sp, sc), L)) {
    const double nan = __builtin_nan("");
    blk[i] = -1;
    cb[3 * i] = cb[3 * i + 1] = cb[3 * i + 2] = nan;
    return;
  }
  const int s = split_index(L.b0, L.b1, L.b2);
  const d3 c = sub_bary(s, L.b0, L.b1, L.b2, L.s12, L.s23, L.s31, L.s4);
  blk[i] = L.tri * 6 + s;
  cb[3 * i] = c.x;
  cb[3 * i + 1] = c.y;
  cb[3 * i + 2] = c.z;
}

// value and ambient gradients of the raw polynomial components at a located
// point: grad p_c = sum_i dp_c/db_i grad b_i, where the sub-triangle
// barycentrics are linear in the macro ones (sub_bary), so their gradients
// are sub_bary applied to the columns of the macro dual basis
__device__ inline void grad_level(const double* __restrict__ packed, int ncomp, const Located& L, int s,
                                  double* val, d3* g) {
  const d3 cb = sub_bary(s, L.b0, L.b1, L.b2, L.s12, L.s23, L.s31, L.s4);
  const d3 gx = sub_bary(s, L.w0.x, L.w1.x, L.w2.x, L.s12, L.s23, L.s31, L.s4);
  const d3 gy = sub_bary(s, L.w0.y, L.w1.y, L.w2.y, L.s12, L.s23, L.s31, L.s4);
  const d3 gz = sub_bary(s, L.w0.z, L.w1.z, L.w2.z, L.s12, L.s23, L.s31, L.s4);
  const double* p = packed + ((int64_t)L.tri * 6 + s) * 6 * ncomp;
  for (int c = 0; c < ncomp; ++c) {
    val[c] = bb6(cb, p + 6 * c);
    const d3 db = bb6_db(cb, p + 6 * c);
    g[c] = d3{dot(db, gx), dot(db, gy), dot(db, gz)};
  }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tri_locate(cmm_tri_mesh M, const double* __restrict__ q, int64_t n, int64_t sp,
                             int64_t sc, int32_t* __restrict__ tri, double* __restrict__ bcc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d3 x = load_pt(q, i, sp, sc);
  Located L;
  bool ok = tri_locate(M, M.geom, M.cells, M.over, x, L);
  tri[i] = ok ? L.tri : -1;
  if (bcc) {
    double nan = __builtin_nan("");
    bcc[3 * i] = ok ? L.b0 : nan;
    bcc[3 * i + 1] = ok ? L.b1 : nan;
    bcc[3 * i + 2] = ok ? L.b2 : nan;
  }
}

__global__ __launch_bounds__(256) void k_tri_eval(cmm_tri_mesh M, const double* __restrict__ packed, int ncomp, int normalize,
                           const double* __restrict__ q, int64_t n, int64_t sp, int64_t sc,
                           double* __restrict__ out, int64_t osp, int64_t osc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d3 x = load_pt(q, i, sp, sc);
  Located L;
  double nan = __builtin_nan("");
  if (!tri_locate(M, M.geom, M.cells, M.over, x, L)) {
    for (int c = 0; c < ncomp; ++c) out[i * osp + c * osc] = nan;
    return;
  }
  int s = split_index(L.b0, L.b1, L.b2);
  d3 cb = sub_bary(s, L.b0, L.b1, L.b2, L.s12, L.s23, L.s31, L.s4);
  const double* p = packed + ((int64_t)L.tri * 6 + s) * 6 * ncomp;
  double r[3];
  for (int c = 0; c < ncomp; ++c) r[c] = bb6(cb, p + 6 * c);
  if (normalize && ncomp == 3) {
    double nn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int c = 0; c < 3; ++c) r[c] = (1 / nn) * r[c];
  }
  for (int c = 0; c < ncomp; ++c) out[i * osp + c * osc] = r[c];
}

// Ring-major grid traversal: with ncol > 0 the points form nring x ncol
// (row-major) and thread t walks them in tiles of tr x (64 / tr) (one wave
// per tile), so a wave covers a compact patch instead of a long strip of one
// ring and touches fewer triangles per level.
__device__ inline int64_t grid_point(int64_t t, int64_t nring, int64_t ncol, int tr) {
  const int tc = 64 / tr;
  const int64_t ntc = (ncol + tc - 1) / tc;
  const int64_t tile = t >> 6;
  const int w = (int)(t & 63);
  const int64_t r = (tile / ntc) * tr + w / tc, c = (tile % ntc) * tc + w % tc;
  return (r < nring && c < ncol) ? r * ncol + c : -1;
}

struct KTriCompose {
  struct Args {
    cmm_tri_mesh M;
    LevelPtrs lv;
    int depth;
    const double* q;
    int64_t n, sp, sc;
    double* out;
    int64_t osp, osc;
    double* inter;
    int64_t nring, ncol;
    int tr;
    // optional: the first level's location of every point, computed once
    // (cmm_tri_split_locate): coefficient block 6 t + s (-1: not located)
    // and sub-triangle barycentrics (n, 3)
    const int32_t* pre_blk;
    const double* pre_cb;
  };
  static constexpr int kThreads = 256;
  // 7 waves per SIMD: the kernel fits 72 VGPRs without spills (78 unbounded,
  // 6 waves) and the composition is 0.8 % faster
  static constexpr int kMinEU = 7;
  static constexpr bool kArgsInPlace = true;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

__device__ __forceinline__ void KTriCompose::run(const Args& a, int bid) {
  const cmm_tri_mesh& M = a.M;
  const LevelPtrs& lv = a.lv;
  const int depth = a.depth;
  const double* __restrict__ q = a.q;
  const int64_t n = a.n, sp = a.sp, sc = a.sc, osp = a.osp, osc = a.osc, nring = a.nring, ncol = a.ncol;
  double* __restrict__ out = a.out;
  double* __restrict__ inter = a.inter;
  const int tr = a.tr;
  int64_t i = (int64_t)bid * blockDim.x + threadIdx.x;
  if (ncol > 0) {
    i = grid_point(i, nring, ncol, tr);
    if (i < 0) return;
  }
  if (i >= n) return;
  d3 x;
  int l0 = 0;
  if (a.pre_blk && depth > 0) {
    // first level at the stored location: the values the search below
    // would produce (same block, same sub-triangle barycentrics)
    const int blk = a.pre_blk[i];
    if (blk < 0) {
      const double nan = __builtin_nan("");
      x = d3{nan, nan, nan};
    } else {
      const d3 c = d3{a.pre_cb[3 * i], a.pre_cb[3 * i + 1], a.pre_cb[3 * i + 2]};
      const d3 r = eval_block3(lv.p[0], blk, c);
      double nn = sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
      double inv = 1 / nn;
      x = d3{inv * r.x, inv * r.y, inv * r.z};
    }
    if (inter) store_pt(inter, i, 3, 1, x);
    l0 = 1;
  } else {
    x = load_pt(q, i, sp, sc);
  }
  for (int l = l0; l < depth; ++l) {
    Located L;
    if (!tri_locate(M, M.geom, M.cells, M.over, x, L)) {
      double nan = __builtin_nan("");
      x = d3{nan, nan, nan};
    } else {
      int s = split_index(L.b0, L.b1, L.b2);
      d3 r = eval_level3(lv.p[l], L, d3{L.b0, L.b1, L.b2}, s);
      double nn = sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
      double inv = 1 / nn;
      x = d3{inv * r.x, inv * r.y, inv * r.z};
    }
    if (inter) store_pt(inter + (int64_t)l * n * 3, i, 3, 1, x);
  }
  store_pt(out, i, osp, osc, x);
}

__global__ __launch_bounds__(256, 7) void k_tri_compose(KTriCompose::Args a) { KTriCompose::run(a, blockIdx.x); }

// map Jacobian (ncomp 3, out (3,3,n) [c][d]) or tangential gradient of a
// scalar spline (ncomp 1, out (3,n)); NaN where the point is not located
__global__ __launch_bounds__(256) void k_tri_eval_grad(cmm_tri_mesh M, const double* __restrict__ packed, int ncomp,
                                                       const double* __restrict__ q, int64_t n, int64_t sp, int64_t sc,
                                                       double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const d3 x = load_pt(q, i, sp, sc);
  Located L;
  const int nout = ncomp == 3 ? 9 : 3;
  if (!tri_locate(M, M.geom, M.cells, M.over, x, L)) {
    for (int k = 0; k < nout; ++k) out[(int64_t)k * n + i] = __builtin_nan("");
    return;
  }
  const int s = split_index(L.b0, L.b1, L.b2);
  double v[3];
  d3 g[3];
  grad_level(packed, ncomp, L, s, v, g);
  if (ncomp == 3) {
    double J[9];
    normalized_jacobian(d3{v[0], v[1], v[2]}, g, x, J);
    for (int k = 0; k < 9; ++k) out[(int64_t)k * n + i] = J[k];
  } else {
    const double gq = dot(g[0], x);
    out[i] = g[0].x - gq * x.x;
    out[n + i] = g[0].y - gq * x.y;
    out[2 * n + i] = g[0].z - gq * x.z;
  }
}

// composed map and its Jacobian: J <- J_level(x) J through the chain
// (levels in application order), out (n,3) positions and jac (3,3,n)
__global__ __launch_bounds__(256) void k_tri_compose_grad(cmm_tri_mesh M, LevelPtrs lv, int depth,
                                                          const double* __restrict__ q, int64_t n, int64_t sp,
                                                          int64_t sc, double* __restrict__ out, int64_t osp,
                                                          int64_t osc, double* __restrict__ jac) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d3 x = load_pt(q, i, sp, sc);
  double J[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  // project the identity on the input tangent plane
  {
    const double qv[3] = {x.x, x.y, x.z};
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) J[3 * c + d] -= qv[c] * qv[d];
  }
  for (int l = 0; l < depth; ++l) {
    Located L;
    if (!tri_locate(M, M.geom, M.cells, M.over, x, L)) {
      const double nan = __builtin_nan("");
      x = d3{nan, nan, nan};
      for (int k = 0; k < 9; ++k) J[k] = nan;
      break;
    }
    const int s = split_index(L.b0, L.b1, L.b2);
    double v[3];
    d3 g[3];
    grad_level(lv.p[l], 3, L, s, v, g);
    double Jl[9], Jn[9];
    normalized_jacobian(d3{v[0], v[1], v[2]}, g, x, Jl);
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) Jn[3 * c + d] = Jl[3 * c] * J[d] + Jl[3 * c + 1] * J[3 + d] + Jl[3 * c + 2] * J[6 + d];
    for (int k = 0; k < 9; ++k) J[k] = Jn[k];
    const double inv = 1 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    x = d3{inv * v[0], inv * v[1], inv * v[2]};
  }
  store_pt(out, i, osp, osc, x);
  for (int k = 0; k < 9; ++k) jac[(int64_t)k * n + i] = J[k];
}

template <int NC>
struct KTriAssemble {
  struct Args {
    cmm_tri_mesh M;
    const double* vals;
    double* coeffs;
    double* packed;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

template <int NC>
__device__ __forceinline__ void KTriAssemble<NC>::run(const Args& a, int bid) {
  const cmm_tri_mesh& M = a.M;
  const double* __restrict__ vals = a.vals;
  double* __restrict__ coeffs = a.coeffs;
  double* __restrict__ packed = a.packed;
  __shared__ __align__(16) double lds[64 * 6 * NC * 6];
  const int64_t t0 = (int64_t)bid * blockDim.x;
  const int64_t t = t0 + threadIdx.x;
  double cf[NC][19];
  if (t < M.ntri) {
    int vi[3] = {M.tris[3 * t], M.tris[3 * t + 1], M.tris[3 * t + 2]};
    // Powell-Sabin geometry recomputed from the three vertices (72 bytes of
    // gathers instead of the 600-byte stored table row)
    PSGeom P;
    ps_geometry(load_pt(M.verts, vi[0], 3, 1), load_pt(M.verts, vi[1], 3, 1), load_pt(M.verts, vi[2], 3, 1), P);
    const int64_t Nv = M.nv;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double f[3];
      d3 gr[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        if (NC == 1) {
          f[i] = vals[vi[i]];
          gr[i] = d3{vals[Nv + vi[i]], vals[2 * Nv + vi[i]], vals[3 * Nv + vi[i]]};
        } else {
          f[i] = vals[(c * 4) * Nv + vi[i]];
          gr[i] = d3{vals[(0 * 4 + 1 + c) * Nv + vi[i]], vals[(1 * 4 + 1 + c) * Nv + vi[i]],
                     vals[(2 * 4 + 1 + c) * Nv + vi[i]]};
        }
      }
      ps_coeffs(P, f, gr, cf[c]);
      if (coeffs)
        for (int k = 0; k < 19; ++k) coeffs[((int64_t)c * 19 + k) * M.ntri + t] = cf[c][k];
    }
  }
  if (packed) store_packed_block<NC>(cf, t0, M.ntri, packed, lds);
}

template <int NC>
__global__ __launch_bounds__(256) void k_tri_assemble(typename KTriAssemble<NC>::Args a) {
  KTriAssemble<NC>::run(a, blockIdx.x);
}

struct KTriStencilJets {
  struct Args {
    cmm_tri_mesh M;
    const double* cur;
    const double* yv;
    int64_t yrow;
    const double* sf;
    int64_t srow;
    const double* tanv;
    double eps;
    double* jets;
    double* st;
  };
  static constexpr int kThreads = 128;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

__device__ __forceinline__ void KTriStencilJets::run(const Args& a, int bid) {
  const cmm_tri_mesh& M = a.M;
  const double* __restrict__ cur = a.cur;
  const double* __restrict__ yv = a.yv;
  const double* __restrict__ sf = a.sf;
  const double* __restrict__ tanv = a.tanv;
  double* __restrict__ jets = a.jets;
  double* __restrict__ st = a.st;
  const int64_t yrow = a.yrow, srow = a.srow;
  const double eps = a.eps;
  int64_t v = (int64_t)bid * blockDim.x + threadIdx.x;
  const int64_t Nv = M.nv;
  if (v >= Nv) return;
  d3 S[4];
  for (int k = 0; k < 4; ++k) {
    int64_t col = k * Nv + v;
    S[k] = d3{sf[col], sf[srow + col], sf[2 * srow + col]};
  }
  if (cur) {
    d3 y = d3{yv[v], yv[yrow + v], yv[2 * yrow + v]};
    Located L;
    double nan = __builtin_nan("");
    if (!tri_locate(M, M.geom, M.cells, M.over, y, L)) {
      for (int k = 0; k < 4; ++k) S[k] = d3{nan, nan, nan};
    } else {
      int s = split_index(L.b0, L.b1, L.b2);
      for (int k = 0; k < 4; ++k) {
        d3 bq = d3{dot(S[k], L.w0), dot(S[k], L.w1), dot(S[k], L.w2)};
        S[k] = eval_level3(cur, L, bq, s);
      }
    }
  }
  if (st)
    for (int k = 0; k < 4; ++k) {
      st[(0 * 4 + k) * Nv + v] = S[k].x;
      st[(1 * 4 + k) * Nv + v] = S[k].y;
      st[(2 * 4 + k) * Nv + v] = S[k].z;
    }
  // spline_proj_sphere (evolution_functions.py:97-110)
  d3 m = d3{(S[0].x + S[1].x + S[2].x + S[3].x) / 4, (S[0].y + S[1].y + S[2].y + S[3].y) / 4,
            (S[0].z + S[1].z + S[2].z + S[3].z) / 4};
  m = unit(m);
  double e4 = 4 * eps;
  d3 d1 = d3{(S[1].x - S[0].x + S[3].x - S[2].x) / e4, (S[1].y - S[0].y + S[3].y - S[2].y) / e4,
             (S[1].z - S[0].z + S[3].z - S[2].z) / e4};
  d3 d2 = d3{(S[2].x - S[0].x + S[3].x - S[1].x) / e4, (S[2].y - S[0].y + S[3].y - S[1].y) / e4,
             (S[2].z - S[0].z + S[3].z - S[1].z) / e4};
  d3 a1 = d3{tanv[3 * v], tanv[3 * v + 1], tanv[3 * v + 2]};
  d3 a2 = d3{tanv[3 * (Nv + v)], tanv[3 * (Nv + v) + 1], tanv[3 * (Nv + v) + 2]};
  jets[(0 * 4 + 0) * Nv + v] = m.x;
  jets[(1 * 4 + 0) * Nv + v] = m.y;
  jets[(2 * 4 + 0) * Nv + v] = m.z;
  double dc1[3] = {d1.x, d1.y, d1.z}, dc2[3] = {d2.x, d2.y, d2.z};
  double A1[3] = {a1.x, a1.y, a1.z}, A2[3] = {a2.x, a2.y, a2.z};
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) jets[(d * 4 + 1 + c) * Nv + v] = dc1[c] * A1[d] + dc2[c] * A2[d];
}

__global__ __launch_bounds__(128) void k_tri_stencil_jets(KTriStencilJets::Args a) { KTriStencilJets::run(a, blockIdx.x); }

// ---------------------------------------------------------------------------
// adjoints (reverse mode).  Coefficient and jet adjoints accumulate with
// atomics (the caller clears them); position adjoints overwrite.
// ---------------------------------------------------------------------------
constexpr int kAdjMaxDepth = 8;

struct LevelBarPtrs {
  double* p[kAdjMaxDepth];
};

// geometry of triangle t as tri_locate leaves it, barycentrics of x
__device__ inline void located_at(const double* __restrict__ geom, int t, d3 x, Located& L) {
  load_geom_test(geom, t, L);
  L.tri = t;
  load_geom_rest(geom, L);
  L.b0 = dot(x, L.w0);
  L.b1 = dot(x, L.w1);
  L.b2 = dot(x, L.w2);
}

// adjoint of the raw evaluation r = PS(b(x)) of one packed level at a located
// point: adds rbar_c basis_j into pb (if non-null), returns xbar
__device__ inline d3 eval_level3_adj(const double* __restrict__ packed, double* pb, const Located& L, d3 cb, int s,
                                     d3 rb) {
  const double* p = packed + ((int64_t)L.tri * 6 + s) * 18;
  const double rv[3] = {rb.x, rb.y, rb.z};
  if (pb) {
    double bas[6];
    bb6_basis(cb, bas);
    double* dst = pb + ((int64_t)L.tri * 6 + s) * 18;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j = 0; j < 6; ++j) unsafeAtomicAdd(dst + 6 * c + j, rv[c] * bas[j]);
  }
  d3 cbar = d3{0, 0, 0};
#pragma unroll
  for (int c = 0; c < 3; ++c) cbar = add(cbar, scl(rv[c], bb6_db(cb, p + 6 * c)));
  const d3 bb = sub_bary_T(s, cbar, L.s12, L.s23, L.s31, L.s4);
  return add(add(scl(bb.x, L.w0), scl(bb.y, L.w1)), scl(bb.z, L.w2));
}

// the 18 coefficients of sub-triangle s of triangle t (9 16-byte loads)
__device__ inline void load_block18(const double* __restrict__ packed, int t, int s, double (&buf)[18]) {
  const double2* p = reinterpret_cast<const double2*>(packed + ((int64_t)t * 6 + s) * 18);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const double2 v = p[i];
    buf[2 * i] = v.x;
    buf[2 * i + 1] = v.y;
  }
}

// eval_level3_adj on a block already in registers (no atomics)
__device__ inline d3 eval_level3_adj_buf(const double (&buf)[18], const Located& L, d3 cb, int s, d3 rb) {
  const double rv[3] = {rb.x, rb.y, rb.z};
  d3 cbar = d3{0, 0, 0};
#pragma unroll
  for (int c = 0; c < 3; ++c) cbar = add(cbar, scl(rv[c], bb6_db(cb, buf + 6 * c)));
  const d3 bb = sub_bary_T(s, cbar, L.s12, L.s23, L.s31, L.s4);
  return add(add(scl(bb.x, L.w0), scl(bb.y, L.w1)), scl(bb.z, L.w2));
}

// adjoint of k_tri_compose: xbar of the output (n,3) -> coefficient adjoints
// of every level (null entries skipped) and, if qb, the input adjoint
// With rec (a workspace) the coefficient adjoints become records
// (slot l n + i, key l 6 Ntri + 6 tri + s) for the sorted scatter-add.
__global__ __launch_bounds__(256) void k_tri_compose_adj(cmm_tri_mesh M, LevelPtrs lv, LevelBarPtrs lb, int depth,
                                                         const double* __restrict__ q, int64_t n, int64_t sp,
                                                         int64_t sc, const double* __restrict__ ob,
                                                         double* __restrict__ qb, ScatterWork sw, int rec) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d3 xs[kAdjMaxDepth];
  int ts[kAdjMaxDepth], ss[kAdjMaxDepth];
  d3 x = load_pt(q, i, sp, sc);
  bool ok = true;
#pragma unroll
  for (int l = 0; l < kAdjMaxDepth; ++l) {
    if (l < depth) {
      xs[l] = x;
      Located L;
      if (ok && tri_locate(M, M.geom, M.cells, M.over, x, L)) {
        const int s = split_index(L.b0, L.b1, L.b2);
        ts[l] = L.tri;
        ss[l] = s;
        const d3 r = eval_level3(lv.p[l], L, d3{L.b0, L.b1, L.b2}, s);
        const double inv = 1 / sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
        x = d3{inv * r.x, inv * r.y, inv * r.z};
      } else {
        ok = false;
        ts[l] = 0;
        ss[l] = 0;
      }
    }
  }
  d3 xb = ok ? d3{ob[3 * i], ob[3 * i + 1], ob[3 * i + 2]} : d3{0, 0, 0};
  const int64_t span = M.ntri * 6;
#pragma unroll
  for (int l = kAdjMaxDepth - 1; l >= 0; --l) {
    if (l < depth) {
      if (ok) {
        Located L;
        located_at(M.geom, ts[l], xs[l], L);
        const int s = ss[l];
        const d3 cb = sub_bary(s, L.b0, L.b1, L.b2, L.s12, L.s23, L.s31, L.s4);
        // the block once for the value (unit_adj) and the position adjoint
        double buf[18];
        load_block18(lv.p[l], L.tri, s, buf);
        const d3 r = d3{bb6(cb, buf), bb6(cb, buf + 6), bb6(cb, buf + 12)};
        const d3 rb = unit_adj(r, xb);
        if (rec) {
          if (lb.p[l])
            scatter_put(sw, (int64_t)l * n + i, (uint32_t)(l * span + (int64_t)L.tri * 6 + s), rb, cb, l);
          else
            scatter_empty(sw, (int64_t)l * n + i);
          xb = eval_level3_adj_buf(buf, L, cb, s, rb);
        } else {
          xb = eval_level3_adj(lv.p[l], lb.p[l], L, cb, s, rb);
        }
      } else if (rec) {
        scatter_empty(sw, (int64_t)l * n + i);
      }
    }
  }
  if (qb) {
    qb[3 * i] = xb.x;
    qb[3 * i + 1] = xb.y;
    qb[3 * i + 2] = xb.z;
  }
}

// transpose of k_tri_assemble<3> (accumulate into the jets vb (3,4,Nv))
__global__ __launch_bounds__(256) void k_tri_assemble_adj(cmm_tri_mesh M, const double* __restrict__ pb,
                                                          double* __restrict__ vb) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M.ntri) return;
  const int vi[3] = {M.tris[3 * t], M.tris[3 * t + 1], M.tris[3 * t + 2]};
  PSGeom P;
  ps_unpack(M.psgeom + t, M.ntri, P);
  const int64_t Nv = M.nv;
  const double* src = pb + t * 108;
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    double cb[19];
#pragma unroll
    for (int k = 0; k < 19; ++k) cb[k] = 0.0;
#pragma unroll
    for (int ss = 0; ss < 6; ++ss)
#pragma unroll
      for (int j = 0; j < 6; ++j) cb[cs_table(ss, j)] += src[(ss * 3 + cc) * 6 + j];
    double fb[3] = {0.0, 0.0, 0.0};
    d3 gb[3] = {d3{0, 0, 0}, d3{0, 0, 0}, d3{0, 0, 0}};
    ps_coeffs_adj(P, cb, fb, gb);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      unsafeAtomicAdd(vb + (cc * 4) * Nv + vi[i], fb[i]);
      unsafeAtomicAdd(vb + (0 * 4 + 1 + cc) * Nv + vi[i], gb[i].x);
      unsafeAtomicAdd(vb + (1 * 4 + 1 + cc) * Nv + vi[i], gb[i].y);
      unsafeAtomicAdd(vb + (2 * 4 + 1 + cc) * Nv + vi[i], gb[i].z);
    }
  }
}

// Two-phase transpose of k_tri_assemble<3> without atomics.  Phase 1, one
// thread per triangle: the jet adjoints of its three corners, corner i row r
// at tmp[(12 i + r) Ntri + t] (rows in the jets order: c 4 + 0 = value of
// component c, d 4 + 1 + c = its derivative along axis d).  Phase 2, one
// thread per vertex: the sum over its incidence list, added into vb (the only
// writer of that vertex).  Deterministic.
__global__ __launch_bounds__(192) void k_tri_assemble_adj_corners(cmm_tri_mesh M, const double* __restrict__ pb,
                                                                  double* __restrict__ tmp) {
  // 64 triangles per block: their packed blocks (contiguous) are staged
  // through LDS with coalesced 16-byte loads, then thread (component cc,
  // triangle tt) forms that component's corner adjoints
  constexpr int W = 108;
  __shared__ double sh[64 * (W + 1)];
  const int64_t t0 = (int64_t)blockIdx.x * 64;
  const int64_t nt = M.ntri;
  const int64_t nvalid = min((int64_t)64, nt - t0) * W;
  const double2* src = reinterpret_cast<const double2*>(pb + t0 * W);
#pragma unroll 4
  for (int q = threadIdx.x; q < 32 * W; q += 192) {
    const int e = 2 * q, tt = e / W, r = e - W * tt;
    const double2 v = e < nvalid ? src[q] : double2{0.0, 0.0};
    sh[tt * (W + 1) + r] = v.x;
    sh[tt * (W + 1) + r + 1] = v.y;
  }
  __syncthreads();
  const int tt = threadIdx.x & 63, cc = threadIdx.x >> 6;
  const int64_t t = t0 + tt;
  if (t >= nt) return;
  PSGeom P;
  ps_unpack(M.psgeom + t, M.ntri, P);
  double cb[19];
#pragma unroll
  for (int k = 0; k < 19; ++k) cb[k] = 0.0;
#pragma unroll
  for (int ss = 0; ss < 6; ++ss)
#pragma unroll
    for (int j = 0; j < 6; ++j) cb[cs_table(ss, j)] += sh[tt * (W + 1) + (ss * 3 + cc) * 6 + j];
  double fb[3] = {0.0, 0.0, 0.0};
  d3 gb[3] = {d3{0, 0, 0}, d3{0, 0, 0}, d3{0, 0, 0}};
  ps_coeffs_adj(P, cb, fb, gb);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    tmp[(12 * i + cc * 4) * nt + t] = fb[i];
    tmp[(12 * i + 0 * 4 + 1 + cc) * nt + t] = gb[i].x;
    tmp[(12 * i + 1 * 4 + 1 + cc) * nt + t] = gb[i].y;
    tmp[(12 * i + 2 * 4 + 1 + cc) * nt + t] = gb[i].z;
  }
}

// thread (row r, vertex v), r-major: a wave covers 64 consecutive vertices
// of one jets row (12 x more threads than one per vertex: the gather is
// latency-bound)
__global__ __launch_bounds__(256) void k_tri_assemble_adj_gather(cmm_tri_mesh M, const double* __restrict__ tmp,
                                                                 double* __restrict__ vb) {
  const int64_t Nv = M.nv;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 12 * Nv) return;
  const int r = (int)(idx / Nv);
  const int64_t v = idx - r * Nv;
  const int64_t nt = M.ntri;
  const int b = M.vstart[v], e = M.vstart[v + 1];
  double acc = 0.0;
  for (int k0 = b; k0 < e; k0 += 8) {  // 8 incident corners per pass, their loads in flight together
    double x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      x[k] = 0.0;
      if (k0 + k < e) {
        const int tc = M.vinc[k0 + k];
        const int t = tc / 3;
        const int i = tc - 3 * t;
        x[k] = tmp[(12 * i + r) * nt + t];
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += x[k];
  }
  vb[r * Nv + v] += acc;
}

// adjoint of k_tri_stencil_jets: jet adjoints (3,4,Nv) -> stencil-feet
// adjoints sfb (3, 4 Nv, overwrite) and, with a current map, its coefficient
// adjoints curb (accumulate)
__global__ __launch_bounds__(128) void k_tri_stencil_jets_adj(cmm_tri_mesh M, const double* __restrict__ cur,
                                                              const double* __restrict__ yv, int64_t yrow,
                                                              const double* __restrict__ sf, int64_t srow,
                                                              const double* __restrict__ tanv, double eps,
                                                              const double* __restrict__ jb, double* __restrict__ sfb,
                                                              int64_t sbrow, double* __restrict__ curb, ScatterWork sw,
                                                              int rec) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t Nv = M.nv;
  if (v >= Nv) return;
  d3 S[4], Sin[4];
  for (int k = 0; k < 4; ++k) {
    const int64_t col = k * Nv + v;
    Sin[k] = d3{sf[col], sf[srow + col], sf[2 * srow + col]};
    S[k] = Sin[k];
  }
  Located L;
  int s = 0;
  bool ok = true;
  d3 cbk[4];
  if (cur) {
    const d3 y = d3{yv[v], yv[yrow + v], yv[2 * yrow + v]};
    ok = tri_locate(M, M.geom, M.cells, M.over, y, L);
    if (ok) {
      s = split_index(L.b0, L.b1, L.b2);
      for (int k = 0; k < 4; ++k) {
        const d3 bq = d3{dot(Sin[k], L.w0), dot(Sin[k], L.w1), dot(Sin[k], L.w2)};
        cbk[k] = sub_bary(s, bq.x, bq.y, bq.z, L.s12, L.s23, L.s31, L.s4);
        S[k] = eval_level3(cur, L, bq, s);
      }
    }
  }
  d3 Sb[4] = {d3{0, 0, 0}, d3{0, 0, 0}, d3{0, 0, 0}, d3{0, 0, 0}};
  if (rec && !(ok && cur)) scatter_empty(sw, v);
  if (ok) {
    const d3 mu = d3{(S[0].x + S[1].x + S[2].x + S[3].x) / 4, (S[0].y + S[1].y + S[2].y + S[3].y) / 4,
                     (S[0].z + S[1].z + S[2].z + S[3].z) / 4};
    const d3 a1 = d3{tanv[3 * v], tanv[3 * v + 1], tanv[3 * v + 2]};
    const d3 a2 = d3{tanv[3 * (Nv + v)], tanv[3 * (Nv + v) + 1], tanv[3 * (Nv + v) + 2]};
    const double A1[3] = {a1.x, a1.y, a1.z}, A2[3] = {a2.x, a2.y, a2.z};
    double d1b[3], d2b[3];
    for (int c = 0; c < 3; ++c) {
      d1b[c] = 0.0;
      d2b[c] = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double g = jb[(d * 4 + 1 + c) * Nv + v];
        d1b[c] += g * A1[d];
        d2b[c] += g * A2[d];
      }
    }
    const d3 mb = unit_adj(mu, d3{jb[(0 * 4) * Nv + v], jb[(1 * 4) * Nv + v], jb[(2 * 4) * Nv + v]});
    const double e4 = 4 * eps;
    const d3 q = scl(0.25, mb), D1 = d3{d1b[0] / e4, d1b[1] / e4, d1b[2] / e4},
             D2 = d3{d2b[0] / e4, d2b[1] / e4, d2b[2] / e4};
    Sb[0] = sub(sub(q, D1), D2);
    Sb[1] = sub(add(q, D1), D2);
    Sb[2] = add(sub(q, D1), D2);
    Sb[3] = add(add(q, D1), D2);
    if (cur) {
      double acc[18];
      for (int e = 0; e < 18; ++e) acc[e] = 0.0;
      const double* p = cur + ((int64_t)L.tri * 6 + s) * 18;
      for (int k = 0; k < 4; ++k) {
        double bas[6];
        bb6_basis(cbk[k], bas);
        const double rv[3] = {Sb[k].x, Sb[k].y, Sb[k].z};
        d3 cbar = d3{0, 0, 0};
        for (int c = 0; c < 3; ++c) {
          for (int j = 0; j < 6; ++j) acc[6 * c + j] += rv[c] * bas[j];
          cbar = add(cbar, scl(rv[c], bb6_db(cbk[k], p + 6 * c)));
        }
        const d3 bb = sub_bary_T(s, cbar, L.s12, L.s23, L.s31, L.s4);
        Sb[k] = add(add(scl(bb.x, L.w0), scl(bb.y, L.w1)), scl(bb.z, L.w2));
      }
      // the four feet share the block: one value record (or 18 atomics)
      // per vertex
      if (rec) {
        scatter_put_values(sw, v, (uint32_t)((int64_t)L.tri * 6 + s), acc, 0);
      } else if (curb) {
        double* dst = curb + ((int64_t)L.tri * 6 + s) * 18;
        for (int e = 0; e < 18; ++e) unsafeAtomicAdd(dst + e, acc[e]);
      }
    }
  }
  for (int k = 0; k < 4; ++k) {
    const int64_t col = k * Nv + v;
    sfb[col] = Sb[k].x;
    sfb[sbrow + col] = Sb[k].y;
    sfb[2 * sbrow + col] = Sb[k].z;
  }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int cmm_tri_locate(const cmm_tri_mesh* m, const double* q, int64_t n, int64_t sp,
                              int64_t sc, int32_t* tri, double* bcc, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_tri_locate: no batched form");
  if (!m || (!q && n > 0) || (!tri && n > 0) || n < 0) return fail(CMM_E_ARG, "cmm_tri_locate: bad arguments");
  if (n == 0) return CMM_OK;
  hipLaunchKernelGGL(k_tri_locate, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), *m, q, n,
                     sp, sc, tri, bcc);
  CMM_LAUNCH_CHECK("cmm_tri_locate");
  return CMM_OK;
}

extern "C" int cmm_tri_assemble(const cmm_tri_mesh* m, const double* vals, int ncomp, double* coeffs,
                                double* packed, void* stream) {
  if (!m || !vals || (ncomp != 1 && ncomp != 3)) return fail(CMM_E_ARG, "cmm_tri_assemble: bad arguments");
  if (!coeffs && !packed) return CMM_OK;
  const unsigned blocks = (unsigned)grid_for(m->ntri, 256);
  if (ncomp == 3)
    return launch<KTriAssemble<3>>(k_tri_assemble<3>, blocks, 256, 0, as_stream(stream),
                                   KTriAssemble<3>::Args{*m, vals, coeffs, packed}, "cmm_tri_assemble");
  return launch<KTriAssemble<1>>(k_tri_assemble<1>, blocks, 256, 0, as_stream(stream),
                                 KTriAssemble<1>::Args{*m, vals, coeffs, packed}, "cmm_tri_assemble");
}

extern "C" int cmm_tri_eval(const cmm_tri_mesh* m, const double* packed, int ncomp, int normalize,
                            const double* q, int64_t n, int64_t sp, int64_t sc, double* out,
                            int64_t osp, int64_t osc, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_tri_eval: no batched form");
  if (!m || !packed || (ncomp != 1 && ncomp != 3) || n < 0 || (n > 0 && (!q || !out)))
    return fail(CMM_E_ARG, "cmm_tri_eval: bad arguments");
  if (n == 0) return CMM_OK;
  hipLaunchKernelGGL(k_tri_eval, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), *m, packed,
                     ncomp, normalize, q, n, sp, sc, out, osp, osc);
  CMM_LAUNCH_CHECK("cmm_tri_eval");
  return CMM_OK;
}

extern "C" int cmm_tri_compose(const cmm_tri_mesh* m, const double* const* levels, int depth,
                               const double* q, int64_t n, int64_t sp, int64_t sc, double* out,
                               int64_t osp, int64_t osc, double* inter, void* stream) {
  if (!m || depth < 0 || n < 0 || (n > 0 && (!q || !out)) || (depth > 0 && !levels))
    return fail(CMM_E_ARG, "cmm_tri_compose: bad arguments");
  if (depth > kMaxDepth) return fail(CMM_E_RANGE, "cmm_tri_compose: depth > 64");
  LevelPtrs lv;
  for (int i = 0; i < depth; ++i) {
    if (!levels[i]) return fail(CMM_E_ARG, "cmm_tri_compose: null level");
    lv.p[i] = levels[i];
  }
  if (n == 0) return CMM_OK;
  return launch<KTriCompose>(k_tri_compose, (unsigned)grid_for(n, 256), 256, 0, as_stream(stream),
                             KTriCompose::Args{*m, lv, depth, q, n, sp, sc, out, osp, osc, inter, 0, 0, 1,
                                               nullptr, nullptr},
                             "cmm_tri_compose");
}

static int compose_grid(const cmm_tri_mesh* m, const double* const* levels, int depth, const double* q,
                        int64_t nring, int64_t ncol, int tile_rows, int64_t q_stride_pt, int64_t q_stride_comp,
                        double* out, int64_t osp, int64_t osc, double* inter, const int32_t* pre_blk,
                        const double* pre_cb, void* stream) {
  if (tile_rows == 0) tile_rows = 4;
  if (!m || depth < 0 || depth > kMaxDepth || (depth > 0 && !levels) || nring < 0 || ncol < 1 ||
      tile_rows < 1 || tile_rows > 64 || (tile_rows & (tile_rows - 1)) || (!pre_blk != !pre_cb))
    return fail(CMM_E_ARG, "cmm_tri_compose_grid: bad arguments");
  const int64_t n = nring * ncol;
  if (n > 0 && (!q || !out)) return fail(CMM_E_ARG, "cmm_tri_compose_grid: bad arguments");
  LevelPtrs lv;
  for (int l = 0; l < depth; ++l) {
    if (!levels[l]) return fail(CMM_E_ARG, "cmm_tri_compose_grid: null level");
    lv.p[l] = levels[l];
  }
  if (n == 0) return CMM_OK;
  const int tc = 64 / tile_rows;
  const int64_t threads = (nring + tile_rows - 1) / tile_rows * ((ncol + tc - 1) / tc) * 64;
  return launch<KTriCompose>(k_tri_compose, (unsigned)grid_for(threads, 256), 256, 0, as_stream(stream),
                             KTriCompose::Args{*m, lv, depth, q, n, q_stride_pt, q_stride_comp, out, osp, osc,
                                               inter, nring, ncol, tile_rows, pre_blk, pre_cb},
                             "cmm_tri_compose_grid");
}

extern "C" int cmm_tri_compose_grid(const cmm_tri_mesh* m, const double* const* levels, int depth,
                                    const double* q, int64_t nring, int64_t ncol, int tile_rows,
                                    int64_t q_stride_pt, int64_t q_stride_comp, double* out, int64_t osp,
                                    int64_t osc, double* inter, void* stream) {
  return compose_grid(m, levels, depth, q, nring, ncol, tile_rows, q_stride_pt, q_stride_comp, out, osp, osc, inter,
                      nullptr, nullptr, stream);
}

extern "C" int cmm_tri_compose_grid_located(const cmm_tri_mesh* m, const double* const* levels, int depth,
                                            const double* q, int64_t nring, int64_t ncol, int tile_rows,
                                            int64_t q_stride_pt, int64_t q_stride_comp, double* out, int64_t osp,
                                            int64_t osc, double* inter, const int32_t* first_block,
                                            const double* first_sub_bary, void* stream) {
  if (!first_block || !first_sub_bary) return fail(CMM_E_ARG, "cmm_tri_compose_grid_located: bad arguments");
  return compose_grid(m, levels, depth, q, nring, ncol, tile_rows, q_stride_pt, q_stride_comp, out, osp, osc, inter,
                      first_block, first_sub_bary, stream);
}

extern "C" int cmm_tri_split_locate(const cmm_tri_mesh* m, const double* q, int64_t n, int64_t sp, int64_t sc,
                                    int32_t* block, double* sub_bary, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_tri_split_locate: no batched form");
  if (!m || n < 0 || (n > 0 && (!q || !block || !sub_bary))) return fail(CMM_E_ARG, "cmm_tri_split_locate: bad arguments");
  if (n > 0x7fffffffll) return fail(CMM_E_RANGE, "cmm_tri_split_locate: n > 2^31 - 1");
  if (n == 0) return CMM_OK;
  hipLaunchKernelGGL(k_tri_split_locate, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), *m, q, n, sp, sc,
                     block, sub_bary);
  CMM_LAUNCH_CHECK("cmm_tri_split_locate");
  return CMM_OK;
}

extern "C" int cmm_tri_eval_grad(const cmm_tri_mesh* m, const double* packed, int ncomp, const double* q,
                                 int64_t n, int64_t sp, int64_t sc, double* out, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_tri_eval_grad: no batched form");
  if (!m || !packed || (ncomp != 1 && ncomp != 3) || n < 0 || (n > 0 && (!q || !out)))
    return fail(CMM_E_ARG, "cmm_tri_eval_grad: bad arguments");
  if (n == 0) return CMM_OK;
  hipLaunchKernelGGL(k_tri_eval_grad, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), *m, packed, ncomp,
                     q, n, sp, sc, out);
  CMM_LAUNCH_CHECK("cmm_tri_eval_grad");
  return CMM_OK;
}

extern "C" int cmm_tri_compose_grad(const cmm_tri_mesh* m, const double* const* levels, int depth,
                                    const double* q, int64_t n, int64_t sp, int64_t sc, double* out,
                                    int64_t osp, int64_t osc, double* jac, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_tri_compose_grad: no batched form");
  if (!m || depth < 0 || depth > kMaxDepth || (depth > 0 && !levels) || n < 0 ||
      (n > 0 && (!q || !out || !jac)))
    return fail(CMM_E_ARG, "cmm_tri_compose_grad: bad arguments");
  LevelPtrs lv;
  for (int l = 0; l < depth; ++l) {
    if (!levels[l]) return fail(CMM_E_ARG, "cmm_tri_compose_grad: null level");
    lv.p[l] = levels[l];
  }
  if (n == 0) return CMM_OK;
  hipLaunchKernelGGL(k_tri_compose_grad, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), *m, lv, depth, q,
                     n, sp, sc, out, osp, osc, jac);
  CMM_LAUNCH_CHECK("cmm_tri_compose_grad");
  return CMM_OK;
}

extern "C" int cmm_tri_stencil_jets(const cmm_tri_mesh* m, const double* cur, const double* yv,
                                    int64_t yrow, const double* sf, int64_t srow, const double* tanv,
                                    double eps, double* jets, double* st, void* stream) {
  if (!m || !sf || !tanv || !jets || !(eps > 0) || (cur && !yv))
    return fail(CMM_E_ARG, "cmm_tri_stencil_jets: bad arguments");
  return launch<KTriStencilJets>(k_tri_stencil_jets, (unsigned)grid_for(m->nv, 128), 128, 0, as_stream(stream),
                                 KTriStencilJets::Args{*m, cur, yv, yrow, sf, srow, tanv, eps, jets, st},
                                 "cmm_tri_stencil_jets");
}

// ---- adjoints -------------------------------------------------------------
extern "C" int cmm_tri_compose_adjoint(const cmm_tri_mesh* m, const double* const* levels,
                                       double* const* level_bars, int depth, const double* q, int64_t n,
                                       int64_t sp, int64_t sc, const double* out_bar, double* q_bar, void* work,
                                       int64_t work_bytes, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_tri_compose_adjoint: no batched form");
  if (!m || depth < 1 || n < 0 || !levels || !level_bars || (n > 0 && (!q || !out_bar)))
    return fail(CMM_E_ARG, "cmm_tri_compose_adjoint: bad arguments");
  if (depth > kAdjMaxDepth) return fail(CMM_E_RANGE, "cmm_tri_compose_adjoint: depth > 8");
  LevelPtrs lv;
  LevelBarPtrs lb;
  for (int l = 0; l < depth; ++l) {
    if (!levels[l]) return fail(CMM_E_ARG, "cmm_tri_compose_adjoint: null level");
    lv.p[l] = levels[l];
    lb.p[l] = level_bars[l];
  }
  for (int l = depth; l < kAdjMaxDepth; ++l) lb.p[l] = nullptr;
  if (n == 0) return CMM_OK;
  ScatterWork sw;
  const int64_t nrec = (int64_t)depth * n;
  if (work) {
    int rc = scatter_layout(work, work_bytes, nrec, (int64_t)depth * m->ntri * 6, sw);
    if (rc) return rc;
  }
  hipStream_t st = as_stream(stream);
  if (work) {
    int rc = scatter_begin(sw, st);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_tri_compose_adj, dim3(grid_for(n, 256)), dim3(256), 0, st, *m, lv, lb, depth, q, n, sp, sc,
                     out_bar, q_bar, sw, work ? 1 : 0);
  CMM_LAUNCH_CHECK("cmm_tri_compose_adjoint");
  if (work) {
    ScatterTargets t;
    t.mode = 0;
    t.nc = 3;
    t.span = m->ntri * 6;
    for (int l = 0; l < depth; ++l) t.T[l] = level_bars[l];
    return scatter_reduce(sw, nrec, t, st);
  }
  return CMM_OK;
}

extern "C" int cmm_tri_assemble_adjoint(const cmm_tri_mesh* m, const double* packed_bar, double* vals_bar,
                                        void* work, int64_t work_bytes, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_tri_assemble_adjoint: no batched form");
  if (!m || !packed_bar || !vals_bar) return fail(CMM_E_ARG, "cmm_tri_assemble_adjoint: bad arguments");
  hipStream_t st = as_stream(stream);
  if (!work) {
    hipLaunchKernelGGL(k_tri_assemble_adj, dim3(grid_for(m->ntri, 256)), dim3(256), 0, st, *m, packed_bar, vals_bar);
    CMM_LAUNCH_CHECK("cmm_tri_assemble_adjoint");
    return CMM_OK;
  }
  if (work_bytes < 36 * 8 * m->ntri) return fail(CMM_E_ARG, "cmm_tri_assemble_adjoint: workspace too small");
  double* tmp = static_cast<double*>(work);
  hipLaunchKernelGGL(k_tri_assemble_adj_corners, dim3(grid_for(m->ntri, 64)), dim3(192), 0, st, *m, packed_bar, tmp);
  CMM_LAUNCH_CHECK("cmm_tri_assemble_adjoint");
  hipLaunchKernelGGL(k_tri_assemble_adj_gather, dim3(grid_for(12 * m->nv, 256)), dim3(256), 0, st, *m, tmp, vals_bar);
  CMM_LAUNCH_CHECK("cmm_tri_assemble_adjoint");
  return CMM_OK;
}

extern "C" int cmm_tri_stencil_jets_adjoint(const cmm_tri_mesh* m, const double* cur, const double* yv,
                                            int64_t yrow, const double* sf, int64_t srow, const double* tanv,
                                            double eps, const double* jets_bar, double* sf_bar, int64_t sbrow,
                                            double* cur_bar, void* work, int64_t work_bytes, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_tri_stencil_jets_adjoint: no batched form");
  if (!m || !sf || !tanv || !jets_bar || !sf_bar || !(eps > 0) || (cur && !yv) || srow < 4 * m->nv ||
      sbrow < 4 * m->nv)
    return fail(CMM_E_ARG, "cmm_tri_stencil_jets_adjoint: bad arguments");
  ScatterWork sw;
  const bool rec = work && cur && cur_bar;
  if (rec) {
    int rc = scatter_layout(work, work_bytes, m->nv, m->ntri * 6, sw);
    if (rc) return rc;
  }
  hipStream_t st = as_stream(stream);
  if (rec) {
    int rc = scatter_begin(sw, st);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_tri_stencil_jets_adj, dim3(grid_for(m->nv, 128)), dim3(128), 0, st, *m, cur, yv, yrow, sf,
                     srow, tanv, eps, jets_bar, sf_bar, sbrow, cur ? cur_bar : nullptr, sw, rec ? 1 : 0);
  CMM_LAUNCH_CHECK("cmm_tri_stencil_jets_adjoint");
  if (rec) {
    // value records of the per-vertex sums (class 0, weight 1)
    ScatterTargets t;
    t.mode = 2;
    t.nc = 3;
    t.nout = 1;
    t.T[0] = cur_bar;
    return scatter_reduce(sw, m->nv, t, st);
  }
  return CMM_OK;
}
