// MWSS spherical-harmonic transforms (real fields) and the spectral
// operators of project_stream.
//
// Algorithm (identical definition in the NumPy oracle):
//   inverse  F_m(theta_t) = sum_l f_lm Pbar_lm(cos theta_t) at the L+1 MWSS
//            rings (equatorially folded: even / odd l+m partial sums give the
//            ring and its mirror), then a complex-to-real FFT of length 2L
//            per ring.
//   forward  ring FFT (scaled by pi/L) -> resample the theta-extended
//            trigonometric polynomial F_m(2 pi - theta) = (-1)^m F_m(theta)
//            at the L Gauss-Legendre colatitudes (one GEMM per parity of m
//            with the plan's Dirichlet-kernel matrices) -> Gauss-Legendre
//            quadrature against Pbar_lm.
//   Pbar_lm  three-term recurrence in l, started from tabulated scaled
//            Pbar_mm = P * 2^(600 s); values still at s < 0 are below
//            2^-300 and are flushed to zero.
// Spectral layout [lm][f] complex with lm = m L - m(m-1)/2 + (l - m).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"

using namespace cmm;

namespace {
constexpr int kScaleExp = 600;
constexpr int kMaxFFT = 2048;
}

struct cmm_sht_plan {
  int L = 0, N = 0;  // N = 2L
  int nrad = 0;
  int radix[24];
  // device tables
  double* x_t = nullptr;     // cos theta_t (L+1)
  double* dP_t = nullptr;    // Pbar_mm mantissa at rings (L, L+1) [m][t]
  int* dS_t = nullptr;       // scale (L, L+1)
  double* x_g = nullptr;     // GL nodes (L)
  double* w_g = nullptr;     // GL weights (L)
  double* dP_g = nullptr;    // (L, L) [m][j]
  int* dS_g = nullptr;
  double* K = nullptr;       // (2, L, L+1) resampling matrices (even, odd)
  double2* tw = nullptr;     // exp(-2 pi i n / N), n < N
  double* dinv = nullptr;    // D_inv[l] (L)
  double2* rec = nullptr;    // recurrence (a_lm, b_lm) per packed lm
  std::vector<double> hx_g, hw_g;
};

// ---------------------------------------------------------------------------
// host setup
// ---------------------------------------------------------------------------
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  auto pn = [&](double z, double& p, double& dp) {
    double p0 = 1, p1 = z;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2.0 * k - 1) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) { p0 = 1; p1 = z; }
    p = p1;
    dp = n * (z * p1 - p0) / (z * z - 1);
  };
  for (int j = 0; j < n; ++j) {
    double z = std::cos(M_PI * (j + 1 - 0.25) / (n + 0.5));
    double p, dp;
    for (int it = 0; it < 30; ++it) {
      pn(z, p, dp);
      double dz = p / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-15) break;
    }
    pn(z, p, dp);
    z -= p / dp;
    pn(z, p, dp);
    x[j] = z;
    w[j] = 2.0 / ((1 - z * z) * dp * dp);
  }
}

static void legendre_diag(int L, const std::vector<double>& sint, std::vector<double>& P,
                          std::vector<int>& S) {
  int nx = (int)sint.size();
  P.assign((size_t)L * nx, 0.0);
  S.assign((size_t)L * nx, 0);
  const double big = std::ldexp(1.0, kScaleExp), lo = std::ldexp(1.0, -300);
  for (int j = 0; j < nx; ++j) {
    double p = 1.0 / std::sqrt(4.0 * M_PI);
    int s = 0;
    P[j] = p;
    for (int m = 1; m < L; ++m) {
      p = p * (-std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * sint[j]);
      if (std::fabs(p) < lo && p != 0) { p *= big; s -= 1; }
      P[(size_t)m * nx + j] = p;
      S[(size_t)m * nx + j] = s;
    }
  }
}

static double dirichlet(double d, int L) {
  double s = std::sin(d / 2);
  double v = std::fabs(s) < 1e-12 ? 2.0 * L - 1.0 : std::sin((L - 0.5) * d) / s;
  return (v + std::cos(L * d)) / (2.0 * L);
}

template <class T>
static int upload(T** dst, const T* src, size_t n) {
  hipError_t e = hipMalloc(dst, n * sizeof(T));
  if (e != hipSuccess) return hip_check(e, "cmm_sht_plan_create: hipMalloc");
  e = hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_check(e, "cmm_sht_plan_create: hipMemcpy");
  return CMM_OK;
}

extern "C" int cmm_sht_plan_destroy(cmm_sht_plan* p) {
  if (!p) return CMM_OK;
  (void)hipFree(p->x_t); (void)hipFree(p->dP_t); (void)hipFree(p->dS_t);
  (void)hipFree(p->x_g); (void)hipFree(p->w_g); (void)hipFree(p->dP_g); (void)hipFree(p->dS_g);
  (void)hipFree(p->K); (void)hipFree(p->tw); (void)hipFree(p->dinv); (void)hipFree(p->rec);
  delete p;
  return CMM_OK;
}

extern "C" int cmm_sht_plan_create(int L, const double* cos_t, const double* sin_t, cmm_sht_plan** out) {
  if (L < 2 || !cos_t || !sin_t || !out) return fail(CMM_E_ARG, "cmm_sht_plan_create: bad arguments");
  if (2 * L > kMaxFFT) return fail(CMM_E_RANGE, "cmm_sht_plan_create: L > 1024");
  cmm_sht_plan* p = new cmm_sht_plan();
  p->L = L;
  p->N = 2 * L;
  int n = p->N;
  while (n % 4 == 0) { p->radix[p->nrad++] = 4; n /= 4; }
  while (n % 2 == 0) { p->radix[p->nrad++] = 2; n /= 2; }
  while (n % 3 == 0) { p->radix[p->nrad++] = 3; n /= 3; }
  if (n != 1) {
    delete p;
    return fail(CMM_E_RANGE, "cmm_sht_plan_create: 2L must be 2^a 3^b");
  }
  std::vector<double> xt(cos_t, cos_t + L + 1), st(sin_t, sin_t + L + 1);
  std::vector<double> xg, wg;
  gauss_legendre(L, xg, wg);
  p->hx_g = xg;
  p->hw_g = wg;
  std::vector<double> sg(L), thg(L);
  for (int j = 0; j < L; ++j) {
    sg[j] = std::sqrt((1 - xg[j]) * (1 + xg[j]));
    thg[j] = std::acos(xg[j]);
  }
  std::vector<double> Pt, Pg;
  std::vector<int> St, Sg;
  legendre_diag(L, st, Pt, St);
  legendre_diag(L, sg, Pg, Sg);
  // resampling matrices: K_par[j][t] = D(th_j - th_t) + (-1)^par D(th_j + th_t), 0 < t < L
  std::vector<double> K((size_t)2 * L * (L + 1));
  for (int j = 0; j < L; ++j)
    for (int t = 0; t <= L; ++t) {
      double tht = 2.0 * t * M_PI / (2.0 * L);
      double d1 = dirichlet(thg[j] - tht, L);
      double d2 = (t > 0 && t < L) ? dirichlet(thg[j] + tht, L) : 0.0;
      K[(size_t)j * (L + 1) + t] = d1 + d2;
      K[(size_t)L * (L + 1) + (size_t)j * (L + 1) + t] = d1 - d2;
    }
  std::vector<double2> tw(p->N);
  for (int k = 0; k < p->N; ++k) {
    double a = -2.0 * M_PI * k / p->N;
    tw[k] = double2{std::cos(a), std::sin(a)};
  }
  std::vector<double2> rec((size_t)L * (L + 1) / 2, double2{0.0, 0.0});
  for (int m = 0; m < L; ++m) {
    size_t lm0 = (size_t)m * L - (size_t)m * (m - 1) / 2;
    for (int l = m + 1; l < L; ++l) {
      double a = std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
      double lm1 = l - 1.0;
      double b = std::sqrt((lm1 * lm1 - (double)m * m) / (4.0 * lm1 * lm1 - 1.0));
      rec[lm0 + (l - m)] = double2{a, b};
    }
  }
  std::vector<double> dinv(L);
  for (int l = 0; l < L; ++l) dinv[l] = l == 0 ? 0.0 : -1.0 / (double)((int64_t)l * (l + 1));
  int rc;
  if ((rc = upload(&p->x_t, xt.data(), xt.size())) || (rc = upload(&p->dP_t, Pt.data(), Pt.size())) ||
      (rc = upload(&p->dS_t, St.data(), St.size())) || (rc = upload(&p->x_g, xg.data(), xg.size())) ||
      (rc = upload(&p->w_g, wg.data(), wg.size())) || (rc = upload(&p->dP_g, Pg.data(), Pg.size())) ||
      (rc = upload(&p->dS_g, Sg.data(), Sg.size())) || (rc = upload(&p->K, K.data(), K.size())) ||
      (rc = upload(&p->tw, tw.data(), tw.size())) || (rc = upload(&p->dinv, dinv.data(), dinv.size())) ||
      (rc = upload(&p->rec, rec.data(), rec.size()))) {
    cmm_sht_plan_destroy(p);
    return rc;
  }
  *out = p;
  return CMM_OK;
}

extern "C" int cmm_sht_plan_gl(const cmm_sht_plan* p, double* x, double* w) {
  if (!p || !x || !w) return fail(CMM_E_ARG, "cmm_sht_plan_gl: bad arguments");
  for (int j = 0; j < p->L; ++j) { x[j] = p->hx_g[j]; w[j] = p->hw_g[j]; }
  return CMM_OK;
}

extern "C" int cmm_sht_workspace_bytes(const cmm_sht_plan* p, int nf, int64_t* bytes) {
  if (!p || nf < 1 || !bytes) return fail(CMM_E_ARG, "cmm_sht_workspace_bytes: bad arguments");
  int64_t L = p->L;
  *bytes = (int64_t)nf * ((L + 1) * L + L * L) * 16;
  return CMM_OK;
}

// ---------------------------------------------------------------------------
// FFT (Stockham, radix 4 / 2 / 3, in LDS; one ring pair per block)
// ---------------------------------------------------------------------------
struct FFTPlanDev {
  int N, nrad;
  int radix[24];
};

__device__ inline double2 cmul(double2 a, double2 b) {
  return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ inline double2 cadd(double2 a, double2 b) { return double2{a.x + b.x, a.y + b.y}; }
__device__ inline double2 csub(double2 a, double2 b) { return double2{a.x - b.x, a.y - b.y}; }
// s * i * a
__device__ inline double2 cmuli(double2 a, double s) { return double2{-s * a.y, s * a.x}; }

template <int R>
__device__ inline void fft_pass(const double2* __restrict__ tw, const double2* in, double2* out, int N,
                                int Ns, double sign) {
  const int NR = N / R;
  const int stride = N / (Ns * R);
  for (int j = threadIdx.x; j < NR; j += blockDim.x) {
    const int k = j % Ns;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double2 a = in[j + r * NR];
      if (r > 0 && k > 0) {
        double2 w = tw[k * r * stride];
        if (sign > 0) w.y = -w.y;
        a = cmul(a, w);
      }
      v[r] = a;
    }
    double2 y[R];
    if constexpr (R == 2) {
      y[0] = cadd(v[0], v[1]);
      y[1] = csub(v[0], v[1]);
    } else if constexpr (R == 4) {
      double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
      double2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
      y[0] = cadd(a0, b0);
      y[2] = csub(a0, b0);
      y[1] = cadd(a1, cmuli(b1, sign));
      y[3] = csub(a1, cmuli(b1, sign));
    } else {
      const double h = 0.86602540378443864676;  // sqrt(3)/2
      double2 t1 = cadd(v[1], v[2]);
      double2 t2 = double2{v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y};
      double2 t3 = cmuli(csub(v[1], v[2]), sign * h);
      y[0] = cadd(v[0], t1);
      y[1] = cadd(t2, t3);
      y[2] = csub(t2, t3);
    }
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[base + r * Ns] = y[r];
  }
}

// FFT of length N held in buf0 (ping-pong with buf1); sign -1 forward,
// +1 inverse (unnormalised).  Returns the buffer holding the result.
__device__ double2* fft_lds(const FFTPlanDev& P, const double2* __restrict__ tw, double2* buf0,
                            double2* buf1, double sign) {
  const int N = P.N;
  int Ns = 1;
  double2* in = buf0;
  double2* out = buf1;
  for (int pass = 0; pass < P.nrad; ++pass) {
    const int R = P.radix[pass];
    if (R == 4) fft_pass<4>(tw, in, out, N, Ns, sign);
    else if (R == 2) fft_pass<2>(tw, in, out, N, Ns, sign);
    else fft_pass<3>(tw, in, out, N, Ns, sign);
    __syncthreads();
    double2* t = in;
    in = out;
    out = t;
    Ns *= R;
  }
  return in;
}

// real rings -> half spectrum F[f][t][m] (m < L), scaled by pi / L
__global__ __launch_bounds__(256) void k_fft_r2c(FFTPlanDev P, const double2* __restrict__ tw, const double* __restrict__ f,
                          int L, double2* __restrict__ F) {
  extern __shared__ double2 lds[];
  double2* b0 = lds;
  double2* b1 = lds + P.N;
  const int N = P.N;
  int pairs = (L + 2) / 2;
  int fi = blockIdx.x / pairs, pr = blockIdx.x % pairs;
  int ta = 2 * pr, tb = 2 * pr + 1;
  bool hasb = tb <= L;
  const double* ra = f + ((int64_t)fi * (L + 1) + ta) * N;
  const double* rb = f + ((int64_t)fi * (L + 1) + (hasb ? tb : ta)) * N;
  for (int n = threadIdx.x; n < N; n += blockDim.x) b0[n] = double2{ra[n], hasb ? rb[n] : 0.0};
  __syncthreads();
  double2* Z = fft_lds(P, tw, b0, b1, -1.0);
  const double sc = M_PI / L;
  for (int m = threadIdx.x; m < L; m += blockDim.x) {
    double2 zk = Z[m], zn = Z[(N - m) % N];
    double2 fa = double2{0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y)};
    double2 fb = double2{0.5 * (zk.y + zn.y), -0.5 * (zk.x - zn.x)};
    F[((int64_t)fi * (L + 1) + ta) * L + m] = double2{sc * fa.x, sc * fa.y};
    if (hasb) F[((int64_t)fi * (L + 1) + tb) * L + m] = double2{sc * fb.x, sc * fb.y};
  }
}

// half spectrum F[f][t][m] -> real rings (irfft * 2L)
__global__ __launch_bounds__(256) void k_fft_c2r(FFTPlanDev P, const double2* __restrict__ tw, const double2* __restrict__ F,
                          int L, double* __restrict__ f) {
  extern __shared__ double2 lds[];
  double2* b0 = lds;
  double2* b1 = lds + P.N;
  const int N = P.N;
  int pairs = (L + 2) / 2;
  int fi = blockIdx.x / pairs, pr = blockIdx.x % pairs;
  int ta = 2 * pr, tb = 2 * pr + 1;
  bool hasb = tb <= L;
  const double2* Fa = F + ((int64_t)fi * (L + 1) + ta) * L;
  const double2* Fb = F + ((int64_t)fi * (L + 1) + (hasb ? tb : ta)) * L;
  for (int n = threadIdx.x; n < N; n += blockDim.x) {
    double2 xa, xb;
    if (n == 0) {
      xa = double2{Fa[0].x, 0.0};
      xb = double2{hasb ? Fb[0].x : 0.0, 0.0};
    } else if (n < L) {
      xa = Fa[n];
      xb = hasb ? Fb[n] : double2{0.0, 0.0};
    } else if (n == L) {
      xa = double2{0.0, 0.0};
      xb = double2{0.0, 0.0};
    } else {
      double2 a = Fa[N - n];
      xa = double2{a.x, -a.y};
      if (hasb) {
        double2 b = Fb[N - n];
        xb = double2{b.x, -b.y};
      } else {
        xb = double2{0.0, 0.0};
      }
    }
    b0[n] = double2{xa.x - xb.y, xa.y + xb.x};  // xa + i xb
  }
  __syncthreads();
  double2* Z = fft_lds(P, tw, b0, b1, +1.0);
  double* ra = f + ((int64_t)fi * (L + 1) + ta) * N;
  double* rb = f + ((int64_t)fi * (L + 1) + (hasb ? tb : ta)) * N;
  for (int n = threadIdx.x; n < N; n += blockDim.x) {
    ra[n] = Z[n].x;
    if (hasb) rb[n] = Z[n].y;
  }
}

// ---------------------------------------------------------------------------
// Legendre recurrence: Pbar_lm = a_lm (x Pbar_{l-1,m} - b_lm Pbar_{l-2,m}),
// a_lm = sqrt((4l^2-1)/(l^2-m^2)), b_lm = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1)),
// tabulated per packed (l, m) so the per-lane work is two FMAs and the
// coefficients arrive by uniform (scalar) loads.  At l = m+1, b = 0.
// ---------------------------------------------------------------------------
__device__ inline void leg_step(double a, double b, double x, double& pp, double& pc, int& s) {
  const double inv_big = 0x1p-600, hi = 0x1p300;
  double pn = a * (x * pc - b * pp);
  pp = pc;
  pc = pn;
  if (s < 0 && fabs(pc) > hi) {
    pc *= inv_big;
    pp *= inv_big;
    s += 1;
  }
}

// ---------------------------------------------------------------------------
// inverse Legendre: F[f][t][m] for all rings from flm [lm][nf].
// One block per pair of orders (m, L-1-m): every block runs L+1 recurrence
// steps, which balances the triangular (l >= m) work over the CUs.  Threads
// own the northern rings t <= L/2; even / odd l+m partial sums give ring t
// and its mirror L-t.  The coefficients of a chunk of l are staged through
// LDS once per block and read back as broadcasts.
// ---------------------------------------------------------------------------
constexpr int kInvChunk = 32;

template <int NFB>
__global__ __launch_bounds__(576) void k_leg_inverse(int L, int nf, int f0, const double* __restrict__ x_t,
                                                     const double* __restrict__ dP, const int* __restrict__ dS,
                                                     const double2* __restrict__ rec,
                                                     const double2* __restrict__ flm, double2* __restrict__ F) {
  __shared__ double2 cs[kInvChunk][NFB];
  __shared__ double2 rs[kInvChunk];
  const int nh = L / 2 + 1;  // northern rings incl. equator (t <= L/2)
  const int t = threadIdx.x;
  const bool active = t < nh;
  const int tt = active ? t : 0;
  const double x = x_t[tt];
  const int nm = ((int)blockIdx.x == L - 1 - (int)blockIdx.x) ? 1 : 2;
  for (int q = 0; q < nm; ++q) {
    const int m = q == 0 ? (int)blockIdx.x : L - 1 - (int)blockIdx.x;
    double2 E[NFB], O[NFB];
#pragma unroll
    for (int k = 0; k < NFB; ++k) { E[k] = double2{0, 0}; O[k] = double2{0, 0}; }
    double pp = 0.0, pc = dP[(int64_t)m * (L + 1) + tt];
    int s = dS[(int64_t)m * (L + 1) + tt];
    const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
    for (int lc = m; lc < L; lc += kInvChunk) {
      const int n = min(kInvChunk, L - lc);
      __syncthreads();
      for (int e = threadIdx.x; e < n * NFB; e += blockDim.x) {
        const int i = e / NFB, k = e % NFB;
        cs[i][k] = flm[(lm0 + (lc - m) + i) * nf + f0 + k];
      }
      for (int i = threadIdx.x; i < n; i += blockDim.x) rs[i] = rec[lm0 + (lc - m) + i];
      __syncthreads();
      for (int i = 0; i < n; ++i) {
        const int l = lc + i;
        if (l > m) {
          double2 ab = rs[i];
          leg_step(ab.x, ab.y, x, pp, pc, s);
        }
        const double pv = (s == 0) ? pc : 0.0;
        if (((l - m) & 1) == 0) {
#pragma unroll
          for (int k = 0; k < NFB; ++k) { E[k].x += pv * cs[i][k].x; E[k].y += pv * cs[i][k].y; }
        } else {
#pragma unroll
          for (int k = 0; k < NFB; ++k) { O[k].x += pv * cs[i][k].x; O[k].y += pv * cs[i][k].y; }
        }
      }
    }
    if (active) {
#pragma unroll
      for (int k = 0; k < NFB; ++k) {
        const int fi = f0 + k;
        F[((int64_t)fi * (L + 1) + t) * L + m] = double2{E[k].x + O[k].x, E[k].y + O[k].y};
        if (L - t != t)
          F[((int64_t)fi * (L + 1) + (L - t)) * L + m] = double2{E[k].x - O[k].x, E[k].y - O[k].y};
      }
    }
  }
}

// ---------------------------------------------------------------------------
// forward: resampling GEMM G[f][j][m] = sum_t K_{m%2}[j][t] F[f][t][m]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_resample(int L, const double* __restrict__ K, const double2* __restrict__ F,
                           double2* __restrict__ G) {
  // block: 16 nodes j x 64 m-columns; grid: (ceil(L/64), ceil(L/16), nf)
  __shared__ double Ks[2][16][17];
  __shared__ double2 Fs[16][64];
  const int fi = blockIdx.z;
  const int j0 = blockIdx.y * 16, m0 = blockIdx.x * 64;
  const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;  // 64 x 4 threads
  double2 acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  const int m = m0 + tx;
  const int par = m & 1;
  for (int t0 = 0; t0 <= L; t0 += 16) {
    for (int e = threadIdx.x; e < 2 * 16 * 16; e += blockDim.x) {
      int p = e / 256, jj = (e / 16) % 16, tt = e % 16;
      int j = j0 + jj, t = t0 + tt;
      Ks[p][jj][tt] = (j < L && t <= L) ? K[(int64_t)p * L * (L + 1) + (int64_t)j * (L + 1) + t] : 0.0;
    }
    for (int e = threadIdx.x; e < 16 * 64; e += blockDim.x) {
      int tt = e / 64, mm = e % 64;
      int t = t0 + tt, mc = m0 + mm;
      Fs[tt][mm] = (t <= L && mc < L) ? F[((int64_t)fi * (L + 1) + t) * L + mc] : double2{0, 0};
    }
    __syncthreads();
    for (int tt = 0; tt < 16; ++tt) {
      double2 fv = Fs[tt][tx];
      for (int r = 0; r < 4; ++r) {
        double kv = Ks[par][ty * 4 + r][tt];
        acc[r].x += kv * fv.x;
        acc[r].y += kv * fv.y;
      }
    }
    __syncthreads();
  }
  if (m >= L) return;
  for (int r = 0; r < 4; ++r) {
    int j = j0 + ty * 4 + r;
    if (j < L) G[((int64_t)fi * L + j) * L + m] = acc[r];
  }
}

// ---------------------------------------------------------------------------
// forward Legendre analysis: flm[lm][f] = sum_j w_j Pbar_lm(x_j) G[f][j][m].
// One block per pair of orders (m, L-1-m) and field; threads own the
// northern Gauss-Legendre nodes.  For a chunk of kAnaChunk degrees every
// thread writes its 2*kAnaChunk products to LDS, then groups of 8 lanes sum
// columns of that tile and finish with 3 butterfly steps.
// ---------------------------------------------------------------------------
constexpr int kAnaChunk = 8;

__global__ __launch_bounds__(512) void k_leg_forward(int L, int nf, const double* __restrict__ xg,
                                                     const double* __restrict__ wg, const double* __restrict__ dP,
                                                     const int* __restrict__ dS, const double2* __restrict__ rec,
                                                     const double2* __restrict__ G, double2* __restrict__ flm) {
  extern __shared__ double tile[];  // [2*kAnaChunk][T]
  __shared__ double2 rs[kAnaChunk];
  const int fi = blockIdx.y;
  const int nh = (L + 1) / 2;  // northern nodes (incl. the middle one when L is odd)
  const int T = blockDim.x;    // multiple of 64 >= nh
  const int j = threadIdx.x;
  const bool active = j < nh;
  const int jj = active ? j : 0;
  const int jm = L - 1 - jj;
  const double x = xg[jj];
  const int nm = ((int)blockIdx.x == L - 1 - (int)blockIdx.x) ? 1 : 2;
  // reduction role: output o = j / 8 (of 2*kAnaChunk), part = j % 8
  const int o = j >> 3, part = j & 7;
  for (int q = 0; q < nm; ++q) {
    const int m = q == 0 ? (int)blockIdx.x : L - 1 - (int)blockIdx.x;
    double2 S = {0, 0}, D = {0, 0};
    if (active) {
      double2 a = G[((int64_t)fi * L + jj) * L + m];
      double w = wg[jj];
      if (jm != jj) {
        double2 b = G[((int64_t)fi * L + jm) * L + m];
        S = double2{w * (a.x + b.x), w * (a.y + b.y)};
        D = double2{w * (a.x - b.x), w * (a.y - b.y)};
      } else {
        S = double2{w * a.x, w * a.y};
        D = S;
      }
    }
    double pp = 0.0, pc = dP[(int64_t)m * L + jj];
    int s = dS[(int64_t)m * L + jj];
    const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
    for (int lc = m; lc < L; lc += kAnaChunk) {
      if (j < kAnaChunk && lc + j < L) rs[j] = rec[lm0 + (lc - m) + j];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kAnaChunk; ++i) {
        const int l = lc + i;
        double vr = 0.0, vi = 0.0;
        if (l < L) {
          if (l > m) {
            double2 ab = rs[i];
            leg_step(ab.x, ab.y, x, pp, pc, s);
          }
          const double pv = (s == 0 && active) ? pc : 0.0;
          const double2 src = (((l - m) & 1) == 0) ? S : D;
          vr = pv * src.x;
          vi = pv * src.y;
        }
        tile[(2 * i) * T + j] = vr;
        tile[(2 * i + 1) * T + j] = vi;
      }
      __syncthreads();
      if (o < 2 * kAnaChunk) {
        double acc = 0.0;
        for (int c = part; c < T; c += 8) acc += tile[o * T + c];
        acc += __shfl_xor(acc, 1, 8);
        acc += __shfl_xor(acc, 2, 8);
        acc += __shfl_xor(acc, 4, 8);
        const int l = lc + (o >> 1);
        if (part == 0 && l < L) {
          double* dst = reinterpret_cast<double*>(flm + (lm0 + (l - m)) * nf + fi);
          dst[o & 1] = acc;
        }
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// spectral operators (spherical_harmonic_tools.py:14-19, 49-62)
// ---------------------------------------------------------------------------
struct Cplx {
  double re, im;
};
__device__ inline Cplx cz() { return Cplx{0.0, 0.0}; }

__device__ inline int64_t lm_index(int L, int l, int m) {
  return (int64_t)m * L - (int64_t)m * (m - 1) / 2 + (l - m);
}

// coefficient (l, m) of a real field from its m >= 0 half (conjugate symmetry)
template <class F>
__device__ inline Cplx coef_any(int l, int m, F get) {
  if (m > l || -m > l) return cz();
  if (m >= 0) return get(l, m);
  Cplx c = get(l, -m);
  double sg = ((-m) & 1) ? -1.0 : 1.0;
  return Cplx{sg * c.re, -sg * c.im};
}

__device__ inline double Lp(int l, int m) {
  double v = (double)l * (l + 1) - (double)m * (m - 1);
  return v >= 0 ? sqrt(v) : 0.0;
}
__device__ inline double Lm(int l, int m) {
  double v = (double)l * (l + 1) - (double)m * (m + 1);
  return v >= 0 ? sqrt(v) : 0.0;
}

// angular_momentum at (l, m): returns [i Lx a, i Ly a, i Lz a]
template <class F>
__device__ inline void angmom(int l, int m, F get, Cplx out[3]) {
  Cplx am1 = coef_any(l, m - 1, get), ap1 = coef_any(l, m + 1, get), a0 = coef_any(l, m, get);
  double lp = Lp(l, m), lmn = Lm(l, m);
  Cplx P = Cplx{lp * am1.re, lp * am1.im};
  Cplx Mi = Cplx{lmn * ap1.re, lmn * ap1.im};
  Cplx ox = Cplx{(P.re + Mi.re) / 2, (P.im + Mi.im) / 2};
  Cplx d = Cplx{Mi.re - P.re, Mi.im - P.im};
  Cplx oy = Cplx{(0.0 * d.re - 1.0 * d.im) / 2, (0.0 * d.im + 1.0 * d.re) / 2};  // 1j*(M-P)/2
  Cplx az = Cplx{(double)m * a0.re, (double)m * a0.im};
  out[0] = Cplx{0.0 * ox.re - 1.0 * ox.im, 0.0 * ox.im + 1.0 * ox.re};
  out[1] = Cplx{0.0 * oy.re - 1.0 * oy.im, 0.0 * oy.im + 1.0 * oy.re};
  out[2] = Cplx{0.0 * az.re - 1.0 * az.im, 0.0 * az.im + 1.0 * az.re};
}

__global__ __launch_bounds__(128) void k_stream_ops(int L, const double2* __restrict__ w, const double* __restrict__ dinv,
                             int apply_dinv, double2* __restrict__ out) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)L * (L + 1) / 2;
  if (idx >= total) return;
  // decode (l, m) from the packed index
  int m = 0;
  {
    int lo = 0, hi = L - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) / 2;
      if (lm_index(L, mid, mid) <= idx) lo = mid; else hi = mid - 1;
    }
    m = lo;
  }
  int l = m + (int)(idx - lm_index(L, m, m));
  auto psi = [&](int ll, int mm) -> Cplx {
    double2 v = w[lm_index(L, ll, mm)];
    double dd = apply_dinv ? dinv[ll] : 1.0;
    return apply_dinv ? Cplx{v.x * dd, v.y * dd} : Cplx{v.x, v.y};
  };
  Cplx u[3];
  angmom(l, m, psi, u);
  for (int i = 0; i < 3; ++i) out[idx * 12 + i] = double2{u[i].re, u[i].im};
  for (int i = 0; i < 3; ++i) {
    auto ui = [&](int ll, int mm) -> Cplx {
      Cplx r[3];
      angmom(ll, mm, psi, r);
      return r[i];
    };
    Cplx g[3];
    angmom(l, m, ui, g);
    for (int k = 0; k < 3; ++k) out[idx * 12 + 3 + 3 * i + k] = double2{g[k].re, g[k].im};
  }
}

__global__ __launch_bounds__(128) void k_scalar_ops(int L, const double2* __restrict__ w, double2* __restrict__ out) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)L * (L + 1) / 2;
  if (idx >= total) return;
  int lo = 0, hi = L - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) / 2;
    if (lm_index(L, mid, mid) <= idx) lo = mid; else hi = mid - 1;
  }
  int m = lo;
  int l = m + (int)(idx - lm_index(L, m, m));
  auto f = [&](int ll, int mm) -> Cplx {
    double2 v = w[lm_index(L, ll, mm)];
    return Cplx{v.x, v.y};
  };
  Cplx u[3];
  angmom(l, m, f, u);
  out[idx * 4] = w[idx];
  for (int i = 0; i < 3; ++i) out[idx * 4 + 1 + i] = double2{u[i].re, u[i].im};
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static FFTPlanDev fft_plan(const cmm_sht_plan* p) {
  FFTPlanDev P;
  P.N = p->N;
  P.nrad = p->nrad;
  for (int i = 0; i < 24; ++i) P.radix[i] = p->radix[i];
  return P;
}

extern "C" int cmm_sht_inverse(const cmm_sht_plan* p, const double* flm, int nf, double* f, void* work,
                               void* stream) {
  if (!p || !flm || !f || !work || nf < 1) return fail(CMM_E_ARG, "cmm_sht_inverse: bad arguments");
  const int L = p->L;
  double2* F = reinterpret_cast<double2*>(work);
  hipStream_t st = as_stream(stream);
  const int nh = L / 2 + 1;
  const int T = ((nh + 63) / 64) * 64;
  if (T > 576) return fail(CMM_E_RANGE, "cmm_sht_inverse: L > 1024");
  const dim3 grid((L + 1) / 2), block(T);
  for (int f0 = 0; f0 < nf;) {
    int left = nf - f0;
    if (left >= 6) {
      hipLaunchKernelGGL(k_leg_inverse<6>, grid, block, 0, st, L, nf, f0, p->x_t, p->dP_t, p->dS_t, p->rec,
                         reinterpret_cast<const double2*>(flm), F);
      f0 += 6;
    } else if (left >= 4) {
      hipLaunchKernelGGL(k_leg_inverse<4>, grid, block, 0, st, L, nf, f0, p->x_t, p->dP_t, p->dS_t, p->rec,
                         reinterpret_cast<const double2*>(flm), F);
      f0 += 4;
    } else {
      hipLaunchKernelGGL(k_leg_inverse<1>, grid, block, 0, st, L, nf, f0, p->x_t, p->dP_t, p->dS_t, p->rec,
                         reinterpret_cast<const double2*>(flm), F);
      f0 += 1;
    }
    CMM_LAUNCH_CHECK("cmm_sht_inverse: legendre");
  }
  int pairs = (L + 2) / 2;
  size_t lds = (size_t)2 * p->N * sizeof(double2);
  hipLaunchKernelGGL(k_fft_c2r, dim3(nf * pairs), dim3(256), lds, st, fft_plan(p), p->tw, F, L, f);
  CMM_LAUNCH_CHECK("cmm_sht_inverse: fft");
  return CMM_OK;
}

extern "C" int cmm_sht_forward(const cmm_sht_plan* p, const double* f, int nf, double* flm, void* work,
                               void* stream) {
  if (!p || !flm || !f || !work || nf < 1) return fail(CMM_E_ARG, "cmm_sht_forward: bad arguments");
  const int L = p->L;
  double2* F = reinterpret_cast<double2*>(work);
  double2* G = F + (int64_t)nf * (L + 1) * L;
  hipStream_t st = as_stream(stream);
  int pairs = (L + 2) / 2;
  size_t lds = (size_t)2 * p->N * sizeof(double2);
  hipLaunchKernelGGL(k_fft_r2c, dim3(nf * pairs), dim3(256), lds, st, fft_plan(p), p->tw, f, L, F);
  CMM_LAUNCH_CHECK("cmm_sht_forward: fft");
  hipLaunchKernelGGL(k_resample, dim3((L + 63) / 64, (L + 15) / 16, nf), dim3(256), 0, st, L, p->K, F, G);
  CMM_LAUNCH_CHECK("cmm_sht_forward: resample");
  int nh = (L + 1) / 2;
  int T = ((nh + 63) / 64) * 64;
  if (T < 128) T = 128;  // 2*kAnaChunk outputs x 8 lanes
  if (T > 512) return fail(CMM_E_RANGE, "cmm_sht_forward: L > 1024");
  size_t lds2 = (size_t)2 * kAnaChunk * T * sizeof(double);
  hipLaunchKernelGGL(k_leg_forward, dim3((L + 1) / 2, nf), dim3(T), lds2, st, L, nf, p->x_g, p->w_g, p->dP_g,
                     p->dS_g, p->rec, G, reinterpret_cast<double2*>(flm));
  CMM_LAUNCH_CHECK("cmm_sht_forward: legendre");
  return CMM_OK;
}

extern "C" int cmm_sht_stream_operators(const cmm_sht_plan* p, const double* w, int apply_dinv,
                                        double* out, void* stream) {
  if (!p || !w || !out) return fail(CMM_E_ARG, "cmm_sht_stream_operators: bad arguments");
  int64_t total = (int64_t)p->L * (p->L + 1) / 2;
  hipLaunchKernelGGL(k_stream_ops, dim3(grid_for(total, 128)), dim3(128), 0, as_stream(stream), p->L,
                     reinterpret_cast<const double2*>(w), p->dinv, apply_dinv, reinterpret_cast<double2*>(out));
  CMM_LAUNCH_CHECK("cmm_sht_stream_operators");
  return CMM_OK;
}

extern "C" int cmm_sht_scalar_operators(const cmm_sht_plan* p, const double* w, double* out, void* stream) {
  if (!p || !w || !out) return fail(CMM_E_ARG, "cmm_sht_scalar_operators: bad arguments");
  int64_t total = (int64_t)p->L * (p->L + 1) / 2;
  hipLaunchKernelGGL(k_scalar_ops, dim3(grid_for(total, 128)), dim3(128), 0, as_stream(stream), p->L,
                     reinterpret_cast<const double2*>(w), reinterpret_cast<double2*>(out));
  CMM_LAUNCH_CHECK("cmm_sht_scalar_operators");
  return CMM_OK;
}
