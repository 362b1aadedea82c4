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
#include <queue>
#include <vector>

#include "batch.h"
#include "common.h"

using namespace cmm;

namespace {
constexpr int kScaleExp = 600;
constexpr int kMaxFFT = 4096;  // ring length 2L: L <= 2048 (the FFT's two LDS buffers take 32 N bytes)
constexpr int kFwdSegHost = 64;   // degrees per forward work unit (kFwdSeg)
constexpr int kOzJ = 16;          // split-BF16 synthesis: degrees of one parity per chunk
constexpr int kOzMaxS = 7;        // split-BF16 synthesis: most digits per operand
constexpr int kI8MaxS = 8;        // int8 digit-table synthesis: most digits per operand
constexpr int kI8Rings = 32;      // int8 synthesis: rings per tile (M of the 32x32x32 instruction)
constexpr int kI8Deg = 32;        // int8 synthesis: degrees of one parity per chunk (its K)
}

// Scratch of cmm_sht_forward_noalloc: a fixed array in the code object, so
// every GPU the library runs on holds its own copy (at its own address)
// until the process ends.  Host code never passes `scratch` itself to a
// kernel (its host-side name is a stand-in, not a device address): the plan
// keeps the address hipGetSymbolAddress returns on the plan's device.
constexpr int64_t kNoAllocScratchBytes = (int64_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct cmm_sht_plan {
  int L = 0, N = 0;  // N = 2L
  void* scratch = nullptr;   // device address of `scratch` on the plan's device
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
  double* K = nullptr;       // (2, tpad, L) resampling matrices (even, odd), transposed, zero rows t > L
  int tpad = 0;
  double2* tw = nullptr;     // exp(-2 pi i n / N), n < N
  double* dinv = nullptr;    // D_inv[l] (L)
  double2* rec = nullptr;    // recurrence (a_lm, a_lm b_lm) per packed lm
  int legendre = 0;          // CMM_LEGENDRE_* arithmetic of the synthesis
  // split-BF16 synthesis: chunk offsets offJ[m] = sum_{m' < m} ceil((L - m') / 32)
  // (L + 1), their total, and the exponent bounding |Pbar| (2^eA > max |Pbar|)
  int* offJ = nullptr;
  int64_t nJtot = 0;
  int eA = 0;
  int nunits = 0;            // forward work units (m, first degree of a segment)
  int2* units = nullptr;
  int* unit0 = nullptr;      // first unit of each m (L)
  double2* st_p = nullptr;   // recurrence state at each unit start (nunits, L/2 nodes)
  int* st_s = nullptr;
  // transposed synthesis (adjoint): the same analysis schedule over the L+1
  // rings, unit weights, per-order scale of the transposed real FFT
  double2* st_pr = nullptr;  // (nunits, L/2 + 1 northern rings)
  int* st_sr = nullptr;
  double* w_r = nullptr;     // ones (L+1)
  double* msc_r = nullptr;   // (L / pi) * (1 for m = 0, else 2)  (L)
  double* K2 = nullptr;      // (2, jpad, tcols) K transposed for the adjoint resampling: K2[p][j][t] = K[p][t][j]
  int jpad = 0, tcols = 0;
  // synthesis schedule: wave w of k_leg_inverse runs the (order, ring block)
  // tasks inv_tasks[inv_off[w] .. inv_off[w + 1])
  int inv_waves = 0;
  int2* inv_tasks = nullptr;
  int* inv_off = nullptr;
  // int8 digit-table synthesis: nrt tiles of 32 northern rings (t = 1 .. L/2);
  // per (order m, parity p) nk(m, p) = ceil(ceil((L - m - p) / 2) / 32) chunks
  // of 32 degrees; the coefficient digits of chunk (m, p, kc) are chunk block
  // i8_boff[2 m + p] + kc of the workspace (i8_nb blocks in all).  The table
  // (i8S digits) holds, per (m, p, tile), the chunks from the first one with
  // a nonzero digit on: i8_tdesc = (that chunk, its block in i8_tab, the
  // tile's exponent).
  int nrt = 0;
  int i8_nb = 0;
  int* i8_boff = nullptr;
  int* i8_cdesc = nullptr;   // (i8_nb): chunk block -> (m << 16) | (p << 15) | kc
  int i8S = 0;
  int64_t i8_blocks = 0;
  int4* i8_tdesc = nullptr;
  unsigned char* i8_tab = nullptr;
  std::vector<double> hx_g, hw_g;
};

static int build_restart(cmm_sht_plan* p, const std::vector<int2>& units, const std::vector<int>& unit0);
static int build_inverse_tasks(cmm_sht_plan* p);

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
  (void)hipFree(p->x_t); (void)hipFree(p->dP_t); (void)hipFree(p->dS_t); (void)hipFree(p->offJ);
  (void)hipFree(p->x_g); (void)hipFree(p->w_g); (void)hipFree(p->dP_g); (void)hipFree(p->dS_g);
  (void)hipFree(p->K); (void)hipFree(p->tw); (void)hipFree(p->dinv); (void)hipFree(p->rec);
  (void)hipFree(p->units); (void)hipFree(p->unit0); (void)hipFree(p->st_p); (void)hipFree(p->st_s);
  (void)hipFree(p->st_pr); (void)hipFree(p->st_sr); (void)hipFree(p->w_r); (void)hipFree(p->msc_r);
  (void)hipFree(p->K2); (void)hipFree(p->inv_tasks); (void)hipFree(p->inv_off);
  (void)hipFree(p->i8_boff); (void)hipFree(p->i8_cdesc); (void)hipFree(p->i8_tdesc); (void)hipFree(p->i8_tab);
  delete p;
  return CMM_OK;
}

extern "C" int cmm_sht_plan_create(int L, const double* cos_t, const double* sin_t, cmm_sht_plan** out) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_sht_plan_create: no batched form");
  if (L < 2 || !cos_t || !sin_t || !out) return fail(CMM_E_ARG, "cmm_sht_plan_create: bad arguments");
  if (2 * L > kMaxFFT) return fail(CMM_E_RANGE, "cmm_sht_plan_create: L > 2048");
  cmm_sht_plan* p = new cmm_sht_plan();
  p->L = L;
  p->N = 2 * L;
  // ring FFT passes: radix 4, 2 and 3 butterflies, then every remaining prime
  // factor as a direct DFT pass (any L: 2L = 2000 for L = 1000 takes three
  // radix-5 passes, 2L = 2 p for a prime L one pass of length L)
  int n = p->N;
  while (n % 4 == 0) { p->radix[p->nrad++] = 4; n /= 4; }
  while (n % 2 == 0) { p->radix[p->nrad++] = 2; n /= 2; }
  while (n % 3 == 0) { p->radix[p->nrad++] = 3; n /= 3; }
  for (int f = 5; n > 1; f += 2)
    while (n % f == 0) { p->radix[p->nrad++] = f; n /= f; }
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
  // resampling matrices K_par[j][t] = D(th_j - th_t) + (-1)^par D(th_j + th_t)
  // (second term for 0 < t < L only), stored transposed K[par][t][j] with
  // zero rows t > L up to a multiple of 4, at least one (the resampling
  // kernel's loads past the last step are clamped onto the last row)
  const int tpad = ((L + 1) / 4 + 1) * 4;
  p->tpad = tpad;
  std::vector<double> K((size_t)2 * L * tpad, 0.0);
  for (int j = 0; j < L; ++j)
    for (int t = 0; t <= L; ++t) {
      double tht = 2.0 * t * M_PI / (2.0 * L);
      double d1 = dirichlet(thg[j] - tht, L);
      double d2 = (t > 0 && t < L) ? dirichlet(thg[j] + tht, L) : 0.0;
      K[(size_t)t * L + j] = d1 + d2;
      K[(size_t)L * tpad + (size_t)t * L + j] = d1 - d2;
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
      rec[lm0 + (l - m)] = double2{a, a * b};
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
  std::vector<int> offJ(L + 1, 0);
  for (int m = 0; m < L; ++m) offJ[m + 1] = offJ[m] + (L - m + 31) / 32;
  p->nJtot = offJ[L];
  p->eA = std::ilogb(1.01 * std::sqrt((2.0 * L - 1.0) / (4.0 * M_PI))) + 1;
  if ((rc = upload(&p->offJ, offJ.data(), offJ.size()))) {
    cmm_sht_plan_destroy(p);
    return rc;
  }
  // int8 synthesis: coefficient chunk blocks per (order, parity)
  p->nrt = (L / 2 + kI8Rings - 1) / kI8Rings;
  std::vector<int> boff(2 * L + 1, 0);
  for (int m = 0; m < L; ++m)
    for (int par = 0; par < 2; ++par) {
      const int nl = (L - m - par + 1) / 2;
      boff[2 * m + par + 1] = boff[2 * m + par] + (nl + kI8Deg - 1) / kI8Deg;
    }
  p->i8_nb = boff[2 * L];
  std::vector<int> cdesc(p->i8_nb);
  for (int m = 0; m < L; ++m)
    for (int par = 0; par < 2; ++par)
      for (int kc = 0; kc < boff[2 * m + par + 1] - boff[2 * m + par]; ++kc)
        cdesc[boff[2 * m + par] + kc] = (m << 16) | (par << 15) | kc;
  if ((rc = upload(&p->i8_boff, boff.data(), boff.size())) || (rc = upload(&p->i8_cdesc, cdesc.data(), cdesc.size()))) {
    cmm_sht_plan_destroy(p);
    return rc;
  }
  std::vector<double> ones(L + 1, 1.0), msc(L);
  for (int m = 0; m < L; ++m) msc[m] = (m == 0 ? 1.0 : 2.0) * L / M_PI;
  // transposed resampling matrices, zero rows j >= L and columns t > L
  p->jpad = (L + 3) / 4 * 4;
  p->tcols = (L + 1 + 15) / 16 * 16;
  std::vector<double> K2((size_t)2 * p->jpad * p->tcols, 0.0);
  for (int par = 0; par < 2; ++par)
    for (int t = 0; t <= L; ++t)
      for (int j = 0; j < L; ++j)
        K2[((size_t)par * p->jpad + j) * p->tcols + t] = K[(size_t)par * L * tpad + (size_t)t * L + j];
  if ((rc = upload(&p->w_r, ones.data(), ones.size())) || (rc = upload(&p->msc_r, msc.data(), msc.size())) ||
      (rc = upload(&p->K2, K2.data(), K2.size())) || (rc = build_inverse_tasks(p))) {
    cmm_sht_plan_destroy(p);
    return rc;
  }
  std::vector<int2> units;
  std::vector<int> unit0(L);
  for (int m = 0; m < L; ++m) {
    unit0[m] = (int)units.size();
    for (int ls = m; ls < L; ls += kFwdSegHost) units.push_back(int2{m, ls});
  }
  if ((rc = build_restart(p, units, unit0))) {
    cmm_sht_plan_destroy(p);
    return rc;
  }
  // the no-allocation forward's scratch on this (current) device
  hipError_t e = hipGetSymbolAddress(&p->scratch, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    cmm_sht_plan_destroy(p);
    return hip_check(e, "cmm_sht_plan_create: hipGetSymbolAddress(scratch)");
  }
  *out = p;
  return CMM_OK;
}

static int build_i8_table(cmm_sht_plan* p, int S);

extern "C" int cmm_sht_set_legendre(cmm_sht_plan* p, int mode) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_sht_set_legendre: no batched form");
  if (!p || mode < CMM_LEGENDRE_FP64 || mode > CMM_LEGENDRE_I8X8)
    return fail(CMM_E_ARG, "cmm_sht_set_legendre: bad arguments");
  if (mode >= CMM_LEGENDRE_I8X4) {
    const int rc = build_i8_table(p, 4 + (mode - CMM_LEGENDRE_I8X4));
    if (rc) return rc;
  }
  p->legendre = mode;
  return CMM_OK;
}

extern "C" int cmm_sht_plan_gl(const cmm_sht_plan* p, double* x, double* w) {
  if (!p || !x || !w) return fail(CMM_E_ARG, "cmm_sht_plan_gl: bad arguments");
  for (int j = 0; j < p->L; ++j) { x[j] = p->hx_g[j]; w[j] = p->hw_g[j]; }
  return CMM_OK;
}

// Workspace: the Fourier coefficients F (nf, L+1, L) complex, then either the
// analysis intermediate G (nf, L, L) complex or, during a split-BF16
// synthesis, the coefficient digits of one group of <= 16 fields (room for
// kOzMaxS digits whatever the plan's mode) and their exponents (L, 32).
static int64_t sht_f_bytes(const cmm_sht_plan* p, int nf) { return (int64_t)nf * (p->L + 1) * p->L * 16; }
static int64_t sht_oz_dstride(const cmm_sht_plan* p, int nfp) { return p->nJtot * 2 * (2 * nfp) * kOzJ; }
// int8 synthesis: the coefficient digits of one group of fields (room for
// kI8MaxS digits), then their exponents (i8_nb chunks, 32 columns)
static int64_t sht_i8_dbytes(const cmm_sht_plan* p) { return ((int64_t)kI8MaxS * p->i8_nb * 1024 + 255) / 256 * 256; }
static int64_t sht_work_bytes(const cmm_sht_plan* p, int nf) {
  const int64_t L = p->L;
  const int64_t g = (int64_t)nf * L * L * 16;
  const int64_t oz = kOzMaxS * sht_oz_dstride(p, std::min(nf, 16)) * 2 + 256 + L * 32 * 4;
  const int64_t i8 = sht_i8_dbytes(p) + (int64_t)p->i8_nb * 32 * 4;
  return sht_f_bytes(p, nf) + std::max(g, std::max(oz, i8));
}

extern "C" int cmm_sht_workspace_bytes(const cmm_sht_plan* p, int nf, int64_t* bytes) {
  if (!p || nf < 1 || !bytes) return fail(CMM_E_ARG, "cmm_sht_workspace_bytes: bad arguments");
  *bytes = sht_work_bytes(p, nf);
  return CMM_OK;
}

// ---------------------------------------------------------------------------
// FFT (Stockham, radix 4 / 2 / 3 butterflies and direct-DFT passes for the
// other prime factors, in LDS; one ring pair per block)
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

// radix-4 butterfly (4-point DFT, root exp(sign 2 pi i / 4))
__device__ inline void bfly4(const double2 (&v)[4], double2 (&y)[4], double sign) {
  double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  double2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
  y[0] = cadd(a0, b0);
  y[2] = csub(a0, b0);
  y[1] = cadd(a1, cmuli(b1, sign));
  y[3] = csub(a1, cmuli(b1, sign));
}

// butterfly j of a Stockham pass of radix R and span Ns: its R inputs
// in[j + r N / R] (twiddled) into the R-point DFT y
template <int R>
__device__ inline void fft_bfly(const double2* __restrict__ tw, const double2* in, int N, int Ns, int j, double sign,
                                double2 (&y)[R]) {
  const int NR = N / R;
  const int stride = N / (Ns * R);
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
  if constexpr (R == 2) {
    y[0] = cadd(v[0], v[1]);
    y[1] = csub(v[0], v[1]);
  } else if constexpr (R == 4) {
    bfly4(v, y, sign);
  } else {
    const double h = 0.86602540378443864676;  // sqrt(3)/2
    double2 t1 = cadd(v[1], v[2]);
    double2 t2 = double2{v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y};
    double2 t3 = cmuli(csub(v[1], v[2]), sign * h);
    y[0] = cadd(v[0], t1);
    y[1] = cadd(t2, t3);
    y[2] = csub(t2, t3);
  }
}

// one Stockham pass of radix R; output element i goes to store(i, value)
template <int R, class Store>
__device__ inline void fft_pass_to(const double2* __restrict__ tw, const double2* in, int N, int Ns, double sign,
                                   Store&& store) {
  const int NR = N / R;
  for (int j = threadIdx.x; j < NR; j += blockDim.x) {
    double2 y[R];
    fft_bfly<R>(tw, in, N, Ns, j, sign, y);
    const int base = (j / Ns) * Ns * R + j % Ns;
#pragma unroll
    for (int r = 0; r < R; ++r) store(base + r * Ns, y[r]);
  }
}

// the same pass in place on one buffer: every thread reads the inputs of its
// (at most KB) butterflies into registers, the block synchronises, then
// writes the outputs; ends synchronised
template <int R, int KB>
__device__ inline void fft_pass_inplace(const double2* __restrict__ tw, double2* buf, int N, int Ns, double sign) {
  const int NR = N / R;
  double2 y[KB][R];
#pragma unroll
  for (int b = 0; b < KB; ++b) {
    const int j = threadIdx.x + b * blockDim.x;
    if (j < NR) fft_bfly<R>(tw, buf, N, Ns, j, sign, y[b]);
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < KB; ++b) {
    const int j = threadIdx.x + b * blockDim.x;
    if (j < NR) {
      const int base = (j / Ns) * Ns * R + j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[base + r * Ns] = y[b][r];
    }
  }
  __syncthreads();
}

// FFTs that run in place on one LDS buffer of N elements (256 threads):
// radices 2, 3 and 4 only, N <= 2048, so a thread holds at most 2 radix-4,
// 3 radix-3 or 4 radix-2 butterflies of a pass
constexpr int kFftThreads = 256;
__host__ __device__ inline bool fft_one_buffer(const FFTPlanDev& P) {
  if (P.N > 2048) return false;
  for (int i = 0; i < P.nrad; ++i)
    if (P.radix[i] < 2 || P.radix[i] > 4) return false;
  return true;
}
// dynamic LDS bytes of the ring FFT kernels
__host__ __device__ inline size_t fft_lds_bytes(const FFTPlanDev& P) {
  return (size_t)(fft_one_buffer(P) ? 1 : 2) * P.N * 16;
}

template <int R>
__device__ inline void fft_pass(const double2* __restrict__ tw, const double2* in, double2* out, int N,
                                int Ns, double sign) {
  fft_pass_to<R>(tw, in, N, Ns, sign, [out](int i, double2 v) { out[i] = v; });
}

// Stockham pass of any radix R as a direct DFT: output q of butterfly group
// j (k = j mod Ns) is sum_r in[j + r N/R] W^(r (k + q Ns) N / (Ns R)), W =
// exp(sign 2 pi i / N) -- the group twiddle and the R-point DFT root in one
// table lookup; one thread per output, R complex products each.
template <class Store>
__device__ inline void fft_pass_dft_to(const double2* __restrict__ tw, const double2* in, int N, int Ns, int R,
                                       double sign, Store&& store) {
  const int NR = N / R;
  const int unit = N / (Ns * R);
  for (int o = threadIdx.x; o < N; o += blockDim.x) {
    const int j = o % NR, q = o / NR;
    const int k = j % Ns;
    const int e1 = ((k + q * Ns) * unit) % N;  // exponent step per r
    double2 acc = {0.0, 0.0};
    int e = 0;
    for (int r = 0; r < R; ++r) {
      double2 w = tw[e];
      if (sign > 0) w.y = -w.y;
      const double2 a = in[j + r * NR];
      acc.x = fma(a.x, w.x, fma(-a.y, w.y, acc.x));
      acc.y = fma(a.x, w.y, fma(a.y, w.x, acc.y));
      e += e1;
      if (e >= N) e -= N;
    }
    store((j / Ns) * Ns * R + k + q * Ns, acc);
  }
}

__device__ inline void fft_pass_dft(const double2* __restrict__ tw, const double2* in, double2* out, int N, int Ns,
                                    int R, double sign) {
  fft_pass_dft_to(tw, in, N, Ns, R, sign, [out](int i, double2 v) { out[i] = v; });
}

// the last pass of an FFT (span N / R), its outputs handed to store(i, value)
// instead of written to LDS
template <class Store>
__device__ inline void fft_last_pass_to(const FFTPlanDev& P, const double2* __restrict__ tw, const double2* in,
                                        double sign, Store&& store) {
  const int R = P.radix[P.nrad - 1];
  const int Ns = P.N / R;
  if (R == 4) fft_pass_to<4>(tw, in, P.N, Ns, sign, store);
  else if (R == 2) fft_pass_to<2>(tw, in, P.N, Ns, sign, store);
  else if (R == 3) fft_pass_to<3>(tw, in, P.N, Ns, sign, store);
  else fft_pass_dft_to(tw, in, P.N, Ns, R, sign, store);
}

// FFT of length N held in buf0 (ping-pong with buf1); sign -1 forward,
// +1 inverse (unnormalised).  Runs passes pass0 .. pass_end - 1 (all by
// default; pass0 > 0 takes the span Ns0 the earlier passes reached).
// Returns the buffer holding the result.
__device__ double2* fft_lds(const FFTPlanDev& P, const double2* __restrict__ tw, double2* buf0,
                            double2* buf1, double sign, int pass0 = 0, int Ns0 = 1, int pass_end = -1) {
  const int N = P.N;
  int Ns = Ns0;
  double2* in = buf0;
  double2* out = buf1;
  const int pe = pass_end < 0 ? P.nrad : pass_end;
  for (int pass = pass0; pass < pe; ++pass) {
    const int R = P.radix[pass];
    if (R == 4) fft_pass<4>(tw, in, out, N, Ns, sign);
    else if (R == 2) fft_pass<2>(tw, in, out, N, Ns, sign);
    else if (R == 3) fft_pass<3>(tw, in, out, N, Ns, sign);
    else fft_pass_dft(tw, in, out, N, Ns, R, sign);
    __syncthreads();
    double2* t = in;
    in = out;
    out = t;
    Ns *= R;
  }
  return in;
}

// passes pass0 .. pass_end - 1 in place on buf (fft_one_buffer plans)
__device__ void fft_lds_inplace(const FFTPlanDev& P, const double2* __restrict__ tw, double2* buf, double sign,
                                int pass0, int Ns0, int pass_end) {
  const int N = P.N;
  int Ns = Ns0;
  for (int pass = pass0; pass < pass_end; ++pass) {
    const int R = P.radix[pass];
    const int kb = (N / R + kFftThreads - 1) / kFftThreads;
    if (R == 4) {
      if (kb <= 1) fft_pass_inplace<4, 1>(tw, buf, N, Ns, sign);
      else fft_pass_inplace<4, 2>(tw, buf, N, Ns, sign);
    } else if (R == 2) {
      if (kb <= 1) fft_pass_inplace<2, 1>(tw, buf, N, Ns, sign);
      else if (kb <= 2) fft_pass_inplace<2, 2>(tw, buf, N, Ns, sign);
      else fft_pass_inplace<2, 4>(tw, buf, N, Ns, sign);
    } else {
      if (kb <= 1) fft_pass_inplace<3, 1>(tw, buf, N, Ns, sign);
      else if (kb <= 2) fft_pass_inplace<3, 2>(tw, buf, N, Ns, sign);
      else fft_pass_inplace<3, 3>(tw, buf, N, Ns, sign);
    }
    Ns *= R;
  }
}

// passes pass0 .. pass_end - 1 of the data in cur (lds or lds + N): in place
// for fft_one_buffer plans, else ping-pong with the other half of lds.
// Returns the buffer holding the result.
__device__ double2* fft_run(const FFTPlanDev& P, const double2* __restrict__ tw, double2* lds, double2* cur,
                            double sign, int pass0, int Ns0, int pass_end) {
  if (fft_one_buffer(P)) {
    fft_lds_inplace(P, tw, cur, sign, pass0, Ns0, pass_end);
    return cur;
  }
  return fft_lds(P, tw, cur, cur == lds ? lds + P.N : lds, sign, pass0, Ns0, pass_end);
}

// real rings -> half spectrum F[f][t][m] (m < L), scaled by pi / L; with
// psplit each ring's orders are stored even m first, then odd m (the
// resampling reads one parity at a time)
struct KFftR2c {
  struct Args {
    FFTPlanDev P;
    const double2* tw;
    const double* f;
    int L;
    double2* F;
    int psplit;  // 1: each ring's orders parity-split (even m, then odd m)
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

__device__ __forceinline__ void KFftR2c::run(const Args& a, int bid) {
  const FFTPlanDev& P = a.P;
  const double2* __restrict__ tw = a.tw;
  const double* __restrict__ f = a.f;
  const int L = a.L;
  double2* __restrict__ F = a.F;
  extern __shared__ double2 lds[];
  double2* b0 = lds;
  double2* b1 = lds + P.N;
  const int N = P.N;
  int pairs = (L + 2) / 2;
  int fi = bid / pairs, pr = bid % pairs;
  int ta = 2 * pr, tb = 2 * pr + 1;
  bool hasb = tb <= L;
  const double* ra = f + ((int64_t)fi * (L + 1) + ta) * N;
  const double* rb = f + ((int64_t)fi * (L + 1) + (hasb ? tb : ta)) * N;
  double2* Z;
  if (P.radix[0] == 4) {
    // the first radix-4 pass (span 1: no twiddles) straight from the rings
    double2* dst = fft_one_buffer(P) ? b0 : b1;
    const int NR = N / 4;
    for (int j = threadIdx.x; j < NR; j += blockDim.x) {
      double2 v[4], y[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = double2{ra[j + r * NR], hasb ? rb[j + r * NR] : 0.0};
      bfly4(v, y, -1.0);
#pragma unroll
      for (int r = 0; r < 4; ++r) dst[4 * j + r] = y[r];
    }
    __syncthreads();
    Z = fft_run(P, tw, lds, dst, -1.0, 1, 4, P.nrad);
  } else {
    for (int n = threadIdx.x; n < N; n += blockDim.x) b0[n] = double2{ra[n], hasb ? rb[n] : 0.0};
    __syncthreads();
    Z = fft_run(P, tw, lds, b0, -1.0, 0, 1, P.nrad);
  }
  const double sc = M_PI / L;
  const int half = (L + 1) / 2;
  for (int m = threadIdx.x; m < L; m += blockDim.x) {
    double2 zk = Z[m], zn = Z[(N - m) % N];
    double2 fa = double2{0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y)};
    double2 fb = double2{0.5 * (zk.y + zn.y), -0.5 * (zk.x - zn.x)};
    const int mi = a.psplit ? (m & 1) * half + (m >> 1) : m;
    F[((int64_t)fi * (L + 1) + ta) * L + mi] = double2{sc * fa.x, sc * fa.y};
    if (hasb) F[((int64_t)fi * (L + 1) + tb) * L + mi] = double2{sc * fb.x, sc * fb.y};
  }
}

__global__ __launch_bounds__(256) void k_fft_r2c(KFftR2c::Args a) { KFftR2c::run(a, blockIdx.x); }

// half spectrum F[f][t][m] -> real rings (irfft * 2L)
// RF (rings fastest): the coefficients come as G[f][m][t] (the int8
// synthesis writes whole rings of one order contiguously) instead of
// F[f][t][m]; element m of ring t is then G[(f L + m)(L + 1) + t].  Blocks
// are mapped so that four consecutive ring pairs (the eight rings of a
// 128-byte line of G) run on one XCD: block b runs on XCD b % 8 and its k-th
// block there (k = b / 8) is pair 4 (8 (k / 4) + b % 8) + k % 4.
template <bool RF>
struct KFftC2r {
  struct Args {
    FFTPlanDev P;
    const double2* tw;
    const double2* F;
    int L;
    double* f;
    int gx;  // RF: blocks per field (the grid's x extent)
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

template <bool RF>
__device__ __forceinline__ void KFftC2r<RF>::run(const Args& a, int bid) {
  const FFTPlanDev& P = a.P;
  const double2* __restrict__ tw = a.tw;
  const double2* __restrict__ F = a.F;
  const int L = a.L;
  double* __restrict__ f = a.f;
  extern __shared__ double2 lds[];
  double2* b0 = lds;
  double2* b1 = lds + P.N;
  const int N = P.N;
  int pairs = (L + 2) / 2;
  int fi, pr;
  if constexpr (RF) {
    const int b = bid % a.gx, k = b >> 3;
    fi = bid / a.gx;
    pr = 4 * (8 * (k >> 2) + (b & 7)) + (k & 3);
    if (pr >= pairs) return;
  } else {
    fi = bid / pairs;
    pr = bid % pairs;
  }
  int ta = 2 * pr, tb = 2 * pr + 1;
  bool hasb = tb <= L;
  // element m of ring a / b: Fa[m * st], Fb[m * st]
  const int64_t st = RF ? (int64_t)L + 1 : 1;
  const double2* Fa = RF ? F + (int64_t)fi * L * (L + 1) + ta : F + ((int64_t)fi * (L + 1) + ta) * L;
  const double2* Fb = RF ? F + (int64_t)fi * L * (L + 1) + (hasb ? tb : ta)
                         : F + ((int64_t)fi * (L + 1) + (hasb ? tb : ta)) * L;
  if constexpr (RF) {
    // each coefficient read once (strided by the ring count): element m fills
    // n = m and its conjugate n = N - m
    for (int n = threadIdx.x; n <= L; n += blockDim.x) {
      if (n == L) {
        b0[L] = double2{0.0, 0.0};
        continue;
      }
      const double2 a = Fa[n * st];
      const double2 b = hasb ? Fb[n * st] : double2{0.0, 0.0};
      if (n == 0) {
        b0[0] = double2{a.x, b.x};
      } else {
        b0[n] = double2{a.x - b.y, a.y + b.x};      // a + i b
        b0[N - n] = double2{a.x + b.y, b.x - a.y};  // conj(a) + i conj(b)
      }
    }
  }
  // element n of the complex sequence xa + i xb (conjugate-symmetric fill)
  auto xval = [&](int n) -> double2 {
    double2 xa, xb;
    if (n == 0) {
      xa = double2{Fa[0].x, 0.0};
      xb = double2{hasb ? Fb[0].x : 0.0, 0.0};
    } else if (n < L) {
      xa = Fa[n * st];
      xb = hasb ? Fb[n * st] : double2{0.0, 0.0};
    } else if (n == L) {
      xa = double2{0.0, 0.0};
      xb = double2{0.0, 0.0};
    } else {
      double2 a = Fa[(N - n) * st];
      xa = double2{a.x, -a.y};
      if (hasb) {
        double2 b = Fb[(N - n) * st];
        xb = double2{b.x, -b.y};
      } else {
        xb = double2{0.0, 0.0};
      }
    }
    return double2{xa.x - xb.y, xa.y + xb.x};  // xa + i xb
  };
  // the last pass stores the real rings itself (element n of the pass's
  // output: ring ta gets its real part, ring tb its imaginary part)
  const bool direct0 = !RF && P.radix[0] == 4;
  const int pend = P.nrad - 1 >= (direct0 ? 1 : 0) ? P.nrad - 1 : P.nrad;
  double2* Z;
  if (direct0) {
    // the first radix-4 pass (span 1: no twiddles) straight from the
    // spectrum into LDS: fft_pass<4>'s arithmetic without staging the input
    double2* dst = fft_one_buffer(P) ? b0 : b1;
    const int NR = N / 4;
    for (int j = threadIdx.x; j < NR; j += blockDim.x) {
      double2 v[4], y[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = xval(j + r * NR);
      bfly4(v, y, +1.0);
#pragma unroll
      for (int r = 0; r < 4; ++r) dst[4 * j + r] = y[r];
    }
    __syncthreads();
    Z = fft_run(P, tw, lds, dst, +1.0, 1, 4, pend);
  } else {
    if constexpr (!RF) {
      for (int n = threadIdx.x; n < N; n += blockDim.x) b0[n] = xval(n);
    }
    __syncthreads();
    Z = fft_run(P, tw, lds, b0, +1.0, 0, 1, pend);
  }
  double* ra = f + ((int64_t)fi * (L + 1) + ta) * N;
  double* rb = f + ((int64_t)fi * (L + 1) + (hasb ? tb : ta)) * N;
  if (pend < P.nrad) {
    fft_last_pass_to(P, tw, Z, +1.0, [=](int n, double2 v) {
      ra[n] = v.x;
      if (hasb) rb[n] = v.y;
    });
  } else {
    for (int n = threadIdx.x; n < N; n += blockDim.x) {
      ra[n] = Z[n].x;
      if (hasb) rb[n] = Z[n].y;
    }
  }
}

template <bool RF>
__global__ __launch_bounds__(256) void k_fft_c2r(typename KFftC2r<RF>::Args a) {
  KFftC2r<RF>::run(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Legendre recurrence
//   Pbar_lm = a_lm (x Pbar_{l-1,m} - b_lm Pbar_{l-2,m}),
//   a_lm = sqrt((4l^2-1)/(l^2-m^2)), b_lm = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1)),
// evaluated as (a x) P_{l-1} - (a b) P_{l-2} from the tabulated pairs
// (a_lm, a_lm b_lm): one FMA on the dependency chain per degree.  Values are
// carried as mantissa * 2^(600 s); the scale check (s < 0 and |P| > 2^300 ->
// multiply by 2^-600, s += 1) runs once per chunk of degrees, and values
// still at s < 0 (below ~2^-280) count as zero.
// ---------------------------------------------------------------------------
__device__ inline void leg_step(double a, double ab, double x, double& pp, double& pc) {
  const double pn = fma(a * x, pc, -ab * pp);
  pp = pc;
  pc = pn;
}
__device__ inline void leg_rescale(double& pp, double& pc, int& s) {
  const bool c = (s < 0) & (fabs(pc) > 0x1p300);
  const double f = c ? 0x1p-600 : 1.0;
  pc *= f;
  pp *= f;
  s += c ? 1 : 0;
}

// value of lane `src` (a compile-time constant) of every lane's v, as a
// wave-uniform (scalar register) double
__device__ inline double readlane_d(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// ---------------------------------------------------------------------------
// inverse Legendre on FP64 matrix cores: F[f][t][m] for all rings from
// flm [lm][nf].  One wave per (order m, block of 64 northern rings), blocks
// ordered by ascending m so the longest recurrences start first.  Rings
// t >= 1 only for m > 0 (Pbar_lm vanishes at the poles; those two rings are
// written as zeros).  Per chunk of 8 degrees every lane advances the
// recurrence of its ring and stores Pbar into an LDS tile [degree][ring];
// the contraction over degrees is then v_mfma_f64_16x16x4f64 with
//   A (16 rings x 4 degrees of one parity of l - m)  read from the tile,
//   B (4 degrees x 16 columns = 8 fields x re/im)     read from flm,
// one accumulator per (ring tile, parity, column tile).  The coefficients
// reach each lane as its own element (no broadcast), loaded one chunk
// ahead; the recurrence pairs of the order are staged in LDS at block
// start; the recurrence of chunk c+1 is issued under the matrix
// instructions of chunk c.  Even / odd l - m sums give ring t and its
// mirror L - t.
// Operand layouts of the f64 16x16x4 instruction (measured on gfx950):
//   A: lane l holds A[l % 16][l / 16]    B: lane l holds B[l / 16][l % 16]
//   D: lane l, element r holds D[4 r + l / 16][l % 16]
// ---------------------------------------------------------------------------
typedef double dx4 __attribute__((ext_vector_type(4)));
constexpr int kInvRings = 64;  // rings per wave
constexpr int kInvDeg = 8;     // degrees per chunk (4 of each parity)

// MIXED: FP32 matrix cores (v_mfma_f32_16x16x4f32, twice the FP64 rate) on
// FP32 copies of Pbar and of the coefficients; the recurrence stays FP64.
// The FP32 accumulators cover kMixedFlush chunks (32 degrees of each parity)
// and are then added into FP64 accumulators, so the FP32 rounding of the
// accumulation is bounded by 32-term partial sums (compensated FP64
// correction).  The f32 instruction's D layout differs from the f64 one:
// lane l, element r holds D[4 (l / 16) + r][l % 16].
typedef float fx4 __attribute__((ext_vector_type(4)));
constexpr int kMixedFlush = 8;

template <int NT, bool MIXED>
struct KLegInverse {
  struct Args {
    int L, nf, f0, nfp;
    const int2* tasks;
    const int* toff;
    const double* x_t;
    const double* dP;
    const int* dS;
    const double2* rec;
    const double* flm;
    double* F;
  };
  static constexpr int kThreads = 64;
  static constexpr int kMinEU = 2;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

template <int NT, bool MIXED>
__device__ __forceinline__ void KLegInverse<NT, MIXED>::run(const Args& a, int bid) {
  const int L = a.L, nf = a.nf, f0 = a.f0, nfp = a.nfp;
  const int2* __restrict__ tasks = a.tasks;
  const int* __restrict__ toff = a.toff;
  const double* __restrict__ x_t = a.x_t;
  const double* __restrict__ dP = a.dP;
  const int* __restrict__ dS = a.dS;
  const double2* __restrict__ rec = a.rec;
  const double* __restrict__ flm = a.flm;
  double* __restrict__ F = a.F;
  __shared__ double Pt[kInvDeg][kInvRings + 2];
  extern __shared__ double2 rsh[];  // recurrence pairs of this order, degrees m .. L-1 (+ padding)
  const int k1 = toff[bid + 1];
  for (int k = toff[bid]; k < k1; ++k) {
  __syncthreads();  // the previous task's reads of rsh and Pt are done
  const int m = tasks[k].x, rb = tasks[k].y;
  const int lane = threadIdx.x;
  const int nh = L / 2 + 1;
  const int tfirst = m == 0 ? 0 : 1;
  const int nr = nh - tfirst;
  if (m > 0 && rb == 0 && lane < nfp) {
    const double2 z = {0.0, 0.0};
    double2* F2 = reinterpret_cast<double2*>(F);
    F2[((int64_t)(f0 + lane) * (L + 1) + 0) * L + m] = z;
    F2[((int64_t)(f0 + lane) * (L + 1) + L) * L + m] = z;
  }
  const int r = rb * kInvRings + lane;
  const bool valid = r < nr;
  const int t = tfirst + (valid ? r : 0);
  const double x = x_t[t];
  double pp = 0.0, pc = dP[(int64_t)m * (L + 1) + t];
  int s = dS[(int64_t)m * (L + 1) + t];
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  const int bk = lane >> 4, bj = lane & 15;

  // recurrence pairs of every degree of this order, staged once
  const int nch = (L - m + kInvDeg - 1) / kInvDeg;          // chunks of degrees
  const int nrs = (nch + (nch & 1) + 1) * kInvDeg;           // padded to whole chunk pairs
  for (int i = lane; i < nrs; i += 64) rsh[i] = rec[lm0 + min(i, L - 1 - m)];
  __syncthreads();
  // B operands of one chunk from clamped (always valid) addresses.  No
  // masking: the tile rows of degrees >= L are zero, and columns of fields
  // >= nfp are never stored.
  auto load = [&](int l0, double (&b)[2][NT]) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int l = min(l0 + 2 * bk + p, L - 1);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const int col = 16 * nt + bj, f = min(col >> 1, nfp - 1);
        b[p][nt] = flm[((lm0 + (l - m)) * nf + f0 + f) * 2 + (col & 1)];
      }
    }
  };
  // recurrence of the chunk starting at degree l0 into the tile
  auto recur = [&](int l0, bool first) {
    leg_rescale(pp, pc, s);
    const bool live = (s == 0) & valid;
#pragma unroll
    for (int i = 0; i < kInvDeg; ++i) {
      const double2 ab = rsh[l0 - m + i];
      if (!(first && i == 0)) leg_step(ab.x, ab.y, x, pp, pc);
      Pt[i][lane] = (live & (l0 + i < L)) ? pc : 0.0;
    }
  };
  dx4 acc[4][2][NT];
  fx4 acc32[4][2][NT];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        acc[mt][p][nt] = dx4{0.0, 0.0, 0.0, 0.0};
        acc32[mt][p][nt] = fx4{0.0f, 0.0f, 0.0f, 0.0f};
      }
  double a[4][2];
  auto readA = [&]() {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int p = 0; p < 2; ++p) a[mt][p] = Pt[2 * bk + p][16 * mt + bj];
  };
  int nmix = 0;
  auto mma = [&](const double (&b)[2][NT]) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          if constexpr (MIXED)
            acc32[mt][p][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a[mt][p], (float)b[p][nt],
                                                                    acc32[mt][p][nt], 0, 0, 0);
          else
            acc[mt][p][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mt][p], b[p][nt], acc[mt][p][nt], 0, 0, 0);
        }
    if constexpr (MIXED) {
      if (++nmix == kMixedFlush) {
        nmix = 0;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
          for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
#pragma unroll
              for (int e = 0; e < 4; ++e) acc[mt][p][nt][e] += (double)acc32[mt][p][nt][e];
              acc32[mt][p][nt] = fx4{0.0f, 0.0f, 0.0f, 0.0f};
            }
      }
    }
  };
  // Two chunks per iteration with no exit in between (an odd chunk count is
  // padded with a chunk whose tile rows and coefficients are zero), so the
  // operand loads stay one chunk ahead of their use.
  double bA[2][NT], bB[2][NT];
  load(m, bA);
  recur(m, true);
  __syncthreads();
  load(m + kInvDeg, bB);
  for (int c = 0; c < nch; c += 2) {
    const int l0 = m + c * kInvDeg;
    readA();
    mma(bA);
    recur(l0 + kInvDeg, false);
    __syncthreads();
    load(l0 + 2 * kInvDeg, bA);
    readA();
    mma(bB);
    recur(l0 + 2 * kInvDeg, false);
    __syncthreads();
    load(l0 + 3 * kInvDeg, bB);
  }
  if constexpr (MIXED) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
          for (int e = 0; e < 4; ++e) acc[mt][p][nt][e] += (double)acc32[mt][p][nt][e];
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int col = 16 * nt + bj, f = col >> 1;
      if (f >= nfp) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int rr = rb * kInvRings + 16 * mt + (MIXED ? 4 * bk + e : 4 * e + bk);
        if (rr >= nr) continue;
        const int tr = tfirst + rr;
        const double ev = acc[mt][0][nt][e], od = acc[mt][1][nt][e];
        F[(((int64_t)(f0 + f) * (L + 1) + tr) * L + m) * 2 + (col & 1)] = ev + od;
        if (L - tr != tr) F[(((int64_t)(f0 + f) * (L + 1) + (L - tr)) * L + m) * 2 + (col & 1)] = ev - od;
      }
    }
  }  // task
}

template <int NT, bool MIXED>
__global__ __launch_bounds__(64, 2) void k_leg_inverse(typename KLegInverse<NT, MIXED>::Args a) {
  KLegInverse<NT, MIXED>::run(a, blockIdx.x);
}

// Static schedule of k_leg_inverse: the (order, 64-ring block) tasks,
// longest first, each to the least loaded of W waves (load: the task's
// chunks of 8 degrees, padded to pairs, plus two for the staging and the
// stores), W = the blocks that are resident at once (at most 2 per SIMD,
// fewer where the staged recurrence pairs fill the LDS).  Orders have L - m
// degrees: with one wave per task, SIMDs that start with two long orders
// carry twice the mean load (12-field synthesis at L=512: 90 -> 87 us; the
// matrix cores stay busy ~35% of the kernel, the rest is the recurrence and
// LDS latency of each wave's chunk chain).
static int build_inverse_tasks(cmm_sht_plan* p) {
  const int L = p->L, nh = L / 2 + 1;
  std::vector<std::pair<int, int2>> tasks;  // (load, (m, rb))
  for (int m = 0; m < L; ++m) {
    const int nr = nh - (m == 0 ? 0 : 1);
    const int nch = (L - m + kInvDeg - 1) / kInvDeg;
    const int load = 2 * ((nch + 1) / 2) + 2;
    for (int rb = 0; rb * kInvRings < nr; ++rb) tasks.push_back({load, int2{m, rb}});
  }
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
    ncu = 256;
  const size_t lds = (size_t)((L + kInvDeg - 1) / kInvDeg + 2) * kInvDeg * sizeof(double2) +
                     sizeof(double) * kInvDeg * (kInvRings + 2);
  const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / lds));
  const int W = std::max(1, std::min<int>(ncu * per_cu, (int)tasks.size()));
  std::stable_sort(tasks.begin(), tasks.end(), [](const std::pair<int, int2>& a, const std::pair<int, int2>& b) {
    return a.first > b.first;
  });
  typedef std::pair<int64_t, int> Slot;  // (load, wave)
  std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
  for (int w = 0; w < W; ++w) heap.push(Slot{0, w});
  std::vector<std::vector<int2>> per(W);
  for (const auto& t : tasks) {
    const Slot sl = heap.top();
    heap.pop();
    per[sl.second].push_back(t.second);
    heap.push(Slot{sl.first + t.first, sl.second});
  }
  std::vector<int2> flat;
  std::vector<int> off(W + 1, 0);
  for (int w = 0; w < W; ++w) {
    off[w + 1] = off[w] + (int)per[w].size();
    flat.insert(flat.end(), per[w].begin(), per[w].end());
  }
  p->inv_waves = W;
  int rc = upload(&p->inv_tasks, flat.data(), flat.size());
  if (!rc) rc = upload(&p->inv_off, off.data(), off.size());
  return rc;
}

// ---------------------------------------------------------------------------
// Split-BF16 synthesis: fixed-point slicing of both operands (the Ozaki
// scheme).  Each operand is a power of two times S base-256 digits,
//   Pbar_lm(x_t) = 2^(eA - 8) sum_i a_i,   flm[col] = 2^(eB[m][col] - 8) sum_j b_j,
// with a_1, b_1 integers in [-256, 256] and a_i = d_i 2^(-8 (i - 1)),
// |d_i| <= 128 (balanced digits, rounded to nearest at every step).  A digit
// has at most 8 significant bits, so it is exact in bf16, and a product
// a_i b_j is an exact multiple of 2^(-8 (i + j - 2)) below 2^16 in magnitude.
// The products are summed on bf16 matrix cores into FP32 accumulators by
// weight: i + j = 2, 3, 4, 5 (multiples of 1, 2^-8, 2^-16, 2^-24 with
// per-degree magnitudes below 2^16, 2^9, 1.25 and 2^-8 + 2^-9) and, from five
// digits on, i + j >= 6.  Over at most 192 degrees of one parity the first
// three stay below 2^24 units, so their FP32 sums are exact, the fourth up to
// 170 degrees (beyond, it rounds at half a unit, 2^-49 of the scale); the
// last carries the FP32 rounding of a quantity 2^-32 below the result.  (With
// i + j = 5 and the smaller weights in one accumulator its rounding measured
// ~7e-13 of the scale in a CPU emulation: the floor of five and more digits.)
// Products with i + j > S + 1 are left out (each below 2^(-8 S) of the
// scale).  Every 192 degrees per parity the sums are scaled and added into
// the FP64 output.  eA bounds |Pbar| over all degrees (orthonormal functions:
// |Pbar_lm| <= sqrt((2 l + 1) / 4 pi)); eB is the exponent of the largest
// coefficient of each order and column.  The recurrence stays FP64.
// ---------------------------------------------------------------------------
typedef short bf4 __attribute__((ext_vector_type(4)));
constexpr int kOzRings = 64;  // rings per wave
constexpr int kOzFlush = 12;  // chunks between FP64 flushes (192 degrees per parity)
// FP32 accumulators by digit weight for S digits
template <int S>
constexpr int oz_acc() { return S >= 5 ? 5 : 4; }
constexpr int kOzRow = kOzJ / 2 + 2;  // dwords per tile row (16 bf16 + padding)

// bf16 bit pattern of a float whose significand has at most 8 bits (exact)
__device__ inline uint32_t bf16_bits(float f) { return __float_as_uint(f) >> 16; }

// the S digits of v (|v| <= 1) multiplied by their weights, as bf16 bits
template <int S>
__device__ inline void oz_digits(double v, uint32_t (&d)[S]) {
  double t = v * 256.0;
  double s = rint(t);
  d[0] = bf16_bits((float)s);
  double r = t - s;  // exact, |r| <= 1/2
  double w = 1.0;
#pragma unroll
  for (int k = 1; k < S; ++k) {
    t = r * 256.0;
    s = rint(t);
    r = t - s;
    w *= 0x1p-8;
    d[k] = bf16_bits((float)(s * w));
  }
}

// exponents eB[m][col] (stride 32) with max_l |flm| < 2^eB for the columns
// col < ncols (field f0 + col / 2, real / imaginary part) of order
// m = blockIdx.x; 32 partial maxima per column, loads unrolled by 4
__global__ __launch_bounds__(1024) void k_oz_exponents(int L, int nf, int f0, int ncols,
                                                       const double* __restrict__ flm, int* __restrict__ eB) {
  __shared__ double mx[32][33];
  const int m = blockIdx.x;
  const int col = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  const int cc = col < ncols ? col : 0;
  const double* base = flm + (lm0 * nf + f0 + (cc >> 1)) * 2 + (cc & 1);
  const int64_t st = (int64_t)nf * 2;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int l = m + g;
  for (; l + 96 < L; l += 128) {
    a0 = fmax(a0, fabs(base[(l - m) * st]));
    a1 = fmax(a1, fabs(base[(l + 32 - m) * st]));
    a2 = fmax(a2, fabs(base[(l + 64 - m) * st]));
    a3 = fmax(a3, fabs(base[(l + 96 - m) * st]));
  }
  for (; l < L; l += 32) a0 = fmax(a0, fabs(base[(l - m) * st]));
  mx[g][col] = fmax(fmax(a0, a1), fmax(a2, a3));
  __syncthreads();
  if (g != 0 || col >= ncols) return;
  double amax = 0.0;
#pragma unroll
  for (int i = 0; i < 32; ++i) amax = fmax(amax, mx[i][col]);
  eB[(int64_t)m * 32 + col] = amax > 0.0 ? ilogb(amax) + 1 : 0;
}

// B operands: the digits of the coefficients of order m = blockIdx.x and
// chunk c = blockIdx.y (16 degrees of each parity), columns col < ncols,
// laid out [digit][((offJ[m] * 2 + p * nJ + c) * ncols + col) * 16 + jj] for
// degree l = m + p + 2 (16 c + jj) (zero for l >= L).  One thread per
// (parity, column, 4 degrees), the 4 degrees fastest (coalesced stores).
template <int S>
__global__ __launch_bounds__(256) void k_oz_coeffs(int L, int nf, int f0, int ncols, const int* __restrict__ offJ,
                                                   int64_t dstride, const double* __restrict__ flm,
                                                   const int* __restrict__ eB, uint16_t* __restrict__ Bd) {
  const int m = blockIdx.x, c = blockIdx.y;
  const int nJ = (L - m + 31) / 32;
  if (c >= nJ) return;
  const int j4 = threadIdx.x & 3, col = (threadIdx.x >> 2) & 31, p = threadIdx.x >> 7;
  if (col >= ncols) return;
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  const double sc = ldexp(1.0, -eB[(int64_t)m * 32 + col]);
  uint32_t d[4][S];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int l = m + p + 2 * (kOzJ * c + 4 * j4 + u);
    const double v = l < L ? flm[((lm0 + (l - m)) * nf + f0 + (col >> 1)) * 2 + (col & 1)] * sc : 0.0;
    oz_digits<S>(v, d[u]);
  }
  const int64_t at = (int64_t)offJ[m] * 2 * ncols * kOzJ + ((int64_t)(p * nJ + c) * ncols + col) * kOzJ + 4 * j4;
#pragma unroll
  for (int k = 0; k < S; ++k)
    *reinterpret_cast<uint2*>(Bd + k * dstride + at) = uint2{d[0][k] | (d[1][k] << 16), d[2][k] | (d[3][k] << 16)};
}

// The S digits of v * 256 (|v| < 1) as floats that are exact in bf16, given
// V = v 2^24, from two integers: H = rint(V) (|H| <= 2^24) and
// Lo = rint((V - H) 2^24) (|Lo| <= 2^23).  With a bias of 128 in its two low bytes, an integer
// X = d0 2^16 + d1 2^8 + d2 has bytes d1 + 128, d2 + 128 and d0 = (X + 0x8080)
// >> 16, so d1, d2 are in [-128, 127] and |d0| <= 256:
//   v 256 = d0 + d1 2^-8 + d2 2^-16 + e0 2^-24 + e1 2^-32 + e2 2^-40 (+ f 2^-48),
// the seventh digit f = rint(((V - H) 2^24 - Lo) 2^8) (|f| <= 128).
template <int S>
__device__ inline void oz_digits_fixed(double V, float (&d)[S]) {
  const double Hd = rint(V);
  const int hb = (int)Hd + 0x8080;
  d[0] = (float)(hb >> 16);
  if (S > 1) d[1] = fmaf((float)((hb >> 8) & 0xff), 0x1p-8f, -0x1p-1f);
  if (S > 2) d[2] = fmaf((float)(hb & 0xff), 0x1p-16f, -0x1p-9f);
  if (S > 3) {
    const double R = (V - Hd) * 0x1p24;
    const double Ld = rint(R);
    const int lb = (int)Ld + 0x8080;
    d[3] = (float)(lb >> 16) * 0x1p-24f;
    if (S > 4) d[4] = fmaf((float)((lb >> 8) & 0xff), 0x1p-32f, -0x1p-25f);
    if (S > 5) d[5] = fmaf((float)(lb & 0xff), 0x1p-40f, -0x1p-33f);
    if (S > 6) d[6] = (float)(rint((R - Ld) * 256.0) * 0x1p-48);
  }
}

// two floats exact in bf16 -> one dword of bf16 bits (lo in the low half)
__device__ inline uint32_t pack_bf16(float lo, float hi) {
  return (__float_as_uint(lo) >> 16) | (__float_as_uint(hi) & 0xffff0000u);
}

// inverse Legendre F[f][t][m] on bf16 matrix cores.  A block is two waves on
// the same 64 northern rings of one order: wave w sums the degrees with
// l - m = w (mod 2).  Both run the FP64 recurrence; per chunk of 32 degrees
// each slices its 16 values into a tile [digit][ring][degree] in LDS, then
// multiplies with the coefficient digits (32 x 32 x 16 per instruction,
// A = 32 rings x 16 degrees, B = 16 degrees x 32 columns = 16 fields re / im),
// loaded at the start of the chunk.  At a flush the two parity sums meet in
// LDS and give ring t (even + odd) and L - t (even - odd).  Blocks are mapped
// so that the ring blocks of one order run on one XCD (its L2 holds the
// coefficient digits), orders ascending.  (Overlapping the matrix
// instructions of one chunk with the recurrence of the next, with two tiles
// per wave, measured slower: one wave per SIMD.)
// Operand layouts (v_mfma_f32_32x32x16_bf16, measured): A lane l holds
// A[l % 32][8 (l / 32) + j], B lane l holds B[8 (l / 32) + j][l % 32],
// D lane l element r holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float fx16 __attribute__((ext_vector_type(16)));
constexpr int kOzRow2 = kOzJ / 2 + 4;  // dwords per tile row (16 bf16, 16-byte aligned rows)
// (six and more digits need the register file of one wave per SIMD)
template <int S>
__global__ __launch_bounds__(128, S >= 6 ? 1 : 2) void k_leg_inverse_oz(int L, int nf, int f0, int nfp, int nrb,
                                                                        int eA,
                                                           const double* __restrict__ x_t,
                                                           const double* __restrict__ dP, const int* __restrict__ dS,
                                                           const double2* __restrict__ rec,
                                                           const int* __restrict__ offJ, int64_t dstride,
                                                           const uint16_t* __restrict__ Bd,
                                                           const int* __restrict__ eB, double* __restrict__ F) {
  constexpr int kTile = 2 * S * kOzRings * kOzRow2;  // dwords
  constexpr int kComb = 2 * kOzRings * 33 * 2;       // dwords of the combine buffer
  __shared__ __align__(16) uint32_t lds[kTile > kComb ? kTile : kComb];
  extern __shared__ double2 rsh[];  // recurrence pairs of this order
  const int b = (int)blockIdx.x;
  const int slot = b >> 3;
  const int m = 8 * (slot / nrb) + (b & 7), rb = slot % nrb;
  if (m >= L) return;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nh = L / 2 + 1;
  const int tfirst = m == 0 ? 0 : 1;
  const int nr = nh - tfirst;
  if (rb * kOzRings >= nr) return;
  if (m > 0 && rb == 0 && threadIdx.x < nfp) {
    const double2 z = {0.0, 0.0};
    double2* F2 = reinterpret_cast<double2*>(F);
    F2[((int64_t)(f0 + threadIdx.x) * (L + 1) + 0) * L + m] = z;
    F2[((int64_t)(f0 + threadIdx.x) * (L + 1) + L) * L + m] = z;
  }
  const int r = rb * kOzRings + lane;
  const bool valid = r < nr;
  const int t = tfirst + (valid ? r : 0);
  const double x = x_t[t];
  double pp = 0.0, pc = dP[(int64_t)m * (L + 1) + t];
  int s = dS[(int64_t)m * (L + 1) + t];
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  const int nJ = (L - m + 31) / 32;
  for (int i = threadIdx.x; i < 32 * nJ; i += 128) rsh[i] = rec[lm0 + min(i, L - 1 - m)];
  const int hi = lane >> 5, lc = lane & 31;
  const int ncols = 2 * nfp;
  const int64_t bat = (int64_t)offJ[m] * 2 * ncols * kOzJ + (int64_t)min(lc, ncols - 1) * kOzJ + 8 * hi;
  constexpr int NQ = oz_acc<S>();
  fx16 acc[NQ][2];
  auto zero = [&]() {
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) acc[q][mt] = fx16{};
  };
  zero();
  const double scA = ldexp(1.0, 24 - eA);  // Pbar -> V = (Pbar / 2^eA) 2^24
  uint32_t* tile = lds + w * S * kOzRings * kOzRow2;
  double* comb = reinterpret_cast<double*>(lds);  // [parity][ring][33]
  auto flush = [&](bool first) {
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        double v = (double)acc[0][mt][e];
#pragma unroll
        for (int q = 1; q < NQ; ++q) v += (double)acc[q][mt][e];
        comb[(w * kOzRings + 32 * mt + 8 * (e >> 2) + 4 * hi + (e & 3)) * 33 + lc] = v;
      }
    zero();
    __syncthreads();
    double2* F2 = reinterpret_cast<double2*>(F);
    for (int e = threadIdx.x; e < kOzRings * nfp; e += 128) {
      const int row = e & (kOzRings - 1), f = e / kOzRings;
      const int rr = rb * kOzRings + row;
      if (rr >= nr) continue;
      const int tr = tfirst + rr;
      const double sr = ldexp(1.0, eA + eB[(int64_t)m * 32 + 2 * f] - 16);
      const double si = ldexp(1.0, eA + eB[(int64_t)m * 32 + 2 * f + 1] - 16);
      const double er = comb[row * 33 + 2 * f], ei = comb[row * 33 + 2 * f + 1];
      const double orr = comb[(kOzRings + row) * 33 + 2 * f], oi = comb[(kOzRings + row) * 33 + 2 * f + 1];
      const double2 va = {(er + orr) * sr, (ei + oi) * si}, vb = {(er - orr) * sr, (ei - oi) * si};
      const int64_t ia = ((int64_t)(f0 + f) * (L + 1) + tr) * L + m;
      F2[ia] = first ? va : double2{F2[ia].x + va.x, F2[ia].y + va.y};
      if (L - tr != tr) {
        const int64_t ib = ((int64_t)(f0 + f) * (L + 1) + (L - tr)) * L + m;
        F2[ib] = first ? vb : double2{F2[ib].x + vb.x, F2[ib].y + vb.y};
      }
    }
    __syncthreads();
  };
  bf8 bcur[S];
  __syncthreads();  // rsh staged
  for (int c = 0; c < nJ; ++c) {
    // coefficient digits of this chunk, consumed after the recurrence below
    const int64_t at = bat + (int64_t)(w * nJ + c) * ncols * kOzJ;
#pragma unroll
    for (int k = 0; k < S; ++k) bcur[k] = *reinterpret_cast<const bf8*>(Bd + k * dstride + at);
    // recurrence of degrees l0 .. l0 + 31; this wave keeps l0 + 2 jj + w
    leg_rescale(pp, pc, s);
    const double f = ((s == 0) & valid) ? scA : 0.0;
    const int l0 = m + 32 * c;
    float lo[S];
#pragma unroll
    for (int i = 0; i < 32; i += 2) {
      const double2 ab0 = rsh[l0 - m + i], ab1 = rsh[l0 - m + i + 1];
      if (!(c == 0 && i == 0)) leg_step(ab0.x, ab0.y, x, pp, pc);
      const double v0 = pc;
      leg_step(ab1.x, ab1.y, x, pp, pc);
      const double v = w == 0 ? v0 : pc;
      const int l = l0 + i + w;
      float d[S];
      oz_digits_fixed<S>(l < L ? v * f : 0.0, d);
      const int jj = i >> 1;
      if (jj & 1) {
#pragma unroll
        for (int k = 0; k < S; ++k) tile[(k * kOzRings + lane) * kOzRow2 + (jj >> 1)] = pack_bf16(lo[k], d[k]);
      } else {
#pragma unroll
        for (int k = 0; k < S; ++k) lo[k] = d[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < S; ++i) {
      bf8 av[2];
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
        av[mt] = *reinterpret_cast<const bf8*>(&tile[(i * kOzRings + 32 * mt + lc) * kOzRow2 + 4 * hi]);
#pragma unroll
      for (int j = 0; i + j < S; ++j) {
        const int q = i + j < NQ - 1 ? i + j : NQ - 1;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
          acc[q][mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[mt], bcur[j], acc[q][mt], 0, 0, 0);
      }
    }
    __syncthreads();
    if ((c + 1) % kOzFlush == 0 || c + 1 == nJ) flush(c < kOzFlush);
  }
}

// ---------------------------------------------------------------------------
// Int8 digit-table synthesis: the Ozaki scheme on the integer matrix cores
// (v_mfma_i32_32x32x32_i8), with the Legendre operand precomputed.
//   Pbar_lm(x_t) = 2^eT sum_{i<S} d_i 2^(-8 (i + 1)),
//   flm[col]     = 2^eB sum_{j<S} c_j 2^(-8 (j + 1)),   d_i, c_j in [-128, 127],
// balanced base-256 digits (round to nearest; a digit of +128 becomes -128
// with one carried into the digit above; the exponents leave |value| 2^-e
// <= 126/256, so the leading digit stays in range after a carry).  eT is one
// exponent per (order m, parity p of l - m, tile of 32 northern rings), eB
// one per chunk of 32 degrees and column.  Products d_i c_j are exact
// integers below 2^14 and are summed over a chunk's degrees in exact int32
// accumulators, one per weight k = i + j < S (at most (k + 1) 2^14 32 < 2^31),
// so the only errors are the digit truncation of both operands and the
// left-out products with i + j >= S, each ~2^(-8 S) of the operands' scales;
// the accumulators enter FP64 sums once per chunk.  The table holds the digits of
// every (m, p, tile, chunk of 32 degrees of parity p) in the instruction's A
// layout, chunks before the first one with a nonzero digit left out (their
// rings lie beyond the turning point of the order, where Pbar_lm is below
// 2^(-8 S - 1) of the tile's scale: ~18 % of the table at L = 512): S x 1 KB
// per chunk, ~150 MB at L = 512 with 5 digits, so a 12-field synthesis
// streams the table once instead of running the FP64 recurrence.  Rings
// t = 1 .. L/2 and their mirrors come from the table; the pole rings (only
// m = 0 is nonzero there) are summed in FP64 with Pbar_l0(+-1) =
// (+-1)^l sqrt((2 l + 1) / 4 pi).
// Operand layouts (v_mfma_i32_32x32x32_i8, measured): A lane l holds
// A[l % 32][16 (l / 32) + j], B lane l holds B[16 (l / 32) + j][l % 32]
// (j < 16, one byte each), D lane l element r holds
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
// ---------------------------------------------------------------------------
typedef int ix4 __attribute__((ext_vector_type(4)));
typedef int ix16 __attribute__((ext_vector_type(16)));

// the S balanced base-256 digits of V, |V| <= 126/256
template <int S>
__device__ inline void i8_digits(double V, int (&d)[S]) {
  double r = V;
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const double t = r * 256.0;
    const double q = rint(t);
    d[i] = (int)q;
    r = t - q;  // exact, |r| <= 1/2
  }
#pragma unroll
  for (int i = S - 1; i > 0; --i)
    if (d[i] == 128) {
      d[i] = -128;
      d[i - 1] += 1;
    }
}

// exponent e with |a| 2^-e <= 126/256 (0 for a = 0)
__host__ __device__ inline int i8_exponent(double a) {
  if (!(a > 0.0)) return 0;
  int e = ilogb(a) + 2;
  if (ldexp(a, -e) > 126.0 / 256.0) ++e;
  return e;
}

// Table set-up, pass 1: max |Pbar_lm| over each (order, parity, tile, chunk)
// -> cmax[((m 2 + p) nrt + rt) nkmax + kc].  Block (pair of tiles, order m);
// lane = ring t = 1 + 32 rt + lane % 32 of tile rt = 2 blockIdx.x + lane / 32,
// running the FP64 recurrence of the synthesis over all degrees.
__global__ __launch_bounds__(64) void k_i8_chunk_max(int L, int nrt, int nkmax, const double* __restrict__ x_t,
                                                     const double* __restrict__ dP, const int* __restrict__ dS,
                                                     const double2* __restrict__ rec, double* __restrict__ cmax) {
  const int m = blockIdx.y, lane = threadIdx.x;
  const int rt = 2 * blockIdx.x + (lane >> 5), lc = lane & 31;
  const int t = 1 + kI8Rings * rt + lc;
  const bool valid = rt < nrt && t <= L / 2;
  const int tt = valid ? t : 1;
  const double x = x_t[tt];
  double pp = 0.0, pc = dP[(int64_t)m * (L + 1) + tt];
  int s = dS[(int64_t)m * (L + 1) + tt];
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  for (int kc = 0; 2 * kI8Deg * kc < L - m; ++kc) {
    double mx0 = 0.0, mx1 = 0.0;
    for (int i = 0; i < 2 * kI8Deg; ++i) {
      const int l = m + 2 * kI8Deg * kc + i;
      if (l >= L) break;
      if (l > m) {
        const double2 ab = rec[lm0 + (l - m)];
        leg_step(ab.x, ab.y, x, pp, pc);
      }
      leg_rescale(pp, pc, s);
      const double v = ((s == 0) & valid) ? fabs(pc) : 0.0;
      if (i & 1) mx1 = fmax(mx1, v);
      else mx0 = fmax(mx0, v);
    }
    for (int o = 16; o > 0; o >>= 1) {
      mx0 = fmax(mx0, __shfl_xor(mx0, o));
      mx1 = fmax(mx1, __shfl_xor(mx1, o));
    }
    if (lc == 0 && rt < nrt) {
      cmax[(((int64_t)m * 2 + 0) * nrt + rt) * nkmax + kc] = mx0;
      cmax[(((int64_t)m * 2 + 1) * nrt + rt) * nkmax + kc] = mx1;
    }
  }
}

// Table set-up, pass 2: the digits.  Same lanes and recurrence as pass 1;
// lane (ring r of tile rt) writes bytes 16 h .. 16 h + 15 of the A operand
// row r for half h of each stored chunk and digit, the padded rings
// (t > L/2) and degrees (l >= L) as zeros.
template <int S>
__global__ __launch_bounds__(64) void k_i8_table(int L, int nrt, const double* __restrict__ x_t,
                                                 const double* __restrict__ dP, const int* __restrict__ dS,
                                                 const double2* __restrict__ rec, const int4* __restrict__ tdesc,
                                                 unsigned char* __restrict__ tab) {
  const int m = blockIdx.y, lane = threadIdx.x;
  const int rt = 2 * blockIdx.x + (lane >> 5), lc = lane & 31;
  if (rt >= nrt) return;  // second tile of an odd last pair
  const int t = 1 + kI8Rings * rt + lc;
  const bool valid = t <= L / 2;
  const int tt = valid ? t : 1;
  const double x = x_t[tt];
  double pp = 0.0, pc = dP[(int64_t)m * (L + 1) + tt];
  int s = dS[(int64_t)m * (L + 1) + tt];
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  const int4 td[2] = {tdesc[((int64_t)m * 2 + 0) * nrt + rt], tdesc[((int64_t)m * 2 + 1) * nrt + rt]};
  const int nk[2] = {((L - m + 1) / 2 + kI8Deg - 1) / kI8Deg, ((L - m) / 2 + kI8Deg - 1) / kI8Deg};
  for (int kc = 0; 2 * kI8Deg * kc < L - m; ++kc) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      uint32_t w[2][S][4];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int i = 0; i < S; ++i)
#pragma unroll
          for (int q = 0; q < 4; ++q) w[p][i][q] = 0u;
#pragma unroll
      for (int b = 0; b < 16; ++b)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const int l = m + 2 * (kI8Deg * kc + 16 * h + b) + p;
          if (l < L) {
            if (l > m) {
              const double2 ab = rec[lm0 + (l - m)];
              leg_step(ab.x, ab.y, x, pp, pc);
            }
            leg_rescale(pp, pc, s);
            const double v = ((s == 0) & valid) ? pc : 0.0;
            int d[S];
            i8_digits<S>(ldexp(v, -td[p].z), d);
#pragma unroll
            for (int i = 0; i < S; ++i) w[p][i][b >> 2] |= (uint32_t)(d[i] & 0xff) << (8 * (b & 3));
          }
        }
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        if (kc < td[p].x || kc >= nk[p]) continue;
        unsigned char* dst = tab + ((int64_t)(td[p].y + kc - td[p].x) * S) * 1024 + (lc + 32 * h) * 16;
#pragma unroll
        for (int i = 0; i < S; ++i)
          *reinterpret_cast<uint4*>(dst + i * 1024) = uint4{w[p][i][0], w[p][i][1], w[p][i][2], w[p][i][3]};
      }
    }
  }
}

// Coefficient digits, one wave per chunk block (the chunk (m, p, kc) of
// cdesc) for the columns col < ncols (field f0 + col / 2, real / imaginary
// part): lane = (column lane % 32, degrees 16 (lane / 32) .. + 15 of the
// chunk), so that byte 16 lane + j of digit i of the chunk's block is degree
// l = m + p + 2 (32 kc + 16 (lane / 32) + j) (zero for l >= L and col >=
// ncols): the B operand.  Each chunk and column has its own exponent
// eB[chunk][col] (max over its 32 degrees, |flm| 2^-eB <= 126/256).  The
// last nfp blocks sum the pole rings of field f0 + (block - nblk): order 0
// (Pbar_l0(+-1) = (+-1)^l sqrt((2 l + 1) / 4 pi), FP64) into G[f][0][0] and
// G[f][0][L]; the first chunk of each order m > 0 zeroes G[f][m][0] and
// G[f][m][L].
template <int S>
struct KI8Coeffs {
  struct Args {
    int L, nf, f0, ncols, nb;
    const int* cdesc;
    const double* flm;
    int* eB;
    unsigned char* Bd;
    double* F;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

template <int S>
__device__ __forceinline__ void KI8Coeffs<S>::run(const Args& a, int bid) {
  const int L = a.L, nf = a.nf, f0 = a.f0, ncols = a.ncols, nb = a.nb;
  const int* __restrict__ cdesc = a.cdesc;
  const double* __restrict__ flm = a.flm;
  int* __restrict__ eB = a.eB;
  unsigned char* __restrict__ Bd = a.Bd;
  double* __restrict__ F = a.F;
  const int64_t st = (int64_t)nf * 2;
  double2* G = reinterpret_cast<double2*>(F);
  const int nblk = (nb + 3) / 4;
  if (bid >= nblk) {
    // pole rings of one field: thread = degree group, both columns
    __shared__ double2 red[2][256];
    const int f = bid - nblk;
    double2 sn = {0.0, 0.0}, ss = {0.0, 0.0};
    for (int l0 = threadIdx.x; l0 < L; l0 += 256 * 4) {
      double2 x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = *reinterpret_cast<const double2*>(flm + ((int64_t)min(l0 + 256 * u, L - 1) * nf + f0 + f) * 2);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int l = l0 + 256 * u;
        const double c = l < L ? sqrt((2.0 * l + 1.0) / (4.0 * M_PI)) : 0.0;
        const double sg = (l & 1) ? -c : c;
        sn.x += c * x[u].x; sn.y += c * x[u].y;
        ss.x += sg * x[u].x; ss.y += sg * x[u].y;
      }
    }
    red[0][threadIdx.x] = sn;
    red[1][threadIdx.x] = ss;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          red[q][threadIdx.x].x += red[q][threadIdx.x + w].x;
          red[q][threadIdx.x].y += red[q][threadIdx.x + w].y;
        }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const int64_t gb = (int64_t)(f0 + f) * L * (L + 1);
      G[gb + 0] = red[0][0];
      G[gb + L] = red[1][0];
    }
    return;
  }
  const int ci = 4 * bid + ((int)threadIdx.x >> 6);
  if (ci >= nb) return;
  const int lane = threadIdx.x & 63, bc = lane & 31, h = lane >> 5;
  const int cd = cdesc[ci];
  const int m = cd >> 16, p = (cd >> 15) & 1, kc = cd & 0x7fff;
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  if (m > 0 && p == 0 && kc == 0 && lane < 2 * (ncols >> 1)) {
    const int64_t gb = ((int64_t)(f0 + (lane >> 1)) * L + m) * (L + 1);
    F[(gb + ((lane & 1) ? L : 0)) * 2 + 0] = 0.0;
    F[(gb + ((lane & 1) ? L : 0)) * 2 + 1] = 0.0;
  }
  const double* bb = flm + (lm0 * nf + f0 + (min(bc, ncols - 1) >> 1)) * 2 + (bc & 1);
  const int l0 = m + p + 2 * (kI8Deg * kc + 16 * h);
  double x[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) x[j] = bb[(min(l0 + 2 * j, L - 1) - m) * st];
  __builtin_amdgcn_sched_barrier(0);  // the 16 loads in flight together
  double mx = 0.0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    x[j] = (l0 + 2 * j < L && bc < ncols) ? x[j] : 0.0;
    mx = fmax(mx, fabs(x[j]));
  }
  mx = fmax(mx, __shfl_xor(mx, 32));
  const int e = i8_exponent(mx);
  if (h == 0) eB[(int64_t)ci * 32 + bc] = e;
  uint32_t word[S][4];
#pragma unroll
  for (int i = 0; i < S; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) word[i][q] = 0u;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    int d[S];
    i8_digits<S>(ldexp(x[j], -e), d);
#pragma unroll
    for (int i = 0; i < S; ++i) word[i][j >> 2] |= (uint32_t)(d[i] & 0xff) << (8 * (j & 3));
  }
  unsigned char* dst = Bd + (int64_t)ci * S * 1024 + lane * 16;
#pragma unroll
  for (int i = 0; i < S; ++i)
    *reinterpret_cast<uint4*>(dst + i * 1024) = uint4{word[i][0], word[i][1], word[i][2], word[i][3]};
}

template <int S>
__global__ __launch_bounds__(256) void k_i8_coeffs(typename KI8Coeffs<S>::Args a) { KI8Coeffs<S>::run(a, blockIdx.x); }

// Synthesis: one block of two waves per (order m, tile rt); wave w sums the
// degrees with l - m = w (mod 2).  Blocks are mapped so that groups of eight
// consecutive orders run on one XCD (its L2 holds the group's coefficient
// digits for all tiles), groups ascending (longest first).  Per chunk: the
// operands of the next chunk are issued before the S (S + 1) / 2 products of
// this one (buffer loads, fenced from the products for the scheduler) into S
// int32 sums, which then enter FP64 sums with the chunk's column scale.  The
// parity sums meet in LDS: E + O and E - O (times 2^eT) give ring t (wave 0)
// and L - t (wave 1), written rings fastest through a wave-local LDS
// transpose (whole 512-byte runs: written as F[f][t][m], one order per wave,
// the 8-byte stores reached DRAM as partial lines, 132 MB for 50 MB of
// output at L = 512, and the kernel took 67 instead of 38 us).
template <int S>
struct KLegInverseI8 {
  struct Args {
    int L, nf, f0, nfp, nrt;
    const int4* tdesc;
    const int* boff;
    const ix4* tab;
    const ix4* Bd;
    const int* eB;
    double* F;
  };
  static constexpr int kThreads = 128;
  static constexpr int kMinEU = S >= 7 ? 1 : 2;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

template <int S>
__device__ __forceinline__ void KLegInverseI8<S>::run(const Args& a, int bid) {
  const int L = a.L, nf = a.nf, f0 = a.f0, nfp = a.nfp, nrt = a.nrt;
  const int4* __restrict__ tdesc = a.tdesc;
  const int* __restrict__ boff = a.boff;
  const ix4* __restrict__ tab = a.tab;
  const ix4* __restrict__ Bd = a.Bd;
  const int* __restrict__ eB = a.eB;
  double* __restrict__ F = a.F;
  __shared__ double xch[2][16][66];  // (64 lanes; or a 32 x 33 tile)
  // block b runs on XCD b % 8; its k-th block there (k = b / 8) is order
  // 8 (8 (k / (8 nrt)) + b % 8) + k % 8 of tile (k / 8) % nrt
  const int b = bid;
  const int k = b >> 3;
  const int m = 8 * (8 * (k / (8 * nrt)) + (b & 7)) + (k & 7), rt = (k >> 3) % nrt;
  if (m >= L) return;
  const int p = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, hi = lane >> 5, lc = lane & 31;
  const int nk = ((L - m - p + 1) / 2 + kI8Deg - 1) / kI8Deg;
  const int4 td = tdesc[((int64_t)m * 2 + p) * nrt + rt];
  const int nc = nk - td.x;
  double v[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) v[e] = 0.0;
  if (nc > 0) {
    const int64_t cb = (int64_t)boff[2 * m + p] + td.x;  // first coefficient chunk block of the tile
    const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(tab + (int64_t)td.y * S * 64), (short)0, nc * S * 1024, 0x00020000);
    const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc((void*)(Bd + cb * S * 64), (short)0,
                                                                        nc * S * 1024, 0x00020000);
    const __amdgpu_buffer_rsrc_t rE = __builtin_amdgcn_make_buffer_rsrc((void*)(eB + cb * 32), (short)0, nc * 128,
                                                                        0x00020000);
    const int lo = lane * 16;
    // two operand sets, each loaded one chunk ahead of its products (no
    // register copies: a copy would wait for the load).  Chunks c >= nc lie
    // beyond the buffer resources' range: those loads return zeros, so the
    // loop runs whole chunk pairs with no condition for the compiler to sink
    // a load into.
    ix4 a0[S], b0[S], a1[S], b1[S];
    int e0, e1;
    auto load = [&](int c, ix4 (&a)[S], ix4 (&b)[S], int& ex) {
      const int o = c * S * 1024 + lo;
#pragma unroll
      for (int i = 0; i < S; ++i) {
        a[i] = __builtin_amdgcn_raw_buffer_load_b128(rA, o + i * 1024, 0, 2);  // streamed once: non-temporal
        b[i] = __builtin_amdgcn_raw_buffer_load_b128(rB, o + i * 1024, 0, 0);
      }
      ex = __builtin_amdgcn_raw_buffer_load_b32(rE, c * 128 + 4 * lc, 0, 0);
    };
    // one chunk: the S (S + 1) / 2 leading products into S exact int32 sums
    // by weight (the first product of each weight starts from zero), then
    // into the FP64 sums with the chunk's column scale 2^(eB - 16)
    auto mma = [&](const ix4 (&a)[S], const ix4 (&b)[S], int ex) {
      ix16 acc[S];
#pragma unroll
      for (int j = 0; j < S; ++j) acc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[0], b[j], ix16{}, 0, 0, 0);
#pragma unroll
      for (int i = 1; i < S; ++i)
#pragma unroll
        for (int j = 0; i + j < S; ++j)
          acc[i + j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i], b[j], acc[i + j], 0, 0, 0);
      const double sc = ldexp(1.0, ex - 16);
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        double x = (double)acc[S - 1][e];
#pragma unroll
        for (int k = S - 2; k >= 0; --k) x = fma(x, 0x1p-8, (double)acc[k][e]);
        v[e] = fma(x, sc, v[e]);
      }
    };
    load(0, a0, b0, e0);
    for (int c = 0; c < nc; c += 2) {
      load(c + 1, a1, b1, e1);
      __builtin_amdgcn_sched_barrier(0);
      mma(a0, b0, e0);
      __builtin_amdgcn_sched_barrier(0);
      load(c + 2, a0, b0, e0);
      __builtin_amdgcn_sched_barrier(0);
      mma(a1, b1, e1);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    v[e] = ldexp(v[e], td.z);
    xch[p][e][lane] = v[e];
  }
  __syncthreads();
  // wave 0: ring t = E + O, wave 1: ring L - t = E - O
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const double o = xch[p ^ 1][e][lane];
    v[e] = p == 0 ? v[e] + o : o - v[e];
  }
  __syncthreads();
  // transpose through this wave's half of xch: tile[row][col], row = ring
  double* tile = &xch[p][0][0];
#pragma unroll
  for (int e = 0; e < 16; ++e) tile[(8 * (e >> 2) + 4 * hi + (e & 3)) * 33 + lc] = v[e];
  // G[f][m][t] complex, rings fastest: 32 consecutive rings of one field
  // are 512 contiguous bytes (lane = ring, lanes 32..63 the next field)
  const int R = L / 2;
  double2* G = reinterpret_cast<double2*>(F);
  for (int i = lane; i < 32 * nfp; i += 64) {
    const int row = i & 31, f = i >> 5;
    const int t = 1 + kI8Rings * rt + row;
    if (t > R || (p == 1 && L - t == t)) continue;
    const int tt = p == 0 ? t : L - t;
    G[((int64_t)(f0 + f) * L + m) * (L + 1) + tt] = double2{tile[row * 33 + 2 * f], tile[row * 33 + 2 * f + 1]};
  }
}

template <int S>
__global__ __launch_bounds__(128, S >= 7 ? 1 : 2) void k_leg_inverse_i8(typename KLegInverseI8<S>::Args a) {
  KLegInverseI8<S>::run(a, blockIdx.x);
}

// Builds the digit table of S digits for plan p (once per mode change).
static int build_i8_table(cmm_sht_plan* p, int S) {
  if (p->i8S == S && p->i8_tab) return CMM_OK;
  (void)hipFree(p->i8_tab);
  (void)hipFree(p->i8_tdesc);
  p->i8_tab = nullptr;
  p->i8_tdesc = nullptr;
  p->i8S = 0;
  const int L = p->L, nrt = p->nrt;
  const int nkmax = ((L + 1) / 2 + kI8Deg - 1) / kI8Deg;
  const size_t ncm = (size_t)L * 2 * nrt * nkmax;
  double* cmax = nullptr;
  CMM_HIP(hipMalloc(&cmax, ncm * sizeof(double)));
  const dim3 grid((unsigned)((nrt + 1) / 2), (unsigned)L);
  hipError_t e = hipMemset(cmax, 0, ncm * sizeof(double));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_i8_chunk_max, grid, dim3(64), 0, 0, L, nrt, nkmax, p->x_t, p->dP_t, p->dS_t, p->rec, cmax);
    e = hipGetLastError();
  }
  std::vector<double> hc(ncm);
  if (e == hipSuccess) e = hipMemcpy(hc.data(), cmax, ncm * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(cmax);
  if (e != hipSuccess) return hip_check(e, "cmm_sht_set_legendre: int8 table maxima");
  std::vector<int4> td((size_t)L * 2 * nrt);
  int64_t blocks = 0;
  for (int m = 0; m < L; ++m)
    for (int par = 0; par < 2; ++par) {
      const int nk = ((L - m - par + 1) / 2 + kI8Deg - 1) / kI8Deg;
      for (int rt = 0; rt < nrt; ++rt) {
        const double* c = &hc[(((size_t)m * 2 + par) * nrt + rt) * nkmax];
        double tmax = 0.0;
        for (int kc = 0; kc < nk; ++kc) tmax = std::max(tmax, c[kc]);
        const int eT = i8_exponent(tmax);
        int k0 = nk;
        for (int kc = 0; kc < nk; ++kc)
          if (tmax > 0.0 && std::ldexp(c[kc], 8 * S - eT) >= 0.25) {
            k0 = kc;
            break;
          }
        td[((size_t)m * 2 + par) * nrt + rt] = int4{k0, (int)blocks, eT, 0};
        blocks += nk - k0;
      }
    }
  if (blocks >= ((int64_t)1 << 31) / S) return fail(CMM_E_RANGE, "cmm_sht_set_legendre: int8 table too large");
  int rc = upload(&p->i8_tdesc, td.data(), td.size());
  if (rc) return rc;
  CMM_HIP(hipMalloc(&p->i8_tab, std::max<int64_t>(1, blocks) * S * 1024));
  switch (S) {
    case 4: hipLaunchKernelGGL(k_i8_table<4>, grid, dim3(64), 0, 0, L, nrt, p->x_t, p->dP_t, p->dS_t, p->rec, p->i8_tdesc, p->i8_tab); break;
    case 5: hipLaunchKernelGGL(k_i8_table<5>, grid, dim3(64), 0, 0, L, nrt, p->x_t, p->dP_t, p->dS_t, p->rec, p->i8_tdesc, p->i8_tab); break;
    case 6: hipLaunchKernelGGL(k_i8_table<6>, grid, dim3(64), 0, 0, L, nrt, p->x_t, p->dP_t, p->dS_t, p->rec, p->i8_tdesc, p->i8_tab); break;
    case 7: hipLaunchKernelGGL(k_i8_table<7>, grid, dim3(64), 0, 0, L, nrt, p->x_t, p->dP_t, p->dS_t, p->rec, p->i8_tdesc, p->i8_tab); break;
    default: hipLaunchKernelGGL(k_i8_table<8>, grid, dim3(64), 0, 0, L, nrt, p->x_t, p->dP_t, p->dS_t, p->rec, p->i8_tdesc, p->i8_tab); break;
  }
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_check(e, "cmm_sht_set_legendre: int8 table digits");
  p->i8S = S;
  p->i8_blocks = blocks;
  return CMM_OK;
}

template <int S>
static int i8_inverse_s(const cmm_sht_plan* p, const double* flm, int nf, int f0, int nfp, double2* F, char* wk,
                        hipStream_t st) {
  const int L = p->L;
  unsigned char* Bd = reinterpret_cast<unsigned char*>(wk);
  int* eB = reinterpret_cast<int*>(wk + sht_i8_dbytes(p));
  int rc = launch<KI8Coeffs<S>>(k_i8_coeffs<S>, (unsigned)((p->i8_nb + 3) / 4 + nfp), 256, 0, st,
                                typename KI8Coeffs<S>::Args{L, nf, f0, 2 * nfp, p->i8_nb, p->i8_cdesc, flm, eB, Bd,
                                                            reinterpret_cast<double*>(F)},
                                "cmm_sht_inverse: int8 coefficient digits");
  if (rc) return rc;
  return launch<KLegInverseI8<S>>(k_leg_inverse_i8<S>, (unsigned)(64 * p->nrt * ((L + 63) / 64)), 128, 0, st,
                                  typename KLegInverseI8<S>::Args{L, nf, f0, nfp, p->nrt, p->i8_tdesc, p->i8_boff,
                                                                  reinterpret_cast<const ix4*>(p->i8_tab),
                                                                  reinterpret_cast<const ix4*>(Bd), eB,
                                                                  reinterpret_cast<double*>(F)},
                                  "cmm_sht_inverse: int8 legendre");
}

static int i8_inverse(const cmm_sht_plan* p, const double* flm, int nf, int f0, int nfp, double2* F, char* wk,
                      hipStream_t st) {
  switch (p->i8S) {
    case 4: return i8_inverse_s<4>(p, flm, nf, f0, nfp, F, wk, st);
    case 5: return i8_inverse_s<5>(p, flm, nf, f0, nfp, F, wk, st);
    case 6: return i8_inverse_s<6>(p, flm, nf, f0, nfp, F, wk, st);
    case 7: return i8_inverse_s<7>(p, flm, nf, f0, nfp, F, wk, st);
    default: return i8_inverse_s<8>(p, flm, nf, f0, nfp, F, wk, st);
  }
}

// ---------------------------------------------------------------------------
// forward: resampling GEMM on FP64 matrix cores,
//   G[f][m][j] = sum_t K_{m%2}[j][t] F[f][t][m],  j < L, t <= L
// (orders-major, so the analysis reads every order's nodes contiguously).
// One wave per (16 nodes j) x (16 orders m of one parity, re/im = 32 real
// columns) x field; K is stored transposed (Kt[p][t][j]) so the A operand
// is a coalesced row read.  Degrees of freedom t advance 4 per MFMA; the
// operands of 8 steps are loaded one block ahead.
// ---------------------------------------------------------------------------
constexpr int kResSteps = 8;

template <bool SPLIT>
struct KResample {
  struct Args {
    int L, nmg, tpad;
    const double* Kt;
    const double* F;
    double* G;
    int gx, gy;  // grid extents: (orders x parities, node tiles, fields) flattened
  };
  static constexpr int kThreads = SPLIT ? 256 : 64;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

// SPLIT: the K dimension (rings) of every output tile is split into
// kResWaves contiguous ranges, one per wave of the block; the partial tiles
// meet in LDS and wave 0 adds them in wave order (deterministic).  Small
// bandlimits have too few tiles to give every SIMD more than one wave
// (L = 512: 1,024 tiles), so they split; from L = 1024 on one wave runs a
// whole tile.
constexpr int kResWaves = 4;
__host__ __device__ constexpr bool res_split(int L) { return L <= 512; }

template <bool SPLIT>
__device__ __forceinline__ void KResample<SPLIT>::run(const Args& a, int bid) {
  const int L = a.L, nmg = a.nmg, tpad = a.tpad;
  const double* __restrict__ Kt = a.Kt;
  const double* __restrict__ F = a.F;
  double* __restrict__ G = a.G;
  __shared__ dx4 part[SPLIT ? kResWaves - 1 : 1][2][64];
  const int bx = bid % a.gx, by = (bid / a.gx) % a.gy, bz = bid / (a.gx * a.gy);
  constexpr int W = SPLIT ? kResWaves : 1;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int bk = lane >> 4, bj = lane & 15;
  const int fi = bz;
  const int j0 = by * 16;
  const int p = bx / nmg, qb = bx % nmg;
  const double* K = Kt + (int64_t)p * tpad * L;
  const double* Ff = F + (int64_t)fi * (L + 1) * L * 2;
  // B column of this lane in N-tile nt: order m = p + 2 (16 qb + 8 nt + bj / 2), part bj & 1
  int mcol[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) mcol[nt] = p + 2 * (16 * qb + 8 * nt + (bj >> 1));
  const int ja = j0 + bj;  // A row of this lane
  const int jac = min(ja, L - 1);
  // F rows parity-split (even orders, then odd): order m at column mi(m)
  const int half = (L + 1) / 2;
  int mcc[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    const int mm = min(mcol[nt], L - 1);
    mcc[nt] = (mm & 1) * half + (mm >> 1);
  }
  // this wave's steps of 4 rings: [sb, se)
  const int nsteps = tpad / 4;
  const int per = (nsteps + W - 1) / W;
  const int sb = w * per, se = min(nsteps, sb + per);
  // unconditional loads from clamped addresses (rows t > L of K are zero
  // padding to tpad rows, so steps past the last read zeros; rows j >= L and
  // columns m >= L are never stored); split: steps outside the wave's range
  // take a zero A operand
  auto load = [&](int s0, double (&a)[kResSteps], double (&b)[kResSteps][2]) {
#pragma unroll
    for (int u = 0; u < kResSteps; ++u) {
      const int tw = 4 * (s0 + u) + bk;
      const double av = K[(int64_t)min(tw, tpad - 1) * L + jac];
      a[u] = (!SPLIT || s0 + u < se) ? av : 0.0;
      const int t = min(tw, L);
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) b[u][nt] = Ff[((int64_t)t * L + mcc[nt]) * 2 + (bj & 1)];
    }
  };
  dx4 acc[2] = {dx4{0.0, 0.0, 0.0, 0.0}, dx4{0.0, 0.0, 0.0, 0.0}};
  auto mma = [&](const double (&a)[kResSteps], const double (&b)[kResSteps][2]) {
#pragma unroll
    for (int u = 0; u < kResSteps; ++u)
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u][nt], acc[nt], 0, 0, 0);
  };
  if (sb < se) {
    double aA[kResSteps], bA[kResSteps][2], aB[kResSteps], bB[kResSteps][2];
    load(sb, aA, bA);
    for (int s0 = sb;;) {
      load(s0 + kResSteps, aB, bB);
      mma(aA, bA);
      s0 += kResSteps;
      if (s0 >= se) break;
      load(s0 + kResSteps, aA, bA);
      mma(aB, bB);
      s0 += kResSteps;
      if (s0 >= se) break;
    }
  }
  if (SPLIT) {
    if (w > 0) {
      part[w - 1][0][lane] = acc[0];
      part[w - 1][1][lane] = acc[1];
    }
    __syncthreads();
    if (w > 0) return;
  }
#pragma unroll
  for (int ww = 0; ww < W - 1; ++ww)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      const dx4 v = part[ww][nt][lane];
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[nt][e] += v[e];
    }
  double* Gf = G + (int64_t)fi * L * L * 2;
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    if (mcol[nt] >= L) continue;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int j = j0 + 4 * e + bk;
      if (j < L) Gf[((int64_t)mcol[nt] * L + j) * 2 + (bj & 1)] = acc[nt][e];
    }
  }
}

template <bool SPLIT>
__global__ __launch_bounds__(KResample<SPLIT>::kThreads) void k_resample(typename KResample<SPLIT>::Args a) {
  KResample<SPLIT>::run(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// forward Legendre analysis: flm[lm][f] = sum_j w_j Pbar_lm(x_j) G[f][j][m].
// Work unit = (order m, segment of at most kFwdSeg degrees); the recurrence
// state at every segment start is tabulated once per plan (k_leg_restart,
// same chunked schedule), so no unit is longer than one segment and the
// outputs of different units are disjoint.  One block per (unit, field);
// each thread owns kFwdNpt northern Gauss-Legendre nodes (folded with their
// mirrors into S / D for even / odd l - m) and adds their products before
// they reach LDS.  Per chunk of C degrees: products -> LDS tile
// [degree][thread]; every wave sums its 64 columns of each row (RQ = 64 / C
// lanes per row, interleaved columns so 8 neighbouring lanes hit 8 bank
// groups); the first C threads add the per-wave sums.  The recurrence
// pairs of the next chunk are staged through LDS during the current one.
// Fixed summation order: deterministic.
// ---------------------------------------------------------------------------
constexpr int kFwdSeg = 64;
constexpr int kFwdChunk = 16;
constexpr int kFwdNpt = 2;

// state (P_{l-2}, P_{l-1}, s) before the chunk at every segment start
// l = m + kFwdSeg k, for every northern node: one thread per (m, node)
__global__ __launch_bounds__(256) void k_leg_restart(int L, int nh, int dstride, const double* __restrict__ xg,
                                                     const double* __restrict__ dP, const int* __restrict__ dS,
                                                     const double2* __restrict__ rec, const int* __restrict__ unit0,
                                                     double2* __restrict__ st_p, int* __restrict__ st_s) {
  const int m = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nh) return;
  const double x = xg[j];
  double pp = 0.0, pc = dP[(int64_t)m * dstride + j];
  int s = dS[(int64_t)m * dstride + j];
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  int u = unit0[m];
  for (int lc = m; lc < L; lc += kFwdChunk) {
    if ((lc - m) % kFwdSeg == 0) {
      st_p[(int64_t)u * nh + j] = double2{pp, pc};
      st_s[(int64_t)u * nh + j] = s;
      ++u;
    }
    leg_rescale(pp, pc, s);
    for (int i = 0; i < kFwdChunk; ++i) {
      const int l = lc + i;
      if (l > m && l < L) {
        const double2 ab = rec[lm0 + (l - m)];
        leg_step(ab.x, ab.y, x, pp, pc);
      }
    }
  }
}

// nn nodes symmetric about the equator (x_{nn-1-j} = -x_j), G [f][j][m] with
// nn rows per field.  The analysis runs it over the L Gauss-Legendre nodes;
// the transposed synthesis over the L+1 rings (unit weights, order scale
// msc, no pole terms for m > 0, as the synthesis skips them).
template <int C>
struct KLegForward {
  struct Args {
    int L, nn, nf;
    const int2* units;
    const double* xg;
    const double* wg;
    const double2* st_p;
    const int* st_s;
    const double2* rec;
    const double2* G;
    double2* flm;
    const double* msc;
    int skip_poles;
    int gx;  // work units (the grid's x extent; y = field)
  };
  static constexpr int kThreads = 512;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

template <int C>
__device__ __forceinline__ void KLegForward<C>::run(const Args& a, int bid) {
  const int L = a.L, nn = a.nn, nf = a.nf, skip_poles = a.skip_poles;
  const int2* __restrict__ units = a.units;
  const double* __restrict__ xg = a.xg;
  const double* __restrict__ wg = a.wg;
  const double2* __restrict__ st_p = a.st_p;
  const int* __restrict__ st_s = a.st_s;
  const double2* __restrict__ rec = a.rec;
  const double2* __restrict__ G = a.G;
  double2* __restrict__ flm = a.flm;
  const double* __restrict__ msc = a.msc;
  constexpr int RQ = 64 / C;
  extern __shared__ double2 tile[];  // [C][T + 4]
  __shared__ double2 part[2][8][C];
  __shared__ double2 rs[kFwdSeg];  // recurrence pairs of the segment
  const int fi = bid / a.gx;
  const int u = bid % a.gx;
  const int m = units[u].x, ls = units[u].y;
  const int le = min(L, ls + kFwdSeg);
  const int nh = (nn + 1) / 2;  // northern nodes (incl. the middle one when nn is odd)
  const int T = blockDim.x, TP = T + 4, nw = T >> 6;
  const double cm = msc ? msc[m] : 1.0;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  double x[kFwdNpt], pp[kFwdNpt], pc[kFwdNpt];
  int s[kFwdNpt];
  double2 S[kFwdNpt], D[kFwdNpt];
  bool act[kFwdNpt];
#pragma unroll
  for (int q = 0; q < kFwdNpt; ++q) {
    const int jw = tid + q * T;
    act[q] = jw < nh;
    const int j = act[q] ? jw : 0;
    const int jm = nn - 1 - j;
    x[q] = xg[j];
    const double2 ps = st_p[(int64_t)u * nh + j];
    pp[q] = ps.x;
    pc[q] = ps.y;
    s[q] = st_s[(int64_t)u * nh + j];
    const double2 ga = G[((int64_t)fi * nn + j) * L + m];
    const double wj = (act[q] && !(skip_poles && m > 0 && j == 0)) ? cm * wg[j] : 0.0;
    if (jm != j) {
      const double2 gb = G[((int64_t)fi * nn + jm) * L + m];
      S[q] = double2{wj * (ga.x + gb.x), wj * (ga.y + gb.y)};
      D[q] = double2{wj * (ga.x - gb.x), wj * (ga.y - gb.y)};
    } else {
      S[q] = double2{wj * ga.x, wj * ga.y};
      D[q] = S[q];
    }
  }
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  const int ri = lane / RQ, rq = lane % RQ;
  for (int i = tid; i < kFwdSeg; i += T) rs[i] = rec[lm0 + (min(ls + i, L - 1) - m)];
  __syncthreads();
  int buf = 0;
  for (int lc = ls; lc < le; lc += C) {
    bool live[kFwdNpt];
#pragma unroll
    for (int q = 0; q < kFwdNpt; ++q) {
      leg_rescale(pp[q], pc[q], s[q]);
      live[q] = (s[q] == 0) & act[q];
    }
#pragma unroll
    for (int i = 0; i < C; ++i) {
      const int l = lc + i;
      const double2 ab = rs[lc - ls + i];
      double2 v = {0.0, 0.0};
#pragma unroll
      for (int q = 0; q < kFwdNpt; ++q) {
        if (l > m) leg_step(ab.x, ab.y, x[q], pp[q], pc[q]);
        const double pv = live[q] ? pc[q] : 0.0;
        const double2 src = ((l - m) & 1) ? D[q] : S[q];
        v.x = fma(pv, src.x, v.x);
        v.y = fma(pv, src.y, v.y);
      }
      tile[i * TP + tid] = v;
    }
    __syncthreads();
    double2 acc = {0.0, 0.0};
    const double2* row = tile + ri * TP + 64 * w + rq;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const double2 v = row[RQ * c];
      acc.x += v.x;
      acc.y += v.y;
    }
#pragma unroll
    for (int o = 1; o < RQ; o <<= 1) {
      acc.x += __shfl_xor(acc.x, o, RQ);
      acc.y += __shfl_xor(acc.y, o, RQ);
    }
    if (rq == 0) part[buf][w][ri] = acc;
    __syncthreads();
    if (tid < C && lc + tid < le) {
      double2 sum = part[buf][0][tid];
      for (int ww = 1; ww < nw; ++ww) {
        sum.x += part[buf][ww][tid].x;
        sum.y += part[buf][ww][tid].y;
      }
      flm[(lm0 + (lc + tid - m)) * nf + fi] = sum;
    }
    buf ^= 1;
  }
}

template <int C>
__global__ __launch_bounds__(512) void k_leg_forward(typename KLegForward<C>::Args a) {
  KLegForward<C>::run(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Forward analysis with wave-level sums (G orders-major, G[f][m][j], as the
// resampling writes it): a block of W waves per (unit, field); lane t of
// wave w owns the northern nodes j = 64 (NP w + q) + t,
// q < NP.  Per chunk of 16 degrees (the restart table's chunk) each lane
// adds its nodes' products into 32 partial sums (degree, re / im), and the
// wave sums them over its 64 lanes by a butterfly that halves the values a
// lane holds at every step (permlane32 / permlane16 swaps, then row
// rotations, swizzles and quad permutations): no LDS tile, no barrier, 63
// exchanges for 32 sums.  Lane t ends with sum number t >> 1 (degree t >> 2,
// re / im by bit 1).  With W > 1 waves the sums of all of a unit's chunks
// meet in LDS once, after the unit.  Fixed summation order: deterministic.
// ---------------------------------------------------------------------------
constexpr int kFwdWDeg = 16;  // degrees per chunk (== kFwdChunk, the restart schedule)
#ifndef CMM_FWD_NP
#define CMM_FWD_NP 4
#endif
#ifndef CMM_FWD_MINW
#define CMM_FWD_MINW 1
#endif
constexpr int kFwdWNp = CMM_FWD_NP;          // nodes per lane
constexpr int kFwdWMinWaves = CMM_FWD_MINW;  // waves per block at least
static_assert(kFwdWDeg == kFwdChunk, "the restart table's chunk");

__device__ __forceinline__ double dpp_ror8(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), 0x128, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x128, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ double dpp_quad(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double swz_xor4(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_swizzle((int)(b & 0xffffffffll), 0x1F | (4 << 10));
  const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), 0x1F | (4 << 10));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// a + b after exchanging the upper 32 lanes of a with the lower 32 lanes of
// b: lanes 0-31 get a's sum over lanes t and t + 32, lanes 32-63 b's
__device__ __forceinline__ double swap32_add(double a, double b) {
  const long long A = __double_as_longlong(a), B = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)(A & 0xffffffffll), (unsigned)(B & 0xffffffffll),
                                                   false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(A >> 32), (unsigned)(B >> 32), false, false);
  const double na = __longlong_as_double(((long long)hi[0] << 32) | (long long)lo[0]);
  const double nb = __longlong_as_double(((long long)hi[1] << 32) | (long long)lo[1]);
  return na + nb;
}
// the same between the odd 16-lane rows of a and the even rows of b: even
// rows get a's sum over lanes t and t ^ 16, odd rows b's
__device__ __forceinline__ double swap16_add(double a, double b) {
  const long long A = __double_as_longlong(a), B = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)(A & 0xffffffffll), (unsigned)(B & 0xffffffffll),
                                                   false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(A >> 32), (unsigned)(B >> 32), false, false);
  const double na = __longlong_as_double(((long long)hi[0] << 32) | (long long)lo[0]);
  const double nb = __longlong_as_double(((long long)hi[1] << 32) | (long long)lo[1]);
  return na + nb;
}

// r[0..31] of every lane -> lane t holds sum over the wave of r[t >> 1]
__device__ __forceinline__ double wave_sum32(double (&r)[32], int lane) {
#pragma unroll
  for (int k = 0; k < 16; ++k) r[k] = swap32_add(r[k], r[k + 16]);
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = swap16_add(r[k], r[k + 8]);
  {
    const bool h = (lane >> 3) & 1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double keep = h ? r[k + 4] : r[k], send = h ? r[k] : r[k + 4];
      r[k] = keep + dpp_ror8(send);
    }
  }
  {
    const bool h = (lane >> 2) & 1;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const double keep = h ? r[k + 2] : r[k], send = h ? r[k] : r[k + 2];
      r[k] = keep + swz_xor4(send);
    }
  }
  {
    const bool h = (lane >> 1) & 1;
    const double keep = h ? r[1] : r[0], send = h ? r[0] : r[1];
    r[0] = keep + dpp_quad<0x4E>(send);  // lane ^ 2
  }
  return r[0] + dpp_quad<0xB1>(r[0]);  // lane ^ 1
}

template <int NP>
struct KLegForwardW {
  using Args = typename KLegForward<kFwdChunk>::Args;
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a, int bid);
};

template <int NP>
__device__ __forceinline__ void KLegForwardW<NP>::run(const Args& a, int bid) {
  const int L = a.L, nn = a.nn, nf = a.nf, skip_poles = a.skip_poles;
  const double* __restrict__ xg = a.xg;
  const double* __restrict__ wg = a.wg;
  const double2* __restrict__ st_p = a.st_p;
  const int* __restrict__ st_s = a.st_s;
  const double2* __restrict__ rec = a.rec;
  const double2* __restrict__ G = a.G;
  double* __restrict__ flm = reinterpret_cast<double*>(a.flm);
  __shared__ double part[3][kFwdSeg / kFwdWDeg][32];
  const int fi = bid / a.gx;
  const int u = bid % a.gx;
  const int2 un = a.units[u];
  const int m = un.x, ls = un.y;
  const int le = min(L, ls + kFwdSeg);
  const int nh = (nn + 1) / 2;
  const double cm = a.msc ? a.msc[m] : 1.0;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, W = blockDim.x >> 6;
  double x[NP], pp[NP], pc[NP];
  int s[NP];
  double2 S[NP], D[NP];
  bool act[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const int jw = 64 * (NP * w + q) + lane;
    act[q] = jw < nh;
    const int j = act[q] ? jw : 0;
    const int jm = nn - 1 - j;
    x[q] = xg[j];
    const double2 ps = st_p[(int64_t)u * nh + j];
    pp[q] = ps.x;
    pc[q] = ps.y;
    s[q] = st_s[(int64_t)u * nh + j];
    const double2 ga = G[((int64_t)fi * L + m) * nn + j];  // orders-major: node j of order m
    const double wj = (act[q] && !(skip_poles && m > 0 && j == 0)) ? cm * wg[j] : 0.0;
    const double2 gb = G[((int64_t)fi * L + m) * nn + jm];
    if (jm != j) {
      S[q] = double2{wj * (ga.x + gb.x), wj * (ga.y + gb.y)};
      D[q] = double2{wj * (ga.x - gb.x), wj * (ga.y - gb.y)};
    } else {
      S[q] = double2{wj * ga.x, wj * ga.y};
      D[q] = S[q];
    }
  }
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  // the recurrence pairs through the constant address space: wave-uniform
  // scalar loads (the table is read-only here)
#if defined(__HIP_DEVICE_COMPILE__)
  typedef const __attribute__((address_space(4))) double2 CD2;
  const CD2* rc = (const CD2*)(rec + lm0 - m);
#else
  const double2* rc = rec + lm0 - m;  // host pass: the body is not compiled for the host
#endif
  double tk[kFwdSeg / kFwdWDeg];  // this wave's sums of the unit's chunks
#pragma unroll
  for (int c = 0; c < kFwdSeg / kFwdWDeg; ++c) {
    const int lc = ls + kFwdWDeg * c;
    if (lc >= le) break;
    // values below the scale floor (s < 0) count as zero: their weights are
    // zeroed for the chunk instead of the values at every degree
    double2 Sl[NP], Dl[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      leg_rescale(pp[q], pc[q], s[q]);
      const bool live = (s[q] == 0) & act[q];
      Sl[q] = live ? S[q] : double2{0.0, 0.0};
      Dl[q] = live ? D[q] : double2{0.0, 0.0};
    }
    double r[32];
#pragma unroll
    for (int i = 0; i < kFwdWDeg; ++i) {
      // lc - m is a multiple of the chunk, so l - m has the parity of i;
      // only the first degree of an order (l == m) takes no step
      const int l = lc + i;
      const double2 ab = rc[min(l, L - 1)];
      if (i > 0 || lc > m) {
#pragma unroll
        for (int q = 0; q < NP; ++q) leg_step(ab.x, ab.y, x[q], pp[q], pc[q]);
      }
      double vx = 0.0, vy = 0.0;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const double2 src = (i & 1) ? Dl[q] : Sl[q];
        vx = fma(pc[q], src.x, vx);
        vy = fma(pc[q], src.y, vy);
      }
      r[2 * i] = vx;
      r[2 * i + 1] = vy;
    }
    tk[c] = wave_sum32(r, lane);
  }
  // waves > 0 hand their sums to wave 0 through LDS, once per unit
  const int nc = (le - ls + kFwdWDeg - 1) / kFwdWDeg;
  constexpr int NC = kFwdSeg / kFwdWDeg;
  if (W > 1) {
    if (w > 0 && !(lane & 1)) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (c < nc) part[w - 1][c][lane >> 1] = tk[c];
    }
    __syncthreads();
    if (w > 0) return;
    for (int ww = 0; ww < W - 1; ++ww) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (c < nc) tk[c] += part[ww][c][lane >> 1];
    }
  }
  const int d = lane >> 2, comp = (lane >> 1) & 1;
  if (!(lane & 1)) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int l = ls + kFwdWDeg * c + d;
      if (c < nc && l < le) flm[2 * ((lm0 + (l - m)) * nf + fi) + comp] = tk[c];
    }
  }
}

template <int NP>
__global__ __launch_bounds__(256) void k_leg_forward_w(typename KLegForwardW<NP>::Args a) {
  KLegForwardW<NP>::run(a, blockIdx.x);
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
__device__ inline void angmom(int l, int m, F get, Cplx out[3]);

// angmom from the three coefficients it reads: (l, m-1), (l, m+1), (l, m)
__device__ inline void angmom_from(int l, int m, Cplx am1, Cplx ap1, Cplx a0, Cplx out[3]) {
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

template <class F>
__device__ inline void angmom(int l, int m, F get, Cplx out[3]) {
  angmom_from(l, m, coef_any(l, m - 1, get), coef_any(l, m + 1, get), coef_any(l, m, get), out);
}

struct KStreamOps {
  struct Args {
    int L;
    const double2* w;
    const double* dinv;
    int apply_dinv;
    double2* out;
  };
  static constexpr int kThreads = 128;
  static __device__ __forceinline__ void run(const Args& a, int bid);
  // the 12 outputs of packed coefficient idx, into res[0..11]
  static __device__ __forceinline__ void compute(const Args& a, int64_t idx, double2* res);
};

__device__ __forceinline__ void KStreamOps::run(const Args& a, int bid) {
  const int L = a.L;
  double2* __restrict__ out = a.out;
  // the block's 12 outputs per coefficient are staged in LDS (row of 13 per
  // thread against bank conflicts) and stored as one contiguous run
  __shared__ double2 stage[kThreads * 13];
  const int64_t base = (int64_t)bid * kThreads;
  const int64_t idx = base + threadIdx.x;
  const int64_t total = (int64_t)L * (L + 1) / 2;
  if (idx < total) compute(a, idx, stage + threadIdx.x * 13);
  __syncthreads();
  const int nv = (int)(total - base < kThreads ? total - base : kThreads);
  double2* o = out + base * 12;
  for (int e = threadIdx.x; e < nv * 12; e += kThreads) o[e] = stage[(e / 12) * 13 + e % 12];
}

__device__ __forceinline__ void KStreamOps::compute(const Args& a, int64_t idx, double2* res) {
  const int L = a.L, apply_dinv = a.apply_dinv;
  const double2* __restrict__ w = a.w;
  const double* __restrict__ dinv = a.dinv;
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
  // u = (r x grad) psi at (l, m - 1), (l, m), (l, m + 1) once each, with the
  // conjugate symmetry of a real field (coef_any) for m - 1 = -1; then the
  // second application to each component reuses them
  Cplx u[3], up[3], um[3];
  angmom(l, m, psi, u);
  const Cplx z = cz();
  if (m + 1 <= l) angmom(l, m + 1, psi, up);
  else up[0] = up[1] = up[2] = z;
  if (m >= 1) {
    angmom(l, m - 1, psi, um);
  } else if (l >= 1) {  // (l, -1) from (l, 1): (-1)^1 conj
    angmom(l, 1, psi, um);
    for (int i = 0; i < 3; ++i) um[i] = Cplx{-um[i].re, um[i].im};
  } else {
    um[0] = um[1] = um[2] = z;
  }
  for (int i = 0; i < 3; ++i) res[i] = double2{u[i].re, u[i].im};
  for (int i = 0; i < 3; ++i) {
    Cplx g[3];
    angmom_from(l, m, um[i], up[i], u[i], g);
    for (int k = 0; k < 3; ++k) res[3 + 3 * i + k] = double2{g[k].re, g[k].im};
  }
}

__global__ __launch_bounds__(128) void k_stream_ops(KStreamOps::Args a) { KStreamOps::run(a, blockIdx.x); }

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
// adjoints (transposes of the linear maps above, for reverse-mode gradients)
// ---------------------------------------------------------------------------
// Transposed Gauss-Legendre quadrature of the analysis:
// G[f][j][m] = w_j sum_l Pbar_lm(x_j) flm[lm][f] at the L nodes, one thread per
// northern node (folded with its mirror), one block per (m, field).
__global__ __launch_bounds__(256) void k_leg_synth_gl(int L, int nf, const double* __restrict__ xg,
                                                      const double* __restrict__ wg, const double* __restrict__ dP,
                                                      const int* __restrict__ dS, const double2* __restrict__ rec,
                                                      const double2* __restrict__ flm, double2* __restrict__ G) {
  // recurrence pairs and this order's coefficients of one field staged in
  // LDS (L <= 512: 2 x 512 x 16 bytes)
  __shared__ double2 srec[512], scf[512];
  const int m = blockIdx.x, fi = blockIdx.y;
  const int nh = (L + 1) / 2;
  const int j = threadIdx.x;
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  for (int i = threadIdx.x; i < L - m; i += blockDim.x) {
    srec[i] = rec[lm0 + i];
    scf[i] = flm[(lm0 + i) * nf + fi];
  }
  __syncthreads();
  if (j >= nh) return;
  const double x = xg[j];
  double pp = 0.0, pc = dP[(int64_t)m * L + j];
  int s = dS[(int64_t)m * L + j];
  double2 S = {0.0, 0.0}, D = {0.0, 0.0};
  leg_rescale(pp, pc, s);
  {
    const double v = s == 0 ? pc : 0.0;
    S.x = v * scf[0].x;
    S.y = v * scf[0].y;
  }
#pragma unroll 4
  for (int i = 1; i < L - m; ++i) {
    const double2 ab = srec[i];
    leg_step(ab.x, ab.y, x, pp, pc);
    leg_rescale(pp, pc, s);
    const double v = s == 0 ? pc : 0.0;
    const double2 a = scf[i];
    if (i & 1) {
      D.x = fma(v, a.x, D.x);
      D.y = fma(v, a.y, D.y);
    } else {
      S.x = fma(v, a.x, S.x);
      S.y = fma(v, a.y, S.y);
    }
  }
  const int jm = L - 1 - j;
  double2* Gf = G + (int64_t)fi * L * L;
  const double wj = wg[j];
  if (jm != j) {
    Gf[(int64_t)j * L + m] = double2{wj * (S.x + D.x), wj * (S.y + D.y)};
    const double wm = wg[jm];
    Gf[(int64_t)jm * L + m] = double2{wm * (S.x - D.x), wm * (S.y - D.y)};
  } else {
    Gf[(int64_t)j * L + m] = double2{wj * S.x, wj * S.y};
  }
}

// Transposed resampling with the input scale of the transposed r2c FFT:
// F[f][t][m] = c_m sum_j K_{m mod 2}[j][t] G[f][j][m], t = 0..L, c_0 = pi / L,
// c_m = pi / (2 L) (k_fft_c2r doubles orders m > 0).  The layout of
// k_resample with the roles of t and j exchanged: one wave per (16 rings) x
// (16 orders of one parity, re/im) x field, K2[p][j][t] = K[p][t][j] zero
// padded (rows j >= L, columns t > L), FP64 matrix cores.
__global__ __launch_bounds__(64) void k_resample_adj(int L, int nmg, int jpad, int tcols, const double* __restrict__ K2,
                                                     const double* __restrict__ G, double* __restrict__ F) {
  const int lane = threadIdx.x;
  const int bk = lane >> 4, bj = lane & 15;
  const int fi = blockIdx.z;
  const int t0 = blockIdx.y * 16;
  const int p = (int)blockIdx.x / nmg, qb = (int)blockIdx.x % nmg;
  const double* K = K2 + (int64_t)p * jpad * tcols;
  const double* Gf = G + (int64_t)fi * L * L * 2;
  int mcol[2], mcc[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    mcol[nt] = p + 2 * (16 * qb + 8 * nt + (bj >> 1));
    mcc[nt] = min(mcol[nt], L - 1);
  }
  const int ta = t0 + bj;  // A row (output ring) of this lane, < tcols
  auto load = [&](int s0, double (&a)[kResSteps], double (&b)[kResSteps][2]) {
#pragma unroll
    for (int u = 0; u < kResSteps; ++u) {
      const int jw = 4 * (s0 + u) + bk;
      a[u] = K[(int64_t)min(jw, jpad - 1) * tcols + ta];
      const int j = min(jw, L - 1);
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) b[u][nt] = Gf[((int64_t)j * L + mcc[nt]) * 2 + (bj & 1)];
    }
  };
  dx4 acc[2] = {dx4{0.0, 0.0, 0.0, 0.0}, dx4{0.0, 0.0, 0.0, 0.0}};
  const int nsteps = jpad / 4;
  auto mma = [&](const double (&a)[kResSteps], const double (&b)[kResSteps][2], int s0) {
#pragma unroll
    for (int u = 0; u < kResSteps; ++u)
      if (s0 + u < nsteps)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u][nt], acc[nt], 0, 0, 0);
  };
  double aA[kResSteps], bA[kResSteps][2], aB[kResSteps], bB[kResSteps][2];
  load(0, aA, bA);
  for (int s0 = 0;;) {
    load(s0 + kResSteps, aB, bB);
    mma(aA, bA, s0);
    s0 += kResSteps;
    if (s0 >= nsteps) break;
    load(s0 + kResSteps, aA, bA);
    mma(aB, bB, s0);
    s0 += kResSteps;
    if (s0 >= nsteps) break;
  }
  double* Ff = F + (int64_t)fi * (L + 1) * L * 2;
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    if (mcol[nt] >= L) continue;
    const double c = (mcol[nt] == 0 ? 1.0 : 0.5) * M_PI / L;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int t = t0 + 4 * e + bk;
      if (t <= L) Ff[((int64_t)t * L + mcol[nt]) * 2 + (bj & 1)] = c * acc[nt][e];
    }
  }
}

__device__ inline void decode_lm(int L, int64_t idx, int& l, int& m) {
  int lo = 0, hi = L - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) / 2;
    if (lm_index(L, mid, mid) <= idx) lo = mid; else hi = mid - 1;
  }
  m = lo;
  l = m + (int)(idx - lm_index(L, m, m));
}

// Transpose of angmom: the adjoint of the stored coefficient a(l, m), m >= 0,
// from the adjoints o(l', m', k) of the three outputs at every stored (l', m')
// that reads it: (l, m+1) through a(l, m'-1), (l, m-1) through a(l, m'+1),
// (l, m) through a(l, m'), and for m = 1 also (l, 0) through
// a(l, -1) = -conj(a(l, 1)).  Complex-scalar factors transpose to their
// conjugates.
template <class O>
__device__ inline Cplx angmom_T(int l, int m, O o) {
  Cplx r = cz();
  auto add_mi = [&](double c, Cplx v) {  // r += (-i c) v
    r.re += c * v.im;
    r.im -= c * v.re;
  };
  if (m + 1 <= l) {
    const double h = Lp(l, m + 1) / 2;
    add_mi(h, o(l, m + 1, 0));
    const Cplx v = o(l, m + 1, 1);
    r.re += h * v.re;
    r.im += h * v.im;
  }
  if (m >= 1) {
    const double h = Lm(l, m - 1) / 2;
    add_mi(h, o(l, m - 1, 0));
    const Cplx v = o(l, m - 1, 1);
    r.re -= h * v.re;
    r.im -= h * v.im;
  }
  add_mi((double)m, o(l, m, 2));
  if (m == 1) {
    const double h = Lp(l, 0) / 2;
    const Cplx v0 = o(l, 0, 0), v1 = o(l, 0, 1);
    const Cplx t = Cplx{h * v0.im + h * v1.re, -h * v0.re + h * v1.im};
    r.re -= t.re;  // transpose of a -> -conj(a)
    r.im += t.im;
  }
  return r;
}

// stream operators, reverse step 1: ubar_i = obar_i + angmom^T of the
// gradient outputs of u_i
__global__ __launch_bounds__(128) void k_stream_ops_adj1(int L, const double2* __restrict__ ob,
                                                         double2* __restrict__ ub) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)L * (L + 1) / 2) return;
  int l, m;
  decode_lm(L, idx, l, m);
  for (int i = 0; i < 3; ++i) {
    const Cplx g = angmom_T(l, m, [&](int ll, int mm, int k) {
      const double2 v = ob[lm_index(L, ll, mm) * 12 + 3 + 3 * i + k];
      return Cplx{v.x, v.y};
    });
    const double2 d = ob[idx * 12 + i];
    ub[idx * 3 + i] = double2{d.x + g.re, d.y + g.im};
  }
}

// reverse step 2: wbar = D_inv * angmom^T(ubar)
__global__ __launch_bounds__(128) void k_stream_ops_adj2(int L, const double2* __restrict__ ub,
                                                         const double* __restrict__ dinv, int apply_dinv,
                                                         double2* __restrict__ wb) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)L * (L + 1) / 2) return;
  int l, m;
  decode_lm(L, idx, l, m);
  const Cplx g = angmom_T(l, m, [&](int ll, int mm, int k) {
    const double2 v = ub[lm_index(L, ll, mm) * 3 + k];
    return Cplx{v.x, v.y};
  });
  const double dd = apply_dinv ? dinv[l] : 1.0;
  wb[idx] = double2{dd * g.re, dd * g.im};
}

__global__ __launch_bounds__(128) void k_scalar_ops_adj(int L, const double2* __restrict__ ob,
                                                        double2* __restrict__ wb) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)L * (L + 1) / 2) return;
  int l, m;
  decode_lm(L, idx, l, m);
  const Cplx g = angmom_T(l, m, [&](int ll, int mm, int k) {
    const double2 v = ob[lm_index(L, ll, mm) * 4 + 1 + k];
    return Cplx{v.x, v.y};
  });
  const double2 d = ob[idx * 4];
  wb[idx] = double2{d.x + g.re, d.y + g.im};
}

static_assert(kFwdSeg == kFwdSegHost, "segment length");

static int build_restart(cmm_sht_plan* p, const std::vector<int2>& units, const std::vector<int>& unit0) {
  const int L = p->L, nh = (L + 1) / 2;
  int rc;
  if ((rc = upload(&p->units, units.data(), units.size())) || (rc = upload(&p->unit0, unit0.data(), unit0.size())))
    return rc;
  p->nunits = (int)units.size();
  hipError_t e = hipMalloc(&p->st_p, sizeof(double2) * (size_t)p->nunits * nh);
  if (e != hipSuccess) return hip_check(e, "cmm_sht_plan_create: hipMalloc");
  e = hipMalloc(&p->st_s, sizeof(int) * (size_t)p->nunits * nh);
  if (e != hipSuccess) return hip_check(e, "cmm_sht_plan_create: hipMalloc");
  hipLaunchKernelGGL(k_leg_restart, dim3((nh + 255) / 256, L), dim3(256), 0, 0, L, nh, L, p->x_g, p->dP_g, p->dS_g,
                     p->rec, p->unit0, p->st_p, p->st_s);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_check(e, "cmm_sht_plan_create: restart kernel");
  const int nhr = L / 2 + 1;
  e = hipMalloc(&p->st_pr, sizeof(double2) * (size_t)p->nunits * nhr);
  if (e != hipSuccess) return hip_check(e, "cmm_sht_plan_create: hipMalloc");
  e = hipMalloc(&p->st_sr, sizeof(int) * (size_t)p->nunits * nhr);
  if (e != hipSuccess) return hip_check(e, "cmm_sht_plan_create: hipMalloc");
  hipLaunchKernelGGL(k_leg_restart, dim3((nhr + 255) / 256, L), dim3(256), 0, 0, L, nhr, L + 1, p->x_t, p->dP_t,
                     p->dS_t, p->rec, p->unit0, p->st_pr, p->st_sr);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_check(e, "cmm_sht_plan_create: restart kernel");
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_check(e, "cmm_sht_plan_create: restart kernel");
  return CMM_OK;
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

template <int S>
static int oz_inverse_s(const cmm_sht_plan* p, const double* flm, int nf, int f0, int nfp, double2* F, char* oz,
                        hipStream_t st) {
  const int L = p->L;
  const int64_t dstride = sht_oz_dstride(p, nfp);
  uint16_t* Bd = reinterpret_cast<uint16_t*>(oz);
  int* eB = reinterpret_cast<int*>(oz + (kOzMaxS * sht_oz_dstride(p, std::min(nf, 16)) * 2 + 255) / 256 * 256);
  hipLaunchKernelGGL(k_oz_exponents, dim3(L), dim3(1024), 0, st, L, nf, f0, 2 * nfp, flm, eB);
  CMM_LAUNCH_CHECK("cmm_sht_inverse: coefficient exponents");
  hipLaunchKernelGGL(k_oz_coeffs<S>, dim3(L, (L + 31) / 32), dim3(256), 0, st, L, nf, f0, 2 * nfp, p->offJ, dstride,
                     flm, eB, Bd);
  CMM_LAUNCH_CHECK("cmm_sht_inverse: coefficient digits");
  const int nh = L / 2 + 1;
  const int nrb = (nh + kOzRings - 1) / kOzRings;
  const dim3 grid((unsigned)(8 * ((L + 7) / 8) * nrb)), block(128);
  const size_t lds_rec = (size_t)32 * ((L + 31) / 32) * sizeof(double2);
  hipLaunchKernelGGL(k_leg_inverse_oz<S>, grid, block, lds_rec, st, L, nf, f0, nfp, nrb, p->eA, p->x_t, p->dP_t, p->dS_t, p->rec,
                     p->offJ, dstride, Bd, eB, reinterpret_cast<double*>(F));
  CMM_LAUNCH_CHECK("cmm_sht_inverse: split-bf16 legendre");
  return CMM_OK;
}

static int oz_inverse(const cmm_sht_plan* p, const double* flm, int nf, int f0, int nfp, double2* F, char* oz,
                      hipStream_t st) {
  switch (p->legendre) {
    case CMM_LEGENDRE_BF16X3: return oz_inverse_s<3>(p, flm, nf, f0, nfp, F, oz, st);
    case CMM_LEGENDRE_BF16X4: return oz_inverse_s<4>(p, flm, nf, f0, nfp, F, oz, st);
    case CMM_LEGENDRE_BF16X5: return oz_inverse_s<5>(p, flm, nf, f0, nfp, F, oz, st);
    case CMM_LEGENDRE_BF16X6: return oz_inverse_s<6>(p, flm, nf, f0, nfp, F, oz, st);
    default: return oz_inverse_s<7>(p, flm, nf, f0, nfp, F, oz, st);
  }
}

extern "C" int cmm_sht_inverse(const cmm_sht_plan* p, const double* flm, int nf, double* f, void* work,
                               void* stream) {
  if (!p || !flm || !f || !work || nf < 1) return fail(CMM_E_ARG, "cmm_sht_inverse: bad arguments");
  const int L = p->L;
  double2* F = reinterpret_cast<double2*>(work);
  hipStream_t st = as_stream(stream);
  // staged recurrence pairs: at most ceil(L / 8) + 2 chunks of 8 degrees
  const size_t lds_rec = (size_t)((L + kInvDeg - 1) / kInvDeg + 2) * kInvDeg * sizeof(double2);
  if (p->legendre >= CMM_LEGENDRE_I8X4) {
    char* wk = reinterpret_cast<char*>(work) + sht_f_bytes(p, nf);
    for (int f0 = 0; f0 < nf;) {
      const int nfp = std::min(16, nf - f0);
      const int rc = i8_inverse(p, flm, nf, f0, nfp, F, wk, st);
      if (rc) return rc;
      f0 += nfp;
    }
  } else if (p->legendre >= CMM_LEGENDRE_BF16X3) {
    if (batch_recording()) return fail(CMM_E_STATE, "cmm_sht_inverse: the split-BF16 forms have no batched form");
    char* oz = reinterpret_cast<char*>(work) + sht_f_bytes(p, nf);
    for (int f0 = 0; f0 < nf;) {
      const int nfp = std::min(16, nf - f0);
      const int rc = oz_inverse(p, flm, nf, f0, nfp, F, oz, st);
      if (rc) return rc;
      f0 += nfp;
    }
  } else {
    for (int f0 = 0; f0 < nf;) {
      const int nfp = std::min(16, nf - f0);
      const unsigned nb = (unsigned)p->inv_waves;
      const char* what = "cmm_sht_inverse: legendre";
      double* Fd = reinterpret_cast<double*>(F);
#define CMM_LEG_INV(NT, MX)                                                                                   \
  launch<KLegInverse<NT, MX>>(k_leg_inverse<NT, MX>, nb, 64, lds_rec, st,                                    \
                              typename KLegInverse<NT, MX>::Args{L, nf, f0, nfp, p->inv_tasks, p->inv_off, p->x_t, \
                                                                 p->dP_t, p->dS_t, p->rec, flm, Fd},           \
                              what)
      const int rc = p->legendre == CMM_LEGENDRE_FP32 ? (nfp > 8 ? CMM_LEG_INV(2, true) : CMM_LEG_INV(1, true))
                                                      : (nfp > 8 ? CMM_LEG_INV(2, false) : CMM_LEG_INV(1, false));
#undef CMM_LEG_INV
      if (rc) return rc;
      f0 += nfp;
    }
  }
  int pairs = (L + 2) / 2;
  size_t lds = fft_lds_bytes(fft_plan(p));
  if (p->legendre >= CMM_LEGENDRE_I8X4) {  // rings-fastest coefficients
    const int gx = 32 * ((pairs + 31) / 32);
    return launch<KFftC2r<true>>(k_fft_c2r<true>, (unsigned)(gx * nf), 256, lds, st,
                                 KFftC2r<true>::Args{fft_plan(p), p->tw, F, L, f, gx}, "cmm_sht_inverse: fft");
  }
  return launch<KFftC2r<false>>(k_fft_c2r<false>, (unsigned)(nf * pairs), 256, lds, st,
                                KFftC2r<false>::Args{fft_plan(p), p->tw, F, L, f, 0}, "cmm_sht_inverse: fft");
}

// analysis workspace: F (nf, L+1, L) then G (nf, L, L), complex
static int64_t sht_forward_bytes(const cmm_sht_plan* p, int nf) {
  return sht_f_bytes(p, nf) + (int64_t)nf * p->L * p->L * 16;
}

static int sht_forward(const cmm_sht_plan* p, const double* f, int nf, double* flm, void* work, void* stream);

extern "C" int cmm_sht_forward(const cmm_sht_plan* p, const double* f, int nf, double* flm, void* work,
                               void* stream) {
  if (!p || !flm || !f || !work || nf < 1) return fail(CMM_E_ARG, "cmm_sht_forward: bad arguments");
  return sht_forward(p, f, nf, flm, work, stream);
}

extern "C" int cmm_sht_forward_noalloc(const cmm_sht_plan* p, const double* f, int nf, double* flm, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_sht_forward_noalloc: no batched form (members would share its scratch)");
  if (!p || !flm || !f || nf < 1) return fail(CMM_E_ARG, "cmm_sht_forward_noalloc: bad arguments");
  if (!p->scratch) return fail(CMM_E_ARG, "cmm_sht_forward_noalloc: plan has no scratch address");
  const int64_t need = sht_forward_bytes(p, nf);
  if (need > kNoAllocScratchBytes)
    return fail(CMM_E_RANGE, "cmm_sht_forward_noalloc: needs " + std::to_string(need) + " bytes of scratch, " +
                                 "the limit is " + std::to_string(kNoAllocScratchBytes) +
                                 " (16 nf L (2L + 1) bytes: L = 1024 with nf = 1, L = 512 with nf <= 7)");
  return sht_forward(p, f, nf, flm, p->scratch, stream);
}

// FFT -> resampling -> Gauss-Legendre analysis; `work` holds F and G (every
// element is written before it is read, so no clearing)
static int sht_forward(const cmm_sht_plan* p, const double* f, int nf, double* flm, void* work, void* stream) {
  const int L = p->L;
  double2* F = reinterpret_cast<double2*>(work);
  double2* G = F + (int64_t)nf * (L + 1) * L;
  hipStream_t st = as_stream(stream);
  int pairs = (L + 2) / 2;
  size_t lds = fft_lds_bytes(fft_plan(p));
  int rc = launch<KFftR2c>(k_fft_r2c, (unsigned)(nf * pairs), 256, lds, st, KFftR2c::Args{fft_plan(p), p->tw, f, L, F, 1},
                           "cmm_sht_forward: fft");
  if (rc) return rc;
  const int nmg = ((L + 1) / 2 + 15) / 16;  // groups of 16 orders of one parity (at most ceil(L / 2) orders)
  const int gy = (L + 15) / 16;
  const KResample<true>::Args ra{L, nmg, p->tpad, p->K, reinterpret_cast<const double*>(F),
                                 reinterpret_cast<double*>(G), 2 * nmg, gy};
  const unsigned rb = (unsigned)(2 * nmg * gy * nf);
  rc = res_split(L) ? launch<KResample<true>>(k_resample<true>, rb, 64 * kResWaves, 0, st, ra,
                                              "cmm_sht_forward: resample")
                    : launch<KResample<false>>(k_resample<false>, rb, 64, 0, st,
                                               KResample<false>::Args{ra.L, ra.nmg, ra.tpad, ra.Kt, ra.F, ra.G,
                                                                      ra.gx, ra.gy},
                                               "cmm_sht_forward: resample");
  if (rc) return rc;
  const int nh = (L + 1) / 2;
  // waves of kFwdWNp x 64 nodes; bandlimits up to 512 share one block size
  // (a batched launch has one), so their sums run in the same order in a
  // batch as alone
  const int W = std::max(kFwdWMinWaves, (nh + 64 * kFwdWNp - 1) / (64 * kFwdWNp));
  if (W > 4) return fail(CMM_E_RANGE, "cmm_sht_forward: L > 2048");
  return launch<KLegForwardW<kFwdWNp>>(k_leg_forward_w<kFwdWNp>, (unsigned)(p->nunits * nf), 64 * W, 0, st,
                                       KLegForwardW<kFwdWNp>::Args{L, L, nf, p->units, p->x_g, p->w_g, p->st_p,
                                                                   p->st_s, p->rec, G,
                                                                   reinterpret_cast<double2*>(flm), nullptr, 0,
                                                                   p->nunits},
                                       "cmm_sht_forward: legendre");
}

extern "C" int cmm_sht_stream_operators(const cmm_sht_plan* p, const double* w, int apply_dinv,
                                        double* out, void* stream) {
  if (!p || !w || !out) return fail(CMM_E_ARG, "cmm_sht_stream_operators: bad arguments");
  int64_t total = (int64_t)p->L * (p->L + 1) / 2;
  return launch<KStreamOps>(k_stream_ops, (unsigned)grid_for(total, 128), 128, 0, as_stream(stream),
                            KStreamOps::Args{p->L, reinterpret_cast<const double2*>(w), p->dinv, apply_dinv,
                                             reinterpret_cast<double2*>(out)},
                            "cmm_sht_stream_operators");
}

extern "C" int cmm_sht_scalar_operators(const cmm_sht_plan* p, const double* w, double* out, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_sht_scalar_operators: no batched form");
  if (!p || !w || !out) return fail(CMM_E_ARG, "cmm_sht_scalar_operators: bad arguments");
  int64_t total = (int64_t)p->L * (p->L + 1) / 2;
  hipLaunchKernelGGL(k_scalar_ops, dim3(grid_for(total, 128)), dim3(128), 0, as_stream(stream), p->L,
                     reinterpret_cast<const double2*>(w), reinterpret_cast<double2*>(out));
  CMM_LAUNCH_CHECK("cmm_sht_scalar_operators");
  return CMM_OK;
}

// ---- adjoints -------------------------------------------------------------
// Transposed synthesis (adjoint of the inverse Legendre stage) on FP64 matrix
// cores: flm_bar[lm][f] = msc[m] sum_t Pbar_lm(x_t) G[f][t][m] over the L + 1
// MWSS rings (the pole rings skipped for m > 0), folded by parity:
// S_t = G_t + G_{L-t} for even l - m, D_t = G_t - G_{L-t} for odd.  One block
// per (order m, chunk c of 32 degrees; blockIdx.x = offJ[m] + c) and group of
// <= 16 fields; wave w takes northern rings 64 w .. 64 w + 63 (lane = ring),
// restarts the recurrence from the analysis restart state of its 64-degree
// unit, writes the 16 even-degree values of the chunk into an LDS tile
// [degree][ring] and multiplies (v_mfma_f64_16x16x4f64: A = 16 degrees x 4
// rings from the tile, B = 4 rings x 16 columns of S from G), then the odd
// degrees with D.  The waves' partial sums meet in LDS.
template <int NT>
__global__ __launch_bounds__(320) void k_leg_adj_mfma(int L, int nf, int f0, int nfp, const int* __restrict__ offJ,
                                                      const int* __restrict__ unit0, const double* __restrict__ xr,
                                                      const double2* __restrict__ st_p, const int* __restrict__ st_s,
                                                      const double2* __restrict__ rec, const double* __restrict__ G,
                                                      const double* __restrict__ msc, double* __restrict__ flm) {
  __shared__ double tile[5][16][65];  // per wave: 16 degrees x 64 rings (+1 padding)
  __shared__ double2 rs[kFwdSeg];     // recurrence pairs of degrees ls .. lc + 31
  const int b = (int)blockIdx.x;
  int lo = 0, hi = L;  // largest m with offJ[m] <= b
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offJ[mid] <= b) lo = mid; else hi = mid;
  }
  const int m = lo, c = b - offJ[m];
  const int nn = L + 1, nh = L / 2 + 1;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lc = m + 32 * c;
  const int su = (32 * c) / kFwdSeg;
  const int u = unit0[m] + su;
  const int ls = m + kFwdSeg * su;
  const int64_t lm0 = (int64_t)m * L - (int64_t)m * (m - 1) / 2;
  // the ring pair values of this wave's B operands, loaded under the recurrence:
  // ring 64 w + 4 k + lane / 16, column 16 nt + lane % 16
  double ga[16][NT], gb[16][NT];
  bool gok[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int jr = 64 * w + 4 * k + (lane >> 4);
    gok[k] = jr < nh && !(m > 0 && jr == 0);
    const int jc = gok[k] ? jr : 0;
    const int jm = nn - 1 - jc;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int col = 16 * nt + (lane & 15);
      const int f = min(col >> 1, nfp - 1);
      ga[k][nt] = G[(((int64_t)(f0 + f) * nn + jc) * L + m) * 2 + (col & 1)];
      gb[k][nt] = jm != jc ? G[(((int64_t)(f0 + f) * nn + jm) * L + m) * 2 + (col & 1)] : 0.0;
    }
  }
  for (int i = threadIdx.x; i < lc + 32 - ls; i += blockDim.x) rs[i] = rec[lm0 + (min(ls + i, L - 1) - m)];
  const int j = 64 * w + lane;
  const bool act = j < nh;
  const int jj = act ? j : 0;
  const double x = xr[jj];
  const double2 ps = st_p[(int64_t)u * nh + jj];
  double pp = ps.x, pc = ps.y;
  int s = st_s[(int64_t)u * nh + jj];
  __syncthreads();
  // advance from the unit start to the chunk (rescale every 16 degrees, as the analysis)
  for (int l = ls; l < lc; l += 16) {
    leg_rescale(pp, pc, s);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int ll = l + i;
      const double2 ab = rs[ll - ls];
      if (ll > m) leg_step(ab.x, ab.y, x, pp, pc);
    }
  }
  double odd[16];
  bool live = false;
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    if ((i & 15) == 0) {
      leg_rescale(pp, pc, s);
      live = (s == 0) & act;
    }
    const int ll = lc + i;
    const double2 ab = rs[ll - ls];
    if (ll > m) leg_step(ab.x, ab.y, x, pp, pc);
    const double pv = (live & (ll < L)) ? pc : 0.0;
    if (i & 1) odd[i >> 1] = pv;
    else tile[w][i >> 1][lane] = pv;
  }
  const double cm = msc[m];
  dx4 acc[2][NT];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[p][nt] = dx4{0.0, 0.0, 0.0, 0.0};
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const double a = tile[w][lane & 15][4 * k + (lane >> 4)];
    const double wgt = gok[k] ? cm : 0.0;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      acc[0][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, wgt * (ga[k][nt] + gb[k][nt]), acc[0][nt], 0, 0, 0);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 16; ++i) tile[w][i][lane] = odd[i];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const double a = tile[w][lane & 15][4 * k + (lane >> 4)];
    const int jr = 64 * w + 4 * k + (lane >> 4);
    const double wgt = gok[k] ? cm : 0.0;
    const bool mid = (nn - 1 - jr) == jr;  // the equator ring: its own mirror
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      acc[1][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, wgt * (mid ? ga[k][nt] : ga[k][nt] - gb[k][nt]),
                                                        acc[1][nt], 0, 0, 0);
  }
  __syncthreads();
  double* red = &tile[w][0][0];  // [p][nt][e][lane]
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int e = 0; e < 4; ++e) red[((p * NT + nt) * 4 + e) * 64 + lane] = acc[p][nt][e];
  __syncthreads();
  const int nw = blockDim.x >> 6;
  for (int o = threadIdx.x; o < 2 * NT * 4 * 64; o += blockDim.x) {
    double v = 0.0;
    for (int ww = 0; ww < nw; ++ww) v += (&tile[ww][0][0])[o];
    const int ln = o & 63, e = (o >> 6) & 3, pn = o >> 8, nt = pn % NT, p = pn / NT;
    const int row = 4 * e + (ln >> 4), col = 16 * nt + (ln & 15);
    const int l = lc + 2 * row + p, f = col >> 1;
    if (l < L && f < nfp) flm[((lm0 + (l - m)) * nf + f0 + f) * 2 + (col & 1)] = v;
  }
}

extern "C" int cmm_sht_inverse_adjoint(const cmm_sht_plan* p, const double* fbar, int nf, double* flm_bar,
                                       void* work, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_sht_inverse_adjoint: no batched form");
  if (!p || !fbar || !flm_bar || !work || nf < 1) return fail(CMM_E_ARG, "cmm_sht_inverse_adjoint: bad arguments");
  const int L = p->L;
  double2* F = reinterpret_cast<double2*>(work);
  hipStream_t st = as_stream(stream);
  const int pairs = (L + 2) / 2;
  const size_t lds = fft_lds_bytes(fft_plan(p));
  hipLaunchKernelGGL(k_fft_r2c, dim3(nf * pairs), dim3(256), lds, st, KFftR2c::Args{fft_plan(p), p->tw, fbar, L, F, 0});
  CMM_LAUNCH_CHECK("cmm_sht_inverse_adjoint: fft");
  const int nhr = L / 2 + 1;
  if (nhr <= 5 * 64) {  // matrix-core form: every field of a group shares the recurrence
    const int nw = (nhr + 63) / 64;
    for (int f0 = 0; f0 < nf; f0 += 16) {
      const int nfp = std::min(16, nf - f0);
      auto kern = nfp > 8 ? k_leg_adj_mfma<2> : k_leg_adj_mfma<1>;
      hipLaunchKernelGGL(kern, dim3((unsigned)p->nJtot), dim3(64 * nw), 0, st, L, nf, f0, nfp, p->offJ, p->unit0,
                         p->x_t, p->st_pr, p->st_sr, p->rec, reinterpret_cast<const double*>(F), p->msc_r, flm_bar);
      CMM_LAUNCH_CHECK("cmm_sht_inverse_adjoint: legendre");
    }
    return CMM_OK;
  }
  const int T = std::max(64, ((nhr + kFwdNpt - 1) / kFwdNpt + 63) / 64 * 64);
  if (T > 512) return fail(CMM_E_RANGE, "cmm_sht_inverse_adjoint: L > 2046");
  const size_t lds2 = (size_t)kFwdChunk * (T + 4) * sizeof(double2);
  hipLaunchKernelGGL(k_leg_forward<kFwdChunk>, dim3(p->nunits * nf), dim3(T), lds2, st,
                     KLegForward<kFwdChunk>::Args{L, L + 1, nf, p->units, p->x_t, p->w_r, p->st_pr, p->st_sr, p->rec, F,
                                                  reinterpret_cast<double2*>(flm_bar), p->msc_r, 1, p->nunits});
  CMM_LAUNCH_CHECK("cmm_sht_inverse_adjoint: legendre");
  return CMM_OK;
}

extern "C" int cmm_sht_forward_adjoint(const cmm_sht_plan* p, const double* flm_bar, int nf, double* fbar,
                                       void* work, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_sht_forward_adjoint: no batched form");
  if (!p || !fbar || !flm_bar || !work || nf < 1) return fail(CMM_E_ARG, "cmm_sht_forward_adjoint: bad arguments");
  const int L = p->L;
  double2* F = reinterpret_cast<double2*>(work);
  double2* G = F + (int64_t)nf * (L + 1) * L;
  hipStream_t st = as_stream(stream);
  const int nh = (L + 1) / 2;
  const int T = (nh + 63) / 64 * 64;
  if (T > 256) return fail(CMM_E_RANGE, "cmm_sht_forward_adjoint: L > 512");
  hipLaunchKernelGGL(k_leg_synth_gl, dim3(L, nf), dim3(T), 0, st, L, nf, p->x_g, p->w_g, p->dP_g, p->dS_g, p->rec,
                     reinterpret_cast<const double2*>(flm_bar), G);
  CMM_LAUNCH_CHECK("cmm_sht_forward_adjoint: legendre");
  const int nmg = ((L + 1) / 2 + 15) / 16;
  hipLaunchKernelGGL(k_resample_adj, dim3(2 * nmg, p->tcols / 16, nf), dim3(64), 0, st, L, nmg, p->jpad, p->tcols,
                     p->K2, reinterpret_cast<const double*>(G), reinterpret_cast<double*>(F));
  CMM_LAUNCH_CHECK("cmm_sht_forward_adjoint: resample");
  const int pairs = (L + 2) / 2;
  const size_t lds = fft_lds_bytes(fft_plan(p));
  hipLaunchKernelGGL(k_fft_c2r<false>, dim3(nf * pairs), dim3(256), lds, st,
                     KFftC2r<false>::Args{fft_plan(p), p->tw, F, L, fbar, 0});
  CMM_LAUNCH_CHECK("cmm_sht_forward_adjoint: fft");
  return CMM_OK;
}

extern "C" int cmm_sht_stream_operators_adjoint(const cmm_sht_plan* p, const double* ops_bar, int apply_dinv,
                                                double* w_bar, double* scratch, void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_sht_stream_operators_adjoint: no batched form");
  if (!p || !ops_bar || !w_bar || !scratch) return fail(CMM_E_ARG, "cmm_sht_stream_operators_adjoint: bad arguments");
  const int64_t total = (int64_t)p->L * (p->L + 1) / 2;
  hipStream_t st = as_stream(stream);
  double2* ub = reinterpret_cast<double2*>(scratch);
  hipLaunchKernelGGL(k_stream_ops_adj1, dim3(grid_for(total, 128)), dim3(128), 0, st, p->L,
                     reinterpret_cast<const double2*>(ops_bar), ub);
  CMM_LAUNCH_CHECK("cmm_sht_stream_operators_adjoint");
  hipLaunchKernelGGL(k_stream_ops_adj2, dim3(grid_for(total, 128)), dim3(128), 0, st, p->L, ub, p->dinv, apply_dinv,
                     reinterpret_cast<double2*>(w_bar));
  CMM_LAUNCH_CHECK("cmm_sht_stream_operators_adjoint");
  return CMM_OK;
}

extern "C" int cmm_sht_scalar_operators_adjoint(const cmm_sht_plan* p, const double* ops_bar, double* w_bar,
                                                void* stream) {
  if (batch_recording()) return fail(CMM_E_STATE, "cmm_sht_scalar_operators_adjoint: no batched form");
  if (!p || !ops_bar || !w_bar) return fail(CMM_E_ARG, "cmm_sht_scalar_operators_adjoint: bad arguments");
  const int64_t total = (int64_t)p->L * (p->L + 1) / 2;
  hipLaunchKernelGGL(k_scalar_ops_adj, dim3(grid_for(total, 128)), dim3(128), 0, as_stream(stream), p->L,
                     reinterpret_cast<const double2*>(ops_bar), reinterpret_cast<double2*>(w_bar));
  CMM_LAUNCH_CHECK("cmm_sht_scalar_operators_adjoint");
  return CMM_OK;
}
