/*
 * cmm_s2.h -- C ABI of the MI355X-native Characteristic Mapping Method engine.
 *
 * Plain pointers and sizes only.  Every `const double* q`, `double* out` ...
 * below is a DEVICE pointer (HBM) unless the comment says "host".  Every
 * function is asynchronous on the caller's HIP stream (`void* stream`, a
 * hipStream_t; NULL = the default stream) and returns 0 on success or a
 * negative CMM_E* code; cmm_last_error() then holds a message (thread local).
 * Mesh / plan handles own their device tables on the device that was current
 * when they were created.
 *
 * Array layouts follow the reference (seth-tl/CMM-S2, cmm_s2/core):
 *   jets      vals[c][k][v]  (ncomp, 4, Nv): k = 0 value, k = 1+d derivative
 *             along Cartesian axis d, stored transposed for vector fields
 *             (vals[d][1+c][v] = d f_c / d x_d), spherical_spline.py:300-337
 *   coeffs    (ncomp, 19, Ntri) Bernstein-Bezier coefficients, same file
 *   packed    (Ntri, 6, ncomp, 6) evaluation layout: for each split
 *             sub-triangle (argmin, argmax of the barycentrics) the 6
 *             coefficients of the Cs table (spherical_spline.py:660-662)
 *   points    strided: element (i, c) at q[i*stride_pt + c*stride_comp], so
 *             both (N,3) and (3,N) arrays are accepted without copies
 *   spectra   [lm][f] complex128 (interleaved re, im), lm = packed (l, m>=0)
 *             index m*L - m*(m-1)/2 + (l - m), f = field index fastest
 *   grids     (nf, L+1, 2L) real, MWSS rings theta_t = 2 t pi / (2L)
 */
#ifndef CMM_S2_H
#define CMM_S2_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CMM_OK 0
#define CMM_E_ARG (-1)      /* bad argument / shape */
#define CMM_E_HIP (-2)      /* HIP runtime error (launch, alloc, copy) */
#define CMM_E_RANGE (-3)    /* size outside the supported range */
#define CMM_E_STATE (-4)    /* call out of order (batch recording) */

#define CMM_ABI_VERSION 1

int cmm_abi_version(void);
const char* cmm_last_error(void);

/* ------------------------------------------------------------------------ */
/* Unstructured spherical triangulations (icosahedral meshes).              */
/* Replaces spherical_triangulation.__init__ / query                        */
/* (cmm_s2/core/mesh_functions.py:29-88): location is a cube-map bucket     */
/* search over exact barycentric tests instead of Qhull's find_simplex.     */
/* ------------------------------------------------------------------------ */
typedef struct cmm_tri_mesh cmm_tri_mesh;

/* vertices: host (nv,3); simplices: host (ntri,3) int64, counter-clockwise
 * (det(v0,v1,v2) > 0).  cells_per_face <= 0 picks a size from ntri. */
int cmm_tri_mesh_create(const double* vertices, int64_t nv, const int64_t* simplices,
                        int64_t ntri, int cells_per_face, cmm_tri_mesh** out);
int cmm_tri_mesh_destroy(cmm_tri_mesh* mesh);
/* host out[6] = {nv, ntri, cells_per_face, total candidates, max candidates
 * per cell, bytes of device tables} */
int cmm_tri_mesh_stats(const cmm_tri_mesh* mesh, int64_t* out);

/* mesh_functions.py:68-88 -- containing triangle (row of simplices) and
 * homogeneous barycentrics bcc (n,3) (bcc may be NULL).  Never misses: a
 * query outside every candidate by round-off gets the most-inside one. */
int cmm_tri_locate(const cmm_tri_mesh* mesh, const double* q, int64_t n,
                   int64_t q_stride_pt, int64_t q_stride_comp, int32_t* tri,
                   double* bcc, void* stream);

/* sphere_diffeomorphism / spherical_spline .assemble_coefficients
 * (spherical_spline.py:115-152, 300-337, 471-508): jets (ncomp,4,Nv) ->
 * coeffs (ncomp,19,Ntri) and/or packed (Ntri,6,ncomp,6); either output may
 * be NULL.  ncomp in {1, 3}. */
int cmm_tri_assemble(const cmm_tri_mesh* mesh, const double* vals, int ncomp,
                     double* coeffs, double* packed, void* stream);

/* sphere_diffeomorphism.__call__ (spherical_spline.py:178-205) when
 * normalize != 0, the raw PS polynomial otherwise: out(i, c) for c < ncomp. */
int cmm_tri_eval(const cmm_tri_mesh* mesh, const double* packed, int ncomp, int normalize,
                 const double* q, int64_t n, int64_t q_stride_pt, int64_t q_stride_comp,
                 double* out, int64_t out_stride_pt, int64_t out_stride_comp, void* stream);

/* composite_sphere_diffeomorphism.__call__ / eval_all
 * (spherical_spline.py:365-437) fused over the chain: levels is a HOST array
 * of `depth` device pointers to packed map coefficients, applied in array
 * order (levels[0] first), each followed by normalisation.  inter (optional,
 * may be NULL): (depth, n, 3) positions after every level. */
int cmm_tri_compose(const cmm_tri_mesh* mesh, const double* const* levels, int depth,
                    const double* q, int64_t n, int64_t q_stride_pt, int64_t q_stride_comp,
                    double* out, int64_t out_stride_pt, int64_t out_stride_comp,
                    double* inter, void* stream);

/* cmm_tri_compose for points forming a ring-major nring x ncol grid (point
 * (r, c) at index r ncol + c, e.g. the MWSS sample grid): identical results;
 * the threads walk the grid in tiles of tile_rows x (64 / tile_rows) points
 * (tile_rows a power of two <= 64, 0 = 4) so a wave covers a compact patch.
 * No counterpart in the reference (a traversal order, not an operation). */
int cmm_tri_compose_grid(const cmm_tri_mesh* mesh, const double* const* levels, int depth,
                         const double* q, int64_t nring, int64_t ncol, int tile_rows, int64_t q_stride_pt,
                         int64_t q_stride_comp, double* out, int64_t out_stride_pt,
                         int64_t out_stride_comp, double* inter, void* stream);

/* Location of points for spline evaluation on the mesh: block[i] = 6 t + s
 * for the containing triangle t (the search of cmm_tri_locate) and its
 * Powell-Sabin split sub-triangle s (spherical_spline.py:660-666 picks the
 * same one), sub_bary (n,3) the barycentrics in that sub-triangle; -1 / NaN
 * where the point is not located.  No counterpart in the reference (it
 * locates the evaluation grid anew in every step, spherical_spline.py:
 * 365-398); these are the values the evaluation computes, stored. */
int cmm_tri_split_locate(const cmm_tri_mesh* mesh, const double* q, int64_t n, int64_t q_stride_pt,
                         int64_t q_stride_comp, int32_t* block, double* sub_bary, void* stream);

/* cmm_tri_compose_grid whose first level (levels[0]) is evaluated at the
 * stored location of every point (first_block, first_sub_bary from
 * cmm_tri_split_locate on the same points and mesh) instead of a search:
 * identical results.  For a grid composed at every step, as the CMM loop's
 * sample grid (spherical_simulations.py:263-294).  No counterpart in the
 * reference. */
int cmm_tri_compose_grid_located(const cmm_tri_mesh* mesh, const double* const* levels, int depth,
                                 const double* q, int64_t nring, int64_t ncol, int tile_rows,
                                 int64_t q_stride_pt, int64_t q_stride_comp, double* out,
                                 int64_t out_stride_pt, int64_t out_stride_comp, double* inter,
                                 const int32_t* first_block, const double* first_sub_bary, void* stream);

/* Differential of an evaluated spline.  ncomp 3: Jacobian of the normalised
 * map x -> p(x)/|p(x)| restricted to the tangent plane at x, out (3,3,n)
 * row-major [c][d] at out[(3c + d) n + i] (sphere_diffeomorphism.eval_grad,
 * spherical_spline.py:208-265 with grad_HBB 700-711; the reference function
 * reads attributes it never sets, so its math is restated, plus the input-side
 * projection).  ncomp 1: tangential gradient of a scalar spline, out (3,n)
 * (spherical_spline.gradient 73-111 with d_Proj 714-720). */
int cmm_tri_eval_grad(const cmm_tri_mesh* mesh, const double* packed, int ncomp, const double* q,
                      int64_t n, int64_t q_stride_pt, int64_t q_stride_comp, double* out, void* stream);

/* composite map and its Jacobian (chain rule through the levels, applied in
 * array order): out(i, c) positions, jac (3,3,n) as cmm_tri_eval_grad.  No
 * counterpart in the reference; used for the area-preservation diagnostic
 * det(Dphi) = 1 of incompressible flows. */
int cmm_tri_compose_grad(const cmm_tri_mesh* mesh, const double* const* levels, int depth,
                         const double* q, int64_t n, int64_t q_stride_pt, int64_t q_stride_comp,
                         double* out, int64_t out_stride_pt, int64_t out_stride_comp, double* jac,
                         void* stream);

/* advect (stencil part) + spline_proj_sphere (evolution_functions.py:90-133,
 * spherical_spline.py:268-298) fused.  yv: (3, Nv) integrated vertex feet,
 * element (c, v) at yv[c*yv_row + v]; st: (3, 4*Nv) stencil feet, element
 * (c, k*Nv + v) at st[c*st_row + k*Nv + v].  cur_packed is the current map:
 * the foot yv_v is located in it and the four stencil feet are evaluated on
 * that split sub-triangle's polynomial (un-normalised extrapolation).  With
 * cur_packed == NULL (identity map) the stencil feet are the stencil values
 * and yv is not read.  tan_vects: (2, Nv, 3) = [a1_n, a2_n].  Writes jets
 * (3,4,Nv) and, if st_evals != NULL, the stencil values (3,4,Nv) before
 * projection. */
int cmm_tri_stencil_jets(const cmm_tri_mesh* mesh, const double* cur_packed, const double* yv,
                         int64_t yv_row, const double* st, int64_t st_row, const double* tan_vects,
                         double eps, double* jets, double* st_evals, void* stream);

/* ------------------------------------------------------------------------ */
/* MWSS structured velocity mesh (mesh_functions.py:169-318, 352-412).       */
/* ------------------------------------------------------------------------ */
typedef struct cmm_mwss_mesh cmm_mwss_mesh;

/* Host tables: thetas (L+1), phis (2L) and their sines / cosines exactly as
 * the caller's numpy computes them (the location tests then agree bit for
 * bit with numpy away from ties). */
int cmm_mwss_mesh_create(int L, const double* thetas, const double* phis,
                         const double* sin_t, const double* cos_t,
                         const double* sin_p, const double* cos_p, cmm_mwss_mesh** out);
int cmm_mwss_mesh_destroy(cmm_mwss_mesh* mesh);

/* structure_spherical_triangulation.query (mesh_functions.py:196-282):
 * numpy floor-division / remainder semantics and the 0.03 / 0.85 band
 * corrections.  tri (n,) int32, bcc (n,3) or NULL. */
int cmm_mwss_locate(const cmm_mwss_mesh* mesh, const double* q, int64_t n,
                    int64_t q_stride_pt, int64_t q_stride_comp, int32_t* tri,
                    double* bcc, void* stream);

/* The location step of spline_interp_velocity.__call__ (spherical_spline.py
 * :570-615 with bary_minmax / new_vpoints, 660-697): tri (n) int32 (-1 on a
 * miss), and where not NULL sub (n) int32 the split sub-triangle, cb (n,3)
 * its barycentrics, w (n,9) the triangle's dual basis w_0, w_1, w_2
 * (homogeneous barycentrics b_i = q . w_i).  known != NULL (device int32
 * (n)): point i is taken to lie in triangle known[i] and its vertices are
 * built in closed form without the angle / cell search (what the RK-feet
 * adjoint does for the stage points it recomputes); when known[i] is the
 * triangle the full search finds, the triangle and sub-triangle are the
 * search's and cb, w agree with it to rounding. */
int cmm_mwss_split_locate(const cmm_mwss_mesh* mesh, const double* q, int64_t n, int64_t q_stride_pt,
                          int64_t q_stride_comp, const int32_t* known, int32_t* tri, int32_t* sub,
                          double* cb, double* w, void* stream);

/* assemble_coefficients on the MWSS mesh; jets (ncomp,4,Nu). */
int cmm_mwss_assemble(const cmm_mwss_mesh* mesh, const double* vals, int ncomp,
                      double* coeffs, double* packed, void* stream);

/* cmm_mwss_assemble of a velocity level (ncomp 3, packed only) on a subset
 * of the Tu triangles.  Bit `slot` (0..7) of vbits[t] says where the level's
 * packed blocks are current (the rest is never written):
 *   mark != NULL: the triangles t with mark[t] != 0, those of the 4 rows of
 *     triangles next to either pole and, if inherit >= 0, those with bit
 *     `inherit` set (the set of an earlier level) are assembled, their bit
 *     set and the bit of every other triangle cleared, and mark is cleared
 *     (mark is what the evaluations of cmm_velocity.mark flagged since the
 *     last call: the triangles the next steps' feet will fall in);
 *   mark == NULL: the triangles whose bit is clear are assembled and their
 *     bit set (the level becomes complete).
 * Replaces the full-mesh assembly of spline_interp_velocity.init_coeffs
 * (spherical_spline.py:565-568) in the loop; evaluations outside the
 * assembled set compute their block from the jets (cmm_velocity.vbits), with
 * the same arithmetic. */
int cmm_mwss_assemble_sparse(const cmm_mwss_mesh* mesh, const double* vals, double* packed,
                             uint8_t* vbits, int slot, int inherit, uint8_t* mark, void* stream);

/* spherical_spline.gradient (spherical_spline.py:73-111) for a scalar spline
 * (packed, ncomp 1) on the MWSS mesh: tangential gradient, out (3, n). */
int cmm_mwss_eval_grad(const cmm_mwss_mesh* mesh, const double* packed, const double* q, int64_t n,
                       int64_t q_stride_pt, int64_t q_stride_comp, double* out, void* stream);

/* spline_interp_velocity.__call__ (spherical_spline.py:570-615) and the
 * scalar spherical_spline (42-70): out(i,c) = sum_k w[k] * P_k(q_i)_c over
 * nlev packed levels (host arrays `levels`, `w`). */
int cmm_mwss_eval(const cmm_mwss_mesh* mesh, const double* const* levels, const double* w,
                  int nlev, int ncomp, const double* q, int64_t n, int64_t q_stride_pt,
                  int64_t q_stride_comp, double* out, int64_t out_stride_pt,
                  int64_t out_stride_comp, void* stream);

/* ------------------------------------------------------------------------ */
/* Backward characteristic integrators (evolution_functions.py:19-80).       */
/* ------------------------------------------------------------------------ */
#define CMM_VEL_SPLINE 0     /* spline_interp_velocity on an MWSS mesh */
#define CMM_VEL_SBR 1        /* dynamical_fields.u_sbr: params[0]=T, [1]=alpha */
#define CMM_VEL_DEFORM_ROT 2 /* dynamical_fields.u_deform_rot: params[1]=alpha */
#define CMM_VEL_RH4 3        /* Rossby-Haurwitz-4: params[0]=w, [1]=K, [2]=Omega */
#define CMM_VEL_DEFORM 4     /* dynamical_fields.u_deform: params[0]=T */
#define CMM_VEL_STATIC_VORTEX 5   /* dynamical_fields.u_static_vortex */
#define CMM_VEL_ROTATING_VORTEX 6 /* dynamical_fields.u_rotating_vortex: params[0]=T */
#define CMM_VEL_EQUATOR_VORTEX 7  /* dynamical_fields.equator_vortex */
#define CMM_VEL_DIV 8        /* dynamical_fields.u_div: params[0]=T */

#define CMM_INT_EULER 0      /* Euler_step_proj */
#define CMM_INT_IMPROVED_EULER 1 /* improved_Euler_proj */
#define CMM_INT_RK4 2        /* RK4_proj */
#define CMM_INT_RK2 3        /* RK2_proj */
#define CMM_INT_EVAL 4       /* no step: y = V(stage_t[0], x) (the velocity itself) */

typedef struct cmm_velocity {
  int kind;
  int nlev;                       /* spline: time levels read (1..3) */
  const cmm_mwss_mesh* mesh;      /* spline */
  const double* levels[3];        /* spline: packed coefficients per level */
  double stage_w[4][3];           /* spline: time weights for each stage */
  double stage_t[4];              /* analytic: time of each stage */
  double params[8];
  /* spline, optional (vbits NULL: every level complete).  Levels assembled
   * on a subset of triangles (cmm_mwss_assemble_sparse): where bit vslot[k]
   * of vbits[t] is clear an evaluation builds level k's block of triangle t
   * from jets[k] (3, 4, Nu) (the wave computes the 57 coefficients together;
   * same arithmetic as the assembly).  mark != NULL: the triangles an
   * evaluation locates are set to 1 there, every one (mark_every != 0) or
   * only those whose block of some level was missing.  Not accepted by the
   * adjoint entry points. */
  const uint8_t* vbits;
  int vslot[3];
  const double* jets[3];
  uint8_t* mark;
  int mark_every;
} cmm_velocity;

/* x, y: (3, m) component-major; y may alias x. */
int cmm_advect_points(const cmm_velocity* vel, int integrator, double dt, const double* x,
                      int64_t m, double* y, void* stream);

/* The integration part of advect (evolution_functions.py:117-133) for one
 * map in one launch: vertex feet verts (3, nv) (row stride verts_row) and
 * stencil feet st (3, 4 nv) (element (c, k nv + v) at st[c st_row + k nv + v])
 * -> yv, yst in the same layouts.  The foot of a vertex and its four stencil
 * points run in neighbouring lanes (they are 1e-5 apart, so their velocity
 * gathers share cache lines).  verts == NULL (identity map): stencil feet
 * only.  Outputs must not alias the inputs. */
int cmm_advect_stencil(const cmm_velocity* vel, int integrator, double dt, const double* verts,
                       int64_t verts_row, const double* st, int64_t st_row, int64_t nv, double* yv,
                       int64_t yv_row, double* yst, int64_t yst_row, void* stream);

/* ------------------------------------------------------------------------ */
/* MWSS spherical-harmonic transforms (replace pyssht forward / inverse,     */
/* spherical_harmonic_tools.py:16,84,91,123,133) and spectral operators.     */
/* ------------------------------------------------------------------------ */
typedef struct cmm_sht_plan cmm_sht_plan;

/* cos_t, sin_t: host (L+1) cos / sin of the MWSS colatitudes. */
int cmm_sht_plan_create(int L, const double* cos_t, const double* sin_t, cmm_sht_plan** out);
int cmm_sht_plan_destroy(cmm_sht_plan* plan);
/* Arithmetic of the Legendre synthesis of cmm_sht_inverse (the analysis and
 * every other kernel stay FP64).  No counterpart in the reference, which is
 * FP64 only.  FP32 runs the contraction on FP32 matrix cores with FP32
 * operands, the recurrence in FP64 and FP32 partial sums of 32 degrees per
 * parity added into FP64 accumulators; relative error ~1e-7 of the field.
 * BF16XS (S = 3..7) splits both operands into S bf16 fixed-point digits
 * (a power of two per ring bound / per order and column) and sums the
 * S (S + 1) / 2 leading digit products on bf16 matrix cores in FP32
 * accumulators that are exact for the four leading weights, flushed into
 * FP64 every 192 degrees per parity; the truncation error is ~2^(4 - 8 S)
 * of the largest term (S = 3: ~1e-6, S = 4: ~1e-8, S = 5: ~4e-11,
 * S = 6: ~1e-13, S = 7: ~1e-14). */
#define CMM_LEGENDRE_FP64 0
#define CMM_LEGENDRE_FP32 1
#define CMM_LEGENDRE_BF16X3 2
#define CMM_LEGENDRE_BF16X4 3
#define CMM_LEGENDRE_BF16X5 4
#define CMM_LEGENDRE_BF16X6 5
#define CMM_LEGENDRE_BF16X7 6
/* I8XS (S = 4..8): the Ozaki scheme on int8 matrix cores.  Pbar at the
 * northern MWSS rings is held as S balanced base-256 int8 digits per value
 * (one exponent per order, parity of l - m and tile of 32 rings), built once
 * by this call (device memory ~S x 30 MB at L = 512, ~S x 240 MB at
 * L = 1024; chunks of 32 degrees whose digits are all zero are not stored);
 * the coefficients are split the same way per call (one exponent per order
 * and column) and the S (S + 1) / 2 leading digit products are summed
 * exactly in int32 accumulators by weight, converted to FP64 once per
 * parity.  Error ~2^(-8 S) of the operands' scales (S = 5: ~1e-12,
 * S = 7: ~1e-16, S = 8: FP64 rounding level).  The pole rings are summed
 * in FP64. */
#define CMM_LEGENDRE_I8X4 7
#define CMM_LEGENDRE_I8X5 8
#define CMM_LEGENDRE_I8X6 9
#define CMM_LEGENDRE_I8X7 10
#define CMM_LEGENDRE_I8X8 11
int cmm_sht_set_legendre(cmm_sht_plan* plan, int mode);
/* host copies of the plan's Gauss-Legendre nodes / weights (L each) */
int cmm_sht_plan_gl(const cmm_sht_plan* plan, double* x, double* w);
/* bytes of device workspace needed by forward / inverse for nf fields */
int cmm_sht_workspace_bytes(const cmm_sht_plan* plan, int nf, int64_t* bytes);

/* synthesis: flm [lm][nf] complex -> f (nf, L+1, 2L) */
int cmm_sht_inverse(const cmm_sht_plan* plan, const double* flm, int nf, double* f,
                    void* work, void* stream);
/* analysis: f (nf, L+1, 2L) -> flm [lm][nf] complex */
int cmm_sht_forward(const cmm_sht_plan* plan, const double* f, int nf, double* flm,
                    void* work, void* stream);

/* cmm_sht_forward with its workspace in a fixed device array of the library
 * (64 MiB per GPU, held until the process ends) instead of `work`: calls no
 * allocator.  No counterpart in the reference; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Needs 16 nf L (2L + 1) bytes <= 64 MiB (L = 1024 with
 * nf = 1, L = 512 with nf <= 7), else CMM_E_RANGE.  Every call on one GPU shares the
 * array: calls that overlap on one GPU (other streams, other threads) race. */
int cmm_sht_forward_noalloc(const cmm_sht_plan* plan, const double* f, int nf, double* flm, void* stream);

/* inv_Laplacian scaling + angular_momentum twice (spherical_harmonic_tools.py
 * :14-19, 49-62, 79-91): from omega_lm [lm] (one field) writes out [lm][12]:
 * fields 0..2 = (r x grad) psi, field 3+3i+k = (r x grad)_k of field i, with
 * psi = D_inv omega when apply_dinv != 0 (psi = omega_lm otherwise). */
int cmm_sht_stream_operators(const cmm_sht_plan* plan, const double* omega_lm,
                             int apply_dinv, double* out, void* stream);
/* angular_momentum once: in [lm] -> out [lm][4] = (f, (r x grad) f) */
int cmm_sht_scalar_operators(const cmm_sht_plan* plan, const double* f_lm, double* out,
                             void* stream);

/* project_stream grid post-processing (spherical_harmonic_tools.py:85-101):
 * fields (12, L+1, 2L) -> velocity jets vals (3,4,Nu) (J^T layout). */
int cmm_mwss_velocity_jets(const cmm_mwss_mesh* mesh, const double* fields, double* vals,
                           void* stream);
/* project_onto_S12 post-processing (spherical_harmonic_tools.py:121-138):
 * fields (4, L+1, 2L) = (f, r x grad f) -> scalar jets vals (4, Nu). */
int cmm_mwss_scalar_jets(const cmm_mwss_mesh* mesh, const double* fields, double* vals,
                         void* stream);

/* Vorticity sampling (spherical_simulations.py:268-273):
 * out_t,p = w0(X) + 2 rot (cos theta(X) - cos theta_t) at the MWSS grid of
 * bandlimit L (theta_t from cos_t, device (L+1)); X (n,3) with n = (L+1)*2L.
 * w0 is a scalar PS spline on `wmesh` (packed) when kind = 0, a
 * spherical-harmonic series (table in wpacked) for CMM_VORT_SH_SERIES, else
 * an analytic field CMM_VORT_*.  rot_rate 0 samples w0 alone (tracers). */
#define CMM_VORT_SPLINE 0
#define CMM_VORT_ROSSBY 1        /* dynamical_fields.rossby_wave, params[0]=t */
#define CMM_VORT_ZONAL_JET 2
#define CMM_VORT_GAUSSIAN 3
#define CMM_VORT_PERTURBED_JET 4
#define CMM_VORT_MULTI_JET 5
#define CMM_VORT_PERTURBED_ROSSBY 6
/* tracer initial conditions (dynamical_fields.py:87-156) */
#define CMM_VORT_COSINE_BELL 7
#define CMM_VORT_SLOTTED_CYLINDER 8        /* IC_SC */
#define CMM_VORT_STATIC_VORTEX_TRACER 9    /* static_vort_IC(t, .), params[0]=t */
#define CMM_VORT_ROTATING_VORTEX_TRACER 10 /* rotating_vort_IC(t, .), params[0]=t */
/* a band-limited real field from its coefficients: wpacked is then a device
 * table [m][l] complex (Ls, Ls) of a_lm (m >= 0, conjugate-symmetric
 * field), params[0] = Ls; e.g. dynamical_fields.random_initial_condition
 * (dynamical_fields.py:159-171) */
#define CMM_VORT_SH_SERIES 11
int cmm_sample_vorticity(int kind, const cmm_mwss_mesh* wmesh, const double* wpacked,
                         const double* params, const double* X, int64_t n, int64_t x_stride_pt,
                         int64_t x_stride_comp, const double* cos_grid_theta, int ring_len,
                         double rot_rate, double* out, void* stream);

/* ------------------------------------------------------------------------ */
/* Adjoints (reverse-mode transposes of the entry points above) for the      */
/* gradient of a final-vorticity misfit with respect to the initial spectral */
/* coefficients.  The reference has no gradient code; each entry point is    */
/* the exact transpose of the forward entry point named in its comment, for  */
/* the same location decisions.  "accumulates" = adds into the output with   */
/* FP64 atomics (clear it first; summation order is not fixed); the other    */
/* outputs are overwritten.                                                  */
/* ------------------------------------------------------------------------ */
/* Coefficient adjoints ("packed_bar", "level_bars", "cur_bar", ...) have the
 * layout of the packed array they belong to, (Ntri, 6, ncomp, 6).
 * The coefficient adjoints of the four scattering entry points below
 * (composition, RK feet, stencil projection, vorticity sampling) either
 * accumulate with FP64 atomics (work == NULL) or go through a grouped
 * scatter-add (counting sort by target block, one writer per block) in a
 * caller workspace of cmm_scatter_workspace_bytes(records, key_range) bytes
 * (records / key_range: depth * n / depth * 6 Ntri for the composition,
 * 16 Nv / 6 Tu of the velocity mesh for the RK feet, Nv / 6 Ntri for the
 * stencil projection, n / 6 Tu of the vorticity mesh for the sampling):
 * several times faster than the FP64 atomics. */
int cmm_scatter_workspace_bytes(int64_t records, int64_t key_range, int64_t* bytes);
/* transpose of cmm_sht_inverse: fbar (nf, L+1, 2L) -> flm_bar [lm][nf].
 * work: cmm_sht_workspace_bytes(nf). */
int cmm_sht_inverse_adjoint(const cmm_sht_plan* plan, const double* fbar, int nf, double* flm_bar,
                            void* work, void* stream);
/* transpose of cmm_sht_forward: flm_bar [lm][nf] -> fbar (nf, L+1, 2L); L <= 512. */
int cmm_sht_forward_adjoint(const cmm_sht_plan* plan, const double* flm_bar, int nf, double* fbar,
                            void* work, void* stream);
/* transpose of cmm_sht_stream_operators: ops_bar [lm][12] -> omega_bar [lm];
 * scratch: 3 * nlm complex. */
int cmm_sht_stream_operators_adjoint(const cmm_sht_plan* plan, const double* ops_bar, int apply_dinv,
                                     double* omega_bar, double* scratch, void* stream);
/* transpose of cmm_sht_scalar_operators: ops_bar [lm][4] -> f_bar [lm]. */
int cmm_sht_scalar_operators_adjoint(const cmm_sht_plan* plan, const double* ops_bar, double* f_bar,
                                     void* stream);
/* transposes of cmm_mwss_velocity_jets (12 fields) / cmm_mwss_scalar_jets
 * (4 fields): vals_bar (.., Nu) -> fields_bar (nf, L+1, 2L). */
int cmm_mwss_velocity_jets_adjoint(const cmm_mwss_mesh* mesh, const double* vals_bar, double* fields_bar,
                                   void* stream);
int cmm_mwss_scalar_jets_adjoint(const cmm_mwss_mesh* mesh, const double* vals_bar, double* fields_bar,
                                 void* stream);
/* transpose of cmm_mwss_assemble (packed only): packed_bar -> vals_bar
 * (accumulates).  With a workspace of 96 * ncomp * Tu bytes it runs as a
 * per-triangle pass and a per-vertex gather (no atomics, deterministic);
 * work == NULL uses atomics. */
int cmm_mwss_assemble_adjoint(const cmm_mwss_mesh* mesh, const double* packed_bar, int ncomp,
                              double* vals_bar, void* work, int64_t work_bytes, void* stream);
/* transpose of cmm_tri_assemble (ncomp 3, packed only): accumulates vals_bar;
 * workspace 288 * Ntri bytes for the gather form, or NULL for atomics. */
int cmm_tri_assemble_adjoint(const cmm_tri_mesh* mesh, const double* packed_bar, double* vals_bar,
                             void* work, int64_t work_bytes, void* stream);
/* adjoint of cmm_advect_stencil for the stencil feet (spline velocity;
 * Euler, improved Euler or RK4): yst_bar (3, 4 nv) -> the packed adjoints of
 * the nlev velocity levels (accumulate; level_bars[k] pairs with
 * vel->levels[k]) and, if st_bar != NULL, the adjoint of the start points
 * (3, 4 nv), overwritten. */
int cmm_advect_stencil_adjoint(const cmm_velocity* vel, int integrator, double dt, const double* st,
                               int64_t st_row, int64_t nv, const double* yst_bar, int64_t yst_bar_row,
                               double* const* level_bars, double* st_bar, int64_t st_bar_row, void* work,
                               int64_t work_bytes, void* stream);
/* cmm_advect_stencil_adjoint with the velocity-level adjoints kept sparse,
 * as corner jets: level_corners[k] (Tu, 3, 12) holds, for every triangle t
 * with level_touched[k][t] != 0, the adjoints of the 12 jets rows
 * (cmm_mwss_velocity_jets order) at its 3 corners; triangles with a zero flag
 * hold no adjoint and are not read.  Adds into the levels (a level's first
 * touch of a triangle stores and sets its flag; zero the flags of a new
 * level).  Equals cmm_advect_stencil_adjoint followed by the per-triangle
 * pass of cmm_mwss_assemble_adjoint, without the (Tu, 108) packed adjoints.
 * work (required): cmm_scatter_workspace_bytes(16 nv, Tu) bytes. */
int cmm_advect_stencil_adjoint_corners(const cmm_velocity* vel, int integrator, double dt, const double* st,
                                       int64_t st_row, int64_t nv, const double* yst_bar, int64_t yst_bar_row,
                                       double* const* level_corners, uint8_t* const* level_touched,
                                       double* st_bar, int64_t st_bar_row, void* work, int64_t work_bytes,
                                       void* stream);
/* second pass of the transposed velocity assembly from the sparse corner jets
 * of cmm_advect_stencil_adjoint_corners: vals_bar (12, Nu) accumulates. */
int cmm_mwss_corner_jets_adjoint(const cmm_mwss_mesh* mesh, const double* corners, const uint8_t* touched,
                                 double* vals_bar, void* stream);
/* adjoint of cmm_tri_stencil_jets: jets_bar (3,4,Nv) -> stencil-feet
 * adjoints sf_bar (3, 4 Nv) and, when cur != NULL, the current map's packed
 * adjoints cur_bar (accumulates). */
int cmm_tri_stencil_jets_adjoint(const cmm_tri_mesh* mesh, const double* cur, const double* yv,
                                 int64_t yrow, const double* sf, int64_t srow, const double* tanv,
                                 double eps, const double* jets_bar, double* sf_bar, int64_t sf_bar_row,
                                 double* cur_bar, void* work, int64_t work_bytes, void* stream);
/* adjoint of cmm_tri_compose (depth <= 8): out_bar (n,3) -> packed adjoints
 * of the levels (accumulate; NULL entries skipped) and q_bar (n,3) or NULL. */
int cmm_tri_compose_adjoint(const cmm_tri_mesh* mesh, const double* const* levels,
                            double* const* level_bars, int depth, const double* q, int64_t n,
                            int64_t q_stride_pt, int64_t q_stride_comp, const double* out_bar,
                            double* q_bar, void* work, int64_t work_bytes, void* stream);
/* adjoint of cmm_sample_vorticity for a spline vorticity: wbar (n) ->
 * x_bar (strided, may be NULL) and wpacked_bar (accumulates, may be NULL). */
int cmm_sample_vorticity_adjoint(const cmm_mwss_mesh* wmesh, const double* wpacked, const double* X,
                                 int64_t n, int64_t x_stride_pt, int64_t x_stride_comp,
                                 const double* wbar, double rot_rate, double* x_bar,
                                 int64_t xb_stride_pt, int64_t xb_stride_comp, double* wpacked_bar,
                                 void* work, int64_t work_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Member-batched launches (ensembles).  No counterpart in the reference,    */
/* which runs one simulation per process; this is the one-launch-per-kernel  */
/* step of an ensemble of simulations with different bandlimits and meshes. */
/* Between cmm_batch_begin and cmm_batch_launch the calling host thread      */
/* records the launches of the step entry points (advect, stencil jets, map  */
/* assembly, composition, vorticity sampling, the transforms, spectral       */
/* operators, velocity jets and assembly) instead of issuing them, member by */
/* member (cmm_batch_member(b, 0), the member's calls, cmm_batch_member(b,   */
/* 1), ...); cmm_batch_launch then issues one grid per recorded position     */
/* for all members on `stream` (the j-th launches of every member must be    */
/* the same kernel with the same block size) and reports the grids issued.   */
/* Each block runs its member's code and arguments, so every member gets the */
/* bits of its own run.  The batch owns `bytes` of device memory for the     */
/* per-member arguments of one step and two pinned staging buffers.  An      */
/* entry point without a batched form returns CMM_E_STATE while recording.   */
/* ------------------------------------------------------------------------ */
typedef struct cmm_batch cmm_batch;
int cmm_batch_create(int64_t bytes, cmm_batch** out);
int cmm_batch_destroy(cmm_batch* batch);
int cmm_batch_begin(cmm_batch* batch);
int cmm_batch_member(cmm_batch* batch, int member);
int cmm_batch_cancel(cmm_batch* batch);
int cmm_batch_launch(cmm_batch* batch, void* stream, int64_t* launches);
/* dst[i] = src[i], i < n, on `stream` (batchable): the remap's copy of the
 * current map's jets into the submap stack (spherical_spline.py:471-473
 * assigns them). */
int cmm_copy_f64(double* dst, const double* src, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif
#endif
