/*
 * bem2d_hip.h -- C-ABI of the MI355X (gfx950) implementation of the bem2d hot path:
 * dense Galerkin assembly of the Helmholtz single-layer, double-layer, adjoint double-layer and
 * combined-field operators with piecewise-constant bases, the plane-wave right-hand sides, the
 * dense complex solve behind SoundSoftScattering, and the potential / far-field evaluation.
 *
 * Plain C types only.  Every entry point returns a status (BEM2D_OK = 0); bem2d_last_error()
 * returns the message of the calling thread's last failure.  The C++ drop-in headers under
 * include/bem2d/ map the statuses back onto the reference's exception tags
 * (lib/bem2d_exceptions.h:10-33).
 *
 * Ownership: the caller owns every buffer it passes.  Matrices, right-hand sides and solutions are
 * device pointers (complex128 = interleaved re, im; column-major with the stated leading
 * dimension) unless a parameter says otherwise.  A context owns a HIP stream (or borrows one via
 * bem2d_ctx_set_stream), a rocSOLVER handle and scratch memory; a mesh owns its device copy of the
 * geometry and the quadrature rules.  Calls on one context are issued in stream order; calls on one
 * context from several host threads must be serialised by the caller.
 */
#ifndef BEM2D_HIP_H
#define BEM2D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------------------------- */
#define BEM2D_OK 0
#define BEM2D_ERR_ARG 1           /* bad argument: ParameterException / SizeError / ArrayMismatch */
#define BEM2D_ERR_SINGULAR 2      /* exactly singular U in the LU: LapackError */
#define BEM2D_ERR_NOMEM 3         /* device allocation failed */
#define BEM2D_ERR_DEVICE 4        /* HIP / rocSOLVER failure */
#define BEM2D_ERR_UNSUPPORTED 5   /* element type, kernel or field kind not implemented */
#define BEM2D_ERR_NOT_CONVERGED 6 /* mixed-precision refinement missed its tolerance */

/* ---- element types (per element of a geometry) ------------------------------------------------
 * FLAT:      ConstElement(p1, p2)        param = (p1.x, p1.y, d.x, d.y, 0, 0, 0, 0), d = p2 - p1
 *            (lib/bem2d_element.h:72-94)
 * others:    AnalyticCurveElement(tstart, tend, curve)  (lib/bem2d_element.h:96-116)
 *            param = (tstart, tend, c0, c1, c2, c3, 0, 0) with the curve's constructor arguments
 *            CIRCLE (none), TREFOIL (none), KITE (none), ELLIPSE (rho), ELLIPSE_ARC (a, b, theta1,
 *            theta2), INV_ELLIPSE (alpha)            (lib/bem2d_curvetypes.h:13-99)            */
#define BEM2D_ELEM_FLAT 0
#define BEM2D_ELEM_CIRCLE 1
#define BEM2D_ELEM_TREFOIL 2
#define BEM2D_ELEM_KITE 3
#define BEM2D_ELEM_ELLIPSE 4
#define BEM2D_ELEM_ELLIPSE_ARC 5
#define BEM2D_ELEM_INV_ELLIPSE 6

/* ---- kernel kinds (lib/bem2d_kernel.h:12-98) -------------------------------------------------- */
#define BEM2D_KERNEL_NONE 0     /* no kernel term: the matrix is mass_scale * M (EvalIdent) */
#define BEM2D_KERNEL_SL 1       /* SingleLayer             (i/4) H0(kd) */
#define BEM2D_KERNEL_DL 2       /* DoubleLayer             -(ik/4) H1(kd)/d n_y.(y-x) */
#define BEM2D_KERNEL_CDL 3      /* ConjDoubleLayer         -(ik/4) H1(kd)/d n_x.(x-y) */
#define BEM2D_KERNEL_COMBINED 4 /* CombinedSingleConjDouble  CDL - i eta SL */
#define BEM2D_KERNEL_ID 5       /* IdKernel                1 */

/* ---- field kinds (lib/bem2d_fun.h:14-93) ------------------------------------------------------ */
#define BEM2D_FIELD_PLANE 0          /* PlaneWave          exp(ik d.x) */
#define BEM2D_FIELD_NORMAL_PLANE 1   /* NormalPlaneWave    ik (n.d) exp(ik d.x) */
#define BEM2D_FIELD_COMBINED_PLANE 2 /* CombinedPlaneWave  ik (n.d) exp(ik d.x) - i eta exp(ik d.x) */

typedef struct bem2d_ctx bem2d_ctx;
typedef struct bem2d_mesh bem2d_mesh;

/* Host-side description of a P0 geometry: element i is matrix row/column i (the FlatMap order of
 * lib/bem2d_geometry.h:48-57 with one PolBasis(0) per element).  index/next/prev mirror
 * Element::index/next/prev (lib/bem2d_element.h:36-58); -1 means "no neighbour". */
typedef struct {
    int64_t n;
    const int32_t* type;   /* [n] BEM2D_ELEM_* */
    const double* param;   /* [8 n] */
    const int64_t* index;  /* [n] */
    const int64_t* next;   /* [n] */
    const int64_t* prev;   /* [n] */
} bem2d_geometry;

/* QuadOption (lib/bem2d_quadrature.h:26-30): N Gauss points per element, L grading levels and
 * grading ratio sigma of the singular rules. */
typedef struct {
    int L;
    int N;
    double sigma;
} bem2d_quad_option;

/* ---- context ----------------------------------------------------------------------------------- */
int bem2d_ctx_create(bem2d_ctx** out, int device);
void bem2d_ctx_destroy(bem2d_ctx* ctx);
/* Issue all further work on a caller stream (hipStream_t; NULL is the null/default stream).
 * bem2d_ctx_reset_stream returns to the context's own non-blocking stream. */
int bem2d_ctx_set_stream(bem2d_ctx* ctx, void* stream);
int bem2d_ctx_reset_stream(bem2d_ctx* ctx);
void* bem2d_ctx_stream(bem2d_ctx* ctx);
int bem2d_synchronize(bem2d_ctx* ctx);
/* Frees the context's scratch memory (factor copies, far-field signatures, pivots), after the work
 * queued on its stream; it is allocated again on demand.  No counterpart in the reference. */
int bem2d_ctx_release_scratch(bem2d_ctx* ctx);
const char* bem2d_last_error(void);

/* Device memory helpers, so that host code above this layer needs no HIP headers. */
int bem2d_device_alloc(bem2d_ctx* ctx, int64_t bytes, void** out);
int bem2d_device_free(bem2d_ctx* ctx, void* p);
int bem2d_copy_to_device(bem2d_ctx* ctx, void* dst, const void* src, int64_t bytes);
int bem2d_copy_to_host(bem2d_ctx* ctx, void* dst, const void* src, int64_t bytes);

/* ---- launch shapes per size bucket (no counterpart in the reference) ----------------------------
 * The context's choices of launch shape, measured by the autotuner (bem2d_amd/autotune.py) and
 * written as a text table; bem2d_ctx_create loads $BEM2D_TUNE_FILE, else tuned/gfx950.tune beside
 * the library when present, else keeps the built-in defaults.
 * Assembly: a regular-pair launch of t tiles (8 x 8 elements each) runs split[b] wavefronts per
 * tile (1, 2 or 4) for the first bucket b with t < upper_tiles[b], one beyond the last bucket; at
 * most 8 buckets, upper_tiles increasing.  bem2d_tune_get_assembly returns the bucket count and
 * fills the arrays (8 entries each) when given.
 * FP32 LU: panels pair into depth-128 trailing updates while the trailing matrix has at least
 * lu_pair_tiles 64 x 64 tiles, group in fours into depth-256 updates while it has at least
 * lu_quad_tiles and in eights into depth-512 updates from lu_oct_tiles (the largest group whose
 * threshold is met wins); trailing matrices of at least lu_split_tiles tiles per side are
 * updated without the fused diagonal block (two workgroups per compute unit), the block then being
 * factorised by a kernel of its own.
 * Table lines: "asm_split <upper_tiles> <split>", "lu_pair_tiles <tiles>", "lu_quad_tiles <tiles>",
 * "lu_oct_tiles <tiles>" and "lu_split_tiles <tiles>"; '#' comments. */
int bem2d_tune_set_assembly(bem2d_ctx* ctx, int nbuckets, const int64_t* upper_tiles,
                            const int* split);
int bem2d_tune_get_assembly(const bem2d_ctx* ctx, int64_t* upper_tiles, int* split);
int bem2d_tune_set_lu_pair_tiles(bem2d_ctx* ctx, int tiles);
int bem2d_tune_get_lu_pair_tiles(const bem2d_ctx* ctx);
int bem2d_tune_set_lu_quad_tiles(bem2d_ctx* ctx, int tiles);
int bem2d_tune_get_lu_quad_tiles(const bem2d_ctx* ctx);
int bem2d_tune_set_lu_oct_tiles(bem2d_ctx* ctx, int tiles);
int bem2d_tune_get_lu_oct_tiles(const bem2d_ctx* ctx);
int bem2d_tune_set_lu_split_tiles(bem2d_ctx* ctx, int tiles);
int bem2d_tune_get_lu_split_tiles(const bem2d_ctx* ctx);
int bem2d_tune_load(bem2d_ctx* ctx, const char* path);

/* ---- mesh: device copy of the geometry, the Gauss-node tables and the singular rules ---------- */
/* Replaces Geometry::FlatMap + the rule construction at the top of DiscreteKernel
 * (lib/bem2d_quadrature.h:235-242): Gauss2D(N), AdaptedGauss1/2/3(N, L, sigma). */
int bem2d_mesh_create(bem2d_ctx* ctx, const bem2d_geometry* geom, bem2d_quad_option opts,
                      bem2d_mesh** out);
void bem2d_mesh_destroy(bem2d_mesh* mesh);
int64_t bem2d_mesh_size(const bem2d_mesh* mesh);
/* Number of coincident / adjacent (singular-rule) entries of the mesh. */
int64_t bem2d_mesh_special_count(const bem2d_mesh* mesh);

/* ---- assembly ---------------------------------------------------------------------------------- */
/* A[i + lda*j] = mass_scale * M_ij + kernel_scale * K_ij for all 0 <= i, j < n, where
 *   K = DiscreteKernel(geom, opts, kernel)   lib/bem2d_quadrature.h:231-294 (rule selection
 *       identical -> AdaptedGauss1 with the Duffy split of IntegrateIdent (:67-109), next ->
 *       AdaptedGauss3, prev -> AdaptedGauss2, otherwise Gauss2D; Integrate :37-65)
 *   M = EvalIdent(geom, opts)                lib/bem2d_quadrature.cpp:79-127
 * SoundSoftScattering::DiscretizeMatrix (lib/bem2d_soundsoftscattering.h:95-103) is
 * kernel = COMBINED, eta = the field's eta, mass_scale = 1, kernel_scale = 2.
 * k is a real wavenumber (freqtype with im == 0); A is a device pointer. */
int bem2d_assemble(bem2d_ctx* ctx, const bem2d_mesh* mesh, int kernel_kind, double k, double eta,
                   double mass_scale, double kernel_scale, void* A, int64_t lda);

/* ---- complex wavenumbers ---------------------------------------------------------------------
 * The kernels' and fields' branches for freqtype k with k.im != 0 (lib/bem2d_kernel.cpp:15-112,
 * lib/bem2d_fun.h:21-24, 45-49, 71-75), whose Hankel values come from AMOS zbesh through
 * BesselH0(complex) / BesselH1(complex) (lib/bem2d_mathroutines.cpp:44-75); here from the
 * library's own complex Hankel routine.  Same meaning of the other arguments as bem2d_assemble,
 * bem2d_rhs and bem2d_eval_single_layer; k_im == 0 calls those (the reference's k.im == 0
 * branch).  Needs k_re > 0 (BEM2D_ERR_UNSUPPORTED otherwise: AMOS covers the whole plane, the
 * resonance scans of the reference use Re k in [4, 8]); assembly needs n <= 8191.  eta stays real
 * as in CombinedSingleConjDouble(freqtype k, double eta). */
int bem2d_assemble_complex(bem2d_ctx* ctx, const bem2d_mesh* mesh, int kernel_kind, double k_re,
                           double k_im, double eta, double mass_scale, double kernel_scale,
                           void* A, int64_t lda);
int bem2d_rhs_complex(bem2d_ctx* ctx, const bem2d_mesh* mesh, int field_kind, double k_re,
                      double k_im, double eta, const double* dirs, int nrhs, double scale,
                      void* B, int64_t ldb);
int bem2d_eval_single_layer_complex(bem2d_ctx* ctx, const bem2d_mesh* mesh, double k_re,
                                    double k_im, const void* coef, const double* pts, int64_t npts,
                                    void* out);

/* ---- right-hand sides -------------------------------------------------------------------------- */
/* B[i + ldb*r] = scale * DotProdBasFuns(geom, opts, field(dirs[r]))_i   (lib/bem2d_quadrature.h:
 * 111-156) for r < nrhs; dirs is a HOST array of nrhs (x, y) directions (the field's direction,
 * used as given), read before the call returns; B is written in stream order (the call does not
 * wait for it).  DiscretizeRhs (lib/bem2d_soundsoftscattering.h:105-110) is scale = 2. */
int bem2d_rhs(bem2d_ctx* ctx, const bem2d_mesh* mesh, int field_kind, double k, double eta,
              const double* dirs, int nrhs, double scale, void* B, int64_t ldb);

/* ---- plain FP64 solve -------------------------------------------------------------------------- */
/* SolveSystem (lib/bem2d_mathroutines.cpp:323-351): LU with partial pivoting of A, then the
 * solve for nrhs right-hand sides; B is overwritten with the solution.  With preserve_A != 0, A
 * is left unchanged (the reference factors a copy); otherwise A is overwritten by its factors.
 * Returns BEM2D_ERR_SINGULAR when U has an exactly zero pivot (zgetrf info > 0). */
int bem2d_solve_fp64(bem2d_ctx* ctx, int64_t n, int64_t nrhs, void* A, int64_t lda, void* B,
                     int64_t ldb, int preserve_A);

/* ---- mixed-precision solve (FP32 LU on the MFMA cores + FP64 iterative refinement) -------------
 * Not in the reference, whose only solve is SolveSystem; it answers the same question to a stated
 * backward error.  bem2d_lu32 holds FP32 factors of an n x n matrix resident on the device
 * (8 n^2 bytes), so that one factorisation serves many right-hand sides and refinement passes.
 * Row order: static pivoting -- before the factorisation, which then makes no interchanges, each
 * column claims the row of its largest entry (contested rows go to the column where the entry is
 * largest, left-over positions take the unclaimed rows in order), so any row permutation of a
 * matrix whose column maxima lie in distinct rows is factorised in the order of those maxima.
 * The refinement reads the element growth of the factors; above 1e4 (or at a zero / non-finite
 * pivot) it reports BEM2D_ERR_NOT_CONVERGED at once; when plain refinement stalls (a step that
 * does not halve the backward error) GMRES-IR preconditioned by the FP32 factors continues. */
typedef struct bem2d_lu32 bem2d_lu32;
typedef struct {
    int iterations;          /* refinement steps taken (max over right-hand-side groups) */
    int fallback_fp64;       /* 1 if bem2d_solve_mixed fell back to the FP64 LU */
    int converged;           /* 1 if every column met tol */
    int gmres_steps;         /* GMRES iterations of GMRES-IR (0: plain refinement sufficed) */
    double backward_error;   /* max over columns of |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) */
    double matrix_norm_inf;  /* |A|_inf of the FP64 matrix */
    double growth;           /* element growth of the FP32 factors, max(max|U| / max|a_ij|, max|L|) */
} bem2d_ir_report;

int bem2d_lu32_create(bem2d_ctx* ctx, int64_t n, bem2d_lu32** out);
void bem2d_lu32_destroy(bem2d_lu32* lu);
int64_t bem2d_lu32_size(const bem2d_lu32* lu);
/* FP32 factorisation of the FP64 device matrix A64 (unchanged).  Factorising the same buffer
 * (address and lda) again replays a hipGraph of the launch sequence, captured on the second call
 * (BEM2D_LU_GRAPH=0 in the environment: never); bem2d_lu32_graph_replays counts the replays. */
int bem2d_lu32_factor(bem2d_ctx* ctx, bem2d_lu32* lu, const void* A64, int64_t lda);
int64_t bem2d_lu32_graph_replays(const bem2d_lu32* lu);
/* X (n x nrhs, device) = A64^{-1} B by refinement from x0 = 0 with the FP32 factors and FP64
 * residuals, until the backward error is <= tol or maxit steps; BEM2D_ERR_NOT_CONVERGED otherwise. */
int bem2d_lu32_refine(bem2d_ctx* ctx, bem2d_lu32* lu, const void* A64, int64_t lda, const void* B,
                      int64_t ldb, void* X, int64_t ldx, int nrhs, double tol, int maxit,
                      bem2d_ir_report* report);
/* One-shot: factor, refine (GMRES-IR when plain refinement stalls), and fall back to
 * bem2d_solve_fp64 -- the pivoted LU of SolveSystem -- if the factors are unusable or the
 * tolerance is missed.  B is overwritten with the solution; A64 is unchanged. */
int bem2d_solve_mixed(bem2d_ctx* ctx, int64_t n, int64_t nrhs, const void* A64, int64_t lda,
                      void* B, int64_t ldb, double tol, bem2d_ir_report* report);

/* No counterpart in the reference: bem2d_solve_mixed without any allocation, for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The FP32 factorisation and the refinement of bem2d_solve_mixed -- the same kernels
 * and step sequence, hence the same solution -- with every buffer in one fixed device workspace of
 * 20 MiB per GPU, held until the process ends: n <= 1024 and nrhs <= 8 (larger sizes return
 * BEM2D_ERR_ARG).  No GMRES-IR and no FP64 fallback (both allocate): factors with a zero pivot or
 * element growth above 1e4, or refinement that stalls or misses tol (<= 0: 1e-12), return
 * BEM2D_ERR_NOT_CONVERGED with B unchanged.  Waits for the solve before returning.  All calls on
 * one GPU share its workspace: calls that overlap on one GPU race (serialise them). */
int bem2d_solve_mixed_static(bem2d_ctx* ctx, int64_t n, int64_t nrhs, const void* A64, int64_t lda,
                             void* B, int64_t ldb, double tol, bem2d_ir_report* report);

/* ---- spectral drivers on assembled matrices -------------------------------------------------------
 * The reference's research routines (lib/bem2d_mathroutines.cpp), on the device with rocSOLVER /
 * rocBLAS.  Matrices are device pointers and are not modified except where stated. */
/* ChangeBasis (lib/bem2d_mathroutines.cpp:534-588): with the Cholesky factor M = U^H U (upper) of the
 * Hermitian positive definite mass matrix M, K is overwritten by U^{-H} K U^{-1} -- the matrix of K
 * in the M-orthonormal basis.  BEM2D_ERR_SINGULAR if M is not positive definite (the reference
 * does not check zpotrf's info). */
int bem2d_change_basis(bem2d_ctx* ctx, int64_t n, void* K, int64_t ldk, const void* M, int64_t ldm);
/* SingularValues (:266-320): s (device, n doubles) = the singular values of A, descending (zgesvd
 * without vectors). */
int bem2d_singular_values(bem2d_ctx* ctx, int64_t n, const void* A, int64_t lda, double* s);
/* L2NormCond (:208-264): *norm = s_max, *cond = s_max / s_min (host outputs). */
int bem2d_l2_norm_cond(bem2d_ctx* ctx, int64_t n, const void* A, int64_t lda, double* norm,
                       double* cond);
/* HermitianEigenvalues (:353-410): w (device, n doubles) = the eigenvalues of the Hermitian A,
 * ascending (zheev without vectors, upper triangle referenced). */
int bem2d_hermitian_eigenvalues(bem2d_ctx* ctx, int64_t n, const void* A, int64_t lda, double* w);
/* Eigenvalues (lib/bem2d_mathroutines.cpp:438-500, zgeev with no eigenvectors there): all n
 * eigenvalues of the general complex matrix A (device, unchanged) into w (device, n complex128), in
 * the order of the diagonal of the computed triangular form (LAPACK's order may differ; the set is
 * the same).  Hessenberg reduction + multishift QR on the GPU.  BEM2D_ERR_NOT_CONVERGED when the QR
 * iteration does not converge. */
int bem2d_eigenvalues(bem2d_ctx* ctx, int64_t n, const void* A, int64_t lda, void* w);
/* NumRange (:590-653) without the file: the extreme eigenvalues of (e^{i t} A + (e^{i t} A)^H) / 2 at
 * t = i pi / nsteps, i < nsteps, and the 2 nsteps + 1 boundary points (re, im) of the numerical
 * range built from them as the reference builds them; points is a HOST array of 2 (2 nsteps + 1)
 * doubles.  n <= 65535. */
int bem2d_numerical_range(bem2d_ctx* ctx, int64_t n, const void* A, int64_t lda, int nsteps,
                          double* points);

/* ---- potential and far field ------------------------------------------------------------------- */
/* out[p] = sum_j coef[j] * int_{elem j} (i/4) H0(k |P_p - y|) ds(y) with Gauss1D(N):
 * EvalKernel<SingleLayer> (lib/bem2d_quadrature.h:184-228) as used by EvalSol
 * (lib/bem2d_soundsoftscattering.h:149-168; EvalSol = u_inc - out).  coef, pts (2 doubles per
 * point) and out are device pointers. */
int bem2d_eval_single_layer(bem2d_ctx* ctx, const bem2d_mesh* mesh, double k, const void* coef,
                            const double* pts, int64_t npts, void* out);

/* Far field of u = u_inc - S phi (not in the reference; derived from EvalSol's representation):
 * out[a + nang*r] = -(exp(i pi/4)/sqrt(8 pi k)) sum_j coef[j + ldc*r] sum_t
 *                   exp(-i k (cos th_a, sin th_a).y_t) |Gamma_j'(t)| w_t.
 * coef, theta and out are device pointers. */
int bem2d_farfield(bem2d_ctx* ctx, const bem2d_mesh* mesh, double k, const void* coef,
                   int64_t ldc, int nrhs, const double* theta, int64_t nang, void* out);

/* ---- several basis functions per element -------------------------------------------------------
 * ConstBasis, PolBasis, LinBasis, WaveBasis, WavePolBasis (lib/bem2d_basistypes.h:16-99) with the
 * AddBasis helpers' lists (lib/bem2d_basistypes.cpp:12-46):
 *   PolBasis::AddBasis(d)        {POL 0, POL 1, ..., POL d}
 *   LinBasis::AddBasis           {LIN +1 (t), LIN -1 (1 - t)}
 *   WaveBasis::AddBasis(k)       {WAVE dir +1, WAVE dir -1}
 *   WavePolBasis::AddBasis(d, k) {WAVEPOL (0, +1), (0, -1), (1, +1), (1, -1), ..., (d, -1)}
 * The same nb bases sit on every element; row / column e * nb + b is basis b of element e
 * (Geometry::FlatMap, lib/bem2d_geometry.h:33-57, when each AddBasis call adds its basis to all
 * elements).  nb <= BEM2D_MAX_BASES. */
#define BEM2D_BASIS_CONST 0   /* 1 */
#define BEM2D_BASIS_POL 1     /* t^degree */
#define BEM2D_BASIS_LIN 2     /* t (degree = sign = +1), 1 - t (sign = -1) */
#define BEM2D_BASIS_WAVE 3    /* exp(i k direction t), k = k_re + i k_im */
#define BEM2D_BASIS_WAVEPOL 4 /* exp(i k direction t) t^degree */
#define BEM2D_MAX_BASES 8
typedef struct {
    int kind;         /* BEM2D_BASIS_* */
    int degree;       /* POL / WAVEPOL degree (0..64); LIN: the sign, -1 for 1 - t */
    double direction; /* WAVE / WAVEPOL */
    double k_re, k_im; /* WAVE / WAVEPOL wavenumber (freqtype) */
} bem2d_basis;

/* The (n nb) x (n nb) Galerkin matrix mass_scale * M + kernel_scale * K with the bases: entries
 * sum_q conj(phi_a(x_q)) K(x_q, y_q) phi_b(y_q) |Gamma_i'| |Gamma_j'| w_q over the same rules as
 * bem2d_assemble (Integrate / IntegrateIdent, lib/bem2d_quadrature.h:37-109, DiscreteKernel
 * :231-294) and the mass matrix sum_t |Gamma'| conj(phi_a) phi_b w_t (EvalIdent,
 * lib/bem2d_quadrature.cpp:79-127).  k = k_re + i k_im (k_im != 0: the complex-k kernels, Re k > 0,
 * n <= 8191).  lda >= n nb. */
int bem2d_assemble_basis(bem2d_ctx* ctx, const bem2d_mesh* mesh, const bem2d_basis* bases, int nb,
                         int kernel_kind, double k_re, double k_im, double eta,
                         double mass_scale, double kernel_scale, void* A, int64_t lda);
/* B[(i nb + b) + ldb r] = scale * <f_r, phi_b> on element i (DotProdBasFuns with bases,
 * lib/bem2d_quadrature.h:111-128); dirs as bem2d_rhs; complex k as bem2d_rhs_complex. */
int bem2d_rhs_basis(bem2d_ctx* ctx, const bem2d_mesh* mesh, const bem2d_basis* bases, int nb,
                    int field_kind, double k_re, double k_im, double eta, const double* dirs,
                    int nrhs, double scale, void* B, int64_t ldb);
/* EvalKernel<SingleLayer> with bases (lib/bem2d_quadrature.h:184-228): out[p] =
 * sum_j int_{elem j} (i/4) H0(k |P_p - y|) sum_b coef[j nb + b] phi_b ds(y). */
int bem2d_eval_single_layer_basis(bem2d_ctx* ctx, const bem2d_mesh* mesh, const bem2d_basis* bases,
                                  int nb, double k_re, double k_im, const void* coef,
                                  const double* pts, int64_t npts, void* out);
/* bem2d_farfield with bases: the density on element j is sum_b coef[j nb + b + ldc r] phi_b. */
int bem2d_farfield_basis(bem2d_ctx* ctx, const bem2d_mesh* mesh, const bem2d_basis* bases, int nb,
                         double k, const void* coef, int64_t ldc, int nrhs, const double* theta,
                         int64_t nang, void* out);
/* Host evaluation of the bases: out[2 (q nb + b) + {0, 1}] = phi_b(t[q]) (Basis::operator()). */
int bem2d_basis_values(const bem2d_basis* bases, int nb, const double* t, int64_t count,
                       double* out);

/* ---- arc-length parameterisation (host) --------------------------------------------------------
 * Curve::Length(t0, t1) (lib/bem2d_curve.cpp:11-25) and the parameters t_i of arc lengths s_i
 * (AnalyticCurve::ParameterizeArc, lib/bem2d_shape.cpp:332-366) of curve type BEM2D_ELEM_CIRCLE ..
 * BEM2D_ELEM_INV_ELLIPSE with constructor arguments args[4].  The reference integrates |Gamma'| by
 * GSL's qag (relative tolerance 1e-4) and dt/ds = 1/|Gamma'| by GSL's adaptive RK2 (absolute
 * tolerance 1e-10); here both are computed to rounding level (adaptive Gauss-Legendre, Newton's
 * method), so parameters agree with the reference's to its ODE tolerance.  s must be
 * non-decreasing in [0, Length]; s_i >= Length gives t_i = 1. */
int bem2d_curve_length(int curve_type, const double* args, double t0, double t1, double* out);
int bem2d_arclength_params(int curve_type, const double* args, const double* s, int64_t count,
                           double* t);

/* ---- serving engine: continuous batching of scattering solves ----------------------------------
 * Not in the reference, whose only driver is one SoundSoftScattering at a time
 * (lib/bem2d_soundsoftscattering.h:74-126).  A request is one such problem -- DiscretizeMatrix
 * (A = M + 2K, combined field), DiscretizeRhs (nrhs plane waves b = 2<f, phi>), Solve, and the far
 * field at nang angles -- solved on the mixed path (FP32 MFMA LU + FP64 refinement to the
 * request's backward-error tolerance) or the plain FP64 path (rocSOLVER).  The engine owns one
 * preallocated HBM arena and admits queued requests first-in first-out as their footprint fits;
 * each engine step advances every admitted request by one stage with one batched launch per
 * kernel per stage, so requests of different sizes and arrival times share launches.  Results
 * (x and u_inf) are copied to host memory; the device memory returns to the arena at once.
 * A mixed request whose FP32 factors are unusable (zero pivot, element growth above 1e4) or whose
 * refinement stalls or misses its tolerance is solved by the FP64 LU instead (fallback_fp64 in
 * its result).  A request whose FP64 LU meets an exactly zero pivot (zgetrf info > 0) finishes
 * with BEM2D_ERR_SINGULAR, the reference's LapackError. */
#define BEM2D_PRECISION_MIXED 0
#define BEM2D_PRECISION_FP64 1

typedef struct bem2d_engine bem2d_engine;

typedef struct {
    bem2d_geometry geom;     /* host arrays, copied at submission */
    bem2d_quad_option opts;
    double k;                /* real wavenumber > 0 */
    double eta;              /* coupling of the combined-field operator and of CombinedPlaneWave */
    int nrhs;                /* plane waves */
    int precision;           /* BEM2D_PRECISION_* */
    const double* dirs;      /* host, 2 nrhs: directions (x, y), used as given */
    int64_t nang;            /* far-field angles (0: none) */
    const double* theta;     /* host, nang */
    double tol;              /* mixed path: backward-error target (<= 0: 1e-12) */
} bem2d_request;

typedef struct {
    int status;              /* BEM2D_OK, or the failure of this request */
    int precision;
    int64_t n;
    int nrhs;
    int iterations;          /* refinement steps (mixed path) */
    double backward_error;   /* mixed path: max over columns (the last refinement step's) */
    double t_arrival, t_admit, t_done; /* engine clock (s): latency = t_done - t_arrival */
    int fallback_fp64;       /* mixed request solved by the FP64 LU after its FP32 factors had a
                                zero pivot or too much growth, or refinement stalled */
} bem2d_result_info;

typedef struct {
    int64_t submitted, admitted, completed, failed;
    int64_t assembly_launches, lu_step_launches, refine_step_launches;
    int64_t arena_bytes, arena_in_use, queued, active;
} bem2d_engine_stats;

/* Device bytes one request holds while admitted. */
int64_t bem2d_engine_footprint(int64_t n, int nrhs, int64_t nang, int precision,
                               bem2d_quad_option opts);
int bem2d_engine_create(bem2d_ctx* ctx, int64_t hbm_bytes, int max_active, bem2d_engine** out);
void bem2d_engine_destroy(bem2d_engine* eng);
double bem2d_engine_now(const bem2d_engine* eng); /* engine clock, seconds since creation */
/* Queues a request (t_arrival < 0: now); *ticket identifies it. */
int bem2d_engine_submit(bem2d_engine* eng, const bem2d_request* req, double t_arrival,
                        int64_t* ticket);
/* One engine step (never blocks on the GPU except for mesh uploads); *unfinished = requests
 * queued or in flight. */
int bem2d_engine_step(bem2d_engine* eng, int* unfinished);
int bem2d_engine_poll(bem2d_engine* eng, int64_t ticket, int* done);
/* Copies a finished request's solution (n x nrhs complex128, column-major) and far field
 * (nang x nrhs) into host memory (either may be NULL) and forgets the request. */
int bem2d_engine_result(bem2d_engine* eng, int64_t ticket, void* x, void* farfield,
                        bem2d_result_info* info);
int bem2d_engine_get_stats(const bem2d_engine* eng, bem2d_engine_stats* out);
/* Replays a trace: request q is submitted when the engine clock reaches (start + arrival[q]) s,
 * and the engine steps until all are done; tickets[q] receives each ticket. */
int bem2d_engine_run_trace(bem2d_engine* eng, const bem2d_request* reqs, const double* arrival,
                           int count, int64_t* tickets);

/* ---- kernel timing ------------------------------------------------------------------------------
 * With profiling on, the context records HIP events around each launch of the instrumented kernels
 * on the stream they run on.  bem2d_profile_read synchronises and returns the summed duration (ms)
 * and the number of launches since the last reset.  Kernel ids: */
#define BEM2D_PROF_ASSEMBLE_REGULAR 0 /* regular-pair assembly kernel */
#define BEM2D_PROF_ASSEMBLE_SPECIAL 1 /* coincident / adjacent assembly kernel */
#define BEM2D_PROF_LU 2               /* LU factorisation (FP64 rocSOLVER or the FP32 MFMA LU) */
#define BEM2D_PROF_LU_UPDATE 3        /* FP32 LU trailing-update kernel (launches made eagerly; a
                                         factorisation replayed from a hipGraph records none) */
#define BEM2D_PROF_COUNT 4
int bem2d_profile_enable(bem2d_ctx* ctx, int on);
int bem2d_profile_reset(bem2d_ctx* ctx);
int bem2d_profile_read(bem2d_ctx* ctx, int kernel_id, double* total_ms, int64_t* launches);

/* ---- quadrature rules (host) ---------------------------------------------------------------------
 * bem2d_gauss_legendre: n-point Gauss-Legendre rule on [0, 1], nodes ascending (replaces Gauss(),
 * lib/bem2d_quadrature.cpp:10-56).  bem2d_quad_rule: node list of Gauss2D (which = 0),
 * AdaptedGauss1 (1), AdaptedGauss2 (2) or AdaptedGauss3 (3) in the reference's node order
 * (lib/bem2d_quadpoints.cpp:166-318); returns the node count, or -1 on bad arguments.  With null
 * x/y/w only the count is returned. */
int bem2d_gauss_legendre(int n, double* x, double* w);
int bem2d_quad_rule(int which, int N, int L, double sigma, double* x, double* y, double* w);

/* ---- special functions ------------------------------------------------------------------------- */
/* BesselH0(double), BesselH1(double) (lib/bem2d_mathroutines.h:19-20, .cpp:20-42) for n real
 * arguments x > 0: out[4 i .. 4 i + 3] = Re H0, Im H0, Re H1, Im H1 of x[i].
 * bem2d_hankel01 runs on the device (x, out device pointers); bem2d_hankel01_host runs the same
 * FP64 routine on the host (host pointers) and needs no GPU. */
int bem2d_hankel01(bem2d_ctx* ctx, const double* x, int64_t n, double* out);
int bem2d_hankel01_host(const double* x, int64_t n, double* out);
/* BesselH0(complex), BesselH1(complex) (lib/bem2d_mathroutines.cpp:44-75, AMOS zbesh with
 * kode = 1, n = 1) at the n complex points z (interleaved re, im; Re z > 0); out as for
 * bem2d_hankel01.  phi (|phi| < pi/2) selects the integration path of the routine's middle region
 * (2 <= |z| < 20) and must be within 0.05 of every arg z -- one argument per call, as in an
 * assembly, where z = k d and phi = arg k.  Device and host versions as above. */
int bem2d_hankel01_complex(bem2d_ctx* ctx, double phi, const double* z, int64_t n, double* out);
int bem2d_hankel01_complex_host(double phi, const double* z, int64_t n, double* out);

#ifdef __cplusplus
}
#endif

#endif /* BEM2D_HIP_H */
