// Library-internal definitions shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/bem2d_hip.h"

constexpr int kMaxNq = 64; // QuadOption.N accepted by the device path (LDS tiles are sized by N)

enum {
    SCRATCH_SMALL = 0,    // right-hand-side directions and other tiny uploads
    SCRATCH_FARFIELD = 1, // far-field signatures, nang x n complex128
    SCRATCH_FACTOR = 2,   // copy of A when the caller's matrix must be preserved
    SCRATCH_PIVOT = 3,    // pivots + info of the FP64 LU
    SCRATCH_CONTOUR = 4,  // Hankel contour table of a complex wavenumber
    SCRATCH_SPECTRAL = 5, // rocSOLVER info, off-diagonals, eigenvalues of the spectral drivers
    SCRATCH_EVAL = 6,     // potential: source-node weights and per-chunk partial sums
    SCRATCH_COUNT = 7
};

struct ScratchBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct MixedState; // mixed-precision solver workspace (mixed.hip)

// Launch-shape choices per size bucket, installed by the autotuner's table (bem2d_tune_*).
constexpr int kMaxTuneBuckets = 8;
struct TuneState {
    // regular assembly: a launch of t tiles runs asm_split[b] wavefronts per tile for the first
    // bucket b with t < asm_upper[b], one wavefront per tile beyond the last bucket
    int nasm = 1;
    int64_t asm_upper[kMaxTuneBuckets] = {16384};
    int asm_split[kMaxTuneBuckets] = {4};
    // FP32 LU: panels pair into depth-128 updates while the trailing matrix has >= this many tiles
    int lu_pair_tiles = 64;
    // four panels into depth-256 updates while it has >= this many, eight into depth-512 from
    int lu_quad_tiles = 96;
    int lu_oct_tiles = 150;
    int lu_split_tiles = 64;   // trailing tiles from which the diagonal block is factorised apart
};

struct ProfSlot {
    std::vector<hipEvent_t> pending; // start/stop pairs not yet read
    double total_ms = 0.0;
    int64_t launches = 0;
};

// Quadrature rule tables of one QuadOption (AdaptedGauss1/2/3 and Gauss1D), built and uploaded once
// per context and shared by every mesh of that option: they depend on (N, L, sigma) only.
struct RuleSet {
    int N = 0, L = 0;
    double sigma = 0.0;
    int m1 = 0, m2 = 0, m3 = 0;          // node counts of AdaptedGauss1/2/3
    double* d_rules = nullptr;           // AG1 x|y|w, AG2 x|y|w, AG3 x|y|w, abscissa blocks
    double* d_g1d = nullptr;             // Gauss1D x | w
};

struct bem2d_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;           // second stream: singular-pair assembly beside the regular
    hipStream_t capture = nullptr;       // stream that hipGraph captures record on (never runs work)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    rocblas_handle blas = nullptr;
    ScratchBuf scratch[SCRATCH_COUNT];
    double* stage = nullptr;             // pinned staging of small host uploads (rhs directions)
    size_t stage_bytes = 0;
    hipEvent_t ev_stage = nullptr;       // recorded after the last copy out of stage
    MixedState* mixed = nullptr;
    bool profiling = false;
    ProfSlot prof[BEM2D_PROF_COUNT];
    TuneState tune;
    std::mutex rules_mu;
    std::vector<RuleSet*> rulesets;      // owned; freed with the context
};

// Records a start event for kernel id when profiling is on; returns the event index (or -1).
void bem2d_internal_prof_begin(bem2d_ctx* c, int id, hipStream_t s = nullptr);
void bem2d_internal_prof_end(bem2d_ctx* c, int id, hipStream_t s = nullptr);

struct DevArena;

struct bem2d_mesh {
    bem2d_ctx* ctx = nullptr;
    DevArena* arena = nullptr;           // source of the device arrays (null: hipMalloc)
    int64_t n = 0;
    bem2d_quad_option opts{};
    int nq = 0;
    int m1 = 0, m2 = 0, m3 = 0;          // node counts of AdaptedGauss1/2/3
    int64_t nspecial = 0;
    std::vector<int32_t> h_type;
    std::vector<double> h_param;
    int32_t* d_type = nullptr;
    double* d_param = nullptr;
    int64_t* d_adj = nullptr;            // index | next | prev, n each
    double* d_nodes = nullptr;           // 5 planes of nq*n (x, y, nx, ny, |Gamma'| w), then
                                         // 3 planes of n (bounding circle cx, cy, r)
    const double* d_rules = nullptr;     // the context's rule tables of opts (RuleSet)
    const double* d_g1d = nullptr;
    int64_t* d_special = nullptr;        // (i, j, rule) triples
    void* d_block = nullptr;             // one allocation holding d_type, d_param, d_adj, d_special
    double* bounds() const { return d_nodes + 5 * (size_t)n * nq; }
};

// Pinned host memory for the uploads of a mesh built inside a stream-ordered pipeline: get(bytes)
// returns a buffer that stays untouched until the work enqueued on the mesh's stream has passed
// the copies out of it (the caller's promise), so mesh creation neither waits nor stages.
struct HostStager {
    virtual void* get(size_t bytes) = 0;
    virtual ~HostStager() {}
};

// Device memory source for meshes, factorisations and refinement workspaces; a null arena means
// hipMalloc / hipFree.
struct DevArena {
    virtual void* alloc(size_t bytes) = 0;
    virtual void release(void* p) = 0;
    virtual ~DevArena() {}
};
void* bem2d_internal_dalloc(DevArena* a, size_t bytes);
void bem2d_internal_dfree(DevArena* a, void* p);

int bem2d_internal_fail(int code, const char* fmt, ...);
void* bem2d_internal_scratch(bem2d_ctx* c, int slot, size_t bytes);
// Contour table of the complex wavenumber kr + i ki in device scratch (SCRATCH_CONTOUR), uploaded in
// stream order on c->stream; *nt = its node count.
int bem2d_internal_upload_contour(bem2d_ctx* c, double kr, double ki, const double** d_tab,
                                  int* nt);
namespace bem2d_dev {
struct BasisSet;
}
// out = -(exp(i pi/4) / sqrt(8 pi k)) F coef on rocBLAS zgemm (F: nang x ncols far-field
// signatures, ld nang), on c->stream.
int bem2d_internal_farfield_gemm(bem2d_ctx* c, int64_t ncols, double k, const void* F, int64_t nang,
                                 const void* coef, int64_t ldc, int nrhs, void* out);
// Single-layer potential of the density sum_b coef[j nb + b] phi_b at npts device points (real
// k_im == 0, or complex k), on c->stream.
int bem2d_internal_potential(bem2d_ctx* c, const bem2d_mesh* m, const bem2d_dev::BasisSet& bs,
                             double k_re, double k_im, const void* coef, const double* pts,
                             int64_t npts, void* out);
void bem2d_internal_mixed_release(bem2d_ctx* c);
void bem2d_internal_gauss01(int n, std::vector<double>& x, std::vector<double>& w);
size_t bem2d_internal_mesh_bytes(int64_t n, bem2d_quad_option q);
// Mesh whose device arrays come from arena (null: hipMalloc), built on stream s.  With a stager the
// uploads leave from its pinned buffer and nothing is waited for; without one they leave from
// pageable memory and the call synchronises s before returning.
int bem2d_internal_mesh_create(bem2d_ctx* c, const bem2d_geometry* g, bem2d_quad_option q,
                               DevArena* arena, hipStream_t s, HostStager* stager,
                               bem2d_mesh** out);
void bem2d_internal_rulesets_release(bem2d_ctx* c);
int bem2d_internal_stage_dirs(bem2d_ctx* c, const double* dirs, int nrhs, const double** out);
int bem2d_internal_rhs(const bem2d_mesh* m, int kind, double k, double eta, const double* d_dirs,
                       int nrhs, double scale, void* B, int64_t ldb, hipStream_t s);
int bem2d_internal_farfield(const bem2d_mesh* m, double k, const void* coef, int64_t ldc,
                            int nrhs, const double* theta, int64_t nang, void* out, void* F,
                            hipStream_t s);
// Assembly of count meshes in batched launches on stream s (see bem2d_assemble for the meaning).
int bem2d_internal_assemble_many(bem2d_ctx* c, hipStream_t s, int kind,
                                 const bem2d_mesh* const* meshes, const double* k,
                                 const double* eta, double mscale, double kscale,
                                 void* const* A, const int64_t* lda, int count);

#define FAIL(code, ...) bem2d_internal_fail((code), __VA_ARGS__)
#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return bem2d_internal_fail(BEM2D_ERR_DEVICE, "%s failed: %s", #expr,                  \
                                       hipGetErrorString(e_));                                    \
    } while (0)

// ---- mixed-precision factorisation and refinement, several problems per launch ------------------
int bem2d_internal_lu32_create(bem2d_ctx* c, int64_t n, DevArena* arena, bem2d_lu32** out);
size_t bem2d_internal_lu32_bytes(int64_t n);
// cast + |A|_inf + diagonal block 0, then one panel step per call (each problem at its own step)
int bem2d_internal_lu_begin(hipStream_t s, bem2d_lu32* const* lus, const void* const* A64,
                            const int64_t* lda, int count);
// split_tiles > 0 overrides the context's lu_split_tiles (the engine splits every update: its LU
// steps share the device with other requests' assemblies, and the split update leaves room on a
// compute unit for their wavefronts)
int bem2d_internal_lu_advance(hipStream_t s, bem2d_lu32* const* lus, int count,
                              int split_tiles = 0);
bool bem2d_internal_lu_done(const bem2d_lu32* f);
int64_t bem2d_internal_lu_n(const bem2d_lu32* f);

struct RefSpec { // x = A^{-1} b by refinement with the FP32 factors lu
    bem2d_lu32* lu;
    const void* A64;
    int64_t lda;
    const void* B;
    int64_t ldb;
    void* X;
    int64_t ldx;
    int nrhs;
    double tol;      // backward-error target
    int maxit;       // refinement steps at most
    DevArena* ws;    // source of this problem's workspace (null: the job's arena)
};
struct RefJob;
size_t bem2d_internal_ref_bytes(int64_t n, int nrhs);
// pinned / dnorms: optional host (pinned) and device buffers of >= kRefNormSlots x (units) 64-bit
// words for the read-backs (null: the job allocates its own, the device one from arena)
constexpr int kRefNormSlots = 32;
int bem2d_internal_ref_units(int nrhs);
int bem2d_internal_ref_create(bem2d_ctx* c, DevArena* arena, const RefSpec* specs, int count,
                              unsigned long long* pinned, unsigned long long* dnorms,
                              RefJob** out);
void bem2d_internal_ref_destroy(RefJob* j);
// enqueue the initial residual (first call) or one correction step, then an asynchronous read-back
int bem2d_internal_ref_step(RefJob* j, hipStream_t s);
bool bem2d_internal_ref_ready(RefJob* j);
int bem2d_internal_ref_collect(RefJob* j, bool* all_done);
bool bem2d_internal_ref_spec_active(const RefJob* j, int sidx);
void bem2d_internal_ref_report(const RefJob* j, int sidx, bem2d_ir_report* r);
bool bem2d_internal_ref_bad_pivot(const RefJob* j, int sidx);
int bem2d_internal_refine_sync(hipStream_t s, RefJob* j);

// Device address of the no-allocation solve's file-scope workspace on `device` (current): taken
// once per device with hipGetSymbolAddress, at the first context creation on it.
int bem2d_internal_static_scratch(int device, void** out);

// Read-only view of FP32 factors for code outside mixed.hip: the n x n leading block of F (ld np)
// holds L (unit diagonal implied) and U of P A, where row i of P A is row perm[i] of A.
struct Lu32View {
    const float2* F;
    int64_t n, np;
    const int* perm;
    const double* anorm; // |A|_inf (device)
};
Lu32View bem2d_internal_lu32_view(const bem2d_lu32* f);
// GMRES-IR (FP64 GMRES on M^{-1} A d = M^{-1} r, M = the FP32 factors) from the current X for the
// columns whose backward error is above tol; rep->backward_error / converged / gmres_steps are
// updated.  Allocates its workspace; runs on the context's stream.
int bem2d_internal_gmres_ir(bem2d_ctx* c, const bem2d_lu32* f, const void* A64, int64_t lda,
                            const void* B, int64_t ldb, void* X, int64_t ldx, int nrhs, double tol,
                            bem2d_ir_report* rep);
