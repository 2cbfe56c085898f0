// Library-internal definitions shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <vector>

#include "../../include/bem2d_hip.h"

constexpr int kMaxNq = 64; // QuadOption.N accepted by the device path (LDS tiles are sized by N)

enum {
    SCRATCH_SMALL = 0,    // right-hand-side directions and other tiny uploads
    SCRATCH_FARFIELD = 1, // far-field signatures, nang x n complex128
    SCRATCH_FACTOR = 2,   // copy of A when the caller's matrix must be preserved
    SCRATCH_PIVOT = 3,    // pivots + info of the FP64 LU
    SCRATCH_COUNT = 4
};

struct ScratchBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct MixedState; // mixed-precision solver workspace (mixed.hip)

struct ProfSlot {
    std::vector<hipEvent_t> pending; // start/stop pairs not yet read
    double total_ms = 0.0;
    int64_t launches = 0;
};

struct bem2d_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    ScratchBuf scratch[SCRATCH_COUNT];
    MixedState* mixed = nullptr;
    bool profiling = false;
    ProfSlot prof[BEM2D_PROF_COUNT];
};

// Records a start event for kernel id when profiling is on; returns the event index (or -1).
void bem2d_internal_prof_begin(bem2d_ctx* c, int id);
void bem2d_internal_prof_end(bem2d_ctx* c, int id);

struct bem2d_mesh {
    bem2d_ctx* ctx = nullptr;
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
    double* d_rules = nullptr;           // AG1 x|y|w, AG2 x|y|w, AG3 x|y|w
    double* d_g1d = nullptr;             // Gauss1D x | w
    int64_t* d_special = nullptr;        // (i, j, rule) triples
    double* bounds() const { return d_nodes + 5 * (size_t)n * nq; }
};

int bem2d_internal_fail(int code, const char* fmt, ...);
void* bem2d_internal_scratch(bem2d_ctx* c, int slot, size_t bytes);
void bem2d_internal_mixed_release(bem2d_ctx* c);
void bem2d_internal_gauss01(int n, std::vector<double>& x, std::vector<double>& w);

#define FAIL(code, ...) bem2d_internal_fail((code), __VA_ARGS__)
#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return bem2d_internal_fail(BEM2D_ERR_DEVICE, "%s failed: %s", #expr,                  \
                                       hipGetErrorString(e_));                                    \
    } while (0)
