// GMRES-IR: the fallback of the mixed-precision solve when plain iterative refinement stalls.
//
// Plain refinement x += M^{-1} (b - A x) with M = the FP32 factors converges only while
// ||I - M^{-1} A|| < 1, roughly cond(A) * growth * u32 < 1.  GMRES-IR (E. Carson and N. J. Higham,
// "Accelerating the solution of linear systems by iterative refinement in three precisions", SIAM
// J. Sci. Comput. 40 (2018)) keeps the outer loop -- the residual r = b - A x in FP64 and the
// update x += d in FP64 -- and solves the correction equation A d = r by GMRES on the
// left-preconditioned system M^{-1} A d = M^{-1} r in FP64, with M^{-1} applied through the FP32
// factors.  GMRES converges in few iterations whenever M^{-1} A has its eigenvalues clustered,
// which holds well beyond the range where plain refinement converges.
//
// One right-hand side at a time; rocBLAS for the FP64 products, the Krylov basis and its norms,
// ctrsv on the FP32 factors for M^{-1}; the Hessenberg least-squares problem (Givens rotations) on
// the host.  This is a fallback path: it favours robustness over launch count.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "../../include/bem2d_hip.h"
#include "internal.h"

namespace {

constexpr int kRestart = 40;       // Krylov dimension per GMRES cycle
constexpr int kMaxOuter = 12;      // outer refinement steps
constexpr double kInnerTol = 1e-8; // relative reduction of the preconditioned residual per cycle

using cd = std::complex<double>;

// y32[i] = (P v)[i] * s in FP32 (row i of P A is row perm[i] of A)
__global__ void gm_demote_kernel(int64_t n, const double2* __restrict__ v, const int* __restrict__ perm,
                                 double s, float2* __restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 a = v[perm[i]];
    y[i] = make_float2((float)(a.x * s), (float)(a.y * s));
}

// out[i] = y32[i] / s in FP64
__global__ void gm_promote_kernel(int64_t n, const float2* __restrict__ y, double inv_s,
                                  double2* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 a = y[i];
    out[i] = make_double2((double)a.x * inv_s, (double)a.y * inv_s);
}

// m[0] = max_i |v_i| (non-negative double bits; NaN counts as infinite), m zeroed beforehand
__global__ void gm_absmax_kernel(int64_t n, const double2* __restrict__ v, unsigned long long* m)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0;
    if (i < n) {
        a = hypot(v[i].x, v[i].y);
        if (!(a == a)) a = INFINITY;
    }
    for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_xor(a, off));
    if ((threadIdx.x & 63) == 0) atomicMax(m, (unsigned long long)__double_as_longlong(a));
}

__global__ void gm_zero_kernel(unsigned long long* m, int k)
{
    if ((int)threadIdx.x < k) m[threadIdx.x] = 0ull;
}

struct Gmres {
    bem2d_ctx* c;
    rocblas_handle h;
    hipStream_t s;
    Lu32View lu;
    const rocblas_double_complex* A;
    int64_t lda;
    double anorm;
    // workspace
    double2* V = nullptr;   // n x (kRestart + 1)
    double2* w = nullptr;   // n
    double2* r = nullptr;   // n
    double2* hv = nullptr;  // kRestart + 1 (projection coefficients)
    float2* y32 = nullptr;  // n
    unsigned long long* m = nullptr;
    int rc = BEM2D_OK;

    unsigned grid() const { return (unsigned)((lu.n + 255) / 256); }

    bool ok(rocblas_status st)
    {
        if (st != rocblas_status_success && rc == BEM2D_OK)
            rc = FAIL(BEM2D_ERR_DEVICE, "GMRES-IR: rocBLAS status %d", (int)st);
        return rc == BEM2D_OK;
    }

    double absmax(const double2* v)
    {
        gm_zero_kernel<<<1, 64, 0, s>>>(m, 1);
        gm_absmax_kernel<<<grid(), 256, 0, s>>>(lu.n, v, m);
        unsigned long long b = 0;
        if (hipMemcpyAsync(&b, m, sizeof b, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            if (rc == BEM2D_OK) rc = FAIL(BEM2D_ERR_DEVICE, "GMRES-IR: read-back failed");
            return INFINITY;
        }
        double d;
        std::memcpy(&d, &b, sizeof d);
        return d;
    }

    double nrm2(const double2* v)
    {
        double x = 0.0;
        ok(rocblas_dznrm2(h, (rocblas_int)lu.n, (const rocblas_double_complex*)v, 1, &x));
        return x;
    }

    // out = M^{-1} v, M = P^T L U (FP32), v scaled into FP32 range by its largest entry
    void precond(const double2* v, double2* out)
    {
        const double vm = absmax(v);
        const double sc = (vm > 1e-300 && vm < INFINITY) ? 1.0 / vm : 1.0;
        gm_demote_kernel<<<grid(), 256, 0, s>>>(lu.n, v, lu.perm, sc, y32);
        if (!ok(rocblas_ctrsv(h, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_unit,
                              (rocblas_int)lu.n, (const rocblas_float_complex*)lu.F,
                              (rocblas_int)lu.np, (rocblas_float_complex*)y32, 1)))
            return;
        if (!ok(rocblas_ctrsv(h, rocblas_fill_upper, rocblas_operation_none,
                              rocblas_diagonal_non_unit, (rocblas_int)lu.n,
                              (const rocblas_float_complex*)lu.F, (rocblas_int)lu.np,
                              (rocblas_float_complex*)y32, 1)))
            return;
        gm_promote_kernel<<<grid(), 256, 0, s>>>(lu.n, y32, 1.0 / sc, out);
    }

    // r = b - A x
    void residual(const double2* b, const double2* x)
    {
        const rocblas_double_complex one = {1.0, 0.0}, mone = {-1.0, 0.0};
        (void)hipMemcpyAsync(r, b, sizeof(double2) * lu.n, hipMemcpyDeviceToDevice, s);
        ok(rocblas_zgemv(h, rocblas_operation_none, (rocblas_int)lu.n, (rocblas_int)lu.n, &mone,
                         A, (rocblas_int)lda, (const rocblas_double_complex*)x, 1, &one,
                         (rocblas_double_complex*)r, 1));
    }

    // One GMRES(kRestart) cycle on M^{-1} A d = M^{-1} r, then x += d.  Returns the iterations.
    int cycle(double2* x)
    {
        const int64_t n = lu.n;
        const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0}, mone = {-1.0, 0.0};
        precond(r, V); // v0 = M^{-1} r / beta
        const double beta = nrm2(V);
        if (rc || !(beta > 0) || !std::isfinite(beta)) return 0;
        const double ib = 1.0 / beta;
        ok(rocblas_zdscal(h, (rocblas_int)n, &ib, (rocblas_double_complex*)V, 1));
        std::vector<cd> H((size_t)(kRestart + 1) * kRestart, cd(0)), g(kRestart + 1, cd(0));
        std::vector<double> cs(kRestart, 0.0);
        std::vector<cd> sn(kRestart, cd(0));
        std::vector<cd> hh(kRestart + 1);
        g[0] = beta;
        int k = 0;
        for (; k < kRestart && rc == BEM2D_OK; ++k) {
            double2* vk = V + (size_t)n * k;
            // w = M^{-1} A v_k
            ok(rocblas_zgemv(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n, &one, A,
                             (rocblas_int)lda, (const rocblas_double_complex*)vk, 1, &zero,
                             (rocblas_double_complex*)r, 1));
            precond(r, w);
            // classical Gram-Schmidt, twice: h = V^H w, w -= V h
            std::fill(hh.begin(), hh.end(), cd(0));
            for (int pass = 0; pass < 2 && rc == BEM2D_OK; ++pass) {
                ok(rocblas_zgemv(h, rocblas_operation_conjugate_transpose, (rocblas_int)n,
                                 (rocblas_int)(k + 1), &one, (const rocblas_double_complex*)V,
                                 (rocblas_int)n, (const rocblas_double_complex*)w, 1, &zero,
                                 (rocblas_double_complex*)hv, 1));
                ok(rocblas_zgemv(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)(k + 1),
                                 &mone, (const rocblas_double_complex*)V, (rocblas_int)n,
                                 (const rocblas_double_complex*)hv, 1, &one,
                                 (rocblas_double_complex*)w, 1));
                std::vector<cd> part(k + 1);
                if (hipMemcpyAsync(part.data(), hv, sizeof(cd) * (k + 1), hipMemcpyDeviceToHost, s) !=
                        hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess) {
                    rc = FAIL(BEM2D_ERR_DEVICE, "GMRES-IR: read-back failed");
                    break;
                }
                for (int i = 0; i <= k; ++i) hh[i] += part[i];
            }
            const double hn = nrm2(w);
            if (rc) break;
            hh[k + 1] = hn;
            if (hn > 0 && std::isfinite(hn)) {
                const double ih = 1.0 / hn;
                (void)hipMemcpyAsync(V + (size_t)n * (k + 1), w, sizeof(double2) * n,
                                     hipMemcpyDeviceToDevice, s);
                ok(rocblas_zdscal(h, (rocblas_int)n, &ih,
                                  (rocblas_double_complex*)(V + (size_t)n * (k + 1)), 1));
            }
            // previous rotations, then a new one that zeroes hh[k + 1]
            for (int i = 0; i < k; ++i) {
                const cd a = hh[i], b = hh[i + 1];
                hh[i] = cs[i] * a + sn[i] * b;
                hh[i + 1] = -std::conj(sn[i]) * a + cs[i] * b;
            }
            const double na = std::abs(hh[k]), nb = std::abs(hh[k + 1]);
            const double rr = std::hypot(na, nb);
            if (rr == 0.0) {
                cs[k] = 1.0;
                sn[k] = 0.0;
            } else if (na == 0.0) {
                cs[k] = 0.0;
                sn[k] = std::conj(hh[k + 1]) / nb;
            } else {
                const cd ph = hh[k] / na;
                cs[k] = na / rr;
                sn[k] = ph * std::conj(hh[k + 1]) / rr;
            }
            hh[k] = cs[k] * hh[k] + sn[k] * hh[k + 1];
            hh[k + 1] = 0.0;
            g[k + 1] = -std::conj(sn[k]) * g[k];
            g[k] = cs[k] * g[k];
            for (int i = 0; i <= k; ++i) H[(size_t)i * kRestart + k] = hh[i];
            if (std::abs(g[k + 1]) <= kInnerTol * beta || !(hn > 0 && std::isfinite(hn))) {
                ++k;
                break;
            }
        }
        if (rc || k == 0) return k;
        // y = H(0:k, 0:k)^{-1} g(0:k), d = V y, x += d
        std::vector<cd> y(k);
        for (int i = k - 1; i >= 0; --i) {
            cd acc = g[i];
            for (int j = i + 1; j < k; ++j) acc -= H[(size_t)i * kRestart + j] * y[j];
            const cd d = H[(size_t)i * kRestart + i];
            y[i] = d != cd(0) ? acc / d : cd(0);
        }
        if (hipMemcpyAsync(hv, y.data(), sizeof(cd) * k, hipMemcpyHostToDevice, s) != hipSuccess) {
            rc = FAIL(BEM2D_ERR_DEVICE, "GMRES-IR: upload failed");
            return k;
        }
        ok(rocblas_zgemv(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)k, &one,
                         (const rocblas_double_complex*)V, (rocblas_int)n,
                         (const rocblas_double_complex*)hv, 1, &one, (rocblas_double_complex*)x, 1));
        // the coefficients must stay valid until the product has read them
        if (hipStreamSynchronize(s) != hipSuccess && rc == BEM2D_OK)
            rc = FAIL(BEM2D_ERR_DEVICE, "GMRES-IR: update failed");
        return k;
    }
};

} // namespace

int bem2d_internal_gmres_ir(bem2d_ctx* c, const bem2d_lu32* f, const void* A64, int64_t lda,
                            const void* B, int64_t ldb, void* X, int64_t ldx, int nrhs, double tol,
                            bem2d_ir_report* rep)
{
    Gmres G{};
    G.c = c;
    G.h = c->blas;
    G.s = c->stream;
    G.lu = bem2d_internal_lu32_view(f);
    G.A = (const rocblas_double_complex*)A64;
    G.lda = lda;
    const int64_t n = G.lu.n;
    if (n > 2000000000 || lda > 2000000000)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "GMRES-IR: matrix too large for rocBLAS");
    if (hipMemcpyAsync(&G.anorm, G.lu.anorm, sizeof(double), hipMemcpyDeviceToHost, G.s) !=
        hipSuccess)
        return FAIL(BEM2D_ERR_DEVICE, "GMRES-IR: read-back failed");
    void* ws = nullptr;
    const size_t bytes = sizeof(double2) * ((size_t)n * (kRestart + 3) + kRestart + 1) +
                         sizeof(float2) * n + 64 + 5 * 256;
    if (hipMalloc(&ws, bytes) != hipSuccess)
        return FAIL(BEM2D_ERR_NOMEM, "GMRES-IR: workspace allocation failed");
    char* p = (char*)ws;
    auto carve = [&](size_t b) {
        char* q = p;
        p += (b + 255) & ~(size_t)255;
        return q;
    };
    G.V = (double2*)carve(sizeof(double2) * n * (kRestart + 1));
    G.w = (double2*)carve(sizeof(double2) * n);
    G.r = (double2*)carve(sizeof(double2) * n);
    G.hv = (double2*)carve(sizeof(double2) * (kRestart + 1));
    G.y32 = (float2*)carve(sizeof(float2) * n);
    G.m = (unsigned long long*)carve(64);
    double worst = 0.0;
    int steps = 0;
    for (int h = 0; h < nrhs && G.rc == BEM2D_OK; ++h) {
        const double2* b = (const double2*)B + ldb * h;
        double2* x = (double2*)X + ldx * h;
        const double bn = G.absmax(b);
        double omega = INFINITY;
        for (int outer = 0; outer <= kMaxOuter && G.rc == BEM2D_OK; ++outer) {
            G.residual(b, x);
            const double rn = G.absmax(G.r), xn = G.absmax(x);
            const double den = G.anorm * xn + bn;
            omega = den > 0 ? rn / den : (rn > 0 ? INFINITY : 0.0);
            if (!std::isfinite(omega)) omega = INFINITY;
            if (omega <= tol || outer == kMaxOuter) break;
            const int k = G.cycle(x);
            steps += k;
            if (k == 0) break;
        }
        worst = std::max(worst, omega);
    }
    (void)hipStreamSynchronize(G.s);
    (void)hipFree(ws);
    if (rep) {
        rep->gmres_steps += steps;
        rep->backward_error = worst;
        rep->converged = G.rc == BEM2D_OK && worst <= tol;
    }
    return G.rc;
}
