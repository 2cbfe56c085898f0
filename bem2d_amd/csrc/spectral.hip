// Spectral drivers on the device matrices: the change to the mass matrix's orthonormal basis, the
// singular values (L2 norm and condition number), Hermitian eigenvalues and the numerical range.
// These are the reference's research routines (lib/bem2d_mathroutines.cpp:208-320, 353-410,
// 534-653), which the normcomps/ numrangecomps/ resonances/ drivers run on the assembled
// matrices; here they run on rocSOLVER (zpotrf, zgesvd, zheev) and rocBLAS (ztrsm) on the GPU.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "../../include/bem2d_hip.h"
#include "internal.h"

namespace {

// dst (ld n) = src (ld lds), n x n complex128
__global__ void spectral_copy_kernel(int64_t n, const double2* __restrict__ src, int64_t lds,
                                     double2* __restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i < n) dst[i + n * j] = src[i + lds * j];
}

// H_b = (e^{i th_b} A + (e^{i th_b} A)^H) / 2 for the angles of the batch (blockIdx.z = b), each n x n
// with ld n at H + b n^2.  cos / sin of the angles come precomputed from the host.
__global__ void numrange_hermitian_kernel(int64_t n, const double2* __restrict__ A, int64_t lda,
                                          const double* __restrict__ cs, double2* __restrict__ H)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    const int b = blockIdx.z;
    if (i >= n) return;
    const double c = cs[2 * b], s = cs[2 * b + 1];
    const double2 a = A[i + lda * j], t = A[j + lda * i];
    // e a = (c a.x - s a.y) + i (c a.y + s a.x); conj(e t) = (c t.x - s t.y) - i (c t.y + s t.x)
    const double re = 0.5 * ((c * a.x - s * a.y) + (c * t.x - s * t.y));
    const double im = 0.5 * ((c * a.y + s * a.x) - (c * t.y + s * t.x));
    H[(int64_t)b * n * n + i + n * j] = make_double2(re, im);
}

// s[0 .. n) into decreasing order in place: odd-even transposition passes by one workgroup until a
// pass swaps nothing (the divide-and-conquer SVD leaves near-equal values out of order by rounding,
// so one or two passes do; n passes at most).
__global__ __launch_bounds__(1024) void sort_desc_kernel(double* __restrict__ s, int64_t n)
{
    __shared__ int swapped;
    for (int64_t pass = 0; pass < n + 2; ++pass) {
        if (threadIdx.x == 0) swapped = 0;
        __syncthreads();
        for (int parity = 0; parity < 2; ++parity) {
            for (int64_t i = 2 * (int64_t)threadIdx.x + parity; i + 1 < n; i += 2 * blockDim.x) {
                const double a = s[i], b = s[i + 1];
                if (a < b) {
                    s[i] = b;
                    s[i + 1] = a;
                    swapped = 1;
                }
            }
            __syncthreads();
        }
        const int any = swapped;
        __syncthreads();
        if (!any) break;
    }
}

int check_args(bem2d_ctx* c, int64_t n, const void* A, int64_t lda, const char* who)
{
    if (!c || !A || n < 1 || lda < n || n > 2000000000 || lda > 2000000000)
        return FAIL(BEM2D_ERR_ARG, "%s: bad argument", who);
    return BEM2D_OK;
}

// rocSOLVER's info (device int) to the host.
int read_info(bem2d_ctx* c, const int* dinfo, int* out)
{
    HIPCHK(hipMemcpyAsync(out, dinfo, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BEM2D_OK;
}

} // namespace

extern "C" int bem2d_change_basis(bem2d_ctx* c, int64_t n, void* K, int64_t ldk, const void* M,
                                  int64_t ldm)
{
    if (int rc = check_args(c, n, M, ldm, "bem2d_change_basis")) return rc;
    if (!K || ldk < n || ldk > 2000000000) return FAIL(BEM2D_ERR_ARG, "bem2d_change_basis: bad K");
    HIPCHK(hipSetDevice(c->device));
    double2* U = (double2*)bem2d_internal_scratch(c, SCRATCH_FACTOR, sizeof(double2) * n * n);
    int* info = (int*)bem2d_internal_scratch(c, SCRATCH_SPECTRAL, 256);
    if (!U || !info) return FAIL(BEM2D_ERR_NOMEM, "bem2d_change_basis: scratch allocation failed");
    spectral_copy_kernel<<<dim3((unsigned)((n + 255) / 256), (unsigned)n), 256, 0, c->stream>>>(
        n, (const double2*)M, ldm, U);
    HIPCHK(hipGetLastError());
    // M = U^H U (upper Cholesky factor), then K U^{-1}, then U^{-H} (K U^{-1})
    rocblas_status st = rocsolver_zpotrf(c->blas, rocblas_fill_upper, (rocblas_int)n,
                                         (rocblas_double_complex*)U, (rocblas_int)n, info);
    if (st != rocblas_status_success) return FAIL(BEM2D_ERR_DEVICE, "rocsolver_zpotrf status %d", (int)st);
    int h_info = 0;
    if (int rc = read_info(c, info, &h_info)) return rc;
    if (h_info > 0)
        return FAIL(BEM2D_ERR_SINGULAR, "bem2d_change_basis: the mass matrix is not positive "
                                        "definite (leading minor %d)", h_info);
    const rocblas_double_complex one = {1.0, 0.0};
    st = rocblas_ztrsm(c->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                       rocblas_diagonal_non_unit, (rocblas_int)n, (rocblas_int)n, &one,
                       (const rocblas_double_complex*)U, (rocblas_int)n,
                       (rocblas_double_complex*)K, (rocblas_int)ldk);
    if (st == rocblas_status_success)
        st = rocblas_ztrsm(c->blas, rocblas_side_left, rocblas_fill_upper,
                           rocblas_operation_conjugate_transpose, rocblas_diagonal_non_unit,
                           (rocblas_int)n, (rocblas_int)n, &one, (const rocblas_double_complex*)U,
                           (rocblas_int)n, (rocblas_double_complex*)K, (rocblas_int)ldk);
    if (st != rocblas_status_success) return FAIL(BEM2D_ERR_DEVICE, "rocblas_ztrsm status %d", (int)st);
    return BEM2D_OK;
}

extern "C" int bem2d_singular_values(bem2d_ctx* c, int64_t n, const void* A, int64_t lda, double* s)
{
    if (int rc = check_args(c, n, A, lda, "bem2d_singular_values")) return rc;
    if (!s) return FAIL(BEM2D_ERR_ARG, "bem2d_singular_values: null output");
    HIPCHK(hipSetDevice(c->device));
    double2* W = (double2*)bem2d_internal_scratch(c, SCRATCH_FACTOR, sizeof(double2) * n * n);
    char* small = (char*)bem2d_internal_scratch(c, SCRATCH_SPECTRAL, sizeof(double) * n + 512);
    if (!W || !small) return FAIL(BEM2D_ERR_NOMEM, "bem2d_singular_values: scratch allocation failed");
    int* info = (int*)small;
    double* E = (double*)(small + 256);
    spectral_copy_kernel<<<dim3((unsigned)((n + 255) / 256), (unsigned)n), 256, 0, c->stream>>>(
        n, (const double2*)A, lda, W);
    HIPCHK(hipGetLastError());
    const rocblas_int N = (rocblas_int)n;
    rocblas_double_complex* Wz = (rocblas_double_complex*)W;
    // Divide and conquer (zgesdd) first; the bidiagonal QR iteration of zgesvd, then one-sided
    // Jacobi (zgesvdj), on a fresh copy if it does not converge (strongly graded operators can
    // stall the QR iteration).  All three return the singular values in decreasing order.
    rocblas_status st = rocsolver_zgesdd(c->blas, rocblas_svect_none, rocblas_svect_none, N, N, Wz,
                                         N, s, nullptr, N, nullptr, N, info);
    if (st != rocblas_status_success) return FAIL(BEM2D_ERR_DEVICE, "rocsolver_zgesdd status %d", (int)st);
    int h_info = 0;
    if (int rc = read_info(c, info, &h_info)) return rc;
    if (h_info == 0) {
        sort_desc_kernel<<<1, 1024, 0, c->stream>>>(s, n);
        HIPCHK(hipGetLastError());
        return BEM2D_OK;
    }
    const int first_info = h_info;
    spectral_copy_kernel<<<dim3((unsigned)((n + 255) / 256), (unsigned)n), 256, 0, c->stream>>>(
        n, (const double2*)A, lda, W);
    st = rocsolver_zgesvd(c->blas, rocblas_svect_none, rocblas_svect_none, N, N, Wz, N, s, nullptr,
                          N, nullptr, N, E, rocblas_outofplace, info);
    if (st != rocblas_status_success) return FAIL(BEM2D_ERR_DEVICE, "rocsolver_zgesvd status %d", (int)st);
    if (int rc = read_info(c, info, &h_info)) return rc;
    if (h_info == 0) return BEM2D_OK;
    spectral_copy_kernel<<<dim3((unsigned)((n + 255) / 256), (unsigned)n), 256, 0, c->stream>>>(
        n, (const double2*)A, lda, W);
    double* residual = (double*)(small + 8);
    int* sweeps = (int*)(small + 16);
    st = rocsolver_zgesvdj(c->blas, rocblas_svect_none, rocblas_svect_none, N, N, Wz, N, 0.0,
                           residual, 100, sweeps, s, nullptr, N, nullptr, N, info);
    if (st != rocblas_status_success) return FAIL(BEM2D_ERR_DEVICE, "rocsolver_zgesvdj status %d", (int)st);
    if (int rc = read_info(c, info, &h_info)) return rc;
    if (h_info != 0)
        return FAIL(BEM2D_ERR_SINGULAR,
                    "bem2d_singular_values: SVD did not converge (zgesdd %d, zgesvd and zgesvdj %d)",
                    first_info, h_info);
    sort_desc_kernel<<<1, 1024, 0, c->stream>>>(s, n);
    HIPCHK(hipGetLastError());
    return BEM2D_OK;
}

extern "C" int bem2d_l2_norm_cond(bem2d_ctx* c, int64_t n, const void* A, int64_t lda,
                                  double* norm, double* cond)
{
    if (int rc = check_args(c, n, A, lda, "bem2d_l2_norm_cond")) return rc;
    HIPCHK(hipSetDevice(c->device));
    // the singular values go to the small scratch slot's tail: [info | E (n) | S (n)]
    char* small = (char*)bem2d_internal_scratch(c, SCRATCH_SPECTRAL, 2 * sizeof(double) * n + 512);
    if (!small) return FAIL(BEM2D_ERR_NOMEM, "bem2d_l2_norm_cond: scratch allocation failed");
    double* S = (double*)(small + 256 + ((sizeof(double) * n + 255) & ~(size_t)255));
    if (int rc = bem2d_singular_values(c, n, A, lda, S)) return rc;
    double s0 = 0.0, s1 = 0.0;
    HIPCHK(hipMemcpyAsync(&s0, S, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&s1, S + (n - 1), sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (norm) *norm = s0;
    if (cond) *cond = s0 / s1;
    return BEM2D_OK;
}

extern "C" int bem2d_hermitian_eigenvalues(bem2d_ctx* c, int64_t n, const void* A, int64_t lda,
                                           double* w)
{
    if (int rc = check_args(c, n, A, lda, "bem2d_hermitian_eigenvalues")) return rc;
    if (!w) return FAIL(BEM2D_ERR_ARG, "bem2d_hermitian_eigenvalues: null output");
    HIPCHK(hipSetDevice(c->device));
    double2* W = (double2*)bem2d_internal_scratch(c, SCRATCH_FACTOR, sizeof(double2) * n * n);
    char* small = (char*)bem2d_internal_scratch(c, SCRATCH_SPECTRAL, sizeof(double) * n + 512);
    if (!W || !small) return FAIL(BEM2D_ERR_NOMEM, "bem2d_hermitian_eigenvalues: scratch allocation failed");
    int* info = (int*)small;
    double* E = (double*)(small + 256);
    spectral_copy_kernel<<<dim3((unsigned)((n + 255) / 256), (unsigned)n), 256, 0, c->stream>>>(
        n, (const double2*)A, lda, W);
    HIPCHK(hipGetLastError());
    const rocblas_status st = rocsolver_zheev(c->blas, rocblas_evect_none, rocblas_fill_upper,
                                              (rocblas_int)n, (rocblas_double_complex*)W,
                                              (rocblas_int)n, w, E, info);
    if (st != rocblas_status_success) return FAIL(BEM2D_ERR_DEVICE, "rocsolver_zheev status %d", (int)st);
    int h_info = 0;
    if (int rc = read_info(c, info, &h_info)) return rc;
    if (h_info != 0)
        return FAIL(BEM2D_ERR_SINGULAR, "bem2d_hermitian_eigenvalues: no convergence (%d)", h_info);
    return BEM2D_OK;
}

extern "C" int bem2d_numerical_range(bem2d_ctx* c, int64_t n, const void* A, int64_t lda,
                                     int nsteps, double* points)
{
    if (int rc = check_args(c, n, A, lda, "bem2d_numerical_range")) return rc;
    if (nsteps < 1 || !points) return FAIL(BEM2D_ERR_ARG, "bem2d_numerical_range: bad argument");
    if (n > 65535) return FAIL(BEM2D_ERR_UNSUPPORTED, "bem2d_numerical_range: n > 65535");
    HIPCHK(hipSetDevice(c->device));
    const double pi = 3.14159265358979323846;
    const double step = pi / nsteps;
    // angles in batches of one batched Hermitian eigenvalue problem each (at most 2 GiB of matrices)
    const int64_t per = std::max<int64_t>(1, std::min<int64_t>(32, (int64_t)(2ll << 30) / (16 * n * n)));
    const int batch = (int)std::min<int64_t>(per, nsteps);
    double2* H = (double2*)bem2d_internal_scratch(c, SCRATCH_FACTOR, sizeof(double2) * n * n * batch);
    const size_t dbytes = sizeof(double) * n * batch;
    const size_t off_e = 4096, off_d = off_e + ((dbytes + 255) & ~(size_t)255),
                 off_i = off_d + ((dbytes + 255) & ~(size_t)255);
    char* small = (char*)bem2d_internal_scratch(c, SCRATCH_SPECTRAL, off_i + sizeof(int) * batch + 256);
    if (!H || !small) return FAIL(BEM2D_ERR_NOMEM, "bem2d_numerical_range: scratch allocation failed");
    double* cs = (double*)small; // 2 batch cos / sin (batch <= 32: 512 bytes)
    double* E = (double*)(small + off_e);
    double* D = (double*)(small + off_d);
    int* info = (int*)(small + off_i);
    std::vector<double> lmin(nsteps), lmax(nsteps), hcs(2 * batch), hD((size_t)n * batch);
    std::vector<int> hinfo(batch);
    for (int i0 = 0; i0 < nsteps; i0 += batch) {
        const int cnt = std::min(batch, nsteps - i0);
        for (int b = 0; b < cnt; ++b) {
            const double th = (i0 + b) * step;
            hcs[2 * b] = std::cos(th);
            hcs[2 * b + 1] = std::sin(th);
        }
        HIPCHK(hipMemcpyAsync(cs, hcs.data(), sizeof(double) * 2 * cnt, hipMemcpyHostToDevice, c->stream));
        numrange_hermitian_kernel<<<dim3((unsigned)((n + 255) / 256), (unsigned)n, (unsigned)cnt), 256,
                                    0, c->stream>>>(n, (const double2*)A, lda, cs, H);
        HIPCHK(hipGetLastError());
        const rocblas_status st = rocsolver_zheev_strided_batched(
            c->blas, rocblas_evect_none, rocblas_fill_upper, (rocblas_int)n,
            (rocblas_double_complex*)H, (rocblas_int)n, (rocblas_stride)(n * n), D, (rocblas_stride)n,
            E, (rocblas_stride)n, info, cnt);
        if (st != rocblas_status_success)
            return FAIL(BEM2D_ERR_DEVICE, "rocsolver_zheev_strided_batched status %d", (int)st);
        HIPCHK(hipMemcpyAsync(hD.data(), D, sizeof(double) * n * cnt, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(hinfo.data(), info, sizeof(int) * cnt, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int b = 0; b < cnt; ++b) {
            if (hinfo[b] != 0)
                return FAIL(BEM2D_ERR_SINGULAR, "bem2d_numerical_range: no convergence (%d)", hinfo[b]);
            lmin[i0 + b] = hD[(size_t)n * b];         // ascending
            lmax[i0 + b] = hD[(size_t)n * b + n - 1];
        }
    }
    // Boundary points: consecutive support lines Re(e^{i th} z) = lambda intersected, in the
    // reference's order (lib/bem2d_mathroutines.cpp:604-632): the minima at th_0 .. th_{m-1}, then
    // the maxima at the same angles, closed by the first point.
    const int m = nsteps;
    std::vector<double> steps(2 * m + 1), eigs(2 * m + 1);
    for (int i = 0; i < m; ++i) {
        steps[i] = steps[m + i] = i * step;
        eigs[i] = lmin[i];
        eigs[m + i] = lmax[i];
    }
    steps[2 * m] = 0.0;
    eigs[2 * m] = eigs[0];
    std::vector<std::complex<double>> nr(2 * m + 1);
    const std::complex<double> im(0.0, 1.0);
    for (int i = 0; i < 2 * m; ++i) {
        const std::complex<double> e1 = std::exp(-im * steps[i]);
        const double c1 = std::cos(steps[i]), c2 = std::cos(steps[i + 1]);
        const double s1 = std::sin(steps[i]), s2 = std::sin(steps[i + 1]);
        const double lam1 = eigs[i], lam2 = eigs[i + 1];
        const double t1 = (lam2 - lam1 * (c1 * c2 + s1 * s2)) / (s1 * c2 - s2 * c1);
        nr[i] = e1 * (lam1 + im * t1);
    }
    nr[2 * m] = nr[0];
    for (int i = 0; i <= 2 * m; ++i) {
        points[2 * i] = nr[i].real();
        points[2 * i + 1] = nr[i].imag();
    }
    return BEM2D_OK;
}
