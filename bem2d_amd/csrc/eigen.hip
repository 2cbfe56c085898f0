// Eigenvalues of a general complex matrix (Eigenvalues, lib/bem2d_mathroutines.cpp:438-500: LAPACK
// zgeev there, ScaLAPACK pzgehrd + pzlahqr in its MPI build) on the GPU, in the same two stages:
//
// 1. Hessenberg reduction Q^H A Q = H by Householder reflectors I - tau v v^H, one column at a time:
//    a one-workgroup kernel forms the reflector of column k (norm, beta, tau, v) and writes the
//    reduced column; rocBLAS zgemv / zgerc then apply it from the left to the trailing block and
//    from the right to the trailing columns.  The scalars stay in device memory (rocBLAS device
//    pointer mode), so the n - 2 steps are queued without a host round trip.
// 2. Shifted QR iteration on H for the eigenvalues only (the active diagonal block is transformed,
//    as zlahqr does with wantt = false).  The bottom unreduced block [l, i] is swept with m shifts
//    at once: m single-shift implicit QR steps whose bulges are chased down the diagonal together,
//    3 rows apart (pipelined bulge chasing: a later step starts at the top once the earlier one has
//    moved on; the left rotations of one step and the right rotations of another commute, so the
//    result is that of the m steps one after another).  Each chase step is two launches: the
//    rotations of all bulges in flight, then every entry they touch updated by exactly one thread
//    (2-entry row or column pieces, and the 2 x 2 blocks where a bulge's rows meet a bulge's
//    columns).  The shifts are the eigenvalues of the trailing m x m block, and blocks of at most
//    kLdsMax rows are finished, by one workgroup holding the block in LDS and running the classic
//    single-shift QR (Wilkinson shift, exceptional shifts after 10 and 20 stalled iterations).
//    Deflation (between sweeps and inside the LDS solver) zeroes a subdiagonal entry h that passes
//    |h| <= ulp (|h_kk| + |h_k+1,k+1|) and the Ahues-Tisseur test for relative accuracy.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/bem2d_hip.h"
#include "internal.h"

namespace {

constexpr int kLdsMax = 64;        // largest block the LDS solver holds (64 x 64 complex = 64 KiB)
constexpr int kLdsThreads = 256;
constexpr double kUlp = 1.1102230246251565e-16;   // 2^-53
constexpr double kSafeMin = 2.2250738585072014e-308;

struct Cx {
    double re, im;
};
__device__ __forceinline__ Cx cx(double2 v) { return {v.x, v.y}; }
__device__ __forceinline__ double2 d2(Cx v) { return make_double2(v.re, v.im); }
__device__ __forceinline__ Cx operator+(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cx operator-(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cx operator*(Cx a, Cx b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ Cx operator*(double s, Cx a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ Cx conj(Cx a) { return {a.re, -a.im}; }
__device__ __forceinline__ double cabs1(Cx a) { return fabs(a.re) + fabs(a.im); }
__device__ __forceinline__ double cabs(Cx a) { return hypot(a.re, a.im); }
__device__ __forceinline__ Cx cdiv(Cx a, Cx b)
{
    // Smith's division
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.im + b.re * r;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__device__ __forceinline__ Cx csqrt(Cx z)
{
    const double m = cabs(z);
    if (m == 0.0) return {0.0, 0.0};
    if (z.re >= 0) {
        const double t = sqrt(0.5 * (m + z.re));
        return {t, z.im / (2 * t)};
    }
    const double t = sqrt(0.5 * (m - z.re));
    return {fabs(z.im) / (2 * t), copysign(t, z.im)};
}

// Unitary G = [[c, s], [-conj(s), c]] (c real) with G [a; b] = [r; 0].
struct Rot {
    double c;
    Cx s;
};
__device__ __forceinline__ Rot make_rot(Cx a, Cx b, Cx& r)
{
    const double ab = cabs(b);
    if (ab == 0.0) {
        r = a;
        return {1.0, {0.0, 0.0}};
    }
    const double aa = cabs(a);
    if (aa == 0.0) {
        r = {ab, 0.0};
        return {0.0, {b.re / ab, -b.im / ab}};
    }
    const double rho = hypot(aa, ab);
    const Cx u = {a.re / aa, a.im / aa}; // a / |a|
    r = rho * u;
    return {aa / rho, (1.0 / rho) * (u * conj(b))};
}
// rows: [x; y] <- G [x; y]
__device__ __forceinline__ void rot_rows(const Rot& g, Cx& x, Cx& y)
{
    const Cx nx = g.c * x + g.s * y;
    const Cx ny = g.c * y - conj(g.s) * x;
    x = nx;
    y = ny;
}
// columns: [x y] <- [x y] G^H, G^H = [[c, -s], [conj(s), c]]
__device__ __forceinline__ void rot_cols(const Rot& g, Cx& x, Cx& y)
{
    const Cx nx = g.c * x + conj(g.s) * y;
    const Cx ny = g.c * y - g.s * x;
    x = nx;
    y = ny;
}

// Deflation test of the subdiagonal entry s = h(k+1, k) between the diagonal entries a = h(k, k),
// d = h(k+1, k+1) and the superdiagonal entry u = h(k, k+1).
__device__ __forceinline__ bool negligible(Cx s, Cx a, Cx d, Cx u)
{
    const double hs = cabs1(s);
    if (hs <= kSafeMin) return true;
    const double tst = cabs1(a) + cabs1(d);
    if (!(hs <= kUlp * tst)) return false;
    // Ahues & Tisseur: small relative to the 2 x 2 block's own scale
    const double ab = fmax(hs, cabs1(u)), ba = fmin(hs, cabs1(u));
    const double aa = fmax(cabs1(d), cabs1(a - d)), bb = fmin(cabs1(d), cabs1(a - d));
    const double sc = aa + ab;
    return ba * (ab / sc) <= fmax(kSafeMin, kUlp * (bb * (aa / sc)));
}

// Eigenvalue of [[a, b], [c, d]] nearest d: d - bc / (x + y), x = (a - d)/2, y = sqrt(x^2 + bc)
// with the sign that keeps x + y away from cancellation.
__device__ __forceinline__ Cx wilkinson(Cx a, Cx b, Cx c, Cx d)
{
    const Cx bc = b * c;
    if (cabs1(bc) == 0.0) return d;
    const Cx x = 0.5 * (a - d);
    Cx y = csqrt(x * x + bc);
    if (x.re * y.re + x.im * y.im < 0) y = {-y.re, -y.im};
    const Cx den = x + y;
    if (cabs1(den) == 0.0) return d;
    return d - cdiv(bc, den);
}

// ---- the LDS solver ------------------------------------------------------------------------------
// Single-shift QR on the nb x nb Hessenberg matrix S (column-major, ld kLdsMax) held in LDS by one
// workgroup, eigenvalues only.  On return S is upper triangular (its diagonal holds the eigenvalues)
// and *ok says whether every eigenvalue converged within 30 nb iterations.
__device__ void hqr_lds(double2* S, int nb, int* ok)
{
    __shared__ int s_ll, s_ii, s_iter, s_its, s_done;
    __shared__ Rot s_rot;
    __shared__ double2 s_shift;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    auto H = [&](int r, int c) -> double2& { return S[r + kLdsMax * c]; };
    if (tid == 0) {
        s_ii = nb - 1;
        s_iter = 0;
        s_its = 0;
        s_done = 0;
        *ok = 1;
    }
    __syncthreads();
    while (true) {
        if (tid == 0) {
            int ii = s_ii;
            // deflate at the bottom and find the active block [ll, ii]
            while (true) {
                if (ii < 0) break;
                int ll = ii;
                while (ll > 0) {
                    Cx s = cx(H(ll, ll - 1));
                    if (negligible(s, cx(H(ll - 1, ll - 1)), cx(H(ll, ll)), cx(H(ll - 1, ll)))) {
                        H(ll, ll - 1) = make_double2(0.0, 0.0);
                        break;
                    }
                    --ll;
                }
                if (ll == ii) { // 1 x 1 block
                    --ii;
                    s_its = 0;
                    continue;
                }
                if (ll == ii - 1) { // 2 x 2 block: its eigenvalues directly, made triangular
                    const Cx a = cx(H(ll, ll)), b = cx(H(ll, ii)), c = cx(H(ii, ll)), d = cx(H(ii, ii));
                    const Cx l2 = wilkinson(a, b, c, d);
                    const Cx l1 = (a + d) - l2;
                    H(ll, ll) = d2(l1);
                    H(ii, ii) = d2(l2);
                    H(ii, ll) = make_double2(0.0, 0.0);
                    ii -= 2;
                    s_its = 0;
                    continue;
                }
                s_ll = ll;
                break;
            }
            s_ii = ii;
            if (ii < 0) {
                s_done = 1;
            } else {
                ++s_its;
                if (++s_iter > 30 * nb) {
                    *ok = 0;
                    s_done = 1;
                }
                Cx sh;
                if (s_its % 10 == 0) { // exceptional shift
                    sh = cx(H(ii, ii)) + Cx{0.75 * cabs1(cx(H(ii, ii - 1))), 0.0};
                } else {
                    sh = wilkinson(cx(H(ii - 1, ii - 1)), cx(H(ii - 1, ii)), cx(H(ii, ii - 1)),
                                   cx(H(ii, ii)));
                }
                s_shift = d2(sh);
            }
        }
        __syncthreads();
        if (s_done) break;
        const int ll = s_ll, ii = s_ii;
        const Cx sh = cx(s_shift);
        for (int k = ll; k < ii; ++k) {
            if (tid == 0) {
                Cx r;
                if (k == ll) {
                    s_rot = make_rot(cx(H(ll, ll)) - sh, cx(H(ll + 1, ll)), r);
                } else {
                    s_rot = make_rot(cx(H(k, k - 1)), cx(H(k + 1, k - 1)), r);
                    H(k, k - 1) = d2(r);
                    H(k + 1, k - 1) = make_double2(0.0, 0.0);
                }
            }
            __syncthreads();
            const Rot g = s_rot;
            for (int c = k + tid; c <= ii; c += nthr) { // rows k, k+1
                Cx x = cx(H(k, c)), y = cx(H(k + 1, c));
                rot_rows(g, x, y);
                H(k, c) = d2(x);
                H(k + 1, c) = d2(y);
            }
            __syncthreads();
            const int rmax = min(k + 2, ii);
            for (int r = ll + tid; r <= rmax; r += nthr) { // columns k, k+1
                Cx x = cx(H(r, k)), y = cx(H(r, k + 1));
                rot_cols(g, x, y);
                H(r, k) = d2(x);
                H(r, k + 1) = d2(y);
            }
            __syncthreads();
        }
    }
    __syncthreads();
}

// Finishes the block [l, l + nb) of H: S = the block, QR in LDS, then the block's diagonal and
// zero subdiagonal written back (the rest of the block is not needed for the eigenvalues).
__global__ __launch_bounds__(kLdsThreads) void hqr_block_kernel(double2* __restrict__ H, int64_t ldh,
                                                                int64_t l, int nb,
                                                                int* __restrict__ status)
{
    __shared__ double2 S[kLdsMax * kLdsMax];
    __shared__ int ok;
    for (int u = threadIdx.x; u < nb * nb; u += blockDim.x) {
        const int r = u % nb, c = u / nb;
        S[r + kLdsMax * c] = H[(l + r) + ldh * (l + c)];
    }
    __syncthreads();
    hqr_lds(S, nb, &ok);
    for (int r = threadIdx.x; r < nb; r += blockDim.x) {
        H[(l + r) + ldh * (l + r)] = S[r + kLdsMax * r];
        if (r + 1 < nb) H[(l + r + 1) + ldh * (l + r)] = make_double2(0.0, 0.0);
    }
    if (threadIdx.x == 0 && !ok) *status = 1;
}

// Shifts of a sweep: the eigenvalues of the trailing m x m block of [.., i] (H untouched).
__global__ __launch_bounds__(kLdsThreads) void shifts_kernel(const double2* __restrict__ H,
                                                             int64_t ldh, int64_t i, int m,
                                                             double2* __restrict__ shifts,
                                                             int* __restrict__ status)
{
    __shared__ double2 S[kLdsMax * kLdsMax];
    __shared__ int ok;
    const int64_t l = i - m + 1;
    for (int u = threadIdx.x; u < m * m; u += blockDim.x) {
        const int r = u % m, c = u / m;
        S[r + kLdsMax * c] = (r <= c + 1) ? H[(l + r) + ldh * (l + c)] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    hqr_lds(S, m, &ok);
    for (int r = threadIdx.x; r < m; r += blockDim.x) shifts[r] = S[r + kLdsMax * r];
    if (threadIdx.x == 0) *status = ok ? 0 : 1;
}

// Exceptional shifts around h(i, i) when sweeps stall: h_ii + 0.75 |h(i, i-1)| e^{2 pi i b / m}.
__global__ void exceptional_shifts_kernel(const double2* __restrict__ H, int64_t ldh, int64_t i,
                                          int m, double2* __restrict__ shifts)
{
    const int b = threadIdx.x;
    if (b >= m) return;
    const Cx h = cx(H[i + ldh * i]);
    const double r = 0.75 * cabs1(cx(H[i + ldh * (i - 1)])) + 1e-3 * cabs1(h);
    double sn, cs;
    sincospi(2.0 * b / m + 0.25, &sn, &cs);
    shifts[b] = make_double2(h.re + r * cs, h.im + r * sn);
}

// ---- pipelined multishift sweep over [l, i] --------------------------------------------------------
// Bulge b (shift b) sits at row k_b = l + t - 3 b at chase step t, in flight while l <= k_b <= i - 1.
__device__ __forceinline__ int64_t bulge_pos(int64_t l, int64_t t, int b) { return l + t - 3 * (int64_t)b; }

// Rotations of the bulges in flight; a bulge past its first row also gets its chased entries
// h(k, k-1) = r, h(k+1, k-1) = 0 (nobody else touches them in this step).
__global__ void sweep_rot_kernel(double2* __restrict__ H, int64_t ldh, int64_t l, int64_t i,
                                 int64_t t, int m, const double2* __restrict__ shifts,
                                 double* __restrict__ rot)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m) return;
    const int64_t k = bulge_pos(l, t, b);
    if (k < l || k > i - 1) return;
    Cx r;
    Rot g;
    if (k == l) {
        g = make_rot(cx(H[l + ldh * l]) - cx(shifts[b]), cx(H[(l + 1) + ldh * l]), r);
    } else {
        g = make_rot(cx(H[k + ldh * (k - 1)]), cx(H[(k + 1) + ldh * (k - 1)]), r);
        H[k + ldh * (k - 1)] = d2(r);
        H[(k + 1) + ldh * (k - 1)] = make_double2(0.0, 0.0);
    }
    rot[3 * b] = g.c;
    rot[3 * b + 1] = g.s.re;
    rot[3 * b + 2] = g.s.im;
}

__device__ __forceinline__ Rot load_rot(const double* rot, int b)
{
    return {rot[3 * b], {rot[3 * b + 1], rot[3 * b + 2]}};
}

// The bulge whose row / column pair starts at index x (x = k_b), or -1.
__device__ __forceinline__ int bulge_at(int64_t l, int64_t i, int64_t t, int m, int64_t x)
{
    const int64_t d = l + t - x; // = 3 b
    if (d < 0 || d % 3 != 0) return -1;
    const int64_t b = d / 3;
    if (b >= m) return -1;
    if (x < l || x > i - 1) return -1;
    return (int)b;
}

// Updates of chase step t: workgroup row blockIdx.y = bulge b, thread x = offset in [l, i].
//  - row piece (rows k_b, k_b + 1; column c = l + x >= k_b): the 2 x 2 block with bulge b' when
//    c = k_b' (k_b' >= k_b: rotation of b from the left, then of b' from the right); skipped when
//    c = k_b' + 1; otherwise the 2-entry column piece;
//  - column piece (columns k_b, k_b + 1; row r = l + x <= min(k_b + 2, i)): skipped when r belongs
//    to the rows of a bulge at or behind b (done as a block); otherwise the 2-entry row piece.
__global__ __launch_bounds__(256) void sweep_update_kernel(double2* __restrict__ H, int64_t ldh,
                                                           int64_t l, int64_t i, int64_t t, int m,
                                                           const double* __restrict__ rot)
{
    const int b = blockIdx.y;
    const int64_t k = bulge_pos(l, t, b);
    if (k < l || k > i - 1) return;
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x > i - l) return;
    const Rot g = load_rot(rot, b);
    const int64_t c = l + x;
    if (c >= k) {
        const int bc = bulge_at(l, i, t, m, c);
        const int bprev = bulge_at(l, i, t, m, c - 1);
        if (bc >= 0) { // 2 x 2 block rows (k, k+1) x columns (c, c+1)
            const Rot gc = load_rot(rot, bc);
            Cx a = cx(H[k + ldh * c]), bb = cx(H[k + ldh * (c + 1)]);
            Cx cc = cx(H[(k + 1) + ldh * c]), dd = cx(H[(k + 1) + ldh * (c + 1)]);
            rot_rows(g, a, cc);
            rot_rows(g, bb, dd);
            rot_cols(gc, a, bb);
            rot_cols(gc, cc, dd);
            H[k + ldh * c] = d2(a);
            H[k + ldh * (c + 1)] = d2(bb);
            H[(k + 1) + ldh * c] = d2(cc);
            H[(k + 1) + ldh * (c + 1)] = d2(dd);
        } else if (bprev < 0) {
            Cx a = cx(H[k + ldh * c]), cc = cx(H[(k + 1) + ldh * c]);
            rot_rows(g, a, cc);
            H[k + ldh * c] = d2(a);
            H[(k + 1) + ldh * c] = d2(cc);
        }
    }
    const int64_t r = l + x;
    if (r <= std::min(k + 2, i)) {
        // rows of a bulge at or behind b (k_b'' <= k) are blocks of that bulge's row pieces
        const int b0 = bulge_at(l, i, t, m, r), b1 = bulge_at(l, i, t, m, r - 1);
        const bool owned = (b0 >= 0 && b0 >= b) || (b1 >= 0 && b1 >= b);
        if (!owned) {
            Cx a = cx(H[r + ldh * k]), bb = cx(H[r + ldh * (k + 1)]);
            rot_cols(g, a, bb);
            H[r + ldh * k] = d2(a);
            H[r + ldh * (k + 1)] = d2(bb);
        }
    }
}

// Deflation scan of [l, i]: zeroes negligible subdiagonal entries; flag[k] = 1 where h(k+1, k) = 0.
__global__ void deflate_kernel(double2* __restrict__ H, int64_t ldh, int64_t l, int64_t i,
                               int* __restrict__ flag)
{
    const int64_t k = l + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= i) return;
    const Cx s = cx(H[(k + 1) + ldh * k]);
    bool z = s.re == 0.0 && s.im == 0.0;
    if (!z && negligible(s, cx(H[k + ldh * k]), cx(H[(k + 1) + ldh * (k + 1)]),
                         cx(H[k + ldh * (k + 1)]))) {
        H[(k + 1) + ldh * k] = make_double2(0.0, 0.0);
        z = true;
    }
    flag[k] = z ? 1 : 0;
}

// ---- Hessenberg reduction ------------------------------------------------------------------------
// Reflector of column k: x = h(k+1 : n, k), alpha = x_0; beta = -sign(Re alpha) ||x||,
// tau = (beta - alpha) / beta, v = [1; x_1: / (alpha - beta)].  Writes v, the scalars
// -conj(tau) (left update) and -tau (right update), and the reduced column (beta, 0, ..., 0).
__global__ __launch_bounds__(1024) void reflector_kernel(double2* __restrict__ H, int64_t ldh,
                                                         int64_t n, int64_t k,
                                                         double2* __restrict__ v,
                                                         double2* __restrict__ scal)
{
    __shared__ double part[16];
    __shared__ double s_norm2;
    const int64_t m = n - k - 1;
    double2* x = H + (k + 1) + ldh * k;
    double acc = 0.0;
    for (int64_t r = 1 + threadIdx.x; r < m; r += blockDim.x) {
        const double2 e = x[r];
        acc += e.x * e.x + e.y * e.y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += part[w];
        s_norm2 = s;
    }
    __syncthreads();
    const double xn2 = s_norm2;
    const Cx alpha = cx(x[0]);
    if (xn2 == 0.0 && alpha.im == 0.0) { // already reduced: H = I
        for (int64_t r = threadIdx.x; r < m; r += blockDim.x)
            v[r] = make_double2(r == 0 ? 1.0 : 0.0, 0.0);
        if (threadIdx.x == 0) {
            scal[2] = make_double2(0.0, 0.0);
            scal[3] = make_double2(0.0, 0.0);
        }
        return;
    }
    const double beta = -copysign(sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2), alpha.re);
    const Cx tau = {(beta - alpha.re) / beta, -alpha.im / beta};
    const Cx inv = cdiv({1.0, 0.0}, alpha - Cx{beta, 0.0});
    __syncthreads(); // every thread has read alpha before it is overwritten
    for (int64_t r = threadIdx.x; r < m; r += blockDim.x) {
        if (r == 0) {
            v[0] = make_double2(1.0, 0.0);
            x[0] = make_double2(beta, 0.0);
        } else {
            v[r] = d2(cx(x[r]) * inv);
            x[r] = make_double2(0.0, 0.0);
        }
    }
    if (threadIdx.x == 0) {
        scal[2] = make_double2(-tau.re, tau.im);  // -conj(tau)
        scal[3] = make_double2(-tau.re, -tau.im); // -tau
    }
}

__global__ void diag_kernel(const double2* __restrict__ H, int64_t ldh, int64_t n,
                            double2* __restrict__ w)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) w[k] = H[k + ldh * k];
}

__global__ void copy_in_kernel(int64_t n, const double2* __restrict__ src, int64_t lds,
                               double2* __restrict__ dst)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (r < n) dst[r + n * c] = src[r + lds * c];
}

// Number of shifts per sweep for an active block of size s.
int shift_count(int64_t s)
{
    int m = s < 150 ? 8 : s < 600 ? 16 : s < 2000 ? 32 : 64;
    while (m > 2 && 4 * m > s) m /= 2;
    return std::min(m, kLdsMax);
}

} // namespace

extern "C" int bem2d_eigenvalues(bem2d_ctx* c, int64_t n, const void* A, int64_t lda, void* w)
{
    if (!c || !A || !w || n < 1 || lda < n || n > 46340)
        return FAIL(BEM2D_ERR_ARG, "bem2d_eigenvalues: bad argument");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const size_t hbytes = sizeof(double2) * (size_t)n * n;
    double2* H = (double2*)bem2d_internal_scratch(c, SCRATCH_FACTOR, hbytes);
    // small workspace: v, u / w vectors (3 n), 4 scalars, shifts, rotations, flags, status
    const size_t small = sizeof(double2) * (3 * (size_t)n + 4 + kLdsMax) +
                         sizeof(double) * 3 * kLdsMax + sizeof(int) * ((size_t)n + 4) + 1024;
    char* ws = (char*)bem2d_internal_scratch(c, SCRATCH_SPECTRAL, small);
    if (!H || !ws) return FAIL(BEM2D_ERR_NOMEM, "scratch allocation failed");
    double2* vv = (double2*)ws;
    double2* wv = vv + n;
    double2* uv = wv + n;
    double2* scal = uv + n; // [0] = 1, [1] = 0, [2] = -conj(tau), [3] = -tau
    double2* shifts = scal + 4;
    double* rot = (double*)(shifts + kLdsMax);
    int* flag = (int*)(rot + 3 * kLdsMax);
    int* status = flag + n + 2;
    {
        dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
        copy_in_kernel<<<grid, 256, 0, st>>>(n, (const double2*)A, lda, H);
        const double2 consts[2] = {make_double2(1.0, 0.0), make_double2(0.0, 0.0)};
        HIPCHK(hipMemcpyAsync(scal, consts, sizeof(consts), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(status, 0, sizeof(int), st));
        HIPCHK(hipStreamSynchronize(st)); // consts lives on this stack
    }

    // ---- stage 1: Hessenberg reduction
    rocblas_pointer_mode pm;
    rocblas_get_pointer_mode(c->blas, &pm);
    rocblas_set_pointer_mode(c->blas, rocblas_pointer_mode_device);
    int rc = BEM2D_OK;
    const rocblas_double_complex* one = (const rocblas_double_complex*)scal;
    const rocblas_double_complex* zero = (const rocblas_double_complex*)(scal + 1);
    const rocblas_double_complex* aleft = (const rocblas_double_complex*)(scal + 2);
    const rocblas_double_complex* aright = (const rocblas_double_complex*)(scal + 3);
    for (int64_t k = 0; k + 2 < n && rc == BEM2D_OK; ++k) {
        const rocblas_int m = (rocblas_int)(n - k - 1);
        reflector_kernel<<<1, 1024, 0, st>>>(H, n, n, k, vv, scal);
        rocblas_double_complex* A2 = (rocblas_double_complex*)(H + (k + 1) + n * (k + 1));
        rocblas_double_complex* A3 = (rocblas_double_complex*)(H + n * (k + 1));
        rocblas_double_complex* v = (rocblas_double_complex*)vv;
        rocblas_double_complex* wl = (rocblas_double_complex*)wv;
        rocblas_double_complex* ur = (rocblas_double_complex*)uv;
        // left: A2 <- A2 - conj(tau) v (A2^H v)^H
        if (rocblas_zgemv(c->blas, rocblas_operation_conjugate_transpose, m, m, one, A2,
                          (rocblas_int)n, v, 1, zero, wl, 1) != rocblas_status_success ||
            rocblas_zgerc(c->blas, m, m, aleft, v, 1, wl, 1, A2, (rocblas_int)n) !=
                rocblas_status_success ||
            // right: A3 <- A3 - tau (A3 v) v^H
            rocblas_zgemv(c->blas, rocblas_operation_none, (rocblas_int)n, m, one, A3,
                          (rocblas_int)n, v, 1, zero, ur, 1) != rocblas_status_success ||
            rocblas_zgerc(c->blas, (rocblas_int)n, m, aright, ur, 1, v, 1, A3, (rocblas_int)n) !=
                rocblas_status_success)
            rc = FAIL(BEM2D_ERR_DEVICE, "bem2d_eigenvalues: rocBLAS call failed");
    }
    rocblas_set_pointer_mode(c->blas, pm);
    if (rc) return rc;
    if (hipGetLastError() != hipSuccess) return FAIL(BEM2D_ERR_DEVICE, "Hessenberg reduction failed");

    // ---- stage 2: QR iteration on [0, n)
    std::vector<int> hflag((size_t)n + 1);
    int64_t i = n - 1;
    int stalled = 0;
    int64_t last_l = -1, last_i = -1;
    int64_t shifts_used = 0; // at most 30 n shifts in all (LAPACK's iteration budget)
    while (i > 0) {
        deflate_kernel<<<(unsigned)((i + 255) / 256), 256, 0, st>>>(H, n, 0, i, flag);
        HIPCHK(hipMemcpyAsync(hflag.data(), flag, sizeof(int) * (size_t)i, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        while (i > 0 && hflag[i - 1]) --i; // 1 x 1 blocks at the bottom
        if (i <= 0) break;
        int64_t l = i;
        while (l > 0 && !hflag[l - 1]) --l;
        const int64_t s = i - l + 1;
        if (s <= kLdsMax) {
            hqr_block_kernel<<<1, kLdsThreads, 0, st>>>(H, n, l, (int)s, status);
            i = l - 1;
            stalled = 0;
            continue;
        }
        if (l == last_l && i == last_i) {
            ++stalled;
        } else {
            stalled = 0;
        }
        last_l = l;
        last_i = i;
        const int m = shift_count(s);
        shifts_used += m;
        if (stalled > 50 || shifts_used > 30 * n)
            return FAIL(BEM2D_ERR_NOT_CONVERGED, "bem2d_eigenvalues: QR iteration did not converge");
        if (stalled > 0 && stalled % 5 == 0)
            exceptional_shifts_kernel<<<1, 64, 0, st>>>(H, n, i, m, shifts);
        else
            shifts_kernel<<<1, kLdsThreads, 0, st>>>(H, n, i, m, shifts, status + 1);
        const int64_t steps = (i - l) + 3 * (int64_t)(m - 1);
        const dim3 ugrid((unsigned)((s + 255) / 256), (unsigned)m);
        for (int64_t t = 0; t < steps; ++t) {
            sweep_rot_kernel<<<1, 64, 0, st>>>(H, n, l, i, t, m, shifts, rot);
            sweep_update_kernel<<<ugrid, 256, 0, st>>>(H, n, l, i, t, m, rot);
        }
        if (hipGetLastError() != hipSuccess) return FAIL(BEM2D_ERR_DEVICE, "QR sweep launch failed");
    }
    diag_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(H, n, n, (double2*)w);
    int hstatus = 0;
    HIPCHK(hipMemcpyAsync(&hstatus, status, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (hstatus) return FAIL(BEM2D_ERR_NOT_CONVERGED, "bem2d_eigenvalues: a block did not converge");
    return BEM2D_OK;
}
