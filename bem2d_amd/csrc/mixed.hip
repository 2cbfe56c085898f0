// Mixed-precision dense complex solve: FP32 LU on the MFMA cores + FP64 iterative refinement.
//
// Factorisation (bem2d_lu32_factor): A32 = (complex64) A64, then a blocked right-looking LU with
// 64-wide panels and no row interchanges.  Per panel: the 64x64 diagonal block is factorised by one
// workgroup with the tile in registers, which also forms the inverses of its unit-lower and upper
// triangles; the block column below and the block row to the right are multiplied by those
// inverses; the trailing matrix receives the rank-64 update A22 -= L21 U12 as a complex GEMM on
// v_mfma_f32_16x16x4f32 with FP32 accumulation (four real MFMA products per complex product).  Without interchanges the panel has
// no sequential pivot search -- the step that makes partial-pivoting panels latency bound on a GPU.
// A zero or non-finite pivot is flagged and the refinement then reports failure, after which the
// one-shot solver falls back to the FP64 LU.
//
// Refinement (bem2d_lu32_refine): x0 = 0, r = b - A64 x in FP64, d = U^{-1} L^{-1} r in FP32
// (two persistent block-triangular sweeps with the inverted diagonal blocks; blocks hand their
// solution on through release/acquire flags), x += d in FP64, until the backward error
// |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) <= tol for every right-hand side.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/bem2d_hip.h"
#include "internal.h"

namespace {

constexpr int NB = 64;             // panel / block width
constexpr int LDT = NB + 1;        // padded LDS row of a 64x64 complex tile (float2 units)
constexpr int kMaxSolveRhs = 8;    // right-hand sides handled per triangular sweep
constexpr int kResidChunks = 64;   // column chunks of the FP64 residual
constexpr long long kSpinLimit = 1ll << 26;

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cfms(float2 acc, float2 a, float2 b) // acc - a b
{
    acc.x -= a.x * b.x - a.y * b.y;
    acc.y -= a.x * b.y + a.y * b.x;
    return acc;
}
__device__ __forceinline__ float2 crecip(float2 a)
{
    const float d = a.x * a.x + a.y * a.y;
    return make_float2(a.x / d, -a.y / d);
}

// ---- cast ---------------------------------------------------------------------------------------
__global__ void cast_kernel(int64_t n, const double2* __restrict__ A, int64_t lda,
                            float2* __restrict__ F, int64_t ldf)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i < n) {
        const double2 v = A[i + lda * j];
        F[i + ldf * j] = make_float2((float)v.x, (float)v.y);
    }
}

// |A|_inf = max_i sum_j |a_ij| (FP64): partial row sums per column chunk, then max.
__global__ void rowsum_kernel(int64_t n, const double2* __restrict__ A, int64_t lda, int nchunk,
                              double* __restrict__ part)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n) return;
    const int64_t j0 = n * c / nchunk, j1 = n * (c + 1) / nchunk;
    double s = 0.0;
    for (int64_t j = j0; j < j1; ++j) {
        const double2 v = A[i + lda * j];
        s += hypot(v.x, v.y);
    }
    part[c * n + i] = s;
}

__global__ void rowsum_max_kernel(int64_t n, int nchunk, const double* __restrict__ part,
                                  double* __restrict__ out)
{
    __shared__ double red[256];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < nchunk; ++c) s += part[c * n + i];
        m = fmax(m, s);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// ---- diagonal block: LU without interchanges + inverses of both triangles -----------------------
// 256 threads.  Thread t = 64 w + l holds T[a][b] = tile(R(a), C(b)) with R(a) = 16 a + (l & 15) and
// C(b) = 16 b + 4 w + (l >> 4): each wavefront holds a 16 x 4 piece of each of the 4 x 4 sub-blocks
// of 16, so at step j the sub-blocks above or left of row/column j drop out on every SIMD alike and
// the work per step shrinks with the active region.
//
// Step j eliminates column j (pivot row j = U[j][:], multipliers L[r][j] = A[r][j] / U[j][j]) and,
// in the same step, advances two forward sweeps that need only row j of U and column j of L:
//   X = L^{-1}:           X = I;  rows r > j:  X[r][:] -= L[r][j] X[j][:];
//   Y = (U^T)^{-1}:       Y = I;  Y[j][:] /= U[j][j];  rows r > j:  Y[r][:] -= U[j][r] Y[j][:];
// so U^{-1} = Y^T comes out of the same 64 steps.  Row and column j+1 of the updated matrix and rows
// j+1 of X and Y (Y unscaled) are published through ping-pong LDS buffers: one barrier per step.
// Linv receives X and UinvT receives Y = (U^{-1})^T (column-major, ld 64; identity-padded when
// kb < 64).
__device__ __forceinline__ float2 fast_crecip(float2 a)
{
    const float d = __builtin_amdgcn_rcpf(a.x * a.x + a.y * a.y);
    return make_float2(a.x * d, -a.y * d);
}

struct DiagLds {
    float2 row[2][NB], col[2][NB];   // pivot row / column of the current matrix
    float2 xrow[2][NB], yrow[2][NB]; // row j of X and of Y (before its scaling by 1/U[j][j])
};

__device__ __forceinline__ int diag_R(int a, int l) { return 16 * a + (l & 15); }
__device__ __forceinline__ int diag_C(int b, int w, int l) { return 16 * b + 4 * w + (l >> 4); }

template <int V> using ic = std::integral_constant<int, V>;

// Lanes holding row j: (l & 15) == (j & 15), sub-block row j >> 4.  Lanes holding column j:
// wavefront (j >> 2) & 3, (l >> 4) == (j & 3), sub-block column j >> 4.
__device__ __forceinline__ bool holds_row(int j, int l) { return (l & 15) == (j & 15); }
__device__ __forceinline__ bool holds_col(int j, int w, int l)
{
    return w == ((j >> 2) & 3) && (l >> 4) == (j & 3);
}

// The 64 steps run as four groups of 16, one per sub-block index js = j >> 4: within a group every
// sub-block index is a compile-time constant, so the register tiles are never indexed at run time
// (that would place them in scratch memory) and the inactive sub-blocks are dropped at compile time.
__device__ __forceinline__ void diag_factor(float2* __restrict__ A, int64_t ld, int64_t k0, int kb,
                                            float2* __restrict__ Linv,
                                            float2* __restrict__ UinvT, int* __restrict__ bad,
                                            DiagLds& S)
{
    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int l = t & 63;
    const float2 zero = make_float2(0.f, 0.f);
    float2 T[4][4], X[4][4], Y[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int r = diag_R(a, l), c = diag_C(b, w, l);
            const float2 e = make_float2(r == c ? 1.f : 0.f, 0.f);
            T[a][b] = (r < kb && c < kb) ? A[(k0 + r) + ld * (k0 + c)] : e;
            X[a][b] = e;
            Y[a][b] = e;
        }
    // row / column j (sub-block index JS = j >> 4) into buffer buf
    auto publish = [&](auto JS_, int j, int buf) __attribute__((always_inline)) {
        constexpr int js = decltype(JS_)::value;
        if (holds_row(j, l)) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int c = diag_C(b, w, l);
                S.row[buf][c] = T[js][b];
                if (b <= js) {
                    S.xrow[buf][c] = X[js][b];
                    S.yrow[buf][c] = Y[js][b];
                }
            }
        }
        if (holds_col(j, w, l)) {
#pragma unroll
            for (int a = 0; a < 4; ++a) S.col[buf][diag_R(a, l)] = T[a][js];
        }
    };
    auto group = [&](auto JS_) __attribute__((always_inline)) {
        constexpr int js = decltype(JS_)::value;
#pragma unroll 1
        for (int jj = 0; jj < 16; ++jj) {
            const int j = 16 * js + jj;
            __syncthreads();
            const int buf = j & 1;
            const float2 p = S.row[buf][j];
            if (t == 0 && j < kb &&
                !(p.x * p.x + p.y * p.y > 0.f && isfinite(p.x) && isfinite(p.y)))
                *bad = 1;
            const float2 ip = fast_crecip(p);
            // m[a] = L[R(a)][j], g[a] = U[j][R(a)] for rows below j; u[b] = U[j][C(b)] right of j
            float2 m[4], g[4], u[4], xr[4], yr[4];
#pragma unroll
            for (int a = js; a < 4; ++a) {
                const int r = diag_R(a, l);
                const bool below = a > js || r > j;
                const float2 cv = S.col[buf][r];
                const float2 rv = S.row[buf][r];
                m[a] = below ? cmul(cv, ip) : zero;
                g[a] = below ? rv : zero;
            }
#pragma unroll
            for (int b = js; b < 4; ++b) {
                const int c = diag_C(b, w, l);
                const float2 rv = S.row[buf][c];
                u[b] = (b > js || c > j) ? rv : zero;
            }
#pragma unroll
            for (int b = 0; b <= js; ++b) {
                const int c = diag_C(b, w, l);
                xr[b] = S.xrow[buf][c];
                yr[b] = cmul(S.yrow[buf][c], ip);
            }
            if (holds_row(j, l)) {
#pragma unroll
                for (int b = 0; b <= js; ++b) Y[js][b] = yr[b];
            }
#pragma unroll
            for (int a = js; a < 4; ++a) {
#pragma unroll
                for (int b = js; b < 4; ++b) T[a][b] = cfms(T[a][b], m[a], u[b]);
#pragma unroll
                for (int b = 0; b <= js; ++b) {
                    X[a][b] = cfms(X[a][b], m[a], xr[b]);
                    Y[a][b] = cfms(Y[a][b], g[a], yr[b]);
                }
            }
            // column j takes the multipliers (u is zero there, so the update left it as it was)
            if (holds_col(j, w, l)) {
#pragma unroll
                for (int a = js; a < 4; ++a)
                    if (a > js || diag_R(a, l) > j) T[a][js] = m[a];
            }
            if (jj < 15)
                publish(ic<js>(), j + 1, buf ^ 1);
            else if (js < 3)
                publish(ic<(js < 3 ? js + 1 : 3)>(), j + 1, buf ^ 1);
        }
    };
    publish(ic<0>(), 0, 0);
    group(ic<0>());
    group(ic<1>());
    group(ic<2>());
    group(ic<3>());
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int r = diag_R(a, l), c = diag_C(b, w, l);
            if (r < kb && c < kb) A[(k0 + r) + ld * (k0 + c)] = T[a][b];
            Linv[r + NB * c] = X[a][b];
            UinvT[r + NB * c] = Y[a][b];
        }
}

__global__ __launch_bounds__(256) void diag_lu_kernel(float2* __restrict__ A, int64_t ld,
                                                      int64_t k0, int kb,
                                                      float2* __restrict__ Linv,
                                                      float2* __restrict__ UinvT,
                                                      int* __restrict__ bad)
{
    __shared__ DiagLds S;
    diag_factor(A, ld, k0, kb, Linv, UinvT, bad, S);
}

// ---- 64 x 64 x 64 complex products on MFMA -----------------------------------------------------
// acc += Aop Bop with Aop in LDS as [row][k] and Bop as [col][k] (pitch KP, real and imaginary
// planes apart).  Wavefront w computes the 32 x 32 quarter at rows 32 (w & 1), cols 32 (w >> 1)
// as 2 x 2 tiles of 16 x 16; each complex product is four real v_mfma_f32_16x16x4f32.
// v_mfma_f32_16x16x4f32 lane layout: A operand lane l holds A[l % 16][l / 16], B operand
// B[l / 16][l % 16]; accumulator register r of lane l is C[4 (l / 16) + r][l % 16].
constexpr int KP = NB + 4; // padded k extent of the LDS operand tiles
constexpr int kTileLdsFloats = 4 * NB * KP;

struct TileAcc {
    v4f re[2][2], im[2][2];
};

__device__ __forceinline__ void cgemm_64(const float* __restrict__ Are,
                                         const float* __restrict__ Aim,
                                         const float* __restrict__ Bre,
                                         const float* __restrict__ Bim, TileAcc& c)
{
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const int li = lane % 16, lk = lane / 16;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            c.re[p][q] = v4f{0.f, 0.f, 0.f, 0.f};
            c.im[p][q] = v4f{0.f, 0.f, 0.f, 0.f};
        }
    for (int ks = 0; ks < NB; ks += 4) {
        float ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            ar[p] = Are[(wr + 16 * p + li) * KP + ks + lk];
            ai[p] = Aim[(wr + 16 * p + li) * KP + ks + lk];
            br[p] = Bre[(wc + 16 * p + li) * KP + ks + lk];
            bi[p] = Bim[(wc + 16 * p + li) * KP + ks + lk];
        }
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                c.re[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[p], br[q], c.re[p][q], 0, 0, 0);
                c.re[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai[p], bi[q], c.re[p][q], 0, 0, 0);
                c.im[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[p], bi[q], c.im[p][q], 0, 0, 0);
                c.im[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[p], br[q], c.im[p][q], 0, 0, 0);
            }
    }
}

// Tile-local coordinates of accumulator element (p, q, r) of this thread.
__device__ __forceinline__ int acc_row(int p, int r)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    return (wave & 1) * 32 + 16 * p + 4 * (lane / 16) + r;
}
__device__ __forceinline__ int acc_col(int q)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    return (wave >> 1) * 32 + 16 * q + lane % 16;
}

// ---- panel: L21 = A21 U11^{-1} and U12 = L11^{-1} A12 on MFMA ---------------------------------
// blockIdx.x < nrow_tiles: a 64-row tile of the block column below the diagonal block;
// otherwise a 64-column tile of the block row to its right.  In place, through LDS.
__global__ __launch_bounds__(256) void panel_trsm_kernel(float2* __restrict__ A, int64_t ld,
                                                         int64_t n, int64_t k0, int kb,
                                                         int nrow_tiles,
                                                         const float2* __restrict__ Linv,
                                                         const float2* __restrict__ UinvT)
{
    __shared__ __align__(16) float lds[kTileLdsFloats];
    float* Are = lds;
    float* Aim = lds + NB * KP;
    float* Bre = lds + 2 * NB * KP;
    float* Bim = lds + 3 * NB * KP;
    const int t = threadIdx.x;
    const int64_t rest0 = k0 + kb;
    const bool rowtile = (int)blockIdx.x < nrow_tiles;
    const int64_t base = rest0 + (int64_t)(rowtile ? blockIdx.x : blockIdx.x - nrow_tiles) * NB;
    const int len = (int)min((int64_t)NB, n - base);
    for (int e = t; e < NB * NB; e += 256) {
        const int x = e % NB, y = e / NB;
        float2 a, b;
        if (rowtile) { // out = S U^{-1}: Aop[x][y] = S[x][y], Bop[x][y] = U^{-1}[y][x] = UinvT[x][y]
            a = (x < len && y < kb) ? A[(base + x) + ld * (k0 + y)] : make_float2(0.f, 0.f);
            b = UinvT[x + NB * y];
            Bre[x * KP + y] = b.x;
            Bim[x * KP + y] = b.y;
        } else {       // out = L^{-1} S: Aop[x][y] = Linv[x][y], Bop[y][x] = S[x][y]
            a = Linv[x + NB * y];
            b = (x < kb && y < len) ? A[(k0 + x) + ld * (base + y)] : make_float2(0.f, 0.f);
            Bre[y * KP + x] = b.x;
            Bim[y * KP + x] = b.y;
        }
        Are[x * KP + y] = a.x;
        Aim[x * KP + y] = a.y;
    }
    __syncthreads();
    TileAcc c;
    cgemm_64(Are, Aim, Bre, Bim, c);
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int x = acc_row(p, r), y = acc_col(q);
                const float2 v = make_float2(c.re[p][q][r], c.im[p][q][r]);
                if (rowtile) {
                    if (x < len && y < kb) A[(base + x) + ld * (k0 + y)] = v;
                } else {
                    if (x < kb && y < len) A[(k0 + x) + ld * (base + y)] = v;
                }
            }
}

// ---- trailing update on MFMA, fused with the next diagonal block ---------------------------------
// C[i][j] -= sum_k L[i][k] U[k][j] for i, j in [rest0, n), k in [k0, k0 + kb); workgroup = one
// 64 x 64 output tile.  Workgroup (0, 0) owns the next diagonal block: after its update it goes on
// to factorise that block (when Linv_next is given), overlapping the rest of the update.
__global__ __launch_bounds__(256) void trailing_update_kernel(float2* __restrict__ A, int64_t ld,
                                                              int64_t n, int64_t k0, int kb,
                                                              float2* __restrict__ Linv_next,
                                                              float2* __restrict__ UinvT_next,
                                                              int* __restrict__ bad)
{
    __shared__ __align__(16) float lds[kTileLdsFloats];
    static_assert(sizeof(DiagLds) <= sizeof(float) * kTileLdsFloats, "LDS reuse");
    float* Are = lds;
    float* Aim = lds + NB * KP;
    float* Bre = lds + 2 * NB * KP;
    float* Bim = lds + 3 * NB * KP;
    const int t = threadIdx.x;
    const int64_t rest0 = k0 + kb;
    const int64_t r0 = rest0 + (int64_t)blockIdx.x * NB;
    const int64_t c0 = rest0 + (int64_t)blockIdx.y * NB;
    for (int e = t; e < NB * NB; e += 256) {
        const int r = e % NB, k = e / NB;
        float2 v = make_float2(0.f, 0.f);
        if (r0 + r < n && k < kb) v = A[(r0 + r) + ld * (k0 + k)];
        Are[r * KP + k] = v.x;
        Aim[r * KP + k] = v.y;
        float2 u = make_float2(0.f, 0.f);
        const int kk = e % NB, c = e / NB;
        if (c0 + c < n && kk < kb) u = A[(k0 + kk) + ld * (c0 + c)];
        Bre[c * KP + kk] = u.x;
        Bim[c * KP + kk] = u.y;
    }
    __syncthreads();
    TileAcc acc;
    cgemm_64(Are, Aim, Bre, Bim, acc);
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t i = r0 + acc_row(p, r);
                const int64_t j = c0 + acc_col(q);
                if (i < n && j < n) {
                    float2 v = A[i + ld * j];
                    v.x -= acc.re[p][q][r];
                    v.y -= acc.im[p][q][r];
                    A[i + ld * j] = v;
                }
            }
    if (Linv_next && blockIdx.x == 0 && blockIdx.y == 0) {
        __syncthreads(); // the tile's stores are visible to the workgroup; the LDS is free
        diag_factor(A, ld, rest0, (int)min((int64_t)NB, n - rest0), Linv_next, UinvT_next, bad,
                    *reinterpret_cast<DiagLds*>(lds));
    }
}

// ---- block-triangular sweeps -------------------------------------------------------------------
// Y (n x nrhs, column-major, FP32) is solved in place.  Lower: unit-lower L from the factors with
// inverted diagonal blocks Linv; block b waits for blocks < b.  Upper: blocks in reverse order, with
// the transposed inverses UinvT.
// Each workgroup owns one 64-row block; the first workgroups dispatched own the blocks without
// dependencies, so every wait is on a block that is already running or done.  A wait gives up
// after kSpinLimit polls and raises *err, so the kernel always terminates.
// Thread (row = t % 64, grp = t / 64) owns 16 consecutive k of each 64-wide column block: the
// matrix block is loaded into registers before the flag wait (it does not depend on it), and the
// partial sums stay in registers until the end, so a dependency costs one wait and one barrier.
template <bool LOWER>
__global__ __launch_bounds__(256) void block_trsv_kernel(const float2* __restrict__ F, int64_t ld,
                                                         int64_t n, int nblk,
                                                         const float2* __restrict__ Dinv,
                                                         float2* __restrict__ Y, int64_t ldy,
                                                         int nrhs, unsigned long long* flags,
                                                         unsigned long long epoch,
                                                         int* __restrict__ err)
{
    __shared__ float2 part[4][NB][kMaxSolveRhs];
    __shared__ float2 accs[NB][kMaxSolveRhs];
    const int t = threadIdx.x;
    const int b = LOWER ? (int)blockIdx.x : nblk - 1 - (int)blockIdx.x;
    const int64_t rb = (int64_t)b * NB;
    const int lenb = (int)min((int64_t)NB, n - rb);
    const int row = t % NB, grp = t / NB, k0 = 16 * grp;
    const float2 zero = make_float2(0.f, 0.f);
    float2 s[kMaxSolveRhs];
#pragma unroll
    for (int h = 0; h < kMaxSolveRhs; ++h) s[h] = zero;
    // the diagonal-block inverse, this thread's 16 columns of row `row` (the upper sweep's
    // inverses are stored transposed)
    float2 dv[16];
    const float2* D = Dinv + (int64_t)b * NB * NB;
#pragma unroll
    for (int q = 0; q < 16; ++q) dv[q] = LOWER ? D[row + NB * (k0 + q)] : D[(k0 + q) + NB * row];
    // Dependencies are consumed in the order they complete: ascending for the lower sweep,
    // descending for the upper sweep (block b + 1 is the last of b's dependencies to finish).
    const int ndep = LOWER ? b : nblk - 1 - b;
    for (int q = 0; q < ndep; ++q) {
        const int c = LOWER ? q : nblk - 1 - q;
        const int64_t cb = (int64_t)c * NB;
        const int lenc = (int)min((int64_t)NB, n - cb);
        float2 fv[16];
#pragma unroll
        for (int m = 0; m < 16; ++m)
            fv[m] = (row < lenb && k0 + m < lenc) ? F[(rb + row) + ld * (cb + k0 + m)] : zero;
        if (t == 0) {
            long long spins = 0;
            while (__hip_atomic_load(&flags[c], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kSpinLimit) {
                    atomicExch(err, 1);
                    break;
                }
            }
        }
        __syncthreads();
        // plain loads: the agent-scope acquire above invalidated this CU's vector L1
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            if (k0 + m < lenc) {
#pragma unroll
                for (int h = 0; h < kMaxSolveRhs; ++h)
                    if (h < nrhs) {
                        const float2 y = Y[(cb + k0 + m) + ldy * h];
                        s[h].x += fv[m].x * y.x - fv[m].y * y.y;
                        s[h].y += fv[m].x * y.y + fv[m].y * y.x;
                    }
            }
        }
    }
    // acc = y_b - sum_c F_bc y_c  (rows of this block)
#pragma unroll
    for (int h = 0; h < kMaxSolveRhs; ++h) part[grp][row][h] = s[h];
    __syncthreads();
    for (int e = t; e < NB * nrhs; e += 256) {
        const int r = e % NB, h = e / NB;
        float2 a = r < lenb ? Y[(rb + r) + ldy * h] : zero;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            a.x -= part[g][r][h].x;
            a.y -= part[g][r][h].y;
        }
        accs[r][h] = a;
    }
    __syncthreads();
    // y_b = Dinv_b acc: partial over this thread's 16 k, then summed over the 4 groups
#pragma unroll
    for (int h = 0; h < kMaxSolveRhs; ++h) {
        float2 v = zero;
        if (h < nrhs) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const float2 a = accs[k0 + q][h];
                v.x += dv[q].x * a.x - dv[q].y * a.y;
                v.y += dv[q].x * a.y + dv[q].y * a.x;
            }
        }
        s[h] = v;
    }
    __syncthreads(); // part[] is reused
#pragma unroll
    for (int h = 0; h < kMaxSolveRhs; ++h) part[grp][row][h] = s[h];
    __syncthreads();
    if (grp == 0 && row < lenb) {
        for (int h = 0; h < nrhs; ++h) {
            float2 v = part[0][row][h];
#pragma unroll
            for (int g = 1; g < 4; ++g) {
                v.x += part[g][row][h].x;
                v.y += part[g][row][h].y;
            }
            Y[(rb + row) + ldy * h] = v;
        }
    }
    __threadfence();
    __syncthreads();
    if (t == 0) __hip_atomic_store(&flags[b], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- FP64 residual and updates -------------------------------------------------------------------
// part[c][h][i] = sum_{j in chunk c} A[i][j] X[j][h]
__global__ void residual_partial_kernel(int64_t n, const double2* __restrict__ A, int64_t lda,
                                        const double2* __restrict__ X, int64_t ldx, int nrhs,
                                        double2* __restrict__ part)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n) return;
    const int64_t j0 = n * c / kResidChunks, j1 = n * (c + 1) / kResidChunks;
    double2 s[kMaxSolveRhs];
#pragma unroll
    for (int h = 0; h < kMaxSolveRhs; ++h) s[h] = make_double2(0.0, 0.0);
#pragma unroll 4
    for (int64_t j = j0; j < j1; ++j) {
        const double2 a = A[i + lda * j];
#pragma unroll
        for (int h = 0; h < kMaxSolveRhs; ++h)
            if (h < nrhs) {
                const double2 x = X[j + ldx * h];
                s[h].x = fma(a.x, x.x, fma(-a.y, x.y, s[h].x));
                s[h].y = fma(a.x, x.y, fma(a.y, x.x, s[h].y));
            }
    }
#pragma unroll
    for (int h = 0; h < kMaxSolveRhs; ++h)
        if (h < nrhs) part[((int64_t)c * nrhs + h) * n + i] = s[h];
}

// R = B - sum_c part[c]; also per-column max |R_i| (atomic max on the bit pattern of a
// non-negative double) into rmax[h].
__global__ void residual_reduce_kernel(int64_t n, const double2* __restrict__ B, int64_t ldb,
                                       int nrhs, const double2* __restrict__ part,
                                       double2* __restrict__ R, unsigned long long* rmax)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    double m = 0.0;
    if (i < n) {
        double2 s = B[i + ldb * h];
        for (int c = 0; c < kResidChunks; ++c) {
            const double2 p = part[((int64_t)c * nrhs + h) * n + i];
            s.x -= p.x;
            s.y -= p.y;
        }
        R[i + n * h] = s;
        m = hypot(s.x, s.y);
        if (!(m == m)) m = INFINITY; // NaN poisons the norm
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(&rmax[h], (unsigned long long)__double_as_longlong(m));
}

// Y (FP32) = R * inv_scale[h]
__global__ void demote_kernel(int64_t n, int nrhs, const double2* __restrict__ R,
                              const double* __restrict__ scale, float2* __restrict__ Y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    if (i >= n) return;
    const double s = scale[h];
    const double2 r = R[i + n * h];
    Y[i + n * h] = make_float2((float)(r.x * s), (float)(r.y * s));
}

// X += Y / inv_scale[h]; per-column max |X_i| into xmax[h].
__global__ void update_kernel(int64_t n, int nrhs, const float2* __restrict__ Y,
                              const double* __restrict__ scale, double2* __restrict__ X,
                              int64_t ldx, unsigned long long* xmax)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    double m = 0.0;
    if (i < n) {
        const double s = 1.0 / scale[h];
        const float2 y = Y[i + n * h];
        double2 x = X[i + ldx * h];
        x.x += (double)y.x * s;
        x.y += (double)y.y * s;
        X[i + ldx * h] = x;
        m = hypot(x.x, x.y);
        if (!(m == m)) m = INFINITY;
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(&xmax[h], (unsigned long long)__double_as_longlong(m));
}

__global__ void zero_u64_kernel(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void zero_complex_kernel(int64_t n, int nrhs, double2* __restrict__ X, int64_t ldx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) X[i + ldx * blockIdx.y] = make_double2(0.0, 0.0);
}

// scale[h] = 1 / max(rmax[h], tiny)
__global__ void scale_kernel(int nrhs, const unsigned long long* __restrict__ rmax,
                             double* __restrict__ scale)
{
    const int h = threadIdx.x;
    if (h < nrhs) {
        const double m = __longlong_as_double((long long)rmax[h]);
        scale[h] = (m > 1e-300 && m < INFINITY) ? 1.0 / m : 1.0;
    }
}

} // namespace

struct MixedState {
};

void bem2d_internal_mixed_release(bem2d_ctx* c)
{
    delete c->mixed;
    c->mixed = nullptr;
}

struct bem2d_lu32 {
    bem2d_ctx* ctx = nullptr;
    int64_t n = 0;
    int nblk = 0;
    bool factored = false;
    float2* F = nullptr;          // n x n factors, ld = n
    float2* Linv = nullptr;       // nblk x 64 x 64
    float2* UinvT = nullptr;      // nblk x (U_bb^{-1})^T
    unsigned long long* flags = nullptr; // 2 * nblk (lower, upper)
    unsigned long long epoch = 0;
    int* dflag = nullptr;         // [0] bad pivot, [1] sweep time-out
    double* anorm = nullptr;      // |A|_inf
    double* rowpart = nullptr;    // 16 x n
    // refinement workspace, grown on demand
    int rhs_cap = 0;
    double2* R = nullptr;         // n x rhs
    double2* part = nullptr;      // chunks x rhs x n
    float2* Y = nullptr;          // n x rhs
    double* scale = nullptr;      // rhs
    unsigned long long* norms = nullptr; // rmax[rhs], xmax[rhs], bmax[rhs]
};

static void lu32_free_ws(bem2d_lu32* f)
{
    (void)hipFree(f->R);
    (void)hipFree(f->part);
    (void)hipFree(f->Y);
    (void)hipFree(f->scale);
    (void)hipFree(f->norms);
    f->R = nullptr;
    f->part = nullptr;
    f->Y = nullptr;
    f->scale = nullptr;
    f->norms = nullptr;
    f->rhs_cap = 0;
}

extern "C" void bem2d_lu32_destroy(bem2d_lu32* f)
{
    if (!f) return;
    (void)hipSetDevice(f->ctx->device);
    (void)hipStreamSynchronize(f->ctx->stream);
    (void)hipFree(f->F);
    (void)hipFree(f->Linv);
    (void)hipFree(f->UinvT);
    (void)hipFree(f->flags);
    (void)hipFree(f->dflag);
    (void)hipFree(f->anorm);
    (void)hipFree(f->rowpart);
    lu32_free_ws(f);
    delete f;
}

extern "C" int bem2d_lu32_create(bem2d_ctx* c, int64_t n, bem2d_lu32** out)
{
    if (!c || !out || n < 1) return FAIL(BEM2D_ERR_ARG, "bem2d_lu32_create: bad argument");
    HIPCHK(hipSetDevice(c->device));
    bem2d_lu32* f = new bem2d_lu32();
    f->ctx = c;
    f->n = n;
    f->nblk = (int)((n + NB - 1) / NB);
    const size_t inv_bytes = sizeof(float2) * (size_t)f->nblk * NB * NB;
    if (hipMalloc(&f->F, sizeof(float2) * n * n) != hipSuccess ||
        hipMalloc(&f->Linv, inv_bytes) != hipSuccess || hipMalloc(&f->UinvT, inv_bytes) != hipSuccess ||
        hipMalloc(&f->flags, sizeof(unsigned long long) * 2 * f->nblk) != hipSuccess ||
        hipMalloc(&f->dflag, sizeof(int) * 2) != hipSuccess ||
        hipMalloc(&f->anorm, sizeof(double)) != hipSuccess ||
        hipMalloc(&f->rowpart, sizeof(double) * 16 * n) != hipSuccess) {
        bem2d_lu32_destroy(f);
        return FAIL(BEM2D_ERR_NOMEM, "bem2d_lu32_create: device allocation of %lld bytes failed",
                    (long long)(8 * n * n));
    }
    zero_u64_kernel<<<(2 * f->nblk + 255) / 256, 256, 0, c->stream>>>(f->flags, 2 * f->nblk);
    HIPCHK(hipGetLastError());
    *out = f;
    return BEM2D_OK;
}

extern "C" int64_t bem2d_lu32_size(const bem2d_lu32* f) { return f ? f->n : -1; }

extern "C" int bem2d_lu32_factor(bem2d_ctx* c, bem2d_lu32* f, const void* A64, int64_t lda)
{
    if (!c || !f || !A64 || lda < f->n) return FAIL(BEM2D_ERR_ARG, "bem2d_lu32_factor: bad argument");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = f->n;
    const double2* A = (const double2*)A64;
    bem2d_internal_prof_begin(c, BEM2D_PROF_LU);
    {
        dim3 g((unsigned)((n + 255) / 256), (unsigned)n);
        cast_kernel<<<g, 256, 0, s>>>(n, A, lda, f->F, n);
        dim3 g2((unsigned)((n + 255) / 256), 16);
        rowsum_kernel<<<g2, 256, 0, s>>>(n, A, lda, 16, f->rowpart);
        rowsum_max_kernel<<<1, 256, 0, s>>>(n, 16, f->rowpart, f->anorm);
        zero_u64_kernel<<<1, 64, 0, s>>>((unsigned long long*)f->dflag, 1);
    }
    // Step b: panel solves of block b, then the trailing update, whose workgroup (0, 0) also
    // factorises diagonal block b + 1 as soon as its own tile is updated.
    diag_lu_kernel<<<1, 256, 0, s>>>(f->F, n, 0, (int)std::min<int64_t>(NB, n), f->Linv, f->UinvT,
                                     f->dflag);
    for (int b = 0; b < f->nblk; ++b) {
        const int64_t k0 = (int64_t)b * NB;
        const int kb = (int)std::min<int64_t>(NB, n - k0);
        const int64_t rest = n - k0 - kb;
        if (rest > 0) {
            const int tiles = (int)((rest + NB - 1) / NB);
            panel_trsm_kernel<<<2 * tiles, 256, 0, s>>>(f->F, n, n, k0, kb, tiles,
                                                        f->Linv + (size_t)b * NB * NB,
                                                        f->UinvT + (size_t)b * NB * NB);
            dim3 g((unsigned)tiles, (unsigned)tiles);
            trailing_update_kernel<<<g, 256, 0, s>>>(f->F, n, n, k0, kb,
                                                     f->Linv + (size_t)(b + 1) * NB * NB,
                                                     f->UinvT + (size_t)(b + 1) * NB * NB,
                                                     f->dflag);
        }
    }
    bem2d_internal_prof_end(c, BEM2D_PROF_LU);
    HIPCHK(hipGetLastError());
    f->factored = true;
    return BEM2D_OK;
}

static int lu32_reserve(bem2d_lu32* f, int nrhs)
{
    if (nrhs <= f->rhs_cap) return BEM2D_OK;
    (void)hipStreamSynchronize(f->ctx->stream);
    lu32_free_ws(f);
    const int64_t n = f->n;
    if (hipMalloc(&f->R, sizeof(double2) * n * nrhs) != hipSuccess ||
        hipMalloc(&f->part, sizeof(double2) * kResidChunks * nrhs * n) != hipSuccess ||
        hipMalloc(&f->Y, sizeof(float2) * n * nrhs) != hipSuccess ||
        hipMalloc(&f->scale, sizeof(double) * nrhs) != hipSuccess ||
        hipMalloc(&f->norms, sizeof(unsigned long long) * 3 * nrhs) != hipSuccess) {
        lu32_free_ws(f);
        return FAIL(BEM2D_ERR_NOMEM, "refinement workspace allocation failed");
    }
    f->rhs_cap = nrhs;
    return BEM2D_OK;
}

// FP32 solve of L U Y = Y in place for up to kMaxSolveRhs columns (ld = n).
static int lu32_apply(bem2d_lu32* f, float2* Y, int nrhs)
{
    hipStream_t s = f->ctx->stream;
    f->epoch++;
    block_trsv_kernel<true><<<f->nblk, 256, 0, s>>>(f->F, f->n, f->n, f->nblk, f->Linv, Y, f->n,
                                                     nrhs, f->flags, f->epoch, f->dflag + 1);
    block_trsv_kernel<false><<<f->nblk, 256, 0, s>>>(f->F, f->n, f->n, f->nblk, f->UinvT, Y, f->n,
                                                      nrhs, f->flags + f->nblk, f->epoch,
                                                      f->dflag + 1);
    return hipGetLastError() == hipSuccess ? BEM2D_OK : BEM2D_ERR_DEVICE;
}

extern "C" int bem2d_lu32_refine(bem2d_ctx* c, bem2d_lu32* f, const void* A64, int64_t lda,
                                 const void* B, int64_t ldb, void* X, int64_t ldx, int nrhs,
                                 double tol, int maxit, bem2d_ir_report* rep)
{
    if (!c || !f || !A64 || !B || !X || nrhs < 1 || lda < f->n || ldb < f->n || ldx < f->n ||
        maxit < 1)
        return FAIL(BEM2D_ERR_ARG, "bem2d_lu32_refine: bad argument");
    if (!f->factored) return FAIL(BEM2D_ERR_ARG, "bem2d_lu32_refine: not factored");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t n = f->n;
    bem2d_ir_report r{};
    double worst = INFINITY;
    for (int h0 = 0; h0 < nrhs; h0 += kMaxSolveRhs) {
        const int nr = std::min(kMaxSolveRhs, nrhs - h0);
        int rc = lu32_reserve(f, nr);
        if (rc) return rc;
        const double2* Bh = (const double2*)B + ldb * h0;
        double2* Xh = (double2*)X + ldx * h0;
        unsigned long long* rmax = f->norms;
        unsigned long long* xmax = f->norms + nr;
        unsigned long long* bmax = f->norms + 2 * nr;
        dim3 gv((unsigned)((n + 255) / 256), (unsigned)nr);
        zero_complex_kernel<<<gv, 256, 0, s>>>(n, nr, Xh, ldx);
        // |b|_inf via the residual of x = 0
        zero_u64_kernel<<<1, 64, 0, s>>>(f->norms, 3 * nr);
        dim3 gp((unsigned)((n + 255) / 256), kResidChunks);
        residual_partial_kernel<<<gp, 256, 0, s>>>(n, (const double2*)A64, lda, Xh, ldx, nr, f->part);
        residual_reduce_kernel<<<gv, 256, 0, s>>>(n, Bh, ldb, nr, f->part, f->R, bmax);
        std::vector<unsigned long long> hb(3 * nr);
        double anorm = 0.0;
        int dflag[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(&anorm, f->anorm, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(dflag, f->dflag, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (dflag[0]) {
            r.iterations = 0;
            r.converged = 0;
            r.backward_error = INFINITY;
            r.matrix_norm_inf = anorm;
            if (rep) *rep = r;
            return FAIL(BEM2D_ERR_NOT_CONVERGED, "FP32 LU met a zero or non-finite pivot");
        }
        HIPCHK(hipMemcpyAsync(hb.data(), f->norms, sizeof(unsigned long long) * 3 * nr,
                              hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::vector<double> bnorm(nr);
        for (int h = 0; h < nr; ++h) {
            double v;
            std::memcpy(&v, &hb[2 * nr + h], sizeof v);
            bnorm[h] = v;
        }
        // R holds b; copy rmax = bmax for the first scaling
        HIPCHK(hipMemcpyAsync(rmax, bmax, sizeof(unsigned long long) * nr, hipMemcpyDeviceToDevice, s));
        double omega = INFINITY;
        int it = 0;
        for (; it < maxit; ++it) {
            scale_kernel<<<1, 64, 0, s>>>(nr, rmax, f->scale);
            demote_kernel<<<gv, 256, 0, s>>>(n, nr, f->R, f->scale, f->Y);
            rc = lu32_apply(f, f->Y, nr);
            if (rc) return FAIL(rc, "triangular sweep launch failed");
            zero_u64_kernel<<<1, 64, 0, s>>>(f->norms, 2 * nr);
            update_kernel<<<gv, 256, 0, s>>>(n, nr, f->Y, f->scale, Xh, ldx, xmax);
            residual_partial_kernel<<<gp, 256, 0, s>>>(n, (const double2*)A64, lda, Xh, ldx, nr, f->part);
            residual_reduce_kernel<<<gv, 256, 0, s>>>(n, Bh, ldb, nr, f->part, f->R, rmax);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(hb.data(), f->norms, sizeof(unsigned long long) * 2 * nr,
                                  hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(dflag, f->dflag, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (dflag[1]) return FAIL(BEM2D_ERR_DEVICE, "triangular sweep wait timed out");
            omega = 0.0;
            for (int h = 0; h < nr; ++h) {
                double rn, xn;
                std::memcpy(&rn, &hb[h], sizeof rn);
                std::memcpy(&xn, &hb[nr + h], sizeof xn);
                const double den = anorm * xn + bnorm[h];
                const double w = den > 0 ? rn / den : (rn > 0 ? INFINITY : 0.0);
                omega = std::max(omega, std::isfinite(w) ? w : INFINITY);
            }
            if (omega <= tol) {
                ++it;
                break;
            }
        }
        r.iterations = std::max(r.iterations, it);
        worst = (h0 == 0) ? omega : std::max(worst, omega);
        r.matrix_norm_inf = anorm;
    }
    r.backward_error = worst;
    r.converged = worst <= tol;
    if (rep) *rep = r;
    if (!r.converged)
        return FAIL(BEM2D_ERR_NOT_CONVERGED, "refinement reached backward error %.3e > %.3e",
                    worst, tol);
    return BEM2D_OK;
}

// One-shot mixed solve with the FP64 fallback: B (n x nrhs) is overwritten with the solution.
extern "C" int bem2d_solve_mixed(bem2d_ctx* c, int64_t n, int64_t nrhs, const void* A64,
                                 int64_t lda, void* B, int64_t ldb, double tol,
                                 bem2d_ir_report* rep)
{
    if (!c || !A64 || !B || n < 1 || nrhs < 1 || lda < n || ldb < n)
        return FAIL(BEM2D_ERR_ARG, "bem2d_solve_mixed: bad argument");
    bem2d_lu32* f = nullptr;
    int rc = bem2d_lu32_create(c, n, &f);
    if (rc) return rc;
    double2* X = nullptr;
    if (hipMalloc(&X, sizeof(double2) * n * nrhs) != hipSuccess) {
        bem2d_lu32_destroy(f);
        return FAIL(BEM2D_ERR_NOMEM, "bem2d_solve_mixed: allocation failed");
    }
    bem2d_ir_report r{};
    rc = bem2d_lu32_factor(c, f, A64, lda);
    if (!rc) rc = bem2d_lu32_refine(c, f, A64, lda, B, ldb, X, n, (int)nrhs, tol, 30, &r);
    if (rc == BEM2D_OK) {
        rc = hipMemcpy2DAsync(B, sizeof(double2) * ldb, X, sizeof(double2) * n,
                              sizeof(double2) * n, nrhs, hipMemcpyDeviceToDevice, c->stream) ==
                     hipSuccess
                 ? BEM2D_OK
                 : FAIL(BEM2D_ERR_DEVICE, "copy of the solution failed");
    } else if (rc == BEM2D_ERR_NOT_CONVERGED) {
        // FP64 fallback on a copy of A (the caller's A is const)
        r.fallback_fp64 = 1;
        rc = bem2d_solve_fp64(c, n, nrhs, const_cast<void*>(A64), lda, B, ldb, 1);
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(X);
    bem2d_lu32_destroy(f);
    if (rep) *rep = r;
    return rc;
}
