// Mixed-precision dense complex solve: FP32 LU on the MFMA cores + FP64 iterative refinement.
//
// Factorisation (bem2d_lu32_factor): A32 = (complex64) A64 in a static pivoting order (rows
// claimed by the columns where they are largest), bordered by the identity to whole 64-blocks,
// then a blocked right-looking LU with 64-wide panels and no row interchanges.  Step b: the block
// column below and the block row right of diagonal block b are multiplied by the inverses of its
// triangles (panel_trsm_kernel, MFMA), and the trailing matrix receives the update
// A22 -= L21 U12 as a complex GEMM on v_mfma_f32_16x16x4f32 with FP32 accumulation (three real
// MFMA products per complex product, Gauss's form).  On large trailing matrices panels go in
// groups of two, four or eight: each step of a group updates only the next panel's block row and
// column, and the group's last step applies all its panels to the rest at depth 128 to 512.  The
// diagonal block b + 1 is factorised -- LU plus the inverses of both triangles, blocked in
// 16-column panels inside one workgroup (diag_factor) -- by workgroup 0 of the update while the
// rest of the update runs (small trailing matrices), or by diag_lu_kernel after it (large ones,
// whose update then fits two workgroups per compute unit).  Without interchanges the panel has no
// sequential pivot search, the step that makes partial-pivoting panels latency bound on a GPU.  A
// zero or non-finite pivot, or element growth above 1e4, is flagged; refinement then reports
// failure and the one-shot solver falls back to the FP64 LU.
//
// Refinement (bem2d_lu32_refine): x0 = 0, r = b - A64 x in FP64, d = U^{-1} L^{-1} r in FP32
// (two persistent block-triangular sweeps with the inverted diagonal blocks; blocks hand their
// solution on through tagged words; rocBLAS ctrsm for 64 right-hand sides and more), x += d in
// FP64, until the backward error |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) <= tol for every
// right-hand side; GMRES-IR preconditioned by the FP32 factors takes over a stalled refinement.
//
// Every kernel takes a by-value batch of up to kMaxLuBatch problems, so that several problems of
// different sizes (each at its own panel step, or its own refinement step) advance in one launch
// per kernel; the single-problem entry points are batches of one.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/bem2d_hip.h"
#include "internal.h"

namespace {

constexpr int NB = 64;             // panel / block width
constexpr int kMaxSolveRhs = 8;    // right-hand sides handled per triangular sweep
constexpr int kResidChunks = 64;   // column chunks of the FP64 residual
constexpr long long kSpinLimit = 1ll << 26;
// Panels group in twos (one depth-128 trailing update per two panels), fours (depth 256) or eights
// (depth 512) while the trailing matrix has at least tune.lu_pair_tiles / lu_quad_tiles /
// lu_oct_tiles 64-wide tiles: the saved passes over it (16 B per entry each) and the longer runs of
// products per output tile then outweigh the narrow cross updates and diagonal factorisations the
// group leaves exposed (~60 us each).  Measured on MI355X with the three-product update: twos
// from 64 tiles, fours from 96, eights from 150 -- N = 4000 3.8 ms (ungrouped: at most 62 tiles),
// N = 8000 18.2 ms ungrouped -> 15.9 ms, N = 12000 45.7 -> 42.5 ms (against twos from 100, fours
// from 200, eights from 250), N = 32000 700 ms with twos only -> 613 ms; sixteens gained nothing.

typedef float v4f __attribute__((ext_vector_type(4)));

// ---- batched launches ----------------------------------------------------------------------------
// Every factorisation and refinement kernel serves up to kMaxLuBatch problems per launch; the
// descriptors travel by value in the kernel arguments and workgroup b belongs to the problem d with
// start[d] <= b < start[d + 1].  A single problem is a batch of one.
constexpr int kMaxLuBatch = 16;

struct LuProblem {
    float2* F;              // n x n FP32 factors (ld = n)
    int64_t n;
    int64_t k0;             // first row / column of the current diagonal block
    int kb;                 // its size
    int tiles;              // 64-wide tiles of the trailing matrix
    int ntile;              // output tiles of the trailing update
    int nwg;                // its persistent workgroups
    int mode;               // trailing update: LU_UPDATE_FULL or LU_UPDATE_CROSS
    const float2* Linv;     // inverses of the current diagonal block (panel solves)
    const float2* UinvT;
    float2* Linv_next;      // inverses to produce for the next diagonal block, or null
    float2* UinvT_next;
    int* bad;               // raised on a zero or non-finite pivot
};

// Trailing update over the full tile grid, or over its first block row and block column only
// (the "cross": the next panel, updated at once while the rest waits for the group's update).
enum { LU_UPDATE_FULL = 0, LU_UPDATE_CROSS = 1 };

struct LuBatch {
    int count;
    int64_t start[kMaxLuBatch + 1];
    LuProblem p[kMaxLuBatch];
};

__device__ __forceinline__ int find_problem(const int64_t* start, int count, int64_t b)
{
    int d = 0;
    while (d + 1 < count && b >= start[d + 1]) ++d;
    return d;
}

// The start of a factorisation (zeroed flags, static pivoting order, cast, |A|_inf) for up to
// kMaxLuBatch problems per launch; workgroup b belongs to problem d with start[d] <= b < start[d+1]
// (each kernel fills start[] with its own per-problem workgroup counts).
struct BeginProblem {
    int64_t n, np, lda;
    const double2* A;
    int* perm;
    float2* F;
    double* part;                 // (np / kCastCols) x n partial row sums
    int* pivwork;                 // colarg | freepos | freerow, n each
    unsigned long long* claim;    // n
    unsigned long long* gmax;     // max |U|, max |L|, max |a_ij|
    unsigned long long* anorm;    // |A|_inf (double bits)
    int* dflag;                   // [0] bad pivot, [1] sweep time-out
};
struct BeginBatch {
    int count;
    int64_t start[kMaxLuBatch + 1];
    BeginProblem p[kMaxLuBatch];
};


__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cfms(float2 acc, float2 a, float2 b) // acc - a b
{
    // two fused multiply-adds per component (written out: acc - (ab - cd) would keep the inner
    // difference and cost a third instruction)
    acc.x = fmaf(-a.x, b.x, acc.x);
    acc.x = fmaf(a.y, b.y, acc.x);
    acc.y = fmaf(-a.x, b.y, acc.y);
    acc.y = fmaf(-a.y, b.x, acc.y);
    return acc;
}
__device__ __forceinline__ float2 cfma(float2 acc, float2 a, float2 b) // acc + a b
{
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
    return acc;
}
// Selections on values.  `c ? u : v` between two float2 lvalues selects between their ADDRESSES
// (a conditional of lvalues is an lvalue), which forces a named constant such as a zero into
// scratch memory and turns the reads into scratch / flat loads inside the step loops.
__device__ __forceinline__ float2 csel(bool c, float2 u, float2 v)
{
    return make_float2(c ? u.x : v.x, c ? u.y : v.y);
}
// A guarded load without a branch: the load always issues (from `safe`, a valid element, when !ok)
// and the value is selected afterwards.  With a branch around each load the compiler may copy the
// result inside the branch and wait for the load there, which turns a batch of guarded loads into
// one memory round trip each.
__device__ __forceinline__ float2 load_or(bool ok, const float2* p, const float2* safe, float2 v)
{
    const float2 x = *(ok ? p : safe);
    return csel(ok, x, v);
}

// ---- cast and |A|_inf ------------------------------------------------------------------------------
// The np x np FP32 matrix: A cast to FP32, bordered by the identity (rows and columns n .. np): the
// factors of the bordered matrix are those of A bordered by the identity, and every 64 x 64 tile of
// the factorisation is full.  Workgroup (x, y) covers rows 256 x .. 256 x + 255 of the kCastCols
// columns of chunk y and leaves the FP64 row sums of |a_ij| over the chunk in part[y][i], so that
// |A|_inf = max_i sum_j |a_ij| costs no second pass over A.
constexpr int kCastCols = 64;

// Workgroups of a problem: ceil(np / 256) row groups x (np / kCastCols) column chunks.
__global__ __launch_bounds__(256) void cast_kernel(const BeginBatch bt)
{
    const int di = find_problem(bt.start, bt.count, blockIdx.x);
    const BeginProblem& pb = bt.p[di];
    const int64_t n = pb.n, np = pb.np, lda = pb.lda;
    const double2* __restrict__ A = pb.A;
    const int* __restrict__ perm = pb.perm;
    float2* __restrict__ F = pb.F;
    double* __restrict__ part = pb.part;
    const int64_t lb = (int64_t)blockIdx.x - bt.start[di], nx = (np + 255) / 256;
    const int64_t bx = lb % nx, by = lb / nx;
    const int64_t i = bx * 256 + threadIdx.x;
    const int64_t j0 = by * kCastCols;
    if (i >= np) return;
    // row i of the factorised matrix is row perm[i] of A (the static pivoting order)
    const int64_t pi = i < n ? (int64_t)perm[i] : 0;
    double s = 0.0;
    constexpr int kBatch = 16;
    for (int c0 = 0; c0 < kCastCols; c0 += kBatch) {
        double2 v[kBatch];
#pragma unroll
        for (int c = 0; c < kBatch; ++c) {
            const int64_t j = j0 + c0 + c;
            v[c] = *(i < n && j < n ? &A[pi + lda * j] : A);
        }
#pragma unroll
        for (int c = 0; c < kBatch; ++c) {
            const int64_t j = j0 + c0 + c;
            const bool in = i < n && j < n;
            const double re = in ? v[c].x : (i == j ? 1.0 : 0.0), im = in ? v[c].y : 0.0;
            F[i + np * j] = make_float2((float)re, (float)im);
            s += in ? sqrt(fma(re, re, im * im)) : 0.0;
        }
    }
    if (i < n) part[by * n + i] = s;
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v)
{
    // the bit patterns of non-negative doubles are ordered like the values
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// Zeroes the claims of the static pivoting, |A|_inf, the growth maxima and the flags;
// ceil(n / 256) workgroups per problem.
__global__ __launch_bounds__(256) void begin_init_kernel(const BeginBatch bt)
{
    const int di = find_problem(bt.start, bt.count, blockIdx.x);
    const BeginProblem& pb = bt.p[di];
    const int64_t lb = (int64_t)blockIdx.x - bt.start[di];
    const int64_t i = lb * blockDim.x + threadIdx.x;
    if (i < pb.n) pb.claim[i] = 0ull;
    if (lb == 0 && threadIdx.x < 3) pb.gmax[threadIdx.x] = 0ull;
    if (lb == 0 && threadIdx.x == 3) *pb.anorm = 0ull;
    if (lb == 0 && threadIdx.x < 2) pb.dflag[threadIdx.x] = 0;
}

// anorm = max_i sum_c part[c][i] (zeroed beforehand); ceil(n / 256) workgroups per problem.
__global__ __launch_bounds__(256) void rowsum_max_kernel(const BeginBatch bt)
{
    __shared__ double red[256];
    const int di = find_problem(bt.start, bt.count, blockIdx.x);
    const BeginProblem& pb = bt.p[di];
    const int64_t n = pb.n;
    const int nchunk = (int)(pb.np / kCastCols);
    const double* __restrict__ part = pb.part;
    unsigned long long* __restrict__ out = pb.anorm;
    const int64_t i = ((int64_t)blockIdx.x - bt.start[di]) * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (i < n) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        int c = 0;
        for (; c + 4 <= nchunk; c += 4) {
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] += part[(int64_t)(c + q) * n + i];
        }
        for (; c < nchunk; ++c) acc[0] += part[(int64_t)c * n + i];
        m = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomic_max_nonneg(out, red[0]);
}

// ---- static pivoting ------------------------------------------------------------------------------
// The factorisation runs without interchanges on the rows of A in an order fixed beforehand from
// the matrix itself (static pivoting, as in SuperLU_DIST: X. S. Li and J. W. Demmel, "Making
// sparse Gaussian elimination scalable by static pivoting", SC 1998, where a row permutation puts
// large entries on the diagonal, the factorisation then never pivots, and iterative refinement
// recovers the accuracy).  Column j claims the row of its largest entry; a row claimed by several
// columns goes to the column in which its entry is largest (ties: the smaller column index); the
// positions left over take the unclaimed rows in increasing order.  So a matrix whose largest
// entry of every column lies in a different row -- a BEM matrix in its natural order, or any row
// permutation of one -- gets exactly those rows on its diagonal, and any matrix gets a
// permutation.  perm[i] = row of A at position i of the factorised matrix (i < n).
// Key of a claim: |a_rj| as float bits above the complement of j (larger value first, then the
// smaller j); never 0, so 0 marks an unclaimed row.

// One workgroup per column: colarg[j] = argmax_i |a_ij| (the smallest i on ties), the claim of
// that row, and amax = max_ij |a_ij| (NaN entries make it NaN).
// n workgroups per problem.
__global__ __launch_bounds__(256) void pivot_colmax_kernel(const BeginBatch bt)
{
    __shared__ double sb[4];
    __shared__ int sa[4];
    const int di = find_problem(bt.start, bt.count, blockIdx.x);
    const BeginProblem& pb = bt.p[di];
    const int64_t n = pb.n, lda = pb.lda;
    const double2* __restrict__ A = pb.A;
    int* __restrict__ colarg = pb.pivwork;
    unsigned long long* __restrict__ claim = pb.claim;
    unsigned long long* __restrict__ amax = pb.gmax + 2;
    const int64_t j = (int64_t)blockIdx.x - bt.start[di];
    const double2* __restrict__ col = A + lda * j;
    double best = -1.0;
    int arg = 0;
    bool nan = false;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double2 a = col[i];
        const double m = fma(a.x, a.x, a.y * a.y);
        nan |= !(m == m);
        if (m > best) { // strictly: the smallest i of this thread's rows on ties
            best = m;
            arg = (int)i;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oa = __shfl_xor(arg, off);
        if (ob > best || (ob == best && oa < arg)) {
            best = ob;
            arg = oa;
        }
    }
    const unsigned long long anynan = __ballot(nan);
    __shared__ int snan[4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sb[w] = best;
        sa[w] = arg;
        snan[w] = anynan != 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bool bad = false;
        for (int q = 0; q < 4; ++q) {
            bad |= snan[q] != 0;
            if (sb[q] > best || (sb[q] == best && sa[q] < arg)) {
                best = sb[q];
                arg = sa[q];
            }
        }
        const double v = best > 0.0 ? sqrt(best) : 0.0;
        colarg[j] = arg;
        const unsigned long long key = ((unsigned long long)__float_as_uint((float)v) << 32) |
                                       (unsigned long long)(0xffffffffu - (unsigned)j);
        atomicMax(&claim[arg], key);
        atomicMax(amax, (unsigned long long)__double_as_longlong(bad ? (double)NAN : v));
    }
}

// Exclusive prefix count of flag over a 1024-thread workgroup (wave ballots + 16 wave totals);
// *total = the count over the workgroup.
__device__ __forceinline__ int block_scan_1024(bool flag, int* wave_tot, int* total)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const unsigned long long b = __ballot(flag);
    const int pre = __popcll(b & ((1ull << lane) - 1ull));
    if (lane == 0) wave_tot[w] = __popcll(b);
    __syncthreads();
    int off = 0, tot = 0;
    for (int q = 0; q < 16; ++q) {
        const int v = wave_tot[q];
        off += q < w ? v : 0;
        tot += v;
    }
    __syncthreads();
    *total = tot;
    return off + pre;
}

// One workgroup of 1024 threads: positions whose claim won keep their row; the others, in
// increasing order, take the unclaimed rows in increasing order (freepos / freerow: n ints each).
// One workgroup per problem (start[] unused).
__global__ __launch_bounds__(1024) void pivot_assign_kernel(const BeginBatch bt)
{
    __shared__ int wave_tot[16];
    const BeginProblem& pb = bt.p[blockIdx.x];
    const int64_t n = pb.n;
    const int* __restrict__ colarg = pb.pivwork;
    const unsigned long long* __restrict__ claim = pb.claim;
    int* __restrict__ perm = pb.perm;
    int* __restrict__ freepos = pb.pivwork + n;
    int* __restrict__ freerow = pb.pivwork + 2 * n;
    const int t = threadIdx.x;
    int nfp = 0, nfr = 0;
    for (int64_t base = 0; base < n; base += 1024) {
        const int64_t j = base + t;
        bool taken = false, unclaimed = false;
        if (j < n) {
            const int r = colarg[j];
            taken = (unsigned)(claim[r] & 0xffffffffull) == 0xffffffffu - (unsigned)j;
            if (taken) perm[j] = r;
            unclaimed = claim[j] == 0ull;
        }
        int tot = 0;
        const int fp = block_scan_1024(j < n && !taken, wave_tot, &tot);
        if (j < n && !taken) freepos[nfp + fp] = (int)j;
        nfp += tot;
        const int fr = block_scan_1024(unclaimed, wave_tot, &tot);
        if (unclaimed) freerow[nfr + fr] = (int)j;
        nfr += tot;
    }
    __syncthreads(); // the lists written above are read by other threads of this workgroup
    for (int q = t; q < nfp; q += 1024) perm[freepos[q]] = freerow[q];
}

// Element growth of the factors: gmax[0] = max |U_ij| (i <= j < n), gmax[1] = max |L_ij|
// (j < i < n); NaN or infinite entries count as infinite.  One pass over the factors after the last
// step, HBM-bound: workgroup (x, y) reads rows 1024 x .. 1024 x + 1023 of the 32 columns of chunk y,
// each thread four consecutive rows of a column per load (two 16-byte accesses), and leaves its two
// maxima in part[2 wg], part[2 wg + 1]; growth_reduce_kernel then takes the maxima over the
// workgroups (no atomics on two hot words).
constexpr int kGrowthRows = 1024, kGrowthCols = 32;
__device__ __forceinline__ float growth_abs(float2 v)
{
    const float a = sqrtf(fmaf(v.x, v.x, v.y * v.y));
    return a <= 3.0e38f ? a : INFINITY; // NaN and overflow
}
__global__ __launch_bounds__(256) void growth_kernel(int64_t n, int64_t np, const float2* __restrict__ F,
                                                     float* __restrict__ part)
{
    __shared__ float red[2][4];
    const int64_t nx = (np + kGrowthRows - 1) / kGrowthRows;
    const int64_t bx = blockIdx.x % nx, by = blockIdx.x / nx;
    const int64_t i0 = bx * kGrowthRows + 4 * threadIdx.x;
    const int64_t j0 = by * kGrowthCols;
    float mu = 0.f, ml = 0.f;
    if (i0 < n) {
        constexpr int kUnroll = 8;
        for (int c0 = 0; c0 < kGrowthCols; c0 += kUnroll) {
            float4 v[kUnroll][2];
#pragma unroll
            for (int c = 0; c < kUnroll; ++c) {
                const int64_t j = min(j0 + c0 + c, n - 1);
                const float4* p = reinterpret_cast<const float4*>(F + i0 + np * j);
                v[c][0] = p[0];
                v[c][1] = p[1];
            }
#pragma unroll
            for (int c = 0; c < kUnroll; ++c) {
                const int64_t j = j0 + c0 + c;
                const float2 e[4] = {make_float2(v[c][0].x, v[c][0].y), make_float2(v[c][0].z, v[c][0].w),
                                     make_float2(v[c][1].x, v[c][1].y), make_float2(v[c][1].z, v[c][1].w)};
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t i = i0 + r;
                    const float a = i < n && j < n ? growth_abs(e[r]) : 0.f;
                    if (i <= j) mu = fmaxf(mu, a);
                    else ml = fmaxf(ml, a);
                }
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        mu = fmaxf(mu, __shfl_xor(mu, off));
        ml = fmaxf(ml, __shfl_xor(ml, off));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = mu;
        red[1][w] = ml;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = fmaxf(fmaxf(red[0][0], red[0][1]), fmaxf(red[0][2], red[0][3]));
        part[2 * blockIdx.x + 1] = fmaxf(fmaxf(red[1][0], red[1][1]), fmaxf(red[1][2], red[1][3]));
    }
}

// gmax[0], gmax[1] = the maxima of the nwg workgroup pairs in part (one workgroup of 1024).
__global__ __launch_bounds__(1024) void growth_reduce_kernel(int64_t nwg, const float* __restrict__ part,
                                                             unsigned long long* __restrict__ gmax)
{
    __shared__ float red[2][16];
    float mu = 0.f, ml = 0.f;
    for (int64_t q = threadIdx.x; q < nwg; q += 1024) {
        mu = fmaxf(mu, part[2 * q]);
        ml = fmaxf(ml, part[2 * q + 1]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mu = fmaxf(mu, __shfl_xor(mu, off));
        ml = fmaxf(ml, __shfl_xor(ml, off));
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = mu;
        red[1][threadIdx.x >> 6] = ml;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 16; ++q) {
            mu = fmaxf(mu, red[0][q]);
            ml = fmaxf(ml, red[1][q]);
        }
        mu = fmaxf(mu, red[0][0]);
        ml = fmaxf(ml, red[1][0]);
        gmax[0] = (unsigned long long)__double_as_longlong((double)mu);
        gmax[1] = (unsigned long long)__double_as_longlong((double)ml);
    }
}

// Reciprocal of a complex number (hardware reciprocal of |a|^2).
__device__ __forceinline__ float2 fast_crecip(float2 a)
{
    const float d = __builtin_amdgcn_rcpf(a.x * a.x + a.y * a.y);
    return make_float2(a.x * d, -a.y * d);
}

// ---- diagonal block, blocked: LU without interchanges + inverses of both triangles ----------------
// 256 threads, the tile in LDS, four panels of 16 columns.  The critical chain runs on wavefront 0:
// factorise panel P (rows 16 P .. 63, a row per lane, pivot row broadcast by lane reads: 16
// dependent steps with no barrier), solve the next block of the U row, U_P,P+1 = L_PP^{-1} T_P,P+1,
// update block column P + 1 (MFMA), factorise panel P + 1.  Meanwhile wavefronts 1 - 3 take the
// rest of phase P: the other blocks of the U row and their block columns, X_PP = L_PP^{-1},
// Z_PP = U_PP^{-1} (the 16-column triangular solves use four lanes per column, rows spread over
// the quad and the solved row broadcast by DPP), and the off-diagonal inverse blocks that have
// become computable,
//   X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj,   Z_ji = -(sum_{k=j}^{i-1} Z_jk U_ki) Z_ii,
// as 16 x 16 products on MFMA.  One barrier per phase.  Linv receives X and UinvT receives Z^T
// (column-major, ld 64).  The tile is always a full 64 x 64 block (the factorised matrix is
// bordered to whole blocks).
constexpr int TP = NB + 1; // padded row pitch of the LDS tiles (float2)
struct DiagLds {
    float2 T[NB][TP];       // the tile; L (unit diagonal implied) and U in place
    float2 W[NB][TP];       // X = L^{-1} strictly below the diagonal, Z = U^{-1} on and above it
    float2 S[4][16][17];    // per-wavefront 16 x 16 scratch
    float2 S3[6][16][17];   // sums of block row 3 of X (0 - 2) and block column 3 of Z (3 - 5)
};

__device__ __forceinline__ float lane_read(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

struct Acc16 {
    v4f re, im;
};
__device__ __forceinline__ void acc16_zero(Acc16& c)
{
    c.re = v4f{0.f, 0.f, 0.f, 0.f};
    c.im = v4f{0.f, 0.f, 0.f, 0.f};
}
// c += A B, A and B 16 x 16 complex given by element functions a(i, k), b(k, j); lane l supplies
// A[l % 16][4 s + l / 16] and B[4 s + l / 16][l % 16] to k-slice s of v_mfma_f32_16x16x4f32, and
// holds C[4 (l / 16) + r][l % 16] in register r.
template <class FA, class FB>
__device__ __forceinline__ void mma16(Acc16& c, FA a, FB b)
{
    const int l = threadIdx.x & 63, li = l & 15, lk = l >> 4;
    float2 av[4], bv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        av[s] = a(li, 4 * s + lk);
        bv[s] = b(4 * s + lk, li);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        c.re = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s].x, bv[s].x, c.re, 0, 0, 0);
        c.im = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s].x, bv[s].y, c.im, 0, 0, 0);
        c.re = __builtin_amdgcn_mfma_f32_16x16x4f32(-av[s].y, bv[s].y, c.re, 0, 0, 0);
        c.im = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s].y, bv[s].x, c.im, 0, 0, 0);
    }
}

// Panel P factorised by wavefront 0: lane l holds row 16 P + l.
template <int P>
__device__ __forceinline__ void diag_a1(DiagLds& S, int l, int* __restrict__ bad)
{
    constexpr int b0 = 16 * P;
    constexpr int nr = NB - b0;
    const float2 zero = make_float2(0.f, 0.f);
    const bool act = l < nr;
    const int row = act ? b0 + l : b0;
    float2 a[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) a[c] = csel(act, S.T[row][b0 + c], zero);
    // pivot check after the loop: the smallest |p|^2 must be positive and their sum finite
    float dmin = INFINITY, dsum = 0.f;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const float pr = lane_read(a[j].x, j), pi = lane_read(a[j].y, j);
        const float d = pr * pr + pi * pi;
        dmin = fminf(dmin, d);
        dsum += d;
        const float rd = __builtin_amdgcn_rcpf(d);
        const float2 ip = make_float2(pr * rd, -pi * rd);
        float2 u[16];
#pragma unroll
        for (int c = j + 1; c < 16; ++c) u[c] = make_float2(lane_read(a[c].x, j), lane_read(a[c].y, j));
        if (l > j) {
            const float2 m = cmul(a[j], ip);
            a[j] = m;
#pragma unroll
            for (int c = j + 1; c < 16; ++c) a[c] = cfms(a[c], m, u[c]);
        }
    }
    if (act) {
#pragma unroll
        for (int c = 0; c < 16; ++c) S.T[row][b0 + c] = a[c];
    }
    if (l == 0 && !(dmin > 0.f && isfinite(dsum))) *bad = 1;
}

// Value of quad lane SRC to all four lanes of each quad (DPP quad_perm [SRC, SRC, SRC, SRC]).
template <int SRC>
__device__ __forceinline__ float2 quad_bcast_c(float2 v)
{
    constexpr int ctrl = SRC * 0x55;
    return make_float2(
        __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v.x), ctrl, 0xF, 0xF, false)),
        __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v.y), ctrl, 0xF, 0xF, false)));
}
// (src is a constant once the caller's loop is unrolled; the switch folds away)
__device__ __forceinline__ float2 quad_bcast(float2 v, int src)
{
    switch (src) {
    case 0: return quad_bcast_c<0>(v);
    case 1: return quad_bcast_c<1>(v);
    case 2: return quad_bcast_c<2>(v);
    default: return quad_bcast_c<3>(v);
    }
}

// x = L_PP^{-1} b for 16 columns by one wavefront: lane l takes column l / 4, rows 4 (l % 4) .. +3.
// b(r, c) gives the right-hand side, st(r, c, v) stores the solution.  The lane's 64 coefficients
// L[4 q + i][j] (zero on and above the diagonal) are read before the 16 dependent steps.
template <int P, class FB, class FS>
__device__ __forceinline__ void quad_lower(DiagLds& S, int l, FB b, FS st)
{
    constexpr int b0 = 16 * P;
    const float2 zero = make_float2(0.f, 0.f);
    const int c = l >> 2, q = l & 3;
    float2 lv[4][16];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 16; ++j) lv[i][j] = S.T[b0 + 4 * q + i][b0 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 16; ++j) lv[i][j] = csel(4 * q + i > j, lv[i][j], zero);
    float2 x[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = b(4 * q + i, c);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const float2 xj = quad_bcast(x[j & 3], j >> 2);
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = cfms(x[i], lv[i][j], xj);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) st(4 * q + i, c, x[i]);
}

// Z_PP = U_PP^{-1} by one wavefront (back substitution, same lane layout), into W on and above
// the diagonal.
template <int P>
__device__ __forceinline__ void quad_zinv(DiagLds& S, int l)
{
    constexpr int b0 = 16 * P;
    const float2 zero = make_float2(0.f, 0.f);
    const int c = l >> 2, q = l & 3;
    float2 uv[4][16], inv[16];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 16; ++j) uv[i][j] = S.T[b0 + 4 * q + i][b0 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 16; ++j) uv[i][j] = csel(4 * q + i < j, uv[i][j], zero);
#pragma unroll
    for (int j = 0; j < 16; ++j) inv[j] = fast_crecip(S.T[b0 + j][b0 + j]);
    float2 x[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = make_float2(4 * q + i == c ? 1.f : 0.f, 0.f);
#pragma unroll
    for (int j = 15; j >= 0; --j) {
        x[j & 3] = csel(q == (j >> 2), cmul(x[j & 3], inv[j]), x[j & 3]);
        const float2 xj = quad_bcast(x[j & 3], j >> 2);
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = cfms(x[i], uv[i][j], xj);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (4 * q + i <= c) S.W[b0 + 4 * q + i][b0 + c] = x[i];
}

// U_P,Q = L_PP^{-1} T_P,Q in place
template <int P>
__device__ __forceinline__ void diag_urow(DiagLds& S, int l, int Q)
{
    constexpr int b0 = 16 * P;
    quad_lower<P>(S, l, [&](int r, int c) { return S.T[b0 + r][16 * Q + c]; },
                  [&](int r, int c, float2 v) { S.T[b0 + r][16 * Q + c] = v; });
}

// X_PP = L_PP^{-1}, its strict lower part into W
template <int P>
__device__ __forceinline__ void diag_xinv(DiagLds& S, int l)
{
    constexpr int b0 = 16 * P;
    quad_lower<P>(S, l, [](int r, int c) { return make_float2(r == c ? 1.f : 0.f, 0.f); },
                  [&](int r, int c, float2 v) {
                      if (r > c) S.W[b0 + r][b0 + c] = v;
                  });
}

// T_bi,bj -= L_bi,P U_P,bj (16 x 16 x 16 on MFMA, one wavefront)
template <int P>
__device__ __forceinline__ void diag_update(DiagLds& S, int l, int bi, int bj)
{
    constexpr int b0 = 16 * P;
    Acc16 c;
    acc16_zero(c);
    mma16(c, [&](int i, int k) { return S.T[16 * bi + i][b0 + k]; },
          [&](int k, int j) { return S.T[b0 + k][16 * bj + j]; });
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float2& v = S.T[16 * bi + 4 * (l >> 4) + r][16 * bj + (l & 15)];
        v = make_float2(v.x - c.re[r], v.y - c.im[r]);
    }
}

// Off-diagonal inverse blocks: kind 0 is X_ij (i > j), kind 1 is Z_ji (j < i).  inv_sum leaves
// S_ij = sum_{k=j}^{i-1} L_ik X_kj (kind 0) or sum_{k=j}^{i-1} Z_jk U_ki (kind 1) in a 16 x 16
// scratch block; inv_apply stores X_ij = -X_ii S_ij or Z_ji = -S_ij Z_ii into W.
__device__ __forceinline__ auto inv_X(DiagLds& S, int k, int jj)
{
    // X_kj with its unit diagonal and zero upper part when k == j
    return [&S, k, jj](int r, int c) {
        const float2 v = S.W[16 * k + r][16 * jj + c];
        if (k > jj || r > c) return v;
        return make_float2(r == c ? 1.f : 0.f, 0.f);
    };
}
__device__ __forceinline__ auto inv_Z(DiagLds& S, int jj, int k)
{
    // Z_jk with its zero lower part when j == k
    return [&S, jj, k](int r, int c) {
        const float2 v = S.W[16 * jj + r][16 * k + c];
        return (jj < k || r <= c) ? v : make_float2(0.f, 0.f);
    };
}
__device__ __forceinline__ void inv_sum(DiagLds& S, int kind, int i, int j, float2 (*sc)[17])
{
    const int l = threadIdx.x & 63;
    auto T = [&S](int bi, int bk) {
        return [&S, bi, bk](int r, int c) { return S.T[16 * bi + r][16 * bk + c]; };
    };
    Acc16 s;
    acc16_zero(s);
    for (int k = j; k < i; ++k) {
        if (kind == 0)
            mma16(s, T(i, k), inv_X(S, k, j));
        else
            mma16(s, inv_Z(S, j, k), T(k, i));
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) sc[4 * (l >> 4) + r][l & 15] = make_float2(s.re[r], s.im[r]);
}
__device__ __forceinline__ void inv_apply(DiagLds& S, int kind, int i, int j, float2 (*sc)[17])
{
    const int l = threadIdx.x & 63;
    auto SC = [sc](int r, int c) { return sc[r][c]; };
    Acc16 o;
    acc16_zero(o);
    if (kind == 0)
        mma16(o, inv_X(S, i, i), SC);
    else
        mma16(o, SC, inv_Z(S, i, i));
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int rr = 4 * (l >> 4) + r, cc = l & 15;
        const float2 v = make_float2(-o.re[r], -o.im[r]);
        if (kind == 0)
            S.W[16 * i + rr][16 * j + cc] = v;
        else
            S.W[16 * j + rr][16 * i + cc] = v;
    }
}
__device__ __forceinline__ void diag_inverse_block(DiagLds& S, int w, int kind, int i, int j)
{
    inv_sum(S, kind, i, j, S.S[w]);
    inv_apply(S, kind, i, j, S.S[w]);
}

__device__ __forceinline__ void diag_factor(float2* __restrict__ A, int64_t ld, int64_t k0, int kb,
                                            float2* __restrict__ Linv,
                                            float2* __restrict__ UinvT, int* __restrict__ bad,
                                            DiagLds& S)
{
    (void)kb; // always NB
    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int l = t & 63;
    float2 v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int e = t + 256 * i, r = e & 63, c = e >> 6;
        v[i] = A[(k0 + r) + ld * (k0 + c)];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int e = t + 256 * i, r = e & 63, c = e >> 6;
        S.T[r][c] = v[i];
        S.W[r][c] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    if (w == 0) diag_a1<0>(S, l, bad);
    __syncthreads();
    // phase 0
    if (w == 0) {
        diag_urow<0>(S, l, 1);
        for (int bi = 1; bi < 4; ++bi) diag_update<0>(S, l, bi, 1);
        diag_a1<1>(S, l, bad);
    } else if (w == 1) {
        diag_urow<0>(S, l, 2);
        for (int bi = 1; bi < 4; ++bi) diag_update<0>(S, l, bi, 2);
    } else if (w == 2) {
        diag_urow<0>(S, l, 3);
        for (int bi = 1; bi < 4; ++bi) diag_update<0>(S, l, bi, 3);
    } else {
        diag_xinv<0>(S, l);
        quad_zinv<0>(S, l);
    }
    __syncthreads();
    // phase 1
    if (w == 0) {
        diag_urow<1>(S, l, 2);
        for (int bi = 2; bi < 4; ++bi) diag_update<1>(S, l, bi, 2);
        diag_a1<2>(S, l, bad);
    } else if (w == 1) {
        diag_urow<1>(S, l, 3);
        for (int bi = 2; bi < 4; ++bi) diag_update<1>(S, l, bi, 3);
    } else if (w == 2) {
        diag_xinv<1>(S, l);
        diag_inverse_block(S, w, 0, 1, 0);
    } else {
        quad_zinv<1>(S, l);
        diag_inverse_block(S, w, 1, 1, 0);
    }
    __syncthreads();
    // phase 2
    if (w == 0) {
        diag_urow<2>(S, l, 3);
        diag_update<2>(S, l, 3, 3);
        diag_a1<3>(S, l, bad);
    } else if (w == 2) {
        diag_xinv<2>(S, l);
        diag_inverse_block(S, w, 0, 2, 1);
        diag_inverse_block(S, w, 0, 2, 0);
        // block row 3 of X needs only X_33 beyond this point: its sums now
        for (int j = 0; j < 3; ++j) inv_sum(S, 0, 3, j, S.S3[j]);
    } else if (w == 3) {
        quad_zinv<2>(S, l);
        diag_inverse_block(S, w, 1, 2, 1);
        diag_inverse_block(S, w, 1, 2, 0);
    }
    __syncthreads();
    // phase 3: the last diagonal inverses and the sums of block column 3 of Z (which need U_23,
    // solved by wavefront 0 in phase 2), then the six blocks of row 3 of X and column 3 of Z
    if (w == 0) {
        diag_xinv<3>(S, l);
    } else if (w == 1) {
        quad_zinv<3>(S, l);
    } else if (w == 2) {
        inv_sum(S, 1, 3, 0, S.S3[3]);
    } else {
        inv_sum(S, 1, 3, 1, S.S3[4]);
        inv_sum(S, 1, 3, 2, S.S3[5]);
    }
    __syncthreads();
    if (w == 0) {
        inv_apply(S, 0, 3, 0, S.S3[0]);
        inv_apply(S, 0, 3, 1, S.S3[1]);
    } else if (w == 1) {
        inv_apply(S, 1, 3, 0, S.S3[3]);
        inv_apply(S, 1, 3, 1, S.S3[4]);
    } else if (w == 2) {
        inv_apply(S, 0, 3, 2, S.S3[2]);
    } else {
        inv_apply(S, 1, 3, 2, S.S3[5]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int e = t + 256 * i, r = e & 63, c = e >> 6;
        A[(k0 + r) + ld * (k0 + c)] = S.T[r][c];
        const float2 x = S.W[r][c];
        Linv[r + NB * c] = r > c ? x : make_float2(r == c ? 1.f : 0.f, 0.f);
        const float2 z = S.W[c][r];
        UinvT[r + NB * c] = c <= r ? z : make_float2(0.f, 0.f);
    }
}

// Diagonal block p.k0 .. p.k0 + p.kb of problem blockIdx.x, inverses into Linv_next / UinvT_next.
__global__ __launch_bounds__(256) void diag_lu_kernel(const LuBatch batch)
{
    __shared__ DiagLds S;
    const LuProblem& p = batch.p[blockIdx.x];
    diag_factor(p.F, p.n, p.k0, p.kb, p.Linv_next, p.UinvT_next, p.bad, S);
}

// ---- 64 x 64 output tiles on MFMA -------------------------------------------------------------------
// Wavefront w of a tile computes the 32 x 32 quarter at rows 32 (w & 1), cols 32 (w >> 1) as 2 x 2
// blocks of 16 x 16; each complex product is four real v_mfma_f32_16x16x4f32.
// v_mfma_f32_16x16x4f32 lane layout: A operand lane l holds A[l % 16][l / 16], B operand
// B[l / 16][l % 16]; accumulator register r of lane l is C[4 (l / 16) + r][l % 16].
// Padded k extent of the [col][k] LDS operand tile.  A wavefront's MFMA operand read takes 16
// columns x 4 k per instruction, served 32 lanes at a time (16 columns x 2 k): with a pitch of
// 2 mod 32 floats those 32 words fall in 32 different banks (a pitch of 4 mod 32 put columns c and
// c + 8 in one bank: two-way conflicts, a quarter of the update kernel's LDS cycles).
constexpr int KP = NB + 2;

// The trailing update forms each complex product from three real products (Gauss's method):
// with A = Ar + i Ai and B = Br + i Bi, t1 = Ar Br, t2 = Ai Bi and t3 = (Ar + Ai)(Br + Bi) give
// Re AB = t1 - t2 and Im AB = t3 - t1 - t2, three MFMAs where the direct form takes four.  Its
// error bound is normwise, not componentwise, in the imaginary part (Higham, "Stability of a
// method for multiplying complex matrices with three real matrix multiplications", SIMAX 13
// (1992)): a backward error of the factorisation of the same order, which refinement removes.
struct TileAcc {
    v4f t1[2][2], t2[2][2], t3[2][2];
};

__device__ __forceinline__ void acc_zero(TileAcc& c)
{
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            c.t1[p][q] = v4f{0.f, 0.f, 0.f, 0.f};
            c.t2[p][q] = v4f{0.f, 0.f, 0.f, 0.f};
            c.t3[p][q] = v4f{0.f, 0.f, 0.f, 0.f};
        }
}

// ---- panel: L21 = A21 U11^{-1} and U12 = L11^{-1} A12 on MFMA ---------------------------------
// Four workgroups per 64 x 64 panel tile, each a strip of 16 rows (tiles of the block column below
// the diagonal block: out = S U^{-1}, row i of out needs row i of S only) or of 16 columns (tiles of
// the block row to its right: out = L^{-1} S, column j of out needs column j of S only), so the
// strips are independent and the tile is updated in place.  Wavefront w computes one 16 x 16
// block of the strip with its MFMA operands loaded straight from memory (all 32 loads of a lane in
// flight at once; no LDS staging), and writes it after a barrier, once every wavefront of the
// workgroup has read the strip.
__global__ __launch_bounds__(256) void panel_trsm_kernel(const LuBatch batch)
{
    const int di = find_problem(batch.start, batch.count, blockIdx.x);
    const LuProblem& pb = batch.p[di];
    float2* __restrict__ A = pb.F;
    // the bordered matrix: every tile full
    const int64_t ld = pb.n, k0 = pb.k0;
    const int nrow_tiles = pb.tiles;
    const float2* __restrict__ Linv = pb.Linv;
    const float2* __restrict__ UinvT = pb.UinvT;
    const int bid = (int)((int64_t)blockIdx.x - batch.start[di]);
    const int tile = bid >> 2, strip = bid & 3;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
    const int li = l & 15, lk = l >> 4;
    const int64_t rest0 = k0 + NB;
    const bool rowtile = tile < nrow_tiles;
    const int64_t base = rest0 + (int64_t)(rowtile ? tile : tile - nrow_tiles) * NB;
    // operand k-slice s of this lane: A(i = li, k = 4 s + lk), B(k = 4 s + lk, j = li)
    float2 av[16], bv[16];
    if (rowtile) {
        // out[16 strip + i][16 w + j] = sum_k S[16 strip + i][k] U^{-1}[k][16 w + j],
        // U^{-1}[k][c] = UinvT[c + 64 k]
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            av[s] = A[(base + 16 * strip + li) + ld * (k0 + 4 * s + lk)];
            bv[s] = UinvT[(16 * w + li) + NB * (4 * s + lk)];
        }
    } else {
        // out[16 w + i][16 strip + j] = sum_k L^{-1}[16 w + i][k] S[k][16 strip + j]
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            av[s] = Linv[(16 * w + li) + NB * (4 * s + lk)];
            bv[s] = A[(k0 + 4 * s + lk) + ld * (base + 16 * strip + li)];
        }
    }
    Acc16 c;
    acc16_zero(c);
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        c.re = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s].x, bv[s].x, c.re, 0, 0, 0);
        c.re = __builtin_amdgcn_mfma_f32_16x16x4f32(-av[s].y, bv[s].y, c.re, 0, 0, 0);
        c.im = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s].x, bv[s].y, c.im, 0, 0, 0);
        c.im = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s].y, bv[s].x, c.im, 0, 0, 0);
    }
    __syncthreads(); // every wavefront has read the strip
    // rows 4 lk .. 4 lk + 3 of column li: 32 contiguous bytes, written as two float4
    float2* dst = rowtile ? &A[(base + 16 * strip + 4 * lk) + ld * (k0 + 16 * w + li)]
                          : &A[(k0 + 16 * w + 4 * lk) + ld * (base + 16 * strip + li)];
    reinterpret_cast<float4*>(dst)[0] = make_float4(c.re[0], c.im[0], c.re[1], c.im[1]);
    reinterpret_cast<float4*>(dst)[1] = make_float4(c.re[2], c.im[2], c.re[3], c.im[3]);
}

// ---- trailing update on MFMA, fused with the next diagonal block ---------------------------------
// C[i][j] -= sum_k L[i][k] U[k][j] for i, j in [rest0, n), k in [k0, k0 + K), K = kb = 64 .. 512
// (staged in chunks of 64), over 64 x 64 output tiles of the full grid, or of its cross (first
// block row and column) for LU_UPDATE_CROSS.  The workgroups are persistent (one per compute unit
// for the fused kernel, whose registers and LDS allow one; two for the unfused) and walk their tiles
// software-pipelined: while the products of one 64-deep chunk run, the operands of the next chunk
// are already in flight to registers, and a tile's output block is read when the tile starts.
// Workgroup 0 of a fused update (Linv_next given) takes tile 0 alone -- the next diagonal block --
// and goes on to factorise it, overlapping the rest of the update.
__device__ __forceinline__ void update_tile_xy(const LuProblem& pb, int idx, int& bx, int& by)
{
    if (pb.mode == LU_UPDATE_CROSS) {
        bx = idx < pb.tiles ? idx : 0;
        by = idx < pb.tiles ? 0 : idx - pb.tiles + 1;
    } else {
        bx = idx % pb.tiles;
        by = idx / pb.tiles;
    }
}

// Raw buffer access to the factor matrix: a descriptor per 64-column slab (base in SGPRs) and
// 32-bit byte offsets, the part common to a thread's loads in a VGPR and the rest in an SGPR, so a
// load costs no address arithmetic.  Offsets stay below 512 (ld + 1) bytes.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slab(const float2* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(base), (short)0, 0x7fffffff,
                                             0x00020000);
}
// four consecutive complex values (32 bytes) as two 16-byte accesses
__device__ __forceinline__ void slab_load4(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff,
                                           float2* f)
{
    const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    const auto b = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff + 16, 0);
    __builtin_memcpy(f, &a, 16);
    __builtin_memcpy(f + 2, &b, 16);
}

// The trailing update stages its first operand k-major ([k][row], pitch KPA): the 64 consecutive
// rows a wavefront writes and the 16 rows x 4 k a MFMA operand reads both fall in distinct banks.
constexpr int KPA = NB + 16;
constexpr int kUpdOpFloats = 2 * NB * KPA + 2 * NB * KP;
// Each wavefront's 32 x 32 output quarter passes through LDS on its way out, column-major with a
// padded column (conflict-free writes of the accumulator rows), so that every store instruction
// covers whole 128-byte column segments: the accumulator layout (four consecutive rows of one
// column per lane) would otherwise give every store only half of each segment.
constexpr int kOutPitch = 34;                            // complex values per LDS column
constexpr int kUpdOutFloats = 4 * 2 * 32 * kOutPitch;    // four wavefronts' quarters
constexpr int kUpdAllFloats = kUpdOpFloats + kUpdOutFloats;
constexpr int kUpdLdsFloats = kUpdAllFloats > (int)(sizeof(DiagLds) / sizeof(float))
                                  ? kUpdAllFloats
                                  : (int)(sizeof(DiagLds) / sizeof(float));

// c += Aop Bop with Aop in LDS as [k][row] (pitch KPA) and Bop as [col][k] (pitch KP), in the
// three-product form of TileAcc.  The MFMAs feeding one accumulator are twelve apart.
__device__ __forceinline__ void cgemm_64t(const float* __restrict__ Are,
                                          const float* __restrict__ Aim,
                                          const float* __restrict__ Bre,
                                          const float* __restrict__ Bim, TileAcc& c)
{
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const int li = lane % 16, lk = lane / 16;
#pragma unroll
    for (int ks = 0; ks < NB; ks += 4) {
        float ar[2], ai[2], br[2], bi[2], as[2], bs[2];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            ar[p] = Are[(ks + lk) * KPA + wr + 16 * p + li];
            ai[p] = Aim[(ks + lk) * KPA + wr + 16 * p + li];
            br[p] = Bre[(wc + 16 * p + li) * KP + ks + lk];
            bi[p] = Bim[(wc + 16 * p + li) * KP + ks + lk];
            as[p] = ar[p] + ai[p];
            bs[p] = br[p] + bi[p];
        }
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q)
                c.t1[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[p], br[q], c.t1[p][q], 0, 0, 0);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q)
                c.t2[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[p], bi[q], c.t2[p][q], 0, 0, 0);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q)
                c.t3[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(as[p], bs[q], c.t3[p][q], 0, 0, 0);
    }
}

template <bool FUSED>
__global__ __launch_bounds__(256, FUSED ? 1 : 2) void trailing_update_kernel(const LuBatch batch)
{
    const int di = find_problem(batch.start, batch.count, blockIdx.x);
    const LuProblem& pb = batch.p[di];
    float2* __restrict__ A = pb.F;
    // the bordered matrix: n is a multiple of 64 and every tile is full
    const int64_t ld = pb.n, k0 = pb.k0;
    const int K = pb.kb;
    const int w = (int)((int64_t)blockIdx.x - batch.start[di]);
    const bool fused = FUSED && pb.Linv_next != nullptr;
    int first = w, stride = pb.nwg;
    if (fused) {
        first = w;
        stride = w == 0 ? pb.ntile : pb.nwg - 1;
    }
    // the unfused update stores straight from the accumulators (no output quarters in LDS), so
    // that two workgroups fit on a compute unit
    __shared__ __align__(16) float lds[FUSED ? kUpdLdsFloats : kUpdOpFloats];
    float* Are = lds;
    float* Aim = lds + NB * KPA;
    float* Bre = lds + 2 * NB * KPA;
    float* Bim = lds + 2 * NB * KPA + NB * KP;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t rest0 = k0 + K;
    constexpr int kPer = NB * NB / 512;
    // Operand element pair i of a chunk: rows x = 2 (t % 32), x + 1 of column y = t / 32 + 8 i of
    // the slab (one 16-byte load per lane, coalesced down the columns), at byte offset
    // voff + i soff_step; staged as two 8-byte LDS writes per operand (real parts, imaginary parts).
    const unsigned voff = (unsigned)((2 * (t & 31) + ld * (t >> 5)) * 8);
    const unsigned soff_step = (unsigned)(8 * ld * 8);
    // output element (p, q, r) of this thread: row wr + 16 p + 4 (lane / 16) + r, column
    // wc + 16 q + lane % 16 of the tile
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const unsigned voffc = (unsigned)(((wr + 4 * (lane >> 4)) + ld * (wc + (lane & 15))) * 8);
    float4 va[kPer], vb[kPer];
    auto fetch = [&](int idx, int kc) __attribute__((always_inline)) {
        int bx, by;
        update_tile_xy(pb, idx, bx, by);
        const int64_t r0 = rest0 + (int64_t)bx * NB, c0 = rest0 + (int64_t)by * NB;
        const __amdgpu_buffer_rsrc_t ra = slab(A + r0 + ld * (k0 + kc));
        const __amdgpu_buffer_rsrc_t rb = slab(A + (k0 + kc) + ld * c0);
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const auto a = __builtin_amdgcn_raw_buffer_load_b128(ra, voff, i * soff_step, 0);
            const auto b = __builtin_amdgcn_raw_buffer_load_b128(rb, voff, i * soff_step, 0);
            __builtin_memcpy(&va[i], &a, 16);
            __builtin_memcpy(&vb[i], &b, 16);
        }
    };
    // A tile's output is stored during the next tile (after its operands are staged), so the
    // stores are never the youngest memory operations when a wait for loads comes.
    float2 res[2][2][4];
    __amdgpu_buffer_rsrc_t rres = slab(A);
    bool pending = false;
    // this wavefront's output quarter in LDS: column c (0..31) at c * kOutPitch complex values
    float2* outq = reinterpret_cast<float2*>(lds + (FUSED ? kUpdOpFloats : 0)) + wave * 32 * kOutPitch;
    auto store_res = [&]() __attribute__((always_inline)) {
        if constexpr (!FUSED) {
            // rows 16 p + 4 (lane / 16) .. + 3 of column 16 q + lane % 16: two 16-byte stores
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const unsigned off = (unsigned)(((wr + 16 * p + 4 * (lane >> 4)) +
                                                     ld * (wc + 16 * q + (lane & 15))) * 8);
                    const float4 v0 = make_float4(res[p][q][0].x, res[p][q][0].y, res[p][q][1].x,
                                                  res[p][q][1].y);
                    const float4 v1 = make_float4(res[p][q][2].x, res[p][q][2].y, res[p][q][3].x,
                                                  res[p][q][3].y);
                    decltype(__builtin_amdgcn_raw_buffer_load_b128(rres, 0u, 0u, 0)) a0, a1;
                    __builtin_memcpy(&a0, &v0, 16);
                    __builtin_memcpy(&a1, &v1, 16);
                    __builtin_amdgcn_raw_buffer_store_b128(a0, rres, off, 0u, 0);
                    __builtin_amdgcn_raw_buffer_store_b128(a1, rres, off + 16u, 0u, 0);
                }
        } else {
        // accumulator rows 16 p + 4 (lane / 16) .. + 3 of column 16 q + lane % 16: 32 bytes each
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                float2* d = outq + (16 * q + (lane & 15)) * kOutPitch + 16 * p + 4 * (lane >> 4);
                reinterpret_cast<float4*>(d)[0] = make_float4(res[p][q][0].x, res[p][q][0].y,
                                                              res[p][q][1].x, res[p][q][1].y);
                reinterpret_cast<float4*>(d)[1] = make_float4(res[p][q][2].x, res[p][q][2].y,
                                                              res[p][q][3].x, res[p][q][3].y);
            }
        // (the quarter is this wavefront's own: its LDS operations complete in order)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // store j: columns 4 j .. 4 j + 3, rows 2 (lane % 16), + 1 -- four whole 256-byte columns
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int c = 4 * j + (lane >> 4), r = 2 * (lane & 15);
            const float4 v = *reinterpret_cast<const float4*>(outq + c * kOutPitch + r);
            decltype(__builtin_amdgcn_raw_buffer_load_b128(rres, 0u, 0u, 0)) a;
            __builtin_memcpy(&a, &v, 16);
            __builtin_amdgcn_raw_buffer_store_b128(
                a, rres, (unsigned)(((wr + r) + ld * (wc + c)) * 8), 0u, 0);
        }
        }
    };
    int idx = first;
    if (idx < pb.ntile) fetch(idx, 0);
    while (idx < pb.ntile) {
        int bx, by;
        update_tile_xy(pb, idx, bx, by);
        const int64_t r0 = rest0 + (int64_t)bx * NB, c0 = rest0 + (int64_t)by * NB;
        const __amdgpu_buffer_rsrc_t rc = slab(A + r0 + ld * c0);
        float2 cv[2][2][4];
        TileAcc acc;
        acc_zero(acc);
        for (int kc = 0; kc < K; kc += NB) {
            __syncthreads(); // the previous chunk's operand reads are done
#pragma unroll
            for (int i = 0; i < kPer; ++i) {
                const int x = 2 * (t & 31), y = (t >> 5) + 8 * i;
                *reinterpret_cast<float2*>(Are + y * KPA + x) = make_float2(va[i].x, va[i].z);
                *reinterpret_cast<float2*>(Aim + y * KPA + x) = make_float2(va[i].y, va[i].w);
                *reinterpret_cast<float2*>(Bre + y * KP + x) = make_float2(vb[i].x, vb[i].z);
                *reinterpret_cast<float2*>(Bim + y * KP + x) = make_float2(vb[i].y, vb[i].w);
            }
            __syncthreads();
            // The previous tile's output goes out and this tile's output block comes in (first
            // chunk); then the next chunk (of this tile or of the next one; past the last, this
            // tile's chunk again, unused: one call site keeps the wait counts exact) goes in
            // flight under the products.
            if (kc == 0) {
                if (pending) store_res();
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q)
                        slab_load4(rc, voffc, (unsigned)((16 * p + ld * 16 * q) * 8), cv[p][q]);
            }
            const bool more = kc + NB < K;
            const int nidx = more ? idx : (idx + stride < pb.ntile ? idx + stride : idx);
            fetch(nidx, more ? kc + NB : 0);
            cgemm_64t(Are, Aim, Bre, Bim, acc);
        }
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    res[p][q][r] = make_float2(
                        cv[p][q][r].x - (acc.t1[p][q][r] - acc.t2[p][q][r]),
                        cv[p][q][r].y - (acc.t3[p][q][r] - acc.t1[p][q][r] - acc.t2[p][q][r]));
        rres = rc;
        pending = true;
        idx += stride;
    }
    if (pending) store_res();
    if constexpr (FUSED) {
        if (fused && w == 0) {
            __syncthreads(); // the tile's stores are visible to the workgroup; the LDS is free
            diag_factor(A, ld, rest0, NB, pb.Linv_next, pb.UinvT_next, pb.bad,
                        *reinterpret_cast<DiagLds*>(lds));
        }
    }
}

// ---- refinement units ------------------------------------------------------------------------------
// A unit is one problem's group of up to kMaxSolveRhs right-hand sides with its own workspace.
struct RefUnit {
    int64_t n;
    int nrhs;                      // <= kMaxSolveRhs
    int nblk;                      // 64-row blocks
    int chunks;                    // column chunks of the residual (<= kResidChunks)
    const double2* A;              // FP64 matrix
    int64_t lda;
    const double2* B;
    int64_t ldb;
    double2* X;
    int64_t ldx;
    double2* R;                    // n x nrhs residual (ld n)
    double2* part;                 // chunks x nrhs x n partial products
    float2* Y;                     // n x nrhs FP32 correction (ld n)
    double* scale;                 // nrhs
    unsigned long long* norms;     // kNormSlots: rmax | xmax | bmax | |A|_inf | pivot flag | time-out
    const float2* F;               // FP32 factors
    int64_t ldf;                   // their leading dimension
    const float2* Linv;
    const float2* UinvT;
    unsigned long long* ty;        // 2 n x nrhs tagged words handed between sweep blocks
    const double* anorm;           // |A|_inf of the factorised matrix
    int* luflags;                  // [0] bad pivot (factorisation), [1] sweep time-out
    const int* perm;               // static pivoting order: factor row i is row perm[i] of A
    const unsigned long long* gmax; // max |U|, max |L|, max |a_ij| (non-negative double bits)
};
constexpr int kNormSlots = 4 * kMaxSolveRhs;
static_assert(kNormSlots == kRefNormSlots, "read-back slots per unit");
constexpr int kNormA = 3 * kMaxSolveRhs, kNormBad = kNormA + 1, kNormTimeout = kNormA + 2;
constexpr int kNormGrowthU = kNormA + 3, kNormGrowthL = kNormA + 4, kNormAmax = kNormA + 5;

struct RefBatch {
    int count;
    unsigned long long epoch;
    int64_t start[kMaxLuBatch + 1];
    RefUnit u[kMaxLuBatch];
};

// ---- block-triangular sweeps -------------------------------------------------------------------
// Y (n x nrhs, column-major, FP32) is solved in place.  Lower: unit-lower L from the factors with
// inverted diagonal blocks Linv; block b waits for blocks < b.  Upper: blocks in reverse order, with
// the transposed inverses UinvT.
// Each workgroup owns one 64-row block of one unit; a unit's workgroups are consecutive and the
// first ones dispatched own the blocks without dependencies, so every wait is on a block that is
// already running or done.
// A finished block hands its rows on through TY, where every FP32 component sits in a 64-bit word
// beside the tag of the sweep that wrote it (2 epoch for the lower sweep, 2 epoch + 1 for the upper
// one).  Words are written and read as relaxed device-scope atomics, so a reader that sees the
// tag it waits for has the value of the same word: no fence and no flag are needed, and with them
// go the cache write-back and invalidation that a release / acquire pair costs on every hand-over.
// A thread gives up waiting after kSpinLimit polls in all and raises the unit's time-out flag, so
// the kernel always terminates (with wrong values, which the refinement then reports).
// Thread t = 4 row + kq owns row `row` of the block and the 16 columns 16 kq .. 16 kq + 15 of each
// 64-wide column block; the four threads of a row are adjacent lanes and combine their partial
// sums with lane exchanges.  The matrix block is loaded into registers before the wait (it does
// not depend on it), the dependency rows go through LDS (double-buffered: one barrier per
// dependency), and the block's own right-hand side is read at the start, so that after the last
// dependency arrives only one more barrier stands between it and the block's result.  The products
// are summed apart from the right-hand side and subtracted once at the end: folding them into an
// accumulator that starts at y_b loses accuracy to cancellation (measured: a first correction about
// 3.5 times less accurate on the kite problem).
__device__ __forceinline__ unsigned long long tagged(unsigned tag, float v)
{
    return ((unsigned long long)tag << 32) | __float_as_uint(v);
}

// LDS column of 64 complex rows with one complex of padding after every 16 rows, so that the four
// 16-row quarters read together by a wavefront fall in different banks
constexpr int kPadCol = 2 * NB + 8; // floats
__device__ __forceinline__ int pad_row(int r) { return 2 * r + 2 * (r >> 4); }

// Sum over each group of four adjacent lanes: lane ^ 1, then lane ^ 2, as DPP quad permutations
// (register-to-register; a shuffle would go through the LDS crossbar).
__device__ __forceinline__ float quad_add(float v)
{
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    return v;
}
__device__ __forceinline__ float2 quad_sum(float2 v) { return make_float2(quad_add(v.x), quad_add(v.y)); }

template <bool LOWER>
__global__ __launch_bounds__(256) void block_trsv_kernel(const RefBatch batch)
{
    __shared__ __align__(16) float ys[2][kMaxSolveRhs][kPadCol]; // dependency rows, (re, im)
    __shared__ __align__(16) float accs[kMaxSolveRhs][kPadCol];  // this block's reduced rhs
    __shared__ __align__(16) float yl[kMaxSolveRhs][kPadCol];    // this block's result
    const int di = find_problem(batch.start, batch.count, blockIdx.x);
    const RefUnit& U = batch.u[di];
    const float2* __restrict__ F = U.F;
    const int64_t ld = U.ldf, n = U.n;
    const int nblk = U.nblk, nrhs = U.nrhs;
    float2* __restrict__ Y = U.Y;
    const int64_t ldy = U.n;
    unsigned long long* TY = U.ty;
    const unsigned tag = (unsigned)(2 * batch.epoch + (LOWER ? 0 : 1));
    const int local = (int)((int64_t)blockIdx.x - batch.start[di]);
    const int t = threadIdx.x;
    const int b = LOWER ? local : nblk - 1 - local;
    const int64_t rb = (int64_t)b * NB;
    const int lenb = (int)min((int64_t)NB, n - rb);
    const int row = t >> 2, kq = t & 3, k0 = 16 * kq;
    const float2 zero = make_float2(0.f, 0.f);
    long long spins = 0;
    // this block's right-hand side (written by the previous kernel)
    float2 s[kMaxSolveRhs], yb[kMaxSolveRhs];
#pragma unroll
    for (int h = 0; h < kMaxSolveRhs; ++h) {
        yb[h] = load_or(h < nrhs && row < lenb, &Y[(rb + row) + ldy * h], Y, zero);
        s[h] = zero;
    }
    // the diagonal-block inverse, this thread's 16 columns of row `row` (the upper sweep's
    // inverses are stored transposed)
    float2 dv[16];
    const float2* D = (LOWER ? U.Linv : U.UinvT) + (int64_t)b * NB * NB;
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
        // (out-of-range entries are loaded from F[0] and zeroed where they are used: a select
        // here would wait for each load before the dependency rows are requested)
#pragma unroll
        for (int m = 0; m < 16; ++m)
            fv[m] = *(row < lenb && k0 + m < lenc ? &F[(rb + row) + ld * (cb + k0 + m)] : F);
        // rows cb .. cb + lenc of every column h: 2 lenc consecutive words at TY + 2 (cb + n h).
        // One wavefront fetches them all (requested at once; only those not yet written are
        // polled): with a single poller the wait ends on the first poll that finds the block
        // written, not on the last of several pollers.
        const int buf = q & 1;
        if (t < 64) {
            constexpr int kWords = 2 * kMaxSolveRhs;
            unsigned long long v[kWords];
#pragma unroll
            for (int i = 0; i < kWords; ++i) {
                const int w = t + 64 * i, h = w / (2 * NB), o = w % (2 * NB);
                v[i] = (unsigned long long)tag << 32;
                if (h < nrhs && o < 2 * lenc)
                    v[i] = __hip_atomic_load(TY + 2 * (cb + n * h) + o, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
            }
#pragma unroll
            for (int i = 0; i < kWords; ++i) {
                const int w = t + 64 * i, h = w / (2 * NB), o = w % (2 * NB);
                if (h < nrhs && o < 2 * lenc) {
                    unsigned long long* p = TY + 2 * (cb + n * h) + o;
                    while ((unsigned)(v[i] >> 32) != tag && spins <= kSpinLimit) {
                        __builtin_amdgcn_s_sleep(1);
                        if (++spins > kSpinLimit) atomicExch(&U.luflags[1], 1);
                        v[i] = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                // rows past the end of block c read as zero (their matrix entries are zero too)
                if (h < nrhs) ys[buf][h][pad_row(o >> 1) + (o & 1)] = __uint_as_float((unsigned)v[i]);
            }
        }
        __syncthreads();
        // straight-line per column: the 16 LDS reads issue together, then the products
#pragma unroll
        for (int h = 0; h < kMaxSolveRhs; ++h) {
            if (h < nrhs) {
                const float2* yc = reinterpret_cast<const float2*>(&ys[buf][h][0]);
                float2 yv[16];
#pragma unroll
                for (int m = 0; m < 16; ++m) yv[m] = yc[pad_row(k0 + m) >> 1];
#pragma unroll
                for (int m = 0; m < 16; ++m)
                    s[h] = cfma(s[h], csel(row < lenb && k0 + m < lenc, fv[m], zero), yv[m]);
            }
        }
    }
    // acc = y_b - sum_c F_bc y_c for row `row`, then y_b = Dinv_b acc
#pragma unroll
    for (int h = 0; h < kMaxSolveRhs; ++h) {
        const float2 p = quad_sum(s[h]);
        const float2 a = make_float2(yb[h].x - p.x, yb[h].y - p.y);
        if (kq == 0 && h < nrhs) {
            accs[h][pad_row(row)] = a.x;
            accs[h][pad_row(row) + 1] = a.y;
        }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < kMaxSolveRhs; ++h) {
        if (h < nrhs) {
            const float2* ac = reinterpret_cast<const float2*>(&accs[h][0]);
            float2 av[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) av[q] = ac[pad_row(k0 + q) >> 1];
            float2 v = zero;
#pragma unroll
            for (int q = 0; q < 16; ++q) v = cfma(v, dv[q], av[q]);
            v = quad_sum(v);
            // the plain copy serves the next kernel (it reads this block's rows only)
            if (kq == 0) {
                yl[h][pad_row(row)] = v.x;
                yl[h][pad_row(row) + 1] = v.y;
                if (row < lenb) Y[(rb + row) + ldy * h] = v;
            }
        }
    }
    // The tagged words serve the blocks of this sweep still waiting.  One wavefront writes them
    // all, back to back: a reader's poll then rarely finds only part of the block written (each
    // word it misses costs it another round trip to memory).
    __syncthreads();
    if (t < 64) {
        for (int h = 0; h < nrhs; ++h) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const int o = t + 64 * half, r = o >> 1;
                if (r < lenb)
                    __hip_atomic_store(TY + 2 * (rb + n * h) + o,
                                       tagged(tag, yl[h][pad_row(r) + (o & 1)]), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// ---- FP64 residual, scaling and updates ------------------------------------------------------------
// Row-parallel kernels: workgroup of 256 rows of one unit (start[] counts ceil(n / 256) per unit),
// blockIdx.y = right-hand side h or residual chunk c.


__device__ __forceinline__ double wave_max(double m)
{
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    return m;
}

// X = 0, R = B, bmax = rmax = max_i |B_i| (the residual of x0 = 0).
__global__ void ref_init_kernel(const RefBatch batch)
{
    const int di = find_problem(batch.start, batch.count, blockIdx.x);
    const RefUnit& U = batch.u[di];
    const int64_t i = ((int64_t)blockIdx.x - batch.start[di]) * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    if (h >= U.nrhs) return;
    double m = 0.0;
    if (i < U.n) {
        const double2 b = U.B[i + U.ldb * h];
        U.X[i + U.ldx * h] = make_double2(0.0, 0.0);
        U.R[i + U.n * h] = b;
        m = hypot(b.x, b.y);
        if (!(m == m)) m = INFINITY; // NaN poisons the norm
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) {
        atomic_max_nonneg(&U.norms[2 * kMaxSolveRhs + h], m);
        atomic_max_nonneg(&U.norms[h], m);
    }
}

// scale[h] = 1 / max(rmax[h], tiny); then rmax = xmax = 0 for the coming step.
__global__ void ref_scale_kernel(const RefBatch batch)
{
    const RefUnit& U = batch.u[blockIdx.x];
    const int h = threadIdx.x;
    if (h < U.nrhs) {
        const double m = __longlong_as_double((long long)U.norms[h]);
        U.scale[h] = (m > 1e-300 && m < INFINITY) ? 1.0 / m : 1.0;
        U.norms[h] = 0ull;
        U.norms[kMaxSolveRhs + h] = 0ull;
    }
}

// Y (FP32) = P R * scale[h]
__global__ void ref_demote_kernel(const RefBatch batch)
{
    const int di = find_problem(batch.start, batch.count, blockIdx.x);
    const RefUnit& U = batch.u[di];
    const int64_t i = ((int64_t)blockIdx.x - batch.start[di]) * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    if (h >= U.nrhs || i >= U.n) return;
    const double s = U.scale[h];
    const double2 r = U.R[(int64_t)U.perm[i] + U.n * h]; // P r: the factors are those of P A
    U.Y[i + U.n * h] = make_float2((float)(r.x * s), (float)(r.y * s));
}

// X += Y / scale[h]; xmax[h] = max_i |X_i|.
__global__ void ref_update_kernel(const RefBatch batch)
{
    const int di = find_problem(batch.start, batch.count, blockIdx.x);
    const RefUnit& U = batch.u[di];
    const int64_t i = ((int64_t)blockIdx.x - batch.start[di]) * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    if (h >= U.nrhs) return;
    double m = 0.0;
    if (i < U.n) {
        const double s = 1.0 / U.scale[h];
        const float2 y = U.Y[i + U.n * h];
        double2 x = U.X[i + U.ldx * h];
        x.x += (double)y.x * s;
        x.y += (double)y.y * s;
        U.X[i + U.ldx * h] = x;
        m = hypot(x.x, x.y);
        if (!(m == m)) m = INFINITY;
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&U.norms[kMaxSolveRhs + h], m);
}

// part[c][h][i] = sum_{j in chunk c} A[i][j] X[j][h]
__global__ void ref_residual_partial_kernel(const RefBatch batch)
{
    const int di = find_problem(batch.start, batch.count, blockIdx.x);
    const RefUnit& U = batch.u[di];
    const int64_t n = U.n;
    const int64_t i = ((int64_t)blockIdx.x - batch.start[di]) * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int nrhs = U.nrhs;
    if (i >= n || c >= U.chunks) return;
    const double2* __restrict__ A = U.A;
    const double2* __restrict__ X = U.X;
    const int64_t lda = U.lda, ldx = U.ldx;
    const int64_t j0 = n * c / U.chunks, j1 = n * (c + 1) / U.chunks;
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
        if (h < nrhs) U.part[((int64_t)c * nrhs + h) * n + i] = s[h];
}

// R = B - sum_c part[c]; rmax[h] = max_i |R_i|.
__global__ void ref_residual_reduce_kernel(const RefBatch batch)
{
    const int di = find_problem(batch.start, batch.count, blockIdx.x);
    const RefUnit& U = batch.u[di];
    const int64_t n = U.n;
    const int64_t i = ((int64_t)blockIdx.x - batch.start[di]) * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    if (h >= U.nrhs) return;
    double m = 0.0;
    if (i < n) {
        double2 s = U.B[i + U.ldb * h];
        for (int c = 0; c < U.chunks; ++c) {
            const double2 p = U.part[((int64_t)c * U.nrhs + h) * n + i];
            s.x -= p.x;
            s.y -= p.y;
        }
        U.R[i + n * h] = s;
        m = hypot(s.x, s.y);
        if (!(m == m)) m = INFINITY;
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&U.norms[h], m);
}

// Copies |A|_inf, the factorisation's pivot / time-out flags and its element growth into the norm
// slots, so that one read-back carries everything the host needs.
__global__ void ref_flags_kernel(const RefBatch batch)
{
    const RefUnit& U = batch.u[blockIdx.x];
    if (threadIdx.x == 0) {
        U.norms[kNormA] = (unsigned long long)__double_as_longlong(*U.anorm);
        U.norms[kNormBad] = (unsigned long long)U.luflags[0];
        U.norms[kNormTimeout] = (unsigned long long)U.luflags[1];
        U.norms[kNormGrowthU] = U.gmax[0];
        U.norms[kNormGrowthL] = U.gmax[1];
        U.norms[kNormAmax] = U.gmax[2];
    }
}

__global__ void zero_u64_kernel(unsigned long long* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ---- many right-hand sides: column-parallel helpers of the library-GEMM refinement ------------------
// m[h] = max_i |V[i + ld h]| (atomic max on the bit pattern of a non-negative double)
__global__ void colmax_kernel(int64_t n, int nrhs, const double2* __restrict__ V, int64_t ld,
                              unsigned long long* m)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    double v = 0.0;
    if (i < n && h < nrhs) {
        const double2 x = V[i + ld * h];
        v = hypot(x.x, x.y);
        if (!(v == v)) v = INFINITY;
    }
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0 && h < nrhs) atomic_max_nonneg(&m[h], v);
}

// R = B, X = 0
__global__ void wide_init_kernel(int64_t n, int nrhs, const double2* __restrict__ B, int64_t ldb,
                                 double2* __restrict__ R, double2* __restrict__ X, int64_t ldx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    if (i >= n) return;
    R[i + n * h] = B[i + ldb * h];
    X[i + ldx * h] = make_double2(0.0, 0.0);
}

// Y = P R / max(rmax, tiny) in FP32; s[h] records the scale.
__global__ void wide_demote_kernel(int64_t n, int nrhs, const double2* __restrict__ R,
                                   const int* __restrict__ perm,
                                   const unsigned long long* __restrict__ rmax,
                                   double* __restrict__ scale, float2* __restrict__ Y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    const double m = __longlong_as_double((long long)rmax[h]);
    const double sc = (m > 1e-300 && m < INFINITY) ? 1.0 / m : 1.0;
    if (i == 0) scale[h] = sc;
    if (i >= n) return;
    const double2 r = R[(int64_t)perm[i] + n * h];
    Y[i + n * h] = make_float2((float)(r.x * sc), (float)(r.y * sc));
}

// X += Y / scale; R = B (the residual is then formed in place by the GEMM)
__global__ void wide_update_kernel(int64_t n, int nrhs, const float2* __restrict__ Y,
                                   const double* __restrict__ scale, double2* __restrict__ X,
                                   int64_t ldx, const double2* __restrict__ B, int64_t ldb,
                                   double2* __restrict__ R)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int h = blockIdx.y;
    if (i >= n) return;
    const double s = 1.0 / scale[h];
    const float2 y = Y[i + n * h];
    double2 x = X[i + ldx * h];
    x.x += (double)y.x * s;
    x.y += (double)y.y * s;
    X[i + ldx * h] = x;
    R[i + n * h] = B[i + ldb * h];
}

} // namespace

// ================================================================================================
// host side
// ================================================================================================
struct MixedState {
};

void bem2d_internal_mixed_release(bem2d_ctx* c)
{
    delete c->mixed;
    c->mixed = nullptr;
}

void* bem2d_internal_dalloc(DevArena* a, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (a) return a->alloc(bytes);
    void* p = nullptr;
    return hipMalloc(&p, bytes) == hipSuccess ? p : nullptr;
}

void bem2d_internal_dfree(DevArena* a, void* p)
{
    if (!p) return;
    if (a)
        a->release(p);
    else
        (void)hipFree(p);
}

// Workspace of the single-problem refinement, kept with the factors so that repeated refine calls
// allocate nothing: one device block handed out by a bump arena (reset per call; a factorisation
// is refined by one call at a time) and a pinned read-back buffer.
struct RefCache final : DevArena {
    char* base = nullptr;
    size_t cap = 0, off = 0;
    int nrhs = 0;                      // right-hand sides the workspace is sized for
    unsigned long long* pinned = nullptr;
    void* alloc(size_t bytes) override
    {
        const size_t a = (off + 255) & ~(size_t)255;
        if (a + bytes > cap) return nullptr;
        off = a + bytes;
        return base + a;
    }
    void release(void*) override {}
};

struct bem2d_lu32 {
    bem2d_ctx* ctx = nullptr;
    DevArena* arena = nullptr;
    int64_t n = 0;
    int nblk = 0;
    int step = -1;                // next panel step; nblk - 1 when factorised; -1 before
    int pend = 0;                 // panels of the current group applied to the cross only so far
    int group = 1;                // panels per trailing update in the current group (1, 2, 4, 8)
    bool factored = false;
    int cus = 256;                // compute units of the device (persistent update workgroups)
    int64_t np = 0;               // n rounded up to whole 64-wide blocks
    float2* F = nullptr;          // np x np factors (A bordered by the identity), ld = np
    float2* Linv = nullptr;       // nblk x 64 x 64
    float2* UinvT = nullptr;      // nblk x (U_bb^{-1})^T
    int* dflag = nullptr;         // [0] bad pivot, [1] sweep time-out
    double* anorm = nullptr;      // |A|_inf
    double* rowpart = nullptr;    // (np / kCastCols) x n partial row sums of |a_ij|
    int* perm = nullptr;          // np: static pivoting order (factor row i = row perm[i] of A)
    int* pivwork = nullptr;       // colarg | freepos | freerow, n each
    unsigned long long* claim = nullptr; // n claims of the rows
    unsigned long long* gmax = nullptr;  // max |U|, max |L|, max |a_ij|
    RefCache* rcache = nullptr;   // refinement workspace of bem2d_lu32_refine
    // hipGraph of the whole factorisation launch sequence for one matrix buffer (A, lda): built on
    // the second factorisation of the same buffer, replayed from then on
    const void* seen_A = nullptr;
    int64_t seen_lda = 0;
    hipGraphExec_t graph = nullptr;
    const void* graph_A = nullptr;
    int64_t graph_lda = 0;
    int graph_pair_tiles = 0;     // the panel-grouping choices the graph was captured with
    int graph_quad_tiles = 0;
    int graph_oct_tiles = 0;
    int graph_split_tiles = 0;    // and the diagonal-split choice
    int64_t graph_replays = 0;
};

size_t bem2d_internal_lu32_bytes(int64_t n)
{
    const int64_t nblk = (n + NB - 1) / NB, np = nblk * NB;
    return sizeof(float2) * (size_t)(np * np + 2 * nblk * NB * NB) +
           sizeof(double) * ((np / kCastCols) * n + 1) + sizeof(int) * (np + 3 * n) +
           sizeof(unsigned long long) * (n + 3) + 12 * 256;
}

extern "C" void bem2d_lu32_destroy(bem2d_lu32* f)
{
    if (!f) return;
    (void)hipSetDevice(f->ctx->device);
    if (!f->arena) (void)hipStreamSynchronize(f->ctx->stream);
    bem2d_internal_dfree(f->arena, f->F);
    bem2d_internal_dfree(f->arena, f->Linv);
    bem2d_internal_dfree(f->arena, f->UinvT);
    bem2d_internal_dfree(f->arena, f->dflag);
    bem2d_internal_dfree(f->arena, f->anorm);
    bem2d_internal_dfree(f->arena, f->rowpart);
    bem2d_internal_dfree(f->arena, f->perm);
    bem2d_internal_dfree(f->arena, f->pivwork);
    bem2d_internal_dfree(f->arena, f->claim);
    bem2d_internal_dfree(f->arena, f->gmax);
    if (f->graph) (void)hipGraphExecDestroy(f->graph);
    if (f->rcache) {
        (void)hipFree(f->rcache->base);
        (void)hipHostFree(f->rcache->pinned);
        delete f->rcache;
    }
    delete f;
}

int bem2d_internal_lu32_create(bem2d_ctx* c, int64_t n, DevArena* arena, bem2d_lu32** out)
{
    if (!c || !out || n < 1) return FAIL(BEM2D_ERR_ARG, "bem2d_lu32_create: bad argument");
    HIPCHK(hipSetDevice(c->device));
    bem2d_lu32* f = new bem2d_lu32();
    f->ctx = c;
    f->arena = arena;
    f->n = n;
    f->nblk = (int)((n + NB - 1) / NB);
    f->np = (int64_t)f->nblk * NB;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess &&
        cus > 0)
        f->cus = cus;
    const size_t inv_bytes = sizeof(float2) * (size_t)f->nblk * NB * NB;
    f->F = (float2*)bem2d_internal_dalloc(arena, sizeof(float2) * f->np * f->np);
    f->Linv = (float2*)bem2d_internal_dalloc(arena, inv_bytes);
    f->UinvT = (float2*)bem2d_internal_dalloc(arena, inv_bytes);
    f->dflag = (int*)bem2d_internal_dalloc(arena, sizeof(int) * 2);
    f->anorm = (double*)bem2d_internal_dalloc(arena, sizeof(double));
    f->rowpart = (double*)bem2d_internal_dalloc(arena, sizeof(double) * (f->np / kCastCols) * n);
    f->perm = (int*)bem2d_internal_dalloc(arena, sizeof(int) * f->np);
    f->pivwork = (int*)bem2d_internal_dalloc(arena, sizeof(int) * 3 * n);
    f->claim = (unsigned long long*)bem2d_internal_dalloc(arena, sizeof(unsigned long long) * n);
    f->gmax = (unsigned long long*)bem2d_internal_dalloc(arena, sizeof(unsigned long long) * 3);
    if (!f->F || !f->Linv || !f->UinvT || !f->dflag || !f->anorm || !f->rowpart || !f->perm ||
        !f->pivwork || !f->claim || !f->gmax) {
        bem2d_lu32_destroy(f);
        return FAIL(BEM2D_ERR_NOMEM, "bem2d_lu32_create: device allocation of %lld bytes failed",
                    (long long)(8 * n * n));
    }
    *out = f;
    return BEM2D_OK;
}

extern "C" int bem2d_lu32_create(bem2d_ctx* c, int64_t n, bem2d_lu32** out)
{
    return bem2d_internal_lu32_create(c, n, nullptr, out);
}

extern "C" int64_t bem2d_lu32_size(const bem2d_lu32* f) { return f ? f->n : -1; }

Lu32View bem2d_internal_lu32_view(const bem2d_lu32* f)
{
    return Lu32View{f->F, f->n, f->np, f->perm, f->anorm};
}

static LuProblem lu_problem(bem2d_lu32* f, int b)
{
    LuProblem p{};
    p.F = f->F;
    p.n = f->np; // the bordered matrix: every tile full
    p.k0 = (int64_t)b * NB;
    p.kb = NB;
    const int64_t rest = f->np - p.k0 - p.kb;
    p.tiles = (int)((rest + NB - 1) / NB);
    p.Linv = f->Linv + (size_t)b * NB * NB;
    p.UinvT = f->UinvT + (size_t)b * NB * NB;
    p.bad = f->dflag;
    return p;
}

// Element growth of finished factors (read by the refinement's first read-back).

// Static pivoting order, cast, |A|_inf and the factorisation of diagonal block 0 for count
// problems.
// Element growth of finished factors (the cast's row-sum buffer, free by then, holds the partials).
static void launch_growth(hipStream_t s, const bem2d_lu32* f)
{
    const int64_t nwg = ((f->np + kGrowthRows - 1) / kGrowthRows) * ((f->n + kGrowthCols - 1) / kGrowthCols);
    float* part = reinterpret_cast<float*>(f->rowpart);
    growth_kernel<<<(unsigned)nwg, 256, 0, s>>>(f->n, f->np, f->F, part);
    growth_reduce_kernel<<<1, 1024, 0, s>>>(nwg, part, f->gmax);
}

int bem2d_internal_lu_begin(hipStream_t s, bem2d_lu32* const* lus, const void* const* A64,
                            const int64_t* lda, int count)
{
    for (int c0 = 0; c0 < count; c0 += kMaxLuBatch) {
        BeginBatch bt{};
        bt.count = std::min(kMaxLuBatch, count - c0);
        for (int d = 0; d < bt.count; ++d) {
            bem2d_lu32* f = lus[c0 + d];
            BeginProblem& p = bt.p[d];
            p.n = f->n;
            p.np = f->np;
            p.lda = lda[c0 + d];
            p.A = (const double2*)A64[c0 + d];
            p.perm = f->perm;
            p.F = f->F;
            p.part = f->rowpart;
            p.pivwork = f->pivwork;
            p.claim = f->claim;
            p.gmax = f->gmax;
            p.anorm = (unsigned long long*)f->anorm;
            p.dflag = f->dflag;
            f->step = 0;
            f->pend = 0;
            f->factored = false;
        }
        // workgroups per problem of each kernel -> start[]
        auto launch = [&](auto per_problem) {
            bt.start[0] = 0;
            for (int d = 0; d < bt.count; ++d) bt.start[d + 1] = bt.start[d] + per_problem(bt.p[d]);
            return (unsigned)bt.start[bt.count];
        };
        const unsigned g_rows = launch([](const BeginProblem& p) { return (p.n + 255) / 256; });
        begin_init_kernel<<<g_rows, 256, 0, s>>>(bt);
        // static pivoting order, then the cast of the rows in that order
        const unsigned g_cols = launch([](const BeginProblem& p) { return p.n; });
        pivot_colmax_kernel<<<g_cols, 256, 0, s>>>(bt);
        pivot_assign_kernel<<<(unsigned)bt.count, 1024, 0, s>>>(bt);
        const unsigned g_cast =
            launch([](const BeginProblem& p) { return ((p.np + 255) / 256) * (p.np / kCastCols); });
        cast_kernel<<<g_cast, 256, 0, s>>>(bt);
        launch([](const BeginProblem& p) { return (p.n + 255) / 256; });
        rowsum_max_kernel<<<g_rows, 256, 0, s>>>(bt);
    }
    for (int c0 = 0; c0 < count; c0 += kMaxLuBatch) {
        LuBatch bt{};
        bt.count = std::min(kMaxLuBatch, count - c0);
        for (int d = 0; d < bt.count; ++d) {
            bem2d_lu32* f = lus[c0 + d];
            LuProblem p = lu_problem(f, 0);
            p.Linv_next = f->Linv;
            p.UinvT_next = f->UinvT;
            bt.p[d] = p;
            bt.start[d + 1] = d + 1;
        }
        diag_lu_kernel<<<bt.count, 256, 0, s>>>(bt);
    }
    for (int d = 0; d < count; ++d)
        if (lus[d]->nblk == 1) {
            lus[d]->factored = true;
            launch_growth(s, lus[d]);
        }
    if (hipGetLastError() != hipSuccess) return FAIL(BEM2D_ERR_DEVICE, "LU launch failed");
    return BEM2D_OK;
}

// One panel step for each of count problems (each at its own step): the panel solves of block b,
// then the trailing update, whose first workgroup also factorises diagonal block b + 1 (or, for
// large trailing matrices, the update followed by that factorisation on its own).
int bem2d_internal_lu_advance(hipStream_t s, bem2d_lu32* const* lus, int count, int split_tiles)
{
    std::vector<bem2d_lu32*> act;
    for (int d = 0; d < count; ++d)
        if (lus[d]->step >= 0 && lus[d]->step < lus[d]->nblk - 1) act.push_back(lus[d]);
    for (size_t c0 = 0; c0 < act.size(); c0 += kMaxLuBatch) {
        LuBatch tr{}, up{}, dg{};
        const int cnt = (int)std::min<size_t>(kMaxLuBatch, act.size() - c0);
        tr.count = up.count = cnt;
        // Trailing matrices of at least tune.lu_split_tiles tiles per side are updated without the
        // fused diagonal block: the update then fits two workgroups per compute unit, and
        // diag_lu_kernel factorises the block after it.  One launch shape for the whole batch:
        // split only when every problem is that large.
        const int split_from = split_tiles > 0 ? split_tiles : act[c0]->ctx->tune.lu_split_tiles;
        bool split = true;
        for (int d = 0; d < cnt; ++d)
            if (lu_problem(act[c0 + d], act[c0 + d]->step).tiles < split_from) split = false;
        for (int d = 0; d < cnt; ++d) {
            bem2d_lu32* f = act[c0 + d];
            const int b = f->step;
            LuProblem p = lu_problem(f, b);
            p.mode = LU_UPDATE_FULL;
            p.Linv_next = f->Linv + (size_t)(b + 1) * NB * NB;
            p.UinvT_next = f->UinvT + (size_t)(b + 1) * NB * NB;
            tr.p[d] = p;
            tr.start[d + 1] = tr.start[d] + 4 * 2 * p.tiles; // four strips per panel tile
            // Panels group in twos, fours or eights: each step of a group but the last updates only
            // the cross (the next panel, which the next step factorises) with every panel of the
            // group so far, and the last step applies the whole group to the rest at depth 128 to
            // 512, so the trailing matrix is read and written once per group.
            LuProblem u = p;
            int64_t blocks;
            int pend = f->pend;
            if (pend == 0) {
                const TuneState& tn = f->ctx->tune;
                f->group = 1;
                if (b + 8 <= f->nblk - 1 && u.tiles >= tn.lu_oct_tiles) f->group = 8;
                else if (b + 4 <= f->nblk - 1 && u.tiles >= tn.lu_quad_tiles) f->group = 4;
                else if (b + 2 <= f->nblk - 1 && u.tiles >= tn.lu_pair_tiles) f->group = 2;
            }
            u.k0 = (int64_t)(b - pend) * NB;
            u.kb = (pend + 1) * NB;
            if (pend + 1 < f->group) {
                u.mode = LU_UPDATE_CROSS;
                f->pend = pend + 1;
                blocks = 2 * (int64_t)u.tiles - 1;
            } else {
                f->pend = 0;
                blocks = (int64_t)u.tiles * u.tiles;
            }
            // persistent workgroups: the compute units shared among the problems of the launch
            // (a fused update needs two: workgroup 0 holds the next diagonal block alone)
            // (two per compute unit for the unfused update, which fits two)
            const int64_t share = std::max<int64_t>(2, (split ? 2 : 1) * act[c0]->cus / cnt);
            u.ntile = (int)blocks;
            u.nwg = (int)std::min<int64_t>(blocks, share);
            if (split) { // the update alone; diag_lu_kernel factorises block b + 1 after it
                u.Linv_next = nullptr;
                u.UinvT_next = nullptr;
                LuProblem q = lu_problem(f, b + 1);
                q.Linv_next = p.Linv_next;
                q.UinvT_next = p.UinvT_next;
                dg.p[d] = q;
                dg.start[d + 1] = d + 1;
            }
            up.p[d] = u;
            up.start[d + 1] = up.start[d] + u.nwg;
        }
        dg.count = split ? cnt : 0;
        if (up.start[cnt] > (int64_t)INT32_MAX) return FAIL(BEM2D_ERR_UNSUPPORTED, "LU grid too large");
        panel_trsm_kernel<<<(unsigned)tr.start[cnt], 256, 0, s>>>(tr);
        // HIP events around the update when the context profiles (not inside a graph capture)
        bem2d_ctx* pc = act[c0]->ctx;
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        const bool prof = pc->profiling && hipStreamIsCapturing(s, &cap) == hipSuccess &&
                          cap == hipStreamCaptureStatusNone;
        if (prof) bem2d_internal_prof_begin(pc, BEM2D_PROF_LU_UPDATE, s);
        if (split)
            trailing_update_kernel<false><<<(unsigned)up.start[cnt], 256, 0, s>>>(up);
        else
            trailing_update_kernel<true><<<(unsigned)up.start[cnt], 256, 0, s>>>(up);
        if (prof) bem2d_internal_prof_end(pc, BEM2D_PROF_LU_UPDATE, s);
        if (split) diag_lu_kernel<<<dg.count, 256, 0, s>>>(dg);
    }
    if (hipGetLastError() != hipSuccess) return FAIL(BEM2D_ERR_DEVICE, "LU launch failed");
    for (bem2d_lu32* f : act) {
        f->step++;
        if (f->step == f->nblk - 1) {
            f->factored = true;
            launch_growth(s, f);
        }
    }
    if (hipGetLastError() != hipSuccess) return FAIL(BEM2D_ERR_DEVICE, "LU launch failed");
    return BEM2D_OK;
}

bool bem2d_internal_lu_done(const bem2d_lu32* f) { return f->factored; }
int64_t bem2d_internal_lu_n(const bem2d_lu32* f) { return f->n; }

static bool lu_graphs_enabled()
{
    static const bool on = [] {
        const char* e = std::getenv("BEM2D_LU_GRAPH");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Enqueues the factorisation's launch sequence (cast and |A|_inf, first diagonal block, then two
// kernels per panel step) on s; every launch is decided on the host from n alone.
static int lu_factor_launches(hipStream_t s, bem2d_lu32* f, const void* A64, int64_t lda)
{
    int rc = bem2d_internal_lu_begin(s, &f, &A64, &lda, 1);
    while (!rc && !f->factored) rc = bem2d_internal_lu_advance(s, &f, 1);
    return rc;
}

// A factorisation is ~2 launches per 64 columns (124 at n = 4000) with nothing decided on the
// device, so a repeated factorisation of the same buffer -- new contents, same address and lda, as
// in a frequency sweep or a benchmark loop -- replays one hipGraph of the sequence instead of
// launching it kernel by kernel.  The graph is captured on the second factorisation of a buffer
// (a one-shot solve never pays for the capture) and holds kernel nodes only.  The capture records
// on a private stream of the context (the context's own stream may be the null stream, which
// cannot capture); the graph then runs on the context's stream.  Set BEM2D_LU_GRAPH=0 to launch
// eagerly always.
extern "C" int bem2d_lu32_factor(bem2d_ctx* c, bem2d_lu32* f, const void* A64, int64_t lda)
{
    if (!c || !f || !A64 || lda < f->n) return FAIL(BEM2D_ERR_ARG, "bem2d_lu32_factor: bad argument");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const bool graphs = lu_graphs_enabled();
    const bool current = f->graph && f->graph_A == A64 && f->graph_lda == lda &&
                         f->graph_pair_tiles == c->tune.lu_pair_tiles &&
                         f->graph_quad_tiles == c->tune.lu_quad_tiles &&
                         f->graph_oct_tiles == c->tune.lu_oct_tiles &&
                         f->graph_split_tiles == c->tune.lu_split_tiles;
    if (graphs && !current && f->seen_A == A64 && f->seen_lda == lda) {
        if (!c->capture) HIPCHK(hipStreamCreateWithFlags(&c->capture, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c->capture, hipStreamCaptureModeThreadLocal));
        int rc = lu_factor_launches(c->capture, f, A64, lda);
        const hipError_t e = hipStreamEndCapture(c->capture, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : FAIL(BEM2D_ERR_DEVICE, "LU graph capture failed: %s", hipGetErrorString(e));
        }
        if (f->graph) (void)hipGraphExecDestroy(f->graph);
        f->graph = nullptr;
        const hipError_t ei = hipGraphInstantiate(&f->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            f->graph = nullptr;
            return FAIL(BEM2D_ERR_DEVICE, "LU graph instantiation failed: %s", hipGetErrorString(ei));
        }
        f->graph_A = A64;
        f->graph_lda = lda;
        f->graph_pair_tiles = c->tune.lu_pair_tiles;
        f->graph_quad_tiles = c->tune.lu_quad_tiles;
        f->graph_oct_tiles = c->tune.lu_oct_tiles;
        f->graph_split_tiles = c->tune.lu_split_tiles;
    }
    bem2d_internal_prof_begin(c, BEM2D_PROF_LU);
    int rc = BEM2D_OK;
    if (graphs && f->graph && f->graph_A == A64 && f->graph_lda == lda &&
        f->graph_pair_tiles == c->tune.lu_pair_tiles &&
        f->graph_quad_tiles == c->tune.lu_quad_tiles &&
        f->graph_oct_tiles == c->tune.lu_oct_tiles &&
        f->graph_split_tiles == c->tune.lu_split_tiles) {
        if (hipGraphLaunch(f->graph, s) != hipSuccess) rc = FAIL(BEM2D_ERR_DEVICE, "LU graph launch failed");
        // the host-side state the launch sequence leaves behind
        f->step = f->nblk - 1;
        f->pend = 0;
        f->factored = true;
        f->graph_replays++;
    } else {
        rc = lu_factor_launches(s, f, A64, lda);
        f->seen_A = A64;
        f->seen_lda = lda;
    }
    bem2d_internal_prof_end(c, BEM2D_PROF_LU);
    return rc;
}

extern "C" int64_t bem2d_lu32_graph_replays(const bem2d_lu32* f) { return f ? f->graph_replays : -1; }

// ---- refinement jobs --------------------------------------------------------------------------
// A job refines x = A^{-1} b for several problems at once: every problem's right-hand sides are
// split into units of <= kMaxSolveRhs columns, and each step is one batched launch per kernel over
// the units still running.  Progress is asynchronous: step() enqueues a step and an asynchronous
// copy of every unit's norms into pinned host memory; ready() polls the step's event; collect()
// reads the norms and retires the units that met the tolerance.  The first step() enqueues the
// initial residual and the first two corrections together: x = 0 meets no tolerance unless b = 0,
// and one correction from the FP32 factors leaves a backward error near the FP32 rounding level
// (~1e-7), far above any FP64 target, so waiting for those norms would only add round trips
// through the host.  (A unit limited to one step gets one.)
struct RefJob {
    bem2d_ctx* ctx = nullptr;
    DevArena* arena = nullptr;
    std::vector<RefUnit> units;
    std::vector<int> unit_spec;           // spec index of each unit
    std::vector<char> active;             // unit still refining
    std::vector<int> iters;               // steps taken per unit
    std::vector<double> omega;            // backward error per unit
    std::vector<double> bnorm;            // |b_h|_inf per unit and column (kMaxSolveRhs each)
    std::vector<double> tol;              // backward-error target per unit
    std::vector<int> maxit;               // step limit per unit
    std::vector<double> prev_omega;       // backward error of the previous read-back per unit
    std::vector<double> growth;           // element growth of the unit's factors
    std::vector<char> bad;                // zero / non-finite pivot or growth above the limit
    std::vector<char> stalled;            // the backward error stopped falling
    bool own_pinned = true;
    std::vector<void*> blocks;            // device allocations of the workspace
    unsigned long long* dnorms = nullptr; // units x kNormSlots
    unsigned long long* hnorms = nullptr; // pinned mirror
    unsigned long long epoch = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
    bool started = false;
    bool bnorm_read = false;              // |b| per column taken from the first read-back
    int last_steps = 1;                   // correction steps behind the pending read-back
    int nspec = 0;
};

static int ref_chunks(int64_t n)
{
    const int64_t rows = (n + 255) / 256;
    return (int)std::max<int64_t>(1, std::min<int64_t>(kResidChunks, (2048 + rows - 1) / rows));
}

size_t bem2d_internal_ref_bytes(int64_t n, int nrhs)
{
    const int64_t units = (nrhs + kMaxSolveRhs - 1) / kMaxSolveRhs;
    const size_t per = sizeof(double2) * n * kMaxSolveRhs * (1 + ref_chunks(n)) +
                       sizeof(float2) * n * kMaxSolveRhs + sizeof(double) * kMaxSolveRhs +
                       sizeof(unsigned long long) * (2 * n * kMaxSolveRhs + kNormSlots) + 6 * 256;
    return units * per;
}

void bem2d_internal_ref_destroy(RefJob* j)
{
    if (!j) return;
    if (j->ev) (void)hipEventDestroy(j->ev);
    if (j->hnorms && j->own_pinned) (void)hipHostFree(j->hnorms);
    for (void* p : j->blocks) bem2d_internal_dfree(j->arena, p);
    delete j;
}

int bem2d_internal_ref_units(int nrhs) { return (nrhs + kMaxSolveRhs - 1) / kMaxSolveRhs; }

int bem2d_internal_ref_create(bem2d_ctx* c, DevArena* arena, const RefSpec* specs, int count,
                              unsigned long long* pinned, unsigned long long* dnorms,
                              RefJob** out)
{
    RefJob* j = new RefJob();
    j->ctx = c;
    j->arena = arena;
    j->nspec = count;
    auto fail = [&](int code, const char* what) {
        bem2d_internal_ref_destroy(j);
        return FAIL(code, "refinement: %s", what);
    };
    for (int sidx = 0; sidx < count; ++sidx) {
        const RefSpec& sp = specs[sidx];
        bem2d_lu32* f = sp.lu;
        const int64_t n = f->n;
        if (sp.lda < n || sp.ldb < n || sp.ldx < n || sp.nrhs < 1)
            return fail(BEM2D_ERR_ARG, "bad leading dimension or column count");
        for (int h0 = 0; h0 < sp.nrhs; h0 += kMaxSolveRhs) {
            RefUnit u{};
            u.n = n;
            u.nrhs = std::min(kMaxSolveRhs, sp.nrhs - h0);
            u.nblk = f->nblk;
            u.chunks = ref_chunks(n);
            u.A = (const double2*)sp.A64;
            u.lda = sp.lda;
            u.B = (const double2*)sp.B + sp.ldb * h0;
            u.ldb = sp.ldb;
            u.X = (double2*)sp.X + sp.ldx * h0;
            u.ldx = sp.ldx;
            u.F = f->F;
            u.ldf = f->np;
            u.Linv = f->Linv;
            u.UinvT = f->UinvT;
            u.anorm = f->anorm;
            u.luflags = f->dflag;
            u.perm = f->perm;
            u.gmax = f->gmax;
            const size_t bytes = sizeof(double2) * n * kMaxSolveRhs * (1 + u.chunks) +
                                 sizeof(float2) * n * kMaxSolveRhs +
                                 sizeof(double) * kMaxSolveRhs +
                                 sizeof(unsigned long long) * 2 * n * kMaxSolveRhs + 4 * 256;
            char* base = (char*)bem2d_internal_dalloc(sp.ws ? sp.ws : arena, bytes);
            if (!base) return fail(BEM2D_ERR_NOMEM, "workspace allocation failed");
            if (!sp.ws) j->blocks.push_back(base); // a spec's own arena is released by its owner
            auto carve = [&](size_t b) {
                char* p = base;
                base += (b + 255) & ~(size_t)255;
                return p;
            };
            u.R = (double2*)carve(sizeof(double2) * n * kMaxSolveRhs);
            u.part = (double2*)carve(sizeof(double2) * n * kMaxSolveRhs * u.chunks);
            u.Y = (float2*)carve(sizeof(float2) * n * kMaxSolveRhs);
            u.scale = (double*)carve(sizeof(double) * kMaxSolveRhs);
            u.ty = (unsigned long long*)carve(sizeof(unsigned long long) * 2 * n * kMaxSolveRhs);
            j->units.push_back(u);
            j->unit_spec.push_back(sidx);
            j->tol.push_back(sp.tol);
            j->maxit.push_back(sp.maxit);
        }
    }
    const size_t nu = j->units.size();
    if (dnorms) {
        j->dnorms = dnorms;
    } else {
        j->dnorms = (unsigned long long*)bem2d_internal_dalloc(arena, sizeof(unsigned long long) *
                                                                          kNormSlots * nu);
        if (!j->dnorms) return fail(BEM2D_ERR_NOMEM, "workspace allocation failed");
        j->blocks.push_back(j->dnorms);
    }
    for (size_t u = 0; u < nu; ++u) j->units[u].norms = j->dnorms + kNormSlots * u;
    if (pinned) {
        j->hnorms = pinned;
        j->own_pinned = false;
    } else if (hipHostMalloc((void**)&j->hnorms, sizeof(unsigned long long) * kNormSlots * nu) !=
               hipSuccess) {
        j->hnorms = nullptr;
        return fail(BEM2D_ERR_NOMEM, "pinned buffer allocation failed");
    }
    if (hipEventCreateWithFlags(&j->ev, hipEventDisableTiming) != hipSuccess)
        return fail(BEM2D_ERR_NOMEM, "event creation failed");
    j->active.assign(nu, 1);
    j->iters.assign(nu, 0);
    j->omega.assign(nu, INFINITY);
    j->prev_omega.assign(nu, INFINITY);
    j->growth.assign(nu, 0.0);
    j->bad.assign(nu, 0);
    j->stalled.assign(nu, 0);
    j->bnorm.assign(nu * kMaxSolveRhs, 0.0);
    *out = j;
    return BEM2D_OK;
}

// Row-parallel grid over the listed units (start[] in workgroups of 256 rows).
static RefBatch ref_batch(const RefJob* j, const std::vector<int>& ids, size_t c0, int per_unit)
{
    RefBatch bt{};
    bt.count = (int)std::min<size_t>(kMaxLuBatch, ids.size() - c0);
    bt.epoch = j->epoch;
    for (int d = 0; d < bt.count; ++d) {
        const RefUnit& u = j->units[ids[c0 + d]];
        bt.u[d] = u;
        int64_t w = per_unit;
        if (per_unit == 0) w = (u.n + 255) / 256;   // rows
        if (per_unit == -1) w = u.nblk;              // sweep blocks
        bt.start[d + 1] = bt.start[d] + w;
    }
    return bt;
}

// Enqueues one refinement step (the first one preceded by the initial residual) for every active
// unit, then the asynchronous copy of the norms and the job's event.
int bem2d_internal_ref_step(RefJob* j, hipStream_t s)
{
    std::vector<int> ids;
    for (size_t u = 0; u < j->units.size(); ++u)
        if (j->active[u]) ids.push_back((int)u);
    int nsteps = 1;
    if (!j->started) {
        nsteps = 2;
        for (int u : ids) nsteps = std::min(nsteps, j->maxit[u]);
        const size_t nu = j->units.size();
        zero_u64_kernel<<<(unsigned)((kNormSlots * nu + 255) / 256), 256, 0, s>>>(
            j->dnorms, (int64_t)(kNormSlots * nu));
        for (size_t u = 0; u < nu; ++u)
            zero_u64_kernel<<<(unsigned)((2 * j->units[u].n * kMaxSolveRhs + 255) / 256), 256, 0,
                              s>>>(j->units[u].ty, 2 * j->units[u].n * kMaxSolveRhs);
        for (size_t c0 = 0; c0 < ids.size(); c0 += kMaxLuBatch) {
            RefBatch bt = ref_batch(j, ids, c0, 0);
            ref_init_kernel<<<dim3((unsigned)bt.start[bt.count], kMaxSolveRhs), 256, 0, s>>>(bt);
        }
        j->started = true;
    }
    for (int st = 0; st < nsteps; ++st) {
        j->epoch++;
        for (size_t c0 = 0; c0 < ids.size(); c0 += kMaxLuBatch) {
            RefBatch one = ref_batch(j, ids, c0, 1);
            RefBatch rows = ref_batch(j, ids, c0, 0);
            RefBatch sweep = ref_batch(j, ids, c0, -1);
            RefBatch chunks = rows;
            const unsigned rgrid = (unsigned)rows.start[rows.count];
            ref_scale_kernel<<<one.count, 64, 0, s>>>(one);
            ref_demote_kernel<<<dim3(rgrid, kMaxSolveRhs), 256, 0, s>>>(rows);
            block_trsv_kernel<true><<<(unsigned)sweep.start[sweep.count], 256, 0, s>>>(sweep);
            block_trsv_kernel<false><<<(unsigned)sweep.start[sweep.count], 256, 0, s>>>(sweep);
            ref_update_kernel<<<dim3(rgrid, kMaxSolveRhs), 256, 0, s>>>(rows);
            ref_residual_partial_kernel<<<dim3(rgrid, kResidChunks), 256, 0, s>>>(chunks);
            ref_residual_reduce_kernel<<<dim3(rgrid, kMaxSolveRhs), 256, 0, s>>>(rows);
            ref_flags_kernel<<<one.count, 64, 0, s>>>(one);
        }
    }
    if (hipMemcpyAsync(j->hnorms, j->dnorms, sizeof(unsigned long long) * kNormSlots * j->units.size(),
                       hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipEventRecord(j->ev, s) != hipSuccess || hipGetLastError() != hipSuccess)
        return FAIL(BEM2D_ERR_DEVICE, "refinement launch failed");
    j->pending = true;
    j->last_steps = nsteps;
    return BEM2D_OK;
}

bool bem2d_internal_ref_ready(RefJob* j)
{
    return !j->pending || hipEventQuery(j->ev) == hipSuccess;
}

static double u64_as_double(unsigned long long v)
{
    double d;
    std::memcpy(&d, &v, sizeof d);
    return d;
}

// Element growth of FP32 factors without interchanges: max(max|U| / max|a_ij|, max|L|).  Above
// kGrowthLimit the FP32 backward error (about n u32 growth |A|) leaves refinement nothing to
// converge from, and the solve goes to the pivoted FP64 LU at once.
constexpr double kGrowthLimit = 1e4;
constexpr double kStallRatio = 0.5;
static double factor_growth(double gu, double gl, double amax)
{
    if (!(gu == gu) || !(gl == gl) || !(amax == amax)) return INFINITY;
    const double ru = amax > 0 ? gu / amax : (gu > 0 ? INFINITY : 0.0);
    return std::max(ru, gl);
}

// Reads the norms of the last step (waits for it) and retires the units that met tol, that failed
// (a bad pivot or a sweep time-out) or that used maxit steps.
// Returns BEM2D_ERR_DEVICE on a sweep time-out.
int bem2d_internal_ref_collect(RefJob* j, bool* all_done)
{
    if (j->pending) {
        if (hipEventSynchronize(j->ev) != hipSuccess)
            return FAIL(BEM2D_ERR_DEVICE, "refinement step failed");
        j->pending = false;
    }
    int rc = BEM2D_OK;
    bool done = true;
    for (size_t u = 0; u < j->units.size(); ++u) {
        if (!j->active[u]) continue;
        const unsigned long long* nm = j->hnorms + kNormSlots * u;
        const RefUnit& U = j->units[u];
        const double g = factor_growth(u64_as_double(nm[kNormGrowthU]),
                                       u64_as_double(nm[kNormGrowthL]),
                                       u64_as_double(nm[kNormAmax]));
        j->growth[u] = g;
        if (nm[kNormTimeout] || nm[kNormBad] || !(g <= kGrowthLimit)) {
            j->active[u] = 0;
            j->omega[u] = INFINITY;
            j->bad[u] = !nm[kNormTimeout];
            if (nm[kNormTimeout]) rc = BEM2D_ERR_DEVICE;
            continue;
        }
        if (!j->bnorm_read) // the initial residual's slots, untouched by the correction steps
            for (int h = 0; h < U.nrhs; ++h)
                j->bnorm[kMaxSolveRhs * u + h] = u64_as_double(nm[2 * kMaxSolveRhs + h]);
        j->iters[u] += j->last_steps;
        const double anorm = u64_as_double(nm[kNormA]);
        double w = 0.0;
        for (int h = 0; h < U.nrhs; ++h) {
            const double rn = u64_as_double(nm[h]);
            const double xn = u64_as_double(nm[kMaxSolveRhs + h]);
            const double den = anorm * xn + j->bnorm[kMaxSolveRhs * u + h];
            const double wh = den > 0 ? rn / den : (rn > 0 ? INFINITY : 0.0);
            w = std::max(w, std::isfinite(wh) ? wh : INFINITY);
        }
        j->omega[u] = w;
        if (w <= j->tol[u] || j->iters[u] >= j->maxit[u]) {
            j->active[u] = 0;
        } else if (j->iters[u] >= 3 && !(w <= kStallRatio * j->prev_omega[u])) {
            // plain refinement converges linearly; a step that no longer halves the backward
            // error marks a factorisation too far from A for it (GMRES-IR or FP64 take over)
            j->active[u] = 0;
            j->stalled[u] = 1;
        }
        j->prev_omega[u] = w;
        if (j->active[u]) done = false;
    }
    j->bnorm_read = true;
    *all_done = done;
    return rc;
}

bool bem2d_internal_ref_spec_active(const RefJob* j, int sidx)
{
    for (size_t u = 0; u < j->units.size(); ++u)
        if (j->unit_spec[u] == sidx && j->active[u]) return true;
    return false;
}

// Report of spec sidx: worst backward error and most steps over its units.
void bem2d_internal_ref_report(const RefJob* j, int sidx, bem2d_ir_report* r)
{
    double tol = INFINITY;
    double worst = 0.0, growth = 0.0;
    int it = 0;
    double anorm = 0.0;
    bool any = false;
    for (size_t u = 0; u < j->units.size(); ++u) {
        if (j->unit_spec[u] != sidx) continue;
        worst = any ? std::max(worst, j->omega[u]) : j->omega[u];
        tol = std::min(tol, j->tol[u]);
        it = std::max(it, j->iters[u]);
        anorm = u64_as_double(j->hnorms[kNormSlots * u + kNormA]);
        growth = std::max(growth, j->growth[u]);
        any = true;
    }
    r->iterations = it;
    r->backward_error = worst;
    r->converged = worst <= tol;
    r->matrix_norm_inf = anorm;
    r->growth = growth;
}

// A zero or non-finite pivot, or element growth above kGrowthLimit, in the factors of spec sidx.
bool bem2d_internal_ref_bad_pivot(const RefJob* j, int sidx)
{
    for (size_t u = 0; u < j->units.size(); ++u)
        if (j->unit_spec[u] == sidx && j->bad[u]) return true;
    return false;
}

// Synchronous refinement of count problems (the engine drives the same steps asynchronously).
int bem2d_internal_refine_sync(hipStream_t s, RefJob* j)
{
    bool done = false;
    int rc = BEM2D_OK;
    while (!rc && !done) {
        rc = bem2d_internal_ref_step(j, s); // correction steps (the first after the initial residual)
        if (!rc) rc = bem2d_internal_ref_collect(j, &done);
    }
    return rc;
}

// Refinement of many right-hand sides at once (nrhs >= kWideRhs): the triangular solves are
// rocBLAS ctrsm on the FP32 factors and the FP64 residual R = B - A X is one zgemm, so A64 and
// the factors are read once per step for all columns instead of once per group of kMaxSolveRhs.
constexpr int kWideRhs = 64;

static int refine_wide(bem2d_ctx* c, bem2d_lu32* f, const double2* A, int64_t lda,
                       const double2* B, int64_t ldb, double2* X, int64_t ldx, int nrhs,
                       double tol, int maxit, bem2d_ir_report* rep)
{
    const int64_t n = f->n;
    hipStream_t s = c->stream;
    double2* R = nullptr;
    float2* Y = nullptr;
    double* scale = nullptr;
    unsigned long long* nm = nullptr; // rmax | xmax | bmax
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);
        (void)hipFree(R);
        (void)hipFree(Y);
        (void)hipFree(scale);
        (void)hipFree(nm);
    };
    if (hipMalloc(&R, sizeof(double2) * n * nrhs) != hipSuccess ||
        hipMalloc(&Y, sizeof(float2) * n * nrhs) != hipSuccess ||
        hipMalloc(&scale, sizeof(double) * nrhs) != hipSuccess ||
        hipMalloc(&nm, sizeof(unsigned long long) * 3 * nrhs) != hipSuccess) {
        cleanup();
        return FAIL(BEM2D_ERR_NOMEM, "refinement workspace allocation failed");
    }
    std::vector<unsigned long long> h(3 * nrhs);
    int dflag[2] = {0, 0};
    double anorm = 0.0;
    double gm[3] = {0.0, 0.0, 0.0};
    const dim3 g((unsigned)((n + 255) / 256), (unsigned)nrhs);
    wide_init_kernel<<<g, 256, 0, s>>>(n, nrhs, B, ldb, R, X, ldx);
    zero_u64_kernel<<<(unsigned)((3 * nrhs + 255) / 256), 256, 0, s>>>(nm, 3 * nrhs);
    colmax_kernel<<<g, 256, 0, s>>>(n, nrhs, R, n, nm + 2 * nrhs);
    (void)hipMemcpyAsync(&anorm, f->anorm, sizeof(double), hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(dflag, f->dflag, sizeof(int) * 2, hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(gm, f->gmax, sizeof(double) * 3, hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(h.data(), nm, sizeof(unsigned long long) * 3 * nrhs,
                         hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) {
        cleanup();
        return FAIL(BEM2D_ERR_DEVICE, "refinement start failed");
    }
    bem2d_ir_report r{};
    r.matrix_norm_inf = anorm;
    r.growth = factor_growth(gm[0], gm[1], gm[2]);
    if (dflag[0] || !(r.growth <= kGrowthLimit)) {
        cleanup();
        r.backward_error = INFINITY;
        if (rep) *rep = r;
        return FAIL(BEM2D_ERR_NOT_CONVERGED, dflag[0] ? "FP32 LU met a zero or non-finite pivot"
                                                      : "FP32 LU element growth %.3g above %.0g",
                    r.growth, kGrowthLimit);
    }
    std::vector<double> bnorm(nrhs);
    for (int q = 0; q < nrhs; ++q) {
        std::memcpy(&bnorm[q], &h[2 * nrhs + q], sizeof(double));
    }
    // rmax = bmax for the first scaling
    (void)hipMemcpyAsync(nm, nm + 2 * nrhs, sizeof(unsigned long long) * nrhs,
                         hipMemcpyDeviceToDevice, s);
    const rocblas_float_complex one32 = {1.f, 0.f};
    const rocblas_double_complex mone = {-1.0, 0.0}, one = {1.0, 0.0};
    double omega = INFINITY, prev = INFINITY;
    int it = 0;
    int rc = BEM2D_OK;
    for (; it < maxit; ++it) {
        wide_demote_kernel<<<g, 256, 0, s>>>(n, nrhs, R, f->perm, nm, scale, Y);
        rocblas_status st = rocblas_ctrsm(c->blas, rocblas_side_left, rocblas_fill_lower,
                                          rocblas_operation_none, rocblas_diagonal_unit,
                                          (rocblas_int)n, (rocblas_int)nrhs, &one32,
                                          (const rocblas_float_complex*)f->F, (rocblas_int)f->np,
                                          (rocblas_float_complex*)Y, (rocblas_int)n);
        if (st == rocblas_status_success)
            st = rocblas_ctrsm(c->blas, rocblas_side_left, rocblas_fill_upper,
                               rocblas_operation_none, rocblas_diagonal_non_unit, (rocblas_int)n,
                               (rocblas_int)nrhs, &one32, (const rocblas_float_complex*)f->F,
                               (rocblas_int)f->np, (rocblas_float_complex*)Y, (rocblas_int)n);
        wide_update_kernel<<<g, 256, 0, s>>>(n, nrhs, Y, scale, X, ldx, B, ldb, R);
        if (st == rocblas_status_success)
            st = rocblas_zgemm(c->blas, rocblas_operation_none, rocblas_operation_none,
                               (rocblas_int)n, (rocblas_int)nrhs, (rocblas_int)n, &mone,
                               (const rocblas_double_complex*)A, (rocblas_int)lda,
                               (const rocblas_double_complex*)X, (rocblas_int)ldx, &one,
                               (rocblas_double_complex*)R, (rocblas_int)n);
        if (st != rocblas_status_success) {
            rc = FAIL(BEM2D_ERR_DEVICE, "rocBLAS status %d in refinement", (int)st);
            break;
        }
        zero_u64_kernel<<<(unsigned)((2 * nrhs + 255) / 256), 256, 0, s>>>(nm, 2 * nrhs);
        colmax_kernel<<<g, 256, 0, s>>>(n, nrhs, R, n, nm);
        colmax_kernel<<<g, 256, 0, s>>>(n, nrhs, X, ldx, nm + nrhs);
        (void)hipMemcpyAsync(h.data(), nm, sizeof(unsigned long long) * 2 * nrhs,
                             hipMemcpyDeviceToHost, s);
        if (hipStreamSynchronize(s) != hipSuccess) {
            rc = FAIL(BEM2D_ERR_DEVICE, "refinement step failed");
            break;
        }
        omega = 0.0;
        for (int q = 0; q < nrhs; ++q) {
            double rn, xn;
            std::memcpy(&rn, &h[q], sizeof rn);
            std::memcpy(&xn, &h[nrhs + q], sizeof xn);
            const double den = anorm * xn + bnorm[q];
            const double w = den > 0 ? rn / den : (rn > 0 ? INFINITY : 0.0);
            omega = std::max(omega, std::isfinite(w) ? w : INFINITY);
        }
        if (omega <= tol || (it >= 2 && !(omega <= kStallRatio * prev))) {
            ++it; // converged, or stalled (GMRES-IR takes over below)
            break;
        }
        prev = omega;
    }
    cleanup();
    r.iterations = it;
    r.backward_error = omega;
    r.converged = omega <= tol;
    if (!rc && !r.converged)
        rc = bem2d_internal_gmres_ir(c, f, A, lda, B, ldb, X, ldx, nrhs, tol, &r);
    if (rep) *rep = r;
    if (rc) return rc;
    if (!r.converged)
        return FAIL(BEM2D_ERR_NOT_CONVERGED, "refinement reached backward error %.3e > %.3e",
                    r.backward_error, tol);
    return BEM2D_OK;
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
    if (nrhs >= kWideRhs && nrhs <= 65535 && f->n <= 2000000000 && lda <= 2000000000 &&
        ldx <= 2000000000)
        return refine_wide(c, f, (const double2*)A64, lda, (const double2*)B, ldb, (double2*)X, ldx,
                           nrhs, tol, maxit, rep);
    // the workspace is kept with the factors and grows to the largest nrhs seen
    RefCache*& rcache = f->rcache;
    if (!rcache || rcache->nrhs < nrhs) {
        if (rcache) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(rcache->base);
            (void)hipHostFree(rcache->pinned);
            delete rcache;
            rcache = nullptr;
        }
        RefCache* rc_new = new RefCache();
        const int units = bem2d_internal_ref_units(nrhs);
        rc_new->cap = bem2d_internal_ref_bytes(f->n, nrhs) +
                      sizeof(unsigned long long) * kRefNormSlots * units + 1024;
        rc_new->nrhs = nrhs;
        if (hipMalloc(&rc_new->base, rc_new->cap) != hipSuccess ||
            hipHostMalloc((void**)&rc_new->pinned,
                          sizeof(unsigned long long) * kRefNormSlots * units) != hipSuccess) {
            (void)hipFree(rc_new->base);
            delete rc_new;
            return FAIL(BEM2D_ERR_NOMEM, "refinement workspace allocation failed");
        }
        rcache = rc_new;
    }
    rcache->off = 0;
    RefSpec sp{f, A64, lda, B, ldb, X, ldx, nrhs, tol, maxit, rcache};
    RefJob* j = nullptr;
    int rc = bem2d_internal_ref_create(c, rcache, &sp, 1, rcache->pinned, nullptr, &j);
    if (rc) return rc;
    rc = bem2d_internal_refine_sync(c->stream, j);
    bem2d_ir_report r{};
    bem2d_internal_ref_report(j, 0, &r);
    const bool bad = bem2d_internal_ref_bad_pivot(j, 0);
    bem2d_internal_ref_destroy(j);
    if (rep) *rep = r;
    if (rc == BEM2D_ERR_DEVICE) return FAIL(BEM2D_ERR_DEVICE, "triangular sweep wait timed out");
    if (bad) {
        if (rep) {
            rep->converged = 0;
            rep->iterations = 0;
            rep->backward_error = INFINITY;
        }
        return FAIL(BEM2D_ERR_NOT_CONVERGED, "FP32 LU met a zero or non-finite pivot");
    }
    if (rc) return rc;
    if (!r.converged) { // plain refinement stalled or ran out of steps: GMRES-IR from here
        rc = bem2d_internal_gmres_ir(c, f, A64, lda, B, ldb, X, ldx, nrhs, tol, &r);
        if (rep) *rep = r;
        if (rc) return rc;
    }
    if (!r.converged)
        return FAIL(BEM2D_ERR_NOT_CONVERGED, "refinement reached backward error %.3e > %.3e",
                    r.backward_error, tol);
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

// ================================================================================================
// mixed solve without allocation
// ================================================================================================
// bem2d_solve_mixed_static: the factorisation and refinement of bem2d_solve_mixed with every
// device buffer carved from one file-scope array, for callers that cannot allocate (code under
// graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its address on each device
// is taken once, with that device current, when the first context on it is created; host code uses
// that address, never the array's name (which on the host is only a stand-in).  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU race, and each call clears what must start
// at zero (row claims, norms, flags, the tagged words of the sweeps) before use.
constexpr int64_t kStaticMaxN = 1024;
constexpr int kStaticMaxRhs = kMaxSolveRhs;
constexpr size_t kStaticScratchBytes = 20u << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kStaticMaxDevices = 64;
std::mutex g_static_mutex;
void* g_static_ptr[kStaticMaxDevices] = {};
}

// Takes the device address of `scratch` on device `device` (current) once; later calls return it.
int bem2d_internal_static_scratch(int device, void** out)
{
    if (device < 0 || device >= kStaticMaxDevices)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "static scratch: device %d beyond %d", device,
                    kStaticMaxDevices);
    std::lock_guard<std::mutex> lock(g_static_mutex);
    if (!g_static_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return FAIL(BEM2D_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        g_static_ptr[device] = p;
    }
    if (out) *out = g_static_ptr[device];
    return BEM2D_OK;
}

extern "C" int bem2d_solve_mixed_static(bem2d_ctx* c, int64_t n, int64_t nrhs, const void* A64,
                                        int64_t lda, void* B, int64_t ldb, double tol,
                                        bem2d_ir_report* rep)
{
    if (!c || !A64 || !B || n < 1 || nrhs < 1 || lda < n || ldb < n)
        return FAIL(BEM2D_ERR_ARG, "bem2d_solve_mixed_static: bad argument");
    if (n > kStaticMaxN || nrhs > kStaticMaxRhs)
        return FAIL(BEM2D_ERR_ARG,
                    "bem2d_solve_mixed_static: n = %lld, nrhs = %lld beyond the fixed workspace "
                    "(n <= %lld, nrhs <= %d)", (long long)n, (long long)nrhs,
                    (long long)kStaticMaxN, kStaticMaxRhs);
    HIPCHK(hipSetDevice(c->device));
    void* base = nullptr;
    if (int rc = bem2d_internal_static_scratch(c->device, &base)) return rc;
    if (!(tol > 0)) tol = 1e-12;
    hipStream_t s = c->stream;
    // the factor object lives on the stack; its buffers are carved from the array
    bem2d_lu32 f;
    f.ctx = c;
    f.n = n;
    f.nblk = (int)((n + NB - 1) / NB);
    f.np = (int64_t)f.nblk * NB;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess &&
        cus > 0)
        f.cus = cus;
    char* p = (char*)base;
    size_t used = 0;
    auto carve = [&](size_t bytes) {
        void* q = p + used;
        used += (bytes + 255) & ~(size_t)255;
        return q;
    };
    const int64_t np = f.np;
    f.F = (float2*)carve(sizeof(float2) * np * np);
    f.Linv = (float2*)carve(sizeof(float2) * (size_t)f.nblk * NB * NB);
    f.UinvT = (float2*)carve(sizeof(float2) * (size_t)f.nblk * NB * NB);
    f.dflag = (int*)carve(sizeof(int) * 2);
    f.anorm = (double*)carve(sizeof(double));
    f.rowpart = (double*)carve(sizeof(double) * (np / kCastCols) * n);
    f.perm = (int*)carve(sizeof(int) * np);
    f.pivwork = (int*)carve(sizeof(int) * 3 * n);
    f.claim = (unsigned long long*)carve(sizeof(unsigned long long) * n);
    f.gmax = (unsigned long long*)carve(sizeof(unsigned long long) * 3);
    RefUnit u{};
    u.n = n;
    u.nrhs = (int)nrhs;
    u.nblk = f.nblk;
    u.chunks = ref_chunks(n);
    u.A = (const double2*)A64;
    u.lda = lda;
    u.B = (const double2*)B;
    u.ldb = ldb;
    u.X = (double2*)carve(sizeof(double2) * n * nrhs);
    u.ldx = n;
    u.R = (double2*)carve(sizeof(double2) * n * kMaxSolveRhs);
    u.part = (double2*)carve(sizeof(double2) * n * kMaxSolveRhs * u.chunks);
    u.Y = (float2*)carve(sizeof(float2) * n * kMaxSolveRhs);
    u.scale = (double*)carve(sizeof(double) * kMaxSolveRhs);
    u.ty = (unsigned long long*)carve(sizeof(unsigned long long) * 2 * n * kMaxSolveRhs);
    u.norms = (unsigned long long*)carve(sizeof(unsigned long long) * kNormSlots);
    u.F = f.F;
    u.ldf = np;
    u.Linv = f.Linv;
    u.UinvT = f.UinvT;
    u.anorm = f.anorm;
    u.luflags = f.dflag;
    u.perm = f.perm;
    u.gmax = f.gmax;
    if (used > kStaticScratchBytes)
        return FAIL(BEM2D_ERR_ARG, "bem2d_solve_mixed_static: workspace of %zu bytes exceeds %zu",
                    used, kStaticScratchBytes);
    // factorisation: the launch sequence of bem2d_lu32_factor, eagerly
    bem2d_lu32* fp = &f;
    int rc = bem2d_internal_lu_begin(s, &fp, &A64, &lda, 1);
    while (!rc && !f.factored) rc = bem2d_internal_lu_advance(s, &fp, 1);
    if (rc) return rc;
    // refinement: the step sequence of bem2d_lu32_refine for one group of right-hand sides (the
    // initial residual and two corrections, then one correction per read-back), so the arithmetic
    // -- and the solution -- is that of bem2d_solve_mixed
    RefBatch one{}, rows{}, sweep{};
    one.count = rows.count = sweep.count = 1;
    one.u[0] = rows.u[0] = sweep.u[0] = u;
    one.start[1] = 1;
    rows.start[1] = (n + 255) / 256;
    sweep.start[1] = f.nblk;
    const unsigned rgrid = (unsigned)rows.start[1];
    zero_u64_kernel<<<1, 64, 0, s>>>(u.norms, kNormSlots);
    zero_u64_kernel<<<(unsigned)((2 * n * kMaxSolveRhs + 255) / 256), 256, 0, s>>>(
        u.ty, 2 * n * kMaxSolveRhs);
    ref_init_kernel<<<dim3(rgrid, kMaxSolveRhs), 256, 0, s>>>(rows);
    constexpr int kMaxit = 30;
    unsigned long long nm[kNormSlots];
    double bnorm[kMaxSolveRhs] = {};
    double omega = INFINITY, prev = INFINITY, growth = 0.0, anorm = 0.0;
    int it = 0;
    bool bad = false, first = true;
    while (it < kMaxit) {
        const int nsteps = first ? 2 : 1;
        for (int st = 0; st < nsteps; ++st) {
            const unsigned long long epoch = (unsigned long long)(it + st + 1);
            one.epoch = rows.epoch = sweep.epoch = epoch;
            ref_scale_kernel<<<1, 64, 0, s>>>(one);
            ref_demote_kernel<<<dim3(rgrid, kMaxSolveRhs), 256, 0, s>>>(rows);
            block_trsv_kernel<true><<<(unsigned)sweep.start[1], 256, 0, s>>>(sweep);
            block_trsv_kernel<false><<<(unsigned)sweep.start[1], 256, 0, s>>>(sweep);
            ref_update_kernel<<<dim3(rgrid, kMaxSolveRhs), 256, 0, s>>>(rows);
            ref_residual_partial_kernel<<<dim3(rgrid, kResidChunks), 256, 0, s>>>(rows);
            ref_residual_reduce_kernel<<<dim3(rgrid, kMaxSolveRhs), 256, 0, s>>>(rows);
            ref_flags_kernel<<<1, 64, 0, s>>>(one);
        }
        it += nsteps;
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(nm, u.norms, sizeof nm, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        growth = factor_growth(u64_as_double(nm[kNormGrowthU]), u64_as_double(nm[kNormGrowthL]),
                               u64_as_double(nm[kNormAmax]));
        anorm = u64_as_double(nm[kNormA]);
        if (nm[kNormTimeout]) return FAIL(BEM2D_ERR_DEVICE, "triangular sweep wait timed out");
        if (nm[kNormBad] || !(growth <= kGrowthLimit)) {
            bad = true;
            omega = INFINITY;
            break;
        }
        if (first)
            for (int h = 0; h < nrhs; ++h) bnorm[h] = u64_as_double(nm[2 * kMaxSolveRhs + h]);
        first = false;
        omega = 0.0;
        for (int h = 0; h < nrhs; ++h) {
            const double rn = u64_as_double(nm[h]), xn = u64_as_double(nm[kMaxSolveRhs + h]);
            const double den = anorm * xn + bnorm[h];
            const double wh = den > 0 ? rn / den : (rn > 0 ? INFINITY : 0.0);
            omega = std::max(omega, std::isfinite(wh) ? wh : INFINITY);
        }
        if (omega <= tol || (it >= 3 && !(omega <= kStallRatio * prev))) break;
        prev = omega;
    }
    if (rep) {
        *rep = bem2d_ir_report{};
        rep->iterations = bad ? 0 : it;
        rep->converged = omega <= tol;
        rep->backward_error = omega;
        rep->matrix_norm_inf = anorm;
        rep->growth = growth;
    }
    if (!(omega <= tol))
        return FAIL(BEM2D_ERR_NOT_CONVERGED,
                    bad ? "FP32 factors unusable (zero pivot or element growth %.3g)"
                        : "refinement reached backward error %.3e > tol",
                    bad ? growth : omega);
    HIPCHK(hipMemcpy2DAsync(B, sizeof(double2) * ldb, u.X, sizeof(double2) * n,
                            sizeof(double2) * n, nrhs, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return BEM2D_OK;
}
