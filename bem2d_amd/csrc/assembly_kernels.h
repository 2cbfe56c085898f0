// Device kernels of the FP64 hot path: Gauss-node tables, regular-pair and singular-pair Galerkin
// assembly, plane-wave right-hand sides and far field.
#pragma once

#include <hip/hip_runtime.h>

#include "geometry.h"
#include "hankel.h"

namespace bem2d_dev {

// Node tables of a mesh: five planes of nq*n doubles, plane[a*n + e] for Gauss node a of element e.
enum { NODE_X = 0, NODE_Y = 1, NODE_NX = 2, NODE_NY = 3, NODE_SW = 4, NODE_PLANES = 5 };

// Kernel functors of lib/bem2d_kernel.cpp:15-112 and IdKernel (lib/bem2d_kernel.h:91-98) for real k.
// (x, y) = test and trial points, n1 / n2 = test and trial normals.  kinv = 1/k, k4 = k/4,
// eta4 = eta/4.  1/d comes from one refined reciprocal square root, which also gives 1/(kd).
template <int KIND>
__host__ __device__ __forceinline__ void kernel_eval(double k, double kinv, double k4, double eta4,
                                                     double xx, double xy, double yx, double yy,
                                                     double n1x, double n1y, double n2x,
                                                     double n2y, double& kr, double& ki)
{
    if (KIND == BEM2D_KERNEL_ID) {
        kr = 1.0;
        ki = 0.0;
        return;
    }
    if (KIND == BEM2D_KERNEL_NONE) {
        kr = 0.0;
        ki = 0.0;
        return;
    }
    double wx, wy, d2, nw;
    {
        // plain multiply/add (no contraction) as in the reference: n.w cancels for near nodes
#pragma clang fp contract(off)
        wx = xx - yx; // x - y
        wy = xy - yy;
        d2 = wx * wx + wy * wy;
        if (KIND == BEM2D_KERNEL_DL)
            nw = n2x * wx + n2y * wy;
        else
            nw = n1x * wx + n1y * wy;
    }
    const double rd = rsqrt_refined(d2); // 1/d
    const double d = d2 * rd;
    double h0r, h0i, h1r, h1i;
    hankel01_inv(k * d, kinv * rd, h0r, h0i, h1r, h1i);
    if (KIND == BEM2D_KERNEL_SL) { // (i/4) H0
        kr = -0.25 * h0i;
        ki = 0.25 * h0r;
        return;
    }
    // -(ik/4) c H1 with c = n.w/d
    // c = n_y.(y - x)/d for DL, n_x.(x - y)/d otherwise
    const double c = (KIND == BEM2D_KERNEL_DL ? -nw : nw) * rd;
    kr = k4 * c * h1i;
    ki = -k4 * c * h1r;
    if (KIND == BEM2D_KERNEL_COMBINED) { // CDL - i eta (i/4) H0 = CDL + (eta/4) H0
        kr += eta4 * h0r;
        ki += eta4 * h0i;
    }
}

// Gauss-node tables: for element e and node a, position, unit normal and |Gamma'| w_a.
static __global__ void prep_nodes_kernel(int64_t n, int nq, const int32_t* __restrict__ type,
                                  const double* __restrict__ param, const double* __restrict__ gx,
                                  const double* __restrict__ gw, double* __restrict__ nodes)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nq) return;
    const int64_t e = t % n;
    const int a = (int)(t / n);
    double p[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) p[q] = param[8 * e + q];
    double mx, my, dx, dy, nx, ny, len;
    elem_eval(type[e], p, gx[a], mx, my, dx, dy);
    normal_len(dx, dy, nx, ny, len);
    const int64_t plane = n * nq;
    nodes[NODE_X * plane + t] = mx;
    nodes[NODE_Y * plane + t] = my;
    nodes[NODE_NX * plane + t] = nx;
    nodes[NODE_NY * plane + t] = ny;
    nodes[NODE_SW * plane + t] = len * gw[a];
}

// Element bounding circles, three planes of n after the node planes: centre of the Gauss nodes
// and the largest node distance from it (rounded up).  For any nodes x of e and y of f,
// |x - y| >= |c_e - c_f| - r_e - r_f.
enum { BOUND_CX = 0, BOUND_CY = 1, BOUND_R = 2, BOUND_PLANES = 3 };

static __global__ void prep_bounds_kernel(int64_t n, int nq, const double* __restrict__ nodes,
                                   double* __restrict__ bounds)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int64_t plane = n * nq;
    double sx = 0.0, sy = 0.0;
    for (int a = 0; a < nq; ++a) {
        sx += nodes[NODE_X * plane + (int64_t)a * n + e];
        sy += nodes[NODE_Y * plane + (int64_t)a * n + e];
    }
    const double cx = sx / nq, cy = sy / nq;
    double r2 = 0.0;
    for (int a = 0; a < nq; ++a) {
        const double dx = nodes[NODE_X * plane + (int64_t)a * n + e] - cx;
        const double dy = nodes[NODE_Y * plane + (int64_t)a * n + e] - cy;
        r2 = fmax(r2, dx * dx + dy * dy);
    }
    bounds[BOUND_CX * n + e] = cx;
    bounds[BOUND_CY * n + e] = cy;
    bounds[BOUND_R * n + e] = sqrt(r2) * (1.0 + 1e-12) + 1e-300;
}

// ---- batched launches --------------------------------------------------------------------------
// An assembly launch serves up to kMaxBatch problems (meshes) of one kernel kind and one nq: the
// descriptors travel by value in the kernel arguments, and workgroup b belongs to the problem d with
// start[d] <= b < start[d + 1].  A single problem is a batch of one.
constexpr int kMaxBatch = 16;

struct AsmProblem {
    int64_t n;
    int nq, m1, m2, m3;           // Gauss points; node counts of AdaptedGauss1/2/3
    const double* nodes;          // NODE_PLANES planes of nq*n, then BOUND_PLANES planes of n
    const int64_t* adj;           // index | next | prev, n each
    const int64_t* special;       // (i, j, rule) triples
    int64_t nspecial;
    const int32_t* type;
    const double* param;
    const double* rules;          // AG1 x|y|w, AG2 x|y|w, AG3 x|y|w, then the abscissa blocks
    const double* g1d;            // Gauss1D x | w
    double k, eta, kscale, mscale;
    double2* A;
    int64_t lda;
};

struct AsmBatch {
    int count;
    int64_t start[kMaxBatch + 1]; // first workgroup of each problem; start[count] = grid size
    AsmProblem p[kMaxBatch];
};

__device__ __forceinline__ int batch_find(const int64_t* start, int count, int64_t b)
{
    int d = 0;
    while (d + 1 < count && b >= start[d + 1]) ++d;
    return d;
}

// Evaluation modes of the regular pairs (kernel_pair).
enum { PAIR_GENERAL = 0, PAIR_MID = 1, PAIR_FAR = 2 };

// Loop constants of the regular pairs: k, 1/k, k/4, eta/4, and for the modes that know
// k|x - y| >= kHankelMidX the amplitude factor sqrt(2/(pi k)) folded into k/4, eta/4 and 1/4
// (H_nu carries sqrt(2/(pi k d)) = sqrt(2/(pi k)) / sqrt(d)) and the operands of hankel_tops.
struct PairConst {
    double k, kinv, k4, eta4;
    double k4a, eta4a, a4;
    HankelTops h;
    const double* tab; // PAIR_GENERAL: the piece tables in LDS (stage_piece_tables)
};

template <int MODE>
__device__ __forceinline__ PairConst pair_const(double k, double eta, const double* tab)
{
    PairConst pc;
    pc.tab = tab;
    pc.k = k;
    pc.kinv = 1.0 / k;
    pc.k4 = 0.25 * k;
    pc.eta4 = 0.25 * eta;
    const double amp = sqrt(2.0 * kInvPi * pc.kinv);
    pc.k4a = pc.k4 * amp;
    pc.eta4a = pc.eta4 * amp;
    pc.a4 = 0.25 * amp;
    if (MODE == PAIR_FAR)
        pc.h = hankel_tops_vgpr<BEM2D_HANKEL_NPIECE>();
    else if (MODE == PAIR_MID)
        pc.h = hankel_tops_vgpr<BEM2D_HANKEL_NPIECE - 1>();
    return pc;
}

// Kernel values of one regular node pair for both orders of its element pair: (kr, ki) with the
// test node x on element i and the trial node y on element j, (mr, mi) with test node y on j and
// trial node x on i.  The two share |x - y| and with it H0 and H1; only the normal factor differs,
// and the negated difference y - x is exact, so each value is the one kernel_eval gives for its
// order.  MODE: PAIR_GENERAL evaluates H0, H1 by the piecewise code; PAIR_MID knows
// k|x - y| >= kHankelMidX and PAIR_FAR knows k|x - y| >= kHankelFarX, and these use the one
// polynomial set that covers the range, without a test.
template <int KIND, int MODE>
__device__ __forceinline__ void kernel_pair(const PairConst& pc, double xx, double xy, double yx,
                                            double yy, double n1x, double n1y, double n2x,
                                            double n2y, double& kr, double& ki, double& mr,
                                            double& mi)
{
    if (KIND == BEM2D_KERNEL_ID || KIND == BEM2D_KERNEL_NONE) {
        kr = mr = (KIND == BEM2D_KERNEL_ID) ? 1.0 : 0.0;
        ki = mi = 0.0;
        return;
    }
    double wx, wy, d2, nw1, nw2;
    if (MODE == PAIR_GENERAL) {
        // near pairs: plain multiply/add as in kernel_eval (n.w cancels for near nodes)
#pragma clang fp contract(off)
        wx = xx - yx; // x - y
        wy = xy - yy;
        d2 = wx * wx + wy * wy;
        nw1 = n1x * wx + n1y * wy; // n_x.(x - y)
        nw2 = n2x * wx + n2y * wy; // n_y.(x - y)
    } else {
        // pairs at k|x - y| >= kHankelMidX: fused (the negation y - x stays exact either way)
        wx = xx - yx;
        wy = xy - yy;
        d2 = fma(wx, wx, wy * wy);
        nw1 = fma(n1x, wx, n1y * wy);
        nw2 = fma(n2x, wx, n2y * wy);
    }
    const double rd = rsqrt_refined(d2);
    const double d = d2 * rd;
    const double x = pc.k * d, invx = pc.kinv * rd;
    if (MODE != PAIR_GENERAL) {
        // H0 = A (u0r + i u0i), H1 = A (u1r + i u1i) as in hankel01_combine, with the amplitude
        // A = +-sqrt(2/(pi k)) / sqrt(d) (the sign that of the half-turn reduction) and its
        // constant factor folded into the scalar factors of the kernel
        double P0, Q0, P1, Q1, S, C;
        if (MODE == PAIR_FAR)
            hankel_pq_far(invx, P0, Q0, P1, Q1, pc.h);
        else
            hankel_pq<BEM2D_HANKEL_NPIECE - 1>(invx, P0, Q0, P1, Q1, pc.h);
        const int nt = sincos_half_turn(x, S, C, pc.h);
        const double y = with_parity_sign(rsqrt_refined(d), nt); // A / sqrt(2/(pi k))
        const double u0r = P0 * C - Q0 * S, u0i = P0 * S + Q0 * C;
        const double u1r = P1 * S + Q1 * C, u1i = Q1 * S - P1 * C;
        if (KIND == BEM2D_KERNEL_SL) {
            const double a4 = pc.a4 * y;
            kr = mr = -a4 * u0i;
            ki = mi = a4 * u0r;
            return;
        }
        const double ka = pc.k4a * y * rd;
        const double c = ka * (KIND == BEM2D_KERNEL_DL ? -nw2 : nw1);
        const double cm = ka * (KIND == BEM2D_KERNEL_DL ? nw1 : -nw2);
        double gr = 0.0, gi = 0.0;
        if (KIND == BEM2D_KERNEL_COMBINED) {
            const double ea = pc.eta4a * y;
            gr = ea * u0r;
            gi = ea * u0i;
        }
        kr = fma(c, u1i, gr);
        ki = fma(-c, u1r, gi);
        mr = fma(cm, u1i, gr);
        mi = fma(-cm, u1r, gi);
        return;
    }
    double h0r, h0i, h1r, h1i;
    hankel01_gather(x, invx, pc.tab, h0r, h0i, h1r, h1i);
    const double k4 = pc.k4, eta4 = pc.eta4;
    if (KIND == BEM2D_KERNEL_SL) {
        kr = mr = -0.25 * h0i;
        ki = mi = 0.25 * h0r;
        return;
    }
    // -(ik/4) c H1: DL c = n_y.(y - x)/d, CDL c = n_x.(x - y)/d; in the swapped order the test
    // and trial normals trade places and x - y changes sign.
    const double c = (KIND == BEM2D_KERNEL_DL ? -nw2 : nw1) * rd;
    const double cm = (KIND == BEM2D_KERNEL_DL ? nw1 : -nw2) * rd;
    kr = k4 * c * h1i;
    ki = -k4 * c * h1r;
    mr = k4 * cm * h1i;
    mi = -k4 * cm * h1r;
    if (KIND == BEM2D_KERNEL_COMBINED) {
        const double gr = eta4 * h0r, gi = eta4 * h0i;
        kr += gr;
        ki += gi;
        mr += gr;
        mi += gi;
    }
}

// Tile shape of the regular assembly: one wavefront (one workgroup) takes 8 test x 8 trial
// elements, lane l the pair (8 p + l % 8, 8 q + l / 8).  The tier of a wavefront (far / mid /
// general, below) is decided by its nearest pair, and a square tile spans |i - j| over 16 elements
// rather than 64: near the diagonal fewer wavefronts mix near and far pairs (kite N = 4000, k = 100:
// about 2% of the wavefronts take the general loop against 5% with 64 x 1 tiles).  Both stores
// then cover whole 128-byte segments (8 consecutive rows of a column).
constexpr int kTileSide = 8;
__host__ __device__ __forceinline__ int node_stride(int nq) { return nq * NODE_PLANES + 1; }

// Tiles of the upper triangle of the P x P grid of 8 x 8 blocks: row block p holds the P - p tiles
// q = p .. P - 1, so tile t lies in the largest p with tiles_before(p) = p P - p (p - 1) / 2 <= t.
__host__ __device__ __forceinline__ int64_t tiles_before(int64_t p, int64_t P)
{
    return p * P - p * (p - 1) / 2;
}
__host__ __device__ __forceinline__ int64_t upper_tile_count(int64_t n)
{
    const int64_t P = (n + kTileSide - 1) / kTileSide;
    return tiles_before(P, P);
}

// Node-pair sums of one element pair (i, j) over test nodes a0 <= a < a1: direct (ar, ai) =
// sum_a w_a sum_b w_b K(x_a, y_b) for entry (i, j), in the order of the reference's Integrate;
// mirror (mr, mi) for entry (j, i).
template <int KIND, int NQ, int MODE>
__device__ __forceinline__ void pair_sums(int nq, int a0, int a1, int64_t n, int64_t i,
                                          const double* __restrict__ nodes, const double* sc,
                                          const double* tab, double k, double eta, double& ar,
                                          double& ai, double& mr, double& mi)
{
    const PairConst pc = pair_const<MODE>(k, eta, tab);
    const int64_t plane = n * nq;
    ar = ai = mr = mi = 0.0;
    for (int a = a0; a < a1; ++a) {
        const int64_t o = (int64_t)a * n + i;
        const double xx = nodes[NODE_X * plane + o], xy = nodes[NODE_Y * plane + o];
        const double n1x = nodes[NODE_NX * plane + o], n1y = nodes[NODE_NY * plane + o];
        const double swa = nodes[NODE_SW * plane + o];
        double br = 0.0, bi = 0.0, cr = 0.0, ci = 0.0;
        for (int b = 0; b < nq; ++b) {
            const double* y = sc + b * 5;
            double kr, ki, kmr, kmi;
            kernel_pair<KIND, MODE>(pc, xx, xy, y[0], y[1], n1x, n1y, y[2], y[3], kr, ki, kmr,
                                    kmi);
            br = fma(kr, y[4], br);
            bi = fma(ki, y[4], bi);
            cr = fma(kmr, y[4], cr);
            ci = fma(kmi, y[4], ci);
        }
        ar = fma(br, swa, ar);
        ai = fma(bi, swa, ai);
        mr = fma(cr, swa, mr);
        mi = fma(ci, swa, mi);
    }
}

// Regular pairs (Integrate with Gauss2D, lib/bem2d_quadrature.h:37-65, :283-286), both orders of
// each element pair from one evaluation of the Hankel functions: workgroup t takes tile t of the
// upper triangle and writes A(i, j) and A(j, i) for its pairs with j > i.
// SPLIT wavefronts share a tile, wavefront w summing the test nodes a in [w nq / SPLIT,
// (w + 1) nq / SPLIT) of every pair; their partial sums meet in LDS and are added in wavefront
// order.  Small problems launch too few tiles to fill the device, and each lane's 256 node pairs
// then run back to back on an otherwise idle SIMD: SPLIT = 4 cuts that chain fourfold.
// The trial elements' node data sit in LDS (one odd-strided row per element: the eight elements of
// a wavefront fall in different banks).  A tile whose pairs are all at k|x - y| >= kHankelFarX by
// the bounding circles takes the branch-free asymptotic loop; all at >= kHankelMidX, the
// branch-free loop of the last piece; otherwise the general loop (every wavefront of the tile sees
// the same pairs and takes the same loop).  Coincident / adjacent pairs are skipped here and
// written by assemble_special_kernel.  NQ = nq at compile time (0: runtime nq).
template <int KIND, int NQ, int SPLIT>
__global__ __launch_bounds__(64 * SPLIT) void assemble_regular_kernel(const AsmBatch batch)
{
    const int di = batch_find(batch.start, batch.count, blockIdx.x);
    const AsmProblem& pb = batch.p[di];
    const int64_t n = pb.n;
    const int nq_rt = pb.nq;
    const double* __restrict__ nodes = pb.nodes;
    const double* __restrict__ bounds = pb.nodes + (int64_t)NODE_PLANES * n * pb.nq;
    const int64_t* __restrict__ index = pb.adj;
    const int64_t* __restrict__ next = pb.adj + n;
    const int64_t* __restrict__ prev = pb.adj + 2 * n;
    const double k = pb.k, eta = pb.eta, kscale = pb.kscale;
    double2* __restrict__ A = pb.A;
    const int64_t lda = pb.lda;
    extern __shared__ double sh[]; // [kTileSide][node_stride(nq)]
    __shared__ double piece_tab[kPieceTableDoubles];
    __shared__ double part[SPLIT > 1 ? SPLIT - 1 : 1][4][64];
    const int nq = NQ ? NQ : nq_rt;
    const int ns = node_stride(nq);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t P = (n + kTileSide - 1) / kTileSide;
    const int64_t t = (int64_t)blockIdx.x - batch.start[di];
    const double qq = (double)P + 0.5; // smaller root of p^2 / 2 - (P + 1/2) p + t = 0
    int64_t p = (int64_t)(qq - sqrt(fmax(qq * qq - 2.0 * (double)t, 0.0)));
    p = p < 0 ? 0 : (p >= P ? P - 1 : p);
    while (p > 0 && tiles_before(p, P) > t) --p;
    while (p + 1 < P && tiles_before(p + 1, P) <= t) ++p;
    const int64_t q = p + (t - tiles_before(p, P));
    const int64_t j0 = q * kTileSide;
    const int64_t plane = n * nq;

    for (int u = threadIdx.x; u < kTileSide * nq; u += 64 * SPLIT) {
        const int c = u / nq, b = u - c * nq;
        const int64_t j = j0 + c;
        if (j < n) {
            const int64_t o = (int64_t)b * n + j;
#pragma unroll
            for (int f = 0; f < NODE_PLANES; ++f) sh[c * ns + b * NODE_PLANES + f] = nodes[f * plane + o];
        }
    }
    __syncthreads();
    const int c = lane / kTileSide;
    const int64_t i = p * kTileSide + lane % kTileSide, j = j0 + c;
    // (every wavefront of the tile leaves, or stays, at the same lanes: the barriers below are
    // reached by all wavefronts or by none)
    if (i >= n || j >= n) return;

    const int64_t idx_i = index[i], nx_i = next[i], pv_i = prev[i];
    const int64_t idx_j = index[j], nx_j = next[j], pv_j = prev[j];
    const bool direct = j > i && !(idx_j == idx_i || nx_i == idx_j || pv_i == idx_j);
    const bool mirror = j > i && !(idx_i == idx_j || nx_j == idx_i || pv_j == idx_i);
    const bool any = direct || mirror;
    bool far = true, mid = true;
    if (any) {
        const double dx = bounds[BOUND_CX * n + i] - bounds[BOUND_CX * n + j];
        const double dy = bounds[BOUND_CY * n + i] - bounds[BOUND_CY * n + j];
        const double kgap =
            k * (sqrt(dx * dx + dy * dy) - bounds[BOUND_R * n + i] - bounds[BOUND_R * n + j]);
        far = kgap >= kHankelFarX * (1.0 + 1e-12);
        mid = kgap >= kHankelMidX * (1.0 + 1e-12);
    }
    const bool all_far = __all(far), all_mid = __all(mid);
    if (!any) return;
    const double* sc = sh + c * ns;
    const int a0 = wave * nq / SPLIT, a1 = (wave + 1) * nq / SPLIT;
    double ar, ai, mr, mi;
    if (all_far) {
        pair_sums<KIND, NQ, PAIR_FAR>(nq, a0, a1, n, i, nodes, sc, nullptr, k, eta, ar, ai, mr, mi);
    } else if (all_mid) {
        pair_sums<KIND, NQ, PAIR_MID>(nq, a0, a1, n, i, nodes, sc, nullptr, k, eta, ar, ai, mr, mi);
    } else {
        // near pairs: the piece tables go to LDS, copied by wavefront 0's remaining lanes
        if (wave == 0) {
            const unsigned long long act = __ballot(1);
            const int rank = __popcll(act & ((1ull << lane) - 1ull));
            stage_piece_tables(piece_tab, rank, __popcll(act));
        }
        __syncthreads();
        pair_sums<KIND, NQ, PAIR_GENERAL>(nq, a0, a1, n, i, nodes, sc, piece_tab, k, eta, ar, ai,
                                          mr, mi);
    }
    if (SPLIT > 1) {
        if (wave > 0) {
            part[wave - 1][0][lane] = ar;
            part[wave - 1][1][lane] = ai;
            part[wave - 1][2][lane] = mr;
            part[wave - 1][3][lane] = mi;
        }
        __syncthreads();
        if (wave > 0) return;
#pragma unroll
        for (int w = 0; w < SPLIT - 1; ++w) {
            ar += part[w][0][lane];
            ai += part[w][1][lane];
            mr += part[w][2][lane];
            mi += part[w][3][lane];
        }
    }
    if (direct) A[i + lda * j] = make_double2(kscale * ar, kscale * ai);
    if (mirror) A[j + lda * i] = make_double2(kscale * mr, kscale * mi);
}

// Abscissa blocks of the singular-pair rules (built with the mesh, after the x|y|w arrays of
// AG1, AG2, AG3): per rule, the counts nx, ny of its distinct x and y, then kRuleAbscissae slots
// of each (the distinct values), then per node the indices (ix, iy) of its x and y, all doubles.
// The element maps at a rule's x (test side) and y (trial side) then need evaluating once per
// distinct value instead of once per node (AdaptedGauss1 has L N distinct values for N^2 L^2
// nodes).  A count of 0 (more distinct values than slots) means: evaluate at every node.
constexpr int kRuleAbscissae = 256;
__host__ __device__ __forceinline__ int64_t rule_block_doubles(int m)
{
    return 2 + 2 * kRuleAbscissae + 2 * (int64_t)m;
}

// Element map at t for the singular-pair kernel: position, unit normal, |Gamma'|.
struct ElemPoint {
    double mx, my, nx, ny, s;
};
__device__ __forceinline__ ElemPoint elem_point(int type, const double* p, double t)
{
    ElemPoint e;
    double dx, dy;
    elem_eval(type, p, t, e.mx, e.my, dx, dy);
    normal_len(dx, dy, e.nx, e.ny, e.s);
    return e;
}

// Coincident / adjacent pairs: one block of 256 threads per entry (i, j, rule).
// rule 1: IntegrateIdent with AdaptedGauss1 (lib/bem2d_quadrature.h:67-109), plus the EvalIdent
//         mass entry (lib/bem2d_quadrature.cpp:108-123) scaled by mscale;
// rule 3: next(i) == index(j), Integrate with AdaptedGauss3; rule 2: prev, AdaptedGauss2.
// Kernel values of the real-k functors for the singular-pair body below.
struct RealKernelEval {
    double k, kinv, k4, eta4;
    __device__ __forceinline__ RealKernelEval(double k_, double eta_)
        : k(k_), kinv(1.0 / k_), k4(0.25 * k_), eta4(0.25 * eta_)
    {
    }
    template <int KIND>
    __device__ __forceinline__ void eval(double xx, double xy, double yx, double yy, double n1x,
                                         double n1y, double n2x, double n2y, double& kr,
                                         double& ki) const
    {
        kernel_eval<KIND>(k, kinv, k4, eta4, xx, xy, yx, yy, n1x, n1y, n2x, n2y, kr, ki);
    }
};

// One singular entry per 256-thread block; ev supplies the kernel values (real or complex k).
template <int KIND, class Eval>
__device__ __forceinline__ void special_entry(const AsmBatch& batch, const Eval& ev)
{
    const int di = batch_find(batch.start, batch.count, blockIdx.x);
    const AsmProblem& pb = batch.p[di];
    const int64_t* __restrict__ special = pb.special;
    const int32_t* __restrict__ type = pb.type;
    const double* __restrict__ param = pb.param;
    const double* __restrict__ rules = pb.rules;
    const int m1 = pb.m1, m2 = pb.m2, m3 = pb.m3, nq = pb.nq;
    const double* __restrict__ gx = pb.g1d;
    const double* __restrict__ gw = pb.g1d + pb.nq;
    const double kscale = pb.kscale, mscale = pb.mscale;
    double2* __restrict__ A = pb.A;
    const int64_t lda = pb.lda;
    __shared__ double red[2][256];
    const int64_t e = (int64_t)blockIdx.x - batch.start[di];
    const int64_t i = special[3 * e], j = special[3 * e + 1];
    const int rule = (int)special[3 * e + 2];
    const int ti = type[i], tj = type[j];
    double pi_[8], pj[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        pi_[q] = param[8 * i + q];
        pj[q] = param[8 * j + q];
    }
    // element maps at the rule's distinct abscissae: the test element at x (and, for the Duffy
    // split of rule 1, at 1 - x), the trial element at y for rules 2 and 3
    __shared__ ElemPoint cache[2][kRuleAbscissae];
    double ar = 0.0, ai = 0.0;
    if (KIND != BEM2D_KERNEL_NONE && kscale != 0.0) {
        const double* rx;
        const double* blk = rules + 3 * (int64_t)(m1 + m2 + m3);
        int m;
        if (rule == 1) {
            rx = rules;
            m = m1;
        } else if (rule == 3) {
            rx = rules + 3 * (int64_t)(m1 + m2);
            m = m3;
            blk += rule_block_doubles(m1) + rule_block_doubles(m2);
        } else {
            rx = rules + 3 * (int64_t)m1;
            m = m2;
            blk += rule_block_doubles(m1);
        }
        const double* ry = rx + m;
        const double* rw = ry + m;
        const int nx = (int)blk[0], ny = (int)blk[1];
        const double* ux = blk + 2;
        const double* uy = ux + kRuleAbscissae;
        const double* ridx = uy + kRuleAbscissae;
        const bool cached = nx > 0 && ny > 0;
        if (cached) {
            for (int u = threadIdx.x; u < nx; u += blockDim.x) {
                cache[0][u] = elem_point(ti, pi_, ux[u]);
                if (rule == 1) cache[1][u] = elem_point(ti, pi_, 1 - ux[u]);
            }
            if (rule != 1)
                for (int u = threadIdx.x; u < ny; u += blockDim.x) cache[1][u] = elem_point(tj, pj, uy[u]);
            __syncthreads();
        }
        for (int q = threadIdx.x; q < m; q += blockDim.x) {
            const double xq = rx[q], yq = ry[q], wq = rw[q];
            const int ix = cached ? (int)ridx[2 * q] : 0, iy = cached ? (int)ridx[2 * q + 1] : 0;
            double kr, ki;
            if (rule == 1) {
                // Duffy split of the coincident pair: (x, (1-y)x) and (1-x, 1-x+yx), Jacobian x.
                // The trial parameters are rounded as the reference rounds them (no fused
                // multiply-add): near-coincident nodes make the entry depend on their last bits.
                double u2a, u2b;
                {
#pragma clang fp contract(off)
                    u2a = (1 - yq) * xq;
                    u2b = 1 - xq + yq * xq;
                }
                ElemPoint p1 = cached ? cache[0][ix] : elem_point(ti, pi_, xq);
                ElemPoint p2 = elem_point(tj, pj, u2a);
                ev.template eval<KIND>(p1.mx, p1.my, p2.mx, p2.my, p1.nx, p1.ny, p2.nx, p2.ny,
                                       kr, ki);
                ar += xq * kr * p1.s * p2.s * wq;
                ai += xq * ki * p1.s * p2.s * wq;
                p1 = cached ? cache[1][ix] : elem_point(ti, pi_, 1 - xq);
                p2 = elem_point(tj, pj, u2b);
                ev.template eval<KIND>(p1.mx, p1.my, p2.mx, p2.my, p1.nx, p1.ny, p2.nx, p2.ny,
                                       kr, ki);
                ar += xq * kr * p1.s * p2.s * wq;
                ai += xq * ki * p1.s * p2.s * wq;
            } else {
                const ElemPoint p1 = cached ? cache[0][ix] : elem_point(ti, pi_, xq);
                const ElemPoint p2 = cached ? cache[1][iy] : elem_point(tj, pj, yq);
                ev.template eval<KIND>(p1.mx, p1.my, p2.mx, p2.my, p1.nx, p1.ny, p2.nx, p2.ny,
                                       kr, ki);
                ar += kr * p1.s * p2.s * wq;
                ai += ki * p1.s * p2.s * wq;
            }
        }
    }
    red[0][threadIdx.x] = ar;
    red[1][threadIdx.x] = ai;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double vr = kscale * red[0][0], vi = kscale * red[1][0];
        if (rule == 1 && mscale != 0.0) {
            double mass = 0.0; // sum_t |Gamma_j'(x_t)| w_t, Gauss1D(N)
            for (int t = 0; t < nq; ++t) {
                double mx, my, dx, dy;
                elem_eval(tj, pj, gx[t], mx, my, dx, dy);
                mass += sqrt(dx * dx + dy * dy) * gw[t];
            }
            vr += mscale * mass;
        }
        A[i + lda * j] = make_double2(vr, vi);
    }
}

template <int KIND>
__global__ __launch_bounds__(256) void assemble_special_kernel(const AsmBatch batch)
{
    const AsmProblem& pb = batch.p[batch_find(batch.start, batch.count, blockIdx.x)];
    special_entry<KIND>(batch, RealKernelEval(pb.k, pb.eta));
}

// DotProdBasFuns (lib/bem2d_quadrature.h:111-156) for the plane-wave field functors
// (lib/bem2d_fun.h:21-24, 45-49, 71-75): thread per (row i, right-hand side r).
static __global__ void rhs_kernel(int64_t n, int nq, const double* __restrict__ nodes, int kind, double k,
                           double eta, const double* __restrict__ dirs, double scale,
                           double2* __restrict__ B, int64_t ldb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n) return;
    const double ddx = dirs[2 * r], ddy = dirs[2 * r + 1];
    const int64_t plane = n * nq;
    double ar = 0.0, ai = 0.0;
    for (int a = 0; a < nq; ++a) {
        const int64_t o = (int64_t)a * n + i;
        const double px = nodes[NODE_X * plane + o], py = nodes[NODE_Y * plane + o];
        const double sw = nodes[NODE_SW * plane + o];
        const double ph = k * (ddx * px + ddy * py);
        const double c = cos(ph), s = sin(ph);
        double fr, fi;
        if (kind == BEM2D_FIELD_PLANE) {
            fr = c;
            fi = s;
        } else {
            const double nd = nodes[NODE_NX * plane + o] * ddx + nodes[NODE_NY * plane + o] * ddy;
            const double g = (kind == BEM2D_FIELD_COMBINED_PLANE) ? k * nd - eta : k * nd;
            fr = -g * s; // i g e^{i ph}
            fi = g * c;
        }
        ar = fma(fr, sw, ar);
        ai = fma(fi, sw, ai);
    }
    B[i + ldb * r] = make_double2(scale * ar, scale * ai);
}

// Far-field signature F[a + nang*j] = sum_t exp(-ik xhat_a.y_t) |Gamma_j'| w_t: thread per (a, j).
static __global__ void farfield_signature_kernel(int64_t n, int nq, const double* __restrict__ nodes,
                                          double k, const double* __restrict__ theta, int64_t nang,
                                          double2* __restrict__ F)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nang) return;
    const int64_t a = t % nang, j = t / nang;
    const double cx = cos(theta[a]), cy = sin(theta[a]);
    const int64_t plane = n * nq;
    double sr = 0.0, si = 0.0;
    for (int q = 0; q < nq; ++q) {
        const int64_t o = (int64_t)q * n + j;
        const double ph = -k * (cx * nodes[NODE_X * plane + o] + cy * nodes[NODE_Y * plane + o]);
        const double sw = nodes[NODE_SW * plane + o];
        sr = fma(cos(ph), sw, sr);
        si = fma(sin(ph), sw, si);
    }
    F[t] = make_double2(sr, si);
}

// out[a + nang*r] = pref * sum_j F[a + nang*j] coef[j + ldc*r]: one 256-thread workgroup per
// (a, r), its threads striding over j; the partial sums meet by lane exchange and in LDS.  (A thread
// per (a, r) walked all n elements alone: ~0.8 ms per request at n = 2500 with 36 angles.)
static __global__ __launch_bounds__(256) void farfield_combine_kernel(
    int64_t n, const double2* __restrict__ F, int64_t nang, const double2* __restrict__ coef,
    int64_t ldc, double pr, double pi_, double2* __restrict__ out)
{
    const int64_t a = blockIdx.x;
    const int r = blockIdx.y;
    double vr = 0.0, vi = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += 256) {
        const double2 f = F[a + nang * j];
        const double2 c = coef[j + ldc * r];
        vr += f.x * c.x - f.y * c.y;
        vi += f.x * c.y + f.y * c.x;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        vr += __shfl_xor(vr, o, 64);
        vi += __shfl_xor(vi, o, 64);
    }
    __shared__ double part[2][4];
    if (threadIdx.x % 64 == 0) {
        part[0][threadIdx.x / 64] = vr;
        part[1][threadIdx.x / 64] = vi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        vr = (part[0][0] + part[0][1]) + (part[0][2] + part[0][3]);
        vi = (part[1][0] + part[1][1]) + (part[1][2] + part[1][3]);
        out[a + nang * r] = make_double2(pr * vr - pi_ * vi, pr * vi + pi_ * vr);
    }
}

} // namespace bem2d_dev
