// Several basis functions per element: ConstBasis, PolBasis, LinBasis, WaveBasis and WavePolBasis
// (lib/bem2d_basistypes.h:16-99) on every element of a mesh, the matrix row / column of basis b on
// element e being e * nb + b (Geometry::FlatMap, lib/bem2d_geometry.h:33-57, when AddBasis adds
// each basis to all elements).  The Galerkin entries carry conj(phi_a(x)) phi_b(y) in the
// quadrature sums (Integrate / IntegrateIdent, lib/bem2d_quadrature.h:37-109), the mass matrix
// conj(phi_a) phi_b (EvalIdent, lib/bem2d_quadrature.cpp:108-123), the right-hand sides
// conj(phi_a) (DotProdBasFuns, lib/bem2d_quadrature.h:111-128) and the potential and far field the
// density sum_b c_b phi_b (EvalKernel, lib/bem2d_quadrature.h:184-228).
//
// A basis value depends on the element-local parameter only, so at the Gauss nodes it is one
// small table shared by every element.  The regular pairs keep the real-k tile kernel's shape (one
// lane per element pair, both orders of the pair from one Hankel evaluation, far / mid / general
// tiers): a lane sums, per test node, the trial-node kernel values weighted by the trial bases,
// then adds them into its block of entries weighted by the test bases.  The bases are taken in
// groups of at most kGroup per launch (group pair (g1, g2): test bases of g1 against trial bases
// of g2), which bounds the accumulators a lane holds.
//
// Also here: the arc-length parameterisation of a curve (Curve::Length, lib/bem2d_curve.cpp:11-25,
// and AnalyticCurve::ParameterizeArc, lib/bem2d_shape.cpp:332-366), host code.
#include <algorithm>
#include <cmath>
#include <vector>

#pragma clang diagnostic ignored "-Wunused-function" // kernels of the shared headers
#include "assembly_kernels.h"
#include "complex_kernels.h"
#include "basis.h"
#include "internal.h"

using namespace bem2d_dev;

namespace {

using bem2d_dev::BasisSet;
using bem2d_dev::basis_value;
constexpr int kMaxBases = BEM2D_MAX_BASES;
constexpr int kGroup = bem2d_dev::kBasisGroup;

// Basis table at the Gauss nodes: bt[2 (a * nbp + b) + {0, 1}], nbp = G * ngroup columns.
__device__ __forceinline__ void stage_basis_table(const BasisSet& s, int nq,
                                                  const double* __restrict__ gx, double* bt,
                                                  int tid, int nthr)
{
    const int nbp = s.G * s.ngroup;
    for (int u = tid; u < nq * nbp; u += nthr) {
        const int a = u / nbp, b = u - a * nbp;
        basis_value(s, b, gx[a], bt[2 * u], bt[2 * u + 1]);
    }
}

// ---- regular pairs, real k -----------------------------------------------------------------------
// Lane sums for the element pair (i, j): direct D[al][be] = entry (i nb + g1 + al, j nb + g2 + be)
// = sum_a conj(phi_{g1+al}(x_a)) sw_a sum_b K(x_a, y_b) sw_b phi_{g2+be}(y_b); mirror
// M[be][al] = entry (j nb + g1 + be, i nb + g2 + al) = sum_a phi_{g2+al}(x_a) sw_a
// sum_b K(y_b, x_a) sw_b conj(phi_{g1+be}(y_b)).
// Pair evaluators: kernel values of one node pair for both orders (kr, ki: test x, trial y;
// mr, mi: test y, trial x).  Real k: kernel_pair of the tier MODE.
template <int KIND, int MODE>
struct RealPair {
    PairConst pc;
    __device__ __forceinline__ RealPair(double k, double eta, const double* tab)
        : pc(pair_const<MODE>(k, eta, tab))
    {
    }
    __device__ __forceinline__ void operator()(double xx, double xy, double yx, double yy,
                                               double n1x, double n1y, double n2x, double n2y,
                                               double& kr, double& ki, double& mr,
                                               double& mi) const
    {
        kernel_pair<KIND, MODE>(pc, xx, xy, yx, yy, n1x, n1y, n2x, n2y, kr, ki, mr, mi);
    }
};

// Complex k = kr + i ki (the kernels' k.im != 0 branches, lib/bem2d_kernel.cpp:15-112): H0(kd),
// H1(kd) once for both orders; the normal factor of the swapped order has the normals traded and
// x - y negated, as in kernel_pair.  tab / nt: the contour table of arg k.
template <int KIND>
struct ComplexPair {
    double kr, ki, eta;
    const double* tab;
    int nt;
    __device__ __forceinline__ void operator()(double xx, double xy, double yx, double yy,
                                               double n1x, double n1y, double n2x, double n2y,
                                               double& vr, double& vi, double& mr,
                                               double& mi) const
    {
        if (KIND == BEM2D_KERNEL_ID || KIND == BEM2D_KERNEL_NONE) {
            vr = mr = (KIND == BEM2D_KERNEL_ID) ? 1.0 : 0.0;
            vi = mi = 0.0;
            return;
        }
        double wx, wy, d2, nw1, nw2;
        {
#pragma clang fp contract(off)
            wx = xx - yx;
            wy = xy - yy;
            d2 = wx * wx + wy * wy;
            nw1 = n1x * wx + n1y * wy; // n_x.(x - y)
            nw2 = n2x * wx + n2y * wy; // n_y.(x - y)
        }
        const double d = sqrt(d2);
        double h0r, h0i, h1r, h1i;
        hankel01_complex(kr * d, ki * d, tab, nt, h0r, h0i, h1r, h1i);
        if (KIND == BEM2D_KERNEL_SL) {
            vr = mr = -0.25 * h0i;
            vi = mi = 0.25 * h0r;
            return;
        }
        const double c = (KIND == BEM2D_KERNEL_DL ? -nw2 : nw1) / d;
        const double cm = (KIND == BEM2D_KERNEL_DL ? nw1 : -nw2) / d;
        const double ar = kr * h1r - ki * h1i, ai = kr * h1i + ki * h1r; // k H1
        vr = 0.25 * c * ai;
        vi = -0.25 * c * ar;
        mr = 0.25 * cm * ai;
        mi = -0.25 * cm * ar;
        if (KIND == BEM2D_KERNEL_COMBINED) {
            vr += 0.25 * eta * h0r;
            vi += 0.25 * eta * h0i;
            mr += 0.25 * eta * h0r;
            mi += 0.25 * eta * h0i;
        }
    }
};

template <int G, class Pair>
__device__ __forceinline__ void pair_sums_basis(int nq, int64_t n, int64_t i,
                                                const double* __restrict__ nodes, const double* sc,
                                                const double* bt, int nbp, int g1, int g2,
                                                const Pair& pair, double (&D)[G][G][2],
                                                double (&M)[G][G][2])
{
    const int64_t plane = n * nq;
#pragma unroll
    for (int u = 0; u < G; ++u)
#pragma unroll
        for (int v = 0; v < G; ++v) D[u][v][0] = D[u][v][1] = M[u][v][0] = M[u][v][1] = 0.0;
    for (int a = 0; a < nq; ++a) {
        const int64_t o = (int64_t)a * n + i;
        const double xx = nodes[NODE_X * plane + o], xy = nodes[NODE_Y * plane + o];
        const double n1x = nodes[NODE_NX * plane + o], n1y = nodes[NODE_NY * plane + o];
        const double swa = nodes[NODE_SW * plane + o];
        double id[G][2], im[G][2];
#pragma unroll
        for (int u = 0; u < G; ++u) id[u][0] = id[u][1] = im[u][0] = im[u][1] = 0.0;
        for (int b = 0; b < nq; ++b) {
            const double* y = sc + b * 5;
            double kr, ki, kmr, kmi;
            pair(xx, xy, y[0], y[1], n1x, n1y, y[2], y[3], kr, ki, kmr, kmi);
            const double w = y[4];
            const double dr = kr * w, di = ki * w, mr = kmr * w, mi = kmi * w;
            const double* row = bt + 2 * b * nbp;
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const double pr = row[2 * (g2 + u)], pi = row[2 * (g2 + u) + 1]; // phi (trial)
                id[u][0] = fma(dr, pr, fma(-di, pi, id[u][0]));
                id[u][1] = fma(dr, pi, fma(di, pr, id[u][1]));
                const double qr = row[2 * (g1 + u)], qi = row[2 * (g1 + u) + 1]; // conj phi (test)
                im[u][0] = fma(mr, qr, fma(mi, qi, im[u][0]));
                im[u][1] = fma(mi, qr, fma(-mr, qi, im[u][1]));
            }
        }
        const double* row = bt + 2 * a * nbp;
#pragma unroll
        for (int u = 0; u < G; ++u) {
            // direct: conj(phi_{g1+u}(x_a)) swa
            const double cr = row[2 * (g1 + u)] * swa, ci = -row[2 * (g1 + u) + 1] * swa;
            // mirror: phi_{g2+u}(x_a) swa
            const double er = row[2 * (g2 + u)] * swa, ei = row[2 * (g2 + u) + 1] * swa;
#pragma unroll
            for (int v = 0; v < G; ++v) {
                D[u][v][0] = fma(cr, id[v][0], fma(-ci, id[v][1], D[u][v][0]));
                D[u][v][1] = fma(cr, id[v][1], fma(ci, id[v][0], D[u][v][1]));
                M[v][u][0] = fma(er, im[v][0], fma(-ei, im[v][1], M[v][u][0]));
                M[v][u][1] = fma(er, im[v][1], fma(ei, im[v][0], M[v][u][1]));
            }
        }
    }
}

// Regular pairs with bases, real k: workgroup t = one wavefront, tile t of the upper triangle of
// 8 x 8 element blocks (as assemble_regular_kernel), blockIdx.y = group pair g1 * ngroup + g2.
// Dynamic LDS: the tile's trial node data, then the basis table.
template <int KIND, int G>
__global__ __launch_bounds__(64) void assemble_regular_basis_kernel(const AsmProblem pb,
                                                                    const BasisSet bs)
{
    const int64_t n = pb.n;
    const int nq = pb.nq;
    const double* __restrict__ nodes = pb.nodes;
    const double* __restrict__ bounds = pb.nodes + (int64_t)NODE_PLANES * n * nq;
    const int64_t* __restrict__ index = pb.adj;
    const int64_t* __restrict__ next = pb.adj + n;
    const int64_t* __restrict__ prev = pb.adj + 2 * n;
    const double k = pb.k, eta = pb.eta, kscale = pb.kscale;
    double2* __restrict__ A = pb.A;
    const int64_t lda = pb.lda;
    const int nb = bs.nb, nbp = G * bs.ngroup;
    const int g1 = (int)(blockIdx.y / bs.ngroup) * G, g2 = (int)(blockIdx.y % bs.ngroup) * G;
    extern __shared__ double sh[];
    __shared__ double piece_tab[kPieceTableDoubles];
    const int ns = node_stride(nq);
    double* bt = sh + kTileSide * ns;
    const int lane = threadIdx.x;
    const int64_t P = (n + kTileSide - 1) / kTileSide;
    const int64_t t = blockIdx.x;
    const double qq = (double)P + 0.5;
    int64_t p = (int64_t)(qq - sqrt(fmax(qq * qq - 2.0 * (double)t, 0.0)));
    p = p < 0 ? 0 : (p >= P ? P - 1 : p);
    while (p > 0 && tiles_before(p, P) > t) --p;
    while (p + 1 < P && tiles_before(p + 1, P) <= t) ++p;
    const int64_t q = p + (t - tiles_before(p, P));
    const int64_t j0 = q * kTileSide;
    const int64_t plane = n * nq;
    for (int u = lane; u < kTileSide * nq; u += 64) {
        const int c = u / nq, b = u - c * nq;
        const int64_t j = j0 + c;
        if (j < n) {
            const int64_t o = (int64_t)b * n + j;
#pragma unroll
            for (int f = 0; f < NODE_PLANES; ++f) sh[c * ns + b * NODE_PLANES + f] = nodes[f * plane + o];
        }
    }
    stage_basis_table(bs, nq, pb.g1d, bt, lane, 64);
    __syncthreads();
    const int c = lane / kTileSide;
    const int64_t i = p * kTileSide + lane % kTileSide, j = j0 + c;
    if (i >= n || j >= n) return; // one wavefront: no barrier follows for the lanes that leave

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
    double D[G][G][2], M[G][G][2];
    if (all_far) {
        pair_sums_basis<G>(nq, n, i, nodes, sc, bt, nbp, g1, g2,
                           RealPair<KIND, PAIR_FAR>(k, eta, nullptr), D, M);
    } else if (all_mid) {
        pair_sums_basis<G>(nq, n, i, nodes, sc, bt, nbp, g1, g2,
                           RealPair<KIND, PAIR_MID>(k, eta, nullptr), D, M);
    } else {
        // near pairs: the piece tables go to LDS, copied by the lanes still here
        const unsigned long long act = __ballot(1);
        const int rank = __popcll(act & ((1ull << lane) - 1ull));
        stage_piece_tables(piece_tab, rank, __popcll(act));
        __syncthreads();
        pair_sums_basis<G>(nq, n, i, nodes, sc, bt, nbp, g1, g2,
                           RealPair<KIND, PAIR_GENERAL>(k, eta, piece_tab), D, M);
    }
#pragma unroll
    for (int u = 0; u < G; ++u)
#pragma unroll
        for (int v = 0; v < G; ++v) {
            if (direct && g1 + u < nb && g2 + v < nb)
                A[(i * nb + g1 + u) + lda * (j * nb + g2 + v)] =
                    make_double2(kscale * D[u][v][0], kscale * D[u][v][1]);
            if (mirror && g1 + u < nb && g2 + v < nb)
                A[(j * nb + g1 + u) + lda * (i * nb + g2 + v)] =
                    make_double2(kscale * M[u][v][0], kscale * M[u][v][1]);
        }
}

// ---- regular pairs, complex k, tiled ---------------------------------------------------------------
// The real-k tile kernel's shape for complex k: one lane per element pair of an 8 x 8 tile, both
// orders of the pair from one complex Hankel evaluation, one loop for all pairs (the complex
// routine picks its region per argument).  Dynamic LDS: the tile's trial node data, the basis
// table, then the contour table of arg k (read by every evaluation in the middle region).
template <int KIND, int G>
__global__ __launch_bounds__(64) void assemble_regular_basis_ckt_kernel(
    const AsmProblem pb, const BasisSet bs, double ki, const double* __restrict__ gtab, int nt)
{
    const int64_t n = pb.n;
    const int nq = pb.nq;
    const double* __restrict__ nodes = pb.nodes;
    const int64_t* __restrict__ index = pb.adj;
    const int64_t* __restrict__ next = pb.adj + n;
    const int64_t* __restrict__ prev = pb.adj + 2 * n;
    double2* __restrict__ A = pb.A;
    const int64_t lda = pb.lda;
    const int nb = bs.nb, nbp = G * bs.ngroup;
    const int g1 = (int)(blockIdx.y / bs.ngroup) * G, g2 = (int)(blockIdx.y % bs.ngroup) * G;
    extern __shared__ double sh[];
    const int ns = node_stride(nq);
    double* bt = sh + kTileSide * ns;
    double* ctab = bt + 2 * nq * nbp;
    const int lane = threadIdx.x;
    const int64_t P = (n + kTileSide - 1) / kTileSide;
    const int64_t t = blockIdx.x;
    const double qq = (double)P + 0.5;
    int64_t p = (int64_t)(qq - sqrt(fmax(qq * qq - 2.0 * (double)t, 0.0)));
    p = p < 0 ? 0 : (p >= P ? P - 1 : p);
    while (p > 0 && tiles_before(p, P) > t) --p;
    while (p + 1 < P && tiles_before(p + 1, P) <= t) ++p;
    const int64_t q = p + (t - tiles_before(p, P));
    const int64_t j0 = q * kTileSide;
    const int64_t plane = n * nq;
    for (int u = lane; u < kTileSide * nq; u += 64) {
        const int c = u / nq, b = u - c * nq;
        const int64_t j = j0 + c;
        if (j < n) {
            const int64_t o = (int64_t)b * n + j;
#pragma unroll
            for (int f = 0; f < NODE_PLANES; ++f) sh[c * ns + b * NODE_PLANES + f] = nodes[f * plane + o];
        }
    }
    stage_basis_table(bs, nq, pb.g1d, bt, lane, 64);
    for (int u = lane; u < nt * kContourDoubles; u += 64) ctab[u] = gtab[u];
    __syncthreads();
    const int c = lane / kTileSide;
    const int64_t i = p * kTileSide + lane % kTileSide, j = j0 + c;
    if (i >= n || j >= n) return;
    const int64_t idx_i = index[i], nx_i = next[i], pv_i = prev[i];
    const int64_t idx_j = index[j], nx_j = next[j], pv_j = prev[j];
    const bool direct = j > i && !(idx_j == idx_i || nx_i == idx_j || pv_i == idx_j);
    const bool mirror = j > i && !(idx_i == idx_j || nx_j == idx_i || pv_j == idx_i);
    if (!(direct || mirror)) return;
    double D[G][G][2], M[G][G][2];
    pair_sums_basis<G>(nq, n, i, nodes, sh + c * ns, bt, nbp, g1, g2,
                       ComplexPair<KIND>{pb.k, ki, pb.eta, ctab, nt}, D, M);
    const double kscale = pb.kscale;
#pragma unroll
    for (int u = 0; u < G; ++u)
#pragma unroll
        for (int v = 0; v < G; ++v) {
            if (direct && g1 + u < nb && g2 + v < nb)
                A[(i * nb + g1 + u) + lda * (j * nb + g2 + v)] =
                    make_double2(kscale * D[u][v][0], kscale * D[u][v][1]);
            if (mirror && g1 + u < nb && g2 + v < nb)
                A[(j * nb + g1 + u) + lda * (i * nb + g2 + v)] =
                    make_double2(kscale * M[u][v][0], kscale * M[u][v][1]);
        }
}

// Launches the tiled complex-k regular kernel (all gp group pairs) on stream s.
template <int KIND, int G>
void launch_ckt(hipStream_t s, const AsmProblem& p, const BasisSet& bs, double ki,
                const double* tab, int nt)
{
    const size_t shm = sizeof(double) * ((size_t)kTileSide * node_stride(p.nq) +
                                         2 * (size_t)p.nq * G * bs.ngroup +
                                         (size_t)kContourDoubles * nt);
    const unsigned gp = (unsigned)(bs.ngroup * bs.ngroup);
    assemble_regular_basis_ckt_kernel<KIND, G>
        <<<dim3((unsigned)upper_tile_count(p.n), gp), 64, shm, s>>>(p, bs, ki, tab, nt);
}

// Launches of fewer workgroups than this take the one-wavefront-per-entry complex-k kernel, whose
// grid of n^2 entries fills the device when the tiles cannot.
constexpr int64_t kCktMinTiles = 2048;

// ---- regular pairs, complex k --------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Entry block (i, j) = (e mod n, e div n) for workgroup e (64 lanes over the nq x nq Gauss2D node
// pairs), group pair blockIdx.y; the singular list's pairs are left to the singular kernel.
template <int KIND, int G>
__global__ __launch_bounds__(64) void assemble_regular_basis_ck_kernel(const AsmProblem p,
                                                                       const BasisSet bs, double ki,
                                                                       const double* __restrict__ tab,
                                                                       int nt)
{
    const int64_t n = p.n;
    const int64_t e = blockIdx.x;
    const int64_t i = e % n, j = e / n;
    const int64_t* __restrict__ index = p.adj;
    const int64_t* __restrict__ next = p.adj + n;
    const int64_t* __restrict__ prev = p.adj + 2 * n;
    const int64_t idx_j = index[j];
    if (index[i] == idx_j || next[i] == idx_j || prev[i] == idx_j) return;
    const int nq = p.nq, nb = bs.nb, nbp = G * bs.ngroup;
    const int g1 = (int)(blockIdx.y / bs.ngroup) * G, g2 = (int)(blockIdx.y % bs.ngroup) * G;
    extern __shared__ double bt[];
    stage_basis_table(bs, nq, p.g1d, bt, threadIdx.x, 64);
    __syncthreads();
    const int64_t plane = n * nq;
    const double* __restrict__ nodes = p.nodes;
    double acc[G][G][2] = {};
    for (int q = threadIdx.x; q < nq * nq; q += 64) {
        const int a = q / nq, b = q % nq; // Gauss2D node (x_a, y_b)
        const int64_t ox = (int64_t)a * n + i, oy = (int64_t)b * n + j;
        double vr, vi;
        kernel_eval_ck<KIND>(p.k, ki, p.eta, tab, nt, nodes[NODE_X * plane + ox],
                             nodes[NODE_Y * plane + ox], nodes[NODE_X * plane + oy],
                             nodes[NODE_Y * plane + oy], nodes[NODE_NX * plane + ox],
                             nodes[NODE_NY * plane + ox], nodes[NODE_NX * plane + oy],
                             nodes[NODE_NY * plane + oy], vr, vi);
        const double s = nodes[NODE_SW * plane + ox] * nodes[NODE_SW * plane + oy];
        vr *= s;
        vi *= s;
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const double cr = bt[2 * (a * nbp + g1 + u)], ci = -bt[2 * (a * nbp + g1 + u) + 1];
            const double tr = cr * vr - ci * vi, ti = cr * vi + ci * vr;
#pragma unroll
            for (int v = 0; v < G; ++v) {
                const double fr = bt[2 * (b * nbp + g2 + v)], fi = bt[2 * (b * nbp + g2 + v) + 1];
                acc[u][v][0] += tr * fr - ti * fi;
                acc[u][v][1] += tr * fi + ti * fr;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < G; ++u)
#pragma unroll
        for (int v = 0; v < G; ++v) {
            const double sr = wave_sum(acc[u][v][0]), si = wave_sum(acc[u][v][1]);
            if (threadIdx.x == 0 && g1 + u < nb && g2 + v < nb)
                p.A[(i * nb + g1 + u) + p.lda * (j * nb + g2 + v)] =
                    make_double2(p.kscale * sr, p.kscale * si);
        }
}

// ---- singular pairs ------------------------------------------------------------------------------
// One 256-thread workgroup per (singular entry e = blockIdx.x, group pair blockIdx.y): the rule /
// Duffy body of the P0 singular kernel with the basis factors conj(phi_{g1+u}(s)) phi_{g2+v}(t) at
// each node's test and trial parameters (s, t); rule 1 adds the mass entries
// sum_t |Gamma_j'(x_t)| conj(phi_{g1+u}(x_t)) phi_{g2+v}(x_t) w_t.
template <int KIND, int G, class Eval>
__device__ __forceinline__ void special_basis_entry(const AsmProblem& pb, const BasisSet& bs,
                                                    const Eval& ev)
{
    const int64_t* __restrict__ special = pb.special;
    const double* __restrict__ rules = pb.rules;
    const int m1 = pb.m1, m2 = pb.m2, m3 = pb.m3, nq = pb.nq, nb = bs.nb;
    const double* __restrict__ gx = pb.g1d;
    const double* __restrict__ gw = pb.g1d + pb.nq;
    const double kscale = pb.kscale, mscale = pb.mscale;
    const int g1 = (int)(blockIdx.y / bs.ngroup) * G, g2 = (int)(blockIdx.y % bs.ngroup) * G;
    const int64_t e = blockIdx.x;
    const int64_t i = special[3 * e], j = special[3 * e + 1];
    const int rule = (int)special[3 * e + 2];
    const int ti = pb.type[i], tj = pb.type[j];
    double pi_[8], pj[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        pi_[q] = pb.param[8 * i + q];
        pj[q] = pb.param[8 * j + q];
    }
    double acc[G][G][2] = {};
    if (KIND != BEM2D_KERNEL_NONE && kscale != 0.0) {
        const double* rx;
        int m;
        if (rule == 1) {
            rx = rules;
            m = m1;
        } else if (rule == 3) {
            rx = rules + 3 * (int64_t)(m1 + m2);
            m = m3;
        } else {
            rx = rules + 3 * (int64_t)m1;
            m = m2;
        }
        const double* ry = rx + m;
        const double* rw = ry + m;
        for (int q = threadIdx.x; q < m; q += blockDim.x) {
            const double xq = rx[q], yq = ry[q], wq = rw[q];
            const int halves = rule == 1 ? 2 : 1;
            for (int h = 0; h < halves; ++h) {
                double s, t, jac;
                if (rule == 1) {
                    // Duffy split: (x, (1-y)x) and (1-x, 1-x+yx), Jacobian x (trial parameters
                    // rounded as the reference rounds them)
#pragma clang fp contract(off)
                    s = h == 0 ? xq : 1 - xq;
                    t = h == 0 ? (1 - yq) * xq : 1 - xq + yq * xq;
                    jac = xq;
                } else {
                    s = xq;
                    t = yq;
                    jac = 1.0;
                }
                const ElemPoint p1 = elem_point(ti, pi_, s);
                const ElemPoint p2 = elem_point(tj, pj, t);
                double kr, ki;
                ev.template eval<KIND>(p1.mx, p1.my, p2.mx, p2.my, p1.nx, p1.ny, p2.nx, p2.ny, kr,
                                       ki);
                const double w = jac * p1.s * p2.s * wq;
                kr *= w;
                ki *= w;
                double fr[G], fi[G];
#pragma unroll
                for (int v = 0; v < G; ++v) basis_value(bs, g2 + v, t, fr[v], fi[v]);
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    double cr, ci;
                    basis_value(bs, g1 + u, s, cr, ci);
                    ci = -ci;
                    const double tr = cr * kr - ci * ki, tim = cr * ki + ci * kr;
#pragma unroll
                    for (int v = 0; v < G; ++v) {
                        acc[u][v][0] += tr * fr[v] - tim * fi[v];
                        acc[u][v][1] += tr * fi[v] + tim * fr[v];
                    }
                }
            }
        }
    }
    __shared__ double red[4][2 * G * G];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int u = 0; u < G; ++u)
#pragma unroll
        for (int v = 0; v < G; ++v)
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const double sum = wave_sum(acc[u][v][r]);
                if (lane == 0) red[wave][2 * (u * G + v) + r] = sum;
            }
    __syncthreads();
    const int tix = threadIdx.x;
    if (tix < G * G) {
        const int u = tix / G, v = tix % G;
        if (g1 + u < nb && g2 + v < nb) {
            double vr = kscale * ((red[0][2 * tix] + red[1][2 * tix]) +
                                  (red[2][2 * tix] + red[3][2 * tix]));
            double vi = kscale * ((red[0][2 * tix + 1] + red[1][2 * tix + 1]) +
                                  (red[2][2 * tix + 1] + red[3][2 * tix + 1]));
            if (rule == 1 && mscale != 0.0) {
                double mr = 0.0, mi = 0.0;
                for (int t = 0; t < nq; ++t) {
                    double mx, my, dx, dy;
                    elem_eval(tj, pj, gx[t], mx, my, dx, dy);
                    const double s = sqrt(dx * dx + dy * dy) * gw[t];
                    double ar, ai, br, bi;
                    basis_value(bs, g1 + u, gx[t], ar, ai);
                    basis_value(bs, g2 + v, gx[t], br, bi);
                    mr += s * (ar * br + ai * bi); // conj(phi_a) phi_b
                    mi += s * (ar * bi - ai * br);
                }
                vr += mscale * mr;
                vi += mscale * mi;
            }
            pb.A[(i * nb + g1 + u) + pb.lda * (j * nb + g2 + v)] = make_double2(vr, vi);
        }
    }
}

template <int KIND, int G>
__global__ __launch_bounds__(256) void assemble_special_basis_kernel(const AsmProblem pb,
                                                                     const BasisSet bs)
{
    special_basis_entry<KIND, G>(pb, bs, RealKernelEval(pb.k, pb.eta));
}

template <int KIND, int G>
__global__ __launch_bounds__(256) void assemble_special_basis_ck_kernel(
    const AsmProblem pb, const BasisSet bs, double ki, const double* __restrict__ tab, int nt)
{
    special_basis_entry<KIND, G>(pb, bs, ComplexKernelEval{pb.k, ki, pb.eta, tab, nt});
}

// ---- right-hand sides, potential, far field ------------------------------------------------------
// B[(i nb + b) + ldb r] = scale sum_a f(x_a) conj(phi_b(t_a)) |Gamma_i'(t_a)| w_a for the plane-wave
// fields (lib/bem2d_fun.h:21-24, 45-49, 71-75) with k = kr + i ki: thread per (i, b), blockIdx.y = r.
__global__ void rhs_basis_kernel(int64_t n, int nq, const double* __restrict__ nodes,
                                 const double* __restrict__ gx, const BasisSet bs, int kind,
                                 double kr, double ki, double eta, const double* __restrict__ dirs,
                                 double scale, double2* __restrict__ B, int64_t ldb)
{
    const int nb = bs.nb;
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (u >= n * nb) return;
    const int64_t i = u / nb;
    const int b = (int)(u - i * nb);
    const double ddx = dirs[2 * r], ddy = dirs[2 * r + 1];
    const int64_t plane = n * nq;
    double ar = 0.0, ai = 0.0;
    for (int a = 0; a < nq; ++a) {
        const int64_t o = (int64_t)a * n + i;
        const double px = nodes[NODE_X * plane + o], py = nodes[NODE_Y * plane + o];
        const double sw = nodes[NODE_SW * plane + o];
        const double s = ddx * px + ddy * py;
        const double mag = ki == 0.0 ? 1.0 : exp(-ki * s);
        double sn, cs;
        sincos(kr * s, &sn, &cs);
        const double fr0 = mag * cs, fi0 = mag * sn;
        double fr, fi;
        if (kind == BEM2D_FIELD_PLANE) {
            fr = fr0;
            fi = fi0;
        } else {
            const double nd = nodes[NODE_NX * plane + o] * ddx + nodes[NODE_NY * plane + o] * ddy;
            double gr = -ki * nd, gi = kr * nd; // i k (n.d) (- i eta)
            if (kind == BEM2D_FIELD_COMBINED_PLANE) gi -= eta;
            fr = gr * fr0 - gi * fi0;
            fi = gr * fi0 + gi * fr0;
        }
        double br, bi;
        basis_value(bs, b, gx[a], br, bi);
        bi = -bi; // conj(phi_b)
        const double tr = fr * br - fi * bi, tim = fr * bi + fi * br;
        ar = fma(tr, sw, ar);
        ai = fma(tim, sw, ai);
    }
    B[u + ldb * r] = make_double2(scale * ar, scale * ai);
}

// Far-field signature with bases: F[a + nang (j nb + b)] = sum_t exp(-i k xhat_a.y_t)
// |Gamma_j'(t)| w_t phi_b(t): thread per (a, j).
__global__ void farfield_signature_basis_kernel(int64_t n, int nq, const double* __restrict__ nodes,
                                                const double* __restrict__ gx, const BasisSet bs,
                                                double k, const double* __restrict__ theta,
                                                int64_t nang, double2* __restrict__ F)
{
    extern __shared__ double bt[];
    const int nb = bs.nb, nbp = bs.G * bs.ngroup;
    stage_basis_table(bs, nq, gx, bt, threadIdx.x, blockDim.x);
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nang) return;
    const int64_t a = t % nang, j = t / nang;
    const double cx = cos(theta[a]), cy = sin(theta[a]);
    const int64_t plane = n * nq;
    double sr[kMaxBases] = {}, si[kMaxBases] = {};
    for (int q = 0; q < nq; ++q) {
        const int64_t o = (int64_t)q * n + j;
        const double ph = -k * (cx * nodes[NODE_X * plane + o] + cy * nodes[NODE_Y * plane + o]);
        const double sw = nodes[NODE_SW * plane + o];
        double sn, cs;
        sincos(ph, &sn, &cs);
        const double er = cs * sw, ei = sn * sw;
#pragma unroll
        for (int b = 0; b < kMaxBases; ++b) {
            if (b < nb) {
                const double fr = bt[2 * (q * nbp + b)], fi = bt[2 * (q * nbp + b) + 1];
                sr[b] += er * fr - ei * fi;
                si[b] += er * fi + ei * fr;
            }
        }
    }
#pragma unroll
    for (int b = 0; b < kMaxBases; ++b)
        if (b < nb) F[a + nang * (j * nb + b)] = make_double2(sr[b], si[b]);
}

// ---- host side -----------------------------------------------------------------------------------
AsmProblem basis_problem(const bem2d_mesh* m, double k, double eta, double mscale, double kscale,
                         void* A, int64_t lda)
{
    AsmProblem p{};
    p.n = m->n;
    p.nq = m->nq;
    p.m1 = m->m1;
    p.m2 = m->m2;
    p.m3 = m->m3;
    p.nodes = m->d_nodes;
    p.adj = m->d_adj;
    p.special = m->d_special;
    p.nspecial = m->nspecial;
    p.type = m->d_type;
    p.param = m->d_param;
    p.rules = m->d_rules;
    p.g1d = m->d_g1d;
    p.k = k;
    p.eta = eta;
    p.kscale = kscale;
    p.mscale = mscale;
    p.A = (double2*)A;
    p.lda = lda;
    return p;
}

size_t basis_table_bytes(const bem2d_mesh* m, const BasisSet& s)
{
    return sizeof(double) * 2 * (size_t)m->nq * s.G * s.ngroup;
}

template <int KIND, int G>
int launch_basis(bem2d_ctx* c, const AsmProblem& p, const BasisSet& s, double ki,
                 const double* tab, int nt, size_t btab)
{
    const unsigned gp = (unsigned)(s.ngroup * s.ngroup);
    hipStream_t st = c->stream;
    // the singular entries run on the second stream beside the regular pairs (disjoint entries),
    // forked from and joined back into the caller's stream
    const bool fork = p.nspecial > 0;
    if (fork) {
        if (hipEventRecord(c->ev_fork, st) != hipSuccess ||
            hipStreamWaitEvent(c->aux, c->ev_fork, 0) != hipSuccess)
            return FAIL(BEM2D_ERR_DEVICE, "stream fork failed");
        bem2d_internal_prof_begin(c, BEM2D_PROF_ASSEMBLE_SPECIAL, c->aux);
        if (ki == 0.0)
            assemble_special_basis_kernel<KIND, G>
                <<<dim3((unsigned)p.nspecial, gp), 256, 0, c->aux>>>(p, s);
        else
            assemble_special_basis_ck_kernel<KIND, G>
                <<<dim3((unsigned)p.nspecial, gp), 256, 0, c->aux>>>(p, s, ki, tab, nt);
        bem2d_internal_prof_end(c, BEM2D_PROF_ASSEMBLE_SPECIAL, c->aux);
    }
    if (ki == 0.0) {
        const size_t shm = sizeof(double) * kTileSide * node_stride(p.nq) + btab;
        bem2d_internal_prof_begin(c, BEM2D_PROF_ASSEMBLE_REGULAR, st);
        assemble_regular_basis_kernel<KIND, G>
            <<<dim3((unsigned)upper_tile_count(p.n), gp), 64, shm, st>>>(p, s);
        bem2d_internal_prof_end(c, BEM2D_PROF_ASSEMBLE_REGULAR, st);
    } else {
        bem2d_internal_prof_begin(c, BEM2D_PROF_ASSEMBLE_REGULAR, st);
        if (upper_tile_count(p.n) * gp >= kCktMinTiles)
            launch_ckt<KIND, G>(st, p, s, ki, tab, nt);
        else
            assemble_regular_basis_ck_kernel<KIND, G>
                <<<dim3((unsigned)(p.n * p.n), gp), 64, btab, st>>>(p, s, ki, tab, nt);
        bem2d_internal_prof_end(c, BEM2D_PROF_ASSEMBLE_REGULAR, st);
    }
    if (fork && (hipEventRecord(c->ev_join, c->aux) != hipSuccess ||
                 hipStreamWaitEvent(st, c->ev_join, 0) != hipSuccess))
        return FAIL(BEM2D_ERR_DEVICE, "stream join failed");
    return hipGetLastError() == hipSuccess
               ? BEM2D_OK
               : FAIL(BEM2D_ERR_DEVICE, "basis assembly launch failed");
}

template <int KIND>
int launch_basis_g(bem2d_ctx* c, const AsmProblem& p, const BasisSet& s, double ki,
                   const double* tab, int nt, size_t btab)
{
    switch (s.G) {
    case 1: return launch_basis<KIND, 1>(c, p, s, ki, tab, nt, btab);
    case 2: return launch_basis<KIND, 2>(c, p, s, ki, tab, nt, btab);
    case 3: return launch_basis<KIND, 3>(c, p, s, ki, tab, nt, btab);
    default: return launch_basis<KIND, 4>(c, p, s, ki, tab, nt, btab);
    }
}

// Arc length of the curve between parameters a < b: adaptive Gauss-Legendre, 20 nodes per panel,
// a panel halved until its two halves agree with it to 1e-15 relative.
struct ArcLength {
    int type;
    double c[4];
    double x[20], w[20];
    explicit ArcLength(int t, const double* args) : type(t)
    {
        for (int i = 0; i < 4; ++i) c[i] = args[i];
        bem2d_gauss_legendre(20, x, w);
    }
    double speed(double t) const
    {
        double mx, my, dx, dy;
        curve_eval(type, c, t, mx, my, dx, dy);
        return std::sqrt(dx * dx + dy * dy);
    }
    double panel(double a, double b) const
    {
        double s = 0.0;
        for (int i = 0; i < 20; ++i) s += w[i] * speed(a + (b - a) * x[i]);
        return s * (b - a);
    }
    double adapt(double a, double b, double whole, int depth) const
    {
        const double m = 0.5 * (a + b);
        const double l = panel(a, m), r = panel(m, b);
        if (depth >= 30 || std::fabs(l + r - whole) <= 1e-15 * std::fabs(l + r)) return l + r;
        return adapt(a, m, l, depth + 1) + adapt(m, b, r, depth + 1);
    }
    double length(double a, double b) const
    {
        if (!(b > a)) return 0.0;
        double s = 0.0; // 16 panels first: a periodic integrand is resolved before the test
        for (int p = 0; p < 16; ++p) {
            const double pa = a + (b - a) * p / 16.0, pb = a + (b - a) * (p + 1) / 16.0;
            s += adapt(pa, pb, panel(pa, pb), 0);
        }
        return s;
    }
};

bool curve_ok(int type) { return type >= BEM2D_ELEM_CIRCLE && type <= BEM2D_ELEM_INV_ELLIPSE; }

} // namespace

// Validated device descriptor of a basis list: groups of at most kGroup bases, G per group.
int bem2d_dev::bem2d_internal_make_basis_set(const bem2d_basis* bases, int nb, BasisSet* out)
{
    if (!bases || nb < 1 || nb > kMaxBases)
        return FAIL(BEM2D_ERR_ARG, "basis list: need 1 <= nb <= %d (got %d)", kMaxBases, nb);
    BasisSet s{};
    s.nb = nb;
    s.ngroup = (nb + kGroup - 1) / kGroup;
    s.G = (nb + s.ngroup - 1) / s.ngroup;
    for (int b = 0; b < nb; ++b) {
        const bem2d_basis& x = bases[b];
        if (x.kind < BEM2D_BASIS_CONST || x.kind > BEM2D_BASIS_WAVEPOL)
            return FAIL(BEM2D_ERR_ARG, "basis %d: unknown kind %d", b, x.kind);
        if ((x.kind == BEM2D_BASIS_POL || x.kind == BEM2D_BASIS_WAVEPOL) &&
            (x.degree < 0 || x.degree > 64))
            return FAIL(BEM2D_ERR_ARG, "basis %d: degree %d outside [0, 64]", b, x.degree);
        if ((x.kind == BEM2D_BASIS_WAVE || x.kind == BEM2D_BASIS_WAVEPOL) &&
            !(std::isfinite(x.direction) && std::isfinite(x.k_re) && std::isfinite(x.k_im)))
            return FAIL(BEM2D_ERR_ARG, "basis %d: direction or wavenumber not finite", b);
        s.kind[b] = x.kind;
        s.deg[b] = (x.kind == BEM2D_BASIS_POL || x.kind == BEM2D_BASIS_WAVEPOL || x.kind == BEM2D_BASIS_LIN)
                       ? x.degree
                       : 0;
        s.dir[b] = x.direction;
        s.kr[b] = x.k_re;
        s.ki[b] = x.k_im;
    }
    *out = s;
    return BEM2D_OK;
}

// The regular pairs of a P0 complex-k assembly through the tiled kernel (one ConstBasis), when the
// mesh gives enough tiles; returns false (nothing launched) otherwise.
bool bem2d_dev::bem2d_internal_regular_ck_tiles(hipStream_t st, const AsmProblem& p, int kind,
                                                double ki, const double* tab, int nt)
{
    if (upper_tile_count(p.n) < kCktMinTiles) return false;
    const bem2d_basis c0{BEM2D_BASIS_CONST, 0, 0.0, 0.0, 0.0};
    BasisSet bs;
    if (bem2d_internal_make_basis_set(&c0, 1, &bs)) return false;
    switch (kind) {
    case BEM2D_KERNEL_SL: launch_ckt<BEM2D_KERNEL_SL, 1>(st, p, bs, ki, tab, nt); return true;
    case BEM2D_KERNEL_DL: launch_ckt<BEM2D_KERNEL_DL, 1>(st, p, bs, ki, tab, nt); return true;
    case BEM2D_KERNEL_CDL: launch_ckt<BEM2D_KERNEL_CDL, 1>(st, p, bs, ki, tab, nt); return true;
    case BEM2D_KERNEL_COMBINED:
        launch_ckt<BEM2D_KERNEL_COMBINED, 1>(st, p, bs, ki, tab, nt);
        return true;
    default: return false;
    }
}

extern "C" int bem2d_assemble_basis(bem2d_ctx* c, const bem2d_mesh* m, const bem2d_basis* bases,
                                    int nb, int kind, double k_re, double k_im, double eta,
                                    double mscale, double kscale, void* A, int64_t lda)
{
    if (!c || !m || !A) return FAIL(BEM2D_ERR_ARG, "bem2d_assemble_basis: null argument");
    BasisSet s;
    if (int rc = bem2d_internal_make_basis_set(bases, nb, &s)) return rc;
    if (lda < m->n * nb)
        return FAIL(BEM2D_ERR_ARG, "bem2d_assemble_basis: lda %lld < n nb %lld", (long long)lda,
                    (long long)(m->n * nb));
    if (kind < BEM2D_KERNEL_NONE || kind > BEM2D_KERNEL_ID)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "unknown kernel kind %d", kind);
    const bool haswave = kind != BEM2D_KERNEL_NONE && kind != BEM2D_KERNEL_ID;
    if (!haswave) { // no wavenumber in these: the real-k kernels with k = 1
        k_re = 1.0;
        k_im = 0.0;
    }
    if (k_im == 0.0 && !(k_re > 0.0))
        return FAIL(BEM2D_ERR_ARG, "bem2d_assemble_basis: wavenumber must be positive (got %g)",
                    k_re);
    if (k_im != 0.0 && !(k_re > 0.0 && std::isfinite(k_re) && std::isfinite(k_im)))
        return FAIL(BEM2D_ERR_UNSUPPORTED, "complex wavenumber needs Re k > 0 (got %g%+gi)", k_re,
                    k_im);
    const int64_t gp = (int64_t)s.ngroup * s.ngroup;
    if (gp > 65535 || m->nspecial > (int64_t)UINT32_MAX ||
        (k_im != 0.0 && m->n * m->n * 64 > (int64_t)UINT32_MAX) ||
        upper_tile_count(m->n) * 64 > (int64_t)UINT32_MAX)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "bem2d_assemble_basis: mesh too large for one launch");
    HIPCHK(hipSetDevice(c->device));
    const double* tab = nullptr;
    int nt = 0;
    if (k_im != 0.0)
        if (int rc = bem2d_internal_upload_contour(c, k_re, k_im, &tab, &nt)) return rc;
    const AsmProblem p = basis_problem(m, k_re, eta, mscale, kscale, A, lda);
    const size_t btab = basis_table_bytes(m, s);
    switch (kind) {
    case BEM2D_KERNEL_NONE: return launch_basis_g<BEM2D_KERNEL_NONE>(c, p, s, 0.0, tab, nt, btab);
    case BEM2D_KERNEL_SL: return launch_basis_g<BEM2D_KERNEL_SL>(c, p, s, k_im, tab, nt, btab);
    case BEM2D_KERNEL_DL: return launch_basis_g<BEM2D_KERNEL_DL>(c, p, s, k_im, tab, nt, btab);
    case BEM2D_KERNEL_CDL: return launch_basis_g<BEM2D_KERNEL_CDL>(c, p, s, k_im, tab, nt, btab);
    case BEM2D_KERNEL_COMBINED:
        return launch_basis_g<BEM2D_KERNEL_COMBINED>(c, p, s, k_im, tab, nt, btab);
    default: return launch_basis_g<BEM2D_KERNEL_ID>(c, p, s, 0.0, tab, nt, btab);
    }
}

extern "C" int bem2d_rhs_basis(bem2d_ctx* c, const bem2d_mesh* m, const bem2d_basis* bases, int nb,
                               int kind, double k_re, double k_im, double eta, const double* dirs,
                               int nrhs, double scale, void* B, int64_t ldb)
{
    if (!c || !m || !B || (!dirs && nrhs > 0))
        return FAIL(BEM2D_ERR_ARG, "bem2d_rhs_basis: null argument");
    BasisSet s;
    if (int rc = bem2d_internal_make_basis_set(bases, nb, &s)) return rc;
    if (nrhs < 0 || nrhs > 65535 || ldb < m->n * nb)
        return FAIL(BEM2D_ERR_ARG, "bem2d_rhs_basis: bad sizes");
    if (kind < BEM2D_FIELD_PLANE || kind > BEM2D_FIELD_COMBINED_PLANE)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "unknown field kind %d", kind);
    if (!std::isfinite(k_re) || !std::isfinite(k_im))
        return FAIL(BEM2D_ERR_ARG, "bem2d_rhs_basis: wavenumber not finite");
    if (nrhs == 0) return BEM2D_OK;
    if ((m->n * nb + 255) / 256 > (int64_t)UINT32_MAX / 256)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "bem2d_rhs_basis: mesh too large for one launch");
    HIPCHK(hipSetDevice(c->device));
    const double* d_dirs = nullptr;
    if (int rc = bem2d_internal_stage_dirs(c, dirs, nrhs, &d_dirs)) return rc;
    dim3 grid((unsigned)((m->n * nb + 255) / 256), (unsigned)nrhs);
    rhs_basis_kernel<<<grid, 256, 0, c->stream>>>(m->n, m->nq, m->d_nodes, m->d_g1d, s, kind, k_re,
                                                  k_im, eta, d_dirs, scale, (double2*)B, ldb);
    HIPCHK(hipGetLastError());
    return BEM2D_OK;
}

extern "C" int bem2d_eval_single_layer_basis(bem2d_ctx* c, const bem2d_mesh* m,
                                             const bem2d_basis* bases, int nb, double k_re,
                                             double k_im, const void* coef, const double* pts,
                                             int64_t npts, void* out)
{
    if (!c || !m || !coef || (!pts && npts > 0) || (!out && npts > 0))
        return FAIL(BEM2D_ERR_ARG, "bem2d_eval_single_layer_basis: null argument");
    BasisSet s;
    if (int rc = bem2d_internal_make_basis_set(bases, nb, &s)) return rc;
    if (k_im == 0.0 ? !(k_re > 0.0) : !(k_re > 0.0 && std::isfinite(k_re) && std::isfinite(k_im)))
        return FAIL(BEM2D_ERR_UNSUPPORTED, "bem2d_eval_single_layer_basis: needs Re k > 0 (got %g%+gi)",
                    k_re, k_im);
    if (npts <= 0) return BEM2D_OK;
    HIPCHK(hipSetDevice(c->device));
    return bem2d_internal_potential(c, m, s, k_re, k_im, coef, pts, npts, out);
}

extern "C" int bem2d_farfield_basis(bem2d_ctx* c, const bem2d_mesh* m, const bem2d_basis* bases,
                                    int nb, double k, const void* coef, int64_t ldc, int nrhs,
                                    const double* theta, int64_t nang, void* out)
{
    if (!c || !m || !coef || !theta || !out)
        return FAIL(BEM2D_ERR_ARG, "bem2d_farfield_basis: null argument");
    BasisSet s;
    if (int rc = bem2d_internal_make_basis_set(bases, nb, &s)) return rc;
    const int64_t nn = m->n * nb;
    if (ldc < nn || nrhs < 1 || nrhs > 65535 || nang < 1 || nang >= (1 << 24) || !(k > 0))
        return FAIL(BEM2D_ERR_ARG, "bem2d_farfield_basis: bad sizes");
    const int64_t tot = nang * m->n;
    if (tot > (int64_t)UINT32_MAX - 256)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "far field: %lld angles x %lld elements exceed one launch",
                    (long long)nang, (long long)m->n);
    HIPCHK(hipSetDevice(c->device));
    double2* F = (double2*)bem2d_internal_scratch(c, SCRATCH_FARFIELD, sizeof(double2) * nang * nn);
    if (!F) return FAIL(BEM2D_ERR_NOMEM, "scratch allocation failed");
    const size_t btab = basis_table_bytes(m, s);
    farfield_signature_basis_kernel<<<(unsigned)((tot + 255) / 256), 256, btab, c->stream>>>(
        m->n, m->nq, m->d_nodes, m->d_g1d, s, k, theta, nang, F);
    HIPCHK(hipGetLastError());
    return bem2d_internal_farfield_gemm(c, nn, k, F, nang, coef, ldc, nrhs, out);
}

extern "C" int bem2d_basis_values(const bem2d_basis* bases, int nb, const double* t, int64_t count,
                                  double* out)
{
    BasisSet s;
    if (int rc = bem2d_internal_make_basis_set(bases, nb, &s)) return rc;
    if (count < 0 || (count > 0 && (!t || !out)))
        return FAIL(BEM2D_ERR_ARG, "bem2d_basis_values: bad arguments");
    for (int64_t q = 0; q < count; ++q)
        for (int b = 0; b < nb; ++b)
            basis_value(s, b, t[q], out[2 * (q * nb + b)], out[2 * (q * nb + b) + 1]);
    return BEM2D_OK;
}

extern "C" int bem2d_curve_length(int curve_type, const double* args, double t0, double t1,
                                  double* out)
{
    if (!args || !out || !curve_ok(curve_type) || !(t0 <= t1))
        return FAIL(BEM2D_ERR_ARG, "bem2d_curve_length: bad arguments");
    *out = ArcLength(curve_type, args).length(t0, t1);
    return BEM2D_OK;
}

extern "C" int bem2d_arclength_params(int curve_type, const double* args, const double* s,
                                      int64_t count, double* t)
{
    if (!args || !curve_ok(curve_type) || count < 0 || (count > 0 && (!s || !t)))
        return FAIL(BEM2D_ERR_ARG, "bem2d_arclength_params: bad arguments");
    const ArcLength arc(curve_type, args);
    const double L = arc.length(0.0, 1.0);
    double tprev = 0.0, sprev = 0.0;
    for (int64_t i = 0; i < count; ++i) {
        const double target = s[i];
        if (!(target >= sprev) || !(target <= L * (1 + 1e-12)))
            return FAIL(BEM2D_ERR_ARG,
                        "bem2d_arclength_params: arc lengths must increase within [0, %.17g] "
                        "(s[%lld] = %.17g)",
                        L, (long long)i, target);
        if (target >= L) {
            t[i] = 1.0;
            continue;
        }
        // Newton on s(x) - target (s' = |Gamma'|), kept inside the bracket [lo, hi] by bisection
        double lo = tprev, hi = 1.0, x = tprev + (target - sprev) / L;
        if (!(x > lo && x < hi)) x = 0.5 * (lo + hi);
        for (int it = 0; it < 200; ++it) {
            const double f = sprev + arc.length(tprev, x) - target;
            if (f > 0)
                hi = x;
            else
                lo = x;
            double xn = x - f / arc.speed(x);
            if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
            if (std::fabs(xn - x) <= 1e-16 * std::fabs(x) + 1e-300) {
                x = xn;
                break;
            }
            x = xn;
        }
        sprev += arc.length(tprev, x);
        tprev = x;
        t[i] = x;
    }
    return BEM2D_OK;
}
