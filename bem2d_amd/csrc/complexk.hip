// Complex wavenumbers: the kernels' `k_.im != 0` branches (lib/bem2d_kernel.cpp:15-112), which call
// BesselH0(complex) / BesselH1(complex) (AMOS zbesh, lib/bem2d_mathroutines.cpp:44-75), and the
// plane-wave fields with complex k (lib/bem2d_fun.h:21-24, 45-49, 71-75).
//
// Regular pairs: one wavefront per matrix entry, the nq x nq Gauss2D node pairs spread over the
// lanes and summed by lane exchange.  Singular pairs: the rule / Duffy body of the real-k kernel
// with complex kernel values.  Every complex Hankel value comes from hankel01_complex; the contour
// table of arg k is built on the host once per call.  This path is written for correctness over
// the domain Re k > 0 (the resonance scans of the reference sit in Re k in [4, 8], Im k in [-2, 0]),
// not for the throughput of the real-k kernels.
#include <cmath>
#include <cstring>

#pragma clang diagnostic ignored "-Wunused-function" // real-k kernels of the shared header
#include "assembly_kernels.h"
#include "basis.h"
#include "complex_kernels.h"
#include "internal.h"

using namespace bem2d_dev;

namespace {

template <int KIND>
__global__ __launch_bounds__(256) void assemble_special_ck_kernel(const AsmBatch batch, double ki,
                                                                  const double* __restrict__ tab,
                                                                  int nt)
{
    const AsmProblem& pb = batch.p[0];
    special_entry<KIND>(batch, ComplexKernelEval{pb.k, ki, pb.eta, tab, nt});
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Regular entry (i, j) = (e mod n, e div n) for workgroup e; entries of the singular list are
// left to assemble_special_ck_kernel (same precedence test as the real-k regular kernel).
template <int KIND>
__global__ __launch_bounds__(64) void assemble_regular_ck_kernel(const AsmProblem p, double ki,
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
    const int nq = p.nq;
    const int64_t plane = n * nq;
    const double* __restrict__ nodes = p.nodes;
    double sr = 0.0, si = 0.0;
    for (int q = threadIdx.x; q < nq * nq; q += 64) {
        const int a = q / nq, b = q % nq; // Gauss2D node (x_a, y_b), lib/bem2d_quadpoints.cpp:43-47
        const int64_t ox = (int64_t)a * n + i, oy = (int64_t)b * n + j;
        double vr, vi;
        kernel_eval_ck<KIND>(p.k, ki, p.eta, tab, nt, nodes[NODE_X * plane + ox],
                             nodes[NODE_Y * plane + ox], nodes[NODE_X * plane + oy],
                             nodes[NODE_Y * plane + oy], nodes[NODE_NX * plane + ox],
                             nodes[NODE_NY * plane + ox], nodes[NODE_NX * plane + oy],
                             nodes[NODE_NY * plane + oy], vr, vi);
        const double s = nodes[NODE_SW * plane + ox] * nodes[NODE_SW * plane + oy];
        sr += vr * s;
        si += vi * s;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (threadIdx.x == 0) p.A[i + p.lda * j] = make_double2(p.kscale * sr, p.kscale * si);
}

// DotProdBasFuns for the plane-wave fields with complex k: f = exp(i k d.p).
__global__ void rhs_ck_kernel(int64_t n, int nq, const double* __restrict__ nodes, int kind,
                              double kr, double ki, double eta, const double* __restrict__ dirs,
                              double scale, double2* __restrict__ B, int64_t ldb)
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
        const double s = ddx * px + ddy * py;
        // i k s = -ki s + i kr s
        const double mag = exp(-ki * s);
        double sn, cs;
        sincos(kr * s, &sn, &cs);
        const double fr0 = mag * cs, fi0 = mag * sn;
        double fr, fi;
        if (kind == BEM2D_FIELD_PLANE) {
            fr = fr0;
            fi = fi0;
        } else {
            const double nd = nodes[NODE_NX * plane + o] * ddx + nodes[NODE_NY * plane + o] * ddy;
            // i k f nd (- i eta f):  i k = (-ki, kr)
            double gr = -ki * nd, gi = kr * nd;
            if (kind == BEM2D_FIELD_COMBINED_PLANE) gi -= eta;
            fr = gr * fr0 - gi * fi0;
            fi = gr * fi0 + gi * fr0;
        }
        ar = fma(fr, sw, ar);
        ai = fma(fi, sw, ai);
    }
    B[i + ldb * r] = make_double2(scale * ar, scale * ai);
}

__global__ void hankel01_complex_kernel(const double* __restrict__ tab, int nt,
                                        const double* __restrict__ z, int64_t n,
                                        double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a, b, c, d;
    hankel01_complex(z[2 * i], z[2 * i + 1], tab, nt, a, b, c, d);
    out[4 * i] = a;
    out[4 * i + 1] = b;
    out[4 * i + 2] = c;
    out[4 * i + 3] = d;
}

bool wavenumber_ok(double kr, double ki) { return kr > 0.0 && std::isfinite(kr) && std::isfinite(ki); }

int upload_contour(bem2d_ctx* c, double kr, double ki, const double** d_tab, int* nt)
{
    return bem2d_internal_upload_contour(c, kr, ki, d_tab, nt);
}

} // namespace

// Contour table of arg k in device scratch, uploaded in stream order on c->stream.
int bem2d_internal_upload_contour(bem2d_ctx* c, double kr, double ki, const double** d_tab,
                                  int* nt)
{
    double tab[kContourMaxNodes * kContourDoubles];
    *nt = contour_table(std::atan2(ki, kr), tab);
    const size_t bytes = sizeof(double) * kContourDoubles * (size_t)*nt;
    double* d = (double*)bem2d_internal_scratch(c, SCRATCH_CONTOUR, sizeof(tab));
    if (!d) return FAIL(BEM2D_ERR_NOMEM, "scratch allocation failed");
    HIPCHK(hipMemcpyAsync(d, tab, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream)); // tab lives on this call's stack
    *d_tab = d;
    return BEM2D_OK;
}

namespace {

template <int KIND>
int launch_ck(bem2d_ctx* c, const bem2d_mesh* m, double kr, double ki, double eta, double mscale,
              double kscale, double2* A, int64_t lda, const double* tab, int nt)
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
    p.k = kr;
    p.eta = eta;
    p.kscale = kscale;
    p.mscale = mscale;
    p.A = A;
    p.lda = lda;
    AsmBatch sp{};
    sp.count = 1;
    sp.p[0] = p;
    sp.start[1] = m->nspecial;
    if (m->nspecial > 0)
        assemble_special_ck_kernel<KIND><<<(unsigned)m->nspecial, 256, 0, c->stream>>>(sp, ki, tab,
                                                                                      nt);
    if (!bem2d_internal_regular_ck_tiles(c->stream, p, KIND, ki, tab, nt))
        assemble_regular_ck_kernel<KIND><<<(unsigned)(m->n * m->n), 64, 0, c->stream>>>(p, ki, tab,
                                                                                      nt);
    HIPCHK(hipGetLastError());
    return BEM2D_OK;
}

} // namespace

extern "C" int bem2d_assemble_complex(bem2d_ctx* c, const bem2d_mesh* m, int kind, double k_re,
                                      double k_im, double eta, double mscale, double kscale,
                                      void* A, int64_t lda)
{
    if (k_im == 0.0) return bem2d_assemble(c, m, kind, k_re, eta, mscale, kscale, A, lda);
    if (!c || !m || !A) return FAIL(BEM2D_ERR_ARG, "bem2d_assemble_complex: null argument");
    if (lda < m->n) return FAIL(BEM2D_ERR_ARG, "bem2d_assemble_complex: lda < n");
    if (kind == BEM2D_KERNEL_NONE || kind == BEM2D_KERNEL_ID) // no wavenumber in these
        return bem2d_assemble(c, m, kind, 1.0, eta, mscale, kscale, A, lda);
    if (!wavenumber_ok(k_re, k_im))
        return FAIL(BEM2D_ERR_UNSUPPORTED, "complex wavenumber needs Re k > 0 (got %g%+gi)", k_re,
                    k_im);
    if (m->n * m->n * 64 > (int64_t)UINT32_MAX) // one 64-lane workgroup per entry
        return FAIL(BEM2D_ERR_UNSUPPORTED, "complex-k assembly is limited to n <= 8191");
    HIPCHK(hipSetDevice(c->device));
    const double* tab = nullptr;
    int nt = 0;
    if (int rc = upload_contour(c, k_re, k_im, &tab, &nt)) return rc;
    double2* a = (double2*)A;
    switch (kind) {
    case BEM2D_KERNEL_SL:
        return launch_ck<BEM2D_KERNEL_SL>(c, m, k_re, k_im, eta, mscale, kscale, a, lda, tab, nt);
    case BEM2D_KERNEL_DL:
        return launch_ck<BEM2D_KERNEL_DL>(c, m, k_re, k_im, eta, mscale, kscale, a, lda, tab, nt);
    case BEM2D_KERNEL_CDL:
        return launch_ck<BEM2D_KERNEL_CDL>(c, m, k_re, k_im, eta, mscale, kscale, a, lda, tab, nt);
    case BEM2D_KERNEL_COMBINED:
        return launch_ck<BEM2D_KERNEL_COMBINED>(c, m, k_re, k_im, eta, mscale, kscale, a, lda, tab,
                                                nt);
    default:
        return FAIL(BEM2D_ERR_UNSUPPORTED, "unknown kernel kind %d", kind);
    }
}

extern "C" int bem2d_rhs_complex(bem2d_ctx* c, const bem2d_mesh* m, int kind, double k_re,
                                 double k_im, double eta, const double* dirs, int nrhs,
                                 double scale, void* B, int64_t ldb)
{
    if (k_im == 0.0) return bem2d_rhs(c, m, kind, k_re, eta, dirs, nrhs, scale, B, ldb);
    if (!c || !m || !B || (!dirs && nrhs > 0))
        return FAIL(BEM2D_ERR_ARG, "bem2d_rhs_complex: null argument");
    if (nrhs < 0 || nrhs > 65535 || ldb < m->n) return FAIL(BEM2D_ERR_ARG, "bem2d_rhs_complex: bad sizes");
    if (kind < BEM2D_FIELD_PLANE || kind > BEM2D_FIELD_COMBINED_PLANE)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "unknown field kind %d", kind);
    if (!std::isfinite(k_re) || !std::isfinite(k_im))
        return FAIL(BEM2D_ERR_ARG, "bem2d_rhs_complex: wavenumber not finite");
    if (nrhs == 0) return BEM2D_OK;
    HIPCHK(hipSetDevice(c->device));
    const double* d_dirs = nullptr;
    if (int rc = bem2d_internal_stage_dirs(c, dirs, nrhs, &d_dirs)) return rc;
    dim3 grid((unsigned)((m->n + 255) / 256), (unsigned)nrhs);
    rhs_ck_kernel<<<grid, 256, 0, c->stream>>>(m->n, m->nq, m->d_nodes, kind, k_re, k_im, eta,
                                               d_dirs, scale, (double2*)B, ldb);
    HIPCHK(hipGetLastError());
    return BEM2D_OK;
}

extern "C" int bem2d_eval_single_layer_complex(bem2d_ctx* c, const bem2d_mesh* m, double k_re,
                                               double k_im, const void* coef, const double* pts,
                                               int64_t npts, void* out)
{
    if (k_im == 0.0) return bem2d_eval_single_layer(c, m, k_re, coef, pts, npts, out);
    if (!c || !m || !coef || (!pts && npts > 0) || (!out && npts > 0))
        return FAIL(BEM2D_ERR_ARG, "bem2d_eval_single_layer_complex: null argument");
    if (!wavenumber_ok(k_re, k_im))
        return FAIL(BEM2D_ERR_UNSUPPORTED, "complex wavenumber needs Re k > 0 (got %g%+gi)", k_re,
                    k_im);
    if (npts <= 0) return BEM2D_OK;
    HIPCHK(hipSetDevice(c->device));
    const bem2d_basis p0{BEM2D_BASIS_CONST, 0, 0.0, 0.0, 0.0};
    BasisSet bs;
    if (int rc = bem2d_internal_make_basis_set(&p0, 1, &bs)) return rc;
    return bem2d_internal_potential(c, m, bs, k_re, k_im, coef, pts, npts, out);
}

extern "C" int bem2d_hankel01_complex(bem2d_ctx* c, double phi, const double* z, int64_t n,
                                      double* out)
{
    if (!c || n < 0 || (n > 0 && (!z || !out)) || !(std::fabs(phi) < 1.5707963267948966))
        return FAIL(BEM2D_ERR_ARG, "bem2d_hankel01_complex: bad argument");
    if (n == 0) return BEM2D_OK;
    HIPCHK(hipSetDevice(c->device));
    const double* tab = nullptr;
    int nt = 0;
    if (int rc = upload_contour(c, std::cos(phi), std::sin(phi), &tab, &nt)) return rc;
    hankel01_complex_kernel<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(tab, nt, z, n, out);
    HIPCHK(hipGetLastError());
    return BEM2D_OK;
}

extern "C" int bem2d_hankel01_complex_host(double phi, const double* z, int64_t n, double* out)
{
    if (n < 0 || (n > 0 && (!z || !out)) || !(std::fabs(phi) < 1.5707963267948966))
        return FAIL(BEM2D_ERR_ARG, "bem2d_hankel01_complex_host: bad argument");
    double tab[kContourMaxNodes * kContourDoubles];
    const int nt = contour_table(phi, tab);
    for (int64_t i = 0; i < n; ++i)
        hankel01_complex(z[2 * i], z[2 * i + 1], tab, nt, out[4 * i], out[4 * i + 1],
                         out[4 * i + 2], out[4 * i + 3]);
    return BEM2D_OK;
}
