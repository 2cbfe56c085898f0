// Single-layer potential at points: EvalKernel<SingleLayer> (lib/bem2d_quadrature.h:184-228) as
// EvalSol uses it (lib/bem2d_soundsoftscattering.h:149-168), for P0, several bases per element and
// complex wavenumbers alike:
//   out[p] = sum_{j, a} (i/4) H0(k |P_p - y_ja|) q_ja,   q_ja = |Gamma_j'(t_a)| w_a sum_b c_{j nb + b} phi_b(t_a)
// over the Gauss nodes y_ja of every element.
//
// It is a dense kernel-times-vector product of npts x (n nq) kernel values generated on the fly:
// pass 1 (density_kernel) folds the coefficients, bases and quadrature weights into one complex
// weight per source node; pass 2 (potential_partial_kernel) gives each thread one point and each
// workgroup one chunk of the source nodes, staged through LDS 256 at a time (every lane reads the
// same LDS word: a broadcast), so that few points still fill the device; pass 3 adds the chunks'
// partial sums in chunk order (deterministic, no atomics).
#include <algorithm>
#include <cmath>

#pragma clang diagnostic ignored "-Wunused-function" // kernels of the shared headers
#include "assembly_kernels.h"
#include "basis.h"
#include "complex_kernels.h"
#include "internal.h"

using namespace bem2d_dev;

namespace {

constexpr int kPotThreads = 256;

// q[a n + j] = |Gamma_j'(t_a)| w_a sum_b coef[j nb + b] phi_b(t_a) (the node-plane order)
__global__ void density_kernel(int64_t n, int nq, const double* __restrict__ nodes,
                               const double* __restrict__ gx, const BasisSet bs,
                               const double2* __restrict__ coef, double2* __restrict__ q)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n * nq) return;
    const int64_t a = s / n, j = s - a * n;
    const int nb = bs.nb;
    double dr = 0.0, di = 0.0;
    for (int b = 0; b < nb; ++b) {
        double fr, fi;
        basis_value(bs, b, gx[a], fr, fi);
        const double2 c = coef[j * nb + b];
        dr += c.x * fr - c.y * fi;
        di += c.x * fi + c.y * fr;
    }
    const double sw = nodes[NODE_SW * n * nq + s];
    q[s] = make_double2(dr * sw, di * sw);
}

// partial[c npts + p] = sum over the source nodes of chunk c of K(P_p, y) q(y)
template <bool CK>
__global__ __launch_bounds__(kPotThreads) void potential_partial_kernel(
    int64_t nsrc, const double* __restrict__ sx, const double* __restrict__ sy,
    const double2* __restrict__ q, int64_t chunk, double kr, double ki,
    const double* __restrict__ tab, int nt, const double* __restrict__ pts, int64_t npts,
    double2* __restrict__ partial)
{
    __shared__ double lx[kPotThreads], ly[kPotThreads], lqr[kPotThreads], lqi[kPotThreads];
    const int tid = threadIdx.x;
    const int64_t p = (int64_t)blockIdx.x * kPotThreads + tid;
    const int64_t c = blockIdx.y;
    const int64_t s0 = c * chunk, s1 = std::min(s0 + chunk, nsrc);
    const bool live = p < npts;
    const double px = live ? pts[2 * p] : 0.0, py = live ? pts[2 * p + 1] : 0.0;
    const double kinv = 1.0 / kr, k4 = 0.25 * kr;
    double vr = 0.0, vi = 0.0;
    for (int64_t base = s0; base < s1; base += kPotThreads) {
        __syncthreads(); // the previous tile has been read
        const int64_t s = base + tid;
        if (s < s1) {
            lx[tid] = sx[s];
            ly[tid] = sy[s];
            const double2 w = q[s];
            lqr[tid] = w.x;
            lqi[tid] = w.y;
        }
        __syncthreads();
        const int cnt = (int)std::min<int64_t>(kPotThreads, s1 - base);
        if (live) {
            for (int u = 0; u < cnt; ++u) {
                double kvr, kvi;
                if (CK)
                    kernel_eval_ck<BEM2D_KERNEL_SL>(kr, ki, 0.0, tab, nt, px, py, lx[u], ly[u], 0,
                                                    0, 0, 0, kvr, kvi);
                else
                    kernel_eval<BEM2D_KERNEL_SL>(kr, kinv, k4, 0.0, px, py, lx[u], ly[u], 0, 0, 0,
                                                 0, kvr, kvi);
                vr = fma(kvr, lqr[u], fma(-kvi, lqi[u], vr));
                vi = fma(kvr, lqi[u], fma(kvi, lqr[u], vi));
            }
        }
    }
    if (live) partial[c * npts + p] = make_double2(vr, vi);
}

__global__ void potential_reduce_kernel(int64_t npts, int nchunk,
                                        const double2* __restrict__ partial,
                                        double2* __restrict__ out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    double vr = 0.0, vi = 0.0;
    for (int c = 0; c < nchunk; ++c) {
        const double2 v = partial[(int64_t)c * npts + p];
        vr += v.x;
        vi += v.y;
    }
    out[p] = make_double2(vr, vi);
}

} // namespace

// The potential of the density sum_b coef[j nb + b] phi_b on the mesh at npts points (device
// arrays), real (k_im == 0) or complex k.  Scratch: SCRATCH_EVAL.
int bem2d_internal_potential(bem2d_ctx* c, const bem2d_mesh* m, const BasisSet& bs, double k_re,
                             double k_im, const void* coef, const double* pts, int64_t npts,
                             void* out)
{
    if (npts <= 0) return BEM2D_OK;
    const int64_t nsrc = m->n * m->nq;
    const int64_t ptiles = (npts + kPotThreads - 1) / kPotThreads;
    if (ptiles > (int64_t)INT32_MAX || nsrc / kPotThreads > (int64_t)INT32_MAX)
        return FAIL(BEM2D_ERR_UNSUPPORTED, "potential: %lld points exceed one launch", (long long)npts);
    // about four workgroups per compute unit: chunks of the sources when the points are few
    const int64_t want = std::max<int64_t>(1, (1024 + ptiles - 1) / ptiles);
    const int64_t maxchunks = std::max<int64_t>(1, (nsrc + kPotThreads - 1) / kPotThreads);
    int64_t nchunk = std::min<int64_t>({want, maxchunks, 4096});
    int64_t chunk = (nsrc + nchunk - 1) / nchunk;
    chunk = (chunk + kPotThreads - 1) / kPotThreads * kPotThreads;
    nchunk = (nsrc + chunk - 1) / chunk;
    const size_t qbytes = sizeof(double2) * (size_t)nsrc;
    const size_t pbytes = sizeof(double2) * (size_t)nchunk * (size_t)npts;
    char* scratch = (char*)bem2d_internal_scratch(c, SCRATCH_EVAL, qbytes + pbytes + 256);
    if (!scratch) return FAIL(BEM2D_ERR_NOMEM, "scratch allocation failed");
    double2* q = (double2*)scratch;
    double2* partial = (double2*)(scratch + (qbytes + 255) / 256 * 256);
    const double* tab = nullptr;
    int nt = 0;
    if (k_im != 0.0)
        if (int rc = bem2d_internal_upload_contour(c, k_re, k_im, &tab, &nt)) return rc;
    hipStream_t s = c->stream;
    density_kernel<<<(unsigned)((nsrc + 255) / 256), 256, 0, s>>>(m->n, m->nq, m->d_nodes,
                                                                   m->d_g1d, bs,
                                                                   (const double2*)coef, q);
    const double* sx = m->d_nodes + (int64_t)NODE_X * nsrc;
    const double* sy = m->d_nodes + (int64_t)NODE_Y * nsrc;
    dim3 grid((unsigned)ptiles, (unsigned)nchunk);
    if (k_im == 0.0)
        potential_partial_kernel<false><<<grid, kPotThreads, 0, s>>>(
            nsrc, sx, sy, q, chunk, k_re, 0.0, nullptr, 0, pts, npts, partial);
    else
        potential_partial_kernel<true><<<grid, kPotThreads, 0, s>>>(
            nsrc, sx, sy, q, chunk, k_re, k_im, tab, nt, pts, npts, partial);
    potential_reduce_kernel<<<(unsigned)((npts + 255) / 256), 256, 0, s>>>(npts, (int)nchunk,
                                                                          partial, (double2*)out);
    return hipGetLastError() == hipSuccess ? BEM2D_OK
                                           : FAIL(BEM2D_ERR_DEVICE, "potential launch failed");
}
