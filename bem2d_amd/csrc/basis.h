// Basis descriptors shared by the basis kernels and the potential: ConstBasis, PolBasis, LinBasis,
// WaveBasis, WavePolBasis (lib/bem2d_basistypes.h:16-99) as plain data for kernel arguments.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/bem2d_hip.h"

namespace bem2d_dev {

constexpr int kBasisGroup = 4; // bases per group in one assembly launch

// The bases of a launch, by value in the kernel arguments; entries past nb (padding of the last
// group) are zero functions and their entries are not written.
struct BasisSet {
    int nb;     // bases per element
    int G;      // bases per group
    int ngroup; // groups (G * ngroup >= nb)
    int kind[BEM2D_MAX_BASES + kBasisGroup];
    int deg[BEM2D_MAX_BASES + kBasisGroup];
    double dir[BEM2D_MAX_BASES + kBasisGroup];
    double kr[BEM2D_MAX_BASES + kBasisGroup];
    double ki[BEM2D_MAX_BASES + kBasisGroup];
};

// phi_b(t), lib/bem2d_basistypes.h:19-94: 1, t^deg, t (sign 1) / 1 - t (sign -1),
// exp(i k dir t), exp(i k dir t) t^deg; zero past nb.
__host__ __device__ inline void basis_value(const BasisSet& s, int b, double t, double& vr,
                                            double& vi)
{
    vi = 0.0;
    if (b >= s.nb) {
        vr = 0.0;
        return;
    }
    const int kind = s.kind[b];
    if (kind == BEM2D_BASIS_CONST) {
        vr = 1.0;
        return;
    }
    if (kind == BEM2D_BASIS_LIN) {
        vr = s.deg[b] == -1 ? 1 - t : t;
        return;
    }
    double r = 1.0;
    for (int i = 0; i < s.deg[b]; ++i) r = r * t;
    if (kind == BEM2D_BASIS_POL) {
        vr = r;
        return;
    }
    // i k dir t = (-ki dir t) + i (kr dir t)
    const double er = -s.ki[b] * s.dir[b] * t, ei = s.kr[b] * s.dir[b] * t;
    double sn, cs;
    sincos(ei, &sn, &cs);
    const double m = exp(er);
    vr = m * cs * r;
    vi = m * sn * r;
}

// Checks a basis list and fills its descriptor (groups of at most kBasisGroup bases).
int bem2d_internal_make_basis_set(const bem2d_basis* bases, int nb, BasisSet* out);

struct AsmProblem;
// The regular pairs of a P0 assembly with complex k (k = p.k + i ki, contour table tab / nt) on the
// tiled kernel, when the mesh has enough tiles to fill the device; false (nothing launched)
// otherwise.
bool bem2d_internal_regular_ck_tiles(hipStream_t st, const AsmProblem& p, int kind, double ki,
                                     const double* tab, int nt);

} // namespace bem2d_dev
