// Drop-in replacement of the operator API of lib/bem2d_quadrature.h:26-294.
// The reference's templates DiscreteKernel<kernel>, DotProdBasFuns<Bem2dfun> and EvalKernel<kernel>
// become one overload per functor class of the reference; each overload flattens the geometry and
// runs on the GPU through the C-ABI (include/bem2d_hip.h).  Geometries must carry one PolBasis(0)
// per element on elements whose curve the GPU evaluates itself (ConstElement and
// AnalyticCurveElement on the reference's curve types); otherwise ParameterException is thrown.
#ifndef BEM2D_DROPIN_QUADRATURE_H
#define BEM2D_DROPIN_QUADRATURE_H

#include <vector>

#include "bem2d_fun.h"
#include "bem2d_geometry.h"
#include "bem2d_kernel.h"
#include "bem2d_mathroutines.h"
#include "bem2d_quadpoints.h"

namespace bem2d {

struct QuadOption {
    int L;
    int N;
    double sigma;
};

// Gauss-Legendre rule on [0,1] (lib/bem2d_quadrature.cpp:10-56).
void Gauss(dvector& x, dvector& w, int N);

// Galerkin matrices (lib/bem2d_quadrature.h:231-294).
pMatrix DiscreteKernel(const Geometry& geom, QuadOption opts, SingleLayer g);
pMatrix DiscreteKernel(const Geometry& geom, QuadOption opts, DoubleLayer g);
pMatrix DiscreteKernel(const Geometry& geom, QuadOption opts, ConjDoubleLayer g);
pMatrix DiscreteKernel(const Geometry& geom, QuadOption opts, CombinedSingleConjDouble g);
pMatrix DiscreteKernel(const Geometry& geom, QuadOption opts, IdKernel g);

// Mass matrix (lib/bem2d_quadrature.cpp:79-127).
pMatrix EvalIdent(const Geometry& geom, QuadOption opts);

// Right-hand sides <f, phi_i> (lib/bem2d_quadrature.h:130-156).
pMatrix DotProdBasFuns(const Geometry& geom, QuadOption ops, PlaneWave fun);
pMatrix DotProdBasFuns(const Geometry& geom, QuadOption ops, NormalPlaneWave fun);
pMatrix DotProdBasFuns(const Geometry& geom, QuadOption ops, CombinedPlaneWave fun);

// vals[j] += sum_i alpha_i int phi_i(y) G(points[j], y) ds(y) for the single layer
// (lib/bem2d_quadrature.h:184-228); throws SizeError if vals and points differ in size.
void EvalKernel(const Geometry::flat_basis_map& basfuns, const std::vector<Point>& points,
                cvector& vals, Matrix& alpha, SingleLayer g, const Gauss1D& g1d);

// New: far field u_inf(theta) of the scattered field -S phi for the coefficient columns of alpha:
// result is theta.size() x alpha.dim[1], column-major.
pMatrix FarField(const Geometry& geom, QuadOption opts, freqtype k, Matrix& alpha,
                 const dvector& theta);

} // namespace bem2d

#endif
