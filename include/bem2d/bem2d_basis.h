// Drop-in replacement of lib/bem2d_basis.h:9-16 and lib/bem2d_basistypes.h:28-99.
// The device path accepts one PolBasis(0) (piecewise constant) per element; geometries with other
// bases are rejected with ParameterException by the GPU operators.
#ifndef BEM2D_DROPIN_BASIS_H
#define BEM2D_DROPIN_BASIS_H

#include "bem2d_defs.h"

namespace bem2d {

class Basis {
public:
    virtual complex operator()(double t) const = 0;
    // Polynomial degree of a PolBasis, -1 for any other basis.
    virtual int pol_degree() const { return -1; }
    virtual ~Basis() {}
};

typedef std::shared_ptr<Basis> pBasis;

} // namespace bem2d

#endif
