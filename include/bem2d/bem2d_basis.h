// Drop-in replacement of lib/bem2d_basis.h:9-16.
// device_desc() is new: a basis the GPU path can evaluate by itself reports its descriptor
// (bem2d_basis, include/bem2d_hip.h); user-defined bases return false and are rejected with
// ParameterException by the GPU operators.
#ifndef BEM2D_DROPIN_BASIS_H
#define BEM2D_DROPIN_BASIS_H

#include "bem2d_defs.h"
#include "../bem2d_hip.h"

namespace bem2d {

class Basis {
public:
    virtual complex operator()(double t) const = 0;
    virtual bool device_desc(bem2d_basis& out) const
    {
        (void)out;
        return false;
    }
    virtual ~Basis() {}
};

typedef std::shared_ptr<Basis> pBasis;

} // namespace bem2d

#endif
