// Drop-in replacement of lib/bem2d_defs.h:15-35 (types and constants of bem2d).
// boost::shared_ptr becomes std::shared_ptr behind the same typedef names.
#ifndef BEM2D_DROPIN_DEFS_H
#define BEM2D_DROPIN_DEFS_H

#include <complex>
#include <memory>
#include <vector>

namespace bem2d {

typedef std::complex<double> complex;
typedef std::vector<double> dvector;
typedef std::vector<complex> cvector;

// Wavenumber k = re + i im (lib/bem2d_defs.h:20-24).
struct freqtype {
    double re;
    double im;
};

typedef std::shared_ptr<cvector> pcvector;
typedef std::shared_ptr<dvector> pdvector;

const double PI = 3.14159265358979323846;

} // namespace bem2d

#endif
