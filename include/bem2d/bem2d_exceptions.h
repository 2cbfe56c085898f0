// Drop-in replacement of lib/bem2d_exceptions.h:10-33: empty exception tags.  DeviceError is new:
// it reports a failure of the GPU runtime underneath (the reference has no device).
#ifndef BEM2D_DROPIN_EXCEPTIONS_H
#define BEM2D_DROPIN_EXCEPTIONS_H

namespace bem2d {

class ArrayMismatch {};
class LapackError {};
class IntervalException {};
class ParameterException {};
class SizeError {};
class DeviceError {};

} // namespace bem2d

#endif
