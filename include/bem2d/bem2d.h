// Umbrella header of the drop-in API (mirrors lib/bem2d.h).
#ifndef BEM2D_DROPIN_H
#define BEM2D_DROPIN_H

#include "bem2d_basis.h"
#include "bem2d_basistypes.h"
#include "bem2d_curve.h"
#include "bem2d_curvetypes.h"
#include "bem2d_defs.h"
#include "bem2d_element.h"
#include "bem2d_exceptions.h"
#include "bem2d_fun.h"
#include "bem2d_geometry.h"
#include "bem2d_kernel.h"
#include "bem2d_mathroutines.h"
#include "bem2d_point.h"
#include "bem2d_quadpoints.h"
#include "bem2d_quadrature.h"
#include "bem2d_shape.h"
#include "bem2d_soundsoftscattering.h"

#endif
