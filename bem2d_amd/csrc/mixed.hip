// Mixed-precision solver (FP32-MFMA complex LU + FP64 refinement): workspace release hook.
#include "internal.h"

struct MixedState {
};

void bem2d_internal_mixed_release(bem2d_ctx* c)
{
    delete c->mixed;
    c->mixed = nullptr;
}
