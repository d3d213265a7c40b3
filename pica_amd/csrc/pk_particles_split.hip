// pk_particles_split.hip -- FP32 split-position particle step (int32 cell + float offset).
#include "pk_internal.h"

int pk_particles_split(pk_ens* e, pk_grid* g, double dt, int shape, int dep)
{
    (void)g; (void)dt; (void)shape; (void)dep;
    return pk_fail(e->ctx, PK_E_UNSUPPORTED, "FP32 split-position step not built yet");
}
