// Adjoint of the converged gamma (placeholder until the CG kernels land).
#include "device_common.h"

namespace pasc {

int engine_adjoint(Engine &, int, const uint8_t *, double, int, double *, double *, double *, int *) {
    return fail(PASC_ERR_STATE, "adjoint: not available in this build");
}

}  // namespace pasc
