// Mixed-precision solve: FP32 Jacobi-PCG with FP64 residual refinement.
#include "kt_internal.h"

namespace kt {

int solveMixed(kt_ctx *ctx, const kt_solve_opts &o, kt_solve_result &res) {
    setError("mixed precision solve not available yet");
    return KT_EINVAL;
}

}  // namespace kt
