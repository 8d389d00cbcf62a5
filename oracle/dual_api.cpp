// dual_api.cpp -- TEST INFRASTRUCTURE ONLY.  The GLM oracle (glm_oracle.c) compiled over
// forward-mode dual numbers (dual.hpp), with a C entry point for ctypes.  Arrays hold (value,
// tangent) pairs: a state is 9 planes of n pairs.  One step maps (U, dU) to (L(U), J_L(U) dU)
// with the step's scalars (dt, ch, psi decay, ramp) held fixed; the forcing field may carry a
// tangent too (its derivative with respect to the forcing amplitude).
#include "dual.hpp"
#define REAL Dual
#include "glm_oracle.c"

extern "C" {

int od_step(const double* uold, double* unew, int nx, int ny, int nz, const o_scheme* sc, const double* afield,
            double* scratch)
{
    return o_step((const Dual*)uold, (Dual*)unew, nx, ny, nz, sc, (const Dual*)afield, (Dual*)scratch);
}

size_t od_step_scratch_pairs(int nx, int ny, int nz) { return o_step_scratch_reals(nx, ny, nz); }

}  // extern "C"
