// adj_launch.hpp -- host side of the reverse step (adj.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "physics.hpp"

namespace glm {

constexpr int ADJ_NT = 256;

// Scratch planes of the reverse step (n doubles each).
constexpr int ADJ_W = 0, ADJ_MH = 9, ADJ_FS = 18, ADJ_LPRE = 72, ADJ_LFS = 81, ADJ_CON = 135, ADJ_MHB = 198,
              ADJ_FLX = 207, ADJ_NPLANES = 234;   // ADJ_FLX: the flux through each cell's three lower faces

struct AdjArgs {
    const double* u0[9];
    const double* lam1[9];
    double* lam0[9];
    int nx, ny, nz, tgs;
    Phys<double> ph;
    double dtdx, half_dtdx, glm_fac, smallc, turb_min_rho, ramp, dt, dx, lx, ly, lz;
    int mh_revert, face_guard, update_floor, do_turb;
    const double* afield;   // 3 planes of tgs^3 (the step's forcing field), or null
    double* scr;            // ADJ_NPLANES planes of n
    double* part;           // per-block partial sums of a_bar . a (pass 3)
};


// Device scratch of one reverse step over n cells, in bytes (ADJ_NPLANES planes plus the
// per-block partial sums).
size_t adj_scratch_bytes(size_t n);
// Launch the reverse step and add the step's sum of (kick acceleration adjoint . acceleration)
// to *fdot (device).  With fused, the face-state and predictor adjoints and the gather run as one
// z-march kernel (adj_march.hpp, z planes per tile: zchunk, 0 = the occupancy model); otherwise as
// the separate passes 5 and 6.  a.scr and a.part must come from a scratch of adj_scratch_bytes(n).
// Returns hipErrorInvalidValue for an unsupported scheme.
hipError_t launch_step_adjoint(AdjArgs a, void* scratch, int pred, int riem, int recon, int slope, double* fdot,
                               bool fused, int zchunk, hipStream_t st);

}  // namespace glm
