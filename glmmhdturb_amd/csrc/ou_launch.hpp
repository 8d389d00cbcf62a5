// ou_launch.hpp -- host side of the band-limited OU forcing (ou.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace glm {

// One box's OU forcing: scalars and device buffers (the kernels are in ou_impl.hpp).
struct OUBox {
    int g, nb, K, pad0;
    double kmin, kmax, alpha, beta, comp, sol, scale;   // scale = amplitude / g^3
    unsigned long long seed;
    double *fre, *fim;        // [3][nb^3] band modes, (ea, eb, ed) at ea + nb*(eb + nb*ed)
    double *w1r, *w1i;        // [3][nb][nb][g] after the x pass
    double *w2r, *w2i;        // [3][nb][g][g]  after the y pass
    const double *cs, *sn;    // [g] cos, sin of 2*pi*m/g
    void* afield;             // [3][g^3] forcing acceleration in the state precision
    int due;                  // ensemble: this box advances in the current step
    int pad1;
    unsigned long long step;  // ensemble: Philox step counter of that advance
};


// Band width nb = min(2*floor(kmax)+1, g) of the active modes per axis.
int ou_band_width(int g, double kmax);

// Allocate and zero one box's OU buffers (one device allocation) and fill its twiddle table;
// the scalars follow glmmhd_turb.jl:2117-2131, 2146-2150.
hipError_t ou_box_alloc(OUBox& b, int g, double kmin, double kmax, double T, int ndt, double comp, double amp,
                        unsigned long long seed, int prec, hipStream_t st);
void ou_box_free(OUBox& b);
// Zero the OU state and the forcing field (a restart from F = 0).
hipError_t ou_box_reset(const OUBox& b, int prec, hipStream_t st);

// One OU advance of nbox boxes (descriptors in device memory): step_all >= 0 advances every box
// with that Philox step counter; step_all < 0 advances the boxes whose `due` flag is set, each
// with its own `step`.  gmax, nbmax bound the boxes' g and nb.
hipError_t launch_ou_advance(int prec, const OUBox* d_boxes, int nbox, int gmax, int nbmax, long long step_all,
                             hipStream_t st);

}  // namespace glm
