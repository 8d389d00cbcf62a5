// Host build of the select-free limiter forms the GPU kernels use (csrc/physics.hpp: MonCen's half
// slope, the Julia PPM edge offsets, the spike ppm1 edges), for CPU checks against the oracle's
// branch forms.  Test infrastructure; no GPU code.
#include "../../glmmhdturb_amd/csrc/physics.hpp"

extern "C" {

void lc_moncen_half(const double* dl, const double* dr, double* out, long n)
{
    for (long i = 0; i < n; ++i) out[i] = glm::moncen_half(dl[i], dr[i]);
}

// glmmhd_turb.jl:917-923 edges: q0 - dl and q0 + dr from ppm_offsets
void lc_ppm_julia(const double* qm, const double* q0, const double* qp, double* lo, double* hi, long n)
{
    for (long i = 0; i < n; ++i) {
        double dl, dr;
        glm::ppm_offsets(qm[i], q0[i], qp[i], dl, dr);
        lo[i] = q0[i] - dl;
        hi[i] = q0[i] + dr;
    }
}

// spike_mhd.cu:164-172 ppm1 edges
void lc_ppm_spike(const double* qm, const double* q0, const double* qp, double* lo, double* hi, long n)
{
    for (long i = 0; i < n; ++i) glm::ppm_edges_spike(qm[i], q0[i], qp[i], lo[i], hi[i]);
}

}
