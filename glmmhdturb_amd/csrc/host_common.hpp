// host_common.hpp -- host-side helpers shared by the C-ABI translation units: the thread-local
// error state (defined in capi.hip), the scheme block -> Phys conversion and parameter checks.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/glmmhd.h"
#include "physics.hpp"

namespace glm {

int set_err(int code, const char* msg);
int hip_err(hipError_t e, const char* where);

#define HIPCHK(call)                                                   \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return hip_err(e_, #call);               \
    } while (0)

template <typename T>
inline Phys<T> make_phys(const glm_params* p)
{
    double pfl = p->pfl > 0.0 ? p->pfl : p->smallr * p->smallc * p->smallc / p->gamma;
    Phys<T> s;
    s.gamma = (T)p->gamma;
    s.igm1 = (T)(1.0 / (p->gamma - 1.0));
    s.smallr = (T)p->smallr;
    s.pfl = (T)pfl;
    s.ch = (T)p->ch;
    s.ch2 = (T)(p->ch * p->ch);
    s.half_ich = (T)(0.5 / p->ch);
    s.llf_dmin = (T)p->llf_dmin;
    s.llf_pmin = (T)p->llf_pmin;
    return s;
}

// Every one of the nv plane addresses is set (a kernel would otherwise fault on the null plane).
inline int check_planes(void* const* q, int nv, const char* who)
{
    if (!q) return set_err(GLM_ERR_ARG, who);
    for (int v = 0; v < nv; ++v)
        if (!q[v]) return set_err(GLM_ERR_ARG, who);
    return GLM_OK;
}

inline int check_params(const glm_params* p)
{
    if (!p) return set_err(GLM_ERR_ARG, "null glm_params");
    if (!(p->gamma > 1.0) || !(p->dx > 0.0) || !(p->ch > 0.0) || !(p->smallr > 0.0))
        return set_err(GLM_ERR_ARG, "glm_params: need gamma > 1, dx > 0, ch > 0, smallr > 0");
    if (p->predictor < 0 || p->predictor > 1 || p->recon < 0 || p->recon > 2 || p->slope < 0 || p->slope > 1 ||
        p->riemann < 0 || p->riemann > 3 || p->face_guard < 0 || p->face_guard > 2)
        return set_err(GLM_ERR_ARG, "glm_params: scheme switch out of range");
    return GLM_OK;
}

}  // namespace glm
