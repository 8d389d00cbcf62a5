// march_f64.hip -- FP64 instantiations of the fused line-march kernel.
#include "launch.hpp"

namespace glm {

template <>
hipError_t launch_march<double>(const MarchArgs<double>& a, int pred, int riem, hipStream_t st)
{
    dim3 grid((unsigned)(a.tiles_x * a.tiles_y * a.nzc)), block(NT);
#define GLM_CASE(P, R)                                                                   \
    if (pred == P && riem == R) {                                                        \
        hipLaunchKernelGGL((march_kernel<double, P, R>), grid, block, 0, st, a);         \
        return hipGetLastError();                                                        \
    }
    GLM_CASE(PRED_FULL, RIEM_LLF) GLM_CASE(PRED_FULL, RIEM_LLFB) GLM_CASE(PRED_FULL, RIEM_HLL) GLM_CASE(PRED_FULL, RIEM_HLLD)
    GLM_CASE(PRED_ONED, RIEM_LLF) GLM_CASE(PRED_ONED, RIEM_LLFB) GLM_CASE(PRED_ONED, RIEM_HLL) GLM_CASE(PRED_ONED, RIEM_HLLD)
#undef GLM_CASE
    return hipErrorInvalidValue;
}

template <>
const void* march_kernel_ptr<double>(int pred, int riem)
{
#define GLM_CASE(P, R) \
    if (pred == P && riem == R) return (const void*)&march_kernel<double, P, R>;
    GLM_CASE(PRED_FULL, RIEM_LLF) GLM_CASE(PRED_FULL, RIEM_LLFB) GLM_CASE(PRED_FULL, RIEM_HLL) GLM_CASE(PRED_FULL, RIEM_HLLD)
    GLM_CASE(PRED_ONED, RIEM_LLF) GLM_CASE(PRED_ONED, RIEM_LLFB) GLM_CASE(PRED_ONED, RIEM_HLL) GLM_CASE(PRED_ONED, RIEM_HLLD)
#undef GLM_CASE
    return nullptr;
}

}  // namespace glm
