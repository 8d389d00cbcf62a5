// gpu_util.hpp -- small device helpers shared by the march kernels (GLM and CT).
#pragma once
#include "physics.hpp"

namespace glm {

constexpr int NXCD = 8;   // XCDs of an MI355X

// i mod n in [0, n) for any int i (periodic wrap of a cell or forcing index).
GLM_DEV int wrapi(int i, int n)
{
    int m = i % n;
    return m < 0 ? m + n : m;
}

// Logical tile of hardware workgroup id w when ids are dealt round-robin over NXCD XCDs: XCD x
// gets the contiguous logical range [start_x, start_x + count_x).  A bijection on [0, total).
GLM_DEV int xcd_block(int w, int total)
{
    const int q = total / NXCD, r = total % NXCD;
    const int xcd = w % NXCD, slot = w / NXCD;
    return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
}

// HBM access through the global address space.  Pointers read back from LDS (the cold arguments)
// have lost their address space, and generic loads are counted together with LDS operations, so
// every LDS wait in the march would also wait for the prefetch of the next planes.
template <typename T>
GLM_DEV T gload(const T* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(1))) T*)p;
#else
    return *p;   // host pass of the kernel templates: never executed
#endif
}

template <typename T>
GLM_DEV void gstore_nt(T v, T* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_nontemporal_store(v, (__attribute__((address_space(1))) T*)p);
#else
    *p = v;
#endif
}

// An argument block as a reference into the constant address space: a kernel reads each field with
// a scalar load where it needs it, instead of holding a copy of the whole block in registers (which
// cost the ensemble kernel 64 B per lane of spills).
template <typename S>
GLM_DEV const auto& constant_ref(const S* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(4))) S*)p;
#else
    return *p;   // host pass of the kernel templates: never executed
#endif
}
// A kernel's own argument block read in place from the kernarg segment, where the explicit
// arguments start: for a kernel whose one argument is an S (at offset 0), each field is then a
// scalar load where it is read, as the ensemble reads its boxes' blocks, instead of coming through
// the by-value parameter (the FP64 PPM + HLLD march at 256^3: 44 -> 36 B of spills per lane, 2%
// faster).
template <typename S>
GLM_DEV const auto& kernarg_ref()
{
    return constant_ref((const S*)__builtin_amdgcn_kernarg_segment_ptr());
}

}  // namespace glm
