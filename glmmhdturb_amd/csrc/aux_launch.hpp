// aux_launch.hpp -- host launchers of the auxiliary kernels (aux.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "physics.hpp"

namespace glm {

constexpr size_t DIAG_CELLS_PER_BLOCK = 16384;

hipError_t launch_clear_u64(unsigned long long* p, int n, hipStream_t st);
template <typename T>
hipError_t launch_ctot(Ptr9<const T> q9dev, size_t n, const Phys<T>& s, unsigned long long* out, hipStream_t st,
                       bool fast = false);
size_t diag_nparts(size_t n);
size_t diag_part_bytes();
template <typename T>
hipError_t launch_diag(Ptr9<const T> q9dev, int nx, int ny, int nz, double dx, double smallr, void* parts,
                       double* out, hipStream_t st);
template <typename T>
hipError_t launch_ic(Ptr9<T> q9dev, int nx, int ny, int nz, int kind, const double* args, hipStream_t st);

}  // namespace glm
