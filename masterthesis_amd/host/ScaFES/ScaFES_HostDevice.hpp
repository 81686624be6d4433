// ScaFES_HostDevice.hpp -- what the host-API headers need to be usable from
// device code when a problem's translation unit is compiled by hipcc.
//
// SCAFES_HD marks the small index/grid/accessor functions a problem's
// update hooks call (Ntuple arithmetic, Grid::neigh, Problem::connect,
// tau, gridsize, ...) as callable on the host and on the GPU.  Compiled by
// g++ it expands to nothing and the headers are plain C++11.
//
// SCAFES_DEVICE_HOOKS_BEGIN / SCAFES_DEVICE_HOOKS_END bracket the #include
// of an unchanged problem header in a hipcc-compiled driver: every function
// declared in between becomes host+device, so the problem's own
// updateInner/updateBorder(/2) templates can run as GPU code (the generic
// device update phase of Problem::iterateOverTime).  Under g++ they expand
// to nothing.
#ifndef SCAFES_HOST_HOSTDEVICE_HPP_
#define SCAFES_HOST_HOSTDEVICE_HPP_

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SCAFES_HD __host__ __device__
#define SCAFES_DEVICE_HOOKS_BEGIN _Pragma("clang force_cuda_host_device begin")
#define SCAFES_DEVICE_HOOKS_END _Pragma("clang force_cuda_host_device end")
#else
#define SCAFES_HD
#define SCAFES_DEVICE_HOOKS_BEGIN
#define SCAFES_DEVICE_HOOKS_END
#endif

#include <cstddef>

namespace ScaFES {
// Data fields the generic device update phase handles (the reference's
// largest example, EMFDTD, has 25).
constexpr std::size_t kMaxDeviceFields = 32;
}  // namespace ScaFES

#endif
