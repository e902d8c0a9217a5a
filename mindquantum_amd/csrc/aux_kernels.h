// aux_kernels.h - small element-wise kernels (aux_kernels.hip) used by the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mqa {

// dst[j] = (re + i im) * src[j] for j < n_amps (dst may equal src).  dtype: 0 = complex64,
// 1 = complex128.
hipError_t launch_scale(int dtype, const void* src, void* dst, uint64_t n_amps, double re, double im, hipStream_t s);

}  // namespace mqa
