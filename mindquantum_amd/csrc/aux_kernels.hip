// aux_kernels.hip - small element-wise kernels for gfx950: complex scaling of a state.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "aux_kernels.h"

namespace mqa {
namespace {

template <typename T>
struct Cplx;
template <>
struct Cplx<float> {
    using V = float2;
};
template <>
struct Cplx<double> {
    using V = double2;
};

// Grid-stride over amplitudes; one 8- or 16-byte load and store per amplitude.
template <typename T>
__global__ void __launch_bounds__(256) k_scale(const typename Cplx<T>::V* src, typename Cplx<T>::V* dst, uint64_t n,
                                               T re, T im) {
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += uint64_t(gridDim.x) * blockDim.x) {
        const typename Cplx<T>::V v = src[j];
        typename Cplx<T>::V o;
        o.x = re * v.x - im * v.y;
        o.y = re * v.y + im * v.x;
        dst[j] = o;
    }
}

}  // namespace

hipError_t launch_scale(int dtype, const void* src, void* dst, uint64_t n_amps, double re, double im, hipStream_t s) {
    if (n_amps == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((n_amps + 255) / 256, uint64_t(1) << 20);
    if (dtype == 0) {
        hipLaunchKernelGGL(k_scale<float>, dim3(unsigned(blocks)), dim3(256), 0, s, static_cast<const float2*>(src),
                           static_cast<float2*>(dst), n_amps, float(re), float(im));
    } else {
        hipLaunchKernelGGL(k_scale<double>, dim3(unsigned(blocks)), dim3(256), 0, s, static_cast<const double2*>(src),
                           static_cast<double2*>(dst), n_amps, re, im);
    }
    return hipGetLastError();
}

}  // namespace mqa
