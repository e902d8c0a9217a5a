// aux_kernels.hip - kernels for gfx950 outside the tile engine: complex scaling of a state (also the
// copy probe), CSR matrix times states, probability block sums and the per-block search of
// measurement sampling.
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

// Grid-stride over amplitudes; one 8- or 16-byte load and store per amplitude.  NT: non-temporal
// accesses (streamed data that is not read again).
template <typename T, bool NT = false>
__global__ void __launch_bounds__(256) k_scale(const typename Cplx<T>::V* src, typename Cplx<T>::V* dst, uint64_t n,
                                               T re, T im) {
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += uint64_t(gridDim.x) * blockDim.x) {
        typename Cplx<T>::V v;
        if constexpr (NT) {
            v.x = __builtin_nontemporal_load(&src[j].x);
            v.y = __builtin_nontemporal_load(&src[j].y);
        } else {
            v = src[j];
        }
        typename Cplx<T>::V o;
        o.x = re * v.x - im * v.y;
        o.y = re * v.y + im * v.x;
        if constexpr (NT) {
            __builtin_nontemporal_store(o.x, &dst[j].x);
            __builtin_nontemporal_store(o.y, &dst[j].y);
        } else {
            dst[j] = o;
        }
    }
}

// CSR matrix-vector product, one row per thread (Hamiltonian rows hold few non-zeros); the gathered
// ket entries of neighbouring rows fall in the same cache lines for Pauli-structured matrices.
template <typename T>
__global__ void __launch_bounds__(256) k_csr_spmv(const int64_t* __restrict__ indptr, const int64_t* __restrict__ indices,
                                                  const typename Cplx<T>::V* __restrict__ data, uint64_t rows,
                                                  const typename Cplx<T>::V* __restrict__ ket, uint64_t ket_stride,
                                                  typename Cplx<T>::V* __restrict__ out, uint64_t out_stride,
                                                  uint64_t out_rows) {
    const uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= out_rows) return;
    const uint64_t b = blockIdx.y;
    const typename Cplx<T>::V* k = ket + b * ket_stride;
    T ar = 0, ai = 0;
    if (r < rows) {
        for (int64_t e = indptr[r]; e < indptr[r + 1]; e++) {
            const typename Cplx<T>::V m = data[e];
            const typename Cplx<T>::V x = k[indices[e]];
            ar += m.x * x.x - m.y * x.y;
            ai += m.x * x.y + m.y * x.x;
        }
    }
    typename Cplx<T>::V o;
    o.x = ar;
    o.y = ai;
    out[b * out_stride + r] = o;
}

// Block sums of |psi|^2: one workgroup of 256 threads per block, double accumulation, fixed order.
template <typename T>
__global__ void __launch_bounds__(256) k_prob_block_sums(const typename Cplx<T>::V* psi, uint32_t block, double* out) {
    __shared__ double part[256];
    const typename Cplx<T>::V* p = psi + uint64_t(blockIdx.x) * block;
    double acc = 0;
    for (uint32_t j = threadIdx.x; j < block; j += blockDim.x) {
        const typename Cplx<T>::V v = p[j];
        acc += double(v.x) * double(v.x) + double(v.y) * double(v.y);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

// Per touched block: inclusive prefix of |psi|^2 over the block in LDS (each thread scans a
// contiguous chunk, a block-wide scan of the chunk totals adds the offsets), then a binary search per
// shot.  block <= kSampleBlock, a multiple of 256 or smaller than 256 (one amplitude per thread).
template <typename T>
__global__ void __launch_bounds__(256) k_sample_blocks(const typename Cplx<T>::V* psi, uint32_t block,
                                                       const uint64_t* block_id, const uint64_t* shot_begin,
                                                       const uint64_t* shot_end, const double* residual,
                                                       uint64_t* out) {
    __shared__ double pre[kSampleBlock];
    __shared__ double tot[256];
    const uint64_t b = block_id[blockIdx.x];
    const typename Cplx<T>::V* p = psi + b * block;
    const uint32_t per = (block + 255u) / 256u;  // amplitudes per thread
    const uint32_t j0 = threadIdx.x * per;
    double run = 0;
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t j = j0 + k;
        if (j < block) {
            const typename Cplx<T>::V v = p[j];
            run += double(v.x) * double(v.x) + double(v.y) * double(v.y);
            pre[j] = run;
        }
    }
    tot[threadIdx.x] = run;
    __syncthreads();
    // inclusive scan of the 256 chunk totals (Hillis-Steele, two buffers' worth of syncs)
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const double add = threadIdx.x >= off ? tot[threadIdx.x - off] : 0.0;
        __syncthreads();
        tot[threadIdx.x] += add;
        __syncthreads();
    }
    const double base = threadIdx.x > 0 ? tot[threadIdx.x - 1] : 0.0;
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t j = j0 + k;
        if (j < block) pre[j] += base;
    }
    __syncthreads();
    for (uint64_t sh = shot_begin[blockIdx.x] + threadIdx.x; sh < shot_end[blockIdx.x]; sh += blockDim.x) {
        const double r = residual[sh];
        uint32_t lo = 0, hi = block - 1;  // first j with pre[j] >= r, at most block - 1
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pre[mid] >= r) hi = mid;
            else lo = mid + 1;
        }
        out[sh] = b * block + lo;
    }
}

}  // namespace

hipError_t launch_csr_spmv(int dtype, const int64_t* indptr, const int64_t* indices, const void* data, uint64_t rows,
                           const void* ket, uint64_t ket_stride, void* out, uint64_t out_stride, uint64_t out_rows,
                           uint64_t batch, hipStream_t s) {
    if (batch == 0 || out_rows == 0) return hipSuccess;
    const uint64_t gx = (out_rows + 255) / 256;
    if (gx > 0x7fffffffull || batch > 65535) return hipErrorInvalidValue;
    if (dtype == 0)
        hipLaunchKernelGGL(k_csr_spmv<float>, dim3(unsigned(gx), unsigned(batch)), dim3(256), 0, s, indptr, indices,
                           static_cast<const float2*>(data), rows, static_cast<const float2*>(ket), ket_stride,
                           static_cast<float2*>(out), out_stride, out_rows);
    else
        hipLaunchKernelGGL(k_csr_spmv<double>, dim3(unsigned(gx), unsigned(batch)), dim3(256), 0, s, indptr, indices,
                           static_cast<const double2*>(data), rows, static_cast<const double2*>(ket), ket_stride,
                           static_cast<double2*>(out), out_stride, out_rows);
    return hipGetLastError();
}

hipError_t launch_prob_block_sums(int dtype, const void* psi, uint64_t dim, uint32_t block, double* out,
                                  hipStream_t s) {
    if (block == 0 || dim % block != 0) return hipErrorInvalidValue;
    const uint64_t nb = dim / block;
    if (nb == 0 || nb > 0x7fffffffull) return hipErrorInvalidValue;
    if (dtype == 0)
        hipLaunchKernelGGL(k_prob_block_sums<float>, dim3(unsigned(nb)), dim3(256), 0, s,
                           static_cast<const float2*>(psi), block, out);
    else
        hipLaunchKernelGGL(k_prob_block_sums<double>, dim3(unsigned(nb)), dim3(256), 0, s,
                           static_cast<const double2*>(psi), block, out);
    return hipGetLastError();
}

hipError_t launch_sample_blocks(int dtype, const void* psi, uint32_t block, const uint64_t* block_id,
                                const uint64_t* shot_begin, const uint64_t* shot_end, uint64_t n_groups,
                                const double* residual, uint64_t* out, hipStream_t s) {
    if (n_groups == 0) return hipSuccess;
    if (block == 0 || block > kSampleBlock || (block > 256 && block % 256 != 0) || n_groups > 0x7fffffffull)
        return hipErrorInvalidValue;
    if (dtype == 0)
        hipLaunchKernelGGL(k_sample_blocks<float>, dim3(unsigned(n_groups)), dim3(256), 0, s,
                           static_cast<const float2*>(psi), block, block_id, shot_begin, shot_end, residual, out);
    else
        hipLaunchKernelGGL(k_sample_blocks<double>, dim3(unsigned(n_groups)), dim3(256), 0, s,
                           static_cast<const double2*>(psi), block, block_id, shot_begin, shot_end, residual, out);
    return hipGetLastError();
}

hipError_t launch_scale(int dtype, const void* src, void* dst, uint64_t n_amps, double re, double im, hipStream_t s) {
    if (n_amps == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((n_amps + 255) / 256, uint64_t(1) << 20);
    if (dtype == 2) {  // complex128 with non-temporal accesses (probe)
        hipLaunchKernelGGL((k_scale<double, true>), dim3(unsigned(blocks)), dim3(256), 0, s,
                           static_cast<const double2*>(src), static_cast<double2*>(dst), n_amps, re, im);
        return hipGetLastError();
    }
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
