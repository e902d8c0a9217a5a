// aux_kernels.h - small element-wise kernels (aux_kernels.hip) used by the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mqa {



// dst[j] = (re + i im) * src[j] for j < n_amps (dst may equal src).  dtype: 0 = complex64,
// 1 = complex128.
hipError_t launch_scale(int dtype, const void* src, void* dst, uint64_t n_amps, double re, double im, hipStream_t s);

// out[b][r] = sum_k data[k] * ket[b][indices[k]] over indptr[r] <= k < indptr[r + 1] for r < rows, and 0
// for rows <= r < out_rows (padding amplitudes), for every sample b < batch (CSR matrix times a
// batch of states; data holds interleaved complex values in the state's precision).  Column
// indices must be < rows (checked where the matrix is created).
hipError_t launch_csr_spmv(int dtype, const int64_t* indptr, const int64_t* indices, const void* data, uint64_t rows,
                           const void* ket, uint64_t ket_stride, void* out, uint64_t out_stride, uint64_t out_rows,
                           uint64_t batch, hipStream_t s);

// Sampling (measure-ending circuits).  Amplitudes are split into blocks of kSampleBlock (or the
// whole state when smaller).
constexpr uint32_t kSampleBlock = 4096;
// out[b] = sum of |psi[j]|^2 over block b (double), for n_blocks = dim / block blocks.
hipError_t launch_prob_block_sums(int dtype, const void* psi, uint64_t dim, uint32_t block, double* out,
                                  hipStream_t s);
// One workgroup per touched block: the block's inclusive prefix of |psi|^2 in LDS, then for every
// shot s in [shot_begin[g], shot_end[g]) the first index j of the block with prefix[j] >= residual[s]
// (lower bound; clamped to the block) -> out[s] = block_id[g] * block + j.
hipError_t launch_sample_blocks(int dtype, const void* psi, uint32_t block, const uint64_t* block_id,
                                const uint64_t* shot_begin, const uint64_t* shot_end, uint64_t n_groups,
                                const double* residual, uint64_t* out, hipStream_t s);

}  // namespace mqa
