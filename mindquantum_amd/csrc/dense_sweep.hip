// dense_sweep.hip - one dense 2^k x 2^k matrix (k <= 5) applied to a state in one streaming sweep
// over HBM, for gfx950.  This is the apply of a fused dense block (a run of gates merged on the host
// into one k-qubit unitary), the GPU form of ApplyNQubitsMatrix
// (ccsrc/lib/simulator/vector/detail/cpu_common/cpu_vector_core_matrix_gate.cpp:28-83): amplitude
// (rest | sum_b bit_b(i) << q_b) is row / column i of the matrix, for every rest index whose control
// bits are all set.
//
// Two forms:
// * VALU (any k, both precisions): one thread per rest index holds the block's 2^k amplitudes in
//   registers and writes each output row as soon as it is summed (4 * 2^k FMA per amplitude).  The
//   matrix is uniform, so its entries come through the scalar cache as SGPR operands.
// * MFMA (complex64, k = 5): a wave takes 32 rest indices at a time as the columns of a 32 x 32
//   complex block X and forms Y = U X as the real products Ur Xr - Ui Xi and Ur Xi + Ui Xr with
//   v_mfma_f32_32x32x2_f32 (K = 2 per instruction, 64 instructions per block).  At k = 5 the VALU
//   form needs 128 FMA per amplitude, more issue time than the sweep's HBM time on MI355X with packed
//   FP32 off; the FP32 matrix cores run the same contraction at twice that rate.  (FP64 matrix and
//   vector rates are equal on MI355X, so complex128 stays on the VALU form.)
#include <hip/hip_runtime.h>

#include "dense_sweep.h"

namespace mqa {
namespace {

template <typename T>
struct Amp;
template <>
struct Amp<float> {
    using V = float2;
};
template <>
struct Amp<double> {
    using V = double2;
};

struct DenseArgs {
    void* psi;
    const void* mat;  // row-major 2^k x 2^k, interleaved (re, im), state precision
    uint64_t n_rest;  // dim >> k
    uint64_t ctrl;    // control mask (0: none)
    uint64_t off[kDenseMaxDim];  // off[i] = sum_b bit_b(i) << q_b
    uint32_t sq[kDenseMaxQubits];  // target qubits, ascending
};

// rest index -> amplitude index with a zero at every target position
template <int K>
__device__ __forceinline__ uint64_t spread(uint64_t r, const DenseArgs& a) {
#pragma unroll
    for (int b = 0; b < K; b++) {
        const uint32_t p = a.sq[b];
        r = (r & ((uint64_t(1) << p) - 1)) | ((r >> p) << (p + 1));
    }
    return r;
}

template <typename T, int K>
__global__ void __launch_bounds__(256) k_dense_sweep(DenseArgs a) {
    constexpr int D = 1 << K;
    using V = typename Amp<T>::V;
    const uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= a.n_rest) return;
    const uint64_t base = spread<K>(r, a);
    if ((base & a.ctrl) != a.ctrl) return;
    V* psi = static_cast<V*>(a.psi);
    const T* m = static_cast<const T*>(a.mat);
    T xr[D], xi[D];
#pragma unroll
    for (int i = 0; i < D; i++) {
        const V v = psi[base + a.off[i]];
        xr[i] = v.x, xi[i] = v.y;
    }
#pragma unroll
    for (int row = 0; row < D; row++) {
        T yr = 0, yi = 0;
#pragma unroll
        for (int col = 0; col < D; col++) {
            const T mr = m[2 * (row * D + col)], mi = m[2 * (row * D + col) + 1];
            yr = __builtin_fma(mr, xr[col], yr);
            yr = __builtin_fma(-mi, xi[col], yr);
            yi = __builtin_fma(mr, xi[col], yi);
            yi = __builtin_fma(mi, xr[col], yi);
        }
        V o;
        o.x = yr, o.y = yi;
        psi[base + a.off[row]] = o;
    }
}

typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int kMfmaTilesPerWave = 4;  // 32-column tiles per wave: the matrix operands load once per 128 columns

// k = 5, complex64.  Operand layouts of v_mfma_f32_32x32x2_f32 (lane l, h = l / 32, j = l % 32):
// A[i][kk] at i = j, kk = h; B[kk][c] at kk = h, c = j; D[row][c] in accumulator v at c = j,
// row = 8 (v / 4) + 4 h + v % 4.  Instruction t contracts matrix columns / block rows 2t and 2t + 1.
__global__ void __launch_bounds__(256) k_dense_sweep_mfma32(DenseArgs a) {
    const int lane = threadIdx.x & 63;
    const int h = lane >> 5, j = lane & 31;
    const uint64_t wave = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const float2* m = static_cast<const float2*>(a.mat);
    float2* psi = static_cast<float2*>(a.psi);
    float ur[16], ui[16], nui[16];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const float2 u = m[j * 32 + 2 * t + h];
        ur[t] = u.x, ui[t] = u.y, nui[t] = -u.y;
    }
    const uint64_t hl = h ? a.off[1] : 0;  // block row 2t + h
    const uint64_t hs = h ? a.off[4] : 0;  // block row 8 (v / 4) + 4 h + v % 4
    for (int tile = 0; tile < kMfmaTilesPerWave; tile++) {
        const uint64_t r0 = (wave * kMfmaTilesPerWave + tile) * 32;
        if (r0 >= a.n_rest) break;  // uniform over the wave (n_rest is a multiple of 32)
        const uint64_t base = spread<5>(r0 + j, a);
        const bool on = (base & a.ctrl) == a.ctrl;
        float xr[16], xi[16];
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const float2 v = psi[base + a.off[2 * t] + hl];
            xr[t] = v.x, xi[t] = v.y;
        }
        f32x16 yr = {}, yi = {};
#pragma unroll
        for (int t = 0; t < 16; t++) {
            yr = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[t], xr[t], yr, 0, 0, 0);
            yi = __builtin_amdgcn_mfma_f32_32x32x2f32(ur[t], xi[t], yi, 0, 0, 0);
            yr = __builtin_amdgcn_mfma_f32_32x32x2f32(nui[t], xi[t], yr, 0, 0, 0);
            yi = __builtin_amdgcn_mfma_f32_32x32x2f32(ui[t], xr[t], yi, 0, 0, 0);
        }
        if (on) {
#pragma unroll
            for (int v = 0; v < 16; v++) {
                float2 o;
                o.x = yr[v], o.y = yi[v];
                psi[base + a.off[8 * (v >> 2) + (v & 3)] + hs] = o;
            }
        }
    }
}

template <typename T>
hipError_t launch_valu(const DenseArgs& a, int k, hipStream_t s) {
    const dim3 block(256);
    const dim3 grid(uint32_t((a.n_rest + 255) / 256));
    switch (k) {
        case 1: hipLaunchKernelGGL((k_dense_sweep<T, 1>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_dense_sweep<T, 2>), grid, block, 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_dense_sweep<T, 3>), grid, block, 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_dense_sweep<T, 4>), grid, block, 0, s, a); break;
        case 5: hipLaunchKernelGGL((k_dense_sweep<T, 5>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

bool dense_sweep_uses_mfma(int dtype, int n_qubits, int k, bool allow_mfma) {
    return allow_mfma && dtype == 0 && k == 5 && n_qubits >= 10;
}

hipError_t launch_dense_sweep(int dtype, void* psi, int n_qubits, const int32_t* qubits, int k, uint64_t ctrl_mask,
                              const void* mat, bool allow_mfma, hipStream_t s) {
    if (k < 1 || k > kDenseMaxQubits || k > n_qubits || n_qubits > 36) return hipErrorInvalidValue;
    if (ctrl_mask >> n_qubits) return hipErrorInvalidValue;
    DenseArgs a{};
    a.psi = psi;
    a.mat = mat;
    a.n_rest = (uint64_t(1) << n_qubits) >> k;
    a.ctrl = ctrl_mask;
    uint32_t sq[kDenseMaxQubits];
    for (int b = 0; b < k; b++) {
        if (qubits[b] < 0 || qubits[b] >= n_qubits) return hipErrorInvalidValue;
        if (ctrl_mask >> qubits[b] & 1) return hipErrorInvalidValue;
        sq[b] = uint32_t(qubits[b]);
        for (int c = 0; c < b; c++)
            if (sq[c] == sq[b]) return hipErrorInvalidValue;
    }
    for (int b = 1; b < k; b++)  // insertion sort (k <= 5)
        for (int c = b; c > 0 && sq[c - 1] > sq[c]; c--) {
            const uint32_t t = sq[c];
            sq[c] = sq[c - 1], sq[c - 1] = t;
        }
    for (int b = 0; b < k; b++) a.sq[b] = sq[b];
    for (int i = 0; i < (1 << k); i++) {
        uint64_t o = 0;
        for (int b = 0; b < k; b++)
            if (i >> b & 1) o |= uint64_t(1) << qubits[b];
        a.off[i] = o;
    }
    if (dense_sweep_uses_mfma(dtype, n_qubits, k, allow_mfma)) {
        const uint64_t cols_per_block = 4 * 32 * kMfmaTilesPerWave;  // 4 waves per block
        const dim3 grid(uint32_t((a.n_rest + cols_per_block - 1) / cols_per_block));
        hipLaunchKernelGGL(k_dense_sweep_mfma32, grid, dim3(256), 0, s, a);
        return hipGetLastError();
    }
    return dtype == 0 ? launch_valu<float>(a, k, s) : launch_valu<double>(a, k, s);
}

}  // namespace mqa
