// dense_sweep.h - streaming apply of one dense k-qubit matrix (dense_sweep.hip), used by the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mqa {

constexpr int kDenseMaxQubits = 5;
constexpr int kDenseMaxDim = 1 << kDenseMaxQubits;

// psi (2^n_qubits amplitudes of dtype 0 = complex64 / 1 = complex128, in place) <- M psi on `qubits`
// (matrix index bit b <-> qubits[b]) for the amplitudes whose bits in ctrl_mask are all set.  mat: a
// device copy of M, row-major, interleaved (re, im), in the state's precision.  allow_mfma: the
// complex64 5-qubit form may use the FP32 matrix cores.  Invalid arguments (k outside 1..5 or above
// n_qubits, n_qubits > 36, a repeated or out-of-range qubit, a control on a target) return
// hipErrorInvalidValue before anything is launched.
hipError_t launch_dense_sweep(int dtype, void* psi, int n_qubits, const int32_t* qubits, int k, uint64_t ctrl_mask,
                              const void* mat, bool allow_mfma, hipStream_t s);

// Whether launch_dense_sweep takes the matrix-core form for these arguments.
bool dense_sweep_uses_mfma(int dtype, int n_qubits, int k, bool allow_mfma);

}  // namespace mqa
