// launch.h - kernel argument blocks and launchers (kernels.hip) used by the C ABI (capi.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "program.h"

namespace mqa {

// dtype: 0 = complex64 (float2 amplitudes), 1 = complex128 (double2 amplitudes).


struct HamArgs {
    const void* ket;
    uint64_t ket_stride;
    const void* bra;  // NULL: no reduction
    uint64_t bra_stride;
    void* out;        // H ket (WRITE kernels)
    uint64_t out_stride;
    uint64_t dim;
    uint32_t amps_per_block;
    uint32_t n_blocks;
    const DevHamGroup* groups;
    int n_groups;
    const DevHamTerm* terms;
    double* partial;  // [batch][n_blocks] (re, im) or NULL
};

// One pass of the Hamiltonian tile engine: out (+)= sum over the pass's flip groups of the
// weighted ket gathered inside LDS tiles; the last pass also writes per-tile <ket|out> partials.
struct HamTileArgs {
    const void* ket;
    void* out;
    uint64_t stride;
    int K;
    int pass_index;
    const DevHamPass* passes;
    const DevHamTileGroup* groups;
    const DevHamTileTerm* terms;
    double* partial;  // [batch][n_tiles] (re, im) or NULL
    uint64_t n_tiles; // tiles per sample (partial row length)
};

struct VdotArgs {
    const void* a;
    uint64_t a_stride;
    const void* b;
    uint64_t b_stride;
    uint64_t dim;
    uint32_t amps_per_block;
    uint32_t n_blocks;
    double* partial;  // [batch][n_blocks] (re, im)
};

struct AngleArgs {
    const double* enc;         // [batch][n_enc]
    const double* ans;         // [n_ans]
    int n_enc;
    int n_slots;
    uint64_t batch;
    const double* const_val;   // [n_slots]
    const int32_t* term_begin; // [n_slots + 1]
    const int32_t* term_param;
    const double* term_coeff;
    const int8_t* half;        // [n_slots]
    double* out;               // [batch][n_slots][2]
};

size_t tile_pass_lds_bytes(int dtype, bool adj, int K, int C, int max_taps);
// LDS of an interpreted pass: back-sweeps without dense blocks share one tile buffer between
// psi and lambda (two only when `full`).
size_t interpreted_pass_lds_bytes(int dtype, bool adj, bool full, int K, int C, int max_taps);
// full: the pass holds dense 4x4 / 8x8 / 16x16 blocks (a heavier instantiation); phase: phase
// blocks (OPK_PHASE)
hipError_t launch_tile_pass(int dtype, bool adj, bool full, bool phase, const TilePassArgs& a, uint64_t batch, int C,
                            int max_taps, hipStream_t s);
hipError_t launch_ham(int dtype, bool write, const HamArgs& a, uint64_t batch, hipStream_t s);
hipError_t launch_ham_tile(int dtype, bool first, bool last, const HamTileArgs& a, int C, uint64_t n_tiles,
                           uint64_t batch, hipStream_t s);
// Several independent single-sample states in one launch: items (device array) with ascending
// first-block offsets prefix (device array, n_items entries); blocks = total tiles.
hipError_t launch_tile_pass_items(int dtype, bool adj, bool full, bool phase, const TilePassArgs* items,
                                  const uint64_t* prefix, int n_items, uint64_t blocks, int K, size_t lds,
                                  hipStream_t s);
hipError_t launch_ham_tile_items(int dtype, bool first, bool last, const HamTileArgs* items, const uint64_t* prefix,
                                 int n_items, uint64_t blocks, int K, int C, hipStream_t s);
hipError_t launch_unit_items(int dtype, void* base, const uint64_t* offset, int n, hipStream_t s);
hipError_t launch_reduce_segments(const double* in, const uint64_t* begin, const uint64_t* len, double* out,
                                  uint64_t n_segments, hipStream_t s);
hipError_t launch_vdot(int dtype, const VdotArgs& a, uint64_t batch, hipStream_t s);
hipError_t launch_fill(int dtype, void* dst, const void* src, uint64_t dim, uint64_t stride, uint64_t src_dim,
                       uint64_t batch, hipStream_t s);
hipError_t launch_cast(int src_dtype, const void* src, void* dst, uint64_t n_reals, hipStream_t s);
hipError_t launch_angles(const AngleArgs& a, hipStream_t s);
// Coefficient table [batch][n_slots][2] (dtype 0: float, 1: double) of an angle table, per the
// slots' CoefForm entries (forms[0, n_forms); slots beyond: the (cos, sin) pair).
hipError_t launch_coefs(int dtype, const double* angles, const uint8_t* forms, int n_forms, int n_slots, uint64_t batch,
                        void* out, hipStream_t s);
// Phase table of a plan's phase blocks ([batch][n_terms][2] in the real type; PhaseCoef).
hipError_t launch_phase_coefs(int dtype, const double* angles, const PhaseCoef* pc, int n_terms, int n_slots,
                              uint64_t batch, void* out, hipStream_t s);
hipError_t launch_reduce_rows(const double* in, uint64_t n, double* out, uint64_t rows, hipStream_t s);
hipError_t launch_reduce_rows1(const double* in, uint64_t n, double* out, uint64_t rows, hipStream_t s);
hipError_t launch_jacobian(const double* taps, int n_tap_ids, const int32_t* jac_begin, const int32_t* jac_tap,
                           const double* jac_coeff, int n_params, double* g, uint64_t batch, hipStream_t s);

}  // namespace mqa
