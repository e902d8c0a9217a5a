// program.h - the tile-pass program format shared by the host scheduler and the gfx950 kernels.
//
// A circuit is executed as a sequence of *passes*.  A pass streams every amplitude of the
// state(s) through LDS exactly once: the amplitude index space is split into tiles of 2^K
// amplitudes whose K "local" qubits are chosen per pass (always including the lowest qubits,
// so that global loads are runs of contiguous amplitudes).  Inside a tile, gates whose target
// qubits are all local are applied in *groups*: a group names 4 local positions; every thread
// gathers the 16 amplitudes spanned by those positions into registers, applies all of the
// group's ops there, and writes them back.  One pass therefore costs one HBM read and one
// HBM write of the state no matter how many gates it holds.
//
// The same program drives the adjoint back-sweep: there every op applies G^dagger to both the
// forward state and the co-state, and parameterised ops first emit a derivative tap.
#pragma once
#include <stdint.h>

namespace mqa {

constexpr int kGroupWidth = 4;                   // register positions per group
constexpr int kAmpsPerThread = 1 << kGroupWidth;  // 16 amplitudes per thread per state
constexpr int kMaxTapsPerPass = 96;               // LDS tap slots per pass
constexpr int kMaxTapsPerGroup = 8;               // register tap accumulators per thread
constexpr int kMinTileQubits = 6;                 // smallest tile: 4 threads
constexpr int kMaxTileQubits = 13;                // largest tile: 512 threads

// Device op kinds.
enum OpKind : int32_t {
    OPK_X = 1,   // swap the pair (X, CNOT, Toffoli...)
    // Diagonal phase block: a run of commuting diagonal gates (RZ, PS, RZZ) merged into one op
    // whose qubits need not be register positions (nor local): DevOp::mat = index of the first
    // DevPhaseTerm of the plan, DevOp::slot = number of terms.  Every amplitude i is multiplied by
    // prod_t z_t^{f_t(i)} (z_t = e^{i phi_t} from the per-sample phase table, conjugated in the
    // back-sweep); the back-sweep taps are Re(tap) += weight_t * sum_{f_t(i) = 1} Im(conj(lambda_i)
    // psi_i), which no phase of the block changes.
    OPK_PHASE = 2,
    OPK_M1 = 5,  // dense 2x2 from the pool (Y, Z, H, S, T, SX, 1-qubit matrices)
    OPK_SWAP = 6,
    OPK_M2 = 7,  // dense 4x4 from the pool (ISWAP, 2-qubit matrices)
    OPK_M3 = 8,  // dense 8x8 over group positions 0..2 (3-qubit matrices)
    OPK_M4 = 9,  // dense 16x16 over group positions 0..3 (4-qubit matrices)
    OPK_CX = 10, // X on q0 controlled by register position q1 (CNOT), plus any other controls
    OPK_CZ = 11, // Z on q0 controlled by register position q1 (CZ), plus any other controls
    OPK_H = 12,  // Hadamard: (a + b, a - b) / sqrt(2), half the operations of a dense 2x2
    // Compiled passes only (never in a plan): the uncontrolled Hadamard without its 1 / sqrt(2),
    // (a + b, a - b); the pass divides the tile by sqrt(2)^n once, before it stores it (n: the
    // pass's count), and each tap by 2^n (n: the count before the tap)
    OPK_HU = 13,
    // parameterised: the angle comes from the per-sample angle table
    OPK_RX = 16,
    OPK_RY = 17,
    OPK_RZ = 18,
    OPK_PS = 19,
    OPK_GP = 20,
    OPK_RXX = 21,
    OPK_RYY = 22,
    OPK_RZZ = 23,
    // Phase-free forms of uncontrolled rotations, for plans whose states only enter quantities
    // that do not see a global phase (the hermitian gradient: <psi|H|psi>, and taps of psi and
    // lambda, which carry the same phase).  Each drops a global phase that the forward op and
    // its back-sweep inverse drop alike:
    //   PF_RX, PF_RY: RX / RY with (cos, sin) of the half angle negated when cos < 0 (a factor
    //                 -1), so the rotation runs as three shears with |tan| <= 1;
    //   PF_RZ:        diag(1, e^{i theta}) = e^{i theta/2} RZ(theta), a phase on one amplitude;
    //   PF_GP:        the identity;
    //   PF_RZZ:       diag(1, e^{i theta}, e^{i theta}, 1) = e^{i theta/2} RZZ(theta), a phase on
    //                 the odd-parity pair.
    // Back-sweep taps: PF_RX / PF_RY as RX / RY; PF_RZ, PF_RZZ and PF_GP rely on <lambda|psi>
    // being real (lambda = H psi with H hermitian), so Re(tap) of RZ is -Im(sum over |1> of
    // conj(lambda) psi) (the PS tap), that of RZZ the same sum over the odd-parity pair, and that
    // of GP zero.
    OPK_PF_RX = 24,
    OPK_PF_RY = 25,
    OPK_PF_RZ = 26,
    OPK_PF_GP = 27,
    OPK_PF_RZZ = 28,
    // Parameterised dense 2x2 / 4x4 (custom gates): the matrices come from the per-sample table
    // TilePassArgs::custom at DevOp::mat (doubles): [G][G^dagger][M] with M = dG G^dagger, the tap
    // matrix (Re(tap) = Re <lambda|M|psi>, taken before G^dagger is applied); DevOp::pad0 holds the
    // target count.  PMTAP: the tap of a further parameter expression only ([M] at DevOp::mat).
    OPK_PM = 29,
    OPK_PMTAP = 30,
};

// Coefficient forms of a slot in the compiled passes' coefficient table (TilePassArgs::coefs),
// derived from the slot's (cos, sin) in the state's real type, so that a pass reads its
// coefficients ready-made instead of deriving them in every lane:
//   COEF_CS:     (cos, sin) of the angle table (parameterised ops in their exact form);
//   COEF_SHEAR:  Shear (t, s) of the sign-normalised pair (PF_RX, PF_RY);
//   COEF_DOUBLE: (cos, sin) of the doubled angle (PF_RZ, PF_RZZ).
// A back-sweep inverse negates the second entry (and t), exactly as it negates sin.  COEF_NONE:
// no entry (the slot has no op in the plan, or ops of two forms).
enum CoefForm : uint8_t { COEF_NONE = 0, COEF_CS = 1, COEF_SHEAR = 2, COEF_DOUBLE = 3 };

inline bool opk_is_param(int k) { return k >= OPK_RX; }
inline bool opk_is_two(int k) {
    return k == OPK_SWAP || k == OPK_M2 || k == OPK_RXX || k == OPK_RYY || k == OPK_RZZ || k == OPK_CX || k == OPK_CZ ||
           k == OPK_PF_RZZ;
}
inline bool opk_is_custom(int k) { return k == OPK_PM || k == OPK_PMTAP; }

// Byte layout of a tile pass's dynamic LDS: [psi tile][lambda tile (adjoint)][offset table]
// [tap partials].  Shared by the kernel and the host launcher so the two cannot disagree.
struct LdsLayout {
    uint32_t psi, lam, offtab, taps, total;
};
#if defined(__HIPCC__)
__host__ __device__
#endif
inline LdsLayout lds_layout(int K, int C, bool adj, int amp_bytes, int max_taps) {
    LdsLayout l;
    const uint32_t tile = (uint32_t(1) << K) * uint32_t(amp_bytes);
    const uint32_t n_threads = uint32_t(1) << (K - kGroupWidth);
    const uint32_t n_waves = (n_threads + 63) / 64;
    l.psi = 0;
    l.lam = tile;
    l.offtab = tile * (adj ? 2u : 1u);
    l.taps = (l.offtab + (uint32_t(1) << (K - C)) * 8u + 15u) & ~15u;
    l.total = l.taps + uint32_t(max_taps) * n_waves * 8u;  // one double (Re tap) per tap and wave
    return l;
}

struct DevOp {             // 32 bytes
    int32_t kind;          // OpKind
    int32_t slot;          // angle slot (param kinds), -1 otherwise
    int32_t tap;           // tap index inside the pass (adjoint), -1: none
    int32_t mat;           // offset in doubles into the matrix pool (fixed kinds)
    uint8_t q0, q1;        // register positions (0..3) of matrix bit 0 / bit 1
    uint8_t pad0, pad1;
    uint32_t ctrl_local;   // control qubits that are local, as a mask over tile positions
    uint64_t ctrl_global;  // control qubits that are not local, as a global index mask
};

// One term of a phase block (OPK_PHASE): its indicator f(i) over the amplitude index is
//   PHT_AND:    every qubit of the term set in i (an empty term: every amplitude);
//   PHT_PARITY: the parity of i over the term's qubits;
// the qubits split into tile positions (ml) and qubits outside the tile (mg, global index bits,
// uniform over a tile).
enum PhaseTermKind : int32_t { PHT_AND = 0, PHT_PARITY = 1 };
struct DevPhaseTerm {      // 32 bytes
    uint64_t mg;           // qubits outside the tile (global index mask)
    uint32_t ml;           // qubits inside the tile (tile-position mask)
    int32_t kind;          // PhaseTermKind
    int32_t coef;          // column of z_t in the per-sample phase table
    int32_t tap;           // pass tap index (back-sweep), -1: none
    double weight;         // Re(tap) += weight * sum_{f_t(i) = 1} Im(conj(lambda_i) psi_i)
};
// Phase of a term from the per-sample angle table: phi = k * alpha_slot + b, where alpha_slot is
// the angle the table holds for the slot ((cos, sin) of the half angle, or of the angle for PS /
// GP); slot < 0: the constant e^{i b}.
struct PhaseCoef {         // 32 bytes
    int32_t slot;
    int32_t k;             // -2..2
    double cos_b, sin_b;
    double pad;
};

// XOR swizzle of a tile index into its LDS slot.  It folds higher index bits into the low bits
// so that register groups whose positions include low qubits do not send every lane of a wave
// to one bank.  It is linear over GF(2): swz(a ^ b) == swz(a) ^ swz(b), which lets the kernel
// combine a per-thread base with per-group, wave-uniform register offsets.
#if defined(__HIPCC__)
__host__ __device__
#endif
inline uint32_t lds_swizzle(uint32_t u, int amp_bytes) {
    if (amp_bytes == 8) return u ^ (((u >> 5) ^ (u >> 10)) & 31u);
    return u ^ (((u >> 4) ^ (u >> 8) ^ (u >> 12)) & 15u);
}

struct DevGroup {          // 64 bytes
    uint8_t pos[4];        // local positions; register index bit i <=> pos[i]
    uint8_t spos[4];       // the same positions sorted ascending
    int32_t op_begin;
    int32_t n_ops;
    int32_t tap_begin;     // pass-level index of the group's first tap (taps are consecutive)
    int32_t n_taps;        // <= kMaxTapsPerGroup
    int32_t pad[2];
    uint16_t off[16];      // lds_swizzle(register offset of register v) for v = 0..15
};

struct DevPass {           // 48 bytes
    uint64_t nonlocal_mask;  // global qubits that index the tiles
    uint64_t local_hi_mask;  // local qubits above the contiguous low run
    int32_t group_begin;
    int32_t n_groups;
    int32_t tap_begin;       // first entry of this pass in the tap -> slot table
    int32_t n_taps;
    int32_t low_run;         // number of contiguous low qubits (C)
    int32_t pad;
};

// Pauli-term Hamiltonian grouped by flip mask (x|y): out[j] = sum_g sum_t coef_t *
// (-1)^{popcount(src & yz_t)} * psi[src], src = j ^ flip_g, coef_t = c_t * i^{|y_t|}.
struct DevHamGroup {
    uint64_t flip;
    int32_t term_begin;
    int32_t n_terms;
};
struct DevHamTerm {
    uint64_t yz;
    double coef_re;
    double coef_im;
    double pad;
};

// Arguments of one tile pass (static and generated kernels).
struct TilePassArgs {
    void* psi;          // [batch][stride] amplitudes
    void* lam;          // co-state (adjoint passes only)
    int K;              // tile qubits
    int pass_index;
    uint64_t stride;    // amplitudes between consecutive samples
    uint64_t n_tiles;   // tiles per sample = 2^(n - K)
    const DevPass* passes;
    const DevGroup* groups;
    const DevOp* ops;
    const void* pool;      // matrix pool in the state's real type (float or double)
    const double* angles;  // [batch][n_slots][2] (cos, sin) (NULL without parameterised ops)
    int n_slots;
    int n_tap_ids;
    const int32_t* tap_ids;
    double* partials;      // [batch][n_tap_ids][n_tiles] Re(tap)
    uint64_t batch;        // states
    const void* coefs;     // [batch][n_slots][2] in the real type, per CoefForm (compiled passes)
    const double* custom;  // per-sample matrices of parameterised dense ops (OPK_PM / OPK_PMTAP)
    uint64_t custom_stride;  // doubles between samples' rows (0: one row for every sample)
    // Forward passes: read the input tiles from src (sample b at src + b * src_stride) instead of
    // psi; src_stride 0 broadcasts one state to every sample (cached loads).  NULL: read psi.
    const void* src;
    uint64_t src_stride;
    // 1: the pass's output states are dead (the last back-sweep pass): only its taps are written.
    int32_t skip_store;
    // phase blocks (OPK_PHASE): the plan's terms and the per-sample table of their z_t
    // ([batch][n_phase][2] in the real type; the row of sample b at phase + 2 * b * n_phase)
    const DevPhaseTerm* terms;
    const void* phase;
    int32_t n_phase;
};

// Hamiltonian tile pass: the flip groups whose flip mask lies inside one tile's local qubits are
// applied from an LDS-staged ket tile; the passes of a plan together cover every group.
constexpr int kMaxHamTileQubits = 14;

struct DevHamPass {          // 32 bytes
    uint64_t nonlocal_mask;  // global qubits that index the tiles
    uint64_t local_hi_mask;  // local qubits above the contiguous low run
    int32_t group_begin;
    int32_t n_groups;
    int32_t low_run;         // C
    int32_t pad;
};
// Weight kinds of a group inside a pass: the sign (-1)^{popcount(src & yz)} of each term splits
// into a tile factor (yz bits outside the tile), a per-thread factor and a per-register factor.
enum HamWeightKind : int32_t {
    HAMW_TILE = 0,    // no term has yz bits inside the tile: one weight per tile
    HAMW_THREAD = 1,  // yz bits inside the tile only on thread-index bits: one weight per thread
    HAMW_FULL = 2,    // a weight per amplitude
    // a weight per amplitude through the register masks: the group's terms are sorted by the
    // 4-bit mask m of their yz bits on register positions (tile bits 0, K-3, K-2, K-1) and
    // seg_end[m] ends each mask's run; a thread sums each run's coefficients with its tile and
    // thread signs, and a 16-point Walsh-Hadamard transform gives every register's weight
    HAMW_WHT = 3,
};
struct DevHamTileGroup {     // 48 bytes
    uint32_t flipl;          // flip mask in tile-index coordinates
    int32_t term_begin;
    int32_t n_terms;
    int32_t kind;            // HamWeightKind
    uint16_t seg_end[16];    // HAMW_WHT: end of the terms of register mask m (relative to term_begin)
};
struct DevHamTileTerm {      // 32 bytes
    uint64_t yz_nl;          // yz bits outside the tile (global index positions)
    uint32_t yzl;            // yz bits inside the tile (tile-index coordinates)
    int32_t pad;
    double coef_re;          // c * i^{|y|}
    double coef_im;
};

}  // namespace mqa
