// schedule.h - host side: compiled circuits and the tile-pass scheduler (no HIP calls).
#pragma once
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "program.h"

namespace mqa {

using cplx = std::complex<double>;

// One validated gate of a circuit.
struct CGate {
    int kind = 0;                              // MQA_GATE_*
    std::vector<int> objs;                     // targets, obj[i] <=> matrix bit i
    uint64_t ctrl = 0;                         // control mask
    int slot = -1;                             // angle slot for parameterised kinds
    bool grad = false;                         // emits a derivative tap in the back-sweep
    double const_val = 0;                      // angle = const + sum coeff * param
    std::vector<std::pair<int, double>> terms;  // (param index, coefficient)
    std::vector<cplx> mat;                     // MATRIX kind, row-major
    int custom_off = -1;                       // PMATRIX kinds: offset (doubles) in a sample's custom row
    uint64_t qubits() const;                   // targets | controls
    uint64_t targets() const;
};

struct SchedOptions {
    int tile_qubits = 12;      // K: amplitudes per tile = 2^K
    int low_run = 5;           // C: lowest qubits always local (contiguous runs of 2^C amplitudes)
    int group_width = 4;       // max target positions per register group (1..4)
    int max_ops_per_pass = 0;  // 0: chosen by the pass cost model (build_plan); 1: one gate per pass
    int taps_per_group = 8;    // derivative taps a register group may hold (<= kMaxTapsPerGroup)
    int amp_bytes = 8;         // 8 (complex64) or 16 (complex128): selects the LDS swizzle
    bool phase_free = false;   // uncontrolled RX / RY / RZ / GP lower to their OPK_PF_* forms
    // RZ, PS and uncontrolled RZZ become terms of phase blocks (OPK_PHASE): they need no local
    // qubit, and a run of them that commutes with the gates around it costs one block
    bool phase_ops = false;
    bool operator==(const SchedOptions& o) const {
        return tile_qubits == o.tile_qubits && low_run == o.low_run && group_width == o.group_width &&
               max_ops_per_pass == o.max_ops_per_pass && taps_per_group == o.taps_per_group &&
               amp_bytes == o.amp_bytes && phase_free == o.phase_free && phase_ops == o.phase_ops;
    }
};

// A term of a phase block before flattening: the gate it comes from, its indicator over global
// qubits, its phase (PhaseCoef) and its tap (index among the pass's phase-block taps, -1: none).
struct HostPhaseTerm {
    int gate = -1;
    int kind = PHT_AND;
    uint64_t mask = 0;
    PhaseCoef coef{};
    int tap = -1;
    double weight = 0;
};
struct HostOp {
    int gate = -1;  // index into the circuit (a phase block: its first gate)
    DevOp dev{};
    std::vector<HostPhaseTerm> terms;  // OPK_PHASE: the block's terms
};
struct HostGroup {
    int pos[4] = {0, 1, 2, 3};
    std::vector<HostOp> ops;
    int tap_begin = 0;
    int n_taps = 0;
};
struct HostPass {
    std::vector<int> local_qubits;  // sorted ascending, size K
    std::vector<HostGroup> groups;
    std::vector<int> tap_slots;     // tap index inside the pass -> global tap id
};

// A complete schedule, flattened into the device arrays.
struct Plan {
    int n_eff = 0;   // qubits of the (padded) state the plan runs on
    int K = 0;       // tile qubits
    int C = 0;       // contiguous low run
    bool adjoint = false;
    SchedOptions opts;
    std::vector<HostPass> passes;
    std::vector<DevPass> dev_passes;
    std::vector<DevGroup> dev_groups;
    std::vector<DevOp> dev_ops;
    std::vector<int32_t> dev_tap_slots;
    std::vector<DevPhaseTerm> dev_terms;  // terms of the phase blocks (OPK_PHASE)
    std::vector<PhaseCoef> phase_coefs;   // phase of term i (column i of the per-sample phase table)
    std::vector<double> pool;  // matrix pool (interleaved complex)
    int max_taps = 0;
    int n_tap_ids = 0;  // gates with a derivative tap; tap ids follow circuit order
};

// Every circuit gate an op of a plan applies (a phase block applies several).
inline void op_gates(const HostOp& op, std::vector<int>& out) {
    if (op.dev.kind != OPK_PHASE) {
        out.push_back(op.gate);
        return;
    }
    for (const HostPhaseTerm& t : op.terms)
        if (out.empty() || out.back() != t.gate) out.push_back(t.gate);
}

// Build the schedule of `gates` (adjoint: the reverse order, each op applying G^dagger).
// n_eff: qubits of the state.  Throws std::invalid_argument on unsupported gates.
Plan build_plan(const std::vector<CGate>& gates, int n_eff, bool adjoint, const SchedOptions& opts);

// Matrix of a fixed (non-parameterised) gate kind; 2^m x 2^m row-major.
std::vector<cplx> fixed_matrix(const CGate& g);

// Passes of the Hamiltonian tile engine.  Tiles have K local qubits: the C lowest qubits and
// K - C higher ones per pass.  Flip groups are packed first-fit (largest first) into passes
// whose local qubits contain their flip mask.
struct HamTilePlan {
    int K = 0;
    int C = 0;
    std::vector<DevHamPass> passes;
    std::vector<DevHamTileGroup> groups;
    std::vector<DevHamTileTerm> terms;
};

// Returns false when a flip mask has more than K - C qubits above the low run (no tile holds
// it); the caller then uses the gather kernel.
bool build_ham_plan(const std::vector<DevHamGroup>& groups, const std::vector<DevHamTerm>& terms, int n_eff, int K,
                    int C, HamTilePlan& out);

}  // namespace mqa
