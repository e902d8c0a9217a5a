// schedule.cpp - cache-blocked tile-pass scheduler.
//
// The reference applies one gate per full sweep of the state (vector_state.tpp:948-960,
// one OpenMP loop per gate).  On MI355X every sweep is an HBM round trip, so this scheduler
// packs as many gates as the dependency order allows into each pass: a pass picks K local
// qubits, and every not-yet-applied gate whose targets are local and which does not depend
// on a deferred gate is applied inside the tiles.  Controls never need to be local.
#include "schedule.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>

#include "../../include/mqamd.h"

namespace mqa {

uint64_t CGate::targets() const {
    uint64_t m = 0;
    for (int q : objs) m |= uint64_t(1) << q;
    return m;
}
uint64_t CGate::qubits() const { return targets() | ctrl; }

static int popc(uint64_t x) { return __builtin_popcountll(x); }

std::vector<cplx> fixed_matrix(const CGate& g) {
    const double r = M_SQRT1_2;
    const cplx I(0, 1);
    switch (g.kind) {
        case MQA_GATE_X: return {0, 1, 1, 0};
        case MQA_GATE_Y: return {0, -I, I, 0};
        case MQA_GATE_Z: return {1, 0, 0, -1};
        case MQA_GATE_H: return {r, r, r, -r};
        case MQA_GATE_S: return {1, 0, 0, I};
        case MQA_GATE_SDAG: return {1, 0, 0, -I};
        case MQA_GATE_T: return {1, 0, 0, cplx(r, r)};
        case MQA_GATE_TDAG: return {1, 0, 0, cplx(r, -r)};
        case MQA_GATE_SX: return {cplx(0.5, 0.5), cplx(0.5, -0.5), cplx(0.5, -0.5), cplx(0.5, 0.5)};
        case MQA_GATE_SXDAG: return {cplx(0.5, -0.5), cplx(0.5, 0.5), cplx(0.5, 0.5), cplx(0.5, -0.5)};
        case MQA_GATE_SWAP: return {1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1};
        case MQA_GATE_ISWAP: return {1, 0, 0, 0, 0, 0, I, 0, 0, I, 0, 0, 0, 0, 0, 1};
        case MQA_GATE_ISWAPDAG: return {1, 0, 0, 0, 0, 0, -I, 0, 0, -I, 0, 0, 0, 0, 0, 1};
        case MQA_GATE_MATRIX: return g.mat;
        default: throw std::invalid_argument("gate kind " + std::to_string(g.kind) + " has no fixed matrix");
    }
}

static std::vector<cplx> dagger(const std::vector<cplx>& m) {
    size_t d = size_t(std::lround(std::sqrt(double(m.size()))));
    std::vector<cplx> h(m.size());
    for (size_t r = 0; r < d; r++)
        for (size_t c = 0; c < d; c++) h[r * d + c] = std::conj(m[c * d + r]);
    return h;
}

// Device op kind of a gate, plus its pool matrix (already daggered for the back-sweep).
static int lower_kind(const CGate& g, bool adjoint, std::vector<cplx>* mat) {
    mat->clear();
    switch (g.kind) {
        case MQA_GATE_X: return OPK_X;
        case MQA_GATE_H: return OPK_H;  // self-inverse: the same op in the back-sweep
        case MQA_GATE_SWAP: return OPK_SWAP;
        case MQA_GATE_Y:
        case MQA_GATE_Z:
        case MQA_GATE_S:
        case MQA_GATE_SDAG:
        case MQA_GATE_T:
        case MQA_GATE_TDAG:
        case MQA_GATE_SX:
        case MQA_GATE_SXDAG:
        case MQA_GATE_ISWAP:
        case MQA_GATE_ISWAPDAG:
        case MQA_GATE_MATRIX: {
            auto m = fixed_matrix(g);
            if (adjoint) m = dagger(m);
            *mat = m;
            switch (g.objs.size()) {
                case 1: return OPK_M1;
                case 2: return OPK_M2;
                case 3: return OPK_M3;
                case 4: return OPK_M4;
                default: throw std::invalid_argument("matrix gate acts on too many qubits");
            }
        }
        case MQA_GATE_RX: return OPK_RX;
        case MQA_GATE_RY: return OPK_RY;
        case MQA_GATE_RZ: return OPK_RZ;
        case MQA_GATE_PS: return OPK_PS;
        case MQA_GATE_GP: return OPK_GP;
        case MQA_GATE_RXX: return OPK_RXX;
        case MQA_GATE_RYY: return OPK_RYY;
        case MQA_GATE_RZZ: return OPK_RZZ;
        case MQA_GATE_PMATRIX: return OPK_PM;
        case MQA_GATE_PMATRIX_TAP: return OPK_PMTAP;
        default: throw std::invalid_argument("Apply of gate " + std::to_string(g.kind) + " not implement.");
    }
}

// Qubits a gate acts on other than diagonally: two gates commute when neither one's non-diagonal
// qubits meet the other's qubits (a diagonal gate commutes with diagonal gates and with gates
// that only use its qubits as controls).
static uint64_t nondiag_targets(const CGate& g) {
    switch (g.kind) {
        case MQA_GATE_Z:
        case MQA_GATE_S:
        case MQA_GATE_SDAG:
        case MQA_GATE_T:
        case MQA_GATE_TDAG:
        case MQA_GATE_RZ:
        case MQA_GATE_PS:
        case MQA_GATE_GP:
        case MQA_GATE_RZZ:
            return 0;
        default:
            return g.targets();
    }
}

// Terms of a gate that joins phase blocks (RZ, PS, uncontrolled RZZ, and the constant phases Z, S,
// T, their inverses and their controlled forms such as CZ; false for other gates).  The phase of a
// parameterised term is k * alpha + 0 with alpha the angle the slot's table entry holds (the half
// angle of RZ / RZZ, the angle of PS), and its tap weight -dphi / dtheta; a constant term has no
// slot (k = 0) and the phase b.  Phase-free plans drop the global phase of uncontrolled RZ / RZZ,
// as their OPK_PF_* forms do.
static bool phase_terms(const CGate& g, bool phase_free, std::vector<HostPhaseTerm>* out) {
    out->clear();
    auto constant = [&](double cb, double sb) {  // e^{i b} = (cb, sb) on the control subspace's |1>
        if (g.objs.size() != 1) return false;
        HostPhaseTerm t;
        t.kind = PHT_AND;
        t.mask = g.ctrl | (uint64_t(1) << g.objs[0]);
        t.coef.slot = -1;
        t.coef.k = 0;
        t.coef.cos_b = cb;
        t.coef.sin_b = sb;
        t.weight = 0.0;
        out->push_back(t);
        return true;
    };
    const double h = 0.70710678118654752440084436210484903928;
    switch (g.kind) {
        case MQA_GATE_Z: return constant(-1.0, 0.0);
        case MQA_GATE_S: return constant(0.0, 1.0);
        case MQA_GATE_SDAG: return constant(0.0, -1.0);
        case MQA_GATE_T: return constant(h, h);
        case MQA_GATE_TDAG: return constant(h, -h);
        default: break;
    }
    if (g.slot < 0) return false;
    auto add = [&](int kind, uint64_t mask, int k, double dalpha) {
        HostPhaseTerm t;
        t.kind = kind;
        t.mask = mask;
        t.coef.slot = g.slot;
        t.coef.k = k;
        t.coef.cos_b = 1.0;
        t.coef.sin_b = 0.0;
        t.weight = -double(k) * dalpha;
        out->push_back(t);
    };
    const uint64_t t0 = uint64_t(1) << g.objs[0];
    switch (g.kind) {
        case MQA_GATE_RZ:  // diag(e^{-i theta / 2}, e^{i theta / 2}) on the control subspace
            if (phase_free && g.ctrl == 0) {
                add(PHT_AND, t0, 2, 0.5);
            } else {
                add(PHT_AND, g.ctrl, -1, 0.5);
                add(PHT_AND, g.ctrl | t0, 2, 0.5);
            }
            return true;
        case MQA_GATE_PS:  // diag(1, e^{i theta})
            add(PHT_AND, g.ctrl | t0, 1, 1.0);
            return true;
        case MQA_GATE_RZZ:  // e^{-i theta / 2} on even parity, e^{i theta / 2} on odd
            if (g.ctrl != 0 || g.objs.size() != 2) return false;
            if (!phase_free) add(PHT_AND, 0, -1, 0.5);
            add(PHT_PARITY, t0 | (uint64_t(1) << g.objs[1]), 2, 0.5);
            return true;
        default:
            return false;
    }
}

namespace {

struct PassBuilder {
    const std::vector<CGate>& gates;
    bool adjoint;
    int W;
    int taps_per_group;
    bool phase_free;
    bool phase_ops;
    std::vector<int> tap_id_of_slot;  // angle slot -> global tap id (-1: no tap)
    std::vector<char> phase_ok;       // gate joins phase blocks (phase_ops plans)
    std::vector<uint64_t> nd;         // nondiag_targets (phase_ops plans), else every qubit

    // Fewest blocks for the pass's register groups: every gate of a phase block may move later,
    // up to the first op that does not commute with it, so its terms can sit anywhere in an
    // interval of op positions; one block at the right end of the interval that ends first covers
    // every interval reaching it (interval stabbing), and so on.
    void merge_blocks(HostPass* pass) const {
        struct Seq {
            int g, k;
            uint64_t nd;
        };
        std::vector<Seq> seq;  // the regular ops in execution order
        struct Item {
            int a, b;  // legal positions: after regular op a .. after regular op b - 1 (-1: before all)
            std::vector<HostPhaseTerm> terms;
        };
        std::vector<Item> items;  // one per gate, in execution order
        for (int g = 0; g < int(pass->groups.size()); g++) {
            for (int k = 0; k < int(pass->groups[size_t(g)].ops.size()); k++) {
                const HostOp& op = pass->groups[size_t(g)].ops[size_t(k)];
                if (op.dev.kind != OPK_PHASE) {
                    seq.push_back(Seq{g, k, nd[size_t(op.gate)]});
                    continue;
                }
                for (const HostPhaseTerm& t : op.terms) {
                    if (items.empty() || items.back().terms.back().gate != t.gate || items.back().a != int(seq.size()) - 1)
                        items.push_back(Item{int(seq.size()) - 1, 0, {}});
                    items.back().terms.push_back(t);
                }
            }
        }
        if (items.empty()) return;
        for (Item& it : items) {
            const uint64_t q = gates[size_t(it.terms.front().gate)].qubits();
            it.b = int(seq.size());
            for (int j = it.a + 1; j < int(seq.size()); j++)
                if (seq[size_t(j)].nd & q) {
                    it.b = j;
                    break;
                }
        }
        std::vector<int> order(items.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = int(i);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return items[size_t(x)].b < items[size_t(y)].b; });
        std::vector<int> at(items.size(), -2);  // chosen position of each item
        for (int i : order) {
            if (at[size_t(i)] != -2) continue;
            const int p = items[size_t(i)].b - 1;
            for (size_t j = 0; j < items.size(); j++)
                if (at[j] == -2 && items[j].a <= p && items[j].b > p) at[j] = p;
        }
        // rebuild the groups: the regular ops, each followed by the block of its position
        std::map<int, HostOp> blocks;
        for (size_t i = 0; i < items.size(); i++) {
            HostOp& op = blocks[at[i]];
            if (op.terms.empty()) {
                op.gate = items[i].terms.front().gate;
                op.dev.kind = OPK_PHASE;
                op.dev.tap = -1;
                op.dev.mat = -1;
            }
            for (auto& t : items[i].terms) op.terms.push_back(t);
        }
        for (auto& kv : blocks) kv.second.dev.slot = int(kv.second.terms.size());
        std::vector<std::vector<HostOp>> ops(pass->groups.size());
        auto emit = [&](int pos, int g) {
            auto f = blocks.find(pos);
            if (f != blocks.end()) ops[size_t(g)].push_back(std::move(f->second)), blocks.erase(f);
        };
        // a block before everything goes into the first group, unless that group is a dense 3- / 4-qubit
        // block (which reads its inputs from the tile): then into a group of its own
        HostOp lead;
        bool has_lead = false;
        for (int i = 0; i < int(seq.size()); i++) {
            const Seq& sq = seq[size_t(i)];
            HostOp& cur = pass->groups[size_t(sq.g)].ops[size_t(sq.k)];
            if (i == 0) {
                if (cur.dev.kind == OPK_M3 || cur.dev.kind == OPK_M4) {
                    auto f = blocks.find(-1);
                    if (f != blocks.end()) lead = std::move(f->second), blocks.erase(f), has_lead = true;
                } else {
                    emit(-1, sq.g);
                }
            }
            ops[size_t(sq.g)].push_back(std::move(cur));
            emit(i, sq.g);
        }
        if (seq.empty()) emit(-1, 0);
        std::vector<HostGroup> kept;
        if (has_lead) {
            HostGroup grp;
            std::vector<int> gpos;
            pad_positions(grp, gpos, int(pass->local_qubits.size()));
            grp.tap_begin = pass->groups.front().tap_begin;
            grp.ops.push_back(std::move(lead));
            kept.push_back(std::move(grp));
        }
        for (size_t g = 0; g < pass->groups.size(); g++) {
            pass->groups[g].ops = std::move(ops[g]);
            if (!pass->groups[g].ops.empty()) kept.push_back(std::move(pass->groups[g]));  // no empty groups
        }
        pass->groups = std::move(kept);
    }

    static void pad_positions(HostGroup& grp, std::vector<int>& gpos, int K) {
        for (int p = 0; int(gpos.size()) < kGroupWidth && p < K; p++)
            if (std::find(gpos.begin(), gpos.end(), p) == gpos.end()) gpos.push_back(p);
        for (int i = 0; i < kGroupWidth; i++) grp.pos[i] = gpos[i];
    }

    // Split the pass's gates (dependency-ordered) into register groups of <= W target positions.
    // Split the pass's gates (dependency-ordered) into register groups of <= W target positions.
    // Phase-block gates (phase_ops) wait as pending terms while everything applied after them
    // commutes with them; a gate that does not commute with a pending term waits for the block,
    // and the block is applied (at the end of the last group) once nothing else can proceed.
    void groups(HostPass* pass, const std::vector<int>& chosen, const std::vector<int>& pos_of, uint64_t nonlocal,
                std::vector<double>* pool) {
        const int K = int(pass->local_qubits.size());
        std::vector<int> todo = chosen;
        std::vector<HostPhaseTerm> pending;
        uint64_t pending_q = 0;
        std::vector<int> phase_taps;  // global tap ids of the pass's phase-block taps, in tap order
        std::vector<HostPhaseTerm> terms;
        // the pending terms as one op at the end of group `into` (null: the pass's last group)
        auto flush = [&](HostGroup* into) {
            if (pending.empty()) return;
            if (!into && pass->groups.empty()) {  // a pass that opens with a block: a group of its own
                HostGroup grp;
                grp.tap_begin = int(pass->tap_slots.size());
                std::vector<int> gpos;
                pad_positions(grp, gpos, K);
                pass->groups.push_back(std::move(grp));
            }
            HostOp op;
            op.gate = pending.front().gate;
            op.dev.kind = OPK_PHASE;
            op.dev.slot = int(pending.size());
            op.dev.tap = -1;
            op.dev.mat = -1;
            op.terms = std::move(pending);
            (into ? into : &pass->groups.back())->ops.push_back(std::move(op));
            pending.clear();
            pending_q = 0;
        };
        while (!todo.empty()) {
            HostGroup grp;
            grp.tap_begin = int(pass->tap_slots.size());
            std::vector<int> gpos;  // positions in register order
            std::vector<int> keep;
            uint64_t bq = 0, bt = 0;  // qubits of the deferred gates, and their non-diagonal qubits
            bool closed = false;
            int consumed = 0;
            for (int gi : todo) {
                const CGate& g = gates[gi];
                uint64_t q = g.qubits();
                const uint64_t ndq = nd[size_t(gi)];
                const bool in_block = phase_ok[size_t(gi)] != 0;
                const bool taps_full =
                    !in_block && adjoint && g.grad && g.slot >= 0 && grp.n_taps >= taps_per_group;
                // a gate waits behind a deferred gate it does not commute with; one that does not
                // commute with pending terms (phase_ops) has them applied first, in this group,
                // when it joins the group, and waits for them otherwise
                const bool after_block = !in_block && (ndq & pending_q) != 0;
                if (closed || (ndq & bq) || (q & bt) || taps_full) {
                    bq |= q;
                    bt |= ndq;
                    keep.push_back(gi);
                    continue;
                }
                if (in_block) {
                    phase_terms(g, phase_free, &terms);
                    int tap = -1;
                    if (adjoint && g.grad && g.slot >= 0) {
                        tap = int(phase_taps.size());
                        phase_taps.push_back(tap_id_of_slot[size_t(g.slot)]);
                    }
                    for (auto& t : terms) {
                        t.gate = gi;
                        t.tap = tap;
                        pending.push_back(t);
                    }
                    pending_q |= q;
                    consumed++;
                    continue;
                }
                std::vector<cplx> m;
                int kind = lower_kind(g, adjoint, &m);
                // a parameterised dense op opens a register group, so that its targets take register
                // positions 0 (and 1): the interpreted pass then needs one instantiation per shape
                if (opk_is_custom(kind) && !grp.ops.empty()) {
                    bq |= q;
                    bt |= ndq;
                    keep.push_back(gi);
                    continue;
                }
                int reg_ctrl = -1;  // control qubit bound to a register position (CX / CZ)
                if (kind == OPK_M3 || kind == OPK_M4) {
                    if (!grp.ops.empty()) {  // dense multi-qubit blocks take a group of their own
                        bq |= q;
                        bt |= ndq;
                        keep.push_back(gi);
                        continue;
                    }
                    for (int t : g.objs) gpos.push_back(pos_of[t]);
                    closed = true;
                } else {
                    std::vector<int> merged = gpos;
                    for (int t : g.objs)
                        if (std::find(merged.begin(), merged.end(), pos_of[t]) == merged.end()) merged.push_back(pos_of[t]);
                    // CNOT / CZ: one local control is bound to a register position, so that the
                    // control bit is a compile-time property of each register (no predicate)
                    const uint64_t lctrl = g.ctrl & ~nonlocal;
                    int want_ctrl = -1;
                    if ((g.kind == MQA_GATE_X || g.kind == MQA_GATE_Z) && lctrl) {
                        for (int c = 0; c < 64 && want_ctrl < 0; c++)
                            if (((lctrl >> c) & 1) && std::find(merged.begin(), merged.end(), pos_of[c]) != merged.end())
                                want_ctrl = c;
                        for (int c = 0; c < 64 && want_ctrl < 0; c++)
                            if ((lctrl >> c) & 1) want_ctrl = c;
                        if (std::find(merged.begin(), merged.end(), pos_of[want_ctrl]) == merged.end())
                            merged.push_back(pos_of[want_ctrl]);
                    }
                    if (int(merged.size()) > W && !grp.ops.empty()) {
                        bq |= q;
                        bt |= ndq;
                        keep.push_back(gi);
                        continue;
                    }
                    if (want_ctrl >= 0) {
                        reg_ctrl = want_ctrl;
                        kind = (g.kind == MQA_GATE_X) ? OPK_CX : OPK_CZ;
                        m.clear();
                    }
                    gpos = merged;
                }
                // dense 3- / 4-qubit blocks read their inputs from the tile, so a block before one
                // goes to the end of the previous group
                if (after_block) flush((kind == OPK_M3 || kind == OPK_M4) ? nullptr : &grp);
                HostOp op;
                op.gate = gi;
                DevOp& d = op.dev;
                if (phase_free && g.ctrl == 0) {
                    if (kind == OPK_RX) kind = OPK_PF_RX;
                    else if (kind == OPK_RY) kind = OPK_PF_RY;
                    else if (kind == OPK_RZ) kind = OPK_PF_RZ;
                    else if (kind == OPK_GP) kind = OPK_PF_GP;
                    else if (kind == OPK_RZZ) kind = OPK_PF_RZZ;
                }
                d.kind = kind;
                d.slot = opk_is_param(kind) ? g.slot : -1;
                d.tap = -1;
                if (adjoint && opk_is_param(kind) && g.grad) {
                    d.tap = int(pass->tap_slots.size());
                    pass->tap_slots.push_back(tap_id_of_slot[g.slot]);
                    grp.n_taps++;
                }
                d.mat = -1;
                d.pad0 = 0;
                if (opk_is_custom(kind)) {  // per-sample matrices: the offset in a sample's row
                    d.mat = g.custom_off;
                    d.pad0 = uint8_t(g.objs.size());
                }
                if (!m.empty()) {
                    d.mat = int(pool->size());
                    for (auto& c : m) {
                        pool->push_back(c.real());
                        pool->push_back(c.imag());
                    }
                }
                auto reg = [&](int t) {
                    return int(std::find(gpos.begin(), gpos.end(), pos_of[t]) - gpos.begin());
                };
                d.q0 = uint8_t(reg(g.objs[0]));
                d.q1 = uint8_t(reg_ctrl >= 0 ? reg(reg_ctrl) : (g.objs.size() > 1 ? reg(g.objs[1]) : 0));
                d.ctrl_global = g.ctrl & nonlocal;
                d.ctrl_local = 0;
                const uint64_t rest = g.ctrl & ~nonlocal & ~(reg_ctrl >= 0 ? (uint64_t(1) << reg_ctrl) : uint64_t(0));
                for (int c = 0; c < 64; c++)
                    if ((rest >> c) & 1) d.ctrl_local |= uint32_t(1) << pos_of[c];
                grp.ops.push_back(op);
                consumed++;
            }
            if (!grp.ops.empty()) {  // pad the group to 4 distinct positions
                pad_positions(grp, gpos, K);
                pass->groups.push_back(std::move(grp));
            }
            if (consumed == 0) flush(nullptr);  // nothing could proceed past the pending terms
            todo = keep;
        }
        flush(nullptr);
        if (phase_ops) merge_blocks(pass);
        // the phase-block taps follow the register groups' taps of the pass
        const int off = int(pass->tap_slots.size());
        for (auto& grp : pass->groups)
            for (auto& op : grp.ops)
                for (auto& t : op.terms)
                    if (t.tap >= 0) t.tap += off;
        pass->tap_slots.insert(pass->tap_slots.end(), phase_taps.begin(), phase_taps.end());
    }
};

}  // namespace

Plan build_plan(const std::vector<CGate>& gates, int n_eff, bool adjoint, const SchedOptions& opts) {
    Plan plan;
    plan.n_eff = n_eff;
    plan.adjoint = adjoint;
    plan.opts = opts;
    if (n_eff < kMinTileQubits) throw std::invalid_argument("state too small for the tile engine");
    const uint64_t all = (n_eff >= 64) ? ~uint64_t(0) : ((uint64_t(1) << n_eff) - 1);
    int K = std::min(std::max(opts.tile_qubits, kMinTileQubits), std::min(n_eff, kMaxTileQubits));
    // at least one contiguous low qubit: the kernels move complex64 amplitudes in pairs
    int C = std::min(std::max(opts.low_run, 1), K - kGroupWidth);
    const bool whole = (n_eff <= K);
    if (whole) C = K;
    const int W = std::min(std::max(opts.group_width, 1), kGroupWidth);
    plan.K = K;
    plan.C = C;

    std::vector<int> remaining;
    for (size_t i = 0; i < gates.size(); i++) {
        const CGate& g = gates[i];
        if (g.kind == MQA_GATE_I) continue;
        if (g.objs.empty() || int(g.objs.size()) > MQA_MAX_OBJ) throw std::invalid_argument("bad target count");
        if ((g.qubits() & ~all) != 0) throw std::invalid_argument("gate acts on a qubit outside the state");
        if (g.targets() & g.ctrl) throw std::invalid_argument("a qubit cannot be both target and control");
        if (popc(g.targets()) != int(g.objs.size())) throw std::invalid_argument("repeated target qubit");
        remaining.push_back(int(i));
    }
    if (adjoint) std::reverse(remaining.begin(), remaining.end());

    PassBuilder pb{gates, adjoint, W, std::min(std::max(opts.taps_per_group, 1), kMaxTapsPerGroup), opts.phase_free,
                   opts.phase_ops, {}, {}, {}};
    pb.phase_ok.assign(gates.size(), 0);
    pb.nd.resize(gates.size());
    {
        std::vector<HostPhaseTerm> scratch;
        for (size_t i = 0; i < gates.size(); i++) {
            // without phase blocks, any shared qubit orders two gates (the original packing)
            pb.nd[i] = opts.phase_ops ? nondiag_targets(gates[i]) : gates[i].qubits();
            pb.phase_ok[i] = (opts.phase_ops && phase_terms(gates[i], opts.phase_free, &scratch)) ? 1 : 0;
        }
    }
    int n_slots = 0;
    for (auto& g : gates) n_slots = std::max(n_slots, g.slot + 1);
    pb.tap_id_of_slot.assign(n_slots, -1);
    for (auto& g : gates)
        if (g.slot >= 0 && g.grad) pb.tap_id_of_slot[g.slot] = plan.n_tap_ids++;
    const uint64_t lowmask = (C >= 64) ? all : ((uint64_t(1) << C) - 1);
    // Greedy partition into passes: each pass takes, in order, every gate whose targets still fit
    // its local set and that no deferred gate blocks; cap > 0 closes a pass after cap gates.
    // Phase-block gates need no local qubit.
    struct Part {
        std::vector<int> chosen;
        uint64_t L;
    };
    auto partition = [&](int cap) {
        std::vector<Part> parts;
        std::vector<int> todo = remaining;
        while (!todo.empty()) {
            uint64_t L = whole ? all : lowmask;
            uint64_t bq = 0, bt = 0;  // qubits of the deferred gates, and their non-diagonal qubits
            std::vector<int> chosen, rest;
            int taps = 0;
            for (int gi : todo) {
                const CGate& g = gates[gi];
                uint64_t q = g.qubits();
                const uint64_t ndq = pb.nd[size_t(gi)];
                bool needs_tap = adjoint && g.grad && g.slot >= 0;
                bool full = (cap > 0 && int(chosen.size()) >= cap) || (needs_tap && taps >= kMaxTapsPerPass);
                if (full || (ndq & bq) || (q & bt)) {
                    bq |= q;
                    bt |= ndq;
                    rest.push_back(gi);
                    continue;
                }
                uint64_t nl = pb.phase_ok[size_t(gi)] ? L : (L | g.targets());
                if (popc(nl) <= K) {
                    L = nl;
                    chosen.push_back(gi);
                    taps += needs_tap ? 1 : 0;
                } else {
                    bq |= q;
                    bt |= ndq;
                    rest.push_back(gi);
                }
            }
            if (chosen.empty()) throw std::runtime_error("scheduler made no progress");
            parts.push_back(Part{std::move(chosen), L});
            todo = std::move(rest);
        }
        return parts;
    };
    // Model of a pass: max(1, work / B) streams of the state, work = parameterised ops + 0.3 per
    // other op (register renames, fixed matrices and phase-block terms), B = the work at which a
    // pass's arithmetic takes as long as its streaming (measured on MI355X: about 32 for a
    // complex64 back-sweep, 48 for a forward sweep; twice that for complex128, which streams twice
    // the bytes).  Greedy packing can leave one arithmetic-bound pass and stragglers; without an
    // explicit cap, the partition of lowest modelled time among a few caps is kept.
    auto cost = [&](const std::vector<Part>& parts) {
        const double B = (adjoint ? 32.0 : 48.0) * (opts.amp_bytes > 8 ? 2.0 : 1.0);
        double t = 0;
        for (const auto& p : parts) {
            double w = 0;
            for (int gi : p.chosen) w += (gates[gi].slot >= 0 && !pb.phase_ok[size_t(gi)]) ? 1.0 : 0.3;
            t += std::max(1.0, w / B);
        }
        return t;
    };
    std::vector<Part> parts = partition(opts.max_ops_per_pass);
    if (opts.max_ops_per_pass == 0) {
        double best = cost(parts);
        for (int cap : {96, 80, 64, 50, 40}) {
            auto alt = partition(cap);
            const double c = cost(alt);
            if (c < best - 1e-9) best = c, parts = std::move(alt);
        }
    }
    for (auto& part : parts) {
        uint64_t L = part.L;
        for (int q = 0; popc(L) < K && q < n_eff; q++) L |= uint64_t(1) << q;
        HostPass pass;
        std::vector<int> pos_of(64, -1);
        for (int q = 0; q < n_eff; q++)
            if ((L >> q) & 1) {
                pos_of[q] = int(pass.local_qubits.size());
                pass.local_qubits.push_back(q);
            }
        pb.groups(&pass, part.chosen, pos_of, all & ~L, &plan.pool);
        plan.passes.push_back(std::move(pass));
    }

    // flatten
    for (auto& pass : plan.passes) {
        DevPass dp{};
        uint64_t L = 0;
        for (int q : pass.local_qubits) L |= uint64_t(1) << q;
        dp.nonlocal_mask = all & ~L;
        dp.local_hi_mask = L & ~lowmask;
        dp.low_run = C;
        dp.group_begin = int(plan.dev_groups.size());
        dp.n_groups = int(pass.groups.size());
        dp.tap_begin = int(plan.dev_tap_slots.size());
        dp.n_taps = int(pass.tap_slots.size());
        plan.max_taps = std::max(plan.max_taps, dp.n_taps);
        for (int s : pass.tap_slots) plan.dev_tap_slots.push_back(s);
        for (auto& grp : pass.groups) {
            DevGroup dg{};
            int sp[4];
            for (int i = 0; i < 4; i++) {
                dg.pos[i] = uint8_t(grp.pos[i]);
                sp[i] = grp.pos[i];
            }
            std::sort(sp, sp + 4);
            for (int i = 0; i < 4; i++) dg.spos[i] = uint8_t(sp[i]);
            dg.op_begin = int(plan.dev_ops.size());
            dg.n_ops = int(grp.ops.size());
            dg.tap_begin = grp.tap_begin;
            dg.n_taps = grp.n_taps;
            for (int v = 0; v < 16; v++) {
                uint32_t u = 0;
                for (int i = 0; i < 4; i++)
                    if ((v >> i) & 1) u |= uint32_t(1) << grp.pos[i];
                dg.off[v] = uint16_t(lds_swizzle(u, opts.amp_bytes));
            }
            for (auto& op : grp.ops) {
                DevOp d = op.dev;
                if (d.kind == OPK_PHASE) {  // terms: tile positions and the qubits outside the tile
                    d.mat = int(plan.dev_terms.size());
                    d.slot = int(op.terms.size());
                    for (const HostPhaseTerm& t : op.terms) {
                        DevPhaseTerm dt{};
                        dt.mg = t.mask & dp.nonlocal_mask;
                        dt.ml = 0;
                        for (size_t p = 0; p < pass.local_qubits.size(); p++)
                            if ((t.mask >> pass.local_qubits[p]) & 1) dt.ml |= uint32_t(1) << p;
                        dt.kind = t.kind;
                        dt.coef = int(plan.phase_coefs.size());
                        dt.tap = t.tap;
                        dt.weight = t.weight;
                        plan.phase_coefs.push_back(t.coef);
                        plan.dev_terms.push_back(dt);
                    }
                }
                plan.dev_ops.push_back(d);
            }
            plan.dev_groups.push_back(dg);
        }
        plan.dev_passes.push_back(dp);
    }
    if (plan.pool.empty()) plan.pool.push_back(0.0);
    if (plan.dev_ops.empty()) plan.dev_ops.push_back(DevOp{});
    if (plan.dev_groups.empty()) plan.dev_groups.push_back(DevGroup{});
    if (plan.dev_tap_slots.empty()) plan.dev_tap_slots.push_back(0);
    return plan;
}


namespace {
uint64_t pext64(uint64_t src, uint64_t mask) {
    uint64_t out = 0;
    int k = 0;
    for (uint64_t m = mask; m; m &= m - 1, k++)
        if (src & (m & (~m + 1))) out |= uint64_t(1) << k;
    return out;
}
}  // namespace

bool build_ham_plan(const std::vector<DevHamGroup>& groups, const std::vector<DevHamTerm>& terms, int n_eff, int K,
                    int C, HamTilePlan& out) {
    if (n_eff < kMinTileQubits) throw std::invalid_argument("state too small for the tile engine");
    K = std::min(std::max(K, kMinTileQubits), std::min(n_eff, kMaxHamTileQubits));
    C = std::min(std::max(C, 1), K - kGroupWidth);
    if (n_eff <= K) C = K;
    const int H = K - C;
    const uint64_t all = (n_eff >= 64) ? ~uint64_t(0) : ((uint64_t(1) << n_eff) - 1);
    const uint64_t low = (uint64_t(1) << C) - 1;

    // first-fit decreasing packing of the high flip bits into passes of H local qubits
    std::vector<int> order(groups.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = int(i);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return __builtin_popcountll(groups[a].flip & ~low) > __builtin_popcountll(groups[b].flip & ~low);
    });
    struct P {
        uint64_t hi = 0;
        std::vector<int> members;
    };
    std::vector<P> ps;
    for (int gi : order) {
        const uint64_t hi = groups[gi].flip & ~low;
        if ((groups[gi].flip & ~all) != 0 || __builtin_popcountll(hi) > H) return false;
        bool placed = false;
        for (auto& p : ps) {
            if (__builtin_popcountll(p.hi | hi) <= H) {
                p.hi |= hi;
                p.members.push_back(gi);
                placed = true;
                break;
            }
        }
        if (!placed) ps.push_back(P{hi, {gi}});
    }
    if (ps.empty()) ps.push_back(P{});

    HamTilePlan plan;
    plan.K = K;
    plan.C = C;
    const uint32_t tmask = ((uint32_t(1) << (K - 3)) - 1u) & ~1u;  // thread-index bits of a tile index
    for (auto& p : ps) {
        // pad the high local set to exactly H qubits with the lowest unused ones
        for (int q = C; q < n_eff && __builtin_popcountll(p.hi) < H; q++) p.hi |= uint64_t(1) << q;
        DevHamPass dp{};
        dp.local_hi_mask = p.hi;
        dp.nonlocal_mask = all & ~low & ~p.hi;
        dp.group_begin = int32_t(plan.groups.size());
        dp.n_groups = int32_t(p.members.size());
        dp.low_run = C;
        std::sort(p.members.begin(), p.members.end());
        for (int gi : p.members) {
            const DevHamGroup& g = groups[gi];
            DevHamTileGroup tg{};
            tg.flipl = uint32_t((g.flip & low) | (pext64(g.flip, p.hi) << C));
            tg.term_begin = int32_t(plan.terms.size());
            tg.n_terms = g.n_terms;
            tg.kind = HAMW_TILE;
            for (int t = g.term_begin; t < g.term_begin + g.n_terms; t++) {
                DevHamTileTerm tt{};
                tt.yz_nl = terms[t].yz & dp.nonlocal_mask;
                tt.yzl = uint32_t((terms[t].yz & low) | (pext64(terms[t].yz, p.hi) << C));
                tt.coef_re = terms[t].coef_re;
                tt.coef_im = terms[t].coef_im;
                if (tt.yzl & ~tmask) tg.kind = HAMW_FULL;
                else if (tt.yzl && tg.kind == HAMW_TILE) tg.kind = HAMW_THREAD;
                plan.terms.push_back(tt);
            }
            if (tg.kind == HAMW_FULL && tg.n_terms <= 0xffff) {
                // sort by register mask (bit 0 <- tile bit 0, bits 1..3 <- tile bits K-3..K-1)
                auto reg_mask = [K](uint32_t yzl) { return (yzl & 1u) | (((yzl >> (K - 3)) & 7u) << 1); };
                auto first = plan.terms.begin() + tg.term_begin;
                std::stable_sort(first, plan.terms.end(), [&](const DevHamTileTerm& a, const DevHamTileTerm& b) {
                    return reg_mask(a.yzl) < reg_mask(b.yzl);
                });
                int cnt = 0;
                for (uint32_t m = 0; m < 16; m++) {
                    for (auto it = first + cnt; it != plan.terms.end() && reg_mask(it->yzl) == m; ++it) cnt++;
                    tg.seg_end[m] = uint16_t(cnt);
                }
                tg.kind = HAMW_WHT;
            }
            plan.groups.push_back(tg);
        }
        plan.passes.push_back(dp);
    }
    out = std::move(plan);
    return true;
}

}  // namespace mqa
