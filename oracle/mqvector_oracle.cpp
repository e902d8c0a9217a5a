// mqvector_oracle.cpp - TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference MindQuantum `mqvector` simulator path (state-vector
// gate application, Pauli-term Hamiltonian, expectation and the hermitian adjoint gradient).
// It is the parity checker for the HIP library and the CPU baseline timed beside it.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product library never links or calls it.
//
// Every routine follows the reference algorithm step by step (file:line relative to the
// mindquantum 0.12.0 tree) but is written from scratch here: same loop structure over
// index masks, same OpenMP threshold (dim >= 2^13, ccsrc/include/core/utils.h:42-49,
// cpu_vector_policy.h:47), same arithmetic type (float for c64, double for c128), same
// hermitian back-sweep (gradient.h:160-197).  It uses the gate / term IR of include/mqamd.h.
//
// Parity pins (tests/test_oracle.py): reference golden values (simulator.py:369 docstring,
// test_operations.py:49-62, test_simulator.py:664-677), dense numpy simulation (n <= 10),
// closed forms, central finite differences, parameter-shift identity, hermitian round trip.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../include/mqamd.h"

namespace oracle {

// OpenMP threads of each ParallelTask worker (0: the OpenMP default).  A std::thread starts with the
// default thread count, not the one omp_set_num_threads set on the caller, so the layouts of the
// reference's thread balance (batch threads x OpenMP threads) are set here explicitly.
static int g_worker_threads = 0;

using index_t = uint64_t;
constexpr index_t kDimTh = index_t(1) << 13;  // cpu_vector_policy.h:47

// Host copy of one gate of the IR.
struct Gate {
    int kind = 0;
    std::vector<int> objs;
    index_t ctrl_mask = 0;
    double const_val = 0;
    std::vector<std::pair<int, double>> terms;  // (param index, coefficient)
    std::vector<std::complex<double>> mat;      // MATRIX kind, row-major
    bool requires_grad = false;
};

static bool is_param_kind(int k) { return k >= MQA_GATE_RX && k <= MQA_GATE_RZZ; }

// Hermitian conjugate of one gate, as mindquantum builds it for herm_circ:
// parameterised gates negate every coefficient (core/gates/basic.py:510-514), S/T/SX/ISWAP
// toggle their dagger flag (basicgate.py:536-614), dense matrices are conjugate-transposed.
static Gate hermitian_gate(const Gate& g) {
    Gate h = g;
    if (is_param_kind(g.kind)) {
        h.const_val = -g.const_val;
        for (auto& t : h.terms) t.second = -t.second;
        return h;
    }
    switch (g.kind) {
        case MQA_GATE_S: h.kind = MQA_GATE_SDAG; break;
        case MQA_GATE_SDAG: h.kind = MQA_GATE_S; break;
        case MQA_GATE_T: h.kind = MQA_GATE_TDAG; break;
        case MQA_GATE_TDAG: h.kind = MQA_GATE_T; break;
        case MQA_GATE_SX: h.kind = MQA_GATE_SXDAG; break;
        case MQA_GATE_SXDAG: h.kind = MQA_GATE_SX; break;
        case MQA_GATE_ISWAP: h.kind = MQA_GATE_ISWAPDAG; break;
        case MQA_GATE_ISWAPDAG: h.kind = MQA_GATE_ISWAP; break;
        case MQA_GATE_MATRIX: {
            size_t m = size_t(1) << g.objs.size();
            for (size_t r = 0; r < m; r++)
                for (size_t c = 0; c < m; c++) h.mat[r * m + c] = std::conj(g.mat[c * m + r]);
        } break;
        default: break;  // X, Y, Z, H, SWAP, I are self-inverse
    }
    return h;
}

template <typename T>
struct Sim {
    using C = std::complex<T>;

    // ---- index masks (ccsrc/lib/simulator/utils.cpp:41-74, simulator/utils.h:54-58) ----
    struct Mask1 {
        index_t obj_mask, ctrl_mask, low, high;
        Mask1(int q, index_t ctrl) : obj_mask(index_t(1) << q), ctrl_mask(ctrl), low(obj_mask - 1), high(~(obj_mask - 1)) {}
        index_t at(index_t l) const { return ((l & high) << 1) + (l & low); }
    };
    struct Mask2 {
        index_t obj_mask, min_mask, max_mask, ctrl_mask, low, rev_low, high, rev_high;
        // as in the reference, "min"/"max" masks are those of obj[0]/obj[1] (not sorted);
        // the zero bits are inserted at the smaller, then at the larger qubit
        Mask2(int q0, int q1, index_t ctrl) : ctrl_mask(ctrl) {
            int qmin = std::min(q0, q1), qmax = std::max(q0, q1);
            min_mask = index_t(1) << q0;
            max_mask = index_t(1) << q1;
            obj_mask = min_mask | max_mask;
            low = (index_t(1) << qmin) - 1;
            rev_low = ~low;
            high = (index_t(1) << qmax) - 1;
            rev_high = ~high;
        }
        index_t at(index_t l) const {  // SHIFT_BIT_TWO
            index_t d = ((l & rev_low) << 1) + (l & low);
            return ((d & rev_high) << 1) + (d & high);
        }
    };

    static bool par(index_t dim) { return dim >= kDimTh; }

    // ---- gate kernels (cpu_common/*.cpp) ----
    static void single_matrix(C* qs, int q, index_t ctrl, const C m[2][2], index_t dim) {
        Mask1 mk(q, ctrl);  // cpu_vector_core_matrix_gate.cpp:150-194
#pragma omp parallel for schedule(static) if (par(dim))
        for (int64_t l = 0; l < int64_t(dim / 2); l++) {
            index_t i = mk.at(l);
            if ((i & ctrl) != ctrl) continue;
            index_t j = i + mk.obj_mask;
            C t1 = m[0][0] * qs[i] + m[0][1] * qs[j];
            C t2 = m[1][0] * qs[i] + m[1][1] * qs[j];
            qs[i] = t1;
            qs[j] = t2;
        }
    }
    static void two_matrix(C* qs, int q0, int q1, index_t ctrl, const C m[4][4], index_t dim) {
        Mask2 mk(q0, q1, ctrl);  // cpu_vector_core_matrix_gate.cpp:85-149
        index_t m1 = index_t(1) << q0, m2 = index_t(1) << q1;
#pragma omp parallel for schedule(static) if (par(dim))
        for (int64_t l = 0; l < int64_t(dim / 4); l++) {
            index_t i = mk.at(l);
            if ((i & ctrl) != ctrl) continue;
            index_t j = i + m1, k = i + m2, w = i + mk.obj_mask;
            C a = qs[i], b = qs[j], c = qs[k], d = qs[w];
            qs[i] = m[0][0] * a + m[0][1] * b + m[0][2] * c + m[0][3] * d;
            qs[j] = m[1][0] * a + m[1][1] * b + m[1][2] * c + m[1][3] * d;
            qs[k] = m[2][0] * a + m[2][1] * b + m[2][2] * c + m[2][3] * d;
            qs[w] = m[3][0] * a + m[3][1] * b + m[3][2] * c + m[3][3] * d;
        }
    }
    static void n_matrix(C* qs, const std::vector<int>& objs, index_t ctrl, const std::vector<C>& m, index_t dim) {
        size_t md = size_t(1) << objs.size();  // cpu_vector_core_matrix_gate.cpp:28-83
        std::vector<index_t> om(md);
        for (size_t i = 0; i < md; i++) {
            index_t mk = 0;
            for (size_t b = 0; b < objs.size(); b++)
                if ((i >> b) & 1) mk |= index_t(1) << objs[b];
            om[i] = mk;
        }
        index_t all = om.back();
#pragma omp parallel for schedule(static) if (par(dim))
        for (int64_t l = 0; l < int64_t(dim); l++) {
            if ((index_t(l) & ctrl) != ctrl || (index_t(l) & all) != 0) continue;
            std::vector<C> tmp(md);
            for (size_t r = 0; r < md; r++) {
                C s = 0;
                for (size_t c = 0; c < md; c++) s += m[r * md + c] * qs[om[c] | index_t(l)];
                tmp[r] = s;
            }
            for (size_t r = 0; r < md; r++) qs[om[r] | index_t(l)] = tmp[r];
        }
    }
    static void x_like(C* qs, int q, index_t ctrl, C v1, C v2, index_t dim) {
        Mask1 mk(q, ctrl);  // cpu_vector_core_x_like.cpp:28-57
#pragma omp parallel for schedule(static) if (par(dim))
        for (int64_t l = 0; l < int64_t(dim / 2); l++) {
            index_t i = mk.at(l);
            if ((i & ctrl) != ctrl) continue;
            index_t j = i | mk.obj_mask;
            C t = qs[i];
            qs[i] = qs[j] * v1;
            qs[j] = t * v2;
        }
    }
    static void z_like(C* qs, int q, index_t ctrl, C v, index_t dim) {
        Mask1 mk(q, ctrl);  // cpu_vector_core_z_like.cpp:28-51
#pragma omp parallel for schedule(static) if (par(dim))
        for (int64_t l = 0; l < int64_t(dim / 2); l++) {
            index_t i = mk.at(l) + mk.obj_mask;
            if ((i & ctrl) != ctrl) continue;
            qs[i] *= v;
        }
    }
    static void swap_like(C* qs, int q0, int q1, index_t ctrl, C f, bool plain, index_t dim) {
        Mask2 mk(q0, q1, ctrl);  // cpu_vector_core_swap_like.cpp:31-110
#pragma omp parallel for schedule(static) if (par(dim))
        for (int64_t l = 0; l < int64_t(dim / 4); l++) {
            index_t i = mk.at(l);
            if ((i & ctrl) != ctrl) continue;
            index_t j = i + mk.min_mask, k = i + mk.max_mask;
            C t = qs[j];
            if (plain) {
                qs[j] = qs[k];
                qs[k] = t;
            } else {
                qs[j] = f * qs[k] * C(0, 1);
                qs[k] = f * t * C(0, 1);
            }
        }
    }
    // Rxx / Ryy / Rzz (cpu_vector_core_rot_pauli.cpp:31-60, 322-345, 379-429): c, s of val/2,
    // or their derivatives when diff is set.
    static void rpp(C* qs, int kind, int q0, int q1, index_t ctrl, T val, bool diff, index_t dim) {
        Mask2 mk(q0, q1, ctrl);
        T c = std::cos(val / 2), s = std::sin(val / 2);
        if (diff) {
            c = -std::sin(val / 2) / 2;
            s = std::cos(val / 2) / 2;
        }
#pragma omp parallel for schedule(static) if (par(dim))
        for (int64_t l = 0; l < int64_t(dim / 4); l++) {
            index_t i = mk.at(l);
            if ((i & ctrl) != ctrl) continue;
            index_t w = i + mk.obj_mask, j = i + mk.min_mask, k = i + mk.max_mask;
            if (kind == MQA_GATE_RZZ) {
                C e(c, s), me(c, -s);
                qs[i] *= me;
                qs[j] *= e;
                qs[k] *= e;
                qs[w] *= me;
            } else if (kind == MQA_GATE_RXX) {
                C sc = C(0, -1) * s;
                C v00 = c * qs[i] + sc * qs[w], v01 = c * qs[j] + sc * qs[k];
                C v10 = c * qs[k] + sc * qs[j], v11 = c * qs[w] + sc * qs[i];
                qs[i] = v00, qs[j] = v01, qs[k] = v10, qs[w] = v11;
            } else {  // RYY
                C sc = C(0, 1) * s;
                C v00 = c * qs[i] + sc * qs[w], v01 = c * qs[j] - sc * qs[k];
                C v10 = c * qs[k] - sc * qs[j], v11 = c * qs[w] + sc * qs[i];
                qs[i] = v00, qs[j] = v01, qs[k] = v10, qs[w] = v11;
            }
        }
        if (diff && ctrl) zero_except(qs, ctrl, dim);
    }
    // SetToZeroExcept (cpu_vector_core_policy.cpp:85-97)
    static void zero_except(C* qs, index_t ctrl, index_t dim) {
#pragma omp parallel for schedule(static) if (par(dim))
        for (int64_t i = 0; i < int64_t(dim); i++)
            if ((index_t(i) & ctrl) != ctrl) qs[i] = 0;
    }

    static T angle(const Gate& g, const double* params) {
        double v = g.const_val;  // ParameterResolver::Combination (parameter_resolver.cpp:481-489)
        for (auto& t : g.terms) v += t.second * params[t.first];
        return static_cast<T>(v);
    }

    // ApplyGate dispatch (vector_state.tpp:165-437), diff == false.
    static void apply_gate(C* qs, const Gate& g, const double* params, index_t dim) {
        const index_t ctrl = g.ctrl_mask;
        const int q = g.objs.empty() ? 0 : g.objs[0];
        const T r2 = static_cast<T>(M_SQRT1_2);
        switch (g.kind) {
            case MQA_GATE_I: break;
            case MQA_GATE_X: x_like(qs, q, ctrl, C(1), C(1), dim); break;
            case MQA_GATE_Y: x_like(qs, q, ctrl, C(0, -1), C(0, 1), dim); break;
            case MQA_GATE_Z: z_like(qs, q, ctrl, C(-1), dim); break;
            case MQA_GATE_S: z_like(qs, q, ctrl, C(0, 1), dim); break;
            case MQA_GATE_SDAG: z_like(qs, q, ctrl, C(0, -1), dim); break;
            case MQA_GATE_T: z_like(qs, q, ctrl, C(1, 1) / static_cast<T>(std::sqrt(2.0)), dim); break;
            case MQA_GATE_TDAG: z_like(qs, q, ctrl, C(1, -1) / static_cast<T>(std::sqrt(2.0)), dim); break;
            case MQA_GATE_H: {
                const C m[2][2] = {{r2, r2}, {r2, -r2}};
                single_matrix(qs, q, ctrl, m, dim);
            } break;
            case MQA_GATE_SX:
            case MQA_GATE_SXDAG: {
                C a = g.kind == MQA_GATE_SX ? C(0.5, 0.5) : C(0.5, -0.5), b = std::conj(a);
                const C m[2][2] = {{a, b}, {b, a}};
                single_matrix(qs, q, ctrl, m, dim);
            } break;
            case MQA_GATE_SWAP: swap_like(qs, g.objs[0], g.objs[1], ctrl, C(1), true, dim); break;
            case MQA_GATE_ISWAP: swap_like(qs, g.objs[0], g.objs[1], ctrl, C(1), false, dim); break;
            case MQA_GATE_ISWAPDAG: swap_like(qs, g.objs[0], g.objs[1], ctrl, C(-1), false, dim); break;
            case MQA_GATE_MATRIX: {
                std::vector<C> m(g.mat.size());
                for (size_t i = 0; i < m.size(); i++) m[i] = C(static_cast<T>(g.mat[i].real()), static_cast<T>(g.mat[i].imag()));
                if (g.objs.size() == 1) {
                    const C mm[2][2] = {{m[0], m[1]}, {m[2], m[3]}};
                    single_matrix(qs, q, ctrl, mm, dim);
                } else if (g.objs.size() == 2) {
                    C mm[4][4];
                    for (int r = 0; r < 4; r++)
                        for (int c = 0; c < 4; c++) mm[r][c] = m[r * 4 + c];
                    two_matrix(qs, g.objs[0], g.objs[1], ctrl, mm, dim);
                } else {
                    n_matrix(qs, g.objs, ctrl, m, dim);
                }
            } break;
            case MQA_GATE_RX: {  // rot_pauli.cpp:431-446
                T v = angle(g, params);
                T a = std::cos(v / 2), b = -std::sin(v / 2);
                const C m[2][2] = {{C(a, 0), C(0, b)}, {C(0, b), C(a, 0)}};
                single_matrix(qs, q, ctrl, m, dim);
            } break;
            case MQA_GATE_RY: {  // rot_pauli.cpp:448-463
                T v = angle(g, params);
                T a = std::cos(v / 2), b = std::sin(v / 2);
                const C m[2][2] = {{C(a, 0), C(-b, 0)}, {C(b, 0), C(a, 0)}};
                single_matrix(qs, q, ctrl, m, dim);
            } break;
            case MQA_GATE_RZ: {  // rot_pauli.cpp:525-540
                T v = angle(g, params);
                T a = std::cos(v / 2), b = std::sin(v / 2);
                const C m[2][2] = {{C(a, -b), C(0, 0)}, {C(0, 0), C(a, b)}};
                single_matrix(qs, q, ctrl, m, dim);
            } break;
            case MQA_GATE_PS: {  // z_like.cpp:84-88
                T v = angle(g, params);
                z_like(qs, q, ctrl, C(std::cos(v), std::sin(v)), dim);
            } break;
            case MQA_GATE_GP: {  // other_gate.cpp:53-58
                T v = angle(g, params);
                C c = std::exp(C(0, -v));
                const C m[2][2] = {{c, C(0)}, {C(0), c}};
                single_matrix(qs, q, ctrl, m, dim);
            } break;
            case MQA_GATE_RXX:
            case MQA_GATE_RYY:
            case MQA_GATE_RZZ: rpp(qs, g.kind, g.objs[0], g.objs[1], ctrl, angle(g, params), false, dim); break;
            default: throw std::invalid_argument("Apply of gate " + std::to_string(g.kind) + " not implement.");
        }
    }

    // ---- derivative expectations <bra| dG(val) |ket> (cpu_vector_core_gate_expect.cpp) ----
    static C expect_diff_single(const C* bra, const C* ket, int q, index_t ctrl, const C m[2][2], index_t dim) {
        Mask1 mk(q, ctrl);  // gate_expect.cpp:243-329
        T rr = 0, ri = 0;
#pragma omp parallel for schedule(static) reduction(+ : rr, ri) if (par(dim))
        for (int64_t l = 0; l < int64_t(dim / 2); l++) {
            index_t i = mk.at(l);
            if ((i & ctrl) != ctrl) continue;
            index_t j = i + mk.obj_mask;
            C t1 = m[0][0] * ket[i] + m[0][1] * ket[j];
            C t2 = m[1][0] * ket[i] + m[1][1] * ket[j];
            C r = std::conj(bra[i]) * t1 + std::conj(bra[j]) * t2;
            rr += r.real();
            ri += r.imag();
        }
        return {rr, ri};
    }
    static C expect_diff_rpp(const C* bra, const C* ket, int kind, int q0, int q1, index_t ctrl, T val, index_t dim) {
        Mask2 mk(q0, q1, ctrl);  // gate_expect.cpp:443-500 (Rxx), 798-860 (Ryy), 871-937 (Rzz)
        T c = -std::sin(val / 2) / 2, s = std::cos(val / 2) / 2;
        T rr = 0, ri = 0;
#pragma omp parallel for schedule(static) reduction(+ : rr, ri) if (par(dim))
        for (int64_t l = 0; l < int64_t(dim / 4); l++) {
            index_t i = mk.at(l);
            if ((i & ctrl) != ctrl) continue;
            index_t w = i + mk.obj_mask, j = i + mk.min_mask, k = i + mk.max_mask;
            C r;
            if (kind == MQA_GATE_RZZ) {
                C e(c, s), me(c, -s);
                r = std::conj(bra[i]) * ket[i] * me + std::conj(bra[j]) * ket[j] * e + std::conj(bra[k]) * ket[k] * e +
                    std::conj(bra[w]) * ket[w] * me;
            } else if (kind == MQA_GATE_RXX) {
                C sc = C(0, -1) * s;
                r = std::conj(bra[i]) * (c * ket[i] + sc * ket[w]) + std::conj(bra[j]) * (c * ket[j] + sc * ket[k]) +
                    std::conj(bra[k]) * (c * ket[k] + sc * ket[j]) + std::conj(bra[w]) * (c * ket[w] + sc * ket[i]);
            } else {
                C sc = C(0, 1) * s;
                r = std::conj(bra[i]) * (c * ket[i] + sc * ket[w]) + std::conj(bra[j]) * (c * ket[j] - sc * ket[k]) +
                    std::conj(bra[k]) * (c * ket[k] - sc * ket[j]) + std::conj(bra[w]) * (c * ket[w] + sc * ket[i]);
            }
            rr += r.real();
            ri += r.imag();
        }
        return {rr, ri};
    }
    static C expect_diff_ps(const C* bra, const C* ket, int q, index_t ctrl, T val, index_t dim) {
        Mask1 mk(q, ctrl);  // gate_expect.cpp:388-440
        C e(-std::sin(val), std::cos(val));
        T rr = 0, ri = 0;
#pragma omp parallel for schedule(static) reduction(+ : rr, ri) if (par(dim))
        for (int64_t l = 0; l < int64_t(dim / 2); l++) {
            index_t i = mk.at(l);
            if ((i & ctrl) != ctrl) continue;
            index_t j = i + mk.obj_mask;
            C r = std::conj(bra[j]) * ket[j] * e;
            rr += r.real();
            ri += r.imag();
        }
        return {rr, ri};
    }
    // ExpectDiffGate dispatch (vector_state.tpp:777-851)
    static C expect_diff_gate(const C* bra, const C* ket, const Gate& g, const double* params, index_t dim) {
        T v = angle(g, params);
        int q = g.objs[0];
        switch (g.kind) {
            case MQA_GATE_RX: {  // gate_expect.cpp:345-353
                T c = static_cast<T>(-0.5 * std::sin(v / 2));
                C is = C(0, -1) * static_cast<T>(0.5 * std::cos(v / 2));
                const C m[2][2] = {{c, is}, {is, c}};
                return expect_diff_single(bra, ket, q, g.ctrl_mask, m, dim);
            }
            case MQA_GATE_RY: {  // gate_expect.cpp:355-363
                T c = -0.5 * std::sin(v / 2), s = 0.5 * std::cos(v / 2);
                const C m[2][2] = {{c, -s}, {s, c}};
                return expect_diff_single(bra, ket, q, g.ctrl_mask, m, dim);
            }
            case MQA_GATE_RZ: {  // gate_expect.cpp:365-375
                T c = -0.5 * std::sin(v / 2), s = 0.5 * std::cos(v / 2);
                const C m[2][2] = {{C(c, -s), C(0)}, {C(0), C(c, s)}};
                return expect_diff_single(bra, ket, q, g.ctrl_mask, m, dim);
            }
            case MQA_GATE_GP: {  // gate_expect.cpp:378-385
                C e = C(0, -1) * std::exp(C(0, -v));
                const C m[2][2] = {{e, C(0)}, {C(0), e}};
                return expect_diff_single(bra, ket, q, g.ctrl_mask, m, dim);
            }
            case MQA_GATE_PS: return expect_diff_ps(bra, ket, q, g.ctrl_mask, v, dim);
            case MQA_GATE_RXX:
            case MQA_GATE_RYY:
            case MQA_GATE_RZZ: return expect_diff_rpp(bra, ket, g.kind, g.objs[0], g.objs[1], g.ctrl_mask, v, dim);
            default: throw std::invalid_argument("Expectation of gate " + std::to_string(g.kind) + " not implement.");
        }
    }

    // ---- Pauli terms (cpu_vector_core_policy.cpp:132-161 ApplyTerms, :211-261 ExpectationOfTerms) ----
    static C polar(const mqa_pauli_term& t, index_t i) {
        static const std::complex<double> P[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};  // mq_base/utils.cc:22
        int ny = __builtin_popcountll(t.y_mask);
        int k = (ny + 2 * __builtin_popcountll(i & t.y_mask) + 2 * __builtin_popcountll(i & t.z_mask)) & 3;
        return C(static_cast<T>(P[k].real()), static_cast<T>(P[k].imag()));
    }
    static void apply_terms(const C* qs, C* out, const mqa_pauli_term* terms, int n_terms, index_t dim) {
        std::fill(out, out + dim, C(0));
        for (int t = 0; t < n_terms; t++) {
            const auto& tm = terms[t];
            index_t f = tm.x_mask | tm.y_mask;
            T coeff = static_cast<T>(tm.coeff);
#pragma omp parallel for schedule(static) if (par(dim))
            for (int64_t ii = 0; ii < int64_t(dim); ii++) {
                index_t i = ii, j = i ^ f;
                if (i > j) continue;
                C c = polar(tm, i);
                out[j] += qs[i] * coeff * c;
                if (i != j) out[i] += qs[j] * coeff / c;
            }
        }
    }
    static C expectation(const C* bra, const C* ket, const mqa_pauli_term* terms, int n_terms, index_t dim) {
        std::complex<double> out = 0;
        for (int t = 0; t < n_terms; t++) {
            const auto& tm = terms[t];
            index_t f = tm.x_mask | tm.y_mask;
            T coeff = static_cast<T>(tm.coeff);
            T rr = 0, ri = 0;
#pragma omp parallel for schedule(static) reduction(+ : rr, ri) if (par(dim))
            for (int64_t ii = 0; ii < int64_t(dim); ii++) {
                index_t i = ii, j = i ^ f;
                if (i > j) continue;
                C c = polar(tm, i);
                C r = std::conj(bra[j]) * ket[i] * coeff * c;
                if (i != j) r += std::conj(bra[i]) * ket[j] * coeff / c;
                rr += r.real();
                ri += r.imag();
            }
            out += std::complex<double>(rr, ri);
        }
        return C(static_cast<T>(out.real()), static_cast<T>(out.imag()));
    }
    // Vdot (cpu_vector_core_dot_like.cpp:32-52)
    static C vdot(const C* bra, const C* ket, index_t dim) {
        T rr = 0, ri = 0;
#pragma omp parallel for schedule(static) reduction(+ : rr, ri) if (par(dim))
        for (int64_t i = 0; i < int64_t(dim); i++) {
            rr += bra[i].real() * ket[i].real() + bra[i].imag() * ket[i].imag();
            ri += bra[i].real() * ket[i].imag() - bra[i].imag() * ket[i].real();
        }
        return {rr, ri};
    }

    // ---- hermitian adjoint gradient (gradient.h:100-121, 160-197, 124-133) ----
    // out: [batch][n_hams][1 + n_params] complex (double pairs).
    static void grad_batch(int n, const C* init, const std::vector<Gate>& circ, const mqa_pauli_term* terms,
                           const int32_t* ham_off, int n_hams, const double* enc, int64_t batch, int n_enc,
                           const double* ans, int n_ans, int batch_threads, double* out) {
        const index_t dim = index_t(1) << n;
        const int n_params = n_enc + n_ans;
        std::vector<Gate> herm;  // circuit.py:831-845: reversed order, each gate hermitian-ed
        for (auto it = circ.rbegin(); it != circ.rend(); ++it) herm.push_back(hermitian_gate(*it));
        auto run = [&](int64_t idx) {
            // pr.SetItems(enc_name, enc_data[idx]); pr.SetItems(ans_name, ans_data)
            std::vector<double> pr(n_params);
            for (int p = 0; p < n_enc; p++) pr[p] = static_cast<T>(enc[idx * n_enc + p]);
            for (int p = 0; p < n_ans; p++) pr[n_enc + p] = static_cast<T>(ans[p]);
            std::vector<C> psi_l(dim);
            if (init) {
                std::copy(init, init + dim, psi_l.begin());
            } else {
                psi_l[0] = 1;
            }
            for (auto& g : circ) apply_gate(psi_l.data(), g, pr.data(), dim);  // ApplyCircuit
            // LeftSizeGrad: psi_r[i] = H_i psi, f = <psi_l|psi_r>
            std::vector<std::vector<C>> psi_r(n_hams, std::vector<C>(dim));
            std::vector<std::vector<std::complex<double>>> fg(n_hams, std::vector<std::complex<double>>(1 + n_params));
            for (int h = 0; h < n_hams; h++) {
                apply_terms(psi_l.data(), psi_r[h].data(), terms + ham_off[h], ham_off[h + 1] - ham_off[h], dim);
                fg[h][0] = vdot(psi_l.data(), psi_r[h].data(), dim);
            }
            for (auto& g : herm) {
                apply_gate(psi_l.data(), g, pr.data(), dim);
                if (g.requires_grad && is_param_kind(g.kind) && !g.terms.empty()) {
                    for (int h = 0; h < n_hams; h++) {
                        C ig = expect_diff_gate(psi_l.data(), psi_r[h].data(), g, pr.data(), dim);
                        // MatMul(intrin_grad[1x1], jac[1 x n_names]): jacobian = PR coefficients
                        for (auto& t : g.terms)
                            fg[h][1 + t.first] += std::complex<double>(ig.real(), ig.imag()) * t.second;
                    }
                }
                for (int h = 0; h < n_hams; h++) apply_gate(psi_r[h].data(), g, pr.data(), dim);
            }
            for (int h = 0; h < n_hams; h++) {
                double* o = out + ((idx * n_hams + h) * (1 + n_params)) * 2;
                // values pass through calc_type like the reference's VT<py_qs_data_t> of complex<T>
                o[0] = static_cast<T>(fg[h][0].real());
                o[1] = static_cast<T>(fg[h][0].imag());
                for (int p = 0; p < n_params; p++) {
                    std::complex<T> v(static_cast<T>(fg[h][1 + p].real()), static_cast<T>(fg[h][1 + p].imag()));
                    v += std::conj(v);  // GradRealDouble
                    o[2 * (1 + p)] = v.real();
                    o[2 * (1 + p) + 1] = v.imag();
                }
            }
        };
        // ParallelTask (gradient.h:270-299)
        if (batch == 1) {
            run(0);
            return;
        }
        if (batch_threads <= 0) throw std::runtime_error("batch_threads cannot be zero.");
        std::vector<std::thread> tasks;
        int64_t end = 0, offset = batch / batch_threads, left = batch % batch_threads;
        for (int64_t i = 0; i < batch_threads; ++i) {
            int64_t start = end;
            end = start + offset + (i < left ? 1 : 0);
            tasks.emplace_back([&, start, end]() {
                if (g_worker_threads > 0) omp_set_num_threads(g_worker_threads);
                for (int64_t b = start; b < end; b++) run(b);
            });
        }
        for (auto& t : tasks) t.join();
    }
};

static thread_local std::string g_err;

static std::vector<Gate> to_gates(const mqa_gate* gates, int n_gates, const int32_t* tp, const double* tc,
                                  const double* mats) {
    std::vector<Gate> out;
    for (int i = 0; i < n_gates; i++) {
        const mqa_gate& s = gates[i];
        Gate g;
        g.kind = s.kind;
        g.objs.assign(s.obj, s.obj + s.n_obj);
        g.ctrl_mask = s.ctrl_mask;
        g.const_val = s.const_val;
        for (int t = 0; t < s.n_term; t++) g.terms.push_back({tp[s.term_begin + t], tc[s.term_begin + t]});
        if (s.kind == MQA_GATE_MATRIX) {
            size_t m = size_t(1) << s.n_obj;
            for (size_t e = 0; e < m * m; e++)
                g.mat.push_back({mats[2 * (s.mat_begin + e)], mats[2 * (s.mat_begin + e) + 1]});
        }
        g.requires_grad = s.requires_grad != 0;
        out.push_back(std::move(g));
    }
    return out;
}

template <typename F>
static int guarded(F&& f) {
    try {
        f();
        return MQA_OK;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return MQA_ERR_ARG;
    } catch (const std::exception& e) {
        g_err = e.what();
        return MQA_ERR_RUNTIME;
    }
}

}  // namespace oracle

extern "C" {

const char* oracle_last_error(void) { return oracle::g_err.c_str(); }

int oracle_num_threads(void) { return omp_get_max_threads(); }

void oracle_set_num_threads(int n) { omp_set_num_threads(n); }

void oracle_set_worker_threads(int n) { oracle::g_worker_threads = n < 0 ? 0 : n; }

// qs: interleaved complex (float pairs for c64, double pairs for c128), length 2^n, in place.
int oracle_apply_circuit(int n, int dtype, void* qs, const mqa_gate* gates, int n_gates, const int32_t* term_param,
                         const double* term_coeff, const double* mats, const double* params) {
    return oracle::guarded([&] {
        auto circ = oracle::to_gates(gates, n_gates, term_param, term_coeff, mats);
        uint64_t dim = uint64_t(1) << n;
        if (dtype == MQA_C64) {
            using S = oracle::Sim<float>;
            auto* q = reinterpret_cast<S::C*>(qs);
            for (auto& g : circ) S::apply_gate(q, g, params, dim);
        } else {
            using S = oracle::Sim<double>;
            auto* q = reinterpret_cast<S::C*>(qs);
            for (auto& g : circ) S::apply_gate(q, g, params, dim);
        }
    });
}

int oracle_apply_hamiltonian(int n, int dtype, const void* qs, void* out, const mqa_pauli_term* terms, int n_terms) {
    return oracle::guarded([&] {
        uint64_t dim = uint64_t(1) << n;
        if (dtype == MQA_C64) {
            using S = oracle::Sim<float>;
            S::apply_terms(reinterpret_cast<const S::C*>(qs), reinterpret_cast<S::C*>(out), terms, n_terms, dim);
        } else {
            using S = oracle::Sim<double>;
            S::apply_terms(reinterpret_cast<const S::C*>(qs), reinterpret_cast<S::C*>(out), terms, n_terms, dim);
        }
    });
}

int oracle_expectation(int n, int dtype, const void* bra, const void* ket, const mqa_pauli_term* terms, int n_terms,
                       double* out) {
    return oracle::guarded([&] {
        uint64_t dim = uint64_t(1) << n;
        if (dtype == MQA_C64) {
            using S = oracle::Sim<float>;
            auto r = S::expectation(reinterpret_cast<const S::C*>(bra), reinterpret_cast<const S::C*>(ket), terms,
                                    n_terms, dim);
            out[0] = r.real(), out[1] = r.imag();
        } else {
            using S = oracle::Sim<double>;
            auto r = S::expectation(reinterpret_cast<const S::C*>(bra), reinterpret_cast<const S::C*>(ket), terms,
                                    n_terms, dim);
            out[0] = r.real(), out[1] = r.imag();
        }
    });
}

int oracle_vdot(int n, int dtype, const void* bra, const void* ket, double* out) {
    return oracle::guarded([&] {
        uint64_t dim = uint64_t(1) << n;
        if (dtype == MQA_C64) {
            using S = oracle::Sim<float>;
            auto r = S::vdot(reinterpret_cast<const S::C*>(bra), reinterpret_cast<const S::C*>(ket), dim);
            out[0] = r.real(), out[1] = r.imag();
        } else {
            using S = oracle::Sim<double>;
            auto r = S::vdot(reinterpret_cast<const S::C*>(bra), reinterpret_cast<const S::C*>(ket), dim);
            out[0] = r.real(), out[1] = r.imag();
        }
    });
}

// Hermitian adjoint gradient over a batch.  init may be NULL (|0...0>).
// ham_off has n_hams + 1 entries delimiting each Hamiltonian's terms.
// out: [batch][n_hams][1 + n_enc + n_ans] complex as double pairs.
int oracle_grad_batch(int n, int dtype, const void* init, const mqa_gate* gates, int n_gates, const int32_t* term_param,
                      const double* term_coeff, const double* mats, const mqa_pauli_term* terms, const int32_t* ham_off,
                      int n_hams, const double* enc, int64_t batch, int n_enc, const double* ans, int n_ans,
                      int batch_threads, double* out) {
    return oracle::guarded([&] {
        auto circ = oracle::to_gates(gates, n_gates, term_param, term_coeff, mats);
        if (dtype == MQA_C64) {
            using S = oracle::Sim<float>;
            S::grad_batch(n, reinterpret_cast<const S::C*>(init), circ, terms, ham_off, n_hams, enc, batch, n_enc, ans,
                          n_ans, batch_threads, out);
        } else {
            using S = oracle::Sim<double>;
            S::grad_batch(n, reinterpret_cast<const S::C*>(init), circ, terms, ham_off, n_hams, enc, batch, n_enc, ans,
                          n_ans, batch_threads, out);
        }
    });
}

}  // extern "C"
