// kernels.hip - gfx950 kernels of the state-vector backend.
//
//  k_tile_pass<T, ADJ>  one cache-blocked pass (program.h): HBM -> LDS tile -> register
//                       groups -> LDS -> HBM.  ADJ: the fused adjoint back-sweep pass, which
//                       streams the forward state psi and the co-state lambda together, emits
//                       the derivative tap <lambda_k| dG_k |psi_{k-1}> of every parameterised
//                       op and applies G_k^dagger to both (reference: gradient.h:179-194 does
//                       this as three separate full-state sweeps per gate).
//  k_ham<T, WRITE>      Pauli-term Hamiltonian gathered by flip mask: out = H psi and/or the
//                       partial sums of <bra|H psi> (reference: cpu_vector_core_policy.cpp:
//                       132-161 ApplyTerms, 211-261 ExpectationOfTerms, one sweep per term).
//  k_vdot<T>            partial sums of <a|b> (dot_like.cpp:32-52).
//  k_fill<T>            broadcast an initial state (or |0>) into a batch of states.
//  k_reduce_rows(1)     deterministic double-precision row sums of (complex / real) partials.
//  k_jacobian           taps -> parameter gradients, g = sum coeff * 2 Re(tap) (gradient.h:
//                       183-188 MatMul with the PR jacobian, :124-133 GradRealDouble).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <utility>

#include "launch.h"
#include "program.h"

namespace mqa {

template <typename T>
struct V2t;
template <>
struct V2t<float> {
    using type = float2;
};
template <>
struct V2t<double> {
    using type = double2;
};
template <typename T>
using V2 = typename V2t<T>::type;

// ---------------------------------------------------------------- small helpers
__device__ __forceinline__ uint64_t pdep64(uint64_t src, uint64_t mask) {
    uint64_t out = 0;
    while (mask) {
        uint64_t low = mask & (~mask + 1);
        if (src & 1) out |= low;
        src >>= 1;
        mask &= mask - 1;
    }
    return out;
}

// LDS address swizzle: XOR-folds higher index bits into the low bits so that register
// groups whose positions include low qubits do not hit one bank with every lane.
template <typename T>
__device__ __forceinline__ uint32_t swz(uint32_t u) {
    return lds_swizzle(u, int(sizeof(T) * 2));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v, int width) {
    for (int off = width >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Wave-wide sum of a full 64-lane wave: DPP within rows of 16 (VALU, no LDS round trip),
// then two cross-row exchanges.  Every lane ends with the total.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(b & 0xffffffffll), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
template <typename T>
__device__ __forceinline__ T wave_sum64(T v) {
    v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_mov<0x141>(v);  // row_half_mirror
    v += dpp_mov<0x140>(v);  // row_mirror
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// Per-thread derivative-tap accumulators of one register group.  Only Re(tap) is kept: every
// parameter gradient is sum coeff * 2 Re(tap) with real coefficients.  The scheduler caps the
// taps of a group at 8 so that the accumulators stay in registers.
template <typename T>
constexpr int kTapSlots = 8;
template <typename T>
struct TapAcc {
    T r[kTapSlots<T>];
};

// Register state of one thread inside a group: 16 amplitudes of psi (and of lambda).
template <typename T>
struct Regs {
    T r[kAmpsPerThread];
    T i[kAmpsPerThread];
};

// Context shared by the op bodies.
struct GroupCtx {
    uint32_t ub;     // local index of register 0 (group positions cleared)
    uint32_t pm[4];  // 1 << pos[i]
};

template <int V>
__device__ __forceinline__ uint32_t pv(const GroupCtx& g) {
    return ((V & 1) ? g.pm[0] : 0u) | ((V & 2) ? g.pm[1] : 0u) | ((V & 4) ? g.pm[2] : 0u) | ((V & 8) ? g.pm[3] : 0u);
}

template <int V>
__device__ __forceinline__ bool ctrl_ok(const GroupCtx& g, uint32_t cl) {
    return ((g.ub | pv<V>(g)) & cl) == cl;
}

// ---------------------------------------------------------------- op bodies
// Compile-time loops over the 8 pairs (one target) or 4 quads (two targets) of a register set.
template <int Q, int V = 0, typename F>
__device__ __forceinline__ void for_pairs(F&& f) {
    if constexpr (V < kAmpsPerThread) {
        if constexpr ((V & (1 << Q)) == 0) f.template operator()<V, (V | (1 << Q))>();
        for_pairs<Q, V + 1>(f);
    }
}
template <int Q0, int Q1, int V = 0, typename F>
__device__ __forceinline__ void for_quads(F&& f) {
    if constexpr (V < kAmpsPerThread) {
        if constexpr ((V & ((1 << Q0) | (1 << Q1))) == 0)
            f.template operator()<V, (V | (1 << Q0)), (V | (1 << Q1)), (V | (1 << Q0) | (1 << Q1))>();
        for_quads<Q0, Q1, V + 1>(f);
    }
}

// a' = m a on one pair, m a row-major 2x2 complex matrix (uniform: lives in SGPRs)
template <typename T>
__device__ __forceinline__ void mat2(const T* __restrict__ m, T& ar, T& ai, T& br, T& bi) {
    T xr = m[0] * ar - m[1] * ai + m[2] * br - m[3] * bi;
    T xi = m[0] * ai + m[1] * ar + m[2] * bi + m[3] * br;
    T yr = m[4] * ar - m[5] * ai + m[6] * br - m[7] * bi;
    T yi = m[4] * ai + m[5] * ar + m[6] * bi + m[7] * br;
    ar = xr, ai = xi, br = yr, bi = yi;
}

// conj(x) * y accumulated as four FMAs (a dependent chain costs nothing extra: a wave64
// VALU op occupies its SIMD for four cycles, which covers the FMA latency)
template <typename T>
__device__ __forceinline__ void cdot_acc(T xr, T xi, T yr, T yi, T& sr, T& si) {
    sr = fma(xr, yr, sr);
    sr = fma(xi, yi, sr);
    si = fma(xr, yi, si);
    si = fma(-xi, yr, si);
}

// One-target op on a register set.  c, s: cos / sin of the half angle (RX, RY, RZ) or of the
// angle (PS, GP); m: dense 2x2 (M1) in the matrix pool.  CTRL: some control qubits are local;
// their predicate is evaluated per pair and applied with selects (no divergent branches).
template <typename T, int KIND, int Q, bool CTRL>
__device__ __forceinline__ void one_target(Regs<T>& a, const GroupCtx& g, uint32_t cl, T c, T s,
                                           const T* __restrict__ m) {
    for_pairs<Q>([&]<int V, int W>() {
        T ar = a.r[V], ai = a.i[V], br = a.r[W], bi = a.i[W];
        T xr, xi, yr, yi;
        if constexpr (KIND == OPK_X) {
            xr = br, xi = bi, yr = ar, yi = ai;
        } else if constexpr (KIND == OPK_RX) {  // [[c, -is], [-is, c]]
            xr = c * ar + s * bi, xi = c * ai - s * br;
            yr = c * br + s * ai, yi = c * bi - s * ar;
        } else if constexpr (KIND == OPK_RY) {  // [[c, -s], [s, c]]
            xr = c * ar - s * br, xi = c * ai - s * bi;
            yr = s * ar + c * br, yi = s * ai + c * bi;
        } else if constexpr (KIND == OPK_RZ) {  // diag(c - is, c + is)
            xr = c * ar + s * ai, xi = c * ai - s * ar;
            yr = c * br - s * bi, yi = c * bi + s * br;
        } else if constexpr (KIND == OPK_PS) {  // diag(1, c + is)
            xr = ar, xi = ai;
            yr = c * br - s * bi, yi = c * bi + s * br;
        } else if constexpr (KIND == OPK_GP) {  // (c - is) I
            xr = c * ar + s * ai, xi = c * ai - s * ar;
            yr = c * br + s * bi, yi = c * bi - s * br;
        } else {  // dense 2x2
            xr = ar, xi = ai, yr = br, yi = bi;
            mat2(m, xr, xi, yr, yi);
        }
        if constexpr (CTRL) {
            const bool ok = ctrl_ok<V>(g, cl);
            xr = ok ? xr : ar, xi = ok ? xi : ai, yr = ok ? yr : br, yi = ok ? yi : bi;
        }
        a.r[V] = xr, a.i[V] = xi, a.r[W] = yr, a.i[W] = yi;
    });
}

// Controlled X / Z whose (first) control is register position QC of the group: the control
// bit is known per register at compile time.  cl: further local controls (predicated).
template <typename T, bool ISX, int QT, int QC, bool CTRL>
__device__ __forceinline__ void controlled_xz(Regs<T>& a, const GroupCtx& g, uint32_t cl) {
    for_pairs<QT>([&]<int V, int W>() {
        if constexpr (((V >> QC) & 1) != 0) {
            T ar = a.r[V], ai = a.i[V], br = a.r[W], bi = a.i[W];
            T xr, xi, yr, yi;
            if constexpr (ISX) {
                xr = br, xi = bi, yr = ar, yi = ai;
            } else {
                xr = ar, xi = ai, yr = -br, yi = -bi;
            }
            if constexpr (CTRL) {
                const bool ok = ctrl_ok<V>(g, cl);
                xr = ok ? xr : ar, xi = ok ? xi : ai, yr = ok ? yr : br, yi = ok ? yi : bi;
            }
            a.r[V] = xr, a.i[V] = xi, a.r[W] = yr, a.i[W] = yi;
        }
    });
}

// Derivative tap of a one-target parameterised op, taken before G^dagger is applied:
// accumulates the raw sum s such that tap = factor * s (factor applied by the caller).
//   RX: s = <l|X|p>, RY: s = <l|Y|p>/i, RZ: s = <l|Z|p> (tap = -i/2 * <l|P|p>)
//   PS: s = sum conj(l_b) p_b (tap = i s);  GP: s = <l|p> (tap = -i s)
template <typename T, int KIND, int Q, bool CTRL>
__device__ __forceinline__ void one_target_tap(const Regs<T>& p, const Regs<T>& l, const GroupCtx& g, uint32_t cl,
                                               T& sr, T& si) {
    for_pairs<Q>([&]<int V, int W>() {
        T ur = 0, ui = 0;
        T& xr = CTRL ? ur : sr;
        T& xi = CTRL ? ui : si;
        if constexpr (KIND == OPK_RX) {
            cdot_acc(l.r[V], l.i[V], p.r[W], p.i[W], xr, xi);
            cdot_acc(l.r[W], l.i[W], p.r[V], p.i[V], xr, xi);
        } else if constexpr (KIND == OPK_RY) {  // <l|Y|p> = i (conj(l_b) p_a - conj(l_a) p_b)
            cdot_acc(l.r[W], l.i[W], p.r[V], p.i[V], xr, xi);
            cdot_acc(l.r[V], l.i[V], -p.r[W], -p.i[W], xr, xi);
        } else if constexpr (KIND == OPK_RZ) {
            cdot_acc(l.r[V], l.i[V], p.r[V], p.i[V], xr, xi);
            cdot_acc(l.r[W], l.i[W], -p.r[W], -p.i[W], xr, xi);
        } else if constexpr (KIND == OPK_PS) {
            cdot_acc(l.r[W], l.i[W], p.r[W], p.i[W], xr, xi);
        } else {  // GP
            cdot_acc(l.r[V], l.i[V], p.r[V], p.i[V], xr, xi);
            cdot_acc(l.r[W], l.i[W], p.r[W], p.i[W], xr, xi);
        }
        if constexpr (CTRL) {
            const bool ok = ctrl_ok<V>(g, cl);
            sr += ok ? ur : T(0), si += ok ? ui : T(0);
        }
    });
}

// Two-target ops; quad order (00, 01, 10, 11) with bit 0 <=> Q0 (matrix bit 0).
template <typename T, int KIND, int Q0, int Q1, bool CTRL>
__device__ __forceinline__ void two_target(Regs<T>& a, const GroupCtx& g, uint32_t cl, T c, T s,
                                           const T* __restrict__ m) {
    for_quads<Q0, Q1>([&]<int V0, int V1, int V2_, int V3>() {
        if constexpr (CTRL) {
            if (!ctrl_ok<V0>(g, cl)) return;
        }
        if constexpr (KIND == OPK_SWAP) {
            T tr = a.r[V1], ti = a.i[V1];
            a.r[V1] = a.r[V2_], a.i[V1] = a.i[V2_];
            a.r[V2_] = tr, a.i[V2_] = ti;
        } else if constexpr (KIND == OPK_RZZ) {  // diag(e-, e+, e+, e-), e-+ = c -+ is
            T r0 = c * a.r[V0] + s * a.i[V0], i0 = c * a.i[V0] - s * a.r[V0];
            T r1 = c * a.r[V1] - s * a.i[V1], i1 = c * a.i[V1] + s * a.r[V1];
            T r2 = c * a.r[V2_] - s * a.i[V2_], i2 = c * a.i[V2_] + s * a.r[V2_];
            T r3 = c * a.r[V3] + s * a.i[V3], i3 = c * a.i[V3] - s * a.r[V3];
            a.r[V0] = r0, a.i[V0] = i0, a.r[V1] = r1, a.i[V1] = i1;
            a.r[V2_] = r2, a.i[V2_] = i2, a.r[V3] = r3, a.i[V3] = i3;
        } else if constexpr (KIND == OPK_RXX || KIND == OPK_RYY) {
            // RXX = c I - is XX; RYY = c I - is YY with YY|00> = -|11>, YY|01> = |10>
            constexpr int sg = (KIND == OPK_RXX) ? 1 : -1;
            T ss = s * T(sg);
            T r0 = c * a.r[V0] + ss * a.i[V3], i0 = c * a.i[V0] - ss * a.r[V3];
            T r3 = c * a.r[V3] + ss * a.i[V0], i3 = c * a.i[V3] - ss * a.r[V0];
            T r1 = c * a.r[V1] + s * a.i[V2_], i1 = c * a.i[V1] - s * a.r[V2_];
            T r2 = c * a.r[V2_] + s * a.i[V1], i2 = c * a.i[V2_] - s * a.r[V1];
            a.r[V0] = r0, a.i[V0] = i0, a.r[V1] = r1, a.i[V1] = i1;
            a.r[V2_] = r2, a.i[V2_] = i2, a.r[V3] = r3, a.i[V3] = i3;
        } else {  // dense 4x4
            T xr[4] = {a.r[V0], a.r[V1], a.r[V2_], a.r[V3]};
            T xi[4] = {a.i[V0], a.i[V1], a.i[V2_], a.i[V3]};
            T yr[4], yi[4];
#pragma unroll
            for (int row = 0; row < 4; row++) {
                T ur = 0, ui = 0;
#pragma unroll
                for (int col = 0; col < 4; col++) {
                    const T mr = m[2 * (row * 4 + col)], mi = m[2 * (row * 4 + col) + 1];
                    ur += mr * xr[col] - mi * xi[col];
                    ui += mr * xi[col] + mi * xr[col];
                }
                yr[row] = ur, yi[row] = ui;
            }
            a.r[V0] = yr[0], a.i[V0] = yi[0], a.r[V1] = yr[1], a.i[V1] = yi[1];
            a.r[V2_] = yr[2], a.i[V2_] = yi[2], a.r[V3] = yr[3], a.i[V3] = yi[3];
        }
    });
}

// raw sum s with tap = -i/2 * s, s = <l|PP|p> for PP = XX, YY, ZZ
template <typename T, int KIND, int Q0, int Q1, bool CTRL>
__device__ __forceinline__ void two_target_tap(const Regs<T>& p, const Regs<T>& l, const GroupCtx& g, uint32_t cl,
                                               T& sr, T& si) {
    for_quads<Q0, Q1>([&]<int V0, int V1, int V2_, int V3>() {
        if constexpr (CTRL) {
            if (!ctrl_ok<V0>(g, cl)) return;
        }
        if constexpr (KIND == OPK_RZZ) {
            cdot_acc(l.r[V0], l.i[V0], p.r[V0], p.i[V0], sr, si);
            cdot_acc(l.r[V1], l.i[V1], -p.r[V1], -p.i[V1], sr, si);
            cdot_acc(l.r[V2_], l.i[V2_], -p.r[V2_], -p.i[V2_], sr, si);
            cdot_acc(l.r[V3], l.i[V3], p.r[V3], p.i[V3], sr, si);
        } else {
            constexpr int sg = (KIND == OPK_RXX) ? 1 : -1;
            cdot_acc(l.r[V0], l.i[V0], T(sg) * p.r[V3], T(sg) * p.i[V3], sr, si);
            cdot_acc(l.r[V3], l.i[V3], T(sg) * p.r[V0], T(sg) * p.i[V0], sr, si);
            cdot_acc(l.r[V1], l.i[V1], p.r[V2_], p.i[V2_], sr, si);
            cdot_acc(l.r[V2_], l.i[V2_], p.r[V1], p.i[V1], sr, si);
        }
    });
}

// Dense 2^M x 2^M (M = 3, 4) over register bits 0..M-1 (group positions in the gate's order).
// Such a gate forms a group of its own.  The inputs are re-read from the LDS tile (which
// still holds them) inside a rolled column loop, so no second register set is needed.
template <typename T, int M>
__device__ __forceinline__ void dense_from_lds(Regs<T>& a, const GroupCtx& g, uint32_t cl, const T* __restrict__ mat,
                                               const V2<T>* __restrict__ tile) {
    constexpr int D = 1 << M;
#pragma unroll
    for (int hi = 0; hi < kAmpsPerThread / D; hi++) {
        // controls are never target positions; the pad bit (M == 3) is register bit 3
        const uint32_t u0 = g.ub | ((M == 3 && hi) ? g.pm[3] : 0u);
        if (cl && (u0 & cl) != cl) continue;
#pragma unroll
        for (int row = 0; row < D; row++) {
            T yr = 0, yi = 0;
#pragma unroll 1
            for (int col = 0; col < D; col++) {
                const uint32_t u = u0 | ((col & 1) ? g.pm[0] : 0u) | ((col & 2) ? g.pm[1] : 0u) |
                                   ((col & 4) ? g.pm[2] : 0u) | ((M == 4 && (col & 8)) ? g.pm[3] : 0u);
                const V2<T> x = tile[swz<T>(u)];
                const T mr = mat[2 * (row * D + col)], mi = mat[2 * (row * D + col) + 1];
                yr += mr * x.x - mi * x.y;
                yi += mr * x.y + mi * x.x;
            }
            a.r[hi * D + row] = yr, a.i[hi * D + row] = yi;
        }
    }
}

#define MQA_SW1(CALL)                 \
    switch (op.q0) {                  \
        case 0: CALL(0); break;       \
        case 1: CALL(1); break;       \
        case 2: CALL(2); break;       \
        default: CALL(3); break;      \
    }
#define MQA_SW2(CALL)                        \
    switch (op.q0 * 4 + op.q1) {             \
        case 1: CALL(0, 1); break;           \
        case 2: CALL(0, 2); break;           \
        case 3: CALL(0, 3); break;           \
        case 4: CALL(1, 0); break;           \
        case 6: CALL(1, 2); break;           \
        case 7: CALL(1, 3); break;           \
        case 8: CALL(2, 0); break;           \
        case 9: CALL(2, 1); break;           \
        case 11: CALL(2, 3); break;          \
        case 12: CALL(3, 0); break;          \
        case 13: CALL(3, 1); break;          \
        default: CALL(3, 2); break;          \
    }

// Apply one op (G forward, G^dagger in the back-sweep) to a register set.  FULL adds the dense
// 4x4 kind; passes without it run a leaner instantiation (fewer live registers).
template <typename T, bool FULL>
__device__ __forceinline__ void apply_op(int kind, const DevOp& op, Regs<T>& a, const GroupCtx& g, uint32_t cl, T c,
                                         T s, const T* __restrict__ m) {
    // lean passes never carry local control predicates (the host picks FULL when they do)
#define ONE(KD, Q)                                                                       \
    ((FULL && cl) ? one_target<T, KD, Q, FULL>(a, g, cl, c, s, m) : one_target<T, KD, Q, false>(a, g, cl, c, s, m))
#define TWO(KD, Q0, Q1)                                                                        \
    ((FULL && cl) ? two_target<T, KD, Q0, Q1, FULL>(a, g, cl, c, s, m) : two_target<T, KD, Q0, Q1, false>(a, g, cl, c, s, m))
#define CXZ(ISX, Q0, Q1)                                                                     \
    ((FULL && cl) ? controlled_xz<T, ISX, Q0, Q1, FULL>(a, g, cl) : controlled_xz<T, ISX, Q0, Q1, false>(a, g, cl))
#define X1(Q) ONE(OPK_X, Q)
#define RX1(Q) ONE(OPK_RX, Q)
#define RY1(Q) ONE(OPK_RY, Q)
#define RZ1(Q) ONE(OPK_RZ, Q)
#define PS1(Q) ONE(OPK_PS, Q)
#define GP1(Q) ONE(OPK_GP, Q)
#define M11(Q) ONE(OPK_M1, Q)
#define SW2(Q0, Q1) TWO(OPK_SWAP, Q0, Q1)
#define ZZ2(Q0, Q1) TWO(OPK_RZZ, Q0, Q1)
#define XX2(Q0, Q1) TWO(OPK_RXX, Q0, Q1)
#define YY2(Q0, Q1) TWO(OPK_RYY, Q0, Q1)
#define M22(Q0, Q1) TWO(OPK_M2, Q0, Q1)
#define CX2(Q0, Q1) CXZ(true, Q0, Q1)
#define CZ2(Q0, Q1) CXZ(false, Q0, Q1)
    switch (kind) {
        case OPK_X: MQA_SW1(X1) break;
        case OPK_RX: MQA_SW1(RX1) break;
        case OPK_RY: MQA_SW1(RY1) break;
        case OPK_RZ: MQA_SW1(RZ1) break;
        case OPK_PS: MQA_SW1(PS1) break;
        case OPK_GP: MQA_SW1(GP1) break;
        case OPK_M1: MQA_SW1(M11) break;
        case OPK_CX: MQA_SW2(CX2) break;
        case OPK_CZ: MQA_SW2(CZ2) break;
        case OPK_SWAP: MQA_SW2(SW2) break;
        case OPK_RZZ: MQA_SW2(ZZ2) break;
        case OPK_RXX: MQA_SW2(XX2) break;
        case OPK_RYY: MQA_SW2(YY2) break;
        case OPK_M2:
            if constexpr (FULL) MQA_SW2(M22)
            break;
        default: break;
    }
#undef ONE
#undef TWO
#undef CXZ
#undef X1
#undef RX1
#undef RY1
#undef RZ1
#undef PS1
#undef GP1
#undef M11
#undef SW2
#undef ZZ2
#undef XX2
#undef YY2
#undef M22
#undef CX2
#undef CZ2
}

// One op of a group.  ADJ: accumulate the derivative tap, then apply G^dagger to psi and lambda.
template <typename T, bool ADJ, bool FULL>
__device__ __forceinline__ void run_op(const DevOp& op, Regs<T>& p, Regs<T>& l, const GroupCtx& g, uint64_t base,
                                       const T* __restrict__ pool, const double* __restrict__ ang_row,
                                       TapAcc<T>& acc) {
    if ((base & op.ctrl_global) != op.ctrl_global) return;  // a non-local control is off: no-op, zero tap
    const uint32_t cl = op.ctrl_local;
    const int kind = op.kind;
    T c = 0, s = 0;
    if (kind >= OPK_RX) {  // (cos, sin) of the half angle, or of the angle for PS / GP
        c = T(ang_row[2 * op.slot]);
        s = T(ang_row[2 * op.slot + 1]);
        if (ADJ) s = -s;  // G^dagger = G(-angle)
    }
    const T* m = (op.mat >= 0) ? pool + op.mat : pool;

    if constexpr (ADJ) {
        if (op.tap >= 0) {
            T sr = 0, si = 0;
#define T1(KD, Q)                                                              \
    ((FULL && cl) ? one_target_tap<T, KD, Q, FULL>(p, l, g, cl, sr, si)          \
                  : one_target_tap<T, KD, Q, false>(p, l, g, cl, sr, si))
#define T2(KD, Q0, Q1)                                                         \
    ((FULL && cl) ? two_target_tap<T, KD, Q0, Q1, FULL>(p, l, g, cl, sr, si)     \
                  : two_target_tap<T, KD, Q0, Q1, false>(p, l, g, cl, sr, si))
#define RXT(Q) T1(OPK_RX, Q)
#define RYT(Q) T1(OPK_RY, Q)
#define RZT(Q) T1(OPK_RZ, Q)
#define PST(Q) T1(OPK_PS, Q)
#define GPT(Q) T1(OPK_GP, Q)
#define XXT(Q0, Q1) T2(OPK_RXX, Q0, Q1)
#define YYT(Q0, Q1) T2(OPK_RYY, Q0, Q1)
#define ZZT(Q0, Q1) T2(OPK_RZZ, Q0, Q1)
            switch (kind) {
                case OPK_RX: MQA_SW1(RXT) break;
                case OPK_RY: MQA_SW1(RYT) break;
                case OPK_RZ: MQA_SW1(RZT) break;
                case OPK_PS: MQA_SW1(PST) break;
                case OPK_GP: MQA_SW1(GPT) break;
                case OPK_RXX: MQA_SW2(XXT) break;
                case OPK_RYY: MQA_SW2(YYT) break;
                default: MQA_SW2(ZZT) break;
            }
#undef T1
#undef T2
#undef RXT
#undef RYT
#undef RZT
#undef PST
#undef GPT
#undef XXT
#undef YYT
#undef ZZT
            // Re(tap), tap = factor * s; only one component of s survives, so the compiler drops
            // the other half of the accumulation
            T tr;
            if (kind == OPK_PS) {  // i s
                tr = -si;
            } else if (kind == OPK_GP) {  // -i s
                tr = si;
            } else if (kind == OPK_RY) {  // -i/2 * (i s)
                tr = T(0.5) * sr;
            } else {  // -i/2 * s
                tr = T(0.5) * si;
            }
            // the group's taps arrive in order: shift them through the register file so that
            // every access uses a constant index (no dynamically indexed, scratch-backed array)
#pragma unroll
            for (int j = kTapSlots<T> - 1; j > 0; j--) acc.r[j] = acc.r[j - 1];
            acc.r[0] = tr;
        }
    }
    apply_op<T, FULL>(kind, op, p, g, cl, c, s, m);
    if constexpr (ADJ) apply_op<T, FULL>(kind, op, l, g, cl, c, s, m);
}
#undef MQA_SW1
#undef MQA_SW2

// ---------------------------------------------------------------- tile I/O
// HBM <-> LDS transfer of one tile.  Every access of a thread is issued before the first is
// waited on (full unroll) so that enough bytes are in flight per CU to stream at HBM rate.
// complex64 moves two adjacent amplitudes (16 B) per access: for u even, u and u + 1 are
// adjacent in HBM (the low run C >= 1 is contiguous) and share the 16-byte LDS slot at
// swz(u) & ~1, since swz(u ^ 1) == swz(u) ^ 1; when the swizzle sets bit 0 of swz(u) the
// two amplitudes sit in that slot in swapped order.
__device__ __forceinline__ float4 swap_halves_if(float4 v, bool odd) {
    return odd ? make_float4(v.z, v.w, v.x, v.y) : v;
}

template <typename T>
__device__ __forceinline__ void tile_load(V2<T>* __restrict__ lds, const V2<T>* __restrict__ g,
                                          const uint64_t* __restrict__ offtab, int C, uint32_t tid, uint32_t NT) {
    const uint32_t cmask = (1u << C) - 1u;
    if constexpr (sizeof(T) == 4) {
        float4 buf[kAmpsPerThread / 2];
#pragma unroll
        for (int j = 0; j < kAmpsPerThread / 2; j++) {
            const uint32_t u = 2u * (tid + NT * j);
            buf[j] = *reinterpret_cast<const float4*>(g + ((u & cmask) | offtab[u >> C]));
        }
#pragma unroll
        for (int j = 0; j < kAmpsPerThread / 2; j++) {
            const uint32_t a = swz<T>(2u * (tid + NT * j));
            *reinterpret_cast<float4*>(lds + (a & ~1u)) = swap_halves_if(buf[j], a & 1u);
        }
    } else {
        // complex128: two batches of 8 x 16 B, the same bytes in flight as complex64
#pragma unroll
        for (int h = 0; h < kAmpsPerThread; h += kAmpsPerThread / 2) {
            V2<T> buf[kAmpsPerThread / 2];
#pragma unroll
            for (int j = 0; j < kAmpsPerThread / 2; j++) {
                const uint32_t u = tid + NT * (h + j);
                buf[j] = g[(u & cmask) | offtab[u >> C]];
            }
#pragma unroll
            for (int j = 0; j < kAmpsPerThread / 2; j++) lds[swz<T>(tid + NT * (h + j))] = buf[j];
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

template <typename T>
__device__ __forceinline__ void tile_store(const V2<T>* __restrict__ lds, V2<T>* __restrict__ g,
                                           const uint64_t* __restrict__ offtab, int C, uint32_t tid, uint32_t NT) {
    const uint32_t cmask = (1u << C) - 1u;
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int j = 0; j < kAmpsPerThread / 2; j++) {
            const uint32_t u = 2u * (tid + NT * j);
            const uint32_t a = swz<T>(u);
            *reinterpret_cast<float4*>(g + ((u & cmask) | offtab[u >> C])) =
                swap_halves_if(*reinterpret_cast<const float4*>(lds + (a & ~1u)), a & 1u);
        }
    } else {
#pragma unroll
        for (int j = 0; j < kAmpsPerThread; j++) {
            const uint32_t u = tid + NT * j;
            g[(u & cmask) | offtab[u >> C]] = lds[swz<T>(u)];
        }
    }
}

// ---------------------------------------------------------------- the tile pass
template <typename T, bool ADJ, bool FULL>
__global__ void __launch_bounds__(512) k_tile_pass(TilePassArgs args) {
    extern __shared__ __align__(16) unsigned char smem[];
    using V = V2<T>;
    const int K = args.K;
    const uint32_t NT = blockDim.x;
    const uint32_t tid = threadIdx.x;
    const DevPass pass = args.passes[args.pass_index];
    const int C = pass.low_run;
    const int H = K - C;

    const LdsLayout lay = lds_layout(K, C, ADJ, int(sizeof(V)), 0);
    V* t_psi = reinterpret_cast<V*>(smem + lay.psi);
    V* t_lam = reinterpret_cast<V*>(smem + lay.lam);  // used by the adjoint pass only
    uint64_t* offtab = reinterpret_cast<uint64_t*>(smem + lay.offtab);
    double* lds_taps = reinterpret_cast<double*>(smem + lay.taps);

    const uint64_t sample = blockIdx.y;
    const uint64_t tile = blockIdx.x;
    const uint64_t base = pdep64(tile, pass.nonlocal_mask);
    for (uint32_t h = tid; h < (1u << H); h += NT) offtab[h] = pdep64(h, pass.local_hi_mask);
    __syncthreads();

    V* g_psi = static_cast<V*>(args.psi) + sample * args.stride + base;
    V* g_lam = ADJ ? (static_cast<V*>(args.lam) + sample * args.stride + base) : nullptr;
    tile_load<T>(t_psi, g_psi, offtab, C, tid, NT);
    if constexpr (ADJ) tile_load<T>(t_lam, g_lam, offtab, C, tid, NT);
    __syncthreads();

    const int wave = int(tid >> 6);
    const int n_waves = int((NT + 63) >> 6);
    const int wave_width = NT < 64 ? int(NT) : 64;
    const double* ang_row = args.angles ? args.angles + 2 * sample * uint64_t(args.n_slots) : nullptr;

    for (int gi = 0; gi < pass.n_groups; gi++) {
        const DevGroup grp = args.groups[pass.group_begin + gi];
        GroupCtx g;
        uint32_t ub = tid;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint32_t q = grp.spos[k];
            ub = ((ub >> q) << (q + 1)) | (ub & ((1u << q) - 1u));
        }
        g.ub = ub;
#pragma unroll
        for (int k = 0; k < 4; k++) g.pm[k] = 1u << grp.pos[k];

        // LDS slot of register v = swz(ub) ^ swz(offset of v): a per-thread base XOR a
        // wave-uniform offset precomputed by the host
        const uint32_t a0 = swz<T>(ub);
        Regs<T> p, l;
#pragma unroll
        for (int v = 0; v < kAmpsPerThread; v++) {
            const V x = t_psi[a0 ^ grp.off[v]];
            p.r[v] = x.x, p.i[v] = x.y;
            if constexpr (ADJ) {
                const V y = t_lam[a0 ^ grp.off[v]];
                l.r[v] = y.x, l.i[v] = y.y;
            }
        }

        TapAcc<T> acc;
        if constexpr (ADJ) {
#pragma unroll
            for (int j = 0; j < kTapSlots<T>; j++) acc.r[j] = T(0);
        }
        const DevOp op0 = args.ops[grp.op_begin];
        if (FULL && (op0.kind == OPK_M3 || op0.kind == OPK_M4)) {  // a dense block alone in its group
            if ((base & op0.ctrl_global) == op0.ctrl_global) {
                const T* m = static_cast<const T*>(args.pool) + op0.mat;
                if (op0.kind == OPK_M3) {
                    dense_from_lds<T, 3>(p, g, op0.ctrl_local, m, t_psi);
                    if constexpr (ADJ) dense_from_lds<T, 3>(l, g, op0.ctrl_local, m, t_lam);
                } else {
                    dense_from_lds<T, 4>(p, g, op0.ctrl_local, m, t_psi);
                    if constexpr (ADJ) dense_from_lds<T, 4>(l, g, op0.ctrl_local, m, t_lam);
                }
            }
        } else {
            for (int oi = 0; oi < grp.n_ops; oi++) {
                const DevOp op = args.ops[grp.op_begin + oi];
                run_op<T, ADJ, FULL>(op, p, l, g, base, static_cast<const T*>(args.pool), ang_row, acc);
            }
        }
        if constexpr (ADJ) {
            if (grp.n_taps > 0) {  // one batched cross-lane reduction per group
                const bool lane0 = (tid & 63) == 0;
                if (wave_width == 64) {
#pragma unroll
                    for (int j = 0; j < kTapSlots<T>; j++) {
                        if (j < 4 || grp.n_taps > 4) acc.r[j] = wave_sum64(acc.r[j]);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < kTapSlots<T>; j++) acc.r[j] = wave_sum(acc.r[j], wave_width);
                }
                if (lane0) {
#pragma unroll
                    for (int j = 0; j < kTapSlots<T>; j++) {
                        if (j < grp.n_taps) {  // acc[j] holds the group's (n_taps - 1 - j)-th tap
                            const int t = grp.tap_begin + grp.n_taps - 1 - j;
                            lds_taps[t * n_waves + wave] = double(acc.r[j]);
                        }
                    }
                }
            }
        }

#pragma unroll
        for (int v = 0; v < kAmpsPerThread; v++) {
            t_psi[a0 ^ grp.off[v]] = V{p.r[v], p.i[v]};
            if constexpr (ADJ) t_lam[a0 ^ grp.off[v]] = V{l.r[v], l.i[v]};
        }
        __syncthreads();
    }

    // launder tid so that the address arithmetic of the load phase is recomputed here rather
    // than kept live in registers through the whole group loop
    uint32_t tid_st = tid;
    asm volatile("" : "+v"(tid_st));
    tile_store<T>(t_psi, g_psi, offtab, C, tid_st, NT);
    if constexpr (ADJ) tile_store<T>(t_lam, g_lam, offtab, C, tid_st, NT);

    if constexpr (ADJ) {
        // one partial per (sample, tap id, tile); summed over tiles by k_reduce_rows
        for (int t = int(tid); t < pass.n_taps; t += int(NT)) {
            double sr = 0;
            for (int w = 0; w < n_waves; w++) sr += lds_taps[t * n_waves + w];
            const int tap_id = args.tap_ids[pass.tap_begin + t];
            args.partials[(sample * uint64_t(args.n_tap_ids) + uint64_t(tap_id)) * args.n_tiles + tile] = sr;
        }
    }
}

// ---------------------------------------------------------------- Hamiltonian / reductions
template <typename T>
__device__ __forceinline__ void kahan_add(T x, T& sum, T& comp) {
    T y = x - comp;
    T t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// Block-wide double sum of (a, b) written to out[0..1] by thread 0 (up to 16 waves).
constexpr size_t kBlockSumLds = 2 * 16 * sizeof(double);
__device__ __forceinline__ void block_sum2(double a, double b, double* out) {
    __shared__ double red[2 * 16];
    a = wave_sum(a, 64);
    b = wave_sum(b, 64);
    const int wave = threadIdx.x >> 6, n_waves = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) red[2 * wave] = a, red[2 * wave + 1] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, sb = 0;
        for (int w = 0; w < n_waves; w++) sa += red[2 * w], sb += red[2 * w + 1];
        out[0] = sa, out[1] = sb;
    }
}

// XCD-aware block order: the hardware deals consecutive workgroups round-robin over the 8
// XCDs; remap so that each XCD walks one contiguous range of the state (its own L2 holds the
// neighbourhood that the flip-mask gathers revisit).
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nblk) {
    if (nblk % 8 != 0) return bid;
    return (bid % 8) * (nblk / 8) + bid / 8;
}

template <typename T, bool WRITE>
__global__ void __launch_bounds__(256) k_ham(HamArgs args) {
    using V = V2<T>;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t sample = blockIdx.y;
    const V* ket = static_cast<const V*>(args.ket) + sample * args.ket_stride;
    const V* bra = args.bra ? static_cast<const V*>(args.bra) + sample * args.bra_stride : nullptr;
    V* out = WRITE ? static_cast<V*>(args.out) + sample * args.out_stride : nullptr;
    T fr = 0, fi = 0, cr = 0, ci = 0;
    const uint64_t first = uint64_t(blk) * args.amps_per_block;
    for (uint32_t m = 0; m < args.amps_per_block; m += blockDim.x) {
        const uint64_t j = first + m + threadIdx.x;
        if (j >= args.dim) break;
        T accr = 0, acci = 0;
        for (int gi = 0; gi < args.n_groups; gi++) {
            const DevHamGroup grp = args.groups[gi];
            const uint64_t src = j ^ grp.flip;
            T wr = 0, wi = 0;
            for (int t = 0; t < grp.n_terms; t++) {
                const DevHamTerm tm = args.terms[grp.term_begin + t];
                const bool neg = (__popcll(src & tm.yz) & 1) != 0;
                const T er = T(tm.coef_re), ei = T(tm.coef_im);
                wr += neg ? -er : er;
                wi += neg ? -ei : ei;
            }
            const V x = ket[src];
            accr += wr * x.x - wi * x.y;
            acci += wr * x.y + wi * x.x;
        }
        if constexpr (WRITE) out[j] = V{accr, acci};
        if (bra) {
            const V y = bra[j];
            kahan_add(y.x * accr + y.y * acci, fr, cr);
            kahan_add(y.x * acci - y.y * accr, fi, ci);
        }
    }
    if (args.partial) {
        block_sum2(double(fr), double(fi), args.partial + 2 * (sample * gridDim.x + blockIdx.x));
    }
}

// ---------------------------------------------------------------- Hamiltonian tile pass
// out (+)= sum_g w_g(u) * ket[u ^ flip_g] over the flip groups of one pass, for every amplitude u
// of an LDS-staged ket tile (reference: cpu_vector_core_policy.cpp:132-161 ApplyTerms, which
// sweeps the whole state once per term).  The term sign (-1)^{popcount(src & yz)} factors over
// the global index bits: bits outside the tile (one parity per tile), the thread-index bits of the
// tile index and the register bits, so most terms cost a few operations per thread, not per
// amplitude.  Each thread owns 16 amplitudes, u(J) = 2 tid | (J & 1) | (J >> 1) << (K - 3),
// read and written in adjacent pairs.  The last pass also reduces <ket|out> per tile.
template <typename T>
__device__ __forceinline__ T flip_sign(T x, uint32_t neg);
template <>
__device__ __forceinline__ float flip_sign<float>(float x, uint32_t neg) {
    return __uint_as_float(__float_as_uint(x) ^ (neg << 31));
}
template <>
__device__ __forceinline__ double flip_sign<double>(double x, uint32_t neg) {
    return __longlong_as_double(__double_as_longlong(x) ^ (static_cast<long long>(neg) << 63));
}

template <typename T>
__device__ __forceinline__ void pair_load(const V2<T>* p, T& r0, T& i0, T& r1, T& i1) {
    if constexpr (sizeof(T) == 4) {
        const float4 v = *reinterpret_cast<const float4*>(p);
        r0 = v.x, i0 = v.y, r1 = v.z, i1 = v.w;
    } else {
        const double2 a = p[0], b = p[1];
        r0 = a.x, i0 = a.y, r1 = b.x, i1 = b.y;
    }
}
template <typename T>
__device__ __forceinline__ void pair_store(V2<T>* p, T r0, T i0, T r1, T i1) {
    if constexpr (sizeof(T) == 4) {
        *reinterpret_cast<float4*>(p) = make_float4(r0, i0, r1, i1);
    } else {
        p[0] = double2{r0, i0};
        p[1] = double2{r1, i1};
    }
}

template <typename T>
constexpr int ham_tile_threads() {
    return sizeof(T) == 4 ? 1024 : 512;
}

template <typename T, bool FIRST, bool LAST>
__global__ void __launch_bounds__(ham_tile_threads<T>()) k_ham_tile(HamTileArgs args) {
    extern __shared__ __align__(16) unsigned char smem[];
    using V = V2<T>;
    constexpr int NA = kAmpsPerThread;
    const int K = args.K;
    const uint32_t NT = blockDim.x;
    const uint32_t tid = threadIdx.x;
    const DevHamPass pass = args.passes[args.pass_index];
    const int C = pass.low_run;
    const int H = K - C;
    const LdsLayout lay = lds_layout(K, C, false, int(sizeof(V)), 0);
    V* t_ket = reinterpret_cast<V*>(smem + lay.psi);
    uint64_t* offtab = reinterpret_cast<uint64_t*>(smem + lay.offtab);

    const uint64_t sample = blockIdx.y;
    const uint64_t tile = blockIdx.x;
    const uint64_t base = pdep64(tile, pass.nonlocal_mask);
    for (uint32_t h = tid; h < (1u << H); h += NT) offtab[h] = pdep64(h, pass.local_hi_mask);
    __syncthreads();

    const V* g_ket = static_cast<const V*>(args.ket) + sample * args.stride + base;
    V* g_out = static_cast<V*>(args.out) + sample * args.stride + base;
    tile_load<T>(t_ket, g_ket, offtab, C, tid, NT);

    const uint32_t cmask = (1u << C) - 1u;
    const int JS = K - 3;  // tile-index position of register bit 1
    T ar[NA], ai[NA];
    if constexpr (FIRST) {
#pragma unroll
        for (int J = 0; J < NA; J++) ar[J] = ai[J] = T(0);
    } else {
#pragma unroll
        for (int P = 0; P < NA / 2; P++) {
            const uint32_t u = (2u * tid) | (uint32_t(P) << JS);
            pair_load<T>(g_out + ((u & cmask) | offtab[u >> C]), ar[2 * P], ai[2 * P], ar[2 * P + 1], ai[2 * P + 1]);
        }
    }
    // LDS slot of owned amplitude J: swz(2 tid) ^ swz(jb(J)), the swizzle being XOR-linear
    const uint32_t a0 = swz<T>(2u * tid);
    auto jb = [JS](int J) { return uint32_t(J & 1) | (uint32_t(J >> 1) << JS); };
    __syncthreads();

    const uint32_t tmask = ((1u << JS) - 1u) & ~1u;
    for (int gi = 0; gi < pass.n_groups; gi++) {
        const DevHamTileGroup grp = args.groups[pass.group_begin + gi];
        const DevHamTileTerm* tm = args.terms + grp.term_begin;
        if (grp.kind == HAMW_TILE) {
            T wr = 0, wi = 0;
            for (int t = 0; t < grp.n_terms; t++) {
                const uint32_t neg = uint32_t(__popcll(base & tm[t].yz_nl)) & 1u;
                wr += flip_sign(T(tm[t].coef_re), neg);
                wi += flip_sign(T(tm[t].coef_im), neg);
            }
#pragma unroll
            for (int J = 0; J < NA; J++) {
                const V x = t_ket[a0 ^ swz<T>(jb(J) ^ grp.flipl)];
                ar[J] += wr * x.x - wi * x.y;
                ai[J] += wr * x.y + wi * x.x;
            }
        } else if (grp.kind == HAMW_THREAD) {
            // source index bits on this thread: (2 tid) ^ flip restricted to the thread bits
            const uint32_t tsrc = ((2u * tid) ^ grp.flipl) & tmask;
            T wr = 0, wi = 0;
            for (int t = 0; t < grp.n_terms; t++) {
                const uint32_t neg = (uint32_t(__popcll(base & tm[t].yz_nl)) + __popc(tsrc & tm[t].yzl)) & 1u;
                wr += flip_sign(T(tm[t].coef_re), neg);
                wi += flip_sign(T(tm[t].coef_im), neg);
            }
#pragma unroll
            for (int J = 0; J < NA; J++) {
                const V x = t_ket[a0 ^ swz<T>(jb(J) ^ grp.flipl)];
                ar[J] += wr * x.x - wi * x.y;
                ai[J] += wr * x.y + wi * x.x;
            }
        } else {
            const uint32_t tsrc = ((2u * tid) ^ grp.flipl) & tmask;
            // two halves of the owned amplitudes bound the weight registers
#pragma unroll
            for (int half = 0; half < 2; half++) {
                constexpr int NH = NA / 2;
                T cr0 = 0, ci0 = 0;
                T wr[NH], wi[NH];
#pragma unroll
                for (int J = 0; J < NH; J++) wr[J] = wi[J] = T(0);
                for (int t = 0; t < grp.n_terms; t++) {
                    const uint32_t yzl = tm[t].yzl;
                    const uint32_t neg = (uint32_t(__popcll(base & tm[t].yz_nl)) + __popc(tsrc & yzl)) & 1u;
                    const T cr = flip_sign(T(tm[t].coef_re), neg);
                    const T ci = flip_sign(T(tm[t].coef_im), neg);
                    const uint32_t jm = yzl & ~tmask;  // yz bits on register bits
                    if (jm == 0) {
                        cr0 += cr;
                        ci0 += ci;
                    } else {
                        const uint32_t p0 = uint32_t(__popc(grp.flipl & jm)) & 1u;
#pragma unroll
                        for (int J = 0; J < NH; J++) {
                            const uint32_t nj = p0 ^ (uint32_t(__popc(jb(half * NH + J) & jm)) & 1u);
                            wr[J] += flip_sign(cr, nj);
                            wi[J] += flip_sign(ci, nj);
                        }
                    }
                }
#pragma unroll
                for (int J = 0; J < NH; J++) {
                    const int JJ = half * NH + J;
                    const V x = t_ket[a0 ^ swz<T>(jb(JJ) ^ grp.flipl)];
                    const T w_r = cr0 + wr[J], w_i = ci0 + wi[J];
                    ar[JJ] += w_r * x.x - w_i * x.y;
                    ai[JJ] += w_r * x.y + w_i * x.x;
                }
            }
        }
    }

#pragma unroll
    for (int P = 0; P < NA / 2; P++) {
        const uint32_t u = (2u * tid) | (uint32_t(P) << JS);
        pair_store<T>(g_out + ((u & cmask) | offtab[u >> C]), ar[2 * P], ai[2 * P], ar[2 * P + 1], ai[2 * P + 1]);
    }
    if constexpr (LAST) {
        if (args.partial) {  // <ket|out> of this tile, summed in double
            double fr = 0, fi = 0;
#pragma unroll
            for (int J = 0; J < NA; J++) {
                const V x = t_ket[a0 ^ swz<T>(jb(J))];
                fr += double(x.x * ar[J] + x.y * ai[J]);
                fi += double(x.x * ai[J] - x.y * ar[J]);
            }
            block_sum2(fr, fi, args.partial + 2 * (sample * gridDim.x + tile));
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_vdot(VdotArgs args) {
    using V = V2<T>;
    const uint64_t sample = blockIdx.y;
    const V* a = static_cast<const V*>(args.a) + sample * args.a_stride;
    const V* b = static_cast<const V*>(args.b) + sample * args.b_stride;
    T fr = 0, fi = 0, cr = 0, ci = 0;
    const uint64_t first = uint64_t(blockIdx.x) * args.amps_per_block;
    for (uint32_t m = 0; m < args.amps_per_block; m += blockDim.x) {
        const uint64_t j = first + m + threadIdx.x;
        if (j >= args.dim) break;
        const V x = a[j], y = b[j];
        kahan_add(x.x * y.x + x.y * y.y, fr, cr);
        kahan_add(x.x * y.y - x.y * y.x, fi, ci);
    }
    block_sum2(double(fr), double(fi), args.partial + 2 * (sample * gridDim.x + blockIdx.x));
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill(V2<T>* dst, const V2<T>* src, uint64_t dim, uint64_t stride,
                                             uint64_t src_dim, uint64_t batch) {
    const uint64_t total = dim * batch;
    for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t b = e / dim, i = e - b * dim;
        V2<T> v;
        if (src) {
            v = (i < src_dim) ? src[i] : V2<T>{T(0), T(0)};
        } else {
            v = V2<T>{T(i == 0 ? 1 : 0), T(0)};
        }
        dst[b * stride + i] = v;
    }
}

template <typename S, typename D>
__global__ void __launch_bounds__(256) k_cast(const S* src, D* dst, uint64_t n) {
    for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += uint64_t(gridDim.x) * blockDim.x)
        dst[e] = D(src[e]);
}

// Per-sample (cos, sin) table of every angle slot: angle = const + sum coeff * param, where
// param p < n_enc reads enc[b][p] and p >= n_enc reads ans[p - n_enc]; half[slot] selects the
// half angle (rotations) or the full angle (PhaseShift, GlobalPhase).  Double precision.
__global__ void __launch_bounds__(256) k_angles(AngleArgs a) {
    const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= a.batch * uint64_t(a.n_slots)) return;
    const uint64_t b = e / a.n_slots;
    const int slot = int(e - b * a.n_slots);
    double th = a.const_val[slot];
    for (int k = a.term_begin[slot]; k < a.term_begin[slot + 1]; k++) {
        const int p = a.term_param[k];
        const double v = (p < a.n_enc) ? a.enc[b * a.n_enc + p] : a.ans[p - a.n_enc];
        th += a.term_coeff[k] * v;
    }
    if (a.half[slot]) th *= 0.5;
    double sn, cs;
    sincos(th, &sn, &cs);
    a.out[2 * e] = cs;
    a.out[2 * e + 1] = sn;
}

// rows x n (re, im) -> rows (re, im), one wave per row, fixed order (deterministic)
__global__ void __launch_bounds__(256) k_reduce_rows(const double* __restrict__ in, uint64_t n, double* __restrict__ out,
                                                    uint64_t rows) {
    const uint64_t row = uint64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= rows) return;
    const int lane = threadIdx.x & 63;
    double a = 0, b = 0;
    for (uint64_t k = lane; k < n; k += 64) {
        a += in[2 * (row * n + k)];
        b += in[2 * (row * n + k) + 1];
    }
    a = wave_sum(a, 64);
    b = wave_sum(b, 64);
    if (lane == 0) out[2 * row] = a, out[2 * row + 1] = b;
}

__global__ void __launch_bounds__(256) k_reduce_rows1(const double* __restrict__ in, uint64_t n,
                                                     double* __restrict__ out, uint64_t rows) {
    const uint64_t row = uint64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= rows) return;
    const int lane = threadIdx.x & 63;
    double a = 0;
    for (uint64_t k = lane; k < n; k += 64) a += in[row * n + k];
    a = wave_sum(a, 64);
    if (lane == 0) out[row] = a;
}

// g[b][p] = sum_{e in param p} coeff_e * 2 Re(tap[b][tap_e]); taps hold Re(tap) only
__global__ void __launch_bounds__(256) k_jacobian(const double* __restrict__ taps, int n_tap_ids,
                                                 const int32_t* __restrict__ jac_begin,
                                                 const int32_t* __restrict__ jac_tap,
                                                 const double* __restrict__ jac_coeff, int n_params,
                                                 double* __restrict__ g, uint64_t batch) {
    const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= batch * uint64_t(n_params)) return;
    const uint64_t b = e / n_params;
    const int p = int(e - b * n_params);
    double acc = 0;
    for (int k = jac_begin[p]; k < jac_begin[p + 1]; k++)
        acc += jac_coeff[k] * 2.0 * taps[b * n_tap_ids + jac_tap[k]];
    g[e] = acc;
}

// ---------------------------------------------------------------- launchers
static constexpr size_t kLdsLimit = 160 * 1024;

size_t tile_pass_lds_bytes(int dtype, bool adj, int K, int C, int max_taps) {
    return lds_layout(K, C, adj, dtype == 0 ? 8 : 16, max_taps).total;
}

template <typename T, bool ADJ, bool FULL>
static hipError_t launch_tp(const TilePassArgs& a, uint64_t batch, size_t lds, hipStream_t s) {
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_pass<T, ADJ, FULL>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(kLdsLimit));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    dim3 grid(uint32_t(a.n_tiles), uint32_t(batch));
    dim3 block(1u << (a.K - kGroupWidth));
    hipLaunchKernelGGL((k_tile_pass<T, ADJ, FULL>), grid, block, lds, s, a);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_tp_t(bool adj, bool full, const TilePassArgs& a, uint64_t batch, size_t lds, hipStream_t s) {
    if (adj) return full ? launch_tp<T, true, true>(a, batch, lds, s) : launch_tp<T, true, false>(a, batch, lds, s);
    return full ? launch_tp<T, false, true>(a, batch, lds, s) : launch_tp<T, false, false>(a, batch, lds, s);
}

hipError_t launch_tile_pass(int dtype, bool adj, bool full, const TilePassArgs& a, uint64_t batch, int C,
                            int max_taps, hipStream_t s) {
    size_t lds = tile_pass_lds_bytes(dtype, adj, a.K, C, max_taps);
    if (lds > kLdsLimit) return hipErrorInvalidValue;
    if (dtype == 0) return launch_tp_t<float>(adj, full, a, batch, lds, s);
    return launch_tp_t<double>(adj, full, a, batch, lds, s);
}

hipError_t launch_ham(int dtype, bool write, const HamArgs& a, uint64_t batch, hipStream_t s) {
    dim3 grid(uint32_t(a.n_blocks), uint32_t(batch));
    if (dtype == 0) {
        if (write) hipLaunchKernelGGL((k_ham<float, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_ham<float, false>), grid, dim3(256), 0, s, a);
    } else {
        if (write) hipLaunchKernelGGL((k_ham<double, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_ham<double, false>), grid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

template <typename T, bool FIRST, bool LAST>
static hipError_t launch_ht(const HamTileArgs& a, uint64_t n_tiles, uint64_t batch, size_t lds, hipStream_t s) {
    static bool attr_set = false;
    if (!attr_set) {
        // the block reduction's static LDS comes out of the same 160 KB
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ham_tile<T, FIRST, LAST>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           int(kLdsLimit - kBlockSumLds));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const uint32_t threads = 1u << (a.K - kGroupWidth);
    if (threads > uint32_t(ham_tile_threads<T>())) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_ham_tile<T, FIRST, LAST>), dim3(uint32_t(n_tiles), uint32_t(batch)), dim3(threads), lds, s,
                       a);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_ht_t(bool first, bool last, const HamTileArgs& a, uint64_t n_tiles, uint64_t batch,
                              size_t lds, hipStream_t s) {
    if (first) return last ? launch_ht<T, true, true>(a, n_tiles, batch, lds, s)
                           : launch_ht<T, true, false>(a, n_tiles, batch, lds, s);
    return last ? launch_ht<T, false, true>(a, n_tiles, batch, lds, s)
                : launch_ht<T, false, false>(a, n_tiles, batch, lds, s);
}

hipError_t launch_ham_tile(int dtype, bool first, bool last, const HamTileArgs& a, int C, uint64_t n_tiles,
                           uint64_t batch, hipStream_t s) {
    const size_t amp = dtype == 0 ? 8 : 16;
    const size_t lds = lds_layout(a.K, C, false, int(amp), 0).total;
    if (a.K < kMinTileQubits || a.K > kMaxHamTileQubits || lds + kBlockSumLds > kLdsLimit) return hipErrorInvalidValue;
    if (dtype == 0) return launch_ht_t<float>(first, last, a, n_tiles, batch, lds, s);
    return launch_ht_t<double>(first, last, a, n_tiles, batch, lds, s);
}

hipError_t launch_vdot(int dtype, const VdotArgs& a, uint64_t batch, hipStream_t s) {
    dim3 grid(uint32_t(a.n_blocks), uint32_t(batch));
    if (dtype == 0) hipLaunchKernelGGL(k_vdot<float>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_vdot<double>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fill(int dtype, void* dst, const void* src, uint64_t dim, uint64_t stride, uint64_t src_dim,
                       uint64_t batch, hipStream_t s) {
    uint64_t total = dim * batch;
    uint32_t blocks = uint32_t(std::min<uint64_t>((total + 255) / 256, 65536));
    if (dtype == 0)
        hipLaunchKernelGGL(k_fill<float>, dim3(blocks), dim3(256), 0, s, static_cast<float2*>(dst),
                           static_cast<const float2*>(src), dim, stride, src_dim, batch);
    else
        hipLaunchKernelGGL(k_fill<double>, dim3(blocks), dim3(256), 0, s, static_cast<double2*>(dst),
                           static_cast<const double2*>(src), dim, stride, src_dim, batch);
    return hipGetLastError();
}

hipError_t launch_cast(int src_dtype, const void* src, void* dst, uint64_t n_reals, hipStream_t s) {
    uint32_t blocks = uint32_t(std::min<uint64_t>((n_reals + 255) / 256, 65536));
    if (src_dtype == 0)
        hipLaunchKernelGGL((k_cast<float, double>), dim3(blocks), dim3(256), 0, s, static_cast<const float*>(src),
                           static_cast<double*>(dst), n_reals);
    else
        hipLaunchKernelGGL((k_cast<double, float>), dim3(blocks), dim3(256), 0, s, static_cast<const double*>(src),
                           static_cast<float*>(dst), n_reals);
    return hipGetLastError();
}

hipError_t launch_angles(const AngleArgs& a, hipStream_t s) {
    uint64_t total = a.batch * uint64_t(a.n_slots);
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_angles, dim3(uint32_t((total + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_reduce_rows(const double* in, uint64_t n, double* out, uint64_t rows, hipStream_t s) {
    uint32_t blocks = uint32_t((rows + 3) / 4);
    hipLaunchKernelGGL(k_reduce_rows, dim3(blocks), dim3(256), 0, s, in, n, out, rows);
    return hipGetLastError();
}

hipError_t launch_reduce_rows1(const double* in, uint64_t n, double* out, uint64_t rows, hipStream_t s) {
    uint32_t blocks = uint32_t((rows + 3) / 4);
    hipLaunchKernelGGL(k_reduce_rows1, dim3(blocks), dim3(256), 0, s, in, n, out, rows);
    return hipGetLastError();
}

hipError_t launch_jacobian(const double* taps, int n_tap_ids, const int32_t* jac_begin, const int32_t* jac_tap,
                           const double* jac_coeff, int n_params, double* g, uint64_t batch, hipStream_t s) {
    uint64_t total = batch * uint64_t(n_params);
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_jacobian, dim3(uint32_t((total + 255) / 256)), dim3(256), 0, s, taps, n_tap_ids, jac_begin,
                       jac_tap, jac_coeff, n_params, g, batch);
    return hipGetLastError();
}

}  // namespace mqa
