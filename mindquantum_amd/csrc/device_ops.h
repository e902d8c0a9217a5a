// device_ops.h - device-side building blocks of the tile passes: LDS swizzle helpers, wave
// reductions, the register-group op bodies, HBM <-> LDS tile transfer and the pass frame.
// Included by kernels.hip and, as source text, by the per-circuit kernels generated at run
// time (jit.cpp), so it depends on nothing but program.h.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <cstdint>

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
// Tables a kernel only reads (angles, matrix pool), seen through the constant address space:
// uniform reads of them stay scalar loads even in a loop that also stores to global memory
// (where the compiler could not otherwise prove them unclobbered).
template <typename T>
using const_ptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ const_ptr<T> as_const(const T* p) {
    return (const_ptr<T>)p;
}

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
// Cross-row steps with gfx950's lane swaps (VALU, no LDS crossbar): permlane16_swap exchanges
// odd rows of its first operand with even rows of its second, permlane32_swap the upper half of
// the first with the lower half of the second; with v in both, the two results sum to the
// xor-16 / xor-32 partner sum in every lane.
template <bool ROWS32>
__device__ __forceinline__ float swap_sum(float v) {
    const unsigned u = __float_as_uint(v);
    if constexpr (ROWS32) {
        const auto r = __builtin_amdgcn_permlane32_swap(u, u, false, false);
        return __uint_as_float(r[0]) + __uint_as_float(r[1]);
    } else {
        const auto r = __builtin_amdgcn_permlane16_swap(u, u, false, false);
        return __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
}
template <bool ROWS32>
__device__ __forceinline__ double swap_sum(double v) {
    const unsigned long long b = __double_as_longlong(v);
    const unsigned lo = unsigned(b), hi = unsigned(b >> 32);
    unsigned a0, a1, b0, b1;
    if constexpr (ROWS32) {
        const auto rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        a0 = rl[0], a1 = rl[1], b0 = rh[0], b1 = rh[1];
    } else {
        const auto rl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto rh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        a0 = rl[0], a1 = rl[1], b0 = rh[0], b1 = rh[1];
    }
    return __longlong_as_double((static_cast<long long>(b0) << 32) | a0) +
           __longlong_as_double((static_cast<long long>(b1) << 32) | a1);
}
template <typename T>
__device__ __forceinline__ T wave_sum64(T v) {
    v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_mov<0x141>(v);  // row_half_mirror
    v += dpp_mov<0x140>(v);  // row_mirror
    v = swap_sum<false>(v);  // rows 0 + 1, 2 + 3
    return swap_sum<true>(v);
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

// x / sqrt(2).  In float, 1 / sqrt(2) rounds down (2 r^2 = 1 - 3.1e-8), so a plain product shrinks
// the norm of the state by that factor at every Hadamard (-4e-5 after 1400 of them); the product
// with r = hi + lo (lo the rounding remainder of hi) rounds without that bias.
template <typename T>
__device__ __forceinline__ T hmul(T x) {
    if constexpr (sizeof(T) == 4) {
        constexpr float hi = 0.70710678118654752440f;
        constexpr float lo = float(0.70710678118654752440084436210484903928 - double(hi));
        return fma(x, hi, x * lo);
    } else {
        return x * T(0.70710678118654752440084436210484903928);
    }
}

// A register set times a constant (a compiled pass's sqrt(2)^-n after n unnormalised Hadamards),
// given as hi + lo (lo: the rounding error of hi, zero for a power of two), so that the product is
// as close to exact as the per-gate 1 / sqrt(2) of hmul.
template <typename T>
__device__ __forceinline__ void scale_regs(Regs<T>& a, T hi, T lo) {
#pragma unroll
    for (int v = 0; v < kAmpsPerThread; v++) {
        if (lo == T(0)) {
            a.r[v] *= hi, a.i[v] *= hi;
        } else {
            a.r[v] = fma(a.r[v], hi, a.r[v] * lo);
            a.i[v] = fma(a.i[v], hi, a.i[v] * lo);
        }
    }
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
        } else if constexpr (KIND == OPK_H) {  // [[1, 1], [1, -1]] / sqrt(2)
            xr = hmul<T>(ar + br), xi = hmul<T>(ai + bi);
            yr = hmul<T>(ar - br), yi = hmul<T>(ai - bi);
        } else if constexpr (KIND == OPK_HU) {  // [[1, 1], [1, -1]]
            xr = ar + br, xi = ai + bi;
            yr = ar - br, yi = ai - bi;
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

// Re(tap) of a parameterised op needs one real component of a complex sum: Re(conj(x) y) or
// Im(conj(x) y).  Accumulates that component only.
template <bool IM, typename T>
__device__ __forceinline__ void cdot_part(T xr, T xi, T yr, T yi, T& acc) {
    if constexpr (IM) {
        acc = fma(xr, yi, acc);
        acc = fma(-xi, yr, acc);
    } else {
        acc = fma(xr, yr, acc);
        acc = fma(xi, yi, acc);
    }
}

// Which component of the raw sum s gives Re(tap), and the factor: tap = factor * s with
//   RX, RZ, RXX, RYY, RZZ: s = <l|P|p>, tap = -i/2 s  -> Re(tap) =  Im(s) / 2
//   RY: s = <l|Y|p> / i, tap = -i/2 (i s)            -> Re(tap) =  Re(s) / 2
//   PS: s = sum conj(l_b) p_b, tap = i s              -> Re(tap) = -Im(s)
//   GP: s = <l|p>, tap = -i s                         -> Re(tap) =  Im(s)
template <int KIND>
constexpr bool kTapIm = KIND != OPK_RY;
template <typename T, int KIND>
__device__ __forceinline__ T tap_factor() {
    return KIND == OPK_PS ? T(-1) : (KIND == OPK_GP ? T(1) : T(0.5));
}

// Re(tap) of a one-target parameterised op, taken before G^dagger is applied.
template <typename T, int KIND, int Q, bool CTRL>
__device__ __forceinline__ T one_target_tap(const Regs<T>& p, const Regs<T>& l, const GroupCtx& g, uint32_t cl) {
    constexpr bool IM = kTapIm<KIND>;
    T s = 0;
    for_pairs<Q>([&]<int V, int W>() {
        T u = 0;
        T& x = CTRL ? u : s;
        if constexpr (KIND == OPK_RX) {
            cdot_part<IM>(l.r[V], l.i[V], p.r[W], p.i[W], x);
            cdot_part<IM>(l.r[W], l.i[W], p.r[V], p.i[V], x);
        } else if constexpr (KIND == OPK_RY) {  // <l|Y|p> = i (conj(l_b) p_a - conj(l_a) p_b)
            cdot_part<IM>(l.r[W], l.i[W], p.r[V], p.i[V], x);
            cdot_part<IM>(l.r[V], l.i[V], -p.r[W], -p.i[W], x);
        } else if constexpr (KIND == OPK_RZ) {
            cdot_part<IM>(l.r[V], l.i[V], p.r[V], p.i[V], x);
            cdot_part<IM>(l.r[W], l.i[W], -p.r[W], -p.i[W], x);
        } else if constexpr (KIND == OPK_PS) {
            cdot_part<IM>(l.r[W], l.i[W], p.r[W], p.i[W], x);
        } else {  // GP
            cdot_part<IM>(l.r[V], l.i[V], p.r[V], p.i[V], x);
            cdot_part<IM>(l.r[W], l.i[W], p.r[W], p.i[W], x);
        }
        if constexpr (CTRL) s += ctrl_ok<V>(g, cl) ? u : T(0);
    });
    return tap_factor<T, KIND>() * s;
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
        } else if constexpr (KIND == OPK_PF_RZZ) {  // diag(1, e, e, 1), e = c + is (full angle)
            T r1 = c * a.r[V1] - s * a.i[V1], i1 = c * a.i[V1] + s * a.r[V1];
            T r2 = c * a.r[V2_] - s * a.i[V2_], i2 = c * a.i[V2_] + s * a.r[V2_];
            a.r[V1] = r1, a.i[V1] = i1, a.r[V2_] = r2, a.i[V2_] = i2;
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

// Re(tap) of RXX / RYY / RZZ: tap = -i/2 s, s = <l|PP|p>
template <typename T, int KIND, int Q0, int Q1, bool CTRL>
__device__ __forceinline__ T two_target_tap(const Regs<T>& p, const Regs<T>& l, const GroupCtx& g, uint32_t cl) {
    T s = 0;
    for_quads<Q0, Q1>([&]<int V0, int V1, int V2_, int V3>() {
        if constexpr (CTRL) {
            if (!ctrl_ok<V0>(g, cl)) return;
        }
        if constexpr (KIND == OPK_PF_RZZ) {  // -Im sum over the odd-parity pair (negated below)
            cdot_part<true>(l.r[V1], l.i[V1], p.r[V1], p.i[V1], s);
            cdot_part<true>(l.r[V2_], l.i[V2_], p.r[V2_], p.i[V2_], s);
        } else if constexpr (KIND == OPK_RZZ) {
            cdot_part<true>(l.r[V0], l.i[V0], p.r[V0], p.i[V0], s);
            cdot_part<true>(l.r[V1], l.i[V1], -p.r[V1], -p.i[V1], s);
            cdot_part<true>(l.r[V2_], l.i[V2_], -p.r[V2_], -p.i[V2_], s);
            cdot_part<true>(l.r[V3], l.i[V3], p.r[V3], p.i[V3], s);
        } else {
            constexpr int sg = (KIND == OPK_RXX) ? 1 : -1;
            cdot_part<true>(l.r[V0], l.i[V0], T(sg) * p.r[V3], T(sg) * p.i[V3], s);
            cdot_part<true>(l.r[V3], l.i[V3], T(sg) * p.r[V0], T(sg) * p.i[V0], s);
            cdot_part<true>(l.r[V1], l.i[V1], p.r[V2_], p.i[V2_], s);
            cdot_part<true>(l.r[V2_], l.i[V2_], p.r[V1], p.i[V1], s);
        }
    });
    return KIND == OPK_PF_RZZ ? -s : T(0.5) * s;
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

// ---------------------------------------------------------------- parameterised dense ops
// Custom gates whose matrix is a host function of their parameters: a sample's row of the custom
// table holds, at the op's offset, [G][G^dagger][M] (D x D interleaved complex doubles, D = 2^NQ) with
// M = dG G^dagger the tap matrix; a tap-only op (a further parameter expression) holds [M].
template <typename T, int NQ>
struct DMat {
    T m[2 << (2 * NQ)];
};
template <typename T, int NQ>
__device__ __forceinline__ DMat<T, NQ> load_dmat(const double* p) {
    DMat<T, NQ> d;
#pragma unroll
    for (int i = 0; i < (2 << (2 * NQ)); i++) d.m[i] = T(p[i]);
    return d;
}

// Re <l|M|p> over the op's control subspace (M a dense 2x2 on register position Q).
template <typename T, int Q, bool CTRL>
__device__ __forceinline__ T dense1_tap(const Regs<T>& p, const Regs<T>& l, const GroupCtx& g, uint32_t cl,
                                        const T* __restrict__ m) {
    T s = 0;
    for_pairs<Q>([&]<int V, int W>() {
        T xr = p.r[V], xi = p.i[V], yr = p.r[W], yi = p.i[W];
        mat2(m, xr, xi, yr, yi);
        T u = 0;
        cdot_part<false>(l.r[V], l.i[V], xr, xi, u);
        cdot_part<false>(l.r[W], l.i[W], yr, yi, u);
        if constexpr (CTRL) s += ctrl_ok<V>(g, cl) ? u : T(0);
        else s += u;
    });
    return s;
}

// The same for a dense 4x4 on register positions (Q0, Q1) (quad order 00, 01, 10, 11, bit 0 <=> Q0).
template <typename T, int Q0, int Q1, bool CTRL>
__device__ __forceinline__ T dense2_tap(const Regs<T>& p, const Regs<T>& l, const GroupCtx& g, uint32_t cl,
                                        const T* __restrict__ m) {
    T s = 0;
    for_quads<Q0, Q1>([&]<int V0, int V1, int V2_, int V3>() {
        if constexpr (CTRL) {
            if (!ctrl_ok<V0>(g, cl)) return;
        }
        const int idx[4] = {V0, V1, V2_, V3};
#pragma unroll
        for (int row = 0; row < 4; row++) {
            T ur = 0, ui = 0;
#pragma unroll
            for (int col = 0; col < 4; col++) {
                const T mr = m[2 * (row * 4 + col)], mi = m[2 * (row * 4 + col) + 1];
                ur += mr * p.r[idx[col]] - mi * p.i[idx[col]];
                ui += mr * p.i[idx[col]] + mi * p.r[idx[col]];
            }
            cdot_part<false>(l.r[idx[row]], l.i[idx[row]], ur, ui, s);
        }
    });
    return s;
}

// One parameterised dense op.  Forward: G on p.  Back-sweep (ADJ): Re(tap) with M (returned), then
// G^dagger on p and l; a tap-only op (TAP_ONLY) changes nothing.
template <typename T, bool ADJ, int NQ, int Q0, int Q1, bool CTRL, bool TAP_ONLY>
__device__ __forceinline__ T pm_op(Regs<T>& p, Regs<T>& l, const GroupCtx& g, uint32_t cl, const double* m) {
    constexpr int DD = 2 << (2 * NQ);  // doubles of one matrix
    T tr = 0;
    if constexpr (ADJ) {
        const DMat<T, NQ> mt = load_dmat<T, NQ>(TAP_ONLY ? m : m + 2 * DD);
        if constexpr (NQ == 1) tr = dense1_tap<T, Q0, CTRL>(p, l, g, cl, mt.m);
        else tr = dense2_tap<T, Q0, Q1, CTRL>(p, l, g, cl, mt.m);
    }
    if constexpr (!TAP_ONLY) {
        const DMat<T, NQ> gm = load_dmat<T, NQ>(ADJ ? m + DD : m);
        if constexpr (NQ == 1) {
            one_target<T, OPK_M1, Q0, CTRL>(p, g, cl, T(0), T(0), gm.m);
            if constexpr (ADJ) one_target<T, OPK_M1, Q0, CTRL>(l, g, cl, T(0), T(0), gm.m);
        } else {
            two_target<T, OPK_M2, Q0, Q1, CTRL>(p, g, cl, T(0), T(0), gm.m);
            if constexpr (ADJ) two_target<T, OPK_M2, Q0, Q1, CTRL>(l, g, cl, T(0), T(0), gm.m);
        }
    }
    return tr;
}

// Calls f.template operator()<KIND, Q0, Q1>() for an op's kind and register positions: one
// switch per op, so the tap and both register sets are updated inside a single case body (one
// control-flow merge per op, which keeps the register allocator from copying the state).
template <bool FULL, typename F>
__device__ __forceinline__ void dispatch_op(int kind, int q0, int q1, F&& f) {
#define MQA_D1(KD)                                  \
    switch (q0) {                                   \
        case 0: f.template operator()<KD, 0, 0>(); break; \
        case 1: f.template operator()<KD, 1, 0>(); break; \
        case 2: f.template operator()<KD, 2, 0>(); break; \
        default: f.template operator()<KD, 3, 0>(); break; \
    }
#define MQA_D2(KD)                                   \
    switch (q0 * 4 + q1) {                           \
        case 1: f.template operator()<KD, 0, 1>(); break;  \
        case 2: f.template operator()<KD, 0, 2>(); break;  \
        case 3: f.template operator()<KD, 0, 3>(); break;  \
        case 4: f.template operator()<KD, 1, 0>(); break;  \
        case 6: f.template operator()<KD, 1, 2>(); break;  \
        case 7: f.template operator()<KD, 1, 3>(); break;  \
        case 8: f.template operator()<KD, 2, 0>(); break;  \
        case 9: f.template operator()<KD, 2, 1>(); break;  \
        case 11: f.template operator()<KD, 2, 3>(); break; \
        case 12: f.template operator()<KD, 3, 0>(); break; \
        case 13: f.template operator()<KD, 3, 1>(); break; \
        default: f.template operator()<KD, 3, 2>(); break; \
    }
    switch (kind) {
        case OPK_X: MQA_D1(OPK_X) break;
        case OPK_H: MQA_D1(OPK_H) break;
        case OPK_RX: MQA_D1(OPK_RX) break;
        case OPK_RY: MQA_D1(OPK_RY) break;
        case OPK_RZ: MQA_D1(OPK_RZ) break;
        case OPK_PS: MQA_D1(OPK_PS) break;
        case OPK_GP: MQA_D1(OPK_GP) break;
        case OPK_M1: MQA_D1(OPK_M1) break;
        case OPK_CX: MQA_D2(OPK_CX) break;
        case OPK_CZ: MQA_D2(OPK_CZ) break;
        case OPK_SWAP: MQA_D2(OPK_SWAP) break;
        case OPK_RZZ: MQA_D2(OPK_RZZ) break;
        case OPK_PF_RZZ: MQA_D2(OPK_PF_RZZ) break;
        case OPK_RXX: MQA_D2(OPK_RXX) break;
        case OPK_RYY: MQA_D2(OPK_RYY) break;
        case OPK_M2:
            if constexpr (FULL) MQA_D2(OPK_M2)
            break;
        default: break;
    }
#undef MQA_D1
#undef MQA_D2
}

template <int KIND>
constexpr bool kTwoTarget = KIND == OPK_CX || KIND == OPK_CZ || KIND == OPK_SWAP || KIND == OPK_RZZ || KIND == OPK_PF_RZZ ||
                            KIND == OPK_RXX || KIND == OPK_RYY || KIND == OPK_M2;

// G (forward) or G^dagger (back-sweep, angle already negated) on one register set.
template <typename T, int KIND, int Q0, int Q1, bool CTRL>
__device__ __forceinline__ void apply_body(Regs<T>& a, const GroupCtx& g, uint32_t cl, T c, T s,
                                           const T* __restrict__ m) {
    if constexpr (KIND == OPK_CX || KIND == OPK_CZ) {
        controlled_xz<T, KIND == OPK_CX, Q0, Q1, CTRL>(a, g, cl);  // Q0 target, Q1 bound control
    } else if constexpr (kTwoTarget<KIND>) {
        two_target<T, KIND, Q0, Q1, CTRL>(a, g, cl, c, s, m);
    } else {
        one_target<T, KIND, Q0, CTRL>(a, g, cl, c, s, m);
    }
}

// One op of a group.  ADJ: Re(tap) of the derivative, then G^dagger on psi and lambda.  Lean
// passes (FULL = false) never carry local control predicates: the host picks FULL when they do.
// The phase-free kinds run as the op they equal up to that phase: PF_RX / PF_RY as RX / RY with
// the (cos, sin) pair negated when cos < 0, PF_RZ as PS of the full angle, PF_GP as nothing.
template <typename T, bool ADJ, bool FULL>
__device__ __forceinline__ void run_op(const DevOp& op, Regs<T>& p, Regs<T>& l, const GroupCtx& g, uint64_t base,
                                       const T* __restrict__ pool, const double* __restrict__ ang_row,
                                       TapAcc<T>& acc, const double* __restrict__ pm_row = nullptr) {
    T tr = 0;  // Re(tap); zero when a non-local control is off (the op is then a no-op)
    int kind = op.kind;
    if constexpr (FULL) {  // parameterised dense ops run in FULL plans only, on register positions 0 (and 1)
        if (kind == OPK_PM || kind == OPK_PMTAP) {
            if ((base & op.ctrl_global) == op.ctrl_global) {
                const uint32_t cl = op.ctrl_local;
                const double* m = pm_row + op.mat;
                const bool tap_only = kind == OPK_PMTAP;
                if (op.pad0 == 2) {
                    if (cl) tr = tap_only ? pm_op<T, ADJ, 2, 0, 1, true, true>(p, l, g, cl, m)
                                          : pm_op<T, ADJ, 2, 0, 1, true, false>(p, l, g, cl, m);
                    else tr = tap_only ? pm_op<T, ADJ, 2, 0, 1, false, true>(p, l, g, cl, m)
                                       : pm_op<T, ADJ, 2, 0, 1, false, false>(p, l, g, cl, m);
                } else {
                    if (cl) tr = tap_only ? pm_op<T, ADJ, 1, 0, 0, true, true>(p, l, g, cl, m)
                                          : pm_op<T, ADJ, 1, 0, 0, true, false>(p, l, g, cl, m);
                    else tr = tap_only ? pm_op<T, ADJ, 1, 0, 0, false, true>(p, l, g, cl, m)
                                       : pm_op<T, ADJ, 1, 0, 0, false, false>(p, l, g, cl, m);
                }
            }
            if constexpr (ADJ) {
                if (op.tap >= 0) {
#pragma unroll
                    for (int j = kTapSlots<T> - 1; j > 0; j--) acc.r[j] = acc.r[j - 1];
                    acc.r[0] = tr;
                }
            }
            return;
        }
    }
    if ((base & op.ctrl_global) == op.ctrl_global && kind != OPK_PF_GP) {
        const uint32_t cl = op.ctrl_local;
        T c = 0, s = 0;
        if (kind >= OPK_RX) {  // (cos, sin) of the half angle, or of the angle for PS / GP
            c = T(ang_row[2 * op.slot]);
            s = T(ang_row[2 * op.slot + 1]);
            if (ADJ) s = -s;  // G^dagger = G(-angle)
        }
        if (kind == OPK_PF_RZ || kind == OPK_PF_RZZ) {  // full angle from the half angle
            const T c2 = c * c - s * s;
            s = T(2) * c * s;
            c = c2;
            if (kind == OPK_PF_RZ) kind = OPK_PS;
        } else if (kind == OPK_PF_RX || kind == OPK_PF_RY) {
            if (c < T(0)) c = -c, s = -s;
            kind = kind == OPK_PF_RX ? OPK_RX : OPK_RY;
        }
        const T* m = (op.mat >= 0) ? pool + op.mat : pool;
        dispatch_op<FULL>(kind, op.q0, op.q1, [&]<int KD, int Q0, int Q1>() {
            auto body = [&]<bool CT>() {
                if constexpr (ADJ && KD >= OPK_RX) {
                    if constexpr (kTwoTarget<KD>) tr = two_target_tap<T, KD, Q0, Q1, CT>(p, l, g, cl);
                    else tr = one_target_tap<T, KD, Q0, CT>(p, l, g, cl);
                }
                apply_body<T, KD, Q0, Q1, CT>(p, g, cl, c, s, m);
                if constexpr (ADJ) apply_body<T, KD, Q0, Q1, CT>(l, g, cl, c, s, m);
            };
            if constexpr (FULL) {
                if (cl) body.template operator()<true>();
                else body.template operator()<false>();
            } else {
                body.template operator()<false>();
            }
        });
    }
    if constexpr (ADJ) {
        if (op.tap >= 0) {
            // the group's taps arrive in order: shift them through the register file so that
            // every access uses a constant index (no dynamically indexed, scratch-backed array)
#pragma unroll
            for (int j = kTapSlots<T> - 1; j > 0; j--) acc.r[j] = acc.r[j - 1];
            acc.r[0] = tr;
        }
    }
}

// X on register position Q when on (a tile-uniform predicate: a non-local control): selects, no
// branch, so the register state has no control-flow merge (a merge in the middle of a group can
// make the compiler keep both register sets alive).
template <typename T, int Q>
__device__ __forceinline__ void x_if(Regs<T>& a, bool on) {
    for_pairs<Q>([&]<int V, int W>() {
        const T ar = a.r[V], ai = a.i[V], br = a.r[W], bi = a.i[W];
        a.r[V] = on ? br : ar, a.i[V] = on ? bi : ai;
        a.r[W] = on ? ar : br, a.i[W] = on ? ai : bi;
    });
}

// ---------------------------------------------------------------- phase-free rotations
// A plane rotation [[c, -s], [s, c]] as three shears (x -= t y; y += s x; x -= t y) with
// t = tan(angle / 2) = s / (1 + c): three FMAs per real pair instead of four products.  The
// coefficients come from the (cos, sin) pair with its sign normalised to cos >= 0 (the OPK_PF_*
// convention), so |t| <= 1.
template <typename T>
struct Shear {
    T t, s;
};
template <typename T>
__device__ __forceinline__ Shear<T> shear_of(T c, T s) {
    if (c < T(0)) c = -c, s = -s;
    Shear<T> h;
    if constexpr (sizeof(T) == 4) h.t = s * __builtin_amdgcn_rcpf(T(1) + c);
    else h.t = s / (T(1) + c);
    h.s = s;
    return h;
}
template <typename T>
__device__ __forceinline__ void shear_rotate(T& x, T& y, T t, T s) {
    x = fma(-t, y, x);
    y = fma(s, x, y);
    x = fma(-t, y, x);
}

// PF_RY: [[c, -s], [s, c]] on the real and on the imaginary parts of each pair.
// PF_RX: [[c, -is], [-is, c]] rotates (re a, im b) by -angle and (im a, re b) by +angle.
template <typename T, int KIND, int Q>
__device__ __forceinline__ void pf_rotation(Regs<T>& a, Shear<T> h) {
    for_pairs<Q>([&]<int V, int W>() {
        if constexpr (KIND == OPK_PF_RY) {
            shear_rotate(a.r[V], a.r[W], h.t, h.s);
            shear_rotate(a.i[V], a.i[W], h.t, h.s);
        } else {
            shear_rotate(a.r[V], a.i[W], -h.t, -h.s);
            shear_rotate(a.i[V], a.r[W], h.t, h.s);
        }
    });
}

// PF_RZ: cos / sin of the full angle from those of the half angle.
template <typename T>
__device__ __forceinline__ void double_angle(T& c, T& s) {
    const T c2 = c * c - s * s;
    s = T(2) * c * s;
    c = c2;
}

// ---------------------------------------------------------------- streaming accesses
// Every amplitude of a pass is read once and written once: non-temporal accesses keep the stream
// from allocating in the caches (on MI355X an in-place read-modify-write stream moves 6.57 TB/s this
// way against 5.91 TB/s with ordinary accesses).
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float4 ld_stream(const float4* p) {
    const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 ld_stream(const float2* p) {
    const nt_f2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f2*>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ double2 ld_stream(const double2* p) {
    const nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(float4 v, float4* p) {
    const nt_f4 w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<nt_f4*>(p));
}
__device__ __forceinline__ void st_stream(float2 v, float2* p) {
    const nt_f2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<nt_f2*>(p));
}
__device__ __forceinline__ void st_stream(double2 v, double2* p) {
    const nt_d2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<nt_d2*>(p));
}

// ---------------------------------------------------------------- tile I/O
// HBM <-> LDS transfer of one tile.  Every access of a thread is issued before the first is
// waited on (full unroll) so that enough bytes are in flight per CU to stream at HBM rate.
// complex64 moves two adjacent amplitudes (16 B) per HBM access: for u even, u and u + 1 are
// adjacent in HBM (the low run C >= 1 is contiguous); in LDS they sit at swz(u) and
// swz(u ^ 1) == swz(u) ^ 1, written and read as two 8-byte accesses (the swizzle may swap the
// pair inside its 16-byte slot, and a per-lane swap in registers costs more than the second
// LDS access).

// NT = false: cached loads (a state that every sample of a launch reads).
template <typename T, bool NT = true>
__device__ __forceinline__ void tile_load(V2<T>* __restrict__ lds, const V2<T>* __restrict__ g,
                                          const uint64_t* __restrict__ offtab, int C, uint32_t tid, uint32_t NT_) {
    const uint32_t cmask = (1u << C) - 1u;
    const uint32_t NT_threads = NT_;
    if constexpr (sizeof(T) == 4) {
        float4 buf[kAmpsPerThread / 2];
#pragma unroll
        for (int j = 0; j < kAmpsPerThread / 2; j++) {
            const uint32_t u = 2u * (tid + NT_threads * j);
            const float4* src = reinterpret_cast<const float4*>(g + ((u & cmask) | offtab[u >> C]));
            if constexpr (NT) buf[j] = ld_stream(src);
            else buf[j] = *src;
        }
#pragma unroll
        for (int j = 0; j < kAmpsPerThread / 2; j++) {
            const uint32_t a = swz<T>(2u * (tid + NT_threads * j));
            lds[a] = make_float2(buf[j].x, buf[j].y);
            lds[a ^ 1u] = make_float2(buf[j].z, buf[j].w);
        }
    } else {
        // complex128: two batches of 8 x 16 B, the same bytes in flight as complex64
#pragma unroll
        for (int h = 0; h < kAmpsPerThread; h += kAmpsPerThread / 2) {
            V2<T> buf[kAmpsPerThread / 2];
#pragma unroll
            for (int j = 0; j < kAmpsPerThread / 2; j++) {
                const uint32_t u = tid + NT_threads * (h + j);
                const V2<T>* src = g + ((u & cmask) | offtab[u >> C]);
                if constexpr (NT) buf[j] = ld_stream(src);
                else buf[j] = *src;
            }
#pragma unroll
            for (int j = 0; j < kAmpsPerThread / 2; j++) lds[swz<T>(tid + NT_threads * (h + j))] = buf[j];
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
            const V2<T> x = lds[a], y = lds[a ^ 1u];
            st_stream(make_float4(x.x, x.y, y.x, y.y), reinterpret_cast<float4*>(g + ((u & cmask) | offtab[u >> C])));
        }
    } else {
#pragma unroll
        for (int j = 0; j < kAmpsPerThread; j++) {
            const uint32_t u = tid + NT * j;
            st_stream(lds[swz<T>(u)], g + ((u & cmask) | offtab[u >> C]));
        }
    }
}

// ---------------------------------------------------------------- pass frame
// State shared by every group of a tile pass: LDS regions, this block's tile, wave geometry.
template <typename T>
struct TileFrame {
    V2<T>* t_psi;
    V2<T>* t_lam;
    uint64_t* offtab;
    double* lds_taps;
    V2<T>* g_psi;
    V2<T>* g_lam;
    uint64_t base;    // global index of tile amplitude 0
    uint64_t sample;
    uint64_t tile;
    int C;
    uint32_t NT, tid;
    int wave, n_waves, wave_width;
};

// Build the offset table and stage the tile (psi, and lambda for the back-sweep) in LDS.
// KC > 0: the tile width is the compile-time constant KC (generated passes), so the LDS layout
// and the thread count (2^(KC - 4), what every launch uses) fold into immediates.
template <typename T, bool ADJ, int KC = 0>
__device__ __forceinline__ TileFrame<T> tile_begin(const TilePassArgs& args, const DevPass& pass, unsigned char* smem,
                                                   uint64_t tile, uint64_t sample) {
    using V = V2<T>;
    TileFrame<T> f;
    const int K = KC > 0 ? KC : args.K;
    f.NT = KC > 0 ? (1u << (KC - kGroupWidth)) : blockDim.x;
    f.tid = threadIdx.x;
    f.C = pass.low_run;
    const int H = K - f.C;
    const LdsLayout lay = lds_layout(K, f.C, ADJ, int(sizeof(V)), 0);
    f.t_psi = reinterpret_cast<V*>(smem + lay.psi);
    f.t_lam = reinterpret_cast<V*>(smem + lay.lam);  // used by the adjoint pass only
    f.offtab = reinterpret_cast<uint64_t*>(smem + lay.offtab);
    f.lds_taps = reinterpret_cast<double*>(smem + lay.taps);
    f.sample = sample;
    f.tile = tile;
    f.base = pdep64(f.tile, pass.nonlocal_mask);
    for (uint32_t h = f.tid; h < (1u << H); h += f.NT) f.offtab[h] = pdep64(h, pass.local_hi_mask);
    __syncthreads();
    f.g_psi = static_cast<V*>(args.psi) + f.sample * args.stride + f.base;
    f.g_lam = ADJ ? (static_cast<V*>(args.lam) + f.sample * args.stride + f.base) : nullptr;
    if (!ADJ && args.src != nullptr) {  // the input state comes from src (broadcast: cached loads)
        const V* g_src = static_cast<const V*>(args.src) + f.sample * args.src_stride + f.base;
        if (args.src_stride == 0) tile_load<T, false>(f.t_psi, g_src, f.offtab, f.C, f.tid, f.NT);
        else tile_load<T>(f.t_psi, g_src, f.offtab, f.C, f.tid, f.NT);
    } else {
        tile_load<T>(f.t_psi, f.g_psi, f.offtab, f.C, f.tid, f.NT);
    }
    if constexpr (ADJ) tile_load<T>(f.t_lam, f.g_lam, f.offtab, f.C, f.tid, f.NT);
    __syncthreads();
    f.wave = int(f.tid >> 6);
    f.n_waves = int((f.NT + 63) >> 6);
    f.wave_width = f.NT < 64 ? int(f.NT) : 64;
    return f;
}

// The usual mapping: blockIdx.x = tile, blockIdx.y = sample.
template <typename T, bool ADJ, int KC = 0>
__device__ __forceinline__ TileFrame<T> tile_begin(const TilePassArgs& args, const DevPass& pass,
                                                   unsigned char* smem) {
    return tile_begin<T, ADJ, KC>(args, pass, smem, blockIdx.x, blockIdx.y);
}

// Launches over the tiles of several independent states (items of different sizes and
// schedules): tile_prefix[i] = first block of item i (ascending, n_items entries).
__device__ __forceinline__ int item_of_block(const uint64_t* __restrict__ tile_prefix, int n_items, uint64_t b) {
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_prefix[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Write the tile back and emit this tile's partial of every tap of the pass.
template <typename T, bool ADJ>
__device__ __forceinline__ void tile_end(const TilePassArgs& args, const DevPass& pass, const TileFrame<T>& f) {
    // launder tid so that the address arithmetic of the load phase is recomputed here rather
    // than kept live in registers through the whole group loop
    uint32_t tid_st = f.tid;
    asm volatile("" : "+v"(tid_st));
    if (!args.skip_store) {
        tile_store<T>(f.t_psi, f.g_psi, f.offtab, f.C, tid_st, f.NT);
        if constexpr (ADJ) tile_store<T>(f.t_lam, f.g_lam, f.offtab, f.C, tid_st, f.NT);
    }
    if constexpr (ADJ) {
        // one partial per (sample, tap id, tile); summed over tiles by k_reduce_rows1
        for (int t = int(f.tid); t < pass.n_taps; t += int(f.NT)) {
            double sr = 0;
            for (int w = 0; w < f.n_waves; w++) sr += f.lds_taps[t * f.n_waves + w];
            const int tap_id = args.tap_ids[pass.tap_begin + t];
            args.partials[(f.sample * uint64_t(args.n_tap_ids) + uint64_t(tap_id)) * args.n_tiles + f.tile] = sr;
        }
    }
}

// ---------------------------------------------------------------- single-buffer back-sweep passes
// Compiled back-sweep passes keep psi and lambda in registers between register groups and re-lay
// them out through ONE LDS tile buffer, one state after the other: a block then holds one tile of
// LDS instead of two (36 instead of 66 KB for 12-qubit complex64 tiles), so twice as many blocks
// share a CU, at the price of four barriers per group transition instead of one.
// KC > 0: compile-time tile width (generated passes); 0: args.K and blockDim.x (interpreted).
template <typename T, int KC>
__device__ __forceinline__ TileFrame<T> sb_begin(const TilePassArgs& args, const DevPass& pass, unsigned char* smem,
                                                 uint64_t tile, uint64_t sample) {
    using V = V2<T>;
    TileFrame<T> f;
    const int K = KC > 0 ? KC : args.K;
    f.NT = KC > 0 ? (1u << (KC - kGroupWidth)) : blockDim.x;
    f.tid = threadIdx.x;
    f.C = pass.low_run;
    const int H = K - f.C;
    const LdsLayout lay = lds_layout(K, f.C, false, int(sizeof(V)), 0);
    f.t_psi = reinterpret_cast<V*>(smem + lay.psi);
    f.t_lam = f.t_psi;
    f.offtab = reinterpret_cast<uint64_t*>(smem + lay.offtab);
    f.lds_taps = reinterpret_cast<double*>(smem + lay.taps);
    f.tile = tile;
    f.sample = sample;
    f.base = pdep64(f.tile, pass.nonlocal_mask);
    for (uint32_t h = f.tid; h < (1u << H); h += f.NT) f.offtab[h] = pdep64(h, pass.local_hi_mask);
    __syncthreads();
    f.g_psi = static_cast<V*>(args.psi) + f.sample * args.stride + f.base;
    f.g_lam = static_cast<V*>(args.lam) + f.sample * args.stride + f.base;
    f.wave = int(f.tid >> 6);
    f.n_waves = int((f.NT + 63) >> 6);
    f.wave_width = f.NT < 64 ? int(f.NT) : 64;
    return f;
}

template <typename T, int KC>
__device__ __forceinline__ TileFrame<T> sb_begin(const TilePassArgs& args, const DevPass& pass, unsigned char* smem) {
    return sb_begin<T, KC>(args, pass, smem, blockIdx.x, blockIdx.y);
}

// One state's registers <-> the LDS tile buffer in a group's layout.
template <typename T>
__device__ __forceinline__ void group_load1(const V2<T>* buf, uint32_t a0, const uint16_t* off, Regs<T>& r) {
#pragma unroll
    for (int v = 0; v < kAmpsPerThread; v++) {
        const V2<T> x = buf[a0 ^ off[v]];
        r.r[v] = x.x, r.i[v] = x.y;
    }
}
template <typename T>
__device__ __forceinline__ void group_store1(V2<T>* buf, uint32_t a0, const uint16_t* off, const Regs<T>& r) {
#pragma unroll
    for (int v = 0; v < kAmpsPerThread; v++) buf[a0 ^ off[v]] = V2<T>{r.r[v], r.i[v]};
}

// This tile's partial of every tap of the pass (after the barrier that follows the last group).
template <typename T>
__device__ __forceinline__ void taps_out(const TilePassArgs& args, const DevPass& pass, const TileFrame<T>& f) {
    for (int t = int(f.tid); t < pass.n_taps; t += int(f.NT)) {
        double sr = 0;
        for (int w = 0; w < f.n_waves; w++) sr += f.lds_taps[t * f.n_waves + w];
        const int tap_id = args.tap_ids[pass.tap_begin + t];
        args.partials[(f.sample * uint64_t(args.n_tap_ids) + uint64_t(tap_id)) * args.n_tiles + f.tile] = sr;
    }
}

// Register context of a group with positions pos (spos: ascending).
__device__ __forceinline__ GroupCtx group_ctx(uint32_t tid, const uint8_t* pos, const uint8_t* spos) {
    GroupCtx g;
    uint32_t ub = tid;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t q = spos[k];
        ub = ((ub >> q) << (q + 1)) | (ub & ((1u << q) - 1u));
    }
    g.ub = ub;
#pragma unroll
    for (int k = 0; k < 4; k++) g.pm[k] = 1u << pos[k];
    return g;
}

// LDS slot of register v is swz(ub) ^ off[v]: a per-thread base XOR a wave-uniform offset.
template <typename T, bool ADJ>
__device__ __forceinline__ void group_load(const TileFrame<T>& f, uint32_t a0, const uint16_t* off, Regs<T>& p,
                                           Regs<T>& l) {
#pragma unroll
    for (int v = 0; v < kAmpsPerThread; v++) {
        const V2<T> x = f.t_psi[a0 ^ off[v]];
        p.r[v] = x.x, p.i[v] = x.y;
        if constexpr (ADJ) {
            const V2<T> y = f.t_lam[a0 ^ off[v]];
            l.r[v] = y.x, l.i[v] = y.y;
        }
    }
}

template <typename T, bool ADJ>
__device__ __forceinline__ void group_store(const TileFrame<T>& f, uint32_t a0, const uint16_t* off,
                                            const Regs<T>& p, const Regs<T>& l) {
#pragma unroll
    for (int v = 0; v < kAmpsPerThread; v++) {
        f.t_psi[a0 ^ off[v]] = V2<T>{p.r[v], p.i[v]};
        if constexpr (ADJ) f.t_lam[a0 ^ off[v]] = V2<T>{l.r[v], l.i[v]};
    }
}

// Wave-reduce one Re(tap) and leave the wave's partial in the LDS tap table.
template <typename T>
__device__ __forceinline__ void tap_to_lds(const TileFrame<T>& f, T v, int tap) {
    v = (f.wave_width == 64) ? wave_sum64(v) : wave_sum(v, f.wave_width);
    if ((f.tid & 63) == 0) f.lds_taps[tap * f.n_waves + f.wave] = double(v);
}

// Lane swaps of two different values (no copies): permlane16_swap(a, b) leaves a's rows 0 / 2
// with b's rows 0 / 2 and the reverse, so a + b holds a's xor-16 sums in rows 0 / 2 and b's in
// rows 1 / 3; permlane32_swap(c, d) then gives c's halves summed in lanes 0-31 and d's in 32-63.
template <bool ROWS32>
__device__ __forceinline__ float swap_pair_sum(float a, float b) {
    const auto r = ROWS32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false)
                          : __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
template <bool ROWS32>
__device__ __forceinline__ double swap_pair_sum(double a, double b) {
    const unsigned long long x = __double_as_longlong(a), y = __double_as_longlong(b);
    const auto rl = ROWS32 ? __builtin_amdgcn_permlane32_swap(unsigned(x), unsigned(y), false, false)
                           : __builtin_amdgcn_permlane16_swap(unsigned(x), unsigned(y), false, false);
    const auto rh = ROWS32 ? __builtin_amdgcn_permlane32_swap(unsigned(x >> 32), unsigned(y >> 32), false, false)
                           : __builtin_amdgcn_permlane16_swap(unsigned(x >> 32), unsigned(y >> 32), false, false);
    return __longlong_as_double((static_cast<long long>(rh[0]) << 32) | rl[0]) +
           __longlong_as_double((static_cast<long long>(rh[1]) << 32) | rl[1]);
}

// Four Re(taps) of a full 64-lane wave reduced together: row sums by DPP, then two pairwise
// swaps (rows 0+1 / 2+3 for (t0, t1) and (t2, t3)) and one across halves.  Lanes 0, 16, 32 and
// 48 end with the totals of t0, t1, t2, t3 and write them to the LDS tap table (index < 0:
// no tap, its value is ignored).
template <typename T>
__device__ __forceinline__ void taps4_to_lds(const TileFrame<T>& f, T t0, int i0, T t1, int i1, T t2, int i2, T t3,
                                             int i3) {
    if (f.wave_width != 64) {
        if (i0 >= 0) tap_to_lds(f, t0, i0);
        if (i1 >= 0) tap_to_lds(f, t1, i1);
        if (i2 >= 0) tap_to_lds(f, t2, i2);
        if (i3 >= 0) tap_to_lds(f, t3, i3);
        return;
    }
    auto rows = [](T v) {
        v += dpp_mov<0xB1>(v);
        v += dpp_mov<0x4E>(v);
        v += dpp_mov<0x141>(v);
        return v + dpp_mov<0x140>(v);
    };
    const T c = swap_pair_sum<false>(rows(t0), rows(t1));
    const T d = swap_pair_sum<false>(rows(t2), rows(t3));
    const T e = swap_pair_sum<true>(c, d);
    const uint32_t lane = f.tid & 63;
    const int idx = lane == 0 ? i0 : lane == 16 ? i1 : lane == 32 ? i2 : lane == 48 ? i3 : -1;
    if (idx >= 0) f.lds_taps[idx * f.n_waves + f.wave] = double(e);
}


// ---------------------------------------------------------------- phase blocks (OPK_PHASE)
template <typename T>
__device__ __forceinline__ void cmul(T& ar, T& ai, T br, T bi) {
    const T r = ar * br - ai * bi;
    ai = ar * bi + ai * br;
    ar = r;
}
// a *= on ? b : 1
template <typename T>
__device__ __forceinline__ void cmul_if(T& ar, T& ai, T br, T bi, bool on) {
    const T r = ar * br - ai * bi, i = ar * bi + ai * br;
    ar = on ? r : ar;
    ai = on ? i : ai;
}

// Interpreted phase block (terms read at run time, uniform over the block): as in the generated
// passes, a term without register bits only updates a per-thread factor E and a term on one
// register bit j a factor F_j; register v's phase, E times the F_j of its bits, is applied once at
// the end.  Terms on two or more register bits are applied to the registers directly.  Back-sweep:
// conjugate phases; each term's tap value from the register sums of m_v = Im(conj(lambda_v) psi_v)
// (over all registers and per bit), which the block's phases leave unchanged; a gate's terms are
// consecutive and share its tap.
template <typename T, bool ADJ>
__device__ __forceinline__ void phase_block(const DevOp& op, Regs<T>& p, Regs<T>& l, const GroupCtx& g,
                                            const DevPhaseTerm* __restrict__ terms, const T* __restrict__ z,
                                            const TileFrame<T>& f) {
    const uint32_t posm = g.pm[0] | g.pm[1] | g.pm[2] | g.pm[3];
    T er = T(1), ei = T(0);
    T f0r = T(1), f0i = T(0), f1r = T(1), f1i = T(0), f2r = T(1), f2i = T(0), f3r = T(1), f3i = T(0);
    T S = T(0), M0 = T(0), M1 = T(0), M2 = T(0), M3 = T(0);  // register sums of m_v (back-sweep taps)
    if constexpr (ADJ) {
#pragma unroll
        for (int v = 0; v < kAmpsPerThread; v++) {
            const T m = l.r[v] * p.i[v] - l.i[v] * p.r[v];
            S += m;
            if (v & 1) M0 += m;
            if (v & 2) M1 += m;
            if (v & 4) M2 += m;
            if (v & 8) M3 += m;
        }
    }
    // F_j *= (ur, ui) for the block-uniform bit j, written on the named factors (a helper taking them
    // by reference lets the compiler pick one through a pointer, which puts them in scratch memory)
#define MQA_FMUL(J, UR, UI)                                                                      \
    switch (J) {                                                                                 \
        case 0: { const T r_ = f0r * (UR) - f0i * (UI); f0i = f0r * (UI) + f0i * (UR); f0r = r_; } break; \
        case 1: { const T r_ = f1r * (UR) - f1i * (UI); f1i = f1r * (UI) + f1i * (UR); f1r = r_; } break; \
        case 2: { const T r_ = f2r * (UR) - f2i * (UI); f2i = f2r * (UI) + f2i * (UR); f2r = r_; } break; \
        default: { const T r_ = f3r * (UR) - f3i * (UI); f3i = f3r * (UI) + f3i * (UR); f3r = r_; } break; \
    }
    T acc = T(0);
    int cur = -1;
    for (int k = 0; k < op.slot; k++) {
        const DevPhaseTerm t = terms[op.mat + k];
        const T zr = z[2 * t.coef];
        const T zi = ADJ ? -z[2 * t.coef + 1] : z[2 * t.coef + 1];
        const uint32_t R = ((t.ml & g.pm[0]) ? 1u : 0u) | ((t.ml & g.pm[1]) ? 2u : 0u) | ((t.ml & g.pm[2]) ? 4u : 0u) |
                           ((t.ml & g.pm[3]) ? 8u : 0u);
        const uint32_t H = t.ml & ~posm;
        const int nr = __popc(R);
        const int j0 = __ffs(int(R)) - 1;
        const uint64_t gm = f.base & t.mg;
        const bool is_and = t.kind == PHT_AND;
        const bool on = is_and ? (((g.ub & H) == H) && gm == t.mg)
                               : (((uint32_t(__popc(g.ub & H)) + uint32_t(__popcll(gm))) & 1u) != 0u);
        T val = T(0);  // this thread's tap value of the term
        if (nr == 0) {
            cmul_if(er, ei, zr, zi, on);
            val = on ? S : T(0);
        } else if (nr == 1) {
            const T Mj = j0 == 0 ? M0 : j0 == 1 ? M1 : j0 == 2 ? M2 : M3;
            if (is_and) {  // z on v_j = 1 where the term's other bits hold
                const T ur = on ? zr : T(1), ui = on ? zi : T(0);
                MQA_FMUL(j0, ur, ui)
                val = on ? Mj : T(0);
            } else {  // z^{v_j ^ s}: z on E and conj(z) on F_j when s = 1
                cmul_if(er, ei, zr, zi, on);
                const T ui = on ? -zi : zi;
                MQA_FMUL(j0, zr, ui)
                val = on ? S - Mj : Mj;
            }
        } else {
#pragma unroll
            for (int v = 0; v < kAmpsPerThread; v++) {
                const uint32_t rv = uint32_t(v) & R;
                const bool hit = is_and ? (on && rv == R) : (((uint32_t(__popc(rv)) + (on ? 1u : 0u)) & 1u) != 0u);
                if constexpr (ADJ) val += hit ? (l.r[v] * p.i[v] - l.i[v] * p.r[v]) : T(0);
                cmul_if(p.r[v], p.i[v], zr, zi, hit);
                if constexpr (ADJ) cmul_if(l.r[v], l.i[v], zr, zi, hit);
            }
        }
        if constexpr (ADJ) {
            if (t.tap >= 0) {
                if (t.tap != cur && cur >= 0) {
                    tap_to_lds(f, acc, cur);
                    acc = T(0);
                }
                cur = t.tap;
                acc += T(t.weight) * val;
            }
        }
    }
#undef MQA_FMUL
    if constexpr (ADJ) {
        if (cur >= 0) tap_to_lds(f, acc, cur);
    }
    // register v: E times the F_j of its bits
#pragma unroll
    for (int v = 0; v < kAmpsPerThread; v++) {
        T qr = er, qi = ei;
        if (v & 1) cmul(qr, qi, f0r, f0i);
        if (v & 2) cmul(qr, qi, f1r, f1i);
        if (v & 4) cmul(qr, qi, f2r, f2i);
        if (v & 8) cmul(qr, qi, f3r, f3i);
        cmul(p.r[v], p.i[v], qr, qi);
        if constexpr (ADJ) cmul(l.r[v], l.i[v], qr, qi);
    }
}

}  // namespace mqa
