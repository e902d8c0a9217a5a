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

#include "device_ops.h"
#include "launch.h"

namespace mqa {

// ---------------------------------------------------------------- the tile pass
// The interpreted pass: walks the group and op tables of any plan.  Circuits that run often
// are also compiled into straight-line kernels of the same structure (jit.cpp).
// Back-sweep without dense blocks: psi and lambda stay in registers between groups and share one
// LDS tile buffer (the single-buffer form of the generated passes, device_ops.h), so twice as many
// blocks fit a CU.
template <typename T, bool PHASE>
__device__ __forceinline__ void tile_pass_adj_one_buffer(const TilePassArgs& args, const DevPass& pass,
                                                         unsigned char* smem, uint64_t tile, uint64_t sample) {
    TileFrame<T> f = sb_begin<T, 0>(args, pass, smem, tile, sample);
    const double* ang_row = args.angles ? args.angles + 2 * f.sample * uint64_t(args.n_slots) : nullptr;
    Regs<T> p, l;
    DevGroup prev = args.groups[pass.group_begin];
    {
        const uint32_t a0 = swz<T>(group_ctx(f.tid, prev.pos, prev.spos).ub);
        tile_load<T>(f.t_psi, f.g_psi, f.offtab, f.C, f.tid, f.NT);
        __syncthreads();
        group_load1<T>(f.t_psi, a0, prev.off, p);
        __syncthreads();
        tile_load<T>(f.t_psi, f.g_lam, f.offtab, f.C, f.tid, f.NT);
        __syncthreads();
        group_load1<T>(f.t_psi, a0, prev.off, l);
    }
    for (int gi = 0; gi < pass.n_groups; gi++) {
        const DevGroup grp = args.groups[pass.group_begin + gi];
        const GroupCtx g = group_ctx(f.tid, grp.pos, grp.spos);
        if (gi > 0) {  // psi, then lambda, from the previous group's layout into this one's
            const uint32_t ap = swz<T>(group_ctx(f.tid, prev.pos, prev.spos).ub), an = swz<T>(g.ub);
            __syncthreads();
            group_store1<T>(f.t_psi, ap, prev.off, p);
            __syncthreads();
            group_load1<T>(f.t_psi, an, grp.off, p);
            __syncthreads();
            group_store1<T>(f.t_psi, ap, prev.off, l);
            __syncthreads();
            group_load1<T>(f.t_psi, an, grp.off, l);
        }
        TapAcc<T> acc;
#pragma unroll
        for (int j = 0; j < kTapSlots<T>; j++) acc.r[j] = T(0);
        for (int oi = 0; oi < grp.n_ops; oi++) {
            const DevOp op = args.ops[grp.op_begin + oi];
            if (PHASE && op.kind == OPK_PHASE)
                phase_block<T, true>(op, p, l, g, args.terms,
                                     static_cast<const T*>(args.phase) + 2 * f.sample * uint64_t(args.n_phase), f);
            else
                run_op<T, true, false>(op, p, l, g, f.base, static_cast<const T*>(args.pool), ang_row, acc);
        }
        if (grp.n_taps > 0) {  // one batched cross-lane reduction per group
            if (f.wave_width == 64) {
#pragma unroll
                for (int j = 0; j < kTapSlots<T>; j++) {
                    if (j < 4 || grp.n_taps > 4) acc.r[j] = wave_sum64(acc.r[j]);
                }
            } else {
#pragma unroll
                for (int j = 0; j < kTapSlots<T>; j++) acc.r[j] = wave_sum(acc.r[j], f.wave_width);
            }
            if ((f.tid & 63) == 0) {
#pragma unroll
                for (int j = 0; j < kTapSlots<T>; j++) {
                    if (j < grp.n_taps) {  // acc[j] holds the group's (n_taps - 1 - j)-th tap
                        const int t = grp.tap_begin + grp.n_taps - 1 - j;
                        f.lds_taps[t * f.n_waves + f.wave] = double(acc.r[j]);
                    }
                }
            }
        }
        prev = grp;
    }
    uint32_t tid_st = f.tid;
    asm volatile("" : "+v"(tid_st));
    const uint32_t ap = swz<T>(group_ctx(tid_st, prev.pos, prev.spos).ub);
    __syncthreads();
    if (!args.skip_store) {  // uniform: dead outputs of the last back-sweep pass are not written
        group_store1<T>(f.t_psi, ap, prev.off, p);
        __syncthreads();
        tile_store<T>(f.t_psi, f.g_psi, f.offtab, f.C, tid_st, f.NT);
        __syncthreads();
        group_store1<T>(f.t_psi, ap, prev.off, l);
        __syncthreads();
        tile_store<T>(f.t_psi, f.g_lam, f.offtab, f.C, tid_st, f.NT);
    }
    taps_out<T>(args, pass, f);
}

// PHASE: the plan has phase blocks (an instantiation of its own, so that plans without them keep
// the register allocation of the plain kernels).
template <typename T, bool ADJ, bool FULL, bool PHASE>
__device__ __forceinline__ void tile_pass_body(const TilePassArgs& args, unsigned char* smem, uint64_t tile,
                                               uint64_t sample) {
    const DevPass pass = args.passes[args.pass_index];
    if constexpr (ADJ && !FULL) {
        tile_pass_adj_one_buffer<T, PHASE>(args, pass, smem, tile, sample);
        return;
    }
    TileFrame<T> f = tile_begin<T, ADJ>(args, pass, smem, tile, sample);
    const double* ang_row = args.angles ? args.angles + 2 * f.sample * uint64_t(args.n_slots) : nullptr;
    const double* pm_row = args.custom ? args.custom + f.sample * args.custom_stride : nullptr;

    for (int gi = 0; gi < pass.n_groups; gi++) {
        const DevGroup grp = args.groups[pass.group_begin + gi];
        const GroupCtx g = group_ctx(f.tid, grp.pos, grp.spos);
        const uint32_t a0 = swz<T>(g.ub);
        Regs<T> p, l;
        group_load<T, ADJ>(f, a0, grp.off, p, l);

        TapAcc<T> acc;
        if constexpr (ADJ) {
#pragma unroll
            for (int j = 0; j < kTapSlots<T>; j++) acc.r[j] = T(0);
        }
        const DevOp op0 = args.ops[grp.op_begin];
        int first = 0;
        if (FULL && (op0.kind == OPK_M3 || op0.kind == OPK_M4)) {  // a dense block opens its group
            if ((f.base & op0.ctrl_global) == op0.ctrl_global) {
                const T* m = static_cast<const T*>(args.pool) + op0.mat;
                if (op0.kind == OPK_M3) {
                    dense_from_lds<T, 3>(p, g, op0.ctrl_local, m, f.t_psi);
                    if constexpr (ADJ) dense_from_lds<T, 3>(l, g, op0.ctrl_local, m, f.t_lam);
                } else {
                    dense_from_lds<T, 4>(p, g, op0.ctrl_local, m, f.t_psi);
                    if constexpr (ADJ) dense_from_lds<T, 4>(l, g, op0.ctrl_local, m, f.t_lam);
                }
            }
            first = 1;  // phase blocks may follow it
        }
        for (int oi = first; oi < grp.n_ops; oi++) {
            const DevOp op = args.ops[grp.op_begin + oi];
            if (PHASE && op.kind == OPK_PHASE)
                phase_block<T, ADJ>(op, p, l, g, args.terms,
                                    static_cast<const T*>(args.phase) + 2 * f.sample * uint64_t(args.n_phase), f);
            else
                run_op<T, ADJ, FULL>(op, p, l, g, f.base, static_cast<const T*>(args.pool), ang_row, acc, pm_row);
        }
        if constexpr (ADJ) {
            if (grp.n_taps > 0) {  // one batched cross-lane reduction per group
                if (f.wave_width == 64) {
#pragma unroll
                    for (int j = 0; j < kTapSlots<T>; j++) {
                        if (j < 4 || grp.n_taps > 4) acc.r[j] = wave_sum64(acc.r[j]);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < kTapSlots<T>; j++) acc.r[j] = wave_sum(acc.r[j], f.wave_width);
                }
                if ((f.tid & 63) == 0) {
#pragma unroll
                    for (int j = 0; j < kTapSlots<T>; j++) {
                        if (j < grp.n_taps) {  // acc[j] holds the group's (n_taps - 1 - j)-th tap
                            const int t = grp.tap_begin + grp.n_taps - 1 - j;
                            f.lds_taps[t * f.n_waves + f.wave] = double(acc.r[j]);
                        }
                    }
                }
            }
        }
        group_store<T, ADJ>(f, a0, grp.off, p, l);
        __syncthreads();
    }
    tile_end<T, ADJ>(args, pass, f);
}

template <typename T, bool ADJ, bool FULL, bool PHASE>
__global__ void __launch_bounds__(512) k_tile_pass(TilePassArgs args) {
    extern __shared__ __align__(16) unsigned char smem[];
    tile_pass_body<T, ADJ, FULL, PHASE>(args, smem, blockIdx.x, blockIdx.y);
}

// One pass step over several independent single-sample states (uneven circuits): block b runs
// tile b - tile_prefix[i] of item i.  Each item carries its own schedule, buffers and pass index.
// complex64: at least three waves per SIMD (at most 168 VGPRs); left alone, the interpreter's
// back-sweep takes 187 and two waves per SIMD, a third of what the LDS of its tiles allows.
template <typename T>
constexpr int kItemsMinWaves = sizeof(T) == 4 ? 3 : 1;

template <typename T, bool ADJ, bool FULL, bool PHASE>
__global__ void __launch_bounds__(512, kItemsMinWaves<T>) k_tile_pass_items(const TilePassArgs* __restrict__ items,
                                                         const uint64_t* __restrict__ tile_prefix, int n_items) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int i = item_of_block(tile_prefix, n_items, blockIdx.x);
    const TilePassArgs args = items[i];
    tile_pass_body<T, ADJ, FULL, PHASE>(args, smem, uint64_t(blockIdx.x) - tile_prefix[i], 0);
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
__device__ __forceinline__ void ham_tile_body(const HamTileArgs& args, unsigned char* smem, uint64_t tile,
                                              uint64_t sample) {
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
        } else if (grp.kind == HAMW_WHT) {
            const uint32_t tsrc = ((2u * tid) ^ grp.flipl) & tmask;
            // register part of the flip, in register-mask coordinates
            const uint32_t freg = (grp.flipl & 1u) | (((grp.flipl >> JS) & 7u) << 1);
            const_ptr<DevHamTileTerm> tc = as_const(tm);
            T wr[NA], wi[NA];
            int t = 0;
#pragma unroll
            for (int m = 0; m < NA; m++) {  // W[m]: terms of register mask m, tile and thread signs
                T sr = 0, si = 0;
                const int end = grp.seg_end[m];
#pragma unroll 4
                for (; t < end; t++) {  // unrolled: four terms' scalar loads share one wait
                    const uint32_t neg =
                        (uint32_t(__popcll(base & tc[t].yz_nl)) + __popc(tsrc & tc[t].yzl)) & 1u;
                    sr += flip_sign(T(tc[t].coef_re), neg);
                    si += flip_sign(T(tc[t].coef_im), neg);
                }
                const uint32_t fneg = uint32_t(__popc(freg & uint32_t(m))) & 1u;
                wr[m] = flip_sign(sr, fneg);
                wi[m] = flip_sign(si, fneg);
            }
            // weight of register J = sum_m W[m] (-1)^{popcount(J & m)}: Walsh-Hadamard transform
#pragma unroll
            for (int b = 1; b < NA; b <<= 1) {
#pragma unroll
                for (int m = 0; m < NA; m++) {
                    if (m & b) continue;
                    const T xr = wr[m], xi = wi[m];
                    wr[m] = xr + wr[m | b], wi[m] = xi + wi[m | b];
                    wr[m | b] = xr - wr[m | b], wi[m | b] = xi - wi[m | b];
                }
            }
#pragma unroll
            for (int J = 0; J < NA; J++) {
                const V x = t_ket[a0 ^ swz<T>(jb(J) ^ grp.flipl)];
                ar[J] += wr[J] * x.x - wi[J] * x.y;
                ai[J] += wr[J] * x.y + wi[J] * x.x;
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
            block_sum2(fr, fi, args.partial + 2 * (sample * args.n_tiles + tile));
        }
    }
}

template <typename T, bool FIRST, bool LAST>
__global__ void __launch_bounds__(ham_tile_threads<T>()) k_ham_tile(HamTileArgs args) {
    extern __shared__ __align__(16) unsigned char smem[];
    ham_tile_body<T, FIRST, LAST>(args, smem, blockIdx.x, blockIdx.y);
}

// One Hamiltonian pass step over several independent single-sample states (see k_tile_pass_items).
template <typename T, bool FIRST, bool LAST>
__global__ void __launch_bounds__(ham_tile_threads<T>()) k_ham_tile_items(const HamTileArgs* __restrict__ items,
                                                                       const uint64_t* __restrict__ tile_prefix,
                                                                       int n_items) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int i = item_of_block(tile_prefix, n_items, blockIdx.x);
    const HamTileArgs args = items[i];
    ham_tile_body<T, FIRST, LAST>(args, smem, uint64_t(blockIdx.x) - tile_prefix[i], 0);
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

// Coefficient table of the compiled passes: per (sample, slot), the slot's CoefForm of its
// (cos, sin) in the real type T, with the same device functions the passes would otherwise run
// in every lane (so the values are the ones those would compute).
template <typename T>
__global__ void __launch_bounds__(256) k_coefs(const double* __restrict__ ang, const uint8_t* __restrict__ forms,
                                               int n_forms, int n_slots, uint64_t total, T* __restrict__ out) {
    const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int slot = int(e % uint64_t(n_slots));
    const uint8_t form = slot < n_forms ? forms[slot] : uint8_t(COEF_NONE);
    T c = T(ang[2 * e]), s = T(ang[2 * e + 1]);
    if (form == COEF_SHEAR) {
        const Shear<T> h = shear_of(c, s);
        c = h.t, s = h.s;
    } else if (form == COEF_DOUBLE) {
        double_angle(c, s);
    }
    out[2 * e] = c;
    out[2 * e + 1] = s;
}

// Phase table of a plan's phase blocks: per (sample, term), e^{i (k alpha + b)} in the real type,
// alpha the angle the sample's angle table holds for the term's slot (PhaseCoef), evaluated in
// double from its (cos, sin) and rounded once.
template <typename T>
__global__ void __launch_bounds__(256) k_phase_coefs(const double* __restrict__ ang, const PhaseCoef* __restrict__ pc,
                                                     int n_terms, int n_slots, uint64_t total, T* __restrict__ out) {
    const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const uint64_t b = e / uint64_t(n_terms);
    const PhaseCoef c = pc[e % uint64_t(n_terms)];
    double x = 1.0, y = 0.0;
    if (c.slot >= 0 && ang != nullptr) {
        const double ca = ang[2 * (b * uint64_t(n_slots) + uint64_t(c.slot))];
        const double sa = ang[2 * (b * uint64_t(n_slots) + uint64_t(c.slot)) + 1];
        const int k = c.k < 0 ? -c.k : c.k;
        if (k == 1) x = ca, y = sa;
        else if (k == 2) x = ca * ca - sa * sa, y = 2.0 * ca * sa;
        if (c.k < 0) y = -y;
    }
    out[2 * e] = T(x * c.cos_b - y * c.sin_b);
    out[2 * e + 1] = T(x * c.sin_b + y * c.cos_b);
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

// base[offset[i]] = 1 for every item (states already zeroed): |0...0> for a set of items
template <typename T>
__global__ void __launch_bounds__(256) k_unit_items(V2<T>* base, const uint64_t* __restrict__ offset, int n) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) base[offset[i]] = V2<T>{T(1), T(0)};
}

// out[k] = sum of in[begin[k] .. begin[k] + len[k]) in double, one wave per segment, fixed order
__global__ void __launch_bounds__(256) k_reduce_segments(const double* __restrict__ in,
                                                         const uint64_t* __restrict__ begin,
                                                         const uint64_t* __restrict__ len, double* __restrict__ out,
                                                         uint64_t n_segments) {
    const uint64_t k = uint64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (k >= n_segments) return;
    const int lane = threadIdx.x & 63;
    double a = 0;
    for (uint64_t j = lane; j < len[k]; j += 64) a += in[begin[k] + j];
    a = wave_sum(a, 64);
    if (lane == 0) out[k] = a;
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
size_t interpreted_pass_lds_bytes(int dtype, bool adj, bool full, int K, int C, int max_taps) {
    return tile_pass_lds_bytes(dtype, adj && full, K, C, max_taps);  // one tile buffer unless dense blocks
}

template <typename T, bool ADJ, bool FULL, bool PHASE>
static hipError_t launch_tp(const TilePassArgs& a, uint64_t batch, size_t lds, hipStream_t s) {
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_pass<T, ADJ, FULL, PHASE>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(kLdsLimit));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    dim3 grid(uint32_t(a.n_tiles), uint32_t(batch));
    dim3 block(1u << (a.K - kGroupWidth));
    hipLaunchKernelGGL((k_tile_pass<T, ADJ, FULL, PHASE>), grid, block, lds, s, a);
    return hipGetLastError();
}

template <typename T, bool ADJ>
static hipError_t launch_tp_m(bool full, bool phase, const TilePassArgs& a, uint64_t batch, size_t lds, hipStream_t s) {
    if (full) return phase ? launch_tp<T, ADJ, true, true>(a, batch, lds, s) : launch_tp<T, ADJ, true, false>(a, batch, lds, s);
    return phase ? launch_tp<T, ADJ, false, true>(a, batch, lds, s) : launch_tp<T, ADJ, false, false>(a, batch, lds, s);
}

hipError_t launch_tile_pass(int dtype, bool adj, bool full, bool phase, const TilePassArgs& a, uint64_t batch, int C,
                            int max_taps, hipStream_t s) {
    size_t lds = interpreted_pass_lds_bytes(dtype, adj, full, a.K, C, max_taps);
    if (lds > kLdsLimit) return hipErrorInvalidValue;
    if (dtype == 0)
        return adj ? launch_tp_m<float, true>(full, phase, a, batch, lds, s)
                   : launch_tp_m<float, false>(full, phase, a, batch, lds, s);
    return adj ? launch_tp_m<double, true>(full, phase, a, batch, lds, s)
               : launch_tp_m<double, false>(full, phase, a, batch, lds, s);
}

template <typename T, bool ADJ, bool FULL, bool PHASE>
static hipError_t launch_tpi(const TilePassArgs* items, const uint64_t* prefix, int n_items, uint64_t blocks,
                             uint32_t threads, size_t lds, hipStream_t s) {
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_pass_items<T, ADJ, FULL, PHASE>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(kLdsLimit));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((k_tile_pass_items<T, ADJ, FULL, PHASE>), dim3(uint32_t(blocks)), dim3(threads), lds, s, items,
                       prefix, n_items);
    return hipGetLastError();
}

template <typename T, bool ADJ>
static hipError_t launch_tpi_m(bool full, bool phase, const TilePassArgs* items, const uint64_t* prefix, int n_items,
                               uint64_t blocks, uint32_t threads, size_t lds, hipStream_t s) {
    if (full)
        return phase ? launch_tpi<T, ADJ, true, true>(items, prefix, n_items, blocks, threads, lds, s)
                     : launch_tpi<T, ADJ, true, false>(items, prefix, n_items, blocks, threads, lds, s);
    return phase ? launch_tpi<T, ADJ, false, true>(items, prefix, n_items, blocks, threads, lds, s)
                 : launch_tpi<T, ADJ, false, false>(items, prefix, n_items, blocks, threads, lds, s);
}

hipError_t launch_tile_pass_items(int dtype, bool adj, bool full, bool phase, const TilePassArgs* items,
                                  const uint64_t* prefix, int n_items, uint64_t blocks, int K, size_t lds,
                                  hipStream_t s) {
    if (lds > kLdsLimit || blocks == 0 || blocks > 0x7fffffffull || K < kMinTileQubits || K > kMaxTileQubits)
        return hipErrorInvalidValue;
    const uint32_t threads = 1u << (K - kGroupWidth);
    if (dtype == 0)
        return adj ? launch_tpi_m<float, true>(full, phase, items, prefix, n_items, blocks, threads, lds, s)
                   : launch_tpi_m<float, false>(full, phase, items, prefix, n_items, blocks, threads, lds, s);
    return adj ? launch_tpi_m<double, true>(full, phase, items, prefix, n_items, blocks, threads, lds, s)
               : launch_tpi_m<double, false>(full, phase, items, prefix, n_items, blocks, threads, lds, s);
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

template <typename T, bool FIRST, bool LAST>
static hipError_t launch_hti(const HamTileArgs* items, const uint64_t* prefix, int n_items, uint64_t blocks,
                             uint32_t threads, size_t lds, hipStream_t s) {
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ham_tile_items<T, FIRST, LAST>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           int(kLdsLimit - kBlockSumLds));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    if (threads > uint32_t(ham_tile_threads<T>())) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_ham_tile_items<T, FIRST, LAST>), dim3(uint32_t(blocks)), dim3(threads), lds, s, items,
                       prefix, n_items);
    return hipGetLastError();
}

hipError_t launch_ham_tile_items(int dtype, bool first, bool last, const HamTileArgs* items, const uint64_t* prefix,
                                 int n_items, uint64_t blocks, int K, int C, hipStream_t s) {
    const size_t amp = dtype == 0 ? 8 : 16;
    const size_t lds = lds_layout(K, C, false, int(amp), 0).total;
    if (K < kMinTileQubits || K > kMaxHamTileQubits || lds + kBlockSumLds > kLdsLimit || blocks == 0 ||
        blocks > 0x7fffffffull)
        return hipErrorInvalidValue;
    const uint32_t threads = 1u << (K - kGroupWidth);
    if (dtype == 0) {
        if (first) return last ? launch_hti<float, true, true>(items, prefix, n_items, blocks, threads, lds, s)
                               : launch_hti<float, true, false>(items, prefix, n_items, blocks, threads, lds, s);
        return last ? launch_hti<float, false, true>(items, prefix, n_items, blocks, threads, lds, s)
                    : launch_hti<float, false, false>(items, prefix, n_items, blocks, threads, lds, s);
    }
    if (first) return last ? launch_hti<double, true, true>(items, prefix, n_items, blocks, threads, lds, s)
                           : launch_hti<double, true, false>(items, prefix, n_items, blocks, threads, lds, s);
    return last ? launch_hti<double, false, true>(items, prefix, n_items, blocks, threads, lds, s)
                : launch_hti<double, false, false>(items, prefix, n_items, blocks, threads, lds, s);
}

hipError_t launch_unit_items(int dtype, void* base, const uint64_t* offset, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const uint32_t blocks = uint32_t((n + 255) / 256);
    if (dtype == 0) hipLaunchKernelGGL(k_unit_items<float>, dim3(blocks), dim3(256), 0, s, static_cast<float2*>(base), offset, n);
    else hipLaunchKernelGGL(k_unit_items<double>, dim3(blocks), dim3(256), 0, s, static_cast<double2*>(base), offset, n);
    return hipGetLastError();
}

hipError_t launch_reduce_segments(const double* in, const uint64_t* begin, const uint64_t* len, double* out,
                                  uint64_t n_segments, hipStream_t s) {
    if (n_segments == 0) return hipSuccess;
    const uint32_t blocks = uint32_t((n_segments + 3) / 4);
    hipLaunchKernelGGL(k_reduce_segments, dim3(blocks), dim3(256), 0, s, in, begin, len, out, n_segments);
    return hipGetLastError();
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

hipError_t launch_coefs(int dtype, const double* angles, const uint8_t* forms, int n_forms, int n_slots, uint64_t batch,
                        void* out, hipStream_t s) {
    const uint64_t total = batch * uint64_t(n_slots);
    if (total == 0) return hipSuccess;
    const dim3 grid(uint32_t((total + 255) / 256));
    if (dtype == 0)
        hipLaunchKernelGGL(k_coefs<float>, grid, dim3(256), 0, s, angles, forms, n_forms, n_slots, total,
                           static_cast<float*>(out));
    else
        hipLaunchKernelGGL(k_coefs<double>, grid, dim3(256), 0, s, angles, forms, n_forms, n_slots, total,
                           static_cast<double*>(out));
    return hipGetLastError();
}

hipError_t launch_phase_coefs(int dtype, const double* angles, const PhaseCoef* pc, int n_terms, int n_slots,
                              uint64_t batch, void* out, hipStream_t s) {
    const uint64_t total = batch * uint64_t(n_terms);
    if (total == 0) return hipSuccess;
    const dim3 grid(uint32_t((total + 255) / 256));
    if (dtype == 0)
        hipLaunchKernelGGL(k_phase_coefs<float>, grid, dim3(256), 0, s, angles, pc, n_terms, n_slots, total,
                           static_cast<float*>(out));
    else
        hipLaunchKernelGGL(k_phase_coefs<double>, grid, dim3(256), 0, s, angles, pc, n_terms, n_slots, total,
                           static_cast<double*>(out));
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
