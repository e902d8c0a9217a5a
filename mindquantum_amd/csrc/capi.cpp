// capi.cpp - the C ABI (include/mqamd.h): simulator handles, device workspaces, cached
// tile-pass plans and the orchestration of forward sweeps, Hamiltonian application,
// expectation reductions and the batched adjoint gradient.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mqamd.h"
#include "aux_kernels.h"
#include "dense_sweep.h"
#include "jit.h"
#include "launch.h"
#include "program.h"
#include "schedule.h"

namespace mqa {
namespace {

thread_local std::string g_last_error;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        throw Error(MQA_ERR_OOM, std::string(what) + ": out of device memory");
    }
    throw Error(MQA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return MQA_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return MQA_ERR_ARG;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return MQA_ERR_OOM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return MQA_ERR_RUNTIME;
    }
}

// ---------------------------------------------------------------- options
struct Options {
    std::atomic<int64_t> tile_c64{12}, tile_c128{11}, adj_tile_c64{12}, adj_tile_c128{11};
    std::atomic<int64_t> group_width{4}, lowq_c64{5}, lowq_c128{4}, max_ops{0};
    std::atomic<int64_t> budget{0};  // bytes of HBM a batched call may use (0: 85% of free memory)
    std::atomic<int64_t> profile{0};  // 1: time kernels with HIP events
    std::atomic<int64_t> ham_tile_c64{14}, ham_tile_c128{13}, ham_engine{1};  // ham_engine 0: gather kernel only
    // per-circuit compiled passes: 0 off, 1 compile in the background and switch over when
    // ready, 2 compile before the first launch; only for launches of >= jit_min_amps amplitudes
    std::atomic<int64_t> jit{1}, jit_min_amps{int64_t(1) << 24};
    // ... and for smaller launches once a plan has run this many times (0: never)
    std::atomic<int64_t> jit_min_runs{8};
    // forward-state checkpoints of the batched gradient (0: none; the back-sweep uncomputes psi)
    std::atomic<int64_t> checkpoints{0};
    // 1: hermitian gradients run phase-free plans (uncontrolled rotations drop global phases)
    std::atomic<int64_t> phase_free{1};
    // 1: compiled back-sweep passes keep psi and lambda in registers and share one LDS tile
    // buffer (twice the resident blocks per CU)
    std::atomic<int64_t> jit_single_buffer{1};
    // 1: complex64 5-qubit dense sweeps run on the FP32 matrix cores
    std::atomic<int64_t> dense_mfma{1};
    // 1: batched gradients and gate application whose launches run compiled passes merge RZ, PS and
    // RZZ gates into phase blocks (OPK_PHASE); 2: every launch of those (interpreted ones too)
    std::atomic<int64_t> phase_ops{1};
    // uneven batches (mqa_grad_items): items of at least this many qubits run one by one on the
    // per-circuit compiled passes (with "jit" on) instead of in the shared interpreted launches; 0:
    // every item shares them
    std::atomic<int64_t> items_jit_qubits{23};
};
Options g_opt;

struct OptEntry {
    const char* key;
    std::atomic<int64_t>* v;
    int64_t def;
};
std::vector<OptEntry> opt_table() {
    return {{"tile_qubits_c64", &g_opt.tile_c64, 12},     {"tile_qubits_c128", &g_opt.tile_c128, 11},
            {"adj_tile_qubits_c64", &g_opt.adj_tile_c64, 12}, {"adj_tile_qubits_c128", &g_opt.adj_tile_c128, 11},
            {"group_width", &g_opt.group_width, 4},       {"lowq_c64", &g_opt.lowq_c64, 5},
            {"lowq_c128", &g_opt.lowq_c128, 4},           {"max_ops_per_pass", &g_opt.max_ops, 0},
            {"hbm_budget_bytes", &g_opt.budget, 0},      {"profile", &g_opt.profile, 0},
            {"ham_tile_qubits_c64", &g_opt.ham_tile_c64, 14}, {"ham_tile_qubits_c128", &g_opt.ham_tile_c128, 13},
            {"ham_tile_engine", &g_opt.ham_engine, 1}, {"jit", &g_opt.jit, 1},
            {"jit_min_amps", &g_opt.jit_min_amps, int64_t(1) << 24}, {"jit_min_runs", &g_opt.jit_min_runs, 8},
            {"checkpoints", &g_opt.checkpoints, 0},
            {"phase_free", &g_opt.phase_free, 1}, {"jit_single_buffer", &g_opt.jit_single_buffer, 1},
            {"dense_mfma", &g_opt.dense_mfma, 1}, {"phase_ops", &g_opt.phase_ops, 1},
            {"items_jit_qubits", &g_opt.items_jit_qubits, 23}};
}

SchedOptions sched_options(int dtype, bool adjoint) {
    SchedOptions o;
    o.tile_qubits = int(dtype == MQA_C64 ? (adjoint ? g_opt.adj_tile_c64 : g_opt.tile_c64)
                                         : (adjoint ? g_opt.adj_tile_c128 : g_opt.tile_c128));
    o.low_run = int(dtype == MQA_C64 ? g_opt.lowq_c64 : g_opt.lowq_c128);
    o.group_width = int(g_opt.group_width);
    o.max_ops_per_pass = int(g_opt.max_ops);
    o.taps_per_group = 8;  // register budget of the adjoint pass (real tap accumulators)
    o.amp_bytes = dtype == MQA_C64 ? 8 : 16;
    return o;
}

// ---------------------------------------------------------------- device memory
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t n) {
        if (n <= bytes) return;
        release();
        if (n == 0) return;
        hip_check(hipMalloc(&p, n), "hipMalloc");
        bytes = n;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

template <typename T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
    size_t n = std::max<size_t>(v.size(), 1) * sizeof(T);
    b.ensure(n);
    if (!v.empty()) hip_check(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}

int current_device() {
    int d = 0;
    hip_check(hipGetDevice(&d), "hipGetDevice");
    return d;
}

// ---------------------------------------------------------------- device copies of plans
struct DevPlan {
    Plan plan;
    DevBuf passes, groups, ops, pool, tap_ids;
    DevBuf coef_forms;  // CoefForm per angle slot (the compiled passes' coefficient table)
    DevBuf terms, phase_coefs;  // phase blocks: DevPhaseTerm and PhaseCoef arrays
    int n_coef_forms = 0;
    bool full = false;
    bool phase = false;  // the plan has phase blocks (the interpreted pass's PHASE instantiation)
    std::mutex jit_mu;
    std::unique_ptr<JitPlan> jit;     // compiled passes (created on first large launch)
    std::unique_ptr<JitPlan> jit_sb;  // the same with one LDS tile buffer (option "jit_single_buffer")
    std::atomic<int64_t> runs{0};     // run_plan calls so far (small plans compile once they recur)
};

// ---------------------------------------------------------------- kernel timing
struct KernelStats {
    int64_t launches = 0;
    double ms = 0;
    double bytes = 0;
};
std::mutex g_prof_mu;
std::map<std::string, KernelStats> g_prof;

// Event pairs recorded on one stream; resolved after the stream is synchronised.
struct Timer {
    struct Pending {
        hipEvent_t a, b;
        std::string name;
        double bytes;
    };
    std::vector<hipEvent_t> pool;
    std::vector<Pending> pending;
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }
    void resolve() {
        if (pending.empty()) return;
        std::lock_guard<std::mutex> lk(g_prof_mu);
        for (auto& p : pending) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                auto& st = g_prof[p.name];
                st.launches++;
                st.ms += ms;
                st.bytes += p.bytes;
            }
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    ~Timer() {
        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace
}  // namespace mqa

// ---------------------------------------------------------------- circuit
struct mqa_circuit {
    std::vector<mqa::CGate> gates;
    int n_params = 0;
    int n_slots = 0;
    int max_qubit = -1;
    bool unitary = true;  // every dense matrix is unitary (the other kinds always are)
    bool dense_ops = false;  // multi-qubit or parameterised matrix gates (plans of the FULL interpreter)
    // parameterised matrices (MQA_GATE_PMATRIX / _TAP): doubles a sample's row needs, and the table
    // of the calls that follow (mqa_circuit_set_custom_matrices)
    int64_t custom_need = 0;
    int64_t custom_rows = 0, custom_stride = 0;
    std::vector<double> custom_table;
    // angle table inputs (per slot)
    std::vector<double> slot_const;
    std::vector<int32_t> slot_term_begin;
    std::vector<int32_t> slot_term_param;
    std::vector<double> slot_term_coeff;
    std::vector<int8_t> slot_half;
    // jacobian CSR by parameter: (tap id, coeff)
    std::vector<int32_t> jac_begin, jac_tap;
    std::vector<double> jac_coeff;
    int n_tap_ids = 0;

    std::mutex mu;
    std::map<std::tuple<int, int, int, int, int, int, int, int, int, int>, std::unique_ptr<mqa::DevPlan>> plans;
    std::map<int, std::unique_ptr<mqa::DevBuf[]>> dev_slots;  // per device: 6 buffers

    // per-circuit tile widths (autotuner), [dtype][adjoint]; < 0: the global option
    std::atomic<int> tile_override[2][2] = {{-1, -1}, {-1, -1}};

    // phase_free: the plan may drop global phases of uncontrolled gates (OPK_PF_* kinds); only for
    // states whose results do not see them (the hermitian gradient's forward and back-sweep)
    mqa::SchedOptions options(int dtype, bool adjoint, bool phase_free = false, bool phase_ops = false) const {
        auto o = mqa::sched_options(dtype, adjoint);
        const int k = tile_override[dtype == MQA_C64 ? 0 : 1][adjoint ? 1 : 0].load();
        if (k > 0) o.tile_qubits = k;
        o.phase_free = phase_free;
        o.phase_ops = phase_ops;
        return o;
    }

    mqa::DevPlan* get_plan(int n_eff, int dtype, bool adjoint, hipStream_t s, bool upload_dev, bool phase_free = false,
                           bool phase_ops = false) {
        auto o = options(dtype, adjoint, phase_free, phase_ops);
        int dev = upload_dev ? mqa::current_device() : -1;
        auto key = std::make_tuple(n_eff, dtype, int(adjoint), o.tile_qubits, o.low_run, o.group_width,
                                   o.max_ops_per_pass, o.taps_per_group, int(o.phase_free) | (int(o.phase_ops) << 1), dev);
        std::lock_guard<std::mutex> lk(mu);
        auto it = plans.find(key);
        if (it != plans.end()) return it->second.get();
        auto dp = make_devplan(mqa::build_plan(gates, n_eff, adjoint, o), dtype, s, upload_dev);
        auto* raw = dp.get();
        plans[key] = std::move(dp);
        return raw;
    }

    static std::unique_ptr<mqa::DevPlan> make_devplan(mqa::Plan&& plan, int dtype, hipStream_t s, bool upload_dev) {
        auto dp = std::make_unique<mqa::DevPlan>();
        dp->plan = std::move(plan);
        // dense blocks or controls that stay per-amplitude predicates need the FULL instantiation,
        // phase blocks the PHASE one
        for (auto& op : dp->plan.dev_ops) {
            if (op.kind == mqa::OPK_M2 || op.kind == mqa::OPK_M3 || op.kind == mqa::OPK_M4 || op.ctrl_local != 0 ||
                mqa::opk_is_custom(op.kind))
                dp->full = true;
            if (op.kind == mqa::OPK_PHASE) dp->phase = true;
        }
        if (upload_dev) {
            mqa::upload(dp->passes, dp->plan.dev_passes, s);
            mqa::upload(dp->groups, dp->plan.dev_groups, s);
            mqa::upload(dp->ops, dp->plan.dev_ops, s);
            mqa::upload(dp->tap_ids, dp->plan.dev_tap_slots, s);
            if (!dp->plan.dev_terms.empty()) {
                mqa::upload(dp->terms, dp->plan.dev_terms, s);
                mqa::upload(dp->phase_coefs, dp->plan.phase_coefs, s);
            }
            const std::vector<uint8_t> forms = mqa::jit_coef_forms(dp->plan);  // alive until the sync
            mqa::upload(dp->coef_forms, forms, s);
            dp->n_coef_forms = int(forms.size());
            std::vector<float> pf;  // kept alive until the copy has completed
            if (dtype == MQA_C64) {
                pf.assign(dp->plan.pool.begin(), dp->plan.pool.end());
                mqa::upload(dp->pool, pf, s);
            } else {
                mqa::upload(dp->pool, dp->plan.pool, s);
            }
            mqa::hip_check(hipStreamSynchronize(s), "plan upload");
        }
        return dp;
    }

    // The forward sweep split at back-sweep pass boundaries, for checkpointed gradients.
    // After bound[i] back-sweep passes, psi must equal the forward state of the gates the
    // back-sweep has not undone yet (a dependency-closed subset, so its circuit-order product
    // is well defined).  Forward segment k (k < n) completes the state of boundary n-1-k (the
    // deepest boundary first); segment n completes the full forward state.
    struct Checkpointed {
        std::vector<int> bound;                           // increasing back-sweep pass counts
        std::vector<std::unique_ptr<mqa::DevPlan>> segs;  // bound.size() + 1 forward plans
    };
    std::map<std::tuple<const mqa::DevPlan*, int, int, int, int, int>, std::unique_ptr<Checkpointed>> ckpt_plans;

    // The segments drop the same phases as the back-sweep plan adj (a phase-free back-sweep
    // continues from phase-free checkpoints, so psi and lambda keep one common phase).
    Checkpointed* checkpointed(int n_eff, int dtype, int n_ck, const mqa::DevPlan* adj, hipStream_t s) {
        auto o = options(dtype, false, adj->plan.opts.phase_free, adj->plan.opts.phase_ops);
        std::lock_guard<std::mutex> lk(mu);
        auto key = std::make_tuple(adj, n_ck, o.tile_qubits, o.low_run, o.group_width, o.max_ops_per_pass);
        auto it = ckpt_plans.find(key);
        if (it != ckpt_plans.end()) return it->second.get();
        auto cp = std::make_unique<Checkpointed>();
        const int P = int(adj->plan.passes.size());
        for (int i = 1; i <= n_ck; i++) {
            const int b = int((int64_t(i) * P) / (n_ck + 1));
            if (b >= 1 && b <= P - 1 && (cp->bound.empty() || b > cp->bound.back())) cp->bound.push_back(b);
        }
        std::vector<int> pass_of(gates.size(), -1);  // back-sweep pass that undoes each gate
        for (int j = 0; j < P; j++)
            for (auto& g : adj->plan.passes[size_t(j)].groups)
                for (auto& op : g.ops) {
                    std::vector<int> gs;
                    mqa::op_gates(op, gs);
                    for (int gi : gs) pass_of[size_t(gi)] = j;
                }
        const int n = int(cp->bound.size());
        for (int k = 0; k <= n; k++) {
            std::vector<mqa::CGate> seg = gates;
            for (size_t gi = 0; gi < gates.size(); gi++) {
                const int p = pass_of[gi];
                int below = 0;  // boundaries at or before the gate's undo pass
                for (int b : cp->bound) below += (b <= p) ? 1 : 0;
                if (p < 0 || n - below != k) seg[gi].kind = MQA_GATE_I;
            }
            cp->segs.push_back(make_devplan(mqa::build_plan(seg, n_eff, false, o), dtype, s, true));
        }
        auto* raw = cp.get();
        ckpt_plans[key] = std::move(cp);
        return raw;
    }

    // device copies of the angle-table inputs and the jacobian (index order fixed below)
    mqa::DevBuf* slots_on_device(hipStream_t s) {
        int dev = mqa::current_device();
        std::lock_guard<std::mutex> lk(mu);
        auto it = dev_slots.find(dev);
        if (it != dev_slots.end()) return it->second.get();
        std::unique_ptr<mqa::DevBuf[]> b(new mqa::DevBuf[8]);
        mqa::upload(b[0], slot_const, s);
        mqa::upload(b[1], slot_term_begin, s);
        mqa::upload(b[2], slot_term_param, s);
        mqa::upload(b[3], slot_term_coeff, s);
        mqa::upload(b[4], slot_half, s);
        mqa::upload(b[5], jac_begin, s);
        mqa::upload(b[6], jac_tap, s);
        mqa::upload(b[7], jac_coeff, s);
        mqa::hip_check(hipStreamSynchronize(s), "slot upload");
        auto* raw = b.get();
        dev_slots[dev] = std::move(b);
        return raw;
    }
};

// ---------------------------------------------------------------- Hamiltonian
struct mqa_ham {
    std::vector<mqa_pauli_term> terms;
    // sum of the identity terms' coefficients, and the same operator without them (Pauli form only;
    // NULL when there is no identity term)
    double identity = 0;
    std::unique_ptr<mqa_ham> traceless;
    std::vector<mqa::DevHamGroup> groups;
    std::vector<mqa::DevHamTerm> dterms;
    uint64_t qubit_mask = 0;
    std::mutex mu;
    std::map<int, std::unique_ptr<mqa::DevBuf[]>> dev;

    // tile-engine plans keyed by (device, n_eff, K, C); ok == false: some flip fits no tile
    struct TilePlan {
        bool ok = false;
        mqa::HamTilePlan plan;
        mqa::DevBuf passes, groups, terms;
    };
    std::map<std::tuple<int, int, int, int>, std::unique_ptr<TilePlan>> tile_plans;

    TilePlan* tile_plan(int n_eff, int K, int C, hipStream_t s) {
        const int d = mqa::current_device();
        std::lock_guard<std::mutex> lk(mu);
        auto key = std::make_tuple(d, n_eff, K, C);
        auto it = tile_plans.find(key);
        if (it != tile_plans.end()) return it->second.get();
        auto tp = std::make_unique<TilePlan>();
        tp->ok = mqa::build_ham_plan(groups, dterms, n_eff, K, C, tp->plan);
        if (tp->ok) {
            mqa::upload(tp->passes, tp->plan.passes, s);
            mqa::upload(tp->groups, tp->plan.groups, s);
            mqa::upload(tp->terms, tp->plan.terms, s);
            mqa::hip_check(hipStreamSynchronize(s), "ham plan upload");
        }
        auto* raw = tp.get();
        tile_plans[key] = std::move(tp);
        return raw;
    }

    // sparse form (a CSR matrix given by the caller): [0] the operator, [1] its conjugate transpose
    bool csr = false;
    bool hermitian = true;  // Pauli terms have real coefficients; a CSR matrix is compared with H^dagger
    uint64_t csr_dim = 0;
    std::vector<int64_t> csr_ptr[2], csr_idx[2];
    std::vector<double> csr_val[2];  // interleaved complex
    std::map<std::pair<int, int>, std::unique_ptr<mqa::DevBuf[]>> csr_dev;  // (device, dtype) -> 6 buffers

    mqa::DevBuf* csr_on_device(int dtype, hipStream_t s) {
        const int d = mqa::current_device();
        std::lock_guard<std::mutex> lk(mu);
        auto key = std::make_pair(d, dtype);
        auto it = csr_dev.find(key);
        if (it != csr_dev.end()) return it->second.get();
        std::unique_ptr<mqa::DevBuf[]> b(new mqa::DevBuf[6]);
        std::vector<float> vf[2];
        for (int w = 0; w < 2; w++) {
            mqa::upload(b[3 * w], csr_ptr[w], s);
            mqa::upload(b[3 * w + 1], csr_idx[w], s);
            if (dtype == MQA_C64) {
                vf[w].assign(csr_val[w].begin(), csr_val[w].end());
                mqa::upload(b[3 * w + 2], vf[w], s);
            } else {
                mqa::upload(b[3 * w + 2], csr_val[w], s);
            }
        }
        mqa::hip_check(hipStreamSynchronize(s), "csr upload");
        auto* raw = b.get();
        csr_dev[key] = std::move(b);
        return raw;
    }

    mqa::DevBuf* on_device(hipStream_t s) {
        int d = mqa::current_device();
        std::lock_guard<std::mutex> lk(mu);
        auto it = dev.find(d);
        if (it != dev.end()) return it->second.get();
        std::unique_ptr<mqa::DevBuf[]> b(new mqa::DevBuf[2]);
        mqa::upload(b[0], groups, s);
        mqa::upload(b[1], dterms, s);
        mqa::hip_check(hipStreamSynchronize(s), "ham upload");
        auto* raw = b.get();
        dev[d] = std::move(b);
        return raw;
    }
};

// ---------------------------------------------------------------- simulator
struct mqa_sim {
    int n = 0;
    int n_eff = 0;
    int dtype = MQA_C64;
    uint32_t seed = 0;
    uint64_t dim = 0, dim_eff = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    mqa::DevBuf state;
    // workspaces (grow-only)
    mqa::DevBuf psi, lam, psi_keep, ckpt, angles, enc, ans, partials, taps, grads, fpart, fsum, scratch_state;
    mqa::DevBuf coefs;  // coefficient table of the compiled passes (rebuilt per run_plan call)
    mqa::DevBuf phase;  // phase table of the running plan's phase blocks (rebuilt per run_plan call)
    bool force_jit = false;  // compiled passes at any size (the solo items of an uneven batch)
    mqa::DevBuf custom[2];  // rows of the custom-gate matrices of the running circuits (two: bra and ket sides)
    mqa::DevBuf dense;      // matrices of the dense blocks of mqa_sim_apply_dense_blocks
    mqa::Timer timer;
    size_t amp_bytes() const { return dtype == MQA_C64 ? 8 : 16; }
    size_t state_bytes() const { return dim_eff * amp_bytes(); }
};

namespace mqa {
namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int d) {
        hip_check(hipGetDevice(&prev), "hipGetDevice");
        if (prev != d) hip_check(hipSetDevice(d), "hipSetDevice");
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

void require(bool ok, const std::string& msg) {
    if (!ok) throw std::invalid_argument(msg);
}

int n_eff_of(int n) { return std::max(n, kMinTileQubits); }

// Launch `fn` on the simulator's stream, bracketed by HIP events when profiling is on.
template <typename F>
void timed(mqa_sim* sim, const char* name, double bytes, F&& fn) {
    if (!g_opt.profile) {
        fn();
        return;
    }
    hipEvent_t a = sim->timer.get(), b = sim->timer.get();
    hip_check(hipEventRecord(a, sim->stream), "hipEventRecord");
    fn();
    hip_check(hipEventRecord(b, sim->stream), "hipEventRecord");
    sim->timer.pending.push_back({a, b, name, bytes});
}

std::atomic<int64_t> g_jit_launches{0};
std::atomic<int64_t> g_jit_sb_launches{0};       // of those, single-buffer kernels
std::atomic<int64_t> g_recomputes{0};  // forward sweeps recomputed because a kept copy did not fit
std::atomic<int64_t> g_ckpt_used{0};    // checkpoints the last batched gradient call held
std::atomic<int64_t> g_pf_used{0};      // 1: the last batched gradient call ran phase-free plans
void count_jit_launch() { g_jit_launches.fetch_add(1); }

void sync(mqa_sim* sim, const char* what) {
    hip_check(hipStreamSynchronize(sim->stream), what);
    sim->timer.resolve();
}

// Run every pass of a plan on `batch` states.
// Passes [p0, p1) of the plan (p1 < 0: to the end).
struct CustomRows {
    const double* p = nullptr;
    uint64_t stride = 0;
};

// src0: the first pass (p0, forward plans) reads its input tiles from src0 instead of psi: sample b
// at src0 + b * src_stride amplitudes (src_stride 0: one state broadcast to every sample).
// skip_last: the outputs of the last pass (p1 - 1) are dead (the last back-sweep pass of a
// gradient): it writes its taps only.
void run_plan(mqa_sim* sim, DevPlan* dp, void* psi, void* lam, uint64_t batch, const double* angles, int n_slots,
              double* partials, int n_tap_ids, int p0 = 0, int p1 = -1, CustomRows custom = {},
              const void* src0 = nullptr, bool skip_last = false, uint64_t src_stride = 0) {
    const Plan& pl = dp->plan;
    if (p1 < 0 || p1 > int(pl.passes.size())) p1 = int(pl.passes.size());
    if (p0 >= p1) return;
    TilePassArgs a{};
    a.psi = psi;
    a.lam = lam;
    a.K = pl.K;
    a.stride = sim->dim_eff;
    a.n_tiles = uint64_t(1) << (pl.n_eff - pl.K);
    a.passes = dp->passes.as<DevPass>();
    a.groups = dp->groups.as<DevGroup>();
    a.ops = dp->ops.as<DevOp>();
    a.pool = dp->pool.p;
    a.angles = angles;
    a.n_slots = n_slots;
    a.n_tap_ids = n_tap_ids;
    a.tap_ids = dp->tap_ids.as<int32_t>();
    a.partials = partials;
    a.batch = batch;
    a.custom = custom.p;
    a.custom_stride = custom.stride;
    const double bytes = double(batch) * double(sim->state_bytes()) * (pl.adjoint ? 4.0 : 2.0);
    JitPlan* jp = nullptr;
    const int64_t runs = dp->runs.fetch_add(1) + 1;
    const bool big = double(batch) * double(sim->dim_eff) >= double(g_opt.jit_min_amps);
    const bool recurring = g_opt.jit_min_runs > 0 && runs >= g_opt.jit_min_runs;
    if (g_opt.jit != 0 && (big || recurring || sim->force_jit)) {
        std::lock_guard<std::mutex> lk(dp->jit_mu);
        JitForm form = JitForm::kTile;
        if (g_opt.jit_single_buffer != 0 && jit_can_single_buffer(pl, sim->dtype)) form = JitForm::kSingleBuffer;
        auto& slot = form == JitForm::kSingleBuffer ? dp->jit_sb : dp->jit;
        if (!slot) slot = std::make_unique<JitPlan>(pl, sim->dtype, form);
        if (g_opt.jit == 2) slot->wait();
        if (slot->ready()) jp = slot.get();  // until then (or if it failed): the interpreted kernel
    }
    // compiled passes read their coefficients ready-made: the table follows this call's angles
    if (jp && angles && n_slots > 0 && dp->n_coef_forms > 0) {
        const size_t real = sim->dtype == MQA_C64 ? sizeof(float) : sizeof(double);
        sim->coefs.ensure(size_t(batch) * size_t(n_slots) * 2 * real);
        hip_check(launch_coefs(sim->dtype, angles, dp->coef_forms.as<uint8_t>(), dp->n_coef_forms, n_slots, batch,
                               sim->coefs.p, sim->stream),
                  "coefficient table");
        a.coefs = sim->coefs.p;
    }
    // phase blocks: the table of the terms' e^{i phi} follows this call's angles
    if (!pl.dev_terms.empty()) {
        const size_t real = sim->dtype == MQA_C64 ? sizeof(float) : sizeof(double);
        const int nt = int(pl.dev_terms.size());
        sim->phase.ensure(size_t(batch) * size_t(nt) * 2 * real);
        hip_check(launch_phase_coefs(sim->dtype, angles, dp->phase_coefs.as<PhaseCoef>(), nt, n_slots, batch,
                                     sim->phase.p, sim->stream),
                  "phase table");
        a.terms = dp->terms.as<DevPhaseTerm>();
        a.phase = sim->phase.p;
        a.n_phase = nt;
    }
    // a single-buffer pass holds one tile of LDS (the forward layout) plus the tap table
    const bool one_tile = jp && jp->form() == JitForm::kSingleBuffer;
    const size_t lds =
        tile_pass_lds_bytes(sim->dtype, pl.adjoint && !one_tile, pl.K, pl.C, std::max(pl.max_taps, 1));
    for (int i = p0; i < p1; i++) {
        a.pass_index = i;
        a.src = (i == p0 && !pl.adjoint) ? src0 : nullptr;
        a.src_stride = a.src ? src_stride : 0;
        a.skip_store = (skip_last && i == p1 - 1) ? 1 : 0;
        // algorithmic bytes: a broadcast input is read once, dead outputs are not written
        double pass_bytes = bytes;
        if (a.src && a.src_stride == 0) pass_bytes = double(batch) * double(sim->state_bytes()) + double(sim->state_bytes());
        if (a.skip_store) pass_bytes = bytes / 2;
        timed(sim, pl.adjoint ? "tile_pass_adj" : "tile_pass_fwd", pass_bytes, [&] {
            if (jp) {
                hip_check(jp->launch(i, a, batch, lds, sim->stream), "compiled tile pass");
            } else {
                hip_check(launch_tile_pass(sim->dtype, pl.adjoint, dp->full, dp->phase, a, batch, pl.C,
                                           std::max(pl.max_taps, 1), sim->stream),
                          "tile pass");
            }
        });
        if (jp) count_jit_launch();
        if (one_tile) g_jit_sb_launches.fetch_add(1);
    }
}

// Phase blocks (option phase_ops) for plans that compile (jit on: large launches at once, smaller
// ones once they recur); the interpreted pass applies a block term by term, which costs more than
// the per-gate ops it replaces, so with jit off the plans keep one op per gate.  Value 2 uses
// them whatever the jit setting.
// Launches below 2^22 amplitudes (a single small state) are latency-bound: there the serial
// products of a block cost more than the per-gate ops, so their plans keep one op per gate; so do
// circuits with multi-qubit or parameterised matrix gates, whose plans run the FULL interpreted
// instantiation until compiled, which phase blocks would make heavier still.
bool want_phase_ops(const mqa_circuit* c, uint64_t batch, uint64_t dim) {
    if (g_opt.phase_ops == 2) return true;
    return g_opt.phase_ops != 0 && g_opt.jit != 0 && !c->dense_ops && double(batch) * double(dim) >= 4194304.0;
}

// Whether a plan's compiled passes are ready, in the form run_plan would launch; starts their
// compile (in the background) if it has not started.  jit 2 waits for it.
bool jit_ready(DevPlan* dp, int dtype) {
    if (g_opt.jit == 0) return false;
    std::lock_guard<std::mutex> lk(dp->jit_mu);
    const Plan& pl = dp->plan;
    JitForm form = JitForm::kTile;
    if (g_opt.jit_single_buffer != 0 && jit_can_single_buffer(pl, dtype)) form = JitForm::kSingleBuffer;
    auto& slot = form == JitForm::kSingleBuffer ? dp->jit_sb : dp->jit;
    if (!slot) slot = std::make_unique<JitPlan>(pl, dtype, form);
    if (g_opt.jit == 2) slot->wait();
    return slot->ready();
}

// Device rows [b0, b0 + nb) of a circuit's custom-gate matrices (a one-row table serves every sample),
// in the simulator's workspace `which` (0 or 1: the two circuits of one call use separate ones).
// one_row: every state of the launch belongs to sample b0 (row b0 for all of them).
CustomRows custom_rows(mqa_sim* sim, mqa_circuit* c, int64_t b0, int64_t nb, int which = 0, bool one_row = false) {
    CustomRows r;
    if (c == nullptr || c->custom_need == 0) return r;
    std::lock_guard<std::mutex> lk(c->mu);
    require(c->custom_rows >= 1, "the circuit has parameterised matrix gates: set their values first "
                                 "(mqa_circuit_set_custom_matrices)");
    const bool shared = c->custom_rows == 1;
    const int64_t need_rows = one_row ? b0 + 1 : b0 + nb;
    require(shared || c->custom_rows >= need_rows, "custom gate matrices hold " + std::to_string(c->custom_rows) +
                                                       " rows, the call needs " + std::to_string(need_rows));
    const bool single = shared || one_row;
    const size_t n = size_t(single ? 1 : nb) * size_t(c->custom_stride);
    DevBuf& buf = sim->custom[which];
    buf.ensure(n * sizeof(double));
    hip_check(hipMemcpyAsync(buf.p, c->custom_table.data() + (shared ? 0 : size_t(b0) * size_t(c->custom_stride)),
                             n * sizeof(double), hipMemcpyHostToDevice, sim->stream),
              "custom matrices");
    // the host table may change after the call returns: the copy completes before that (every entry
    // point synchronises its stream before returning)
    r.p = buf.as<double>();
    r.stride = single ? 0 : uint64_t(c->custom_stride);
    return r;
}

// (cos, sin) table for `batch` rows: params p < n_enc from enc rows, the rest from ans.
const double* build_angles(mqa_sim* sim, mqa_circuit* circ, const double* enc_host, uint64_t batch, int n_enc,
                           const double* ans_host, int n_ans, DevBuf* out = nullptr) {
    if (circ->n_slots == 0) return nullptr;
    DevBuf& table = out ? *out : sim->angles;
    DevBuf* sl = circ->slots_on_device(sim->stream);
    hipStream_t s = sim->stream;
    sim->enc.ensure(std::max<size_t>(batch * n_enc, 1) * sizeof(double));
    sim->ans.ensure(std::max<size_t>(n_ans, 1) * sizeof(double));
    if (n_enc > 0)
        hip_check(hipMemcpyAsync(sim->enc.p, enc_host, batch * n_enc * sizeof(double), hipMemcpyHostToDevice, s),
                  "enc upload");
    if (n_ans > 0)
        hip_check(hipMemcpyAsync(sim->ans.p, ans_host, n_ans * sizeof(double), hipMemcpyHostToDevice, s), "ans upload");
    table.ensure(batch * circ->n_slots * 2 * sizeof(double));
    AngleArgs aa{};
    aa.enc = sim->enc.as<double>();
    aa.ans = sim->ans.as<double>();
    aa.n_enc = n_enc;
    aa.n_slots = circ->n_slots;
    aa.batch = batch;
    aa.const_val = sl[0].as<double>();
    aa.term_begin = sl[1].as<int32_t>();
    aa.term_param = sl[2].as<int32_t>();
    aa.term_coeff = sl[3].as<double>();
    aa.half = sl[4].as<int8_t>();
    aa.out = table.as<double>();
    hip_check(launch_angles(aa, s), "angles");
    return table.as<double>();
}

// Tile plan of a Hamiltonian: tiles of the option's width, or one qubit narrower when that fits in
// no more passes (half the LDS per block, so twice the resident blocks per CU).  ok == false: no
// tile holds some flip group.
mqa_ham::TilePlan* ham_plan_for(mqa_ham* ham, int n_eff, int dtype, hipStream_t s) {
    // complex128 tiles above 13 qubits exceed the LDS of a CU
    const int K = int(dtype == MQA_C64 ? g_opt.ham_tile_c64.load() : std::min<int64_t>(g_opt.ham_tile_c128, 13));
    const int C = int(dtype == MQA_C64 ? g_opt.lowq_c64 : g_opt.lowq_c128);
    mqa_ham::TilePlan* wide = ham->tile_plan(n_eff, K, C, s);
    if (K - 1 - C >= kGroupWidth && n_eff > K - 1) {
        mqa_ham::TilePlan* narrow = ham->tile_plan(n_eff, K - 1, C, s);
        if (narrow->ok && (!wide->ok || narrow->plan.passes.size() <= wide->plan.passes.size())) return narrow;
    }
    return wide;
}

// (cos, sin) table of a circuit's angle slots for one parameter vector (the half angle for RX /
// RY / RZ and the two-qubit rotations, the angle for PS / GP), as k_angles computes it.
std::vector<double> host_angles(const mqa_circuit* c, const double* pv) {
    std::vector<double> tab(2 * size_t(c->n_slots));
    for (int sl = 0; sl < c->n_slots; sl++) {
        double th = c->slot_const[size_t(sl)];
        for (int k = c->slot_term_begin[size_t(sl)]; k < c->slot_term_begin[size_t(sl) + 1]; k++)
            th += c->slot_term_coeff[size_t(k)] * pv[c->slot_term_param[size_t(k)]];
        if (c->slot_half[size_t(sl)]) th *= 0.5;
        tab[2 * size_t(sl)] = std::cos(th);
        tab[2 * size_t(sl) + 1] = std::sin(th);
    }
    return tab;
}

// The global phase a phase-free plan of circuit c drops (exact state = phase x phase-free state),
// from the angle table the kernels read: -1 for an uncontrolled RX / RY whose half-angle cosine
// is negative in the state's precision, e^{-i theta/2} for an uncontrolled RZ or RZZ, e^{-i theta}
// for an uncontrolled GP (the OPK_PF_* conventions of program.h).
std::complex<double> dropped_phase(const mqa_circuit* c, const std::vector<double>& tab, int dtype) {
    std::complex<double> ph(1.0, 0.0);
    for (const auto& g : c->gates) {
        if (g.ctrl != 0 || g.slot < 0) continue;
        const double cs = tab[2 * size_t(g.slot)], sn = tab[2 * size_t(g.slot) + 1];
        if (g.kind == MQA_GATE_RX || g.kind == MQA_GATE_RY) {
            const bool neg = dtype == MQA_C64 ? (float(cs) < 0.0f) : (cs < 0.0);
            if (neg) ph = -ph;
        } else if (g.kind == MQA_GATE_RZ || g.kind == MQA_GATE_RZZ || g.kind == MQA_GATE_GP) {
            ph *= std::complex<double>(cs, -sn);
        }
    }
    return ph;
}

// out = H ket over a batch with the tile engine; with `expect`, per-sample <ket|H|ket> in
// sim->fsum.  Returns false (nothing launched) when the Hamiltonian has no tile plan.
bool ham_apply_tiles(mqa_sim* sim, mqa_ham* ham, const void* ket, void* out, bool expect, uint64_t batch) {
    if (!g_opt.ham_engine) return false;
    mqa_ham::TilePlan* tp = ham_plan_for(ham, sim->n_eff, sim->dtype, sim->stream);
    if (!tp->ok) return false;
    const HamTilePlan& pl = tp->plan;
    const uint64_t n_tiles = uint64_t(1) << (sim->n_eff - pl.K);
    HamTileArgs h{};
    h.ket = ket;
    h.out = out;
    h.stride = sim->dim_eff;
    h.K = pl.K;
    h.passes = tp->passes.as<DevHamPass>();
    h.groups = tp->groups.as<DevHamTileGroup>();
    h.terms = tp->terms.as<DevHamTileTerm>();
    h.n_tiles = n_tiles;
    if (expect) {
        sim->fpart.ensure(batch * n_tiles * 2 * sizeof(double));
        sim->fsum.ensure(batch * 2 * sizeof(double));
    }
    const int np = int(pl.passes.size());
    for (int p = 0; p < np; p++) {
        h.pass_index = p;
        h.partial = (expect && p == np - 1) ? sim->fpart.as<double>() : nullptr;
        const double bytes = double(batch) * double(sim->state_bytes()) * (p == 0 ? 2.0 : 3.0);
        timed(sim, "hamiltonian", bytes, [&] {
            hip_check(launch_ham_tile(sim->dtype, p == 0, p == np - 1, h, pl.C, n_tiles, batch, sim->stream),
                      "hamiltonian tile pass");
        });
    }
    if (expect)
        hip_check(launch_reduce_rows(sim->fpart.as<double>(), uint32_t(n_tiles), sim->fsum.as<double>(), batch,
                                     sim->stream),
                  "reduce");
    return true;
}

// Sparse Hamiltonian: out = H ket (or H^dagger ket) by CSR products, then <bra|out> per sample.
void csr_apply(mqa_sim* sim, mqa_ham* ham, const void* ket, const void* bra, void* out, uint64_t batch, bool dagger) {
    DevBuf* d = ham->csr_on_device(sim->dtype, sim->stream);
    const int w = dagger ? 1 : 0;
    void* target = out;
    if (!target) {
        sim->scratch_state.ensure(sim->state_bytes() * batch);
        target = sim->scratch_state.p;
    }
    timed(sim, "hamiltonian", double(batch) * double(sim->state_bytes()) * 2.0, [&] {
        hip_check(launch_csr_spmv(sim->dtype, d[3 * w].as<int64_t>(), d[3 * w + 1].as<int64_t>(), d[3 * w + 2].p,
                                  ham->csr_dim, ket, sim->dim_eff, target, sim->dim_eff, sim->dim_eff, batch,
                                  sim->stream),
                  "sparse hamiltonian");
    });
    if (!bra) return;
    VdotArgs v{};
    v.a = bra;
    v.a_stride = sim->dim_eff;
    v.b = target;
    v.b_stride = sim->dim_eff;
    v.dim = sim->dim_eff;
    v.amps_per_block = uint32_t(std::min<uint64_t>(sim->dim_eff, 4096));
    v.n_blocks = uint32_t((sim->dim_eff + v.amps_per_block - 1) / v.amps_per_block);
    sim->fpart.ensure(batch * v.n_blocks * 2 * sizeof(double));
    sim->fsum.ensure(batch * 2 * sizeof(double));
    v.partial = sim->fpart.as<double>();
    hip_check(launch_vdot(sim->dtype, v, batch, sim->stream), "vdot");
    hip_check(launch_reduce_rows(sim->fpart.as<double>(), v.n_blocks, sim->fsum.as<double>(), batch, sim->stream),
              "reduce");
}

// <bra|H|ket> or H ket over a batch; returns per-sample sums in sim->fsum (device, [batch][2]).
// dagger: apply H^dagger (differs from H only for a sparse Hamiltonian given as a matrix).
void ham_apply(mqa_sim* sim, mqa_ham* ham, const void* ket, const void* bra, void* out, uint64_t batch,
               bool dagger = false) {
    if (ham->csr) {
        csr_apply(sim, ham, ket, bra, out, batch, dagger);
        return;
    }
    if (out && (bra == nullptr || bra == ket) && ham_apply_tiles(sim, ham, ket, out, bra != nullptr, batch)) return;
    DevBuf* hd = ham->on_device(sim->stream);
    HamArgs h{};
    h.ket = ket;
    h.ket_stride = sim->dim_eff;
    h.bra = bra;
    h.bra_stride = sim->dim_eff;
    h.out = out;
    h.out_stride = sim->dim_eff;
    h.dim = sim->dim_eff;
    h.amps_per_block = uint32_t(std::min<uint64_t>(sim->dim_eff, 2048));
    h.n_blocks = uint32_t((sim->dim_eff + h.amps_per_block - 1) / h.amps_per_block);
    h.groups = hd[0].as<DevHamGroup>();
    h.n_groups = int(ham->groups.size());
    h.terms = hd[1].as<DevHamTerm>();
    sim->fpart.ensure(batch * h.n_blocks * 2 * sizeof(double));
    sim->fsum.ensure(batch * 2 * sizeof(double));
    h.partial = bra ? sim->fpart.as<double>() : nullptr;
    const double passes = 1.0 + (out ? 1.0 : 0.0) + ((bra && bra != ket) ? 1.0 : 0.0);
    timed(sim, "hamiltonian", double(batch) * double(sim->state_bytes()) * passes, [&] {
        hip_check(launch_ham(sim->dtype, out != nullptr, h, batch, sim->stream), "hamiltonian");
    });
    if (bra)
        hip_check(launch_reduce_rows(sim->fpart.as<double>(), h.n_blocks, sim->fsum.as<double>(), batch, sim->stream),
                  "reduce");
}

void vdot(mqa_sim* sim, const void* a, const void* b, double out[2]) {
    VdotArgs v{};
    v.a = a;
    v.a_stride = sim->dim_eff;
    v.b = b;
    v.b_stride = sim->dim_eff;
    v.dim = sim->dim_eff;
    v.amps_per_block = uint32_t(std::min<uint64_t>(sim->dim_eff, 4096));
    v.n_blocks = uint32_t((sim->dim_eff + v.amps_per_block - 1) / v.amps_per_block);
    sim->fpart.ensure(v.n_blocks * 2 * sizeof(double));
    sim->fsum.ensure(2 * sizeof(double));
    v.partial = sim->fpart.as<double>();
    hip_check(launch_vdot(sim->dtype, v, 1, sim->stream), "vdot");
    hip_check(launch_reduce_rows(sim->fpart.as<double>(), v.n_blocks, sim->fsum.as<double>(), 1, sim->stream), "reduce");
    hip_check(hipMemcpyAsync(out, sim->fsum.p, 2 * sizeof(double), hipMemcpyDeviceToHost, sim->stream), "d2h");
    hip_check(hipStreamSynchronize(sim->stream), "sync");
}

// <state|state> of the simulator's own state.
double state_norm2(mqa_sim* sim) {
    double out[2];
    vdot(sim, sim->state.p, sim->state.p, out);
    return out[0];
}

void check_circuit(const mqa_sim* sim, const mqa_circuit* c, int n_params) {
    require(c != nullptr, "circuit is NULL");
    require(c->max_qubit < sim->n, "Simulator has " + std::to_string(sim->n) + " qubits, but circuit has " +
                                       std::to_string(c->max_qubit + 1) + " qubits.");
    require(n_params >= c->n_params, "circuit needs " + std::to_string(c->n_params) + " parameters, got " +
                                         std::to_string(n_params));
}

void check_ham(const mqa_sim* sim, const mqa_ham* h) {
    require(h != nullptr, "hamiltonian is NULL");
    if (h->csr) {  // cpu_vector_core_dot_like.cpp:134-137
        require(h->csr_dim == sim->dim, "Sparse hamiltonian size not match with quantum state size.");
        return;
    }
    require((h->qubit_mask >> sim->n) == 0 || sim->n >= 64, "hamiltonian acts on more qubits than the simulator has");
}

// A Pauli-term Hamiltonian grouped by flip mask x|y; the coefficient carries i^{|y|} (the
// reference's POLAR phase, mq_base/utils.cc:22): P|src> = i^{|y|} (-1)^{popcount(src & (y|z))}
// |src ^ (x|y)>.
std::unique_ptr<mqa_ham> pauli_ham(const mqa_pauli_term* terms, int32_t n_terms) {
    auto h = std::make_unique<mqa_ham>();
    std::map<uint64_t, std::vector<int>> by_flip;
    for (int t = 0; t < n_terms; t++) {
        const auto& tm = terms[t];
        h->terms.push_back(tm);
        h->qubit_mask |= tm.x_mask | tm.y_mask | tm.z_mask;
        by_flip[tm.x_mask | tm.y_mask].push_back(t);
    }
    for (auto& [flip, ids] : by_flip) {
        DevHamGroup g{};
        g.flip = flip;
        g.term_begin = int32_t(h->dterms.size());
        g.n_terms = int32_t(ids.size());
        for (int t : ids) {
            const auto& tm = terms[t];
            const int ny = __builtin_popcountll(tm.y_mask) & 3;
            static const double pr[4] = {1, 0, -1, 0}, pi[4] = {0, 1, 0, -1};
            DevHamTerm d{};
            d.yz = tm.y_mask | tm.z_mask;
            d.coef_re = tm.coeff * pr[ny];
            d.coef_im = tm.coeff * pi[ny];
            h->dterms.push_back(d);
        }
        h->groups.push_back(g);
    }
    if (h->groups.empty()) {  // zero operator: one empty group keeps the kernels uniform
        DevHamGroup g{};
        g.flip = 0;
        g.term_begin = 0;
        g.n_terms = 0;
        h->groups.push_back(g);
    }
    return h;
}

// The observable whose product with psi is the hermitian gradient's co-state lambda: for a unitary
// circuit, H without its identity part c I, which adds c <psi|psi> (the initial state's norm) to f
// and nothing to any derivative; keeping it out of lambda keeps its rounding out of the taps (a
// MaxCut cost has c = |E| / 2 beside terms of size 1 / 2).  *shift gets c (0: H itself is used).
mqa_ham* costate_ham(mqa_ham* h, const mqa_circuit* c, double* shift) {
    *shift = 0;
    if (h->csr || !h->traceless || !c->unitary) return h;
    *shift = h->identity;
    return h->traceless.get();
}

// Workspace budget of a call: the option / environment value, else 85% of the free memory plus
// `held`, the bytes of the simulator's grow-only workspaces that the call may reuse (a repeated
// call must see the same budget as the first, which allocated them).
size_t hbm_budget(size_t held = 0) {
    int64_t b = g_opt.budget;
    if (const char* env = std::getenv("MQAMD_HBM_BUDGET_BYTES")) b = std::atoll(env);
    if (b > 0) return size_t(b);
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    return size_t(double(free_b + held) * 0.85);
}

// Device workspaces of the batched gradient for a workspace budget: samples per chunk, whether the
// forward state is kept for further Hamiltonians, and the bytes of psi, lambda, the kept state and
// the checkpoints (n_ck already fitted to the budget).  side: per-sample bytes of the small tables.
struct GradWorkspace {
    int64_t chunk = 1;
    bool keep = false;
    size_t bytes[4] = {0, 0, 0, 0};
};
GradWorkspace plan_grad_workspace(size_t budget, size_t sb, size_t side, int n_hams, int n_ck, int64_t batch) {
    GradWorkspace w;
    w.keep = n_ck == 0 && n_hams > 1 && budget >= sb * 3 + side;
    const size_t per_sample = sb * size_t(2 + n_ck + (w.keep ? 1 : 0)) + side;
    w.chunk = std::min<int64_t>({int64_t(std::max<size_t>(budget / per_sample, 1)), batch, 65535});
    const size_t one = sb * size_t(w.chunk);
    w.bytes[0] = w.bytes[1] = one;
    w.bytes[2] = w.keep ? one : 0;
    w.bytes[3] = one * size_t(n_ck);
    return w;
}

// The budget counted the held workspaces as free memory.  A workspace that must grow is released
// before anything is allocated; one that is larger than needed is kept only while the kept and the
// new sizes together stay within the budget (so a repeated call keeps its buffers, and a call after
// a larger one does not allocate on top of stale ones).  release[i] tells which are released.
void fit_workspaces(const size_t* held, const size_t* need, int n, size_t budget, bool* release) {
    size_t total = 0;
    for (int i = 0; i < n; i++) total += std::max(held[i], need[i]);
    for (int i = 0; i < n; i++)
        release[i] = held[i] != need[i] && (held[i] < need[i] || need[i] == 0 || total > budget);
}

void fit_and_ensure(DevBuf* const* bufs, const size_t* need, int n, size_t budget) {
    std::vector<size_t> held(size_t(n), 0);
    for (int i = 0; i < n; i++) held[size_t(i)] = bufs[i]->bytes;
    std::unique_ptr<bool[]> rel(new bool[size_t(n)]);
    fit_workspaces(held.data(), need, n, budget, rel.get());
    for (int i = 0; i < n; i++)
        if (rel[size_t(i)]) bufs[i]->release();
    for (int i = 0; i < n; i++)
        if (need[i] > 0) bufs[i]->ensure(need[i]);
}

}  // namespace
}  // namespace mqa

using namespace mqa;

extern "C" {

const char* mqa_last_error(void) { return g_last_error.c_str(); }

int mqa_grad_workspace_plan(uint64_t budget_bytes, uint64_t state_bytes, uint64_t side_bytes, int32_t n_hams,
                            int32_t n_checkpoints, int64_t batch, const uint64_t held[4], int64_t* chunk, int32_t* keep,
                            uint64_t need[4], int32_t release[4]) {
    return guarded([&] {
        require(state_bytes > 0 && n_hams >= 1 && n_checkpoints >= 0 && batch >= 1, "bad argument");
        require(held != nullptr && chunk != nullptr && keep != nullptr && need != nullptr && release != nullptr,
                "NULL argument");
        const GradWorkspace w = plan_grad_workspace(size_t(budget_bytes), size_t(state_bytes), size_t(side_bytes),
                                                    n_hams, n_checkpoints, batch);
        size_t h[4], nd[4];
        bool rel[4];
        for (int i = 0; i < 4; i++) h[i] = size_t(held[i]), nd[i] = w.bytes[i];
        fit_workspaces(h, nd, 4, size_t(budget_bytes), rel);
        *chunk = w.chunk;
        *keep = w.keep ? 1 : 0;
        for (int i = 0; i < 4; i++) need[i] = nd[i], release[i] = rel[i] ? 1 : 0;
    });
}

const char* mqa_version(void) { return "mqamd 0.1.0 (gfx950)"; }

int mqa_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int mqa_set_option(const char* key, int64_t value) {
    return guarded([&] {
        require(key != nullptr, "option key is NULL");
        for (auto& e : opt_table()) {
            if (std::strcmp(e.key, key) == 0) {
                *e.v = value < 0 ? e.def : value;
                return;
            }
        }
        throw std::invalid_argument(std::string("unknown option ") + key);
    });
}

int mqa_get_option(const char* key, int64_t* value) {
    return guarded([&] {
        require(key != nullptr && value != nullptr, "NULL argument");
        if (std::strcmp(key, "jit_launches") == 0) {  // read-only: passes launched from compiled kernels
            *value = g_jit_launches.load();
            return;
        }
        if (std::strcmp(key, "jit_single_buffer_launches") == 0) {  // read-only: of those, single-buffer ones
            *value = g_jit_sb_launches.load();
            return;
        }
        if (std::strcmp(key, "grad_recomputes") == 0) {  // read-only: see mqa_sim_grad_batch
            *value = g_recomputes.load();
            return;
        }
        if (std::strcmp(key, "checkpoints_used") == 0) {  // read-only: see mqa_sim_grad_batch
            *value = g_ckpt_used.load();
            return;
        }
        if (std::strcmp(key, "phase_free_used") == 0) {  // read-only: see mqa_sim_grad_batch
            *value = g_pf_used.load();
            return;
        }
        if (std::strcmp(key, "jit_compile_ms") == 0) {  // read-only: time spent compiling generated passes
            *value = int64_t(jit_compile_seconds_total() * 1e3);
            return;
        }
        for (auto& e : opt_table()) {
            if (std::strcmp(e.key, key) == 0) {
                *value = *e.v;
                return;
            }
        }
        throw std::invalid_argument(std::string("unknown option ") + key);
    });
}

int mqa_profile_get(const char* kernel, int64_t* launches, double* total_ms, double* bytes) {
    return guarded([&] {
        require(kernel != nullptr, "kernel name is NULL");
        std::lock_guard<std::mutex> lk(g_prof_mu);
        KernelStats st;
        auto it = g_prof.find(kernel);
        if (it != g_prof.end()) st = it->second;
        if (launches) *launches = st.launches;
        if (total_ms) *total_ms = st.ms;
        if (bytes) *bytes = st.bytes;
    });
}

int mqa_profile_reset(void) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        g_prof.clear();
    });
}

int mqa_circuit_create(const mqa_gate* gates, int32_t n_gates, const int32_t* term_param, const double* term_coeff,
                       int32_t n_terms, const double* mats, int64_t n_mat_entries, int32_t n_params,
                       const uint8_t* param_requires_grad, mqa_circuit** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        require(n_gates >= 0 && (n_gates == 0 || gates != nullptr), "bad gate array");
        require(n_params >= 0, "bad parameter count");
        auto c = std::make_unique<mqa_circuit>();
        c->n_params = n_params;
        std::vector<std::vector<std::pair<int, double>>> by_param(n_params);
        for (int i = 0; i < n_gates; i++) {
            const mqa_gate& s = gates[i];
            CGate g;
            g.kind = s.kind;
            require(s.n_obj >= 1 && s.n_obj <= MQA_MAX_OBJ, "gate " + std::to_string(i) + ": bad target count");
            for (int k = 0; k < s.n_obj; k++) {
                require(s.obj[k] >= 0 && s.obj[k] < 64, "gate " + std::to_string(i) + ": bad qubit");
                g.objs.push_back(s.obj[k]);
                c->max_qubit = std::max(c->max_qubit, s.obj[k]);
            }
            g.ctrl = s.ctrl_mask;
            for (int q = 0; q < 64; q++)
                if ((s.ctrl_mask >> q) & 1) c->max_qubit = std::max(c->max_qubit, q);
            require(!(g.targets() & g.ctrl), "gate " + std::to_string(i) + ": a qubit is both target and control");
            const bool custom_kind = s.kind == MQA_GATE_PMATRIX || s.kind == MQA_GATE_PMATRIX_TAP;
            bool param_kind = (s.kind >= MQA_GATE_RX && s.kind <= MQA_GATE_RZZ) || custom_kind;
            bool fixed_kind = (s.kind >= MQA_GATE_I && s.kind <= MQA_GATE_MATRIX);
            if (!param_kind && !fixed_kind)
                throw Error(MQA_ERR_UNSUPPORTED, "Apply of gate " + std::to_string(s.kind) + " not implement.");
            int want = 1;
            if (s.kind == MQA_GATE_SWAP || s.kind == MQA_GATE_ISWAP || s.kind == MQA_GATE_ISWAPDAG ||
                s.kind == MQA_GATE_RXX || s.kind == MQA_GATE_RYY || s.kind == MQA_GATE_RZZ)
                want = 2;
            if (custom_kind) {
                require(s.n_obj <= 2, "gate " + std::to_string(i) +
                                          ": parameterised matrix gates act on one or two target qubits");
                require(s.mat_begin >= 0, "gate " + std::to_string(i) + ": bad custom matrix offset");
                const int64_t dd = int64_t(2) << (2 * s.n_obj);  // doubles of one D x D matrix
                g.custom_off = s.mat_begin;
                c->custom_need = std::max(c->custom_need, int64_t(s.mat_begin) + (s.kind == MQA_GATE_PMATRIX ? 3 : 1) * dd);
            } else if (s.kind != MQA_GATE_MATRIX) {
                require(s.n_obj == want, "gate " + std::to_string(i) + ": wrong number of target qubits");
            }
            if ((s.kind == MQA_GATE_MATRIX && s.n_obj >= 2) || custom_kind) c->dense_ops = true;
            if (s.kind == MQA_GATE_MATRIX) {
                int64_t m = int64_t(1) << s.n_obj;
                require(mats != nullptr && s.mat_begin >= 0 && s.mat_begin + m * m <= n_mat_entries,
                        "gate " + std::to_string(i) + ": matrix out of range");
                for (int64_t e = 0; e < m * m; e++)
                    g.mat.push_back({mats[2 * (s.mat_begin + e)], mats[2 * (s.mat_begin + e) + 1]});
                // M^dagger M == I to 1e-12 (a projector or a derivative matrix is not unitary)
                for (int64_t r = 0; r < m && c->unitary; r++)
                    for (int64_t q = 0; q < m; q++) {
                        std::complex<double> acc = 0;
                        for (int64_t k = 0; k < m; k++)
                            acc += std::conj(g.mat[size_t(k * m + r)]) * g.mat[size_t(k * m + q)];
                        if (std::abs(acc - (r == q ? 1.0 : 0.0)) > 1e-12) {
                            c->unitary = false;
                            break;
                        }
                    }
            }
            if (param_kind) {
                g.slot = c->n_slots++;
                g.const_val = s.const_val;
                require(s.n_term >= 0 && s.term_begin >= 0 && s.term_begin + s.n_term <= n_terms,
                        "gate " + std::to_string(i) + ": parameter terms out of range");
                c->slot_const.push_back(s.const_val);
                c->slot_term_begin.push_back(int32_t(c->slot_term_param.size()));
                c->slot_half.push_back((s.kind == MQA_GATE_PS || s.kind == MQA_GATE_GP) ? 0 : 1);
                for (int t = 0; t < s.n_term; t++) {
                    int p = term_param[s.term_begin + t];
                    require(p >= 0 && p < n_params, "gate " + std::to_string(i) + ": parameter index out of range");
                    g.terms.push_back({p, term_coeff[s.term_begin + t]});
                    c->slot_term_param.push_back(p);
                    c->slot_term_coeff.push_back(term_coeff[s.term_begin + t]);
                }
                g.grad = s.requires_grad != 0 && s.n_term > 0;
                if (g.grad) {
                    int tap = c->n_tap_ids++;
                    for (auto& t : g.terms) by_param[t.first].push_back({tap, t.second});
                }
            }
            c->gates.push_back(std::move(g));
        }
        c->slot_term_begin.push_back(int32_t(c->slot_term_param.size()));
        c->jac_begin.push_back(0);
        for (int p = 0; p < n_params; p++) {
            if (param_requires_grad && !param_requires_grad[p]) by_param[p].clear();
            for (auto& e : by_param[p]) {
                c->jac_tap.push_back(e.first);
                c->jac_coeff.push_back(e.second);
            }
            c->jac_begin.push_back(int32_t(c->jac_tap.size()));
        }
        *out = c.release();
    });
}

int mqa_circuit_set_tile_qubits(mqa_circuit* circ, int32_t dtype, int32_t adjoint, int32_t tile_qubits) {
    return guarded([&] {
        require(circ != nullptr, "circuit is NULL");
        require(dtype == MQA_C64 || dtype == MQA_C128, "bad dtype");
        require(tile_qubits < 0 || (tile_qubits >= kMinTileQubits && tile_qubits <= kMaxTileQubits),
                "tile qubits must be in [" + std::to_string(kMinTileQubits) + ", " + std::to_string(kMaxTileQubits) +
                    "] (or negative for the global option)");
        circ->tile_override[dtype == MQA_C64 ? 0 : 1][adjoint ? 1 : 0].store(tile_qubits < 0 ? -1 : tile_qubits);
    });
}

int mqa_circuit_get_tile_qubits(const mqa_circuit* circ, int32_t dtype, int32_t adjoint, int32_t* tile_qubits) {
    return guarded([&] {
        require(circ != nullptr && tile_qubits != nullptr, "NULL argument");
        require(dtype == MQA_C64 || dtype == MQA_C128, "bad dtype");
        *tile_qubits = circ->options(dtype, adjoint != 0).tile_qubits;
    });
}

int mqa_circuit_set_custom_matrices(mqa_circuit* circ, int64_t rows, int64_t stride, const double* table) {
    return guarded([&] {
        require(circ != nullptr, "circuit is NULL");
        require(rows >= 1 && stride >= circ->custom_need && table != nullptr,
                "custom matrix table: need rows >= 1 and a row of at least " + std::to_string(circ->custom_need) +
                    " doubles");
        std::lock_guard<std::mutex> lk(circ->mu);
        circ->custom_table.assign(table, table + rows * stride);
        circ->custom_rows = rows;
        circ->custom_stride = stride;
    });
}

int mqa_circuit_destroy(mqa_circuit* circ) {
    return guarded([&] { delete circ; });
}

int mqa_circuit_schedule(const mqa_circuit* circ_c, int32_t n_qubits, int32_t dtype, int32_t adjoint, int64_t* out,
                         int64_t cap, int64_t* len) {
    return guarded([&] {
        require(circ_c != nullptr && len != nullptr, "NULL argument");
        require(dtype == MQA_C64 || dtype == MQA_C128, "bad dtype");
        require(n_qubits > circ_c->max_qubit && n_qubits <= 40, "bad qubit count");
        require(adjoint >= 0 && (adjoint & ~19) == 0, "adjoint must be 0..3, optionally + 16");
        auto* circ = const_cast<mqa_circuit*>(circ_c);
        DevPlan* dp = circ->get_plan(n_eff_of(n_qubits), dtype, (adjoint & 1) != 0, nullptr, false, (adjoint & 2) != 0,
                                     (adjoint & 16) != 0);
        std::vector<int64_t> w;
        w.push_back(int64_t(dp->plan.passes.size()));
        for (auto& ps : dp->plan.passes) {
            uint64_t L = 0;
            for (int q : ps.local_qubits) L |= uint64_t(1) << q;
            w.push_back(int64_t(L));
            std::vector<int> gs;
            for (auto& g : ps.groups)
                for (auto& op : g.ops) op_gates(op, gs);
            w.push_back(int64_t(gs.size()));
            for (int gi : gs) w.push_back(gi);
        }
        *len = int64_t(w.size());
        if (out && cap >= int64_t(w.size())) std::memcpy(out, w.data(), w.size() * sizeof(int64_t));
    });
}

int mqa_circuit_jit_compile(const mqa_circuit* circ, int32_t n_qubits, int32_t dtype, int32_t adjoint,
                            double* seconds) {
    return guarded([&] {
        require(circ != nullptr, "circuit is NULL");
        require(dtype == MQA_C64 || dtype == MQA_C128, "bad dtype");
        require(n_qubits > circ->max_qubit && n_qubits <= 40, "bad qubit count");
        require(adjoint >= 0 && (adjoint & ~27) == 0, "adjoint must be 0..3, optionally + 8 and + 16");
        const bool adj = (adjoint & 1) != 0, pf = (adjoint & 2) != 0, one_tile = (adjoint & 8) != 0;
        const bool ops = (adjoint & 16) != 0;
        const Plan pl = build_plan(circ->gates, n_eff_of(n_qubits), adj, circ->options(dtype, adj, pf, ops));
        require(!one_tile || jit_can_single_buffer(pl, dtype),
                "single-buffer passes are back-sweeps without dense 3- or 4-qubit blocks");
        JitPlan jp(pl, dtype, one_tile ? JitForm::kSingleBuffer : JitForm::kTile);
        jp.wait();
        if (jp.failed()) throw Error(MQA_ERR_RUNTIME, "code generation failed: " + jp.error());
        if (seconds) *seconds = jp.compile_seconds();
    });
}

int mqa_ham_create(const mqa_pauli_term* terms, int32_t n_terms, mqa_ham** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        require(n_terms >= 0 && (n_terms == 0 || terms != nullptr), "bad term array");
        for (int t = 0; t < n_terms; t++) {
            const auto& tm = terms[t];
            require((tm.x_mask & tm.y_mask) == 0 && (tm.x_mask & tm.z_mask) == 0 && (tm.y_mask & tm.z_mask) == 0,
                    "pauli masks overlap");
        }
        std::unique_ptr<mqa_ham> h = pauli_ham(terms, n_terms);
        std::vector<mqa_pauli_term> rest;
        bool has_identity = false;
        for (int t = 0; t < n_terms; t++) {
            if ((terms[t].x_mask | terms[t].y_mask | terms[t].z_mask) == 0) {
                h->identity += terms[t].coeff;
                has_identity = true;
            } else {
                rest.push_back(terms[t]);
            }
        }
        if (has_identity) h->traceless = pauli_ham(rest.data(), int32_t(rest.size()));
        *out = h.release();
    });
}

int mqa_ham_create_csr(uint64_t dim, uint64_t nnz, const int64_t* indptr, const int64_t* indices, const double* data,
                       mqa_ham** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        require(dim >= 1 && (dim & (dim - 1)) == 0 && dim <= (uint64_t(1) << 36),
                "size of hamiltonian sparse matrix should be power of 2");
        require(indptr != nullptr && (nnz == 0 || (indices != nullptr && data != nullptr)), "NULL argument");
        require(indptr[0] == 0 && uint64_t(indptr[dim]) == nnz, "bad CSR row pointer");
        for (uint64_t r = 0; r < dim; r++) require(indptr[r + 1] >= indptr[r], "bad CSR row pointer");
        for (uint64_t e = 0; e < nnz; e++)
            require(indices[e] >= 0 && uint64_t(indices[e]) < dim, "CSR column index out of range");
        auto h = std::make_unique<mqa_ham>();
        h->csr = true;
        h->csr_dim = dim;
        h->qubit_mask = dim - 1;
        h->csr_ptr[0].assign(indptr, indptr + dim + 1);
        h->csr_idx[0].assign(indices, indices + nnz);
        h->csr_val[0].assign(data, data + 2 * nnz);
        // conjugate transpose (counting sort by column)
        auto& tp = h->csr_ptr[1];
        auto& ti = h->csr_idx[1];
        auto& tv = h->csr_val[1];
        tp.assign(dim + 1, 0);
        for (uint64_t e = 0; e < nnz; e++) tp[size_t(indices[e]) + 1]++;
        for (uint64_t r = 0; r < dim; r++) tp[r + 1] += tp[r];
        ti.assign(nnz, 0);
        tv.assign(2 * nnz, 0.0);
        std::vector<int64_t> fill(tp.begin(), tp.end() - 1);
        for (uint64_t r = 0; r < dim; r++)
            for (int64_t e = indptr[r]; e < indptr[r + 1]; e++) {
                const int64_t at = fill[size_t(indices[e])]++;
                ti[size_t(at)] = int64_t(r);
                tv[2 * size_t(at)] = data[2 * e];
                tv[2 * size_t(at) + 1] = -data[2 * e + 1];
            }
        // hermitian iff every row equals the same row of the conjugate transpose (entries sorted by
        // column, duplicates summed, zeros dropped); exact comparison
        auto row_of = [](const int64_t* ptr, const int64_t* idx, const double* val, uint64_t r) {
            std::vector<std::tuple<int64_t, double, double>> e;
            for (int64_t k = ptr[r]; k < ptr[r + 1]; k++) e.emplace_back(idx[k], val[2 * k], val[2 * k + 1]);
            std::sort(e.begin(), e.end(), [](auto& a, auto& b) { return std::get<0>(a) < std::get<0>(b); });
            std::vector<std::tuple<int64_t, double, double>> m;
            for (auto& [c, re, im] : e) {
                if (!m.empty() && std::get<0>(m.back()) == c) {
                    std::get<1>(m.back()) += re;
                    std::get<2>(m.back()) += im;
                } else {
                    m.emplace_back(c, re, im);
                }
            }
            m.erase(std::remove_if(m.begin(), m.end(),
                                   [](auto& x) { return std::get<1>(x) == 0.0 && std::get<2>(x) == 0.0; }),
                    m.end());
            return m;
        };
        for (uint64_t r = 0; r < dim && h->hermitian; r++)
            h->hermitian = row_of(indptr, indices, data, r) == row_of(tp.data(), ti.data(), tv.data(), r);
        DevHamGroup g{};  // the Pauli tables stay empty
        h->groups.push_back(g);
        *out = h.release();
    });
}

int mqa_ham_destroy(mqa_ham* ham) {
    return guarded([&] { delete ham; });
}

int mqa_sim_create(int32_t n_qubits, int32_t dtype, uint32_t seed, mqa_sim** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        require(dtype == MQA_C64 || dtype == MQA_C128, "dtype must be MQA_C64 or MQA_C128");
        require(n_qubits >= 1 && n_qubits <= 36, "Dimension too large.");
        if (mqa_device_count() <= 0) throw Error(MQA_ERR_HIP, "no HIP device is visible");
        auto s = std::make_unique<mqa_sim>();
        s->n = n_qubits;
        s->n_eff = n_eff_of(n_qubits);
        s->dtype = dtype;
        s->seed = seed;
        s->dim = uint64_t(1) << n_qubits;
        s->dim_eff = uint64_t(1) << s->n_eff;
        s->device = current_device();
        hip_check(hipStreamCreate(&s->stream), "hipStreamCreate");
        s->state.ensure(s->state_bytes());
        hip_check(launch_fill(dtype, s->state.p, nullptr, s->dim_eff, s->dim_eff, 0, 1, s->stream), "init");
        hip_check(hipStreamSynchronize(s->stream), "sync");
        *out = s.release();
    });
}

int mqa_sim_destroy(mqa_sim* sim) {
    return guarded([&] {
        if (!sim) return;
        {
            DeviceGuard dg(sim->device);
            if (sim->stream) (void)hipStreamSynchronize(sim->stream);
            if (sim->stream) (void)hipStreamDestroy(sim->stream);
            sim->stream = nullptr;
        }
        delete sim;
    });
}

int mqa_sim_reset(mqa_sim* sim) {
    return guarded([&] {
        require(sim != nullptr, "sim is NULL");
        DeviceGuard dg(sim->device);
        hip_check(launch_fill(sim->dtype, sim->state.p, nullptr, sim->dim_eff, sim->dim_eff, 0, 1, sim->stream), "reset");
        hip_check(hipStreamSynchronize(sim->stream), "sync");
    });
}

int mqa_sim_copy(const mqa_sim* sim, mqa_sim** out) {
    return guarded([&] {
        require(sim != nullptr && out != nullptr, "NULL argument");
        DeviceGuard dg(sim->device);
        mqa_sim* c = nullptr;
        int st = mqa_sim_create(sim->n, sim->dtype, sim->seed, &c);
        if (st != MQA_OK) throw Error(st, g_last_error);
        hip_check(hipMemcpyAsync(c->state.p, sim->state.p, sim->state_bytes(), hipMemcpyDeviceToDevice, c->stream), "copy");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        *out = c;
    });
}

int mqa_sim_n_qubits(const mqa_sim* sim, int32_t* n_qubits, int32_t* dtype) {
    return guarded([&] {
        require(sim != nullptr, "sim is NULL");
        if (n_qubits) *n_qubits = sim->n;
        if (dtype) *dtype = sim->dtype;
    });
}

int mqa_sim_get_qs(const mqa_sim* sim, void* host_out, uint64_t dim) {
    return guarded([&] {
        require(sim != nullptr && host_out != nullptr, "NULL argument");
        require(dim == sim->dim, "state size not match");
        DeviceGuard dg(sim->device);
        hip_check(hipMemcpyAsync(host_out, sim->state.p, sim->dim * sim->amp_bytes(), hipMemcpyDeviceToHost, sim->stream),
                  "get_qs");
        hip_check(hipStreamSynchronize(sim->stream), "sync");
    });
}

int mqa_sim_set_qs(mqa_sim* sim, const void* host_in, uint64_t dim) {
    return guarded([&] {
        require(sim != nullptr && host_in != nullptr, "NULL argument");
        require(dim == sim->dim, "state size not match");
        DeviceGuard dg(sim->device);
        if (sim->dim_eff != sim->dim)
            hip_check(hipMemsetAsync(sim->state.p, 0, sim->state_bytes(), sim->stream), "set_qs clear");
        hip_check(hipMemcpyAsync(sim->state.p, host_in, sim->dim * sim->amp_bytes(), hipMemcpyHostToDevice, sim->stream),
                  "set_qs");
        hip_check(hipStreamSynchronize(sim->stream), "sync");
    });
}

int mqa_sim_apply_circuit(mqa_sim* sim, const mqa_circuit* circ_c, const double* params, int32_t n_params) {
    return guarded([&] {
        require(sim != nullptr, "sim is NULL");
        check_circuit(sim, circ_c, n_params);
        auto* circ = const_cast<mqa_circuit*>(circ_c);
        if (circ->gates.empty()) return;
        DeviceGuard dg(sim->device);
        // phase-free plan (uncontrolled RX / RY / RZ / GP drop their global phase), then the dropped
        // phase restored by one scaling pass; the angle table is computed here, so the phase is
        // decided on exactly the values the kernels read
        const bool pf = g_opt.phase_free.load() != 0 && circ->n_slots > 0;
        DevPlan* dp = circ->get_plan(sim->n_eff, sim->dtype, false, sim->stream, true, pf, want_phase_ops(circ, 1, sim->dim_eff));
        const double* ang = nullptr;
        std::vector<double> tab;
        std::complex<double> phase(1.0, 0.0);
        if (pf) {
            tab = host_angles(circ, params);
            phase = dropped_phase(circ, tab, sim->dtype);
            sim->angles.ensure(tab.size() * sizeof(double));
            hip_check(hipMemcpyAsync(sim->angles.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice,
                                     sim->stream),
                      "angle upload");
            ang = sim->angles.as<double>();
        } else {
            ang = build_angles(sim, circ, nullptr, 1, 0, params, n_params);
        }
        run_plan(sim, dp, sim->state.p, nullptr, 1, ang, circ->n_slots, nullptr, 0, 0, -1, custom_rows(sim, circ, 0, 1));
        if (phase != std::complex<double>(1.0, 0.0))
            hip_check(launch_scale(sim->dtype, sim->state.p, sim->state.p, sim->dim_eff, phase.real(), phase.imag(),
                                   sim->stream),
                      "global phase");
        sync(sim, "apply_circuit");
    });
}

int mqa_sim_apply_dense_blocks(mqa_sim* sim, int32_t n_blocks, const int32_t* block_qubits, const int32_t* qubits,
                               const double* matrices) {
    return guarded([&] {
        require(sim != nullptr, "sim is NULL");
        require(n_blocks >= 0, "n_blocks must be >= 0");
        if (n_blocks == 0) return;
        require(block_qubits != nullptr && qubits != nullptr && matrices != nullptr, "NULL block arrays");
        // validate everything before anything runs; offsets of each block's qubits and matrix
        std::vector<size_t> q_off(n_blocks), m_off(n_blocks);
        size_t nq = 0, nm = 0;
        for (int b = 0; b < n_blocks; b++) {
            const int k = block_qubits[b];
            require(k >= 1 && k <= kDenseMaxQubits, "a dense block acts on 1 to 5 qubits");
            require(k <= sim->n, "a dense block acts on more qubits than the simulator has");
            uint64_t seen = 0;
            for (int i = 0; i < k; i++) {
                const int q = qubits[nq + i];
                require(q >= 0 && q < sim->n, "dense block qubit out of range");
                require(!(seen >> q & 1), "dense block qubits must be distinct");
                seen |= uint64_t(1) << q;
            }
            q_off[b] = nq, m_off[b] = nm;
            nq += size_t(k);
            nm += size_t(2) << (2 * k);
        }
        DeviceGuard dg(sim->device);
        // matrices in the state's precision, uploaded once for the whole sequence
        const size_t real = sim->dtype == MQA_C64 ? 4 : 8;
        std::vector<unsigned char> host(nm * real);
        if (sim->dtype == MQA_C64) {
            float* f = reinterpret_cast<float*>(host.data());
            for (size_t i = 0; i < nm; i++) f[i] = float(matrices[i]);
        } else {
            std::memcpy(host.data(), matrices, nm * real);
        }
        sim->dense.ensure(host.size());
        hip_check(hipMemcpyAsync(sim->dense.p, host.data(), host.size(), hipMemcpyHostToDevice, sim->stream),
                  "dense matrices upload");
        const bool mfma = g_opt.dense_mfma.load() != 0;
        for (int b = 0; b < n_blocks; b++) {
            const int k = block_qubits[b];
            const void* mat = static_cast<const unsigned char*>(sim->dense.p) + m_off[b] * real;
            const char* name = dense_sweep_uses_mfma(sim->dtype, sim->n_eff, k, mfma) ? "dense_sweep_mfma" : "dense_sweep";
            timed(sim, name, 2.0 * double(sim->state_bytes()), [&] {
                hip_check(launch_dense_sweep(sim->dtype, sim->state.p, sim->n_eff, qubits + q_off[b], k, 0, mat, mfma,
                                             sim->stream),
                          "dense sweep");
            });
        }
        sync(sim, "apply_dense_blocks");
    });
}

int mqa_sim_apply_hamiltonian(mqa_sim* sim, const mqa_ham* ham_c) {
    return guarded([&] {
        require(sim != nullptr, "sim is NULL");
        check_ham(sim, ham_c);
        auto* ham = const_cast<mqa_ham*>(ham_c);
        DeviceGuard dg(sim->device);
        sim->scratch_state.ensure(sim->state_bytes());
        ham_apply(sim, ham, sim->state.p, nullptr, sim->scratch_state.p, 1);
        hip_check(hipMemcpyAsync(sim->state.p, sim->scratch_state.p, sim->state_bytes(), hipMemcpyDeviceToDevice,
                                 sim->stream),
                  "copy");
        sync(sim, "apply_hamiltonian");
    });
}

int mqa_sim_expectation_lr(const mqa_sim* sim_c, const mqa_sim* left_c, const mqa_ham* ham_c,
                           const mqa_circuit* circ_right, const mqa_circuit* circ_left, const double* params,
                           int32_t n_params, double out[2]) {
    return guarded([&] {
        require(sim_c != nullptr && out != nullptr, "NULL argument");
        auto* sim = const_cast<mqa_sim*>(sim_c);
        auto* left = const_cast<mqa_sim*>(left_c);
        check_ham(sim, ham_c);
        if (left) {
            require(left->n == sim->n && left->dtype == sim->dtype, "simulator_left does not match this simulator");
        }
        if (circ_right) check_circuit(sim, circ_right, n_params);
        if (circ_left) check_circuit(sim, circ_left, n_params);
        auto* ham = const_cast<mqa_ham*>(ham_c);
        DeviceGuard dg(sim->device);
        const size_t sb = sim->state_bytes();
        // ket = Ur |sim>, bra = Ul |left> in two workspace states
        sim->psi.ensure(sb);
        sim->lam.ensure(sb);
        hip_check(hipMemcpyAsync(sim->psi.p, sim->state.p, sb, hipMemcpyDeviceToDevice, sim->stream), "copy");
        if (circ_right && !circ_right->gates.empty()) {
            auto* cr = const_cast<mqa_circuit*>(circ_right);
            DevPlan* dp = cr->get_plan(sim->n_eff, sim->dtype, false, sim->stream, true);
            const double* ang = build_angles(sim, cr, nullptr, 1, 0, params, n_params);
            run_plan(sim, dp, sim->psi.p, nullptr, 1, ang, cr->n_slots, nullptr, 0, 0, -1, custom_rows(sim, cr, 0, 1, 0));
        }
        const void* bra = sim->psi.p;
        bool two = (left != nullptr && left != sim) || (circ_left != nullptr && circ_left != circ_right);
        if (two) {
            const mqa_sim* src = left ? left : sim;
            hip_check(hipMemcpyAsync(sim->lam.p, src->state.p, sb, hipMemcpyDeviceToDevice, sim->stream), "copy");
            if (left && left->stream != sim->stream) hip_check(hipStreamSynchronize(left->stream), "sync");
            const mqa_circuit* cl = circ_left ? circ_left : circ_right;
            if (cl && !cl->gates.empty()) {
                auto* clm = const_cast<mqa_circuit*>(cl);
                DevPlan* dp = clm->get_plan(sim->n_eff, sim->dtype, false, sim->stream, true);
                const double* ang = build_angles(sim, clm, nullptr, 1, 0, params, n_params);
                run_plan(sim, dp, sim->lam.p, nullptr, 1, ang, clm->n_slots, nullptr, 0, 0, -1,
                         custom_rows(sim, clm, 0, 1, 1));
            }
            bra = sim->lam.p;
        }
        ham_apply(sim, ham, sim->psi.p, bra, nullptr, 1);
        hip_check(hipMemcpyAsync(out, sim->fsum.p, 2 * sizeof(double), hipMemcpyDeviceToHost, sim->stream), "d2h");
        sync(sim, "expectation");
    });
}

int mqa_sim_expectation(const mqa_sim* sim, const mqa_ham* ham, const mqa_circuit* circ, const double* params,
                        int32_t n_params, double out[2]) {
    return mqa_sim_expectation_lr(sim, nullptr, ham, circ, nullptr, params, n_params, out);
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free expectation
// Device scratch of mqa_sim_expectation_noalloc, one copy per GPU for the life of the process:
// [the reduced (re, im) sum][per-block (re, im) partials][flip groups][terms].  In host code the
// array's name is not a device address: each device's address comes from hipGetSymbolAddress.
namespace mqa {
namespace {
constexpr size_t kNoAllocBytes = size_t(4) << 20;
constexpr size_t kNoAllocPartialsAt = 256;
constexpr uint64_t kNoAllocMaxBlocks = uint64_t(1) << 17;
constexpr size_t kNoAllocGroupsAt = kNoAllocPartialsAt + kNoAllocMaxBlocks * 16;
constexpr size_t kNoAllocMaxGroups = 8192;
constexpr size_t kNoAllocTermsAt = kNoAllocGroupsAt + kNoAllocMaxGroups * sizeof(DevHamGroup);
constexpr size_t kNoAllocMaxTerms = (kNoAllocBytes - kNoAllocTermsAt) / sizeof(DevHamTerm);
static_assert(kNoAllocMaxTerms >= 60000, "scratch layout");
}  // namespace
}  // namespace mqa

static __device__ __align__(256) unsigned char scratch[mqa::kNoAllocBytes];

namespace mqa {
namespace {
// Address of `scratch` on the current device (looked up once per device, with it current).
unsigned char* noalloc_scratch(int device) {
    static std::mutex mu;
    static std::map<int, unsigned char*> ptrs;
    std::lock_guard<std::mutex> lk(mu);
    auto it = ptrs.find(device);
    if (it != ptrs.end()) return it->second;
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    ptrs[device] = static_cast<unsigned char*>(p);
    return static_cast<unsigned char*>(p);
}
}  // namespace
}  // namespace mqa

extern "C" {

int mqa_sim_expectation_noalloc(const mqa_sim* sim_c, const mqa_ham* ham, double out[2]) {
    return guarded([&] {
        require(sim_c != nullptr && ham != nullptr && out != nullptr, "NULL argument");
        auto* sim = const_cast<mqa_sim*>(sim_c);
        check_ham(sim, ham);
        require(!ham->csr, "mqa_sim_expectation_noalloc takes Pauli-term hamiltonians only (a CSR matrix needs a "
                           "workspace state)");
        require(ham->groups.size() <= kNoAllocMaxGroups,
                "hamiltonian has " + std::to_string(ham->groups.size()) + " flip groups; mqa_sim_expectation_noalloc "
                "holds at most " + std::to_string(kNoAllocMaxGroups));
        require(ham->dterms.size() <= kNoAllocMaxTerms,
                "hamiltonian has " + std::to_string(ham->dterms.size()) + " terms; mqa_sim_expectation_noalloc "
                "holds at most " + std::to_string(kNoAllocMaxTerms));
        DeviceGuard dg(sim->device);
        unsigned char* base = noalloc_scratch(sim->device);
        hipStream_t s = sim->stream;
        hip_check(hipMemcpyAsync(base + kNoAllocGroupsAt, ham->groups.data(), ham->groups.size() * sizeof(DevHamGroup),
                                 hipMemcpyHostToDevice, s),
                  "groups to scratch");
        if (!ham->dterms.empty())
            hip_check(hipMemcpyAsync(base + kNoAllocTermsAt, ham->dterms.data(), ham->dterms.size() * sizeof(DevHamTerm),
                                     hipMemcpyHostToDevice, s),
                      "terms to scratch");
        // the gather kernel of mqa_sim_expectation on the state itself (no circuit, so no copy); blocks
        // of 2048 amplitudes, or wider so that the partials fit the scratch
        uint64_t per_block = std::min<uint64_t>(sim->dim_eff, 2048);
        while ((sim->dim_eff + per_block - 1) / per_block > kNoAllocMaxBlocks) per_block *= 2;
        require(per_block <= 0xffffffffull, "state too large for mqa_sim_expectation_noalloc");
        HamArgs h{};
        h.ket = sim->state.p;
        h.ket_stride = sim->dim_eff;
        h.bra = sim->state.p;
        h.bra_stride = sim->dim_eff;
        h.out = nullptr;
        h.out_stride = sim->dim_eff;
        h.dim = sim->dim_eff;
        h.amps_per_block = uint32_t(per_block);
        h.n_blocks = uint32_t((sim->dim_eff + per_block - 1) / per_block);
        h.groups = reinterpret_cast<const DevHamGroup*>(base + kNoAllocGroupsAt);
        h.n_groups = int(ham->groups.size());
        h.terms = reinterpret_cast<const DevHamTerm*>(base + kNoAllocTermsAt);
        h.partial = reinterpret_cast<double*>(base + kNoAllocPartialsAt);  // every block writes its entry
        hip_check(launch_ham(sim->dtype, false, h, 1, s), "hamiltonian");
        hip_check(launch_reduce_rows(h.partial, h.n_blocks, reinterpret_cast<double*>(base), 1, s), "reduce");
        hip_check(hipMemcpyAsync(out, base, 2 * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
        hip_check(hipStreamSynchronize(s), "expectation (no allocation)");
    });
}

int mqa_sim_grad_batch(const mqa_sim* sim_c, const mqa_ham* const* hams, int32_t n_hams, const mqa_circuit* circ_c,
                       const double* enc, int64_t batch, int32_t n_enc, const double* ans, int32_t n_ans, double* f_out,
                       double* g_out) {
    return guarded([&] {
        require(sim_c != nullptr && circ_c != nullptr, "NULL argument");
        require(n_hams >= 1 && hams != nullptr, "need at least one hamiltonian");
        require(batch >= 1, "batch must be positive");
        require(n_enc >= 0 && n_ans >= 0, "bad parameter counts");
        require(n_enc == 0 || enc != nullptr, "enc is NULL");
        require(n_ans == 0 || ans != nullptr, "ans is NULL");
        require(f_out != nullptr && g_out != nullptr, "output is NULL");
        auto* sim = const_cast<mqa_sim*>(sim_c);
        auto* circ = const_cast<mqa_circuit*>(circ_c);
        const int n_params = n_enc + n_ans;
        check_circuit(sim, circ, n_params);
        for (int h = 0; h < n_hams; h++) check_ham(sim, hams[h]);
        DeviceGuard dg(sim->device);
        hipStream_t s = sim->stream;

        // <psi|H|psi> and the taps <lambda|dG|psi> (lambda = H psi) do not see a global phase
        // common to psi and lambda: phase-free plans; their RZ / GP taps also need <lambda|psi>
        // real, i.e. every observable hermitian
        bool pf = g_opt.phase_free.load() != 0;
        for (int h = 0; h < n_hams; h++) pf = pf && hams[h]->hermitian;
        g_pf_used.store(pf ? 1 : 0);
        const bool ops = want_phase_ops(circ, uint64_t(batch), sim->dim_eff);
        DevPlan* fwd = circ->get_plan(sim->n_eff, sim->dtype, false, s, true, pf, ops);
        DevPlan* adj = circ->get_plan(sim->n_eff, sim->dtype, true, s, true, pf, ops);
        DevBuf* sl = circ->slots_on_device(s);
        const int n_tap_ids = circ->n_tap_ids;
        const uint64_t n_tiles = uint64_t(1) << (adj->plan.n_eff - adj->plan.K);
        const size_t sb = sim->state_bytes();

        // samples per chunk from the HBM budget: psi + lam, plus a kept copy of the forward state
        // when several hamiltonians share it.  If one sample with that copy exceeds the budget,
        // the forward sweep is recomputed from the initial state for every further hamiltonian
        // instead (the back-sweep consumes psi, so it must be restored one way or the other).
        // With checkpoints, the forward sweep also stores psi at back-sweep pass boundaries and
        // the back-sweep continues from each stored state (a pointer switch, no copy), so
        // rounding drift of the uncomputed psi cannot build up over more than one segment; as
        // many checkpoints are held as the budget allows for one sample.
        const size_t side = size_t(n_tap_ids) * n_tiles * 8 + size_t(circ->n_slots) * 16 +
                            size_t(n_tap_ids + n_params) * 16 + 4096;
        const size_t budget = hbm_budget(sim->psi.bytes + sim->lam.bytes + sim->psi_keep.bytes + sim->ckpt.bytes);
        int n_ck = 0;
        if (g_opt.checkpoints > 0 && n_tap_ids > 0) {
            n_ck = int(std::min<int64_t>(g_opt.checkpoints, int64_t(adj->plan.passes.size()) - 1));
            while (n_ck > 0 && budget < sb * size_t(2 + n_ck) + side) n_ck--;
        }
        mqa_circuit::Checkpointed* ck = n_ck > 0 ? circ->checkpointed(sim->n_eff, sim->dtype, n_ck, adj, s) : nullptr;
        if (ck) n_ck = int(ck->bound.size());
        g_ckpt_used.store(n_ck);
        const GradWorkspace ws = plan_grad_workspace(budget, sb, side, n_hams, n_ck, batch);
        const bool keep = ws.keep;
        const int64_t chunk = ws.chunk;
        {
            DevBuf* bufs[4] = {&sim->psi, &sim->lam, &sim->psi_keep, &sim->ckpt};
            fit_and_ensure(bufs, ws.bytes, 4, budget);
        }
        sim->partials.ensure(std::max<size_t>(size_t(chunk) * n_tap_ids * n_tiles * 8, 16));
        sim->taps.ensure(std::max<size_t>(size_t(chunk) * n_tap_ids * 8, 16));
        sim->grads.ensure(std::max<size_t>(size_t(chunk) * n_params * 8, 8));
        std::vector<double> g_host(size_t(chunk) * n_params), f_host(size_t(chunk) * 2);
        auto ckpt_at = [&](int i) { return static_cast<char*>(sim->ckpt.p) + size_t(i) * sb * size_t(chunk); };
        double norm0 = -1;  // <state|state> of the initial state, when an identity part needs it

        for (int64_t b0 = 0; b0 < batch; b0 += chunk) {
            const int64_t nb = std::min<int64_t>(chunk, batch - b0);
            const double* ang = build_angles(sim, circ, n_enc ? enc + b0 * n_enc : nullptr, nb, n_enc, ans, n_ans);
            const CustomRows cust = custom_rows(sim, circ, b0, nb);
            // forward sweep into psi (and the checkpoints)
            // the first forward pass reads the initial state itself (every sample starts from it), so
            // psi is not filled first; a plan without passes still needs the copies
            auto forward = [&] {
                if (!ck) {
                    const bool direct = !fwd->plan.passes.empty();
                    if (!direct)
                        hip_check(launch_fill(sim->dtype, sim->psi.p, sim->state.p, sim->dim_eff, sim->dim_eff,
                                              sim->dim_eff, nb, s),
                                  "fill");
                    run_plan(sim, fwd, sim->psi.p, nullptr, nb, ang, circ->n_slots, nullptr, 0, 0, -1, cust,
                             direct ? sim->state.p : nullptr);
                    return;
                }
                // segment k reads the previous one's output and writes its own buffer (its first
                // pass out of place): segment k < n_ck fills checkpoint slot n_ck - 1 - k, the last
                // one psi, so the checkpoints cost no copies
                for (int k = 0; k <= n_ck; k++) {
                    DevPlan* seg = ck->segs[size_t(k)].get();
                    void* out = k < n_ck ? ckpt_at(n_ck - 1 - k) : sim->psi.p;
                    if (seg->plan.passes.empty()) {
                        if (k == 0)
                            hip_check(launch_fill(sim->dtype, out, sim->state.p, sim->dim_eff, sim->dim_eff,
                                                  sim->dim_eff, nb, s),
                                      "fill");
                        else
                            hip_check(hipMemcpyAsync(out, ckpt_at(n_ck - k), sb * nb, hipMemcpyDeviceToDevice, s),
                                      "checkpoint");
                        continue;
                    }
                    run_plan(sim, seg, out, nullptr, nb, ang, circ->n_slots, nullptr, 0, 0, -1, cust,
                             k == 0 ? sim->state.p : ckpt_at(n_ck - k), false, k == 0 ? 0 : sim->dim_eff);
                }
            };
            forward();
            if (keep)
                hip_check(hipMemcpyAsync(sim->psi_keep.p, sim->psi.p, sb * nb, hipMemcpyDeviceToDevice, s), "keep");
            for (int h = 0; h < n_hams; h++) {
                if (h > 0 && keep) {
                    hip_check(hipMemcpyAsync(sim->psi.p, sim->psi_keep.p, sb * nb, hipMemcpyDeviceToDevice, s), "restore");
                } else if (h > 0) {  // recompute the forward state (and the checkpoints)
                    forward();
                    g_recomputes.fetch_add(1);
                }
                double shift = 0;
                mqa_ham* ham = costate_ham(const_cast<mqa_ham*>(hams[h]), circ, &shift);
                // lam = H psi, f = <psi|H psi> (+ the identity part times the initial norm)
                ham_apply(sim, ham, sim->psi.p, sim->psi.p, sim->lam.p, nb);
                hip_check(hipMemcpyAsync(f_host.data(), sim->fsum.p, nb * 2 * sizeof(double), hipMemcpyDeviceToHost, s),
                          "d2h f");
                if (n_tap_ids > 0) {
                    int p0 = 0;
                    void* cur = sim->psi.p;
                    for (int i = 0; i < n_ck; i++) {  // back-sweep segment, then continue from checkpoint i
                        run_plan(sim, adj, cur, sim->lam.p, nb, ang, circ->n_slots, sim->partials.as<double>(),
                                 n_tap_ids, p0, ck->bound[size_t(i)], cust);
                        p0 = ck->bound[size_t(i)];
                        cur = ckpt_at(i);
                    }
                    run_plan(sim, adj, cur, sim->lam.p, nb, ang, circ->n_slots, sim->partials.as<double>(), n_tap_ids,
                             p0, -1, cust, nullptr, true);
                    hip_check(launch_reduce_rows1(sim->partials.as<double>(), n_tiles, sim->taps.as<double>(),
                                                  uint64_t(nb) * n_tap_ids, s),
                              "reduce taps");
                }
                if (n_params > 0) {
                    if (n_tap_ids > 0) {
                        hip_check(launch_jacobian(sim->taps.as<double>(), n_tap_ids, sl[5].as<int32_t>(),
                                                  sl[6].as<int32_t>(), sl[7].as<double>(), n_params,
                                                  sim->grads.as<double>(), nb, s),
                                  "jacobian");
                        hip_check(hipMemcpyAsync(g_host.data(), sim->grads.p, size_t(nb) * n_params * sizeof(double),
                                                 hipMemcpyDeviceToHost, s),
                                  "d2h g");
                    } else {
                        std::fill(g_host.begin(), g_host.end(), 0.0);
                    }
                }
                sync(sim, "grad batch");
                if (shift != 0 && norm0 < 0) norm0 = state_norm2(sim);
                for (int64_t b = 0; b < nb; b++) {
                    double* fo = f_out + ((b0 + b) * n_hams + h) * 2;
                    fo[0] = f_host[2 * b] + (shift != 0 ? shift * norm0 : 0.0);
                    fo[1] = f_host[2 * b + 1];
                    double* go = g_out + ((b0 + b) * n_hams + h) * size_t(n_params);
                    for (int p = 0; p < n_params; p++) go[p] = g_host[size_t(b) * n_params + p];
                }
            }
        }
    });
}

}  // extern "C"

namespace {

// One circuit of an uneven batch and its place in the chunk's device arrays.
struct GradItem {
    int index = 0;  // position in the caller's arrays
    int n_eff = 0;
    uint64_t dim = 0;
    DevPlan* fwd = nullptr;
    DevPlan* adj = nullptr;
    mqa_ham::TilePlan* hp = nullptr;  // NULL: gather kernel
    uint64_t psi = 0, lam = 0;        // amplitude offsets in the chunk's arenas
    uint64_t ang = 0;                 // offset (doubles) in the angle table
    uint64_t taps_part = 0, taps = 0; // offsets in the tap partials / reduced taps
    uint64_t fpart = 0, f_len = 0;    // offset (doubles) and (re, im) pairs of the <psi|H psi> partials
    uint64_t ph_fwd = 0, ph_adj = 0;  // offsets (reals) of its plans' phase tables
    size_t bytes = 0;                 // device bytes this item needs
};

// Row of a plan's phase table for one sample from its angle table (k_phase_coefs on the host).
template <typename T>
void phase_row(const Plan& pl, const double* ang, T* out) {
    for (size_t t = 0; t < pl.phase_coefs.size(); t++) {
        const PhaseCoef& c = pl.phase_coefs[t];
        double x = 1.0, y = 0.0;
        if (c.slot >= 0) {
            const double ca = ang[2 * c.slot], sa = ang[2 * c.slot + 1];
            const int k = c.k < 0 ? -c.k : c.k;
            if (k == 1) x = ca, y = sa;
            else if (k == 2) x = ca * ca - sa * sa, y = 2.0 * ca * sa;
            if (c.k < 0) y = -y;
        }
        out[2 * t] = T(x * c.cos_b - y * c.sin_b);
        out[2 * t + 1] = T(x * c.sin_b + y * c.cos_b);
    }
}

struct StreamOwner {
    hipStream_t s = nullptr;
    StreamOwner() { hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"); }
    ~StreamOwner() {
        if (s) (void)hipStreamDestroy(s);
    }
};

// Launch groups of one pass step: items that share the launch shape.
template <typename A>
struct StepGroup {
    int K = 0, C = 0, max_taps = 0;
    bool full = false, phase = false, first = false, last = false;
    std::vector<A> args;
    std::vector<uint64_t> prefix;
    uint64_t blocks = 0;
    size_t args_at = 0, prefix_at = 0;  // element offsets in the uploaded tables
};

void run_grad_chunk(std::vector<GradItem>& items, const mqa_circuit* const* circs, const mqa_ham* const* hams,
                    const double* params, const int64_t* param_offset, int dtype, hipStream_t s, double* f_out,
                    double* g_out) {
    const size_t amp = dtype == MQA_C64 ? 8 : 16;
    uint64_t amps = 0, n_ang = 0, n_part = 0, n_taps = 0, n_fpart = 0;
    for (auto& it : items) {
        const mqa_circuit* c = circs[it.index];
        it.psi = amps;
        it.lam = amps + it.dim;
        amps += 2 * it.dim;
        it.ang = n_ang;
        n_ang += uint64_t(c->n_slots) * 2;
        const uint64_t adj_tiles = uint64_t(1) << (it.n_eff - it.adj->plan.K);
        it.taps_part = n_part;
        n_part += uint64_t(c->n_tap_ids) * adj_tiles;
        it.taps = n_taps;
        n_taps += uint64_t(c->n_tap_ids);
        it.fpart = n_fpart;
        it.f_len = it.hp ? (uint64_t(1) << (it.n_eff - it.hp->plan.K))
                         : (it.dim + 2047) / 2048;  // gather kernel: one partial per 2048 amplitudes
        n_fpart += 2 * it.f_len;
    }
    DevBuf arena, ang_d, part_d, taps_d, fpart_d, table_d;
    arena.ensure(amps * amp);
    part_d.ensure(std::max<uint64_t>(n_part, 1) * 8);
    taps_d.ensure(std::max<uint64_t>(n_taps, 1) * 8);
    fpart_d.ensure(std::max<uint64_t>(n_fpart, 1) * 8);

    // angle table, computed on the host (a few doubles per gate)
    std::vector<double> ang(std::max<uint64_t>(n_ang, 1));
    for (auto& it : items) {
        const mqa_circuit* c = circs[it.index];
        const double* pv = params + param_offset[it.index];
        for (int sl = 0; sl < c->n_slots; sl++) {
            double th = c->slot_const[size_t(sl)];
            for (int k = c->slot_term_begin[size_t(sl)]; k < c->slot_term_begin[size_t(sl) + 1]; k++)
                th += c->slot_term_coeff[size_t(k)] * pv[c->slot_term_param[size_t(k)]];
            if (c->slot_half[size_t(sl)]) th *= 0.5;
            ang[it.ang + 2 * uint64_t(sl)] = std::cos(th);
            ang[it.ang + 2 * uint64_t(sl) + 1] = std::sin(th);
        }
    }
    upload(ang_d, ang, s);
    // phase tables of the items' plans (phase blocks), in the state's real type
    DevBuf ph_d;
    {
        uint64_t n_ph = 0;
        for (auto& it : items) {
            it.ph_fwd = n_ph;
            n_ph += 2 * it.fwd->plan.phase_coefs.size();
            it.ph_adj = n_ph;
            n_ph += 2 * it.adj->plan.phase_coefs.size();
        }
        std::vector<float> pf32;
        std::vector<double> pf64;
        if (dtype == MQA_C64) pf32.assign(std::max<uint64_t>(n_ph, 1), 0.0f);
        else pf64.assign(std::max<uint64_t>(n_ph, 1), 0.0);
        for (auto& it : items) {
            const double* a = ang.data() + it.ang;
            if (dtype == MQA_C64) {
                phase_row(it.fwd->plan, a, pf32.data() + it.ph_fwd);
                phase_row(it.adj->plan, a, pf32.data() + it.ph_adj);
            } else {
                phase_row(it.fwd->plan, a, pf64.data() + it.ph_fwd);
                phase_row(it.adj->plan, a, pf64.data() + it.ph_adj);
            }
        }
        if (dtype == MQA_C64) upload(ph_d, pf32, s);
        else upload(ph_d, pf64, s);
        hip_check(hipStreamSynchronize(s), "phase tables");  // the host tables go out of scope
    }

    // |0...0> for every item
    std::vector<uint64_t> unit;
    for (auto& it : items) unit.push_back(it.psi);
    DevBuf unit_d;
    upload(unit_d, unit, s);
    hip_check(hipMemsetAsync(arena.p, 0, amps * amp, s), "memset");
    hip_check(launch_unit_items(dtype, arena.p, unit_d.as<uint64_t>(), int(unit.size()), s), "unit");

    char* base = static_cast<char*>(arena.p);
    auto tile_args = [&](const GradItem& it, DevPlan* dp, int pass, bool adjoint) {
        TilePassArgs a{};
        a.psi = base + it.psi * amp;
        a.lam = adjoint ? base + it.lam * amp : nullptr;
        a.K = dp->plan.K;
        a.pass_index = pass;
        a.stride = it.dim;
        a.n_tiles = uint64_t(1) << (it.n_eff - dp->plan.K);
        a.passes = dp->passes.as<DevPass>();
        a.groups = dp->groups.as<DevGroup>();
        a.ops = dp->ops.as<DevOp>();
        a.pool = dp->pool.p;
        a.angles = ang_d.as<double>() + it.ang;
        a.n_slots = circs[it.index]->n_slots;
        a.n_tap_ids = circs[it.index]->n_tap_ids;
        a.tap_ids = dp->tap_ids.as<int32_t>();
        a.partials = adjoint ? part_d.as<double>() + it.taps_part : nullptr;
        if (!dp->plan.dev_terms.empty()) {
            a.terms = dp->terms.as<DevPhaseTerm>();
            a.phase = static_cast<char*>(ph_d.p) + (adjoint ? it.ph_adj : it.ph_fwd) * (amp / 2);
            a.n_phase = int32_t(dp->plan.dev_terms.size());
        }
        return a;
    };

    // every step of the forward sweep, the Hamiltonian passes and the back-sweep: tables first
    std::vector<std::vector<StepGroup<TilePassArgs>>> fwd_steps, adj_steps;
    std::vector<std::vector<StepGroup<HamTileArgs>>> ham_steps;
    auto build_tile_steps = [&](bool adjoint, std::vector<std::vector<StepGroup<TilePassArgs>>>& steps) {
        size_t max_p = 0;
        for (auto& it : items) max_p = std::max(max_p, (adjoint ? it.adj : it.fwd)->plan.passes.size());
        for (size_t st = 0; st < max_p; st++) {
            std::map<std::tuple<int, int, int, int>, StepGroup<TilePassArgs>> by;
            for (auto& it : items) {
                DevPlan* dp = adjoint ? it.adj : it.fwd;
                if (st >= dp->plan.passes.size()) continue;
                if (adjoint && circs[it.index]->n_tap_ids == 0) continue;  // nothing to differentiate
                auto& g = by[std::make_tuple(dp->plan.K, dp->plan.C, int(dp->full), int(dp->phase))];
                g.K = dp->plan.K;
                g.C = dp->plan.C;
                g.full = dp->full;
                g.phase = dp->phase;
                g.max_taps = std::max(g.max_taps, std::max(dp->plan.max_taps, 1));
                g.prefix.push_back(g.blocks);
                g.args.push_back(tile_args(it, dp, int(st), adjoint));
                g.blocks += uint64_t(1) << (it.n_eff - dp->plan.K);
            }
            std::vector<StepGroup<TilePassArgs>> v;
            for (auto& [k, g] : by) v.push_back(std::move(g));
            steps.push_back(std::move(v));
        }
    };
    build_tile_steps(false, fwd_steps);
    build_tile_steps(true, adj_steps);
    {
        size_t max_p = 0;
        for (auto& it : items)
            if (it.hp) max_p = std::max(max_p, it.hp->plan.passes.size());
        for (size_t st = 0; st < max_p; st++) {
            std::map<std::tuple<int, int, int, int>, StepGroup<HamTileArgs>> by;
            for (auto& it : items) {
                if (!it.hp || st >= it.hp->plan.passes.size()) continue;
                const HamTilePlan& pl = it.hp->plan;
                const bool first = st == 0, last = st + 1 == pl.passes.size();
                auto& g = by[std::make_tuple(pl.K, pl.C, int(first), int(last))];
                g.K = pl.K;
                g.C = pl.C;
                g.first = first;
                g.last = last;
                HamTileArgs h{};
                h.ket = base + it.psi * amp;
                h.out = base + it.lam * amp;
                h.stride = it.dim;
                h.K = pl.K;
                h.pass_index = int(st);
                h.passes = it.hp->passes.as<DevHamPass>();
                h.groups = it.hp->groups.as<DevHamTileGroup>();
                h.terms = it.hp->terms.as<DevHamTileTerm>();
                h.partial = last ? fpart_d.as<double>() + it.fpart : nullptr;
                h.n_tiles = uint64_t(1) << (it.n_eff - pl.K);
                g.prefix.push_back(g.blocks);
                g.args.push_back(h);
                g.blocks += h.n_tiles;
            }
            std::vector<StepGroup<HamTileArgs>> v;
            for (auto& [k, g] : by) v.push_back(std::move(g));
            ham_steps.push_back(std::move(v));
        }
    }
    // one upload of every table: [tile args][ham args][prefixes]
    std::vector<TilePassArgs> targs;
    std::vector<HamTileArgs> hargs;
    std::vector<uint64_t> prefix;
    for (auto* steps : {&fwd_steps, &adj_steps})
        for (auto& v : *steps)
            for (auto& g : v) {
                g.args_at = targs.size();
                targs.insert(targs.end(), g.args.begin(), g.args.end());
                g.prefix_at = prefix.size();
                prefix.insert(prefix.end(), g.prefix.begin(), g.prefix.end());
            }
    for (auto& v : ham_steps)
        for (auto& g : v) {
            g.args_at = hargs.size();
            hargs.insert(hargs.end(), g.args.begin(), g.args.end());
            g.prefix_at = prefix.size();
            prefix.insert(prefix.end(), g.prefix.begin(), g.prefix.end());
        }
    const size_t t_bytes = targs.size() * sizeof(TilePassArgs), h_bytes = hargs.size() * sizeof(HamTileArgs);
    const size_t h_at = (t_bytes + 255) & ~size_t(255), p_at = (h_at + h_bytes + 255) & ~size_t(255);
    std::vector<char> table(p_at + prefix.size() * sizeof(uint64_t) + 8);
    std::memcpy(table.data(), targs.data(), t_bytes);
    std::memcpy(table.data() + h_at, hargs.data(), h_bytes);
    std::memcpy(table.data() + p_at, prefix.data(), prefix.size() * sizeof(uint64_t));
    upload(table_d, table, s);
    char* tab = static_cast<char*>(table_d.p);
    const auto* targs_d = reinterpret_cast<const TilePassArgs*>(tab);
    const auto* hargs_d = reinterpret_cast<const HamTileArgs*>(tab + h_at);
    const auto* prefix_d = reinterpret_cast<const uint64_t*>(tab + p_at);

    // kernel timing (option "profile"): HIP events on this call's stream, with the algorithmic bytes
    // of each shared launch (every item's state read and written once per pass; psi and lambda in
    // back-sweep passes; the Hamiltonian passes read the ket and write H ket)
    Timer timer;
    auto prof = [&](const char* name, double bytes, auto&& fn) {
        if (!g_opt.profile) {
            fn();
            return;
        }
        hipEvent_t a = timer.get(), b = timer.get();
        hip_check(hipEventRecord(a, s), "hipEventRecord");
        fn();
        hip_check(hipEventRecord(b, s), "hipEventRecord");
        timer.pending.push_back({a, b, name, bytes});
    };
    auto run_tile_steps = [&](const std::vector<std::vector<StepGroup<TilePassArgs>>>& steps, bool adjoint) {
        for (auto& v : steps)
            for (auto& g : v) {
                const size_t lds = interpreted_pass_lds_bytes(dtype, adjoint, g.full, g.K, g.C, g.max_taps);
                double bytes = 0;
                for (const auto& a : g.args) bytes += double(a.stride) * double(amp) * (adjoint ? 4.0 : 2.0);
                prof(adjoint ? "items_tile_pass_adj" : "items_tile_pass_fwd", bytes, [&] {
                    hip_check(launch_tile_pass_items(dtype, adjoint, g.full, g.phase, targs_d + g.args_at,
                                                     prefix_d + g.prefix_at, int(g.args.size()), g.blocks, g.K, lds, s),
                              "tile pass (items)");
                });
            }
    };
    run_tile_steps(fwd_steps, false);
    for (auto& v : ham_steps)
        for (auto& g : v) {
            double bytes = 0;
            for (const auto& h : g.args) bytes += double(h.stride) * double(amp) * 2.0;
            prof("items_hamiltonian", bytes, [&] {
                hip_check(launch_ham_tile_items(dtype, g.first, g.last, hargs_d + g.args_at, prefix_d + g.prefix_at,
                                                int(g.args.size()), g.blocks, g.K, g.C, s),
                          "hamiltonian (items)");
            });
        }
    for (auto& it : items) {  // flip masks wider than a tile: gather kernel, one launch per item
        if (it.hp) continue;
        auto* ham = const_cast<mqa_ham*>(hams[it.index]);
        DevBuf* hd = ham->on_device(s);
        HamArgs h{};
        h.ket = base + it.psi * amp;
        h.ket_stride = it.dim;
        h.bra = h.ket;
        h.bra_stride = it.dim;
        h.out = base + it.lam * amp;
        h.out_stride = it.dim;
        h.dim = it.dim;
        h.amps_per_block = 2048;
        h.n_blocks = uint32_t(it.f_len);
        h.groups = hd[0].as<DevHamGroup>();
        h.n_groups = int(ham->groups.size());
        h.terms = hd[1].as<DevHamTerm>();
        h.partial = fpart_d.as<double>() + it.fpart;
        hip_check(launch_ham(dtype, true, h, 1, s), "hamiltonian (gather)");
    }
    run_tile_steps(adj_steps, true);

    // taps: one segment per (item, tap) over that item's tiles
    std::vector<uint64_t> seg;  // [begin..., len...]
    for (auto& it : items) {
        const uint64_t tiles = uint64_t(1) << (it.n_eff - it.adj->plan.K);
        for (int t = 0; t < circs[it.index]->n_tap_ids; t++) seg.push_back(it.taps_part + uint64_t(t) * tiles);
    }
    const size_t n_seg = seg.size();
    for (auto& it : items) {
        const uint64_t tiles = uint64_t(1) << (it.n_eff - it.adj->plan.K);
        for (int t = 0; t < circs[it.index]->n_tap_ids; t++) seg.push_back(tiles);
    }
    DevBuf seg_d;
    upload(seg_d, seg, s);
    hip_check(launch_reduce_segments(part_d.as<double>(), seg_d.as<uint64_t>(), seg_d.as<uint64_t>() + n_seg,
                                     taps_d.as<double>(), n_seg, s),
              "reduce taps");
    std::vector<double> taps_h(std::max<uint64_t>(n_taps, 1)), fpart_h(std::max<uint64_t>(n_fpart, 1));
    hip_check(hipMemcpyAsync(taps_h.data(), taps_d.p, n_taps * 8, hipMemcpyDeviceToHost, s), "d2h taps");
    hip_check(hipMemcpyAsync(fpart_h.data(), fpart_d.p, n_fpart * 8, hipMemcpyDeviceToHost, s), "d2h f");
    hip_check(hipStreamSynchronize(s), "grad items");
    timer.resolve();

    // f = sum of the partials; g = jacobian^T * 2 Re(taps) (gradient.h:124-133, 183-188)
    for (auto& it : items) {
        const mqa_circuit* c = circs[it.index];
        double fr = 0, fi = 0;
        for (uint64_t k = 0; k < it.f_len; k++) fr += fpart_h[it.fpart + 2 * k], fi += fpart_h[it.fpart + 2 * k + 1];
        f_out[2 * it.index] = fr;
        f_out[2 * it.index + 1] = fi;
        double* go = g_out + param_offset[it.index];
        for (int p = 0; p < c->n_params; p++) {
            double acc = 0;
            for (int k = c->jac_begin[size_t(p)]; k < c->jac_begin[size_t(p) + 1]; k++)
                acc += c->jac_coeff[size_t(k)] * 2.0 * taps_h[it.taps + uint64_t(c->jac_tap[size_t(k)])];
            go[p] = acc;
        }
    }
}

}  // namespace

extern "C" {

int mqa_grad_items(int32_t n_items, const mqa_circuit* const* circs, const mqa_ham* const* hams,
                   const int32_t* n_qubits, const double* params, const int64_t* param_offset, int32_t dtype,
                   double* f_out, double* g_out) {
    return guarded([&] {
        require(n_items >= 0, "n_items must be >= 0");
        if (n_items == 0) return;
        require(circs && hams && n_qubits && param_offset && f_out && g_out, "NULL argument");
        require(dtype == MQA_C64 || dtype == MQA_C128, "dtype must be MQA_C64 or MQA_C128");
        if (mqa_device_count() <= 0) throw Error(MQA_ERR_HIP, "no HIP device is visible");
        StreamOwner so;
        const size_t amp = dtype == MQA_C64 ? 8 : 16;
        std::vector<GradItem> all;
        std::vector<const mqa_ham*> costate(size_t(n_items), nullptr);  // the co-state observable of each item
        std::vector<double> shift(size_t(n_items), 0.0);
        // large items run one by one on their compiled passes (a launch of one such state fills the
        // device, and the compiled passes outrun the shared interpreted ones); the rest share launches
        const int64_t solo_q = g_opt.jit != 0 ? g_opt.items_jit_qubits.load() : 0;
        std::vector<int> solo;
        for (int i = 0; i < n_items; i++) {
            auto* c = const_cast<mqa_circuit*>(circs[i]);
            require(c != nullptr && hams[i] != nullptr, "item " + std::to_string(i) + ": NULL circuit or hamiltonian");
            const mqa_ham* h = costate_ham(const_cast<mqa_ham*>(hams[i]), c, &shift[size_t(i)]);
            costate[size_t(i)] = h;
            const int n = n_qubits[i];
            require(n >= 1 && n <= 36, "item " + std::to_string(i) + ": bad qubit count");
            require(c->max_qubit < n, "item " + std::to_string(i) + ": circuit is wider than its qubit count");
            require(!h->csr, "item " + std::to_string(i) + ": sparse (CSR) hamiltonians are not supported in item batches");
            require(c->custom_need == 0, "item " + std::to_string(i) + ": parameterised matrix gates are not supported "
                                                                         "in item batches");
            require((h->qubit_mask >> n) == 0, "item " + std::to_string(i) + ": hamiltonian is wider than its qubit count");
            require(c->n_params == 0 || params != nullptr, "params is NULL");
            if (solo_q > 0 && n >= solo_q) {
                // on their own only once their compiled passes are ready (jit 2 compiles them here;
                // jit 1 in the background, the item sharing the interpreted launches meanwhile) -
                // the plans the batched gradient of one sample runs
                const bool pf1 = g_opt.phase_free.load() != 0 && h->hermitian && hams[i]->hermitian;
                const bool ops1 = want_phase_ops(c, 1, uint64_t(1) << n_eff_of(n));
                DevPlan* f1 = c->get_plan(n_eff_of(n), dtype, false, so.s, true, pf1, ops1);
                DevPlan* a1 = c->get_plan(n_eff_of(n), dtype, true, so.s, true, pf1, ops1);
                const bool rf = jit_ready(f1, dtype), ra = jit_ready(a1, dtype);
                if (rf && ra) {
                    solo.push_back(i);
                    continue;
                }
            }
            GradItem it;
            it.index = i;
            it.n_eff = n_eff_of(n);
            it.dim = uint64_t(1) << it.n_eff;
            // Pauli observables (the only kind items take) are hermitian: phase-free plans
            const bool pf = g_opt.phase_free.load() != 0;
            // the item passes always run interpreted: phase blocks unless the option is off or the
            // circuit's plans need the FULL instantiation
            const bool ops = g_opt.phase_ops != 0 && !c->dense_ops;
            it.fwd = c->get_plan(it.n_eff, dtype, false, so.s, true, pf, ops);
            it.adj = c->get_plan(it.n_eff, dtype, true, so.s, true, pf, ops);
            if (g_opt.ham_engine) {
                auto* tp = ham_plan_for(const_cast<mqa_ham*>(h), it.n_eff, dtype, so.s);
                it.hp = tp->ok ? tp : nullptr;
            }
            const uint64_t adj_tiles = uint64_t(1) << (it.n_eff - it.adj->plan.K);
            it.bytes = 2 * it.dim * amp + uint64_t(c->n_tap_ids) * adj_tiles * 8 + uint64_t(c->n_tap_ids) * 8 +
                       uint64_t(c->n_slots) * 16 + 2 * ((it.dim + 2047) / 2048 + 1) * 8 + 4 * sizeof(TilePassArgs) *
                       (it.fwd->plan.passes.size() + it.adj->plan.passes.size() + 4);
            all.push_back(it);
        }
        // the solo items: the batched gradient of one sample on a simulator of their size (one per
        // size, released at the end of the call; the compiled passes stay with the circuits)
        {
            std::map<int, std::unique_ptr<mqa_sim, int (*)(mqa_sim*)>> sims;
            for (int i : solo) {
                const int n = n_qubits[i];
                auto it = sims.find(n);
                if (it == sims.end()) {
                    mqa_sim* s = nullptr;
                    if (mqa_sim_create(n, dtype, 0, &s) != MQA_OK) throw Error(MQA_ERR_HIP, mqa_last_error());
                    s->force_jit = true;
                    it = sims.emplace(n, std::unique_ptr<mqa_sim, int (*)(mqa_sim*)>(s, &mqa_sim_destroy)).first;
                }
                const mqa_circuit* c = circs[i];
                const mqa_ham* h = hams[i];
                const int st = mqa_sim_grad_batch(it->second.get(), &h, 1, c, nullptr, 1, 0,
                                                  c->n_params ? params + param_offset[i] : nullptr, c->n_params,
                                                  f_out + 2 * i, g_out + param_offset[i]);
                if (st != MQA_OK) throw Error(st, "item " + std::to_string(i) + ": " + mqa_last_error());
                shift[size_t(i)] = 0.0;  // the batched gradient adds the identity part itself
            }
        }
        // chunks that fit the HBM budget, largest states first
        std::stable_sort(all.begin(), all.end(), [](const GradItem& a, const GradItem& b) { return a.n_eff > b.n_eff; });
        const size_t budget = hbm_budget();
        for (size_t a = 0; a < all.size();) {
            size_t b = a, used = 0;
            while (b < all.size() && (b == a || used + all[b].bytes <= budget)) used += all[b++].bytes;
            std::vector<GradItem> chunk(all.begin() + long(a), all.begin() + long(b));
            run_grad_chunk(chunk, circs, costate.data(), params, param_offset, dtype, so.s, f_out, g_out);
            a = b;
        }
        for (int i = 0; i < n_items; i++) f_out[2 * i] += shift[size_t(i)];  // every item starts from |0...0>
    });
}

int mqa_sim_grad_nonherm(const mqa_sim* sim_c, const mqa_sim* left_c, const mqa_ham* const* hams, int32_t n_hams,
                         const mqa_circuit* circ_r_c, const mqa_circuit* circ_l_c, const double* enc, int64_t batch,
                         int32_t n_enc, const double* ans, int32_t n_ans, double* f_out, double* g_out) {
    return guarded([&] {
        require(sim_c != nullptr && circ_r_c != nullptr && circ_l_c != nullptr, "NULL argument");
        require(n_hams >= 1 && hams != nullptr, "need at least one hamiltonian");
        require(batch >= 1, "batch must be positive");
        require(n_enc >= 0 && n_ans >= 0, "bad parameter counts");
        require(n_enc == 0 || enc != nullptr, "enc is NULL");
        require(n_ans == 0 || ans != nullptr, "ans is NULL");
        require(f_out != nullptr && g_out != nullptr, "output is NULL");
        auto* sim = const_cast<mqa_sim*>(sim_c);
        auto* left = const_cast<mqa_sim*>(left_c ? left_c : sim_c);
        require(left->n == sim->n && left->dtype == sim->dtype, "simulator_left does not match this simulator");
        auto* cr = const_cast<mqa_circuit*>(circ_r_c);
        auto* cl = const_cast<mqa_circuit*>(circ_l_c);
        const int n_params = n_enc + n_ans;
        check_circuit(sim, cr, n_params);
        check_circuit(sim, cl, n_params);
        for (int h = 0; h < n_hams; h++) check_ham(sim, hams[h]);
        DeviceGuard dg(sim->device);
        hipStream_t s = sim->stream;
        if (left->stream != s) hip_check(hipStreamSynchronize(left->stream), "sync");
        const size_t sb = sim->state_bytes();
        const uint64_t dim = sim->dim_eff;

        // f = <phi_l|H|psi_r> with phi_l = Ul|left>, psi_r = Ur|sim>;
        // df/dp = <phi_l|H|d psi_r> + conj(<H psi_r|d phi_l>)  (gradient.h:199-268, without the
        // hermitian doubling).  Each term is a back-sweep with complex taps; the kernels accumulate
        // Re<lam|dG|psi>, and Re<i lam|dG|psi> = Im<lam|dG|psi>, so every sweep runs as a
        // two-state batch with co-states (lam, i lam).  Buffers: psi = [psi_r, psi_r, phi_l, phi_l],
        // lam = [H phi_l, i H phi_l, H psi_r, i H psi_r].
        sim->psi.ensure(sb * 4);
        sim->lam.ensure(sb * 4);
        char* P = static_cast<char*>(sim->psi.p);
        char* L = static_cast<char*>(sim->lam.p);
        struct Side {
            mqa_circuit* c;
            DevPlan* fwd;
            DevPlan* adj;
            uint64_t n_tiles;
            DevBuf ang, part, taps;
        };
        Side side[2];
        side[0].c = cr;
        side[1].c = cl;
        for (auto& sd : side) {
            sd.fwd = sd.c->get_plan(sim->n_eff, sim->dtype, false, s, true);
            sd.adj = sd.c->get_plan(sim->n_eff, sim->dtype, true, s, true);
            sd.n_tiles = uint64_t(1) << (sd.adj->plan.n_eff - sd.adj->plan.K);
            sd.part.ensure(std::max<size_t>(2 * size_t(sd.c->n_tap_ids) * sd.n_tiles * 8, 16));
            sd.taps.ensure(std::max<size_t>(2 * size_t(sd.c->n_tap_ids) * 8, 16));
        }
        std::vector<double> enc2(size_t(2) * std::max(n_enc, 1));
        std::vector<double> taps_h[2];
        double f_h[2];
        for (int64_t b = 0; b < batch; b++) {
            // the same angles for both states of a pair
            for (int k = 0; k < n_enc; k++) enc2[size_t(k)] = enc2[size_t(n_enc + k)] = enc[b * n_enc + k];
            const double* ang[2];
            for (int i = 0; i < 2; i++)
                ang[i] = build_angles(sim, side[i].c, n_enc ? enc2.data() : nullptr, 2, n_enc, ans, n_ans, &side[i].ang);
            CustomRows cust[2];
            for (int i = 0; i < 2; i++) cust[i] = custom_rows(sim, side[i].c, b, 1, i, true);
            // forward states: psi_r twice (slots 0, 1), phi_l twice (slots 2, 3)
            hip_check(launch_fill(sim->dtype, P, sim->state.p, dim, dim, dim, 2, s), "fill");
            run_plan(sim, side[0].fwd, P, nullptr, 2, ang[0], cr->n_slots, nullptr, 0, 0, -1, cust[0]);
            hip_check(launch_fill(sim->dtype, P + 2 * sb, left->state.p, dim, dim, dim, 2, s), "fill");
            run_plan(sim, side[1].fwd, P + 2 * sb, nullptr, 2, ang[1], cl->n_slots, nullptr, 0, 0, -1, cust[1]);
            for (int h = 0; h < n_hams; h++) {
                // the back-sweeps of the previous hamiltonian consumed the forward pairs: redo them
                for (int i = 0; h > 0 && i < 2; i++) {
                    if (side[i].c->n_tap_ids == 0) continue;
                    char* psi_pair = P + size_t(2 * i) * sb;
                    hip_check(launch_fill(sim->dtype, psi_pair, (i == 0 ? sim : left)->state.p, dim, dim, dim, 2, s),
                              "fill");
                    run_plan(sim, side[i].fwd, psi_pair, nullptr, 2, ang[i], side[i].c->n_slots, nullptr, 0, 0, -1,
                             cust[i]);
                }
                auto* ham = const_cast<mqa_ham*>(hams[h]);
                ham_apply(sim, ham, P + 2 * sb, nullptr, L, 1, true);  // H^dagger phi_l
                ham_apply(sim, ham, P, nullptr, L + 2 * sb, 1);        // H psi_r
                hip_check(launch_scale(sim->dtype, L, L + sb, dim, 0.0, 1.0, s), "scale");
                hip_check(launch_scale(sim->dtype, L + 2 * sb, L + 3 * sb, dim, 0.0, 1.0, s), "scale");
                vdot(sim, L, P, f_h);  // <H^dagger phi_l|psi_r> = <phi_l|H|psi_r> (synchronises)
                for (int i = 0; i < 2; i++) {  // the back-sweeps (they consume the forward pairs)
                    if (side[i].c->n_tap_ids == 0) {
                        taps_h[i].clear();
                        continue;
                    }
                    char* psi_pair = P + size_t(2 * i) * sb;
                    char* lam_pair = L + size_t(2 * i) * sb;  // right sweep: H phi_l; left: H psi_r
                    run_plan(sim, side[i].adj, psi_pair, lam_pair, 2, ang[i], side[i].c->n_slots,
                             side[i].part.as<double>(), side[i].c->n_tap_ids, 0, -1, cust[i]);
                    hip_check(launch_reduce_rows1(side[i].part.as<double>(), side[i].n_tiles, side[i].taps.as<double>(),
                                                  uint64_t(2) * side[i].c->n_tap_ids, s),
                              "reduce taps");
                    taps_h[i].assign(size_t(2) * side[i].c->n_tap_ids, 0.0);
                    hip_check(hipMemcpyAsync(taps_h[i].data(), side[i].taps.p, taps_h[i].size() * 8,
                                             hipMemcpyDeviceToHost, s),
                              "d2h taps");
                }
                sync(sim, "grad nonherm");
                double* fo = f_out + (b * n_hams + h) * 2;
                fo[0] = f_h[0];
                fo[1] = f_h[1];
                // jacobian: right taps enter as z, left taps as conj(z); z = taps[0][t] + i taps[1][t]
                double* go = g_out + (b * n_hams + h) * size_t(n_params) * 2;
                for (int p = 0; p < n_params; p++) {
                    double gr = 0, gi = 0;
                    for (int i = 0; i < 2; i++) {
                        const mqa_circuit* c = side[i].c;
                        if (taps_h[i].empty()) continue;
                        const size_t nt = size_t(c->n_tap_ids);
                        for (int k = c->jac_begin[size_t(p)]; k < c->jac_begin[size_t(p) + 1]; k++) {
                            const size_t t = size_t(c->jac_tap[size_t(k)]);
                            const double co = c->jac_coeff[size_t(k)];
                            gr += co * taps_h[i][t];
                            gi += co * (i == 0 ? taps_h[i][nt + t] : -taps_h[i][nt + t]);
                        }
                    }
                    go[2 * p] = gr;
                    go[2 * p + 1] = gi;
                }
            }
        }
    });
}

int mqa_sim_sample(const mqa_sim* sim_c, const mqa_circuit* circ_c, const double* params, int32_t n_params,
                   const double* uniforms, uint64_t shots, uint64_t* out_index) {
    return guarded([&] {
        require(sim_c != nullptr && out_index != nullptr, "NULL argument");
        require(shots == 0 || uniforms != nullptr, "uniforms is NULL");
        auto* sim = const_cast<mqa_sim*>(sim_c);
        if (circ_c) check_circuit(sim, circ_c, n_params);
        for (uint64_t i = 0; i < shots; i++) {
            require(uniforms[i] >= 0.0 && uniforms[i] < 1.0, "uniform samples must lie in [0, 1)");
            require(i == 0 || uniforms[i] >= uniforms[i - 1], "uniform samples must be sorted");
        }
        if (shots == 0) return;
        DeviceGuard dg(sim->device);
        hipStream_t s = sim->stream;
        const size_t sb = sim->state_bytes();
        // the circuit runs on a copy: the simulator's own state is left unchanged
        sim->psi.ensure(sb);
        hip_check(hipMemcpyAsync(sim->psi.p, sim->state.p, sb, hipMemcpyDeviceToDevice, s), "copy");
        auto* circ = const_cast<mqa_circuit*>(circ_c);
        if (circ && !circ->gates.empty()) {
            DevPlan* dp = circ->get_plan(sim->n_eff, sim->dtype, false, s, true);
            const double* ang = build_angles(sim, circ, nullptr, 1, 0, params, n_params);
            run_plan(sim, dp, sim->psi.p, nullptr, 1, ang, circ->n_slots, nullptr, 0, 0, -1, custom_rows(sim, circ, 0, 1));
        }
        // cumulative probability by blocks (device sums, host prefix), then the search inside the
        // blocks the shots land in (GetCumulativeProbs / LowerBound, vector_state.tpp:1357-1361)
        const uint32_t block = uint32_t(std::min<uint64_t>(kSampleBlock, sim->dim_eff));
        const uint64_t nb = sim->dim_eff / block;
        sim->fpart.ensure(nb * sizeof(double));
        hip_check(launch_prob_block_sums(sim->dtype, sim->psi.p, sim->dim_eff, block, sim->fpart.as<double>(), s),
                  "probability sums");
        std::vector<double> bsum(nb);
        hip_check(hipMemcpyAsync(bsum.data(), sim->fpart.p, nb * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
        hip_check(hipStreamSynchronize(s), "sync");
        std::vector<double> cum(nb + 1, 0.0);
        for (uint64_t b = 0; b < nb; b++) cum[b + 1] = cum[b] + bsum[b];
        const double total = cum[nb];
        if (!(total > 0)) throw Error(MQA_ERR_RUNTIME, "the state has zero norm");
        std::vector<uint64_t> gid, gbeg, gend;
        std::vector<double> resid(shots);
        for (uint64_t i = 0; i < shots; i++) {
            const double t = uniforms[i] * total;
            // block b with cum[b] <= t < cum[b + 1]; blocks of zero mass are never chosen
            uint64_t b = uint64_t(std::upper_bound(cum.begin(), cum.end(), t) - cum.begin());
            b = (b == 0) ? 0 : b - 1;
            if (b >= nb) b = nb - 1;
            while (b > 0 && bsum[b] <= 0) b--;
            resid[i] = t - cum[b];
            if (gid.empty() || gid.back() != b) {
                if (!gid.empty()) gend.push_back(i);
                gid.push_back(b);
                gbeg.push_back(i);
            }
        }
        gend.push_back(shots);
        DevBuf d_gid, d_gbeg, d_gend, d_res, d_out;
        upload(d_gid, gid, s);
        upload(d_gbeg, gbeg, s);
        upload(d_gend, gend, s);
        upload(d_res, resid, s);
        d_out.ensure(shots * sizeof(uint64_t));
        hip_check(launch_sample_blocks(sim->dtype, sim->psi.p, block, d_gid.as<uint64_t>(), d_gbeg.as<uint64_t>(),
                                       d_gend.as<uint64_t>(), uint64_t(gid.size()), d_res.as<double>(),
                                       d_out.as<uint64_t>(), s),
                  "sample");
        hip_check(hipMemcpyAsync(out_index, d_out.p, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, s), "d2h");
        sync(sim, "sample");
    });
}

int mqa_sim_expectation_angles(const mqa_sim* sim_c, const mqa_ham* const* hams, int32_t n_hams,
                               const mqa_circuit* circ_c, const double* angles, int64_t batch, double* f_out) {
    return guarded([&] {
        require(sim_c != nullptr && circ_c != nullptr && f_out != nullptr, "NULL argument");
        require(n_hams >= 1 && hams != nullptr, "need at least one hamiltonian");
        require(batch >= 1, "batch must be positive");
        auto* sim = const_cast<mqa_sim*>(sim_c);
        auto* circ = const_cast<mqa_circuit*>(circ_c);
        require(circ->max_qubit < sim->n, "circuit is wider than the simulator");
        require(circ->n_slots == 0 || angles != nullptr, "angles is NULL");
        require(circ->custom_need == 0, "angle settings do not reach parameterised matrix gates (their matrices come "
                                        "from mqa_circuit_set_custom_matrices)");
        for (int h = 0; h < n_hams; h++) check_ham(sim, hams[h]);
        DeviceGuard dg(sim->device);
        hipStream_t s = sim->stream;
        const size_t sb = sim->state_bytes();
        DevPlan* fwd = circ->gates.empty() ? nullptr : circ->get_plan(sim->n_eff, sim->dtype, false, s, true);
        const size_t budget = hbm_budget(sim->psi.bytes + sim->lam.bytes);
        int64_t chunk = int64_t(std::max<size_t>(budget / (2 * sb + 4096), 1));
        chunk = std::min<int64_t>({chunk, batch, 65535});
        {
            DevBuf* bufs[2] = {&sim->psi, &sim->lam};
            const size_t need[2] = {sb * size_t(chunk), sb * size_t(chunk)};
            fit_and_ensure(bufs, need, 2, budget);
        }
        std::vector<double> tab(size_t(chunk) * std::max(circ->n_slots, 1) * 2), f_host(size_t(chunk) * 2);
        for (int64_t b0 = 0; b0 < batch; b0 += chunk) {
            const int64_t nb = std::min<int64_t>(chunk, batch - b0);
            // (cos, sin) of the half angle (rotations) or of the angle (PhaseShift, GlobalPhase)
            for (int64_t b = 0; b < nb; b++)
                for (int sl = 0; sl < circ->n_slots; sl++) {
                    double th = angles[(b0 + b) * circ->n_slots + sl];
                    if (circ->slot_half[size_t(sl)]) th *= 0.5;
                    tab[size_t(b * circ->n_slots + sl) * 2] = std::cos(th);
                    tab[size_t(b * circ->n_slots + sl) * 2 + 1] = std::sin(th);
                }
            if (circ->n_slots > 0) {
                sim->angles.ensure(tab.size() * sizeof(double));
                hip_check(hipMemcpyAsync(sim->angles.p, tab.data(), size_t(nb) * circ->n_slots * 2 * sizeof(double),
                                         hipMemcpyHostToDevice, s),
                          "angles");
            }
            hip_check(launch_fill(sim->dtype, sim->psi.p, sim->state.p, sim->dim_eff, sim->dim_eff, sim->dim_eff, nb, s),
                      "fill");
            if (fwd)
                run_plan(sim, fwd, sim->psi.p, nullptr, nb, circ->n_slots ? sim->angles.as<double>() : nullptr,
                         circ->n_slots, nullptr, 0);
            for (int h = 0; h < n_hams; h++) {
                ham_apply(sim, const_cast<mqa_ham*>(hams[h]), sim->psi.p, sim->psi.p, sim->lam.p, nb);
                hip_check(hipMemcpyAsync(f_host.data(), sim->fsum.p, nb * 2 * sizeof(double), hipMemcpyDeviceToHost, s),
                          "d2h f");
                sync(sim, "expectation batch");
                for (int64_t b = 0; b < nb; b++) {
                    f_out[((b0 + b) * n_hams + h) * 2] = f_host[2 * b];
                    f_out[((b0 + b) * n_hams + h) * 2 + 1] = f_host[2 * b + 1];
                }
            }
        }
    });
}

int mqa_copy_probe(uint64_t bytes, int32_t reps, int32_t in_place, double* ms_per_copy) {
    return guarded([&] {
        require(ms_per_copy != nullptr && reps >= 1 && bytes >= 16, "bad argument");
        if (mqa_device_count() <= 0) throw Error(MQA_ERR_HIP, "no HIP device is visible");
        StreamOwner so;
        const uint64_t n = bytes / 16;  // complex128 elements
        DevBuf a, b;
        a.ensure(n * 16);
        b.ensure(n * 16);
        hip_check(hipMemsetAsync(a.p, 0, n * 16, so.s), "memset");
        hip_check(launch_scale(1, a.p, b.p, n, 1.0, 0.0, so.s), "copy");  // first touch
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        hip_check(hipEventRecord(e0, so.s), "event");
        void* dst = in_place ? a.p : b.p;  // in place: every element read and written back (a gate pass's pattern)
        const int kind = in_place == 2 ? 2 : 1;  // 2: in place with non-temporal accesses
        for (int r = 0; r < reps; r++) hip_check(launch_scale(kind, a.p, dst, n, 1.0, 0.0, so.s), "copy");
        hip_check(hipEventRecord(e1, so.s), "event");
        hip_check(hipEventSynchronize(e1), "sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms_per_copy = double(ms) / reps;
    });
}

int mqa_sim_astype(const mqa_sim* sim, int32_t dtype, mqa_sim** out) {
    return guarded([&] {
        require(sim != nullptr && out != nullptr, "NULL argument");
        require(dtype == MQA_C64 || dtype == MQA_C128, "bad dtype");
        DeviceGuard dg(sim->device);
        mqa_sim* c = nullptr;
        int st = mqa_sim_create(sim->n, dtype, sim->seed, &c);
        if (st != MQA_OK) throw Error(st, g_last_error);
        if (dtype == sim->dtype) {
            hip_check(hipMemcpyAsync(c->state.p, sim->state.p, sim->state_bytes(), hipMemcpyDeviceToDevice, c->stream),
                      "copy");
        } else {
            hip_check(hipStreamSynchronize(sim->stream), "sync");
            hip_check(launch_cast(sim->dtype, sim->state.p, c->state.p, sim->dim_eff * 2, c->stream), "cast");
        }
        hip_check(hipStreamSynchronize(c->stream), "sync");
        *out = c;
    });
}

int mqa_inner_product(const mqa_sim* bra, const mqa_sim* ket, double out[2]) {
    return guarded([&] {
        require(bra != nullptr && ket != nullptr && out != nullptr, "NULL argument");
        require(bra->n == ket->n && bra->dtype == ket->dtype, "states do not match");
        auto* b = const_cast<mqa_sim*>(bra);
        DeviceGuard dg(b->device);
        if (ket->stream != b->stream) hip_check(hipStreamSynchronize(ket->stream), "sync");
        vdot(b, b->state.p, ket->state.p, out);
    });
}

}  // extern "C"
