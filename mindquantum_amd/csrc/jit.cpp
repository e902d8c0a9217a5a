// jit.cpp - source generation, hipRTC compilation and launch of per-circuit tile passes.
#include "jit.h"

#include <dlfcn.h>
#include <hip/hiprtc.h>
#include <spawn.h>
#include <sys/wait.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cctype>
#include <cinttypes>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>

#include "launch.h"

extern char** environ;

namespace mqa {
namespace {

// Text of program.h and device_ops.h, embedded at build time (kJitHeaderNames / kJitHeaders).
#include "jit_headers.inc"

const char* kind_name(int k) {
    switch (k) {
        case OPK_X: return "OPK_X";
        case OPK_H: return "OPK_H";
        case OPK_HU: return "OPK_HU";
        case OPK_M1: return "OPK_M1";
        case OPK_SWAP: return "OPK_SWAP";
        case OPK_M2: return "OPK_M2";
        case OPK_CX: return "OPK_CX";
        case OPK_CZ: return "OPK_CZ";
        case OPK_RX: return "OPK_RX";
        case OPK_RY: return "OPK_RY";
        case OPK_RZ: return "OPK_RZ";
        case OPK_PS: return "OPK_PS";
        case OPK_GP: return "OPK_GP";
        case OPK_RXX: return "OPK_RXX";
        case OPK_RYY: return "OPK_RYY";
        case OPK_RZZ: return "OPK_RZZ";
        case OPK_PF_RZZ: return "OPK_PF_RZZ";
        default: return nullptr;
    }
}

bool two_target_kind(int k) {
    return k == OPK_CX || k == OPK_CZ || k == OPK_SWAP || k == OPK_RZZ || k == OPK_RXX || k == OPK_RYY ||
           k == OPK_M2 || k == OPK_PF_RZZ;
}

// The compile helper sits next to this library (mqamd_jitc); empty if it is not there.
std::string helper_path() {
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&helper_path), &info) == 0 || info.dli_fname == nullptr) return {};
    std::string lib = info.dli_fname;
    const size_t slash = lib.rfind('/');
    const std::string path = (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/mqamd_jitc";
    return access(path.c_str(), X_OK) == 0 ? path : std::string();
}

bool read_file(const std::string& path, std::string& out) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    std::stringstream ss;
    ss << in.rdbuf();
    out = ss.str();
    return true;
}

// Body of a phase-free op (OPK_PF_*; see program.h), closing the op's brace.  Taps are taken
// before the op, as for every kind: PF_RX / PF_RY with the RX / RY tap, PF_RZ with the PS tap of
// the full angle, PF_GP with a zero tap (its variable keeps its initial zero).
// With `ready` (the slot's CoefForm entry exists), the coefficients come from the coefficient
// table instead of being derived from (cos, sin) in every lane.
void emit_phase_free(std::ostringstream& b, const DevOp& op, bool adj, bool ready,
                     std::vector<std::pair<int, int>>& taps) {
    const int q = int(op.q0);
    int var = -1;
    if (adj && op.tap >= 0) {
        var = int(taps.size());
        taps.push_back({var, op.tap});
    }
    if (op.kind == OPK_PF_GP) {
        b << "    }\n";
        return;
    }
    const char* neg = adj ? "-" : "";
    if (ready && op.kind != OPK_PF_RX && op.kind != OPK_PF_RY)  // COEF_DOUBLE
        b << "      const T c = coef_row[" << 2 * op.slot << "];\n"
          << "      const T s = " << neg << "coef_row[" << 2 * op.slot + 1 << "];\n";
    else if (!ready)
        b << "      T c = T(ang_row[" << 2 * op.slot << "]);\n"
          << "      T s = " << neg << "T(ang_row[" << 2 * op.slot + 1 << "]);\n";
    const char* dbl = ready ? "" : "      double_angle(c, s);\n";
    if (op.kind == OPK_PF_RZZ) {
        const int q1 = int(op.q1);
        b << dbl;
        if (var >= 0)
            b << "      tap" << var << " = two_target_tap<T, OPK_PF_RZZ, " << q << ", " << q1
              << ", false>(p, l, g, 0u);\n";
        b << "      apply_body<T, OPK_PF_RZZ, " << q << ", " << q1 << ", false>(p, g, 0u, c, s, pool);\n";
        if (adj) b << "      apply_body<T, OPK_PF_RZZ, " << q << ", " << q1 << ", false>(l, g, 0u, c, s, pool);\n";
    } else if (op.kind == OPK_PF_RZ) {
        b << dbl;
        if (var >= 0) b << "      tap" << var << " = one_target_tap<T, OPK_PS, " << q << ", false>(p, l, g, 0u);\n";
        b << "      apply_body<T, OPK_PS, " << q << ", 0, false>(p, g, 0u, c, s, pool);\n";
        if (adj) b << "      apply_body<T, OPK_PS, " << q << ", 0, false>(l, g, 0u, c, s, pool);\n";
    } else {
        const char* base = op.kind == OPK_PF_RX ? "OPK_RX" : "OPK_RY";
        const char* pf = op.kind == OPK_PF_RX ? "OPK_PF_RX" : "OPK_PF_RY";
        if (var >= 0)
            b << "      tap" << var << " = one_target_tap<T, " << base << ", " << q << ", false>(p, l, g, 0u);\n";
        if (ready)  // COEF_SHEAR
            b << "      const Shear<T> h{" << neg << "coef_row[" << 2 * op.slot << "], " << neg << "coef_row["
              << 2 * op.slot + 1 << "]};\n";
        else
            b << "      const Shear<T> h = shear_of(c, s);\n";
        b << "      pf_rotation<T, " << pf << ", " << q << ">(p, h);\n";
        if (adj) b << "      pf_rotation<T, " << pf << ", " << q << ">(l, h);\n";
    }
    b << "    }\n";
}

// Body of a phase block (OPK_PHASE) in a group with register positions grp.pos.  Each term's
// indicator splits into register bits R (compile-time per register), thread bits (g.ub) and
// tile bits (f.base): terms without register bits multiply a per-thread factor E, terms on one
// register bit j a factor F_j, and register v's phase is E * prod_{j in v} F_j (times the rare
// multi-register terms), built once per register and applied to psi (and lambda).  Back-sweep
// taps: sums of m_v = Im(conj(lambda_v) psi_v) over each term's registers, weighted and reduced
// four at a time.
void emit_phase_block(std::ostringstream& b, const Plan& pl, const DevOp& op, const DevGroup& grp, bool adj) {
    uint32_t posmask = 0;
    for (int j = 0; j < 4; j++) posmask |= uint32_t(1) << grp.pos[j];
    auto hex32 = [](uint32_t x) {
        std::ostringstream o;
        o << "0x" << std::hex << x << "u";
        return o.str();
    };
    auto hex64 = [](uint64_t x) {
        std::ostringstream o;
        o << "0x" << std::hex << x << "ull";
        return o.str();
    };
    auto num = [](double x) {
        std::ostringstream o;
        o << std::setprecision(17) << x;
        return o.str();
    };
    struct Factor {
        bool id = true;
        std::string r, i;
    };
    Factor E{true, "er", "ei"};
    Factor F[4];
    for (int j = 0; j < 4; j++) F[j] = Factor{true, "f" + std::to_string(j) + "r", "f" + std::to_string(j) + "i"};
    struct Extra {
        uint32_t R;
        bool parity;
        std::string s;  // PARITY: runtime flip of the term's parity ("" when the flip is 0)
        std::string x;  // variable name stem (x##r, x##i)
    };
    std::vector<Extra> extras;
    std::ostringstream decl, body, taps_code;
    body << "    {  // phase block, " << op.slot << " terms\n";
    // multiply factor X by (zr, zi), when cond holds (cond empty: always)
    auto mul = [&](Factor& X, const std::string& zr, const std::string& zi, const std::string& cond) {
        if (X.id) {
            if (cond.empty())
                body << "      " << X.r << " = " << zr << ", " << X.i << " = " << zi << ";\n";
            else
                body << "      " << X.r << " = (" << cond << ") ? " << zr << " : T(1), " << X.i << " = (" << cond
                     << ") ? " << zi << " : T(0);\n";
            X.id = false;
        } else if (cond.empty()) {
            body << "      cmul<T>(" << X.r << ", " << X.i << ", " << zr << ", " << zi << ");\n";
        } else {
            body << "      cmul_if<T>(" << X.r << ", " << X.i << ", " << zr << ", " << zi << ", " << cond << ");\n";
        }
    };
    // taps: pass tap index -> accumulator variable
    std::vector<int> tap_idx;
    std::map<std::string, bool> sums;  // named register sums of m_v emitted so far
    bool want_m = false;
    for (int k = 0; k < op.slot; k++) want_m = want_m || (adj && pl.dev_terms.at(size_t(op.mat + k)).tap >= 0);
    if (want_m)
        for (int v = 0; v < 16; v++)
            taps_code << "      const T m" << v << " = l.r[" << v << "] * p.i[" << v << "] - l.i[" << v << "] * p.r["
                      << v << "];\n";
    auto sum_of = [&](const std::string& name, auto pred) {
        if (!sums.count(name)) {
            taps_code << "      const T " << name << " = ";
            bool first = true;
            for (int v = 0; v < 16; v++)
                if (pred(v)) taps_code << (first ? "" : " + ") << "m" << v, first = false;
            if (first) taps_code << "T(0)";
            taps_code << ";\n";
            sums[name] = true;
        }
        return name;
    };
    for (int k = 0; k < op.slot; k++) {
        const DevPhaseTerm& t = pl.dev_terms.at(size_t(op.mat + k));
        uint32_t R = 0;
        for (int j = 0; j < 4; j++)
            if ((t.ml >> grp.pos[j]) & 1u) R |= 1u << j;
        const uint32_t H = t.ml & ~posmask;
        const std::string zr = "zrow[" + std::to_string(2 * t.coef) + "]";
        const std::string zi = std::string(adj ? "-" : "") + "zrow[" + std::to_string(2 * t.coef + 1) + "]";
        const std::string zrn = "z" + std::to_string(k) + "r", zin = "z" + std::to_string(k) + "i";
        body << "      const T " << zrn << " = " << zr << ", " << zin << " = " << zi << ";\n";
        std::string cond;  // AND: the term holds on this thread's registers covering R
        std::string sflip;  // PARITY: the term's parity over thread and tile bits
        if (t.kind == PHT_AND) {
            if (H) cond = "(g.ub & " + hex32(H) + ") == " + hex32(H);
            if (t.mg) cond += std::string(cond.empty() ? "" : " && ") + "(f.base & " + hex64(t.mg) + ") == " + hex64(t.mg);
        } else {
            std::string par;
            if (H) par = "__popc(g.ub & " + hex32(H) + ")";
            if (t.mg) par += std::string(par.empty() ? "" : " + ") + "__popcll(f.base & " + hex64(t.mg) + ")";
            if (!par.empty()) sflip = "((" + par + ") & 1) != 0";
        }
        const std::string condv = "c" + std::to_string(k);
        if (!cond.empty()) body << "      const bool " << condv << " = " << cond << ";\n";
        if (!sflip.empty()) body << "      const bool " << condv << " = " << sflip << ";\n";
        const std::string cnd = (cond.empty() && sflip.empty()) ? std::string() : condv;
        const int nr = __builtin_popcount(R);
        if (t.kind == PHT_AND) {
            if (nr == 0) mul(E, zrn, zin, cnd);
            else if (nr == 1) mul(F[__builtin_ctz(R)], zrn, zin, cnd);
            else {
                const std::string x = "x" + std::to_string(k);
                if (cnd.empty()) body << "      const T " << x << "r = " << zrn << ", " << x << "i = " << zin << ";\n";
                else
                    body << "      const T " << x << "r = " << cnd << " ? " << zrn << " : T(1), " << x << "i = " << cnd
                         << " ? " << zin << " : T(0);\n";
                extras.push_back(Extra{R, false, "", x});
            }
        } else {
            if (nr == 0) {
                mul(E, zrn, zin, cnd);
            } else if (nr == 1) {
                // parity = v_j ^ s: z on v_j = 1 when s = 0; when s = 1, z on every register and
                // conj(z) on v_j = 1
                Factor& Fj = F[__builtin_ctz(R)];
                if (cnd.empty()) {
                    mul(Fj, zrn, zin, "");
                } else {
                    mul(E, zrn, zin, cnd);
                    const std::string zjn = "zj" + std::to_string(k);
                    body << "      const T " << zjn << " = " << cnd << " ? -" << zin << " : " << zin << ";\n";
                    mul(Fj, zrn, zjn, "");
                }
            } else {
                const std::string x = "x" + std::to_string(k);
                body << "      const T " << x << "r = " << zrn << ", " << x << "i = " << zin << ";\n";
                extras.push_back(Extra{R, true, cnd, x});
            }
        }
        // back-sweep tap of the term: weight * sum of m_v over the registers where it holds
        if (adj && t.tap >= 0) {
            std::string val;
            if (t.kind == PHT_AND) {
                const std::string M =
                    R == 0 ? sum_of("sall", [](int) { return true; })
                           : sum_of("ma" + std::to_string(R), [R](int v) { return (uint32_t(v) & R) == R; });
                val = cnd.empty() ? M : "(" + cnd + " ? " + M + " : T(0))";
            } else if (R == 0) {
                const std::string S = sum_of("sall", [](int) { return true; });
                val = "(" + cnd + " ? " + S + " : T(0))";
            } else {
                const std::string O =
                    sum_of("mo" + std::to_string(R), [R](int v) { return (__builtin_popcount(uint32_t(v) & R) & 1) != 0; });
                if (cnd.empty()) {
                    val = O;
                } else {
                    const std::string S = sum_of("sall", [](int) { return true; });
                    val = "(" + cnd + " ? (" + S + " - " + O + ") : " + O + ")";
                }
            }
            size_t a = 0;
            while (a < tap_idx.size() && tap_idx[a] != t.tap) a++;
            if (a == tap_idx.size()) tap_idx.push_back(t.tap);
            body << "      ptap" << a << " += T(" << num(t.weight) << ") * " << val << ";\n";
        }
    }
    // register phases: Q[v] = Q[v without its highest bit] * F_hb (* multi-register AND terms),
    // each applied as soon as it is built, in depth-first order over that tree (few live values)
    std::vector<Factor> Q(16);
    Q[0] = E;
    auto apply_reg = [&](int v) {
        Factor a = Q[size_t(v)];
        bool copied = false;
        for (const Extra& x : extras) {
            if (!x.parity) continue;
            const bool odd = (__builtin_popcount(uint32_t(v) & x.R) & 1) != 0;
            if (x.s.empty() && !odd) continue;
            if (!copied) {  // the multi-register parity terms of register v on a copy
                const std::string nr = "qa" + std::to_string(v) + "r", ni = "qa" + std::to_string(v) + "i";
                if (a.id) body << "      T " << nr << " = T(1), " << ni << " = T(0);\n";
                else body << "      T " << nr << " = " << a.r << ", " << ni << " = " << a.i << ";\n";
                a = Factor{false, nr, ni};
                copied = true;
            }
            if (x.s.empty())
                body << "      cmul<T>(" << a.r << ", " << a.i << ", " << x.x << "r, " << x.x << "i);\n";
            else
                body << "      cmul_if<T>(" << a.r << ", " << a.i << ", " << x.x << "r, " << x.x << "i, "
                     << (odd ? "!" : "") << x.s << ");\n";
        }
        if (a.id) return;
        body << "      cmul<T>(p.r[" << v << "], p.i[" << v << "], " << a.r << ", " << a.i << ");\n";
        if (adj) body << "      cmul<T>(l.r[" << v << "], l.i[" << v << "], " << a.r << ", " << a.i << ");\n";
    };
    auto build = [&](auto&& self, int v) -> void {
        if (v > 0) {
            const int hb = 31 - __builtin_clz(uint32_t(v));
            const Factor& P = Q[size_t(v & ~(1 << hb))];
            std::vector<std::pair<std::string, std::string>> mults;
            if (!F[hb].id) mults.push_back({F[hb].r, F[hb].i});
            for (const Extra& x : extras)
                if (!x.parity && ((x.R >> hb) & 1u) && (uint32_t(v) & x.R) == x.R) mults.push_back({x.x + "r", x.x + "i"});
            if (mults.empty()) {
                Q[size_t(v)] = P;
            } else {
                Factor q{false, "q" + std::to_string(v) + "r", "q" + std::to_string(v) + "i"};
                size_t m0 = 0;
                if (P.id) {
                    body << "      T " << q.r << " = " << mults[0].first << ", " << q.i << " = " << mults[0].second
                         << ";\n";
                    m0 = 1;
                } else {
                    body << "      T " << q.r << " = " << P.r << ", " << q.i << " = " << P.i << ";\n";
                }
                for (size_t m = m0; m < mults.size(); m++)
                    body << "      cmul<T>(" << q.r << ", " << q.i << ", " << mults[m].first << ", " << mults[m].second
                         << ");\n";
                Q[size_t(v)] = q;
            }
        }
        apply_reg(v);
        const int lo = v == 0 ? 0 : 32 - __builtin_clz(uint32_t(v));  // children add a higher bit
        for (int k = lo; k < 4; k++) self(self, v | (1 << k));
    };
    build(build, 0);
    // factor declarations, tap accumulators; taps before the phases (m_v does not change)
    if (!E.id) decl << "      T er, ei;\n";
    for (int j = 0; j < 4; j++)
        if (!F[j].id) decl << "      T " << F[j].r << ", " << F[j].i << ";\n";
    for (size_t a = 0; a < tap_idx.size(); a++) decl << "      T ptap" << a << " = T(0);\n";
    std::string bs = body.str();
    const std::string head = "    {  // phase block, " + std::to_string(op.slot) + " terms\n";
    b << head << decl.str() << taps_code.str() << bs.substr(head.size());
    for (size_t a = 0; a < tap_idx.size(); a += 4) {
        b << "      taps4_to_lds<T>(f";
        for (size_t j = a; j < a + 4; j++) {
            if (j < tap_idx.size()) b << ", ptap" << j << ", " << tap_idx[j];
            else b << ", T(0), -1";
        }
        b << ");\n";
    }
    b << "    }\n";
}

}  // namespace

std::vector<uint8_t> jit_coef_forms(const Plan& pl) {
    std::vector<uint8_t> forms;
    std::vector<bool> clash;
    for (const DevOp& op : pl.dev_ops) {
        if (!opk_is_param(op.kind) || op.slot < 0 || op.kind == OPK_PF_GP || opk_is_custom(op.kind)) continue;
        const uint8_t f = op.kind == OPK_PF_RX || op.kind == OPK_PF_RY     ? COEF_SHEAR
                          : op.kind == OPK_PF_RZ || op.kind == OPK_PF_RZZ ? COEF_DOUBLE
                                                                          : COEF_CS;
        const size_t k = size_t(op.slot);
        if (k >= forms.size()) forms.resize(k + 1, COEF_NONE), clash.resize(k + 1, false);
        if (forms[k] != COEF_NONE && forms[k] != f) clash[k] = true;
        forms[k] = f;
    }
    for (size_t k = 0; k < forms.size(); k++)
        if (clash[k]) forms[k] = COEF_NONE;
    return forms;
}

const char* jit_kernel_name(bool adjoint) { return adjoint ? "mqa_jit_tile_pass_adj" : "mqa_jit_tile_pass_fwd"; }

namespace {
std::atomic<int64_t> g_compile_us{0};
}
double jit_compile_seconds_total() { return double(g_compile_us.load()) * 1e-6; }

std::string jit_pass_source(const Plan& pl, int pass_index, int dtype, JitForm form) {
    const bool sb = form == JitForm::kSingleBuffer;
    if (sb && !jit_can_single_buffer(pl, dtype))
        throw std::invalid_argument(
            "code generation: single-buffer passes are back-sweeps without dense 3- or 4-qubit blocks");
    const DevPass& ps = pl.dev_passes.at(size_t(pass_index));
    const bool adj = pl.adjoint;
    const char* T = dtype == 0 ? "float" : "double";
    const char* A = adj ? "true" : "false";
    const unsigned nt = 1u << (pl.K - kGroupWidth);
    // waves per SIMD the LDS of a complex64 pass allows (forward and single-buffer tiles: four;
    // two tile buffers: two), as the register budget the compiler must meet; complex128 passes
    // hold twice the registers and are left to the compiler
    const int min_waves = dtype != 0 ? 1 : (adj && !sb ? 2 : 4);
    std::ostringstream o;
    o << "#include \"device_ops.h\"\n"
      << "namespace mqa {\n"
      << "extern \"C\" __global__ void __launch_bounds__(" << nt << ", " << min_waves << ") " << jit_kernel_name(adj)
      << "(TilePassArgs args) {\n"
      << "  using T = " << T << ";\n"
      << "  extern __shared__ __align__(16) unsigned char smem[];\n"
      << "  DevPass pass{};\n"
      << "  pass.nonlocal_mask = 0x" << std::hex << ps.nonlocal_mask << "ull;\n"
      << "  pass.local_hi_mask = 0x" << ps.local_hi_mask << "ull;\n"
      << std::dec << "  pass.group_begin = " << ps.group_begin << ";\n"
      << "  pass.n_groups = " << ps.n_groups << ";\n"
      << "  pass.tap_begin = " << ps.tap_begin << ";\n"
      << "  pass.n_taps = " << ps.n_taps << ";\n"
      << "  pass.low_run = " << ps.low_run << ";\n"
      << "  const T* pool = (const T*)as_const(static_cast<const T*>(args.pool));\n"
      << "  (void)pool;\n";
    if (sb) {
        o << "  TileFrame<T> f = sb_begin<T, " << pl.K << ">(args, pass, smem);\n";
    } else {
        o << "  TileFrame<T> f = tile_begin<T, " << A << ", " << pl.K << ">(args, pass, smem);\n";
    }
    o << "  const_ptr<double> ang_row = args.angles ? as_const(args.angles + 2 * f.sample * uint64_t(args.n_slots)) "
         ": nullptr;\n"
      << "  (void)ang_row;\n";
    bool has_custom = false;
    for (const DevOp& op : pl.dev_ops) has_custom = has_custom || opk_is_custom(op.kind);
    if (has_custom)
        o << "  const double* pm_row = args.custom + f.sample * args.custom_stride;\n";
    const std::vector<uint8_t> forms = jit_coef_forms(pl);
    auto ready = [&](const DevOp& op) {
        return op.slot >= 0 && size_t(op.slot) < forms.size() && forms[size_t(op.slot)] != COEF_NONE;
    };
    if (!forms.empty())
        o << "  const_ptr<T> coef_row = as_const(static_cast<const T*>(args.coefs) + 2 * f.sample * "
             "uint64_t(args.n_slots));\n"
          << "  (void)coef_row;\n";
    if (!pl.dev_terms.empty())
        o << "  const_ptr<T> zrow = as_const(static_cast<const T*>(args.phase) + 2 * f.sample * "
             "uint64_t(args.n_phase));\n"
          << "  (void)zrow;\n";
    // single buffer: every group's layout at function scope; psi, then lambda, into group 0's layout
    auto sb_move = [&](const char* what, int gi, const char* reg) {
        o << "  group_" << what << "1<T>(f.t_psi, swz<T>(group_ctx(f.tid, pos_" << gi << ", spos_" << gi
          << ").ub), off_" << gi << ", " << reg << ");\n";
    };
    if (sb) {
        for (int gi = 0; gi < ps.n_groups; gi++) {
            const DevGroup& g = pl.dev_groups.at(size_t(ps.group_begin + gi));
            o << "  constexpr uint8_t pos_" << gi << "[4] = {" << int(g.pos[0]) << ", " << int(g.pos[1]) << ", "
              << int(g.pos[2]) << ", " << int(g.pos[3]) << "};\n"
              << "  constexpr uint8_t spos_" << gi << "[4] = {" << int(g.spos[0]) << ", " << int(g.spos[1]) << ", "
              << int(g.spos[2]) << ", " << int(g.spos[3]) << "};\n"
              << "  constexpr uint16_t off_" << gi << "[16] = {";
            for (int v = 0; v < 16; v++) o << (v ? ", " : "") << g.off[v];
            o << "};\n";
        }
        o << "  Regs<T> p, l;\n"
          << "  tile_load<T>(f.t_psi, f.g_psi, f.offtab, f.C, f.tid, f.NT);\n  __syncthreads();\n";
        sb_move("load", 0, "p");
        o << "  __syncthreads();\n  tile_load<T>(f.t_psi, f.g_lam, f.offtab, f.C, f.tid, f.NT);\n  __syncthreads();\n";
        sb_move("load", 0, "l");
    }
    // unnormalised Hadamards applied so far (OPK_HU): the tile is sqrt(2)^n_hu times the state, so
    // each tap taken from here on is 2^n_hu times its value
    int n_hu = 0;
    auto pow2 = [](int e) {
        std::ostringstream s;
        s << "T(" << std::setprecision(17) << std::ldexp(1.0, e) << ")";
        return s.str();
    };
    auto unscale = [&](const char* indent) {  // sqrt(2)^-n_hu on the registers, before the tile leaves
        if (n_hu == 0) return;
        // hi + lo: the factor in the state's real type and its rounding error (complex64: a float
        // pair, as hmul; zero for even counts, a power of two)
        const double c = std::ldexp(n_hu % 2 ? 0.70710678118654752440084436210484903928 : 1.0, -(n_hu / 2));
        const double hi = dtype == 0 ? double(float(c)) : c;
        const double lo = dtype == 0 ? double(float(c - hi)) : 0.0;
        std::ostringstream s;
        s << std::setprecision(17) << indent << "{\n" << indent << "  const T hh = T(" << hi << "), hl = T(" << lo
          << ");\n" << indent << "  scale_regs<T>(p, hh, hl);\n";
        if (adj) s << indent << "  scale_regs<T>(l, hh, hl);\n";
        s << indent << "}\n";
        o << s.str();
    };
    for (int gi = 0; gi < ps.n_groups; gi++) {
        const DevGroup& g = pl.dev_groups.at(size_t(ps.group_begin + gi));
        if (sb) {
            if (gi > 0) {  // re-lay psi, then lambda, from group gi - 1's layout into group gi's
                o << "  __syncthreads();\n";
                sb_move("store", gi - 1, "p");
                o << "  __syncthreads();\n";
                sb_move("load", gi, "p");
                o << "  __syncthreads();\n";
                sb_move("store", gi - 1, "l");
                o << "  __syncthreads();\n";
                sb_move("load", gi, "l");
            }
            o << "  {  // group " << gi << "\n"
              << "    const GroupCtx g = group_ctx(f.tid, pos_" << gi << ", spos_" << gi << ");\n";
        } else {
        o << "  {  // group " << gi << "\n"
          << "    constexpr uint8_t pos[4] = {" << int(g.pos[0]) << ", " << int(g.pos[1]) << ", " << int(g.pos[2])
          << ", " << int(g.pos[3]) << "};\n"
          << "    constexpr uint8_t spos[4] = {" << int(g.spos[0]) << ", " << int(g.spos[1]) << ", " << int(g.spos[2])
          << ", " << int(g.spos[3]) << "};\n"
          << "    constexpr uint16_t off[16] = {";
        for (int v = 0; v < 16; v++) o << (v ? ", " : "") << g.off[v];
        o << "};\n"
          << "    const GroupCtx g = group_ctx(f.tid, pos, spos);\n"
          << "    const uint32_t a0 = swz<T>(g.ub);\n"
          << "    Regs<T> p, l;\n"
          << "    group_load<T, " << A << ">(f, a0, off, p, l);\n";
        }
        std::vector<std::pair<int, int>> taps;  // (variable, pass-level tap index)
        std::ostringstream b;                   // the group's ops
        for (int oi = 0; oi < g.n_ops; oi++) {
            const DevOp& op = pl.dev_ops.at(size_t(g.op_begin + oi));
            if (op.kind == OPK_PHASE) {
                emit_phase_block(b, pl, op, g, adj);
                continue;
            }
            const bool gated = op.ctrl_global != 0;
            const char* ctrl = op.ctrl_local ? "true" : "false";
            // CNOT across the tile in a back-sweep pass: selects (a branch there can make the
            // compiler keep both register sets of psi and lambda alive); forward passes branch,
            // which costs register moves only on tiles where the control is on
            if (adj && gated && op.kind == OPK_X && op.ctrl_local == 0) {
                b << "    {\n      const bool on = (f.base & 0x" << std::hex << op.ctrl_global << "ull) == 0x"
                  << op.ctrl_global << "ull;\n" << std::dec << "      x_if<T, " << int(op.q0) << ">(p, on);\n"
                  << "      x_if<T, " << int(op.q0) << ">(l, on);\n    }\n";
                continue;
            }
            b << "    ";
            if (gated) b << "if ((f.base & 0x" << std::hex << op.ctrl_global << "ull) == 0x" << op.ctrl_global << "ull) " << std::dec;
            b << "{\n";
            if (op.kind == OPK_M3 || op.kind == OPK_M4) {
                const int m = op.kind == OPK_M3 ? 3 : 4;
                b << "      dense_from_lds<T, " << m << ">(p, g, " << op.ctrl_local << "u, pool + " << op.mat
                  << ", f.t_psi);\n";
                if (adj)
                    b << "      dense_from_lds<T, " << m << ">(l, g, " << op.ctrl_local << "u, pool + " << op.mat
                      << ", f.t_lam);\n";
                b << "    }\n";
                continue;
            }
            if (opk_is_custom(op.kind)) {  // parameterised dense ops: matrices from the per-sample table
                const int nq = op.pad0 == 2 ? 2 : 1;
                b << "      ";
                if (adj && op.tap >= 0) {
                    const int var = int(taps.size());
                    taps.push_back({var, op.tap});
                    b << "tap" << var << " = ";
                }
                b << "pm_op<T, " << A << ", " << nq << ", " << int(op.q0) << ", " << (nq == 2 ? int(op.q1) : 0) << ", "
                  << ctrl << ", " << (op.kind == OPK_PMTAP ? "true" : "false") << ">(p, l, g, " << op.ctrl_local
                  << "u, pm_row + " << op.mat << ");\n    }\n";
                continue;
            }
            if (op.kind >= OPK_PF_RX) {  // phase-free forms: never controlled
                emit_phase_free(b, op, adj, ready(op), taps);
                continue;
            }
            // an uncontrolled Hadamard runs without its 1 / sqrt(2) (the pass rescales the tile once)
            const bool hu = op.kind == OPK_H && op.ctrl_local == 0 && op.ctrl_global == 0;
            const char* kn = kind_name(hu ? int(OPK_HU) : op.kind);
            if (!kn) throw std::invalid_argument("code generation: unknown op kind " + std::to_string(op.kind));
            const bool two = two_target_kind(op.kind);
            const int q1 = two ? int(op.q1) : 0;
            if (opk_is_param(op.kind) && ready(op)) {  // COEF_CS
                b << "      const T c = coef_row[" << 2 * op.slot << "];\n"
                  << "      const T s = " << (adj ? "-" : "") << "coef_row[" << 2 * op.slot + 1 << "];\n";
            } else if (opk_is_param(op.kind)) {
                b << "      const T c = T(ang_row[" << 2 * op.slot << "]);\n"
                  << "      const T s = " << (adj ? "-" : "") << "T(ang_row[" << 2 * op.slot + 1 << "]);\n";
            } else {
                b << "      const T c = T(0), s = T(0);\n";
            }
            if (adj && opk_is_param(op.kind) && op.tap >= 0) {
                const int var = int(taps.size());
                taps.push_back({var, op.tap});
                b << "      tap" << var << " = ";
                if (two)
                    b << "two_target_tap<T, " << kn << ", " << int(op.q0) << ", " << q1 << ", " << ctrl << ">(p, l, g, "
                      << op.ctrl_local << "u);\n";
                else
                    b << "one_target_tap<T, " << kn << ", " << int(op.q0) << ", " << ctrl << ">(p, l, g, "
                      << op.ctrl_local << "u);\n";
            }
            const std::string mat = op.mat >= 0 ? ("pool + " + std::to_string(op.mat)) : std::string("pool");
            b << "      apply_body<T, " << kn << ", " << int(op.q0) << ", " << q1 << ", " << ctrl << ">(p, g, "
              << op.ctrl_local << "u, c, s, " << mat << ");\n";
            if (adj)
                b << "      apply_body<T, " << kn << ", " << int(op.q0) << ", " << q1 << ", " << ctrl << ">(l, g, "
                  << op.ctrl_local << "u, c, s, " << mat << ");\n";
            b << "    }\n";
        }
        if (!taps.empty()) {  // tap variables, declared before the ops that assign them
            o << "    T";
            for (size_t k = 0; k < taps.size(); k++) o << (k ? ", " : " ") << "tap" << k << " = T(0)";
            o << ";\n";
        }
        // each tap is pinned where it is taken: left free, the compiler sinks the tap products to the
        // group's reduction, past the rotations that follow, and keeps both register states alive
        {
            std::istringstream in(b.str());
            std::string line;
            while (std::getline(in, line)) {
                if (n_hu > 0 && line.find("taps4_to_lds<T>(f, ptap") != std::string::npos) {
                    // phase-block taps: each accumulator over 2^n_hu
                    std::string out;
                    for (size_t i = 0; i < line.size();) {
                        if (line.compare(i, 4, "ptap") == 0) {
                            size_t e = i + 4;
                            while (e < line.size() && std::isdigit(static_cast<unsigned char>(line[e]))) e++;
                            out += line.substr(i, e - i) + " * " + pow2(-n_hu);
                            i = e;
                        } else {
                            out += line[i++];
                        }
                    }
                    line = out;
                }
                o << line << "\n";
                if (line.find("apply_body<T, OPK_HU,") != std::string::npos && line.find("(p, g") != std::string::npos)
                    n_hu++;
                const size_t t = line.find_first_not_of(' ');
                if (t != std::string::npos && line.compare(t, 3, "tap") == 0) {
                    const size_t eq = line.find(" = ", t);
                    if (eq != std::string::npos && line.back() == ';') {
                        const std::string var = line.substr(t, eq - t);
                        if (n_hu > 0) o << std::string(t, ' ') << var << " *= " << pow2(-n_hu) << ";\n";
                        o << std::string(t, ' ') << "asm volatile(\"\" : \"+v\"(" << var << "));\n";
                    }
                }
            }
        }
        for (size_t k = 0; k < taps.size(); k += 4) {  // four taps per batched cross-lane reduction
            o << "    taps4_to_lds<T>(f";
            for (size_t j = k; j < k + 4; j++) {
                if (j < taps.size()) o << ", tap" << taps[j].first << ", " << taps[j].second;
                else o << ", T(0), -1";
            }
            o << ");\n";
        }
        if (sb) {
            o << "  }\n";
            continue;
        }
        if (gi == ps.n_groups - 1) unscale("    ");
        // the next group reads the tile in its own layout
        o << "    group_store<T, " << A << ">(f, a0, off, p, l);\n"
          << "    __syncthreads();\n"
          << "  }\n";
    }
    if (sb) {  // psi, then lambda, back through the buffer to HBM; the tap partials
        const int last = ps.n_groups - 1;
        o << "  uint32_t tid_st = f.tid;\n  asm volatile(\"\" : \"+v\"(tid_st));\n  __syncthreads();\n"
          << "  if (!args.skip_store) {\n";
        unscale("  ");
        sb_move("store", last, "p");
        o << "  __syncthreads();\n  tile_store<T>(f.t_psi, f.g_psi, f.offtab, f.C, tid_st, f.NT);\n"
          << "  __syncthreads();\n";
        sb_move("store", last, "l");
        o << "  __syncthreads();\n  tile_store<T>(f.t_psi, f.g_lam, f.offtab, f.C, tid_st, f.NT);\n"
          << "  }\n"
          << "  taps_out<T>(args, pass, f);\n";
    } else {
        o << "  tile_end<T, " << A << ">(args, pass, f);\n";
    }
    o << "}\n"
      << "}  // namespace mqa\n";
    return o.str();
}

bool jit_compile(const std::string& src, std::string& code, std::string& log) {
    hiprtcProgram prog;
    const int n_headers = int(sizeof(kJitHeaders) / sizeof(kJitHeaders[0]));
    if (hiprtcCreateProgram(&prog, src.c_str(), "mqa_jit_pass.hip", n_headers, kJitHeaders, kJitHeaderNames) !=
        HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    // the same code generation flags as the static kernels (packed FP32 disabled)
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++20", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, int(sizeof(opts) / sizeof(opts[0])), opts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    log.assign(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, log.data());
    bool ok = r == HIPRTC_SUCCESS;
    if (ok) {
        size_t cs = 0;
        ok = hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS && cs > 0;
        if (ok) {
            code.assign(cs, '\0');
            ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
        }
    }
    hiprtcDestroyProgram(&prog);
    return ok;
}

JitPlan::JitPlan(const Plan& pl, int dtype, JitForm form) : plan_(pl), dtype_(dtype), form_(form) {
    if (form_ == JitForm::kSingleBuffer && !jit_can_single_buffer(pl, dtype)) form_ = JitForm::kTile;
    code_.resize(plan_.dev_passes.size());
    worker_ = std::thread([this] { compile_all(); });
}

JitPlan::~JitPlan() {
    if (worker_.joinable()) worker_.join();
    for (auto& [dev, mods] : modules_) {
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(dev);
        for (auto m : mods)
            if (m) (void)hipModuleUnload(m);
        if (prev >= 0) (void)hipSetDevice(prev);
    }
}

void JitPlan::compile_all() {
    const auto t0 = std::chrono::steady_clock::now();
    const int n = int(code_.size());
    std::vector<std::string> src(static_cast<size_t>(n)), err(static_cast<size_t>(n));
    std::vector<char> done(static_cast<size_t>(n), 0);
    try {
        for (int p = 0; p < n; p++) src[size_t(p)] = jit_pass_source(plan_, p, dtype_, form_);
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mu_);
        error_ = e.what();
        state_.store(kFailed);
        return;
    }
    if (const char* dir = std::getenv("MQAMD_JIT_DUMP")) {  // keep the generated sources
        for (int p = 0; p < n; p++) std::ofstream(std::string(dir) + "/pass" + std::to_string(p) + ".hip") << src[size_t(p)];
    }

    // parallel: helper processes, each compiling every W-th pass; one plan at a time holds them
    // (plans compiling in the background - many circuits at once - queue here instead of each
    // starting its own set of processes)
    const std::string helper = helper_path();
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int workers = std::max(1, std::min<int>({n, int(hw), 16}));
    static std::mutex helpers_mu;
    std::unique_lock<std::mutex> helpers_lock(helpers_mu, std::defer_lock);
    if (!helper.empty() && workers > 1) helpers_lock.lock();
    if (!helper.empty() && workers > 1) {
        const char* tmp = std::getenv("TMPDIR");
        std::string tmpl = std::string(tmp && *tmp ? tmp : "/tmp") + "/mqamd_jit_XXXXXX";
        if (mkdtemp(tmpl.data()) != nullptr) {
            const std::string dir = tmpl;
            auto src_path = [&](int p) { return dir + "/pass" + std::to_string(p) + ".hip"; };
            auto out_path = [&](int p) { return dir + "/pass" + std::to_string(p) + ".co"; };
            for (int p = 0; p < n; p++) std::ofstream(src_path(p)) << src[size_t(p)];
            std::vector<pid_t> pids;
            for (int w = 0; w < workers; w++) {
                std::vector<std::string> args{helper};
                for (int p = w; p < n; p += workers) {
                    args.push_back(out_path(p));
                    args.push_back(src_path(p));
                }
                std::vector<char*> argv;
                for (auto& a : args) argv.push_back(a.data());
                argv.push_back(nullptr);
                pid_t pid = 0;
                if (posix_spawn(&pid, helper.c_str(), nullptr, nullptr, argv.data(), environ) == 0) pids.push_back(pid);
            }
            for (pid_t pid : pids) {
                int st = 0;
                while (waitpid(pid, &st, 0) < 0 && errno == EINTR) {
                }
            }
            for (int p = 0; p < n; p++) {
                done[size_t(p)] = read_file(out_path(p), code_[size_t(p)]) && !code_[size_t(p)].empty();
                std::remove(out_path(p).c_str());
                std::remove(src_path(p).c_str());
                std::remove((out_path(p) + ".log").c_str());
            }
            rmdir(dir.c_str());
        }
    }
    // serial, in this process: whatever the helpers did not produce
    for (int p = 0; p < n; p++) {
        if (done[size_t(p)]) continue;
        std::string log;
        if (!jit_compile(src[size_t(p)], code_[size_t(p)], log)) err[size_t(p)] = log.empty() ? "compile failed" : log;
    }
    seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    g_compile_us.fetch_add(int64_t(seconds_ * 1e6));
    std::string first;
    for (int p = 0; p < n && first.empty(); p++)
        if (!err[size_t(p)].empty()) first = "pass " + std::to_string(p) + ": " + err[size_t(p)];
    {
        std::lock_guard<std::mutex> lk(mu_);
        error_ = first;
    }
    state_.store(first.empty() ? kReady : kFailed);
}

void JitPlan::wait() {
    if (worker_.joinable()) worker_.join();
}

std::string JitPlan::error() const {
    std::lock_guard<std::mutex> lk(mu_);
    return error_;
}

hipError_t JitPlan::launch(int pass, const TilePassArgs& a, uint64_t batch, size_t lds, hipStream_t s) {
    if (!ready()) return hipErrorNotReady;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    hipFunction_t fn = nullptr;
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto& fns = functions_[dev];
        if (fns.empty()) {
            auto& mods = modules_[dev];
            mods.assign(code_.size(), nullptr);
            std::vector<hipFunction_t> loaded(code_.size(), nullptr);
            for (size_t p = 0; p < code_.size(); p++) {
                e = hipModuleLoadData(&mods[p], code_[p].data());
                if (e == hipSuccess) e = hipModuleGetFunction(&loaded[p], mods[p], jit_kernel_name(plan_.adjoint));
                if (e != hipSuccess) {
                    // unload what did load and mark the plan failed: callers use the interpreted
                    // kernel from now on instead of retrying the load
                    for (auto& m : mods)
                        if (m) (void)hipModuleUnload(m), m = nullptr;
                    mods.clear();
                    error_ = std::string("module load: ") + hipGetErrorString(e);
                    state_.store(kFailed);
                    return e;
                }
            }
            fns = std::move(loaded);
        }
        fn = fns.at(size_t(pass));
    }
    TilePassArgs args = a;
    args.batch = batch;
    void* params[] = {&args};
    const unsigned nt = 1u << (plan_.K - kGroupWidth);
    return hipModuleLaunchKernel(fn, unsigned(a.n_tiles), unsigned(batch), 1, nt, 1, 1, unsigned(lds), s, params,
                                 nullptr);
}

}  // namespace mqa
