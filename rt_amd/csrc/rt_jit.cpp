// rt_jit.cpp -- generation, compilation (hiprtc) and caching of per-scheme
// trace kernels.
//
// The generic trace kernel branches on every surface's class, type, shape
// and number of aspheric terms at run time.  For a given scheme that
// structure is fixed, so the engine writes a kernel whose surface loop is a
// straight-line sequence of rt::surface_step<StatSurf<...>> calls (see
// rt_device.h) and compiles it for the device once per structure.
// Parameter values still come from the device table, so perturbed instances
// of one scheme share the kernel.
#include "rt_jit.h"

#include <hip/hiprtc.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

#include "rt_device.h"
#include "rt_adjoint.h"
#include "rt_embed.inc"   // kEmbedDevice, kEmbedAbi, kEmbedAdjoint: text of rt_device.h, rt_hip.h, rt_adjoint.h

namespace {

std::mutex g_mu;
int g_mode = -1;   // -1: not yet read from the environment
int g_ltab = -1;   // per-wavelength tables; -1: not yet read from the environment
struct Entry {
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
    bool failed = false;
    std::string log;
};
std::map<std::string, Entry> g_cache;

int surface_ikind(const double* p) {
    int c = (int)p[RT_S_CLASS], n = (int)p[RT_S_NPOLY];
    if (c == RT_PLANE || (c == RT_CONIC && n == 0 && p[RT_S_C] == 0.0)) return rt::IK_PLANE;
    if (c == RT_CONIC) return n == 0 ? rt::IK_QUAD : rt::IK_CONIC_NEWTON;
    if (c == RT_TORIC) return rt::IK_TORIC_NEWTON;
    return rt::IK_NONE;
}

}  // namespace

std::string rt_jit_signature(const double* dtable, int n_surf) {
    std::ostringstream os;
    double prev_n2 = -1.0;   // device pool offset of the n2 table the previous action evaluated
    for (int s = 0; s < n_surf; ++s) {
        const double* p = dtable + (size_t)s * rt::DEV_STRIDE;
        int type = (int)p[RT_S_TYPE];
        bool opt = (type == RT_LENS || type == RT_GRATING);
        int grt = type == RT_GRATING ? (int)p[RT_S_GRTYPE] : 0;
        int ordm = type == RT_GRATING ? (int)p[RT_S_ORDMODE] : 0;
        bool n1s = p[RT_S_N1KIND] == (double)RT_FN_SPLINE, n2s = p[RT_S_N2KIND] == (double)RT_FN_SPLINE;
        bool ccd = type == RT_DETECTOR && p[RT_S_CCD_XDIM] > 0.0;
        // bit 4: n1 is the table the previous non-auxiliary surface's action evaluated as
        // its n2, for the same ray and wavelength, so its value is reused
        int fnmask = opt ? ((n1s ? 1 : 0) | (n2s ? 2 : 0) | (n1s && p[RT_S_N1VAL] == prev_n2 ? 4 : 0)) : 0;
        // bits 8, 16: the detector's QE and spectrum are tables
        if (ccd)
            fnmask |= (p[RT_S_QEKIND] == (double)RT_FN_SPLINE ? 8 : 0) | (p[RT_S_SPKIND] == (double)RT_FN_SPLINE ? 16 : 0);
        if (type != RT_AUXTYPE) prev_n2 = (opt && n2s) ? p[RT_S_N2VAL] : -1.0;
        int rot = (p[RT_S_ANG + 0] != 0.0 ? 1 : 0) | (p[RT_S_ANG + 1] != 0.0 ? 2 : 0) |
                  (p[RT_S_ANG + 2] != 0.0 ? 4 : 0);
        bool trans = p[RT_S_DIST + 0] != 0.0 || p[RT_S_DIST + 1] != 0.0 || p[RT_S_DIST + 2] != 0.0;
        os << "rt::StatSurf<" << (int)p[RT_S_CLASS] << "," << type << "," << (int)p[RT_S_SHAPE] << ","
           << (int)p[RT_S_NPOLY] << "," << grt << "," << ordm << "," << surface_ikind(p) << "," << fnmask
           << "," << rot << "," << (trans ? "true" : "false") << "," << (ccd ? "true" : "false") << ">;";
    }
    return os.str();
}

// The columns of the per-wavelength table of a scheme with this signature:
// per surface, in order, n1 (a table not reused from the previous surface),
// n2, the detector QE, the detector spectrum, for the functions given as
// tables (StatSurf's FNMASK, the 8th template argument).
void rt_jit_ltab_layout(const std::string& sig, std::vector<int>& col_base, std::vector<int>& col_surf,
                        std::vector<int>& col_role) {
    col_base.clear();
    col_surf.clear();
    col_role.clear();
    size_t pos = 0;
    int s = 0;
    while (pos < sig.size()) {
        const size_t end = sig.find(';', pos);
        const std::string t = sig.substr(pos, end - pos);
        const size_t lt = t.find('<');
        int fnmask = 0;
        if (lt != std::string::npos) {
            // the 8th comma-separated template argument
            size_t q = lt + 1;
            for (int k = 0; k < 7 && q != std::string::npos; ++k) {
                q = t.find(',', q);
                if (q != std::string::npos) ++q;
            }
            if (q != std::string::npos) fnmask = atoi(t.c_str() + q);
        }
        const bool roles[4] = {(fnmask & 1) != 0 && (fnmask & 4) == 0, (fnmask & 2) != 0, (fnmask & 8) != 0,
                               (fnmask & 16) != 0};
        col_base.push_back(-1);
        for (int r = 0; r < 4; ++r) {
            if (!roles[r]) continue;
            if (col_base.back() < 0) col_base.back() = (int)col_surf.size();
            col_surf.push_back(s);
            col_role.push_back(r);
        }
        pos = end + 1;
        ++s;
    }
}

// The columns of the derived per-wavelength table (single-scheme traces):
// per surface, in order, mu = n1/n2 for a lens or a grating, then the
// diffraction term L = m lambda / (n2 d) for a grating, and the flux weight
// for a CCD detector (StatSurf's TYPE and CCD, its 2nd and 11th template
// arguments).
void rt_jit_dtab_layout(const std::string& sig, std::vector<int>& col_base, std::vector<int>& col_surf,
                        std::vector<int>& col_role) {
    col_base.clear();
    col_surf.clear();
    col_role.clear();
    size_t pos = 0;
    int s = 0;
    while (pos < sig.size()) {
        const size_t end = sig.find(';', pos);
        const std::string t = sig.substr(pos, end - pos);
        int type = -1;
        bool ccd = false;
        size_t q = t.find('<');
        for (int k = 0; q != std::string::npos && k <= 10; ++k) {
            ++q;
            if (k == 1) type = atoi(t.c_str() + q);
            if (k == 10) ccd = t.compare(q, 4, "true") == 0;
            q = t.find(k < 10 ? ',' : '>', q);
        }
        col_base.push_back(-1);
        int roles[2], nr = 0;
        if (type == RT_LENS || type == RT_GRATING) roles[nr++] = RT_DTAB_MU;
        if (type == RT_GRATING) roles[nr++] = RT_DTAB_L;
        if (type == RT_DETECTOR && ccd) roles[nr++] = RT_DTAB_W;
        for (int r = 0; r < nr; ++r) {
            if (col_base.back() < 0) col_base.back() = (int)col_surf.size();
            col_surf.push_back(s);
            col_role.push_back(roles[r]);
        }
        pos = end + 1;
        ++s;
    }
}

bool rt_jit_enabled() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_mode < 0) {
        const char* e = getenv("RT_HIP_JIT");
        g_mode = (e && e[0] == '0') ? 0 : 1;
    }
    return g_mode == 1;
}

bool rt_jit_ltab_enabled() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ltab < 0) {
        const char* e = getenv("RT_HIP_LTAB");
        g_ltab = (e && e[0] == '0') ? 0 : 1;
    }
    return g_ltab == 1;
}

extern "C" int rt_hip_set_ltab(int mode) {
    std::lock_guard<std::mutex> lk(g_mu);
    int prev = g_ltab < 0 ? 1 : g_ltab;
    g_ltab = mode ? 1 : 0;
    return prev;
}

extern "C" int rt_hip_set_jit(int mode) {
    std::lock_guard<std::mutex> lk(g_mu);
    int prev = g_mode < 0 ? 1 : g_mode;
    g_mode = mode ? 1 : 0;
    return prev;
}

// variant: RT_JIT_VAR_DUMP (per-surface state dump), RT_JIT_VAR_LOSS and
// RT_JIT_VAR_EXT (the recording modes rt::kRecLoss, rt::kRecExt),
// RT_JIT_VAR_SPOTS (group spot partials per tile, rt::kRecSpots)
static std::string generate_source(const std::string& sig, int beam_mode, int variant, int block, int waves,
                                   bool batch = false) {
    const bool dump = (variant & RT_JIT_VAR_DUMP) != 0;
    const int rec = ((variant & RT_JIT_VAR_LOSS) ? rt::kRecLoss : 0) | ((variant & RT_JIT_VAR_EXT) ? rt::kRecExt : 0) |
                    ((variant & RT_JIT_VAR_SPOTS) ? rt::kRecSpots : 0);
    // RT_JIT_LTAB: per-wavelength values come from a table (TraceArgs::ltab),
    // each column loaded at the surface that reads it: for a single scheme the
    // derived table (mu, the grating term, the flux weight), for a batch the
    // spline-valued functions themselves
    const bool ltab = beam_mode >= 0 && (beam_mode & RT_JIT_LTAB) != 0;
    if (beam_mode >= 0) beam_mode &= ~RT_JIT_LTAB;
    const bool derived = ltab && !batch;
    std::vector<int> col_base, col_surf, col_role;
    if (derived) rt_jit_dtab_layout(sig, col_base, col_surf, col_role);
    else rt_jit_ltab_layout(sig, col_base, col_surf, col_role);
    const bool use_tab = ltab && !col_surf.empty();
    std::ostringstream os;
    os << "#include \"rt_device.h\"\n"
       << "extern \"C\" __global__ __launch_bounds__(" << block;
    if (waves > 0) os << ", " << waves;      // minimum waves per SIMD the register allocation must allow
    os << ") void rt_trace_jit(rt::TraceArgs a) {\n"
       << "  rt::" << (batch ? "trace_batch_body<" : "trace_body<") << beam_mode;
    if (rec) os << ", " << rec;
    os << ">(a, [&](rt::Ray& r, bool& alive, rt::TraceCtx& c, uint64_t i) {\n";
    if (use_tab) os << "    double lv[" << col_surf.size() << "];\n";
    size_t pos = 0;
    int s = 0;
    while (pos < sig.size()) {
        size_t end = sig.find(';', pos);
        std::string type = sig.substr(pos, end - pos);
        if (use_tab)
            for (size_t q = 0; q < col_surf.size(); ++q)
                if (col_surf[q] == s) os << "    rt::load_lval(lv, c, " << q << ");\n";
        os << "    rt::surface_step";
        if (rec || (derived && use_tab)) os << "<" << type << ", " << rec << (derived && use_tab ? ", true" : "") << ">";
        os << "(" << type << "{c.table + " << s << " * rt::DEV_STRIDE}, " << s << ", r, alive, c";
        if (use_tab && col_base[(size_t)s] >= 0) os << ", lv, " << col_base[(size_t)s];
        os << ");\n";
        if (dump) os << "    rt::dump_state(a, " << s << ", r, alive, i);\n";
        pos = end + 1;
        ++s;
    }
    os << "  });\n}\n";
    return os.str();
}

// The adjoint kernel of a scheme structure and parameter list: the forward
// surface steps and the backward steps as straight-line sequences, each
// backward step adding exactly the asked-for slots of that surface to the
// wave's register accumulators, in groups of up to four slots.
// The n-th (0-based) integer template argument of one signature entry.
static int sig_arg(const std::string& t, int n) {
    size_t q = t.find('<');
    if (q == std::string::npos) return 0;
    ++q;
    for (int k = 0; k < n && q != std::string::npos; ++k) {
        q = t.find(',', q);
        if (q != std::string::npos) ++q;
    }
    return q == std::string::npos ? 0 : atoi(t.c_str() + q);
}

// Tape rows of one surface in the generated adjoints: kTape, plus the Newton
// iterates the forward records where the intersection iterates (StatSurf's
// IK, its 7th template argument)
static int adjoint_tape_rows_of(const std::string& type) {
    const int ik = sig_arg(type, 6);
    return rt::adj::kTape + ((ik == rt::IK_CONIC_NEWTON || ik == rt::IK_TORIC_NEWTON) ? rt::adj::kMaxNewtonAdj : 0);
}

int rt_jit_adjoint_tape_rows(const std::string& sig) {
    int rows = 0;
    size_t pos = 0;
    while (pos < sig.size()) {
        const size_t end = sig.find(';', pos);
        rows += adjoint_tape_rows_of(sig.substr(pos, end - pos));
        pos = end + 1;
    }
    return rows;
}

static std::string generate_adjoint_source(const std::string& sig, int beam_type, const std::vector<int>& rows,
                                           int n_surf, int block, int waves = 0, bool batch = false,
                                           bool ltab = false) {
    std::vector<std::string> types;
    size_t pos = 0;
    while (pos < sig.size()) {
        size_t end = sig.find(';', pos);
        types.push_back(sig.substr(pos, end - pos));
        pos = end + 1;
    }
    // per-wavelength table (RT_JIT_LTAB).  One scheme: the derived table, whose
    // mu and grating columns the forward steps read as the trace does and the
    // backward steps read again (the detector's flux weight is not read: no
    // flux in the adjoint).  A batch: the spline-valued functions' columns,
    // read by the forward steps as the trace does; the backward steps read n1
    // and n2 from the columns the forward read them from (a reused n1 is the
    // previous non-auxiliary surface's n2 column)
    const bool derived = ltab && !batch;
    std::vector<int> col_base, col_surf, col_role;
    if (derived) rt_jit_dtab_layout(sig, col_base, col_surf, col_role);
    else rt_jit_ltab_layout(sig, col_base, col_surf, col_role);
    const bool use_tab = ltab && !col_surf.empty();
    std::vector<int> n1col(n_surf, -1), n2col(n_surf, -1);
    if (!derived) {
        int prev_n2 = -1;
        for (int s = 0; s < n_surf; ++s) {
            const int type = sig_arg(types[s], 1), fnmask = sig_arg(types[s], 7);
            for (size_t q = 0; q < col_surf.size(); ++q)
                if (col_surf[q] == s) {
                    if (col_role[q] == 0) n1col[s] = (int)q;
                    if (col_role[q] == 1) n2col[s] = (int)q;
                }
            if ((fnmask & 4) != 0) n1col[s] = prev_n2;
            if (type != RT_AUXTYPE) prev_n2 = n2col[s];
        }
    }
    // gradient sums in registers (rt::adj::GroupAcc): each surface's slots
    // reduced in groups of up to four; group GRP's inputs x0..x3 end in rows
    // 0, 2, 1, 3 of register GRP / 16 at lane offset GRP % 16
    struct Group { int surf; int sl[4]; int par[4]; };
    std::vector<Group> groups;
    for (int s = 0; s < n_surf; ++s) {
        Group g{s, {-1, -1, -1, -1}, {-1, -1, -1, -1}};
        int k = 0;
        for (int sl = 0; sl < rt::adj::PG_N; ++sl) {
            const int j = rows[(size_t)s * rt::adj::PG_N + sl];
            if (j < 0) continue;
            g.sl[k] = sl;
            g.par[k] = j;
            if (++k == 4) {
                groups.push_back(g);
                g = Group{s, {-1, -1, -1, -1}, {-1, -1, -1, -1}};
                k = 0;
            }
        }
        if (k > 0) groups.push_back(g);
    }
    const int n_reg = std::max(1, ((int)groups.size() + 15) / 16);
    std::vector<int> map((size_t)n_reg * 64, -1);
    static const int kRowOf[4] = {0, 2, 1, 3};
    for (size_t gi = 0; gi < groups.size(); ++gi)
        for (int k = 0; k < 4; ++k)
            if (groups[gi].par[k] >= 0) map[(gi / 16) * 64 + 16 * kRowOf[k] + gi % 16] = groups[gi].par[k];
    std::ostringstream os;
    const size_t nc = col_surf.size();
    os << "#include \"rt_adjoint.h\"\n"
       << "__constant__ int rt_adj_slot_map[" << map.size() << "] = {";
    for (size_t k = 0; k < map.size(); ++k) os << (k ? "," : "") << map[k];
    os << "};\n"
       << "struct RtAdjMap { __device__ static int slot(int k) { return rt_adj_slot_map[k]; } };\n"
       << "extern \"C\" __global__ __launch_bounds__(" << block;
    if (waves > 0) os << ", " << waves;
    os << ") void rt_adjoint_jit(rt::adj::AdjArgs a) {\n"
       << "  rt::adj::" << (batch ? "adjoint_batch_body<" : "adjoint_body<") << beam_type
       << ", rt::adj::GroupAcc<" << n_reg << ", RtAdjMap>>(a,\n"
       << "    [&](rt::Ray& r, bool& alive, rt::TraceCtx& c, const rt::adj::LaneRows& t) {\n";
    if (use_tab) os << "      double lv[" << nc << "];\n";
    // tape rows of each surface: the entry state and Newton count (kTape),
    // and the Newton iterates where the intersection iterates
    std::vector<int> k0(n_surf, 0);
    for (int s = 1; s < n_surf; ++s) k0[s] = k0[s - 1] + adjoint_tape_rows_of(types[s - 1]);
    for (int s = 0; s < n_surf; ++s) {
        // refractive-index columns only: the adjoint accumulates no flux, so
        // the detector's QE and spectrum are never read
        bool tab_s = false;
        if (use_tab)
            for (size_t q = 0; q < nc; ++q)
                if (col_surf[q] == s && (derived ? col_role[q] != RT_DTAB_W : col_role[q] <= 1)) {
                    os << "      rt::adj::load_lval_adj(lv, c, " << q << ");\n";
                    tab_s = true;
                }
        os << "      rt::adj::fwd_step<" << (derived && use_tab ? "true" : "false") << ", true>(" << types[s]
           << "{c.table + " << s << " * rt::DEV_STRIDE}, " << s << ", r, alive, c, t, " << k0[s];
        if (tab_s) os << ", lv, " << col_base[(size_t)s];
        os << ");\n";
    }
    os << "    },\n"
       << "    [&](const rt::Ray& r, rt::adj::St& ad, rt::TraceCtx& c, const rt::adj::LaneRows& t, "
          "auto& g, bool ok) -> bool {\n";
    if (use_tab) os << "      double lb[" << nc << "];\n";
    for (int s = n_surf - 1; s >= 0; --s) {
        if (derived && use_tab) {
            // the surface's mu (and grating) columns, loaded again for its backward step
            bool tab_s = false;
            os << "      {";
            for (size_t q = 0; q < nc; ++q)
                if (col_surf[q] == s && col_role[q] != RT_DTAB_W) {
                    os << " rt::adj::load_lval_adj(lb, c, " << q << ");";
                    tab_s = true;
                }
            os << " double gp[rt::adj::PG_N]; const bool use = rt::adj::bwd_step_d<true>(" << types[s]
               << "{c.table + " << s << " * rt::DEV_STRIDE}, r, ad, c, t, " << k0[s] << ", gp, lb, "
               << (tab_s ? col_base[(size_t)s] : -1) << ") && ok;";
            for (size_t gi = 0; gi < groups.size(); ++gi) {
                if (groups[gi].surf != s) continue;
                os << " g.template add4<" << gi << ">(";
                for (int k = 0; k < 4; ++k) {
                    if (k) os << ", ";
                    if (groups[gi].sl[k] >= 0) os << "use ? gp[" << groups[gi].sl[k] << "] : 0.0";
                    else os << "0.0";
                }
                os << ");";
            }
            os << " ok = use; }\n";
            continue;
        }
        const bool tab_s = use_tab && (n1col[(size_t)s] >= 0 || n2col[(size_t)s] >= 0);
        os << "      {";
        if (tab_s) {
            if (n1col[(size_t)s] >= 0) os << " rt::adj::load_lval_adj(lb, c, " << n1col[(size_t)s] << ");";
            if (n2col[(size_t)s] >= 0 && n2col[(size_t)s] != n1col[(size_t)s])
                os << " rt::adj::load_lval_adj(lb, c, " << n2col[(size_t)s] << ");";
        }
        os << " double gp[rt::adj::PG_N]; const bool use = rt::adj::bwd_step<true>(" << types[s] << "{c.table + "
           << s << " * rt::DEV_STRIDE}, r, ad, c, t, " << k0[s] << ", gp";
        if (tab_s) os << ", lb, " << n1col[(size_t)s] << ", " << n2col[(size_t)s];
        else if (use_tab) os << ", lb, -1, -1";
        os << ") && ok;";
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            if (groups[gi].surf != s) continue;
            os << " g.template add4<" << gi << ">(";
            for (int k = 0; k < 4; ++k) {
                if (k) os << ", ";
                if (groups[gi].sl[k] >= 0) os << "use ? gp[" << groups[gi].sl[k] << "] : 0.0";
                else os << "0.0";
            }
            os << ");";
        }
        os << " ok = use; }\n";
    }
    os << "      return ok;\n    });\n}\n";
    return os.str();
}

// hiprtc compilation of `src` to a gfx950 code object (no device needed)
static bool compile_code(const std::string& src, std::vector<char>& code, std::string& log) {
    hiprtcProgram prog;
    const char* headers[] = {kEmbedDevice, kEmbedAbi, kEmbedAdjoint};
    const char* names[] = {"rt_device.h", "rt_hip.h", "rt_adjoint.h"};
    if (hiprtcCreateProgram(&prog, src.c_str(), "rt_trace_jit.hip", 3, headers, names) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    // the atomic optimizer is off: it wraps every statistics add (already made
    // by one lane of a ballot count) in a second lane election and a multiply
    // by the active-lane count, about 130 VALU instructions per pass over the
    // echelle's surfaces; the counts are integer sums, the same either way.
    // Multiply-adds are fused within a source expression only (-ffp-contract=on),
    // so every kernel built from the stage code forms the same values: fused
    // across statements, which products become FMAs depends on the code
    // around them
    std::vector<const char*> opts = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-munsafe-fp-atomics",
                                     "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                                     "-mllvm", "-amdgpu-atomic-optimizer-strategy=None", "-ffp-contract=on"};
    hiprtcResult rc = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    if (log_size > 1) {
        std::vector<char> l(log_size + 1, 0);
        hiprtcGetProgramLog(prog, l.data());
        log = l.data();
    }
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    code.resize(code_size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return true;
}

static bool compile(const std::string& src, Entry& e, const char* name = "rt_trace_jit") {
    // RT_HIP_JIT_SOURCE_DIR: each generated source is also written there
    // (rt_jit_<n>.hip, in compilation order), for inspection
    if (const char* dir = getenv("RT_HIP_JIT_SOURCE_DIR")) {
        static int n_written = 0;
        char path[4096];
        snprintf(path, sizeof path, "%s/rt_jit_%03d.hip", dir, n_written++);
        if (FILE* f = fopen(path, "w")) {
            fputs(src.c_str(), f);
            fclose(f);
        }
    }
    std::vector<char> code;
    if (!compile_code(src, code, e.log)) return false;
    if (hipModuleLoadData(&e.module, code.data()) != hipSuccess) {
        e.log += "\nhipModuleLoadData failed";
        return false;
    }
    if (hipModuleGetFunction(&e.fn, e.module, name) != hipSuccess) {
        e.log += "\nhipModuleGetFunction failed";
        return false;
    }
    return true;
}

hipFunction_t rt_jit_kernel(const std::string& sig, int beam_mode, int variant, int block, int waves, int device,
                            bool batch) {
    std::ostringstream key;
    key << device << "|" << beam_mode << "|" << block << "|" << waves << "|v" << variant << "|"
        << (batch ? "batch|" : "") << sig;
    std::lock_guard<std::mutex> lk(g_mu);
    Entry& e = g_cache[key.str()];
    if (e.fn) return e.fn;
    if (e.failed) return nullptr;
    if (!compile(generate_source(sig, beam_mode, variant, block, waves, batch), e)) {
        e.failed = true;
        fprintf(stderr, "rt_hip: generated kernel failed to build, using the generic kernel:\n%s\n",
                e.log.c_str());
        return nullptr;
    }
    return e.fn;
}

hipFunction_t rt_jit_adjoint_kernel(const std::string& sig, int beam_type, const std::vector<int>& rows, int n_surf,
                                    int block, int waves, int device, bool batch, bool ltab) {
    std::ostringstream key;
    key << device << "|adjoint|" << (batch ? "batch|" : "") << (ltab ? "ltab|" : "") << beam_type << "|" << block
        << "|" << waves << "|";
    for (size_t k = 0; k < rows.size(); ++k)
        if (rows[k] >= 0) key << k << ":" << rows[k] << ",";
    key << "|" << sig;
    std::lock_guard<std::mutex> lk(g_mu);
    Entry& e = g_cache[key.str()];
    if (e.fn) return e.fn;
    if (e.failed) return nullptr;
    if (!compile(generate_adjoint_source(sig, beam_type, rows, n_surf, block, waves, batch, ltab), e,
                 "rt_adjoint_jit")) {
        e.failed = true;
        fprintf(stderr, "rt_hip: generated adjoint kernel failed to build, using the generic kernel:\n%s\n",
                e.log.c_str());
        return nullptr;
    }
    return e.fn;
}

// Generate and compile (without loading) the kernel of a device table; the
// source and the compiler log are returned in `text`.  Host only.
extern "C" int rt_jit_check_dtable(const double* dtable, int n_surf, int beam_mode, char* text, size_t len,
                                   size_t* code_size) {
    // a negative mode (rays given) carries no flags
    const bool batch = beam_mode >= 0 && (beam_mode & RT_JIT_BATCH) != 0;
    const bool adjoint = beam_mode >= 0 && (beam_mode & RT_JIT_ADJOINT) != 0;
    const bool record = beam_mode >= 0 && (beam_mode & RT_JIT_RECORD) != 0;
    if (beam_mode >= 0) beam_mode &= ~(RT_JIT_BATCH | RT_JIT_ADJOINT | RT_JIT_RECORD);
    const bool ltab = beam_mode >= 0 && (beam_mode & RT_JIT_LTAB) != 0;
    if (adjoint) beam_mode &= ~RT_JIT_LTAB;
    std::string src;
    if (adjoint) {
        // every slot a surface has (aspheric terms up to its count) asked for
        std::vector<int> rows((size_t)n_surf * rt::adj::PG_N, -1);
        int j = 0;
        for (int s = 0; s < n_surf; ++s) {
            const int npoly = (int)dtable[(size_t)s * rt::DEV_STRIDE + RT_S_NPOLY];
            for (int sl = 0; sl < rt::adj::PG_N; ++sl)
                if (sl < rt::adj::PG_A || sl >= rt::adj::PG_CR || sl - rt::adj::PG_A < npoly)
                    rows[(size_t)s * rt::adj::PG_N + sl] = j++;
        }
        src = generate_adjoint_source(rt_jit_signature(dtable, n_surf), beam_mode, rows, n_surf, rt::kBlock, 0, batch,
                                      ltab);
    } else {
        src = generate_source(rt_jit_signature(dtable, n_surf), beam_mode,
                              record ? (RT_JIT_VAR_LOSS | RT_JIT_VAR_EXT) : 0, rt::kBlock, 0, batch);
    }
    std::vector<char> code;
    std::string log;
    bool ok = compile_code(src, code, log);
    if (text && len) snprintf(text, len, "%s\n%s", src.c_str(), log.c_str());
    if (code_size) *code_size = code.size();
    return ok ? 0 : RT_ERR_FAILED;
}
