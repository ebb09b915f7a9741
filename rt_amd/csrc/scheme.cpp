// scheme.cpp -- librt_scheme_hip.so: the Scheme class of common-lib/scheme.h
// on the MI355X engine.
//
// load_from_file reads the reference's XML scheme format with rt_amd's own
// reader (rt_xml.h; the reference needs mxml) and applies the reference's
// attribute rules in the reference's order, so a malformed file fails with
// the same error (scheme.cpp:272-857; Surface/Lens/Grating setters and
// constructors, surface.cpp:85-202,400-560; Beam setters, beam.cpp:67-120).
// The surfaces and the beam are packed into the engine's table
// (include/rt_hip.h).
//
// run traces one wavelength range per device launch: rt_hip_trace_record
// generates the beam and runs every surface's transform, normalisation,
// intersection, aperture and action (scheme.cpp:983-1046) with the loss code
// of every ray and the logged extents; rt_hip_rearrange replays
// Beam::Rearrange (beam.cpp:574-598) after every stage that lost rays;
// rt_hip_beam_order lays the arrays out in Beam::Save's order with the lost
// rays' cosines renormalised as the reference's later Normalize calls leave
// them; the host writes the .dat file.  The log has the reference's lines
// (scheme.cpp:899-1096); its extents cover every good ray (the reference's
// GetBeamDim / GetBeamCosins, beam.cpp:490-523, stop one ray short).
//
// ComputeQE (scheme.cpp:1135-1338) runs the efficiency products on the
// engine (surface_QE, grating_energy_distr).
#include "rt/scheme.h"
#include "rt/rt_engine_func.h"
#include "rt_ascii.h"
#include "rt_hip.h"
#include "rt_xml.h"

#include <hip/hip_runtime.h>

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <map>
#include <sstream>

// the scheme format version this implementation reads (common-lib/version.h)
static const int kPackageMajor = 0;
static const int kPackageMinor = 2;

// ------------------------------------------------------------ exceptions

Scheme::bad_scheme::bad_scheme(SchemeError err) : std::exception(), Error(err) {}
Scheme::SchemeError Scheme::bad_scheme::scheme_error() const { return Error; }

Beam::bad_beam::bad_beam(BeamError err) : Error(err) {}
Beam::BeamError Beam::bad_beam::beam_error() const { return Error; }

Surface::bad_surface::bad_surface(SurfaceError err) : Error(err) {}
Surface::SurfaceError Surface::bad_surface::surface_error() const { return Error; }

Lens::bad_lens::bad_lens(Lens::LensError err) : Error(err) {}
Lens::LensError Lens::bad_lens::bad_lens_error() const { return Error; }

Grating::bad_grating::bad_grating(Grating::GratingError err) : Error(err) {}
Grating::GratingError Grating::bad_grating::bad_grating_error() const { return Error; }

namespace rt_scheme_detail {

// a refractive index (or QE) given as a number or a tabulated file
struct Fn {
    bool file = false;
    std::string path;
    TabulatedFunction tab;
};

struct Surf {
    Surface::SurfaceClass cls = Surface::Plane;
    Surface::SurfaceType type = Surface::Mirror;
    Surface::SurfaceShape shape = Surface::Circle;
    std::vector<real_t> params, dist, ang;
    real_t size[2] = {0.0, 0.0}, center[2] = {0.0, 0.0};
    std::string comment;
    Grating::GratingType grtype = Grating::Reflective;
    std::vector<long> order;
    real_t grconst = 1.0;
    Fn n1, n2;
};

struct BeamDesc {
    Beam::BeamType type = Beam::Parallel;
    Beam::BeamShape shape = Beam::Circle;
    Beam::BeamProfile profile = Beam::Random;
    Beam::BeamRangeDistr distr = Beam::UniformDistr;
    std::vector<real_t> params, center, range;
    bool composite = false;
    size_t nlambda = 1;
};

struct State {
    std::string file, name = "Unknown", result_root, log_name;
    std::vector<Surf> surfaces;
    BeamDesc beam;
    std::vector<std::vector<real_t>> ranges;
    std::unique_ptr<rtxml::Node> tree;          // kept by load_from_file(..., true) for ComputeQE
    uint64_t seed = 0;                          // first beam's seed (RT_SEED)
    uint64_t beams_made = 0;                    // beam k is drawn with seed + k
    bool loaded = false;
};

namespace {

// leading and trailing spaces and tabs (scheme.cpp:27-41)
void trim_spaces(std::string& str) {
    const char* ws = " \t";
    const size_t b = str.find_first_not_of(ws);
    if (b == std::string::npos) {
        str.clear();
        return;
    }
    const size_t e = str.find_last_not_of(ws);
    str = str.substr(b, e - b + 1);
}

void toupcase(std::string& s) {
    for (char& c : s) c = (char)toupper((unsigned char)c);
}

// the attribute accessors of scheme.cpp:131-212 over one element
struct Attrs {
    const rtxml::Node* node;

    const char* raw(const std::string& name) const {
        const rtxml::Node* e = rtxml::find_path(node, name);
        if (!e) throw Scheme::bad_scheme(Scheme::ParseFailure);
        const std::string* v = rtxml::value(e);
        if (!v) throw Scheme::bad_scheme(Scheme::BadAttrValue);
        return v->c_str();
    }

    std::string str(const std::string& name, bool upper = true) const {
        std::string v = raw(name);
        trim_spaces(v);
        if (v.empty()) throw Scheme::bad_scheme(Scheme::BadAttrValue);
        if (upper) toupcase(v);
        return v;
    }

    std::vector<real_t> num(const std::string& name, bool strict = false) const {
        const std::string v = str(name);
        const char* s = v.c_str();
        std::vector<real_t> nums;
        for (;;) {
            char* p = nullptr;
            const real_t x = strtod(s, &p);
            if (p == s) break;
            nums.push_back(x);
            s = p;
        }
        if (strict && s[0] != '\0') throw Scheme::bad_scheme(Scheme::NanAttrValue);
        if (nums.empty()) throw Scheme::bad_scheme(Scheme::BadAttrValue);
        return nums;
    }
};

// get_range (scheme.cpp:215-264)
std::vector<std::vector<real_t>> get_range(const Attrs& a, const std::string& name, unsigned n_gratings) {
    std::vector<std::vector<real_t>> ranges;
    try {
        std::vector<real_t> val = a.num(name, true);
        if (val.size() < 2 || val[0] >= val[1] || val[1] <= 0.0) throw Scheme::bad_scheme(Scheme::BadRange);
        ranges.push_back(val);
    } catch (Scheme::bad_scheme& ex) {
        if (ex.scheme_error() != Scheme::NanAttrValue) throw;
        const std::string file = a.str(name, false);
        rtascii::Reader rd(file);
        if (!rd.ok()) throw Scheme::bad_scheme(Scheme::NoRangeFile);
        int flag = 0;
        std::vector<real_t> v;
        while (rd.next(2 + n_gratings, flag, v)) {
            if (flag == rtascii::kInvalid) throw Scheme::bad_scheme(Scheme::BadRange);
            if (flag == rtascii::kData) {
                if (v[0] >= v[1] || v[0] <= 0.0 || v[1] <= 0.0) throw Scheme::bad_scheme(Scheme::BadRange);
                ranges.push_back(v);
            }
        }
        if (ranges.empty()) throw Scheme::bad_scheme(Scheme::BadRange);
    }
    return ranges;
}

// a number (strict) or else a file name (scheme.cpp:496-517)
void read_fn_attr(const Attrs& a, const char* name, bool& is_file, real_t& value, std::string& path) {
    try {
        value = a.num(name, true)[0];
        is_file = false;
    } catch (Scheme::bad_scheme&) {
        path = a.str(name, false);
        is_file = true;
    }
}

// the TabulatedFunction members of Lens / Grating: constants and files are
// set in member order (n1, then n2); the constructors that take a file catch
// bad_tab_func and throw their own BadRefIndexFile (surface.cpp:410-470,
// 500-560), the one with two constants lets it through
template <class Bad, class Err>
void make_fns(Fn& n1, Fn& n2, bool f1, real_t v1, const std::string& p1, bool f2, real_t v2,
              const std::string& p2, Err bad_file) {
    try {
        n1.file = f1;
        n1.path = p1;
        if (f1) n1.tab.Load(p1.c_str()); else n1.tab.Set(v1);
        n2.file = f2;
        n2.path = p2;
        if (f2) n2.tab.Load(p2.c_str()); else n2.tab.Set(v2);
    } catch (TabulatedFunction::bad_tab_func&) {
        if (f1 || f2) throw Bad(bad_file);
        throw;
    }
}

double surface_angle_rad(double deg) {
    // Surface::SetAngles, GetAngles and Beam::Transform (surface.cpp:201,227;
    // beam.cpp:609-611): deg *= deg2rad, /= deg2rad, *= deg2rad
    double a = deg;
    a *= deg2rad;
    a /= deg2rad;
    a *= deg2rad;
    return a;
}

// ---- packing into the engine table --------------------------------------

struct Packed {
    std::vector<double> table, pool;
    std::map<std::string, double> pool_at;   // file tables already in the pool
};

void fn_ref(Packed& pk, const Fn& f, double& kind, double& val) {
    if (!f.file || f.tab.IsSingleValue()) {
        kind = RT_FN_CONST;
        val = f.tab.Values().empty() ? 1.0 : f.tab.Values()[0];
        return;
    }
    auto it = pk.pool_at.find(f.path);
    if (it == pk.pool_at.end()) {
        const auto& x = f.tab.Abscissae();
        const auto& y = f.tab.Values();
        const auto& y2 = f.tab.SecondDerivatives();
        const double at = (double)pk.pool.size();
        pk.pool.push_back((double)x.size());
        pk.pool.insert(pk.pool.end(), x.begin(), x.end());
        pk.pool.insert(pk.pool.end(), y.begin(), y.end());
        pk.pool.insert(pk.pool.end(), y2.begin(), y2.end());
        it = pk.pool_at.emplace(f.path, at).first;
    }
    kind = RT_FN_SPLINE;
    val = it->second;
}

void pack_surface(Packed& pk, const Surf& s, long order) {
    double r[RT_SURF_STRIDE] = {0};
    r[RT_S_CLASS] = s.cls;
    r[RT_S_TYPE] = s.type;
    r[RT_S_SHAPE] = s.shape;
    if (s.cls == Surface::Conic || s.cls == Surface::Toric) {
        r[RT_S_C] = s.params[0];
        r[RT_S_K] = s.params[1];
        const size_t np = (size_t)s.params[2];
        if (np > RT_MAX_POLY) throw Surface::bad_surface(Surface::InvalidSurfaceParam);
        r[RT_S_NPOLY] = (double)np;
        for (size_t i = 0; i < np; ++i) r[RT_S_A + i] = s.params[3 + i];
        if (s.cls == Surface::Toric) r[RT_S_CR] = s.params[3 + np];
    }
    r[RT_S_SIZE0] = s.size[0];
    r[RT_S_SIZE1] = s.size[1];
    r[RT_S_CEN0] = s.center[0];
    r[RT_S_CEN1] = s.center[1];
    for (int j = 0; j < 3; ++j) {
        r[RT_S_DIST + j] = s.dist[j];
        const double a = surface_angle_rad(s.ang[j]);
        r[RT_S_ANG + j] = a;
        r[RT_S_SIN + j] = sin(a);
        r[RT_S_COS + j] = cos(a);
    }
    r[RT_S_GRTYPE] = s.grtype;
    r[RT_S_ORDER] = s.type == Surface::Grating ? (double)order : 0.0;
    r[RT_S_GRCONST] = s.grconst;
    r[RT_S_ORDMODE] = RT_ORDER_FIXED;
    r[RT_S_ORDMIN] = -1000000.0;
    r[RT_S_ORDMAX] = 1000000.0;
    r[RT_S_N1KIND] = RT_FN_CONST;
    r[RT_S_N1VAL] = 1.0;
    r[RT_S_N2KIND] = RT_FN_CONST;
    r[RT_S_N2VAL] = 1.0;
    if (s.type == Surface::Lens || s.type == Surface::Grating) {
        fn_ref(pk, s.n1, r[RT_S_N1KIND], r[RT_S_N1VAL]);
        fn_ref(pk, s.n2, r[RT_S_N2KIND], r[RT_S_N2VAL]);
    }
    // no CCD frame: the detector's QE and spectrum fields stay zero
    pk.table.insert(pk.table.end(), r, r + RT_SURF_STRIDE);
}

void pack_beam(const BeamDesc& b, const std::vector<real_t>& range, double* out) {
    for (int i = 0; i < RT_BEAM_STRIDE; ++i) out[i] = 0.0;
    const std::vector<real_t>& p = b.params;
    out[RT_B_TYPE] = b.type;
    out[RT_B_SHAPE] = b.shape;
    out[RT_B_RANGE_DISTR] = b.distr;
    out[RT_B_NPER] = (double)(size_t)p[0];
    out[RT_B_P1] = p[1];
    out[RT_B_P2] = p[2];
    if (b.shape == Beam::Rectangle) {
        out[RT_B_P3] = p[3];
        if (b.type != Beam::Parallel) out[RT_B_ANGPAR] = p[4];
    } else if (b.type != Beam::Parallel) {
        out[RT_B_ANGPAR] = p[3];
    }
    for (int j = 0; j < 3; ++j) out[RT_B_CEN + j] = b.center[j];
    out[RT_B_RANGE] = range[0];
    out[RT_B_RANGE + 1] = range.size() > 1 ? range[1] : range[0];
    out[RT_B_NLAMBDA] = (double)b.nlambda;
    out[RT_B_COMPOSITE] = b.composite ? 1.0 : 0.0;
}

size_t beam_rays(const BeamDesc& b) {
    const size_t n = (size_t)b.params[0];
    return b.composite ? n * b.nlambda : n;
}

// ---- device buffers -------------------------------------------------------

struct DevMem {
    void* p = nullptr;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T* get() const { return (T*)p; }
    void alloc(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Beam::bad_beam(Beam::MemAllocFailure);
        }
    }
};

struct HostMem {
    void* p = nullptr;
    HostMem() = default;
    HostMem(const HostMem&) = delete;
    HostMem& operator=(const HostMem&) = delete;
    ~HostMem() {
        if (p) (void)hipHostFree(p);
    }
    void alloc(size_t bytes) {
        if (hipHostMalloc(&p, bytes ? bytes : 8, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Beam::bad_beam(Beam::MemAllocFailure);
        }
    }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() {
        if (hipStreamCreate(&s) != hipSuccess) throw Beam::bad_beam(Beam::CreationFailure);
    }
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
};

std::string engine_error() {
    char buf[1024];
    rt_hip_last_error(buf, sizeof buf);
    return buf;
}

// Beam::Save's layout (beam.cpp:636-715) from a 7 x n device array in file order
void save_beam(const std::string& path, size_t n, size_t n_good, size_t n_lambda, const double* ordered_dev,
               hipStream_t st) {
    HostMem h;
    h.alloc(7 * n * sizeof(double));
    if (n && (hipMemcpyAsync(h.p, ordered_dev, 7 * n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
              hipStreamSynchronize(st) != hipSuccess))
        throw Beam::bad_beam(Beam::SaveFailure);
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) throw Beam::bad_beam(Beam::SaveFailure);
    const double* a = (const double*)h.p;
    bool ok = fwrite(&n, sizeof(size_t), 1, f) == 1 && fwrite(&n_good, sizeof(size_t), 1, f) == 1 &&
              fwrite(&n_lambda, sizeof(size_t), 1, f) == 1;
    for (int k = 0; ok && n_good && k < 7; ++k) ok = fwrite(a + (size_t)k * n, sizeof(double), n_good, f) == n_good;
    const size_t rest = n - n_good;
    for (int k = 0; ok && rest && k < 7; ++k)
        ok = fwrite(a + (size_t)k * n + n_good, sizeof(double), rest, f) == rest;
    if (fclose(f) != 0) ok = false;
    if (!ok) throw Beam::bad_beam(Beam::SaveFailure);
}

std::string num_str(double v) {
    std::ostringstream ss;
    ss << v;
    return ss.str();
}

}  // namespace

// the first range's packed table, pool and beam of a loaded scheme
int pack_for_tests(const Scheme& scheme, double* table, int max_surf, int* n_surf, double* pool, size_t max_pool,
                   size_t* pool_len, double* beam_desc, uint64_t* n_rays) {
    const State& S = *scheme.st;
    Packed pk;
    size_t i_grating = 0;
    for (const Surf& sf : S.surfaces) {
        long order = sf.order.empty() ? 0 : sf.order[0];
        if (sf.type == Surface::Grating && S.ranges.size() > 1) order = (long)S.ranges.front()[2 + i_grating++];
        pack_surface(pk, sf, order);
    }
    if (pk.pool.empty()) pk.pool.push_back(0.0);
    if ((int)S.surfaces.size() > max_surf || pk.pool.size() > max_pool) return 900;
    *n_surf = (int)S.surfaces.size();
    std::copy(pk.table.begin(), pk.table.end(), table);
    std::copy(pk.pool.begin(), pk.pool.end(), pool);
    *pool_len = pk.pool.size();
    pack_beam(S.beam, S.ranges.front(), beam_desc);
    *n_rays = beam_rays(S.beam);
    return 0;
}
}  // namespace rt_scheme_detail

using namespace rt_scheme_detail;

// ------------------------------------------------------------------ Scheme

Scheme::Scheme() : st(new State) {}

Scheme::Scheme(const std::string& filename) : st(new State) {
    st->file = filename;
    try {
        load_from_file(filename);
    } catch (bad_scheme& ex) {
        // scheme.cpp:77-115
        switch (ex.scheme_error()) {
            case FileOpenFailure: std::cerr << "Can not open input scheme file!!!\n"; break;
            case ParseFailure: std::cerr << "Can not parse input scheme desciption!!!\n"; break;
            case NoClassAttr: std::cerr << "There is no surface class description!!!\n"; break;
            case BadAttrValue: std::cerr << "Invalid surface attribute in the input scheme description!!!\n"; break;
            case MemAllocFailure: std::cerr << "Memory allocation failure!!!\n"; break;
            default: std::cerr << "Unknown error!!!\n";
        }
        destroy();
    }
}

Scheme::~Scheme() = default;

void Scheme::destroy() {
    const uint64_t seed = st->seed, made = st->beams_made;
    st.reset(new State);
    st->seed = seed;
    st->beams_made = made;
}

size_t Scheme::GetNumberOfSurfaces() const { return st->surfaces.size(); }

void Scheme::load_from_file(const std::string& filename, bool QE_flag) {
    if (!st->surfaces.empty() || st->tree) destroy();
    st->file = filename;
    if (const char* s = getenv("RT_SEED")) st->seed = strtoull(s, nullptr, 0);

    {
        FILE* f = fopen(filename.c_str(), "r");
        if (!f) throw bad_scheme(FileOpenFailure);
        fclose(f);
    }
    std::unique_ptr<rtxml::Node> tree;
    try {
        tree = rtxml::parse_file(filename);
    } catch (rtxml::ParseError&) {
        std::cerr << "XML PARSING FAILURE!\n";
        throw bad_scheme(ParseFailure);
    }
    if (QE_flag) {          // only the tree, for ComputeQE (scheme.cpp:361-363)
        st->tree = std::move(tree);
        return;
    }

    // general description (scheme.cpp:367-402)
    std::vector<const rtxml::Node*> gen = rtxml::elements(tree.get(), "general");
    if (gen.empty()) throw bad_scheme(NoSchemeDesc);
    Attrs ga{gen.front()};
    const std::string sname = ga.str("name", false);
    const std::string resroot = ga.str("result_file", false);
    std::string logname;
    try {
        logname = ga.str("log_file", false);
    } catch (bad_scheme&) {
        logname = "";
    }

    // surfaces (scheme.cpp:406-664), in document order
    std::vector<Surf> surfs;
    unsigned n_gratings = 0;
    for (const rtxml::Node* el : rtxml::elements(tree.get(), "surface")) {
        Attrs a{el};
        Surf s;
        std::string sclass = a.str("class"), stype, sshape;
        std::vector<real_t> spars, ssize, scen;
        if (sclass != "AUX") {
            stype = a.str("type");
            sshape = a.str("shape");
            spars = a.num("params");
            ssize = a.num("size");
            scen = a.num("center");
        } else {
            stype = "AUX";
            sshape = "CIRC";
        }
        std::vector<real_t> sdist = a.num("distance");
        std::vector<real_t> sangle = a.num("angles");

        if (sclass == "PLANE") s.cls = Surface::Plane;
        else if (sclass == "CONIC") s.cls = Surface::Conic;
        else if (sclass == "TORIC") s.cls = Surface::Toric;
        else if (sclass == "AUX") s.cls = Surface::AuxClass;
        else throw bad_scheme(BadAttrValue);

        if (stype == "MIRROR") s.type = Surface::Mirror;
        else if (stype == "LENS") s.type = Surface::Lens;
        else if (stype == "GRATING") s.type = Surface::Grating;
        else if (stype == "STOP") s.type = Surface::Stop;
        else if (stype == "SHADE") s.type = Surface::Shade;
        else if (stype == "DETECTOR") s.type = Surface::Detector;
        else if (stype == "AUX") s.type = Surface::AuxType;
        else throw bad_scheme(BadAttrValue);

        if (!sshape.compare(0, 4, "CIRC")) s.shape = Surface::Circle;
        else if (!sshape.compare(0, 4, "RECT")) s.shape = Surface::Rectangle;
        else throw bad_scheme(BadAttrValue);

        try {
            s.comment = a.str("comment", false);
        } catch (bad_scheme&) {
            s.comment = "";
        }

        bool f1 = false, f2 = false;
        real_t v1 = 1.0, v2 = 1.0;
        std::string p1, p2;
        if (s.type == Surface::Grating) {
            try {
                const std::string g = a.str("grating_type");
                if (g == "REFLECTIVE") s.grtype = Grating::Reflective;
                else if (g == "TRANSPARENT") s.grtype = Grating::Transparent;
                else throw Grating::BadType;            // an enumerator, as scheme.cpp:456 throws it
            } catch (bad_scheme&) {
                s.grtype = Grating::Reflective;
            }
            for (real_t o : a.num("order")) s.order.push_back((long)o);
            s.grconst = a.num("const")[0];
            read_fn_attr(a, "n1", f1, v1, p1);
            read_fn_attr(a, "n2", f2, v2, p2);
        } else if (s.type == Surface::Lens) {
            read_fn_attr(a, "n1", f1, v1, p1);
            read_fn_attr(a, "n2", f2, v2, p2);
        }

        // constructors (scheme.cpp:571-635; surface.cpp:400-560)
        if (s.type == Surface::Lens) {
            make_fns<Lens::bad_lens>(s.n1, s.n2, f1, v1, p1, f2, v2, p2, Lens::BadRefIndexFile);
        } else if (s.type == Surface::Grating) {
            make_fns<Grating::bad_grating>(s.n1, s.n2, f1, v1, p1, f2, v2, p2, Grating::BadRefIndexFile);
            if (s.order.empty()) throw Grating::bad_grating(Grating::BadOrder);
            if (s.order.size() > 1 && s.order[0] > s.order[1]) throw Grating::bad_grating(Grating::BadOrder);
            ++n_gratings;
        }
        if (s.type == Surface::AuxType) s.cls = Surface::Plane;     // Aux() is a plane (surface.cpp:702-705)

        // setters (scheme.cpp:640-650; surface.cpp:85-202)
        if (sdist.size() < 3) throw Surface::bad_surface(Surface::InvalidSurfaceDistance);
        s.dist = sdist;
        if (sangle.size() < 3) throw Surface::bad_surface(Surface::InvalidSurfaceAngles);
        s.ang = sangle;
        if (s.type != Surface::AuxType) {
            if (spars.size() < 2) throw Surface::bad_surface(Surface::InvalidSurfaceParam);
            if (s.cls == Surface::Toric) {
                if (spars.size() < 4 || spars.size() < 4 + (size_t)spars[2])
                    throw Surface::bad_surface(Surface::InvalidSurfaceParam);
                s.params = spars;
            } else if (s.cls == Surface::Conic) {
                if (spars.size() < 3 || spars.size() < 3 + (size_t)spars[2])
                    throw Surface::bad_surface(Surface::InvalidSurfaceParam);
                s.params = spars;
            }
            if (s.shape == Surface::Circle) {
                if (ssize.empty()) throw Surface::bad_surface(Surface::InvalidSurfaceSize);
                s.size[0] = ssize[0];
            } else {
                if (ssize.size() < 2) throw Surface::bad_surface(Surface::InvalidSurfaceSize);
                s.size[0] = ssize[0];
                s.size[1] = ssize[1];
            }
            if (scen.size() < 2) throw Surface::bad_surface(Surface::InvalidSurfaceCenter);
            s.center[0] = scen[0];
            s.center[1] = scen[1];
        }
        surfs.push_back(std::move(s));
    }
    if (surfs.empty()) throw bad_scheme(NoSurfaceDesc);

    // beam (scheme.cpp:668-856)
    std::vector<const rtxml::Node*> beams = rtxml::elements(tree.get(), "beam");
    if (beams.empty()) throw bad_scheme(NoBeamDesc);
    Attrs ba{beams.front()};
    const std::string btype = ba.str("type");
    const std::string bshape = ba.str("shape");
    const std::string bpr = ba.str("profile");
    const std::string bdistr = ba.str("range_distr");
    BeamDesc b;
    b.params = ba.num("params");
    b.center = ba.num("center");
    std::vector<std::vector<real_t>> ranges = get_range(ba, "range", n_gratings);

    if (btype == "PARALLEL") b.type = Beam::Parallel;
    else if (btype == "CONIC") b.type = Beam::Conic;
    else if (btype == "GAUSS") b.type = Beam::Gauss;
    else throw Beam::bad_beam(Beam::InvalidParams);
    if (!bshape.compare(0, 4, "CIRC")) b.shape = Beam::Circle;
    else if (!bshape.compare(0, 4, "RECT")) b.shape = Beam::Rectangle;
    else throw Beam::bad_beam(Beam::InvalidParams);
    if (bpr == "RANDOM") b.profile = Beam::Random;
    else if (bpr == "CONCENTRIC") b.profile = Beam::Concentric;
    else throw Beam::bad_beam(Beam::InvalidParams);
    if (bdistr == "RANDOM") b.distr = Beam::RandomDistr;
    else if (bdistr == "UNIFORM") b.distr = Beam::UniformDistr;
    else if (bdistr == "MONOCHROMATIC") b.distr = Beam::Monochromatic;
    else throw Beam::bad_beam(Beam::InvalidParams);

    // the reference reads an uninitialised nlambda for a non-composite,
    // non-monochromatic beam (scheme.cpp:822-834); here it is 1
    b.nlambda = 1;
    b.composite = false;
    if (b.distr != Beam::Monochromatic) {
        try {
            const std::vector<real_t> nl = ba.num("nlambda", true);
            b.nlambda = (size_t)nl[0];
            b.composite = true;
        } catch (bad_scheme&) {
            b.composite = false;
        }
    }
    // Beam::SetParams / SetCenter / SetRange (beam.cpp:67-120)
    const std::vector<real_t>& p = b.params;
    if (p.size() < 3) throw Beam::bad_beam(Beam::InvalidParams);
    if (b.shape == Beam::Rectangle && p.size() < 4) throw Beam::bad_beam(Beam::InvalidParams);
    if (b.type != Beam::Parallel) {
        if (b.shape == Beam::Circle && p.size() < 4) throw Beam::bad_beam(Beam::InvalidParams);
        if (b.shape == Beam::Rectangle && p.size() < 5) throw Beam::bad_beam(Beam::InvalidParams);
    }
    if (p[0] <= 0) throw Beam::bad_beam(Beam::InvalidParams);
    if (p[1] < 0 || p[2] <= 0) throw Beam::bad_beam(Beam::InvalidParams);
    if (b.shape == Beam::Circle && p[1] >= p[2]) throw Beam::bad_beam(Beam::InvalidParams);
    if (b.shape == Beam::Rectangle && p[3] <= 0) throw Beam::bad_beam(Beam::InvalidParams);
    if (b.nlambda == 0) b.nlambda = 1;
    if (b.center.size() < 3) throw Beam::bad_beam(Beam::InvalidParams);
    const std::vector<real_t>& r0 = ranges.front();
    if (b.distr != Beam::Monochromatic) {
        if (r0.size() < 2 || r0[0] >= r0[1] || r0[0] <= 0.0 || r0[1] <= 0.0)
            throw Beam::bad_beam(Beam::InvalidParams);
    } else if (r0[0] <= 0.0) {
        throw Beam::bad_beam(Beam::InvalidParams);
    }
    // a concentric profile gets a ray count but no positions in the
    // reference (beam.cpp:276-290,354-356): not generated here
    if (b.profile == Beam::Concentric) {
        std::cerr << "rt_amd: concentric beams are not generated (the reference leaves their positions unset)\n";
        throw Beam::bad_beam(Beam::InvalidParams);
    }

    st->name = sname;
    st->result_root = resroot;
    st->log_name = logname;
    st->surfaces = std::move(surfs);
    st->beam = b;
    st->ranges = ranges;
    st->loaded = true;
}

void Scheme::run(std::list<std::string>& resultfile_list, std::ostream& log_file, const size_t start_surface) {
    State& S = *st;
    if (S.surfaces.empty()) {
        log_file << "The scheme was not loaded!\n";
        throw Scheme::EmptyScheme;                 // an enumerator, as scheme.cpp:888 throws it
    }
    if (start_surface > S.surfaces.size()) throw Scheme::BadStartIndex;

    std::stringstream ss;
    ss << "\nRAY-TRAYCING PACKAGE:  version " << kPackageMajor << "." << kPackageMinor << "\n";
    log_file << ss.str();
    log_file << "\n  RAY-TRACING ENGINE INFO:\n";
    std::list<std::string> info;
    if (rt_engine_info(info) == ENGINE_ERROR_OK)
        for (const std::string& l : info) log_file << "    " << l << std::endl;
    log_file << "\n  The scheme was loaded from: " + S.file + "\n";
    log_file << "  Scheme description: " + S.name << std::endl;

    time_t rawtime, endtime;
    time(&rawtime);
    const std::string date = asctime(localtime(&rawtime));
    resultfile_list.clear();
    ss.str("");
    ss << "\nStarting ray-tracing procedure (" << date.substr(0, date.length() - 1) << ") ...\n";
    log_file << ss.str();

    try {
        if (rt_engine_init() != ENGINE_ERROR_OK) throw Beam::bad_beam(Beam::CreationFailure);
        Stream stream;
        hipStream_t hs = stream.s;
        const size_t n = beam_rays(S.beam);
        const size_t n_lambda = S.beam.nlambda;

        // the input beam (scheme.cpp:950-955)
        {
            double bd[RT_BEAM_STRIDE];
            pack_beam(S.beam, S.ranges.front(), bd);
            const uint64_t seed = S.seed + S.beams_made++;
            DevMem arr;
            arr.alloc(7 * n * sizeof(double));
            double* a = arr.get<double>();
            if (rt_hip_beam_positions(bd, seed, 0, n, a, a + n, a + 2 * n, hs) != RT_OK ||
                rt_hip_beam_directions(bd, seed, 0, n, a, a + n, a + 3 * n, a + 4 * n, a + 5 * n, hs) != RT_OK ||
                rt_hip_beam_wavelengths(bd, seed, 0, n, a + 6 * n, hs) != RT_OK) {
                std::cerr << "rt_amd: " << engine_error() << "\n";
                throw Beam::bad_beam(Beam::CreationFailure);
            }
            save_beam(S.result_root + "_input_beam.dat", n, n, n_lambda, a, hs);
        }
        ss.str("");
        ss << "\n  Number of wavelength ranges: " << S.ranges.size() << "\n";
        log_file << ss.str();
        ss.str("");
        ss << "\n  Total number of rays in beam: " << n << "\n";
        log_file << ss.str();

        const size_t n_surf = S.surfaces.size() - start_surface;
        DevMem rays, flag, loss, perm, ordered, ext;
        rays.alloc(7 * n * sizeof(double));
        flag.alloc(n);
        loss.alloc(n);
        perm.alloc(n * sizeof(uint64_t));
        ordered.alloc(7 * n * sizeof(double));
        ext.alloc((n_surf ? n_surf : 1) * RT_EXT_PER_SURF * sizeof(double));
        std::vector<double> exts(n_surf * RT_EXT_PER_SURF);

        // later Normalize calls a ray lost at each stage meets (scheme.cpp:993,1040)
        unsigned char renorm[256] = {0};
        for (size_t s = 0; s < n_surf; ++s) {
            unsigned later = 0;
            for (size_t t = s + 1; t < n_surf; ++t)
                later += S.surfaces[start_surface + t].type == Surface::AuxType ? 1u : 2u;
            for (int k = 0; k < 3; ++k) renorm[RT_LOSS_CODE(s, k)] = (unsigned char)std::min(255u, 1u + later);
        }

        for (size_t i_range = 0; i_range < S.ranges.size(); ++i_range) {
            const std::vector<real_t>& range = S.ranges[i_range];
            ss.str("");
            ss << "\n  *** Current range: [" << range[0] << "," << range[1] << "] ***\n";
            log_file << ss.str();

            // per-range diffraction orders (scheme.cpp:1024-1030) of the traced gratings
            Packed pk;
            size_t i_grating = 0;
            for (size_t s = start_surface; s < S.surfaces.size(); ++s) {
                const Surf& sf = S.surfaces[s];
                long order = sf.order.empty() ? 0 : sf.order[0];
                if (sf.type == Surface::Grating && S.ranges.size() > 1) order = (long)range[2 + i_grating++];
                pack_surface(pk, sf, order);
            }
            if (pk.pool.empty()) pk.pool.push_back(0.0);
            rt_hip_scheme* dev_scheme = nullptr;
            if (n_surf == 0 || rt_hip_scheme_create(pk.table.data(), (int)n_surf, pk.pool.data(), pk.pool.size(),
                                                    &dev_scheme) != RT_OK) {
                std::cerr << "rt_amd: " << engine_error() << "\n";
                throw Surface::bad_surface(Surface::InvalidSurfaceParam);
            }
            std::unique_ptr<rt_hip_scheme, int (*)(rt_hip_scheme*)> guard(dev_scheme, rt_hip_scheme_destroy);

            double bd[RT_BEAM_STRIDE];
            pack_beam(S.beam, range, bd);
            const uint64_t seed = S.seed + S.beams_made++;
            double* a = rays.get<double>();
            std::vector<unsigned long long> stats(RT_STATS_LEN, 0);
            DevMem stats_dev;
            stats_dev.alloc(RT_STATS_LEN * sizeof(unsigned long long));
            if (hipMemsetAsync(stats_dev.p, 0, RT_STATS_LEN * sizeof(unsigned long long), hs) != hipSuccess)
                throw Beam::bad_beam(Beam::CreationFailure);
            int rc = rt_hip_trace_record(dev_scheme, bd, seed, n, a, a + n, a + 2 * n, a + 3 * n, a + 4 * n,
                                         a + 5 * n, a + 6 * n, flag.get<signed char>(),
                                         RT_TRACE_GENERATE | RT_TRACE_STORE, stats_dev.get<unsigned long long>(),
                                         loss.get<unsigned char>(), ext.get<double>(), hs);
            const std::string trace_error = rc == RT_OK ? std::string() : engine_error();
            if (hipMemcpyAsync(stats.data(), stats_dev.p, RT_STATS_LEN * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, hs) != hipSuccess ||
                hipMemcpyAsync(exts.data(), ext.p, exts.size() * sizeof(double), hipMemcpyDeviceToHost, hs) !=
                    hipSuccess ||
                hipStreamSynchronize(hs) != hipSuccess)
                throw Surface::bad_surface(Surface::IntersectionFailure);
            // a wavelength outside a refractive-index table: the reference's
            // TabulatedFunction throws from inside the engine call
            if (rc == RT_ERR_BAD_VALUE && stats[RT_ST_SPLINE_OOB])
                throw TabulatedFunction::bad_tab_func(TabulatedFunction::OutOfRange);
            if (rc != RT_OK) {
                std::cerr << "rt_amd: " << trace_error << "\n";
                throw Surface::bad_surface(Surface::IntersectionFailure);
            }

            uint64_t counts[256];
            if (rt_hip_rearrange(loss.get<unsigned char>(), n, perm.get<uint64_t>(), counts, hs) != RT_OK) {
                std::cerr << "rt_amd: " << engine_error() << "\n";
                throw Beam::bad_beam(Beam::TransformationFailure);
            }

            // the log of the stage loop (scheme.cpp:983-1046), stopping where
            // the reference throws because a stage lost every ray
            uint64_t good = n;
            for (size_t s = 0; s < n_surf; ++s) {
                const Surf& sf = S.surfaces[start_surface + s];
                const double* e = exts.data() + s * RT_EXT_PER_SURF;
                ss.str("");
                ss << "\n  Current surface index: " << start_surface + s << " ( " << sf.comment << ")\n";
                log_file << ss.str();
                log_file << "    Coordinate system transformation ...  ";
                log_file << "OK\n";
                ss.str("");
                ss << "      Z-cosins range: [" << e[0] << "," << e[1] << "]" << std::endl;
                log_file << ss.str();
                if (sf.type == Surface::AuxType) continue;
                const uint64_t l_int = counts[RT_LOSS_CODE(s, RT_LOSS_INTERSECTION)];
                const uint64_t l_ap = counts[RT_LOSS_CODE(s, RT_LOSS_APERTURE)];
                const uint64_t l_act = counts[RT_LOSS_CODE(s, RT_LOSS_ACTION)];
                log_file << "    Computation of beam-to-surface intersection coordinates ...  ";
                if (l_int && l_int == good) throw Surface::bad_surface(Surface::NoIntersections);
                good -= l_int;
                ss.str("");
                ss << "OK  ( number of good rays: " << good << ")\n";
                log_file << ss.str();
                ss.str("");
                ss << "      Incident ray dimension: x = [" << e[2] << "," << e[3] << "], y = [" << e[4] << ","
                   << e[5] << "]" << std::endl;
                log_file << ss.str();
                log_file << "    Apply surface geometrical constrains ...  ";
                if (l_ap && l_ap == good) throw Surface::bad_surface(Surface::NoIntersections);
                good -= l_ap;
                ss.str("");
                ss << "OK  ( number of good rays: " << good << ")\n";
                log_file << ss.str();
                ss.str("");
                ss << "      Ray dimension on the surface: x = [" << e[6] << "," << e[7] << "], y = [" << e[8]
                   << "," << e[9] << "]" << std::endl;
                log_file << ss.str();
                log_file << "    Surface action ...  ";
                if (l_act && l_act == good) throw Surface::bad_surface(Surface::NoGoodRays);
                good -= l_act;
                ss.str("");
                ss << "OK  ( number of good rays: " << good << ")\n";
                log_file << ss.str();
                ss.str("");
                ss << "      Z-cosins range: [" << e[10] << "," << e[11] << "]" << std::endl;
                log_file << ss.str();
            }

            if (rt_hip_beam_order(n, perm.get<uint64_t>(), loss.get<unsigned char>(), renorm, a, a + n, a + 2 * n,
                                  a + 3 * n, a + 4 * n, a + 5 * n, a + 6 * n, ordered.get<double>(), hs) != RT_OK) {
                std::cerr << "rt_amd: " << engine_error() << "\n";
                throw Beam::bad_beam(Beam::SaveFailure);
            }
            ss.str("");
            if (S.ranges.size() > 1)
                ss << S.result_root << "_range_(" << range[0] << "-" << range[1] << ").dat";
            else
                ss << S.result_root << ".dat";
            save_beam(ss.str(), n, (size_t)counts[0], n_lambda, ordered.get<double>(), hs);
            resultfile_list.push_back(ss.str());
        }
    } catch (...) {
        log_file << "ERROR!\n";
        throw;
    }

    time(&endtime);
    const double elapsed = difftime(endtime, rawtime);
    unsigned long hours = 0;
    unsigned minutes = 0;
    double seconds = elapsed;
    if (elapsed > 3600.0) {
        hours = (unsigned long)(seconds / 3600.0);
        seconds -= hours * 3600.0;
    }
    minutes = (unsigned)(seconds / 60.0);
    seconds -= minutes * 60.0;
    ss.str("");
    ss << "\nFinished! ( elapsed time: ";
    if (hours) ss << hours << "hours, ";
    ss << minutes << " minutes, " << seconds << " secs )\n";
    log_file << ss.str();
}

void Scheme::ComputeQE(size_t Nlambda, char* QE_filename) {
    State& S = *st;
    if (!S.tree) throw bad_scheme(ParseFailure);
    struct QeSurf {
        bool grating = false;
        long order = 0;
        real_t alpha = 0, gamma = 0, blaze = 0, grconst = 1, rule_start = 0, rule_stop = 1;
        TabulatedFunction qe;
    };
    std::vector<QeSurf> surfs;
    unsigned n_gratings = 0;
    std::vector<std::vector<real_t>> ranges;
    {
        std::unique_ptr<rtxml::Node> tree = std::move(S.tree);     // released like mxmlDelete (scheme.cpp:1237-1246)
        // surfaces (scheme.cpp:1160-1222)
        for (const rtxml::Node* el : rtxml::elements(tree.get(), "surface")) {
            Attrs a{el};
            std::string sclass;
            try {
                sclass = a.str("class");
            } catch (bad_scheme&) {
                sclass = "";
            }
            if (sclass == "AUX") continue;
            const std::string stype = a.str("type");
            if (stype != "MIRROR" && stype != "LENS" && stype != "GRATING" && stype != "DETECTOR") continue;
            QeSurf q;
            if (stype == "GRATING") {
                q.grating = true;
                q.alpha = a.num("alpha")[0];
                q.gamma = a.num("gamma")[0];
                q.blaze = a.num("blaze_angle")[0];
                q.grconst = a.num("const")[0];
                q.order = (long)a.num("order")[0];
                std::vector<real_t> rp;
                try {
                    rp = a.num("rule_params", true);
                    if (rp.size() == 1) rp.insert(rp.begin(), 0.0);
                } catch (bad_scheme&) {
                    rp = {0.0, 1.0};
                }
                // Grating::SetRuleParams (surface.cpp:595-611)
                q.blaze = q.blaze * deg2rad;
                q.alpha = q.alpha * deg2rad;
                q.gamma = q.gamma * deg2rad;
                q.rule_start = std::min(1.0, std::max(0.0, rp[0]));
                q.rule_stop = std::min(1.0, std::max(0.0, rp[1]));
                if (q.rule_start >= q.rule_stop) throw Grating::bad_grating(Grating::BadRuleParams);
                ++n_gratings;
            }
            try {
                q.qe.Set(a.num("qe_curve", true)[0]);
            } catch (bad_scheme& ex) {
                if (ex.scheme_error() == ParseFailure) throw;
                q.qe.Load(a.str("qe_curve", false).c_str());
            }
            surfs.push_back(std::move(q));
        }
        if (surfs.empty()) throw bad_scheme(NoSurfaceDesc);
        std::vector<const rtxml::Node*> beams = rtxml::elements(tree.get(), "beam");
        if (beams.empty()) throw bad_scheme(NoBeamDesc);
        ranges = get_range(Attrs{beams.front()}, "range", n_gratings);
    }

    FILE* file = fopen(QE_filename, "wb");
    if (!file) throw bad_scheme(FileSaveFailure);
    try {
        std::vector<real_t> lambda(Nlambda), qe;
        for (const std::vector<real_t>& val : ranges) {
            const real_t d_lambda = (val[1] - val[0]) / (Nlambda - 1);
            for (size_t i = 0; i < Nlambda; ++i) lambda[i] = val[0] + d_lambda * i;
            qe.assign(Nlambda, 1.0);
            unsigned i_grating = 0;
            for (QeSurf& q : surfs) {
                // surface_QE, then for a grating its blaze function (surface.cpp:331-340,614-636)
                if (surface_QE(qe.size(), lambda.data(), qe.data(), q.qe) != ENGINE_ERROR_OK)
                    throw Surface::bad_surface(Surface::BadQE);
                if (!q.grating) continue;
                long order = q.order;
                if (ranges.size() > 1) order = (long)val[2 + i_grating++];
                if (grating_energy_distr(lambda.size(), lambda.data(), labs(order), q.blaze, q.alpha, q.gamma,
                                         q.grconst, q.rule_start, q.rule_stop, qe.data()) != ENGINE_ERROR_OK)
                    throw Surface::bad_surface(Surface::BadQE);
            }
            if (fwrite(&Nlambda, sizeof(size_t), 1, file) != 1 ||
                fwrite(lambda.data(), sizeof(real_t), Nlambda, file) != Nlambda ||
                fwrite(qe.data(), sizeof(real_t), Nlambda, file) != Nlambda)
                throw bad_scheme(FileSaveFailure);
        }
    } catch (...) {
        fclose(file);
        throw;
    }
    fclose(file);
}

// ------------------------------------------------------------ C entry points

extern "C" int rt_scheme_run_file(const char* scheme_file, const char* rt_files_list, const char* log_path) {
    Scheme scheme;
    std::list<std::string> results;
    const char* what = nullptr;
    std::string detail;
    try {
        scheme.load_from_file(scheme_file);
        if (log_path) {
            std::ofstream log(log_path);
            scheme.run(results, log);
        } else {
            scheme.run(results);
        }
        std::ofstream out(rt_files_list);
        for (const std::string& r : results) out << r << std::endl;
        return 0;
    } catch (std::bad_alloc&) {
        what = "Memory allocation error was detected!";
    } catch (Scheme::bad_scheme& ex) {
        what = "scheme error";
        detail = std::to_string((int)ex.scheme_error());
    } catch (Scheme::SchemeError e) {
        what = "scheme error";
        detail = std::to_string((int)e);
    } catch (Beam::bad_beam& ex) {
        what = "beam error";
        detail = std::to_string((int)ex.beam_error());
    } catch (Surface::bad_surface& ex) {
        what = "surface error";
        detail = std::to_string((int)ex.surface_error());
    } catch (Lens::bad_lens& ex) {
        what = "lens error";
        detail = std::to_string((int)ex.bad_lens_error());
    } catch (Grating::bad_grating& ex) {
        what = "grating error";
        detail = std::to_string((int)ex.bad_grating_error());
    } catch (Grating::GratingError e) {
        what = "grating error";
        detail = std::to_string((int)e);
    } catch (TabulatedFunction::bad_tab_func& ex) {
        what = "tabulated function error";
        detail = std::to_string((int)ex.bad_tab_func_error());
    } catch (...) {
        what = "Something is wrong!!!";
    }
    std::cerr << what << (detail.empty() ? "" : " ") << detail << "\n";
    return 1;
}

extern "C" int rt_scheme_compute_qe(const char* scheme_file, size_t nlambda, const char* qe_file) {
    Scheme scheme;
    try {
        scheme.load_from_file(scheme_file, true);
        std::vector<char> name(qe_file, qe_file + strlen(qe_file) + 1);
        scheme.ComputeQE(nlambda, name.data());
        return 0;
    } catch (Scheme::bad_scheme& ex) {
        std::cerr << "Error: " << ex.scheme_error() << "\n";
        return 1 + (int)ex.scheme_error();
    } catch (...) {
        std::cerr << "Something is wrong!\n";
        return 100;
    }
}

extern "C" int rt_scheme_pack_file(const char* scheme_file, double* table, int max_surf, int* n_surf, double* pool,
                                   size_t max_pool, size_t* pool_len, double* beam_desc, uint64_t* n_rays) {
    Scheme scheme;
    try {
        scheme.load_from_file(scheme_file);
    } catch (Scheme::bad_scheme& ex) {
        return 100 + (int)ex.scheme_error();
    } catch (Beam::bad_beam& ex) {
        return 200 + (int)ex.beam_error();
    } catch (Surface::bad_surface& ex) {
        return 300 + (int)ex.surface_error();
    } catch (Lens::bad_lens& ex) {
        return 400 + (int)ex.bad_lens_error();
    } catch (Grating::bad_grating& ex) {
        return 500 + (int)ex.bad_grating_error();
    } catch (Grating::GratingError e) {
        return 700 + (int)e;
    } catch (TabulatedFunction::bad_tab_func& ex) {
        return 600 + (int)ex.bad_tab_func_error();
    }
    return rt_scheme_detail::pack_for_tests(scheme, table, max_surf, n_surf, pool, max_pool, pool_len, beam_desc,
                                            n_rays);
}
