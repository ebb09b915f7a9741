// engine_api.cpp -- librt_engine_hip.so: the rt engine plugin API
// (common-lib/rt_engine_func.h:25-104) on MI355X.
//
// Each entry point keeps the reference's contract -- host arrays owned by
// the caller and updated in place, N_bad accumulated, flag only ever set to
// 0, RT_engine_error returned -- and runs its stage as a HIP kernel through
// the C ABI of librt_hip.so (include/rt_hip.h).  Differences from the CUDA
// engine: the whole array goes to the device at once (288 GB of HBM instead
// of 75%-of-free-memory chunks), N_bad is exact, a TabulatedFunction
// argument out of range is caught and reported as ENGINE_ERROR_BAD_VALUE,
// and the beam generators are a seeded Philox stream: the k-th call of
// beam_random in a process uses key RT_ENGINE_SEED + k (environment, default
// 0), and the direction/wavelength generators that follow use the same key,
// so one Beam::Create draws the rays the fused trace draws for that seed.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <list>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_set>
#include <vector>

#include "rt/rt_engine_func.h"
#include "rt_hip.h"

namespace {

const double kHC = 1.98644560233e-12;   // rt_cuda_engine_ccd.cu:19 (erg um)

std::mutex g_mu;
bool g_init = false;
uint64_t g_seed = 0;
uint64_t g_beam_calls = 0;   // beam_random calls so far
uint64_t g_beam_key = 0;     // key of the current beam

RT_engine_error ensure_init() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_init) return ENGINE_ERROR_OK;
    const char* s = getenv("RT_ENGINE_SEED");
    g_seed = s ? strtoull(s, nullptr, 0) : 0;
    g_beam_key = g_seed;
    const char* d = getenv("RT_ENGINE_DEVICE");
    if (rt_hip_init(d ? atoi(d) : 0) != RT_OK) return ENGINE_ERROR_INIT_FAILED;
    g_init = true;
    return ENGINE_ERROR_OK;
}

RT_engine_error status(int rc) { return (RT_engine_error)rc; }

// device buffers, grown on demand, kept for the life of the thread
struct DeviceArena {
    struct Slot { void* p = nullptr; size_t cap = 0; };
    Slot slot[16];
    void* get(int i, size_t bytes) {
        if (bytes == 0) bytes = 8;
        if (slot[i].cap < bytes) {
            if (slot[i].p) (void)hipFree(slot[i].p);
            slot[i].p = nullptr;
            slot[i].cap = 0;
            if (hipMalloc(&slot[i].p, bytes) != hipSuccess) return nullptr;
            slot[i].cap = bytes;
        }
        return slot[i].p;
    }
};
thread_local DeviceArena g_arena;

struct Rays {
    double *X = nullptr, *Y = nullptr, *Z = nullptr, *cX = nullptr, *cY = nullptr, *cZ = nullptr;
    signed char* flag = nullptr;
    unsigned long long* counters = nullptr;   // [0] N_bad, [1] Newton iterations
};

bool alloc_rays(size_t n, Rays& r, bool with_flag) {
    size_t b = n * sizeof(double);
    r.X = (double*)g_arena.get(0, b);
    r.Y = (double*)g_arena.get(1, b);
    r.Z = (double*)g_arena.get(2, b);
    r.cX = (double*)g_arena.get(3, b);
    r.cY = (double*)g_arena.get(4, b);
    r.cZ = (double*)g_arena.get(5, b);
    if (with_flag) r.flag = (signed char*)g_arena.get(6, n);
    r.counters = (unsigned long long*)g_arena.get(7, 2 * sizeof(unsigned long long));
    return r.X && r.Y && r.Z && r.cX && r.cY && r.cZ && (!with_flag || r.flag) && r.counters;
}

bool h2d(void* dst, const void* src, size_t bytes) {
    return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
}
bool d2h(void* dst, const void* src, size_t bytes) {
    return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
}

// packed record of a surface from the plugin arguments (surface.cpp:64-81 layout)
void make_record(Surface::SurfaceClass cls, const real_t* params, double* rec) {
    memset(rec, 0, RT_SURF_STRIDE * sizeof(double));
    rec[RT_S_CLASS] = (double)cls;
    rec[RT_S_TYPE] = RT_STOP;
    if (!params || cls == Surface::Plane || cls == Surface::AuxClass) return;
    rec[RT_S_C] = params[0];
    rec[RT_S_K] = params[1];
    size_t N = (size_t)params[2];
    if (N > RT_MAX_POLY) N = RT_MAX_POLY + 1;     // rejected by validation
    rec[RT_S_NPOLY] = (double)N;
    for (size_t i = 0; i < N && i < RT_MAX_POLY; ++i) rec[RT_S_A + i] = params[3 + i];
    if (cls == Surface::Toric) rec[RT_S_CR] = params[3 + N];
}

// n1(lambda)/n2(lambda) and m lambda/n2(lambda)/d per ray on the host, as the
// reference does (surface.cu:1094-1097,1565-1569); evaluated once per run of
// equal wavelengths.  false when a wavelength is outside a table.
bool host_indices(size_t n, const real_t* lambda, TabulatedFunction& n1, TabulatedFunction& n2, double* mu,
                  double* L, long order, double gr_const) {
    try {
        double last = NAN, m = 0.0, l = 0.0;
        for (size_t i = 0; i < n; ++i) {
            if (!(lambda[i] == last)) {
                last = lambda[i];
                double a = n1[last];
                double b = n2[last];
                m = a;
                m /= b;
                l = order * last / b / gr_const;
            }
            if (mu) mu[i] = m;
            if (L) L[i] = l;
        }
    } catch (...) {
        return false;
    }
    return true;
}

}  // namespace

// ------------------------------------------------------- init and info

RT_engine_error rt_engine_init() { return ensure_init(); }

RT_engine_error rt_engine_info(std::list<std::string>& info_str) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    char buf[4096];
    if (rt_hip_info(buf, sizeof buf) != RT_OK) return ENGINE_ERROR_FAILED;
    try {
        info_str.clear();
        std::istringstream ss(buf);
        std::string line;
        while (std::getline(ss, line))
            if (!line.empty()) info_str.push_back(line);
    } catch (std::bad_alloc&) {
        return ENGINE_ERROR_BAD_ALLOC;
    }
    return ENGINE_ERROR_OK;
}

// ------------------------------------------------------- beam

RT_engine_error beam_random(Beam::BeamShape shape, real_t* params, real_t* center, size_t Nrays, real_t* X,
                            real_t* Y, real_t* Z) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    if (!params || !center || !X || !Y || !Z) return ENGINE_ERROR_BAD_VALUE;
    if (shape != Beam::Circle && shape != Beam::Rectangle) return ENGINE_ERROR_FAILED;   // beam.cu:92
    double b[RT_BEAM_STRIDE] = {0};
    b[RT_B_TYPE] = RT_BEAM_PARALLEL;
    b[RT_B_SHAPE] = shape == Beam::Circle ? RT_CIRCLE : RT_RECTANGLE;
    b[RT_B_NPER] = (double)Nrays;
    b[RT_B_P1] = params[1];
    b[RT_B_P2] = params[2];
    if (shape == Beam::Rectangle) b[RT_B_P3] = params[3];
    for (int k = 0; k < 3; ++k) b[RT_B_CEN + k] = center[k];
    uint64_t key;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        key = g_seed + g_beam_calls++;
        g_beam_key = key;
    }
    Rays r;
    if (!alloc_rays(Nrays, r, false)) return ENGINE_ERROR_BAD_ALLOC;
    int rc = rt_hip_beam_positions(b, key, 0, Nrays, r.X, r.Y, r.Z, nullptr);
    if (rc != RT_OK) return status(rc);
    size_t bytes = Nrays * sizeof(double);
    if (!d2h(X, r.X, bytes) || !d2h(Y, r.Y, bytes) || !d2h(Z, r.Z, bytes)) return ENGINE_ERROR_FAILED;
    return ENGINE_ERROR_OK;
}

static RT_engine_error directions(size_t Nrays, int type, double angpar, real_t* X, real_t* Y, real_t* cX,
                                  real_t* cY, real_t* cZ) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    if (!cX || !cY || !cZ || (type == RT_BEAM_CONIC && (!X || !Y))) return ENGINE_ERROR_BAD_VALUE;
    double b[RT_BEAM_STRIDE] = {0};
    b[RT_B_TYPE] = type;
    b[RT_B_ANGPAR] = angpar;
    b[RT_B_ANGUNIT] = type == RT_BEAM_CONIC ? 1.0 : 0.0;   // the plugin passes the cone angle in radians
    uint64_t key;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        key = g_beam_key;
    }
    Rays r;
    if (!alloc_rays(Nrays, r, false)) return ENGINE_ERROR_BAD_ALLOC;
    size_t bytes = Nrays * sizeof(double);
    if (type == RT_BEAM_CONIC && (!h2d(r.X, X, bytes) || !h2d(r.Y, Y, bytes))) return ENGINE_ERROR_FAILED;
    int rc = rt_hip_beam_directions(b, key, 0, Nrays, r.X, r.Y, r.cX, r.cY, r.cZ, nullptr);
    if (rc != RT_OK) return status(rc);
    if (!d2h(cX, r.cX, bytes) || !d2h(cY, r.cY, bytes) || !d2h(cZ, r.cZ, bytes)) return ENGINE_ERROR_FAILED;
    return ENGINE_ERROR_OK;
}

RT_engine_error beam_conic_cosins(size_t Nrays, real_t conic_angle, real_t* X, real_t* Y, real_t* cX, real_t* cY,
                                  real_t* cZ) {
    return directions(Nrays, RT_BEAM_CONIC, conic_angle, X, Y, cX, cY, cZ);
}

RT_engine_error beam_gauss_cosins(size_t Nrays, real_t fwhm, real_t* cX, real_t* cY, real_t* cZ) {
    return directions(Nrays, RT_BEAM_GAUSS, fwhm, nullptr, nullptr, cX, cY, cZ);
}

static RT_engine_error wavelengths(size_t Nrays, int distr, const real_t* range, real_t* lambda) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    if (!range || !lambda) return ENGINE_ERROR_BAD_VALUE;
    double b[RT_BEAM_STRIDE] = {0};
    b[RT_B_RANGE_DISTR] = distr;
    b[RT_B_NPER] = (double)Nrays;
    b[RT_B_RANGE] = range[0];
    b[RT_B_RANGE + 1] = range[1];
    uint64_t key;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        key = g_beam_key;
    }
    double* d = (double*)g_arena.get(8, Nrays * sizeof(double));
    if (!d) return ENGINE_ERROR_BAD_ALLOC;
    int rc = rt_hip_beam_wavelengths(b, key, 0, Nrays, d, nullptr);
    if (rc != RT_OK) return status(rc);
    return d2h(lambda, d, Nrays * sizeof(double)) ? ENGINE_ERROR_OK : ENGINE_ERROR_FAILED;
}

RT_engine_error uniform_range_beam(size_t Nrays, real_t* range, real_t* lambda) {
    return wavelengths(Nrays, RT_RANGE_UNIFORM, range, lambda);
}

RT_engine_error random_range_beam(size_t Nrays, real_t* range, real_t* lambda) {
    return wavelengths(Nrays, RT_RANGE_RANDOM, range, lambda);
}

RT_engine_error beam_transform(size_t Nrays, real_t* angles, real_t* dist, real_t* X, real_t* Y, real_t* Z,
                               real_t* cX, real_t* cY, real_t* cZ) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    if (!angles || !dist || !X || !Y || !Z || !cX || !cY || !cZ) return ENGINE_ERROR_BAD_VALUE;
    if (Nrays == 0) return ENGINE_ERROR_OK;
    Rays r;
    if (!alloc_rays(Nrays, r, false)) return ENGINE_ERROR_BAD_ALLOC;
    size_t b = Nrays * sizeof(double);
    if (!h2d(r.X, X, b) || !h2d(r.Y, Y, b) || !h2d(r.Z, Z, b) || !h2d(r.cX, cX, b) || !h2d(r.cY, cY, b) ||
        !h2d(r.cZ, cZ, b))
        return ENGINE_ERROR_FAILED;
    int rc = rt_hip_stage_transform(angles, dist, Nrays, r.X, r.Y, r.Z, r.cX, r.cY, r.cZ, nullptr);
    if (rc != RT_OK) return status(rc);
    if (!d2h(X, r.X, b) || !d2h(Y, r.Y, b) || !d2h(Z, r.Z, b) || !d2h(cX, r.cX, b) || !d2h(cY, r.cY, b) ||
        !d2h(cZ, r.cZ, b))
        return ENGINE_ERROR_FAILED;
    return ENGINE_ERROR_OK;
}

// ------------------------------------------------------- surfaces

RT_engine_error surface_intersection(Surface::SurfaceClass surf_class, real_t* params, size_t Nrays, real_t* X,
                                     real_t* Y, real_t* Z, real_t* cX, real_t* cY, real_t* cZ, size_t* N_bad,
                                     beam_flag_t* flag) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    if (!N_bad || !flag || !X || !Y || !Z || !cX || !cY || !cZ) return ENGINE_ERROR_BAD_VALUE;
    if (surf_class == Surface::AuxClass || Nrays == 0) return ENGINE_ERROR_OK;   // surface.cu:634-635
    double rec[RT_SURF_STRIDE];
    make_record(surf_class, params, rec);
    Rays r;
    if (!alloc_rays(Nrays, r, true)) return ENGINE_ERROR_BAD_ALLOC;
    size_t b = Nrays * sizeof(double);
    if (!h2d(r.X, X, b) || !h2d(r.Y, Y, b) || !h2d(r.Z, Z, b) || !h2d(r.cX, cX, b) || !h2d(r.cY, cY, b) ||
        !h2d(r.cZ, cZ, b) || !h2d(r.flag, flag, Nrays) || hipMemset(r.counters, 0, 16) != hipSuccess)
        return ENGINE_ERROR_FAILED;
    int rc = rt_hip_stage_intersect(rec, Nrays, r.X, r.Y, r.Z, r.cX, r.cY, r.cZ, r.flag, r.counters, r.counters + 1,
                                    nullptr);
    if (rc != RT_OK) return status(rc);
    unsigned long long cnt[2];
    if (!d2h(X, r.X, b) || !d2h(Y, r.Y, b) || !d2h(Z, r.Z, b) || !d2h(flag, r.flag, Nrays) || !d2h(cnt, r.counters, 16))
        return ENGINE_ERROR_FAILED;
    *N_bad += (size_t)cnt[0];
    return ENGINE_ERROR_OK;
}

RT_engine_error surface_reflection(Surface::SurfaceClass surf_class, real_t* params, size_t Nrays, real_t* X,
                                   real_t* Y, real_t* Z, real_t* cX, real_t* cY, real_t* cZ) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    if (!X || !Y || !Z || !cX || !cY || !cZ) return ENGINE_ERROR_BAD_VALUE;
    if (surf_class == Surface::AuxClass || Nrays == 0) return ENGINE_ERROR_OK;
    double rec[RT_SURF_STRIDE];
    make_record(surf_class, params, rec);
    Rays r;
    if (!alloc_rays(Nrays, r, false)) return ENGINE_ERROR_BAD_ALLOC;
    size_t b = Nrays * sizeof(double);
    if (!h2d(r.X, X, b) || !h2d(r.Y, Y, b) || !h2d(r.Z, Z, b) || !h2d(r.cX, cX, b) || !h2d(r.cY, cY, b) ||
        !h2d(r.cZ, cZ, b))
        return ENGINE_ERROR_FAILED;
    int rc = rt_hip_stage_reflect(rec, Nrays, r.X, r.Y, r.Z, r.cX, r.cY, r.cZ, nullptr);
    if (rc != RT_OK) return status(rc);
    if (!d2h(cX, r.cX, b) || !d2h(cY, r.cY, b) || !d2h(cZ, r.cZ, b)) return ENGINE_ERROR_FAILED;
    return ENGINE_ERROR_OK;
}

static RT_engine_error refract_or_diffract(bool grating, Surface::SurfaceClass surf_class, int gr_type,
                                           real_t* params, long order, real_t gr_constant, size_t Nrays, real_t* X,
                                           real_t* Y, real_t* Z, real_t* cX, real_t* cY, real_t* cZ, real_t* lambda,
                                           TabulatedFunction& n1, TabulatedFunction& n2, size_t* N_bad,
                                           beam_flag_t* flag) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    if (!N_bad || !flag || !lambda || !X || !Y || !Z || !cX || !cY || !cZ) return ENGINE_ERROR_BAD_VALUE;
    if (grating && gr_type != RT_REFLECTIVE && gr_type != RT_TRANSPARENT) return ENGINE_ERROR_BAD_VALUE;
    if (surf_class == Surface::AuxClass || Nrays == 0) return ENGINE_ERROR_OK;
    double rec[RT_SURF_STRIDE];
    make_record(surf_class, params, rec);
    rec[RT_S_GRTYPE] = grating ? gr_type : 0;
    std::vector<double> mu, L;
    bool need_mu = !grating || gr_type == RT_TRANSPARENT;
    try {
        if (need_mu) mu.resize(Nrays);
        if (grating) L.resize(Nrays);
    } catch (std::bad_alloc&) {
        return ENGINE_ERROR_BAD_ALLOC;
    }
    if (!host_indices(Nrays, lambda, n1, n2, need_mu ? mu.data() : nullptr, grating ? L.data() : nullptr, order,
                      gr_constant))
        return ENGINE_ERROR_BAD_VALUE;
    Rays r;
    if (!alloc_rays(Nrays, r, true)) return ENGINE_ERROR_BAD_ALLOC;
    size_t b = Nrays * sizeof(double);
    double* dmu = need_mu ? (double*)g_arena.get(9, b) : nullptr;
    double* dL = grating ? (double*)g_arena.get(10, b) : nullptr;
    if ((need_mu && !dmu) || (grating && !dL)) return ENGINE_ERROR_BAD_ALLOC;
    if (!h2d(r.X, X, b) || !h2d(r.Y, Y, b) || !h2d(r.Z, Z, b) || !h2d(r.cX, cX, b) || !h2d(r.cY, cY, b) ||
        !h2d(r.cZ, cZ, b) || !h2d(r.flag, flag, Nrays) || hipMemset(r.counters, 0, 16) != hipSuccess ||
        (need_mu && !h2d(dmu, mu.data(), b)) || (grating && !h2d(dL, L.data(), b)))
        return ENGINE_ERROR_FAILED;
    int rc = grating ? rt_hip_stage_diffract(rec, Nrays, r.X, r.Y, r.Z, r.cX, r.cY, r.cZ, dmu, dL, r.flag, r.counters,
                                             nullptr)
                     : rt_hip_stage_refract(rec, Nrays, r.X, r.Y, r.Z, r.cX, r.cY, r.cZ, dmu, r.flag, r.counters,
                                            nullptr);
    if (rc != RT_OK) return status(rc);
    unsigned long long cnt[2];
    if (!d2h(cX, r.cX, b) || !d2h(cY, r.cY, b) || !d2h(cZ, r.cZ, b) || !d2h(flag, r.flag, Nrays) ||
        !d2h(cnt, r.counters, 16))
        return ENGINE_ERROR_FAILED;
    *N_bad += (size_t)cnt[0];
    return ENGINE_ERROR_OK;
}

RT_engine_error surface_refraction(Surface::SurfaceClass surf_class, real_t* params, size_t Nrays, real_t* X,
                                   real_t* Y, real_t* Z, real_t* cX, real_t* cY, real_t* cZ, real_t* lambda,
                                   TabulatedFunction& n1, TabulatedFunction& n2, size_t* N_bad, beam_flag_t* flag) {
    return refract_or_diffract(false, surf_class, 0, params, 0, 1.0, Nrays, X, Y, Z, cX, cY, cZ, lambda, n1, n2,
                               N_bad, flag);
}

RT_engine_error surface_diffration(Surface::SurfaceClass surf_class, Grating::GratingType gr_type, real_t* params,
                                   long diff_order, real_t gr_constant, size_t Nrays, real_t* X, real_t* Y,
                                   real_t* Z, real_t* cX, real_t* cY, real_t* cZ, real_t* lambda,
                                   TabulatedFunction& n1, TabulatedFunction& n2, size_t* N_bad, beam_flag_t* flag) {
    int gt = gr_type == Grating::Reflective ? RT_REFLECTIVE : gr_type == Grating::Transparent ? RT_TRANSPARENT : -1;
    return refract_or_diffract(true, surf_class, gt, params, diff_order, gr_constant, Nrays, X, Y, Z, cX, cY, cZ,
                               lambda, n1, n2, N_bad, flag);
}

// ------------------------------------------------------- efficiency

RT_engine_error surface_QE(size_t N, real_t* lambda, real_t* spec, TabulatedFunction& QE) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    if (!lambda || !spec) return ENGINE_ERROR_BAD_VALUE;
    if (N == 0) return ENGINE_ERROR_OK;
    std::vector<double> q(N);
    try {
        for (size_t i = 0; i < N; ++i) q[i] = QE[lambda[i]];       // surface.cu:1810
    } catch (...) {
        return ENGINE_ERROR_BAD_VALUE;
    }
    size_t b = N * sizeof(double);
    double* dv = (double*)g_arena.get(0, b);
    double* df = (double*)g_arena.get(1, b);
    if (!dv || !df) return ENGINE_ERROR_BAD_ALLOC;
    if (!h2d(dv, spec, b) || !h2d(df, q.data(), b)) return ENGINE_ERROR_FAILED;
    int rc = rt_hip_scale(N, dv, df, nullptr);
    if (rc != RT_OK) return status(rc);
    return d2h(spec, dv, b) ? ENGINE_ERROR_OK : ENGINE_ERROR_FAILED;
}

RT_engine_error grating_energy_distr(size_t N, real_t* lambda, long order, real_t blaze_angle, real_t alpha,
                                     real_t gamma, real_t gr_const, real_t rule_start, real_t rule_stop,
                                     real_t* energy_distr) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    if (!lambda || !energy_distr) return ENGINE_ERROR_BAD_VALUE;
    if (N == 0) return ENGINE_ERROR_OK;
    size_t b = N * sizeof(double);
    double* dl = (double*)g_arena.get(0, b);
    double* de = (double*)g_arena.get(1, b);
    if (!dl || !de) return ENGINE_ERROR_BAD_ALLOC;
    if (!h2d(dl, lambda, b) || !h2d(de, energy_distr, b)) return ENGINE_ERROR_FAILED;
    int rc = rt_hip_grating_energy(N, dl, order, blaze_angle, alpha, gamma, gr_const, rule_start, rule_stop, de,
                                   nullptr);
    if (rc != RT_OK) return status(rc);
    return d2h(energy_distr, de, b) ? ENGINE_ERROR_OK : ENGINE_ERROR_FAILED;
}

// ------------------------------------------------------- memory

// Pinned host memory, so the caller's beam arrays (beam.cpp:306-325) move to
// and from HBM at full PCIe rate; plain malloc when pinning fails.
static std::mutex g_pin_mu;
static std::unordered_set<void*> g_pinned;

real_t* rt_engine_alocate_vector(size_t N_elem) {
    void* p = nullptr;
    if (ensure_init() == ENGINE_ERROR_OK && hipHostMalloc(&p, N_elem * sizeof(real_t), hipHostMallocDefault) == hipSuccess) {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        g_pinned.insert(p);
        return (real_t*)p;
    }
    return (real_t*)malloc(sizeof(real_t) * N_elem);
}

void rt_engine_free_vector(real_t* data) {
    if (!data) return;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        auto it = g_pinned.find(data);
        if (it != g_pinned.end()) {
            g_pinned.erase(it);
            (void)hipHostFree(data);
            return;
        }
    }
    free(data);
}

// ------------------------------------------------------- splines

RT_engine_error spline_init(size_t N, real_t* x, real_t* y, real_t* y2) {
    return status(rt_hip_spline_init(N, x, y, y2));
}

// ------------------------------------------------------- CCD

RT_engine_error ccd_coordinates(size_t Nrays, real_t* X, real_t* Y, real_t ccd_xsize, real_t ccd_ysize,
                                real_t ccd_xpix, real_t ccd_ypix, unsigned int* ccd_X, unsigned int* ccd_Y) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    if (!X || !Y || !ccd_X || !ccd_Y) return ENGINE_ERROR_BAD_VALUE;
    if (Nrays == 0) return ENGINE_ERROR_OK;
    size_t b = Nrays * sizeof(double), u = Nrays * sizeof(unsigned);
    double* dX = (double*)g_arena.get(0, b);
    double* dY = (double*)g_arena.get(1, b);
    unsigned* cx = (unsigned*)g_arena.get(11, u);
    unsigned* cy = (unsigned*)g_arena.get(12, u);
    if (!dX || !dY || !cx || !cy) return ENGINE_ERROR_BAD_ALLOC;
    if (!h2d(dX, X, b) || !h2d(dY, Y, b)) return ENGINE_ERROR_FAILED;
    int rc = rt_hip_ccd_coordinates(Nrays, dX, dY, ccd_xsize, ccd_ysize, ccd_xpix, ccd_ypix, cx, cy, nullptr);
    if (rc != RT_OK) return status(rc);
    return (d2h(ccd_X, cx, u) && d2h(ccd_Y, cy, u)) ? ENGINE_ERROR_OK : ENGINE_ERROR_FAILED;
}

RT_engine_error compute_ccd_image(size_t Nrays, unsigned int* ccd_X, unsigned int* ccd_Y, real_t* lambda,
                                  TabulatedFunction& QE, TabulatedFunction& incident_spec, size_t ccd_xdim,
                                  size_t ccd_ydim, real_t ccd_xpix, real_t ccd_ypix, real_t* image) {
    RT_engine_error e = ensure_init();
    if (e != ENGINE_ERROR_OK) return e;
    if (!ccd_X || !ccd_Y || !lambda || !image) return ENGINE_ERROR_BAD_VALUE;
    if (Nrays == 0) return ENGINE_ERROR_OK;
    // photons per ray: spec(lambda) lambda / hc QE(lambda) pixel area (ccd.cu:228,286-287,360)
    std::vector<double> w(Nrays);
    const double pix_area = ccd_xpix * ccd_ypix * 1.0E-2;
    try {
        for (size_t i = 0; i < Nrays; ++i) {
            double q = QE[lambda[i]];
            double s = incident_spec[lambda[i]] * lambda[i];
            w[i] = s / kHC * q * pix_area;
        }
    } catch (...) {
        return ENGINE_ERROR_BAD_VALUE;
    }
    size_t u = Nrays * sizeof(unsigned), b = Nrays * sizeof(double), ib = ccd_xdim * ccd_ydim * sizeof(double);
    unsigned* cx = (unsigned*)g_arena.get(11, u);
    unsigned* cy = (unsigned*)g_arena.get(12, u);
    double* dw = (double*)g_arena.get(0, b);
    double* dimg = (double*)g_arena.get(13, ib);
    if (!cx || !cy || !dw || !dimg) return ENGINE_ERROR_BAD_ALLOC;
    if (!h2d(cx, ccd_X, u) || !h2d(cy, ccd_Y, u) || !h2d(dw, w.data(), b) || !h2d(dimg, image, ib))
        return ENGINE_ERROR_FAILED;
    int rc = rt_hip_ccd_accumulate(Nrays, cx, cy, dw, ccd_xdim, ccd_ydim, dimg, nullptr, nullptr, nullptr);
    if (rc != RT_OK) return status(rc);
    return d2h(image, dimg, ib) ? ENGINE_ERROR_OK : ENGINE_ERROR_FAILED;
}
