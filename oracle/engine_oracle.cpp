// engine_oracle.cpp -- TEST INFRASTRUCTURE.  The 19 entry points of the rt
// engine plugin API (include/rt/rt_engine_func.h) computed on the host by the
// oracle's per-stage functions, with the same conventions as
// librt_engine_hip.so (seeded beams: key RT_ENGINE_SEED + k for the k-th
// beam_random call; N_bad exact; out-of-range tables -> ENGINE_ERROR_BAD_VALUE).
// It lets the reference's own host code (beam.cpp, surface.cpp, ...) drive a
// CPU engine in the tests, beside the same code driving the HIP engine.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <list>
#include <mutex>
#include <string>
#include <vector>

#include "rt/rt_engine_func.h"
#include "rt_hip.h"
#include "rt_oracle.h"

namespace {

const double kHC = 1.98644560233e-12;   // rt_cuda_engine_ccd.cu:19

std::mutex g_mu;
bool g_init = false;
uint64_t g_seed = 0, g_calls = 0, g_key = 0;

void init_once() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_init) return;
    const char* s = getenv("RT_ENGINE_SEED");
    g_seed = s ? strtoull(s, nullptr, 0) : 0;
    g_key = g_seed;
    g_init = true;
}

void make_record(Surface::SurfaceClass cls, const real_t* params, double* rec) {
    memset(rec, 0, RT_SURF_STRIDE * sizeof(double));
    rec[RT_S_CLASS] = (double)cls;
    rec[RT_S_TYPE] = RT_STOP;
    if (!params || cls == Surface::Plane || cls == Surface::AuxClass) return;
    rec[RT_S_C] = params[0];
    rec[RT_S_K] = params[1];
    size_t N = (size_t)params[2];
    if (N > RT_MAX_POLY) N = RT_MAX_POLY;
    rec[RT_S_NPOLY] = (double)N;
    for (size_t i = 0; i < N; ++i) rec[RT_S_A + i] = params[3 + i];
    if (cls == Surface::Toric) rec[RT_S_CR] = params[3 + (size_t)params[2]];
}

}  // namespace

RT_engine_error rt_engine_init() { init_once(); return ENGINE_ERROR_OK; }

RT_engine_error rt_engine_info(std::list<std::string>& info_str) {
    info_str.clear();
    info_str.push_back("rt oracle engine (host, FP64)");
    return ENGINE_ERROR_OK;
}

RT_engine_error beam_random(Beam::BeamShape shape, real_t* params, real_t* center, size_t Nrays, real_t* X,
                            real_t* Y, real_t* Z) {
    init_once();
    if (!params || !center || !X || !Y || !Z) return ENGINE_ERROR_BAD_VALUE;
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
        key = g_seed + g_calls++;
        g_key = key;
    }
    oracle_beam_positions(b, key, 0, Nrays, X, Y, Z);
    return ENGINE_ERROR_OK;
}

static RT_engine_error directions(size_t n, int type, double angpar, real_t* X, real_t* Y, real_t* cX, real_t* cY,
                                  real_t* cZ) {
    init_once();
    double b[RT_BEAM_STRIDE] = {0};
    b[RT_B_TYPE] = type;
    b[RT_B_ANGPAR] = angpar;
    b[RT_B_ANGUNIT] = type == RT_BEAM_CONIC ? 1.0 : 0.0;
    oracle_beam_directions(b, g_key, 0, n, X, Y, cX, cY, cZ);
    return ENGINE_ERROR_OK;
}

RT_engine_error beam_conic_cosins(size_t Nrays, real_t conic_angle, real_t* X, real_t* Y, real_t* cX, real_t* cY,
                                  real_t* cZ) {
    return directions(Nrays, RT_BEAM_CONIC, conic_angle, X, Y, cX, cY, cZ);
}

RT_engine_error beam_gauss_cosins(size_t Nrays, real_t fwhm, real_t* cX, real_t* cY, real_t* cZ) {
    return directions(Nrays, RT_BEAM_GAUSS, fwhm, nullptr, nullptr, cX, cY, cZ);
}

static RT_engine_error wavelengths(size_t n, int distr, const real_t* range, real_t* lambda) {
    init_once();
    double b[RT_BEAM_STRIDE] = {0};
    b[RT_B_RANGE_DISTR] = distr;
    b[RT_B_NPER] = (double)n;
    b[RT_B_RANGE] = range[0];
    b[RT_B_RANGE + 1] = range[1];
    oracle_beam_wavelengths(b, g_key, 0, n, lambda);
    return ENGINE_ERROR_OK;
}

RT_engine_error uniform_range_beam(size_t Nrays, real_t* range, real_t* lambda) {
    return wavelengths(Nrays, RT_RANGE_UNIFORM, range, lambda);
}

RT_engine_error random_range_beam(size_t Nrays, real_t* range, real_t* lambda) {
    return wavelengths(Nrays, RT_RANGE_RANDOM, range, lambda);
}

RT_engine_error beam_transform(size_t Nrays, real_t* angles, real_t* dist, real_t* X, real_t* Y, real_t* Z,
                               real_t* cX, real_t* cY, real_t* cZ) {
    oracle_beam_transform(Nrays, angles, dist, X, Y, Z, cX, cY, cZ);
    return ENGINE_ERROR_OK;
}

RT_engine_error surface_intersection(Surface::SurfaceClass surf_class, real_t* params, size_t Nrays, real_t* X,
                                     real_t* Y, real_t* Z, real_t* cX, real_t* cY, real_t* cZ, size_t* N_bad,
                                     beam_flag_t* flag) {
    if (surf_class == Surface::AuxClass) return ENGINE_ERROR_OK;
    double rec[RT_SURF_STRIDE];
    make_record(surf_class, params, rec);
    uint64_t nb = 0;
    oracle_surface_intersection(rec, Nrays, X, Y, Z, cX, cY, cZ, &nb, (signed char*)flag, nullptr);
    *N_bad += (size_t)nb;
    return ENGINE_ERROR_OK;
}

RT_engine_error surface_reflection(Surface::SurfaceClass surf_class, real_t* params, size_t Nrays, real_t* X,
                                   real_t* Y, real_t* Z, real_t* cX, real_t* cY, real_t* cZ) {
    if (surf_class == Surface::AuxClass) return ENGINE_ERROR_OK;
    double rec[RT_SURF_STRIDE];
    make_record(surf_class, params, rec);
    oracle_surface_reflection(rec, Nrays, X, Y, Z, cX, cY, cZ);
    return ENGINE_ERROR_OK;
}

static bool indices(size_t n, const real_t* lambda, TabulatedFunction& n1, TabulatedFunction& n2,
                    std::vector<double>& mu, std::vector<double>* L, long order, double d) {
    try {
        for (size_t i = 0; i < n; ++i) {
            double a = n1[lambda[i]], b = n2[lambda[i]];
            mu[i] = a / b;
            if (L) (*L)[i] = order * lambda[i] / b / d;
        }
    } catch (...) {
        return false;
    }
    return true;
}

RT_engine_error surface_refraction(Surface::SurfaceClass surf_class, real_t* params, size_t Nrays, real_t* X,
                                   real_t* Y, real_t* Z, real_t* cX, real_t* cY, real_t* cZ, real_t* lambda,
                                   TabulatedFunction& n1, TabulatedFunction& n2, size_t* N_bad, beam_flag_t* flag) {
    if (surf_class == Surface::AuxClass) return ENGINE_ERROR_OK;
    double rec[RT_SURF_STRIDE];
    make_record(surf_class, params, rec);
    std::vector<double> mu(Nrays);
    if (!indices(Nrays, lambda, n1, n2, mu, nullptr, 0, 1.0)) return ENGINE_ERROR_BAD_VALUE;
    uint64_t nb = 0;
    oracle_surface_refraction(rec, Nrays, X, Y, Z, cX, cY, cZ, mu.data(), &nb, (signed char*)flag);
    *N_bad += (size_t)nb;
    return ENGINE_ERROR_OK;
}

RT_engine_error surface_diffration(Surface::SurfaceClass surf_class, Grating::GratingType gr_type, real_t* params,
                                   long diff_order, real_t gr_constant, size_t Nrays, real_t* X, real_t* Y,
                                   real_t* Z, real_t* cX, real_t* cY, real_t* cZ, real_t* lambda,
                                   TabulatedFunction& n1, TabulatedFunction& n2, size_t* N_bad, beam_flag_t* flag) {
    if (surf_class == Surface::AuxClass) return ENGINE_ERROR_OK;
    double rec[RT_SURF_STRIDE];
    make_record(surf_class, params, rec);
    rec[RT_S_GRTYPE] = gr_type == Grating::Reflective ? RT_REFLECTIVE : RT_TRANSPARENT;
    std::vector<double> mu(Nrays), L(Nrays);
    if (!indices(Nrays, lambda, n1, n2, mu, &L, diff_order, gr_constant)) return ENGINE_ERROR_BAD_VALUE;
    uint64_t nb = 0;
    oracle_surface_diffraction(rec, Nrays, X, Y, Z, cX, cY, cZ, mu.data(), L.data(), &nb, (signed char*)flag);
    *N_bad += (size_t)nb;
    return ENGINE_ERROR_OK;
}

RT_engine_error surface_QE(size_t N, real_t* lambda, real_t* spec, TabulatedFunction& QE) {
    try {
        for (size_t i = 0; i < N; ++i) spec[i] *= QE[lambda[i]];
    } catch (...) {
        return ENGINE_ERROR_BAD_VALUE;
    }
    return ENGINE_ERROR_OK;
}

RT_engine_error grating_energy_distr(size_t N, real_t* lambda, long order, real_t blaze_angle, real_t alpha,
                                     real_t gamma, real_t gr_const, real_t rule_start, real_t rule_stop,
                                     real_t* energy_distr) {
    oracle_grating_energy_distr(N, lambda, order, blaze_angle, alpha, gamma, gr_const, rule_start, rule_stop,
                                energy_distr);
    return ENGINE_ERROR_OK;
}

void rt_engine_free_vector(real_t* data) { free(data); }

real_t* rt_engine_alocate_vector(size_t N_elem) { return (real_t*)malloc(sizeof(real_t) * N_elem); }

RT_engine_error spline_init(size_t N, real_t* x, real_t* y, real_t* y2) {
    return oracle_spline_init(N, x, y, y2) == RT_OK ? ENGINE_ERROR_OK : ENGINE_ERROR_BAD_VALUE;
}

RT_engine_error ccd_coordinates(size_t Nrays, real_t* X, real_t* Y, real_t ccd_xsize, real_t ccd_ysize,
                                real_t ccd_xpix, real_t ccd_ypix, unsigned int* ccd_X, unsigned int* ccd_Y) {
    oracle_ccd_coordinates(Nrays, X, Y, ccd_xsize, ccd_ysize, ccd_xpix, ccd_ypix, ccd_X, ccd_Y);
    return ENGINE_ERROR_OK;
}

RT_engine_error compute_ccd_image(size_t Nrays, unsigned int* ccd_X, unsigned int* ccd_Y, real_t* lambda,
                                  TabulatedFunction& QE, TabulatedFunction& incident_spec, size_t ccd_xdim,
                                  size_t ccd_ydim, real_t ccd_xpix, real_t ccd_ypix, real_t* image) {
    const double pix_area = ccd_xpix * ccd_ypix * 1.0E-2;
    try {
        for (size_t i = 0; i < Nrays; ++i) {
            if (ccd_X[i] >= ccd_xdim || ccd_Y[i] >= ccd_ydim) continue;
            double q = QE[lambda[i]];
            double s = incident_spec[lambda[i]] * lambda[i];
            image[(size_t)ccd_Y[i] * ccd_xdim + ccd_X[i]] += s / kHC * q * pix_area;
        }
    } catch (...) {
        return ENGINE_ERROR_BAD_VALUE;
    }
    return ENGINE_ERROR_OK;
}
