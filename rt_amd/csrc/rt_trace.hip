// rt_trace.hip -- fused per-scheme trace kernel and the C ABI of include/rt_hip.h.
//
// One launch replaces the reference's per-surface engine calls of
// Scheme::run (common-lib/scheme.cpp:983-1046): every lane owns one ray in
// registers from generation to the CCD, the packed surface table is read
// through uniform (scalar) loads and the spline pool is staged in LDS.  No
// host round trip, no Rearrange: a lost ray is masked, not compacted.
//
// Build decisions (differences from cuda_engine, each restated in the CPU
// oracle so parity is exact up to rounding):
//   D2  lost-ray counts are exact (wave ballots), not racy ++
//   D3  spline second derivatives by the serial recurrence (host)
//   D4  toric intersection by Newton on the documented F (surface.cpp:56-61)
//   D5  toric grating groove vector uses dFdZ*dFdZ (surface.cu:1438,1487)
//   D6  c == 0 aspheres go through the conic Newton; counters reset per ray
//   D7  one numerically stable quadratic root for every conic (incl. k = -1)
//   D8  the conic root is chosen by sign(c) as in surface.cu:240-244
//   D10 CCD accumulation uses atomics with a bounds check
//   D12 wavelength grids use the global ray index
//   D14 Philox4x32-10 keyed by a 64-bit seed, counter = ray index
//   D15 rectangular random beams are generated
//   D18 toric asphere derivative keeps its factor Y
//   D19 a one-sample uniform wavelength grid is range[0]
//   D20 a lost ray is frozen at the state it was lost in
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "rt_device.h"

using rt::Ray;

// --------------------------------------------------------------- errors
static thread_local std::string g_last_error;

static int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr, code)                                                                 \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return set_error(code, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                     \
    } while (0)

// ----------------------------------------------------------------- kernel

constexpr int kBlock = 256;               // 4 waves of 64
constexpr int kMaxPoolLds = 12288;        // doubles of spline pool staged in LDS (96 KiB)
constexpr int kLdsStats = RT_STATS_LEN;   // per-block counters

struct BeamArg {
    double v[RT_BEAM_STRIDE];
    rt::BeamUniform u;
};

// BEAM < 0: rays are loaded; otherwise the beam is generated with direction type BEAM
constexpr int kLoad = -1;

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// per-lane count added into an LDS counter (no cross-lane reads, so lanes
// that already left the grid-stride loop do not matter)
__device__ __forceinline__ void wave_add(unsigned* counter, unsigned v) {
    if (v) atomicAdd(counter, v);
}

template <int BEAM>
__global__ __launch_bounds__(kBlock) void rt_trace_kernel(
    const double* __restrict__ table, int n_surf, const double* __restrict__ pool_g, int pool_len,
    BeamArg beam, uint64_t seed, uint64_t ray_begin, uint64_t n,
    double* __restrict__ X, double* __restrict__ Y, double* __restrict__ Z,
    double* __restrict__ cX, double* __restrict__ cY, double* __restrict__ cZ,
    double* __restrict__ L, signed char* __restrict__ F, int flags,
    unsigned* __restrict__ hist, double* __restrict__ flux,
    unsigned long long* __restrict__ stats,
    double* __restrict__ dump, signed char* __restrict__ dump_flag) {
    extern __shared__ double s_pool[];
    __shared__ unsigned s_stats[kLdsStats];

    for (int i = threadIdx.x; i < kLdsStats; i += blockDim.x) s_stats[i] = 0u;
    for (int i = threadIdx.x; i < pool_len; i += blockDim.x) s_pool[i] = pool_g[i];
    __syncthreads();

    const bool store = (flags & RT_TRACE_STORE) != 0;
    const bool do_hist = (flags & RT_TRACE_HIST) && hist;
    const bool do_flux = (flags & RT_TRACE_FLUX) && flux;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;

    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        Ray r;
        bool alive;
        if (BEAM >= 0) {
            rt::gen_ray<BEAM>(beam.v, beam.u, seed, ray_begin + i, r);
            alive = true;
        } else {
            r.X = X[i]; r.Y = Y[i]; r.Z = Z[i];
            r.cX = cX[i]; r.cY = cY[i]; r.cZ = cZ[i];
            r.lam = L[i];
            alive = F[i] != 0;
        }
        unsigned n_oob = 0, n_hit = 0, n_off = 0, n_edge = 0;

        for (int s = 0; s < n_surf; ++s) {
            const double* p = table + (size_t)s * rt::DEV_STRIDE;
            const int type = (int)p[RT_S_TYPE];
            if (alive) {
                rt::transform_ray(p, r);                                 // scheme.cpp:990
                rt::normalize_ray(r.cX, r.cY, r.cZ);                     // :993
            }
            if (type != RT_AUXTYPE) {
                unsigned long long entering = __ballot(alive);
                if (lane_id() == 0 && entering) atomicAdd(&s_stats[RT_ST_SURF + 4 * s], (unsigned)__popcll(entering));
                unsigned it = 0;
                bool lost_int = false, lost_act = false;
                if (alive) {
                    alive = rt::intersect_ray(p, r, it);                 // :1002
                    lost_int = !alive;
                }
                if (alive) {
                    alive = rt::mask_ray(p, r.X, r.Y);                   // :1014
                    if (alive) {                                         // :1035
                        if (type == RT_MIRROR) {
                            rt::reflect_ray(p, r);
                        } else if (type == RT_LENS || type == RT_GRATING) {
                            int e1 = 0, e2 = 0;
                            double n1 = rt::fn_eval(s_pool, p[RT_S_N1KIND], p[RT_S_N1VAL], r.lam, e1);
                            double n2 = rt::fn_eval(s_pool, p[RT_S_N2KIND], p[RT_S_N2VAL], r.lam, e2);
                            if (e1 | e2) {
                                alive = false;
                                ++n_oob;
                            } else {
                                // mu = n1/n2 (surface.cu:1095-1096), L = m lambda / n2 / d (:1568)
                                const double in2 = rt::rcp(n2);
                                double mu = n1 * in2;
                                if (type == RT_LENS) {
                                    alive = rt::refract_ray(p, r, mu);
                                } else {
                                    double m = rt::grating_order(p, r.lam);
                                    double Lg = m * r.lam * in2 * p[rt::D_INV_D];
                                    alive = rt::diffract_ray(p, r, mu, Lg);
                                }
                            }
                        }
                        if (alive) rt::normalize_ray(r.cX, r.cY, r.cZ);  // :1040
                    }
                    lost_act = !alive;
                }
                if (__ballot(it != 0)) wave_add(&s_stats[RT_ST_SURF + 4 * s + 1], it);
                unsigned long long li = __ballot(lost_int), la = __ballot(lost_act);
                if (lane_id() == 0) {
                    if (li) atomicAdd(&s_stats[RT_ST_SURF + 4 * s + 2], (unsigned)__popcll(li));
                    if (la) atomicAdd(&s_stats[RT_ST_SURF + 4 * s + 3], (unsigned)__popcll(la));
                }
                if (type == RT_DETECTOR && p[RT_S_CCD_XDIM] > 0.0 && (do_hist || do_flux) && alive) {
                    // ccd_coordinates_kernel (rt_cuda_engine_ccd.cu:30-45) with a bounds check (D10)
                    const uint32_t xdim = (uint32_t)p[RT_S_CCD_XDIM], ydim = (uint32_t)p[RT_S_CCD_YDIM];
                    double vx = (r.X + p[rt::D_HALF_X]) / p[RT_S_CCD_XPIX];
                    double vy = (r.Y + p[rt::D_HALF_Y]) / p[RT_S_CCD_YPIX];
                    uint32_t px = rt::trunc_u32(vx), py = rt::trunc_u32(vy);
                    if (px < xdim && py < ydim) {
                        size_t pix = (size_t)py * xdim + px;
                        ++n_hit;
                        if (rt::near_edge(vx) || rt::near_edge(vy)) ++n_edge;
                        if (do_hist) atomicAdd(&hist[pix], 1u);
                        if (do_flux) {
                            // compute_ccd_image_kernel (ccd.cu:219-232), weights of ccd.cu:286-287,360
                            int eq = 0, es = 0;
                            double qe = rt::fn_eval(s_pool, p[RT_S_QEKIND], p[RT_S_QEVAL], r.lam, eq);
                            double sp = rt::fn_eval(s_pool, p[RT_S_SPKIND], p[RT_S_SPVAL], r.lam, es);
                            if (eq | es) {
                                ++n_oob;
                            } else {
                                double spec = sp * r.lam;
                                double pix_area = p[RT_S_CCD_XPIX] * p[RT_S_CCD_YPIX] * 1.0E-2;
                                double w = spec * (1.0 / rt::kHC) * qe * pix_area;
                                unsafeAtomicAdd(&flux[pix], w);
                            }
                        }
                    } else {
                        ++n_off;
                    }
                }
            }
            if (dump) {
                double* d = dump + (size_t)s * 7 * n;
                d[0 * n + i] = r.X; d[1 * n + i] = r.Y; d[2 * n + i] = r.Z;
                d[3 * n + i] = r.cX; d[4 * n + i] = r.cY; d[5 * n + i] = r.cZ;
                d[6 * n + i] = r.lam;
                dump_flag[(size_t)s * n + i] = (signed char)alive;
            }
        }

        if (store || BEAM >= 0) {
            if (X) {
                X[i] = r.X; Y[i] = r.Y; Z[i] = r.Z;
                cX[i] = r.cX; cY[i] = r.cY; cZ[i] = r.cZ;
                L[i] = r.lam;
            }
            if (F) F[i] = (signed char)alive;
        }
        unsigned long long good = __ballot(alive), all = __ballot(true);
        if (lane_id() == 0) {
            atomicAdd(&s_stats[RT_ST_RAYS], (unsigned)__popcll(all));
            if (good) atomicAdd(&s_stats[RT_ST_GOOD], (unsigned)__popcll(good));
        }
        if (__ballot(n_oob | n_hit | n_off | n_edge)) {
            wave_add(&s_stats[RT_ST_SPLINE_OOB], n_oob);
            wave_add(&s_stats[RT_ST_CCD_HITS], n_hit);
            wave_add(&s_stats[RT_ST_CCD_OFFCHIP], n_off);
            wave_add(&s_stats[RT_ST_CCD_EDGE], n_edge);
        }
    }

    __syncthreads();
    if (stats) {
        for (int k = threadIdx.x; k < kLdsStats; k += blockDim.x)
            if (s_stats[k]) atomicAdd(&stats[k], (unsigned long long)s_stats[k]);
    }
}

// ------------------------------------------------------------------- C ABI

struct rt_hip_scheme {
    int device;
    int n_surf;
    size_t pool_len;                 // device pool (doubles)
    double* table_dev;               // n_surf records of rt::DEV_STRIDE
    double* pool_dev;
    std::vector<double> table;       // public packed table, as given
    std::vector<double> pool;        // public spline pool, as given
    std::vector<double> dtable;      // device records
    std::vector<double> dpool;       // device spline records
};

namespace {
struct DeviceInfo {
    bool valid = false;
    int device = 0;
    int cus = 0;
    hipDeviceProp_t prop;
};
thread_local DeviceInfo g_dev;

int ensure_device() {
    if (g_dev.valid) return RT_OK;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), RT_ERR_INIT_FAILED);
    HIP_TRY(hipGetDeviceProperties(&g_dev.prop, dev), RT_ERR_INIT_FAILED);
    g_dev.device = dev;
    g_dev.cus = g_dev.prop.multiProcessorCount;
    g_dev.valid = true;
    return RT_OK;
}

bool fn_ref_ok(double kind, double val, size_t pool_len, const double* pool) {
    if (kind == (double)RT_FN_CONST) return true;
    if (kind != (double)RT_FN_SPLINE) return false;
    if (!(val >= 0) || val != (double)(size_t)val) return false;
    size_t o = (size_t)val;
    if (o >= pool_len) return false;
    double k = pool[o];
    if (!(k >= 2) || k > RT_MAX_SPLINE_KNOTS || k != (double)(size_t)k) return false;
    return o + 1 + 3 * (size_t)k <= pool_len;
}
}  // namespace

extern "C" int rt_hip_init(int device) {
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count), RT_ERR_INIT_FAILED);
    if (device < 0 || device >= count)
        return set_error(RT_ERR_INIT_FAILED, "device %d out of range (%d devices)", device, count);
    HIP_TRY(hipSetDevice(device), RT_ERR_INIT_FAILED);
    g_dev.valid = false;
    return ensure_device();
}

extern "C" int rt_hip_info(char* buf, size_t len) {
    int rc = ensure_device();
    if (rc != RT_OK) return rc;
    const hipDeviceProp_t& p = g_dev.prop;
    int n = snprintf(buf, len,
                     "AMD HIP engine (gfx950 FP64)\nDevice name: %s\nArchitecture: %s\n"
                     "Global memory: %zu\nNumber of compute units: %d\nClock (kHz): %d\n",
                     p.name, p.gcnArchName, (size_t)p.totalGlobalMem, p.multiProcessorCount, p.clockRate);
    if (n < 0 || (size_t)n >= len) return set_error(RT_ERR_BAD_ALLOC, "info buffer too small");
    return RT_OK;
}

extern "C" int rt_hip_last_error(char* buf, size_t len) {
    if (!buf || !len) return RT_ERR_BAD_VALUE;
    snprintf(buf, len, "%s", g_last_error.c_str());
    return RT_OK;
}

extern "C" int rt_hip_validate_scheme(const double* t, int n_surf, const double* pool, size_t pool_len) {
    if (!t || n_surf < 1 || n_surf > RT_MAX_SURF)
        return set_error(RT_ERR_BAD_VALUE, "n_surf %d out of range [1,%d]", n_surf, RT_MAX_SURF);
    if (pool_len && !pool) return set_error(RT_ERR_BAD_VALUE, "pool is NULL");
    for (int s = 0; s < n_surf; ++s) {
        const double* p = t + (size_t)s * RT_SURF_STRIDE;
        int cls = (int)p[RT_S_CLASS], type = (int)p[RT_S_TYPE], shape = (int)p[RT_S_SHAPE];
        if (cls < RT_PLANE || cls > RT_AUXCLASS || (double)cls != p[RT_S_CLASS])
            return set_error(RT_ERR_BAD_VALUE, "surface %d: bad class", s);
        if (type < RT_MIRROR || type > RT_AUXTYPE || (double)type != p[RT_S_TYPE])
            return set_error(RT_ERR_BAD_VALUE, "surface %d: bad type", s);
        if (shape != RT_CIRCLE && shape != RT_RECTANGLE)
            return set_error(RT_ERR_BAD_VALUE, "surface %d: bad shape", s);
        double np = p[RT_S_NPOLY];
        if (!(np >= 0 && np <= RT_MAX_POLY) || np != (double)(int)np)
            return set_error(RT_ERR_BAD_VALUE, "surface %d: bad number of aspheric terms", s);
        if (type == RT_GRATING) {
            int gt = (int)p[RT_S_GRTYPE];
            if (gt != RT_REFLECTIVE && gt != RT_TRANSPARENT)   // surface.cu:1637 BAD_VALUE
                return set_error(RT_ERR_BAD_VALUE, "surface %d: bad grating type", s);
            if (!(p[RT_S_GRCONST] != 0.0))
                return set_error(RT_ERR_BAD_VALUE, "surface %d: zero grating constant", s);
            int om = (int)p[RT_S_ORDMODE];
            if (om != RT_ORDER_FIXED && om != RT_ORDER_BLAZE)
                return set_error(RT_ERR_BAD_VALUE, "surface %d: bad order mode", s);
        }
        if (type == RT_LENS || type == RT_GRATING) {
            if (!fn_ref_ok(p[RT_S_N1KIND], p[RT_S_N1VAL], pool_len, pool) ||
                !fn_ref_ok(p[RT_S_N2KIND], p[RT_S_N2VAL], pool_len, pool))
                return set_error(RT_ERR_BAD_VALUE, "surface %d: bad refractive index reference", s);
        }
        if (type == RT_DETECTOR && p[RT_S_CCD_XDIM] > 0.0) {
            double xd = p[RT_S_CCD_XDIM], yd = p[RT_S_CCD_YDIM];
            if (!(xd >= 1 && yd >= 1 && xd * yd <= 4294967296.0) || xd != (double)(uint64_t)xd ||
                yd != (double)(uint64_t)yd)
                return set_error(RT_ERR_BAD_VALUE, "surface %d: bad CCD dimensions", s);
            if (!(p[RT_S_CCD_XPIX] > 0.0) || !(p[RT_S_CCD_YPIX] > 0.0))
                return set_error(RT_ERR_BAD_VALUE, "surface %d: bad CCD pixel size", s);
            if (!fn_ref_ok(p[RT_S_QEKIND], p[RT_S_QEVAL], pool_len, pool) ||
                !fn_ref_ok(p[RT_S_SPKIND], p[RT_S_SPVAL], pool_len, pool))
                return set_error(RT_ERR_BAD_VALUE, "surface %d: bad QE/spectrum reference", s);
        }
    }
    if (pool_len > (size_t)kMaxPoolLds - 64)
        return set_error(RT_ERR_BAD_VALUE, "spline pool of %zu doubles exceeds %d", pool_len, kMaxPoolLds - 64);
    return RT_OK;
}

// Device tables: each record gets its host-derived constants (rt::DevField)
// and each referenced spline is re-laid out as [K, x0, (K-1)/(x[K-1]-x0),
// x, y, y2] with the references patched to the new offsets.
static void build_device_tables(rt_hip_scheme* s) {
    const int n = s->n_surf;
    s->dtable.assign((size_t)n * rt::DEV_STRIDE, 0.0);
    s->dpool.clear();
    std::vector<std::pair<size_t, size_t>> moved;   // (public offset, device offset)
    auto remap = [&](double& kind, double& val) {
        if (kind != (double)RT_FN_SPLINE) return;
        size_t o = (size_t)val;
        for (auto& m : moved)
            if (m.first == o) { val = (double)m.second; return; }
        const double* sp = s->pool.data() + o;
        size_t K = (size_t)sp[0];
        size_t d = s->dpool.size();
        s->dpool.push_back((double)K);
        s->dpool.push_back(sp[1]);
        s->dpool.push_back((double)(K - 1) / (sp[K] - sp[1]));
        s->dpool.insert(s->dpool.end(), sp + 1, sp + 1 + 3 * K);
        moved.push_back({o, d});
        val = (double)d;
    };
    for (int i = 0; i < n; ++i) {
        double* d = s->dtable.data() + (size_t)i * rt::DEV_STRIDE;
        const double* p = s->table.data() + (size_t)i * RT_SURF_STRIDE;
        for (int k = 0; k < RT_SURF_STRIDE; ++k) d[k] = p[k];
        int type = (int)p[RT_S_TYPE];
        if (type == RT_LENS || type == RT_GRATING) {
            remap(d[RT_S_N1KIND], d[RT_S_N1VAL]);
            remap(d[RT_S_N2KIND], d[RT_S_N2VAL]);
        }
        if (type == RT_DETECTOR && p[RT_S_CCD_XDIM] > 0.0) {
            remap(d[RT_S_QEKIND], d[RT_S_QEVAL]);
            remap(d[RT_S_SPKIND], d[RT_S_SPVAL]);
        }
        d[rt::D_INV_C] = p[RT_S_C] != 0.0 ? 1.0 / p[RT_S_C] : 0.0;
        d[rt::D_INV_D] = p[RT_S_GRCONST] != 0.0 ? 1.0 / p[RT_S_GRCONST] : 0.0;
        d[rt::D_HALF_X] = p[RT_S_CCD_XSIZE] / 2.0;
        d[rt::D_HALF_Y] = p[RT_S_CCD_YSIZE] / 2.0;
    }
    s->pool_len = s->dpool.size();
}

extern "C" int rt_hip_scheme_create(const double* surf_table, int n_surf, const double* spline_pool,
                                    size_t pool_len, rt_hip_scheme** out) {
    if (!out) return set_error(RT_ERR_BAD_VALUE, "out is NULL");
    *out = nullptr;
    int rc = rt_hip_validate_scheme(surf_table, n_surf, spline_pool, pool_len);
    if (rc != RT_OK) return rc;
    rc = ensure_device();
    if (rc != RT_OK) return rc;
    rt_hip_scheme* s = new (std::nothrow) rt_hip_scheme();
    if (!s) return set_error(RT_ERR_BAD_ALLOC, "out of host memory");
    s->device = g_dev.device;
    s->n_surf = n_surf;
    s->table.assign(surf_table, surf_table + (size_t)n_surf * RT_SURF_STRIDE);
    if (pool_len) s->pool.assign(spline_pool, spline_pool + pool_len);
    build_device_tables(s);
    s->table_dev = nullptr;
    s->pool_dev = nullptr;
    size_t tb = s->dtable.size() * sizeof(double);
    size_t pb = (s->dpool.empty() ? 1 : s->dpool.size()) * sizeof(double);
    if (hipMalloc(&s->table_dev, tb) != hipSuccess || hipMalloc(&s->pool_dev, pb) != hipSuccess) {
        (void)hipFree(s->table_dev);
        (void)hipFree(s->pool_dev);
        delete s;
        return set_error(RT_ERR_BAD_ALLOC, "hipMalloc of scheme tables failed");
    }
    if (hipMemcpy(s->table_dev, s->dtable.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
        (!s->dpool.empty() && hipMemcpy(s->pool_dev, s->dpool.data(), s->dpool.size() * sizeof(double),
                                        hipMemcpyHostToDevice) != hipSuccess)) {
        (void)hipFree(s->table_dev);
        (void)hipFree(s->pool_dev);
        delete s;
        return set_error(RT_ERR_FAILED, "upload of scheme tables failed");
    }
    *out = s;
    return RT_OK;
}

extern "C" int rt_hip_scheme_destroy(rt_hip_scheme* s) {
    if (!s) return RT_OK;
    (void)hipFree(s->table_dev);
    (void)hipFree(s->pool_dev);
    delete s;
    return RT_OK;
}

static int validate_beam(const double* b) {
    if (!b) return set_error(RT_ERR_BAD_VALUE, "beam description is NULL");
    int type = (int)b[RT_B_TYPE], shape = (int)b[RT_B_SHAPE], distr = (int)b[RT_B_RANGE_DISTR];
    if (type < RT_BEAM_PARALLEL || type > RT_BEAM_GAUSS) return set_error(RT_ERR_BAD_VALUE, "bad beam type");
    if (shape != RT_CIRCLE && shape != RT_RECTANGLE) return set_error(RT_ERR_BAD_VALUE, "bad beam shape");
    if (distr < RT_RANGE_RANDOM || distr > RT_RANGE_MONO) return set_error(RT_ERR_BAD_VALUE, "bad range distribution");
    // Beam::SetParams checks (beam.cpp:78-91)
    if (!(b[RT_B_NPER] > 0)) return set_error(RT_ERR_BAD_VALUE, "number of rays must be > 0");
    if (b[RT_B_P1] < 0 || !(b[RT_B_P2] > 0)) return set_error(RT_ERR_BAD_VALUE, "beam sizes must be > 0");
    if (shape == RT_CIRCLE && b[RT_B_P1] >= b[RT_B_P2])
        return set_error(RT_ERR_BAD_VALUE, "inner radius must be below the outer one");
    if (shape == RT_RECTANGLE && !(b[RT_B_P3] > 0)) return set_error(RT_ERR_BAD_VALUE, "size ratio must be > 0");
    if (b[RT_B_COMPOSITE] != 0.0 && !(b[RT_B_NLAMBDA] >= 1))
        return set_error(RT_ERR_BAD_VALUE, "composite beam needs nlambda >= 1");
    return RT_OK;
}

extern "C" int rt_hip_trace(const rt_hip_scheme* sch, const double* beam_desc, uint64_t seed,
                            size_t ray_begin, size_t n_rays, double* X, double* Y, double* Z,
                            double* cX, double* cY, double* cZ, double* lambda, signed char* flag,
                            int flags, unsigned int* ccd_hist, double* ccd_flux,
                            unsigned long long* stats, double* dump, signed char* dump_flag,
                            void* stream) {
    if (!sch) return set_error(RT_ERR_BAD_VALUE, "scheme is NULL");
    const bool gen = (flags & RT_TRACE_GENERATE) != 0;
    if (gen) {
        int rc = validate_beam(beam_desc);
        if (rc != RT_OK) return rc;
    } else if (!X || !Y || !Z || !cX || !cY || !cZ || !lambda || !flag) {
        return set_error(RT_ERR_BAD_VALUE, "input rays missing");
    }
    if ((flags & RT_TRACE_STORE) && (!X || !Y || !Z || !cX || !cY || !cZ || !lambda || !flag))
        return set_error(RT_ERR_BAD_VALUE, "RT_TRACE_STORE needs all ray buffers");
    if (X && (!Y || !Z || !cX || !cY || !cZ || !lambda))
        return set_error(RT_ERR_BAD_VALUE, "ray buffers must be given together");
    if ((dump == nullptr) != (dump_flag == nullptr))
        return set_error(RT_ERR_BAD_VALUE, "dump and dump_flag must be given together");
    if (n_rays == 0) return RT_OK;
    int rc = ensure_device();
    if (rc != RT_OK) return rc;

    BeamArg b;
    memset(&b, 0, sizeof b);
    if (beam_desc) memcpy(b.v, beam_desc, sizeof b.v);
    const int cus = g_dev.cus > 0 ? g_dev.cus : 256;
    uint64_t want = (n_rays + kBlock - 1) / kBlock;
    uint64_t cap = (uint64_t)cus * 8;      // persistent grid: 8 blocks (32 waves) per CU
    unsigned grid = (unsigned)(want < cap ? want : cap);
    size_t lds = sch->pool_len * sizeof(double);
    hipStream_t st = (hipStream_t)stream;
    if (gen) {
        // uniform values of the direction generators, on the host (beam.cu:592-598, 1034, 1099)
        double ang = b.v[RT_B_ANGPAR] * rt::kDeg2Rad;
        b.u.cone_cos = cos(ang / 2.0);
        b.u.cone_sin = sin(ang / 2.0);
        double stddev = b.v[RT_B_ANGPAR] / 2.35482 * rt::kPi / 180 / 3600;
        b.u.gauss_tg = tan(stddev);
    }
    auto kern = rt_trace_kernel<kLoad>;
    if (gen) {
        int type = (int)b.v[RT_B_TYPE];
        kern = type == RT_BEAM_PARALLEL ? rt_trace_kernel<RT_BEAM_PARALLEL>
             : type == RT_BEAM_CONIC    ? rt_trace_kernel<RT_BEAM_CONIC>
                                        : rt_trace_kernel<RT_BEAM_GAUSS>;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, st, sch->table_dev, sch->n_surf, sch->pool_dev,
                       (int)sch->pool_len, b, seed, (uint64_t)ray_begin, (uint64_t)n_rays, X, Y, Z, cX, cY,
                       cZ, lambda, flag, flags, ccd_hist, ccd_flux, stats, dump, dump_flag);
    HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    return RT_OK;
}
