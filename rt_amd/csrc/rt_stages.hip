// rt_stages.hip -- per-stage kernels behind the rt_engine_func.h plugin API.
//
// The reference engine runs one kernel per engine call (beam_random,
// beam_transform, surface_intersection, ... ; rt_cuda_engine_*.cu).  These
// are the same stages over device-resident SoA arrays, built from the
// per-ray stage functions of rt_device.h that the fused trace uses, so the
// per-stage API and the fused path compute identical arithmetic.  Lost rays
// are counted exactly (wave ballots, one atomic per wave) instead of the
// reference's racy ++dev_N_bad.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <vector>

#include "rt_device.h"
#include "rt_host.h"

using rt::cptr;
using rt::Ray;

namespace {

constexpr int kThreads = 256;

enum StageOp { OP_TRANSFORM, OP_NORMALIZE, OP_INTERSECT, OP_MASK, OP_REFLECT, OP_REFRACT, OP_DIFFRACT };

struct StageArgs {
    double rec[rt::DEV_STRIDE];        // device record of the surface (first: kernarg offset 0)
    uint64_t n;
    int op, pad;
    double *X, *Y, *Z, *cX, *cY, *cZ;
    const double* mu;                  // per-ray n1/n2 (refraction, transmissive gratings)
    const double* L;                   // per-ray m lambda / n2 / d (gratings)
    signed char* flag;
    unsigned long long* n_bad;
    unsigned long long* iters;
};

__device__ __forceinline__ cptr kernarg_record() {
    return (cptr)__builtin_amdgcn_kernarg_segment_ptr();   // StageArgs::rec is at offset 0
}

// Adds v summed over the workgroup into *dst with one atomic (none for a
// zero sum).  Every thread of the workgroup calls it, after its ray loop.
__device__ void block_add(unsigned long long* dst, unsigned long long v) {
    __shared__ unsigned long long part[kThreads / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (__lane_id() == 0) part[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (unsigned w = 0; w < blockDim.x / 64; ++w) t += part[w];
        if (t) atomicAdd(dst, t);
    }
    __syncthreads();
}

__global__ __launch_bounds__(kThreads) void rt_stage_kernel(StageArgs a) {
    const rt::DynSurf d{kernarg_record()};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long it_sum = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        Ray r;
        r.X = a.X ? a.X[i] : 0.0; r.Y = a.Y ? a.Y[i] : 0.0; r.Z = a.Z ? a.Z[i] : 0.0;
        r.cX = a.cX[i]; r.cY = a.cY[i]; r.cZ = a.cZ[i];
        r.lam = 0.0;
        bool lost = false;
        unsigned it = 0;
        switch (a.op) {
            case OP_TRANSFORM:                 // beam_transform, beam.cu:416-579
                rt::transform_ray(d, r);
                a.X[i] = r.X; a.Y[i] = r.Y; a.Z[i] = r.Z;
                a.cX[i] = r.cX; a.cY[i] = r.cY; a.cZ[i] = r.cZ;
                break;
            case OP_NORMALIZE:                 // Beam::Normalize, beam.cpp:527-571
                rt::normalize_ray(r.cX, r.cY, r.cZ);
                a.cX[i] = r.cX; a.cY[i] = r.cY; a.cZ[i] = r.cZ;
                break;
            case OP_INTERSECT:                 // surface_intersection, surface.cu:655-760
                lost = !rt::intersect_ray(d, r, it);
                a.X[i] = r.X; a.Y[i] = r.Y; a.Z[i] = r.Z;
                break;
            case OP_MASK:                      // Surface/Shade::ApplyConstrains, surface.cpp:264-321,652-699
                lost = !rt::mask_ray(d, r.X, r.Y);
                break;
            case OP_REFLECT:                   // surface_reflection, surface.cu:883-951
                rt::reflect_ray(d, r);
                a.cX[i] = r.cX; a.cY[i] = r.cY; a.cZ[i] = r.cZ;
                break;
            case OP_REFRACT:                   // surface_refraction, surface.cu:1155-1242
                lost = !rt::refract_ray(d, r, a.mu[i]);
                if (!lost) { a.cX[i] = r.cX; a.cY[i] = r.cY; a.cZ[i] = r.cZ; }
                break;
            case OP_DIFFRACT:                  // surface_diffration, surface.cu:1655-1768
                lost = !rt::diffract_ray(d, r, a.mu ? a.mu[i] : 1.0, a.L[i]);
                if (!lost) { a.cX[i] = r.cX; a.cY[i] = r.cY; a.cZ[i] = r.cZ; }
                break;
        }
        if (lost) a.flag[i] = 0;
        unsigned long long b = __ballot(lost);
        if (__lane_id() == 0 && b && a.n_bad) atomicAdd(a.n_bad, (unsigned long long)__popcll(b));
        it_sum += it;
    }
    if (a.iters) block_add(a.iters, it_sum);
}

enum BeamOp { GEN_POSITIONS, GEN_DIRECTIONS, GEN_WAVELENGTHS };

struct BeamGenArgs {
    double beam[rt::kBeamRec];         // beam fields + BeamUniform (first: kernarg offset 0)
    uint64_t seed, begin, n;
    int op, dir_type;
    double *X, *Y, *Z, *cX, *cY, *cZ, *L;
};

__global__ __launch_bounds__(kThreads) void rt_beam_kernel(BeamGenArgs a) {
    cptr B = (cptr)__builtin_amdgcn_kernarg_segment_ptr();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        Ray r;
        const uint64_t g = a.begin + i;
        if (a.op == GEN_POSITIONS) {           // beam_random, beam.cu:224-312
            rt::gen_position(B, a.seed, g, r);
            a.X[i] = r.X; a.Y[i] = r.Y; a.Z[i] = r.Z;
        } else if (a.op == GEN_DIRECTIONS) {   // beam_conic_cosins / beam_gauss_cosins
            r.X = a.X ? a.X[i] : 0.0;
            r.Y = a.Y ? a.Y[i] : 0.0;
            if (a.dir_type == RT_BEAM_CONIC) rt::gen_direction<RT_BEAM_CONIC>(B, a.seed, g, r);
            else if (a.dir_type == RT_BEAM_GAUSS) rt::gen_direction<RT_BEAM_GAUSS>(B, a.seed, g, r);
            else rt::gen_direction<RT_BEAM_PARALLEL>(B, a.seed, g, r);
            a.cX[i] = r.cX; a.cY[i] = r.cY; a.cZ[i] = r.cZ;
        } else {                               // uniform_range_beam / random_range_beam
            uint64_t j;
            a.L[i] = rt::gen_lambda(B, a.seed, g, j);
        }
    }
}

struct CcdArgs {
    uint64_t n;
    double half_x, half_y, xpix, ypix;
    uint32_t xdim, ydim;
    const double *X, *Y;
    unsigned *ccdX, *ccdY;
    const double* w;
    double* image;
    unsigned* hist;
    unsigned long long* offchip;
    int op;
};

// ccd_coordinates_kernel (rt_cuda_engine_ccd.cu:30-45) and
// compute_ccd_image_kernel (:219-232) with atomics and a bounds check (D10)
__global__ __launch_bounds__(kThreads) void rt_ccd_kernel(CcdArgs a) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        if (a.op == 0) {
            a.ccdX[i] = rt::trunc_u32((a.X[i] + a.half_x) / a.xpix);
            a.ccdY[i] = rt::trunc_u32((a.Y[i] + a.half_y) / a.ypix);
        } else {
            uint32_t px = a.ccdX[i], py = a.ccdY[i];
            bool off = !(px < a.xdim && py < a.ydim);
            uint32_t pix = off ? 0u : py * a.xdim + px;     // xdim * ydim <= 2^32 (checked on the host)
            if (!off && a.image) unsafeAtomicAdd(&a.image[pix], a.w ? a.w[i] : 1.0);
            if (a.hist) rt::hist_add_wave(a.hist, pix, !off);
            unsigned long long b = __ballot(off);
            if (__lane_id() == 0 && b && a.offchip) atomicAdd(a.offchip, (unsigned long long)__popcll(b));
        }
    }
}

struct ScaleArgs {
    uint64_t n;
    double* v;
    const double* f;
};

__global__ __launch_bounds__(kThreads) void rt_scale_kernel(ScaleArgs a) {   // surface_QE_kernel
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) a.v[i] *= a.f[i];
}

struct GratingArgs {
    uint64_t n;
    double blaze, alpha, gamma, d, rule_size;
    long order;
    const double* lambda;
    double* energy;
};

// I1_func_Tarasov (surface.cu:1988-2019)
__device__ double I1_tarasov(double lam, double beta, double c, double k1, double eps, double rule_size) {
    double I1 = rt::kPi * eps * rule_size / lam * cos(beta) / cos(c - beta) * (sin(c - k1) + sin(c - beta));
    double term = cos(beta) * cos(c - k1) / cos(c - beta);
    if (fabs(I1) < rt::kPi / 2.0) {
        I1 *= I1;
        I1 = (0.00761 * I1 - 0.16605) * I1 + 1;
    } else {
        I1 = sin(I1) / (I1);
    }
    I1 *= term;
    I1 *= I1;
    return I1;
}

// grating_energy_distr_kernel (surface.cu:2098-2153)
__global__ __launch_bounds__(kThreads) void rt_grating_energy_kernel(GratingArgs a) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const double c = a.blaze, k1 = a.alpha, cR = a.gamma, eps = a.d;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        double lam = a.lambda[i];
        double sin_alpha = sin(k1), cos_gamma = cos(cR);
        long long omin = (long long)(eps / lam * (sin_alpha - 1) * cos_gamma);
        long long omax = (long long)(eps / lam * (sin_alpha + 1) * cos_gamma);
        double norm = 0.0;
        for (long long m = omin; m <= omax; ++m)
            norm += I1_tarasov(lam, asin((double)m * lam / eps / cos_gamma - sin_alpha), c, k1, eps, a.rule_size);
        double beta = asin((double)a.order * lam / (eps * cos_gamma) - sin_alpha);
        a.energy[i] *= I1_tarasov(lam, beta, c, k1, eps, a.rule_size) / norm;
    }
}

struct SplineArgs {
    uint64_t n;
    int pool_len, pad;
    const double* spline;              // device spline record [K, x0, inv, x, y, y2]
    const double* x;
    double* y;
    unsigned long long* oob;
};

// TabulatedFunction::operator[] over an array, record staged in LDS
__global__ __launch_bounds__(kThreads) void rt_spline_kernel(SplineArgs a) {
    extern __shared__ double s_sp[];
    for (int i = threadIdx.x; i < a.pool_len; i += blockDim.x) s_sp[i] = a.spline[i];
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        int err = 0;
        double v = rt::spline_eval(s_sp, a.x[i], err);
        a.y[i] = err ? 0.0 : v;
        unsigned long long b = __ballot(err != 0);
        if (__lane_id() == 0 && b && a.oob) atomicAdd(a.oob, (unsigned long long)__popcll(b));
    }
}

unsigned grid_for(size_t n) {
    uint64_t want = (n + kThreads - 1) / kThreads;
    uint64_t cap = (uint64_t)rt_device_cus() * 16;
    return (unsigned)(want < cap ? (want ? want : 1) : cap);
}

int launch_stage(const double* rec, size_t n, int op, double* X, double* Y, double* Z, double* cX, double* cY,
                 double* cZ, const double* mu, const double* L, signed char* flag, unsigned long long* n_bad,
                 unsigned long long* iters, void* stream) {
    if (n == 0) return RT_OK;
    int rc = rt_ensure_device();
    if (rc != RT_OK) return rc;
    StageArgs a;
    memset(&a, 0, sizeof a);
    if (rec) rt_device_record(rec, a.rec);
    a.n = n; a.op = op;
    a.X = X; a.Y = Y; a.Z = Z; a.cX = cX; a.cY = cY; a.cZ = cZ;
    a.mu = mu; a.L = L; a.flag = flag; a.n_bad = n_bad; a.iters = iters;
    hipLaunchKernelGGL(rt_stage_kernel, dim3(grid_for(n)), dim3(kThreads), 0, (hipStream_t)stream, a);
    HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    return RT_OK;
}

bool rays_ok(const void* a, const void* b, const void* c, const void* d, const void* e, const void* f) {
    return a && b && c && d && e && f;
}

}  // namespace

extern "C" int rt_hip_stage_transform(const double* angles, const double* dist, size_t n, double* X, double* Y,
                                      double* Z, double* cX, double* cY, double* cZ, void* stream) {
    if (!angles || !dist || !rays_ok(X, Y, Z, cX, cY, cZ)) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    double rec[RT_SURF_STRIDE];
    memset(rec, 0, sizeof rec);
    for (int j = 0; j < 3; ++j) {
        rec[RT_S_DIST + j] = dist[j];
        rec[RT_S_ANG + j] = angles[j];
        rec[RT_S_SIN + j] = sin(angles[j]);          // host libm, as beam.cu:520-521
        rec[RT_S_COS + j] = cos(angles[j]);
    }
    return launch_stage(rec, n, OP_TRANSFORM, X, Y, Z, cX, cY, cZ, nullptr, nullptr, nullptr, nullptr, nullptr,
                        stream);
}

extern "C" int rt_hip_stage_normalize(size_t n, double* cX, double* cY, double* cZ, void* stream) {
    if (!cX || !cY || !cZ) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    return launch_stage(nullptr, n, OP_NORMALIZE, nullptr, nullptr, nullptr, cX, cY, cZ, nullptr, nullptr,
                        nullptr, nullptr, nullptr, stream);
}

extern "C" int rt_hip_stage_intersect(const double* rec, size_t n, double* X, double* Y, double* Z, double* cX,
                                      double* cY, double* cZ, signed char* flag, unsigned long long* n_bad,
                                      unsigned long long* iters, void* stream) {
    if (!rec || !flag || !rays_ok(X, Y, Z, cX, cY, cZ)) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    // the intersection reads the geometry only: validate it without the record's
    // tabulated-function and CCD fields, which refer to a scheme's spline pool
    double geo[RT_SURF_STRIDE];
    memcpy(geo, rec, sizeof geo);
    geo[RT_S_TYPE] = RT_STOP;
    int rc = rt_hip_validate_scheme(geo, 1, nullptr, 0);
    if (rc != RT_OK) return rc;
    return launch_stage(rec, n, OP_INTERSECT, X, Y, Z, cX, cY, cZ, nullptr, nullptr, flag, n_bad, iters, stream);
}

extern "C" int rt_hip_stage_mask(const double* rec, size_t n, double* X, double* Y, signed char* flag,
                                 unsigned long long* n_bad, void* stream) {
    if (!rec || !X || !Y || !flag) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    // the stage kernel reads direction cosines; give it the positions (unused by the mask)
    return launch_stage(rec, n, OP_MASK, X, Y, nullptr, X, Y, X, nullptr, nullptr, flag, n_bad, nullptr, stream);
}

extern "C" int rt_hip_stage_reflect(const double* rec, size_t n, double* X, double* Y, double* Z, double* cX,
                                    double* cY, double* cZ, void* stream) {
    if (!rec || !rays_ok(X, Y, Z, cX, cY, cZ)) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    return launch_stage(rec, n, OP_REFLECT, X, Y, Z, cX, cY, cZ, nullptr, nullptr, nullptr, nullptr, nullptr,
                        stream);
}

extern "C" int rt_hip_stage_refract(const double* rec, size_t n, double* X, double* Y, double* Z, double* cX,
                                    double* cY, double* cZ, const double* mu, signed char* flag,
                                    unsigned long long* n_bad, void* stream) {
    if (!rec || !mu || !flag || !rays_ok(X, Y, Z, cX, cY, cZ)) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    return launch_stage(rec, n, OP_REFRACT, X, Y, Z, cX, cY, cZ, mu, nullptr, flag, n_bad, nullptr, stream);
}

extern "C" int rt_hip_stage_diffract(const double* rec, size_t n, double* X, double* Y, double* Z, double* cX,
                                     double* cY, double* cZ, const double* mu, const double* L, signed char* flag,
                                     unsigned long long* n_bad, void* stream) {
    if (!rec || !L || !flag || !rays_ok(X, Y, Z, cX, cY, cZ)) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    int gt = (int)rec[RT_S_GRTYPE];
    if (gt != RT_REFLECTIVE && gt != RT_TRANSPARENT)           // surface.cu:1637,1706
        return rt_set_error(RT_ERR_BAD_VALUE, "bad grating type");
    if (gt == RT_TRANSPARENT && !mu) return rt_set_error(RT_ERR_BAD_VALUE, "transmissive grating needs mu");
    return launch_stage(rec, n, OP_DIFFRACT, X, Y, Z, cX, cY, cZ, mu, L, flag, n_bad, nullptr, stream);
}

static int launch_beam(const double* beam_desc, uint64_t seed, size_t begin, size_t n, int op, double* X,
                       double* Y, double* Z, double* cX, double* cY, double* cZ, double* L, void* stream) {
    if (!beam_desc) return rt_set_error(RT_ERR_BAD_VALUE, "beam description is NULL");
    if (n == 0) return RT_OK;
    int rc = rt_ensure_device();
    if (rc != RT_OK) return rc;
    BeamGenArgs a;
    memset(&a, 0, sizeof a);
    memcpy(a.beam, beam_desc, RT_BEAM_STRIDE * sizeof(double));
    // params[3]*deg2rad as the macro expands (beam.cpp:383): (params[3]*RT_PI)/180.0
    double ang = a.beam[RT_B_ANGUNIT] != 0.0 ? a.beam[RT_B_ANGPAR] : a.beam[RT_B_ANGPAR] * rt::kPi / 180.0;
    a.beam[rt::BU_CONE_COS] = cos(ang / 2.0);                         // beam.cu:593-594
    a.beam[rt::BU_CONE_SIN] = sin(ang / 2.0);
    a.beam[rt::BU_GAUSS_TG] = tan(a.beam[RT_B_ANGPAR] / 2.35482 * rt::kPi / 180 / 3600);   // beam.cu:1034,1099
    a.beam[rt::BU_LAM_STEP] = rt::beam_lam_step(a.beam);
    a.seed = seed; a.begin = begin; a.n = n; a.op = op; a.dir_type = (int)beam_desc[RT_B_TYPE];
    a.X = X; a.Y = Y; a.Z = Z; a.cX = cX; a.cY = cY; a.cZ = cZ; a.L = L;
    hipLaunchKernelGGL(rt_beam_kernel, dim3(grid_for(n)), dim3(kThreads), 0, (hipStream_t)stream, a);
    HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    return RT_OK;
}

extern "C" int rt_hip_beam_positions(const double* beam_desc, uint64_t seed, size_t begin, size_t n, double* X,
                                     double* Y, double* Z, void* stream) {
    if (!X || !Y || !Z) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    return launch_beam(beam_desc, seed, begin, n, GEN_POSITIONS, X, Y, Z, nullptr, nullptr, nullptr, nullptr, stream);
}

extern "C" int rt_hip_beam_directions(const double* beam_desc, uint64_t seed, size_t begin, size_t n,
                                      const double* X, const double* Y, double* cX, double* cY, double* cZ,
                                      void* stream) {
    if (!cX || !cY || !cZ) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    if (beam_desc && (int)beam_desc[RT_B_TYPE] == RT_BEAM_CONIC && (!X || !Y))
        return rt_set_error(RT_ERR_BAD_VALUE, "conic directions need positions");
    return launch_beam(beam_desc, seed, begin, n, GEN_DIRECTIONS, (double*)X, (double*)Y, nullptr, cX, cY, cZ,
                       nullptr, stream);
}

extern "C" int rt_hip_beam_wavelengths(const double* beam_desc, uint64_t seed, size_t begin, size_t n,
                                       double* lambda, void* stream) {
    if (!lambda) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    return launch_beam(beam_desc, seed, begin, n, GEN_WAVELENGTHS, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, lambda, stream);
}

extern "C" int rt_hip_ccd_coordinates(size_t n, const double* X, const double* Y, double xsize, double ysize,
                                      double xpix, double ypix, unsigned* ccdX, unsigned* ccdY, void* stream) {
    if (!X || !Y || !ccdX || !ccdY) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    if (n == 0) return RT_OK;
    int rc = rt_ensure_device();
    if (rc != RT_OK) return rc;
    CcdArgs a;
    memset(&a, 0, sizeof a);
    a.n = n; a.half_x = xsize / 2.0; a.half_y = ysize / 2.0; a.xpix = xpix; a.ypix = ypix;
    a.X = X; a.Y = Y; a.ccdX = ccdX; a.ccdY = ccdY; a.op = 0;
    hipLaunchKernelGGL(rt_ccd_kernel, dim3(grid_for(n)), dim3(kThreads), 0, (hipStream_t)stream, a);
    HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    return RT_OK;
}

extern "C" int rt_hip_ccd_accumulate(size_t n, const unsigned* ccdX, const unsigned* ccdY, const double* w,
                                     size_t xdim, size_t ydim, double* image, unsigned* hist,
                                     unsigned long long* offchip, void* stream) {
    if (!ccdX || !ccdY || (!image && !hist)) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    if (xdim == 0 || ydim == 0 || xdim > 0xffffffffu || ydim > 0xffffffffu ||
        (unsigned long long)xdim * ydim > 4294967296ull)
        return rt_set_error(RT_ERR_BAD_VALUE, "bad CCD dimensions");
    if (n == 0) return RT_OK;
    int rc = rt_ensure_device();
    if (rc != RT_OK) return rc;
    CcdArgs a;
    memset(&a, 0, sizeof a);
    a.n = n; a.xdim = (uint32_t)xdim; a.ydim = (uint32_t)ydim;
    a.ccdX = (unsigned*)ccdX; a.ccdY = (unsigned*)ccdY; a.w = w; a.image = image; a.hist = hist;
    a.offchip = offchip; a.op = 1;
    hipLaunchKernelGGL(rt_ccd_kernel, dim3(grid_for(n)), dim3(kThreads), 0, (hipStream_t)stream, a);
    HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    return RT_OK;
}

extern "C" int rt_hip_scale(size_t n, double* v, const double* f, void* stream) {
    if (!v || !f) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    if (n == 0) return RT_OK;
    int rc = rt_ensure_device();
    if (rc != RT_OK) return rc;
    ScaleArgs a{n, v, f};
    hipLaunchKernelGGL(rt_scale_kernel, dim3(grid_for(n)), dim3(kThreads), 0, (hipStream_t)stream, a);
    HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    return RT_OK;
}

extern "C" int rt_hip_grating_energy(size_t n, const double* lambda, long order, double blaze, double alpha,
                                     double gamma, double gr_const, double rule_start, double rule_stop,
                                     double* energy, void* stream) {
    if (!lambda || !energy) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    if (n == 0) return RT_OK;
    int rc = rt_ensure_device();
    if (rc != RT_OK) return rc;
    GratingArgs a;
    a.n = n; a.blaze = blaze; a.alpha = alpha; a.gamma = gamma; a.d = gr_const;
    a.rule_size = rule_stop - rule_start; a.order = order; a.lambda = lambda; a.energy = energy;
    hipLaunchKernelGGL(rt_grating_energy_kernel, dim3(grid_for(n)), dim3(kThreads), 0, (hipStream_t)stream, a);
    HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    return RT_OK;
}

// Natural cubic spline second derivatives: the serial Numerical Recipes
// recurrence (rt_cuda_engine_spline.cu:42-145 semantics; its kernel runs the
// recurrence in parallel, a race).  Host code: at most 1024 knots.
extern "C" int rt_hip_spline_init(size_t n, const double* x, const double* y, double* y2) {
    if (!x || !y || !y2 || n < 2) return rt_set_error(RT_ERR_BAD_VALUE, "bad spline arguments");
    std::vector<double> u(n, 0.0);
    y2[0] = 0.0;
    y2[n - 1] = 0.0;
    for (size_t i = 1; i + 1 < n; ++i) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    for (size_t k = n - 2; k > 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
    return RT_OK;
}

// Evaluate a spline [K, x[K], y[K], y2[K]] (host record) at n device points.
extern "C" int rt_hip_spline_eval(const double* spline, size_t n, const double* x, double* y,
                                  unsigned long long* oob, void* stream) {
    if (!spline || !x || !y) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    size_t K = (size_t)spline[0];
    if (K < 2 || K > RT_MAX_SPLINE_KNOTS) return rt_set_error(RT_ERR_BAD_VALUE, "bad knot count");
    if (n == 0) return RT_OK;
    int rc = rt_ensure_device();
    if (rc != RT_OK) return rc;
    std::vector<double> rec(3 + 3 * K);
    rec[0] = (double)K;
    rec[1] = spline[1];
    rec[2] = (double)(K - 1) / (spline[K] - spline[1]);
    memcpy(rec.data() + 3, spline + 1, 3 * K * sizeof(double));
    double* dev = nullptr;
    HIP_TRY(hipMallocAsync(&dev, rec.size() * sizeof(double), (hipStream_t)stream), RT_ERR_BAD_ALLOC);
    HIP_TRY(hipMemcpyAsync(dev, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice,
                           (hipStream_t)stream), RT_ERR_FAILED);
    SplineArgs a;
    memset(&a, 0, sizeof a);
    a.n = n; a.pool_len = (int)rec.size(); a.spline = dev; a.x = x; a.y = y; a.oob = oob;
    hipLaunchKernelGGL(rt_spline_kernel, dim3(grid_for(n)), dim3(kThreads), rec.size() * sizeof(double),
                       (hipStream_t)stream, a);
    hipError_t e = hipGetLastError();
    (void)hipFreeAsync(dev, (hipStream_t)stream);
    // the host record must outlive the asynchronous copy
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream), RT_ERR_FAILED);
    if (e != hipSuccess) return rt_set_error(RT_ERR_FAILED, "spline kernel: %s", hipGetErrorString(e));
    return RT_OK;
}

// ------------------------------------------------------------ ccd_frame flux

struct FluxArgs {
    uint64_t n;
    const double *X, *Y, *L;
    const double* qe;          // device spline record [K, x0, inv, x, y, y2], or null (qe_c)
    const double* sp;          // the same for the incident spectrum (sp_c)
    double qe_c, sp_c;
    double half_x, half_y, xpix, ypix, area;
    uint32_t xdim, ydim;
    double* image;
    unsigned long long* counts;   // [0] rays on the frame, [1] off the frame, [2] lambda outside a table
};

// ccd_frame's per-file work on the device (ccd/ccd_frame.cpp:190-195): the
// pixel of every ray (ccd_coordinates_kernel, rt_cuda_engine_ccd.cu:30-45),
// its photon weight spec(lambda) lambda / hc QE(lambda) pixel area (ccd.cu:
// 286-287,360) and the atomic add into the flux image (compute_ccd_image,
// ccd.cu:219-232, with a bounds check, D10)
__global__ __launch_bounds__(kThreads) void rt_ccd_flux_kernel(FluxArgs a) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned on_n = 0, off_n = 0, oob_n = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        const uint32_t px = rt::trunc_u32((a.X[i] + a.half_x) / a.xpix);
        const uint32_t py = rt::trunc_u32((a.Y[i] + a.half_y) / a.ypix);
        if (!(px < a.xdim && py < a.ydim)) {
            ++off_n;
            continue;
        }
        const double lam = a.L[i];
        int eq = 0, es = 0;
        const double q = a.qe ? rt::spline_eval(a.qe, lam, eq) : a.qe_c;
        const double s = (a.sp ? rt::spline_eval(a.sp, lam, es) : a.sp_c) * lam;
        if (eq | es) {
            ++oob_n;
            continue;
        }
        ++on_n;
        unsafeAtomicAdd(&a.image[(uint64_t)py * a.xdim + px], s / rt::kHC * q * a.area);
    }
    if (a.counts) {
        block_add(&a.counts[0], on_n);
        block_add(&a.counts[1], off_n);
        block_add(&a.counts[2], oob_n);
    }
}

namespace {
// device record of a host spline [K, x, y, y2] (the layout spline_eval reads)
std::vector<double> spline_record(const double* spline) {
    const size_t K = (size_t)spline[0];
    std::vector<double> rec(3 + 3 * K);
    rec[0] = (double)K;
    rec[1] = spline[1];
    rec[2] = (double)(K - 1) / (spline[K] - spline[1]);
    memcpy(rec.data() + 3, spline + 1, 3 * K * sizeof(double));
    return rec;
}
}  // namespace

extern "C" int rt_hip_ccd_flux(size_t n, const double* X, const double* Y, const double* lambda,
                               const double* qe_spline, double qe_const, const double* spec_spline,
                               double spec_const, double xsize, double ysize, double xpix, double ypix, size_t xdim,
                               size_t ydim, double* image, unsigned long long* counts, void* stream) {
    if (!image || (n > 0 && (!X || !Y || !lambda))) return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    if (xdim == 0 || ydim == 0 || xdim > 0xffffffffu || ydim > 0xffffffffu ||
        (unsigned long long)xdim * ydim > 4294967296ull)
        return rt_set_error(RT_ERR_BAD_VALUE, "bad CCD dimensions");
    for (const double* sp : {qe_spline, spec_spline})
        if (sp && ((size_t)sp[0] < 2 || !(sp[(size_t)sp[0]] > sp[1])))
            return rt_set_error(RT_ERR_BAD_VALUE, "bad spline table");
    if (n == 0) return RT_OK;
    int rc = rt_ensure_device();
    if (rc != RT_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    std::vector<double> rq = qe_spline ? spline_record(qe_spline) : std::vector<double>();
    std::vector<double> rs = spec_spline ? spline_record(spec_spline) : std::vector<double>();
    std::vector<double> both(rq);
    both.insert(both.end(), rs.begin(), rs.end());
    double* dev = nullptr;
    if (!both.empty()) {
        HIP_TRY(hipMallocAsync(&dev, both.size() * sizeof(double), st), RT_ERR_BAD_ALLOC);
        HIP_TRY(hipMemcpyAsync(dev, both.data(), both.size() * sizeof(double), hipMemcpyHostToDevice, st),
                RT_ERR_FAILED);
    }
    FluxArgs a;
    memset(&a, 0, sizeof a);
    a.n = n; a.X = X; a.Y = Y; a.L = lambda;
    a.qe = qe_spline ? dev : nullptr;
    a.sp = spec_spline ? dev + rq.size() : nullptr;
    a.qe_c = qe_const; a.sp_c = spec_const;
    a.half_x = xsize / 2.0; a.half_y = ysize / 2.0; a.xpix = xpix; a.ypix = ypix;
    a.area = xpix * ypix * 1.0E-2;
    a.xdim = (uint32_t)xdim; a.ydim = (uint32_t)ydim;
    a.image = image; a.counts = counts;
    hipLaunchKernelGGL(rt_ccd_flux_kernel, dim3(grid_for(n)), dim3(kThreads), 0, st, a);
    hipError_t e = hipGetLastError();
    if (dev) (void)hipFreeAsync(dev, st);
    // the host records must outlive the asynchronous copy
    HIP_TRY(hipStreamSynchronize(st), RT_ERR_FAILED);
    if (e != hipSuccess) return rt_set_error(RT_ERR_FAILED, "ccd flux kernel: %s", hipGetErrorString(e));
    return RT_OK;
}
