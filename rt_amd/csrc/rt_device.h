// rt_device.h -- per-ray stage arithmetic of the rt tracing path for gfx950.
//
// Each function restates one engine kernel of amirych/rt (file:line in its
// comment) as a register-resident per-ray step, so the fused trace, the
// per-stage kernels behind rt_engine_func.h and the batched trace share one
// implementation.  FP64 throughout.  Build decisions D2-D20 are listed in
// rt_trace.hip.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#else
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::uint32_t uint32_t;
#endif

#include "rt_hip.h"

namespace rt {

constexpr double kPi = 3.14159265359;            // common-lib/num_defs.h:20
constexpr double kDeg2Rad = kPi / 180.0;         // num_defs.h:31
constexpr double kPi2Beam = 6.28318531;          // rt_cuda_engine_beam.cu:41
constexpr double kTwoPi = 6.283185307179586;     // Box-Muller angle scale
constexpr double kNewtonEps = 1.0e-7;            // rt_cuda_engine_surface.cu:667
constexpr unsigned kNewtonMaxIter = 100000u;     // rt_cuda_engine_surface.cu:668
constexpr double kHC = 1.98644560233e-12;        // rt_cuda_engine_ccd.cu:19

struct Ray {
    double X, Y, Z, cX, cY, cZ, lam;
};

// Read-only kernel data (surface records, beam description) is read through
// the constant address space, so wave-uniform reads become scalar loads.
typedef const __attribute__((address_space(4))) double* cptr;
typedef const __attribute__((address_space(4))) uint64_t* cptr64;

// Device-side record: the public RT_SURF_STRIDE fields plus per-surface
// constants derived on the host when the scheme is uploaded.
enum DevField {
    D_INV_C = RT_SURF_STRIDE + 0,      // 1/c (0 when c == 0)
    D_INV_D = RT_SURF_STRIDE + 1,      // 1/grating constant
    D_HALF_X = RT_SURF_STRIDE + 2,     // CCD xsize/2
    D_HALF_Y = RT_SURF_STRIDE + 3,     // CCD ysize/2
    D_K1C2 = RT_SURF_STRIDE + 4,       // (k+1) c c, the product the reference forms per ray
    D_INV_XPIX = RT_SURF_STRIDE + 5,   // 1/CCD x pixel size
    D_INV_YPIX = RT_SURF_STRIDE + 6,   // 1/CCD y pixel size
    // aperture mask constants (mask_ray): circle, size0^2; rectangle of a
    // Shade, -size0/2, size0/2, -size1/2, size1/2; other rectangle, its edges
    // cen0 -+ size0/2, cen1 -+ size1/2
    D_MASK = RT_SURF_STRIDE + 18,
    D_IA = RT_SURF_STRIDE + 8,         // i a_i, i = 1..RT_MAX_POLY (aspheric derivative terms)
    DEV_STRIDE = RT_SURF_STRIDE + 24
};

// Intersection algorithm of a surface (engine_intersection_cycle dispatch,
// rt_cuda_engine_surface.cu:593-636)
enum IntersectKind { IK_PLANE = 0, IK_QUAD = 1, IK_CONIC_NEWTON = 2, IK_TORIC_NEWTON = 3, IK_NONE = 4 };

// Surface descriptors.  Every stage is a template over one of these: DynSurf
// reads the structure of a surface (class, type, shape, number of aspheric
// terms ...) from its record at run time; StatSurf carries it as compile-time
// constants so a kernel generated for one scheme folds every structural
// branch and unrolls the polynomial loops.  Parameter values are always read
// from the record.
struct DynSurf {
    cptr p;
    __device__ int cls() const { return (int)p[RT_S_CLASS]; }
    __device__ int type() const { return (int)p[RT_S_TYPE]; }
    __device__ int shape() const { return (int)p[RT_S_SHAPE]; }
    __device__ int npoly() const { return (int)p[RT_S_NPOLY]; }
    __device__ int grtype() const { return (int)p[RT_S_GRTYPE]; }
    __device__ int ordmode() const { return (int)p[RT_S_ORDMODE]; }
    __device__ bool n1spline() const { return p[RT_S_N1KIND] == (double)RT_FN_SPLINE; }
    __device__ bool n2spline() const { return p[RT_S_N2KIND] == (double)RT_FN_SPLINE; }
    __device__ static constexpr bool n1reuse() { return false; }
    __device__ bool qespline() const { return p[RT_S_QEKIND] == (double)RT_FN_SPLINE; }
    __device__ bool spspline() const { return p[RT_S_SPKIND] == (double)RT_FN_SPLINE; }
    __device__ bool trans() const {
        return (p[RT_S_DIST + 2] != 0.0) || (p[RT_S_DIST + 0] != 0.0) || (p[RT_S_DIST + 1] != 0.0);
    }
    __device__ bool rot(int j) const { return p[RT_S_ANG + j] != 0.0; }
    __device__ bool ccd() const { return type() == RT_DETECTOR && p[RT_S_CCD_XDIM] > 0.0; }
    __device__ int ikind() const {
        const int c = cls(), n = npoly();
        if (c == RT_PLANE || (c == RT_CONIC && n == 0 && p[RT_S_C] == 0.0)) return IK_PLANE;
        if (c == RT_CONIC) return n == 0 ? IK_QUAD : IK_CONIC_NEWTON;
        if (c == RT_TORIC) return IK_TORIC_NEWTON;
        return IK_NONE;
    }
};

template <int CLS, int TYPE, int SHAPE, int NPOLY, int GRT, int ORDM, int IK, int FNMASK, int ROTMASK,
          bool TRANS, bool CCD>
struct StatSurf {
    cptr p;
    __device__ static constexpr int cls() { return CLS; }
    __device__ static constexpr int type() { return TYPE; }
    __device__ static constexpr int shape() { return SHAPE; }
    __device__ static constexpr int npoly() { return NPOLY; }
    __device__ static constexpr int grtype() { return GRT; }
    __device__ static constexpr int ordmode() { return ORDM; }
    __device__ static constexpr bool n1spline() { return (FNMASK & 1) != 0; }
    __device__ static constexpr bool n2spline() { return (FNMASK & 2) != 0; }
    // n1 is the n2 table of the previous surface's action: its value is reused
    __device__ static constexpr bool n1reuse() { return (FNMASK & 4) != 0; }
    // detector QE and spectrum given as tables (set on CCD surfaces only)
    __device__ static constexpr bool qespline() { return (FNMASK & 8) != 0; }
    __device__ static constexpr bool spspline() { return (FNMASK & 16) != 0; }
    __device__ static constexpr bool trans() { return TRANS; }
    __device__ static constexpr bool rot(int j) { return ((ROTMASK >> j) & 1) != 0; }
    __device__ static constexpr bool ccd() { return CCD; }
    __device__ static constexpr int ikind() { return IK; }
};

// ------------------------------------------------------------- FP64 math
// The hardware reciprocal / reciprocal square root estimates refined by two
// Newton-Raphson steps each (quadratic convergence: any estimate better than
// 2^-14 reaches the FP64 rounding floor), about 1 ulp from the IEEE result
// at a third of the instructions of the IEEE division sequence.
// Reciprocals and quotients take their special values from v_div_fixup_f64,
// the instruction that ends the IEEE division sequence: for a zero, infinite
// or NaN operand it returns the IEEE quotient (x/0 = inf, x/inf = 0, 0/0 and
// inf/inf NaN), otherwise the refined quotient with the operands' sign.
// Quotients beyond the FP64 range and denormal arguments are outside the
// domain of every caller.

__device__ __forceinline__ bool finite_nonzero(double r) {
    return __builtin_isfinite(r) && r != 0.0;
}

// 1/x refined, for x finite and away from zero (a zero or infinite x gives NaN)
__device__ __forceinline__ double rcp_fin(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    double q = fma(r, e, r);
    e = fma(-x, q, 1.0);
    return fma(q, e, q);
}

__device__ __forceinline__ double rcp(double x) {
    return __builtin_amdgcn_div_fixup(rcp_fin(x), x, 1.0);
}

// a / b: the refined reciprocal, the product and one residual correction
__device__ __forceinline__ double fdiv(double a, double b) {
    const double r = rcp_fin(b);
    const double q = a * r;
    return __builtin_amdgcn_div_fixup(fma(fma(-b, q, a), r, q), b, a);
}

__device__ __forceinline__ double rsqrt(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * x;
    double e = fma(-h * y, y, 0.5);
    double z = fma(y, e, y);
    e = fma(-h * z, z, 0.5);
    z = fma(z, e, z);
    return finite_nonzero(y) ? z : y;
}

// sqrt(x) from x * rsqrt(x) with one residual correction; x = +-0 and +inf
// return x (the IEEE results), negative and NaN x give NaN through ry
__device__ __forceinline__ double fsqrt_with(double x, double ry) {
    double s = x * ry;
    double t = fma(-s, s, x);
    s = fma(t, 0.5 * ry, s);
    constexpr int kZeroOrPosInf = (1 << 5) | (1 << 6) | (1 << 9);    // v_cmp_class: -0, +0, +inf
    return __builtin_amdgcn_class(x, kZeroOrPosInf) ? x : s;
}

__device__ __forceinline__ double fsqrt(double x) { return fsqrt_with(x, rsqrt(x)); }

// a / b for b finite, nonzero (the same sequence as fdiv)
__device__ __forceinline__ double fdiv_fin(double a, double b) { return fdiv(a, b); }

// 1/sqrt(x) for x > 0 finite; x <= 0 gives NaN (the Newton loops end on it)
__device__ __forceinline__ double rsqrt_pos(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * x;
    double e = fma(-h * y, y, 0.5);
    double z = fma(y, e, y);
    e = fma(-h * z, z, 0.5);
    return fma(z, e, z);
}

__device__ __forceinline__ double fsqrt_pos(double x, double ry) {
    double s = x * ry;
    double t = fma(-s, s, x);
    return fma(t, 0.5 * ry, s);
}

// sin and cos of an angle of moderate size (|x| < 2^20; the beam generators
// use |x| <= 2 pi): Cody-Waite reduction to r = x - k pi/2, |r| <= pi/4, with
// pi/2 split in two parts (the first of 31 significant bits, so k times it is
// exact), then the minimax polynomials of fdlibm's __kernel_sin and
// __kernel_cos (Sun Microsystems, 1993; coefficients checked against their
// published bit patterns), and the quadrant k mod 4.  Within 1.5 ulp of the
// correctly rounded values; no large-argument path, so few registers.
// A constant operand made at its use: an empty asm keeps the compiler from
// hoisting it out of the ray loop, where a polynomial's coefficients would
// otherwise hold about 20 registers for the whole loop (and push the trace
// into scratch at 6 waves per SIMD); it is rematerialised in scalar registers
// (scalar moves, issued beside the vector work).
__device__ __forceinline__ double kconst(double v) {
    __asm__ volatile("" : "+s"(v));
    return v;
}

__device__ __forceinline__ void sincos_bounded(double x, double& s_out, double& c_out) {
    const double kd = rint(x * kconst(6.36619772367581382433e-01));          // 2/pi
    const int k = (int)kd;
    const double r = fma(-kd, kconst(6.07710050650619224932e-11), fma(-kd, kconst(1.57079632673412561417e+00), x));
    const double z = r * r;
    // __kernel_sin: r + r^3 (S1 + z (S2 + ... + z^4 S6))
    const double ps = fma(z, fma(z, fma(z, fma(z, kconst(1.58969099521155010221e-10), kconst(-2.50507602534068634195e-08)),
                                        kconst(2.75573137070700676789e-06)),
                                 kconst(-1.98412698298579493134e-04)),
                          kconst(8.33333333332248946124e-03));
    const double sr = fma(z * r, fma(z, ps, kconst(-1.66666666666666324348e-01)), r);
    // __kernel_cos: w + (((1 - w) - z/2) + z^2 (C1 + z (C2 + ... + z^5 C6))), w = 1 - z/2
    const double pc = z * fma(z, fma(z, fma(z, fma(z, fma(z, kconst(-1.13596475577881948265e-11), kconst(2.08757232129817482790e-09)),
                                                    kconst(-2.75573143513906633035e-07)),
                                             kconst(2.48015872894767294178e-05)),
                                      kconst(-1.38888888888741095749e-03)),
                               kconst(4.16666666666666019037e-02));
    const double hz = 0.5 * z, w = 1.0 - hz;
    const double cr = w + (((1.0 - w) - hz) + z * pc);
    // quadrant: (s, c) = (sr, cr), (cr, -sr), (-sr, -cr), (-cr, sr) for k mod 4 = 0..3
    const bool swap = (k & 1) != 0;
    const double s0 = swap ? cr : sr, c0 = swap ? sr : cr;
    const uint64_t sgn_s = (uint64_t)(k & 2) << 62, sgn_c = (uint64_t)((k + 1) & 2) << 62;
    s_out = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, s0) ^ sgn_s);
    c_out = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, c0) ^ sgn_c);
}

// Natural logarithm of a normal x in (0, 1] (the Box-Muller radius): fdlibm's
// __ieee754_log (Sun Microsystems, 1993): x = 2^k (1 + f) with 1 + f in
// [sqrt(2)/2, sqrt(2)), s = f / (2 + f), log(1 + f) = f - s (f - R(s^2))
// with the minimax polynomial R of coefficients Lg1..Lg7 (checked against
// their published bit patterns), plus k ln 2 in two parts.  Within 1 ulp of
// the correctly rounded value; log(1) = 0.  Coefficients rematerialised at
// use (kconst).
__device__ __forceinline__ double log_unit(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    int hx = (int)(uint32_t)(u >> 32);
    int k = (hx >> 20) - 1023;
    hx &= 0x000fffff;
    const int i = (hx + 0x95f64) & 0x100000;
    const double xr = __builtin_bit_cast(double, ((uint64_t)(uint32_t)(hx | (i ^ 0x3ff00000)) << 32) | (uint32_t)u);
    k += i >> 20;
    const double f = xr - 1.0;
    const double s = fdiv_fin(f, 2.0 + f);
    const double dk = (double)k;
    const double z = s * s, w = z * z;
    const double t1 = w * fma(w, fma(w, kconst(1.531383769920937332e-01), kconst(2.222219843214978396e-01)),
                              kconst(3.999999999940941908e-01));
    const double t2 = z * fma(w, fma(w, fma(w, kconst(1.479819860511658591e-01), kconst(1.818357216161805012e-01)),
                                     kconst(2.857142874366239149e-01)),
                              kconst(6.666666666666735130e-01));
    const double R = t2 + t1;
    const double ln2_hi = kconst(6.93147180369123816490e-01), ln2_lo = kconst(1.90821492927058770002e-10);
    if (((hx - 0x6147a) | (0x6b851 - hx)) > 0) {
        const double hfsq = 0.5 * f * f;
        return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
    }
    return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 product per word (v_mad_u64_u32) gives both halves
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        // hi ^ c ^ k in one three-input bitwise operation (gfx950 v_bitop3_b32,
        // truth table 0x96: set where an odd number of the inputs is)
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// (0,1] from 64 random bits: (top 53 bits + 1) * 2^-53
__device__ __forceinline__ double u01(uint32_t hi, uint32_t lo) {
    uint64_t v = ((((uint64_t)hi) << 32) | lo) >> 11;
    return (double)(v + 1) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ void philox_stream(uint64_t seed, uint64_t idx, uint32_t stream,
                                              uint32_t out[4]) {
    philox4x32_10((uint32_t)idx, (uint32_t)(idx >> 32), stream, 0u,
                  (uint32_t)seed, (uint32_t)(seed >> 32), out);
}

// ----------------------------------------------------------- beam generation
// circle: random_circle_kernel (rt_cuda_engine_beam.cu:126-152);
// rectangle: uniform over the outer rectangle minus the inner one
// (beam.cpp:232-244 semantics; the reference kernel :154-163 is a stub).
__device__ __forceinline__ void gen_position(cptr B, uint64_t seed, uint64_t r, Ray& ray) {
    uint32_t o[4];
    philox_stream(seed, r, 0u, o);
    double u1 = u01(o[0], o[1]);
    double u2 = u01(o[2], o[3]);
    if ((int)B[RT_B_SHAPE] == RT_CIRCLE) {
        double phi = kPi2Beam * u2;
        double R = (B[RT_B_P2] - B[RT_B_P1]) * fsqrt(u1) + B[RT_B_P1];
        double s, c;
        sincos_bounded(phi, s, c);
        ray.X = R * c + B[RT_B_CEN + 0];
        ray.Y = R * s + B[RT_B_CEN + 1];
    } else {
        double W2 = B[RT_B_P2], H2 = B[RT_B_P2] / B[RT_B_P3];
        double W1 = B[RT_B_P1], H1 = B[RT_B_P1] / B[RT_B_P3];
        double x = (u1 - 0.5) * W2, y = (u2 - 0.5) * H2;
        for (uint32_t k = 1; k < 16 && W1 > 0.0 && fabs(x) < 0.5 * W1 && fabs(y) < 0.5 * H1; ++k) {
            philox_stream(seed, r, 3u + k, o);
            x = (u01(o[0], o[1]) - 0.5) * W2;
            y = (u01(o[2], o[3]) - 0.5) * H2;
        }
        ray.X = x + B[RT_B_CEN + 0];
        ray.Y = y + B[RT_B_CEN + 1];
    }
    ray.Z = B[RT_B_CEN + 2];
}

// Uniform (per-beam) values of the direction generators, evaluated once on
// the host as the reference does: the cone half-angle sin/cos of
// conic_cosins_kernel (beam.cu:592-598) and tan(sigma) of gauss_cosins_cycle
// (beam.cu:1034).
// lam_step: the spacing (range[1] - range[0]) / (N - 1) of a uniform
// wavelength grid of N > 1 wavelengths (N: the composite beam's NLAMBDA, else
// the ray count), the IEEE quotient formed once on the host (0 otherwise).
struct BeamUniform {
    double cone_cos, cone_sin, gauss_tg, lam_step;
};
enum {
    BU_CONE_COS = RT_BEAM_STRIDE, BU_CONE_SIN = RT_BEAM_STRIDE + 1, BU_GAUSS_TG = RT_BEAM_STRIDE + 2,
    BU_LAM_STEP = RT_BEAM_STRIDE + 3
};

// the uniform wavelength grid's size and spacing of a beam description (host
// and device; the division is IEEE, so the step equals grid_lambda's)
__host__ __device__ inline uint64_t beam_grid_size(const double* B) {
    return B[RT_B_COMPOSITE] != 0.0 ? (uint64_t)B[RT_B_NLAMBDA] : (uint64_t)B[RT_B_NPER];
}
__host__ __device__ inline double beam_lam_step(const double* B) {
    const uint64_t nl = beam_grid_size(B);
    return nl > 1 ? (B[RT_B_RANGE + 1] - B[RT_B_RANGE + 0]) / (double)(nl - 1) : 0.0;
}

template <int TYPE>
__device__ __forceinline__ void gen_direction(cptr B, uint64_t seed, uint64_t r, Ray& ray) {
    if (TYPE == RT_BEAM_PARALLEL) {                    // beam.cpp:363-377
        ray.cX = 0.0; ray.cY = 0.0; ray.cZ = 1.0;
    } else if (TYPE == RT_BEAM_CONIC) {                // conic_cosins_kernel, beam.cu:586-611
        double t = atan2(ray.Y, ray.X);
        double st, ct;
        sincos_bounded(t, st, ct);
        ray.cZ = B[BU_CONE_COS];
        ray.cX = B[BU_CONE_SIN] * ct;
        ray.cY = B[BU_CONE_SIN] * st;
    } else {                                           // gauss_cosins_kernel, beam.cu:952-991
        const double tg = B[BU_GAUSS_TG];
        uint32_t o[4];
        philox_stream(seed, r, 1u, o);
        double u3 = u01(o[0], o[1]);
        double u4 = u01(o[2], o[3]);
        double rad = fsqrt(-2.0 * log_unit(u3));
        double sb, cb;
        sincos_bounded(kTwoPi * u4, sb, cb);
        double tX = tg * (rad * cb);
        double tY = tg * (rad * sb);
        // cos = 1/sqrt(1+t^2); sin = sign(t)/sqrt(1+1/t^2) = t * cos.  For a
        // narrow beam, tan sigma < 1e-5, every |t| is below 1e-4 (the
        // Box-Muller radius is at most 8.6), and 1 - x/2 + 3x^2/8 in x = t^2
        // is within 1e-24 of 1/sqrt(1+x): no reciprocal square root.  The
        // beam's uniform value picks the branch for the whole wave.
        double csx, csy;
        if (tg < 1e-5) {
            const double x = tX * tX, y = tY * tY;
            csx = fma(x, fma(0.375, x, -0.5), 1.0);
            csy = fma(y, fma(0.375, y, -0.5), 1.0);
        } else {
            csx = rsqrt(1.0 + tX * tX);
            csy = rsqrt(1.0 + tY * tY);
        }
        double cX, cY, cZ = 1.0, sn;
        sn = tX * csx;
        cY = -sn * cZ;
        cZ = csx * cZ;
        sn = tY * csy;
        cX = -sn * cZ;
        cZ = csy * cZ;
        ray.cX = cX; ray.cY = cY; ray.cZ = cZ;
    }
}

// wavelengths: beam.cpp:425-484; uniform_range_beam_kernel (beam.cu:701-710,
// global index), random_range_beam_kernel (beam.cu:801-812).  Products and
// sums rounded separately (no FMA contraction), so every wavelength is the
// correctly rounded value of the written expression, bit for bit.
//
// grid_lambda: the wavelength of grid index j of a uniform range (j < nl, or
// any j when nl <= 1); shared by the per-ray generator and the per-wavelength
// tables, so both see the same value.
__device__ __forceinline__ double grid_lambda(cptr B, uint64_t nl, uint64_t j) {
#pragma clang fp contract(off)
    const double r0 = B[RT_B_RANGE + 0], r1 = B[RT_B_RANGE + 1];
    if (nl <= 1) return r0;
    const double d_l = (r1 - r0) / (double)(nl - 1);
    return r0 + (double)j * d_l;
}

__device__ __forceinline__ double mono_lambda(cptr B) {
#pragma clang fp contract(off)
    return (B[RT_B_RANGE + 1] + B[RT_B_RANGE + 0]) / 2.0;
}

// the wavelength of ray r and its grid index j (0 for one wavelength; the
// ray's own draw index for random wavelengths)
__device__ __forceinline__ double gen_lambda(cptr B, uint64_t seed, uint64_t r, uint64_t& j) {
#pragma clang fp contract(off)
    int distr = (int)B[RT_B_RANGE_DISTR];
    double r0 = B[RT_B_RANGE + 0], r1 = B[RT_B_RANGE + 1];
    uint64_t nper = (uint64_t)B[RT_B_NPER];
    bool composite = B[RT_B_COMPOSITE] != 0.0;
    uint64_t nl = composite ? (uint64_t)B[RT_B_NLAMBDA] : nper;
    // r mod nl in 32 bits where both fit (the 64-bit remainder is a long
    // instruction sequence)
    j = r;
    if (composite) {
        if (nl <= 1) j = 0;
        else if (((r | nl) >> 32) == 0) j = (uint32_t)r % (uint32_t)nl;
        else j = r % nl;
    }
    if (distr == RT_RANGE_MONO) { j = 0; return mono_lambda(B); }
    // grid_lambda with the host's step: r0 + j * ((r1 - r0) / (nl - 1))
    if (distr == RT_RANGE_UNIFORM) return nl <= 1 ? r0 : r0 + (double)j * B[BU_LAM_STEP];
    uint32_t o[4];
    philox_stream(seed, j, 2u, o);
    return (r1 - r0) * u01(o[0], o[1]) + r0;
}

// B: RT_BEAM_STRIDE beam fields followed by the BeamUniform values
template <int TYPE>
__device__ __forceinline__ void gen_ray(cptr B, uint64_t seed, uint64_t r, Ray& ray, uint64_t& j) {
    gen_position(B, seed, r, ray);
    gen_direction<TYPE>(B, seed, r, ray);
    ray.lam = gen_lambda(B, seed, r, j);
}

template <int TYPE>
__device__ __forceinline__ void gen_ray(cptr B, uint64_t seed, uint64_t r, Ray& ray) {
    uint64_t j;
    gen_ray<TYPE>(B, seed, r, ray, j);
}

// ------------------------------------------------------------------ splines
// TabulatedFunction::operator[] (tabulated_function.cpp:173-204).  Device
// spline record: [K, x0, (K-1)/(x[K-1]-x0), x[K], y[K], y2[K]] (staged in
// LDS).  The interval is guessed from the knot density and verified; a
// wrong guess (non-uniform knots) falls back to the reference's bisection.
// Any interval with x[i] <= xv <= x[i+1] yields the same value as the
// bisection, exact knots included (a or b is then exactly 0 and 1).
// err: 1 argument out of range, 2 zero interval.
__device__ __forceinline__ double spline_eval(const double* s, double xv, int& err) {
    // no multiply-add contraction: the kernels that evaluate a table (the
    // per-ray trace, the per-wavelength table kernels) must form the same
    // value bit for bit, and which products the compiler would fuse depends
    // on the code around each copy
#pragma clang fp contract(off)
    const int K = (int)s[0];
    const double* x = s + 3;
    const double* y = x + K;
    const double* y2 = y + K;
    if ((xv < x[0]) || (xv > x[K - 1])) { err = 1; return 0.0; }
    int il = (int)((xv - s[1]) * s[2]);
    il = il < 0 ? 0 : (il > K - 2 ? K - 2 : il);
    if (!(x[il] <= xv && xv <= x[il + 1])) {
        int lo = 0, hi = K;
        while ((hi - lo) > 1) {
            int i = (lo + hi) >> 1;
            if (x[i] > xv) hi = i; else lo = i;
        }
        il = lo < K - 1 ? lo : K - 2;
    }
    const int ir = il + 1;
    double h = x[ir] - x[il];
    if (h == 0.0) { err = 2; return 0.0; }
    double rh = rcp(h);
    double a = (x[ir] - xv) * rh;
    double b = (xv - x[il]) * rh;
    return a * y[il] + b * y[ir] + ((a * a * a - a) * y2[il] + (b * b * b - b) * y2[ir]) * (h * h) * (1.0 / 6.0);
}

__device__ __forceinline__ double fn_eval(const double* pool, bool spline, double val, double xv, int& err) {
    if (!spline) return val;
    return spline_eval(pool + (size_t)val, xv, err);
}

// ------------------------------------------------------------------- stages

// beam_transform (beam.cu:416-579): translation when any d != 0, then the
// rotations about Y, X, Z that have a non-zero angle.  sin/cos come from the
// table (evaluated on the host, as beam.cu:520-521 does).
__device__ __forceinline__ void rot2(double sA, double cA, double& v1, double& v2) {
    double e = v1 * cA - v2 * sA;
    v2 = v1 * sA + v2 * cA;
    v1 = e;
}

template <class SD>
__device__ __forceinline__ void transform_ray(const SD& d, Ray& r) {
    const cptr p = d.p;
    if (d.trans()) { r.X -= p[RT_S_DIST + 0]; r.Y -= p[RT_S_DIST + 1]; r.Z -= p[RT_S_DIST + 2]; }
    if (d.rot(0)) {
        double s = p[RT_S_SIN + 0], c = p[RT_S_COS + 0];
        rot2(s, c, r.X, r.Z); rot2(s, c, r.cX, r.cZ);
    }
    if (d.rot(1)) {
        double s = p[RT_S_SIN + 1], c = p[RT_S_COS + 1];
        rot2(s, c, r.Y, r.Z); rot2(s, c, r.cY, r.cZ);
    }
    if (d.rot(2)) {
        double s = p[RT_S_SIN + 2], c = p[RT_S_COS + 2];
        rot2(s, c, r.X, r.Y); rot2(s, c, r.cX, r.cY);
    }
}

// Beam::Normalize (beam.cpp:541-569).  The reference divides by a norm
// scaled by the largest component (7 divisions and a square root); direction
// cosines are O(1), so the unscaled norm cannot overflow and one reciprocal
// square root gives the same vector to about 1 ulp.
//
// Between stages the vector is already of unit length to rounding (rotations
// and the reflection, Snell and grating laws preserve it), |c|^2 = 1 + d with
// |d| ~ 1e-16: then 1/sqrt(1+d) = 1 - d/2 + 3d^2/8 - ..., and 1 - d/2 is
// within 3d^2/8 < 4e-17 of it for |d| < 1e-8, below half an ulp.  A wave
// whose rays are all that close takes that form; otherwise the reciprocal
// square root.
__device__ __forceinline__ void normalize_ray(double& cX, double& cY, double& cZ) {
    const double s = cX * cX + cY * cY + cZ * cZ;
    const double d = s - 1.0;
    const bool near = fabs(d) < 1e-8;
    double inv;
    if (__all(near)) inv = fma(-0.5, d, 1.0);
    else inv = near ? fma(-0.5, d, 1.0) : rsqrt(s);
    cX *= inv; cY *= inv; cZ *= inv;
}

// The even-asphere polynomial and its derivative in r2 = r^2 by one Horner
// pass over both: P = sum_{i=1..N} a_i r2^i, Q = sum_{i=1..N} i a_i r2^(i-1)
// (the reference sums the same terms by ascending powers, surface.cu:45-61,
// 257-333; the value differs only by rounding).
template <class SD>
__device__ __forceinline__ void asphere_pq(const SD& d, double r2, double& P, double& Q) {
    const cptr p = d.p;
    const int N = d.npoly();
    if (N == 0) { P = 0.0; Q = 0.0; return; }
    double pa = p[RT_S_A + N - 1], qa = p[D_IA + N - 1];
    for (int i = N - 1; i >= 1; --i) {
        pa = fma(pa, r2, p[RT_S_A + i - 1]);
        qa = fma(qa, r2, p[D_IA + i - 1]);
    }
    P = pa * r2;
    Q = qa;
}

// conic_derivs (surface.cu:45-61)
template <class SD>
__device__ __forceinline__ void conic_derivs(const SD& d, double X, double Y, double& dFdX, double& dFdY) {
    const cptr p = d.p;
    const double c = p[RT_S_C];
    double r2 = X * X + Y * Y;
    double E = c * rsqrt(1.0 - p[D_K1C2] * r2);
    double P, Q;
    asphere_pq(d, r2, P, Q);
    E += 2.0 * Q;
    dFdX = -E * X;
    dFdY = -E * Y;
}

// toric_derivs (surface.cu:65-94), asphere derivative with its factor Y (D18)
template <class SD>
__device__ __forceinline__ void toric_derivs(const SD& d, double X, double Y, double Z,
                                             double& dFdX, double& dFdY, double& dFdZ) {
    const cptr p = d.p;
    const double c = p[RT_S_C], cR = p[RT_S_CR];
    double Y2 = Y * Y;
    double t = 1.0 - p[D_K1C2] * Y2;
    double irt = rsqrt(t);
    double Fy = fdiv(c * Y2, 1.0 + fsqrt_with(t, irt));
    double P, Q;
    asphere_pq(d, Y2, P, Q);
    Fy += P;
    double dFy = c * Y * irt;
    dFy += 2.0 * Q * Y;
    dFdX = -cR * X;
    dFdY = (cR * Fy - 1.0) * dFy;
    dFdZ = 1.0 - cR * Z;
}

// Receiver of the Newton iterates of an intersection, sink(k, s^(k)) at the
// start of iteration k: none in the traces; the adjoint's forward records
// them on its tape.
struct NoIterSink {
    __device__ void operator()(unsigned, double) const {}
};

// Intersection with the surface (engine_intersection_cycle dispatch,
// surface.cu:593-636).  Returns false for a lost ray; `iters` counts Newton
// iterations.  A lost Newton ray keeps its last iterate, as the reference
// writes Xi,Yi,Zi before flagging (surface.cu:319-326).
template <class SD, class Sink = NoIterSink>
__device__ __forceinline__ bool intersect_ray(const SD& d, Ray& r, unsigned& iters, const Sink& sink = Sink{}) {
    const cptr p = d.p;
    const int ik = d.ikind();
    const double c = p[RT_S_C];
    iters = 0;
    if (ik == IK_PLANE) {
        // plane_intersection (surface.cu:154-178)
        double s0 = fdiv(-r.Z, r.cZ);
        if (s0 < 0.0) return false;
        r.X += r.cX * s0;
        r.Y += r.cY * s0;
        r.Z = 0.0;
        return true;
    }
    if (ik == IK_QUAD) {
        // conic_intersection (surface.cu:216-253), stable root (D7), root by sign(c) (D8)
        const double k1 = p[RT_S_K] + 1;
        double A = k1 * r.cZ * r.cZ + r.cX * r.cX + r.cY * r.cY;
        const double ic = p[D_INV_C];
        double B = 2.0 * (k1 * r.Z * r.cZ + r.X * r.cX + r.Y * r.cY - r.cZ * ic);
        double C = k1 * r.Z * r.Z + r.X * r.X + r.Y * r.Y - 2.0 * r.Z * ic;
        double D = B * B - 4.0 * A * C;
        if (D < 0.0) return false;
        double sq = fsqrt(D);
        double sgnB = (B < 0.0) ? -1.0 : 1.0;
        double sigma = (c > 0) ? -1.0 : 1.0;
        double q = -0.5 * (B + sgnB * sq);
        double s;
        if (sigma == sgnB) s = (q != 0.0) ? fdiv(C, q) : fdiv(-B + sigma * sq, 2.0 * A);
        else s = fdiv(q, A);
        r.X += s * r.cX;
        r.Y += s * r.cY;
        r.Z += s * r.cZ;
        return true;
    }
    if (ik == IK_CONIC_NEWTON) {
        // iterative_conic_intersection (surface.cu:257-333), counters per ray (D6)
        // Special values inside the loop: t <= 0 (outside the conic) gives NaN
        // through rsqrt_pos, and NaN or infinite F, dF end the loop at this
        // iterate exactly as the IEEE forms do (s2 is not used after the loop).
        // The step F/dF is F times a refined reciprocal, within 2 ulp of the
        // quotient: the iterates agree with the IEEE ones to rounding, and the
        // exit test |s2 - s1| >= eps only differs for a step within rounding
        // of eps.
        const double K = p[D_K1C2];
        const double X0 = r.X, Y0 = r.Y, cX = r.cX, cY = r.cY, cZ = r.cZ;
        double s0 = fdiv(-r.Z, cZ);
        if (!(s0 >= 0)) return false;
        bool alive = true;
        unsigned it = 0;
        double s1, s2 = 0.0, Xi = 0, Yi = 0, Zi = 0;
        do {
            s1 = s2;
            sink(it, s1);
            Xi = X0 + (s1 + s0) * cX;
            Yi = Y0 + (s1 + s0) * cY;
            Zi = s1 * cZ;
            double r2 = Xi * Xi + Yi * Yi;
            double t = 1.0 - K * r2;
            double irt = rsqrt_pos(t);
            const double u = 1.0 + fsqrt_pos(t, irt);
            double P, Q;
            asphere_pq(d, r2, P, Q);
            double E = c * irt + 2.0 * Q;
            double dF = -E * Xi * cX - E * Yi * cY + cZ;
            if (fabs(dF) < kNewtonEps) { alive = false; break; }
            // one reciprocal for both quotients: w = 1/(u dF), 1/u = dF w,
            // 1/dF = u w (each within 2 ulp)
            const double w = rcp_fin(u * dF);
            double F = Zi - (c * r2) * (dF * w);
            F -= P;
            s2 = fma(-F, u * w, s1);
            ++it;
        } while ((fabs(s2 - s1) >= kNewtonEps) && (it <= kNewtonMaxIter));
        r.X = Xi; r.Y = Yi; r.Z = Zi;
        iters = it;
        if (it > kNewtonMaxIter) alive = false;
        return alive;
    }
    if (ik == IK_TORIC_NEWTON) {
        // iterative_toric_intersection (surface.cu:414-492) with the documented
        // F of surface.cpp:56-61 (D4) and the asphere derivative factor (D18)
        const double K = p[D_K1C2], cR = p[RT_S_CR];
        const double X0 = r.X, Y0 = r.Y, cX = r.cX, cY = r.cY, cZ = r.cZ;
        double s0 = fdiv(-r.Z, cZ);
        if (!(s0 >= 0)) return false;
        bool alive = true;
        unsigned it = 0;
        double s1, s2 = 0.0, Xi = 0, Yi = 0, Zi = 0;
        do {
            s1 = s2;
            sink(it, s1);
            Xi = X0 + (s1 + s0) * cX;
            Yi = Y0 + (s1 + s0) * cY;
            Zi = s1 * cZ;
            double Y2 = Yi * Yi;
            double t = 1.0 - K * Y2;
            double irt = rsqrt_pos(t);
            double P, Q;
            asphere_pq(d, Y2, P, Q);
            double Fy = (c * Y2) * rcp_fin(1.0 + fsqrt_pos(t, irt)) + P;
            double F = Zi - Fy - 0.5 * cR * (Xi * Xi + Zi * Zi - Fy * Fy);
            double dFy = c * Yi * irt + 2.0 * Q * Yi;
            double dF = -cR * Xi * cX + (cR * Fy - 1.0) * dFy * cY + (1.0 - cR * Zi) * cZ;
            if (fabs(dF) < kNewtonEps) { alive = false; break; }
            s2 = fma(-F, rcp_fin(dF), s1);
            ++it;
        } while ((fabs(s2 - s1) >= kNewtonEps) && (it <= kNewtonMaxIter));
        r.X = Xi; r.Y = Yi; r.Z = Zi;
        iters = it;
        if (it > kNewtonMaxIter) alive = false;
        return alive;
    }
    return true;   // AuxClass
}

// Surface::ApplyConstrains (surface.cpp:264-321), Shade::ApplyConstrains
// (:652-699).  The bounds (s0 * s0, -+s0/2, c0 -+ s0/2, ...) are formed once
// on the host (D_MASK), the same IEEE values the reference forms per ray.
template <class SD>
__device__ __forceinline__ bool mask_ray(const SD& d, double X, double Y) {
    const cptr p = d.p;
    const int type = d.type(), shape = d.shape();
    const double c0 = p[RT_S_CEN0], c1 = p[RT_S_CEN1];
    double xr = X - c0, yr = Y - c1;
    if (type == RT_SHADE) {
        if (shape == RT_CIRCLE) return !(xr * xr + yr * yr <= p[D_MASK]);
        return !((xr >= p[D_MASK]) && (xr <= p[D_MASK + 1]) && (yr >= p[D_MASK + 2]) && (yr <= p[D_MASK + 3]));
    }
    if (shape == RT_CIRCLE) return !(xr * xr + yr * yr > p[D_MASK]);
    return !((X < p[D_MASK]) || (X > p[D_MASK + 1]) || (Y < p[D_MASK + 2]) || (Y > p[D_MASK + 3]));
}

// surface_reflection kernels (surface.cu:771-827)
template <class SD>
__device__ __forceinline__ void reflect_ray(const SD& d, Ray& r) {
    const int cls = d.cls();
    if (cls == RT_PLANE) { r.cZ = -r.cZ; return; }
    if (cls == RT_CONIC) {
        double dFdX, dFdY;
        conic_derivs(d, r.X, r.Y, dFdX, dFdY);
        double a2 = fdiv(2.0 * (r.cX * dFdX + r.cY * dFdY + r.cZ), dFdX * dFdX + dFdY * dFdY + 1.0);
        r.cX -= a2 * dFdX;
        r.cY -= a2 * dFdY;
        r.cZ -= a2;
    } else if (cls == RT_TORIC) {
        double dFdX, dFdY, dFdZ;
        toric_derivs(d, r.X, r.Y, r.Z, dFdX, dFdY, dFdZ);
        double a2 = fdiv(2.0 * (r.cX * dFdX + r.cY * dFdY + r.cZ * dFdZ), dFdX * dFdX + dFdY * dFdY + dFdZ * dFdZ);
        r.cX -= a2 * dFdX;
        r.cY -= a2 * dFdY;
        r.cZ -= a2 * dFdZ;
    }
}

// refraction kernels (surface.cu:956-1066); false on total internal reflection
template <class SD>
__device__ __forceinline__ bool refract_ray(const SD& d, Ray& r, double mu) {
    const int cls = d.cls();
    double a, b, a2, G;
    if (cls == RT_PLANE) {
        a = mu * r.cZ;
        b = mu * mu - 1.0;
        a2 = a * a;
        if (b > a2) return false;
        b = a2 - b;
        G = -a + fsqrt(b);
        r.cX *= mu;
        r.cY *= mu;
        r.cZ = mu * r.cZ + G;
        return true;
    }
    double dFdX, dFdY, dFdZ = 1.0;
    if (cls == RT_CONIC) conic_derivs(d, r.X, r.Y, dFdX, dFdY);
    else if (cls == RT_TORIC) toric_derivs(d, r.X, r.Y, r.Z, dFdX, dFdY, dFdZ);
    else return true;
    G = dFdX * dFdX + dFdY * dFdY + dFdZ * dFdZ;
    const double iG = rcp(G);
    a = mu * (r.cX * dFdX + r.cY * dFdY + r.cZ * dFdZ) * iG;
    b = (mu * mu - 1.0) * iG;
    a2 = a * a;
    if (b > a2) return false;
    b = a2 - b;
    G = -a + fsqrt(b);
    r.cX = mu * r.cX + G * dFdX;
    r.cY = mu * r.cY + G * dFdY;
    r.cZ = mu * r.cZ + G * dFdZ;
    return true;
}

// grating kernels (surface.cu:1247-1520); false for an evanescent order
template <class SD>
__device__ __forceinline__ bool diffract_ray(const SD& d, Ray& r, double mu, double L) {
    const int cls = d.cls();
    const bool refl = (d.grtype() == RT_REFLECTIVE);
    double a, b, a2, G;
    if (cls == RT_PLANE) {
        if (!refl) {                                   // :1247-1282
            a = mu * r.cZ;
            b = mu * mu - 1.0 + L * L - 2.0 * mu * L * r.cX;
            a2 = a * a;
            if (b > a2) return false;
            b = a2 - b;
            G = -a + fsqrt(b);
            r.cX = mu * r.cX - L;
            r.cY *= mu;
            r.cZ = mu * r.cZ + G;
        } else {                                       // :1284-1318
            a = r.cZ;
            b = L * L - 2.0 * L * r.cX;
            a2 = a * a;
            if (b > a2) return false;
            b = a2 - b;
            G = -a - fsqrt(b);
            r.cX -= L;
            r.cZ += G;
        }
        return true;
    }
    double dFdX, dFdY, dFdZ = 1.0, u, v, w;
    if (cls == RT_CONIC) {
        conic_derivs(d, r.X, r.Y, dFdX, dFdY);
        G = fdiv(dFdX, dFdY * dFdY + 1.0);
        u = rsqrt(1.0 + dFdX * G);
        v = -dFdY * u * G;
        w = -u * G;
    } else if (cls == RT_TORIC) {
        toric_derivs(d, r.X, r.Y, r.Z, dFdX, dFdY, dFdZ);
        G = fdiv(dFdX, dFdY * dFdY + dFdZ * dFdZ);    // D5
        u = rsqrt(1.0 + dFdX * G);
        v = -dFdY * u * G;
        w = -dFdZ * u * G;
    } else {
        return true;
    }
    G = dFdX * dFdX + dFdY * dFdY + dFdZ * dFdZ;
    const double iG = rcp(G);
    if (!refl) {                                       // :1320-1368, :1423-1470
        a = mu * (r.cX * dFdX + r.cY * dFdY + r.cZ * dFdZ) * iG;
        b = (mu * mu - 1.0 + L * L - 2.0 * L * mu * (r.cX * u + r.cY * v + r.cZ * w)) * iG;
        a2 = a * a;
        if (b > a2) return false;
        b = a2 - b;
        G = -a + fsqrt(b);
        r.cX = mu * r.cX - L * u + G * dFdX;
        r.cY = mu * r.cY - L * v + G * dFdY;
        r.cZ = mu * r.cZ - L * w + G * dFdZ;
    } else {                                           // :1370-1421, :1472-1520
        a = (r.cX * dFdX + r.cY * dFdY + r.cZ * dFdZ) * iG;
        b = (L * L - 2.0 * L * (r.cX * u + r.cY * v + r.cZ * w)) * iG;
        a2 = a * a;
        if (b > a2) return false;
        b = a2 - b;
        G = -a - fsqrt(b);
        if (cls == RT_CONIC) {
            r.cX = r.cX - L * u + G * dFdX;
            r.cY = r.cY - L * v + G * dFdY;
            r.cZ = r.cZ - L * w + G * dFdZ;
        } else {
            r.cX += -L * u + G * dFdX;
            r.cY += -L * v + G * dFdY;
            r.cZ += -L * w + G * dFdZ;
        }
    }
    return true;
}

// A derived-table value whose wavelength lies outside a function's table: a
// NaN with its own payload (compared bit for bit, so it differs from any NaN
// the arithmetic makes).
constexpr uint64_t kOutOfTableBits = 0x7ff8000000010001ull;
__device__ __forceinline__ double out_of_table_value() { return __builtin_bit_cast(double, kOutOfTableBits); }
__device__ __forceinline__ bool out_of_table(double v) { return __builtin_bit_cast(uint64_t, v) == kOutOfTableBits; }

// grating order of a ray: Order[0] (surface.cpp:570) or the echelle blaze rule
template <class SD>
__device__ __forceinline__ double grating_order(const SD& d, double lam) {
    const cptr p = d.p;
    if (d.ordmode() == RT_ORDER_FIXED) return p[RT_S_ORDER];
    double m = floor(p[RT_S_BLAZEK] / lam + 0.5);
    m = fmax(m, p[RT_S_ORDMIN]);
    m = fmin(m, p[RT_S_ORDMAX]);
    return m;
}

// truncation toward zero to uint32 (ccd_coordinates_kernel, ccd.cu:30-45):
// negative and NaN -> 0, >= 2^32 -> 0xffffffff
__device__ __forceinline__ uint32_t trunc_u32(double v) {
    if (!(v >= 0.0)) return 0u;
    if (v >= 4294967296.0) return 0xffffffffu;
    return (uint32_t)v;
}

__device__ __forceinline__ bool near_edge(double v) {
    double f = v - floor(v);
    return (f < 1e-9) || (1.0 - f < 1e-9);
}

// ------------------------------------------------------------ fused trace

constexpr int kBlock = 256;               // 4 waves of 64
constexpr int kMaxPoolLds = 12288;        // doubles of device spline pool staged in LDS (96 KiB)

struct BeamArg {
    double v[RT_BEAM_STRIDE];
    BeamUniform u;
};

// kernel arguments (one struct, so generated kernels share the launch code)
struct TraceArgs {
    const double* table;          // n_surf device records of DEV_STRIDE
    const double* pool;           // device spline pool
    int n_surf, pool_len, flags, pad;
    BeamArg beam;
    uint64_t seed, ray_begin, n;
    double *X, *Y, *Z, *cX, *cY, *cZ, *L;
    signed char* F;
    unsigned* hist;
    double* flux;
    unsigned long long* stats;
    double* dump;
    signed char* dump_flag;
    // batched traces (trace_batch_body): instance i owns records
    // table[i * inst_stride ..], rays [ray_start[i], ray_start[i+1]) of the
    // concatenated outputs and tiles [tile_start[i], tile_start[i+1]) of one
    // workgroup each; its beam is generated from seed + i
    const uint64_t* tile_start;
    const uint64_t* ray_start;
    double* tile_moments;         // per tile: n, mean x, mean y, M2 x, M2 y of the rays alive at the end
    uint64_t n_inst, n_tiles, inst_stride;
    // grid-wavelength beams: per-wavelength values, ltab[q * lt_nk + j] for
    // column q at grid index j (generated kernels built for it only): the
    // derived table (dtab_body) for a single scheme, the spline tables
    // (ltab_body) for a batch
    const double* ltab;
    uint64_t lt_nk;
    const uint32_t* tile_inst;    // batched traces: the instance of each tile
    // batched traces with a beam per instance: n_inst records of kBeamRec
    // doubles (beam fields, then the BeamUniform values), or null (the beam
    // above for every instance)
    const double* inst_beams;
    // recording traces (kernels built with kRecLoss / kRecExt): the loss code
    // of every ray, and the extents keys (n_surf x kExtPerSurf, ord_key
    // encoded, combined with atomic min / max)
    unsigned char* loss;
    unsigned long long* ext;
    // group spot partials (kernels built with kRecSpots, and batched traces
    // when set): per group and tile of blockDim.x rays (ray r in group
    // r mod n_groups; r counted from the instance's first ray in a batch),
    // group-major, group_part[(g * tiles + tile) * 5 ..] = n, mean x, mean y,
    // M2 x, M2 y (tiles: of the launch, or the batch's n_tiles)
    double* group_part;
    uint64_t n_groups;
};
constexpr int kBeamRec = (int)(sizeof(BeamArg) / sizeof(double));
// copies of the statistics counters the trace kernels add to (TraceArgs::stats
// points at kStatCopies x RT_STATS_LEN counters, workgroup b adding to copy
// b % kStatCopies)
constexpr int kStatCopies = 64;

// stats[k] += sum of the copies' counter k, and the copies cleared for the
// next launch (one workgroup).  The caller's counters are added to
// atomically: traces of distinct schemes on distinct streams may fold into
// one caller buffer at the same time.
__device__ __forceinline__ void stats_fold_body(unsigned long long* copies, unsigned long long* stats) {
    for (int k = threadIdx.x; k < RT_STATS_LEN; k += blockDim.x) {
        unsigned long long t = 0;
        for (int c = 0; c < kStatCopies; ++c) {
            t += copies[(size_t)c * RT_STATS_LEN + k];
            copies[(size_t)c * RT_STATS_LEN + k] = 0;
        }
        if (t) atomicAdd(&stats[k], t);
    }
}

// Recording modes of the trace kernels (template argument MODE of
// surface_step and trace_body), for Scheme::run's saved beam and log:
// kRecLoss keeps every ray's loss code (0 alive at the end, 1 + 3 s + k for a
// ray lost at surface s in stage k: 0 intersection, 1 aperture mask, 2 action;
// kLossDead for a ray that entered dead); kRecExt keeps, per surface, the
// extents the reference logs (scheme.cpp:993-1042): min and max of cZ after
// the first normalisation, of X and Y after the intersection, of X and Y after
// the mask, of cZ after the action's normalisation, over the rays alive there.
constexpr int kRecLoss = 1;
constexpr int kRecExt = 2;
// kRecSpots: the trace walks its rays a workgroup tile at a time and writes
// each tile's group spot partials (TraceArgs::group_part)
constexpr int kRecSpots = 4;
constexpr unsigned char kLossDead = 255;
constexpr int kExtPerSurf = 12;

// order-preserving integer key of a double (unsigned compare = double compare)
__device__ __forceinline__ unsigned long long ord_key(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double key_value(unsigned long long k) {
    const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __builtin_bit_cast(double, u);
}

struct TraceCtx {
    cptr table;                   // device records (per-iteration copy of TraceArgs::table)
    const double* pool;           // LDS copy
    unsigned* st;                 // per-block LDS counters (RT_STATS_LEN)
    unsigned long long* ext;      // per-block LDS extents keys (recording kernels)
    unsigned char loss;           // loss code of the current ray (recording kernels)
    unsigned* hist;
    double* flux;
    bool do_hist, do_flux;
    unsigned n_oob, n_hit, n_off, n_edge;
    double n_prev;                // n2 of the last action (reused where the next n1 is the same table)
    unsigned last_it;             // Newton iterations of the last intersection (adjoint tape)
    uint32_t lk;                  // wavelength grid index of the current ray
};

__device__ __forceinline__ void lds_add(unsigned* counter, unsigned v) {
    if (v) atomicAdd(counter, v);
}

// Hit counts pre-aggregated within the wave before the global atomics: the
// first pending lane's pixel takes every lane of the wave that hit the same
// pixel in one atomic add.  The peel is bounded: it stops after kPeelRounds
// rounds, or after a round that merged a single lane (hits scattered over the
// wave, e.g. one wavelength per lane), and the lanes left add their own hit.
constexpr int kPeelRounds = 8;

__device__ __forceinline__ void hist_add_wave(unsigned* hist, uint32_t pix, bool on) {
    unsigned long long pending = __ballot(on);
    bool mine = on;
    for (int round = 0; round < kPeelRounds && pending; ++round) {
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t lp = (uint32_t)__builtin_amdgcn_readlane((int)pix, leader);
        const bool match = mine && pix == lp;
        const unsigned long long m = __ballot(match);
        if ((int)__lane_id() == leader) atomicAdd(&hist[lp], (unsigned)__popcll(m));
        mine = mine && !match;
        pending &= ~m;
        if (__popcll(m) == 1) break;
    }
    if (mine) atomicAdd(&hist[pix], 1u);
}

// Extents: the wave's min and max of v over the lanes with `on`, folded into
// the block's keys e[2 q] (min) and e[2 q + 1] (max).  Called by every lane
// of the wave (converged code).
__device__ __forceinline__ void ext_add(unsigned long long* e, int q, bool on, double v) {
    double lo = on ? v : __builtin_inf(), hi = on ? v : -__builtin_inf();
    for (int m = 32; m >= 1; m >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, m));
        hi = fmax(hi, __shfl_xor(hi, m));
    }
    const bool any = __ballot(on) != 0;
    if (any && __lane_id() == 0) {
        atomicMin(&e[2 * q], ord_key(lo));
        atomicMax(&e[2 * q + 1], ord_key(hi));
    }
}

// One surface of the Scheme::run stage loop (common-lib/scheme.cpp:983-1046)
// for one ray: transform, normalize, intersection, aperture mask, action
// (reflection, refraction or diffraction), normalize, then the CCD frame on
// a detector.  Counters are wave-aggregated into LDS.
//
// lv, col: the ray's per-wavelength table values (generated kernels of
// grid-wavelength beams).  DERIV false: this surface's spline-valued
// functions read lv[col], lv[col + 1], ... in the order n1 (unless reused),
// n2, QE, spectrum, a NaN meaning a wavelength outside the table; col < 0
// evaluates the splines here.  DERIV true: the derived table's columns
// (dtab_body) give mu, then the grating term, or the detector's flux weight,
// kOutOfTable where a function's wavelength is outside its table.
template <class SD, int MODE = 0, bool DERIV = false, class Sink = NoIterSink>
__device__ __forceinline__ void surface_step(const SD& d0, int s, Ray& r, bool& alive, TraceCtx& c,
                                             const double* lv = nullptr, int col = -1, const Sink& sink = Sink{}) {
    // the record's loads are issued at this surface, not hoisted ahead of
    // the previous surfaces' work (that would keep every surface's
    // parameters live in SGPRs at once)
    SD d = d0;
    __asm__ volatile("" : "+s"(d.p));
    const cptr p = d.p;
    if (alive) {
        transform_ray(d, r);                                     // scheme.cpp:990
        // :993.  A frame change without rotation leaves the direction the
        // previous surface's last normalisation produced; normalising it again
        // (as the reference does) changes a component by at most one ulp, so
        // only the first surface and rotated frames normalise here
        if (s == 0 || d.rot(0) || d.rot(1) || d.rot(2)) normalize_ray(r.cX, r.cY, r.cZ);
    }
    if constexpr ((MODE & kRecExt) != 0) ext_add(c.ext + s * kExtPerSurf, 0, alive, r.cZ);   // :995
    if (d.type() == RT_AUXTYPE) return;                          // :1000
    unsigned long long entering = __ballot(alive);
    if (__lane_id() == 0 && entering) atomicAdd(&c.st[RT_ST_SURF + 4 * s], (unsigned)__popcll(entering));
    unsigned it = 0;
    bool lost_int = false, lost_mask = false, lost_act = false;
    if (alive) {
        alive = intersect_ray(d, r, it, sink);                   // :1002
        lost_int = !alive;
    }
    c.last_it = it;
    const bool met = alive;           // alive after the intersection
    const unsigned long long met_mask = __ballot(met);
    bool passed = false;              // alive after the aperture mask
    if (alive) {
        alive = mask_ray(d, r.X, r.Y);                           // :1014
        passed = alive;
        lost_mask = !alive;
        if (alive) {                                             // :1035
            if (d.type() == RT_MIRROR) {
                reflect_ray(d, r);
            } else if (DERIV && (d.type() == RT_LENS || d.type() == RT_GRATING)) {
                const double mu = lv[col];
                if (out_of_table(mu)) {
                    alive = false;
                    ++c.n_oob;
                } else if (d.type() == RT_LENS) {
                    alive = refract_ray(d, r, mu);
                } else {
                    alive = diffract_ray(d, r, mu, lv[col + 1]);
                }
            } else if (d.type() == RT_LENS || d.type() == RT_GRATING) {
                int e1 = 0, e2 = 0;
                double n1, n2;
                if (col >= 0) {
                    int q = col;
                    const bool t1 = !d.n1reuse() && d.n1spline();
                    n1 = d.n1reuse() ? c.n_prev : (t1 ? lv[q++] : p[RT_S_N1VAL]);
                    n2 = d.n2spline() ? lv[q] : p[RT_S_N2VAL];
                    e1 = t1 && n1 != n1;
                    e2 = d.n2spline() && n2 != n2;
                } else {
                    n1 = d.n1reuse() ? c.n_prev : fn_eval(c.pool, d.n1spline(), p[RT_S_N1VAL], r.lam, e1);
                    n2 = fn_eval(c.pool, d.n2spline(), p[RT_S_N2VAL], r.lam, e2);
                }
                c.n_prev = n2;
                if (e1 | e2) {
                    alive = false;
                    ++c.n_oob;
                } else {
                    // mu = n1/n2 (surface.cu:1095-1096), L = m lambda / n2 / d (:1568)
                    const double in2 = rcp(n2);
                    double mu = n1 * in2;
                    if (d.type() == RT_LENS) {
                        alive = refract_ray(d, r, mu);
                    } else {
                        double m = grating_order(d, r.lam);
                        double Lg = m * r.lam * in2 * p[D_INV_D];
                        alive = diffract_ray(d, r, mu, Lg);
                    }
                }
            }
            // :1040.  Only an action changes the direction since the entry
            // normalisation: a stop, a Shade or a detector keeps it (the
            // reference's renormalisation would move it by one ulp at most)
            if (alive && (d.type() == RT_MIRROR || d.type() == RT_LENS || d.type() == RT_GRATING))
                normalize_ray(r.cX, r.cY, r.cZ);
            lost_act = !alive;
        }
    }
    if constexpr ((MODE & kRecExt) != 0) {
        // the action changes only the cosines, so X, Y are still those of
        // the intersection (:1005-1008) and of the mask (:1017-1020)
        unsigned long long* e = c.ext + s * kExtPerSurf;
        ext_add(e, 1, met, r.X);
        ext_add(e, 2, met, r.Y);
        ext_add(e, 3, passed, r.X);
        ext_add(e, 4, passed, r.Y);
        ext_add(e, 5, alive, r.cZ);                              // :1042
    }
    if constexpr ((MODE & kRecLoss) != 0) {
        if (lost_int) c.loss = (unsigned char)(1 + 3 * s);
        if (lost_mask) c.loss = (unsigned char)(2 + 3 * s);
        if (lost_act) c.loss = (unsigned char)(3 + 3 * s);
    }
    if (d.ikind() == IK_CONIC_NEWTON || d.ikind() == IK_TORIC_NEWTON) lds_add(&c.st[RT_ST_SURF + 4 * s + 1], it);
    // lost at the intersection: entered, not met; at the mask or the action:
    // met, not alive at the end (one ballot each of the alive flag)
    const unsigned long long li = entering & ~met_mask, la = met_mask & ~__ballot(alive);
    if (__lane_id() == 0) {
        if (li) atomicAdd(&c.st[RT_ST_SURF + 4 * s + 2], (unsigned)__popcll(li));
        if (la) atomicAdd(&c.st[RT_ST_SURF + 4 * s + 3], (unsigned)__popcll(la));
    }
    if (d.ccd() && (c.do_hist || c.do_flux)) {
        // ccd_coordinates_kernel (rt_cuda_engine_ccd.cu:30-45) with a bounds check (D10);
        // xdim * ydim <= 2^32 (validated), so a pixel index fits 32 bits.  The
        // coordinate divided by the pixel size is taken as a product with the
        // reciprocal (within 1 ulp of the quotient: a ray can change pixel
        // only within about 1e-12 px of an edge, and rays within 1e-9 px of
        // one are counted in RT_ST_CCD_EDGE)
        const uint32_t xdim = (uint32_t)p[RT_S_CCD_XDIM], ydim = (uint32_t)p[RT_S_CCD_YDIM];
        bool on = false;
        uint32_t pix = 0;
        if (alive) {
            double vx = (r.X + p[D_HALF_X]) * p[D_INV_XPIX];
            double vy = (r.Y + p[D_HALF_Y]) * p[D_INV_YPIX];
            uint32_t px = trunc_u32(vx), py = trunc_u32(vy);
            on = px < xdim && py < ydim;
            if (on) {
                pix = py * xdim + px;
                ++c.n_hit;
                if (near_edge(vx) || near_edge(vy)) ++c.n_edge;
                if (DERIV && c.do_flux) {
                    const double w = lv[col];
                    if (out_of_table(w)) ++c.n_oob;
                    else unsafeAtomicAdd(&c.flux[pix], w);
                } else if (c.do_flux) {
                    // compute_ccd_image_kernel (ccd.cu:219-232), weights of ccd.cu:286-287,360
                    int eq = 0, es = 0;
                    double qe, sp;
                    if (col >= 0) {
                        int q = col;
                        qe = d.qespline() ? lv[q++] : p[RT_S_QEVAL];
                        sp = d.spspline() ? lv[q] : p[RT_S_SPVAL];
                        eq = d.qespline() && qe != qe;
                        es = d.spspline() && sp != sp;
                    } else {
                        qe = fn_eval(c.pool, d.qespline(), p[RT_S_QEVAL], r.lam, eq);
                        sp = fn_eval(c.pool, d.spspline(), p[RT_S_SPVAL], r.lam, es);
                    }
                    if (eq | es) {
                        ++c.n_oob;
                    } else {
                        double spec = sp * r.lam;
                        double pix_area = p[RT_S_CCD_XPIX] * p[RT_S_CCD_YPIX] * 1.0E-2;
                        double w = spec * (1.0 / kHC) * qe * pix_area;
                        unsafeAtomicAdd(&c.flux[pix], w);
                    }
                }
            } else {
                ++c.n_off;
            }
        }
        if (c.do_hist) hist_add_wave(c.hist, pix, on);
    }
}

// TraceArgs is the only argument of the trace kernels, at offset 0 of the
// kernel argument segment; a pointer to it through the constant address
// space, re-derived at each use (the empty asm keeps the loads at the use).
typedef const __attribute__((address_space(4))) TraceArgs* cargs;
__device__ __forceinline__ cargs kernel_args() {
    cargs ka = (cargs)__builtin_amdgcn_kernarg_segment_ptr();
    __asm__ volatile("" : "+s"(ka));
    return ka;
}

__device__ __forceinline__ void dump_state(const TraceArgs& a, int s, const Ray& r, bool alive, uint64_t i) {
    const uint64_t n = a.n;
    double* d = a.dump + (size_t)s * 7 * n;
    d[0 * n + i] = r.X; d[1 * n + i] = r.Y; d[2 * n + i] = r.Z;
    d[3 * n + i] = r.cX; d[4 * n + i] = r.cY; d[5 * n + i] = r.cZ;
    d[6 * n + i] = r.lam;
    a.dump_flag[(size_t)s * n + i] = (signed char)alive;
}

// Column q of the current ray's per-wavelength table values, loaded at the
// start of the surface that uses it (its latency hidden behind that
// surface's transform and intersection).  The grid index is clamped into the
// table.
template <int NC>
__device__ __forceinline__ void load_lval(double (&lv)[NC], const TraceCtx& c, int q) {
    const cargs ka = kernel_args();
    const uint64_t nk = ka->lt_nk;
    const uint64_t k = (uint64_t)c.lk < nk ? (uint64_t)c.lk : nk - 1;
    lv[q] = ka->ltab[(uint64_t)q * nk + k];
}

// Arguments of the table kernel; the beam first (offset 0 of the argument
// segment, read through the constant address space as in the traces).
struct LtabArgs {
    BeamArg beam;
    const double* pool;           // device spline pool
    const double* col_off;        // pool offset of each column's table
    double* out;                  // n_cols x nk
    uint64_t nk;
    int n_cols, pad;
};

// out[q * nk + j]: column q's table at the wavelength of grid index j of the
// beam (one wavelength, or a composite uniform grid), evaluated by the same
// spline code and wavelength formula as the per-ray path; NaN outside the
// table.
__device__ __forceinline__ void ltab_body(const LtabArgs& a) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)a.n_cols * a.nk) return;
    const uint64_t q = idx / a.nk, j = idx % a.nk;
    cptr B = (cptr)__builtin_amdgcn_kernarg_segment_ptr();
    const double lam = (int)B[RT_B_RANGE_DISTR] == RT_RANGE_MONO
                           ? mono_lambda(B)
                           : grid_lambda(B, (uint64_t)B[RT_B_NLAMBDA], j);
    int err = 0;
    const double v = spline_eval(a.pool + (size_t)a.col_off[q], lam, err);
    a.out[idx] = err ? __builtin_nan("") : v;
}

// Arguments of the derived-table kernel; the beam first (offset 0 of the
// argument segment, read through the constant address space).
struct DtabArgs {
    BeamArg beam;
    const double* table;          // n_surf device records
    const double* pool;           // device spline pool
    double* out;                  // n_cols x nk
    uint64_t nk;
    int n_surf, pad;
};

// out[q * nk + j] at the wavelength of grid index j for the surface of
// workgroup row blockIdx.y, in the column order of rt_jit_dtab_layout
// (surfaces in order): for a lens or a grating mu = n1 / n2, and for a
// grating L = m lambda / (n2 d), as the per-ray action forms them
// (surface.cu:1095-1096,1568); for a CCD detector the flux weight of
// compute_ccd_image_kernel (ccd.cu:286-287,360) as the per-ray flux forms
// it.  Functions evaluated by the per-ray spline code on the same
// wavelength, so every value is the one the per-ray path computes;
// kOutOfTable where the per-ray path finds a wavelength outside a table.
__device__ __forceinline__ int dtab_cols(const DynSurf& d) {
    const int type = d.type();
    return type == RT_GRATING ? 2 : (type == RT_LENS || d.ccd()) ? 1 : 0;
}

__device__ __forceinline__ void dtab_body(const DtabArgs& a) {
    const int s = blockIdx.y;
    const DynSurf d{(cptr)(a.table + (size_t)s * DEV_STRIDE)};
    const int nc = dtab_cols(d);
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nc == 0 || j >= a.nk) return;
    uint64_t q = 0;                                   // the surface's first column
    for (int k = 0; k < s; ++k) q += dtab_cols(DynSurf{(cptr)(a.table + (size_t)k * DEV_STRIDE)});
    cptr B = (cptr)__builtin_amdgcn_kernarg_segment_ptr();
    const double lam = (int)B[RT_B_RANGE_DISTR] == RT_RANGE_MONO
                           ? mono_lambda(B)
                           : grid_lambda(B, (uint64_t)B[RT_B_NLAMBDA], j);
    const cptr p = d.p;
    const int type = d.type();
    if (type == RT_LENS || type == RT_GRATING) {
        int e1 = 0, e2 = 0;
        const double n1 = fn_eval(a.pool, d.n1spline(), p[RT_S_N1VAL], lam, e1);
        const double n2 = fn_eval(a.pool, d.n2spline(), p[RT_S_N2VAL], lam, e2);
        double mu = out_of_table_value(), Lg = out_of_table_value();
        if (!(e1 | e2)) {
            const double in2 = rcp(n2);
            mu = n1 * in2;
            if (type == RT_GRATING) {
                const double m = grating_order(d, lam);
                Lg = m * lam * in2 * p[D_INV_D];
            }
        }
        a.out[q * a.nk + j] = mu;
        if (type == RT_GRATING) a.out[(q + 1) * a.nk + j] = Lg;
    } else {
        int eq = 0, es = 0;
        const double qe = fn_eval(a.pool, d.qespline(), p[RT_S_QEVAL], lam, eq);
        const double sp = fn_eval(a.pool, d.spspline(), p[RT_S_SPVAL], lam, es);
        double w = out_of_table_value();
        if (!(eq | es)) {
            double spec = sp * lam;
            double pix_area = p[RT_S_CCD_XPIX] * p[RT_S_CCD_YPIX] * 1.0E-2;
            w = spec * (1.0 / kHC) * qe * pix_area;
        }
        a.out[q * a.nk + j] = w;
    }
}

// Group spot partials of one tile of blockDim.x consecutive rays, the first
// of index r0 (ray r in group r mod G): per group, the moments (n, mean x,
// mean y, M2 x, M2 y) of the tile's rays alive at the end (two passes over
// them: sums, then squared deviations from the mean), written group-major to
// part[(g * tiles + tile) * 5 ..] (a group's rows are consecutive for the
// merge that reads them).  Called by every thread of the workgroup, with
// `buf` alternating between consecutive calls (the LDS copies are double
// buffered, so one barrier per tile suffices).
__device__ __forceinline__ void tile_group_moments(double* part, uint64_t tile, uint64_t tiles, uint64_t r0,
                                                   uint64_t G, bool alive, double x, double y, unsigned buf) {
    __shared__ double s_gx[2][kBlock], s_gy[2][kBlock];
    __shared__ unsigned char s_ga[2][kBlock];
    const int B = blockDim.x;                         // <= kBlock
    const unsigned h = buf & 1u;
    s_gx[h][threadIdx.x] = x;
    s_gy[h][threadIdx.x] = y;
    s_ga[h][threadIdx.x] = alive ? 1 : 0;
    __syncthreads();
    const uint64_t shift = r0 % G;
    for (uint64_t g = threadIdx.x; g < G; g += (uint64_t)B) {
        const uint64_t q0 = (g + G - shift) % G;
        double n = 0.0, sx = 0.0, sy = 0.0;
        for (uint64_t q = q0; q < (uint64_t)B; q += G)
            if (s_ga[h][q]) {
                n += 1.0;
                sx += s_gx[h][q];
                sy += s_gy[h][q];
            }
        double mx = 0.0, my = 0.0, qx = 0.0, qy = 0.0;
        if (n > 0.0) {
            mx = sx / n;
            my = sy / n;
            for (uint64_t q = q0; q < (uint64_t)B; q += G)
                if (s_ga[h][q]) {
                    const double dx = s_gx[h][q] - mx, dy = s_gy[h][q] - my;
                    qx += dx * dx;
                    qy += dy * dy;
                }
        }
        double* o = part + (g * tiles + tile) * 5;
        o[0] = n; o[1] = mx; o[2] = my; o[3] = qx; o[4] = qy;
    }
}

// Kernel body: the spline pool goes to LDS, every lane of a persistent grid
// walks rays in a grid-stride loop (generated in-kernel when BEAM >= 0,
// loaded otherwise), `surfaces` applies the scheme, counters are flushed per
// block.  `surfaces(r, alive, ctx, i)` is the runtime loop over the table or
// a generated straight-line sequence of StatSurf steps.
template <int BEAM, int MODE = 0, class Surfaces>
__device__ __forceinline__ void trace_body(const TraceArgs& a, Surfaces surfaces) {
    extern __shared__ double s_pool[];
    __shared__ unsigned s_stats[RT_STATS_LEN];
    __shared__ unsigned long long s_ext[(MODE & kRecExt) ? RT_MAX_SURF * kExtPerSurf : 1];
    for (int i = threadIdx.x; i < RT_STATS_LEN; i += blockDim.x) s_stats[i] = 0u;
    for (int i = threadIdx.x; i < a.pool_len; i += blockDim.x) s_pool[i] = a.pool[i];
    if constexpr ((MODE & kRecExt) != 0)
        for (int i = threadIdx.x; i < RT_MAX_SURF * kExtPerSurf; i += blockDim.x) s_ext[i] = (i & 1) ? 0ull : ~0ull;
    __syncthreads();

    TraceCtx c;
    c.pool = s_pool;
    c.st = s_stats;
    c.ext = s_ext;
    c.loss = 0;
    c.hist = a.hist;
    c.flux = a.flux;
    c.do_hist = (a.flags & RT_TRACE_HIST) && a.hist;
    c.do_flux = (a.flags & RT_TRACE_FLUX) && a.flux;
    c.n_oob = c.n_hit = c.n_off = c.n_edge = 0;
    c.n_prev = 0.0;
    const bool store = (a.flags & RT_TRACE_STORE) != 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;

    // Recording kernels walk the rays a wave at a time: every lane stays in
    // the loop while the wave's first lane has a ray, a lane past the end
    // carrying a dead ray, so the extents' wave reductions read live lanes
    // only; spot kernels likewise a workgroup at a time (the tile's partials
    // are formed by the whole workgroup).  Other kernels let each lane leave
    // after its own last ray.
    constexpr bool kTiles = (MODE & kRecSpots) != 0;
    unsigned tile_iter = 0;                           // the workgroup's tiles so far (kTiles)
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
         MODE == 0 ? i < a.n : (kTiles ? i - threadIdx.x : i - __lane_id()) < a.n; i += stride) {
        const bool valid = MODE == 0 || i < a.n;
        // Surface parameters stay per-iteration scalar loads: the table
        // pointer is passed through an empty asm so the loads cannot be
        // hoisted out of this loop (hoisted, they outnumber the SGPRs and
        // spill to VGPR lanes).
        cptr tbl = (cptr)a.table;
        __asm__ volatile("" : "+s"(tbl));
        c.table = tbl;
        Ray r;
        bool alive;
        if constexpr (BEAM >= 0) {
            // beam fields and seed re-read per ray from the kernel argument
            // segment (TraceArgs is the kernel's only argument, at offset 0)
            typedef const __attribute__((address_space(4))) char* cbyte;
            cbyte karg = (cbyte)__builtin_amdgcn_kernarg_segment_ptr();
            __asm__ volatile("" : "+s"(karg));
            cptr bm = (cptr)(karg + __builtin_offsetof(TraceArgs, beam));
            cptr64 sd = (cptr64)(karg + __builtin_offsetof(TraceArgs, seed));
            uint64_t j;
            gen_ray<BEAM>(bm, *sd, a.ray_begin + i, r, j);
            c.lk = (uint32_t)j;
            alive = valid;
        } else {
            c.lk = 0;
            const cargs ka = kernel_args();
            const uint64_t k = valid ? i : 0;
            r.X = ka->X[k]; r.Y = ka->Y[k]; r.Z = ka->Z[k];
            r.cX = ka->cX[k]; r.cY = ka->cY[k]; r.cZ = ka->cZ[k];
            r.lam = ka->L[k];
            alive = valid && ka->F[k] != 0;
        }
        if constexpr ((MODE & kRecLoss) != 0) c.loss = alive ? 0 : kLossDead;
        surfaces(r, alive, c, i);
        if constexpr ((MODE & kRecLoss) != 0)
            if (valid) kernel_args()->loss[i] = c.loss;
        if ((store || BEAM >= 0) && valid) {
            // the ray pointers are read from the argument segment where they are
            // used, not held in SGPRs across the surface chain
            const cargs ka = kernel_args();
            if (ka->X) {
                ka->X[i] = r.X; ka->Y[i] = r.Y; ka->Z[i] = r.Z;
                ka->cX[i] = r.cX; ka->cY[i] = r.cY; ka->cZ[i] = r.cZ;
                ka->L[i] = r.lam;
            }
            if (ka->F) ka->F[i] = (signed char)alive;
        }
        unsigned long long good = __ballot(alive), all = __ballot(valid);
        if (__lane_id() == 0) {
            atomicAdd(&s_stats[RT_ST_RAYS], (unsigned)__popcll(all));
            if (good) atomicAdd(&s_stats[RT_ST_GOOD], (unsigned)__popcll(good));
        }
        if constexpr (kTiles)
            tile_group_moments(kernel_args()->group_part, (i - threadIdx.x) / blockDim.x,
                               (a.n + blockDim.x - 1) / blockDim.x, a.ray_begin + (i - threadIdx.x), a.n_groups,
                               alive, r.X, r.Y, tile_iter++);
    }
    lds_add(&s_stats[RT_ST_SPLINE_OOB], c.n_oob);
    lds_add(&s_stats[RT_ST_CCD_HITS], c.n_hit);
    lds_add(&s_stats[RT_ST_CCD_OFFCHIP], c.n_off);
    lds_add(&s_stats[RT_ST_CCD_EDGE], c.n_edge);
    __syncthreads();
    if (a.stats) {
        // one of kStatCopies copies of the counters per workgroup (folded
        // afterwards): thousands of workgroups adding to the same few lines at
        // the end of a launch would serialise on them
        unsigned long long* dst = a.stats + (size_t)(blockIdx.x % kStatCopies) * RT_STATS_LEN;
        for (int k = threadIdx.x; k < RT_STATS_LEN; k += blockDim.x)
            if (s_stats[k]) atomicAdd(&dst[k], (unsigned long long)s_stats[k]);
    }
    if constexpr ((MODE & kRecExt) != 0) {
        for (int k = threadIdx.x; k < a.n_surf * kExtPerSurf; k += blockDim.x) {
            const unsigned long long v = s_ext[k];
            if (k & 1) {
                if (v != 0ull) atomicMax(&a.ext[k], v);
            } else if (v != ~0ull) {
                atomicMin(&a.ext[k], v);
            }
        }
    }
}

// A double moved between lanes by a DPP pattern inside rows of 16 lanes
// (every source lane active).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// x and y exchanged by v_permlane16_swap (H = 16: odd rows of 16 lanes of x
// with even rows of y) or v_permlane32_swap (H = 32: upper half of x with
// lower half of y), the two results added: lanes of row pair (2k, 2k+1), or of
// the two halves, get x's rows summed in the even row (half) and y's rows
// summed in the odd one, the lower row's value first.
template <int H>
__device__ __forceinline__ double swap_sum(double x, double y) {
    const uint64_t ux = __builtin_bit_cast(uint64_t, x), uy = __builtin_bit_cast(uint64_t, y);
    uint32_t l0, l1, h0, h1;
    if (H == 16) {
        const auto a = __builtin_amdgcn_permlane16_swap((uint32_t)ux, (uint32_t)uy, false, false);
        const auto b = __builtin_amdgcn_permlane16_swap((uint32_t)(ux >> 32), (uint32_t)(uy >> 32), false, false);
        l0 = a[0]; l1 = a[1]; h0 = b[0]; h1 = b[1];
    } else {
        const auto a = __builtin_amdgcn_permlane32_swap((uint32_t)ux, (uint32_t)uy, false, false);
        const auto b = __builtin_amdgcn_permlane32_swap((uint32_t)(ux >> 32), (uint32_t)(uy >> 32), false, false);
        l0 = a[0]; l1 = a[1]; h0 = b[0]; h1 = b[1];
    }
    return __builtin_bit_cast(double, ((uint64_t)h0 << 32) | l0) +
           __builtin_bit_cast(double, ((uint64_t)h1 << 32) | l1);
}

// sum over the 16 lanes of each row (pairs, quads, 8, 16), the same value in
// every lane of the row
__device__ __forceinline__ double row_sum(double v) {
    v += dpp_f64<0xB1>(v);                   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);                   // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);                  // row_half_mirror
    v += dpp_f64<0x140>(v);                  // row_mirror
    return v;
}

// Sum over the 64 lanes of the wave in a fixed tree (pairs, quads, 8, 16
// lanes by DPP, then rows by lane swaps), the same value in every lane and on
// every run; VALU only (no LDS permutes).  Every lane must be active.
__device__ __forceinline__ double wave_sum(double v) {
    v = row_sum(v);
    v = swap_sum<16>(v, v);
    return swap_sum<32>(v, v);
}

// Sum over the workgroup of K values per lane, the same in every lane: a
// fixed-tree sum within each wave (wave_sum), then the wave partials in wave
// order, so the result is the same bit for bit on every run.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* s_red) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (__lane_id() == 0)
        for (int k = 0; k < K; ++k) s_red[w * K + k] = v[k];
    __syncthreads();
    for (int k = 0; k < K; ++k) {
        double t = 0.0;
        for (int j = 0; j < nw; ++j) t += s_red[j * K + k];
        v[k] = t;
    }
}

// Batched kernel body: a persistent grid walks workgroup tiles; a tile holds
// up to blockDim rays of one instance, so that instance's surface records are
// wave-uniform (scalar loads) while instances of uneven ray counts share one
// launch.  Each tile leaves the spot moments of its surviving rays (position on
// the last surface) for a fixed-order per-instance combination afterwards.
template <int BEAM, class Surfaces>
__device__ __forceinline__ void trace_batch_body(const TraceArgs& a, Surfaces surfaces) {
    extern __shared__ double s_pool[];
    __shared__ unsigned s_stats[RT_STATS_LEN];
    __shared__ double s_red[16 * 3];
    for (int i = threadIdx.x; i < RT_STATS_LEN; i += blockDim.x) s_stats[i] = 0u;
    for (int i = threadIdx.x; i < a.pool_len; i += blockDim.x) s_pool[i] = a.pool[i];
    __syncthreads();

    TraceCtx c;
    c.pool = s_pool;
    c.st = s_stats;
    c.hist = nullptr;
    c.flux = nullptr;
    c.do_hist = c.do_flux = false;
    c.n_oob = c.n_hit = c.n_off = c.n_edge = 0;
    c.n_prev = 0.0;
    const bool store = (a.flags & RT_TRACE_STORE) != 0;

    unsigned tile_iter = 0;                           // the workgroup's tiles so far
    for (uint64_t t = blockIdx.x; t < a.n_tiles; t += gridDim.x) {
        // instance of tile t, from the per-tile map (one scalar load; uniform)
        const cargs ka = kernel_args();
        const cptr64 ts = (cptr64)ka->tile_start;
        const uint64_t inst = (uint32_t)__builtin_amdgcn_readfirstlane(
            (int)((const __attribute__((address_space(4))) uint32_t*)ka->tile_inst)[t]);
        const cptr64 rs = (cptr64)ka->ray_start;
        const uint64_t r0 = rs[inst], n_i = rs[inst + 1] - r0;
        const uint64_t r = (t - ts[inst]) * blockDim.x + threadIdx.x;
        const bool valid = r < n_i;
        cptr tbl = (cptr)a.table + inst * a.inst_stride;
        __asm__ volatile("" : "+s"(tbl));
        c.table = tbl;
        Ray ray;
        bool alive = valid;
        c.lk = 0;
        if (valid) {
            cptr bm = ka->inst_beams
                          ? (cptr)ka->inst_beams + inst * (uint64_t)kBeamRec
                          : (cptr)((const __attribute__((address_space(4))) char*)ka + __builtin_offsetof(TraceArgs, beam));
            uint64_t j;
            gen_ray<BEAM>(bm, ka->seed + inst, r, ray, j);
            c.lk = (uint32_t)j;
        } else {
            ray.X = ray.Y = ray.Z = ray.cX = ray.cY = ray.lam = 0.0;
            ray.cZ = 1.0;
        }
        surfaces(ray, alive, c, r0 + r);
        if (store && valid) {
            const cargs kb = kernel_args();
            const uint64_t g = r0 + r;
            if (kb->X) {
                kb->X[g] = ray.X; kb->Y[g] = ray.Y; kb->Z[g] = ray.Z;
                kb->cX[g] = ray.cX; kb->cY[g] = ray.cY; kb->cZ[g] = ray.cZ;
                kb->L[g] = ray.lam;
            }
            if (kb->F) kb->F[g] = (signed char)alive;
        }
        unsigned long long good = __ballot(alive), all = __ballot(valid);
        if (__lane_id() == 0) {
            if (all) atomicAdd(&s_stats[RT_ST_RAYS], (unsigned)__popcll(all));
            if (good) atomicAdd(&s_stats[RT_ST_GOOD], (unsigned)__popcll(good));
        }
        // spot moments of the tile (asked for): count and means, then
        // squared deviations
        if (a.tile_moments) {
            double m1[3] = {alive ? 1.0 : 0.0, alive ? ray.X : 0.0, alive ? ray.Y : 0.0};
            block_sum<3>(m1, s_red);
            const double cnt = m1[0];
            const double mx = cnt > 0.0 ? m1[1] / cnt : 0.0, my = cnt > 0.0 ? m1[2] / cnt : 0.0;
            double m2[2] = {alive ? (ray.X - mx) * (ray.X - mx) : 0.0, alive ? (ray.Y - my) * (ray.Y - my) : 0.0};
            block_sum<2>(m2, s_red);
            if (threadIdx.x == 0) {
                double* o = a.tile_moments + t * 5;
                o[0] = cnt; o[1] = mx; o[2] = my; o[3] = m2[0]; o[4] = m2[1];
            }
        }
        // group spot partials of the tile (asked for)
        if (a.group_part)
            tile_group_moments(a.group_part, t, a.n_tiles, (t - ts[inst]) * blockDim.x, a.n_groups, alive, ray.X,
                               ray.Y, tile_iter++);
    }
    lds_add(&s_stats[RT_ST_SPLINE_OOB], c.n_oob);
    __syncthreads();
    if (a.stats) {
        // one of kStatCopies copies of the counters per workgroup (folded
        // afterwards): thousands of workgroups adding to the same few lines at
        // the end of a launch would serialise on them
        unsigned long long* dst = a.stats + (size_t)(blockIdx.x % kStatCopies) * RT_STATS_LEN;
        for (int k = threadIdx.x; k < RT_STATS_LEN; k += blockDim.x)
            if (s_stats[k]) atomicAdd(&dst[k], (unsigned long long)s_stats[k]);
    }
}

}  // namespace rt
