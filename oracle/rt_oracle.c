/*
 * rt_oracle.c -- CPU oracle (plain C99, FP64, -ffp-contract=off) for the rt
 * per-ray tracing path.  TEST INFRASTRUCTURE ONLY: the parity checker for
 * the HIP engine and the timed CPU baseline.  The product never links it.
 *
 * Every stage restates the formula of the amirych/rt file:line cited in its
 * comment, in the same evaluation order, with the build decisions below
 * (numbered as in the engine sources):
 *   D2  N_bad counted exactly (reference: racy ++ in every kernel)
 *   D3  spline second derivatives by the serial NR recurrence
 *   D4  toric intersection: documented F (surface.cpp:56-61) and the Newton of
 *       iterative_toric_intersection (surface.cu:414-492), not iter_toric
 *   D5  toric grating groove vector uses dFdZ*dFdZ (surface.cu:1438,1487 typo)
 *   D6  c == 0 aspheres use the conic Newton; Newton counters reset per ray
 *   D7  one numerically stable quadratic root for every conic (incl. k = -1)
 *   D8  the root is chosen by sign(c), as in surface.cu:240-244
 *   D10 CCD accumulation with bounds check; off-chip rays counted
 *   D12 wavelength grids use the global ray index
 *   D14 seeded Philox4x32-10 counter RNG instead of clock-seeded cuRAND
 *   D15 rectangular random beams generated (beam.cu:154-163 is a stub)
 *   D18 toric asphere derivative carries the factor Y of d(Y^2i)/dY
 *       (surface.cu:87,457 drop it)
 *   D19 a one-sample uniform wavelength grid is range[0] (beam.cu:726 divides by 0)
 *   D20 a ray that dies is frozen: later surfaces neither move nor normalise it
 */
#include "rt_oracle.h"
#include "../include/rt_hip.h"

#include <math.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RT_PI 3.14159265359           /* common-lib/num_defs.h:20 */
#define DEG2RAD (RT_PI / 180.0)       /* num_defs.h:31 (macro RT_PI/180.0) */
#define PI2_BEAM 6.28318531           /* cuda_engine/rt_cuda_engine_beam.cu:41 */
#define TWO_PI 6.283185307179586      /* Box-Muller angle scale */
#define NEWTON_EPS 1.0e-7             /* surface.cu:667 */
#define NEWTON_MAX_ITER 100000ULL     /* surface.cu:668 */
#define HC 1.98644560233e-12          /* rt_cuda_engine_ccd.cu:19, erg*um */

/* ------------------------------------------------------------------ RNG */

void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double oracle_u01(uint32_t hi, uint32_t lo)
{
    uint64_t v = (((uint64_t)hi << 32) | lo) >> 11;
    return (double)(v + 1) * (1.0 / 9007199254740992.0);   /* 2^-53 */
}

static void philox_stream(uint64_t seed, uint64_t idx, uint32_t stream, uint32_t out[4])
{
    uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), stream, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    oracle_philox4x32_10(ctr, key, out);
}

/* -------------------------------------------------------------- splines */

/* Numerical Recipes spline() with natural ends; the reference's
   spline_init (rt_cuda_engine_spline.cu:42-145) sets dy2[0]=dy2[N-1]=0,
   u[0]=0 and back-substitutes dy2[k]=dy2[k]*dy2[k+1]+u[k] (:138). D3. */
int oracle_spline_init(size_t n, const double *x, const double *y, double *y2)
{
    if (n < 2) return RT_ERR_BAD_VALUE;
    double *u = (double *)malloc(n * sizeof(double));
    if (!u) return RT_ERR_BAD_ALLOC;
    y2[0] = 0.0; u[0] = 0.0; y2[n - 1] = 0.0;
    for (size_t i = 1; i + 1 < n; ++i) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    for (size_t k = n - 2; k > 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
    free(u);
    return RT_OK;
}

/* tabulated_function.cpp:173-204 */
double oracle_spline_eval(size_t n, const double *x, const double *y, const double *y2,
                          double xv, int *err)
{
    *err = 0;
    if ((xv < x[0]) || (xv > x[n - 1])) { *err = 1; return 0.0; }
    size_t il = 0, ir = n, i;
    while ((ir - il) > 1) {
        i = (il + ir) >> 1;
        if (x[i] == xv) return y[i];
        if (x[i] > xv) ir = i; else il = i;
    }
    double h = x[ir] - x[il];
    if (h == 0.0) { *err = 2; return 0.0; }
    double a = (x[ir] - xv) / h;
    double b = (xv - x[il]) / h;
    return a * y[il] + b * y[ir] + ((a * a * a - a) * y2[il] + (b * b * b - b) * y2[ir]) * (h * h) / 6.0;
}

/* (kind, value) function reference of the packed table */
static double fn_eval(const double *pool, double kind, double val, double xv, int *err)
{
    *err = 0;
    if (kind == (double)RT_FN_CONST) return val;
    const double *s = pool + (size_t)val;
    size_t k = (size_t)s[0];
    return oracle_spline_eval(k, s + 1, s + 1 + k, s + 1 + 2 * k, xv, err);
}

/* ------------------------------------------------------------ beam gen */

typedef struct { double X, Y, Z, cX, cY, cZ, lam; } ray_t;

/* beam.cu:126-152 (circle) and beam.cpp:232-244 semantics for rectangles (D15) */
static void gen_position(const double *B, uint64_t seed, uint64_t r, ray_t *ray)
{
    uint32_t o[4];
    philox_stream(seed, r, 0u, o);
    double u1 = oracle_u01(o[0], o[1]);
    double u2 = oracle_u01(o[2], o[3]);
    if ((int)B[RT_B_SHAPE] == RT_CIRCLE) {
        double phi = PI2_BEAM * u2;
        double R = (B[RT_B_P2] - B[RT_B_P1]) * sqrt(u1) + B[RT_B_P1];
        ray->X = R * cos(phi) + B[RT_B_CEN + 0];
        ray->Y = R * sin(phi) + B[RT_B_CEN + 1];
    } else {
        double W2 = B[RT_B_P2], H2 = B[RT_B_P2] / B[RT_B_P3];
        double W1 = B[RT_B_P1], H1 = B[RT_B_P1] / B[RT_B_P3];
        double x = (u1 - 0.5) * W2, y = (u2 - 0.5) * H2;
        for (uint32_t k = 1; k < 16 && W1 > 0.0 && fabs(x) < 0.5 * W1 && fabs(y) < 0.5 * H1; ++k) {
            philox_stream(seed, r, 3u + k, o);
            x = (oracle_u01(o[0], o[1]) - 0.5) * W2;
            y = (oracle_u01(o[2], o[3]) - 0.5) * H2;
        }
        ray->X = x + B[RT_B_CEN + 0];
        ray->Y = y + B[RT_B_CEN + 1];
    }
    ray->Z = B[RT_B_CEN + 2];
}

static void gen_direction(const double *B, uint64_t seed, uint64_t r, ray_t *ray)
{
    int type = (int)B[RT_B_TYPE];
    if (type == RT_BEAM_PARALLEL) {            /* beam.cpp:363-377 */
        ray->cX = 0.0; ray->cY = 0.0; ray->cZ = 1.0;
    } else if (type == RT_BEAM_CONIC) {        /* beam.cu:586-611, angle: beam.cpp:383 */
        /* params[3]*deg2rad with the macro RT_PI/180.0 expanded: (params[3]*RT_PI)/180.0;
           RT_B_ANGUNIT 1: the angle is already in radians (engine API argument) */
        double ang = B[RT_B_ANGUNIT] != 0.0 ? B[RT_B_ANGPAR] : B[RT_B_ANGPAR] * RT_PI / 180.0;
        double cs = cos(ang / 2.0), sn = sin(ang / 2.0);
        double t = atan2(ray->Y, ray->X);
        ray->cZ = cs;
        ray->cX = sn * cos(t);
        ray->cY = sn * sin(t);
    } else {                                   /* beam.cu:952-991, 1099, 1034-1043 */
        double stddev = B[RT_B_ANGPAR] / 2.35482 * RT_PI / 180 / 3600;
        double tg = tan(stddev);
        uint32_t o[4];
        philox_stream(seed, r, 1u, o);
        double u3 = oracle_u01(o[0], o[1]);
        double u4 = oracle_u01(o[2], o[3]);
        double rad = sqrt(-2.0 * log(u3));
        double tX = tg * (rad * cos(TWO_PI * u4));
        double tY = tg * (rad * sin(TWO_PI * u4));
        double cX = 0.0, cY = 0.0, cZ = 1.0, cs, sn, sig;
        cs = 1.0 / sqrt(1.0 + tX * tX);
        sig = signbit(tX) ? -1.0 : 1.0;
        sn = sig * (1.0 / sqrt(1.0 + 1.0 / tX / tX));
        cY = -sn * cZ;
        cZ = cs * cZ;
        cs = 1.0 / sqrt(1.0 + tY * tY);
        sig = signbit(tY) ? -1.0 : 1.0;
        sn = sig * (1.0 / sqrt(1.0 + 1.0 / tY / tY));
        cX = -sn * cZ;
        cZ = cs * cZ;
        ray->cX = cX; ray->cY = cY; ray->cZ = cZ;
    }
}

/* wavelengths: beam.cpp:425-484, uniform beam.cu:701-726 (D12, D19), random beam.cu:801-812 */
static double gen_lambda(const double *B, uint64_t seed, uint64_t r)
{
    int distr = (int)B[RT_B_RANGE_DISTR];
    double r0 = B[RT_B_RANGE + 0], r1 = B[RT_B_RANGE + 1];
    uint64_t nper = (uint64_t)B[RT_B_NPER];
    int composite = B[RT_B_COMPOSITE] != 0.0;
    uint64_t nl = composite ? (uint64_t)B[RT_B_NLAMBDA] : nper;
    uint64_t j = composite ? (r % nl) : r;
    if (distr == RT_RANGE_MONO) return (r1 + r0) / 2.0;
    if (distr == RT_RANGE_UNIFORM) {
        if (nl <= 1) return r0;
        double d_l = (r1 - r0) / (double)(nl - 1);
        return r0 + (double)j * d_l;
    }
    uint32_t o[4];
    philox_stream(seed, j, 2u, o);
    return (r1 - r0) * oracle_u01(o[0], o[1]) + r0;
}

void oracle_gen_ray(const double *beam, uint64_t seed, uint64_t r,
                    double *X, double *Y, double *Z, double *cX, double *cY, double *cZ,
                    double *lambda)
{
    ray_t ray;
    gen_position(beam, seed, r, &ray);
    gen_direction(beam, seed, r, &ray);
    ray.lam = gen_lambda(beam, seed, r);
    *X = ray.X; *Y = ray.Y; *Z = ray.Z; *cX = ray.cX; *cY = ray.cY; *cZ = ray.cZ; *lambda = ray.lam;
}

/* ------------------------------------------------------------ stages */

/* beam_transform: translation (beam.cu:320-331, applied when any d != 0, :437)
   then rotations about Y (X,Z), X (Y,Z), Z (X,Y), each when its angle != 0
   (beam.cu:333-350, 486-576).  sin/cos are evaluated on the host (:520-521). */
static void rot2(double sA, double cA, double *v1, double *v2)
{
    double e = *v1 * cA - *v2 * sA;
    *v2 = *v1 * sA + *v2 * cA;
    *v1 = e;
}

static void transform_ray(const double *ang, const double *dist, ray_t *r)
{
    if ((dist[2] != 0.0) || (dist[0] != 0.0) || (dist[1] != 0.0)) {
        r->X -= dist[0]; r->Y -= dist[1]; r->Z -= dist[2];
    }
    if (ang[0] != 0) { double s = sin(ang[0]), c = cos(ang[0]); rot2(s, c, &r->X, &r->Z); rot2(s, c, &r->cX, &r->cZ); }
    if (ang[1] != 0) { double s = sin(ang[1]), c = cos(ang[1]); rot2(s, c, &r->Y, &r->Z); rot2(s, c, &r->cY, &r->cZ); }
    if (ang[2] != 0) { double s = sin(ang[2]), c = cos(ang[2]); rot2(s, c, &r->X, &r->Y); rot2(s, c, &r->cX, &r->cY); }
}

/* Beam::Normalize, beam.cpp:541-569 (scaled norm) */
static void normalize_ray(double *cX, double *cY, double *cZ)
{
    double acX = fabs(*cX), acY = fabs(*cY), acZ = fabs(*cZ), norm;
    if (acX >= acY) {
        if (acX >= acZ) norm = acX * sqrt(1.0 + acY * acY / acX / acX + acZ * acZ / acX / acX);
        else norm = acZ * sqrt(1.0 + acX * acX / acZ / acZ + acY * acY / acZ / acZ);
    } else {
        if (acY >= acZ) norm = acY * sqrt(1.0 + acX * acX / acY / acY + acZ * acZ / acY / acY);
        else norm = acZ * sqrt(1.0 + acX * acX / acZ / acZ + acY * acY / acZ / acZ);
    }
    *cX /= norm; *cY /= norm; *cZ /= norm;
}

/* conic_derivs, surface.cu:45-61 */
static void conic_derivs(const double *p, double X, double Y, double *dFdX, double *dFdY)
{
    double c = p[RT_S_C], k1 = p[RT_S_K] + 1, r2, E, sum, rr;
    int N = (int)p[RT_S_NPOLY];
    r2 = X * X + Y * Y;
    E = c / sqrt(1.0 - k1 * c * c * r2);
    sum = 0.0; rr = 1.0;
    for (int i = 1; i <= N; ++i, rr *= r2) sum += (double)i * p[RT_S_A + i - 1] * rr;
    E += 2.0 * sum;
    *dFdX = -E * X;
    *dFdY = -E * Y;
}

/* toric_derivs, surface.cu:65-94 (with D18) */
static void toric_derivs(const double *p, double X, double Y, double Z,
                         double *dFdX, double *dFdY, double *dFdZ)
{
    double c = p[RT_S_C], k1 = p[RT_S_K] + 1, cR = p[RT_S_CR], Y2, Fy, dFy, sum, rr;
    int N = (int)p[RT_S_NPOLY];
    Y2 = Y * Y;
    Fy = c * Y2 / (1.0 + sqrt(1.0 - k1 * c * c * Y2));
    sum = 0.0; rr = Y2;
    for (int i = 0; i < N; ++i, rr *= Y2) sum += p[RT_S_A + i] * rr;
    Fy += sum;
    dFy = c * Y / sqrt(1.0 - k1 * c * c * Y2);
    sum = 0.0; rr = 1.0;
    for (int i = 1; i <= N; ++i, rr *= Y2) sum += (double)i * p[RT_S_A + i - 1] * rr;
    dFy += 2.0 * sum * Y;
    *dFdX = -cR * X;
    *dFdY = (cR * Fy - 1.0) * dFy;
    *dFdZ = 1.0 - cR * Z;
}

/* return 1 if alive after intersection; *niter counts Newton iterations.
   forced > 0 (test infrastructure for finite differences at fixed iteration
   counts): the Newton loops run exactly `forced` iterations instead of
   stopping on the step size, so the returned iterate is a smooth function of
   the scheme parameters; forced == 0 is the reference's stopping rule. */
static int intersect_ray(const double *p, ray_t *r, uint64_t *niter, uint64_t forced)
{
    int cls = (int)p[RT_S_CLASS];
    int N = (int)p[RT_S_NPOLY];
    double c = p[RT_S_C];
    *niter = 0;
    if (cls == RT_PLANE || (cls == RT_CONIC && N == 0 && c == 0.0)) {
        /* plane_intersection, surface.cu:154-178 (dispatch :594-603) */
        double s0 = -r->Z / r->cZ;
        if (s0 < 0.0) return 0;
        r->X += r->cX * s0;
        r->Y += r->cY * s0;
        r->Z = 0.0;
        return 1;
    }
    if (cls == RT_CONIC && N == 0) {
        /* conic_intersection, surface.cu:216-253; D7 stable root, D8 root by sign(c) */
        double k1 = p[RT_S_K] + 1;
        double A = k1 * r->cZ * r->cZ + r->cX * r->cX + r->cY * r->cY;
        double B = 2.0 * (k1 * r->Z * r->cZ + r->X * r->cX + r->Y * r->cY - r->cZ / c);
        double C = k1 * r->Z * r->Z + r->X * r->X + r->Y * r->Y - 2.0 * r->Z / c;
        double D = B * B - 4.0 * A * C;
        if (D < 0.0) return 0;
        double sq = sqrt(D);
        double sgnB = (B < 0.0) ? -1.0 : 1.0;
        double sigma = (c > 0) ? -1.0 : 1.0;      /* c > 0: (-B - sqrt(D))/2A */
        double q = -0.5 * (B + sgnB * sq);
        double s;
        if (sigma == sgnB) s = (q != 0.0) ? C / q : (-B + sigma * sq) / 2.0 / A;
        else s = q / A;
        r->X += s * r->cX;
        r->Y += s * r->cY;
        r->Z += s * r->cZ;
        return 1;
    }
    if (cls == RT_CONIC) {
        /* iterative_conic_intersection, surface.cu:257-333 (D6) */
        double k1 = p[RT_S_K] + 1;
        double X0 = r->X, Y0 = r->Y, cX = r->cX, cY = r->cY, cZ = r->cZ;
        double s0 = -r->Z / cZ, s1, s2, Xi = 0, Yi = 0, Zi = 0, F, dF, r2, rr, E, sum;
        uint64_t it = 0;
        if (!(s0 >= 0)) return 0;
        int alive = 1;
        s2 = 0.0;
        do {
            s1 = s2;
            Xi = X0 + (s1 + s0) * cX;
            Yi = Y0 + (s1 + s0) * cY;
            Zi = s1 * cZ;
            r2 = Xi * Xi + Yi * Yi;
            F = Zi - c * r2 / (1.0 + sqrt(1.0 - k1 * c * c * r2));
            sum = 0.0; rr = r2;
            for (int i = 0; i < N; ++i, rr *= r2) sum += p[RT_S_A + i] * rr;
            F -= sum;
            E = c / sqrt(1.0 - k1 * c * c * r2);
            sum = 0.0; rr = 1.0;
            for (int i = 1; i <= N; ++i, rr *= r2) sum += (double)i * p[RT_S_A + i - 1] * rr;
            E += 2.0 * sum;
            dF = -E * Xi * cX - E * Yi * cY + cZ;
            if (fabs(dF) < NEWTON_EPS) { alive = 0; break; }
            s2 = s1 - F / dF;
            ++it;
        } while (forced ? (it < forced) : ((fabs(s2 - s1) >= NEWTON_EPS) && (it <= NEWTON_MAX_ITER)));
        r->X = Xi; r->Y = Yi; r->Z = Zi;
        *niter = it;
        if (it > NEWTON_MAX_ITER) alive = 0;
        return alive;
    }
    if (cls == RT_TORIC) {
        /* iterative_toric_intersection, surface.cu:414-492 (D4, D18) */
        double k1 = p[RT_S_K] + 1, cR = p[RT_S_CR];
        double X0 = r->X, Y0 = r->Y, cX = r->cX, cY = r->cY, cZ = r->cZ;
        double s0 = -r->Z / cZ, s1, s2, Xi = 0, Yi = 0, Zi = 0, F, Fy, dF, rr, dFy, sum, Y2;
        uint64_t it = 0;
        if (!(s0 >= 0)) return 0;
        int alive = 1;
        s2 = 0.0;
        do {
            s1 = s2;
            Xi = X0 + (s1 + s0) * cX;
            Yi = Y0 + (s1 + s0) * cY;
            Zi = s1 * cZ;
            Y2 = Yi * Yi;
            Fy = c * Y2 / (1.0 + sqrt(1.0 - k1 * c * c * Y2));
            sum = 0.0; rr = Y2;
            for (int i = 0; i < N; ++i, rr *= Y2) sum += p[RT_S_A + i] * rr;
            Fy += sum;
            F = Zi - Fy - 0.5 * cR * (Xi * Xi + Zi * Zi - Fy * Fy);
            dFy = c * Yi / sqrt(1.0 - k1 * c * c * Y2);
            sum = 0.0; rr = 1.0;
            for (int i = 1; i <= N; ++i, rr *= Y2) sum += (double)i * p[RT_S_A + i - 1] * rr;
            dFy += 2.0 * sum * Yi;
            dF = -cR * Xi * cX + (cR * Fy - 1.0) * dFy * cY + (1.0 - cR * Zi) * cZ;
            if (fabs(dF) < NEWTON_EPS) { alive = 0; break; }
            s2 = s1 - F / dF;
            ++it;
        } while (forced ? (it < forced) : ((fabs(s2 - s1) >= NEWTON_EPS) && (it <= NEWTON_MAX_ITER)));
        r->X = Xi; r->Y = Yi; r->Z = Zi;
        *niter = it;
        if (it > NEWTON_MAX_ITER) alive = 0;
        return alive;
    }
    return 1;   /* AuxClass: no intersection (surface.cu:634-635) */
}

/* Surface::ApplyConstrains (surface.cpp:264-321) and Shade::ApplyConstrains (:652-699) */
static int mask_ray(const double *p, double X, double Y)
{
    int type = (int)p[RT_S_TYPE], shape = (int)p[RT_S_SHAPE];
    double s0 = p[RT_S_SIZE0], s1 = p[RT_S_SIZE1], c0 = p[RT_S_CEN0], c1 = p[RT_S_CEN1];
    double R2 = s0 * s0;
    double xr = X - c0, yr = Y - c1;
    if (type == RT_SHADE) {
        if (shape == RT_CIRCLE) return !(xr * xr + yr * yr <= R2);
        return !((xr >= -s0 / 2.0) && (xr <= s0 / 2.0) && (yr >= -s1 / 2.0) && (yr <= s1 / 2.0));
    }
    if (shape == RT_CIRCLE) return !(xr * xr + yr * yr > R2);
    return !((X < (c0 - s0 / 2.0)) || (X > (c0 + s0 / 2.0)) || (Y < (c1 - s1 / 2.0)) || (Y > (c1 + s1 / 2.0)));
}

/* surface_reflection kernels, surface.cu:771-827 */
static void reflect_ray(const double *p, ray_t *r)
{
    int cls = (int)p[RT_S_CLASS];
    if (cls == RT_PLANE) { r->cZ = -r->cZ; return; }
    if (cls == RT_CONIC) {
        double dFdX, dFdY, a2;
        conic_derivs(p, r->X, r->Y, &dFdX, &dFdY);
        a2 = 2.0 * (r->cX * dFdX + r->cY * dFdY + r->cZ) / (dFdX * dFdX + dFdY * dFdY + 1.0);
        r->cX -= a2 * dFdX;
        r->cY -= a2 * dFdY;
        r->cZ -= a2;
        return;
    }
    if (cls == RT_TORIC) {
        double dFdX, dFdY, dFdZ, a2;
        toric_derivs(p, r->X, r->Y, r->Z, &dFdX, &dFdY, &dFdZ);
        a2 = 2.0 * (r->cX * dFdX + r->cY * dFdY + r->cZ * dFdZ) / (dFdX * dFdX + dFdY * dFdY + dFdZ * dFdZ);
        r->cX -= a2 * dFdX;
        r->cY -= a2 * dFdY;
        r->cZ -= a2 * dFdZ;
    }
}

/* surface refraction kernels, surface.cu:956-1066; returns 0 on TIR */
static int refract_ray(const double *p, ray_t *r, double mu)
{
    int cls = (int)p[RT_S_CLASS];
    double a, b, a2, G;
    if (cls == RT_PLANE) {
        a = mu * r->cZ;
        b = mu * mu - 1.0;
        a2 = a * a;
        if (b > a2) return 0;
        b = a2 - b;
        G = -a + sqrt(b);
        r->cX *= mu;
        r->cY *= mu;
        r->cZ = mu * r->cZ + G;
        return 1;
    }
    if (cls == RT_CONIC) {
        double dFdX, dFdY;
        conic_derivs(p, r->X, r->Y, &dFdX, &dFdY);
        G = dFdX * dFdX + dFdY * dFdY + 1.0;
        a = mu * (r->cX * dFdX + r->cY * dFdY + r->cZ) / G;
        b = (mu * mu - 1.0) / G;
        a2 = a * a;
        if (b > a2) return 0;
        b = a2 - b;
        G = -a + sqrt(b);
        r->cX = mu * r->cX + G * dFdX;
        r->cY = mu * r->cY + G * dFdY;
        r->cZ = mu * r->cZ + G;
        return 1;
    }
    if (cls == RT_TORIC) {
        double dFdX, dFdY, dFdZ;
        toric_derivs(p, r->X, r->Y, r->Z, &dFdX, &dFdY, &dFdZ);
        G = dFdX * dFdX + dFdY * dFdY + dFdZ * dFdZ;
        a = mu * (r->cX * dFdX + r->cY * dFdY + r->cZ * dFdZ) / G;
        b = (mu * mu - 1.0) / G;
        a2 = a * a;
        if (b > a2) return 0;
        b = a2 - b;
        G = -a + sqrt(b);
        r->cX = mu * r->cX + G * dFdX;
        r->cY = mu * r->cY + G * dFdY;
        r->cZ = mu * r->cZ + G * dFdZ;
        return 1;
    }
    return 1;
}

/* grating kernels, surface.cu:1247-1520; returns 0 for an evanescent order */
static int diffract_ray(const double *p, ray_t *r, double mu, double L)
{
    int cls = (int)p[RT_S_CLASS];
    int refl = ((int)p[RT_S_GRTYPE] == RT_REFLECTIVE);
    double a, b, a2, G;
    if (cls == RT_PLANE) {
        if (!refl) {                                 /* :1247-1282 */
            a = mu * r->cZ;
            b = mu * mu - 1.0 + L * L - 2.0 * mu * L * r->cX;
            a2 = a * a;
            if (b > a2) return 0;
            b = a2 - b;
            G = -a + sqrt(b);
            r->cX = mu * r->cX - L;
            r->cY *= mu;
            r->cZ = mu * r->cZ + G;
        } else {                                     /* :1284-1318 */
            a = r->cZ;
            b = L * L - 2.0 * L * r->cX;
            a2 = a * a;
            if (b > a2) return 0;
            b = a2 - b;
            G = -a - sqrt(b);
            r->cX -= L;
            r->cZ += G;
        }
        return 1;
    }
    if (cls == RT_CONIC) {
        double dFdX, dFdY, u, v, w;
        conic_derivs(p, r->X, r->Y, &dFdX, &dFdY);
        G = dFdX / (dFdY * dFdY + 1.0);
        u = 1.0 / sqrt(1.0 + dFdX * G);
        v = -dFdY * u * G;
        w = -u * G;
        G = dFdX * dFdX + dFdY * dFdY + 1.0;
        if (!refl) {                                 /* :1320-1368 */
            a = mu * (r->cX * dFdX + r->cY * dFdY + r->cZ) / G;
            b = (mu * mu - 1.0 + L * L - 2.0 * L * mu * (r->cX * u + r->cY * v + r->cZ * w)) / G;
            a2 = a * a;
            if (b > a2) return 0;
            b = a2 - b;
            G = -a + sqrt(b);
            r->cX = mu * r->cX - L * u + G * dFdX;
            r->cY = mu * r->cY - L * v + G * dFdY;
            r->cZ = mu * r->cZ - L * w + G;
        } else {                                     /* :1370-1421 */
            a = (r->cX * dFdX + r->cY * dFdY + r->cZ) / G;
            b = (L * L - 2.0 * L * (r->cX * u + r->cY * v + r->cZ * w)) / G;
            a2 = a * a;
            if (b > a2) return 0;
            b = a2 - b;
            G = -a - sqrt(b);
            r->cX = r->cX - L * u + G * dFdX;
            r->cY = r->cY - L * v + G * dFdY;
            r->cZ = r->cZ - L * w + G;
        }
        return 1;
    }
    if (cls == RT_TORIC) {
        double dFdX, dFdY, dFdZ, u, v, w;
        toric_derivs(p, r->X, r->Y, r->Z, &dFdX, &dFdY, &dFdZ);
        G = dFdX / (dFdY * dFdY + dFdZ * dFdZ);      /* D5 */
        u = 1.0 / sqrt(1.0 + dFdX * G);
        v = -dFdY * u * G;
        w = -dFdZ * u * G;
        G = dFdX * dFdX + dFdY * dFdY + dFdZ * dFdZ;
        if (!refl) {                                 /* :1423-1470 */
            a = mu * (r->cX * dFdX + r->cY * dFdY + r->cZ * dFdZ) / G;
            b = (mu * mu - 1.0 + L * L - 2.0 * L * mu * (r->cX * u + r->cY * v + r->cZ * w)) / G;
            a2 = a * a;
            if (b > a2) return 0;
            b = a2 - b;
            G = -a + sqrt(b);
            r->cX = mu * r->cX - L * u + G * dFdX;
            r->cY = mu * r->cY - L * v + G * dFdY;
            r->cZ = mu * r->cZ - L * w + G * dFdZ;
        } else {                                     /* :1472-1520 */
            a = (r->cX * dFdX + r->cY * dFdY + r->cZ * dFdZ) / G;
            b = (L * L - 2.0 * L * (r->cX * u + r->cY * v + r->cZ * w)) / G;
            a2 = a * a;
            if (b > a2) return 0;
            b = a2 - b;
            G = -a - sqrt(b);
            r->cX += -L * u + G * dFdX;
            r->cY += -L * v + G * dFdY;
            r->cZ += -L * w + G * dFdZ;
        }
        return 1;
    }
    return 1;
}

/* the grating order of a ray: Order[0] (surface.cpp:570) or the echelle rule */
static double grating_order(const double *p, double lam)
{
    if ((int)p[RT_S_ORDMODE] == RT_ORDER_FIXED) return p[RT_S_ORDER];
    double m = floor(p[RT_S_BLAZEK] / lam + 0.5);
    if (m < p[RT_S_ORDMIN]) m = p[RT_S_ORDMIN];
    if (m > p[RT_S_ORDMAX]) m = p[RT_S_ORDMAX];
    return m;
}

/* __double2uint_rz restated: truncation toward zero, negatives and NaN -> 0,
   values >= 2^32 -> 0xffffffff */
static uint32_t trunc_u32(double v)
{
    if (!(v >= 0.0)) return 0u;
    if (v >= 4294967296.0) return 0xffffffffu;
    return (uint32_t)v;
}

static int near_edge(double v)
{
    double f = v - floor(v);
    return (f < 1e-9) || (1.0 - f < 1e-9);
}

/* ------------------------------------------------------------ fused trace */

typedef struct {
    uint64_t v[RT_STATS_LEN];
} stats_t;

#define SET_FN_ERR(e) do { if (e) { spl_oob = 1; } } while (0)

/* extents of the scheme log (scheme.cpp:993-1042): e[2q] min, e[2q+1] max */
static void ext_update(double *e, int q, double v)
{
    if (v < e[2 * q]) e[2 * q] = v;
    if (v > e[2 * q + 1]) e[2 * q + 1] = v;
}

static void trace_one(const double *table, int n_surf, const double *pool,
                      ray_t *r, int *alive_io, int flags, uint64_t ray_local,
                      size_t n, uint32_t *hist, double *flux, stats_t *st,
                      double *dump, signed char *dump_flag, uint32_t *iters, int iters_mode,
                      unsigned char *loss, double *ext)
{
    int alive = *alive_io;
    if (loss) *loss = alive ? 0 : RT_LOSS_DEAD;
    for (int s = 0; s < n_surf; ++s) {
        const double *p = table + (size_t)s * RT_SURF_STRIDE;
        int type = (int)p[RT_S_TYPE];
        double *e = ext ? ext + (size_t)s * RT_EXT_PER_SURF : NULL;
        if (alive) {
            transform_ray(p + RT_S_ANG, p + RT_S_DIST, r);          /* scheme.cpp:990 */
            normalize_ray(&r->cX, &r->cY, &r->cZ);                   /* :993 */
            if (e) ext_update(e, 0, r->cZ);                          /* :995 */
            if (type != RT_AUXTYPE) {
                uint64_t it = 0;
                st->v[RT_ST_SURF + 4 * s + 0] += 1;
                uint64_t forced = (iters && iters_mode == 2) ? iters[(size_t)s * n + ray_local] : 0;
                alive = intersect_ray(p, r, &it, forced);             /* :1002 */
                if (iters && iters_mode == 1) iters[(size_t)s * n + ray_local] = (uint32_t)it;
                st->v[RT_ST_SURF + 4 * s + 1] += it;
                if (!alive) {
                    st->v[RT_ST_SURF + 4 * s + 2] += 1;
                    if (loss) *loss = (unsigned char)RT_LOSS_CODE(s, RT_LOSS_INTERSECTION);
                } else {
                    if (e) { ext_update(e, 1, r->X); ext_update(e, 2, r->Y); }   /* :1005-1008 */
                    alive = mask_ray(p, r->X, r->Y);                  /* :1014 */
                    if (!alive && loss) *loss = (unsigned char)RT_LOSS_CODE(s, RT_LOSS_APERTURE);
                    if (alive && e) { ext_update(e, 3, r->X); ext_update(e, 4, r->Y); }   /* :1017-1020 */
                    if (alive) {                                      /* :1035 */
                        int spl_oob = 0, e1 = 0, e2 = 0;
                        if (type == RT_MIRROR) {
                            reflect_ray(p, r);
                        } else if (type == RT_LENS) {
                            double n1 = fn_eval(pool, p[RT_S_N1KIND], p[RT_S_N1VAL], r->lam, &e1);
                            double n2 = fn_eval(pool, p[RT_S_N2KIND], p[RT_S_N2VAL], r->lam, &e2);
                            SET_FN_ERR(e1 | e2);
                            if (!spl_oob) {
                                double mu = n1;                       /* surface.cu:1095-1096 */
                                mu /= n2;
                                alive = refract_ray(p, r, mu);
                            }
                        } else if (type == RT_GRATING) {
                            double n1 = fn_eval(pool, p[RT_S_N1KIND], p[RT_S_N1VAL], r->lam, &e1);
                            double n2 = fn_eval(pool, p[RT_S_N2KIND], p[RT_S_N2VAL], r->lam, &e2);
                            SET_FN_ERR(e1 | e2);
                            if (!spl_oob) {
                                double mu = n1;                       /* surface.cu:1566-1568 */
                                mu /= n2;
                                double m = grating_order(p, r->lam);
                                double L = m * r->lam / n2 / p[RT_S_GRCONST];
                                alive = diffract_ray(p, r, mu, L);
                            }
                        }
                        if (spl_oob) { alive = 0; st->v[RT_ST_SPLINE_OOB] += 1; }
                        if (alive) normalize_ray(&r->cX, &r->cY, &r->cZ);   /* :1040 */
                        if (!alive && loss) *loss = (unsigned char)RT_LOSS_CODE(s, RT_LOSS_ACTION);
                        if (alive && e) ext_update(e, 5, r->cZ);             /* :1042 */
                    }
                    if (!alive) st->v[RT_ST_SURF + 4 * s + 3] += 1;
                }
                if (alive && type == RT_DETECTOR && p[RT_S_CCD_XDIM] > 0 && (flags & (RT_TRACE_HIST | RT_TRACE_FLUX))) {
                    /* ccd_coordinates_kernel (rt_cuda_engine_ccd.cu:30-45), D10 */
                    uint64_t xdim = (uint64_t)p[RT_S_CCD_XDIM], ydim = (uint64_t)p[RT_S_CCD_YDIM];
                    double vx = (r->X + p[RT_S_CCD_XSIZE] / 2.0) / p[RT_S_CCD_XPIX];
                    double vy = (r->Y + p[RT_S_CCD_YSIZE] / 2.0) / p[RT_S_CCD_YPIX];
                    uint32_t px = trunc_u32(vx), py = trunc_u32(vy);
                    if (px < xdim && py < ydim) {
                        size_t pix = (size_t)py * xdim + px;
                        st->v[RT_ST_CCD_HITS] += 1;
                        if (near_edge(vx) || near_edge(vy)) st->v[RT_ST_CCD_EDGE] += 1;
                        if ((flags & RT_TRACE_HIST) && hist) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                            hist[pix] += 1u;
                        }
                        if ((flags & RT_TRACE_FLUX) && flux) {
                            /* compute_ccd_image_kernel (ccd.cu:219-232) with the host
                               weights of ccd.cu:286-287 and pixel area of :360 */
                            int eq = 0, es = 0;
                            double qe = fn_eval(pool, p[RT_S_QEKIND], p[RT_S_QEVAL], r->lam, &eq);
                            double sp = fn_eval(pool, p[RT_S_SPKIND], p[RT_S_SPVAL], r->lam, &es);
                            if (eq | es) {
                                st->v[RT_ST_SPLINE_OOB] += 1;
                            } else {
                                double spec = sp * r->lam;
                                double pix_area = p[RT_S_CCD_XPIX] * p[RT_S_CCD_YPIX] * 1.0E-2;
                                double w = spec / HC * qe * pix_area;
#ifdef _OPENMP
#pragma omp atomic
#endif
                                flux[pix] += w;
                            }
                        }
                    } else {
                        st->v[RT_ST_CCD_OFFCHIP] += 1;
                    }
                }
            }
        }
        if (dump) {
            double *d = dump + (size_t)s * 7 * n;
            d[0 * n + ray_local] = r->X; d[1 * n + ray_local] = r->Y; d[2 * n + ray_local] = r->Z;
            d[3 * n + ray_local] = r->cX; d[4 * n + ray_local] = r->cY; d[5 * n + ray_local] = r->cZ;
            d[6 * n + ray_local] = r->lam;
            if (dump_flag) dump_flag[(size_t)s * n + ray_local] = (signed char)alive;
        }
    }
    *alive_io = alive;
}

int oracle_trace(const double *table, int n_surf, const double *pool,
                 const double *beam, uint64_t seed, size_t ray_begin, size_t n,
                 double *X, double *Y, double *Z, double *cX, double *cY, double *cZ,
                 double *lambda, signed char *flag, int flags,
                 uint32_t *hist, double *flux, uint64_t *stats,
                 double *dump, signed char *dump_flag, int nthreads)
{
    return oracle_trace_iters(table, n_surf, pool, beam, seed, ray_begin, n, X, Y, Z, cX, cY, cZ, lambda, flag,
                              flags, hist, flux, stats, dump, dump_flag, nthreads, NULL, 0);
}

int oracle_trace_iters(const double *table, int n_surf, const double *pool,
                       const double *beam, uint64_t seed, size_t ray_begin, size_t n,
                       double *X, double *Y, double *Z, double *cX, double *cY, double *cZ,
                       double *lambda, signed char *flag, int flags,
                       uint32_t *hist, double *flux, uint64_t *stats,
                       double *dump, signed char *dump_flag, int nthreads,
                       uint32_t *iters, int iters_mode)
{
    if (n_surf < 0 || n_surf > RT_MAX_SURF) return RT_ERR_BAD_VALUE;
    int gen = (flags & RT_TRACE_GENERATE) != 0;
    int store = (flags & RT_TRACE_STORE) != 0;
    if (gen && !beam) return RT_ERR_BAD_VALUE;
    if (!gen && (!X || !Y || !Z || !cX || !cY || !cZ || !lambda || !flag)) return RT_ERR_BAD_VALUE;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    stats_t total;
    memset(&total, 0, sizeof(total));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        stats_t st;
        memset(&st, 0, sizeof(st));
#ifdef _OPENMP
#pragma omp for schedule(static, 4096)
#endif
        for (long long ii = 0; ii < (long long)n; ++ii) {
            size_t i = (size_t)ii;
            ray_t r;
            int alive;
            if (gen) {
                oracle_gen_ray(beam, seed, (uint64_t)(ray_begin + i), &r.X, &r.Y, &r.Z, &r.cX, &r.cY, &r.cZ, &r.lam);
                alive = 1;
            } else {
                r.X = X[i]; r.Y = Y[i]; r.Z = Z[i]; r.cX = cX[i]; r.cY = cY[i]; r.cZ = cZ[i]; r.lam = lambda[i];
                alive = flag[i] != 0;
            }
            st.v[RT_ST_RAYS] += 1;
            trace_one(table, n_surf, pool, &r, &alive, flags, i, n, hist, flux, &st, dump, dump_flag, iters,
                      iters_mode, NULL, NULL);
            if (alive) st.v[RT_ST_GOOD] += 1;
            if (store || (gen && X)) {
                if (X) { X[i] = r.X; Y[i] = r.Y; Z[i] = r.Z; cX[i] = r.cX; cY[i] = r.cY; cZ[i] = r.cZ; lambda[i] = r.lam; }
                if (flag) flag[i] = (signed char)alive;
            }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        for (int k = 0; k < RT_STATS_LEN; ++k) total.v[k] += st.v[k];
    }
    if (stats) for (int k = 0; k < RT_STATS_LEN; ++k) stats[k] += total.v[k];
    return total.v[RT_ST_SPLINE_OOB] ? RT_ERR_BAD_VALUE : RT_OK;
}

/* oracle_trace with the loss code of every ray and the logged extents
   (rt_hip_trace_record semantics: beam generated from `seed`, rays 0..n-1) */
int oracle_trace_record(const double *table, int n_surf, const double *pool,
                        const double *beam, uint64_t seed, size_t n,
                        double *X, double *Y, double *Z, double *cX, double *cY, double *cZ,
                        double *lambda, signed char *flag, uint64_t *stats,
                        unsigned char *loss, double *extents, int nthreads)
{
    if (n_surf < 0 || n_surf > RT_MAX_SURF || !beam || !loss) return RT_ERR_BAD_VALUE;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const size_t ne = (size_t)n_surf * RT_EXT_PER_SURF;
    double ext_all[RT_MAX_SURF * RT_EXT_PER_SURF];
    for (size_t k = 0; k < ne; ++k) ext_all[k] = (k & 1) ? -INFINITY : INFINITY;
    stats_t total;
    memset(&total, 0, sizeof(total));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        stats_t st;
        memset(&st, 0, sizeof(st));
        double ext[RT_MAX_SURF * RT_EXT_PER_SURF];
        for (size_t k = 0; k < ne; ++k) ext[k] = (k & 1) ? -INFINITY : INFINITY;
#ifdef _OPENMP
#pragma omp for schedule(static, 4096)
#endif
        for (long long ii = 0; ii < (long long)n; ++ii) {
            size_t i = (size_t)ii;
            ray_t r;
            int alive = 1;
            oracle_gen_ray(beam, seed, (uint64_t)i, &r.X, &r.Y, &r.Z, &r.cX, &r.cY, &r.cZ, &r.lam);
            st.v[RT_ST_RAYS] += 1;
            trace_one(table, n_surf, pool, &r, &alive, RT_TRACE_GENERATE, i, n, NULL, NULL, &st, NULL, NULL, NULL,
                      0, &loss[i], ext);
            if (alive) st.v[RT_ST_GOOD] += 1;
            if (X) { X[i] = r.X; Y[i] = r.Y; Z[i] = r.Z; cX[i] = r.cX; cY[i] = r.cY; cZ[i] = r.cZ; lambda[i] = r.lam; }
            if (flag) flag[i] = (signed char)alive;
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (int k = 0; k < RT_STATS_LEN; ++k) total.v[k] += st.v[k];
            for (size_t k = 0; k < ne; ++k) {
                if (k & 1) { if (ext[k] > ext_all[k]) ext_all[k] = ext[k]; }
                else if (ext[k] < ext_all[k]) ext_all[k] = ext[k];
            }
        }
    }
    if (stats) for (int k = 0; k < RT_STATS_LEN; ++k) stats[k] += total.v[k];
    if (extents)
        for (size_t k = 0; k < ne; ++k) extents[k] = isinf(ext_all[k]) ? NAN : ext_all[k];
    return total.v[RT_ST_SPLINE_OOB] ? RT_ERR_BAD_VALUE : RT_OK;
}

/* Beam::Rearrange (beam.cpp:574-598) after every stage of Scheme::run that
   lost rays, in stage order: the reference's two-pointer loop on the flags of
   [0, N_good), restated with indices (the reference walks char pointers).
   perm[p] is the ray at position p; counts (256) the rays of every code. */
int oracle_rearrange(const unsigned char *loss, size_t n, uint64_t *perm, uint64_t *counts)
{
    memset(counts, 0, 256 * sizeof(uint64_t));
    for (size_t i = 0; i < n; ++i) counts[loss[i]] += 1;
    if (counts[RT_LOSS_DEAD]) return RT_ERR_BAD_VALUE;
    for (size_t i = 0; i < n; ++i) perm[i] = i;
    char *fl = (char *)malloc(n ? n : 1);
    if (!fl) return RT_ERR_BAD_ALLOC;
    size_t N_good = n;
    for (int code = 1; code < RT_LOSS_DEAD; ++code) {
        if (!counts[code]) continue;
        for (size_t p = 0; p < N_good; ++p) fl[p] = (loss[perm[p]] == code) ? 0 : 1;
        long long start = 0, end = (long long)N_good - 1;
        size_t i = 0;
        while (start <= end) {
            while (start <= end && fl[start] == 1) { ++start; ++i; }        /* find a zero from the front */
            while (start <= end && fl[end] == 0) { --end; --N_good; }       /* find a one from the back */
            if (start < end) {
                fl[start] = 1;
                fl[end] = 0;
                uint64_t t = perm[i]; perm[i] = perm[N_good - 1]; perm[N_good - 1] = t;   /* swap_elements */
            }
        }
    }
    free(fl);
    return RT_OK;
}

/* The arrays of Beam::Save (beam.cpp:636-715) in file order, a lost ray's
   cosines first normalised renorm[code] more times (Beam::Normalize,
   beam.cpp:527-571, runs over every ray); out: 7 x n */
void oracle_beam_order(size_t n, const uint64_t *perm, const unsigned char *loss, const unsigned char *renorm,
                       const double *X, const double *Y, const double *Z, const double *cX, const double *cY,
                       const double *cZ, const double *lambda, double *out)
{
    for (size_t p = 0; p < n; ++p) {
        const uint64_t r = perm[p];
        double x = cX[r], y = cY[r], z = cZ[r];
        for (int j = 0; j < renorm[loss[r]]; ++j) normalize_ray(&x, &y, &z);
        out[0 * n + p] = X[r]; out[1 * n + p] = Y[r]; out[2 * n + p] = Z[r];
        out[3 * n + p] = x; out[4 * n + p] = y; out[5 * n + p] = z;
        out[6 * n + p] = lambda[r];
    }
}

/* ------------------------------------------------------ per-stage API */

int oracle_beam_transform(size_t n, const double *angles, const double *dist,
                          double *X, double *Y, double *Z,
                          double *cX, double *cY, double *cZ)
{
    for (size_t i = 0; i < n; ++i) {
        ray_t r = {X[i], Y[i], Z[i], cX[i], cY[i], cZ[i], 0.0};
        transform_ray(angles, dist, &r);
        X[i] = r.X; Y[i] = r.Y; Z[i] = r.Z; cX[i] = r.cX; cY[i] = r.cY; cZ[i] = r.cZ;
    }
    return RT_OK;
}

void oracle_normalize(size_t n, double *cX, double *cY, double *cZ)
{
    for (size_t i = 0; i < n; ++i) normalize_ray(cX + i, cY + i, cZ + i);
}

int oracle_surface_intersection(const double *rec, size_t n,
                                double *X, double *Y, double *Z,
                                double *cX, double *cY, double *cZ,
                                uint64_t *n_bad, signed char *flag, uint64_t *newton_iters)
{
    for (size_t i = 0; i < n; ++i) {
        ray_t r = {X[i], Y[i], Z[i], cX[i], cY[i], cZ[i], 0.0};
        uint64_t it = 0;
        int ok = intersect_ray(rec, &r, &it, 0);
        X[i] = r.X; Y[i] = r.Y; Z[i] = r.Z;
        if (newton_iters) *newton_iters += it;
        if (!ok) { flag[i] = 0; if (n_bad) ++*n_bad; }
    }
    return RT_OK;
}

int oracle_apply_mask(const double *rec, size_t n, const double *X, const double *Y,
                      uint64_t *n_bad, signed char *flag)
{
    for (size_t i = 0; i < n; ++i)
        if (!mask_ray(rec, X[i], Y[i])) { flag[i] = 0; if (n_bad) ++*n_bad; }
    return RT_OK;
}

int oracle_surface_reflection(const double *rec, size_t n,
                              const double *X, const double *Y, const double *Z,
                              double *cX, double *cY, double *cZ)
{
    for (size_t i = 0; i < n; ++i) {
        ray_t r = {X[i], Y[i], Z[i], cX[i], cY[i], cZ[i], 0.0};
        reflect_ray(rec, &r);
        cX[i] = r.cX; cY[i] = r.cY; cZ[i] = r.cZ;
    }
    return RT_OK;
}

int oracle_surface_refraction(const double *rec, size_t n,
                              const double *X, const double *Y, const double *Z,
                              double *cX, double *cY, double *cZ, const double *mu,
                              uint64_t *n_bad, signed char *flag)
{
    for (size_t i = 0; i < n; ++i) {
        ray_t r = {X[i], Y[i], Z[i], cX[i], cY[i], cZ[i], 0.0};
        if (!refract_ray(rec, &r, mu[i])) { flag[i] = 0; if (n_bad) ++*n_bad; continue; }
        cX[i] = r.cX; cY[i] = r.cY; cZ[i] = r.cZ;
    }
    return RT_OK;
}

int oracle_surface_diffraction(const double *rec, size_t n,
                               const double *X, const double *Y, const double *Z,
                               double *cX, double *cY, double *cZ,
                               const double *mu, const double *L,
                               uint64_t *n_bad, signed char *flag)
{
    for (size_t i = 0; i < n; ++i) {
        ray_t r = {X[i], Y[i], Z[i], cX[i], cY[i], cZ[i], 0.0};
        if (!diffract_ray(rec, &r, mu ? mu[i] : 1.0, L[i])) { flag[i] = 0; if (n_bad) ++*n_bad; continue; }
        cX[i] = r.cX; cY[i] = r.cY; cZ[i] = r.cZ;
    }
    return RT_OK;
}

int oracle_ccd_coordinates(size_t n, const double *X, const double *Y,
                           double xsize, double ysize, double xpix, double ypix,
                           uint32_t *ccdX, uint32_t *ccdY)
{
    for (size_t i = 0; i < n; ++i) {
        ccdX[i] = trunc_u32((X[i] + xsize / 2.0) / xpix);
        ccdY[i] = trunc_u32((Y[i] + ysize / 2.0) / ypix);
    }
    return RT_OK;
}

/* I1_func_Tarasov, surface.cu:1988-2019 */
static double I1_tarasov(double lam, double beta, double c, double k1, double eps, double rule_size)
{
    double term, I1;
    I1 = RT_PI * eps * rule_size / lam * cos(beta) / cos(c - beta) * (sin(c - k1) + sin(c - beta));
    term = cos(beta) * cos(c - k1) / cos(c - beta);
    if (fabs(I1) < RT_PI / 2.0) {
        I1 *= I1;
        I1 = (0.00761 * I1 - 0.16605) * I1 + 1;
    } else I1 = sin(I1) / (I1);
    I1 *= term;
    I1 *= I1;
    return I1;
}

/* grating_energy_distr_kernel, surface.cu:2098-2153 (blaze in c, alpha in k1,
   gamma in cR, grating constant in eps; surface.cu:2213-2241) */
int oracle_grating_energy_distr(size_t n, const double *lambda, long order,
                                double blaze, double alpha, double gamma, double gr_const,
                                double rule_start, double rule_stop, double *energy)
{
    double c = blaze, k1 = alpha, cR = gamma, eps = gr_const;
    double rule_size = rule_stop - rule_start;
    for (size_t i = 0; i < n; ++i) {
        double lam = lambda[i];
        double sin_alpha = sin(k1), cos_gamma = cos(cR);
        long long omin = (long long)(eps / lam * (sin_alpha - 1) * cos_gamma);
        long long omax = (long long)(eps / lam * (sin_alpha + 1) * cos_gamma);
        double norm = 0.0, beta;
        for (long long m = omin; m <= omax; ++m) {
            beta = asin((double)m * lam / eps / cos_gamma - sin_alpha);
            norm += I1_tarasov(lam, beta, c, k1, eps, rule_size);
        }
        beta = asin((double)order * lam / (eps * cos_gamma) - sin_alpha);
        double I1 = I1_tarasov(lam, beta, c, k1, eps, rule_size);
        energy[i] *= I1 / norm;
    }
    return RT_OK;
}

/* ------------------------------------------------- per-range beam generators */

int oracle_beam_positions(const double *beam, uint64_t seed, size_t begin, size_t n,
                          double *X, double *Y, double *Z)
{
    for (size_t i = 0; i < n; ++i) {
        ray_t r;
        gen_position(beam, seed, begin + i, &r);
        X[i] = r.X; Y[i] = r.Y; Z[i] = r.Z;
    }
    return RT_OK;
}

int oracle_beam_directions(const double *beam, uint64_t seed, size_t begin, size_t n,
                           const double *X, const double *Y, double *cX, double *cY, double *cZ)
{
    for (size_t i = 0; i < n; ++i) {
        ray_t r;
        r.X = X ? X[i] : 0.0;
        r.Y = Y ? Y[i] : 0.0;
        gen_direction(beam, seed, begin + i, &r);
        cX[i] = r.cX; cY[i] = r.cY; cZ[i] = r.cZ;
    }
    return RT_OK;
}

int oracle_beam_wavelengths(const double *beam, uint64_t seed, size_t begin, size_t n, double *lambda)
{
    for (size_t i = 0; i < n; ++i) lambda[i] = gen_lambda(beam, seed, begin + i);
    return RT_OK;
}
