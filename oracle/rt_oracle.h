/*
 * rt_oracle.h -- CPU oracle for the rt per-ray tracing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * amirych/rt engine arithmetic (cuda_engine/rt_cuda_engine_{beam,surface,spline,ccd}.cu, common-lib/beam.cpp,
 * surface.cpp, tabulated_function.cpp) used as the parity checker for the
 * HIP engine and as the timed CPU baseline.  Only tests/, the smoke check
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * It reads the same packed scheme / beam layout as the engine (include/rt_hip.h).
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 block (Salmon et al., SC'11): ctr[4], key[2] -> out[4] */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* (0,1] double from two 32-bit words: ((hi:lo) >> 11) + 1) * 2^-53 */
double oracle_u01(uint32_t hi, uint32_t lo);

/* Natural cubic spline second derivatives (Numerical Recipes spline()). */
int oracle_spline_init(size_t n, const double *x, const double *y, double *y2);
/* TabulatedFunction::operator[] (tabulated_function.cpp:173-204).
   *err = 1 for an argument outside [x0, x_{n-1}], 2 for a zero interval. */
double oracle_spline_eval(size_t n, const double *x, const double *y, const double *y2,
                          double xv, int *err);

/* generate ray r of a beam (beam.cpp:257-485 with the engine kernels) */
void oracle_gen_ray(const double *beam, uint64_t seed, uint64_t r,
                    double *X, double *Y, double *Z, double *cX, double *cY, double *cZ,
                    double *lambda);

/* Fused per-ray trace through a packed scheme (Scheme::run stage loop,
   scheme.cpp:983-1046).  Same contract as rt_hip_trace, host memory,
   OpenMP over rays (nthreads <= 0: library default). */
int oracle_trace(const double *table, int n_surf, const double *pool,
                 const double *beam, uint64_t seed, size_t ray_begin, size_t n,
                 double *X, double *Y, double *Z, double *cX, double *cY, double *cZ,
                 double *lambda, signed char *flag, int flags,
                 uint32_t *hist, double *flux, uint64_t *stats,
                 double *dump, signed char *dump_flag, int nthreads);

/* oracle_trace with the Newton iteration counts of every (surface, ray):
   iters is n_surf x n uint32 (row s at s * n); iters_mode 1 records them,
   2 forces each intersection to run exactly the given number of iterations
   (finite differences at fixed iteration counts), 0 ignores iters. */
int oracle_trace_iters(const double *table, int n_surf, const double *pool,
                       const double *beam, uint64_t seed, size_t ray_begin, size_t n,
                       double *X, double *Y, double *Z, double *cX, double *cY, double *cZ,
                       double *lambda, signed char *flag, int flags,
                       uint32_t *hist, double *flux, uint64_t *stats,
                       double *dump, signed char *dump_flag, int nthreads,
                       uint32_t *iters, int iters_mode);

/* per-range beam generators (the engine's beam_random, beam_*_cosins and
   *_range_beam split by component) */
int oracle_beam_positions(const double *beam, uint64_t seed, size_t begin, size_t n,
                          double *X, double *Y, double *Z);
int oracle_beam_directions(const double *beam, uint64_t seed, size_t begin, size_t n,
                           const double *X, const double *Y, double *cX, double *cY, double *cZ);
int oracle_beam_wavelengths(const double *beam, uint64_t seed, size_t begin, size_t n, double *lambda);

/* ---- per-stage restatements of the rt_engine_func.h entry points ---- */
int oracle_beam_transform(size_t n, const double *angles, const double *dist,
                          double *X, double *Y, double *Z,
                          double *cX, double *cY, double *cZ);
void oracle_normalize(size_t n, double *cX, double *cY, double *cZ);
int oracle_surface_intersection(const double *rec, size_t n,
                                double *X, double *Y, double *Z,
                                double *cX, double *cY, double *cZ,
                                uint64_t *n_bad, signed char *flag, uint64_t *newton_iters);
int oracle_apply_mask(const double *rec, size_t n, const double *X, const double *Y,
                      uint64_t *n_bad, signed char *flag);
int oracle_surface_reflection(const double *rec, size_t n,
                              const double *X, const double *Y, const double *Z,
                              double *cX, double *cY, double *cZ);
int oracle_surface_refraction(const double *rec, size_t n,
                              const double *X, const double *Y, const double *Z,
                              double *cX, double *cY, double *cZ, const double *mu,
                              uint64_t *n_bad, signed char *flag);
int oracle_surface_diffraction(const double *rec, size_t n,
                               const double *X, const double *Y, const double *Z,
                               double *cX, double *cY, double *cZ,
                               const double *mu, const double *L,
                               uint64_t *n_bad, signed char *flag);
int oracle_ccd_coordinates(size_t n, const double *X, const double *Y,
                           double xsize, double ysize, double xpix, double ypix,
                           uint32_t *ccdX, uint32_t *ccdY);
int oracle_grating_energy_distr(size_t n, const double *lambda, long order,
                                double blaze, double alpha, double gamma, double gr_const,
                                double rule_start, double rule_stop, double *energy);

/* rt_hip_trace_record semantics: loss code per ray, extents (n_surf x 12) */
int oracle_trace_record(const double *table, int n_surf, const double *pool,
                        const double *beam, uint64_t seed, size_t n,
                        double *X, double *Y, double *Z, double *cX, double *cY, double *cZ,
                        double *lambda, signed char *flag, uint64_t *stats,
                        unsigned char *loss, double *extents, int nthreads);

/* Beam::Rearrange after every losing stage (two-pointer loop, beam.cpp:574-598) */
int oracle_rearrange(const unsigned char *loss, size_t n, uint64_t *perm, uint64_t *counts);

/* the saved beam in file order with the reference's renormalisation of lost rays */
void oracle_beam_order(size_t n, const uint64_t *perm, const unsigned char *loss, const unsigned char *renorm,
                       const double *X, const double *Y, const double *Z, const double *cX, const double *cY,
                       const double *cZ, const double *lambda, double *out);

#ifdef __cplusplus
}
#endif

#endif
