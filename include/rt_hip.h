/*
 * rt_hip.h -- C ABI of the MI355X (gfx950) per-ray tracing engine.
 *
 * This is the thin extern "C" boundary underneath the rt engine plugin API
 * (common-lib/rt_engine_func.h:25-104 of amirych/rt).  The 19 C++ functions
 * of that header are implemented in librt_engine_hip.so on top of the
 * entry points below; callers that keep rays resident in HBM (a fused
 * Scheme::run, the Python tests, bench.py) call these directly.
 *
 * Conventions
 *   - Every function returns an int status that mirrors RT_engine_error
 *     (common-lib/rt_engine_errors.h:4-5): RT_OK .. RT_ERR_BAD_VALUE.
 *   - Buffers named *_dev are device (HBM) pointers owned by the caller.
 *     Pointers without the suffix are host pointers.
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *   - All floating-point data is FP64.  A ray is alive while flag != 0;
 *     the engine only ever writes flag = 0 (rt_engine_func.h semantics).
 *   - No function throws; no function keeps per-call global state, so
 *     distinct schemes may be traced concurrently on distinct streams.  A
 *     scheme (or batch) object holds working buffers of its traces (the
 *     per-wavelength table, the copies its kernels add statistics to before
 *     they are folded into the caller's counters): the traces of one scheme
 *     run in the order of one stream.
 *
 * Packed scheme table
 *   A scheme is `n_surf` records of RT_SURF_STRIDE doubles (layout: enum
 *   rt_surf_field) plus a spline pool.  A spline stored at pool offset o is
 *   [K, x[0..K-1], y[0..K-1], y2[0..K-1]] (natural cubic spline; y2 from
 *   the Numerical Recipes recurrence).  A refractive index / QE / spectrum
 *   reference is a (kind, value) pair: kind RT_FN_CONST -> the constant
 *   `value`; kind RT_FN_SPLINE -> spline at pool offset `value`.
 */
#ifndef RT_HIP_H
#define RT_HIP_H

#ifndef __HIPCC_RTC__
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: same values as enum RT_engine_error (rt_engine_errors.h:4-5) */
enum rt_status {
    RT_OK = 0,
    RT_ERR_INIT_FAILED = 1,
    RT_ERR_BAD_ALLOC = 2,
    RT_ERR_FAILED = 3,
    RT_ERR_BAD_VALUE = 4
};

#define RT_MAX_SURF 32        /* surfaces per scheme */
#define RT_MAX_POLY 10        /* MAX_POLY_COEFFS, common-lib/surface.h:10 */
#define RT_MAX_SPLINE_KNOTS 1024

/* Surface::SurfaceClass / SurfaceType / SurfaceShape (common-lib/surface.h:25-27) */
enum rt_surf_class { RT_PLANE = 0, RT_CONIC = 1, RT_TORIC = 2, RT_AUXCLASS = 3 };
enum rt_surf_type  { RT_MIRROR = 0, RT_LENS = 1, RT_GRATING = 2, RT_STOP = 3,
                     RT_SHADE = 4, RT_DETECTOR = 5, RT_AUXTYPE = 6 };
enum rt_surf_shape { RT_CIRCLE = 0, RT_RECTANGLE = 1 };
/* Grating::GratingType (common-lib/surface.h:136) */
enum rt_grating_type { RT_REFLECTIVE = 0, RT_TRANSPARENT = 1 };
/* order rule of a grating: the reference's single order per grating, or the
   echelle rule m = clamp(floor(K/lambda + 0.5), m_min, m_max), K = 2 d sin(blaze) */
enum rt_order_mode { RT_ORDER_FIXED = 0, RT_ORDER_BLAZE = 1 };
/* (kind, value) function references */
enum rt_fn_kind { RT_FN_CONST = 0, RT_FN_SPLINE = 1 };

/* one packed surface record, RT_SURF_STRIDE doubles */
enum rt_surf_field {
    RT_S_CLASS = 0, RT_S_TYPE = 1, RT_S_SHAPE = 2,
    RT_S_C = 3,          /* vertex curvature 1/R                          */
    RT_S_K = 4,          /* conic constant k                              */
    RT_S_NPOLY = 5,      /* N aspheric coefficients                       */
    RT_S_A = 6,          /* a1..a10 at RT_S_A + 0..9 (r^2 .. r^20 terms)  */
    RT_S_CR = 16,        /* toric 1/Rr                                    */
    RT_S_SIZE0 = 17,     /* circle radius | rectangle width               */
    RT_S_SIZE1 = 18,     /* rectangle height                              */
    RT_S_CEN0 = 19, RT_S_CEN1 = 20,
    RT_S_DIST = 21,      /* 21..23 translation dx, dy, dz                 */
    RT_S_ANG = 24,       /* 24..26 rotation about Y, X, Z in radians      */
    RT_S_SIN = 27,       /* 27..29 sin of the three angles (host libm)    */
    RT_S_COS = 30,       /* 30..32 cos of the three angles (host libm)    */
    RT_S_GRTYPE = 33, RT_S_ORDER = 34, RT_S_GRCONST = 35,   /* d in um   */
    RT_S_ORDMODE = 36, RT_S_BLAZEK = 37, RT_S_ORDMIN = 38, RT_S_ORDMAX = 39,
    RT_S_N1KIND = 40, RT_S_N1VAL = 41, RT_S_N2KIND = 42, RT_S_N2VAL = 43,
    /* detector frame (used on the surface of type RT_DETECTOR when XDIM > 0) */
    RT_S_CCD_XDIM = 44, RT_S_CCD_YDIM = 45, RT_S_CCD_XPIX = 46, RT_S_CCD_YPIX = 47,
    RT_S_CCD_XSIZE = 48, RT_S_CCD_YSIZE = 49,
    RT_S_QEKIND = 50, RT_S_QEVAL = 51, RT_S_SPKIND = 52, RT_S_SPVAL = 53,
    RT_SURF_STRIDE = 56
};

/* beam description, RT_BEAM_STRIDE doubles (Beam::SetParams/Create, beam.cpp:67-97,257-485) */
enum rt_beam_type { RT_BEAM_PARALLEL = 0, RT_BEAM_CONIC = 1, RT_BEAM_GAUSS = 2 };
enum rt_range_distr { RT_RANGE_RANDOM = 0, RT_RANGE_UNIFORM = 1, RT_RANGE_MONO = 2 };
enum rt_beam_field {
    RT_B_TYPE = 0, RT_B_SHAPE = 1, RT_B_RANGE_DISTR = 2,
    RT_B_NPER = 3,       /* params[0]: rays (per wavelength for composite beams) */
    RT_B_P1 = 4,         /* params[1]: inner radius | inner width           */
    RT_B_P2 = 5,         /* params[2]: outer radius | outer width           */
    RT_B_P3 = 6,         /* params[3] (rectangle): width/height ratio       */
    RT_B_ANGPAR = 7,     /* cone angle (deg) or Gaussian FWHM (arcsec)      */
    RT_B_CEN = 8,        /* 8..10 beam center                               */
    RT_B_RANGE = 11,     /* 11..12 wavelength range (um)                    */
    RT_B_NLAMBDA = 13,   /* wavelengths of a composite beam                 */
    RT_B_COMPOSITE = 14, /* 1: lambda[r] = grid[r mod NLAMBDA]              */
    RT_B_ANGUNIT = 15,   /* 1: a conic beam's ANGPAR is in radians (plugin API) */
    RT_BEAM_STRIDE = 16
};

/* trace flags */
enum rt_trace_flag {
    RT_TRACE_GENERATE = 1,   /* generate the beam in-kernel from (beam_desc, seed) */
    RT_TRACE_STORE = 2,      /* write the final ray state to X..lambda, flag       */
    RT_TRACE_HIST = 4,       /* accumulate integer hits into ccd_hist_dev          */
    RT_TRACE_FLUX = 8        /* accumulate photon flux into ccd_flux_dev           */
};

/* statistics block: RT_STATS_LEN uint64 counters */
enum rt_stat_field {
    RT_ST_RAYS = 0,          /* rays traced                                       */
    RT_ST_GOOD = 1,          /* rays alive after the last surface                 */
    RT_ST_CCD_HITS = 2,      /* rays counted into the CCD frame                   */
    RT_ST_CCD_OFFCHIP = 3,   /* alive rays whose pixel falls outside the frame    */
    RT_ST_CCD_EDGE = 4,      /* counted rays within 1e-9 px of a pixel edge       */
    RT_ST_SPLINE_OOB = 5,    /* rays killed by a spline argument out of range     */
    RT_ST_ADJ_SKIP = 6,      /* rays the adjoint could not differentiate          */
    RT_ST_SURF = 16,         /* per surface s: 4 counters at RT_ST_SURF + 4*s:    */
    /*   +0 rays alive entering the intersection stage (the ray.surface count)     */
    /*   +1 Newton iterations summed over rays                                    */
    /*   +2 rays lost at intersection, +3 rays lost at mask or action             */
    RT_STATS_LEN = RT_ST_SURF + 4 * RT_MAX_SURF
};

#ifndef __HIPCC_RTC__   /* the entry points are host API; device code needs only the layouts */

/* ---- device / context ------------------------------------------------- */

/* Select `device` for the calling thread and cache its properties. */
int rt_hip_init(int device);

/* Human-readable engine/device description, one item per line
   (replaces rt_engine_info, cuda_engine/rt_cuda_engine_init.cu:14-56). */
int rt_hip_info(char *buf, size_t len);

/* Text of the last error recorded on the calling thread. */
int rt_hip_last_error(char *buf, size_t len);

/* ---- scheme ----------------------------------------------------------- */

typedef struct rt_hip_scheme rt_hip_scheme;

/* Host-side validation of a packed scheme (no device needed): RT_OK or
   RT_ERR_BAD_VALUE with the reason in rt_hip_last_error. */
int rt_hip_validate_scheme(const double *surf_table, int n_surf,
                           const double *spline_pool, size_t pool_len);

/* Validate and upload a packed scheme (host arrays) to device memory. */
int rt_hip_scheme_create(const double *surf_table, int n_surf,
                         const double *spline_pool, size_t pool_len,
                         rt_hip_scheme **out);
int rt_hip_scheme_destroy(rt_hip_scheme *scheme);

/* ---- fused trace (replaces the per-surface loop of Scheme::run,
        common-lib/scheme.cpp:983-1046, one launch for all surfaces) ---- */

/*
 * Trace rays [ray_begin, ray_begin + n_rays) through every surface.
 * With RT_TRACE_GENERATE the beam is generated in-kernel (Philox4x32-10,
 * key = seed, counter = global ray index) and X..flag are outputs;
 * otherwise X..lambda, flag are the input rays and are updated in place
 * when RT_TRACE_STORE is set.  ccd_hist_dev (uint32, YDIM*XDIM) and
 * ccd_flux_dev (FP64, YDIM*XDIM) are accumulated, never cleared.
 * stats_dev (RT_STATS_LEN uint64) is accumulated, never cleared.
 * dump_dev, when not NULL, receives the state after every surface:
 * [n_surf][7][n_rays] doubles (X,Y,Z,cX,cY,cZ,lambda) and
 * dump_flag_dev [n_surf][n_rays].
 */
int rt_hip_trace(const rt_hip_scheme *scheme, const double *beam_desc, uint64_t seed,
                 size_t ray_begin, size_t n_rays,
                 double *X_dev, double *Y_dev, double *Z_dev,
                 double *cX_dev, double *cY_dev, double *cZ_dev,
                 double *lambda_dev, signed char *flag_dev,
                 int flags,
                 unsigned int *ccd_hist_dev, double *ccd_flux_dev,
                 unsigned long long *stats_dev,
                 double *dump_dev, signed char *dump_flag_dev,
                 void *stream);

/* ---- Scheme::run's saved beam and log (common-lib/scheme.cpp:983-1066) -- */

/* Loss codes of a recording trace: 0 for a ray alive after the last surface,
   RT_LOSS_CODE(s, k) for a ray lost at surface s in stage k, RT_LOSS_DEAD
   for a ray that entered dead.  Extents: RT_EXT_PER_SURF doubles per
   surface, min and max of cZ after the transform's normalisation, of X and Y
   after the intersection, of X and Y after the aperture, of cZ after the
   action's normalisation, over the rays alive at that point (NaN where none
   is): what Scheme::run logs (scheme.cpp:993-1042). */
enum rt_loss_stage { RT_LOSS_INTERSECTION = 0, RT_LOSS_APERTURE = 1, RT_LOSS_ACTION = 2 };
#define RT_LOSS_CODE(s, k) (1 + 3 * (s) + (k))
#define RT_LOSS_DEAD 255
enum { RT_EXT_PER_SURF = 12 };

/*
 * rt_hip_trace with RT_TRACE_GENERATE / RT_TRACE_STORE semantics for rays
 * [0, n_rays), plus the loss code of every ray (loss, n bytes, device) and,
 * when `extents` (device, n_surf x RT_EXT_PER_SURF doubles) is not NULL, the
 * logged extents.  No CCD frame (RT_TRACE_HIST / RT_TRACE_FLUX are refused).
 * The extents are combined per workgroup with integer min / max on
 * order-preserving keys (bit-exact whatever the order).  Replaces the
 * stage loop of Scheme::run (scheme.cpp:983-1046) with its log
 * (GetBeamDim / GetBeamCosins, beam.cpp:490-523, over all good rays).
 */
int rt_hip_trace_record(const rt_hip_scheme *scheme, const double *beam_desc, uint64_t seed, size_t n_rays,
                        double *X_dev, double *Y_dev, double *Z_dev, double *cX_dev, double *cY_dev,
                        double *cZ_dev, double *lambda_dev, signed char *flag_dev, int flags,
                        unsigned long long *stats_dev, unsigned char *loss_dev, double *extents_dev,
                        void *stream);

/*
 * The order Beam::Rearrange (common-lib/beam.cpp:574-598) leaves a beam in
 * after every stage of Scheme::run that lost rays (in stage order), from the
 * loss codes of a recording trace: perm_dev (device, n uint64) receives the
 * ray at each position, the good rays first; counts (host, 256 uint64)
 * receives the number of rays of every loss code (counts[0]: good rays).
 * Waits for the stream (the counts drive the stage launches).  Refuses
 * RT_LOSS_DEAD rays.
 */
int rt_hip_rearrange(const unsigned char *loss_dev, size_t n, uint64_t *perm_dev, uint64_t *counts, void *stream);

/*
 * The arrays of Beam::Save (beam.cpp:636-715) in file order: out_dev[k * n +
 * p] = array k (X, Y, Z, cX, cY, cZ, lambda) of ray perm_dev[p].  A lost
 * ray's cosines first go renorm[loss code] times through Beam::Normalize's
 * scaled-norm formula (beam.cpp:541-569, reference operation order, IEEE
 * division and square root, no contraction): the reference normalises every
 * ray, lost ones included, twice per later surface (once for an AUX one).
 * renorm: host, 256 entries.
 */
int rt_hip_beam_order(size_t n, const uint64_t *perm_dev, const unsigned char *loss_dev,
                      const unsigned char *renorm, const double *X_dev, const double *Y_dev,
                      const double *Z_dev, const double *cX_dev, const double *cY_dev, const double *cZ_dev,
                      const double *lambda_dev, double *out_dev, void *stream);

/* ---- per-stage kernels (the plugin API of rt_engine_func.h runs on these) --
   All ray arrays are device SoA buffers of n elements; `rec` is one packed
   surface record on the host; n_bad_dev / iters_dev (nullable) are device
   uint64 counters that are incremented, never cleared. */

/* beam_transform (rt_engine_func.h:41-42; rt_cuda_engine_beam.cu:416-579);
   angles in radians, sin/cos evaluated on the host */
int rt_hip_stage_transform(const double *angles, const double *dist, size_t n,
                           double *X, double *Y, double *Z, double *cX, double *cY, double *cZ,
                           void *stream);
/* Beam::Normalize (common-lib/beam.cpp:527-571) */
int rt_hip_stage_normalize(size_t n, double *cX, double *cY, double *cZ, void *stream);
/* surface_intersection (rt_engine_func.h:46-49; rt_cuda_engine_surface.cu:655-760) */
int rt_hip_stage_intersect(const double *rec, size_t n, double *X, double *Y, double *Z,
                           double *cX, double *cY, double *cZ, signed char *flag,
                           unsigned long long *n_bad_dev, unsigned long long *iters_dev, void *stream);
/* Surface::ApplyConstrains / Shade::ApplyConstrains (surface.cpp:264-321,652-699) */
int rt_hip_stage_mask(const double *rec, size_t n, double *X, double *Y, signed char *flag,
                      unsigned long long *n_bad_dev, void *stream);
/* surface_reflection (rt_engine_func.h:52-54; rt_cuda_engine_surface.cu:883-951) */
int rt_hip_stage_reflect(const double *rec, size_t n, double *X, double *Y, double *Z,
                         double *cX, double *cY, double *cZ, void *stream);
/* surface_refraction (rt_engine_func.h:57-61): mu = n1(lambda)/n2(lambda) per ray */
int rt_hip_stage_refract(const double *rec, size_t n, double *X, double *Y, double *Z,
                         double *cX, double *cY, double *cZ, const double *mu, signed char *flag,
                         unsigned long long *n_bad_dev, void *stream);
/* surface_diffration (rt_engine_func.h:64-70): L = m lambda / n2(lambda) / d per ray,
   mu per ray for transmissive gratings (NULL for reflective ones) */
int rt_hip_stage_diffract(const double *rec, size_t n, double *X, double *Y, double *Z,
                          double *cX, double *cY, double *cZ, const double *mu, const double *L,
                          signed char *flag, unsigned long long *n_bad_dev, void *stream);

/* beam_random (rt_engine_func.h:30-31): positions of rays [begin, begin+n) */
int rt_hip_beam_positions(const double *beam_desc, uint64_t seed, size_t begin, size_t n,
                          double *X, double *Y, double *Z, void *stream);
/* beam_conic_cosins / beam_gauss_cosins (rt_engine_func.h:33,35) by beam type;
   X, Y are read for conic beams */
int rt_hip_beam_directions(const double *beam_desc, uint64_t seed, size_t begin, size_t n,
                           const double *X, const double *Y, double *cX, double *cY, double *cZ,
                           void *stream);
/* uniform_range_beam / random_range_beam (rt_engine_func.h:37,39) by range distribution */
int rt_hip_beam_wavelengths(const double *beam_desc, uint64_t seed, size_t begin, size_t n,
                            double *lambda, void *stream);

/* ccd_coordinates (rt_engine_func.h:96-99; rt_cuda_engine_ccd.cu:96-214) */
int rt_hip_ccd_coordinates(size_t n, const double *X, const double *Y, double xsize, double ysize,
                           double xpix, double ypix, unsigned *ccdX, unsigned *ccdY, void *stream);
/* compute_ccd_image accumulation (rt_engine_func.h:101-104; ccd.cu:219-232):
   image[y*xdim+x] += w (w NULL: 1) and/or hist += 1, bounds-checked, the
   off-chip rays counted */
int rt_hip_ccd_accumulate(size_t n, const unsigned *ccdX, const unsigned *ccdY, const double *w,
                          size_t xdim, size_t ydim, double *image, unsigned *hist,
                          unsigned long long *offchip_dev, void *stream);
/* surface_QE multiply (rt_engine_func.h:73; rt_cuda_engine_surface.cu:1775-1784): v *= f */
int rt_hip_scale(size_t n, double *v, const double *f, void *stream);

/* ccd_frame's work per result file (ccd/ccd_frame.cpp:190-195) in one
   kernel: each ray's pixel (ccd_coordinates, rt_cuda_engine_ccd.cu:30-45),
   its weight spec(lambda) lambda / hc QE(lambda) xpix ypix 1e-2 (ccd.cu:
   286-287,360; QE and spectrum as host splines [K, x, y, y2], or NULL for
   the constants) added into image_dev (ydim x xdim FP64, device) with atomics
   and a bounds check (compute_ccd_image, ccd.cu:219-232).  counts_dev
   (device, 3 uint64, added to, nullable): rays added, rays off the frame,
   rays whose lambda is outside a table (the reference's bad_tab_func; their
   weight is not added).  Waits for the stream. */
int rt_hip_ccd_flux(size_t n, const double *X_dev, const double *Y_dev, const double *lambda_dev,
                    const double *qe_spline, double qe_const, const double *spec_spline, double spec_const,
                    double xsize, double ysize, double xpix, double ypix, size_t xdim, size_t ydim,
                    double *image_dev, unsigned long long *counts_dev, void *stream);
/* grating_energy_distr (rt_engine_func.h:75-78; rt_cuda_engine_surface.cu:2098-2290) */
int rt_hip_grating_energy(size_t n, const double *lambda, long order, double blaze, double alpha,
                          double gamma, double gr_const, double rule_start, double rule_stop,
                          double *energy, void *stream);
/* spline_init (rt_engine_func.h:91): host, serial natural-spline recurrence */
int rt_hip_spline_init(size_t n, const double *x, const double *y, double *y2);
/* TabulatedFunction::operator[] (tabulated_function.cpp:173-204) at n device
   points, spline [K, x, y, y2] given on the host and staged in LDS; points out
   of range give 0 and are counted in *oob_dev */
int rt_hip_spline_eval(const double *spline, size_t n, const double *x, double *y,
                       unsigned long long *oob_dev, void *stream);

/* ---- generated kernels ------------------------------------------------ */

/* The trace launches a kernel generated for the scheme's structure (surface
   classes, types, shapes, numbers of aspheric terms) and compiled once per
   structure with hiprtc; mode 0 forces the generic kernel that reads the
   structure from the table at run time.  Returns the previous mode.  The
   environment variable RT_HIP_JIT=0 sets mode 0 at start-up. */
int rt_hip_set_jit(int mode);

/* For beams of one wavelength or a composite uniform grid (at most 16384
   wavelengths), the generated traces read per-wavelength values from a table
   formed by a small kernel with the same formulas, spline code and wavelength
   formula as the per-ray evaluation (the results are the same bit for bit):
   for one scheme n1/n2 of each lens and grating, the grating term and the
   CCD flux weight (refilled when the wavelength grid or the stream changes,
   and on every trace once one was captured in a graph); for a batch its
   spline-valued functions (refractive indices, detector QE and spectrum),
   filled before each trace.  Mode 0 evaluates everything per ray.  Returns
   the previous mode.  The environment variable RT_HIP_LTAB=0 sets mode 0 at
   start-up.  The table is per scheme (or batch): traces of one scheme with
   different wavelength grids must not run on two streams at once.  No
   counterpart in the reference. */
int rt_hip_set_ltab(int mode);

/* Launch of a scheme's generated kernel: lanes per workgroup (64..1024,
   powers of two), the minimum number of waves per SIMD the compiler must fit
   its registers to (0: its own choice), and persistent workgroups per CU as a
   multiple of those that are resident at once.  Default 256, 7, 16 (for
   schemes and batches).  No counterpart in the reference (its launches are
   fixed per kernel). */
int rt_hip_scheme_set_launch(rt_hip_scheme *scheme, int block, int waves, int grid_mult);
int rt_hip_scheme_get_launch(const rt_hip_scheme *scheme, int *block, int *waves, int *grid_mult);

/* Surface unroll depth of a scheme's traces: full != 0 (default) traces with
   the kernel generated for the scheme's structure, the whole surface chain
   unrolled into straight-line code; full == 0 with the generic kernel, a
   runtime loop over the surface records (smaller code, for long schemes).
   The autotuner times both.  No counterpart in the reference. */
int rt_hip_scheme_set_unroll(rt_hip_scheme *scheme, int full);
int rt_hip_scheme_get_unroll(const rt_hip_scheme *scheme, int *full);

/* Autotuner: traces n_sample rays of the beam (generated from seed, seed + 1,
   ... for successive launches, with `flags`) under each candidate launch
   (block 128/256, waves 0/5/6/7/8, grid multiple 1/2/4/8/16 of the fully
   unrolled generated kernel, and the runtime surface loop of the generic
   kernel; those spilling more than 64 bytes per lane to scratch are set
   aside when others exist) in
   3 interleaved rounds of 3 back-to-back launches, re-times the four fastest
   in 5 rounds of 20 (a sustained load runs at a lower clock than isolated
   launches, by an amount that depends on the launch), keeps the one with the
   lowest median (within 0.5%, the one with less scratch memory per lane) in
   the scheme and returns that median (per launch) in *best_ms.
   RT_HIP_AUTOTUNE_LOG=1 prints every candidate's times on stderr.  Allocates
   its sample buffers; call it once after rt_hip_scheme_create, outside any
   timed or captured region. */
int rt_hip_autotune(rt_hip_scheme *scheme, const double *beam_desc, uint64_t seed, size_t n_sample,
                    int flags, double *best_ms, void *stream);

/* ---- batched traces ---------------------------------------------------- */

/* Many instances of one scheme structure (same classes, types, shapes,
   polynomial counts, tables; any parameter values, e.g. Monte-Carlo
   tolerance perturbations) with uneven ray counts, traced in one persistent
   launch.  tables: n_inst packed tables of n_surf records; one spline pool
   for all; ray_counts[n_inst].  Instance i's beam is generated from
   seed + i, rays 0 .. ray_counts[i]-1, exactly as rt_hip_trace would for
   that instance alone.  No counterpart in the reference, which traces one
   scheme per run (common-lib/scheme.cpp:878-1099). */
typedef struct rt_hip_batch rt_hip_batch;
enum { RT_SPOT_STRIDE = 6 };   /* per instance: n, centroid x, y, variance x, y, RMS radius */

int rt_hip_batch_create(const double *tables, int n_inst, int n_surf, const double *spline_pool,
                        size_t pool_len, const uint64_t *ray_counts, rt_hip_batch **out);
/* new parameter values for every instance (same structure and pool layout) */
int rt_hip_batch_update(rt_hip_batch *batch, const double *tables, const double *spline_pool,
                        size_t pool_len);
int rt_hip_batch_destroy(rt_hip_batch *batch);

/* A beam per instance: beam_descs holds n_inst beam descriptions
   (RT_BEAM_STRIDE doubles each, one beam type for all: the generated kernel
   is specialised on it), e.g. a wavelength set per instance; instance i's
   rays are generated from its own description (and seed + i) by the batched
   trace and the batched adjoint, whose beam_desc argument must then have the
   same type.  NULL returns to one beam for every instance (a graph captured
   before keeps reading the beams it was captured with).  Such batches
   evaluate their splines per ray (no per-wavelength table).  No counterpart
   in the reference. */
int rt_hip_batch_set_beams(rt_hip_batch *batch, const double *beam_descs);
int rt_hip_batch_info(const rt_hip_batch *batch, uint64_t *total_rays, uint64_t *n_tiles);
int rt_hip_batch_set_launch(rt_hip_batch *batch, int block, int waves, int grid_mult);
/* Trace every instance (flags: RT_TRACE_STORE optional; CCD frames are not
   formed in batches).  With STORE the rays go to X..flag concatenated by
   instance.  spots (device, n_inst x RT_SPOT_STRIDE) receives each
   instance's spot on the last surface from its surviving rays, combined in a
   fixed order (bit-reproducible).  stats is accumulated over all instances. */
int rt_hip_batch_trace(const rt_hip_batch *batch, const double *beam_desc, uint64_t seed, int flags,
                       double *X, double *Y, double *Z, double *cX, double *cY, double *cZ,
                       double *lambda, signed char *flag, double *spots,
                       unsigned long long *stats, void *stream);

/* The same batched trace recorded as a hipGraph on `stream` (not the null
   stream) and replayed with rt_hip_graph_launch; table contents may change
   between replays (rt_hip_batch_update).  The graph holds no memset nodes:
   stats keep accumulating over replays. */
typedef struct rt_hip_graph rt_hip_graph;
int rt_hip_batch_graph_create(const rt_hip_batch *batch, const double *beam_desc, uint64_t seed, int flags,
                              double *X, double *Y, double *Z, double *cX, double *cY, double *cZ,
                              double *lambda, signed char *flag, double *spots,
                              unsigned long long *stats, void *stream, rt_hip_graph **out);
int rt_hip_graph_launch(rt_hip_graph *graph, void *stream);
int rt_hip_graph_destroy(rt_hip_graph *graph);

/* ---- adjoint ------------------------------------------------------------ */

/* Gradient of a loss on the surviving rays' final positions with respect to
   scheme parameters.  The beam (beam_desc, seed, rays 0 .. n_rays-1) is traced
   forward and the surface chain reversed per ray, recomputing each surface
   from its entry state.  The loss enters through group_seeds (device,
   n_groups x 3): for a ray r alive at the end, of group g = r mod n_groups,
   with final position (X, Y), dLoss/dX = ax_g + b_g X and dLoss/dY =
   ay_g + b_g Y (centroids, RMS radii and squared deviations of them are of
   this form).  Parameters: param_surface[j], param_field[j] (host) name
   RT_S_C, RT_S_K, RT_S_A + i, RT_S_CR, RT_S_DIST + j, RT_S_ANG + j (radians,
   the record's value) or RT_S_GRCONST (um) of a surface; grad (device,
   n_params) receives dLoss/dparameter, reduced in a fixed order.  Plane
   gratings only.  No counterpart in the reference (it has no derivatives). */
int rt_hip_adjoint(const rt_hip_scheme *scheme, const double *beam_desc, uint64_t seed, size_t n_rays,
                   const double *group_seeds, size_t n_groups, const int *param_surface,
                   const int *param_field, int n_params, double *grad, unsigned long long *stats,
                   void *stream);

/* Spots of a traced beam per ray group (ray r belongs to group r mod
   n_groups): spots (device, n_groups x RT_SPOT_STRIDE) receives, from the
   rays with flag 1, n, centroid x, y, variance x, y and RMS radius
   sqrt(var x + var y), summed in a fixed order (bit-reproducible).  X, Y,
   flag: device arrays of n rays.  No counterpart in the reference; it feeds
   the spot loss of the adjoint. */
int rt_hip_group_spots(const double *X, const double *Y, const signed char *flag, size_t n, size_t n_groups,
                       double *spots, void *stream);

/* The generated beam of `seed` (rays 0 .. n_rays - 1 of beam_desc) traced
   through the scheme with the group spots of the rays alive at the end (ray
   r in group r mod n_groups) formed in the trace itself: each workgroup tile
   of rays adds its per-group moments, merged in a fixed order afterwards.
   spots (device, n_groups x RT_SPOT_STRIDE) receives what
   rt_hip_trace(RT_TRACE_GENERATE | RT_TRACE_STORE) followed by
   rt_hip_group_spots gives, to rounding (another summation order), without
   storing the rays; stats as for rt_hip_trace.  No counterpart in the
   reference; it is the loss forward of the adjoint. */
int rt_hip_trace_spots(const rt_hip_scheme *scheme, const double *beam_desc, uint64_t seed, size_t n_rays,
                       size_t n_groups, double *spots, unsigned long long *stats, void *stream);

/* rt_hip_group_spots without any allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The partial moments go to a static 8 MiB device array per GPU
   (2^20 doubles, 5 per (chunk of 8 rays of a group, group)): inputs needing
   more, roughly n > 1.6e6 rays, are rejected with RT_ERR_BAD_VALUE.  Same
   kernels and summation order as rt_hip_group_spots, so the spots are equal
   bit for bit.  Calls on one GPU share the array: calls that overlap on one
   GPU (different streams) race.  rt_hip_init looks the array up for its
   device; otherwise the first call on a device does. */
int rt_hip_group_spots_noalloc(const double *X, const double *Y, const signed char *flag, size_t n,
                               size_t n_groups, double *spots, void *stream);

/* The spot loss over groups with live rays, sum of rms + weight x squared
   centroid offset from target (device, n_groups x 2, or NULL for 0), into
   *loss (device), and the adjoint seeds of rt_hip_adjoint (device, n_groups
   x 3: ax, ay, b with dloss/dX_r = ax + b X_r, dloss/dY_r = ay + b Y_r),
   from rt_hip_group_spots' output; fixed order. */
int rt_hip_spot_loss(const double *spots, size_t n_groups, const double *target, double weight, double *seeds,
                     double *loss, void *stream);

/* The adjoint over every instance of a batch (rt_hip_batch_create), each
   with its own beam (seed + i, rays 0 .. ray_counts[i]-1) and group seeds
   (device, n_inst x n_groups x 3: group g of instance i's ray r is r mod
   n_groups); grad (device, n_inst x n_params) receives each instance's
   gradient with respect to its own parameters, reduced in a fixed order.
   The batch must use workgroups of 256 lanes.  No counterpart in the
   reference. */
int rt_hip_batch_adjoint(const rt_hip_batch *batch, const double *beam_desc, uint64_t seed,
                         const double *group_seeds, size_t n_groups, const int *param_surface,
                         const int *param_field, int n_params, double *grad, unsigned long long *stats,
                         void *stream);

/* Spots per (instance, group) of a batch's stored rays (rt_hip_batch_trace
   with RT_TRACE_STORE, rays concatenated by instance; instance i's ray r is in
   group r mod n_groups): spots (device) receives n_inst x n_groups x
   RT_SPOT_STRIDE, with the summation order of rt_hip_group_spots on each
   instance's rays alone. */
int rt_hip_batch_group_spots(const rt_hip_batch *batch, const double *X, const double *Y,
                             const signed char *flag, size_t n_groups, double *spots, void *stream);

/* The batched trace of generated beams (rt_hip_batch_trace with
   RT_TRACE_GENERATE) with the spots per (instance, group) of the rays alive
   at the end formed in the trace itself (instance i's ray r in group
   r mod n_groups): spots (device, n_inst x n_groups x RT_SPOT_STRIDE)
   receives what RT_TRACE_STORE followed by rt_hip_batch_group_spots gives,
   to rounding, without storing the rays.  No counterpart in the reference. */
int rt_hip_batch_trace_group_spots(const rt_hip_batch *batch, const double *beam_desc, uint64_t seed,
                                   size_t n_groups, double *spots, unsigned long long *stats, void *stream);

/* One gradient-descent step on the device tables of every instance of a
   batch: field -= lr[j] * grad[i][j] for the parameters (param_surface,
   param_field) (grad: device, n_inst x n_params, as rt_hip_batch_adjoint
   returns it; lr: device, n_params), with the derived record constants and
   the sin/cos of stepped tilts re-formed on the device.  Steppable fields:
   those of the adjoint; a tilt or decentre must be non-zero on that surface
   in some instance of the batch.  The host copy of the tables is left as it
   was (rt_hip_batch_update replaces both).  After its first call with a
   parameter list it only launches a kernel (capturable in a hipGraph).  No
   counterpart in the reference. */
int rt_hip_batch_descend(rt_hip_batch *batch, const int *param_surface, const int *param_field, int n_params,
                         const double *grad, const double *lr, void *stream);

/* "generated" or "generic": the kernel the calling thread's last trace used. */
int rt_hip_last_kernel(char *buf, size_t len);

/* Generate and compile (without loading, no device needed) the kernel of a
   packed scheme for beam mode (-1: rays given, else RT_BEAM_*, or'ed with
   RT_JIT_BATCH for the batched kernel, or with RT_JIT_ADJOINT for the adjoint
   kernel with every slot of every surface asked for, or with RT_JIT_LTAB for
   the trace that reads its spline-valued functions from a per-wavelength
   table); `text` receives the generated source
   and the compiler log, *code_size the code object size.  RT_OK when it
   compiles. */
enum { RT_JIT_BATCH = 0x100, RT_JIT_ADJOINT = 0x200, RT_JIT_LTAB = 0x400, RT_JIT_RECORD = 0x800 };
int rt_hip_jit_check(const double *surf_table, int n_surf, const double *spline_pool,
                     size_t pool_len, int beam_mode, char *text, size_t len, size_t *code_size);

#endif /* __HIPCC_RTC__ */

#ifdef __cplusplus
}
#endif

#endif /* RT_HIP_H */
