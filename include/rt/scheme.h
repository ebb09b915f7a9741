/* scheme.h -- the Scheme class of common-lib/scheme.h (public interface of
 * scheme.h:21-66: the same members, signatures, enumerations and exception)
 * run on the MI355X engine.
 *
 * load_from_file reads the reference's XML scheme format without mxml
 * (rt_amd's own reader); run traces every wavelength range of the beam in one
 * fused device launch per range (beam generation, every surface's transform,
 * normalisation, intersection, aperture and action, the reference's
 * Rearrange order and the logged extents), then writes the result beams in
 * Beam::Save's .dat layout; ComputeQE writes the efficiency file of
 * scheme.cpp:1135-1338.  The include guard is common-lib's own, so this
 * header takes the place of common-lib/scheme.h for callers such as
 * ray-tracing/main.cpp.  Definitions: librt_scheme_hip.so.
 */
#ifndef SCHEME_H
#define SCHEME_H

#include "num_defs.h"
#include "beam.h"
#include "surface.h"
#include "tabulated_function.h"

#include <stddef.h>
#include <stdint.h>

#include <exception>
#include <fstream>
#include <iostream>
#include <list>
#include <memory>
#include <string>
#include <vector>

#define ATTR_NAME_MAXLEN 1024

class Scheme;
namespace rt_scheme_detail {
struct State;   /* the loaded scheme: surfaces, beam, ranges, XML tree */
int pack_for_tests(const Scheme &, double *, int, int *, double *, size_t, size_t *, double *, uint64_t *);
}

class Scheme {
public:
    enum SchemeError {FileOpenFailure, ParseFailure, NoClassAttr, BadAttrValue, NanAttrValue, MemAllocFailure,
                      NoSchemeDesc, NoBeamDesc, NoSurfaceDesc, NoRange, NoRangeFile, BadRange,
                      EmptyScheme, BadStartIndex, FileSaveFailure};

    Scheme();
    Scheme(const std::string &filename);
    ~Scheme();

    /* QE_flag: keep the XML tree for ComputeQE only (scheme.cpp:361-363) */
    void load_from_file(const std::string &filename, bool QE_flag = false);
    void destroy();
    void run(std::list<std::string> &resultfile_list, std::ostream &log_file = std::cout,
             const size_t start_surface = 0);

    size_t GetNumberOfSurfaces() const;

    void ComputeQE(size_t Nlambda, char *QE_filename);

    class bad_scheme : public std::exception {
    public:
        bad_scheme(SchemeError err);
        SchemeError scheme_error() const;
    private:
        SchemeError Error;
    };

private:
    friend int rt_scheme_detail::pack_for_tests(const Scheme &, double *, int, int *, double *, size_t, size_t *,
                                                double *, uint64_t *);
    Scheme(const Scheme &) = delete;
    Scheme &operator=(const Scheme &) = delete;
    std::unique_ptr<rt_scheme_detail::State> st;
};

/* C entry points over Scheme for callers outside C++ (ctypes, cgo ...). */
extern "C" {
/* ray-tracing/main.cpp's flow: load the scheme, run it (log to log_file, or
   stdout when NULL), write the result file names, one per line, to
   rt_files_list.  Returns 0, or 1 with the reason on stderr, as main.cpp. */
int rt_scheme_run_file(const char *scheme_file, const char *rt_files_list, const char *log_file);
/* efficiency/efficiency.cpp's flow (qe_comp): Scheme::ComputeQE with
   nlambda wavelengths per range into qe_file.  0, or the SchemeError + 1. */
int rt_scheme_compute_qe(const char *scheme_file, size_t nlambda, const char *qe_file);
/* load_from_file alone, with what run would hand the engine for the first
   wavelength range: the packed surface table (n_surf x RT_SURF_STRIDE
   doubles into table, room for max_surf surfaces), the spline pool (room for
   max_pool doubles) and the beam description (RT_BEAM_STRIDE doubles); no
   device is used.  Returns 0, or the exception load_from_file threw:
   100 + Scheme::SchemeError, 200 + Beam::BeamError, 300 +
   Surface::SurfaceError, 400 + Lens::LensError, 500 +
   Grating::GratingError (700 + it when thrown as a bare enumerator),
   600 + TabulatedFunction::TabulatedFunctionError; 900 if an output does
   not fit. */
int rt_scheme_pack_file(const char *scheme_file, double *table, int max_surf, int *n_surf, double *pool,
                        size_t max_pool, size_t *pool_len, double *beam_desc, uint64_t *n_rays);
}

#endif /* SCHEME_H */
