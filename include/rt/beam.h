/* beam.h -- the caller-facing part of common-lib/beam.h for code that runs
 * schemes through rt_amd's Scheme (include/rt/scheme.h) and reads the saved
 * beams: the Beam enumerations (names and values of beam.h:31-36), the
 * Beam::bad_beam exception (beam.h:79-87) the Scheme path throws, and the
 * .dat reader load_beam_data (beam.h:18-19, beam.cpp:829-1036).
 *
 * The include guard is common-lib's own: in a build that uses rt_amd's
 * Scheme this header takes the place of common-lib/beam.h.  The ray state of
 * a beam lives in HBM and is handled by the engine, so the Beam class itself
 * carries only the types its callers name.  Definitions: librt_scheme_hip.so.
 */
#ifndef BEAM_H
#define BEAM_H

#include "num_defs.h"

#include <stddef.h>

#include <exception>

/* Read a beam saved by Scheme::run (Beam::Save layout, beam.cpp:636-715):
 * three size_t (total rays, good rays, wavelengths), then X, Y, Z, cX, cY,
 * cZ, lambda over the good rays, then the same seven arrays over the lost
 * rays.  Each non-NULL pointer receives a malloc'ed array of the good rays
 * (all rays when `all`); NULL pointers skip that array.  Returns 0, or the
 * reference's codes: 1 open failure, 2 short file, 3 allocation failure. */
int load_beam_data(const char *rt_filename, size_t *total_Nrays, size_t *N_good, size_t *N_lambda,
                   real_t **X, real_t **Y, real_t **Z, real_t **cX, real_t **cY, real_t **cZ,
                   real_t **lambda, bool all = false);

class Beam {
public:
    enum BeamType {Parallel, Conic, Gauss};
    enum BeamShape {Circle, Rectangle};
    enum BeamProfile {Random, Concentric};
    enum BeamRangeDistr {RandomDistr, UniformDistr, Monochromatic};

    enum BeamError {MemAllocFailure, InvalidParams, SaveFailure, CreationFailure, TransformationFailure};

    class bad_beam : public std::exception {
    public:
        bad_beam(BeamError err);
        BeamError beam_error() const;
    private:
        BeamError Error;
    };
};

#endif /* BEAM_H */
