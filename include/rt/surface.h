/* surface.h -- the surface enumerations the plugin API takes and the
 * surface exception types of the Scheme path (names and values of
 * common-lib/surface.h:25-32,62-70,136-137,163-176 and :117-126).
 *
 * The include guard is common-lib's own: with rt_amd's Scheme this header
 * takes the place of common-lib/surface.h for the callers (the surface
 * classes themselves are packed into the engine's surface table).  The
 * plugin library needs only the enumerations; the exceptions' definitions
 * are in librt_scheme_hip.so. */
#ifndef SURFACE_H
#define SURFACE_H

#include "num_defs.h"
#include "beam.h"
#include "tabulated_function.h"

#include <exception>

class Surface {
public:
    enum SurfaceClass {Plane, Conic, Toric, AuxClass};
    enum SurfaceType {Mirror, Lens, Grating, Stop, Shade, Detector, AuxType};
    enum SurfaceShape {Circle, Rectangle};

    enum SurfaceError {Ok, InvalidSurfaceParam, InvalidSurfaceSize, InvalidSurfaceCenter,
                       InvalidSurfaceDistance, InvalidSurfaceAngles,
                       MemAllocFailure, NoIntersections, IntersectionFailure,
                       ActionFailure, NoGoodRays, BadQE};

    class bad_surface : public std::exception {
    public:
        bad_surface(SurfaceError err);
        SurfaceError surface_error() const;
    private:
        SurfaceError Error;
    };
};

class Lens : public Surface {
public:
    enum LensError {BadRefIndex, BadRefIndexFile};

    class bad_lens : public std::exception {
    public:
        bad_lens(Lens::LensError err);
        Lens::LensError bad_lens_error() const;
    private:
        Lens::LensError Error;
    };
};

class Grating : public Surface {
public:
    enum GratingType {Reflective, Transparent};
    enum GratingError {BadRefIndex, BadRefIndexFile, BadOrder, BadType, BadRuleParams};

    class bad_grating : public std::exception {
    public:
        bad_grating(Grating::GratingError err);
        Grating::GratingError bad_grating_error() const;
    private:
        Grating::GratingError Error;
    };
};

#endif /* SURFACE_H */
