/* surface.h -- the surface and beam enumerations the plugin API takes
 * (names and values of common-lib/beam.h:31-34 and surface.h:25-27,136).
 * Only the nested enumerations are part of the engine ABI; the full Beam,
 * Surface and Grating classes belong to the caller's common library. */
#ifndef RT_PLUGIN_TYPES_H
#define RT_PLUGIN_TYPES_H

#include "num_defs.h"
#include "tabulated_function.h"

#if !defined(BEAM_H)
class Beam {
public:
    enum BeamType {Parallel, Conic, Gauss};
    enum BeamShape {Circle, Rectangle};
    enum BeamProfile {Random, Concentric};
    enum BeamRangeDistr {RandomDistr, UniformDistr, Monochromatic};
};
#endif

#if !defined(SURFACE_H)
class Surface {
public:
    enum SurfaceClass {Plane, Conic, Toric, AuxClass};
    enum SurfaceType {Mirror, Lens, Grating, Stop, Shade, Detector, AuxType};
    enum SurfaceShape {Circle, Rectangle};
};

class Grating : public Surface {
public:
    enum GratingType {Reflective, Transparent};
};
#endif

#endif /* RT_PLUGIN_TYPES_H */
