// Host build of rt_amd/csrc/rt_adjoint.h for the CPU tests: the surface
// adjoint is checked against central finite differences of the oracle.
#include "rt_adjoint.h"

extern "C" int adj_host_pg_n() { return rt::adj::PG_N; }
extern "C" int adj_host_slot(int field) { return rt::adj::slot_of_field(field); }

// one surface backwards: entry[6] state on entry, a6[6] adjoint after (in) /
// at entry (out), gp[PG_N] parameter derivatives (accumulated).  0: ok
extern "C" int adj_host_surface(const double* rec, const double* entry, int n_fwd, double mu, double L,
                                double* a6, double* gp) {
    rt::adj::St e = {entry[0], entry[1], entry[2], entry[3], entry[4], entry[5]};
    rt::adj::St a = {a6[0], a6[1], a6[2], a6[3], a6[4], a6[5]};
    bool ok = rt::adj::surface_adjoint(rt::adj::DynRec{rec}, e, n_fwd, mu, L, a, gp);
    a6[0] = a.X; a6[1] = a.Y; a6[2] = a.Z; a6[3] = a.cX; a6[4] = a.cY; a6[5] = a.cZ;
    return ok ? 0 : 1;
}
