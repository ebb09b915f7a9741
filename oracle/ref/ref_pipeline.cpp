// ref_pipeline.cpp -- TEST INFRASTRUCTURE.  Drives the reference's own host
// code (common-lib beam.cpp, surface.cpp, tabulated_function.cpp,
// ascii_file.cpp, compiled unchanged from the reference tree) through an
// engine plugin library linked at build time: librt_engine_hip.so (the
// product) or librt_engine_oracle.so (the oracle).  It replays the per-surface
// loop of Scheme::run (common-lib/scheme.cpp:983-1046) on a scheme read from
// a plain text description (written by tests/refpipe.py; the XML loader of
// scheme.cpp needs mxml, absent here), saves the beam with Beam::Save, and
// exercises the CCD and efficiency entry points the way ccd_frame.cpp:184-195
// and Surface/Grating::ApplyQE (surface.cpp:335-347,614-640) call them.
//
// usage: ref_pipeline <scheme.txt> <out prefix>
//   writes <prefix>.dat (Beam::Save), <prefix>.ccd (FP64 image, if the last
//   surface is a detector with a CCD frame), <prefix>.qe (efficiency of every
//   grating and of the detector on the beam's wavelength grid), and prints
//   "info", "trace_seconds" (wall time of Beam::Create and the surface loop)
//   and "ngood" lines.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <fstream>
#include <stdexcept>
#include <iostream>
#include <list>
#include <sstream>
#include <string>
#include <vector>

#include "beam.h"
#include "rt_engine_errors.h"
#include "rt_engine_func.h"
#include "surface.h"
#include "tabulated_function.h"

using namespace std;

namespace {

struct Fn {                 // "v:<value>" or "f:<file>"
    bool file = false;
    double value = 1.0;
    string path;
};

Fn parse_fn(const string& s) {
    Fn f;
    if (s.compare(0, 2, "f:") == 0) {
        f.file = true;
        f.path = s.substr(2);
    } else {
        f.value = atof(s.c_str() + 2);
    }
    return f;
}

struct SurfDesc {
    int cls, type, shape;
    vector<real_t> params, size, center, dist, ang;
    Fn n1, n2, qe;
    int grtype = 0;
    long order = 1;
    double grconst = 1.0;
    double blaze = 0.0, alpha = 0.0, gamma = 0.0;   // rule parameters (degrees) for ApplyQE
    // detector frame
    size_t xdim = 0, ydim = 0;
    double xpix = 0.0, ypix = 0.0, xsize = 0.0, ysize = 0.0;
    Fn spec;
};

vector<real_t> read_vec(istringstream& is, size_t n) {
    vector<real_t> v(n);
    for (size_t i = 0; i < n; ++i) is >> v[i];
    return v;
}

Surface* make_surface(const SurfDesc& d) {
    Surface::SurfaceClass cl = (Surface::SurfaceClass)d.cls;
    Surface::SurfaceShape sh = (Surface::SurfaceShape)d.shape;
    Surface* s = nullptr;
    switch ((Surface::SurfaceType)d.type) {
        case Surface::Mirror: s = new Mirror(cl, sh); break;
        case Surface::Lens:
            if (!d.n1.file && !d.n2.file) s = new Lens(cl, sh, d.n1.value, d.n2.value);
            else if (!d.n1.file) s = new Lens(cl, sh, d.n1.value, d.n2.path.c_str());
            else if (!d.n2.file) s = new Lens(cl, sh, d.n1.path.c_str(), d.n2.value);
            else s = new Lens(cl, sh, d.n1.path.c_str(), d.n2.path.c_str());
            break;
        case Surface::Grating: {
            vector<long> ord(1, d.order);
            Grating::GratingType gt = (Grating::GratingType)d.grtype;
            Grating* g;
            if (!d.n1.file && !d.n2.file) g = new Grating(cl, sh, gt, ord, d.grconst, d.n1.value, d.n2.value);
            else if (!d.n1.file) g = new Grating(cl, sh, gt, ord, d.grconst, d.n1.value, d.n2.path.c_str());
            else if (!d.n2.file) g = new Grating(cl, sh, gt, ord, d.grconst, d.n1.path.c_str(), d.n2.value);
            else g = new Grating(cl, sh, gt, ord, d.grconst, d.n1.path.c_str(), d.n2.path.c_str());
            g->SetRuleParams(d.blaze, d.alpha, d.gamma);
            s = g;
            break;
        }
        case Surface::Stop: s = new Stop(cl, sh); break;
        case Surface::Shade: s = new Shade(cl, sh); break;
        case Surface::Detector: s = new Surface(cl, Surface::Detector, sh); break;
        default: s = new Aux(); return s;
    }
    s->SetParams(d.params);
    s->SetGeometry(d.size, d.center);
    s->SetDistance(d.dist);
    s->SetAngles(d.ang);
    if (d.type == Surface::Detector && d.qe.file) s->SetQE(d.qe.path.c_str());
    return s;
}

// Beam::Save layout (beam.cpp:636-715): Nrays, N_good, N_lambda, then the 7
// arrays of the good rays, then those of the others
bool read_dat(const string& path, size_t& n, size_t& ngood, vector<vector<double>>& a) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    size_t hdr[3];
    if (fread(hdr, sizeof(size_t), 3, f) != 3) { fclose(f); return false; }
    n = hdr[0];
    ngood = hdr[1];
    a.assign(7, vector<double>(n));
    for (int k = 0; k < 7; ++k)
        if (ngood && fread(a[k].data(), sizeof(double), ngood, f) != ngood) { fclose(f); return false; }
    fclose(f);
    return true;
}

TabulatedFunction make_fn(const Fn& f) {
    return f.file ? TabulatedFunction(f.path.c_str()) : TabulatedFunction(f.value);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s scheme.txt out_prefix\n", argv[0]);
        return 2;
    }
    ifstream in(argv[1]);
    string line, prefix = argv[2];
    Beam* beam = nullptr;
    vector<real_t> bparams, bcenter, brange;
    bool composite = false;
    size_t nlambda = 1;
    vector<SurfDesc> descs;
    try {
        while (getline(in, line)) {
            istringstream is(line);
            string kw;
            is >> kw;
            if (kw == "beam") {
                int type, shape, distr, comp, np;
                is >> type >> shape >> distr >> comp >> nlambda >> np;
                composite = comp != 0;
                bparams = read_vec(is, np);
                bcenter = read_vec(is, 3);
                brange = read_vec(is, 2);
                beam = new Beam((Beam::BeamType)type, (Beam::BeamShape)shape, Beam::Random,
                                (Beam::BeamRangeDistr)distr);
            } else if (kw == "surface") {
                SurfDesc d;
                int np;
                string n1, n2, qe, sp;
                is >> d.cls >> d.type >> d.shape >> np;
                d.params = read_vec(is, np);
                d.size = read_vec(is, 2);
                d.center = read_vec(is, 2);
                d.dist = read_vec(is, 3);
                d.ang = read_vec(is, 3);
                is >> n1 >> n2 >> d.grtype >> d.order >> d.grconst >> d.blaze >> d.alpha >> d.gamma;
                is >> d.xdim >> d.ydim >> d.xpix >> d.ypix >> d.xsize >> d.ysize >> qe >> sp;
                d.n1 = parse_fn(n1);
                d.n2 = parse_fn(n2);
                d.qe = parse_fn(qe);
                d.spec = parse_fn(sp);
                descs.push_back(d);
            }
        }
        if (!beam) throw runtime_error("no beam");
        if (rt_engine_init() != ENGINE_ERROR_OK) throw runtime_error("rt_engine_init");
        list<string> info;
        if (rt_engine_info(info) == ENGINE_ERROR_OK)
            for (const string& s : info) printf("info %s\n", s.c_str());

        vector<Surface*> scheme;
        for (const SurfDesc& d : descs) scheme.push_back(make_surface(d));

        const auto t0 = std::chrono::steady_clock::now();
        beam->Create(bparams, bcenter, brange, composite, nlambda);
        // Scheme::run per-surface loop (scheme.cpp:983-1046), one wavelength range
        for (size_t i = 0; i < scheme.size(); ++i) {
            Surface* s = scheme[i];
            beam->Transform(*s);
            beam->Normalize();
            if (s->GetType() != Surface::AuxType) {
                s->Intersection(*beam);
                s->ApplyConstrains(*beam);
                s->Action(*beam);
                beam->Normalize();
            }
        }
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        printf("trace_seconds %.6f\n", secs);
        beam->Save(prefix + ".dat");
        printf("ngood %zu of %zu\n", beam->GetNumberOfGoodRays(), beam->GetTotalNumberOfRays());

        // CCD frame of the good rays (ccd_frame.cpp:184-195)
        const SurfDesc& last = descs.back();
        if (last.type == Surface::Detector && last.xdim) {
            size_t n = 0, ngood = 0;
            vector<vector<double>> a;
            if (!read_dat(prefix + ".dat", n, ngood, a)) throw runtime_error("read back");
            vector<unsigned> cx(ngood), cy(ngood);
            if (ccd_coordinates(ngood, a[0].data(), a[1].data(), last.xsize, last.ysize, last.xpix, last.ypix,
                                cx.data(), cy.data()) != ENGINE_ERROR_OK)
                throw runtime_error("ccd_coordinates");
            TabulatedFunction qe = make_fn(last.qe), spec = make_fn(last.spec);
            vector<double> image(last.xdim * last.ydim, 0.0);
            if (compute_ccd_image(ngood, cx.data(), cy.data(), a[6].data(), qe, spec, last.xdim, last.ydim,
                                  last.xpix, last.ypix, image.data()) != ENGINE_ERROR_OK)
                throw runtime_error("compute_ccd_image");
            FILE* f = fopen((prefix + ".ccd").c_str(), "wb");
            fwrite(image.data(), sizeof(double), image.size(), f);
            fclose(f);
        }
        // efficiency of the gratings and the detector on the wavelength grid
        {
            vector<real_t> lam(64), out;
            for (size_t k = 0; k < lam.size(); ++k) lam[k] = brange[0] + (brange[1] - brange[0]) * k / 63.0;
            for (size_t i = 0; i < scheme.size(); ++i) {
                if (descs[i].type != Surface::Grating && !(descs[i].type == Surface::Detector && descs[i].qe.file))
                    continue;
                vector<real_t> spec(lam.size(), 1.0);
                scheme[i]->ApplyQE(lam, spec);
                out.insert(out.end(), spec.begin(), spec.end());
            }
            FILE* f = fopen((prefix + ".qe").c_str(), "wb");
            if (!out.empty()) fwrite(out.data(), sizeof(double), out.size(), f);
            fclose(f);
        }
        // the surfaces live until exit: Surface has no virtual destructor
        delete beam;
    } catch (Beam::bad_beam& e) {
        fprintf(stderr, "bad_beam %d\n", (int)e.beam_error());
        return 1;
    } catch (Surface::bad_surface& e) {
        fprintf(stderr, "bad_surface %d\n", (int)e.surface_error());
        return 1;
    } catch (exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
