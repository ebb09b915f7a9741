// ccd_frame.cpp -- librt_scheme_hip.so: ccd/ccd_frame.cpp's program on the
// MI355X engine.  The flux image lives in HBM for the whole list: per result
// file the good rays' X, Y, lambda go to the device once and rt_hip_ccd_flux
// maps, weighs and accumulates them in one kernel (the reference copies the
// 128 MiB image both ways in every compute_ccd_image call).  The image is
// written as FITS by a small writer (cfitsio is not available here): a
// primary HDU with BITPIX -64, big-endian data, 2880-byte records (FITS
// Standard 4.0, sections 3-4).
#include "rt/beam.h"
#include "rt/ccd_frame.h"
#include "rt/tabulated_function.h"
#include "rt_ascii.h"
#include "rt_hip.h"

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

// one 80-character header card
void card(std::string& hdr, const std::string& key, const std::string& value, const std::string& comment = "") {
    std::string c = key;
    c.resize(8, ' ');
    c += "= ";
    c += value;
    if (!comment.empty()) c += " / " + comment;
    c.resize(80, ' ');
    hdr += c;
}

std::string right(const std::string& v) {          // fixed-format value, right-justified to column 30
    return v.size() >= 20 ? v : std::string(20 - v.size(), ' ') + v;
}

std::string quoted(const char* s) {                 // string value: quotes doubled, at least 8 characters
    std::string v;
    for (const char* p = s; *p && v.size() < 68; ++p) {
        if (*p == '\'') {
            if (v.size() + 2 > 68) break;
            v += "''";
        } else {
            v += *p;
        }
    }
    if (v.size() < 8) v.resize(8, ' ');
    return "'" + v + "'";
}

struct Dev {
    void* p = nullptr;
    size_t have = 0;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    bool fit(size_t bytes) {
        if (have >= bytes && p) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        have = 0;
        if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        have = bytes;
        return true;
    }
};

// [K, x, y, y2] of a tabulated function, or empty for a single value
std::vector<double> spline_of(const TabulatedFunction& f, double& constant) {
    std::vector<double> s;
    constant = f.Values().empty() ? 0.0 : f.Values()[0];
    if (f.IsSingleValue()) return s;
    const auto& x = f.Abscissae();
    s.push_back((double)x.size());
    s.insert(s.end(), x.begin(), x.end());
    s.insert(s.end(), f.Values().begin(), f.Values().end());
    s.insert(s.end(), f.SecondDerivatives().begin(), f.SecondDerivatives().end());
    return s;
}

}  // namespace

extern "C" int rt_write_fits_image(const char* path, size_t xdim, size_t ydim, const double* image, int n_keys,
                                   const char* const* keys, const char* const* values) {
    std::string hdr;
    card(hdr, "SIMPLE", right("T"), "file does conform to FITS standard");
    card(hdr, "BITPIX", right("-64"), "number of bits per data pixel");
    card(hdr, "NAXIS", right("2"), "number of data axes");
    card(hdr, "NAXIS1", right(std::to_string(xdim)), "length of data axis 1");
    card(hdr, "NAXIS2", right(std::to_string(ydim)), "length of data axis 2");
    card(hdr, "EXTEND", right("T"), "FITS dataset may contain extensions");
    for (int i = 0; i < n_keys; ++i) card(hdr, keys[i], quoted(values[i]));
    std::string end = "END";
    end.resize(80, ' ');
    hdr += end;
    hdr.resize((hdr.size() + 2879) / 2880 * 2880, ' ');

    FILE* f = fopen(path, "wb");
    if (!f) return 1;
    bool ok = fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    const size_t n = xdim * ydim;
    std::vector<unsigned char> buf;
    const size_t chunk = 1 << 16;
    for (size_t b = 0; ok && b < n; b += chunk) {
        const size_t m = n - b < chunk ? n - b : chunk;
        buf.resize(m * 8);
        for (size_t i = 0; i < m; ++i) {
            uint64_t u;
            memcpy(&u, image + b + i, 8);
            for (int k = 0; k < 8; ++k) buf[i * 8 + k] = (unsigned char)(u >> (56 - 8 * k));   // big-endian
        }
        ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    const size_t pad = (2880 - (n * 8) % 2880) % 2880;
    if (ok && pad) {
        std::vector<unsigned char> z(pad, 0);
        ok = fwrite(z.data(), 1, pad, f) == pad;
    }
    if (fclose(f) != 0) ok = false;
    return ok ? 0 : 1;
}

extern "C" int rt_ccd_frame(const char* rt_files_list, const char* qe_file, const char* spec_file,
                            const char* ccd_spec_file, const char* fits_file) {
    // CCD specification: linear size, pixels, pixel size (ccd_frame.cpp:42-75)
    double spec3[3][2];
    {
        rtascii::Reader rd(ccd_spec_file);
        if (!rd.ok()) {
            std::cerr << "Can not open CCD specification file!\n";
            return 2;
        }
        for (int k = 0; k < 3; ++k) {
            int flag = 0;
            std::vector<double> v;
            for (;;) {
                if (!rd.next(2, flag, v) || flag == rtascii::kInvalid) {
                    std::cerr << "Invalid CCD linear size specification!\n";
                    return 2;
                }
                if (flag == rtascii::kData) break;
            }
            spec3[k][0] = v[0];
            spec3[k][1] = v[1];
        }
    }
    const double xsize = spec3[0][0], ysize = spec3[0][1];
    const double ccd_xdim = spec3[1][0], ccd_ydim = spec3[1][1];
    const double xpix = spec3[2][0], ypix = spec3[2][1];
    std::cout << "CCD specifications:\n";
    std::cout << "Linear size: [" << xsize << ":" << ysize << "]" << std::endl;
    std::cout << "Pixel linear size: [" << xpix << ":" << ypix << "]" << std::endl;
    std::cout << "Number of pixels: [" << ccd_xdim << ":" << ccd_ydim << "]" << std::endl;
    const size_t xdim = (size_t)(long)ccd_xdim, ydim = (size_t)(long)ccd_ydim;
    const size_t image_size = xdim * ydim;

    FILE* qef = fopen(qe_file, "rb");
    if (!qef) {
        std::cerr << "Can not open QE file!\n";
        return 3;
    }
    TabulatedFunction spec;
    try {
        spec.Load(spec_file);
    } catch (TabulatedFunction::bad_tab_func& ex) {
        std::cerr << "Can not load input spectrum function! Error code: " << ex.bad_tab_func_error() << std::endl;
        fclose(qef);
        return 3;
    }
    double spec_c = 0.0;
    const std::vector<double> spec_sp = spline_of(spec, spec_c);

    Dev image, dx, dy, dl, dcounts;
    hipStream_t st = nullptr;
    if (!image.fit(image_size * sizeof(double)) || !dcounts.fit(3 * sizeof(unsigned long long)) ||
        hipStreamCreate(&st) != hipSuccess) {
        std::cerr << "Can not allocate memory for the result image!\n";
        fclose(qef);
        return 100;
    }
    (void)hipMemsetAsync(image.p, 0, image_size * sizeof(double), st);

    std::ifstream list(rt_files_list);
    if (!list) {
        std::cerr << "Can not open file with list of ray-tracing result files!\n";
        fclose(qef);
        (void)hipStreamDestroy(st);
        return 5;
    }
    std::cout << "\nSTART CCD IMAGE CREATION ...\n";
    int flag = 0;
    std::string name;
    while (std::getline(list, name)) {
        if (list.eof()) break;              // a last line without a newline is not read (:140-142)
        std::cout << "   Processing file: " << name << "...  \n";
        size_t Npoints = 0, Nrays = 0, N_good = 0, N_lambda = 0;
        real_t *X = nullptr, *Y = nullptr, *lambda = nullptr;
        std::vector<double> ql, qv;
        int code = 0;
        try {
            // this file's block of the QE file (:160-175)
            if (fread(&Npoints, sizeof(size_t), 1, qef) != 1) throw 30;
            ql.resize(Npoints);
            qv.resize(Npoints);
            if (fread(ql.data(), sizeof(double), Npoints, qef) != Npoints) throw 30;
            if (fread(qv.data(), sizeof(double), Npoints, qef) != Npoints) throw 30;
            TabulatedFunction QE;
            QE.Set(Npoints, ql.data(), qv.data());
            double qe_c = 0.0;
            const std::vector<double> qe_sp = spline_of(QE, qe_c);

            if (load_beam_data(name.c_str(), &Nrays, &N_good, &N_lambda, &X, &Y, nullptr, nullptr, nullptr, nullptr,
                               &lambda)) throw 1;
            const size_t b = N_good * sizeof(double);
            if (!dx.fit(b) || !dy.fit(b) || !dl.fit(b)) throw 50;
            if (hipMemcpyAsync(dx.p, X, b, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(dy.p, Y, b, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(dl.p, lambda, b, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemsetAsync(dcounts.p, 0, 3 * sizeof(unsigned long long), st) != hipSuccess)
                throw 10;
            if (rt_hip_ccd_flux(N_good, (const double*)dx.p, (const double*)dy.p, (const double*)dl.p,
                                qe_sp.empty() ? nullptr : qe_sp.data(), qe_c,
                                spec_sp.empty() ? nullptr : spec_sp.data(), spec_c, xsize, ysize, xpix, ypix, xdim,
                                ydim, (double*)image.p, (unsigned long long*)dcounts.p, st) != RT_OK)
                throw 20;
            unsigned long long counts[3];
            if (hipMemcpy(counts, dcounts.p, sizeof counts, hipMemcpyDeviceToHost) != hipSuccess) throw 20;
            if (counts[2]) throw 20;          // a wavelength outside the QE or spectrum table
        } catch (int c) {
            code = c;
        } catch (TabulatedFunction::bad_tab_func& ex) {
            flag = 1;
            std::cerr << "\nError occured in QE data creation! Error code: " << ex.bad_tab_func_error() << std::endl;
        }
        switch (code) {
            case 0: break;
            case 1: std::cerr << "\nAn error occured while read the ray-tracing output file!\n"; break;
            case 10: std::cerr << "\nAn error occured while compute CCD coordinates!\n"; break;
            case 20: std::cerr << "\nAn error occured while compute CCD image!\n"; break;
            case 30: std::cerr << "\nAn error occured while read QE file!\n"; break;
            default: std::cerr << "\nMemory allocation failure!\n"; break;
        }
        if (code) flag = 1;
        // the host arrays may be in use by the copies until the stream drains
        (void)hipStreamSynchronize(st);
        free(X);
        free(Y);
        free(lambda);
        if (flag) break;
    }
    if (!flag && list.bad()) {
        std::cerr << "\nAn error occured while read the file!!!\n";
        flag = 1;
    }
    fclose(qef);
    if (flag) {
        (void)hipStreamDestroy(st);
        return 5;
    }
    std::cout << "DONE!\n";

    std::cout << "\nSAVE IMAGE TO FITS FILE ...  ";
    std::vector<double> host(image_size);
    if (hipMemcpyAsync(host.data(), image.p, image_size * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        (void)hipStreamDestroy(st);
        std::cerr << "Can not create result FITS image! Error code " << 1 << std::endl;
        return 4;
    }
    (void)hipStreamDestroy(st);
    const char* keys[] = {"RT_LIST", "QE_FILE", "SPECFILE", "CCDINFO"};
    const char* vals[] = {rt_files_list, qe_file, spec_file, ccd_spec_file};
    if (rt_write_fits_image(fits_file, xdim, ydim, host.data(), 4, keys, vals) != 0) {
        std::cerr << "Can not create result FITS file! Error code " << 1 << std::endl;
        return 4;
    }
    std::cout << "OK!\n";
    return 0;
}
