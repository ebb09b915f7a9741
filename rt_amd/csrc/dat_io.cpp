// dat_io.cpp -- load_beam_data: the reader of the beams Scheme::run saves
// (Beam::Save's layout, common-lib/beam.cpp:636-715).  Interface and return
// codes of beam.cpp:829-1036, so ccd_frame, losses_comp and the tools read
// rt_amd's files with it unchanged.
#include "rt/beam.h"

#include <stdio.h>
#include <stdlib.h>

namespace {

// one array of the file: `count` doubles into dst (NULL: skipped)
bool read_or_skip(FILE* f, real_t* dst, size_t count) {
    if (dst) return fread(dst, sizeof(real_t), count, f) == count;
    return fseek(f, (long)(count * sizeof(real_t)), SEEK_CUR) == 0;
}

}  // namespace

int load_beam_data(const char* rt_filename, size_t* total_Nrays, size_t* N_good, size_t* N_lambda, real_t** X,
                   real_t** Y, real_t** Z, real_t** cX, real_t** cY, real_t** cZ, real_t** lambda, bool all) {
    FILE* f = fopen(rt_filename, "rb");
    if (!f) return 1;
    real_t** outs[7] = {X, Y, Z, cX, cY, cZ, lambda};
    int rc = 0;
    if (fread(total_Nrays, sizeof(size_t), 1, f) != 1 || fread(N_good, sizeof(size_t), 1, f) != 1 ||
        fread(N_lambda, sizeof(size_t), 1, f) != 1) {
        rc = 2;
    } else if (*N_good > *total_Nrays) {
        rc = 2;
    } else {
        const size_t len = all ? *total_Nrays : *N_good;
        for (int k = 0; k < 7 && rc == 0; ++k) {
            if (!outs[k]) continue;
            *outs[k] = (real_t*)malloc((len ? len : 1) * sizeof(real_t));
            if (!*outs[k]) rc = 3;
        }
        // the good rays of every array, then (all) the rest of every array
        for (int k = 0; k < 7 && rc == 0; ++k)
            if (!read_or_skip(f, outs[k] ? *outs[k] : nullptr, *N_good)) rc = 2;
        if (all) {
            const size_t rest = *total_Nrays - *N_good;
            for (int k = 0; k < 7 && rc == 0; ++k)
                if (!read_or_skip(f, outs[k] ? *outs[k] + *N_good : nullptr, rest)) rc = 2;
        }
    }
    fclose(f);
    return rc;
}
