/* ccd_frame.h -- the ccd_frame program of amirych/rt (ccd/ccd_frame.cpp) as
 * a library call on the MI355X engine, and the FITS writer it needs (the
 * reference writes with cfitsio).  Definitions: librt_scheme_hip.so. */
#ifndef RT_CCD_FRAME_H
#define RT_CCD_FRAME_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/*
 * ccd_frame rt_files_list QE_filename spec_filename ccd_spec_filename
 * result_FITS_filename (ccd_frame.cpp:22-319): the flux image of the good
 * rays of every result file in the list, each with its block of the QE file
 * (size_t N, N wavelengths, N values: Scheme::ComputeQE's output), weighted
 * by the incident spectrum, written as a FITS image (BITPIX -64, NAXIS1 =
 * xdim, NAXIS2 = ydim, keywords RT_LIST, QE_FILE, SPECFILE, CCDINFO).  The
 * image stays on the device across the files.  Returns ccd_frame's exit
 * codes: 0, 2 bad CCD specification, 3 QE or spectrum file, 4 FITS file,
 * 5 list or result file, 100 out of memory; messages as the reference's on
 * stdout / stderr.
 */
int rt_ccd_frame(const char *rt_files_list, const char *qe_file, const char *spec_file,
                 const char *ccd_spec_file, const char *fits_file);

/* A primary-HDU FITS file of one xdim x ydim FP64 image (row y at
 * image[y * xdim]), string keywords keys[i] = values[i] (values longer than
 * 68 characters are cut), overwriting `path`.  0 on success. */
int rt_write_fits_image(const char *path, size_t xdim, size_t ydim, const double *image, int n_keys,
                        const char *const *keys, const char *const *values);

#ifdef __cplusplus
}
#endif

#endif /* RT_CCD_FRAME_H */
