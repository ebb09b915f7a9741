// rt_jit.h -- per-scheme kernels generated at run time (hiprtc).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

// Structural key of a scheme: one StatSurf type per device record.
std::string rt_jit_signature(const double* dtable, int n_surf);

// The columns of the per-wavelength table of a scheme with this signature:
// col_base[s] is surface s's first column or -1; column q holds the function
// of role col_role[q] (0 n1, 1 n2, 2 detector QE, 3 detector spectrum) of
// surface col_surf[q].  Kernels generated with RT_JIT_LTAB read that layout.
void rt_jit_ltab_layout(const std::string& signature, std::vector<int>& col_base, std::vector<int>& col_surf,
                        std::vector<int>& col_role);

// The columns of the derived per-wavelength table that generated
// single-scheme traces read (RT_JIT_LTAB without batch): column q holds role
// col_role[q] (RT_DTAB_MU n1/n2, RT_DTAB_L the grating's m lambda / (n2 d),
// RT_DTAB_W the detector's flux weight) of surface col_surf[q]; col_base[s]
// is surface s's first column or -1.  rt::dtab_body writes the same order.
enum { RT_DTAB_MU = 0, RT_DTAB_L = 1, RT_DTAB_W = 2 };
void rt_jit_dtab_layout(const std::string& signature, std::vector<int>& col_base, std::vector<int>& col_surf,
                        std::vector<int>& col_role);

// Whether generated kernels are used (environment RT_HIP_JIT=0 disables;
// rt_hip_set_jit overrides).
bool rt_jit_enabled();

// Whether generated traces of grid-wavelength beams read the per-wavelength
// table (environment RT_HIP_LTAB=0 disables; rt_hip_set_ltab overrides).
bool rt_jit_ltab_enabled();

// The kernel generated for (signature, beam mode, per-surface state dump or
// not, lanes per workgroup, minimum waves per SIMD or 0) on `device`,
// compiled on first use and cached for the life of the process; nullptr
// when the runtime compiler is unavailable or fails (the caller then
// launches the generic kernel, which reads the structure from the records).
// kernel variants of rt_jit_kernel (bit mask)
enum { RT_JIT_VAR_DUMP = 1, RT_JIT_VAR_LOSS = 2, RT_JIT_VAR_EXT = 4, RT_JIT_VAR_SPOTS = 8 };

hipFunction_t rt_jit_kernel(const std::string& signature, int beam_mode, int variant, int block, int waves,
                            int device, bool batch = false);

// The adjoint kernel generated for (signature, beam type, gradient rows,
// lanes per workgroup, minimum waves per SIMD or 0, single scheme or batch of
// instances, reading the per-wavelength table or not) on `device`: rows holds
// n_surf x PG_N entries, the caller's parameter index of each (surface, slot)
// or -1.  Compiled on first use and cached; nullptr on failure (the caller
// launches the generic adjoint kernel).
// Tape rows per lane of the generated adjoint kernels of a signature (entry
// state and Newton count per surface, the Newton iterates where the
// intersection iterates); the generic adjoint kernel uses n_surf x kTape.
int rt_jit_adjoint_tape_rows(const std::string& signature);

hipFunction_t rt_jit_adjoint_kernel(const std::string& signature, int beam_type, const std::vector<int>& rows,
                                    int n_surf, int block, int waves, int device, bool batch,
                                    bool ltab = false);
