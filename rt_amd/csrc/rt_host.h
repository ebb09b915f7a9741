// rt_host.h -- host-side helpers shared by the translation units of librt_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

// Record an error message for rt_hip_last_error and return `code`.
int rt_set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define HIP_TRY(expr, code)                                                                    \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return rt_set_error(code, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                                __LINE__);                                                     \
    } while (0)

// Device working memory of a scheme or batch (or of one host thread, for the
// calls that belong to neither): the adjoint's tape, gradient partials and
// row table, the spot partials, the descent map.  A buffer that must grow is
// replaced by a new allocation and the old one is kept until release() (the
// owner's destruction): a graph captured before may still hold its address,
// so no buffer a call has handed to a kernel is ever freed while its owner
// lives.  The calls of one owner share its buffers and therefore run in the
// order of one stream.
struct WorkBufs {
    enum Slot { TAPE, GACC, PART, PARAMS, SPOT_PART, SPOT_CS, DESCEND, EXT_KEYS, SAVE_HIST, SAVE_LPOS, SAVE_TILES,
                SAVE_GPOS, TRACE_PART, N_SLOTS };
    void* p[N_SLOTS] = {};
    size_t have[N_SLOTS] = {};
    std::vector<void*> owned;
    std::vector<int> params_key;     // contents of the PARAMS buffer
    // p[s] holds at least `bytes` afterwards (RT_OK, or RT_ERR_BAD_ALLOC
    // with p[s] unchanged)
    int ensure(Slot s, size_t bytes);
    // p[s] is a new buffer of `bytes` afterwards (for contents that change
    // between calls while a captured graph may still read the old ones)
    int fresh(Slot s, size_t bytes);
    // frees every buffer ever allocated (the owner is being destroyed)
    void release();
    template <class T>
    T* get(Slot s) const { return (T*)p[s]; }
};

// Current device selected and its properties cached (RT_OK or an error).
int rt_ensure_device();
// Compute units of the current device (after rt_ensure_device).
int rt_device_cus();
// Looks up, once per device, the device address of the scratch array of
// rt_hip_group_spots_noalloc (dev must be the current device).
int rt_scratch_setup(int dev);

// Device record (rt::DEV_STRIDE doubles) of one public packed record: the
// public fields plus the host-derived constants; spline references are
// copied unchanged.
void rt_device_record(const double* pub, double* dev);

// Structural key of a scheme's generated kernels and the device it lives on.
struct rt_hip_scheme;
const std::string& rt_scheme_signature(const rt_hip_scheme* s);
int rt_scheme_device(const rt_hip_scheme* s);
// the working buffers owned by a scheme / a batch
WorkBufs* rt_scheme_work(const rt_hip_scheme* s);

// What the batched adjoint needs of a batch (device pointers and structure).
struct rt_hip_batch;
struct BatchView {
    int device, n_inst, n_surf, block;
    const std::string* signature;
    const double* tables_dev;        // n_inst x n_surf device records
    const double* pool_dev;
    size_t pool_len;
    const uint64_t* ray_start_dev;   // n_inst + 1
    const uint64_t* tile_start_dev;  // n_inst + 1 (tiles of `block` rays)
    const uint32_t* tile_inst_dev;   // n_tiles: the instance of each tile
    uint64_t n_tiles;
    const double* dtables;           // host copy of the device records
    const uint64_t* ray_start;       // host copy
    const double* beams_dev;         // a beam per instance (kBeamRec doubles each), or null
    int beam_type;                   // their beam type (-1: none)
};
int rt_batch_view(const rt_hip_batch* b, BatchView* v);
WorkBufs* rt_batch_work(const rt_hip_batch* b);

// The per-wavelength table of a scheme or batch (its spline-valued functions
// at the grid wavelengths of a beam): the number of grid wavelengths it
// serves for this beam description, 0 when it does not (random wavelengths,
// no spline-valued function, tables disabled); and its filling on `stream`
// (*tab: the device table, column q at q * nk).
uint64_t rt_scheme_ltab_nk(const rt_hip_scheme* s, const double* beam_desc);
int rt_scheme_ltab_fill(const rt_hip_scheme* s, const double* beam_desc, uint64_t nk, hipStream_t stream,
                        const double** tab);
// The derived per-wavelength table of a scheme (mu, grating term, flux
// weight; the layout of rt_jit_dtab_layout), likewise; kept while the grid
// and the stream stay those of its last fill.
uint64_t rt_scheme_dtab_nk(const rt_hip_scheme* s, const double* beam_desc);
int rt_scheme_dtab_fill(const rt_hip_scheme* s, const double* beam_desc, uint64_t nk, hipStream_t stream,
                        const double** tab);
uint64_t rt_batch_ltab_nk(const rt_hip_batch* b, const double* beam_desc);
int rt_batch_ltab_fill(const rt_hip_batch* b, const double* beam_desc, uint64_t nk, hipStream_t stream,
                       const double** tab);
