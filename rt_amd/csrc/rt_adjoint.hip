// rt_adjoint.hip -- gradient of a loss on the surviving rays' final positions
// with respect to scheme parameters, by a hand-written adjoint of the trace.
//
// One persistent kernel per call.  Each lane traces its rays forward through
// the scheme (the same per-ray stages as the trace), keeping only the state on
// entry to every surface and that surface's Newton iteration count in a tape
// slot of its own (n_surf x 7 doubles per lane, reused ray after ray), then
// walks the surfaces backwards with rt::adj::surface_adjoint, which recomputes
// each surface's internals from its entry state.  The loss enters as a seed per
// ray group g = ray mod n_groups: adjoint of the final position (X, Y) =
// (ax_g + b_g X, ay_g + b_g Y), which covers spot centroids, RMS radii and any
// quadratic of them.  Parameter derivatives are summed per lane into slots of
// their own and reduced afterwards in a fixed order (bit-reproducible).
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "rt_adjoint.h"
#include "rt_device.h"
#include "rt_host.h"

using rt::cptr;
using rt::Ray;

namespace {

constexpr int kAdjBlock = 256;
constexpr int kTape = 7;                 // X Y Z cX cY cZ, Newton iterations

struct AdjArgs {
    double beam[RT_BEAM_STRIDE + 3];     // beam fields + BeamUniform (first: kernarg offset 0)
    const double* table;                 // device records
    const double* pool;
    int n_surf, pool_len, n_par, pad;
    uint64_t seed, n, n_groups;
    const double* seeds;                 // n_groups x 3: ax, ay, b
    const int* prow;                     // n_surf x PG_N: the caller's index of (surface, slot), or -1
    double* tape;                        // (n_surf * kTape) x lanes
    double* gacc;                        // n_par x lanes
    unsigned long long* stats;
};

template <int BEAM>
__global__ __launch_bounds__(kAdjBlock) void rt_adjoint_kernel(AdjArgs a) {
    extern __shared__ double s_pool[];
    __shared__ unsigned s_stats[RT_STATS_LEN];
    for (int i = threadIdx.x; i < RT_STATS_LEN; i += blockDim.x) s_stats[i] = 0u;
    for (int i = threadIdx.x; i < a.pool_len; i += blockDim.x) s_pool[i] = a.pool[i];
    __syncthreads();

    rt::TraceCtx c;
    c.pool = s_pool;
    c.st = s_stats;
    c.hist = nullptr;
    c.flux = nullptr;
    c.do_hist = c.do_flux = false;
    c.n_oob = c.n_hit = c.n_off = c.n_edge = 0;
    c.n_prev = 0.0;
    c.last_it = 0;
    const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int j = 0; j < a.n_par; ++j) a.gacc[(uint64_t)j * G + lane] = 0.0;
    unsigned skipped = 0;

    for (uint64_t i = lane; i < a.n; i += G) {
        cptr tbl = (cptr)a.table;
        __asm__ volatile("" : "+s"(tbl));
        c.table = tbl;
        Ray r;
        cptr bm = (cptr)__builtin_amdgcn_kernarg_segment_ptr();
        rt::gen_ray<BEAM>(bm, a.seed, i, r);
        bool alive = true;
        for (int s = 0; s < a.n_surf; ++s) {
            double* t = a.tape + (uint64_t)s * kTape * G + lane;
            if (alive) {
                t[0] = r.X; t[G] = r.Y; t[2 * G] = r.Z;
                t[3 * G] = r.cX; t[4 * G] = r.cY; t[5 * G] = r.cZ;
            }
            rt::surface_step(rt::DynSurf{c.table + (size_t)s * rt::DEV_STRIDE}, s, r, alive, c);
            if (alive) t[6 * G] = (double)c.last_it;
        }
        unsigned long long good = __ballot(alive), all = __ballot(true);
        if (__lane_id() == 0) {
            atomicAdd(&s_stats[RT_ST_RAYS], (unsigned)__popcll(all));
            if (good) atomicAdd(&s_stats[RT_ST_GOOD], (unsigned)__popcll(good));
        }
        if (!alive) continue;
        // seed: adjoint of the final position
        const uint64_t g = i % a.n_groups;
        const double bx = a.seeds[3 * g], by = a.seeds[3 * g + 1], b = a.seeds[3 * g + 2];
        rt::adj::St ad = {bx + b * r.X, by + b * r.Y, 0.0, 0.0, 0.0, 0.0};
        bool ok = true;
        for (int s = a.n_surf - 1; s >= 0 && ok; --s) {
            const double* p = a.table + (size_t)s * rt::DEV_STRIDE;
            const rt::DynSurf d{c.table + (size_t)s * rt::DEV_STRIDE};
            const double* t = a.tape + (uint64_t)s * kTape * G + lane;
            const rt::adj::St entry = {t[0], t[G], t[2 * G], t[3 * G], t[4 * G], t[5 * G]};
            const int nit = (int)t[6 * G];
            // mu and L exactly as the forward pass formed them
            double mu = 0.0, L = 0.0;
            const int type = d.type();
            if (type == RT_LENS || type == RT_GRATING) {
                int e1 = 0, e2 = 0;
                const double n1 = rt::fn_eval(c.pool, d.n1spline(), p[RT_S_N1VAL], r.lam, e1);
                const double n2 = rt::fn_eval(c.pool, d.n2spline(), p[RT_S_N2VAL], r.lam, e2);
                const double in2 = rt::rcp(n2);
                mu = n1 * in2;
                if (type == RT_GRATING) L = rt::grating_order(d, r.lam) * r.lam * in2 * p[rt::D_INV_D];
            }
            double gp[rt::adj::PG_N];
            for (int k = 0; k < rt::adj::PG_N; ++k) gp[k] = 0.0;
            ok = rt::adj::surface_adjoint(p, entry, nit, mu, L, ad, gp);
            if (!ok) break;
            // the slot loop is unrolled, so gp stays in registers; the row table is uniform
            const __attribute__((address_space(4))) int* rows =
                (const __attribute__((address_space(4))) int*)a.prow + s * rt::adj::PG_N;
#pragma unroll
            for (int sl = 0; sl < rt::adj::PG_N; ++sl) {
                const int j = rows[sl];
                if (j >= 0) a.gacc[(uint64_t)j * G + lane] += gp[sl];
            }
        }
        if (!ok) ++skipped;
    }
    rt::lds_add(&s_stats[RT_ST_SPLINE_OOB], c.n_oob);
    rt::lds_add(&s_stats[RT_ST_ADJ_SKIP], skipped);
    __syncthreads();
    if (a.stats) {
        for (int k = threadIdx.x; k < RT_STATS_LEN; k += blockDim.x)
            if (s_stats[k]) atomicAdd(&a.stats[k], (unsigned long long)s_stats[k]);
    }
}

// grad[j] = sum over lanes of gacc[j][lane], one workgroup per parameter, in a
// fixed order
__global__ __launch_bounds__(256) void rt_adjoint_reduce(const double* gacc, uint64_t G, double* grad) {
    __shared__ double s[256];
    const uint64_t j = blockIdx.x;
    double v = 0.0;
    for (uint64_t k = threadIdx.x; k < G; k += blockDim.x) v += gacc[j * G + k];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad[j] = s[0];
}

struct AdjBuffers {
    int device = -1;
    double* tape = nullptr;
    size_t tape_len = 0;
    double* gacc = nullptr;
    size_t gacc_len = 0;
    int* params = nullptr;
    size_t params_len = 0;
};
thread_local AdjBuffers g_adj;

bool grow(double*& p, size_t& have, size_t want) {
    if (have >= want) return true;
    (void)hipFree(p);
    p = nullptr;
    have = 0;
    if (hipMalloc(&p, want * sizeof(double)) != hipSuccess) return false;
    have = want;
    return true;
}

}  // namespace

// defined in rt_trace.hip
extern "C" int rt_hip_scheme_device_tables(const rt_hip_scheme* s, const double** table_dev, const double** pool_dev,
                                           int* n_surf, size_t* pool_len, const double** host_table);

extern "C" int rt_hip_adjoint(const rt_hip_scheme* sch, const double* beam_desc, uint64_t seed, size_t n_rays,
                              const double* group_seeds, size_t n_groups, const int* param_surface,
                              const int* param_field, int n_params, double* grad, unsigned long long* stats,
                              void* stream) {
    const double *table = nullptr, *pool = nullptr, *host_table = nullptr;
    int n_surf = 0;
    size_t pool_len = 0;
    int rc = rt_hip_scheme_device_tables(sch, &table, &pool, &n_surf, &pool_len, &host_table);
    if (rc != RT_OK) return rc;
    if (!beam_desc) return rt_set_error(RT_ERR_BAD_VALUE, "the adjoint traces a generated beam: beam_desc is NULL");
    const int beam_type = (int)beam_desc[RT_B_TYPE];
    if (beam_type != RT_BEAM_PARALLEL && beam_type != RT_BEAM_CONIC && beam_type != RT_BEAM_GAUSS)
        return rt_set_error(RT_ERR_BAD_VALUE, "bad beam type");
    if (!group_seeds || n_groups == 0) return rt_set_error(RT_ERR_BAD_VALUE, "group seeds missing");
    for (int s = 0; s < n_surf; ++s) {
        const double* p = host_table + (size_t)s * RT_SURF_STRIDE;
        if ((int)p[RT_S_TYPE] == RT_GRATING && (int)p[RT_S_CLASS] != RT_PLANE)
            return rt_set_error(RT_ERR_BAD_VALUE, "surface %d: the adjoint covers plane gratings only", s);
    }
    if (n_params < 0 || n_params > 4096 || (n_params > 0 && (!param_surface || !param_field || !grad)))
        return rt_set_error(RT_ERR_BAD_VALUE, "bad parameter list");
    // row table: the caller's index of each (surface, slot), -1 where not asked for
    std::vector<int> host((size_t)n_surf * rt::adj::PG_N, -1);
    for (int j = 0; j < n_params; ++j) {
        if (param_surface[j] < 0 || param_surface[j] >= n_surf)
            return rt_set_error(RT_ERR_BAD_VALUE, "parameter %d: surface %d out of range", j, param_surface[j]);
        int slot = rt::adj::slot_of_field(param_field[j]);
        if (slot < 0) return rt_set_error(RT_ERR_BAD_VALUE, "parameter %d: field %d is not differentiable", j,
                                          param_field[j]);
        if (slot >= rt::adj::PG_A && slot < rt::adj::PG_CR &&
            slot - rt::adj::PG_A >= (int)host_table[(size_t)param_surface[j] * RT_SURF_STRIDE + RT_S_NPOLY])
            return rt_set_error(RT_ERR_BAD_VALUE, "parameter %d: aspheric term beyond the surface's N", j);
        int& row = host[(size_t)param_surface[j] * rt::adj::PG_N + slot];
        if (row >= 0) return rt_set_error(RT_ERR_BAD_VALUE, "parameter %d repeats parameter %d", j, row);
        row = j;
    }
    if (n_rays == 0) {
        if (n_params) HIP_TRY(hipMemsetAsync(grad, 0, n_params * sizeof(double), (hipStream_t)stream), RT_ERR_FAILED);
        return RT_OK;
    }
    rc = rt_ensure_device();
    if (rc != RT_OK) return rc;

    AdjArgs a;
    memset(&a, 0, sizeof a);
    memcpy(a.beam, beam_desc, RT_BEAM_STRIDE * sizeof(double));
    double ang = a.beam[RT_B_ANGUNIT] != 0.0 ? a.beam[RT_B_ANGPAR] : a.beam[RT_B_ANGPAR] * rt::kPi / 180.0;
    a.beam[rt::BU_CONE_COS] = cos(ang / 2.0);                           // beam.cu:593-594
    a.beam[rt::BU_CONE_SIN] = sin(ang / 2.0);
    a.beam[rt::BU_GAUSS_TG] = tan(a.beam[RT_B_ANGPAR] / 2.35482 * rt::kPi / 180 / 3600);   // beam.cu:1034,1099
    a.table = table;
    a.pool = pool;
    a.n_surf = n_surf;
    a.pool_len = (int)pool_len;
    a.n_par = n_params;
    a.seed = seed;
    a.n = n_rays;
    a.n_groups = n_groups;
    a.seeds = group_seeds;
    a.stats = stats;

    auto kern = beam_type == RT_BEAM_PARALLEL ? rt_adjoint_kernel<RT_BEAM_PARALLEL>
              : beam_type == RT_BEAM_CONIC    ? rt_adjoint_kernel<RT_BEAM_CONIC>
                                              : rt_adjoint_kernel<RT_BEAM_GAUSS>;
    const size_t lds = pool_len * sizeof(double);
    int resident = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kern, kAdjBlock, lds) != hipSuccess || resident < 1)
        resident = 1;
    uint64_t want = (n_rays + kAdjBlock - 1) / kAdjBlock;
    uint64_t cap = (uint64_t)rt_device_cus() * resident;
    const unsigned grid = (unsigned)(want < cap ? want : cap);
    const uint64_t G = (uint64_t)grid * kAdjBlock;

    AdjBuffers& B = g_adj;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), RT_ERR_FAILED);
    if (B.device != dev) {
        (void)hipFree(B.tape); (void)hipFree(B.gacc); (void)hipFree(B.params);
        B = AdjBuffers();
        B.device = dev;
    }
    if (!grow(B.tape, B.tape_len, (size_t)n_surf * kTape * G) ||
        !grow(B.gacc, B.gacc_len, (size_t)std::max(n_params, 1) * G))
        return rt_set_error(RT_ERR_BAD_ALLOC, "adjoint buffers");
    if (B.params_len < host.size()) {
        (void)hipFree(B.params);
        B.params = nullptr;
        B.params_len = 0;
        HIP_TRY(hipMalloc(&B.params, host.size() * sizeof(int)), RT_ERR_BAD_ALLOC);
        B.params_len = host.size();
    }
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(B.params, host.data(), host.size() * sizeof(int), hipMemcpyHostToDevice, st),
            RT_ERR_FAILED);
    a.prow = B.params;
    a.tape = B.tape;
    a.gacc = B.gacc;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kAdjBlock), lds, st, a);
    HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    if (n_params > 0) {
        hipLaunchKernelGGL(rt_adjoint_reduce, dim3(n_params), dim3(256), 0, st, (const double*)B.gacc, G, grad);
        HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    }
    // the host copy of the parameter list must outlive the asynchronous upload
    HIP_TRY(hipStreamSynchronize(st), RT_ERR_FAILED);
    return RT_OK;
}
