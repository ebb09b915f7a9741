// rt_adjoint.hip -- gradient of a loss on the surviving rays' final positions
// with respect to scheme parameters, by a hand-written adjoint of the trace.
//
// One persistent kernel per call (body: rt::adj::adjoint_body in
// rt_adjoint.h): forward with a per-lane tape, then the surfaces backwards.
// The kernel is generated per scheme structure and parameter list (rt_jit.cpp:
// compile-time surface descriptors, so every structural branch folds away and
// the gradient slots are registers); the generic kernel below reads the
// structure at run time and is the fallback.  Per-lane accumulators are
// reduced afterwards in two fixed-order stages (bit-reproducible).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "rt_adjoint.h"
#include "rt_device.h"
#include "rt_host.h"
#include "rt_jit.h"

using rt::cptr;
using rt::Ray;
using rt::adj::AdjArgs;
using rt::adj::kTape;

namespace {

constexpr int kAdjBlock = 256;

// p[0 .. n) = 0 by a kernel (a caller may capture the call in a hipGraph,
// where memset nodes are avoided)
__global__ __launch_bounds__(256) void rt_zero_kernel(double* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

int zero_fill(double* p, uint64_t n, hipStream_t st) {
    hipLaunchKernelGGL(rt_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    return RT_OK;
}
constexpr int kReduceChunks = 64;        // first-stage workgroups per parameter

template <int BEAM>
__global__ __launch_bounds__(kAdjBlock) void rt_adjoint_kernel(AdjArgs a) {
    using rt::adj::LaneRows;
    rt::adj::adjoint_body<BEAM, rt::adj::WaveAcc>(
        a,
        [&](Ray& r, bool& alive, rt::TraceCtx& c, const LaneRows& t) {
            for (int s = 0; s < a.n_surf; ++s)
                rt::adj::fwd_step(rt::DynSurf{c.table + (size_t)s * rt::DEV_STRIDE}, s, r, alive, c, t, s * kTape);
        },
        [&](const Ray& r, rt::adj::St& ad, rt::TraceCtx& c, const LaneRows& t, const auto& g, bool ok) {
            for (int s = a.n_surf - 1; s >= 0; --s) {
                // the row table is uniform: read through the constant address space
                const __attribute__((address_space(4))) int* rows =
                    (const __attribute__((address_space(4))) int*)a.prow + s * rt::adj::PG_N;
                double gp[rt::adj::PG_N];
                const bool res =
                    rt::adj::bwd_step(rt::DynSurf{c.table + (size_t)s * rt::DEV_STRIDE}, r, ad, c, t, s * kTape, gp);
                const bool use = ok && res;
#pragma unroll
                for (int sl = 0; sl < rt::adj::PG_N; ++sl) {
                    const int j = rows[sl];
                    if (j >= 0) g.add(j, use ? gp[sl] : 0.0);
                }
                ok = use;
            }
            return ok;
        });
}

// the same surface steps over a batch's tiles (adjoint_batch_body)
template <int BEAM>
__global__ __launch_bounds__(kAdjBlock) void rt_adjoint_batch_kernel(AdjArgs a) {
    using rt::adj::LaneRows;
    rt::adj::adjoint_batch_body<BEAM, rt::adj::TileOut>(
        a,
        [&](Ray& r, bool& alive, rt::TraceCtx& c, const LaneRows& t) {
            for (int s = 0; s < a.n_surf; ++s)
                rt::adj::fwd_step(rt::DynSurf{c.table + (size_t)s * rt::DEV_STRIDE}, s, r, alive, c, t, s * kTape);
        },
        [&](const Ray& r, rt::adj::St& ad, rt::TraceCtx& c, const LaneRows& t, const auto& g, bool ok) {
            for (int s = a.n_surf - 1; s >= 0; --s) {
                // the row table is uniform: read through the constant address space
                const __attribute__((address_space(4))) int* rows =
                    (const __attribute__((address_space(4))) int*)a.prow + s * rt::adj::PG_N;
                double gp[rt::adj::PG_N];
                const bool res =
                    rt::adj::bwd_step(rt::DynSurf{c.table + (size_t)s * rt::DEV_STRIDE}, r, ad, c, t, s * kTape, gp);
                const bool use = ok && res;
#pragma unroll
                for (int sl = 0; sl < rt::adj::PG_N; ++sl) {
                    const int j = rows[sl];
                    if (j >= 0) g.add(j, use ? gp[sl] : 0.0);
                }
                ok = use;
            }
            return ok;
        });
}

// Fixed-order reduction of the per-wave accumulators: stage 1 sums chunk b of
// parameter j's waves (blockIdx = (b, j)) into part[j][b], stage 2 sums the
// chunks of each parameter; both with the same tree over the workgroup.
__device__ __forceinline__ double block_sum_256(double v) {
    __shared__ double s[256];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    return s[0];
}

__global__ __launch_bounds__(256) void rt_adjoint_reduce_chunks(const double* gacc, uint64_t G, double* part) {
    const uint64_t j = blockIdx.y, b = blockIdx.x;
    const uint64_t per = (G + kReduceChunks - 1) / kReduceChunks;
    const uint64_t lo = b * per, hi = lo + per < G ? lo + per : G;
    double v = 0.0;
    for (uint64_t k = lo + threadIdx.x; k < hi; k += blockDim.x) v += gacc[j * G + k];
    v = block_sum_256(v);
    if (threadIdx.x == 0) part[j * kReduceChunks + b] = v;
}

// grad[i][j] = sum over instance i's tile-waves tw of part[tw][j], one
// workgroup per instance (blockIdx.x = i), fixed order. The columns are taken
// in chunks of C = min(n_par, 256); lane k adds column c0 + k % C of rows
// lo + k / C, lo + k / C + R, ... (R = 256 / C rows per pass, so a pass reads
// R whole rows, coalesced) into 8 sums, row m into sum m mod 8 (8 loads in
// flight per lane: one workgroup reads the largest instance's rows), the 8
// sums are added as a fixed tree, then the R partial sums of each column in
// row order.
__global__ __launch_bounds__(256) void rt_adjoint_batch_reduce(const double* part, const uint64_t* tile_start,
                                                              uint64_t n_par, uint64_t wpg, double* grad) {
    __shared__ double s[256];
    const uint64_t i = blockIdx.x, k = threadIdx.x;
    const uint64_t lo = tile_start[i] * wpg, hi = tile_start[i + 1] * wpg;
    for (uint64_t c0 = 0; c0 < n_par; c0 += 256) {
        const uint64_t C = n_par - c0 < 256 ? n_par - c0 : 256, R = 256 / C;
        const uint64_t col = c0 + k % C, r0 = k / C;
        double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (r0 < R) {
            uint64_t tw = lo + r0;
            for (; tw + 7 * R < hi; tw += 8 * R) {
#pragma unroll
                for (int u = 0; u < 8; ++u) acc[u] += part[(tw + u * R) * n_par + col];
            }
#pragma unroll
            for (int u = 0; u < 7; ++u)
                if (tw + u * R < hi) acc[u] += part[(tw + u * R) * n_par + col];
        }
        s[k] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        __syncthreads();
        if (k < C) {
            double t = 0.0;
            for (uint64_t r = 0; r < R; ++r) t += s[k + r * C];
            grad[i * n_par + col] = t;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void rt_adjoint_reduce_parts(const double* part, double* grad) {
    const uint64_t j = blockIdx.x;
    double v = (int)threadIdx.x < kReduceChunks ? part[j * kReduceChunks + threadIdx.x] : 0.0;
    v = block_sum_256(v);
    if (threadIdx.x == 0) grad[j] = v;
}

// The tape, the per-wave gradient accumulators, the chunk partials and the
// row table of an adjoint, in its scheme's (or batch's) working buffers: a
// buffer only grows, and the one it replaced stays allocated for the owner's
// life (a graph captured with it may be replayed).  The row table is re-made
// in a new buffer when its contents change.
struct AdjBufs {
    double* tape;
    double* gacc;
    double* part;
    int* params;
};

int adj_buffers(WorkBufs* w, size_t tape, size_t gacc, size_t part, const std::vector<int>& rows, bool upload_rows,
                AdjBufs& out) {
    std::vector<int>& rows_key = w->params_key;
    int rc = w->ensure(WorkBufs::TAPE, tape * sizeof(double));
    if (rc == RT_OK) rc = w->ensure(WorkBufs::GACC, gacc * sizeof(double));
    if (rc == RT_OK && part) rc = w->ensure(WorkBufs::PART, part * sizeof(double));
    if (rc != RT_OK) return rc;
    if (upload_rows && (rows != rows_key || !w->p[WorkBufs::PARAMS])) {
        rows_key.clear();
        rc = w->fresh(WorkBufs::PARAMS, rows.size() * sizeof(int));
        if (rc != RT_OK) return rc;
        // read by the generic kernel only (the generated one has the rows in
        // its code): a synchronous upload, so no host buffer outlives the call
        HIP_TRY(hipMemcpy(w->p[WorkBufs::PARAMS], rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice),
                RT_ERR_FAILED);
        rows_key = rows;
    }
    out.tape = w->get<double>(WorkBufs::TAPE);
    out.gacc = w->get<double>(WorkBufs::GACC);
    out.part = w->get<double>(WorkBufs::PART);
    out.params = w->get<int>(WorkBufs::PARAMS);
    return RT_OK;
}

}  // namespace

// defined in rt_trace.hip
extern "C" int rt_hip_scheme_device_tables(const rt_hip_scheme* s, const double** table_dev, const double** pool_dev,
                                           int* n_surf, size_t* pool_len, const double** host_table);

namespace {

// Validation shared by the single-scheme and batched adjoints: beam type,
// plane gratings only, and the row table (the caller's index of each
// (surface, slot), -1 where not asked for).  rec(s) is surface s's record
// (public or device layout: the fields read here sit at the same offsets).
template <class Rec>
int adjoint_rows(Rec rec, int n_surf, const double* beam_desc, const double* group_seeds, size_t n_groups,
                 const int* param_surface, const int* param_field, int n_params, double* grad,
                 std::vector<int>& host) {
    if (!beam_desc) return rt_set_error(RT_ERR_BAD_VALUE, "the adjoint traces a generated beam: beam_desc is NULL");
    const int beam_type = (int)beam_desc[RT_B_TYPE];
    if (beam_type != RT_BEAM_PARALLEL && beam_type != RT_BEAM_CONIC && beam_type != RT_BEAM_GAUSS)
        return rt_set_error(RT_ERR_BAD_VALUE, "bad beam type");
    if (!group_seeds || n_groups == 0) return rt_set_error(RT_ERR_BAD_VALUE, "group seeds missing");
    for (int s = 0; s < n_surf; ++s) {
        const double* p = rec(s);
        if ((int)p[RT_S_TYPE] == RT_GRATING && (int)p[RT_S_CLASS] != RT_PLANE)
            return rt_set_error(RT_ERR_BAD_VALUE, "surface %d: the adjoint covers plane gratings only", s);
    }
    if (n_params < 0 || n_params > 4096 || (n_params > 0 && (!param_surface || !param_field || !grad)))
        return rt_set_error(RT_ERR_BAD_VALUE, "bad parameter list");
    host.assign((size_t)n_surf * rt::adj::PG_N, -1);
    for (int j = 0; j < n_params; ++j) {
        if (param_surface[j] < 0 || param_surface[j] >= n_surf)
            return rt_set_error(RT_ERR_BAD_VALUE, "parameter %d: surface %d out of range", j, param_surface[j]);
        int slot = rt::adj::slot_of_field(param_field[j]);
        if (slot < 0) return rt_set_error(RT_ERR_BAD_VALUE, "parameter %d: field %d is not differentiable", j,
                                          param_field[j]);
        if (slot >= rt::adj::PG_A && slot < rt::adj::PG_CR &&
            slot - rt::adj::PG_A >= (int)rec(param_surface[j])[RT_S_NPOLY])
            return rt_set_error(RT_ERR_BAD_VALUE, "parameter %d: aspheric term beyond the surface's N", j);
        int& row = host[(size_t)param_surface[j] * rt::adj::PG_N + slot];
        if (row >= 0) return rt_set_error(RT_ERR_BAD_VALUE, "parameter %d repeats parameter %d", j, row);
        row = j;
    }
    return RT_OK;
}

// beam fields + the host-evaluated uniforms of the direction generators
void fill_beam(AdjArgs& a, const double* beam_desc) {
    memcpy(a.beam, beam_desc, RT_BEAM_STRIDE * sizeof(double));
    double ang = a.beam[RT_B_ANGUNIT] != 0.0 ? a.beam[RT_B_ANGPAR] : a.beam[RT_B_ANGPAR] * rt::kPi / 180.0;
    a.beam[rt::BU_CONE_COS] = cos(ang / 2.0);                           // beam.cu:593-594
    a.beam[rt::BU_CONE_SIN] = sin(ang / 2.0);
    a.beam[rt::BU_GAUSS_TG] = tan(a.beam[RT_B_ANGPAR] / 2.35482 * rt::kPi / 180 / 3600);   // beam.cu:1034,1099
    a.beam[rt::BU_LAM_STEP] = rt::beam_lam_step(a.beam);
}

// minimum waves per SIMD asked of the generated kernels' register allocation:
// 3 (168 VGPRs) measured fastest on the echelle's 48 parameters against 2 and
// 4 and the compiler's own choice; RT_HIP_ADJOINT_WAVES overrides (0: the
// compiler's choice)
int adjoint_waves() {
    const char* w = getenv("RT_HIP_ADJOINT_WAVES");
    return w ? atoi(w) : 3;
}

// workgroups launched per resident slot (the persistent grid is CUs x
// resident x this, within the tape's 2 GiB buffer): 16 for batches, where the
// hardware's dealing of the extra workgroups balances instances of uneven
// cost (1024 echelle instances: 95.2 ms at 1, 85.7 at 8, 84.1 at 16), 1 for
// one scheme, where more gains nothing and the tape stays small;
// RT_HIP_ADJOINT_GRID_MULT overrides both
uint64_t adjoint_grid_mult(uint64_t dflt) {
    const char* m = getenv("RT_HIP_ADJOINT_GRID_MULT");
    const int v = m ? atoi(m) : (int)dflt;
    return v >= 1 && v <= 64 ? (uint64_t)v : dflt;
}

}  // namespace

extern "C" int rt_hip_adjoint(const rt_hip_scheme* sch, const double* beam_desc, uint64_t seed, size_t n_rays,
                              const double* group_seeds, size_t n_groups, const int* param_surface,
                              const int* param_field, int n_params, double* grad, unsigned long long* stats,
                              void* stream) {
    const double *table = nullptr, *pool = nullptr, *host_table = nullptr;
    int n_surf = 0;
    size_t pool_len = 0;
    int rc = rt_hip_scheme_device_tables(sch, &table, &pool, &n_surf, &pool_len, &host_table);
    if (rc != RT_OK) return rc;
    std::vector<int> host;
    rc = adjoint_rows([&](int s) { return host_table + (size_t)s * RT_SURF_STRIDE; }, n_surf, beam_desc, group_seeds,
                      n_groups, param_surface, param_field, n_params, grad, host);
    if (rc != RT_OK) return rc;
    const int beam_type = (int)beam_desc[RT_B_TYPE];
    if (n_rays == 0) {
        if (n_params) return zero_fill(grad, (uint64_t)n_params, (hipStream_t)stream);
        return RT_OK;
    }
    rc = rt_ensure_device();
    if (rc != RT_OK) return rc;

    AdjArgs a;
    memset(&a, 0, sizeof a);
    fill_beam(a, beam_desc);
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

    // the kernel generated for this structure and parameter list, else the generic one
    // (grid-wavelength beams: the generated kernel reads the per-wavelength table)
    const uint64_t nk = rt_scheme_dtab_nk(sch, beam_desc);
    hipFunction_t fn = nullptr;
    if (rt_jit_enabled())
        fn = rt_jit_adjoint_kernel(rt_scheme_signature(sch), beam_type, host, n_surf, kAdjBlock, adjoint_waves(),
                                   rt_scheme_device(sch), false, nk != 0);
    auto kern = beam_type == RT_BEAM_PARALLEL ? rt_adjoint_kernel<RT_BEAM_PARALLEL>
              : beam_type == RT_BEAM_CONIC    ? rt_adjoint_kernel<RT_BEAM_CONIC>
                                              : rt_adjoint_kernel<RT_BEAM_GAUSS>;
    const size_t lds = pool_len * sizeof(double);
    int resident = 0;
    hipError_t oe = fn ? hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&resident, fn, kAdjBlock, lds)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kern, kAdjBlock, lds);
    if (oe != hipSuccess || resident < 1) resident = 1;
    uint64_t want = (n_rays + kAdjBlock - 1) / kAdjBlock;
    uint64_t cap = (uint64_t)rt_device_cus() * resident * adjoint_grid_mult(1);
    // the tape (tape_rows rows: n_surf * kTape, plus the Newton iterates the
    // generated kernel records) and the accumulators (n_params rows) of G
    // lanes are addressed as buffer resources of at most 2^31 - 1 bytes
    a.tape_rows = fn ? (uint64_t)rt_jit_adjoint_tape_rows(rt_scheme_signature(sch)) : (uint64_t)n_surf * kTape;
    const uint64_t rows_max = std::max<uint64_t>(a.tape_rows, (uint64_t)std::max(n_params, 1));
    const uint64_t lanes_max = 0x7fffffffull / (rows_max * sizeof(double));
    if (cap > lanes_max / kAdjBlock) cap = lanes_max / kAdjBlock;
    if (cap < 1) return rt_set_error(RT_ERR_BAD_VALUE, "adjoint: too many parameters");
    const unsigned grid = (unsigned)(want < cap ? want : cap);
    const uint64_t G = (uint64_t)grid * kAdjBlock;

    AdjBufs B;
    rc = adj_buffers(rt_scheme_work(sch), (size_t)a.tape_rows * G, (size_t)std::max(n_params, 1) * (G / 64),
                     (size_t)std::max(n_params, 1) * kReduceChunks, host, !fn, B);
    if (rc != RT_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    a.prow = B.params;
    a.tape = B.tape;
    a.gacc = B.gacc;
    if (fn && nk) {
        rc = rt_scheme_dtab_fill(sch, beam_desc, nk, st, &a.ltab);
        if (rc != RT_OK) return rc;
        a.lt_nk = nk;
    }
    if (fn) {
        void* params[] = {&a};
        HIP_TRY(hipModuleLaunchKernel(fn, grid, 1, 1, kAdjBlock, 1, 1, (unsigned)lds, st, params, nullptr),
                RT_ERR_FAILED);
    } else {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kAdjBlock), lds, st, a);
    }
    HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    if (n_params > 0) {
        hipLaunchKernelGGL(rt_adjoint_reduce_chunks, dim3(kReduceChunks, n_params), dim3(256), 0, st,
                           (const double*)B.gacc, G / 64, B.part);
        HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
        hipLaunchKernelGGL(rt_adjoint_reduce_parts, dim3(n_params), dim3(256), 0, st, (const double*)B.part, grad);
        HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    }
    return RT_OK;
}

extern "C" int rt_hip_batch_adjoint(const rt_hip_batch* b, const double* beam_desc, uint64_t seed,
                                    const double* group_seeds, size_t n_groups, const int* param_surface,
                                    const int* param_field, int n_params, double* grad, unsigned long long* stats,
                                    void* stream) {
    BatchView v;
    int rc = rt_batch_view(b, &v);
    if (rc != RT_OK) return rc;
    std::vector<int> host;
    rc = adjoint_rows([&](int s) { return v.dtables + (size_t)s * rt::DEV_STRIDE; }, v.n_surf, beam_desc, group_seeds,
                      n_groups, param_surface, param_field, n_params, grad, host);
    if (rc != RT_OK) return rc;
    if ((uint64_t)v.n_inst * v.n_surf * rt::DEV_STRIDE >= (1ull << 32))
        return rt_set_error(RT_ERR_BAD_VALUE, "batched adjoint: too many instance records");
    if (v.block != kAdjBlock)
        return rt_set_error(RT_ERR_BAD_VALUE, "the batched adjoint takes workgroups of %d lanes (the batch has %d)",
                            kAdjBlock, v.block);
    const int beam_type = (int)beam_desc[RT_B_TYPE];
    rc = rt_ensure_device();
    if (rc != RT_OK) return rc;
    if (n_params == 0 || v.n_tiles == 0) {
        if (n_params) return zero_fill(grad, (uint64_t)v.n_inst * (uint64_t)n_params, (hipStream_t)stream);
        return RT_OK;
    }

    AdjArgs a;
    memset(&a, 0, sizeof a);
    fill_beam(a, beam_desc);
    a.table = v.tables_dev;
    a.pool = v.pool_dev;
    a.n_surf = v.n_surf;
    a.pool_len = (int)v.pool_len;
    a.n_par = n_params;
    a.seed = seed;
    a.n = v.ray_start[v.n_inst];
    a.n_groups = n_groups;
    a.seeds = group_seeds;
    a.stats = stats;
    a.tile_start = v.tile_start_dev;
    a.tile_inst = v.tile_inst_dev;
    a.ray_start = v.ray_start_dev;
    a.n_inst = (uint64_t)v.n_inst;
    a.n_tiles = v.n_tiles;
    a.inst_stride = (uint64_t)v.n_surf * rt::DEV_STRIDE;
    if (v.beams_dev) {
        if (beam_type != v.beam_type)
            return rt_set_error(RT_ERR_BAD_VALUE, "beam type %d differs from the instances' beams (%d)", beam_type,
                                v.beam_type);
        a.inst_beams = v.beams_dev;
    }

    const uint64_t nk = rt_batch_ltab_nk(b, beam_desc);
    hipFunction_t fn = nullptr;
    if (rt_jit_enabled())
        fn = rt_jit_adjoint_kernel(*v.signature, beam_type, host, v.n_surf, kAdjBlock, adjoint_waves(), v.device,
                                   true, nk != 0);
    auto kern = beam_type == RT_BEAM_PARALLEL ? rt_adjoint_batch_kernel<RT_BEAM_PARALLEL>
              : beam_type == RT_BEAM_CONIC    ? rt_adjoint_batch_kernel<RT_BEAM_CONIC>
                                              : rt_adjoint_batch_kernel<RT_BEAM_GAUSS>;
    const size_t lds = v.pool_len * sizeof(double);
    int resident = 0;
    hipError_t oe = fn ? hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&resident, fn, kAdjBlock, lds)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kern, kAdjBlock, lds);
    if (oe != hipSuccess || resident < 1) resident = 1;
    uint64_t cap = (uint64_t)rt_device_cus() * resident * adjoint_grid_mult(16);
    a.tape_rows = fn ? (uint64_t)rt_jit_adjoint_tape_rows(*v.signature) : (uint64_t)v.n_surf * kTape;
    const uint64_t lanes_max = 0x7fffffffull / (a.tape_rows * sizeof(double));
    if (cap > lanes_max / kAdjBlock) cap = lanes_max / kAdjBlock;
    const unsigned grid = (unsigned)(v.n_tiles < cap ? v.n_tiles : cap);
    const uint64_t G = (uint64_t)grid * kAdjBlock, wpg = kAdjBlock / 64;
    const uint64_t part_len = v.n_tiles * wpg * (uint64_t)n_params;
    if (part_len * sizeof(double) > 0x7fffffffull)
        return rt_set_error(RT_ERR_BAD_VALUE, "batched adjoint: %llu tiles x %d parameters exceed one buffer",
                            (unsigned long long)v.n_tiles, n_params);

    AdjBufs B;
    rc = adj_buffers(rt_batch_work(b), (size_t)a.tape_rows * G, part_len, 0, host, !fn, B);
    if (rc != RT_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    a.prow = B.params;
    a.tape = B.tape;
    a.gacc = B.gacc;
    if (fn && nk) {
        rc = rt_batch_ltab_fill(b, beam_desc, nk, st, &a.ltab);
        if (rc != RT_OK) return rc;
        a.lt_nk = nk;
    }
    if (fn) {
        void* params[] = {&a};
        HIP_TRY(hipModuleLaunchKernel(fn, grid, 1, 1, kAdjBlock, 1, 1, (unsigned)lds, st, params, nullptr),
                RT_ERR_FAILED);
    } else {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kAdjBlock), lds, st, a);
    }
    HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    if (n_params > 0) {
        hipLaunchKernelGGL(rt_adjoint_batch_reduce, dim3((unsigned)v.n_inst), dim3(256), 0, st,
                           (const double*)B.gacc, v.tile_start_dev, (uint64_t)n_params, wpg, grad);
        HIP_TRY(hipGetLastError(), RT_ERR_FAILED);
    }
    return RT_OK;
}
