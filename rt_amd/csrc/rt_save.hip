// rt_save.hip -- the saved beam of Scheme::run on the device: the order in
// which the reference's Beam::Rearrange calls leave the rays, and the state
// of the rays lost on the way, from the loss codes of a recording trace.
//
// Scheme::run (common-lib/scheme.cpp:983-1046) keeps the good rays at the
// front of the beam arrays: after every stage that loses rays (intersection,
// surface.cpp:256-259; aperture, :316-320 and :694-697; action, :467-471 and
// :579-583) Beam::Rearrange (beam.cpp:574-598) partitions [0, N_good) with
// two pointers: the k-th lost ray from the front is swapped with the k-th
// good ray from the back while the first lies before the second.  With G the
// good count before the stage and G' after it, those pairs are exactly the
// lost rays in [0, G') in ascending order and the good rays in [G', G) in
// descending order, so one stage is a prefix count of its lost rays and one
// swap per pair, all in parallel.  The stages run in Scheme::run's order.
//
// Beam::Save (beam.cpp:636-715) writes the arrays in that order, good rays
// first.  Transform and the engine's stages touch only [0, N_good), so a lost
// ray keeps the state it was lost in, but Beam::Normalize runs over every ray
// (beam.cpp:527-571): its cosines pass through the scaled-norm formula once
// more at every later Normalize call.  rt_hip_beam_order applies them, in
// the reference's operation order with no contraction, IEEE division and
// square root, so the lost rays' cosines come out as the reference's would
// from the same state.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "rt_device.h"
#include "rt_host.h"

namespace {

constexpr int kTileThreads = 256;
constexpr int kPerThread = 16;
constexpr uint64_t kTile = (uint64_t)kTileThreads * kPerThread;   // positions per tile

__global__ void rt_save_zero_kernel(unsigned long long* v, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = 0ull;
}

// counts of every loss code
__global__ __launch_bounds__(256) void rt_loss_hist_kernel(const unsigned char* loss, uint64_t n,
                                                           unsigned long long* counts) {
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(&h[loss[i]], 1u);
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&counts[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// position p holds ray p before the first stage
__global__ void rt_order_init_kernel(const unsigned char* loss, uint64_t n, uint64_t* perm, unsigned char* lpos) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        perm[p] = p;
        lpos[p] = loss[p];
    }
}

// lost rays of this stage (code) per tile of [0, G)
__global__ __launch_bounds__(kTileThreads) void rt_stage_count_kernel(const unsigned char* lpos, uint64_t G,
                                                                      unsigned char code, uint64_t* tile_bad) {
    __shared__ unsigned s[kTileThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kPerThread;
    unsigned c = 0;
    for (int k = 0; k < kPerThread; ++k) {
        const uint64_t p = base + k;
        c += (p < G && lpos[p] == code) ? 1u : 0u;
    }
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m);
    if (__lane_id() == 0) s[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (int w = 0; w < kTileThreads / 64; ++w) t += s[w];
        tile_bad[blockIdx.x] = t;
    }
}

// exclusive prefix over the tiles (one workgroup, chunks of 1024 with a carry)
__global__ __launch_bounds__(1024) void rt_tile_scan_kernel(uint64_t* tile_bad, uint64_t n_tiles) {
    __shared__ uint64_t s[1024];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint64_t b0 = 0; b0 < n_tiles; b0 += 1024) {
        const uint64_t b = b0 + threadIdx.x;
        const uint64_t v = b < n_tiles ? tile_bad[b] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {       // inclusive Hillis-Steele scan
            const uint64_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        const uint64_t c = carry;
        if (b < n_tiles) tile_bad[b] = c + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c + s[1023];
        __syncthreads();
    }
}

// per position of tile `tile`: lost rays of this stage before it (exclusive,
// counted from the start of [0, G), tile_off being the tile's offset) and
// whether it is lost; the workgroup is the tile.  Every read of the tile
// completes before the function returns (barriers of the scan).
__device__ __forceinline__ void tile_prefix(const unsigned char* lpos, uint64_t G, unsigned char code,
                                            uint64_t tile, uint64_t tile_off, uint64_t (&pre)[kPerThread],
                                            bool (&bad)[kPerThread]) {
    __shared__ unsigned s[kTileThreads];
    const uint64_t base = tile * kTile + (uint64_t)threadIdx.x * kPerThread;
    unsigned c = 0;
    for (int k = 0; k < kPerThread; ++k) {
        const uint64_t p = base + k;
        bad[k] = p < G && lpos[p] == code;
        c += bad[k] ? 1u : 0u;
    }
    s[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < kTileThreads; off <<= 1) {
        const unsigned t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    uint64_t run = tile_off + s[threadIdx.x] - c;
    for (int k = 0; k < kPerThread; ++k) {
        pre[k] = run;
        run += bad[k] ? 1u : 0u;
    }
}

// good rays of [G', G): the k-th from the back (k = 1, 2, ...) at gpos[k - 1];
// workgroup b takes tile first_tile + b
__global__ __launch_bounds__(kTileThreads) void rt_stage_tail_kernel(const unsigned char* lpos, uint64_t G,
                                                                     uint64_t Gn, uint64_t nb, unsigned char code,
                                                                     const uint64_t* tile_off, uint64_t first_tile,
                                                                     uint64_t* gpos) {
    const uint64_t tile = first_tile + blockIdx.x;
    uint64_t pre[kPerThread];
    bool bad[kPerThread];
    tile_prefix(lpos, G, code, tile, tile_off[tile], pre, bad);
    const uint64_t base = tile * kTile + (uint64_t)threadIdx.x * kPerThread;
    for (int k = 0; k < kPerThread; ++k) {
        const uint64_t q = base + k;
        if (q < Gn || q >= G || bad[k]) continue;
        // lost rays after q: all of them but the pre[k] before q (q is good)
        const uint64_t bad_after = nb - pre[k];
        const uint64_t good_after = (G - 1 - q) - bad_after;
        gpos[good_after] = q;
    }
}

// lost rays of [0, G'): the k-th (k = 1, 2, ...) swaps places with gpos[k - 1]
__global__ __launch_bounds__(kTileThreads) void rt_stage_head_kernel(unsigned char* lpos, uint64_t G, uint64_t Gn,
                                                                     unsigned char code, const uint64_t* tile_off,
                                                                     const uint64_t* gpos, uint64_t* perm) {
    uint64_t pre[kPerThread];
    bool bad[kPerThread];
    tile_prefix(lpos, G, code, blockIdx.x, tile_off[blockIdx.x], pre, bad);
    const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kPerThread;
    for (int k = 0; k < kPerThread; ++k) {
        const uint64_t p = base + k;
        if (p >= Gn || !bad[k]) continue;
        const uint64_t q = gpos[pre[k]];
        const uint64_t rp = perm[p], rq = perm[q];
        perm[p] = rq;
        perm[q] = rp;
        const unsigned char lq = lpos[q];
        lpos[q] = lpos[p];
        lpos[p] = lq;
    }
}

// Beam::Normalize on one ray (beam.cpp:541-569): the scaled norm, in the
// reference's operation order with no contraction
__device__ __attribute__((noinline)) void normalize_reference(double& cX, double& cY, double& cZ) {
#pragma clang fp contract(off)
    const double acX = fabs(cX), acY = fabs(cY), acZ = fabs(cZ);
    double norm;
    if (acX >= acY) {
        if (acX >= acZ) norm = acX * sqrt(1.0 + acY * acY / acX / acX + acZ * acZ / acX / acX);
        else norm = acZ * sqrt(1.0 + acX * acX / acZ / acZ + acY * acY / acZ / acZ);
    } else {
        if (acY >= acZ) norm = acY * sqrt(1.0 + acX * acX / acY / acY + acZ * acZ / acY / acY);
        else norm = acZ * sqrt(1.0 + acX * acX / acZ / acZ + acY * acY / acZ / acZ);
    }
    cX /= norm;
    cY /= norm;
    cZ /= norm;
}

struct OrderArgs {
    const double* in[7];
    double* out;
    const uint64_t* perm;
    const unsigned char* loss;
    uint64_t n;
    unsigned char renorm[256];
};

__global__ __launch_bounds__(256) void rt_beam_order_kernel(OrderArgs a) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.n; p += stride) {
        const uint64_t r = a.perm[p];
        double v[7];
        for (int k = 0; k < 7; ++k) v[k] = a.in[k][r];
        const int m = a.renorm[a.loss[r]];
        for (int j = 0; j < m; ++j) normalize_reference(v[3], v[4], v[5]);
        for (int k = 0; k < 7; ++k) a.out[(uint64_t)k * a.n + p] = v[k];
    }
}

unsigned grid_for(uint64_t n, unsigned block) {
    const uint64_t want = (n + block - 1) / block;
    const uint64_t cap = (uint64_t)rt_device_cus() * 16;
    return (unsigned)std::max<uint64_t>(1, std::min(want, cap));
}

struct SaveWork {
    int device = -1;
    WorkBufs w;
    ~SaveWork() { w.release(); }
};
thread_local SaveWork g_save_work;

}  // namespace

extern "C" int rt_hip_rearrange(const unsigned char* loss_dev, size_t n, uint64_t* perm_dev, uint64_t* counts,
                                void* stream) {
    if (n > 0 && (!loss_dev || !perm_dev)) return rt_set_error(RT_ERR_BAD_VALUE, "NULL loss or perm buffer");
    if (!counts) return rt_set_error(RT_ERR_BAD_VALUE, "counts is NULL");
    memset(counts, 0, 256 * sizeof(uint64_t));
    if (n == 0) return RT_OK;
    int rc = rt_ensure_device();
    if (rc != RT_OK) return rc;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return rt_set_error(RT_ERR_FAILED, "hipGetDevice failed");
    if (g_save_work.device != dev) {
        g_save_work.w.release();
        g_save_work.device = dev;
    }
    WorkBufs& w = g_save_work.w;
    const uint64_t n_tiles = (n + kTile - 1) / kTile;
    if ((rc = w.ensure(WorkBufs::SAVE_HIST, 256 * sizeof(unsigned long long))) != RT_OK ||
        (rc = w.ensure(WorkBufs::SAVE_LPOS, n)) != RT_OK ||
        (rc = w.ensure(WorkBufs::SAVE_TILES, n_tiles * sizeof(uint64_t))) != RT_OK ||
        (rc = w.ensure(WorkBufs::SAVE_GPOS, n * sizeof(uint64_t))) != RT_OK)
        return rc;
    hipStream_t st = (hipStream_t)stream;
    unsigned long long* hist = w.get<unsigned long long>(WorkBufs::SAVE_HIST);
    unsigned char* lpos = w.get<unsigned char>(WorkBufs::SAVE_LPOS);
    uint64_t* tiles = w.get<uint64_t>(WorkBufs::SAVE_TILES);
    uint64_t* gpos = w.get<uint64_t>(WorkBufs::SAVE_GPOS);

    hipLaunchKernelGGL(rt_save_zero_kernel, dim3(1), dim3(256), 0, st, hist, 256);
    hipLaunchKernelGGL(rt_loss_hist_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, loss_dev, (uint64_t)n, hist);
    hipLaunchKernelGGL(rt_order_init_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, loss_dev, (uint64_t)n,
                       perm_dev, lpos);
    if (hipGetLastError() != hipSuccess) return rt_set_error(RT_ERR_FAILED, "rearrange launch failed");
    if (hipMemcpyAsync(counts, hist, 256 * sizeof(uint64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return rt_set_error(RT_ERR_FAILED, "rearrange: reading the loss counts failed");

    // stages in Scheme::run's order: codes 1, 2, 3, ... (kLossDead rays never
    // were good and sit outside [0, G) from the start)
    uint64_t G = n - counts[rt::kLossDead];
    if (counts[rt::kLossDead]) return rt_set_error(RT_ERR_BAD_VALUE, "rays that entered dead cannot be ordered");
    for (int code = 1; code < rt::kLossDead; ++code) {
        const uint64_t nb = counts[code];
        if (nb == 0) continue;
        const uint64_t Gn = G - nb;
        const uint64_t tiles_G = (G + kTile - 1) / kTile;
        hipLaunchKernelGGL(rt_stage_count_kernel, dim3((unsigned)tiles_G), dim3(kTileThreads), 0, st,
                           (const unsigned char*)lpos, G, (unsigned char)code, tiles);
        hipLaunchKernelGGL(rt_tile_scan_kernel, dim3(1), dim3(1024), 0, st, tiles, tiles_G);
        const uint64_t t0 = Gn / kTile;
        hipLaunchKernelGGL(rt_stage_tail_kernel, dim3((unsigned)(tiles_G - t0)), dim3(kTileThreads), 0, st,
                           (const unsigned char*)lpos, G, Gn, nb, (unsigned char)code, (const uint64_t*)tiles, t0,
                           gpos);
        const uint64_t tiles_Gn = (Gn + kTile - 1) / kTile;
        if (tiles_Gn)
            hipLaunchKernelGGL(rt_stage_head_kernel, dim3((unsigned)tiles_Gn), dim3(kTileThreads), 0, st, lpos, G, Gn,
                               (unsigned char)code, (const uint64_t*)tiles, (const uint64_t*)gpos, perm_dev);
        if (hipGetLastError() != hipSuccess) return rt_set_error(RT_ERR_FAILED, "rearrange stage launch failed");
        G = Gn;
    }
    return RT_OK;
}

extern "C" int rt_hip_beam_order(size_t n, const uint64_t* perm_dev, const unsigned char* loss_dev,
                                 const unsigned char* renorm, const double* X, const double* Y, const double* Z,
                                 const double* cX, const double* cY, const double* cZ, const double* lambda,
                                 double* out_dev, void* stream) {
    if (n == 0) return RT_OK;
    if (!perm_dev || !loss_dev || !renorm || !X || !Y || !Z || !cX || !cY || !cZ || !lambda || !out_dev)
        return rt_set_error(RT_ERR_BAD_VALUE, "NULL argument");
    int rc = rt_ensure_device();
    if (rc != RT_OK) return rc;
    OrderArgs a;
    memset(&a, 0, sizeof a);
    a.in[0] = X; a.in[1] = Y; a.in[2] = Z; a.in[3] = cX; a.in[4] = cY; a.in[5] = cZ; a.in[6] = lambda;
    a.out = out_dev;
    a.perm = perm_dev;
    a.loss = loss_dev;
    a.n = n;
    memcpy(a.renorm, renorm, 256);
    hipLaunchKernelGGL(rt_beam_order_kernel, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, a);
    if (hipGetLastError() != hipSuccess) return rt_set_error(RT_ERR_FAILED, "beam order launch failed");
    return RT_OK;
}
