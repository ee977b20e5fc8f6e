// ctfp_device.hip -- gfx950 kernels and the device half of the C ABI (include/ctfp.h).
//
// Data layout in HBM
//   volume : the n_mat density channels in groups of up to four, each group
//            interleaved per voxel ([z][y][x][m]; 16-bit storage holds the
//            x-pair (v(x), v(x+1)) per dword), padded with two zero voxels on
//            both sides of every axis so the 8-corner gather of any integer
//            cell index clamped to [-2, N] stays in bounds and reads only
//            zeros outside the support; read with raw buffer loads.
//   occupancy: per group, one byte per brick of 8^3 padded cells (any
//            non-zero cell), for skipping chunks that would add only zeros.
//   weights: W[e][v][u] = I0Val * S_e * T_e(u) * eta_e(v,u) for the active
//            (non-zero) bins; the v stride is 0 when eta does not vary per pixel.
//   outputs: I, IAbsorb [view][v][u]; optional L [view][m][v][u].
//
// The march evaluates the reference's sample lattice t_k = SOD - R + dx/2 + k dx
// (CBCTPolyProjGPU.cu:1846-1848) with zero-padded trilinear interpolation
// (CBCTPolyProjGPU.cu:1920, 2200-2205).  Per-ray geometry is set up in FP64:
// the sample index position is affine in k, so each ray carries its FP64-derived
// start and step and its support interval [k_lo, k_hi) where any channel can
// be non-zero; samples outside it are exact zeros in the reference and are
// skipped.  The start and step become 32.32 fixed point: every 16th sample is
// an anchor reached by exact integer steps, and the 16 samples after it are
// FP32 offsets from it (error ~3e-6 voxel, against ~1e-4 for the reference's
// FP32 t * d - SOD), so no FP64 runs in the march itself.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "ctfp.h"
#include "ctfp_internal.h"

using ctfp::set_error;

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return set_error(e_ == hipErrorOutOfMemory ? CTFP_ERR_NOMEM : CTFP_ERR_DEVICE,  \
                             "%s failed: %s", #expr, hipGetErrorString(e_));                \
    } while (0)

// --------------------------------------------------------------- context ---

// Materials per scan, and channels per interleaved volume group (one march
// handles up to four).
constexpr int kMaxMaterials = CTFP_MAX_MATERIALS;
constexpr int kGroupMats = 4;
constexpr int kMaxGroups = (kMaxMaterials + kGroupMats - 1) / kGroupMats;

struct ctfp_context {
    int device = 0;
    ctfp_scan_params scan{};
    ctfp_system_info sys{};
    // volume: the n_mat channels in groups of up to four, each group its own
    // interleaved layout (one march per group)
    int n_mat = 0;
    int dtype = CTFP_F32;
    void *vol = nullptr;
    size_t vol_bytes = 0;
    int64_t px = 0, py = 0, pz = 0; // padded extents
    int n_groups = 0;
    int group_mats[kMaxGroups] = {};  // channels of each group
    int group_moff[kMaxGroups] = {};  // its first channel
    size_t group_off[kMaxGroups] = {}; // its byte offset in vol
    size_t group_bytes[kMaxGroups] = {};
    int *d_flag = nullptr; // device flag: a source voxel was not finite
    // occupancy of each group's layout: one byte per brick of 8^3 padded
    // cells, non-zero when a cell of the brick holds a non-zero channel
    uint8_t *occ = nullptr;
    int64_t ob_x = 0, ob_y = 0, ob_z = 0; // bricks per axis
    int32_t *obox = nullptr; // [n_groups][6] bounds of the occupied bricks
    // per group, the Chebyshev distance in bricks from each brick to the
    // nearest occupied one (capped at kDistCap), then one scratch grid:
    // [n_groups + 1][bricks]
    uint8_t *odist = nullptr;
    // line-integral workspace of grouped forwards and adjoints
    float *l_ws = nullptr;
    size_t l_ws_bytes = 0;
    // spectral
    int n_act = 0;
    int spec_n_mat = 0;  // materials the spectral tables were built for
    float *A = nullptr;  // [n_act][n_mat]
    _Float16 *Aq = nullptr; // [n_act rounded up to 32][16]: MFMA operand of 2^s A
    float aq_inv_scale = 1.0f; // 2^-s
    float *W = nullptr;  // [n_act][nv or 1][nu]
    int64_t w_stride_e = 0;
    int32_t w_stride_v = 0;
    float *I0 = nullptr; // [nv][nu]
    // per-view rotation (cos, sin) of the reference's FP32 angle
    double2 *view_cs = nullptr;
    // per-view focal-spot source offsets (mm), or null
    double2 *focal = nullptr;
    // host-mode output staging
    float *dI = nullptr, *dIA = nullptr;
    int64_t staged_views = 0;
    // launch tuning
    ctfp_tuning tune{CTFP_MARCH_LDS, 16, 40 * 1024, 8, 2, CTFP_ADJOINT_CELL, 1};
    // dl/dL workspace of the gather adjoint
    float *adj_ws = nullptr;
    size_t adj_ws_bytes = 0;
    // corner sums of the cell gather, [brick][M][corners]
    float *cell_part = nullptr;
    size_t cell_part_bytes = 0;
    // bumped by every change of volumes or spectral tables (batches check it)
    uint64_t generation = 0;
};

// The workspace of ctfp_backward_noalloc: the gather adjoint's dl/dL of one
// group of views, in a static device array (one copy per GPU, held until the
// process ends) so that the call allocates nothing.  Host code never passes
// `scratch` itself to a kernel (on the host it names a stand-in, not device
// memory): static_scratch() resolves its device address per GPU.
constexpr size_t kStaticScratchBytes = (size_t)CTFP_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) float scratch[kStaticScratchBytes / sizeof(float)];

namespace {

// A roctx range over one entry point (visible to rocprofv3 --marker-trace;
// a call that returns at once when no profiler is attached).
struct Range {
    explicit Range(const char *name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
};

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

template <typename T> void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Device allocations that still succeed before every allocation fails as out
// of memory (ctfp_debug_fail_allocations; -1: no failures).  Process-wide.
std::atomic<long long> g_alloc_budget{-1};

// hipMalloc through the failure hook.
template <typename T> hipError_t dev_malloc(T **p, size_t bytes)
{
    long long b = g_alloc_budget.load();
    while (b >= 0) {
        if (b == 0) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
        if (g_alloc_budget.compare_exchange_weak(b, b - 1)) break;
    }
    return hipMalloc(p, bytes);
}

// The status code of a failed HIP call.
int hip_status(hipError_t e) { return e == hipErrorOutOfMemory ? CTFP_ERR_NOMEM : CTFP_ERR_DEVICE; }

// ------------------------------------------------------------- kernels ---

constexpr int kVoxPadLo = 2; // zero voxels below index 0 on each axis
constexpr int kVoxPadHi = 2; // zero voxels above index N-1 on each axis

// The FP32 source volumes [z][y][x] (device pointers) of one group of up to
// four channels.
struct VolPtrs {
    const float *p[4];
};

// Source voxel of channel m, zero outside the volume; a value that is not
// finite raises *bad (when given).
__device__ __forceinline__ float src_voxel(const VolPtrs &src, int m, int64_t x, int64_t y, int64_t z,
                                           int64_t nx, int64_t ny, int64_t nz, int *bad)
{
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0.0f;
    const float v = src.p[m][(z * ny + y) * nx + x];
    if (bad && !isfinite(v)) *bad = 1;
    return v;
}

// FP32 resident layout: channels interleaved per voxel, zero padding.
__global__ void layout_voxels_f32(VolPtrs src, float *__restrict__ dst, int n_mat, int64_t nx,
                                  int64_t ny, int64_t nz, int64_t px, int64_t py, int64_t pz, int *bad)
{
    const int64_t total = px * py * pz;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t x = i % px - kVoxPadLo, y = (i / px) % py - kVoxPadLo, z = i / (px * py) - kVoxPadLo;
        for (int m = 0; m < n_mat; m++) dst[i * n_mat + m] = src_voxel(src, m, x, y, z, nx, ny, nz, bad);
    }
}

// 32-bit integer hash (xor-shift-multiply; the "lowbias32" constants of
// C. Wellons' hash search).
__device__ __forceinline__ uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Dither threshold in [0, 1) of voxel (x, y, z) >= 0 of channel `mat`,
// stratified over 2 x 2 x 2 blocks: the block's eight voxels take the ranks
// 0..7 of a pseudo-random permutation plus one shared uniform offset, so each
// block rounds up close to 8 f of its voxels when all hold a fraction f of a
// step (an unbiased rounding whose error nearly cancels within a block).
__device__ __forceinline__ float dither_threshold(int64_t x, int64_t y, int64_t z, int mat)
{
    const uint32_t key =
        hash32((uint32_t)(x >> 1) * 0x9E3779B1u ^
               hash32((uint32_t)(y >> 1) * 0x85EBCA77u ^ hash32((uint32_t)(z >> 1) * 0xC2B2AE3Du ^
                                                                  (uint32_t)mat * 0x27D4EB2Fu)));
    const uint32_t self = (uint32_t)((x & 1) | ((y & 1) << 1) | ((z & 1) << 2));
    const uint32_t hs = hash32(key + self);
    uint32_t rank = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) {
        const uint32_t hj = hash32(key + j);
        rank += (hj < hs || (hj == hs && j < self)) ? 1u : 0u;
    }
    const float u = (float)(hash32(key ^ 0xA511E9B3u) >> 8) * 0x1p-24f;
    return ((float)rank + u) * 0.125f;
}

// v rounded to one of its two neighbouring BF16 values: the one toward zero,
// or the next one away from zero when the fraction of the step beyond the
// first exceeds the threshold t (exact values stay exact).  Over uniform
// thresholds the mean is v; the error stays below one step (2^-7 relative).
template <bool kBF16> __device__ __forceinline__ uint32_t quantize16(float v, float t)
{
    static_assert(kBF16, "FP16 storage rounds to nearest");
    const uint32_t b = __float_as_uint(v), lo = b & 0xFFFF0000u;
    if (lo == b || !isfinite(v)) return b >> 16;
    const float flo = __uint_as_float(lo), fhi = __uint_as_float(lo + 0x10000u);
    const float f = (v - flo) / (fhi - flo);
    return (f > t ? lo + 0x10000u : lo) >> 16;
}

// 16-bit resident layout: cell (x, y, z) holds, per channel, the pair
// (v(x), v(x+1)) in one dword, so one row load serves both x corners.
// FP16 rounds to nearest (|error| <= 2^-11 relative).  BF16 (8 significant
// bits) rounds each voxel up or down by the stratified dither, the same way
// in both cells that hold it: a uniform insert then keeps its mean along a
// ray instead of one systematic offset of up to 2^-8 relative.
// mat0: index of the group's first channel in the scan (seeds its dither).
template <bool kBF16>
__global__ void layout_pairs_16(VolPtrs src, uint32_t *__restrict__ dst, int n_mat, int mat0, int64_t nx,
                                int64_t ny, int64_t nz, int64_t px, int64_t py, int64_t pz, int *bad)
{
    const int64_t total = px * py * pz;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t x = i % px - kVoxPadLo, y = (i / px) % py - kVoxPadLo, z = i / (px * py) - kVoxPadLo;
        for (int m = 0; m < n_mat; m++) {
            const float lo = src_voxel(src, m, x, y, z, nx, ny, nz, bad);
            const float hi = src_voxel(src, m, x + 1, y, z, nx, ny, nz, nullptr);
            uint32_t bits;
            if constexpr (kBF16) {
                const uint32_t qlo = lo == 0.0f ? 0u : quantize16<true>(lo, dither_threshold(x, y, z, mat0 + m));
                const uint32_t qhi =
                    hi == 0.0f ? 0u : quantize16<true>(hi, dither_threshold(x + 1, y, z, mat0 + m));
                bits = (qlo & 0xFFFFu) | (qhi << 16);
            } else {
                bits = __builtin_bit_cast(uint32_t, __floats2half2_rn(lo, hi));
            }
            dst[i * n_mat + m] = bits;
        }
    }
}

// Occupancy of a group's resident layout: one byte per brick of
// kOccBrick^3 padded cells, 1 when any dword of any cell in it is non-zero
// (a 16-bit cell holds v(x) and v(x+1)).  One workgroup per brick at a
// time, bricks strided by the grid (the grid stays below the launch limit for
// any volume).
constexpr int kOccBrick = 8;
constexpr int kBlock256Waves = 4; // wavefronts of a 256-thread workgroup
constexpr int64_t kOccGrid = 1 << 20;
__global__ __launch_bounds__(256) void occupancy_kernel(const uint32_t *__restrict__ vol, int words, int64_t px,
                                                        int64_t py, int64_t pz, int64_t obx, int64_t oby,
                                                        int64_t obn, uint8_t *__restrict__ occ)
{
    for (int64_t b = blockIdx.x; b < obn; b += gridDim.x) { // block-uniform
        const int64_t bx = b % obx, by = (b / obx) % oby, bz = b / (obx * oby);
        int any = 0;
        for (int c = threadIdx.x; c < kOccBrick * kOccBrick * kOccBrick; c += blockDim.x) {
            const int64_t x = bx * kOccBrick + (c % kOccBrick), y = by * kOccBrick + (c / kOccBrick) % kOccBrick;
            const int64_t z = bz * kOccBrick + c / (kOccBrick * kOccBrick);
            if (x >= px || y >= py || z >= pz) continue;
            const uint32_t *cell = vol + ((z * py + y) * px + x) * words;
            for (int w = 0; w < words; w++) any |= cell[w] != 0u;
        }
        any = __syncthreads_or(any);
        if (threadIdx.x == 0) occ[b] = any ? 1 : 0;
        __syncthreads(); // the reduction's scratch is free for the next brick
    }
}

// Brick distances: dist(b) = min over occupied bricks o of max_axis |b - o|,
// capped at kDistCap.  The cap-limited distance separates into three 1-D
// passes, each dst(i) = min over |j - i| <= kDistCap along the axis of
// max(|i - j|, src(j)) (src of the first pass: 0 where occupied, the cap
// elsewhere).  Bricks strided by the grid.
constexpr int kDistCap = 15;
__global__ __launch_bounds__(256) void brick_distance_kernel(const uint8_t *__restrict__ src,
                                                             uint8_t *__restrict__ dst, int64_t obx, int64_t oby,
                                                             int64_t obz, int axis, int from_occ)
{
    const int64_t obn = obx * oby * obz;
    const int64_t n = axis == 0 ? obx : axis == 1 ? oby : obz;
    const int64_t step = axis == 0 ? 1 : axis == 1 ? obx : obx * oby;
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < obn; b += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = axis == 0 ? b % obx : axis == 1 ? (b / obx) % oby : b / (obx * oby);
        const int64_t j0 = i - kDistCap < 0 ? 0 : i - kDistCap, j1 = i + kDistCap >= n ? n - 1 : i + kDistCap;
        int best = kDistCap;
        for (int64_t j = j0; j <= j1; j++) {
            const uint8_t v = src[b + (j - i) * step];
            const int f = from_occ ? (v ? 0 : kDistCap) : (int)v;
            const int d = (int)(j > i ? j - i : i - j);
            best = min(best, max(d, f));
        }
        dst[b] = (uint8_t)best;
    }
}

// Bounds of the occupied bricks of every group: box[6 g + (0..5)] = min / max
// brick x, y, z (INT_MAX / INT_MIN when no brick is occupied).
__global__ void occupied_box_init(int32_t *box, int n)
{
    if ((int)threadIdx.x < n) box[threadIdx.x] = (threadIdx.x & 1) ? INT_MIN : INT_MAX;
}

__global__ __launch_bounds__(256) void occupied_box_kernel(const uint8_t *__restrict__ occ, int64_t obx,
                                                           int64_t oby, int64_t obn, int32_t *box)
{
    __shared__ int32_t part[kBlock256Waves][6];
    int32_t v[6] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN};
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < obn; b += (int64_t)gridDim.x * blockDim.x) {
        if (!occ[b]) continue;
        const int32_t c[3] = {(int32_t)(b % obx), (int32_t)((b / obx) % oby), (int32_t)(b / (obx * oby))};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            v[2 * a] = min(v[2 * a], c[a]);
            v[2 * a + 1] = max(v[2 * a + 1], c[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
        for (int sh = 32; sh >= 1; sh >>= 1) {
            const int32_t o = __shfl_xor(v[a], sh);
            v[a] = (a & 1) ? max(v[a], o) : min(v[a], o);
        }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int a = 0; a < 6; a++) part[threadIdx.x >> 6][a] = v[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        int32_t r = part[0][a];
        for (int w = 1; w < kBlock256Waves; w++) r = (a & 1) ? max(r, part[w][a]) : min(r, part[w][a]);
        if (a & 1) {
            if (r != INT_MIN) atomicMax(box + a, r);
        } else if (r != INT_MAX) {
            atomicMin(box + a, r);
        }
    }
}

// Density channels [m][z][y][x] from a label volume.
struct Densities {
    float d[kMaxMaterials];
};
__global__ void labels_to_raw(const uint8_t *__restrict__ labels, float *__restrict__ raw, int n_mat,
                              int single_material, Densities rho, int64_t nvox)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nvox;
         i += (int64_t)gridDim.x * blockDim.x) {
        unsigned lab = labels[i];
        if (single_material) {
            raw[i] = (float)lab * rho.d[0];
        } else {
            for (int m = 0; m < n_mat; m++) raw[m * nvox + i] = (lab == (unsigned)(m + 1)) ? rho.d[m] : 0.0f;
        }
    }
}

// W[e][v][u] = ((I0Val * S_e) * T_e(u)) * eta_e(v,u) for active bins; the same
// product order as the reference's addI0 (CBCTPolyForwardProjGrid.cpp:336).
struct WeightArgs {
    const float *S;     // [n_act] spectrum of active bins (already energy-weighted)
    const float *T;     // [n_act][nu] or null
    const float *eta;   // [n_act][npix] or null
    const float *mu_c;  // [n_act] or null (scintillator model)
    const float *thick; // [npix]
    float factor;
    float i0;
    int n_act;
    int64_t nu, rows; // rows = nv when per pixel, else 1
};
__global__ void build_weights(WeightArgs a, float *__restrict__ W)
{
    const int64_t per = a.rows * a.nu;
    const int64_t total = per * a.n_act;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int e = (int)(i / per);
        int64_t p = i % per;
        int64_t u = p % a.nu;
        float w = a.i0 * a.S[e];
        if (a.T) w = w * a.T[(int64_t)e * a.nu + u];
        if (a.eta) w = w * a.eta[(int64_t)e * per + p];
        else if (a.mu_c) w = w * (a.factor * (1.0f - expf(-a.mu_c[e] * a.thick[p])));
        W[i] = w;
    }
}

// I0[v][u] = sum_e W[e][v][u] in bin order.
// Compensated (Kahan) FP32 sum: c keeps the rounding error of each addition,
// so a sum of ~100 positive terms is accurate to a few FP32 ulps of the total
// (a plain FP32 sum can lose ~n ulps, visible in ln(I0/I) of thin rays).
struct KSum {
    float s, c;
};
__device__ __forceinline__ KSum ksum_zero() { return KSum{0.0f, 0.0f}; }
__device__ __forceinline__ void ksum_add(KSum &k, float x)
{
#pragma clang fp contract(off)
    const float y = x - k.c;
    const float t = k.s + y;
    k.c = (t - k.s) - y;
    k.s = t;
}
__device__ __forceinline__ float ksum_value(const KSum &k) { return k.s - k.c; }

__global__ void flat_field_kernel(const float *__restrict__ W, int n_act, int64_t stride_e,
                                  int stride_v, int64_t nu, int64_t nv, float *__restrict__ I0)
{
    const int64_t npix = nu * nv;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npix;
         p += (int64_t)gridDim.x * blockDim.x) {
        int64_t u = p % nu, v = p / nu;
        KSum acc = ksum_zero(); // bin order, compensated like the per-ray sums
        for (int e = 0; e < n_act; e++) ksum_add(acc, W[e * stride_e + v * stride_v + u]);
        I0[p] = ksum_value(acc);
    }
}

__global__ void postprocess_kernel(int64_t n_views, int64_t npix, const float *__restrict__ I,
                                   const float *__restrict__ IA, const float *__restrict__ I0,
                                   uint32_t *__restrict__ Iu, uint32_t *__restrict__ IAu,
                                   float *__restrict__ proj)
{
    const int64_t total = n_views * npix;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        float i = I[q];
        uint32_t iu = (uint32_t)floorf(i);
        if (Iu) Iu[q] = iu;
        if (IAu) IAu[q] = (uint32_t)floorf(IA[q]);
        if (proj) proj[q] = logf(I0[q % npix] / (float)iu);
    }
}

__global__ void copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------ the march ---

struct MarchArgs {
    // volume (padded, interleaved, bytes)
    const void *vol;
    uint32_t vol_bytes;    // buffer-load range (all of it unless wide)
    uint32_t row_y, row_z; // byte strides of one y row / one z slice
    int32_t wide;          // the volume spans 4 GiB or more
    float nx, ny, nz;      // clamp limits (= N per axis)
    // geometry (FP64, reference scalars)
    double sod, sdd, dsize, half_u, half_v, t0, dx;
    double half_x, half_y, half_z, inv_px, inv_py, inv_pz;
    int32_t nu, nv, K;
    const double2 *view_cs;
    const double2 *focal; // per-view source offsets (mm, detector u / v directions), or null
    int64_t view_begin;
    int32_t n_views;
    int32_t tiles_u, tiles_v;
    uint32_t n_blocks; // logical blocks (n_views * tiles_u * tiles_v)
    uint32_t per_xcd;  // launched blocks / 8
    float dx_f;
    // spectrum
    int32_t n_act;
    const float *A; // [n_act][M]
    const _Float16 *Aq; // [n_act rounded up to 32][16]: 2^s A split into FP16 hi / lo (MFMA operand)
    float aq_inv_scale; // 2^-s
    const float *W;
    int64_t w_stride_e;
    int32_t w_stride_v;
    // outputs; L rows [view][l_mats][v][u], this launch's channels from l_moff
    float *I, *IA, *L;
    int32_t l_mats, l_moff;
    // LDS staging (march_lds_kernel)
    int32_t chunk;     // samples per staged chunk
    int32_t lds_cells; // staged cells that fit the LDS budget
    // empty chunks: occupancy of 8^3-cell bricks of the padded volume (null:
    // no skipping), its bricks along x and y, and the padded cells per axis
    const uint8_t *occ;
    int32_t ob_x, ob_y;
    int32_t pcx, pcy, pcz;
    // bounds of the occupied bricks (x lo, x hi, y lo, y hi, z lo, z hi, in
    // bricks), or null: the LDS march clips each ray's window to them
    const int32_t *obox;
    // Chebyshev distance (in bricks, capped) from each brick to the nearest
    // occupied one, or null: the LDS march also steps each ray's window ends
    // through it
    const uint8_t *odist;
    // ray tile: wavefronts of (1 << wu_log2) x (64 >> wu_log2) pixels,
    // (1 << bu_log2) x (4 >> bu_log2) wavefronts per workgroup
    int32_t wu_log2, bu_log2;
};

constexpr int kBlock = 256; // 4 wavefronts per workgroup
// Zero rows after the last active bin of the spectral weight table.
constexpr int kWeightPadBins = 8;

template <int M>
__device__ __forceinline__ void load_vox(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                         float (&c)[M])
{
    if constexpr (M == 1) {
        c[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    } else if constexpr (M == 2) {
        auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        c[0] = __uint_as_float(v[0]);
        c[1] = __uint_as_float(v[1]);
    } else if constexpr (M == 3) {
        auto v = __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, 0);
        c[0] = __uint_as_float(v[0]);
        c[1] = __uint_as_float(v[1]);
        c[2] = __uint_as_float(v[2]);
    } else {
        auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
        c[0] = __uint_as_float(v[0]);
        c[1] = __uint_as_float(v[1]);
        c[2] = __uint_as_float(v[2]);
        c[3] = __uint_as_float(v[3]);
    }
}

// Support interval of index s + k*d inside (a, b): intersect into [lo, hi].
// (-1, n) is the support of a volume of n voxels: a sample outside it has no
// corner inside the volume.
__device__ __forceinline__ void clip_axis(double s, double d, double a, double b, double &lo, double &hi)
{
    // no fused multiply-adds beyond the explicit ones: every kernel that
    // inlines this evaluates the same operations
#pragma clang fp contract(off)
    if (fabs(d) < 1e-300) {
        if (!(s > a && s < b)) { lo = 1.0; hi = 0.0; }
        return;
    }
    double k1 = (a - s) / d, k2 = (b - s) / d;
    lo = fmax(lo, fmin(k1, k2));
    hi = fmin(hi, fmax(k1, k2));
}

// Sample positions.  Every kAnchor-th sample of the ray (global sample index
// k a multiple of kAnchor) is an anchor whose position is exact in 32.32
// fixed point and split into an integer cell and an FP32 offset; the
// kAnchor samples from it on are offset + i * step in FP32 (|error| <
// ~3e-6 * max(1, |step|) voxel, against ~1e-4 for the reference's FP32
// t * d - SOD).  The anchor grid is fixed, so every launch option sees the
// same positions.
constexpr int kAnchor = 16;

// Ray (u, v) of view `vw`: FP64 set-up of the index-space start (at k_lo) and
// step, and the sample range [k_lo, k_hi).
// A ray's samples in index space.  Positions are 32.32 fixed point (units of
// 2^-32 voxel) so that anchors are exact integer steps: every path that
// reaches an anchor, by whatever steps, gets the same value, and the march
// needs no FP64 after the set-up.
struct RaySetup {
    int64_t p0[3]; // position of sample k_ref
    int64_t d[3];  // step per sample
    float df[3];   // step per sample in FP32 (samples after an anchor)
    int k_lo, n;   // the samples marched, [k_lo, k_lo + n)
    int k_ref;     // the start of the volume's support window (k_ref <= k_lo)
};

constexpr double kFix = 4294967296.0; // 2^32

// FP64 index-space coordinate to 32.32 fixed point (|x| < 2^31).
__device__ __forceinline__ int64_t to_fix(double x) { return (int64_t)rint(x * kFix); }

// Index-space line of ray (u, v) in the view with rotation cs = (cos, sin):
// position(t) = src + t * dir, t in mm from the source (FP64).
struct RayLine {
    double sa, sb, sc; // source (t = 0)
    double da, db, dc; // d(index)/dt
};
__device__ __forceinline__ RayLine ray_line(const MarchArgs &P, int u, int v, int vw)
{
#pragma clang fp contract(off)
    const double2 cs = P.view_cs[vw];
    // computeIntPointCoordinatesKernel (CBCTPolyProjGPU.cu:1813-1829), FP64.
    double detU = P.half_u - P.dsize * 0.5 - (double)u * P.dsize;
    double detV = P.half_v - P.dsize * 0.5 - (double)v * P.dsize;
    double sdd2 = P.sdd * P.sdd;
    double r = sqrt(sdd2 + detU * detU + detV * detV);
    double inv_r = 1.0 / r;
    double sg = detV * inv_r;
    double cg = sqrt(sdd2 + detU * detU) * inv_r;
    double sb = detU * inv_r;
    double cb = sqrt(sdd2 + detV * detV) * inv_r;
    double dX = cg * cb, dY = cg * sb, dZ = sg;
    // transformKernel (CBCTPolyProjGPU.cu:1915-1917) is affine in t.
    RayLine L;
    if (!P.focal) {
        L.da = (dX * cs.x - dY * cs.y) * P.inv_px;
        L.db = -(dX * cs.y + dY * cs.x) * P.inv_py;
        L.dc = -dZ * P.inv_pz;
        L.sa = (P.half_x - P.sod * cs.x) * P.inv_px;
        L.sb = (P.half_y + P.sod * cs.y) * P.inv_py;
        L.sc = P.half_z * P.inv_pz;
        return L;
    }
    // Focal spot: the source moves to (-SOD, su, sv) in its frame and the ray
    // still reaches the point where the nominal ray meets the detector plane
    // x = SDD - SOD; samples are measured from the displaced source.
    const double2 f = P.focal[vw];
    const double qx = P.sdd, qy = P.sdd * dY / dX - f.x, qz = P.sdd * dZ / dX - f.y;
    const double inv_n = 1.0 / sqrt(qx * qx + qy * qy + qz * qz);
    const double ex = qx * inv_n, ey = qy * inv_n, ez = qz * inv_n;
    L.da = (ex * cs.x - ey * cs.y) * P.inv_px;
    L.db = -(ex * cs.y + ey * cs.x) * P.inv_py;
    L.dc = -ez * P.inv_pz;
    // source (-SOD, su, sv) rotated like every sample
    L.sa = (P.half_x - P.sod * cs.x - f.x * cs.y) * P.inv_px;
    L.sb = (P.half_y + P.sod * cs.y - f.x * cs.x) * P.inv_py;
    L.sc = (P.half_z - f.y) * P.inv_pz;
    return L;
}

// Continuous detector coordinates (u, v) of the ray through the index-space
// point (a, b, c) in the view cs: the inverse of ray_line's direction model
// (azimuth tan = detU / sqrt(SDD^2 + detV^2), elevation sin = detV / r).
__device__ __forceinline__ void project_point(const MarchArgs &P, int vw, double a, double b, double c,
                                              double &u, double &v)
{
#pragma clang fp contract(off)
    const double2 cs = P.view_cs[vw];
    const double qx = a / P.inv_px - P.half_x + P.sod * cs.x; // world, relative to the source
    const double qy = b / P.inv_py - P.half_y - P.sod * cs.y;
    const double qz = c / P.inv_pz - P.half_z;
    double X = qx * cs.x - qy * cs.y, Y = -qx * cs.y - qy * cs.x, Z = -qz;
    if (P.focal) {
        // through the displaced source to the detector plane, then the nominal
        // direction from the nominal source to that point
        const double2 f = P.focal[vw];
        const double lam = P.sdd / X;
        Y = f.x + lam * (Y - f.x);
        Z = f.y + lam * (Z - f.y);
        X = P.sdd;
    }
    const double al = Z / sqrt(X * X + Y * Y), be = Y / X;
    const double den = 1.0 - al * al * be * be;
    const double detV = al * P.sdd * sqrt((1.0 + be * be) / den);
    const double detU = be * P.sdd * sqrt((1.0 + al * al) / den);
    u = (P.half_u - P.dsize * 0.5 - detU) / P.dsize;
    v = (P.half_v - P.dsize * 0.5 - detV) / P.dsize;
}

// Steps both ends of the window [k_lo, k_hi) of the ray s + k d (index
// space, FP64) inward past samples whose cells are all zero, through the
// brick distances: at sample k, whose cells (floor(p) + pad and the next on
// each axis) lie in brick B (the brick of floor(p) + pad), every brick within
// dist(B) - 1 of B is unoccupied, so every cell within 8 (dist(B) - 1) - 1 of
// floor(p) + pad on every axis is zero.  A sample at most 8 (dist(B) - 1) - 2
// voxels away on every axis (the march's own positions are within 1e-5
// voxel of these) reads only such cells: the step is that many samples of
// the largest axis component, less one voxel for rounding.  A sample outside
// the padded volume, or a brick within one of an occupied brick, ends the
// stepping.  Every sample stepped over adds exact zeros.
constexpr int kTraceSteps = 64;
__device__ __forceinline__ void trace_window_ends(const MarchArgs &P, double sa, double sb, double sc, double da,
                                                  double db, double dc, int &k_lo, int &k_hi)
{
#pragma clang fp contract(off)
    const double amax = fmax(fabs(da), fmax(fabs(db), fabs(dc)));
    if (!(amax > 0.0)) return;
    const double inv = 1.0 / amax;
    auto dist_at = [&](int k) -> int {
        const double kk = (double)k;
        const double x = floor(sa + da * kk) + kVoxPadLo, y = floor(sb + db * kk) + kVoxPadLo;
        const double z = floor(sc + dc * kk) + kVoxPadLo;
        if (!(x >= 0.0 && y >= 0.0 && z >= 0.0 && x < (double)P.pcx && y < (double)P.pcy && z < (double)P.pcz))
            return 0;
        const int64_t bx = (int64_t)x / kOccBrick, by = (int64_t)y / kOccBrick, bz = (int64_t)z / kOccBrick;
        return P.odist[(bz * P.ob_y + by) * P.ob_x + bx];
    };
    // the step for distance d, at most `limit` samples (converted only once
    // it fits an int)
    auto step_of = [&](int d, int limit) {
        const double n = floor((8.0 * (d - 1) - 3.0) * inv);
        return n >= (double)limit ? limit : (int)n;
    };
    int k = k_lo;
    for (int it = 0; it < kTraceSteps && k < k_hi; it++) {
        const int d = dist_at(k);
        const int n = d >= 2 ? step_of(d, k_hi - k) : 0;
        if (n < 1) break;
        k += n; // samples k .. k + n read only zero cells
    }
    if (k >= k_hi) {
        k_hi = k_lo; // the whole window reads zeros
        return;
    }
    k_lo = max(k_lo, k - 1);
    int e = k_hi - 1;
    for (int it = 0; it < kTraceSteps && e > k_lo; it++) {
        const int d = dist_at(e);
        const int n = d >= 2 ? step_of(d, e - k_lo + 1) : 0;
        if (n < 1) break;
        e -= n;
    }
    k_hi = min(k_hi, e + 2);
}

// kBox: also clip the window to P.obox, the occupied bricks, when set.  The
// positions stay anchored at the volume window's start, so a clipped ray
// samples the same positions and only leaves out samples whose corners are
// all zero: the line integrals keep every bit.  (Only the forward quantities
// may be clipped; the scatter adjoint needs every sample in the volume.)
template <bool kBox = false>
__device__ __forceinline__ RaySetup setup_ray(const MarchArgs &P, int u, int v, int vw)
{
#pragma clang fp contract(off)
    const RayLine Ln = ray_line(P, u, v, vw);
    double sa = Ln.sa + Ln.da * P.t0, sb2 = Ln.sb + Ln.db * P.t0, sc = Ln.sc + Ln.dc * P.t0; // k = 0
    double da = Ln.da * P.dx, db = Ln.db * P.dx, dc = Ln.dc * P.dx;
    double lo = 0.0, hi = (double)P.K - 1.0;
    clip_axis(sa, da, -1.0, (double)P.nx, lo, hi);
    clip_axis(sb2, db, -1.0, (double)P.ny, lo, hi);
    clip_axis(sc, dc, -1.0, (double)P.nz, lo, hi);
    RaySetup R;
    int k_lo = 0, k_hi = 0;
    if (hi >= lo) {
        // widen by one sample each side; extra samples evaluate to exact zeros
        k_lo = max(0, (int)floor(lo) - 1);
        k_hi = min(P.K, (int)ceil(hi) + 2);
    }
    const int k_ref = k_lo;
    if (kBox && P.obox && k_hi > k_lo) {
        // Voxels outside [8 lo - 2, 8 hi + 5] on an axis are zero: a brick
        // covers padded cells 8 b .. 8 b + 7, padded cell c holds voxel c - 2
        // (and, in 16-bit storage, c - 1 as the pair's second half), and
        // every cell outside the occupied bricks is zero.
        const int32_t *ob = P.obox;
        if (ob[0] > ob[1]) {
            k_hi = k_lo; // nothing occupied
        } else {
            clip_axis(sa, da, 8.0 * ob[0] - 3.0, 8.0 * ob[1] + 6.0, lo, hi);
            clip_axis(sb2, db, 8.0 * ob[2] - 3.0, 8.0 * ob[3] + 6.0, lo, hi);
            clip_axis(sc, dc, 8.0 * ob[4] - 3.0, 8.0 * ob[5] + 6.0, lo, hi);
            if (hi >= lo) {
                k_lo = max(k_lo, (int)floor(lo) - 1);
                k_hi = min(k_hi, (int)ceil(hi) + 2);
            } else {
                k_hi = k_lo;
            }
        }
        if (P.odist && k_hi > k_lo) trace_window_ends(P, sa, sb2, sc, da, db, dc, k_lo, k_hi);
        if (k_hi > k_lo) {
            // Whole anchor segments: the window grows to the anchor grid
            // (up to kAnchor - 1 samples each side, outside the window
            // above, so exact zeros; cells past the padded volume stage as
            // zeros), and every segment of the ray but one at the lattice's
            // end takes the unrolled path of the march instead of the
            // sample-by-sample loop.
            k_lo &= ~(kAnchor - 1);
            k_hi = min(P.K, (k_hi + kAnchor - 1) & ~(kAnchor - 1));
        }
    }
    R.k_lo = k_lo;
    R.n = max(0, k_hi - k_lo);
    R.k_ref = k_ref;
    R.p0[0] = to_fix(sa + da * k_ref);
    R.p0[1] = to_fix(sb2 + db * k_ref);
    R.p0[2] = to_fix(sc + dc * k_ref);
    R.d[0] = to_fix(da);
    R.d[1] = to_fix(db);
    R.d[2] = to_fix(dc);
    R.df[0] = (float)da;
    R.df[1] = (float)db;
    R.df[2] = (float)dc;
    return R;
}

// A lane without a ray (past the detector edge): no samples.
__device__ __forceinline__ RaySetup empty_ray()
{
    RaySetup R;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        R.p0[a] = 0;
        R.d[a] = 0;
        R.df[a] = 0.0f;
    }
    R.k_lo = 0;
    R.n = 0;
    R.k_ref = 0;
    return R;
}

// Storage of the resident volume: FP32 voxels, or FP16 / BF16 x-pairs.
enum Storage { kStF32 = 0, kStF16 = 1, kStBF16 = 2 };

template <int M>
__device__ __forceinline__ void load_raw(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                         uint32_t (&c)[M])
{
    if constexpr (M == 1) {
        c[0] = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
    } else if constexpr (M == 2) {
        auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        c[0] = v[0];
        c[1] = v[1];
    } else if constexpr (M == 3) {
        auto v = __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, 0);
        c[0] = v[0];
        c[1] = v[1];
        c[2] = v[2];
    } else {
        auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
        c[0] = v[0];
        c[1] = v[1];
        c[2] = v[2];
        c[3] = v[3];
    }
}

// The M dwords of one cell through a 64-bit address (global loads; a cell
// is dword-aligned), for volumes past the 4 GiB buffer-load range.
template <int M>
__device__ __forceinline__ void load_glob(const char *p, uint32_t (&c)[M])
{
    const uint32_t *q = (const uint32_t *)p;
#pragma unroll
    for (int m = 0; m < M; m++) c[m] = q[m];
}

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef short bf16x2_t __attribute__((ext_vector_type(2)));

// acc += w.x * lo(d) + w.y * hi(d), FP32 accumulate (v_dot2_f32_f16 / _bf16).
template <int ST>
__device__ __forceinline__ float dot2_acc(uint32_t w, uint32_t d, float acc)
{
    if constexpr (ST == kStF16)
        return __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, w), __builtin_bit_cast(half2_t, d), acc,
                                      false);
    else
        return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2_t, w),
                                               __builtin_bit_cast(bf16x2_t, d), acc, false);
}

typedef __bf16 bf16x2_v __attribute__((ext_vector_type(2)));
typedef float f32x2_v __attribute__((ext_vector_type(2)));

// Round-to-nearest-even pack of two FP32 weights into the storage's 16-bit
// pair (v_cvt_pk_f16_f32 / v_cvt_pk_bf16_f32).
template <int ST> __device__ __forceinline__ uint32_t pack_w(float lo, float hi)
{
    if constexpr (ST == kStF16) {
        return __builtin_bit_cast(uint32_t, __floats2half2_rn(lo, hi));
    } else {
        return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_v){lo, hi}, bf16x2_v));
    }
}

// The x-pair of weights (w * (1 - ax), w * ax) in the storage's 16-bit pair:
// FP32 products (the first as the fma w - w * ax), then one pack.
template <int ST> __device__ __forceinline__ uint32_t x_pair(float w, float ax)
{
    return pack_w<ST>(fmaf(-w, ax, w), w * ax);
}

struct Anchor {
    int ia, ib, ic;       // cell of the anchor
    float fa, fb, fc;     // offset of the anchor inside its cell, [0, 1)
    float da, db, dc;     // per-sample step (FP32)
};

// A fixed-point position (three axes).
struct FixPos {
    int64_t p[3];
};

// Position of global sample k: one 64-bit product, for loop starts.
__device__ __forceinline__ FixPos fix_at(const RaySetup &R, int k)
{
    const int64_t j = (int64_t)(k - R.k_ref);
    FixPos q;
#pragma unroll
    for (int a = 0; a < 3; a++) q.p[a] = R.p0[a] + j * R.d[a];
    return q;
}

// n samples' worth of steps.
__device__ __forceinline__ FixPos fix_steps(const RaySetup &R, int n)
{
    FixPos q;
#pragma unroll
    for (int a = 0; a < 3; a++) q.p[a] = (int64_t)n * R.d[a];
    return q;
}

__device__ __forceinline__ void fix_add(FixPos &q, const FixPos &s)
{
#pragma unroll
    for (int a = 0; a < 3; a++) q.p[a] += s.p[a];
}

// The anchor at fixed-point position q: the cell is the high word (a floor),
// the offset the top 24 bits of the low word (exact in FP32, < 1).
__device__ __forceinline__ Anchor anchor_from(const FixPos &q, const RaySetup &R)
{
    Anchor A;
    A.ia = (int)(q.p[0] >> 32);
    A.ib = (int)(q.p[1] >> 32);
    A.ic = (int)(q.p[2] >> 32);
    A.fa = (float)((uint32_t)q.p[0] >> 8) * 0x1p-24f;
    A.fb = (float)((uint32_t)q.p[1] >> 8) * 0x1p-24f;
    A.fc = (float)((uint32_t)q.p[2] >> 8) * 0x1p-24f;
    A.da = R.df[0];
    A.db = R.df[1];
    A.dc = R.df[2];
    return A;
}

// Sample `fi` (0..kAnchor-1) after an anchor: cell relative to the anchor's
// cell and the FP32 interpolation fractions.
struct Cell {
    int x, y, z;
    float ax, ay, az;
};
// floor(x) as an int in one instruction (v_cvt_flr_i32_f32; |x| < 2^24 here).
__device__ __forceinline__ int floor_to_int(float x)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

__device__ __forceinline__ Cell offset_cell(const Anchor &A, float fi)
{
    const float x = fmaf(fi, A.da, A.fa), y = fmaf(fi, A.db, A.fb), z = fmaf(fi, A.dc, A.fc);
    Cell c;
    // fract(x) = x - floor(x), held below 1.0 (v_fract_f32)
    c.ax = __builtin_amdgcn_fractf(x);
    c.ay = __builtin_amdgcn_fractf(y);
    c.az = __builtin_amdgcn_fractf(z);
    c.x = floor_to_int(x);
    c.y = floor_to_int(y);
    c.z = floor_to_int(z);
    return c;
}

// Absolute cell clamped to [-2, N] (every corner outside the volume is then a
// padding zero, as is every corner of an unclamped cell outside [-1, N)).
__device__ __forceinline__ Cell absolute_cell(const Anchor &A, const Cell &o, int nx, int ny, int nz)
{
    Cell c = o;
    c.x = min(max(A.ia + o.x, -2), nx);
    c.y = min(max(A.ib + o.y, -2), ny);
    c.z = min(max(A.ic + o.z, -2), nz);
    return c;
}

// Cells that samples i0 .. i1-1 after anchor A use, [lo, hi] with the +1
// corner included, merged into lo / hi.  A sample's position is one rounding
// of an affine function of its index (offset_cell), so positions, and their
// cells, are monotone in the index: the two end samples bound the rest
// exactly.
__device__ __forceinline__ void segment_box(const Anchor &A, int i0, int i1, int (&lo)[3], int (&hi)[3])
{
    const float f0 = (float)i0, f1 = (float)(i1 - 1);
    const Cell a = offset_cell(A, f0), b = offset_cell(A, f1);
    lo[0] = min(lo[0], A.ia + min(a.x, b.x));
    lo[1] = min(lo[1], A.ib + min(a.y, b.y));
    lo[2] = min(lo[2], A.ic + min(a.z, b.z));
    hi[0] = max(hi[0], A.ia + max(a.x, b.x) + 1);
    hi[1] = max(hi[1], A.ib + max(a.y, b.y) + 1);
    hi[2] = max(hi[2], A.ic + max(a.z, b.z) + 1);
}

// Trilinear accumulation of one sample into the M per-material sums from
// 8 corners (FP32 voxels) or 4 rows of x-pairs (16-bit).  `fetch(row, out)`
// returns the M dwords at corner row `row` (0: y0z0, 1: y1z0, 2: y0z1,
// 3: y1z1) and, for FP32, `fetch(row + 4, out)` the x+1 corner.
template <int M, int ST, typename Fetch>
__device__ __forceinline__ void accumulate(const Cell &c, Fetch fetch, float (&acc)[M])
{
    if constexpr (ST == kStF32) {
        const float bx = 1.0f - c.ax, by = 1.0f - c.ay, bz = 1.0f - c.az;
        uint32_t d[8][M];
#pragma unroll
        for (int r = 0; r < 8; r++) fetch(r, d[r]);
        const float wr[4] = {by * bz, c.ay * bz, by * c.az, c.ay * c.az};
#pragma unroll
        for (int m = 0; m < M; m++) {
            float a = acc[m];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                a = fmaf(bx * wr[r], __uint_as_float(d[r][m]), a);
                a = fmaf(c.ax * wr[r], __uint_as_float(d[r + 4][m]), a);
            }
            acc[m] = a;
        }
    } else {
        uint32_t d[4][M];
#pragma unroll
        for (int r = 0; r < 4; r++) fetch(r, d[r]);
        // (y, z) row weights; (1 - ay) * b as the fma b - ay * b
        const float bz = 1.0f - c.az;
        const float w00 = fmaf(-c.ay, bz, bz), w10 = c.ay * bz, w01 = fmaf(-c.ay, c.az, c.az),
                    w11 = c.ay * c.az;
        const uint32_t p[4] = {x_pair<ST>(w00, c.ax), x_pair<ST>(w10, c.ax), x_pair<ST>(w01, c.ax),
                               x_pair<ST>(w11, c.ax)};
#pragma unroll
        for (int m = 0; m < M; m++) {
            float a = acc[m];
#pragma unroll
            for (int r = 0; r < 4; r++) a = dot2_acc<ST>(p[r], d[r][m], a);
            acc[m] = a;
        }
    }
}

// Global-memory gather of one sample (raw buffer loads; one VGPR offset).
template <int M, int ST, bool W>
__device__ __forceinline__ void accumulate_global(__amdgpu_buffer_rsrc_t rsrc, const char *vol, const Cell &c,
                                                  uint32_t oy, uint32_t oz, float (&acc)[M])
{
    constexpr uint32_t CB = 4u * M;
    // Byte offset with full-rate 24-bit multiply-adds: the masks (no-ops on
    // the clamped cells) tell the compiler every factor is < 2^24; the
    // strides are < 2^24 and the offset < 2^32 (checked at allocation).
    const uint32_t xi = (uint32_t)(c.x + kVoxPadLo) & 0xFFFFFu;
    const uint32_t yi = (uint32_t)(c.y + kVoxPadLo) & 0xFFFFFu;
    const uint32_t zi = (uint32_t)(c.z + kVoxPadLo) & 0xFFFFFu;
    const uint32_t soff[4] = {0u, oy, oz, oy + oz};
    if constexpr (W) {
        // a volume of 4 GiB or more: a 64-bit z offset; within a slice
        // (< 16 MiB) the y and x offsets stay 32-bit
        const char *p = vol + ((uint64_t)zi * oz + (yi * oy + xi * CB));
        accumulate<M, ST>(c, [&](int r, uint32_t(&out)[M]) {
            load_glob<M>(p + soff[r & 3] + (r < 4 ? 0u : CB), out);
        }, acc);
    } else {
        const uint32_t voff = zi * oz + (yi * oy + xi * CB);
        accumulate<M, ST>(c, [&](int r, uint32_t(&out)[M]) {
            load_raw<M>(rsrc, r < 4 ? voff : voff + CB, soff[r & 3], out);
        }, acc);
    }
}

// Bytes of one staged cell in LDS: M dwords, three padded to four so every
// cell read is one aligned ds_read.
template <int M> constexpr uint32_t lds_cell_bytes() { return M == 3 ? 16u : 4u * M; }

template <int M>
__device__ __forceinline__ void lds_read(const char *lds, uint32_t off, uint32_t (&c)[M])
{
    if constexpr (M == 1) {
        c[0] = *(const uint32_t *)(lds + off);
    } else if constexpr (M == 2) {
        uint2 v = *(const uint2 *)(lds + off);
        c[0] = v.x;
        c[1] = v.y;
    } else {
        // All 16 bytes, padding included: ds_read_b128 serves a wavefront
        // in fewer LDS cycles than ds_read_b96 (1.44x the lane rate on
        // gfx950).  The volatile access keeps the compiler from narrowing
        // the load to the three dwords that are used; the explicit LDS
        // address space keeps it a ds_read (volatile accesses are not
        // moved from generic to LDS addressing by the compiler).
        typedef uint32_t u32x4_v __attribute__((ext_vector_type(4)));
        typedef const volatile __attribute__((address_space(3))) u32x4_v lds_u32x4;
        const u32x4_v v = *(lds_u32x4 *)(lds + off);
        c[0] = v.x;
        c[1] = v.y;
        c[2] = v.z;
        if constexpr (M == 4) c[3] = v.w;
    }
}

template <int M>
__device__ __forceinline__ void lds_write(char *lds, uint32_t off, const uint32_t (&c)[M])
{
    if constexpr (M == 1) {
        *(uint32_t *)(lds + off) = c[0];
    } else if constexpr (M == 2) {
        *(uint2 *)(lds + off) = make_uint2(c[0], c[1]);
    } else if constexpr (M == 3) {
        *(uint4 *)(lds + off) = make_uint4(c[0], c[1], c[2], 0u);
    } else {
        *(uint4 *)(lds + off) = make_uint4(c[0], c[1], c[2], c[3]);
    }
}

// Ray geometry of the lane's pixel (or an empty range for lanes past the edge).
struct TileLane {
    int vw_local, u, v;
    bool active;
};
__device__ __forceinline__ TileLane tile_lane(const MarchArgs &P, uint32_t logical)
{
    const uint32_t tiles = (uint32_t)(P.tiles_u * P.tiles_v);
    TileLane t;
    t.vw_local = (int)(logical / tiles);
    const uint32_t q = logical % tiles;
    const int tile_v = (int)(q / P.tiles_u), tile_u = (int)(q % P.tiles_u);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wu = P.wu_log2, bu = P.bu_log2;
    const int tile_pu = 1 << (wu + bu), tile_pv = (64 >> wu) * (4 >> bu);
    t.u = tile_u * tile_pu + ((wave & ((1 << bu) - 1)) << wu) + (lane & ((1 << wu) - 1));
    t.v = tile_v * tile_pv + (wave >> bu) * (64 >> wu) + (lane >> wu);
    t.active = (t.u < P.nu) && (t.v < P.nv);
    return t;
}

// e^-tau and 1 - e^-tau in FP32.  e^-tau is one v_exp_f32 of -tau log2(e)
// (1 ulp, plus the rounding of the product: relative error < 1e-6 up to
// tau = 20).  1 - e^-tau is 1 - e^-tau for |tau| >= 0.35 (relative error
// < 2e-7 there) and below it the Taylor series to tau^7 in Horner form
// (remainder < tau^8 / 8! : 2e-8 relative at 0.35), so thin rays keep their
// absorbed fraction without cancellation.
__device__ __forceinline__ float absorb_series(float tau)
{
    float q = fmaf(tau, 1.0f / 5040.0f, -1.0f / 720.0f);
    q = fmaf(tau, q, 1.0f / 120.0f);
    q = fmaf(tau, q, -1.0f / 24.0f);
    q = fmaf(tau, q, 1.0f / 6.0f);
    q = fmaf(tau, q, -0.5f);
    q = fmaf(tau, q, 1.0f);
    return tau * q;
}
__device__ __forceinline__ void attenuate(float tau, float &ex, float &ab)
{
    ex = __builtin_amdgcn_exp2f(-tau * 1.44269504088896341f);
    ab = fabsf(tau) < 0.35f ? absorb_series(tau) : 1.0f - ex;
}

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float16_t __attribute__((ext_vector_type(16)));

// The spectrum of the wavefront's 64 rays with the attenuation exponents
// from the matrix cores: tau[e][ray] = sum_k A'[e][k] L'[k][ray] with
// v_mfma_f32_32x32x16_f16 (energies x rays, 32 x 32 per instruction).  Both
// operands are split into FP16 hi + lo parts and K holds, per material,
// (A_hi, L_hi), (A_hi, L_lo), (A_lo, L_hi): the products are exact in the FP32
// accumulator and only A_lo L_lo (~2^-22 relative) is left out.  A is scaled
// by a power of two 2^s (P.Aq, [energy rows padded to 32][16] FP16) chosen
// from the largest attenuation so that 2^s A stays below 2^15 (no FP16
// overflow) and A_lo of the large entries stays normal; tau is the FP32 sum
// times 2^-s (exact).
// Lane l (half h = l / 32, column c = l % 32) receives tau for rays c and
// 32 + c and the energy rows 8b + 4h + q (b, q < 4) of each 32-row block, runs
// the exponentials and the weighted sums for those (each group of four bins
// in FP32, then one compensated add), and one exchange with lane l ^ 32
// completes the sums of its own ray.  Every lane
// of the wavefront must call it; `active` marks rays inside the detector.
template <int M>
__device__ __forceinline__ void spectrum_mfma(const MarchArgs &P, const TileLane &t, const float (&Lm)[M],
                                              float &sI_out, float &sA_out)
{
    static_assert(3 * M <= 16, "K = 3M must fit one 32x32x16 MFMA");
    const int lane = threadIdx.x & 63, h = lane >> 5;
    // the other ray of this lane's column lives in lane ^ 32
    float Lx[M];
#pragma unroll
    for (int m = 0; m < M; m++) Lx[m] = __shfl_xor(Lm[m], 32);
    const int pix_own = t.active ? t.v * (int)P.w_stride_v + t.u : -1;
    const int pix_x = __shfl_xor(pix_own, 32);
    float sI[2] = {0.0f, 0.0f}, sA[2] = {0.0f, 0.0f};
    const float16_t zero = {};
    const int n_act = P.n_act;
    // exp2 argument and the series test straight from the scaled sum d =
    // 2^s tau: d * (-2^-s log2 e) and |d| < 0.35 * 2^s round and compare
    // exactly as -tau * log2 e and |tau| < 0.35 (power-of-two scalings)
    const float arg_scale = -P.aq_inv_scale * 1.44269504088896341f;
    const float small_d = 0.35f / P.aq_inv_scale;
#pragma unroll 1
    for (int c = 0; c < 2; c++) {
        // ray c * 32 + (lane % 32): this lane's own ray when c == h
        const bool own = c == h;
        _Float16 kv[16];
#pragma unroll
        for (int k = 0; k < 16; k++) kv[k] = (_Float16)0.0f;
#pragma unroll
        for (int m = 0; m < M; m++) {
            const float L = own ? Lm[m] : Lx[m];
            const _Float16 hi = (_Float16)L;
            const _Float16 lo = (_Float16)(L - (float)hi);
            kv[3 * m] = hi;
            kv[3 * m + 1] = lo;
            kv[3 * m + 2] = hi;
        }
        // lanes 0-31 hold k = 0..7, lanes 32-63 k = 8..15 (a dword-wise select)
        const uint32_t sel = 0u - (uint32_t)h;
        uint32_t bw[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t lo = (uint32_t)__builtin_bit_cast(uint16_t, kv[2 * j]) |
                                ((uint32_t)__builtin_bit_cast(uint16_t, kv[2 * j + 1]) << 16);
            const uint32_t hi = (uint32_t)__builtin_bit_cast(uint16_t, kv[8 + 2 * j]) |
                                ((uint32_t)__builtin_bit_cast(uint16_t, kv[8 + 2 * j + 1]) << 16);
            bw[j] = (hi & sel) | (lo & ~sel);
        }
        const half8_t b = __builtin_bit_cast(half8_t, bw);
        const int pix = own ? pix_own : pix_x;
        // The weight of bin e = 32 rb + 8 bb + 4 h + q for this lane's ray
        // sits at W + e * stride + pix.  Buffer loads from a descriptor at
        // row 32 rb take the uniform (8 bb + q) * stride in a scalar offset
        // and this lane's (pix + 4 h stride) as its vector offset, so no
        // weight load needs address arithmetic (both offsets stay below
        // 32 * stride * 4 bytes, which the host keeps under 4 GiB).
        const uint32_t woff = ((uint32_t)(pix < 0 ? 0 : pix) + (uint32_t)(4 * h) * (uint32_t)P.w_stride_e) * 4u;
        const uint32_t wrow = (uint32_t)P.w_stride_e * 4u;
        KSum si = ksum_zero(), sa = ksum_zero();
#pragma unroll 1
        for (int rb = 0; rb * 32 < n_act; rb++) {
            const half8_t a = *(const half8_t *)(P.Aq + ((rb * 32 + (lane & 31)) * 16 + 8 * h));
            const float16_t d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, zero, 0, 0, 0);
            const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<float *>(P.W + (int64_t)rb * 32 * P.w_stride_e), 0, (int)0xFFFFFFFFu, 0x00020000);
#pragma unroll
            for (int bb = 0; bb < 4; bb++) {
                if (rb * 32 + 8 * bb >= n_act) break; // uniform: the rest of the block is past the bins
                // four rows at a time: keeps the weight loads and exp temporaries few
                __builtin_amdgcn_sched_barrier(0);
                // the four weights first, so their loads overlap; bins past
                // the last active one read the table's zero rows and add 0
                float wv[4];
#pragma unroll
                for (int q = 0; q < 4; q++)
                    wv[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(wr, woff, (uint32_t)(8 * bb + q) * wrow, 0));
                float ex[4], ab[4];
                bool small = false;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    ex[q] = __builtin_amdgcn_exp2f(d[4 * bb + q] * arg_scale);
                    ab[q] = 1.0f - ex[q];
                    small |= fabsf(d[4 * bb + q]) < small_d;
                }
                // the series only when a lane of the wavefront needs it
                // (a uniform branch; the same values as attenuate)
                if (__builtin_amdgcn_ballot_w64(small) != 0) {
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        if (fabsf(d[4 * bb + q]) < small_d)
                            ab[q] = absorb_series(d[4 * bb + q] * P.aq_inv_scale);
                }
                // the four bins summed in FP32, then one compensated add each
                // (a few ulps of the total, as for bin-by-bin compensation)
                float gI = wv[0] * ex[0], gA = wv[0] * ab[0];
#pragma unroll
                for (int q = 1; q < 4; q++) {
                    gI = fmaf(wv[q], ex[q], gI);
                    gA = fmaf(wv[q], ab[q], gA);
                }
                ksum_add(si, gI);
                ksum_add(sa, gA);
            }
        }
        if (c == 0) {
            sI[0] = ksum_value(si);
            sA[0] = ksum_value(sa);
        } else {
            sI[1] = ksum_value(si);
            sA[1] = ksum_value(sa);
        }
    }
    // partial sums of the partner's own ray go to the partner
    const float rI = __shfl_xor(h ? sI[0] : sI[1], 32);
    const float rA = __shfl_xor(h ? sA[0] : sA[1], 32);
    sI_out = (h ? sI[1] : sI[0]) + rI;
    sA_out = (h ? sA[1] : sA[0]) + rA;
}

// Writes L and, with the spectrum, I and IAbsorb of one ray.  Every lane of
// the wavefront calls it; lanes past the detector edge write nothing.
template <int M, int ST, bool kSpectrum>
__device__ __forceinline__ void finish_ray(const MarchArgs &P, const TileLane &t, const float (&acc)[M])
{
    const int64_t npix = (int64_t)P.nu * P.nv;
    const int64_t pix = (int64_t)t.v * P.nu + t.u;
    float Lm[M];
#pragma unroll
    for (int m = 0; m < M; m++) Lm[m] = acc[m] * P.dx_f;
    if (P.L && t.active) {
#pragma unroll
        for (int m = 0; m < M; m++) P.L[((int64_t)t.vw_local * P.l_mats + P.l_moff + m) * npix + pix] = Lm[m];
    }
    if constexpr (kSpectrum && ST != kStF32 && 3 * M <= 16) {
        // the mixed-precision projector: exponents on the matrix cores
        float sI, sA;
        spectrum_mfma<M>(P, t, Lm, sI, sA);
        if (t.active) {
            P.I[(int64_t)t.vw_local * npix + pix] = sI;
            P.IA[(int64_t)t.vw_local * npix + pix] = sA;
        }
    } else if constexpr (kSpectrum) {
        if (!t.active) return;
        // Per-ray spectrum: tau_e = sum_m A[e][m] L_m (FP32), e^-tau and
        // 1 - e^-tau (attenuate: no cancellation at small tau) in FP32, and the
        // energy sums of W e^-tau and W (1 - e^-tau) in bin order, compensated
        // (a plain FP32 sum of ~100 terms would cost ~1e-6 of I, visible in
        // ln(I0/I) of thin rays).
        const float *wp = P.W + (int64_t)t.v * P.w_stride_v + t.u;
        KSum sI = ksum_zero(), sA = ksum_zero();
        for (int e = 0; e < P.n_act; e++) {
            float tau = 0.0f;
#pragma unroll
            for (int m = 0; m < M; m++) tau = fmaf(P.A[e * M + m], Lm[m], tau);
            float ex, ab;
            attenuate(tau, ex, ab);
            const float w = wp[(int64_t)e * P.w_stride_e];
            ksum_add(sI, w * ex);
            ksum_add(sA, w * ab);
        }
        P.I[(int64_t)t.vw_local * npix + pix] = ksum_value(sI);
        P.IA[(int64_t)t.vw_local * npix + pix] = ksum_value(sA);
    }
}

__device__ __forceinline__ uint32_t xcd_logical(const MarchArgs &P)
{
    // XCD-aware block order: consecutive logical blocks (adjacent tiles of one
    // view) are kept on one XCD so they share its L2.
    const uint32_t bid = blockIdx.x;
    return (bid & 7u) * P.per_xcd + (bid >> 3);
}

// One ray's samples, every corner gathered from the volume in HBM / L2 (W:
// the volume spans 4 GiB or more).
template <int M, int ST, bool W>
__device__ __forceinline__ void march_global(const MarchArgs &P, const RaySetup &R, float (&acc)[M])
{
    __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(P.vol), 0, (int)P.vol_bytes, 0x00020000);
    const uint32_t oy = P.row_y & 0xFFFFFFu, oz = P.row_z & 0xFFFFFFu;
    const int nx = (int)P.nx, ny = (int)P.ny, nz = (int)P.nz;
    const int k_hi = R.k_lo + R.n;
    const FixPos d16 = fix_steps(R, kAnchor);
    FixPos q = fix_at(R, R.k_lo & ~(kAnchor - 1));
    for (int ka = R.k_lo & ~(kAnchor - 1); ka < k_hi; ka += kAnchor, fix_add(q, d16)) {
        const Anchor A = anchor_from(q, R);
        const int i1 = min(ka + kAnchor, k_hi) - ka;
        int i = max(ka, R.k_lo) - ka;
        for (float fi = (float)i; i < i1; i++, fi += 1.0f)
            accumulate_global<M, ST, W>(rsrc, (const char *)P.vol, absolute_cell(A, offset_cell(A, fi), nx, ny, nz),
                                        oy, oz, acc);
    }
}

// Ray-march with direct global gathers: 8 (FP32) or 4 (16-bit) loads per sample.
template <int M, int ST, bool kSpectrum>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void march_kernel(MarchArgs P)
{
    const uint32_t logical = xcd_logical(P);
    if (logical >= P.n_blocks) return;
    const TileLane t = tile_lane(P, logical);
    const RaySetup R = t.active ? setup_ray(P, t.u, t.v, (int)(P.view_begin + t.vw_local)) : empty_ray();
    float acc[M];
#pragma unroll
    for (int m = 0; m < M; m++) acc[m] = 0.0f;
    if (P.wide)
        march_global<M, ST, true>(P, R, acc);
    else
        march_global<M, ST, false>(P, R, acc);
    finish_ray<M, ST, kSpectrum>(P, t, acc);
}

// Block-wide min / max of six ints (cell box of a chunk).
__device__ __forceinline__ void block_box(int (&lo)[3], int (&hi)[3], int *boxbuf)
{
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) {
            lo[d] = min(lo[d], __shfl_xor(lo[d], s));
            hi[d] = max(hi[d], __shfl_xor(hi[d], s));
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            boxbuf[wave * 8 + d] = lo[d];
            boxbuf[wave * 8 + 4 + d] = hi[d];
        }
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int a = boxbuf[d], b = boxbuf[4 + d];
#pragma unroll
        for (int w = 1; w < kBlock / 64; w++) {
            a = min(a, boxbuf[w * 8 + d]);
            b = max(b, boxbuf[w * 8 + 4 + d]);
        }
        lo[d] = a;
        hi[d] = b;
    }
}

// Min / max over each 16-lane row, in every lane of the row (DPP quad
// permutes and row mirrors fused into v_min / v_max; all lanes active).
__device__ __forceinline__ int row_min(int v)
{
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true));  // quad_perm [1, 0, 3, 2]
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true));  // quad_perm [2, 3, 0, 1]
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true)); // row_half_mirror
    return min(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true)); // row_mirror
}
__device__ __forceinline__ int row_max(int v)
{
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true));
    return max(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true));
}

// Ints of the march's box scratch: three buffers of 8 (lo x, y, z at 0..2,
// hi x, y, z at 4..6).
constexpr int kBoxBufs = 3;

// Every box buffer back to (INT_MAX, INT_MIN); thread 0 only.
__device__ __forceinline__ void box_reset(int *buf)
{
#pragma unroll
    for (int d = 0; d < 3; d++) {
        buf[d] = INT_MAX;
        buf[4 + d] = INT_MIN;
    }
}

// Block-wide min of lo[] and max of hi[] with one barrier, as scalars.  Each
// 16-lane row reduces its lanes with DPP, one lane per row folds the row
// into buffer `call` with LDS atomic min / max, and after the barrier every
// lane reads it.  The buffer must be reset before the call; the call then
// resets buffer call - 1 (mod 3), whose last reads all came before this
// barrier, for use two calls later.
template <int D>
__device__ __forceinline__ void block_box_atomic(int (&lo)[D], int (&hi)[D], int *boxbuf, int call)
{
    int *buf = boxbuf + call * 8;
#pragma unroll
    for (int d = 0; d < D; d++) {
        lo[d] = row_min(lo[d]);
        hi[d] = row_max(hi[d]);
    }
    if ((threadIdx.x & 15) == 0) {
#pragma unroll
        for (int d = 0; d < D; d++) {
            atomicMin(buf + d, lo[d]);
            atomicMax(buf + 4 + d, hi[d]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < D; d++) {
        lo[d] = __builtin_amdgcn_readfirstlane(buf[d]);
        hi[d] = __builtin_amdgcn_readfirstlane(buf[4 + d]);
    }
    if (threadIdx.x == 0) box_reset(boxbuf + (call == 0 ? kBoxBufs - 1 : call - 1) * 8);
}

// Exact q = n / d for 0 <= n < 2^22, 1 <= d < 2^10, from a float reciprocal.
__device__ __forceinline__ int small_div(int n, int d, float inv_d)
{
    int q = (int)((float)n * inv_d);
    q -= (__mul24(q, d) > n);
    q += (__mul24(q + 1, d) <= n);
    return q;
}

// 1 / d to within an ulp (v_rcp_f32); small_div corrects the quotient.
__device__ __forceinline__ float small_rcp(int d) { return __builtin_amdgcn_rcpf((float)d); }

// Rows a thread loads at once in fill_box.
constexpr int kFillBatch = 4;

// An empty asm that reads one batch row's loaded dwords, placed after every
// load of the batch is issued: the compiler waits for the batch there, once,
// and cannot sink a load into the conditional LDS write that consumes it
// (which it did, making the rows of a batch separate memory round trips:
// 4.3 % of the march's time on the 3-material 512^3 scan).
template <int M> __device__ __forceinline__ void loads_landed(const uint32_t (&c)[M])
{
    if constexpr (M == 1)
        asm volatile("" ::"v"(c[0]));
    else if constexpr (M == 2)
        asm volatile("" ::"v"(c[0]), "v"(c[1]));
    else if constexpr (M == 3)
        asm volatile("" ::"v"(c[0]), "v"(c[1]), "v"(c[2]));
    else
        asm volatile("" ::"v"(c[0]), "v"(c[1]), "v"(c[2]), "v"(c[3]));
}

// Cooperative, row-coalesced load of the cell box [lo, lo + (sx, sy, sz)) into
// LDS (cell (lx, row) of the box at byte (row * px + lx) * LB, x fastest;
// row pitch px >= sx).  Cells outside the padded volume are zeros.  A thread
// keeps one x column and walks rows (y, z) with a fixed stride, so the x test
// and offset are hoisted and the loop holds no division.
template <int M, bool W>
__device__ __forceinline__ void fill_box(__amdgpu_buffer_rsrc_t rsrc, const char *vol, char *lds,
                                         const int (&lo)[3], int sx, int px, int sy, int sz, int nx, int ny,
                                         int nz, uint32_t oy, uint32_t oz)
{
    constexpr uint32_t LB = lds_cell_bytes<M>();
    const int rows = sy * sz;
    // A volume of 4 GiB or more (W): the buffer starts at plane zb, the box's
    // first plane inside the padded volume, so every row offset stays
    // 32-bit; offset 0 is still a padding (zero) cell.
    uint32_t zb = 0;
    if constexpr (W) {
        const int pzn = nz + kVoxPadLo + kVoxPadHi;
        zb = (uint32_t)min(max(lo[2] + kVoxPadLo, 0), pzn - 1);
        const uint64_t rest = (uint64_t)(pzn - (int)zb) * oz;
        rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(vol) + (uint64_t)zb * oz, 0,
                                                 (int)(rest < 0xFFFFFFFFull ? rest : 0xFFFFFFFFull), 0x00020000);
    }
    if (sx <= kBlock) {
        const float inv_sx = small_rcp(sx), inv_sy = small_rcp(sy);
        // rows per pass and its (y, z) step: block-uniform
        const int per = __builtin_amdgcn_readfirstlane(small_div(kBlock, sx, inv_sx));
        const int dz = __builtin_amdgcn_readfirstlane(small_div(per, sy, inv_sy)), dy = per - dz * sy;
        const int tid = threadIdx.x;
        if (tid >= per * sx) return;
        const int r0 = small_div(tid, sx, inv_sx);
        const int lx = tid - __mul24(r0, sx);
        const int x = lo[0] + lx;
        int lz = small_div(r0, sy, inv_sy);
        int ly = r0 - __mul24(lz, sy);
        const bool xin = (uint32_t)(x + kVoxPadLo) < (uint32_t)(nx + kVoxPadLo + kVoxPadHi);
        const uint32_t xoff = ((uint32_t)(x + kVoxPadLo) & 0xFFFFFu) * (4u * M);
        const uint32_t ymax = (uint32_t)(ny + kVoxPadLo + kVoxPadHi), zmax = (uint32_t)(nz + kVoxPadLo + kVoxPadHi);
        uint32_t loff = (uint32_t)(__mul24(r0, px) + lx) * LB;
        const uint32_t lstep = (uint32_t)(per * px) * LB;
        // kFillBatch rows per pass of the loop: their loads are all in flight
        // before the first LDS write.  A row outside the padded volume (or
        // past the box) loads the padded cell at offset 0, which is a zero.
        const bool inside = lo[0] >= -kVoxPadLo && lo[0] + sx <= nx + kVoxPadHi && lo[1] >= -kVoxPadLo &&
                            lo[1] + sy <= ny + kVoxPadHi && lo[2] >= -kVoxPadLo && lo[2] + sz <= nz + kVoxPadHi;
        if (inside) { // block-uniform: no row needs a bounds test
            uint32_t off = __umul24((uint32_t)(lo[2] + lz + kVoxPadLo) - zb, oz) +
                           (__umul24((uint32_t)(lo[1] + ly + kVoxPadLo), oy) + xoff);
            const uint32_t step = (uint32_t)dy * oy + (uint32_t)dz * oz, wrap = oz - (uint32_t)sy * oy;
            for (int r = r0; r < rows; r += kFillBatch * per) {
                uint32_t d[kFillBatch][M];
#pragma unroll
                for (int b = 0; b < kFillBatch; b++) {
                    load_raw<M>(rsrc, r + b * per < rows ? off : 0u, 0, d[b]);
                    ly += dy;
                    off += step;
                    if (ly >= sy) {
                        ly -= sy;
                        off += wrap;
                    }
                }
#pragma unroll
                for (int b = 0; b < kFillBatch; b++) loads_landed<M>(d[b]);
#pragma unroll
                for (int b = 0; b < kFillBatch; b++) {
                    if (r + b * per < rows) lds_write<M>(lds, loff, d[b]);
                    loff += lstep;
                }
            }
            return;
        }
        for (int r = r0; r < rows; r += kFillBatch * per) {
            uint32_t d[kFillBatch][M];
#pragma unroll
            for (int b = 0; b < kFillBatch; b++) {
                const uint32_t yi = (uint32_t)(lo[1] + ly + kVoxPadLo), zi = (uint32_t)(lo[2] + lz + kVoxPadLo);
                const bool in = xin && yi < ymax && zi < zmax;
                load_raw<M>(rsrc, in ? __umul24(zi - zb, oz) + (__umul24(yi, oy) + xoff) : 0u, 0, d[b]);
                ly += dy;
                lz += dz;
                if (ly >= sy) {
                    ly -= sy;
                    lz++;
                }
            }
#pragma unroll
            for (int b = 0; b < kFillBatch; b++) loads_landed<M>(d[b]);
#pragma unroll
            for (int b = 0; b < kFillBatch; b++) {
                if (r + b * per < rows) lds_write<M>(lds, loff, d[b]);
                loff += lstep;
            }
        }
    } else {
        const int cells = rows * sx;
        const float inv_sx = small_rcp(sx), inv_sy = small_rcp(sy);
        for (int i = threadIdx.x; i < cells; i += kBlock) {
            const int row = small_div(i, sx, inv_sx), lx = i - row * sx;
            const int lz = small_div(row, sy, inv_sy), ly = row - lz * sy;
            const int x = lo[0] + lx, y = lo[1] + ly, z = lo[2] + lz;
            uint32_t d[M];
            if (x >= -kVoxPadLo && x < nx + kVoxPadHi && y >= -kVoxPadLo && y < ny + kVoxPadHi &&
                z >= -kVoxPadLo && z < nz + kVoxPadHi) {
                const uint32_t xi = (uint32_t)(x + kVoxPadLo) & 0xFFFFFu;
                const uint32_t yi = (uint32_t)(y + kVoxPadLo) & 0xFFFFFu;
                const uint32_t zi = (uint32_t)(z + kVoxPadLo) & 0xFFFFFu;
                load_raw<M>(rsrc, (zi - zb) * oz + (yi * oy + xi * (4u * M)), 0, d);
            } else {
#pragma unroll
                for (int m = 0; m < M; m++) d[m] = 0u;
            }
            lds_write<M>(lds, (uint32_t)(row * px + lx) * LB, d);
        }
    }
}

// a * b + c for |a|, |b| < 2^23 (v_mad_i32_i24, one full-rate instruction;
// written out because the compiler otherwise emits two multiplies and a
// three-way add for the cell index).
__device__ __forceinline__ int mad_i24(int a, int b, int c)
{
    int r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// The samples of one chunk from its staged box (row pitch PX cells, or px
// when PX is 0; plane pitch pitch * sy).  A sample's cell offset from the
// anchor's cell becomes one cell index and then bytes; the y + 1 rows lie
// one row pitch further (an immediate offset when PX is fixed) and the
// z + 1 rows one plane further.
template <int M, int ST, int PX>
__device__ __forceinline__ void march_staged(const RaySetup &R, const Anchor &A0, FixPos q, const FixPos &d16,
                                             int k0, int chunk, int c_lo, int c_hi, const int (&lo)[3], int px,
                                             int sy, const char *lds, float (&acc)[M])
{
    constexpr uint32_t LB = lds_cell_bytes<M>();
    const int pitch = PX > 0 ? PX : px;
    const int plane = pitch * sy;
    const uint32_t ry = (uint32_t)pitch * LB, rz = (uint32_t)plane * LB;
    for (int ka = k0; ka < k0 + chunk; ka += kAnchor, fix_add(q, d16)) {
        const int i1 = min(ka + kAnchor, c_hi) - ka;
        int i = max(ka, c_lo) - ka;
        if (i >= i1) continue;
        const Anchor A = ka == k0 ? A0 : anchor_from(q, R);
        // the anchor's cell in the box
        const char *pa = lds + (((A.ic - lo[2]) * sy + (A.ib - lo[1])) * pitch + (A.ia - lo[0])) * (int)LB;
        auto sample = [&](float fi) {
            const Cell c = offset_cell(A, fi);
            const int ci = mad_i24(c.z, plane, mad_i24(c.y, pitch, c.x));
            const char *p0 = pa + ci * (int)LB;
            const char *p2 = p0 + rz;
            accumulate<M, ST>(c, [&](int r, uint32_t(&out)[M]) {
                lds_read<M>((r & 2) ? p2 : p0, ((r & 1) ? ry : 0u) + (r < 4 ? 0u : LB), out);
            }, acc);
        };
        if (i == 0 && i1 == kAnchor) { // a whole segment (most of them)
#pragma unroll
            for (int s = 0; s < kAnchor; s++) sample((float)s);
        } else {
            for (float fi = (float)i; i < i1; i++, fi += 1.0f) sample(fi);
        }
    }
}

// True when every cell the chunk's box [lo, hi] stages (16-bit cells hold
// the x + 1 voxel, so one cell fewer in x) is zero in every channel: all
// bricks of the occupancy grid that overlap the box, clipped to the padded
// volume, are empty (cells outside it read the padding's zeros).  The chunk
// would add exact zeros to every ray, so it can be skipped.  Every wave
// computes the same answer from the block-uniform box, one brick per lane.
template <int ST>
__device__ __forceinline__ bool chunk_empty(const MarchArgs &P, const int (&lo)[3], const int (&hi)[3])
{
    const int x0 = max(lo[0] + kVoxPadLo, 0), x1 = min(hi[0] + kVoxPadLo - (ST == kStF32 ? 0 : 1), P.pcx - 1);
    const int y0 = max(lo[1] + kVoxPadLo, 0), y1 = min(hi[1] + kVoxPadLo, P.pcy - 1);
    const int z0 = max(lo[2] + kVoxPadLo, 0), z1 = min(hi[2] + kVoxPadLo, P.pcz - 1);
    if (x0 > x1 || y0 > y1 || z0 > z1) return true;
    const int bx0 = x0 / kOccBrick, by0 = y0 / kOccBrick, bz0 = z0 / kOccBrick;
    const int nbx = x1 / kOccBrick - bx0 + 1, nby = y1 / kOccBrick - by0 + 1, nbz = z1 / kOccBrick - bz0 + 1;
    if (nbx > 4 || nby > 4 || nbz > 4) return false; // a box this large is marched
    const int lane = threadIdx.x & 63;
    const int jx = lane & 3, jy = (lane >> 2) & 3, jz = lane >> 4;
    bool any = false;
    // 24-bit products: a padded volume has fewer than 2^24 bricks per axis
    // and fewer than 2^32 in all (2^41 cells)
    if (jx < nbx && jy < nby && jz < nbz)
        any = P.occ[__umul24(__umul24((uint32_t)(bz0 + jz), (uint32_t)P.ob_y) + (uint32_t)(by0 + jy),
                             (uint32_t)P.ob_x) +
                    (uint32_t)(bx0 + jx)] != 0;
    return __builtin_amdgcn_ballot_w64(any) == 0;
}

// Ray-march of one block's rays with the volume staged in LDS per chunk of
// samples: the rays move together along the common sample lattice; for each
// chunk of `chunk` samples the box of cells they touch is loaded once,
// row-coalesced, and all 8 (or 4) corner reads come from LDS.  A chunk whose
// box exceeds the LDS budget reads from global memory instead.  Each ray still
// accumulates in ascending k, so the result equals march_kernel's bit for bit.
// Every thread of the block must call it (it synchronises the block).
// kCount: no fill and no march; `marched` receives the lane's samples in the
// chunks that would be marched (the decisions are the march's own).
template <int M, int ST, bool W, bool kCount = false>
__device__ __forceinline__ void march_chunks(const MarchArgs &P, const RaySetup &R,
                                             __amdgpu_buffer_rsrc_t rsrc, char *lds, int *boxbuf,
                                             float (&acc)[M], int *marched = nullptr)
{
    const uint32_t oy = P.row_y & 0xFFFFFFu, oz = P.row_z & 0xFFFFFFu;
    const int nx = (int)P.nx, ny = (int)P.ny, nz = (int)P.nz;
    const int k_hi = R.k_lo + R.n;
    int kr_lo[1] = {R.n > 0 ? R.k_lo : INT_MAX}, kr_hi[1] = {R.n > 0 ? k_hi : INT_MIN};
    // the block's previous users of the scratch have finished two barriers
    // ago; the barrier below publishes the reset buffers
    if (threadIdx.x == 0) {
#pragma unroll
        for (int b = 0; b < kBoxBufs; b++) box_reset(boxbuf + 8 * b);
    }
    __syncthreads();
    int call = 0;
    block_box_atomic<1>(kr_lo, kr_hi, boxbuf, call);
    const int kb_hi = kr_hi[0];
    // the chunk's first anchor walks in exact fixed-point steps
    const int kstart = kr_lo[0] & ~(kAnchor - 1);
    const FixPos d16 = fix_steps(R, kAnchor), dchunk = fix_steps(R, P.chunk);
    FixPos pk = fix_at(R, kstart);
    // chunks start on the anchor grid (chunk is a multiple of kAnchor)
    for (int k0 = kstart; k0 < kb_hi; k0 += P.chunk, fix_add(pk, dchunk)) {
        const int c_lo = max(k0, R.k_lo), c_hi = min(k0 + P.chunk, k_hi);
        // the chunk's cell box from the lane's FP32 sample positions; the
        // first anchor is kept for the march
        int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
        const Anchor A0 = anchor_from(pk, R);
        {
            const int i0 = max(k0, c_lo) - k0, i1 = min(k0 + kAnchor, c_hi) - k0;
            if (i0 < i1) segment_box(A0, i0, i1, lo, hi);
        }
        {
            FixPos q = pk;
            for (int ka = k0 + kAnchor; ka < k0 + P.chunk; ka += kAnchor) {
                fix_add(q, d16);
                const int i0 = max(ka, c_lo) - ka, i1 = min(ka + kAnchor, c_hi) - ka;
                if (i0 < i1) segment_box(anchor_from(q, R), i0, i1, lo, hi);
            }
        }
        // every wave has finished the previous chunk past this barrier
        call = call == kBoxBufs - 1 ? 0 : call + 1;
        block_box_atomic<3>(lo, hi, boxbuf, call);
        if (hi[0] < lo[0]) continue; // no lane of the block samples this chunk (block-uniform)
        if (P.occ && chunk_empty<ST>(P, lo, hi)) continue; // all zeros (block-uniform)
        if constexpr (kCount) {
            *marched += max(0, c_hi - c_lo);
            continue;
        }
        // 16-bit cells are x-pairs that already hold the +1 x corner
        const int sx = hi[0] - lo[0] + (ST == kStF32 ? 1 : 0), sy = hi[1] - lo[1] + 1, sz = hi[2] - lo[2] + 1;
        // (a box of a volume past 4 GiB is staged only if its planes span
        // less than 4 GiB, so its row offsets fit 32 bits)
        const bool staged = (int64_t)sx * sy * sz <= (int64_t)P.lds_cells &&
                            (!W || (uint64_t)sz * oz < 0xFFFFFFFFull);
        if (staged) {
            // Row pitch: 12 cells for boxes up to 12 wide when that still
            // fits (then the y + 1 rows are an immediate LDS offset away),
            // sx otherwise (block-uniform).
            const int px = sx <= 12 && (int64_t)12 * sy * sz <= (int64_t)P.lds_cells ? 12 : sx;
            fill_box<M, W>(rsrc, (const char *)P.vol, lds, lo, sx, px, sy, sz, nx, ny, nz, oy, oz);
            __syncthreads();
            if (px == 12)
                march_staged<M, ST, 12>(R, A0, pk, d16, k0, P.chunk, c_lo, c_hi, lo, px, sy, lds, acc);
            else
                march_staged<M, ST, 0>(R, A0, pk, d16, k0, P.chunk, c_lo, c_hi, lo, px, sy, lds, acc);
        } else {
            FixPos q = pk;
            for (int ka = k0; ka < k0 + P.chunk; ka += kAnchor, fix_add(q, d16)) {
                const int i1 = min(ka + kAnchor, c_hi) - ka;
                int i = max(ka, c_lo) - ka;
                if (i >= i1) continue;
                const Anchor A = ka == k0 ? A0 : anchor_from(q, R);
                for (float fi = (float)i; i < i1; i++, fi += 1.0f)
                    accumulate_global<M, ST, W>(rsrc, (const char *)P.vol,
                                                absolute_cell(A, offset_cell(A, fi), nx, ny, nz), oy, oz, acc);
            }
        }
    }
}

// kBox: clip to the occupied bricks (the LDS march and what mirrors it).
template <bool kBox = false>
__device__ __forceinline__ RaySetup lane_ray(const MarchArgs &P, const TileLane &t)
{
    return t.active ? setup_ray<kBox>(P, t.u, t.v, (int)(P.view_begin + t.vw_local)) : empty_ray();
}

// One ray tile (logical block `logical` of P) through the LDS-staged march.
// The first 256 bytes of `lds_raw` are block scratch, the rest staged cells.
// kWide: also instantiate the march for volumes of 4 GiB or more (the batch
// kernel, which inlines every material count, leaves it out).
template <int M, int ST, bool kSpectrum, bool kWide = true>
__device__ __forceinline__ void march_tile(const MarchArgs &P, uint32_t logical, char *lds_raw)
{
    const TileLane t = tile_lane(P, logical);
    const RaySetup R = lane_ray<true>(P, t);
    __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(P.vol), 0, (int)P.vol_bytes, 0x00020000);
    float acc[M];
#pragma unroll
    for (int m = 0; m < M; m++) acc[m] = 0.0f;
    if (kWide && P.wide)
        march_chunks<M, ST, kWide>(P, R, rsrc, lds_raw + 256, (int *)lds_raw, acc);
    else
        march_chunks<M, ST, false>(P, R, rsrc, lds_raw + 256, (int *)lds_raw, acc);
    finish_ray<M, ST, kSpectrum>(P, t, acc);
}

template <int M, int ST, bool kSpectrum>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void march_lds_kernel(MarchArgs P)
{
    extern __shared__ char lds_raw[];
    const uint32_t logical = xcd_logical(P);
    if (logical >= P.n_blocks) return;
    march_tile<M, ST, kSpectrum>(P, logical, lds_raw);
}

// Samples the march evaluates (each ray's clipped window [k_lo, k_hi)),
// summed per workgroup into *count.
__global__ __launch_bounds__(kBlock) void count_kernel(MarchArgs P, unsigned long long *count)
{
    __shared__ unsigned long long part[kBlock / 64];
    const uint32_t logical = blockIdx.x;
    if (logical >= P.n_blocks) return;
    const TileLane t = tile_lane(P, logical);
    const RaySetup R = lane_ray(P, t);
    unsigned long long n = (unsigned long long)R.n;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) n += __shfl_xor(n, s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int w = 0; w < kBlock / 64; w++) tot += part[w];
        atomicAdd(count, tot);
    }
}

// Samples in the chunks the LDS march evaluates (rays' windows, less the
// chunks it skips as empty), summed per workgroup into *count.
template <int ST>
__global__ __launch_bounds__(kBlock) void count_marched_kernel(MarchArgs P, unsigned long long *count)
{
    __shared__ int boxbuf[64];
    __shared__ unsigned long long part[kBlock / 64];
    const uint32_t logical = xcd_logical(P);
    if (logical >= P.n_blocks) return;
    const TileLane t = tile_lane(P, logical);
    const RaySetup R = lane_ray<true>(P, t);
    __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(P.vol), 0, (int)P.vol_bytes, 0x00020000);
    float acc[1] = {0.0f};
    int marched = 0;
    march_chunks<1, ST, false, true>(P, R, rsrc, nullptr, boxbuf, acc, &marched);
    unsigned long long n = (unsigned long long)marched;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) n += __shfl_xor(n, s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int w = 0; w < kBlock / 64; w++) tot += part[w];
        atomicAdd(count, tot);
    }
}

// ------------------------------------------------------------ batches ---
//
// Several scans (contexts, each with its own volume, detector, views and
// spectrum) in one persistent launch.  The work units are ray tiles; unit
// queue x (one per XCD) holds, for every item, the tiles of the views
// v = x (mod 8) in item order, so the tiles of one view stay on one XCD
// (one L2).  Each workgroup drains the queue of its own XCD, then the other
// seven in turn; it leaves when every queue is empty, so all waves end.

constexpr int kXcds = 8;

struct BatchItem {
    MarchArgs P;
    int32_t n_mat;
};

struct BatchQueues {
    uint32_t *taken;            // [kXcds] units handed out per queue (zeroed before a launch)
    const uint32_t *unit_base;  // [kXcds][n_items + 1] first unit of each item in queue x
    int32_t n_items;
};

__global__ void batch_reset(uint32_t *taken, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) taken[i] = 0u;
}

template <int ST>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void batch_kernel(const BatchItem *__restrict__ items, BatchQueues Q)
{
    extern __shared__ char lds_raw[];
    uint32_t *slot = (uint32_t *)lds_raw + 60; // beside the march's box scratch (ints 0..23)
    const uint32_t home = blockIdx.x & (kXcds - 1);
    const int stride = Q.n_items + 1;
    for (uint32_t qi = 0; qi < (uint32_t)kXcds; qi++) {
        const uint32_t x = (home + qi) & (kXcds - 1);
        const uint32_t *base = Q.unit_base + x * stride;
        const uint32_t total = base[Q.n_items];
        for (;;) {
            __syncthreads(); // the previous unit has finished with the LDS
            if (threadIdx.x == 0) *slot = atomicAdd(&Q.taken[x], 1u);
            __syncthreads();
            const uint32_t u = *slot;
            if (u >= total) break;
            int i = 0;
            while (u >= base[i + 1]) i++;
            const BatchItem &it = items[i];
            const uint32_t r = u - base[i];
            const uint32_t tiles = (uint32_t)(it.P.tiles_u * it.P.tiles_v);
            const uint32_t logical = ((r / tiles) * kXcds + x) * tiles + r % tiles;
            switch (it.n_mat) {
            case 1: march_tile<1, ST, true, false>(it.P, logical, lds_raw); break;
            case 2: march_tile<2, ST, true, false>(it.P, logical, lds_raw); break;
            case 3: march_tile<3, ST, true, false>(it.P, logical, lds_raw); break;
            default: march_tile<4, ST, true, false>(it.P, logical, lds_raw); break;
            }
        }
    }
}

using BatchFn = void (*)(const BatchItem *, BatchQueues);

// ------------------------------------------------------------ the adjoint ---

struct AdjointArgs {
    const float *gI, *gIA; // [views][nv][nu] or null (zero)
    const float *L;        // [views][M][nv][nu] or null (recomputed)
    float *grad;           // [M][nz][ny][nx], accumulated
};

// Backprojection of one block's rays: every sample adds g_m * w_corner to
// the 8 corners of its cell.  Per chunk the block accumulates into an LDS copy
// of the chunk's cell box (ds_add_f32), then adds the box's non-zero cells to
// the gradient with one hardware FP32 atomic per cell and material.  A chunk
// whose box exceeds the LDS budget adds straight to global memory.
template <int M>
__device__ __forceinline__ void scatter_chunks(const MarchArgs &P, const RaySetup &R, const float (&g)[M],
                                               float *grad, char *lds, int *boxbuf)
{
    constexpr uint32_t LW = lds_cell_bytes<M>() / 4; // floats per staged cell
    float *box = (float *)lds;
    const int nx = (int)P.nx, ny = (int)P.ny, nz = (int)P.nz;
    const int64_t nvox = (int64_t)nx * ny * nz;
    const int k_hi = R.k_lo + R.n;
    int kr_lo[3] = {R.n > 0 ? R.k_lo : INT_MAX, 0, 0}, kr_hi[3] = {R.n > 0 ? k_hi : INT_MIN, 0, 0};
    __syncthreads();
    block_box(kr_lo, kr_hi, boxbuf);
    const int kb_hi = kr_hi[0];
    auto add_global = [&](int x, int y, int z, float w) {
        if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz || w == 0.0f) return;
        const int64_t q = ((int64_t)z * ny + y) * nx + x;
#pragma unroll
        for (int m = 0; m < M; m++)
            if (g[m] != 0.0f) unsafeAtomicAdd(grad + m * nvox + q, g[m] * w);
    };
    const int kstart = kr_lo[0] & ~(kAnchor - 1);
    const FixPos d16 = fix_steps(R, kAnchor), dchunk = fix_steps(R, P.chunk);
    FixPos pk = fix_at(R, kstart);
    for (int k0 = kstart; k0 < kb_hi; k0 += P.chunk, fix_add(pk, dchunk)) {
        const int c_lo = max(k0, R.k_lo), c_hi = min(k0 + P.chunk, k_hi);
        int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
        {
            FixPos q = pk;
            for (int ka = k0; ka < k0 + P.chunk; ka += kAnchor, fix_add(q, d16)) {
                const int i0 = max(ka, c_lo) - ka, i1 = min(ka + kAnchor, c_hi) - ka;
                if (i0 < i1) segment_box(anchor_from(q, R), i0, i1, lo, hi);
            }
        }
        __syncthreads();
        block_box(lo, hi, boxbuf);
        if (hi[0] < lo[0]) continue; // no lane of the block samples this chunk (block-uniform)
        const int sx = hi[0] - lo[0] + 1, sy = hi[1] - lo[1] + 1, sz = hi[2] - lo[2] + 1;
        const bool staged = (int64_t)sx * sy * sz <= (int64_t)P.lds_cells;
        if (staged) {
            const int cells = sx * sy * sz;
            for (int i = threadIdx.x; i < cells * (int)LW; i += kBlock) box[i] = 0.0f;
            __syncthreads();
            const int ry = sx * (int)LW, rz = sx * sy * (int)LW;
            const int roff[4] = {0, ry, rz, ry + rz};
            FixPos q = pk;
            for (int ka = k0; ka < k0 + P.chunk; ka += kAnchor, fix_add(q, d16)) {
                const int i1 = min(ka + kAnchor, c_hi) - ka;
                int i = max(ka, c_lo) - ka;
                if (i >= i1) continue;
                const Anchor A = anchor_from(q, R);
                const int base = (((A.ic - lo[2]) * sy + (A.ib - lo[1])) * sx + (A.ia - lo[0])) * (int)LW;
                for (float fi = (float)i; i < i1; i++, fi += 1.0f) {
                    const Cell c = offset_cell(A, fi);
                    const float bx = 1.0f - c.ax, by = 1.0f - c.ay, bz = 1.0f - c.az;
                    const float wr[4] = {by * bz, c.ay * bz, by * c.az, c.ay * c.az};
                    float *p = box + (base + c.x * (int)LW + __mul24(c.y, ry) + __mul24(c.z, rz));
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        float *p0 = p + roff[r];
#pragma unroll
                        for (int m = 0; m < M; m++) {
                            atomicAdd(p0 + m, g[m] * (bx * wr[r]));
                            atomicAdd(p0 + LW + m, g[m] * (c.ax * wr[r]));
                        }
                    }
                }
            }
            __syncthreads();
            const float inv_sx = small_rcp(sx), inv_sy = small_rcp(sy);
            for (int i = threadIdx.x; i < cells; i += kBlock) {
                const int row = small_div(i, sx, inv_sx), lx = i - row * sx;
                const int lz = small_div(row, sy, inv_sy), ly = row - lz * sy;
                const int x = lo[0] + lx, y = lo[1] + ly, z = lo[2] + lz;
                if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) continue;
                const int64_t q = ((int64_t)z * ny + y) * nx + x;
#pragma unroll
                for (int m = 0; m < M; m++) {
                    const float v = box[(uint32_t)i * LW + m];
                    if (v != 0.0f) unsafeAtomicAdd(grad + m * nvox + q, v);
                }
            }
        } else {
            FixPos q = pk;
            for (int ka = k0; ka < k0 + P.chunk; ka += kAnchor, fix_add(q, d16)) {
                const int i1 = min(ka + kAnchor, c_hi) - ka;
                int i = max(ka, c_lo) - ka;
                if (i >= i1) continue;
                const Anchor A = anchor_from(q, R);
                for (float fi = (float)i; i < i1; i++, fi += 1.0f) {
                    const Cell o = offset_cell(A, fi);
                    const int x = A.ia + o.x, y = A.ib + o.y, z = A.ic + o.z;
                    const float bx = 1.0f - o.ax, by = 1.0f - o.ay, bz = 1.0f - o.az;
                    const float wr[4] = {by * bz, o.ay * bz, by * o.az, o.ay * o.az};
                    const int dy[4] = {0, 1, 0, 1}, dz[4] = {0, 0, 1, 1};
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        add_global(x, y + dy[r], z + dz[r], bx * wr[r]);
                        add_global(x + 1, y + dy[r], z + dz[r], o.ax * wr[r]);
                    }
                }
            }
        }
    }
}

// Adjoint of ctfp_forward for upstream gradients gI, gIA: per ray, the line
// integrals (read or recomputed), then dl/dL_m = (gIA - gI) sum_e A[e][m]
// W_e e^-tau_e, then the backprojection of dx * dl/dL_m.
// dl/dL_m * dx of the lane's ray: the line integrals (read from G.L or
// recomputed by the march), then (gIA - gI) sum_e A[e][m] W_e e^-tau_e dx.
// Every thread of the block must call it (the march synchronises).
template <int M, int ST>
__device__ __forceinline__ void ray_gradient(const MarchArgs &P, const AdjointArgs &G, const TileLane &t,
                                             const RaySetup &R, char *lds_raw, float (&g)[M])
{
    const int64_t npix = (int64_t)P.nu * P.nv;
    const int64_t pix = (int64_t)t.v * P.nu + t.u;
    float Lm[M];
    if (G.L) {
#pragma unroll
        for (int m = 0; m < M; m++)
            Lm[m] = t.active ? G.L[((int64_t)t.vw_local * M + m) * npix + pix] : 0.0f;
    } else {
        __amdgpu_buffer_rsrc_t rsrc =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(P.vol), 0, (int)P.vol_bytes, 0x00020000);
        float acc[M];
#pragma unroll
        for (int m = 0; m < M; m++) acc[m] = 0.0f;
        if (P.wide)
            march_chunks<M, ST, true>(P, R, rsrc, lds_raw + 256, (int *)lds_raw, acc);
        else
            march_chunks<M, ST, false>(P, R, rsrc, lds_raw + 256, (int *)lds_raw, acc);
#pragma unroll
        for (int m = 0; m < M; m++) Lm[m] = acc[m] * P.dx_f;
    }
#pragma unroll
    for (int m = 0; m < M; m++) g[m] = 0.0f;
    if (t.active) {
        const int64_t q = (int64_t)t.vw_local * npix + pix;
        const float c = (G.gIA ? G.gIA[q] : 0.0f) - (G.gI ? G.gI[q] : 0.0f);
        if (c != 0.0f) {
            const float *wp = P.W + (int64_t)t.v * P.w_stride_v + t.u;
            KSum sm[M];
#pragma unroll
            for (int m = 0; m < M; m++) sm[m] = ksum_zero();
            // each group of four bins summed in FP32, then one compensated
            // add.  The weight rows past the last active bin are zeros
            // (kWeightPadBins of them), so a group past the end adds exact
            // zeros (its A rows are the last bin's).
            const int last = P.n_act - 1;
            for (int e0 = 0; e0 < P.n_act; e0 += 4) {
                float w4[4], part[M];
#pragma unroll
                for (int k = 0; k < 4; k++) w4[k] = wp[(int64_t)(e0 + k) * P.w_stride_e];
#pragma unroll
                for (int m = 0; m < M; m++) part[m] = 0.0f;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int e = min(e0 + k, last);
                    float tau = 0.0f;
#pragma unroll
                    for (int m = 0; m < M; m++) tau = fmaf(P.A[e * M + m], Lm[m], tau);
                    const float we = w4[k] * __builtin_amdgcn_exp2f(-tau * 1.44269504088896341f);
#pragma unroll
                    for (int m = 0; m < M; m++) part[m] = fmaf(P.A[e * M + m], we, part[m]);
                }
#pragma unroll
                for (int m = 0; m < M; m++) ksum_add(sm[m], part[m]);
            }
#pragma unroll
            for (int m = 0; m < M; m++) g[m] = c * ksum_value(sm[m]) * P.dx_f;
        }
    }
}

// Ray pass of the gather adjoint: writes dl/dL_m * dx per ray into
// gL [view - view_begin][m][v][u].
template <int M, int ST>
__global__ __launch_bounds__(kBlock) void ray_grad_kernel(MarchArgs P, AdjointArgs G, float *gL)
{
    extern __shared__ char lds_raw[];
    const uint32_t logical = xcd_logical(P);
    if (logical >= P.n_blocks) return;
    const TileLane t = tile_lane(P, logical);
    // the ray's samples only when its line integrals are recomputed
    const RaySetup R = G.L ? empty_ray() : lane_ray<true>(P, t);
    float g[M];
    ray_gradient<M, ST>(P, G, t, R, lds_raw, g);
    if (t.active) {
        const int64_t npix = (int64_t)P.nu * P.nv;
        const int64_t pix = (int64_t)t.v * P.nu + t.u;
#pragma unroll
        for (int m = 0; m < M; m++) gL[((int64_t)t.vw_local * M + m) * npix + pix] = g[m];
    }
}

// Voxel pass of the gather adjoint.  Each thread owns a 2 x 2 quad of voxels
// (in x and y) of a 16 x 16 x 4 brick and sums, over the views, every sample
// of every ray whose trilinear support (the open cube of side 2 around a
// voxel) holds it:
//   grad_m(c) += sum_rays gL_m(ray) sum_k prod_axes max(0, 1 - |p_k - c|).
// Per view the block stages, for each detector pixel of the brick's shadow,
// the ray's closest approach to the brick centre (FP64-derived offset e and
// sample index k*), its per-sample step and the ray's gL, so that the voxel
// loop runs in FP32 on small numbers (positions good to ~1e-6 voxel).  The
// quad's four voxels see nearly the same pixels, so they are walked together:
// one pixel window (their four joined), one sample interval per pixel (the
// union of theirs) and the z hat shared.  Each voxel writes its own
// gradient: no atomics, and the result is deterministic.
struct GatherArgs {
    MarchArgs P;       // geometry of the scan (views [view_begin, +n_views))
    const float *gL;   // [n_views][gl_mats][nv][nu]
    float *grad;       // [gl_mats][nz][ny][nx], accumulated
    int32_t gl_mats;   // materials of the scan
    int32_t mat_off;   // this launch's first material (M of them)
    int32_t bricks_x, bricks_y;
    uint32_t brick_base; // brick of this launch's first workgroup
    int32_t cap;       // staged pixels that fit the LDS budget
    float *part;       // cell gather: [brick][M][corners] sums, or null
};

// Each thread owns a quad of 2 x 2 voxels (x, y); a wavefront holds one z
// layer of 8 x 8 quads, a workgroup the brick's four layers.
constexpr int kBrickX = 16, kBrickY = 16, kBrickZ = 4;
constexpr int kGatherBlock = 256;
// LDS of a gather workgroup: the view table and the staged pixel records
constexpr int kGatherLds = 40 * 1024;

// A brick's shadow in one view: the projection of its centre, the Jacobian
// (pixels per voxel step), the half-extent of one voxel's support in pixels
// and the shadow's pixel rectangle.
struct BrickView {
    float u0, v0, jua, jub, juc, jva, jvb, jvc, hu, hv;
    int ru0, ru1, rv0, rv1;
};
// Views whose brick shadows a workgroup computes at once (one lane each).
constexpr int kGatherViews = 64;

// project_point in FP32, for bounds only: detector coordinates to ~1e-4
// pixel, so differences of them resolve curvature to about that.
__device__ __forceinline__ void project_point_f32(const MarchArgs &P, float cs_x, float cs_y, float fx, float fy,
                                                  float a, float b, float c, float &u, float &v)
{
    const float qx = a * (float)(1.0 / P.inv_px) - (float)P.half_x + (float)P.sod * cs_x;
    const float qy = b * (float)(1.0 / P.inv_py) - (float)P.half_y - (float)P.sod * cs_y;
    const float qz = c * (float)(1.0 / P.inv_pz) - (float)P.half_z;
    float X = qx * cs_x - qy * cs_y, Y = -qx * cs_y - qy * cs_x, Z = -qz;
    if (P.focal) {
        const float sdd = (float)P.sdd, lam = sdd / X;
        Y = fx + lam * (Y - fx);
        Z = fy + lam * (Z - fy);
        X = sdd;
    }
    const float al = Z / sqrtf(X * X + Y * Y), be = Y / X;
    const float den = 1.0f - al * al * be * be;
    const float detV = al * (float)P.sdd * sqrtf((1.0f + be * be) / den);
    const float detU = be * (float)P.sdd * sqrtf((1.0f + al * al) / den);
    const float ds = (float)P.dsize;
    u = ((float)P.half_u - 0.5f * ds - detU) / ds;
    v = ((float)P.half_v - 0.5f * ds - detV) / ds;
}

// The shadow of a brick in view `view` for elements (voxel supports or
// cells) of half-width `s` voxels whose centres lie within (ha, hb, hc) of
// the brick centre (ca, cb, cc): the centre's projection, the Jacobian per
// voxel step (central differences, FP64), an element's window half-extent
// and the brick's pixel rectangle.  Both widen by a bound on how far the
// projection departs from that linear model over the brick (reach
// a_i = h_i + s): |e(r)| <= 1/2 sum_ij |H_ij| a_i a_j, the Hessian entries from central differences of projections at the six
// face centres (H_ii a_i^2) and twelve edge midpoints (4 H_ij a_i a_j from
// the four sign pairs), times 1.25 for the cubic terms, plus 1e-3 pixel for
// the FP32 projections.  At magnification 4 with 1.5 mm voxels and 0.6 mm
// pixels the bound reaches about a pixel; for common scans it is a few
// hundredths.
__device__ __forceinline__ BrickView brick_view(const MarchArgs &P, int view, double ca, double cb, double cc,
                                                double ha, double hb, double hc, double s)
{
    double u0, v0, ju[3], jv[3];
    project_point(P, view, ca, cb, cc, u0, v0);
#pragma nounroll
    for (int i = 0; i < 3; i++) {
        double up, vp, um, vm;
        project_point(P, view, ca + (i == 0 ? 1.0 : 0.0), cb + (i == 1 ? 1.0 : 0.0), cc + (i == 2 ? 1.0 : 0.0), up, vp);
        project_point(P, view, ca - (i == 0 ? 1.0 : 0.0), cb - (i == 1 ? 1.0 : 0.0), cc - (i == 2 ? 1.0 : 0.0), um, vm);
        const double du = 0.5 * (up - um), dv = 0.5 * (vp - vm);
        ju[0] = i == 0 ? du : ju[0];
        ju[1] = i == 1 ? du : ju[1];
        ju[2] = i == 2 ? du : ju[2];
        jv[0] = i == 0 ? dv : jv[0];
        jv[1] = i == 1 ? dv : jv[1];
        jv[2] = i == 2 ? dv : jv[2];
    }
    const double a[3] = {ha + s, hb + s, hc + s};
    const float cx = (float)P.view_cs[view].x, cy = (float)P.view_cs[view].y;
    const float fx = P.focal ? (float)P.focal[view].x : 0.0f, fy = P.focal ? (float)P.focal[view].y : 0.0f;
    const float fa = (float)ca, fb = (float)cb, fc = (float)cc;
    float pu0, pv0;
    project_point_f32(P, cx, cy, fx, fy, fa, fb, fc, pu0, pv0);
    const float a0 = (float)a[0], a1 = (float)a[1], a2 = (float)a[2];
    // one projection at a time (loops kept rolled: the kernels' registers
    // belong to their sample loops)
    float bu = 0.0f, bv = 0.0f;
#pragma nounroll
    for (int i = 0; i < 3; i++) {
        const float ai = i == 0 ? a0 : i == 1 ? a1 : a2;
        const float ox = i == 0 ? ai : 0.0f, oy = i == 1 ? ai : 0.0f, oz = i == 2 ? ai : 0.0f;
        float pu, pv, mu, mv;
        project_point_f32(P, cx, cy, fx, fy, fa + ox, fb + oy, fc + oz, pu, pv);
        project_point_f32(P, cx, cy, fx, fy, fa - ox, fb - oy, fc - oz, mu, mv);
        // 1/2 |H_ii| a_i^2
        bu += 0.5f * fabsf(pu + mu - 2.0f * pu0);
        bv += 0.5f * fabsf(pv + mv - 2.0f * pv0);
    }
#pragma nounroll
    for (int q = 0; q < 3; q++) {
        // axis pairs (0, 1), (0, 2), (1, 2)
        const int i = q == 2 ? 1 : 0, j = q == 0 ? 1 : 2;
        const float ai = i == 0 ? a0 : a1, aj = j == 1 ? a1 : a2;
        float hu = 0.0f, hv = 0.0f;
#pragma nounroll
        for (int k = 0; k < 4; k++) {
            const float si = (k & 1) ? -ai : ai, sj = (k & 2) ? -aj : aj;
            float du, dv;
            project_point_f32(P, cx, cy, fx, fy, fa + (i == 0 ? si : 0.0f), fb + (i == 1 ? si : 0.0f) + (j == 1 ? sj : 0.0f),
                              fc + (j == 2 ? sj : 0.0f), du, dv);
            const float sg = ((k & 1) ^ ((k >> 1) & 1)) ? -1.0f : 1.0f;
            hu += sg * du;
            hv += sg * dv;
        }
        // |H_ij| a_i a_j, a quarter of the four-point sum
        bu += 0.25f * fabsf(hu);
        bv += 0.25f * fabsf(hv);
    }
    bu = 1.25f * bu + 1e-3f;
    bv = 1.25f * bv + 1e-3f;
    BrickView w;
    w.u0 = (float)u0;
    w.v0 = (float)v0;
    w.jua = (float)ju[0];
    w.jub = (float)ju[1];
    w.juc = (float)ju[2];
    w.jva = (float)jv[0];
    w.jvb = (float)jv[1];
    w.jvc = (float)jv[2];
    // an element spans +-s voxels around its centre: +-h pixels under the
    // linear model, plus the bound and 0.01 pixel for the FP32 centre and
    // Jacobian the kernels use
    w.hu = (float)(s * (fabs(ju[0]) + fabs(ju[1]) + fabs(ju[2])) + 0.01 + bu);
    w.hv = (float)(s * (fabs(jv[0]) + fabs(jv[1]) + fabs(jv[2])) + 0.01 + bv);
    const double eu = fabs(ju[0]) * a[0] + fabs(ju[1]) * a[1] + fabs(ju[2]) * a[2] + 0.01 + bu;
    const double ev = fabs(jv[0]) * a[0] + fabs(jv[1]) * a[1] + fabs(jv[2]) * a[2] + 0.01 + bv;
    w.ru0 = max(0, (int)ceil(u0 - eu));
    w.ru1 = min(P.nu - 1, (int)floor(u0 + eu));
    w.rv0 = max(0, (int)ceil(v0 - ev));
    w.rv1 = min(P.nv - 1, (int)floor(v0 + ev));
    return w;
}

// Staged dwords per shadow pixel, one record per pixel: the support bounds
// at the brick centre in j = k - floor(k*) (on x and y already widened to a
// quad: the union over offsets 0 and 1) and the inverse step (3 axes each),
// the lattice's range in j (-floor(k*) and K - 1 - floor(k*), as ints; read
// for every pixel of a quad's window), then k* - floor(k*), the
// closest-approach offset and the step (3 axes each) and dl/dL_m * dx (read
// for the pixels that touch the quad).  An odd record length keeps the
// lanes' reads of different records on different LDS banks.
__host__ __device__ constexpr int gather_pixel_words(int m) { return (18 + m) | 1; }
constexpr int kGLo = 0, kGHi = 3, kGId = 6, kGJ0 = 9, kGJ1 = 10, kGKr = 11, kGE = 12, kGD = 15, kGG = 18;

// hat(r) = max(0, 1 - |r|), one clamped subtraction
__device__ __forceinline__ float hat(float r) { return __builtin_amdgcn_fmed3f(1.0f - fabsf(r), 0.0f, 1.0f); }

// The thread's quad of voxels at offsets (ox + {0, 1}, oy + {0, 1}, oz) from
// the brick centre: for each staged pixel of rows [q0, q1] x columns
// [w0, w1] (records from `rec0`, row stride rs records), the samples inside
// any of the four voxels' supports (the staged x and y bounds hold the union
// of the two shifted intervals, so only the quad's offset is added),
// weighted into every voxel and then times the pixel's dl/dL into
// acc[y][x].  A sample outside a voxel's support has a zero hat there.  Per
// sample: three positions, five hats, two (y, z) products and four
// accumulations for the four voxels.
template <int M>
__device__ __forceinline__ void quad_pixels(const float *rec0, int rs, int q0, int q1, int w0, int w1, float ox,
                                            float oy, float oz, float (&acc)[2][2][M])
{
    constexpr int PW = gather_pixel_words(M);
    for (int pv = q0; pv <= q1; pv++) {
        const float *rec = rec0 + (pv - q0) * rs * PW;
        for (int pu = w0; pu <= w1; pu++, rec += PW) {
            const float iX = rec[kGId + 0], iY = rec[kGId + 1], iZ = rec[kGId + 2];
            // support in j = k - kf: the staged bounds shifted by o / d on
            // each axis
            const float jlo = fmaxf(fmaxf(fmaf(ox, iX, rec[kGLo + 0]), fmaf(oy, iY, rec[kGLo + 1])),
                                    fmaf(oz, iZ, rec[kGLo + 2]));
            const float jhi = fminf(fminf(fmaf(ox, iX, rec[kGHi + 0]), fmaf(oy, iY, rec[kGHi + 1])),
                                    fmaf(oz, iZ, rec[kGHi + 2]));
            // j over the open interval, and 0 <= k < K
            const int j0 = max((int)ceilf(jlo), __float_as_int(rec[kGJ0]));
            const int j1 = min((int)floorf(jhi), __float_as_int(rec[kGJ1]));
            if (j0 > j1) continue;
            const float kr = rec[kGKr];
            const float fX = rec[kGE + 0] - ox, fY = rec[kGE + 1] - oy, fZ = rec[kGE + 2] - oz;
            const float dX = rec[kGD + 0], dY = rec[kGD + 1], dZ = rec[kGD + 2];
            float w00 = 0.0f, w01 = 0.0f, w10 = 0.0f, w11 = 0.0f, sj = (float)j0 - kr;
            for (int j = j0; j <= j1; j++, sj += 1.0f) {
                const float hz = hat(fmaf(sj, dZ, fZ));
                const float ty = fmaf(sj, dY, fY), tx = fmaf(sj, dX, fX);
                const float hy0 = hat(ty) * hz, hy1 = hat(ty - 1.0f) * hz;
                const float hx0 = hat(tx), hx1 = hat(tx - 1.0f);
                w00 = fmaf(hy0, hx0, w00);
                w01 = fmaf(hy0, hx1, w01);
                w10 = fmaf(hy1, hx0, w10);
                w11 = fmaf(hy1, hx1, w11);
            }
#pragma unroll
            for (int m = 0; m < M; m++) {
                const float g = rec[kGG + m];
                acc[0][0][m] = fmaf(w00, g, acc[0][0][m]);
                acc[0][1][m] = fmaf(w01, g, acc[0][1][m]);
                acc[1][0][m] = fmaf(w10, g, acc[1][0][m]);
                acc[1][1][m] = fmaf(w11, g, acc[1][1][m]);
            }
        }
    }
}

template <int M>
__global__ __launch_bounds__(kGatherBlock) void gather_kernel(GatherArgs A)
{
    extern __shared__ char lds_raw[];
    const MarchArgs &P = A.P;
    const int cap = A.cap;
    constexpr int PW = gather_pixel_words(M);
    BrickView *s_view = (BrickView *)(lds_raw + 256); // kGatherViews entries
    // staged pixel records, `cap` of them
    float *s_px = (float *)(lds_raw + 256 + kGatherViews * sizeof(BrickView));
    const uint32_t brick = A.brick_base + blockIdx.x;
    const int bx = (int)(brick % (uint32_t)A.bricks_x);
    const int by = (int)((brick / (uint32_t)A.bricks_x) % (uint32_t)A.bricks_y);
    const int bz = (int)(brick / ((uint32_t)A.bricks_x * (uint32_t)A.bricks_y));
    // the thread's quad: voxels (x0 + {0, 1}, y0 + {0, 1}, z)
    const int lx = 2 * (threadIdx.x & 7), ly = 2 * ((threadIdx.x >> 3) & 7), lz = threadIdx.x >> 6;
    const int x0 = bx * kBrickX + lx, y0 = by * kBrickY + ly, z = bz * kBrickZ + lz;
    const int nx = (int)P.nx, ny = (int)P.ny, nz = (int)P.nz;
    const bool inside = x0 < nx && y0 < ny && z < nz; // the quad's first voxel
    // brick centre (index space) and the quad's first voxel's offset from it
    const double ca = bx * kBrickX + 0.5 * (kBrickX - 1), cb = by * kBrickY + 0.5 * (kBrickY - 1);
    const double cc = bz * kBrickZ + 0.5 * (kBrickZ - 1);
    const float ox = (float)(lx - 0.5 * (kBrickX - 1)), oy = (float)(ly - 0.5 * (kBrickY - 1));
    const float oz = (float)(lz - 0.5 * (kBrickZ - 1));
    const int64_t npix = (int64_t)P.nu * P.nv;
    float acc[2][2][M]; // [y - y0][x - x0][m]
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int m = 0; m < M; m++) acc[i][j][m] = 0.0f;
    for (int vt = 0; vt < P.n_views; vt += kGatherViews) {
        const int nvt = min(kGatherViews, P.n_views - vt);
        __syncthreads(); // the previous views' table is consumed
        if ((int)threadIdx.x < nvt) {
            // the brick's shadow in view vt + lane (one lane per view):
            // voxel centres within (7.5, 7.5, 1.5) of the brick centre,
            // supports +-1 voxel
            const BrickView bv = brick_view(P, (int)(P.view_begin + vt + threadIdx.x), ca, cb, cc,
                                            0.5 * (kBrickX - 1), 0.5 * (kBrickY - 1), 0.5 * (kBrickZ - 1), 1.0);
            s_view[threadIdx.x] = bv;
        }
        __syncthreads();
        for (int vw = vt; vw < vt + nvt; vw++) {
            const BrickView bv = s_view[vw - vt];
            const int ru0 = bv.ru0, ru1 = bv.ru1, rv0 = bv.rv0, rv1 = bv.rv1;
            if (ru0 > ru1 || rv0 > rv1) continue; // block-uniform
            // the quad's pixel window: its four voxels' windows joined
            // (inside the brick's)
            const float hu = bv.hu + 0.5f * (fabsf(bv.jua) + fabsf(bv.jub));
            const float hv = bv.hv + 0.5f * (fabsf(bv.jva) + fabsf(bv.jvb));
            const float cu = bv.u0 + bv.jua * (ox + 0.5f) + bv.jub * (oy + 0.5f) + bv.juc * oz;
            const float cv = bv.v0 + bv.jva * (ox + 0.5f) + bv.jvb * (oy + 0.5f) + bv.jvc * oz;
            const int pu0 = max(ru0, (int)ceilf(cu - hu)), pu1 = min(ru1, (int)floorf(cu + hu));
            const int pv0 = max(rv0, (int)ceilf(cv - hv)), pv1 = min(rv1, (int)floorf(cv + hv));
            // the brick's shadow in pieces of at most `cap` pixels: column
            // segments no wider than cap, each in bands of whole rows
            for (int c0 = ru0; c0 <= ru1; c0 += cap) {
                const int c1 = min(ru1, c0 + cap - 1);
                const int rw = c1 - c0 + 1;
                const int band = max(1, cap / rw);
                for (int b0 = rv0; b0 <= rv1; b0 += band) {
                    const int b1 = min(rv1, b0 + band - 1);
                    const int count = (b1 - b0 + 1) * rw;
                    __syncthreads(); // the previous band's entries are consumed
                    for (int i = threadIdx.x; i < count; i += kGatherBlock) {
                        const int pv = b0 + i / rw, pu = c0 + i % rw;
                        const RayLine Ln = ray_line(P, pu, pv, (int)(P.view_begin + vw));
                        // position at sample k: q0 + k * d, relative to the brick centre
                        const double dxa = Ln.da * P.dx, dxb = Ln.db * P.dx, dxc = Ln.dc * P.dx;
                        const double qa = Ln.sa + Ln.da * P.t0 - ca, qb = Ln.sb + Ln.db * P.t0 - cb;
                        const double qc = Ln.sc + Ln.dc * P.t0 - cc;
                        const double ks =
                            -(qa * dxa + qb * dxb + qc * dxc) / (dxa * dxa + dxb * dxb + dxc * dxc);
                        const double kf = floor(ks);
                        const float e[3] = {(float)(qa + ks * dxa), (float)(qb + ks * dxb), (float)(qc + ks * dxc)};
                        const float d[3] = {(float)dxa, (float)dxb, (float)dxc};
                        float *rec = s_px + i * PW;
                        const float kr = (float)(ks - kf);
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            // a zero step leaves that axis unconstrained (or
                            // empty) through +-huge bounds.  1/d to an ulp:
                            // a bound off by that moves a sample whose hat is
                            // ~1e-6 across the support's edge
                            const float id = __builtin_amdgcn_rcpf(fabsf(d[a]) > 1e-30f ? d[a] : 1e-30f);
                            // samples s = k - k* with |e - o + s d| < 1 lie
                            // between (+-1 - e) / d + o / d: the bounds at o = 0,
                            // stored as k - kf = s + kr
                            const float l = (-1.0f - e[a]) * id, h = (1.0f - e[a]) * id;
                            // on x and y, the union with the interval one
                            // voxel further (a quad's second column / row)
                            const float wl = a < 2 ? fminf(id, 0.0f) : 0.0f, wh = a < 2 ? fmaxf(id, 0.0f) : 0.0f;
                            rec[kGLo + a] = fminf(l, h) + kr + wl;
                            rec[kGHi + a] = fmaxf(l, h) + kr + wh;
                            rec[kGId + a] = id;
                            rec[kGE + a] = e[a];
                            rec[kGD + a] = d[a];
                        }
                        rec[kGJ0] = __int_as_float(-(int)kf);
                        rec[kGJ1] = __int_as_float(P.K - 1 - (int)kf);
                        rec[kGKr] = kr;
                        const int64_t pix = (int64_t)pv * P.nu + pu;
#pragma unroll
                        for (int m = 0; m < M; m++)
                            rec[kGG + m] = A.gL[((int64_t)vw * A.gl_mats + A.mat_off + m) * npix + pix];
                    }
                    __syncthreads();
                    if (!inside) continue;
                    const int q0 = max(pv0, b0), q1 = min(pv1, b1);
                    const int w0 = max(pu0, c0), w1 = min(pu1, c1);
                    if (q0 > q1 || w0 > w1) continue;
                    quad_pixels<M>(s_px + ((q0 - b0) * rw + (w0 - c0)) * PW, rw, q0, q1, w0, w1, ox, oy, oz, acc);
                }
            }
        }
    }
    const int64_t nvox = (int64_t)nx * ny * nz;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int x = x0 + j, y = y0 + i;
            if (x >= nx || y >= ny || z >= nz) continue;
            const int64_t q = ((int64_t)z * ny + y) * nx + x;
#pragma unroll
            for (int m = 0; m < M; m++) A.grad[(int64_t)(A.mat_off + m) * nvox + q] += acc[i][j][m];
        }
}

// Cell-driven gather.  Each thread owns one cell of an 8 x 8 x 4-cell brick:
// the cube between voxels (cx, cy, cz) and (cx + 1, cy + 1, cz + 1); a
// wavefront holds one z layer of the brick.  Every lattice sample lies in
// exactly one cell, whose thread weights it into the cell's eight corners
// with the trilinear weights (1 - t, t) per axis, so no weight is evaluated
// twice or found to be zero.  The samples of a cell come from the ray's
// plane crossings: on each axis they are the k with
// min(B(X), B(X + 1)) <= k < max(B(X), B(X + 1)), B(X) = fma(X, 1/d, Q) the
// (fractional) sample index at which the ray crosses plane X (absolute index
// space).  Every cell and brick that shares a plane computes its B from the
// same three operands, so the cells split the ray's samples without a gap or
// an overlap.  A sample within rounding of a plane may go to either cell;
// both weight it alike (t = 1 in one cell is t = 0 in the other) up to that
// rounding (the weights come from the sample's offset in the brick's frame).
// Bricks start at cell -1 (cx = bx * 8 - 1 + lx), so every
// cell [-1, n - 1] of each axis is someone's.  A brick's corner sums
// (9 x 9 x 5 voxels, summed in LDS at the end) go to `part`, and
// cell_combine_kernel adds the (up to eight) bricks' sums of each voxel to
// grad in a fixed order: no atomics, a deterministic result.
constexpr int kCellX = 8, kCellY = 8, kCellZ = 4;
constexpr int kCellBlock = kCellX * kCellY * kCellZ; // one cell per thread
constexpr int kCellCorners = (kCellX + 1) * (kCellY + 1) * (kCellZ + 1);

// Staged dwords per shadow pixel of the cell gather: the plane-crossing base
// Q = -q / d and 1 / d (3 axes each; q the ray's position at k = 0, index
// space, so every brick stages the same values and a plane X is crossed at
// sample k = fma(X, 1 / d, Q) in all of them), floor(k*) as an int and
// k* - floor(k*) (k* the sample closest to the brick centre), the offset e of
// that closest approach from the centre and the step d (3 axes each) and
// dl/dL_m * dx.  Odd length, as for the voxel gather's records.
__host__ __device__ constexpr int cell_pixel_words(int m) { return (14 + m) | 1; }
constexpr int kCQ = 0, kCId = 3, kCKf = 6, kCKr = 7, kCE = 8, kCD = 11, kCG = 14;

// The thread's cell, lower corner (cx, cy, cz) = (X, Y, Z) and (ox, oy, oz)
// from the brick centre: for each staged pixel of rows [q0, q1] x columns
// [w0, w1], the lattice samples between the ray's crossings of the cell's
// planes X, X + 1 (and Y, Z), weighted into the eight corners, times the
// pixel's dl/dL into acc[corner].  Per sample: three positions, three
// complements, four (y, z) products and eight accumulations.
template <int M>
__device__ __forceinline__ void cell_pixels(const float *rec0, int rs, int q0, int q1, int w0, int w1, float X,
                                            float Y, float Z, float ox, float oy, float oz, float Kf,
                                            float (&acc)[8][M])
{
    constexpr int PW = cell_pixel_words(M);
    for (int pv = q0; pv <= q1; pv++) {
        const float *rec = rec0 + (pv - q0) * rs * PW;
        for (int pu = w0; pu <= w1; pu++, rec += PW) {
            const float iX = rec[kCId + 0], iY = rec[kCId + 1], iZ = rec[kCId + 2];
            const float bX0 = fmaf(X, iX, rec[kCQ + 0]), bX1 = fmaf(X + 1.0f, iX, rec[kCQ + 0]);
            const float bY0 = fmaf(Y, iY, rec[kCQ + 1]), bY1 = fmaf(Y + 1.0f, iY, rec[kCQ + 1]);
            const float bZ0 = fmaf(Z, iZ, rec[kCQ + 2]), bZ1 = fmaf(Z + 1.0f, iZ, rec[kCQ + 2]);
            // k in [max of the entries, min of the exits), inside [0, K)
            const float lo = fmaxf(fmaxf(fminf(bX0, bX1), fminf(bY0, bY1)), fmaxf(fminf(bZ0, bZ1), 0.0f));
            const float hi = fminf(fminf(fmaxf(bX0, bX1), fmaxf(bY0, bY1)), fminf(fmaxf(bZ0, bZ1), Kf));
            const int k0 = (int)ceilf(lo), k1 = (int)ceilf(hi);
            if (k0 >= k1) continue;
            const float fX = rec[kCE + 0] - ox, fY = rec[kCE + 1] - oy, fZ = rec[kCE + 2] - oz;
            const float dX = rec[kCD + 0], dY = rec[kCD + 1], dZ = rec[kCD + 2];
            // s = k - k*, from the brick's closest approach
            float sj = (float)(k0 - __float_as_int(rec[kCKf])) - rec[kCKr];
            float w[8];
            {
                // the first sample (k0 < k1) sets the weights
                const float tx = fmaf(sj, dX, fX), ty = fmaf(sj, dY, fY), tz = fmaf(sj, dZ, fZ);
                const float ax = 1.0f - tx, ay = 1.0f - ty, az = 1.0f - tz;
                const float y0z0 = ay * az, y1z0 = ty * az, y0z1 = ay * tz, y1z1 = ty * tz;
                w[0] = y0z0 * ax;
                w[1] = y0z0 * tx;
                w[2] = y1z0 * ax;
                w[3] = y1z0 * tx;
                w[4] = y0z1 * ax;
                w[5] = y0z1 * tx;
                w[6] = y1z1 * ax;
                w[7] = y1z1 * tx;
                sj += 1.0f;
            }
            for (int k = k0 + 1; k < k1; k++, sj += 1.0f) {
                const float tx = fmaf(sj, dX, fX), ty = fmaf(sj, dY, fY), tz = fmaf(sj, dZ, fZ);
                const float ax = 1.0f - tx, ay = 1.0f - ty, az = 1.0f - tz;
                const float y0z0 = ay * az, y1z0 = ty * az, y0z1 = ay * tz, y1z1 = ty * tz;
                w[0] = fmaf(y0z0, ax, w[0]);
                w[1] = fmaf(y0z0, tx, w[1]);
                w[2] = fmaf(y1z0, ax, w[2]);
                w[3] = fmaf(y1z0, tx, w[3]);
                w[4] = fmaf(y0z1, ax, w[4]);
                w[5] = fmaf(y0z1, tx, w[5]);
                w[6] = fmaf(y1z1, ax, w[6]);
                w[7] = fmaf(y1z1, tx, w[7]);
            }
#pragma unroll
            for (int m = 0; m < M; m++) {
                const float g = rec[kCG + m];
#pragma unroll
                for (int c = 0; c < 8; c++) acc[c][m] = fmaf(w[c], g, acc[c][m]);
            }
        }
    }
}

template <int M>
__global__ __launch_bounds__(kCellBlock) __attribute__((amdgpu_waves_per_eu(4))) void cell_gather_kernel(GatherArgs A)
{
    extern __shared__ char lds_raw[];
    const MarchArgs &P = A.P;
    const int cap = A.cap;
    constexpr int PW = cell_pixel_words(M);
    BrickView *s_view = (BrickView *)(lds_raw + 256); // kGatherViews entries
    float *s_px = (float *)(lds_raw + 256 + kGatherViews * sizeof(BrickView));
    const uint32_t brick = A.brick_base + blockIdx.x;
    const int bx = (int)(brick % (uint32_t)A.bricks_x);
    const int by = (int)((brick / (uint32_t)A.bricks_x) % (uint32_t)A.bricks_y);
    const int bz = (int)(brick / ((uint32_t)A.bricks_x * (uint32_t)A.bricks_y));
    const int lx = threadIdx.x % kCellX, ly = (threadIdx.x / kCellX) % kCellY, lz = threadIdx.x / (kCellX * kCellY);
    // the cell's lower corner; cells -1 .. n - 1 of each axis take samples
    const int cx = bx * kCellX - 1 + lx, cy = by * kCellY - 1 + ly, cz = bz * kCellZ - 1 + lz;
    const bool inside = cx < (int)P.nx && cy < (int)P.ny && cz < (int)P.nz;
    // the brick's centre (index space) and the cell's lower corner from it;
    // the planes in absolute index space (exact floats, the same in every
    // cell and brick that shares one)
    const double ca = bx * kCellX - 1 + 0.5 * kCellX, cb = by * kCellY - 1 + 0.5 * kCellY;
    const double cc = bz * kCellZ - 1 + 0.5 * kCellZ;
    const float ox = (float)(lx - kCellX / 2), oy = (float)(ly - kCellY / 2), oz = (float)(lz - kCellZ / 2);
    const float X = (float)cx, Y = (float)cy, Z = (float)cz, Kf = (float)P.K;
    const int64_t npix = (int64_t)P.nu * P.nv;
    float acc[8][M];
#pragma unroll
    for (int c = 0; c < 8; c++)
#pragma unroll
        for (int m = 0; m < M; m++) acc[c][m] = 0.0f;
    for (int vt = 0; vt < P.n_views; vt += kGatherViews) {
        const int nvt = min(kGatherViews, P.n_views - vt);
        __syncthreads(); // the previous views' table is consumed
        if ((int)threadIdx.x < nvt) {
            // cell centres within (3.5, 3.5, 1.5) of the brick centre, cells
            // +-0.5 voxel
            const BrickView bv = brick_view(P, (int)(P.view_begin + vt + threadIdx.x), ca, cb, cc,
                                            0.5 * (kCellX - 1), 0.5 * (kCellY - 1), 0.5 * (kCellZ - 1), 0.5);
            s_view[threadIdx.x] = bv;
        }
        __syncthreads();
        for (int vw = vt; vw < vt + nvt; vw++) {
            const BrickView bv = s_view[vw - vt];
            const int ru0 = bv.ru0, ru1 = bv.ru1, rv0 = bv.rv0, rv1 = bv.rv1;
            if (ru0 > ru1 || rv0 > rv1) continue; // block-uniform
            // the cell's pixel window (inside the brick's)
            const float cu = bv.u0 + bv.jua * (ox + 0.5f) + bv.jub * (oy + 0.5f) + bv.juc * (oz + 0.5f);
            const float cv = bv.v0 + bv.jva * (ox + 0.5f) + bv.jvb * (oy + 0.5f) + bv.jvc * (oz + 0.5f);
            const int pu0 = max(ru0, (int)ceilf(cu - bv.hu)), pu1 = min(ru1, (int)floorf(cu + bv.hu));
            const int pv0 = max(rv0, (int)ceilf(cv - bv.hv)), pv1 = min(rv1, (int)floorf(cv + bv.hv));
            for (int c0 = ru0; c0 <= ru1; c0 += cap) {
                const int c1 = min(ru1, c0 + cap - 1);
                const int rw = c1 - c0 + 1;
                const int band = max(1, cap / rw);
                for (int b0 = rv0; b0 <= rv1; b0 += band) {
                    const int b1 = min(rv1, b0 + band - 1);
                    const int count = (b1 - b0 + 1) * rw;
                    __syncthreads(); // the previous band's entries are consumed
                    for (int i = threadIdx.x; i < count; i += kCellBlock) {
                        const int pv = b0 + i / rw, pu = c0 + i % rw;
                        const RayLine Ln = ray_line(P, pu, pv, (int)(P.view_begin + vw));
                        const double dxa = Ln.da * P.dx, dxb = Ln.db * P.dx, dxc = Ln.dc * P.dx;
                        // position at k = 0: absolute, and from the brick centre
                        const double pa = Ln.sa + Ln.da * P.t0, pb = Ln.sb + Ln.db * P.t0, pc = Ln.sc + Ln.dc * P.t0;
                        const double qa = pa - ca, qb = pb - cb, qc = pc - cc;
                        const double ks =
                            -(qa * dxa + qb * dxb + qc * dxc) / (dxa * dxa + dxb * dxb + dxc * dxc);
                        const double kf = floor(ks);
                        const double p[3] = {pa, pb, pc};
                        const float e[3] = {(float)(qa + ks * dxa), (float)(qb + ks * dxb), (float)(qc + ks * dxc)};
                        const float d[3] = {(float)dxa, (float)dxb, (float)dxc};
                        float *rec = s_px + i * PW;
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            // a zero step makes that axis's crossings +-huge:
                            // all of the ray's samples in the cells holding
                            // it, none in the others
                            const float id = __builtin_amdgcn_rcpf(fabsf(d[a]) > 1e-30f ? d[a] : 1e-30f);
                            rec[kCQ + a] = (float)(-p[a] * (double)id);
                            rec[kCId + a] = id;
                            rec[kCE + a] = e[a];
                            rec[kCD + a] = d[a];
                        }
                        rec[kCKf] = __int_as_float((int)kf);
                        rec[kCKr] = (float)(ks - kf);
                        const int64_t pix = (int64_t)pv * P.nu + pu;
#pragma unroll
                        for (int m = 0; m < M; m++)
                            rec[kCG + m] = A.gL[((int64_t)vw * A.gl_mats + A.mat_off + m) * npix + pix];
                    }
                    __syncthreads();
                    if (!inside) continue;
                    const int q0 = max(pv0, b0), q1 = min(pv1, b1);
                    const int w0 = max(pu0, c0), w1 = min(pu1, c1);
                    if (q0 > q1 || w0 > w1) continue;
                    cell_pixels<M>(s_px + ((q0 - b0) * rw + (w0 - c0)) * PW, rw, q0, q1, w0, w1, X, Y, Z, ox, oy,
                                   oz, Kf, acc);
                }
            }
        }
    }
    // the brick's corner sums, one corner of every cell at a time (the cells
    // of one pass add to distinct voxels), then out to `part`
    __syncthreads();
    float *V = (float *)lds_raw; // [M][kCellZ + 1][kCellY + 1][kCellX + 1]
    for (int i = threadIdx.x; i < M * kCellCorners; i += kCellBlock) V[i] = 0.0f;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int vi = ((lz + (c >> 2)) * (kCellY + 1) + ly + ((c >> 1) & 1)) * (kCellX + 1) + lx + (c & 1);
#pragma unroll
        for (int m = 0; m < M; m++) V[m * kCellCorners + vi] += acc[c][m];
        __syncthreads();
    }
    float *dst = A.part + (int64_t)brick * (M * kCellCorners);
    for (int i = threadIdx.x; i < M * kCellCorners; i += kCellBlock) dst[i] = V[i];
}

// grad[mat_off + m] += the cell gather's corner sums of each voxel: voxel x
// is corner (x + 1) % 8 of brick (x + 1) / 8 and, when that corner is 0,
// also corner 8 of the brick before (4 for z).  Fixed order: deterministic.
template <int M>
__global__ __launch_bounds__(256) void cell_combine_kernel(const float *__restrict__ part, float *grad, int nx, int ny,
                                                           int nz, int bricks_x, int bricks_y, int mat_off)
{
    const int64_t nvox = (int64_t)nx * ny * nz;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nvox;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(q % nx), y = (int)((q / nx) % ny), z = (int)(q / ((int64_t)nx * ny));
        const int bx = (x + 1) / kCellX, ix = (x + 1) % kCellX;
        const int by = (y + 1) / kCellY, iy = (y + 1) % kCellY;
        const int bz = (z + 1) / kCellZ, iz = (z + 1) % kCellZ;
        float s[M];
#pragma unroll
        for (int m = 0; m < M; m++) s[m] = 0.0f;
        for (int sz = 0; sz < (iz == 0 && bz > 0 ? 2 : 1); sz++)
            for (int sy = 0; sy < (iy == 0 && by > 0 ? 2 : 1); sy++)
                for (int sx = 0; sx < (ix == 0 && bx > 0 ? 2 : 1); sx++) {
                    const int64_t b = ((int64_t)(bz - sz) * bricks_y + (by - sy)) * bricks_x + (bx - sx);
                    const int vi = ((iz + sz * kCellZ) * (kCellY + 1) + iy + sy * kCellY) * (kCellX + 1) + ix +
                                   sx * kCellX;
                    const float *src = part + b * (M * kCellCorners) + vi;
#pragma unroll
                    for (int m = 0; m < M; m++) s[m] += src[m * kCellCorners];
                }
#pragma unroll
        for (int m = 0; m < M; m++) grad[(int64_t)(mat_off + m) * nvox + q] += s[m];
    }
}

template <int M, int ST>
__global__ __launch_bounds__(kBlock) void adjoint_kernel(MarchArgs P, AdjointArgs G)
{
    extern __shared__ char lds_raw[];
    const uint32_t logical = xcd_logical(P);
    if (logical >= P.n_blocks) return;
    const TileLane t = tile_lane(P, logical);
    const RaySetup R = lane_ray(P, t);
    float g[M];
    ray_gradient<M, ST>(P, G, t, R, lds_raw, g);
    scatter_chunks<M>(P, R, g, G.grad, lds_raw + 256, (int *)lds_raw);
}

// The spectrum of every ray from its line integrals L [view][n_mat][v][u]
// (scans of more than four materials, whose groups are marched one after the
// other): the arithmetic of finish_ray's per-ray FP32 spectrum for any
// n_mat.  The line integrals of the block's rays wait in LDS.
__global__ __launch_bounds__(kBlock) void spectrum_kernel(MarchArgs P, const float *__restrict__ L, int n_mat)
{
    __shared__ float sL[kMaxMaterials][kBlock];
    const int64_t npix = (int64_t)P.nu * P.nv;
    const int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (q >= (int64_t)P.n_views * npix) return;
    const int64_t vw = q / npix, pix = q % npix;
    const int u = (int)(pix % P.nu), v = (int)(pix / P.nu);
    for (int m = 0; m < n_mat; m++) sL[m][threadIdx.x] = L[(vw * n_mat + m) * npix + pix];
    const float *wp = P.W + (int64_t)v * P.w_stride_v + u;
    KSum sI = ksum_zero(), sA = ksum_zero();
    for (int e = 0; e < P.n_act; e++) {
        float tau = 0.0f;
        for (int m = 0; m < n_mat; m++) tau = fmaf(P.A[e * n_mat + m], sL[m][threadIdx.x], tau);
        float ex, ab;
        attenuate(tau, ex, ab);
        const float w = wp[(int64_t)e * P.w_stride_e];
        ksum_add(sI, w * ex);
        ksum_add(sA, w * ab);
    }
    P.I[q] = ksum_value(sI);
    P.IA[q] = ksum_value(sA);
}

// dl/dL_m * dx of every ray from its line integrals L [view][n_mat][v][u] and
// the upstream gradients (the arithmetic of ray_gradient for any n_mat), into
// gL [view][n_mat][v][u].  The per-material sums wait in LDS.
__global__ __launch_bounds__(kBlock) void ray_grad_from_l_kernel(MarchArgs P, AdjointArgs G, const float *__restrict__ L,
                                                                 int n_mat, float *__restrict__ gL)
{
    __shared__ float sL[kMaxMaterials][kBlock];
    __shared__ KSum sm[kMaxMaterials][kBlock];
    const int64_t npix = (int64_t)P.nu * P.nv;
    const int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (q >= (int64_t)P.n_views * npix) return;
    const int64_t vw = q / npix, pix = q % npix;
    const int u = (int)(pix % P.nu), v = (int)(pix / P.nu);
    const float c = (G.gIA ? G.gIA[q] : 0.0f) - (G.gI ? G.gI[q] : 0.0f);
    for (int m = 0; m < n_mat; m++) {
        sL[m][threadIdx.x] = L[(vw * n_mat + m) * npix + pix];
        sm[m][threadIdx.x] = ksum_zero();
    }
    if (c != 0.0f) {
        const float *wp = P.W + (int64_t)v * P.w_stride_v + u;
        for (int e = 0; e < P.n_act; e++) {
            float tau = 0.0f;
            for (int m = 0; m < n_mat; m++) tau = fmaf(P.A[e * n_mat + m], sL[m][threadIdx.x], tau);
            float ex, ab;
            attenuate(tau, ex, ab);
            const float we = wp[(int64_t)e * P.w_stride_e] * ex;
            for (int m = 0; m < n_mat; m++) ksum_add(sm[m][threadIdx.x], P.A[e * n_mat + m] * we);
        }
    }
    for (int m = 0; m < n_mat; m++)
        gL[(vw * n_mat + m) * npix + pix] = c != 0.0f ? c * ksum_value(sm[m][threadIdx.x]) * P.dx_f : 0.0f;
}

using MarchFn = void (*)(MarchArgs);

template <int ST> MarchFn pick_march_st(int n_mat, bool spectrum, bool lds)
{
#define CTFP_PICK(Mv)                                                                   \
    case Mv:                                                                            \
        if (lds) return spectrum ? march_lds_kernel<Mv, ST, true> : march_lds_kernel<Mv, ST, false>; \
        return spectrum ? march_kernel<Mv, ST, true> : march_kernel<Mv, ST, false>;
    switch (n_mat) {
        CTFP_PICK(1)
        CTFP_PICK(2)
        CTFP_PICK(3)
        CTFP_PICK(4)
    default: return nullptr;
    }
#undef CTFP_PICK
}

using AdjointFn = void (*)(MarchArgs, AdjointArgs);
using RayGradFn = void (*)(MarchArgs, AdjointArgs, float *);
using GatherFn = void (*)(GatherArgs);

template <int ST> RayGradFn pick_ray_grad_st(int n_mat)
{
    switch (n_mat) {
    case 1: return ray_grad_kernel<1, ST>;
    case 2: return ray_grad_kernel<2, ST>;
    case 3: return ray_grad_kernel<3, ST>;
    case 4: return ray_grad_kernel<4, ST>;
    default: return nullptr;
    }
}

RayGradFn pick_ray_grad(int n_mat, int dtype)
{
    switch (dtype) {
    case CTFP_F32: return pick_ray_grad_st<kStF32>(n_mat);
    case CTFP_F16: return pick_ray_grad_st<kStF16>(n_mat);
    case CTFP_BF16: return pick_ray_grad_st<kStBF16>(n_mat);
    default: return nullptr;
    }
}

GatherFn pick_gather(int n_mat)
{
    switch (n_mat) {
    case 1: return gather_kernel<1>;
    case 2: return gather_kernel<2>;
    case 3: return gather_kernel<3>;
    case 4: return gather_kernel<4>;
    default: return nullptr;
    }
}

GatherFn pick_cell_gather(int n_mat)
{
    switch (n_mat) {
    case 1: return cell_gather_kernel<1>;
    case 2: return cell_gather_kernel<2>;
    case 3: return cell_gather_kernel<3>;
    case 4: return cell_gather_kernel<4>;
    default: return nullptr;
    }
}

using CombineFn = void (*)(const float *, float *, int, int, int, int, int, int);
CombineFn pick_cell_combine(int n_mat)
{
    switch (n_mat) {
    case 1: return cell_combine_kernel<1>;
    case 2: return cell_combine_kernel<2>;
    case 3: return cell_combine_kernel<3>;
    case 4: return cell_combine_kernel<4>;
    default: return nullptr;
    }
}

template <int ST> AdjointFn pick_adjoint_st(int n_mat)
{
    switch (n_mat) {
    case 1: return adjoint_kernel<1, ST>;
    case 2: return adjoint_kernel<2, ST>;
    case 3: return adjoint_kernel<3, ST>;
    case 4: return adjoint_kernel<4, ST>;
    default: return nullptr;
    }
}

AdjointFn pick_adjoint(int n_mat, int dtype)
{
    switch (dtype) {
    case CTFP_F32: return pick_adjoint_st<kStF32>(n_mat);
    case CTFP_F16: return pick_adjoint_st<kStF16>(n_mat);
    case CTFP_BF16: return pick_adjoint_st<kStBF16>(n_mat);
    default: return nullptr;
    }
}

MarchFn pick_march(int n_mat, int dtype, bool spectrum, bool lds)
{
    switch (dtype) {
    case CTFP_F32: return pick_march_st<kStF32>(n_mat, spectrum, lds);
    case CTFP_F16: return pick_march_st<kStF16>(n_mat, spectrum, lds);
    case CTFP_BF16: return pick_march_st<kStBF16>(n_mat, spectrum, lds);
    default: return nullptr;
    }
}

int grid_for(int64_t n, int block = 256)
{
    int64_t g = (n + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

int check_ctx(const ctfp_context *ctx)
{
    if (!ctx) return set_error(CTFP_ERR_INVALID, "null context");
    return CTFP_OK;
}

// Device address of `scratch` on `device`, which must be current: looked up
// with hipGetSymbolAddress on the first call per device and kept.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};

int static_scratch(int device, float **out)
{
    if (device < 0 || device >= kMaxDevices)
        return set_error(CTFP_ERR_UNSUPPORTED, "device %d: the static workspace covers devices 0..%d", device,
                         kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(CTFP_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                             hipGetErrorString(e));
        g_scratch_ptr[device] = (float *)p;
    }
    *out = g_scratch_ptr[device];
    return CTFP_OK;
}

// Channel groups of a scan of n_mat materials: fours, then the rest.
void plan_groups(int n_mat, int &n_groups, int (&mats)[kMaxGroups], int (&moff)[kMaxGroups])
{
    n_groups = 0;
    for (int m = 0; m < n_mat; m += kGroupMats) {
        mats[n_groups] = std::min(kGroupMats, n_mat - m);
        moff[n_groups] = m;
        n_groups++;
    }
}

int check_n_mat(int n_mat)
{
    if (n_mat < 1 || n_mat > kMaxMaterials)
        return set_error(CTFP_ERR_UNSUPPORTED, "n_mat=%d: 1..%d materials are supported", n_mat, kMaxMaterials);
    return CTFP_OK;
}

// The context holds no volume (after an upload that failed part way: its
// layout, occupancy grid and material count no longer agree).
void drop_volume(ctfp_context *ctx)
{
    dfree(ctx->vol);
    dfree(ctx->occ);
    dfree(ctx->obox);
    dfree(ctx->odist);
    ctx->vol_bytes = 0;
    ctx->n_mat = 0;
    ctx->n_groups = 0;
}

static int alloc_volume_buffers(ctfp_context *ctx, size_t bytes, int64_t obx, int64_t oby, int64_t obz,
                                int n_groups)
{
    if (ctx->vol && (ctx->vol_bytes != bytes)) dfree(ctx->vol);
    if (!ctx->vol) HIP_TRY(dev_malloc(&ctx->vol, bytes));
    ctx->vol_bytes = bytes;
    if (ctx->occ && (ctx->ob_x != obx || ctx->ob_y != oby || ctx->ob_z != obz || ctx->n_groups != n_groups)) {
        dfree(ctx->occ);
        dfree(ctx->obox);
        dfree(ctx->odist);
    }
    if (!ctx->occ) HIP_TRY(dev_malloc(&ctx->occ, (size_t)(obx * oby * obz) * n_groups));
    if (!ctx->obox) HIP_TRY(dev_malloc(&ctx->obox, (size_t)n_groups * 6 * sizeof(int32_t)));
    if (!ctx->odist) HIP_TRY(dev_malloc(&ctx->odist, (size_t)(obx * oby * obz) * (n_groups + 1)));
    return CTFP_OK;
}

int alloc_volume(ctfp_context *ctx, int n_mat)
{
    if (check_n_mat(n_mat)) return CTFP_ERR_UNSUPPORTED;
    const ctfp_scan_params &s = ctx->scan;
    int64_t px = s.p_num_x + kVoxPadLo + kVoxPadHi;
    int64_t py = s.p_num_y + kVoxPadLo + kVoxPadHi;
    int64_t pz = s.p_num_z + kVoxPadLo + kVoxPadHi;
    int n_groups = 0, mats[kMaxGroups] = {}, moff[kMaxGroups] = {};
    plan_groups(n_mat, n_groups, mats, moff);
    size_t off[kMaxGroups] = {}, gbytes[kMaxGroups] = {};
    size_t bytes = 0;
    for (int g = 0; g < n_groups; g++) {
        if ((size_t)(px * py) * mats[g] * sizeof(float) >= (1u << 24))
            return set_error(CTFP_ERR_UNSUPPORTED, "one padded z slice exceeds 16 MiB");
        off[g] = bytes;
        gbytes[g] = (size_t)(px * py * pz) * (size_t)mats[g] * sizeof(float);
        bytes += (gbytes[g] + 255) / 256 * 256;
    }
    const int64_t obx = (px + kOccBrick - 1) / kOccBrick, oby = (py + kOccBrick - 1) / kOccBrick;
    const int64_t obz = (pz + kOccBrick - 1) / kOccBrick;
    // the march indexes the occupancy grid with 24-bit products
    if (obz * oby >= (1ll << 24) || obx >= (1ll << 24))
        return set_error(CTFP_ERR_UNSUPPORTED, "volume of %lld x %lld padded planes exceeds the occupancy index",
                         (long long)pz, (long long)py);
    if (int rc = alloc_volume_buffers(ctx, bytes, obx, oby, obz, n_groups)) {
        drop_volume(ctx);
        return rc;
    }
    ctx->ob_x = obx;
    ctx->ob_y = oby;
    ctx->ob_z = obz;
    ctx->n_mat = n_mat;
    ctx->n_groups = n_groups;
    for (int g = 0; g < n_groups; g++) {
        ctx->group_mats[g] = mats[g];
        ctx->group_moff[g] = moff[g];
        ctx->group_off[g] = off[g];
        ctx->group_bytes[g] = gbytes[g];
    }
    ctx->px = px;
    ctx->py = py;
    ctx->pz = pz;
    return CTFP_OK;
}

// Workgroups of one launch: the grid's work-items (blocks x 256) must stay
// below 2^32, and the march launches a multiple of 8 blocks.
constexpr uint64_t kMaxLaunchBlocks = ((1ull << 32) - 1) / kBlock / 8 * 8;
// the gathers' workgroups hold kGatherBlock or kCellBlock work-items
constexpr uint64_t kMaxGatherBricks = ((1ull << 32) - 1) / (kGatherBlock > kCellBlock ? kGatherBlock : kCellBlock);

// Views per march launch so that no launch exceeds kMaxLaunchBlocks (every
// view has the same tiles).
int64_t views_per_launch(const ctfp_context *ctx)
{
    const int64_t tile_pu = (int64_t)ctx->tune.wave_u * ctx->tune.block_u;
    const int64_t tile_pv = (int64_t)(64 / ctx->tune.wave_u) * (4 / ctx->tune.block_u);
    const int64_t tiles = ((ctx->scan.d_num_u + tile_pu - 1) / tile_pu) * ((ctx->scan.d_num_v + tile_pv - 1) / tile_pv);
    return std::max<int64_t>(1, (int64_t)kMaxLaunchBlocks / std::max<int64_t>(1, tiles));
}

int ilog2(int v)
{
    int r = 0;
    while ((1 << (r + 1)) <= v) r++;
    return r;
}

// Launch arguments of the march over views [view_begin, +view_count) of
// channel group `group` (its line integrals go to rows group_moff.. of L).
int fill_march_args(ctfp_context *ctx, int64_t view_begin, int64_t view_count, MarchArgs &P, int group = 0)
{
    const ctfp_scan_params &s = ctx->scan;
    const ctfp_system_info &y = ctx->sys;
    if (!ctx->vol) return set_error(CTFP_ERR_STATE, "no volumes uploaded");
    if (view_begin < 0 || view_count < 1 || view_begin + view_count > s.proj_num)
        return set_error(CTFP_ERR_INVALID, "view range [%lld, %lld) outside [0, %lld)",
                         (long long)view_begin, (long long)(view_begin + view_count),
                         (long long)s.proj_num);
    if (y.int_num < 1 || y.int_num > (1 << 30))
        return set_error(CTFP_ERR_INVALID, "samples per ray %lld out of range", (long long)y.int_num);
    if (group < 0 || group >= ctx->n_groups) return set_error(CTFP_ERR_INVALID, "material group %d", group);
    std::memset(&P, 0, sizeof P);
    const uint32_t VB = (uint32_t)(ctx->group_mats[group] * sizeof(float));
    P.vol = (const char *)ctx->vol + ctx->group_off[group];
    // volumes of 4 GiB and more read through per-box buffers and 64-bit
    // addresses; below that, one buffer covers the volume
    P.wide = ctx->group_bytes[group] >= 0xFFFFFFFFull ? 1 : 0;
    P.vol_bytes = P.wide ? 0xFFFFFFFFu : (uint32_t)ctx->group_bytes[group];
    P.l_mats = ctx->n_mat;
    P.l_moff = ctx->group_moff[group];
    P.occ = ctx->tune.skip_empty && ctx->occ ? ctx->occ + (size_t)group * ctx->ob_x * ctx->ob_y * ctx->ob_z : nullptr;
    P.obox = P.occ ? ctx->obox + 6 * group : nullptr;
    P.odist = P.occ ? ctx->odist + (size_t)group * ctx->ob_x * ctx->ob_y * ctx->ob_z : nullptr;
    P.ob_x = (int32_t)ctx->ob_x;
    P.ob_y = (int32_t)ctx->ob_y;
    P.pcx = (int32_t)ctx->px;
    P.pcy = (int32_t)ctx->py;
    P.pcz = (int32_t)ctx->pz;
    P.row_y = (uint32_t)(ctx->px * VB);
    P.row_z = (uint32_t)(ctx->px * ctx->py * VB);
    P.nx = (float)s.p_num_x;
    P.ny = (float)s.p_num_y;
    P.nz = (float)s.p_num_z;
    P.sod = s.sod;
    P.sdd = s.sdd;
    P.dsize = s.d_size;
    P.half_u = y.d_half_lu;
    P.half_v = y.d_half_lv;
    P.t0 = (double)s.sod - (double)y.fov_r + (double)y.dx / 2.0;
    P.dx = y.dx;
    P.half_x = y.p_half_x;
    P.half_y = y.p_half_y;
    P.half_z = y.p_half_z;
    P.inv_px = 1.0 / (double)y.p_size_x;
    P.inv_py = 1.0 / (double)y.p_size_y;
    P.inv_pz = 1.0 / (double)y.p_size_z;
    P.nu = (int32_t)s.d_num_u;
    P.nv = (int32_t)s.d_num_v;
    P.K = (int32_t)y.int_num;
    P.view_cs = ctx->view_cs;
    P.focal = ctx->focal;
    P.view_begin = view_begin;
    P.n_views = (int32_t)view_count;
    P.wu_log2 = ilog2(ctx->tune.wave_u);
    P.bu_log2 = ilog2(ctx->tune.block_u);
    const int64_t tile_pu = (int64_t)ctx->tune.wave_u * ctx->tune.block_u;
    const int64_t tile_pv = (int64_t)(64 / ctx->tune.wave_u) * (4 / ctx->tune.block_u);
    P.tiles_u = (int32_t)((s.d_num_u + tile_pu - 1) / tile_pu);
    P.tiles_v = (int32_t)((s.d_num_v + tile_pv - 1) / tile_pv);
    uint64_t nb = (uint64_t)view_count * P.tiles_u * P.tiles_v;
    if (nb > kMaxLaunchBlocks)
        return set_error(CTFP_ERR_INVALID, "too many tiles in one launch (%llu > %llu)", (unsigned long long)nb,
                         (unsigned long long)kMaxLaunchBlocks);
    P.n_blocks = (uint32_t)nb;
    P.per_xcd = (uint32_t)((nb + 7) / 8);
    P.dx_f = y.dx;
    return CTFP_OK;
}

} // namespace

// ------------------------------------------------------------- C ABI ---

extern "C" {


int ctfp_create(const ctfp_scan_params *scan, int device, ctfp_context **out)
{
    if (!scan || !out) return set_error(CTFP_ERR_INVALID, "null argument");
    *out = nullptr;
    ctfp_system_info sys;
    int rc = ctfp_compute_paras(scan, &sys);
    if (rc != CTFP_OK) return rc;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return set_error(CTFP_ERR_INVALID, "device %d not present (%d devices)", device, ndev);
    DeviceGuard g(device);
    ctfp_context *ctx = new (std::nothrow) ctfp_context();
    if (!ctx) return set_error(CTFP_ERR_NOMEM, "context allocation failed");
    ctx->device = device;
    ctx->scan = *scan;
    ctx->sys = sys;
    // Rotation of view i: the reference's FP32 angle (CBCTPolyProjGPU.cu:51),
    // cos/sin evaluated in FP64.
    std::vector<double2> cs((size_t)scan->proj_num);
    for (int64_t i = 0; i < scan->proj_num; i++) {
        double th = (double)ctfp::view_angle(scan, &sys, i);
        cs[(size_t)i] = make_double2(cos(th), sin(th));
    }
    hipError_t e = dev_malloc(&ctx->view_cs, cs.size() * sizeof(double2));
    if (e == hipSuccess)
        e = hipMemcpy(ctx->view_cs, cs.data(), cs.size() * sizeof(double2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ctfp_destroy(ctx);
        return set_error(hip_status(e), "view table: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return CTFP_OK;
}

int ctfp_destroy(ctfp_context *ctx)
{
    if (!ctx) return CTFP_OK;
    DeviceGuard g(ctx->device);
    // hipFree waits for the device's work before it releases memory; the
    // caller must not free a context that other threads still launch on
    dfree(ctx->vol);
    dfree(ctx->occ);
    dfree(ctx->obox);
    dfree(ctx->odist);
    dfree(ctx->d_flag);
    dfree(ctx->l_ws);
    dfree(ctx->adj_ws);
    dfree(ctx->cell_part);
    dfree(ctx->focal);
    dfree(ctx->Aq);
    dfree(ctx->A);
    dfree(ctx->W);
    dfree(ctx->I0);
    dfree(ctx->view_cs);
    dfree(ctx->dI);
    dfree(ctx->dIA);
    delete ctx;
    return CTFP_OK;
}

int ctfp_get_system_info(const ctfp_context *ctx, ctfp_system_info *out)
{
    if (check_ctx(ctx) || !out) return set_error(CTFP_ERR_INVALID, "null argument");
    *out = ctx->sys;
    return CTFP_OK;
}

// Writes the resident layout for `dtype` from n_mat FP32 device volumes
// (chan[m], [z][y][x]) on `stream`, one interleaved volume per channel group.
// Unless `async` it waits for the layout (on `stream` only) and rejects
// volumes holding a NaN or an infinity (the context then holds no volume).
static int write_layout_impl(ctfp_context *ctx, const float *const *chan, int n_mat, int dtype, hipStream_t stream,
                             bool async)
{
    int rc = alloc_volume(ctx, n_mat);
    if (rc) return rc;
    if (!async && !ctx->d_flag) HIP_TRY(dev_malloc(&ctx->d_flag, sizeof(int)));
    int *bad = async ? nullptr : ctx->d_flag;
    if (bad) HIP_TRY(hipMemsetAsync(bad, 0, sizeof(int), stream));
    const ctfp_scan_params &s = ctx->scan;
    const int64_t total = ctx->px * ctx->py * ctx->pz;
    for (int g = 0; g < ctx->n_groups; g++) {
        VolPtrs vp{};
        const int mg = ctx->group_mats[g], m0 = ctx->group_moff[g];
        for (int m = 0; m < mg; m++) vp.p[m] = chan[m0 + m];
        char *dst = (char *)ctx->vol + ctx->group_off[g];
        if (dtype == CTFP_F32)
            layout_voxels_f32<<<grid_for(total), 256, 0, stream>>>(vp, (float *)dst, mg, s.p_num_x, s.p_num_y,
                                                                   s.p_num_z, ctx->px, ctx->py, ctx->pz, bad);
        else if (dtype == CTFP_F16)
            layout_pairs_16<false><<<grid_for(total), 256, 0, stream>>>(vp, (uint32_t *)dst, mg, m0, s.p_num_x,
                                                                        s.p_num_y, s.p_num_z, ctx->px, ctx->py,
                                                                        ctx->pz, bad);
        else
            layout_pairs_16<true><<<grid_for(total), 256, 0, stream>>>(vp, (uint32_t *)dst, mg, m0, s.p_num_x,
                                                                       s.p_num_y, s.p_num_z, ctx->px, ctx->py,
                                                                       ctx->pz, bad);
        HIP_TRY(hipGetLastError());
        const int64_t obn = ctx->ob_x * ctx->ob_y * ctx->ob_z;
        occupancy_kernel<<<(unsigned)std::min<int64_t>(obn, kOccGrid), 256, 0, stream>>>(
            (const uint32_t *)dst, mg, ctx->px, ctx->py, ctx->pz, ctx->ob_x, ctx->ob_y, obn, ctx->occ + g * obn);
        HIP_TRY(hipGetLastError());
        if (g == 0) {
            occupied_box_init<<<1, 64, 0, stream>>>(ctx->obox, 6 * ctx->n_groups);
            HIP_TRY(hipGetLastError());
        }
        occupied_box_kernel<<<(unsigned)std::min<int64_t>(1024, (obn + 255) / 256), 256, 0, stream>>>(
            ctx->occ + g * obn, ctx->ob_x, ctx->ob_y, obn, ctx->obox + 6 * g);
        HIP_TRY(hipGetLastError());
        // brick distances: x from the occupancy into the group's grid, y into
        // the scratch grid, z back into the group's grid
        uint8_t *dg = ctx->odist + g * obn, *tmp = ctx->odist + ctx->n_groups * obn;
        const unsigned dgrid = (unsigned)std::min<int64_t>(kOccGrid, (obn + 255) / 256);
        brick_distance_kernel<<<dgrid, 256, 0, stream>>>(ctx->occ + g * obn, dg, ctx->ob_x, ctx->ob_y, ctx->ob_z, 0, 1);
        brick_distance_kernel<<<dgrid, 256, 0, stream>>>(dg, tmp, ctx->ob_x, ctx->ob_y, ctx->ob_z, 1, 0);
        brick_distance_kernel<<<dgrid, 256, 0, stream>>>(tmp, dg, ctx->ob_x, ctx->ob_y, ctx->ob_z, 2, 0);
        HIP_TRY(hipGetLastError());
    }
    ctx->dtype = dtype;
    ctx->generation++;
    if (!async) {
        int h = 0;
        HIP_TRY(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        if (h) return set_error(CTFP_ERR_INVALID, "volumes: a density is NaN or infinite");
    }
    return CTFP_OK;
}

// Any failure leaves the context with no volume (a forward then reports
// CTFP_ERR_STATE instead of reading a layout that does not match).
static int write_layout(ctfp_context *ctx, const float *const *chan, int n_mat, int dtype,
                        hipStream_t stream = nullptr, bool async = false)
{
    const int rc = write_layout_impl(ctx, chan, n_mat, dtype, stream, async);
    if (rc) drop_volume(ctx);
    return rc;
}

static int check_dtype(int dtype)
{
    if (dtype != CTFP_F32 && dtype != CTFP_F16 && dtype != CTFP_BF16)
        return set_error(CTFP_ERR_INVALID, "storage dtype %d is not CTFP_F32/F16/BF16", dtype);
    return CTFP_OK;
}

int ctfp_set_volumes(ctfp_context *ctx, int n_mat, const float *const *volumes,
                     int volumes_on_device, int storage_dtype)
{
    Range range_("ctfp_set_volumes");
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    if (!volumes) return set_error(CTFP_ERR_INVALID, "null volume list");
    if (check_dtype(storage_dtype)) return CTFP_ERR_INVALID;
    if (check_n_mat(n_mat)) return CTFP_ERR_UNSUPPORTED;
    for (int m = 0; m < n_mat; m++)
        if (!volumes[m]) return set_error(CTFP_ERR_INVALID, "volume %d is null", m);
    DeviceGuard g(ctx->device);
    const ctfp_scan_params &s = ctx->scan;
    const int64_t nvox = s.p_num_x * s.p_num_y * s.p_num_z;
    const float *chan[kMaxMaterials] = {};
    float *stage = nullptr;
    if (volumes_on_device) {
        for (int m = 0; m < n_mat; m++) chan[m] = volumes[m];
    } else {
        HIP_TRY(dev_malloc(&stage, (size_t)nvox * n_mat * sizeof(float)));
        for (int m = 0; m < n_mat; m++) {
            hipError_t e = hipMemcpy(stage + m * nvox, volumes[m], (size_t)nvox * sizeof(float),
                                     hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                (void)hipFree(stage);
                return set_error(hip_status(e), "volume upload: %s", hipGetErrorString(e));
            }
            chan[m] = stage + m * nvox;
        }
    }
    int rc = write_layout(ctx, chan, n_mat, storage_dtype);
    if (stage) (void)hipFree(stage);
    return rc;
}

int ctfp_set_volumes_async(ctfp_context *ctx, int n_mat, const float *const *volumes,
                           int storage_dtype, void *stream)
{
    Range range_("ctfp_set_volumes_async");
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    if (!volumes) return set_error(CTFP_ERR_INVALID, "null volume list");
    if (check_dtype(storage_dtype)) return CTFP_ERR_INVALID;
    if (check_n_mat(n_mat)) return CTFP_ERR_UNSUPPORTED;
    for (int m = 0; m < n_mat; m++)
        if (!volumes[m]) return set_error(CTFP_ERR_INVALID, "volume %d is null", m);
    DeviceGuard g(ctx->device);
    return write_layout(ctx, volumes, n_mat, storage_dtype, (hipStream_t)stream, true);
}

int ctfp_set_focal_spot(ctfp_context *ctx, const float *offset_u, const float *offset_v)
{
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    if ((offset_u == nullptr) != (offset_v == nullptr))
        return set_error(CTFP_ERR_INVALID, "focal spot: give both offset arrays or neither");
    DeviceGuard g(ctx->device);
    ctx->generation++;
    if (!offset_u) {
        dfree(ctx->focal);
        return CTFP_OK;
    }
    const int64_t n = ctx->scan.proj_num;
    std::vector<double2> f((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        if (!std::isfinite(offset_u[i]) || !std::isfinite(offset_v[i]))
            return set_error(CTFP_ERR_INVALID, "focal spot: offset of view %lld is not finite", (long long)i);
        f[(size_t)i] = make_double2((double)offset_u[i], (double)offset_v[i]);
    }
    if (!ctx->focal) HIP_TRY(dev_malloc(&ctx->focal, (size_t)n * sizeof(double2)));
    HIP_TRY(hipMemcpy(ctx->focal, f.data(), (size_t)n * sizeof(double2), hipMemcpyHostToDevice));
    return CTFP_OK;
}

int ctfp_get_scan(const ctfp_context *ctx, ctfp_scan_params *out)
{
    if (check_ctx(ctx) || !out) return set_error(CTFP_ERR_INVALID, "null argument");
    *out = ctx->scan;
    return CTFP_OK;
}

int ctfp_device_bytes(const ctfp_context *ctx, int64_t *bytes)
{
    if (check_ctx(ctx) || !bytes) return set_error(CTFP_ERR_INVALID, "null argument");
    const int64_t npix = ctx->scan.d_num_u * ctx->scan.d_num_v;
    int64_t b = ctx->vol ? (int64_t)ctx->vol_bytes : 0;
    if (ctx->A) b += (int64_t)ctx->n_act * ctx->spec_n_mat * (int64_t)sizeof(float);
    if (ctx->W) b += (int64_t)(std::max(ctx->n_act, 1) + kWeightPadBins) * ctx->w_stride_e * (int64_t)sizeof(float);
    if (ctx->I0) b += npix * (int64_t)sizeof(float);
    if (ctx->view_cs) b += ctx->scan.proj_num * (int64_t)sizeof(double2);
    if (ctx->dI) b += 2 * ctx->staged_views * npix * (int64_t)sizeof(float);
    if (ctx->adj_ws) b += (int64_t)ctx->adj_ws_bytes;
    if (ctx->cell_part) b += (int64_t)ctx->cell_part_bytes;
    if (ctx->l_ws) b += (int64_t)ctx->l_ws_bytes;
    if (ctx->occ) b += ctx->ob_x * ctx->ob_y * ctx->ob_z * ctx->n_groups;
    *bytes = b;
    return CTFP_OK;
}

int ctfp_get_device(const ctfp_context *ctx, int *device)
{
    if (check_ctx(ctx) || !device) return set_error(CTFP_ERR_INVALID, "null argument");
    *device = ctx->device;
    return CTFP_OK;
}

int ctfp_get_volumes(const ctfp_context *ctx, int *n_mat, int *storage_dtype)
{
    if (check_ctx(ctx) || !n_mat || !storage_dtype) return set_error(CTFP_ERR_INVALID, "null argument");
    if (!ctx->vol) return set_error(CTFP_ERR_STATE, "no volumes uploaded");
    *n_mat = ctx->n_mat;
    *storage_dtype = ctx->dtype;
    return CTFP_OK;
}

int ctfp_set_labels(ctfp_context *ctx, const uint8_t *labels, int labels_on_device, int n_mat,
                    const float *density, int single_material, int storage_dtype)
{
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    if (!labels || !density) return set_error(CTFP_ERR_INVALID, "null labels or densities");
    if (check_dtype(storage_dtype)) return CTFP_ERR_INVALID;
    if (check_n_mat(n_mat)) return CTFP_ERR_UNSUPPORTED;
    DeviceGuard g(ctx->device);
    const int channels = single_material ? 1 : n_mat;
    const ctfp_scan_params &s = ctx->scan;
    const int64_t nvox = s.p_num_x * s.p_num_y * s.p_num_z;
    uint8_t *stage = nullptr;
    float *raw = nullptr;
    const uint8_t *dl = labels;
    hipError_t e = hipSuccess;
    if (!labels_on_device) {
        e = dev_malloc(&stage, (size_t)nvox);
        if (e == hipSuccess) e = hipMemcpy(stage, labels, (size_t)nvox, hipMemcpyHostToDevice);
        dl = stage;
    }
    if (e == hipSuccess) e = dev_malloc(&raw, (size_t)nvox * channels * sizeof(float));
    if (e == hipSuccess) {
        Densities rho{};
        for (int m = 0; m < channels; m++) rho.d[m] = density[m];
        labels_to_raw<<<grid_for(nvox), 256>>>(dl, raw, channels, single_material, rho, nvox);
        e = hipGetLastError();
    }
    int rc = CTFP_OK;
    if (e == hipSuccess) {
        const float *chan[kMaxMaterials] = {};
        for (int m = 0; m < channels; m++) chan[m] = raw + m * nvox;
        rc = write_layout(ctx, chan, channels, storage_dtype);
    }
    if (stage) (void)hipFree(stage);
    if (raw) (void)hipFree(raw);
    if (e != hipSuccess)
        return set_error(hip_status(e), "label channels: %s",
                         hipGetErrorString(e));
    return rc;
}

static int set_spectral_impl(ctfp_context *ctx, const ctfp_spectral *sp)
{
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    if (!sp || !sp->spectrum || !sp->atten)
        return set_error(CTFP_ERR_INVALID, "spectral: spectrum and atten are required");
    if (sp->n_energy < 1 || sp->n_energy > 4096)
        return set_error(CTFP_ERR_INVALID, "spectral: n_energy=%d", sp->n_energy);
    if (sp->n_mat < 1 || sp->n_mat > kMaxMaterials)
        return set_error(CTFP_ERR_UNSUPPORTED, "spectral: n_mat=%d (1..%d)", sp->n_mat, kMaxMaterials);
    if (ctx->vol && sp->n_mat != ctx->n_mat)
        return set_error(CTFP_ERR_INVALID, "spectral: n_mat=%d but volumes have %d channels",
                         sp->n_mat, ctx->n_mat);
    if (sp->scint_mu && !sp->det && !sp->scint_thickness)
        return set_error(CTFP_ERR_INVALID, "spectral: scint_mu needs scint_thickness");
    DeviceGuard g(ctx->device);
    const ctfp_scan_params &s = ctx->scan;
    const int64_t nu = s.d_num_u, nv = s.d_num_v, npix = nu * nv;
    // every table entry must be a finite number (a NaN or an infinity would
    // reach every ray through the energy sums)
    {
        const int64_t E = sp->n_energy;
        struct Table {
            const char *name;
            const float *p;
            int64_t n;
        } tables[] = {{"spectrum", sp->spectrum, E},
                      {"atten", sp->atten, E * sp->n_mat},
                      {"grid", sp->grid, E * nu},
                      {"det", sp->det, E * npix},
                      {"scint_mu", sp->det ? nullptr : sp->scint_mu, E},
                      {"scint_thickness", sp->det ? nullptr : (sp->scint_mu ? sp->scint_thickness : nullptr), npix}};
        for (const Table &t : tables) {
            if (!t.p) continue;
            for (int64_t i = 0; i < t.n; i++)
                if (!std::isfinite(t.p[i]))
                    return set_error(CTFP_ERR_INVALID, "spectral: %s[%lld] is not finite", t.name, (long long)i);
        }
        if (!std::isfinite(sp->i0_val) || !std::isfinite(sp->det_factor))
            return set_error(CTFP_ERR_INVALID, "spectral: i0_val or det_factor is not finite");
    }
    // Active bins: a zero spectrum entry contributes exactly zero.
    std::vector<int> act;
    for (int e = 0; e < sp->n_energy; e++)
        if (sp->spectrum[e] != 0.0f) act.push_back(e);
    const int na = (int)act.size();
    std::vector<float> S(std::max(na, 1)), A((size_t)std::max(na, 1) * sp->n_mat);
    for (int i = 0; i < na; i++) {
        int e = act[i];
        S[i] = sp->spectrum[e] * (sp->energy_weighting ? (float)(e + 1) : 1.0f);
        for (int m = 0; m < sp->n_mat; m++) A[(size_t)i * sp->n_mat + m] = sp->atten[(size_t)e * sp->n_mat + m];
    }
    // A response that does not vary along v (a scintillator of one thickness
    // per detector column, or a response table whose rows repeat) keeps one
    // row of weights per bin: the same values, read from a table that stays
    // in the caches instead of one row per pixel row.
    bool v_uniform = true;
    if (sp->det) {
        for (int i = 0; i < na && v_uniform; i++) {
            const float *d = sp->det + (size_t)act[i] * npix;
            for (int64_t q = nu; q < npix && v_uniform; q++) v_uniform = d[q] == d[q % nu];
        }
    } else if (sp->scint_mu) {
        for (int64_t q = nu; q < npix && v_uniform; q++)
            v_uniform = sp->scint_thickness[q] == sp->scint_thickness[q % nu];
    }
    const bool per_pixel = (sp->det != nullptr || sp->scint_mu != nullptr) && !v_uniform;
    const int64_t rows = per_pixel ? nv : 1;
    ctx->generation++; // the tables are replaced from here on
    dfree(ctx->A);
    dfree(ctx->Aq);
    dfree(ctx->W);
    if (!ctx->I0) HIP_TRY(dev_malloc(&ctx->I0, (size_t)npix * sizeof(float)));
    HIP_TRY(dev_malloc(&ctx->A, A.size() * sizeof(float)));
    HIP_TRY(hipMemcpy(ctx->A, A.data(), A.size() * sizeof(float), hipMemcpyHostToDevice));
    {
        // MFMA operand: row e, k = 3m + (0, 1, 2) = (hi, hi, lo) of 2^s A[e][m],
        // 2^s the largest power of two with 2^s max|A| <= 2^15 (FP16 holds
        // 2^s A without overflow; the split's lo part of the largest entries
        // stays normal)
        double amax = 0.0;
        for (float a : A) amax = std::max(amax, (double)std::fabs(a));
        const int sc = amax > 0.0 ? std::min(40, std::max(-40, (int)std::floor(std::log2(32768.0 / amax)))) : 0;
        const float scale = std::ldexp(1.0f, sc);
        ctx->aq_inv_scale = std::ldexp(1.0f, -sc);
        const int rows = std::max(32, (na + 31) / 32 * 32);
        std::vector<_Float16> Aq((size_t)rows * 16, (_Float16)0.0f);
        if (3 * sp->n_mat <= 16) {
            for (int i = 0; i < na; i++) {
                for (int m = 0; m < sp->n_mat; m++) {
                    const float a = A[(size_t)i * sp->n_mat + m] * scale;
                    const _Float16 hi = (_Float16)a;
                    Aq[(size_t)i * 16 + 3 * m] = hi;
                    Aq[(size_t)i * 16 + 3 * m + 1] = hi;
                    Aq[(size_t)i * 16 + 3 * m + 2] = (_Float16)(a - (float)hi);
                }
            }
        }
        HIP_TRY(dev_malloc(&ctx->Aq, Aq.size() * sizeof(_Float16)));
        HIP_TRY(hipMemcpy(ctx->Aq, Aq.data(), Aq.size() * sizeof(_Float16), hipMemcpyHostToDevice));
    }
    // kWeightPadBins zero rows follow the last active bin (the matrix-core
    // spectrum reads bins in groups that may run past it)
    HIP_TRY(dev_malloc(&ctx->W, (size_t)(std::max(na, 1) + kWeightPadBins) * rows * nu * sizeof(float)));
    HIP_TRY(hipMemset(ctx->W + (size_t)std::max(na, 1) * rows * nu, 0,
                      (size_t)kWeightPadBins * rows * nu * sizeof(float)));
    // stage the per-bin tables of the active bins
    std::vector<float> T, eta, mu;
    float *dS = nullptr, *dT = nullptr, *dEta = nullptr, *dMu = nullptr, *dThick = nullptr;
    auto cleanup = [&]() {
        dfree(dS); dfree(dT); dfree(dEta); dfree(dMu); dfree(dThick);
    };
    hipError_t e = dev_malloc(&dS, S.size() * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(dS, S.data(), S.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess && sp->grid && na > 0) {
        T.resize((size_t)na * nu);
        for (int i = 0; i < na; i++)
            std::memcpy(&T[(size_t)i * nu], sp->grid + (size_t)act[i] * nu, nu * sizeof(float));
        e = dev_malloc(&dT, T.size() * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(dT, T.data(), T.size() * sizeof(float), hipMemcpyHostToDevice);
    }
    // rows * nu response entries per bin: every pixel, or row 0 when the
    // response does not vary along v
    const int64_t per = rows * nu;
    if (e == hipSuccess && sp->det && na > 0) {
        e = dev_malloc(&dEta, (size_t)na * per * sizeof(float));
        for (int i = 0; i < na && e == hipSuccess; i++)
            e = hipMemcpy(dEta + (size_t)i * per, sp->det + (size_t)act[i] * npix, (size_t)per * sizeof(float),
                          hipMemcpyHostToDevice);
    } else if (e == hipSuccess && sp->scint_mu && na > 0) {
        mu.resize(na);
        for (int i = 0; i < na; i++) mu[i] = sp->scint_mu[act[i]];
        e = dev_malloc(&dMu, mu.size() * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(dMu, mu.data(), mu.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = dev_malloc(&dThick, (size_t)per * sizeof(float));
        if (e == hipSuccess)
            e = hipMemcpy(dThick, sp->scint_thickness, (size_t)per * sizeof(float), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && na > 0) {
        WeightArgs wa{};
        wa.S = dS; wa.T = dT; wa.eta = dEta; wa.mu_c = dMu; wa.thick = dThick;
        wa.factor = sp->det_factor;
        wa.i0 = sp->i0_val;
        wa.n_act = na;
        wa.nu = nu;
        wa.rows = rows;
        build_weights<<<grid_for((int64_t)na * rows * nu), 256>>>(wa, ctx->W);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        ctx->n_act = na;
        ctx->spec_n_mat = sp->n_mat;
        ctx->w_stride_e = rows * nu;
        ctx->w_stride_v = per_pixel ? (int32_t)nu : 0;
        flat_field_kernel<<<grid_for(npix), 256>>>(ctx->W, na, ctx->w_stride_e, ctx->w_stride_v, nu,
                                                   nv, ctx->I0);
        e = hipGetLastError();
    }
    // the tables are built on the null stream; wait for that stream only
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    cleanup();
    if (e != hipSuccess) return set_error(hip_status(e), "spectral set-up: %s", hipGetErrorString(e));
    return CTFP_OK;
}

// A failure after the old tables were released leaves the context with no
// spectral tables (forwards then report CTFP_ERR_STATE), not with tables that
// disagree with its bin count; invalid arguments leave the old tables.
int ctfp_set_spectral(ctfp_context *ctx, const ctfp_spectral *sp)
{
    Range range_("ctfp_set_spectral");
    const uint64_t gen = ctx ? ctx->generation : 0;
    const int rc = set_spectral_impl(ctx, sp);
    if (rc != CTFP_OK && ctx && ctx->generation != gen) {
        DeviceGuard g(ctx->device);
        dfree(ctx->A);
        dfree(ctx->Aq);
        dfree(ctx->W);
        dfree(ctx->I0);
        ctx->n_act = 0;
        ctx->spec_n_mat = 0;
    }
    return rc;
}

int ctfp_debug_fail_allocations(long long after)
{
    g_alloc_budget.store(after < 0 ? -1 : after);
    return CTFP_OK;
}

// The spectral tables of the context, checked against its volumes.
static int spectral_args(ctfp_context *ctx, MarchArgs &P)
{
    if (!ctx->W || !ctx->A) return set_error(CTFP_ERR_STATE, "spectral weights not set");
    if (ctx->spec_n_mat != ctx->n_mat)
        return set_error(CTFP_ERR_STATE, "spectral tables are for %d materials, volumes have %d", ctx->spec_n_mat,
                         ctx->n_mat);
    // the matrix-core spectrum of 16-bit volumes addresses 32 weight rows
    // from one buffer descriptor with 32-bit offsets
    if (ctx->dtype != CTFP_F32 && ctx->w_stride_e * 4 * 32 >= (int64_t)(1ll << 32))
        return set_error(CTFP_ERR_UNSUPPORTED, "16-bit volumes: a weight row of %lld entries exceeds the %lld the "
                         "matrix-core spectrum addresses", (long long)ctx->w_stride_e, (long long)((1ll << 32) / 128));
    P.n_act = ctx->n_act;
    P.A = ctx->A;
    P.Aq = ctx->Aq;
    P.aq_inv_scale = ctx->aq_inv_scale;
    P.W = ctx->W;
    P.w_stride_e = ctx->w_stride_e;
    P.w_stride_v = ctx->w_stride_v;
    return CTFP_OK;
}

// Launch arguments of the march of channel group `group` over views
// [view_begin, view_begin + count): with the spectrum (I, IA; one group
// only) and / or the line integrals into L [count][n_mat][v][u].
static int forward_args(ctfp_context *ctx, int64_t view_begin, int64_t view_count, float *I,
                        float *IA, float *L, bool spectrum, MarchArgs &P, int group = 0)
{
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    int rc = fill_march_args(ctx, view_begin, view_count, P, group);
    if (rc) return rc;
    if (spectrum) {
        if ((rc = spectral_args(ctx, P))) return rc;
        if (!I || !IA) return set_error(CTFP_ERR_INVALID, "null output");
    } else if (!L) {
        return set_error(CTFP_ERR_INVALID, "null output");
    }
    P.I = I;
    P.IA = IA;
    P.L = L;
    const int mg = ctx->group_mats[group];
    const uint32_t lb = mg == 3 ? 16u : 4u * (uint32_t)mg;
    P.chunk = ctx->tune.chunk;
    P.lds_cells = (int32_t)((ctx->tune.lds_bytes - 256) / lb);
    return CTFP_OK;
}

// The march of one channel group, in launches of at most views_per_launch
// views.  I / IA / L point at the first view's rows.
static int launch_group(ctfp_context *ctx, int64_t view_begin, int64_t view_count, float *I, float *IA, float *L,
                        void *stream, bool spectrum, int group)
{
    const bool lds = ctx->tune.march == CTFP_MARCH_LDS;
    if (!ctx->vol) return set_error(CTFP_ERR_STATE, "no volumes uploaded");
    MarchFn fn = pick_march(ctx->group_mats[group], ctx->dtype, spectrum, lds);
    if (!fn) return set_error(CTFP_ERR_UNSUPPORTED, "n_mat=%d", ctx->group_mats[group]);
    const size_t shmem = lds ? (size_t)ctx->tune.lds_bytes : 0;
    const int64_t npix = ctx->scan.d_num_u * ctx->scan.d_num_v;
    const int64_t per = views_per_launch(ctx);
    DeviceGuard g(ctx->device);
    for (int64_t v0 = 0; v0 < std::max<int64_t>(view_count, 1); v0 += per) {
        const int64_t n = std::min(per, view_count - v0);
        MarchArgs P;
        int rc = forward_args(ctx, view_begin + v0, n, I ? I + v0 * npix : nullptr, IA ? IA + v0 * npix : nullptr,
                              L ? L + v0 * ctx->n_mat * npix : nullptr, spectrum, P, group);
        if (rc) return rc;
        hipLaunchKernelGGL(fn, dim3(P.per_xcd * 8), dim3(kBlock), shmem, (hipStream_t)stream, P);
        HIP_TRY(hipGetLastError());
    }
    return CTFP_OK;
}

// Line-integral workspace of the context for `views` views (grouped scans).
static int ensure_l_ws(ctfp_context *ctx, int64_t views)
{
    const size_t need = (size_t)views * ctx->n_mat * ctx->scan.d_num_u * ctx->scan.d_num_v * sizeof(float);
    if (ctx->l_ws_bytes >= need) return CTFP_OK;
    dfree(ctx->l_ws);
    ctx->l_ws_bytes = 0;
    HIP_TRY(dev_malloc(&ctx->l_ws, need));
    ctx->l_ws_bytes = need;
    return CTFP_OK;
}

// Views of a grouped scan per pass through a line-integral buffer of `bytes`.
static int64_t l_views(const ctfp_context *ctx, size_t bytes, int64_t view_count)
{
    const int64_t per_view = (int64_t)ctx->n_mat * ctx->scan.d_num_u * ctx->scan.d_num_v * (int64_t)sizeof(float);
    return std::min<int64_t>({view_count, (int64_t)bytes / per_view, views_per_launch(ctx)});
}

constexpr size_t kLWorkspaceBytes = 512ull << 20;
// The gather's dl/dL workspace per context: 1 GiB holds 300 views of a
// 3-material 512 x 512 detector (3 MiB each), so a 180-view scan of that
// size runs as one group (every brick then walks all views in one pass).
constexpr size_t kAdjWorkspaceBytes = 1ull << 30;

static int grouped_forward(ctfp_context *ctx, int64_t view_begin, int64_t view_count, float *I, float *IA, float *L,
                           void *stream, float *lws, size_t lws_bytes);

// ctfp_forward (spectrum) and ctfp_line_integrals.  Up to four materials:
// one march per launch with the spectrum fused.  More: every group marched
// into L (the caller's, or the context's workspace in view chunks), then
// spectrum_kernel from L.
static int launch_march(ctfp_context *ctx, int64_t view_begin, int64_t view_count, float *I,
                        float *IA, float *L, void *stream, bool spectrum)
{
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    if (!ctx->vol) return set_error(CTFP_ERR_STATE, "no volumes uploaded");
    if (view_begin < 0 || view_count < 1 || view_begin + view_count > ctx->scan.proj_num)
        return set_error(CTFP_ERR_INVALID, "view range [%lld, %lld) outside [0, %lld)", (long long)view_begin,
                         (long long)(view_begin + view_count), (long long)ctx->scan.proj_num);
    if (ctx->n_groups <= 1) return launch_group(ctx, view_begin, view_count, I, IA, L, stream, spectrum, 0);
    int rc;
    if (!spectrum) {
        for (int g = 0; g < ctx->n_groups; g++)
            if ((rc = launch_group(ctx, view_begin, view_count, nullptr, nullptr, L, stream, false, g))) return rc;
        return CTFP_OK;
    }
    if (!L) {
        const int64_t chunk = l_views(ctx, kLWorkspaceBytes, view_count);
        if (chunk < 1) return set_error(CTFP_ERR_UNSUPPORTED, "line integrals of one view exceed the workspace");
        if ((rc = ensure_l_ws(ctx, chunk))) return rc;
    }
    return grouped_forward(ctx, view_begin, view_count, I, IA, L, stream, ctx->l_ws, ctx->l_ws_bytes);
}

// The spectrum of a scan of more than four materials: per chunk of views,
// every group's line integrals into L (the caller's, else `lws` of
// lws_bytes), then spectrum_kernel.  Allocates nothing.
static int grouped_forward(ctfp_context *ctx, int64_t view_begin, int64_t view_count, float *I, float *IA, float *L,
                           void *stream, float *lws, size_t lws_bytes)
{
    if (!I || !IA) return set_error(CTFP_ERR_INVALID, "null output");
    int rc;
    MarchArgs S;
    if ((rc = fill_march_args(ctx, view_begin, std::min(view_count, views_per_launch(ctx)), S))) return rc;
    if ((rc = spectral_args(ctx, S))) return rc;
    const int64_t npix = ctx->scan.d_num_u * ctx->scan.d_num_v;
    const int64_t chunk = L ? std::min<int64_t>(view_count, views_per_launch(ctx)) : l_views(ctx, lws_bytes, view_count);
    if (chunk < 1 || (!L && !lws))
        return set_error(CTFP_ERR_UNSUPPORTED, "line integrals of one view exceed the workspace");
    DeviceGuard dg(ctx->device);
    for (int64_t v0 = 0; v0 < view_count; v0 += chunk) {
        const int64_t n = std::min(chunk, view_count - v0);
        float *Lc = L ? L + v0 * ctx->n_mat * npix : lws;
        for (int g = 0; g < ctx->n_groups; g++)
            if ((rc = launch_group(ctx, view_begin + v0, n, nullptr, nullptr, Lc, stream, false, g))) return rc;
        MarchArgs P = S;
        P.view_begin = view_begin + v0;
        P.n_views = (int32_t)n;
        P.I = I + v0 * npix;
        P.IA = IA + v0 * npix;
        const int64_t blocks = (n * npix + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(spectrum_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, (hipStream_t)stream, P,
                           (const float *)Lc, ctx->n_mat);
        HIP_TRY(hipGetLastError());
    }
    return CTFP_OK;
}

int ctfp_set_tuning(ctfp_context *ctx, const ctfp_tuning *t)
{
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    if (!t) return set_error(CTFP_ERR_INVALID, "null tuning");
    if (t->march != CTFP_MARCH_GLOBAL && t->march != CTFP_MARCH_LDS)
        return set_error(CTFP_ERR_INVALID, "tuning: march mode %d", t->march);
    if (t->chunk < kAnchor || t->chunk > 4096 || t->chunk % kAnchor != 0)
        return set_error(CTFP_ERR_INVALID, "tuning: chunk %d is not a multiple of %d in [%d, 4096]",
                         t->chunk, kAnchor, kAnchor);
    if (t->lds_bytes < 1024 || t->lds_bytes > 64 * 1024)
        return set_error(CTFP_ERR_INVALID, "tuning: lds_bytes %d outside [1 KiB, 64 KiB]", t->lds_bytes);
    if (t->wave_u != 4 && t->wave_u != 8 && t->wave_u != 16 && t->wave_u != 32)
        return set_error(CTFP_ERR_INVALID, "tuning: wave_u %d is not 4, 8, 16 or 32", t->wave_u);
    if (t->block_u != 1 && t->block_u != 2 && t->block_u != 4)
        return set_error(CTFP_ERR_INVALID, "tuning: block_u %d is not 1, 2 or 4", t->block_u);
    if (t->adjoint != CTFP_ADJOINT_GATHER && t->adjoint != CTFP_ADJOINT_SCATTER && t->adjoint != CTFP_ADJOINT_CELL)
        return set_error(CTFP_ERR_INVALID, "tuning: adjoint mode %d", t->adjoint);
    if (t->skip_empty != 0 && t->skip_empty != 1)
        return set_error(CTFP_ERR_INVALID, "tuning: skip_empty %d is not 0 or 1", t->skip_empty);
    ctx->tune = *t;
    return CTFP_OK;
}

int ctfp_get_tuning(const ctfp_context *ctx, ctfp_tuning *t)
{
    if (check_ctx(ctx) || !t) return set_error(CTFP_ERR_INVALID, "null argument");
    *t = ctx->tune;
    return CTFP_OK;
}

int ctfp_forward(ctfp_context *ctx, int64_t view_begin, int64_t view_count, float *I,
                 float *IAbsorb, float *L, void *stream)
{
    Range range_("ctfp_forward");
    return launch_march(ctx, view_begin, view_count, I, IAbsorb, L, stream, true);
}

// The adjoint of both backward entry points.  ws == nullptr: the gather's
// dl/dL workspace (and, for more than four materials without L, the line
// integrals) are the context's, allocated on first use for up to
// kAdjWorkspaceBytes of views per group; otherwise `ws` of ws_bytes holds them and nothing is
// allocated.
static int backward_impl(ctfp_context *ctx, int64_t view_begin, int64_t view_count, const float *gI,
                         const float *gIAbsorb, const float *L, float *grad, void *stream, float *ws,
                         size_t ws_bytes)
{
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    MarchArgs P;
    int rc = fill_march_args(ctx, view_begin, view_count, P);
    if (rc) return rc;
    if ((rc = spectral_args(ctx, P))) return rc;
    if (!grad) return set_error(CTFP_ERR_INVALID, "null gradient output");
    // a caller's workspace (the call allocates nothing): the voxel gather,
    // which needs no buffer of corner sums
    const bool cell = ctx->tune.adjoint == CTFP_ADJOINT_CELL && !ws;
    const bool grouped = ctx->n_groups > 1;
    // CTFP_MARCH_GLOBAL: no chunk is staged; every sample gathers and
    // scatters through global memory (the LDS then only holds the box scratch)
    const bool global = ctx->tune.march == CTFP_MARCH_GLOBAL;
    const size_t shmem = global ? 256 : (size_t)ctx->tune.lds_bytes;
    const ctfp_scan_params &s = ctx->scan;
    const int64_t npix = s.d_num_u * s.d_num_v;
    DeviceGuard g(ctx->device);
    if (ctx->tune.adjoint == CTFP_ADJOINT_SCATTER) {
        if (grouped)
            return set_error(CTFP_ERR_UNSUPPORTED, "the scatter adjoint takes up to %d materials (%d given); use "
                             "the gather", kGroupMats, ctx->n_mat);
        AdjointFn fn = pick_adjoint(ctx->n_mat, ctx->dtype);
        if (!fn) return set_error(CTFP_ERR_UNSUPPORTED, "n_mat=%d", ctx->n_mat);
        const int64_t per = views_per_launch(ctx);
        for (int64_t v0 = 0; v0 < view_count; v0 += per) {
            const int64_t n = std::min(per, view_count - v0);
            MarchArgs Q;
            if ((rc = fill_march_args(ctx, view_begin + v0, n, Q))) return rc;
            if ((rc = spectral_args(ctx, Q))) return rc;
            const uint32_t lb = ctx->n_mat == 3 ? 16u : 4u * (uint32_t)ctx->n_mat;
            Q.chunk = ctx->tune.chunk;
            Q.lds_cells = global ? 0 : (int32_t)((ctx->tune.lds_bytes - 256) / lb);
            AdjointArgs G{gI ? gI + v0 * npix : nullptr, gIAbsorb ? gIAbsorb + v0 * npix : nullptr,
                          L ? L + v0 * ctx->n_mat * npix : nullptr, grad};
            hipLaunchKernelGGL(fn, dim3(Q.per_xcd * 8), dim3(kBlock), shmem, (hipStream_t)stream, Q, G);
            HIP_TRY(hipGetLastError());
        }
        return CTFP_OK;
    }
    // gather: per group of views, the ray pass writes dl/dL * dx into the
    // workspace and the voxel pass adds its backprojection to grad
    RayGradFn rfn = grouped ? nullptr : pick_ray_grad(ctx->n_mat, ctx->dtype);
    if (!grouped && !rfn) return set_error(CTFP_ERR_UNSUPPORTED, "n_mat=%d", ctx->n_mat);
    const int64_t per_view = (int64_t)ctx->n_mat * npix * (int64_t)sizeof(float);
    // more than four materials without L: the line integrals of the group
    // wait in a second buffer of the same size
    const bool need_l = grouped && !L;
    const int64_t per_view_ws = per_view * (need_l ? 2 : 1);
    int64_t group;
    float *lbuf = nullptr;
    if (ws) {
        group = std::min<int64_t>({view_count, (int64_t)ws_bytes / per_view_ws, views_per_launch(ctx)});
        if (group < 1)
            return set_error(CTFP_ERR_UNSUPPORTED,
                             "dl/dL of one view (%lld bytes) exceeds the %lld-byte static workspace",
                             (long long)per_view_ws, (long long)ws_bytes);
        if (need_l) lbuf = ws + group * (per_view / (int64_t)sizeof(float));
    } else {
        group = std::max<int64_t>(
            1, std::min<int64_t>({view_count, (int64_t)kAdjWorkspaceBytes / per_view, views_per_launch(ctx)}));
        const size_t need = (size_t)(group * per_view);
        if (ctx->adj_ws_bytes < need) {
            dfree(ctx->adj_ws);
            ctx->adj_ws_bytes = 0;
            HIP_TRY(dev_malloc(&ctx->adj_ws, need));
            ctx->adj_ws_bytes = need;
        }
        ws = ctx->adj_ws;
        if (need_l) {
            if ((rc = ensure_l_ws(ctx, group))) return rc;
            lbuf = ctx->l_ws;
        }
    }
    GatherArgs A;
    A.grad = grad;
    A.gL = ws;
    A.gl_mats = ctx->n_mat;
    A.part = nullptr;
    // voxel gather: bricks of 16 x 16 x 4 voxels; cell gather: 8 x 8 x 4
    // cells from cell -1, so n + 1 cells per axis
    const int64_t ex = cell ? 1 : 0;
    const int bxd = cell ? kCellX : kBrickX, byd = cell ? kCellY : kBrickY, bzd = cell ? kCellZ : kBrickZ;
    A.bricks_x = (int32_t)((s.p_num_x + ex + bxd - 1) / bxd);
    A.bricks_y = (int32_t)((s.p_num_y + ex + byd - 1) / byd);
    const int64_t bricks_z = (s.p_num_z + ex + bzd - 1) / bzd;
    const int64_t bricks = (int64_t)A.bricks_x * A.bricks_y * bricks_z;
    if (bricks >= (1ll << 31)) return set_error(CTFP_ERR_UNSUPPORTED, "volume too large for the gather");
    if (cell) {
        int gmax = 1;
        for (int gi = 0; gi < ctx->n_groups; gi++) gmax = std::max(gmax, ctx->group_mats[gi]);
        const size_t need = (size_t)bricks * (size_t)gmax * kCellCorners * sizeof(float);
        if (ctx->cell_part_bytes < need) {
            dfree(ctx->cell_part);
            ctx->cell_part_bytes = 0;
            HIP_TRY(dev_malloc(&ctx->cell_part, need));
            ctx->cell_part_bytes = need;
        }
        A.part = ctx->cell_part;
    }
    for (int64_t v0 = 0; v0 < view_count; v0 += group) {
        const int64_t n = std::min(group, view_count - v0);
        AdjointArgs G{gI ? gI + v0 * npix : nullptr, gIAbsorb ? gIAbsorb + v0 * npix : nullptr,
                      L ? L + v0 * ctx->n_mat * npix : nullptr, grad};
        if (!grouped) {
            MarchArgs Q;
            if ((rc = fill_march_args(ctx, view_begin + v0, n, Q))) return rc;
            if ((rc = spectral_args(ctx, Q))) return rc;
            const uint32_t lb = ctx->n_mat == 3 ? 16u : 4u * (uint32_t)ctx->n_mat;
            Q.chunk = ctx->tune.chunk;
            Q.lds_cells = global ? 0 : (int32_t)((ctx->tune.lds_bytes - 256) / lb);
            hipLaunchKernelGGL(rfn, dim3(Q.per_xcd * 8), dim3(kBlock), shmem, (hipStream_t)stream, Q, G, ws);
            HIP_TRY(hipGetLastError());
        } else {
            const float *Lc = G.L;
            if (need_l) {
                for (int gi = 0; gi < ctx->n_groups; gi++)
                    if ((rc = launch_group(ctx, view_begin + v0, n, nullptr, nullptr, lbuf, stream, false, gi)))
                        return rc;
                Lc = lbuf;
            }
            MarchArgs Q;
            if ((rc = fill_march_args(ctx, view_begin + v0, n, Q))) return rc;
            if ((rc = spectral_args(ctx, Q))) return rc;
            const int64_t blocks = (n * npix + kBlock - 1) / kBlock;
            hipLaunchKernelGGL(ray_grad_from_l_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, (hipStream_t)stream,
                               Q, G, Lc, ctx->n_mat, ws);
            HIP_TRY(hipGetLastError());
        }
        // the voxel pass, per channel group, in launches of at most
        // kMaxGatherBricks bricks
        for (int gi = 0; gi < ctx->n_groups; gi++) {
            const int gm = ctx->group_mats[gi];
            GatherFn gfn = cell ? pick_cell_gather(gm) : pick_gather(gm);
            if (!gfn) return set_error(CTFP_ERR_UNSUPPORTED, "n_mat=%d", gm);
            if ((rc = fill_march_args(ctx, view_begin + v0, n, A.P, gi))) return rc;
            A.mat_off = ctx->group_moff[gi];
            A.cap = (int32_t)((kGatherLds - 256 - kGatherViews * (int)sizeof(BrickView)) /
                              ((cell ? cell_pixel_words(gm) : gather_pixel_words(gm)) * 4));
            for (int64_t b0 = 0; b0 < bricks; b0 += (int64_t)kMaxGatherBricks) {
                A.brick_base = (uint32_t)b0;
                const int64_t nb = std::min<int64_t>((int64_t)kMaxGatherBricks, bricks - b0);
                hipLaunchKernelGGL(gfn, dim3((unsigned)nb), dim3(cell ? kCellBlock : kGatherBlock), (size_t)kGatherLds,
                                   (hipStream_t)stream, A);
                HIP_TRY(hipGetLastError());
            }
            if (cell) {
                // the bricks' corner sums into grad, one pass over the voxels
                const int64_t nvox = s.p_num_x * s.p_num_y * s.p_num_z;
                const int64_t blocks = std::min<int64_t>((nvox + 255) / 256, 1ll << 20);
                hipLaunchKernelGGL(pick_cell_combine(gm), dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                                   (const float *)A.part, grad, (int)s.p_num_x, (int)s.p_num_y, (int)s.p_num_z,
                                   A.bricks_x, A.bricks_y, A.mat_off);
                HIP_TRY(hipGetLastError());
            }
        }
    }
    return CTFP_OK;
}

int ctfp_backward(ctfp_context *ctx, int64_t view_begin, int64_t view_count, const float *gI,
                  const float *gIAbsorb, const float *L, float *grad, void *stream)
{
    Range range_("ctfp_backward");
    return backward_impl(ctx, view_begin, view_count, gI, gIAbsorb, L, grad, stream, nullptr, 0);
}

int ctfp_backward_noalloc(ctfp_context *ctx, int64_t view_begin, int64_t view_count, const float *gI,
                          const float *gIAbsorb, const float *L, float *grad, void *stream)
{
    Range range_("ctfp_backward_noalloc");
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    DeviceGuard g(ctx->device);
    float *ws = nullptr;
    int rc = static_scratch(ctx->device, &ws);
    if (rc) return rc;
    return backward_impl(ctx, view_begin, view_count, gI, gIAbsorb, L, grad, stream, ws, kStaticScratchBytes);
}

int ctfp_line_integrals(ctfp_context *ctx, int64_t view_begin, int64_t view_count, float *L,
                        void *stream)
{
    Range range_("ctfp_line_integrals");
    return launch_march(ctx, view_begin, view_count, nullptr, nullptr, L, stream, false);
}

int ctfp_flat_field(ctfp_context *ctx, float *I0, void *stream)
{
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    if (!ctx->I0) return set_error(CTFP_ERR_STATE, "spectral weights not set");
    if (!I0) return set_error(CTFP_ERR_INVALID, "null output");
    DeviceGuard g(ctx->device);
    const int64_t npix = ctx->scan.d_num_u * ctx->scan.d_num_v;
    HIP_TRY(hipMemcpyAsync(I0, ctx->I0, (size_t)npix * sizeof(float), hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
    return CTFP_OK;
}

int ctfp_forward_host(ctfp_context *ctx, float *I, float *IAbsorb, float *I0)
{
    Range range_("ctfp_forward_host");
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    if (!I || !IAbsorb) return set_error(CTFP_ERR_INVALID, "null output");
    DeviceGuard g(ctx->device);
    const int64_t npix = ctx->scan.d_num_u * ctx->scan.d_num_v;
    // Bounded device staging: views in chunks of at most ~1 GiB of output.
    int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(ctx->scan.proj_num,
                                                           (int64_t)(1ll << 27) / std::max<int64_t>(npix, 1)));
    if (ctx->staged_views < chunk) {
        dfree(ctx->dI);
        dfree(ctx->dIA);
        ctx->staged_views = 0;
        HIP_TRY(dev_malloc(&ctx->dI, (size_t)chunk * npix * sizeof(float)));
        HIP_TRY(dev_malloc(&ctx->dIA, (size_t)chunk * npix * sizeof(float)));
        ctx->staged_views = chunk;
    }
    for (int64_t v0 = 0; v0 < ctx->scan.proj_num; v0 += chunk) {
        int64_t n = std::min(chunk, ctx->scan.proj_num - v0);
        int rc = ctfp_forward(ctx, v0, n, ctx->dI, ctx->dIA, nullptr, nullptr);
        if (rc) return rc;
        HIP_TRY(hipMemcpy(I + v0 * npix, ctx->dI, (size_t)n * npix * sizeof(float), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(IAbsorb + v0 * npix, ctx->dIA, (size_t)n * npix * sizeof(float),
                          hipMemcpyDeviceToHost));
    }
    if (I0) {
        if (!ctx->I0) return set_error(CTFP_ERR_STATE, "spectral weights not set");
        HIP_TRY(hipMemcpy(I0, ctx->I0, (size_t)npix * sizeof(float), hipMemcpyDeviceToHost));
    }
    return CTFP_OK;
}

int ctfp_postprocess(int64_t n_views, int64_t npix, const float *I, const float *IAbsorb,
                     const float *I0, uint32_t *I_u32, uint32_t *IAbsorb_u32, float *proj,
                     void *stream)
{
    if (n_views < 0 || npix < 0) return set_error(CTFP_ERR_INVALID, "negative size");
    if (!I || (IAbsorb_u32 && !IAbsorb) || (proj && !I0))
        return set_error(CTFP_ERR_INVALID, "null input for a requested output");
    if (n_views * npix == 0) return CTFP_OK;
    postprocess_kernel<<<grid_for(n_views * npix), 256, 0, (hipStream_t)stream>>>(
        n_views, npix, I, IAbsorb, I0, I_u32, IAbsorb_u32, proj);
    HIP_TRY(hipGetLastError());
    return CTFP_OK;
}

int ctfp_synchronize(void *stream)
{
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return CTFP_OK;
}

int ctfp_count_samples(ctfp_context *ctx, int64_t view_begin, int64_t view_count, int64_t *marched)
{
    if (check_ctx(ctx) || !marched) return set_error(CTFP_ERR_INVALID, "null argument");
    MarchArgs P;
    int rc = fill_march_args(ctx, view_begin, std::min(view_count, views_per_launch(ctx)), P);
    if (rc) return rc;
    if (view_count > 0 && (rc = fill_march_args(ctx, view_begin + view_count - 1, 1, P))) return rc;
    DeviceGuard g(ctx->device);
    unsigned long long *d = nullptr;
    HIP_TRY(dev_malloc(&d, sizeof(unsigned long long)));
    unsigned long long h = 0;
    hipError_t e = hipMemcpy(d, &h, sizeof h, hipMemcpyHostToDevice);
    const int64_t per = views_per_launch(ctx);
    for (int64_t v0 = 0; v0 < view_count && e == hipSuccess; v0 += per) {
        if ((rc = fill_march_args(ctx, view_begin + v0, std::min(per, view_count - v0), P))) {
            (void)hipFree(d);
            return rc;
        }
        count_kernel<<<P.n_blocks, kBlock>>>(P, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return set_error(hip_status(e), "count: %s", hipGetErrorString(e));
    *marched = (int64_t)h;
    return CTFP_OK;
}

int ctfp_count_marched(ctfp_context *ctx, int64_t view_begin, int64_t view_count, int64_t *marched)
{
    if (check_ctx(ctx) || !marched) return set_error(CTFP_ERR_INVALID, "null argument");
    if (!ctx->vol) return set_error(CTFP_ERR_STATE, "no volumes uploaded");
    if (view_begin < 0 || view_count < 1 || view_begin + view_count > ctx->scan.proj_num)
        return set_error(CTFP_ERR_INVALID, "view range [%lld, %lld) outside [0, %lld)", (long long)view_begin,
                         (long long)(view_begin + view_count), (long long)ctx->scan.proj_num);
    DeviceGuard g(ctx->device);
    unsigned long long *d = nullptr;
    HIP_TRY(dev_malloc(&d, sizeof(unsigned long long)));
    unsigned long long h = 0;
    hipError_t e = hipMemcpy(d, &h, sizeof h, hipMemcpyHostToDevice);
    const int64_t per = views_per_launch(ctx);
    for (int gi = 0; gi < ctx->n_groups; gi++) {
        for (int64_t v0 = 0; v0 < view_count && e == hipSuccess; v0 += per) {
            MarchArgs P;
            int rc = forward_args(ctx, view_begin + v0, std::min(per, view_count - v0), nullptr, nullptr,
                                  (float *)d, false, P, gi);
            if (rc) {
                (void)hipFree(d);
                return rc;
            }
            if (ctx->dtype == CTFP_F32)
                count_marched_kernel<kStF32><<<P.per_xcd * 8, kBlock>>>(P, d);
            else
                count_marched_kernel<kStF16><<<P.per_xcd * 8, kBlock>>>(P, d);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return set_error(hip_status(e), "count: %s", hipGetErrorString(e));
    *marched = (int64_t)h;
    return CTFP_OK;
}

// ---- autotuner ----------------------------------------------------------

int ctfp_autotune(ctfp_context *ctx, int max_lds_bytes, ctfp_tuning *best)
{
    Range range_("ctfp_autotune");
    if (check_ctx(ctx)) return CTFP_ERR_INVALID;
    if (!ctx->vol || !ctx->W) return set_error(CTFP_ERR_STATE, "autotune needs volumes and spectral tables");
    const ctfp_scan_params &s = ctx->scan;
    const int64_t npix = s.d_num_u * s.d_num_v;
    // time enough views to fill the GPU several times (>= 8192 tiles of
    // 256 rays), in groups at 0, 1/8, 1/4 and 3/8 of the scan's turn: rays
    // along the volume's axes and across its diagonals stage chunk boxes of
    // very different sizes, so a sample of one kind alone misjudges the
    // staging step and the LDS budget
    constexpr int kGroups = 4;
    const int64_t tiles_per_view = (npix + 255) / 256;
    const int64_t want = std::max<int64_t>(1, (8192 + tiles_per_view - 1) / tiles_per_view);
    const int64_t per_group =
        std::max<int64_t>(1, std::min<int64_t>((want + kGroups - 1) / kGroups, s.proj_num / kGroups));
    int64_t starts[kGroups];
    for (int gi = 0; gi < kGroups; gi++) starts[gi] = std::min<int64_t>(gi * s.proj_num / 8, s.proj_num - per_group);
    DeviceGuard g(ctx->device);
    float *I = nullptr, *IA = nullptr;
    HIP_TRY(dev_malloc(&I, (size_t)per_group * npix * sizeof(float)));
    hipError_t e = dev_malloc(&IA, (size_t)per_group * npix * sizeof(float));
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&ev0);
    if (e == hipSuccess) e = hipEventCreate(&ev1);
    const ctfp_tuning saved = ctx->tune;
    auto cleanup = [&]() {
        ctx->tune = saved;
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        dfree(I);
        dfree(IA);
    };
    if (e != hipSuccess) {
        cleanup();
        return set_error(hip_status(e), "autotune set-up: %s", hipGetErrorString(e));
    }
    // time one candidate (ms over all groups)
    auto time_it = [&](const ctfp_tuning &t, float &ms) -> int {
        ctx->tune = t;
        for (int rep = 0; rep < 2; rep++) { // rep 0 warms caches and code
            hipError_t r = hipEventRecord(ev0, nullptr);
            for (int gi = 0; gi < kGroups && r == hipSuccess; gi++) {
                int rc = launch_march(ctx, starts[gi], per_group, I, IA, nullptr, nullptr, true);
                if (rc) return rc;
            }
            if (r == hipSuccess) r = hipEventRecord(ev1, nullptr);
            if (r == hipSuccess) r = hipEventSynchronize(ev1);
            if (r == hipSuccess) r = hipEventElapsedTime(&ms, ev0, ev1);
            if (r != hipSuccess) return set_error(CTFP_ERR_DEVICE, "autotune: %s", hipGetErrorString(r));
        }
        return CTFP_OK;
    };
    ctfp_tuning cur = saved;
    cur.march = CTFP_MARCH_LDS;
    if (max_lds_bytes > 0 && cur.lds_bytes > max_lds_bytes) cur.lds_bytes = std::max(1024, max_lds_bytes);
    float best_ms = 0.0f;
    int rc = time_it(cur, best_ms);
    // ray-tile shapes (wave_u, block_u), then staging step and LDS budget
    const int shapes[][2] = {{8, 2}, {16, 1}, {4, 4}, {16, 2}, {8, 1}, {32, 1}, {8, 4}};
    ctfp_tuning pick = cur;
    for (const auto &sh : shapes) {
        if (rc) break;
        ctfp_tuning t = cur;
        t.wave_u = sh[0];
        t.block_u = sh[1];
        float ms = 0.0f;
        rc = time_it(t, ms);
        if (!rc && ms < best_ms) { best_ms = ms; pick = t; }
    }
    const int chunks[] = {16, 32};
    const int budgets[] = {24 * 1024, 32 * 1024, 40 * 1024, 48 * 1024, 64 * 1024};
    const ctfp_tuning shape = pick;
    for (int c : chunks) {
        for (int b : budgets) {
            if (rc) break;
            if (max_lds_bytes > 0 && b > max_lds_bytes) continue;
            ctfp_tuning t = shape;
            t.chunk = c;
            t.lds_bytes = b;
            float ms = 0.0f;
            rc = time_it(t, ms);
            if (!rc && ms < best_ms) { best_ms = ms; pick = t; }
        }
    }
    cleanup();
    if (rc) return rc;
    ctx->tune = pick;
    if (best) *best = pick;
    return CTFP_OK;
}

// ---- batches -------------------------------------------------------------

} // extern "C"

struct ctfp_batch {
    // items of one LDS budget share a persistent launch (and its occupancy)
    struct Group {
        BatchItem *d_items = nullptr;
        uint32_t *d_base = nullptr; // [kXcds][n + 1]
        int n = 0;
        size_t shmem = 0;
        unsigned grid = 0;
    };
    int device = 0;
    int dtype = CTFP_F32;
    int n_items = 0;
    std::vector<ctfp_batch_item> items;
    std::vector<uint64_t> gens; // generation of each context at creation (checked per run)
    std::vector<Group> groups;
    uint32_t *d_taken = nullptr; // [groups][kXcds]
    // items of more than four materials: their group marches and spectrum
    // pass follow the persistent launches in the graph, through l_ws
    std::vector<int> grouped;
    float *l_ws = nullptr;
    size_t l_ws_bytes = 0;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t samples = 0;
};

namespace {
void batch_free(ctfp_batch *b)
{
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    for (auto &g : b->groups) {
        dfree(g.d_items);
        dfree(g.d_base);
    }
    dfree(b->d_taken);
    dfree(b->l_ws);
    delete b;
}

BatchFn pick_batch(int dtype)
{
    switch (dtype) {
    case CTFP_F32: return batch_kernel<kStF32>;
    case CTFP_F16: return batch_kernel<kStF16>;
    case CTFP_BF16: return batch_kernel<kStBF16>;
    default: return nullptr;
    }
}
} // namespace

extern "C" {

int ctfp_batch_create(const ctfp_batch_item *items, int n_items, ctfp_batch **out)
{
    if (!out) return set_error(CTFP_ERR_INVALID, "null output");
    *out = nullptr;
    if (!items || n_items < 1 || n_items > 4096)
        return set_error(CTFP_ERR_INVALID, "batch: %d items (1..4096)", n_items);
    const ctfp_context *c0 = items[0].ctx;
    if (check_ctx(c0)) return CTFP_ERR_INVALID;
    std::vector<BatchItem> h(n_items);
    std::vector<int> grouped;
    size_t l_ws_bytes = 0;
    int64_t samples = 0;
    for (int i = 0; i < n_items; i++) {
        ctfp_context *c = items[i].ctx;
        if (check_ctx(c)) return CTFP_ERR_INVALID;
        if (c->device != c0->device || c->dtype != c0->dtype)
            return set_error(CTFP_ERR_INVALID, "batch: item %d is on another device or storage type", i);
        if (c->tune.march != CTFP_MARCH_LDS)
            return set_error(CTFP_ERR_INVALID, "batch: item %d is tuned for the global march", i);
        if (c->spec_n_mat != c->n_mat || !c->W)
            return set_error(CTFP_ERR_STATE, "batch: item %d has no spectral tables for its volumes", i);
        if (c->n_groups > 1) {
            // more than four materials: checked here, run after the
            // persistent launches through a line-integral workspace
            MarchArgs P;
            const int64_t vb = items[i].view_begin, vc = items[i].view_count;
            if (vb < 0 || vc < 1 || vb + vc > c->scan.proj_num)
                return set_error(CTFP_ERR_INVALID, "batch: item %d's view range is outside its scan", i);
            int rc = fill_march_args(c, vb, 1, P);
            if (!rc) rc = spectral_args(c, P);
            if (rc) return rc;
            if (!items[i].I || !items[i].IAbsorb) return set_error(CTFP_ERR_INVALID, "batch: null output");
            const size_t per_view = (size_t)c->n_mat * c->scan.d_num_u * c->scan.d_num_v * sizeof(float);
            l_ws_bytes = std::max(l_ws_bytes, per_view * (size_t)std::max<int64_t>(1, l_views(c, kLWorkspaceBytes, vc)));
            grouped.push_back(i);
        } else {
            if (c->vol_bytes >= 0xFFFFFFFFull)
                return set_error(CTFP_ERR_UNSUPPORTED,
                                 "batch: item %d's volume spans 4 GiB or more (use ctfp_forward for it)", i);
            int rc = forward_args(c, items[i].view_begin, items[i].view_count, items[i].I, items[i].IAbsorb,
                                  nullptr, true, h[i].P);
            if (rc) return rc;
            h[i].n_mat = c->n_mat;
        }
        samples += items[i].view_count * c->scan.d_num_u * c->scan.d_num_v * c->sys.int_num;
    }
    auto is_grouped = [&](int i) { return std::find(grouped.begin(), grouped.end(), i) != grouped.end(); };
    // groups by LDS budget, in order of first appearance
    std::vector<int> budgets;
    for (int i = 0; i < n_items; i++)
        if (!is_grouped(i) &&
            std::find(budgets.begin(), budgets.end(), items[i].ctx->tune.lds_bytes) == budgets.end())
            budgets.push_back(items[i].ctx->tune.lds_bytes);
    DeviceGuard g(c0->device);
    BatchFn fn = pick_batch(c0->dtype);
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, c0->device));
    ctfp_batch *b = new ctfp_batch();
    b->device = c0->device;
    b->dtype = c0->dtype;
    b->n_items = n_items;
    b->items.assign(items, items + n_items);
    for (int i = 0; i < n_items; i++) b->gens.push_back(items[i].ctx->generation);
    b->samples = samples;
    hipError_t e = hipSuccess;
    auto fail = [&](int code, const char *what) {
        batch_free(b);
        return set_error(code, "batch: %s", what);
    };
    b->grouped = grouped;
    if (l_ws_bytes) {
        e = dev_malloc(&b->l_ws, l_ws_bytes);
        if (e != hipSuccess) return fail(hip_status(e), hipGetErrorString(e));
        b->l_ws_bytes = l_ws_bytes;
    }
    for (int budget : budgets) {
        std::vector<BatchItem> gi;
        for (int i = 0; i < n_items; i++)
            if (!is_grouped(i) && items[i].ctx->tune.lds_bytes == budget) gi.push_back(h[i]);
        const int n = (int)gi.size();
        std::vector<uint32_t> base((size_t)kXcds * (n + 1), 0u);
        for (int i = 0; i < n; i++) {
            const uint64_t tiles = (uint64_t)gi[i].P.tiles_u * gi[i].P.tiles_v;
            const int64_t vc = gi[i].P.n_views;
            for (int x = 0; x < kXcds; x++) {
                const int64_t nv = vc > x ? (vc - x + kXcds - 1) / kXcds : 0;
                const uint64_t next = base[(size_t)x * (n + 1) + i] + (uint64_t)nv * tiles;
                if (next >= (1ull << 31)) return fail(CTFP_ERR_INVALID, "too many tiles");
                base[(size_t)x * (n + 1) + i + 1] = (uint32_t)next;
            }
        }
        ctfp_batch::Group grp;
        grp.n = n;
        grp.shmem = (size_t)std::max(budget, 256);
        int per_cu = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, grp.shmem);
        if (e != hipSuccess) return fail(CTFP_ERR_DEVICE, hipGetErrorString(e));
        if (per_cu < 1) return fail(CTFP_ERR_UNSUPPORTED, "kernel does not fit a CU");
        grp.grid = std::max(8u, (unsigned)(per_cu * prop.multiProcessorCount) / 8 * 8);
        e = dev_malloc(&grp.d_items, gi.size() * sizeof(BatchItem));
        if (e == hipSuccess) e = hipMemcpy(grp.d_items, gi.data(), gi.size() * sizeof(BatchItem), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = dev_malloc(&grp.d_base, base.size() * sizeof(uint32_t));
        if (e == hipSuccess)
            e = hipMemcpy(grp.d_base, base.data(), base.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
        b->groups.push_back(grp);
        if (e != hipSuccess) return fail(hip_status(e),
                                         hipGetErrorString(e));
    }
    const int ng = (int)b->groups.size();
    if (ng > 0) e = dev_malloc(&b->d_taken, (size_t)ng * kXcds * sizeof(uint32_t));
    // the launch sequence (zero every queue counter, one persistent kernel
    // per group, then the items of more than four materials) is captured
    // once on a private stream and replayed
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
    int grc = CTFP_OK;
    if (e == hipSuccess) {
        hipError_t launch = hipSuccess;
        if (ng > 0) {
            batch_reset<<<1, 256, 0, b->cap_stream>>>(b->d_taken, ng * kXcds);
            launch = hipGetLastError();
        }
        for (int k = 0; k < ng && launch == hipSuccess; k++) {
            const ctfp_batch::Group &grp = b->groups[k];
            BatchQueues Q{b->d_taken + k * kXcds, grp.d_base, grp.n};
            hipLaunchKernelGGL(fn, dim3(grp.grid), dim3(kBlock), grp.shmem, b->cap_stream, grp.d_items, Q);
            launch = hipGetLastError();
        }
        for (size_t k = 0; k < grouped.size() && launch == hipSuccess && grc == CTFP_OK; k++) {
            const ctfp_batch_item &it = items[grouped[k]];
            grc = grouped_forward(it.ctx, it.view_begin, it.view_count, it.I, it.IAbsorb, nullptr, b->cap_stream,
                                  b->l_ws, b->l_ws_bytes);
        }
        e = hipStreamEndCapture(b->cap_stream, &b->graph);
        if (e == hipSuccess) e = launch;
    }
    if (grc != CTFP_OK) {
        const std::string msg = ctfp::last_error();
        batch_free(b);
        return set_error(grc, "batch: %s", msg.c_str());
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess)
        return fail(hip_status(e), hipGetErrorString(e));
    *out = b;
    return CTFP_OK;
}

int ctfp_batch_run(ctfp_batch *b, void *stream)
{
    Range range_("ctfp_batch_run");
    if (!b) return set_error(CTFP_ERR_INVALID, "null batch");
    for (int i = 0; i < b->n_items; i++)
        if (b->items[i].ctx->generation != b->gens[i])
            return set_error(CTFP_ERR_STATE,
                             "batch: the volumes or spectral tables of item %d changed; rebuild the batch", i);
    DeviceGuard g(b->device);
    HIP_TRY(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return CTFP_OK;
}

int ctfp_batch_samples(const ctfp_batch *b, int64_t *samples)
{
    if (!b || !samples) return set_error(CTFP_ERR_INVALID, "null argument");
    *samples = b->samples;
    return CTFP_OK;
}

int ctfp_batch_destroy(ctfp_batch *b)
{
    if (!b) return CTFP_OK;
    DeviceGuard g(b->device);
    batch_free(b);
    return CTFP_OK;
}

int ctfp_bench_copy(const void *src, void *dst, int64_t bytes, void *stream)
{
    if (!src || !dst || bytes < 0 || bytes % 16)
        return set_error(CTFP_ERR_INVALID, "copy: pointers and a multiple of 16 bytes required");
    int64_t n = bytes / 16;
    copy_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>((const float4 *)src, (float4 *)dst, n);
    HIP_TRY(hipGetLastError());
    return CTFP_OK;
}

} // extern "C"
