// ctfp_host.cpp -- host half of libctfp: scan scalars, the focused-grid
// transmission model, the scintillator thickness map and response.
//
// These are host-side (per-scan or per-energy, O(pixels)) computations of the
// reference's physics; they produce the tables the device kernels consume.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "ctfp.h"
#include "ctfp_internal.h"

using ctfp::set_error;

namespace ctfp {

float view_angle(const ctfp_scan_params *scan, const ctfp_system_info *sys, int64_t view)
{
    const float pi = acosf(-1.0f);
    float ri = scan->rotated_direction >= 0
                   ? (float)((uint64_t)scan->rotated_direction * (uint64_t)view)
                   : -(float)((int64_t)(-scan->rotated_direction) * view);
    return ri * sys->theta_step / 180 * pi;
}

} // namespace ctfp

extern "C" {

int ctfp_compute_paras(const ctfp_scan_params *s, ctfp_system_info *o)
{
    if (!s || !o) return set_error(CTFP_ERR_INVALID, "null argument");
    if (s->p_num_x < 1 || s->p_num_y < 1 || s->p_num_z < 1 || s->p_num_x > 8192 ||
        s->p_num_y > 8192 || s->p_num_z > 8192)
        return set_error(CTFP_ERR_INVALID, "phantom voxels %lldx%lldx%lld out of range",
                         (long long)s->p_num_x, (long long)s->p_num_y, (long long)s->p_num_z);
    if (!(s->p_length_x > 0) || !(s->p_length_y > 0) || !(s->p_length_z > 0))
        return set_error(CTFP_ERR_INVALID, "phantom lengths must be positive");
    if (s->d_num_u < 1 || s->d_num_v < 1 || s->d_num_u > 65536 || s->d_num_v > 65536)
        return set_error(CTFP_ERR_INVALID, "detector %lldx%lld out of range",
                         (long long)s->d_num_u, (long long)s->d_num_v);
    if (s->proj_num < 1 || s->proj_num > (1 << 24))
        return set_error(CTFP_ERR_INVALID, "proj_num %lld out of range", (long long)s->proj_num);
    if (!(s->d_size > 0) || !(s->sod > 0) || !(s->sdd > 0))
        return set_error(CTFP_ERR_INVALID, "d_size, sod and sdd must be positive");
    if (s->rotated_direction != 1 && s->rotated_direction != -1)
        return set_error(CTFP_ERR_INVALID, "rotated_direction must be +1 or -1");
    std::memset(o, 0, sizeof *o);
    // CBCTForwardProj::computeParas, CBCTForwardProj.cpp:23-62, in FP32.
    o->rotated_direction = s->rotated_direction;
    o->theta_step = (float)360 / (float)s->proj_num;
    o->d_half_lu = (float)s->d_num_u * s->d_size / 2;
    o->d_half_lv = (float)s->d_num_v * s->d_size / 2;
    o->fov_r = s->sod * o->d_half_lu / sqrtf(o->d_half_lu * o->d_half_lu + s->sdd * s->sdd);
    o->fov_h = (s->sod - o->fov_r) * o->d_half_lv / s->sdd;
    o->p_size_x = s->p_length_x / (float)s->p_num_x;
    o->p_size_y = s->p_length_y / (float)s->p_num_y;
    o->p_size_z = s->p_length_z / (float)s->p_num_z;
    o->p_half_x = s->p_length_x / 2;
    o->p_half_y = s->p_length_y / 2;
    o->p_half_z = s->p_length_z / 2;
    o->dx = (float)(0.5 * (double)o->p_size_x);
    o->dy = (float)(0.5 * (double)o->p_size_y);
    o->dz = (float)(0.5 * (double)o->p_size_z);
    float k = roundf(2 * o->fov_r / o->dx);
    if (!(k >= 1.0f) || k > (float)(1 << 30))
        return set_error(CTFP_ERR_INVALID, "samples per ray %g out of range", (double)k);
    o->int_num = (int64_t)k;
    return CTFP_OK;
}

// ---------------------------------------------------------------- detector

int ctfp_scint_thickness(int64_t npix, float thickness, float thickness_err, float *out)
{
    if (npix < 0 || !out) return set_error(CTFP_ERR_INVALID, "bad thickness arguments");
    if (!(thickness_err >= 0) || thickness_err > 1.0e4f)
        return set_error(CTFP_ERR_INVALID, "thickness_err out of range");
    // computePerScinThinckness (CBCTForwardProj.cpp:99-111): srand(100),
    // err = -e + (rand() % (end - beg + 1)) / 10000 with beg/end = -/+ e*1e4.
    uint32_t x = 100u;
    int beg = (int)(-thickness_err * 10000);
    int endd = (int)(thickness_err * 10000);
    int span = endd - beg + 1;
    for (int64_t i = 0; i < npix; i++) {
        x = x * 214013u + 2531011u;
        int r = (int)((x >> 16) & 0x7fff);
        float err = -thickness_err + (float)(r % span) / 10000.0f;
        out[i] = thickness + err;
    }
    return CTFP_OK;
}

int ctfp_det_response(int64_t npix, float factor, float mu_c, const float *thickness, float *out)
{
    if (npix < 0 || !thickness || !out) return set_error(CTFP_ERR_INVALID, "bad response arguments");
    for (int64_t i = 0; i < npix; i++) out[i] = factor * (1 - expf(-mu_c * thickness[i]));
    return CTFP_OK;
}

// -------------------------------------------------------------------- grid
//
// The reference discretises one detector row into G = d_num * D units of 1 um
// (D = pixel pitch in whole um) and fills them with four loop variants chosen
// by the parities of G and of the interspace width w (Grid.cpp:90-270).  Every
// variant reduces to the same picture, evaluated here per unit:
//   - a centre interspace run [c_lo, c_hi] where each unit uses its own slant
//     height H_i;
//   - below it, periods of P = strip + w units counted down from i0: strip
//     units use their own H_i, interspace units use H at the period start;
//   - above it, a mirror of the lower half (src = mirror_sum - i).
// Pixel u is the mean of its D units, summed in double in unit order
// (binning1D, Grid.cpp:297-309), so the result matches the loop form exactly.

namespace {

struct GridLayout {
    int64_t G, D;
    int w, sw, P;
    int64_t c_lo, c_hi, i0, mirror_from, mirror_sum;
};

int grid_layout(const ctfp_grid_params *g, GridLayout &L)
{
    if (!g || g->d_num < 1) return set_error(CTFP_ERR_INVALID, "grid: d_num must be positive");
    // mGridAndDetectorSystem.dSize = dSize * 1000 (float), stored to size_t (Grid.h:126)
    float d_um = g->d_size_um;
    if (!(d_um >= 1.0f) || d_um > 1.0e6f) return set_error(CTFP_ERR_INVALID, "grid: pixel pitch out of range");
    L.D = (int64_t)d_um;
    L.G = g->d_num * L.D;
    L.w = g->interspace_width;
    L.sw = g->strip_width;
    L.P = L.w + L.sw;
    if (L.w < 1 || L.sw < 0 || L.P < 1)
        return set_error(CTFP_ERR_INVALID, "grid: interspace width must be >= 1 um and strip >= 0");
    if (!(g->height_mm >= 0) || !(g->focal_mm > 0))
        return set_error(CTFP_ERR_INVALID, "grid: height >= 0 and focal distance > 0 required");
    if (L.G > (int64_t)1 << 31) return set_error(CTFP_ERR_INVALID, "grid: too many 1-um units");
    const int64_t h = L.G / 2, hw = L.w / 2;
    if (hw + 2 > h) return set_error(CTFP_ERR_INVALID, "grid: interspace wider than the detector");
    const bool godd = L.G % 2, wodd = L.w % 2;
    if (!godd && !wodd) { // uEdE, Grid.cpp:90-130
        L.c_lo = h - hw; L.c_hi = h - 1; L.i0 = h - 1 - hw; L.mirror_from = h; L.mirror_sum = L.G - 1;
    } else if (!godd && wodd) { // uEdO, Grid.cpp:133-174
        L.c_lo = h - hw; L.c_hi = h; L.i0 = h - 1 - hw; L.mirror_from = h + 1; L.mirror_sum = L.G;
    } else if (godd && !wodd) { // uOdE, Grid.cpp:177-219
        L.c_lo = h - hw + 1; L.c_hi = h; L.i0 = h - hw; L.mirror_from = h + 1; L.mirror_sum = L.G;
    } else { // uOdO, Grid.cpp:222-270
        L.c_lo = h - hw; L.c_hi = h; L.i0 = h - 1 - hw; L.mirror_from = h + 1; L.mirror_sum = L.G - 1;
    }
    return CTFP_OK;
}

} // namespace

int ctfp_grid_transmission(const ctfp_grid_params *g, int n_energy, const float *mu_strip,
                           const float *mu_inter, float *T)
{
    GridLayout L;
    int rc = grid_layout(g, L);
    if (rc) return rc;
    if (n_energy < 0 || (n_energy > 0 && (!mu_strip || !mu_inter || !T)))
        return set_error(CTFP_ERR_INVALID, "grid: bad energy tables");
    // Slant heights (um), Grid.cpp:76-87.
    const float h = g->height_mm * 1000, FD = g->focal_mm * 1000;
    std::vector<float> H((size_t)L.G);
    const double half = (double)((float)L.G / 2);
    for (int64_t i = 0; i < L.G; i++) {
        float dc = (float)((double)i + 0.5 - half);
        H[(size_t)i] = h / FD * sqrtf(FD * FD + dc * dc);
    }
    auto one_energy = [&](int e) {
        const float mus = mu_strip[e] / 1000, mui = mu_inter[e] / 1000; // 1/mm -> 1/um
        float *out = T + (int64_t)e * g->d_num;
        int64_t cached_period = -1;
        float cached_value = 0.0f;
        for (int64_t px = 0; px < g->d_num; px++) {
            double sum = 0.0;
            for (int64_t q = 0; q < L.D; q++) {
                int64_t i = px * L.D + q;
                if (i >= L.mirror_from) i = L.mirror_sum - i;
                float val;
                if (i >= L.c_lo && i <= L.c_hi) {
                    val = expf(-mui * H[(size_t)i]);
                } else {
                    int64_t d = L.i0 - i, j = d / L.P, r = d % L.P;
                    if (r < L.sw) {
                        val = expf(-mus * H[(size_t)i]);
                    } else {
                        if (j != cached_period) {
                            cached_period = j;
                            cached_value = expf(-mui * H[(size_t)(L.i0 - j * L.P)]);
                        }
                        val = cached_value;
                    }
                }
                sum += val;
            }
            out[px] = (float)(sum / (double)L.D);
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nt = std::min<unsigned>(nt, (unsigned)std::max(n_energy, 1));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++)
        pool.emplace_back([&, t]() {
            for (int e = (int)t; e < n_energy; e += (int)nt) one_energy(e);
        });
    for (auto &th : pool) th.join();
    return CTFP_OK;
}

int ctfp_grid_unit_heights(const ctfp_grid_params *g, float *H, int64_t capacity, int64_t *count)
{
    GridLayout L;
    int rc = grid_layout(g, L);
    if (rc) return rc;
    if (!count || (capacity > 0 && !H)) return set_error(CTFP_ERR_INVALID, "grid heights: bad output");
    *count = L.G;
    // Slant heights (um), Grid.cpp:76-87, in the same FP32 arithmetic as the
    // transmission's units.
    const float h = g->height_mm * 1000, FD = g->focal_mm * 1000;
    const double half = (double)((float)L.G / 2);
    for (int64_t i = 0; i < std::min(capacity, L.G); i++) {
        float dc = (float)((double)i + 0.5 - half);
        H[i] = h / FD * sqrtf(FD * FD + dc * dc);
    }
    return CTFP_OK;
}

int ctfp_grid_period_index(int64_t d_num, const float *T, int64_t *out, int64_t capacity, int64_t *count)
{
    if (d_num < 2 || !T || !count) return set_error(CTFP_ERR_INVALID, "period index: bad arguments");
    int64_t n = 0;
    auto push = [&](int64_t i) {
        if (out && n < capacity) out[n] = i;
        n++;
    };
    push(0);
    for (int64_t i = 1; i < d_num - 1; i++)
        if (T[i - 1] > T[i] && T[i] < T[i + 1]) push(i);
    push(d_num - 1);
    *count = n;
    return CTFP_OK;
}

/* computeFoSpOffset (CBCTForwardProj.cpp:122-148) with the disc test fixed:
 * MSVC rand() after srand(seed) (the reference uses 10), offsets
 * -r + (rand() % (size * 10000)) / 10000 mm, drawn again until
 * u^2 + v^2 <= r^2 (the reference compares against pi r^2). */
int ctfp_focal_spot_offsets(int64_t n_views, float spot_mm, uint32_t seed, float *offset_u, float *offset_v)
{
    if (n_views < 0 || !(spot_mm > 0.0f) || spot_mm > 100.0f || !offset_u || !offset_v)
        return set_error(CTFP_ERR_INVALID, "focal spot offsets: %lld views, spot %g mm",
                         (long long)n_views, (double)spot_mm);
    uint32_t x = seed;
    auto rnd = [&x]() {
        x = x * 214013u + 2531011u;
        return (int)((x >> 16) & 0x7FFF);
    };
    const float r = spot_mm / 2.0f;
    const int span = (int)(spot_mm * 10000);
    if (span < 1) return set_error(CTFP_ERR_INVALID, "focal spot below 0.1 um");
    for (int64_t i = 0; i < n_views; i++) {
        float u, v;
        do {
            u = -r + (float)(rnd() % span) / 10000.0f;
            v = -r + (float)(rnd() % span) / 10000.0f;
        } while (u * u + v * v > r * r);
        offset_u[i] = u;
        offset_v[i] = v;
    }
    return CTFP_OK;
}

} // extern "C"
