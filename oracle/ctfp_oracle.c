/*
 * ctfp_oracle.c -- CPU restatement of the reference projector's method.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product path (libctfp.so / the HIP kernels) never links,
 * loads or calls it.
 *
 * Two flavours of every numeric routine:
 *   *_f32  -- an FP32 mirror of the reference's arithmetic, operation by
 *             operation (same sample lattice, same rounding points, serial
 *             k-order FP32 sums, energy-ordered FP32 accumulation).  This is
 *             the 1e-4 parity target for the FP32 GPU path and the OpenMP CPU
 *             baseline.
 *   *_f64  -- the same method evaluated in FP64 (scalars set up exactly as the
 *             reference does in FP32, then every per-sample / per-ray
 *             operation in double).  Error reference for precision claims.
 *
 * Reference files are cited as path:line relative to the reference repository
 * (Source/...).  Per-function pinning lives in tests/test_oracle.py.
 *
 * Known unpinnable details (reference behaviour that cannot be reproduced
 * bit-for-bit without running the reference's CUDA 11.4 build):
 *   - tex3D linear filtering uses 9-bit fixed-point weights in hardware; this
 *     restatement interpolates with exact FP32 weights.
 *   - CUDA device powf/sqrtf/expf/cosf/sinf ULP behaviour; glibc (or x*x for
 *     powf(x,2)) is used here.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI_F ((float)acos(-1.0)) /* PI acosf(-1) in CBCTPolyProjGPU.cuh:21 */

typedef struct {
    int64_t nx, ny, nz;    /* pNumX/Y/Z      DataType.h:35 */
    float lx, ly, lz;      /* pLengthX/Y/Z   DataType.h:36 */
    int64_t n_views;       /* projNum        DataType.h:38 */
    int64_t nu, nv;        /* dNumU, dNumV   DataType.h:39 */
    float sod, sdd;        /* DataType.h:40 */
    int32_t rot;           /* rotatedDirection DataType.h:41 */
    float dsize;           /* DataType.h:43 */
} orc_geom;

typedef struct {
    float theta_step, half_u, half_v, fov_r, fov_h;
    float half_x, half_y, half_z;
    float px, py, pz;
    float dx, dy, dz;
    int64_t n_steps; /* intNum */
} orc_paras;

/* CBCTForwardProj::computeParas, CBCTForwardProj.cpp:23-62 (FP32). */
void orc_compute_paras(const orc_geom *g, orc_paras *s)
{
    s->theta_step = (float)360 / (float)g->n_views;
    s->half_u = (float)g->nu * g->dsize / 2;
    s->half_v = (float)g->nv * g->dsize / 2;
    s->fov_r = g->sod * s->half_u / sqrtf(s->half_u * s->half_u + g->sdd * g->sdd);
    s->fov_h = (g->sod - s->fov_r) * s->half_v / g->sdd;
    s->px = g->lx / (float)g->nx;
    s->py = g->ly / (float)g->ny;
    s->pz = g->lz / (float)g->nz;
    s->half_x = g->lx / 2;
    s->half_y = g->ly / 2;
    s->half_z = g->lz / 2;
    /* 0.5 * pSize is evaluated in double and stored to float (:55-57) */
    s->dx = (float)(0.5 * (double)s->px);
    s->dy = (float)(0.5 * (double)s->py);
    s->dz = (float)(0.5 * (double)s->pz);
    s->n_steps = (int64_t)roundf(2 * s->fov_r / s->dx); /* :60 */
}

/* View angle: CBCTPolyProjGPU.cu:51 `rotatedDirection * i * thetaStep / 180 * PI`.
 * For rot=+1 the product is (float)i.  For rot=-1 the reference's int*size_t
 * product wraps (a degenerate scan); the intended -i is used instead. */
float orc_view_angle(const orc_geom *g, const orc_paras *s, int64_t view)
{
    float ri = (g->rot >= 0) ? (float)((uint64_t)g->rot * (uint64_t)view)
                             : -(float)((int64_t)(-g->rot) * view);
    return ri * s->theta_step / 180 * ORC_PI_F;
}

/* ---------------------------------------------------------------- rays --- */

typedef struct { float cgcb, cgsb, sg; } ray_f32;
typedef struct { double cgcb, cgsb, sg; } ray_f64;

/* computeIntPointCoordinatesKernel, CBCTPolyProjGPU.cu:1813-1829: the mixed
 * gamma/beta convention is reproduced (rays miss pixel centres off-axis). */
static ray_f32 ray_dir_f32(const orc_geom *g, const orc_paras *s, int64_t u, int64_t v)
{
    float detU = s->half_u - g->dsize / 2 - (float)u * g->dsize;
    float detV = s->half_v - g->dsize / 2 - (float)v * g->dsize;
    float sdd2 = g->sdd * g->sdd;
    float r = sqrtf(sdd2 + detU * detU + detV * detV);
    float sg = detV / r;
    float cg = sqrtf(sdd2 + detU * detU) / r;
    float sb = detU / r;
    float cb = sqrtf(sdd2 + detV * detV) / r;
    ray_f32 d = {cg * cb, cg * sb, sg};
    return d;
}

static ray_f64 ray_dir_f64(const orc_geom *g, const orc_paras *s, int64_t u, int64_t v)
{
    double ds = g->dsize;
    double detU = (double)s->half_u - ds / 2 - (double)u * ds;
    double detV = (double)s->half_v - ds / 2 - (double)v * ds;
    double sdd2 = (double)g->sdd * (double)g->sdd;
    double r = sqrt(sdd2 + detU * detU + detV * detV);
    double sg = detV / r;
    double cg = sqrt(sdd2 + detU * detU) / r;
    double sb = detU / r;
    double cb = sqrt(sdd2 + detV * detV) / r;
    ray_f64 d = {cg * cb, cg * sb, sg};
    return d;
}

/* ------------------------------------------------------------ trilinear --- */

static inline float vox_f32(const float *vol, int64_t nx, int64_t ny, int64_t nz,
                            int64_t ix, int64_t iy, int64_t iz)
{
    if (ix < 0 || iy < 0 || iz < 0 || ix >= nx || iy >= ny || iz >= nz)
        return 0.0f; /* cudaAddressModeBorder, CBCTPolyProjGPU.cu:2200-2202 */
    return vol[(iz * ny + iy) * nx + ix];
}

/* tex3D<float> with cudaFilterModeLinear and unnormalised coordinates
 * (CBCTPolyProjGPU.cu:1920, 2203-2205): x_B = x - 0.5, i = floor(x_B),
 * alpha = frac(x_B), exact FP32 weights. */
static float tex3d_f32(const float *vol, int64_t nx, int64_t ny, int64_t nz,
                       float x, float y, float z)
{
    float xb = x - 0.5f, yb = y - 0.5f, zb = z - 0.5f;
    if (!(xb > -2.0f && xb < (float)nx + 1.0f && yb > -2.0f && yb < (float)ny + 1.0f &&
          zb > -2.0f && zb < (float)nz + 1.0f))
        return 0.0f;
    float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
    float ax = xb - fx, ay = yb - fy, az = zb - fz;
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    float c000 = vox_f32(vol, nx, ny, nz, ix, iy, iz);
    float c100 = vox_f32(vol, nx, ny, nz, ix + 1, iy, iz);
    float c010 = vox_f32(vol, nx, ny, nz, ix, iy + 1, iz);
    float c110 = vox_f32(vol, nx, ny, nz, ix + 1, iy + 1, iz);
    float c001 = vox_f32(vol, nx, ny, nz, ix, iy, iz + 1);
    float c101 = vox_f32(vol, nx, ny, nz, ix + 1, iy, iz + 1);
    float c011 = vox_f32(vol, nx, ny, nz, ix, iy + 1, iz + 1);
    float c111 = vox_f32(vol, nx, ny, nz, ix + 1, iy + 1, iz + 1);
    float c00 = (1.0f - ax) * c000 + ax * c100;
    float c10 = (1.0f - ax) * c010 + ax * c110;
    float c01 = (1.0f - ax) * c001 + ax * c101;
    float c11 = (1.0f - ax) * c011 + ax * c111;
    float c0 = (1.0f - ay) * c00 + ay * c10;
    float c1 = (1.0f - ay) * c01 + ay * c11;
    return (1.0f - az) * c0 + az * c1;
}

static inline double vox_f64(const float *vol, int64_t nx, int64_t ny, int64_t nz,
                             int64_t ix, int64_t iy, int64_t iz)
{
    if (ix < 0 || iy < 0 || iz < 0 || ix >= nx || iy >= ny || iz >= nz)
        return 0.0;
    return (double)vol[(iz * ny + iy) * nx + ix];
}

/* Same filter evaluated exactly in double at continuous index (x,y,z) where
 * voxel i sits at integer i (the +0.5 / -0.5 pair cancels exactly). */
static double tri_f64(const float *vol, int64_t nx, int64_t ny, int64_t nz,
                      double x, double y, double z)
{
    if (!(x > -2.0 && x < (double)nx + 1.0 && y > -2.0 && y < (double)ny + 1.0 &&
          z > -2.0 && z < (double)nz + 1.0))
        return 0.0;
    double fx = floor(x), fy = floor(y), fz = floor(z);
    double ax = x - fx, ay = y - fy, az = z - fz;
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double c000 = vox_f64(vol, nx, ny, nz, ix, iy, iz);
    double c100 = vox_f64(vol, nx, ny, nz, ix + 1, iy, iz);
    double c010 = vox_f64(vol, nx, ny, nz, ix, iy + 1, iz);
    double c110 = vox_f64(vol, nx, ny, nz, ix + 1, iy + 1, iz);
    double c001 = vox_f64(vol, nx, ny, nz, ix, iy, iz + 1);
    double c101 = vox_f64(vol, nx, ny, nz, ix + 1, iy, iz + 1);
    double c011 = vox_f64(vol, nx, ny, nz, ix, iy + 1, iz + 1);
    double c111 = vox_f64(vol, nx, ny, nz, ix + 1, iy + 1, iz + 1);
    double c00 = (1.0 - ax) * c000 + ax * c100;
    double c10 = (1.0 - ax) * c010 + ax * c110;
    double c01 = (1.0 - ax) * c001 + ax * c101;
    double c11 = (1.0 - ax) * c011 + ax * c111;
    double c0 = (1.0 - ay) * c00 + ay * c10;
    double c1 = (1.0 - ay) * c01 + ay * c11;
    return (1.0 - az) * c0 + az * c1;
}

/* Exposed for the trilinear known-answer tests. */
float orc_tex3d_f32(const float *vol, int64_t nx, int64_t ny, int64_t nz, float x, float y, float z)
{
    return tex3d_f32(vol, nx, ny, nz, x, y, z);
}
double orc_tri_f64(const float *vol, int64_t nx, int64_t ny, int64_t nz, double x, double y, double z)
{
    return tri_f64(vol, nx, ny, nz, x, y, z);
}

/* ----------------------------------------------------------- the march --- */

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* Sample position of step k on ray (u,v) at view angle theta, as texture
 * coordinates: computeIntPointCoordinatesKernel (.cu:1846-1848) followed by
 * transformKernel (.cu:1915-1920). */
static inline void sample_tex_f32(const orc_geom *g, const orc_paras *s, ray_f32 d,
                                  float t0, int64_t k, float cth, float sth,
                                  float *a, float *b, float *c)
{
    (void)g;
    float t = t0 + (float)k * s->dx;
    float X = d.cgcb * t - g->sod;
    float Y = d.cgsb * t;
    float Z = d.sg * t;
    float tu = X * cth - Y * sth + s->half_x;
    float tv = -(X * sth + Y * cth) + s->half_y;
    float tw = -Z + s->half_z;
    *a = tu / s->px + 0.5f;
    *b = tv / s->py + 0.5f;
    *c = tw / s->pz + 0.5f;
}

/* Per-material line integrals L_m = dx * sum_k f_m(P_k), FP32 mirror of
 * transformKernel + computeSinMatIndensityProjKernel (.cu:1902-1924,
 * 2005-2020): serial FP32 sum in ascending k, then * dx.
 * vols: n_mat volumes [nz][ny][nx] back to back.
 * L:    [n_views][n_mat][nv][nu] for views view0 .. view0+n_views-1. */
void orc_march_f32(const orc_geom *g, const orc_paras *s, int n_mat, const float *vols,
                   int64_t view0, int64_t n_views, float *L, int nthreads)
{
    set_threads(nthreads);
    const int64_t nvox = g->nx * g->ny * g->nz;
    const int64_t npix = g->nu * g->nv;
    const float t0 = g->sod - s->fov_r + s->dx / 2;
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int64_t iv = 0; iv < n_views; iv++) {
        for (int64_t v = 0; v < g->nv; v++) {
            float th = orc_view_angle(g, s, view0 + iv);
            float cth = cosf(th), sth = sinf(th);
            float acc[16];
            for (int64_t u = 0; u < g->nu; u++) {
                ray_f32 d = ray_dir_f32(g, s, u, v);
                for (int m = 0; m < n_mat; m++) acc[m] = 0.0f;
                for (int64_t k = 0; k < s->n_steps; k++) {
                    float a, b, c;
                    sample_tex_f32(g, s, d, t0, k, cth, sth, &a, &b, &c);
                    for (int m = 0; m < n_mat; m++)
                        acc[m] += tex3d_f32(vols + m * nvox, g->nx, g->ny, g->nz, a, b, c);
                }
                for (int m = 0; m < n_mat; m++)
                    L[((iv * n_mat + m) * g->nv + v) * g->nu + u] = acc[m] * s->dx;
            }
        }
    }
}

/* A ray in the source frame (x along the central ray, y with detector u,
 * z with detector v): position(t) = (x0, y0, z0) + t * (dx, dy, dz). */
typedef struct { double x0, y0, z0, dx, dy, dz; } line_f64;

/* The nominal ray of pixel (u, v) leaves the source at (-SOD, 0, 0).  With a
 * focal-spot offset (su, sv) (mm, along y and z) the source moves to
 * (-SOD, su, sv) while the detector stays: the ray still reaches the point
 * where the nominal ray meets the detector plane x = SDD - SOD, and the
 * sample lattice t_k is measured from the displaced source.  (The reference's
 * focal-spot path instead shifts every sample by the offset,
 * CBCTPolyProjGPU.cu:1944-1945, which moves the whole ray bundle.) */
static line_f64 ray_line_f64(const orc_geom *g, const orc_paras *s, int64_t u, int64_t v,
                             const double *focal)
{
    ray_f64 d = ray_dir_f64(g, s, u, v);
    line_f64 r;
    if (!focal) {
        r.x0 = -(double)g->sod; r.y0 = 0.0; r.z0 = 0.0;
        r.dx = d.cgcb; r.dy = d.cgsb; r.dz = d.sg;
        return r;
    }
    const double sdd = g->sdd, su = focal[0], sv = focal[1];
    const double qx = sdd, qy = sdd * d.cgsb / d.cgcb - su, qz = sdd * d.sg / d.cgcb - sv;
    const double n = sqrt(qx * qx + qy * qy + qz * qz);
    r.x0 = -(double)g->sod; r.y0 = su; r.z0 = sv;
    r.dx = qx / n; r.dy = qy / n; r.dz = qz / n;
    return r;
}

/* FP64 evaluation of the same lattice and filter.  focal: per-view source
 * offsets [n_views][2] (mm) indexed by absolute view, or NULL. */
void orc_march_focal_f64(const orc_geom *g, const orc_paras *s, int n_mat, const float *vols,
                         int64_t view0, int64_t n_views, const double *focal, double *L, int nthreads)
{
    set_threads(nthreads);
    const int64_t nvox = g->nx * g->ny * g->nz;
    const double t0 = (double)g->sod - (double)s->fov_r + (double)s->dx / 2;
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int64_t iv = 0; iv < n_views; iv++) {
        for (int64_t v = 0; v < g->nv; v++) {
            double th = (double)orc_view_angle(g, s, view0 + iv);
            double cth = cos(th), sth = sin(th);
            double acc[16];
            const double *f = focal ? focal + 2 * (view0 + iv) : NULL;
            for (int64_t u = 0; u < g->nu; u++) {
                line_f64 d = ray_line_f64(g, s, u, v, f);
                for (int m = 0; m < n_mat; m++) acc[m] = 0.0;
                for (int64_t k = 0; k < s->n_steps; k++) {
                    double t = t0 + (double)k * (double)s->dx;
                    double X = d.dx * t + d.x0;
                    double Y = d.dy * t + d.y0;
                    double Z = d.dz * t + d.z0;
                    double a = (X * cth - Y * sth + s->half_x) / s->px;
                    double b = (-(X * sth + Y * cth) + s->half_y) / s->py;
                    double c = (-Z + s->half_z) / s->pz;
                    for (int m = 0; m < n_mat; m++)
                        acc[m] += tri_f64(vols + m * nvox, g->nx, g->ny, g->nz, a, b, c);
                }
                for (int m = 0; m < n_mat; m++)
                    L[((iv * n_mat + m) * g->nv + v) * g->nu + u] = acc[m] * (double)s->dx;
            }
        }
    }
}

void orc_march_f64(const orc_geom *g, const orc_paras *s, int n_mat, const float *vols,
                   int64_t view0, int64_t n_views, double *L, int nthreads)
{
    orc_march_focal_f64(g, s, n_mat, vols, view0, n_views, NULL, L, nthreads);
}

/* Focal-spot offsets of the reference's computeFoSpOffset
 * (CBCTForwardProj.cpp:122-148): MSVC rand() after srand(10), offsets
 * -r + (rand() % (size * 10000)) / 10000 mm, drawn again until inside the
 * spot.  The reference's test compares against pi r^2 (and so keeps points
 * outside the disc); here the test is the disc, u^2 + v^2 <= r^2. */
void orc_focal_offsets(int64_t n_views, float spot_mm, uint32_t seed, float *ou, float *ov)
{
    uint32_t x = seed;
    const float r = spot_mm / 2.0f;
    const int span = (int)(spot_mm * 10000);
    for (int64_t i = 0; i < n_views; i++) {
        float u, v;
        do {
            x = x * 214013u + 2531011u;
            u = -r + (float)((int)((x >> 16) & 0x7FFF) % span) / 10000.0f;
            x = x * 214013u + 2531011u;
            v = -r + (float)((int)((x >> 16) & 0x7FFF) % span) / 10000.0f;
        } while (u * u + v * v > r * r);
        ou[i] = u;
        ov[i] = v;
    }
}

/* Texture coordinates of sample k on ray (u,v) at a view, for geometry tests. */
void orc_sample_position_f32(const orc_geom *g, const orc_paras *s, int64_t view,
                             int64_t u, int64_t v, int64_t k, float *abc)
{
    float th = orc_view_angle(g, s, view);
    ray_f32 d = ray_dir_f32(g, s, u, v);
    float t0 = g->sod - s->fov_r + s->dx / 2;
    sample_tex_f32(g, s, d, t0, k, cosf(th), sinf(th), abc, abc + 1, abc + 2);
}

void orc_ray_direction_f64(const orc_geom *g, const orc_paras *s, int64_t u, int64_t v, double *dir)
{
    ray_f64 d = ray_dir_f64(g, s, u, v);
    dir[0] = d.cgcb; dir[1] = d.cgsb; dir[2] = d.sg;
}

/* -------------------------------------------------------- the spectrum --- */

/* Energy-ordered FP32 accumulation of the polyenergetic intensities from the
 * per-material line integrals, mirroring (per energy E, bin order):
 *   forwardProjGridKernel .cu:1994-1995 / forwardSinMatProjGridKernel
 *   .cu:2055-2056 and their NoGrid/NoResponse variants:
 *     I_E  = I0Val * S_E * expf(-tau) * T_E[u] * eta_E[v,u]
 *     IA_E = I0Val * S_E * (1 - expf(-tau)) * T_E[u] * eta_E[v,u]
 *   addIandIAbsorb CBCTForwardProj.cpp:65-76 (I += I_E, IA += IA_E)
 *   addI0 CBCTPolyForwardProjGrid.cpp:330-339 (I0 += I0Val*S_E*T_E*eta_E)
 * with tau = sum_m atten[E][m] * L_m.  For one material this is exactly the
 * reference's single-material form expf(-a * proj); for several materials it
 * is the per-material decomposition of the reference's per-energy re-march
 * (checked against orc_forward_literal_f32).
 * grid (T) [n_energy][nu] and det (eta) [n_energy][nv*nu] may be NULL.
 * L: [n_views][n_mat][nv][nu];  I, IA: [n_views][nv][nu];  I0: [nv][nu]. */
void orc_spectrum_f32(int64_t nu, int64_t nv, int64_t n_views, int n_mat, int n_energy,
                      const float *L, const float *atten, const float *spectrum, float i0val,
                      const float *grid, const float *det, float *I, float *IA, float *I0,
                      int nthreads)
{
    set_threads(nthreads);
    const int64_t npix = nu * nv;
#pragma omp parallel for schedule(static)
    for (int64_t iv = 0; iv < n_views; iv++) {
        for (int64_t p = 0; p < npix; p++) {
            int64_t u = p % nu;
            float accI = 0.0f, accA = 0.0f;
            for (int e = 0; e < n_energy; e++) {
                float tau = 0.0f;
                for (int m = 0; m < n_mat; m++)
                    tau += atten[e * n_mat + m] * L[(iv * n_mat + m) * npix + p];
                float ex = expf(-tau);
                float wi = i0val * spectrum[e] * ex;
                float wa = i0val * spectrum[e] * (1 - ex);
                if (grid) { wi = wi * grid[e * nu + u]; wa = wa * grid[e * nu + u]; }
                if (det) { wi = wi * det[(int64_t)e * npix + p]; wa = wa * det[(int64_t)e * npix + p]; }
                accI += wi;
                accA += wa;
            }
            I[iv * npix + p] = accI;
            IA[iv * npix + p] = accA;
        }
    }
    if (I0) {
        for (int64_t p = 0; p < npix; p++) {
            int64_t u = p % nu;
            float acc = 0.0f;
            for (int e = 0; e < n_energy; e++) {
                float w = i0val * spectrum[e];
                if (grid) w = w * grid[e * nu + u];
                if (det) w = w * det[(int64_t)e * npix + p];
                acc += w;
            }
            I0[p] = acc;
        }
    }
}

void orc_spectrum_f64(int64_t nu, int64_t nv, int64_t n_views, int n_mat, int n_energy,
                      const double *L, const float *atten, const float *spectrum, double i0val,
                      const float *grid, const float *det, double *I, double *IA, double *I0,
                      int nthreads)
{
    set_threads(nthreads);
    const int64_t npix = nu * nv;
#pragma omp parallel for schedule(static)
    for (int64_t iv = 0; iv < n_views; iv++) {
        for (int64_t p = 0; p < npix; p++) {
            int64_t u = p % nu;
            double accI = 0.0, accA = 0.0;
            for (int e = 0; e < n_energy; e++) {
                double tau = 0.0;
                for (int m = 0; m < n_mat; m++)
                    tau += (double)atten[e * n_mat + m] * L[(iv * n_mat + m) * npix + p];
                double ex = exp(-tau);
                double w = i0val * (double)spectrum[e];
                if (grid) w *= (double)grid[e * nu + u];
                if (det) w *= (double)det[(int64_t)e * npix + p];
                accI += w * ex;
                accA += w * (1.0 - ex);
            }
            I[iv * npix + p] = accI;
            IA[iv * npix + p] = accA;
        }
    }
    if (I0) {
        for (int64_t p = 0; p < npix; p++) {
            int64_t u = p % nu;
            double acc = 0.0;
            for (int e = 0; e < n_energy; e++) {
                double w = i0val * (double)spectrum[e];
                if (grid) w *= (double)grid[e * nu + u];
                if (det) w *= (double)det[(int64_t)e * npix + p];
                acc += w;
            }
            I0[p] = acc;
        }
    }
}

/* The reference's literal multi-material method, kept to prove the
 * decomposition: for each energy, build the linear-attenuation volume from
 * the label volume (PhantomObject::updataLineAtten, PhantomObject.cpp:38-58),
 * march it (transformKernel + forwardProj{Grid,NoGrid}Kernel,
 * .cu:1902-1924, 1980-2003, 2079-2108) and accumulate on the host
 * (CBCTForwardProj.cpp:65-76).  Small problems only: it re-marches per energy.
 * mass_atten: [n_energy][n_mat] in cm^2/g; density: [n_mat] in g/cm^3. */
void orc_forward_literal_f32(const orc_geom *g, const orc_paras *s, const uint8_t *labels,
                             int n_mat, const float *density, int n_energy,
                             const float *mass_atten, const float *spectrum, float i0val,
                             const float *grid, const float *det, float *I, float *IA,
                             float *I0, int nthreads)
{
    set_threads(nthreads);
    const int64_t nvox = g->nx * g->ny * g->nz;
    const int64_t npix = g->nu * g->nv;
    float *mu = (float *)calloc((size_t)nvox, sizeof(float));
    float *tau = (float *)malloc((size_t)(g->n_views * npix) * sizeof(float));
    memset(I, 0, sizeof(float) * (size_t)(g->n_views * npix));
    memset(IA, 0, sizeof(float) * (size_t)(g->n_views * npix));
    if (I0) memset(I0, 0, sizeof(float) * (size_t)npix);
    for (int e = 0; e < n_energy; e++) {
        /* updataLineAtten: labels 1..n_mat get (mu/rho)*rho/10; label 0 gets
         * 0; other labels keep the buffer's previous value (zero-initialised,
         * CBCTPolyForwardProjGrid.cpp:114) */
        for (int64_t i = 0; i < nvox; i++) {
            for (int j = 0; j < n_mat; j++) {
                if (labels[i] == j + 1) {
                    mu[i] = mass_atten[e * n_mat + j] * density[j] / 10.f;
                    break;
                } else if (labels[i] == 0) {
                    mu[i] = 0;
                    break;
                }
            }
        }
        orc_march_f32(g, s, 1, mu, 0, g->n_views, tau, nthreads);
        const float sv = spectrum[e];
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < g->n_views * npix; q++) {
            int64_t p = q % npix, u = p % g->nu;
            /* tau[] holds temp*dx; the kernel evaluates expf(-temp*dx) */
            float ex = expf(-tau[q]);
            float wi = i0val * sv * ex;
            float wa = i0val * sv * (1 - ex);
            if (grid) { wi = wi * grid[e * g->nu + u]; wa = wa * grid[e * g->nu + u]; }
            if (det) { wi = wi * det[(int64_t)e * npix + p]; wa = wa * det[(int64_t)e * npix + p]; }
            I[q] += wi;
            IA[q] += wa;
        }
        if (I0) {
            for (int64_t p = 0; p < npix; p++) {
                int64_t u = p % g->nu;
                float w = i0val * sv;
                if (grid) w = w * grid[e * g->nu + u];
                if (det) w = w * det[(int64_t)e * npix + p];
                I0[p] += w;
            }
        }
    }
    free(mu);
    free(tau);
}

/* ---------------------------------------------------------- phantoms ---- */

/* Per-material density channels from a label volume: channel m holds
 * density[m] where label == m+1 (PhantomObject.cpp:43-56 with the (mu/rho)/10
 * factor moved into the spectrum stage).  Labels above n_mat stay 0 (Q9). */
void orc_labels_to_density_channels(const uint8_t *labels, int64_t nvox, int n_mat,
                                    const float *density, float *out)
{
    for (int m = 0; m < n_mat; m++)
        for (int64_t i = 0; i < nvox; i++)
            out[m * nvox + i] = (labels[i] == m + 1) ? density[m] : 0.0f;
}

/* PhantomObject::computeDesity, PhantomObject.cpp:65-71: label * rho0. */
void orc_labels_to_density_single(const uint8_t *labels, int64_t nvox, float rho0, float *out)
{
    for (int64_t i = 0; i < nvox; i++) out[i] = labels[i] * rho0;
}

/* ------------------------------------------------------------- detector --- */

/* MSVC CRT rand(): the published LCG x = x*214013 + 2531011 (mod 2^32),
 * output bits 30..16 (RAND_MAX 32767). */
static uint32_t msvc_state = 1;
static void msvc_srand(uint32_t s) { msvc_state = s; }
static int msvc_rand(void)
{
    msvc_state = msvc_state * 214013u + 2531011u;
    return (int)((msvc_state >> 16) & 0x7fff);
}

void orc_msvc_rand_sequence(uint32_t seed, int n, int *out)
{
    msvc_srand(seed);
    for (int i = 0; i < n; i++) out[i] = msvc_rand();
}

/* CBCTForwardProj::computePerScinThinckness, CBCTForwardProj.cpp:99-111. */
void orc_scint_thickness(int64_t npix, float thickness, float thickness_err, float *out)
{
    msvc_srand(100);
    int beg = (int)(-thickness_err * 10000);
    int endd = (int)(thickness_err * 10000);
    for (int64_t i = 0; i < npix; i++) {
        float err = -thickness_err + (float)(msvc_rand() % (endd - beg + 1)) / 10000.0f;
        out[i] = thickness + err;
    }
}

/* CBCTForwardProj::updateDetResponse, CBCTForwardProj.cpp:114-120:
 * eta = f * (1 - expf(-mu_c * t_j)), mu_c in 1/mm. */
void orc_det_response(int64_t npix, float factor, float mu_c, const float *thickness, float *out)
{
    for (int64_t i = 0; i < npix; i++)
        out[i] = factor * (1 - expf(-mu_c * thickness[i]));
}

/* ----------------------------------------------------------------- grid --- */

/* Literal restatement of the focused-grid model, Grid.cpp:76-87 (unit
 * heights), :90-270 (four parity variants), :297-309 (double binning),
 * :373-443 (defGrid/updataGrid).  One energy per call.
 *   d_num: pixels along u; d_size_um: pixel pitch in um as the reference
 *   truncates it (size_t, Q15); w, sw: interspace / strip widths (um);
 *   h_mm, fd_mm: height / focal distance; mu_s, mu_i: 1/mm.
 * Writes T[d_num].  Returns 0, or -1 if the layout is degenerate. */
int orc_grid_transmission(int64_t d_num, int64_t d_size_um, int w, int sw, float h_mm,
                          float fd_mm, float mu_s_mm, float mu_i_mm, float *T)
{
    size_t G = (size_t)d_num * (size_t)d_size_um;
    if (G < 4 || w < 1 || sw < 0 || (size_t)(w / 2 + 2) > G / 2) return -1;
    float h = h_mm * 1000, FD = fd_mm * 1000;
    float mus = mu_s_mm / 1000, mui = mu_i_mm / 1000;
    int P = w + sw;
    float *dc = (float *)malloc(G * sizeof(float));
    float *H = (float *)malloc(G * sizeof(float));
    float *g1 = (float *)malloc(G * sizeof(float));
    for (size_t i = 0; i < G; i++) dc[i] = (float)((double)i + 0.5 - (double)((float)G / 2));
    for (size_t i = 0; i < G; i++) H[i] = h / FD * sqrtf(FD * FD + dc[i] * dc[i]);
    for (size_t i = 0; i < G; i++) g1[i] = NAN;
    int uflag = (int)(G % 2), wflag = w % 2;
    if (uflag == 0 && wflag == 0) { /* uEdE :90-130 */
        for (size_t i = G / 2 - 1; i > G / 2 - 1 - w / 2; i--) g1[i] = expf(-mui * H[i]);
        for (int i = (int)(G / 2 - 1 - w / 2); i >= 0; i -= P) {
            for (int si = 0; si < sw; si++) { if (i - si < 0) break; g1[i - si] = expf(-mus * H[i - si]); }
            for (int ii = 0; ii < w; ii++) { if (i - sw - ii < 0) break; g1[i - sw - ii] = expf(-mui * H[i]); }
        }
        for (size_t i = G / 2; i < G; i++) g1[i] = g1[G - i - 1];
    } else if (uflag == 0 && wflag == 1) { /* uEdO :133-174 */
        g1[G / 2] = expf(-mui * H[G / 2]);
        for (size_t i = G / 2 - 1; i > G / 2 - 1 - w / 2; i--) g1[i] = expf(-mui * H[i]);
        for (int i = (int)(G / 2 - 1 - w / 2); i >= 0; i -= P) {
            for (int si = 0; si < sw; si++) { if (i - si < 0) break; g1[i - si] = expf(-mus * H[i - si]); }
            for (int ii = 0; ii < w; ii++) { if (i - sw - ii < 0) break; g1[i - sw - ii] = expf(-mui * H[i]); }
            if (i - P < 0) break;
        }
        for (size_t i = G / 2 + 1; i < G; i++) g1[i] = g1[G - i];
    } else if (uflag == 1 && wflag == 0) { /* uOdE :177-219 */
        for (size_t i = G / 2; i > G / 2 - w / 2; i--) g1[i] = expf(-mui * H[i]);
        for (int i = (int)(G / 2 - w / 2); i >= 0; i -= P) {
            for (int si = 0; si < sw; si++) { if (i - si < 0) break; g1[i - si] = expf(-mus * H[i - si]); }
            for (int ii = 0; ii < w; ii++) { if (i - sw - ii < 0) break; g1[i - sw - ii] = expf(-mui * H[i]); }
            if (i - P < 0) break;
        }
        for (size_t i = G / 2 + 1; i < G; i++) g1[i] = g1[G - i];
    } else { /* uOdO :222-270 */
        g1[G / 2] = expf(-mui * H[G / 2]);
        for (size_t i = G / 2 - 1; i > G / 2 - 1 - w / 2; i--) g1[i] = expf(-mui * H[i]);
        for (int i = (int)(G / 2 - 1 - w / 2); i >= 0; i -= P) {
            for (int si = 0; si < sw; si++) { if (i - si < 0) break; g1[i - si] = expf(-mus * H[i - si]); }
            for (int ii = 0; ii < w; ii++) { if (i - sw - ii < 0) break; g1[i - sw - ii] = expf(-mui * H[i]); }
            if (i - P < 0) break;
        }
        for (size_t i = G / 2 + 1; i < G; i++) g1[i] = g1[G - i - 1];
    }
    /* binning1D :297-309 */
    for (size_t i = 0; i < G; i += (size_t)d_size_um) {
        double sum = 0.0;
        for (size_t j = 0; j < (size_t)d_size_um; j++) sum += g1[i + j];
        T[i / (size_t)d_size_um] = (float)(sum / (double)d_size_um);
    }
    free(dc); free(H); free(g1);
    return 0;
}

/* Unit heights H_i (um) of the 1-um discretised lamellae, Grid.cpp:76-87. */
void orc_grid_unit_heights(int64_t n_units, float h_mm, float fd_mm, float *H)
{
    size_t G = (size_t)n_units;
    float h = h_mm * 1000, FD = fd_mm * 1000;
    for (size_t i = 0; i < G; i++) {
        float dc = (float)((double)i + 0.5 - (double)((float)G / 2));
        H[i] = h / FD * sqrtf(FD * FD + dc * dc);
    }
}

/* ------------------------------------------------------ post-processing --- */

/* convertIandIAbsorbDt CBCTForwardProj.cpp:79-90 and computeProj :376-387. */
void orc_postprocess(int64_t n_views, int64_t npix, const float *I, const float *IA,
                     const float *I0, uint32_t *Iu, uint32_t *IAu, float *proj)
{
    for (int64_t q = 0; q < n_views * npix; q++) {
        Iu[q] = (uint32_t)floorf(I[q]);
        if (IAu) IAu[q] = (uint32_t)floorf(IA[q]);
    }
    if (proj)
        for (int64_t j = 0; j < n_views; j++)
            for (int64_t i = 0; i < npix; i++)
                proj[j * npix + i] = logf(I0[i] / (float)Iu[j * npix + i]);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------- adjoint --- */

/* Exact adjoint of orc_march_f64's linear map rho_m -> L_m: every sample k of
 * ray (view, v, u) adds dx * gL * w_corner to the 8 corners of its cell
 * (zero-padded border: corners outside the volume are dropped).
 * gL: [n_views][n_mat][nv][nu]; grad: [n_mat][nz][ny][nx], accumulated.
 * OpenMP over detector rows with atomic adds: the order in which threads
 * meet at a voxel varies, which moves only the last bits of its FP64 sum. */
/* The same accumulated into the voxels of the box [box0, box0 + boxn)
 * (z, y, x) only: grad [n_mat][boxn_z][boxn_y][boxn_x].  For volumes whose
 * full gradient would not fit in memory. */
void orc_backproject_box_f64(const orc_geom *g, const orc_paras *s, int n_mat, int64_t view0,
                             int64_t n_views, const double *focal, const double *gL, const int64_t *box0,
                             const int64_t *boxn, double *grad);

void orc_backproject_focal_f64(const orc_geom *g, const orc_paras *s, int n_mat, int64_t view0,
                               int64_t n_views, const double *focal, const double *gL, double *grad)
{
    const int64_t box0[3] = {0, 0, 0}, boxn[3] = {g->nz, g->ny, g->nx};
    orc_backproject_box_f64(g, s, n_mat, view0, n_views, focal, gL, box0, boxn, grad);
}

void orc_backproject_box_f64(const orc_geom *g, const orc_paras *s, int n_mat, int64_t view0,
                             int64_t n_views, const double *focal, const double *gL, const int64_t *box0,
                             const int64_t *boxn, double *grad)
{
    const int64_t nvox = boxn[0] * boxn[1] * boxn[2];
    const int64_t npix = g->nu * g->nv;
    const double t0 = (double)g->sod - (double)s->fov_r + (double)s->dx / 2;
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int64_t iv = 0; iv < n_views; iv++) {
        for (int64_t v = 0; v < g->nv; v++) {
            const double th = (double)orc_view_angle(g, s, view0 + iv);
            const double cth = cos(th), sth = sin(th);
            for (int64_t u = 0; u < g->nu; u++) {
                line_f64 d = ray_line_f64(g, s, u, v, focal ? focal + 2 * (view0 + iv) : NULL);
                double gm[16];
                int any = 0;
                for (int m = 0; m < n_mat; m++) {
                    gm[m] = gL[(iv * n_mat + m) * npix + v * g->nu + u] * (double)s->dx;
                    any |= gm[m] != 0.0;
                }
                if (!any) continue;
                for (int64_t k = 0; k < s->n_steps; k++) {
                    double t = t0 + (double)k * (double)s->dx;
                    double X = d.dx * t + d.x0, Y = d.dy * t + d.y0, Z = d.dz * t + d.z0;
                    double a = (X * cth - Y * sth + s->half_x) / s->px;
                    double b = (-(X * sth + Y * cth) + s->half_y) / s->py;
                    double c = (-Z + s->half_z) / s->pz;
                    if (!(a > -1.0 && a < (double)g->nx && b > -1.0 && b < (double)g->ny && c > -1.0 &&
                          c < (double)g->nz))
                        continue;
                    double fx = floor(a), fy = floor(b), fz = floor(c);
                    double w[2][3] = {{1.0 - (a - fx), 1.0 - (b - fy), 1.0 - (c - fz)},
                                      {a - fx, b - fy, c - fz}};
                    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
                    for (int dz = 0; dz < 2; dz++)
                        for (int dy = 0; dy < 2; dy++)
                            for (int dx_ = 0; dx_ < 2; dx_++) {
                                int64_t x = ix + dx_, y = iy + dy, z = iz + dz;
                                if (x < 0 || y < 0 || z < 0 || x >= g->nx || y >= g->ny || z >= g->nz)
                                    continue;
                                x -= box0[2];
                                y -= box0[1];
                                z -= box0[0];
                                if (x < 0 || y < 0 || z < 0 || x >= boxn[2] || y >= boxn[1] || z >= boxn[0])
                                    continue;
                                double wc = w[dx_][0] * w[dy][1] * w[dz][2];
                                int64_t q = (z * boxn[1] + y) * boxn[2] + x;
                                for (int m = 0; m < n_mat; m++) {
#pragma omp atomic
                                    grad[m * nvox + q] += gm[m] * wc;
                                }
                            }
                }
            }
        }
    }
}

void orc_backproject_f64(const orc_geom *g, const orc_paras *s, int n_mat, int64_t view0,
                         int64_t n_views, const double *gL, double *grad)
{
    orc_backproject_focal_f64(g, s, n_mat, view0, n_views, NULL, gL, grad);
}

/* Chain rule of the spectrum, per ray: with tau_e = sum_m atten[e][m] L_m,
 * I = sum_e W_e e^-tau_e and IA = sum_e W_e (1 - e^-tau_e),
 * dl/dL_m = (gIA - gI) * sum_e atten[e][m] W_e e^-tau_e,
 * W_e = i0val * S_e * T_e(u) * eta_e(v,u) (grid / det may be NULL). */
void orc_spectrum_grad_f64(int64_t nu, int64_t nv, int64_t n_views, int n_mat, int n_energy,
                           const double *L, const float *atten, const float *spectrum, double i0val,
                           const float *grid, const float *det, const double *gI, const double *gIA,
                           double *gL)
{
    const int64_t npix = nu * nv;
    for (int64_t iv = 0; iv < n_views; iv++) {
        for (int64_t p = 0; p < npix; p++) {
            int64_t u = p % nu;
            double s[16] = {0};
            for (int e = 0; e < n_energy; e++) {
                double tau = 0.0;
                for (int m = 0; m < n_mat; m++)
                    tau += (double)atten[e * n_mat + m] * L[(iv * n_mat + m) * npix + p];
                double w = i0val * (double)spectrum[e];
                if (grid) w *= (double)grid[e * nu + u];
                if (det) w *= (double)det[(int64_t)e * npix + p];
                double we = w * exp(-tau);
                for (int m = 0; m < n_mat; m++) s[m] += (double)atten[e * n_mat + m] * we;
            }
            double c = (gIA ? gIA[iv * npix + p] : 0.0) - (gI ? gI[iv * npix + p] : 0.0);
            for (int m = 0; m < n_mat; m++) gL[(iv * n_mat + m) * npix + p] = c * s[m];
        }
    }
}
