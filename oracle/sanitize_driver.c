/* sanitize_driver.c -- runs every function of the oracle (ctfp_oracle.c) once
 * on a small scan, for the CPU build with AddressSanitizer and
 * UndefinedBehaviorSanitizer (`make -C oracle sanitize`).  TEST
 * INFRASTRUCTURE: the CPU tests run it; a sanitizer report or a failed sanity
 * check ends it with a non-zero status.  It checks the memory and arithmetic
 * of the oracle, not its numbers (tests/test_oracle.py pins those). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int64_t nx, ny, nz;
    float lx, ly, lz;
    int64_t n_views;
    int64_t nu, nv;
    float sod, sdd;
    int32_t rot;
    float dsize;
} orc_geom;

typedef struct {
    float theta_step, half_u, half_v, fov_r, fov_h;
    float half_x, half_y, half_z;
    float px, py, pz;
    float dx, dy, dz;
    int64_t n_steps;
} orc_paras;

void orc_compute_paras(const orc_geom *g, orc_paras *s);
float orc_view_angle(const orc_geom *g, const orc_paras *s, int64_t view);
float orc_tex3d_f32(const float *vol, int64_t nx, int64_t ny, int64_t nz, float x, float y, float z);
double orc_tri_f64(const float *vol, int64_t nx, int64_t ny, int64_t nz, double x, double y, double z);
void orc_march_f32(const orc_geom *g, const orc_paras *s, int n_mat, const float *vols, int64_t view0,
                   int64_t n_views, float *L, int nthreads);
void orc_march_f64(const orc_geom *g, const orc_paras *s, int n_mat, const float *vols, int64_t view0,
                   int64_t n_views, double *L, int nthreads);
void orc_march_focal_f64(const orc_geom *g, const orc_paras *s, int n_mat, const float *vols, int64_t view0,
                         int64_t n_views, const double *focal, double *L, int nthreads);
void orc_focal_offsets(int64_t n_views, float spot_mm, uint32_t seed, float *ou, float *ov);
void orc_sample_position_f32(const orc_geom *g, const orc_paras *s, int64_t view, int64_t u, int64_t v,
                             int64_t k, float *abc);
void orc_ray_direction_f64(const orc_geom *g, const orc_paras *s, int64_t u, int64_t v, double *dir);
void orc_spectrum_f32(int64_t nu, int64_t nv, int64_t n_views, int n_mat, int n_energy, const float *L,
                      const float *atten, const float *spectrum, float i0val, const float *grid,
                      const float *det, float *I, float *IA, float *I0, int nthreads);
void orc_spectrum_f64(int64_t nu, int64_t nv, int64_t n_views, int n_mat, int n_energy, const double *L,
                      const float *atten, const float *spectrum, double i0val, const float *grid,
                      const float *det, double *I, double *IA, double *I0, int nthreads);
void orc_forward_literal_f32(const orc_geom *g, const orc_paras *s, const uint8_t *labels, int n_mat,
                             const float *density, int n_energy, const float *mass_atten,
                             const float *spectrum, float i0val, const float *grid, const float *det,
                             float *I, float *IA, float *I0, int nthreads);
void orc_labels_to_density_channels(const uint8_t *labels, int64_t nvox, int n_mat, const float *density,
                                    float *out);
void orc_labels_to_density_single(const uint8_t *labels, int64_t nvox, float rho0, float *out);
void orc_msvc_rand_sequence(uint32_t seed, int n, int *out);
void orc_scint_thickness(int64_t npix, float thickness, float thickness_err, float *out);
void orc_det_response(int64_t npix, float factor, float mu_c, const float *thickness, float *out);
int orc_grid_transmission(int64_t d_num, int64_t d_size_um, int w, int sw, float h_mm, float fd_mm,
                          float mu_s_mm, float mu_i_mm, float *T);
void orc_grid_unit_heights(int64_t n_units, float h_mm, float fd_mm, float *H);
void orc_postprocess(int64_t n_views, int64_t npix, const float *I, const float *IA, const float *I0,
                     uint32_t *Iu, uint32_t *IAu, float *proj);
void orc_backproject_f64(const orc_geom *g, const orc_paras *s, int n_mat, int64_t view0, int64_t n_views,
                         const double *gL, double *grad);
void orc_backproject_focal_f64(const orc_geom *g, const orc_paras *s, int n_mat, int64_t view0,
                               int64_t n_views, const double *focal, const double *gL, double *grad);
void orc_spectrum_grad_f64(int64_t nu, int64_t nv, int64_t n_views, int n_mat, int n_energy,
                           const double *L, const float *atten, const float *spectrum, double i0val,
                           const float *grid, const float *det, const double *gI, const double *gIA,
                           double *gL);

static void check(int ok, const char *what)
{
    if (!ok) {
        fprintf(stderr, "sanity check failed: %s\n", what);
        exit(1);
    }
}

int main(void)
{
    enum { N = 10, NU = 12, NV = 9, V = 3, M = 2, E = 6 };
    const int64_t nvox = (int64_t)N * N * N, npix = (int64_t)NU * NV;
    orc_geom g = {N, N, N, 10.0f, 10.0f, 10.0f, V, NU, NV, 60.0f, 90.0f, 1, 1.5f};
    orc_paras p;
    orc_compute_paras(&g, &p);
    check(p.n_steps > 0 && p.n_steps < 1000, "samples per ray");
    const int64_t K = p.n_steps;
    uint8_t *labels = malloc(nvox);
    float *vols = malloc(sizeof(float) * M * nvox);
    for (int64_t i = 0; i < nvox; i++) labels[i] = (uint8_t)((i * 7919) % 4);
    float density[M] = {1.0f, 1.9f};
    orc_labels_to_density_channels(labels, nvox, M, density, vols);
    float *single = malloc(sizeof(float) * nvox);
    orc_labels_to_density_single(labels, nvox, 0.5f, single);
    check(fabsf(orc_tex3d_f32(vols, N, N, N, 4.3f, 5.1f, 2.7f)) < 10.0f, "tex3d");
    check(fabs(orc_tri_f64(vols, N, N, N, 4.3, 5.1, 2.7)) < 10.0, "tri");
    (void)orc_view_angle(&g, &p, 2);
    float abc[3];
    orc_sample_position_f32(&g, &p, 1, 3, 4, K / 2, abc);
    double dir[3];
    orc_ray_direction_f64(&g, &p, 0, NV - 1, dir);
    check(fabs(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2] - 1.0) < 1e-12, "unit direction");

    float *L32 = malloc(sizeof(float) * V * M * npix);
    double *L64 = malloc(sizeof(double) * V * M * npix);
    orc_march_f32(&g, &p, M, vols, 0, V, L32, 2);
    orc_march_f64(&g, &p, M, vols, 0, V, L64, 2);
    float ou[V], ov[V];
    orc_focal_offsets(V, 0.8f, 10u, ou, ov);
    double focal[2 * V];
    for (int i = 0; i < V; i++) {
        focal[2 * i] = ou[i];
        focal[2 * i + 1] = ov[i];
    }
    double *Lf = malloc(sizeof(double) * V * M * npix);
    orc_march_focal_f64(&g, &p, M, vols, 0, V, focal, Lf, 2);

    float spectrum[E], atten[E * M], mass[E * M], grid[E * NU], det[E * NU * NV], thick[NU * NV];
    orc_scint_thickness(npix, 0.4f, 0.05f, thick);
    for (int e = 0; e < E; e++) {
        spectrum[e] = (float)(e + 1) / 21.0f;
        for (int m = 0; m < M; m++) {
            mass[e * M + m] = 0.5f + 0.3f * (float)(e + m);
            atten[e * M + m] = mass[e * M + m] / 10.0f;
        }
        check(orc_grid_transmission(NU, 400, 214, 36, 1.5f, 90.0f, 5.0f, 0.0f, grid + e * NU) == 0, "grid");
        orc_det_response(npix, 1.0f, 0.3f + 0.1f * (float)e, thick, det + e * npix);
    }
    float *H = malloc(sizeof(float) * NU * 400);
    orc_grid_unit_heights(NU * 400, 1.5f, 90.0f, H);
    int seq[5];
    orc_msvc_rand_sequence(1u, 5, seq);
    check(seq[0] == 41, "MSVC rand");

    float *I = malloc(sizeof(float) * V * npix), *IA = malloc(sizeof(float) * V * npix);
    float *I0 = malloc(sizeof(float) * npix);
    orc_spectrum_f32(NU, NV, V, M, E, L32, atten, spectrum, 1e5f, grid, det, I, IA, I0, 2);
    double *I64 = malloc(sizeof(double) * V * npix), *IA64 = malloc(sizeof(double) * V * npix);
    double *I064 = malloc(sizeof(double) * npix);
    orc_spectrum_f64(NU, NV, V, M, E, L64, atten, spectrum, 1e5, grid, det, I64, IA64, I064, 2);
    for (int64_t q = 0; q < V * npix; q++)
        check(fabs(I64[q] + IA64[q] - I064[q % npix]) <= 1e-9 * I064[q % npix], "I + IA = I0");
    float *Il = malloc(sizeof(float) * V * npix), *IAl = malloc(sizeof(float) * V * npix);
    float *I0l = malloc(sizeof(float) * npix);
    orc_forward_literal_f32(&g, &p, labels, M, density, E, mass, spectrum, 1e5f, grid, det, Il, IAl, I0l, 2);

    uint32_t *Iu = malloc(sizeof(uint32_t) * V * npix), *IAu = malloc(sizeof(uint32_t) * V * npix);
    float *proj = malloc(sizeof(float) * V * npix);
    orc_postprocess(V, npix, I, IA, I0, Iu, IAu, proj);

    double *gI = malloc(sizeof(double) * V * npix), *gIA = malloc(sizeof(double) * V * npix);
    for (int64_t q = 0; q < V * npix; q++) {
        gI[q] = (double)((q * 31) % 7) - 3.0;
        gIA[q] = (double)((q * 17) % 5) - 2.0;
    }
    double *gL = malloc(sizeof(double) * V * M * npix);
    orc_spectrum_grad_f64(NU, NV, V, M, E, L64, atten, spectrum, 1e5, grid, det, gI, gIA, gL);
    double *grad = calloc((size_t)M * nvox, sizeof(double));
    orc_backproject_f64(&g, &p, M, 0, V, gL, grad);
    orc_backproject_focal_f64(&g, &p, M, 0, V, focal, gL, grad);
    /* <A x, y> = <x, A^T y> on the linear march: a cheap end-to-end check */
    double lhs = 0.0, rhs = 0.0;
    double *gradx = calloc((size_t)M * nvox, sizeof(double));
    orc_backproject_f64(&g, &p, M, 0, V, gL, gradx);
    for (int64_t q = 0; q < V * M * npix; q++) lhs += L64[q] * gL[q];
    for (int64_t q = 0; q < M * nvox; q++) rhs += (double)vols[q] * gradx[q];
    check(fabs(lhs - rhs) <= 1e-9 * (fabs(lhs) + 1.0), "adjoint dot product");

    free(labels); free(vols); free(single); free(L32); free(L64); free(Lf); free(H);
    free(I); free(IA); free(I0); free(I64); free(IA64); free(I064); free(Il); free(IAl); free(I0l);
    free(Iu); free(IAu); free(proj); free(gI); free(gIA); free(gL); free(grad); free(gradx);
    printf("oracle sanitizer run: ok\n");
    return 0;
}
