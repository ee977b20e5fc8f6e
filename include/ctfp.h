/*
 * ctfp.h -- C ABI of the MI355X polyenergetic cone-beam forward projector.
 *
 * This is the drop-in boundary for the reference's GPU layer
 * (Source/CUDAFun/CBCTPolyProjGPU.cuh:33-64).  The reference exposes C++
 * functions taking non-POD structs by reference/value and keeps its texture in
 * hidden globals; this ABI replaces them with plain structs, plain pointers,
 * an opaque per-scan context and integer status codes.  Every entry point
 * names the reference interface it replaces.
 *
 * Conventions
 *   - Status: every function returns CTFP_OK (0) or a negative CTFP_ERR_*;
 *     ctfp_last_error() returns the calling thread's last message.  Nothing
 *     here calls exit() (the reference does, e.g. CBCTForwardProj.cpp:155-160).
 *   - Layouts follow the reference: volumes [z][y][x] (x fastest,
 *     CBCTPolyProjGPU.cu:2184), projections [view][v][u]
 *     (CBCTPolyProjGPU.cu:1994), line integrals [view][material][v][u].
 *   - "Device pointer" arguments are HIP device allocations on the context's
 *     device; "stream" arguments are hipStream_t passed as void* (NULL = the
 *     null stream).
 *   - Units: mm, g/cm^3, 1/mm.  Energy bin k (0-based) is (k+1) keV
 *     (CBCTPolyForwardProjGrid.cpp:40, 387).
 *   - A context is not thread-safe; use one context per host thread.  Several
 *     contexts may coexist on one device.
 */
#ifndef CTFP_H
#define CTFP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTFP_OK 0
#define CTFP_ERR_INVALID -1     /* bad argument or geometry */
#define CTFP_ERR_DEVICE -2      /* HIP runtime error */
#define CTFP_ERR_NOMEM -3       /* host or device allocation failed */
#define CTFP_ERR_STATE -4       /* call out of order (e.g. forward before volumes) */
#define CTFP_ERR_IO -5          /* file could not be opened / read / written */
#define CTFP_ERR_CONFIG -6      /* configuration entry missing or invalid */
#define CTFP_ERR_UNSUPPORTED -7 /* valid request this build does not handle */

/* Materials per scan (the reference reads any MaterialNum,
 * Source/ForwardProjection.cpp:290-316).  Up to four are marched in one pass;
 * more run in groups of four, their line integrals joined by a separate
 * spectrum pass. */
#define CTFP_MAX_MATERIALS 16

/* Storage precision of the resident material volumes.  FP16 rounds each
 * density to nearest.  BF16 (8 significant bits) rounds each density to one
 * of its two neighbouring BF16 values with a fixed pseudo-random threshold
 * stratified over 2x2x2 voxel blocks (an unbiased dither), so that line
 * integrals through uniform regions keep their mean; a single voxel may then
 * be off by up to one BF16 step (2^-7 relative). */
#define CTFP_F32 0
#define CTFP_F16 1
#define CTFP_BF16 2

/* Scan description.  Mirrors the geometry members of CTScanParas
 * (Source/DataType.h:33-55) as a POD struct (no std::string, Q24). */
typedef struct ctfp_scan_params {
    int64_t p_num_x, p_num_y, p_num_z;     /* phantom voxels (pNumX/Y/Z) */
    float p_length_x, p_length_y, p_length_z; /* phantom size, mm (pLengthX/Y/Z) */
    int64_t proj_num;                       /* views over 360 degrees (projNum) */
    int64_t d_num_u, d_num_v;               /* detector pixels (dNumU, dNumV) */
    float sod, sdd;                         /* mm */
    int32_t rotated_direction;              /* +1 or -1 */
    float d_size;                           /* detector pitch, mm (dSize) */
} ctfp_scan_params;

/* Derived scalars.  Mirrors CTScanSystemInfo (Source/DataType.h:78-111). */
typedef struct ctfp_system_info {
    int32_t rotated_direction;
    float theta_step;                  /* degrees */
    float d_half_lu, d_half_lv;        /* half detector extents, mm */
    float fov_r, fov_h;                /* mm */
    float p_half_x, p_half_y, p_half_z;
    float p_size_x, p_size_y, p_size_z;
    float dx, dy, dz;                  /* integration step; dx is used */
    int64_t int_num;                   /* samples per ray, K */
} ctfp_system_info;

/* Spectral weights of one scan.  Replaces the per-energy uploads of the
 * reference's forward*GPU wrappers (CBCTPolyProjGPU.cu:25-30, 204-207, 393)
 * and the scalars spectrumVal / phantomMAtten (Source/DataType.h:109-110).
 * All pointers are host pointers and are copied; they may be freed after the
 * call.  Detector response, by priority: `det` table, else the scintillator
 * model (scint_mu with scint_thickness), else none (NoResponse variants). */
typedef struct ctfp_spectral {
    int32_t n_energy;              /* bins (specEnergyNum) */
    int32_t n_mat;                 /* materials; must match the volumes */
    const float *spectrum;         /* [n_energy] normalised spectrum S_E */
    const float *atten;            /* [n_energy][n_mat] attenuation per unit density,
                                      1/mm per g/cm^3 = (mu/rho)/10 */
    float i0_val;                  /* photons per pixel before response (I0Val) */
    const float *grid;             /* [n_energy][d_num_u] grid transmission T_E(u), or NULL */
    const float *det;              /* [n_energy][d_num_v*d_num_u] response eta, or NULL */
    const float *scint_mu;         /* [n_energy] scintillator mu, 1/mm, or NULL */
    const float *scint_thickness;  /* [d_num_v*d_num_u] scintillator thickness, mm */
    float det_factor;              /* photoelectric conversion factor f */
    int32_t energy_weighting;      /* 0: photon counting (the reference); 1: energy
                                      integrating, weights scaled by (k+1) keV */
} ctfp_spectral;

/* Focused anti-scatter grid.  Mirrors GridAndDetectorSystem (Source/Grid.h:42-66). */
typedef struct ctfp_grid_params {
    int64_t d_num;            /* pixels along u (dNums) */
    float d_size_um;          /* pixel pitch in um (dSize * 1000 in FP32); truncated to whole um
                                 as the reference does (Grid.h:126) */
    int32_t interspace_width; /* um */
    int32_t strip_width;      /* um */
    float height_mm;          /* lamella height h */
    float focal_mm;           /* focal distance FD */
} ctfp_grid_params;

/* Launch tuning of the ray-march (no reference counterpart: the reference
 * launches fixed 8x8x8 blocks, CBCTPolyProjGPU.cuh:22-24).  Tuning never
 * changes results: every setting accumulates each ray in the same order. */
#define CTFP_MARCH_GLOBAL 0 /* corners gathered from global memory */
#define CTFP_MARCH_LDS 1    /* per-chunk volume boxes staged in LDS (default) */
typedef struct ctfp_tuning {
    int32_t march;     /* CTFP_MARCH_GLOBAL or CTFP_MARCH_LDS */
    int32_t chunk;     /* samples per staged chunk (LDS mode): a multiple of 16 in [16, 4096]; default 16 */
    int32_t lds_bytes; /* LDS per workgroup (LDS mode), 1 KiB..64 KiB; default 40 KiB */
    int32_t wave_u;    /* ray tile: pixels along u per wavefront, 4/8/16/32 (64 / wave_u along v); default 8 */
    int32_t block_u;   /* ray tile: wavefronts along u per workgroup, 1/2/4 (4 / block_u along v); default 2 */
    int32_t adjoint;   /* ctfp_backward: CTFP_ADJOINT_CELL (default), CTFP_ADJOINT_GATHER or CTFP_ADJOINT_SCATTER */
    int32_t skip_empty; /* 1 (default): a staged chunk whose cell box is zero in every channel
                           (checked on an occupancy grid of 8^3-cell bricks built with the
                           volumes) is skipped; it would add exact zeros.  0: every chunk marched */
} ctfp_tuning;
#define CTFP_ADJOINT_GATHER 0  /* ray pass to dl/dL, then a voxel-driven gather: no atomics, deterministic */
#define CTFP_ADJOINT_SCATTER 1 /* ray-driven scatter through LDS boxes and float atomics */
#define CTFP_ADJOINT_CELL 2    /* ray pass to dl/dL, then a cell-driven gather (each sample's eight corner weights
                                  once) and a combine pass: no atomics, deterministic; holds a per-brick buffer of
                                  corner sums (about 1.4 x M floats per voxel) */

typedef struct ctfp_context ctfp_context;

/* ---- library ------------------------------------------------------------ */
const char *ctfp_last_error(void);
const char *ctfp_version(void);

/* Replaces CBCTForwardProj::computeParas (Source/CBCTForwardProj.cpp:23-62). */
int ctfp_compute_paras(const ctfp_scan_params *scan, ctfp_system_info *out);

/* ---- scan context --------------------------------------------------------- */
/* Replaces the allocation half of initDeviceGrid / initDeviceNoGrid /
 * initDeviceSinMat{,NoResponse}{Grid,NoGrid} (CBCTPolyProjGPU.cu:478-571,
 * 1112-1208, 573-745, 1210-1380 ...).  No cudaDeviceReset, no globals. */
int ctfp_create(const ctfp_scan_params *scan, int device, ctfp_context **out);
/* Replaces freeDeviceMemory (CBCTPolyProjGPU.cu:1774-1797). */
int ctfp_destroy(ctfp_context *ctx);
int ctfp_get_system_info(const ctfp_context *ctx, ctfp_system_info *out);

/* Upload n_mat density volumes ([z][y][x] FP32 each), kept resident in
 * `storage_dtype`.  Replaces initTexture3D / updateTex
 * (CBCTPolyProjGPU.cu:2175-2221), which re-uploaded mu(E) every energy.
 * volumes_on_device: 0 = host pointers, 1 = device pointers. */
int ctfp_set_volumes(ctfp_context *ctx, int n_mat, const float *const *volumes,
                     int volumes_on_device, int storage_dtype);

/* Focal spot (the reference's FoSp path, computeFoSpOffset /
 * transformFocalSpotKernel, CBCTForwardProj.cpp:122-148,
 * CBCTPolyProjGPU.cu:1927-1954, made physical): per view, the source moves
 * by (offset_u[view], offset_v[view]) mm along the detector's u and v
 * directions while the detector stays, so every ray still reaches the point
 * where its nominal ray meets the detector, from the displaced source, and
 * the sample lattice is measured from there.  (The reference instead shifts
 * every sample by the offset, moving the whole ray bundle.)  Host arrays of
 * proj_num floats, copied; NULL, NULL restores the point source. */
int ctfp_set_focal_spot(ctfp_context *ctx, const float *offset_u, const float *offset_v);

/* Per-view offsets for a focal spot of diameter spot_mm: the reference's
 * generator (MSVC rand() after srand(seed), the reference seeds 10; 0.1 um
 * steps) with the disc test u^2 + v^2 <= r^2 in place of its pi r^2. */
int ctfp_focal_spot_offsets(int64_t n_views, float spot_mm, uint32_t seed, float *offset_u,
                            float *offset_v);

/* Stream-ordered form of ctfp_set_volumes for FP32 device volumes (no
 * reference counterpart: the reference copies host volumes into its texture
 * arrays synchronously, initTexture3D / updateTex, CBCTPolyProjGPU.cu:2175-2221).  The layout pass is queued on `stream` and
 * the call returns without waiting; the volumes must stay valid until it has
 * run.  The first call, or a call with another n_mat, allocates. */
int ctfp_set_volumes_async(ctfp_context *ctx, int n_mat, const float *const *volumes,
                           int storage_dtype, void *stream);

/* The scan parameters and device of a context (accessors for bindings). */
int ctfp_get_scan(const ctfp_context *ctx, ctfp_scan_params *out);
int ctfp_get_device(const ctfp_context *ctx, int *device);
/* The material count and storage dtype of the volumes the context holds
 * (CTFP_ERR_STATE before the first upload).  An accessor for bindings that
 * keep a scan's line integrals for ctfp_backward and so need not upload the
 * same densities again. */
int ctfp_get_volumes(const ctfp_context *ctx, int *n_mat, int *storage_dtype);
/* Device memory the context holds (volume layout, spectral tables, staging). */
int ctfp_device_bytes(const ctfp_context *ctx, int64_t *bytes);

/* Build the density channels from a uint8 label volume on the device.
 * single_material = 0: channel m = density[m] where label == m+1, else 0
 *   (PhantomObject::updataLineAtten, PhantomObject.cpp:38-58, with (mu/rho)/10
 *   moved to ctfp_spectral.atten);
 * single_material = 1: one channel = label * density[0]
 *   (PhantomObject::computeDesity, PhantomObject.cpp:65-71). */
int ctfp_set_labels(ctfp_context *ctx, const uint8_t *labels, int labels_on_device,
                    int n_mat, const float *density, int single_material,
                    int storage_dtype);

/* Replaces the per-energy host uploads and scalars; see ctfp_spectral. */
int ctfp_set_spectral(ctfp_context *ctx, const ctfp_spectral *spec);

int ctfp_set_tuning(ctfp_context *ctx, const ctfp_tuning *tuning);
int ctfp_get_tuning(const ctfp_context *ctx, ctfp_tuning *tuning);

/* Polyenergetic forward projection of views [view_begin, view_begin+view_count):
 * ray-march of every material once, then the whole spectrum per ray.
 * Replaces forward{Proj,SinMatProj,SinMatNoResponseProj}{Grid,NoGrid}GPU
 * (CBCTPolyProjGPU.cu:12-474) together with the host energy loops that
 * called them per energy and accumulated (CBCTForwardProj.cpp:65-76).
 * I, IAbsorb: device [view_count][d_num_v][d_num_u]; L: device
 * [view_count][n_mat][d_num_v][d_num_u] or NULL.  Asynchronous on `stream`. */
int ctfp_forward(ctfp_context *ctx, int64_t view_begin, int64_t view_count, float *I,
                 float *IAbsorb, float *L, void *stream);

/* Line integrals only (the density march of initDeviceSinMat*, replacing
 * computeSinMatIndensityProjKernel, CBCTPolyProjGPU.cu:2005-2020).
 * L: device [view_count][n_mat][d_num_v][d_num_u]. */
int ctfp_line_integrals(ctfp_context *ctx, int64_t view_begin, int64_t view_count, float *L,
                        void *stream);

/* Adjoint of ctfp_forward (no reference counterpart: the reference has no
 * gradient).  For upstream gradients gI = dl/dI and gIAbsorb = dl/dIAbsorb
 * (device [view_count][d_num_v][d_num_u]; either may be NULL = zero), adds
 * dl/d(density) to `grad` (device FP32 [n_mat][p_num_z][p_num_y][p_num_x];
 * the caller zeroes it).  L (device [view_count][n_mat][d_num_v][d_num_u],
 * from ctfp_forward) may be NULL, then the line integrals are recomputed per
 * ray.  With CTFP_ADJOINT_CELL (the default) or CTFP_ADJOINT_GATHER the
 * result is deterministic; the call holds a workspace of dl/dL for up to
 * 1 GiB of views, and for the cell gather a buffer of per-brick corner sums
 * (about 1.4 x n_mat floats per voxel, at most 4 materials' worth), both
 * kept by the context.  With CTFP_ADJOINT_SCATTER accumulation uses FP32
 * atomics, so the last bits can differ between runs. */
int ctfp_backward(ctfp_context *ctx, int64_t view_begin, int64_t view_count, const float *gI,
                  const float *gIAbsorb, const float *L, float *grad, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight training loops, harnesses that
 * forbid hipMalloc): ctfp_backward with the gather adjoint's dl/dL workspace
 * in a static device array of CTFP_NOALLOC_SCRATCH_BYTES per GPU (held until
 * the process ends) instead of the context's allocation.  Calls no hipMalloc
 * or hipFree.  Under CTFP_ADJOINT_CELL it runs the voxel gather
 * (CTFP_ADJOINT_GATHER), which needs no corner-sum buffer: the result then
 * equals ctfp_backward's under CTFP_ADJOINT_GATHER bit for bit and its
 * default result to FP32 rounding.  Views run in groups whose dl/dL (n_mat * d_num_v * d_num_u * 4
 * bytes per view) fits the array; a scan one view of which needs more is
 * rejected with CTFP_ERR_UNSUPPORTED.  Every call on one GPU shares the
 * array: calls that overlap on one GPU (several streams or host threads)
 * race.  Arguments and results as ctfp_backward (the caller zeroes grad). */
#define CTFP_NOALLOC_SCRATCH_BYTES (64ll << 20)
int ctfp_backward_noalloc(ctfp_context *ctx, int64_t view_begin, int64_t view_count, const float *gI,
                          const float *gIAbsorb, const float *L, float *grad, void *stream);

/* Flat field I0 = sum_E I0Val*S_E*T_E*eta_E in bin order.  Replaces addI0 /
 * addI0NoResponse (CBCTPolyForwardProjGrid.cpp:330-350,
 * CBCTPolyForwardProjNoGrid.cpp:290-310).  I0: device [d_num_v][d_num_u]. */
int ctfp_flat_field(ctfp_context *ctx, float *I0, void *stream);

/* Whole scan into host buffers (convenience for the reference-shaped classes):
 * I, IAbsorb [proj_num][d_num_v][d_num_u], I0 [d_num_v][d_num_u] (NULL skips). */
int ctfp_forward_host(ctfp_context *ctx, float *I, float *IAbsorb, float *I0);

/* ---- launch tuning and batches (no reference counterpart) -------------- */

/* Number of samples the march evaluates for views [view_begin, +view_count):
 * each ray's window clipped to the volume's support, against the fixed
 * K per ray of the reference's lattice (the rest are exact zeros).
 * Synchronous. */
int ctfp_count_samples(ctfp_context *ctx, int64_t view_begin, int64_t view_count, int64_t *marched);

/* Samples the LDS march evaluates with the context's tuning: those of
 * ctfp_count_samples less the samples of staged chunks skipped as empty
 * (tuning.skip_empty), summed over the material groups (each group of a scan
 * of more than four materials is marched on its own).  Synchronous. */
int ctfp_count_marched(ctfp_context *ctx, int64_t view_begin, int64_t view_count, int64_t *marched);

/* Test hook; no counterpart in the reference.  After `after` more device
 * allocations by this library (by any context, any thread), every device
 * allocation fails as out of memory (CTFP_ERR_NOMEM) until the hook is reset
 * with a negative value.  For tests of the error paths: a failed call leaves
 * its context usable (an upload or table update that failed part way leaves
 * no volume or no spectral tables, and calls that need them report
 * CTFP_ERR_STATE). */
int ctfp_debug_fail_allocations(long long after);


/* Times candidate launch options (ray-tile shape, samples per staged chunk,
 * LDS budget) on four groups of the scan's views at 0, 1/8, 1/4 and 3/8 of
 * its turn (rays along the volume's axes and across its diagonals), keeps
 * the fastest in the context and returns it in `best` (may be NULL).
 * max_lds_bytes > 0 caps the LDS budget (0: no cap).
 * Needs volumes and spectral tables; results are bit-identical for every
 * candidate, so tuning only changes speed.  Synchronous; allocates scratch
 * output for the timed views. */
int ctfp_autotune(ctfp_context *ctx, int max_lds_bytes, ctfp_tuning *best);

/* Several scans in one persistent launch.  Each item is a context (its own
 * volume, detector, views and spectral tables, LDS march mode) and a view
 * range with device outputs I, IAbsorb [view_count][d_num_v][d_num_u].  All
 * items must share one device and one storage type.  Items with the same LDS
 * budget share one persistent launch; the launches (one per budget) are
 * captured once into a HIP graph at creation and replayed by ctfp_batch_run.
 * The contexts must outlive the batch, and their volumes and spectral tables
 * must not be replaced while it exists (ctfp_batch_run checks the volumes). */
typedef struct ctfp_batch ctfp_batch;
typedef struct ctfp_batch_item {
    ctfp_context *ctx;
    int64_t view_begin, view_count;
    float *I, *IAbsorb;
} ctfp_batch_item;
int ctfp_batch_create(const ctfp_batch_item *items, int n_items, ctfp_batch **out);
int ctfp_batch_run(ctfp_batch *batch, void *stream);
int ctfp_batch_samples(const ctfp_batch *batch, int64_t *samples);
int ctfp_batch_destroy(ctfp_batch *batch);

/* Floor to uint32 and p = log(I0 / floor(I)).  Replaces convertIandIAbsorbDt
 * and computeProj (CBCTForwardProj.cpp:79-90, 376-387).  Device pointers;
 * any output may be NULL. */
int ctfp_postprocess(int64_t n_views, int64_t npix, const float *I, const float *IAbsorb,
                     const float *I0, uint32_t *I_u32, uint32_t *IAbsorb_u32, float *proj,
                     void *stream);

int ctfp_synchronize(void *stream);

/* ---- host-side detector and grid models ------------------------------------ */
/* Scintillator thickness map with the reference's jitter: seed 100, MSVC CRT
 * rand() (x = 214013x + 2531011 mod 2^32, bits 30..16).  Replaces
 * CBCTForwardProj::computePerScinThinckness (CBCTForwardProj.cpp:99-111). */
int ctfp_scint_thickness(int64_t npix, float thickness, float thickness_err, float *out);

/* eta = f (1 - exp(-mu_c t_j)).  Replaces CBCTForwardProj::updateDetResponse
 * (CBCTForwardProj.cpp:114-120). */
int ctfp_det_response(int64_t npix, float factor, float mu_c, const float *thickness,
                      float *out);

/* Grid transmission T_E(u) for n_energy energies, 1-um lamella model.
 * Replaces Grid::defGrid / updataGrid / binning1D (Grid.cpp:373-443, 297-309).
 * mu_strip, mu_inter: [n_energy] in 1/mm; T: [n_energy][d_num]. */
int ctfp_grid_transmission(const ctfp_grid_params *grid, int n_energy, const float *mu_strip,
                           const float *mu_inter, float *T);

/* Slant heights (um) of the grid's 1-um lamella units, h sqrt(FD^2 + d_i^2) / FD
 * for unit centres d_i = i + 0.5 - G/2, G = d_num * (whole-um pitch).
 * Replaces Grid::computeunitH (Grid.cpp:76-87), whose table Grid::showGridH
 * writes (Grid.cpp:506-523).  H holds up to `capacity` entries; *count
 * receives G. */
int ctfp_grid_unit_heights(const ctfp_grid_params *grid, float *H, int64_t capacity, int64_t *count);

/* Grid period indices: 0, every strict local minimum of T, d_num-1.
 * Replaces Grid::RetrievePeriodIndex (Grid.cpp:605-620).  `out` holds up to
 * `capacity` entries; *count receives the number found. */
int ctfp_grid_period_index(int64_t d_num, const float *T, int64_t *out, int64_t capacity,
                           int64_t *count);

/* ---- measurement helper ----------------------------------------------------- */
/* Streaming device copy of `bytes` (a multiple of 16) on `stream`.  No
 * counterpart in the reference: a kernel of known traffic used to calibrate
 * hardware byte counters (FETCH_SIZE / WRITE_SIZE) against the projector. */
int ctfp_bench_copy(const void *src, void *dst, int64_t bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CTFP_H */
