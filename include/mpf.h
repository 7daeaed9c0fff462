/*
 * mpf.h -- C ABI of the MI355X multi-phase-field solidification solver.
 *
 * Two groups of entry points:
 *
 * 1. The resident solver (handle API).  The library owns the device state in
 *    its own compact layout (dense liquid + per-cell grain slots) and advances
 *    it with four to six HIP launches per step (a tile pass, a classification
 *    pass that finishes most cells, a merge pass over queued cells, and the
 *    interface passes).  It replaces the reference's host
 *    time loop with its three launches per step
 *      run_randommode_3d.py:267-287   (temp -> APT -> update -> swap)
 *      tests/verification_3d.py:830-905 (same, with the checked APT and a
 *                                        KMaxOverflowError on overflow)
 *    Inputs and outputs cross the boundary in the reference's dense layouts
 *    (phi[N][nx][ny][nz], mf[depth][nx][ny][nz], nf[nx][ny][nz]).
 *
 * 2. Drop-in kernels on caller-owned device arrays in the reference's dense
 *    layout, one per reference kernel (src/gpu_kernels_3d.py,
 *    src/gpu_kernels.py).
 *
 * Conventions: every function returns an int; 0 (or a non-negative count) is
 * success, negative values are errors (MPF_E*), and mpf_last_error() returns a
 * message for the calling thread.  No exceptions cross the ABI.  A handle is
 * not thread-safe and owns one HIP stream; handle calls are synchronous.
 */
#ifndef MPF_H
#define MPF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPF_ABI_VERSION 3

enum { MPF_FP32 = 0, MPF_FP64 = 1 };
enum {
    MPF_OK = 0,
    MPF_EINVAL = -1,   /* bad argument or configuration */
    MPF_EHIP = -2,     /* HIP runtime error */
    MPF_ENOMEM = -3,   /* device allocation failed */
    MPF_ECAPACITY = -4 /* a cell holds more grains than the slot capacity */
};

typedef struct mpf_handle mpf_handle;

/* Scalars of the update kernel, in the order of its signature
 * (src/gpu_kernels_3d.py:531-537, :684-692). */
typedef struct mpf_scalars {
    double dx, dt, T_melt, Sf, eps0_sl, w0_sl;
    double a0, delta_a, mu_a, p_round, g2_floor, ksi, theta_c_rad;
} mpf_scalars;

typedef struct mpf_config {
    int nx, ny, nz;        /* grid; 2D uses nz = 1 */
    int dim;               /* 3: x,y periodic, z mirror; 2: x periodic, y mirror */
    int n_phase;           /* N = n_solid + 1; phase 0 is the liquid */
    int kmax;              /* KMAX, 1..64 (src/gpu_kernels_3d.py:20) */
    int precision;         /* MPF_FP32 or MPF_FP64 */
    int temp_dense;        /* 0: temperature is uniform in x,y (stored per z); 1: per cell */
    int temp_f32;          /* 1: the temperature is a float32 array (tests/verification_3d.py:817) */
    int enable_anisotropy; /* switchable kernel toggles (src/gpu_kernels_3d.py:692) */
    int enable_torque;
    int halt_on_kmax;      /* stop at the first step whose APT exceeds KMAX */
    int slot_capacity;     /* grains stored per cell; 0 = min(KMAX, N-1).  Steps never
                              store more than KMAX; larger values only admit uploads
                              with more non-zero grains per cell. */
    mpf_scalars s;
    double cooling_rate;   /* subtracted from the temperature before every update */
} mpf_config;

typedef struct mpf_stats {
    int max_active;                 /* largest active-phase count seen (status[1]) */
    int kmax_overflow;              /* 1 if some cell needed more than KMAX (status[0]) */
    unsigned long long overflow_cells;
    unsigned long long general_cells; /* cells that took the many-phase path */
    int halt_step;                  /* -1, or the step index that overflowed */
    int steps_done;                 /* total steps advanced by this handle */
    double last_step_ms;            /* GPU time of the last mpf_step call (hipEvents) */
    unsigned long long merge_cells; /* cells that needed the full APT merge (summed over steps) */
    unsigned long long front_cells; /* solid-liquid interface cells (summed over steps) */
    int class_tables;               /* 1: the tables have the class form of
                                       build_interaction_matrices (src/seed_modes.py:333-380)
                                       and the step selects their values instead of loading them */
    int kernel_steps;               /* steps timed per kernel (mpf_set_kernel_timing) */
    double kernel_ms[4];            /* summed GPU time of the step's kernels over those steps:
                                       classification (with its tile pass), merge, one-solid
                                       interface cells, many-phase (the other interface cells are
                                       not timed on their own) */
    unsigned long long light_cells; /* cells of the tile planes the classification pass computed,
                                       summed over steps; the other planes held liquid 1 with liquid
                                       1 around them and were left unchanged */
    int n111_pairs;                 /* 1: every grain's n111 rows come in opposite pairs (row 7-t = -row t,
                                       src/orientation_utils.py:182-212) and the facet search scans four */
} mpf_stats;

/* ---- resident solver ---------------------------------------------------- */

/* Replaces the to_device allocations of run_randommode_3d.py:244-252. */
int mpf_create(const mpf_config *cfg, mpf_handle **out);
void mpf_destroy(mpf_handle *h);

/* wij, aij, mij: [N][N]; n111: [N][8][3] (src/seed_modes.py:333-380,
 * src/orientation_utils.py:182-212).  Converted to the solver precision. */
int mpf_set_tables(mpf_handle *h, const double *wij, const double *aij, const double *mij,
                   const double *n111);

/* phi in the reference's dense layout [N][nx][ny][nz] (2D: [N][nx][ny]);
 * dtype MPF_FP32 or MPF_FP64 (run_randommode_3d.py:244). */
int mpf_upload_phi(mpf_handle *h, const void *phi, int dtype);

/* Columnar initial state without a dense array: cell (l,m,k) holds grain
 * grain_map[l][m] with value solid[profile[l][m]][k] and liquid
 * liquid[profile[l][m]][k] (src/seed_modes_3d.py:38-57, :117-166). */
int mpf_init_columns(mpf_handle *h, const int32_t *grain_map, const int32_t *profile, int nprofiles,
                     const double *solid, const double *liquid);

/* Temperature: [nz] values when temp_dense == 0, else [nx][ny][nz]. */
int mpf_upload_temp(mpf_handle *h, const double *temp);
int mpf_download_temp(mpf_handle *h, double *temp);

/* Advance nsteps steps.  Returns the number of steps completed (< nsteps
 * only when halt_on_kmax stopped at an overflowing step; the state is then the
 * one that step started from) or a negative error. */
int mpf_step(mpf_handle *h, int nsteps);

int mpf_download_phi(mpf_handle *h, void *phi, int dtype);

/* Checked APT of the current state in the reference layout
 * (kernel_update_nfmf_3d_checked, src/gpu_kernels_3d.py:489-527):
 * mf[depth][nx][ny][nz], nf[nx][ny][nz], status[2]. */
int mpf_download_apt(mpf_handle *h, int32_t *mf, int depth, int32_t *nf, int32_t *status);

int mpf_get_stats(mpf_handle *h, mpf_stats *st);
int mpf_reset_stats(mpf_handle *h);
/* ---- batches of independent 3D grids (no counterpart in the reference's
 * kernels; replaces the per-case loops of the bicrystal and orientation
 * sweeps, run_twomode_3d.py:200-322 and run_orientation_scan_3d.py:495-841,
 * which step one grid at a time).  The grids share the configuration's
 * physics, N and KMAX; each has its own size, tables and temperature.
 * mpf_step advances all of them together, one launch per kernel.  The
 * single-grid entry points above that move grid data (upload, download,
 * observables) refuse a batch handle; mpf_set_tables sets every grid.
 * dims: [n_inst][3] = nx, ny, nz.  cfg->nx/ny/nz are ignored; cfg->dim must
 * be 3, temp_dense and halt_on_kmax 0. */
int mpf_batch_create(const mpf_config *cfg, int n_inst, const int32_t *dims, mpf_handle **out);
int mpf_batch_size(mpf_handle *h); /* number of grids (1 for a single-grid handle) */
int mpf_batch_set_tables(mpf_handle *h, int inst, const double *wij, const double *aij, const double *mij,
                         const double *n111);
int mpf_batch_upload_phi(mpf_handle *h, int inst, const void *phi, int dtype);
int mpf_batch_init_columns(mpf_handle *h, int inst, const int32_t *grain_map, const int32_t *profile,
                           int nprofiles, const double *solid, const double *liquid);
int mpf_batch_upload_temp(mpf_handle *h, int inst, const double *temp); /* [nz] of that grid */
int mpf_batch_download_temp(mpf_handle *h, int inst, double *temp);
int mpf_batch_download_phi(mpf_handle *h, int inst, void *phi, int dtype);
int mpf_batch_download_apt(mpf_handle *h, int inst, int32_t *mf, int depth, int32_t *nf, int32_t *status);
int mpf_batch_observables(mpf_handle *h, int inst, double *grain_volume, int32_t *front, double *out);

/* Observables of every grid of a batch at once (also accepts a single-grid
 * handle): grain_volume[n*N] phase sums, front (optional) the topmost solid
 * cell of each column of each grid in turn (sum of nx*ny), out[n*7] as in
 * mpf_batch_observables.  Replaces the per-grid host reductions of the
 * sweeps (run_orientation_scan_3d.py:354-394, tests/verification_3d.py:612-654)
 * with one launch sequence. */
int mpf_batch_observables_all(mpf_handle *h, double *grain_volume, int32_t *front, double *out);

/* Ray probes of the orientation scan's directional radii
 * (run_orientation_scan_3d.py:236-331, trilinear_sample,
 * max_distance_to_boundary, directional_radius_one_side): ray r starts at
 * origin[3r..] (cells) in grid inst[r] (0 for a single grid), follows the
 * unit vector dir[3r..] in steps of step_cells, samples phase phase[r], and
 * out[r] is the distance (cells) to the first downward crossing of threshold
 * or NaN.  3D only; computed in double with the reference's rounding. */
int mpf_probe_rays(mpf_handle *h, int n_ray, const int32_t *inst, const int32_t *phase, const double *origin,
                   const double *dir, double threshold, double step_cells, double *out);

/* Dominant phase id of every cell of one plane, for the phase-map slices of
 * save_phase_map_slice_3d (src/plot_utils_3d.py:36-45: np.argmax over
 * phases, lowest id on ties).  axis 0: xy plane k = index, out[nx][ny];
 * axis 1: xz plane m = index, out[nx][nz]; axis 2: yz plane l = index,
 * out[ny][nz].  3D only. */
int mpf_phase_map(mpf_handle *h, int axis, int index, int32_t *out);
int mpf_batch_phase_map(mpf_handle *h, int grid, int axis, int index, int32_t *out);

/* Dense phi exchanged through device memory (no counterpart in the
 * reference; for device-side consumers such as the surrogate): the same
 * layouts and checks as mpf_upload_phi / mpf_download_phi and their batch
 * forms, with phi a device pointer. */
int mpf_upload_phi_device(mpf_handle *h, const void *phi, int dtype);
int mpf_download_phi_device(mpf_handle *h, void *phi, int dtype);
int mpf_batch_download_phi_device(mpf_handle *h, int grid, void *phi, int dtype);

/* ---- launch shape, graphs and autotuning (no counterpart in the reference,
 * which launches each kernel once per step from Python).
 * mpf_set_launch: blocks of the merge, interface and many-phase kernels (0
 * keeps the default for that kernel) and how the passes after the merge pass
 * run (side_stream): 0 one after another on the solver stream, 1 the
 * many-phase and multi-solid interface kernels on side streams concurrent
 * with the one-solid interface kernel (default from 4096 classification
 * tiles, 512x512x128 cells, up), 2 all three as one launch (default below).
 * mpf_get_launch: out[4] = the three block counts in use and that mode.
 * mpf_set_tiling: the classification pass's work items -- ranges of
 * planes_per_item z-planes (4, 8 or 16) of a 64 x 8 column tile (0: by grid
 * size, 16 from 4096 tiles up, 8 from 1024, else 4; 2D grids have one
 * plane) -- and its persistent grid (light_blocks, 0: 1024).
 * mpf_get_tiling: out[2] = the planes per item and blocks in use.
 * mpf_set_graphs: replay each step from a HIP graph captured on first use
 * (one per state parity); not with halt_on_kmax, and kernel timing bypasses
 * it.  mpf_autotune: times launch shapes on the current state (trial_steps
 * per shape, coordinate descent over that mode, the tiling and the three
 * block counts), keeps the fastest and restores the state; out (optional, 5
 * ints): the chosen mode and block counts and the time per step in
 * microseconds (the chosen tiling: mpf_get_tiling). */
int mpf_set_launch(mpf_handle *h, int merge_blocks, int front_blocks, int general_blocks, int side_stream);
int mpf_get_launch(mpf_handle *h, int *out);
int mpf_set_tiling(mpf_handle *h, int planes_per_item, int light_blocks);
int mpf_get_tiling(mpf_handle *h, int *out);
int mpf_set_graphs(mpf_handle *h, int on);
int mpf_autotune(mpf_handle *h, int trial_steps, int *out);

/* Time every kernel of every step with HIP events (on: 1) and sum the times
 * into mpf_stats.kernel_ms.  Adds a few microseconds per step.  No
 * counterpart in the reference (it has no timing). */
int mpf_set_kernel_timing(mpf_handle *h, int on);

/* FP64 observables of the current state (tests/verification_3d.py:409-416,
 * :612-654): grain_volume[N] = sum of phi_g over cells (times dx^3 by the
 * caller), front[nx*ny] = topmost k with 1 - phi_L > 0.5 (or -1), and
 * out[0..6] = max|sum phi - 1|, min phi, max phi, NaN count, solid sum, cells,
 * stored grain slots. */
int mpf_observables(mpf_handle *h, double *grain_volume, int32_t *front, double *out);

/* Difference of the current states of two handles on the same grid and phase
 * count (any precisions), reduced on the device in FP64 with grain slots
 * matched by id: out[0] = max over cells and phases of |phi_a - phi_b|,
 * out[1] = their sum, out[2] = cells whose sets of non-zero grains differ,
 * out[3] = cells with a difference above tol.  States the FP32 solver's error
 * against the FP64 one (run_randommode_3d.py:267-300 advances both) without
 * copying the dense phi to the host.  No counterpart in the reference. */
int mpf_compare_phi(mpf_handle *a, mpf_handle *b, double tol, double *out);

/* Bytes of device memory held by the handle. */
size_t mpf_device_bytes(mpf_handle *h);

/* ---- drop-in kernels on the reference's dense layout ---------------------
 * All pointers are device pointers; stream may be NULL (default stream). */

/* kernel_update_nfmf_3d (checked = 0, src/gpu_kernels_3d.py:462-486) and
 * kernel_update_nfmf_3d_checked (checked = 1, :489-527); dim 2 is
 * kernel_update_nfmf (src/gpu_kernels.py:337-356).  phi dtype MPF_FP32/64. */
int mpf_dense_apt(const void *phi, int dtype, int32_t *mf, int32_t *nf, int32_t *status, int n_phase,
                  int nx, int ny, int nz, int dim, int depth, int kmax, int checked, void *stream);

/* kernel_update_phasefield_active_3d(_switchable) (src/gpu_kernels_3d.py:530-830)
 * and kernel_update_phasefield_active (src/gpu_kernels.py:361-515).  phi,
 * phi_new and the tables share dtype; temp has temp_dtype; mf has depth rows.
 * The tables are gathered into a block allocated and freed in stream order
 * (hipMallocAsync/hipFreeAsync); the call returns without waiting for the
 * kernel, as the reference's launch does. */
int mpf_dense_update(const void *phi, void *phi_new, int dtype, const void *temp, int temp_dtype,
                     const int32_t *mf, const int32_t *nf, const void *wij, const void *aij,
                     const void *mij, const void *n111, const mpf_scalars *s, int n_phase, int nx, int ny,
                     int nz, int dim, int depth, int kmax, int enable_anisotropy, int enable_torque,
                     void *stream);

/* mpf_dense_update without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  The same update and arguments as mpf_dense_update;
 * the tables are gathered into a file-scope device array of the library (one
 * per GPU, held until the process ends, 1.6 MB: n_phase <= 256, larger values
 * are rejected with MPF_EINVAL).  Nothing is allocated or freed and the call
 * does not wait for the kernel.  Calls that overlap on one GPU race on that
 * array. */
int mpf_dense_update_noalloc(const void *phi, void *phi_new, int dtype, const void *temp, int temp_dtype,
                             const int32_t *mf, const int32_t *nf, const void *wij, const void *aij,
                             const void *mij, const void *n111, const mpf_scalars *s, int n_phase, int nx,
                             int ny, int nz, int dim, int depth, int kmax, int enable_anisotropy,
                             int enable_torque, void *stream);

/* kernel_update_temp_3d (src/gpu_kernels_3d.py:833-838) / kernel_update_temp
 * (src/gpu_kernels.py:520-525): temp[i] -= cooling_rate over n values. */
int mpf_dense_temp(void *temp, int temp_dtype, double cooling_rate, size_t n, void *stream);

/* ---- device memory for callers without another allocator ---------------- */
int mpf_malloc(void **ptr, size_t bytes);
int mpf_free(void *ptr);
int mpf_memcpy_htod(void *dst, const void *src, size_t bytes);
int mpf_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int mpf_memcpy_dtod(void *dst, const void *src, size_t bytes);
int mpf_memset(void *ptr, int value, size_t bytes);
int mpf_device_synchronize(void);

const char *mpf_last_error(void);
int mpf_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MPF_H */
