/*
 * boostfractor_hip.h -- C ABI of the MI355X (gfx950) engine for BoostFractor.jl's
 * two propagation solvers.
 *
 * The reference is the Julia package BoostFractor.jl; every entry point below
 * replaces one of its Julia functions and is what its `ccall` binding would
 * target.  Conventions shared by all entry points:
 *
 *   - complex numbers are interleaved (re, im) doubles, binary compatible with
 *     Julia's Complex{Float64};
 *   - a field is [ny][nx] with x fastest, i.e. exactly Julia's column-major
 *     E[x, y] (Array{ComplexF64,2}(undef, Nx, Ny));
 *   - stacked objects put their stack index slowest: surfaces [nreg][ny][nx],
 *     dancer fields [nreg-1][2][ny][nx], mode patterns [n][ncomp][ny][nx] with
 *     mode index (m-1)(2L+1) + l + L, n x n matrices column-major;
 *   - all buffers are caller-owned host memory; nothing is retained after a
 *     call returns (plans below keep their own device copies);
 *   - functions return a status (BF_OK or BF_ERR_*); the message of the last
 *     failure is bf_last_error(ctx) (bf_last_error(NULL) for failures raised
 *     before a context was available).  Argument and dimension checks run on
 *     the host before any device work.
 *
 * A context owns one HIP stream on one device.  Contexts are independent: use
 * one per concurrent caller thread.  Calls on one context are serialised by
 * the caller.
 */
#ifndef BOOSTFRACTOR_HIP_H
#define BOOSTFRACTOR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define BF_OK 0
#define BF_ERR_DIM 1         /* Julia DimensionMismatch (src/boost_fractor.jl:105) */
#define BF_ERR_ARG 2         /* Julia ArgumentError / error(...) */
#define BF_ERR_UNSUPPORTED 3 /* request outside the engine's limits (grid > 1024, n > 128, ...) */
#define BF_ERR_HIP 4         /* HIP runtime failure */
#define BF_ERR_NOMEM 5       /* device allocation failed */

/* the reference's propagator plug-ins (src/boost_fractor.jl:137,157,187,218) */
#define BF_PROP_PROPAGATOR 0 /* propagator: mask, FFT, kz phase, IFFT, tilt, surface */
#define BF_PROP_NOTILTS 1    /* propagatorNoTilts */
#define BF_PROP_1D 2         /* propagator1D: scalar phase only */
#define BF_PROP_MOMENTUM 3   /* propagatorMomentumSpace: field held in (fftshifted) k-space */

/* arithmetic of the hot path */
#define BF_PREC_FP64 0  /* everything in FP64 */
#define BF_PREC_MIXED 1 /* fields stored in FP32 in HBM, in-register FFT/phase/update arithmetic in FP64,
                           FP64 outputs and transfer-matrix chains */
#define BF_PREC_FP32 2  /* FP32 storage and FP32 arithmetic (accuracy study; not the mixed path) */

#define BF_MAX_GRID 1024
#define BF_MAX_MODES 128

typedef struct bf_ctx bf_ctx;

/* CoordinateSystem (src/boost_fractor.jl:21-28).  kX/kY are passed in, not
 * recomputed, so that the engine uses the exact doubles the caller's grid holds. */
typedef struct bf_coords {
    int32_t nx, ny;
    const double *X, *Y;   /* [nx], [ny] */
    const double *kX, *kY; /* [nx], [ny] */
} bf_coords;

/* SetupBoundaries (src/boost_fractor.jl:63-71), mirror region first. */
typedef struct bf_setup {
    int32_t nreg;                  /* number of regions = length(distance) */
    const double *distance;        /* [nreg] */
    const double *r;               /* complex [nreg-1] boundary reflectivities */
    const double *eps;             /* complex [nreg]; real values carry imag -0.0 */
    const double *tilt_x, *tilt_y; /* [nreg] */
    const double *surfaces;        /* complex [nreg][ny][nx], or NULL for flat */
} bf_setup;

/* Modes (src/boost_fractor_transformer.jl:31-38) */
typedef struct bf_modes {
    int32_t n;              /* number of modes M*(2L+1) */
    int32_t ncomp;          /* field components per pattern (1: scalar field) */
    const double *patterns; /* complex [n][ncomp][ny][nx] */
    const double *mode_kt;  /* complex [n] transverse wavenumbers */
} bf_modes;

/* ---- context ----------------------------------------------------------- */
int bf_version(void);
const char *bf_last_error(const bf_ctx *ctx);
int bf_device_count(int32_t *count);
int bf_ctx_create(int32_t device, bf_ctx **out);
int bf_ctx_destroy(bf_ctx *ctx);
/* HIP stream of the context (hipStream_t), for callers that time or order work on it */
int bf_ctx_stream(bf_ctx *ctx, void **stream);
int bf_ctx_synchronize(bf_ctx *ctx);
/* Per-kernel-class timing on the context's stream (no reference counterpart;
 * for measurement): enable != 0 clears the counters and starts recording HIP
 * events around the launches of each class, 0 stops.  bf_ctx_kernel_time
 * returns the accumulated milliseconds, launch count and algorithmic flops of
 * one class: "wgram" (propagation-matrix Gram products, 8 n^2 K flop per
 * region and frequency), "inverse" (batched LU inverses), "chain"
 * (transfer-matrix chain and solves, 64 n^3 flop per region and frequency). */
int bf_ctx_timing(bf_ctx *ctx, int32_t enable);
int bf_ctx_kernel_time(bf_ctx *ctx, const char *kernel, double *ms, int64_t *launches, double *flops);
/* Engine options of a context (no reference counterpart; the nearest is the
 * planner flags FFTW's plan_fft takes, src/boost_fractor.jl:163-164).  The
 * defaults are the product configuration; the library never reads the process
 * environment.  Switches (0/1): "fft_lds" (LDS FFT passes instead of the
 * register-resident ones), "prune" (band-pruned transforms), "qtab" (quarter
 * k-space tables for complex64 fields), "band" (skip the rows the aperture
 * zeroes), "chfuse", "chxf", "chwave" (cheerleader pair fusion, stored row
 * transforms, wavefront order), "graph" (dancer plans replay HIP graphs),
 * "parseval_groups", "disk_expansion", "parseval_series" (transformer work
 * reductions: k-bin groups, disks from the Gram matrix, air gaps by a series
 * in the spacing), "series_inverse" (the gaps' inverses by the inverse's
 * series in flat sweep chunks; a run whose series fails its convergence check
 * is redone with LU inverses, which 2 forces), "chain_tile" (register-tiled
 * chain kernels for n = 9, 15, 25, 45), "chain_mfma" (chain on the matrix
 * cores: 1 where it pays, 4 for every n <= 32, 2 / 3 for every n <= 32 in
 * the earlier one-wave-per-tile form with four / three MFMAs per complex
 * product), "chain_order" (chain blocks of a batch start with the longest
 * chains; 0 keeps job order), "pgemm_herm" (the Parseval product over the
 * upper triangles when the mode set's patterns vanish outside the aperture,
 * so that its group Grams are Hermitian; 0 forms every entry), "sinv_mfma"
 * (the inverse series on the matrix cores for 8 < n <= 32; 0: FP64 VALU),
 * "autotune" (below).  Shapes and budgets (0 = default): "col_tx", "rows"
 * (dancer launch shapes), "pg_wn" (2 or 4: Parseval product column blocks of
 * 128 or 256 matrix entries), "batch_bytes" (bytes of P and inv(P) per
 * chunk of a batched sweep; default 16 GiB).  Unknown names and negative
 * values return BF_ERR_ARG.  No option changes a result beyond rounding. */
int bf_ctx_set_option(bf_ctx *ctx, const char *name, int64_t value);
int bf_ctx_get_option(bf_ctx *ctx, const char *name, int64_t *value);
/* Autotuner (option "autotune" = 1).  The reference's FFTs
 * (src/boost_fractor.jl:163-164, FFTW.fft!) are planned by FFTW's estimate on
 * every call; this is the engine's counterpart of a measured plan: on the first use of a
 * shape key the engine times candidate launch shapes on the call's own inputs
 * and keeps the fastest in the context: dancer plans (frequency batch,
 * column-pass tile, row-pass rows; short runs with nmax cut to 20) and the
 * Parseval product of the mode-matching sweep (column-block width).  Explicit
 * options (col_tx, rows, pg_wn, a plan's freq_batch) take precedence.
 * bf_ctx_set_tuning_file loads entries from a text file (key, tab, values) and
 * appends every new winner to it; bf_ctx_tuning_report writes the context's
 * entries as such lines into buf (len bytes) and the milliseconds spent tuning. */
int bf_ctx_set_tuning_file(bf_ctx *ctx, const char *path);
int bf_ctx_tuning_report(bf_ctx *ctx, char *buf, int64_t len, double *tune_ms);

/* ---- propagators --------------------------------------------------------
 * replaces prop(E0, dz, diskR, eps, tilt_x, tilt_y, surface, lambda, coords)
 *   propagator              src/boost_fractor.jl:137-149
 *   propagatorNoTilts       src/boost_fractor.jl:157-179
 *   propagatorMomentumSpace src/boost_fractor.jl:187-210
 *   propagator1D            src/boost_fractor.jl:218-227
 * E: complex [ny][nx], replaced by the propagated field.  surface: complex
 * [ny][nx] or NULL (flat).  eps is complex; propagator (kind 0) casts a real eps
 * to re-0.0im as the reference does, the other kinds use it as given. */
int bf_propagate(bf_ctx *ctx, int32_t prop_kind, const bf_coords *coords, double *E, double dz, double diskR,
                 double eps_re, double eps_im, double tilt_x, double tilt_y, const double *surface, double lambda);

/* ---- recursive Fourier propagation ----------------------------------- */
/* dance_intro (src/boost_fractor_recurprop.jl:153-201).  nd = length(bdry.distance),
 * eps complex [nd], r1 = bdry.r[1].  bfield: complex [nd+1][ny][nx], or NULL for the
 * default real mask (1 on x^2 + y^2 <= diskR^2).  velocity_x != 0 multiplies bfield in
 * place by exp(-1im*Ma*(-velocity_x)*x), Ma = 2 pi f/c (:162-166), as the reference does
 * to the caller's array; it needs a bfield (with the default real mask the reference
 * throws InexactError: BF_ERR_ARG here).  out: complex [nd][2][ny][nx]. */
int bf_dance_intro(bf_ctx *ctx, const bf_coords *coords, int32_t nd, const double *eps, double r1_re,
                   double r1_im, double diskR, double *bfield, double velocity_x, double f, double *out);

/* gauss_profile (src/beam_gaussian.jl:24-37): normalised Gaussian beam at distance z
 * from its waist omega0, frequency f.  out: complex [ny][nx]. */
int bf_gauss_profile(bf_ctx *ctx, const bf_coords *coords, double z, double omega0, double f, double *out);

/* unnormalized_match (src/beam_coupling.jl:13-15): out[b] = sum(Eout[b] .* conj(Eantenna)) * dx * dy
 * for nfield fields Eout complex [nfield][n] against one Eantenna complex [n]; out complex [nfield]. */
int bf_unnormalized_match(bf_ctx *ctx, int32_t nfield, int64_t n, const double *Eout, const double *Eantenna,
                          double dx, double dy, double *out);

/* dancer (src/boost_fractor_recurprop.jl:25-146) for nfreq frequencies of one
 * setup.  setup is the untrimmed SetupBoundaries (mirror region first).
 * emit: complex [nreg-1][2][ny][nx] or NULL (axion source from dance_intro);
 * reflect: complex [ny][nx] or NULL.  Output Eout:
 *   returnsum != 0           -> [nfreq][ny][nx]
 *   returnsum == 0           -> [nfreq][S][ny][nx], S = immediatesum ? 1 : nmax+1
 * power (optional, [nfreq]): sum |Eout|^2 over the grid of the summed output.
 * bounces (optional, int32 [nfreq]): iterations executed (nmax+1 unless amin stops early). */
int bf_dancer(bf_ctx *ctx, const bf_setup *setup, const bf_coords *coords, double amin, int32_t nmax,
              const double *freqs, int32_t nfreq, int32_t prop_kind, const double *emit, const double *reflect,
              double diskR, int32_t returnsum, int32_t immediatesum, int32_t precision, double *Eout,
              double *power, int32_t *bounces);
/* dancer over many geometries at the same frequencies (the reference sweeps them
 * point by point: test/runtests.jl:101-114, examples/Single Disk.ipynb); setups may
 * differ in disk count.  emits / reflects: NULL or one pointer per geometry (each as
 * bf_dancer's emit / reflect, may be NULL).  Geometry c's output, in bf_dancer's
 * Eout layout for that geometry, starts at complex element eout_offset[c] of Eout;
 * power and bounces (optional) are [nconf][nfreq].  Results equal bf_dancer's per
 * geometry. */
int bf_dancer_batch(bf_ctx *ctx, int32_t nconf, const bf_setup *setups, const bf_coords *coords, double amin,
                    int32_t nmax, const double *freqs, int32_t nfreq, int32_t prop_kind,
                    const double *const *emits, const double *const *reflects, double diskR, int32_t returnsum,
                    int32_t immediatesum, int32_t precision, const int64_t *eout_offset, double *Eout,
                    double *power, int32_t *bounces);

/* cheerleader (src/boost_fractor_recurprop.jl:226-363) for nfreq frequencies.
 * emit: complex [nreg][2][ny][nx] or NULL; reflect: complex [ny][nx] or NULL.
 * out_right: complex [nfreq][ny][nx] (fields leaving on the right);
 * out_left: complex [nfreq][ny][nx] or NULL (returnboth=true). */
int bf_cheerleader(bf_ctx *ctx, const bf_setup *setup, const bf_coords *coords, double amin, int32_t nmax,
                   const double *freqs, int32_t nfreq, int32_t prop_kind, const double *emit,
                   const double *reflect, double diskR, int32_t precision, double *out_right, double *out_left);

/* Resident dancer sweep: device buffers and phase tables are built once; run
 * executes all frequencies with inputs already in HBM; fetch copies results.
 * Same semantics as bf_dancer with emit = reflect = NULL, returnsum = immediatesum = 1. */
typedef struct bf_dancer_plan bf_dancer_plan;
int bf_dancer_plan_create(bf_ctx *ctx, const bf_setup *setup, const bf_coords *coords, double amin, int32_t nmax,
                          const double *freqs, int32_t nfreq, int32_t prop_kind, double diskR, int32_t precision,
                          int32_t freq_batch, bf_dancer_plan **out);
int bf_dancer_plan_run(bf_dancer_plan *plan);
int bf_dancer_plan_fetch(bf_dancer_plan *plan, double *Eout, double *power);
/* Per-kernel timing over the next runs (HIP events on the plan's stream):
 * enable != 0 starts collecting; query returns accumulated milliseconds and
 * launch counts for the row pass (index 0) and the column pass (index 1). */
int bf_dancer_plan_timing(bf_dancer_plan *plan, int32_t enable);
int bf_dancer_plan_kernel_times(bf_dancer_plan *plan, double *ms, int64_t *launches);
/* Plan layout chosen at create time (no reference counterpart; for tuning and
 * reports): frequencies per batch, and the row band [band_lo, band_hi) the
 * fused bounce processes at every boundary (rows outside it are zeroed by the
 * aperture and only feed Eout at the exit boundary). */
int bf_dancer_plan_info(bf_dancer_plan *plan, int32_t *freq_batch, int32_t *band_lo, int32_t *band_hi);
int bf_dancer_plan_destroy(bf_dancer_plan *plan);

/* Stream-bandwidth probe: copies `bytes` between two device buffers `reps`
 * times with the engine's copy kernel (k_stream_copy) and returns the mean
 * milliseconds per copy.  Used to state measured HBM bandwidth next to the
 * roofline fractions and to calibrate profiler byte counters. */
int bf_stream_copy_probe(bf_ctx *ctx, int64_t bytes, int32_t reps, double *ms_per_copy);

/* ---- mode matching ----------------------------------------------------- */
/* SeedModes Bessel patterns (src/boost_fractor_transformer.jl:46-107, mode :88):
 * J_l(j_{|l|,m} r / R) exp(-i l phi), zero for r > R, normalised to sum |.|^2 = 1.
 * patterns: complex [M(2L+1)][ny][nx]; mode_kt: complex [M(2L+1)]. */
int bf_seed_modes(bf_ctx *ctx, const bf_coords *coords, int32_t M, int32_t L, double diskR, double *patterns,
                  double *mode_kt);

/* axion_induced_modes / field2modes (src/boost_fractor_transformer.jl:115-156):
 * B: complex [ncomp][ny][nx] or NULL (uniform); masked to r <= diskR,
 * normalised, projected on the modes.  coeffs: complex [n]. */
int bf_field2modes(bf_ctx *ctx, const bf_coords *coords, const bf_modes *modes, const double *B,
                   double velocity_x, double diskR, double f, double *coeffs);

/* modes2field (src/boost_fractor_transformer.jl:161-167); field complex [ncomp][ny][nx] */
int bf_modes2field(bf_ctx *ctx, const bf_coords *coords, const bf_modes *modes, const double *coeffs,
                   double *field);

/* calc_propagation_matrices (src/boost_fractor_transformer.jl:274-280) -> propagation_matrix
 * (:174-227) for every region of the setup at each frequency.
 * out: complex [nfreq][nreg][n][n] (column-major n x n). */
int bf_propagation_matrices(bf_ctx *ctx, const bf_setup *setup, const bf_coords *coords, const bf_modes *modes,
                            const double *freqs, int32_t nfreq, int32_t prop_kind, double diskR,
                            int32_t precision, double *out);

/* transformer (src/boost_fractor_transformer.jl:285-359) at nfreq frequencies.
 * NaN permittivities are treated as 1e30 (the reference mutates bdry.eps so).
 * prop_mats: complex [nfreq][nreg][n][n] or NULL (computed on the device).
 * emit: complex [n]; reflect: complex [n] or NULL.
 * boost: complex [nfreq][n]; refl: complex [nfreq][n] (only if reflect != NULL). */
int bf_transformer(bf_ctx *ctx, const bf_setup *setup, const bf_coords *coords, const bf_modes *modes,
                   const double *freqs, int32_t nfreq, int32_t prop_kind, const double *prop_mats, double diskR,
                   const double *emit, const double *reflect, int32_t precision, double *boost, double *refl);

/* A resident transformer sweep (no reference counterpart; the callers' frequency
 * sweeps over one geometry, e.g. examples/transformer_optim_utilities.jl:99-109):
 * create uploads the setup, the modes and emit/reflect once; every run
 * recomputes transformer (src/boost_fractor_transformer.jl:285-359) at the
 * plan's frequencies from them on the device; fetch copies boost (and refl)
 * out, complex [nfreq][n].  NaN permittivities are read as 1e30; the caller's
 * setup is not modified. */
typedef struct bf_transformer_plan bf_transformer_plan;
int bf_transformer_plan_create(bf_ctx *ctx, const bf_setup *setup, const bf_coords *coords, const bf_modes *modes,
                               const double *freqs, int32_t nfreq, int32_t prop_kind, double diskR,
                               const double *emit, const double *reflect, bf_transformer_plan **out);
int bf_transformer_plan_run(bf_transformer_plan *plan);
int bf_transformer_plan_fetch(bf_transformer_plan *plan, double *boost, double *refl);
int bf_transformer_plan_destroy(bf_transformer_plan *plan);

/* transformer (src/boost_fractor_transformer.jl:285-359) for a batch of geometries
 * at the same nfreq frequencies: the disk-position sweep of
 * examples/transformer_optim_utilities.jl:105,121,202 (calc_propagation_matrices
 * then transformer(...; propagation_matrices=...) per candidate geometry) in one
 * call.  Geometry c uses mode set mode_set_of[c] (0 <= . < nsets) with axion
 * source emits[mode_set_of[c]] (complex [n of that set]); setups may differ in
 * region count, spacing, tilt and surfaces.  Propagation matrices are computed
 * on the device (prop_kind); precision as for bf_transformer.
 * boost: complex, geometry c's [nfreq][n] block starts at complex element
 * boost_offset[c].  No reflect/refl outputs (the sweep uses the boost only). */
int bf_transformer_batch(bf_ctx *ctx, int32_t nconf, const bf_setup *setups, const bf_coords *coords,
                         int32_t nsets, const bf_modes *modesets, const double *const *emits,
                         const int32_t *mode_set_of, const double *freqs, int32_t nfreq, int32_t prop_kind,
                         double diskR, int32_t precision, const int64_t *boost_offset, double *boost);
/* bf_transformer_batch with the grid, the mode sets, their axion sources and the
 * frequencies kept on the device (the same sweep, examples/
 * transformer_optim_utilities.jl:105,121,202, for an optimiser that evaluates
 * many candidate geometries against fixed modes): create uploads them once;
 * each run takes the candidate setups (and their mode set and output offset, as
 * bf_transformer_batch) from the host and computes every boost vector on the
 * device, recomputing the mode sets' pattern FFTs and Gram tables it needs;
 * fetch copies the last run's geometry blocks to the same offsets of `boost`
 * (elements between blocks are not written).  Results equal
 * bf_transformer_batch's. */
typedef struct bf_transformer_batch_plan bf_transformer_batch_plan;
int bf_transformer_batch_plan_create(bf_ctx *ctx, const bf_coords *coords, int32_t nsets, const bf_modes *modesets,
                                     const double *const *emits, const double *freqs, int32_t nfreq,
                                     int32_t prop_kind, double diskR, int32_t precision,
                                     bf_transformer_batch_plan **out);
/* bf_transformer_batch_plan_create with a grid per mode set (no reference
 * counterpart: the reference sweeps one grid at a time): mode set m and the
 * geometries run on it (their surface maps) are sampled on coords[m], so one
 * plan, and one run, covers geometries on grids of different sizes.  coords:
 * [nsets]; the rest as bf_transformer_batch_plan_create.  Each geometry's
 * results equal those of a plan on its set's grid alone. */
int bf_transformer_batch_plan_create_grids(bf_ctx *ctx, int32_t nsets, const bf_coords *coords,
                                           const bf_modes *modesets, const double *const *emits, const double *freqs,
                                           int32_t nfreq, int32_t prop_kind, double diskR, int32_t precision,
                                           bf_transformer_batch_plan **out);
int bf_transformer_batch_plan_run(bf_transformer_batch_plan *plan, int32_t nconf, const bf_setup *setups,
                                  const int32_t *mode_set_of, const int64_t *boost_offset);
int bf_transformer_batch_plan_fetch(bf_transformer_batch_plan *plan, double *boost);
int bf_transformer_batch_plan_destroy(bf_transformer_batch_plan *plan);

/* ---- 1D analytical model (src/Analytical1D.py) ------------------------- */
/* disk_system (src/Analytical1D.py:82-131), ray-tracing form, for ncand geometries
 * at nfreq frequencies in one launch: mirror (eps 1e20, 0.1 m) when `mirror`, then
 * gap 0, (disk, gap i+1) x num_disk; spacings: [ncand][num_disk+1] (m); eps_disk,
 * eps_gap: complex (as lossy_eps gives them); surfaceloss: [2 num_disk + 1] or NULL
 * (non_uniform_surfaceloss).  gamma (amplitude reflectivity) and axion (amplitude
 * boost factor): complex [ncand][nfreq]. */
int bf_disk_system_1d(bf_ctx *ctx, const double *freqs, int32_t nfreq, int32_t ncand, int32_t num_disk,
                      const double *spacings, double disk_thickness, const double *eps_disk, const double *eps_gap,
                      int32_t mirror, const double *surfaceloss, double *gamma, double *axion);
/* transform_surface (src/Analytical1D.py:5-67) over n frequency points: gamma_in,
 * gamma_out complex [n]; axion complex [n] updated in place, or NULL; eps_i, eps_m
 * complex; eps_m_real != 0 when the caller's eps_m is a real number (the
 * propagation phase is then formed in real arithmetic, as numpy does); use_ray 0
 * selects the impedance transformation, which (as in the reference) cannot carry
 * an axion field: BF_ERR_ARG with axion != NULL. */
int bf_transform_surface_1d(bf_ctx *ctx, const double *freqs, int32_t n, const double *gamma_in, double *axion,
                            const double *eps_i, const double *eps_m, int32_t eps_m_real, double l,
                            double surfaceloss, int32_t use_ray, double *gamma_out);

/* ---- disk-position sweep (examples/transformer_optim_utilities.jl) ------ */
/* calc_propagation_matrices_grid (:11-52): the setup split into its gaps (0-based
 * regions 1, 3, ..., nreg-1) and solids (0, 2, ..., nreg-2; nreg even); every gap
 * distance shifted by spacing_grid[s] and every gap tilt by tilt_x_grid[tx],
 * tilt_y_grid[ty]; solids computed once per frequency.
 * out: complex [nspacing][ntilt_x][ntilt_y][nfreq][nreg][n][n] (column-major n x n). */
int bf_propagation_matrices_grid(bf_ctx *ctx, const bf_setup *setup, const bf_coords *coords,
                                 const bf_modes *modes, const double *spacing_grid, int32_t nspacing,
                                 const double *tilt_x_grid, int32_t ntilt_x, const double *tilt_y_grid,
                                 int32_t ntilt_y, const double *freqs, int32_t nfreq, int32_t prop_kind,
                                 double diskR, double *out);

/* construct_prop_matrix_interpolation (:58-74): cubic B-spline (Interpolations.jl
 * BSpline(Cubic(Natural(OnCell())))) over a uniform spacing grid spacing_first +
 * k*spacing_step, k < nspacing (>= 2), kept on the device.
 * grid: complex [nspacing][nfreq][nreg][n][n] (the tilt-index-(1,1) slice of
 * bf_propagation_matrices_grid's output). */
typedef struct bf_spacing_itp bf_spacing_itp;
int bf_spacing_itp_create(bf_ctx *ctx, const double *grid, int32_t nreg, int32_t nfreq, int32_t nspacing,
                          int32_t n, double spacing_first, double spacing_step, bf_spacing_itp **out);
/* Interpolations.bounds along the spacing axis: first - step/2, last + step/2 */
int bf_spacing_itp_bounds(const bf_spacing_itp *itp, double *lo, double *hi);
/* interpolate_prop_matrix (:79-91) for ncand candidates; dist_shift: [ncand][(nreg-2)/2]
 * (one shift per gap except the last; solids and the last gap are evaluated at 0).
 * A shift outside the bounds is an error (BF_ERR_ARG, Interpolations' BoundsError).
 * out: complex [ncand][nfreq][nreg][n][n]. */
int bf_spacing_itp_matrices(bf_spacing_itp *itp, int32_t ncand, const double *dist_shift, double *out);
/* calc_boostfactor_modes / calc_modes (:99-125) on the interpolated matrices of
 * ncand candidates: transformer (src/boost_fractor_transformer.jl:285-359) with
 * propagation_matrices given.  eps: complex [nreg] (NaN read as 1e30); emit:
 * complex [n]; reflect: complex [n] or NULL.  boost, refl: complex [ncand][nfreq][n]. */
int bf_spacing_itp_transformer(bf_spacing_itp *itp, int32_t ncand, const double *dist_shift, const double *eps,
                               const double *emit, const double *reflect, double *boost, double *refl);
int bf_spacing_itp_destroy(bf_spacing_itp *itp);

/* transformer_trace_back (src/boost_fractor_transformer.jl:372-421) at one frequency.
 * reflected, input: complex [n]; emit: complex [n] (used when includes_axion);
 * prop_mats: complex [nreg][n][n] or NULL.  fields: complex [nreg][2][n]. */
int bf_transformer_trace_back(bf_ctx *ctx, const double *reflected, const double *input, const bf_setup *setup,
                              const bf_coords *coords, const bf_modes *modes, double f, int32_t prop_kind,
                              const double *prop_mats, double diskR, int32_t includes_axion, const double *emit,
                              double *fields);

/* ---- allocation-free chain --------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops such as an
 * optimiser's candidate evaluations, harnesses that forbid hipMalloc): the same
 * computation as bf_transformer with precomputed propagation matrices
 * (transformer(...; propagation_matrices=P), src/boost_fractor_transformer.jl:
 * 285-359, run by the same kernels), with every input and output already in
 * device memory.  It calls no hipMalloc/hipFree: its working memory (the
 * inverses, selection list and chain descriptors) is a fixed array of
 * BF_NOALLOC_SCRATCH_BYTES per GPU that the library holds until the process
 * ends.  Every call on one GPU shares that array: calls that overlap on one GPU
 * (several contexts or threads) race; calls on one context are serialised.
 *   eps: HOST complex [nreg] (NaN taken as 1e30, as the reference's mutation);
 *   prop_mats: DEVICE complex [nfreq][nreg][n][n] column-major (region 0 unused);
 *   emit: DEVICE complex [n]; reflect: DEVICE complex [n] or NULL;
 *   boost: DEVICE complex [nfreq][n]; refl: DEVICE complex [nfreq][n] (with reflect).
 * Limits: 1 <= n <= 48, 2 <= nreg <= BF_NOALLOC_MAX_REGIONS, and the call's
 * working memory nfreq * nreg * n^2 * 16 bytes (plus descriptors) within
 * BF_NOALLOC_SCRATCH_BYTES; larger calls return BF_ERR_UNSUPPORTED before any
 * device work.  Returns after the results are written (the context's stream is
 * synchronised). */
#define BF_NOALLOC_SCRATCH_BYTES (32 << 20)
#define BF_NOALLOC_MAX_REGIONS 512
int bf_transformer_chain_noalloc(bf_ctx *ctx, int32_t nreg, const double *eps, int32_t n, int32_t nfreq,
                                 const double *prop_mats, const double *emit, const double *reflect, double *boost,
                                 double *refl);

#ifdef __cplusplus
}
#endif

#endif /* BOOSTFRACTOR_HIP_H */
