/*
 * diffsims_amd — C ABI of the MI355X (gfx950) kinematical diffraction path.
 *
 * Every pointer argument named *_d is a device pointer (HIP global memory,
 * typically a torch tensor's data_ptr()); every other pointer is host memory.
 * `stream` is a hipStream_t passed as void* (NULL = the default stream).  All
 * launches are asynchronous on that stream; the library never allocates, frees
 * or synchronises unless an entry point says so.  Functions return 0 on
 * success or a negative DSX_E* code; dsx_last_error() then describes the
 * failure (thread-local string).
 *
 * The reference (diffsims 0.7.0) is pure Python; each entry point names the
 * reference function whose work it replaces.  A Python caller binds them with
 * ctypes (see diffsims_amd/_lib.py).
 */
#ifndef DIFFSIMS_AMD_H
#define DIFFSIMS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSX_OK 0
#define DSX_EINVAL (-1)    /* bad argument (shape, size, enum)            */
#define DSX_ELAUNCH (-2)   /* HIP launch or runtime error                 */
#define DSX_ECAPACITY (-3) /* caller buffer too small (see the entry)     */

/* Shape-factor models (diffsims/utils/shape_factor_models.py). */
#define DSX_SF_LINEAR 0     /* :52-73                                      */
#define DSX_SF_LORENTZIAN 1 /* :154-180                                    */
#define DSX_SF_SINC 2       /* :76-101                                     */
#define DSX_SF_SIN2C 3      /* :104-123                                    */
#define DSX_SF_ATANC 4      /* :126-151                                    */
#define DSX_SF_BINARY 5     /* :33-49                                      */
/* precession forms (simulation_generator.py:396-411) */
#define DSX_SF_LORENTZIAN_PRECESSION 6 /* shape_factor_models.py:183-219   */
/* Full precession average of model m (approximate_precession=False,
 * _shape_factor_precession, shape_factor_models.py:222-268): DSX_SF_PRECESSION_AVG + m
 * for m in DSX_SF_LINEAR..DSX_SF_BINARY; needs precession_angle > 0. */
#define DSX_SF_PRECESSION_AVG 16

/* Scattering-factor parameterisations (sim_utils.py:227-253, :281-286). */
#define DSX_SCATTER_NONE 0    /* f = 1                                     */
#define DSX_SCATTER_LOBATO 1  /* sum a (2 + b g^2) / (1 + b g^2)^2          */
#define DSX_SCATTER_XTABLES 2 /* sum a exp(-b g^2 / 4)                      */

/* Library version (major*10000 + minor*100 + patch). */
int dsx_version(void);

/* Description of the last failure on the calling thread ("" if none). */
const char *dsx_last_error(void);

/*
 * hkl enumeration — replaces ReciprocalLatticeVector.from_min_dspacing
 * (diffsims/crystallography/reciprocal_lattice_vector.py:1077-1142) minus the
 * zero-vector handling, which the caller appends.
 *
 * Enumerates the box [-H,H]x[-K,K]x[-L,L] (highest[] = {H,K,L}) in descending
 * lexicographic (h,k,l) order, keeps 1/sqrt(q) >= d_min with
 * q = h^2 G00 + k^2 G11 + l^2 G22 + 2 (hk G01 + hl G02 + kl G12) evaluated in
 * IEEE FP64 in exactly that order, drops (000), and writes the survivors,
 * order preserved, as int4 {h,k,l,0} to hkl_d.
 *   metric[6] = {G00, G11, G22, G01, G02, G12} (host, FP64 reciprocal metric)
 *   hkl_d     : device int32 [capacity][4]
 *   count_d   : device int32 [1], receives the number of survivors
 *   work_d    : device int32 scratch of dsx_hkl_workspace_ints(highest) ints
 * If the number of survivors exceeds capacity only `capacity` rows are
 * written; *count_d still receives the full count.
 */
size_t dsx_hkl_workspace_ints(const int highest[3]);
int dsx_hkl_enumerate(const int highest[3], const double metric[6], double d_min,
                      int *hkl_d, int capacity, int *count_d, int *work_d,
                      void *stream);

/*
 * Reciprocal lattice points in a sphere — replaces get_points_in_sphere
 * (diffsims/utils/sim_utils.py:436-474) as the legacy DiffractionGenerator
 * (diffsims/generators/diffraction_generator.py:192-331) and
 * calculate_profile_data (:334-417) use it: the box [-H,H]x[-K,K]x[-L,L]
 * (H = floor(radius / |a*|), etc., from the caller) in ascending
 * lexicographic (h,k,l) order (itertools.product of ascending ranges), kept
 * where sqrt(q) < radius with q the quadratic form of dsx_hkl_enumerate;
 * (000) is kept.  Arguments as for dsx_hkl_enumerate (same workspace).
 */
int dsx_hkl_points_in_sphere(const int highest[3], const double metric[6], double radius,
                             int *hkl_d, int capacity, int *count_d, int *work_d, void *stream);

/*
 * Per-reflection tables — replaces the per-orientation structure-factor work of
 * get_kinematical_intensities / _get_kinematical_structure_factor
 * (diffsims/utils/sim_utils.py:256-354), computed once per phase.
 *
 * For each of n rows of hkl_d (int4 {h,k,l,_}):
 *   g (crystal Cartesian, 1/A) = rec * hkl in FP64 (rec row-major 3x3 host),
 *   g4_d[i] = {gx, gy, gz, |g|^2} as FP32,
 *   F = sum_j f_e(j)(|g|^2) occ_j exp(-B_j |g|^2 / 4) exp(2 pi i hkl.x_j) in
 *       FP32 with FP32 accumulation; the phase hkl.x_j is reduced exactly with
 *       32-bit fixed-point fractional positions (xyz_fx_d, uint32 [n_atoms][4],
 *       value = round(frac * 2^32) mod 2^32),
 *   F2_d[i] = |F|^2,  F_d[i] = {Re F, Im F} (F_d may be NULL).
 * f_e is evaluated and F accumulated in FP64 (the Lobato sums cancel by up to
 * two digits); the per-atom sincos and Debye-Waller exp are FP32.
 * Atoms: elem_d int32 [n_atoms] (row into coeff_d), occ_d, B_d float
 * [n_atoms]; coeff_d double [n_elem][10] = {a1,b1,...,a5,b5};
 * scatter = DSX_SCATTER_*.
 */
int dsx_structure_factor(const int *hkl_d, int n, const double rec[9],
                         const uint32_t *xyz_fx_d, const int *elem_d,
                         const float *occ_d, const float *B_d, int n_atoms,
                         const double *coeff_d, int n_elem, int scatter,
                         float *g4_d, float *F2_d, float *F_d, void *stream);

/*
 * Line index of a reflection table, for dsx_simulate / dsx_simulate_ragged.
 * No counterpart in the reference: it lets the spot kernel skip rows that
 * get_intersecting_reflections (simulation_generator.py:319-412) would
 * reject, without changing its output.
 *
 * A line is a maximal run of consecutive rows with equal (h, k) and l falling
 * by one per row (the rows of one (h, k) column of dsx_hkl_enumerate's box;
 * any table works, irregular ones just have short lines).  Entry j is int32
 * [8] {first row, row count, h, k, l of the first row, min |g|^2, max |g|^2
 * (float bits of g4.w over the line's rows), 0}, lines in row order.
 *   hkl_d, g4_d : the table (n_refl rows; g4 from dsx_structure_factor)
 *   lines_d     : device int32 [capacity][8]
 *   count_d     : device int32 [1], receives the number of lines (all of them,
 *                 also when more than capacity; only capacity are written)
 *   work_d      : device int32 scratch of dsx_hkl_lines_workspace_ints ints
 */
long long dsx_hkl_lines_workspace_ints(int n_refl);
/* dsx_simulate / dsx_simulate_ragged take the line sweep for a table whose
 * lines average at least this many rows (n_refl >= DSX_LINE_SWEEP_MIN_ROWS *
 * n_lines); shorter lines take the full sweep.  The output is the same. */
#define DSX_LINE_SWEEP_MIN_ROWS 10
int dsx_hkl_lines(const int *hkl_d, const float *g4_d, int n_refl, int *lines_d,
                  int capacity, int *count_d, int *work_d, long long work_ints,
                  void *stream);

/*
 * Atom staging for dsx_structure_factor when positions and Debye-Waller
 * factors live on the device (gradient-based refinement; no counterpart in
 * the reference, whose Phase keeps them on the host).  Row i of the outputs
 * takes atom order_d[i] (int64 [n_atoms]):
 *   xyz_fx_d[4 i + c] = round((f - floor(f)) * 2^32) mod 2^32 (half to even)
 *                       of xyz_d[3 order + c] (double fractional), c < 3; the
 *                       fourth word is 0
 *   B_out_d[i]        = (float) B_d[order] (A^2)
 */
int dsx_stage_atoms(const double *xyz_d, const double *B_d, const long long *order_d,
                    int n_atoms, uint32_t *xyz_fx_d, float *B_out_d, void *stream);

/*
 * Simulation parameters (SimulationGenerator.__init__ / calculate_diffraction2d
 * arguments, diffsims/generators/simulation_generator.py:74-146).
 */
typedef struct dsx_sim_params {
  double k;                 /* 1 / wavelength (1/A)                          */
  double max_excitation_error;
  double shape_width;       /* shape_factor_width or max_excitation_error     */
  int shape_model;          /* DSX_SF_*                                       */
  double minima_number;     /* sinc / sin2c / atanc parameter (default 5)     */
  double precession_angle;  /* radians; 0 = no precession                     */
  double minimum_intensity; /* relative cut (default 1e-20)                   */
  double g_max;             /* max |g| over the table (pretest bound)         */
  double rec[9];            /* FP64 reciprocal basis, row-major, g = rec hkl  */
} dsx_sim_params;

/*
 * Per-orientation excitation error, shape factor, relative cut and ordered
 * compaction — replaces the orientation loop of calculate_diffraction2d with
 * get_intersecting_reflections (simulation_generator.py:211-241, :319-412).
 *
 *   hkl_d, g4_d, F2_d : tables (n_refl rows; int4 hkl as enumerated, with
 *                the direct-beam rows appended by the caller when
 *                with_direct_beam; g4/F2 from dsx_structure_factor)
 *   lines_d    : optional line index of the table (dsx_hkl_lines, n_lines
 *                entries) or NULL.  With it the kernel visits only the rows
 *                of each line that can pass the pretest for an orientation
 *                instead of every row; the outputs are identical.  It is
 *                used when the lines average at least 16 rows (shorter ones
 *                cost more than the rows they skip) and not under the
 *                averaged-precession models.
 *   Rt_d       : double [n_orient][9] row-major R = M^T (g_lab = R g), M the
 *                orix to_matrix() of each orientation
 * The FP32 g4 table only drives a conservative pretest; every candidate is
 * re-evaluated in FP64 from its integer hkl, rec and Rt, so the excitation
 * error, the |s| < max_excitation_error test and the shape factor are FP64.
 * Outputs, per orientation o, in the reference's reflection order:
 *   idx_d[o*cap + j]   int32 row index into the tables
 *   spot_d[o*cap + j]  float4 {x, y, z, I} lab frame (1/A) and intensity
 *   count_d[o]         survivors after the relative cut (<= cap written)
 *   ncand_d[o]         intersecting reflections before the cut (may exceed cap)
 *   maxI_d[o]          max intensity over the intersecting set
 * An orientation whose intersecting set exceeds cap writes cap rows and sets
 * ncand_d[o] > cap; the caller re-runs it with a larger cap.
 */
int dsx_simulate(const int *hkl_d, const float *g4_d, const float *F2_d, int n_refl,
                 const int *lines_d, int n_lines, const double *Rt_d, int n_orient,
                 const dsx_sim_params *params, int cap, int *idx_d, float *spot_d,
                 int *count_d, int *ncand_d, float *maxI_d, void *stream);

/*
 * One phase of a ragged multi-phase launch (device-resident array of these).
 * Each phase has its own reflection table (as for dsx_simulate), reciprocal
 * basis, largest |g|, orientations and output buffers with capacity cap.
 */
typedef struct dsx_phase_desc {
  const int *hkl;          /* int32 [n_refl][4]                                */
  const float *g4;         /* float [n_refl][4]                                */
  const float *F2;         /* float [n_refl]                                   */
  int n_refl;
  int cap;
  double rec[9];
  double g_max;
  const double *Rt;        /* double [n_orient][9]                             */
  int n_orient;
  int n_lines;             /* entries of lines (0 when lines is NULL)          */
  int *idx;                /* int32 [n_orient][cap]                            */
  float *spots;            /* float [n_orient][cap][4]                         */
  int *count;              /* int32 [n_orient]                                 */
  int *ncand;              /* int32 [n_orient]                                 */
  float *maxI;             /* float [n_orient]                                 */
  const int *lines;        /* optional line index (dsx_hkl_lines) or NULL      */
} dsx_phase_desc;

/* Orientations per work item (tile) of dsx_simulate_ragged. */
int dsx_simulate_tile_size(void);

/*
 * Ragged multi-phase form of dsx_simulate — replaces the per-phase loop of
 * calculate_diffraction2d (simulation_generator.py:198-242) with one
 * persistent launch over every phase.
 *   phases_d  device array of n_phases dsx_phase_desc
 *   items_d   device int32 [n_items][2] {phase, first orientation}: one work
 *             item per dsx_simulate_tile_size() orientations, processed in
 *             the given order (put the most expensive first)
 *   params    physics shared by all phases (rec and g_max are ignored: each
 *             phase carries its own)
 *   blocks_per_cu  launch width in blocks of one wave per CU (0 = default)
 * Outputs per phase as for dsx_simulate.
 */
int dsx_simulate_ragged(const dsx_phase_desc *phases_d, int n_phases, const int *items_d,
                        int n_items, const dsx_sim_params *params, int blocks_per_cu,
                        void *stream);

/*
 * Render parameters (Simulation2D.get_diffraction_pattern,
 * diffsims/simulations/simulation2d.py:357-442).
 */
typedef struct dsx_render_params {
  int height, width;
  float calibration;       /* 1/A per pixel                                   */
  float center_x, center_y;/* direct-beam pixel position                      */
  float cos_angle, sin_angle; /* in-plane rotation                            */
  int mirrored;
  int normalize;
  float sigma;             /* Gaussian sigma in pixels                        */
  float clip_threshold;    /* sub-pixel mode only                             */
  int radius;              /* fast mode: int(4 sigma + 0.5)                   */
  float taps[65];          /* fast mode: taps[0..R] = w(0..R) of scipy's
                              normalised Gaussian kernel, R <= 64        */
} dsx_render_params;

/*
 * fast=True rendering of n patterns — replaces get_diffraction_pattern with
 * get_pattern_from_pixel_coordinates_and_intensities' integer path
 * (simulation2d.py:405-442, detector_functions.py:293-300): truncate to integer
 * pixels, last write wins per pixel, separable Gaussian with reflect
 * boundaries and truncate 4 (scipy.ndimage.gaussian_filter), divide by max.
 *   spot_d/count_d : as written by dsx_simulate (stride cap per pattern)
 *   out_d          : float [n][height][width]
 *   status_d       : device int32 [n] (may be NULL); 0 = rendered, 1 = the
 *                    pattern has more than dsx_render_max_spots() distinct
 *                    in-frame spots and was not rendered completely.
 */
int dsx_render_max_spots(void);
int dsx_render_fast(const float *spot_d, const int *count_d, int cap, int n,
                    const dsx_render_params *params, float *out_d, int *status_d,
                    void *stream);

/*
 * fast=True rendering of n patterns, the same output as dsx_render_fast bit
 * for bit — replaces the same reference path (simulation2d.py:405-442,
 * detector_functions.py:293-300) — in two kernels: a plan kernel (one wave per
 * pattern: pixels, last write wins, the exact maximum) hands each band of
 * rows its spots through work_d, then a one-shot grid writes every band
 * (two 16-byte stores per thread).  Patterns the plan cannot hold are rendered
 * by dsx_render_fast's kernel in the same call.  Shapes outside the band
 * path (tap radius above 8 or not below both sides, width not a multiple of 4
 * or above 1024, height above 1024) run dsx_render_fast's kernel alone and
 * leave work_d unused.
 *   work_d : dsx_render_fast_bands_workspace_ints(n, height, width) ints
 *            (patterns are taken in rounds of at most 16,384, which reuse
 *            it in stream order); overlapping
 *            calls must not share it
 * Other arguments as for dsx_render_fast.
 */
long long dsx_render_fast_bands_workspace_ints(int n, int height, int width);
int dsx_render_fast_bands(const float *spot_d, const int *count_d, int cap, int n,
                          const dsx_render_params *params, int *work_d, long long work_ints,
                          float *out_d, int *status_d, void *stream);

/*
 * fast=True rendering of large patterns with many spots (the same result as
 * dsx_render_fast, no limit on spots per pattern): each output strip rebuilds
 * the spot image of its window in LDS and filters it; in-frame spots are
 * binned by source row strip in work_d, dsx_render_fast_large_workspace_ints(cap)
 * ints.  The tap radius must be smaller than both pattern dimensions.
 */
long long dsx_render_fast_large_workspace_ints(int cap);
int dsx_render_fast_large(const float *spot_d, const int *count_d, int cap, int n,
                          const dsx_render_params *params, int *work_d, long long work_ints,
                          float *out_d, int *status_d, void *stream);

/*
 * fast=False rendering — replaces _subpixel_gaussian
 * (detector_functions.py:314-359): per spot, the clipped box of
 * I/(2 pi s^2) exp(-d^2 / (2 s^2)) (skipped when I/(2 pi s^2) < clip_threshold),
 * centre ((W-1)/2, (H-1)/2) unless given, optional normalisation.  Arguments
 * as for dsx_render_fast.
 */
int dsx_render_subpixel(const float *spot_d, const int *count_d, int cap, int n,
                        const dsx_render_params *params, float *out_d,
                        int *status_d, void *stream);

/*
 * Ragged fast rendering: the patterns of many phases, each with its own spot
 * capacity, from one spot pool — the render half of the multi-phase
 * calculate_diffraction2d + get_diffraction_pattern loop
 * (simulation_generator.py:198-242, simulation2d.py:357-442).
 *   spot_d      float4 spot pool; pattern p's list starts at spot_off_d[p]
 *               (int64, in spots) and holds min(count_d[p], cap_d[p]) spots
 *   list_d      int32 [n] pattern ids to render; out_d/status_d are indexed by
 *               pattern id (out_d float [*][height][width])
 *   method      DSX_RENDER_SPLAT: footprint splatting (at most
 *               dsx_render_max_spots() distinct in-frame spots per pattern);
 *               DSX_RENDER_DENSE: whole pattern in LDS, owner map + separable
 *               filter, any number of spots, at most
 *               dsx_render_dense_max_pixels() pixels;
 *               DSX_RENDER_DENSE_INPLACE: the same in one buffer (below)
 *   blocks_per_cu  launch width (0 = default)
 * Same results as dsx_render_fast.
 */
#define DSX_RENDER_SPLAT 0
#define DSX_RENDER_DENSE 1
#define DSX_RENDER_DENSE_INPLACE 2
int dsx_render_dense_max_pixels(void);
/* DSX_RENDER_DENSE_INPLACE: as DSX_RENDER_DENSE (the same bits) from one LDS
 * buffer per pattern (owner map, spot image and both filter passes in place),
 * two patterns per CU; 1 when it takes this shape and tap radius
 * (128 x 128, radius 1..5), else 0. */
int dsx_render_dense_inplace_fits(int height, int width, int radius);
int dsx_render_fast_ragged(const float *spot_d, const long long *spot_off_d, const int *cap_d,
                           const int *count_d, const int *list_d, int n,
                           const dsx_render_params *params, int method, int blocks_per_cu,
                           float *out_d, int *status_d, void *stream);

/*
 * Adjoint of dsx_simulate + dsx_render_subpixel.  No counterpart in the
 * reference (diffsims computes no gradients); it serves gradient-based
 * refinement of orientation, lattice, positions and Debye-Waller factors.
 *
 * For each pattern the sub-pixel pattern P (normalised when params->normalize)
 * is recomputed from the spot lists exactly as dsx_render_subpixel renders it,
 * and the upstream gradient G = dL/dP is
 *   mode DSX_ADJ_L2            G = 2 (P - T), loss_d[o] = sum (P - T)^2
 *   mode DSX_ADJ_PATTERN_GRAD  G = T (a caller-supplied dL/dP)
 *   mode DSX_ADJ_FORWARD       no G: only P is written (target_d unused)
 * where T = target_d + o * target_stride (float [H][W]; stride 0 compares
 * every pattern with one target).  G is pulled back through the render, the
 * intensity |F|^2 f(s, r) and the excitation error, which are recomputed from
 * idx_d, hkl_d, F2_d, Rt_d and sim->rec (the inputs dsx_simulate used):
 *   dRt_d   double [n][9]    dL/dRt of each pattern (overwritten); NULL: no
 *                            gradients (loss / pattern only)
 *   dF2_d   float  [n_refl]  dL/d|F(hkl)|^2, ACCUMULATED (zero it first); may
 *                            be NULL (skipped)
 *   drec_d  double [9]       sum over spots of dL/dg (crystal frame) hkl^T,
 *                            i.e. dL/drec through the spot positions and
 *                            excitation errors, ACCUMULATED; may be NULL
 * dF2_d and drec_d need dRt_d.  pattern_d
 * (float [n][H][W], may be NULL) receives P; loss_d (float [n], may be NULL).
 * work_d is a 16-byte aligned workspace of dsx_render_adjoint_workspace_ints(
 * cap, n, H, W, gradients) ints (gradients: nonzero when dRt_d is passed).
 * With cap <= dsx_render_max_spots() it is needed only for gradients
 * (per-spot sums handed from the render pass to the chain-rule pass; EINVAL
 * without it) and every status_d entry is 0.  With a larger cap it holds
 * the strip bins and clip boxes of the patterns, their fixed-point pixels
 * between the splat and the read-out when pattern_d is NULL, and with
 * gradients n * cap per-spot sums (EINVAL without it); when it is missing
 * without gradients, or too small for a pattern's bins, that pattern's
 * status_d entry is 1.  At most 2^22 pixels per pattern.
 * This entry point also serves as the sub-pixel renderer for such large
 * patterns (mode DSX_ADJ_FORWARD).
 */
#define DSX_ADJ_L2 0
#define DSX_ADJ_PATTERN_GRAD 1
#define DSX_ADJ_FORWARD 2
long long dsx_render_adjoint_workspace_ints(int cap, int n, int height, int width,
                                            int gradients);
int dsx_render_adjoint(const int *idx_d, const float *spot_d, const int *count_d, int cap,
                       int n, const int *hkl_d, const float *F2_d, int n_refl,
                       const double *Rt_d, const dsx_sim_params *sim,
                       const dsx_render_params *params, int mode, const float *target_d,
                       long long target_stride, float *pattern_d, float *loss_d,
                       double *dRt_d, float *dF2_d, double *drec_d, int *work_d,
                       long long work_ints, int *status_d, void *stream);

/*
 * Allocation-free dsx_render_adjoint.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight refinement loops, harnesses that forbid hipMalloc).  The
 * arguments and results are dsx_render_adjoint's without work_d / work_ints:
 * the workspace is a static device array of
 * dsx_render_adjoint_noalloc_max_ints() ints (64 MiB), one per GPU, held until
 * the process ends; its address is looked up on the first call on each
 * device.  The call allocates and frees nothing.  Sizes whose
 * dsx_render_adjoint_workspace_ints(cap, n, H, W, dRt_d != NULL) exceeds the maximum
 * are rejected
 * with DSX_EINVAL.  Every call on one GPU shares that GPU's array, which
 * keeps the last call's contents (each call overwrites what it reads, so no
 * clearing is needed): calls that overlap in time on one GPU (other streams
 * or threads) race and must be serialised by the caller.
 */
long long dsx_render_adjoint_noalloc_max_ints(void);
int dsx_render_adjoint_noalloc(const int *idx_d, const float *spot_d, const int *count_d, int cap,
                               int n, const int *hkl_d, const float *F2_d, int n_refl,
                               const double *Rt_d, const dsx_sim_params *sim,
                               const dsx_render_params *params, int mode, const float *target_d,
                               long long target_stride, float *pattern_d, float *loss_d,
                               double *dRt_d, float *dF2_d, double *drec_d, int *status_d,
                               void *stream);

/*
 * Adjoint of dsx_structure_factor.  No counterpart in the reference.  Given
 * dF2_d = dL/d|F|^2 per reflection (float [n], as accumulated by
 * dsx_render_adjoint), recomputes F(hkl) and ACCUMULATES (zero first):
 *   dxyz_d  float  [n_atoms][3]  dL/d(fractional position), atoms in the order
 *                                given (the order dsx_structure_factor took)
 *   dB_d    float  [n_atoms]     dL/dB (A^2)
 *   drec_d  double [9]           dL/d rec (g = rec hkl, row-major) through
 *                                |F|^2: the |g|-dependence of f and of
 *                                exp(-B g^2/4) (the geometric part comes from
 *                                dsx_render_adjoint's drec_d)
 * Other arguments as for dsx_structure_factor.
 */
int dsx_structure_factor_adjoint(const int *hkl_d, int n, const double rec[9],
                                 const uint32_t *xyz_fx_d, const int *elem_d,
                                 const float *occ_d, const float *B_d, int n_atoms,
                                 const double *coeff_d, int n_elem, int scatter,
                                 const float *dF2_d, float *dxyz_d, float *dB_d,
                                 double *drec_d, void *stream);

/*
 * Beam-direction grid of a template library — replaces get_beam_directions_grid
 * (diffsims/generators/rotation_list_generators.py:176-267) for the meshes of
 * diffsims/generators/sphere_mesh_generators.py: the UV sphere (:42-93) and the
 * normalized, spherified-edge and spherified-corner cube meshes (:96-197).
 * Writes, in the reference's row order, the Bunge angles in degrees
 * {0, Phi, phi2} (beam_directions_grid_to_euler, :486-526) of every mesh
 * vertex v with planes[4 i + 3] * (planes[4 i .. 4 i + 2] . v) >= -1e-13 for
 * each of the n_planes (<= 3) crop planes (the stereographic triangle: normal
 * = cross of two corners, sign = that normal . the third corner; the caller
 * applies the reference's hexagonal resolution / sqrt(2)).
 *   euler_d  double [capacity][3]; rows past capacity are not written
 *   count_d  int64 [1]: rows of the full result
 *   work_d   dsx_beam_directions_workspace_ints(mesh, resolution) ints
 * dsx_beam_directions_candidates() is the vertex count (an upper bound of the
 * rows); both return -1 for an unknown mesh or a bad resolution.
 */
#define DSX_MESH_UV_SPHERE 0
#define DSX_MESH_NORMALIZED_CUBE 1
#define DSX_MESH_SPHERIFIED_EDGE 2
#define DSX_MESH_SPHERIFIED_CORNER 3
long long dsx_beam_directions_candidates(int mesh, double resolution);
long long dsx_beam_directions_workspace_ints(int mesh, double resolution);
int dsx_beam_directions(int mesh, double resolution, int n_planes, const int *planes,
                        double *euler_d, long long capacity, long long *count_d, int *work_d,
                        long long work_ints, void *stream);

/*
 * Workspace of the grid entry points below for n_vertices mesh vertices
 * (candidates), in ints; -1 for a negative count.
 */
long long dsx_grid_workspace_ints(long long n_vertices);

/*
 * Icosahedral mesh — replaces get_icosahedral_mesh_vertices
 * (diffsims/generators/sphere_mesh_generators.py:376-449) with
 * _compose_from_faces (:200-325) and _get_max_grid_angle (:328-372).
 *
 * The mesh at refinement level n >= 1 has dsx_icosahedral_vertices(n) =
 * 10 n^2 + 2 vertices: the 12 corners of the reference's icosahedron, then
 * for each of its 30 edges in edge_order (60 host ints: sorted corner pairs
 * {p, q}, p < q, in the order the reference's edge set iterates) the n-1
 * points (1-t) c_p + t c_q, t = linspace(1/n, 1, n-1, endpoint=False), then
 * for each face in the reference's face order the interior points
 * f0 (1 - i/n - j/n) + f1 j/n + f2 i/n for i = 1..n-2, j = 1..n-i-1.
 *
 * dsx_icosahedral_nn_dots writes, for every level L in [level_first,
 * level_last], min over vertices v of max over the other vertices w of
 * u(v).u(w), u = row / |row| (the cosine of the largest nearest-neighbour
 * angle), to min_dot_d[L - level_first] (double).  The nearest neighbour is
 * searched among the vertex's triangulation neighbours (projection distorts
 * the triangles by at most 1.26, a vertex two steps away is >= sqrt(3)
 * steps off).  The reference refines level 1, 2, ... until
 * degrees(arccos(min_dot)) <= resolution; the caller does the same.
 *
 * dsx_beam_directions_icosahedral writes the level's vertices divided by
 * their norms, cropped as dsx_beam_directions crops (n_planes, planes), in
 * vertex order: Bunge angles in degrees {0, Phi, phi2} per kept vertex, or,
 * with xyz_out != 0, the unit vertex itself (out_d double [capacity][3];
 * count_d int64 [1]; work_d dsx_grid_workspace_ints(vertices) ints).
 */
long long dsx_icosahedral_vertices(int level);
int dsx_icosahedral_nn_dots(int level_first, int level_last, const int *edge_order,
                            double *min_dot_d, void *stream);
int dsx_beam_directions_icosahedral(int level, const int *edge_order, int n_planes,
                                    const int *planes, int xyz_out, double *out_d,
                                    long long capacity, long long *count_d, int *work_d,
                                    long long work_ints, void *stream);

/*
 * Crop and Euler conversion of given unit vectors (double [n][3]): the part of
 * get_beam_directions_grid (rotation_list_generators.py:233-267) after the
 * mesh, with beam_directions_grid_to_euler (sphere_mesh_generators.py:486-526)
 * — for the random mesh (get_random_sphere_vertices, :452-483, whose normal
 * draws are numpy's generator) and for beam_directions_grid_to_euler itself
 * (n_planes = 0).  Arguments as for dsx_beam_directions.
 */
int dsx_beam_directions_from_vertices(const double *xyz_d, long long n, int n_planes,
                                      const int *planes, double *euler_d, long long capacity,
                                      long long *count_d, int *work_d, long long work_ints,
                                      void *stream);

/*
 * SO(3) rotation grids — replace get_fundamental_zone_grid and get_local_grid
 * (diffsims/generators/rotation_list_generators.py:85-134), whose sampling the
 * reference takes from orix (get_sample_fundamental / get_sample_local on the
 * cubochoric grid of Rosca, Morawiec & De Graef 2014).  The grid has
 * (2 semi_edge_steps)^3 points i a / n, i = -n+1 .. n, a = pi^(2/3) / 2,
 * x outermost; each becomes a unit quaternion (w >= 0) and is kept
 *   DSX_SO3_FUNDAMENTAL_ZONE: if |w| - max_s |q . s| >= -1e-12 over the n_ops
 *     rotations s of ops (quaternions, [n_ops][4], identity excluded, <= 24);
 *   DSX_SO3_LOCAL: if acos(w) < half_angle; the kept quaternion is then
 *     premultiplied by *center ([4], or NULL for none).
 * Kept rows are written in grid order as Bunge angles (radians, double
 * [capacity][3]; Rowenhorst et al. 2015 eq. A.14 with P = +1).  *count_d gets
 * the number of kept points; with euler_d NULL (capacity 0) only the count is
 * formed.  work_d: dsx_so3_grid_workspace_ints(semi_edge_steps) ints.
 * semi_edge_steps is 1..1024; dsx_so3_grid_candidates returns -1 outside.
 */
#define DSX_SO3_FUNDAMENTAL_ZONE 0
#define DSX_SO3_LOCAL 1
long long dsx_so3_grid_candidates(int semi_edge_steps);
long long dsx_so3_grid_workspace_ints(int semi_edge_steps);
int dsx_so3_grid(int semi_edge_steps, int mode, int n_ops, const double *ops, double half_angle,
                 const double *center, double *euler_d, long long capacity, long long *count_d,
                 int *work_d, long long work_ints, void *stream);

/*
 * Polar templates from spot lists — replaces Simulation2D.polar_flatten_simulations
 * (simulation2d.py:313-355) with DiffractingVector.to_flat_polar
 * (_diffracting_vector.py:186-194) and get_closest (simulation2d.py:767-781).
 * For orientation o, the spots of dsx_simulate give r = |(x, y)| and
 * theta = atan2(y, x); rows of `width` entries (zero-padded, in spot order):
 *   without axes (radial_d or azimuthal_d NULL): r_d, theta_d (float [n][width])
 *   with both axes (ascending double arrays): r_idx_d, theta_idx_d (int32
 *     [n][width]) hold the nearest axis indices, entries with an index on or
 *     past the last axis entry are dropped
 *   I_d float [n][width] intensities; n_out_d int32 [n] (may be NULL) entries
 *   kept per orientation.
 */
int dsx_polar_templates(const float *spot_d, const int *count_d, int cap, int n, int width,
                        const double *radial_d, int n_radial, const double *azimuthal_d,
                        int n_azimuthal, float *r_d, float *theta_d, int *r_idx_d,
                        int *theta_idx_d, float *I_d, int *n_out_d, void *stream);

/*
 * Bunge Euler angles (phi1, Phi, phi2 in radians, double [n][3]) -> rows of
 * M^T (double [n][9]) with M the passive Bunge matrix (orix
 * Rotation.from_euler(...).to_matrix(); the reference rotates with
 * ~rotation * miller, _diffracting_vector.py:157-160), and its adjoint
 * deuler[n][3] = sum_ij drt[n][ij] d rt_ij / d angle.  No counterpart in the
 * reference beyond the matrix itself; serves the differentiable path.
 */
int dsx_euler_to_rt(const double *euler_d, int n, double *rt_d, void *stream);
int dsx_euler_to_rt_backward(const double *euler_d, const double *drt_d, int n,
                             double *deuler_d, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* DIFFSIMS_AMD_H */
