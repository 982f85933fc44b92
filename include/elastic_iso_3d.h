/*
 * elastic_iso_3d.h -- C ABI of the MI355X (gfx950) elastic isotropic 3-D
 * staggered-grid propagator.
 *
 * This is the GPU layer that the reference's C++ operator classes call. Each
 * entry point below names the reference function it replaces (paths relative
 * to elastic_iso_lib_3d/):
 *
 *   ela_create / ela_destroy      initNonlinearElasticGpu_3D / initBornElasticGpu_3D
 *                                 (elastic_iso_float/include/nonlinearPropElasticGpuFunctions_3D.h:9,
 *                                  elastic_iso_float/include/BornElasticGpuFunctions_3D.h:9)
 *                                 and deallocate*ElasticGpu_3D (same headers, :11)
 *   ela_set_medium                allocateNonlinearElasticGpu_3D / allocateBornElasticGpu_3D
 *                                 (nonlinearPropElasticGpuFunctions_3D.h:10, BornElasticGpuFunctions_3D.h:10)
 *   ela_nl_forward                propElasticFwdGpu_3D   (nonlinearPropElasticGpuFunctions_3D.h:14)
 *   ela_nl_adjoint                propElasticAdjGpu_3D   (nonlinearPropElasticGpuFunctions_3D.h:18)
 *   ela_born_forward              BornElasticFwdGpu_3D   (BornElasticGpuFunctions_3D.h:17)
 *   ela_born_adjoint              BornElasticAdjGpu_3D   (BornElasticGpuFunctions_3D.h:22)
 *   ela_scale_medium              the staggering/scaling of fdParamElastic_3D's constructor
 *                                 (elastic_iso_float/src/fdParamElastic_3D.cpp:130-183)
 *   ela_interp_linear             spaceInterpGpu_3D's constructor, linear method
 *                                 (elastic_iso_float/include/spaceInterpGpu_3D.h:41)
 *   ela_interp_weights            the same constructor with the sinc method and dipoles
 *                                 (elastic_iso_float/src/spaceInterpGpu_3D.cpp:13-79, 214-480)
 *   ela_born_stagger              perturbation staggering/scaling of BornElasticShotsGpu_3D::forward
 *                                 (elastic_iso_float/src/BornElasticShotsGpu_3D.cpp:167-235)
 *   ela_born_unstagger            its adjoint in BornElasticShotsGpu_3D::adjoint
 *                                 (elastic_iso_float/src/BornElasticShotsGpu_3D.cpp:541-607)
 *
 * Differences from the reference interface, all deliberate:
 *   - state lives in an opaque handle (no file-scope globals indexed by iGpu);
 *   - errors are returned as status codes with ela_last_error(), never exit();
 *   - FP32 and FP64 are one library (precision chosen per handle), not two builds;
 *   - pointers may be host or device memory (ELA_MEM_HOST / ELA_MEM_DEVICE);
 *     with device pointers the call is asynchronous on the handle's stream.
 *
 * Conventions shared with the reference:
 *   - wavefield and medium arrays are [ny][nx][nz] with z fastest, padded grid
 *     including the FAT=4 layer;
 *   - the 8 scaled medium arrays are, in order, rhoxDtw rhoyDtw rhozDtw
 *     lamb2MuDtw lambDtw muxzDtw muxyDtw muyzDtw (2*dtw/rho_x, ..., 2*dtw*(lambda+2mu),
 *     2*dtw*lambda, 2*dtw*mu_xz, ...);
 *   - regular-grid traces are [nReg][nts] (time fastest), one array per
 *     component in the order vx vy vz sxx syy szz sxz sxy syz;
 *   - regular-grid positions are flat int64 indices, one list per grid in the
 *     order centre X Y Z XZ XY YZ; components map to grids as
 *     vx->X vy->Y vz->Z sxx,syy,szz->centre sxz->XZ sxy->XY syz->YZ;
 *   - positions must be unique within a grid (as produced by the reference's
 *     spaceInterpGpu_3D::convertIrregToReg) and inside the FD interior;
 *   - every array must be smaller than 4 GiB (32-bit element offsets);
 *   - `sub` is the number of fine steps per coarse sample, i.e. the doubled
 *     value the reference propagates with (fdParamElastic_3D.cpp:29-36).
 */
#ifndef ELASTIC_ISO_3D_H
#define ELASTIC_ISO_3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define ELA_OK 0
#define ELA_ERR_ARG 1     /* invalid argument or inconsistent sizes */
#define ELA_ERR_HIP 2     /* a HIP runtime call failed */
#define ELA_ERR_MEM 3     /* device or host allocation failed */
#define ELA_ERR_STATE 4   /* call out of order (e.g. no medium set) */

#define ELA_MEM_HOST 0
#define ELA_MEM_DEVICE 1

#define ELA_FP32 4
#define ELA_FP64 8

typedef struct ela_config {
  int64_t nz, nx, ny;   /* padded grid sizes (including FAT) */
  int64_t nts;          /* number of coarse time samples */
  int64_t sub;          /* fine steps per coarse sample (2 * par "sub") */
  double dz, dx, dy;    /* spacings [m] (y derivatives use dx, as the reference) */
  double dts;           /* coarse sampling [s]; dtw = dts / sub */
  int64_t min_pad;      /* sponge width: min(zPadMinus, zPadPlus, xPadMinus, xPadPlus) */
  double alpha_cos;     /* sponge parameter alphaCos (reference default 0.99) */
  int32_t precision;    /* ELA_FP32 or ELA_FP64 */
  int32_t device;       /* HIP device ordinal; the library sets it for its own calls */
} ela_config;

/* Regular-grid point lists for the 7 grids. */
typedef struct ela_points {
  const int64_t* pos[7];
  int64_t n[7];
} ela_points;

/* One trace array per component; element type follows the handle precision. */
typedef struct ela_traces {
  void* tr[9];
} ela_traces;

typedef struct ela_handle ela_handle;

/* Last error message of the calling thread ("" if none). */
const char* ela_last_error(void);

/* Library version string. */
const char* ela_version(void);

/* Create / destroy a propagator for one grid and time axis. */
int ela_create(const ela_config* cfg, ela_handle** out);
int ela_destroy(ela_handle* h);

/* Use a caller-owned hipStream_t for device-memory calls. NULL selects the
 * handle's own stream, a blocking stream that is ordered with HIP's null
 * (legacy default) stream, so inputs written and outputs read by work on the
 * null stream need no extra synchronisation. */
int ela_set_stream(ela_handle* h, void* hip_stream);
/* Wait for all work queued on the handle's stream. */
int ela_synchronize(ela_handle* h);

/* Upload the 8 scaled medium arrays (8 * nz*nx*ny values, element type of
 * the handle precision). */
int ela_set_medium(ela_handle* h, const void* med8, int mem);

/* Upload the elastic parameters (rho, lambda, mu) [3][ny][nx][nz] (element
 * type of the handle precision) instead of the 8 scaled arrays: the library
 * derives the 8 arrays on the device with the staggering and scaling of
 * ela_scale_medium (same operations and order; dtw = dts / sub as given there),
 * and its forward and nonlinear-adjoint steps form the staggered, scaled
 * values from the three parameters as they go, reading 3 medium arrays per
 * cell instead of 8 (120 instead of 140 bytes per FP32 cell-update). The other
 * steps read the derived 8 arrays. Replaces the medium of ela_set_medium (and
 * is replaced by a later ela_set_medium). No counterpart in the reference,
 * whose allocate functions take the 8 scaled arrays
 * (nonlinearPropElasticGpuFunctions_3D.h:10). */
int ela_set_elastic(ela_handle* h, const void* rlm3, double dtw, int mem);

/* Copy the handle's 8 scaled medium arrays out ([8][ny][nx][nz], element
 * type of the handle precision): the arrays of ela_set_medium, or those
 * ela_set_elastic derived. In FP32 the derived buoyancies and normal moduli
 * are the host scaling (ela_scale_medium) of the same values rounded to FP32,
 * bit for bit; the shear moduli are FP32 half-cell sums. No counterpart in the
 * reference (diagnostics and tests). */
int ela_get_medium(ela_handle* h, void* med8, int mem);

/* Host-side staggering and scaling of (rho, lambda, mu) into the 8 arrays of
 * ela_set_medium (fdParamElastic_3D.cpp:130-183). Double precision in and out;
 * dtw = dts / sub. */
int ela_scale_medium(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* rho,
                     const double* lame, const double* mu, double* med8);

/* Born model perturbations (drho, dlambda, dmu) [3][ny][nx][nz] -> the 8
 * staggered, scaled arrays of ela_born_forward (BornElasticShotsGpu_3D.cpp:167-235).
 * med8 is the output of ela_scale_medium for the background. Double precision. */
int ela_born_stagger(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* med8,
                     const double* dm3, double* out8);

/* Adjoint of ela_born_stagger: the 8 images of ela_born_adjoint -> (drho,
 * dlambda, dmu), overwritten (BornElasticShotsGpu_3D.cpp:541-607). */
int ela_born_unstagger(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* med8,
                       const double* img8, double* out3);

/* Chain rule through ela_scale_medium: g8 = dJ/d(the 8 scaled arrays) ->
 * dJ/d(rho, lambda, mu), overwritten. No counterpart in the reference (its
 * inversion uses the Born adjoint as the gradient); used by the adjoint-state
 * gradient. */
int ela_medium_gradient(int64_t nz, int64_t nx, int64_t ny, double dtw, const double* med8,
                        const double* g8, double* out3);

/* Nonlinear forward: source traces -> receiver traces. `rec` is overwritten. */
int ela_nl_forward(ela_handle* h, const ela_points* src_pos, const ela_traces* src,
                   const ela_points* rec_pos, ela_traces* rec, int mem);

/* Allocation-free nonlinear forward. No counterpart in the reference: it is
 * for callers that cannot allocate device memory inside the call (serving
 * loops, code that forbids hipMalloc between set-up and teardown). Same
 * result as ela_nl_forward with ELA_MEM_DEVICE (positions, traces and output
 * in device memory, asynchronous on the handle's stream) on the same handle,
 * but the entry point calls no hipMalloc / hipFree or allocating library
 * routine: its scratch (converted point lists, an error flag, and the
 * time-major receiver buffer of nrec_total x nts values) is a file-scope
 * device array of ELA_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends, and its time loop is launched eagerly (no graph). Calls needing more
 * scratch than that return ELA_ERR_ARG; stacked windows whose chunk table is
 * stale (ela_set_tiling after ela_set_windows) return ELA_ERR_STATE. The
 * handle (wavefields, medium) is allocated beforehand by ela_create /
 * ela_set_medium / ela_set_elastic. Every call on one GPU shares that GPU's
 * scratch array: calls that overlap on one GPU (several handles or threads on
 * different streams) race and must be serialised by the caller. */
#define ELA_NOALLOC_SCRATCH_BYTES (64ull << 20)
int ela_nl_forward_noalloc(ela_handle* h, const ela_points* src_pos, const ela_traces* src,
                           const ela_points* rec_pos, ela_traces* rec);

/* Nonlinear adjoint: receiver traces -> source traces. `src` is overwritten. */
int ela_nl_adjoint(ela_handle* h, const ela_points* src_pos, ela_traces* src,
                   const ela_points* rec_pos, const ela_traces* rec, int mem);

/* Born forward: the 8 scaled, staggered perturbations (drhox drhoy drhoz
 * dlame dmu dmuxz dmuxy dmuyz, each nz*nx*ny) -> receiver traces
 * (overwritten). `src` are the medium-scaled source traces of the background
 * shot. */
int ela_born_forward(ela_handle* h, const ela_points* src_pos, const ela_traces* src,
                     const void* dmod8, const ela_points* rec_pos, ela_traces* rec, int mem);

/* Born adjoint: receiver traces -> the 8 perturbation images, ACCUMULATED
 * into dmod8 (as the reference, which copies the incoming arrays to the
 * device and adds to them). */
int ela_born_adjoint(ela_handle* h, const ela_points* src_pos, const ela_traces* src,
                     void* dmod8, const ela_points* rec_pos, const ela_traces* rec, int mem);

/* Storage of the background coarse wavefield (vx, vy, vz at every coarse
 * sample) that the Born adjoint images against. The reference keeps all of it
 * in pinned host memory (BornElasticGpuFunctions_3D.cu:2147-2200), 3*nts
 * arrays; here it stays in HBM when it fits, otherwise the forward sweep
 * stores exact checkpoints of the background state every `segment` coarse
 * samples and keeps the last segment's samples, and every other segment is
 * recomputed before its adjoint sweep (one extra background propagation
 * less the last segment; results are bit-identical).
 *   segment = 0   automatic: all samples if they fit, else the longest
 *                 segment that fits, within budget_bytes of device memory
 *                 (0: the free device memory at call time);
 *   segment = -1  always keep all samples;
 *   segment = K   checkpoint every K coarse samples.
 * The Born forward never stores the coarse wavefield: it propagates the
 * background two coarse samples ahead of the scattered field. */
int ela_set_born_storage(ela_handle* h, int64_t segment, int64_t budget_bytes);
/* Plan used by the last Born adjoint: -1 all samples kept, K > 0 checkpoint
 * segment length. */
int ela_born_storage_used(const ela_handle* h, int64_t* segment);

/* Form of the yz shear term of the Born operators' secondary source and
 * imaging. ELA_BORN_REFERENCE (default) reproduces the reference, whose yz
 * term multiplies dmuyz by the xz derivative pattern D-z vx + D-x vz
 * (elastic_iso_float/src/kernelsElaGpu_3D.cu:1419-1431, :1622-1634); the
 * forward and adjoint stay transposes of each other, but the operator is not
 * the linearisation of the scheme in mu. ELA_BORN_PHYSICS uses the yz pattern
 * D-y vz + D-z vy of the stepper's syz update, making the Born forward the
 * first-order Taylor term of the nonlinear operator and the Born adjoint's
 * mu image consistent with the adjoint-state gradient. Nothing else differs.
 * No counterpart in the reference. */
#define ELA_BORN_REFERENCE 0
#define ELA_BORN_PHYSICS   1
int ela_set_born_mode(ela_handle* h, int mode);

/* ---- Adjoint-state gradient (no counterpart in the reference, whose
 * inversion uses the Born adjoint as its gradient) ----------------------------
 *
 * Gradient of a data functional J(d) of the nonlinear forward data with
 * respect to the 8 scaled medium arrays: the exact transpose of the
 * linearised fine-step scheme. Two calls per shot:
 *   ela_gradient_forward   = ela_nl_forward that also stores checkpoints of the
 *                            propagation state (kept in the handle);
 *   ela_gradient_backward  = the adjoint propagation driven by adj = dJ/d(rec)
 *                            (regular-grid receiver traces, same layout as rec),
 *                            recomputing the forward between checkpoints and
 *                            imaging at every fine step. grad8 (8 arrays of
 *                            nz*nx*ny doubles, host or device per `mem`) is
 *                            ACCUMULATED in FP64; src_grad (the layout of src)
 *                            receives dJ/d(scaled source traces), from which
 *                            the caller adds the source-scaling term.
 * The backward call must pass the same source positions and traces as the
 * forward call; ela_set_medium invalidates the stored checkpoints. */
int ela_gradient_forward(ela_handle* h, const ela_points* src_pos, const ela_traces* src,
                         const ela_points* rec_pos, ela_traces* rec, int mem);
int ela_gradient_backward(ela_handle* h, const ela_points* src_pos, const ela_traces* src,
                          const ela_points* rec_pos, const ela_traces* adj, double* grad8,
                          ela_traces* src_grad, int mem);
/* Storage formats of the gradient's outer checkpoints (2 or 4/8 bytes per value). */
#define ELA_CK_NATIVE 0   /* the handle's precision, exact */
#define ELA_CK_BF16   1   /* bfloat16, one power-of-two scale per array and snapshot */
#define ELA_CK_S16    2   /* int16, one scale per 64 consecutive values (2.06 bytes per value) */

/* Checkpointing of the gradient forward: every `interval` fine steps the
 * backward pass recomputes a segment and stores its derivatives; outer
 * checkpoints are taken every `outer_interval` steps (a multiple of
 * `interval`; 0: the smallest memory-feasible choice, equal to `interval` when
 * it fits), stored in `format` (ELA_CK_*). bfloat16 keeps 8 significant bits
 * of each value; the block-scaled int16 format keeps every value to 2^-16 of
 * the largest one of its 64-value block. With outer_interval > interval the
 * outer segments are recomputed once more to store exact inner checkpoints.
 * budget_bytes bounds the automatic choice (0: free device memory).
 * Defaults: 20, 0, ELA_CK_S16, 0 -- the int16 format keeps the FP32 gradient
 * within 1e-3 of the FP64 one (measured 5.4e-5 on a 256x256x160 model over
 * 4000 steps) in the memory of bfloat16, which misses it (1.8e-3: 8
 * significant bits restart every segment, whatever the scale). No
 * counterpart in the reference (which has no gradient). */
int ela_set_gradient_checkpoints(ela_handle* h, int64_t interval, int64_t outer_interval, int format,
                                 int64_t budget_bytes);
/* Plan of the last gradient forward (format = ELA_CK_*) and the bytes its outer checkpoints hold. */
int ela_gradient_plan_used(const ela_handle* h, int64_t* interval, int64_t* outer_interval, int* format,
                           int64_t* checkpoint_bytes);
/* Storage of the 9 derivative combinations per cell and step that the
 * backward's recomputation writes and its transposed steps read: ELA_CK_NATIVE
 * (exact; the default) or ELA_CK_S16 (FP32 handles only): each value within
 * 2^-16 of the largest of its 32-value block (one 128-byte line of a z row),
 * 2.125 bytes per value instead of 4, which halves the backward's largest
 * buffer (interval x 9 arrays) for grids where memory is short. The values
 * enter the gradient linearly and are never propagated (measured: 1.9e-6
 * from the FP64 gradient against 1.1e-6 exact on a 256x256x160 model), but
 * the block maxima cost the recomputation more than the bytes save (about 5%
 * slower backward). Takes effect from the next ela_gradient_forward. No
 * counterpart in the reference. */
int ela_set_gradient_derivatives(ela_handle* h, int format);
/* The derivative storage the next gradient uses (ELA_CK_NATIVE or ELA_CK_S16). */
int ela_gradient_derivatives_used(const ela_handle* h, int* format);

/* Axes of one acquisition grid (its origin already shifted by half a cell on
 * the staggered axes, as python_float/Elastic_iso_float_3D.py:97-110 does),
 * with the padding parameters of the par file. */
typedef struct ela_interp_axes {
  int64_t nz, nx, ny;
  double oz, ox, oy;
  double dz, dx, dy;
  int64_t fat, z_pad_minus, z_pad_plus, x_pad_minus, x_pad_plus, y_pad;
  double tolerance;     /* par "errorTolerance" (reference default 1e-6) */
} ela_interp_axes;

/* Trilinear interpolation weights of n_dev devices on one grid
 * (spaceInterpGpu_3D linear method: elastic_iso_float/src/spaceInterpGpu_3D.cpp:13-79,
 * 113-138, 177-279). Outputs: nbr[8*n_dev] (index of each neighbour in the
 * unique list), weight[8*n_dev], unique_pos[<= 8*n_dev] (flat grid indices in
 * first-seen order), *n_unique. Returns ELA_ERR_ARG if a device or one of its
 * neighbours falls outside the padded interior. */
int ela_interp_linear(const ela_interp_axes* ax, int64_t n_dev, const double* z, const double* x,
                      const double* y, int64_t* nbr, double* weight, int64_t* unique_pos,
                      int64_t* n_unique);

#define ELA_INTERP_LINEAR 0
#define ELA_INTERP_SINC   1

/* Interpolation weights of n_dev devices on one grid with the linear or the
 * sinc method and an optional dipole: spaceInterpGpu_3D's constructor with
 * interpMethod, hFilter1d, dipole and the dipole shifts
 * (elastic_iso_float/src/spaceInterpGpu_3D.cpp:13-79, 214-480).
 * Each device has nf = (dipole ? 2 : 1) * (2*half_length)^3 entries
 * (linear: half_length must be 1, 8 corners in the reference's order; sinc:
 * the (2h)^3 support y slowest, weights sinc(dz)sinc(dx)sinc(dy)); with a
 * dipole the second nf/2 entries are the negative pole at the device shifted
 * by (z_shift, x_shift, y_shift). Outputs nbr[nf*n_dev], weight[nf*n_dev],
 * unique_pos[<= nf*n_dev], *n_unique as for ela_interp_linear.
 * Corrections to the reference (its sinc and dipole paths are otherwise
 * unusable): the sinc support must lie inside the stepped interior (the
 * reference's inverted check rejects every device), the dipole y weight
 * divides by dy (not dz), and the sinc dipole pole is negative as the linear
 * one is. Returns ELA_ERR_ARG for a device or support point out of bounds or
 * a bad method/half-length/dipole flag. */
int ela_interp_weights(const ela_interp_axes* ax, int64_t n_dev, const double* z, const double* x,
                       const double* y, int method, int half_length, int dipole, double z_shift,
                       double x_shift, double y_shift, int64_t* nbr, double* weight, int64_t* unique_pos,
                       int64_t* n_unique);

/* Time of the last propagation's fine-step loop on the GPU, measured with
 * HIP events on the handle's stream [ms], and the number of fine steps and
 * interior cells it advanced. */
int ela_last_timing(const ela_handle* h, double* ms, int64_t* fine_steps, int64_t* cells);

/* Several grids ("windows": e.g. the shot-local windows of a batch of
 * shots) stacked along y in the handle's nz x nx x ny grid: window w occupies
 * planes [y0[w], y0[w] + ny[w]) and x columns [0, nx[w]); windows do not
 * overlap and each has its own FAT layers and sponge (distances are measured
 * to the window's own faces), so that every propagation advances all windows
 * in the same launches and each window's result equals a propagation on that
 * window alone. Point positions must lie in a window's interior. The caller
 * stacks media, perturbations and positions (the Python WindowStack does).
 * n = 0 removes the windows. No counterpart in the reference (its shot-window
 * option `ginsu` is read but unused, nonlinearPropElasticShotsGpu_3D.cpp:12). */
int ela_set_windows(ela_handle* h, int n, const int64_t* y0, const int64_t* ny, const int64_t* nx);

/* Replay the time loops as HIP graphs: the coarse step of ela_nl_forward /
 * ela_nl_adjoint (sub fine steps, sub even; launched nts - 1 times, the
 * coarse-sample index advancing in device memory), pairs of forward steps of
 * ela_gradient_forward and of the gradient's inner recomputation, and each
 * full checkpoint segment of ela_gradient_backward (the forward steps with
 * derivative store, then the transposed steps with imaging; interval even),
 * their fine-step index and time-filter weights read from device memory.
 * A graph is captured and instantiated once and kept by the handle; a later
 * call replays it when its launches (kernels, launch shapes and argument
 * bytes) are the same, and captures a new one otherwise. Results are
 * bit-identical to host launches. No counterpart in the reference, which
 * launches every kernel from the host. */
int ela_set_graph(ela_handle* h, int on);

/* Tile shape of the stencil kernels: x extent of a 64-wide z tile (4, 8 or
 * 16 rows) and the number of y planes one block marches through. Any
 * setting gives the same result; it only changes speed. */
int ela_set_tiling(ela_handle* h, int tile_x, int y_chunk);
int ela_get_tiling(const ela_handle* h, int* tile_x, int* y_chunk);
/* Block-to-tile order of the stencil launches (speed only): 0 (default)
 * the hardware's round-robin dealing of blocks to the 8 XCDs over tiles
 * enumerated z, x, y; 1 each XCD a contiguous range of that enumeration; 2
 * each XCD a contiguous range of a y-major enumeration in groups of x_group
 * x columns, so that concurrently running neighbours share halo planes in
 * the XCD's L2; 3 the order-0 sweep with whole z columns (x_group adjacent
 * columns) dealt to one XCD, so that z and x neighbours share its L2. */
int ela_set_tile_order(ela_handle* h, int order, int x_group);

/* Diagnostic timing: launch the velocity and the stress stencil kernels of a
 * forward fine step `reps` times each on the handle's current wavefields and
 * return the mean duration of one launch of each [ms] (HIP events on the
 * handle's stream). The wavefields are overwritten; no counterpart in the
 * reference. */
int ela_time_step_kernels(ela_handle* h, int reps, double* ms_velocity, double* ms_stress);
/* Mean milliseconds of the three kernel pairs of an adjoint-state gradient
 * step on the handle's grid and launch shape, over `reps` launches each:
 * ms3[0] the forward step, ms3[1] the recomputing step with derivative store
 * (in the handle's derivative format), ms3[2] the transposed step with
 * imaging. Overwrites the wavefields and allocates the gradient's step
 * buffers (kept, as after a gradient; the Born adjoint's automatic storage
 * plan counts them as free and releases them); stored checkpoints are kept.
 * For the launch-shape tuner of the gradient. No counterpart in the
 * reference. */
int ela_time_gradient_kernels(ela_handle* h, int reps, double* ms3);
/* x extent (4, 8 or 16) of the tiles of the premultiplied transposed steps
 * (Born adjoint and gradient backward), which prefer another tile than the
 * forward steps on some grids; 0 (the default) uses the forward tile of
 * ela_set_tiling. The y chunk and block order are shared; results do not
 * depend on it. The getter returns the tile in use. No counterpart in the
 * reference. */
int ela_set_tiling_transposed(ela_handle* h, int tile_x);
int ela_get_tiling_transposed(const ela_handle* h, int* tile_x);

/* Diagnostic streaming reference for the roofline: a plain kernel that reads
 * `reads` arrays of n FP32 values (16-byte accesses) and rewrites the first
 * `writes` of them in place -- the byte mix of the velocity (12, 3) and stress
 * (14, 6) kernels without neighbour accesses, or (10, 3) and (11, 6) when they
 * form the medium from (rho, lambda, mu) -- timed over `reps` launches on the
 * handle's stream; *gbs = (reads + writes) * 4 * n bytes per launch / time.
 * Mixes: (1, 1), (12, 3), (14, 6), (10, 3), (11, 6), (12, 9), (12, 0), (2, 0);
 * n a multiple of 4. (12, 9) is the mix of a step that advances one time
 * sub-lattice of the leapfrog by a velocity and a stress half-step in one pass
 * (v_old, stress, rho, lambda, mu read; velocity and stress rewritten);
 * (12, 0) and (2, 0) only read. With reads * n * 4 bytes well above the
 * 256 MB Infinity Cache the figures are HBM's own.
 * Allocates and frees reads * n * 4 bytes. No counterpart in the reference. */
int ela_stream_probe(ela_handle* h, int reads, int writes, int64_t n, int reps, double* gbs);

#ifdef __cplusplus
}
#endif

#endif /* ELASTIC_ISO_3D_H */
