/*
 * dpt -- C ABI of the MI355X cone-beam projector.
 *
 * Plain pointers and sizes only: device pointers come from the caller's
 * allocator (PyTorch, hipMalloc, ...), streams are hipStream_t passed as
 * void*.  Every entry point returns 0 on success or a negative DPT_E_* code;
 * dpt_last_error() then names the failure (thread-local).
 *
 * The reference binds this path through Warp launches behind the Torch custom
 * operators dpt::project / dpt::project_vjp / dpt::project_out
 * (python/dpt/torch/_projection_ops.py:256-411).  Each entry point below
 * replaces one piece of that stack; the citation is given per declaration.
 *
 * Layouts (identical to the reference):
 *   field      (nz, ny, nx) contiguous, x fastest, mm^-1
 *              (python/dpt/volumes.py:17-25)
 *   poses      (views, 12) float64: row-major object-to-world rotation, then
 *              translation in mm (python/dpt/geometry.py:96-98)
 *   images     (views, rows, cols) optical depth, row-major
 *   grad_pose  (views, 12) float64 packed-pose cotangent, or (views, 6)
 *              local-right tangent (tx,ty,tz,rx,ry,rz) with DPT_POSE_TANGENT
 */
#ifndef DPT_AMD_H
#define DPT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPT_ABI_VERSION 1

/* return codes */
#define DPT_OK 0
#define DPT_E_INVALID -1  /* an argument violates the contract */
#define DPT_E_DEVICE -2   /* a HIP runtime call failed */
#define DPT_E_MEMORY -3   /* device or pinned allocation failed */

/* sample arithmetic (python/dpt/torch/modules.py:123-134) */
#define DPT_ARITH_FP32 0  /* kernels/projection_fp32.py */
#define DPT_ARITH_FP64 1  /* kernels/projection.py, sampling_precision=float64 */

/* resident field storage; FP32 sampling reads every kind */
#define DPT_STORE_F32 0
#define DPT_STORE_BF16 1
#define DPT_STORE_F16 2
/* BF16/FP16 field buffers must stay readable for this many bytes past the
 * last voxel (x-adjacent pairs are fetched with one 8-byte load). */
#define DPT_HALF_PAD_BYTES 8
/* Cell layout, OR-ed into a DPT_STORE_* code of a resident (fit) volume:
 * the field is stored as interpolation cells, cell (i, j, k) for
 * i < max(nx-1, 1), j < max(ny-1, 1), k < max(nz-1, 1) (x fastest) holding
 * its 8 corner values contiguously, corner q = (x + (q&1), y + (q>>1&1),
 * z + (q>>2&1)) clamped to the grid.  One cell is 32 bytes (FP32) or 16
 * bytes (BF16/FP16): a trilinear sample is one or two 16-byte loads.  Build
 * it with dpt_convert_cells. */
#define DPT_STORE_CELLS 8
/* dpt_plan_create `needs` option: FP32 sampling launches gather corners
 * from a plan-owned cell-layout copy of the field, rebuilt from the linear
 * field at each dpt_forward / dpt_vjp call (the caller's API is unchanged). */
#define DPT_GATHER_CELLS 64u
/* dpt_plan_create `needs` option (DPT_ARITH_FP64 only): exact cell-Gauss
 * integration (kernels/projection.py:522-746, ProjectionSpec.integration =
 * "cell_gauss", python/dpt/projection.py:107-135) instead of midpoints: the
 * clipped ray is cut at every interpolation plane and each piece integrated
 * with two Gauss nodes; `samples` is then unused. */
#define DPT_CELL_GAUSS 128u

/* dpt_vjp flags */
#define DPT_VOLUME 1u        /* produce grad_field */
#define DPT_POSE 2u          /* produce grad_pose */
#define DPT_POSE_TANGENT 4u  /* 6-component local-right tangent instead of 12 */
#define DPT_ACCUMULATE 8u    /* add into grad_pose / grad_field instead of writing */
#define DPT_NO_CHECKS 16u    /* skip input validation (caller validated once) */
/* grad_field is float64 (volume_gradient_precision="float64",
 * python/dpt/projection.py:871-1018): FP32-sampled increments RN(f * w) are
 * not flushed, FP64-sampled ones are not rounded to FP32 */
#define DPT_VOLUME_FP64 32u

/* status bits (python/dpt/contracts.py:229-234, projection.py:505-515) */
#define DPT_STATUS_POSE 1    /* invalid rigid pose -> ContractError */
#define DPT_STATUS_RANGE 2   /* non-finite output/gradient -> NumericalError */
#define DPT_STATUS_FAILURE 4 /* numerical execution failure -> NumericalError */
#define DPT_STATUS_DOMAIN 8  /* field not finite/non-negative -> TrialDomainError */

/* Grid: kernels/projection.py:51-65 Configuration.{shape, object_to_grid,
 * grid_origin}; object_to_grid = diag(1/spacing) * orientation^T
 * (python/dpt/projection.py:571-576). */
typedef struct {
    int32_t nx, ny, nz;
    double object_to_grid[9];
    double grid_origin[3];
} dpt_grid_desc;

typedef struct dpt_plan dpt_plan;

/*
 * Prepare a multiview plan: geometry metadata on the device, the status words,
 * pose-reduction partials and the volume accumulator.  Replaces
 * prepare_projection (python/dpt/projection.py:526-624) called once per view
 * by _projection_ops.prepare (python/dpt/torch/_projection_ops.py:127-175).
 *
 * panels: host float64 (views, 12) = source, detector origin (pixel (0,0)
 *   centre), column step (u * column pitch), row step (v * row pitch), or NULL.
 * ray_sources / ray_endpoints: host float64 (views, rows*cols, 3) explicit
 *   world segments (RayBundleGeometry, python/dpt/geometry.py:225-300), used
 *   when panels is NULL.
 * needs: DPT_VOLUME | DPT_POSE, the gradients later calls may request.
 */
int dpt_plan_create(const dpt_grid_desc *grid, int32_t views, int32_t rows, int32_t cols,
                    const double *panels, const double *ray_sources, const double *ray_endpoints,
                    int32_t samples, int32_t arithmetic, int32_t storage, uint32_t needs,
                    int32_t device, dpt_plan **out);

/*
 * Prepare a multiview plan from per-view panel records, expanded into one
 * world segment per pixel on the device (flat cone, parallel and cylindrical
 * panels may share one batch).  Replaces prepare_geometry
 * (python/dpt/geometry_device.py:190-233) with the generate kernel of
 * kernels/batched_geometry.py:9-39, for BatchedGeometry views of
 * DetectorGeometry / ParallelGeometry / CurvedDetectorGeometry
 * (python/dpt/geometry.py:101-460).
 * panel_records: host float64 (views, 18) = source (for parallel panels the
 *   unit beam direction), pixel (0,0) endpoint, u, v, column pitch, row
 *   pitch, length (parallel: source distance behind each pixel; cylindrical:
 *   radius), radial vector (cylindrical: pixel (0,0) minus source, with its
 *   v component removed).
 * panel_modes: host int32 (views): 0 flat cone, 1 parallel, 2 cylindrical.
 * The image is (views, rows, cols) as for dpt_plan_create; other arguments
 * as there.  Plans made here gather with every pixel of a view in the
 * adjoint gather (no projected-window culling).
 */
int dpt_plan_create_panels(const dpt_grid_desc *grid, int32_t views, int32_t rows, int32_t cols,
                           const double *panel_records, const int32_t *panel_modes, int32_t samples,
                           int32_t arithmetic, int32_t storage, uint32_t needs, int32_t device,
                           dpt_plan **out);

/* Release every device and pinned buffer of the plan. */
int dpt_plan_destroy(dpt_plan *plan);

/*
 * Optical depth of every view: out (views, rows, cols), float32 for
 * DPT_ARITH_FP32 and float64 for DPT_ARITH_FP64.  Enqueues the field and pose
 * checks and one multiview ray-march launch on `stream`; no host sync.
 * Replaces the per-view loop of _forward (python/dpt/torch/_projection_ops.py:
 * 213-253) around project_optical_depth (python/dpt/projection.py:747-809)
 * and the kernels of kernels/projection_fp32.py:334-374 and
 * kernels/projection.py:802-846.
 * field: float32 (nz,ny,nx) for DPT_STORE_F32, else bf16/f16 storage.
 */
int dpt_forward(dpt_plan *plan, const void *field, const double *poses, void *out, void *stream);

/*
 * Vector-Jacobian product for a cotangent (views, rows, cols) of the output
 * dtype.  grad_field: float32 (nz,ny,nx) (DPT_VOLUME), float64 with
 * DPT_VOLUME_FP64; grad_pose: float64 (views, 12) or (views, 6) with
 * DPT_POSE_TANGENT (DPT_POSE).
 * Replaces project_vjp (python/dpt/torch/_projection_ops.py:273-353) around
 * projection_vjp (python/dpt/projection.py:871-1018), the VJP kernels
 * kernels/projection_fp32.py:377-617 / kernels/projection.py:873-1237 and the
 * reduction tree (python/dpt/_reductions.py:27-86).
 */
int dpt_vjp(dpt_plan *plan, const void *field, const double *poses, const void *cotangent,
            void *grad_field, double *grad_pose, uint32_t flags, void *stream);

/*
 * Ray-coordinate VJP: out_rays (views, rows, cols, 6) float64 = cotangent[p]
 * x d(optical depth)/d(world source xyz, world endpoint xyz) of each ray,
 * stored per ray (no reduction), through the plan's arithmetic and
 * integration rule (clipping, sample motion and the segment norm included).
 * cotangent as for dpt_vjp; flags: DPT_NO_CHECKS.  Replaces
 * projection_ray_vjp (python/dpt/batched_projection.py:37-111) over the
 * ray_coordinates branch of kernels/projection_fp32.py:570-583 and
 * kernels/projection.py:1184-1200.
 */
int dpt_ray_vjp(dpt_plan *plan, const void *field, const double *poses, const void *cotangent,
                double *out_rays, uint32_t flags, void *stream);

/*
 * Per-ray pose sensitivities: out_jacobian (views, rows, cols, 6) float64 =
 * cotangent[p] x d(optical depth p)/d(delta) for pose @ exp(delta^), rows
 * ordered tx, ty, tz, rx, ry, rz (mm, rad), one row per ray with no
 * reduction; unit cotangents export the unweighted Jacobian, missed rays and
 * zero cotangents give zero rows.  Through the plan's arithmetic and
 * integration rule (the clip-face and interpolation branches included).
 * cotangent as for dpt_vjp; flags: DPT_NO_CHECKS.  Replaces
 * projection_pose_sensitivities (python/dpt/projection.py:1021-1084), the
 * per_ray branch of the VJP kernels (kernels/projection_fp32.py:377-617,
 * kernels/projection.py:873-1237).
 */
int dpt_pose_sensitivities(dpt_plan *plan, const void *field, const double *poses, const void *cotangent,
                           double *out_jacobian, uint32_t flags, void *stream);

/*
 * out (views, components) float64 = per view, sum over its pixels of
 * rays[v, p, :] . jacobians[v, k, :] (rays from dpt_ray_vjp; jacobians
 * device float64 (views, components, 6): d(source, endpoint)/d(parameter k)
 * in world mm), in a fixed order; DPT_ACCUMULATE adds to out.  With the cone
 * intrinsic Jacobians (SOD, SDD, column offset, row offset) this is the
 * intrinsics VJP of batched_projection_vjp (python/dpt/batched_projection.py:
 * 178-209, 433-539; kernels/batched_projection.py:15-89).
 */
int dpt_reduce_rays(dpt_plan *plan, const double *rays, const double *jacobians, int32_t components,
                    double *out, uint32_t flags, void *stream);

/*
 * Deterministic adjoint gather: out[v] = sum over views, pixels (row-major),
 * midpoint samples and cell corners, in that order, of A(poses)[p, v] *
 * cotangent[p], summed in FP64 by one thread per voxel, so the result is the
 * same bit for bit on every call (unlike the atomic volume VJP).  Requires a
 * DPT_ARITH_FP64 plan.  cotangent: float64 (views, rows, cols); out: (nz, ny,
 * nx) float64 (out_precision DPT_ARITH_FP64) or float32 (DPT_ARITH_FP32),
 * rounded once from the FP64 sum.  flags: DPT_ACCUMULATE adds the previous
 * out[v] after the sum; DPT_NO_CHECKS skips the cotangent/pose validation.
 * A non-finite cotangent (DPT_STATUS_RANGE) or a non-rigid pose
 * (DPT_STATUS_POSE) leaves out untouched.  The plan allocates a ray table of
 * 112 bytes per pixel at the first call.  Enqueued on `stream`, no host sync.
 * Replaces prepare_projection_adjoint + projection_adjoint_gather
 * (python/dpt/projection.py:1087-1405) and the kernels of
 * kernels/projection_adjoint.py:73-170 and :296-635.
 */
int dpt_adjoint_gather(dpt_plan *plan, const double *poses, const double *cotangent, void *out,
                       int32_t out_precision, uint32_t flags, void *stream);

/*
 * dpt_adjoint_gather without any allocation: the same kernels and the same
 * result bit for bit, with the ray table, view table and guard word in a
 * fixed device workspace of DPT_GATHER_NOALLOC_BYTES that the library holds
 * on every GPU it runs on until the process ends (no hipMalloc, hipFree or
 * allocating library call, at the first call or later).  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The plan needs 112 bytes per ray plus 256-byte-aligned view and guard
 * records within the workspace (one 512 x 512 view fits; larger plans get
 * DPT_E_INVALID and should use dpt_adjoint_gather).  All calls on one GPU
 * share the workspace: calls that overlap on one GPU (other streams, other
 * threads) race and must be serialised by the caller.  Arguments, flags and
 * status behaviour as for dpt_adjoint_gather.
 */
#define DPT_GATHER_NOALLOC_BYTES (32u << 20)
int dpt_adjoint_gather_noalloc(dpt_plan *plan, const double *poses, const double *cotangent, void *out,
                               int32_t out_precision, uint32_t flags, void *stream);

/*
 * Copy the per-view status words to host_codes (views int32) with one D2H
 * transfer and one stream synchronisation; clear them when `clear` != 0.
 * Replaces StatusCheckpoint.check/consume (python/dpt/_runtime.py:317-445).
 */
int dpt_status(dpt_plan *plan, int32_t *host_codes, int32_t clear, void *stream);

/* Convert float32 values to bf16 (storage DPT_STORE_BF16) or f16 with
 * round-to-nearest-even, flagging non-finite or negative inputs with
 * DPT_STATUS_DOMAIN in status[0] when status is not NULL.  No reference
 * counterpart: the reference has no FP32-sampling half-storage form
 * (python/dpt/projection.py:131-136 forbids it). */
int dpt_convert_field(const float *src, void *dst, int64_t count, int32_t storage, int32_t *status,
                      void *stream);

/* Cells of a grid in the DPT_STORE_CELLS layout:
 * max(nx-1,1) * max(ny-1,1) * max(nz-1,1); each holds 8 values. */
int64_t dpt_cell_count(int32_t nx, int32_t ny, int32_t nz);

/* Rewrite a linear (nz,ny,nx) field of `storage` (DPT_STORE_F32/BF16/F16)
 * into the DPT_STORE_CELLS layout of the same element type: dst holds
 * 8 * dpt_cell_count(nx, ny, nz) elements.  Enqueued on `stream`.  No
 * reference counterpart: the reference gathers the 8 corners from the linear
 * array (kernels/projection_fp32.py:157-168). */
int dpt_convert_cells(const void *src, void *dst, int32_t nx, int32_t ny, int32_t nz, int32_t storage,
                      void *stream);

/* Kernel time of the most recent launch of `kind` (0 forward, 1 VJP,
 * 2 adjoint gather including its ray preparation) on the
 * plan, in milliseconds, measured with HIP events on the launch stream;
 * negative when no timed launch has completed.  Enable with dpt_set_timing. */
int dpt_set_timing(dpt_plan *plan, int32_t enabled);
double dpt_kernel_ms(dpt_plan *plan, int32_t kind);

/* Number of rays whose FP32 clipped interval is non-empty for these poses
 * (effective samples = hits * samples).  Synchronises the stream; returns a
 * negative DPT_E_* code on failure.  Measurement helper, no reference
 * counterpart. */
int64_t dpt_count_hit_rays(dpt_plan *plan, const double *poses, void *stream);

/* VALU issue-rate probe: 16 independent chains of v_fma_f32 (kind 0) or
 * v_add_u32 (kind 1) per lane, 8 waves per SIMD; *rate = wave64 VALU
 * instructions per second over the device (one warm and one timed launch,
 * HIP events on `stream`, synchronised).  scratch_out: device float buffer
 * of 2048 * 256 elements.  The ceiling the projector kernels' instruction
 * rates are reported against; measurement helper, no reference
 * counterpart. */
int dpt_valu_probe(int32_t kind, void *scratch_out, double *rate, void *stream);

/* SE(3) chart G = anchor @ exp(hat(xi)) for `count` pose pairs on
 * `stream`: anchor 12 FP64 per pair (row-major rotation, translation mm),
 * tangent 6 FP64 (tx, ty, tz, rx, ry, rz), out 12 FP64; jet receives the
 * increment exp(hat(xi)) and its six partial derivatives (84 FP64 per
 * pair) for dpt_compose_pose_vjp.  Replaces the matrix_exp chart of
 * python/dpt/torch/geometry.py:19-63 (compose_pose) on the device. */
int dpt_compose_pose(const double *anchor, const double *tangent, int64_t count, double *out, double *jet,
                     void *stream);

/* Gradients of that chart for a cotangent grad_out (12 FP64 per pair) into
 * grad_anchor (12 per pair) and grad_tangent (6 per pair), from the jet of
 * dpt_compose_pose (the backward of python/dpt/torch/geometry.py:19-63). */
int dpt_compose_pose_vjp(const double *anchor, const double *jet, const double *grad_out, int64_t count,
                         double *grad_anchor, double *grad_tangent, void *stream);

/* The launch layout a plan chose: out[0] FP64 LDS brick entries of the
 * volume adjoint, out[1] 1 when the FP64 volume adjoint chunks by slabs of
 * grid samples instead of runs of sample indices, out[2] samples per chunk,
 * out[3] views per workgroup of the VJP launches, out[4] chunks per ray of
 * the cell-Gauss volume adjoint.  Inspection helper, no reference
 * counterpart. */
int dpt_plan_layout(const dpt_plan *plan, int32_t *out);

const char *dpt_last_error(void);
int32_t dpt_abi_version(void);

/* ------------------------------------------------------------------------
 * Pose-fit service (batched solves on resident volumes).  No reference
 * counterpart: the reference runs one solve per process
 * (python/dpt/pose_optimization.py:753-1006, python/dpt/registration.py).
 * ---------------------------------------------------------------------- */

/* A resident volume: device field in DPT_STORE_* storage plus its grid. */
typedef struct {
    const void *field;
    int32_t storage;
    int32_t nx, ny, nz;
    double object_to_grid[9];
    double grid_origin[3];
} dpt_volume_desc;

/* One active solve: `views` cone panels (device float64 (views, 12), as for
 * dpt_plan_create), measured optical depth `targets` (device float32
 * (views, rows, cols)) and its device solver state (DPT_FIT_STATE_SIZE
 * float64, layout below; the caller initialises the pose and zeroes the
 * rest). */
typedef struct {
    int32_t volume; /* index into the volume table passed with the step */
    int32_t views, rows, cols, samples;
    const double *panels;
    const float *targets;
    double *state;
} dpt_fit_request;

#define DPT_FIT_POSE 0        /* 12: pose evaluated by the next step */
#define DPT_FIT_BEST_POSE 12  /* 12: lowest-loss pose so far (the result) */
#define DPT_FIT_BEST_LOSS 24  /* 1/2 sum of squared residuals at BEST_POSE */
#define DPT_FIT_LAMBDA 25     /* Levenberg-Marquardt damping */
#define DPT_FIT_EVALS 26      /* evaluations done */
#define DPT_FIT_ACCEPTED 27   /* accepted steps */
#define DPT_FIT_JTJ 28        /* 21: J^T J (upper, row-major) at BEST_POSE */
#define DPT_FIT_JTR 49        /* 6: J^T r at BEST_POSE */
#define DPT_FIT_LAST_LOSS 55  /* loss of the last evaluation */
#define DPT_FIT_STATE_SIZE 56

typedef struct dpt_fit dpt_fit;

/* Allocate the device tables for up to max_requests (<= 256) solves per step
 * and max_tiles 16x16 detector tiles in total. */
int dpt_fit_create(int32_t device, int32_t max_requests, int64_t max_tiles, dpt_fit **out);
int dpt_fit_destroy(dpt_fit *fit);

/* Advance every listed solve by one Levenberg-Marquardt iteration: one FP32
 * march per ray gives the optical depth (residual against the target) and the
 * local-right pose Jacobian row; normal equations are summed in FP64 in a
 * fixed order per request and solved on the device.  With use_graph != 0 the
 * step runs as a captured hipGraph (one per grid size and stream; the stream
 * must not be the legacy default stream).  No host synchronisation. */
int dpt_fit_step(dpt_fit *fit, const dpt_volume_desc *volumes, int32_t n_volumes,
                 const dpt_fit_request *requests, int32_t n_requests, int32_t use_graph,
                 void *stream);
const char *dpt_fit_last_error(void);

/* ------------------------------------------------------------------------
 * Monte Carlo photon transport (analogue Woodcock tracking).
 *
 * Replaces the reference's prepare_transport / simulate pair
 * (python/dpt/transport/forward.py:146-264, transport/simulation.py:186-316)
 * with the walk of transport/kernels.py:941-1563 and the point-cone source of
 * transport/kernels.py:2339-2366.  Everything is binary64; random numbers are
 * the reference's counter-addressed Philox4x32-10 (kernels/random.py:27-60),
 * so a history's path depends only on (seed, history identity).
 * ---------------------------------------------------------------------- */

#define DPT_TRANSPORT_MAX_MATERIALS 8
#define DPT_TRANSPORT_MAX_TABLE_BYTES 49152 /* all tables, staged in LDS */

/* Terminal history statuses (transport/model.py:36-45). */
#define DPT_HISTORY_ESCAPED 0
#define DPT_HISTORY_ABSORBED 1
#define DPT_HISTORY_EVENT_BUDGET 2
#define DPT_HISTORY_CROSSING_BUDGET 3
#define DPT_HISTORY_NUMERICAL_FAILURE 4
#define DPT_HISTORY_ENERGY_SUPPORT 5
#define DPT_HISTORY_ANGLE_BUDGET 6

/* Host-side model; every table pointer is host memory, copied at create.
 * Field meanings follow the reference's Parameters / PhysicalParameters
 * structs (transport/kernels.py:105-133, physical_kernels.py:20-45). */
typedef struct dpt_transport_model {
    int32_t nx, ny, nz, materials, energy_bins, spectrum_bins;
    int32_t max_events, max_crossings, max_angle_trials;
    int32_t compton;        /* 0 isotropic elastic, 1 free-electron Klein-Nishina */
    int32_t compton_packed; /* two proposals per Philox block */
    int32_t energy_score;   /* score keV instead of photon counts */
    int32_t physics;        /* Rayleigh channel, spectrum and majorant bound */
    int32_t woodcock;       /* majorant present (requires physics) */
    int32_t det_shape[2];   /* width (columns), height (rows) */
    double origin[3], spacing[3]; /* lower cell faces and cell size, mm */
    double det_corner[3], det_u[3], det_v[3], det_normal[3], det_spacing[2]; /* object frame */
    double inv_rot[9], inv_trans[3]; /* world -> object (inverse object pose) */
    double source_energy, cutoff;    /* keV */
    const uint8_t *material_ids;     /* nx*ny*nz labels, x fastest */
    const double *energies;          /* K increasing nodes, keV */
    const double *absorption;        /* M*K, energy fastest, mm^-1 per unit density */
    const double *scattering;        /* M*K (coherent + incoherent) */
    const double *rayleigh;          /* M*K coherent part (physics) */
    const double *majorant;          /* K (woodcock) */
    const double *majorant_max_density; /* M: 1 where the material occurs, else 0 */
    const int32_t *q_offsets;        /* M+1 offsets into q / form_factor (physics) */
    const double *q, *form_factor;   /* momentum transfer (1/angstrom), F(q) */
    const double *atomic_number;     /* M */
    const double *spectrum_energies, *spectrum_cdf; /* spectrum_bins each */
} dpt_transport_model;

/* Point emission aimed at a uniform point of the detector rectangle
 * corner + U*extent[0]*u + V*extent[1]*v (transport/source.py:54-182). */
typedef struct dpt_point_source {
    double source[3], corner[3], u[3], v[3], extent[2], weight;
} dpt_point_source;

/* Optional per-history device outputs (trace_histories, kernels.py:1944-1994,
 * plus the real-interaction count that tells primary from scatter). */
typedef struct dpt_history_records {
    int32_t *pixel;
    double *score;
    double *energy;
    int32_t *events;
    int32_t *real_events;
    int32_t *status;
} dpt_history_records;

typedef struct dpt_transport dpt_transport;

/* Validate the model and upload labels (as 8-bit) and tables. */
int dpt_transport_create(const dpt_transport_model *model, int32_t device, dpt_transport **out);
int dpt_transport_destroy(dpt_transport *t);

/* Trace histories first_history .. first_history+count-1 (count < 2^31).
 * density: host array of `materials` scales (rate = density*table).  With a
 * majorant, every occurring material must satisfy density*(a+s) < M at every
 * node (transport/kernels.py:1850-1872), else DPT_E_INVALID.  Scores of
 * detected histories accumulate into tally (device, 4*pixels doubles: primary
 * sum, primary sum of squares, scatter sum, scatter sum of squares) when it
 * is non-NULL; records (device pointers, each optional) receive per-history
 * results.  Terminal-status counts accumulate on the device until read. */
int dpt_transport_trace(dpt_transport *t, const double *density, const dpt_point_source *source,
                        uint64_t seed, uint64_t first_history, int64_t count, double amplitude,
                        const dpt_history_records *records, double *tally, void *stream);

/* Per-pixel mean and variance of the mean over `histories` original histories
 * (transport/simulation.py:61-66): mean = sum/N, var = (sumsq - sum^2/N) /
 * (N(N-1)).  Outputs are device arrays of 3*pixels (total, primary, scatter). */
int dpt_transport_moments(dpt_transport *t, const double *tally, int64_t histories, double *mean,
                          double *variance, void *stream);

/* Synchronise the stream, read the seven terminal-status counts and the OR of
 * 1 << status over failed histories (status >= 2, plus 1 << 4 for a
 * non-finite score) accumulated since the last clear. */
int dpt_transport_status(dpt_transport *t, int64_t *histogram, int32_t *failure_bits, int32_t clear,
                         void *stream);
const char *dpt_transport_last_error(void);


#ifdef __cplusplus
}
#endif

#endif /* DPT_AMD_H */
