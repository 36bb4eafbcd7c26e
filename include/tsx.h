/*
 * libtsx: tomographic forward projection and backprojection on AMD MI355X
 * (gfx950), behind a plain C ABI.
 *
 * This library takes the place of the ASTRA projector that tomosipo drives.
 * tomosipo reaches ASTRA through exactly two calls, and each entry point
 * below names the call it replaces:
 *
 *   astra.create_projector("cuda3d", proj_geom, vol_geom, options)
 *       tomosipo/astra.py:80-98   -> tsx_plan_create / tsx_plan_destroy
 *   astra.experimental.direct_FPBP3D(projector, vol, proj, mode, "FP"|"BP")
 *       tomosipo/astra.py:104-153 -> tsx_direct_fpbp3d, tsx_forward, tsx_backward
 *
 * Geometry crosses the boundary in the form tomosipo serialises it for ASTRA:
 * the volume dictionary (geometry/volume.py:286-301: GridColCount = x,
 * GridRowCount = y, GridSliceCount = z, WindowMin/Max per axis) and the
 * vector projection dictionary (geometry/cone_vec.py:174-191,
 * geometry/parallel_vec.py:175-192: DetectorRowCount, DetectorColCount and
 * n x 12 Vectors [ray|src, det, u, v], each 3-vector in (x, y, z) order).
 * Parametric "parallel3d" / "cone" geometries are converted to vector form by
 * the caller, as tomosipo's to_vec() does.
 *
 * Data layout (tomosipo links/base.py:21-31, links/torch.py:94-99): volume
 * float32 [z][y][x], projections float32 [row][angle][col], both C-contiguous
 * in device memory on the plan's device. The library never allocates or frees
 * caller buffers. Work is enqueued on the caller's HIP stream and is
 * asynchronous; the plan is immutable after creation and may be shared by
 * threads.
 *
 * Every function returns TSX_OK (0) or a negative tsx_status; the message of
 * the last failure on the calling thread is returned by tsx_last_error().
 *
 * Tracing and failure injection: every entry point that enqueues device work
 * opens a roctx range named after itself (rocprofv3 --marker-trace shows
 * them). With TSX_FAIL_LAUNCH set to a comma-separated list of call numbers,
 * those calls of these entry points fail with TSX_ERR_HIP before enqueueing
 * anything (calls numbered from 1 while the variable keeps its value; the
 * count restarts when the value changes) -- for tests of failure handling.
 * With TSX_TIMING=1 each of these entry points times the device work it
 * enqueues with HIP events and writes "tsx: <entry point> <ms> ms" to stderr,
 * waiting for that work (a debugging aid; not while the stream is captured).
 *
 * Launch limits: batches, angle ranges and fused instance lists larger than
 * one kernel launch takes (65535 per grid dimension, 2^32 work-items per
 * grid) are split over several launches inside the call; results do not
 * depend on the split.
 *
 * Size limits: the fast kernels address their scratch layouts with 32-bit
 * byte offsets. A volume whose padded forward layouts would reach 4 GiB
 * (e.g. 1024^3) is forward-projected in z-slabs, each slab's ray sums added to
 * the projections; projections whose transposed, length-weighted copy would
 * reach 4 GiB are backprojected in ranges of angles, each range's sums added
 * to the volume. Chunked results equal the unchunked ones up to float32
 * rounding of the extra additions. TSX_CHUNK_BYTES, read by tsx_plan_create,
 * lowers the 4 GiB bound (for tests of these paths at small sizes).
 */
#ifndef TSX_H
#define TSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TSX_OK = 0,
    TSX_ERR_INVALID_ARGUMENT = -1, /* tomosipo raises ValueError/TypeError for these */
    TSX_ERR_UNSUPPORTED = -2,      /* a geometry or option this library does not implement */
    TSX_ERR_HIP = -3,              /* a HIP runtime call failed */
    TSX_ERR_NO_DEVICE = -4,        /* compute requested on a host-only plan */
    TSX_ERR_OUT_OF_MEMORY = -5
} tsx_status;

/* Projection geometry kinds (the ASTRA "type" strings). */
enum { TSX_PARALLEL3D_VEC = 0, TSX_CONE_VEC = 1 };

/* Mode constants of direct_FPBP3D (tomosipo/astra.py:132-135). */
enum { TSX_MODE_ADD = 0, TSX_MODE_SET = 1 };

/* Operation selector of direct_FPBP3D: "FP" / "BP". */
enum { TSX_OP_FP = 0, TSX_OP_BP = 1 };

typedef struct {
    /* projection dictionary */
    int32_t type;              /* TSX_PARALLEL3D_VEC or TSX_CONE_VEC */
    int32_t det_row_count;     /* DetectorRowCount (v, rows) */
    int32_t det_col_count;     /* DetectorColCount (u, columns) */
    int32_t n_angles;          /* rows of Vectors */
    const double *vectors;     /* n_angles x 12, row-major, ASTRA order and axes */
    /* volume dictionary */
    int32_t grid_col_count;    /* x */
    int32_t grid_row_count;    /* y */
    int32_t grid_slice_count;  /* z */
    double window_min[3];      /* (x, y, z) */
    double window_max[3];      /* (x, y, z) */
    /* projector options (astra.py:94-97) */
    int32_t det_supersampling; /* DetectorSuperSampling, >= 1 */
    int32_t vox_supersampling; /* VoxelSuperSampling; only 1 is supported (BP is the exact transpose) */
} tsx_geometry;

typedef struct tsx_plan tsx_plan;

/* Replaces astra.create_projector (tomosipo/astra.py:80-98). Validates the
 * geometry, computes the per-angle ray parameters in FP64 and uploads them to
 * `device`. device = -1 builds a host-only plan: validation and
 * tsx_plan_info work, compute calls return TSX_ERR_NO_DEVICE. */
int tsx_plan_create(const tsx_geometry *geometry, int device, tsx_plan **out);

/* Frees the plan's device tables. tomosipo never frees its projector id
 * (Operator.py:63-174); here the owner calls this. NULL is accepted. */
int tsx_plan_destroy(tsx_plan *plan);

typedef struct {
    int32_t device;
    int32_t n_angles, det_rows, det_cols;
    int64_t vol_shape[3];      /* (z, y, x) */
    int64_t plan_bytes;        /* device bytes held by the plan */
    int32_t axis_count[3];     /* number of angles stepping along x, y, z */
    int32_t fp_kernel;         /* 0 = general, 1 = padded per-axis layouts */
    int32_t bp_kernel;         /* 0 = general, 1 = separable (detector v parallel to a volume
                                  axis: z, or x or y through bp_axis_swap; any detector
                                  supersampling through bp_refined_detector) */
    int32_t fp_variant;        /* padded forward projector: 2 = per-lane slice ranges,
                                  3 = wave-synchronous slices, 4 = as 3 on row-paired
                                  layouts (one 16-byte load per sample), 5 = as 3 with
                                  each slice's window staged through LDS */
    int32_t bp_brick_z;        /* z-extent of the separable backprojector's bricks (16, 24 or 32) */
    int32_t tuned;             /* 1 once tsx_plan_autotune has chosen the variants */
    int32_t fp_slabs;          /* padded forward projector: z-slabs it projects the volume in
                                  (1 = whole volume; more when the layouts reach 4 GiB) */
    int32_t bp_launch_angles;  /* separable backprojector: angles per launch (all of them
                                  unless their transposed projections reach 4 GiB) */
    int32_t fp_tile_rows;      /* padded forward projector: detector rows per block (4, 8, 16) */
    int32_t bp_axis_swap;      /* separable backprojector run on the plan with volume axes swapped:
                                  0 = none, 1 = x and z (detector v parallel to x), 2 = y and z */
    int32_t fp_axis_swap;      /* padded forward projector run on that plan (the volume moved
                                  into its layout first): 0 = none, 1 or 2 as bp_axis_swap; the
                                  fp_* fields above then describe that plan's kernels */
    int32_t fp_shear_rows;     /* padded forward projector on a detector whose rows are not
                                  parallel to the slices' rows: each wave's rays follow the
                                  detector direction that keeps them on one row of the slices
                                  (sheared tiles, rows of a tile column taken cyclically), a
                                  lane's row offset at most this many rows; 0 for tiles along
                                  the detector rows */
    int32_t fp_column_waves;   /* padded forward projector: 1 if a wave's rays run along a
                                  detector column (every angle's rays keep to one row of the
                                  slices along a direction closer to v than to u; the shear
                                  of fp_shear_rows is then in columns), 0 along a row */
    int32_t fp_refined_detector; /* detector supersampling ss > 1 with the padded forward
                                  projector on the detector refined ss times: ss (each pixel
                                  the mean of its ss x ss fine pixels); 0 otherwise. The fp_*
                                  fields above then describe that plan's kernels */
    int32_t bp_refined_detector; /* detector supersampling ss > 1 with the separable
                                  backprojector: ss, the backprojection run on the detector
                                  refined ss times (its sub-rays as pixels); 0 otherwise. The
                                  bp_* fields above then describe that plan's kernels */
} tsx_plan_info;

int tsx_plan_info_get(const tsx_plan *plan, tsx_plan_info *info);

/* No counterpart in the reference (ASTRA has one kernel per geometry type).
 * Times the kernel variants of the plan's forward projector and
 * backprojector on its own geometry and keeps the fastest for later launches
 * of this plan: the forward march (row-paired, LDS-staged, plain or per-lane
 * on tiles of 8 detector rows, then tiles of 4 rows for the chosen march
 * where it takes them, kept if more than 3 % faster) on every angle over bands of 64 and 32
 * detector rows, compared on the predicted time of the whole plan (fixed
 * layout-building cost plus cost per row); for a detector whose rows cross
 * the slices' rows, sheared against plain tiles on the whole plan (sheared
 * kept unless plain ones are more than 2 % faster); the backprojector's brick depth
 * (32, 24 or 16 slices) on at most 64 of its angles, in runs of consecutive
 * angles from four quadrants of the scan. Trials use zero-filled scratch
 * buffers allocated and freed on `stream`. Choices are cached per (type,
 * detector, volume, angle count rounded up to a power of two) shape, so a
 * second plan of the same shape takes them without timing. Synchronises
 * `stream`. Must not run while launches of the same plan are being issued
 * from another thread. The TSX_FP_VARIANT, TSX_FP_TILE_ROWS, TSX_FP_SHEAR
 * (0 or 1) and TSX_BP_BRICK_Z environment variables, when set, override any
 * choice (TSX_BP_PARTS=1: no angle parts in the backprojector). Plans held by the plan (axis-swapped, refined detector) are tuned
 * too. */
int tsx_plan_autotune(tsx_plan *plan, void *stream);

/* Replaces astra.experimental.accumulate_FDK(projector, vol, proj)
 * (tomosipo/astra.py:374-406): FDK reconstruction of the projections `proj`
 * ([row][angle][col], float32) of a cone_vec plan into `vol` ([z][y][x]),
 * added to it when `additive` is non-zero, written otherwise. Cosine
 * pre-weighting on a virtual detector through the volume centre, Ram-Lak
 * ramp filter along detector rows (linear convolution of the zero-padded
 * row), voxel-driven bilinear backprojection weighted by (dbeta / 2) *
 * (D / U)^2, dbeta being the source's angular step around the z axis through
 * the volume centre. `workspace` holds the filtered projections
 * (tsx_fdk_workspace_size bytes); NULL takes stream-ordered device memory.
 * TSX_ERR_UNSUPPORTED for parallel-beam plans, detector supersampling or more
 * than 4096 detector columns. */
int tsx_fdk_workspace_size(const tsx_plan *plan, int64_t *bytes);
int tsx_fdk(const tsx_plan *plan, float *vol, const float *proj, int additive, void *workspace,
            int64_t workspace_bytes, void *stream);

/* Dominant (stepping) axis of every angle: 0 = x, 1 = y, 2 = z. */
int tsx_plan_axes(const tsx_plan *plan, int32_t *axes_out);

/* Replaces astra.experimental.direct_FPBP3D(projector, vol, proj, mode, op)
 * (tomosipo/astra.py:147-153). mode: TSX_MODE_SET overwrites the output,
 * TSX_MODE_ADD accumulates into it. op: TSX_OP_FP writes `proj` from `vol`,
 * TSX_OP_BP writes `vol` from `proj`. */
int tsx_direct_fpbp3d(const tsx_plan *plan, float *vol, float *proj, int mode, int op,
                      void *hip_stream);

/* Forward projection: proj (=|+=) A vol. Same kernel as direct_fpbp3d(FP). */
int tsx_forward(const tsx_plan *plan, const float *vol, float *proj, int additive,
                void *hip_stream);

/* Backprojection, the exact transpose of tsx_forward: vol (=|+=) A^T proj. */
int tsx_backward(const tsx_plan *plan, float *vol, const float *proj, int additive,
                 void *hip_stream);

/* Device scratch bytes that one launch of `op` over `batch` entries needs
 * (padded volume copies for the padded FP kernels, length-weighted
 * projections for BP without detector supersampling, plus, for volumes of
 * few bricks, the sums of the backprojector's angle parts; 0 for the general
 * FP kernel and for supersampled BP). */
int tsx_workspace_size(const tsx_plan *plan, int op, int64_t batch, int64_t *bytes);

/* Batched variants: `batch` independent volumes / projection sets with the
 * given element strides, in one launch. Replaces the per-slice loop of
 * tomosipo's autograd wrapper (torch_support.py:49-53, 70-74). `workspace`
 * (device memory of at least tsx_workspace_size bytes) may be NULL, in which
 * case the library takes stream-ordered memory for the call; passing it keeps
 * all allocation with the caller, e.g. under HIP graph capture. */
int tsx_forward_batched(const tsx_plan *plan, const float *vol, int64_t vol_stride, float *proj,
                        int64_t proj_stride, int64_t batch, int additive, void *workspace,
                        int64_t workspace_bytes, void *hip_stream);
int tsx_backward_batched(const tsx_plan *plan, float *vol, int64_t vol_stride, const float *proj,
                         int64_t proj_stride, int64_t batch, int additive, void *workspace,
                         int64_t workspace_bytes, void *hip_stream);

/* Fused launches over n independent instances (each its own plan, volume
 * and projections; shapes may differ; all plans on one device). Instances on
 * the fast kernels share one launch per MAX 16 instances, so the device never
 * idles between instances; the others run one launch each. Results equal
 * those of n separate tsx_forward / tsx_backward calls bit for bit. No
 * counterpart in the reference (tomosipo makes one synchronous projector call
 * per operator application, tomosipo/astra.py:147-153); for callers that
 * project many volumes at once, such as multi-volume throughput jobs and the
 * request server. `workspace` as for the batched calls, sized by
 * tsx_multi_workspace_size. */
int tsx_multi_workspace_size(const tsx_plan *const *plans, int n, int op, int64_t *bytes);
int tsx_multi_forward(const tsx_plan *const *plans, const float *const *vols, float *const *projs,
                      int n, int additive, void *workspace, int64_t workspace_bytes,
                      void *hip_stream);
int tsx_multi_backward(const tsx_plan *const *plans, float *const *vols, const float *const *projs,
                       int n, int additive, void *workspace, int64_t workspace_bytes,
                       void *hip_stream);

/* Epilogue of one instance of a fused call: how the projector's result r is
 * written to that instance's output o.
 *   TSX_EP_SET:        o = r
 *   TSX_EP_ADD:        o += r
 *   TSX_EP_RESIDUAL:   o = w * (y - r)     (forward calls only; y, w like o)
 *   TSX_EP_SCALED_ADD: o += w * r          (w like o)
 * The residual and scaled-add forms are the two halves of a SIRT iteration
 * (see tsx_forward_residual / tsx_backward_scaled_add). */
enum { TSX_EP_SET = 0, TSX_EP_ADD = 1, TSX_EP_RESIDUAL = 2, TSX_EP_SCALED_ADD = 3 };
typedef struct {
    int32_t mode;
    const float *y;
    const float *w;
} tsx_epilogue;

/* As tsx_multi_forward / tsx_multi_backward with a per-instance epilogue, so
 * that one fused launch carries instances that overwrite, accumulate, or run
 * a SIRT half-step. No counterpart in the reference; for the request server,
 * whose fused launches mix plain projections with SIRT iterations. */
int tsx_multi_forward_ep(const tsx_plan *const *plans, const float *const *vols, float *const *projs,
                         const tsx_epilogue *eps, int n, void *workspace, int64_t workspace_bytes,
                         void *hip_stream);
int tsx_multi_backward_ep(const tsx_plan *const *plans, float *const *vols, const float *const *projs,
                          const tsx_epilogue *eps, int n, void *workspace, int64_t workspace_bytes,
                          void *hip_stream);

/* The two halves of one SIRT iteration x += C * A^T(R * (y - A x)) with the
 * elementwise work fused into the projection kernels. They replace the
 * projector calls plus elementwise operations of tomosipo's SIRT loops
 * (README.md:150-164, doc/intro/fast_reconstruction.rst:92-104,
 * notebooks/sirt_benchmark.py:131-136).
 *   tsx_forward_residual:    proj = w * (y - A vol)      (w = R, same layout as proj)
 *   tsx_backward_scaled_add: vol += w * (A^T proj)       (w = C, same layout as vol)
 * y and w use the batch stride of the output. */
int tsx_forward_residual(const tsx_plan *plan, const float *vol, int64_t vol_stride, const float *y,
                         const float *w, float *proj, int64_t proj_stride, int64_t batch,
                         void *workspace, int64_t workspace_bytes, void *hip_stream);
int tsx_backward_scaled_add(const tsx_plan *plan, float *vol, int64_t vol_stride,
                            const float *proj, int64_t proj_stride, const float *w, int64_t batch,
                            void *workspace, int64_t workspace_bytes, void *hip_stream);

/* FP64 forward and backprojection (general kernels, any geometry). No
 * counterpart in the reference: ASTRA projects float32 only. For callers that
 * need the operator and its transpose to FP64 rounding, such as
 * finite-difference gradient checks of the autograd wrapper. */
int tsx_forward_batched_f64(const tsx_plan *plan, const double *vol, int64_t vol_stride,
                            double *proj, int64_t proj_stride, int64_t batch, int additive,
                            void *hip_stream);
int tsx_backward_batched_f64(const tsx_plan *plan, double *vol, int64_t vol_stride,
                             const double *proj, int64_t proj_stride, int64_t batch, int additive,
                             void *hip_stream);

/* Bytes of the static device scratch buffer of tsx_direct_fpbp3d_noalloc. */
#define TSX_NOALLOC_SCRATCH_BYTES (256ll << 20)

/* No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. Same arguments, kernels and results as
 * tsx_direct_fpbp3d, but the kernels' workspace (tsx_workspace_size(plan,
 * op, 1) bytes) is taken from a static device buffer of
 * TSX_NOALLOC_SCRATCH_BYTES that the library holds on every device it runs on
 * until the process ends; the call allocates and frees no device memory.
 * Plans whose workspace exceeds that buffer fail with TSX_ERR_UNSUPPORTED
 * (e.g. a forward projection of a 256^3 volume with two stepping axes).
 * Every call on a device shares that device's buffer: calls that overlap on
 * one device (several host threads, or several streams in flight) race, so
 * the caller serialises them. */
int tsx_direct_fpbp3d_noalloc(const tsx_plan *plan, float *vol, float *proj, int mode, int op,
                              void *hip_stream);

/* Message describing the last failure on this thread ("" if none). */
const char *tsx_last_error(void);

/* Library version string. */
const char *tsx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* TSX_H */
