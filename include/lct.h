/*
 * lct.h -- C ABI of the MI355X helical cone-beam FDK library (liblct.so).
 *
 * This is the boundary below the reference's C++ stage classes: the drop-in
 * classes of listctrecons_amd/cpp (FileParams, ReconData, Rebiner, Filter,
 * Backprojector, PreWeighting, FDKReconstructor, ReconstrctorAbstractFactory)
 * call only these functions, and so do the Python tests and bench.py.
 * Plain pointers and sizes; no HIP, CUDA or torch types appear here.
 *
 * Conventions
 *  - Every function returns LCT_OK (0) or a negative LCT_ERR_* code; the
 *    message of the last failure on the calling thread is lct_last_error().
 *  - d_* arguments are device pointers (from lct_device_malloc or any HIP
 *    allocation on the context's device); h_* arguments are host pointers.
 *  - Work is stream-ordered on the context's stream (the null stream unless
 *    lct_ctx_set_stream was called) and asynchronous; lct_synchronize waits.
 *    Host arrays passed to a call (angles, couch, taps, parameters) are
 *    copied into the context's pinned staging ring (16 MB) before the call
 *    returns, so the caller may reuse them at once; the copies reach the
 *    device in stream order.  Device buffers must stay valid until the
 *    stream has finished with them.
 *  - Array layouts follow the reference: projections [view][row][channel],
 *    volumes [z][y][x], all row-major, channel / x fastest.
 */
#ifndef LCT_H
#define LCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define LCT_OK 0
#define LCT_ERR_INVALID (-1)     /* bad argument or shape */
#define LCT_ERR_HIP (-2)         /* HIP runtime error (message names the call) */
#define LCT_ERR_NOMEM (-3)       /* device allocation failed */
#define LCT_ERR_RANGE (-4)       /* requested reconstruction outside the data */
#define LCT_ERR_UNSUPPORTED (-5) /* option not available for this geometry */

/* ------------------------------------------------------------- enums */
enum { LCT_F32 = 0, LCT_F16 = 1, LCT_BF16 = 2 };
enum { LCT_REBIN_NULL = 0, LCT_REBIN_FAN2PAR = 1, LCT_REBIN_PI_ORI = 2 };
enum { LCT_WEIGHT_NONE = 0, LCT_WEIGHT_PI_ORI = 1 };
enum { LCT_WINDOW_NONE = 0, LCT_WINDOW_HANN = 1, LCT_WINDOW_SHEPP_LOGAN = 2 };
enum { LCT_FILTER_DIRECT_F32 = 0, LCT_FILTER_TOEPLITZ_BF16X3 = 1, LCT_FILTER_TOEPLITZ_BF16X6 = 2 };
enum { LCT_BP_HELICAL = 0, LCT_BP_PI_ORI = 1 };
/* Bilinear taps of the backprojector: LDS = view windows staged in LDS and
 * interpolated explicitly; GATHER = taps loaded from HBM through the vector
 * L1/L2 caches.  (gfx950 exposes no texture sampler to HIP: the image API is
 * compiled out for this target.) */
enum { LCT_INTERP_LDS = 0, LCT_INTERP_GATHER = 1 };
enum { LCT_PIPE_BLOCKS = 0, LCT_PIPE_RESIDENT = 1 };

/* ------------------------------------------------------------ geometry */
/* CTGeom, Include/DataStructs.h:9-100 (same fields, same meaning). */
typedef struct {
    float orbit, orbit_start;
    int ns, nt;
    float ws, wt;
    int n_angle, n_angle_add;
    float ds, dt, pitch, source_z0, dso;
    int n_proj_turn, z_dir, add_projections;
} lct_ct_geom;

/* ImgGeom, Include/DataStructs.h:102-156. */
typedef struct {
    int nx, ny, nz;
    float wx, wy, wz;
    float dx, dy, dz;
} lct_img_geom;

/* The numeric keys of a .prm file, Src/FileParams.cpp:10. */
typedef struct {
    int ns, nt;
    float ds, dt, pitch_value;
    int nx, ny;
    float dx, dz, offset_x, offset_y;
    int readings, view_offset, data_offset;
    float r_src_to_iso, central_channel;
    int n_proj_turn;
} lct_prm;

/* ReconInfo (Include/DataStructs.h:225-273) with the derived geometry. */
typedef struct {
    lct_ct_geom ctg;
    lct_img_geom ig;
    int n_readings;
    float data_begin_pos, data_end_pos, allow_begin, allow_end;
    float start_pos, end_pos, recon_start_pos, recon_end_pos;
    int n_slices_recon, n_slices_block, n_blocks, data_offset;
    int recon_direction;
} lct_plan;

/* ------------------------------------------------------------- context */
typedef struct lct_ctx lct_ctx;

/* Creates a context on HIP device `device`.  The reference has no explicit
 * context: its state lives in globals (texture references Rebiner.cu:11,
 * Backprojector.cu:10) and in CTDataGPU (DataStructs.cpp:57-66). */
int lct_ctx_create(int device, lct_ctx **out);
int lct_ctx_destroy(lct_ctx *ctx);
/* stream: a hipStream_t passed as void* (NULL = null stream). */
int lct_ctx_set_stream(lct_ctx *ctx, void *stream);
int lct_synchronize(lct_ctx *ctx);
const char *lct_last_error(void);
/* Library version string, e.g. "lct 1.0 gfx950". */
const char *lct_version(void);

/* Device memory (backs GPUMemoryCtrl<T>, Include/GPUMemoryCtrl.h:6-62). */
int lct_device_malloc(lct_ctx *ctx, void **d_ptr, size_t bytes);
int lct_device_free(lct_ctx *ctx, void *d_ptr);
int lct_memcpy_h2d(lct_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int lct_memcpy_d2h(lct_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
int lct_memcpy_d2d(lct_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);
int lct_memset_zero(lct_ctx *ctx, void *d_dst, size_t bytes);

/* ------------------------------------------------------------ planner */
/* ReconData::convert_from_file_params, Src/DataStructs.cpp:121-250, host
 * only.  h_angles / h_couch hold prm->readings values already past
 * ViewOffset.  Returns LCT_ERR_RANGE where the reference prints "Requested
 * reconstruction is outside of allowed data range" and exits. */
int lct_plan_helical(const lct_prm *prm, const float *h_angles, const float *h_couch,
                     lct_plan *out);
/* BlockInfo::update, Src/DataStructs.cpp:77-108 (block 0: :222-238).
 * ref_view: the view whose table position the block's slice 0 is aligned to
 * (array_search(blockStart)), for mapping slices to table positions. */
int lct_block_info(const lct_plan *plan, const float *h_couch, int block, int *proj_start,
                   int *proj_end, float *block_start, float *block_end, int *ref_view);
/* Projection-file views [*view_begin, *view_end) that blocks
 * [block_begin, block_end) read: the union of their BlockInfo windows
 * (Src/DataStructs.cpp:77-108) with DataOffset applied (RawDataReader,
 * Src/DataReader.cpp:14).  Host only.  What a caller streaming a scan slab
 * by slab uploads for lct_reconstruct_slabs. */
int lct_slab_views(const lct_plan *plan, const float *h_couch, int block_begin, int block_end,
                   long *view_begin, long *view_end);

/* -------------------------------------------------------- stage kernels
 * Per-block semantics of the reference stages: one call processes one
 * 32-slice block on the arrays ReconData holds for it. */

/* RampFilterGPU::calculate_filter, Src/Filter.cpp:8-20 (window NONE gives
 * the reference taps); Hann / Shepp-Logan apodisation replaces the empty
 * Include/SmoothWindow.h.  h_taps: 2*ns floats, host. */
int lct_ramp_taps(int window, int ns, float ds, float *h_taps);

/* Rebiner::rebin, CUDA/Rebiner.cu:13-29 (NULL), :82-103 (FAN2PAR),
 * :156-179 (PI_ORI), with PiOriPreWeighting (CUDA/PreWeighting.cu:8-29)
 * fused into the epilogue when weight == LCT_WEIGHT_PI_ORI.
 * d_fan [n_angle_add][nt][ns] f32 -> d_par [n_angle][nt][ns] f32. */
int lct_rebin(lct_ctx *ctx, int kind, int weight, const lct_ct_geom *g, const float *d_fan,
              float *d_par);

/* PiOriPreWeighting::weighting, CUDA/PreWeighting.cu:31-44, in place on
 * d_par [n_angle][nt][ns].  kind LCT_WEIGHT_NONE is NULLPreWeighting. */
int lct_weighting(lct_ctx *ctx, int kind, const lct_ct_geom *g, float *d_par);

/* Filter::filter, CUDA/Filter.cu:52-73: q[r][is] = sum_k h[is-k+ns] row_r[k]
 * over rows = n_angle*nt rows of ns channels, in place, FP32. */
int lct_filter(lct_ctx *ctx, const lct_ct_geom *g, const float *h_taps, float *d_rows,
               long rows);

/* The ramp convolution of Filter::filter (CUDA/Filter.cu:10-50) written in
 * the resident backprojector's layout: d_rows [views][nt][ns] f32 ->
 * d_q [views][ns + 3][nt + 8] of `storage` type, channel-major, channel c at
 * slot c + 1 and row r at slot r + 4, every other slot 0 (the caller clears
 * d_q).  algo LCT_FILTER_DIRECT_F32 (FP32 FMA, k ascending),
 * LCT_FILTER_TOEPLITZ_BF16X3 (Toeplitz GEMM on BF16 MFMA, taps and data split
 * in two bf16 parts, three products, FP32 accumulation) or
 * LCT_FILTER_TOEPLITZ_BF16X6 (both split in three parts, six products:
 * FP32-level accuracy).  Needs nt % 8 == 0 (and ns % 4 == 0 for the Toeplitz
 * algorithms). */
int lct_filter_to_bp_layout(lct_ctx *ctx, int algo, int storage, int ns, int nt,
                            const float *h_taps, const float *d_rows, long views, void *d_q);

/* Backprojector::backproject, CUDA/Backprojector.cu:101-169 (HELICAL) and
 * :252-307 (PI_ORI): d_q [n_angle][nt][ns] f32, h_betas the block's tube
 * angles (n_angle + n_proj_turn/2 + 1 values; see the conjugate-index note in
 * the implementation), d_img [nz][ny][nx] f32. */
int lct_backproject(lct_ctx *ctx, int kind, const lct_ct_geom *g, const lct_img_geom *ig,
                    const float *d_q, const float *h_betas, float *d_img);

/* lct_backproject without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device or pinned memory (code under HIP
 * graph capture, tight serving loops).  Same arguments, kernels and results
 * as lct_backproject; its tables live in a fixed device array of the library
 * (128.4 KiB per GPU, held until the process ends) and reach it through
 * kernel arguments.  Limit: n_angle + n_proj_turn/2 + 1 <= 16384, else
 * LCT_ERR_UNSUPPORTED.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams or contexts) race; calls in one stream
 * are ordered. */
int lct_backproject_noalloc(lct_ctx *ctx, int kind, const lct_ct_geom *g, const lct_img_geom *ig,
                            const float *d_q, const float *h_betas, float *d_img);

/* ------------------------------------------------ whole-scan pipeline */
typedef struct {
    int pipeline;    /* LCT_PIPE_BLOCKS: the reference's per-block loop
                        (Src/Reconstructor.cpp:33-48) on device; LCT_PIPE_RESIDENT:
                        every view rebinned and filtered once, volume
                        backprojected slab by slab from HBM-resident views */
    int rebin_kind;  /* LCT_REBIN_* */
    int weight_kind; /* LCT_WEIGHT_* */
    int window;      /* LCT_WINDOW_* */
    int filter_algo; /* LCT_FILTER_* (RESIDENT only; BLOCKS uses DIRECT_F32) */
    int storage;     /* LCT_F32 / LCT_F16 / LCT_BF16: filtered-projection
                        storage of the RESIDENT backprojector */
    int bp_kind;     /* LCT_BP_*; on RESIDENT both kinds fold the view pairs
                        (j, j + npt/2): HELICAL with the conjugate weights,
                        PI_ORI taking view j where its ray lies in the voxel's
                        PI window and view j + npt/2 otherwise (no INTERP_GATHER
                        form: PI_ORI with INTERP_GATHER backprojects block by
                        block, F32 only) */
    int interp;      /* LCT_INTERP_* (RESIDENT only) */
} lct_recon_opts;

/* Stage times of the last lct_reconstruct on this context (HIP events). */
typedef struct {
    float rebin_ms, filter_ms, backproject_ms, total_ms;
    int n_backproject_launches;
    long long backproject_updates; /* voxel-view updates (one per view in a
                                      voxel's window; for LCT_BP_PI_ORI half
                                      a turn of views per voxel) */
    float backproject_kernel_ms;   /* of backproject_ms, the backprojector
                                      launches alone (without the pass that
                                      adds the partial sums of split blocks) */
} lct_timings;

/* FDKReconstructor::reconstruct (Src/Reconstructor.cpp:30-49) for the whole
 * planned range.  d_fan: the projection file's views [file_views][nt][ns]
 * resident in device memory (view 0 = file view 0; the plan's DataOffset is
 * applied here).  d_vol: [n_slices_recon][ny][nx] f32.  Device scratch is
 * owned by the context and reused across calls. */
int lct_reconstruct(lct_ctx *ctx, const lct_plan *plan, const lct_recon_opts *opts,
                    const float *h_angles, const float *h_couch, const float *d_fan,
                    long file_views, float *d_vol);
/* lct_reconstruct for blocks [block_begin, block_end) of the plan only (the
 * reference's per-block loop, Src/Reconstructor.cpp:33-48, over a range).
 * d_fan holds projection-file views [fan_view0, fan_view0 + fan_views),
 * which must cover lct_slab_views of the range (LCT_ERR_RANGE otherwise);
 * d_vol receives (block_end - block_begin) * 32 slices, its slice 0 being
 * slice 32 * block_begin of the plan.  Every slice equals the same slice of
 * lct_reconstruct: a long scan can be streamed through HBM slab by slab and
 * checkpointed between slabs. */
int lct_reconstruct_slabs(lct_ctx *ctx, const lct_plan *plan, const lct_recon_opts *opts,
                          const float *h_angles, const float *h_couch, const float *d_fan,
                          long fan_view0, long fan_views, int block_begin, int block_end,
                          float *d_vol);
/* The filtered views [n_angle][nt][ns] (f32) of the last block the last
 * lct_reconstruct / lct_reconstruct_slabs call on this context reconstructed:
 * what Filter::filter leaves in CTDataCPU::rebin (CUDA/Filter.cu:69) and
 * RawImageWriter writes to rebin.raw (Src/ImageWriter.cpp:27-40).  With
 * f16 / bf16 storage they carry that rounding.  Valid until the next
 * reconstruct call on the context. */
int lct_last_block_filtered(lct_ctx *ctx, float *d_rows);
int lct_get_timings(lct_ctx *ctx, lct_timings *out);
/* Enables HIP-event timing of every stage launch (off by default). */
int lct_set_profiling(lct_ctx *ctx, int enable);

/* ------------------------------------------- synthetic scan generator */
/* Exact line integrals of additive ellipsoids (16 doubles each: centre,
 * semi-axes, row-major body->world rotation, attenuation) for a curved-
 * detector helical scan; the source of view v sits at
 * (-dso sin b, dso cos b, couch[v]), b = angles[v].  Not part of the
 * reference (it ships no data); it produces every test and benchmark input.
 * d_out [n_views][nt][ns] f32. */
int lct_project_ellipsoids(lct_ctx *ctx, const lct_ct_geom *g, const float *h_angles,
                           const float *h_couch, long view0, long n_views,
                           const double *h_ell, int n_ell, float *d_out);
/* In place p -> -ln(max(n,1)/I0), n ~ Poisson(I0 exp(-p)); counter-based,
 * reproducible for a given seed. */
int lct_poisson_noise(lct_ctx *ctx, float *d_proj, size_t n, float i0, uint64_t seed);

/* The same ellipsoids sampled at voxel centres x = (ix - wx) dx,
 * y = (iy - wy) dy, z = z0 + iz dz (supersample^3 sub-samples averaged):
 * d_vol [nz][ny][nx] f32 (ig->nz is ignored).  Training targets of the
 * learned correction. */
int lct_voxelize_ellipsoids(lct_ctx *ctx, const lct_img_geom *ig, double z0, int nz,
                            const double *h_ell, int n_ell, int supersample, float *d_vol);

/* Reprojection of a voxel volume d_vol [vol_nz][ny][nx] (voxel centres
 * x = (ix - wx) dx, y = (iy - wy) dy, z = vol_z0 + iz dz; nz of vg is
 * ignored) along the rays of lct_project_ellipsoids: trilinear samples at the
 * midpoints of equal steps of at most `step` mm inside the voxel-centre box.
 * d_out [n_views][nt][ns].  Not part of the reference: it scores the learned
 * correction's reprojection residual. */
int lct_forward_project(lct_ctx *ctx, const lct_ct_geom *g, const float *h_angles,
                        const float *h_couch, long view0, long n_views, const float *d_vol,
                        const lct_img_geom *vg, float vol_z0, int vol_nz, float step,
                        float *d_out);

/* Adjoint (transpose) of lct_forward_project over the same rays, samples and
 * trilinear weights: d_vol [vol_nz][ny][nx] = A^T d_proj, d_proj
 * [n_views][nt][ns].  d_vol is overwritten.  Each sample's weighted value is
 * added with float atomics, so the summation order, and the last bits, vary
 * from run to run.  Not part of the reference: with lct_forward_project it
 * forms a matched projector pair, <A x, y> = <x, A^T y>. */
int lct_forward_project_adjoint(lct_ctx *ctx, const lct_ct_geom *g, const float *h_angles,
                                const float *h_couch, long view0, long n_views,
                                const float *d_proj, const lct_img_geom *vg, float vol_z0,
                                int vol_nz, float step, float *d_vol);

/* ------------------------------------------- learned artifact correction
 * Not part of the reference.  A 3D residual CNN, y = x + net(x), net = six
 * 3x3x3 convolutions 1 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1 with ReLU after the
 * first five, zero 'same' padding.  FP32 master weights packed as
 * [W1 27x32 | b1 32 | (W 27x32x32 | b 32) x 4 | W6 27x32 | b6 1],
 * W[tap][co][ci], tap = (dz+1)*9 + (dy+1)*3 + (dx+1); bf16 operand copies
 * are refreshed after every update.  Volumes are f32 [B][D][H][W]. */
typedef struct lct_cnn lct_cnn;
long lct_cnn_param_count(void);
int lct_cnn_create(lct_ctx *ctx, uint64_t seed, lct_cnn **out);
int lct_cnn_destroy(lct_cnn *net);
int lct_cnn_get_params(lct_cnn *net, float *h_params);
int lct_cnn_set_params(lct_cnn *net, const float *h_params);
/* Gradients of the last lct_cnn_loss_grad, same packing as the parameters. */
int lct_cnn_get_grads(lct_cnn *net, float *h_grads);
int lct_cnn_forward(lct_cnn *net, const float *d_x, int B, int D, int H, int W, float *d_y);
/* Forward, loss = mean((y - t)^2), and backward into the gradient buffer. */
int lct_cnn_loss_grad(lct_cnn *net, const float *d_x, const float *d_t, int B, int D, int H,
                      int W, float *h_loss);
/* One Adam update of the master weights from the stored gradients. */
int lct_cnn_adam_step(lct_cnn *net, float lr, float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif
#endif /* LCT_H */
