/*
 * lct_oracle.h -- FP64 CPU restatement of the helical FDK hot path of
 * LISTCTRecons (planner, rebin, pre-weighting, ramp filter, backprojection)
 * plus an exact analytic ellipsoid projector used to make test inputs.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (liblct.so, the C++ drop-in classes, the CLI) never links or calls
 * it and has no CPU fallback.
 *
 * Arithmetic convention: every discrete decision of the reference (planner
 * integer casts, ceil/floor, array_search, the roundf() that picks a voxel's
 * central view) is evaluated in float32 exactly as the reference evaluates it,
 * so the oracle selects the same views and texels.  Every continuous quantity
 * (trigonometry, interpolation weights, convolution and backprojection sums)
 * is evaluated in double.  The CUDA texture fetch tex3D(x+0.5,...) with
 * cudaFilterModeLinear / cudaAddressModeBorder is emulated with exact weights
 * and zero outside the array.
 *
 * Struct layouts are plain C and match the C-ABI structs of include/lct.h
 * field for field, so the same bytes can be handed to both libraries.
 */
#ifndef LCT_ORACLE_H
#define LCT_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors CTGeom (Include/DataStructs.h:9-100). */
typedef struct {
    float orbit, orbit_start;
    int ns, nt;
    float ws, wt;
    int n_angle, n_angle_add;
    float ds, dt, pitch, source_z0, dso;
    int n_proj_turn, z_dir, add_projections;
} lco_ct_geom;

/* Mirrors ImgGeom (Include/DataStructs.h:102-156). */
typedef struct {
    int nx, ny, nz;
    float wx, wy, wz;
    float dx, dy, dz;
} lco_img_geom;

/* The numeric .prm keys of FileParams (Src/FileParams.cpp:10). */
typedef struct {
    int ns, nt;
    float ds, dt, pitch_value;
    int nx, ny;
    float dx, dz, offset_x, offset_y;
    int readings, view_offset, data_offset;
    float r_src_to_iso, central_channel;
    int n_proj_turn;
} lco_prm;

/* Mirrors ReconInfo (Include/DataStructs.h:225-273) and block 0 of BlockInfo. */
typedef struct {
    lco_ct_geom ctg;
    lco_img_geom ig;
    int n_readings;
    float data_begin_pos, data_end_pos, allow_begin, allow_end;
    float start_pos, end_pos, recon_start_pos, recon_end_pos;
    int n_slices_recon, n_slices_block, n_blocks, data_offset;
    int recon_direction;
} lco_plan;

/* status codes */
#define LCO_OK 0
#define LCO_ERR_RANGE (-2)   /* requested range outside the allowed data */
#define LCO_ERR_ARG (-1)

/* Src/DataStructs.cpp:12-34 (bounds checked before the read). */
int lco_array_search(float key, const float *array, int n, int search_type);

/* Src/DataStructs.cpp:121-250.  angles/couch hold `readings` float32 values
 * (already past ViewOffset). */
int lco_plan_helical(const lco_prm *prm, const float *angles, const float *couch,
                     lco_plan *out);

/* Src/DataStructs.cpp:77-108 (and :222-238 for block 0). */
int lco_block_info(const lco_plan *plan, const float *couch, int block_id,
                   int *proj_start, int *proj_end, float *block_start, float *block_end);

/* Src/Filter.cpp:8-20 with an apodisation window (Include/SmoothWindow.h is
 * empty in the reference; window 0 reproduces the reference taps).
 * window: 0 = none (Ram-Lak), 1 = Hann, 2 = Shepp-Logan.  h has 2*ns taps. */
void lco_ramp_taps(int ns, double ds, int window, double *h);

/* CUDA/Rebiner.cu:82-103.  fan [NA+][nt][ns] float32, par [NA][nt][ns]. */
void lco_rebin_fan2par(const lco_ct_geom *g, const float *fan, double *par);
/* CUDA/Rebiner.cu:13-29. */
void lco_rebin_null(const lco_ct_geom *g, const float *fan, double *par);
/* CUDA/Rebiner.cu:156-179. */
void lco_rebin_pi_ori(const lco_ct_geom *g, const float *fan, double *par);
/* CUDA/PreWeighting.cu:8-29, in place on [NA][nt][ns]. */
void lco_weight_pi_ori(const lco_ct_geom *g, double *par);

/* CUDA/Filter.cu:10-50: q[r][is] = sum_{k<ns} h[is-k+ns] row_r[k]. */
void lco_filter_rows(int ns, const double *h, const double *in, double *out, long rows);

/* CUDA/Backprojector.cu:12-99.  q [NA][nt][ns], betas [NA] float32.
 * Computes slices iz_list[0..n_iz) and rows [iy0,iy1) of img [nz][ny][nx]. */
void lco_bp_helical(const lco_ct_geom *g, const lco_img_geom *ig, const double *q,
                    const float *betas, const int *iz_list, int n_iz, int iy0, int iy1,
                    double *img);
/* CUDA/Backprojector.cu:171-250. */
void lco_bp_pi_ori(const lco_ct_geom *g, const lco_img_geom *ig, const double *q,
                   const float *betas, const int *iz_list, int n_iz, int iy0, int iy1,
                   double *img);

/* Whole per-block pipeline of Src/Reconstructor.cpp:33-48.
 * fan_file: the whole projection file as float32 [views][nt][ns] (views
 *   counted from file start, i.e. DataOffset is applied here).
 * angles: the tube angles (Readings values past ViewOffset).
 * rebin_kind: 0 null, 1 fan2par, 2 pi_ori;  weight_kind: 0 none, 1 pi_ori;
 * bp_kind: 0 helical (conjugate-ray), 1 pi_ori; window as lco_ramp_taps.
 * Blocks [b0,b1) and, inside each block, slices iz_list / rows [iy0,iy1) are
 * computed into vol [n_slices_recon][ny][nx].  Returns 0 or a negative code. */
int lco_reconstruct(const lco_plan *plan, const float *fan_file, long file_views,
                    const float *angles, const float *couch, int rebin_kind,
                    int weight_kind, int window, int bp_kind, int b0, int b1,
                    const int *iz_list, int n_iz, int iy0, int iy1, double *vol);

/* ---------------- analytic phantom (test-input generator) ----------------
 * Ellipsoid record: cx cy cz (mm), ax ay az (semi-axes, mm), R (row-major
 * 3x3 rotation, body->world), mu (1/mm, additive): 16 doubles each. */
#define LCO_ELL_STRIDE 16

/* Curved (cylindrical, source-centred) detector helical scan.
 * Ray of fan view g, row it, channel is: source at
 *   S = (-dso sin b, dso cos b, z_g),  b = angles[g], z_g = couch[g];
 * in-plane unit direction (sin(b-a), -cos(b-a)),  a = (is - ws) * ds / dso;
 * slope dz/dl = (it - wt) * dt / dso (row index grows with +z).
 * out [n_views][nt][ns] line integrals in double. */
void lco_project_ellipsoids(const lco_ct_geom *g, const float *angles, const float *couch,
                            long view0, long n_views, const double *ell, int n_ell,
                            double *out);

/* Phantom sampled at voxel centres: x = (ix - wx)dx, y = (iy - wy)dy, z = z0 + iz*dz. */
void lco_voxelize(const lco_img_geom *ig, double z0, int nz, const double *ell, int n_ell,
                  int supersample, double *vol);

/* Reprojection of a voxel volume vol[nz][ny][nx] (voxel centres as above,
 * trilinear interpolation, zero outside) along the rays of
 * lco_project_ellipsoids: midpoint rule over n = ceil(L / step) equal steps of
 * the ray's chord L through the box of voxel centres.  Not in the reference:
 * the definition the build's lct_forward_project implements (reprojection
 * residual of the learned correction).  out [n_views][nt][ns]. */
void lco_forward_project(const lco_ct_geom *g, const float *angles, const float *couch,
                         long view0, long n_views, const double *vol, const lco_img_geom *ig,
                         double z0, int nz, double step, double *out);

/* Adjoint (transpose) of lco_forward_project: vol = A^T proj for the same rays,
 * samples and trilinear weights.  vol [nz][ny][nx] is overwritten. */
void lco_forward_project_adjoint(const lco_ct_geom *g, const float *angles, const float *couch,
                                 long view0, long n_views, const double *proj,
                                 const lco_img_geom *ig, double z0, int nz, double step,
                                 double *vol);

int lco_num_threads(void);
void lco_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
