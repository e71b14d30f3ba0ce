/*
 * plenvdb_hip.h -- C ABI of libplenvdb_hip.so, the MI355X (gfx950) library
 * behind the PlenVDB sparse-VDB radiance-field hot path.
 *
 * Conventions
 *  - Every pointer argument marked "dev" is a device pointer on the current
 *    HIP device; "host" pointers are read synchronously.  `stream` is a
 *    hipStream_t (NULL = the legacy default stream).  Calls are asynchronous
 *    on that stream unless stated otherwise.
 *  - Every function returns 0 on success or a negative code:
 *    -1 invalid argument, -2 HIP error, -3 bad handle, -4 out of memory,
 *    -5 size out of range.  pvdb_last_error() returns the calling thread's
 *    last message.
 *  - Payloads are slot-major: value c of slot s is at vals[s*C + c], where a
 *    slot is leaf*512 + ((x&7)<<6 | (y&7)<<3 | (z&7)) and leaves are numbered
 *    in NanoVDB depth-first order (the reference's `n`, densityvdb.cu:41-42).
 *    A 12-channel colour payload keeps channel c of the reference's Vec3f
 *    grid c/3, component c%3 (colorvdb.cu:59,96-108).
 *
 * Reference interfaces cited below are paths in wolfball/PlenVDB.
 */
#ifndef PLENVDB_HIP_H
#define PLENVDB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PVDB_ABI_VERSION 1

typedef struct pvdb_topology pvdb_topology;

/* Thread-local message of the last failing call (plenvdb/lib/vdb/plenvdb.cu:3-17
 * prints CUDA errors; this library returns them). */
const char* pvdb_last_error(void);
int pvdb_abi_version(void);

/* ---- topology ------------------------------------------------------------
 * Replaces ScaleGrid/VectorGrid::create (plenvdb/lib/vdb/plenvdb.h:117-125,
 * 197-210: OpenVDB denseFill -> openToNanoVDB -> deviceUpload).  `mask` is a
 * host dense [rx][ry][rz] array of active voxels; a leaf exists for every 8^3
 * block holding at least one active voxel (an all-true mask gives denseFill's
 * topology).  Synchronous; allocates the device leaf table. */
int pvdb_topo_create_from_mask(const uint8_t* mask /*host*/, int32_t rx, int32_t ry, int32_t rz,
                               pvdb_topology** out);
/* Leaf list (origins multiples of 8, 8 value-mask words per leaf); the
 * library sorts it into NanoVDB order (openvdb/nanovdb/nanovdb/util/
 * OpenToNanoVDB.h:509-533).  layout: 0 automatic, 1 one dense leaf table,
 * 2 lower-node table + per-lower-node child tables (for leaves spread over a
 * large extent).  Synchronous. */
int pvdb_topo_create_from_leaves(const int32_t* origins /*host [n][3]*/,
                                 const uint64_t* masks /*host [n][8]*/, int64_t nleaf,
                                 int32_t layout, pvdb_topology** out);
int pvdb_topo_destroy(pvdb_topology* topo);
/* tree().nodeCount(0) and active voxel count (plenvdb.h:123: nleafCount). */
int pvdb_topo_info(const pvdb_topology* topo, int64_t* nleaf, int64_t* nactive,
                   int32_t* leaf_lo3, int32_t* leaf_dim3, int32_t* two_level);
/* Leaf origins and value masks in NanoVDB order (tree().getFirstNode<0>()+i). */
int pvdb_topo_leaves(const pvdb_topology* topo, int32_t* origins /*host*/,
                     uint64_t* masks /*host*/);
/* Host slot query: slot of each voxel (or -1 when its leaf does not exist) and
 * its active flag (LeafNode::CoordToOffset, NanoVDB.h:3893-3900;
 * ReadAccessor::isActive, NanoVDB.h:4660-4674). */
int pvdb_topo_slots(const pvdb_topology* topo, const int32_t* ijk /*host [n][3]*/, int64_t n,
                    int64_t* slots /*host*/, uint8_t* active /*host, may be NULL*/);
int pvdb_topo_device_arrays(const pvdb_topology* topo, const int32_t** table,
                            const uint64_t** masks, const int32_t** origins);
/* No counterpart in the reference (it steps through empty space voxel by
 * voxel, render_utils_kernel.cu:373-424, renderer.cu:222-367): the empty-run
 * map the marchers skip with, one byte per 8^3 cell of the leaf grid box
 * (pvdb_topo_info's leaf_lo3 / leaf_dim3, C order): 0 where a leaf exists,
 * else r >= 1 with no leaf within Chebyshev distance r - 1 (the smallest such
 * r, capped at 64).  Copies it to `out` (host, may be NULL) and returns its
 * size, 0 for the two-level layout, which has none. */
int64_t pvdb_topo_empty_run(const pvdb_topology* topo, uint8_t* out /*host*/);

/* ---- grid payload kernels ------------------------------------------------
 * Trilinear forward.  order 0: density rule of density_forward_kernel
 * (plenvdb/lib/vdb/densityvdb.cu:101-125, replaces DensityVDB::forward
 * plenvdb.h:441-466 / plenvdb.cpp:11-19); order 1: colour rule of
 * color_forward_kernel (colorvdb.cu:81-111, ColorVDB::forward plenvdb.h:535-556).
 * xs/ys/zs are index-space points (grid.py:77-85); out is [n][channels]. */
int pvdb_trilinear_fwd(const pvdb_topology* topo, const float* vals /*dev*/, int32_t channels,
                       int32_t order, const float* xs /*dev*/, const float* ys /*dev*/,
                       const float* zs /*dev*/, int64_t n, float* out /*dev*/, void* stream);
/* Atomic scatter with the "leaf must exist" rule: density_backward_kernel
 * (densityvdb.cu:143-181, DensityVDB::backward plenvdb.h:468-485) and
 * color_backward_kernel (colorvdb.cu:130-175, ColorVDB::backward plenvdb.h:557-573). */
int pvdb_trilinear_bwd(const pvdb_topology* topo, float* grad /*dev*/, int32_t channels,
                       int32_t order, const float* xs, const float* ys, const float* zs,
                       const float* g /*dev [n][channels]*/, int64_t n, void* stream);
/* copyFromDense (densityvdb.cu:31-61, colorvdb.cu:42-77; plenvdb.cpp:37-39,72-76). */
int pvdb_copy_from_dense(const pvdb_topology* topo, float* vals, int32_t channels,
                         const float* dense /*dev [rx][ry][rz][C]*/, int32_t rx, int32_t ry,
                         int32_t rz, void* stream);
/* get_dense_grid (plenvdb.h:149-160,242-262; plenvdb.cpp:33). */
int pvdb_copy_to_dense(const pvdb_topology* topo, const float* vals, int32_t channels,
                       float* dense /*dev*/, int32_t rx, int32_t ry, int32_t rz, float background,
                       void* stream);
/* setValuesOn_bymask (densityvdb.cu:64-97; plenvdb.cpp:7-9). */
int pvdb_set_values_by_mask(const pvdb_topology* topo, float* vals, const uint8_t* mask /*dev*/,
                            float value, int32_t rx, int32_t ry, int32_t rz, void* stream);
/* Total-variation gradient added into grad [nslots][channels]: the dense
 * kernel total_variation_add_grad_cuda_kernel (total_variation_kernel.cu:13-35,
 * weights divided by 6 as in :45-47; the x axis takes wz there and here) on the
 * grid's dense view over [0,rx)x[0,ry)x[0,rz), for every slot in that box.
 * Replaces the VDB entry the reference stubs (DensityVDB/ColorVDB::
 * total_variation_add_grad, plenvdb.h:497-501,574-578). */
int pvdb_tv_add_grad(const pvdb_topology* topo, const float* vals /*dev*/, float* grad /*dev*/,
                     int32_t channels, int32_t rx, int32_t ry, int32_t rz, float wx, float wy,
                     float wz, int32_t dense_mode, void* stream);
/* zero_grad (densityvdb.cu:353-368, colorvdb.cu:355-371; plenvdb.cpp:94,120). */
int pvdb_zero_active(const pvdb_topology* topo, float* grad, int32_t channels, void* stream);
/* DensityOpt/ColorOpt::step_optimizer kernels (plenvdb.h:751-789;
 * densityvdb.cu:185-349, colorvdb.cu:179-351).  group = values per VDB value
 * (1 density, 3 colour Vec3f); mode 0 plain, 1 skip-zero-grad, 2 per-voxel lr. */
int pvdb_vdb_adam(const pvdb_topology* topo, float* param, const float* grad, float* exp_avg,
                  float* exp_avg_sq, const float* per_lr /*dev [slots] or NULL*/,
                  int32_t channels, int32_t group, int32_t mode, float stepsz, float eps,
                  float beta0, float beta1, void* stream);
/* adam_upd / masked_adam_upd / adam_upd_with_perlr
 * (plenvdb/lib/cuda/adam_upd_kernel.cu:8-132, adam_upd.cpp:79-86). */
int pvdb_dense_adam(float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                    const float* per_lr, int64_t n, int32_t mode, float step_size, float eps,
                    float beta0, float beta1, void* stream);

/* ---- FP64 twins of the hand-written adjoint pairs ------------------------
 * Same formulas as pvdb_trilinear_fwd/_bwd (densityvdb.cu:101-181,
 * colorvdb.cu:81-175; channels 1, 3 or 12, slot-major), pvdb_raw2alpha/_bwd
 * (render_utils_kernel.cu:430-443,506-517) and pvdb_alpha2weight/_bwd
 * (:576-603,653-676) with every quantity in double, for central
 * finite-difference checks of the adjoints on the device.  No counterpart in
 * the reference, which checks none of its adjoints. */
int pvdb_trilinear_fwd_f64(const pvdb_topology* topo, const double* vals, int32_t channels,
                           const double* xs, const double* ys, const double* zs, int64_t n,
                           double* out, void* stream);
int pvdb_trilinear_bwd_f64(const pvdb_topology* topo, double* grad, int32_t channels,
                           const double* xs, const double* ys, const double* zs, const double* g,
                           int64_t n, void* stream);
int pvdb_raw2alpha_f64(const double* sigma, int64_t n, double shift, double interval,
                       double* exp_out, double* alpha_out, void* stream);
int pvdb_raw2alpha_bwd_f64(const double* exp_d, const double* grad_alpha, int64_t n,
                           double interval, double* grad_sigma, void* stream);
int pvdb_alpha2weight_f64(const double* alpha, int64_t n_rays, double* weight, double* T,
                          double* alphainv_last, const int64_t* i_start, int64_t* i_end,
                          void* stream);
int pvdb_alpha2weight_bwd_f64(const double* alpha, const double* weight, const double* T,
                              const double* alphainv_last, const int64_t* i_start,
                              const int64_t* i_end, int64_t n_rays, const double* grad_weights,
                              const double* grad_last, double* grad_alpha, void* stream);

/* ---- ray operations (plenvdb/lib/cuda/render_utils_kernel.cu) ------------ */
/* get_rays_of_a_view (plenvdb/lib/dvgo.py:522-527) with get_rays (:471-499) and
 * ndc_rays (:502-519, near = 1, focal = K[0][0]): rays_o, rays_d and viewdirs
 * [H][W][3] of one camera.  K: host 3x3 row-major (double, as the loaders hand
 * it over; rounded to float32 where torch rounds it); c2w: host 3x4 row-major
 * float32.  mode 0 = 'center', 1 = 'lefttop', 2 = 'random': pixel (h, w)
 * of the unflipped grids sits at (w + jitter[0][h][w], h + jitter[1][h][w]),
 * jitter being dev [2][H][W] float32 drawn by the caller (the reference draws
 * it with torch.rand_like, dvgo.py:482-484); NULL for modes 0 and 1. */
int pvdb_rays_of_view(int32_t H, int32_t W, const double* K /*host[9]*/,
                      const float* c2w /*host[12]*/, int32_t ndc, int32_t inverse_y,
                      int32_t flip_x, int32_t flip_y, int32_t mode, float* rays_o,
                      float* rays_d, float* viewdirs, const float* jitter, void* stream);
/* infer_t_minmax + infer_n_samples (:11-55).  far_per_ray may be NULL. */
int pvdb_ray_bounds(const float* rays_o, const float* rays_d, int64_t n,
                    const float* xyz_min /*host[3]*/, const float* xyz_max /*host[3]*/,
                    float near_, float far_, const float* far_per_ray, float stepdist,
                    float* t_min, float* t_max, int64_t* n_steps, void* stream);
/* sample_pts_on_rays (:144-242) given the inclusive prefix sum of n_steps. */
int pvdb_sample_pts(const float* rays_o, const float* rays_d, const float* t_min,
                    const int64_t* steps_incl, int64_t n_rays, int64_t total, float stepdist,
                    const float* xyz_min /*host*/, const float* xyz_max /*host*/, float* pts,
                    int64_t* ray_id, int64_t* step_id, uint8_t* inbbox, void* stream);
/* maskcache_lookup (:373-424). */
int pvdb_mask_lookup(const uint8_t* mask, int32_t sx, int32_t sy, int32_t sz, const float* pts,
                     int64_t m, const float* scale /*host*/, const float* shift /*host*/,
                     uint8_t* out, void* stream);
/* maskcache_lookup's rule evaluated on the VDB value masks (MaskGrid of the
 * topology's active voxels, plenvdb/lib/grid.py:207-243). */
int pvdb_active_lookup(const pvdb_topology* topo, const float* pts, int64_t m,
                       const float* scale /*host*/, const float* shift /*host*/, uint8_t* out,
                       void* stream);
/* VDBGrid.wld2idx (plenvdb/lib/grid.py:77-78), SoA output. */
int pvdb_wld2idx(const float* pts, int64_t m, const float* xyz_min /*host*/,
                 const float* xyz_max /*host*/, const int32_t* reso /*host*/, float* xs,
                 float* ys, float* zs, void* stream);
/* raw2alpha / raw2alpha_backward (:430-443, :506-517). */
int pvdb_raw2alpha(const float* sigma, int64_t n, float shift, float interval, float* exp_out,
                   float* alpha_out, void* stream);
int pvdb_raw2alpha_bwd(const float* exp_d, const float* grad_alpha, int64_t n, float interval,
                       float* grad_sigma, void* stream);
/* __set_i_for_segment_start_end (:625-634); i_start/i_end must be zeroed. */
int pvdb_segment_bounds(const int64_t* ray_id, int64_t n, int64_t* i_start, int64_t* i_end,
                        void* stream);
/* alpha2weight / alpha2weight_backward (:576-707). */
int pvdb_alpha2weight(const float* alpha, int64_t n_rays, float* weight, float* T,
                      float* alphainv_last, const int64_t* i_start, int64_t* i_end,
                      void* stream);
int pvdb_alpha2weight_bwd(const float* alpha, const float* weight, const float* T,
                          const float* alphainv_last, const int64_t* i_start,
                          const int64_t* i_end, int64_t n_rays, const float* grad_weights,
                          const float* grad_last, float* grad_alpha, void* stream);

/* ---- fused (batched) training step ---------------------------------------
 * One iteration of plenvdb/run.py:541-598 on the fine-stage model
 * (dvgo.py:296-388 forward with the 12-ch feature grid and the 39-128-128-3
 * rgbnet, MSE loss, backward through Raw2Alpha / Alphas2Weights /
 * QueryVerticalInVDB, MaskedAdam on the MLP and VDBAdam on both grids) or on
 * a model without the MLP (k0_dim, the coarse stage: rgb = sigmoid(k0),
 * dvgo.py:72-75,344-346) as one fixed launch sequence on `stream`: no host
 * synchronisation, no allocation (all scratch in `workspace`), so it can be
 * captured in a hipGraph.
 * Replaces, for that loop, the per-call round trips of DensityVDB/ColorVDB
 * forward/backward (plenvdb.cpp:11-29,46-64) plus render_utils_cuda and
 * adam_upd_cuda.  Payloads: density [slots] fp32, colour [slots][12] fp32 or
 * bf16 (color_bf16), gradients and Adam moments fp32; MLP parameters fp32 in
 * the order W0[39][128] b0[128] W1[128][128] b1[128] W2[128][3] b2[3] (W as
 * Linear.weight.T, run.py:98-104).  Grid gradients are accumulated (callers
 * zero them once); with do_adam the step applies Adam and re-zeroes the
 * gradients of the leaves it touched.  The MLP gradient is overwritten by
 * every step.  Alignment: color 16 bytes (8 for a bf16 payload), grad_color,
 * color_m and color_v 16 bytes, workspace 256 bytes (PVDB_EINVAL otherwise). */
typedef struct pvdb_step_desc {
    const float* rays_o;      /* dev [n][3] */
    const float* rays_d;      /* dev [n][3] */
    const float* viewdirs;    /* dev [n][3] */
    const float* target;      /* dev [n][3] */
    const float* far_per_ray; /* dev [n] or NULL (far = 1e9, dvgo.py:284) */
    int32_t n_rays;
    int32_t color_bf16;
    float xyz_min[3];
    float xyz_max[3];
    int32_t reso[3];
    float near_, stepdist, act_shift, interval, fast_color_thres, bg;
    float* density;
    float* grad_density;
    float* density_m;
    float* density_v;
    void* color;
    float* grad_color;
    float* color_m;
    float* color_v;
    float* mlp;
    float* mlp_grad;
    float* mlp_m;
    float* mlp_v;
    float stepsz_density, stepsz_color, step_size_mlp, eps, beta0, beta1;
    /* 0: forward + backward (gradients accumulate); 1: forward + backward +
     * Adam; 2: Adam alone on the gradients accumulated by earlier calls (the
     * reference's separate VDBAdam.step / MaskedAdam.step after a backward,
     * masked_adam.py:70-80,127-158), over the leaves those calls touched */
    int32_t do_adam;
    /* 0: the MLP model above (12 colour features; mlp* required).  3 or 12: no
     * MLP, the colour grid holds k0_dim fp32 channels per slot and
     * rgb = sigmoid(k0) (dvgo.py:72-75,344-346); target then holds k0_dim
     * values per ray and ray_out k0_dim + 1 (rgb, T_last); mlp* are ignored. */
    int32_t k0_dim;
    void* workspace; /* dev, pvdb_fused_workspace_bytes() bytes, 256-byte aligned */
    int64_t workspace_bytes;
    /* max records (occupancy-kept samples) per step, in [1, 2^31 - 2].  Records
     * past it are dropped: the step then covers the first `capacity` records in
     * ray order (no memory is touched beyond it) and pvdb_fused_records' total,
     * the full count saturated at 2^31 - 1, exceeds it.  The sum of the rays'
     * step counts (pvdb_ray_bounds) bounds every step. */
    int64_t capacity;
    /* dev [n][4] rgb + T_last, 16-byte aligned ([n][k0_dim + 1] with k0_dim), or
     * NULL; written in place by the step */
    float* ray_out;
    float* loss;      /* dev [1] or NULL */
    /* dev [3] or NULL: stepsz_density, stepsz_color, step_size_mlp read on the
     * device instead of the three fields above, so one captured graph of the
     * step can be replayed with each step's values */
    const float* adam_scalars;
    /* launch shapes, 0 = built-in default (results do not depend on them):
     * [0] forward-MLP workgroups, [1] scatter workgroups, [2] scatter records
     * per chunk (32 or 64), [3] rays per workgroup of the wave-per-ray kernels
     * (marches and composite: 1, 2, 4 or 8); pvdb_fused_launch_defaults()
     * fills the defaults for this device */
    int32_t launch[4];
    /* loss weights of plenvdb/run.py:551-574 (cfg_train.weight_*): the MSE
     * term (1.0 in every configuration), the background entropy of T_last
     * (clamped to [1e-6, 1 - 1e-6], mean over rays) and the per-sample rgb
     * loss sum(w.detach() * |rgb_i - target|^2) / N; 0 switches a term off */
    float weight_main, weight_entropy_last, weight_rgbper;
    /* 0: grad_color is [slots][C] and grad_density [slots].  Otherwise both are
     * rows grad_stride floats apart (C <= grad_stride, a multiple of 4 for
     * C = 12), and grad_density may point inside grad_color's rows (e.g.
     * grad_color + C): a voxel's gradients then share one cache line, and the
     * step's atomics for it are one memory request */
    int32_t grad_stride;
    /* dev [slots][12] fp32 or NULL: with a bf16 colour payload, the fp32
     * master the colour Adam updates; the payload then receives its
     * round-to-nearest bf16 value after each update (NULL: the bf16 payload is
     * updated in place).  Ignored for an fp32 payload. */
    float* color_master;
    /* dev [slots] per-voxel rate of density stepmode 2 (set_pervoxel_lr,
     * masked_adam.py:43-46), or NULL */
    const float* density_lr;
    /* VDBAdam stepmodes (masked_adam.py:56-68; densityvdb.cu:185-349,
     * colorvdb.cu:179-351): density 1 skip-zero-grad (the fine stage), 0 every
     * active voxel, 2 every active voxel with density_lr; colour 1 (a Vec3
     * group is skipped when its three gradients are zero) or 0.  With both at
     * 1 the Adam visits only the leaves the backward touched. */
    int32_t stepmode_density;
    int32_t stepmode_color;
} pvdb_step_desc;

/* The default launch shapes (pvdb_step_desc.launch) on the current device. */
int pvdb_fused_launch_defaults(int32_t* launch4);
/* Workspace bytes for n_rays rays, `capacity` records and nleaf leaves (-1 on bad sizes). */
int64_t pvdb_fused_workspace_bytes(int64_t n_rays, int64_t capacity, int64_t nleaf);
/* Clears the workspace (touched-leaf flags, buffers); call once after allocating it. */
int pvdb_fused_init(const pvdb_topology* topo, const pvdb_step_desc* desc, void* stream);
int pvdb_fused_step(const pvdb_topology* topo, const pvdb_step_desc* desc, void* stream);
/* Optional device timing of the fused step's stages (HIP events on its stream).
 * No counterpart in the reference (which times only the renderer, with CPU
 * clock(), plenvdb.h:1030-1035).  Collects pending events into per-stage totals
 * and copies up to n of them to total_ms / calls; enable = 1 starts (and
 * resets), 0 stops (and resets), -1 only reads.  Must stay off while a caller
 * captures the step into a graph. */
int pvdb_fused_timing(int32_t enable, double* total_ms, int32_t* calls, int32_t n);
const char* pvdb_fused_stage_name(int32_t i);
/* Device pointers of the per-sample records and per-ray counts inside the
 * workspace (for inspection and tests).  Per record: xyze float4 (index-space
 * position, exp(sigma + shift)), ab float2 (alpha, ray id and flags; the
 * transmittance and weight are recomputed from the alphas where they are
 * needed, not stored), rgb float4, adj float4 (dL/dlogit x3, dL/dsigma). */
int pvdb_fused_records(const pvdb_topology* topo, const pvdb_step_desc* desc, void** xyze,
                       void** ab, void** rgb, void** adj, int32_t** count, int32_t** offset,
                       int32_t** total);

/* ---- merged-VDB renderer ----------------------------------------------------
 * MGRenderer (plenvdb/lib/vdb/plenvdb.h:933-1068, plenvdb.cpp:135-171;
 * render_an_image_cuda renderer.cu:370-424) as three kernels on `stream`,
 * with no host synchronisation and no allocation: a march kernel runs both of
 * the reference's marches per pixel (first_look_merged's count and t_min /
 * t_max rewrite, then ray_marching_merged from the rewritten t_min, each with
 * its own summation order, renderer.cu:222-367) and writes one record per
 * kept sample; an MLP kernel evaluates the RGB MLP on fp16 MFMA over the
 * records; a composite kernel adds each pixel's contributions to T * bg in
 * sample order.  The merged index grid (vdb_compression.py:19-59) is a
 * topology whose active voxels are the non-zero indices plus `slot_idx` (the
 * index of every slot, 0 when inactive).  Differences from the reference: the
 * early-stop threshold t_stop is a parameter (the reference hard-codes 1e-3,
 * renderer.cu:265); the MLP runs in fp16 with fp32 accumulation; a pixel whose
 * second march keeps more samples than its first counted contributes the first
 * `count` of them (the reference writes past its segment there, renderer.cu:
 * 363 comments the check out). */
typedef struct pvdb_render_desc {
    float c2w[16];    /* host values: camera-to-world, row-major (input_a_c2w) */
    float K[9];       /* intrinsics, row-major (setScene) */
    float xyz_min[3];
    float xyz_max[3];
    int32_t reso[3];
    int32_t H, W, inverse_y;
    float near_, stepdist, act_shift, interval, fast_color_thres, bg, t_stop;
    int32_t payload_f32;   /* payload slots hold fp32 (64 B) instead of fp16 (32 B) */
    const void* payload;   /* dev, pvdb_render_payload_bytes() bytes */
    const void* weights;   /* dev, pvdb_render_weights_bytes() bytes */
    float* out;            /* dev [H*W*3] (output_an_image layout) */
    /* dev [3] or NULL: first-march density gathers, written samples, records the
     * frame reserved (above `capacity` when the frame did not fit) */
    uint64_t* counters;
    /* dev, pvdb_render_workspace_bytes(H, W, capacity) bytes, 256-byte aligned */
    void* workspace;
    int64_t workspace_bytes;
    /* records (kept samples) per frame, in [0, 2^31 - 2]; a frame with more
     * renders the records that fit and pvdb_render_frame_info's total, the
     * full count, exceeds it */
    int64_t capacity;
} pvdb_render_desc;

/* Bytes of the packed MLP operands. */
int64_t pvdb_render_weights_bytes(void);
/* MGRenderer::load_params (plenvdb.h:960-990): w* are [in][out] row-major
 * (Linear.weight.T, run.py:98-104), device fp32.  Only dcol 12, dpe 27,
 * dhid 128, dout 3 are supported. */
int pvdb_render_pack_weights(const float* w0, const float* b0, const float* w1, const float* b1,
                             const float* w2, const float* b2, int32_t dcol, int32_t dpe,
                             int32_t dhid, int32_t dout, void* packed /*dev*/, void* stream);
/* Payload bytes for the topology ((nslots + 1) slots; -1 on a null handle). */
int64_t pvdb_render_payload_bytes(const pvdb_topology* topo, int32_t payload_f32);
/* MGRenderer::load_data (plenvdb.h:992-1005): den [nidx], col [nidx][12] (dev
 * fp32, index 0 = the value of inactive voxels) gathered into the slot payload
 * through slot_idx (dev [nslots], values in [0, nidx)). */
int pvdb_render_build_payload(const pvdb_topology* topo, const int32_t* slot_idx,
                              const float* den, const float* col, int64_t nidx,
                              int32_t payload_f32, void* payload, void* stream);
/* Workspace bytes of an H x W frame with `capacity` records (-1 on bad sizes). */
int64_t pvdb_render_workspace_bytes(int32_t H, int32_t W, int64_t capacity);
/* MGRenderer::render_an_image (plenvdb.h:1029-1035): writes out [H*W*3]. */
int pvdb_render_image(const pvdb_topology* topo, const pvdb_render_desc* desc, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): pvdb_render_image with the frame workspace taken from a fixed
 * 64 MiB device array inside the library instead of desc->workspace (ignored,
 * with workspace_bytes).  Allocates nothing.  The frame must fit:
 * pvdb_render_workspace_bytes(H, W, capacity) <= pvdb_render_noalloc_scratch_bytes(),
 * else PVDB_EINVAL before any device work.  Every call on one device shares that
 * device's array (held until the process ends), so calls that overlap on one
 * device race; pvdb_render_frame_info does not apply to these frames. */
int pvdb_render_image_noalloc(const pvdb_topology* topo, const pvdb_render_desc* desc,
                              void* stream);
/* Bytes of the scratch pvdb_render_image_noalloc renders in (64 MiB). */
int64_t pvdb_render_noalloc_scratch_bytes(void);
/* The last frame's record total (synchronous copy to the host; may exceed the
 * capacity) and device pointers of its per-pixel first-march counts, written
 * samples and first record (the reference keeps the counts in n_samples,
 * renderer.cu:236-262).  No counterpart in the reference for the rest. */
int pvdb_render_frame_info(const pvdb_render_desc* desc, uint64_t* total, int32_t** pix_count,
                           int32_t** pix_written, int32_t** pix_start);
/* Optional device timing of the renderer's march, MLP and composite kernels
 * (HIP events around each launch; no counterpart in the reference, which sums
 * CPU clock() time, plenvdb.h:1030-1035).  Collects pending events and copies
 * the three totals (ms) to total_ms3; enable = 1 starts (and resets), 0 stops
 * (and resets), -1 only reads.  Must stay off under graph capture. */
int pvdb_render_timing(int32_t enable, double* total_ms3);

/* ---- self test --------------------------------------------------------------
 * No counterpart in the reference: multiplies bf16 matrices with the MFMA
 * fragment layouts the MLP kernels assume (A [16][32], B [32][16] -> D [16][16];
 * A16 [16][16], B16 [16][16] -> D16), so a test can check them on the device. */
int pvdb_mfma_selftest(const uint16_t* A, const uint16_t* B, float* D, const uint16_t* A16,
                       const uint16_t* B16, float* D16, void* stream);
/* The same check for v_mfma_f32_16x16x32_f16 (fp16 A [16][32], B [32][16]). */
int pvdb_mfma_selftest_f16(const uint16_t* A, const uint16_t* B, float* D, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PLENVDB_HIP_H */
