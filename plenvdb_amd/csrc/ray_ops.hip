// Stage-per-kernel ray operations with the reference's semantics: ray/bbox
// bounds and sample counts, materialised sample points, occupancy lookups,
// raw2alpha and Alphas2Weights (forward and backward).
//
// They follow plenvdb/lib/cuda/render_utils_kernel.cu operation by operation
// (including its float/double promotions) and back both the reference-shaped
// FP32 training step and the render_utils-compatible Python ops.  The library
// is compiled with -ffp-contract=off, so every expression rounds as written.
#include "pvdb_common.h"
#include "plenvdb_hip.h"

namespace pvdb {
TopoView view_of(const pvdb_topology* t);
}

namespace {

// infer_t_minmax + infer_n_samples (render_utils_kernel.cu:11-55).
__global__ void k_ray_bounds(const float* __restrict__ ro, const float* __restrict__ rd, int64_t n,
                             float mn0, float mn1, float mn2, float mx0, float mx1, float mx2,
                             float near_, float far_, const float* __restrict__ far_per_ray,
                             float stepdist, float* __restrict__ t_min, float* __restrict__ t_max,
                             int64_t* __restrict__ n_steps) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float ox = ro[3 * r], oy = ro[3 * r + 1], oz = ro[3 * r + 2];
    const float dx = rd[3 * r], dy = rd[3 * r + 1], dz = rd[3 * r + 2];
    const float vx = dx == 0.f ? 1e-6f : dx;
    const float vy = dy == 0.f ? 1e-6f : dy;
    const float vz = dz == 0.f ? 1e-6f : dz;
    const float ax = (mx0 - ox) / vx, ay = (mx1 - oy) / vy, az = (mx2 - oz) / vz;
    const float bx = (mn0 - ox) / vx, by = (mn1 - oy) / vy, bz = (mn2 - oz) / vz;
    const float fr = far_per_ray ? far_per_ray[r] : far_;
    const float lo = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    const float hi = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    const float tmn = fmaxf(fminf(lo, fr), near_);
    const float tmx = fmaxf(fminf(hi, fr), near_);
    t_min[r] = tmn;
    t_max[r] = tmx;
    const float rn = sqrtf(dx * dx + dy * dy + dz * dz);
    const float q = ceilf(((tmx - tmn) * rn) / stepdist);
    n_steps[r] = (int64_t)(q > 1.f ? q : 1.f);
}

// sample_pts_on_rays (render_utils_kernel.cu:57-79,144-194).  One thread per
// sample; its ray is found by binary search over the inclusive prefix sum of
// n_steps (the reference materialises ray_id with a scatter plus cumsum).
__global__ void k_sample_pts(const float* __restrict__ ro, const float* __restrict__ rd,
                             const float* __restrict__ t_min, const int64_t* __restrict__ incl,
                             int64_t n_rays, int64_t total, float stepdist, float mn0, float mn1,
                             float mn2, float mx0, float mx1, float mx2, float* __restrict__ pts,
                             int64_t* __restrict__ ray_id, int64_t* __restrict__ step_id,
                             uint8_t* __restrict__ inbbox) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= total) return;
    int64_t lo = 0, hi = n_rays - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (incl[mid] > m)
            hi = mid;
        else
            lo = mid + 1;
    }
    const int64_t r = lo;
    const int64_t i = m - (r > 0 ? incl[r - 1] : 0);
    const float dx = rd[3 * r], dy = rd[3 * r + 1], dz = rd[3 * r + 2];
    const float rn = sqrtf(dx * dx + dy * dy + dz * dz);
    const float tm = t_min[r];
    const float sx = ro[3 * r] + dx * tm, sy = ro[3 * r + 1] + dy * tm, sz = ro[3 * r + 2] + dz * tm;
    const float ux = dx / rn, uy = dy / rn, uz = dz / rn;
    const float dist = stepdist * (float)i;
    const float px = sx + ux * dist, py = sy + uy * dist, pz = sz + uz * dist;
    pts[3 * m] = px;
    pts[3 * m + 1] = py;
    pts[3 * m + 2] = pz;
    ray_id[m] = r;
    step_id[m] = i;
    inbbox[m] = !((mn0 > px) | (mn1 > py) | (mn2 > pz) | (mx0 < px) | (mx1 < py) | (mx2 < pz));
}

// maskcache_lookup (render_utils_kernel.cu:373-390): round half away from zero.
__global__ void k_mask_lookup(const uint8_t* __restrict__ mask, int sx, int sy, int sz,
                              const float* __restrict__ pts, int64_t m, float s0, float s1,
                              float s2, float b0, float b1, float b2, uint8_t* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const int i = (int)roundf(pts[3 * p] * s0 + b0);
    const int j = (int)roundf(pts[3 * p + 1] * s1 + b1);
    const int k = (int)roundf(pts[3 * p + 2] * s2 + b2);
    const bool in = (0 <= i) && (i < sx) && (0 <= j) && (j < sy) && (0 <= k) && (k < sz);
    out[p] = in ? mask[((int64_t)i * sy + j) * sz + k] : 0;
}

// The same lookup against the VDB value masks instead of a dense bool grid:
// the occupancy of a leaf-granular topology is its set of active voxels.
__global__ void k_active_lookup(pvdb::TopoView t, const float* __restrict__ pts, int64_t m,
                                float s0, float s1, float s2, float b0, float b1, float b2,
                                uint8_t* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const int i = (int)roundf(pts[3 * p] * s0 + b0);
    const int j = (int)roundf(pts[3 * p + 1] * s1 + b1);
    const int k = (int)roundf(pts[3 * p + 2] * s2 + b2);
    out[p] = pvdb::voxel_active(t, i, j, k) ? 1 : 0;
}

// wld2idx (plenvdb/lib/grid.py:77-78), SoA output.
__global__ void k_wld2idx(const float* __restrict__ pts, int64_t m, float mn0, float mn1,
                          float mn2, float ln0, float ln1, float ln2, float r0, float r1, float r2,
                          float* __restrict__ xs, float* __restrict__ ys, float* __restrict__ zs) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    xs[p] = ((pts[3 * p] - mn0) / ln0) * r0;
    ys[p] = ((pts[3 * p + 1] - mn1) / ln1) * r1;
    zs[p] = ((pts[3 * p + 2] - mn2) / ln2) * r2;
}

// raw2alpha (render_utils_kernel.cu:430-443).
__global__ void k_raw2alpha(const float* __restrict__ sigma, int64_t n, float shift,
                            float interval, float* __restrict__ e_out, float* __restrict__ a_out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float e = expf(sigma[p] + shift);
    e_out[p] = e;
    a_out[p] = 1.f - powf(1.f + e, -interval);
}

// raw2alpha backward (render_utils_kernel.cu:506-517); min(float, 1e10) is a
// double in the reference, so the product is formed in double.
__global__ void k_raw2alpha_bwd(const float* __restrict__ e, const float* __restrict__ ga,
                                int64_t n, float interval, float* __restrict__ gs) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double m = fmin((double)e[p], 1e10);
    const float pw = powf(1.f + e[p], -interval - 1.f);
    gs[p] = (float)(m * (double)pw * (double)interval * (double)ga[p]);
}

// __set_i_for_segment_start_end (render_utils_kernel.cu:625-634) plus the
// i_end of the last ray.
__global__ void k_segment_bounds(const int64_t* __restrict__ ray_id, int64_t n,
                                 int64_t* __restrict__ i_start, int64_t* __restrict__ i_end) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (p > 0 && ray_id[p] != ray_id[p - 1]) {
        i_start[ray_id[p]] = p;
        i_end[ray_id[p - 1]] = p;
    }
    if (p == n - 1) i_end[ray_id[p]] = n;
}

// alpha2weight (render_utils_kernel.cu:576-603): float T_cum times the double
// (1. - alpha), stop after the first T < 1e-3.
__global__ void k_alpha2weight(const float* __restrict__ alpha, int64_t n_rays,
                               float* __restrict__ w, float* __restrict__ T,
                               float* __restrict__ t_last, const int64_t* __restrict__ i_start,
                               int64_t* __restrict__ i_end) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rays) return;
    const int64_t is = i_start[r], ie = i_end[r];
    float Tc = 1.f;
    int64_t i;
    for (i = is; i < ie; ++i) {
        T[i] = Tc;
        w[i] = Tc * alpha[i];
        Tc = (float)((double)Tc * (1.0 - (double)alpha[i]));
        if ((double)Tc < 1e-3) {
            i += 1;
            break;
        }
    }
    i_end[r] = i;
    t_last[r] = Tc;
}

// alpha2weight backward (render_utils_kernel.cu:653-676).
__global__ void k_alpha2weight_bwd(const float* __restrict__ alpha, const float* __restrict__ w,
                                   const float* __restrict__ T, const float* __restrict__ t_last,
                                   const int64_t* __restrict__ i_start,
                                   const int64_t* __restrict__ i_end, int64_t n_rays,
                                   const float* __restrict__ gw, const float* __restrict__ glast,
                                   float* __restrict__ ga) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rays) return;
    float back = glast[r] * t_last[r];
    for (int64_t i = i_end[r] - 1; i >= i_start[r]; --i) {
        const double den = (double)(1.f - alpha[i]) + 1e-10;
        ga[i] = (float)((double)(gw[i] * T[i]) - (double)back / den);
        back += gw[i] * w[i];
    }
}

// get_rays + ndc_rays + get_rays_of_a_view (dvgo.py:471-527), one pixel per
// thread.  The scalars that torch rounds to float32 (K's entries, the NDC
// constants formed in double on the host) arrive rounded; every tensor
// operation is one float32 operation in the reference's order.
struct ViewParams {
    int H, W;
    float fx, fy, cx, cy;  // K[0][0], K[1][1], K[0][2], K[1][2]
    float R[9];            // c2w[:3,:3], row-major
    float o[3];            // c2w[:3,3]
    int ndc, inverse_y, flip_x, flip_y;
    float off;           // 0.5 ('center') or 0 ('lefttop')
    float near_;         // NDC near plane (1.0)
    float ax, ay;        // -1./(W/(2.*focal)), -1./(H/(2.*focal))
    float two_near;      // 2. * near
    float neg_two_near;  // -2. * near
};

__global__ void k_rays_of_view(ViewParams V, const float* __restrict__ jitter,
                               float* __restrict__ ro, float* __restrict__ rd,
                               float* __restrict__ vd) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)V.H * V.W) return;
    const int h = (int)(p / V.W), w = (int)(p % V.W);
    // linspace(0, n-1, n) holds the exact integers; flips reverse an axis of
    // the (jittered) grids, so a flipped pixel takes its mirror's jitter
    const int wi = V.flip_x ? V.W - 1 - w : w, hj = V.flip_y ? V.H - 1 - h : h;
    const float oi = jitter ? jitter[(int64_t)h * V.W + wi] : V.off;
    const float oj = jitter ? jitter[(int64_t)V.H * V.W + (int64_t)hj * V.W + w] : V.off;
    const float i = (float)wi + oi;
    const float j = (float)hj + oj;
    const float d0 = (i - V.cx) / V.fx;
    float d1, d2;
    if (V.inverse_y) {
        d1 = (j - V.cy) / V.fy;
        d2 = 1.f;
    } else {
        d1 = -((j - V.cy) / V.fy);
        d2 = -1.f;
    }
    // rays_d = sum(dirs[..., None, :] * c2w[:3,:3], -1)
    float dx = (d0 * V.R[0] + d1 * V.R[1]) + d2 * V.R[2];
    float dy = (d0 * V.R[3] + d1 * V.R[4]) + d2 * V.R[5];
    float dz = (d0 * V.R[6] + d1 * V.R[7]) + d2 * V.R[8];
    const float nrm = sqrtf((dx * dx + dy * dy) + dz * dz);
    vd[3 * p] = dx / nrm;
    vd[3 * p + 1] = dy / nrm;
    vd[3 * p + 2] = dz / nrm;
    float ox = V.o[0], oy = V.o[1], oz = V.o[2];
    if (V.ndc) {
        const float t = -(V.near_ + oz) / dz;
        ox = ox + t * dx;
        oy = oy + t * dy;
        oz = oz + t * dz;
        const float o0 = (V.ax * ox) / oz, o1 = (V.ay * oy) / oz;
        const float o2 = 1.f + V.two_near / oz;
        const float e0 = V.ax * (dx / dz - ox / oz), e1 = V.ay * (dy / dz - oy / oz);
        const float e2 = V.neg_two_near / oz;
        ox = o0;
        oy = o1;
        oz = o2;
        dx = e0;
        dy = e1;
        dz = e2;
    }
    ro[3 * p] = ox;
    ro[3 * p + 1] = oy;
    ro[3 * p + 2] = oz;
    rd[3 * p] = dx;
    rd[3 * p + 1] = dy;
    rd[3 * p + 2] = dz;
}

}  // namespace

extern "C" {

int pvdb_rays_of_view(int32_t H, int32_t W, const double* K, const float* c2w, int32_t ndc,
                      int32_t inverse_y, int32_t flip_x, int32_t flip_y, int32_t mode,
                      float* rays_o, float* rays_d, float* viewdirs, const float* jitter,
                      void* stream) {
    PVDB_CHECK_ARG(H >= 0 && W >= 0 && K && c2w, "bad arguments");
    PVDB_CHECK_ARG(mode == 0 || mode == 1 || mode == 2,
                   "mode must be 0 ('center'), 1 ('lefttop') or 2 ('random', with jitter)");
    if ((int64_t)H * W == 0) return PVDB_OK;
    PVDB_CHECK_ARG(rays_o && rays_d && viewdirs, "null pointer");
    PVDB_CHECK_ARG((mode == 2) == (jitter != nullptr), "mode 2 ('random') takes the jitter array");
    PVDB_CHECK_ARG(K[0] != 0.0 && K[4] != 0.0, "focal lengths must be non-zero");
    ViewParams V;
    V.H = H;
    V.W = W;
    V.fx = (float)K[0];
    V.fy = (float)K[4];
    V.cx = (float)K[2];
    V.cy = (float)K[5];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) V.R[3 * r + c] = c2w[4 * r + c];
        V.o[r] = c2w[4 * r + 3];
    }
    V.ndc = ndc != 0;
    V.inverse_y = inverse_y != 0;
    V.flip_x = flip_x != 0;
    V.flip_y = flip_y != 0;
    V.off = mode == 0 ? 0.5f : 0.f;
    const double focal = K[0], near_ = 1.0;  // ndc_rays(H, W, K[0][0], 1., ...)
    V.near_ = (float)near_;
    V.ax = (float)(-1. / ((double)W / (2. * focal)));
    V.ay = (float)(-1. / ((double)H / (2. * focal)));
    V.two_near = (float)(2. * near_);
    V.neg_two_near = (float)(-2. * near_);
    const int64_t n = (int64_t)H * W;
    hipLaunchKernelGGL(k_rays_of_view, dim3(pvdb_blocks(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, V, jitter, rays_o, rays_d, viewdirs);
    PVDB_LAUNCHED("k_rays_of_view");
    return PVDB_OK;
}

int pvdb_ray_bounds(const float* rays_o, const float* rays_d, int64_t n, const float* xyz_min,
                    const float* xyz_max, float near_, float far_, const float* far_per_ray,
                    float stepdist, float* t_min, float* t_max, int64_t* n_steps, void* stream) {
    PVDB_CHECK_ARG(n >= 0 && xyz_min && xyz_max, "bad arguments");
    PVDB_CHECK_ARG(stepdist > 0.f, "stepdist must be positive");
    if (n == 0) return PVDB_OK;
    PVDB_CHECK_ARG(rays_o && rays_d && t_min && t_max && n_steps, "null pointer");
    hipLaunchKernelGGL(k_ray_bounds, dim3(pvdb_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       rays_o, rays_d, n, xyz_min[0], xyz_min[1], xyz_min[2], xyz_max[0],
                       xyz_max[1], xyz_max[2], near_, far_, far_per_ray, stepdist, t_min, t_max,
                       n_steps);
    PVDB_LAUNCHED("k_ray_bounds");
    return PVDB_OK;
}

int pvdb_sample_pts(const float* rays_o, const float* rays_d, const float* t_min,
                    const int64_t* steps_incl, int64_t n_rays, int64_t total, float stepdist,
                    const float* xyz_min, const float* xyz_max, float* pts, int64_t* ray_id,
                    int64_t* step_id, uint8_t* inbbox, void* stream) {
    PVDB_CHECK_ARG(n_rays >= 0 && total >= 0 && xyz_min && xyz_max, "bad arguments");
    if (total == 0) return PVDB_OK;
    PVDB_CHECK_ARG(n_rays > 0, "samples without rays");
    PVDB_CHECK_ARG(rays_o && rays_d && t_min && steps_incl && pts && ray_id && step_id && inbbox,
                   "null pointer");
    hipLaunchKernelGGL(k_sample_pts, dim3(pvdb_blocks(total, 256)), dim3(256), 0,
                       (hipStream_t)stream, rays_o, rays_d, t_min, steps_incl, n_rays, total,
                       stepdist, xyz_min[0], xyz_min[1], xyz_min[2], xyz_max[0], xyz_max[1],
                       xyz_max[2], pts, ray_id, step_id, inbbox);
    PVDB_LAUNCHED("k_sample_pts");
    return PVDB_OK;
}

int pvdb_mask_lookup(const uint8_t* mask, int32_t sx, int32_t sy, int32_t sz, const float* pts,
                     int64_t m, const float* scale, const float* shift, uint8_t* out,
                     void* stream) {
    PVDB_CHECK_ARG(m >= 0 && sx > 0 && sy > 0 && sz > 0 && scale && shift, "bad arguments");
    if (m == 0) return PVDB_OK;
    PVDB_CHECK_ARG(mask && pts && out, "null pointer");
    hipLaunchKernelGGL(k_mask_lookup, dim3(pvdb_blocks(m, 256)), dim3(256), 0, (hipStream_t)stream,
                       mask, sx, sy, sz, pts, m, scale[0], scale[1], scale[2], shift[0], shift[1],
                       shift[2], out);
    PVDB_LAUNCHED("k_mask_lookup");
    return PVDB_OK;
}

int pvdb_active_lookup(const pvdb_topology* topo, const float* pts, int64_t m,
                       const float* scale, const float* shift, uint8_t* out, void* stream) {
    if (!topo) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    PVDB_CHECK_ARG(m >= 0 && scale && shift, "bad arguments");
    if (m == 0) return PVDB_OK;
    PVDB_CHECK_ARG(pts && out, "null pointer");
    hipLaunchKernelGGL(k_active_lookup, dim3(pvdb_blocks(m, 256)), dim3(256), 0,
                       (hipStream_t)stream, pvdb::view_of(topo), pts, m, scale[0], scale[1],
                       scale[2], shift[0], shift[1], shift[2], out);
    PVDB_LAUNCHED("k_active_lookup");
    return PVDB_OK;
}

int pvdb_wld2idx(const float* pts, int64_t m, const float* xyz_min, const float* xyz_max,
                 const int32_t* reso, float* xs, float* ys, float* zs, void* stream) {
    PVDB_CHECK_ARG(m >= 0 && xyz_min && xyz_max && reso, "bad arguments");
    if (m == 0) return PVDB_OK;
    PVDB_CHECK_ARG(pts && xs && ys && zs, "null pointer");
    hipLaunchKernelGGL(k_wld2idx, dim3(pvdb_blocks(m, 256)), dim3(256), 0, (hipStream_t)stream,
                       pts, m, xyz_min[0], xyz_min[1], xyz_min[2], xyz_max[0] - xyz_min[0],
                       xyz_max[1] - xyz_min[1], xyz_max[2] - xyz_min[2], (float)(reso[0] - 1),
                       (float)(reso[1] - 1), (float)(reso[2] - 1), xs, ys, zs);
    PVDB_LAUNCHED("k_wld2idx");
    return PVDB_OK;
}

int pvdb_raw2alpha(const float* sigma, int64_t n, float shift, float interval, float* exp_out,
                   float* alpha_out, void* stream) {
    PVDB_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PVDB_OK;
    PVDB_CHECK_ARG(sigma && exp_out && alpha_out, "null pointer");
    hipLaunchKernelGGL(k_raw2alpha, dim3(pvdb_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       sigma, n, shift, interval, exp_out, alpha_out);
    PVDB_LAUNCHED("k_raw2alpha");
    return PVDB_OK;
}

int pvdb_raw2alpha_bwd(const float* exp_d, const float* grad_alpha, int64_t n, float interval,
                       float* grad_sigma, void* stream) {
    PVDB_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PVDB_OK;
    PVDB_CHECK_ARG(exp_d && grad_alpha && grad_sigma, "null pointer");
    hipLaunchKernelGGL(k_raw2alpha_bwd, dim3(pvdb_blocks(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, exp_d, grad_alpha, n, interval, grad_sigma);
    PVDB_LAUNCHED("k_raw2alpha_bwd");
    return PVDB_OK;
}

int pvdb_segment_bounds(const int64_t* ray_id, int64_t n, int64_t* i_start, int64_t* i_end,
                        void* stream) {
    PVDB_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PVDB_OK;
    PVDB_CHECK_ARG(ray_id && i_start && i_end, "null pointer");
    hipLaunchKernelGGL(k_segment_bounds, dim3(pvdb_blocks(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, ray_id, n, i_start, i_end);
    PVDB_LAUNCHED("k_segment_bounds");
    return PVDB_OK;
}

int pvdb_alpha2weight(const float* alpha, int64_t n_rays, float* weight, float* T,
                      float* alphainv_last, const int64_t* i_start, int64_t* i_end,
                      void* stream) {
    PVDB_CHECK_ARG(n_rays >= 0, "n_rays must be >= 0");
    if (n_rays == 0) return PVDB_OK;
    PVDB_CHECK_ARG(alpha && weight && T && alphainv_last && i_start && i_end, "null pointer");
    hipLaunchKernelGGL(k_alpha2weight, dim3(pvdb_blocks(n_rays, 256)), dim3(256), 0,
                       (hipStream_t)stream, alpha, n_rays, weight, T, alphainv_last, i_start,
                       i_end);
    PVDB_LAUNCHED("k_alpha2weight");
    return PVDB_OK;
}

int pvdb_alpha2weight_bwd(const float* alpha, const float* weight, const float* T,
                          const float* alphainv_last, const int64_t* i_start,
                          const int64_t* i_end, int64_t n_rays, const float* grad_weights,
                          const float* grad_last, float* grad_alpha, void* stream) {
    PVDB_CHECK_ARG(n_rays >= 0, "n_rays must be >= 0");
    if (n_rays == 0) return PVDB_OK;
    PVDB_CHECK_ARG(alpha && weight && T && alphainv_last && i_start && i_end && grad_weights &&
                       grad_last && grad_alpha,
                   "null pointer");
    hipLaunchKernelGGL(k_alpha2weight_bwd, dim3(pvdb_blocks(n_rays, 256)), dim3(256), 0,
                       (hipStream_t)stream, alpha, weight, T, alphainv_last, i_start, i_end,
                       n_rays, grad_weights, grad_last, grad_alpha);
    PVDB_LAUNCHED("k_alpha2weight_bwd");
    return PVDB_OK;
}

}  // extern "C"
