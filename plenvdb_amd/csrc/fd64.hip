// FP64 twins of the hand-written adjoint pairs of the plain path, for central
// finite-difference checks on the device: trilinear gather / "leaf must exist"
// scatter (densityvdb.cu:101-181, colorvdb.cu:81-175), raw2alpha and its
// backward (render_utils_kernel.cu:430-443, 506-517), Alphas2Weights and its
// backward (:576-603, 653-676).  Each follows its FP32 kernel's formula with
// every quantity in double (the same corner walk, the same early stop after
// the first T < 1e-3); nothing on the training or render path calls them.
#include "pvdb_common.h"
#include "gather.h"
#include "plenvdb_hip.h"

namespace pvdb {
TopoView view_of(const pvdb_topology* t);
}

using pvdb::kCx;
using pvdb::kCy;
using pvdb::kCz;
using pvdb::LeafCache;
using pvdb::TopoView;

namespace {

struct Cell64 {
    int i0, j0, k0;
    double u, v, w;
};

__device__ __forceinline__ Cell64 make_cell64(double x, double y, double z) {
    Cell64 c;
    c.i0 = (int)floor(x);
    c.j0 = (int)floor(y);
    c.k0 = (int)floor(z);
    c.u = x - (double)c.i0;
    c.v = y - (double)c.j0;
    c.w = z - (double)c.k0;
    return c;
}

template <int C>
__global__ void __launch_bounds__(256) k_tri_fwd64(TopoView t, const double* __restrict__ vals,
                                                   const double* __restrict__ xs,
                                                   const double* __restrict__ ys,
                                                   const double* __restrict__ zs, int64_t n,
                                                   double* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const Cell64 cell = make_cell64(xs[p], ys[p], zs[p]);
    LeafCache cache;
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int dx = (kCx >> q) & 1, dy = (kCy >> q) & 1, dz = (kCz >> q) & 1;
        const double wq = (dx ? cell.u : 1.0 - cell.u) * (dy ? cell.v : 1.0 - cell.v) *
                          (dz ? cell.w : 1.0 - cell.w);
        const int64_t s = cache.slot(t, cell.i0 + dx, cell.j0 + dy, cell.k0 + dz);
        if (s < 0) continue;
#pragma unroll
        for (int c = 0; c < C; ++c) acc[c] += vals[s * C + c] * wq;
    }
#pragma unroll
    for (int c = 0; c < C; ++c) out[p * C + c] = acc[c];
}

template <int C>
__global__ void __launch_bounds__(256) k_tri_bwd64(TopoView t, double* __restrict__ grad,
                                                   const double* __restrict__ xs,
                                                   const double* __restrict__ ys,
                                                   const double* __restrict__ zs,
                                                   const double* __restrict__ g, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const Cell64 cell = make_cell64(xs[p], ys[p], zs[p]);
    LeafCache cache;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int dx = (kCx >> q) & 1, dy = (kCy >> q) & 1, dz = (kCz >> q) & 1;
        const double wq = (dx ? cell.u : 1.0 - cell.u) * (dy ? cell.v : 1.0 - cell.v) *
                          (dz ? cell.w : 1.0 - cell.w);
        const int64_t s = cache.slot(t, cell.i0 + dx, cell.j0 + dy, cell.k0 + dz);
        if (s < 0) continue;  // the leaf must exist
#pragma unroll
        for (int c = 0; c < C; ++c) atomicAdd(grad + s * C + c, g[p * C + c] * wq);
    }
}

__global__ void k_raw2alpha64(const double* __restrict__ sigma, int64_t n, double shift,
                              double interval, double* __restrict__ e_out,
                              double* __restrict__ a_out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double e = exp(sigma[p] + shift);
    e_out[p] = e;
    a_out[p] = 1.0 - pow(1.0 + e, -interval);
}

__global__ void k_raw2alpha_bwd64(const double* __restrict__ e, const double* __restrict__ ga,
                                  int64_t n, double interval, double* __restrict__ gs) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    gs[p] = fmin(e[p], 1e10) * pow(1.0 + e[p], -interval - 1.0) * interval * ga[p];
}

__global__ void k_alpha2weight64(const double* __restrict__ alpha, int64_t n_rays,
                                 double* __restrict__ w, double* __restrict__ T,
                                 double* __restrict__ t_last, const int64_t* __restrict__ i_start,
                                 int64_t* __restrict__ i_end) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rays) return;
    const int64_t is = i_start[r], ie = i_end[r];
    double Tc = 1.0;
    int64_t i;
    for (i = is; i < ie; ++i) {
        T[i] = Tc;
        w[i] = Tc * alpha[i];
        Tc = Tc * (1.0 - alpha[i]);
        if (Tc < 1e-3) {
            i += 1;
            break;
        }
    }
    i_end[r] = i;
    t_last[r] = Tc;
}

__global__ void k_alpha2weight_bwd64(const double* __restrict__ alpha,
                                     const double* __restrict__ w, const double* __restrict__ T,
                                     const double* __restrict__ t_last,
                                     const int64_t* __restrict__ i_start,
                                     const int64_t* __restrict__ i_end, int64_t n_rays,
                                     const double* __restrict__ gw,
                                     const double* __restrict__ glast, double* __restrict__ ga) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rays) return;
    double back = glast[r] * t_last[r];
    for (int64_t i = i_end[r] - 1; i >= i_start[r]; --i) {
        ga[i] = gw[i] * T[i] - back / ((1.0 - alpha[i]) + 1e-10);
        back += gw[i] * w[i];
    }
}

}  // namespace

extern "C" {

int pvdb_trilinear_fwd_f64(const pvdb_topology* topo, const double* vals, int32_t channels,
                           const double* xs, const double* ys, const double* zs, int64_t n,
                           double* out, void* stream) {
    if (!topo) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    PVDB_CHECK_ARG(channels == 1 || channels == 3 || channels == 12, "channels must be 1, 3 or 12");
    PVDB_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PVDB_OK;
    PVDB_CHECK_ARG(vals && xs && ys && zs && out, "null pointer");
    const TopoView t = pvdb::view_of(topo);
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid(pvdb_blocks(n, 256)), block(256);
    if (channels == 1)
        hipLaunchKernelGGL(k_tri_fwd64<1>, grid, block, 0, s, t, vals, xs, ys, zs, n, out);
    else if (channels == 3)
        hipLaunchKernelGGL(k_tri_fwd64<3>, grid, block, 0, s, t, vals, xs, ys, zs, n, out);
    else
        hipLaunchKernelGGL(k_tri_fwd64<12>, grid, block, 0, s, t, vals, xs, ys, zs, n, out);
    PVDB_LAUNCHED("k_tri_fwd64");
    return PVDB_OK;
}

int pvdb_trilinear_bwd_f64(const pvdb_topology* topo, double* grad, int32_t channels,
                           const double* xs, const double* ys, const double* zs, const double* g,
                           int64_t n, void* stream) {
    if (!topo) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    PVDB_CHECK_ARG(channels == 1 || channels == 3 || channels == 12, "channels must be 1, 3 or 12");
    PVDB_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PVDB_OK;
    PVDB_CHECK_ARG(grad && xs && ys && zs && g, "null pointer");
    const TopoView t = pvdb::view_of(topo);
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid(pvdb_blocks(n, 256)), block(256);
    if (channels == 1)
        hipLaunchKernelGGL(k_tri_bwd64<1>, grid, block, 0, s, t, grad, xs, ys, zs, g, n);
    else if (channels == 3)
        hipLaunchKernelGGL(k_tri_bwd64<3>, grid, block, 0, s, t, grad, xs, ys, zs, g, n);
    else
        hipLaunchKernelGGL(k_tri_bwd64<12>, grid, block, 0, s, t, grad, xs, ys, zs, g, n);
    PVDB_LAUNCHED("k_tri_bwd64");
    return PVDB_OK;
}

int pvdb_raw2alpha_f64(const double* sigma, int64_t n, double shift, double interval,
                       double* exp_out, double* alpha_out, void* stream) {
    PVDB_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PVDB_OK;
    PVDB_CHECK_ARG(sigma && exp_out && alpha_out, "null pointer");
    hipLaunchKernelGGL(k_raw2alpha64, dim3(pvdb_blocks(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, sigma, n, shift, interval, exp_out, alpha_out);
    PVDB_LAUNCHED("k_raw2alpha64");
    return PVDB_OK;
}

int pvdb_raw2alpha_bwd_f64(const double* exp_d, const double* grad_alpha, int64_t n,
                           double interval, double* grad_sigma, void* stream) {
    PVDB_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PVDB_OK;
    PVDB_CHECK_ARG(exp_d && grad_alpha && grad_sigma, "null pointer");
    hipLaunchKernelGGL(k_raw2alpha_bwd64, dim3(pvdb_blocks(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, exp_d, grad_alpha, n, interval, grad_sigma);
    PVDB_LAUNCHED("k_raw2alpha_bwd64");
    return PVDB_OK;
}

int pvdb_alpha2weight_f64(const double* alpha, int64_t n_rays, double* weight, double* T,
                          double* alphainv_last, const int64_t* i_start, int64_t* i_end,
                          void* stream) {
    PVDB_CHECK_ARG(n_rays >= 0, "n_rays must be >= 0");
    if (n_rays == 0) return PVDB_OK;
    PVDB_CHECK_ARG(alpha && weight && T && alphainv_last && i_start && i_end, "null pointer");
    hipLaunchKernelGGL(k_alpha2weight64, dim3(pvdb_blocks(n_rays, 256)), dim3(256), 0,
                       (hipStream_t)stream, alpha, n_rays, weight, T, alphainv_last, i_start,
                       i_end);
    PVDB_LAUNCHED("k_alpha2weight64");
    return PVDB_OK;
}

int pvdb_alpha2weight_bwd_f64(const double* alpha, const double* weight, const double* T,
                              const double* alphainv_last, const int64_t* i_start,
                              const int64_t* i_end, int64_t n_rays, const double* grad_weights,
                              const double* grad_last, double* grad_alpha, void* stream) {
    PVDB_CHECK_ARG(n_rays >= 0, "n_rays must be >= 0");
    if (n_rays == 0) return PVDB_OK;
    PVDB_CHECK_ARG(alpha && weight && T && alphainv_last && i_start && i_end && grad_weights &&
                       grad_last && grad_alpha,
                   "null pointer");
    hipLaunchKernelGGL(k_alpha2weight_bwd64, dim3(pvdb_blocks(n_rays, 256)), dim3(256), 0,
                       (hipStream_t)stream, alpha, weight, T, alphainv_last, i_start, i_end,
                       n_rays, grad_weights, grad_last, grad_alpha);
    PVDB_LAUNCHED("k_alpha2weight_bwd64");
    return PVDB_OK;
}

}  // extern "C"
