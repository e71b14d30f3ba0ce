// Per-point and per-slot grid kernels: trilinear gather (forward) and atomic
// scatter (backward) of a VDB payload, dense <-> VDB copies, the VDB Adam step
// and the dense Adam step of the MLP parameters.
//
// These are the kernels behind the reference-compatible plenvdb API (one
// launch per call, reference operation order).  The fused training step lives
// in fused_step.hip.
#include <type_traits>

#include "pvdb_common.h"
#include "gather.h"
#include "plenvdb_hip.h"

namespace pvdb {
TopoView view_of(const pvdb_topology* t);
}

using pvdb::TopoView;
using pvdb::Cell;
using pvdb::LeafCache;
using pvdb::kCx;
using pvdb::kCy;
using pvdb::kCz;
using pvdb::make_cell;

namespace {

// ORDER 0: out += ((val*a)*b)*c          (densityvdb.cu:115-122)
// ORDER 1: out[c] += val[c]*((a*b)*c)    (colorvdb.cu:16-26,96-103)
template <int C, int ORDER>
__global__ void __launch_bounds__(256) k_tri_fwd(TopoView t, const float* __restrict__ vals,
                                                 int cdyn, const float* __restrict__ xs,
                                                 const float* __restrict__ ys,
                                                 const float* __restrict__ zs, int64_t n,
                                                 float* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int CC = C > 0 ? C : cdyn;
    const Cell cell = make_cell(xs[p], ys[p], zs[p]);
    LeafCache cache;
    float acc[C > 0 ? C : 48];
#pragma unroll
    for (int c = 0; c < (C > 0 ? C : 48); ++c) acc[c] = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int dx = (kCx >> q) & 1, dy = (kCy >> q) & 1, dz = (kCz >> q) & 1;
        const float a = dx ? cell.u : 1.f - cell.u;
        const float b = dy ? cell.v : 1.f - cell.v;
        const float d = dz ? cell.w : 1.f - cell.w;
        const int64_t s = cache.slot(t, cell.i0 + dx, cell.j0 + dy, cell.k0 + dz);
        if (ORDER == 0) {
            const float val = s >= 0 ? vals[s * CC] : 0.f;
            acc[0] = acc[0] + ((val * a) * b) * d;
        } else {
            const float sc = (a * b) * d;
            const float* src = vals + (s >= 0 ? s : 0) * CC;
#pragma unroll
            for (int c = 0; c < (C > 0 ? C : 48); ++c) {
                if (c < CC) {
                    const float val = s >= 0 ? src[c] : 0.f;
                    acc[c] = acc[c] + val * sc;
                }
            }
        }
    }
    float* o = out + p * CC;
#pragma unroll
    for (int c = 0; c < (C > 0 ? C : 48); ++c)
        if (c < CC) o[c] = acc[c];
}

// Scatter with the "leaf must exist" rule (densityvdb.cu:16-27,143-181;
// colorvdb.cu:28-37,130-175): corners in a missing leaf are dropped, voxels of
// existing leaves receive the add whether active or not.
template <int C, int ORDER>
__global__ void __launch_bounds__(256) k_tri_bwd(TopoView t, float* __restrict__ grad, int cdyn,
                                                 const float* __restrict__ xs,
                                                 const float* __restrict__ ys,
                                                 const float* __restrict__ zs,
                                                 const float* __restrict__ g, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int CC = C > 0 ? C : cdyn;
    const Cell cell = make_cell(xs[p], ys[p], zs[p]);
    const float* gp = g + p * CC;
    float gv[C > 0 ? C : 48];
#pragma unroll
    for (int c = 0; c < (C > 0 ? C : 48); ++c) gv[c] = c < CC ? gp[c] : 0.f;
    LeafCache cache;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int dx = (kCx >> q) & 1, dy = (kCy >> q) & 1, dz = (kCz >> q) & 1;
        const float a = dx ? cell.u : 1.f - cell.u;
        const float b = dy ? cell.v : 1.f - cell.v;
        const float d = dz ? cell.w : 1.f - cell.w;
        const int64_t s = cache.slot(t, cell.i0 + dx, cell.j0 + dy, cell.k0 + dz);
        if (s < 0) continue;
        float* dst = grad + s * CC;
        if (ORDER == 0) {
            atomicAdd(dst, ((gv[0] * a) * b) * d);
        } else {
            const float sc = (a * b) * d;
#pragma unroll
            for (int c = 0; c < (C > 0 ? C : 48); ++c)
                if (c < CC) atomicAdd(dst + c, gv[c] * sc);
        }
    }
}

__device__ __forceinline__ bool slot_active(const TopoView& t, int64_t slot) {
    const int64_t leaf = slot >> 9;
    const int off = (int)(slot & 511);
    return (t.masks[leaf * 8 + (off >> 6)] >> (off & 63)) & 1ull;
}

__device__ __forceinline__ void slot_coord(const TopoView& t, int64_t slot, int& x, int& y,
                                           int& z) {
    const int64_t leaf = slot >> 9;
    const int off = (int)(slot & 511);
    x = t.origins[leaf * 3 + 0] + (off >> 6);
    y = t.origins[leaf * 3 + 1] + ((off >> 3) & 7);
    z = t.origins[leaf * 3 + 2] + (off & 7);
}

// copyFromDense (densityvdb.cu:31-48, colorvdb.cu:42-62): active slot <-
// dense[x][y][z][:] with z fastest.
__global__ void k_copy_from_dense(TopoView t, float* __restrict__ vals, int C,
                                  const float* __restrict__ dense, int rx, int ry, int rz) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= t.nleaf * 512) return;
    if (!slot_active(t, s)) return;
    int x, y, z;
    slot_coord(t, s, x, y, z);
    if (x < 0 || y < 0 || z < 0 || x >= rx || y >= ry || z >= rz) return;
    const float* src = dense + (((int64_t)x * ry + y) * rz + z) * C;
    for (int c = 0; c < C; ++c) vals[s * C + c] = src[c];
}

// Dense copy of the grid values over [0,rx)x[0,ry)x[0,rz): every voxel of an
// existing leaf (active or not) reads its stored value, others the background
// (OpenVDB tools::copyToDense semantics used by plenvdb.h:149-160,242-262).
__global__ void k_copy_to_dense(TopoView t, const float* __restrict__ vals, int C,
                                float* __restrict__ dense, int rx, int ry, int rz,
                                float background) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nvox = (int64_t)rx * ry * rz;
    if (q >= nvox) return;
    const int z = (int)(q % rz);
    const int y = (int)((q / rz) % ry);
    const int x = (int)(q / ((int64_t)ry * rz));
    const int64_t s = pvdb::slot_of(t, x, y, z);
    for (int c = 0; c < C; ++c) dense[q * C + c] = s >= 0 ? vals[s * C + c] : background;
}

// setValuesOn_bymask (densityvdb.cu:64-83): active slot <- val where mask.
__global__ void k_set_by_mask(TopoView t, float* __restrict__ vals,
                              const uint8_t* __restrict__ mask, float val, int rx, int ry,
                              int rz) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= t.nleaf * 512) return;
    if (!slot_active(t, s)) return;
    int x, y, z;
    slot_coord(t, s, x, y, z);
    if (x < 0 || y < 0 || z < 0 || x >= rx || y >= ry || z >= rz) return;
    if (mask[((int64_t)x * ry + y) * rz + z]) vals[s] = val;
}

// zero_grad (densityvdb.cu:353-361, colorvdb.cu:355-364): active slots only.
__global__ void k_zero_active(TopoView t, float* __restrict__ grad, int C) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= t.nleaf * 512) return;
    if (!slot_active(t, s)) return;
    for (int c = 0; c < C; ++c) grad[s * C + c] = 0.f;
}

// Total-variation gradient of a VDB payload: the dense kernel's rule
// (total_variation_kernel.cu:13-35) on the grid's dense view over
// [0,rx)x[0,ry)x[0,rz) (stored value in an existing leaf, 0 elsewhere), applied
// to every slot inside that box.  Terms in the dense kernel's order z-1, z+1,
// y-1, y+1, x-1, x+1; the x axis takes wz, as there (lines 31-32).  One thread
// per (slot, V consecutive channels), so neighbouring threads read neighbouring
// 4V-byte pieces; a thread reads its own and its neighbours' values and writes
// only its own gradient, so no two threads touch one element.
template <int V>
__global__ void __launch_bounds__(256) k_tv_add_grad(TopoView t, const float* __restrict__ vals,
                                                     float* __restrict__ grad, int C, int rx,
                                                     int ry, int rz, float wy6, float wz6,
                                                     int dense_mode) {
    const int G = C / V;  // channel groups per slot
    // XCD-aware block order: workgroups are dealt to the 8 XCDs in turn, so
    // block b of the first 8*(n/8) is given the run (b % 8)*(n/8) + b/8 and each
    // XCD's L2 holds a contiguous range of slots with their in-leaf neighbours
    // (a bijection; the last n % 8 blocks keep their index)
    const unsigned nblk = gridDim.x, per = nblk / 8, bid = blockIdx.x;
    const unsigned lblk = bid < 8 * per ? (bid % 8) * per + bid / 8 : bid;
    const int64_t tid = (int64_t)lblk * blockDim.x + threadIdx.x;
    if (tid >= t.nleaf * 512 * G) return;
    const int64_t s = tid / G;
    const int q = (int)(tid - s * G);
    int x, y, z;
    slot_coord(t, s, x, y, z);
    if (x < 0 || y < 0 || z < 0 || x >= rx || y >= ry || z >= rz) return;
    using Vec = std::conditional_t<V == 4, float4, float>;
    auto ld = [&](const float* p, float* o) {
        const Vec w = *reinterpret_cast<const Vec*>(p);
        if constexpr (V == 4) {
            o[0] = w.x; o[1] = w.y; o[2] = w.z; o[3] = w.w;
        } else {
            o[0] = w;
        }
    };
    const int64_t e = s * C + (int64_t)q * V;
    float g0[V];
    ld(grad + e, g0);
    bool any = dense_mode != 0;
#pragma unroll
    for (int c = 0; c < V; ++c) any = any || g0[c] != 0.f;
    if (!any) return;
    // neighbour slots: -2 past the box (no term), -1 in a missing leaf (value 0);
    // inside this leaf the neighbour is slot +-1, +-8, +-64 (offset x<<6|y<<3|z),
    // only a leaf face goes through the leaf table
    const int lx = x & 7, ly = y & 7, lz = z & 7;
    int64_t nb[6];
    nb[0] = z > 0 ? (lz != 0 ? s - 1 : pvdb::slot_of(t, x, y, z - 1)) : -2;
    nb[1] = z < rz - 1 ? (lz != 7 ? s + 1 : pvdb::slot_of(t, x, y, z + 1)) : -2;
    nb[2] = y > 0 ? (ly != 0 ? s - 8 : pvdb::slot_of(t, x, y - 1, z)) : -2;
    nb[3] = y < ry - 1 ? (ly != 7 ? s + 8 : pvdb::slot_of(t, x, y + 1, z)) : -2;
    nb[4] = x > 0 ? (lx != 0 ? s - 64 : pvdb::slot_of(t, x - 1, y, z)) : -2;
    nb[5] = x < rx - 1 ? (lx != 7 ? s + 64 : pvdb::slot_of(t, x + 1, y, z)) : -2;
    const float w[6] = {wz6, wz6, wy6, wy6, wz6, wz6};
    float v[V], acc[V];
    ld(vals + e, v);
#pragma unroll
    for (int c = 0; c < V; ++c) acc[c] = 0.f;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
        if (nb[n] == -2) continue;
        float u[V];
        if (nb[n] >= 0) {
            ld(vals + nb[n] * C + (int64_t)q * V, u);
        } else {
#pragma unroll
            for (int c = 0; c < V; ++c) u[c] = 0.f;
        }
#pragma unroll
        for (int c = 0; c < V; ++c) acc[c] += w[n] * fminf(fmaxf(v[c] - u[c], -1.f), 1.f);
    }
    float r[V];
#pragma unroll
    for (int c = 0; c < V; ++c) r[c] = (dense_mode || g0[c] != 0.f) ? g0[c] + acc[c] : g0[c];
    if constexpr (V == 4)
        *reinterpret_cast<float4*>(grad + e) = make_float4(r[0], r[1], r[2], r[3]);
    else
        grad[e] = r[0];
}

// VDB Adam over all slots, one thread per (slot, group of `group` channels)
// (densityvdb.cu:185-349, colorvdb.cu:179-351).
__global__ void k_vdb_adam(TopoView t, float* __restrict__ param, const float* __restrict__ grad,
                           float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ per_lr, int C, int group, int mode,
                           float stepsz, float eps, float b0, float b1) {
    const int ngroups = C / group;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= t.nleaf * 512 * ngroups) return;
    const int64_t s = q / ngroups;
    const int grp = (int)(q - s * ngroups);
    if (!slot_active(t, s)) return;
    const int64_t base = s * C + (int64_t)grp * group;
    if (mode == 1) {
        bool all0 = true;
        for (int c = 0; c < group; ++c) all0 &= grad[base + c] == 0.f;
        if (all0) return;
    }
    float sz = stepsz;
    if (mode == 2) sz = per_lr[s] * stepsz;
    for (int c = 0; c < group; ++c) {
        const float g = grad[base + c];
        const float nm = b0 * m[base + c] + (1.f - b0) * g;
        const float nv = b1 * v[base + c] + ((1.f - b1) * g) * g;
        m[base + c] = nm;
        v[base + c] = nv;
        param[base + c] = param[base + c] - (sz * nm) / (eps + sqrtf(nv));
    }
}

// Dense Adam (plenvdb/lib/cuda/adam_upd_kernel.cu:8-58).
__global__ void k_dense_adam(float* __restrict__ p, const float* __restrict__ g,
                             float* __restrict__ m, float* __restrict__ v,
                             const float* __restrict__ per_lr, int64_t n, int mode,
                             float step_size, float eps, float b0, float b1) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gg = g[q];
    if (mode == 1 && gg == 0.f) return;
    const float nm = b0 * m[q] + (1.f - b0) * gg;
    const float nv = b1 * v[q] + ((1.f - b1) * gg) * gg;
    m[q] = nm;
    v[q] = nv;
    float sz = step_size;
    if (mode == 2) sz = sz * per_lr[q];
    p[q] = p[q] - (sz * nm) / (sqrtf(nv) + eps);
}

template <int C, int ORDER>
void launch_fwd(const TopoView& t, const float* vals, int cdyn, const float* xs, const float* ys,
                const float* zs, int64_t n, float* out, hipStream_t s) {
    hipLaunchKernelGGL((k_tri_fwd<C, ORDER>), dim3(pvdb_blocks(n, 256)), dim3(256), 0, s, t, vals,
                       cdyn, xs, ys, zs, n, out);
}
template <int C, int ORDER>
void launch_bwd(const TopoView& t, float* grad, int cdyn, const float* xs, const float* ys,
                const float* zs, const float* g, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_tri_bwd<C, ORDER>), dim3(pvdb_blocks(n, 256)), dim3(256), 0, s, t, grad,
                       cdyn, xs, ys, zs, g, n);
}

int check_topo(const pvdb_topology* t) {
    if (!t) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    return PVDB_OK;
}

}  // namespace

extern "C" {

int pvdb_trilinear_fwd(const pvdb_topology* topo, const float* vals, int32_t channels,
                       int32_t order, const float* xs, const float* ys, const float* zs,
                       int64_t n, float* out, void* stream) {
    if (int rc = check_topo(topo)) return rc;
    PVDB_CHECK_ARG(channels >= 1 && channels <= 48, "channels must be in [1, 48]");
    PVDB_CHECK_ARG(order == 0 || order == 1, "order must be 0 (density) or 1 (color)");
    PVDB_CHECK_ARG(order == 1 || channels == 1, "density order needs one channel");
    PVDB_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PVDB_OK;
    PVDB_CHECK_ARG(vals && xs && ys && zs && out, "null pointer");
    const TopoView t = pvdb::view_of(topo);
    hipStream_t s = (hipStream_t)stream;
    if (order == 0)
        launch_fwd<1, 0>(t, vals, 1, xs, ys, zs, n, out, s);
    else if (channels == 1)
        launch_fwd<1, 1>(t, vals, 1, xs, ys, zs, n, out, s);
    else if (channels == 3)
        launch_fwd<3, 1>(t, vals, 3, xs, ys, zs, n, out, s);
    else if (channels == 12)
        launch_fwd<12, 1>(t, vals, 12, xs, ys, zs, n, out, s);
    else
        launch_fwd<0, 1>(t, vals, channels, xs, ys, zs, n, out, s);
    PVDB_LAUNCHED("k_tri_fwd");
    return PVDB_OK;
}

int pvdb_trilinear_bwd(const pvdb_topology* topo, float* grad, int32_t channels, int32_t order,
                       const float* xs, const float* ys, const float* zs, const float* g,
                       int64_t n, void* stream) {
    if (int rc = check_topo(topo)) return rc;
    PVDB_CHECK_ARG(channels >= 1 && channels <= 48, "channels must be in [1, 48]");
    PVDB_CHECK_ARG(order == 0 || order == 1, "order must be 0 (density) or 1 (color)");
    PVDB_CHECK_ARG(order == 1 || channels == 1, "density order needs one channel");
    PVDB_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PVDB_OK;
    PVDB_CHECK_ARG(grad && xs && ys && zs && g, "null pointer");
    const TopoView t = pvdb::view_of(topo);
    hipStream_t s = (hipStream_t)stream;
    if (order == 0)
        launch_bwd<1, 0>(t, grad, 1, xs, ys, zs, g, n, s);
    else if (channels == 1)
        launch_bwd<1, 1>(t, grad, 1, xs, ys, zs, g, n, s);
    else if (channels == 3)
        launch_bwd<3, 1>(t, grad, 3, xs, ys, zs, g, n, s);
    else if (channels == 12)
        launch_bwd<12, 1>(t, grad, 12, xs, ys, zs, g, n, s);
    else
        launch_bwd<0, 1>(t, grad, channels, xs, ys, zs, g, n, s);
    PVDB_LAUNCHED("k_tri_bwd");
    return PVDB_OK;
}

int pvdb_copy_from_dense(const pvdb_topology* topo, float* vals, int32_t channels,
                         const float* dense, int32_t rx, int32_t ry, int32_t rz, void* stream) {
    if (int rc = check_topo(topo)) return rc;
    PVDB_CHECK_ARG(channels >= 1 && rx > 0 && ry > 0 && rz > 0, "bad sizes");
    PVDB_CHECK_ARG(vals && dense, "null pointer");
    const TopoView t = pvdb::view_of(topo);
    if (t.nleaf == 0) return PVDB_OK;
    hipLaunchKernelGGL(k_copy_from_dense, dim3(pvdb_blocks(t.nleaf * 512, 256)), dim3(256), 0,
                       (hipStream_t)stream, t, vals, channels, dense, rx, ry, rz);
    PVDB_LAUNCHED("k_copy_from_dense");
    return PVDB_OK;
}

int pvdb_copy_to_dense(const pvdb_topology* topo, const float* vals, int32_t channels,
                       float* dense, int32_t rx, int32_t ry, int32_t rz, float background,
                       void* stream) {
    if (int rc = check_topo(topo)) return rc;
    PVDB_CHECK_ARG(channels >= 1 && rx > 0 && ry > 0 && rz > 0, "bad sizes");
    PVDB_CHECK_ARG(vals && dense, "null pointer");
    const TopoView t = pvdb::view_of(topo);
    const int64_t nvox = (int64_t)rx * ry * rz;
    hipLaunchKernelGGL(k_copy_to_dense, dim3(pvdb_blocks(nvox, 256)), dim3(256), 0,
                       (hipStream_t)stream, t, vals, channels, dense, rx, ry, rz, background);
    PVDB_LAUNCHED("k_copy_to_dense");
    return PVDB_OK;
}

int pvdb_set_values_by_mask(const pvdb_topology* topo, float* vals, const uint8_t* mask,
                            float value, int32_t rx, int32_t ry, int32_t rz, void* stream) {
    if (int rc = check_topo(topo)) return rc;
    PVDB_CHECK_ARG(rx > 0 && ry > 0 && rz > 0, "bad sizes");
    PVDB_CHECK_ARG(vals && mask, "null pointer");
    const TopoView t = pvdb::view_of(topo);
    if (t.nleaf == 0) return PVDB_OK;
    hipLaunchKernelGGL(k_set_by_mask, dim3(pvdb_blocks(t.nleaf * 512, 256)), dim3(256), 0,
                       (hipStream_t)stream, t, vals, mask, value, rx, ry, rz);
    PVDB_LAUNCHED("k_set_by_mask");
    return PVDB_OK;
}

int pvdb_tv_add_grad(const pvdb_topology* topo, const float* vals, float* grad,
                     int32_t channels, int32_t rx, int32_t ry, int32_t rz, float wx, float wy,
                     float wz, int32_t dense_mode, void* stream) {
    if (int rc = check_topo(topo)) return rc;
    PVDB_CHECK_ARG(channels >= 1 && rx > 0 && ry > 0 && rz > 0, "bad sizes");
    PVDB_CHECK_ARG(vals && grad, "null pointer");
    (void)wx;  // the dense kernel weighs the x axis with wz
    const TopoView t = pvdb::view_of(topo);
    if (t.nleaf == 0) return PVDB_OK;
    // weights scaled on the host as total_variation_kernel.cu:45-47 does; four
    // channels per thread when the row splits into 16-byte pieces
    const bool vec4 = channels % 4 == 0 && reinterpret_cast<uintptr_t>(vals) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(grad) % 16 == 0;
    const int64_t nthreads = t.nleaf * 512 * (vec4 ? channels / 4 : channels);
    if (vec4)
        hipLaunchKernelGGL(k_tv_add_grad<4>, dim3(pvdb_blocks(nthreads, 256)), dim3(256), 0,
                           (hipStream_t)stream, t, vals, grad, channels, rx, ry, rz, wy / 6.f,
                           wz / 6.f, dense_mode != 0 ? 1 : 0);
    else
        hipLaunchKernelGGL(k_tv_add_grad<1>, dim3(pvdb_blocks(nthreads, 256)), dim3(256), 0,
                           (hipStream_t)stream, t, vals, grad, channels, rx, ry, rz, wy / 6.f,
                           wz / 6.f, dense_mode != 0 ? 1 : 0);
    PVDB_LAUNCHED("k_tv_add_grad");
    return PVDB_OK;
}

int pvdb_zero_active(const pvdb_topology* topo, float* grad, int32_t channels, void* stream) {
    if (int rc = check_topo(topo)) return rc;
    PVDB_CHECK_ARG(channels >= 1 && grad, "bad arguments");
    const TopoView t = pvdb::view_of(topo);
    if (t.nleaf == 0) return PVDB_OK;
    hipLaunchKernelGGL(k_zero_active, dim3(pvdb_blocks(t.nleaf * 512, 256)), dim3(256), 0,
                       (hipStream_t)stream, t, grad, channels);
    PVDB_LAUNCHED("k_zero_active");
    return PVDB_OK;
}

int pvdb_vdb_adam(const pvdb_topology* topo, float* param, const float* grad, float* exp_avg,
                  float* exp_avg_sq, const float* per_lr, int32_t channels, int32_t group,
                  int32_t mode, float stepsz, float eps, float beta0, float beta1,
                  void* stream) {
    if (int rc = check_topo(topo)) return rc;
    PVDB_CHECK_ARG(channels >= 1 && group >= 1 && channels % group == 0, "bad channel grouping");
    PVDB_CHECK_ARG(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
    PVDB_CHECK_ARG(mode != 2 || per_lr, "mode 2 needs per-voxel lr");
    PVDB_CHECK_ARG(param && grad && exp_avg && exp_avg_sq, "null pointer");
    const TopoView t = pvdb::view_of(topo);
    if (t.nleaf == 0) return PVDB_OK;
    const int64_t n = t.nleaf * 512 * (channels / group);
    hipLaunchKernelGGL(k_vdb_adam, dim3(pvdb_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       t, param, grad, exp_avg, exp_avg_sq, per_lr, channels, group, mode, stepsz,
                       eps, beta0, beta1);
    PVDB_LAUNCHED("k_vdb_adam");
    return PVDB_OK;
}

int pvdb_dense_adam(float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                    const float* per_lr, int64_t n, int32_t mode, float step_size, float eps,
                    float beta0, float beta1, void* stream) {
    PVDB_CHECK_ARG(n >= 0, "n must be >= 0");
    PVDB_CHECK_ARG(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
    PVDB_CHECK_ARG(mode != 2 || per_lr, "mode 2 needs per-element lr");
    if (n == 0) return PVDB_OK;
    PVDB_CHECK_ARG(param && grad && exp_avg && exp_avg_sq, "null pointer");
    hipLaunchKernelGGL(k_dense_adam, dim3(pvdb_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       param, grad, exp_avg, exp_avg_sq, per_lr, n, mode, step_size, eps, beta0,
                       beta1);
    PVDB_LAUNCHED("k_dense_adam");
    return PVDB_OK;
}

}  // extern "C"
