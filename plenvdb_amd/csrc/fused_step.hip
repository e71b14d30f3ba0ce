// Fused (batched) training step: C-ABI entry point, workspace layout, the
// sparse VDB Adam over touched leaves and the per-step fills.
//
// pvdb_fused_step runs one iteration of the reference's training loop
// (plenvdb/run.py:541-598: forward, MSE loss, backward, MaskedAdam on the MLP,
// VDBAdam on the grids) as a fixed sequence of kernel launches on one stream
// with no host synchronisation and no allocation, so a caller can capture it
// in a hipGraph.  All scratch lives in the caller's workspace.
#include <cmath>
#include <utility>
#include <vector>

#include "fused_common.h"
#include "plenvdb_hip.h"

namespace pvdb {
TopoView view_of(const pvdb_topology* t);
int launch_ray_count(const StepParams& P, const TopoView& t, const RayBuffers& R,
                     const PackArgs& pk, hipStream_t s);
int launch_ray_march(const StepParams& P, const TopoView& t, const float* density,
                     const RayBuffers& R, const Records& rec, hipStream_t s);
int launch_ray_composite(const StepParams& P, const RayBuffers& R, const Records& rec,
                         hipStream_t s);
int launch_direct_composite(int nc, const StepParams& P, const TopoView& t, const float* color,
                            const RayBuffers& R, const Records& rec, hipStream_t s);
int launch_mlp(bool bwd, bool bf16col, int grid, const StepParams& P, const TopoView& t,
               const void* color, const MlpPacked& W, const RayBuffers& R, const Records& rec,
               float* dw_partial, hipStream_t s);
int launch_dw_reduce(const float* part, int nparts, float* dw, const MlpAdam& adam,
                     hipStream_t s);
int launch_scatter(int nc, const TopoView& t, const RayBuffers& R, const Records& rec,
                   int64_t capacity, const GradRows& G, uint8_t* leaf_touched, int grid_req,
                   int chunk_records, const ReduceArgs& red, hipStream_t s);
}  // namespace pvdb

using namespace pvdb;

namespace {

constexpr int kMaxMlpGrid = 512;   // workgroups of the MLP kernels at most (dW partials)
constexpr int kFwdGrid = 512;      // default forward-MLP workgroups
constexpr int kScatterChunk = 64;  // default scatter records per chunk
constexpr int kScatterGrid = 1024; // default scatter workgroups
constexpr int kRayWaves = 4;       // default rays per workgroup of the wave-per-ray kernels
// Records are indexed through int32 offsets (saturating at INT32_MAX), so a
// capacity below that bound keeps every stored offset of a dropped record
// outside the record arrays.
constexpr int64_t kMaxCapacity = 0x7fffffffll;

inline bool aligned(const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

struct Layout {
    size_t count, offset, total, t_last, ray_out, ray_grad, ray_ctot, ray_loss, ray_rgb, loss, steps,
        pe, setup;
    size_t xyze, ab, rgb, adj, xrow, relu, grad;
    size_t touched, dw_part, w0t, w1t, w2t, w2, w1, w0;
    size_t bytes;
};

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

Layout layout_for(int64_t n_rays, int64_t capacity, int64_t nleaf) {
    Layout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.count = take(4 * n_rays);
    L.offset = take(4 * n_rays);
    L.total = take(4);
    L.t_last = take(4 * n_rays);
    L.ray_out = take(16 * n_rays);
    L.ray_grad = take(16 * n_rays);
    L.ray_ctot = take(16 * n_rays);
    L.ray_loss = take(4 * n_rays);
    L.ray_rgb = take(4 * 13 * n_rays);
    L.loss = take(4);
    L.steps = take(2 * (size_t)kStepCache * n_rays);
    L.pe = take(2 * 96 * (size_t)n_rays);
    L.setup = take(32 * (size_t)n_rays);
    L.xyze = take(16 * capacity);
    L.ab = take(8 * capacity);
    L.rgb = take(16 * capacity);
    L.adj = take(16 * capacity);
    L.xrow = take(80 * capacity);
    L.relu = take(32 * capacity);
    L.grad = take(64 * capacity);
    L.touched = take(nleaf > 0 ? nleaf : 1);
    L.dw_part = take((size_t)4 * kPartStride * kMaxMlpGrid);
    L.w0t = take(2 * 128 * 64);
    L.w1t = take(2 * 128 * 128);
    L.w2t = take(2 * 16 * 128);
    L.w2 = take(2 * 128 * 32);
    L.w1 = take(2 * 128 * 128);
    L.w0 = take(2 * 16 * 128);
    L.bytes = o;
    return L;
}

__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ void adam1(float& p, float g, float& m, float& v, float sz, float eps,
                                      float b0, float b1) {
    const float nm = b0 * m + (1.f - b0) * g;
    const float nv = b1 * v + ((1.f - b1) * g) * g;
    m = nm;
    v = nv;
    p = p - (sz * nm) / (eps + sqrtf(nv));
}

// A row of NC fp32 values of slot s, rows `stride` floats apart (16-byte
// vectors when NC is a multiple of 4; the stride is then one too).
template <int NC>
__device__ __forceinline__ void ld_row(const float* base, int64_t s, float (&v)[NC],
                                       int64_t stride = NC) {
    if constexpr (NC % 4 == 0) {
        const float4* p = reinterpret_cast<const float4*>(base + s * stride);
#pragma unroll
        for (int k = 0; k < NC / 4; ++k) {
            const float4 q = p[k];
            v[4 * k] = q.x;
            v[4 * k + 1] = q.y;
            v[4 * k + 2] = q.z;
            v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = base[s * stride + c];
    }
}

template <int NC>
__device__ __forceinline__ void st_row(float* base, int64_t s, const float (&v)[NC],
                                       int64_t stride = NC) {
    if constexpr (NC % 4 == 0) {
        float4* p = reinterpret_cast<float4*>(base + s * stride);
#pragma unroll
        for (int k = 0; k < NC / 4; ++k)
            p[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
    } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) base[s * stride + c] = v[c];
    }
}

// VDB Adam of the step (densityvdb.cu:185-349, colorvdb.cu:179-351), then the
// gradients and touched flags re-zeroed.  One workgroup per leaf.
//   density stepmode 1 (skip-zero-grad) updates a voxel whose gradient is
//   nonzero; 0 updates every active voxel; 2 every active voxel with its step
//   scaled by the per-voxel rate, (rate * step) * m / (eps + sqrt(v)).
//   colour stepmode 1 skips a Vec3 group only when all three components are
//   zero; 0 updates every active voxel.
// With both grids in stepmode 1 only the leaves the backward touched are
// visited: the others have all-zero gradients, which stepmode 1 skips, so the
// result equals the reference's all-slot pass.  NC colour channels per slot.
// With a bf16 colour payload and an fp32 master copy (cmaster != nullptr) the
// update runs on the master and the payload receives its round-to-nearest
// bf16 value, so updates below half a bf16 step accumulate instead of being
// lost; without one the bf16 payload is updated in place.
struct GridAdamModes {
    int density, color;
    const float* rate;  // [slots] per-voxel rate of density stepmode 2
};

template <int NC, bool BF16COL>
__global__ void __launch_bounds__(256) k_sparse_adam(TopoView t, uint8_t* __restrict__ touched,
                                                     float* __restrict__ den, GradRows G,
                                                     float* __restrict__ mden,
                                                     float* __restrict__ vden, void* col,
                                                     float* __restrict__ cmaster,
                                                     float* __restrict__ mcol,
                                                     float* __restrict__ vcol, GridAdamModes md_,
                                                     float sz_d_arg, float sz_c_arg,
                                                     const float* __restrict__ dev_sz, float eps,
                                                     float b0, float b1) {
    static_assert(!BF16COL || NC == 12, "bf16 colour payloads hold 12 channels");
    const float sz_d = dev_sz ? dev_sz[0] : sz_d_arg;
    const float sz_c = dev_sz ? dev_sz[1] : sz_c_arg;
    const bool every = md_.density != 1 || md_.color != 1;
    // a voxel's operands, loaded for both of a thread's voxels before either
    // is updated (two round trips in flight per thread)
    struct Vox {
        float gd, pd, md, vd;
        float gg[NC], mm[NC], vv[NC], pp[NC];
    };
    auto load = [&](int64_t s, Vox& v) {
        v.pd = den[s];
        v.md = mden[s];
        v.vd = vden[s];
        if (G.whole) {  // the row's vectors: colour, then the density
            float r[NC + 1 + (NC == 12 ? 3 : 0)];
            const float4* p = reinterpret_cast<const float4*>(G.color + s * G.cs);
#pragma unroll
            for (int kv = 0; kv < (NC + 1 + 3) / 4; ++kv) {
                const float4 q = p[kv];
                r[4 * kv] = q.x;
                r[4 * kv + 1] = q.y;
                r[4 * kv + 2] = q.z;
                r[4 * kv + 3] = q.w;
            }
#pragma unroll
            for (int c = 0; c < NC; ++c) v.gg[c] = r[c];
            v.gd = r[NC];
        } else {
            v.gd = G.density[s * G.ds];
            ld_row<NC>(G.color, s, v.gg, G.cs);
        }
        ld_row<NC>(mcol, s, v.mm);
        ld_row<NC>(vcol, s, v.vv);
        if constexpr (BF16COL) {
            if (!cmaster) {
                const uint2* c2 = reinterpret_cast<const uint2*>(col) + s * 3;
#pragma unroll
                for (int k4 = 0; k4 < 3; ++k4) {
                    const uint2 q = c2[k4];
                    v.pp[4 * k4] = bf_bits2f((uint16_t)(q.x & 0xffffu));
                    v.pp[4 * k4 + 1] = bf_bits2f((uint16_t)(q.x >> 16));
                    v.pp[4 * k4 + 2] = bf_bits2f((uint16_t)(q.y & 0xffffu));
                    v.pp[4 * k4 + 3] = bf_bits2f((uint16_t)(q.y >> 16));
                }
            } else {
                ld_row<NC>(cmaster, s, v.pp);
            }
        } else {
            ld_row<NC>(reinterpret_cast<const float*>(col), s, v.pp);
        }
    };
    // the update (a voxel whose gradients are all zero writes nothing back in
    // stepmode 1)
    auto update = [&](int64_t s, Vox& v) {
        if (md_.density != 1 || v.gd != 0.f) {
            const float sz = md_.density == 2 ? md_.rate[s] * sz_d : sz_d;
            adam1(v.pd, v.gd, v.md, v.vd, sz, eps, b0, b1);
            den[s] = v.pd;
            mden[s] = v.md;
            vden[s] = v.vd;
        }
        bool any = false;
#pragma unroll
        for (int grp = 0; grp < NC / 3; ++grp) {
            const int c0 = 3 * grp;
            if (md_.color == 1 && v.gg[c0] == 0.f && v.gg[c0 + 1] == 0.f && v.gg[c0 + 2] == 0.f)
                continue;
            any = true;
#pragma unroll
            for (int c = c0; c < c0 + 3; ++c)
                adam1(v.pp[c], v.gg[c], v.mm[c], v.vv[c], sz_c, eps, b0, b1);
        }
        if (any) {
            st_row<NC>(mcol, s, v.mm);
            st_row<NC>(vcol, s, v.vv);
            if constexpr (BF16COL) {
                uint2* c2 = reinterpret_cast<uint2*>(col) + s * 3;
#pragma unroll
                for (int k4 = 0; k4 < 3; ++k4)
                    c2[k4] = make_uint2((uint32_t)f2bf_bits(v.pp[4 * k4]) | ((uint32_t)f2bf_bits(v.pp[4 * k4 + 1]) << 16),
                                        (uint32_t)f2bf_bits(v.pp[4 * k4 + 2]) | ((uint32_t)f2bf_bits(v.pp[4 * k4 + 3]) << 16));
                if (cmaster) st_row<NC>(cmaster, s, v.pp);
            } else {
                st_row<NC>(reinterpret_cast<float*>(col), s, v.pp);
            }
        }
    };
    auto zero = [&](int64_t s) {
        if (G.whole) {  // whole-line stores
            float4* p = reinterpret_cast<float4*>(G.color + s * G.cs);
#pragma unroll
            for (int kv = 0; kv < (NC + 1 + 3) / 4; ++kv) p[kv] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            G.density[s * G.ds] = 0.f;
            float zr[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) zr[c] = 0.f;
            st_row<NC>(G.color, s, zr, G.cs);
        }
    };
    for (int64_t leaf = blockIdx.x; leaf < t.nleaf; leaf += gridDim.x) {
        if (!every && !touched[leaf]) continue;  // uniform over the workgroup
        // voxels k and k + 256 of the leaf (a 256-thread workgroup)
        const int k = threadIdx.x;
        const int64_t s0 = leaf * 512 + k, s1 = s0 + 256;
        const bool a0 = (t.masks[leaf * 8 + (k >> 6)] >> (k & 63)) & 1ull;
        const bool a1 = (t.masks[leaf * 8 + 4 + (k >> 6)] >> (k & 63)) & 1ull;
        Vox v0, v1;
        if (a0) load(s0, v0);
        if (a1) load(s1, v1);
        if (a0) update(s0, v0);
        if (a1) update(s1, v1);
        zero(s0);
        zero(s1);
        __syncthreads();
        if (threadIdx.x == 0) touched[leaf] = 0;
    }
}


// Optional per-stage device timing (HIP events on the step's stream).
constexpr int kStages = 9;
const char* const kStageNames[kStages] = {"pack_mlp", "ray_count", "ray_march",
                                          "mlp_fwd",  "composite", "mlp_bwd",
                                          "scatter",  "dw_reduce", "adam"};
struct StageTimer {
    bool enabled = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> open;
    double total_ms[kStages] = {0};
    int calls[kStages] = {0};
};
StageTimer g_timer;

struct StageScope {
    int stage;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    StageScope(int st, hipStream_t str) : stage(st), s(str) {
        if (!g_timer.enabled) return;
        hipEventCreate(&a);
        hipEventCreate(&b);
        hipEventRecord(a, s);
    }
    ~StageScope() {
        if (!g_timer.enabled || !a) return;
        hipEventRecord(b, s);
        g_timer.open.push_back({stage, {a, b}});
    }
};

// Colour channels per slot of the step's model (12 features, or k0_dim).
inline int slot_channels(const pvdb_step_desc* d) { return d->k0_dim == 3 ? 3 : 12; }

// The gradient arrays as rows: grad_stride 0 means separate arrays.
inline GradRows grad_rows_of(const pvdb_step_desc* d) {
    GradRows G;
    G.density = d->grad_density;
    G.color = d->grad_color;
    G.ds = d->grad_stride > 0 ? d->grad_stride : 1;
    G.cs = d->grad_stride > 0 ? d->grad_stride : slot_channels(d);
    const int nc = slot_channels(d);
    G.whole = d->grad_stride == (nc == 12 ? 16 : 4) && d->grad_density == d->grad_color + nc &&
              aligned(d->grad_color, 16);
    return G;
}

// VDB Adam of the grids (over the touched leaves when both run stepmode 1).
int launch_grid_adam(const pvdb_step_desc* d, const TopoView& t, uint8_t* touched, hipStream_t s) {
    const int ag = (int)std::min<int64_t>(t.nleaf > 0 ? t.nleaf : 1, 4096);
    float* master = d->color_bf16 ? d->color_master : nullptr;
    const GridAdamModes md{d->stepmode_density, d->stepmode_color, d->density_lr};
    const GradRows G = grad_rows_of(d);
#define PVDB_ADAM(NC, BF)                                                                        \
    hipLaunchKernelGGL((k_sparse_adam<NC, BF>), dim3(ag), dim3(256), 0, s, t, touched,          \
                       d->density, G, d->density_m, d->density_v, d->color, master, d->color_m, \
                       d->color_v, md, d->stepsz_density, d->stepsz_color, d->adam_scalars,     \
                       d->eps, d->beta0, d->beta1)
    if (d->k0_dim == 3)
        PVDB_ADAM(3, false);
    else if (d->color_bf16)
        PVDB_ADAM(12, true);
    else
        PVDB_ADAM(12, false);
#undef PVDB_ADAM
    PVDB_LAUNCHED("k_sparse_adam");
    return PVDB_OK;
}

}  // namespace

extern "C" {

int pvdb_fused_launch_defaults(int32_t* launch4) {
    PVDB_CHECK_ARG(launch4 != nullptr, "launch4 is NULL");
    launch4[0] = kFwdGrid;
    launch4[1] = kScatterGrid;
    launch4[2] = kScatterChunk;
    launch4[3] = kRayWaves;
    return PVDB_OK;
}

int pvdb_fused_timing(int32_t enable, double* total_ms, int32_t* calls, int32_t n) {
    for (auto& e : g_timer.open) {
        hipEventSynchronize(e.second.second);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e.second.first, e.second.second);
        g_timer.total_ms[e.first] += ms;
        g_timer.calls[e.first] += 1;
        hipEventDestroy(e.second.first);
        hipEventDestroy(e.second.second);
    }
    g_timer.open.clear();
    for (int i = 0; i < n && i < kStages; ++i) {
        if (total_ms) total_ms[i] = g_timer.total_ms[i];
        if (calls) calls[i] = g_timer.calls[i];
    }
    if (enable >= 0) {
        g_timer.enabled = enable != 0;
        for (int i = 0; i < kStages; ++i) {
            g_timer.total_ms[i] = 0;
            g_timer.calls[i] = 0;
        }
    }
    return PVDB_OK;
}

const char* pvdb_fused_stage_name(int32_t i) { return (i >= 0 && i < kStages) ? kStageNames[i] : ""; }

int64_t pvdb_fused_workspace_bytes(int64_t n_rays, int64_t capacity, int64_t nleaf) {
    if (n_rays < 0 || capacity < 0 || nleaf < 0) return -1;
    return (int64_t)layout_for(n_rays, capacity, nleaf).bytes;
}

int pvdb_fused_init(const pvdb_topology* topo, const pvdb_step_desc* d, void* stream) {
    if (!topo) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    PVDB_CHECK_ARG(d && d->workspace, "null descriptor or workspace");
    PVDB_CHECK_ARG(aligned(d->workspace, 256), "workspace must be 256-byte aligned");
    const TopoView t = view_of(topo);
    const Layout L = layout_for(d->n_rays, d->capacity, t.nleaf);
    PVDB_CHECK_ARG(d->workspace_bytes >= (int64_t)L.bytes, "workspace too small");
    uint8_t* ws = reinterpret_cast<uint8_t*>(d->workspace);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_fill_u8, dim3(pvdb_blocks(L.bytes, 256)), dim3(256), 0, s, ws,
                       (int64_t)L.bytes, (uint8_t)0);
    PVDB_LAUNCHED("k_fill_u8");
    return PVDB_OK;
}

int pvdb_fused_step(const pvdb_topology* topo, const pvdb_step_desc* d, void* stream) {
    if (!topo) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    PVDB_CHECK_ARG(d != nullptr, "null descriptor");
    PVDB_CHECK_ARG(d->n_rays > 0, "n_rays must be positive");
    PVDB_CHECK_ARG(d->capacity > 0 && d->capacity < kMaxCapacity,
                   "record capacity must be in [1, 2^31 - 2]");
    PVDB_CHECK_ARG(d->rays_o && d->rays_d && d->viewdirs && d->target, "null ray arrays");
    PVDB_CHECK_ARG(d->density && d->grad_density && d->color && d->grad_color, "null grid arrays");
    PVDB_CHECK_ARG(d->k0_dim == 0 || d->k0_dim == 3 || d->k0_dim == 12, "k0_dim must be 0, 3 or 12");
    const bool direct = d->k0_dim != 0;
    PVDB_CHECK_ARG(direct || (d->mlp && d->mlp_grad), "null MLP arrays");
    PVDB_CHECK_ARG(!direct || !d->color_bf16, "a model without the MLP takes an fp32 colour grid");
    PVDB_CHECK_ARG(d->do_adam >= 0 && d->do_adam <= 2, "do_adam must be 0, 1 or 2");
    PVDB_CHECK_ARG(!d->do_adam || (d->density_m && d->density_v && d->color_m && d->color_v &&
                                   (direct || (d->mlp_m && d->mlp_v))),
                   "Adam needs moment arrays");
    PVDB_CHECK_ARG(d->stepmode_density >= 0 && d->stepmode_density <= 2 &&
                       d->stepmode_color >= 0 && d->stepmode_color <= 1,
                   "stepmode_density must be 0, 1 or 2 and stepmode_color 0 or 1");
    PVDB_CHECK_ARG(d->stepmode_density != 2 || d->density_lr,
                   "density stepmode 2 needs density_lr");
    PVDB_CHECK_ARG(d->grad_stride == 0 ||
                       (d->grad_stride >= slot_channels(d) &&
                        (slot_channels(d) % 4 != 0 || d->grad_stride % 4 == 0)),
                   "grad_stride must be 0, or at least the colour channels (a multiple of 4 "
                   "for 12 channels)");
    PVDB_CHECK_ARG(d->stepdist > 0.f, "stepdist must be positive");
    PVDB_CHECK_ARG(d->launch[0] >= 0 && d->launch[0] <= 4096 && d->launch[1] >= 0 &&
                       d->launch[1] <= 65536,
                   "launch[0] must be in [0, 4096] and launch[1] in [0, 65536]");
    PVDB_CHECK_ARG(d->launch[2] == 0 || d->launch[2] == 32 || d->launch[2] == 64,
                   "launch[2]: scatter chunk must be 32 or 64 records");
    PVDB_CHECK_ARG(d->launch[3] == 0 || d->launch[3] == 1 || d->launch[3] == 2 ||
                       d->launch[3] == 4 || d->launch[3] == kMaxRayWaves,
                   "launch[3]: rays per workgroup must be 1, 2, 4 or 8");
    PVDB_CHECK_ARG(d->weight_main >= 0.f && d->weight_entropy_last >= 0.f &&
                       d->weight_rgbper >= 0.f && d->weight_main < INFINITY &&
                       d->weight_entropy_last < INFINITY && d->weight_rgbper < INFINITY,
                   "loss weights must be finite and >= 0");
    // the colour payload, its gradient and moments are moved as 16-byte vectors
    // (bf16 colour as 8-byte ones), the workspace's sub-arrays assume a
    // 256-byte aligned base
    PVDB_CHECK_ARG(aligned(d->grad_color, 16) && aligned(d->color_m, 16) &&
                       aligned(d->color_v, 16) && aligned(d->color, d->color_bf16 ? 8 : 16) &&
                       aligned(d->color_master, 16),
                   "colour arrays must be 16-byte aligned (8-byte for a bf16 payload)");
    PVDB_CHECK_ARG(aligned(d->workspace, 256), "workspace must be 256-byte aligned");
    // the composite writes ray_out in place, a 16-byte vector per ray with the MLP
    PVDB_CHECK_ARG(aligned(d->ray_out, direct ? 4 : 16),
                   "ray_out must be 16-byte aligned (4-byte without the MLP)");
    const TopoView t = view_of(topo);
    const Layout L = layout_for(d->n_rays, d->capacity, t.nleaf);
    PVDB_CHECK_ARG(d->workspace && d->workspace_bytes >= (int64_t)L.bytes, "workspace too small");
    uint8_t* ws = reinterpret_cast<uint8_t*>(d->workspace);
    hipStream_t s = (hipStream_t)stream;

    StepParams P;
    P.rays_o = d->rays_o;
    P.rays_d = d->rays_d;
    P.viewdirs = d->viewdirs;
    P.target = d->target;
    P.far_per_ray = d->far_per_ray;
    P.n_rays = d->n_rays;
    P.color_bf16 = d->color_bf16;
    for (int a = 0; a < 3; ++a) {
        P.mn[a] = d->xyz_min[a];
        P.mx[a] = d->xyz_max[a];
        P.len[a] = d->xyz_max[a] - d->xyz_min[a];
        P.rm1[a] = (float)(d->reso[a] - 1);
        // MaskGrid: scale = (shape-1)/(max-min), shift = -min*scale (grid.py:225-227)
        P.ms[a] = ((float)d->reso[a] - 1.f) / P.len[a];
        P.mb[a] = -d->xyz_min[a] * P.ms[a];
    }
    P.near_ = d->near_;
    P.far_ = 1e9f;
    P.stepdist = d->stepdist;
    P.act_shift = d->act_shift;
    P.interval = d->interval;
    P.thres = d->fast_color_thres;
    P.bg = d->bg;
    const int nc = direct ? d->k0_dim : 3;  // rendered channels
    P.loss_scale = d->weight_main / ((float)nc * (float)d->n_rays);
    P.ent_scale = d->weight_entropy_last / (float)d->n_rays;
    P.rgbper_scale = d->weight_rgbper / (float)d->n_rays;
    P.capacity = d->capacity;
    P.ray_waves = d->launch[3] > 0 ? d->launch[3] : kRayWaves;

    RayBuffers R;
    R.count = reinterpret_cast<int32_t*>(ws + L.count);
    R.offset = reinterpret_cast<int32_t*>(ws + L.offset);
    R.total = reinterpret_cast<int32_t*>(ws + L.total);
    R.t_last = reinterpret_cast<float*>(ws + L.t_last);
    // the composite writes the caller's per-ray output when there is one
    R.ray_out = d->ray_out && !direct ? reinterpret_cast<float4*>(d->ray_out)
                                      : reinterpret_cast<float4*>(ws + L.ray_out);
    R.ray_grad = reinterpret_cast<float4*>(ws + L.ray_grad);
    R.ray_ctot = reinterpret_cast<float4*>(ws + L.ray_ctot);
    R.ray_loss = reinterpret_cast<float*>(ws + L.ray_loss);
    R.ray_rgb = d->ray_out && direct ? d->ray_out : reinterpret_cast<float*>(ws + L.ray_rgb);
    R.loss = d->loss ? d->loss : reinterpret_cast<float*>(ws + L.loss);
    R.steps = reinterpret_cast<uint16_t*>(ws + L.steps);
    R.pe = direct ? nullptr : reinterpret_cast<uint16_t*>(ws + L.pe);
    R.setup = reinterpret_cast<float4*>(ws + L.setup);
    Records rec;
    rec.xyze = reinterpret_cast<float4*>(ws + L.xyze);
    rec.ab = reinterpret_cast<float2*>(ws + L.ab);
    rec.rgb = reinterpret_cast<float4*>(ws + L.rgb);
    rec.adj = reinterpret_cast<float4*>(ws + L.adj);
    rec.xrow = reinterpret_cast<uint4*>(ws + L.xrow);
    rec.relu = reinterpret_cast<uint32_t*>(ws + L.relu);
    rec.grad = reinterpret_cast<float4*>(ws + L.grad);
    uint8_t* touched = ws + L.touched;
    float* dw_part = reinterpret_cast<float*>(ws + L.dw_part);
    uint16_t* w0t = reinterpret_cast<uint16_t*>(ws + L.w0t);
    uint16_t* w1t = reinterpret_cast<uint16_t*>(ws + L.w1t);
    uint16_t* w2t = reinterpret_cast<uint16_t*>(ws + L.w2t);
    uint16_t* w2 = reinterpret_cast<uint16_t*>(ws + L.w2);
    uint16_t* w1 = reinterpret_cast<uint16_t*>(ws + L.w1);
    uint16_t* w0 = reinterpret_cast<uint16_t*>(ws + L.w0);
    MlpPacked W;
    W.w0t = w0t;
    W.w1t = w1t;
    W.w2t = w2t;
    W.w2 = w2;
    W.w1 = w1;
    W.w0 = w0;
    W.b0 = direct ? nullptr : d->mlp + kB0;
    W.b1 = direct ? nullptr : d->mlp + kB1;
    W.b2 = direct ? nullptr : d->mlp + kB2;

    int rc;
    if (d->do_adam == 2) {  // the optimizer alone, on the gradients already accumulated
        if (!direct) {
            StageScope sc(7, s);
            const MlpAdam adam{d->mlp,           d->mlp_m, d->mlp_v, d->adam_scalars,
                               d->step_size_mlp, d->eps,   d->beta0, d->beta1};
            if ((rc = launch_dw_reduce(dw_part, 0, d->mlp_grad, adam, s))) return rc;
        }
        StageScope sc(8, s);
        return launch_grid_adam(d, t, touched, s);
    }
    {
        // the counting march also packs the bf16 MLP operands and zeroes the
        // MLP gradient (accumulated by k_dw_reduce later in the step)
        StageScope sc(1, s);
        PackArgs pk{d->mlp, w0t, w1t, w2t, w2, w1, w0, d->mlp_grad};
        if (direct) pk.prm = nullptr;
        if ((rc = launch_ray_count(P, t, R, pk, s))) return rc;
    }
    {
        StageScope sc(2, s);
        if ((rc = launch_ray_march(P, t, d->density, R, rec, s))) return rc;
    }
    if (direct) {
        // no MLP: the composite gathers k0, and its gradient rows go straight
        // to the scatter
        {
            StageScope sc(4, s);
            if ((rc = launch_direct_composite(d->k0_dim, P, t, reinterpret_cast<const float*>(d->color),
                                              R, rec, s)))
                return rc;
        }
        {
            StageScope sc(6, s);
            const MlpAdam none{nullptr, nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f};
            const ReduceArgs red{nullptr, 0, nullptr, none, 0, R.ray_loss, R.loss, d->n_rays};
            if ((rc = launch_scatter(slot_channels(d), t, R, rec, d->capacity, grad_rows_of(d),
                                     touched, d->launch[1] > 0 ? d->launch[1] : kScatterGrid,
                                     d->launch[2] > 0 ? d->launch[2] : kScatterChunk, red, s)))
                return rc;
        }
        if (d->do_adam) {
            StageScope sc(8, s);
            if ((rc = launch_grid_adam(d, t, touched, s))) return rc;
        }
        return PVDB_OK;
    }
    // The sample-side kernels are persistent over the records; their grid is
    // fixed (the record count stays on the device).  The backward holds one
    // workgroup per CU (registers), so it runs one per CU and writes that many
    // dW partials.
    const int grid = d->launch[0] > 0 ? d->launch[0] : kFwdGrid;
    static int ncu = 0;
    if (ncu == 0) {
        int dev = 0;
        PVDB_HIP(hipGetDevice(&dev));
        PVDB_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    const int grid_bwd = ncu < kMaxMlpGrid ? (ncu < 1 ? 1 : ncu) : kMaxMlpGrid;
    {
        StageScope sc(3, s);
        if ((rc = launch_mlp(false, d->color_bf16 != 0, grid, P, t, d->color, W, R, rec, nullptr,
                             s)))
            return rc;
    }
    {
        StageScope sc(4, s);
        if ((rc = launch_ray_composite(P, R, rec, s))) return rc;
    }
    {
        StageScope sc(5, s);
        if ((rc = launch_mlp(true, d->color_bf16 != 0, grid_bwd, P, t, d->color, W, R, rec,
                             dw_part, s)))
            return rc;
    }
    {
        // the scatter's launch also carries the dW reduction (its first
        // workgroups), and the MLP Adam rides on the reduction (each parameter
        // once its gradient is final)
        StageScope sc(6, s);
        MlpAdam adam{nullptr, nullptr, nullptr, nullptr, 0.f, 0.f, 0.f, 0.f};
        if (d->do_adam)
            adam = MlpAdam{d->mlp,           d->mlp_m, d->mlp_v, d->adam_scalars,
                           d->step_size_mlp, d->eps,   d->beta0, d->beta1};
        const ReduceArgs red{dw_part, grid_bwd, d->mlp_grad, adam, kReduceBlocks,
                             R.ray_loss, R.loss, d->n_rays};
        if ((rc = launch_scatter(12, t, R, rec, d->capacity, grad_rows_of(d), touched,
                                 d->launch[1] > 0 ? d->launch[1] : kScatterGrid,
                                 d->launch[2] > 0 ? d->launch[2] : kScatterChunk, red, s)))
            return rc;
    }
    if (d->do_adam) {
        StageScope sc(8, s);
        if ((rc = launch_grid_adam(d, t, touched, s))) return rc;
    }
    return PVDB_OK;
}

int pvdb_fused_records(const pvdb_topology* topo, const pvdb_step_desc* d, void** xyze,
                       void** ab, void** rgb, void** adj, int32_t** count, int32_t** offset,
                       int32_t** total) {
    if (!topo) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    PVDB_CHECK_ARG(d && d->workspace, "null descriptor or workspace");
    const Layout L = layout_for(d->n_rays, d->capacity, view_of(topo).nleaf);
    uint8_t* ws = reinterpret_cast<uint8_t*>(d->workspace);
    if (xyze) *xyze = ws + L.xyze;
    if (ab) *ab = ws + L.ab;
    if (rgb) *rgb = ws + L.rgb;
    if (adj) *adj = ws + L.adj;
    if (count) *count = reinterpret_cast<int32_t*>(ws + L.count);
    if (offset) *offset = reinterpret_cast<int32_t*>(ws + L.offset);
    if (total) *total = reinterpret_cast<int32_t*>(ws + L.total);
    return PVDB_OK;
}

}  // extern "C"
