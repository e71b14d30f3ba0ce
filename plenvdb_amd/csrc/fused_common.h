// Shared definitions of the fused (batched) training step.
//
// Pipeline of one step (all on one stream, no host synchronisation):
//   k_ray_count    wave per ray: slab test + march + occupancy -> samples per ray;
//                  its extra workgroups pack the bf16 MLP operands and zero the
//                  MLP gradient
//   k_scan         exclusive scan of the counts -> record offsets, total
//   k_ray_march    wave per ray: march again, density gather, raw2alpha, alpha cull,
//                  transmittance product-scan with early stop, weights, weight cull;
//                  one record per occupancy-kept sample (contiguous per ray)
//   k_mlp_fwd      workgroup per 32-record group: colour gather + MFMA MLP -> rgb;
//                  the features (bf16 hi, lo) and the ReLU masks kept per record
//   k_ray_composite wave per ray: prefix of w*rgb, ray colour, the loss terms and
//                  their gradient per ray; a second pass forms each sample's
//                  dL/dlogit and, through the alpha/raw2alpha adjoints, dL/dsigma
//   k_loss_reduce  the per-ray loss terms summed in a fixed order
//   k_mlp_bwd      workgroup per CU over 32-record groups: recompute from the
//                  kept features under the kept masks, MFMA MLP backward, dW
//                  partials; per-record feature and sigma gradients
//   k_scatter_agg  workgroups over 32- or 64-record chunks: corner contributions
//                  summed per slot in an LDS hash, device atomics per distinct slot
//   k_dw_reduce    the dW partials summed in a fixed order (+ the MLP Adam)
//   k_sparse_adam  VDB Adam over the touched leaves (every leaf in stepmodes 0
//                  and 2), gradients re-zeroed
// A model without the MLP (k0_dim 3 or 12: rgb = sigmoid(k0)) runs k_ray_count,
// k_scan, k_ray_march, then k_direct_composite (the k0 gather, composite, loss
// and per-sample gradients in one wave-per-ray kernel), k_loss_reduce,
// k_scatter_agg over its k0 channels and k_sparse_adam.
#pragma once

#include "mfma_util.h"
#include "pvdb_common.h"

namespace pvdb {

// Float index of dL/dsigma in a record's gradient row: after the colour
// gradients (12 features, or the 3 channels of a direct-colour grid).
template <int NC>
constexpr int kSigmaAt = NC == 12 ? 12 : 3;

// Per-record flags (bits 30..31 of the packed ray field).
constexpr uint32_t kAliveA = 1u << 30;  // in the alpha-kept set, before the early stop
constexpr uint32_t kAliveW = 1u << 31;  // also weight-kept: gets colour, MLP and composite
constexpr uint32_t kRayMask = (1u << 30) - 1;

struct StepParams {
    const float* rays_o;
    const float* rays_d;
    const float* viewdirs;
    const float* target;       // [N][3]
    const float* far_per_ray;  // may be null
    int32_t n_rays;
    int32_t color_bf16;        // colour payload stored as bf16
    float mn[3], mx[3];        // scene bbox
    float len[3], rm1[3];      // wld2idx: (p - mn) / len * rm1
    float ms[3], mb[3];        // occupancy lookup: round(p * ms + mb)
    float near_, far_, stepdist;
    float act_shift, interval, thres, bg;
    float loss_scale;          // weight_main / (3 N): MSE mean over N rays x 3 channels
    float ent_scale;           // weight_entropy_last / N: entropy of T_last, mean over rays
    float rgbper_scale;        // weight_rgbper / N: sum of w * |rgb_i - target|^2 over samples
    int64_t capacity;          // record capacity
    int32_t ray_waves;         // rays (wavefronts) per workgroup of the wave-per-ray kernels
};

// The grid gradients the scatter adds into and the sparse Adam consumes:
// voxel s's density gradient at density[s * ds], its colour gradient row at
// color[s * cs .. + NC).  Separate arrays have ds = 1 and cs = NC; rows of one
// buffer (ds = cs = the row length, the density inside the row) put a voxel's
// gradients in one cache line, so the scatter's atomics for a voxel are one
// memory request instead of two or three.
struct GradRows {
    float* density;
    float* color;
    int32_t ds, cs;
    bool whole;  // rows of 16 (4) floats with the density right after the 12 (3)
                 // colour values: a row is moved as whole 16-byte vectors
};

// Rays per workgroup of the wave-per-ray kernels (k_ray_count, k_ray_march and
// the composites) at most; the launch chooses 1, 2, 4 or 8.
constexpr int kMaxRayWaves = 8;

// Records: structure of arrays, index = record id.
struct Records {
    float4* xyze;      // index-space x, y, z; e = exp(sigma + shift)
    float2* ab;        // alpha, ray|flags (as bits); T and w are recomputed from the alphas
    float4* rgb;       // MLP colour (rgb), written by the forward MLP pass
    float4* adj;       // per-sample adjoints: dL/dlogit (3), dL/dsigma
    uint4* xrow;       // [5 per record] the backward's MLP input row, X columns 0..39 as bf16
                       // hi parts: the 12 gathered colour features, the 27 view-direction
                       // encoding inputs and the bias input (a zero row unless weight-kept)
    uint32_t* relu;    // [8 per record] forward ReLU masks: word 4l + j = layer l's
                       // hidden features 32j + 16m + 4g + i (m, i < 2, 4; g < 4) at
                       // bit 8g + 4m + i
    float4* grad;      // [4 per record] dL/dfeature (12) and dL/dsigma; in the direct-colour
                       // configurations rgb (NC) between the composite's two passes, then
                       // dL/dk0 (NC) and dL/dsigma at float kSigmaAt<NC>
};

struct RayBuffers {
    int32_t* count;    // occupancy-kept samples per ray
    int32_t* offset;   // exclusive prefix of count
    int32_t* total;    // [1] total records
    float* t_last;     // [N]
    float4* ray_out;   // [N] rendered rgb + T_last
    float4* ray_grad;  // [N] dL/drgb (3) + dL/dT_last
    float4* ray_ctot;  // [N] sum of w*rgb (3)
    float* ray_loss;   // [N] the ray's share of the loss (all weighted terms)
    float* ray_rgb;    // [N][NC + 1] direct-colour configurations: rendered rgb (NC) + T_last
    float* loss;       // [1]
    uint16_t* steps;   // [N][kStepCache] the counting march's kept step indices, in step order
    float4* setup;     // [N][2] the counting march's ray set-up for the forward march:
                       // (o + d t_min, .), (d / |d|, .), the step count in the .w pair
    uint16_t* pe;      // [N][3][32] with the MLP: the ray's MLP input columns 12..39 (view
                       // direction, its sin/cos encoding, the bias column) as bf16 hi,
                       // mid and lo rows, written by the counting march; null without
};

// Kept step indices the counting march stores per ray (16-bit: rays of fewer
// than 65536 steps).  A ray with more kept steps than this, or more steps, is
// marched again by the forward march.
constexpr int kStepCache = 256;

// Packed bf16 MLP weights for MFMA (row-major, zero padded):
//   w0t [128][64]  W0^T (in 0..38 used)       w1t [128][128] W1^T
//   w2t [16][128]  W2^T (rows 0..2 used)      w2  [128][32]  W2 (cols 0..2 used)
//   w1  [128][128] W1                          w0  [16][128]  W0 rows 0..15 (0..11 used)
// and fp32 biases b0[128], b1[128], b2[3].
struct MlpPacked {
    const uint16_t* w0t;
    const uint16_t* w1t;
    const uint16_t* w2t;
    const uint16_t* w2;
    const uint16_t* w1;
    const uint16_t* w0;
    const float* b0;
    const float* b1;
    const float* b2;
};

// fp32 gradient layout of the MLP (the layout of the fp32 master parameters):
//   W0 [39][128] | b0 [128] | W1 [128][128] | b1 [128] | W2 [128][3] | b2 [3]
constexpr int kW0 = 0;
constexpr int kB0 = kW0 + 39 * 128;
constexpr int kW1 = kB0 + 128;
constexpr int kB1 = kW1 + 128 * 128;
constexpr int kW2 = kB1 + 128;
constexpr int kB2 = kW2 + 128 * 3;
constexpr int kMlpParams = kB2 + 3;  // 22019
constexpr int kPartStride = (kMlpParams + 3) & ~3;  // dW partial rows padded to 16 bytes

// Adam of the MLP parameters (stepmode 0, adam_upd_kernel.cu:8-58), applied by
// the dW reduction to each parameter once its gradient is final.  p == nullptr
// skips it; the step size comes from the device (dev_sz[2]) when given.
// The fp32 master parameters (kMlpParams layout) packed into the bf16 MFMA
// operands of the step, and the MLP gradient zeroed (k_dw_reduce accumulates
// it later in the step): item t of kPackItems.
struct PackArgs {
    const float* prm;
    uint16_t *w0t, *w1t, *w2t, *w2, *w1, *w0;
    float* grad;
};
constexpr int kPackItems = kMlpParams > 128 * 128 ? kMlpParams : 128 * 128;

__device__ __forceinline__ void pack_mlp_item(int t, const PackArgs& A) {
    if (t < kMlpParams) A.grad[t] = 0.f;
    if (t < 128 * 128) {  // w1t / w1: 128 x 128
        const int a = t >> 7, b = t & 127;
        A.w1t[a * 128 + b] = f2bf_bits(A.prm[kW1 + b * 128 + a]);  // W1^T[out a][in b]
        A.w1[a * 128 + b] = f2bf_bits(A.prm[kW1 + a * 128 + b]);   // W1[in a][out b]
    }
    if (t < 128 * 64) {  // w0t [128 out][64 in]
        const int o = t >> 6, i = t & 63;
        A.w0t[o * 64 + i] = i < 39 ? f2bf_bits(A.prm[kW0 + i * 128 + o]) : (uint16_t)0;
    }
    if (t < 16 * 128) {  // w2t [16 out][128 in], w0 [16 in][128 out]
        const int o = t >> 7, i = t & 127;
        A.w2t[o * 128 + i] = o < 3 ? f2bf_bits(A.prm[kW2 + i * 3 + o]) : (uint16_t)0;
        A.w0[o * 128 + i] = o < 12 ? f2bf_bits(A.prm[kW0 + o * 128 + i]) : (uint16_t)0;
    }
    if (t < 128 * 32) {  // w2 [128 in][32 out]
        const int i = t >> 5, o = t & 31;
        A.w2[i * 32 + o] = o < 3 ? f2bf_bits(A.prm[kW2 + i * 3 + o]) : (uint16_t)0;
    }
}

struct MlpAdam {
    float* p;
    float* m;
    float* v;
    const float* dev_sz;
    float sz, eps, b0, b1;
};

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Two floats -> two bf16 (round to nearest even) in one dword, element 0 in
// the low half (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t pk_bf(float a, float b) {
    typedef __bf16 bf2_t __attribute__((ext_vector_type(2)));
    typedef float f2_t __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f2_t){a, b}, bf2_t));
}
__device__ __forceinline__ float bf_lo(uint32_t p) { return __uint_as_float(p << 16); }
__device__ __forceinline__ float bf_hi(uint32_t p) { return __uint_as_float(p & 0xffff0000u); }

// x = hi + lo with hi = bf16(x) and lo = bf16(x - hi): the pair carries about
// 16 significant bits, so products of a bf16 weight with (hi, lo) keep the
// activation's precision.  Two values per call, packed as above.
__device__ __forceinline__ void split2(float a, float b, uint32_t& h, uint32_t& l) {
    h = pk_bf(a, b);
    l = pk_bf(a - bf_lo(h), b - bf_hi(h));
}
// x = hi + mid + lo, each bf16: about 24 significant bits, as many as fp32
// holds (x - hi and mid's remainder are exact in fp32).  The forward pass feeds
// its products with triples: with pairs, about one pre-activation in a million
// rounds to the other side of zero at the benchmarked sizes, and each such
// ReLU flip moves its sample's feature gradient by a whole term.
__device__ __forceinline__ void split3(float a, float b, uint32_t& h, uint32_t& m, uint32_t& l) {
    h = pk_bf(a, b);
    const float ra = a - bf_lo(h), rb = b - bf_hi(h);
    m = pk_bf(ra, rb);
    l = pk_bf(ra - bf_lo(m), rb - bf_hi(m));
}

// Wavefront scans on the DPP datapath, for a fully active wavefront: four row
// shifts (1, 2, 4, 8 lanes within each 16-lane row), then lane 15's value
// broadcast into rows 1 and 3 and lane 31's into rows 2 and 3.  A lane whose
// source lies outside its row reads the identity.  Each step is one VALU
// operation, where a shuffle is an LDS permute round trip.
template <int kCtrl, int kRows = 0xf>
__device__ __forceinline__ float dpp_mov(float x, float ident) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
        __builtin_bit_cast(int, ident), __builtin_bit_cast(int, x), kCtrl, kRows, 0xf, false));
}
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114,
              kDppRowShr8 = 0x118, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143,
              kDppWaveShr1 = 0x138;

// inclusive product over lanes 0..l
__device__ __forceinline__ float wave_incl_prod(float x) {
    x *= dpp_mov<kDppRowShr1>(x, 1.f);
    x *= dpp_mov<kDppRowShr2>(x, 1.f);
    x *= dpp_mov<kDppRowShr4>(x, 1.f);
    x *= dpp_mov<kDppRowShr8>(x, 1.f);
    x *= dpp_mov<kDppRowBcast15, 0xa>(x, 1.f);
    x *= dpp_mov<kDppRowBcast31, 0xc>(x, 1.f);
    return x;
}

// inclusive sums over lanes 0..l of N values at once (their steps interleave)
template <int N>
__device__ __forceinline__ void wave_incl_sum(float (&x)[N]) {
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] += dpp_mov<kDppRowShr1>(x[c], 0.f);
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] += dpp_mov<kDppRowShr2>(x[c], 0.f);
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] += dpp_mov<kDppRowShr4>(x[c], 0.f);
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] += dpp_mov<kDppRowShr8>(x[c], 0.f);
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] += dpp_mov<kDppRowBcast15, 0xa>(x[c], 0.f);
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] += dpp_mov<kDppRowBcast31, 0xc>(x[c], 0.f);
}

// lane l - 1's value (ident on lane 0)
__device__ __forceinline__ float wave_shr1(float x, float ident) {
    return dpp_mov<kDppWaveShr1>(x, ident);
}

// lane `src`'s value (src uniform over the wavefront)
__device__ __forceinline__ float wave_read(float x, int src) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), src));
}

// The dW reduction a launch carries in its first `blocks` workgroups (0: none),
// and the loss sum it carries in the workgroup after them (when ray_loss is set).
struct ReduceArgs {
    const float* part;
    int nparts;
    float* dw;
    MlpAdam adam;
    int blocks;
    const float* ray_loss;  // [n_rays] per-ray loss terms, or null
    float* loss;            // [1] their sum
    int n_rays;
};

// loss[0] = sum of v[0..n), by one wavefront in a fixed order (lane l adds
// v[4l..4l+3], v[4(l+64)..], ... in turn, then a butterfly over the lanes), so
// the sum does not depend on the launch that carries it.  v is 16-byte aligned.
__device__ __forceinline__ void loss_sum_wave(const float* __restrict__ v, int n,
                                              float* __restrict__ loss) {
    const int lane = lane_id();
    const int n4 = n >> 2;
    const float4* v4 = reinterpret_cast<const float4*>(v);
    float s = 0.f;
    int i = lane;
    for (; i + 192 < n4; i += 256) {  // four 16-byte loads in flight per lane
        const float4 a = v4[i], b = v4[i + 64], c = v4[i + 128], d = v4[i + 192];
        s += ((a.x + a.y) + a.z) + a.w;
        s += ((b.x + b.y) + b.z) + b.w;
        s += ((c.x + c.y) + c.z) + c.w;
        s += ((d.x + d.y) + d.z) + d.w;
    }
    for (; i < n4; i += 64) {
        const float4 a = v4[i];
        s += ((a.x + a.y) + a.z) + a.w;
    }
    if (lane < (n & 3)) s += v[4 * n4 + lane];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (lane == 0) loss[0] = s;
}

// Sum of the per-workgroup dW partials, added into dw without atomics: workgroup
// `block` owns kReduceVec four-parameter vectors and splits the partials among
// kReduceGrp lane groups (16-byte loads, eight in flight per lane); the groups'
// sums are combined through LDS (`red`, kReduceGrp * kReduceVec float4) in a
// fixed order, so the result is reproducible.  With Adam on (A.p), the same
// thread then updates the parameters (adam_upd_kernel.cu:8-58, stepmode 0, on
// the final gradient).  Needs 256 threads.
constexpr int kReduceVec = 16, kReduceGrp = 16;
static_assert(kReduceVec * kReduceGrp == 256, "one thread per (vector, group)");
constexpr int kReduceBlocks = (kPartStride / 4 + kReduceVec - 1) / kReduceVec;
__device__ __forceinline__ void dw_reduce_block(int block, const float* __restrict__ part,
                                                int nparts, float* __restrict__ dw,
                                                const MlpAdam& A, float4* red) {
    const int lv = threadIdx.x % kReduceVec, grp = threadIdx.x / kReduceVec;
    const int v = block * kReduceVec + lv;  // float4 index; rows are padded to 16 B
    const int q = (nparts + kReduceGrp - 1) / kReduceGrp, b0 = grp * q, b1 = min(nparts, b0 + q);
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    if (4 * v < kMlpParams) {
        for (int b = b0; b < b1; b += 8) {
            float4 x[8];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                x[k] = b + k < b1
                           ? reinterpret_cast<const float4*>(part + (int64_t)(b + k) * kPartStride)[v]
                           : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                s.x += x[k].x;
                s.y += x[k].y;
                s.z += x[k].z;
                s.w += x[k].w;
            }
        }
    }
    red[grp * kReduceVec + lv] = s;
    __syncthreads();
    if (grp == 0 && 4 * v < kMlpParams) {
        float4 t = red[lv];
#pragma unroll
        for (int g2 = 1; g2 < kReduceGrp; ++g2) {
            const float4 u = red[g2 * kReduceVec + lv];
            t.x += u.x;
            t.y += u.y;
            t.z += u.z;
            t.w += u.w;
        }
        const float tv[4] = {t.x, t.y, t.z, t.w};
        const float sz = A.dev_sz ? A.dev_sz[2] : A.sz;
#pragma unroll
        for (int k = 0; k < 4; ++k) {  // the padding element of the last vector is skipped
            const int qq = 4 * v + k;
            if (qq >= kMlpParams) continue;
            const float gg = dw[qq] + tv[k];
            dw[qq] = gg;
            if (A.p) {
                const float nm = A.b0 * A.m[qq] + (1.f - A.b0) * gg;
                const float nv = A.b1 * A.v[qq] + ((1.f - A.b1) * gg) * gg;
                A.m[qq] = nm;
                A.v[qq] = nv;
                A.p[qq] = A.p[qq] - (sz * nm) / (sqrtf(nv) + A.eps);
            }
        }
    }
}

}  // namespace pvdb
