// Merged-VDB renderer: MGRenderer::render_an_image (plenvdb/lib/vdb/plenvdb.h:
// 1029-1035 -> renderer.cu:370-424) in four kernels and no host round trip.
//
// The reference renders in two full marches per pixel (first_look_merged
// counts the kept samples and rewrites t_min / t_max, renderer.cu:222-268;
// ray_marching_merged re-marches from the rewritten t_min and gathers them,
// :312-367), a host read-back of the sample total, four SGEMMs over the
// materialised sample list (cuda_rgbnet, :83-109) and an atomic composite
// (final_render, :112-119).  Here:
//
//   k_render_count  a wavefront per 8x8 pixel tile, a lane per pixel: the
//                   first march with trigetDensity's summation order; stores
//                   the kept count and the rewritten (t_min, t_max) per pixel.
//                   Light on registers, so it runs at full occupancy.
//   k_render_fill   the second march from the rewritten t_min with
//                   trigetDensity2's order: the tile reserves its records with
//                   one atomic (counts scanned across the wave) and writes one
//                   32-byte record per kept sample (12 colour features in fp16,
//                   the weight, the pixel), at most the first march's count of
//                   them, contiguous per pixel.  So the sample set is the
//                   reference's, including the second march's restart
//                   (renderer.cu:263).
//   k_render_mlp    persistent waves over 64-record chunks: the RGB MLP on
//                   v_mfma_f32_16x16x32_f16 (weights in LDS, shared by the
//                   workgroup's 8 waves) and each record's contribution
//                   w / (1 + exp(-raw)) (final_render).
//   k_render_composite  per pixel: T_last * bg plus its records' contributions,
//                   eight lanes per pixel reading its run of records together
//                   (a fixed order, so frames are reproducible).
//
// Data layout.  The merged index grid (vdb_compression.py:19-59: index 1 +
// rank of an active voxel, 0 elsewhere) is a topology whose value masks are
// the non-zero indices, plus a slot-major payload holding, per slot, density
// and the 12 colour channels of the voxel's index (index 0's values for
// inactive voxels; slot `nslots` holds index 0's values for corners whose leaf
// does not exist).  A slot is 32 bytes (fp16: exact for the fp16 data of
// mergeddata.npz) or 64 bytes (fp32); the densities follow once more as a
// compact array, which the marches' corner loads read.
//
// MLP.  Layer outputs are computed transposed (features x samples) and stay in
// registers: the D fragment of one layer (lane (g,n): features 16m+4g+i of
// sample n) is the B fragment of the next once the next layer's weights are
// stored with their K index permuted accordingly (see kperm).  Weights are
// fp16 in LDS, read as 16-byte fragments.  The first layer's input row is
// [feature 0..11 | 0 x4 | PE 0..26 | 1 | 0 ...] (64 halves), so its bias is a
// weight column and a record's 16-byte halves and the pixel's PE row land on
// 16-byte boundaries.
#include <cmath>
#include <mutex>
#include <utility>
#include <vector>

#include "gather.h"
#include "mfma_util.h"
#include "plenvdb_hip.h"
#include "pvdb_common.h"

namespace pvdb {
TopoView view_of(const pvdb_topology* t);
}

using namespace pvdb;

namespace {

constexpr int kHid = 128;
constexpr int kWaves = 8;   // MLP waves per workgroup: they share the LDS weights
constexpr int XS = 72;      // halves per staged input row (64 + 8)
constexpr int kPeIn = 16;   // first-layer input of PE component 0
constexpr int kW0Frag = 8 * 2 * 64 * 8;  // [m-tile 8][k-step 2][lane][8]
constexpr int kW1Frag = 8 * 4 * 64 * 8;  // [m-tile 8][k-step 4][lane][8]
constexpr int kW2Frag = 4 * 64 * 8;      // [k-step 4][lane][8]
constexpr int kPackedHalves = kW0Frag + kW1Frag + kW2Frag;
constexpr int kBiasFloats = 128 + 4;  // b1, b2 (padded)
constexpr int kPackedBytes = kPackedHalves * 2 + kBiasFloats * 4;
constexpr int kMarchThreads = 256;    // 4 waves per march workgroup
constexpr int64_t kMaxRecords = 0x7fffffffll;  // record indices are int32 in the pixel table

// Position j (0..7) of k-step s in lane group g holds feature kperm(s, g, j):
// the features lane group g of the previous layer's D fragments carries.
__host__ __device__ __forceinline__ int kperm(int s, int g, int j) {
    return 32 * s + 4 * g + (j & 3) + ((j >> 2) << 4);
}

__global__ void k_render_pack(const float* __restrict__ w0, const float* __restrict__ b0,
                              const float* __restrict__ w1, const float* __restrict__ b1,
                              const float* __restrict__ w2, const float* __restrict__ b2,
                              int dcol, int dpe, int dout, _Float16* __restrict__ outh,
                              float* __restrict__ outb) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < kW0Frag) {
        const int j = idx & 7, lane = (idx >> 3) & 63, ms = idx >> 9;
        const int s = ms & 1, m = ms >> 1, g = lane >> 4, r = lane & 15;
        const int o = 16 * m + r, in = 32 * s + 8 * g + j;
        float v = 0.f;
        if (in < dcol)
            v = w0[in * kHid + o];
        else if (in >= kPeIn && in < kPeIn + dpe)
            v = w0[(dcol + in - kPeIn) * kHid + o];
        else if (in == kPeIn + dpe)
            v = b0[o];
        outh[idx] = (_Float16)v;
    } else if (idx < kW0Frag + kW1Frag) {
        const int q = idx - kW0Frag;
        const int j = q & 7, lane = (q >> 3) & 63, ms = q >> 9;
        const int s = ms & 3, m = ms >> 2, g = lane >> 4, r = lane & 15;
        outh[idx] = (_Float16)w1[kperm(s, g, j) * kHid + 16 * m + r];
    } else if (idx < kPackedHalves) {
        const int q = idx - kW0Frag - kW1Frag;
        const int j = q & 7, lane = (q >> 3) & 63, s = q >> 9;
        const int g = lane >> 4, r = lane & 15;
        outh[idx] = (_Float16)(r < dout ? w2[kperm(s, g, j) * dout + r] : 0.f);
    } else if (idx < kPackedHalves + kBiasFloats) {
        const int q = idx - kPackedHalves;
        outb[q] = q < 128 ? b1[q] : (q - 128 < dout ? b2[q - 128] : 0.f);
    }
}

// Slot payload: density and 12 colour channels of the voxel's merged index.
template <typename PT>
__global__ void k_render_payload(const int32_t* __restrict__ slot_idx, int64_t nslots,
                                 const float* __restrict__ den, const float* __restrict__ col,
                                 int64_t nidx, PT* __restrict__ pay, PT* __restrict__ dens) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nslots) return;
    int64_t id = s < nslots ? slot_idx[s] : 0;  // slot nslots: index 0 (missing leaves)
    if (id < 0 || id >= nidx) id = 0;           // the host checks the range
    PT* p = pay + s * 16;
    p[0] = (PT)den[id];
    dens[s] = (PT)den[id];
#pragma unroll
    for (int c = 0; c < 12; ++c) p[1 + c] = (PT)col[id * 12 + c];
    p[13] = p[14] = p[15] = (PT)0.f;
}

// Workspace of one frame (pvdb_render_workspace_bytes), 256-byte aligned parts.
struct FrameBuffers {
    unsigned long long* total;       // [1] records reserved (may exceed the capacity)
    int32_t* pix_start;              // [HW] first record of the pixel (saturated)
    int32_t* pix_count;              // [HW] pass-1 kept count (records reserved)
    float2* tspan;                   // [HW] pass 2's t range (rewritten t_min, t_max)
    int32_t* pix_written;            // [HW] records pass 2 filled (<= count)
    float* t_last;                   // [HW] transmittance after pass 2
    uint4* pe;                       // [HW][4] PE row: 27 halves, 1, 0 x4
    uint4* rec;                      // [capacity][2] records
    float4* contrib;                 // [capacity] w * sigmoid(raw), per channel
    int64_t capacity;
};

struct RenderArgs {
    TopoView topo;
    const void* payload;
    const void* den;  // [nslots + 1] densities alone (the first element of each slot)
    int64_t nslots;
    const uint4* packed;
    float* out;
    unsigned long long* counters;  // [3]: pass-1 density gathers, written samples, records
                                   // reserved (may be null)
    FrameBuffers fb;
    float c2w[12];
    float fx, cx, fy, cy;
    float mn[3], ext[3], ws[3], inv_ws[3];
    int32_t H, W, inverse_y, ntx, ntiles;
    float near_, stepdist, act_shift, interval, thres, bg, t_stop;
};

// Two floats -> two fp16 (round to nearest even) in one dword, a in the low
// half: one v_cvt_pk_f16_f32.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack2(float a, float b) {
    typedef float f2_t __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f2_t){a, b}, half2_t));
}
// relu of two floats as packed fp16: converted first, then one v_pk_max_f16
// (rounding to nearest is monotonic and keeps 0, so this equals converting
// the relu'd values, up to the sign of a zero)
__device__ __forceinline__ uint32_t relu_pack2(float a, float b) {
    const half2_t z = {(_Float16)0.f, (_Float16)0.f};
    const half2_t h = __builtin_bit_cast(half2_t, pack2(a, b));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(h, z));
}

__device__ __forceinline__ float h2f(uint16_t h) {
    return (float)__builtin_bit_cast(_Float16, h);
}

// Density of a slot from the compact density array (2 or 4 bytes a slot, so
// the marches' corner loads share cache lines; 2 MB at 1M slots in fp16).
template <bool F32PAY>
__device__ __forceinline__ float load_den(const void* den, int64_t slot) {
    if (F32PAY) return reinterpret_cast<const float*>(den)[slot];
    return h2f(reinterpret_cast<const uint16_t*>(den)[slot]);
}

template <bool F32PAY>
__device__ __forceinline__ void load_col(const void* pay, int64_t slot, float* c12) {
    if (F32PAY) {
        const float4* p = reinterpret_cast<const float4*>(pay) + slot * 4;
        const float4 a = p[0], b = p[1], c = p[2], d = p[3];
        c12[0] = a.y; c12[1] = a.z; c12[2] = a.w;
        c12[3] = b.x; c12[4] = b.y; c12[5] = b.z; c12[6] = b.w;
        c12[7] = c.x; c12[8] = c.y; c12[9] = c.z; c12[10] = c.w;
        c12[11] = d.x;
    } else {
        const uint4* p = reinterpret_cast<const uint4*>(pay) + slot * 2;
        const uint4 a = p[0], b = p[1];
        const uint32_t w[7] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z};
#pragma unroll
        for (int c = 0; c < 12; ++c) {
            const int h = c + 1;  // half index: 0 is the density
            c12[c] = h2f((uint16_t)(w[h >> 1] >> (16 * (h & 1))));
        }
    }
}

__device__ __forceinline__ f16x8 as_h8(const uint4& v) { return __builtin_bit_cast(f16x8, v); }

// RGB logits of the 64 samples staged in this wave's rows of sX (row =
// lane); lanes 0..15 write them to sL (the rows of empty tiles hold zeros, so
// their logits are the bias and are not read).  All four 16-sample tiles go
// through each layer together, so every weight fragment read from LDS feeds
// four MFMAs and the tiles' activations stay in registers.
__device__ __forceinline__ void wave_mlp(const uint4* __restrict__ sW, const float* __restrict__ sB,
                                         const _Float16* __restrict__ sX, float* __restrict__ sL,
                                         int lane) {
    const int g = lane >> 4, r16 = lane & 15;
    const uint4* W0 = sW;
    const uint4* W1 = sW + kW0Frag / 8;
    const uint4* W2 = W1 + kW1Frag / 8;
    uint32_t h0[4][4][4];  // [tile][k-step][dword]: layer-1 B fragments (fp16 pairs)
    {
        f16x8 xb[4][2];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int s = 0; s < 2; ++s)
                xb[u][s] = as_h8(*reinterpret_cast<const uint4*>(sX + (16 * u + r16) * XS +
                                                                 32 * s + 8 * g));
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const f16x8 a0 = as_h8(W0[(m * 2 + 0) * 64 + lane]);
            const f16x8 a1 = as_h8(W0[(m * 2 + 1) * 64 + lane]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma16x16x32h(a0, xb[u][0], acc);
                acc = mfma16x16x32h(a1, xb[u][1], acc);
                const int s = m >> 1, d = (m & 1) * 2;
                h0[u][s][d] = relu_pack2(acc[0], acc[1]);
                h0[u][s][d + 1] = relu_pack2(acc[2], acc[3]);
            }
        }
    }
    // layer 1, and layer 2 as its outputs appear: layer 2's k-step s takes
    // layer 1's output tiles 2s and 2s + 1, so only one k-step of H1 is live
    const float b2x = sB[128], b2y = sB[129], b2z = sB[130];
    f32x4 acc2[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc2[u] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) {
        uint32_t hk[4][4];  // [tile][dword]: layer 2's B fragments of k-step s2
#pragma unroll
        for (int mm = 0; mm < 2; ++mm) {
            const int m = 2 * s2 + mm;
            // the accumulators start at the bias (the MFMA adds it)
            const float4 bias = *reinterpret_cast<const float4*>(sB + 16 * m + 4 * g);
            f32x4 acc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[u] = f32x4{bias.x, bias.y, bias.z, bias.w};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const f16x8 a = as_h8(W1[(m * 4 + s) * 64 + lane]);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint4 bv = make_uint4(h0[u][s][0], h0[u][s][1], h0[u][s][2], h0[u][s][3]);
                    acc[u] = mfma16x16x32h(a, as_h8(bv), acc[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                hk[u][2 * mm] = relu_pack2(acc[u][0], acc[u][1]);
                hk[u][2 * mm + 1] = relu_pack2(acc[u][2], acc[u][3]);
            }
        }
        const f16x8 a2 = as_h8(W2[s2 * 64 + lane]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            acc2[u] = mfma16x16x32h(a2, as_h8(make_uint4(hk[u][0], hk[u][1], hk[u][2], hk[u][3])),
                                    acc2[u]);
    }
    if (g == 0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float* dst = sL + (16 * u + r16) * 4;
            dst[0] = acc2[u][0] + b2x;
            dst[1] = acc2[u][1] + b2y;
            dst[2] = acc2[u][2] + b2z;
        }
    }
}

// Leaf index of (i,j,k) through a one-entry cache of its 8^3 block.
struct BlockCache {
    int a = INT32_MIN, b = 0, c = 0;
    int32_t leaf = -1;
    __device__ __forceinline__ int32_t get(const TopoView& t, int i, int j, int k) {
        const int x = i >> 3, y = j >> 3, z = k >> 3;
        if (x != a || y != b || z != c) {
            a = x;
            b = y;
            c = z;
            leaf = leaf_of(t, i, j, k);
        }
        return leaf;
    }
};

// A pixel's ray in the renderer's normalised box coordinates (get_rays,
// renderer.cu:122-167; get_tminmax :50-64).
struct PixelRay {
    float o[3], d[3], step, tmin, tmax;
    float inv_d[3];  // 1 / d (INFINITY for d = 0), for the empty-space exits only
};

__device__ __forceinline__ void pixel_dir(const RenderArgs& A, int px, int py, float* rd) {
    const float pi = (float)px + 0.5f, pj = (float)py + 0.5f;
    const float dx = (pi - A.cx) / A.fx;
    const float dy0 = (pj - A.cy) / A.fy;
    const float dy = A.inverse_y ? dy0 : -dy0;
    const float dz = A.inverse_y ? 1.f : -1.f;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        rd[a] = dx * A.c2w[4 * a] + dy * A.c2w[4 * a + 1] + dz * A.c2w[4 * a + 2];
}

// The pixel's PE row (27 components of the view direction, dvgo.py:91-109
// order, then the bias input 1 and four zeros) as 16 packed fp16 pairs.
__device__ __forceinline__ void pixel_pe(const RenderArgs& A, int px, int py, uint32_t* pe) {
    float rd[3];
    pixel_dir(A, px, py, rd);
    const float len = sqrtf(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]);
    const float inv = 1.f / len;
    const float vd[3] = {rd[0] * inv, rd[1] * inv, rd[2] * inv};
    float p27[28];
    p27[0] = vd[0];
    p27[1] = vd[1];
    p27[2] = vd[2];
    float base = 1.f;
#pragma unroll
    for (int q = 3; q < 7; ++q) {
        p27[q] = sinf(vd[0] * base);
        p27[q + 4] = sinf(vd[1] * base);
        p27[q + 8] = sinf(vd[2] * base);
        p27[q + 12] = cosf(vd[0] * base);
        p27[q + 16] = cosf(vd[1] * base);
        p27[q + 20] = cosf(vd[2] * base);
        base *= 2.f;
    }
    p27[27] = 1.f;  // the bias input
#pragma unroll
    for (int q = 0; q < 14; ++q) pe[q] = pack2(p27[2 * q], p27[2 * q + 1]);
    pe[14] = pe[15] = 0u;
}

__device__ __forceinline__ PixelRay pixel_ray(const RenderArgs& A, int px, int py) {
    PixelRay R;
    float rd[3];
    pixel_dir(A, px, py, rd);
#pragma unroll
    for (int a = 0; a < 3; ++a) R.o[a] = (A.c2w[4 * a + 3] - A.mn[a]) / A.ext[a];
    const float len = sqrtf(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]);
    R.step = A.stepdist / len;
    float lo = -INFINITY, hi = INFINITY;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        R.d[a] = rd[a] / A.ext[a];
        R.inv_d[a] = 1.f / R.d[a];
        const float v = R.d[a] == 0.f ? 1e-6f : R.d[a];
        const float ta = (1.f - R.o[a]) / v, tb = (-R.o[a]) / v;
        lo = a == 0 ? fminf(ta, tb) : fmaxf(lo, fminf(ta, tb));
        hi = a == 0 ? fmaxf(ta, tb) : fminf(hi, fmaxf(ta, tb));
    }
    R.tmin = fmaxf(fminf(lo, 1e9f), A.near_);  // far is forced to 1e9 (plenvdb.h:1008)
    R.tmax = fmaxf(fminf(hi, 1e9f), A.near_);
    return R;
}

// One step of a march: advances t, applies the box and isActive(Round(xyz))
// tests and, for a sample that passes, forms its 8 corner slots (int()
// truncation, renderer.cu:271-300) and trilinear weights.  Returns false for a
// rejected step.  Missing leaves are stepped over with the t += step
// accumulation alone (see the comment inside), so t stays the reference's
// running sum and the set of tested steps is unchanged.
// Takes steps t <- t + s (float, round to nearest) while the next t stays
// below tsafe, t < tend and iter < max_iter -- the loop
//     while (t + s < tsafe && t < tend && iter < max_iter) { t += s; ++iter; }
// with the same result, but runs of steps inside one binade [2^(e-1), 2^e)
// of t are taken in O(1): there t is an integer multiple T of the binade's
// ulp u, and t + s rounds to (T + D) u with D = s / u rounded to nearest,
// the same D for every step, unless s / u lies exactly half way between two
// integers (round-half-even would then depend on T's parity) or t + s
// reaches the binade's top.  So k steps give (T + k D) u exactly, with k
// bounded below the binade's top and below the stopping conditions (counted
// in double, one step of margin); the loop finishes the rest.
__device__ __forceinline__ void advance_steps(float& t, float s, float tsafe, float tend,
                                              int& iter, int max_iter) {
#pragma unroll 1
    for (int pass = 0; pass < 4; ++pass) {
        if (!(t >= 1e-20f) || !(s > 0.f) || iter >= max_iter) break;
        int e;
        (void)frexpf(t, &e);                      // t in [2^(e-1), 2^e)
        const float u = ldexpf(1.f, e - 24);      // ulp of the binade
        const float q = s / u;                    // exact: a power-of-two scale
        if (!(q >= 1.f) || !(q < 8388608.f) || q - floorf(q) == 0.5f) break;
        const double D = (double)rintf(q);
        const double T = (double)(t / u);         // integer in [2^23, 2^24)
        // steps j -> j + 1 inside the binade while T + (j + 1) D <= 2^24 - 1
        const double kb = floor((16777215.0 - T) / D);
        // the loop's conditions for step j: T + (j + 1) D < tsafe / u, T + j D < tend / u
        const double ks = ceil(((double)tsafe / (double)u - T) / D) - 1.0;
        const double ke = ceil(((double)tend / (double)u - T) / D);
        const double k = fmin(fmin(kb, ks), fmin(ke, (double)(max_iter - iter))) - 1.0;
        if (!(k >= 4.0)) break;
        const int ki = (int)k;
        t = ldexpf((float)(T + (double)ki * D), e - 24);
        iter += ki;
        if (kb - 1.0 > k) break;  // stopped by tsafe, tend or the cap: the loop ends it
        // at the binade's top: plain steps into the next binade, then again
        for (int z = 0; z < 4 && t + s < tsafe && t < tend && iter < max_iter; ++z) {
            t += s;
            ++iter;
        }
    }
    while (t + s < tsafe && t < tend && iter < max_iter) {
        t += s;
        ++iter;
    }
}

struct Corners {
    int32_t slot[8];  // the host keeps nslots + 1 below 2^31
    float u, v, w;
    int i0, j0, k0;   // the cell's lower corner
};

// Slots of the 8 corners of the cell with lower corner (i0, j0, k0) in the
// corner order 000,001,011,010,110,111,101,100 (slot nslots, index 0's
// values, for a corner whose leaf does not exist).
__device__ __forceinline__ void corner_slots(const RenderArgs& A, int i0, int j0, int k0,
                                             BlockCache& bc, int32_t* slot) {
    if (((i0 & 7) != 7) & ((j0 & 7) != 7) & ((k0 & 7) != 7)) {
        // the cell lies inside one 8^3 block: one lookup, slots by in-leaf offset
        const int32_t l0 = bc.get(A.topo, i0, j0, k0);
        const int32_t s0 = l0 * 512 + voxel_offset(i0, j0, k0);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int cx = (kCx >> q) & 1, cy = (kCy >> q) & 1, cz = (kCz >> q) & 1;
            slot[q] = l0 < 0 ? (int32_t)A.nslots : s0 + ((cx << 6) | (cy << 3) | cz);
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int cx = (kCx >> q) & 1, cy = (kCy >> q) & 1, cz = (kCz >> q) & 1;
        const int32_t lq = bc.get(A.topo, i0 + cx, j0 + cy, k0 + cz);
        slot[q] = lq < 0 ? (int32_t)A.nslots : lq * 512 + voxel_offset(i0 + cx, j0 + cy, k0 + cz);
    }
}

__device__ __forceinline__ bool march_step(const RenderArgs& A, const PixelRay& R, float& t,
                                           float tend, int& iter, int max_iter, BlockCache& bc,
                                           Corners& c) {
    t += R.step;
    ++iter;
    const float p0 = R.o[0] + t * R.d[0], p1 = R.o[1] + t * R.d[1], p2 = R.o[2] + t * R.d[2];
    if (!(p0 >= 0.f && p1 >= 0.f && p2 >= 0.f && p0 <= 1.f && p1 <= 1.f && p2 <= 1.f))
        return false;
    const float x = p0 * A.ws[0], y = p1 * A.ws[1], z = p2 * A.ws[2];
    // isActive(Round(xyz)): rintf, half to even (NanoVDB.h:774-779)
    const int vi = (int)rintf(x), vj = (int)rintf(y), vk = (int)rintf(z);
    const int32_t l = bc.get(A.topo, vi, vj, vk);
    if (l < 0) {
        // Empty space: the rounded position is in a missing leaf, and by the
        // empty-run map no leaf lies within r - 1 cells of it, so every step
        // until the ray leaves the rounding box of that cube of cells,
        // [8(L - r + 1) - 0.5, 8(L + r) - 0.5), is rejected by this same test.
        // Those steps are taken with the accumulation alone (in closed form,
        // advance_steps), stopping two steps short of the computed exit, which
        // covers its rounding (the exit is formed with reciprocals: a relative
        // error of a few ulps, far inside the two steps).
        const int L[3] = {vi >> 3, vj >> 3, vk >> 3};
        const int r = empty_run(A.topo, L[0], L[1], L[2]);
        float te = INFINITY;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (R.d[a] == 0.f) continue;
            const float xb = R.d[a] > 0.f ? 8.f * (float)(L[a] + r) - 0.5f
                                          : 8.f * (float)(L[a] - r + 1) - 0.5f;
            te = fminf(te, (xb * A.inv_ws[a] - R.o[a]) * R.inv_d[a]);
        }
        advance_steps(t, R.step, te - 2.f * R.step, tend, iter, max_iter);
        return false;
    }
    const int off = voxel_offset(vi, vj, vk);
    if (!((A.topo.masks[(int64_t)l * 8 + (off >> 6)] >> (off & 63)) & 1ull)) return false;
    const int i0 = (int)x, j0 = (int)y, k0 = (int)z;  // x >= 0 here: int() == floor
    c.u = x - (float)i0;
    c.v = y - (float)j0;
    c.w = z - (float)k0;
    c.i0 = i0;
    c.j0 = j0;
    c.k0 = k0;
    corner_slots(A, i0, j0, k0, bc, c.slot);
    return true;
}

__device__ __forceinline__ int march_cap(const PixelRay& R, float t0, float t1) {
    // the reference loop ends when t reaches its bound; the cap only guards
    // steps too small to advance t
    const float span = (t1 - t0) / R.step;
    return (span >= 0.f) ? (int)fminf(span * 1.0625f + 16.f, 16777216.f) : 0;
}

// The pixel of lane `lane` of tile `tile` (8 x 8 pixels, row-major tiles).
struct TilePixel {
    int px, py;
    bool valid;
    int64_t pix;
};
__device__ __forceinline__ TilePixel tile_pixel(const RenderArgs& A, int tile, int lane) {
    TilePixel P;
    P.px = (tile % A.ntx) * 8 + (lane & 7);
    P.py = (tile / A.ntx) * 8 + (lane >> 3);
    P.valid = P.px < A.W && P.py < A.H;
    P.pix = (int64_t)P.py * A.W + P.px;
    return P;
}

__device__ __forceinline__ void wave_count_add(unsigned long long* dst, unsigned long long v,
                                               int lane) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0 && v) atomicAdd(dst, v);
}

// Pass 1, first_look_merged (renderer.cu:222-268): a wave per 8 x 8 tile, a
// lane per pixel.  Counts the kept samples with trigetDensity's summation
// order and records the rewritten t range: the first kept t minus one step,
// and the t of the early stop (or the box exit).
template <bool F32PAY>
__global__ void __launch_bounds__(kMarchThreads) k_render_count(RenderArgs A) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * (kMarchThreads / 64) + (threadIdx.x >> 6);
    if (tile >= A.ntiles) return;  // uniform over the wave
    const FrameBuffers& F = A.fb;
    const TilePixel tp = tile_pixel(A, tile, lane);
    unsigned long long n_gather = 0;
    if (tp.valid) {
        const PixelRay R = pixel_ray(A, tp.px, tp.py);
        BlockCache bc;
        float t = R.tmin, T = 1.f, tmin2 = R.tmin, tmax2 = R.tmax;
        int cnt = 0, iter = 0;
        const int cap1 = march_cap(R, R.tmin, R.tmax);
        bool first = false;
        while (t < R.tmax && iter < cap1) {
            Corners c;
            if (!march_step(A, R, t, R.tmax, iter, cap1, bc, c)) continue;
            float dv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) dv[q] = load_den<F32PAY>(A.den, c.slot[q]);
            // trigetDensity order (:191-220): res += den * (1-u) * (1-v) * (1-w), ...
            float sig = 0.f;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float a = ((kCx >> q) & 1) ? c.u : 1.f - c.u;
                const float b = ((kCy >> q) & 1) ? c.v : 1.f - c.v;
                const float e = ((kCz >> q) & 1) ? c.w : 1.f - c.w;
                sig = sig + ((dv[q] * a) * b) * e;
            }
            ++n_gather;
            const float alpha = 1.f - powf(1.f + expf(sig + A.act_shift), -A.interval);
            if (alpha <= A.thres) continue;
            const float wt = T * alpha;
            T = T * (1.f - alpha);
            if (wt <= A.thres) continue;
            ++cnt;
            if (!first) {
                tmin2 = t - R.step;
                first = true;
            }
            if (T < A.t_stop) {
                tmax2 = t;
                break;
            }
        }
        F.pix_count[tp.pix] = cnt;
        F.tspan[tp.pix] = make_float2(tmin2, tmax2);
    }
    if (A.counters) wave_count_add(A.counters, n_gather, lane);
}

// Pass 2, ray_marching_merged (renderer.cu:312-367), from the rewritten t_min
// to t_max with trigetDensity2's summation order: the tile reserves its records
// with one atomic (pass 1's counts scanned across the wave) and writes one
// 32-byte record per kept sample (12 colour features in fp16, the weight, the
// pixel), at most pass 1's count of them, contiguous per pixel.
template <bool F32PAY>
__global__ void __launch_bounds__(kMarchThreads) k_render_fill(RenderArgs A) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * (kMarchThreads / 64) + (threadIdx.x >> 6);
    if (tile >= A.ntiles) return;  // uniform over the wave
    const FrameBuffers& F = A.fb;
    const TilePixel tp = tile_pixel(A, tile, lane);
    const int cnt = tp.valid ? F.pix_count[tp.pix] : 0;
    int incl = cnt;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
        const int o = __shfl_up(incl, dd, 64);
        if (lane >= dd) incl += o;
    }
    const int tile_total = __shfl(incl, 63, 64);
    unsigned long long base = 0;
    if (lane == 0 && tile_total > 0) base = atomicAdd(F.total, (unsigned long long)tile_total);
    base = __shfl(base, 0, 64);
    const int64_t start = (int64_t)base + (incl - cnt);

    int r = 0;
    float T = 1.f;
    if (cnt > 0) {
        const PixelRay R = pixel_ray(A, tp.px, tp.py);
        const float2 span = F.tspan[tp.pix];
        const float tmax2 = span.y;
        BlockCache bc;
        float t = span.x;
        int iter = 0;
        const int cap2 = march_cap(R, span.x, tmax2);
        while (t < tmax2 && iter < cap2) {
            Corners c;
            if (!march_step(A, R, t, tmax2, iter, cap2, bc, c)) continue;
            float dv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) dv[q] = load_den<F32PAY>(A.den, c.slot[q]);
            float sc[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float a = ((kCx >> q) & 1) ? c.u : 1.f - c.u;
                const float b = ((kCy >> q) & 1) ? c.v : 1.f - c.v;
                const float e = ((kCz >> q) & 1) ? c.w : 1.f - c.w;
                sc[q] = (a * b) * e;
            }
            // trigetDensity2 order (:271-300): sum of den * scale, left to right
            float sig = dv[0] * sc[0];
#pragma unroll
            for (int q = 1; q < 8; ++q) sig = sig + dv[q] * sc[q];
            const float alpha = 1.f - powf(1.f + expf(sig + A.act_shift), -A.interval);
            if (alpha <= A.thres) continue;
            const float wt = T * alpha;
            T = T * (1.f - alpha);
            if (wt <= A.thres) continue;
            if (r < cnt) {
                // the sample's cell and fractions; k_render_gather adds its colour
                const int64_t q = start + r;
                if (q < F.capacity) {
                    F.rec[2 * q] = make_uint4((uint32_t)c.i0, (uint32_t)c.j0, (uint32_t)c.k0,
                                              __float_as_uint(c.u));
                    F.rec[2 * q + 1] = make_uint4(__float_as_uint(c.v), __float_as_uint(c.w),
                                                  __float_as_uint(wt), (uint32_t)tp.pix);
                }
            }
            ++r;  // pass 2 may keep more than pass 1 counted; those are not written
        }
    }
    const int written = r < cnt ? r : cnt;
    // reserved records pass 2 did not fill: no cell (i0 = -1), zero weight
    for (int k = written; k < cnt; ++k) {
        const int64_t q = start + k;
        if (q < F.capacity) {
            F.rec[2 * q] = make_uint4(0xffffffffu, 0u, 0u, 0u);
            F.rec[2 * q + 1] = make_uint4(0u, 0u, 0u, (uint32_t)tp.pix);
        }
    }
    if (tp.valid) {
        F.pix_start[tp.pix] = (int32_t)(start < kMaxRecords ? start : kMaxRecords);
        F.pix_written[tp.pix] = written;
        F.t_last[tp.pix] = cnt > 0 ? T : 1.f;
        if (cnt > 0) {
            uint32_t pe[16];
            pixel_pe(A, tp.px, tp.py, pe);
            uint4* row = F.pe + 4 * tp.pix;
            row[0] = make_uint4(pe[0], pe[1], pe[2], pe[3]);
            row[1] = make_uint4(pe[4], pe[5], pe[6], pe[7]);
            row[2] = make_uint4(pe[8], pe[9], pe[10], pe[11]);
            row[3] = make_uint4(pe[12], pe[13], pe[14], pe[15]);
        }
    }
    if (A.counters) wave_count_add(A.counters + 1, (unsigned long long)written, lane);
}

// Colour of each record (trigetColor, renderer.cu:303-310: corner 0's row
// times its weight, then the others added in corner order, fp32), written over
// the record's cell as 12 fp16 features: a thread per record, so the colour
// loads run at full lane occupancy instead of inside the marches' divergent
// loops.  Records: in (i0, j0, k0, u | v, w, weight, pixel), out (features
// 0..7 | 8..11, weight, pixel); i0 = -1 marks a reserved record without a
// sample (zero features).
template <bool F32PAY>
__global__ void __launch_bounds__(256) k_render_gather(RenderArgs A) {
    const FrameBuffers& F = A.fb;
    const unsigned long long tot = *F.total;
    const int64_t n = (int64_t)(tot < (unsigned long long)F.capacity ? tot : F.capacity);
    // one thread per record (the grid covers the capacity): each thread's
    // chain of dependent loads (record, leaf table, colour rows) is short, and
    // the waves in flight overlap them
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) {
        const uint4 a = F.rec[2 * q], b = F.rec[2 * q + 1];
        uint4 o0 = make_uint4(0u, 0u, 0u, 0u), o1 = make_uint4(0u, 0u, b.z, b.w);
        const int i0 = (int)a.x, j0 = (int)a.y, k0 = (int)a.z;
        if (i0 >= 0) {
            const float u = __uint_as_float(a.w), v = __uint_as_float(b.x), w = __uint_as_float(b.y);
            BlockCache bc;
            int32_t slot[8];
            corner_slots(A, i0, j0, k0, bc, slot);
            float sc[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const float wa = ((kCx >> c) & 1) ? u : 1.f - u;
                const float wb = ((kCy >> c) & 1) ? v : 1.f - v;
                const float we = ((kCz >> c) & 1) ? w : 1.f - w;
                sc[c] = (wa * wb) * we;
            }
            float f[12], cq[12];
            load_col<F32PAY>(A.payload, slot[0], cq);
#pragma unroll
            for (int ch = 0; ch < 12; ++ch) f[ch] = cq[ch] * sc[0];
#pragma unroll
            for (int c = 1; c < 8; ++c) {
                load_col<F32PAY>(A.payload, slot[c], cq);
#pragma unroll
                for (int ch = 0; ch < 12; ++ch) f[ch] = f[ch] + cq[ch] * sc[c];
            }
            o0 = make_uint4(pack2(f[0], f[1]), pack2(f[2], f[3]), pack2(f[4], f[5]),
                            pack2(f[6], f[7]));
            o1.x = pack2(f[8], f[9]);
            o1.y = pack2(f[10], f[11]);
        }
        F.rec[2 * q] = o0;
        F.rec[2 * q + 1] = o1;
    }
}

// The MLP over the records, 64 per wave-chunk; each record's contribution
// w / (1 + exp(-raw)) per channel (final_render, renderer.cu:112-119).
__global__ void __launch_bounds__(64 * kWaves) k_render_mlp(RenderArgs A) {
    extern __shared__ uint4 smem[];
    uint4* sW = smem;  // weights + biases
    float* sB = reinterpret_cast<float*>(reinterpret_cast<uint8_t*>(smem) + kPackedHalves * 2);
    _Float16* sXall = reinterpret_cast<_Float16*>(reinterpret_cast<uint8_t*>(smem) + kPackedBytes);
    float* sLall = reinterpret_cast<float*>(sXall + kWaves * 64 * XS);

    for (int i = threadIdx.x; i < kPackedBytes / 16; i += blockDim.x) sW[i] = A.packed[i];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    _Float16* sX = sXall + wv * 64 * XS;
    float* sL = sLall + wv * 64 * 4;
    uint4* row = reinterpret_cast<uint4*>(sX + lane * XS);
    row[6] = row[7] = make_uint4(0u, 0u, 0u, 0u);  // inputs 48..63 stay zero
    __syncthreads();

    const FrameBuffers& F = A.fb;
    const unsigned long long tot = *F.total;
    if (A.counters && blockIdx.x == 0 && threadIdx.x == 0) A.counters[2] = tot;
    const int64_t n = (int64_t)(tot < (unsigned long long)F.capacity ? tot : F.capacity);
    const int64_t nchunks = (n + 63) / 64;
    const int64_t stride = (int64_t)gridDim.x * kWaves;
    // Two-stage prefetch: a chunk's records arrive one chunk ahead and its
    // pixels' PE rows (which need the records) are loaded while the chunk
    // before it runs the MLP, so neither load's latency sits in front of the
    // MFMAs.
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    auto load_rec = [&](int64_t ch, uint4& a, uint4& b) {
        const int64_t q = ch * 64 + lane;
        a = b = z;
        if (ch < nchunks && q < n) {
            a = F.rec[2 * q];
            b = F.rec[2 * q + 1];
        }
    };
    auto load_pe = [&](int64_t ch, const uint4& b, uint4* pe) {
        const int64_t q = ch * 64 + lane;
        pe[0] = pe[1] = pe[2] = pe[3] = z;
        if (ch < nchunks && q < n) {
            const uint4* pr = F.pe + 4 * (int64_t)b.w;
            pe[0] = pr[0];
            pe[1] = pr[1];
            pe[2] = pr[2];
            pe[3] = pr[3];
        }
    };
    int64_t ch = (int64_t)blockIdx.x * kWaves + wv;
    uint4 ca, cb, cpe[4], na, nb;
    load_rec(ch, ca, cb);
    load_pe(ch, cb, cpe);
    load_rec(ch + stride, na, nb);
    for (; ch < nchunks; ch += stride) {
        const int64_t q = ch * 64 + lane;
        const bool live = q < n;
        row[0] = ca;
        row[1] = make_uint4(cb.x, cb.y, 0u, 0u);
        row[2] = cpe[0];
        row[3] = cpe[1];
        row[4] = cpe[2];
        row[5] = cpe[3];
        // the next chunk's PE rows and the records of the one after
        uint4 npe[4], fa, fb;
        load_pe(ch + stride, nb, npe);
        load_rec(ch + 2 * stride, fa, fb);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        wave_mlp(sW, sB, sX, sL, lane);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (live) {
            const float w = __uint_as_float(cb.z);
            const float* lg = sL + lane * 4;
            // w * sigmoid(raw) with the hardware exp and reciprocal (a few ulps,
            // far inside the fp16 MLP's own rounding)
            F.contrib[q] = make_float4(w * __builtin_amdgcn_rcpf(1.f + __expf(-lg[0])),
                                       w * __builtin_amdgcn_rcpf(1.f + __expf(-lg[1])),
                                       w * __builtin_amdgcn_rcpf(1.f + __expf(-lg[2])), 0.f);
        }
        __builtin_amdgcn_wave_barrier();  // sL and the rows are rewritten by the next chunk
        ca = na;
        cb = nb;
        cpe[0] = npe[0];
        cpe[1] = npe[1];
        cpe[2] = npe[2];
        cpe[3] = npe[3];
        na = fa;
        nb = fb;
    }
}

// Pixel colour: T_last * bg + the pixel's contributions in sample order.
constexpr int kCompLanes = 8;  // lanes per pixel in k_render_composite

__global__ void __launch_bounds__(256) k_render_composite(RenderArgs A) {
    // lane j of a pixel's eight sums the pixel's records j, j + 8, ... (the
    // eight read 128 contiguous bytes per step), then the eight add up their
    // sums; every lane takes part in the exchange (no early exit)
    const int64_t p = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kCompLanes;
    const int j = (int)(threadIdx.x & (kCompLanes - 1));
    const bool on = p < (int64_t)A.H * A.W;  // uniform over a pixel's lanes
    const FrameBuffers& F = A.fb;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    if (on) {
        const unsigned long long tot = *F.total;
        const int64_t n = (int64_t)(tot < (unsigned long long)F.capacity ? tot : F.capacity);
        const int64_t s = F.pix_start[p];
        const int wn = F.pix_written[p];
        for (int k = j; k < wn; k += kCompLanes) {
            const int64_t q = s + k;
            if (q >= n) break;
            const float4 c = F.contrib[q];
            a0 = a0 + c.x;
            a1 = a1 + c.y;
            a2 = a2 + c.z;
        }
    }
#pragma unroll
    for (int d = 1; d < kCompLanes; d <<= 1) {
        a0 += __shfl_xor(a0, d, 64);
        a1 += __shfl_xor(a1, d, 64);
        a2 += __shfl_xor(a2, d, 64);
    }
    if (on && j == 0) {
        const float last = F.t_last[p] * A.bg;
        A.out[3 * p] = last + a0;
        A.out[3 * p + 1] = last + a1;
        A.out[3 * p + 2] = last + a2;
    }
}

__global__ void k_render_reset(unsigned long long* total, unsigned long long* counters) {
    if (threadIdx.x == 0) *total = 0ull;
    if (counters && threadIdx.x < 3) counters[threadIdx.x] = 0ull;
}

constexpr size_t kRenderSmem =
    (size_t)kPackedBytes + (size_t)kWaves * 64 * XS * 2 + (size_t)kWaves * 64 * 4 * 4;

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Optional device timing of the three kernels (HIP events around each launch).
struct RenderTimer {
    bool enabled = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> open;
    double total_ms[3] = {0, 0, 0};
};
RenderTimer g_rtimer;

struct RenderStage {
    int stage;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    RenderStage(int st, hipStream_t str) : stage(st), s(str) {
        if (!g_rtimer.enabled) return;
        hipEventCreate(&a);
        hipEventCreate(&b);
        hipEventRecord(a, s);
    }
    ~RenderStage() {
        if (!g_rtimer.enabled || !a) return;
        hipEventRecord(b, s);
        g_rtimer.open.push_back({stage, {a, b}});
    }
};

// Payload of (nslots + 1) slots: the slot rows (32 or 64 bytes: density and
// 12 colour channels), then the densities alone (2 or 4 bytes a slot).
struct PayloadLayout {
    size_t den, bytes;
};
PayloadLayout payload_layout(const TopoView& t, int32_t payload_f32) {
    const size_t n = (size_t)t.nleaf * 512 + 1, es = payload_f32 ? 4 : 2;
    PayloadLayout L;
    L.den = align256(n * 16 * es);
    L.bytes = align256(L.den + n * es);
    return L;
}

struct FrameLayout {
    size_t total, start, count, tspan, written, tlast, pe, rec, contrib, bytes;
};

FrameLayout frame_layout(int64_t hw, int64_t capacity) {
    FrameLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.total = take(8);
    L.start = take(4 * (size_t)hw);
    L.count = take(4 * (size_t)hw);
    L.tspan = take(8 * (size_t)hw);
    L.written = take(4 * (size_t)hw);
    L.tlast = take(4 * (size_t)hw);
    L.pe = take(64 * (size_t)hw);
    L.rec = take(32 * (size_t)capacity);
    L.contrib = take(16 * (size_t)capacity);
    L.bytes = o;
    return L;
}

FrameBuffers frame_buffers(void* ws, int64_t hw, int64_t capacity) {
    const FrameLayout L = frame_layout(hw, capacity);
    uint8_t* b = reinterpret_cast<uint8_t*>(ws);
    FrameBuffers F;
    F.total = reinterpret_cast<unsigned long long*>(b + L.total);
    F.pix_start = reinterpret_cast<int32_t*>(b + L.start);
    F.pix_count = reinterpret_cast<int32_t*>(b + L.count);
    F.tspan = reinterpret_cast<float2*>(b + L.tspan);
    F.pix_written = reinterpret_cast<int32_t*>(b + L.written);
    F.t_last = reinterpret_cast<float*>(b + L.tlast);
    F.pe = reinterpret_cast<uint4*>(b + L.pe);
    F.rec = reinterpret_cast<uint4*>(b + L.rec);
    F.contrib = reinterpret_cast<float4*>(b + L.contrib);
    F.capacity = capacity;
    return F;
}

}  // namespace

extern "C" {

int64_t pvdb_render_weights_bytes(void) { return kPackedBytes; }

int pvdb_render_pack_weights(const float* w0, const float* b0, const float* w1, const float* b1,
                             const float* w2, const float* b2, int32_t dcol, int32_t dpe,
                             int32_t dhid, int32_t dout, void* packed, void* stream) {
    PVDB_CHECK_ARG(w0 && b0 && w1 && b1 && w2 && b2 && packed, "null pointer");
    PVDB_CHECK_ARG(dcol == 12 && dpe == 27 && dhid == kHid && dout == 3,
                   "unsupported MLP shape: the renderer takes dcol 12, dpe 27, dhid 128, dout 3");
    uint8_t* p = reinterpret_cast<uint8_t*>(packed);
    const int n = kPackedHalves + kBiasFloats;
    hipLaunchKernelGGL(k_render_pack, dim3(pvdb_blocks(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, w0, b0, w1, b1, w2, b2, dcol, dpe, dout,
                       reinterpret_cast<_Float16*>(p),
                       reinterpret_cast<float*>(p + kPackedHalves * 2));
    PVDB_LAUNCHED("k_render_pack");
    return PVDB_OK;
}

int64_t pvdb_render_payload_bytes(const pvdb_topology* topo, int32_t payload_f32) {
    if (!topo) return -1;
    return (int64_t)payload_layout(view_of(topo), payload_f32).bytes;
}

int pvdb_render_build_payload(const pvdb_topology* topo, const int32_t* slot_idx,
                              const float* den, const float* col, int64_t nidx,
                              int32_t payload_f32, void* payload, void* stream) {
    if (!topo) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    PVDB_CHECK_ARG(slot_idx && den && col && payload, "null pointer");
    PVDB_CHECK_ARG(nidx >= 1, "the data arrays need at least index 0");
    const TopoView tv = view_of(topo);
    const int64_t nslots = tv.nleaf * 512;
    const PayloadLayout L = payload_layout(tv, payload_f32);
    uint8_t* p = reinterpret_cast<uint8_t*>(payload);
    if (payload_f32)
        hipLaunchKernelGGL(k_render_payload<float>, dim3(pvdb_blocks(nslots + 1, 256)), dim3(256),
                           0, (hipStream_t)stream, slot_idx, nslots, den, col, nidx,
                           reinterpret_cast<float*>(p), reinterpret_cast<float*>(p + L.den));
    else
        hipLaunchKernelGGL(k_render_payload<_Float16>, dim3(pvdb_blocks(nslots + 1, 256)),
                           dim3(256), 0, (hipStream_t)stream, slot_idx, nslots, den, col, nidx,
                           reinterpret_cast<_Float16*>(p), reinterpret_cast<_Float16*>(p + L.den));
    PVDB_LAUNCHED("k_render_payload");
    return PVDB_OK;
}

int64_t pvdb_render_workspace_bytes(int32_t H, int32_t W, int64_t capacity) {
    if (H <= 0 || W <= 0 || capacity < 0 || capacity >= kMaxRecords) return -1;
    return (int64_t)frame_layout((int64_t)H * W, capacity).bytes;
}

int pvdb_render_frame_info(const pvdb_render_desc* d, uint64_t* total, int32_t** pix_count,
                           int32_t** pix_written, int32_t** pix_start) {
    PVDB_CHECK_ARG(d && d->workspace, "null descriptor or workspace");
    const FrameBuffers F = frame_buffers(d->workspace, (int64_t)d->H * d->W, d->capacity);
    if (total) {
        unsigned long long v = 0;
        PVDB_HIP(hipMemcpy(&v, F.total, sizeof(v), hipMemcpyDeviceToHost));
        *total = v;
    }
    if (pix_count) *pix_count = F.pix_count;
    if (pix_written) *pix_written = F.pix_written;
    if (pix_start) *pix_start = F.pix_start;
    return PVDB_OK;
}

}  // extern "C"

namespace {

// The frame on `ws` (a 256-byte aligned workspace of at least
// pvdb_render_workspace_bytes(H, W, capacity) bytes; the callers check it).
int render_image(const pvdb_topology* topo, const pvdb_render_desc* d, void* ws,
                 hipStream_t stream) {
    for (int a = 0; a < 3; ++a) {
        PVDB_CHECK_ARG(d->reso[a] >= 2, "reso must be at least 2");
        PVDB_CHECK_ARG(d->xyz_max[a] > d->xyz_min[a], "empty bounding box");
    }
    const TopoView t = view_of(topo);
    PVDB_CHECK_ARG(t.nleaf * 512 + 1 < kMaxRecords, "too many leaves for the renderer's slot indices");
    RenderArgs A;
    A.topo = t;
    A.payload = d->payload;
    const PayloadLayout PL = payload_layout(t, d->payload_f32);
    A.den = reinterpret_cast<const uint8_t*>(d->payload) + PL.den;
    A.nslots = t.nleaf * 512;
    A.packed = reinterpret_cast<const uint4*>(d->weights);
    A.out = d->out;
    A.counters = reinterpret_cast<unsigned long long*>(d->counters);
    A.fb = frame_buffers(ws, (int64_t)d->H * d->W, d->capacity);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) A.c2w[4 * r + c] = d->c2w[4 * r + c];
    A.fx = d->K[0];
    A.cx = d->K[2];
    A.fy = d->K[4];
    A.cy = d->K[5];
    for (int a = 0; a < 3; ++a) {
        A.mn[a] = d->xyz_min[a];
        A.ext[a] = d->xyz_max[a] - d->xyz_min[a];
        A.ws[a] = (float)(d->reso[a] - 1);
        A.inv_ws[a] = 1.f / A.ws[a];
    }
    A.H = d->H;
    A.W = d->W;
    A.inverse_y = d->inverse_y;
    A.ntx = (d->W + 7) / 8;
    A.ntiles = A.ntx * ((d->H + 7) / 8);
    A.near_ = d->near_;
    A.stepdist = d->stepdist;
    A.act_shift = d->act_shift;
    A.interval = d->interval;
    A.thres = d->fast_color_thres;
    A.bg = d->bg;
    A.t_stop = d->t_stop;
    hipStream_t s = stream;
    hipLaunchKernelGGL(k_render_reset, dim3(1), dim3(64), 0, s, A.fb.total, A.counters);
    PVDB_LAUNCHED("k_render_reset");
    static int ncu = 0;
    if (ncu == 0) {
        int dev = 0;
        PVDB_HIP(hipGetDevice(&dev));
        PVDB_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    // the two marches: a wave per 8 x 8 tile
    const int mgrid = (A.ntiles + kMarchThreads / 64 - 1) / (kMarchThreads / 64);
    {
        RenderStage st(0, s);
        if (d->payload_f32) {
            hipLaunchKernelGGL(k_render_count<true>, dim3(mgrid), dim3(kMarchThreads), 0, s, A);
            PVDB_LAUNCHED("k_render_count");
            hipLaunchKernelGGL(k_render_fill<true>, dim3(mgrid), dim3(kMarchThreads), 0, s, A);
            PVDB_LAUNCHED("k_render_fill");
            hipLaunchKernelGGL(k_render_gather<true>, dim3(pvdb_blocks(A.fb.capacity, 256)),
                               dim3(256), 0, s, A);
        } else {
            hipLaunchKernelGGL(k_render_count<false>, dim3(mgrid), dim3(kMarchThreads), 0, s, A);
            PVDB_LAUNCHED("k_render_count");
            hipLaunchKernelGGL(k_render_fill<false>, dim3(mgrid), dim3(kMarchThreads), 0, s, A);
            PVDB_LAUNCHED("k_render_fill");
            hipLaunchKernelGGL(k_render_gather<false>, dim3(pvdb_blocks(A.fb.capacity, 256)),
                               dim3(256), 0, s, A);
        }
        PVDB_LAUNCHED("k_render_gather");
    }
    static bool smem_set = false;
    if (!smem_set) {
        PVDB_HIP(hipFuncSetAttribute((const void*)k_render_mlp,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRenderSmem));
        smem_set = true;
    }
    {
        RenderStage st(1, s);
        hipLaunchKernelGGL(k_render_mlp, dim3(ncu), dim3(64 * kWaves), kRenderSmem, s, A);
        PVDB_LAUNCHED("k_render_mlp");
    }
    {
        RenderStage st(2, s);
        hipLaunchKernelGGL(k_render_composite,
                           dim3(pvdb_blocks((int64_t)A.H * A.W * kCompLanes, 256)), dim3(256), 0,
                           s, A);
        PVDB_LAUNCHED("k_render_composite");
    }
    return PVDB_OK;
}

}  // namespace

// Frame workspace of the no-allocation entry point.  The array's name is only
// a host-side handle: host code uses the device address hipGetSymbolAddress
// gives for it, kept per device (each device holds its own copy).
constexpr int64_t kScratchBytes = int64_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

// Checks shared by both entry points (everything but the workspace).
int check_frame(const pvdb_topology* topo, const pvdb_render_desc* d) {
    if (!topo) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    PVDB_CHECK_ARG(d && d->payload && d->weights && d->out, "null pointer");
    PVDB_CHECK_ARG(d->H > 0 && d->W > 0 && (int64_t)d->H * d->W < (int64_t(1) << 31),
                   "bad image size");
    PVDB_CHECK_ARG(d->stepdist > 0.f && std::isfinite(d->stepdist), "stepdist must be positive");
    PVDB_CHECK_ARG(d->capacity >= 0 && d->capacity < kMaxRecords,
                   "record capacity must be in [0, 2^31 - 2]");
    return PVDB_OK;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(void** out) {
    int dev = 0;
    PVDB_HIP(hipGetDevice(&dev));
    PVDB_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        PVDB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PVDB_CHECK_ARG(p && reinterpret_cast<uintptr_t>(p) % 256 == 0,
                       "scratch address not 256-byte aligned");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PVDB_OK;
}

}  // namespace

extern "C" {

int pvdb_render_image(const pvdb_topology* topo, const pvdb_render_desc* d, void* stream) {
    const int rc = check_frame(topo, d);
    if (rc != PVDB_OK) return rc;
    PVDB_CHECK_ARG(d->workspace, "null workspace");
    PVDB_CHECK_ARG(reinterpret_cast<uintptr_t>(d->workspace) % 256 == 0,
                   "workspace must be 256-byte aligned");
    PVDB_CHECK_ARG(d->workspace_bytes >= pvdb_render_workspace_bytes(d->H, d->W, d->capacity),
                   "workspace too small");
    return render_image(topo, d, d->workspace, (hipStream_t)stream);
}

int64_t pvdb_render_noalloc_scratch_bytes(void) { return kScratchBytes; }

int pvdb_render_image_noalloc(const pvdb_topology* topo, const pvdb_render_desc* d,
                              void* stream) {
    const int rc = check_frame(topo, d);
    if (rc != PVDB_OK) return rc;
    const int64_t need = pvdb_render_workspace_bytes(d->H, d->W, d->capacity);
    PVDB_CHECK_ARG(need > 0 && need <= kScratchBytes,
                   "frame does not fit the 64 MiB scratch of pvdb_render_image_noalloc "
                   "(pvdb_render_workspace_bytes(H, W, capacity) must be at most "
                   "pvdb_render_noalloc_scratch_bytes())");
    void* ws = nullptr;
    const int rs = scratch_of_current_device(&ws);
    if (rs != PVDB_OK) return rs;
    return render_image(topo, d, ws, (hipStream_t)stream);
}

int pvdb_render_timing(int32_t enable, double* total_ms3) {
    for (auto& e : g_rtimer.open) {
        hipEventSynchronize(e.second.second);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e.second.first, e.second.second);
        g_rtimer.total_ms[e.first] += ms;
        hipEventDestroy(e.second.first);
        hipEventDestroy(e.second.second);
    }
    g_rtimer.open.clear();
    for (int i = 0; i < 3; ++i) {
        if (total_ms3) total_ms3[i] = g_rtimer.total_ms[i];
        if (enable >= 0) g_rtimer.total_ms[i] = 0;
    }
    if (enable >= 0) g_rtimer.enabled = enable != 0;
    return PVDB_OK;
}

}  // extern "C"
