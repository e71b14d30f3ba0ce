// Merged-VDB renderer: MGRenderer::render_an_image (plenvdb/lib/vdb/plenvdb.h:
// 1029-1035 -> renderer.cu:370-424) as one persistent kernel.
//
// The reference renders in two full marches per pixel (first_look_merged
// counts the kept samples, renderer.cu:222-268; ray_marching_merged re-marches
// and gathers them, :312-367), a host read-back of the sample total, four
// SGEMMs over the materialised sample list (cuda_rgbnet, :83-109) and an
// atomic composite (final_render, :112-119).  Here each wavefront takes an
// 8x8 pixel tile, one pixel per lane, and marches all 64 rays at once with the
// reference's sample rule (first sample at tmin+step, the [0,1]^3 box test,
// isActive(Round(xyz)), raw2alpha, the two fast_color_thres culls, stop after a
// kept sample that leaves T < t_stop).  When every live lane holds its next
// kept sample, the wave evaluates the RGB MLP for those (up to 64) samples on
// v_mfma_f32_16x16x32_f16 and composites them in registers.  No sample list,
// no host synchronisation, no allocation.
//
// Data layout.  The merged index grid (vdb_compression.py:19-59: index 1 +
// rank of an active voxel, 0 elsewhere) is turned into a topology whose value
// masks are the non-zero indices, plus a slot-major payload holding, per
// slot, density and the 12 colour channels of the voxel's index (index 0's
// values for inactive voxels; slot `nslots` holds index 0's values for
// corners whose leaf does not exist).  A slot is 32 bytes (fp16: exact for
// the fp16 data of mergeddata.npz) or 64 bytes (fp32).
//
// MLP.  Layer outputs are computed transposed (features x samples) and stay in
// registers: the D fragment of one layer (lane (g,n): features 16m+4g+i of
// sample n) is the B fragment of the next once the next layer's weights are
// stored with their K index permuted accordingly (see kperm).  Weights are
// fp16 in LDS, read as 16-byte fragments; the first layer takes
// [feature(12) | PE(27) | 1] so its bias is a weight column.
#include <cmath>

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
constexpr int kWaves = 8;   // waves per workgroup: they share the LDS weights
constexpr int XS = 72;      // halves per staged input row (64 + 8)
constexpr int kW0Frag = 8 * 2 * 64 * 8;  // [m-tile 8][k-step 2][lane][8]
constexpr int kW1Frag = 8 * 4 * 64 * 8;  // [m-tile 8][k-step 4][lane][8]
constexpr int kW2Frag = 4 * 64 * 8;      // [k-step 4][lane][8]
constexpr int kPackedHalves = kW0Frag + kW1Frag + kW2Frag;
constexpr int kBiasFloats = 128 + 4;  // b1, b2 (padded)
constexpr int kPackedBytes = kPackedHalves * 2 + kBiasFloats * 4;

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
    const int nin = dcol + dpe;  // the bias column follows the inputs
    if (idx < kW0Frag) {
        const int j = idx & 7, lane = (idx >> 3) & 63, ms = idx >> 9;
        const int s = ms & 1, m = ms >> 1, g = lane >> 4, r = lane & 15;
        const int o = 16 * m + r, in = 32 * s + 8 * g + j;
        float v = 0.f;
        if (in < nin)
            v = w0[in * kHid + o];
        else if (in == nin)
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
                                 int64_t nidx, PT* __restrict__ pay) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nslots) return;
    int64_t id = s < nslots ? slot_idx[s] : 0;  // slot nslots: index 0 (missing leaves)
    if (id < 0 || id >= nidx) id = 0;           // the host checks the range
    PT* p = pay + s * 16;
    p[0] = (PT)den[id];
#pragma unroll
    for (int c = 0; c < 12; ++c) p[1 + c] = (PT)col[id * 12 + c];
    p[13] = p[14] = p[15] = (PT)0.f;
}

struct RenderArgs {
    TopoView topo;
    const void* payload;
    int64_t nslots;
    const uint4* packed;
    float* out;
    unsigned long long* counters;  // [2]: density gathers, kept samples (may be null)
    int32_t* tile_counter;
    float c2w[12];
    float fx, cx, fy, cy;
    float mn[3], ext[3], ws[3];
    int32_t H, W, inverse_y, ntx, ntiles;
    float near_, stepdist, act_shift, interval, thres, bg, t_stop;
};

__device__ __forceinline__ uint32_t pack2(float a, float b) {
    const _Float16 ha = (_Float16)a, hb = (_Float16)b;
    return (uint32_t)__builtin_bit_cast(uint16_t, ha) |
           ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
}

__device__ __forceinline__ float h2f(uint16_t h) {
    return (float)__builtin_bit_cast(_Float16, h);
}

template <bool F32PAY>
__device__ __forceinline__ float load_den(const void* pay, int64_t slot) {
    if (F32PAY) return reinterpret_cast<const float*>(pay)[slot * 16];
    return h2f(reinterpret_cast<const uint16_t*>(pay)[slot * 16]);
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

// RGB logits of the samples staged in this wave's rows of sX (row = lane),
// for the 16-sample tiles flagged in tmask; lanes 0..15 write them to sL.
// Tiles go two at a time, so each weight fragment read from LDS feeds two
// MFMAs and the two tiles' activations stay in registers.
__device__ __forceinline__ void wave_mlp(const uint4* __restrict__ sW, const float* __restrict__ sB,
                                         const _Float16* __restrict__ sX, float* __restrict__ sL,
                                         unsigned tmask, int lane) {
    const int g = lane >> 4, r16 = lane & 15;
    const uint4* W0 = sW;
    const uint4* W1 = sW + kW0Frag / 8;
    const uint4* W2 = W1 + kW1Frag / 8;
    const float b2x = sB[128], b2y = sB[129], b2z = sB[130];
    for (int tp = 0; tp < 2; ++tp) {
        const unsigned pm = (tmask >> (2 * tp)) & 3u;
        if (!pm) continue;  // uniform over the wave
        uint32_t h0[2][4][4];  // [tile][k-step][dword]: layer-1 B fragments (fp16 pairs)
        uint32_t h1[2][4][4];
        {
            f16x8 xb[2][2];
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    xb[u][s] = as_h8(*reinterpret_cast<const uint4*>(
                        sX + (16 * (2 * tp + u) + r16) * XS + 32 * s + 8 * g));
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const f16x8 a0 = as_h8(W0[(m * 2 + 0) * 64 + lane]);
                const f16x8 a1 = as_h8(W0[(m * 2 + 1) * 64 + lane]);
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                    acc = mfma16x16x32h(a0, xb[u][0], acc);
                    acc = mfma16x16x32h(a1, xb[u][1], acc);
                    const int s = m >> 1, d = (m & 1) * 2;
                    h0[u][s][d] = pack2(fmaxf(acc[0], 0.f), fmaxf(acc[1], 0.f));
                    h0[u][s][d + 1] = pack2(fmaxf(acc[2], 0.f), fmaxf(acc[3], 0.f));
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const float4 bias = *reinterpret_cast<const float4*>(sB + 16 * m + 4 * g);
            f32x4 acc[2];
            acc[0] = acc[1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const f16x8 a = as_h8(W1[(m * 4 + s) * 64 + lane]);
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const uint4 bv = make_uint4(h0[u][s][0], h0[u][s][1], h0[u][s][2], h0[u][s][3]);
                    acc[u] = mfma16x16x32h(a, as_h8(bv), acc[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int s = m >> 1, d = (m & 1) * 2;
                h1[u][s][d] = pack2(fmaxf(acc[u][0] + bias.x, 0.f), fmaxf(acc[u][1] + bias.y, 0.f));
                h1[u][s][d + 1] =
                    pack2(fmaxf(acc[u][2] + bias.z, 0.f), fmaxf(acc[u][3] + bias.w, 0.f));
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const uint4 bv = make_uint4(h1[u][s][0], h1[u][s][1], h1[u][s][2], h1[u][s][3]);
                acc = mfma16x16x32h(as_h8(W2[s * 64 + lane]), as_h8(bv), acc);
            }
            if (g == 0 && ((pm >> u) & 1u)) {
                float* dst = sL + (16 * (2 * tp + u) + r16) * 4;
                dst[0] = acc[0] + b2x;
                dst[1] = acc[1] + b2y;
                dst[2] = acc[2] + b2z;
            }
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

template <bool F32PAY>
__global__ void __launch_bounds__(64 * kWaves) k_render(RenderArgs A) {
    extern __shared__ uint4 smem[];
    uint4* sW = smem;  // weights + biases
    float* sB = reinterpret_cast<float*>(reinterpret_cast<uint8_t*>(smem) + kPackedHalves * 2);
    _Float16* sXall = reinterpret_cast<_Float16*>(reinterpret_cast<uint8_t*>(smem) + kPackedBytes);
    float* sLall = reinterpret_cast<float*>(sXall + kWaves * 64 * XS);

    for (int i = threadIdx.x; i < kPackedBytes / 16; i += blockDim.x) sW[i] = A.packed[i];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    _Float16* sX = sXall + wv * 64 * XS;
    float* sL = sLall + wv * 64 * 4;
    {
        // input columns 40..63 stay zero
        uint4* row = reinterpret_cast<uint4*>(sX + lane * XS + 40);
        row[0] = row[1] = row[2] = make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();

    const TopoView& topo = A.topo;
    const int64_t miss = A.nslots;  // payload slot of index 0
    unsigned long long n_gather = 0, n_kept = 0;
    while (true) {
        int tile = 0;
        if (lane == 0) tile = atomicAdd(A.tile_counter, 1);
        tile = __shfl(tile, 0, 64);
        if (tile >= A.ntiles) break;  // uniform over the wave
        const int px = (tile % A.ntx) * 8 + (lane & 7), py = (tile / A.ntx) * 8 + (lane >> 3);
        const bool valid = px < A.W && py < A.H;

        // ---- ray of the pixel (get_rays, renderer.cu:122-167; get_tminmax :50-64) ----
        float o[3], d[3], step = 0.f, tmin = 0.f, tmax = 0.f;
        uint32_t pe[14];
        {
            const float pi = (float)px + 0.5f, pj = (float)py + 0.5f;
            const float dx = (pi - A.cx) / A.fx;
            const float dy0 = (pj - A.cy) / A.fy;
            const float dy = A.inverse_y ? dy0 : -dy0;
            const float dz = A.inverse_y ? 1.f : -1.f;
            float rd[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                o[a] = (A.c2w[4 * a + 3] - A.mn[a]) / A.ext[a];
                rd[a] = dx * A.c2w[4 * a] + dy * A.c2w[4 * a + 1] + dz * A.c2w[4 * a + 2];
            }
            const float len = sqrtf(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]);
            step = A.stepdist / len;
            float lo = -INFINITY, hi = INFINITY;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                d[a] = rd[a] / A.ext[a];
                const float v = d[a] == 0.f ? 1e-6f : d[a];
                const float ta = (1.f - o[a]) / v, tb = (-o[a]) / v;
                lo = a == 0 ? fminf(ta, tb) : fmaxf(lo, fminf(ta, tb));
                hi = a == 0 ? fmaxf(ta, tb) : fminf(hi, fmaxf(ta, tb));
            }
            tmin = fmaxf(fminf(lo, 1e9f), A.near_);  // far is forced to 1e9 (plenvdb.h:1008)
            tmax = fmaxf(fminf(hi, 1e9f), A.near_);
            const float inv = 1.f / len;
            const float vd[3] = {rd[0] * inv, rd[1] * inv, rd[2] * inv};
            float p27[27];
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
#pragma unroll
            for (int q = 0; q < 13; ++q) pe[q] = pack2(p27[2 * q], p27[2 * q + 1]);
            pe[13] = pack2(p27[26], 1.f);  // PE 26 and the bias input
        }
        // march bound: the reference loop ends when t reaches tmax; the cap only
        // guards steps too small to advance t
        const float span = (tmax - tmin) / step;
        const int max_iter = (span >= 0.f) ? (int)fminf(span * 1.0625f + 16.f, 16777216.f) : 0;

        float t = tmin, T = 1.f, acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, wt_keep = 0.f;
        bool done = !valid || !(t < tmax), have = false, stop_after = false;
        int iter = 0;
        BlockCache bc;
        while (true) {
            while (!done && !have) {
                if (!(t < tmax) || iter >= max_iter) {
                    done = true;
                    break;
                }
                t += step;
                ++iter;
                const float p0 = o[0] + t * d[0], p1 = o[1] + t * d[1], p2 = o[2] + t * d[2];
                if (!(p0 >= 0.f && p1 >= 0.f && p2 >= 0.f && p0 <= 1.f && p1 <= 1.f && p2 <= 1.f))
                    continue;
                const float x = p0 * A.ws[0], y = p1 * A.ws[1], z = p2 * A.ws[2];
                // isActive(Round(xyz)): rintf, half to even (NanoVDB.h:774-779)
                {
                    const int vi = (int)rintf(x), vj = (int)rintf(y), vk = (int)rintf(z);
                    const int32_t l = bc.get(topo, vi, vj, vk);
                    if (l < 0) {
                        // Empty space: the rounded position is in a missing leaf, so
                        // every step until the ray leaves that leaf's rounding box
                        // [8L - 0.5, 8L + 7.5) is skipped by this same test.  Take
                        // those steps with the accumulation alone (t stays the
                        // reference's running sum) and stop two steps short of the
                        // computed exit, which covers its rounding.
                        const int L[3] = {vi >> 3, vj >> 3, vk >> 3};
                        float te = INFINITY;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            if (d[a] == 0.f) continue;
                            const float xb = d[a] > 0.f ? 8.f * (float)L[a] + 7.5f
                                                        : 8.f * (float)L[a] - 0.5f;
                            te = fminf(te, (xb / A.ws[a] - o[a]) / d[a]);
                        }
                        const float tsafe = te - 2.f * step;
                        while (t + step < tsafe && t < tmax && iter < max_iter) {
                            t += step;
                            ++iter;
                        }
                        continue;
                    }
                    const int off = voxel_offset(vi, vj, vk);
                    if (!((topo.masks[(int64_t)l * 8 + (off >> 6)] >> (off & 63)) & 1ull)) continue;
                }
                // corners with int() truncation (renderer.cu:271-300), x >= 0 here
                const int i0 = (int)x, j0 = (int)y, k0 = (int)z;
                const float u = x - (float)i0, v = y - (float)j0, w = z - (float)k0;
                int64_t slot[8];
                float sc[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int cx = (kCx >> q) & 1, cy = (kCy >> q) & 1, cz = (kCz >> q) & 1;
                    const int32_t l = bc.get(topo, i0 + cx, j0 + cy, k0 + cz);
                    slot[q] = l < 0 ? miss
                                    : (int64_t)l * 512 + voxel_offset(i0 + cx, j0 + cy, k0 + cz);
                    const float a = cx ? u : 1.f - u, b = cy ? v : 1.f - v, c = cz ? w : 1.f - w;
                    sc[q] = (a * b) * c;
                }
                // trigetDensity2 order: sum of den*scale, left to right
                float sig = load_den<F32PAY>(A.payload, slot[0]) * sc[0];
#pragma unroll
                for (int q = 1; q < 8; ++q) sig = sig + load_den<F32PAY>(A.payload, slot[q]) * sc[q];
                ++n_gather;
                const float e = expf(sig + A.act_shift);
                const float alpha = 1.f - powf(1.f + e, -A.interval);
                if (alpha <= A.thres) continue;
                const float wt = T * alpha;
                T = T * (1.f - alpha);
                if (wt <= A.thres) continue;
                // kept: colour features (trigetColor order) staged as this lane's input row
                float f[12], cq[12];
                load_col<F32PAY>(A.payload, slot[0], cq);
#pragma unroll
                for (int c = 0; c < 12; ++c) f[c] = cq[c] * sc[0];
#pragma unroll
                for (int q = 1; q < 8; ++q) {
                    load_col<F32PAY>(A.payload, slot[q], cq);
#pragma unroll
                    for (int c = 0; c < 12; ++c) f[c] = f[c] + cq[c] * sc[q];
                }
                uint4* row = reinterpret_cast<uint4*>(sX + lane * XS);
                row[0] = make_uint4(pack2(f[0], f[1]), pack2(f[2], f[3]), pack2(f[4], f[5]),
                                    pack2(f[6], f[7]));
                row[1] = make_uint4(pack2(f[8], f[9]), pack2(f[10], f[11]), pe[0], pe[1]);
                row[2] = make_uint4(pe[2], pe[3], pe[4], pe[5]);
                row[3] = make_uint4(pe[6], pe[7], pe[8], pe[9]);
                row[4] = make_uint4(pe[10], pe[11], pe[12], pe[13]);
                wt_keep = wt;
                have = true;
                ++n_kept;
                // first_look_merged stops after a kept sample with T below the threshold
                if (T < A.t_stop) stop_after = true;
            }
            const unsigned long long hm = __ballot(have);
            if (hm == 0ull) break;  // every lane is done
            unsigned tmask = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if ((hm >> (16 * q)) & 0xffffull) tmask |= 1u << q;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            wave_mlp(sW, sB, sX, sL, tmask, lane);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (have) {
                // final_render: weight / (1 + exp(-raw)) (renderer.cu:112-119)
                const float* lg = sL + lane * 4;
                acc0 = acc0 + wt_keep / (1.f + expf(-lg[0]));
                acc1 = acc1 + wt_keep / (1.f + expf(-lg[1]));
                acc2 = acc2 + wt_keep / (1.f + expf(-lg[2]));
                have = false;
                if (stop_after) done = true;
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (valid) {
            const float last = T * A.bg;
            float* dst = A.out + ((int64_t)py * A.W + px) * 3;
            dst[0] = last + acc0;
            dst[1] = last + acc1;
            dst[2] = last + acc2;
        }
    }
    if (A.counters) {
        // wave sums, one atomic per wave
        for (int off = 32; off > 0; off >>= 1) {
            n_gather += __shfl_xor(n_gather, off, 64);
            n_kept += __shfl_xor(n_kept, off, 64);
        }
        if (lane == 0) {
            atomicAdd(A.counters, n_gather);
            atomicAdd(A.counters + 1, n_kept);
        }
    }
}

__global__ void k_render_reset(int32_t* tile_counter, unsigned long long* counters) {
    if (threadIdx.x == 0) *tile_counter = 0;
    if (counters && threadIdx.x < 2) counters[threadIdx.x] = 0ull;
}

constexpr size_t kRenderSmem =
    (size_t)kPackedBytes + (size_t)kWaves * 64 * XS * 2 + (size_t)kWaves * 64 * 4 * 4;

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
    return (view_of(topo).nleaf * 512 + 1) * (payload_f32 ? 64 : 32);
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
    const int64_t nslots = view_of(topo).nleaf * 512;
    if (payload_f32)
        hipLaunchKernelGGL(k_render_payload<float>, dim3(pvdb_blocks(nslots + 1, 256)), dim3(256),
                           0, (hipStream_t)stream, slot_idx, nslots, den, col, nidx,
                           reinterpret_cast<float*>(payload));
    else
        hipLaunchKernelGGL(k_render_payload<_Float16>, dim3(pvdb_blocks(nslots + 1, 256)),
                           dim3(256), 0, (hipStream_t)stream, slot_idx, nslots, den, col, nidx,
                           reinterpret_cast<_Float16*>(payload));
    PVDB_LAUNCHED("k_render_payload");
    return PVDB_OK;
}

int pvdb_render_image(const pvdb_topology* topo, const pvdb_render_desc* d, void* stream) {
    if (!topo) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    PVDB_CHECK_ARG(d && d->payload && d->weights && d->out && d->workspace, "null pointer");
    PVDB_CHECK_ARG(d->H > 0 && d->W > 0 && (int64_t)d->H * d->W < (int64_t(1) << 31),
                   "bad image size");
    PVDB_CHECK_ARG(d->stepdist > 0.f && std::isfinite(d->stepdist), "stepdist must be positive");
    for (int a = 0; a < 3; ++a) {
        PVDB_CHECK_ARG(d->reso[a] >= 2, "reso must be at least 2");
        PVDB_CHECK_ARG(d->xyz_max[a] > d->xyz_min[a], "empty bounding box");
    }
    const TopoView t = view_of(topo);
    RenderArgs A;
    A.topo = t;
    A.payload = d->payload;
    A.nslots = t.nleaf * 512;
    A.packed = reinterpret_cast<const uint4*>(d->weights);
    A.out = d->out;
    A.counters = reinterpret_cast<unsigned long long*>(d->counters);
    A.tile_counter = reinterpret_cast<int32_t*>(d->workspace);
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
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_render_reset, dim3(1), dim3(64), 0, s, A.tile_counter, A.counters);
    PVDB_LAUNCHED("k_render_reset");
    int dev = 0, ncu = 0;
    PVDB_HIP(hipGetDevice(&dev));
    PVDB_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int want = (A.ntiles + kWaves - 1) / kWaves;
    const int grid = want < ncu ? (want < 1 ? 1 : want) : ncu;
    if (d->payload_f32) {
        PVDB_HIP(hipFuncSetAttribute((const void*)k_render<true>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRenderSmem));
        hipLaunchKernelGGL(k_render<true>, dim3(grid), dim3(64 * kWaves), kRenderSmem, s, A);
    } else {
        PVDB_HIP(hipFuncSetAttribute((const void*)k_render<false>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRenderSmem));
        hipLaunchKernelGGL(k_render<false>, dim3(grid), dim3(64 * kWaves), kRenderSmem, s, A);
    }
    PVDB_LAUNCHED("k_render");
    return PVDB_OK;
}

}  // extern "C"
