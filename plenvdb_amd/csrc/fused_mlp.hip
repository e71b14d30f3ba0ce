// Fused step, sample side: the view-dependent RGB MLP (39-128-128-3, ReLU,
// sigmoid; plenvdb/lib/dvgo.py:91-109,354-362) on MFMA, forward and backward,
// over groups of 32 consecutive records (two 16-sample MFMA tiles).
//
// A workgroup of 4 wavefronts takes a group; wave w owns hidden features
// [32w, 32w+32) of both hidden layers and exchanges activations with the
// other waves through LDS.  Activations are computed transposed (features x
// samples):
//   H0^T = relu(W0^T X^T + b0),  H1^T = relu(W1^T H0^T + b1),  L^T = W2^T H1^T + b2
// with v_mfma_f32_16x16x32_bf16 (samples are the N dimension).  The forward
// weight slices stay in registers; the backward weights sit in LDS.  The
// backward pass recomputes the forward, forms dL/dlogit from the per-ray loss
// gradient, runs g_H1^T = W2 g_L^T, g_H0^T = W1 g_H1^T (masked by the ReLUs,
// kept as bit masks), g_X^T = W0[0:12] g_H0^T, and accumulates dW over the
// group's 32 samples (K = 32) into per-wave accumulator tiles written once per
// workgroup.  The same pass evaluates the Alphas2Weights / raw2alpha adjoints
// (render_utils_kernel.cu:653-676, 506-517) and writes each record's feature
// and sigma gradients; k_scatter_agg (or k_scatter) then adds them to the
// grids with the "leaf must exist" rule (densityvdb.cu:16-27,
// colorvdb.cu:28-37).  The forward pass keeps the gathered BF16 features per
// record, so the backward reads 32 bytes per record instead of 8 gathers.
// Each group's records, features and ray data are loaded into registers while
// the previous group computes.
#include <type_traits>

#include "fused_common.h"
#include "gather.h"
#include "plenvdb_hip.h"

namespace pvdb {
TopoView view_of(const pvdb_topology* t);
}

using namespace pvdb;

namespace {

constexpr int GS = 32;       // records per group
constexpr int NT = GS / 16;  // 16-sample MFMA tiles per group
constexpr int XM = 72;       // sample-major stride of X (64 + 8)
constexpr int SM = 136;      // sample-major stride of 128-wide activations (128 + 8)
constexpr int FM = 40;       // feature-major stride (32 samples + 8; rows stay 16-byte aligned)
constexpr int GM = 40;       // sample-major stride of dL/dlogit (32 + 8)

__device__ __forceinline__ bf16x8 ld_frag(const uint16_t* p) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ bf16x8 as_bf8(const uint4& v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ void st4(uint16_t* p, float a, float b, float c, float d) {
    uint2 v;
    v.x = (uint32_t)f2bf_bits(a) | ((uint32_t)f2bf_bits(b) << 16);
    v.y = (uint32_t)f2bf_bits(c) | ((uint32_t)f2bf_bits(d) << 16);
    *reinterpret_cast<uint2*>(p) = v;
}

struct RecInfo {
    float x, y, z, e;
    float alpha, T, w;
    uint32_t bits;
};

struct alignas(16) SmemFwd {
    uint16_t Xs[GS * XM];
    uint16_t H0s[GS * SM];
    uint16_t H1s[GS * SM];
    RecInfo ri[GS];
};

// Backward weights as MFMA A fragments, [tile][k-step][lane] x 16 bytes:
//   W2  [8 row tiles of 16 hidden][1][64]   (K = 3 logits, padded to 32)
//   W1  [8 row tiles][4][64]                (K = 128 hidden outputs)
//   W0  [1 (12 feature rows)][4 k-steps of 32 hidden][64]
constexpr int kBw2Frags = 8 * 64, kBw1Frags = 8 * 4 * 64, kBw0Frags = 4 * 64;

struct alignas(16) SmemBwd {
    uint4 Bw2[kBw2Frags];
    uint4 Bw1[kBw1Frags];
    uint4 Bw0[kBw0Frags];
    uint16_t Xs[GS * XM];
    uint16_t Xf[64 * FM];
    uint16_t H0s[GS * SM];
    uint16_t H0f[128 * FM];
    uint16_t H1s[GS * SM];
    uint16_t H1f[128 * FM];
    uint16_t GLs[GS * GM];
    uint16_t GLf[16 * FM];
    uint16_t G1s[GS * SM];
    uint16_t G1f[128 * FM];
    uint16_t G0s[GS * SM];
    uint16_t G0f[128 * FM];
    float GXp[4][16][GS];
    float gsig[GS];
    float rgb[GS][4];
    RecInfo ri[GS];
    float4 rayg[GS];
    float4 ctot[GS];
    float4 cinc[GS];
    float tlast[GS];
};

__device__ __forceinline__ float sigmoidf_(float x) { return 1.f / (1.f + expf(-x)); }

// Register prefetch of one group's per-record inputs (thread = record).
struct Prefetch {
    float4 xyze, atwr;          // stage A: the record
    float4 rayg, ctot, cinc;    // stage B (backward): ray data, prefix
    float tlast;
    uint4 f0, f1;               // stage B (backward): kept features
};

template <bool BWD, bool BF16COL>
__global__ void __launch_bounds__(256) k_mlp(StepParams P, TopoView topo, const void* color,
                                             MlpPacked W, RayBuffers R, Records rec,
                                             float* __restrict__ dw_partial) {
    using Smem = std::conditional_t<BWD, SmemBwd, SmemFwd>;
    __shared__ Smem sm;
    const int tid = threadIdx.x;
    const int wv = tid >> 6, lane = tid & 63, r16 = lane & 15, g = lane >> 4;
    const int f0 = 32 * wv;  // first hidden feature owned by this wave

    // ---- forward weight slices -> registers (bf16 fragments) ----
    bf16x8 A0[2][2], A1[2][4], A2[4];
    float b0v[2][4], b1v[2][4], b2v[3];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        const int row = f0 + 16 * mt + r16;
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) A0[mt][kt] = ld_frag(W.w0t + row * 64 + 32 * kt + 8 * g);
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) A1[mt][kt] = ld_frag(W.w1t + row * 128 + 32 * kt + 8 * g);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            b0v[mt][i] = W.b0[f0 + 16 * mt + 4 * g + i];
            b1v[mt][i] = W.b1[f0 + 16 * mt + 4 * g + i];
        }
    }
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) A2[kt] = ld_frag(W.w2t + r16 * 128 + 32 * kt + 8 * g);
    b2v[0] = W.b2[0];
    b2v[1] = W.b2[1];
    b2v[2] = W.b2[2];

    // ---- backward weights -> LDS, dW accumulators ----
    f32x4 dW0[4][2], dW1[8][2], dW2[2];
    float db1[2][4], db2[3];
    if constexpr (BWD) {
        for (int f = tid; f < kBw2Frags + kBw1Frags + kBw0Frags; f += 256) {
            const int ln = f & 63, gg = ln >> 4, rr = ln & 15;
            uint4 v;
            if (f < kBw2Frags) {  // W2 [128 in][32 out (3 used)]
                const int t = f >> 6;
                v = *reinterpret_cast<const uint4*>(W.w2 + (16 * t + rr) * 32 + 8 * gg);
                sm.Bw2[f] = v;
            } else if (f < kBw2Frags + kBw1Frags) {  // W1 [128 in][128 out]
                const int q = f - kBw2Frags, t = q >> 8, kt = (q >> 6) & 3;
                v = *reinterpret_cast<const uint4*>(W.w1 + (16 * t + rr) * 128 + 32 * kt + 8 * gg);
                sm.Bw1[q] = v;
            } else {  // W0 rows 0..15 [16][128]
                const int q = f - kBw2Frags - kBw1Frags, kt = q >> 6;
                v = *reinterpret_cast<const uint4*>(W.w0 + rr * 128 + 32 * kt + 8 * gg);
                sm.Bw0[q] = v;
            }
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) dW0[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) dW1[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        dW2[0] = dW2[1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int i = 0; i < 4; ++i) db1[a][i] = 0.f;
        db2[0] = db2[1] = db2[2] = 0.f;
    }

    const int64_t total = min((int64_t)R.total[0], P.capacity);
    const int64_t ngroups = (total + GS - 1) / GS;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);

    // ---- prefetch of the first group ----
    Prefetch pf;
    pf.xyze = pf.atwr = pf.rayg = pf.ctot = pf.cinc = z4;
    pf.tlast = 0.f;
    pf.f0 = pf.f1 = make_uint4(0u, 0u, 0u, 0u);
    auto stage_a = [&](int64_t grp) {
        const int64_t q = grp * GS + tid;
        pf.xyze = pf.atwr = z4;
        if (tid < GS && grp < ngroups && q < total) {
            pf.xyze = rec.xyze[q];
            pf.atwr = rec.atwr[q];
        }
    };
    auto stage_b = [&](int64_t grp) {  // backward only; needs stage A's atwr
        const int64_t q = grp * GS + tid;
        pf.rayg = pf.ctot = pf.cinc = z4;
        pf.tlast = 0.f;
        pf.f0 = pf.f1 = make_uint4(0u, 0u, 0u, 0u);
        if (tid < GS && grp < ngroups && q < total) {
            const uint32_t bits = __float_as_uint(pf.atwr.w);
            if (bits & (kAliveA | kAliveW)) {
                const uint32_t ray = bits & kRayMask;
                pf.rayg = R.ray_grad[ray];
                pf.ctot = R.ray_ctot[ray];
                pf.tlast = R.t_last[ray];
                pf.cinc = rec.cinc[q];
            }
            if (bits & kAliveW) {
                pf.f0 = rec.feat[2 * q];
                pf.f1 = rec.feat[2 * q + 1];
            }
        }
    };
    stage_a(blockIdx.x);
    if constexpr (BWD) stage_b(blockIdx.x);

    for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {  // uniform bound
        const int64_t base = grp * GS;
        // ---- 1. commit the prefetched group to LDS; start loading the next one ----
        if (tid < GS) {
            RecInfo ri;
            ri.x = pf.xyze.x;
            ri.y = pf.xyze.y;
            ri.z = pf.xyze.z;
            ri.e = pf.xyze.w;
            ri.alpha = pf.atwr.x;
            ri.T = pf.atwr.y;
            ri.w = pf.atwr.z;
            ri.bits = __float_as_uint(pf.atwr.w);  // zero for records past the end
            sm.ri[tid] = ri;
            if constexpr (BWD) {
                sm.rayg[tid] = pf.rayg;
                sm.ctot[tid] = pf.ctot;
                sm.cinc[tid] = pf.cinc;
                sm.tlast[tid] = pf.tlast;
                const uint32_t w6[6] = {pf.f0.x, pf.f0.y, pf.f0.z, pf.f0.w, pf.f1.x, pf.f1.y};
                uint32_t* xs = reinterpret_cast<uint32_t*>(&sm.Xs[tid * XM]);
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    xs[c] = w6[c];
                    sm.Xf[(2 * c) * FM + tid] = (uint16_t)(w6[c] & 0xffffu);
                    sm.Xf[(2 * c + 1) * FM + tid] = (uint16_t)(w6[c] >> 16);
                }
            }
        }
        const int64_t next = grp + gridDim.x;
        stage_a(next);
        __syncthreads();
        // ---- 2. X = [colour features (12) | PE (27) | 1 | 0...] ----
        {
            const int s = tid >> 3, part = tid & 7;  // 8 threads per record
            const RecInfo ri = sm.ri[s];
            const bool alive = (ri.bits & kAliveW) != 0u;
            if constexpr (!BWD) {
                // colour gather: thread `part` takes corner `part`
                float v[12];
#pragma unroll
                for (int c = 0; c < 12; ++c) v[c] = 0.f;
                if (alive) {
                    const Cell cell = make_cell(ri.x, ri.y, ri.z);
                    const int dx = (kCx >> part) & 1, dy = (kCy >> part) & 1, dz = (kCz >> part) & 1;
                    const float a = dx ? cell.u : 1.f - cell.u;
                    const float b = dy ? cell.v : 1.f - cell.v;
                    const float d = dz ? cell.w : 1.f - cell.w;
                    const float sc = (a * b) * d;
                    const int64_t slot = slot_of(topo, cell.i0 + dx, cell.j0 + dy, cell.k0 + dz);
                    if (slot >= 0) {
                        if (BF16COL) {
                            const uint16_t* src =
                                reinterpret_cast<const uint16_t*>(color) + slot * 12;
                            const uint2 p0 = *reinterpret_cast<const uint2*>(src);
                            const uint2 p1 = *reinterpret_cast<const uint2*>(src + 4);
                            const uint2 p2 = *reinterpret_cast<const uint2*>(src + 8);
                            const uint32_t w6[6] = {p0.x, p0.y, p1.x, p1.y, p2.x, p2.y};
#pragma unroll
                            for (int c = 0; c < 6; ++c) {
                                v[2 * c] = bf_bits2f((uint16_t)(w6[c] & 0xffffu)) * sc;
                                v[2 * c + 1] = bf_bits2f((uint16_t)(w6[c] >> 16)) * sc;
                            }
                        } else {
                            const float4* src = reinterpret_cast<const float4*>(
                                reinterpret_cast<const float*>(color) + slot * 12);
                            const float4 p0 = src[0], p1 = src[1], p2 = src[2];
                            v[0] = p0.x * sc; v[1] = p0.y * sc; v[2] = p0.z * sc; v[3] = p0.w * sc;
                            v[4] = p1.x * sc; v[5] = p1.y * sc; v[6] = p1.z * sc; v[7] = p1.w * sc;
                            v[8] = p2.x * sc; v[9] = p2.y * sc; v[10] = p2.z * sc; v[11] = p2.w * sc;
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < 12; ++c) {
                    v[c] += __shfl_xor(v[c], 1, 64);
                    v[c] += __shfl_xor(v[c], 2, 64);
                    v[c] += __shfl_xor(v[c], 4, 64);
                }
                if (part == 0) {
                    uint32_t w6[6];
#pragma unroll
                    for (int c = 0; c < 6; ++c)
                        w6[c] = (uint32_t)f2bf_bits(v[2 * c]) |
                                ((uint32_t)f2bf_bits(v[2 * c + 1]) << 16);
                    uint32_t* xs = reinterpret_cast<uint32_t*>(&sm.Xs[s * XM]);
#pragma unroll
                    for (int c = 0; c < 6; ++c) xs[c] = w6[c];
                    const int64_t qq = base + s;
                    if (qq < total && alive) {
                        rec.feat[2 * qq] = make_uint4(w6[0], w6[1], w6[2], w6[3]);
                        rec.feat[2 * qq + 1] = make_uint4(w6[4], w6[5], 0u, 0u);
                    }
                }
            }
            // positional encoding of the ray's view direction (dvgo.py:354-358) and the
            // bias input; columns 40..63 are zero
            float vd[3] = {0.f, 0.f, 0.f};
            if (alive) {
                const uint32_t ray = ri.bits & kRayMask;
                vd[0] = P.viewdirs[3 * ray];
                vd[1] = P.viewdirs[3 * ray + 1];
                vd[2] = P.viewdirs[3 * ray + 2];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int p = part * 4 + k;  // 0..27: PE 0..26, then the bias column
                if (p < 28) {
                    float val = 0.f;
                    if (alive) {
                        if (p < 3) {
                            val = vd[p];
                        } else if (p < 27) {
                            const int idx = p - 3, axis = (idx % 12) / 4, fr = idx % 4;
                            const float arg = vd[axis] * (float)(1 << fr);
                            val = idx < 12 ? sinf(arg) : cosf(arg);
                        } else {
                            val = 1.f;
                        }
                    }
                    const uint16_t h = f2bf_bits(val);
                    sm.Xs[s * XM + 12 + p] = h;
                    if constexpr (BWD) sm.Xf[(12 + p) * FM + s] = h;
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int col = 40 + part * 3 + k;
                sm.Xs[s * XM + col] = 0;
                if constexpr (BWD) sm.Xf[col * FM + s] = 0;
            }
        }
        __syncthreads();
        // ---- 3. H0^T slice = relu(W0^T X^T + b0) ----
        uint32_t m0 = 0u, m1 = 0u;  // ReLU masks: bit (mt*NT + nt)*4 + i
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            const int fr = f0 + 16 * mt + 4 * g;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kt = 0; kt < 2; ++kt)
                    acc = mfma16x16x32(A0[mt][kt],
                                       ld_frag(&sm.Xs[(16 * nt + r16) * XM + 32 * kt + 8 * g]), acc);
                float h[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    h[i] = fmaxf(acc[i] + b0v[mt][i], 0.f);
                    if (h[i] > 0.f) m0 |= 1u << ((mt * NT + nt) * 4 + i);
                }
                st4(&sm.H0s[(16 * nt + r16) * SM + fr], h[0], h[1], h[2], h[3]);
                if constexpr (BWD) {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        sm.H0f[(fr + i) * FM + 16 * nt + r16] = f2bf_bits(h[i]);
                }
            }
        }
        __syncthreads();
        // ---- 4. H1^T slice = relu(W1^T H0^T + b1) ----
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            const int fr = f0 + 16 * mt + 4 * g;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kt = 0; kt < 4; ++kt)
                    acc = mfma16x16x32(A1[mt][kt],
                                       ld_frag(&sm.H0s[(16 * nt + r16) * SM + 32 * kt + 8 * g]), acc);
                float h[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    h[i] = fmaxf(acc[i] + b1v[mt][i], 0.f);
                    if (h[i] > 0.f) m1 |= 1u << ((mt * NT + nt) * 4 + i);
                }
                st4(&sm.H1s[(16 * nt + r16) * SM + fr], h[0], h[1], h[2], h[3]);
                if constexpr (BWD) {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        sm.H1f[(fr + i) * FM + 16 * nt + r16] = f2bf_bits(h[i]);
                }
            }
        }
        __syncthreads();
        // ---- 5. logits (wave nt takes tile nt): L^T = W2^T H1^T + b2, rgb = sigmoid ----
        if (wv < NT) {
            const int nt = wv;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt)
                acc = mfma16x16x32(A2[kt], ld_frag(&sm.H1s[(16 * nt + r16) * SM + 32 * kt + 8 * g]),
                                   acc);
            if (g == 0) {
                const int s = 16 * nt + r16;
                const float c0 = sigmoidf_(acc[0] + b2v[0]);
                const float c1 = sigmoidf_(acc[1] + b2v[1]);
                const float c2 = sigmoidf_(acc[2] + b2v[2]);
                if constexpr (BWD) {
                    sm.rgb[s][0] = c0;
                    sm.rgb[s][1] = c1;
                    sm.rgb[s][2] = c2;
                } else {
                    const int64_t q = base + s;
                    if (q < total) {
                        const bool alive = (sm.ri[s].bits & kAliveW) != 0u;
                        rec.rgb[q] = alive ? make_float4(c0, c1, c2, 0.f) : z4;
                    }
                }
            }
        }
        if constexpr (!BWD) {
            __syncthreads();
            continue;
        } else {
            // the next group's stage-A loads have landed by now: start stage B
            stage_b(next);
            __syncthreads();
            // ---- 6. per-sample adjoints: dL/dlogit, alpha, sigma ----
            if (tid < GS) {
                const RecInfo ri = sm.ri[tid];
                const float4 gr = sm.rayg[tid];
                float gl0 = 0.f, gl1 = 0.f, gl2 = 0.f, gw = 0.f;
                if (ri.bits & kAliveW) {
                    const float c0 = sm.rgb[tid][0], c1 = sm.rgb[tid][1], c2 = sm.rgb[tid][2];
                    gl0 = ri.w * gr.x * c0 * (1.f - c0);
                    gl1 = ri.w * gr.y * c1 * (1.f - c1);
                    gl2 = ri.w * gr.z * c2 * (1.f - c2);
                    gw = gr.x * c0 + gr.y * c1 + gr.z * c2;
                }
                float gs = 0.f;
                if (ri.bits & kAliveA) {
                    const float4 ct = sm.ctot[tid], ci = sm.cinc[tid];
                    const float back = gr.w * sm.tlast[tid] + gr.x * (ct.x - ci.x) +
                                       gr.y * (ct.y - ci.y) + gr.z * (ct.z - ci.z);
                    const float ga = gw * ri.T - back / ((1.f - ri.alpha) + 1e-10f);
                    gs = fminf(ri.e, 1e10f) * powf(1.f + ri.e, -P.interval - 1.f) * P.interval * ga;
                }
                sm.gsig[tid] = gs;
                db2[0] += gl0;
                db2[1] += gl1;
                db2[2] += gl2;
                uint32_t* gls = reinterpret_cast<uint32_t*>(&sm.GLs[tid * GM]);
                gls[0] = (uint32_t)f2bf_bits(gl0) | ((uint32_t)f2bf_bits(gl1) << 16);
                gls[1] = (uint32_t)f2bf_bits(gl2);
#pragma unroll
                for (int c = 2; c < 16; ++c) gls[c] = 0u;
#pragma unroll
                for (int o = 0; o < 16; ++o) {
                    const float v = o == 0 ? gl0 : o == 1 ? gl1 : o == 2 ? gl2 : 0.f;
                    sm.GLf[o * FM + tid] = f2bf_bits(v);
                }
            }
            __syncthreads();
            // ---- 7. g_H1^T slice = W2[slice] g_L^T, masked; dW2 slice ----
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                const bf16x8 a = as_bf8(sm.Bw2[(wv * 2 + mt) * 64 + lane]);
                const int fr = f0 + 16 * mt + 4 * g;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                    acc = mfma16x16x32(a, ld_frag(&sm.GLs[(16 * nt + r16) * GM + 8 * g]), acc);
                    float gv[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        gv[i] = (m1 >> ((mt * NT + nt) * 4 + i)) & 1u ? acc[i] : 0.f;
                        db1[mt][i] += gv[i];
                    }
                    st4(&sm.G1s[(16 * nt + r16) * SM + fr], gv[0], gv[1], gv[2], gv[3]);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        sm.G1f[(fr + i) * FM + 16 * nt + r16] = f2bf_bits(gv[i]);
                }
                // dW2[in = f0+16mt+row][out] += sum over the group's samples H1[s][in] gL[s][out]
                dW2[mt] = mfma16x16x32(ld_frag(&sm.H1f[(f0 + 16 * mt + r16) * FM + 8 * g]),
                                       ld_frag(&sm.GLf[r16 * FM + 8 * g]), dW2[mt]);
            }
            __syncthreads();
            // ---- 8. g_H0^T slice = W1[slice] g_H1^T, masked; dW1 columns of this wave ----
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                const int fr = f0 + 16 * mt + 4 * g;
                f32x4 acc[NT];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kt = 0; kt < 4; ++kt) {
                    const bf16x8 a = as_bf8(sm.Bw1[((wv * 2 + mt) * 4 + kt) * 64 + lane]);
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[nt] = mfma16x16x32(
                            a, ld_frag(&sm.G1s[(16 * nt + r16) * SM + 32 * kt + 8 * g]), acc[nt]);
                }
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    float gv[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        gv[i] = (m0 >> ((mt * NT + nt) * 4 + i)) & 1u ? acc[nt][i] : 0.f;
                    st4(&sm.G0s[(16 * nt + r16) * SM + fr], gv[0], gv[1], gv[2], gv[3]);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        sm.G0f[(fr + i) * FM + 16 * nt + r16] = f2bf_bits(gv[i]);
                }
            }
#pragma unroll
            for (int mi = 0; mi < 8; ++mi)
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
                    dW1[mi][nt] = mfma16x16x32(ld_frag(&sm.H0f[(16 * mi + r16) * FM + 8 * g]),
                                               ld_frag(&sm.G1f[(f0 + 16 * nt + r16) * FM + 8 * g]),
                                               dW1[mi][nt]);
            __syncthreads();
            // ---- 9. dW0 columns of this wave; partial g_X^T over this wave's K slice ----
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
                    dW0[mi][nt] = mfma16x16x32(ld_frag(&sm.Xf[(16 * mi + r16) * FM + 8 * g]),
                                               ld_frag(&sm.G0f[(f0 + 16 * nt + r16) * FM + 8 * g]),
                                               dW0[mi][nt]);
            {
                const bf16x8 a = as_bf8(sm.Bw0[wv * 64 + lane]);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                    acc = mfma16x16x32(a, ld_frag(&sm.G0s[(16 * nt + r16) * SM + f0 + 8 * g]), acc);
#pragma unroll
                    for (int i = 0; i < 4; ++i) sm.GXp[wv][4 * g + i][16 * nt + r16] = acc[i];
                }
            }
            __syncthreads();
            // ---- 10. per-record gradients: dL/dfeature (12) and dL/dsigma ----
            if (tid < 4 * GS) {
                const int s = tid >> 2, part = tid & 3;  // 4 threads write a record's 4 float4
                const int64_t q = base + s;
                if (q < total) {
                    const RecInfo ri = sm.ri[s];
                    float4 v = z4;
                    if (part < 3) {
                        if (ri.bits & kAliveW) {
                            float g4[4];
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                const int c = 4 * part + k;
                                g4[k] = sm.GXp[0][c][s] + sm.GXp[1][c][s] + sm.GXp[2][c][s] +
                                        sm.GXp[3][c][s];
                            }
                            v = make_float4(g4[0], g4[1], g4[2], g4[3]);
                        }
                    } else {
                        v.x = sm.gsig[s];
                    }
                    rec.grad[4 * q + part] = v;
                }
            }
            __syncthreads();
        }
    }

    if (!BWD) return;
    // ---- flush this workgroup's dW partial ----
    float* out = dw_partial + (int64_t)blockIdx.x * kMlpParams;
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int in = 16 * mi + 4 * g + i, o = f0 + 16 * nt + r16;
                if (in < 39)
                    out[kW0 + in * 128 + o] = dW0[mi][nt][i];
                else if (in == 39)
                    out[kB0 + o] = dW0[mi][nt][i];
            }
#pragma unroll
    for (int mi = 0; mi < 8; ++mi)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                out[kW1 + (16 * mi + 4 * g + i) * 128 + f0 + 16 * nt + r16] = dW1[mi][nt][i];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (r16 < 3) out[kW2 + (f0 + 16 * mt + 4 * g + i) * 3 + r16] = dW2[mt][i];
            float v = db1[mt][i];
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            v += __shfl_xor(v, 8, 64);
            if (r16 == 0) out[kB1 + f0 + 16 * mt + 4 * g + i] = v;
        }
    if (wv == 0) {
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            float v = tid < GS ? db2[o] : 0.f;  // step 6 ran on threads 0..GS-1
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            v += __shfl_xor(v, 8, 64);
            v += __shfl_xor(v, 16, 64);
            if (tid == 0) out[kB2 + o] = v;
        }
    }
}

// Sum of the per-workgroup dW partials, accumulated into dw: blockIdx.y takes
// a chunk of kReduceChunk partials, so the reduction runs on enough
// workgroups to stream the partials at HBM rate.
constexpr int kReduceChunk = 16;
__global__ void k_dw_reduce(const float* __restrict__ part, int nparts, float* __restrict__ dw) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const int b0 = blockIdx.y * kReduceChunk, b1 = min(nparts, b0 + kReduceChunk);
    float s = 0.f;
    for (int b = b0; b < b1; ++b) s += part[(int64_t)b * kMlpParams + p];
    atomicAdd(dw + p, s);
}

// Gradients into the grids: one thread per (record, corner), with the
// "leaf must exist" rule (densityvdb.cu:16-27, colorvdb.cu:28-37); the density
// term keeps the reference's ((g*a)*b)*c order.  Touched leaves are flagged
// for the sparse Adam pass.
__global__ void __launch_bounds__(256) k_scatter(TopoView topo, RayBuffers R, Records rec,
                                                 int64_t capacity, float* __restrict__ grad_density,
                                                 float* __restrict__ grad_color,
                                                 uint8_t* __restrict__ leaf_touched) {
    const int64_t total = min((int64_t)R.total[0], capacity);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total * 8;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = t >> 3;
        const int corner = (int)(t & 7);
        const uint32_t bits = __float_as_uint(rec.atwr[q].w);
        if (!(bits & (kAliveA | kAliveW))) continue;
        const float4 p = rec.xyze[q];
        const Cell cell = make_cell(p.x, p.y, p.z);
        const int dx = (kCx >> corner) & 1, dy = (kCy >> corner) & 1, dz = (kCz >> corner) & 1;
        const int i = cell.i0 + dx, j = cell.j0 + dy, k = cell.k0 + dz;
        const int32_t leaf = leaf_of(topo, i, j, k);
        if (leaf < 0) continue;
        const float a = dx ? cell.u : 1.f - cell.u;
        const float b = dy ? cell.v : 1.f - cell.v;
        const float d = dz ? cell.w : 1.f - cell.w;
        const int64_t slot = (int64_t)leaf * 512 + voxel_offset(i, j, k);
        leaf_touched[leaf] = 1;
        const float4* g = rec.grad + 4 * q;
        if (bits & kAliveA) atomicAdd(grad_density + slot, ((g[3].x * a) * b) * d);
        if (bits & kAliveW) {
            const float sc = (a * b) * d;
            float* dst = grad_color + slot * 12;
#pragma unroll
            for (int part = 0; part < 3; ++part) {
                const float4 v = g[part];
                atomicAdd(dst + 4 * part + 0, v.x * sc);
                atomicAdd(dst + 4 * part + 1, v.y * sc);
                atomicAdd(dst + 4 * part + 2, v.z * sc);
                atomicAdd(dst + 4 * part + 3, v.w * sc);
            }
        }
    }
}

// fp32 master parameters (kMlpParams layout) -> packed bf16 MFMA operands.
__global__ void k_pack_mlp(const float* __restrict__ prm, uint16_t* __restrict__ w0t,
                           uint16_t* __restrict__ w1t, uint16_t* __restrict__ w2t,
                           uint16_t* __restrict__ w2, uint16_t* __restrict__ w1,
                           uint16_t* __restrict__ w0) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    // w1t / w1: 128 x 128
    if (t < 128 * 128) {
        const int a = t >> 7, b = t & 127;
        w1t[a * 128 + b] = f2bf_bits(prm[kW1 + b * 128 + a]);  // W1^T[out a][in b]
        w1[a * 128 + b] = f2bf_bits(prm[kW1 + a * 128 + b]);   // W1[in a][out b]
    }
    if (t < 128 * 64) {  // w0t [128 out][64 in]
        const int o = t >> 6, i = t & 63;
        w0t[o * 64 + i] = i < 39 ? f2bf_bits(prm[kW0 + i * 128 + o]) : (uint16_t)0;
    }
    if (t < 16 * 128) {  // w2t [16 out][128 in], w0 [16 in][128 out]
        const int o = t >> 7, i = t & 127;
        w2t[o * 128 + i] = o < 3 ? f2bf_bits(prm[kW2 + i * 3 + o]) : (uint16_t)0;
        w0[o * 128 + i] = o < 12 ? f2bf_bits(prm[kW0 + o * 128 + i]) : (uint16_t)0;
    }
    if (t < 128 * 32) {  // w2 [128 in][32 out]
        const int i = t >> 5, o = t & 31;
        w2[i * 32 + o] = o < 3 ? f2bf_bits(prm[kW2 + i * 3 + o]) : (uint16_t)0;
    }
}

}  // namespace

namespace pvdb {



int launch_mlp(bool bwd, bool bf16col, int grid, const StepParams& P, const TopoView& t,
               const void* color, const MlpPacked& W, const RayBuffers& R, const Records& rec,
               float* dw_partial, hipStream_t s) {
    if (bwd) {
        if (bf16col)
            hipLaunchKernelGGL((k_mlp<true, true>), dim3(grid), dim3(256), 0, s, P, t, color, W, R,
                               rec, dw_partial);
        else
            hipLaunchKernelGGL((k_mlp<true, false>), dim3(grid), dim3(256), 0, s, P, t, color, W, R,
                               rec, dw_partial);
    } else {
        if (bf16col)
            hipLaunchKernelGGL((k_mlp<false, true>), dim3(grid), dim3(256), 0, s, P, t, color, W, R,
                               rec, dw_partial);
        else
            hipLaunchKernelGGL((k_mlp<false, false>), dim3(grid), dim3(256), 0, s, P, t, color, W,
                               R, rec, dw_partial);
    }
    PVDB_LAUNCHED("k_mlp");
    return PVDB_OK;
}

// Same result as k_scatter with far fewer global atomics: a workgroup takes
// kAggRecords consecutive records (a few rays; consecutive samples of a ray
// share most of their corners), sums their (slot, gradient) contributions in
// an LDS hash table keyed by the 32-bit slot, then adds each distinct slot's
// 13 sums to the grids once.  The flush gives 16 lanes to a slot, so one
// atomic instruction covers four slots' contiguous 12-float colour rows.
constexpr int kAggRecords = 128;
constexpr int kAggSlots = 8 * kAggRecords;  // every contribution fits: probing ends
constexpr uint32_t kEmptyKey = 0xffffffffu;
__global__ void __launch_bounds__(256) k_scatter_agg(TopoView topo, RayBuffers R, Records rec,
                                                     int64_t capacity,
                                                     float* __restrict__ grad_density,
                                                     float* __restrict__ grad_color,
                                                     uint8_t* __restrict__ leaf_touched) {
    __shared__ uint32_t key[kAggSlots];
    __shared__ float val[13][kAggSlots];
    const int tid = threadIdx.x;
    const int64_t total = min((int64_t)R.total[0], capacity);
    const int64_t nchunk = (total + kAggRecords - 1) / kAggRecords;
    for (int64_t chunk = blockIdx.x; chunk < nchunk; chunk += gridDim.x) {
        for (int e = tid; e < kAggSlots; e += 256) {
            key[e] = kEmptyKey;
#pragma unroll
            for (int c = 0; c < 13; ++c) val[c][e] = 0.f;
        }
        __syncthreads();
        for (int p = tid; p < 8 * kAggRecords; p += 256) {
            const int64_t q = chunk * kAggRecords + (p >> 3);
            if (q >= total) continue;
            const int corner = p & 7;
            const uint32_t bits = __float_as_uint(rec.atwr[q].w);
            if (!(bits & (kAliveA | kAliveW))) continue;
            const float4 pt = rec.xyze[q];
            const Cell cell = make_cell(pt.x, pt.y, pt.z);
            const int dx = (kCx >> corner) & 1, dy = (kCy >> corner) & 1, dz = (kCz >> corner) & 1;
            const int i = cell.i0 + dx, j = cell.j0 + dy, k = cell.k0 + dz;
            const int32_t leaf = leaf_of(topo, i, j, k);
            if (leaf < 0) continue;
            const uint32_t slot = (uint32_t)leaf * 512u + (uint32_t)voxel_offset(i, j, k);
            const float a = dx ? cell.u : 1.f - cell.u;
            const float b = dy ? cell.v : 1.f - cell.v;
            const float d = dz ? cell.w : 1.f - cell.w;
            uint32_t h = (slot * 2654435761u) >> (32 - 10);  // kAggSlots == 1 << 10
            while (true) {
                const uint32_t prev = atomicCAS(&key[h], kEmptyKey, slot);
                if (prev == kEmptyKey || prev == slot) break;
                h = (h + 1) & (kAggSlots - 1);
            }
            const float4* g = rec.grad + 4 * q;
            if (bits & kAliveA) atomicAdd(&val[12][h], ((g[3].x * a) * b) * d);
            if (bits & kAliveW) {
                const float sc = (a * b) * d;
#pragma unroll
                for (int part = 0; part < 3; ++part) {
                    const float4 v = g[part];
                    atomicAdd(&val[4 * part + 0][h], v.x * sc);
                    atomicAdd(&val[4 * part + 1][h], v.y * sc);
                    atomicAdd(&val[4 * part + 2][h], v.z * sc);
                    atomicAdd(&val[4 * part + 3][h], v.w * sc);
                }
            }
        }
        __syncthreads();
        const int c = tid & 15;
        for (int e = tid >> 4; e < kAggSlots; e += 16) {
            const uint32_t slot = key[e];
            if (slot == kEmptyKey) continue;
            if (c == 13) leaf_touched[slot >> 9] = 1;
            if (c > 12) continue;
            const float v = val[c][e];
            if (v == 0.f) continue;  // adding zero changes nothing
            if (c < 12)
                atomicAdd(grad_color + (int64_t)slot * 12 + c, v);
            else
                atomicAdd(grad_density + slot, v);
        }
        __syncthreads();
    }
}

int launch_dw_reduce(const float* part, int nparts, float* dw, hipStream_t s) {
    const unsigned chunks = (unsigned)((nparts + kReduceChunk - 1) / kReduceChunk);
    hipLaunchKernelGGL(k_dw_reduce, dim3(pvdb_blocks(kMlpParams, 256), chunks), dim3(256), 0, s,
                       part, nparts, dw);
    PVDB_LAUNCHED("k_dw_reduce");
    return PVDB_OK;
}

int launch_scatter(const TopoView& t, const RayBuffers& R, const Records& rec, int64_t capacity,
                   float* grad_density, float* grad_color, uint8_t* leaf_touched, hipStream_t s) {
    if (t.nleaf < (int64_t(1) << 23) - 1) {  // every slot < 2^32 - 1, the empty key
        const int64_t chunks = (capacity + kAggRecords - 1) / kAggRecords;
        const unsigned grid = (unsigned)(chunks < 2048 ? (chunks < 1 ? 1 : chunks) : 2048);
        hipLaunchKernelGGL(k_scatter_agg, dim3(grid), dim3(256), 0, s, t, R, rec, capacity,
                           grad_density, grad_color, leaf_touched);
        PVDB_LAUNCHED("k_scatter_agg");
        return PVDB_OK;
    }
    // enough threads for a few (record, corner) pairs each at the largest sizes
    const int64_t want = (capacity * 8 + 255) / 256;
    const unsigned grid = (unsigned)(want < 16384 ? (want < 1 ? 1 : want) : 16384);
    hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, s, t, R, rec, capacity, grad_density,
                       grad_color, leaf_touched);
    PVDB_LAUNCHED("k_scatter");
    return PVDB_OK;
}

int launch_pack_mlp(const float* prm, uint16_t* w0t, uint16_t* w1t, uint16_t* w2t, uint16_t* w2,
                    uint16_t* w1, uint16_t* w0, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_mlp, dim3(pvdb_blocks(128 * 128, 256)), dim3(256), 0, s, prm, w0t,
                       w1t, w2t, w2, w1, w0);
    PVDB_LAUNCHED("k_pack_mlp");
    return PVDB_OK;
}

}  // namespace pvdb
