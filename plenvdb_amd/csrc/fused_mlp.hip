// Fused step, sample side: the view-dependent RGB MLP (39-128-128-3, ReLU,
// sigmoid; plenvdb/lib/dvgo.py:91-109,354-362) on MFMA, forward and backward,
// over groups of 32 consecutive records (two 16-sample MFMA tiles).
//
// Activations are computed transposed (features x samples):
//   H0^T = relu(W0^T X^T + b0),  H1^T = relu(W1^T H0^T + b1),  L^T = W2^T H1^T + b2
// with v_mfma_f32_16x16x32_bf16 (samples are the N dimension); each wavefront
// owns a slice of the hidden features and exchanges activations with the
// others through LDS.  Weights are held in bf16; gradients and the last
// hidden layer travel as (hi, lo) bf16 pairs, so each product with a weight is
// two MFMAs, which keeps about 16 significant bits where plain bf16 keeps 8
// (plain bf16 inputs put the colour-feature gradients several percent off an
// FP64 evaluation).  The forward's first two products take (hi, mid, lo)
// triples (three MFMAs, fp32's 24 bits): their pre-activations decide the ReLU
// masks the backward applies, and a mask bit that rounds the other way moves a
// sample's feature gradient by a whole term.  dW products (activation x
// gradient) take the hi parts only.
//
// k_mlp_fwd gathers the colour features, keeps them per record as (hi, lo)
// bf16 (so the backward reads 48 bytes per record instead of redoing 8
// gathers) and writes rgb.  k_mlp_bwd recomputes H0 and H1 (for the ReLU
// masks and the dW operands), takes each sample's dL/dlogit and dL/dsigma from
// the composite pass, runs g_H1^T = W2 g_L^T, g_H0^T = W1 g_H1^T (masked by
// the ReLUs, kept as bit masks), g_X^T = W0[0:12] g_H0^T, accumulates dW over
// each group's 32 samples (K = 32) in registers (one partial per workgroup),
// and writes each record's feature and sigma gradients; the scatter
// (fused_scatter.hip) then adds them to the grids with the "leaf must exist"
// rule (densityvdb.cu:16-27, colorvdb.cu:28-37).  Both kernels load a group's
// records into registers while the previous group computes.

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
constexpr int GM = 24;       // sample-major stride of dL/dlogit (16 + 8)
constexpr int FM = 40;       // feature-major stride (32 samples + 8; rows 16-byte aligned)
constexpr int XF_ROWS = 48;  // feature-major X rows: inputs 0..39 are used

__device__ __forceinline__ bf16x8 ld_frag(const uint16_t* p) {  // 16-byte aligned
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ s16x4 ld_frag4(const uint16_t* p) {
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    return __builtin_bit_cast(s16x4, v);
}

__device__ __forceinline__ bf16x8 as_bf8(const uint4& v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ void split_bf(float x, uint16_t& h, uint16_t& l) {
    uint32_t hh, ll;
    split2(x, 0.f, hh, ll);
    h = (uint16_t)hh;
    l = (uint16_t)ll;
}

// DPP permutation within rows (quad permutes, the half-row mirror), all lanes
// active: every lane reads a valid source, so the move needs no old value
// (update_dpp's would be materialised with a v_mov).
template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
}

// A lane holds four consecutive features (rows) of one sample; the four lanes of
// a quad hold four consecutive samples.  Returns row (lane & 3) of the quad's
// four samples, as four bf16 in two dwords: two quad swaps instead of four
// 2-byte stores per lane.
__device__ __forceinline__ uint2 quad_transpose(uint2 h, int q) {
    const uint32_t v0 = h.x & 0xffffu, v1 = h.x >> 16, v2 = h.y & 0xffffu, v3 = h.y >> 16;
    // exchange 1 (lanes q, q^1): even lanes gather rows 0/2, odd lanes rows 1/3, of the pair
    const uint32_t send = (q & 1) ? (v0 | (v2 << 16)) : (v1 | (v3 << 16));
    const uint32_t recv = dpp<0xB1>(send);  // quad_perm [1,0,3,2]
    const uint32_t A = (q & 1) ? ((recv & 0xffffu) | (v1 << 16)) : (v0 | (recv << 16));
    const uint32_t B = (q & 1) ? ((recv >> 16) | (v3 << 16)) : (v2 | (recv & 0xffff0000u));
    // exchange 2 (lanes q, q^2): rows (q & 1) and (q & 1) + 2 of the two pairs
    const uint32_t send2 = (q & 2) ? A : B;
    const uint32_t recv2 = dpp<0x4E>(send2);  // quad_perm [2,3,0,1]
    return (q & 2) ? make_uint2(recv2, B) : make_uint2(A, recv2);
}

// Four consecutive features fr..fr+3 of sample `col`: sample-major (hi, lo)
// pairs into sh/sl, and (FMC) the hi parts feature-major into fm.
template <bool FMC>
__device__ __forceinline__ void st_act(uint16_t* sh, uint16_t* sl, uint16_t* fm, int col, int fr,
                                       int lane, const float* v) {
    uint2 hi, lo;
    split2(v[0], v[1], hi.x, lo.x);
    split2(v[2], v[3], hi.y, lo.y);
    *reinterpret_cast<uint2*>(&sh[col * SM + fr]) = hi;
    *reinterpret_cast<uint2*>(&sl[col * SM + fr]) = lo;
    if constexpr (FMC) {
        const int q = lane & 3;
        *reinterpret_cast<uint2*>(&fm[(fr + q) * FM + (col & ~3)]) = quad_transpose(hi, q);
    }
}

// The hi parts of four consecutive features fr..fr+3 of sample `col`:
// sample-major into sh and feature-major into fm.
__device__ __forceinline__ void st_hi_fm(uint16_t* sh, uint16_t* fm, int col, int fr, int lane,
                                         const float* v) {
    uint2 hi, lo;
    split2(v[0], v[1], hi.x, lo.x);
    split2(v[2], v[3], hi.y, lo.y);
    *reinterpret_cast<uint2*>(&sh[col * SM + fr]) = hi;
    const int q = lane & 3;
    *reinterpret_cast<uint2*>(&fm[(fr + q) * FM + (col & ~3)]) = quad_transpose(hi, q);
}

// The hi parts of four consecutive features fr..fr+3 of sample `col`,
// feature-major into fm (the feature-major half of st_act<true>).
__device__ __forceinline__ void st_fm(uint16_t* fm, int col, int fr, int lane, const float* v) {
    uint2 hi, lo;
    split2(v[0], v[1], hi.x, lo.x);
    split2(v[2], v[3], hi.y, lo.y);
    const int q = lane & 3;
    *reinterpret_cast<uint2*>(&fm[(fr + q) * FM + (col & ~3)]) = quad_transpose(hi, q);
}

// A wave's 32-bit slice of one sample's ReLU mask is written a byte per lane:
// lane (g, n) holds bit i of nib[mt] for feature 16 mt + 4g + i and stores
// byte g of sample n's word, nib[0] | nib[1] << 4, so the bit of feature
// 16 mt + 4g + i is bit 8g + 4mt + i (no exchange between the lanes).
__device__ __forceinline__ uint8_t relu_byte(uint32_t nib0, uint32_t nib1) {
    return (uint8_t)(nib0 | (nib1 << 4));
}

// a (bf16 weight) times a split activation
__device__ __forceinline__ f32x4 mfma_w(const bf16x8& a, const bf16x8& bh, const bf16x8& bl,
                                        f32x4 c) {
    c = mfma16x16x32(a, bh, c);
    return mfma16x16x32(a, bl, c);
}

// Four consecutive features fr..fr+3 of sample `col` as (hi, mid, lo) triples,
// sample-major.
__device__ __forceinline__ void st_act3(uint16_t* sh, uint16_t* sm_, uint16_t* sl, int col,
                                        int fr, const float* v) {
    uint2 hi, mi, lo;
    split3(v[0], v[1], hi.x, mi.x, lo.x);
    split3(v[2], v[3], hi.y, mi.y, lo.y);
    *reinterpret_cast<uint2*>(&sh[col * SM + fr]) = hi;
    *reinterpret_cast<uint2*>(&sm_[col * SM + fr]) = mi;
    *reinterpret_cast<uint2*>(&sl[col * SM + fr]) = lo;
}

// dW products (activation x gradient) use the hi parts only, and the backward
// recomputes the activations they take from hi parts: the sums over a step's
// samples average that rounding (a device check keeps every MLP gradient
// within 6e-4 of FP64), so the gradients' lo parts stay sample-major.

// x where bit k of m is set, else +0: the bit sign-extended by one v_bfe_i32
// and ANDed with x's bits (two instructions; a compare and a select take
// three, with the bit extraction)
__device__ __forceinline__ float keep_if(uint32_t m, int k, float x) {
    const uint32_t all = (uint32_t)__builtin_amdgcn_sbfe((int)m, k, 1);
    return __uint_as_float(__float_as_uint(x) & all);
}

// ---- forward -----------------------------------------------------------------
// A workgroup of 4 wavefronts takes a group of 32 records; wave w owns hidden
// features [32w, 32w + 32) of both hidden layers (up to three workgroups per
// CU).  X and H0 enter their products as (hi, mid, lo) triples, so the
// pre-activations, and with them the ReLU masks the backward applies, carry
// fp32 precision; H1 enters the logits as (hi, lo) pairs.  Per group, three
// phases with a barrier after each:
//   A  rgb of the previous group from its partial logits; X = [colour
//      features | PE | 1] from the gather and view directions loaded during
//      the previous group; the features are kept per record
//   B  H0, then the next group's topology lookups and view-direction loads
//   C  H1, then each wave's partial logits over its own 32 hidden features
//      (it reads back only the H1 it wrote: no workgroup barrier between),
//      then the next group's colour loads
// The records are staged two groups ahead (flags and positions double-buffered
// in LDS); W0^T and W1^T slices stay in registers, W2^T in LDS.  The ReLU
// masks of both hidden layers are stored per record for the backward.
struct RecInfo {
    float x, y, z;
    uint32_t bits;
};

struct alignas(16) SmemFwd {
    uint16_t W2t[16 * 128];  // W2^T [16 logits (3 used)][128 hidden]
    // X is read in phase B only and H1 in phase D only (a barrier between D
    // and the next group's A), so they share storage
    union {
        struct {
            uint16_t Xh[GS * XM], Xm[GS * XM], Xl[GS * XM];
        } x;
        struct {
            uint16_t A2h[GS * SM], A2l[GS * SM];  // H1
        } h1;
    } u;
    uint16_t A1h[GS * SM], A1m[GS * SM], A1l[GS * SM];  // H0
    RecInfo ri[2][GS];
    float4 plog[4][GS];  // each wave's partial logits (0..2) over its 32 hidden features
    uint32_t pbits[GS];  // the flags of the group whose partial logits these are
};

__device__ __forceinline__ float sigmoidf_(float x) { return 1.f / (1.f + expf(-x)); }

template <bool BF16COL>
__global__ void __launch_bounds__(256, 3) k_mlp_fwd(StepParams P, TopoView topo, const void* color,
                                                 MlpPacked W, RayBuffers R, Records rec) {
    constexpr int MT = 2;  // 16-feature tiles per wave
    using Raw = std::conditional_t<BF16COL, uint2, float4>;  // one third of a colour row
    __shared__ SmemFwd sm;
    const int tid = threadIdx.x;
    const int wv = tid >> 6, lane = tid & 63, r16 = lane & 15, g = lane >> 4;
    const int f0 = 32 * wv;
    const int pr = lane < GS / 4 ? wv * (GS / 4) + lane : -1;  // staging lanes
    const int gs = tid >> 3, gpart = tid & 7;                  // gather: record, corner
    const int ps = tid & (GS - 1), pp = tid >> 5;              // encoding: sample, part

    // this wave's W0^T and W1^T slices in registers; W2^T in LDS
    bf16x8 A0[MT][2], A1[MT][4];
    float b0v[MT][4], b1v[MT][4], b2v[3];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
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
    b2v[0] = W.b2[0];
    b2v[1] = W.b2[1];
    b2v[2] = W.b2[2];
    for (int e = tid; e < 16 * 128 / 8; e += 256)
        reinterpret_cast<uint4*>(sm.W2t)[e] = reinterpret_cast<const uint4*>(W.w2t)[e];

    const int64_t total = min((int64_t)R.total[0], P.capacity);
    const int64_t ngroups = (total + GS - 1) / GS;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    // staging registers: a group's records (two groups ahead of the one computed)
    float4 p_xyze = z4;
    float2 p_ab = make_float2(0.f, 0.f);
    auto stage = [&](int64_t grp) {
        const int64_t q = grp * GS + pr;
        p_xyze = z4;
        p_ab = make_float2(0.f, 0.f);
        if (pr >= 0 && grp < ngroups && q < total) {
            p_xyze = rec.xyze[q];
            p_ab = rec.ab[q];
        }
    };
    auto commit = [&](int b) {
        if (pr >= 0) {
            RecInfo ri;
            ri.x = p_xyze.x;
            ri.y = p_xyze.y;
            ri.z = p_xyze.z;
            ri.bits = __float_as_uint(p_ab.y);  // zero for records past the end
            sm.ri[b][pr] = ri;
        }
    };
    // gather prefetch of the next group: this thread's corner of record gs
    int64_t c_slot = -1;
    float c_sc = 0.f;
    Raw c_raw[3];
    uint2 pe_h = make_uint2(0u, 0u), pe_m = pe_h, pe_l = pe_h;  // sample ps's X columns 12 + 4pp ..
    auto lookup = [&](int b) {  // topology and view direction (after the records' barrier)
        const RecInfo ri = sm.ri[b][gs];
        c_slot = -1;
        c_sc = 0.f;
        if (ri.bits & kAliveW) {
            const Cell cell = make_cell(ri.x, ri.y, ri.z);
            const int dx = (kCx >> gpart) & 1, dy = (kCy >> gpart) & 1, dz = (kCz >> gpart) & 1;
            const float a = dx ? cell.u : 1.f - cell.u;
            const float bb = dy ? cell.v : 1.f - cell.v;
            const float d = dz ? cell.w : 1.f - cell.w;
            c_sc = (a * bb) * d;
            c_slot = slot_of(topo, cell.i0 + dx, cell.j0 + dy, cell.k0 + dz);
        }
        // the encoding columns of sample ps from its ray's row (the counting
        // march wrote them), zero for a sample that is not weight-kept
        const uint32_t pb = sm.ri[b][ps].bits;
        pe_h = pe_m = pe_l = make_uint2(0u, 0u);
        if ((pb & kAliveW) && pp < 7) {
            const uint2* row = reinterpret_cast<const uint2*>(R.pe + 96 * (int64_t)(pb & kRayMask));
            pe_h = row[pp];
            pe_m = row[8 + pp];
            pe_l = row[16 + pp];
        }
    };
    auto fetch = [&]() {  // the colour row of the looked-up corner
        if constexpr (BF16COL) {
            c_raw[0] = c_raw[1] = c_raw[2] = make_uint2(0u, 0u);
            if (c_slot >= 0) {
                const uint2* src = reinterpret_cast<const uint2*>(
                    reinterpret_cast<const uint16_t*>(color) + c_slot * 12);
                c_raw[0] = src[0];
                c_raw[1] = src[1];
                c_raw[2] = src[2];
            }
        } else {
            c_raw[0] = c_raw[1] = c_raw[2] = z4;
            if (c_slot >= 0) {
                const float4* src = reinterpret_cast<const float4*>(
                    reinterpret_cast<const float*>(color) + c_slot * 12);
                c_raw[0] = src[0];
                c_raw[1] = src[1];
                c_raw[2] = src[2];
            }
        }
    };

    stage(blockIdx.x);
    commit(0);
    stage((int64_t)blockIdx.x + gridDim.x);
    __syncthreads();
    lookup(0);
    fetch();
    int cb = 0;  // buffer of the current group's records
    // rgb of a group from the four waves' partial logits (summed in wave
    // order), written by the threads of its 32 samples
    auto finish_rgb = [&](int64_t pbase) {
        if (tid < GS) {
            const int64_t q = pbase + tid;
            if (q < total) {
                const float4 l0 = sm.plog[0][tid], l1 = sm.plog[1][tid], l2 = sm.plog[2][tid],
                             l3 = sm.plog[3][tid];
                const bool alive = (sm.pbits[tid] & kAliveW) != 0u;
                rec.rgb[q] = alive ? make_float4(sigmoidf_(((l0.x + l1.x) + l2.x) + l3.x + b2v[0]),
                                                 sigmoidf_(((l0.y + l1.y) + l2.y) + l3.y + b2v[1]),
                                                 sigmoidf_(((l0.z + l1.z) + l2.z) + l3.z + b2v[2]),
                                                 0.f)
                                   : z4;
            }
        }
    };
    int64_t pbase = -1;  // the group whose partial logits wait in plog (-1: none)

    for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {  // uniform bound
        const int64_t base = grp * GS;
        // ---- A. the previous group's rgb; X = [colour features (12) | PE (27) | 1 | 0...];
        //         stage the records ----
        if (pbase >= 0) finish_rgb(pbase);
        {
            // trilinear colour: corner gpart's row times its weight, summed over
            // the record's 8 threads (zero where the record or corner is absent)
            float v[12];
            if constexpr (BF16COL) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    v[4 * k + 0] = bf_lo(c_raw[k].x) * c_sc;
                    v[4 * k + 1] = bf_hi(c_raw[k].x) * c_sc;
                    v[4 * k + 2] = bf_lo(c_raw[k].y) * c_sc;
                    v[4 * k + 3] = bf_hi(c_raw[k].y) * c_sc;
                }
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    v[4 * k + 0] = c_raw[k].x * c_sc;
                    v[4 * k + 1] = c_raw[k].y * c_sc;
                    v[4 * k + 2] = c_raw[k].z * c_sc;
                    v[4 * k + 3] = c_raw[k].w * c_sc;
                }
            }
            // butterfly over the record's eight lanes on the DPP path (folded
            // into the adds): lanes l^1 and l^2 by quad permutes, then the
            // other quad of the eight by the half-row mirror (lane j of the
            // eight reads lane 7 - j, which holds the other quad's sum)
#pragma unroll
            for (int c = 0; c < 12; ++c)
                v[c] += __uint_as_float(dpp<0xB1>(__float_as_uint(v[c])));  // quad_perm [1,0,3,2]
#pragma unroll
            for (int c = 0; c < 12; ++c)
                v[c] += __uint_as_float(dpp<0x4E>(__float_as_uint(v[c])));  // quad_perm [2,3,0,1]
#pragma unroll
            for (int c = 0; c < 12; ++c)
                v[c] += __uint_as_float(dpp<0x141>(__float_as_uint(v[c])));  // row_half_mirror
            // the butterfly leaves the same twelve sums (bit for bit: each step
            // adds the same two operands on every lane) in all eight lanes of
            // the record, so lane gpart < 6 splits and stores channel pair gpart
            if (gpart < 6) {
                float a = v[0], b = v[1];
#pragma unroll
                for (int c = 1; c < 6; ++c) {
                    a = gpart == c ? v[2 * c] : a;
                    b = gpart == c ? v[2 * c + 1] : b;
                }
                uint32_t wh, wm, wl;
                split3(a, b, wh, wm, wl);
                reinterpret_cast<uint32_t*>(&sm.u.x.Xh[gs * XM])[gpart] = wh;
                reinterpret_cast<uint32_t*>(&sm.u.x.Xm[gs * XM])[gpart] = wm;
                reinterpret_cast<uint32_t*>(&sm.u.x.Xl[gs * XM])[gpart] = wl;
                // the backward's input row, features part (X columns 0..11, hi)
                const int64_t qq = base + gs;
                if (qq < total) reinterpret_cast<uint32_t*>(rec.xrow + 5 * qq)[gpart] = wh;
            }
        }
        {
            // positional encoding (dvgo.py:354-358) and the bias column: thread
            // (sample ps, part pp < 7) writes X columns 12 + 4pp .. +3, as (hi,
            // mid, lo) bf16 from its ray's encoding row
            const uint2 hi = pe_h, mi = pe_m, lo = pe_l;
            if (pp < 7) {
                *reinterpret_cast<uint2*>(&sm.u.x.Xh[ps * XM + 12 + 4 * pp]) = hi;
                *reinterpret_cast<uint2*>(&sm.u.x.Xm[ps * XM + 12 + 4 * pp]) = mi;
                *reinterpret_cast<uint2*>(&sm.u.x.Xl[ps * XM + 12 + 4 * pp]) = lo;
                // the backward's input row, encoding part (X columns 12 + 4pp .. +3)
                const int64_t qq = base + ps;
                if (qq < total)
                    reinterpret_cast<uint2*>(rec.xrow + 5 * qq)[3 + pp] = hi;
            } else {
                // X columns 40..63 (zero weights) are zeroed for every group: H1
                // of the previous group shares their storage
                const uint4 zz = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    *reinterpret_cast<uint4*>(&sm.u.x.Xh[ps * XM + 40 + 8 * k]) = zz;
                    *reinterpret_cast<uint4*>(&sm.u.x.Xm[ps * XM + 40 + 8 * k]) = zz;
                    *reinterpret_cast<uint4*>(&sm.u.x.Xl[ps * XM + 40 + 8 * k]) = zz;
                }
            }
        }
        commit(cb ^ 1);                 // the next group's records
        stage(grp + 2 * (int64_t)gridDim.x);  // and the one after's loads
        __syncthreads();
        // ---- B. H0^T slice = relu(W0^T X^T + b0) -> A1; next group's lookups ----
        // k-step outer: each B fragment is read once and the (mt, nt) chains interleave
        // the accumulators start at the bias (the MFMA adds it, no separate pass)
        f32x4 acc[MT][NT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[mt][nt] = f32x4{b0v[mt][0], b0v[mt][1], b0v[mt][2], b0v[mt][3]};
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int xo = (16 * nt + r16) * XM + 32 * kt + 8 * g;
                const bf16x8 bp[3] = {ld_frag(&sm.u.x.Xh[xo]), ld_frag(&sm.u.x.Xm[xo]),
                                      ld_frag(&sm.u.x.Xl[xo])};
#pragma unroll
                for (int part = 0; part < 3; ++part)
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
                        acc[mt][nt] = mfma16x16x32(A0[mt][kt], bp[part], acc[mt][nt]);
            }
        }
        uint32_t nib[MT][NT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                float h[4];
                nib[mt][nt] = 0u;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    // relu as an integer select on one integer compare, which
                    // also gives the mask bit: a float is > 0 exactly when its
                    // bits are > 0 as an int (a float max would first
                    // canonicalize the MFMA output, one more instruction per
                    // value; +NaN passes through, as torch's relu passes NaN)
                    const uint32_t hb = __float_as_uint(acc[mt][nt][i]);
                    const bool on = (int32_t)hb > 0;
                    h[i] = __uint_as_float(on ? hb : 0u);
                    nib[mt][nt] |= (on ? 1u : 0u) << i;
                }
                st_act3(sm.A1h, sm.A1m, sm.A1l, 16 * nt + r16, f0 + 16 * mt + 4 * g, h);
            }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int64_t q = base + 16 * nt + r16;
            if (q < total)
                reinterpret_cast<uint8_t*>(rec.relu + 8 * q + wv)[g] = relu_byte(nib[0][nt], nib[1][nt]);
        }
        lookup(cb ^ 1);
        __syncthreads();
        // ---- C. H1^T slice = relu(W1^T H0^T + b1) -> A2; next group's colour loads ----
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[mt][nt] = f32x4{b1v[mt][0], b1v[mt][1], b1v[mt][2], b1v[mt][3]};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int ho = (16 * nt + r16) * SM + 32 * kt + 8 * g;
                const bf16x8 bp[3] = {ld_frag(&sm.A1h[ho]), ld_frag(&sm.A1m[ho]),
                                      ld_frag(&sm.A1l[ho])};
#pragma unroll
                for (int part = 0; part < 3; ++part)
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
                        acc[mt][nt] = mfma16x16x32(A1[mt][kt], bp[part], acc[mt][nt]);
            }
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                float h[4];
                nib[mt][nt] = 0u;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    // relu as an integer select on one integer compare, which
                    // also gives the mask bit: a float is > 0 exactly when its
                    // bits are > 0 as an int (a float max would first
                    // canonicalize the MFMA output, one more instruction per
                    // value; +NaN passes through, as torch's relu passes NaN)
                    const uint32_t hb = __float_as_uint(acc[mt][nt][i]);
                    const bool on = (int32_t)hb > 0;
                    h[i] = __uint_as_float(on ? hb : 0u);
                    nib[mt][nt] |= (on ? 1u : 0u) << i;
                }
                st_act<false>(sm.u.h1.A2h, sm.u.h1.A2l, nullptr, 16 * nt + r16, f0 + 16 * mt + 4 * g,
                              lane, h);
            }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int64_t q = base + 16 * nt + r16;
            if (q < total)
                reinterpret_cast<uint8_t*>(rec.relu + 8 * q + 4 + wv)[g] =
                    relu_byte(nib[0][nt], nib[1][nt]);
        }
        // partial logits over this wave's hidden features f0 .. f0 + 31 (one
        // k-step of L^T = W2^T H1^T), read back from the H1 this wave just wrote
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        {
            const bf16x8 w2 = ld_frag(&sm.W2t[r16 * 128 + f0 + 8 * g]);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int ho = (16 * nt + r16) * SM + f0 + 8 * g;
                f32x4 a5 = {0.f, 0.f, 0.f, 0.f};
                a5 = mfma_w(w2, ld_frag(&sm.u.h1.A2h[ho]), ld_frag(&sm.u.h1.A2l[ho]), a5);
                if (g == 0) sm.plog[wv][16 * nt + r16] = make_float4(a5[0], a5[1], a5[2], 0.f);
            }
        }
        if (tid < GS) sm.pbits[tid] = sm.ri[cb][tid].bits;
        pbase = base;
        fetch();
        __syncthreads();
        cb ^= 1;
    }
    if (pbase >= 0) finish_rgb(pbase);  // the loop ended on a barrier
}

// Backward weights as MFMA A fragments in LDS, [tile][k-step][lane] (each
// wave's W1 rows for the g_H0 product, 32 x 128, sit in its registers):
//   W2  [8 row tiles of 16 hidden][64] x 8 bytes   (K = 16 logits, 3 used)
//   W0  [4 k-steps of 32 hidden][64] x 16 bytes    (rows: the 12 features)
constexpr int kBw2Frags = 8 * 64, kBw0Frags = 4 * 64;

// ---- backward ----------------------------------------------------------------
// One workgroup of 8 wavefronts per CU (two per SIMD); wave w owns hidden
// features [16w, 16w + 16) of both hidden layers, except in the g_H0 product
// (phase 8), where it owns features [32(w/2), +32) of samples [16(w%2), +16).
// Per group of 32 records three stages of work, each pairing products that
// do not depend on one another:
//   A  H0^T = W0^T X^T + b0 under the forward's ReLU mask m0    -> A1h, H0f
//      g_H1^T = W2 g_L^T masked by m1                           -> G1h, G1l, G1f
//   B  H1^T = W1^T H0^T + b1 under the forward's mask m1         -> F2
//      g_H0^T = W1 g_H1^T masked by m0                          -> A2h, A2l, F3
//      dW1 += H0^T g_H1; the next group's records are parked
//   C  dW2 += H1^T g_L; dW0 += X^T g_H0; waves 0-1: g_X^T = W0[0:12] g_H0^T
//      over the whole hidden layer, written from registers as the records'
//      feature gradients
// run as two barrier-separated steps: group k's A with group k-1's C, then
// group k's B.  The input rows X come whole from the forward pass (features,
// encoding, bias input), so group k+1's A needs nothing computed after group
// k's B.
// H0 and H1 are recomputed for the dW operands only (from hi parts, one MFMA
// per product): the masks come from the forward pass (per-record words), so the
// backward applies exactly the ReLU pattern the forward applied.
// X (features, encoding, bias), the per-sample adjoints and the flags are
// double-buffered, so the next group's inputs are staged while this one runs.
constexpr int kBwdWaves = 8;
struct alignas(16) SmemBwd {
    uint4 Bw0[kBw0Frags];
    uint2 Bw2[kBw2Frags];
    uint32_t relu[2][GS * 8];  // the forward's ReLU mask words of each record
    uint16_t Xh[2][GS * XM];
    uint16_t Xf[2][XF_ROWS * FM];
    uint16_t A1h[GS * SM];               // H0 (hi), sample-major
    uint16_t G1h[GS * SM], G1l[GS * SM]; // g_H1, sample-major
    uint16_t A2h[GS * SM], A2l[GS * SM]; // g_H0, sample-major
    uint16_t GLh[2][GS * GM], GLl[2][GS * GM];
    uint16_t H0f[128 * FM];              // H0, feature-major
    uint16_t F2[128 * FM];               // H1, feature-major
    uint16_t G1f[128 * FM];              // g_H1, feature-major
    uint16_t F3[128 * FM];               // g_H0, feature-major
    uint16_t GLf[2][16 * FM];
    float gsig[2][GS];
    uint32_t bits[2][GS];
};

__global__ void __launch_bounds__(64 * kBwdWaves) k_mlp_bwd(StepParams P, MlpPacked W,
                                                            RayBuffers R, Records rec,
                                                            float* __restrict__ dw_partial) {
    constexpr int NTH = 64 * kBwdWaves;
    __shared__ SmemBwd sm;
    const int tid = threadIdx.x;
    const int wv = tid >> 6, lane = tid & 63, r16 = lane & 15, g = lane >> 4;
    const int f0 = 16 * wv;
    const int fr = f0 + 4 * g;  // the 4 features of this lane's D fragment rows
    const int pr = lane < GS / kBwdWaves ? wv * (GS / kBwdWaves) + lane : -1;  // staging lanes

    // forward weight slices in registers (this wave's rows of W0^T and W1^T)
    bf16x8 A0[2], A1[4];
    float b0v[4], b1v[4];
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) A0[kt] = ld_frag(W.w0t + (f0 + r16) * 64 + 32 * kt + 8 * g);
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) A1[kt] = ld_frag(W.w1t + (f0 + r16) * 128 + 32 * kt + 8 * g);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        b0v[i] = W.b0[fr + i];
        b1v[i] = W.b1[fr + i];
    }
    // the g_H0 product's tile of this wave: hidden features 32 fb .. +31 (two
    // row tiles) of samples 16 sh .. +15; W1 rows [32 in][128 out] as MFMA A
    // fragments, in registers
    const int fb = wv >> 1, sh = wv & 1;
    bf16x8 Bw1r[2][4];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
            Bw1r[mt][kt] = ld_frag(W.w1 + (32 * fb + 16 * mt + r16) * 128 + 32 * kt + 8 * g);
    // the other backward weights -> LDS
    for (int f = tid; f < kBw2Frags + kBw0Frags; f += NTH) {
        const int ln = f & 63, gg = ln >> 4, rr = ln & 15;
        if (f < kBw2Frags) {  // W2 [128 in][32 out (3 used)], K = 16
            const int t = f >> 6;
            sm.Bw2[f] = *reinterpret_cast<const uint2*>(W.w2 + (16 * t + rr) * 32 + 4 * gg);
        } else {  // W0 rows 0..15 [16][128]
            const int q = f - kBw2Frags, kt = q >> 6;
            sm.Bw0[q] = *reinterpret_cast<const uint4*>(W.w0 + rr * 128 + 32 * kt + 8 * gg);
        }
    }
    f32x4 dW0[3], dW1[8], dW2;
    float db1[4] = {0.f, 0.f, 0.f, 0.f}, db2[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 3; ++a) dW0[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 8; ++a) dW1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
    dW2 = f32x4{0.f, 0.f, 0.f, 0.f};

    // constant parts of the double buffers: X columns 40..63, feature-major X
    // rows 40..47, logits 3..15
    for (int e = tid; e < 2 * GS * 24; e += NTH) {
        const int b = e / (GS * 24), row = (e / 24) % GS, col = 40 + e % 24;
        sm.Xh[b][row * XM + col] = 0;
    }
    for (int e = tid; e < 2 * (XF_ROWS - 40) * FM; e += NTH) {
        const int b = e / ((XF_ROWS - 40) * FM);
        sm.Xf[b][40 * FM + e % ((XF_ROWS - 40) * FM)] = 0;
    }
    for (int e = tid; e < 2 * GS * GM; e += NTH) {
        if (e % GM >= 3) {
            (&sm.GLh[0][0])[e] = 0;
            (&sm.GLl[0][0])[e] = 0;
        }
    }
    for (int e = tid; e < 2 * 16 * FM; e += NTH)
        if ((e % (16 * FM)) >= 3 * FM) (&sm.GLf[0][0])[e] = 0;

    const int64_t total = min((int64_t)R.total[0], P.capacity);
    const int64_t ngroups = (total + GS - 1) / GS;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const uint4 zu4 = make_uint4(0u, 0u, 0u, 0u);

    // register prefetch of a group's records: the staging lanes (4 per wave, a
    // record each) take its flags, adjoints and ReLU masks; the row lanes (20
    // per wave) one 16-byte fifth of its input row each.  Every address
    // follows from the record index alone, so the loads of the next group are
    // issued at the start of the current one and land during its products.
    float4 p_adj = z4;
    uint32_t p_bits = 0u;
    uint4 p_r0 = zu4, p_r1 = zu4, p_x = zu4;
    const int xe = (lane >= 32 && lane < 52) ? wv * 20 + (lane - 32) : -1;
    const int xr = xe / 5, xj = xe % 5;  // row lane: record, 8-column part
    auto stage = [&](int64_t grp) {
        const int64_t q = grp * GS + pr;
        p_adj = z4;
        p_bits = 0u;
        p_r0 = p_r1 = zu4;
        if (pr >= 0 && grp < ngroups && q < total) {
            p_bits = __float_as_uint(rec.ab[q].y);
            p_adj = rec.adj[q];
            p_r0 = reinterpret_cast<const uint4*>(rec.relu)[2 * q];  // ReLU masks
            p_r1 = reinterpret_cast<const uint4*>(rec.relu)[2 * q + 1];
        }
        const int64_t qx = grp * GS + xr;
        p_x = zu4;
        if (xe >= 0 && grp < ngroups && qx < total) p_x = rec.xrow[5 * qx + xj];
    };
    // staged records -> buffer b: flags, adjoints (dL/dlogit as (hi, lo) pairs,
    // dL/dsigma), ReLU masks, and the input rows (X columns 0..39, sample-major
    // and feature-major)
    auto park = [&](int b) {
        if (pr >= 0) {
            sm.bits[b][pr] = p_bits;  // zero for records past the end
            sm.gsig[b][pr] = p_adj.w;
            reinterpret_cast<uint4*>(&sm.relu[b][8 * pr])[0] = p_r0;
            reinterpret_cast<uint4*>(&sm.relu[b][8 * pr])[1] = p_r1;
            const float gl[3] = {p_adj.x, p_adj.y, p_adj.z};
            db2[0] += gl[0];
            db2[1] += gl[1];
            db2[2] += gl[2];
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                uint16_t h, l;
                split_bf(gl[o], h, l);
                sm.GLh[b][pr * GM + o] = h;
                sm.GLl[b][pr * GM + o] = l;
                sm.GLf[b][o * FM + pr] = h;
            }
        }
        if (xe >= 0) {
            *reinterpret_cast<uint4*>(&sm.Xh[b][xr * XM + 8 * xj]) = p_x;  // 16-byte aligned
            const uint32_t w4[4] = {p_x.x, p_x.y, p_x.z, p_x.w};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                sm.Xf[b][(8 * xj + 2 * c) * FM + xr] = (uint16_t)(w4[c] & 0xffffu);
                sm.Xf[b][(8 * xj + 2 * c + 1) * FM + xr] = (uint16_t)(w4[c] >> 16);
            }
        }
    };
    // stage C of a group whose g_H0 and H1 sit in A2h/A2l, F3 and F2 (its
    // inputs in buffer b, its records from pbase): dW2, dW0 and the records'
    // feature gradients
    auto finish = [&](int b, int64_t pbase) {
        // dW2[in = f0 + row][out] += sum over the group's samples H1[s][in] gL[s][out]
        dW2 = mfma16x16x32(ld_frag(&sm.F2[(f0 + r16) * FM + 8 * g]),
                           ld_frag(&sm.GLf[b][r16 * FM + 8 * g]), dW2);
        {
            const bf16x8 bh = ld_frag(&sm.F3[(f0 + r16) * FM + 8 * g]);
#pragma unroll
            for (int mi = 0; mi < 3; ++mi)
                dW0[mi] = mfma16x16x32(ld_frag(&sm.Xf[b][(16 * mi + r16) * FM + 8 * g]), bh, dW0[mi]);
        }
        if (wv < NT) {
            const int nt = wv;
            f32x4 gx = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                const int go = (16 * nt + r16) * SM + 32 * kt + 8 * g;
                gx = mfma_w(as_bf8(sm.Bw0[kt * 64 + lane]), ld_frag(&sm.A2h[go]),
                            ld_frag(&sm.A2l[go]), gx);
            }
            // lane (g, r16): features 4g..4g+3 of sample 16 nt + r16 = part g of its
            // record's gradient; part 3 carries dL/dsigma
            const int sidx = 16 * nt + r16;
            const int64_t q = pbase + sidx;
            if (q < total) {
                const uint32_t bits = sm.bits[b][sidx];
                float4 v = z4;
                if (g < 3) {
                    if (bits & kAliveW) v = make_float4(gx[0], gx[1], gx[2], gx[3]);
                } else {
                    v.x = sm.gsig[b][sidx];
                }
                rec.grad[4 * q + g] = v;
            }
        }
    };

    stage(blockIdx.x);
    park(0);
    __syncthreads();
    int cb = 0;          // buffer of the current group
    int64_t pbase = -1;  // the previous group, whose stage C is pending (-1: none)

    for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {  // uniform bound
        const int64_t base = grp * GS, next = grp + gridDim.x;
        stage(next);  // in flight during this group's products, parked at its end
        // ---- 1. this group's stage A: H0^T slice = W0^T X^T + b0 under the
        //         forward's mask -> A1h, H0f; g_H1^T slice = W2[slice] g_L^T under
        //         the mask -> G1h, G1l, G1f.  The previous group's stage C ----
        uint32_t m0 = 0u, m1 = 0u;  // the forward's ReLU masks: bit nt * 4 + i
        {
            const int sh = 8 * g + 4 * (wv & 1);  // features fr..fr+3 in word wv/2
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const uint32_t* mw = &sm.relu[cb][8 * (16 * nt + r16)];
                m0 |= ((mw[wv >> 1] >> sh) & 0xFu) << (nt * 4);
                m1 |= ((mw[4 + (wv >> 1)] >> sh) & 0xFu) << (nt * 4);
            }
        }
        f32x4 acc[NT];  // starting at the bias, as in the forward
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4{b0v[0], b0v[1], b0v[2], b0v[3]};
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int xo = (16 * nt + r16) * XM + 32 * kt + 8 * g;
                acc[nt] = mfma16x16x32(A0[kt], ld_frag(&sm.Xh[cb][xo]), acc[nt]);
            }
        }
        {
            const s16x4 a = __builtin_bit_cast(s16x4, sm.Bw2[wv * 64 + lane]);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                f32x4 c4 = {0.f, 0.f, 0.f, 0.f};
                const int go = (16 * nt + r16) * GM + 4 * g;
                c4 = mfma16x16x16(a, ld_frag4(&sm.GLh[cb][go]), c4);
                c4 = mfma16x16x16(a, ld_frag4(&sm.GLl[cb][go]), c4);
                float gv[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    gv[i] = keep_if(m1, nt * 4 + i, c4[i]);
                    db1[i] += gv[i];
                }
                st_act<true>(sm.G1h, sm.G1l, sm.G1f, 16 * nt + r16, fr, lane, gv);
            }
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            float h[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) h[i] = keep_if(m0, nt * 4 + i, acc[nt][i]);
            st_hi_fm(sm.A1h, sm.H0f, 16 * nt + r16, fr, lane, h);
        }
        if (pbase >= 0) finish(cb ^ 1, pbase);
        __syncthreads();
        // ---- 2. B. H1^T slice = W1^T H0^T + b1 under the mask -> F2 (dW2 only);
        //         g_H0^T tile = W1[rows] g_H1^T under the mask -> A2h, A2l, F3;
        //         dW1 += H0^T g_H1; the next group's records are parked ----
        // (g_H0 in 32 features x 16 samples per wave: half the g_H1 reads of 16 x 32)
        f32x4 acc2[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) acc2[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            const int go = (16 * sh + r16) * SM + 32 * kt + 8 * g;
            const bf16x8 bh = ld_frag(&sm.G1h[go]), bl = ld_frag(&sm.G1l[go]);
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) acc2[mt] = mfma_w(Bw1r[mt][kt], bh, bl, acc2[mt]);
        }
        {
            // dW1[in][out] += sum over samples H0[s][in] g_H1[s][out]; this wave's
            // block: inputs 64 (wv & 1) .. +63, outputs 32 (wv >> 1) .. +31 (fewer
            // LDS reads than a full column of 128 inputs x 16 outputs)
            const int i0 = 64 * (wv & 1), o0 = 32 * (wv >> 1);
            const bf16x8 b0 = ld_frag(&sm.G1f[(o0 + r16) * FM + 8 * g]);
            const bf16x8 b1 = ld_frag(&sm.G1f[(o0 + 16 + r16) * FM + 8 * g]);
#pragma unroll
            for (int mi = 0; mi < 4; ++mi) {
                const bf16x8 a = ld_frag(&sm.H0f[(i0 + 16 * mi + r16) * FM + 8 * g]);
                dW1[2 * mi] = mfma16x16x32(a, b0, dW1[2 * mi]);
                dW1[2 * mi + 1] = mfma16x16x32(a, b1, dW1[2 * mi + 1]);
            }
        }
        {
            // g_H0 masked by the layer-0 word of features 32 fb .. +31
            const uint32_t mw = sm.relu[cb][8 * (16 * sh + r16) + fb];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                float gv[4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    gv[i] = keep_if(mw, 8 * g + 4 * mt + i, acc2[mt][i]);
                st_act<true>(sm.A2h, sm.A2l, sm.F3, 16 * sh + r16, 32 * fb + 16 * mt + 4 * g, lane,
                             gv);
            }
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4{b1v[0], b1v[1], b1v[2], b1v[3]};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int ho = (16 * nt + r16) * SM + 32 * kt + 8 * g;
                acc[nt] = mfma16x16x32(A1[kt], ld_frag(&sm.A1h[ho]), acc[nt]);
            }
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            float h[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) h[i] = keep_if(m1, nt * 4 + i, acc[nt][i]);
            st_fm(sm.F2, 16 * nt + r16, fr, lane, h);
        }
        park(cb ^ 1);  // the next group's records, loaded at this group's start
        __syncthreads();
        pbase = base;
        cb ^= 1;
    }
    if (pbase >= 0) finish(cb ^ 1, pbase);  // the loop ended on a barrier

    // ---- flush this workgroup's dW partial ----
    float* out = dw_partial + (int64_t)blockIdx.x * kPartStride;
#pragma unroll
    for (int mi = 0; mi < 3; ++mi)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int in = 16 * mi + 4 * g + i, o = f0 + r16;
            if (in < 39)
                out[kW0 + in * 128 + o] = dW0[mi][i];
            else if (in == 39)
                out[kB0 + o] = dW0[mi][i];
        }
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int in = 64 * (wv & 1) + 16 * (t >> 1) + 4 * g + i;
            const int o = 32 * (wv >> 1) + 16 * (t & 1) + r16;
            out[kW1 + in * 128 + o] = dW1[t][i];
        }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (r16 < 3) out[kW2 + (fr + i) * 3 + r16] = dW2[i];
        float v = db1[i];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        if (r16 == 0) out[kB1 + fr + i] = v;
    }
    // db2 was accumulated by the staging lanes (park); the loop ended on a barrier,
    // so gsig[0] is free for the per-wave sums
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        float v = pr >= 0 ? db2[o] : 0.f;
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        if (lane == 0) sm.gsig[0][4 * wv + o] = v;
    }
    __syncthreads();
    if (tid < 3) {
        float v = 0.f;
#pragma unroll
        for (int w = 0; w < kBwdWaves; ++w) v += sm.gsig[0][4 * w + tid];
        out[kB2 + tid] = v;
    }
}

// Sum of the per-workgroup dW partials (dw_reduce_block): one workgroup per
// kReduceVec four-parameter vectors.  The fused step runs these workgroups
// inside the scatter's launch instead (launch_scatter), where they overlap the
// scatter's atomics; this launch serves the optimizer-only step.
__global__ void __launch_bounds__(256) k_dw_reduce(const float* __restrict__ part, int nparts,
                                                   float* __restrict__ dw, MlpAdam A) {
    __shared__ float4 red[kReduceGrp * kReduceVec];
    dw_reduce_block(blockIdx.x, part, nparts, dw, A, red);
}


}  // namespace

namespace pvdb {



int launch_mlp(bool bwd, bool bf16col, int grid, const StepParams& P, const TopoView& t,
               const void* color, const MlpPacked& W, const RayBuffers& R, const Records& rec,
               float* dw_partial, hipStream_t s) {
    if (bwd) {
        hipLaunchKernelGGL(k_mlp_bwd, dim3(grid), dim3(64 * kBwdWaves), 0, s, P, W, R, rec,
                           dw_partial);
        PVDB_LAUNCHED("k_mlp_bwd");
    } else {
        if (bf16col)
            hipLaunchKernelGGL((k_mlp_fwd<true>), dim3(grid), dim3(256), 0, s, P, t, color, W, R, rec);
        else
            hipLaunchKernelGGL((k_mlp_fwd<false>), dim3(grid), dim3(256), 0, s, P, t, color, W, R, rec);
        PVDB_LAUNCHED("k_mlp_fwd");
    }
    return PVDB_OK;
}

int launch_dw_reduce(const float* part, int nparts, float* dw, const MlpAdam& adam,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_dw_reduce, dim3(kReduceBlocks), dim3(256), 0, s, part, nparts, dw, adam);
    PVDB_LAUNCHED("k_dw_reduce");
    return PVDB_OK;
}


}  // namespace pvdb
