// Fused step, ray side: counting march, offsets scan, the forward march that
// writes per-sample records, and the per-ray composite / loss pass.
//
// One wavefront marches one ray, 64 consecutive steps per iteration (lane l
// takes step base+l), so sample positions are the reference's
// p_i = (o + d*t_min) + (d/|d|)*(stepdist*i) (render_utils_kernel.cu:57-79,
// 166-194) and the occupancy decision is the MaskGrid rule round(p*s+b)
// (render_utils_kernel.cu:373-390) on the VDB value masks.  Live lanes are
// compacted with ballot + mbcnt.  Transmittance is an exclusive product scan
// over the wave carried from chunk to chunk; the early stop after the first
// sample with T < 1e-3 and the alpha / weight culls follow
// render_utils_kernel.cu:576-603 and dvgo.py:319-335.
#include "fused_common.h"
#include "gather.h"
#include "plenvdb_hip.h"

namespace pvdb {
TopoView view_of(const pvdb_topology* t);
}

using namespace pvdb;

namespace {

struct RaySetup {
    float sx, sy, sz;  // o + d * t_min
    float ux, uy, uz;  // d / |d|
    int64_t n_steps;
    float t_min, t_max;
};

// Ray bounds with the reference's operation order (render_utils_kernel.cu:11-79).
__device__ __forceinline__ RaySetup ray_setup(const StepParams& P, int r) {
    const float ox = P.rays_o[3 * r], oy = P.rays_o[3 * r + 1], oz = P.rays_o[3 * r + 2];
    const float dx = P.rays_d[3 * r], dy = P.rays_d[3 * r + 1], dz = P.rays_d[3 * r + 2];
    const float vx = dx == 0.f ? 1e-6f : dx;
    const float vy = dy == 0.f ? 1e-6f : dy;
    const float vz = dz == 0.f ? 1e-6f : dz;
    const float ax = (P.mx[0] - ox) / vx, ay = (P.mx[1] - oy) / vy, az = (P.mx[2] - oz) / vz;
    const float bx = (P.mn[0] - ox) / vx, by = (P.mn[1] - oy) / vy, bz = (P.mn[2] - oz) / vz;
    const float fr = P.far_per_ray ? P.far_per_ray[r] : P.far_;
    const float lo = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    const float hi = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    RaySetup s;
    s.t_min = fmaxf(fminf(lo, fr), P.near_);
    s.t_max = fmaxf(fminf(hi, fr), P.near_);
    const float rn = sqrtf(dx * dx + dy * dy + dz * dz);
    const float q = ceilf(((s.t_max - s.t_min) * rn) / P.stepdist);
    s.n_steps = (int64_t)(q > 1.f ? q : 1.f);
    s.sx = ox + dx * s.t_min;
    s.sy = oy + dy * s.t_min;
    s.sz = oz + dz * s.t_min;
    s.ux = dx / rn;
    s.uy = dy / rn;
    s.uz = dz / rn;
    return s;
}

struct StepPoint {
    float px, py, pz;
    bool keep;  // inside the bbox and in an active voxel
};

// Step i of n (I = int32_t when n < 2^24, where (float)i is the same as from
// int64, so the common case needs no 64-bit index arithmetic).
template <typename I>
__device__ __forceinline__ StepPoint step_point(const StepParams& P, const TopoView& t,
                                                const RaySetup& s, I i, I n) {
    StepPoint sp;
    const float dist = P.stepdist * (float)i;
    sp.px = s.sx + s.ux * dist;
    sp.py = s.sy + s.uy * dist;
    sp.pz = s.sz + s.uz * dist;
    const bool out = (P.mn[0] > sp.px) | (P.mn[1] > sp.py) | (P.mn[2] > sp.pz) |
                     (P.mx[0] < sp.px) | (P.mx[1] < sp.py) | (P.mx[2] < sp.pz);
    sp.keep = false;
    if (i < n && !out) {
        const int a = (int)roundf(sp.px * P.ms[0] + P.mb[0]);
        const int b = (int)roundf(sp.py * P.ms[1] + P.mb[1]);
        const int c = (int)roundf(sp.pz * P.ms[2] + P.mb[2]);
        sp.keep = voxel_active(t, a, b, c);
    }
    return sp;
}

// Empty-space skip in two levels.  A coarse pass gives each lane a segment of
// kSeg consecutive steps (a wave covers 64 * kSeg steps) and looks up the
// topology's empty-run radius r of the 8^3 cell holding the segment's middle
// step's occupancy voxel round(p*ms + mb): no leaf lies within r - 1 cells of
// that cell, so the voxel sits at least 8 (r - 1) voxels inside an empty box
// on every axis.  Positions are p_i = s + u*(stepdist*i) (from i, not from a
// running sum), so the segment's steps lie within |g| (kSeg/2 + 1) voxels of
// the middle one, g the per-step move in voxels; when that stays below
// 8 (r - 1) - 1 on every axis, every step of the segment falls in a missing
// leaf and is rejected by the occupancy test: the segment is certified empty.
// The fine pass then tests the remaining segments' steps, four segments (64
// steps) per window, in step order.  Without an empty-run map (the two-level
// layout) no segment is certified and every step is tested.
constexpr int kSeg = 16;

template <typename I>
__device__ __forceinline__ bool segment_empty(const StepParams& P, const TopoView& t,
                                              const RaySetup& s, I mid, float reach) {
    const float dist = P.stepdist * (float)mid;
    const int v0 = (int)roundf((s.sx + s.ux * dist) * P.ms[0] + P.mb[0]);
    const int v1 = (int)roundf((s.sy + s.uy * dist) * P.ms[1] + P.mb[1]);
    const int v2 = (int)roundf((s.sz + s.uz * dist) * P.ms[2] + P.mb[2]);
    const int a = (v0 >> 3) - t.lo0, b = (v1 >> 3) - t.lo1, c = (v2 >> 3) - t.lo2;
    const int oa = a < 0 ? -a : (a >= t.d0 ? a - t.d0 + 1 : 0);
    const int ob = b < 0 ? -b : (b >= t.d1 ? b - t.d1 + 1 : 0);
    const int oc = c < 0 ? -c : (c >= t.d2 ? c - t.d2 + 1 : 0);
    int r = max(oa, max(ob, oc));
    if (r == 0) r = (int)t.erun[((int64_t)a * t.d1 + b) * t.d2 + c];  // 0: a leaf exists
    return reach < 8.f * (float)(r - 1) - 1.f;
}

// Runs window(step, n) over the ray's steps, lane l taking step `step` (n for
// none; steps increase with the lane), skipping certified-empty segments; with
// 32-bit indices when the ray's step count allows.
template <typename I, typename F>
__device__ __forceinline__ void windows_of(const StepParams& P, const TopoView& t,
                                           const RaySetup& s, I n, F&& window) {
    const int lane = lane_id();
    if (!t.erun) {
        for (I base = 0; base < n; base += 64) window(base + lane, n);
        return;
    }
    // voxels a step moves on the worst axis, times the half segment (+1 for rounding)
    const float g = fmaxf(fmaxf(fabsf(s.ux * P.ms[0]), fabsf(s.uy * P.ms[1])),
                          fabsf(s.uz * P.ms[2])) * P.stepdist;
    const float reach = g * (float)(kSeg / 2 + 1);
    for (I base = 0; base < n; base += I(64 * kSeg)) {
        const I seg0 = base + (I)(kSeg * lane);
        const bool live = seg0 < n && !segment_empty(P, t, s, seg0 + kSeg / 2, reach);
        uint64_t m = __ballot(live);
        while (m) {  // uniform: four segments per window
            int pick[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                pick[q] = m ? __ffsll((unsigned long long)m) - 1 : -1;
                m &= m - 1;
            }
            const int sel = pick[lane >> 4];
            const I step = sel >= 0 ? base + (I)(kSeg * sel + (lane & 15)) : n;
            window(step < n ? step : n, n);
        }
    }
}
template <typename F>
__device__ __forceinline__ void for_windows(const StepParams& P, const TopoView& t,
                                            const RaySetup& s, F&& window) {
    if (s.n_steps < (int64_t(1) << 24))
        windows_of(P, t, s, (int32_t)s.n_steps, window);
    else
        windows_of(P, t, s, s.n_steps, window);
}

__device__ __forceinline__ int mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Counting march: occupancy-kept samples per ray, and the first kStepCache
// kept step indices of each ray of fewer than 65536 steps, for the forward
// march.  The workgroups past the rays' pack the step's MLP operands
// (independent work, one launch fewer).
__global__ void __launch_bounds__(kMaxRayWaves * 64) k_ray_count(StepParams P, TopoView t,
                                                                 RayBuffers R, PackArgs pk,
                                                                 int ray_blocks) {
    if ((int)blockIdx.x >= ray_blocks) {  // uniform per workgroup
        const int item = ((int)blockIdx.x - ray_blocks) * (int)blockDim.x + (int)threadIdx.x;
        if (item < kPackItems) pack_mlp_item(item, pk);
        return;
    }
    const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= P.n_rays) return;
    const int lane = lane_id();
    // the view direction, loaded now and encoded after the march (its loads
    // then add no latency to the wave)
    float vd0 = 0.f, vd1 = 0.f, vd2 = 0.f;
    if (R.pe) {
        vd0 = P.viewdirs[3 * (int64_t)r];
        vd1 = P.viewdirs[3 * (int64_t)r + 1];
        vd2 = P.viewdirs[3 * (int64_t)r + 2];
    }
    const RaySetup s = ray_setup(P, r);
    if (lane < 2) {  // the set-up, for the forward march (its divisions run once)
        const uint64_t ns = (uint64_t)s.n_steps;
        R.setup[2 * (int64_t)r + lane] =
            lane == 0 ? make_float4(s.sx, s.sy, s.sz, __uint_as_float((uint32_t)ns))
                      : make_float4(s.ux, s.uy, s.uz, __uint_as_float((uint32_t)(ns >> 32)));
    }
    int count = 0;
    uint16_t* cache = R.steps + (int64_t)r * kStepCache;
    const bool cached = s.n_steps < 65536;  // uniform over the wave
    for_windows(P, t, s, [&](auto step, auto n) {
        const StepPoint sp = step_point(P, t, s, step, n);
        const uint64_t km = __ballot(sp.keep);
        if (cached && sp.keep) {
            const int at = count + mbcnt(km);
            if (at < kStepCache) cache[at] = (uint16_t)step;
        }
        count += __popcll(km);
    });
    if (lane == 0) R.count[r] = count;
    if (R.pe && lane < 28) {
        // the MLP's input columns 12..39 for this ray's samples (dvgo.py:354-358):
        // the view direction, sin then cos of it at 1, 2, 4, 8 (hardware sin/cos,
        // error about 1e-6 for |arg| <= 8), and the bias column; one column per lane
        float v = 1.f;
        if (lane < 3) {
            v = lane == 0 ? vd0 : (lane == 1 ? vd1 : vd2);
        } else if (lane < 27) {
            const int idx = lane - 3, axis = (idx % 12) / 4, fq = idx % 4;
            const float arg = (axis == 0 ? vd0 : (axis == 1 ? vd1 : vd2)) * (float)(1 << fq);
            v = idx < 12 ? __sinf(arg) : __cosf(arg);
        }
        uint32_t h, m, l;
        split3(v, 0.f, h, m, l);
        uint16_t* row = R.pe + 96 * (int64_t)r;
        row[lane] = (uint16_t)h;
        row[32 + lane] = (uint16_t)m;
        row[64 + lane] = (uint16_t)l;
    }
}

// Exclusive scan of the per-ray counts by one workgroup of 1024 threads, in
// chunks of 1024 * kScanPer: each thread takes kScanPer consecutive counts
// (16-byte loads, all issued before any is used),
// wavefronts scan their thread sums with shuffles, one LDS pass combines the
// 16 wavefronts, and a carry runs from chunk to chunk.  The sums run in 64 bits
// and every stored offset (and the total) saturates at INT32_MAX: a batch with
// more samples than that keeps offsets >= the record capacity (which the step
// caps below INT32_MAX), so its excess records are dropped, never wrapped.
constexpr int kScanPer = 16;  // counts per thread and chunk (65,536 rays in four chunks)

__global__ void __launch_bounds__(1024) k_scan(RayBuffers R, int n) {
    __shared__ long long wsum[16];
    __shared__ long long carry_s;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const long long kSat = 0x7fffffffll;
    long long carry = 0;
    for (int c0 = 0; c0 < n; c0 += 1024 * kScanPer) {
        const int b = c0 + kScanPer * tid;
        int v[kScanPer];
        if (b + kScanPer <= n) {  // 64-byte aligned (the workspace arrays are 256-byte aligned)
#pragma unroll
            for (int k = 0; k < kScanPer / 4; ++k) {
                const int4 a = *reinterpret_cast<const int4*>(R.count + b + 4 * k);
                v[4 * k] = a.x;
                v[4 * k + 1] = a.y;
                v[4 * k + 2] = a.z;
                v[4 * k + 3] = a.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < kScanPer; ++k) v[k] = b + k < n ? R.count[b + k] : 0;
        }
        long long sum = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) sum += v[k];
        long long inc = sum;  // inclusive scan of the thread sums within the wavefront
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const long long o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) wsum[wv] = inc;
        __syncthreads();
        long long before = carry;
        for (int w = 0; w < wv; ++w) before += wsum[w];
        if (tid == 1023) carry_s = before + inc;
        long long run = before + inc - sum;
        int off[kScanPer];
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            off[k] = (int)(run < kSat ? run : kSat);
            run += v[k];
        }
        if (b + kScanPer <= n) {
#pragma unroll
            for (int k = 0; k < kScanPer / 4; ++k)
                *reinterpret_cast<int4*>(R.offset + b + 4 * k) =
                    make_int4(off[4 * k], off[4 * k + 1], off[4 * k + 2], off[4 * k + 3]);
        } else {
#pragma unroll
            for (int k = 0; k < kScanPer; ++k)
                if (b + k < n) R.offset[b + k] = off[k];
        }
        __syncthreads();
        carry = carry_s;
        __syncthreads();  // wsum and carry_s are reused by the next chunk
    }
    if (tid == 0) R.total[0] = (int)(carry < kSat ? carry : kSat);
}

// Forward march: one record per occupancy-kept sample.  The sample work --
// density gather, raw2alpha, the transmittance scan, records -- runs on 64
// kept steps at a time, in step order, so lanes are not spent on rejected
// steps.  The kept step indices come from the counting march's per-ray store
// when it holds them all; otherwise the occupancy test runs again over windows
// of 64 consecutive steps and the kept indices are queued in LDS.
__global__ void __launch_bounds__(kMaxRayWaves * 64) k_ray_march(StepParams P, TopoView t,
                                                                 const float* density,
                                                                 RayBuffers R, Records rec) {
    extern __shared__ int64_t queue_all[];  // 128 entries per wave (dynamic: waves x 1 KB)
    const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= P.n_rays) return;  // uniform over the wave
    const int lane = lane_id();
    int64_t* queue = queue_all + 128 * (threadIdx.x >> 6);
    RaySetup s;  // the counting march's set-up (positions and step count only)
    {
        const float4 a = R.setup[2 * (int64_t)r], b = R.setup[2 * (int64_t)r + 1];
        s.sx = a.x;
        s.sy = a.y;
        s.sz = a.z;
        s.ux = b.x;
        s.uy = b.y;
        s.uz = b.z;
        s.n_steps = (int64_t)(((uint64_t)__float_as_uint(b.w) << 32) | __float_as_uint(a.w));
        s.t_min = s.t_max = 0.f;  // unused by the forward march
    }
    const int64_t off = R.offset[r];
    const int cnt = R.count[r];  // the counting march's total bounds this ray's records
    float Tc = 1.f;  // transmittance carried from batch to batch (FP32, as the reference's T_cum)
    bool stopped = false;
    int written = 0, queued = 0;
    // sample work on the next k (<= 64) kept steps, lane l holding step i of
    // the l-th, in step order
    auto flush = [&](int k, int64_t i) __attribute__((always_inline)) {
        const bool mine = lane < k;
        float x = 0.f, y = 0.f, z = 0.f, e = 0.f, alpha = 0.f;
        const bool live = mine && !stopped;
        if (live) {
            const float dist = P.stepdist * (float)i;
            const float px = s.sx + s.ux * dist, py = s.sy + s.uy * dist, pz = s.sz + s.uz * dist;
            x = ((px - P.mn[0]) / P.len[0]) * P.rm1[0];
            y = ((py - P.mn[1]) / P.len[1]) * P.rm1[1];
            z = ((pz - P.mn[2]) / P.len[2]) * P.rm1[2];
            const Cell cell = make_cell(x, y, z);
            LeafCache cache;
            const float sigma = gather_density(t, density, cell, cache);
            e = expf(sigma + P.act_shift);
            alpha = 1.f - powf(1.f + e, -P.interval);
        }
        const bool akeep = live && (P.thres > 0.f ? alpha > P.thres : true);
        // inclusive product of (1 - alpha) over alpha-kept lanes, in step order
        // (a log-depth scan: each T is a product of the same factors as the
        // reference's running product, grouped differently)
        const float incl = wave_incl_prod(akeep ? 1.f - alpha : 1.f);
        const float excl = wave_shr1(incl, 1.f);
        const float T = Tc * excl, Ta = Tc * incl;
        const uint64_t stopm = __ballot(akeep && Ta < 1e-3f);
        const int stop_lane = stopm ? __ffsll((unsigned long long)stopm) - 1 : 64;
        const bool alive_a = akeep && lane <= stop_lane;
        const float w = T * alpha;
        const bool alive_w = alive_a && (P.thres > 0.f ? w > P.thres : true);
        const int idx = written + lane;
        if (mine && idx < cnt) {
            const int64_t q = off + idx;
            if (q >= 0 && q < P.capacity) {
                uint32_t bits = (uint32_t)r | (alive_a ? kAliveA : 0u) | (alive_w ? kAliveW : 0u);
                rec.xyze[q] = make_float4(x, y, z, alive_a ? e : 0.f);
                rec.ab[q] = make_float2(alive_a ? alpha : 0.f, __uint_as_float(bits));
            }
        }
        written += k;
        if (!stopped) {
            if (stop_lane < 64) {
                Tc = wave_read(Ta, stop_lane);
                stopped = true;
            } else {
                Tc = wave_read(Ta, 63);
            }
        }
    };
    if (cnt <= kStepCache && s.n_steps < 65536) {
        // the counting march stored every kept step of this ray (uniform branch)
        const uint16_t* cache = R.steps + (int64_t)r * kStepCache;
        for (int b = 0; b < cnt; b += 64) {
            const int k = min(64, cnt - b);
            flush(k, lane < k ? (int64_t)cache[b + lane] : int64_t(0));
        }
    } else {
        for_windows(P, t, s, [&](auto step, auto n) {
            const StepPoint sp = step_point(P, t, s, step, n);
            const uint64_t km = __ballot(sp.keep);
            if (km == 0) return;
            if (sp.keep) queue[queued + mbcnt(km)] = (int64_t)step;
            queued += __popcll(km);
            if (queued >= 64) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                flush(64, queue[lane]);
                // the overflow (fewer than 64 entries) moves to the queue's front
                const int64_t rest = lane + 64 < queued ? queue[lane + 64] : 0;
                __builtin_amdgcn_wave_barrier();
                if (lane + 64 < queued) queue[lane] = rest;
                queued -= 64;
            }
        });
        if (queued > 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            flush(queued, queue[lane]);
        }
    }
    if (lane == 0) R.t_last[r] = Tc;
}

// Transmittance before each record of a 64-record chunk, recomputed from the
// stored alphas rather than kept per sample: the same inclusive product scan
// of (1 - alpha) over the alpha-alive lanes, carried from chunk to chunk in Tc,
// that the forward march ran.  The chunks are the march's 64-sample batches
// (its record index advances by 64 per batch), and the product for a lane only
// involves lanes at or before it, so every alive record's T is bit-identical
// to the march's.
__device__ __forceinline__ float chunk_transmittance(float alpha, uint32_t bits, int lane,
                                                     float& Tc) {
    const float incl = wave_incl_prod((bits & kAliveA) ? 1.f - alpha : 1.f);
    const float T = Tc * wave_shr1(incl, 1.f);
    Tc = wave_read(Tc * incl, 63);
    return T;
}

// Per-ray composite: inclusive prefix of w*rgb (kept for the alpha adjoint),
// ray colour = sum + T_last*bg (dvgo.py:365-370), the loss of run.py:551-574
// (MSE; background entropy of T_last; per-sample rgb loss with the weights
// detached) and its gradient.
__global__ void __launch_bounds__(kMaxRayWaves * 64) k_ray_composite(StepParams P, RayBuffers R,
                                                                     Records rec) {
    const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= P.n_rays) return;
    const int lane = lane_id();
    const int64_t off = R.offset[r];
    const int cnt = R.count[r];
    const float t0 = P.target[3 * r], t1 = P.target[3 * r + 1], t2 = P.target[3 * r + 2];
    const bool rgbper = P.rgbper_scale > 0.f;
    float cr = 0.f, cg = 0.f, cb = 0.f;
    float sq = 0.f;  // this lane's sum of w * |rgb - target|^2 (rgbper term)
    float Tc = 1.f;
    for (int base = 0; base < cnt; base += 64) {
        const int j = base + lane;
        float vr = 0.f, vg = 0.f, vb = 0.f;
        const int64_t q = off + j;
        const bool valid = j < cnt && q >= 0 && q < P.capacity;
        float2 ab = make_float2(0.f, 0.f);
        if (valid) ab = rec.ab[q];
        const uint32_t bits = __float_as_uint(ab.y);
        const float w = chunk_transmittance(ab.x, bits, lane, Tc) * ab.x;
        if (bits & kAliveW) {
            const float4 c = rec.rgb[q];
            vr = w * c.x;
            vg = w * c.y;
            vb = w * c.z;
            if (rgbper) {
                const float e0 = c.x - t0, e1 = c.y - t1, e2 = c.z - t2;
                sq += w * ((e0 * e0 + e1 * e1) + e2 * e2);
            }
        }
        float v3[3] = {vr, vg, vb};
        wave_incl_sum(v3);
        cr += wave_read(v3[0], 63);
        cg += wave_read(v3[1], 63);
        cb += wave_read(v3[2], 63);
    }
    // the ray's colour, loss and dL/drgb (every lane holds them)
    const float tl = R.t_last[r];
    const float o0 = cr + tl * P.bg, o1 = cg + tl * P.bg, o2 = cb + tl * P.bg;
    const float d0 = o0 - t0, d1 = o1 - t1, d2 = o2 - t2;
    const float g0 = 2.f * d0 * P.loss_scale, g1 = 2.f * d1 * P.loss_scale,
                g2 = 2.f * d2 * P.loss_scale;
    float loss = (d0 * d0 + d1 * d1 + d2 * d2) * P.loss_scale;
    float gent = 0.f;  // dL/dT_last of the entropy term
    if (P.ent_scale > 0.f) {
        // pout = clamp(T_last, 1e-6, 1 - 1e-6); -(p log p + (1-p) log(1-p)), whose
        // derivative log(1-p) - log(p) passes the clamp only inside its bounds
        const float p = fminf(fmaxf(tl, 1e-6f), 0.999999f);
        const float lp = logf(p), lq = logf(1.f - p);
        loss += -(p * lp + (1.f - p) * lq) * P.ent_scale;
        if (tl >= 1e-6f && tl <= 0.999999f) gent = (lq - lp) * P.ent_scale;
    }
    if (rgbper) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) sq += __shfl_xor(sq, d, 64);
        loss += sq * P.rgbper_scale;
    }
    const float gbg = P.bg * (g0 + g1 + g2) + gent;  // dL/dT_last
    const float grp = 2.f * P.rgbper_scale;          // dL/drgb_i = w_i (g + grp (rgb_i - target))
    if (lane == 0) {
        R.ray_out[r] = make_float4(o0, o1, o2, tl);
        R.ray_grad[r] = make_float4(g0, g1, g2, gbg);
        R.ray_ctot[r] = make_float4(cr, cg, cb, 0.f);
        R.ray_loss[r] = loss;
    }
    // second pass: each sample's adjoints.  dL/dlogit = dL/drgb_i * rgb(1 - rgb)
    // (sigmoid, dvgo.py:362), dL/drgb_i = w (g + grp (rgb_i - target)) with the
    // weight detached in the rgbper term; dL/dsigma through Alphas2Weights and raw2alpha
    // (render_utils_kernel.cu:653-676, 506-517), with the exclusive suffix of
    // w*rgb taken as the ray total minus the inclusive prefix
    float pr = 0.f, pg = 0.f, pb = 0.f;
    Tc = 1.f;
    for (int base = 0; base < cnt; base += 64) {
        const int j = base + lane;
        const int64_t q = off + j;
        const bool valid = j < cnt && q >= 0 && q < P.capacity;
        float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
        float2 ab = make_float2(0.f, 0.f);
        float e = 0.f;
        float vr = 0.f, vg = 0.f, vb = 0.f;
        if (valid) ab = rec.ab[q];
        const uint32_t bits = __float_as_uint(ab.y);
        const float T = chunk_transmittance(ab.x, bits, lane, Tc), w = T * ab.x;
        if (bits & kAliveW) {
            c = rec.rgb[q];
            vr = w * c.x;
            vg = w * c.y;
            vb = w * c.z;
        }
        if (bits & kAliveA) e = rec.xyze[q].w;
        {  // the same scan as the first pass
            float v3[3] = {vr, vg, vb};
            wave_incl_sum(v3);
            vr = v3[0];
            vg = v3[1];
            vb = v3[2];
        }
        float gl0 = 0.f, gl1 = 0.f, gl2 = 0.f, gw = 0.f, gs = 0.f;
        if (bits & kAliveW) {
            gl0 = w * (g0 + grp * (c.x - t0)) * c.x * (1.f - c.x);
            gl1 = w * (g1 + grp * (c.y - t1)) * c.y * (1.f - c.y);
            gl2 = w * (g2 + grp * (c.z - t2)) * c.z * (1.f - c.z);
            gw = g0 * c.x + g1 * c.y + g2 * c.z;
        }
        if (bits & kAliveA) {
            const float back = gbg * tl + g0 * (cr - (pr + vr)) + g1 * (cg - (pg + vg)) +
                               g2 * (cb - (pb + vb));
            const float ga = gw * T - back / ((1.f - ab.x) + 1e-10f);
            gs = fminf(e, 1e10f) * powf(1.f + e, -P.interval - 1.f) * P.interval * ga;
        }
        if (valid) rec.adj[q] = make_float4(gl0, gl1, gl2, gs);
        pr += wave_read(vr, 63);
        pg += wave_read(vg, 63);
        pb += wave_read(vb, 63);
    }
}

// Composite of the configurations without the MLP (dvgo.py:72-75,344-346: the
// colour grid holds k0_dim = NC channels and rgb = sigmoid(k0)).  One wave per
// ray, as k_ray_composite.  Pass 1 gathers k0 for the weight-kept records (the
// colour rule of colorvdb.cu:16-26,96-103), keeps rgb in the record's
// gradient row and sums w*rgb; the ray colour, loss and dL/drgb follow
// (dvgo.py:365-370, run.py:551-574).  Pass 2 turns each row into the
// sample's gradients, dL/dk0 = dL/drgb * rgb(1 - rgb) and dL/dsigma through
// Alphas2Weights and raw2alpha, which the scatter reads: no backward pass
// over the records sits in between.
template <int NC>
__global__ void __launch_bounds__(kMaxRayWaves * 64) k_direct_composite(StepParams P, TopoView topo,
                                                          const float* __restrict__ color,
                                                          RayBuffers R, Records rec) {
    static_assert(NC == 3 || NC == 12, "direct colour channels");
    constexpr int kSig = kSigmaAt<NC>;
    const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= P.n_rays) return;
    const int lane = lane_id();
    const int64_t off = R.offset[r];
    const int cnt = R.count[r];
    float tg[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) tg[c] = P.target[(int64_t)NC * r + c];
    const bool rgbper = P.rgbper_scale > 0.f;
    float ctot[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) ctot[c] = 0.f;
    float sq = 0.f;  // this lane's sum of w * |rgb - target|^2 (rgbper term)
    float Tc = 1.f;
    for (int base = 0; base < cnt; base += 64) {
        const int j = base + lane;
        const int64_t q = off + j;
        const bool valid = j < cnt && q >= 0 && q < P.capacity;
        float2 ab = make_float2(0.f, 0.f);
        if (valid) ab = rec.ab[q];
        const uint32_t bits = __float_as_uint(ab.y);
        const float w = chunk_transmittance(ab.x, bits, lane, Tc) * ab.x;
        float v[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = 0.f;
        if (bits & kAliveW) {
            const float4 p = rec.xyze[q];
            LeafCache cache;
            float rgb[NC + 1];
            float k0[NC];
            gather_color<NC>(topo, color, make_cell(p.x, p.y, p.z), cache, k0);
#pragma unroll
            for (int c = 0; c < NC; ++c) rgb[c] = 1.f / (1.f + expf(-k0[c]));
            rgb[NC] = 0.f;
            float4* row = rec.grad + 4 * q;
#pragma unroll
            for (int k = 0; k < (NC + 3) / 4; ++k)
                row[k] = make_float4(rgb[4 * k], rgb[4 * k + 1], rgb[4 * k + 2],
                                     4 * k + 3 < NC ? rgb[4 * k + 3] : 0.f);
            float e2 = 0.f;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                v[c] = w * rgb[c];
                const float e = rgb[c] - tg[c];
                e2 += e * e;
            }
            if (rgbper) sq += w * e2;
        }
        wave_incl_sum(v);
#pragma unroll
        for (int c = 0; c < NC; ++c) ctot[c] += wave_read(v[c], 63);
    }
    // the ray's colour, loss and dL/drgb (every lane holds them)
    const float tl = R.t_last[r];
    float g[NC];
    float loss = 0.f, gsum = 0.f;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const float o = ctot[c] + tl * P.bg, dd = o - tg[c];
        g[c] = 2.f * dd * P.loss_scale;
        loss += dd * dd;
        gsum += g[c];
    }
    loss *= P.loss_scale;
    float gent = 0.f;  // dL/dT_last of the entropy term
    if (P.ent_scale > 0.f) {
        const float p = fminf(fmaxf(tl, 1e-6f), 0.999999f);
        const float lp = logf(p), lq = logf(1.f - p);
        loss += -(p * lp + (1.f - p) * lq) * P.ent_scale;
        if (tl >= 1e-6f && tl <= 0.999999f) gent = (lq - lp) * P.ent_scale;
    }
    if (rgbper) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) sq += __shfl_xor(sq, d, 64);
        loss += sq * P.rgbper_scale;
    }
    const float gbg = P.bg * gsum + gent;      // dL/dT_last
    const float grp = 2.f * P.rgbper_scale;    // dL/drgb_i = w_i (g + grp (rgb_i - target))
    if (lane == 0) {
        float* out = R.ray_rgb + (int64_t)(NC + 1) * r;
#pragma unroll
        for (int c = 0; c < NC; ++c) out[c] = ctot[c] + tl * P.bg;
        out[NC] = tl;
        R.ray_loss[r] = loss;
    }
    float pre[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) pre[c] = 0.f;
    Tc = 1.f;
    for (int base = 0; base < cnt; base += 64) {
        const int j = base + lane;
        const int64_t q = off + j;
        const bool valid = j < cnt && q >= 0 && q < P.capacity;
        float2 ab = make_float2(0.f, 0.f);
        if (valid) ab = rec.ab[q];
        const uint32_t bits = __float_as_uint(ab.y);
        const float T = chunk_transmittance(ab.x, bits, lane, Tc), w = T * ab.x;
        float rgb[(NC + 3) / 4 * 4];
#pragma unroll
        for (int c = 0; c < (NC + 3) / 4 * 4; ++c) rgb[c] = 0.f;
        float v[NC];
        float e = 0.f;
        float4* row = rec.grad + 4 * q;
        if (bits & kAliveW) {
#pragma unroll
            for (int k = 0; k < (NC + 3) / 4; ++k) {
                const float4 x = row[k];
                rgb[4 * k] = x.x;
                rgb[4 * k + 1] = x.y;
                rgb[4 * k + 2] = x.z;
                rgb[4 * k + 3] = x.w;
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = w * rgb[c];
        if (bits & kAliveA) e = rec.xyze[q].w;
        wave_incl_sum(v);
        float gl[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) gl[c] = 0.f;
        float gw = 0.f;
        if (bits & kAliveW) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                gl[c] = w * (g[c] + grp * (rgb[c] - tg[c])) * rgb[c] * (1.f - rgb[c]);
                gw += g[c] * rgb[c];
            }
        }
        if (bits & kAliveA) {
            float back = gbg * tl;
#pragma unroll
            for (int c = 0; c < NC; ++c) back += g[c] * (ctot[c] - (pre[c] + v[c]));
            const float ga = gw * T - back / ((1.f - ab.x) + 1e-10f);
            gl[kSig] = fminf(e, 1e10f) * powf(1.f + e, -P.interval - 1.f) * P.interval * ga;
        }
        if (bits & (kAliveA | kAliveW)) {
#pragma unroll
            for (int k = 0; k <= kSig / 4; ++k)
                row[k] = make_float4(gl[4 * k], gl[4 * k + 1], gl[4 * k + 2], gl[4 * k + 3]);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) pre[c] += wave_read(v[c], 63);
    }
}

}  // namespace

namespace pvdb {

int launch_ray_count(const StepParams& P, const TopoView& t, const RayBuffers& R,
                     const PackArgs& pk, hipStream_t s) {
    const int w = P.ray_waves;
    const int ray_blocks = (int)pvdb_blocks(P.n_rays, w);
    // no pack workgroups without an MLP (pk.prm == nullptr)
    const int grid = ray_blocks + (pk.prm ? (int)pvdb_blocks(kPackItems, 64 * w) : 0);
    hipLaunchKernelGGL(k_ray_count, dim3(grid), dim3(64 * w), 0, s, P, t, R, pk, ray_blocks);
    PVDB_LAUNCHED("k_ray_count");
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, R, P.n_rays);
    PVDB_LAUNCHED("k_scan");
    return PVDB_OK;
}

int launch_ray_march(const StepParams& P, const TopoView& t, const float* density,
                     const RayBuffers& R, const Records& rec, hipStream_t s) {
    const int w = P.ray_waves;
    hipLaunchKernelGGL(k_ray_march, dim3(pvdb_blocks(P.n_rays, w)), dim3(64 * w),
                       (unsigned)(w * 128 * sizeof(int64_t)), s, P, t, density,
                       R, rec);
    PVDB_LAUNCHED("k_ray_march");
    return PVDB_OK;
}

int launch_ray_composite(const StepParams& P, const RayBuffers& R, const Records& rec,
                         hipStream_t s) {
    const int w = P.ray_waves;
    hipLaunchKernelGGL(k_ray_composite, dim3(pvdb_blocks(P.n_rays, w)), dim3(64 * w), 0, s, P, R,
                       rec);
    PVDB_LAUNCHED("k_ray_composite");
    return PVDB_OK;
}

int launch_direct_composite(int nc, const StepParams& P, const TopoView& t, const float* color,
                            const RayBuffers& R, const Records& rec, hipStream_t s) {
    PVDB_CHECK_ARG(nc == 3 || nc == 12, "direct colour channels must be 3 or 12");
    const int w = P.ray_waves;
    const dim3 grid(pvdb_blocks(P.n_rays, w));
    if (nc == 3)
        hipLaunchKernelGGL(k_direct_composite<3>, grid, dim3(64 * w), 0, s, P, t, color, R, rec);
    else
        hipLaunchKernelGGL(k_direct_composite<12>, grid, dim3(64 * w), 0, s, P, t, color, R, rec);
    PVDB_LAUNCHED("k_direct_composite");
    return PVDB_OK;
}

}  // namespace pvdb
