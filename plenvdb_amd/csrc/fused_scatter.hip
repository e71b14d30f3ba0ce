// Fused step: grid gradients from the per-record gradients, without float
// atomics to HBM.
//
// The reference adds every sample's 8 corner contributions with atomicAdd
// (densityvdb.cu:3-27,143-181; colorvdb.cu:28-37,130-175), which on a
// multi-XCD device turns into device-scope atomics per value.  Here records are
// binned by the leaf holding their rounded position (the active voxel the
// occupancy test found, so the leaf exists); every corner of such a record
// lies within one voxel of that leaf, in the 10^3 box [-1, 8]^3 around it.
// Bins are cut into chunks of at most kChunk records, so a leaf crossed by
// many rays is spread over many workgroups:
//   k_bin_count   wave-aggregated counts per leaf
//   k_bin_scan    bin and chunk offsets, the chunk table
//   k_bin_fill    record ids in bin order
//   k_chunk_accum workgroup per chunk: its corner contributions summed in an
//                 LDS 10^3 box, written out whole
//   k_leaf_reduce workgroup per (leaf, channel): each voxel sums the boxes of
//                 its leaf's chunks and, on the leaf's faces, the overlapping
//                 boxes of the (up to 26) neighbouring leaves' chunks, in a fixed
//                 order, and adds the sum to the gradient (its only writer)
//   k_scatter_overflow  records whose rounded position has no leaf (only where
//                 the index-space rounding disagrees with the occupancy test's)
//                 keep per-corner atomics
// The "leaf must exist" rule holds: a voxel's sum is formed only for leaves
// that exist.  Within a chunk the sums run in LDS with float atomics, so their
// order is not fixed (as in the reference); across chunks it is.
#include "fused_common.h"
#include "gather.h"
#include "plenvdb_hip.h"

using namespace pvdb;

namespace {

constexpr int kBox = 10 * 10 * 10;  // the leaf's voxels and a one-voxel shell
constexpr int kChunk = 1024;        // records per chunk

__device__ __forceinline__ int box_index(int x, int y, int z) {  // x, y, z in [-1, 8]
    return ((x + 1) * 10 + (y + 1)) * 10 + (z + 1);
}

// Leaf holding the record's rounded position, -1 when none does, -2 when the
// record carries no gradient.
__device__ __forceinline__ int32_t bin_leaf(const TopoView& t, const Records& rec, int64_t q) {
    const uint32_t bits = __float_as_uint(rec.atwr[q].w);
    if (!(bits & (kAliveA | kAliveW))) return -2;
    const float4 p = rec.xyze[q];
    return leaf_of(t, (int)roundf(p.x), (int)roundf(p.y), (int)roundf(p.z));
}

// Position of this lane's item among the items of equal key, via one atomicAdd
// per distinct key in the wave.  Lanes with active == false get -1.
__device__ __forceinline__ int wave_claim(int32_t* counters, int key, bool active) {
    const int lane = threadIdx.x & 63;
    uint64_t pending = __ballot(active);
    int out = -1;
    while (pending) {  // uniform
        const int leader = __ffsll((long long)pending) - 1;
        const int lkey = __shfl(key, leader, 64);
        const bool mine = ((pending >> lane) & 1ull) && key == lkey;
        const uint64_t mm = __ballot(mine);
        int base = 0;
        if (lane == leader) base = atomicAdd(counters + lkey, __popcll(mm));
        base = __shfl(base, leader, 64);
        if (mine) out = base + __popcll(mm & ((1ull << lane) - 1ull));
        pending &= ~mm;
    }
    return out;
}

struct BinBuffers {
    int32_t* count;    // [nleaf] records per bin (zeroed before k_bin_count)
    int32_t* start;    // [nleaf] exclusive record offsets
    int32_t* cursor;   // [nleaf] fill cursors
    int32_t* cstart;   // [nleaf] first chunk of each bin
    int32_t* misc;     // [4]: chunks, overflow count, overflow cursor
    int32_t* binned;   // [capacity] record ids in bin order
    int32_t* overflow; // [capacity] records whose rounded position has no leaf
    int4* chunks;      // [max chunks] (first record position, records, bin, 0)
    float* boxes;      // [max chunks][13][kBox]
};

__global__ void __launch_bounds__(256) k_bin_count(TopoView t, RayBuffers R, Records rec,
                                                   int64_t capacity, BinBuffers B) {
    const int64_t total = min((int64_t)R.total[0], capacity);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t span = (total + 63) / 64 * 64;  // whole waves take part in wave_claim
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < span; q += stride) {
        const int32_t key = q < total ? bin_leaf(t, rec, q) : -2;
        wave_claim(B.count, key < 0 ? 0 : key, key >= 0);
        if (key == -1) atomicAdd(B.misc + 1, 1);
    }
}

// Single workgroup: record offsets, chunk offsets and the chunk table; cursors
// start at the record offsets.
__global__ void __launch_bounds__(1024) k_bin_scan(int64_t nleaf, BinBuffers B) {
    __shared__ int32_t s_cnt[1024], s_chk[1024];
    __shared__ int32_t carry_cnt, carry_chk;
    if (threadIdx.x == 0) carry_cnt = carry_chk = 0;
    __syncthreads();
    for (int64_t b0 = 0; b0 < nleaf; b0 += 1024) {
        const int64_t i = b0 + threadIdx.x;
        const int32_t c = i < nleaf ? B.count[i] : 0;
        const int32_t k = (c + kChunk - 1) / kChunk;
        s_cnt[threadIdx.x] = c;
        s_chk[threadIdx.x] = k;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
            const int32_t vc = threadIdx.x >= (unsigned)off ? s_cnt[threadIdx.x - off] : 0;
            const int32_t vk = threadIdx.x >= (unsigned)off ? s_chk[threadIdx.x - off] : 0;
            __syncthreads();
            s_cnt[threadIdx.x] += vc;
            s_chk[threadIdx.x] += vk;
            __syncthreads();
        }
        if (i < nleaf) {
            const int32_t st = carry_cnt + s_cnt[threadIdx.x] - c;
            const int32_t cs = carry_chk + s_chk[threadIdx.x] - k;
            B.start[i] = st;
            B.cursor[i] = st;
            B.cstart[i] = cs;
            for (int j = 0; j < k; ++j) {
                const int32_t n = c - j * kChunk < kChunk ? c - j * kChunk : kChunk;
                B.chunks[cs + j] = make_int4(st + j * kChunk, n, (int32_t)i, 0);
            }
        }
        __syncthreads();
        if (threadIdx.x == 1023) {
            carry_cnt += s_cnt[1023];
            carry_chk += s_chk[1023];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        B.misc[0] = carry_chk;
        B.misc[2] = 0;
    }
}

__global__ void __launch_bounds__(256) k_bin_fill(TopoView t, RayBuffers R, Records rec,
                                                  int64_t capacity, BinBuffers B) {
    const int64_t total = min((int64_t)R.total[0], capacity);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t span = (total + 63) / 64 * 64;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < span; q += stride) {
        const int32_t key = q < total ? bin_leaf(t, rec, q) : -2;
        const int pos = wave_claim(B.cursor, key < 0 ? 0 : key, key >= 0);
        if (key >= 0) B.binned[pos] = (int32_t)q;
        if (key == -1) B.overflow[atomicAdd(B.misc + 2, 1)] = (int32_t)q;
    }
}

// Workgroup per chunk: the chunk's contributions in an LDS box, written out.
__global__ void __launch_bounds__(256) k_chunk_accum(TopoView t, Records rec, BinBuffers B) {
    __shared__ alignas(16) float acc[13][kBox];
    const int nchunks = B.misc[0];
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int4 ch = B.chunks[c];
        const int32_t leaf = ch.z;
        const int ox = t.origins[3 * leaf], oy = t.origins[3 * leaf + 1],
                  oz = t.origins[3 * leaf + 2];
        for (int v = threadIdx.x; v < 13 * kBox; v += blockDim.x) (&acc[0][0])[v] = 0.f;
        __syncthreads();
        // thread per record: its 8 corners' 13 contributions.  Records are dealt
        // out with a stride so the lanes of one atomic instruction hold samples
        // of different rays (consecutive samples of a ray share corners, and
        // lanes on one LDS address serialise).
        const int per = (ch.y + (int)blockDim.x - 1) / (int)blockDim.x;
        for (int i = 0; i < per; ++i) {
            const int r = (int)threadIdx.x * per + i;
            if (r >= ch.y) break;
            const int64_t q = B.binned[ch.x + r];
            const uint32_t bits = __float_as_uint(rec.atwr[q].w);
            const float4 pt = rec.xyze[q];
            const float4* gp = rec.grad + 4 * q;
            const float4 g0 = gp[0], g1 = gp[1], g2 = gp[2], g3 = gp[3];
            const bool la = (bits & kAliveA) != 0u, lw = (bits & kAliveW) != 0u;
            const Cell cell = make_cell(pt.x, pt.y, pt.z);
            const int bx = cell.i0 - ox, by = cell.j0 - oy, bz = cell.k0 - oz;  // in [-1, 7]
#pragma unroll
            for (int corner = 0; corner < 8; ++corner) {
                const int dx = (kCx >> corner) & 1, dy = (kCy >> corner) & 1,
                          dz = (kCz >> corner) & 1;
                const int vox = box_index(bx + dx, by + dy, bz + dz);
                const float a_ = dx ? cell.u : 1.f - cell.u;
                const float b_ = dy ? cell.v : 1.f - cell.v;
                const float d_ = dz ? cell.w : 1.f - cell.w;
                if (la) atomicAdd(&acc[12][vox], ((g3.x * a_) * b_) * d_);
                if (lw) {
                    const float sc = (a_ * b_) * d_;
                    atomicAdd(&acc[0][vox], g0.x * sc);
                    atomicAdd(&acc[1][vox], g0.y * sc);
                    atomicAdd(&acc[2][vox], g0.z * sc);
                    atomicAdd(&acc[3][vox], g0.w * sc);
                    atomicAdd(&acc[4][vox], g1.x * sc);
                    atomicAdd(&acc[5][vox], g1.y * sc);
                    atomicAdd(&acc[6][vox], g1.z * sc);
                    atomicAdd(&acc[7][vox], g1.w * sc);
                    atomicAdd(&acc[8][vox], g2.x * sc);
                    atomicAdd(&acc[9][vox], g2.y * sc);
                    atomicAdd(&acc[10][vox], g2.z * sc);
                    atomicAdd(&acc[11][vox], g2.w * sc);
                }
            }
        }
        __syncthreads();
        float4* out = reinterpret_cast<float4*>(B.boxes + (int64_t)c * 13 * kBox);
        const float4* src = reinterpret_cast<const float4*>(&acc[0][0]);
        for (int e = threadIdx.x; e < 13 * kBox / 4; e += blockDim.x) out[e] = src[e];
        __syncthreads();
    }
}

// Workgroup per (leaf, channel), thread per voxel: the leaf's chunks plus the
// overlapping faces of the neighbours' chunks, summed in chunk order.
// Neighbour d = (dx+1)*9 + (dy+1)*3 + (dz+1) sits at origin + 8*(dx,dy,dz);
// voxel v of this leaf is in its box iff v_a = 0 where d_a = -1 and v_a = 7
// where d_a = +1, at box coordinate v - 8*d.
__global__ void __launch_bounds__(512) k_leaf_reduce(TopoView t, BinBuffers B,
                                                     float* __restrict__ grad_density,
                                                     float* __restrict__ grad_color,
                                                     uint8_t* __restrict__ leaf_touched) {
    __shared__ int32_t ncs[27], ncn[27];
    __shared__ int any;
    const int64_t work = t.nleaf * 13;
    for (int64_t wi = blockIdx.x; wi < work; wi += gridDim.x) {
        const int64_t leaf = wi / 13;
        const int chn = (int)(wi - leaf * 13);
        if (threadIdx.x == 0) any = 0;
        __syncthreads();
        if (threadIdx.x < 27) {
            const int d = threadIdx.x;
            const int ddx = d / 9 - 1, ddy = (d / 3) % 3 - 1, ddz = d % 3 - 1;
            const int32_t nl = d == 13 ? (int32_t)leaf
                                       : leaf_of(t, t.origins[3 * leaf] + 8 * ddx,
                                                 t.origins[3 * leaf + 1] + 8 * ddy,
                                                 t.origins[3 * leaf + 2] + 8 * ddz);
            int32_t cs = 0, cn = 0;
            if (nl >= 0) {
                cs = B.cstart[nl];
                cn = (B.count[nl] + kChunk - 1) / kChunk;
            }
            ncs[d] = cs;
            ncn[d] = cn;
            if (cn > 0) atomicOr(&any, 1);
        }
        __syncthreads();
        if (any) {  // uniform
            const int v = threadIdx.x;
            const int vx = v >> 6, vy = (v >> 3) & 7, vz = v & 7;
            const int ax = vx == 0 ? -1 : (vx == 7 ? 1 : 0);
            const int ay = vy == 0 ? -1 : (vy == 7 ? 1 : 0);
            const int az = vz == 0 ? -1 : (vz == 7 ? 1 : 0);
            float sum = 0.f;
            for (int m = 0; m < 8; ++m) {  // the leaf itself (m = 0) and the face subsets
                if (((m & 4) && !ax) || ((m & 2) && !ay) || ((m & 1) && !az)) continue;
                const int ddx = (m & 4) ? ax : 0, ddy = (m & 2) ? ay : 0, ddz = (m & 1) ? az : 0;
                const int d = (ddx + 1) * 9 + (ddy + 1) * 3 + (ddz + 1);
                const int cn = ncn[d];
                if (cn == 0) continue;
                const float* bx = B.boxes + (int64_t)ncs[d] * 13 * kBox + (int64_t)chn * kBox +
                                  box_index(vx - 8 * ddx, vy - 8 * ddy, vz - 8 * ddz);
#pragma unroll 4
                for (int j = 0; j < cn; ++j) sum += bx[(int64_t)j * 13 * kBox];
            }
            if (sum != 0.f) {
                const int64_t slot = leaf * 512 + v;
                if (chn == 12)
                    grad_density[slot] += sum;
                else
                    grad_color[slot * 12 + chn] += sum;
            }
            if (threadIdx.x == 0 && chn == 0) leaf_touched[leaf] = 1;
        }
        __syncthreads();
    }
}

// Records whose rounded position has no leaf: per-corner atomics with the
// "leaf must exist" rule.
__global__ void __launch_bounds__(256) k_scatter_overflow(TopoView topo, Records rec, BinBuffers B,
                                                          float* __restrict__ grad_density,
                                                          float* __restrict__ grad_color,
                                                          uint8_t* __restrict__ leaf_touched) {
    const int n = B.misc[1];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 8 * n; t += gridDim.x * blockDim.x) {
        const int64_t q = B.overflow[t >> 3];
        const int corner = t & 7;
        const uint32_t bits = __float_as_uint(rec.atwr[q].w);
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

__global__ void k_bin_reset(int32_t* count, int64_t nleaf, int32_t* misc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nleaf) count[i] = 0;
    if (i < 4) misc[i] = 0;
}

}  // namespace

namespace pvdb {

static int64_t round256(int64_t b) { return (b + 255) & ~int64_t(255); }

static int64_t max_chunks(int64_t nleaf, int64_t capacity) {
    // every bin with records has a partial chunk at most; the rest are full
    return (capacity + kChunk - 1) / kChunk + (nleaf < capacity ? nleaf : capacity);
}

int64_t scatter_workspace_bytes(int64_t nleaf, int64_t capacity) {
    const int64_t mc = max_chunks(nleaf, capacity);
    return 4 * round256(4 * nleaf) + round256(16) + 2 * round256(4 * capacity) +
           round256(16 * mc) + round256(4 * 13 * kBox * mc);
}

int launch_scatter(const TopoView& t, const RayBuffers& R, const Records& rec, int64_t capacity,
                   float* grad_density, float* grad_color, uint8_t* leaf_touched, void* ws,
                   hipStream_t s) {
    const int64_t nleaf = t.nleaf;
    if (nleaf == 0) return PVDB_OK;
    uint8_t* p = reinterpret_cast<uint8_t*>(ws);
    auto take = [&](int64_t bytes) {
        uint8_t* at = p;
        p += round256(bytes);
        return at;
    };
    const int64_t mc = max_chunks(nleaf, capacity);
    BinBuffers B;
    B.count = reinterpret_cast<int32_t*>(take(4 * nleaf));
    B.start = reinterpret_cast<int32_t*>(take(4 * nleaf));
    B.cursor = reinterpret_cast<int32_t*>(take(4 * nleaf));
    B.cstart = reinterpret_cast<int32_t*>(take(4 * nleaf));
    B.misc = reinterpret_cast<int32_t*>(take(16));
    B.binned = reinterpret_cast<int32_t*>(take(4 * capacity));
    B.overflow = reinterpret_cast<int32_t*>(take(4 * capacity));
    B.chunks = reinterpret_cast<int4*>(take(16 * mc));
    B.boxes = reinterpret_cast<float*>(take(4 * 13 * kBox * mc));

    hipLaunchKernelGGL(k_bin_reset, dim3(pvdb_blocks(nleaf > 4 ? nleaf : 4, 256)), dim3(256), 0,
                       s, B.count, nleaf, B.misc);
    PVDB_LAUNCHED("k_bin_reset");
    const int64_t want = (capacity + 255) / 256;
    const unsigned grid = (unsigned)(want < 2048 ? (want < 1 ? 1 : want) : 2048);
    hipLaunchKernelGGL(k_bin_count, dim3(grid), dim3(256), 0, s, t, R, rec, capacity, B);
    PVDB_LAUNCHED("k_bin_count");
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, nleaf, B);
    PVDB_LAUNCHED("k_bin_scan");
    hipLaunchKernelGGL(k_bin_fill, dim3(grid), dim3(256), 0, s, t, R, rec, capacity, B);
    PVDB_LAUNCHED("k_bin_fill");
    const unsigned cgrid = (unsigned)(mc < 4096 ? mc : 4096);
    hipLaunchKernelGGL(k_chunk_accum, dim3(cgrid), dim3(256), 0, s, t, rec, B);
    PVDB_LAUNCHED("k_chunk_accum");
    const int64_t rw = nleaf * 13;
    const unsigned rgrid = (unsigned)(rw < 65536 ? rw : 65536);
    hipLaunchKernelGGL(k_leaf_reduce, dim3(rgrid), dim3(512), 0, s, t, B, grad_density,
                       grad_color, leaf_touched);
    PVDB_LAUNCHED("k_leaf_reduce");
    hipLaunchKernelGGL(k_scatter_overflow, dim3(256), dim3(256), 0, s, t, rec, B, grad_density,
                       grad_color, leaf_touched);
    PVDB_LAUNCHED("k_scatter_overflow");
    return PVDB_OK;
}

}  // namespace pvdb
