// Fused step: grid gradients from the per-record gradients.
//
// The reference adds every sample's 8 corner contributions with atomicAdd
// (densityvdb.cu:3-27,143-181; colorvdb.cu:28-37,130-175): 8 x 13 device-scope
// float atomics per sample.  Here a workgroup takes 128 consecutive records
// (a few rays, which share most of their corners), sums their contributions
// per slot in an LDS hash table, and adds each distinct slot's 13 sums to the
// grids with one atomic each (k_scatter_agg).  k_scatter_direct is the
// per-corner form, for topologies too large for the hash key.
//
// An atomic-free variant was measured as well: records binned by leaf, summed
// per chunk of a bin in an LDS box, the boxes reduced per leaf.  On the 160^3
// training workload, whose rays all cross the centre, the chunk sums were
// bound by LDS atomics on a few hot leaves (0.37 ms against 0.18 ms here), and
// the result was no more reproducible than this one (LDS float atomics within
// a chunk), so it is not kept.
#include "fused_common.h"
#include "gather.h"
#include "plenvdb_hip.h"

using namespace pvdb;

namespace {

// Gradients into the grids without binning: one thread per (record, corner), with the
// "leaf must exist" rule (densityvdb.cu:16-27, colorvdb.cu:28-37); the density
// term keeps the reference's ((g*a)*b)*c order.  Touched leaves are flagged
// for the sparse Adam pass.  NC colour channels per slot (12 features, or a
// direct-colour grid's 3).
template <int NC>
__global__ void __launch_bounds__(256) k_scatter_direct(TopoView topo, RayBuffers R, Records rec,
                                                        int64_t capacity, GradRows G,
                                                        uint8_t* __restrict__ leaf_touched) {
    const int64_t total = min((int64_t)R.total[0], capacity);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total * 8;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = t >> 3;
        const int corner = (int)(t & 7);
        const uint32_t bits = __float_as_uint(rec.ab[q].y);
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
        const float* g = reinterpret_cast<const float*>(rec.grad + 4 * q);
        if (bits & kAliveA) atomicAdd(G.density + slot * G.ds, ((g[kSigmaAt<NC>] * a) * b) * d);
        if (bits & kAliveW) {
            const float sc = (a * b) * d;
            float* dst = G.color + slot * G.cs;
#pragma unroll
            for (int c = 0; c < NC; ++c) atomicAdd(dst + c, g[c] * sc);
        }
    }
}

// Same result as k_scatter_direct with far fewer global atomics: a workgroup takes
// kAggRecords consecutive records (two or three rays; consecutive samples of a
// ray share most of their corners) and adds each distinct slot's 13 sums to the
// grids once.  Within the chunk, every (record, corner) pair finds its slot's
// entry in an LDS hash table (compare-and-swap on the 32-bit slot key), stores
// its 13 contributions with plain stores, and links itself into the entry's
// list (one exchange); the flush gives each listed entry 16 lanes, which walk
// the list and sum channel c in lane c.  Two LDS atomics per pair instead of
// fourteen: LDS float atomics run at well under one lane per clock.  Every
// record load is issued before any is used, and a chunk's records are loaded
// during the previous chunk's flush.
// kAggRecords (32 or 64, a launch choice) records per chunk; NC colour channels
// per slot (12 features, or a direct-colour grid's 3).
constexpr uint32_t kEmptyKey = 0xffffffffu;
// component k of a float4 (k a compile-time constant after unrolling)
__device__ __forceinline__ float lane4(const float4& v, int k) {
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}
constexpr uint32_t kEnd = 0xffffu;  // end of a pair list
// The launch's first `red.blocks` workgroups sum the MLP's dW partials
// instead (dw_reduce_block, independent of the scatter), and the next one the
// loss when red.ray_loss is set: they overlap the scatter's atomics rather than
// following them in launches of their own.
template <int kAggRecords, int NC>
__global__ void __launch_bounds__(256) k_scatter_agg(TopoView topo, RayBuffers R, Records rec,
                                                     int64_t capacity, GradRows G,
                                                     uint8_t* __restrict__ leaf_touched,
                                                     ReduceArgs red) {
    constexpr int kAggPairsAll = 8 * kAggRecords;  // (record, corner) pairs of a chunk
    constexpr int kAggSlots = kAggPairsAll;        // every pair fits: probing ends
    constexpr int kAggLog2 = kAggRecords == 64 ? 9 : 8;  // kAggSlots == 1 << kAggLog2
    constexpr int kAggPairs = kAggPairsAll / 256;  // pairs per thread
    static_assert(kAggSlots == 1 << kAggLog2 && kAggPairs >= 1, "chunk size");
    constexpr int kRow = NC + 1;                   // a pair's contributions: colour, density
    constexpr int kG = NC == 12 ? 4 : 1;           // float4s of a record's gradient row read
    constexpr int kSig = kSigmaAt<NC>;
    constexpr int kLanes = NC == 12 ? 16 : 4;      // flush lanes per entry
    constexpr int kTouchLane = NC == 12 ? 13 : NC; // the lane that flags the leaf
    __shared__ uint32_t key[kAggSlots];
    __shared__ uint32_t head[kAggSlots];          // first pair of the entry's list
    __shared__ uint16_t next[kAggPairsAll];       // next pair of the same entry
    __shared__ __align__(16) float pv[kAggPairsAll * kRow];  // each pair's contributions
    __shared__ uint16_t listed[kAggSlots];        // the entries in use
    __shared__ int nlisted[2];
    static_assert(kAggPairsAll * kRow >= 4 * kReduceGrp * kReduceVec, "reduction scratch");
    if ((int)blockIdx.x < red.blocks) {  // uniform over the workgroup
        dw_reduce_block(blockIdx.x, red.part, red.nparts, red.dw, red.adam,
                        reinterpret_cast<float4*>(pv));
        return;
    }
    const int lead = red.blocks + (red.ray_loss ? 1 : 0);  // workgroups before the scatter's
    if ((int)blockIdx.x < lead) {  // the loss workgroup
        if (threadIdx.x < 64) loss_sum_wave(red.ray_loss, red.n_rays, red.loss);
        return;
    }
    const int wg = (int)blockIdx.x - lead, nwg = (int)gridDim.x - lead;
    const int tid = threadIdx.x;
    for (int e = tid; e < kAggSlots; e += 256) {
        key[e] = kEmptyKey;
        head[e] = kEnd;
    }
    if (tid < 2) nlisted[tid] = 0;
    __syncthreads();
    const int64_t total = min((int64_t)R.total[0], capacity);
    const int64_t nchunk = (total + kAggRecords - 1) / kAggRecords;
    // this thread's pairs of a chunk: the records are loaded one chunk ahead
    // (issued before the previous chunk's flush)
    uint32_t bits[kAggPairs];
    float4 pt[kAggPairs], gr[kAggPairs][kG];
    auto load = [&](int64_t chunk) {
#pragma unroll
        for (int it = 0; it < kAggPairs; ++it) {
            const int64_t q = chunk * kAggRecords + ((tid + 256 * it) >> 3);
            bits[it] = 0u;
            if (chunk < nchunk && q < total) {
                bits[it] = __float_as_uint(rec.ab[q].y);
                pt[it] = rec.xyze[q];
#pragma unroll
                for (int k = 0; k < kG; ++k) gr[it][k] = rec.grad[4 * q + k];
            }
        }
    };
    load(wg);
    int b = 0;
    for (int64_t chunk = wg; chunk < nchunk; chunk += nwg, b ^= 1) {
        // ---- topology of this thread's pairs ----
        int32_t leaf[kAggPairs];
        uint32_t off[kAggPairs];
        float wgt[kAggPairs][3];
#pragma unroll
        for (int it = 0; it < kAggPairs; ++it) {
            const int corner = (tid + 256 * it) & 7;  // 256 is a multiple of 8
            leaf[it] = -1;
            off[it] = 0u;
            wgt[it][0] = wgt[it][1] = wgt[it][2] = 0.f;
            if (bits[it] & (kAliveA | kAliveW)) {
                const Cell cell = make_cell(pt[it].x, pt[it].y, pt[it].z);
                const int dx = (kCx >> corner) & 1, dy = (kCy >> corner) & 1,
                          dz = (kCz >> corner) & 1;
                const int i = cell.i0 + dx, j = cell.j0 + dy, k = cell.k0 + dz;
                leaf[it] = leaf_of(topo, i, j, k);
                off[it] = (uint32_t)voxel_offset(i, j, k);
                wgt[it][0] = dx ? cell.u : 1.f - cell.u;
                wgt[it][1] = dy ? cell.v : 1.f - cell.v;
                wgt[it][2] = dz ? cell.w : 1.f - cell.w;
            }
        }
        // ---- contributions into LDS, pairs linked per slot ----
#pragma unroll
        for (int it = 0; it < kAggPairs; ++it) {
            if (leaf[it] < 0) continue;
            const int pair = tid + 256 * it;
            const float a = wgt[it][0], bb = wgt[it][1], d = wgt[it][2];
            float* v = pv + pair * kRow;
            const float sc = (a * bb) * d;
            const bool w_ = (bits[it] & kAliveW) != 0u;
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c] = w_ ? lane4(gr[it][c >> 2], c & 3) * sc : 0.f;
            v[NC] = (bits[it] & kAliveA) ? ((lane4(gr[it][kSig >> 2], kSig & 3) * a) * bb) * d
                                         : 0.f;
            const uint32_t s = (uint32_t)leaf[it] * 512u + off[it];
            uint32_t h = (s * 2654435761u) >> (32 - kAggLog2);
            while (true) {
                const uint32_t prev = atomicCAS(&key[h], kEmptyKey, s);
                if (prev == kEmptyKey) {
                    listed[atomicAdd(&nlisted[b], 1)] = (uint16_t)h;
                    break;
                }
                if (prev == s) break;
                h = (h + 1) & (kAggSlots - 1);
            }
            next[pair] = (uint16_t)atomicExch(&head[h], (uint32_t)pair);
        }
        __syncthreads();
        load(chunk + nwg);  // the next chunk's records, in flight during the flush
        // ---- flush: walk each listed entry, add its sums, clear it; reset the
        // other list counter.  16 lanes per entry for 12 channels, so one atomic
        // instruction covers four slots' contiguous 12-float colour rows (fewer
        // cache lines than narrower groups, measured); 4 lanes for 3 ----
        const int n = nlisted[b];
        if (tid == 0) nlisted[b ^ 1] = 0;
        const int c = tid & (kLanes - 1);
        for (int e = tid / kLanes; e < n; e += 256 / kLanes) {
            const int h = listed[e];
            const uint32_t s = key[h];
            float sum = 0.f;
            for (uint32_t p = head[h]; p != kEnd; p = next[p])  // same walk in the entry's lanes
                if (c < kRow) sum += pv[p * kRow + c];
            if (c < kRow && sum != 0.f) {  // adding zero changes nothing
                if (c < NC)
                    atomicAdd(G.color + (int64_t)s * G.cs + c, sum);
                else
                    atomicAdd(G.density + (int64_t)s * G.ds, sum);
            }
            if (c == kTouchLane) leaf_touched[s >> 9] = 1;
        }
        __syncthreads();  // every walk is done: clear the entries
        for (int e = tid; e < n; e += 256) {
            const int h = listed[e];
            key[h] = kEmptyKey;
            head[h] = kEnd;
        }
        __syncthreads();
    }
}

// Loss = sum of the per-ray terms, for the launches that carry no loss
// workgroup (one wavefront, the order of loss_sum_wave).
__global__ void __launch_bounds__(64) k_loss_reduce(const float* ray_loss, int n, float* loss) {
    loss_sum_wave(ray_loss, n, loss);
}

}  // namespace

namespace pvdb {

int launch_dw_reduce(const float* part, int nparts, float* dw, const MlpAdam& adam,
                     hipStream_t s);

// The reductions riding on a scatter launch, as launches of their own (the
// scatter paths that do not carry them).
int launch_reductions(const ReduceArgs& red, hipStream_t s) {
    if (red.ray_loss) {
        hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(64), 0, s, red.ray_loss, red.n_rays,
                           red.loss);
        PVDB_LAUNCHED("k_loss_reduce");
    }
    return red.blocks ? launch_dw_reduce(red.part, red.nparts, red.dw, red.adam, s) : PVDB_OK;
}

template <int NC>
int launch_scatter_nc(const TopoView& t, const RayBuffers& R, const Records& rec, int64_t capacity,
                      const GradRows& G, uint8_t* leaf_touched, int grid_req,
                      int chunk_records, const ReduceArgs& red_in, hipStream_t s) {
    if (t.nleaf == 0) return launch_reductions(red_in, s);
    const ReduceArgs red = red_in;
    if (t.nleaf < (int64_t(1) << 23) - 1) {  // every slot < 2^32 - 1, the empty key
        PVDB_CHECK_ARG(chunk_records == 32 || chunk_records == 64,
                       "launch[2]: scatter chunk must be 32 or 64 records");
        const int64_t chunks = (capacity + chunk_records - 1) / chunk_records;
        // each workgroup takes chunks in a grid-stride loop; four (64-record) or
        // eight (32-record) workgroups fit a CU by LDS
        const int64_t want = grid_req > 0 ? grid_req : 1;
        const unsigned grid = (unsigned)(chunks < want ? (chunks < 1 ? 1 : chunks) : want) +
                              (unsigned)red.blocks + (red.ray_loss ? 1u : 0u);
        if (chunk_records == 64)
            hipLaunchKernelGGL((k_scatter_agg<64, NC>), dim3(grid), dim3(256), 0, s, t, R, rec,
                               capacity, G, leaf_touched, red);
        else
            hipLaunchKernelGGL((k_scatter_agg<32, NC>), dim3(grid), dim3(256), 0, s, t, R, rec,
                               capacity, G, leaf_touched, red);
        PVDB_LAUNCHED("k_scatter_agg");
        return PVDB_OK;
    }
    // enough threads for a few (record, corner) pairs each at the largest sizes
    const int64_t want = (capacity * 8 + 255) / 256;
    const unsigned grid = (unsigned)(want < 16384 ? (want < 1 ? 1 : want) : 16384);
    hipLaunchKernelGGL(k_scatter_direct<NC>, dim3(grid), dim3(256), 0, s, t, R, rec, capacity, G,
                       leaf_touched);
    PVDB_LAUNCHED("k_scatter_direct");
    return launch_reductions(red, s);
}

// nc: colour channels per slot, 12 (the MLP's features) or 3 (a direct-colour grid)
int launch_scatter(int nc, const TopoView& t, const RayBuffers& R, const Records& rec,
                   int64_t capacity, const GradRows& G, uint8_t* leaf_touched, int grid_req,
                   int chunk_records, const ReduceArgs& red, hipStream_t s) {
    PVDB_CHECK_ARG(nc == 3 || nc == 12, "colour channels must be 3 or 12");
    return nc == 12 ? launch_scatter_nc<12>(t, R, rec, capacity, G, leaf_touched, grid_req,
                                            chunk_records, red, s)
                    : launch_scatter_nc<3>(t, R, rec, capacity, G, leaf_touched, grid_req,
                                           chunk_records, red, s);
}

}  // namespace pvdb
