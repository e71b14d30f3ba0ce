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
// for the sparse Adam pass.
__global__ void __launch_bounds__(256) k_scatter_direct(TopoView topo, RayBuffers R, Records rec,
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

// Same result as k_scatter_direct with far fewer global atomics: a workgroup takes
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

}  // namespace

namespace pvdb {

int launch_scatter(const TopoView& t, const RayBuffers& R, const Records& rec, int64_t capacity,
                   float* grad_density, float* grad_color, uint8_t* leaf_touched, hipStream_t s) {
    if (t.nleaf == 0) return PVDB_OK;
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
    hipLaunchKernelGGL(k_scatter_direct, dim3(grid), dim3(256), 0, s, t, R, rec, capacity,
                       grad_density, grad_color, leaf_touched);
    PVDB_LAUNCHED("k_scatter_direct");
    return PVDB_OK;
}

}  // namespace pvdb
