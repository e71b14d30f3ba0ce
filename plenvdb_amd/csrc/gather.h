// Trilinear sampling helpers shared by the per-point kernels and the fused
// step: the reference's corner walk, the cell decomposition and a one-entry
// leaf cache (the analogue of the ReadAccessor's leaf-level cache).
#pragma once

#include "pvdb_common.h"

namespace pvdb {

// Corner walk of the reference: 000,001,011,010,110,111,101,100
// (plenvdb/lib/vdb/densityvdb.cu:115-122).  Bit q of each mask is the corner's
// x/y/z offset.
constexpr unsigned kCx = 0xF0u, kCy = 0x3Cu, kCz = 0x66u;

struct Cell {
    int i0, j0, k0;
    float u, v, w;
};

__device__ __forceinline__ Cell make_cell(float x, float y, float z) {
    Cell c;
    c.i0 = (int)floorf(x);
    c.j0 = (int)floorf(y);
    c.k0 = (int)floorf(z);
    c.u = x - (float)c.i0;
    c.v = y - (float)c.j0;
    c.w = z - (float)c.k0;
    return c;
}

// One-entry leaf cache, the analogue of the accessor's leaf-level cache.
struct LeafCache {
    int key0, key1, key2;
    int32_t leaf;
    __device__ __forceinline__ LeafCache() : key0(INT32_MIN), key1(0), key2(0), leaf(-1) {}
    __device__ __forceinline__ int64_t slot(const TopoView& t, int i, int j, int k) {
        const int a = i >> 3, b = j >> 3, c = k >> 3;
        if (a != key0 || b != key1 || c != key2) {
            key0 = a;
            key1 = b;
            key2 = c;
            leaf = pvdb::leaf_of(t, i, j, k);
        }
        return leaf < 0 ? int64_t(-1) : (int64_t)leaf * 512 + pvdb::voxel_offset(i, j, k);
    }
};


// Density at an index-space point, density rule of densityvdb.cu:115-122.
__device__ __forceinline__ float gather_density(const TopoView& t, const float* __restrict__ vals,
                                                const Cell& cell, LeafCache& cache) {
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int dx = (kCx >> q) & 1, dy = (kCy >> q) & 1, dz = (kCz >> q) & 1;
        const float a = dx ? cell.u : 1.f - cell.u;
        const float b = dy ? cell.v : 1.f - cell.v;
        const float d = dz ? cell.w : 1.f - cell.w;
        const int64_t s = cache.slot(t, cell.i0 + dx, cell.j0 + dy, cell.k0 + dz);
        const float val = s >= 0 ? vals[s] : 0.f;
        acc = acc + ((val * a) * b) * d;
    }
    return acc;
}

// NC colour channels at an index-space point, colour rule of
// colorvdb.cu:16-26,96-103: the corner weight (a*b)*c first, then
// out[c] += val[c] * weight, corners in the order above (a missing leaf adds 0).
template <int NC>
__device__ __forceinline__ void gather_color(const TopoView& t, const float* __restrict__ vals,
                                             const Cell& cell, LeafCache& cache,
                                             float (&out)[NC]) {
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int dx = (kCx >> q) & 1, dy = (kCy >> q) & 1, dz = (kCz >> q) & 1;
        const float a = dx ? cell.u : 1.f - cell.u;
        const float b = dy ? cell.v : 1.f - cell.v;
        const float d = dz ? cell.w : 1.f - cell.w;
        const float sc = (a * b) * d;
        const int64_t s = cache.slot(t, cell.i0 + dx, cell.j0 + dy, cell.k0 + dz);
        float v[NC];
        if constexpr (NC % 4 == 0) {  // 16-byte rows
            const float4* src = reinterpret_cast<const float4*>(vals) + (s >= 0 ? s : 0) * (NC / 4);
#pragma unroll
            for (int k = 0; k < NC / 4; ++k) {
                const float4 r = s >= 0 ? src[k] : make_float4(0.f, 0.f, 0.f, 0.f);
                v[4 * k] = r.x;
                v[4 * k + 1] = r.y;
                v[4 * k + 2] = r.z;
                v[4 * k + 3] = r.w;
            }
        } else {
            const float* src = vals + (s >= 0 ? s : 0) * NC;
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c] = s >= 0 ? src[c] : 0.f;
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) out[c] = out[c] + v[c] * sc;
    }
}

}  // namespace pvdb
