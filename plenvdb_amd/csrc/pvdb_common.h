// Shared definitions of the PlenVDB MI355X library: the flattened VDB topology
// as the kernels see it, error plumbing of the C-ABI and small device helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

// ---------------------------------------------------------------------------
// Flattened topology.  The reference walks a NanoVDB root/upper/lower/leaf
// tree through a per-thread ReadAccessor (openvdb/nanovdb/nanovdb/NanoVDB.h:
// 4514-4680).  Here the tree is flattened once on the host into
//   * a dense leaf table over the bounding box of the leaf grid (one int32 per
//     8^3 block: the leaf's index in NanoVDB depth-first order, or -1), which
//     replaces the root scan plus the upper/lower child lookups by one load
//     that stays L2-resident (40^3 * 4 B = 256 KB at 320^3);
//   * when that table would exceed 2^24 cells (leaves spread far apart), a
//     two-level layout instead: a dense table of 128^3 lower nodes and a
//     16^3 child table per lower node;
//   * a dense lower-node table in both layouts (the upper node's child mask),
//     used by the marchers to skip empty 128^3 blocks;
//   * in the one-level layout, an empty-run map: per 8^3 cell without a leaf,
//     the radius (in cells, up to 64) of the empty cube of cells around it,
//     which lets the marchers step over empty space in one jump;
//   * per-leaf 512-bit value masks (NanoVDB LeafNode::mValueMask);
//   * per-leaf origins.
// Leaf indices and the in-leaf offset (x&7)<<6 | (y&7)<<3 | (z&7) equal the
// reference's (NanoVDB.h:3893-3900), so a payload slot leaf*512+offset is the
// reference's `n` in densityvdb.cu:41-42.
// ---------------------------------------------------------------------------
namespace pvdb {

struct TopoView {
    const int32_t* table;        // one-level: [d0*d1*d2] leaf index or -1
    const uint64_t* masks;       // [nleaf*8] value masks, bit (offset & 63) of word offset>>6
    const int32_t* origins;      // [nleaf*3]
    const int32_t* lower_table;  // [e0*e1*e2] lower-node (128^3 block) index or -1
    const int32_t* lower_leaves; // two-level: [nlower*4096] leaf index per 8^3 block or -1
    const uint8_t* erun;         // one-level: [d0*d1*d2] empty-run map (see empty_run), else null
    int32_t lo0, lo1, lo2;       // one-level: leaf-grid origin (voxel >> 3)
    int32_t d0, d1, d2;          // one-level: leaf-grid extent
    int32_t ulo0, ulo1, ulo2;    // lower-node grid origin (voxel >> 7)
    int32_t e0, e1, e2;          // lower-node grid extent
    int32_t two_level;           // 1: look leaves up through lower_table + lower_leaves
    int32_t pad_;
    int64_t nleaf;
};

__device__ __forceinline__ int voxel_offset(int i, int j, int k) {
    return ((i & 7) << 6) | ((j & 7) << 3) | (k & 7);
}

// Index of the 128^3 lower node holding voxel (i,j,k), or -1.
__device__ __forceinline__ int32_t lower_of(const TopoView& t, int i, int j, int k) {
    const int a = (i >> 7) - t.ulo0, b = (j >> 7) - t.ulo1, c = (k >> 7) - t.ulo2;
    if ((unsigned)a >= (unsigned)t.e0 || (unsigned)b >= (unsigned)t.e1 ||
        (unsigned)c >= (unsigned)t.e2)
        return -1;
    return t.lower_table[(a * t.e1 + b) * t.e2 + c];
}

// Leaf index of voxel (i,j,k) or -1 when no leaf holds it.  One-level layout:
// one load from the dense leaf table.  Two-level layout: the lower node's
// child table, indexed like NanoVDB's InternalNode::CoordToOffset
// (openvdb/nanovdb/nanovdb/NanoVDB.h:3272-3283).
__device__ __forceinline__ int32_t leaf_of(const TopoView& t, int i, int j, int k) {
    if (!t.two_level) {
        const int a = (i >> 3) - t.lo0, b = (j >> 3) - t.lo1, c = (k >> 3) - t.lo2;
        if ((unsigned)a >= (unsigned)t.d0 || (unsigned)b >= (unsigned)t.d1 ||
            (unsigned)c >= (unsigned)t.d2)
            return -1;
        return t.table[(a * t.d1 + b) * t.d2 + c];
    }
    const int32_t low = lower_of(t, i, j, k);
    if (low < 0) return -1;
    const int off = (((i & 127) >> 3) << 8) | (((j & 127) >> 3) << 4) | ((k & 127) >> 3);
    return t.lower_leaves[(int64_t)low * 4096 + off];
}

// Slot (leaf*512 + offset) of voxel (i,j,k) or -1 when its leaf does not exist.
__device__ __forceinline__ int64_t slot_of(const TopoView& t, int i, int j, int k) {
    const int32_t l = leaf_of(t, i, j, k);
    return l < 0 ? int64_t(-1) : (int64_t)l * 512 + voxel_offset(i, j, k);
}

// For the 8^3 cell L = voxel >> 3 that holds no leaf: r >= 1 such that no
// leaf lies within r - 1 cells of L (Chebyshev distance).  Outside the leaf
// grid's box that is L's distance to the box; inside it the empty-run map's
// entry; 1 without a map (two-level layout).
__device__ __forceinline__ int empty_run(const TopoView& t, int L0, int L1, int L2) {
    if (!t.erun) return 1;
    const int a = L0 - t.lo0, b = L1 - t.lo1, c = L2 - t.lo2;
    const int oa = a < 0 ? -a : (a >= t.d0 ? a - t.d0 + 1 : 0);
    const int ob = b < 0 ? -b : (b >= t.d1 ? b - t.d1 + 1 : 0);
    const int oc = c < 0 ? -c : (c >= t.d2 ? c - t.d2 + 1 : 0);
    const int out = max(oa, max(ob, oc));
    if (out > 0) return out;
    return max(1, (int)t.erun[((int64_t)a * t.d1 + b) * t.d2 + c]);
}

// NanoVDB isActive for a tree without active tiles: the value-mask bit of an
// existing leaf, false otherwise.
__device__ __forceinline__ bool voxel_active(const TopoView& t, int i, int j, int k) {
    const int32_t l = leaf_of(t, i, j, k);
    if (l < 0) return false;
    const int off = voxel_offset(i, j, k);
    return (t.masks[(int64_t)l * 8 + (off >> 6)] >> (off & 63)) & 1ull;
}

}  // namespace pvdb

// ---------------------------------------------------------------------------
// C-ABI error plumbing
// ---------------------------------------------------------------------------
enum {
    PVDB_OK = 0,
    PVDB_EINVAL = -1,
    PVDB_EHIP = -2,
    PVDB_EHANDLE = -3,
    PVDB_ENOMEM = -4,
    PVDB_ERANGE = -5,
};

namespace pvdb {
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
}  // namespace pvdb

#define PVDB_CHECK_ARG(cond, msg)          \
    do {                                    \
        if (!(cond)) {                      \
            pvdb::set_error(msg);           \
            return PVDB_EINVAL;             \
        }                                   \
    } while (0)

#define PVDB_HIP(call)                                         \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return pvdb::hip_fail(e_, #call); \
    } while (0)

#define PVDB_LAUNCHED(name)                                            \
    do {                                                               \
        hipError_t e_ = hipGetLastError();                             \
        if (e_ != hipSuccess) return pvdb::hip_fail(e_, "launch " name); \
    } while (0)

static inline unsigned pvdb_blocks(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : b);
}
