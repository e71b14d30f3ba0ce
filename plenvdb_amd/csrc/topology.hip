// Host-side VDB topology builder and its C-ABI.
//
// Replaces the reference's OpenVDB denseFill -> openToNanoVDB -> deviceUpload
// pipeline (plenvdb/lib/vdb/plenvdb.h:117-125,197-210) by a direct build of
// the flattened device layout described in pvdb_common.h.  Leaves are ordered
// exactly as NanoVDB orders them (depth first: upper nodes in lexicographic
// Coord order of their origins, then lower nodes by InternalNode::CoordToOffset,
// then leaves by CoordToOffset; openvdb/nanovdb/nanovdb/util/GridBuilder.h:362-376,
// util/OpenToNanoVDB.h:509-533, NanoVDB.h:932-935,3272-3283).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "pvdb_common.h"
#include "plenvdb_hip.h"

namespace pvdb {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return PVDB_EHIP;
}

}  // namespace pvdb

struct pvdb_topology {
    int64_t nleaf = 0;
    int64_t nactive = 0;
    int64_t nlower = 0;
    int two_level = 0;
    int32_t lo[3] = {0, 0, 0};    // leaf grid (voxel >> 3)
    int32_t dim[3] = {0, 0, 0};
    int32_t ulo[3] = {0, 0, 0};   // lower-node grid (voxel >> 7)
    int32_t udim[3] = {0, 0, 0};
    std::vector<int32_t> origins;  // host copies
    std::vector<uint64_t> masks;
    std::vector<int32_t> table;         // one-level leaf table
    std::vector<int32_t> lower_table;   // lower-node table
    std::vector<int32_t> lower_leaves;  // two-level child tables
    std::vector<uint8_t> erun;          // one-level empty-run map
    int32_t* d_table = nullptr;
    uint64_t* d_masks = nullptr;
    int32_t* d_origins = nullptr;
    int32_t* d_lower = nullptr;
    int32_t* d_lower_leaves = nullptr;
    uint8_t* d_erun = nullptr;
    int device = 0;
};

namespace {

struct LeafRec {
    int32_t o[3];
    uint64_t m[8];
};

// NanoVDB depth-first order of leaf origins.
bool nano_less(const LeafRec& a, const LeafRec& b) {
    int32_t ua[3], ub[3];
    for (int d = 0; d < 3; ++d) {
        ua[d] = a.o[d] & ~4095;
        ub[d] = b.o[d] & ~4095;
    }
    for (int d = 0; d < 3; ++d)
        if (ua[d] != ub[d]) return ua[d] < ub[d];  // Coord::operator< on upper origins
    auto lower_off = [](const int32_t* o) {
        return (((o[0] & 4095) >> 7) << 10) | (((o[1] & 4095) >> 7) << 5) | ((o[2] & 4095) >> 7);
    };
    auto leaf_off = [](const int32_t* o) {
        return (((o[0] & 127) >> 3) << 8) | (((o[1] & 127) >> 3) << 4) | ((o[2] & 127) >> 3);
    };
    const int la = lower_off(a.o), lb = lower_off(b.o);
    if (la != lb) return la < lb;
    return leaf_off(a.o) < leaf_off(b.o);
}

int upload(const std::vector<int32_t>& h, int32_t** d) {
    const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(int32_t);
    hipError_t e = hipMalloc(d, bytes);
    if (e == hipSuccess && !h.empty()) e = hipMemcpy(*d, h.data(), h.size() * 4, hipMemcpyHostToDevice);
    return e == hipSuccess ? PVDB_OK : pvdb::hip_fail(e, "topology upload");
}

// Empty-run map over the one-level leaf grid, one byte per 8^3 cell: 0 where
// a leaf exists, otherwise the smallest r >= 1 such that a leaf lies within
// Chebyshev distance r of the cell (so every cell within r - 1 is empty), or
// `cap` when none lies within cap - 1.  Built by cube dilations (each one a
// pass along x, y and z) until every cell is reached or cap passes were made.
// Empty for the two-level layout.  The passes are limited so the build stays
// within about 2^28 cell updates (the map only has to be a lower bound).
std::vector<uint8_t> empty_run_map(const pvdb_topology* t, int cap) {
    std::vector<uint8_t> dist;
    if (t->two_level) return dist;
    const int n0 = t->dim[0], n1 = t->dim[1], n2 = t->dim[2];
    const size_t n = (size_t)n0 * n1 * n2;
    cap = (int)std::max<size_t>(2, std::min<size_t>((size_t)cap, (size_t(1) << 28) / (3 * n + 1)));
    dist.assign(n, (uint8_t)cap);
    std::vector<uint8_t> cur(n), tmp(n);
    size_t left = n;
    for (size_t c = 0; c < n; ++c) {
        cur[c] = t->table[c] >= 0;
        if (cur[c]) {
            dist[c] = 0;
            --left;
        }
    }
    auto idx = [&](int a, int b, int c) { return ((size_t)a * n1 + b) * n2 + c; };
    for (int r = 1; r < cap && left > 0; ++r) {
        // tmp = cur dilated by one cell along z, then y, then x (back into cur)
        for (int a = 0; a < n0; ++a)
            for (int b = 0; b < n1; ++b)
                for (int c = 0; c < n2; ++c)
                    tmp[idx(a, b, c)] = cur[idx(a, b, c)] | (c > 0 && cur[idx(a, b, c - 1)]) |
                                        (c + 1 < n2 && cur[idx(a, b, c + 1)]);
        std::vector<uint8_t> nb(n);
        for (int a = 0; a < n0; ++a)
            for (int b = 0; b < n1; ++b)
                for (int c = 0; c < n2; ++c)
                    nb[idx(a, b, c)] = tmp[idx(a, b, c)] | (b > 0 && tmp[idx(a, b - 1, c)]) |
                                       (b + 1 < n1 && tmp[idx(a, b + 1, c)]);
        for (int a = 0; a < n0; ++a)
            for (int b = 0; b < n1; ++b)
                for (int c = 0; c < n2; ++c) {
                    const size_t i = idx(a, b, c);
                    const uint8_t v = nb[i] | (a > 0 && nb[idx(a - 1, b, c)]) |
                                      (a + 1 < n0 && nb[idx(a + 1, b, c)]);
                    if (v && !cur[i]) {
                        dist[i] = (uint8_t)r;
                        --left;
                    }
                    tmp[i] = v;
                }
        cur.swap(tmp);
    }
    return dist;
}
// layout: 0 automatic, 1 one-level (dense leaf table), 2 two-level.
int finalize(std::vector<LeafRec>& leaves, int layout, pvdb_topology** out) {
    std::sort(leaves.begin(), leaves.end(), nano_less);
    for (size_t i = 1; i < leaves.size(); ++i)
        if (leaves[i].o[0] == leaves[i - 1].o[0] && leaves[i].o[1] == leaves[i - 1].o[1] &&
            leaves[i].o[2] == leaves[i - 1].o[2]) {
            pvdb::set_error("duplicate leaf origin");
            return PVDB_EINVAL;
        }
    pvdb_topology* t = new (std::nothrow) pvdb_topology();
    if (!t) return PVDB_ENOMEM;
    t->nleaf = (int64_t)leaves.size();
    int32_t lmin[3] = {0, 0, 0}, lmax[3] = {-1, -1, -1};
    if (!leaves.empty()) {
        for (int d = 0; d < 3; ++d) lmin[d] = lmax[d] = leaves[0].o[d] >> 3;
        for (auto& L : leaves)
            for (int d = 0; d < 3; ++d) {
                lmin[d] = std::min(lmin[d], L.o[d] >> 3);
                lmax[d] = std::max(lmax[d], L.o[d] >> 3);
            }
    }
    int64_t cells = 1, ucells = 1;
    for (int d = 0; d < 3; ++d) {
        t->lo[d] = lmin[d];
        t->dim[d] = lmax[d] - lmin[d] + 1;
        t->ulo[d] = lmin[d] >> 4;
        t->udim[d] = (lmax[d] >> 4) - (lmin[d] >> 4) + 1;
        cells *= t->dim[d];
        ucells *= t->udim[d];
    }
    if (ucells > (int64_t(1) << 26)) {
        delete t;
        pvdb::set_error("leaf bounding box too large (lower-node table > 2^26 cells)");
        return PVDB_ERANGE;
    }
    if (layout == 0) layout = cells <= (int64_t(1) << 24) ? 1 : 2;
    if (layout == 1 && cells > (int64_t(1) << 28)) {
        delete t;
        pvdb::set_error("leaf bounding box too large for a one-level leaf table (> 2^28 cells)");
        return PVDB_ERANGE;
    }
    t->two_level = layout == 2;
    t->origins.resize(3 * leaves.size());
    t->masks.resize(8 * leaves.size());
    t->lower_table.assign((size_t)ucells, -1);
    if (!t->two_level) t->table.assign((size_t)std::max<int64_t>(cells, 1), -1);
    int64_t nact = 0;
    for (size_t i = 0; i < leaves.size(); ++i) {
        const LeafRec& L = leaves[i];
        const int32_t lx = L.o[0] >> 3, ly = L.o[1] >> 3, lz = L.o[2] >> 3;
        const size_t u = ((size_t)((lx >> 4) - t->ulo[0]) * t->udim[1] + ((ly >> 4) - t->ulo[1])) *
                             t->udim[2] + ((lz >> 4) - t->ulo[2]);
        if (t->lower_table[u] < 0) {
            t->lower_table[u] = (int32_t)t->nlower++;
            if (t->two_level) t->lower_leaves.resize((size_t)t->nlower * 4096, -1);
        }
        if (t->two_level) {
            const int off = ((lx & 15) << 8) | ((ly & 15) << 4) | (lz & 15);
            t->lower_leaves[(size_t)t->lower_table[u] * 4096 + off] = (int32_t)i;
        } else {
            t->table[((size_t)(lx - t->lo[0]) * t->dim[1] + (ly - t->lo[1])) * t->dim[2] +
                     (lz - t->lo[2])] = (int32_t)i;
        }
        for (int d = 0; d < 3; ++d) t->origins[3 * i + d] = L.o[d];
        for (int w = 0; w < 8; ++w) {
            t->masks[8 * i + w] = L.m[w];
            nact += __builtin_popcountll(L.m[w]);
        }
    }
    t->nactive = nact;
    t->erun = empty_run_map(t, 64);
    int rc = hipGetDevice(&t->device) == hipSuccess ? PVDB_OK : PVDB_EHIP;
    if (rc == PVDB_OK) rc = upload(t->table, &t->d_table);
    if (rc == PVDB_OK) rc = upload(t->origins, &t->d_origins);
    if (rc == PVDB_OK) rc = upload(t->lower_table, &t->d_lower);
    if (rc == PVDB_OK) rc = upload(t->lower_leaves, &t->d_lower_leaves);
    if (rc == PVDB_OK) {
        hipError_t e = hipMalloc(&t->d_masks, std::max<size_t>(t->masks.size(), 8) * 8);
        if (e == hipSuccess && !t->masks.empty())
            e = hipMemcpy(t->d_masks, t->masks.data(), t->masks.size() * 8, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = pvdb::hip_fail(e, "topology upload");
    }
    if (rc == PVDB_OK && !t->erun.empty()) {
        hipError_t e = hipMalloc(&t->d_erun, t->erun.size());
        if (e == hipSuccess)
            e = hipMemcpy(t->d_erun, t->erun.data(), t->erun.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = pvdb::hip_fail(e, "topology upload");
    }
    if (rc != PVDB_OK) {
        pvdb_topo_destroy(t);
        return rc;
    }
    *out = t;
    return PVDB_OK;
}

}  // namespace

namespace pvdb {
TopoView view_of(const pvdb_topology* t) {
    TopoView v;
    v.table = t->d_table;
    v.masks = t->d_masks;
    v.origins = t->d_origins;
    v.lower_table = t->d_lower;
    v.lower_leaves = t->d_lower_leaves;
    v.erun = t->d_erun;
    v.lo0 = t->lo[0];
    v.lo1 = t->lo[1];
    v.lo2 = t->lo[2];
    v.d0 = t->dim[0];
    v.d1 = t->dim[1];
    v.d2 = t->dim[2];
    v.ulo0 = t->ulo[0];
    v.ulo1 = t->ulo[1];
    v.ulo2 = t->ulo[2];
    v.e0 = t->udim[0];
    v.e1 = t->udim[1];
    v.e2 = t->udim[2];
    v.two_level = t->two_level;
    v.pad_ = 0;
    v.nleaf = t->nleaf;
    return v;
}

}  // namespace pvdb

extern "C" {

const char* pvdb_last_error(void) { return pvdb::g_last_error.c_str(); }

int pvdb_abi_version(void) { return PVDB_ABI_VERSION; }

int pvdb_topo_create_from_mask(const uint8_t* mask, int32_t rx, int32_t ry, int32_t rz,
                               pvdb_topology** out) {
    PVDB_CHECK_ARG(out != nullptr, "out is null");
    PVDB_CHECK_ARG(rx > 0 && ry > 0 && rz > 0, "resolution must be positive");
    PVDB_CHECK_ARG(mask != nullptr, "mask is null");
    const int lx = (rx + 7) / 8, ly = (ry + 7) / 8, lz = (rz + 7) / 8;
    std::vector<LeafRec> leaves;
    for (int a = 0; a < lx; ++a)
        for (int b = 0; b < ly; ++b)
            for (int c = 0; c < lz; ++c) {
                LeafRec L;
                std::memset(&L, 0, sizeof(L));
                L.o[0] = 8 * a;
                L.o[1] = 8 * b;
                L.o[2] = 8 * c;
                bool any = false;
                for (int i = 0; i < 8; ++i) {
                    const int x = 8 * a + i;
                    if (x >= rx) break;
                    for (int j = 0; j < 8; ++j) {
                        const int y = 8 * b + j;
                        if (y >= ry) break;
                        const uint8_t* row = mask + ((size_t)x * ry + y) * rz;
                        for (int k = 0; k < 8; ++k) {
                            const int z = 8 * c + k;
                            if (z >= rz) break;
                            if (row[z]) {
                                const int off = (i << 6) | (j << 3) | k;
                                L.m[off >> 6] |= 1ull << (off & 63);
                                any = true;
                            }
                        }
                    }
                }
                if (any) leaves.push_back(L);
            }
    return finalize(leaves, 0, out);
}

int pvdb_topo_create_from_leaves(const int32_t* origins, const uint64_t* masks, int64_t nleaf,
                                 int32_t layout, pvdb_topology** out) {
    PVDB_CHECK_ARG(out != nullptr, "out is null");
    PVDB_CHECK_ARG(nleaf >= 0, "nleaf must be >= 0");
    PVDB_CHECK_ARG(nleaf == 0 || (origins && masks), "null leaf arrays");
    PVDB_CHECK_ARG(layout >= 0 && layout <= 2, "layout must be 0 (auto), 1 or 2");
    std::vector<LeafRec> leaves((size_t)nleaf);
    for (int64_t i = 0; i < nleaf; ++i) {
        for (int d = 0; d < 3; ++d) {
            PVDB_CHECK_ARG((origins[3 * i + d] & 7) == 0, "leaf origins must be multiples of 8");
            leaves[i].o[d] = origins[3 * i + d];
        }
        for (int w = 0; w < 8; ++w) leaves[i].m[w] = masks[8 * i + w];
    }
    return finalize(leaves, layout, out);
}

int pvdb_topo_destroy(pvdb_topology* t) {
    if (!t) return PVDB_OK;
    int dev = 0;
    hipGetDevice(&dev);
    if (dev != t->device) hipSetDevice(t->device);
    if (t->d_table) hipFree(t->d_table);
    if (t->d_masks) hipFree(t->d_masks);
    if (t->d_origins) hipFree(t->d_origins);
    if (t->d_lower) hipFree(t->d_lower);
    if (t->d_lower_leaves) hipFree(t->d_lower_leaves);
    if (t->d_erun) hipFree(t->d_erun);
    if (dev != t->device) hipSetDevice(dev);
    delete t;
    return PVDB_OK;
}

int pvdb_topo_info(const pvdb_topology* t, int64_t* nleaf, int64_t* nactive, int32_t* lo3,
                   int32_t* dim3, int32_t* two_level) {
    if (!t) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    if (nleaf) *nleaf = t->nleaf;
    if (nactive) *nactive = t->nactive;
    if (two_level) *two_level = t->two_level;
    for (int d = 0; d < 3; ++d) {
        if (lo3) lo3[d] = t->lo[d];
        if (dim3) dim3[d] = t->dim[d];
    }
    return PVDB_OK;
}

int pvdb_topo_leaves(const pvdb_topology* t, int32_t* origins, uint64_t* masks) {
    if (!t) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    if (origins && !t->origins.empty())
        std::memcpy(origins, t->origins.data(), t->origins.size() * sizeof(int32_t));
    if (masks && !t->masks.empty()) std::memcpy(masks, t->masks.data(), t->masks.size() * 8);
    return PVDB_OK;
}

int64_t pvdb_topo_empty_run(const pvdb_topology* t, uint8_t* out) {
    if (!t) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    if (out && !t->erun.empty()) std::memcpy(out, t->erun.data(), t->erun.size());
    return (int64_t)t->erun.size();
}

int pvdb_topo_slots(const pvdb_topology* t, const int32_t* ijk, int64_t n, int64_t* slots,
                    uint8_t* active) {
    if (!t) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    PVDB_CHECK_ARG(n >= 0 && (n == 0 || (ijk && slots)), "bad query arrays");
    for (int64_t q = 0; q < n; ++q) {
        const int32_t i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
        int64_t s = -1;
        bool act = false;
        int32_t l = -1;
        const int64_t ua = (int64_t)(i >> 7) - t->ulo[0], ub = (int64_t)(j >> 7) - t->ulo[1],
                      uc = (int64_t)(k >> 7) - t->ulo[2];
        if (ua >= 0 && ub >= 0 && uc >= 0 && ua < t->udim[0] && ub < t->udim[1] && uc < t->udim[2]) {
            const int32_t low = t->lower_table[((size_t)ua * t->udim[1] + ub) * t->udim[2] + uc];
            if (low >= 0) {
                if (t->two_level) {
                    const int off = (((i & 127) >> 3) << 8) | (((j & 127) >> 3) << 4) | ((k & 127) >> 3);
                    l = t->lower_leaves[(size_t)low * 4096 + off];
                } else {
                    const int64_t a = (int64_t)(i >> 3) - t->lo[0], b = (int64_t)(j >> 3) - t->lo[1],
                                  c = (int64_t)(k >> 3) - t->lo[2];
                    l = t->table[((size_t)a * t->dim[1] + b) * t->dim[2] + c];
                }
            }
        }
        if (l >= 0) {
            const int off = ((i & 7) << 6) | ((j & 7) << 3) | (k & 7);
            s = (int64_t)l * 512 + off;
            act = (t->masks[8 * (size_t)l + (off >> 6)] >> (off & 63)) & 1ull;
        }
        slots[q] = s;
        if (active) active[q] = act ? 1 : 0;
    }
    return PVDB_OK;
}

int pvdb_topo_device_arrays(const pvdb_topology* t, const int32_t** table, const uint64_t** masks,
                            const int32_t** origins) {
    if (!t) {
        pvdb::set_error("null topology");
        return PVDB_EHANDLE;
    }
    if (table) *table = t->d_table;
    if (masks) *masks = t->d_masks;
    if (origins) *origins = t->d_origins;
    return PVDB_OK;
}

}  // extern "C"
