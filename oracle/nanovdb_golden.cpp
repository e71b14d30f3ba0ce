// Golden-vector generator built against the reference's own header-only NanoVDB
// (openvdb/nanovdb/nanovdb, version 32.3.3 as vendored by the reference).
//
// TEST INFRASTRUCTURE ONLY: compiled by oracle/Makefile (target `ref`) into
// oracle/_ref/nanovdb_golden, run by tests/golden/make_golden.py to write the
// fixtures under tests/golden/.  The product never links or runs it.
//
// It builds a NanoGrid<float> with GridBuilder (the same depth-first node order
// as OpenToNanoVDB, openvdb/nanovdb/nanovdb/util/GridBuilder.h:362-376 and
// util/OpenToNanoVDB.h:509-533) from a list of active voxels and reports:
//   * the leaf origins in NanoVDB order (tree().getFirstNode<0>() + i),
//   * each leaf's value mask,
//   * for integer query coordinates: ReadAccessor getValue / isActive, whether
//     the accessor holds the leaf afterwards (isCached<Leaf>, the scatter rule
//     of plenvdb/lib/vdb/densityvdb.cu:16-27) and the slot index
//     leaf_index*512 + LeafNode::CoordToOffset (densityvdb.cu:3-14),
//   * for float query points: the density trilinear rule of
//     densityvdb.cu:101-125 evaluated through the NanoVDB accessor, and
//     NanoVDB's own TrilinearSampler (util/SampleFromVoxels.h) for comparison.
//
// Input file (little-endian):  int32 nvox; nvox x {int32 i,j,k; float v}
//                              int32 nq;   nq   x {int32 i,j,k}
//                              int32 np;   np   x {float x,y,z}
// Output file: int32 nleaf; nleaf x {int32 ox,oy,oz; uint64 mask[8]}
//              nq x {float value; int32 active; int32 cached; int64 slot}
//              np x {float tri_ref_order; float tri_nanovdb_sampler}
#include <nanovdb/NanoVDB.h>
#include <nanovdb/util/GridBuilder.h>
#include <nanovdb/util/SampleFromVoxels.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

template <typename T>
static bool rd(FILE* f, T* p, size_t n = 1) { return fread(p, sizeof(T), n, f) == n; }
template <typename T>
static void wr(FILE* f, const T* p, size_t n = 1) { fwrite(p, sizeof(T), n, f); }

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 3;
    int32_t nvox = 0;
    if (!rd(in, &nvox)) return 4;
    nanovdb::GridBuilder<float> builder(0.0f);
    auto bacc = builder.getAccessor();
    for (int32_t n = 0; n < nvox; ++n) {
        int32_t c[3];
        float v;
        if (!rd(in, c, 3) || !rd(in, &v)) return 5;
        bacc.setValue(nanovdb::Coord(c[0], c[1], c[2]), v);
    }
    int32_t nq = 0;
    if (!rd(in, &nq)) return 6;
    std::vector<int32_t> q(3 * (size_t)nq);
    if (nq && !rd(in, q.data(), q.size())) return 7;
    int32_t np = 0;
    if (!rd(in, &np)) return 8;
    std::vector<float> p(3 * (size_t)np);
    if (np && !rd(in, p.data(), p.size())) return 9;
    fclose(in);

    auto handle = builder.getHandle<>();
    const nanovdb::NanoGrid<float>* grid = handle.grid<float>();
    if (!grid) return 10;
    using LeafT = nanovdb::NanoLeaf<float>;
    const auto& tree = grid->tree();
    const int32_t nleaf = (int32_t)tree.nodeCount(0);
    if (nleaf && !grid->isSequential<0>()) return 11;
    const LeafT* first = tree.getFirstNode<0>();

    FILE* out = fopen(argv[2], "wb");
    if (!out) return 12;
    wr(out, &nleaf);
    for (int32_t i = 0; i < nleaf; ++i) {
        const LeafT* leaf = first + i;
        const nanovdb::Coord o = leaf->origin();
        int32_t oc[3] = {o[0], o[1], o[2]};
        wr(out, oc, 3);
        const auto& mask = leaf->valueMask();
        for (int w = 0; w < 8; ++w) {
            uint64_t word = 0;
            for (int b = 0; b < 64; ++b)
                if (mask.isOn(64 * w + b)) word |= uint64_t(1) << b;
            wr(out, &word);
        }
    }
    for (int32_t n = 0; n < nq; ++n) {
        auto acc = grid->getAccessor();  // fresh accessor: no stale cache
        nanovdb::Coord ijk(q[3 * n], q[3 * n + 1], q[3 * n + 2]);
        float v = acc.getValue(ijk);
        int32_t cached = acc.isCached<LeafT>(ijk) ? 1 : 0;
        int32_t active = acc.isActive(ijk) ? 1 : 0;
        int64_t slot = -1;
        if (cached) {
            const LeafT* leaf = acc.getNode<LeafT>();
            slot = (int64_t)(leaf - first) * 512 + LeafT::CoordToOffset(ijk);
        }
        wr(out, &v);
        wr(out, &active);
        wr(out, &cached);
        wr(out, &slot);
    }
    auto smp = nanovdb::createSampler<1, nanovdb::DefaultReadAccessor<float>, false>(grid->getAccessor());
    for (int32_t n = 0; n < np; ++n) {
        auto acc = grid->getAccessor();
        nanovdb::Vec3f xyz(p[3 * n], p[3 * n + 1], p[3 * n + 2]);
        nanovdb::Coord ijk = xyz.floor();
        nanovdb::Vec3f uvw = xyz - ijk.asVec3s();
        // corner walk and per-term product order of densityvdb.cu:115-122
        float r = 0;
        r += acc.getValue(ijk) * (1 - uvw[0]) * (1 - uvw[1]) * (1 - uvw[2]); ijk[2] += 1;
        r += acc.getValue(ijk) * (1 - uvw[0]) * (1 - uvw[1]) * uvw[2]; ijk[1] += 1;
        r += acc.getValue(ijk) * (1 - uvw[0]) * uvw[1] * uvw[2]; ijk[2] -= 1;
        r += acc.getValue(ijk) * (1 - uvw[0]) * uvw[1] * (1 - uvw[2]); ijk[0] += 1;
        r += acc.getValue(ijk) * uvw[0] * uvw[1] * (1 - uvw[2]); ijk[2] += 1;
        r += acc.getValue(ijk) * uvw[0] * uvw[1] * uvw[2]; ijk[1] -= 1;
        r += acc.getValue(ijk) * uvw[0] * (1 - uvw[1]) * uvw[2]; ijk[2] -= 1;
        r += acc.getValue(ijk) * uvw[0] * (1 - uvw[1]) * (1 - uvw[2]);
        float s = smp(xyz);
        wr(out, &r);
        wr(out, &s);
    }
    fclose(out);
    return 0;
}
