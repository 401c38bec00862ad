// Index of faces by their sorted node set, for the mesh builders. Faces are
// grouped by their smallest node, in buckets sized by a first counting pass:
// the cells of a block are walked in order and neighbouring cells share nodes,
// so the buckets a run of cells touches sit close together in memory, where a
// hash over all faces would touch it at random (cache and TLB misses dominate
// on meshes of millions of cells). Lookups return the id stored at insertion;
// the builders assign ids in their own (first-appearance) order, so results do
// not depend on the index.
#pragma once
#include <climits>
#include <cstddef>
#include <vector>

namespace e3d {

template <int N>  // nodes per key, sorted ascending; unused trailing slots hold INT_MAX
class FaceIndex {
   public:
    struct Key {
        int v[N];
    };

    // nnode: node ids are in [0, nnode). for_each_key(visit) must call
    // visit(key) once for every insertion that will be tried (a key tried
    // twice is counted twice); bucket sizes come from these counts.
    template <class ForEachKey>
    FaceIndex(int nnode, ForEachKey for_each_key) : off_((size_t)nnode + 1, 0), fill_(nnode, 0) {
        for_each_key([&](const Key& k) { ++off_[(size_t)k.v[0] + 1]; });
        for (int n = 0; n < nnode; ++n) off_[n + 1] += off_[n];
        ent_.resize(off_[nnode]);
    }

    // The value stored for key; if absent, stores `value` for it, sets
    // *inserted and returns `value`. The key must have been counted.
    int insert_or_get(const Key& k, int value, bool* inserted) {
        const size_t b = off_[k.v[0]];
        const int n = fill_[k.v[0]];
        for (int e = 0; e < n; ++e)
            if (same(ent_[b + e], k)) {
                *inserted = false;
                return ent_[b + e].val;
            }
        Entry& t = ent_[b + n];
        for (int q = 1; q < N; ++q) t.rest[q - 1] = k.v[q];
        t.val = value;
        ++fill_[k.v[0]];
        *inserted = true;
        return value;
    }

    // Value of key, or -1 when absent (or when its smallest node is out of range).
    int find(const Key& k) const {
        if (k.v[0] < 0 || (size_t)k.v[0] + 1 >= off_.size()) return -1;
        const size_t b = off_[k.v[0]];
        for (int e = 0; e < fill_[k.v[0]]; ++e)
            if (same(ent_[b + e], k)) return ent_[b + e].val;
        return -1;
    }

   private:
    struct Entry {
        int rest[N - 1];
        int val;
    };
    static bool same(const Entry& t, const Key& k) {
        for (int q = 1; q < N; ++q)
            if (t.rest[q - 1] != k.v[q]) return false;
        return true;
    }
    std::vector<size_t> off_;
    std::vector<int> fill_;
    std::vector<Entry> ent_;
};

}  // namespace e3d
