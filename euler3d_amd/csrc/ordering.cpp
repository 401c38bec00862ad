// Device memory order of a block's cells and faces.
//
// The reference numbers cells as the mesh file lists them and faces by first
// appearance over those cells (BlockBuilder/src/BlockBuilder.cpp:147-458). That
// order fixes every per-cell summation order, so it is never changed: each
// cell's face slots stay in ascending reference face id. What may change is
// where a cell and a face sit in HBM. Reverse Cuthill-McKee (over the face
// graph) or a Morton curve (over the cell centroids) places cells that share
// faces close together, so the neighbour gathers of a wavefront touch few
// cache lines whatever order the mesh file came in.
#include "ordering.hpp"

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <utility>

namespace e3d {

namespace {

// Breadth-first level structure of the component holding `root`, visiting
// unvisited neighbours by ascending degree (ties: ascending id). Appends the
// visited cells to `out` and returns the number of levels; `last` receives the
// cells of the deepest level.
int bfs_levels(int root, const std::vector<int>& off, const std::vector<int>& adj, const std::vector<int>& deg,
               std::vector<char>& seen, std::vector<int>& out, std::vector<int>* last) {
    const size_t start = out.size();
    out.push_back(root);
    seen[root] = 1;
    size_t level_begin = start, level_end = out.size();
    int levels = 0;
    std::vector<int> nb;
    while (level_begin < level_end) {
        ++levels;
        for (size_t i = level_begin; i < level_end; ++i) {
            const int v = out[i];
            nb.clear();
            for (int j = off[v]; j < off[v + 1]; ++j)
                if (!seen[adj[j]]) {
                    seen[adj[j]] = 1;
                    nb.push_back(adj[j]);
                }
            std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
            out.insert(out.end(), nb.begin(), nb.end());
        }
        if (out.size() == level_end && last) last->assign(out.begin() + level_begin, out.begin() + level_end);
        level_begin = level_end;
        level_end = out.size();
    }
    return levels;
}

}  // namespace

std::vector<int> rcm_order(const HostBlock& B) {
    const int n = B.nreal;
    // real-to-real adjacency (ghost neighbours are not part of the ordering)
    std::vector<int> off(n + 1, 0), adj;
    adj.reserve((size_t)B.cell_off[n]);
    for (int q = 0; q < n; ++q) {
        for (int j = B.cell_off[q]; j < B.cell_off[q + 1]; ++j)
            if (B.cell2cell[j] < n) adj.push_back(B.cell2cell[j]);
        off[q + 1] = (int)adj.size();
    }
    std::vector<int> deg(n);
    for (int q = 0; q < n; ++q) deg[q] = off[q + 1] - off[q];

    std::vector<int> order;
    order.reserve(n);
    std::vector<char> seen(n, 0), probe_seen;
    std::vector<int> probe, last;
    for (int s = 0; s < n; ++s) {
        if (seen[s]) continue;
        // pseudo-peripheral root of this component (George-Liu): start at the
        // lowest-degree unvisited cell, move to a lowest-degree cell of the
        // deepest BFS level while the depth grows
        int root = s;
        for (int q = s; q < n; ++q)
            if (!seen[q] && deg[q] < deg[root]) root = q;
        int depth = -1;
        for (int it = 0; it < 8; ++it) {
            probe_seen = seen;
            probe.clear();
            const int d = bfs_levels(root, off, adj, deg, probe_seen, probe, &last);
            if (d <= depth) break;
            depth = d;
            int next = last[0];
            for (int v : last)
                if (deg[v] < deg[next] || (deg[v] == deg[next] && v < next)) next = v;
            if (next == root) break;
            root = next;
        }
        bfs_levels(root, off, adj, deg, seen, order, nullptr);
    }
    std::reverse(order.begin(), order.end());
    return order;
}

namespace {
// Bits of v (< 2^21) spread to every third bit.
uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
}  // namespace

std::vector<int> morton_order(const HostBlock& B) {
    const int n = B.nreal;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int q = 0; q < n; ++q)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], B.ccenter[3 * q + d]);
            hi[d] = std::max(hi[d], B.ccenter[3 * q + d]);
        }
    // one scale for all three axes, so the curve's cells are cubes
    const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    const double sc = ext > 0.0 ? (double)((1 << 21) - 1) / ext : 0.0;
    std::vector<std::pair<uint64_t, int>> key(n);
    for (int q = 0; q < n; ++q) {
        uint64_t k = 0;
        for (int d = 0; d < 3; ++d) k |= spread3((uint64_t)((B.ccenter[3 * q + d] - lo[d]) * sc)) << d;
        key[q] = {k, q};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> order(n);
    for (int p = 0; p < n; ++p) order[p] = key[p].second;
    return order;
}

void block_layout(const HostBlock& B, int mode, std::vector<int>& cell_pos, std::vector<int>& face_pos) {
    const int n = B.nreal;
    cell_pos.resize(n);
    face_pos.resize(B.nface);
    if (mode == ORDER_REFERENCE) {
        std::iota(cell_pos.begin(), cell_pos.end(), 0);
        std::iota(face_pos.begin(), face_pos.end(), 0);
        return;
    }
    const std::vector<int> order = mode == ORDER_MORTON ? morton_order(B) : rcm_order(B);
    for (int p = 0; p < n; ++p) cell_pos[order[p]] = p;
    // faces by first appearance over the cells in their new order
    std::fill(face_pos.begin(), face_pos.end(), -1);
    int next = 0;
    for (int p = 0; p < n; ++p) {
        const int q = order[p];
        for (int j = B.cell_off[q]; j < B.cell_off[q + 1]; ++j)
            if (face_pos[B.cell2face[j]] < 0) face_pos[B.cell2face[j]] = next++;
    }
    for (int f = 0; f < B.nface; ++f)  // a face of ghosts only (none in a valid block) goes last
        if (face_pos[f] < 0) face_pos[f] = next++;
}

}  // namespace e3d
