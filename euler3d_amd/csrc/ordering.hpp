// Device memory order of a block's real cells and faces (see ordering.cpp).
#pragma once
#include <vector>

#include "mesh.hpp"

namespace e3d {

constexpr int ORDER_REFERENCE = 0;  // cells and faces in the reference's numbering
constexpr int ORDER_RCM = 1;        // reverse Cuthill-McKee cells, faces by first appearance over them
constexpr int ORDER_MORTON = 2;     // cells along a Morton (Z-order) curve of their centroids, faces likewise

// Reverse Cuthill-McKee order of the block's real cells over their real
// neighbours: order[p] = reference id of the cell at position p.
std::vector<int> rcm_order(const HostBlock& B);

// Real cells along the Morton curve of their centroids (one cubic grid of
// 2^21 steps per axis over the block's bounding box; ties by id).
std::vector<int> morton_order(const HostBlock& B);

// Position in device memory of every real cell (cell_pos[q], q < nreal) and
// face (face_pos[f]) of the block, for an ORDER_* mode.
void block_layout(const HostBlock& B, int mode, std::vector<int>& cell_pos, std::vector<int>& face_pos);

}  // namespace e3d
