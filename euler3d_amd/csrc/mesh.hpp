// Host-side mesh layer: raw SU2-dialect blocks, the multi-block topology, the
// bump-channel generator and partitioner, and the "built" block (faces,
// connectivity and geometry in the reference's numbering).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace e3d {

enum BcType : int { BC_WALL = 0, BC_FARFIELD = 1, BC_SYMMETRY = 2, BC_CONNEXION = 3 };

// One SU2 marker: element type per entry (5 triangle, 9 quad) and node lists.
struct Marker {
    BcType type;
    std::vector<int> etype;
    std::vector<int> off{0};
    std::vector<int> nodes;
    int size() const { return (int)etype.size(); }
};

// Contents of one block file in the reference's SU2 dialect
// (BlockBuilder/src/ConcreteBlockBuilder.cpp:22-495).
struct RawBlock {
    std::vector<double> xyz;            // 3 * npoin
    std::vector<int> etype;             // 10 tet, 12 hex ("prism"), 14 pyramid
    std::vector<int> eoff{0};
    std::vector<int> enodes;
    std::vector<Marker> markers;        // physical markers first, CONNEXION last
    int npoin() const { return (int)xyz.size() / 3; }
    int nelem() const { return (int)etype.size(); }
};

// One inter-block interface as the topology file lists it
// (DataStructure/src/CompleteMesh.cpp:198-257): ghosts of the destination block
// numbered from ghost_count, filled from real cells of the origin block.
struct Interface {
    int origin_block;
    std::vector<int> origin_cells;
};

struct BlockTopo {
    int ghost_count = 0;                // first interface-ghost id of the block
    std::vector<Interface> interfaces;
};

struct MultiBlockMesh {
    std::vector<RawBlock> blocks;
    std::vector<BlockTopo> topo;
};

// Built block: everything the solver path reads, in the reference's numbering.
// Cells [0, nreal) are real, [nreal, nall) ghosts in marker order.
struct HostBlock {
    int nreal = 0, nall = 0, nface = 0, nnode = 0;
    std::vector<int> face2cell;         // 2F: [left, right], left = lower cell index
    std::vector<int> face_nnode;        // F
    std::vector<int> face_nodes;        // 4F (unused slots -1)
    std::vector<double> S, area, fcenter, rL, rR;   // 3F, F, 3F, 3F, 3F
    std::vector<int> cell_off;          // nall+1, CSR into cell2face / cell2cell
    std::vector<int> cell2face, cell2cell;
    std::vector<int> cell_nnode, cell_nodes_off, cell_nodes;
    std::vector<double> volume, ccenter;            // nall, 3*nall
    std::vector<double> lsqw;           // 3 * cell_off[nreal] (real cells only)
    struct Boundary { BcType type; std::vector<int> ghosts; };
    std::vector<Boundary> bnds;         // one per marker, file order
    int n_real_boundaries = 0;          // markers before the first CONNEXION
    std::vector<int> wall_faces, farfield_faces, symmetry_faces, connexion_faces;
};

struct BumpParams {
    int nx = 48, ny = 16, nz = 16;
    double h = 0.1;                     // bump thickness / chord
    int brick = 4;                      // hex brick edge for cell ordering (0 = lexicographic,
                                        // < 0: tets in a pseudo-random order seeded by -brick)
    int px = 1, py = 1, pz = 1;         // partition into px*py*pz blocks
};

// Generators / IO (mesh.cpp)
MultiBlockMesh make_bump_channel(const BumpParams& p);
void write_su2(const RawBlock& b, const std::string& path);
RawBlock read_su2(const std::string& path);
void write_topology(const MultiBlockMesh& m, const std::vector<std::string>& block_paths,
                    const std::string& path);
MultiBlockMesh read_topology(const std::string& path);

// Builder (block_builder.cpp)
HostBlock build_block(const RawBlock& raw);

}  // namespace e3d
