// Definition of the opaque e3d_mesh handle, shared by the host-mesh ABI and
// the device context (which uploads it).
#pragma once
#include <vector>

#include "mesh.hpp"

struct e3d_mesh {
    e3d::MultiBlockMesh raw;
    std::vector<e3d::HostBlock> built;
    // flattened per-block views of boundaries and interfaces
    struct Flat {
        std::vector<int> bnd_type, bnd_off, bnd_ghosts;
        std::vector<int> itf_origin_block, itf_off, itf_origin_cells;
    };
    std::vector<Flat> flat;
};
