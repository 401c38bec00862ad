// Bump-channel generator, block partitioner and the reference's file formats.
//
// Formats follow the reference's readers: the SU2 dialect read line by line in
// BlockBuilder/src/ConcreteBlockBuilder.cpp:22-225 (pre-read) and :227-495
// (read) -- blank lines after NDIME and after the node and element sections
// are required; the topology file read in DataStructure/src/CompleteMesh.cpp:
// 45-76 (block file names) and :170-292 (per-block GhostCount and interfaces).
#include "face_index.hpp"
#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace e3d {

namespace {

// Six tets per hex split along the v0-v6 diagonal; hex vertices
// v0=(i,j,k) v1=(i+1,j,k) v2=(i+1,j+1,k) v3=(i,j+1,k) and v4..v7 at k+1.
// Every hex uses the same diagonal, so the split is conforming.
const int kHexTets[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6},
                            {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};

// Boundary triangles of each hex side, as the tet faces that lie on it.
// side: 0 x-, 1 x+, 2 y-, 3 y+, 4 z-, 5 z+
const int kSideTris[6][2][3] = {{{0, 3, 7}, {0, 7, 4}}, {{1, 2, 6}, {5, 1, 6}},
                                {{0, 4, 5}, {0, 5, 1}}, {{2, 3, 6}, {3, 7, 6}},
                                {{0, 1, 2}, {0, 2, 3}}, {{7, 4, 6}, {4, 5, 6}}};

// Local face table of a tetrahedron in the reference
// (FactoryMethod/src/TempTetrahedralFaceCreator.cpp:51).
const int kTetFaces[4][3] = {{0, 2, 1}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}};

double bump_y(double x, double h) {
    if (h <= 0.0 || x <= 1.0 || x >= 2.0) return 0.0;
    const double R = (0.25 + h * h) / (2.0 * h);  // circular arc through (1,0),(1.5,h),(2,0)
    const double dx = x - 1.5;
    double y = std::sqrt(R * R - dx * dx) - (R - h);
    return y > 0.0 ? y : 0.0;
}

using TriKey = FaceIndex<3>::Key;
TriKey tri_key(int a, int b, int c) {
    int v[3] = {a, b, c};
    std::sort(v, v + 3);
    return {{v[0], v[1], v[2]}};
}

}  // namespace

MultiBlockMesh make_bump_channel(const BumpParams& p) {
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("bump channel: non-positive size");
    if (p.px < 1 || p.py < 1 || p.pz < 1 || p.px > nx || p.py > ny || p.pz > nz)
        throw std::invalid_argument("bump channel: bad partition");
    const int NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    auto nid = [&](int i, int j, int k) { return i + NX * (j + NY * k); };

    // Global nodes: x uniform on [0,3], z uniform on [0,1], y sheared from the bump to 1.
    std::vector<double> xyz(3 * (size_t)NX * NY * NZ);
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                const double x = 3.0 * i / nx;
                const double yb = bump_y(x, p.h);
                const size_t n = nid(i, j, k);
                xyz[3 * n] = x;
                xyz[3 * n + 1] = yb + (1.0 - yb) * j / ny;
                xyz[3 * n + 2] = (double)k / nz;
            }

    // Hex order: bricks of brick^3 hexes in lexicographic brick order,
    // lexicographic inside each brick (brick = 0: plain lexicographic).
    const int bs = p.brick > 0 ? p.brick : std::max(nx, std::max(ny, nz));
    std::vector<int> hex_order;  // hex linear ids i + nx*(j + ny*k)
    hex_order.reserve((size_t)nx * ny * nz);
    for (int bk = 0; bk < nz; bk += bs)
        for (int bj = 0; bj < ny; bj += bs)
            for (int bi = 0; bi < nx; bi += bs)
                for (int k = bk; k < std::min(bk + bs, nz); ++k)
                    for (int j = bj; j < std::min(bj + bs, ny); ++j)
                        for (int i = bi; i < std::min(bi + bs, nx); ++i)
                            hex_order.push_back(i + nx * (j + ny * k));

    auto block_of_hex = [&](int i, int j, int k) {
        const int bi = (int)((long)i * p.px / nx), bj = (int)((long)j * p.py / ny),
                  bk = (int)((long)k * p.pz / nz);
        return bi + p.px * (bj + p.py * bk);
    };
    const int nblk = p.px * p.py * p.pz;

    // Global tets in hex order; each block keeps the global order of its tets.
    const size_t ntet = (size_t)nx * ny * nz * 6;
    std::vector<int> tet_nodes(4 * ntet), tet_block(ntet), tet_local(ntet);
    std::vector<int> block_ntet(nblk, 0);
    for (size_t h = 0; h < hex_order.size(); ++h) {
        const int hid = hex_order[h];
        const int i = hid % nx, j = (hid / nx) % ny, k = hid / (nx * ny);
        const int v[8] = {nid(i, j, k),         nid(i + 1, j, k),     nid(i + 1, j + 1, k),
                          nid(i, j + 1, k),     nid(i, j, k + 1),     nid(i + 1, j, k + 1),
                          nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)};
        const int b = block_of_hex(i, j, k);
        for (int t = 0; t < 6; ++t) {
            const size_t g = 6 * h + t;
            for (int q = 0; q < 4; ++q) tet_nodes[4 * g + q] = v[kHexTets[t][q]];
            tet_block[g] = b;
            tet_local[g] = block_ntet[b]++;
        }
    }

    if (p.brick < 0) {
        // A badly numbered mesh (memory-order studies): the tets in a
        // pseudo-random order, Fisher-Yates driven by splitmix64 seeded with -brick.
        uint64_t st = (uint64_t)(-(long long)p.brick);
        auto next = [&st]() {
            uint64_t z = (st += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        std::vector<size_t> perm(ntet);
        for (size_t g = 0; g < ntet; ++g) perm[g] = g;
        for (size_t g = ntet; g > 1; --g) std::swap(perm[g - 1], perm[next() % g]);
        std::vector<int> nodes2(4 * ntet), block2(ntet);
        for (size_t g = 0; g < ntet; ++g) {
            for (int q = 0; q < 4; ++q) nodes2[4 * g + q] = tet_nodes[4 * perm[g] + q];
            block2[g] = tet_block[perm[g]];
        }
        tet_nodes.swap(nodes2);
        tet_block.swap(block2);
        std::fill(block_ntet.begin(), block_ntet.end(), 0);
        for (size_t g = 0; g < ntet; ++g) tet_local[g] = block_ntet[tet_block[g]]++;
    }

    // Physical boundary triangles, with the marker they belong to.
    struct BTri { int n[3]; BcType type; };
    std::vector<BTri> btris;
    auto add_side = [&](int i, int j, int k, int side, BcType t) {
        const int v[8] = {nid(i, j, k),         nid(i + 1, j, k),     nid(i + 1, j + 1, k),
                          nid(i, j + 1, k),     nid(i, j, k + 1),     nid(i + 1, j, k + 1),
                          nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)};
        for (int q = 0; q < 2; ++q)
            btris.push_back({{v[kSideTris[side][q][0]], v[kSideTris[side][q][1]], v[kSideTris[side][q][2]]}, t});
    };
    for (int k = 0; k < nz; ++k)  // lower wall with the bump
        for (int i = 0; i < nx; ++i) add_side(i, 0, k, 2, BC_WALL);
    for (int k = 0; k < nz; ++k)  // top
        for (int i = 0; i < nx; ++i) add_side(i, ny - 1, k, 3, BC_SYMMETRY);
    for (int j = 0; j < ny; ++j)  // side walls
        for (int i = 0; i < nx; ++i) add_side(i, j, 0, 4, BC_SYMMETRY);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) add_side(i, j, nz - 1, 5, BC_SYMMETRY);
    for (int k = 0; k < nz; ++k)  // inlet and outlet
        for (int j = 0; j < ny; ++j) add_side(0, j, k, 0, BC_FARFIELD);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) add_side(nx - 1, j, k, 1, BC_FARFIELD);

    // Face adjacency: sorted triangle -> up to two tets.
    auto tet_face = [&](size_t g, int f) {
        const int* tn = &tet_nodes[4 * g];
        return tri_key(tn[kTetFaces[f][0]], tn[kTetFaces[f][1]], tn[kTetFaces[f][2]]);
    };
    FaceIndex<3> owner(NX * NY * NZ, [&](auto&& visit) {  // key -> 4 * first tet + its local face
        for (size_t g = 0; g < ntet; ++g)
            for (int f = 0; f < 4; ++f) visit(tet_face(g, f));
    });
    std::vector<int> nbr(4 * ntet, -1);
    for (size_t g = 0; g < ntet; ++g)
        for (int f = 0; f < 4; ++f) {
            bool is_new;
            const int o = owner.insert_or_get(tet_face(g, f), (int)(4 * g + f), &is_new);
            if (!is_new) {
                nbr[o] = (int)g;
                nbr[4 * g + f] = o / 4;
            }
        }
    std::vector<int> btri_tet(btris.size());
    for (size_t q = 0; q < btris.size(); ++q) {
        const int o = owner.find(tri_key(btris[q].n[0], btris[q].n[1], btris[q].n[2]));
        if (o < 0) throw std::logic_error("bump channel: boundary triangle without a tet");
        btri_tet[q] = o / 4;
    }

    MultiBlockMesh m;
    m.blocks.resize(nblk);
    m.topo.resize(nblk);
    // Local node numbering: ascending global id among the block's nodes.
    std::vector<int> node_local(xyz.size() / 3, -1);
    for (int b = 0; b < nblk; ++b) {
        RawBlock& rb = m.blocks[b];
        std::vector<char> used(xyz.size() / 3, 0);
        for (size_t g = 0; g < ntet; ++g)
            if (tet_block[g] == b)
                for (int q = 0; q < 4; ++q) used[tet_nodes[4 * g + q]] = 1;
        int cnt = 0;
        for (size_t n = 0; n < used.size(); ++n)
            if (used[n]) {
                node_local[n] = cnt++;
                rb.xyz.insert(rb.xyz.end(), &xyz[3 * n], &xyz[3 * n] + 3);
            }
        rb.etype.reserve(block_ntet[b]);
        for (size_t g = 0; g < ntet; ++g)
            if (tet_block[g] == b) {
                rb.etype.push_back(10);
                for (int q = 0; q < 4; ++q) rb.enodes.push_back(node_local[tet_nodes[4 * g + q]]);
                rb.eoff.push_back((int)rb.enodes.size());
            }
        for (BcType t : {BC_WALL, BC_SYMMETRY, BC_FARFIELD}) {
            Marker mk;
            mk.type = t;
            for (size_t q = 0; q < btris.size(); ++q)
                if (btris[q].type == t && tet_block[btri_tet[q]] == b) {
                    mk.etype.push_back(5);
                    for (int r = 0; r < 3; ++r) mk.nodes.push_back(node_local[btris[q].n[r]]);
                    mk.off.push_back((int)mk.nodes.size());
                }
            if (mk.size() > 0) rb.markers.push_back(std::move(mk));
        }
        // Cut faces: one CONNEXION marker per neighbouring block, ascending id,
        // elements in (local tet, local face) order; the topology lists the tet
        // across each cut face, in the neighbour's numbering.
        std::map<int, std::pair<Marker, Interface>> cuts;
        for (size_t g = 0; g < ntet; ++g) {
            if (tet_block[g] != b) continue;
            for (int f = 0; f < 4; ++f) {
                const int o = nbr[4 * g + f];
                if (o < 0 || tet_block[o] == b) continue;
                auto& e = cuts[tet_block[o]];
                e.first.type = BC_CONNEXION;
                e.first.etype.push_back(5);
                for (int r = 0; r < 3; ++r) e.first.nodes.push_back(node_local[tet_nodes[4 * g + kTetFaces[f][r]]]);
                e.first.off.push_back((int)e.first.nodes.size());
                e.second.origin_block = tet_block[o];
                e.second.origin_cells.push_back(tet_local[o]);
            }
        }
        int nphys = 0;
        for (auto& mk : rb.markers) nphys += mk.size();
        m.topo[b].ghost_count = rb.nelem() + nphys;
        for (auto& kv : cuts) {
            rb.markers.push_back(kv.second.first);
            m.topo[b].interfaces.push_back(kv.second.second);
        }
        for (size_t n = 0; n < used.size(); ++n) node_local[n] = -1;
    }
    return m;
}

static const char* bc_name(BcType t) {
    switch (t) {
        case BC_WALL: return "WALL";
        case BC_FARFIELD: return "FARFIELD";
        case BC_SYMMETRY: return "SYMMETRY";
        default: return "CONNEXION";
    }
}

void write_su2(const RawBlock& b, const std::string& path) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    fprintf(f, "NDIME= 3\n\nNPOIN= %d\n", b.npoin());
    for (int n = 0; n < b.npoin(); ++n)
        fprintf(f, "%.17g %.17g %.17g %d\n", b.xyz[3 * n], b.xyz[3 * n + 1], b.xyz[3 * n + 2], n);
    fprintf(f, "\nNELEM= %d\n", b.nelem());
    for (int e = 0; e < b.nelem(); ++e) {
        fprintf(f, "%d", b.etype[e]);
        for (int q = b.eoff[e]; q < b.eoff[e + 1]; ++q) fprintf(f, " %d", b.enodes[q]);
        fprintf(f, "\n");
    }
    fprintf(f, "\nNMARK= %d\n", (int)b.markers.size());
    for (const Marker& m : b.markers) {
        fprintf(f, "MARKER_TAG= %s\nMARKER_ELEMS= %d\n", bc_name(m.type), m.size());
        for (int e = 0; e < m.size(); ++e) {
            fprintf(f, "%d", m.etype[e]);
            for (int q = m.off[e]; q < m.off[e + 1]; ++q) fprintf(f, " %d", m.nodes[q]);
            fprintf(f, "\n");
        }
    }
    fclose(f);
}

namespace {
bool next_line(std::ifstream& in, std::string& line) { return (bool)std::getline(in, line); }
int header_int(const std::string& line, const char* what) {
    char tag[256];
    int v = 0;
    if (sscanf(line.c_str(), "%255s %d", tag, &v) != 2)
        throw std::runtime_error(std::string("SU2: expected '") + what + "' line, got '" + line + "'");
    return v;
}
}  // namespace

RawBlock read_su2(const std::string& path) {
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("cannot open block file " + path);
    RawBlock b;
    std::string line;
    next_line(in, line);  // NDIME= 3
    next_line(in, line);  // blank
    next_line(in, line);
    const int npoin = header_int(line, "NPOIN=");
    b.xyz.resize(3 * (size_t)npoin);
    for (int n = 0; n < npoin; ++n) {
        if (!next_line(in, line)) throw std::runtime_error("SU2: truncated node section");
        if (sscanf(line.c_str(), "%lf %lf %lf", &b.xyz[3 * n], &b.xyz[3 * n + 1], &b.xyz[3 * n + 2]) != 3)
            throw std::runtime_error("SU2: bad node line '" + line + "'");
    }
    next_line(in, line);  // blank
    next_line(in, line);
    const int nelem = header_int(line, "NELEM=");
    for (int e = 0; e < nelem; ++e) {
        if (!next_line(in, line)) throw std::runtime_error("SU2: truncated element section");
        std::istringstream ss(line);
        int t;
        ss >> t;
        const int nn = t == 10 ? 4 : t == 12 ? 8 : t == 14 ? 5 : -1;
        if (nn < 0) throw std::runtime_error("SU2: unsupported element type in '" + line + "'");
        b.etype.push_back(t);
        for (int q = 0; q < nn; ++q) {
            int v;
            ss >> v;
            b.enodes.push_back(v);
        }
        b.eoff.push_back((int)b.enodes.size());
    }
    next_line(in, line);  // blank
    next_line(in, line);
    const int nmark = header_int(line, "NMARK=");
    for (int k = 0; k < nmark; ++k) {
        // Some partitioned files (the reference's ONERA M6 blocks) declare more
        // markers than they hold; the list then simply ends with the file.
        if (!next_line(in, line)) break;
        char tag[256], name[256];
        if (sscanf(line.c_str(), "%255s %255s", tag, name) != 2) throw std::runtime_error("SU2: bad marker tag");
        Marker m;
        const std::string s(name);
        if (s == "WALL") m.type = BC_WALL;
        else if (s == "FARFIELD") m.type = BC_FARFIELD;
        else if (s == "SYMMETRY") m.type = BC_SYMMETRY;
        else if (s == "CONNEXION") m.type = BC_CONNEXION;
        else throw std::runtime_error("SU2: unknown marker " + s);
        next_line(in, line);
        const int ne = header_int(line, "MARKER_ELEMS=");
        for (int e = 0; e < ne; ++e) {
            next_line(in, line);
            std::istringstream ss(line);
            int t;
            ss >> t;
            const int nn = t == 5 ? 3 : t == 9 ? 4 : -1;
            if (nn < 0) throw std::runtime_error("SU2: unsupported marker element '" + line + "'");
            m.etype.push_back(t);
            for (int q = 0; q < nn; ++q) {
                int v;
                ss >> v;
                m.nodes.push_back(v);
            }
            m.off.push_back((int)m.nodes.size());
        }
        b.markers.push_back(std::move(m));
    }
    return b;
}

void write_topology(const MultiBlockMesh& m, const std::vector<std::string>& block_paths,
                    const std::string& path) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    fprintf(f, "NBLOCK= %d\n", (int)m.blocks.size());
    for (auto& p : block_paths) fprintf(f, "%s\n", p.c_str());
    for (size_t b = 0; b < m.blocks.size(); ++b) {
        fprintf(f, "Block= %d\nGhostCount= %d\nNconnexion= %d\n", (int)b, m.topo[b].ghost_count,
                (int)m.topo[b].interfaces.size());
        for (auto& it : m.topo[b].interfaces) {
            fprintf(f, "Nmark= %d\nNelems= %d\n", it.origin_block, (int)it.origin_cells.size());
            for (int c : it.origin_cells) fprintf(f, "%d\n", c);
        }
    }
    fclose(f);
}

MultiBlockMesh read_topology(const std::string& path) {
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("cannot open topology file " + path);
    std::string line;
    next_line(in, line);
    const int nb = header_int(line, "NBLOCK=");
    std::vector<std::string> files(nb);
    for (int b = 0; b < nb; ++b) {
        next_line(in, line);
        std::istringstream ss(line);
        ss >> files[b];
    }
    MultiBlockMesh m;
    m.blocks.resize(nb);
    m.topo.resize(nb);
    for (int b = 0; b < nb; ++b) {
        next_line(in, line);
        const int id = header_int(line, "Block=");
        if (id < 0 || id >= nb) throw std::runtime_error("topology: bad block id");
        next_line(in, line);
        m.topo[id].ghost_count = header_int(line, "GhostCount=");
        next_line(in, line);
        const int nc = header_int(line, "Nconnexion=");
        for (int c = 0; c < nc; ++c) {
            Interface it;
            next_line(in, line);
            it.origin_block = header_int(line, "Nmark=");
            next_line(in, line);
            const int ne = header_int(line, "Nelems=");
            for (int e = 0; e < ne; ++e) {
                next_line(in, line);
                it.origin_cells.push_back(atoi(line.c_str()));
            }
            m.topo[id].interfaces.push_back(std::move(it));
        }
    }
    // Block file paths are relative to the working directory in the reference;
    // here relative paths are also tried against the topology file's directory.
    std::string dir;
    const size_t slash = path.find_last_of('/');
    if (slash != std::string::npos) dir = path.substr(0, slash + 1);
    for (int b = 0; b < nb; ++b) {
        std::ifstream probe(files[b]);
        m.blocks[b] = read_su2(probe.good() || files[b][0] == '/' ? files[b] : dir + files[b]);
    }
    return m;
}

}  // namespace e3d
