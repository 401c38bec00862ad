// Builds faces, connectivity and geometry of one block in the reference's
// numbering and with the reference's floating-point expression order, so that
// every array equals, bit for bit, what the reference's builder and metrics
// produce from the same SU2 file.
//
//   face numbering  BlockBuilder/src/ConcreteBlockBuilder.cpp:497-742
//                   (first appearance over real cells 0..N-1 and the cell type's
//                   local face table; node order from that first cell). The
//                   reference compares every new face with every existing one
//                   (O(F^2)); a hash on the sorted node set gives the same ids.
//   face->cells     BlockBuilder/src/BlockBuilder.cpp:243-334 (cells holding the
//                   face's first three nodes, ascending, so left < right)
//   cell->faces/cells BlockBuilder/src/BlockBuilder.cpp:350-377 (face order)
//   BC face lists   BlockBuilder/src/BlockBuilder.cpp:379-440
//   metrics         Metrics/src/MetricsInitializer.cpp:27-800
//
// Compiled with -ffp-contract=off; every expression below keeps the operand
// order of the reference statement it restates.
#include <algorithm>
#include <climits>
#include <string>
#include <cmath>
#include <stdexcept>
#include "face_index.hpp"
#include "mesh.hpp"
#include "parallel.hpp"

namespace e3d {

namespace {

// Local face tables (FactoryMethod/src/TempTetrahedralFaceCreator.cpp:51,
// TempPrismFaceCreator.cpp:22 (8-node hexahedron), TempPyramidFaceCreator.cpp:43).
const int kTet[4][4] = {{0, 2, 1, -1}, {0, 1, 3, -1}, {1, 2, 3, -1}, {2, 0, 3, -1}};
const int kHex[6][4] = {{1, 0, 3, 2}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {0, 1, 5, 4}, {4, 5, 6, 7}};
const int kPyr[5][4] = {{0, 3, 2, 1}, {0, 1, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {3, 0, 4, -1}};

using FaceKey = FaceIndex<4>::Key;
FaceKey make_key(const int* n, int nn) {
    FaceKey k{{INT_MAX, INT_MAX, INT_MAX, INT_MAX}};
    for (int i = 0; i < nn; ++i) k.v[i] = n[i];
    std::sort(k.v, k.v + nn);
    return k;
}

}  // namespace

HostBlock build_block(const RawBlock& raw) {
    HostBlock B;
    const int nreal = raw.nelem();
    int nghost = 0;
    for (const Marker& m : raw.markers) nghost += m.size();
    B.nreal = nreal;
    B.nall = nreal + nghost;
    B.nnode = raw.npoin();

    // Cells: real elements, then one ghost per marker element (marker order).
    B.cell_nnode.resize(B.nall);
    B.cell_nodes_off.assign(B.nall + 1, 0);
    std::vector<int> nfaces(B.nall);
    for (int c = 0; c < nreal; ++c) {
        const int t = raw.etype[c];
        B.cell_nnode[c] = raw.eoff[c + 1] - raw.eoff[c];
        nfaces[c] = t == 10 ? 4 : t == 12 ? 6 : 5;
        B.cell_nodes.insert(B.cell_nodes.end(), raw.enodes.begin() + raw.eoff[c], raw.enodes.begin() + raw.eoff[c + 1]);
        B.cell_nodes_off[c + 1] = (int)B.cell_nodes.size();
    }
    int g = nreal;
    B.n_real_boundaries = 0;
    for (const Marker& m : raw.markers) {
        HostBlock::Boundary bd;
        bd.type = m.type;
        for (int e = 0; e < m.size(); ++e, ++g) {
            B.cell_nnode[g] = m.off[e + 1] - m.off[e];
            nfaces[g] = 1;
            B.cell_nodes.insert(B.cell_nodes.end(), m.nodes.begin() + m.off[e], m.nodes.begin() + m.off[e + 1]);
            B.cell_nodes_off[g + 1] = (int)B.cell_nodes.size();
            bd.ghosts.push_back(g);
        }
        if (m.type != BC_CONNEXION) ++B.n_real_boundaries;
        B.bnds.push_back(std::move(bd));
    }

    for (int v : B.cell_nodes)
        if (v < 0 || v >= B.nnode) throw std::runtime_error("mesh: element node index out of range");

    // Faces by first appearance over real cells.
    auto face_nodes_of = [&](int c, int j, int fn[4]) {
        const int t = raw.etype[c];
        const int(*tab)[4] = t == 10 ? kTet : t == 12 ? kHex : kPyr;
        const int* cn = &B.cell_nodes[B.cell_nodes_off[c]];
        int nn = 0;
        for (int q = 0; q < 4; ++q)
            if (tab[j][q] >= 0) fn[nn++] = cn[tab[j][q]];
        return nn;
    };
    FaceIndex<4> face_of(B.nnode, [&](auto&& visit) {
        for (int c = 0; c < nreal; ++c)
            for (int j = 0; j < nfaces[c]; ++j) {
                int fn[4];
                const int nn = face_nodes_of(c, j, fn);
                visit(make_key(fn, nn));
            }
    });
    std::vector<int> occ_cells;   // per face: cells holding it (at most 2 expected)
    std::vector<int> occ_count;
    for (int c = 0; c < nreal; ++c) {
        for (int j = 0; j < nfaces[c]; ++j) {
            int fn[4];
            const int nn = face_nodes_of(c, j, fn);
            const FaceKey key = make_key(fn, nn);
            bool is_new;
            const int f = face_of.insert_or_get(key, B.nface, &is_new);
            if (is_new) {
                ++B.nface;
                B.face_nnode.push_back(nn);
                for (int q = 0; q < 4; ++q) B.face_nodes.push_back(q < nn ? fn[q] : -1);
                occ_cells.push_back(c);
                occ_cells.push_back(-1);
                occ_count.push_back(1);
            } else {
                if (occ_count[f] >= 2) throw std::runtime_error("mesh: face shared by more than two cells");
                occ_cells[2 * f + occ_count[f]++] = c;
            }
        }
    }
    for (int gc = nreal; gc < B.nall; ++gc) {
        const FaceKey key = make_key(&B.cell_nodes[B.cell_nodes_off[gc]], B.cell_nnode[gc]);
        const int f = face_of.find(key);
        if (f < 0) throw std::runtime_error("mesh: boundary element matches no cell face");
        if (occ_count[f] >= 2) throw std::runtime_error("mesh: boundary element on an interior face");
        occ_cells[2 * f + occ_count[f]++] = gc;
    }
    // face -> cells: ascending cell index. A face with a single cell (a boundary
    // face without a marker element) would get cell 0 as its right cell in the
    // reference (zero-initialised array); it is rejected here instead.
    B.face2cell.resize(2 * (size_t)B.nface);
    for (int f = 0; f < B.nface; ++f) {
        if (occ_count[f] == 1)
            throw std::runtime_error("mesh: face " + std::to_string(f) + " has unmatched faces (open boundary: no marker element)");
        const int a = occ_cells[2 * f], b = occ_cells[2 * f + 1];
        B.face2cell[2 * f] = std::min(a, b);
        B.face2cell[2 * f + 1] = std::max(a, b);
    }

    // cell -> faces and cell -> cells in ascending face order.
    B.cell_off.assign(B.nall + 1, 0);
    for (int c = 0; c < B.nall; ++c) B.cell_off[c + 1] = B.cell_off[c] + nfaces[c];
    B.cell2face.assign(B.cell_off[B.nall], -1);
    B.cell2cell.assign(B.cell_off[B.nall], 0);
    std::vector<int> fill(B.nall, 0);
    for (int f = 0; f < B.nface; ++f) {
        const int L = B.face2cell[2 * f], R = B.face2cell[2 * f + 1];
        if (fill[L] < nfaces[L]) {
            B.cell2face[B.cell_off[L] + fill[L]] = f;
            B.cell2cell[B.cell_off[L] + fill[L]] = R;
            ++fill[L];
        }
        if (fill[R] < nfaces[R]) {
            B.cell2face[B.cell_off[R] + fill[R]] = f;
            B.cell2cell[B.cell_off[R] + fill[R]] = L;
            ++fill[R];
        }
    }
    for (int c = 0; c < B.nall; ++c)
        if (fill[c] != nfaces[c]) throw std::runtime_error("mesh: cell with unmatched faces (open boundary?)");

    // Boundary face lists in ghost order.
    for (const auto& bd : B.bnds)
        for (int gc : bd.ghosts) {
            const int f = B.cell2face[B.cell_off[gc]];
            switch (bd.type) {
                case BC_WALL: B.wall_faces.push_back(f); break;
                case BC_FARFIELD: B.farfield_faces.push_back(f); break;
                case BC_SYMMETRY: B.symmetry_faces.push_back(f); break;
                default: B.connexion_faces.push_back(f); break;
            }
        }

    // ---- metrics (MetricsInitializer::doInit order) ----
    const double* X = raw.xyz.data();
    B.ccenter.assign(3 * (size_t)B.nall, 0.0);
    parallel_for(nreal, [&](size_t c0, size_t c1) {
        for (int c = (int)c0; c < (int)c1; ++c) {  // computeCenterCells :64-88
            double sx = 0.0, sy = 0.0, sz = 0.0;
            const unsigned nn = (unsigned)B.cell_nnode[c];
            for (unsigned j = 0; j < nn; ++j) {
                const int n = B.cell_nodes[B.cell_nodes_off[c] + j];
                sx += X[3 * n];
                sy += X[3 * n + 1];
                sz += X[3 * n + 2];
            }
            B.ccenter[3 * c] = sx / nn;
            B.ccenter[3 * c + 1] = sy / nn;
            B.ccenter[3 * c + 2] = sz / nn;
        }
    });
    for (int c = nreal; c < B.nall; ++c) {  // :90-106
        const int l = B.cell2cell[B.cell_off[c]];
        for (int d = 0; d < 3; ++d) B.ccenter[3 * c + d] = B.ccenter[3 * l + d];
    }
    B.fcenter.assign(3 * (size_t)B.nface, 0.0);
    parallel_for(B.nface, [&](size_t f0, size_t f1) {
        for (int f = (int)f0; f < (int)f1; ++f) {  // computeCenterFaces :116-142
            double sx = 0.0, sy = 0.0, sz = 0.0;
            const unsigned nn = (unsigned)B.face_nnode[f];
            for (unsigned j = 0; j < nn; ++j) {
                const int n = B.face_nodes[4 * f + j];
                sx += X[3 * n];
                sy += X[3 * n + 1];
                sz += X[3 * n + 2];
            }
            B.fcenter[3 * f] = sx / nn;
            B.fcenter[3 * f + 1] = sy / nn;
            B.fcenter[3 * f + 2] = sz / nn;
        }
    });
    B.S.assign(3 * (size_t)B.nface, 0.0);
    parallel_for(B.nface, [&](size_t f0, size_t f1) {
        for (int f = (int)f0; f < (int)f1; ++f) {  // computeNormalFaces :337-448
            const int* fn = &B.face_nodes[4 * f];
            double s_x, s_y, s_z;
            if (B.face_nnode[f] == 3) {
                const double x1 = X[3 * fn[0]], x2 = X[3 * fn[1]], x3 = X[3 * fn[2]];
                const double y1 = X[3 * fn[0] + 1], y2 = X[3 * fn[1] + 1], y3 = X[3 * fn[2] + 1];
                const double z1 = X[3 * fn[0] + 2], z2 = X[3 * fn[1] + 2], z3 = X[3 * fn[2] + 2];
                const double dxyA = (x1 - x2) * (y1 + y2), dxyB = (x2 - x3) * (y2 + y3), dxyC = (x3 - x1) * (y3 + y1);
                const double dyzA = (y1 - y2) * (z1 + z2), dyzB = (y2 - y3) * (z2 + z3), dyzC = (y3 - y1) * (z3 + z1);
                const double dzxA = (z1 - z2) * (x1 + x2), dzxB = (z2 - z3) * (x2 + x3), dzxC = (z3 - z1) * (x3 + x1);
                s_x = 0.5 * (dyzA + dyzB + dyzC);
                s_y = 0.5 * (dzxA + dzxB + dzxC);
                s_z = 0.5 * (dxyA + dxyB + dxyC);
            } else {
                const double x1 = X[3 * fn[0]], x2 = X[3 * fn[1]], x3 = X[3 * fn[2]], x4 = X[3 * fn[3]];
                const double y1 = X[3 * fn[0] + 1], y2 = X[3 * fn[1] + 1], y3 = X[3 * fn[2] + 1], y4 = X[3 * fn[3] + 1];
                const double z1 = X[3 * fn[0] + 2], z2 = X[3 * fn[1] + 2], z3 = X[3 * fn[2] + 2], z4 = X[3 * fn[3] + 2];
                const double dxA = x4 - x2, dxB = x3 - x1, dyA = y4 - y2, dyB = y3 - y1, dzA = z4 - z2, dzB = z3 - z1;
                s_x = 0.5 * (dyA * dzB - dzA * dyB);
                s_y = 0.5 * (dzA * dxB - dxA * dzB);
                s_z = 0.5 * (dxA * dyB - dyA * dxB);
            }
            const int L = B.face2cell[2 * f];
            const double c0 = B.fcenter[3 * f] - B.ccenter[3 * L];
            const double c1 = B.fcenter[3 * f + 1] - B.ccenter[3 * L + 1];
            const double c2 = B.fcenter[3 * f + 2] - B.ccenter[3 * L + 2];
            const double dot = c0 * s_x + c1 * s_y + c2 * s_z;
            if (dot < 0) {
                s_x *= -1.0;
                s_y *= -1.0;
                s_z *= -1.0;
            }
            B.S[3 * f] = s_x;
            B.S[3 * f + 1] = s_y;
            B.S[3 * f + 2] = s_z;
        }
    });
    B.area.resize(B.nface);
    parallel_for(B.nface, [&](size_t f0, size_t f1) {
        for (int f = (int)f0; f < (int)f1; ++f) {  // computeAreaFaces :548-561
            const double s_x = B.S[3 * f], s_y = B.S[3 * f + 1], s_z = B.S[3 * f + 2];
            B.area[f] = std::sqrt(s_x * s_x + s_y * s_y + s_z * s_z);
        }
    });
    B.volume.assign(B.nall, 0.0);
    parallel_for(nreal, [&](size_t c0, size_t c1) {
        for (int c = (int)c0; c < (int)c1; ++c) {  // computeVolumeCells :575-624
            double volume = 0.0;
            for (int k = B.cell_off[c]; k < B.cell_off[c + 1]; ++k) {
                const int f = B.cell2face[k];
                double s_x = B.S[3 * f], s_y = B.S[3 * f + 1], s_z = B.S[3 * f + 2];
                const double cv0 = B.fcenter[3 * f] - B.ccenter[3 * c];
                const double cv1 = B.fcenter[3 * f + 1] - B.ccenter[3 * c + 1];
                const double cv2 = B.fcenter[3 * f + 2] - B.ccenter[3 * c + 2];
                const double dot = cv0 * s_x + cv1 * s_y + cv2 * s_z;
                if (dot < 0) {
                    s_x *= -1.0;
                    s_y *= -1.0;
                    s_z *= -1.0;
                }
                volume += (B.fcenter[3 * f] * s_x + B.fcenter[3 * f + 1] * s_y + B.fcenter[3 * f + 2] * s_z) / 3;
            }
            B.volume[c] = volume;
        }
    });
    for (int c = nreal; c < B.nall; ++c) B.volume[c] = B.volume[B.cell2cell[B.cell_off[c]]];  // :628-636
    B.rL.resize(3 * (size_t)B.nface);
    B.rR.resize(3 * (size_t)B.nface);
    parallel_for(B.nface, [&](size_t f0, size_t f1) {
        for (int f = (int)f0; f < (int)f1; ++f) {  // computeInterpVect :460-504
            const int L = B.face2cell[2 * f], R = B.face2cell[2 * f + 1];
            for (int d = 0; d < 3; ++d) {
                B.rL[3 * f + d] = B.fcenter[3 * f + d] - B.ccenter[3 * L + d];
                B.rR[3 * f + d] = B.fcenter[3 * f + d] - B.ccenter[3 * R + d];
            }
        }
    });
    for (int c = nreal; c < B.nall; ++c) {  // :509-540 reflected right vector of ghost faces
        const int f = B.cell2face[B.cell_off[c]];
        const double r_x = B.rL[3 * f], r_y = B.rL[3 * f + 1], r_z = B.rL[3 * f + 2];
        const double area = B.area[f];
        const double n_x = B.S[3 * f] / area, n_y = B.S[3 * f + 1] / area, n_z = B.S[3 * f + 2] / area;
        const double prod = r_x * n_x + r_y * n_y + r_z * n_z;
        B.rR[3 * f] = r_x - 2 * prod * n_x;
        B.rR[3 * f + 1] = r_y - 2 * prod * n_y;
        B.rR[3 * f + 2] = r_z - 2 * prod * n_z;
    }
    for (int c = nreal; c < B.nall; ++c) {  // computeCenterGhostCells :648-669
        const int f = B.cell2face[B.cell_off[c]];
        B.ccenter[3 * c] += 2.0 * B.rL[3 * f];
        B.ccenter[3 * c + 1] += 2.0 * B.rL[3 * f + 1];
        B.ccenter[3 * c + 2] += 2.0 * B.rL[3 * f + 2];
    }
    B.lsqw.assign(3 * (size_t)B.cell_off[nreal], 0.0);
    if (nreal > 1) {  // computeWLS :673-800 (Gram-Schmidt weights)
        parallel_for(nreal, [&](size_t c0, size_t c1) {
            std::vector<double> dx, dy, dz;
            for (int c = (int)c0; c < (int)c1; ++c) {
                const int k0 = B.cell_off[c], nb = B.cell_off[c + 1] - k0;
                dx.resize(nb);
                dy.resize(nb);
                dz.resize(nb);
                for (int j = 0; j < nb; ++j) {
                    const int o = B.cell2cell[k0 + j];
                    dx[j] = B.ccenter[3 * o] - B.ccenter[3 * c];
                    dy[j] = B.ccenter[3 * o + 1] - B.ccenter[3 * c + 1];
                    dz[j] = B.ccenter[3 * o + 2] - B.ccenter[3 * c + 2];
                }
                double sxx = 0.0, sxy = 0.0, sxz = 0.0, syy = 0.0, syz = 0.0, szz = 0.0;
                for (int j = 0; j < nb; ++j) {
                    sxx += dx[j] * dx[j];
                    sxy += dx[j] * dy[j];
                    sxz += dx[j] * dz[j];
                }
                const double r11 = std::sqrt(sxx);
                const double r12 = (1 / r11) * sxy;
                const double r13 = (1 / r11) * sxz;
                for (int j = 0; j < nb; ++j) {
                    syy += dy[j] * dy[j];
                    syz += dy[j] * dz[j];
                }
                const double r22 = std::sqrt(syy - r12 * r12);
                const double r23 = (1 / r22) * (syz - r12 * r13);
                for (int j = 0; j < nb; ++j) szz += dz[j] * dz[j];
                const double r33 = std::sqrt(szz - (r13 * r13 + r23 * r23));
                const double beta = (r12 * r23 - r13 * r22) / (r11 * r22);
                for (int j = 0; j < nb; ++j) {
                    const double alpha3 = (1 / (r33 * r33)) * (dz[j] - (r23 / r22) * dy[j] + beta * dx[j]);
                    const double alpha2 = (1 / (r22 * r22)) * (dy[j] - (r12 / r11) * dx[j]);
                    const double alpha1 = dx[j] / (r11 * r11);
                    B.lsqw[3 * (k0 + j)] = alpha1 - (r12 / r11) * alpha2 + beta * alpha3;
                    B.lsqw[3 * (k0 + j) + 1] = alpha2 - (r23 / r22) * alpha3;
                    B.lsqw[3 * (k0 + j) + 2] = alpha3;
                }
            }
        });
    }
    return B;
}

}  // namespace e3d
