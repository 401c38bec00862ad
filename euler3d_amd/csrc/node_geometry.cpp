// Per-cell node copies and slot codes of tet blocks (tet_geometry.hpp), built
// and checked on the host.
#include <atomic>
#include <cstring>

#include "mesh.hpp"
#include "parallel.hpp"
#include "tet_geometry.hpp"

namespace e3d {

int tet_slot_codes(const HostBlock& B, const RawBlock& R, double* node_xyz, unsigned* slot_code) {
    const double* X = R.xyz.data();
    std::atomic<int> status{0};
    auto same = [](const double* a, const double* b) { return std::memcmp(a, b, 3 * sizeof(double)) == 0; };
    parallel_for(B.nreal, [&](size_t q0, size_t q1) {
        for (int q = (int)q0; q < (int)q1; ++q) {
            if (status.load(std::memory_order_relaxed)) return;
            if (B.cell_nnode[q] != 4 || B.cell_off[q + 1] - B.cell_off[q] != 4) {
                status = 1;
                return;
            }
            const int* cn = &B.cell_nodes[B.cell_nodes_off[q]];
            double x[4], y[4], z[4], cc[3];
            for (int j = 0; j < 4; ++j) {
                x[j] = X[3 * cn[j]];
                y[j] = X[3 * cn[j] + 1];
                z[j] = X[3 * cn[j] + 2];
                node_xyz[12 * (size_t)q + 3 * j] = x[j];
                node_xyz[12 * (size_t)q + 3 * j + 1] = y[j];
                node_xyz[12 * (size_t)q + 3 * j + 2] = z[j];
            }
            tet_centre(x, y, z, cc);
            unsigned code = 0;
            for (int k = 0; k < 4; ++k) {
                const int f = B.cell2face[B.cell_off[q] + k];
                if (B.face_nnode[f] != 3) {
                    status = 1;
                    return;
                }
                unsigned byte = 0;
                for (int i = 0; i < 3; ++i) {
                    int l = -1;
                    for (int j = 0; j < 4; ++j)
                        if (cn[j] == B.face_nodes[4 * f + i]) l = j;
                    if (l < 0) {
                        status = 2;
                        return;
                    }
                    byte |= (unsigned)l << (2 * i);
                }
                const bool right = B.face2cell[2 * f] != q;
                double Ss[3], S3[3], r3[3];
                for (int d = 0; d < 3; ++d) Ss[d] = right ? B.S[3 * f + d] * -1.0 : B.S[3 * f + d];
                const double* Rs = right ? &B.rR[3 * f] : &B.rL[3 * f];
                tet_slot_geometry(byte, x, y, z, cc, S3, r3);
                if (!same(S3, Ss)) {
                    byte |= 64;
                    tet_slot_geometry(byte, x, y, z, cc, S3, r3);
                }
                if (!same(S3, Ss) || !same(r3, Rs)) {
                    status = 2;
                    return;
                }
                code |= byte << (8 * k);
            }
            slot_code[q] = code;
        }
    });
    return status.load();
}

}  // namespace e3d
