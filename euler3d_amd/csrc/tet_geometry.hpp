// Face geometry of a tetrahedral cell rebuilt from its four node coordinates,
// for kernels that would otherwise stream a stored copy per face slot. Every
// expression is the one the block builder (and the reference's metrics) uses,
// in the same order, so the doubles are the stored ones bit for bit:
//   cell centre  the cell's nodes summed from 0.0 in the cell's node order,
//                divided by their count (Metrics/src/MetricsInitializer.cpp:64-88);
//   face centre  the face's nodes summed from 0.0 in the face's node order,
//                divided by 3 (MetricsInitializer.cpp:116-142);
//   area vector  the triangle formula of MetricsInitializer.cpp:337-448 over
//                the face's nodes in the face's order; the builder may flip it
//                (:432-446) and the slot of the face's right cell holds it
//                negated, both exact negations that the slot code records;
//   r vector     face centre - cell centre (MetricsInitializer.cpp:460-504).
// Compiled with -ffp-contract=off on host and device alike, so no expression
// is fused into an FMA on either side.
//
// Slot code: one byte per face slot k (bits 8k .. 8k+7) of a 32-bit word per
// cell: the local indices (0-3, into the cell's node order) of the face's
// first, second and third node in bits 0-1, 2-3, 4-5, and bit 6 set when the
// slot's S is the formula's value negated. A real tet's slot byte is never
// zero (its three indices differ), so a zero word marks a padding cell.
#pragma once
#include <cmath>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define E3D_HD __host__ __device__
#else
#define E3D_HD
#endif

namespace e3d {

E3D_HD inline void tet_centre(const double x[4], const double y[4], const double z[4], double cc[3]) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int j = 0; j < 4; ++j) {
        sx += x[j];
        sy += y[j];
        sz += z[j];
    }
    cc[0] = sx / 4.0;
    cc[1] = sy / 4.0;
    cc[2] = sz / 4.0;
}

// Area vector of the triangle (p1, p2, p3) in that node order, negated when neg.
E3D_HD inline void tri_area_vector(double x1, double y1, double z1, double x2, double y2, double z2,
                                                double x3, double y3, double z3, bool neg, double S[3]) {
    const double dxyA = (x1 - x2) * (y1 + y2), dxyB = (x2 - x3) * (y2 + y3), dxyC = (x3 - x1) * (y3 + y1);
    const double dyzA = (y1 - y2) * (z1 + z2), dyzB = (y2 - y3) * (z2 + z3), dyzC = (y3 - y1) * (z3 + z1);
    const double dzxA = (z1 - z2) * (x1 + x2), dzxB = (z2 - z3) * (x2 + x3), dzxC = (z3 - z1) * (x3 + x1);
    const double sg = neg ? -1.0 : 1.0;  // exact: multiplying by -1 negates
    S[0] = 0.5 * (dyzA + dyzB + dyzC) * sg;
    S[1] = 0.5 * (dzxA + dzxB + dzxC) * sg;
    S[2] = 0.5 * (dxyA + dxyB + dxyC) * sg;
}

// x / 3 correctly rounded in three FMA-unit operations instead of a full
// division: q = x * RN(1/3), the exact remainder x - 3q (one FMA), and
// q + remainder * RN(1/3) rounded once. This is Markstein's correction step
// (P. Markstein, "Computation of elementary functions on the IBM RISC
// System/6000 processor", IBM J. Res. Dev. 34(1), 1990): with RN(1/3) within
// half an ulp of 1/3 and q within one ulp of x/3, the result is x/3 rounded
// to nearest. Exception: x = -0.0 gives +0.0; the face-centre sums below start
// from +0.0 and so are never -0.0.
E3D_HD inline double div3(double x) {
    const double z = 1.0 / 3.0;
    const double q = x * z;
    const double r = fma(-q, 3.0, x);
    return fma(r, z, q);
}

// r vector of the triangle (p1, p2, p3) (its centre, nodes summed in that
// order) from the cell centre cc.
E3D_HD inline void tri_r_vector(double x1, double y1, double z1, double x2, double y2, double z2,
                                             double x3, double y3, double z3, const double cc[3], double r[3]) {
    double fx = 0.0, fy = 0.0, fz = 0.0;
    fx += x1;
    fy += y1;
    fz += z1;
    fx += x2;
    fy += y2;
    fz += z2;
    fx += x3;
    fy += y3;
    fz += z3;
    r[0] = div3(fx) - cc[0];
    r[1] = div3(fy) - cc[1];
    r[2] = div3(fz) - cc[2];
}

// S (oriented as the slot stores it) and r of the face slot whose code byte is
// `code`, given the cell's node coordinates and centre.
E3D_HD inline void tet_slot_geometry(unsigned code, const double x[4], const double y[4],
                                                  const double z[4], const double cc[3], double S[3], double r[3]) {
    const int a = code & 3, b = (code >> 2) & 3, c = (code >> 4) & 3;
    tri_area_vector(x[a], y[a], z[a], x[b], y[b], z[b], x[c], y[c], z[c], (code & 64) != 0, S);
    tri_r_vector(x[a], y[a], z[a], x[b], y[b], z[b], x[c], y[c], z[c], cc, r);
}

struct HostBlock;
struct RawBlock;
// Host side: the node copy (12 doubles per real cell: x0 y0 z0 x1 .. z3, the
// cell's node order) and slot code of every real cell of a built block, each
// slot rebuilt with the functions above and compared bit for bit with the
// block's stored S (negated on the face's right cell) and rL / rR. Returns 0
// when every slot agrees, 1 when the block has a cell that is not a tet or a
// face that is not a triangle, 2 when a rebuilt value differs.
int tet_slot_codes(const HostBlock& B, const RawBlock& R, double* node_xyz, unsigned* slot_code);

}  // namespace e3d
