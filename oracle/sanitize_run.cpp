// Address/undefined-behaviour sanitizer run of the host code -- TEST
// INFRASTRUCTURE, built by `make -C oracle sanitize` with g++/gcc and
// -fsanitize=address,undefined (host code only; nothing here touches a GPU).
//
// Exercises the product's host mesh layer (bump-channel generator, SU2 and
// topology writer and reader, block builder, memory-order permutations) and
// the oracle's whole solver loop on the meshes it builds: every scheme
// (Roe/AUSM, first/second order, Green-Gauss/least squares,
// Venkatakrishnan/Barth-Jespersen, residual smoothing on and off) for a few
// iterations, single- and multi-block, threaded and not. Any invalid access or
// undefined operation aborts with the sanitizer's report; a clean run prints
// "sanitize ok".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../euler3d_amd/csrc/mesh.hpp"
#include "../euler3d_amd/csrc/ordering.hpp"
#include "e3d_oracle.h"

using namespace e3d;

namespace {

struct State {
    std::vector<double> W, conv, diss, W0, grad, lim, dt, specrad;
    std::vector<int> bnd_type, bnd_off, bnd_ghosts;
};

e3do_block make_block(const HostBlock& B, State& s) {
    const size_t na = B.nall, nr = B.nreal;
    s.W.assign(5 * na, 0.0);
    s.conv.assign(5 * na, 0.0);
    s.diss.assign(5 * na, 0.0);
    s.W0.assign(5 * na, 0.0);
    s.grad.assign(15 * na, 0.0);
    s.lim.assign(5 * na, 0.0);
    s.dt.assign(nr ? nr : 1, 0.0);
    s.specrad.assign(nr ? nr : 1, 0.0);
    s.bnd_off.assign(1, 0);
    for (const auto& bd : B.bnds) {
        s.bnd_type.push_back((int)bd.type);
        s.bnd_ghosts.insert(s.bnd_ghosts.end(), bd.ghosts.begin(), bd.ghosts.end());
        s.bnd_off.push_back((int)s.bnd_ghosts.size());
    }
    e3do_block b{};
    b.nreal = B.nreal;
    b.nall = B.nall;
    b.nface = B.nface;
    b.face2cell = B.face2cell.data();
    b.S = B.S.data();
    b.area = B.area.data();
    b.fcenter = B.fcenter.data();
    b.rL = B.rL.data();
    b.rR = B.rR.data();
    b.cell_off = B.cell_off.data();
    b.cell2face = B.cell2face.data();
    b.cell2cell = B.cell2cell.data();
    b.volume = B.volume.data();
    b.lsqw = B.lsqw.data();
    b.n_bnd = (int)B.bnds.size();
    b.n_real_bnd = B.n_real_boundaries;
    b.bnd_type = s.bnd_type.data();
    b.bnd_off = s.bnd_off.data();
    b.bnd_ghosts = s.bnd_ghosts.data();
    b.n_wall = (int)B.wall_faces.size();
    b.wall_faces = B.wall_faces.data();
    b.W = s.W.data();
    b.conv = s.conv.data();
    b.diss = s.diss.data();
    b.W0 = s.W0.data();
    b.grad = s.grad.data();
    b.lim = s.lim.data();
    b.dt = s.dt.data();
    b.specrad = s.specrad.data();
    return b;
}

int fail(const char* what) {
    std::fprintf(stderr, "sanitize: %s\n", what);
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string tmp = argc > 1 ? argv[1] : "/tmp";
    const BumpParams cases[] = {{6, 3, 2, 0.1, 4, 1, 1, 1}, {8, 3, 3, 0.06, -3, 2, 1, 1}, {8, 4, 4, 0.12, 0, 2, 2, 1}};
    int ci = 0;
    for (const BumpParams& bp : cases) {
        // generator -> files -> reader -> builder
        MultiBlockMesh gen = make_bump_channel(bp);
        std::vector<std::string> paths;
        for (size_t b = 0; b < gen.blocks.size(); ++b) {
            paths.push_back(tmp + "/san" + std::to_string(ci) + "_" + std::to_string(b) + ".su2");
            write_su2(gen.blocks[b], paths.back());
        }
        const std::string top = tmp + "/san" + std::to_string(ci) + ".top";
        write_topology(gen, paths, top);
        MultiBlockMesh m = read_topology(top);
        if (m.blocks.size() != gen.blocks.size()) return fail("topology round trip");
        std::vector<HostBlock> built;
        for (const RawBlock& rb : m.blocks) built.push_back(build_block(rb));
        for (const HostBlock& B : built)
            for (int mode : {ORDER_REFERENCE, ORDER_RCM, ORDER_MORTON}) {
                std::vector<int> cp, fp;
                block_layout(B, mode, cp, fp);
                std::vector<char> seen(B.nreal, 0);
                for (int p : cp) {
                    if (p < 0 || p >= B.nreal || seen[p]) return fail("cell layout is not a permutation");
                    seen[p] = 1;
                }
            }
        // interfaces: destination ghosts numbered from GhostCount in topology order
        std::vector<std::vector<int>> dst(built.size());
        std::vector<e3do_interface> itf;
        for (size_t d = 0; d < built.size(); ++d) {
            int k = 0;
            for (const Interface& I : m.topo[d].interfaces)
                for (size_t q = 0; q < I.origin_cells.size(); ++q) dst[d].push_back(m.topo[d].ghost_count + k++);
        }
        for (size_t d = 0; d < built.size(); ++d) {
            int pos = 0;
            for (const Interface& I : m.topo[d].interfaces) {
                const int n = (int)I.origin_cells.size();
                itf.push_back({(int)d, I.origin_block, n, dst[d].data() + pos, I.origin_cells.data()});
                pos += n;
            }
        }
        // every scheme of the reference's factories, a few iterations each
        for (int flux = 0; flux < 2; ++flux)
            for (int order = 1; order <= 2; ++order)
                for (int grad = 0; grad < 2; ++grad)
                    for (int lim = 0; lim < 2; ++lim)
                        for (int smooth = 0; smooth < 2; ++smooth) {
                            if (order == 1 && (grad || lim)) continue;
                            std::vector<State> st(built.size());
                            std::vector<e3do_block> blocks;
                            for (size_t b = 0; b < built.size(); ++b) blocks.push_back(make_block(built[b], st[b]));
                            e3do_params p{1.4, 1.5, 2.0, 0.6, 1.0, 1e-16, 5.0, -15.0, 5, order, flux, grad, lim, 50,
                                          smooth};
                            double fs[5];
                            e3do_freestream(&p, fs);
                            for (auto& b : blocks) e3do_init_flow(&b, fs);
                            int it = 0;
                            double rms0[5];
                            e3do_monitor hist[3];
                            const int done = e3do_iterate_mt(blocks.data(), (int)blocks.size(), itf.data(),
                                                             (int)itf.size(), &p, 3, &it, rms0, hist, 1 + (ci & 1));
                            if (done != 3 || !std::isfinite(hist[2].rms[0])) return fail("oracle run");
                        }
        ++ci;
    }
    std::printf("sanitize ok\n");
    return 0;
}
