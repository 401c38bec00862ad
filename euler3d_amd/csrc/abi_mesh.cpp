// C ABI: errors, host mesh building and the input-file parser. No HIP here.
#include <algorithm>
#include <atomic>
#include <exception>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <thread>

#include "../../include/euler3d.h"
#include "abi_internal.hpp"
#include "mesh.hpp"
#include "ordering.hpp"
#include "mesh_internal.hpp"
#include "tet_geometry.hpp"

namespace e3d {
thread_local std::string g_thread_error;
void set_thread_error(const std::string& s) { g_thread_error = s; }
// The device part of the library stores the message on its context; this weak
// default serves builds without it.
__attribute__((weak)) void ctx_record_error(e3d_ctx*, const std::string&) {}
__attribute__((weak)) const char* ctx_error_text(const e3d_ctx*) { return nullptr; }
int set_error(e3d_ctx* ctx, int code, const std::string& msg) {
    g_thread_error = msg;
    if (ctx) ctx_record_error(ctx, msg);
    return code;
}
}  // namespace e3d

using namespace e3d;

namespace {

// Builds every block, several at a time on host threads (blocks are
// independent); the first failure, in block order, is rethrown.
void build_blocks(e3d_mesh& m) {
    const size_t nb = m.raw.blocks.size();
    m.built.resize(nb);
    std::vector<std::exception_ptr> err(nb);
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t b; (b = next.fetch_add(1)) < nb;) {
            try {
                m.built[b] = build_block(m.raw.blocks[b]);
            } catch (...) {
                err[b] = std::current_exception();
            }
        }
    };
    const size_t nt = std::min<size_t>(nb, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    try {
        for (size_t t = 1; t < nt; ++t) pool.emplace_back(work);
    } catch (...) {  // no more threads: the ones started and this one take the remaining blocks
    }
    work();
    for (auto& t : pool) t.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

int finish_mesh(std::unique_ptr<e3d_mesh>& m, e3d_mesh** out) {
    const size_t nb = m->raw.blocks.size();
    build_blocks(*m);
    m->flat.resize(nb);
    for (size_t b = 0; b < nb; ++b) {
        const HostBlock& B = m->built[b];
        auto& F = m->flat[b];
        F.bnd_off.push_back(0);
        for (auto& bd : B.bnds) {
            F.bnd_type.push_back(bd.type);
            F.bnd_ghosts.insert(F.bnd_ghosts.end(), bd.ghosts.begin(), bd.ghosts.end());
            F.bnd_off.push_back((int)F.bnd_ghosts.size());
        }
        F.itf_off.push_back(0);
        for (auto& it : m->raw.topo[b].interfaces) {
            F.itf_origin_block.push_back(it.origin_block);
            F.itf_origin_cells.insert(F.itf_origin_cells.end(), it.origin_cells.begin(), it.origin_cells.end());
            F.itf_off.push_back((int)F.itf_origin_cells.size());
        }
        // The topology numbers interface ghosts from GhostCount; that must be
        // the first CONNEXION ghost of the block file.
        int first_connexion = B.nall, nconn = 0;
        for (auto& bd : B.bnds)
            if (bd.type == BC_CONNEXION) {
                if (!bd.ghosts.empty()) first_connexion = std::min(first_connexion, bd.ghosts.front());
                nconn += (int)bd.ghosts.size();
            }
        if (nconn != (int)F.itf_origin_cells.size() || (nconn > 0 && first_connexion != m->raw.topo[b].ghost_count)) {
            set_thread_error("topology: block " + std::to_string(b) + " interfaces do not match its CONNEXION markers");
            return E3D_EIO;
        }
        for (size_t i = 0; i < F.itf_origin_block.size(); ++i) {
            const int o = F.itf_origin_block[i];
            if (o < 0 || o >= (int)nb) {
                set_thread_error("topology: bad origin block");
                return E3D_EIO;
            }
        }
    }
    for (size_t b = 0; b < nb; ++b)
        for (size_t i = 0; i < m->flat[b].itf_origin_block.size(); ++i) {
            const int o = m->flat[b].itf_origin_block[i];
            for (int k = m->flat[b].itf_off[i]; k < m->flat[b].itf_off[i + 1]; ++k) {
                const int c = m->flat[b].itf_origin_cells[k];
                if (c < 0 || c >= m->built[o].nreal) {
                    set_thread_error("topology: interface origin cell out of range");
                    return E3D_EIO;
                }
            }
        }
    *out = m.release();
    return E3D_OK;
}

}  // namespace

extern "C" {

const char* e3d_last_error(const e3d_ctx* ctx) {
    if (ctx) {
        const char* t = ctx_error_text(ctx);
        if (t) return t;
    }
    return g_thread_error.c_str();
}

const char* e3d_version(void) { return "euler3d_amd 0.1 (gfx950, fp64)"; }

int e3d_mesh_read(const char* topology_path, e3d_mesh** out) {
    if (!topology_path || !out) return set_error(nullptr, E3D_EINVAL, "e3d_mesh_read: null argument");
    try {
        auto m = std::make_unique<e3d_mesh>();
        m->raw = read_topology(topology_path);
        return finish_mesh(m, out);
    } catch (const std::bad_alloc&) {
        return set_error(nullptr, E3D_ENOMEM, "e3d_mesh_read: out of memory");
    } catch (const std::exception& e) {
        return set_error(nullptr, E3D_EIO, e.what());
    }
}

int e3d_mesh_bump(const e3d_bump_desc* d, e3d_mesh** out) {
    if (!d || !out) return set_error(nullptr, E3D_EINVAL, "e3d_mesh_bump: null argument");
    try {
        BumpParams p;
        p.nx = d->nx; p.ny = d->ny; p.nz = d->nz; p.h = d->h; p.brick = d->brick;
        p.px = d->px; p.py = d->py; p.pz = d->pz;
        auto m = std::make_unique<e3d_mesh>();
        m->raw = make_bump_channel(p);
        return finish_mesh(m, out);
    } catch (const std::bad_alloc&) {
        return set_error(nullptr, E3D_ENOMEM, "e3d_mesh_bump: out of memory");
    } catch (const std::exception& e) {
        return set_error(nullptr, E3D_EINVAL, e.what());
    }
}

int e3d_mesh_write(const e3d_mesh* m, const char* dir, const char* stem) {
    if (!m || !dir || !stem) return set_error(nullptr, E3D_EINVAL, "e3d_mesh_write: null argument");
    try {
        std::vector<std::string> paths;
        for (size_t b = 0; b < m->raw.blocks.size(); ++b) {
            std::string p = std::string(dir) + "/" + stem + "_" + std::to_string(b) + ".su2";
            write_su2(m->raw.blocks[b], p);
            paths.push_back(p);
        }
        write_topology(m->raw, paths, std::string(dir) + "/" + stem + ".top");
        return E3D_OK;
    } catch (const std::exception& e) {
        return set_error(nullptr, E3D_EIO, e.what());
    }
}

int e3d_mesh_nblocks(const e3d_mesh* m) { return m ? (int)m->built.size() : E3D_EINVAL; }

int e3d_mesh_layout(const e3d_mesh* m, int b, int mode, int* cell_pos, int* face_pos) {
    if (!m || b < 0 || b >= (int)m->built.size() || !cell_pos || !face_pos)
        return set_error(nullptr, E3D_EINVAL, "e3d_mesh_layout: bad argument");
    if (mode != E3D_ORDER_REFERENCE && mode != E3D_ORDER_RCM && mode != E3D_ORDER_MORTON)
        return set_error(nullptr, E3D_EINVAL, "e3d_mesh_layout: unknown mode " + std::to_string(mode));
    const HostBlock& B = m->built[b];
    std::vector<int> cp, fp;
    block_layout(B, mode, cp, fp);
    std::copy(cp.begin(), cp.end(), cell_pos);
    std::copy(fp.begin(), fp.end(), face_pos);
    return E3D_OK;
}

int e3d_mesh_tet_geometry(const e3d_mesh* m, int b, double* node_xyz, unsigned* slot_code) {
    if (!m || b < 0 || b >= (int)m->built.size() || !node_xyz || !slot_code)
        return set_error(nullptr, E3D_EINVAL, "e3d_mesh_tet_geometry: bad argument");
    const int rc = tet_slot_codes(m->built[b], m->raw.blocks[b], node_xyz, slot_code);
    if (rc == 1) return set_error(nullptr, E3D_EINVAL, "e3d_mesh_tet_geometry: the block has cells other than tets");
    if (rc == 2)
        return set_error(nullptr, E3D_EINVAL, "e3d_mesh_tet_geometry: a rebuilt slot differs from the stored metrics");
    return E3D_OK;
}

int e3d_mesh_block(const e3d_mesh* m, int b, e3d_block_view* v) {
    if (!m || !v || b < 0 || b >= (int)m->built.size())
        return set_error(nullptr, E3D_EINVAL, "e3d_mesh_block: bad argument");
    const HostBlock& B = m->built[b];
    const auto& F = m->flat[b];
    std::memset(v, 0, sizeof(*v));
    v->nreal = B.nreal; v->nall = B.nall; v->nface = B.nface; v->nnode = B.nnode;
    v->n_boundaries = (int)B.bnds.size();
    v->n_real_boundaries = B.n_real_boundaries;
    v->n_wall_faces = (int)B.wall_faces.size();
    v->face2cell = B.face2cell.data(); v->face_nnode = B.face_nnode.data(); v->face_nodes = B.face_nodes.data();
    v->S = B.S.data(); v->area = B.area.data(); v->fcenter = B.fcenter.data();
    v->rL = B.rL.data(); v->rR = B.rR.data();
    v->cell_off = B.cell_off.data(); v->cell2face = B.cell2face.data(); v->cell2cell = B.cell2cell.data();
    v->volume = B.volume.data(); v->ccenter = B.ccenter.data(); v->lsqw = B.lsqw.data();
    v->bnd_type = F.bnd_type.data(); v->bnd_off = F.bnd_off.data(); v->bnd_ghosts = F.bnd_ghosts.data();
    v->wall_faces = B.wall_faces.data();
    v->ghost_count = m->raw.topo[b].ghost_count;
    v->n_interfaces = (int)F.itf_origin_block.size();
    v->itf_origin_block = F.itf_origin_block.data();
    v->itf_off = F.itf_off.data();
    v->itf_origin_cells = F.itf_origin_cells.data();
    v->cell_node_off = B.cell_nodes_off.data();
    v->cell_nodes = B.cell_nodes.data();
    v->node_xyz = m->raw.blocks[b].xyz.data();
    return E3D_OK;
}

void e3d_mesh_free(e3d_mesh* m) { delete m; }

int e3d_read_input(const char* path, e3d_params* p, char* topo, int topo_len) {
    if (!path || !p) return set_error(nullptr, E3D_EINVAL, "e3d_read_input: null argument");
    std::ifstream in(path);
    if (!in.is_open()) return set_error(nullptr, E3D_EIO, std::string("cannot open input file ") + path);
    // Positional layout of Solver/src/Interface.cpp:33-101.
    std::string line, topology, flux, grad, lim;
    int smoothing = 0;
    in.ignore(256, '\n');
    std::getline(in, line);
    in >> topology;
    in.ignore(256, '\n');
    in.ignore(256, '\n');
    std::getline(in, line);
    in >> p->cfl >> p->gamma >> p->aoa_deg;
    in.ignore(256, '\n');
    std::getline(in, line);
    in >> p->stages >> p->mach >> p->cmac;
    in.ignore(256, '\n');
    std::getline(in, line);
    in >> p->max_iter >> p->criterion;
    in.ignore(256, '\n');
    std::getline(in, line);
    in >> smoothing;
    in.ignore(256, '\n');
    std::getline(in, line);
    in >> flux >> p->order;
    in.ignore(256, '\n');
    std::getline(in, line);
    in >> grad >> lim;
    in.ignore(256, '\n');
    std::getline(in, line);
    in >> p->omega >> p->k;
    if (in.fail()) return set_error(nullptr, E3D_EIO, std::string("malformed input file ") + path);
    p->smoothing = smoothing;
    // Factory semantics: "roe"/"ausm" (anything else is an error here; the
    // reference leaves a null scheme), order 2 = second order else first,
    // "leastsquares" else Green-Gauss, "venkatakrishnan" else Barth-Jespersen.
    if (flux == "roe") p->flux = E3D_FLUX_ROE;
    else if (flux == "ausm") p->flux = E3D_FLUX_AUSM;
    else return set_error(nullptr, E3D_EINVAL, "unknown flux scheme '" + flux + "'");
    p->gradient = grad == "leastsquares" ? E3D_GRAD_LEAST_SQUARES : E3D_GRAD_GREEN_GAUSS;
    p->limiter = lim == "venkatakrishnan" ? E3D_LIM_VENKATAKRISHNAN : E3D_LIM_BARTH_JESPERSEN;
    if (topo && topo_len > 0) {
        std::snprintf(topo, (size_t)topo_len, "%s", topology.c_str());
        if ((int)topology.size() >= topo_len) return set_error(nullptr, E3D_EINVAL, "topology path buffer too small");
    }
    return E3D_OK;
}

}  // extern "C"
