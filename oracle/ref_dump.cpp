// Golden-vector dumper. Container-only test tooling: links the reference's own
// classes (built from the reference's sources by oracle/Makefile into _ref/)
// and drives them exactly as the reference's main.cpp:17-101 and
// Solver::solve (Solver/src/Solver.cpp:36-98) do, writing raw arrays at chosen
// points. Never part of the product library, and nothing the product calls.
//
// usage: ref_dump <input_file> <out_dir> <iters> <stage_iter>
//   geometry of every block                 -> <out>/geom_b<k>_<name>.<f64|i32>
//   state after the residual of every stage of iteration stage_iter
//                                           -> <out>/it<stage_iter>s<s>_b<k>_<name>.f64
//   state after each iteration listed in iters (comma separated)
//                                           -> <out>/it<i>_b<k>_<name>.f64
//   monitor history every iteration         -> <out>/monitor.f64 (rows of 15)
#include "Initializer.h"
#include "Interface.h"
#include "Solver.h"
#include "BlockCommunicator.h"
#include "CompleteMesh.h"
#include "WallCellIds.h"
#include "FarfieldCellIds.h"
#include "SymmetryCellIds.h"
#include "ConnexionBlockCellIds.h"
#include <malloc.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include <sys/stat.h>
#include <unistd.h>

static std::string g_out;

static void put_f64(const std::string& name, const double* p, size_t n) {
    std::string path = g_out + "/" + name + ".f64";
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { perror(path.c_str()); exit(3); }
    fwrite(p, sizeof(double), n, f);
    fclose(f);
}
static void put_i32(const std::string& name, const int* p, size_t n) {
    std::string path = g_out + "/" + name + ".i32";
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { perror(path.c_str()); exit(3); }
    fwrite(p, sizeof(int), n, f);
    fclose(f);
}

static void dump_geometry(Block* b) {
    const std::string pre = "geom_b" + std::to_string(b->block_id_) + "_";
    const int nr = b->n_real_cells_in_block_, na = b->n_all_cells_in_block_, nf = b->n_faces_in_block_;
    int counts[8] = {nr, na, nf, b->n_nodes_in_block_, b->n_real_boundaries_in_block_,
                     b->n_all_boundaries_in_block_, b->n_wall_faces_, 0};
    put_i32(pre + "counts", counts, 8);
    std::vector<int> f2c(2 * nf), fnodes(5 * nf);
    std::vector<double> S(3 * nf), area(nf), fc(3 * nf), rl(3 * nf), rr(3 * nf);
    for (int f = 0; f < nf; ++f) {
        Face* F = b->block_faces_[f];
        f2c[2 * f] = F->face_2_cells_connectivity_[0];
        f2c[2 * f + 1] = F->face_2_cells_connectivity_[1];
        fnodes[5 * f] = F->n_nodes_per_face_;
        for (int k = 0; k < 4; ++k)
            fnodes[5 * f + 1 + k] = k < F->n_nodes_per_face_ ? F->face_2_nodes_connectivity_[k] : -1;
        for (int d = 0; d < 3; ++d) {
            S[3 * f + d] = F->face_normals_[d];
            fc[3 * f + d] = F->face_center_[d];
            rl[3 * f + d] = F->left_cell_r_vector_[d];
            rr[3 * f + d] = F->right_cell_r_vector_[d];
        }
        area[f] = F->face_area_;
    }
    put_i32(pre + "face2cell", f2c.data(), f2c.size());
    put_i32(pre + "face2node", fnodes.data(), fnodes.size());
    put_f64(pre + "S", S.data(), S.size());
    put_f64(pre + "area", area.data(), area.size());
    put_f64(pre + "fcenter", fc.data(), fc.size());
    put_f64(pre + "rL", rl.data(), rl.size());
    put_f64(pre + "rR", rr.data(), rr.size());
    std::vector<int> off(na + 1, 0), c2f, c2c;
    std::vector<double> vol(na), cc(3 * na), w;
    for (int c = 0; c < na; ++c) {
        Cell* C = b->block_cells_[c];
        off[c + 1] = off[c] + C->n_faces_per_cell_;
        for (int k = 0; k < C->n_faces_per_cell_; ++k) {
            c2f.push_back(C->cell_2_faces_connectivity_[k]);
            c2c.push_back(C->cell_2_cells_connectivity_[k]);
            if (c < nr)
                for (int d = 0; d < 3; ++d) w.push_back(C->cell_weights_[k][d]);
        }
        vol[c] = C->cell_volume_;
        for (int d = 0; d < 3; ++d) cc[3 * c + d] = C->cell_coordinates_[d];
    }
    put_i32(pre + "cell_off", off.data(), off.size());
    put_i32(pre + "cell2face", c2f.data(), c2f.size());
    put_i32(pre + "cell2cell", c2c.data(), c2c.size());
    put_f64(pre + "volume", vol.data(), vol.size());
    put_f64(pre + "ccenter", cc.data(), cc.size());
    put_f64(pre + "lsqw", w.data(), w.size());
    // boundaries: [type, count, ids...] for every marker, type 0 wall 1 farfield 2 symmetry 3 connexion
    std::vector<int> bnd;
    for (int j = 0; j < b->n_all_boundaries_in_block_; ++j) {
        BoundaryCellIds* B = b->block_boundary_cell_ids_[j];
        int t = -1;
        if (dynamic_cast<WallCellIds*>(B)) t = 0;
        else if (dynamic_cast<FarfieldCellIds*>(B)) t = 1;
        else if (dynamic_cast<SymmetryCellIds*>(B)) t = 2;
        else if (dynamic_cast<ConnexionBlockCellIds*>(B)) t = 3;
        bnd.push_back(t);
        bnd.push_back(B->n_cell_in_boundary_);
        for (int i = 0; i < B->n_cell_in_boundary_; ++i) bnd.push_back(B->cell_ids_in_boundary_[i]);
    }
    put_i32(pre + "boundaries", bnd.data(), bnd.size());
    put_i32(pre + "wallfaces", b->block_wall_face_ids_, b->n_wall_faces_);
    PrimitiveVariables* P = b->block_primitive_variables_;
    double fs[5] = {P->ro_free_, P->uu_free_, P->vv_free_, P->ww_free_, P->pp_free_};
    put_f64(pre + "freestream", fs, 5);
}

static void dump_state(Block* b, const std::string& tag) {
    const std::string pre = tag + "_b" + std::to_string(b->block_id_) + "_";
    const int na = b->n_all_cells_in_block_, nr = b->n_real_cells_in_block_;
    PrimitiveVariables* P = b->block_primitive_variables_;
    InterpolationVariables* I = b->block_interpolation_variables_;
    std::vector<double> v;
    auto cat5 = [&](double* a, double* bb, double* c, double* d, double* e, int n) {
        v.clear();
        for (double* x : {a, bb, c, d, e}) v.insert(v.end(), x, x + n);
    };
    cat5(P->ro_, P->uu_, P->vv_, P->ww_, P->pp_, na); put_f64(pre + "W", v.data(), v.size());
    cat5(P->conv_res_ro_, P->conv_res_uu_, P->conv_res_vv_, P->conv_res_ww_, P->conv_res_pp_, nr);
    put_f64(pre + "conv", v.data(), v.size());
    cat5(P->diss_res_ro_, P->diss_res_uu_, P->diss_res_vv_, P->diss_res_ww_, P->diss_res_pp_, nr);
    put_f64(pre + "diss", v.data(), v.size());
    cat5(P->ro_0_, P->ru_0_, P->rv_0_, P->rw_0_, P->re_0_, nr); put_f64(pre + "W0", v.data(), v.size());
    cat5(I->limiter_ro_, I->limiter_uu_, I->limiter_vv_, I->limiter_ww_, I->limiter_pp_, na);
    put_f64(pre + "lim", v.data(), v.size());
    v.clear();
    for (double** g : {I->grad_ro_, I->grad_uu_, I->grad_vv_, I->grad_ww_, I->grad_pp_})
        for (int d = 0; d < 3; ++d)
            for (int c = 0; c < na; ++c) v.push_back(g[c][d]);
    put_f64(pre + "grad", v.data(), v.size());
    put_f64(pre + "dt", b->block_time_variables_->dt_, nr);
    put_f64(pre + "specrad", b->block_time_variables_->spectral_radius_, nr);
}

int main(int argc, char* argv[]) {
    if (argc < 5) { fprintf(stderr, "usage: ref_dump input out_dir iters stage_iter\n"); return 2; }
    // Every allocation starts as zero bytes (glibc fills it with the complement
    // of 0xff). The reference reads one array it never writes, the w Jacobi
    // array of CentralIrsSmoother (Solver/src/CentralIrsSmoother.cpp:90); this
    // makes that read the zero a fresh allocation holds, on every call.
    mallopt(M_PERTURB, 0xff);
    g_out = argv[2];
    mkdir(g_out.c_str(), 0755);
    std::vector<int> dump_iters;
    for (const char* q = argv[3]; *q;) {
        dump_iters.push_back(atoi(q));
        while (*q && *q != ',') ++q;
        if (*q == ',') ++q;
    }
    const int stage_iter = atoi(argv[4]);
    mkdir("bin", 0755);  // the reference writes ./bin/Convergence.dat etc.

    Interface* interface = new Interface(argv[1], false);
    Initializer* initializer = new Initializer();
    int n_blocks = interface->n_blocks_interface_, n_mine;
    int* mine;
    BlockCommunicator* comm = new BlockCommunicator(n_blocks);
    comm->getMyBlocks(n_mine, mine);
    CompleteMesh* mesh = new CompleteMesh(n_blocks, n_mine, mine, interface->topology_file_name_interface_);
    mesh->InitializeMyBlocks(interface, comm);
    comm->initializeBuffers();
    for (int i = 0; i < n_mine; ++i) dump_geometry(mesh->all_blocks_[mine[i]]);

    Solver* solver = initializer->initializeSolver(interface);
    PostProcessing* pp = solver->post_processing_;
    RungeKutta* rk = solver->runge_kutta_;
    std::vector<double> mon;
    comm->updateBoundaries(mesh);  // Solver.cpp:48
    int iter = 0;
    while (!pp->stop_solver_) {    // Solver.cpp:76-98, with dump hooks
        for (int i = 0; i < n_mine; ++i) {
            Block* b = mesh->all_blocks_[mine[i]];
            solver->saveW0(b);
            solver->timestep_->computeSpectralRadius(b);
            solver->timestep_->computeTimestep(b);
            for (int s = 0; s < rk->stage_number_; ++s) {  // RungeKutta.cpp:80-94
                rk->residual_calculator_->current_stage_ = s;
                rk->residual_calculator_->computeResidual(b);
                if (iter == stage_iter) dump_state(b, "it" + std::to_string(iter) + "s" + std::to_string(s));
                rk->updater_->current_stage_ = s;
                rk->updater_->updateInternalBlock(b);
                rk->updater_->updateBoundary(b);
            }
            pp->computeFlowData(b);
        }
        comm->updateBoundaries(mesh);
        pp->process(mesh, comm);
        double row[15] = {(double)iter, pp->ro_rms_mesh_, pp->uu_rms_mesh_, pp->vv_rms_mesh_,
                          pp->ww_rms_mesh_, pp->pp_rms_mesh_, pp->cl_geometry_mesh_, pp->cd_geometry_mesh_,
                          pp->cmx_geometry_mesh_, pp->cmy_geometry_mesh_, pp->cmz_geometry_mesh_,
                          pp->ro_convergence_, pp->uu_convergence_, pp->pp_convergence_,
                          pp->stop_solver_ ? 1.0 : 0.0};
        mon.insert(mon.end(), row, row + 15);
        for (int b = 0; b < n_blocks; ++b) {
            double brow[10] = {pp->ro_rms_blocks_[b], pp->uu_rms_blocks_[b], pp->vv_rms_blocks_[b],
                               pp->ww_rms_blocks_[b], pp->pp_rms_blocks_[b], pp->cl_geometry_blocks_[b],
                               pp->cd_geometry_blocks_[b], pp->cmx_geometry_blocks_[b],
                               pp->cmy_geometry_blocks_[b], pp->cmz_geometry_blocks_[b]};
            if (iter < 3) put_f64("blockmon_it" + std::to_string(iter) + "_b" + std::to_string(b), brow, 10);
        }
        bool want = false;
        for (int d : dump_iters) want = want || d == iter;
        if (want)
            for (int i = 0; i < n_mine; ++i) dump_state(mesh->all_blocks_[mine[i]], "it" + std::to_string(iter));
        ++iter;
    }
    put_f64("monitor", mon.data(), mon.size());
    return 0;
}
