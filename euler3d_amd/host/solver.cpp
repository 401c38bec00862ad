// Host classes of euler3d_solver.h: the reference's Solver-layer objects, each
// compute method one call into the C ABI on the block it is given.
#include "euler3d_solver.h"

#include <sys/stat.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <fstream>
#include <iostream>

int e3d_check(int rc, e3d_ctx* ctx) {
    if (rc < 0) throw Euler3dError(rc, e3d_last_error(ctx));
    return rc;
}

// ------------------------------------------------------------------ Block

Block::~Block() {
    if (block_boundary_cell_ids_) {
        for (int j = 0; j < n_all_boundaries_in_block_; ++j) delete block_boundary_cell_ids_[j];
        delete[] block_boundary_cell_ids_;
    }
}

std::vector<double> Block::get(int which) const {
    size_t n = 0;
    switch (which) {
        case E3D_ARR_W: case E3D_ARR_LIM: n = 5 * (size_t)n_all_cells_in_block_; break;
        case E3D_ARR_GRAD: n = 15 * (size_t)n_all_cells_in_block_; break;
        case E3D_ARR_DT: case E3D_ARR_SPECRAD: n = (size_t)n_real_cells_in_block_; break;
        default: n = 5 * (size_t)n_real_cells_in_block_; break;
    }
    std::vector<double> out(n);
    e3d_check(e3d_get_array(ctx_, block_id_, which, out.data()), ctx_);
    return out;
}

void Block::set(int which, const std::vector<double>& values) {
    e3d_check(e3d_set_array(ctx_, block_id_, which, values.data()), ctx_);
}

void WallCellIds::updateBoundary() { e3d_check(e3d_bc_marker(owner_block_->ctx_, owner_block_->block_id_, marker_), owner_block_->ctx_); }
void SymmetryCellIds::updateBoundary() { e3d_check(e3d_bc_marker(owner_block_->ctx_, owner_block_->block_id_, marker_), owner_block_->ctx_); }
void FarfieldCellIds::updateBoundary() { e3d_check(e3d_bc_marker(owner_block_->ctx_, owner_block_->block_id_, marker_), owner_block_->ctx_); }

// -------------------------------------------------------------- Interface

Interface::Interface(const char* input_file, bool i_write) {
    char topo[4096];
    e3d_check(e3d_read_input(input_file, &params_, topo, sizeof topo));
    topology_file_name_interface_ = topo;
    cfl_interface_ = params_.cfl;
    gamma_interface_ = params_.gamma;
    aoa_deg_interface_ = params_.aoa_deg;
    stage_number_interface_ = params_.stages;
    mach_aircraft_interface_ = params_.mach;
    cmac_interface_ = params_.cmac;
    max_iter_interface_ = params_.max_iter;
    convergence_criterion_interface_ = params_.criterion;
    residual_smoother_choice_interface_ = params_.smoothing != 0;
    flux_scheme_choice_interface_ = params_.flux == E3D_FLUX_AUSM ? "ausm" : "roe";
    interpolation_choice_interface_ = params_.order;
    gradient_choice_interface_ = params_.gradient == E3D_GRAD_LEAST_SQUARES ? "leastsquares" : "greengauss";
    limiter_choice_interface_ = params_.limiter == E3D_LIM_VENKATAKRISHNAN ? "venkatakrishnan" : "barthjespersen";
    omega_interface_ = params_.omega;
    k_interface_ = params_.k;
    // Interface::readNBlocks (Solver/src/Interface.cpp:239-266): first line of the topology file
    std::ifstream top(topology_file_name_interface_);
    std::string tag;
    if (!(top >> tag >> n_blocks_interface_))
        throw Euler3dError(E3D_EIO, "cannot read NBLOCK= from " + topology_file_name_interface_);
    if (i_write)
        std::cout << "Reading Input File...\ntopology " << topology_file_name_interface_ << ", " << n_blocks_interface_
                  << " blocks\n";
}

// ----------------------------------------------------------- CompleteMesh

CompleteMesh::CompleteMesh(int n_blocks, int n_blocks_in_process, int* my_blocks, std::string topology_file_name,
                           int device)
    : n_blocks_(n_blocks), n_blocks_in_process_(n_blocks_in_process), topology_file_name_(topology_file_name) {
    my_blocks_ = new int[n_blocks_in_process_];
    for (int i = 0; i < n_blocks_in_process_; ++i) my_blocks_[i] = my_blocks[i];
    all_blocks_ = new Block*[n_blocks_];
    for (int j = 0; j < n_blocks_; ++j) all_blocks_[j] = new Block(j);
    e3d_check(e3d_create(device, &ctx_));
}

CompleteMesh::~CompleteMesh() {
    for (int j = 0; j < n_blocks_; ++j) delete all_blocks_[j];
    delete[] all_blocks_;
    delete[] my_blocks_;
    if (mesh_) e3d_mesh_free(mesh_);
    if (ctx_) e3d_destroy(ctx_);
}

void CompleteMesh::InitializeMyBlocks(Interface* interface, BlockCommunicator*) {
    e3d_check(e3d_mesh_read(topology_file_name_.c_str(), &mesh_));
    if (e3d_mesh_nblocks(mesh_) != n_blocks_) throw Euler3dError(E3D_EIO, "topology block count mismatch");
    e3d_check(e3d_set_params(ctx_, &interface->params_), ctx_);
    e3d_check(e3d_set_ordering(ctx_, memory_order_), ctx_);
    e3d_check(e3d_upload_mesh(ctx_, mesh_), ctx_);
    for (int b = 0; b < n_blocks_; ++b) {
        e3d_block_view v;
        e3d_check(e3d_mesh_block(mesh_, b, &v));
        Block* B = all_blocks_[b];
        B->ctx_ = ctx_;
        B->n_real_cells_in_block_ = v.nreal;
        B->n_all_cells_in_block_ = v.nall;
        B->n_ghost_cells_in_block_ = v.nall - v.nreal;
        B->n_nodes_in_block_ = v.nnode;
        B->n_faces_in_block_ = v.nface;
        B->n_real_boundaries_in_block_ = v.n_real_boundaries;
        B->n_all_boundaries_in_block_ = v.n_boundaries;
        B->n_wall_faces_ = v.n_wall_faces;
        B->view_ = v;
        B->block_boundary_cell_ids_ = new BoundaryCellIds*[v.n_boundaries];
        for (int j = 0; j < v.n_boundaries; ++j) {
            BoundaryCellIds* bc;
            switch (v.bnd_type[j]) {
                case 0: bc = new WallCellIds; break;
                case 1: bc = new FarfieldCellIds; break;
                case 2: bc = new SymmetryCellIds; break;
                default: bc = new ConnexionBlockCellIds; break;
            }
            bc->marker_ = j;
            bc->n_cell_in_boundary_ = v.bnd_off[j + 1] - v.bnd_off[j];
            bc->owner_block_ = B;
            B->block_boundary_cell_ids_[j] = bc;
        }
    }
    // calculateFreeVariables + initializeFlowField (DataStructure/src/CompleteMesh.cpp:157-162)
    e3d_check(e3d_init_freestream(ctx_), ctx_);
}

// ------------------------------------------------------ BlockCommunicator

void BlockCommunicator::updateBoundaries(CompleteMesh* mesh) const { e3d_check(e3d_exchange(mesh->ctx_), mesh->ctx_); }

void BlockCommunicator::getMyBlocks(int& n_blocks_in_process, int*& my_blocks) const {
    n_blocks_in_process = n_blocks_;  // one device holds every block
    my_blocks = new int[n_blocks_];
    for (int i = 0; i < n_blocks_; ++i) my_blocks[i] = i;
}

void BlockCommunicator::getGlobal(CompleteMesh* mesh, PostProcessing* postprocess) {
    e3d_monitor m;
    e3d_check(e3d_monitors(mesh->ctx_, &m), mesh->ctx_);
    postprocess->last_ = m;
}

// ------------------------------------------------------------ flux schemes

void FluxScheme::computeFluxConv(Block* b) { e3d_check(e3d_flux_conv(b->ctx_, b->block_id_, scheme_), b->ctx_); }

RoeScheme::RoeScheme(double gamma) {
    gamma_ = gamma;
    current_beta_ = 1.0;
    scheme_ = E3D_FLUX_ROE;
}
void RoeScheme::computeFluxDiss(Block* b) {
    e3d_check(e3d_flux_diss(b->ctx_, b->block_id_, E3D_FLUX_ROE, current_beta_), b->ctx_);
}

AusmScheme::AusmScheme(double gamma) {
    gamma_ = gamma;
    current_beta_ = 1.0;
    scheme_ = E3D_FLUX_AUSM;
}
void AusmScheme::computeFluxConv(Block* b) { e3d_check(e3d_flux_conv(b->ctx_, b->block_id_, E3D_FLUX_AUSM), b->ctx_); }
void AusmScheme::computeFluxDiss(Block* b) {
    e3d_check(e3d_flux_diss(b->ctx_, b->block_id_, E3D_FLUX_AUSM, current_beta_), b->ctx_);
}

void GreenGauss::computeGradients(Block* b) {
    e3d_check(e3d_gradients(b->ctx_, b->block_id_, E3D_GRAD_GREEN_GAUSS), b->ctx_);
}
void LeastSquares::computeGradients(Block* b) {
    e3d_check(e3d_gradients(b->ctx_, b->block_id_, E3D_GRAD_LEAST_SQUARES), b->ctx_);
}
void BarthJespersen::computeLimiters(Block* b) {
    e3d_check(e3d_limiters(b->ctx_, b->block_id_, E3D_LIM_BARTH_JESPERSEN, omega_), b->ctx_);
}
void Venkatakrishnan::computeLimiters(Block* b) {
    e3d_check(e3d_limiters(b->ctx_, b->block_id_, E3D_LIM_VENKATAKRISHNAN, k_), b->ctx_);
}

// Factories of Solver/src/SecondOrder.cpp:50-61: "leastsquares" else Green-Gauss,
// "venkatakrishnan" else Barth-Jespersen.
SecondOrder::SecondOrder(std::string gradient_choice, std::string limiter_choice, double omega, double k) {
    setGradient(gradient_choice);
    setLimiter(limiter_choice, omega, k);
}
SecondOrder::~SecondOrder() {
    delete gradient_;
    delete limiter_;
}
void SecondOrder::setGradient(std::string g) {
    delete gradient_;
    gradient_ = g == "leastsquares" ? (Gradient*)new LeastSquares() : (Gradient*)new GreenGauss();
}
void SecondOrder::setLimiter(std::string l, double omega, double k) {
    delete limiter_;
    limiter_ = l == "venkatakrishnan" ? (Limiter*)new Venkatakrishnan(k) : (Limiter*)new BarthJespersen(omega);
}
void SecondOrder::computeInterpolation(Block* b) {  // Solver/src/SecondOrder.cpp:20-30
    gradient_->computeGradients(b);
    limiter_->computeLimiters(b);
}

// ------------------------------------------------------- residual + RK

ResidualCalculator::ResidualCalculator(double gamma, std::vector<double> beta_rk, int interpolation_choice,
                                       std::string gradient_choice, std::string limiter_choice,
                                       std::string flux_scheme_choice, bool residual_smoother_choice, double omega,
                                       double k)
    : beta_rk_(beta_rk) {
    setInterpolation(interpolation_choice, gradient_choice, limiter_choice, omega, k);
    setFluxScheme(gamma, flux_scheme_choice);
    setResidualSmoother(residual_smoother_choice);
}
ResidualCalculator::~ResidualCalculator() {
    delete interpolation_;
    delete flux_scheme_;
    delete residual_smoother_;
}
void ResidualCalculator::setInterpolation(int choice, std::string g, std::string l, double omega, double k) {
    delete interpolation_;
    interpolation_ = choice == 2 ? (Interpolation*)new SecondOrder(g, l, omega, k) : (Interpolation*)new FirstOrder();
}
void ResidualCalculator::setFluxScheme(double gamma, std::string choice) {
    delete flux_scheme_;
    flux_scheme_ = nullptr;
    if (choice == "roe") flux_scheme_ = new RoeScheme(gamma);
    else if (choice == "ausm") flux_scheme_ = new AusmScheme(gamma);
    else throw Euler3dError(E3D_EINVAL, "unknown flux scheme '" + choice + "'");  // the reference leaves a null scheme
}
void ResidualCalculator::setResidualSmoother(bool choice) {  // Solver/src/ResidualCalculator.cpp:90-95
    delete residual_smoother_;
    residual_smoother_ = choice ? (ResidualSmoother*)new CentralIrsSmoother() : (ResidualSmoother*)new NoSmoother();
}
void CentralIrsSmoother::smoothResidual(Block* b) { e3d_check(e3d_smooth_residual(b->ctx_, b->block_id_), b->ctx_); }
// Solver/src/ResidualCalculator.cpp:24-41; `10^-16` is an integer XOR (-6), so the
// dissipation always runs, as in the reference.
void ResidualCalculator::computeResidual(Block* b) {
    interpolation_->computeInterpolation(b);
    flux_scheme_->computeFluxConv(b);
    flux_scheme_->current_beta_ = beta_rk_[current_stage_];
    if (beta_rk_[current_stage_] > (10 ^ -16)) flux_scheme_->computeFluxDiss(b);
    residual_smoother_->smoothResidual(b);
}

void Updater::updateInternalBlock(Block* b) { e3d_check(e3d_update(b->ctx_, b->block_id_, current_stage_), b->ctx_); }
void Updater::updateBoundary(Block* b) {  // Solver/src/Updater.cpp:80-88
    for (int i = 0; i < b->n_real_boundaries_in_block_; ++i) b->block_boundary_cell_ids_[i]->updateBoundary();
}

// Coefficient tables of Solver/src/RungeKutta.cpp:23-68.
RungeKutta::RungeKutta(double gamma, int stage_number, int interpolation_choice, std::string gradient_choice,
                       std::string limiter_choice, std::string flux_scheme_choice, bool residual_smoother_choice,
                       double omega, double k)
    : stage_number_(stage_number) {
    std::vector<double> a, bt;
    switch (stage_number_) {
        case 1: a = {1.0}; bt = {1.0}; break;
        case 2: a = {0.6, 1.0}; bt = {1.0, 0.0}; break;
        case 3: a = {0.6, 0.6, 1.0}; bt = {1.0, 0.0, 0.0}; break;
        case 4: a = {0.25, 1.0 / 3.0, 0.5, 1.0}; bt = {1.0, 0.5, 0.0, 0.0}; break;
        case 5: a = {0.2742, 0.2067, 0.5020, 0.5142, 1.0}; bt = {1.0, 0.0, 0.56, 0.0, 0.44}; break;
        default: throw Euler3dError(E3D_EINVAL, "Invalid number of RK stages");
    }
    updater_ = new Updater(gamma, a);
    residual_calculator_ = new ResidualCalculator(gamma, bt, interpolation_choice, gradient_choice, limiter_choice,
                                                  flux_scheme_choice, residual_smoother_choice, omega, k);
}
RungeKutta::~RungeKutta() {
    delete updater_;
    delete residual_calculator_;
}
void RungeKutta::computeRungeKutta(Block* b) {  // Solver/src/RungeKutta.cpp:80-94
    for (int s = 0; s < stage_number_; ++s) {
        residual_calculator_->current_stage_ = s;
        residual_calculator_->computeResidual(b);
        updater_->current_stage_ = s;
        updater_->updateInternalBlock(b);
        updater_->updateBoundary(b);
    }
}

void Timestep::computeSpectralRadius(Block* b) { e3d_check(e3d_spectral_radius(b->ctx_, b->block_id_), b->ctx_); }
void Timestep::computeTimestep(Block* b) { e3d_check(e3d_timestep(b->ctx_, b->block_id_), b->ctx_); }

// ----------------------------------------------------------- OutputTecplot

std::string OutputTecplot::blockFileName(const std::string& stem, int block_id) {
    char id[16];
    std::snprintf(id, sizeof id, "%04d", block_id);  // OUTPUT_ZERO_PADDING 4 (OutputTecplot.cpp:14)
    return stem + id + ".dat";
}

namespace {

// x, y and z of every node, one value per line (OutputTecplot.cpp:57-84).
void print_nodes(std::ostream& o, const e3d_block_view& v) {
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < v.nnode; ++i) o << v.node_xyz[3 * i + d] << std::endl;
}

void print_zone_header(std::ostream& o, const char* title, int nnode, int nelem, const char* zonetype) {
    o << "TTILE = \"" << title << "\"" << std::endl;
    o << "VARIABLES=\"X\",\"Y\",\"Z\",\"RO\",\"UU\",\"VV\",\"WW\",\"PP\",\"CP\",\"MACH\"" << std::endl;
    o << "ZONE T=\"FLOW_FIELD\"" << std::endl;
    o << "Nodes=" << nnode << ", " << "Elements=" << nelem << ", " << "ZONETYPE=" << zonetype << std::endl;
    o << "DATAPACKING=BLOCK" << std::endl;
    o << "VARLOCATION=([4,5,6,7,8,9,10]=CELLCENTERED)" << std::endl;
}

std::ofstream open_output(const std::string& path) {
    std::ofstream f(path);
    if (!f.is_open()) throw Euler3dError(E3D_EIO, "Failed opening file " + path);
    return f;
}

}  // namespace

// OutputTecplot::printFlowData (PostProcessing/src/OutputTecplot.cpp:32-189)
void OutputTecplot::printFlowData(Block* block) {
    const e3d_block_view& v = block->view_;
    const int nr = v.nreal, na = v.nall;
    const std::vector<double> W = block->get(E3D_ARR_W);
    const double *ro = &W[0], *uu = &W[(size_t)na], *vv = &W[2 * (size_t)na], *ww = &W[3 * (size_t)na],
                 *pp = &W[4 * (size_t)na];
    std::ofstream o = open_output(output_dir_ + blockFileName("FlowData_Block", block->block_id_));
    print_zone_header(o, "Vizualisation of the volumetric solution", v.nnode, nr, "FEBRICK");
    print_nodes(o, v);
    for (const double* a : {ro, uu, vv, ww, pp})
        for (int i = 0; i < nr; ++i) o << a[i] << std::endl;
    const double dyn_head = 0.5 * gamma_ * mach_aircraft_ * mach_aircraft_;
    for (int i = 0; i < nr; ++i) o << (pp[i] - 1.) / dyn_head << std::endl;
    for (int i = 0; i < nr; ++i) {
        const double a_cell = std::pow(gamma_ * pp[i] / ro[i], 0.5);
        const double velocity = std::pow(std::pow(uu[i], 2) + std::pow(vv[i], 2) + std::pow(ww[i], 2), 0.5);
        o << velocity / a_cell << std::endl;
    }
    for (int i = 0; i < nr; ++i) {
        for (int j = v.cell_node_off[i]; j < v.cell_node_off[i + 1]; ++j) o << v.cell_nodes[j] + 1 << " ";
        o << std::endl;
    }
}

// OutputTecplot::printSurfaceFlowData (PostProcessing/src/OutputTecplot.cpp:191-411), wall faces
// only, face values the mean of the interior cell and its ghost. The Mach number pairs the
// interior pressure with the ghost's density and velocity, as the reference does (:377-381).
void OutputTecplot::printSurfaceFlowData(Block* block) {
    const e3d_block_view& v = block->view_;
    if (v.n_wall_faces == 0) return;
    const int na = v.nall, nw = v.n_wall_faces;
    const std::vector<double> W = block->get(E3D_ARR_W);
    auto var = [&](int q, int cell) { return W[(size_t)q * na + cell]; };
    std::ofstream o = open_output(output_dir_ + blockFileName("SurfaceFlowData_Block", block->block_id_));
    print_zone_header(o, "Vizualisation of the surface solution", v.nnode, nw, "FEQUADRILATERAL");
    print_nodes(o, v);
    for (int q = 0; q < 5; ++q)
        for (int i = 0; i < nw; ++i) {
            const int f = v.wall_faces[i], c0 = v.face2cell[2 * f], c1 = v.face2cell[2 * f + 1];
            o << (var(q, c0) + var(q, c1)) * 0.5 << std::endl;
        }
    const double dyn_head = 0.5 * gamma_ * mach_aircraft_ * mach_aircraft_;
    for (int i = 0; i < nw; ++i) {
        const int f = v.wall_faces[i], c0 = v.face2cell[2 * f], c1 = v.face2cell[2 * f + 1];
        o << ((var(4, c0) + var(4, c1)) * 0.5 - 1.) / dyn_head << std::endl;
    }
    for (int i = 0; i < nw; ++i) {
        const int f = v.wall_faces[i], c0 = v.face2cell[2 * f], c1 = v.face2cell[2 * f + 1];
        const double p0 = var(4, c0);
        const double pp = (p0 + var(4, c1)) * 0.5, ro = (p0 + var(0, c1)) * 0.5, uu = (p0 + var(1, c1)) * 0.5,
                     vv = (p0 + var(2, c1)) * 0.5, ww = (p0 + var(3, c1)) * 0.5;
        const double a_face = std::pow(gamma_ * pp / ro, 0.5);
        const double velocity = std::pow(std::pow(uu, 2) + std::pow(vv, 2) + std::pow(ww, 2), 0.5);
        o << velocity / a_face << std::endl;
    }
    for (int i = 0; i < nw; ++i) {
        const int f = v.wall_faces[i], nn = v.face_nnode[f];
        for (int j = 0; j < nn; ++j) {
            o << v.face_nodes[4 * f + j] + 1;
            if (j == nn - 1) o << std::endl;
            else o << " ";
        }
    }
}

// OutputTecplot::printConvergence (PostProcessing/src/OutputTecplot.cpp:413-446)
void OutputTecplot::printConvergence(int iter, double cl, double cd, double cmx, double cmy, double cmz,
                                     double ro_convergence, double uu_convergence, double vv_convergence,
                                     double ww_convergence, double pp_convergence) {
    std::ofstream o(output_dir_ + "Convergence.dat", iter == 0 ? std::ios::trunc : std::ios::app);
    if (!o.is_open()) throw Euler3dError(E3D_EIO, "Failed opening file " + output_dir_ + "Convergence.dat");
    if (iter == 0)
        o << "Iteration" << " " << "Cl" << " " << "Cd" << " " << "Cmx" << " " << "Cmy" << " " << "Cmz" << " "
          << "Density_Convergence" << " " << "Uu_Convergence" << " " << "Vv_Convergence" << " " << "Ww_Convergence"
          << " " << "Pressure_Convergence" << " " << std::endl;
    o << iter << " " << cl << " " << cd << " " << cmx << " " << cmy << " " << cmz << " " << ro_convergence << " "
      << uu_convergence << " " << vv_convergence << " " << ww_convergence << " " << pp_convergence << std::endl;
}

// OutputTecplot::printAerodynamicCoefficients (PostProcessing/src/OutputTecplot.cpp:448-473)
void OutputTecplot::printAerodynamicCoefficients(double cl, double cd, double cmx, double cmy, double cmz) {
    std::ofstream o = open_output(output_dir_ + "AerodynamicCoefficients.dat");
    o << "Angle of attack" << " " << "Cl" << " " << "Cd" << " " << "Cmx" << " " << "Cmy" << " " << "Cmz" << std::endl;
    o << aoa_deg_ << " " << cl << " " << cd << " " << cmx << " " << cmy << " " << cmz << std::endl;
}

namespace {
const char kRestartMagic[8] = {'E', '3', 'D', 'R', 'S', 'T', '0', '1'};
}

void OutputTecplot::printRestartFile(Block* block) {
    const e3d_block_view& v = block->view_;
    int32_t head[4] = {block->block_id_, v.nreal, v.nall, 0};
    double rms0[5];
    e3d_check(e3d_monitor_state(block->ctx_, &head[3], rms0), block->ctx_);
    const std::vector<double> W = block->get(E3D_ARR_W), D = block->get(E3D_ARR_DISS);
    std::ofstream o(output_dir_ + blockFileName("RestartFile", block->block_id_), std::ios::binary);
    if (!o.is_open()) throw Euler3dError(E3D_EIO, "Failed opening restart file of block " + std::to_string(head[0]));
    o.write(kRestartMagic, 8);
    o.write((const char*)head, sizeof head);
    o.write((const char*)rms0, sizeof rms0);
    o.write((const char*)W.data(), sizeof(double) * W.size());
    o.write((const char*)D.data(), sizeof(double) * D.size());
    if (!o) throw Euler3dError(E3D_EIO, "short write of restart file of block " + std::to_string(head[0]));
}

void readRestartFiles(const std::string& directory, CompleteMesh* mesh) {
    const std::string dir = directory.empty() || directory.back() == '/' ? directory : directory + "/";
    for (int b = 0; b < mesh->n_blocks_; ++b) {
        Block* B = mesh->all_blocks_[b];
        const std::string path = dir + OutputTecplot::blockFileName("RestartFile", b);
        std::ifstream in(path, std::ios::binary);
        if (!in.is_open()) throw Euler3dError(E3D_EIO, "cannot open restart file " + path);
        char magic[8];
        int32_t head[4];
        double rms0[5];
        in.read(magic, 8);
        in.read((char*)head, sizeof head);
        in.read((char*)rms0, sizeof rms0);
        if (!in || std::memcmp(magic, kRestartMagic, 8) != 0)
            throw Euler3dError(E3D_EIO, path + " is not a restart file");
        if (head[0] != b || head[1] != B->view_.nreal || head[2] != B->view_.nall)
            throw Euler3dError(E3D_EINVAL, path + " does not match block " + std::to_string(b) + " of the mesh");
        std::vector<double> W(5 * (size_t)head[2]), D(5 * (size_t)head[1]);
        in.read((char*)W.data(), sizeof(double) * W.size());
        in.read((char*)D.data(), sizeof(double) * D.size());
        if (!in) throw Euler3dError(E3D_EIO, "truncated restart file " + path);
        B->set(E3D_ARR_W, W);
        B->set(E3D_ARR_DISS, D);
        if (b == 0) e3d_check(e3d_set_monitor_state(mesh->ctx_, head[3], rms0), mesh->ctx_);
    }
}

// ---------------------------------------------------------- PostProcessing

PostProcessing::PostProcessing(int, int max_iter, double convergence_criterion, double, double mach_aircraft,
                               double aoa_deg, double gamma)
    : max_iter_(max_iter), convergence_criterion_(convergence_criterion), aoa_deg_(aoa_deg) {
    output_tecplot_ = new OutputTecplot(mach_aircraft, aoa_deg, gamma);
}

PostProcessing::~PostProcessing() { delete output_tecplot_; }

void PostProcessing::checkStopSolver() {  // PostProcessing/src/PostProcessing.cpp:59-83
    struct stat buffer;
    const bool file_exist_flag = (stat(stop_file_name_.c_str(), &buffer) == 0);
    if ((current_iter_ + 1 == max_iter_) || (ro_convergence_ <= convergence_criterion_) || file_exist_flag)
        stop_solver_ = true;
}

void PostProcessing::process(CompleteMesh* complete_mesh, BlockCommunicator* communicator) {
    int rc = e3d_monitors(complete_mesh->ctx_, &last_);
    if (rc < 0 && rc != E3D_ENONFINITE) e3d_check(rc, complete_mesh->ctx_);
    (void)communicator;
    if (rc == E3D_ENONFINITE) stop_solver_ = true;
    record(last_, complete_mesh);
}

// PostProcessing::process (PostProcessing/src/PostProcessing.cpp:222-308) for one monitor row.
void PostProcessing::record(const e3d_monitor& m, CompleteMesh* complete_mesh) {
    last_ = m;
    current_iter_ = m.iter;
    ro_rms_mesh_ = m.rms[0]; uu_rms_mesh_ = m.rms[1]; vv_rms_mesh_ = m.rms[2]; ww_rms_mesh_ = m.rms[3]; pp_rms_mesh_ = m.rms[4];
    ro_convergence_ = m.drop[0]; uu_convergence_ = m.drop[1]; vv_convergence_ = m.drop[2];
    ww_convergence_ = m.drop[3]; pp_convergence_ = m.drop[4];
    cl_geometry_mesh_ = m.cl; cd_geometry_mesh_ = m.cd;
    cmx_geometry_mesh_ = m.cmx; cmy_geometry_mesh_ = m.cmy; cmz_geometry_mesh_ = m.cmz;
    checkStopSolver();
    if (stop_solver_) {
        if (verbose_) std::cout << "STOP ACTIVATED" << std::endl;
        std::remove(stop_file_name_.c_str());
    }
    if (verbose_)
        std::cout << "Iter: " << current_iter_ << " Convergence Ro: " << ro_convergence_ << " cl: " << cl_geometry_mesh_
                  << " cd: " << cd_geometry_mesh_ << std::endl;
    mkdir(output_dir_.c_str(), 0755);
    output_tecplot_->output_dir_ = output_dir_;
    output_tecplot_->printConvergence(current_iter_, cl_geometry_mesh_, cd_geometry_mesh_, cmx_geometry_mesh_,
                                      cmy_geometry_mesh_, cmz_geometry_mesh_, ro_convergence_, uu_convergence_,
                                      vv_convergence_, ww_convergence_, pp_convergence_);
    if (stop_solver_) {
        if (write_solution_ && complete_mesh) {
            if (verbose_) std::cout << "Writing Solution......................................................" << std::endl;
            for (int i = 0; i < complete_mesh->n_blocks_in_process_; ++i) {
                Block* b = complete_mesh->all_blocks_[complete_mesh->my_blocks_[i]];
                output_tecplot_->printFlowData(b);
                output_tecplot_->printSurfaceFlowData(b);
                output_tecplot_->printRestartFile(b);
            }
        }
        output_tecplot_->printAerodynamicCoefficients(cl_geometry_mesh_, cd_geometry_mesh_, cmx_geometry_mesh_,
                                                      cmy_geometry_mesh_, cmz_geometry_mesh_);
        if (verbose_) std::cout << "========================END OF PROGRAM========================" << std::endl;
    }
    ++current_iter_;
}

// ------------------------------------------------------------------ Solver

Solver::Solver(double gamma, double cfl, int stage_number, int interpolation_choice, std::string gradient_choice,
               std::string limiter_choice, std::string flux_scheme_choice, bool residual_smoother_choice, int n_blocks,
               int max_iter, double convergence_criterion, double cmac, double aoa_deg, double mach_aircraft,
               double omega, double k)
    : gamma_(gamma) {
    timestep_ = new Timestep(gamma, cfl);
    runge_kutta_ = new RungeKutta(gamma, stage_number, interpolation_choice, gradient_choice, limiter_choice,
                                  flux_scheme_choice, residual_smoother_choice, omega, k);
    post_processing_ =
        new PostProcessing(n_blocks, max_iter, convergence_criterion, cmac, mach_aircraft, aoa_deg, gamma);
}

Solver::~Solver() {
    delete timestep_;
    delete runge_kutta_;
    delete post_processing_;
}

void Solver::saveW0(Block* b) { e3d_check(e3d_save_w0(b->ctx_, b->block_id_), b->ctx_); }

// Solver/src/Solver.cpp:36-98
void Solver::solve(CompleteMesh* complete_mesh, BlockCommunicator* communicator) {
    communicator->updateBoundaries(complete_mesh);
    PostProcessing* pp = post_processing_;
    if (!granular_) {
        while (!pp->stop_solver_) {
            e3d_monitor m;
            const int n = e3d_iterate(complete_mesh->ctx_, 1, &m);
            if (n == E3D_ENONFINITE) {
                pp->record(m, complete_mesh);
                pp->stop_solver_ = true;
                break;
            }
            e3d_check(n, complete_mesh->ctx_);
            if (n == 0) break;
            pp->record(m, complete_mesh);
        }
        return;
    }
    while (!pp->stop_solver_) {
        for (int i = 0; i < complete_mesh->n_blocks_in_process_; ++i) {
            Block* b = complete_mesh->all_blocks_[complete_mesh->my_blocks_[i]];
            saveW0(b);
            timestep_->computeSpectralRadius(b);
            timestep_->computeTimestep(b);
            runge_kutta_->computeRungeKutta(b);
            pp->computeFlowData(b);
        }
        communicator->updateBoundaries(complete_mesh);
        pp->process(complete_mesh, communicator);
    }
}

Solver* Initializer::initializeSolver(Interface* in) {  // DataStructure/src/Initializer.cpp:25-50
    return new Solver(in->gamma_interface_, in->cfl_interface_, in->stage_number_interface_,
                      in->interpolation_choice_interface_, in->gradient_choice_interface_,
                      in->limiter_choice_interface_, in->flux_scheme_choice_interface_,
                      in->residual_smoother_choice_interface_, in->n_blocks_interface_, in->max_iter_interface_,
                      in->convergence_criterion_interface_, in->cmac_interface_, in->aoa_deg_interface_,
                      in->mach_aircraft_interface_, in->omega_interface_, in->k_interface_);
}
