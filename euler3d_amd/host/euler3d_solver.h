// C++ host classes with the reference's names, members and virtual methods,
// driving the MI355X kernels through the C ABI (include/euler3d.h). A program
// written against the reference's Solver layer (Solver/head/*.h,
// DataStructure/head/{Block,CompleteMesh,BoundaryCellIds}.h, MPI/head/
// BlockCommunicator.h, PostProcessing/head/PostProcessing.h) keeps its shape:
// the same factories pick the same scheme/gradient/limiter objects, and each
// object's compute method runs one device operator on the block it is given.
//
// Differences to the reference, by design:
//   * Block is a handle on a block resident in device memory (no Cell**/Face**
//     arrays on the host); fields are read with Block::get().
//   * errors throw Euler3dError instead of exit(42) or a null dereference.
//   * Solver::solve runs the stage-fused device path (e3d_iterate); set
//     Solver::granular_ = true to run the operator-by-operator loop of
//     Solver/src/Solver.cpp:76-98 instead (same results).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/euler3d.h"

class Euler3dError : public std::runtime_error {
   public:
    Euler3dError(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code_;
};

// Throws Euler3dError when rc < 0, with the library's message.
int e3d_check(int rc, e3d_ctx* ctx = nullptr);

class BoundaryCellIds;

// DataStructure/head/Block.h:17-60, as a device-resident block handle.
class Block {
   public:
    e3d_ctx* ctx_ = nullptr;
    int block_id_ = 0;
    int n_real_cells_in_block_ = 0;
    int n_ghost_cells_in_block_ = 0;
    int n_all_cells_in_block_ = 0;
    int n_nodes_in_block_ = 0;
    int n_faces_in_block_ = 0;
    int n_real_boundaries_in_block_ = 0;
    int n_all_boundaries_in_block_ = 0;
    int n_wall_faces_ = 0;
    BoundaryCellIds** block_boundary_cell_ids_ = nullptr;

    e3d_block_view view_{};             // host mesh arrays of the block (nodes, connectivity, faces)

    explicit Block(int block_id) : block_id_(block_id) {}
    ~Block();
    // Host copy of one array (E3D_ARR_*), reference layout (variable-major).
    std::vector<double> get(int which) const;
    void set(int which, const std::vector<double>& values);
};

// DataStructure/head/BoundaryCellIds.h:5-18 and its subclasses.
class BoundaryCellIds {
   public:
    int n_cell_in_boundary_ = 0;
    int marker_ = 0;
    Block* owner_block_ = nullptr;
    virtual void updateBoundary() = 0;
    virtual ~BoundaryCellIds() {}
};
class WallCellIds : public BoundaryCellIds {
   public:
    void updateBoundary() override;
};
class SymmetryCellIds : public BoundaryCellIds {
   public:
    void updateBoundary() override;
};
class FarfieldCellIds : public BoundaryCellIds {
   public:
    void updateBoundary() override;
};
class ConnexionBlockCellIds : public BoundaryCellIds {  // DataStructure/src/ConnexionBlockCellIds.cpp:10-13: no-op
   public:
    void updateBoundary() override {}
};

// Solver/head/Interface.h: the positional input file.
class Interface {
   public:
    std::string topology_file_name_interface_;
    double cfl_interface_, gamma_interface_, aoa_deg_interface_;
    int stage_number_interface_;
    double mach_aircraft_interface_, cmac_interface_;
    int max_iter_interface_;
    double convergence_criterion_interface_;
    bool residual_smoother_choice_interface_;
    std::string flux_scheme_choice_interface_;
    int interpolation_choice_interface_;
    std::string gradient_choice_interface_, limiter_choice_interface_;
    double omega_interface_, k_interface_;
    int n_blocks_interface_ = 0;
    e3d_params params_{};
    Interface(const char* input_file, bool i_write);
};

class BlockCommunicator;

// DataStructure/head/CompleteMesh.h: all blocks of the mesh, on one device.
class CompleteMesh {
   public:
    int n_blocks_ = 0;
    int n_blocks_in_process_ = 0;
    int* my_blocks_ = nullptr;
    Block** all_blocks_ = nullptr;
    std::string topology_file_name_;
    e3d_ctx* ctx_ = nullptr;
    e3d_mesh* mesh_ = nullptr;
    int memory_order_ = E3D_ORDER_REFERENCE;  // device memory order of cells and faces (e3d_set_ordering)
    CompleteMesh(int n_blocks, int n_blocks_in_process, int* my_blocks, std::string topology_file_name, int device = 0);
    ~CompleteMesh();
    // Reads, builds and uploads every block, then freestream everywhere
    // (DataStructure/src/CompleteMesh.cpp:79-293).
    void InitializeMyBlocks(Interface* interface, BlockCommunicator* communicator);
};

// MPI/head/BlockCommunicator.h: on one device the interface exchange is a
// device-side copy and the global sums are device reductions.
class PostProcessing;
class BlockCommunicator {
   public:
    int number_of_processes_ = 1;
    int process_id_ = 0;
    int n_blocks_ = 0;
    explicit BlockCommunicator(int nblocks) : n_blocks_(nblocks) {}
    void updateBoundaries(CompleteMesh* mesh) const;
    void getMyBlocks(int& n_blocks_in_process, int*& my_blocks) const;
    void getGlobal(CompleteMesh* mesh, PostProcessing* postprocess);
    void initializeBuffers() {}  // interface buffers live on the device, set up by e3d_upload_mesh
    void sync() {}
};

// Solver/head/FluxScheme.h and subclasses.
class FluxScheme {
   public:
    double gamma_ = 1.4;
    double current_beta_ = 1.0;
    int scheme_ = E3D_FLUX_ROE;
    virtual ~FluxScheme() {}
    virtual void computeFluxDiss(Block* block) = 0;
    virtual void computeFluxConv(Block* block);  // central flux (FluxScheme.cpp:16-211)
};
class RoeScheme : public FluxScheme {
   public:
    explicit RoeScheme(double gamma);
    void computeFluxDiss(Block* block) override;
};
class AusmScheme : public FluxScheme {
   public:
    explicit AusmScheme(double gamma);
    void computeFluxConv(Block* block) override;
    void computeFluxDiss(Block* block) override;
};

class Gradient {
   public:
    virtual ~Gradient() {}
    virtual void computeGradients(Block* block) = 0;
};
class GreenGauss : public Gradient {
   public:
    void computeGradients(Block* block) override;
};
class LeastSquares : public Gradient {
   public:
    void computeGradients(Block* block) override;
};

class Limiter {
   public:
    virtual ~Limiter() {}
    virtual void computeLimiters(Block* block) = 0;
};
class BarthJespersen : public Limiter {
   public:
    double omega_;
    explicit BarthJespersen(double omega) : omega_(omega) {}
    void computeLimiters(Block* block) override;
};
class Venkatakrishnan : public Limiter {
   public:
    double k_;
    explicit Venkatakrishnan(double k) : k_(k) {}
    void computeLimiters(Block* block) override;
};

class Interpolation {
   public:
    virtual ~Interpolation() {}
    virtual void computeInterpolation(Block* block) = 0;
};
class FirstOrder : public Interpolation {  // Solver/src/FirstOrder.cpp:12-15: no-op
   public:
    void computeInterpolation(Block*) override {}
};
class SecondOrder : public Interpolation {
   public:
    Gradient* gradient_ = nullptr;
    Limiter* limiter_ = nullptr;
    SecondOrder(std::string gradient_choice, std::string limiter_choice, double omega, double k);
    ~SecondOrder();
    void setGradient(std::string gradient_choice);
    void setLimiter(std::string limiter_choice, double omega, double k);
    void computeInterpolation(Block* block) override;
};

class ResidualSmoother {
   public:
    virtual ~ResidualSmoother() {}
    virtual void smoothResidual(Block* block) = 0;
};
class NoSmoother : public ResidualSmoother {
   public:
    void smoothResidual(Block*) override {}
};
class CentralIrsSmoother : public ResidualSmoother {  // Solver/src/CentralIrsSmoother.cpp:12-183
   public:
    void smoothResidual(Block* block) override;
};

// Solver/head/ResidualCalculator.h
class ResidualCalculator {
   public:
    Interpolation* interpolation_ = nullptr;
    FluxScheme* flux_scheme_ = nullptr;
    ResidualSmoother* residual_smoother_ = nullptr;
    int current_stage_ = 0;
    std::vector<double> beta_rk_;
    ResidualCalculator(double gamma, std::vector<double> beta_rk, int interpolation_choice, std::string gradient_choice,
                       std::string limiter_choice, std::string flux_scheme_choice, bool residual_smoother_choice,
                       double omega, double k);
    ~ResidualCalculator();
    void setInterpolation(int interpolation_choice, std::string gradient_choice, std::string limiter_choice,
                          double omega, double k);
    void setFluxScheme(double gamma, std::string flux_scheme_choice);
    void setResidualSmoother(bool residual_smoother_choice);
    void computeResidual(Block* block);
};

class Updater {
   public:
    double gamma_;
    int current_stage_ = 0;
    std::vector<double> alpha_rk_;
    Updater(double gamma, std::vector<double> alpha_rk) : gamma_(gamma), alpha_rk_(alpha_rk) {}
    void updateInternalBlock(Block* block);
    void updateBoundary(Block* block);
};

class RungeKutta {
   public:
    int stage_number_;
    ResidualCalculator* residual_calculator_ = nullptr;
    Updater* updater_ = nullptr;
    RungeKutta(double gamma, int stage_number, int interpolation_choice, std::string gradient_choice,
               std::string limiter_choice, std::string flux_scheme_choice, bool residual_smoother_choice,
               double omega, double k);
    ~RungeKutta();
    void computeRungeKutta(Block* block);
};

class Timestep {
   public:
    double gamma_, cfl_;
    Timestep(double gamma, double cfl) : gamma_(gamma), cfl_(cfl) {}
    void computeSpectralRadius(Block* block);
    void computeTimestep(Block* block);
};

// PostProcessing/head/OutputTecplot.h: the solution files written when the
// solver stops, in the reference's text format (default-formatted streams),
// and a restart file that holds the state (the reference's is empty).
class OutputTecplot {
   public:
    double mach_aircraft_, aoa_deg_, gamma_;
    std::string output_dir_ = "./bin/";
    OutputTecplot(double mach_aircraft, double aoa_deg, double gamma)
        : mach_aircraft_(mach_aircraft), aoa_deg_(aoa_deg), gamma_(gamma) {}
    void printFlowData(Block* block);
    void printSurfaceFlowData(Block* block);
    void printConvergence(int iter, double cl, double cd, double cmx, double cmy, double cmz, double ro_convergence,
                          double uu_convergence, double vv_convergence, double ww_convergence, double pp_convergence);
    void printAerodynamicCoefficients(double cl, double cd, double cmx, double cmy, double cmz);
    // Binary: "E3DRST01", int32 block, nreal, nall, iteration; double rms0[5];
    // W [5][nall] (primitive, all cells); diss [5][nreal].
    void printRestartFile(Block* block);
    static std::string blockFileName(const std::string& stem, int block_id);
};

// Restores the state, iteration counter and reference residual written by
// OutputTecplot::printRestartFile for every block of the mesh.
void readRestartFiles(const std::string& directory, CompleteMesh* complete_mesh);

// PostProcessing/head/PostProcessing.h: monitors, stop test and output files.
class PostProcessing {
   public:
    int current_iter_ = 0;
    int max_iter_;
    double convergence_criterion_;
    bool stop_solver_ = false;
    std::string stop_file_name_ = "STOP";
    std::string output_dir_ = "./bin/";
    double ro_rms_mesh_ = 0, uu_rms_mesh_ = 0, vv_rms_mesh_ = 0, ww_rms_mesh_ = 0, pp_rms_mesh_ = 0;
    double ro_convergence_ = 0, uu_convergence_ = 0, vv_convergence_ = 0, ww_convergence_ = 0, pp_convergence_ = 0;
    double cl_geometry_mesh_ = 0, cd_geometry_mesh_ = 0, cmx_geometry_mesh_ = 0, cmy_geometry_mesh_ = 0,
           cmz_geometry_mesh_ = 0;
    double aoa_deg_;
    bool verbose_ = true;
    bool write_solution_ = true;      // FlowData / SurfaceFlowData / RestartFile at the stop
    e3d_monitor last_{};
    OutputTecplot* output_tecplot_ = nullptr;
    PostProcessing(int n_blocks, int max_iter, double convergence_criterion, double cmac, double mach_aircraft,
                   double aoa_deg, double gamma);
    ~PostProcessing();
    void computeFlowData(Block*) {}  // per-block norms are computed on the device by process()
    void process(CompleteMesh* complete_mesh, BlockCommunicator* communicator);
    // One monitor row produced elsewhere (the fused path): stop test, log, files.
    void record(const e3d_monitor& m, CompleteMesh* complete_mesh);
    void checkStopSolver();
};

// Solver/head/Solver.h
class Solver {
   public:
    double gamma_;
    Timestep* timestep_ = nullptr;
    RungeKutta* runge_kutta_ = nullptr;
    PostProcessing* post_processing_ = nullptr;
    bool granular_ = false;
    Solver(double gamma, double cfl, int stage_number, int interpolation_choice, std::string gradient_choice,
           std::string limiter_choice, std::string flux_scheme_choice, bool residual_smoother_choice, int n_blocks,
           int max_iter, double convergence_criterion, double cmac, double aoa_deg, double mach_aircraft,
           double omega, double k);
    ~Solver();
    void solve(CompleteMesh* complete_mesh, BlockCommunicator* communicator);
    void saveW0(Block* block);
};

// DataStructure/head/Initializer.h
class Initializer {
   public:
    Solver* initializeSolver(Interface* interface);
};
