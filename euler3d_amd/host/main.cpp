// euler3d command line: the program flow of the reference's main.cpp:17-101
// (input file → Interface → blocks → CompleteMesh → Solver::solve → elapsed
// time) over the MI355X device path. Usage:
//   euler3d <input_file> [--granular] [--quiet] [--device N] [--out DIR]
//           [--restart DIR] [--no-solution] [--order reference|rcm|morton]
// --granular runs the operator-by-operator loop of Solver/src/Solver.cpp:76-98
// instead of the stage-fused kernels (same results). --order renumbers cells
// and faces in device memory (reverse Cuthill-McKee or a Morton curve) for
// meshes whose file order has little locality (same results). --restart resumes from the
// RestartFile<block>.dat files a previous run wrote into DIR when it stopped
// (state, iteration counter and reference residual); --no-solution skips the
// FlowData / SurfaceFlowData / RestartFile files at the stop. Exit status 42 on
// an I/O error (the reference's convention), 1 on any other library error.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "euler3d_solver.h"

int main(int argc, char* argv[]) {
    const char* input = nullptr;
    bool granular = false, quiet = false, solution = true;
    int device = 0, memory_order = E3D_ORDER_REFERENCE;
    std::string out_dir = "./bin/", restart_dir;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--granular")) granular = true;
        else if (!std::strcmp(argv[i], "--quiet")) quiet = true;
        else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) device = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) out_dir = std::string(argv[++i]) + "/";
        else if (!std::strcmp(argv[i], "--restart") && i + 1 < argc) restart_dir = argv[++i];
        else if (!std::strcmp(argv[i], "--no-solution")) solution = false;
        else if (!std::strcmp(argv[i], "--order") && i + 1 < argc) {
            const char* o = argv[++i];
            if (!std::strcmp(o, "reference")) memory_order = E3D_ORDER_REFERENCE;
            else if (!std::strcmp(o, "rcm")) memory_order = E3D_ORDER_RCM;
            else if (!std::strcmp(o, "morton")) memory_order = E3D_ORDER_MORTON;
            else {
                std::cerr << "unknown memory order " << o << " (reference, rcm or morton)" << std::endl;
                return 2;
            }
        }
        else if (!input) input = argv[i];
        else {
            std::cerr << "unexpected argument " << argv[i] << std::endl;
            return 2;
        }
    }
    std::cout << "========================STARTING PROGRAM========================" << std::endl;
    if (!input) {
        std::cout << "ERROR: No input file" << std::endl;
        return -1;  // exit(-1), as main.cpp:57
    }
    std::cout << "Input file is : " << input << std::endl;
    std::cout << e3d_version() << std::endl;

    try {
        Interface interface(input, true);
        Initializer initializer;
        const int n_blocks = interface.n_blocks_interface_;
        int n_blocks_in_process;
        int* my_blocks;
        BlockCommunicator communicator(n_blocks);
        communicator.getMyBlocks(n_blocks_in_process, my_blocks);
        std::cout << "I am process " << communicator.process_id_ << " and I have " << n_blocks_in_process
                  << " blocks." << std::endl;
        CompleteMesh complete_mesh(n_blocks, n_blocks_in_process, my_blocks, interface.topology_file_name_interface_,
                                   device);
        delete[] my_blocks;
        complete_mesh.memory_order_ = memory_order;
        complete_mesh.InitializeMyBlocks(&interface, &communicator);
        if (!restart_dir.empty()) readRestartFiles(restart_dir, &complete_mesh);
        communicator.initializeBuffers();
        std::cout << "Initialisation du Solver" << std::endl << std::endl;
        Solver* solver = initializer.initializeSolver(&interface);
        solver->granular_ = granular;
        solver->post_processing_->verbose_ = !quiet;
        solver->post_processing_->output_dir_ = out_dir;
        solver->post_processing_->write_solution_ = solution;

        const auto t_start = std::chrono::high_resolution_clock::now();
        solver->solve(&complete_mesh, &communicator);
        const auto t_end = std::chrono::high_resolution_clock::now();
        const PostProcessing* pp = solver->post_processing_;
        std::cout << "Iterations: " << pp->current_iter_ << "  Cl: " << pp->cl_geometry_mesh_
                  << "  Cd: " << pp->cd_geometry_mesh_ << "  density drop: " << pp->ro_convergence_ << std::endl;
        std::cout << "Time elapsed solving: "
                  << std::chrono::duration<double, std::milli>(t_end - t_start).count() / 1000.0 << "s." << std::endl;
        delete solver;
    } catch (const Euler3dError& e) {
        std::cerr << "ERROR: " << e.what() << std::endl;
        return e.code_ == E3D_EIO ? 42 : 1;
    }
    return 0;
}
