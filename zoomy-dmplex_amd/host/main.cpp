// zdm_cli_<model>: the reference's solver executable (main.cpp of solver_cpu) on the MI355X
// library: reads -settings (default settings.json), picks ChorinVAMSolver for
// solver.method == "chorin" and MUSCLSolver otherwise, and runs to t_end.
//   -o <file.npy>      final state, float64 (ncells, n_q), mesh-file cell order
//   -no_output         no VTK snapshots (otherwise io.snapshots of them, as the reference)
//   -max_steps <n>     stop after n steps
//   -print-settings    print the parsed settings as JSON and exit
//   -ksp_* / -pc_type  pressure-solve options (Chorin), as the reference reads them
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>

#include "zoomy/solver.hpp"

static void print_settings(const zoomy::Settings& s) {
    auto q = [](const std::string& v) {
        std::string o = "\"";
        for (const char c : v) {
            if (c == '"' || c == '\\') {
                o += '\\';
                o += c;
            } else if (c == '\n') {
                o += "\\n";
            } else {
                o += c;
            }
        }
        return o + "\"";
    };
    std::cout << "{\"name\": " << q(s.name) << ", \"io\": {\"mesh_path\": " << q(s.io.mesh_path)
              << ", \"initial_condition_file\": " << q(s.io.initial_condition_file)
              << ", \"initial_condition_mask\": [";
    for (size_t i = 0; i < s.io.initial_condition_mask.size(); ++i)
        std::cout << (i ? ", " : "") << s.io.initial_condition_mask[i];
    std::cout << "], \"snapshots\": " << s.io.snapshots << ", \"directory\": " << q(s.io.directory)
              << "}, \"solver\": {\"t_end\": " << s.solver.t_end << ", \"cfl\": " << s.solver.cfl
              << ", \"reconstruction_order\": " << s.solver.reconstruction_order
              << ", \"min_dt\": " << s.solver.min_dt << ", \"limiter\": " << q(s.solver.limiter)
              << ", \"time_integration\": " << q(s.solver.time_integration)
              << ", \"method\": " << q(s.solver.method) << ", \"positivity\": " << q(s.solver.positivity)
              << ", \"ssp_stages\": " << s.solver.ssp_stages << ", \"gradient\": " << q(s.solver.gradient)
              << ", \"well_balanced\": " << (s.solver.well_balanced ? "true" : "false")
              << "}, \"parameters\": {";
    bool first = true;
    for (const auto& kv : s.model.parameters) {
        std::cout << (first ? "" : ", ") << q(kv.first) << ": " << kv.second;
        first = false;
    }
    std::cout << "}}" << std::endl;
}

int main(int argc, char** argv) {
    std::string spath = "settings.json", out;
    int max_steps = 1 << 30;
    bool print = false, snapshots = true, ic_only = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-settings" && i + 1 < argc) spath = argv[i + 1];
        if (a == "-o" && i + 1 < argc) out = argv[i + 1];
        if (a == "-max_steps" && i + 1 < argc) max_steps = std::stoi(argv[i + 1]);
        if (a == "-print-settings") print = true;
        if (a == "-no_output") snapshots = false;
        if (a == "-ic_only") ic_only = true;   // load the mesh and initial condition, write it (-o), stop
    }
    zoomy::Settings st;
    try {
        st = zoomy::Settings::from_json(spath);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    if (print) {
        print_settings(st);
        return 0;
    }
    std::unique_ptr<zoomy::MUSCLSolver> solver;
    if (st.solver.method == "chorin")
        solver = std::make_unique<zoomy::ChorinVAMSolver>();
    else
        solver = std::make_unique<zoomy::MUSCLSolver>();
    solver->max_steps = max_steps;
    solver->write_snapshots = snapshots;
    if (ic_only) {
        // the host setup only (settings, mesh, initial condition): no device work
        std::vector<double> Xo;
        if (solver->Initialize(argc, argv)) {
            std::cerr << "Error: " << zoomy::LastError() << std::endl;
            return 2;
        }
        solver->HostStateInInputOrder(Xo);
        if (!out.empty() && zoomy::WriteNpy(out, Xo, (size_t)solver->mesh.ncells(), (size_t)solver->info.n_q)) {
            std::cerr << "Error: " << zoomy::LastError() << std::endl;
            return 3;
        }
        return 0;
    }
    if (solver->Run(argc, argv)) {
        std::cerr << "Error: " << zoomy::LastError() << std::endl;
        return 2;
    }
    if (!out.empty()) {
        std::vector<double> Xo;
        if (solver->StateInInputOrder(Xo) ||
            zoomy::WriteNpy(out, Xo, (size_t)solver->mesh.ncells(), (size_t)solver->info.n_q)) {
            std::cerr << "Error: " << zoomy::LastError() << std::endl;
            return 3;
        }
    }
    return 0;
}
