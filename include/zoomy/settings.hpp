// Settings of a run: the JSON keys and defaults of the reference's Settings.hpp:13-138
// (SolverSettings, IOSettings, ModelSettings, Settings::from_json), so its settings files load
// unchanged.  Two extension keys: solver.gradient ("green_gauss" | "lsq") and
// solver.well_balanced (bool); unknown keys are ignored, as the reference ignores them.
#pragma once
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "zoomy/json.hpp"

namespace zoomy {

struct SolverSettings {
    double t_end = 1.0;
    double cfl = 0.5;
    bool use_deep_adjacency = true;
    int reconstruction_order = 1;
    double min_dt = 1.0e-12;
    std::string limiter = "venkatakrishnan";   // any value but "none" = Vacuum + TVD
    bool refresh_derivative_aux = true;
    std::string positivity = "none";
    double wet_dry_eps = 1.0e-2;
    std::string time_integration = "splitting";
    std::string method = "muscl";               // "muscl" | "chorin"
    int ssp_stages = 5;
    // extension keys
    std::string gradient = "green_gauss";
    bool well_balanced = false;
};

struct IOSettings {
    std::string directory = "output";
    std::string filename = "sol";
    int snapshots = 10;
    std::string snapshot_logic = "snap";
    bool clean_directory = false;
    std::string mesh_path = "";
    std::string mesh_label = "Face Sets";
    bool restart = false;
    std::string restart_file = "";
    std::string initial_condition_file = "";
    std::vector<int> initial_condition_mask;
    bool write_3d = false;
    int n_layers_3d = 10;
    std::string output_3d_name = "sol_3d";
};

struct ModelSettings {
    std::map<std::string, double> parameters;
};

struct Settings {
    std::string name = "Simulation";
    IOSettings io;
    SolverSettings solver;
    ModelSettings model;

    static Settings from_text(const std::string& text) {
        const JsonValue j = parse_json(text);
        Settings s;
        s.name = j.value("name", "Simulation");
        if (j.contains("io")) {
            const JsonValue& jio = j["io"];
            s.io.directory = jio.value("directory", "output");
            s.io.filename = jio.value("filename", "sol");
            s.io.snapshots = jio.value("snapshots", 10);
            s.io.snapshot_logic = jio.value("snapshot_logic", "snap");
            s.io.clean_directory = jio.value("clean_directory", false);
            s.io.mesh_path = jio.value("mesh_path", "");
            s.io.mesh_label = jio.value("mesh_label", "Face Sets");
            s.io.initial_condition_file = jio.value("initial_condition_file", "");
            if (jio.contains("initial_condition_mask"))
                for (const JsonValue& e : jio["initial_condition_mask"].arr)
                    if (e.is_number_integer()) s.io.initial_condition_mask.push_back(e.get<int>());
            s.io.write_3d = jio.value("write_3d", false);
            s.io.n_layers_3d = jio.value("n_layers_3d", 10);
            s.io.output_3d_name = jio.value("output_3d_name", "sol_3d");
        }
        if (j.contains("solver")) {
            const JsonValue& js = j["solver"];
            s.solver.t_end = js.value("t_end", 1.0);
            s.solver.cfl = js.value("cfl", 0.5);
            s.solver.use_deep_adjacency = js.value("use_deep_adjacency", true);
            s.solver.reconstruction_order = js.value("reconstruction_order", 1);
            s.solver.min_dt = js.value("min_dt", 1.0e-12);
            s.solver.limiter = js.value("limiter", "venkatakrishnan");
            s.solver.time_integration = js.value("time_integration", "splitting");
            s.solver.method = js.value("method", "muscl");
            s.solver.refresh_derivative_aux = js.value("refresh_derivative_aux", true);
            s.solver.positivity = js.value("positivity", "none");
            s.solver.wet_dry_eps = js.value("wet_dry_eps", 1.0e-2);
            s.solver.ssp_stages = js.value("ssp_stages", 5);
            s.solver.gradient = js.value("gradient", "green_gauss");
            s.solver.well_balanced = js.value("well_balanced", false);
        }
        if (j.contains("model") && j["model"].contains("parameters"))
            for (const auto& kv : j["model"]["parameters"].obj)
                if (kv.second.is_number()) s.model.parameters[kv.first] = kv.second.get<double>();
        return s;
    }

    // Settings::from_json; a missing file is an error (the reference exits with status 1)
    static Settings from_json(const std::string& path) {
        std::ifstream f(path);
        if (!f.is_open()) throw std::runtime_error("Could not open settings file: " + path);
        std::stringstream ss;
        ss << f.rdbuf();
        return from_text(ss.str());
    }
};

}  // namespace zoomy
