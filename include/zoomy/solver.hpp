// Host C++ mirror of the reference's solver classes over the C ABI (include/zdm.h).
//
// Same class and method names as the reference, with the bodies calling the MI355X library
// instead of PETSc loops:
//   VirtualSolver   VirtualSolver.hpp:53-525   (Initialize, SetupInitialConditions,
//                                               ComputeTimeStep, PostStep, Run)
//   ModularSolver   ModularSolver.hpp:13-76    (SetReconstruction, SetLimiters,
//                                               SetGradientMethod, SetFluxKernel,
//                                               SetNonConsFluxKernel, SetStrategy,
//                                               InitializeComponents)
//   MUSCLSolver     MUSCLSolver.hpp:14-254     (Run, EnforcePhysicalConstraints)
//   ChorinVAMSolver ChorinVAM.hpp:86-129       (Run; KSP options from -ksp_* arguments)
//   TransportStep   TransportStep.hpp:43-565   (FormRHS, UpdateState, setters)
//   SourceStep      SourceStep.hpp:56-122      (Solve)
//   SolverStrategy  SolverStrategies.hpp:8-195 (SplittingStrategy, IMEXStrategy, FullyImplicitStrategy)
//   GradientCalculator / Reconstructor / Limiter: the configuration they carry
// Error handling mirrors PetscErrorCode: every call returns 0 on success and a non-zero
// code otherwise, with the message in LastError().  The model is the one the linked
// libzdm_<model>.so was built on (as solver_cpu is built on the generated headers).
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "zdm.h"
#include "zdm_mesh.h"
#include "zoomy/io.hpp"
#include "zoomy/settings.hpp"

namespace zoomy {

using ErrorCode = int;

inline std::string& last_error_slot() {
    static std::string e;
    return e;
}
inline const char* LastError() { return last_error_slot().c_str(); }
inline ErrorCode Fail(const std::string& m) {
    last_error_slot() = m;
    return 1;
}
#define ZOOMY_CALL(x)                                                   \
    do {                                                                \
        const int rc_ = (x);                                            \
        if (rc_ != 0) return ::zoomy::Fail(zdm_last_error());           \
    } while (0)
#define ZOOMY_TRY(x)                                                    \
    do {                                                                \
        const int rc_ = (x);                                            \
        if (rc_ != 0) return rc_;                                       \
    } while (0)

enum ReconstructionType { PCM, LINEAR };
enum class GradientMethod { GREEN_GAUSS, LEAST_SQUARES };
enum class LimiterType { TVD, VENKATAKRISHNAN, NONE };

// ---------------------------------------------------------------- model description
struct ModelInfo {
    int n_q = 0, n_aux = 0, n_parameters = 0, dimension = 0;
    std::string name;
    std::vector<std::string> boundary_tags, parameter_names;
    std::vector<double> default_parameters;

    static std::vector<std::string> names64(int (*fn)(char*, int)) {
        std::vector<char> buf(64 * 64);
        const int n = fn(buf.data(), 64);
        std::vector<std::string> out;
        for (int i = 0; i < n && i < 64; ++i) out.emplace_back(buf.data() + 64 * i);
        return out;
    }
    static ModelInfo query() {
        ModelInfo m;
        zdm_model_info(&m.n_q, &m.n_aux, &m.n_parameters, &m.dimension);
        m.name = zdm_model_name();
        m.boundary_tags = names64(zdm_model_boundary_tags);
        m.parameter_names = names64(zdm_model_parameter_names);
        m.default_parameters.assign(m.n_parameters > 0 ? m.n_parameters : 1, 0.0);
        zdm_model_default_params(m.default_parameters.data());
        m.default_parameters.resize(m.n_parameters);
        return m;
    }
};

// ---------------------------------------------------------------- mesh (DMPlex services)
class Mesh {
public:
    Mesh() = default;
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;
    ~Mesh() {
        if (m_) zdm_mesh_destroy(m_);
    }
    ErrorCode ReadMsh(const std::string& path, int renumber = 1) {
        if (m_) zdm_mesh_destroy(m_);
        m_ = nullptr;
        if (zdm_mesh_read_msh(path.c_str(), renumber, &m_)) return Fail(zdm_mesh_last_error());
        return 0;
    }
    void Adopt(zdm_mesh* m) {
        if (m_) zdm_mesh_destroy(m_);
        m_ = m;
    }
    const zdm_mesh* get() const { return m_; }
    int ncells() const { return m_ ? m_->ncells : 0; }

    // VirtualSolver.hpp:394-417 (model tags must be physical names of the boundary
    // dimension, "default" excepted) and 500-511 (a mesh name maps its tag, else "default"
    // maps tag 1): model BC index per face, -1 for interior / unmapped faces.
    ErrorCode BoundaryMap(const std::vector<std::string>& model_tags, std::vector<int>& face_bc) const {
        struct Phys {
            std::string name;
            int dim, tag;
        };
        std::vector<Phys> ph;
        for (int i = 0; i < m_->nphysical; ++i)
            ph.push_back({std::string(m_->physical_names + 64 * i), m_->physical_dim[i], m_->physical_tag[i]});
        auto lookup = [&](const std::string& n) -> const Phys* {
            const Phys* r = nullptr;
            for (const Phys& p : ph)
                if (p.name == n) r = &p;
            return r;
        };
        for (const std::string& t : model_tags) {
            if (t == "default") continue;
            const Phys* p = lookup(t);
            if (!p || p->dim != m_->dim - 1)
                return Fail("GMSH boundaries must exactly match model tags (excluding 'default').");
        }
        std::vector<std::pair<int, int>> bmap;   // tag -> bc index, later entries win
        for (size_t i = 0; i < model_tags.size(); ++i) {
            int tag = -1;
            if (const Phys* p = lookup(model_tags[i]))
                tag = p->tag;
            else if (model_tags[i] == "default")
                tag = 1;
            if (tag == -1) continue;
            bool set = false;
            for (auto& kv : bmap)
                if (kv.first == tag) {
                    kv.second = (int)i;
                    set = true;
                }
            if (!set) bmap.emplace_back(tag, (int)i);
        }
        face_bc.assign(m_->nfaces, -1);
        for (int f = 0; f < m_->nfaces; ++f) {
            if (m_->face_cells[2 * f + 1] >= 0) continue;
            for (const auto& kv : bmap)
                if (kv.first == m_->face_tag[f]) face_bc[f] = kv.second;
        }
        return 0;
    }

private:
    zdm_mesh* m_ = nullptr;
};

// ------------------------------------------------------- HDF5 initial conditions (float64)
// The "state" dataset (IOManager.hpp:122-192): ncells * nq values, flat or (ncells, nq).
inline ErrorCode ReadH5State(const std::string& path, std::vector<double>& data, int ncells, int nq) {
    if (!std::ifstream(path).good()) return Fail("initial condition file " + path + " cannot be opened");
    int nd = 0;
    uint64_t dims[4] = {0, 0, 0, 0};
    if (zdm_h5_read(path.c_str(), "state", nullptr, 0, &nd, dims)) return Fail(zdm_mesh_last_error());
    uint64_t n = 1;
    for (int i = 0; i < nd; ++i) n *= dims[i];
    const uint64_t want = (uint64_t)ncells * (uint64_t)nq;
    if (n != want || (nd == 2 && dims[1] != (uint64_t)nq) || nd > 2)
        return Fail(path + ": dataset \"state\" has " + std::to_string(n) + " values; the mesh and model need " +
                    std::to_string(ncells) + " x " + std::to_string(nq));
    data.resize(n);
    if (zdm_h5_read(path.c_str(), "state", data.data(), n, &nd, dims)) return Fail(zdm_mesh_last_error());
    return 0;
}

// ---------------------------------------------------------------- .npy arrays (float64)
inline ErrorCode ReadNpy(const std::string& path, std::vector<double>& data, std::vector<size_t>& shape) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return Fail("cannot open " + path);
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) return Fail(path + ": not a .npy file");
    uint32_t hlen = 0;
    if (magic[6] == 1) {
        unsigned char b[2];
        f.read((char*)b, 2);
        hlen = b[0] | (b[1] << 8);
    } else {
        unsigned char b[4];
        f.read((char*)b, 4);
        hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
    }
    std::string h(hlen, ' ');
    f.read(&h[0], hlen);
    if (h.find("'<f8'") == std::string::npos) return Fail(path + ": float64 little-endian array expected");
    if (h.find("'fortran_order': True") != std::string::npos) return Fail(path + ": C order expected");
    const size_t a = h.find('('), b = h.find(')');
    if (a == std::string::npos || b == std::string::npos) return Fail(path + ": bad header");
    shape.clear();
    std::stringstream ss(h.substr(a + 1, b - a - 1));
    std::string tok;
    size_t n = 1;
    while (std::getline(ss, tok, ','))
        if (tok.find_first_of("0123456789") != std::string::npos) {
            shape.push_back(std::stoul(tok));
            n *= shape.back();
        }
    data.resize(n);
    f.read((char*)data.data(), sizeof(double) * n);
    if (!f) return Fail(path + ": truncated data");
    return 0;
}

inline ErrorCode WriteNpy(const std::string& path, const std::vector<double>& data, size_t rows, size_t cols) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return Fail("cannot write " + path);
    std::string h = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + std::to_string(rows) + ", " +
                    std::to_string(cols) + "), }";
    while ((10 + h.size() + 1) % 64) h += ' ';
    h += '\n';
    const uint16_t hl = (uint16_t)h.size();
    f.write("\x93NUMPY\x01\x00", 8);
    f.write((const char*)&hl, 2);
    f.write(h.data(), h.size());
    f.write((const char*)data.data(), sizeof(double) * rows * cols);
    return f ? 0 : Fail("write failed: " + path);
}

// ---------------------------------------------------------------- strategy / component config
class ModularSolver;
class SolverStrategy {
public:
    virtual ~SolverStrategy() = default;
    virtual const char* Name() const = 0;
    // SolverStrategies.hpp:11: the time integrator of the step; here the choice of kernels
    virtual ErrorCode SetupTS(ModularSolver* solver, zdm_options& o) = 0;
};
// explicit transport (TSSSP rks2 / TSEULER) + split implicit source (SolverStrategies.hpp:17-57)
class SplittingStrategy : public SolverStrategy {
public:
    const char* Name() const override { return "splitting"; }
    ErrorCode SetupTS(ModularSolver*, zdm_options& o) override {
        o.source = 1;
        o.time_integration = 0;
        return 0;
    }
};
// explicit transport + implicit source in one additive Runge-Kutta step (TSARKIMEX "2e",
// SolverStrategies.hpp:62-84)
class IMEXStrategy : public SolverStrategy {
public:
    const char* Name() const override { return "imex"; }
    ErrorCode SetupTS(ModularSolver*, zdm_options& o) override {
        o.source = 1;
        o.time_integration = 1;
        return 0;
    }
};
// transport and source implicit together: TSBDF order 2, Newton-Krylov with a matrix-free
// Jacobian (SolverStrategies.hpp:88-195)
class FullyImplicitStrategy : public SolverStrategy {
public:
    const char* Name() const override { return "implicit"; }
    ErrorCode SetupTS(ModularSolver*, zdm_options& o) override {
        o.source = 1;
        o.time_integration = 2;
        return 0;
    }
};

template <typename T>
class Limiter {
public:
    explicit Limiter(LimiterType t) : type(t) {}
    virtual ~Limiter() = default;
    LimiterType type;
};
template <typename T>
class GradientCalculator {
public:
    virtual ~GradientCalculator() = default;
    virtual GradientMethod Method() const = 0;
};
template <typename T>
class GreenGaussGradient : public GradientCalculator<T> {
public:
    GradientMethod Method() const override { return GradientMethod::GREEN_GAUSS; }
};
template <typename T>
class LeastSquaresGradient : public GradientCalculator<T> {
public:
    GradientMethod Method() const override { return GradientMethod::LEAST_SQUARES; }
};
template <typename T>
class Reconstructor {
public:
    virtual ~Reconstructor() = default;
    virtual int Order() const = 0;
    bool limited = false;
};
template <typename T>
class PCMReconstructor : public Reconstructor<T> {
public:
    int Order() const override { return 1; }
};
template <typename T>
class LinearReconstructor : public Reconstructor<T> {
public:
    explicit LinearReconstructor(bool with_limiters) { this->limited = with_limiters; }
    int Order() const override { return 2; }
};
// WBPositivityReconstructor (Reconstruction.hpp:246-287): eta-WB faces from the effective
// gradients of TransportStep::LimitGradientsWB (positivity "mood" / "zhang_shu" at order 2)
template <typename T>
class WBPositivityReconstructor : public Reconstructor<T> {
public:
    explicit WBPositivityReconstructor(T eps) : eps_wet(eps) { this->limited = true; }
    int Order() const override { return 2; }
    T eps_wet;
};

// ---------------------------------------------------------------- the step components
template <typename T>
class TransportStep {
public:
    TransportStep(zdm_solver* h, const ModelInfo& mi) : h_(h), mi_(mi) {}
    void SetReconstruction(std::shared_ptr<Reconstructor<T>> r) { reconstructor = r; }
    void SetGradient(std::shared_ptr<GradientCalculator<T>> g) { gradient = g; }
    // the flux kernels are the model's Numerics<T>, compiled into the library
    template <typename K>
    void SetFluxKernel(K) {}
    template <typename K>
    void SetNonConsFlux(K) {}
    // TransportStep::FormRHS (TransportStep.hpp:325-357) on host arrays, AoS n_q per cell
    ErrorCode FormRHS(double time, const std::vector<T>& X, std::vector<T>& F) {
        F.resize(X.size());
        ZOOMY_CALL(zdm_transport_rhs(h_, time, X.data(), nullptr, F.data()));
        return 0;
    }
    // EnforcePhysicalConstraints / UpdateState on the resident state
    ErrorCode UpdateState() {
        ZOOMY_CALL(zdm_enforce_constraints(h_));
        return 0;
    }
    std::shared_ptr<Reconstructor<T>> reconstructor;
    std::shared_ptr<GradientCalculator<T>> gradient;

private:
    zdm_solver* h_;
    const ModelInfo& mi_;
};

template <typename T>
class SourceStep {
public:
    explicit SourceStep(zdm_solver* h) : h_(h) {}
    // SourceStep::Solve(dt, X, A) (SourceStep.hpp:72-122) on the resident state
    ErrorCode Solve(T dt) {
        ZOOMY_CALL(zdm_source_step(h_, dt));
        return 0;
    }

private:
    zdm_solver* h_;
};

// ---------------------------------------------------------------- solvers
class VirtualSolver {
public:
    virtual ~VirtualSolver() {
        if (h_) zdm_solver_destroy(h_);
    }
    virtual ErrorCode Run(int argc, char** argv) = 0;

    // VirtualSolver::Initialize (VirtualSolver.hpp:194-232): settings from -settings (default
    // settings.json), parameter overrides by name with the reference's messages, mesh, tags,
    // initial condition
    ErrorCode Initialize(int argc, char** argv) {
        std::string path = "settings.json";
        for (int i = 1; i + 1 < argc; ++i)
            if (std::string(argv[i]) == "-settings") path = argv[i + 1];
        try {
            settings = Settings::from_json(path);
        } catch (const std::exception& e) {
            return Fail(e.what());
        }
        info = ModelInfo::query();
        parameters = info.default_parameters;
        for (const auto& kv : settings.model.parameters) {
            bool found = false;
            for (size_t i = 0; i < info.parameter_names.size(); ++i)
                if (info.parameter_names[i] == kv.first) {
                    parameters[i] = kv.second;
                    if (verbose) std::cout << "[INFO] Overriding parameter '" << kv.first << "' = " << kv.second << std::endl;
                    found = true;
                    break;
                }
            if (!found && verbose)
                std::cout << "[WARNING] Settings contained parameter '" << kv.first
                          << "' which is NOT defined in the Model. Ignoring." << std::endl;
        }
        ZOOMY_TRY(SetupArchitecture());
        ZOOMY_TRY(SetupInitialConditions());
        return 0;
    }

    // VirtualSolver::SetupArchitecture (VirtualSolver.hpp:386-525): mesh and BC tag map
    ErrorCode SetupArchitecture() {
        if (settings.io.mesh_path.empty()) return Fail("io.mesh_path is not set");
        ZOOMY_TRY(mesh.ReadMsh(settings.io.mesh_path, 1));
        if (mesh.get()->dim != info.dimension) return Fail("mesh dimension does not match the model");
        return mesh.BoundaryMap(info.boundary_tags, face_bc);
    }

    // VirtualSolver::SetupInitialConditions (VirtualSolver.hpp:234-281): the IC file when set,
    // else Model::initial_condition at the centroids.  The file is read as
    // IOManager::LoadSolution reads it (IOManager.hpp:122-192): an HDF5 dataset "state" of
    // ncells * n_q float64 values, cell-major in the mesh file's cell order (flat, or
    // (ncells, n_q)), mapped to the solver's cell order by cell_perm; without a mask every
    // component is taken, with initial_condition_mask only the listed components (the others
    // stay 0).  A float64 .npy of shape (ncells, k) is read the same way when k = n_q; with
    // k < n_q its columns are the mask's components in order.
    virtual ErrorCode SetupInitialConditions() {
        const zdm_mesh* m = mesh.get();
        const int nc = m->ncells, nq = info.n_q;
        X.assign((size_t)nc * nq, 0.0);
        A.assign((size_t)nc * info.n_aux, 0.0);
        const std::string& icf = settings.io.initial_condition_file;
        if (!icf.empty()) {
            std::vector<double> d;
            size_t cols = 0;
            auto ends = [&](const char* suf) {
                const size_t n = std::strlen(suf);
                return icf.size() >= n && icf.compare(icf.size() - n, n, suf) == 0;
            };
            if (ends(".npy")) {
                std::vector<size_t> sh;
                ZOOMY_TRY(ReadNpy(icf, d, sh));
                if (sh.size() != 2 || (int)sh[0] != nc) return Fail(icf + ": expected shape (ncells, k)");
                cols = sh[1];
            } else {
                ZOOMY_TRY(ReadH5State(icf, d, nc, nq));
                cols = (size_t)nq;
            }
            if (verbose) std::cout << "[INFO] Loading initial condition from " << icf << "..." << std::endl;
            std::vector<int> fields;
            for (int k : settings.io.initial_condition_mask)
                if (k >= 0 && k < nq) fields.push_back(k);
            const bool all = settings.io.initial_condition_mask.empty();
            if (all)
                for (int i = 0; i < nq; ++i) fields.push_back(i);
            if (cols != (size_t)nq && cols < fields.size()) return Fail(icf + ": fewer columns than fields");
            for (int c = 0; c < nc; ++c) {
                const size_t src = (size_t)m->cell_perm[c] * cols;
                for (size_t k = 0; k < fields.size(); ++k)
                    X[(size_t)c * nq + fields[k]] = d[src + (cols == (size_t)nq ? (size_t)fields[k] : k)];
            }
            return 0;
        }
        for (int c = 0; c < nc; ++c)
            if (zdm_model_initial_condition(m->cell_centroid + 3 * (size_t)c, parameters.data(), &X[(size_t)c * nq]))
                return Fail(zdm_last_error());
        return 0;
    }

    // the host state X (solver cell order) in the mesh file's cell order, AoS n_q per cell
    void HostStateInInputOrder(std::vector<double>& out) const {
        const zdm_mesh* m = mesh.get();
        const int nc = m->ncells, nq = info.n_q;
        out.assign(X.size(), 0.0);
        for (int c = 0; c < nc; ++c)
            for (int i = 0; i < nq; ++i) out[(size_t)m->cell_perm[c] * nq + i] = X[(size_t)c * nq + i];
    }

    // VirtualSolver::ComputeTimeStep (VirtualSolver.hpp:285-357) on the resident state
    double ComputeTimeStep() {
        double dt = 0;
        if (zdm_cfl_dt(h_, &dt)) Fail(zdm_last_error());
        return dt;
    }

    // the state in the mesh file's cell order, AoS n_q per cell
    ErrorCode StateInInputOrder(std::vector<double>& out) {
        const zdm_mesh* m = mesh.get();
        const int nc = m->ncells, nq = info.n_q;
        ZOOMY_CALL(zdm_state_download(h_, X.data(), A.data()));
        out.assign(X.size(), 0.0);
        for (int c = 0; c < nc; ++c)
            for (int i = 0; i < nq; ++i) out[(size_t)m->cell_perm[c] * nq + i] = X[(size_t)c * nq + i];
        return 0;
    }

    Settings settings;
    ModelInfo info;
    std::vector<double> parameters;
    Mesh mesh;
    std::vector<int> face_bc;
    std::vector<double> X, A;   // solver cell order
    bool verbose = true;
    bool write_snapshots = true;   // io.snapshots VTK output (the CLI's -no_output turns it off)
    int max_steps = 1 << 30;
    double time = 0;
    int steps = 0;

protected:
    zdm_solver* h_ = nullptr;
};

class ModularSolver : public VirtualSolver {
public:
    void SetStrategy(std::shared_ptr<SolverStrategy> s) { strategy = s; }
    void SetReconstruction(ReconstructionType t) { config_reconstruction_order = t == LINEAR ? 2 : 1; }
    void SetLimiters(bool active) { config_use_limiters = active; }
    void SetGradientMethod(GradientMethod m) { config_grad_method = m; }
    template <typename K>
    void SetFluxKernel(K) {}
    template <typename K>
    void SetNonConsFluxKernel(K) {}

    // ModularSolver::InitializeComponents (ModularSolver.hpp:39-76): the options of the step
    // and the device handle
    ErrorCode InitializeComponents() {
        const SolverSettings& ss = settings.solver;
        if (!strategy) {
            if (ss.time_integration == "splitting")
                strategy = std::make_shared<SplittingStrategy>();
            else if (ss.time_integration == "imex")
                strategy = std::make_shared<IMEXStrategy>();
            else if (ss.time_integration == "implicit")
                strategy = std::make_shared<FullyImplicitStrategy>();
            else
                return Fail("time_integration '" + ss.time_integration + "' is not provided by this build "
                            "(splitting, imex, implicit)");
        }
        int positivity = 0;
        if (ss.positivity == "mood")
            positivity = 1;
        else if (ss.positivity == "zhang_shu")
            positivity = 2;
        else if (ss.positivity != "none")
            return Fail("positivity '" + ss.positivity + "' is not one of none, mood, zhang_shu");
        zdm_options o;
        std::memset(&o, 0, sizeof(o));
        o.order = config_reconstruction_order;
        o.ssp_stages = ss.ssp_stages;
        o.use_limiters = config_use_limiters ? 1 : 0;
        o.gradient = config_grad_method == GradientMethod::LEAST_SQUARES ? 1 : 0;
        o.cfl = ss.cfl;
        o.t_end = ss.t_end;
        o.min_dt = ss.min_dt;
        o.block_size = 0;
        o.well_balanced = ss.well_balanced ? 1 : 0;
        o.positivity = positivity;
        o.wet_dry_eps = ss.wet_dry_eps;
        o.refresh_derivative_aux = ss.refresh_derivative_aux ? 1 : 0;   // ModularSolver.hpp:41
        ZOOMY_TRY(strategy->SetupTS(this, o));
        if (h_) zdm_solver_destroy(h_);
        h_ = nullptr;
        ZOOMY_CALL(zdm_solver_create(mesh.get(), face_bc.data(), parameters.data(), &o, &h_));
        transport = std::make_unique<TransportStep<double>>(h_, info);
        if (o.order >= 2 && positivity != 0) {
            transport->SetReconstruction(std::make_shared<WBPositivityReconstructor<double>>(ss.wet_dry_eps));
            transport->SetGradient(std::make_shared<GreenGaussGradient<double>>());
        } else if (o.order >= 2) {
            transport->SetReconstruction(std::make_shared<LinearReconstructor<double>>(config_use_limiters));
            if (o.gradient)
                transport->SetGradient(std::make_shared<LeastSquaresGradient<double>>());
            else
                transport->SetGradient(std::make_shared<GreenGaussGradient<double>>());
        } else {
            transport->SetReconstruction(std::make_shared<PCMReconstructor<double>>());
        }
        source = std::make_unique<SourceStep<double>>(h_);
        return 0;
    }

    std::shared_ptr<SolverStrategy> strategy;
    std::unique_ptr<TransportStep<double>> transport;
    std::unique_ptr<SourceStep<double>> source;

protected:
    int config_reconstruction_order = 1;
    bool config_use_limiters = true;
    GradientMethod config_grad_method = GradientMethod::GREEN_GAUSS;

    // settings -> component setters (MUSCLSolver.hpp:40-43, Settings extension key gradient)
    void ConfigureFromSettings() {
        SetReconstruction(settings.solver.reconstruction_order >= 2 ? LINEAR : PCM);
        SetLimiters(settings.solver.limiter != "none");
        SetGradientMethod(settings.solver.gradient == "lsq" || settings.solver.gradient == "least_squares"
                              ? GradientMethod::LEAST_SQUARES
                              : GradientMethod::GREEN_GAUSS);
    }
};

class MUSCLSolver : public ModularSolver {
public:
    // MUSCLSolver::Run (MUSCLSolver.hpp:22-194): constraints + first dt, then steps to t_end
    ErrorCode Run(int argc, char** argv) override {
        ZOOMY_TRY(Initialize(argc, argv));
        ConfigureFromSettings();
        ZOOMY_TRY(InitializeComponents());
        ZOOMY_CALL(zdm_state_upload(h_, X.data(), A.data()));
        ZOOMY_CALL(zdm_initialize(h_));
        IOManager io(settings);
        double dt = 0;
        if (write_snapshots && io.enabled()) {
            // the reference's Monitor after every step (MUSCLSolver.hpp:122, 187)
            io.PrepareDirectory();
            ZOOMY_TRY(Monitor(io, 0.0));
            while (steps < max_steps) {
                ZOOMY_CALL(zdm_step(h_, 1));
                int n = 0;
                ZOOMY_CALL(zdm_get_time(h_, &time, &dt, &n));
                if (n == steps) break;   // t_end reached
                steps = n;
                ZOOMY_TRY(Monitor(io, time));
            }
        } else {
            int done = 0;
            ZOOMY_CALL(zdm_run(h_, max_steps, 16, &done));
        }
        ZOOMY_CALL(zdm_get_time(h_, &time, &dt, &steps));
        if (settings.solver.positivity == "mood") {
            long long nt = 0;
            ZOOMY_CALL(zdm_troubled_cells(h_, &nt));
            if (verbose) std::cout << "[MOOD] troubled cells overridden: " << nt << std::endl;
        }
        if (verbose) std::cout << "[INFO] Finished: " << steps << " steps, time " << time << std::endl;
        return 0;
    }
    // MUSCLSolver::EnforcePhysicalConstraints (MUSCLSolver.hpp:234-254)
    ErrorCode EnforcePhysicalConstraints() { return transport->UpdateState(); }

protected:
    // VirtualSolver::Monitor (VirtualSolver.hpp:181-192): a VTK snapshot when one is due
    ErrorCode Monitor(IOManager& io, double t) {
        if (!io.ShouldWrite(t)) return 0;
        ZOOMY_CALL(zdm_state_download(h_, X.data(), A.data()));
        std::string err;
        if (!io.WriteVTK(mesh.get(), X, A, info.n_q, info.n_aux, t, &err)) return Fail(err);
        io.AdvanceSnapshot();
        return 0;
    }
};

class ChorinVAMSolver : public MUSCLSolver {
public:
    // ChorinVAMSolver::Run (ChorinVAM.hpp:86-129).  KSP options as the reference reads them
    // through KSPSetFromOptions: -ksp_gmres_restart, -ksp_rtol, -ksp_atol, -ksp_divtol,
    // -ksp_max_it, -pc_type (pbjacobi only); -zdm_pressure_operator assembled|matrix_free.
    ErrorCode Run(int argc, char** argv) override {
        ZOOMY_TRY(Initialize(argc, argv));
        SetReconstruction(settings.solver.reconstruction_order >= 2 ? LINEAR : PCM);
        SetLimiters(settings.solver.limiter != "none");
        ZOOMY_TRY(InitializeComponents());
        zdm_krylov_options k{40, 400, 1e-9, 1e-12, 1e5, 0};
        for (int i = 1; i + 1 < argc; ++i) {
            const std::string a = argv[i], v = argv[i + 1];
            if (a == "-ksp_gmres_restart") k.restart = std::stoi(v);
            else if (a == "-ksp_rtol") k.rtol = std::stod(v);
            else if (a == "-ksp_atol") k.atol = std::stod(v);
            else if (a == "-ksp_divtol") k.dtol = std::stod(v);
            else if (a == "-ksp_max_it") k.max_it = std::stoi(v);
            else if (a == "-pc_type" && v != "pbjacobi")
                return Fail("-pc_type " + v + " is not provided by this build (pbjacobi only)");
            else if (a == "-zdm_pressure_operator") k.op_mode = v == "assembled" ? 1 : 0;
        }
        ZOOMY_CALL(zdm_chorin_setup(h_, &k));
        ZOOMY_CALL(zdm_state_upload(h_, X.data(), A.data()));
        ZOOMY_CALL(zdm_initialize(h_));
        int n = 0, its = 0, reason = 0;
        double dt = 0;
        IOManager io(settings);
        if (write_snapshots && io.enabled()) {
            // WriteSnapshot before the loop and after every step (ChorinVAM.hpp:101, 124)
            io.PrepareDirectory();
            ZOOMY_TRY(Monitor(io, 0.0));
            while (steps < max_steps) {
                int n1 = 0, its1 = 0, r1 = 0;
                ZOOMY_CALL(zdm_chorin_step(h_, 1, &n1, &its1, &r1));
                if (n1 == 0) break;   // t_end reached
                its += its1;
                if (r1 < 0 || reason >= 0) reason = r1;
                ZOOMY_CALL(zdm_get_time(h_, &time, &dt, &steps));
                ZOOMY_TRY(Monitor(io, time));
            }
        } else {
            ZOOMY_CALL(zdm_chorin_step(h_, max_steps, &n, &its, &reason));
        }
        ZOOMY_CALL(zdm_get_time(h_, &time, &dt, &steps));
        last_reason = reason;
        total_its = its;
        if (verbose) {
            if (reason < 0)
                std::cout << "[KSP] pressure NOT converged (reason " << reason << ", " << its << " its in all)" << std::endl;
            std::cout << "[INFO] Finished (Chorin VAM): " << steps << " steps, time " << time << std::endl;
        }
        return 0;
    }
    int last_reason = 0, total_its = 0;
};

}  // namespace zoomy
