// ife_model.hpp -- header-only C++17 host wrapper over the C-ABI of ife.h, shaped like
// IceFEM's Physics / BaseModel plugin interface so that IceFEM-style host code drives the
// MI355X engine with the same life cycle and exception behaviour:
//
//   ife::Physics physics(mesh, {"ux","uy","uz","phase","T"}, {0,0,0,1,2});
//   ife::BaseModel& pf = physics.CreateModel("Fracture/PhaseField", "PhaseField1", mat, &pf, &opts);
//   ife::BaseModel& th = physics.CreateModel("Thermal/ThermalModel", "Thermal", mat);
//   physics.Setup();                 // DofSpace::SyncDofs + InitKf (physics.cpp:40-45)
//   physics.Assemble(step);          // Physics::Assemble (physics.cpp:255-303)
//   pf.Commit(CommitTypes::TIMEDEP); // BaseModel::Commit (BaseModel.h:83)
//   pf.ResetStep();                  // BaseModel::ResetStep (BaseModel.h:86)
//
// Errors become ife::Error (a std::runtime_error carrying the ife_status_code), the way
// the reference throws std::invalid_argument / runtime_error and TimeSolver catches
// std::exception for its dt retry (TimeSolver.cpp:133-171).
#ifndef IFE_MODEL_HPP
#define IFE_MODEL_HPP

#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ife.h"

namespace ife {

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

   private:
    int code_;
};

inline void check(int rc) {
    if (rc != IFE_OK) throw Error(rc, ife_last_error());
}

namespace CommitTypes {
constexpr int TIMEDEP = 1;     // TIMEDEP_COMMIT_TYPE
constexpr int PATHDEP = 2;     // path-dependent commit
}  // namespace CommitTypes

struct Pattern {
    int64_t n_rows = 0, nnz = 0;
    const int64_t* row_ptr = nullptr;  // host
    const int32_t* col = nullptr;      // host
    const int64_t* row_ptr_dev = nullptr;
    const int32_t* col_dev = nullptr;
};

class Physics;

// One registered model (BaseModel, Models/BaseModel/BaseModel.h:32-118).  Assembly is
// driven per step by Physics::Assemble, as in the reference.
class BaseModel {
   public:
    BaseModel(Physics* ph, int index, std::string type, std::string name)
        : physics_(ph), index_(index), ModelName(std::move(type)), MyName(std::move(name)) {}
    void Commit(int commit_type, void* stream = nullptr);
    void ResetStep(void* stream = nullptr);
    // device history array, layout [component][element][ip]
    double* History(const std::string& name, int64_t* n_components = nullptr, int64_t* n_per_component = nullptr);
    int index() const { return index_; }

   private:
    Physics* physics_;
    int index_;

   public:
    const std::string ModelName;  // registered type, e.g. "Fracture/PhaseField"
    const std::string MyName;     // instance name, e.g. "PhaseField1"
};

class Physics {
   public:
    Physics(const ife_mesh_desc& mesh, const std::vector<std::string>& dof_names, const std::vector<int32_t>& dof_steps) {
        if (dof_names.size() != dof_steps.size()) throw Error(IFE_ERR_INVALID, "DofNames/DofStep size mismatch");
        std::vector<const char*> nm;
        for (auto& s : dof_names) nm.push_back(s.c_str());
        ife_dofs_desc d{(int32_t)nm.size(), nm.data(), dof_steps.data()};
        check(ife_physics_create(&mesh, &d, &h_));
    }
    ~Physics() {
        if (h_) ife_physics_destroy(h_);
    }
    Physics(const Physics&) = delete;
    Physics& operator=(const Physics&) = delete;

    // CreateModel + init (ModelRegister.cpp:57-76); models run in creation order
    BaseModel& CreateModel(const std::string& type, const std::string& name, const ife_material& mat,
                           const ife_phasefield_opts* pf = nullptr, const ife_model_opts* opts = nullptr,
                           int32_t elem_group = 0) {
        int32_t idx = -1;
        check(ife_physics_add_model(h_, type.c_str(), name.c_str(), elem_group, &mat, pf, opts, &idx));
        models_.emplace_back(new BaseModel(this, idx, type, name));
        return *models_.back();
    }

    void Setup() { check(ife_physics_setup(h_)); }
    void Plan() { check(ife_physics_plan(h_)); }

    int32_t NumSteps() const {
        int32_t n = 0;
        check(ife_physics_num_steps(h_, &n));
        return n;
    }
    Pattern GetPattern(int32_t step) const {
        Pattern p;
        check(ife_physics_pattern(h_, step, &p.n_rows, &p.nnz, &p.row_ptr, &p.col, &p.row_ptr_dev, &p.col_dev));
        return p;
    }
    std::vector<int64_t> DofNumbers(int32_t field, int64_t n_nodes) const {
        std::vector<int64_t> out((size_t)n_nodes);
        check(ife_physics_dof_numbers(h_, field, out.data()));
        return out;
    }
    double* State(int32_t step, ife_state_kind kind) {
        double* p = nullptr;
        check(ife_physics_state(h_, step, kind, &p));
        return p;
    }
    void SetTime(double t, double dt, double du_dt, double ddu_dt) { check(ife_physics_set_time(h_, t, dt, du_dt, ddu_dt)); }
    void SetPrecision(ife_precision p) { check(ife_physics_set_precision(h_, p)); }

    // Physics::Assemble(step): K/f zeroed, models in order; nullptr = engine-owned buffers
    void Assemble(int32_t step, double* K = nullptr, double* f = nullptr, void* stream = nullptr) {
        check(ife_physics_assemble(h_, step, K, f, stream));
    }
    void AssembleAll(void* stream = nullptr) { check(ife_physics_assemble_all(h_, nullptr, nullptr, stream)); }
    void Kf(int32_t step, double** K, double** f) { check(ife_physics_Kf(h_, step, K, f)); }

    // Constrainer::AddConstraint/Assemble and Physics::Constraint (physics.cpp:329-352):
    // Dirichlet (dof, value) pairs, then K_con = U^T K U, f_con = -U^T (f - K C c) on the device
    void SetConstraints(int32_t step, const std::vector<int64_t>& dofs, const std::vector<double>& values) {
        if (dofs.size() != values.size()) throw Error(IFE_ERR_INVALID, "dofs and values differ in length");
        check(ife_physics_set_constraints(h_, step, (int64_t)dofs.size(), dofs.data(), values.data()));
    }
    void Constraint(int32_t step, void* stream = nullptr) { check(ife_physics_constrain(h_, step, stream)); }
    // Device CSR of the constrained system and its f_con
    void Constrained(int32_t step, int64_t* n_free, int64_t* nnz, const int64_t** row_ptr_dev, const int32_t** col_dev,
                     double** K_con, double** f_con) const {
        check(ife_physics_constrained(h_, step, n_free, nnz, nullptr, nullptr, row_ptr_dev, col_dev, K_con, f_con));
    }
    // TimeScheme and the state vectors: Set_dt, UpdateState + UpdateVelAcc, Commit/ResetStep
    void SetTimeScheme(ife_time_scheme scheme, double c0 = 0.0, double c1 = 0.0) {
        check(ife_physics_set_time_scheme(h_, scheme, c0, c1));
    }
    void SetDt(double t, double dt) { check(ife_physics_set_dt(h_, t, dt)); }
    void UpdateState(int32_t step, const double* dx_dev, void* stream = nullptr) {
        check(ife_physics_update_state(h_, step, dx_dev, stream));
    }
    void CommitState(void* stream = nullptr) { check(ife_physics_commit_state(h_, stream)); }
    void ResetState(void* stream = nullptr) { check(ife_physics_reset_state(h_, stream)); }
    // NonLinSolver::CheckConvergence's three norms over the free dofs (divided by their count)
    // ... into device memory, stream-ordered (capturable in a hipGraph)
    void ResidualNormsAsync(int32_t step, const double* dx_dev, double* out_dev, void* stream = nullptr) {
        check(ife_physics_residual_norms_async(h_, step, dx_dev, out_dev, stream));
    }
    std::array<double, 3> ResidualNorms(int32_t step, const double* dx_dev, void* stream = nullptr) {
        std::array<double, 3> e{};
        check(ife_physics_residual_norms(h_, step, dx_dev, stream, e.data()));
        return e;
    }

    // Throws like NonLinSolver::CheckConvergence on non-finite results (NonLinSolver.cpp:250-269)
    ife_assembly_status Status(void* stream = nullptr, bool throw_on_nonfinite = true) {
        ife_assembly_status s{};
        check(ife_physics_status(h_, stream, &s));
        if (throw_on_nonfinite && s.nonfinite_seen) throw Error(IFE_ERR_INVALID, "NaN/Inf in the assembled system");
        return s;
    }

    ife_physics* handle() { return h_; }

   private:
    ife_physics* h_ = nullptr;
    std::vector<std::unique_ptr<BaseModel>> models_;
    friend class BaseModel;
};

inline void BaseModel::Commit(int commit_type, void* stream) {
    check(ife_model_commit(physics_->h_, index_, commit_type, stream));
}
inline void BaseModel::ResetStep(void* stream) { check(ife_model_reset_step(physics_->h_, index_, stream)); }
inline double* BaseModel::History(const std::string& name, int64_t* nc, int64_t* npc) {
    double* p = nullptr;
    check(ife_model_history(physics_->h_, index_, name.c_str(), &p, nc, npc));
    return p;
}

}  // namespace ife

#endif  // IFE_MODEL_HPP
