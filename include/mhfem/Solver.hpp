// Solver.hpp -- C++ facade of the reference's TNL::MHFEM::Solver
// (Solver.h:28-208, Solver.hpp) over the C ABI of libmhfem_amd.so.
// Method names, call order and error messages follow the reference; every
// kernel runs on the MI355X behind mhfem.h.
#pragma once

#include <cstdint>
#include <iomanip>
#include <iostream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mhfem.h"
#include "BaseModel.hpp"
#include "BoundaryModels.hpp"
#include "Mesh.hpp"
#include "ParameterContainer.hpp"

namespace mhfem {

template <typename Model, typename BoundaryModel = BoundaryModels::Stationary>
class Solver {
public:
    using ModelType = Model;
    static constexpr int n = Model::NumberOfEquations;

    ~Solver()
    {
        if (ctx_) mhfem_destroy(ctx_);
    }

    // Solver::setMesh (Solver.hpp:26-81)
    void setMesh(const SimplexMesh &mesh)
    {
        mesh_ = mesh;
        const mhfem_mesh_desc md = mesh_.desc();
        mhfem_model_desc mo{n, Model::lumping == MassLumping::enabled ? 1 : 0, (int)Model::Advection,
                            Model::do_mobility_upwind ? 1 : 0};
        if (ctx_) mhfem_destroy(ctx_);
        ctx_ = mhfem_create_cells(&md, mesh_.vertices(), &mo);
        if (!ctx_) throw std::runtime_error(std::string("setMesh: ") + mhfem_last_error(nullptr));
        mhfem_get_counts(ctx_, &cells_, &faces_, nullptr);
        model_.attach(ctx_, &mesh_);
    }

    // Solver::setup (Solver.hpp:83-253): solver options by name (configSetup.h:31-62)
    bool setup(const ParameterContainer &p)
    {
        const std::string solver = p.getParameter<std::string>("linear-solver", std::string("bicgstab"));
        const std::string prec = p.getParameter<std::string>("preconditioner", std::string("none"));
        cfg_ = mhfem_solver_config{};
        if (solver == "bicgstab") cfg_.solver = MHFEM_SOLVER_BICGSTAB;
        else if (solver == "bicgstabl") cfg_.solver = MHFEM_SOLVER_BICGSTABL;
        else if (solver == "gmres") cfg_.solver = MHFEM_SOLVER_GMRES;
        else if (solver == "cg") cfg_.solver = MHFEM_SOLVER_CG;
        else if (solver == "tfqmr") cfg_.solver = MHFEM_SOLVER_TFQMR;
        else if (solver == "jacobi") cfg_.solver = MHFEM_SOLVER_JACOBI;
        else if (solver == "idrs") cfg_.solver = MHFEM_SOLVER_IDRS;
        else if (solver == "sor") cfg_.solver = MHFEM_SOLVER_SOR;
        else return false;
        cfg_.bicgstab_ell = p.getParameter<int>("bicgstab-ell", 1);
        cfg_.idr_s = p.getParameter<int>("idr-s", 4);
        cfg_.sor_omega = p.getParameter<double>("sor-omega", 1.0);
        cfg_.ilut_p = p.getParameter<int>("ilut-p", 10);
        cfg_.ilut_threshold = p.getParameter<double>("ilut-threshold", 1e-4);
        if (prec == "none") cfg_.preconditioner = MHFEM_PREC_NONE;
        else if (prec == "diagonal") cfg_.preconditioner = MHFEM_PREC_DIAGONAL;
        else if (prec == "ilu0") cfg_.preconditioner = MHFEM_PREC_ILU0;
        // the reference's Ginkgo "AMGX" preconditioner (Solver.hpp:1025-1075)
        else if (prec == "amg" || prec == "AMGX") cfg_.preconditioner = MHFEM_PREC_AMG;
        // its "ILU_ISAI" and "PARILU_ISAI" (Solver.hpp:1089-1130); PARILUT_ISAI is not built
        else if (prec == "ilu-isai" || prec == "ILU_ISAI") cfg_.preconditioner = MHFEM_PREC_ILU_ISAI;
        else if (prec == "parilu-isai" || prec == "PARILU_ISAI") cfg_.preconditioner = MHFEM_PREC_PARILU_ISAI;
        else if (prec == "ilut") cfg_.preconditioner = MHFEM_PREC_ILUT;
        else return false;
        cfg_.convergence_residue = p.getParameter<double>("convergence-residue", 1e-6);
        cfg_.max_iterations = p.getParameter<int>("max-iterations", 100000);
        cfg_.gmres_restart = p.getParameter<int>("gmres-restarting-max", 30);
        const std::string precision = p.getParameter<std::string>("precision", std::string("fp64"));
        if (precision == "fp64") cfg_.precision = MHFEM_PRECISION_FP64;
        else if (precision == "fp32-refined") cfg_.precision = MHFEM_PRECISION_FP32_REFINED;
        else if (precision == "fp32-reliable") cfg_.precision = MHFEM_PRECISION_FP32_RELIABLE;
        else return false;
        cfg_.inner_residue = p.getParameter<double>("inner-residue", 1e-3);
        cfg_.check_every = p.getParameter<int>("check-every", 8);
        guess_.enabled = p.getParameter<std::string>("initial-guess", std::string("previous")) == "learned";
        guess_.train_steps = p.getParameter<int>("guess-train-steps", 100);
        guess_.epochs = p.getParameter<int>("guess-epochs", 20);
        guess_.learning_rate = (float)p.getParameter<double>("guess-learning-rate", 1e-3);
        guess_.seed = (unsigned)p.getParameter<int>("guess-seed", 1);
        guess_.combine = p.getParameter<int>("guess-combine", MHFEM_GUESS_COMBINED);
        guess_.sample_stride = p.getParameter<int>("guess-sample-stride", 1);
        guess_.reject_after = p.getParameter<int>("guess-reject-after", 3);
        guess_.history_depth = p.getParameter<int>("guess-history-depth", 8);
        // Solver.hpp:97: snapshots and the dump of a failed solve go there
        outputDirectory_ = p.getParameter<std::string>("output-directory", std::string());
        model_.setOutputDirectory(outputDirectory_);
        allIterations_ = allTimeSteps_ = 0;
        return true;
    }

    // boundary conditions before setInitialCondition (BoundaryConditions::init, BoundaryConditions.hpp:272-281)
    void setBoundaryConditions(const BoundaryConditionsStorage &s) { bc_ = s; }
    // A time-dependent boundary model evaluated on the device inside the
    // assembly (mhfem_set_boundary_time_function): the stored Dirichlet
    // (which = MHFEM_BC_DIRICHLET) or Neumann (MHFEM_BC_VALUES) values of
    // species i times f(time + tau).  Call after setMesh; nullptr restores
    // the stationary values.  It needs no per-step upload, unlike a
    // BoundaryModel with time_dependent = true, which is evaluated on the
    // host before each assembly.
    void setBoundaryTimeFunction(int which, int species, const mhfem_time_function *f)
    {
        check(mhfem_set_boundary_time_function(ctx_, which, species, f));
    }

    // Solver::setInitialCondition (Solver.hpp:255-344).  The boundary
    // conditions come from "boundary-conditions-file" (TNL's binary format)
    // or from setBoundaryConditions, else zero Dirichlet everywhere; the
    // initial cell values from the "InitialCondition[Z<i>]" cell arrays of
    // "initial-condition", or from the model's setInitialCondition(i, vector)
    // called beforehand, else Z_iK = 0 (Solver.hpp:307-309).
    bool setInitialCondition(const ParameterContainer &p)
    {
        const long dofs = (long)n * faces_;
        if (p.checkParameter("boundary-conditions-file")) {
            const std::string file = p.getParameter<std::string>("boundary-conditions-file");
            BoundaryConditionsStorage storage;
            storage.load(file);
            if (storage.dofSize != dofs) {
                errorMessage_ = "Wrong dofSize in BoundaryConditionsStorage loaded from file " + file +
                                ". Expected " + std::to_string(dofs) + " elements, got " +
                                std::to_string(storage.dofSize) + ".";
                std::cerr << errorMessage_ << std::endl;
                return false;
            }
            bc_ = std::move(storage);
        }
        if (p.checkParameter("initial-condition")) {
            const MeshReader reader(p.getParameter<std::string>("initial-condition"));
            for (int i = 0; i < n; i++)
                model_.setInitialCondition(i, reader.readCellData("InitialCondition[Z" + std::to_string(i) + "]"));
        }
        if (bc_.dofSize == 0) {
            // default: zero Dirichlet everywhere (Solver.hpp:281-294)
            bc_.dofSize = dofs;
            bc_.tags.assign(dofs, (uint8_t)BoundaryConditionsType::FixedValue);
            bc_.values.assign(dofs, 0.0);
            bc_.dirichletValues.assign(dofs, 0.0);
        }
        if (bc_.dofSize != dofs || (long)bc_.tags.size() != dofs) {
            errorMessage_ = "Wrong dofSize in BoundaryConditionsStorage. Expected " + std::to_string(dofs) +
                            " elements, got " + std::to_string(bc_.dofSize) + ".";
            return false;
        }
        check(mhfem_set_boundary_conditions(ctx_, bc_.tags.data(), bc_.values.data(), bc_.dirichletValues.data()));
        if (!model_.init(p)) return false;
        check(mhfem_init(ctx_));
        return true;
    }

    // Solver::setupLinearSystem (Solver.hpp:346-485)
    void setupLinearSystem()
    {
        check(mhfem_setup_linear_system(ctx_, &cfg_));
        if (guess_.enabled) check(mhfem_guess_configure(ctx_, &guess_));
    }

    // ---- one time step (control.h:83-98)
    void preIterate(double time, double tau)
    {
        check(mhfem_pre_iterate(ctx_, time, tau));
        model_.preIterate(time, tau);  // Solver.hpp:873-875
    }
    void assembleLinearSystem(double time, double tau)
    {
        if constexpr (BoundaryModel::time_dependent) {
            // Dirichlet and Neumann values at time + tau (Solver.hpp:948)
            std::vector<double> dir(bc_.dirichletValues), val(bc_.values);
            for (int i = 0; i < n; i++)
                for (long E = 0; E < faces_; E++) {
                    const long k = (long)i * faces_ + E;
                    dir[k] = BoundaryModel::getDirichletValue(bc_, i, E, time + tau, tau);
                    val[k] = BoundaryModel::getNeumannValue(bc_, i, E, time + tau, tau);
                }
            check(mhfem_set_array(ctx_, MHFEM_BC_DIRICHLET, dir.data(), (long)dir.size()));
            check(mhfem_set_array(ctx_, MHFEM_BC_VALUES, val.data(), (long)val.size()));
        }
        check(mhfem_assemble(ctx_, time, tau));
    }
    void solveLinearSystem()
    {
        mhfem_solve_stats st{};
        const int rc = mhfem_solve(ctx_, &st);
        allTimeSteps_++;
        allIterations_ += st.iterations;
        last_ = st;
        if (rc == MHFEM_E_NOT_CONVERGED && !outputDirectory_.empty()) {
            // save the linear system for debugging, then throw (Solver.hpp:1259-1264)
            const std::string msg = mhfem_last_error(ctx_);
            if (mhfem_save_linear_system(ctx_, outputDirectory_.c_str()) == MHFEM_OK)
                std::cerr << "The linear system has been saved to " << outputDirectory_
                          << "/{matrix.mtx,dof.vec.tnl,rhs.vec.tnl}" << std::endl;
            throw std::runtime_error(msg);
        }
        if (rc != MHFEM_OK) throw std::runtime_error(mhfem_last_error(ctx_));
    }
    void postIterate(double time, double tau)
    {
        check(mhfem_post_iterate(ctx_, time, tau));
        model_.postIterate(time, tau);  // Solver.hpp:1315-1317
    }
    void makeSnapshot(double time, long step) { model_.makeSnapshot(time, step, model_.getArray(MHFEM_Z_iK)); }

    // Solver::writeEpilog (Solver.hpp:1326-1391): the reference's timer names
    void writeEpilog(std::ostream &out) const
    {
        mhfem_timers t{};
        mhfem_get_timers(ctx_, &t);
        auto line = [&](const char *name, double ms) {
            out << std::left << std::setw(48) << name << std::scientific << ms * 1e-3 << " s\n";
        };
        out << std::left << std::setw(48) << "Total number of linear solver iterations:" << allIterations_ << "\n";
        out << std::left << std::setw(48) << "Total number of time steps:" << allTimeSteps_ << "\n";
        line("Pre-iterate time:", t.pre_iterate);
        line("  update_b / update_R / update_Q time:", t.update_b_R_Q);
        line("  upwind update time:", t.upwind);
        line("Linear system assembler time:", t.assemble);
        line("Linear preconditioner update time:", t.preconditioner);
        line("Linear system solver time:", t.linear_solver);
        line("Post-iterate time:", t.post_iterate);
    }
    void enableTimers(bool on = true) { mhfem_set_timing(ctx_, on ? 1 : 0); }

    // getters (Solver.h:76-94)
    Model &getMeshDependentData() { return model_; }
    const BoundaryConditionsStorage &getBoundaryConditions() const { return bc_; }
    const SimplexMesh &getMesh() const { return mesh_; }
    const std::string &getOutputDirectory() const { return outputDirectory_; }
    long getDofs() const { return (long)n * faces_; }
    long getLocalDofs() const { return getDofs(); }
    long getGlobalDofs() const { return getDofs(); }
    long getDofsOffset() const { return 0; }
    const mhfem_solve_stats &lastSolve() const { return last_; }
    mhfem_ctx *context() { return ctx_; }
    const std::string &errorMessage() const { return errorMessage_; }

private:
    void check(int rc) const
    {
        if (rc != MHFEM_OK) throw std::runtime_error(mhfem_last_error(ctx_));
    }
    mhfem_ctx *ctx_ = nullptr;
    SimplexMesh mesh_;
    std::string outputDirectory_;
    Model model_;
    BoundaryConditionsStorage bc_;
    mhfem_solver_config cfg_{};
    mhfem_guess_config guess_{};
    mhfem_solve_stats last_{};
    int cells_ = 0, faces_ = 0;
    long long allIterations_ = 0, allTimeSteps_ = 0;
    std::string errorMessage_;
};

}  // namespace mhfem
