// BaseModel.hpp -- the model-coefficient interface of the reference's
// TNL::MHFEM::BaseModel (BaseModel.h:127-292) for the MI355X solver.
//
// A model derives from BaseModel<n, ...>, sets its coefficients in init()
// (arrays in the CUDA permutations documented in mhfem.h, or uniform values
// as UniformCoefficient does) and may override the host hooks.  The
// device-callable hooks of the reference become host hooks that update the
// coefficient arrays between steps (through deviceArray() on stream()), and
// getBoundaryMobility becomes an array (BC_MOBILITY, i*F + E).
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mhfem.h"
#include "Mesh.hpp"
#include "ParameterContainer.hpp"

namespace mhfem {

enum class MassLumping { enabled, disabled };                                          // MassMatrix.h:16-19
enum class AdvectionDiscretization { explicit_upwind, implicit_upwind, implicit_trace };  // BaseModel.h:12-16

template <int NumberOfEquations_, MassLumping massLumping = MassLumping::enabled,
          AdvectionDiscretization advection = AdvectionDiscretization::explicit_upwind>
class BaseModel {
public:
    static constexpr int NumberOfEquations = NumberOfEquations_;
    static constexpr MassLumping lumping = massLumping;
    static constexpr AdvectionDiscretization Advection = advection;
    // can be overridden by a model with constant mobility (BaseModel.h:163)
    static constexpr bool do_mobility_upwind = true;

    virtual ~BaseModel() = default;

    // called by Solver::setInitialCondition (Solver.hpp:311)
    virtual bool init(const ParameterContainer &) { return true; }
    // hooks (BaseModel.h:248-250)
    virtual void preIterate(double, double) {}
    virtual void postIterate(double, double) {}
    // makeSnapshot (HeatEquationModel.h:164-231): the solver passes the cell values
    virtual void makeSnapshot(double, long, const std::vector<double> &) const {}

    // ---- coefficient access
    void setCoefficient(int which, double uniform) { check(mhfem_set_array(ctx_, which, &uniform, 1)); }
    void setCoefficient(int which, const std::vector<double> &values)
    {
        check(mhfem_set_array(ctx_, which, values.data(), (long)values.size()));
    }
    std::vector<double> getArray(int which) const
    {
        const long n = mhfem_array_size(ctx_, which);
        std::vector<double> out(n > 0 ? n : 0);
        check(mhfem_get_array(ctx_, which, out.data(), n));
        return out;
    }
    // device view for in-place updates by hooks (launch on stream())
    double *deviceArray(int which) const { return static_cast<double *>(mhfem_device_array(ctx_, which)); }
    void *stream() const { return mhfem_stream(ctx_); }

    // BaseModel::setInitialCondition (BaseModel.h:316-326)
    void setInitialCondition(int i, const std::vector<double> &values)
    {
        check(mhfem_set_initial_condition(ctx_, i, values.data(), (long)values.size()));
    }

    // DOF index of (i, E) (BaseModel.h:252-263): Z_iF has the (F, i) permutation
    static long getDofIndex(int i, long E) { return E * NumberOfEquations + i; }
    static long getRowIndex(int i, long E) { return E * NumberOfEquations + i; }

    int cells() const { return cells_; }
    int faces() const { return faces_; }

    // the mesh and the output directory, for makeSnapshot
    const SimplexMesh &getMesh() const { return *mesh_; }
    const std::string &outputDirectory() const { return outputDirectory_; }

    // set by the Solver once the context exists
    void attach(mhfem_ctx *ctx, const SimplexMesh *mesh)
    {
        ctx_ = ctx;
        mesh_ = mesh;
        mhfem_get_counts(ctx, &cells_, &faces_, nullptr);
    }
    void setOutputDirectory(const std::string &dir) { outputDirectory_ = dir; }

protected:
    void check(int rc) const
    {
        if (rc != MHFEM_OK) throw std::runtime_error(mhfem_last_error(ctx_));
    }
    mhfem_ctx *ctx_ = nullptr;
    const SimplexMesh *mesh_ = nullptr;
    std::string outputDirectory_;
    int cells_ = 0, faces_ = 0;
};

}  // namespace mhfem
