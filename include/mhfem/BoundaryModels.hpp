// BoundaryModels.hpp -- boundary models (BoundaryModels/Stationary.h:8-38).
// A boundary model supplies the Neumann and Dirichlet values of (i, E) at
// time + tau.  A model with time_dependent = true is evaluated on the host
// before each assembly (the reference evaluates it inside the assembly
// kernel, Solver.hpp:948); a separable model (stored values times a function
// of time per species) is evaluated on the device instead, through
// Solver::setBoundaryTimeFunction.
#pragma once

#include <vector>

namespace mhfem::BoundaryModels {

// values stored once in the boundary-conditions storage, time independent
struct Stationary {
    static constexpr bool time_dependent = false;
};

}  // namespace mhfem::BoundaryModels
