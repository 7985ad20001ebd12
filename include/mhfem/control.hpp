// control.hpp -- the time-step driver of the reference (control.h:15-149).
#pragma once

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "ParameterContainer.hpp"
#include "Solver.hpp"

namespace mhfem {

// control.h::init (control.h:15-62), mesh already loaded
template <typename Problem>
void init(Problem &problem, const SimplexMesh &mesh, const ParameterContainer &parameters)
{
    problem.setMesh(mesh);
    if (!problem.setup(parameters)) throw std::runtime_error("Failed to set up the MHFEM solver.");
    if (!problem.setInitialCondition(parameters)) throw std::runtime_error("Failed to set up the initial condition.");
    problem.setupLinearSystem();
}

// control.h::init (control.h:15-62): the mesh from the "mesh" file
// ("mesh-format" auto, vtu or vtk; a .pvtu distributed mesh is not supported)
template <typename Problem>
void init(Problem &problem, const ParameterContainer &parameters)
{
    const SimplexMesh mesh = loadMesh(parameters.getParameter<std::string>("mesh"),
                                      parameters.getParameter<std::string>("mesh-format", std::string("auto")));
    init(problem, mesh, parameters);
}

// control.h::solve(problem, startTime, stopTime, timeStep) (control.h:64-102)
template <typename Problem>
void solve(Problem &problem, double startTime, double stopTime, double timeStep)
{
    double t = startTime;
    // the loop ends when what is left of the interval is below 1e-6 of a step:
    // such a remainder comes from rounding in the accumulated t, not from the interval
    while (stopTime - t > timeStep * 1e-6) {
        const double currentTau = std::min(timeStep, stopTime - t);
        problem.preIterate(t, currentTau);
        problem.assembleLinearSystem(t, currentTau);
        problem.solveLinearSystem();
        problem.postIterate(t, currentTau);
        t += currentTau;
    }
}

// control.h::solve(problem, parameters, ...) with snapshots (control.h:104-149)
template <typename Problem>
void solve(Problem &problem, const ParameterContainer &parameters)
{
    const double finalTime = parameters.getParameter<double>("final-time");
    const double initialTime = parameters.getParameter<double>("initial-time", 0.0);
    const double snapshotPeriod = parameters.getParameter<double>("snapshot-period");
    const double timeStep = parameters.getParameter<double>("time-step", 1.0);
    if (finalTime <= initialTime) throw std::runtime_error("Final time must larger than the initial time.");
    if (snapshotPeriod <= 0) throw std::runtime_error("Snapshot period must be positive value.");
    if (timeStep <= 0) throw std::runtime_error("Time step must be positive value.");
    double t = initialTime;
    long step = 0;
    const long allSteps = (long)std::ceil((finalTime - initialTime) / snapshotPeriod);
    problem.makeSnapshot(t, step);
    while (step < allSteps) {
        const double tau = std::min(snapshotPeriod, finalTime - t);
        solve(problem, t, t + tau, timeStep);
        step++;
        t += tau;
        problem.makeSnapshot(t, step);
    }
}

}  // namespace mhfem
