// heat_equation.cpp -- the reference's HeatEquation example
// (examples/HeatEquation/HeatEquationModel.h, config.ini) on the MI355X path:
// 2-D heat equation on the unit square, s x s squares split into triangles,
// zero Dirichlet boundary, Gaussian initial condition, BiCGStab + diagonal.
//
//   heat_equation [--hook] [out.bin] [s] [steps] [residue]
// writes the final cell values Z_iK (float64) to out.bin when given; --hook
// runs the model with a preIterate hook that rewrites the diffusivity.
//
//   heat_equation --config config.ini
// runs the reference's file-driven example instead: the mesh (.vtk or .vtu),
// "boundary-conditions-file" (TNL binary), "initial-condition" (.vtu with
// InitialCondition[Z0]) and the solver options come from the INI file (the
// format of examples/HeatEquation/config.ini), and every snapshot is written
// to <output-directory>/data_<step>.vtu with the cell array "Z".
#include <cmath>
#include <cstdio>
#include <iostream>
#include <string>
#include <type_traits>
#include <vector>

#include "mhfem/control.hpp"

using namespace mhfem;

// HeatEquationModel (HeatEquationModel.h:34-235): uniform coefficients, constant mobility
class HeatEquationModel : public BaseModel<1, MassLumping::disabled> {
public:
    static constexpr bool do_mobility_upwind = false;
    std::vector<double> lastSnapshot;

    bool init(const ParameterContainer &p) override
    {
        const double diffusivity = p.getParameter<double>("diffusivity", 1.0);
        setCoefficient(MHFEM_N_ijK, 1.0);
        setCoefficient(MHFEM_u_ijKe, 0.0);
        setCoefficient(MHFEM_m_iK, 1.0);
        setCoefficient(MHFEM_D_ijK, diffusivity);
        setCoefficient(MHFEM_w_iKe, 0.0);
        setCoefficient(MHFEM_a_ijKe, 0.0);
        setCoefficient(MHFEM_r_ijK, 0.0);
        setCoefficient(MHFEM_f_iK, 0.0);
        return true;
    }
    // HeatEquationModel::makeSnapshot (HeatEquationModel.h:164-231): data_<step>.vtu
    // (zlib-compressed VTU) with the cell values as "Z"
    void makeSnapshot(double time, long step, const std::vector<double> &Z) const override
    {
        double mx = 0;
        for (double z : Z) mx = std::max(mx, z);
        std::printf("snapshot %ld  t = %.4e  max Z = %.10e\n", step, time, mx);
        const_cast<HeatEquationModel *>(this)->lastSnapshot = Z;
        if (!outputDirectory().empty()) {
            char name[32];
            std::snprintf(name, sizeof name, "/data_%05ld.vtu", step);
            writeVTU(outputDirectory() + name, getMesh(), step, time, {{"Z", &Z}});
        }
    }
};

// The same model with a hook that rewrites a coefficient every step: the
// reference's preIterate hook (BaseModel.h:248-250), which Solver::preIterate
// calls after the scheme's update (Solver.hpp:873-875), so the new values
// enter the next step's update.  Here the diffusivity grows in time and
// varies in x, D_K = (1 + 50 (t + tau)) (1 + 0.5 x_K), uploaded as a full
// array in the D_ijK layout.
class HookedHeatModel : public HeatEquationModel {
public:
    std::vector<double> centroid_x;
    void preIterate(double time, double tau) override
    {
        std::vector<double> D(centroid_x.size());
        for (size_t K = 0; K < D.size(); K++) D[K] = (1.0 + 50.0 * (time + tau)) * (1.0 + 0.5 * centroid_x[K]);
        setCoefficient(MHFEM_D_ijK, D);
    }
};

static SimplexMesh unit_square(int s)
{
    SimplexMesh m;
    m.dim = 2;
    const int n1 = s + 1;
    for (int iy = 0; iy < n1; iy++)
        for (int ix = 0; ix < n1; ix++) {
            m.points.push_back((double)ix / s);
            m.points.push_back((double)iy / s);
        }
    for (int qy = 0; qy < s; qy++)
        for (int qx = 0; qx < s; qx++) {
            const int v00 = qy * n1 + qx, v10 = v00 + 1, v01 = v00 + n1, v11 = v01 + 1;
            for (int v : {v00, v10, v11, v00, v11, v01}) m.cells.push_back(v);
        }
    return m;
}

// examples/HeatEquation driven by its INI file (control.h::init, solve, writeEpilog)
static int run_config(const char *config)
{
    try {
        ParameterContainer p;
        p.parseINI(config);
        Solver<HeatEquationModel> problem;
        init(problem, p);
        problem.enableTimers();
        solve(problem, p);
        problem.writeEpilog(std::cout);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

// heat_equation [--hook] [out.bin] [s] [steps] [residue]: the heat example on
// an s x s triangulated square (--hook: with HookedHeatModel's coefficient hook)
template <typename Model>
static int run(const char *out, int s, int steps, double residue)
{
    ParameterContainer p;
    p.setParameter("linear-solver", std::string("bicgstab"));
    p.setParameter("preconditioner", std::string("diagonal"));
    p.setParameter("convergence-residue", residue);
    p.setParameter("time-step", 1e-4);
    p.setParameter("initial-time", 0.0);
    p.setParameter("final-time", 1e-4 * steps);
    p.setParameter("snapshot-period", 1e-4 * steps);
    p.setParameter("diffusivity", 1.0);
    try {
        const SimplexMesh mesh = unit_square(s);
        Solver<Model> problem;
        problem.setMesh(mesh);
        // initial condition at cell centroids
        std::vector<double> z0(mesh.numCells()), cxs(mesh.numCells());
        for (int K = 0; K < mesh.numCells(); K++) {
            double cx = 0, cy = 0;
            for (int k = 0; k < 3; k++) {
                cx += mesh.points[2 * mesh.cells[3 * K + k]] / 3;
                cy += mesh.points[2 * mesh.cells[3 * K + k] + 1] / 3;
            }
            z0[K] = std::exp(-((cx - 0.5) * (cx - 0.5) + (cy - 0.5) * (cy - 0.5)) / 0.01);
            cxs[K] = cx;
        }
        if constexpr (std::is_same_v<Model, HookedHeatModel>) problem.getMeshDependentData().centroid_x = cxs;
        problem.getMeshDependentData().setInitialCondition(0, z0);
        if (!problem.setup(p)) throw std::runtime_error("Failed to set up the MHFEM solver.");
        if (!problem.setInitialCondition(p)) throw std::runtime_error("Failed to set up the initial condition.");
        problem.setupLinearSystem();
        problem.enableTimers();
        solve(problem, p);
        problem.writeEpilog(std::cout);
        if (out) {
            const auto &Z = problem.getMeshDependentData().lastSnapshot;
            FILE *f = std::fopen(out, "wb");
            if (!f || std::fwrite(Z.data(), sizeof(double), Z.size(), f) != Z.size()) return 3;
            std::fclose(f);
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 3 && std::string(argv[1]) == "--config") return run_config(argv[2]);
    const bool hook = argc > 1 && std::string(argv[1]) == "--hook";
    const int a = hook ? 1 : 0;
    const char *out = argc > 1 + a ? argv[1 + a] : nullptr;
    const int s = argc > 2 + a ? std::atoi(argv[2 + a]) : 64;
    const int steps = argc > 3 + a ? std::atoi(argv[3 + a]) : 100;
    const double residue = argc > 4 + a ? std::atof(argv[4 + a]) : 1e-10;
    return hook ? run<HookedHeatModel>(out, s, steps, residue) : run<HeatEquationModel>(out, s, steps, residue);
}
