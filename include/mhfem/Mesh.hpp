// Mesh.hpp -- simplex meshes and the files on either side of the time-step
// path, over the C ABI of libmhfem_amd.so:
//   * loadMesh / MeshReader: .vtu or .vtk meshes with their cell data
//     (TNL::Meshes::loadMesh, control.h:20-48; readCellData, Solver.hpp:296-306);
//   * BoundaryConditionsStorage with save/load in TNL's binary format
//     (BoundaryConditionsStorage.h:9-72);
//   * writeVTU: the snapshot file of HeatEquationModel::makeSnapshot
//     (HeatEquationModel.h:164-231).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../mhfem.h"

namespace mhfem {

// mesh: points (dim per vertex) and cells of verticesPerCell vertices:
// simplices (dim + 1; local face e is opposite local vertex e), edges (2) or
// axis-aligned quadrangles (4 in 2-D, counter-clockwise from the lower-left
// vertex; faces (0,1), (1,2), (2,3), (3,0))
struct SimplexMesh {
    int dim = 2;
    int verticesPerCell = 0;  // 0: dim + 1
    std::vector<double> points;
    std::vector<int32_t> cells;
    int vertices() const { return verticesPerCell ? verticesPerCell : dim + 1; }
    int numPoints() const { return (int)(points.size() / dim); }
    int numCells() const { return (int)(cells.size() / vertices()); }
    mhfem_mesh_desc desc() const { return mhfem_mesh_desc{dim, numPoints(), numCells(), points.data(), cells.data()}; }
};

enum class BoundaryConditionsType : uint8_t {  // BoundaryConditionsType.h:7-31
    FixedValue = 0,
    FixedFlux = 1,
    AdvectiveOutflow = 2,
    FixedFluxNoAdvection = 3
};

namespace detail {
inline void io_check(long rc)
{
    if (rc < 0) throw std::runtime_error(mhfem_last_error(nullptr));
}
}  // namespace detail

// BoundaryConditionsStorage (BoundaryConditionsStorage.h:9-72): arrays of n*F
// in the i*F+E layout; save/load write and read TNL's binary object format
struct BoundaryConditionsStorage {
    std::int64_t dofSize = 0;
    std::vector<uint8_t> tags;
    std::vector<double> values, dirichletValues;

    void save(const std::string &fileName) const
    {
        if ((std::int64_t)tags.size() != dofSize || (std::int64_t)values.size() != dofSize ||
            (std::int64_t)dirichletValues.size() != dofSize)
            throw std::runtime_error("BoundaryConditionsStorage: arrays must have dofSize elements");
        detail::io_check(mhfem_bc_save(fileName.c_str(), dofSize, tags.data(), values.data(), dirichletValues.data()));
    }
    void load(const std::string &fileName)
    {
        std::int64_t n = 0;
        detail::io_check(mhfem_bc_load(fileName.c_str(), &n, nullptr, nullptr, nullptr, 0));
        tags.resize(n);
        values.resize(n);
        dirichletValues.resize(n);
        detail::io_check(mhfem_bc_load(fileName.c_str(), &n, tags.data(), values.data(), dirichletValues.data(), n));
        dofSize = n;
    }
};

// a mesh file and its cell data (getMeshReader(file, format), Solver.hpp:298-306)
class MeshReader {
public:
    explicit MeshReader(const std::string &fileName, const std::string &format = "auto")
        : f_(mhfem_mesh_file_open(fileName.c_str(), format.c_str()))
    {
        if (!f_) throw std::runtime_error(mhfem_last_error(nullptr));
    }
    ~MeshReader() { mhfem_mesh_file_close(f_); }
    MeshReader(const MeshReader &) = delete;
    MeshReader &operator=(const MeshReader &) = delete;

    SimplexMesh mesh() const
    {
        int dim = 0, nvc = 0;
        long np = 0, nc = 0;
        detail::io_check(mhfem_mesh_file_info(f_, &dim, &nvc, &np, &nc));
        SimplexMesh m;
        m.dim = dim;
        m.verticesPerCell = nvc;
        m.points.resize((size_t)np * dim);
        m.cells.resize((size_t)nc * nvc);
        detail::io_check(mhfem_mesh_file_read(f_, m.points.data(), m.cells.data()));
        return m;
    }
    std::vector<double> readCellData(const std::string &name) const
    {
        const long n = mhfem_mesh_file_cell_data(f_, name.c_str(), nullptr, 0);
        detail::io_check(n);
        std::vector<double> v(n);
        detail::io_check(mhfem_mesh_file_cell_data(f_, name.c_str(), v.data(), n));
        return v;
    }
    std::vector<std::string> cellDataNames() const
    {
        std::vector<std::string> out;
        const int k = mhfem_mesh_file_cell_data_count(f_);
        for (int i = 0; i < k; i++) out.emplace_back(mhfem_mesh_file_cell_data_name(f_, i));
        return out;
    }

private:
    mhfem_mesh_file *f_;
};

// TNL::Meshes::loadMesh(mesh, file, format) (control.h:46)
inline SimplexMesh loadMesh(const std::string &fileName, const std::string &format = "auto")
{
    return MeshReader(fileName, format).mesh();
}

// a .vtu file of the mesh with cell arrays (VTUWriter: writeMetadata(cycle, time),
// writeEntities, writeCellData); format MHFEM_VTK_ZLIB as the reference's snapshots
inline void writeVTU(const std::string &fileName, const SimplexMesh &mesh, long cycle, double time,
                     const std::vector<std::pair<std::string, const std::vector<double> *>> &cellData,
                     int format = MHFEM_VTK_ZLIB)
{
    std::vector<const char *> names;
    std::vector<const double *> values;
    for (const auto &c : cellData) {
        if ((long)c.second->size() != mesh.numCells())
            throw std::runtime_error("writeVTU: cell data '" + c.first + "' does not have one value per cell");
        names.push_back(c.first.c_str());
        values.push_back(c.second->data());
    }
    const mhfem_mesh_desc d = mesh.desc();
    detail::io_check(mhfem_write_vtu(fileName.c_str(), &d, mesh.vertices(), cycle, time, (int)names.size(),
                                     names.data(), values.data(), format));
}

}  // namespace mhfem
