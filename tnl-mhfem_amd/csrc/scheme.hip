// scheme.hip -- MHFEM scheme kernels for gfx950: local mass matrices and
// eliminations (preIterate), face-parallel hybridised assembly, recovery of
// the cell averages and velocities (postIterate).
//
// Kernels are templated on the number of equations N and faces per cell FPC
// (3 = triangles, 4 = tetrahedra); every per-element quantity of one cell
// lives in registers of one lane.  Arrays are in the reference's CUDA
// permutations, i.e. the cell (or face) index is the fastest one, so the
// cell-parallel kernels read and write with unit stride across a wavefront.
#include "common.hpp"

namespace mhfem {

constexpr int WG = 256;  // work-group size (4 wavefronts)

template <int FPC>
struct Simplex {
    static constexpr int DIM = FPC - 1;
    static constexpr int NP = FPC * (FPC + 1) / 2;  // packed upper triangle
    static constexpr int S = NP + FPC;              // + row sums (MassMatrix::size)
};

// packed index of (e,f) (MassMatrix.h:613-616)
__host__ __device__ constexpr int pk(int e, int f) { return e <= f ? e + f * (f + 1) / 2 : f + e * (e + 1) / 2; }

// LU without pivoting and the in-place solve (the method of TNL's
// LU_sequential_factorize / LU_sequential_solve_inplace, used at
// MassMatrix.h:571-592 and Solver.hpp:849-866).
template <int M>
__device__ __forceinline__ void lu_factorize(double (&A)[M][M])
{
#pragma unroll
    for (int k = 0; k < M; k++)
#pragma unroll
        for (int i = k + 1; i < M; i++) {
            A[i][k] /= A[k][k];
#pragma unroll
            for (int j = k + 1; j < M; j++) A[i][j] -= A[i][k] * A[k][j];
        }
}
template <int M>
__device__ __forceinline__ void lu_solve(const double (&A)[M][M], double (&x)[M])
{
#pragma unroll
    for (int i = 1; i < M; i++)
#pragma unroll
        for (int j = 0; j < i; j++) x[i] -= A[i][j] * x[j];
#pragma unroll
    for (int i = M - 1; i >= 0; i--) {
#pragma unroll
        for (int j = i + 1; j < M; j++) x[i] -= A[i][j] * x[j];
        x[i] /= A[i][i];
    }
}

// Geometric part of the RTN0 local mass matrix of one simplex and its inverse
// (packed upper triangle, column by column).  The closed forms are those of
// MassMatrix.h:543-569 (triangle), :666-695 (triangle, lumped), :793-831
// (tetrahedron), :938-978 (tetrahedron, lumped).
template <int FPC>
__device__ __forceinline__ void geom_minv(const double (&P)[FPC][FPC - 1], double meas, int lumping,
                                          double (&out)[Simplex<FPC>::NP])
{
    double M[FPC][FPC];
    if constexpr (FPC == 3) {
        if (!lumping) {
            const double p0x = P[0][0] - P[2][0], p0y = P[0][1] - P[2][1];
            const double p1x = P[1][0] - P[2][0], p1y = P[1][1] - P[2][1];
            const double P00 = p0x * p0x + p0y * p0y, P01 = p0x * p1x + p0y * p1y,
                         P11 = p1x * p1x + p1y * p1y;
            const double den = 24 * meas;
            M[0][0] = (3 * P00 + P11 - 3 * P01) / den;
            M[1][1] = (P00 + 3 * P11 - 3 * P01) / den;
            M[2][2] = (P00 + P11 + P01) / den;
            M[0][1] = (3 * P01 - P00 - P11) / den;
            M[0][2] = (P11 - P00 - P01) / den;
            M[1][2] = (P00 - P11 - P01) / den;
        } else {
            const double a01[2] = {P[0][0] - P[1][0], P[0][1] - P[1][1]};
            const double a02[2] = {P[0][0] - P[2][0], P[0][1] - P[2][1]};
            const double a12[2] = {P[1][0] - P[2][0], P[1][1] - P[2][1]};
            auto d2 = [](const double *x, const double *y) { return x[0] * y[0] + x[1] * y[1]; };
            const double den = 12 * meas;
            M[0][0] = (d2(a01, a01) + d2(a02, a02)) / den;
            M[1][1] = (d2(a01, a01) + d2(a12, a12)) / den;
            M[2][2] = (d2(a02, a02) + d2(a12, a12)) / den;
            M[0][1] = (d2(a02, a12)) / den;
            M[0][2] = (-d2(a01, a12)) / den;
            M[1][2] = (d2(a01, a02)) / den;
        }
    } else {
        auto d3 = [](const double *x, const double *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
        if (!lumping) {
            double p0[3], p1[3], p2[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                p0[k] = P[0][k] - P[3][k];
                p1[k] = P[1][k] - P[3][k];
                p2[k] = P[2][k] - P[3][k];
            }
            const double P00 = d3(p0, p0), P11 = d3(p1, p1), P22 = d3(p2, p2);
            const double P01 = d3(p0, p1), P02 = d3(p0, p2), P12 = d3(p1, p2);
            const double den = 180 * meas;
            M[0][0] = (12 * P00 + 2 * P11 + 2 * P22 - 8 * P01 - 8 * P02 + 2 * P12) / den;
            M[1][1] = (2 * P00 + 12 * P11 + 2 * P22 - 8 * P01 + 2 * P02 - 8 * P12) / den;
            M[2][2] = (2 * P00 + 2 * P11 + 12 * P22 + 2 * P01 - 8 * P02 - 8 * P12) / den;
            M[3][3] = 2 * (P00 + P11 + P22 + P01 + P02 + P12) / den;
            M[0][1] = (-3 * P00 - 3 * P11 + 2 * P22 + 12 * P01 - 3 * P02 - 3 * P12) / den;
            M[0][2] = (-3 * P00 + 2 * P11 - 3 * P22 - 3 * P01 + 12 * P02 - 3 * P12) / den;
            M[1][2] = (2 * P00 - 3 * P11 - 3 * P22 - 3 * P01 - 3 * P02 + 12 * P12) / den;
            M[0][3] = (-3 * P00 + 2 * P11 + 2 * P22 - 3 * P01 - 3 * P02 + 2 * P12) / den;
            M[1][3] = (2 * P00 - 3 * P11 + 2 * P22 - 3 * P01 + 2 * P02 - 3 * P12) / den;
            M[2][3] = (2 * P00 + 2 * P11 - 3 * P22 + 2 * P01 - 3 * P02 - 3 * P12) / den;
        } else {
            double a01[3], a02[3], a03[3], a12[3], a13[3], a23[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                a01[k] = P[0][k] - P[1][k];
                a02[k] = P[0][k] - P[2][k];
                a03[k] = P[0][k] - P[3][k];
                a12[k] = P[1][k] - P[2][k];
                a13[k] = P[1][k] - P[3][k];
                a23[k] = P[2][k] - P[3][k];
            }
            const double den = 36 * meas;
            M[0][0] = (d3(a01, a01) + d3(a02, a02) + d3(a03, a03)) / den;
            M[1][1] = (d3(a01, a01) + d3(a12, a12) + d3(a13, a13)) / den;
            M[2][2] = (d3(a02, a02) + d3(a12, a12) + d3(a23, a23)) / den;
            M[3][3] = (d3(a03, a03) + d3(a13, a13) + d3(a23, a23)) / den;
            M[0][1] = (d3(a02, a12) + d3(a03, a13)) / den;
            M[0][2] = (-d3(a01, a12) + d3(a03, a23)) / den;
            M[0][3] = (-d3(a01, a13) - d3(a02, a23)) / den;
            M[1][2] = (d3(a01, a02) + d3(a13, a23)) / den;
            M[1][3] = (d3(a01, a03) - d3(a12, a23)) / den;
            M[2][3] = (d3(a02, a03) + d3(a12, a13)) / den;
        }
    }
#pragma unroll
    for (int i = 0; i < FPC; i++)
#pragma unroll
        for (int j = 0; j < i; j++) M[i][j] = M[j][i];
    lu_factorize<FPC>(M);
#pragma unroll
    for (int c = 0; c < FPC; c++) {
        double vv[FPC];
#pragma unroll
        for (int k = 0; k < FPC; k++) vv[k] = (k == c) ? 1.0 : 0.0;
        lu_solve<FPC>(M, vv);
#pragma unroll
        for (int r = 0; r <= c; r++) out[pk(r, c)] = vv[r];
    }
}

// One entry of b_ijK = M^{-1} D_ij (MassMatrix.h:586-592), rounded once as a
// stored b_ijK_storage value is: contraction is off, so the product never
// fuses with the operation that consumes it.
__device__ __forceinline__ double b_entry(double minv_ef, double D)
{
#pragma clang fp contract(off)
    return minv_ef * D;
}

// Row sum of row e of b_ijK, ascending f (MassMatrix.h:594-597, 864-868), from
// mrow[f] = M^{-1}_ef.  Contraction is off: every kernel that forms the sum
// gets the same bits.
template <int FPC>
__device__ __forceinline__ double row_sum(const double (&mrow)[FPC], double D)
{
#pragma clang fp contract(off)
    double s = mrow[0] * D;
#pragma unroll
    for (int f = 1; f < FPC; f++) s += mrow[f] * D;
    return s;
}

// M^{-1} of every cell is kept twice (both fixed, written once):
//  - s.minv: the packed upper triangle in NP planes, s.minv[q*C + K], for
//    the cell-parallel kernels (unit stride across a wavefront);
//  - s.minv_rows: row-expanded, s.minv_rows[K*MS + e*FPC + f] = M^{-1}_ef
//    (MS = FPC*FPC rounded up to an even count, 16-byte aligned cells), for
//    the face kernels, which gather row e of one cell as one piece.
template <int FPC>
struct MinvLayout {
    static constexpr int MS = (FPC * FPC + 1) & ~1;
};

template <int FPC>
__device__ __forceinline__ void load_minv(const SchemeView &s, int K, double (&m)[FPC][FPC])
{
#pragma unroll
    for (int f = 0; f < FPC; f++)
#pragma unroll
        for (int e = 0; e <= f; e++) {
            const double v = s.minv[(long)pk(e, f) * s.C + K];
            m[e][f] = v;
            m[f][e] = v;
        }
}

// row e (a runtime index) of cell K's M^{-1}
template <int FPC>
__device__ __forceinline__ void load_minv_row(const SchemeView &s, int K, int e, double (&mrow)[FPC])
{
    const double *p = s.minv_rows + (long)K * MinvLayout<FPC>::MS + e * FPC;
    if constexpr (FPC == 4) {
        const Pack<double, 2> a = ld<double, 2>(p), b = ld<double, 2>(p + 2);
        mrow[0] = a.v[0];
        mrow[1] = a.v[1];
        mrow[2] = b.v[0];
        mrow[3] = b.v[1];
    } else {
#pragma unroll
        for (int f = 0; f < FPC; f++) mrow[f] = p[f];
    }
}

// Inputs of the face assembly that k_cell produces for cell K, in the order
// the face kernels read them, one contiguous 16-byte aligned record per cell
// (reading them from the reference-layout arrays, one array per index
// combination, would fetch a partial cache line per value):
//   [(j*FPC + f)*N + x]  R_xjKf
//   [RK_OFF + j]         R_jK
//   [D_OFF + i*N + j]    D_ijK
template <int N, int FPC>
struct CellRec {
    static constexpr int RK_OFF = N * N * FPC, D_OFF = RK_OFF + N;
    static constexpr int SIZE = (D_OFF + N * N + 1) & ~1;
    // The cell-parallel kernels move a work-group's records through LDS
    // (lane t's record at t*LS, one double of padding against bank
    // conflicts) so that global memory sees unit-stride lanes; a record per
    // lane straight from global memory would touch a cache line per lane
    // and instruction.  Only where the block fits in 64 KB of LDS.
    static constexpr int LS = SIZE + 1;
    static constexpr bool STAGE = LS * 256 * 8 <= 65536;
};

// Copies between the records of cells [first, first + count) in global
// memory (contiguous) and their LDS slots; called by the whole work-group.
template <int SIZE, int LS>
__device__ __forceinline__ void records_to_global(double *g, const double *lds, int count)
{
    const int total = count * SIZE;
    for (int q = threadIdx.x; q < total; q += blockDim.x) g[q] = lds[(q / SIZE) * LS + q % SIZE];
}
template <int SIZE, int LS>
__device__ __forceinline__ void records_from_global(const double *g, double *lds, int count)
{
    const int total = count * SIZE;
    for (int q = threadIdx.x; q < total; q += blockDim.x) lds[(q / SIZE) * LS + q % SIZE] = g[q];
}

// ---------------------------------------------------------------------------
// Inverse of the geometric RTN0 mass matrix of every cell (Solver.hpp:627-639,
// MassMatrix::update).  It depends only on the mesh and the lumping, both
// fixed when the context is created, so it is computed once; b_ijK =
// M^{-1} D_ij is formed from it where it is used.
template <int FPC>
__global__ void __launch_bounds__(WG) k_geometry(SchemeView s, double *minv_out, double *rows_out)
{
    constexpr int DIM = FPC - 1, NP = Simplex<FPC>::NP, MS = MinvLayout<FPC>::MS;
    const int K = blockIdx.x * WG + threadIdx.x;
    if (K >= s.C) return;
    double P[FPC][DIM];
#pragma unroll
    for (int e = 0; e < FPC; e++) {
        const int vtx = s.cell_vertices[(long)K * FPC + e];
#pragma unroll
        for (int c = 0; c < DIM; c++) P[e][c] = s.points[(long)vtx * DIM + c];
    }
    double minv[NP];
    geom_minv<FPC>(P, s.cell_meas[K], s.lumping, minv);
#pragma unroll
    for (int q = 0; q < NP; q++) minv_out[(long)q * s.C + K] = minv[q];
    double *o = rows_out + (long)K * MS;
#pragma unroll
    for (int e = 0; e < FPC; e++)
#pragma unroll
        for (int f = 0; f < FPC; f++) o[e * FPC + f] = minv[pk(e, f)];
    if constexpr (MS > FPC * FPC) o[FPC * FPC] = 0.0;
}

// Box cells: the direction of local face e (the slot of MassMatrix::size it
// reads) and the two faces of direction d.  Edges: one direction; quadrangles:
// horizontal faces e = 0, 2 and vertical e = 1, 3 (MassMatrix.h:178-184);
// hexahedra: n_x faces e = 2, 4, n_y e = 1, 3, n_z e = 0, 5 (MassMatrix.h:430-436).
template <int FPC>
__host__ __device__ constexpr int box_dir(int e)
{
    return FPC == 2 ? 0 : FPC == 4 ? e % 2 : (e == 2 || e == 4) ? 0 : (e == 1 || e == 3) ? 1 : 2;
}
template <int FPC>
__host__ __device__ constexpr int box_face(int d, int second)
{
    return FPC == 2 ? second : FPC == 4 ? d + 2 * second : (d == 0 ? 2 + 2 * second : d == 1 ? 1 + 2 * second : 5 * second);
}

// The same for box cells (MassMatrix.h:28-508): an edge (h_x = |K|), an
// axis-aligned rectangle (h_x, h_y from local vertices 0 and 2) or a voxel
// (h_x, h_y, h_z from local vertices 0 and 6).  Each face direction d has one
// value: g = 2/h_x for an edge; 2 h_x/h_y (horizontal faces) and 2 h_y/h_x
// (vertical faces) for a rectangle; 2 h_y h_z/h_x, 2 h_x h_z/h_y and
// 2 h_x h_y/h_z (n_x, n_y, n_z faces) for a voxel.  M^{-1} holds g_d on the
// diagonal when lumped, else 2 g_d on the diagonal and g_d between the two
// faces of direction d, so that its row sums are MassMatrix::b_ijKe (g_d or
// 3 g_d).
template <int FPC>
__global__ void __launch_bounds__(WG) k_geometry_box(SchemeView s, double *minv_out, double *rows_out)
{
    constexpr int DIM = FPC / 2, NVC = 1 << DIM, NP = Simplex<FPC>::NP, MS = MinvLayout<FPC>::MS;
    const int K = blockIdx.x * WG + threadIdx.x;
    if (K >= s.C) return;
    double g[DIM];
    if constexpr (DIM == 1) {
        g[0] = 2.0 / s.cell_meas[K];
    } else {
        // the opposite corner: vertex 2 of a rectangle, 6 of a voxel
        const int v0 = s.cell_vertices[(long)K * NVC], vc = s.cell_vertices[(long)K * NVC + (DIM == 2 ? 2 : 6)];
        double h[DIM];
#pragma unroll
        for (int a = 0; a < DIM; a++) h[a] = s.points[(long)vc * DIM + a] - s.points[(long)v0 * DIM + a];
        if constexpr (DIM == 2) {
            g[0] = 2.0 * h[0] / h[1];
            g[1] = 2.0 * h[1] / h[0];
        } else {
            g[0] = 2.0 * h[1] * h[2] / h[0];
            g[1] = 2.0 * h[0] * h[2] / h[1];
            g[2] = 2.0 * h[0] * h[1] / h[2];
        }
    }
    double minv[NP];
#pragma unroll
    for (int f = 0; f < FPC; f++)
#pragma unroll
        for (int e = 0; e <= f; e++) {
            const int d = box_dir<FPC>(e);
            double v = 0.0;
            if (e == f) v = s.lumping ? g[d] : 2.0 * g[d];
            else if (box_dir<FPC>(f) == d && !s.lumping) v = g[d];
            minv[pk(e, f)] = v;
        }
#pragma unroll
    for (int q = 0; q < NP; q++) minv_out[(long)q * s.C + K] = minv[q];
    double *o = rows_out + (long)K * MS;
#pragma unroll
    for (int e = 0; e < FPC; e++)
#pragma unroll
        for (int f = 0; f < FPC; f++) o[e * FPC + f] = minv[pk(e, f)];
    if constexpr (MS > FPC * FPC) o[FPC * FPC] = 0.0;
}

// b_ijK_storage of box cells: one value per face direction,
// b_ijK_storage(i,j,K,d) = g_d D_ijK (MassMatrix.h:52-53, 172-175).
template <int N, int FPC>
__global__ void __launch_bounds__(WG) k_fill_b_box(SchemeView s, double *b)
{
    constexpr int DIM = FPC / 2;
    const int K = blockIdx.x * WG + threadIdx.x;
    if (K >= s.C) return;
    const long C = s.C;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            const double D = s.D[((long)i * N + j) * C + K];
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                // g_d: the lumped diagonal, or the entry between the two faces of direction d
                const int e0 = box_face<FPC>(d, 0), e1 = box_face<FPC>(d, 1);
                const double g = s.minv[(long)(s.lumping ? pk(e0, e0) : pk(e0, e1)) * C + K];
                b[((long)(i * N + j) * DIM + d) * C + K] = b_entry(g, D);
            }
        }
}

// b_ijK_storage ((i*n+j)*S + q)*C + K: the packed b_ijK followed by its row
// sums (MassMatrix.h:586-597), written only when a caller asks for the array.
template <int N, int FPC>
__global__ void __launch_bounds__(WG) k_fill_b(SchemeView s, double *b)
{
    constexpr int NP = Simplex<FPC>::NP, S = Simplex<FPC>::S;
    const int K = blockIdx.x * WG + threadIdx.x;
    if (K >= s.C) return;
    const long C = s.C;
    double m[FPC][FPC];
    load_minv<FPC>(s, K, m);
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) {
            const double D = s.D[((long)i * N + j) * C + K];
            double *bp = b + ((long)(i * N + j) * S) * C + K;
#pragma unroll
            for (int f = 0; f < FPC; f++)
#pragma unroll
                for (int e = 0; e <= f; e++) bp[(long)pk(e, f) * C] = b_entry(m[e][f], D);
#pragma unroll
            for (int e = 0; e < FPC; e++) bp[(long)(NP + e) * C] = row_sum<FPC>(m[e], D);
        }
}

// ---------------------------------------------------------------------------
// Solver::setInitialCondition face averages (Solver.hpp:319-341)
template <int N>
__global__ void __launch_bounds__(WG) k_face_average(SchemeView s)
{
    const int E = blockIdx.x * WG + threadIdx.x;
    if (E >= s.F) return;
    const int K0 = s.face_cells[2 * E], K1 = s.face_cells[2 * E + 1];
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double z0 = s.Z_iK[(long)i * s.C + K0];
        s.Z_iF[(long)E * N + i] = (K1 < 0) ? z0 : 0.5 * (z0 + s.Z_iK[(long)i * s.C + K1]);
    }
}

// ---------------------------------------------------------------------------
// Upwind coefficients over faces: K4 mobility upwind (Solver.hpp:663-727) and
// K5 explicit upwind of Z (Solver.hpp:729-769).  One lane per face handles all
// components.
template <int N, int FPC>
__global__ void __launch_bounds__(WG) k_upwind(SchemeView s)
{
    const int E = blockIdx.x * WG + threadIdx.x;
    if (E >= s.F) return;
    const uint64_t info = s.face_slots[E];
    const int K1 = s.face_cells[2 * E], K2 = s.face_cells[2 * E + 1];
    const int e1 = (info >> 48) & 7, e2 = (info >> 51) & 7;
    const long C = s.C;
    if (s.do_mob) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            double mu;
            if (K2 < 0) {
                bool inflow = false;
#pragma unroll
                for (int j = 0; j < N; j++) inflow |= s.v[((long)j * FPC + e1) * C + K1] < 0;
                mu = inflow ? s.bc_mobility[(long)i * s.F + E] : s.m[(long)i * C + K1];
            } else {
                const double vel = s.v[((long)i * FPC + e1) * C + K1] - s.v[((long)i * FPC + e2) * C + K2];
                if (vel == 0) mu = 0.5 * (s.m[(long)i * C + K1] + s.m[(long)i * C + K2]);
                else if (vel > 0) mu = s.m[(long)i * C + K1];
                else mu = s.m[(long)i * C + K2];
            }
            s.m_upw[(long)E * N + i] = mu;
        }
    }
    if (s.advection == MHFEM_EXPLICIT_UPWIND) {
#pragma unroll
        for (int j = 0; j < N; j++)
#pragma unroll
            for (int i = 0; i < N; i++) {
                const long idx = (((long)i * N + j) * FPC + e1) * C + K1;
                const double apu = s.a[idx] + s.u[idx];
                double z;
                if (apu > 0.0) z = s.Z_iK[(long)j * C + K1];
                else if (apu == 0.0) z = 0;
                else if (K2 >= 0) z = s.Z_iK[(long)j * C + K2];
                else z = s.Z_iF[(long)E * N + j];
                s.Z_upw[((long)E * N + j) * N + i] = z;
            }
    }
}

// ---------------------------------------------------------------------------
// Per-cell part of preIterate, fused: K2 local mass matrix and its inverse in
// registers (Solver.hpp:627-639), K6 R_ijKe / R_iK (Solver.hpp:793-815,
// SecondaryCoefficients.h:50-127) and K7 Q assembly, singular-row fix, LU and
// in-place Q^{-1} solves (Solver.hpp:817-871, SecondaryCoefficients.h:129-159).
// The model hooks updateNonLinearTerms / updateVectorCoefficients run on the
// host before this kernel; K4/K5 do not depend on b, so they run first.
// The work of one cell; its record goes to rec (an LDS slot or global memory).
template <int N, int FPC>
__device__ __forceinline__ void cell_step(const SchemeView &s, double tau, int K, double *rec)
{
    using Rec = CellRec<N, FPC>;
    const long C = s.C;
    int fi[FPC];
#pragma unroll
    for (int e = 0; e < FPC; e++) fi[e] = s.cell_faces[(long)K * FPC + e];
    const double meas = s.cell_meas[K];
    double minv[FPC][FPC];
    load_minv<FPC>(s, K, minv);
    double Dij[N][N];
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) Dij[i][j] = s.D[((long)i * N + j) * C + K];
    double mu[N][FPC];
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int e = 0; e < FPC; e++) mu[i][e] = s.m_upw[(long)fi[e] * N + i];

    auto vel_of = [&](int i, int j, int e) {
        const long idx = (((long)i * N + j) * FPC + e) * C + K;
        return s.a[idx] + s.u[idx];
    };

    // K6: R_iK (SecondaryCoefficients.h:88-127)
    double RK[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        double r = 0;
#pragma unroll
        for (int j = 0; j < N; j++) r += s.N[((long)i * N + j) * C + K] * s.Z_iK[(long)j * C + K];
        r *= meas / tau;
        r += meas * s.f[(long)i * C + K];
#pragma unroll
        for (int e = 0; e < FPC; e++) r -= mu[i][e] * s.w[((long)i * FPC + e) * C + K];
        if (s.advection == MHFEM_EXPLICIT_UPWIND) {
            double aux = 0;
#pragma unroll
            for (int j = 0; j < N; j++)
#pragma unroll
                for (int e = 0; e < FPC; e++)
                    aux += vel_of(i, j, e) * s.Z_upw[((long)fi[e] * N + j) * N + i];
            r -= aux;
        }
        RK[i] = r;
    }

    // K7: Q (SecondaryCoefficients.h:129-159) with the row sums of b_ijK,
    // singular rows (Solver.hpp:829-847)
    double Q[N][N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        bool singular = true;
#pragma unroll
        for (int j = 0; j < N; j++) {
            double q = 0;
#pragma unroll
            for (int e = 0; e < FPC; e++) {
                q += mu[i][e] * row_sum<FPC>(minv[e], Dij[i][j]) - s.u[(((long)i * N + j) * FPC + e) * C + K];
                if (s.advection == MHFEM_IMPLICIT_UPWIND) {
                    const double vel = vel_of(i, j, e);
                    if (vel >= 0) q += vel;
                    else q -= vel;
                }
            }
            q += meas / tau * s.N[((long)i * N + j) * C + K] + meas * s.r[((long)i * N + j) * C + K];
            Q[i][j] = q;
            if (q != 0.0) singular = false;
        }
        if (singular) {
            Q[i][i] = 1.0;
            RK[i] += s.Z_iK[(long)i * C + K];
        }
    }
    lu_factorize<N>(Q);
    lu_solve<N>(Q, RK);

    // K6: R_ijKe (SecondaryCoefficients.h:50-86), one column (j, e) at a
    // time, then Q^{-1} applied to it; written to the cell's record only
    // (the reference-layout arrays are formed on request, k_fill_R)
#pragma unroll
    for (int j = 0; j < N; j++)
#pragma unroll
        for (int e = 0; e < FPC; e++) {
            double x[N];
#pragma unroll
            for (int i = 0; i < N; i++) {
                double r = 0;
#pragma unroll
                for (int f = 0; f < FPC; f++) r += mu[i][f] * b_entry(minv[f][e], Dij[i][j]);
                if (s.advection == MHFEM_IMPLICIT_UPWIND) {
                    const double vel = vel_of(i, j, e);
                    if (vel < 0) r -= 2 * vel;
                }
                if (s.advection == MHFEM_IMPLICIT_TRACE) r -= vel_of(i, j, e);
                x[i] = r;
            }
            lu_solve<N>(Q, x);
#pragma unroll
            for (int i = 0; i < N; i++) rec[(j * FPC + e) * N + i] = x[i];
        }
#pragma unroll
    for (int i = 0; i < N; i++) rec[Rec::RK_OFF + i] = RK[i];
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) rec[Rec::D_OFF + i * N + j] = Dij[i][j];
    if constexpr (Rec::SIZE > Rec::D_OFF + N * N) rec[Rec::D_OFF + N * N] = 0.0;
}

template <int N, int FPC>
__global__ void __launch_bounds__(WG) k_cell(SchemeView s, double tau)
{
    using Rec = CellRec<N, FPC>;
    const int first = blockIdx.x * WG, K = first + threadIdx.x;
    if constexpr (Rec::STAGE) {
        __shared__ double stage[WG * Rec::LS];
        if (K < s.C) cell_step<N, FPC>(s, tau, K, stage + threadIdx.x * Rec::LS);
        __syncthreads();
        records_to_global<Rec::SIZE, Rec::LS>(s.cell_rec + (long)first * Rec::SIZE, stage, min(WG, s.C - first));
    } else {
        if (K < s.C) cell_step<N, FPC>(s, tau, K, s.cell_rec + (long)K * Rec::SIZE);
    }
}

// R_ijKe ((i*n+j)*FPC + e)*C + K and R_iK i*C + K in the reference's layout,
// from the cell records; written only when a caller asks for the arrays.
template <int N, int FPC>
__global__ void __launch_bounds__(WG) k_fill_R(SchemeView s, double *R_ijKe, double *R_iK)
{
    using Rec = CellRec<N, FPC>;
    const int K = blockIdx.x * WG + threadIdx.x;
    if (K >= s.C) return;
    const long C = s.C;
    const double *rec = s.cell_rec + (long)K * Rec::SIZE;
#pragma unroll
    for (int x = 0; x < N; x++)
#pragma unroll
        for (int j = 0; j < N; j++)
#pragma unroll
            for (int f = 0; f < FPC; f++) R_ijKe[(((long)x * N + j) * FPC + f) * C + K] = rec[(j * FPC + f) * N + x];
#pragma unroll
    for (int j = 0; j < N; j++) R_iK[(long)j * C + K] = rec[Rec::RK_OFF + j];
}

// ---------------------------------------------------------------------------
// Face-parallel hybridised assembly (Solver.hpp:933-959).
//
// CellRow<N, FPC>::compute evaluates, for one cell K, its local face e (the
// face E whose rows are assembled) and one row i, the cell's contribution to
// row (E, i):
//   A[j][f] = A_ijKEF against unknown j of local face f,  Ae[j] = A[j][e],
//   rhs     = RHS_iKE.
// Every register array is indexed by compile-time constants (the runtime
// local indices and the row only enter load and store addresses), so nothing
// is demoted to scratch memory.  Ae is selected while the values are
// computed: a select over an array after the fact is folded by the compiler
// into a load from a selected address, which keeps the whole array in
// scratch memory.  The entries of b_ijK are formed from row e of M^{-1} and
// D_ij (b_entry, row_sum), rounded as the stored values would be.

// Values of cell K that every row of its local face e uses: R_xjKf, R_jK
// (from the cell's record) and row e of M^{-1}.
template <int N, int FPC>
struct CellData {
    double Rv[N][N][FPC], RK[N], mrow[FPC];
    __device__ __forceinline__ void load(const SchemeView &s, int K, int e)
    {
        using Rec = CellRec<N, FPC>;
        const double *rec = s.cell_rec + (long)K * Rec::SIZE;
        if constexpr (Rec::RK_OFF % 2 == 0) {
#pragma unroll
            for (int q = 0; q < Rec::RK_OFF; q += 2) {
                const Pack<double, 2> v = ld<double, 2>(rec + q);
                Rv[q % N][q / (N * FPC)][(q / N) % FPC] = v.v[0];
                Rv[(q + 1) % N][(q + 1) / (N * FPC)][((q + 1) / N) % FPC] = v.v[1];
            }
        } else {
#pragma unroll
            for (int q = 0; q < Rec::RK_OFF; q++) Rv[q % N][q / (N * FPC)][(q / N) % FPC] = rec[q];
        }
#pragma unroll
        for (int j = 0; j < N; j++) RK[j] = rec[Rec::RK_OFF + j];
        load_minv_row<FPC>(s, K, e, mrow);
    }
};

template <int N, int FPC>
struct CellRow {
    double A[N][FPC];
    double Ae[N];
    double Aii;  // A[i][e], the diagonal entry before scaling
    double rhs;

    // Row inputs: Di[j] = D_ijK (from the cell's record), wv = w_iKe.
    __device__ __forceinline__ static void load_row(const SchemeView &s, int K, int e, int i, double (&Di)[N], double &wv)
    {
        const double *rec = s.cell_rec + (long)K * CellRec<N, FPC>::SIZE + CellRec<N, FPC>::D_OFF + i * N;
#pragma unroll
        for (int j = 0; j < N; j++) Di[j] = rec[j];
        wv = s.w[((long)i * FPC + e) * s.C + K];
    }

    // use_adv: A_ijKEF_advection / RHS_iKE_advection (SecondaryCoefficients.h:215-259,
    // 301-333); otherwise the no-advection forms (:189-211, 282-297)
    __device__ __forceinline__ void compute(const SchemeView &s, const CellData<N, FPC> &cd, const double (&Di)[N],
                                            double wv, int adv, bool use_adv, int K, int e, int E, int i)
    {
        const long C = s.C;
        double bf[N][FPC], be[N];
#pragma unroll
        for (int j = 0; j < N; j++) {
#pragma unroll
            for (int f = 0; f < FPC; f++) bf[j][f] = b_entry(cd.mrow[f], Di[j]);
            be[j] = row_sum<FPC>(cd.mrow, Di[j]);
        }
        if (!use_adv) {
#pragma unroll
            for (int j = 0; j < N; j++)
#pragma unroll
                for (int f = 0; f < FPC; f++) {
                    double SQinvR = 0;
#pragma unroll
                    for (int x = 0; x < N; x++) SQinvR += be[x] * cd.Rv[x][j][f];
                    A[j][f] = bf[j][f] - SQinvR;
                    if (f == e) Ae[j] = A[j][f];
                    if (f == e && j == i) Aii = A[j][f];
                }
            double v = wv;
#pragma unroll
            for (int j = 0; j < N; j++) v += be[j] * cd.RK[j];
            rhs = v;
        } else {
            const double mu = s.m_upw[(long)E * N + i];
            double vel[N], Sx[N];
#pragma unroll
            for (int j = 0; j < N; j++) {
                const long idx = (((long)i * N + j) * FPC + e) * C + K;
                vel[j] = s.a[idx] + s.u[idx];
            }
#pragma unroll
            for (int x = 0; x < N; x++) {
                double Sv = mu * be[x];
                if (adv == MHFEM_IMPLICIT_UPWIND) {
                    if (vel[x] >= 0) Sv += vel[x];
                    else Sv -= vel[x];
                }
                Sx[x] = Sv;
            }
#pragma unroll
            for (int j = 0; j < N; j++)
#pragma unroll
                for (int f = 0; f < FPC; f++) {
                    double SQinvR = 0;
#pragma unroll
                    for (int x = 0; x < N; x++) SQinvR += Sx[x] * cd.Rv[x][j][f];
                    double T = mu * bf[j][f];
                    if (f == e && adv == MHFEM_IMPLICIT_UPWIND && vel[j] < 0) T -= 2 * vel[j];
                    if (f == e && adv == MHFEM_IMPLICIT_TRACE) T -= vel[j];
                    A[j][f] = T - SQinvR;
                    if (f == e) Ae[j] = A[j][f];
                    if (f == e && j == i) Aii = A[j][f];
                }
            double v = mu * wv;
#pragma unroll
            for (int j = 0; j < N; j++) {
                v += Sx[j] * cd.RK[j];
                if (adv == MHFEM_EXPLICIT_UPWIND) v += s.Z_upw[((long)E * N + j) * N + i] * vel[j];
            }
            rhs = v;
        }
    }
};

template <int N>
__device__ __forceinline__ void store_block_row(const SchemeView &s, int E, int k, int i, const double (&v)[N])
{
    Pack<double, N> p;
#pragma unroll
    for (int j = 0; j < N; j++) p.v[j] = v[j];
    st<double, N>(s.A.val + sell_block(s.A.slice_ptr, E, k) * (N * N) + i * N, p);
}

// The assembly's per-face store map (built on the host): a 32-bit word for up
// to four faces per cell (block-index fields of 3 bits, K1's at bit 12, the
// local indices of the face at bits 24 and 27), a 64-bit word for six
// (fields of 4 bits, K1's at bit 24, local indices at bits 48 and 51).  The
// top bit is always clear.
template <int FPC>
struct FaceMap {
    using T = uint32_t;
    static constexpr int FB = 3, CB = 12, EB = 24;
};
template <>
struct FaceMap<6> {
    using T = uint64_t;
    static constexpr int FB = 4, CB = 24, EB = 48;
};
template <int FPC>
__device__ __forceinline__ typename FaceMap<FPC>::T face_map(const SchemeView &s, int E)
{
    return reinterpret_cast<const typename FaceMap<FPC>::T *>(s.face_blocks)[E];
}
// block index of local face f of cell c (0: K0, 1: K1) in the SELL row of the face
template <int FPC>
__device__ __forceinline__ int map_block(typename FaceMap<FPC>::T map, int c, int f)
{
    using M = FaceMap<FPC>;
    return (int)(map >> (M::CB * c + M::FB * f)) & ((1 << M::FB) - 1);
}
// local index of the face in cell c
template <int FPC>
__device__ __forceinline__ int map_local(typename FaceMap<FPC>::T map, int c)
{
    return (int)(map >> (FaceMap<FPC>::EB + 3 * c)) & 7;
}
// the top bit of the map as a sign bit (always clear; read so that the exit
// test waits for the map load, see k_assemble)
template <int FPC>
__device__ __forceinline__ int map_top(typename FaceMap<FPC>::T map)
{
    return (int)((uint32_t)(map >> (8 * sizeof(typename FaceMap<FPC>::T) - 32)) & 0x80000000u);
}

// Boundary row (E, i): BoundaryConditions::setMatrixElements
// (BoundaryConditions.hpp:335-508), one lane per row.  The row's tag selects
// the form of its cell's contribution (BoundaryCoefficients_*,
// BoundaryConditions.hpp:11-130): FixedFlux the advective one, the others
// the plain one; an inflowing AdvectiveOutflow row is the identity with the
// cell average as right-hand side; a FixedValue row is the (implicit) unit
// diagonal alone (:373-380).
template <int N, int FPC>
__device__ __forceinline__ void boundary_row(const SchemeView &s, int q, int i)
{
    const int E = s.bnd_faces[q];
    const long bi = (long)i * s.F + E;
    const int tag = s.tags[bi];
    if (tag == MHFEM_BC_FIXED_VALUE) {
        // BoundaryModel::getDirichletValue at time + tau (BoundaryConditions.hpp:379)
        double value = s.bc_dirichlet[bi];
        if (s.bc_fn) value *= time_factor(s.bc_fn[i], s.clock[0]);
        s.rhs[(long)E * N + i] = value;
        return;
    }
    const int K0 = s.face_cells[2 * E];
    const auto map = face_map<FPC>(s, E);
    const int e0 = map_local<FPC>(map, 0);
    const int adv = s.advection;
    const long C = s.C;
    const bool use_adv = tag == MHFEM_BC_FIXED_FLUX;
    CellData<N, FPC> cd;
    cd.load(s, K0, e0);
    double Di[N], wv;
    CellRow<N, FPC>::load_row(s, K0, e0, i, Di, wv);
    CellRow<N, FPC> r;
    r.compute(s, cd, Di, wv, adv, use_adv, K0, e0, E, i);
    const long iidx = (((long)i * N + i) * FPC + e0) * C + K0;
    const bool inflow = tag == MHFEM_BC_ADVECTIVE_OUTFLOW && (s.a[iidx] + s.u[iidx]) < 0;
    // RowSetter::setRow (BoundaryConditions.hpp:132-222); unit diagonal not stored (:200-203)
    const double d = inflow ? 1.0 : r.Aii;
#pragma unroll
    for (int f = 0; f < FPC; f++) {
        if (f == e0) continue;
        double v[N];
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = (inflow ? 0.0 : r.A[j][f]) / d;
        store_block_row<N>(s, E, map_block<FPC>(map, 0, f), i, v);
    }
#pragma unroll
    for (int j = 0; j < N; j++)
        if (j != i) s.A.val[sell_diag(s.A.diag_base, N, E, i, j)] = (inflow ? 0.0 : r.Ae[j]) / d;
    double bval;
    if (tag == MHFEM_BC_ADVECTIVE_OUTFLOW) {
        bval = 0;
        bval += inflow ? s.Z_iK[(long)i * C + K0] : r.rhs;
    } else {
        // BoundaryModel::getNeumannValue at time + tau (BoundaryConditions.hpp:414, 455)
        double J = s.bc_values[bi];
        if (s.bc_fn) J *= time_factor(s.bc_fn[N + i], s.clock[0]);
        bval = -J * s.face_meas[E];
        bval += r.rhs;
    }
    s.rhs[(long)E * N + i] = bval / d;
}

// Interior faces: LinearSystem::setMatrixElements (LinearSystem.h:66-256)
// and RightHandSide::getValue (RightHandSide.h:18-45).  Every column face
// other than E belongs to exactly one of the two cells; the diagonal block
// and the right-hand side add both cells' contributions (K0 first).
//
// 2*NR lanes per face, NR = N rounded up to a power of two: lane
// (E, c, i) evaluates row i of cell K_c and stores that row of the blocks of
// K_c's faces, so the two rows of a block leave as one contiguous write.  The
// lanes of one row in the two cells swap their diagonal columns and
// right-hand sides with a lane shuffle; a sum of two terms is the same in
// either order, so both form the identical scale d.  Few registers per lane
// keep many waves in flight to cover the gathers of the cell data.
// The face map gives the block index of every local face of K0 and K1 in the
// SELL row of E, and the local index of E in K0 and K1.
template <int N, int FPC>
__global__ void __launch_bounds__(WG) k_assemble(SchemeView s)
{
    constexpr int NR = N == 3 ? 4 : N;
    // work-groups past the interior ones take the boundary rows, NR lanes
    // per boundary face: one launch, and the few boundary rows run beside
    // the interior ones
    const unsigned interior_blocks = (unsigned)(((long)s.F * 2 * NR + WG - 1) / WG);
    if (blockIdx.x >= interior_blocks) {
        const long tb = (long)(blockIdx.x - interior_blocks) * WG + threadIdx.x;
        const int q = (int)(tb / NR), ib = (int)(tb % NR);
        if (q < s.nbnd && ib < N) boundary_row<N, FPC>(s, q, ib);
        return;
    }
    const long t = (long)blockIdx.x * WG + threadIdx.x;
    const int E = (int)(t / (2 * NR)), c = (int)(t / NR) & 1, i = (int)(t % NR);
    if (E >= s.F || i >= N) return;  // the partner lane (same E and i) returns too
    // Both cells (one 8-byte load) and the store map in one round trip: the
    // exit test reads the map too (its bit 31 is always clear), so neither
    // load waits behind the branch.
    const uint64_t cells = reinterpret_cast<const uint64_t *>(s.face_cells)[E];
    const auto map = face_map<FPC>(s, E);
    const int K0 = (int)(uint32_t)cells, K1 = (int)(uint32_t)(cells >> 32);
    if ((K1 | map_top<FPC>(map)) < 0) return;  // boundary faces: the boundary rows
    const int K = c ? K1 : K0;
    const int e = map_local<FPC>(map, c);
    const int adv = s.advection;
    const bool use_adv = adv != MHFEM_EXPLICIT_UPWIND;  // SecondaryCoefficients.h:261-278
    CellData<N, FPC> cd;
    cd.load(s, K, e);
    double Di[N], wv;
    CellRow<N, FPC>::load_row(s, K, e, i, Di, wv);
    CellRow<N, FPC> r;
    r.compute(s, cd, Di, wv, adv, use_adv, K, e, E, i);
    double other[N];
#pragma unroll
    for (int j = 0; j < N; j++) other[j] = __shfl_xor(r.Ae[j], NR);
    const double other_rhs = __shfl_xor(r.rhs, NR);
    const double d = r.Aii + __shfl_xor(r.Aii, NR);
#pragma unroll
    for (int f = 0; f < FPC; f++) {
        if (f == e) continue;
        double v[N];
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = r.A[j][f] / d;
        store_block_row<N>(s, E, map_block<FPC>(map, c, f), i, v);
    }
    if (c == 0) {
        // the diagonal entry itself is exactly 1 (LinearSystem.h:196-201) and is not stored
#pragma unroll
        for (int j = 0; j < N; j++)
            if (j != i) s.A.val[sell_diag(s.A.diag_base, N, E, i, j)] = (r.Ae[j] + other[j]) / d;
    } else {
        double rhs = 0.0;
        rhs += other_rhs;  // K0
        rhs += r.rhs;      // K1
        s.rhs[(long)E * N + i] = rhs / d;
    }
}

// ---------------------------------------------------------------------------
// postIterate: K9 Z_iK recovery (Solver.hpp:1283-1292, SecondaryCoefficients.h:163-185)
// and K10 conservative velocities (Solver.hpp:1298-1313, SecondaryCoefficients.h:16-48)
template <int N, int FPC>
__device__ __forceinline__ void post_step(const SchemeView &s, int K, const double *rec)
{
    using Rec = CellRec<N, FPC>;
    const long C = s.C;
    int fi[FPC];
#pragma unroll
    for (int e = 0; e < FPC; e++) fi[e] = s.cell_faces[(long)K * FPC + e];
    double ZF[N][FPC];
#pragma unroll
    for (int f = 0; f < FPC; f++)
#pragma unroll
        for (int j = 0; j < N; j++) ZF[j][f] = s.Z_iF[(long)fi[f] * N + j];
    double ZK[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        double Z = 0;
#pragma unroll
        for (int f = 0; f < FPC; f++)
#pragma unroll
            for (int j = 0; j < N; j++) Z += rec[(j * FPC + f) * N + i] * ZF[j][f];
        Z += rec[Rec::RK_OFF + i];
        ZK[i] = Z;
        s.Z_iK[(long)i * C + K] = Z;
    }
    if (!s.do_mob) return;
    double minv[FPC][FPC];
    load_minv<FPC>(s, K, minv);
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int e = 0; e < FPC; e++) {
            double sK = 0, sE = 0;
#pragma unroll
            for (int j = 0; j < N; j++) {
                const double D = rec[Rec::D_OFF + i * N + j];
                sK += row_sum<FPC>(minv[e], D) * ZK[j];
#pragma unroll
                for (int f = 0; f < FPC; f++) sE += b_entry(minv[e][f], D) * ZF[j][f];
            }
            s.v[((long)i * FPC + e) * C + K] = sK - sE + s.w[((long)i * FPC + e) * C + K];
        }
}

template <int N, int FPC>
__global__ void __launch_bounds__(WG) k_post(SchemeView s)
{
    using Rec = CellRec<N, FPC>;
    const int first = blockIdx.x * WG, K = first + threadIdx.x;
    if constexpr (Rec::STAGE) {
        __shared__ double stage[WG * Rec::LS];
        records_from_global<Rec::SIZE, Rec::LS>(s.cell_rec + (long)first * Rec::SIZE, stage, min(WG, s.C - first));
        __syncthreads();
        if (K < s.C) post_step<N, FPC>(s, K, stage + threadIdx.x * Rec::LS);
    } else {
        if (K < s.C) post_step<N, FPC>(s, K, s.cell_rec + (long)K * Rec::SIZE);
    }
}

__global__ void k_set_clock(double *clock, double t)
{
    if (threadIdx.x == 0) clock[0] = t;
}

__global__ void k_fill(double *p, size_t n, double v)
{
    const size_t i = (size_t)blockIdx.x * WG + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
static inline unsigned blocks(long n) { return (unsigned)((n + WG - 1) / WG); }

template <template <int, int> class KernelTag, typename... Args>
static void dispatch(Ctx &c, Args &&...args);

#define MHFEM_DISPATCH_BS(c, KERNEL, GRID, BS, ...)                                                 \
    do {                                                                                            \
        const unsigned g_ = (GRID);                                                                 \
        if (g_ == 0) break;                                                                         \
        if ((c).fpc == 3) {                                                                         \
            if ((c).n == 1) hipLaunchKernelGGL((KERNEL<1, 3>), g_, BS, 0, (c).stream, __VA_ARGS__); \
            else if ((c).n == 2) hipLaunchKernelGGL((KERNEL<2, 3>), g_, BS, 0, (c).stream, __VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<3, 3>), g_, BS, 0, (c).stream, __VA_ARGS__);           \
        } else if ((c).fpc == 2) {                                                                  \
            if ((c).n == 1) hipLaunchKernelGGL((KERNEL<1, 2>), g_, BS, 0, (c).stream, __VA_ARGS__); \
            else if ((c).n == 2) hipLaunchKernelGGL((KERNEL<2, 2>), g_, BS, 0, (c).stream, __VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<3, 2>), g_, BS, 0, (c).stream, __VA_ARGS__);           \
        } else if ((c).fpc == 6) {                                                                  \
            if ((c).n == 1) hipLaunchKernelGGL((KERNEL<1, 6>), g_, BS, 0, (c).stream, __VA_ARGS__); \
            else if ((c).n == 2) hipLaunchKernelGGL((KERNEL<2, 6>), g_, BS, 0, (c).stream, __VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<3, 6>), g_, BS, 0, (c).stream, __VA_ARGS__);           \
        } else {                                                                                    \
            if ((c).n == 1) hipLaunchKernelGGL((KERNEL<1, 4>), g_, BS, 0, (c).stream, __VA_ARGS__); \
            else if ((c).n == 2) hipLaunchKernelGGL((KERNEL<2, 4>), g_, BS, 0, (c).stream, __VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<3, 4>), g_, BS, 0, (c).stream, __VA_ARGS__);           \
        }                                                                                           \
        MHFEM_HIP(hipGetLastError());                                                               \
    } while (0)
#define MHFEM_DISPATCH(c, KERNEL, GRID, ...) MHFEM_DISPATCH_BS(c, KERNEL, GRID, WG, __VA_ARGS__)

void launch_face_average(Ctx &c)
{
    SchemeView s = c.view();
    if (c.F == 0) return;
    if (c.n == 1) hipLaunchKernelGGL(k_face_average<1>, blocks(c.F), WG, 0, c.stream, s);
    else if (c.n == 2) hipLaunchKernelGGL(k_face_average<2>, blocks(c.F), WG, 0, c.stream, s);
    else hipLaunchKernelGGL(k_face_average<3>, blocks(c.F), WG, 0, c.stream, s);
    MHFEM_HIP(hipGetLastError());
}

void launch_pre_iterate(Ctx &c, double time, double tau)
{
    (void)time;
    SchemeView s = c.view();
    if (c.timing) MHFEM_HIP(hipEventRecord(c.ev[0], c.stream));
    if (c.do_mob || c.advection == MHFEM_EXPLICIT_UPWIND) MHFEM_DISPATCH(c, k_upwind, blocks(c.F), s);
    if (c.timing) MHFEM_HIP(hipEventRecord(c.ev[1], c.stream));
    MHFEM_DISPATCH(c, k_cell, blocks(c.C), s, tau);
    if (c.timing) MHFEM_HIP(hipEventRecord(c.ev[2], c.stream));
}

void launch_assemble(Ctx &c, double time, double tau)
{
    (void)time;
    (void)tau;
    SchemeView s = c.view();
    // interior rows (2*NR lanes per face), then the boundary rows (NR lanes
    // per boundary face) in the same launch
    const long nr = c.n == 3 ? 4 : c.n;
    MHFEM_DISPATCH(c, k_assemble, blocks(2 * nr * c.F) + blocks(nr * c.nbnd), s);
}

void launch_geometry(Ctx &c)
{
    SchemeView s = c.view();
    if (c.box && c.fpc == 2) hipLaunchKernelGGL(k_geometry_box<2>, blocks(c.C), WG, 0, c.stream, s, c.minv.p, c.minv_rows.p);
    else if (c.box && c.fpc == 6) hipLaunchKernelGGL(k_geometry_box<6>, blocks(c.C), WG, 0, c.stream, s, c.minv.p, c.minv_rows.p);
    else if (c.box) hipLaunchKernelGGL(k_geometry_box<4>, blocks(c.C), WG, 0, c.stream, s, c.minv.p, c.minv_rows.p);
    else if (c.fpc == 3) hipLaunchKernelGGL(k_geometry<3>, blocks(c.C), WG, 0, c.stream, s, c.minv.p, c.minv_rows.p);
    else hipLaunchKernelGGL(k_geometry<4>, blocks(c.C), WG, 0, c.stream, s, c.minv.p, c.minv_rows.p);
    MHFEM_HIP(hipGetLastError());
}

void launch_fill_b(Ctx &c)
{
    SchemeView s = c.view();
    if (c.box) MHFEM_DISPATCH(c, k_fill_b_box, blocks(c.C), s, c.arr[MHFEM_b_ijK_storage].p);
    else MHFEM_DISPATCH(c, k_fill_b, blocks(c.C), s, c.arr[MHFEM_b_ijK_storage].p);
}

void launch_fill_R(Ctx &c)
{
    SchemeView s = c.view();
    MHFEM_DISPATCH(c, k_fill_R, blocks(c.C), s, c.arr[MHFEM_R_ijKe].p, c.arr[MHFEM_R_iK].p);
}

void launch_post_iterate(Ctx &c)
{
    SchemeView s = c.view();
    MHFEM_DISPATCH(c, k_post, blocks(c.C), s);
}

void launch_set_clock(Ctx &c, double t)
{
    hipLaunchKernelGGL(k_set_clock, 1, 64, 0, c.stream, c.clock.p, t);
    MHFEM_HIP(hipGetLastError());
}

void launch_zero_double(Ctx &c, double *p, size_t n)
{
    if (!n) return;
    hipLaunchKernelGGL(k_fill, (unsigned)((n + WG - 1) / WG), WG, 0, c.stream, p, n, 0.0);
    MHFEM_HIP(hipGetLastError());
}

}  // namespace mhfem
