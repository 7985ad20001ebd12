// capi.cpp -- the C ABI of libmhfem_amd.so: context lifetime, mesh topology,
// the fixed CSR pattern and its SELL-64 device layout, model-array transfers
// and the per-step entry points (include/mhfem.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <unordered_map>

#include "common.hpp"

using namespace mhfem;

// Staging of mhfem_solve_rhs_noalloc: the right-hand side and the iterate of
// one call, in a fixed device array of the library instead of buffers
// allocated per call.  Its name is a host-side stand-in: host code uses only
// the device address hipGetSymbolAddress returns (one copy per GPU).
static __device__ __align__(256) double scratch[2 * MHFEM_NOALLOC_MAX_ROWS];

namespace {

thread_local std::string g_create_error;

// ---------------------------------------------------------------------------
// Face numbering and incidence.  Faces are numbered in the order in which
// they are first met when the cells are visited in index order and the faces
// of each cell in local order; the cells of a face are kept in ascending
// order (mesh_helpers.h:49-59 returns them in that order: cellIndexes[0] is
// the "main" element of Solver.hpp:670-671).  Local face e of a simplex is
// the face opposite local vertex e (MassMatrix.h:538,661,787,932); the faces
// of an edge are its vertices, those of a quadrangle its edges (0,1), (1,2),
// (2,3), (3,0), those of a hexahedron (VTK vertex order) (0,1,2,3),
// (0,1,5,4), (1,2,6,5), (2,3,7,6), (3,0,4,7), (4,5,6,7) (TNL's subentity
// order, on which the face directions of MassMatrix.h:178-184, 430-436 rest).
int face_vertices(int dim, int box, int e, int *lv)
{
    static const int quad[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    static const int hex[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                  {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};
    if (!box) {
        int t = 0;
        for (int v = 0; v <= dim; v++)
            if (v != e) lv[t++] = v;
        return t;
    }
    if (dim == 1) {
        lv[0] = e;
        return 1;
    }
    if (dim == 2) {
        lv[0] = quad[e][0];
        lv[1] = quad[e][1];
        return 2;
    }
    for (int t = 0; t < 4; t++) lv[t] = hex[e][t];
    return 4;
}

struct FaceKey {
    int v[4];
    bool operator==(const FaceKey &o) const
    {
        return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2] && v[3] == o.v[3];
    }
};
struct FaceKeyHash {
    size_t operator()(const FaceKey &k) const
    {
        uint64_t h = (uint64_t)(uint32_t)k.v[0] * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)k.v[1] + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)k.v[2] * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)k.v[3] * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
        return (size_t)(h ^ (h >> 31));
    }
};

void build_topology(Ctx &c, const int *cv)
{
    const int fpc = c.fpc, C = c.C;
    c.h_cell_faces.assign((size_t)C * fpc, -1);
    c.h_face_cells.clear();
    c.h_face_cells.reserve((size_t)C * fpc);
    std::unordered_map<FaceKey, int, FaceKeyHash> seen;
    seen.reserve((size_t)C * fpc);
    int nf = 0;
    for (int K = 0; K < C; K++) {
        for (int v = 0; v < c.nvc; v++) {
            const int vid = cv[(size_t)K * c.nvc + v];
            if (vid < 0 || vid >= c.nv) throw Error{MHFEM_E_MESH, "cell vertex index out of range"};
        }
        for (int e = 0; e < fpc; e++) {
            FaceKey key{{-1, -1, -1, -1}};
            int lv[4];
            const int t = face_vertices(c.dim, c.box, e, lv);
            for (int q = 0; q < t; q++) key.v[q] = cv[(size_t)K * c.nvc + lv[q]];
            std::sort(key.v, key.v + t);
            auto it = seen.find(key);
            int id;
            if (it == seen.end()) {
                id = nf++;
                seen.emplace(key, id);
                c.h_face_cells.push_back(K);
                c.h_face_cells.push_back(-1);
            } else {
                id = it->second;
                if (c.h_face_cells[2 * (size_t)id + 1] != -1)
                    throw Error{MHFEM_E_MESH, "non-conforming mesh: a face belongs to more than two cells"};
                c.h_face_cells[2 * (size_t)id + 1] = K;
            }
            c.h_cell_faces[(size_t)K * fpc + e] = id;
        }
    }
    c.F = nf;
    c.nbnd = 0;
    for (int E = 0; E < nf; E++) c.nbnd += c.h_face_cells[2 * (size_t)E + 1] < 0;
}

double simplex_measure(int dim, const double *P)
{
    if (dim == 2) {
        const double ax = P[2] - P[0], ay = P[3] - P[1], bx = P[4] - P[0], by = P[5] - P[1];
        return 0.5 * std::fabs(ax * by - ay * bx);
    }
    double a[3], b[3], d[3];
    for (int k = 0; k < 3; k++) {
        a[k] = P[3 + k] - P[k];
        b[k] = P[6 + k] - P[k];
        d[k] = P[9 + k] - P[k];
    }
    const double det = a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                       a[2] * (b[0] * d[1] - b[1] * d[0]);
    return std::fabs(det) / 6.0;
}

// measure of a box cell: an edge's length; a quadrangle's area as half the
// cross product of its diagonals; a voxel's volume from vertices 0 and 6
double box_measure(int dim, const double *P)
{
    if (dim == 1) return std::fabs(P[1] - P[0]);
    if (dim == 3) return std::fabs((P[18] - P[0]) * (P[19] - P[1]) * (P[20] - P[2]));
    const double d1x = P[4] - P[0], d1y = P[5] - P[1], d2x = P[6] - P[2], d2y = P[7] - P[3];
    return 0.5 * std::fabs(d1x * d2y - d1y * d2x);
}

// a hexahedron's face (a planar quadrangle in 3-D): half the cross product of its diagonals
double quad_facet_measure(const double *P)
{
    double a[3], b[3];
    for (int k = 0; k < 3; k++) {
        a[k] = P[6 + k] - P[k];
        b[k] = P[9 + k] - P[3 + k];
    }
    const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
    return 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
}

double facet_measure(int dim, const double *P)
{
    if (dim == 1) return 1.0;  // a point: boundary fluxes are not scaled
    if (dim == 2) return std::hypot(P[2] - P[0], P[3] - P[1]);
    double a[3], b[3];
    for (int k = 0; k < 3; k++) {
        a[k] = P[3 + k] - P[k];
        b[k] = P[6 + k] - P[k];
    }
    const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
    return 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
}

// local faces of K sorted by global face index (LinearSystem.h:108-136)
void sorted_local(const Ctx &c, int K, int *perm)
{
    const int *fi = &c.h_cell_faces[(size_t)K * c.fpc];
    for (int j = 0; j < c.fpc; j++) perm[j] = j;
    std::sort(perm, perm + c.fpc, [fi](int a, int b) { return fi[a] < fi[b]; });
}

int local_index(const Ctx &c, int K, int E)
{
    for (int e = 0; e < c.fpc; e++)
        if (c.h_cell_faces[(size_t)K * c.fpc + e] == E) return e;
    return 0;
}

// Per-face column order of the hybridised rows: slot t names the local face
// of K0 (f0[t]) and of K1 (f1[t]), 7 = absent, and e0 / e1 are the local
// indices of the face itself in K0 / K1.  Interior rows merge the two sorted
// face lists (LinearSystem.h:181-245), boundary rows list the sorted faces of
// the one cell (BoundaryConditions.hpp:196-218).
struct FaceSlots {
    int8_t f0[12], f1[12];
    int ns, e0, e1;
};

std::vector<FaceSlots> build_face_slots(const Ctx &c)
{
    std::vector<FaceSlots> slots(c.F);
    const int fpc = c.fpc;
    for (int E = 0; E < c.F; E++) {
        const int K0 = c.h_face_cells[2 * (size_t)E], K1 = c.h_face_cells[2 * (size_t)E + 1];
        FaceSlots &fs = slots[E];
        fs.ns = 0;
        fs.e1 = 0;
        int p0[6], p1[6];
        sorted_local(c, K0, p0);
        const int *f0 = &c.h_cell_faces[(size_t)K0 * fpc];
        auto put = [&](int a, int b) {
            fs.f0[fs.ns] = (int8_t)a;
            fs.f1[fs.ns] = (int8_t)b;
            fs.ns++;
        };
        fs.e0 = local_index(c, K0, E);
        if (K1 < 0) {
            for (int g = 0; g < fpc; g++) put(p0[g], 7);
        } else {
            sorted_local(c, K1, p1);
            const int *f1 = &c.h_cell_faces[(size_t)K1 * fpc];
            int g0 = 0, g1 = 0;
            while (g0 < fpc || g1 < fpc) {
                const int F0 = g0 < fpc ? f0[p0[g0]] : INT32_MAX, F1 = g1 < fpc ? f1[p1[g1]] : INT32_MAX;
                const int Fm = std::min(F0, F1);
                put(F0 == Fm ? p0[g0] : 7, F1 == Fm ? p1[g1] : 7);
                if (F0 == Fm) g0++;
                if (F1 == Fm) g1++;
            }
            fs.e1 = local_index(c, K1, E);
        }
    }
    return slots;
}

// The per-face word the upwind kernel reads: the local index of the face in
// K0 (bits 48-50) and K1 (bits 51-53).
std::vector<uint64_t> face_slot_words(const std::vector<FaceSlots> &slots)
{
    std::vector<uint64_t> out(slots.size());
    for (size_t E = 0; E < slots.size(); E++) out[E] = (uint64_t)slots[E].e0 << 48 | (uint64_t)slots[E].e1 << 51;
    return out;
}

// Store map of the assembly, per face E: the block index, in the SELL row of
// E, of the column face at local index f of K0 and of K1 (the slots of the
// merged column order, counted without the slot of E itself, which has no
// stored block; its field is 0, unused), and the local index of E in K0 and
// K1 (0 on the boundary).  Up to four faces per cell, one 32-bit word: fields
// of 3 bits at 3f (K0) and 12+3f (K1), local indices at bits 24 and 27.  Six
// faces (hexahedra), one 64-bit word: fields of 4 bits at 4f and 24+4f, local
// indices at bits 48 and 51.  The top bit is always clear.
std::vector<uint32_t> build_face_blocks(const std::vector<FaceSlots> &slots, int fpc)
{
    const bool wide = fpc > 4;
    const int fb = wide ? 4 : 3, cb = wide ? 24 : 12, eb = wide ? 48 : 24;
    std::vector<uint32_t> out(slots.size() * (wide ? 2 : 1));
    for (size_t E = 0; E < slots.size(); E++) {
        const FaceSlots &fs = slots[E];
        uint64_t w = (uint64_t)fs.e0 << eb | (uint64_t)fs.e1 << (eb + 3);
        int k = 0;
        for (int t = 0; t < fs.ns; t++) {
            const int f0 = fs.f0[t], f1 = fs.f1[t];
            if (f0 == fs.e0) continue;
            if (f0 < fpc) w |= (uint64_t)k << (fb * f0);
            if (f1 < fpc) w |= (uint64_t)k << (cb + fb * f1);
            k++;
        }
        if (wide) std::memcpy(&out[2 * E], &w, 8);
        else out[E] = (uint32_t)w;
    }
    return out;
}

template <typename T>
void upload(DBuf<T> &d, const T *h, size_t n)
{
    d.alloc(n);
    if (n) MHFEM_HIP(hipMemcpy(d.p, h, n * sizeof(T), hipMemcpyHostToDevice));
}

bool is_coefficient(int which) { return which >= MHFEM_N_ijK && which <= MHFEM_f_iK; }

template <typename F>
int guarded(mhfem_ctx *ctx, F &&f)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return MHFEM_E_ARGUMENT;
    try {
        f(*c);
        return MHFEM_OK;
    } catch (const Error &e) {
        c->err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        c->err = e.what();
        return MHFEM_E_STATE;
    }
}

void phase_begin(Ctx &c, int slot)
{
    if (c.timing) MHFEM_HIP(hipEventRecord(c.ev[slot], c.stream));
}
void phase_end(Ctx &c, int slot, double &acc)
{
    if (!c.timing) return;
    MHFEM_HIP(hipEventRecord(c.ev[slot + 1], c.stream));
    MHFEM_HIP(hipEventSynchronize(c.ev[slot + 1]));
    float ms = 0;
    MHFEM_HIP(hipEventElapsedTime(&ms, c.ev[slot], c.ev[slot + 1]));
    acc += ms;
}

}  // namespace

// message of mhfem_last_error(NULL) for calls without a context (io.cpp)
void mhfem::set_global_error(const std::string &msg) { g_create_error = msg; }

namespace mhfem {

long Ctx::array_size(int which) const
{
    const long n_ = n, C_ = C, F_ = F, f = fpc;
    switch (which) {
    case MHFEM_Z_iF: return n_ * F_;
    case MHFEM_Z_iK: return n_ * C_;
    case MHFEM_N_ijK: case MHFEM_D_ijK: case MHFEM_r_ijK: return n_ * n_ * C_;
    case MHFEM_u_ijKe: case MHFEM_a_ijKe: case MHFEM_R_ijKe: return n_ * n_ * f * C_;
    case MHFEM_m_iK: case MHFEM_f_iK: case MHFEM_R_iK: return n_ * C_;
    case MHFEM_w_iKe: case MHFEM_v_iKe: return n_ * f * C_;
    case MHFEM_m_iE_upw: return n_ * F_;
    case MHFEM_Z_ijE_upw: return n_ * n_ * F_;
    case MHFEM_b_ijK_storage: return n_ * n_ * S * C_;
    case MHFEM_BC_VALUES: case MHFEM_BC_DIRICHLET: case MHFEM_BC_MOBILITY: return n_ * F_;
    case MHFEM_RHS: return rows;
    case MHFEM_CELL_MEASURE: return C_;
    case MHFEM_FACE_MEASURE: return F_;
    default: return -1;
    }
}

SchemeView Ctx::view()
{
    SchemeView s{};
    s.C = C;
    s.F = F;
    s.dim = dim;
    s.fpc = fpc;
    s.S = S;
    s.lumping = lumping;
    s.advection = advection;
    s.do_mob = do_mob;
    s.cell_faces = cell_faces.p;
    s.cell_vertices = cell_vertices.p;
    s.points = points.p;
    s.face_cells = face_cells.p;
    s.face_slots = face_slots.p;
    s.face_blocks = face_blocks.p;
    s.bnd_faces = bnd_faces.p;
    s.nbnd = nbnd;
    s.cell_meas = arr[MHFEM_CELL_MEASURE].p;
    s.face_meas = arr[MHFEM_FACE_MEASURE].p;
    auto coef = [&](int w) { return Coef{arr[w].p, uniform[w] ? 0L : ~0L}; };
    s.N = coef(MHFEM_N_ijK);
    s.u = coef(MHFEM_u_ijKe);
    s.m = coef(MHFEM_m_iK);
    s.D = coef(MHFEM_D_ijK);
    s.w = coef(MHFEM_w_iKe);
    s.a = coef(MHFEM_a_ijKe);
    s.r = coef(MHFEM_r_ijK);
    s.f = coef(MHFEM_f_iK);
    s.Z_iF = arr[MHFEM_Z_iF].p;
    s.Z_iK = arr[MHFEM_Z_iK].p;
    s.v = arr[MHFEM_v_iKe].p;
    s.m_upw = arr[MHFEM_m_iE_upw].p;
    s.Z_upw = arr[MHFEM_Z_ijE_upw].p;
    s.minv = minv.p;
    s.minv_rows = minv_rows.p;
    s.cell_rec = cell_rec.p;
    s.tags = tags.p;
    s.bc_values = arr[MHFEM_BC_VALUES].p;
    s.bc_dirichlet = arr[MHFEM_BC_DIRICHLET].p;
    s.bc_mobility = arr[MHFEM_BC_MOBILITY].p;
    s.bc_fn = bc_timed ? bc_fn.p : nullptr;
    s.clock = clock.p;
    s.A = SellView<double>{F, slice_ptr.p, slice_width.p, sell_col.p, sell_val.p, diag_base};
    s.rhs = arr[MHFEM_RHS].p;
    return s;
}

}  // namespace mhfem

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double *g_scratch[NOALLOC_MAX_DEVICES] = {};

// the device address of `scratch` on `device`: looked up on the first call
// on that device, with that device current, then kept
double *scratch_on(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw Error{MHFEM_E_ARGUMENT, "device index out of range"};
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        int current = 0;
        MHFEM_HIP(hipGetDevice(&current));
        if (current != device) MHFEM_HIP(hipSetDevice(device));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (current != device) (void)hipSetDevice(current);
        if (e != hipSuccess)
            throw Error{MHFEM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        g_scratch[device] = static_cast<double *>(ptr);
    }
    return g_scratch[device];
}

}  // namespace

extern "C" {

mhfem_ctx *mhfem_create(const mhfem_mesh_desc *mesh, const mhfem_model_desc *model)
{
    if (mesh && (mesh->dim == 2 || mesh->dim == 3)) return mhfem_create_cells(mesh, mesh->dim + 1, model);
    g_create_error = mesh ? "mhfem_create: dim must be 2 or 3" : "mhfem_create: null argument";
    return nullptr;
}

mhfem_ctx *mhfem_create_cells(const mhfem_mesh_desc *mesh, int vertices_per_cell, const mhfem_model_desc *model)
{
    Ctx *c = nullptr;
    try {
        if (!mesh || !model || !mesh->points || !mesh->cell_vertices)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_create: null argument"};
        const int dim = mesh->dim, vpc = vertices_per_cell;
        const bool simplex = (dim == 2 && vpc == 3) || (dim == 3 && vpc == 4);
        const bool box = (dim == 1 && vpc == 2) || (dim == 2 && vpc == 4) || (dim == 3 && vpc == 8);
        if (!simplex && !box)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_create: cells of " + std::to_string(vpc) + " vertices in " +
                                              std::to_string(dim) + "-D are not supported"};
        if (model->num_equations < 1 || model->num_equations > 3)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_create: num_equations must be 1..3"};
        if (mesh->num_cells < 1 || mesh->num_points < vpc)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_create: empty mesh"};
        if (model->advection < 0 || model->advection > 2)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_create: unknown advection discretization"};
        c = new Ctx();
        c->dim = dim;
        c->box = box ? 1 : 0;
        c->nvc = vpc;
        c->fpc = box ? 2 * dim : dim + 1;
        // MassMatrix::size: packed upper triangle + row sums for simplices,
        // one value per face direction for boxes (MassMatrix.h:40, 101, 160, 224)
        c->S = box ? dim : dim == 2 ? 9 : 14;
        c->n = model->num_equations;
        c->nv = mesh->num_points;
        c->C = mesh->num_cells;
        c->lumping = model->mass_lumping ? 1 : 0;
        c->advection = model->advection;
        c->do_mob = model->do_mobility_upwind ? 1 : 0;
        MHFEM_HIP(hipGetDevice(&c->device));
        MHFEM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto &e : c->ev) MHFEM_HIP(hipEventCreate(&e));
        build_topology(*c, mesh->cell_vertices);
        c->h_seg_cells = {0, c->C};
        c->h_seg_faces = {0, c->F};
        const size_t Cn = c->C, fpc = c->fpc, nvc = c->nvc;
        // measures
        std::vector<double> cm(Cn), fm(c->F, 0.0);
        c->h_face_pos.assign(3 * (size_t)c->F, 0.0f);
        for (size_t K = 0; K < Cn; K++) {
            double P[24], lowc[3] = {0, 0, 0};
            for (size_t k = 0; k < nvc; k++)
                for (int d = 0; d < c->dim; d++) {
                    P[k * c->dim + d] = mesh->points[(size_t)mesh->cell_vertices[K * nvc + k] * c->dim + d];
                    lowc[d] = k == 0 ? P[d] : std::min(lowc[d], P[k * c->dim + d]);
                }
            cm[K] = c->box ? box_measure(c->dim, P) : simplex_measure(c->dim, P);
            for (size_t e = 0; e < fpc; e++) {
                const int E = c->h_cell_faces[K * fpc + e];
                if (c->h_face_cells[2 * (size_t)E] != (int)K) continue;
                for (int d = 0; d < 3; d++) c->h_face_pos[3 * (size_t)E + d] = (float)lowc[d];
                double Q[12];
                int lv[4];
                const int t = face_vertices(c->dim, c->box, (int)e, lv);
                for (int q = 0; q < t; q++)
                    for (int d = 0; d < c->dim; d++) Q[q * c->dim + d] = P[lv[q] * c->dim + d];
                fm[E] = t == 4 ? quad_facet_measure(Q) : facet_measure(c->dim, Q);
            }
        }
        upload(c->cell_faces, c->h_cell_faces.data(), c->h_cell_faces.size());
        upload(c->face_cells, c->h_face_cells.data(), c->h_face_cells.size());
        upload(c->cell_vertices, mesh->cell_vertices, Cn * nvc);
        upload(c->points, mesh->points, (size_t)c->nv * c->dim);
        const std::vector<FaceSlots> slots = build_face_slots(*c);
        const std::vector<uint64_t> words = face_slot_words(slots);
        upload(c->face_slots, words.data(), words.size());
        const std::vector<uint32_t> fblocks = build_face_blocks(slots, c->fpc);
        upload(c->face_blocks, fblocks.data(), fblocks.size());
        std::vector<int> bfaces;
        for (int E = 0; E < c->F; E++)
            if (c->h_face_cells[2 * (size_t)E + 1] < 0) bfaces.push_back(E);
        upload(c->bnd_faces, bfaces.data(), bfaces.size());
        upload(c->arr[MHFEM_CELL_MEASURE], cm.data(), cm.size());
        upload(c->arr[MHFEM_FACE_MEASURE], fm.data(), fm.size());
        // model arrays (BaseModel::allocate, BaseModel.h:168-195)
        for (int w = MHFEM_Z_iF; w <= MHFEM_BC_MOBILITY; w++) {
            if (is_coefficient(w)) {
                c->arr[w].alloc(1);
                c->uniform[w] = true;
                MHFEM_HIP(hipMemset(c->arr[w].p, 0, sizeof(double)));
            } else {
                c->arr[w].alloc(c->array_size(w));
                MHFEM_HIP(hipMemset(c->arr[w].p, 0, c->array_size(w) * sizeof(double)));
            }
        }
        // mobility upwind values default to 1 (constant-mobility models, HeatEquationModel.h:130)
        {
            std::vector<double> ones((size_t)c->n * c->F, 1.0);
            MHFEM_HIP(hipMemcpy(c->arr[MHFEM_m_iE_upw].p, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice));
            MHFEM_HIP(hipMemcpy(c->arr[MHFEM_BC_MOBILITY].p, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        // default boundary conditions: zero Dirichlet (Solver.hpp:281-294)
        c->h_tags.assign((size_t)c->n * c->F, MHFEM_BC_FIXED_VALUE);
        upload(c->tags, c->h_tags.data(), c->h_tags.size());
        // stationary boundary model (BoundaryModels/Stationary.h:8-38) until
        // mhfem_set_boundary_time_function
        for (TimeFn &f : c->h_bc_fn) f = TimeFn{-1, 0, 0.0, 0.0, 0.0, 0.0, nullptr, nullptr};
        upload(c->bc_fn, c->h_bc_fn, 2 * (size_t)c->n);
        c->clock.alloc(2);
        MHFEM_HIP(hipMemset(c->clock.p, 0, 2 * sizeof(double)));
        MHFEM_HIP(hipDeviceSynchronize());
        // fixed geometry: M^{-1} of every cell, once (packed planes, and
        // row-expanded with FPC*FPC values rounded up to an even count per cell)
        c->minv.alloc((size_t)(c->fpc * (c->fpc + 1) / 2) * Cn);
        c->minv_rows.alloc((size_t)((c->fpc * c->fpc + 1) & ~1) * Cn);
        // per-cell assembly records: R_ijKe, R_iK, D_ijK, padded to an even count
        const size_t rec = (size_t)((c->n * c->n * c->fpc + c->n + c->n * c->n + 1) & ~1);
        c->cell_rec.alloc(rec * Cn);
        MHFEM_HIP(hipMemset(c->cell_rec.p, 0, rec * Cn * sizeof(double)));
        launch_geometry(*c);
        MHFEM_HIP(hipStreamSynchronize(c->stream));
        return reinterpret_cast<mhfem_ctx *>(c);
    } catch (const Error &e) {
        g_create_error = e.msg;
    } catch (const std::exception &e) {
        g_create_error = e.what();
    }
    if (c) mhfem_destroy(reinterpret_cast<mhfem_ctx *>(c));
    return nullptr;
}

mhfem_ctx *mhfem_create_batch(const mhfem_mesh_desc *meshes, int count, const mhfem_model_desc *model)
{
    try {
        if (!meshes || count < 1 || !model) throw Error{MHFEM_E_ARGUMENT, "mhfem_create_batch: bad argument"};
        const int dim = meshes[0].dim;
        std::vector<double> pts;
        std::vector<int32_t> cells;
        std::vector<int> cell_off(1, 0);
        int voff = 0;
        for (int k = 0; k < count; k++) {
            const mhfem_mesh_desc &m = meshes[k];
            if (m.dim != dim || !m.points || !m.cell_vertices || m.num_cells < 1)
                throw Error{MHFEM_E_ARGUMENT, "mhfem_create_batch: instance " + std::to_string(k) + " is invalid"};
            pts.insert(pts.end(), m.points, m.points + (size_t)m.num_points * dim);
            for (size_t q = 0; q < (size_t)m.num_cells * (dim + 1); q++) {
                const int v = m.cell_vertices[q];
                if (v < 0 || v >= m.num_points) throw Error{MHFEM_E_MESH, "cell vertex index out of range"};
                cells.push_back(v + voff);
            }
            voff += m.num_points;
            cell_off.push_back((int)(cells.size() / (dim + 1)));
        }
        mhfem_mesh_desc u{dim, voff, cell_off.back(), pts.data(), cells.data()};
        mhfem_ctx *ctx = mhfem_create(&u, model);
        if (!ctx) return nullptr;
        Ctx *c = reinterpret_cast<Ctx *>(ctx);
        // faces are numbered by first appearance over the cells, so the faces
        // of each instance form a contiguous range
        std::vector<int> face_off(1, 0);
        int fmax = -1;
        for (int k = 0; k < count; k++) {
            int lo = INT32_MAX;
            for (size_t q = (size_t)cell_off[k] * c->fpc; q < (size_t)cell_off[k + 1] * c->fpc; q++) {
                fmax = std::max(fmax, c->h_cell_faces[q]);
                lo = std::min(lo, c->h_cell_faces[q]);
            }
            if (lo != face_off.back()) {
                mhfem_destroy(ctx);
                throw Error{MHFEM_E_MESH, "mhfem_create_batch: instances share faces"};
            }
            face_off.push_back(fmax + 1);
        }
        c->h_seg_cells = cell_off;
        c->h_seg_faces = face_off;
        return ctx;
    } catch (const Error &e) {
        g_create_error = e.msg;
    } catch (const std::exception &e) {
        g_create_error = e.what();
    }
    return nullptr;
}

int mhfem_batch_info(const mhfem_ctx *ctx, int *count, int32_t *cell_offsets, int32_t *face_offsets)
{
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    if (!c) return MHFEM_E_ARGUMENT;
    const int nseg = (int)c->h_seg_cells.size() - 1;
    if (count) *count = nseg;
    if (cell_offsets) std::copy(c->h_seg_cells.begin(), c->h_seg_cells.end(), cell_offsets);
    if (face_offsets) std::copy(c->h_seg_faces.begin(), c->h_seg_faces.end(), face_offsets);
    return MHFEM_OK;
}

void mhfem_destroy(mhfem_ctx *ctx)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    krylov_destroy(c->krylov);
    guess_destroy(c->guess);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c->kt_ev) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *mhfem_last_error(const mhfem_ctx *ctx)
{
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    return c ? c->err.c_str() : g_create_error.c_str();
}

int mhfem_get_counts(const mhfem_ctx *ctx, int *cells, int *faces, int *boundary_faces)
{
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    if (!c) return MHFEM_E_ARGUMENT;
    if (cells) *cells = c->C;
    if (faces) *faces = c->F;
    if (boundary_faces) *boundary_faces = c->nbnd;
    return MHFEM_OK;
}

int mhfem_get_topology(const mhfem_ctx *ctx, int32_t *cell_faces, int32_t *face_cells)
{
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    if (!c) return MHFEM_E_ARGUMENT;
    if (cell_faces) std::memcpy(cell_faces, c->h_cell_faces.data(), c->h_cell_faces.size() * sizeof(int));
    if (face_cells) std::memcpy(face_cells, c->h_face_cells.data(), c->h_face_cells.size() * sizeof(int));
    return MHFEM_OK;
}

long mhfem_array_size(const mhfem_ctx *ctx, int which)
{
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    if (!c || which < 0 || which >= MHFEM_NUM_ARRAYS) return -1;
    if (is_coefficient(which) && c->uniform[which]) return 1;
    return c->array_size(which);
}

int mhfem_set_array(mhfem_ctx *ctx, int which, const double *host, long count)
{
    return guarded(ctx, [&](Ctx &c) {
        if (which < 0 || which >= MHFEM_NUM_ARRAYS || !host) throw Error{MHFEM_E_ARGUMENT, "mhfem_set_array: bad argument"};
        if (which == MHFEM_CELL_MEASURE || which == MHFEM_FACE_MEASURE)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_set_array: measures are derived from the mesh"};
        const long full = c.array_size(which);
        if (which == MHFEM_RHS && !c.system_ready) throw Error{MHFEM_E_STATE, "linear system not set up"};
        const bool make_uniform = is_coefficient(which) && count == 1;
        if (!make_uniform && count != full)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_set_array: count must be the full size" +
                                              std::string(is_coefficient(which) ? " or 1" : "") + " (expected " +
                                              std::to_string(full) + ", got " + std::to_string(count) + ")"};
        // An allocation of the right size is kept, so that kernels queued (or
        // captured in step graphs) keep valid pointers; a new one is made only
        // after the work queued on the old one has finished.  Step graphs are
        // re-captured when a pointer or a uniform mask of the scheme changes.
        const size_t need = make_uniform ? 1 : (size_t)full;
        if (c.arr[which].n != need || !c.arr[which].p) {
            MHFEM_HIP(hipStreamSynchronize(c.stream));
            c.arr[which].alloc(need);
        }
        if (is_coefficient(which)) c.uniform[which] = make_uniform;
        MHFEM_HIP(hipMemcpyAsync(c.arr[which].p, host, count * sizeof(double), hipMemcpyHostToDevice, c.stream));
        MHFEM_HIP(hipStreamSynchronize(c.stream));
    });
}

int mhfem_get_array(mhfem_ctx *ctx, int which, double *host, long count)
{
    return guarded(ctx, [&](Ctx &c) {
        if (which < 0 || which >= MHFEM_NUM_ARRAYS || !host) throw Error{MHFEM_E_ARGUMENT, "mhfem_get_array: bad argument"};
        const long have = (is_coefficient(which) && c.uniform[which]) ? 1 : c.array_size(which);
        if (count != have || !c.arr[which].p)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_get_array: count mismatch (expected " + std::to_string(have) + ")"};
        // derived arrays the step does not keep in this layout: formed on request
        if (which == MHFEM_b_ijK_storage) launch_fill_b(c);
        if (which == MHFEM_R_ijKe || which == MHFEM_R_iK) launch_fill_R(c);
        MHFEM_HIP(hipMemcpyAsync(host, c.arr[which].p, count * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        MHFEM_HIP(hipStreamSynchronize(c.stream));
    });
}

void *mhfem_device_array(mhfem_ctx *ctx, int which)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || which < 0 || which >= MHFEM_NUM_ARRAYS) return nullptr;
    if (is_coefficient(which) && c->uniform[which]) return nullptr;
    if (which == MHFEM_b_ijK_storage || which == MHFEM_R_ijKe || which == MHFEM_R_iK) {
        // formed on request (a snapshot): b from M^{-1} and the current D,
        // R_ijKe / R_iK from the per-cell records of the last preIterate
        try {
            if (which == MHFEM_b_ijK_storage) launch_fill_b(*c);
            else launch_fill_R(*c);
        } catch (const Error &e) {
            c->err = e.msg;
            return nullptr;
        }
    }
    return c->arr[which].p;
}

void *mhfem_stream(mhfem_ctx *ctx)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    return c ? (void *)c->stream : nullptr;
}

int mhfem_set_boundary_conditions(mhfem_ctx *ctx, const uint8_t *tags, const double *values, const double *dirichlet)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!tags || !values || !dirichlet) throw Error{MHFEM_E_ARGUMENT, "mhfem_set_boundary_conditions: null array"};
        const size_t nF = (size_t)c.n * c.F;
        for (size_t q = 0; q < nF; q++)
            if (tags[q] > 3) throw Error{MHFEM_E_ARGUMENT, "unknown boundary condition type"};
        std::vector<uint8_t> old = c.h_tags;
        c.h_tags.assign(tags, tags + nF);
        MHFEM_HIP(hipMemcpy(c.tags.p, tags, nF, hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(c.arr[MHFEM_BC_VALUES].p, values, nF * sizeof(double), hipMemcpyHostToDevice));
        MHFEM_HIP(hipMemcpy(c.arr[MHFEM_BC_DIRICHLET].p, dirichlet, nF * sizeof(double), hipMemcpyHostToDevice));
        c.bc_set = true;
        // the row lengths depend on the tags (BoundaryConditions.hpp:283-299)
        if (c.system_ready && old != c.h_tags) c.system_ready = false;
    });
}

int mhfem_set_boundary_time_function(mhfem_ctx *ctx, int which, int species, const mhfem_time_function *f)
{
    return guarded(ctx, [&](Ctx &c) {
        if (which != MHFEM_BC_DIRICHLET && which != MHFEM_BC_VALUES)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_set_boundary_time_function: which must be MHFEM_BC_DIRICHLET or "
                                          "MHFEM_BC_VALUES"};
        if (species < 0 || species >= c.n)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_set_boundary_time_function: species out of range"};
        TimeFn t{-1, 0, 0.0, 0.0, 0.0, 0.0, nullptr, nullptr};
        if (f) {
            if (f->kind < MHFEM_TIME_CONSTANT || f->kind > MHFEM_TIME_TABLE)
                throw Error{MHFEM_E_ARGUMENT, "mhfem_set_boundary_time_function: unknown kind"};
            t.kind = f->kind;
            t.offset = f->offset;
            t.amplitude = f->amplitude;
            t.omega = f->angular_frequency;
            t.phase = f->phase;
            if (f->kind == MHFEM_TIME_TABLE) {
                if (f->points < 1 || !f->times || !f->values)
                    throw Error{MHFEM_E_ARGUMENT, "mhfem_set_boundary_time_function: a table needs points >= 1"};
                for (int k = 0; k < f->points; k++)
                    if (!std::isfinite(f->times[k]) || !std::isfinite(f->values[k]) ||
                        (k > 0 && !(f->times[k] > f->times[k - 1])))
                        throw Error{MHFEM_E_ARGUMENT,
                                    "mhfem_set_boundary_time_function: table times must be finite and ascending"};
                t.points = f->points;
            }
        }
        const int slot = (which == MHFEM_BC_DIRICHLET ? 0 : c.n) + species;
        // the work queued so far may still read the old function and table
        MHFEM_HIP(hipStreamSynchronize(c.stream));
        if (t.kind == MHFEM_TIME_TABLE) {
            std::vector<double> tab(f->times, f->times + f->points);
            tab.insert(tab.end(), f->values, f->values + f->points);
            upload(c.bc_tab[slot], tab.data(), tab.size());
            t.times = c.bc_tab[slot].p;
            t.values = c.bc_tab[slot].p + f->points;
        } else {
            c.bc_tab[slot].release();
        }
        c.h_bc_fn[slot] = t;
        MHFEM_HIP(hipMemcpy(c.bc_fn.p, c.h_bc_fn, 2 * (size_t)c.n * sizeof(TimeFn), hipMemcpyHostToDevice));
        c.bc_timed = false;
        for (int q = 0; q < 2 * c.n; q++) c.bc_timed |= c.h_bc_fn[q].kind >= 0;
    });
}

int mhfem_set_initial_condition(mhfem_ctx *ctx, int i, const double *Z_K, long count)
{
    return guarded(ctx, [&](Ctx &c) {
        if (i < 0 || i >= c.n || !Z_K) throw Error{MHFEM_E_ARGUMENT, "mhfem_set_initial_condition: bad argument"};
        if (count != c.C)
            throw Error{MHFEM_E_ARGUMENT, "wrong vector length for the initial condition: expected " +
                                              std::to_string(c.C) + " elements, got " + std::to_string(count)};
        MHFEM_HIP(hipMemcpy(c.arr[MHFEM_Z_iK].p + (size_t)i * c.C, Z_K, count * sizeof(double), hipMemcpyHostToDevice));
    });
}

int mhfem_init(mhfem_ctx *ctx)
{
    return guarded(ctx, [&](Ctx &c) {
        launch_face_average(c);
        launch_zero_double(c, c.arr[MHFEM_v_iKe].p, c.arr[MHFEM_v_iKe].n);
        MHFEM_HIP(hipStreamSynchronize(c.stream));
    });
}

int mhfem_setup_linear_system(mhfem_ctx *ctx, const mhfem_solver_config *cfg)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!cfg) throw Error{MHFEM_E_ARGUMENT, "mhfem_setup_linear_system: null config"};
        if (cfg->solver < 0 || cfg->solver > MHFEM_SOLVER_SOR) throw Error{MHFEM_E_ARGUMENT, "unknown linear solver"};
        if (cfg->preconditioner < 0 || cfg->preconditioner > MHFEM_PREC_ILUT)
            throw Error{MHFEM_E_ARGUMENT, "unknown preconditioner"};
        const bool reused = cfg->preconditioner == MHFEM_PREC_AMG || is_isai(cfg->preconditioner) ||
                            cfg->preconditioner == MHFEM_PREC_ILUT;
        const char *pname = cfg->preconditioner == MHFEM_PREC_AMG    ? "AMG"
                            : cfg->preconditioner == MHFEM_PREC_ILUT ? "ILUT"
                                                                     : "ILU-ISAI";
        if (cfg->solver == MHFEM_SOLVER_SOR && cfg->preconditioner != MHFEM_PREC_NONE &&
            cfg->preconditioner != MHFEM_PREC_DIAGONAL)
            throw Error{MHFEM_E_ARGUMENT, std::string("SOR uses no preconditioner: take none with it, not ") +
                                              (cfg->preconditioner == MHFEM_PREC_ILUT ? "ILUT" : "this one")};
        if (c.h_seg_faces.size() > 2 && (cfg->solver != MHFEM_SOLVER_BICGSTAB || cfg->precision != MHFEM_PRECISION_FP64))
            throw Error{MHFEM_E_ARGUMENT, "batched systems are solved with FP64 BiCGStab"};
        if (reused && cfg->precision != MHFEM_PRECISION_FP64)
            throw Error{MHFEM_E_ARGUMENT, std::string("the ") + pname + " preconditioner works with the FP64 solvers"};
        if (reused && c.h_seg_faces.size() > 2)
            throw Error{MHFEM_E_ARGUMENT, std::string("the ") + pname + " preconditioner is for single systems, not batches"};
        if (cfg->precision < 0 || cfg->precision > 2) throw Error{MHFEM_E_ARGUMENT, "unknown precision"};
        if (cfg->precision != MHFEM_PRECISION_FP64 && cfg->solver != MHFEM_SOLVER_BICGSTAB)
            throw Error{MHFEM_E_ARGUMENT, "FP32 refinement is implemented for BiCGStab only"};
        c.cfg = *cfg;
        if (c.cfg.max_iterations <= 0) c.cfg.max_iterations = 100000;
        if (c.cfg.check_every <= 0) c.cfg.check_every = 1;
        if (c.cfg.gmres_restart <= 0) c.cfg.gmres_restart = 30;
        if (c.cfg.bicgstab_ell <= 0) c.cfg.bicgstab_ell = 1;
        if (c.cfg.idr_s <= 0) c.cfg.idr_s = 4;
        if (c.cfg.bicgstab_ell > 8 || c.cfg.idr_s > 8) throw Error{MHFEM_E_ARGUMENT, "bicgstab_ell and idr_s are at most 8"};
        if (c.cfg.sor_omega == 0.0) c.cfg.sor_omega = 1.0;
        if (c.cfg.ilut_p <= 0) c.cfg.ilut_p = 10;
        if (c.cfg.ilut_threshold <= 0.0) c.cfg.ilut_threshold = 1e-4;
        if (!(c.cfg.sor_omega > 0.0 && c.cfg.sor_omega < 2.0)) throw Error{MHFEM_E_ARGUMENT, "sor_omega must lie in (0, 2)"};
        // row lengths (Solver.hpp:449-474)
        const int n = c.n, fpc = c.fpc;
        c.rows = n * c.F;
        c.h_rowptr.assign((size_t)c.rows + 1, 0);
        for (int E = 0; E < c.F; E++)
            for (int i = 0; i < n; i++) {
                int len;
                if (c.h_face_cells[2 * (size_t)E + 1] < 0)
                    len = c.h_tags[(size_t)i * c.F + E] == MHFEM_BC_FIXED_VALUE ? 1 : fpc * n;
                else
                    len = (2 * fpc - 1) * n;
                c.h_rowptr[(size_t)E * n + i + 1] = len;
            }
        for (int r = 0; r < c.rows; r++) c.h_rowptr[r + 1] += c.h_rowptr[r];
        c.nnz = c.h_rowptr[c.rows];
        c.h_seg_rows.clear();
        for (int f : c.h_seg_faces) c.h_seg_rows.push_back(f * n);
        if ((long)c.nnz >= (long)INT32_MAX) throw Error{MHFEM_E_ARGUMENT, "matrix too large for int32 row pointers"};
        // columns (LinearSystem.h:181-245, BoundaryConditions.hpp:196-218)
        c.h_col.assign(c.nnz, 0);
        for (int E = 0; E < c.F; E++) {
            const int K0 = c.h_face_cells[2 * (size_t)E], K1 = c.h_face_cells[2 * (size_t)E + 1];
            int p0[6], p1[6];
            sorted_local(c, K0, p0);
            const int *f0 = &c.h_cell_faces[(size_t)K0 * fpc];
            for (int i = 0; i < n; i++) {
                int *cc = &c.h_col[c.h_rowptr[(size_t)E * n + i]];
                int t = 0;
                if (K1 < 0) {
                    if (c.h_tags[(size_t)i * c.F + E] == MHFEM_BC_FIXED_VALUE) {
                        cc[0] = E * n + i;
                        continue;
                    }
                    for (int g = 0; g < fpc; g++)
                        for (int j = 0; j < n; j++) cc[t++] = f0[p0[g]] * n + j;
                } else {
                    sorted_local(c, K1, p1);
                    const int *f1 = &c.h_cell_faces[(size_t)K1 * fpc];
                    int g0 = 0, g1 = 0;
                    while (g0 < fpc || g1 < fpc) {
                        const int F0 = g0 < fpc ? f0[p0[g0]] : INT32_MAX, F1 = g1 < fpc ? f1[p1[g1]] : INT32_MAX;
                        const int Fm = std::min(F0, F1);
                        for (int j = 0; j < n; j++) cc[t++] = Fm * n + j;
                        if (F0 == Fm) g0++;
                        if (F1 == Fm) g1++;
                    }
                }
            }
        }
        // block SELL-64 layout (common.hpp): per face the ascending list of
        // the other faces its rows couple to (union over its n rows), one
        // n x n block each; csr_to_sell maps every CSR entry to its value
        // position, -1 for the unit diagonal
        const int nslices = (c.F + 63) / 64;
        std::vector<int64_t> bptr((size_t)c.F + 1, 0);
        std::vector<int> bface;
        bface.reserve((size_t)c.F * (2 * fpc - 2));
        for (int E = 0; E < c.F; E++) {
            const size_t start = bface.size();
            for (int i = 0; i < n; i++)
                for (int q = c.h_rowptr[(size_t)E * n + i]; q < c.h_rowptr[(size_t)E * n + i + 1]; q++) {
                    const int Fc = c.h_col[q] / n;
                    if (Fc != E) bface.push_back(Fc);
                }
            std::sort(bface.begin() + start, bface.end());
            bface.erase(std::unique(bface.begin() + start, bface.end()), bface.end());
            bptr[E + 1] = (int64_t)bface.size();
        }
        c.h_slice_ptr.assign((size_t)nslices + 1, 0);
        c.h_slice_width.assign(nslices, 0);
        for (int s = 0; s < nslices; s++) {
            int64_t w = 0;
            for (int E = s * 64; E < std::min(c.F, s * 64 + 64); E++) w = std::max(w, bptr[E + 1] - bptr[E]);
            c.h_slice_width[s] = (int)w;
            c.h_slice_ptr[s + 1] = c.h_slice_ptr[s] + w * 64;
        }
        const int nn = n * n;
        c.sell_blocks = c.h_slice_ptr[nslices];
        c.diag_base = c.sell_blocks * nn;
        c.sell_size = c.diag_base + (int64_t)c.F * n * (n - 1);
        std::vector<int> scol(std::max<int64_t>(c.sell_blocks, 1));
        for (int s = 0; s < nslices; s++)
            for (int lane = 0; lane < 64; lane++) {
                const int E = s * 64 + lane;
                const int64_t w = E < c.F ? bptr[E + 1] - bptr[E] : 0;
                for (int k = 0; k < c.h_slice_width[s]; k++)
                    scol[sell_block(c.h_slice_ptr.data(), E, k)] = k < w ? bface[bptr[E] + k] : (E < c.F ? E : 0);
            }
        std::vector<int64_t> c2s(c.nnz);
        for (int E = 0; E < c.F; E++)
            for (int i = 0; i < n; i++) {
                const int r = E * n + i;
                int k = 0;
                for (int q = c.h_rowptr[r]; q < c.h_rowptr[r + 1]; q++) {
                    const int Fc = c.h_col[q] / n, j = c.h_col[q] % n;
                    if (Fc == E) {
                        c2s[q] = j == i ? -1 : sell_diag(c.diag_base, n, E, i, j);
                        continue;
                    }
                    while (bface[bptr[E] + k] != Fc) k++;  // the row's faces are a sorted subset
                    c2s[q] = sell_block(c.h_slice_ptr.data(), E, k) * nn + i * n + j;
                }
            }
        c.h_bptr = std::move(bptr);
        c.h_bface = std::move(bface);
        upload(c.slice_ptr, c.h_slice_ptr.data(), c.h_slice_ptr.size());
        upload(c.slice_width, c.h_slice_width.data(), c.h_slice_width.size());
        upload(c.sell_col, scol.data(), scol.size());
        upload(c.csr_to_sell, c2s.data(), c2s.size());
        upload(c.rowptr, c.h_rowptr.data(), c.h_rowptr.size());
        upload(c.csr_col, c.h_col.data(), c.h_col.size());
        c.sell_val.alloc(std::max<int64_t>(c.sell_size, 1));
        MHFEM_HIP(hipMemset(c.sell_val.p, 0, std::max<int64_t>(c.sell_size, 1) * sizeof(double)));
        c.arr[MHFEM_RHS].alloc(c.rows);
        MHFEM_HIP(hipMemset(c.arr[MHFEM_RHS].p, 0, c.rows * sizeof(double)));
        // the old solver's buffers may still be read by queued work
        MHFEM_HIP(hipStreamSynchronize(c.stream));
        krylov_destroy(c.krylov);
        c.krylov = nullptr;
        guess_destroy(c.guess);
        c.guess = nullptr;
        c.guess_on_solve = false;
        c.system_ready = true;
        c.krylov = krylov_create(c);
        MHFEM_HIP(hipDeviceSynchronize());
    });
}

int mhfem_pre_iterate(mhfem_ctx *ctx, double time, double tau)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!(tau > 0)) throw Error{MHFEM_E_ARGUMENT, "time step must be positive"};
        phase_begin(c, 6);
        launch_pre_iterate(c, time, tau);
        phase_end(c, 6, c.timers.pre_iterate);
        if (c.timing) {
            float a = 0, b = 0;
            MHFEM_HIP(hipEventElapsedTime(&a, c.ev[0], c.ev[1]));
            MHFEM_HIP(hipEventElapsedTime(&b, c.ev[1], c.ev[2]));
            c.timers.upwind += a;
            c.timers.update_b_R_Q += b;
        }
    });
}

int mhfem_assemble(mhfem_ctx *ctx, double time, double tau)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!c.system_ready) throw Error{MHFEM_E_STATE, "mhfem_setup_linear_system must be called first"};
        phase_begin(c, 8);
        // the boundary model sees time + tau (Solver.hpp:948)
        launch_set_clock(c, time + tau);
        launch_assemble(c, time, tau);
        phase_end(c, 8, c.timers.assemble);
    });
}

int mhfem_solve(mhfem_ctx *ctx, mhfem_solve_stats *stats)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!c.system_ready) throw Error{MHFEM_E_STATE, "mhfem_setup_linear_system must be called first"};
        mhfem_solve_stats st{};
        krylov_solve(c, c.arr[MHFEM_RHS].p, c.arr[MHFEM_Z_iF].p, &st);
        c.timers.iterations += st.iterations;
        c.timers.time_steps += 1;
        if (stats) *stats = st;
        if (!st.converged)
            throw Error{MHFEM_E_NOT_CONVERGED, "MHFEM error: the linear system solver did not converge (" +
                                                   std::to_string(st.iterations) + " iterations)."};
    });
}

int mhfem_post_iterate(mhfem_ctx *ctx, double time, double tau)
{
    (void)time;
    (void)tau;
    return guarded(ctx, [&](Ctx &c) {
        phase_begin(c, 10);
        launch_post_iterate(c);
        phase_end(c, 10, c.timers.post_iterate);
    });
}

int mhfem_solve_interval(mhfem_ctx *ctx, double t0, double t1, double dt, mhfem_solve_stats *last)
{
    if (!(dt > 0)) return MHFEM_E_ARGUMENT;
    double t = t0;
    int rc = MHFEM_OK;
    // control.h:74-101; the loop stops when less than 1e-6 of a step is left (rounding in t)
    while (t1 - t > dt * 1e-6) {
        const double tau = std::min(dt, t1 - t);
        if ((rc = mhfem_pre_iterate(ctx, t, tau))) return rc;
        if ((rc = mhfem_assemble(ctx, t, tau))) return rc;
        if ((rc = mhfem_solve(ctx, last))) return rc;
        if ((rc = mhfem_post_iterate(ctx, t, tau))) return rc;
        t += tau;
    }
    return rc;
}

int mhfem_system_size(const mhfem_ctx *ctx, int *rows, long *nnz)
{
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    if (!c || !c->system_ready) return MHFEM_E_STATE;
    if (rows) *rows = c->rows;
    if (nnz) *nnz = c->nnz;
    return MHFEM_OK;
}

int mhfem_export_system(mhfem_ctx *ctx, int32_t *rowptr, int32_t *col, double *val, double *rhs)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!c.system_ready) throw Error{MHFEM_E_STATE, "linear system not set up"};
        if (rowptr) std::memcpy(rowptr, c.h_rowptr.data(), c.h_rowptr.size() * sizeof(int));
        if (col) std::memcpy(col, c.h_col.data(), c.h_col.size() * sizeof(int));
        if (val) {
            std::vector<double> sv(c.sell_size);
            std::vector<int64_t> c2s(c.nnz);
            MHFEM_HIP(hipMemcpyAsync(sv.data(), c.sell_val.p, c.sell_size * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            MHFEM_HIP(hipMemcpyAsync(c2s.data(), c.csr_to_sell.p, c.nnz * sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
            MHFEM_HIP(hipStreamSynchronize(c.stream));
            for (long q = 0; q < c.nnz; q++) val[q] = c2s[q] < 0 ? 1.0 : sv[c2s[q]];
        }
        if (rhs) {
            MHFEM_HIP(hipMemcpyAsync(rhs, c.arr[MHFEM_RHS].p, c.rows * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            MHFEM_HIP(hipStreamSynchronize(c.stream));
        }
    });
}

int mhfem_solve_rhs(mhfem_ctx *ctx, const double *rhs, double *x, mhfem_solve_stats *stats)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!c.system_ready) throw Error{MHFEM_E_STATE, "linear system not set up"};
        if (!rhs || !x) throw Error{MHFEM_E_ARGUMENT, "null argument"};
        DBuf<double> b, xx;
        upload(b, rhs, c.rows);
        upload(xx, x, c.rows);
        mhfem_solve_stats st{};
        krylov_solve(c, b.p, xx.p, &st);
        MHFEM_HIP(hipMemcpyAsync(x, xx.p, c.rows * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        MHFEM_HIP(hipStreamSynchronize(c.stream));
        if (stats) *stats = st;
    });
}

int mhfem_solve_rhs_noalloc(mhfem_ctx *ctx, const double *rhs, double *x, mhfem_solve_stats *stats)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!c.system_ready) throw Error{MHFEM_E_STATE, "linear system not set up"};
        if (!rhs || !x) throw Error{MHFEM_E_ARGUMENT, "null argument"};
        if ((long)c.rows > (long)MHFEM_NOALLOC_MAX_ROWS)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_solve_rhs_noalloc: the system has " + std::to_string(c.rows) +
                                              " rows; the staging array holds systems of at most " +
                                              std::to_string((long)MHFEM_NOALLOC_MAX_ROWS) + " rows"};
        // only configurations whose solve allocates nothing: the BiCGStab
        // vectors, the FP32 copies and the ILU(0) factors are allocated when
        // the linear system is set up
        const int p = c.cfg.preconditioner;
        if (c.cfg.solver != MHFEM_SOLVER_BICGSTAB ||
            (p != MHFEM_PREC_NONE && p != MHFEM_PREC_DIAGONAL && p != MHFEM_PREC_ILU0))
            throw Error{MHFEM_E_ARGUMENT, "mhfem_solve_rhs_noalloc: BiCGStab with no, diagonal or ILU(0) "
                                          "preconditioning only"};
        if (c.guess && c.guess_on_solve)
            throw Error{MHFEM_E_ARGUMENT, "mhfem_solve_rhs_noalloc: the learned guess must be off"};
        double *const base = scratch_on(c.device);
        double *const b = base, *const xx = base + c.rows;
        MHFEM_HIP(hipMemcpyAsync(b, rhs, c.rows * sizeof(double), hipMemcpyHostToDevice, c.stream));
        MHFEM_HIP(hipMemcpyAsync(xx, x, c.rows * sizeof(double), hipMemcpyHostToDevice, c.stream));
        mhfem_solve_stats st{};
        krylov_solve(c, b, xx, &st);
        MHFEM_HIP(hipMemcpyAsync(x, xx, c.rows * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        MHFEM_HIP(hipStreamSynchronize(c.stream));
        if (stats) *stats = st;
    });
}

int mhfem_guess_configure(mhfem_ctx *ctx, const mhfem_guess_config *cfg)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!cfg) throw Error{MHFEM_E_ARGUMENT, "null guess config"};
        if (!c.system_ready) throw Error{MHFEM_E_STATE, "mhfem_setup_linear_system must be called first"};
        if (!c.guess) c.guess = guess_create(c);
        guess_configure(c, *c.guess, *cfg);
        c.guess_on_solve = cfg->enabled != 0;
    });
}

int mhfem_guess_stats_last(const mhfem_ctx *ctx, mhfem_guess_stats *out)
{
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    if (!c || !out) return MHFEM_E_ARGUMENT;
    if (!c->guess) return MHFEM_E_STATE;
    *out = guess_last_stats(*c->guess);
    return MHFEM_OK;
}

int mhfem_guess_num_params(void) { return guess_num_params(); }

static void need_guess(Ctx &c)
{
    if (!c.system_ready) throw Error{MHFEM_E_STATE, "mhfem_setup_linear_system must be called first"};
    if (!c.guess) c.guess = guess_create(c);
}

int mhfem_guess_get_weights(mhfem_ctx *ctx, float *params, int count)
{
    return guarded(ctx, [&](Ctx &c) {
        need_guess(c);
        if (!params || count != guess_num_params()) throw Error{MHFEM_E_ARGUMENT, "weight count mismatch"};
        guess_get_weights(c, *c.guess, params);
    });
}

int mhfem_guess_set_weights(mhfem_ctx *ctx, const float *params, int count)
{
    return guarded(ctx, [&](Ctx &c) {
        need_guess(c);
        if (!params || count != guess_num_params()) throw Error{MHFEM_E_ARGUMENT, "weight count mismatch"};
        guess_set_weights(c, *c.guess, params);
    });
}

int mhfem_guess_forward(mhfem_ctx *ctx, const uint16_t *features, long rows, float *y)
{
    return guarded(ctx, [&](Ctx &c) {
        need_guess(c);
        if (!features || !y || rows <= 0) throw Error{MHFEM_E_ARGUMENT, "bad arguments"};
        guess_forward(c, *c.guess, features, rows, y);
    });
}

int mhfem_guess_gradient(mhfem_ctx *ctx, const uint16_t *features, const float *targets, long rows, float *grad)
{
    return guarded(ctx, [&](Ctx &c) {
        need_guess(c);
        if (!features || !targets || !grad || rows <= 0) throw Error{MHFEM_E_ARGUMENT, "bad arguments"};
        guess_gradient(c, *c.guess, features, targets, rows, grad);
    });
}

int mhfem_batch_stats(mhfem_ctx *ctx, int32_t *iterations, double *residues, int count)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!c.system_ready) throw Error{MHFEM_E_STATE, "linear system not set up"};
        krylov_segment_stats(c, iterations, residues, count);
    });
}

int mhfem_step(mhfem_ctx *ctx, double time, double tau, mhfem_solve_stats *stats)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!c.system_ready) throw Error{MHFEM_E_STATE, "mhfem_setup_linear_system must be called first"};
        if (!(tau > 0)) throw Error{MHFEM_E_ARGUMENT, "time step must be positive"};
        if (c.guess_on_solve) throw Error{MHFEM_E_STATE, "graph steps do not run the learned guess; use the four calls"};
        mhfem_solve_stats st{};
        krylov_graph_step(c, time, tau, &st);
        c.timers.iterations += st.iterations;
        c.timers.time_steps += 1;
        if (stats) *stats = st;
        if (!st.converged)
            throw Error{MHFEM_E_NOT_CONVERGED, "MHFEM error: the linear system solver did not converge (" +
                                                   std::to_string(st.iterations) + " iterations)."};
    });
}

int mhfem_autotune(mhfem_ctx *ctx, int *rows_per_block, double *kernel_ms)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!c.system_ready) throw Error{MHFEM_E_STATE, "linear system not set up"};
        double ms = 0;
        const int rpb = krylov_autotune(c, &ms);
        c.cfg.rows_per_block = rpb;
        if (rows_per_block) *rows_per_block = rpb;
        if (kernel_ms) *kernel_ms = ms;
    });
}

int mhfem_benchmark_spmv(mhfem_ctx *ctx, int precision, int reps, double *avg_ms)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!c.system_ready || !avg_ms) throw Error{MHFEM_E_STATE, "linear system not set up"};
        krylov_bench_spmv(c, precision, reps, avg_ms);
    });
}

int mhfem_amg_info(const mhfem_ctx *ctx, int *levels, int *rebuilds)
{
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    if (!c || !c->krylov || !krylov_amg(*c->krylov)) return MHFEM_E_STATE;
    const Amg &a = *krylov_amg(*c->krylov);
    if (levels) *levels = amg_num_levels(a);
    if (rebuilds) *rebuilds = amg_rebuilds(a);
    return MHFEM_OK;
}

int mhfem_amg_level(const mhfem_ctx *ctx, int level, int *rows, long *nnz, int32_t *aggregates)
{
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    if (!c || !c->krylov || !krylov_amg(*c->krylov)) return MHFEM_E_STATE;
    const Amg &a = *krylov_amg(*c->krylov);
    if (level < 0 || level >= amg_num_levels(a)) return MHFEM_E_ARGUMENT;
    amg_level_info(a, level, rows, nnz, aggregates);
    return MHFEM_OK;
}

int mhfem_amg_apply(mhfem_ctx *ctx, const double *b, double *x)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!c.krylov || !krylov_amg(*c.krylov)) throw Error{MHFEM_E_STATE, "no AMG preconditioner"};
        Amg &a = *krylov_amg(*c.krylov);
        if (amg_num_levels(a) == 0) throw Error{MHFEM_E_STATE, "the AMG hierarchy is built by the first solve"};
        DBuf<double> db, dx;
        db.alloc(c.rows);
        dx.alloc(c.rows);
        MHFEM_HIP(hipMemcpyAsync(db.p, b, c.rows * sizeof(double), hipMemcpyHostToDevice, c.stream));
        amg_apply(c, a, db.p, dx.p, nullptr);
        MHFEM_HIP(hipMemcpyAsync(x, dx.p, c.rows * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        MHFEM_HIP(hipStreamSynchronize(c.stream));
    });
}

int mhfem_isai_info(const mhfem_ctx *ctx, long *nnz_lower, long *nnz_upper, int *rebuilds)
{
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    if (!c || !c->krylov || !krylov_isai_info(*c->krylov, nnz_lower, nnz_upper, rebuilds)) return MHFEM_E_STATE;
    return MHFEM_OK;
}

int mhfem_isai_matrix(mhfem_ctx *ctx, int upper, int32_t *rowptr, int32_t *col, double *val)
{
    return guarded(ctx, [&](Ctx &c) {
        if (!c.krylov || !krylov_isai_info(*c.krylov, nullptr, nullptr, nullptr))
            throw Error{MHFEM_E_STATE, "no ILU-ISAI preconditioner"};
        krylov_isai_export(c, upper ? 1 : 0, rowptr, col, val);
    });
}

int mhfem_get_timers(const mhfem_ctx *ctx, mhfem_timers *out)
{
    Ctx *c = const_cast<Ctx *>(reinterpret_cast<const Ctx *>(ctx));
    if (!c || !out) return MHFEM_E_ARGUMENT;
    try {
        kernel_timing_flush(*c);
    } catch (const Error &e) {
        c->err = e.msg;
        return e.code;
    }
    *out = c->timers;
    kernel_timing_summary(*c, out->spmv_ms, out->spmv_timed);
    return MHFEM_OK;
}

int mhfem_reset_timers(mhfem_ctx *ctx)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return MHFEM_E_ARGUMENT;
    c->kt_used = 0;
    c->kt_samples.clear();
    c->timers = mhfem_timers{};
    return MHFEM_OK;
}

int mhfem_set_timing(mhfem_ctx *ctx, int enabled)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return MHFEM_E_ARGUMENT;
    if (enabled < 0 || enabled > 2) return MHFEM_E_ARGUMENT;
    c->timing = enabled != 0;
    c->timing_level = enabled;
    return MHFEM_OK;
}

int mhfem_synchronize(mhfem_ctx *ctx)
{
    return guarded(ctx, [&](Ctx &c) { MHFEM_HIP(hipStreamSynchronize(c.stream)); });
}

}  // extern "C"
