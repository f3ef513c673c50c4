// FEM mesh fields on the host (geometry/mesh_field.hpp of the reference): construction of
// tetrahedral and triangle meshes (explicit, Cartesian grids, the two-file loader), their
// neighbour tables, inverse edge matrices and cell fields, point location and the field /
// potential queries of the C ABI, and the device image a model uploads. Compiled with
// -ffp-contract=off: the cell fields and inverses are the reference's to the last bit.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>

#include "capi_internal.h"
#include "mesh_host.hpp"

using cmpmc_internal::fail;

namespace {

constexpr double kInsideTol = -1e-10;     // mesh_field.hpp:73 kCartTolerance

// mesh_field.hpp:257-284 / 519-544. Facets (3D) or edges (2D) are keyed by their sorted
// vertices; two cells sharing a key are neighbours across the vertex opposite to it.
void build_neighbours(cmpmc_mesh& m) {
    const int nv = m.dim + 1;
    const int64_t n = m.n_cells();
    m.nbr.assign(static_cast<size_t>(n) * nv, -1);
    struct Facet {
        std::array<int32_t, 3> key;
        int32_t cell, slot;
    };
    std::vector<Facet> f;
    f.reserve(static_cast<size_t>(n) * nv);
    for (int64_t c = 0; c < n; ++c) {
        const int32_t* t = &m.cells[static_cast<size_t>(c) * nv];
        for (int s = 0; s < nv; ++s) {
            Facet e{{-1, -1, -1}, static_cast<int32_t>(c), s};
            int k = 0;
            for (int i = 0; i < nv; ++i)
                if (i != s) e.key[k++] = t[i];
            f.push_back(e);
        }
    }
    std::sort(f.begin(), f.end(), [](const Facet& a, const Facet& b) {
        if (a.key != b.key) return a.key < b.key;
        return a.cell != b.cell ? a.cell < b.cell : a.slot < b.slot;
    });
    for (size_t i = 0; i + 1 < f.size(); ++i) {
        if (f[i].key == f[i + 1].key) {
            m.nbr[static_cast<size_t>(f[i].cell) * nv + f[i].slot] = f[i + 1].cell;
            m.nbr[static_cast<size_t>(f[i + 1].cell) * nv + f[i + 1].slot] = f[i].cell;
            ++i;
        }
    }
}

// mesh_field.hpp:286-328: T has columns x_i - x_3; its inverse by cofactors; the cell
// field E = -(T^-1)^T (V_0 - V_3, V_1 - V_3, V_2 - V_3); |det| <= 1e-30 gives zeros.
void build_tets(cmpmc_mesh& m) {
    const int64_t n = m.n_cells();
    m.tinv.assign(static_cast<size_t>(n) * 9, 0.0);
    m.grad.assign(static_cast<size_t>(n) * 3, 0.0);
    for (int64_t c = 0; c < n; ++c) {
        const int32_t* t = &m.cells[static_cast<size_t>(c) * 4];
        const double* r = &m.x[static_cast<size_t>(t[3]) * 3];
        double e[3][3];                                    // e[j] = x_j - x_3
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) e[j][i] = m.x[static_cast<size_t>(t[j]) * 3 + i] - r[i];
        const double a = e[0][0], b = e[1][0], cc = e[2][0];
        const double d = e[0][1], ee = e[1][1], f = e[2][1];
        const double g = e[0][2], h = e[1][2], k = e[2][2];
        const double det = a * (ee * k - f * h) - b * (d * k - f * g) + cc * (d * h - ee * g);
        if (std::abs(det) <= 1e-30) continue;
        const double s = 1.0 / det;
        double* T = &m.tinv[static_cast<size_t>(c) * 9];
        T[0] = (ee * k - f * h) * s;
        T[1] = -(b * k - cc * h) * s;
        T[2] = (b * f - cc * ee) * s;
        T[3] = -(d * k - f * g) * s;
        T[4] = (a * k - cc * g) * s;
        T[5] = -(a * f - cc * d) * s;
        T[6] = (d * h - ee * g) * s;
        T[7] = -(a * h - b * g) * s;
        T[8] = (a * ee - b * d) * s;
        const double v3 = m.v[static_cast<size_t>(t[3])];
        const double dv[3] = {m.v[static_cast<size_t>(t[0])] - v3, m.v[static_cast<size_t>(t[1])] - v3,
                              m.v[static_cast<size_t>(t[2])] - v3};
        double* E = &m.grad[static_cast<size_t>(c) * 3];
        for (int i = 0; i < 3; ++i) E[i] = (T[i] * dv[0] + T[3 + i] * dv[1] + T[6 + i] * dv[2]) * (-1.0);
    }
}

// mesh_field.hpp:546-578
void build_tris(cmpmc_mesh& m) {
    const int64_t n = m.n_cells();
    m.tinv.assign(static_cast<size_t>(n) * 4, 0.0);
    m.grad.assign(static_cast<size_t>(n) * 2, 0.0);
    for (int64_t c = 0; c < n; ++c) {
        const int32_t* t = &m.cells[static_cast<size_t>(c) * 3];
        const double* r = &m.x[static_cast<size_t>(t[2]) * 2];
        const double a = m.x[static_cast<size_t>(t[0]) * 2] - r[0];
        const double b = m.x[static_cast<size_t>(t[1]) * 2] - r[0];
        const double cc = m.x[static_cast<size_t>(t[0]) * 2 + 1] - r[1];
        const double d = m.x[static_cast<size_t>(t[1]) * 2 + 1] - r[1];
        const double det = a * d - b * cc;
        if (std::abs(det) <= 1e-30) continue;
        const double s = 1.0 / det;
        double* T = &m.tinv[static_cast<size_t>(c) * 4];
        T[0] = d * s;
        T[1] = -b * s;
        T[2] = -cc * s;
        T[3] = a * s;
        const double dv0 = m.v[static_cast<size_t>(t[0])] - m.v[static_cast<size_t>(t[2])];
        const double dv1 = m.v[static_cast<size_t>(t[1])] - m.v[static_cast<size_t>(t[2])];
        double* E = &m.grad[static_cast<size_t>(c) * 2];
        E[0] = -(T[0] * dv0 + T[2] * dv1);
        E[1] = -(T[1] * dv0 + T[3] * dv1);
    }
}

// Validates and finishes a mesh whose x, v and cells are filled.
int finish(cmpmc_mesh* m, cmpmc_mesh** out) {
    const int nv = m->dim + 1;
    const int64_t n_vert = static_cast<int64_t>(m->v.size());
    const int64_t n = m->n_cells();
    if (n > INT32_MAX - 8) {
        delete m;
        return fail(CMPMC_ERR_INVALID, "mesh: %lld cells exceed the 32-bit cell index", static_cast<long long>(n));
    }
    for (size_t i = 0; i < m->x.size(); ++i)
        if (!std::isfinite(m->x[i])) {
            delete m;
            return fail(CMPMC_ERR_INVALID, "mesh: vertex coordinate %zu is not finite", i);
        }
    for (int64_t c = 0; c < n; ++c) {
        int32_t* t = &m->cells[static_cast<size_t>(c) * nv];
        for (int i = 0; i < nv; ++i)
            if (t[i] < 0 || t[i] >= n_vert) {
                const long long bad = t[i];
                delete m;
                return fail(CMPMC_ERR_INVALID, "mesh: cell %lld names vertex %lld of %lld", static_cast<long long>(c),
                            bad, static_cast<long long>(n_vert));
            }
        std::sort(t, t + nv);
    }
    build_neighbours(*m);
    if (m->dim == 3) build_tets(*m);
    else build_tris(*m);
    *out = m;
    return CMPMC_OK;
}

void bary_of(const cmpmc_mesh& m, const double* p, int32_t c, double* b) {
    if (m.dim == 3) {
        const double* T = &m.tinv[static_cast<size_t>(c) * 9];
        const double* r = &m.x[static_cast<size_t>(m.cells[static_cast<size_t>(c) * 4 + 3]) * 3];
        const double dx = p[0] - r[0], dy = p[1] - r[1], dz = p[2] - r[2];
        b[0] = T[0] * dx + T[1] * dy + T[2] * dz;
        b[1] = T[3] * dx + T[4] * dy + T[5] * dz;
        b[2] = T[6] * dx + T[7] * dy + T[8] * dz;
        b[3] = 1.0 - b[0] - b[1] - b[2];
    } else {
        const double* T = &m.tinv[static_cast<size_t>(c) * 4];
        const double* r = &m.x[static_cast<size_t>(m.cells[static_cast<size_t>(c) * 3 + 2]) * 2];
        const double du = p[0] - r[0], dv = p[1] - r[1];
        b[0] = T[0] * du + T[1] * dv;
        b[1] = T[2] * du + T[3] * dv;
        b[2] = 1.0 - b[0] - b[1];
    }
}

// the mesh-space point of a 3D position (mesh_field.hpp:491-496)
void project(const cmpmc_mesh& m, const double p[3], double q[3]) {
    if (m.dim == 3) { q[0] = p[0]; q[1] = p[1]; q[2] = p[2]; return; }
    if (m.axisymmetric_rz) { q[0] = std::hypot(p[0], p[1]); q[1] = p[2]; }
    else { q[0] = p[0]; q[1] = p[1]; }
    q[2] = 0.0;
}

}  // namespace

namespace cmpmc_internal {

bool mesh_locate(const cmpmc_mesh& m, const double* p, int32_t* hint, double* b) {
    const int64_t n = m.n_cells();
    if (n == 0) return false;
    if (*hint < 0 || *hint >= n) *hint = static_cast<int32_t>(n / 2);
    const int nv = m.dim + 1;
    const int64_t max_steps = n + 4;
    for (int64_t step = 0; step < max_steps; ++step) {
        bary_of(m, p, *hint, b);
        int mi = 0;
        for (int i = 1; i < nv; ++i)
            if (b[i] < b[mi]) mi = i;
        if (b[mi] >= kInsideTol) return true;
        const int32_t next = m.nbr[static_cast<size_t>(*hint) * nv + mi];
        if (next < 0) return false;
        *hint = next;
    }
    return false;
}

// charge_kernels.hpp:82-99 (the rate's field-dependent branch)
double iv_field_rate(const IvParams& p, const double E[3]) {
    const double n2 = E[0] * E[0] + E[1] * E[1] + E[2] * E[2];
    if (p.model == 0) {
        const double e0 = p.quad_field;
        return p.quad_rate * std::pow(e0 * e0 + n2, 0.5 * p.quad_power);
    }
    const double evcm = std::sqrt(n2) * 0.01;
    return p.lin_rate0 + p.lin_rate1 * std::pow(evcm, p.lin_power);
}

void mesh_image(const cmpmc_mesh& m, const IvParams& iv, bool one_valley, MeshImage& out) {
    const int64_t n = m.n_cells();
    const int nv = m.dim + 1;
    const int w64 = (m.dim == 3) ? 16 : 8, w32 = (m.dim == 3) ? 16 : 12;
    out.rec64.assign(static_cast<size_t>(n) * w64, 0.0);
    out.rec32.assign(static_cast<size_t>(n) * w32, 0.0f);
    out.nbr4.assign(static_cast<size_t>(n) * 4, -1);
    const double zero[3] = {0.0, 0.0, 0.0};
    out.iv_zero = iv_field_rate(iv, zero);
    for (int64_t c = 0; c < n; ++c) {
        const int nt = (m.dim == 3) ? 9 : 4;
        double* r = &out.rec64[static_cast<size_t>(c) * w64];
        float* f = &out.rec32[static_cast<size_t>(c) * w32];
        for (int i = 0; i < nt; ++i) r[i] = m.tinv[static_cast<size_t>(c) * nt + i];
        const int32_t ref = m.cells[static_cast<size_t>(c) * nv + m.dim];
        for (int i = 0; i < m.dim; ++i) r[nt + i] = m.x[static_cast<size_t>(ref) * m.dim + i];
        for (int i = 0; i < m.dim; ++i) r[nt + m.dim + i] = m.grad[static_cast<size_t>(c) * m.dim + i];
        double E[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < m.dim; ++i) E[i] = m.grad[static_cast<size_t>(c) * m.dim + i];
        const double ivc = iv_field_rate(iv, E);
        if (m.dim == 3) r[15] = ivc;
        const int nrec = nt + 2 * m.dim;
        for (int i = 0; i < nrec; ++i) f[i] = static_cast<float>(r[i]);
        f[nrec] = one_valley ? 0.0f : static_cast<float>(ivc);
        if (m.dim == 2) f[nrec + 1] = static_cast<float>(std::sqrt(E[0] * E[0] + E[1] * E[1]));
        for (int i = 0; i < nv; ++i) out.nbr4[static_cast<size_t>(c) * 4 + i] = m.nbr[static_cast<size_t>(c) * nv + i];
    }
    // FP32 walk tolerance: a barycentric coordinate is a row of T^-1 (or minus their sum)
    // times p - x_ref; in FP32 the difference carries an error of about one unit in the
    // last place of the largest coordinate and each product and sum a relative 2^-24, so
    // four times (3 max row 1-norm * max |x| * 2^-24) plus 1e-6 bounds the error, and of
    // two cells sharing a face at least one accepts a point on it.
    double max_row = 0.0, max_abs = 0.0;
    const int nt = (m.dim == 3) ? 9 : 4;
    for (int64_t c = 0; c < n; ++c)
        for (int i = 0; i < m.dim; ++i) {
            double r = 0.0;
            for (int j = 0; j < m.dim; ++j) r += std::abs(m.tinv[static_cast<size_t>(c) * nt + i * m.dim + j]);
            max_row = std::max(max_row, r);
        }
    for (double xv : m.x) max_abs = std::max(max_abs, std::abs(xv));
    out.tol_f = static_cast<float>(4.0 * (3.0 * max_row * max_abs * 0x1.0p-24) + 1e-6);
    // start cells: a regular grid of boxes over the bounding box, about one box per few
    // cells (at most 2^21 boxes), each holding the cell its centre lies in (or the hull
    // cell the walk to it stopped at)
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    const size_t n_vert = m.v.size();
    for (int i = 0; i < m.dim; ++i) {
        lo[i] = hi[i] = n_vert ? m.x[static_cast<size_t>(i)] : 0.0;
        for (size_t k = 0; k < n_vert; ++k) {
            lo[i] = std::min(lo[i], m.x[k * m.dim + i]);
            hi[i] = std::max(hi[i], m.x[k * m.dim + i]);
        }
    }
    double size = 1.0;
    for (int i = 0; i < m.dim; ++i) size *= std::max(hi[i] - lo[i], 1e-300);
    const double boxes = std::clamp(static_cast<double>(n) / (m.dim == 3 ? 6.0 : 2.0), 1.0, 2097152.0);
    const double edge = std::pow(size / boxes, 1.0 / m.dim);
    for (int i = 0; i < 3; ++i) {
        out.seed_n[i] = 1;
        out.lo[i] = lo[i];
        out.inv_h[i] = 0.0;
        if (i < m.dim) {
            const double ext = hi[i] - lo[i];
            out.seed_n[i] = static_cast<int>(std::clamp(std::ceil(ext / edge), 1.0, 256.0));
            out.inv_h[i] = ext > 0.0 ? out.seed_n[i] / ext : 0.0;
        }
    }
    const int nx = out.seed_n[0], ny = out.seed_n[1], nz = out.seed_n[2];
    out.seed.assign(static_cast<size_t>(nx) * ny * nz, 0);
    int32_t hint = 0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int jj = (i & 1) ? ny - 1 - j : j;           // serpentine: neighbouring boxes in turn
                const int kk = ((i * ny + j) & 1) ? nz - 1 - k : k;
                const int idx[3] = {i, jj, kk};
                double p[3] = {0, 0, 0}, b[4];
                for (int d = 0; d < m.dim; ++d)
                    p[d] = lo[d] + (idx[d] + 0.5) * ((hi[d] - lo[d]) / out.seed_n[d]);
                if (n > 0) mesh_locate(m, p, &hint, b);
                out.seed[(static_cast<size_t>(i) * ny + jj) * nz + kk] = hint;
            }
}

}  // namespace cmpmc_internal

extern "C" {

int cmpmc_mesh3d_create(const double* xyz, uint64_t n_vertices, const double* potentials, const int32_t* tets,
                        uint64_t n_tets, double v_min_clamp, double v_max_clamp, cmpmc_mesh** out) {
    if (!out || (n_vertices && (!xyz || !potentials)) || (n_tets && !tets))
        return fail(CMPMC_ERR_INVALID, "cmpmc_mesh3d_create: null argument");
    *out = nullptr;
    cmpmc_mesh* m = new (std::nothrow) cmpmc_mesh();
    if (!m) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh3d_create: out of host memory");
    m->dim = 3;
    m->x.assign(xyz, xyz + 3 * n_vertices);
    m->v.resize(n_vertices);
    for (uint64_t i = 0; i < n_vertices; ++i) m->v[i] = std::clamp(potentials[i], v_min_clamp, v_max_clamp);
    m->cells.assign(tets, tets + 4 * n_tets);
    return finish(m, out);
}

int cmpmc_mesh2d_create(const double* uv, uint64_t n_vertices, const double* potentials, const int32_t* tris,
                        uint64_t n_tris, int axisymmetric_rz, cmpmc_mesh** out) {
    if (!out || (n_vertices && (!uv || !potentials)) || (n_tris && !tris))
        return fail(CMPMC_ERR_INVALID, "cmpmc_mesh2d_create: null argument");
    *out = nullptr;
    cmpmc_mesh* m = new (std::nothrow) cmpmc_mesh();
    if (!m) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh2d_create: out of host memory");
    m->dim = 2;
    m->axisymmetric_rz = axisymmetric_rz ? 1 : 0;
    m->x.assign(uv, uv + 2 * n_vertices);
    m->v.assign(potentials, potentials + n_vertices);
    m->cells.assign(tris, tris + 3 * n_tris);
    return finish(m, out);
}

int cmpmc_mesh3d_grid_vertices(double hx, double hy, double hz, int nx, int ny, int nz, double* xyz) {
    if (!xyz) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh3d_grid_vertices: null argument");
    nx = std::max(1, nx); ny = std::max(1, ny); nz = std::max(1, nz);
    size_t id = 0;
    for (int i = 0; i <= nx; ++i) {
        const double x = -hx + 2.0 * hx * (static_cast<double>(i) / nx);
        for (int j = 0; j <= ny; ++j) {
            const double y = -hy + 2.0 * hy * (static_cast<double>(j) / ny);
            for (int k = 0; k <= nz; ++k, ++id) {
                xyz[3 * id] = x;
                xyz[3 * id + 1] = y;
                xyz[3 * id + 2] = -hz + 2.0 * hz * (static_cast<double>(k) / nz);
            }
        }
    }
    return CMPMC_OK;
}

int cmpmc_mesh3d_grid(double hx, double hy, double hz, int nx, int ny, int nz, const double* potentials,
                      cmpmc_mesh** out) {
    if (!out || !potentials) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh3d_grid: null argument");
    nx = std::max(1, nx); ny = std::max(1, ny); nz = std::max(1, nz);
    const int64_t n_vert = static_cast<int64_t>(nx + 1) * (ny + 1) * (nz + 1);
    if (n_vert > INT32_MAX || static_cast<int64_t>(nx) * ny * nz * 6 > INT32_MAX - 8)
        return fail(CMPMC_ERR_INVALID, "cmpmc_mesh3d_grid: %d x %d x %d voxels exceed the 32-bit indices", nx, ny, nz);
    std::vector<double> xyz(static_cast<size_t>(n_vert) * 3);
    cmpmc_mesh3d_grid_vertices(hx, hy, hz, nx, ny, nz, xyz.data());
    // mesh_field.hpp:154-180: six tetrahedra per voxel around its (0,0,0)-(1,1,1) diagonal
    std::vector<int32_t> tets;
    tets.reserve(static_cast<size_t>(nx) * ny * nz * 24);
    const int vy = ny + 1, vz = nz + 1;
    auto at = [vy, vz](int i, int j, int k) { return static_cast<int32_t>((i * vy + j) * vz + k); };
    static const int kCorner[6][2][3] = {{{1, 0, 0}, {1, 1, 0}}, {{1, 0, 0}, {1, 0, 1}}, {{0, 1, 0}, {1, 1, 0}},
                                         {{0, 1, 0}, {0, 1, 1}}, {{0, 0, 1}, {1, 0, 1}}, {{0, 0, 1}, {0, 1, 1}}};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                for (const auto& c : kCorner) {
                    tets.push_back(at(i, j, k));
                    tets.push_back(at(i + c[0][0], j + c[0][1], k + c[0][2]));
                    tets.push_back(at(i + c[1][0], j + c[1][1], k + c[1][2]));
                    tets.push_back(at(i + 1, j + 1, k + 1));
                }
    return cmpmc_mesh3d_create(xyz.data(), static_cast<uint64_t>(n_vert), potentials, tets.data(), tets.size() / 4,
                               -1e30, 1e30, out);
}

int cmpmc_mesh2d_grid_vertices(double u_min, double u_max, double v_min, double v_max, int nu, int nv, double* uv) {
    if (!uv) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh2d_grid_vertices: null argument");
    nu = std::max(1, nu); nv = std::max(1, nv);
    size_t id = 0;
    for (int i = 0; i <= nu; ++i) {
        const double u = u_min + (u_max - u_min) * (static_cast<double>(i) / nu);
        for (int j = 0; j <= nv; ++j, ++id) {
            uv[2 * id] = u;
            uv[2 * id + 1] = v_min + (v_max - v_min) * (static_cast<double>(j) / nv);
        }
    }
    return CMPMC_OK;
}

int cmpmc_mesh2d_grid(double u_min, double u_max, double v_min, double v_max, int nu, int nv,
                      const double* potentials, int axisymmetric_rz, cmpmc_mesh** out) {
    if (!out || !potentials) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh2d_grid: null argument");
    nu = std::max(1, nu); nv = std::max(1, nv);
    const int64_t n_vert = static_cast<int64_t>(nu + 1) * (nv + 1);
    if (n_vert > INT32_MAX || static_cast<int64_t>(nu) * nv * 2 > INT32_MAX - 8)
        return fail(CMPMC_ERR_INVALID, "cmpmc_mesh2d_grid: %d x %d cells exceed the 32-bit indices", nu, nv);
    std::vector<double> uv(static_cast<size_t>(n_vert) * 2);
    cmpmc_mesh2d_grid_vertices(u_min, u_max, v_min, v_max, nu, nv, uv.data());
    // mesh_field.hpp:464-477: two triangles per cell along its (0,0)-(1,1) diagonal
    std::vector<int32_t> tris;
    tris.reserve(static_cast<size_t>(nu) * nv * 6);
    const int vv = nv + 1;
    for (int i = 0; i < nu; ++i)
        for (int j = 0; j < nv; ++j) {
            const int32_t v00 = i * vv + j, v10 = (i + 1) * vv + j, v01 = i * vv + j + 1, v11 = (i + 1) * vv + j + 1;
            const int32_t two[6] = {v00, v10, v11, v00, v01, v11};
            tris.insert(tris.end(), two, two + 6);
        }
    return cmpmc_mesh2d_create(uv.data(), static_cast<uint64_t>(n_vert), potentials, tris.data(), tris.size() / 3,
                               axisymmetric_rz, out);
}

// mesh_field.hpp:188-230: "x y z V" and "i0 i1 i2 i3" lines; lines starting with '#' or '%'
// and lines that do not parse are skipped.
int cmpmc_mesh3d_load(const char* coords_file, const char* tets_file, double coord_unit_m, double potential_unit_V,
                      cmpmc_mesh** out) {
    if (!coords_file || !tets_file || !out) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh3d_load: null argument");
    std::ifstream fc(coords_file), ft(tets_file);
    if (!fc) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh3d_load: cannot open %s", coords_file);
    if (!ft) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh3d_load: cannot open %s", tets_file);
    std::vector<double> xyz, pot;
    std::vector<int32_t> tets;
    std::string line;
    while (std::getline(fc, line)) {
        if (line.empty() || line[0] == '#' || line[0] == '%') continue;
        std::istringstream is(line);
        double x, y, z, v;
        if (is >> x >> y >> z >> v) {
            xyz.insert(xyz.end(), {x * coord_unit_m, y * coord_unit_m, z * coord_unit_m});
            pot.push_back(v * potential_unit_V);
        }
    }
    while (std::getline(ft, line)) {
        if (line.empty() || line[0] == '#' || line[0] == '%') continue;
        std::istringstream is(line);
        int a, b, c, d;
        if (is >> a >> b >> c >> d) tets.insert(tets.end(), {a, b, c, d});
    }
    return cmpmc_mesh3d_create(xyz.data(), pot.size(), pot.data(), tets.data(), tets.size() / 4, -1e30, 1e30, out);
}

int cmpmc_mesh_destroy(cmpmc_mesh* m) {
    delete m;
    return CMPMC_OK;
}

int cmpmc_mesh_info(const cmpmc_mesh* m, int* dim, int* axisymmetric_rz, uint64_t* n_vertices, uint64_t* n_cells) {
    if (!m) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh_info: null mesh");
    if (dim) *dim = m->dim;
    if (axisymmetric_rz) *axisymmetric_rz = m->axisymmetric_rz;
    if (n_vertices) *n_vertices = m->v.size();
    if (n_cells) *n_cells = static_cast<uint64_t>(m->n_cells());
    return CMPMC_OK;
}

int cmpmc_mesh_locate(const cmpmc_mesh* m, const double* p, int32_t* hint, double* bary) {
    if (!m || !p || !hint || !bary) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh_locate: null argument");
    return cmpmc_internal::mesh_locate(*m, p, hint, bary) ? 1 : 0;
}

int cmpmc_mesh_field(const cmpmc_mesh* m, const double p[3], int32_t* hint, double E[3]) {
    if (!m || !p || !hint || !E) return fail(CMPMC_ERR_INVALID, "cmpmc_mesh_field: null argument");
    double q[3], b[4];
    E[0] = E[1] = E[2] = 0.0;
    project(*m, p, q);
    if (!cmpmc_internal::mesh_locate(*m, q, hint, b)) return CMPMC_OK;
    const double* g = &m->grad[static_cast<size_t>(*hint) * m->dim];
    if (m->dim == 3) { E[0] = g[0]; E[1] = g[1]; E[2] = g[2]; }
    else if (!m->axisymmetric_rz) { E[0] = g[0]; E[1] = g[1]; }
    else if (q[0] > 1e-15) { E[0] = g[0] * (p[0] / q[0]); E[1] = g[0] * (p[1] / q[0]); E[2] = g[1]; }
    else E[2] = g[1];
    return CMPMC_OK;
}

double cmpmc_mesh_potential(const cmpmc_mesh* m, const double p[3], int32_t* hint) {
    if (!m || !p || !hint) return 0.0;
    double q[3], b[4];
    project(*m, p, q);
    if (!cmpmc_internal::mesh_locate(*m, q, hint, b)) return 0.0;
    const int nv = m->dim + 1;
    const int32_t* t = &m->cells[static_cast<size_t>(*hint) * nv];
    double s = 0.0;
    for (int i = 0; i < nv; ++i) s = (i == 0) ? b[0] * m->v[static_cast<size_t>(t[0])] : s + b[i] * m->v[static_cast<size_t>(t[i])];
    return s;
}

}  // extern "C"
