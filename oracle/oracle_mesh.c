/* CPU oracle (TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg,
 * never by the product) - electric fields on finite-element meshes, restated in C from
 * the reference:
 *   geometry/mesh_field.hpp:42-59    TriLinearMesh3D construction (potential clamp,
 *                                    neighbours, inverse edge matrices and fields)
 *   geometry/mesh_field.hpp:63-93    find_tetrahedron: barycentric walk from a hint,
 *                                    tolerance -1e-10, at most n_tets + 4 steps
 *   geometry/mesh_field.hpp:96-124   potential_at / field_at / at (thread-local hint)
 *   geometry/mesh_field.hpp:128-181  build_cartesian_grid (6 Kuhn tetrahedra per voxel)
 *   geometry/mesh_field.hpp:188-230  load_from_files ("x y z V" and "i0 i1 i2 i3" lines)
 *   geometry/mesh_field.hpp:240-329  bary_coord, fill_neighbors, fill_tinv_and_grad
 *   geometry/mesh_field.hpp:340-441  BiLinearMesh2D: the same on triangles, with the
 *                                    axisymmetric (r, z) projection and E_r rotated back
 *   geometry/mesh_field.hpp:444-478  build_grid (2 triangles per cell)
 *   geometry/mesh_field.hpp:506-579  2D neighbours, inverses and fields
 *   transport/engine.hpp:151-155     DetectorModel::electric_field_at (3D mesh, else 2D
 *                                    mesh, else the uniform field)
 * Cells keep their vertex indices sorted ascending (the reference sorts them in place
 * before building neighbours), so the last vertex is the reference vertex of a cell. */
#include "oracle_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static double mclamp(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

static void sort_small(int32_t* a, int n) {
    for (int i = 1; i < n; ++i) {
        const int32_t x = a[i];
        int j = i - 1;
        while (j >= 0 && a[j] > x) { a[j + 1] = a[j]; --j; }
        a[j + 1] = x;
    }
}

/* A facet (3D) or edge (2D) of one cell: its vertices (sorted), the cell and the slot of
 * the vertex opposite to it. */
typedef struct { int32_t v[3]; int32_t cell, slot; } facet_t;
static int facet_cmp(const void* pa, const void* pb) {
    const facet_t* a = (const facet_t*)pa;
    const facet_t* b = (const facet_t*)pb;
    for (int i = 0; i < 3; ++i)
        if (a->v[i] != b->v[i]) return a->v[i] < b->v[i] ? -1 : 1;
    if (a->cell != b->cell) return a->cell < b->cell ? -1 : 1;
    return a->slot - b->slot;
}

/* mesh_field.hpp:257-284 / 519-544: two cells sharing a facet are each other's neighbour
 * across the vertex opposite to it; hull facets keep -1. */
static void fill_neighbors(orc_mesh* m) {
    const int nv = m->dim + 1;
    const size_t n = m->n_cells;
    for (size_t i = 0; i < n * (size_t)nv; ++i) m->nbr[i] = -1;
    facet_t* f = (facet_t*)malloc(sizeof(facet_t) * (n * (size_t)nv + 1));
    size_t nf = 0;
    for (size_t c = 0; c < n; ++c) {
        const int32_t* t = m->cells + c * (size_t)nv;
        for (int s = 0; s < nv; ++s) {
            facet_t e;
            int k = 0;
            for (int i = 0; i < nv; ++i) if (i != s) e.v[k++] = t[i];
            if (nv == 3) e.v[2] = -1;
            e.cell = (int32_t)c;
            e.slot = s;
            f[nf++] = e;
        }
    }
    qsort(f, nf, sizeof(facet_t), facet_cmp);
    for (size_t i = 0; i + 1 < nf; ++i) {
        if (f[i].v[0] == f[i + 1].v[0] && f[i].v[1] == f[i + 1].v[1] && f[i].v[2] == f[i + 1].v[2]) {
            m->nbr[(size_t)f[i].cell * nv + f[i].slot] = f[i + 1].cell;
            m->nbr[(size_t)f[i + 1].cell * nv + f[i + 1].slot] = f[i].cell;
            ++i;
        }
    }
    free(f);
}

/* mesh_field.hpp:286-328: T = [x0 - x3, x1 - x3, x2 - x3] by columns, its inverse by
 * cofactors, and the cell's field E = -(T^-1)^T (V0 - V3, V1 - V3, V2 - V3). Degenerate
 * cells (|det| <= 1e-30) get a zero inverse and a zero field. */
static void fill_tinv_grad3(orc_mesh* m) {
    for (size_t c = 0; c < m->n_cells; ++c) {
        const int32_t* t = m->cells + 4 * c;
        const double* x3 = m->x + 3 * (size_t)t[3];
        double col[3][3];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) col[j][i] = m->x[3 * (size_t)t[j] + i] - x3[i];
        const double a = col[0][0], b = col[1][0], cc = col[2][0];
        const double d = col[0][1], e = col[1][1], f = col[2][1];
        const double g = col[0][2], h = col[1][2], k = col[2][2];
        double* inv = m->tinv + 9 * c;
        double* E = m->grad + 3 * c;
        const double det = a * (e * k - f * h) - b * (d * k - f * g) + cc * (d * h - e * g);
        if (fabs(det) <= 1e-30) {
            memset(inv, 0, 9 * sizeof(double));
            E[0] = E[1] = E[2] = 0.0;
            continue;
        }
        const double id = 1.0 / det;
        inv[0] =  (e * k - f * h) * id;
        inv[1] = -(b * k - cc * h) * id;
        inv[2] =  (b * f - cc * e) * id;
        inv[3] = -(d * k - f * g) * id;
        inv[4] =  (a * k - cc * g) * id;
        inv[5] = -(a * f - cc * d) * id;
        inv[6] =  (d * h - e * g) * id;
        inv[7] = -(a * h - b * g) * id;
        inv[8] =  (a * e - b * d) * id;
        const double v3 = m->v[t[3]];
        const double dv0 = m->v[t[0]] - v3, dv1 = m->v[t[1]] - v3, dv2 = m->v[t[2]] - v3;
        E[0] = (inv[0] * dv0 + inv[3] * dv1 + inv[6] * dv2) * (-1.0);
        E[1] = (inv[1] * dv0 + inv[4] * dv1 + inv[7] * dv2) * (-1.0);
        E[2] = (inv[2] * dv0 + inv[5] * dv1 + inv[8] * dv2) * (-1.0);
    }
}

/* mesh_field.hpp:546-578 */
static void fill_tinv_grad2(orc_mesh* m) {
    for (size_t c = 0; c < m->n_cells; ++c) {
        const int32_t* t = m->cells + 3 * c;
        const double* x2 = m->x + 2 * (size_t)t[2];
        const double a = m->x[2 * (size_t)t[0]] - x2[0];
        const double b = m->x[2 * (size_t)t[1]] - x2[0];
        const double cc = m->x[2 * (size_t)t[0] + 1] - x2[1];
        const double d = m->x[2 * (size_t)t[1] + 1] - x2[1];
        double* inv = m->tinv + 4 * c;
        double* E = m->grad + 2 * c;
        const double det = a * d - b * cc;
        if (fabs(det) <= 1e-30) {
            inv[0] = inv[1] = inv[2] = inv[3] = 0.0;
            E[0] = E[1] = 0.0;
            continue;
        }
        const double id = 1.0 / det;
        inv[0] = d * id;
        inv[1] = -b * id;
        inv[2] = -cc * id;
        inv[3] = a * id;
        const double dv0 = m->v[t[0]] - m->v[t[2]];
        const double dv1 = m->v[t[1]] - m->v[t[2]];
        E[0] = -(inv[0] * dv0 + inv[2] * dv1);
        E[1] = -(inv[1] * dv0 + inv[3] * dv1);
    }
}

static orc_mesh* mesh_alloc(int dim, size_t nv, size_t nc) {
    orc_mesh* m = (orc_mesh*)calloc(1, sizeof(orc_mesh));
    m->dim = dim;
    m->n_vertices = nv;
    m->n_cells = nc;
    m->x = (double*)malloc(sizeof(double) * (size_t)dim * (nv + 1));
    m->v = (double*)malloc(sizeof(double) * (nv + 1));
    m->cells = (int32_t*)malloc(sizeof(int32_t) * (size_t)(dim + 1) * (nc + 1));
    m->nbr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(dim + 1) * (nc + 1));
    m->tinv = (double*)malloc(sizeof(double) * (size_t)(dim == 3 ? 9 : 4) * (nc + 1));
    m->grad = (double*)malloc(sizeof(double) * (size_t)dim * (nc + 1));
    return m;
}

orc_mesh* orc_mesh3d_new(const double* xyz, size_t nv, const double* V, const int32_t* tets, size_t nt,
                         double v_min_clamp, double v_max_clamp) {
    orc_mesh* m = mesh_alloc(3, nv, nt);
    memcpy(m->x, xyz, sizeof(double) * 3 * nv);
    for (size_t i = 0; i < nv; ++i) m->v[i] = mclamp(V[i], v_min_clamp, v_max_clamp);
    memcpy(m->cells, tets, sizeof(int32_t) * 4 * nt);
    for (size_t c = 0; c < nt; ++c) sort_small(m->cells + 4 * c, 4);
    fill_neighbors(m);
    fill_tinv_grad3(m);
    return m;
}

orc_mesh* orc_mesh2d_new(const double* uv, size_t nv, const double* V, const int32_t* tris, size_t nt,
                         int axisymmetric_rz) {
    orc_mesh* m = mesh_alloc(2, nv, nt);
    m->axisymmetric_rz = axisymmetric_rz ? 1 : 0;
    memcpy(m->x, uv, sizeof(double) * 2 * nv);
    memcpy(m->v, V, sizeof(double) * nv);
    memcpy(m->cells, tris, sizeof(int32_t) * 3 * nt);
    for (size_t c = 0; c < nt; ++c) sort_small(m->cells + 3 * c, 3);
    fill_neighbors(m);
    fill_tinv_grad2(m);
    return m;
}

void orc_mesh_free(orc_mesh* m) {
    if (!m) return;
    free(m->x); free(m->v); free(m->cells); free(m->nbr); free(m->tinv); free(m->grad);
    free(m);
}

static int imax1(int n) { return n < 1 ? 1 : n; }

/* mesh_field.hpp:133-151: vertex (i, j, k) at -h + 2 h (i / n) per axis, index (i vy + j) vz + k */
void orc_mesh3d_grid_vertices(double hx, double hy, double hz, int nx, int ny, int nz, double* xyz) {
    nx = imax1(nx); ny = imax1(ny); nz = imax1(nz);
    size_t id = 0;
    for (int i = 0; i <= nx; ++i) {
        const double x = -hx + 2.0 * hx * ((double)i / nx);
        for (int j = 0; j <= ny; ++j) {
            const double y = -hy + 2.0 * hy * ((double)j / ny);
            for (int k = 0; k <= nz; ++k) {
                const double z = -hz + 2.0 * hz * ((double)k / nz);
                xyz[3 * id] = x; xyz[3 * id + 1] = y; xyz[3 * id + 2] = z;
                ++id;
            }
        }
    }
}

/* mesh_field.hpp:154-180: every voxel split into 6 tetrahedra around its v000-v111 diagonal */
orc_mesh* orc_mesh3d_grid(double hx, double hy, double hz, int nx, int ny, int nz, const double* V) {
    nx = imax1(nx); ny = imax1(ny); nz = imax1(nz);
    const int vy = ny + 1, vz = nz + 1;
    const size_t nv = (size_t)(nx + 1) * vy * vz;
    double* xyz = (double*)malloc(sizeof(double) * 3 * nv);
    orc_mesh3d_grid_vertices(hx, hy, hz, nx, ny, nz, xyz);
    const size_t nt = (size_t)nx * ny * nz * 6;
    int32_t* tets = (int32_t*)malloc(sizeof(int32_t) * 4 * nt);
    size_t t = 0;
#define VI(i, j, k) ((int32_t)(((i) * vy + (j)) * vz + (k)))
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int32_t v000 = VI(i, j, k), v100 = VI(i + 1, j, k), v010 = VI(i, j + 1, k),
                              v110 = VI(i + 1, j + 1, k), v001 = VI(i, j, k + 1), v101 = VI(i + 1, j, k + 1),
                              v011 = VI(i, j + 1, k + 1), v111 = VI(i + 1, j + 1, k + 1);
                const int32_t six[6][4] = {{v000, v100, v110, v111}, {v000, v100, v101, v111},
                                           {v000, v010, v110, v111}, {v000, v010, v011, v111},
                                           {v000, v001, v101, v111}, {v000, v001, v011, v111}};
                memcpy(tets + 4 * t, six, sizeof(six));
                t += 6;
            }
#undef VI
    orc_mesh* m = orc_mesh3d_new(xyz, nv, V, tets, nt, -1e30, 1e30);
    free(xyz);
    free(tets);
    return m;
}

/* mesh_field.hpp:448-462 */
void orc_mesh2d_grid_vertices(double u_min, double u_max, double v_min, double v_max, int nu, int nv,
                              double* uv) {
    nu = imax1(nu); nv = imax1(nv);
    size_t id = 0;
    for (int i = 0; i <= nu; ++i) {
        const double u = u_min + (u_max - u_min) * ((double)i / nu);
        for (int j = 0; j <= nv; ++j) {
            uv[2 * id] = u;
            uv[2 * id + 1] = v_min + (v_max - v_min) * ((double)j / nv);
            ++id;
        }
    }
}

/* mesh_field.hpp:464-477: every cell split along its (0,0)-(1,1) diagonal */
orc_mesh* orc_mesh2d_grid(double u_min, double u_max, double v_min, double v_max, int nu, int nv,
                          const double* V, int axisymmetric_rz) {
    nu = imax1(nu); nv = imax1(nv);
    const int vv = nv + 1;
    const size_t nvert = (size_t)(nu + 1) * vv;
    double* uv = (double*)malloc(sizeof(double) * 2 * nvert);
    orc_mesh2d_grid_vertices(u_min, u_max, v_min, v_max, nu, nv, uv);
    const size_t nt = (size_t)nu * nv * 2;
    int32_t* tris = (int32_t*)malloc(sizeof(int32_t) * 3 * nt);
    size_t t = 0;
    for (int i = 0; i < nu; ++i)
        for (int j = 0; j < nv; ++j) {
            const int32_t v00 = i * vv + j, v10 = (i + 1) * vv + j, v01 = i * vv + j + 1, v11 = (i + 1) * vv + j + 1;
            tris[3 * t] = v00; tris[3 * t + 1] = v10; tris[3 * t + 2] = v11; ++t;
            tris[3 * t] = v00; tris[3 * t + 1] = v01; tris[3 * t + 2] = v11; ++t;
        }
    orc_mesh* m = orc_mesh2d_new(uv, nvert, V, tris, nt, axisymmetric_rz);
    free(uv);
    free(tris);
    return m;
}

/* mesh_field.hpp:193-229. Returns NULL when a file cannot be opened. */
orc_mesh* orc_mesh3d_load(const char* coords_file, const char* tets_file, double coord_unit_m,
                          double potential_unit_V) {
    FILE* fc = fopen(coords_file, "r");
    if (!fc) return NULL;
    FILE* ft = fopen(tets_file, "r");
    if (!ft) { fclose(fc); return NULL; }
    size_t nv = 0, cap_v = 64, nt = 0, cap_t = 64;
    double* xyz = (double*)malloc(sizeof(double) * 3 * cap_v);
    double* V = (double*)malloc(sizeof(double) * cap_v);
    int32_t* tets = (int32_t*)malloc(sizeof(int32_t) * 4 * cap_t);
    char line[4096];
    while (fgets(line, sizeof(line), fc)) {
        if (line[0] == '\n' || line[0] == '\0' || line[0] == '#' || line[0] == '%') continue;
        double x, y, z, v;
        if (sscanf(line, "%lf %lf %lf %lf", &x, &y, &z, &v) == 4) {
            if (nv == cap_v) {
                cap_v *= 2;
                xyz = (double*)realloc(xyz, sizeof(double) * 3 * cap_v);
                V = (double*)realloc(V, sizeof(double) * cap_v);
            }
            xyz[3 * nv] = x * coord_unit_m; xyz[3 * nv + 1] = y * coord_unit_m; xyz[3 * nv + 2] = z * coord_unit_m;
            V[nv] = v * potential_unit_V;
            ++nv;
        }
    }
    while (fgets(line, sizeof(line), ft)) {
        if (line[0] == '\n' || line[0] == '\0' || line[0] == '#' || line[0] == '%') continue;
        int a, b, c, d;
        if (sscanf(line, "%d %d %d %d", &a, &b, &c, &d) == 4) {
            if (nt == cap_t) {
                cap_t *= 2;
                tets = (int32_t*)realloc(tets, sizeof(int32_t) * 4 * cap_t);
            }
            tets[4 * nt] = a; tets[4 * nt + 1] = b; tets[4 * nt + 2] = c; tets[4 * nt + 3] = d;
            ++nt;
        }
    }
    fclose(fc);
    fclose(ft);
    orc_mesh* m = orc_mesh3d_new(xyz, nv, V, tets, nt, -1e30, 1e30);
    free(xyz); free(V); free(tets);
    return m;
}

/* mesh_field.hpp:240-247 / 500-508 */
static void bary(const orc_mesh* m, const double* p, int32_t c, double* b) {
    if (m->dim == 3) {
        const double* x3 = m->x + 3 * (size_t)m->cells[4 * (size_t)c + 3];
        const double* T = m->tinv + 9 * (size_t)c;
        const double dx = p[0] - x3[0], dy = p[1] - x3[1], dz = p[2] - x3[2];
        b[0] = T[0] * dx + T[1] * dy + T[2] * dz;
        b[1] = T[3] * dx + T[4] * dy + T[5] * dz;
        b[2] = T[6] * dx + T[7] * dy + T[8] * dz;
        b[3] = 1.0 - b[0] - b[1] - b[2];
    } else {
        const double* x2 = m->x + 2 * (size_t)m->cells[3 * (size_t)c + 2];
        const double* T = m->tinv + 4 * (size_t)c;
        const double du = p[0] - x2[0], dv = p[1] - x2[1];
        b[0] = T[0] * du + T[1] * dv;
        b[1] = T[2] * du + T[3] * dv;
        b[2] = 1.0 - b[0] - b[1];
    }
}

/* mesh_field.hpp:63-93 / 358-387: walk across the facet opposite the most negative
 * barycentric coordinate until all are >= -1e-10 (inside), a hull facet is crossed
 * (outside) or n_cells + 4 steps are done. The hint is left at the last cell. */
int orc_mesh_locate(const orc_mesh* m, const double* p, int32_t* hint, double* b) {
    if (m->n_cells == 0) return 0;
    if (*hint < 0 || *hint >= (int32_t)m->n_cells) *hint = (int32_t)(m->n_cells / 2);
    const int nv = m->dim + 1;
    const int max_steps = (int)m->n_cells + 4;
    for (int step = 0; step < max_steps; ++step) {
        bary(m, p, *hint, b);
        double mb = b[0];
        int mi = 0;
        for (int i = 1; i < nv; ++i)
            if (b[i] < mb) { mb = b[i]; mi = i; }
        if (mb >= -1e-10) return 1;
        const int32_t nx = m->nbr[(size_t)(*hint) * nv + mi];
        if (nx < 0) return 0;
        *hint = nx;
    }
    return 0;
}

static void project(const orc_mesh* m, const double p[3], double q[2]) {
    if (m->axisymmetric_rz) { q[0] = hypot(p[0], p[1]); q[1] = p[2]; }
    else { q[0] = p[0]; q[1] = p[1]; }
}

/* mesh_field.hpp:96-104 / 389-396 (zero outside the mesh) */
double orc_mesh_potential(const orc_mesh* m, const double p[3], int32_t* hint) {
    double b[4], q[2];
    const double* pp = p;
    if (m->dim == 2) { project(m, p, q); pp = q; }
    if (!orc_mesh_locate(m, pp, hint, b)) return 0.0;
    const int32_t* t = m->cells + (size_t)(m->dim + 1) * (size_t)(*hint);
    if (m->dim == 3) return b[0] * m->v[t[0]] + b[1] * m->v[t[1]] + b[2] * m->v[t[2]] + b[3] * m->v[t[3]];
    return b[0] * m->v[t[0]] + b[1] * m->v[t[1]] + b[2] * m->v[t[2]];
}

/* mesh_field.hpp:107-111 / 398-418: the cell's constant field; in the axisymmetric case
 * E_r is rotated back to (x, y) (only E_z on the axis, r <= 1e-15); zero outside. */
void orc_mesh_field(const orc_mesh* m, const double p[3], int32_t* hint, double E[3]) {
    double b[4], q[2];
    E[0] = E[1] = E[2] = 0.0;
    if (m->dim == 3) {
        if (!orc_mesh_locate(m, p, hint, b)) return;
        const double* g = m->grad + 3 * (size_t)(*hint);
        E[0] = g[0]; E[1] = g[1]; E[2] = g[2];
        return;
    }
    project(m, p, q);
    if (!orc_mesh_locate(m, q, hint, b)) return;
    const double* g = m->grad + 2 * (size_t)(*hint);
    if (!m->axisymmetric_rz) { E[0] = g[0]; E[1] = g[1]; return; }
    const double r = q[0];
    if (r > 1e-15) { E[0] = g[0] * (p[0] / r); E[1] = g[0] * (p[1] / r); E[2] = g[1]; }
    else E[2] = g[1];
}

/* mesh_field.hpp:114-117 / 420-423: at() keeps one hint per thread for each mesh class */
static _Thread_local int32_t t_hint3 = 0;
static _Thread_local int32_t t_hint2 = 0;

void orc_mesh_at(const orc_mesh* m, const double p[3], double E[3]) {
    orc_mesh_field(m, p, m->dim == 3 ? &t_hint3 : &t_hint2, E);
}

/* engine.hpp:151-155 */
void orc_field_at(const orc_detector* d, const double p[3], double E[3]) {
    if (d->mesh) { orc_mesh_at(d->mesh, p, E); return; }
    E[0] = d->efield[0]; E[1] = d->efield[1]; E[2] = d->efield[2];
}
