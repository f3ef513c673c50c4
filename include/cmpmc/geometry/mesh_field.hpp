// Drop-in FEM mesh-field API (geometry/mesh_field.hpp of the reference): TriLinearMesh3D
// (tetrahedra) and BiLinearMesh2D (triangles, planar or axisymmetric (r, z)) with the
// reference's constructors, grid builders, file loader and walk / field / potential queries.
// The mesh lives in the library (cmpmc_mesh: neighbours, inverse edge matrices and cell
// fields built there, bit for bit the reference's); copies of these objects share it, as
// the reference's value copies share their immutable tables in effect. A DetectorModel
// holding one hands it to the model, which copies it to the device, where the charge step
// looks the field up once per step as the reference's engine does.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "cmpmc/core/core.hpp"
#include "cmpmc_amd.h"

namespace cmpmc::geometry {

namespace detail {
inline std::shared_ptr<cmpmc_mesh> own_mesh(cmpmc_mesh* m) {
    return std::shared_ptr<cmpmc_mesh>(m, [](cmpmc_mesh* p) { cmpmc_mesh_destroy(p); });
}
inline uint64_t mesh_cells(const cmpmc_mesh* m) {
    uint64_t n = 0;
    if (m) cmpmc_mesh_info(m, nullptr, nullptr, nullptr, &n);
    return n;
}
inline uint64_t mesh_vertices(const cmpmc_mesh* m) {
    uint64_t n = 0;
    if (m) cmpmc_mesh_info(m, nullptr, nullptr, &n, nullptr);
    return n;
}
}  // namespace detail

// mesh_field.hpp:42-330
class TriLinearMesh3D {
public:
    using Tet = std::array<int, 4>;

    TriLinearMesh3D() = default;

    TriLinearMesh3D(std::vector<Vec3> vertices, std::vector<double> potentials, std::vector<Tet> tetrahedra,
                    double V_min_clamp = -1e30, double V_max_clamp = 1e30) {
        std::vector<double> xyz;
        xyz.reserve(vertices.size() * 3);
        for (const Vec3& v : vertices) xyz.insert(xyz.end(), {v.x, v.y, v.z});
        std::vector<int32_t> t;
        t.reserve(tetrahedra.size() * 4);
        for (const Tet& q : tetrahedra) t.insert(t.end(), {q[0], q[1], q[2], q[3]});
        cmpmc_mesh* m = nullptr;
        check_status(cmpmc_mesh3d_create(xyz.data(), vertices.size(), potentials.data(), t.data(), tetrahedra.size(),
                                         V_min_clamp, V_max_clamp, &m));
        mesh_ = detail::own_mesh(m);
    }

    bool empty() const noexcept { return detail::mesh_cells(mesh_.get()) == 0; }
    size_t num_vertices() const noexcept { return detail::mesh_vertices(mesh_.get()); }
    size_t num_tetrahedra() const noexcept { return detail::mesh_cells(mesh_.get()); }

    bool find_tetrahedron(const Vec3& p, int& tet_hint, std::array<double, 4>& bary) const noexcept {
        if (!mesh_) return false;
        const double q[3] = {p.x, p.y, p.z};
        int32_t h = tet_hint;
        const int r = cmpmc_mesh_locate(mesh_.get(), q, &h, bary.data());
        tet_hint = h;
        return r == 1;
    }
    double potential_at(const Vec3& p, int& tet_hint) const noexcept {
        if (!mesh_) return 0.0;
        const double q[3] = {p.x, p.y, p.z};
        int32_t h = tet_hint;
        const double v = cmpmc_mesh_potential(mesh_.get(), q, &h);
        tet_hint = h;
        return v;
    }
    Vec3 field_at(const Vec3& p, int& tet_hint) const noexcept {
        if (!mesh_) return {0.0, 0.0, 0.0};
        const double q[3] = {p.x, p.y, p.z};
        double E[3];
        int32_t h = tet_hint;
        cmpmc_mesh_field(mesh_.get(), q, &h, E);
        tet_hint = h;
        return {E[0], E[1], E[2]};
    }
    // mesh_field.hpp:114-123: one walk hint per thread
    Vec3 at(const Vec3& p) const noexcept {
        thread_local int cached_tet = 0;
        return field_at(p, cached_tet);
    }
    double potential_at(const Vec3& p) const noexcept {
        thread_local int cached_tet = 0;
        return potential_at(p, cached_tet);
    }

    // mesh_field.hpp:128-181: 6 tetrahedra per voxel of the box [-h, h] per axis
    template <typename PotentialFn>
    static TriLinearMesh3D build_cartesian_grid(double half_x, double half_y, double half_z, int nx, int ny, int nz,
                                                PotentialFn&& potential_fn) {
        nx = nx < 1 ? 1 : nx;
        ny = ny < 1 ? 1 : ny;
        nz = nz < 1 ? 1 : nz;
        const size_t nv = static_cast<size_t>(nx + 1) * (ny + 1) * (nz + 1);
        std::vector<double> xyz(nv * 3), pot(nv);
        check_status(cmpmc_mesh3d_grid_vertices(half_x, half_y, half_z, nx, ny, nz, xyz.data()));
        for (size_t i = 0; i < nv; ++i) pot[i] = potential_fn(Vec3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
        cmpmc_mesh* m = nullptr;
        check_status(cmpmc_mesh3d_grid(half_x, half_y, half_z, nx, ny, nz, pot.data(), &m));
        TriLinearMesh3D out;
        out.mesh_ = detail::own_mesh(m);
        return out;
    }

    // mesh_field.hpp:188-230
    static TriLinearMesh3D load_from_files(const std::string& coords_file, const std::string& tets_file,
                                           double coord_unit_m = 1.0, double potential_unit_V = 1.0) {
        cmpmc_mesh* m = nullptr;
        check_status(cmpmc_mesh3d_load(coords_file.c_str(), tets_file.c_str(), coord_unit_m, potential_unit_V, &m));
        TriLinearMesh3D out;
        out.mesh_ = detail::own_mesh(m);
        return out;
    }

    const cmpmc_mesh* handle() const noexcept { return mesh_.get(); }

private:
    std::shared_ptr<cmpmc_mesh> mesh_;
};

// mesh_field.hpp:340-579
class BiLinearMesh2D {
public:
    using Tri = std::array<int, 3>;
    struct Point2D {
        double u = 0.0;   // r (or x)
        double v = 0.0;   // z (or y)
    };

    BiLinearMesh2D() = default;

    BiLinearMesh2D(std::vector<Point2D> vertices, std::vector<double> potentials, std::vector<Tri> triangles,
                   bool axisymmetric_rz = true) {
        std::vector<double> uv;
        uv.reserve(vertices.size() * 2);
        for (const Point2D& v : vertices) uv.insert(uv.end(), {v.u, v.v});
        std::vector<int32_t> t;
        t.reserve(triangles.size() * 3);
        for (const Tri& q : triangles) t.insert(t.end(), {q[0], q[1], q[2]});
        cmpmc_mesh* m = nullptr;
        check_status(cmpmc_mesh2d_create(uv.data(), vertices.size(), potentials.data(), t.data(), triangles.size(),
                                         axisymmetric_rz ? 1 : 0, &m));
        mesh_ = detail::own_mesh(m);
    }

    bool empty() const noexcept { return detail::mesh_cells(mesh_.get()) == 0; }
    size_t num_vertices() const noexcept { return detail::mesh_vertices(mesh_.get()); }
    size_t num_triangles() const noexcept { return detail::mesh_cells(mesh_.get()); }

    bool find_triangle(Point2D p, int& tri_hint, std::array<double, 3>& bary) const noexcept {
        if (!mesh_) return false;
        const double q[2] = {p.u, p.v};
        int32_t h = tri_hint;
        const int r = cmpmc_mesh_locate(mesh_.get(), q, &h, bary.data());
        tri_hint = h;
        return r == 1;
    }
    double potential_at(const Vec3& p, int& tri_hint) const noexcept {
        if (!mesh_) return 0.0;
        const double q[3] = {p.x, p.y, p.z};
        int32_t h = tri_hint;
        const double v = cmpmc_mesh_potential(mesh_.get(), q, &h);
        tri_hint = h;
        return v;
    }
    Vec3 field_at(const Vec3& p, int& tri_hint) const noexcept {
        if (!mesh_) return {0.0, 0.0, 0.0};
        const double q[3] = {p.x, p.y, p.z};
        double E[3];
        int32_t h = tri_hint;
        cmpmc_mesh_field(mesh_.get(), q, &h, E);
        tri_hint = h;
        return {E[0], E[1], E[2]};
    }
    Vec3 at(const Vec3& p) const noexcept {
        thread_local int cached_tri = 0;
        return field_at(p, cached_tri);
    }
    double potential_at(const Vec3& p) const noexcept {
        thread_local int cached_tri = 0;
        return potential_at(p, cached_tri);
    }

    // mesh_field.hpp:444-478: 2 triangles per cell of [u_min, u_max] x [v_min, v_max]
    template <typename PotentialFn2D>
    static BiLinearMesh2D build_grid(double u_min, double u_max, double v_min, double v_max, int nu, int nv,
                                     PotentialFn2D&& potential_fn, bool axisymmetric_rz = true) {
        nu = nu < 1 ? 1 : nu;
        nv = nv < 1 ? 1 : nv;
        const size_t n = static_cast<size_t>(nu + 1) * (nv + 1);
        std::vector<double> uv(n * 2), pot(n);
        check_status(cmpmc_mesh2d_grid_vertices(u_min, u_max, v_min, v_max, nu, nv, uv.data()));
        for (size_t i = 0; i < n; ++i) pot[i] = potential_fn(uv[2 * i], uv[2 * i + 1]);
        cmpmc_mesh* m = nullptr;
        check_status(cmpmc_mesh2d_grid(u_min, u_max, v_min, v_max, nu, nv, pot.data(), axisymmetric_rz ? 1 : 0, &m));
        BiLinearMesh2D out;
        out.mesh_ = detail::own_mesh(m);
        return out;
    }

    const cmpmc_mesh* handle() const noexcept { return mesh_.get(); }

private:
    std::shared_ptr<cmpmc_mesh> mesh_;
};

}  // namespace cmpmc::geometry
