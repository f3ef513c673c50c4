// FEM mesh fields on the device (geometry/mesh_field.hpp of the reference): the cell that
// holds a carrier, found by the barycentric walk of find_tetrahedron / find_triangle
// (mesh_field.hpp:63-93, 358-387), and that cell's constant field (field_at,
// mesh_field.hpp:107-111, 398-418).
//
// The walk starts from the carrier's cell of its previous step (a new carrier: the cell of
// its seed box). The reference starts from the last cell its thread found, whatever carrier
// that was; for a point inside one cell the walk ends in that cell from any start, so the
// two agree except for points within the tolerance of a cell boundary. The FP64 form walks
// with the reference's tolerance (-1e-10) and step limit (cells + 4) on the reference's
// operations; the FP32 form walks on FP32 records with a tolerance sized from the mesh
// (DevMesh::tol_f: the FP32 rounding of a barycentric coordinate, so that of two cells
// sharing a face at least one accepts a point on it) and at most 4096 steps.
#pragma once

#include "device_common.hpp"

namespace cmpmc_dev {

constexpr int kMeshWalkCap32 = 4096;

// The field at a carrier's position (DetectorModel::electric_field_at, engine.hpp:151-155)
// with the field-dependent intervalley rate there (iv < 0: to be computed from E) and, in
// the FP32 form, |E|.
template <class P> struct CField {
    V3<typename P::R> E;
    typename P::R iv, emag;
};

// The walk's start: the stored cell (1 + index) when valid, else the seed box's cell.
template <class R> __device__ inline int32_t mesh_start(const DevMesh& g, int32_t cell1, R q0, R q1, R q2) {
    if (cell1 >= 1 && cell1 <= g.n_cells) return cell1 - 1;
    int32_t ix[3];
    const R q[3] = {q0, q1, q2};
    for (int i = 0; i < 3; ++i) {
        R t;
        if constexpr (sizeof(R) == 8) t = (q[i] - R(g.lo[i])) * R(g.inv_h[i]);
        else t = (q[i] - g.lo_f[i]) * g.inv_h_f[i];
        // clamp in R before converting (t may be huge or NaN)
        t = (t >= R(0)) ? t : R(0);
        const R hi = R(g.seed_n[i] - 1);
        t = (t <= hi) ? t : hi;
        ix[i] = static_cast<int32_t>(t);
    }
    return g.seed[(static_cast<int64_t>(ix[0]) * g.seed_n[1] + ix[1]) * g.seed_n[2] + ix[2]];
}

// dim 3, FP64 records: mesh_field.hpp:240-247 (bary_coord) and 63-93 (walk)
__device__ inline bool mesh_walk3(const DevMesh& g, double px, double py, double pz, int32_t& c) {
    const int64_t max_steps = static_cast<int64_t>(g.n_cells) + 4;
    for (int64_t step = 0; step < max_steps; ++step) {
        const double2* r = reinterpret_cast<const double2*>(g.rec) + static_cast<int64_t>(c) * 8;
        const double2 t01 = r[0], t23 = r[1], t45 = r[2], t67 = r[3], t8x = r[4], yz = r[5];
        const double dx = px - t8x.y, dy = py - yz.x, dz = pz - yz.y;
        const double b0 = t01.x * dx + t01.y * dy + t23.x * dz;
        const double b1 = t23.y * dx + t45.x * dy + t45.y * dz;
        const double b2 = t67.x * dx + t67.y * dy + t8x.x * dz;
        const double b3 = 1.0 - b0 - b1 - b2;
        int mi = 0;
        double mb = b0;
        if (b1 < mb) { mb = b1; mi = 1; }
        if (b2 < mb) { mb = b2; mi = 2; }
        if (b3 < mb) { mb = b3; mi = 3; }
        if (mb >= -1e-10) return true;
        const int32_t next = g.nbr[static_cast<int64_t>(c) * 4 + mi];
        if (next < 0) return false;
        c = next;
    }
    return false;
}

// dim 2, FP64 records: mesh_field.hpp:500-508 and 358-387
__device__ inline bool mesh_walk2(const DevMesh& g, double pu, double pv, int32_t& c) {
    const int64_t max_steps = static_cast<int64_t>(g.n_cells) + 4;
    for (int64_t step = 0; step < max_steps; ++step) {
        const double2* r = reinterpret_cast<const double2*>(g.rec) + static_cast<int64_t>(c) * 4;
        const double2 t01 = r[0], t23 = r[1], x = r[2];
        const double du = pu - x.x, dv = pv - x.y;
        const double b0 = t01.x * du + t01.y * dv;
        const double b1 = t23.x * du + t23.y * dv;
        const double b2 = 1.0 - b0 - b1;
        int mi = 0;
        double mb = b0;
        if (b1 < mb) { mb = b1; mi = 1; }
        if (b2 < mb) { mb = b2; mi = 2; }
        if (mb >= -1e-10) return true;
        const int32_t next = g.nbr[static_cast<int64_t>(c) * 4 + mi];
        if (next < 0) return false;
        c = next;
    }
    return false;
}

// dim 3, FP32 records [t0..t3 | t4..t7 | t8 x y z | Ex Ey Ez iv]
__device__ inline bool mesh_walk3f(const DevMesh& g, float px, float py, float pz, int32_t& c) {
    const float tol = -g.tol_f;
    const int32_t cap = (g.n_cells < kMeshWalkCap32 - 4) ? g.n_cells + 4 : kMeshWalkCap32;
    for (int32_t step = 0; step < cap; ++step) {
        const float4* r = reinterpret_cast<const float4*>(g.rec_f) + static_cast<int64_t>(c) * 4;
        const float4 a = r[0], b = r[1], d = r[2];
        const float dx = px - d.y, dy = py - d.z, dz = pz - d.w;
        const float b0 = a.x * dx + a.y * dy + a.z * dz;
        const float b1 = a.w * dx + b.x * dy + b.y * dz;
        const float b2 = b.z * dx + b.w * dy + d.x * dz;
        const float b3 = 1.0f - b0 - b1 - b2;
        int mi = 0;
        float mb = b0;
        if (b1 < mb) { mb = b1; mi = 1; }
        if (b2 < mb) { mb = b2; mi = 2; }
        if (b3 < mb) { mb = b3; mi = 3; }
        if (mb >= tol) return true;
        const int32_t next = g.nbr[static_cast<int64_t>(c) * 4 + mi];
        if (next < 0) return false;
        c = next;
    }
    return false;
}

// dim 2, FP32 records [t0..t3 | x y Eu Ev | iv |E| 0 0]
__device__ inline bool mesh_walk2f(const DevMesh& g, float pu, float pv, int32_t& c) {
    const float tol = -g.tol_f;
    const int32_t cap = (g.n_cells < kMeshWalkCap32 - 4) ? g.n_cells + 4 : kMeshWalkCap32;
    for (int32_t step = 0; step < cap; ++step) {
        const float4* r = reinterpret_cast<const float4*>(g.rec_f) + static_cast<int64_t>(c) * 3;
        const float4 a = r[0], b = r[1];
        const float du = pu - b.x, dv = pv - b.y;
        const float b0 = a.x * du + a.y * dv;
        const float b1 = a.z * du + a.w * dv;
        const float b2 = 1.0f - b0 - b1;
        int mi = 0;
        float mb = b0;
        if (b1 < mb) { mb = b1; mi = 1; }
        if (b2 < mb) { mb = b2; mi = 2; }
        if (mb >= tol) return true;
        const int32_t next = g.nbr[static_cast<int64_t>(c) * 4 + mi];
        if (next < 0) return false;
        c = next;
    }
    return false;
}

// The field at pos and the carrier's cell (cell1 = 1 + index, updated; inside: the walk
// ended inside the mesh). Outside the mesh the field is zero (mesh_field.hpp:109, 401).
template <class P>
__device__ inline CField<P> mesh_field(const DevModel& m, const V3<typename P::R>& pos, int32_t& cell1, bool& inside) {
    using R = typename P::R;
    const DevMesh& g = m.mesh;
    CField<P> f;
    f.E = V3<R>{R(0), R(0), R(0)};
    f.emag = R(0);
    if constexpr (P::exact) f.iv = R(g.iv_zero);
    else f.iv = g.iv_zero_f;
    inside = false;
    if (g.n_cells <= 0) return f;
    R q0, q1, q2 = R(0);
    if (g.dim == 3) { q0 = pos.x; q1 = pos.y; q2 = pos.z; }
    else if (g.axisym) { q0 = P::hypot_(pos.x, pos.y); q1 = pos.z; }
    else { q0 = pos.x; q1 = pos.y; }
    int32_t c = mesh_start<R>(g, cell1, q0, q1, q2);
    if constexpr (P::exact) inside = (g.dim == 3) ? mesh_walk3(g, q0, q1, q2, c) : mesh_walk2(g, q0, q1, c);
    else inside = (g.dim == 3) ? mesh_walk3f(g, q0, q1, q2, c) : mesh_walk2f(g, q0, q1, c);
    cell1 = c + 1;
    if (!inside) return f;
    if constexpr (P::exact) {
        if (g.dim == 3) {
            const double2* r = reinterpret_cast<const double2*>(g.rec) + static_cast<int64_t>(c) * 8;
            const double2 e01 = r[6], e2i = r[7];
            f.E = V3<R>{e01.x, e01.y, e2i.x};
            f.iv = e2i.y;
            return f;
        }
        const double2 gu = reinterpret_cast<const double2*>(g.rec)[static_cast<int64_t>(c) * 4 + 3];
        f.iv = R(-1);                      // from E at the position (it varies within the cell)
        if (!g.axisym) f.E = V3<R>{gu.x, gu.y, R(0)};
        else if (q0 > R(1e-15)) f.E = V3<R>{gu.x * (pos.x / q0), gu.x * (pos.y / q0), gu.y};
        else f.E = V3<R>{R(0), R(0), gu.y};
        return f;
    } else {
        if (g.dim == 3) {
            const float4 e = reinterpret_cast<const float4*>(g.rec_f)[static_cast<int64_t>(c) * 4 + 3];
            f.E = V3<R>{e.x, e.y, e.z};
            f.iv = e.w;
            f.emag = P::sqrt_(e.x * e.x + e.y * e.y + e.z * e.z);
            return f;
        }
        const float4* r = reinterpret_cast<const float4*>(g.rec_f) + static_cast<int64_t>(c) * 3;
        const float4 b = r[1], e = r[2];
        f.iv = e.x;
        f.emag = e.y;
        if (!g.axisym) f.E = V3<R>{b.z, b.w, R(0)};
        else if (q0 > R(1e-15)) {
            const R s = P::rcp(q0);
            f.E = V3<R>{b.z * (pos.x * s), b.z * (pos.y * s), b.w};
        } else f.E = V3<R>{R(0), R(0), b.w};
        return f;
    }
}

template <class P> __device__ inline CField<P> mesh_field(const DevModel& m, const V3<typename P::R>& pos, int32_t& cell1) {
    bool inside;
    return mesh_field<P>(m, pos, cell1, inside);
}

// The field evaluation of the charge step at n points, for tests and tools: the walk from
// start[i] (a cell index; out of range: cell n_cells / 2, the reference's rule; start NULL:
// the seed box of the point). Without a mesh: the uniform field, inside 1, cell -1.
template <class P>
__global__ void mesh_query_kernel(const DevModel m, uint64_t n, const double* pts, const int32_t* start,
                                  int32_t* inside, int32_t* cell, double* E) {
    using R = typename P::R;
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (m.mesh.dim == 0) {
        inside[i] = 1;
        cell[i] = -1;
        for (int k = 0; k < 3; ++k) E[3 * i + k] = m.efield[k];
        return;
    }
    const V3<R> p{R(pts[3 * i]), R(pts[3 * i + 1]), R(pts[3 * i + 2])};
    int32_t c1 = 0;
    if (start) {
        const int32_t s = start[i];
        c1 = (s >= 0 && s < m.mesh.n_cells) ? s + 1 : m.mesh.n_cells / 2 + 1;
    }
    bool in;
    const CField<P> f = mesh_field<P>(m, p, c1, in);
    inside[i] = in ? 1 : 0;
    cell[i] = c1 - 1;
    E[3 * i] = static_cast<double>(f.E.x);
    E[3 * i + 1] = static_cast<double>(f.E.y);
    E[3 * i + 2] = static_cast<double>(f.E.z);
}

}  // namespace cmpmc_dev
