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

// A walk's result besides the cell: the distance the point can move before one of its
// barycentric coordinates may fall to `margin` (face i is (lambda_i - margin) / |grad
// lambda_i| away in that sense; negative when some lambda_i <= margin).
template <class R> struct WalkOut {
    R safe;
};
template <class R> __device__ inline R face_room(R lam, R n2, R margin) {
    if (!(lam > margin) || !(n2 > R(0))) return R(-1);
    if constexpr (sizeof(R) == 8) return (lam - margin) / sqrt(n2);
    else return (lam - margin) * __builtin_amdgcn_rsqf(n2);
}

// dim 3, FP64 records: mesh_field.hpp:240-247 (bary_coord) and 63-93 (walk)
__device__ inline bool mesh_walk3(const DevMesh& g, double px, double py, double pz, int32_t& c, WalkOut<double>& w,
                                  double margin) {
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
        if (mb >= -1e-10) {
            const double sx = t01.x + t23.y + t67.x, sy = t01.y + t45.x + t67.y, sz = t23.x + t45.y + t8x.x;
            const double n0 = t01.x * t01.x + t01.y * t01.y + t23.x * t23.x;
            const double n1 = t23.y * t23.y + t45.x * t45.x + t45.y * t45.y;
            const double n2 = t67.x * t67.x + t67.y * t67.y + t8x.x * t8x.x;
            const double n3 = sx * sx + sy * sy + sz * sz;
            w.safe = tmin(tmin(face_room(b0, n0, margin), face_room(b1, n1, margin)),
                          tmin(face_room(b2, n2, margin), face_room(b3, n3, margin)));
            return true;
        }
        const int32_t next = g.nbr[static_cast<int64_t>(c) * 4 + mi];
        if (next < 0) return false;
        c = next;
    }
    return false;
}

// dim 2, FP64 records: mesh_field.hpp:500-508 and 358-387
__device__ inline bool mesh_walk2(const DevMesh& g, double pu, double pv, int32_t& c, WalkOut<double>& w,
                                  double margin) {
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
        if (mb >= -1e-10) {
            const double su = t01.x + t23.x, sv = t01.y + t23.y;
            w.safe = tmin(tmin(face_room(b0, t01.x * t01.x + t01.y * t01.y, margin),
                               face_room(b1, t23.x * t23.x + t23.y * t23.y, margin)),
                          face_room(b2, su * su + sv * sv, margin));
            return true;
        }
        const int32_t next = g.nbr[static_cast<int64_t>(c) * 4 + mi];
        if (next < 0) return false;
        c = next;
    }
    return false;
}

// dim 3, FP32 records [t0..t3 | t4..t7 | t8 x y z | Ex Ey Ez iv]
__device__ inline bool mesh_walk3f(const DevMesh& g, float px, float py, float pz, int32_t& c, WalkOut<float>& w,
                                   float margin) {
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
        if (mb >= tol) {
            const float sx = a.x + a.w + b.z, sy = a.y + b.x + b.w, sz = a.z + b.y + d.x;
            const float n0 = a.x * a.x + a.y * a.y + a.z * a.z, n1 = a.w * a.w + b.x * b.x + b.y * b.y;
            const float n2 = b.z * b.z + b.w * b.w + d.x * d.x, n3 = sx * sx + sy * sy + sz * sz;
            w.safe = tmin(tmin(face_room(b0, n0, margin), face_room(b1, n1, margin)),
                          tmin(face_room(b2, n2, margin), face_room(b3, n3, margin)));
            return true;
        }
        const int32_t next = g.nbr[static_cast<int64_t>(c) * 4 + mi];
        if (next < 0) return false;
        c = next;
    }
    return false;
}

// dim 2, FP32 records [t0..t3 | x y Eu Ev | iv |E| 0 0]
__device__ inline bool mesh_walk2f(const DevMesh& g, float pu, float pv, int32_t& c, WalkOut<float>& w,
                                   float margin) {
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
        if (mb >= tol) {
            const float su = a.x + a.z, sv = a.y + a.w;
            w.safe = tmin(tmin(face_room(b0, a.x * a.x + a.y * a.y, margin), face_room(b1, a.z * a.z + a.w * a.w, margin)),
                          face_room(b2, su * su + sv * sv, margin));
            return true;
        }
        const int32_t next = g.nbr[static_cast<int64_t>(c) * 4 + mi];
        if (next < 0) return false;
        c = next;
    }
    return false;
}

// A carrier's view of the mesh between walks: `safe`, the distance it can still move
// without leaving its cell (the cell itself is CTrack::cell). The walk sets safe =
// min_i (lambda_i - margin) / |grad lambda_i| (each barycentric coordinate falls by at most
// |grad lambda_i| per metre moved) and every step subtracts the distance moved; while it
// is positive the point is inside the same cell by more than the margin (1e-9 in the FP64
// form, twice the walk tolerance in the FP32 form), where every walk (from any start)
// ends in that cell, so the walk is skipped and the cell's field read directly. Outside
// the mesh, or within the margin of a face, the next step walks again. (Only the distance
// is kept in registers: the field is re-read from the cell's record, which stays in cache.)
template <class P> struct MeshCache {
    typename P::R safe;
};

template <class P> __device__ inline void mesh_cache_reset(MeshCache<P>& mc) { mc.safe = typename P::R(-1); }

// The walk from the carrier's cell (cell1 = 1 + index, updated) and its new safe distance.
// Returns whether the point is inside the mesh.
template <class P>
__device__ inline bool mesh_lookup(const DevModel& m, const V3<typename P::R>& pos, int32_t& cell1, MeshCache<P>& mc) {
    using R = typename P::R;
    const DevMesh& g = m.mesh;
    mc.safe = R(-1);
    if (g.n_cells <= 0) return false;
    R q0, q1, q2 = R(0);
    if (g.dim == 3) { q0 = pos.x; q1 = pos.y; q2 = pos.z; }
    else if (g.axisym) { q0 = P::hypot_(pos.x, pos.y); q1 = pos.z; }
    else { q0 = pos.x; q1 = pos.y; }
    int32_t c = mesh_start<R>(g, cell1, q0, q1, q2);
    WalkOut<R> w;
    const R margin = P::exact ? R(1e-9) : R(2) * R(g.tol_f);
    bool inside;
    if constexpr (P::exact) {
        inside = (g.dim == 3) ? mesh_walk3(g, q0, q1, q2, c, w, margin) : mesh_walk2(g, q0, q1, c, w, margin);
    } else {
        inside = (g.dim == 3) ? mesh_walk3f(g, q0, q1, q2, c, w, margin) : mesh_walk2f(g, q0, q1, c, w, margin);
    }
    cell1 = c + 1;
    if (inside) mc.safe = w.safe;
    return inside;
}

// The field at pos in cell c (mesh_field.hpp:107-111, 398-418), or zero outside the mesh.
template <class P> __device__ inline CField<P> mesh_cell_field(const DevModel& m, const V3<typename P::R>& pos,
                                                               int32_t c, bool inside) {
    using R = typename P::R;
    const DevMesh& g = m.mesh;
    CField<P> f;
    if (!inside) {
        f.E = V3<R>{R(0), R(0), R(0)};
        f.emag = R(0);
        if constexpr (P::exact) f.iv = R(g.iv_zero);
        else f.iv = g.iv_zero_f;
        return f;
    }
    R e0, e1;
    if constexpr (P::exact) {
        f.emag = R(0);
        if (g.dim == 3) {
            const double2* r = reinterpret_cast<const double2*>(g.rec) + static_cast<int64_t>(c) * 8;
            const double2 e01 = r[6], e2i = r[7];
            f.E = V3<R>{e01.x, e01.y, e2i.x};
            f.iv = e2i.y;
            return f;
        }
        const double2 gu = reinterpret_cast<const double2*>(g.rec)[static_cast<int64_t>(c) * 4 + 3];
        e0 = gu.x;
        e1 = gu.y;
        f.iv = R(-1);                      // from E at the position (it varies within the cell)
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
        e0 = b.z;
        e1 = b.w;
        f.iv = e.x;
        f.emag = e.y;
    }
    if (!g.axisym) {
        f.E = V3<R>{e0, e1, R(0)};
        return f;
    }
    const R r = P::hypot_(pos.x, pos.y);
    if (r > R(1e-15)) {
        if constexpr (P::exact) f.E = V3<R>{e0 * (pos.x / r), e0 * (pos.y / r), e1};
        else {
            const R s = P::rcp(r);
            f.E = V3<R>{e0 * (pos.x * s), e0 * (pos.y * s), e1};
        }
    } else {
        f.E = V3<R>{R(0), R(0), e1};
    }
    return f;
}

// One field evaluation without a carrier cache (a fresh walk).
template <class P>
__device__ inline CField<P> mesh_field(const DevModel& m, const V3<typename P::R>& pos, int32_t& cell1, bool& inside) {
    MeshCache<P> mc;
    inside = mesh_lookup<P>(m, pos, cell1, mc);
    return mesh_cell_field<P>(m, pos, cell1 - 1, inside);
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
