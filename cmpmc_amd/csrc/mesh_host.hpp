// Host representation of an FEM mesh field (the C ABI's opaque cmpmc_mesh) and the image
// of it that a model uploads to the device. Shared by mesh_field.cpp (construction and
// host queries) and capi.hip (upload at model creation).
#pragma once

#include <cstdint>
#include <vector>

#include "cmpmc_amd.h"

// geometry/mesh_field.hpp TriLinearMesh3D (dim 3) / BiLinearMesh2D (dim 2). Cells keep
// their vertex indices sorted ascending, so the last one is the cell's reference vertex.
struct cmpmc_mesh {
    int dim = 3;
    int axisymmetric_rz = 0;              // dim 2: (r = hypot(x, y), z) projection
    std::vector<double> x;                // dim coordinates per vertex
    std::vector<double> v;                // potential per vertex (dim 3: clamped)
    std::vector<int32_t> cells;           // dim + 1 per cell
    std::vector<int32_t> nbr;             // across the facet opposite each vertex, -1 on the hull
    std::vector<double> tinv;             // inverse edge matrix per cell, row-major (9 or 4)
    std::vector<double> grad;             // cell field E = -grad V (3 or 2)
    int64_t n_cells() const { return static_cast<int64_t>(cells.size()) / (dim + 1); }
};

namespace cmpmc_internal {

// Field-dependent part of the intervalley rate (charge_kernels.hpp:82-99) for the crystal's
// model: iv_quad_rate (e0^2 + |E|^2)^(p/2) or iv_lin_rate0 + iv_lin_rate1 (|E|/100)^p.
struct IvParams {
    int model;
    double quad_rate, quad_field, quad_power, lin_rate0, lin_rate1, lin_power;
};
double iv_field_rate(const IvParams& p, const double E[3]);

// Device image of a mesh. Per cell, FP64: dim 3 [tinv 9 | x_ref 3 | E 3 | iv 1] (16),
// dim 2 [tinv 4 | x_ref 2 | E_u E_v] (8); FP32: dim 3 the same 16, dim 2
// [tinv 4 | x_ref 2 | E_u E_v | iv |E| 0 0] (12). nbr4: 4 neighbours per cell (-1 past
// dim + 1). seed: a start cell per box of a regular grid over the mesh's bounding box.
struct MeshImage {
    std::vector<double> rec64;
    std::vector<float> rec32;
    std::vector<int32_t> nbr4;
    std::vector<int32_t> seed;
    int seed_n[3] = {1, 1, 1};
    double lo[3] = {0, 0, 0}, inv_h[3] = {0, 0, 0};
    double iv_zero = 0.0;                 // intervalley field rate at E = 0 (outside the mesh)
    float tol_f = 0.0f;                   // FP32 walk tolerance (see mesh_image)
};
// one_valley: the crystal has a single valley (the FP32 records then carry iv 0)
void mesh_image(const cmpmc_mesh& m, const IvParams& iv, bool one_valley, MeshImage& out);

// Barycentric walk of mesh_field.hpp:63-93 from *hint (reset to n_cells / 2 when out of
// range); p holds dim coordinates. Returns true inside.
bool mesh_locate(const cmpmc_mesh& m, const double* p, int32_t* hint, double* bary);

}  // namespace cmpmc_internal
