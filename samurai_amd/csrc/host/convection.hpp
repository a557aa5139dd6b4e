// Flux-based linear convection (make_convection_upwind) as per-cell ordered
// contribution lists, so that a gather kernel reproduces the reference's
// accumulation order exactly.
#pragma once

#include "mr_mesh.hpp"

#include <cstdint>
#include <vector>

namespace smr
{
    // Row r adds, in order, the terms [ptr[r], ptr[r+1]) into out[row[r]]
    // (which starts at 0): term k is coef[k] * u[src[2k]] when src[2k+1] < 0,
    // else coef[k] * u[src[2k]] + coef[k] * u[src[2k+1]].
    struct ConvTerms
    {
        std::vector<int64_t> row;
        std::vector<int64_t> ptr{0};
        std::vector<int64_t> src;
        std::vector<double> coef;
        std::vector<int8_t> dir; // direction of each term (heterogeneous operator only)
        int64_t nterms() const { return int64_t(coef.size()); }
    };

    // make_convection_upwind(velocity)(u) (schemes/fv/operators/convection_lin.hpp:15-89,
    // FV_scheme.hpp:202-210) on mesh m: for each direction, the same-level
    // interfaces of every level, the level jumps in the direction and in the
    // opposite direction, then the boundary interfaces
    // (flux_based_scheme__lin_hom.hpp:74-207, interface.hpp:35-509), in the
    // order explicit_flux_based_scheme__lin_hom.hpp:39-119, 233-319 adds them.
    // Terms with a zero coefficient are left out: adding 0 * u to a sum that
    // started at +0 never changes it (for finite u).
    ConvTerms build_convection_terms(const MRMeshData& m, const double* velocity);

    // make_convection_upwind<Field>()(u) for a scalar field, the non-linear
    // upwind of f(u) = u * u (schemes/fv/operators/convection_nonlin.hpp:24-79,
    // explicit_flux_based_scheme__nonlin.hpp:35-86): the same interfaces in the
    // same order, each adding flux * left_factor to its left cell and
    // (-flux) * right_factor to its right cell
    // (flux_based_scheme__nonlin.hpp:334-398, 408-590).  Term k is
    // coef[k] * F(u[src[2k]], u[src[2k+1]]) with F(l, r) = (0.5 * (l + r) >= 0 ?
    // l * l : r * r); coef carries the sign, (-F) * rf being F * (-rf) exactly.
    ConvTerms build_burgers_terms(const MRMeshData& m);

    // make_convection_upwind<Field>(velocity_field)(u) for a scalar field
    // (convection_lin.hpp:184-264, explicit_flux_based_scheme__lin_het.hpp:
    // 35-110): the same interfaces; term k, of direction d = dir[k], is
    // (coef[k] * w) * (w >= 0 ? u[s0] : u[s1]) with w = 0.5 * (v_d[s0] + v_d[s1])
    // (s0, s1 = src[2k], src[2k+1]), v_d the storage of velocity component d.
    ConvTerms build_heterogeneous_terms(const MRMeshData& m);

    // make_convection_weno5(velocity)(u) (convection_lin.hpp:95-178) on a mesh
    // whose cells are all at one level, with ghost layers of width 3: for
    // each direction the interfaces between neighbouring cells, then the
    // boundary interfaces in the direction and in the opposite direction
    // (explicit_flux_based_scheme__nonlin.hpp:35-86; interface.hpp:441-509).
    // Term k, of direction d = dir[k], is coef[k] * F_d(u[src[6k]], ...,
    // u[src[6k+5]]) with F_d the WENO5 flux of velocity v_d over the six
    // stencil cells -2..3 along d from the interface's left cell (-3..2 from
    // the inner cell at a boundary in the opposite direction); coef is the
    // h-factor, negative for the right cell.
    ConvTerms build_weno5_terms(const MRMeshData& m);

    // the level of a mesh whose cells fill the domain at one level and whose
    // storage is the dense box of that level widened by 3 ghost layers (every
    // reference row one run [-3, n + 3), rows (z, y) in order), else -1
    int weno5_dense_level(const MRMeshData& m);
} // namespace smr
