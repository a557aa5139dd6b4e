// Host side of samurai's MRMesh: the five mesh ids, the storage numbering,
// the domain corners, and the tag-driven rebuild with graduation.
#pragma once

#include "interval_set.hpp"

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace smr
{
    constexpr int MAXL = 21; // CellArray max_size 20 -> levels 0..20 (samurai_config.hpp:50)

    enum MeshId
    {
        CELLS            = 0,
        CELLS_AND_GHOSTS = 1,
        PROJ_CELLS       = 2,
        UNION_CELLS      = 3,
        REFERENCE        = 4,
        N_MESH_IDS       = 5
    };

    enum CellFlag : uint8_t
    {
        FLAG_KEEP    = 1,
        FLAG_COARSEN = 2,
        FLAG_REFINE  = 4
    };

    using Dir       = std::array<int, 3>;
    using CellArray = std::vector<LevelSet>; // MAXL levels

    struct MeshConfig
    {
        int dim        = 2;
        int min_level  = 0;
        int max_level  = 6;
        int pred_s     = 1; // prediction stencil radius
        int radius     = 1; // scheme stencil radius
        int grad_width = 1;
    };

    CellArray empty_cell_array(int dim);
    int ca_min_level(const CellArray& ca);
    int ca_max_level(const CellArray& ca);
    bool ca_equal(const CellArray& a, const CellArray& b);

    std::vector<Dir> cartesian_directions(int dim);
    std::vector<Dir> diagonal_directions(int dim);

    struct MRMeshData
    {
        MeshConfig cfg;
        CellArray ids[N_MESH_IDS];
        std::vector<LevelSet> corners;
        std::vector<Dir> corner_dirs;
        // reference storage: value index of x in run k of level l is ridx[l][k] + x
        std::vector<std::vector<int64_t>> ridx;
        std::vector<int64_t> level_base; // first storage index of each level
        int64_t nref = 0;
        // the cells are known to satisfy the graduation (make_graduation leaves them unchanged)
        bool graduated = false;

        const LevelSet& get(int id, int level) const { return ids[id][size_t(level)]; }
        // storage index of a reference cell; throws std::out_of_range if absent
        int64_t index(int level, int x, int y, int z = 0) const;
        // run base (value index = base + x) of the reference run holding x; false if absent
        bool run_base(int level, int x, int y, int z, int64_t& base) const;
        bool exists(int level, int x, int y, int z = 0) const;
        int64_t nb_cells(int id) const;
    };

    // Mesh_base(ca, cfg) -> finalize_mesh (mesh.hpp:386-439, mr/mesh.hpp:221-455)
    MRMeshData build_mesh(const MeshConfig& cfg, const CellArray& cells);
    MRMeshData build_mesh(const MeshConfig& cfg, CellArray&& cells);
    // corners and renumbering of a mesh whose five ids are set (the ids built
    // elsewhere, e.g. on the device, as build_mesh would build them)
    void complete_mesh(MRMeshData& m);
    // mra::make_mesh(unit box, cfg): all cells at max_level (mr/mesh.hpp:510-518)
    MRMeshData build_box_mesh(const MeshConfig& cfg);

    // update_cell_array_from_tag (algorithm/graduation.hpp:743-843); tags are
    // indexed by the mesh's reference storage
    CellArray update_cell_array_from_tag(const MRMeshData& mesh, const uint8_t* tags);
    // make_graduation (algorithm/graduation.hpp:573-726), stencil radius 1
    int make_graduation(CellArray& ca, int dim, int grad_width);
} // namespace smr
