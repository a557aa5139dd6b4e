// Restart files in the reference's layout (io/restart.hpp), through the HDF5
// C library and the public C ABI of libsamurai_amd only.
//
// The mesh is stored as samurai's LevelCellArray arrays of the leaves, level
// by level (level_cell_array.hpp:465-522 builds them): per dimension d a list
// of intervals {start, end, step, index} and, for d >= 1, the offsets into the
// intervals of dimension d - 1.  In 2D a row y of a y-interval [ys, ye) is
// entry y + index of the offsets, whose consecutive values bound the row's
// x-intervals; in 3D the z-intervals index the planes the same way, and the
// plane offsets bound the plane's y-intervals.  Consecutive rows (planes)
// share one interval; every row opens new x-intervals.
#include "h5file.hpp"

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
    using namespace smr_io;

    // samurai::Interval<int, long long> (interval.hpp:50-100)
    struct Interval
    {
        int32_t start;
        int32_t end;
        int32_t step;
        int32_t pad;
        int64_t index;
    };
    static_assert(sizeof(Interval) == 24, "Interval layout");

    Handle interval_type()
    {
        Handle t(H5Tcreate(H5T_COMPOUND, sizeof(Interval)), H5Tclose, "compound type");
        h5check(H5Tinsert(t, "start", offsetof(Interval, start), H5T_NATIVE_INT), "insert start");
        h5check(H5Tinsert(t, "end", offsetof(Interval, end), H5T_NATIVE_INT), "insert end");
        h5check(H5Tinsert(t, "step", offsetof(Interval, step), H5T_NATIVE_INT), "insert step");
        h5check(H5Tinsert(t, "index", offsetof(Interval, index), H5T_NATIVE_LLONG), "insert index");
        return t;
    }

    std::string level_dim(int level, int d, const char* what)
    {
        return "/mesh/level/" + std::to_string(level) + "/dim/" + std::to_string(d) + "/" + what;
    }

    // one level's leaves: rows (z, y) in order, each with its x-runs and the
    // storage offset of every run (value index of x = index + x)
    struct LevelRows
    {
        std::vector<int32_t> y, z;
        std::vector<uint64_t> rp{0};
        std::vector<Interval> runs;
    };

    // leaves as rows (level, x, y, z[, storage]) in for_each_cell order
    // (levels ascending, rows (z, y), x) -> per-level rows of maximal runs
    std::vector<LevelRows> group_rows(const std::function<std::array<int64_t, 5>(int64_t)>& cell, int64_t n, bool has_storage)
    {
        std::vector<LevelRows> levels;
        int64_t prev_level = -1;
        for (int64_t i = 0; i < n; ++i)
        {
            const auto c = cell(i);
            require(c[0] >= 0 && c[0] < 21, "level out of range");
            if (c[0] != prev_level)
            {
                require(c[0] > prev_level, "cells not in for_each_cell order (levels)");
                levels.resize(size_t(c[0]) + 1);
                prev_level = c[0];
            }
            LevelRows& L       = levels[size_t(c[0])];
            const int32_t x    = int32_t(c[1]), y = int32_t(c[2]), z = int32_t(c[3]);
            // without storage the cells are numbered in for_each_cell order,
            // as CellArray's update_index does (cell_array.hpp:484-493)
            const int64_t base = (has_storage ? c[4] : i) - x;
            const bool same_row = !L.y.empty() && L.y.back() == y && L.z.back() == z;
            if (same_row && L.runs.back().end == x && L.runs.back().index == base)
            {
                ++L.runs.back().end;
                continue;
            }
            if (!same_row)
            {
                require(L.y.empty() || L.z.back() < z || (L.z.back() == z && L.y.back() < y), "cells not in for_each_cell order (rows)");
                L.y.push_back(y);
                L.z.push_back(z);
                L.rp.push_back(L.rp.back());
            }
            else
            {
                require(L.runs.back().end < x, "cells not in for_each_cell order (x)");
            }
            L.runs.push_back({x, x + 1, 1, 0, base});
            ++L.rp.back();
        }
        return levels;
    }

    void dump_levels(File& f, int dim, const std::vector<LevelRows>& levels)
    {
        const Handle itype = interval_type();
        for (size_t l = 0; l < levels.size(); ++l)
        {
            const LevelRows& L = levels[l];
            if (L.y.empty())
            {
                continue;
            }
            const int level = int(l);
            const size_t nrows = L.y.size();
            // dimension 0: the x-runs, row by row
            f.partitioned(level_dim(level, 0, "intervals"), itype, L.runs);
            if (dim == 2)
            {
                // y-intervals over consecutive rows; index + y = row position
                std::vector<Interval> yi;
                for (size_t p = 0; p < nrows; ++p)
                {
                    if (!yi.empty() && yi.back().end == L.y[p])
                    {
                        ++yi.back().end;
                    }
                    else
                    {
                        yi.push_back({L.y[p], L.y[p] + 1, 1, 0, int64_t(p) - L.y[p]});
                    }
                }
                f.partitioned(level_dim(level, 1, "intervals"), itype, yi);
                f.partitioned(level_dim(level, 1, "offsets"), H5T_NATIVE_ULONG, std::vector<unsigned long>(L.rp.begin(), L.rp.end()));
                continue;
            }
            // 3D: planes of rows with equal z, y-intervals within each plane,
            // z-intervals over consecutive planes
            std::vector<Interval> zi, yi;
            std::vector<unsigned long> plane_off{0};
            size_t q = 0; // plane position
            for (size_t p = 0; p < nrows; ++p)
            {
                const bool new_plane = p == 0 || L.z[p] != L.z[p - 1];
                if (new_plane)
                {
                    if (p > 0)
                    {
                        plane_off.push_back(static_cast<unsigned long>(yi.size()));
                        ++q;
                    }
                    if (!zi.empty() && zi.back().end == L.z[p])
                    {
                        ++zi.back().end;
                    }
                    else
                    {
                        zi.push_back({L.z[p], L.z[p] + 1, 1, 0, int64_t(q) - L.z[p]});
                    }
                }
                if (!new_plane && yi.back().end == L.y[p])
                {
                    ++yi.back().end;
                }
                else
                {
                    yi.push_back({L.y[p], L.y[p] + 1, 1, 0, int64_t(p) - L.y[p]});
                }
            }
            plane_off.push_back(static_cast<unsigned long>(yi.size()));
            f.partitioned(level_dim(level, 1, "intervals"), itype, yi);
            f.partitioned(level_dim(level, 2, "intervals"), itype, zi);
            f.partitioned(level_dim(level, 1, "offsets"), H5T_NATIVE_ULONG, std::vector<unsigned long>(L.rp.begin(), L.rp.end()));
            f.partitioned(level_dim(level, 2, "offsets"), H5T_NATIVE_ULONG, plane_off);
        }
    }

    void dump_header(File& f, int dim, int min_level, int max_level)
    {
        f.scalar<unsigned long>("/n_process", H5T_NATIVE_ULONG, 1);
        f.scalar<unsigned long>("/mesh/dim", H5T_NATIVE_ULONG, static_cast<unsigned long>(dim));
        f.scalar<unsigned long>("/mesh/min_level", H5T_NATIVE_ULONG, static_cast<unsigned long>(min_level));
        f.scalar<unsigned long>("/mesh/max_level", H5T_NATIVE_ULONG, static_cast<unsigned long>(max_level));
        f.vector<double>("/mesh/origin_point", H5T_NATIVE_DOUBLE, std::vector<double>(size_t(dim), 0.0));
        f.scalar<double>("/mesh/scaling_factor", H5T_NATIVE_DOUBLE, 1.0);
    }

    // the leaves of a file as rows (level, x, y, z), for_each_cell order
    std::vector<int32_t> load_leaves(const File& f, int& dim, int& min_level, int& max_level)
    {
        const unsigned long nproc = f.read_scalar<unsigned long>("/n_process", H5T_NATIVE_ULONG);
        require(nproc == 1, "the restart file was written by " + std::to_string(nproc) + " processes, this is one");
        dim       = int(f.read_scalar<unsigned long>("/mesh/dim", H5T_NATIVE_ULONG));
        min_level = int(f.read_scalar<unsigned long>("/mesh/min_level", H5T_NATIVE_ULONG));
        max_level = int(f.read_scalar<unsigned long>("/mesh/max_level", H5T_NATIVE_ULONG));
        require(dim == 2 || dim == 3, "dimension " + std::to_string(dim) + " is not supported");
        require(0 <= min_level && min_level <= max_level && max_level <= 20, "bad levels in the restart file");
        const Handle itype = interval_type();
        std::vector<int32_t> out;
        auto emit_row = [&](int level, int32_t y, int32_t z, const std::vector<Interval>& xi, const std::vector<unsigned long>& off, int64_t p)
        {
            require(p >= 0 && size_t(p) + 1 < off.size(), "row outside the offsets");
            const unsigned long a = off[size_t(p)], b = off[size_t(p) + 1];
            require(a <= b && b <= xi.size(), "offsets outside the x-intervals");
            for (unsigned long k = a; k < b; ++k)
            {
                require(xi[k].start < xi[k].end, "empty x-interval");
                for (int32_t x = xi[k].start; x < xi[k].end; ++x)
                {
                    out.insert(out.end(), {level, x, y, z});
                }
            }
        };
        for (int level = min_level; level <= max_level; ++level)
        {
            if (!f.exists("/mesh/level/" + std::to_string(level)))
            {
                continue;
            }
            const auto xi  = f.read_partitioned<Interval>(level_dim(level, 0, "intervals"), itype);
            const auto yi  = f.read_partitioned<Interval>(level_dim(level, 1, "intervals"), itype);
            const auto yoff = f.read_partitioned<unsigned long>(level_dim(level, 1, "offsets"), H5T_NATIVE_ULONG);
            if (dim == 2)
            {
                for (const Interval& iy : yi)
                {
                    for (int32_t y = iy.start; y < iy.end; ++y)
                    {
                        emit_row(level, y, 0, xi, yoff, int64_t(y) + iy.index);
                    }
                }
                continue;
            }
            const auto zi   = f.read_partitioned<Interval>(level_dim(level, 2, "intervals"), itype);
            const auto zoff = f.read_partitioned<unsigned long>(level_dim(level, 2, "offsets"), H5T_NATIVE_ULONG);
            for (const Interval& iz : zi)
            {
                for (int32_t z = iz.start; z < iz.end; ++z)
                {
                    const int64_t q = int64_t(z) + iz.index;
                    require(q >= 0 && size_t(q) + 1 < zoff.size(), "plane outside the offsets");
                    require(zoff[size_t(q)] <= zoff[size_t(q) + 1] && zoff[size_t(q) + 1] <= yi.size(), "offsets outside the y-intervals");
                    for (unsigned long j = zoff[size_t(q)]; j < zoff[size_t(q) + 1]; ++j)
                    {
                        for (int32_t y = yi[j].start; y < yi[j].end; ++y)
                        {
                            emit_row(level, y, z, xi, yoff, int64_t(y) + yi[j].index);
                        }
                    }
                }
            }
        }
        return out;
    }

    // load_field (io/restart.hpp:341-386): the values of /fields/{name}, for_each_cell
    // order, into a field living on `mesh`
    void load_field(const File& f, const smr_mesh* mesh, smr_field* field)
    {
        const int on = smr_field_on_mesh(field, mesh);
        check_smr(on, "field");
        require(on == 1, "the field does not live on the given mesh");
        const char* name = smr_field_name(field);
        require(name && *name, "The field has no name.");
        const std::string base = std::string("/fields/") + name;
        require(f.exists(base), std::string("The field ") + name + " does not exist in the file.");
        const unsigned long n_comp = f.read_scalar<unsigned long>(base + "/n_comp", H5T_NATIVE_ULONG);
        require(n_comp == 1, "the field in the file has " + std::to_string(n_comp) + " components, a scalar field has 1");
        check_smr(smr_field_resize(field), "field resize");
        std::vector<double> v;
        if (f.exists(base + "/data"))
        {
            v = f.read_partitioned<double>(base + "/data", H5T_NATIVE_DOUBLE);
        }
        const int64_t n = smr_mesh_nb_cells(mesh, SMR_CELLS);
        check_smr(n, "mesh size");
        require(int64_t(v.size()) == n, "the field in the file has " + std::to_string(v.size()) + " values, the mesh " + std::to_string(n) +
                                            " cells");
        check_smr(smr_field_set_cells(field, v.data()), "field values");
    }

    struct Silence
    {
        Silence() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
    };
    const Silence g_silence;
} // namespace

extern "C"
{
    const char* smr_io_last_error(void)
    {
        return g_io_error.c_str();
    }

    int smr_restart_dump(const char* file, const smr_mesh* mesh, smr_field* const* fields, int nfields)
    {
        return guarded(
            [&]
            {
                require(file && mesh && nfields >= 0 && (nfields == 0 || fields), "null argument");
                smr_mesh_config cfg;
                check_smr(smr_mesh_get_config(mesh, &cfg), "mesh config");
                const int64_t n = smr_mesh_nb_cells(mesh, SMR_CELLS);
                check_smr(n, "mesh size");
                std::vector<int64_t> cells(size_t(n) * 5);
                check_smr(smr_mesh_get_cells(mesh, SMR_CELLS, cells.data()), "mesh cells");
                const auto levels = group_rows(
                    [&](int64_t i)
                    {
                        const int64_t* c = &cells[size_t(i) * 5];
                        return std::array<int64_t, 5>{c[0], c[1], c[2], c[3], c[4]};
                    },
                    n, true);
                std::vector<std::vector<double>> values(static_cast<size_t>(nfields));
                std::vector<std::string> names(static_cast<size_t>(nfields));
                for (int k = 0; k < nfields; ++k)
                {
                    require(fields[k] != nullptr, "null field");
                    const int on = smr_field_on_mesh(fields[k], mesh);
                    check_smr(on, "field");
                    require(on == 1, "a field does not live on the dumped mesh");
                    const char* name = smr_field_name(fields[k]);
                    require(name && *name, "a field has no name");
                    names[size_t(k)] = name;
                    values[size_t(k)].resize(size_t(n));
                    check_smr(smr_field_get_cells(fields[k], values[size_t(k)].data()), "field values");
                }
                File f(file, true);
                dump_header(f, cfg.dim, cfg.min_level, cfg.max_level);
                dump_levels(f, cfg.dim, levels);
                for (int k = 0; k < nfields; ++k)
                {
                    f.scalar<unsigned long>("/fields/" + names[size_t(k)] + "/n_comp", H5T_NATIVE_ULONG, 1);
                    f.partitioned("/fields/" + names[size_t(k)] + "/data", H5T_NATIVE_DOUBLE, values[size_t(k)]);
                }
            });
    }

    int smr_restart_load_mesh(const char* file, smr_ctx* ctx, const smr_mesh_config* cfg, smr_mesh** out)
    {
        return guarded(
            [&]
            {
                require(file && ctx && cfg && out, "null argument");
                File f(file, false);
                int dim = 0, min_level = 0, max_level = 0;
                const auto leaves = load_leaves(f, dim, min_level, max_level);
                require(dim == cfg->dim, "the restart file's dimension (" + std::to_string(dim) + ") is not the configuration's");
                smr_mesh_config c = *cfg;
                c.min_level       = min_level;
                c.max_level       = max_level;
                check_smr(smr_mesh_create_from_cells(ctx, &c, leaves.data(), int64_t(leaves.size() / 4), out), "mesh from the restart file");
            });
    }

    int smr_restart_load_field(const char* file, const smr_mesh* mesh, smr_field* field)
    {
        return guarded(
            [&]
            {
                require(file && mesh && field, "null argument");
                File f(file, false);
                load_field(f, mesh, field);
            });
    }

    int smr_restart_load(const char* file, smr_mesh* mesh, smr_field* const* fields, int nfields)
    {
        return guarded(
            [&]
            {
                require(file && mesh && nfields >= 0 && (nfields == 0 || fields), "null argument");
                for (int k = 0; k < nfields; ++k)
                {
                    require(fields[k] != nullptr, "null field");
                    const int on = smr_field_on_mesh(fields[k], mesh);
                    check_smr(on, "field");
                    require(on == 1, "a field does not live on the given mesh");
                    const char* name = smr_field_name(fields[k]);
                    require(name && *name, "The field has no name.");
                }
                File f(file, false);
                smr_mesh_config cfg;
                check_smr(smr_mesh_get_config(mesh, &cfg), "mesh config");
                int dim = 0, min_level = 0, max_level = 0;
                const auto leaves = load_leaves(f, dim, min_level, max_level);
                require(dim == cfg.dim, "The dimension of the mesh is not the same as the one of the mesh to be loaded. " +
                                            std::to_string(dim) + " != " + std::to_string(cfg.dim));
                // every field is in the file before the mesh is replaced
                for (int k = 0; k < nfields; ++k)
                {
                    const std::string name = smr_field_name(fields[k]);
                    require(f.exists("/fields/" + name), "The field " + name + " does not exist in the file.");
                }
                check_smr(smr_mesh_assign_cells(mesh, min_level, max_level, leaves.data(), int64_t(leaves.size() / 4)),
                          "mesh from the restart file");
                for (int k = 0; k < nfields; ++k)
                {
                    load_field(f, mesh, fields[k]);
                }
            });
    }

    int smr_restart_dump_cells(const char* file, int dim, const int32_t* cells, int64_t n)
    {
        return guarded(
            [&]
            {
                require(file && (n == 0 || cells) && n >= 0 && (dim == 2 || dim == 3), "bad argument");
                const auto levels = group_rows(
                    [&](int64_t i)
                    {
                        const int32_t* c = cells + 4 * i;
                        return std::array<int64_t, 5>{c[0], c[1], c[2], dim == 3 ? c[3] : 0, 0};
                    },
                    n, false);
                int lo = 21, hi = -1;
                for (size_t l = 0; l < levels.size(); ++l)
                {
                    if (!levels[l].y.empty())
                    {
                        lo = std::min(lo, int(l));
                        hi = std::max(hi, int(l));
                    }
                }
                File f(file, true);
                dump_header(f, dim, hi < 0 ? 0 : lo, hi < 0 ? 0 : hi);
                dump_levels(f, dim, levels);
            });
    }

    int64_t smr_restart_load_cells(const char* file, int* dim, int32_t* cells, int64_t capacity)
    {
        int64_t count = 0;
        const int rc  = guarded(
            [&]
            {
                require(file && dim && capacity >= 0 && (capacity == 0 || cells), "bad argument");
                File f(file, false);
                int min_level = 0, max_level = 0;
                const auto leaves = load_leaves(f, *dim, min_level, max_level);
                count             = int64_t(leaves.size() / 4);
                const int64_t m   = std::min(count, capacity);
                if (m > 0)
                {
                    std::memcpy(cells, leaves.data(), size_t(m) * 4 * sizeof(int32_t));
                }
            });
        return rc < 0 ? rc : count;
    }
}
