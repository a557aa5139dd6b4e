// samurai's output and restart functions on the C++ drop-in layer, over
// libsamurai_amd_io (link with -lsamurai_amd_io -lsamurai_amd):
//
//   samurai::save(path, filename, [Hdf5Options,] mesh, fields...)  io/hdf5.hpp:1280-1347
//   samurai::dump(path, filename, mesh, fields...)                 io/restart.hpp:189-237
//   samurai::load(path, filename, mesh, fields...)                 io/restart.hpp:423-462
//
// and the overloads without `path` (the current directory).  `filename` has
// no extension: ".h5" (and ".xdmf" for save) are appended, as in samurai.
#pragma once

#include "../samurai_amd_io.h"
#include "samurai.hpp"

#include <filesystem>
#include <stdexcept>
#include <string>
#include <vector>

namespace samurai_amd
{
    namespace fs = std::filesystem;

    // Hdf5Options (io/hdf5.hpp:146-158): one grid per level and/or per mesh id
    template <class Mesh>
    struct Hdf5Options
    {
        Hdf5Options(bool level = false, bool mesh_id = false)
            : by_level(level)
            , by_mesh_id(mesh_id)
        {
        }
        bool by_level   = false;
        bool by_mesh_id = false;
    };

    namespace io_detail
    {
        inline void check(int rc)
        {
            if (rc < 0)
            {
                const std::string msg = smr_io_last_error();
                if (rc == SMR_ERR_ARGUMENT)
                {
                    throw std::invalid_argument(msg);
                }
                throw std::runtime_error(msg);
            }
        }

        template <class... F>
        std::vector<smr_field*> handles(F&... fields)
        {
            return {fields.handle()...};
        }

        inline std::string h5(const fs::path& path, const std::string& filename)
        {
            return (path / filename).string() + ".h5";
        }
    } // namespace io_detail

    template <std::size_t dim, class... F>
    void save(const fs::path& path, const std::string& filename, const Hdf5Options<MRMesh<dim>>& options, MRMesh<dim>& mesh, F&... fields)
    {
        auto h        = io_detail::handles(fields...);
        const int opt = (options.by_level ? SMR_SAVE_BY_LEVEL : 0) | (options.by_mesh_id ? SMR_SAVE_BY_MESH_ID : 0);
        io_detail::check(smr_save(path.string().c_str(), filename.c_str(), mesh.handle(), h.data(), int(h.size()), opt));
    }

    template <std::size_t dim, class... F>
    void save(const fs::path& path, const std::string& filename, MRMesh<dim>& mesh, F&... fields)
    {
        save(path, filename, Hdf5Options<MRMesh<dim>>{}, mesh, fields...);
    }

    template <std::size_t dim, class... F>
    void save(const std::string& filename, MRMesh<dim>& mesh, F&... fields)
    {
        save(fs::current_path(), filename, Hdf5Options<MRMesh<dim>>{}, mesh, fields...);
    }

    template <std::size_t dim, class... F>
    void dump(const fs::path& path, const std::string& filename, MRMesh<dim>& mesh, F&... fields)
    {
        auto h = io_detail::handles(fields...);
        io_detail::check(smr_restart_dump(io_detail::h5(path, filename).c_str(), mesh.handle(), h.data(), int(h.size())));
    }

    template <std::size_t dim, class... F>
    void dump(const std::string& filename, MRMesh<dim>& mesh, F&... fields)
    {
        dump(fs::current_path(), filename, mesh, fields...);
    }

    // the mesh is replaced in place (its fields stay attached to it) and each
    // field gets the file's values of its name
    template <std::size_t dim, class... F>
    void load(const fs::path& path, const std::string& filename, MRMesh<dim>& mesh, F&... fields)
    {
        auto h = io_detail::handles(fields...);
        io_detail::check(smr_restart_load(io_detail::h5(path, filename).c_str(), mesh.handle(), h.data(), int(h.size())));
        mesh.refresh();
        (fields.touched(), ...);
    }

    template <std::size_t dim, class... F>
    void load(const std::string& filename, MRMesh<dim>& mesh, F&... fields)
    {
        load(fs::current_path(), filename, mesh, fields...);
    }
} // namespace samurai_amd
