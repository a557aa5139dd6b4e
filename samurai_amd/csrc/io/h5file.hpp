// HDF5 plumbing shared by the restart and XDMF writers of libsamurai_amd_io:
// RAII handles, a file with the dataset shapes samurai's files use, and the
// error convention of the C ABI (negative status, message per thread).
#pragma once

#include "samurai_amd_io.h"

#include <hdf5.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace smr_io
{
    inline thread_local std::string g_io_error;

    template <class F>
    int guarded(F&& f)
    {
        try
        {
            f();
            return SMR_OK;
        }
        catch (const std::invalid_argument& e)
        {
            g_io_error = e.what();
            return SMR_ERR_ARGUMENT;
        }
        catch (const std::exception& e)
        {
            g_io_error = e.what();
            return SMR_ERR_RUNTIME;
        }
    }

    inline void require(bool ok, const std::string& what)
    {
        if (!ok)
        {
            throw std::invalid_argument(what);
        }
    }

    inline void check_smr(int64_t rc, const char* what)
    {
        if (rc < 0)
        {
            throw std::runtime_error(std::string(what) + ": " + smr_last_error());
        }
    }

    inline void h5check(herr_t e, const std::string& what)
    {
        if (e < 0)
        {
            throw std::runtime_error("HDF5: " + what);
        }
    }

    // an HDF5 handle closed by its own close function
    class Handle
    {
      public:
        Handle(hid_t id, herr_t (*close)(hid_t), const std::string& what)
            : id_(id)
            , close_(close)
        {
            if (id_ < 0)
            {
                throw std::runtime_error("HDF5: " + what);
            }
        }
        ~Handle()
        {
            if (id_ >= 0)
            {
                close_(id_);
            }
        }
        Handle(Handle&& o) noexcept
            : id_(o.id_)
            , close_(o.close_)
        {
            o.id_ = -1;
        }
        Handle(const Handle&)            = delete;
        Handle& operator=(const Handle&) = delete;
        Handle& operator=(Handle&&)      = delete;
        operator hid_t() const { return id_; }

      private:
        hid_t id_;
        herr_t (*close_)(hid_t);
    };

    class File
    {
      public:
        // create (truncating, as HighFive's File::Overwrite) or open read-only
        File(const std::string& path, bool create)
            : f_(create ? H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT) : H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT),
                 H5Fclose, std::string(create ? "cannot create " : "cannot open ") + path)
            , lcpl_(H5Pcreate(H5P_LINK_CREATE), H5Pclose, "link property list")
        {
            h5check(H5Pset_create_intermediate_group(lcpl_, 1), "intermediate groups");
        }

        // H5Easy::dump of a scalar (H5Easy.hpp): a scalar dataspace
        template <class T>
        void scalar(const std::string& name, hid_t type, T v)
        {
            Handle sp(H5Screate(H5S_SCALAR), H5Sclose, "scalar space");
            Handle ds(H5Dcreate2(f_, name.c_str(), type, sp, lcpl_, H5P_DEFAULT, H5P_DEFAULT), H5Dclose, "create " + name);
            h5check(H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), "write " + name);
        }

        // a dataset of shape {rows} (cols == 0) or {rows, cols}, row-major data
        template <class T>
        void array(const std::string& name, hid_t type, const T* data, hsize_t rows, hsize_t cols = 0)
        {
            const hsize_t shape[2] = {rows, cols};
            Handle sp(H5Screate_simple(cols == 0 ? 1 : 2, shape, nullptr), H5Sclose, "space");
            Handle ds(H5Dcreate2(f_, name.c_str(), type, sp, lcpl_, H5P_DEFAULT, H5P_DEFAULT), H5Dclose, "create " + name);
            if (rows > 0)
            {
                h5check(H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "write " + name);
            }
        }

        template <class T>
        void vector(const std::string& name, hid_t type, const std::vector<T>& v)
        {
            array(name, type, v.data(), v.size());
        }

        // dump(file, name, data) (io/restart.hpp:52-87): {name}/partition, the
        // cumulative sizes over the processes, and {name}/data; nothing for
        // empty data
        template <class T>
        void partitioned(const std::string& name, hid_t type, const std::vector<T>& v)
        {
            if (v.empty())
            {
                return;
            }
            vector<unsigned long>(name + "/partition", H5T_NATIVE_ULONG, {0UL, static_cast<unsigned long>(v.size())});
            vector<T>(name + "/data", type, v);
        }

        bool exists(const std::string& name) const
        {
            // every link on the path, from the root
            size_t pos = 1;
            while (true)
            {
                const size_t next      = name.find('/', pos);
                const std::string part = name.substr(0, next);
                if (H5Lexists(f_, part.c_str(), H5P_DEFAULT) <= 0)
                {
                    return false;
                }
                if (next == std::string::npos)
                {
                    return true;
                }
                pos = next + 1;
            }
        }

        template <class T>
        T read_scalar(const std::string& name, hid_t type) const
        {
            require(exists(name), "restart file has no " + name);
            Handle ds(H5Dopen2(f_, name.c_str(), H5P_DEFAULT), H5Dclose, "open " + name);
            Handle sp(H5Dget_space(ds), H5Sclose, "space of " + name);
            require(H5Sget_simple_extent_npoints(sp) == 1, name + " is not a scalar");
            T v{};
            h5check(H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), "read " + name);
            return v;
        }

        template <class T>
        std::vector<T> read_vector(const std::string& name, hid_t type) const
        {
            require(exists(name), "restart file has no " + name);
            Handle ds(H5Dopen2(f_, name.c_str(), H5P_DEFAULT), H5Dclose, "open " + name);
            Handle sp(H5Dget_space(ds), H5Sclose, "space of " + name);
            const hssize_t n = H5Sget_simple_extent_npoints(sp);
            require(n >= 0, "bad extent of " + name);
            std::vector<T> v(static_cast<size_t>(n));
            if (n > 0)
            {
                h5check(H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "read " + name);
            }
            return v;
        }

        // load(file, name) (io/restart.hpp:214-234), one process: partition
        // [0, n] then n values
        template <class T>
        std::vector<T> read_partitioned(const std::string& name, hid_t type) const
        {
            const auto part = read_vector<unsigned long>(name + "/partition", H5T_NATIVE_ULONG);
            require(part.size() == 2 && part[0] == 0, name + "/partition is not a single-process partition");
            auto v = read_vector<T>(name + "/data", type);
            require(v.size() == part[1], name + "/data does not match its partition");
            return v;
        }

      private:
        Handle f_;
        Handle lcpl_;
    };
} // namespace smr_io
