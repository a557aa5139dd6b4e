// samurai::save (io/hdf5.hpp): the leaves of a mesh as an unstructured grid of
// quadrilaterals (hexahedra in 3D) in {path}/{filename}.h5, with one
// cell-centred dataset per field, and the XDMF document that points into it
// in {path}/{filename}.xdmf, for ParaView and VisIt.
//
// Points are numbered in the order they are first met: for_each_cell order of
// the cells, then the corners of each cell in the element order
// (io/hdf5.hpp:63-82, 84-145).  A corner is an exact binary fraction of the
// unit box, so corners are identified by their integer coordinates on the
// finest level of the mesh, which is the same as the reference's comparison of
// the coordinates as doubles.
#include "h5file.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <list>
#include <string>
#include <unordered_map>
#include <vector>

namespace
{
    using namespace smr_io;

    // --- a minimal XML tree, written the way pugixml's default format writes
    // it (declaration, 4-space indentation, text-only elements on one line,
    // childless elements self-closed with " />")
    struct Node
    {
        std::string name;
        std::vector<std::pair<std::string, std::string>> attrs;
        std::string text;
        std::list<Node> children; // a list: references to children stay valid as siblings are added

        Node& child(const std::string& n)
        {
            children.push_back(Node{n, {}, {}, {}});
            return children.back();
        }
        Node& attr(const std::string& k, const std::string& v)
        {
            attrs.emplace_back(k, v);
            return *this;
        }
        Node& attr(const std::string& k, uint64_t v) { return attr(k, std::to_string(v)); }
    };

    std::string escape(const std::string& s, bool attribute)
    {
        std::string o;
        for (char c : s)
        {
            switch (c)
            {
                case '&': o += "&amp;"; break;
                case '<': o += "&lt;"; break;
                case '>': o += "&gt;"; break;
                case '"': o += attribute ? "&quot;" : "\""; break;
                default: o += c;
            }
        }
        return o;
    }

    void write_node(std::string& out, const Node& n, int depth)
    {
        out.append(size_t(4 * depth), ' ');
        out += "<" + n.name;
        for (const auto& [k, v] : n.attrs)
        {
            out += " " + k + "=\"" + escape(v, true) + "\"";
        }
        if (n.children.empty() && n.text.empty())
        {
            out += " />\n";
            return;
        }
        if (n.children.empty())
        {
            out += ">" + escape(n.text, false) + "</" + n.name + ">\n";
            return;
        }
        out += ">\n";
        for (const Node& c : n.children)
        {
            write_node(out, c, depth + 1);
        }
        out.append(size_t(4 * depth), ' ');
        out += "</" + n.name + ">\n";
    }

    // --- a sub-mesh: its cells (level, x, y, z) in for_each_cell order and,
    // per cell, the index of its value in the fields' arrays
    struct Submesh
    {
        std::vector<std::array<int64_t, 4>> cells;
        std::vector<int64_t> value_index;
    };

    struct FieldData
    {
        std::string name;
        std::vector<double> values;
    };

    // io/hdf5.hpp:70-82
    constexpr int kQuad[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    constexpr int kHexa[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    const char* element_type(int dim) { return dim == 2 ? "Quadrilateral" : "Hexahedron"; }

    struct KeyHash
    {
        size_t operator()(const std::array<int64_t, 3>& k) const noexcept
        {
            uint64_t h = uint64_t(k[0]) * 0x9E3779B97F4A7C15ull;
            h ^= uint64_t(k[1]) + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
            h ^= uint64_t(k[2]) + 0x94D049BB133111EBull + (h << 6) + (h >> 2);
            return size_t(h);
        }
    };

    // extract_coords_and_connectivity (io/hdf5.hpp:84-145)
    void coords_and_connectivity(const Submesh& sm, int dim, std::vector<double>& coords, std::vector<uint64_t>& conn)
    {
        int fine = 0;
        for (const auto& c : sm.cells)
        {
            fine = std::max(fine, int(c[0]));
        }
        const int npc = 1 << dim;
        std::unordered_map<std::array<int64_t, 3>, uint64_t, KeyHash> ids;
        ids.reserve(sm.cells.size() * 2);
        std::vector<std::array<int64_t, 3>> points;
        conn.assign(sm.cells.size() * size_t(npc), 0);
        for (size_t i = 0; i < sm.cells.size(); ++i)
        {
            const auto& c   = sm.cells[i];
            const int shift = fine - int(c[0]);
            for (int e = 0; e < npc; ++e)
            {
                std::array<int64_t, 3> key{0, 0, 0};
                for (int d = 0; d < dim; ++d)
                {
                    const int off = dim == 2 ? kQuad[e][d] : kHexa[e][d];
                    key[size_t(d)] = (c[size_t(d) + 1] + off) * (int64_t(1) << shift);
                }
                auto it = ids.find(key);
                if (it == ids.end())
                {
                    it = ids.emplace(key, uint64_t(points.size())).first;
                    points.push_back(key);
                }
                conn[i * size_t(npc) + size_t(e)] = it->second;
            }
        }
        coords.assign(points.size() * 3, 0.0);
        for (size_t p = 0; p < points.size(); ++p)
        {
            for (int d = 0; d < dim; ++d)
            {
                coords[p * 3 + size_t(d)] = std::ldexp(double(points[p][size_t(d)]), -fine);
            }
        }
    }

    class Saver
    {
      public:
        Saver(const std::string& path, const std::string& filename, int dim)
            : path_(path)
            , filename_(filename)
            , dim_(dim)
            , h5_(join(path, filename) + ".h5", true)
        {
            root_.name = "Xdmf";
            root_.child("Domain");
        }

        Node& domain() { return root_.children.front(); }

        // Hdf5::save_on_mesh (io/hdf5.hpp:672-834), one process
        void save_on_mesh(Node& parent, const std::string& prefix, const Submesh& sm, const std::string& mesh_name,
                          const std::vector<FieldData>& fields)
        {
            std::vector<double> coords;
            std::vector<uint64_t> conn;
            coords_and_connectivity(sm, dim_, coords, conn);
            if (coords.empty())
            {
                return;
            }
            const uint64_t ncells = sm.cells.size(), npts = coords.size() / 3, npc = uint64_t(1) << dim_;
            h5_.array(prefix + "/connectivity", H5T_NATIVE_UINT64, conn.data(), ncells, npc);
            h5_.array(prefix + "/points", H5T_NATIVE_DOUBLE, coords.data(), npts, 3);

            Node& grid = parent.child("Grid");
            grid.attr("Name", mesh_name);
            Node& topo = grid.child("Topology");
            topo.attr("TopologyType", element_type(dim_)).attr("NumberOfElements", ncells);
            Node& td = topo.child("DataItem");
            td.attr("Dimensions", ncells * npc).attr("Format", "HDF");
            td.text = filename_ + ".h5:" + prefix + "/connectivity";
            Node& geom = grid.child("Geometry");
            geom.attr("GeometryType", "XYZ");
            Node& gd = geom.child("DataItem");
            gd.attr("Dimensions", npts * 3).attr("Format", "HDF");
            gd.text = filename_ + ".h5:" + prefix + "/points";

            // Hdf5::save_field (io/hdf5.hpp:837-905), scalar fields
            for (const FieldData& f : fields)
            {
                std::vector<double> v(ncells);
                for (size_t i = 0; i < ncells; ++i)
                {
                    v[i] = f.values[size_t(sm.value_index[i])];
                }
                const std::string dset = prefix + "/fields/" + f.name;
                h5_.array(dset, H5T_NATIVE_DOUBLE, v.data(), ncells);
                Node& a = grid.child("Attribute");
                a.attr("Name", f.name).attr("Center", "Cell");
                Node& ad = a.child("DataItem");
                ad.attr("Dimensions", ncells).attr("Format", "HDF").attr("Precision", "8");
                ad.text = filename_ + ".h5:" + dset;
            }
        }

        // the XDMF document, written when the save ends (Hdf5::~Hdf5,
        // io/hdf5.hpp:655-665)
        void write_xdmf() const
        {
            std::string out = "<?xml version=\"1.0\"?>\n";
            write_node(out, root_, 0);
            const std::string name = join(path_, filename_) + ".xdmf";
            FILE* fp = std::fopen(name.c_str(), "wb");
            if (!fp)
            {
                throw std::runtime_error("cannot create " + name + ": " + std::strerror(errno));
            }
            const size_t w = std::fwrite(out.data(), 1, out.size(), fp);
            const int rc   = std::fclose(fp);
            if (w != out.size() || rc != 0)
            {
                throw std::runtime_error("cannot write " + name);
            }
        }

        static std::string join(const std::string& path, const std::string& filename)
        {
            if (path.empty())
            {
                return filename;
            }
            return path.back() == '/' ? path + filename : path + "/" + filename;
        }

      private:
        std::string path_, filename_;
        int dim_;
        File h5_;
        Node root_;
    };

    // the directory of the files, created if missing (fs::create_directory,
    // io/hdf5.hpp:1287-1290: one level)
    void make_dir(const std::string& path)
    {
        if (path.empty())
        {
            return;
        }
        struct stat st;
        if (stat(path.c_str(), &st) == 0)
        {
            require(S_ISDIR(st.st_mode), path + " is not a directory");
            return;
        }
        if (mkdir(path.c_str(), 0777) != 0 && errno != EEXIST)
        {
            throw std::runtime_error("cannot create directory " + path + ": " + std::strerror(errno));
        }
    }

    // the cells of one level of a sub-mesh (mesh[level], a LevelCellArray)
    Submesh level_of(const Submesh& sm, int level)
    {
        Submesh out;
        for (size_t i = 0; i < sm.cells.size(); ++i)
        {
            if (sm.cells[i][0] == level)
            {
                out.cells.push_back(sm.cells[i]);
                out.value_index.push_back(sm.value_index[i]);
            }
        }
        return out;
    }

    // SaveCellArray::save (io/hdf5.hpp:428-522) over named sub-meshes (the
    // mesh ids of an MRMesh, or the one "cell_array"); `main` is the sub-mesh
    // saved when the sub-meshes are not asked for
    void save_cell_array(Saver& s, int options, int min_level, int max_level, const Submesh& main,
                         const std::vector<std::pair<std::string, Submesh>>& subs, const std::vector<FieldData>& fields)
    {
        const bool by_level = options & SMR_SAVE_BY_LEVEL, by_id = options & SMR_SAVE_BY_MESH_ID;
        if (by_level)
        {
            // two coarser levels than the minimum, for the projection ghosts
            const int lo = std::max(0, min_level - 2);
            for (int level = lo; level <= max_level; ++level)
            {
                Node& g = s.domain().child("Grid");
                g.attr("Name", "Level " + std::to_string(level)).attr("GridType", "Collection").attr("CollectionType", "Spatial");
                if (by_id)
                {
                    for (const auto& [name, sm] : subs)
                    {
                        const std::string mesh_name = "level_" + std::to_string(level) + "_" + name;
                        s.save_on_mesh(g, "/level/" + std::to_string(level) + "/mesh/" + mesh_name, level_of(sm, level), mesh_name, fields);
                    }
                }
                else
                {
                    s.save_on_mesh(g, "/level/" + std::to_string(level) + "/mesh", level_of(main, level), "mesh", fields);
                }
            }
            return;
        }
        if (by_id)
        {
            for (const auto& [name, sm] : subs)
            {
                Node& g = s.domain().child("Grid");
                g.attr("Name", name).attr("GridType", "Collection");
                s.save_on_mesh(g, "/mesh/" + name, sm, name, fields);
            }
            return;
        }
        s.save_on_mesh(s.domain(), "/mesh", main, "mesh", fields);
    }

    Submesh mesh_cells(const smr_mesh* mesh, int id)
    {
        const int64_t n = smr_mesh_nb_cells(mesh, id);
        check_smr(n, "mesh size");
        std::vector<int64_t> rows(size_t(n) * 5);
        check_smr(smr_mesh_get_cells(mesh, id, rows.data()), "mesh cells");
        Submesh sm;
        sm.cells.resize(size_t(n));
        sm.value_index.resize(size_t(n));
        for (size_t i = 0; i < size_t(n); ++i)
        {
            sm.cells[i]       = {rows[i * 5], rows[i * 5 + 1], rows[i * 5 + 2], rows[i * 5 + 3]};
            sm.value_index[i] = rows[i * 5 + 4];
        }
        return sm;
    }

    void check_name(const char* name)
    {
        require(name && *name, "a field has no name");
        require(std::strchr(name, '/') == nullptr, std::string("field name ") + name + " holds a '/'");
    }
} // namespace

extern "C"
{
    int smr_save(const char* path, const char* filename, const smr_mesh* mesh, smr_field* const* fields, int nfields, int options)
    {
        return guarded(
            [&]
            {
                require(path && filename && *filename && mesh && nfields >= 0 && (nfields == 0 || fields), "null argument");
                require((options & ~(SMR_SAVE_BY_LEVEL | SMR_SAVE_BY_MESH_ID)) == 0, "unknown save option");
                smr_mesh_config cfg;
                check_smr(smr_mesh_get_config(mesh, &cfg), "mesh config");
                std::vector<FieldData> fd(static_cast<size_t>(nfields));
                for (int k = 0; k < nfields; ++k)
                {
                    require(fields[k] != nullptr, "null field");
                    const int on = smr_field_on_mesh(fields[k], mesh);
                    check_smr(on, "field");
                    require(on == 1, "a field does not live on the saved mesh");
                    check_name(smr_field_name(fields[k]));
                    fd[size_t(k)].name = smr_field_name(fields[k]);
                    const int64_t size = smr_field_size(fields[k]);
                    check_smr(size, "field size");
                    fd[size_t(k)].values.resize(size_t(size));
                    check_smr(smr_field_get_storage(fields[k], fd[size_t(k)].values.data()), "field values");
                }
                // MRMeshId names (mr/mesh.hpp:533-563)
                static const char* const kIdNames[] = {"cells", "cells and ghosts", "projection cells", "union cells", "reference"};
                std::vector<std::pair<std::string, Submesh>> subs;
                if (options & SMR_SAVE_BY_MESH_ID)
                {
                    for (int id = SMR_CELLS; id <= SMR_REFERENCE; ++id)
                    {
                        subs.emplace_back(kIdNames[id], mesh_cells(mesh, id));
                    }
                }
                const Submesh main = (options & SMR_SAVE_BY_MESH_ID) ? Submesh{} : mesh_cells(mesh, SMR_CELLS);
                make_dir(path);
                Saver s(path, filename, cfg.dim);
                save_cell_array(s, options, cfg.min_level, cfg.max_level, main, subs, fd);
                s.write_xdmf();
            });
    }

    int smr_save_cell_array(const char* path, const char* filename, int dim, const int32_t* cells, int64_t n, const double* const* values,
                            const char* const* names, int nfields, int options)
    {
        return guarded(
            [&]
            {
                require(path && filename && *filename && (dim == 2 || dim == 3) && n >= 0 && (n == 0 || cells) && nfields >= 0 &&
                            (nfields == 0 || (values && names)),
                        "bad argument");
                require((options & ~(SMR_SAVE_BY_LEVEL | SMR_SAVE_BY_MESH_ID)) == 0, "unknown save option");
                Submesh sm;
                sm.cells.resize(size_t(n));
                sm.value_index.resize(size_t(n));
                int lo = 21, hi = -1;
                for (int64_t i = 0; i < n; ++i)
                {
                    const int32_t* c = cells + 4 * i;
                    require(c[0] >= 0 && c[0] <= 20, "level out of range");
                    sm.cells[size_t(i)]       = {c[0], c[1], c[2], dim == 3 ? c[3] : 0};
                    sm.value_index[size_t(i)] = i;
                    if (i > 0)
                    {
                        const auto& p = sm.cells[size_t(i) - 1];
                        const auto& q = sm.cells[size_t(i)];
                        // for_each_cell order: level, z, y, x ascending
                        const bool ordered = p[0] < q[0] ||
                                             (p[0] == q[0] && (p[3] < q[3] || (p[3] == q[3] && (p[2] < q[2] || (p[2] == q[2] && p[1] < q[1])))));
                        require(ordered, "cells not in for_each_cell order");
                    }
                    lo = std::min(lo, int(c[0]));
                    hi = std::max(hi, int(c[0]));
                }
                std::vector<FieldData> fd(static_cast<size_t>(nfields));
                for (int k = 0; k < nfields; ++k)
                {
                    check_name(names[k]);
                    require(n == 0 || values[k], "null field values");
                    fd[size_t(k)].name = names[k];
                    fd[size_t(k)].values.assign(values[k], values[k] + n);
                }
                make_dir(path);
                Saver s(path, filename, dim);
                // CellArray::min_level/max_level (cell_array.hpp:431-452): the
                // non-empty levels
                save_cell_array(s, options, hi < 0 ? 21 : lo, hi < 0 ? 0 : hi, sm, {{"cell_array", sm}}, fd);
                s.write_xdmf();
            });
    }
}
