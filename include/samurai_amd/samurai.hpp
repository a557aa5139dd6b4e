// C++ drop-in layer over the C ABI (include/samurai_amd.h) with the names and
// call shapes of samurai's API for the adaptive time step, so that the loop of
// demos/FiniteVolume/advection_2d.cpp compiles against it almost unchanged:
//
//   samurai::initialize(argc, argv);                          samurai.hpp:25-113
//   auto mesh = samurai::mra::make_mesh(box, config);         mr/mesh.hpp:481-529
//   auto u = samurai::make_scalar_field<double>("u", mesh);   scalar_field.hpp:171-176
//   samurai::for_each_cell(mesh, [&](auto& cell) {...});      algorithm.hpp:319-323
//   samurai::make_bc<samurai::Dirichlet<1>>(u, 0.);           bc/bc.hpp:771-795
//   auto MRadaptation = samurai::make_MRAdapt(u);             mr/adapt.hpp:391-397
//   MRadaptation(samurai::mra_config().epsilon(e).regularity(r));
//   samurai::update_ghost_mr(u);                              update_ghost_mr.hpp:260-270
//   unp1 = u - dt * samurai::upwind(a, u);                    stencil_field.hpp:175-179
//   unp1 = u - dt * samurai::upwind_scalar_burgers(k, u);     stencil_field.hpp:181-255
//   auto conv = samurai::make_convection_upwind<decltype(u)>(v);
//   unp1 = u - dt * conv(u);                                  convection_lin.hpp:15-89
//   samurai::swap(u, unp1);                                   field/swap.hpp:12-19
//
// Everything runs through the ABI on the GPU; values written through u[cell]
// live in a host mirror that is uploaded before the next device operation,
// and reading u[cell] after a device operation downloads the storage first.
// The domain is the unit box [0, 1]^dim (the box of the demos).
#pragma once

#include "../samurai_amd.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace samurai_amd
{
    inline void check(int rc)
    {
        if (rc >= 0)
        {
            return;
        }
        const std::string msg = smr_last_error();
        if (rc == SMR_ERR_RANGE)
        {
            throw std::out_of_range(msg);
        }
        if (rc == SMR_ERR_ARGUMENT)
        {
            throw std::invalid_argument(msg);
        }
        throw std::runtime_error(msg);
    }

    // ------------------------------------------------------------------
    // context: samurai::initialize / finalize
    // ------------------------------------------------------------------
    // the library's command-line options (arguments.hpp:11-89) that the path
    // reads: mesh_config and mra_config take them over their own values unless
    // unset (the reference's parse_args, mesh_config.hpp:356-397,
    // mr/config.hpp:47-61); initialize(argc, argv) fills them
    namespace args
    {
        inline int min_level          = -1; // --min-level
        inline int max_level          = -1; // --max-level
        inline int graduation_width   = -1; // --graduation-width
        inline int max_stencil_radius = -1; // --max-stencil-radius
        inline bool refine_boundary   = false; // --refine-boundary
        inline bool timers            = false; // --timers
        inline double epsilon         = std::numeric_limits<double>::infinity(); // --mr-eps
        inline double regularity      = std::numeric_limits<double>::infinity(); // --mr-reg
        inline bool rel_detail        = false; // --mr-rel-detail

        // --name value, --name=value, or a bare flag; other arguments are left
        // to the program
        inline void parse(int argc, char** argv)
        {
            for (int i = 1; i < argc; ++i)
            {
                std::string a = argv[i], v;
                const auto eq = a.find('=');
                if (eq != std::string::npos)
                {
                    v = a.substr(eq + 1);
                    a = a.substr(0, eq);
                }
                auto value = [&]() -> std::string
                {
                    if (!v.empty())
                    {
                        return v;
                    }
                    if (i + 1 >= argc)
                    {
                        throw std::invalid_argument(a + " needs a value");
                    }
                    return argv[++i];
                };
                if (a == "--min-level") min_level = std::stoi(value());
                else if (a == "--max-level") max_level = std::stoi(value());
                else if (a == "--graduation-width") graduation_width = std::stoi(value());
                else if (a == "--max-stencil-radius") max_stencil_radius = std::stoi(value());
                else if (a == "--mr-eps") epsilon = std::stod(value());
                else if (a == "--mr-reg") regularity = std::stod(value());
                else if (a == "--refine-boundary") refine_boundary = true;
                else if (a == "--mr-rel-detail") rel_detail = true;
                else if (a == "--timers") timers = true;
            }
        }
    } // namespace args

    inline std::shared_ptr<smr_ctx>& context_slot()
    {
        static std::shared_ptr<smr_ctx> ctx;
        return ctx;
    }

    inline void initialize(int device = 0)
    {
        smr_ctx* c = nullptr;
        check(smr_ctx_create(device, &c));
        context_slot().reset(c, [](smr_ctx* p) { smr_ctx_destroy(p); });
    }

    inline void initialize(int& argc, char** argv)
    {
        args::parse(argc, argv);
        initialize(0);
        if (args::timers)
        {
            check(smr_timers_enable(1));
        }
    }

    inline void finalize()
    {
        if (args::timers)
        {
            // the timers' table at the end of the program (samurai.hpp's finalize)
            std::string buf(1 << 16, '\0');
            const int64_t n = smr_timers_report(buf.data(), int64_t(buf.size()));
            if (n > 0)
            {
                std::fwrite(buf.data(), 1, size_t(std::min<int64_t>(n, int64_t(buf.size()))), stdout);
                std::fputc('\n', stdout);
            }
        }
        context_slot().reset();
    }

    inline smr_ctx* context()
    {
        if (!context_slot())
        {
            initialize(0);
        }
        return context_slot().get();
    }

    // shared ownership of the context: meshes and fields keep it alive, so
    // finalize() before they go out of scope destroys it after them
    inline std::shared_ptr<smr_ctx> context_holder()
    {
        context();
        return context_slot();
    }

    // ------------------------------------------------------------------
    // configuration
    // ------------------------------------------------------------------
    template <class T, std::size_t dim>
    struct Box
    {
        std::array<T, dim> min_corner, max_corner;
        Box(const std::array<T, dim>& a, const std::array<T, dim>& b)
            : min_corner(a)
            , max_corner(b)
        {
        }
    };

    enum class mesh_id_t : int
    {
        cells            = SMR_CELLS,
        cells_and_ghosts = SMR_CELLS_AND_GHOSTS,
        proj_cells       = SMR_PROJ_CELLS,
        union_cells      = SMR_UNION_CELLS,
        reference        = SMR_REFERENCE
    };

    // mesh_config<dim> (mesh_config.hpp:16-432): the options this path uses
    template <std::size_t dim_, int prediction_stencil_radius_ = 1>
    class mesh_config
    {
      public:
        static constexpr std::size_t dim = dim_;
        mesh_config& min_level(int l)
        {
            c_.min_level = l;
            return *this;
        }
        mesh_config& max_level(int l)
        {
            c_.max_level = l;
            return *this;
        }
        mesh_config& max_stencil_size(int s)
        {
            c_.max_stencil_radius = s / 2 + s % 2; // mesh_config.hpp:71-78
            return *this;
        }
        mesh_config& max_stencil_radius(int r)
        {
            c_.max_stencil_radius = r;
            return *this;
        }
        mesh_config& disable_minimal_ghost_width()
        {
            return *this;
        }
        mesh_config& graduation_width(int w)
        {
            c_.graduation_width = w;
            return *this;
        }
        mesh_config& disable_args_parse()
        {
            args_ = false;
            return *this;
        }
        // the configuration with the command-line options taken over
        smr_mesh_config raw() const
        {
            smr_mesh_config c = c_;
            if (args_)
            {
                if (args::min_level >= 0) c.min_level = args::min_level;
                if (args::max_level >= 0) c.max_level = args::max_level;
                if (args::graduation_width >= 0) c.graduation_width = args::graduation_width;
                if (args::max_stencil_radius >= 0) c.max_stencil_radius = args::max_stencil_radius;
            }
            return c;
        }

      private:
        smr_mesh_config c_{int(dim_), 0, 6, prediction_stencil_radius_, 1, 1};
        bool args_ = true;
    };

    // ------------------------------------------------------------------
    // mesh
    // ------------------------------------------------------------------
    template <std::size_t dim_>
    struct Cell
    {
        static constexpr std::size_t dim = dim_;
        int level;
        std::array<int, dim_> indices;
        int64_t index; // storage index
        double length;
        std::array<double, dim_> center() const
        {
            std::array<double, dim_> c;
            for (std::size_t d = 0; d < dim_; ++d)
            {
                c[d] = length * (indices[d] + 0.5);
            }
            return c;
        }
    };

    template <std::size_t dim_>
    class MRMesh
    {
      public:
        static constexpr std::size_t dim = dim_;
        explicit MRMesh(const smr_mesh_config& cfg)
            : ctx_(context_holder())
        {
            smr_mesh* m = nullptr;
            check(smr_mesh_create_box(ctx_.get(), &cfg, &m));
            h_.reset(m, [](smr_mesh* p) { smr_mesh_destroy(p); });
        }
        // mra::make_empty_mesh(cfg): no cells, to be assigned or loaded
        struct empty_tag
        {
        };
        MRMesh(const smr_mesh_config& cfg, empty_tag)
            : ctx_(context_holder())
        {
            smr_mesh* m = nullptr;
            check(smr_mesh_create_from_cells(ctx_.get(), &cfg, nullptr, 0, &m));
            h_.reset(m, [](smr_mesh* p) { smr_mesh_destroy(p); });
        }
        // a copy is a new mesh with the same cells and configuration (as
        // samurai's MRMesh copy), independent of the original
        MRMesh(const MRMesh& o)
            : ctx_(o.ctx_)
        {
            smr_mesh_config cfg{};
            check(smr_mesh_get_config(o.h_.get(), &cfg));
            const int64_t n = smr_mesh_nb_cells(o.h_.get(), SMR_CELLS);
            std::vector<int64_t> raw(size_t(std::max<int64_t>(n, 0)) * 5);
            check(smr_mesh_get_cells(o.h_.get(), SMR_CELLS, raw.data()));
            std::vector<int32_t> rows(size_t(std::max<int64_t>(n, 0)) * 4);
            for (int64_t k = 0; k < n; ++k)
            {
                for (int d = 0; d < 4; ++d)
                {
                    rows[size_t(4 * k + d)] = int32_t(raw[size_t(5 * k + d)]);
                }
            }
            smr_mesh* m = nullptr;
            check(smr_mesh_create_from_cells(ctx_.get(), &cfg, rows.data(), n, &m));
            h_.reset(m, [](smr_mesh* p) { smr_mesh_destroy(p); });
        }
        MRMesh(MRMesh&&)                 = default;
        MRMesh& operator=(const MRMesh&) = delete;
        // mesh = mra::make_mesh(...): the new contents are swapped into this
        // mesh, so the fields made on it stay attached to it (as in samurai,
        // where fields hold a reference to the mesh object); resize them
        MRMesh& operator=(MRMesh&& o)
        {
            if (h_ && o.h_)
            {
                check(smr_mesh_swap(h_.get(), o.h_.get()));
            }
            else
            {
                ctx_ = o.ctx_;
                h_   = std::move(o.h_);
            }
            return *this;
        }
        // kept for callers of the earlier interface: the configuration is
        // always read from the mesh itself
        void refresh()
        {
        }
        smr_mesh* handle() const
        {
            return h_.get();
        }
        // shared ownership of the mesh for the objects that use it
        const std::shared_ptr<smr_mesh>& shared_handle() const
        {
            return h_;
        }
        const std::shared_ptr<smr_ctx>& shared_context() const
        {
            return ctx_;
        }
        std::size_t nb_cells(mesh_id_t id = mesh_id_t::cells) const
        {
            return std::size_t(smr_mesh_nb_cells(h_.get(), int(id)));
        }
        smr_mesh_config config() const
        {
            smr_mesh_config c{};
            check(smr_mesh_get_config(h_.get(), &c));
            return c;
        }
        int min_level() const
        {
            return config().min_level;
        }
        int max_level() const
        {
            return config().max_level;
        }
        double cell_length(int level) const
        {
            return 1.0 / double(1 << level); // cell.hpp:17-20 on the unit box
        }
        double min_cell_length() const
        {
            return cell_length(max_level());
        }
        std::vector<Cell<dim_>> cells(mesh_id_t id = mesh_id_t::cells) const
        {
            const int64_t n = smr_mesh_nb_cells(h_.get(), int(id));
            std::vector<int64_t> raw(size_t(std::max<int64_t>(n, 0)) * 5);
            check(smr_mesh_get_cells(h_.get(), int(id), raw.data()));
            std::vector<Cell<dim_>> out(size_t(std::max<int64_t>(n, 0)));
            for (int64_t k = 0; k < n; ++k)
            {
                const int64_t* r = &raw[size_t(5 * k)];
                Cell<dim_>& c    = out[size_t(k)];
                c.level          = int(r[0]);
                for (std::size_t d = 0; d < dim_; ++d)
                {
                    c.indices[d] = int(r[1 + d]);
                }
                c.index  = r[4];
                c.length = cell_length(c.level);
            }
            return out;
        }

      private:
        // the context outlives the mesh (members are destroyed last to first)
        std::shared_ptr<smr_ctx> ctx_;
        std::shared_ptr<smr_mesh> h_;
    };

    namespace mra
    {
        template <class T, std::size_t dim, class Config>
        MRMesh<dim> make_mesh(const Box<T, dim>& box, const Config& cfg)
        {
            for (std::size_t d = 0; d < dim; ++d)
            {
                if (box.min_corner[d] != 0 || box.max_corner[d] != 1)
                {
                    throw std::invalid_argument("make_mesh: the domain must be the unit box");
                }
            }
            return MRMesh<dim>(cfg.raw());
        }

        template <class Config>
        MRMesh<Config::dim> make_empty_mesh(const Config& cfg)
        {
            return MRMesh<Config::dim>(cfg.raw(), typename MRMesh<Config::dim>::empty_tag{});
        }
    } // namespace mra

    template <class Mesh, class F>
    void for_each_cell(const Mesh& mesh, F&& f)
    {
        for (auto& c : mesh.cells())
        {
            f(c);
        }
    }

    // ------------------------------------------------------------------
    // fields and expressions
    // ------------------------------------------------------------------
    template <std::size_t dim_>
    class ScalarField;

    template <std::size_t dim_>
    struct UpwindExpr
    {
        std::array<double, 3> a;
        ScalarField<dim_>* u;
        bool scalar_burgers = false; // upwind_scalar_burgers(k, u) instead of upwind(a, u)
    };
    template <std::size_t dim_>
    struct ScaledUpwind
    {
        double dt;
        UpwindExpr<dim_> e;
    };
    template <std::size_t dim_>
    struct UpwindStep
    {
        ScalarField<dim_>* u;
        ScaledUpwind<dim_> s;
    };
    template <std::size_t dim_>
    struct ScaledField
    {
        double a;
        ScalarField<dim_>* f;
    };
    template <std::size_t dim_>
    struct FieldStep
    {
        ScalarField<dim_>* u;
        ScaledField<dim_> s;
    };

    template <std::size_t dim_>
    class ScalarField
    {
      public:
        static constexpr std::size_t dim = dim_;
        ScalarField(const ScalarField&)            = delete;
        ScalarField& operator=(const ScalarField&) = delete;
        ScalarField(ScalarField&&)                 = default;
        ScalarField& operator=(ScalarField&&)      = default;
        ScalarField(std::string name, MRMesh<dim_>& mesh)
            : name_(std::move(name))
            , mesh_(&mesh)
            , ctx_(mesh.shared_context())
            , mesh_h_(mesh.shared_handle())
        {
            smr_field* f = nullptr;
            check(smr_field_create(mesh.handle(), name_.c_str(), &f));
            h_.reset(f, [](smr_field* p) { smr_field_destroy(p); });
        }
        smr_field* handle()
        {
            flush();
            return h_.get();
        }
        const std::shared_ptr<smr_field>& shared_handle() const
        {
            return h_;
        }
        MRMesh<dim_>& mesh()
        {
            return *mesh_;
        }
        const std::string& name() const
        {
            return name_;
        }
        void resize()
        {
            check(smr_field_resize(h_.get()));
            host_.clear();
            dirty_ = false;
        }
        void fill(double v)
        {
            pull();
            std::fill(host_.begin(), host_.end(), v);
            dirty_ = true;
        }
        // host access by cell (storage index), as u[cell]
        double& operator[](const Cell<dim_>& c)
        {
            pull();
            dirty_ = true;
            return host_.at(size_t(c.index));
        }
        // the device storage changed: the host mirror is stale
        void touched()
        {
            host_.clear();
        }
        ScalarField& operator=(const UpwindStep<dim_>& st)
        {
            if (st.u == this)
            {
                throw std::invalid_argument("u = u - dt * upwind(a, u) needs a second field");
            }
            if (st.s.e.scalar_burgers)
            {
                check(smr_upwind_scalar_burgers(st.s.e.a.data(), st.s.dt, st.u->handle(), handle()));
            }
            else
            {
                check(smr_upwind(st.s.e.a.data(), st.s.dt, st.u->handle(), handle()));
            }
            touched();
            return *this;
        }
        ScalarField& operator=(const FieldStep<dim_>& st)
        {
            check(smr_field_sub_scaled(st.u->handle(), st.s.a, st.s.f->handle(), handle()));
            touched();
            return *this;
        }

      private:
        void pull()
        {
            if (host_.empty())
            {
                const int64_t n = smr_field_size(h_.get());
                host_.assign(size_t(std::max<int64_t>(n, 0)), 0.0);
                check(smr_field_get_storage(h_.get(), host_.data()));
            }
        }
        void flush()
        {
            if (dirty_)
            {
                check(smr_field_set_storage(h_.get(), host_.data()));
                dirty_ = false;
            }
        }
        std::string name_;
        MRMesh<dim_>* mesh_;
        // the context and the mesh outlive the field
        std::shared_ptr<smr_ctx> ctx_;
        std::shared_ptr<smr_mesh> mesh_h_;
        std::shared_ptr<smr_field> h_;
        std::vector<double> host_;
        bool dirty_ = false;
    };

    template <class T, std::size_t dim>
    ScalarField<dim> make_scalar_field(const std::string& name, MRMesh<dim>& mesh)
    {
        return ScalarField<dim>(name, mesh);
    }

    // make_field<double, 1> (README.md:104)
    template <class T, std::size_t n_comp, std::size_t dim>
    ScalarField<dim> make_field(const std::string& name, MRMesh<dim>& mesh)
    {
        static_assert(n_comp == 1, "scalar fields only");
        return ScalarField<dim>(name, mesh);
    }

    template <std::size_t order>
    struct Dirichlet
    {
    };

    template <class BC, std::size_t dim>
    void make_bc(ScalarField<dim>& u, double value)
    {
        check(smr_field_attach_dirichlet(u.handle(), value));
    }

    template <std::size_t dim>
    void update_ghost_mr(ScalarField<dim>& u)
    {
        check(smr_update_ghost_mr(u.handle()));
        u.touched();
    }

    // update_ghost_mr(u, v...) (algorithm/update_ghost_mr.hpp:260-270)
    template <std::size_t dim, class... Fields>
    void update_ghost_mr(ScalarField<dim>& u, ScalarField<dim>& v, Fields&... others)
    {
        update_ghost_mr(u);
        update_ghost_mr(v, others...);
    }

    template <std::size_t dim>
    void swap(ScalarField<dim>& u, ScalarField<dim>& v)
    {
        check(smr_field_swap(u.handle(), v.handle()));
        u.touched();
        v.touched();
    }

    template <class Vec, std::size_t dim>
    UpwindExpr<dim> upwind(const Vec& a, ScalarField<dim>& u)
    {
        UpwindExpr<dim> e{{0, 0, 0}, &u};
        for (std::size_t d = 0; d < dim; ++d)
        {
            e.a[d] = a[d];
        }
        return e;
    }
    // upwind_scalar_burgers(k, u) (stencil_field.hpp:181-255), 2D
    template <class Vec>
    UpwindExpr<2> upwind_scalar_burgers(const Vec& k, ScalarField<2>& u)
    {
        UpwindExpr<2> e{{k[0], k[1], 0}, &u, true};
        return e;
    }
    template <std::size_t dim>
    ScaledUpwind<dim> operator*(double dt, const UpwindExpr<dim>& e)
    {
        return {dt, e};
    }
    template <std::size_t dim>
    UpwindStep<dim> operator-(ScalarField<dim>& u, const ScaledUpwind<dim>& s)
    {
        return {&u, s};
    }
    template <std::size_t dim>
    ScaledField<dim> operator*(double a, ScalarField<dim>& f)
    {
        return {a, &f};
    }
    template <std::size_t dim>
    FieldStep<dim> operator-(ScalarField<dim>& u, const ScaledField<dim>& s)
    {
        return {&u, s};
    }

    // make_convection_upwind<Field>(velocity) (convection_lin.hpp:15-89)
    template <class Field>
    class ConvectionUpwind
    {
      public:
        explicit ConvectionUpwind(const std::array<double, Field::dim>& v)
            : v_(v)
        {
        }
        Field operator()(Field& u) const
        {
            Field out("convection(" + u.name() + ")", u.mesh());
            check(smr_convection_upwind(v_.data(), u.handle(), out.handle()));
            u.touched();
            out.touched();
            return out;
        }

      private:
        std::array<double, Field::dim> v_;
    };
    // make_convection_upwind<Field>() (convection_nonlin.hpp:24-79), scalar fields
    template <class Field>
    class ConvectionUpwindNonLinear
    {
      public:
        Field operator()(Field& u) const
        {
            Field out("convection(" + u.name() + ")", u.mesh());
            check(smr_convection_upwind_nonlinear(u.handle(), out.handle()));
            u.touched();
            out.touched();
            return out;
        }
    };
    template <class Field>
    ConvectionUpwindNonLinear<Field> make_convection_upwind()
    {
        return {};
    }

    template <class Field, class Vec>
    ConvectionUpwind<Field> make_convection_upwind(const Vec& v)
    {
        std::array<double, Field::dim> a{};
        for (std::size_t d = 0; d < Field::dim; ++d)
        {
            a[d] = v[d];
        }
        return ConvectionUpwind<Field>(a);
    }

    // make_convection_weno5<Field>(velocity) (convection_lin.hpp:95-178): a mesh
    // of one level built with max_stencil_size(6)
    template <class Field>
    class ConvectionWeno5
    {
      public:
        explicit ConvectionWeno5(const std::array<double, Field::dim>& v)
            : v_(v)
        {
        }
        Field operator()(Field& u) const
        {
            Field out("convection(" + u.name() + ")", u.mesh());
            double v[3] = {0, 0, 0};
            for (std::size_t d = 0; d < Field::dim; ++d)
            {
                v[d] = v_[d];
            }
            check(smr_convection_weno5(v, u.handle(), out.handle()));
            u.touched();
            out.touched();
            return out;
        }

      private:
        std::array<double, Field::dim> v_;
    };

    template <class Field, class Vec>
    ConvectionWeno5<Field> make_convection_weno5(const Vec& v)
    {
        std::array<double, Field::dim> a{};
        for (std::size_t d = 0; d < Field::dim; ++d)
        {
            a[d] = v[d];
        }
        return ConvectionWeno5<Field>(a);
    }

    // mra_config (mr/config.hpp:10-68)
    class mra_config
    {
      public:
        mra_config& epsilon(double e)
        {
            eps_ = e;
            return *this;
        }
        mra_config& regularity(double r)
        {
            reg_ = r;
            return *this;
        }
        double epsilon() const
        {
            return eps_;
        }
        double regularity() const
        {
            return reg_;
        }
        mra_config& relative_detail(bool r)
        {
            rel_ = r;
            return *this;
        }
        bool relative_detail() const
        {
            return rel_;
        }

      private:
        double eps_ = 1e-4, reg_ = 1.0; // mr/config.hpp defaults
        bool rel_   = false;
    };

    // Adapt (mr/adapt.hpp:96-146) over one or more fields on one mesh
    template <std::size_t dim>
    class MRAdapt
    {
      public:
        explicit MRAdapt(std::vector<ScalarField<dim>*> fields)
            : fields_(std::move(fields))
        {
            std::vector<smr_field*> hs;
            for (ScalarField<dim>* f : fields_)
            {
                hs.push_back(f->handle());
                field_h_.push_back(f->shared_handle());
            }
            smr_adapt* a = nullptr;
            check(smr_adapt_create_fields(hs.data(), int(hs.size()), &a));
            h_.reset(a, [](smr_adapt* p) { smr_adapt_destroy(p); });
        }
        void operator()(const mra_config& cfg)
        {
            // mra_config::parse_args (mr/config.hpp:47-61)
            const double eps = args::epsilon != std::numeric_limits<double>::infinity() ? args::epsilon : cfg.epsilon();
            const double reg = args::regularity != std::numeric_limits<double>::infinity() ? args::regularity : cfg.regularity();
            check(smr_adapt_set_relative_detail(h_.get(), (cfg.relative_detail() || args::rel_detail) ? 1 : 0));
            run(eps, reg);
        }
        void operator()(double eps, double regularity)
        {
            check(smr_adapt_set_relative_detail(h_.get(), 0));
            run(eps, regularity);
        }

      private:
        void run(double eps, double regularity)
        {
            check(smr_adapt_set_refine_boundary(h_.get(), args::refine_boundary ? 1 : 0));
            for (ScalarField<dim>* f : fields_)
            {
                f->handle(); // host-side writes reach the device first
            }
            check(smr_adapt_run(h_.get(), eps, regularity));
            for (ScalarField<dim>* f : fields_)
            {
                f->touched();
            }
        }

        std::vector<ScalarField<dim>*> fields_;
        // the fields outlive the adaptation object
        std::vector<std::shared_ptr<smr_field>> field_h_;
        std::shared_ptr<smr_adapt> h_;
    };

    // make_MRAdapt(u, v...) (mr/adapt.hpp:391-397)
    template <std::size_t dim, class... Fields>
    MRAdapt<dim> make_MRAdapt(ScalarField<dim>& u, Fields&... others)
    {
        return MRAdapt<dim>(std::vector<ScalarField<dim>*>{&u, &others...});
    }
} // namespace samurai_amd

#ifdef SAMURAI_AMD_AS_SAMURAI
namespace samurai = samurai_amd;
#endif
