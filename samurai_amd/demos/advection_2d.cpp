// The 2D advection demo on the C++ drop-in layer: a disc of radius 0.2 at
// (0.3, 0.3) advected with velocity (1, 1) on an adaptive MR mesh, Dirichlet
// 0, first-order upwind, from Ti until Tf with the last step shortened to land
// on Tf (demos/FiniteVolume/advection_2d.cpp:22-153; same defaults: levels
// 4..10, eps 2e-4, regularity 1, cfl 0.5).
//
//   advection_2d [--min-level L] [--max-level L] [--mr-eps e] [--mr-reg r]
//                [--Ti t] [--Tf t] [--cfl c] [--prediction-order 0|1]
//                [--path dir] [--filename name] [--nfiles n]
//                [--restart-file file] [--out prefix]
//
// As the reference demo, it writes {path}/{filename}_pred_{order}{suffix}.h5
// and .xdmf (samurai::save) and the restart files {..}_restart{suffix}.h5
// (samurai::dump) after the initial adaptation (suffix _init) and at nfiles
// times up to Tf, and starts from a restart file when one is given.  With
// --out, the cells (level, i, j) and values after the initial adaptation and
// at Tf are also written as text to <prefix>_init.txt and <prefix>_final.txt.
#include <samurai_amd/samurai.hpp>
#ifdef SAMURAI_AMD_WITH_IO
#include <samurai_amd/io.hpp>
#endif

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <string>

namespace sa = samurai_amd;

template <class Field>
static void dump_text(Field& u, const std::string& path)
{
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f)
    {
        throw std::runtime_error("cannot write " + path);
    }
    sa::for_each_cell(u.mesh(),
                      [&](const auto& cell)
                      {
                          std::fprintf(f, "%d %d %d %.17g\n", cell.level, cell.indices[0], cell.indices[1], u[cell]);
                      });
    std::fclose(f);
}

// demos/FiniteVolume/advection_2d.cpp:22-45
template <class Field>
static void init(Field& u)
{
    auto& mesh = u.mesh();
    u.resize();
    sa::for_each_cell(mesh,
                      [&](const auto& cell)
                      {
                          const auto c   = cell.center();
                          const double r = 0.2, xc = 0.3, yc = 0.3;
                          u[cell]        = ((c[0] - xc) * (c[0] - xc) + (c[1] - yc) * (c[1] - yc) <= r * r) ? 1.0 : 0.0;
                      });
}

// demos/FiniteVolume/advection_2d.cpp:47-59
template <class Field>
static void save(const std::string& path, const std::string& filename, Field& u, const std::string& suffix = "")
{
#ifdef SAMURAI_AMD_WITH_IO
    auto& mesh = u.mesh();
    sa::save(path, filename + suffix, mesh, u);
    sa::dump(path, filename + "_restart" + suffix, mesh, u);
#else
    (void)path, (void)filename, (void)u, (void)suffix;
#endif
}

struct Options
{
    int min_level = 4, max_level = 10;
    double eps = 2e-4, reg = 1., Ti = 0., Tf = 0.01, cfl = 0.5;
    std::string path = std::filesystem::current_path().string(), filename = "FV_advection_2d", restart_file, out;
    int nfiles = 1;
};

template <int S>
static int run(const Options& o)
{
    constexpr std::size_t dim = 2;
    const sa::Box<double, dim> box({0., 0.}, {1., 1.});
    auto config = sa::mesh_config<dim, S>().min_level(o.min_level).max_level(o.max_level).max_stencil_size(2).disable_minimal_ghost_width();
    auto mesh   = sa::mra::make_empty_mesh(config);
    auto u      = sa::make_scalar_field<double>("u", mesh);
    const std::string filename = o.filename + "_pred_" + std::to_string(S);

    if (o.restart_file.empty())
    {
        mesh = sa::mra::make_mesh(box, config);
        init(u);
    }
    else
    {
#ifdef SAMURAI_AMD_WITH_IO
        sa::load(o.restart_file, mesh, u);
#else
        throw std::runtime_error("--restart-file needs the restart I/O library (built with HDF5)");
#endif
    }
    sa::make_bc<sa::Dirichlet<1>>(u, 0.);

    double dt            = o.cfl * mesh.min_cell_length();
    const double dt_save = o.Tf / static_cast<double>(o.nfiles);
    auto unp1            = sa::make_scalar_field<double>("unp1", mesh);
    const std::array<double, dim> a{1., 1.};

    auto MRadaptation = sa::make_MRAdapt(u);
    auto mra_config   = sa::mra_config().epsilon(o.eps).regularity(o.reg);
    MRadaptation(mra_config);
    save(o.path, filename, u, "_init");
    if (!o.out.empty())
    {
        dump_text(u, o.out + "_init.txt");
    }

    double t  = o.Ti;
    int nsave = 1, nstep = 0;
    while (t != o.Tf)
    {
        MRadaptation(mra_config);
        t += dt;
        if (t > o.Tf)
        {
            dt += o.Tf - t;
            t = o.Tf;
        }
        sa::update_ghost_mr(u);
        unp1.resize();
        unp1 = u - dt * sa::upwind(a, u);
        sa::swap(u, unp1);
        ++nstep;
        if (t >= static_cast<double>(nsave) * dt_save || t == o.Tf)
        {
            const std::string suffix = (o.nfiles != 1) ? "_ite_" + std::to_string(nsave++) : "";
            save(o.path, filename, u, suffix);
        }
    }
    std::printf("steps %d cells %zu\n", nstep, mesh.nb_cells());
    if (!o.out.empty())
    {
        dump_text(u, o.out + "_final.txt");
    }
    return 0;
}

int main(int argc, char** argv)
{
    Options o;
    int order = 1;
    for (int i = 1; i + 1 < argc; i += 2)
    {
        const std::string k = argv[i];
        const char* v       = argv[i + 1];
        if (k == "--min-level")
        {
            o.min_level = std::atoi(v);
        }
        else if (k == "--max-level")
        {
            o.max_level = std::atoi(v);
        }
        else if (k == "--mr-eps")
        {
            o.eps = std::atof(v);
        }
        else if (k == "--mr-reg")
        {
            o.reg = std::atof(v);
        }
        else if (k == "--Ti")
        {
            o.Ti = std::atof(v);
        }
        else if (k == "--Tf")
        {
            o.Tf = std::atof(v);
        }
        else if (k == "--cfl")
        {
            o.cfl = std::atof(v);
        }
        else if (k == "--prediction-order")
        {
            order = std::atoi(v);
        }
        else if (k == "--path")
        {
            o.path = v;
        }
        else if (k == "--filename")
        {
            o.filename = v;
        }
        else if (k == "--nfiles")
        {
            o.nfiles = std::atoi(v);
        }
        else if (k == "--restart-file")
        {
            o.restart_file = v;
        }
        else if (k == "--out")
        {
            o.out = v;
        }
        else
        {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            return 2;
        }
    }
    try
    {
        sa::initialize(argc, argv);
        const int rc = order == 0 ? run<0>(o) : run<1>(o);
        sa::finalize();
        return rc;
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
