// The scalar Burgers demo on the C++ drop-in layer
// (demos/FiniteVolume/scalar_burgers_2d.cpp:20-170): +1 in the disc of radius
// 0.1 at (0.5, 0.5), -1 in the one at (0.2, 0.2), Dirichlet 0, velocity
// k = (sqrt(2)/2, sqrt(2)/2), unp1 = u - dt * upwind_scalar_burgers(k, u) on an
// adaptive MR mesh (levels 4..10, eps 2e-4, cfl 0.05, Tf 0.1 by default).
//
//   scalar_burgers_2d [--min-level L] [--max-level L] [--Tf t] [--cfl c]
//                     [--path dir] [--filename name] [--nfiles n]
//                     [--restart-file file] [--out prefix]
//
// It saves {path}/{filename}{suffix}.h5/.xdmf after the initial adaptation
// (_init) and at nfiles times (samurai::save), and with --out writes the
// cells (level, i, j) and values at Tf as text to <prefix>_final.txt.
#include <samurai_amd/samurai.hpp>
#ifdef SAMURAI_AMD_WITH_IO
#include <samurai_amd/io.hpp>
#endif

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <string>

namespace sa = samurai_amd;

template <class Field>
static void init(Field& u)
{
    auto& mesh = u.mesh();
    u.resize();
    sa::for_each_cell(mesh,
                      [&](const auto& cell)
                      {
                          const auto c = cell.center();
                          double v     = ((c[0] - 0.5) * (c[0] - 0.5) + (c[1] - 0.5) * (c[1] - 0.5) <= 0.1 * 0.1) ? 1.0 : 0.0;
                          if ((c[0] - 0.2) * (c[0] - 0.2) + (c[1] - 0.2) * (c[1] - 0.2) <= 0.1 * 0.1)
                          {
                              v = -1.0;
                          }
                          u[cell] = v;
                      });
}

template <class Field>
static void save(const std::string& path, const std::string& filename, Field& u, const std::string& suffix = "")
{
#ifdef SAMURAI_AMD_WITH_IO
    sa::save(path, filename + suffix, u.mesh(), u);
#else
    (void)path, (void)filename, (void)u, (void)suffix;
#endif
}

int main(int argc, char** argv)
{
    int min_level = 4, max_level = 10, nfiles = 1;
    double Tf = 0.1, cfl = 0.05;
    std::string path = std::filesystem::current_path().string(), filename = "FV_scalar_burgers_2d", restart_file, out;
    for (int i = 1; i + 1 < argc; i += 2)
    {
        const std::string k = argv[i];
        const char* v       = argv[i + 1];
        if (k == "--min-level")
        {
            min_level = std::atoi(v);
        }
        else if (k == "--max-level")
        {
            max_level = std::atoi(v);
        }
        else if (k == "--Tf")
        {
            Tf = std::atof(v);
        }
        else if (k == "--cfl")
        {
            cfl = std::atof(v);
        }
        else if (k == "--path")
        {
            path = v;
        }
        else if (k == "--filename")
        {
            filename = v;
        }
        else if (k == "--nfiles")
        {
            nfiles = std::atoi(v);
        }
        else if (k == "--restart-file")
        {
            restart_file = v;
        }
        else if (k == "--out")
        {
            out = v;
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
        {
            constexpr std::size_t dim = 2;
            const sa::Box<double, dim> box({0., 0.}, {1., 1.});
            const std::array<double, 2> k{std::sqrt(2.) / 2., std::sqrt(2.) / 2.};
            auto config = sa::mesh_config<dim>().min_level(min_level).max_level(max_level).disable_minimal_ghost_width();
            auto mesh   = sa::mra::make_empty_mesh(config);
            auto u      = sa::make_scalar_field<double>("u", mesh);
            if (restart_file.empty())
            {
                mesh = sa::mra::make_mesh(box, config);
                init(u);
            }
            else
            {
#ifdef SAMURAI_AMD_WITH_IO
                sa::load(restart_file, mesh, u);
#else
                throw std::runtime_error("--restart-file needs the restart I/O library (built with HDF5)");
#endif
            }
            sa::make_bc<sa::Dirichlet<1>>(u, 0.);

            double dt            = cfl * mesh.min_cell_length();
            const double dt_save = Tf / static_cast<double>(nfiles);
            auto unp1            = sa::make_scalar_field<double>("unp1", mesh);
            auto MRadaptation    = sa::make_MRAdapt(u);
            auto mra_config      = sa::mra_config().epsilon(2.e-4);
            MRadaptation(mra_config);
            save(path, filename, u, "_init");

            double t  = 0;
            int nsave = 1, nstep = 0;
            while (t != Tf)
            {
                MRadaptation(mra_config);
                t += dt;
                if (t > Tf)
                {
                    dt += Tf - t;
                    t = Tf;
                }
                sa::update_ghost_mr(u);
                unp1.resize();
                unp1 = u - dt * sa::upwind_scalar_burgers(k, u);
                sa::swap(u, unp1);
                ++nstep;
                if (t >= static_cast<double>(nsave) * dt_save || t == Tf)
                {
                    const std::string suffix = (nfiles != 1) ? "_ite_" + std::to_string(nsave++) : "";
                    save(path, filename, u, suffix);
                }
            }
            std::printf("steps %d cells %zu dt %.17g\n", nstep, mesh.nb_cells(), dt);
            if (!out.empty())
            {
                FILE* f = std::fopen((out + "_final.txt").c_str(), "w");
                if (!f)
                {
                    throw std::runtime_error("cannot write " + out + "_final.txt");
                }
                sa::for_each_cell(mesh,
                                  [&](const auto& cell)
                                  { std::fprintf(f, "%d %d %d %.17g\n", cell.level, cell.indices[0], cell.indices[1], u[cell]); });
                std::fclose(f);
            }
        }
        sa::finalize();
        return 0;
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
