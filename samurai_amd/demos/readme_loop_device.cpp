// The README's example (README.md:86-154) on the C++ drop-in layer, with its
// hand-written upwind loop run on the GPU by samurai::for_each_interval and a
// device lambda: levels 2..8 (or --min-level/--max-level), u = 1 in
// [0.4, 0.6]^2, Dirichlet 0, MRadaptation(1e-4, 2), update_ghost_mr, the loop,
// swap; --steps n (default 50).  --out file writes the final leaves and values
// as text (level i j value).  --probe-missing runs one loop that reads
// u(level, i - 3, j), which is absent from the mesh for some leaves: the
// std::out_of_range it raises is printed and the program exits with status 3.
#define SAMURAI_AMD_AS_SAMURAI
#include <samurai_amd/device.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

int main(int argc, char** argv)
{
    int min_level = 2, max_level = 8, steps = 50;
    bool probe    = false;
    std::string out;
    for (int k = 1; k < argc; ++k)
    {
        const std::string a = argv[k];
        if (a == "--min-level" && k + 1 < argc)
        {
            min_level = std::atoi(argv[++k]);
        }
        else if (a == "--max-level" && k + 1 < argc)
        {
            max_level = std::atoi(argv[++k]);
        }
        else if (a == "--steps" && k + 1 < argc)
        {
            steps = std::atoi(argv[++k]);
        }
        else if (a == "--out" && k + 1 < argc)
        {
            out = argv[++k];
        }
        else if (a == "--probe-missing")
        {
            probe = true;
        }
    }
    try
    {
        samurai::initialize(argc, argv);
        {
            constexpr std::size_t dim = 2;
            const samurai::Box<double, dim> box({0., 0.}, {1., 1.});
            auto mesh = samurai::mra::make_mesh(box, samurai::mesh_config<dim>().min_level(min_level).max_level(max_level));
            auto u    = samurai::make_field<double, 1>("u", mesh);
            samurai::make_bc<samurai::Dirichlet<1>>(u, 0.);
            samurai::for_each_cell(mesh,
                                   [&](const auto& cell)
                                   {
                                       const double length = 0.2;
                                       const auto c        = cell.center();
                                       u[cell] = (std::abs(c[0] - 0.5) <= 0.5 * length && std::abs(c[1] - 0.5) <= 0.5 * length) ? 1.0 : 0.0;
                                   });
            auto MRadaptation = samurai::make_MRAdapt(u);
            const double dx   = mesh.cell_length(max_level);
            const double dt   = 0.5 * dx;
            auto unp1         = samurai::make_field<double, 1>("u", mesh);
            for (int nite = 0; nite < steps; ++nite)
            {
                MRadaptation(1e-4, 2);
                samurai::update_ghost_mr(u);
                unp1.resize();
                auto uv = samurai::device::view(u);
                auto vv = samurai::device::view(unp1);
                auto mv = samurai::device::view(mesh);
                samurai::for_each_interval(mesh,
                                           [=] __device__(int level, const auto& i, const auto& index)
                                           {
                                               double dx = mv.cell_length(level);
                                               auto j    = index[0];
                                               vv(level, i, j) = uv(level, i, j) - dt / dx * (uv(level, i, j) - uv(level, i - 1, j)
                                                                                              + uv(level, i, j) - uv(level, i, j - 1));
                                           });
                samurai::swap(u, unp1);
            }
            if (probe)
            {
                auto uv = samurai::device::view(u);
                auto vv = samurai::device::view(unp1);
                samurai::for_each_interval(mesh,
                                           [=] __device__(int level, const auto& i, const auto& index)
                                           {
                                               vv(level, i, index[0]) = uv(level, i - 3, index[0]);
                                           });
                std::printf("no exception\n");
                return 1;
            }
            if (!out.empty())
            {
                FILE* f = std::fopen(out.c_str(), "w");
                if (!f)
                {
                    throw std::runtime_error("cannot write " + out);
                }
                samurai::for_each_cell(mesh,
                                       [&](const auto& cell)
                                       {
                                           std::fprintf(f, "%d %d %d %.17g\n", cell.level, cell.indices[0], cell.indices[1], u[cell]);
                                       });
                std::fclose(f);
            }
            std::printf("%zu leaves after %d steps\n", mesh.nb_cells(), steps);
        }
        samurai::finalize();
    }
    catch (const std::out_of_range& e)
    {
        std::printf("out_of_range: %s\n", e.what());
        samurai::finalize();
        return 3;
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        samurai::finalize();
        return 2;
    }
    return 0;
}
