// The 2D advection demo on the C++ drop-in layer: a disc of radius 0.2 at
// (0.3, 0.3) advected with velocity (1, 1) on an adaptive MR mesh, Dirichlet
// 0, first-order upwind, until Tf with the last step shortened to land on Tf
// (demos/FiniteVolume/advection_2d.cpp:22-153; same defaults: levels 4..10,
// eps 2e-4, regularity 1, cfl 0.5, Tf 0.01).
//
//   advection_2d [--min-level L] [--max-level L] [--mr-eps e] [--mr-reg r]
//                [--Tf t] [--cfl c] [--prediction-order 0|1] [--out prefix]
//
// With --out, the cells (level, i, j) and values after the initial
// adaptation and at Tf are written to <prefix>_init.txt and <prefix>_final.txt.
#include <samurai_amd/samurai.hpp>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace sa = samurai_amd;

template <class Field>
static void dump(Field& u, const std::string& path)
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

template <int S>
static int run(int min_level, int max_level, double eps, double reg, double Tf, double cfl, const std::string& out)
{
    constexpr std::size_t dim = 2;
    const sa::Box<double, dim> box({0., 0.}, {1., 1.});
    auto config = sa::mesh_config<dim, S>().min_level(min_level).max_level(max_level).max_stencil_size(2).disable_minimal_ghost_width();
    auto mesh   = sa::mra::make_mesh(box, config);

    auto u = sa::make_scalar_field<double>("u", mesh);
    sa::for_each_cell(mesh,
                      [&](const auto& cell)
                      {
                          const auto c   = cell.center();
                          const double r = 0.2, xc = 0.3, yc = 0.3;
                          u[cell]        = ((c[0] - xc) * (c[0] - xc) + (c[1] - yc) * (c[1] - yc) <= r * r) ? 1.0 : 0.0;
                      });
    sa::make_bc<sa::Dirichlet<1>>(u, 0.);
    auto unp1 = sa::make_scalar_field<double>("unp1", mesh);

    const std::array<double, dim> a{1., 1.};
    double dt = cfl * mesh.min_cell_length();

    auto MRadaptation = sa::make_MRAdapt(u);
    auto mra_config   = sa::mra_config().epsilon(eps).regularity(reg);
    MRadaptation(mra_config);
    if (!out.empty())
    {
        dump(u, out + "_init.txt");
    }

    double t  = 0;
    int nstep = 0;
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
        unp1 = u - dt * sa::upwind(a, u);
        sa::swap(u, unp1);
        ++nstep;
    }
    std::printf("steps %d cells %zu\n", nstep, mesh.nb_cells());
    if (!out.empty())
    {
        dump(u, out + "_final.txt");
    }
    return 0;
}

int main(int argc, char** argv)
{
    int min_level = 4, max_level = 10, order = 1;
    double eps = 2e-4, reg = 1., Tf = 0.01, cfl = 0.5;
    std::string out;
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
        else if (k == "--mr-eps")
        {
            eps = std::atof(v);
        }
        else if (k == "--mr-reg")
        {
            reg = std::atof(v);
        }
        else if (k == "--Tf")
        {
            Tf = std::atof(v);
        }
        else if (k == "--cfl")
        {
            cfl = std::atof(v);
        }
        else if (k == "--prediction-order")
        {
            order = std::atoi(v);
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
        const int rc = order == 0 ? run<0>(min_level, max_level, eps, reg, Tf, cfl, out)
                                  : run<1>(min_level, max_level, eps, reg, Tf, cfl, out);
        sa::finalize();
        return rc;
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
