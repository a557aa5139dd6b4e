"""The library command-line options of the C++ drop-in header
(arguments.hpp:11-89): samurai_amd::args::parse fills them, mesh_config and
mra_config take them over their own values the way the reference's
parse_args does (mesh_config.hpp:356-397, mr/config.hpp:47-61), and
disable_args_parse() opts a configuration out.  Host only: nothing here
touches a device."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = r'''
#include "samurai_amd/samurai.hpp"
#include <cstdio>
int main(int argc, char** argv)
{
    samurai_amd::args::parse(argc, argv);
    auto c  = samurai_amd::mesh_config<2>().min_level(2).max_level(5).max_stencil_size(2).raw();
    auto c2 = samurai_amd::mesh_config<2>().min_level(2).max_level(5).disable_args_parse().raw();
    std::printf("%d %d %d %d | %d %d | %g %g %d %d\n", c.min_level, c.max_level, c.max_stencil_radius, c.graduation_width,
                c2.min_level, c2.max_level,
                samurai_amd::args::epsilon, samurai_amd::args::regularity, int(samurai_amd::args::refine_boundary),
                int(samurai_amd::args::rel_detail));
}
'''


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    src, exe = d / "cli.cpp", d / "cli"
    src.write_text(SRC)
    lib = os.path.join(ROOT, "samurai_amd")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe), f"-L{lib}", "-lsamurai_amd",
                    f"-Wl,-rpath,{lib}"], check=True)
    return str(exe)


def run(prog, *argv):
    return subprocess.run([prog, *argv], check=True, capture_output=True, text=True).stdout.split()


def test_defaults_untouched(prog):
    assert run(prog) == ["2", "5", "1", "1", "|", "2", "5", "|", "inf", "inf", "0", "0"]


def test_options_override_the_configuration(prog):
    out = run(prog, "--min-level", "3", "--max-level=7", "--graduation-width", "2", "--mr-eps", "2e-3", "--mr-reg=1.5",
              "--refine-boundary", "--mr-rel-detail", "--other", "x")
    assert out == ["3", "7", "1", "2", "|", "2", "5", "|", "0.002", "1.5", "1", "1"]
