"""Builds libsamurai_amd.so in-tree: host C++ (interval algebra, mesh, work
lists, runtime, C ABI) with g++, the kernels with hipcc for gfx950.

All code is compiled with -ffp-contract=off so that no a*b+c is fused: flags
and fields are compared bit for bit with the reference's operation order.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsamurai_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HOST_SOURCES = ["host/interval_set.cpp", "host/mr_mesh.cpp", "host/programs.cpp", "host/convection.cpp", "runtime.cpp", "batch.cpp", "adjoint.cpp", "capi.cpp"]
HIP_SOURCES = ["device/kernels.hip", "device/program.hip", "device/noalloc.hip", "device/rebuild.hip", "device/rebuild_batch.hip", "device/submesh.hip",
               "device/submesh_batch.hip", "device/lists.hip"]

CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra",
            "-Wno-unused-parameter", f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "--offload-arch=gfx950",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def _obj(src):
    return os.path.join(BUILD, src.replace("/", "_").rsplit(".", 1)[0] + ".o")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HOST_SOURCES:
        jobs.append(["g++", *CXXFLAGS, "-c", os.path.join(CSRC, src), "-o", _obj(src)])
    for src in HIP_SOURCES:
        jobs.append([os.path.join(ROCM, "bin", "hipcc"), *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", _obj(src)])
    procs = [(cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for cmd in jobs]
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 target feature: harmless
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and "is not a recognized feature" not in ln]
        if p.returncode != 0:
            failed = True
            sys.stderr.write(" ".join(cmd) + "\n" + "\n".join(lines) + "\n")
        elif verbose and lines:
            sys.stderr.write("\n".join(lines) + "\n")
    if failed:
        raise RuntimeError("samurai_amd: compilation failed")
    objs = [_obj(s) for s in HOST_SOURCES + HIP_SOURCES]
    link = [os.path.join(ROCM, "bin", "hipcc"), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", LIB + ".tmp", *objs,
            f"-L{ROCM}/lib", "-lamdhip64"]
    r = subprocess.run(link, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("samurai_amd: link failed")
    os.replace(LIB + ".tmp", LIB)
    io = build_io()
    # the C++ drop-in demos, built on include/samurai_amd/samurai.hpp (and
    # io.hpp for their output and restart files when the I/O library is there)
    for name in ("advection_2d", "scalar_burgers_2d"):
        demo = os.path.join(HERE, "demos", name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}", os.path.join(HERE, "demos", name + ".cpp"),
               "-o", demo + ".tmp", f"-L{HERE}"]
        if io:
            cmd += ["-DSAMURAI_AMD_WITH_IO", "-lsamurai_amd_io", f"-Wl,-rpath-link,{os.path.join(HDF5_ROOT, 'lib')}"]
        cmd += ["-lsamurai_amd", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            sys.stderr.write(r.stdout.decode(errors="replace"))
            raise RuntimeError(f"samurai_amd: demo {name} build failed")
        os.replace(demo + ".tmp", demo)
    # the README example with its loop as a device lambda
    # (include/samurai_amd/device.hpp), compiled by hipcc
    demo = os.path.join(HERE, "demos", "readme_loop_device")
    cmd = [os.path.join(ROCM, "bin", "hipcc"), "-x", "hip", *[f for f in HIPFLAGS if f != "-fPIC"], f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(HERE, "demos", "readme_loop_device.cpp"), "-o", demo + ".tmp", f"-L{HERE}", "-lsamurai_amd",
           "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("samurai_amd: demo readme_loop_device build failed")
    os.replace(demo + ".tmp", demo)
    return LIB


HDF5_ROOT = os.environ.get("SMR_HDF5_ROOT", "/opt/conda")
IO_LIB = os.path.join(HERE, "libsamurai_amd_io.so")


def build_io():
    """libsamurai_amd_io.so: restart files and HDF5/XDMF output over the C ABI; skipped when
    the HDF5 headers and library are absent (the compute library does not
    need it)."""
    inc, lib = os.path.join(HDF5_ROOT, "include"), os.path.join(HDF5_ROOT, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(os.path.join(lib, "libhdf5.so"))):
        sys.stderr.write(f"samurai_amd: HDF5 not found under {HDF5_ROOT}; restart I/O not built\n")
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-unused-parameter",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{inc}", os.path.join(CSRC, "io", "restart.cpp"), os.path.join(CSRC, "io", "xdmf.cpp"),
           "-o", IO_LIB + ".tmp",
           f"-L{HERE}", "-lsamurai_amd", f"-L{lib}", "-lhdf5",
           # the HDF5 directory holds an older libstdc++: this library carries its
           # own (static) and finds HDF5 through a RUNPATH, which serves its own
           # dependencies only, so libsamurai_amd and HIP keep the system one
           "-static-libstdc++", "-static-libgcc", "-Wl,--enable-new-dtags", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("samurai_amd: restart I/O build failed")
    os.replace(IO_LIB + ".tmp", IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    print(build(verbose=True))
