"""Build the native parts in-tree.

  libppsk.so            HIP kernels + C ABI (include/ppsk.h), gfx950
  pp_sketchlib*.so      pybind11 module with the reference's Python API,
                        linked against libppsk.so and HDF5

Run `python build.py` (or call build_all()).  Every call recompiles.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
HDF5_PREFIX = os.environ.get("PPSK_HDF5_PREFIX", "/opt/conda")

LIB = os.path.join(HERE, "libppsk.so")

HIP_SOURCES = ["dist_kernels.hip", "select_kernels.hip", "sketch_kernels.hip", "ppsk_capi.cpp",
               "batch.cpp", "serve.cpp", "sketch_capi.cpp", "noalloc.cpp"]
HOST_SOURCES = ["synth.cpp"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # IEEE double for the Jaccard / random-match arithmetic: no FMA contraction
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd))


def _check_pipelined_spills(remarks):
    """Reject a build in which a pipelined tile configuration spills.

    The PB = 2 matching loop issues scalar loads from inline asm and waits for
    them later; if the compiler spilled or copied one of those SGPR tuples in
    between, it would move data that has not arrived yet.  The kernel
    resource remarks list spills per function: every frame that holds the
    pipelined matching must report none -- the dist_tiles kernels (matching
    inlined) and batch_tile_k (the persistent kernels' out-of-line matching,
    the same code as the dist_tiles kernel of its shape).  The persistent
    kernels' own frames (dist_batch, dist_batch_mixed) only take work items
    and call out: no asm operand is in flight there, so a spill in them is
    only a cost."""
    name, bad = None, []
    for line in remarks.splitlines():
        if "Function Name:" in line:
            name = line.split("Function Name:")[1].split()[0]
        elif name and ("SGPRs Spill:" in line or "VGPRs Spill:" in line):
            pipelined = (("dist_tiles" in name or "batch_tile_k" in name) and "ELi2EE" in name
                         or "dist_tiles_diagonal" in name)
            if pipelined and int(line.rsplit(":", 1)[1].split()[0]) != 0:
                bad.append(name + ": " + line.split("remark:")[-1].strip())
    if bad:
        raise RuntimeError("pipelined tile kernels spill registers:\n  " + "\n  ".join(bad))


def build_lib():
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(HERE, "_build", os.path.splitext(src)[0] + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = [HIPCC, "-x", "hip"] + HIP_FLAGS + ["-I", INCLUDE, "-c",
                                                  os.path.join(CSRC, src), "-o", obj]
        if src == "dist_kernels.hip":
            cmd.append("-Rpass-analysis=kernel-resource-usage")
            print("+", " ".join(cmd), flush=True)
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                print(r.stdout)
                raise RuntimeError("build step failed: " + " ".join(cmd))
            _check_pipelined_spills(r.stdout)
        else:
            _run(cmd)
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(HERE, "_build", os.path.splitext(src)[0] + ".o")
        _run(["g++", "-O3", "-std=c++17", "-fPIC", "-pthread", "-I", INCLUDE, "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


BENCH_LIB = os.path.join(HERE, "libppsk_bench.so")


def build_bench_lib():
    """Measurement kernels (VALU peak, copy calibration) used by bench.py."""
    tmp = BENCH_LIB + ".tmp"
    _run([HIPCC, "-x", "hip"] + HIP_FLAGS + ["-shared", os.path.join(CSRC, "peak_kernels.hip"),
                                             "-o", tmp])
    os.replace(tmp, BENCH_LIB)
    return BENCH_LIB


def module_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "pp_sketchlib" + suffix)


def _vendor_hdf5():
    """Copy libhdf5 and its zlib next to the module (RUNPATH $ORIGIN/_hdf5).

    Pointing the module's RUNPATH at the HDF5 installation prefix would also
    resolve libstdc++ from there, shadowing the one the HIP runtime needs."""
    import shutil
    dst = os.path.join(HERE, "_hdf5")
    os.makedirs(dst, exist_ok=True)
    for name in ("libhdf5.so.103", "libz.so.1"):
        src = os.path.realpath(os.path.join(HDF5_PREFIX, "lib", name))
        shutil.copy2(src, os.path.join(dst, name))
    return dst


def build_module():
    import pybind11

    out = module_path()
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    srcs = [os.path.join(CSRC, "host", f) for f in
            ("sketch_db.cpp", "random_match.cpp", "random_build.cpp", "api.cpp", "sketch_build.cpp",
             "bindings.cpp")]
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
           "-ffp-contract=off", "-fvisibility=hidden",
           "-I", INCLUDE, "-I", CSRC, "-I", pybind11.get_include(), "-I", py_inc,
           "-I", os.path.join(HDF5_PREFIX, "include"),
           "-DPPSK_VERSION_INFO=\"2.1.5\""] + srcs + [
           "-L", HERE, "-lppsk",
           os.path.join(_vendor_hdf5(), "libhdf5.so.103"), "-lz",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/_hdf5",
           "-o", tmp]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_all(module=True):
    build_lib()
    build_bench_lib()
    if module and os.path.isdir(os.path.join(CSRC, "host")):
        build_module()


if __name__ == "__main__":
    build_all(module="--lib-only" not in sys.argv)
