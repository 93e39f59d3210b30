"""Builds libmdx.so (HIP, gfx950 only) in-tree with hipcc.

    python -m lammps_amd.build        # or lammps_amd.build.build()

Every HIP translation unit is compiled with packed-FP32 instruction
selection disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
Objects go to lammps_amd/_build/, the library to lammps_amd/libmdx.so.

libmdx.so needs the HIP runtime by its soname (libamdhip64.so.7).  PyTorch's
ROCm wheel carries its own copy under that soname; when torch is imported
first, the dynamic loader binds libmdx.so to torch's copy, so the process
holds one HIP runtime (two copies in one process do not both see the GPU).
The Python layer imports torch before loading the library for that reason.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmdx.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

SOURCES = ["api.cpp", "host_setup.cpp", "neigh.hip", "pair.hip", "integrate.hip", "batch.hip", "halo.hip",
           "noalloc.hip"]
ARCH = "gfx950"


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libmdx)")


def common_flags():
    return ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-I" + INCLUDE, "-I" + CSRC,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-Wno-unused-result"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + common_flags() + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    # the LAMMPS-style script runner: C++ plugin classes over the C ABI
    exe = os.path.join(HERE, "lmp_mdx")
    cmd = ["g++", "-O2", "-std=c++17", "-I" + INCLUDE, os.path.join(HERE, "host", "plugin.cpp"),
           os.path.join(HERE, "host", "output.cpp"), os.path.join(HERE, "host", "lmp_mdx.cpp"), "-o", exe + ".tmp", "-L" + HERE, "-l:libmdx.so",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(exe + ".tmp", exe)
    ext = build_torch_ext()
    if verbose:
        print("built", LIB, exe, "and", ext)
    return LIB


TORCH_EXT = os.path.join(HERE, "_mdx_torch.so")


def build_torch_ext():
    """The autograd Functions over the batch ABI (csrc/torch_ext.cpp) as an
    in-tree Python extension module linked against libmdx.so and libtorch."""
    import sysconfig

    import torch
    import torch.utils.cpp_extension as ce

    libdir = ce.library_paths()[0]
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-std=c++17",
           "-D_GLIBCXX_USE_CXX11_ABI=%d" % int(torch._C._GLIBCXX_USE_CXX11_ABI),
           "-DTORCH_EXTENSION_NAME=_mdx_torch", "-DTORCH_API_INCLUDE_EXTENSION_H", "-Wno-deprecated-declarations"]
    cmd += ["-I" + p for p in ce.include_paths() + [sysconfig.get_paths()["include"], INCLUDE]]
    cmd += [os.path.join(CSRC, "torch_ext.cpp"), "-o", TORCH_EXT + ".tmp", "-L" + libdir, "-lc10", "-ltorch",
            "-ltorch_cpu", "-ltorch_python", "-L" + HERE, "-l:libmdx.so", "-Wl,-rpath," + libdir,
            "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("torch extension build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(TORCH_EXT + ".tmp", TORCH_EXT)
    return TORCH_EXT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
