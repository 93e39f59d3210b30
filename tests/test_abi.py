"""CPU checks of the C ABI: libmdx.so loads and exports every entry point
that include/*.h declares (no device calls are made here)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mdx_\w+)\s*\(", text, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    from lammps_amd import mdx
    lib = ctypes.CDLL(mdx.LIB_PATH)
    names = declared_symbols()
    assert "mdx_compute" in names and "mdx_run" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(mdx.exported_symbols()) <= names


def test_one_hip_runtime_in_a_process_with_torch():
    """torch's wheel carries a HIP runtime under libmdx.so's soname: loaded
    after torch, the library binds to that copy (one runtime mapped)."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from lammps_amd import mdx\n"
            "mdx.lib()\n"
            "import torch\n"
            "print(mdx.LOADED_FLAVOUR, len(mdx.hip_runtime_paths()), mdx.hip_runtime_paths()[0])\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    flavour, n, path = out.stdout.split()
    assert flavour == "torch" and n == "1" and "torch" in path, out.stdout
    env = dict(os.environ, MDX_HIP_RUNTIME="rocm")
    code2 = ("import sys; sys.path.insert(0, %r)\n"
             "from lammps_amd import mdx\n"
             "mdx.lib()\n"
             "print(mdx.LOADED_FLAVOUR, 'torch' in sys.modules, len(mdx.hip_runtime_paths()))\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code2], capture_output=True, text=True, timeout=300, env=env)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["rocm", "False", "1"], out.stdout


def test_version_and_error_string_without_gpu():
    from lammps_amd import mdx
    L = mdx.lib()
    assert L.mdx_version() == 1
    assert isinstance(L.mdx_last_error(), bytes)


def test_batch_header_is_covered_and_create_fails_cleanly_without_gpu():
    names = declared_symbols()
    for n in ("mdx_batch_create", "mdx_batch_lj_vjp", "mdx_batch_eam_vjp", "mdx_batch_autotune"):
        assert n in names
    from lammps_amd import batch
    from lammps_amd.mdx import ERR_HIP, MdxError
    import numpy as np
    try:
        b = batch.Batch([2], [[0, 0, 0]], [[5, 5, 5]], [[1, 1, 1]], 1, np.ones(2, np.int32))
    except MdxError as e:
        # without a HIP device: a clean error, not a crash
        assert e.code == ERR_HIP
    else:
        b.close()


def test_torch_extension_loads_and_exposes_the_autograd_functions():
    from lammps_amd import fit
    m = fit.ext()
    assert m.version() == 1
    assert hasattr(m, "lj_forces") and hasattr(m, "eam_forces")
