"""The C-ABI library loads and exports every entry point include/fame6d.h declares (no GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fame6d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fame6d_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "fame6d_mass_apply_scalar" in names and "fame6d_eigs_method_c" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header(pkg):
    from importlib import import_module
    sigs = import_module("fame6d_amd._lib").SIGNATURES
    missing = [n for n in declared_functions() if n not in sigs and n != "fame6d_last_error"]
    assert not missing, missing


def test_errors_without_gpu_are_loud(pkg):
    import torch
    if torch.cuda.is_available():
        return
    import numpy as np
    try:
        pkg.MassBackend(1, np.ones(64))
    except RuntimeError as e:
        assert "GPU" in str(e)
    else:
        raise AssertionError("MassBackend must refuse to run without a GPU")


def test_mex_gateway_compiles_against_header():
    """The MATLAB MEX gateway (mex/fame6d_mex.c) type-checks against include/fame6d.h: a C
    compiler in syntax-only mode, with a stub of the MEX API calls the gateway makes
    (tests/mex_stub/mex.h) since MATLAB is not installed.  A change of an entry point's
    signature that the gateway does not follow fails here."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    assert cc, "no C compiler"
    src = os.path.join(ROOT, "null--space-free-6d-projection-method_amd", "mex", "fame6d_mex.c")
    cmd = [cc, "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-std=c11", "-D__HIP_PLATFORM_AMD__",
           "-I" + os.path.join(ROOT, "tests", "mex_stub"), "-I" + os.path.join(ROOT, "include"),
           "-I/opt/rocm/include", src]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-4000:]
