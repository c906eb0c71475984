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
