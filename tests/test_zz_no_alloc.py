"""The allocation-free scalar mass apply (fame6d_mass_apply_scalar_noalloc) against the
allocating entry point fame6d_mass_apply_scalar, called through the C ABI."""
import ctypes as C
from importlib import import_module

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _lib():
    return import_module("fame6d_amd._lib")


def _noalloc(mass, prec, x, y, nrhs):
    lib = _lib()
    return lib.call("fame6d_mass_apply_scalar_noalloc", mass.handle, prec, lib.ptr(x), lib.ptr(y), nrhs)


def _alloc(mass, prec, x, y, nrhs):
    lib = _lib()
    return lib.call("fame6d_mass_apply_scalar", mass.handle, prec, lib.ptr(x), lib.ptr(y), nrhs)


def _rand(rng, nrhs, NF, dtype):
    return torch.as_tensor(rng.standard_normal((nrhs, NF)) + 1j * rng.standard_normal((nrhs, NF)),
                           device="cuda").to(dtype)


@pytest.mark.parametrize("prec,dtype", [(0, torch.complex128), (1, torch.complex64)])
def test_noalloc_matches_allocating_apply(pkg, oracle, prec, dtype):
    """Two calls in a row, the second with a smaller input (N = 5 with 3 columns, then N = 3
    with 1 column on another handle that shares the array): each equals the allocating entry
    point bit for bit (same kernels, same tiles) and the oracle to the FP64 / FP32 bound."""
    torch.cuda.set_device(0)
    rng = np.random.default_rng(70 + prec)
    tol = 1e-13 if prec == 0 else 2e-6
    for N, nrhs in ((5, 3), (3, 1)):
        print("case: noalloc apply N=%d nrhs=%d precision %d" % (N, nrhs, prec), flush=True)
        eps = oracle.eps_reference(N)
        mass = pkg.MassBackend(N, eps)
        x = _rand(rng, nrhs, mass.NF, dtype)
        y0 = torch.empty_like(x)
        y1 = torch.full_like(x, float("nan"))
        _alloc(mass, prec, x, y0, nrhs)
        _noalloc(mass, prec, x, y1, nrhs)
        torch.cuda.synchronize()
        assert torch.equal(y0, y1)
        ref = oracle.mass_apply(eps, x.cpu().numpy().astype(np.complex128).T, N).T
        err = np.linalg.norm(y1.cpu().numpy() - ref) / np.linalg.norm(ref)
        assert err < tol, err
        mass.close()


def test_noalloc_allocates_nothing_and_checks_limit(pkg, oracle):
    """On a fresh handle (no workspace yet) the call leaves the device's free memory as it was,
    y may alias x, and a request beyond the fixed 64 MiB array is refused with FAME6D_ESIZE
    before anything runs."""
    torch.cuda.set_device(0)
    N = 5
    eps = oracle.eps_reference(N)
    mass = pkg.MassBackend(N, eps)
    rng = np.random.default_rng(80)
    x = _rand(rng, 4, mass.NF, torch.complex128)  # 4 x 16 MB = the whole array
    ref = torch.empty_like(x)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    _noalloc(mass, 0, x, ref, 4)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    print("case: noalloc free memory before/after %d %d" % (free0, free1), flush=True)
    assert free1 == free0
    xa = x.clone()
    _noalloc(mass, 0, xa, xa, 4)  # in place
    torch.cuda.synchronize()
    assert torch.equal(xa, ref)
    big = torch.empty((5, mass.NF), dtype=torch.complex128, device="cuda")
    with pytest.raises(pkg.Fame6dError) as e:
        _noalloc(mass, 0, big, big, 5)
    assert e.value.name == "ESIZE"
    mass.close()
