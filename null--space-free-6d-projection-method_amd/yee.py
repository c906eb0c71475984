"""Physical-space closure of a reconstructed eigenfield on a Yee window (MI355X kernels).

Mirrors compute_yee_lwrq_metrics_exp4 (exp3/compute_yee_lwrq_metrics_exp4.m:1-100),
apply_cropped_A_exp3 (exp3/apply_cropped_A_exp3.m:1-56), sample_scalar_yee_material_exp4
(exp3/sample_scalar_yee_material_exp4.m:1-21), compute_lwrq_exp4 (exp3/compute_lwrq_exp4.m:1-57)
and crop_plus_field_exp4 (exp3/crop_plus_field_exp4.m:1-7).

Fields are flat complex128 device tensors in the column-major order of the reference
(x fastest): plus-grid fields have (2n+3)^3 entries (the output of
taylor.taylor_evaluate_yee / fame6d_taylor_evaluate_yee), interior arrays (2n+1)^3.
Every computation runs in libfame6d.so; nothing here falls back to the CPU.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr

# cfg.lwrq and the material / projection of exp3/make_config_experiment4.m:17-32,72-74
LWRQ_RADIUS = 3
LWRQ_SIGMA_FACTOR = 2.0
LWRQ_MASS_FLOOR = 1e-14
EPS_C = 4.0
ALPHA = 1.2
P_EXP4 = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0],
                   [np.sqrt(2.0), 0.0, 0.0], [0.0, np.sqrt(3.0), 0.0], [0.0, 0.0, np.sqrt(5.0)]])


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _field(u, count, dtype, name):
    if not isinstance(u, torch.Tensor) or not u.is_cuda:
        raise TypeError("%s must be a CUDA tensor" % name)
    u = u.reshape(-1)
    if u.dtype != dtype or u.numel() != count:
        raise ValueError("%s must hold %d %s values" % (name, count, dtype))
    return u.contiguous()


def workspace(n, device="cuda"):
    """Device workspace for the entry points below (reused across calls of the same n)."""
    nbytes = _lib.load().fame6d_yee_workspace_bytes(int(n))
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


def apply_cropped_A(u1, u2, u3, n, h):
    """Cropped Yee curl-curl of a plus-grid field; returns three (2n+1)^3 complex128 tensors."""
    m3, q3 = (2 * n + 3) ** 3, (2 * n + 1) ** 3
    u = [_field(x, m3, torch.complex128, "u%d" % (i + 1)) for i, x in enumerate((u1, u2, u3))]
    A = [torch.empty(q3, dtype=torch.complex128, device=u[0].device) for _ in range(3)]
    call("fame6d_yee_apply_cropped_A", int(n), float(h), ptr(u[0]), ptr(u[1]), ptr(u[2]), ptr(A[0]), ptr(A[1]),
         ptr(A[2]), _stream())
    return tuple(A)


def sample_scalar_yee_material(n, h, P=P_EXP4, eps_c=EPS_C, alpha=ALPHA, device="cuda", work=None):
    """eps on the three edge families of the interior; three (2n+1)^3 float64 tensors."""
    Pm = np.ascontiguousarray(np.asarray(P, dtype=np.float64))
    q3 = (2 * n + 1) ** 3
    e = [torch.empty(q3, dtype=torch.float64, device=device) for _ in range(3)]
    call("fame6d_yee_sample_eps", int(n), float(h), Pm.ctypes.data_as(C.c_void_p), int(Pm.shape[0]), float(eps_c),
         float(alpha), ptr(e[0]), ptr(e[1]), ptr(e[2]), ptr(work), _stream())
    return tuple(e)


def compute_lwrq(a, b, n, h, radius=LWRQ_RADIUS, sigma_factor=LWRQ_SIGMA_FACTOR, mass_floor_rel=LWRQ_MASS_FLOOR,
                 return_rho=False, work=None):
    """LWRQ scalars of densities a, b ((2n+1)^3 float64 device tensors) and optionally rho."""
    q3 = (2 * n + 1) ** 3
    a = _field(a, q3, torch.float64, "a")
    b = _field(b, q3, torch.float64, "b")
    rho = torch.empty(q3, dtype=torch.float64, device=a.device) if return_rho else None
    out = _lib.LwrqResult()
    call("fame6d_lwrq", int(n), float(h), ptr(a), ptr(b), int(radius), float(sigma_factor), float(mass_floor_rel),
         ptr(rho), ptr(work), C.byref(out), _stream())
    d = out.as_dict()
    d["rho"] = rho
    return d


def compute_yee_lwrq_metrics(u1, u2, u3, lambda6d, n, h, P=P_EXP4, eps_c=EPS_C, alpha=ALPHA, radius=LWRQ_RADIUS,
                             sigma_factor=LWRQ_SIGMA_FACTOR, mass_floor_rel=LWRQ_MASS_FLOOR, return_rho=False,
                             work=None):
    """The metrics struct of compute_yee_lwrq_metrics_exp4 for plus-grid fields u1..u3."""
    m3, q3 = (2 * n + 3) ** 3, (2 * n + 1) ** 3
    u = [_field(x, m3, torch.complex128, "u%d" % (i + 1)) for i, x in enumerate((u1, u2, u3))]
    Pm = np.ascontiguousarray(np.asarray(P, dtype=np.float64))
    rho = torch.empty(q3, dtype=torch.float64, device=u[0].device) if return_rho else None
    out = _lib.YeeMetrics()
    call("fame6d_yee_lwrq_metrics", int(n), float(h), ptr(u[0]), ptr(u[1]), ptr(u[2]), Pm.ctypes.data_as(C.c_void_p),
         int(Pm.shape[0]), float(eps_c), float(alpha), float(lambda6d), int(radius), float(sigma_factor),
         float(mass_floor_rel), ptr(rho), ptr(work), C.byref(out), _stream())
    d = out.as_dict()
    d.update(h=h, n=n, half_width=n * h, lambda6d=lambda6d, rho=rho)
    return d


def crop_plus_field(u, n_large, n_small):
    """Centred crop of a flat plus-grid field to a smaller window of the same h."""
    if n_small > n_large:
        raise ValueError("nSmall must not exceed nLarge")
    m = 2 * n_large + 3
    c = n_large + 1
    r = slice(c - (n_small + 1), c + n_small + 2)
    # column-major (x fastest) flat vector = row-major [z, y, x] array
    return u.reshape(m, m, m)[r, r, r].contiguous().reshape(-1)
