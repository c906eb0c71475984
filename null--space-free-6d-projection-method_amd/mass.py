"""Scalar-permittivity Fourier mass backend on the MI355X.

Mirrors build_mass_backend_large_scalar (exp1/build_mass_backend_large_scalar.m:1-49):
the returned object exposes the handles of the reference's `mass` struct,
MscalarOp / Mop / Mcol3op / Mcols12op, plus `info`.  The operator is
M_s x = fftn(eps .* ifftn(x)) on the (2N)^6 torus (exp1/applyMij_fft.m:11), the
same matrix as the reference's exact rank-1 lattice backend
(exp1/applyMij_rank1_exact.m:4-5).

Vector layout: a grid vector is a 1-D complex tensor of NF = (2N)^6 entries in the
reference's column-major order; r right-hand sides are a (r, NF) tensor (row j is
MATLAB column j).  3-component vectors are (3NF,) or (r, 3NF), component-first.
"""
import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from ._lib import FP32, FP64, MIXED, call, ptr

# Projection of every reference configuration (exp1/make_config_exp1bc.m:20-26).
P_DEFAULT = np.array([[1.0, 0.0, 0.0],
                      [0.0, 1.0, 0.0],
                      [0.0, 0.0, 1.0],
                      [math.sqrt(2.0), 0.0, 0.0],
                      [0.0, math.sqrt(3.0), 0.0],
                      [0.0, 0.0, math.sqrt(5.0)]])


def precision_code(precision):
    if isinstance(precision, int):
        return precision
    return {"double": FP64, "fp64": FP64, "single": FP32, "fp32": FP32, "mixed": MIXED}[precision.lower()]


def vector_dtype(precision):
    return torch.complex64 if precision_code(precision) == FP32 else torch.complex128


def grid_coords(N):
    """x_k = 2*pi*j/nGrid (exp1/build_experiment1_common.m:10-11)."""
    n = 2 * N
    return np.arange(n) / n * 2 * np.pi


def _grid_sum(N, f):
    n = 2 * N
    c = grid_coords(N)
    out = np.zeros((n,) * 6)
    for k in range(6):
        shape = [1] * 6
        shape[k] = n
        out = out + f(k, c).reshape(shape)
    return out.reshape(-1, order="F")


def smooth_scalar_material(N, epsC=4.0, alpha=1.2):
    """eps = epsC*exp((alpha/6)*sum_k cos x_k) (exp1/build_smooth_scalar_material_compact.m:3-7)."""
    return epsC * np.exp((alpha / 6.0) * _grid_sum(N, lambda k, c: np.cos(c)))


def cosine_material(N, amp, seed):
    """eps = 1 + (amp/6)*sum_k cos(2*pi*y_k + phi_k), phases phi_k ~ U[0, 2*pi) from `seed`.
    The benchmark's cosine family with the amplitude split over the six terms, which keeps
    eps in [1-amp, 1+amp] (positive) so the mass operator stays positive definite."""
    phases = np.random.default_rng(seed).uniform(0.0, 2.0 * np.pi, 6)
    eps = 1.0 + (amp / 6.0) * _grid_sum(N, lambda k, c: np.cos(c + phases[k]))
    return eps, phases


def start_vector(n, seed):
    """Seeded complex Gaussian Lanczos start vector (the role of
    rng(seed,'twister'); randn(n,1)+1i*randn(n,1) in exp1/run_method_C_fig1bc.m:5-6)."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def current_stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class MassBackend:
    """Handle of one scalar mass operator (one permittivity, one grid size, one stream)."""

    def __init__(self, N, eps, device=None, autotune=False):
        if not torch.cuda.is_available():
            raise RuntimeError("the MI355X mass backend needs a GPU (torch.cuda.is_available() is False)")
        _lib.load()
        self.N = int(N)
        self.nGrid = 2 * self.N
        self.NF = self.nGrid ** 6
        eps = np.ascontiguousarray(np.asarray(eps, dtype=np.float64).reshape(-1))
        if eps.size != self.NF:
            raise ValueError("eps must have NF = (2N)^6 = %d samples" % self.NF)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self._eps = eps
        self._eps_dev = None
        self._h = C.c_void_p()
        with torch.cuda.device(self.device):
            self.stream = torch.cuda.current_stream(self.device)
            call("fame6d_mass_create", self.N, eps.ctypes.data_as(C.c_void_p),
                 C.c_void_p(self.stream.cuda_stream), C.byref(self._h))
        if autotune:
            self.autotune()
        self.name = "mi355x-circulant-scalar"

    # -- lifecycle -------------------------------------------------------------------
    def close(self):
        if self._h:
            call("fame6d_mass_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_eps(self, eps):
        """Replaces the permittivity samples (same grid).  A tensor on the backend's device is
        validated and scaled on the device (fame6d_mass_set_eps_device); anything else is
        taken as host samples."""
        if isinstance(eps, torch.Tensor) and eps.device == self.device:
            e = eps.detach().reshape(-1).to(torch.float64).contiguous()
            if e.numel() != self.NF:
                raise ValueError("eps must have NF samples")
            e = e.clone() if e.data_ptr() == eps.data_ptr() else e  # the backend keeps its own copy
            call("fame6d_mass_set_eps_device", self._h, ptr(e))
            self._eps_dev = e
            self._eps = None
            return
        if isinstance(eps, torch.Tensor):
            eps = eps.detach().cpu().numpy()
        eps = np.ascontiguousarray(np.asarray(eps, dtype=np.float64).reshape(-1))
        if eps.size != self.NF:
            raise ValueError("eps must have NF samples")
        call("fame6d_mass_set_eps", self._h, eps.ctypes.data_as(C.c_void_p))
        self._eps = eps
        self._eps_dev = None

    def eps_tensor(self):
        """The current permittivity samples as a float64 tensor on the backend's device."""
        if self._eps_dev is None:
            self._eps_dev = torch.as_tensor(self._eps, device=self.device)
        return self._eps_dev

    def using_eps(self, eps):
        """Context manager: the backend computes with `eps` inside the block and gets its
        previous permittivity back on exit, so callers that share one backend see no side
        effect."""
        return _EpsScope(self, eps)

    @property
    def handle(self):
        return self._h

    @property
    def eps(self):
        if self._eps is None:
            self._eps = self._eps_dev.cpu().numpy()
        return self._eps

    @property
    def info(self):
        i = _lib.MassInfo()
        call("fame6d_mass_info", self._h, C.byref(i))
        return dict(name=self.name, N=i.N, nGrid=i.nGrid, NF=i.NF, M=i.NF, MOverNF=1.0,
                    epsMin=i.eps_min, epsMax=i.eps_max, epsMean=i.eps_mean,
                    tiles=dict(a64=i.tile_a64, b64=i.tile_b64, a32=i.tile_a32, b32=i.tile_b32),
                    variants=dict(double=(i.variant_ac64, i.variant_b64), single=(i.variant_ac32, i.variant_b32)))

    def autotune(self, reps=5):
        """Project-written tile autotuner for both precisions."""
        call("fame6d_mass_autotune", self._h, FP64, reps)
        call("fame6d_mass_autotune", self._h, FP32, reps)
        return self.info["tiles"]

    def set_tiles(self, precision, va, vb):
        call("fame6d_mass_set_tiles", self._h, precision_code(precision), va, vb)

    def profile(self, stride=1):
        """Time every `stride`-th launch of the timed kernel families with HIP events on
        the library stream (stride 0 switches timing off); resets the sums."""
        call("fame6d_mass_profile", self._h, int(stride))

    def kernel_times(self):
        """{family: {launches, ms, bytes, avg_us, GBs}} accumulated since profile()."""
        from ._lib import KernelTime, NUM_TIMER_SLOTS, TIMER_SLOTS
        arr = (KernelTime * NUM_TIMER_SLOTS)()
        call("fame6d_mass_kernel_times", self._h, arr)
        out = {}
        for name, t in zip(TIMER_SLOTS, arr):
            d = dict(launches=int(t.launches), ms=float(t.milliseconds), bytes=float(t.bytes))
            d["avg_us"] = 1e3 * d["ms"] / d["launches"] if d["launches"] else None
            d["GBs"] = d["bytes"] / (d["ms"] * 1e-3) / 1e9 if d["ms"] > 0 else None
            out[name] = d
        return out

    # -- applies ---------------------------------------------------------------------
    def _prep(self, x, rows):
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.asarray(x))
        if x.dtype not in (torch.complex128, torch.complex64):
            x = x.to(torch.complex128)
        x = x.to(self.device)
        if not x.is_contiguous():
            x = x.contiguous()
        nrhs = 1 if x.dim() == 1 else x.shape[0]
        if x.shape[-1] != rows:
            raise _lib.Fame6dError(1, "x must have %d rows (BadSize)" % rows)
        prec = FP32 if x.dtype == torch.complex64 else FP64
        return x, nrhs, prec

    def MscalarOp(self, x):
        """y = M_s x (applyMij_rank1_exact_lowmem / applyMij_fft)."""
        x, nrhs, prec = self._prep(x, self.NF)
        y = torch.empty_like(x)
        call("fame6d_mass_apply_scalar", self._h, prec, ptr(x), ptr(y), nrhs)
        return y

    def Mop(self, U):
        """MU = (I3 (x) M_s) U on 3NF vectors (applyM_component_first_rank1_scalar)."""
        U, nrhs, prec = self._prep(U, 3 * self.NF)
        MU = torch.empty_like(U)
        call("fame6d_mass_apply_3comp", self._h, prec, ptr(U), ptr(MU), nrhs)
        return MU

    def Mcols12op(self, x12):
        """2NF -> 3NF with block 3 = 0 (applyM_cols12_rank1_scalar)."""
        x12, nrhs, prec = self._prep(x12, 2 * self.NF)
        shape = (3 * self.NF,) if x12.dim() == 1 else (nrhs, 3 * self.NF)
        Y = torch.empty(shape, dtype=x12.dtype, device=self.device)
        call("fame6d_mass_apply_cols12", self._h, prec, ptr(x12), ptr(Y), nrhs)
        return Y

    def Mcol3op(self, x):
        """NF -> 3NF with blocks 1-2 = 0 (applyM_col3_rank1_scalar)."""
        x, nrhs, prec = self._prep(x, self.NF)
        shape = (3 * self.NF,) if x.dim() == 1 else (nrhs, 3 * self.NF)
        Y = torch.empty(shape, dtype=x.dtype, device=self.device)
        call("fame6d_mass_apply_col3", self._h, prec, ptr(x), ptr(Y), nrhs)
        return Y

    def adjoint_eps(self, x, gy):
        """Gradient of Re<gy, M_s x> with respect to the permittivity samples (FP64)."""
        x, nrhs, _ = self._prep(x, self.NF)
        gy, nrhs2, _ = self._prep(gy, self.NF)
        if nrhs != nrhs2:
            raise ValueError("x and gy must have the same number of columns")
        x = x.to(torch.complex128)
        gy = gy.to(torch.complex128)
        g = torch.zeros(self.NF, dtype=torch.float64, device=self.device)
        call("fame6d_mass_adjoint_eps", self._h, ptr(x), ptr(gy), nrhs, ptr(g))
        return g


class _EpsScope:
    def __init__(self, mass, eps):
        self.mass, self.eps, self.prev = mass, eps, None

    def __enter__(self):
        self.prev = self.mass.eps_tensor()
        self.mass.set_eps(self.eps)
        return self.mass

    def __exit__(self, *exc):
        self.mass.set_eps(self.prev)
        return False


def build_mass_backend_large_scalar(mat, fourier, massCfg=None, gpuInfo=None):
    """Reference-shaped factory: mat['scalarField'] holds eps, fourier['N'] the grid size
    (exp1/build_mass_backend_large_scalar.m:1)."""
    eps = mat["scalarField"] if isinstance(mat, dict) else mat
    N = fourier["N"] if isinstance(fourier, dict) else int(fourier)
    return MassBackend(N, np.asarray(eps).reshape(-1, order="F") if np.ndim(eps) > 1 else eps)
