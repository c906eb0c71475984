"""Method-C eigensolver on the MI355X: q-system, Mhat, K_r^{-1}, Lanczos, recovery.

Mirrors build_experiment1_q_system_C_compact (exp1/build_experiment1_q_system_C_compact.m:7-38),
make_A_eigs_top_preconditioned / run_method_A_fig1a (exp1/make_A_eigs_top_preconditioned.m:1-113,
exp1/run_method_A_fig1a.m:1-50),
make_explicit_krinv_compact (exp1/make_explicit_krinv_compact.m:1-98),
run_method_C_fig1bc (exp1/run_method_C_fig1bc.m:1-36),
recover_original_eigenvectors_exp2 (exp2/recover_original_eigenvectors_exp2.m:1-57) and
compute_fourier_diagnostics_exp2 (exp2/compute_fourier_diagnostics_exp2.m:1-27).
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import FP32, FP64, MIXED, call, ptr
from .mass import P_DEFAULT, MassBackend, precision_code, start_vector


def _start_vectors(v0, rows, n):
    """(owner, pointer) of rows x n complex128 start vectors: a CUDA tensor stays on the
    device (the library copies it device to device), anything else becomes a host array."""
    if isinstance(v0, torch.Tensor) and v0.is_cuda:
        if v0.numel() != rows * n:
            raise _lib.Fame6dError(1, "start vector must have %d entries" % (rows * n))
        v = v0.to(torch.complex128).reshape(rows, n).contiguous()
        return v, ptr(v)
    v = np.ascontiguousarray(np.asarray(v0, dtype=np.complex128).reshape(-1))
    if v.size != rows * n:
        raise _lib.Fame6dError(1, "start vector must have %d entries" % (rows * n))
    return v, v.ctypes.data_as(C.c_void_p)


class QSystem:
    """Null-space-free system of one Bloch point on a mass backend."""

    def __init__(self, mass: MassBackend, q, P=P_DEFAULT):
        self.mass = mass
        self.q = np.asarray(q, dtype=np.float64).reshape(3)
        self.P = np.ascontiguousarray(np.asarray(P, dtype=np.float64).reshape(6, 3))
        self.NF = mass.NF
        self.nr = 2 * mass.NF
        self.ndof = 3 * mass.NF
        self._h = C.c_void_p()
        call("fame6d_qsys_create", mass.handle, self.q.ctypes.data_as(C.c_void_p),
             self.P.ctypes.data_as(C.c_void_p), C.byref(self._h))

    def close(self):
        if self._h:
            call("fame6d_qsys_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def info(self):
        i = _lib.QsysInfo()
        call("fame6d_qsys_info", self._h, C.byref(i))
        return dict(NF=i.NF, minAbsDetV1=i.min_abs_det_v1, minQq=i.min_qq, nSingular=i.n_singular,
                    kLift=np.array(list(i.k_lift)))

    def diagonals(self):
        """Host copies of V11, V12, V21, V22, V31, V32, W1, W2, invC."""
        names = ["V11", "V12", "V21", "V22", "V31", "V32", "W1", "W2", "invC"]
        arrs = [np.empty(self.NF) for _ in names]
        call("fame6d_qsys_diagonals", self._h, *[a.ctypes.data_as(C.c_void_p) for a in arrs])
        return dict(zip(names, arrs))

    def set_mixed(self, inner_tol=3e-5, max_cycles=30, precondition=True):
        call("fame6d_qsys_set_mixed", self._h, float(inner_tol), int(max_cycles), int(bool(precondition)))

    def set_reorth(self, eta=2 ** -0.5):
        """DGKS threshold of the Lanczos re-orthogonalisation: a second Gram-Schmidt sweep runs
        when the first leaves less than eta of ||r||; eta >= 1 runs it at every step."""
        call("fame6d_qsys_set_reorth", self._h, float(eta))

    def _dev(self, x, rows, dtype=torch.complex128):
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.asarray(x))
        x = x.to(device=self.mass.device, dtype=dtype).contiguous()
        if x.numel() != rows:
            raise _lib.Fame6dError(1, "expected %d entries (BadSize)" % rows)
        return x

    def Mhat(self, x, precision="double"):
        """apply_Mhat_C_scalar_compact (exp1/apply_Mhat_C_scalar_compact.m:8-21)."""
        pc = precision_code(precision)
        x = self._dev(x, self.NF, torch.complex64 if pc == FP32 else torch.complex128)
        y = torch.empty_like(x)
        call("fame6d_mhat_apply", self._h, pc, ptr(x), ptr(y))
        return y

    def apply_V(self, y):
        """apply_V_blockdiag (exp1/apply_V_blockdiag.m:9-16)."""
        y = self._dev(y, self.nr)
        out = torch.empty(self.ndof, dtype=torch.complex128, device=self.mass.device)
        call("fame6d_apply_v", self._h, ptr(y), ptr(out))
        return out

    def Kop(self, x):
        """Kop (exp1/assemble_K_and_nullspace_free_basis.m:110-119)."""
        x = self._dev(x, self.ndof)
        out = torch.empty_like(x)
        call("fame6d_apply_k", self._h, ptr(x), ptr(out))
        return out

    def krinv(self, q, tol=1e-10, maxit=1200, precision="double"):
        """One K_r^{-1} action; returns (y, stats)."""
        q = self._dev(q, self.nr)
        y = torch.empty_like(q)
        st = _lib.KrinvStats()
        call("fame6d_krinv_apply", self._h, precision_code(precision), ptr(q), ptr(y), float(tol),
             int(maxit), C.byref(st))
        return y, st.as_dict()


def build_experiment1_q_system_C_compact(common, qPhys, cfg=None):
    """Reference-shaped constructor: common['mass'] is a MassBackend."""
    mass = common["mass"] if isinstance(common, dict) else common
    P = common.get("P", P_DEFAULT) if isinstance(common, dict) else P_DEFAULT
    return QSystem(mass, qPhys, P)


def make_explicit_krinv_compact(sys: QSystem, tolCG, maxitCG, precision="double"):
    """Returns (KrInvOp, getStats) as exp1/make_explicit_krinv_compact.m:1-10."""
    acc = dict(inverseCalls=0, totalMhatIterations=0, maxMhatIterations=0, maxMhatRelativeResidual=0.0,
               iterations=[], relativeResiduals=[], flags=[])

    def op(q):
        qq = q if (isinstance(q, torch.Tensor) and q.dim() == 1) or np.ndim(q) == 1 else None
        if qq is None:
            return torch.stack([op(q[j]) for j in range(q.shape[0])])
        y, st = sys.krinv(q, tolCG, maxitCG, precision)
        acc["inverseCalls"] += 1
        acc["totalMhatIterations"] += st["total_iterations"]
        acc["maxMhatIterations"] = max(acc["maxMhatIterations"], st["last_iterations"])
        acc["maxMhatRelativeResidual"] = max(acc["maxMhatRelativeResidual"], st["last_relres"])
        acc["iterations"].append(st["last_iterations"])
        acc["relativeResiduals"].append(st["last_relres"])
        acc["flags"].append(st["last_flag"])
        return y

    def get_stats():
        s = dict(acc)
        s["averageMhatIterations"] = (s["totalMhatIterations"] / s["inverseCalls"]) if s["inverseCalls"] else float("nan")
        return s

    return op, get_stats


def run_method_C(sys: QSystem, k=10, p=40, tol=1e-10, max_restart=50, tol_cg=1e-10, maxit_cg=1200,
                 seed=20260905, v0=None, precision="double", want_vectors=False):
    """run_method_C_fig1bc + Lanczos_hybrid_identity_gpuop with a device-resident basis.
    Returns dict(lambda (ascending), Y (reduced Ritz vectors, (k, 2NF) tensor or None), stats)."""
    if v0 is None:
        v0 = start_vector(sys.nr, seed)
    v0, v0p = _start_vectors(v0, 1, sys.nr)
    lam = np.empty(k)
    Y = torch.empty((k, sys.nr), dtype=torch.complex128, device=sys.mass.device) if want_vectors else None
    st = _lib.LanczosStats()
    call("fame6d_eigs_method_c", sys.handle, precision_code(precision), int(k), int(p), float(tol),
         int(max_restart), float(tol_cg), int(maxit_cg), v0p,
         lam.ctypes.data_as(C.c_void_p), ptr(Y), C.byref(st))
    return dict(lambda_=lam, Y=Y, stats=st.as_dict())


def solve_tracking_scan(sys: QSystem, k=140, p=None, tol=1e-10, max_restart=30, tol_cg=1e-10, maxit_cg=3000,
                        fixed_indices=None, target_lambda=None, half_width=0, seed=None, v0=None,
                        precision="double", want_vectors=True):
    """solve_tracking_scan + Lanczos_tracking (exp3/solve_tracking_scan.m:1-56,
    exp3/Lanczos_tracking.m:1-188): the k lowest eigenvalues (p = k + 48 by default, the
    N = 5 scan of exp3/make_config_experiment4.m:48-50) and only the selected Ritz vectors.
    Select by fixed_indices (1-based) or by target_lambda (+/- half_width).  The default start
    vector seed is the reference's rngSeed + 1000 N + k (exp3/solve_tracking_scan.m:4; the
    generator itself is this package's)."""
    p = k + 48 if p is None else p
    if v0 is None:
        v0 = start_vector(sys.nr, (20260831 + 1000 * sys.mass.N + k) if seed is None else seed)
    v0, v0p = _start_vectors(v0, 1, sys.nr)
    fixed = None if fixed_indices is None else np.ascontiguousarray(np.asarray(fixed_indices, dtype=np.int32))
    tgt = None if target_lambda is None else np.ascontiguousarray(np.asarray(target_lambda, dtype=np.float64))
    cap = len(fixed) if fixed is not None else len(tgt) * (2 * half_width + 1)
    lam = np.empty(k)
    rel = np.empty(k)
    sel = np.zeros(cap, dtype=np.int32)
    nsel = C.c_int(0)
    Y = torch.empty((cap, sys.nr), dtype=torch.complex128, device=sys.mass.device) if want_vectors else None
    st = _lib.LanczosStats()
    call("fame6d_eigs_tracking", sys.handle, precision_code(precision), int(k), int(p), float(tol), int(max_restart),
         float(tol_cg), int(maxit_cg), v0p, 0 if fixed is None else len(fixed),
         None if fixed is None else fixed.ctypes.data_as(C.c_void_p), 0 if tgt is None else len(tgt),
         None if tgt is None else tgt.ctypes.data_as(C.c_void_p), int(half_width), lam.ctypes.data_as(C.c_void_p),
         rel.ctypes.data_as(C.c_void_p), sel.ctypes.data_as(C.c_void_p), cap, C.byref(nsel), ptr(Y), C.byref(st))
    ns = nsel.value
    idx = sel[:ns].copy()
    return dict(kScan=k, pScan=p, lambda_=lam, Yselected=None if Y is None else Y[:ns], selectedIndices=idx,
                selectedLambda=lam[idx - 1], selectedRelRitzResidual=rel[idx - 1], relRitzResiduals=rel,
                stats=st.as_dict())


def solve_tracked_modes(sys: QSystem, modes=(11, 114), lambda_reference=(1.200090550277e-02, 4.462939906027e-02),
                        lambda_check_rel_tol=5e-7, required_rel_ritz_residual=1e-10, k=140, p_extra=48,
                        tol_recover=1e-12, maxit_recover=3000, **kw):
    """solve_tracked_modes_exp4 (exp3/solve_tracked_modes_exp4.m:1-58): the scan, its checks
    (convergence, the selected indices, the selected relative inverse-Ritz residuals, the
    eigenvalues against the reference values, cfg.tracking of exp3/make_config_experiment4.m:37-45)
    and the recovery of the selected eigenvectors (exp3/recover_fixed_vectors_tracking.m)."""
    if max(modes) > k:
        raise ValueError("kScan=%d does not include tracked mode %d." % (k, max(modes)))
    scan = solve_tracking_scan(sys, k=k, p=k + p_extra, fixed_indices=modes, **kw)
    if not scan["stats"]["converged"]:
        raise _lib.Fame6dError(3, "selected-mode Lanczos did not satisfy the requested convergence test")
    if list(scan["selectedIndices"]) != list(modes):
        raise _lib.Fame6dError(3, "Selected Ritz indices do not match requested tracked modes.")
    if np.any(scan["selectedRelRitzResidual"] > required_rel_ritz_residual):
        raise _lib.Fame6dError(3, "Selected relative inverse-Ritz residual exceeds required threshold.")
    X, MX, info = recover_original_eigenvectors(scan["Yselected"], sys, tol_recover, maxit_recover)
    lam = scan["selectedLambda"]
    if lambda_reference is not None:
        ref = np.asarray(lambda_reference, dtype=float)
        rel = np.abs(lam - ref) / np.maximum(np.abs(ref), np.finfo(float).eps)
        if np.any(rel > lambda_check_rel_tol):
            raise _lib.Fame6dError(3, "Tracked-mode eigenvalue sanity check failed: %s" % rel)
    return dict(modes=list(modes), lambda6D=lam, X=X, MX=MX, scan=scan, recovery=info)


def krinv_nested(sys: QSystem, b, tol_kr=1e-10, tol_m=1e-12, maxit_kr=600, maxit_m=1600, eps_c=None):
    """Method-B K_r^{-1} b (make_nested_krinv_instrumented); eps_c defaults to the mean permittivity."""
    eps_c = sys.mass.info["epsMean"] if eps_c is None else eps_c
    b = sys._dev(b, sys.nr)
    x = torch.empty_like(b)
    st = _lib.KrinvStats()
    call("fame6d_krinv_nested_apply", sys.handle, ptr(b), ptr(x), float(tol_kr), float(tol_m), int(maxit_kr),
         int(maxit_m), float(eps_c), C.byref(st))
    return x, st.as_dict()


def run_method_B(sys: QSystem, k=10, p=40, tol=1e-10, max_restart=50, tol_kr=1e-10, tol_m=1e-12, maxit_kr=600,
                 maxit_m=1600, eps_c=None, seed=20260905, v0=None, want_vectors=False):
    """Method B (nested PCG) eigensolve with the same Lanczos as run_method_C
    (cfg.nested of exp1/make_config_exp1bc.m:83-87)."""
    eps_c = sys.mass.info["epsMean"] if eps_c is None else eps_c
    if v0 is None:
        v0 = start_vector(sys.nr, seed)
    v0, v0p = _start_vectors(v0, 1, sys.nr)
    lam = np.empty(k)
    Y = torch.empty((k, sys.nr), dtype=torch.complex128, device=sys.mass.device) if want_vectors else None
    st = _lib.LanczosStats()
    call("fame6d_eigs_method_b", sys.handle, int(k), int(p), float(tol), int(max_restart), float(tol_kr),
         float(tol_m), int(maxit_kr), int(maxit_m), float(eps_c), v0p,
         lam.ctypes.data_as(C.c_void_p), ptr(Y), C.byref(st))
    return dict(lambda_=lam, Y=Y, stats=st.as_dict())


def method_a_top(sys: QSystem, x, sigma=1e-4, tol=1e-10, maxit=4000):
    """The Method-A operator y = (K - sigma M)^{-1} M x (make_A_eigs_top_preconditioned applyT)
    by preconditioned MINRES; x, y: 3NF."""
    x = sys._dev(x, sys.ndof)
    y = torch.empty_like(x)
    st = _lib.MinresStats()
    call("fame6d_method_a_apply", sys.handle, float(sigma), float(tol), int(maxit), ptr(x), ptr(y), C.byref(st))
    return y, st.as_dict()


def run_method_A(sys: QSystem, k=10, p=40, tol=1e-10, max_restart=300, sigma=1e-4, tol_minres=1e-10,
                 maxit_minres=4000, seed=20260906, v0=None, want_vectors=False):
    """run_method_A_fig1a: the k largest theta of T = (K - sigma M)^{-1} M on the full 3NF space,
    lambda = sigma + 1/theta ascending (cfg.eigs / cfg.shift of exp1/make_config_exp1bc.m:36-72).
    Returns dict(lambda (ascending), Y ((k, 3NF) Ritz vectors or None), stats, minres)."""
    if v0 is None:
        v0 = start_vector(sys.ndof, seed)
    v0, v0p = _start_vectors(v0, 1, sys.ndof)
    lam = np.empty(k)
    Y = torch.empty((k, sys.ndof), dtype=torch.complex128, device=sys.mass.device) if want_vectors else None
    st = _lib.LanczosStats()
    ms = _lib.MinresStats()
    call("fame6d_eigs_method_a", sys.handle, int(k), int(p), float(tol), int(max_restart), float(sigma),
         float(tol_minres), int(maxit_minres), v0p, lam.ctypes.data_as(C.c_void_p),
         ptr(Y), C.byref(st), C.byref(ms))
    return dict(lambda_=lam, Y=Y, stats=st.as_dict(), minres=ms.as_dict())


def run_method_C_batch(mass: MassBackend, qs, k=10, p=40, tol=1e-10, max_restart=50, tol_cg=1e-10,
                       maxit_cg=1200, seed=20260905, v0=None, precision="double", P=P_DEFAULT,
                       want_vectors=False):
    """All Bloch points of a path solved together; one batched K_r^{-1} per Lanczos step."""
    qs = np.ascontiguousarray(np.asarray(qs, dtype=np.float64).reshape(-1, 3))
    nq = qs.shape[0]
    nr = 2 * mass.NF
    if v0 is None:
        v0 = np.stack([start_vector(nr, seed + i) for i in range(nq)])
    v0, v0p = _start_vectors(v0, nq, nr)
    Pc = np.ascontiguousarray(np.asarray(P, dtype=np.float64).reshape(6, 3))
    lam = np.empty((nq, k))
    Y = torch.empty((nq, k, nr), dtype=torch.complex128, device=mass.device) if want_vectors else None
    st = (_lib.LanczosStats * nq)()
    call("fame6d_eigs_method_c_batch", mass.handle, nq, qs.ctypes.data_as(C.c_void_p),
         Pc.ctypes.data_as(C.c_void_p), precision_code(precision), int(k), int(p), float(tol),
         int(max_restart), float(tol_cg), int(maxit_cg), v0p,
         lam.ctypes.data_as(C.c_void_p), ptr(Y), st)
    return dict(lambda_=lam, Y=Y, stats=[s.as_dict() for s in st])


def recover_original_eigenvectors(Y, sys: QSystem, tol=1e-12, maxit=3000):
    """X = M^{-1} V Y, M-normalised; returns (X, MX, info) with (k, 3NF) tensors."""
    Y = Y.to(sys.mass.device, torch.complex128).contiguous()
    k = Y.shape[0]
    X = torch.empty((k, sys.ndof), dtype=torch.complex128, device=sys.mass.device)
    MX = torch.empty_like(X)
    st = _lib.RecoverStats()
    call("fame6d_recover_eigenvectors", sys.handle, ptr(Y), k, float(tol), int(maxit), ptr(X), ptr(MX),
         C.byref(st))
    info = dict(iterations=list(st.iterations)[:k], relativeResiduals=list(st.relres)[:k],
                massNorm2=list(st.mass_norm2)[:k])
    return X, MX, info


def compute_fourier_diagnostics(X, MX, lam, sys: QSystem):
    k = X.shape[0]
    lam = np.ascontiguousarray(np.asarray(lam, dtype=np.float64).reshape(-1))
    res, div, ray = np.empty(k), np.empty(k), np.empty(k)
    call("fame6d_fourier_diagnostics", sys.handle, ptr(X.contiguous()), ptr(MX.contiguous()),
         lam.ctypes.data_as(C.c_void_p), k, res.ctypes.data_as(C.c_void_p),
         div.ctypes.data_as(C.c_void_p), ray.ctypes.data_as(C.c_void_p))
    return dict(resGEVP=res, divRelative2=div, rayleigh=ray, maxResGEVP=res.max(), maxDivRelative2=div.max())
