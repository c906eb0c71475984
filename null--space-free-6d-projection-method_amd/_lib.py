"""ctypes binding of libfame6d.so (the C ABI declared in include/fame6d.h).

torch is imported first so that the library binds to the HIP runtime torch
already loaded (both carry the soname libamdhip64.so.7): device pointers and
streams are then shared between torch tensors and the library.
"""
import ctypes as C
import os

import torch  # noqa: F401  (must precede the library load, see above)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfame6d.so")

FP64, FP32, MIXED = 0, 1, 2
STATUS = {0: "OK", 1: "ESIZE", 2: "ENONFINITE", 3: "ENOCONV", 4: "EBUDGET", 5: "ENOMEM",
          6: "EHIP", 7: "EARG", 8: "EGAMMA", 9: "EMASSNORM"}


class Fame6dError(RuntimeError):
    """A non-zero status of the C ABI; .code is the FAME6D_E* value."""

    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code
        self.name = STATUS.get(code, "?")


class MassInfo(C.Structure):
    _fields_ = [("N", C.c_int), ("nGrid", C.c_int), ("NF", C.c_longlong),
                ("eps_min", C.c_double), ("eps_max", C.c_double), ("eps_mean", C.c_double),
                ("tile_a64", C.c_int), ("tile_b64", C.c_int), ("tile_a32", C.c_int), ("tile_b32", C.c_int),
                ("variant_ac64", C.c_int), ("variant_b64", C.c_int), ("variant_ac32", C.c_int),
                ("variant_b32", C.c_int)]


class QsysInfo(C.Structure):
    _fields_ = [("NF", C.c_longlong), ("min_abs_det_v1", C.c_double), ("min_qq", C.c_double),
                ("n_singular", C.c_int), ("k_lift", C.c_double * 6)]


class KrinvStats(C.Structure):
    _fields_ = [("calls", C.c_longlong), ("total_iterations", C.c_longlong),
                ("total_fp64_iterations", C.c_longlong), ("max_iterations", C.c_int),
                ("last_iterations", C.c_int), ("last_flag", C.c_int), ("last_relres", C.c_double),
                ("max_relres", C.c_double), ("refinement_cycles", C.c_longlong),
                ("mass_applies_fp64", C.c_longlong), ("mass_applies_fp32", C.c_longlong),
                ("seconds", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LanczosStats(C.Structure):
    _fields_ = [("converged", C.c_int), ("num_converged", C.c_int), ("outer_steps", C.c_int),
                ("restarts", C.c_int), ("hit_max_restart", C.c_int), ("max_ritz_residual", C.c_double),
                ("seconds", C.c_double), ("krinv", KrinvStats), ("second_sweeps", C.c_int)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "krinv"}
        d["krinv"] = self.krinv.as_dict()
        return d


class RecoverStats(C.Structure):
    _fields_ = [("iterations", C.c_int * 64), ("relres", C.c_double * 64), ("mass_norm2", C.c_double * 64)]


class MinresStats(C.Structure):
    _fields_ = [("solves", C.c_longlong), ("iterations", C.c_longlong), ("segments", C.c_longlong),
                ("restarts", C.c_longlong), ("max_iterations", C.c_int), ("plateau_accepts", C.c_int),
                ("max_relres", C.c_double), ("seconds", C.c_double), ("corrections", C.c_longlong),
                ("correction_iterations", C.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelTime(C.Structure):
    _fields_ = [("launches", C.c_longlong), ("milliseconds", C.c_double), ("bytes", C.c_double)]


class LwrqResult(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("V", C.c_double), ("nu", C.c_double), ("partition_error", C.c_double),
                ("energy_partition_error", C.c_double), ("mass_partition_error", C.c_double),
                ("sum_a", C.c_double), ("sum_b", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class YeeMetrics(C.Structure):
    _fields_ = [("lambda_crop", C.c_double), ("lambda_3d", C.c_double), ("eta_3d", C.c_double),
                ("eta_6d", C.c_double), ("delta_lambda", C.c_double), ("V", C.c_double), ("nu", C.c_double),
                ("delta_part", C.c_double), ("delta_identity", C.c_double), ("cropped_mass", C.c_double),
                ("energy_partition_error", C.c_double), ("mass_partition_error", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


NUM_TIMER_SLOTS = 5
TIMER_SLOTS = ("mhat_pass_a", "mhat_pass_b", "mhat_pass_c", "lanczos_gemv_c", "lanczos_gemv_n")

P = C.c_void_p
I = C.c_int
D = C.c_double
LL = C.c_longlong
DP = C.POINTER(C.c_double)

# name -> argtypes (restype is always int except where noted)
SIGNATURES = {
    "fame6d_version": [],
    "fame6d_mass_create": [I, P, P, C.POINTER(P)],
    "fame6d_mass_destroy": [P],
    "fame6d_mass_set_eps": [P, P],
    "fame6d_mass_set_eps_device": [P, P],
    "fame6d_mass_info": [P, C.POINTER(MassInfo)],
    "fame6d_mass_autotune": [P, I, I],
    "fame6d_mass_set_tiles": [P, I, I, I],
    "fame6d_mass_profile": [P, I],
    "fame6d_mass_kernel_times": [P, C.POINTER(KernelTime)],
    "fame6d_mass_apply_scalar": [P, I, P, P, I],
    "fame6d_mass_apply_scalar_noalloc": [P, I, P, P, I],
    "fame6d_mass_apply_3comp": [P, I, P, P, I],
    "fame6d_mass_apply_cols12": [P, I, P, P, I],
    "fame6d_mass_apply_col3": [P, I, P, P, I],
    "fame6d_qsys_create": [P, P, P, C.POINTER(P)],
    "fame6d_qsys_destroy": [P],
    "fame6d_qsys_info": [P, C.POINTER(QsysInfo)],
    "fame6d_qsys_diagonals": [P, P, P, P, P, P, P, P, P, P],
    "fame6d_qsys_set_mixed": [P, D, I, I],
    "fame6d_qsys_set_reorth": [P, D],
    "fame6d_mhat_apply": [P, I, P, P],
    "fame6d_apply_v": [P, P, P],
    "fame6d_apply_k": [P, P, P],
    "fame6d_krinv_apply": [P, I, P, P, D, I, C.POINTER(KrinvStats)],
    "fame6d_eigs_method_c": [P, I, I, I, D, I, D, I, P, P, P, C.POINTER(LanczosStats)],
    "fame6d_eigs_tracking": [P, I, I, I, D, I, D, I, P, I, P, I, P, I, P, P, P, I, P, P, C.POINTER(LanczosStats)],
    "fame6d_eigs_method_c_batch": [P, I, P, P, I, I, I, D, I, D, I, P, P, P, C.POINTER(LanczosStats)],
    "fame6d_krinv_nested_apply": [P, P, P, D, D, I, I, D, C.POINTER(KrinvStats)],
    "fame6d_eigs_method_b": [P, I, I, D, I, D, D, I, I, D, P, P, P, C.POINTER(LanczosStats)],
    "fame6d_method_a_apply": [P, D, D, I, P, P, C.POINTER(MinresStats)],
    "fame6d_eigs_method_a": [P, I, I, D, I, D, D, I, P, P, P, C.POINTER(LanczosStats), C.POINTER(MinresStats)],
    "fame6d_recover_eigenvectors": [P, P, I, D, I, P, P, C.POINTER(RecoverStats)],
    "fame6d_fourier_diagnostics": [P, P, P, P, I, P, P, P],
    "fame6d_taylor_moments": [LL, P, P, I, I, P, P, P],
    "fame6d_taylor_evaluate": [LL, P, P, I, I, P, P, P, P],
    "fame6d_taylor_evaluate_yee": [I, D, I, I, I, I, P, P, P, P],
    "fame6d_direct_reconstruct": [LL, P, P, LL, P, P, P],
    "fame6d_field_errors": [LL, P, P, P, P, P, P, P, P, P],
    "fame6d_mass_adjoint_eps": [P, P, P, I, P],
    "fame6d_eig_grad_eps": [P, P, I, P, P, P],
    "fame6d_yee_workspace_bytes": [I],
    "fame6d_yee_apply_cropped_A": [I, D, P, P, P, P, P, P, P],
    "fame6d_yee_sample_eps": [I, D, P, I, D, D, P, P, P, P, P],
    "fame6d_lwrq": [I, D, P, P, I, D, D, P, P, C.POINTER(LwrqResult), P],
    "fame6d_yee_lwrq_metrics": [I, D, P, P, P, P, I, D, D, D, I, D, D, P, P, C.POINTER(YeeMetrics), P],
}
RESTYPES = {"fame6d_yee_workspace_bytes": C.c_size_t}

_lib = None


def load(path=LIB_PATH):
    """Loads the library (once) and declares every prototype.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libfame6d.so is not built (%s); run __graft_entry__.build()" % path)
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    lib.fame6d_last_error.restype = C.c_char_p
    lib.fame6d_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def check(code):
    if code != 0:
        raise Fame6dError(code, load().fame6d_last_error().decode(errors="replace"))


def call(name, *args):
    check(getattr(load(), name)(*args))


def ptr(t):
    """Device (or host) address of a torch tensor / numpy array, None -> NULL."""
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        return C.c_void_p(t.data_ptr())
    return C.c_void_p(t.ctypes.data)
