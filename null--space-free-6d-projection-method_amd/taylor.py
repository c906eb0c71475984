"""Multi-center Taylor reconstruction of Fourier fields on Yee points (MI355X kernels).

Mirrors reconstruct_component_taylor_batch_exp4 (exp3/reconstruct_component_taylor_batch_exp4.m:1-43),
build_multi_index_list_3d_exp3 (exp2/build_multi_index_list_3d_exp3.m:1-15),
build_yee_edge_points_plus (exp2/build_yee_edge_points_plus.m:1-36),
build_yee_block_partition_exp3 (exp2/build_yee_block_partition_exp3.m:1-31),
reconstruct_field_direct_blocked_exp3 (exp2/reconstruct_field_direct_blocked_exp3.m:1-29) and
compute_field_errors_exp3 (exp2/compute_field_errors_exp3.m:1-11).
"""
import ctypes as C

import numpy as np
import torch

from ._lib import call, ptr


def build_multi_index_list_3d(p):
    """All alpha in N_0^3 with |alpha| <= p, a1 outermost (exp2/build_multi_index_list_3d_exp3.m:3-14)."""
    return np.array([(a1, a2, a3) for a1 in range(p + 1) for a2 in range(p - a1 + 1)
                     for a3 in range(p - a1 - a2 + 1)], dtype=int)


def yee_points(n, h, comp):
    """Plus-grid points h*(I + 0.5 e_comp), I in {-(n+1)..n+1}^3, column-major (M, 3)."""
    idx = np.arange(-(n + 1), n + 2, dtype=float)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    pts = np.stack([a.reshape(-1, order="F") for a in (I, J, K)], 1)
    pts[:, comp] += 0.5
    return h * pts


def block_partition(m, nblocks):
    """Block id of every point of an m^3 plus-grid split into nblocks^3 nearly equal blocks
    (the first mod(m, nblocks) ranges one longer; block order i fastest)."""
    base, rem = divmod(m, nblocks)
    lab = np.concatenate([np.full(base + (1 if b < rem else 0), b) for b in range(nblocks)])
    I, J, K = np.meshgrid(lab, lab, lab, indexing="ij")
    return (I + nblocks * J + nblocks * nblocks * K).reshape(-1, order="F").astype(np.int32)


def block_centers(points, bid, nb3):
    """Midpoint of each block's bounding box (exp3/build_taylor_geometry_exp4.m:10-11)."""
    C_ = np.zeros((nb3, 3))
    for b in range(nb3):
        R = points[bid == b]
        C_[b] = (R.min(0) + R.max(0)) / 2
    return C_


def _dev(x, dtype, device):
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(np.asarray(x))
    return x.to(device=device, dtype=dtype).contiguous()


def taylor_moments(qmodes, coeff, p, centers, device="cuda"):
    """moments (L, C) complex128 on the device; qmodes (NF, 3), coeff (NF,)."""
    q = _dev(torch.as_tensor(np.asarray(qmodes, dtype=np.float64)).T.contiguous()
             if not isinstance(qmodes, torch.Tensor) else qmodes.T.contiguous(), torch.float64, device)
    c = _dev(coeff, torch.complex128, device)
    NF = c.numel()
    cen = np.ascontiguousarray(np.asarray(centers, dtype=np.float64).reshape(-1, 3))
    L = (p + 1) * (p + 2) * (p + 3) // 6
    mom = torch.empty((cen.shape[0], L), dtype=torch.complex128, device=device)  # column-major L x C
    call("fame6d_taylor_moments", NF, ptr(q), ptr(c), int(p), cen.shape[0], cen.ctypes.data_as(C.c_void_p),
         ptr(mom), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return mom.T  # (L, C) view


def taylor_evaluate(points, center_id, centers, moments, p, device="cuda"):
    pts = _dev(torch.as_tensor(np.asarray(points, dtype=np.float64)).T.contiguous()
               if not isinstance(points, torch.Tensor) else points.T.contiguous(), torch.float64, device)
    M = pts.shape[1]
    cid = _dev(center_id, torch.int32, device)
    cen = np.ascontiguousarray(np.asarray(centers, dtype=np.float64).reshape(-1, 3))
    mom = moments.T.contiguous() if moments.shape[0] != cen.shape[0] else moments.contiguous()
    u = torch.empty(M, dtype=torch.complex128, device=device)
    call("fame6d_taylor_evaluate", M, ptr(pts), ptr(cid), int(p), cen.shape[0], cen.ctypes.data_as(C.c_void_p),
         ptr(mom), ptr(u), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return u


def direct_reconstruct(qmodes, coeff, points, device="cuda"):
    q = _dev(torch.as_tensor(np.asarray(qmodes, dtype=np.float64)).T.contiguous(), torch.float64, device)
    c = _dev(coeff, torch.complex128, device)
    pts = _dev(torch.as_tensor(np.asarray(points, dtype=np.float64)).T.contiguous(), torch.float64, device)
    M = pts.shape[1]
    u = torch.empty(M, dtype=torch.complex128, device=device)
    call("fame6d_direct_reconstruct", c.numel(), ptr(q), ptr(c), M, ptr(pts), ptr(u),
         C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return u


def compute_field_errors(test, ref):
    """Relative 2-norm and max-norm errors of a three-component field, the norms taken over
    all three components together (exp2/compute_field_errors_exp3.m:3-10): test and ref are
    sequences (u1, u2, u3) of equally long complex128 device tensors.  Returns
    {'e2': ..., 'eInf': ...}."""
    t = [u.reshape(-1).to(torch.complex128).contiguous() for u in test]
    r = [u.reshape(-1).to(torch.complex128).contiguous() for u in ref]
    if len(t) != 3 or len(r) != 3:
        raise ValueError("test and ref must each hold three components")
    n = t[0].numel()
    if any(u.numel() != n for u in t + r):
        raise ValueError("all six components must have the same length")
    e2, einf = C.c_double(), C.c_double()
    call("fame6d_field_errors", n, *[ptr(u) for u in t + r], C.byref(e2), C.byref(einf),
         C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return dict(e2=e2.value, eInf=einf.value)


def projected_q_modes(N, k_lift, P):
    """q_xi = P^T (xi + k_lift) for every mode of the (2N)^6 grid in column-major order, (NF, 3)
    (exp3/build_projected_q_modes.m:14-20)."""
    f = np.concatenate([np.arange(N), np.arange(-N, 0)]).astype(float)
    NF = (2 * N) ** 6
    qm = np.zeros((NF, 3))
    P = np.asarray(P, dtype=float)
    for ell in range(6):
        # xi_ell of mode j: digit ell of j in base 2N (column-major vectorisation)
        xi = np.tile(np.repeat(f, (2 * N) ** ell), (2 * N) ** (5 - ell)) + k_lift[ell]
        for c in range(3):
            qm[:, c] += P[ell, c] * xi
    return qm


def auto_taylor_blocks(m, h, q_l1_max, target=5.20, max_blocks=32, slack=2e-12):
    """Blocks per dimension of build_auto_taylor_partition_exp4
    (exp3/build_auto_taylor_partition_exp4.m:9-25): the smallest nb whose conservative phase
    radius max||q||_1 * (ceil(m/nb) - 1) h/2 is within the target (cfg.taylor,
    exp3/make_config_experiment4.m:65-67)."""
    for nb in range(1, max_blocks + 1):
        half = max(0.0, (-(-m // nb) - 1) * h / 2)
        if q_l1_max * half <= target * (1 + slack):
            return nb, q_l1_max * half
    raise ValueError("No Taylor partition up to %d^3 centers satisfies rho<=%.3f (m=%d, h=%g, qL1Max=%g)"
                     % (max_blocks, target, m, h, q_l1_max))


def yee_block_centers(n, h, comp, nb):
    """Midpoints of the bounding boxes of the nb^3 blocks of component comp's plus-grid points
    (exp2/build_yee_block_partition_exp3.m:5-30: the first mod(m, nb) ranges one longer; block
    order i fastest), (nb^3, 3)."""
    m = 2 * n + 3
    base, rem = divmod(m, nb)
    mid, s = [], 0
    for b in range(nb):
        ln = base + (1 if b < rem else 0)
        mid.append(h * ((s + s + ln - 1) / 2.0 - (n + 1)))
        s += ln
    sh = [0.5 * h if comp == d else 0.0 for d in range(3)]
    return np.array([[mid[i] + sh[0], mid[j] + sh[1], mid[k] + sh[2]]
                     for k in range(nb) for j in range(nb) for i in range(nb)])


def reconstruct_yee_field(X, qmodes, n, h, p=10, nblocks=None, target=5.20):
    """The three plus-grid components of the 6D field X (3NF complex128, component-first) on
    the Yee window of half-width n h, by Taylor moments and Yee evaluation per component
    (exp3/reconstruct_tracked_modes_exp4.m:10-30).  qmodes: (NF, 3) array or a (3, NF)
    device tensor.  nblocks None chooses the automatic phase-radius partition.  Returns
    ([u1, u2, u3] flat (2n+3)^3 device tensors, nblocks)."""
    if isinstance(qmodes, torch.Tensor):
        q_d = qmodes.contiguous()
        q_l1 = float(q_d.abs().sum(0).max())
    else:
        qm = np.asarray(qmodes, dtype=np.float64)
        q_d = torch.as_tensor(qm.T.copy(), device="cuda")
        q_l1 = float(np.abs(qm).sum(1).max())
    NF = q_d.shape[1]
    m = 2 * n + 3
    if nblocks is None:
        nblocks, _ = auto_taylor_blocks(m, h, q_l1, target)
    L = (p + 1) * (p + 2) * (p + 3) // 6
    X = X.reshape(-1)
    if X.numel() != 3 * NF:
        raise ValueError("X must hold 3 NF values")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    mom = torch.empty((nblocks ** 3, L), dtype=torch.complex128, device=q_d.device)
    out = []
    for comp in range(3):
        cen = np.ascontiguousarray(yee_block_centers(n, h, comp, nblocks))
        coeff = X[comp * NF:(comp + 1) * NF].contiguous()
        call("fame6d_taylor_moments", NF, ptr(q_d), ptr(coeff), int(p), nblocks ** 3, cen.ctypes.data_as(C.c_void_p),
             ptr(mom), stream)
        u = torch.empty(m ** 3, dtype=torch.complex128, device=q_d.device)
        call("fame6d_taylor_evaluate_yee", int(n), float(h), comp, int(nblocks), int(p), nblocks ** 3,
             cen.ctypes.data_as(C.c_void_p), ptr(mom), ptr(u), stream)
        out.append(u)
    return out, nblocks
