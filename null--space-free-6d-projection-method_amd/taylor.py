"""Multi-center Taylor reconstruction of Fourier fields on Yee points (MI355X kernels).

Mirrors reconstruct_component_taylor_batch_exp4 (exp3/reconstruct_component_taylor_batch_exp4.m:1-43),
build_multi_index_list_3d_exp3 (exp2/build_multi_index_list_3d_exp3.m:1-15),
build_yee_edge_points_plus (exp2/build_yee_edge_points_plus.m:1-36),
build_yee_block_partition_exp3 (exp2/build_yee_block_partition_exp3.m:1-31) and
reconstruct_field_direct_blocked_exp3 (exp2/reconstruct_field_direct_blocked_exp3.m:1-29).
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
