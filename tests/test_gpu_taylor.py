"""GPU parity of the Taylor reconstruction (FP64 MFMA moments, Yee evaluation, direct sum)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(pkg, oracle):
    torch.cuda.set_device(0)
    rng = np.random.default_rng(21)
    N = 2
    qm = oracle.projected_q_modes(N, (0.13, 0.21, 0.17))
    NF = qm.shape[0]
    coeff = (rng.standard_normal(NF) + 1j * rng.standard_normal(NF)) * np.exp(-0.5 * (qm ** 2).sum(1))
    return dict(N=N, qm=qm, coeff=coeff, NF=NF)


@pytest.mark.parametrize("p,C", [(0, 1), (3, 5), (10, 8), (10, 27)])
def test_moments_vs_oracle(pkg, oracle, setup, p, C):
    print("case: taylor moments p=%d C=%d" % (p, C), flush=True)
    rng = np.random.default_rng(p + C)
    cen = rng.uniform(-0.4, 0.4, (C, 3))
    mom = pkg.taylor_moments(setup["qm"], setup["coeff"], p, cen).cpu().numpy()
    ref = oracle.taylor_moments(setup["qm"], setup["coeff"], p, cen)
    assert mom.shape == ref.shape
    err = np.abs(mom - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


def test_moments_many_chunks_per_workgroup(pkg, oracle):
    """N = 3 (46656 modes) with 216 centers: every workgroup walks several chunks of modes,
    so the double-buffered tables of the next chunk are built while the current one is
    consumed; p = 10 and p = 14 (the one-column-tile instances)."""
    rng = np.random.default_rng(33)
    qm = oracle.projected_q_modes(3, (0.05, 0.04, 0.03))
    NF = qm.shape[0]
    coeff = (rng.standard_normal(NF) + 1j * rng.standard_normal(NF)) * np.exp(-0.5 * (qm ** 2).sum(1))
    cen = rng.uniform(-0.6, 0.6, (216, 3))
    for p in (10, 14):
        print("case: taylor moments N=3 p=%d C=216" % p, flush=True)
        mom = pkg.taylor_moments(qm, coeff, p, cen).cpu().numpy()
        ref = oracle.taylor_moments(qm, coeff, p, cen)
        err = np.abs(mom - ref).max() / np.abs(ref).max()
        assert err < 1e-12, (p, err)


def test_yee_evaluation_vs_oracle_and_direct(pkg, oracle, setup):
    n, h, nb, p = 4, 0.0125, 2, 10
    print("case: taylor yee n=%d nb=%d p=%d" % (n, nb, p), flush=True)
    m = 2 * n + 3
    for comp in range(3):
        pts = oracle.yee_points(n, h, comp)
        bid = oracle.block_partition(m, nb)
        cen = oracle.block_centers(pts, bid, nb ** 3)
        mom = pkg.taylor_moments(setup["qm"], setup["coeff"], p, cen)
        u = torch.empty(m ** 3, dtype=torch.complex128, device="cuda")
        import ctypes as C
        from importlib import import_module
        lib = import_module("fame6d_amd._lib")
        cen_c = np.ascontiguousarray(cen)
        lib.call("fame6d_taylor_evaluate_yee", n, h, comp, nb, p, nb ** 3, cen_c.ctypes.data_as(C.c_void_p),
                 lib.ptr(mom.T.contiguous()), lib.ptr(u), C.c_void_p(torch.cuda.current_stream().cuda_stream))
        u = u.cpu().numpy()
        ref = oracle.taylor_evaluate(pts, bid, cen, mom.cpu().numpy(), p)
        assert np.abs(u - ref).max() < 1e-11 * np.abs(ref).max()
        # the generic evaluator on explicit points agrees
        u2 = pkg.taylor_evaluate(pts, bid, cen, mom, p).cpu().numpy()
        assert np.abs(u2 - ref).max() < 1e-11 * np.abs(ref).max()
        # Taylor vs the direct Fourier sum (small phase radius: error well below 1e-6)
        ud = pkg.direct_reconstruct(setup["qm"], setup["coeff"], pts).cpu().numpy()
        ud_ref = oracle.direct_reconstruct(setup["qm"], setup["coeff"], pts)
        assert np.abs(ud - ud_ref).max() < 1e-10 * np.abs(ud_ref).max()
        assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-6


@pytest.mark.parametrize("N,p,nb", [(2, 0, 2), (2, 3, 3), (3, 10, 3), (3, 10, 6), (3, 14, 4)])
def test_moments_tensor_grid_centres(pkg, oracle, N, p, nb):
    """Centres on a tensor grid (the block centres of a Yee window, c = i + nb j + nb^2 k)
    take the per-axis phase tables; the moments equal the oracle's one-phase-per-centre sum.
    The same centres with one of them moved off the grid take the per-centre phases and
    also match.  p = 14 runs the round-robin tile form, the others the balanced one."""
    print("case: taylor moments tensor-grid centres N=%d p=%d nb=%d" % (N, p, nb), flush=True)
    rng = np.random.default_rng(100 + 10 * N + p)
    qm = oracle.projected_q_modes(N, (0.05, 0.04, 0.03))
    NF = qm.shape[0]
    coeff = (rng.standard_normal(NF) + 1j * rng.standard_normal(NF)) * np.exp(-0.5 * (qm ** 2).sum(1))
    cen = np.ascontiguousarray(pkg.yee_block_centers(4 * nb, 0.05, 1, nb))
    for moved in (False, True):
        c = cen.copy()
        if moved:
            c[nb ** 3 // 2, 0] += 0.01
        mom = pkg.taylor_moments(qm, coeff, p, c).cpu().numpy()
        ref = oracle.taylor_moments(qm, coeff, p, c)
        err = np.abs(mom - ref).max() / np.abs(ref).max()
        assert err < 1e-12, (moved, err)
    # a mode count that is not a whole number of 64-mode chunks: the tensor-grid tables are
    # read from global memory instead of being staged in LDS ahead of each chunk
    k = NF - 37
    mom = pkg.taylor_moments(qm[:k], coeff[:k], p, cen).cpu().numpy()
    ref = oracle.taylor_moments(qm[:k], coeff[:k], p, cen)
    assert np.abs(mom - ref).max() / np.abs(ref).max() < 1e-12


def test_reconstruct_yee_field_vs_oracle(pkg, oracle, setup):
    """reconstruct_yee_field (exp3/reconstruct_tracked_modes_exp4.m:10-30): the three plus-grid
    components of a 6D field on the n = 4 window, with the automatic partition and with two
    blocks per axis, equal the oracle's moments and Taylor sums on the oracle's Yee points,
    block ids and centres; with the small phase radius of this window they also equal the
    direct Fourier sum."""
    rng = np.random.default_rng(77)
    qm, NF = setup["qm"], setup["NF"]
    X = (rng.standard_normal(3 * NF) + 1j * rng.standard_normal(3 * NF)) * np.tile(np.exp(-0.5 * (qm ** 2).sum(1)), 3)
    n, h, p = 4, 0.0125, 10
    m = 2 * n + 3
    Xd = torch.as_tensor(X, device="cuda")
    for nb_req in (None, 2):
        print("case: reconstruct_yee_field n=%d nblocks=%s" % (n, nb_req), flush=True)
        comps, nb = pkg.reconstruct_yee_field(Xd, qm, n, h, p=p, nblocks=nb_req)
        if nb_req is not None:
            assert nb == nb_req
        assert len(comps) == 3
        for comp in range(3):
            coeff = X[comp * NF:(comp + 1) * NF]
            pts = oracle.yee_points(n, h, comp)
            bid = oracle.block_partition(m, nb)
            cen = oracle.block_centers(pts, bid, nb ** 3)
            mom = oracle.taylor_moments(qm, coeff, p, cen)
            ref = oracle.taylor_evaluate(pts, bid, cen, mom, p)
            u = comps[comp].cpu().numpy()
            assert u.shape == (m ** 3,)
            assert np.abs(u - ref).max() < 1e-11 * np.abs(ref).max(), (nb, comp)
            ud = oracle.direct_reconstruct(qm, coeff, pts)
            assert np.linalg.norm(u - ud) / np.linalg.norm(ud) < 1e-6, (nb, comp)
