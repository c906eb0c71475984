"""Pins of the Yee / Taylor geometry (no GPU): the oracle's plus-grid points, block partition,
block centres and cosine permittivity, each by what the mathematics or the reference's rule
fixes, and the product's host helpers (taylor.py, mass.py) against them."""
import itertools
import math

import numpy as np
import pytest

Q_EXP1 = (0.05, 0.04, 0.03)


def _plane_wave_at(points_per_comp, kvec, p, m):
    """u_c = p_c exp(i k . x) sampled at the given points of each component, as (m, m, m)
    arrays (column-major point order)."""
    return [(p[c] * np.exp(1j * (pts @ kvec))).reshape((m, m, m), order="F")
            for c, pts in enumerate(points_per_comp)]


def _eig_defect(oracle, u, h, lam4):
    A = oracle.apply_cropped_A(*u, h)
    return max(np.max(np.abs(a - lam4 * x[1:-1, 1:-1, 1:-1])) / (lam4 * np.max(np.abs(x)))
               for a, x in zip(A, u))


def test_yee_points_stagger_fixed_by_curl_curl(oracle):
    """A transverse staggered plane wave sampled at oracle.yee_points is an exact eigenvector of
    the cropped Yee curl-curl (pinned independently: it equals curl^T curl and annihilates
    discrete gradients, tests/test_oracle.py), A u = |sigma|^2 u with
    sigma_d = (2/h) sin(k_d h / 2).  The same wave sampled with the half-cell shift on another
    axis is not, so a wrong stagger fails here."""
    n, h = 4, 0.1
    m = 2 * n + 3
    kvec = np.array([1.3, -0.7, 2.1])
    sigma = 2.0 / h * np.sin(kvec * h / 2)
    p = np.cross(sigma, [0.3, 1.0, -0.4]) * (1.0 + 0.5j)
    lam4 = float(sigma @ sigma)
    pts = [oracle.yee_points(n, h, c) for c in range(3)]
    for c in range(3):
        assert pts[c].shape == (m ** 3, 3)
        # plus-grid index range -(n+1)..n+1 on every axis, the half shift on axis c only
        for d in range(3):
            vals = np.unique(np.round(pts[c][:, d] / h, 12))
            lo = -(n + 1) + (0.5 if d == c else 0.0)
            assert np.allclose(vals, lo + np.arange(m))
        # column-major order: x fastest
        assert np.allclose(pts[c][1, :] - pts[c][0, :], [h, 0, 0])
        assert np.allclose(pts[c][m, :] - pts[c][0, :], [0, h, 0])
        assert np.allclose(pts[c][m * m, :] - pts[c][0, :], [0, 0, h])
    assert _eig_defect(oracle, _plane_wave_at(pts, kvec, p, m), h, lam4) < 1e-11
    # planted mistake: the half-cell shift of component c taken along axis c+1
    wrong = []
    for c in range(3):
        q = pts[c].copy()
        q[:, c] -= 0.5 * h
        q[:, (c + 1) % 3] += 0.5 * h
        wrong.append(q)
    assert _eig_defect(oracle, _plane_wave_at(wrong, kvec, p, m), h, lam4) > 1e-3


def _reference_ranges(m, nb):
    """split_one_dimension of exp2/build_yee_block_partition_exp3.m:21-31, 0-based ranges."""
    base, remn = m // nb, m % nb
    out, s = [], 0
    for b in range(1, nb + 1):
        ln = base + (1 if b <= remn else 0)
        out.append(list(range(s, s + ln)))
        s += ln
    return out


@pytest.mark.parametrize("m,nb", [(11, 2), (11, 3), (13, 4), (83, 6), (7, 7), (9, 1)])
def test_block_partition(oracle, pkg, m, nb):
    bid = oracle.block_partition(m, nb)
    assert bid.shape == (m ** 3,)
    # per-axis labels read back from the column-major ids
    ids = bid.reshape((m, m, m), order="F")
    lab_x = ids[:, 0, 0]
    lab_y = ids[0, :, 0] // nb
    lab_z = ids[0, 0, :] // (nb * nb)
    for lab in (lab_x, lab_y, lab_z):
        assert np.all(np.diff(lab) >= 0) and lab[0] == 0 and lab[-1] == nb - 1  # contiguous, ordered
        sizes = np.bincount(lab, minlength=nb)
        assert sizes.sum() == m and sizes.max() - sizes.min() <= 1
        assert np.all(np.diff(sizes) <= 0)  # the longer ranges first
        assert sizes[0] == -(-m // nb)
    # block id = i + nb j + nb^2 k with i fastest, every point in exactly one block
    I, J, K = np.meshgrid(lab_x, lab_x, lab_x, indexing="ij")
    assert np.array_equal(ids, I + nb * J + nb * nb * K)
    assert np.array_equal(np.bincount(bid, minlength=nb ** 3).sum(), m ** 3)
    # the reference's construction: blocks{c} for k, j, i loops (i innermost), sub2ind order
    rng = _reference_ranges(m, nb)
    c = 0
    for k in range(nb):
        for j in range(nb):
            for i in range(nb):
                lin = [a + m * b + m * m * cc for cc in rng[k] for b in rng[j] for a in rng[i]]
                assert np.all(bid[lin] == c)
                c += 1
    # the product's partition is the same
    assert np.array_equal(pkg.block_partition(m, nb), bid)


def test_block_centers_bruteforce(oracle, pkg):
    """Centres are the bounding-box midpoints (exp3/build_taylor_geometry_exp4.m:10-11), found
    here by a loop over every point; the product's yee_block_centers gives them without
    forming the points."""
    for n, nb, h in ((2, 3, 0.1), (4, 2, 0.0125), (5, 4, 0.05)):
        m = 2 * n + 3
        bid = oracle.block_partition(m, nb)
        for comp in range(3):
            pts = oracle.yee_points(n, h, comp)
            lo = np.full((nb ** 3, 3), np.inf)
            hi = np.full((nb ** 3, 3), -np.inf)
            for r, b in zip(pts, bid):
                for d in range(3):
                    lo[b, d] = min(lo[b, d], r[d])
                    hi[b, d] = max(hi[b, d], r[d])
            brute = (lo + hi) / 2
            assert np.allclose(oracle.block_centers(pts, bid, nb ** 3), brute, rtol=0, atol=1e-15)
            assert np.allclose(pkg.block_centers(pts, bid, nb ** 3), brute, rtol=0, atol=1e-15)
            assert np.allclose(pkg.yee_block_centers(n, h, comp, nb), brute, rtol=0, atol=1e-15)
            assert np.allclose(pkg.yee_points(n, h, comp), pts, rtol=0, atol=0)


def test_eps_cosine(oracle, pkg):
    """The cosine family eps = 1 + (a/6) sum_k cos(2 pi y_k + phi_k): its grid mean is exactly 1
    (each cosine sums to zero over a full period of n >= 2 samples), it is a sum of six 1-D
    profiles (separable), its value at a point is the closed form, its range lies in
    [1 - a, 1 + a], and the product's cosine_material gives the same samples and phases."""
    for N, amp, seed in ((1, 0.4, 0), (2, 0.8, 3), (3, 0.4, 0)):
        phases = np.random.default_rng(seed).uniform(0.0, 2.0 * np.pi, 6)
        e = oracle.eps_cosine(N, amp, phases)
        n = 2 * N
        assert e.shape == (n ** 6,)
        assert abs(e.mean() - 1.0) < 1e-14
        assert e.min() >= 1 - amp - 1e-15 and e.max() <= 1 + amp + 1e-15
        E = oracle.unvec(e, N)
        rng = np.random.default_rng(5)
        for _ in range(20):
            j = rng.integers(0, n, 6)
            val = 1.0 + amp / 6.0 * sum(math.cos(2 * math.pi * j[k] / n + phases[k]) for k in range(6))
            assert abs(E[tuple(j)] - val) < 1e-15
            # column-major index of the same point
            lin = sum(int(j[k]) * n ** k for k in range(6))
            assert abs(e[lin] - val) < 1e-15
        # separable: second differences across two axes vanish
        D = E - E.mean()
        mixed = D[1:, 1:] - D[1:, :-1] - D[:-1, 1:] + D[:-1, :-1] if n > 1 else D
        assert np.max(np.abs(mixed)) < 1e-14
        ep, ph = pkg.cosine_material(N, amp, seed)
        assert np.allclose(ph, phases, rtol=0, atol=0)
        assert np.max(np.abs(ep - e)) < 1e-15


@pytest.mark.parametrize("N,q", [(1, Q_EXP1), (2, (0.13, 0.21, 0.17)), (3, (0.5, 0.4, 0.3))])
def test_projected_q_modes_product_vs_oracle(oracle, pkg, N, q):
    """taylor.projected_q_modes (digit arithmetic) against the oracle's ndgrid form, and the
    closed form q_xi = P^T xi + q at sampled modes (k_lift = P (P^T P)^{-1} q, so
    P^T k_lift = q)."""
    kl = oracle.klift(q)
    qp = pkg.projected_q_modes(N, kl, oracle.P_REF)
    qo = oracle.projected_q_modes(N, q)
    assert qp.shape == qo.shape == ((2 * N) ** 6, 3)
    assert np.max(np.abs(qp - qo)) < 1e-13
    f = oracle.fourier_freq(N)
    n = 2 * N
    rng = np.random.default_rng(N)
    for lin in rng.integers(0, n ** 6, 25):
        digits = [(int(lin) // n ** k) % n for k in range(6)]
        xi = np.array([f[d] for d in digits])
        assert np.allclose(qp[lin], oracle.P_REF.T @ xi + np.asarray(q), rtol=0, atol=1e-13)


def test_auto_taylor_blocks_reference_rule(pkg):
    """auto_taylor_blocks follows build_auto_taylor_partition_exp4 (exp3/build_auto_taylor_partition_exp4.m:9-25):
    the smallest nb with qL1Max * (ceil(m/nb) - 1) h / 2 <= target (1 + slack), the bound it
    returns, the error when no nb up to the maximum qualifies; and it agrees with
    choose_partition_for_radius_exp2b's search (exp2/choose_partition_for_radius_exp2b.m:22-46)
    where both apply."""
    for m, h, ql1, target in ((83, 0.0125, 33.0, 5.2), (163, 0.0125, 33.0, 5.2), (403, 0.0125, 20.0, 2.0),
                              (11, 0.1, 3.0, 100.0), (83, 0.0125, 0.0, 1.0)):
        nb, rho = pkg.auto_taylor_blocks(m, h, ql1, target)
        brute = next(k for k in range(1, 33)
                     if ql1 * max(0.0, (math.ceil(m / k) - 1) * h / 2) <= target * (1 + 2e-12))
        assert nb == brute
        assert abs(rho - ql1 * max(0.0, (math.ceil(m / nb) - 1) * h / 2)) < 1e-15 * max(rho, 1)
        # exp2b's search over k (relTol 1e-12) lands on the same k
        k = 1
        while 0.5 * h * (math.ceil(m / k) - 1) * ql1 > target * (1 + 1e-12):
            k += 1
        assert k == nb
    with pytest.raises(ValueError):
        pkg.auto_taylor_blocks(403, 0.0125, 1e4, 0.1, max_blocks=8)


def test_multi_index_order_product_vs_oracle(oracle, pkg):
    for p in (0, 1, 4, 10):
        a = pkg.build_multi_index_list_3d(p)
        assert np.array_equal(a, oracle.multi_index_list(p))
        assert len(a) == math.comb(p + 3, 3)
        # a1 outermost, a3 innermost: lexicographic order of (a1, a2, a3)
        assert [tuple(r) for r in a] == sorted(tuple(r) for r in a)
        assert set(tuple(r) for r in a) == {t for t in itertools.product(range(p + 1), repeat=3) if sum(t) <= p}


def test_field_errors_oracle(oracle):
    """compute_field_errors: t = (1 + d) r gives e2 = eInf = |d|; a brute-force loop over
    every entry on a small field; a zero reference takes the eps floor of the denominators."""
    rng = np.random.default_rng(9)
    ref = [rng.standard_normal(50) + 1j * rng.standard_normal(50) for _ in range(3)]
    d = 3e-4 - 2e-4j
    out = oracle.compute_field_errors([(1 + d) * r for r in ref], ref)
    assert abs(out["e2"] - abs(d)) < 1e-15 and abs(out["eInf"] - abs(d)) < 1e-15
    test = [r + 1e-3 * (rng.standard_normal(50) + 1j * rng.standard_normal(50)) for r in ref]
    s_d = s_r = m_d = m_r = 0.0
    for t_c, r_c in zip(test, ref):
        for a, b in zip(t_c, r_c):
            s_d += abs(a - b) ** 2
            s_r += abs(b) ** 2
            m_d = max(m_d, abs(a - b))
            m_r = max(m_r, abs(b))
    out = oracle.compute_field_errors(test, ref)
    assert abs(out["e2"] - math.sqrt(s_d) / math.sqrt(s_r)) < 1e-15
    assert abs(out["eInf"] - m_d / m_r) < 1e-15
    z = [np.zeros(4, complex)] * 3
    out = oracle.compute_field_errors([np.full(4, 1e-20 + 0j)] * 3, z)
    assert abs(out["e2"] - math.sqrt(12) * 1e-20 / np.finfo(float).eps) < 1e-25
