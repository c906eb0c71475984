"""Pins the CPU oracle (no GPU): reference-held golden eigenvalues, the reference's own
smoke-test identities, dense brute force and closed forms."""
import math

import numpy as np
import pytest

GOLDEN_N2 = (1.201276300237e-02, 4.496454358674e-02)  # exp3/make_config_experiment4.m:40
Q_EXP4 = (0.13, 0.21, 0.17)  # exp3/make_config_experiment4.m:26
Q_EXP1 = (0.05, 0.04, 0.03)  # exp1/make_config_exp1bc.m:34


def test_fourier_order(oracle):
    assert list(oracle.fourier_freq(3)) == [0, 1, 2, -3, -2, -1]


def test_klift_closed_form(oracle):
    # P(P^T P)^{-1} q = (q1/3, q2/4, q3/6, sqrt2 q1/3, sqrt3 q2/4, sqrt5 q3/6)
    q = np.array(Q_EXP1)
    kl = oracle.klift(q)
    ref = np.array([q[0] / 3, q[1] / 4, q[2] / 6, math.sqrt(2) * q[0] / 3, math.sqrt(3) * q[1] / 4,
                    math.sqrt(5) * q[2] / 6])
    assert np.allclose(kl, ref, rtol=0, atol=1e-16)


@pytest.mark.parametrize("N", [1, 2])
def test_rank1_lattice_equals_circulant(oracle, N):
    # exp1/smoke_test_rank1_lowmem.m:13-23 requires < 1e-11
    eps = oracle.eps_reference(N)
    plan = oracle.rank1_plan(N)
    kf = oracle.rank1_kernel_fft(eps, plan)
    rng = np.random.default_rng(20260909)
    x = rng.standard_normal(plan["NF"]) + 1j * rng.standard_normal(plan["NF"])
    y1 = oracle.rank1_apply(x, kf, plan)
    y2 = oracle.mass_apply(eps, x, N)
    assert np.linalg.norm(y1 - y2) / np.linalg.norm(y2) < 1e-11


def test_rank1_plan_constants_n5(oracle):
    # exp2/run_experiment3_static_check.m:14-18 / exp3/make_config_experiment4.m:34-35
    n, B = 10, 19
    assert 2 * 5 * B ** 5 == 24760990
    assert [B ** j for j in range(6)] == [1, 19, 361, 6859, 130321, 2476099]


def test_mass_closed_forms(oracle):
    N = 2
    NF = (2 * N) ** 6
    rng = np.random.default_rng(3)
    x = rng.standard_normal(NF) + 1j * rng.standard_normal(NF)
    # constant eps -> c*I
    assert np.allclose(oracle.mass_apply(np.full(NF, 2.5), x, N), 2.5 * x, rtol=1e-14, atol=1e-13)
    # delta input gives the circulant column: M e_0 = fftn(eps/NF) (eps real)
    eps = oracle.eps_reference(N)
    e0 = np.zeros(NF, complex)
    e0[0] = 1
    col = oracle.mass_apply(eps, e0, N)
    ref = oracle.vec(np.fft.fftn(oracle.unvec(eps, N))) / NF
    assert np.allclose(col, ref, atol=1e-14)
    # Hermitian positive definite (run_experiment3_preflight_n5.m:17-24)
    y = oracle.mass_apply(eps, x, N)
    z = rng.standard_normal(NF) + 1j * rng.standard_normal(NF)
    assert abs(np.vdot(z, y) - np.vdot(oracle.mass_apply(eps, z, N), x)) < 1e-10 * np.linalg.norm(x) ** 2
    assert np.vdot(x, y).real > 0
    # inverse closed form
    assert np.allclose(oracle.mass_inverse_apply(eps, y, N), x, atol=1e-12)


def test_qsystem_closed_forms(oracle):
    N = 2
    s = oracle.qsystem(N, Q_EXP1)
    q = np.stack([s["q1"], s["q2"], s["q3"]])
    V = np.array([[s["V11"], s["V12"]], [s["V21"], s["V22"]], [s["V31"], s["V32"]]])  # 3x2xNF
    # V V^T = K = |q|^2 I - q q^T per mode; V^T q = 0; V^T V = |q|^2 I2
    K = np.einsum("ik,jk->ijk", q, q)
    Qq = s["Qq"]
    for i in range(3):
        for j in range(3):
            kij = (Qq if i == j else 0) - K[i, j]
            assert np.allclose(np.einsum("ak,ak->k", V[i], V[j]), kij, atol=1e-12)
    assert np.allclose(np.einsum("iak,ik->ak", V, q), 0, atol=1e-12)
    assert np.allclose(np.einsum("iak,ibk->abk", V, V)[0, 0], Qq, rtol=1e-12)
    # W = V1^{-*} [V31; V32]:  V1^T W = [V31; V32]
    assert np.allclose(s["V11"] * s["W1"] + s["V21"] * s["W2"], s["V31"], atol=1e-12)
    assert np.allclose(s["V12"] * s["W1"] + s["V22"] * s["W2"], s["V32"], atol=1e-12)
    assert np.allclose(s["invC"], 1 / np.sqrt(1 + s["W1"] ** 2 + s["W2"] ** 2))


def test_qsystem_singular_branch_closed_forms(oracle):
    """q = (c, c, c) puts the modes xi = (m, m, m, 0, 0, 0) on the (1,1,1) line, where the regular
    basis formula divides by 3|q|^2 - (q1+q2+q3)^2 = 0; the singular branch
    (exp1/assemble_K_and_nullspace_free_basis.m:75-91) takes a = e_i for the smallest |qhat_i|
    (the three tie up to rounding, so i is whichever rounding makes smallest),
    z1 = (a - qhat qhat^T a)/|.| = (3 e_i - (1,1,1))/sqrt6 and z2 = qhat x z1.
    Every mode, singular or not, keeps V V^T = K, V^T q = 0 and V^T V = |q|^2 I."""
    N, c = 2, 0.1
    s = oracle.qsystem(N, (c, c, c))
    assert s["nsing"] == 2 * N
    q = np.stack([s["q1"], s["q2"], s["q3"]])
    V = np.array([[s["V11"], s["V12"]], [s["V21"], s["V22"]], [s["V31"], s["V32"]]])
    Qq = s["Qq"]
    for i in range(3):
        for j in range(3):
            kij = (Qq if i == j else 0) - q[i] * q[j]
            assert np.allclose(np.einsum("ak,ak->k", V[i], V[j]), kij, atol=1e-12)
    assert np.allclose(np.einsum("iak,ik->ak", V, q), 0, atol=1e-12)
    VtV = np.einsum("iak,ibk->abk", V, V)
    assert np.allclose(VtV[0, 0], Qq, rtol=1e-12) and np.allclose(VtV[1, 1], Qq, rtol=1e-12)
    assert np.allclose(VtV[0, 1], 0, atol=1e-12)
    n = 2 * N
    f = oracle.fourier_freq(N)
    for m in range(-N, N):
        j = int(np.nonzero(f == m)[0][0])
        lin = j + n * j + n * n * j
        sq = math.sqrt(3.0) * abs(m + c)
        qv = q[:, lin]
        hq = qv / np.linalg.norm(qv)
        i = int(np.argmin(np.abs(hq)))
        z1 = (3.0 * np.eye(3)[i] - 1.0) / math.sqrt(6)
        assert np.allclose(V[:, 0, lin], sq * z1, atol=1e-14)
        assert np.allclose(V[:, 1, lin], sq * np.cross(np.sign(m + c) * np.ones(3) / math.sqrt(3), z1), atol=1e-14)
    assert np.all(np.abs(s["det"]) > 1e-12)


def test_qsystem_gamma_point_rejected(oracle):
    with pytest.raises(ValueError):
        oracle.qsystem(1, (0.0, 0.0, 0.0))


def test_krinv_inverts_kr(oracle):
    # K_r (K_r^{-1} q) = q with K_r = V^* M^{-1} V applied through the closed-form inverse mass
    N = 2
    eps = oracle.eps_reference(N)
    s = oracle.qsystem(N, Q_EXP1)
    rng = np.random.default_rng(5)
    q = rng.standard_normal(2 * s["NF"]) + 1j * rng.standard_normal(2 * s["NF"])
    y, st = oracle.krinv_apply(s, eps, q, N, 1e-12, 3000)
    back = oracle.kr_apply(s, eps, y, N)
    assert st["flag"] == 0
    assert np.linalg.norm(back - q) / np.linalg.norm(q) < 1e-9


def test_constant_eps_eigenvalues(oracle):
    # eps = c: K_r = Sigma / c -> lambda = |q_xi|^2 / c, each twice
    N = 1
    c = 3.0
    NF = (2 * N) ** 6
    eps = np.full(NF, c)
    v0 = np.random.default_rng(7).standard_normal(2 * NF) + 0j
    lam, Y, info, s = oracle.eigs_method_c(N, eps, Q_EXP1, v0, k=6, p=30, tol=1e-10, max_restart=50)
    ref = np.sort(np.concatenate([s["Qq"], s["Qq"]]))[:6] / c
    assert np.allclose(lam, ref, rtol=1e-10)


def test_lanczos_toy_operator(oracle):
    # exp1/smoke_test_rank1_lowmem.m:44-54: diag(linspace(100,1,80)), top 5 < 1e-7
    d = np.linspace(100, 1, 80)
    v0 = np.random.default_rng(20260910).standard_normal(80) + 1j * np.random.default_rng(1).standard_normal(80)
    theta, Y, info = oracle.lanczos_hybrid(lambda x: d * x, 80, 5, 1e-10, 40, 24, v0)
    exact = np.sort(d)[::-1][:5]
    assert info["converged"]
    assert np.linalg.norm(np.sort(theta)[::-1] - exact) / np.linalg.norm(exact) < 1e-7


def test_dense_bruteforce_n1(oracle):
    # smallest eigenvalues of the dense K_r at N=1 (128 x 128) vs the Method-C Lanczos
    N = 1
    eps = oracle.eps_reference(N)
    s = oracle.qsystem(N, Q_EXP4)
    n2 = 2 * s["NF"]
    Kr = np.stack([oracle.kr_apply(s, eps, np.eye(n2)[:, j].astype(complex), N) for j in range(n2)], 1)
    assert np.allclose(Kr, Kr.conj().T, atol=1e-12)
    ev = np.linalg.eigvalsh(Kr)[:10]
    v0 = np.random.default_rng(11).standard_normal(n2) + 1j * np.random.default_rng(12).standard_normal(n2)
    lam, Y, info, _ = oracle.eigs_method_c(N, eps, Q_EXP4, v0, k=10, p=40, tol=1e-10, max_restart=50)
    assert np.allclose(lam, ev, rtol=1e-9)


def test_golden_eigenvalues_n2(oracle):
    """Reference-held golden values (exp3/make_config_experiment4.m:38-45), reference check 5e-7."""
    N = 2
    eps = oracle.eps_reference(N)
    v0 = np.random.default_rng(20260831)
    n = 2 * (2 * N) ** 6
    v0 = v0.standard_normal(n) + 1j * v0.standard_normal(n)
    lam, Y, info, s = oracle.eigs_method_c(N, eps, Q_EXP4, v0, k=10, p=40, tol=1e-10, max_restart=30,
                                           tol_cg=1e-10, maxit_cg=3000)
    assert info["converged"]
    assert abs(lam[0] - GOLDEN_N2[0]) / GOLDEN_N2[0] < 1e-10
    assert abs(lam[9] - GOLDEN_N2[1]) / GOLDEN_N2[1] < 1e-10
    # recovery + GEVP residual.  The residual of the recovered pair is bounded by
    # tol_cg * ||K_r|| / lambda (the inexact K_r^{-1} enters the Krylov relation), about
    # 2e-7 here; the reference only warns above 1e-10 (make_config_experiment2_allgpu.m:51).
    X, MX, its = oracle.recover_eigenvectors(s, eps, Y[:, :2], N)
    res, div, ray = oracle.fourier_diagnostics(s, X, MX, lam[:2])
    kr_norm = s["Qq"].max() / eps.min()
    assert res.max() < 10 * 1e-10 * kr_norm / lam[0]
    assert div.max() < 1e-9
    assert np.allclose(ray, lam[:2], rtol=1e-9)


def test_taylor_single_mode_and_direct(oracle):
    # one Fourier mode: u(r) = c exp(i q.r); the Taylor series converges to it as p grows
    rng = np.random.default_rng(2)
    q = rng.uniform(-2, 2, (1, 3))
    c = np.array([0.7 - 0.2j])
    pts = oracle.yee_points(2, 0.05, 1)
    bid = oracle.block_partition(7, 2)
    cen = oracle.block_centers(pts, bid, 8)
    exact = c[0] * np.exp(1j * pts @ q[0])
    assert np.allclose(oracle.direct_reconstruct(q, c, pts), exact, atol=1e-14)
    errs = []
    for p in (2, 4, 8, 12):
        mom = oracle.taylor_moments(q, c, p, cen)
        u = oracle.taylor_evaluate(pts, bid, cen, mom, p)
        errs.append(np.abs(u - exact).max())
    assert errs[-1] < 1e-13 and errs[0] > errs[1] > errs[2]


def test_taylor_moments_bruteforce(oracle):
    rng = np.random.default_rng(4)
    NF = 50
    q = rng.uniform(-1, 1, (NF, 3))
    c = rng.standard_normal(NF) + 1j * rng.standard_normal(NF)
    cen = rng.uniform(-0.3, 0.3, (3, 3))
    p = 3
    mom = oracle.taylor_moments(q, c, p, cen)
    al = oracle.multi_index_list(p)
    assert len(al) == math.comb(p + 3, 3)
    for ia, a in enumerate(al):
        for ic in range(3):
            s = sum(c[j] * np.exp(1j * q[j] @ cen[ic]) * np.prod(q[j] ** a) for j in range(NF))
            ref = (1j) ** a.sum() / np.prod([math.factorial(v) for v in a]) * s
            assert abs(mom[ia, ic] - ref) < 1e-12 * max(1, abs(ref))


def test_multi_index_count(oracle):
    assert len(oracle.multi_index_list(10)) == 286


def test_mass_adjoint_fd(oracle):
    # Re<gy, M(eps) x> is linear in eps: central differences are exact up to rounding
    N = 1
    NF = 64
    rng = np.random.default_rng(9)
    eps = oracle.eps_reference(N)
    x = rng.standard_normal(NF) + 1j * rng.standard_normal(NF)
    gy = rng.standard_normal(NF) + 1j * rng.standard_normal(NF)
    g = oracle.mass_adjoint_eps(eps, x, gy, N)
    for _ in range(5):
        d = rng.standard_normal(NF)
        h = 1e-3
        f = lambda e: np.real(np.vdot(gy, oracle.mass_apply(e, x, N)))
        fd = (f(eps + h * d) - f(eps - h * d)) / (2 * h)
        assert abs(fd - g @ d) < 1e-10 * (1 + abs(fd))


def test_hellmann_feynman_fd_n1(oracle):
    # d lambda/d eps from the Hellmann-Feynman formula vs central FD of a dense solve at N=1
    N = 1
    eps = oracle.eps_reference(N)
    s = oracle.qsystem(N, Q_EXP4)
    n2 = 2 * s["NF"]

    def lam_and_vec(e):
        Kr = np.stack([oracle.kr_apply(s, e, np.eye(n2)[:, j].astype(complex), N) for j in range(n2)], 1)
        w, U = np.linalg.eigh(Kr)
        return w, U

    w, U = lam_and_vec(eps)
    # use a simple (non-degenerate) eigenvalue
    j = int(np.argmax(np.diff(w[:10]) > 1e-6 * w[1]) + 1) if False else 2
    gaps = np.diff(w)
    j = next(i for i in range(1, 10) if gaps[i - 1] > 1e-4 * w[i] and gaps[i] > 1e-4 * w[i])
    X, MX, _ = oracle.recover_eigenvectors(s, eps, U[:, [j]], N)
    g = oracle.eig_grad_eps(X, [w[j]], N)[:, 0]
    rng = np.random.default_rng(13)
    for _ in range(3):
        d = rng.standard_normal(len(eps))
        h = 1e-6
        fd = (lam_and_vec(eps + h * d)[0][j] - lam_and_vec(eps - h * d)[0][j]) / (2 * h)
        assert abs(fd - g @ d) < 1e-6 * abs(fd) + 1e-12


def test_method_b_matches_method_c(oracle):
    # both realise K_r^{-1}: the nested solve and the explicit Method-C inverse agree
    N = 1
    eps = oracle.eps_reference(N)
    s = oracle.qsystem(N, Q_EXP4)
    rng = np.random.default_rng(31)
    b = rng.standard_normal(2 * s["NF"]) + 1j * rng.standard_normal(2 * s["NF"])
    xb, stb = oracle.krinv_nested(s, eps, b, N, 1e-12, 1e-13, 600, 1600, 4.0)
    xc, stc = oracle.krinv_apply(s, eps, b, N, 1e-12, 3000)
    assert np.linalg.norm(xb - xc) / np.linalg.norm(xc) < 1e-9
    assert np.linalg.norm(oracle.kr_apply(s, eps, xb, N) - b) / np.linalg.norm(b) < 1e-10


def test_shift_preconditioner_closed_form(oracle):
    # P_A^{-1} = (K + tau I)^{-1} per mode (build_original_shift_preconditioner_A_only.m:5-12)
    N = 1
    s = oracle.qsystem(N, Q_EXP4)
    n = s["NF"]
    tau = 3e-4
    rng = np.random.default_rng(41)
    x = rng.standard_normal(3 * n) + 1j * rng.standard_normal(3 * n)
    y = oracle.shift_precond_apply(s, tau, x)
    for i in (0, 7, n - 1):
        q = np.array([s["q1"][i], s["q2"][i], s["q3"][i]])
        Kq = (q @ q) * np.eye(3) - np.outer(q, q)
        ref = np.linalg.solve(Kq + tau * np.eye(3), x[[i, n + i, 2 * n + i]])
        assert np.allclose(y[[i, n + i, 2 * n + i]], ref, rtol=1e-9, atol=0)  # cond(K_q + tau I) ~ 1e4


def test_method_a_minres_and_dense_lambda_n1(oracle):
    # T = (K - sigma M)^{-1} M at N = 1 (192 unknowns): MINRES meets the true residual, and
    # sigma + 1/theta of the 10 largest theta are the 10 smallest eigenvalues of the dense
    # K_r of Method C (the reference's cross-method check, exp1/check_fig1bc_results.m:23-25)
    N = 1
    eps = oracle.eps_reference(N)
    s = oracle.qsystem(N, Q_EXP4)
    sigma = 1e-4
    rng = np.random.default_rng(43)
    x = rng.standard_normal(3 * s["NF"]) + 1j * rng.standard_normal(3 * s["NF"])
    y, its = oracle.method_a_top(s, eps, x, N, sigma, 1e-10, 4000)
    b = oracle.mass3_apply(eps, x, N)
    r = oracle.apply_K(s, y) - sigma * oracle.mass3_apply(eps, y, N) - b
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10 and its > 0
    lamA = oracle.method_a_dense_lambda(s, eps, N, sigma, 10)
    n2 = 2 * s["NF"]
    Kr = np.stack([oracle.kr_apply(s, eps, np.eye(n2)[:, j].astype(complex), N) for j in range(n2)], 1)
    lamC = np.linalg.eigvalsh(Kr)[:10]
    assert np.allclose(lamA, lamC, rtol=1e-9)


# ----------------------------------------------------------------------------- Yee closure
def _yee_random_fields(rng, n):
    m = 2 * n + 3
    return [rng.standard_normal((m, m, m)) + 1j * rng.standard_normal((m, m, m)) for _ in range(3)]


def _curl_transpose_curl(u1, u2, u3, h):
    """Independent form of the cropped operator: face curls w = C u of the plus-grid field,
    then C^T w on the interior edges (u1 at (i+1/2, j, k), u2 at (i, j+1/2, k), u3 at
    (i, j, k+1/2); w1 at (i, j+1/2, k+1/2), w2 at (i+1/2, j, k+1/2), w3 at (i+1/2, j+1/2, k),
    each stored at its lower corner index)."""
    m = u1.shape[0]
    W1 = np.zeros_like(u1)
    W2 = np.zeros_like(u1)
    W3 = np.zeros_like(u1)
    W1[:, :-1, :-1] = (u3[:, 1:, :-1] - u3[:, :-1, :-1]) - (u2[:, :-1, 1:] - u2[:, :-1, :-1])
    W2[:-1, :, :-1] = (u1[:-1, :, 1:] - u1[:-1, :, :-1]) - (u3[1:, :, :-1] - u3[:-1, :, :-1])
    W3[:-1, :-1, :] = (u2[1:, :-1, :] - u2[:-1, :-1, :]) - (u1[:-1, 1:, :] - u1[:-1, :-1, :])

    def S(W, di, dj, dk):
        return W[1 + di:m - 1 + di, 1 + dj:m - 1 + dj, 1 + dk:m - 1 + dk]

    A1 = -S(W2, 0, 0, 0) + S(W2, 0, 0, -1) + S(W3, 0, 0, 0) - S(W3, 0, -1, 0)
    A2 = S(W1, 0, 0, 0) - S(W1, 0, 0, -1) - S(W3, 0, 0, 0) + S(W3, -1, 0, 0)
    A3 = -S(W1, 0, 0, 0) + S(W1, 0, -1, 0) + S(W2, 0, 0, 0) - S(W2, -1, 0, 0)
    return A1 / h ** 2, A2 / h ** 2, A3 / h ** 2


def test_cropped_A_is_yee_curl_transpose_curl(oracle):
    rng = np.random.default_rng(41)
    n, h = 3, 0.1
    u = _yee_random_fields(rng, n)
    A = oracle.apply_cropped_A(*u, h)
    B = _curl_transpose_curl(*u, h)
    for a, b in zip(A, B):
        assert a.shape == (2 * n + 1,) * 3
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def test_cropped_A_annihilates_discrete_gradients(oracle):
    rng = np.random.default_rng(42)
    n, h = 4, 0.05
    m = 2 * n + 3
    phi = rng.standard_normal((m + 1,) * 3) + 1j * rng.standard_normal((m + 1,) * 3)
    u1 = (phi[1:, :m, :m] - phi[:m, :m, :m]) / h
    u2 = (phi[:m, 1:, :m] - phi[:m, :m, :m]) / h
    u3 = (phi[:m, :m, 1:] - phi[:m, :m, :m]) / h
    scale = max(np.max(np.abs(x)) for x in (u1, u2, u3)) / h ** 2
    for a in oracle.apply_cropped_A(u1, u2, u3, h):
        assert np.max(np.abs(a)) <= 1e-13 * scale


def _plane_wave(n, h, kvec, p):
    """Staggered plane wave u_c = p_c exp(i k . x_c) on the plus grid and its discrete symbol
    sigma_d = (2/h) sin(k_d h/2)."""
    I = np.arange(-(n + 1), n + 2) * h
    out = []
    for c in range(3):
        ax = [I + (h / 2 if d == c else 0.0) for d in range(3)]
        ph = kvec[0] * ax[0][:, None, None] + kvec[1] * ax[1][None, :, None] + kvec[2] * ax[2][None, None, :]
        out.append(p[c] * np.exp(1j * ph))
    sigma = 2.0 / h * np.sin(np.asarray(kvec) * h / 2)
    return out, sigma


def test_plane_wave_is_exact_yee_eigenpair(oracle):
    """A transverse staggered plane wave in constant eps: A u = |sigma|^2 u exactly, so the
    metrics reduce to lambda = |sigma|^2/epsC, rho = lambda, V = 0 and zero residuals."""
    n, h = 5, 0.1
    kvec = np.array([1.3, -0.7, 2.1])
    sigma = 2.0 / h * np.sin(kvec * h / 2)
    p = np.cross(sigma, [0.3, 1.0, -0.4]) * (1.0 + 0.5j)
    u, sigma = _plane_wave(n, h, kvec, p)
    lam = float(sigma @ sigma) / 4.0
    A = oracle.apply_cropped_A(*u, h)
    for a, x in zip(A, u):
        xi = x[1:-1, 1:-1, 1:-1]
        assert np.max(np.abs(a - 4.0 * lam * xi)) <= 1e-11 * 4.0 * lam * np.max(np.abs(xi))
    r = oracle.compute_yee_lwrq_metrics(*u, lam, h, n, epsC=4.0, alpha=0.0, return_rho=True)
    assert abs(r["lambda_3d"] - lam) <= 1e-12 * lam
    assert r["eta_3d"] < 1e-12 and r["eta_6d"] < 1e-12 and r["delta_lambda"] < 1e-12
    assert r["V"] < 1e-24 and r["delta_identity"] < 1e-13 and r["delta_part"] < 1e-13
    assert np.max(np.abs(r["rho"] - lam)) <= 1e-12 * lam


def test_separable_convolution_is_3d_convolution(oracle):
    from scipy import ndimage
    rng = np.random.default_rng(43)
    X = rng.standard_normal((9, 11, 10))
    k = oracle.lwrq_kernel(0.05, 3, 2.0)
    K3 = k[:, None, None] * k[None, :, None] * k[None, None, :]
    ref = ndimage.convolve(X, K3, mode="constant", cval=0.0)
    assert np.max(np.abs(oracle.separable_conv3_same(X, k) - ref)) <= 1e-13 * np.max(np.abs(ref))
    # den = conv(ones) is the product of one profile per axis
    s = np.array([sum(k[t + 3] for t in range(-3, 4) if 0 <= i - t < 13) for i in range(13)])
    den = oracle.separable_conv3_same(np.ones((13, 13, 13)), k)
    assert np.max(np.abs(den - s[:, None, None] * s[None, :, None] * s[None, None, :])) < 1e-14 * den.max()


def test_lwrq_constant_quotient_and_partition(oracle):
    rng = np.random.default_rng(44)
    b = 1.0 + rng.random((9, 9, 9))
    r = oracle.compute_lwrq(2.5 * b, b, 0.1, return_rho=True)
    assert abs(r["lambda_"] - 2.5) < 1e-14 and r["V"] < 1e-28
    assert np.max(np.abs(r["rho"] - 2.5)) < 1e-13
    a = rng.standard_normal((9, 9, 9))
    r = oracle.compute_lwrq(a, b, 0.1)
    assert r["partition_error"] < 1e-13  # column normalisation preserves both sums
    with pytest.raises(ValueError):
        oracle.compute_lwrq(a, np.zeros_like(b), 0.1)


def test_yee_material_closed_forms(oracle):
    n, h = 3, 0.2
    e = oracle.sample_scalar_yee_material(oracle.P_REF, 4.0, 0.0, h, n)
    assert all(np.all(x == 4.0) for x in e)
    e = oracle.sample_scalar_yee_material(oracle.P_REF, 4.0, 1.2, h, n)
    for comp, (i, j, k) in [(0, (0, 3, 6)), (1, (5, 1, 2)), (2, (6, 6, 0))]:
        x = [(i - n) * h, (j - n) * h, (k - n) * h]
        x[comp] += h / 2
        S = sum(math.cos(sum(oracle.P_REF[l, d] * x[d] for d in range(3))) for l in range(6))
        assert abs(e[comp][i, j, k] - 4.0 * math.exp(0.2 * S)) < 1e-14 * 4.0 * math.exp(0.2 * S)


def test_crop_plus_field(oracle):
    n_large, n_small = 6, 2
    m = 2 * n_large + 3
    I = np.arange(-(n_large + 1), n_large + 2)
    u = I[:, None, None] * 10000 + I[None, :, None] * 100 + I[None, None, :] + 0.0
    c = oracle.crop_plus_field(u, n_large, n_small)
    J = np.arange(-(n_small + 1), n_small + 2)
    assert np.array_equal(c, J[:, None, None] * 10000 + J[None, :, None] * 100 + J[None, None, :])
    assert m == u.shape[0]
