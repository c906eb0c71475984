"""GPU parity of the q-system, Mhat, K_r^{-1}, Lanczos, recovery and batching."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN_N2 = (1.201276300237e-02, 4.496454358674e-02)  # exp3/make_config_experiment4.m:40
Q_EXP4 = (0.13, 0.21, 0.17)
Q_EXP1 = (0.05, 0.04, 0.03)


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.fixture(scope="module")
def n2(pkg, oracle):
    torch.cuda.set_device(0)
    N = 2
    eps = oracle.eps_reference(N)
    m = pkg.MassBackend(N, eps)
    s = pkg.QSystem(m, Q_EXP4)
    yield dict(N=N, eps=eps, m=m, s=s, osys=oracle.qsystem(N, Q_EXP4))
    s.close()
    m.close()


def test_qsystem_diagonals(n2):
    print("case: qsys diagonals N=2", flush=True)
    d = n2["s"].diagonals()
    o = n2["osys"]
    for k in ("V11", "V12", "V21", "V22", "V31", "V32", "W1", "W2", "invC"):
        assert np.allclose(d[k], o[k], rtol=1e-13, atol=1e-13), k
    assert np.allclose(n2["s"].info["kLift"], o["klift"], atol=1e-16)


@pytest.mark.parametrize("N,q", [(3, Q_EXP1), (4, Q_EXP4), (5, Q_EXP1)])
def test_qsystem_device_build(pkg, oracle, N, q):
    """The q-system diagonals formed on the device (one thread per mode) against the oracle's
    per-mode formulas at every mode, and the minima the handle reports (min |q|^2, min
    |det V1|) against the oracle's arrays."""
    print("case: device q-system build N=%d" % N, flush=True)
    torch.cuda.set_device(0)
    eps = oracle.eps_reference(N)
    m = pkg.MassBackend(N, eps)
    s = pkg.QSystem(m, q)
    try:
        o = oracle.qsystem(N, q)
        d = s.diagonals()
        for k in ("V11", "V12", "V21", "V22", "V31", "V32", "W1", "W2", "invC"):
            assert np.allclose(d[k], o[k], rtol=1e-13, atol=1e-13), k
        info = s.info
        assert info["minQq"] == pytest.approx(o["Qq"].min(), rel=1e-14)
        det = np.abs(o["V11"] * o["V22"] - o["V12"] * o["V21"])
        assert info["minAbsDetV1"] == pytest.approx(det.min(), rel=1e-12)
        assert info["nSingular"] == o["nsing"] == 0
    finally:
        s.close()
        m.close()


def test_batched_two_points_n5(pkg, n5):
    """Two Bloch points solved together at N = 5 (the several-group tile choice, tuned here,
    and the batched Lanczos steps) against each point's own solve, FP64 to 1e-10."""
    print("case: batched two points N=5", flush=True)
    m = n5["m"]
    m.autotune(reps=2)
    kw = dict(k=6, p=20, tol=1e-10, max_restart=60, tol_cg=1e-10, maxit_cg=1200)
    qs = np.array([Q_EXP1, Q_EXP4])
    v2 = np.stack([pkg.start_vector(2 * m.NF, 31), pkg.start_vector(2 * m.NF, 32)])
    try:
        out = pkg.run_method_C_batch(m, qs, v0=torch.as_tensor(v2, device="cuda"), **kw)
        for i, q in enumerate(qs):
            s_i = pkg.QSystem(m, q)
            try:
                lam = pkg.run_method_C(s_i, v0=v2[i], **kw)["lambda_"]
            finally:
                s_i.close()
            assert np.max(np.abs(out["lambda_"][i] - lam) / lam) < 1e-10, i
    finally:
        m.set_tiles("double", 0, 0)
        m.set_tiles("single", 0, 0)


def test_start_vector_on_device(pkg, n2):
    """The eigensolvers take the start vector from host or device memory: the same vector from
    a CUDA tensor gives the same eigenvalues, bit for bit (single run and batched path)."""
    print("case: device start vector N=2", flush=True)
    v0 = pkg.start_vector(n2["s"].nr, 9)
    kw = dict(k=6, p=30, tol=1e-10, max_restart=40, tol_cg=1e-10, maxit_cg=3000)
    lh = pkg.run_method_C(n2["s"], v0=v0, **kw)["lambda_"]
    ld = pkg.run_method_C(n2["s"], v0=torch.as_tensor(v0, device="cuda"), **kw)["lambda_"]
    assert np.array_equal(lh, ld)
    qs = np.array([Q_EXP4, Q_EXP1])
    v2 = np.stack([v0, pkg.start_vector(n2["s"].nr, 10)])
    bh = pkg.run_method_C_batch(n2["m"], qs, v0=v2, **kw)["lambda_"]
    bd = pkg.run_method_C_batch(n2["m"], qs, v0=torch.as_tensor(v2, device="cuda"), **kw)["lambda_"]
    assert np.array_equal(bh, bd)
    # the batched launches may use another tile choice than the single run: equal to 1e-10
    assert np.max(np.abs(bh[0] - lh) / lh) < 1e-10


@pytest.mark.parametrize("v3", [0, 1])
def test_mhat(n2, oracle, v3):
    print("case: Mhat N=2 variant %d" % v3, flush=True)
    rng = np.random.default_rng(1)
    x = rand_c(rng, n2["m"].NF)
    ref = oracle.mhat_apply(n2["osys"], n2["eps"], x, n2["N"])
    n2["m"].set_tiles("double", 36 * v3, 0)
    n2["m"].set_tiles("single", 36 * v3, 1)
    y = n2["s"].Mhat(torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13
    y32 = n2["s"].Mhat(torch.as_tensor(x, device="cuda").to(torch.complex64), "single").cpu().numpy()
    assert np.linalg.norm(y32 - ref) / np.linalg.norm(ref) < 5e-6
    n2["m"].set_tiles("double", 0, 0)
    n2["m"].set_tiles("single", 0, 0)


@pytest.fixture(scope="module")
def n5(pkg, oracle):
    torch.cuda.set_device(0)
    N = 5
    eps = oracle.eps_reference(N)
    m = pkg.MassBackend(N, eps)
    s = pkg.QSystem(m, Q_EXP1)
    yield dict(N=N, eps=eps, m=m, s=s)
    s.close()
    m.close()


# Mhat pass A and C variants at N=5, where a persistent workgroup walks several tiles:
# codes 36*v3 + 8*c3p + 64*a3p + 256*a3s (v3: blocks per workgroup of passes A and C; c3p,
# a3p: 0 one workgroup per tile, k > 0 persistent with k workgroups per compute unit; a3s:
# the persistent pass A shares its input through LDS).
C3P_CODES = (0, 36, 8, 44, 16 + 64, 52 + 64, 24 + 128, 60 + 128, 192, 36 + 192, 256 + 64, 256 + 36 + 128)


def test_mhat_pass_c_variants_n5(n5, oracle):
    print("case: Mhat N=5 pass-C variants", flush=True)
    rng = np.random.default_rng(5)
    x = rand_c(rng, n5["m"].NF)
    ref = oracle.mhat_apply(oracle.qsystem(5, Q_EXP1), n5["eps"], x, 5)
    xd = torch.as_tensor(x, device="cuda")
    try:
        for code in C3P_CODES:
            n5["m"].set_tiles("double", code, 0)
            n5["m"].set_tiles("single", code, 0)
            y = n5["s"].Mhat(xd).cpu().numpy()
            assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13, code
            y32 = n5["s"].Mhat(xd.to(torch.complex64), "single").cpu().numpy()
            assert np.linalg.norm(y32 - ref) / np.linalg.norm(ref) < 5e-6, code
    finally:
        n5["m"].set_tiles("double", 0, 0)
        n5["m"].set_tiles("single", 0, 0)


@pytest.mark.parametrize("N", [5, 6])
def test_mhat_fp32_pass_b_two_positions(pkg, oracle, N):
    """FP32 pass B with two inner positions per thread (16-byte accesses; pass-B variant 1 in
    FP32 at N = 5 and 6) against the oracle: the three-vector Mhat apply (variant code vb = 1)
    and the single-vector scalar apply (vb = 2), and the mixed K_r^-1 of the same system
    against the FP64 one."""
    print("case: FP32 pass B two positions per thread N=%d" % N, flush=True)
    torch.cuda.set_device(0)
    rng = np.random.default_rng(50 + N)
    eps = oracle.eps_reference(N)
    m = pkg.MassBackend(N, eps)
    s = pkg.QSystem(m, Q_EXP1)
    try:
        x = rand_c(rng, m.NF)
        xd = torch.as_tensor(x, device="cuda")
        ref = oracle.mhat_apply(oracle.qsystem(N, Q_EXP1), eps, x, N)
        for vb in (1, 2, 3):
            m.set_tiles("single", 0, vb)
            y32 = s.Mhat(xd.to(torch.complex64), "single").cpu().numpy()
            assert np.linalg.norm(y32 - ref) / np.linalg.norm(ref) < 5e-6, vb
            ys = m.MscalarOp(xd.to(torch.complex64)).cpu().numpy()
            rs = oracle.mass_apply(eps, x, N)
            assert np.linalg.norm(ys - rs) / np.linalg.norm(rs) < 2e-6, vb
        # mixed K_r^-1 with either FP32 pass B against FP64: both solve Mhat u = bhat to 1e-10
        # relative residual, so they differ from the FP64 solution by the same order (the
        # reference material at N = 6 has a permittivity contrast of e^2.4)
        q = torch.as_tensor(rand_c(rng, 2 * m.NF), device="cuda")
        y64, st64 = s.krinv(q, 1e-10, 1200, "double")
        assert st64["last_flag"] == 0
        errs = []
        for vb in (0, 1):
            m.set_tiles("single", 0, vb)
            ymx, stmx = s.krinv(q, 1e-10, 1200, "mixed")
            assert stmx["last_flag"] == 0, vb
            errs.append((torch.linalg.norm(ymx - y64) / torch.linalg.norm(y64)).item())
        print("  mixed vs FP64 K_r^-1, FP32 pass-B variants 0 / 1: %.2e / %.2e" % tuple(errs), flush=True)
        assert max(errs) < 2e-7, errs
        assert errs[1] < 10 * errs[0] + 1e-9, errs
    finally:
        s.close()
        m.close()


def test_krinv_pass_c_variants_n5(n5):
    """The CG form of the persistent pass C (alpha from the pass-B partials, x and r updated
    over several tiles per workgroup, ||r||^2 partials per workgroup) against the
    one-workgroup-per-tile kernel, FP64 and mixed."""
    print("case: krinv N=5 pass-C variants", flush=True)
    rng = np.random.default_rng(6)
    q = torch.as_tensor(rand_c(rng, 2 * n5["m"].NF), device="cuda")
    try:
        n5["m"].set_tiles("double", 36, 0)
        n5["m"].set_tiles("single", 36, 0)
        y0, st0 = n5["s"].krinv(q, 1e-10, 1200)
        # the mixed solve meets the same residual tolerance; its solution differs from the
        # FP64 one by about cond * tol
        ym0, stm0 = n5["s"].krinv(q, 1e-10, 1200, "mixed")
        assert (torch.linalg.norm(ym0 - y0) / torch.linalg.norm(y0)).item() < 1e-8
        for code in C3P_CODES[2:]:
            n5["m"].set_tiles("double", code, 0)
            n5["m"].set_tiles("single", code, 0)
            y, st = n5["s"].krinv(q, 1e-10, 1200)
            assert st["last_flag"] == 0 and abs(st["last_iterations"] - st0["last_iterations"]) <= 1, code
            assert (torch.linalg.norm(y - y0) / torch.linalg.norm(y0)).item() < 1e-9, code
            ym, stm = n5["s"].krinv(q, 1e-10, 1200, "mixed")
            assert (torch.linalg.norm(ym - y0) / torch.linalg.norm(y0)).item() < 1e-8, code
    finally:
        n5["m"].set_tiles("double", 0, 0)
        n5["m"].set_tiles("single", 0, 0)


@pytest.fixture(scope="module")
def n4(pkg, oracle):
    torch.cuda.set_device(0)
    N = 4
    eps = oracle.eps_reference(N)
    m = pkg.MassBackend(N, eps)
    s = pkg.QSystem(m, Q_EXP1)
    yield dict(N=N, eps=eps, m=m, s=s)
    s.close()
    m.close()


def test_krinv_pass_a_fold_small_workgroups_n4(n4):
    """From the second iteration of a chunk on, pass A applies the previous pcg step, summing
    the pass-C ||r||^2 partials in k_cg_beta's order.  At N=4 with one block per workgroup
    pass A has 192 threads and pass C writes 512 partials, more than one per thread: the
    iteration counts and solutions agree with those of the 384-thread forms."""
    print("case: krinv N=4 pass-A fold with 192-thread workgroups", flush=True)
    rng = np.random.default_rng(14)
    m = n4["m"]
    q = torch.as_tensor(rand_c(rng, 2 * m.NF), device="cuda")
    try:
        m.set_tiles("double", 0, 0)
        y0, st0 = n4["s"].krinv(q, 1e-10, 1200)
        assert st0["last_flag"] == 0
        # 36: one block per workgroup in passes A and C (512 pass-C partials); + 64: persistent
        # pass A; + 256: persistent pass A sharing its input through LDS; + 8: persistent pass C
        for code in (36, 36 + 64, 36 + 256 + 64, 36 + 8, 4):
            m.set_tiles("double", code, 0)
            y, st = n4["s"].krinv(q, 1e-10, 1200)
            assert st["last_flag"] == 0 and abs(st["last_iterations"] - st0["last_iterations"]) <= 1, code
            assert (torch.linalg.norm(y - y0) / torch.linalg.norm(y0)).item() < 1e-9, code
    finally:
        m.set_tiles("double", 0, 0)


def test_krinv_fp64_and_mixed(n2, oracle):
    print("case: krinv N=2", flush=True)
    rng = np.random.default_rng(2)
    q = rand_c(rng, 2 * n2["m"].NF)
    ref, st_ref = oracle.krinv_apply(n2["osys"], n2["eps"], q, n2["N"], 1e-10, 3000)
    y, st = n2["s"].krinv(torch.as_tensor(q, device="cuda"), 1e-10, 3000)
    y = y.cpu().numpy()
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-9
    assert st["last_flag"] == 0 and abs(st["last_iterations"] - st_ref["iter"]) <= 2
    for precond in (True, False):
        n2["s"].set_mixed(3e-5, 30, precond)
        ym, stm = n2["s"].krinv(torch.as_tensor(q, device="cuda"), 1e-10, 3000, "mixed")
        ym = ym.cpu().numpy()
        assert np.linalg.norm(ym - ref) / np.linalg.norm(ref) < 1e-9, precond
        assert stm["refinement_cycles"] >= 1
    n2["s"].set_mixed(3e-5, 30, True)


@pytest.fixture(scope="module")
def n2_eigs(pkg, n2):
    """The ten lowest eigenpairs of the reference's N = 2 configuration (its own solve, so
    every test that uses them runs alone or in any order)."""
    v0 = pkg.start_vector(n2["s"].nr, 20260831)
    out = pkg.run_method_C(n2["s"], k=10, p=40, tol=1e-10, max_restart=30, tol_cg=1e-10, maxit_cg=3000,
                           v0=v0, want_vectors=True)
    return out


def test_eigs_golden_n2(n2_eigs):
    print("case: eigs golden N=2 fp64", flush=True)
    out = n2_eigs
    lam = out["lambda_"]
    assert out["stats"]["converged"]
    assert abs(lam[0] - GOLDEN_N2[0]) / GOLDEN_N2[0] < 1e-10
    assert abs(lam[9] - GOLDEN_N2[1]) / GOLDEN_N2[1] < 1e-10
    # Ritz vectors of an orthonormal basis: one Gram-Schmidt sweep per step (plus the
    # DGKS-triggered second ones) keeps Q^H Q = I to rounding
    Y = out["Y"].cpu().numpy()
    assert np.max(np.abs(Y.conj() @ Y.T - np.eye(10))) < 1e-12
    assert 0 <= out["stats"]["second_sweeps"] <= out["stats"]["outer_steps"]


def test_eigs_every_step_second_sweep(pkg, n2, n2_eigs):
    """Re-orthogonalisation threshold eta = 1 (exp1/Lanczos_hybrid_identity_gpuop.m:55-63 runs
    two sweeps at every step): the DGKS test asks for the second sweep at every step away from
    a restart, which the single-run loop issues from the host after reading the step and then
    forms the step's next vector again; the eigenvalues and the basis equal the default run's."""
    print("case: eigs N=2 second sweep at every step", flush=True)
    s = n2["s"]
    try:
        s.set_reorth(1.0)
        out = pkg.run_method_C(s, k=10, p=40, tol=1e-10, max_restart=30, tol_cg=1e-10, maxit_cg=3000,
                               v0=pkg.start_vector(s.nr, 20260831), want_vectors=True)
    finally:
        s.set_reorth()
    st = out["stats"]
    assert st["converged"] and st["second_sweeps"] >= st["outer_steps"] - st["restarts"] - 1 > 0
    ref = n2_eigs["lambda_"]
    assert np.max(np.abs(out["lambda_"] - ref) / ref) < 1e-12
    Y = out["Y"].cpu().numpy()
    assert np.max(np.abs(Y.conj() @ Y.T - np.eye(10))) < 1e-12
    # the mixed form steps through the general loop, where the host also issues the second
    # sweep after the step's read: with eta = 1 at every step, with the default only when asked
    print("case: eigs N=2 mixed, second sweep at every step and by the DGKS test", flush=True)
    v0 = pkg.start_vector(s.nr, 20260831)
    outm = pkg.run_method_C(s, k=10, p=40, tol=1e-10, max_restart=30, tol_cg=1e-10, maxit_cg=3000, v0=v0,
                            precision="mixed", want_vectors=True)
    try:
        s.set_reorth(1.0)
        outm1 = pkg.run_method_C(s, k=10, p=40, tol=1e-10, max_restart=30, tol_cg=1e-10, maxit_cg=3000, v0=v0,
                                 precision="mixed", want_vectors=True)
    finally:
        s.set_reorth()
    st1 = outm1["stats"]
    assert st1["converged"] and st1["second_sweeps"] >= st1["outer_steps"] - st1["restarts"] - 1 > 0
    assert outm["stats"]["second_sweeps"] <= st1["second_sweeps"]
    for o in (outm, outm1):
        assert np.max(np.abs(o["lambda_"] - ref) / ref) < 1e-8
        Ym = o["Y"].cpu().numpy()
        assert np.max(np.abs(Ym.conj() @ Ym.T - np.eye(10))) < 1e-10


def test_eigs_vs_oracle_and_mixed(pkg, oracle, n2):
    print("case: eigs oracle + mixed N=2", flush=True)
    v0 = pkg.start_vector(n2["s"].nr, 5)
    lam_o, Y_o, info, _ = oracle.eigs_method_c(n2["N"], n2["eps"], Q_EXP4, v0, k=10, p=40, tol=1e-10,
                                               max_restart=30, tol_cg=1e-10, maxit_cg=3000)
    out = pkg.run_method_C(n2["s"], k=10, p=40, tol=1e-10, max_restart=30, tol_cg=1e-10, maxit_cg=3000, v0=v0)
    assert np.max(np.abs(out["lambda_"] - lam_o) / lam_o) < 1e-10
    outm = pkg.run_method_C(n2["s"], k=10, p=40, tol=1e-10, max_restart=30, tol_cg=1e-10, maxit_cg=3000,
                            v0=v0, precision="mixed")
    assert np.max(np.abs(outm["lambda_"] - lam_o) / lam_o) < 1e-8


def test_mixed_vs_fp64_n6_amplitude_08(pkg):
    """The mixed-precision configuration at its largest size (N = 6, 12^6 modes, cosine
    permittivity of amplitude 0.8): the ten eigenvalues of the mixed solve within 1e-8 of the
    FP64 solve, both converged with small Ritz residuals."""
    print("case: mixed vs FP64 N=6 amplitude 0.8", flush=True)
    N = 6
    eps, _ = pkg.cosine_material(N, 0.8, 0)
    m = pkg.MassBackend(N, eps)
    s = pkg.QSystem(m, Q_EXP1)
    try:
        v0 = pkg.start_vector(s.nr, 20260905)
        lz = dict(k=10, p=40, tol=1e-10, max_restart=50, tol_cg=1e-10, maxit_cg=1200, v0=v0)
        o64 = pkg.run_method_C(s, **lz)
        omx = pkg.run_method_C(s, precision="mixed", **lz)
        assert o64["stats"]["converged"] and omx["stats"]["converged"]
        assert o64["stats"]["max_ritz_residual"] < 1e-10 and omx["stats"]["max_ritz_residual"] < 1e-10
        assert np.all(o64["lambda_"] > 0)
        assert np.max(np.abs(omx["lambda_"] - o64["lambda_"]) / o64["lambda_"]) < 1e-8
    finally:
        s.close()
        m.close()


def test_recovery_and_diagnostics(pkg, oracle, n2, n2_eigs):
    """Recovery x = M^{-1} V y (exp2/recover_original_eigenvectors_exp2.m:18-51) against the
    oracle's, and the diagnostics of compute_fourier_diagnostics_exp2.m:10-26: the Rayleigh
    quotient equals lambda; the divergence |q . M x| / |M x| vanishes (M x = V y is transverse)
    to the recovery tolerance; the device values equal the oracle's evaluated on the same
    vectors.  The GEVP residual is bounded by what the inner tolerance allows: a Ritz pair of
    K_r^{-1} applied to PCG tolerance tol_cg has |K_r y - lambda y| <~ |K_r| tol_cg, i.e. a
    relative residual of order |K_r| tol_cg / (2 lambda) ~ 1e-7 here (|K_r| = max|q_xi|^2 ~ 30,
    lambda ~ 0.012-0.045); the reference only warns when its 1e-10 request
    (exp2/make_config_experiment2_allgpu.m:51, exp2/run_experiment2_allgpu.m:47-51) is missed."""
    print("case: recovery N=2", flush=True)
    lam = n2_eigs["lambda_"][:3]
    Y = n2_eigs["Y"][:3]
    X, MX, info = pkg.recover_original_eigenvectors(Y, n2["s"])
    Xo, MXo, _ = oracle.recover_eigenvectors(n2["osys"], n2["eps"], Y.cpu().numpy().T, n2["N"])
    # eigenvectors agree up to a phase
    for j in range(3):
        a = X[j].cpu().numpy()
        b = Xo[:, j]
        ph = np.vdot(b, a) / abs(np.vdot(b, a))
        assert np.linalg.norm(a - ph * b) / np.linalg.norm(b) < 1e-9
    assert np.allclose(info["massNorm2"], lam, rtol=1e-10)  # x'Mx before normalisation = 1/theta
    d = pkg.compute_fourier_diagnostics(X, MX, lam, n2["s"])
    res_o, div_o, ray_o = oracle.fourier_diagnostics(n2["osys"], X.cpu().numpy().T, MX.cpu().numpy().T, lam)
    assert np.allclose(d["rayleigh"], lam, rtol=1e-10)
    assert np.allclose(d["rayleigh"], ray_o, rtol=1e-12)
    qmax2 = float(np.max(n2["osys"]["Qq"]))
    bound = 10 * qmax2 * 1e-10 / (2 * lam)
    assert np.all(d["resGEVP"] < bound) and np.all(res_o < bound), (d["resGEVP"], bound)
    # the residual is a difference of nearly equal terms (~1e-7 of either); the same inputs
    # give it to ~1e-9 relative, so the device and oracle values must agree far inside 1e-6
    assert np.allclose(d["resGEVP"], res_o, rtol=1e-6, atol=0)
    # q . Mx sums three terms of size |q||Mx| to ~1e-12 of that: only the leading digits of
    # the divergence are determined by the inputs
    assert np.all(d["divRelative2"] < 1e-10) and np.all(div_o < 1e-10)
    assert np.allclose(d["divRelative2"], div_o, rtol=1e-2, atol=1e-15)


def test_batched_bloch_points(pkg, oracle, n2):
    print("case: batched 3 Bloch points N=2", flush=True)
    qs = np.array([[0.05, 0.04, 0.03], [0.13, 0.21, 0.17], [0.2, 0.1, 0.05]])
    nr = n2["s"].nr
    v0 = np.stack([pkg.start_vector(nr, 10 + i) for i in range(3)])
    outb = pkg.run_method_C_batch(n2["m"], qs, k=6, p=30, tol=1e-10, max_restart=40, v0=v0)
    for i in range(3):
        s = pkg.QSystem(n2["m"], qs[i])
        out = pkg.run_method_C(s, k=6, p=30, tol=1e-10, max_restart=40, v0=v0[i])
        assert np.max(np.abs(outb["lambda_"][i] - out["lambda_"]) / out["lambda_"]) < 1e-11
        s.close()


def test_gamma_point_rejected(pkg, n2):
    with pytest.raises(pkg.Fame6dError) as e:
        pkg.QSystem(n2["m"], (0.0, 0.0, 0.0))
    assert e.value.name == "EGAMMA"


def test_method_b_nested(pkg, oracle, n2):
    print("case: Method B nested N=2", flush=True)
    rng = np.random.default_rng(12)
    b = rand_c(rng, 2 * n2["m"].NF)
    ref, _ = oracle.krinv_nested(n2["osys"], n2["eps"], b, n2["N"], 1e-10, 1e-12, 600, 1600, 4.0)
    x, st = pkg.krinv_nested(n2["s"], torch.as_tensor(b, device="cuda"), 1e-10, 1e-12, 600, 1600, 4.0)
    assert np.linalg.norm(x.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-8
    # cross-method agreement of the eigenvalues (exp1/check_fig1bc_results.m:23-25 warns above 1e-7)
    v0 = pkg.start_vector(n2["s"].nr, 3)
    lb = pkg.run_method_B(n2["s"], k=4, p=24, tol=1e-10, max_restart=40, eps_c=4.0, v0=v0)["lambda_"]
    lc = pkg.run_method_C(n2["s"], k=4, p=24, tol=1e-10, max_restart=40, v0=v0)["lambda_"]
    assert np.max(np.abs(lb - lc) / lc) < 1e-7


def test_batched_twelve_point_path(pkg, n2):
    """12 Bloch points on t*(0.5,0.4,0.3), t = k/12 (q = 0 excluded): one batched solve equals
    the per-point solves.  Also exercises per-group buffers sized by the point count."""
    print("case: batched 12-point path N=2", flush=True)
    qs = np.array([[0.5 * t, 0.4 * t, 0.3 * t] for t in np.arange(1, 13) / 12.0])
    nr = n2["s"].nr
    v0 = np.stack([pkg.start_vector(nr, 40 + i) for i in range(12)])
    outb = pkg.run_method_C_batch(n2["m"], qs, k=4, p=20, tol=1e-10, max_restart=60, v0=v0)
    for i in (0, 5, 11):
        s = pkg.QSystem(n2["m"], qs[i])
        out = pkg.run_method_C(s, k=4, p=20, tol=1e-10, max_restart=60, v0=v0[i])
        assert np.max(np.abs(outb["lambda_"][i] - out["lambda_"]) / out["lambda_"]) < 1e-11
        s.close()


def test_kernel_family_timer(n2, pkg):
    """The in-library HIP-event timer sees every timed family of a solve, with the stated
    algorithmic bytes per launch (Mhat passes at N=2: NF*(6*16+3*8), NF*(6*16+8), NF*(8*16+2*8);
    the FP64 pass C recomputes invC from W1 and W2 instead of reading it)."""
    print("case: kernel timer N=2", flush=True)
    m, s = n2["m"], n2["s"]
    NF = m.NF
    m.profile(1)
    pkg.run_method_C(s, k=2, p=8, tol=1e-10, max_restart=50, tol_cg=1e-10, maxit_cg=3000,
                     v0=pkg.start_vector(s.nr, 7), precision="double")
    kt = m.kernel_times()
    m.profile(0)
    per = {"mhat_pass_a": NF * (6 * 16 + 3 * 8), "mhat_pass_b": NF * (6 * 16 + 8),
           "mhat_pass_c": NF * (8 * 16 + 2 * 8)}
    for fam, b in per.items():
        d = kt[fam]
        assert d["launches"] > 0 and d["ms"] > 0, fam
        assert abs(d["bytes"] / d["launches"] - b) < 1e-6 * b, fam
    for fam in ("lanczos_gemv_c", "lanczos_gemv_n"):
        assert kt[fam]["launches"] > 0 and kt[fam]["bytes"] > 0, fam
    # skipped launches after convergence are discarded, so the three passes agree
    assert kt["mhat_pass_a"]["launches"] == kt["mhat_pass_b"]["launches"] == kt["mhat_pass_c"]["launches"]


def test_method_a_top(n2, oracle, pkg):
    """Method-A operator y = (K - sigma M)^{-1} M x: the GPU solution meets the true relative
    residual 1e-10 of the shifted system, checked in numpy (make_A_eigs_top_preconditioned.m)."""
    print("case: Method A top N=2", flush=True)
    N, eps, s, osys = n2["N"], n2["eps"], n2["s"], n2["osys"]
    sigma = 1e-4
    rng = np.random.default_rng(51)
    x = rand_c(rng, 3 * s.NF)
    y, st = pkg.method_a_top(s, torch.as_tensor(x, device="cuda"), sigma, 1e-10, 4000)
    y = y.cpu().numpy()
    b = oracle.mass3_apply(eps, x, N)
    r = oracle.apply_K(osys, y) - sigma * oracle.mass3_apply(eps, y, N) - b
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.5e-10
    ref, its = oracle.method_a_top(osys, eps, x, N, sigma, 1e-10, 4000)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-5  # both at residual 1e-10, cond(A) ~ 1e5
    assert st["solves"] == 1 and st["iterations"] > 0


def test_method_a_eigs_golden_n2(n2, pkg):
    """Method A reaches the reference-held eigenvalues (exp3/make_config_experiment4.m:40) and
    agrees with Method C (exp1/check_fig1bc_results.m:23-25 warns above 1e-7)."""
    print("case: Method A eigs N=2", flush=True)
    s = n2["s"]
    outA = pkg.run_method_A(s, k=10, p=40, tol=1e-10, max_restart=300, sigma=1e-4, tol_minres=1e-10,
                            maxit_minres=4000, v0=pkg.start_vector(s.ndof, 20260906))
    lamA = outA["lambda_"]
    assert outA["stats"]["converged"] == 1
    assert abs(lamA[0] - GOLDEN_N2[0]) / GOLDEN_N2[0] < 1e-9
    assert abs(lamA[9] - GOLDEN_N2[1]) / GOLDEN_N2[1] < 1e-9
    lamC = pkg.run_method_C(s, v0=pkg.start_vector(s.nr, 7), precision="double", maxit_cg=3000)["lambda_"]
    assert np.max(np.abs(lamA - lamC) / lamC) < 1e-9


def test_tracking_scan_n2_selection_and_wide_basis(pkg):
    """Lanczos_tracking's selected-vector extraction on the narrow (p <= 64, in-place restart)
    and wide (out-of-place restart into a second basis) paths, against the full extraction of
    the same recurrence; target selection widened by one index on each side."""
    print("case: tracking scan N=2", flush=True)
    N = 2
    mass = pkg.MassBackend(N, pkg.smooth_scalar_material(N))
    sq = pkg.QSystem(mass, Q_EXP4)
    v0 = pkg.start_vector(sq.nr, 7)
    full = pkg.run_method_C(sq, k=12, p=40, tol=1e-10, max_restart=30, tol_cg=1e-10, maxit_cg=3000, v0=v0,
                            want_vectors=True)
    assert abs(full["lambda_"][0] - GOLDEN_N2[0]) / GOLDEN_N2[0] < 1e-10
    assert abs(full["lambda_"][9] - GOLDEN_N2[1]) / GOLDEN_N2[1] < 1e-10
    Yf = full["Y"].cpu().numpy()
    for k, p in ((12, 40), (70, 118)):
        t = pkg.solve_tracking_scan(sq, k=k, p=p, v0=v0, fixed_indices=[1, 10])
        assert list(t["selectedIndices"]) == [1, 10]
        assert np.max(np.abs(t["lambda_"][:12] - full["lambda_"]) / full["lambda_"]) < 1e-12
        assert np.all(t["selectedRelRitzResidual"] < 1e-10)
        Ys = t["Yselected"].cpu().numpy()
        for a, i in enumerate((1, 10)):
            cos = abs(np.vdot(Ys[a], Yf[i - 1])) / np.linalg.norm(Ys[a]) / np.linalg.norm(Yf[i - 1])
            assert abs(cos - 1.0) < 1e-10
    t = pkg.solve_tracking_scan(sq, k=70, p=118, v0=v0, target_lambda=list(GOLDEN_N2), half_width=1,
                                want_vectors=False)
    assert list(t["selectedIndices"]) == [1, 2, 9, 10, 11]
    with pytest.raises(pkg.Fame6dError):
        pkg.solve_tracking_scan(sq, k=12, p=40, v0=v0, fixed_indices=[13], want_vectors=False)
    sq.close()
    mass.close()


def test_tracked_modes_n5_reference_eigenvalues(pkg):
    """The N = 5 selected-mode scan of experiment 4 (k = 140, p = 188; modes 11 and 114,
    exp3/make_config_experiment4.m:37-50): both eigenvalues against the values the reference
    holds (1.200090550277e-02, 4.462939906027e-02, quoted to 13 digits; the reference accepts
    5e-7), and the recovered eigenvectors' M-norm Rayleigh quotients."""
    print("case: tracked modes N=5 k=140", flush=True)
    N = 5
    mass = pkg.MassBackend(N, pkg.smooth_scalar_material(N))
    sq = pkg.QSystem(mass, Q_EXP4)
    sel = pkg.solve_tracked_modes(sq, lambda_check_rel_tol=1e-11)
    ref = np.array([1.200090550277e-02, 4.462939906027e-02])
    assert np.max(np.abs(sel["lambda6D"] - ref) / ref) < 1e-11
    assert np.allclose(sel["recovery"]["massNorm2"], sel["lambda6D"], rtol=1e-10)
    assert sel["scan"]["stats"]["num_converged"] == 140
    sq.close()
    mass.close()


def test_singular_branch_qsystem_and_eigs(pkg, oracle):
    """q = (0.1, 0.1, 0.1) puts 2N modes on the (1,1,1) line, where the transverse basis takes
    the singular branch (exp1/assemble_K_and_nullspace_free_basis.m:75-91): the device q-system
    reports them, every mode keeps V V^T = K and V^T q = 0, the regular modes equal the
    oracle's diagonals, and the eigenvalues equal the oracle's solve."""
    print("case: singular q-system branch N=2", flush=True)
    N = 2
    q = (0.1, 0.1, 0.1)
    eps = oracle.eps_reference(N)
    m = pkg.MassBackend(N, eps)
    s = pkg.QSystem(m, q)
    try:
        osys = oracle.qsystem(N, q)
        assert s.info["nSingular"] == osys["nsing"] == 2 * N
        d = s.diagonals()
        qv = np.stack([osys["q1"], osys["q2"], osys["q3"]])
        V = np.array([[d["V11"], d["V12"]], [d["V21"], d["V22"]], [d["V31"], d["V32"]]])
        for i in range(3):
            for j in range(3):
                kij = (osys["Qq"] if i == j else 0) - qv[i] * qv[j]
                assert np.allclose(np.einsum("ak,ak->k", V[i], V[j]), kij, atol=1e-12)
        assert np.allclose(np.einsum("iak,ik->ak", V, qv), 0, atol=1e-12)
        den = 3 * osys["Qq"] - qv.sum(0) ** 2
        reg = den > 1e-13
        for key in ("V11", "V12", "V21", "V22", "V31", "V32", "W1", "W2", "invC"):
            assert np.allclose(d[key][reg], osys[key][reg], rtol=1e-13, atol=1e-13), key
        v0 = pkg.start_vector(s.nr, 3)
        lam_o, _, _, _ = oracle.eigs_method_c(N, eps, q, v0, k=6, p=30, tol=1e-10, max_restart=40,
                                              tol_cg=1e-10, maxit_cg=3000)
        out = pkg.run_method_C(s, k=6, p=30, tol=1e-10, max_restart=40, tol_cg=1e-10, maxit_cg=3000, v0=v0)
        assert np.max(np.abs(out["lambda_"] - lam_o) / lam_o) < 1e-9
    finally:
        s.close()
        m.close()


def test_failed_solve_frees_device_memory(pkg, n5):
    """A Method-C solve that stops with FAME6D_ENOCONV (maxit_cg = 1) releases its Lanczos basis
    and workspaces: after the first failure (which may grow the handles' own workspaces) two
    more leave the device's free memory unchanged (to the runtime's small-block caching)."""
    print("case: ENOCONV releases device memory N=5", flush=True)
    s = n5["s"]
    v0 = pkg.start_vector(s.nr, 11)

    def fail():
        with pytest.raises(pkg.Fame6dError) as e:
            pkg.run_method_C(s, k=10, p=40, tol=1e-10, max_restart=50, tol_cg=1e-10, maxit_cg=1, v0=v0)
        assert e.value.name == "ENOCONV"
        torch.cuda.synchronize()

    fail()
    free0 = torch.cuda.mem_get_info()[0]
    fail()
    fail()
    free1 = torch.cuda.mem_get_info()[0]
    # a leaked basis alone would be 42 x 32 MB; the runtime may keep a few small blocks cached
    assert abs(free1 - free0) < (64 << 20), (free0, free1)
