"""GPU parity of the Yee curl-curl / LWRQ closure (through the C ABI) with the oracle."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def flat(x):
    """[x, y, z] array -> column-major flat device tensor (the library layout)."""
    return torch.as_tensor(np.asarray(x).ravel(order="F"), device="cuda")


def unflat(t, m):
    return t.cpu().numpy().reshape((m, m, m), order="F")


def rand_fields(rng, n):
    m = 2 * n + 3
    return [rng.standard_normal((m, m, m)) + 1j * rng.standard_normal((m, m, m)) for _ in range(3)]


def smooth_fields(n, h):
    """A smooth non-eigen field (sums of a few staggered plane waves), closer to what the
    Taylor reconstruction hands over than white noise."""
    I = np.arange(-(n + 1), n + 2) * h
    out = []
    for c in range(3):
        ax = [I + (h / 2 if d == c else 0.0) for d in range(3)]
        X, Y, Z = ax[0][:, None, None], ax[1][None, :, None], ax[2][None, None, :]
        out.append((1.0 + 0.3 * c) * np.exp(1j * (1.1 * X - 0.4 * Y + 0.9 * Z))
                   + 0.2 * np.exp(1j * (-0.5 * X + 2.0 * Y + 0.3 * Z + c))
                   + 0.05 * np.cos(3.0 * X) * np.sin(2.0 * Y + Z))
    return out


def test_apply_cropped_A(pkg, oracle):
    print("case: cropped A n=5", flush=True)
    rng = np.random.default_rng(7)
    n, h = 5, 0.1
    u = rand_fields(rng, n)
    A = pkg.apply_cropped_A(*[flat(x) for x in u], n, h)
    R = oracle.apply_cropped_A(*u, h)
    q = 2 * n + 1
    for a, r in zip(A, R):
        assert np.max(np.abs(unflat(a, q) - r)) <= 1e-13 * np.max(np.abs(r))


def test_sample_eps(pkg, oracle):
    print("case: Yee material n=6", flush=True)
    n, h = 6, 0.05
    E = pkg.sample_scalar_yee_material(n, h)
    R = oracle.sample_scalar_yee_material(oracle.P_REF, 4.0, 1.2, h, n)
    for e, r in zip(E, R):
        assert np.max(np.abs(unflat(e, 2 * n + 1) - r)) <= 1e-14 * np.max(r)


def test_lwrq(pkg, oracle):
    print("case: LWRQ n=7", flush=True)
    rng = np.random.default_rng(8)
    n, h = 7, 0.025
    q = 2 * n + 1
    a = rng.standard_normal((q, q, q)) + 0.5
    b = 1.0 + rng.random((q, q, q))
    g = pkg.compute_lwrq(flat(a), flat(b), n, h, return_rho=True)
    r = oracle.compute_lwrq(a, b, h, return_rho=True)
    assert abs(g["lambda_"] - r["lambda_"]) <= 1e-13 * abs(r["lambda_"])
    assert abs(g["V"] - r["V"]) <= 1e-11 * r["V"]
    assert g["partition_error"] < 1e-13 and r["partition_error"] < 1e-13
    assert np.max(np.abs(unflat(g["rho"], q) - r["rho"])) <= 1e-13 * np.max(np.abs(r["rho"]))
    # a zero mass is refused (exp3/compute_lwrq_exp4.m:31-33)
    with pytest.raises(pkg.Fame6dError) as e:
        pkg.compute_lwrq(flat(a), flat(np.zeros_like(b)), n, h)
    assert e.value.name == "ENOCONV"


@pytest.mark.parametrize("kind", ["random", "smooth"])
def test_metrics_vs_oracle(pkg, oracle, kind):
    print("case: Yee metrics n=6 %s" % kind, flush=True)
    n, h = 6, 0.05
    u = rand_fields(np.random.default_rng(9), n) if kind == "random" else smooth_fields(n, h)
    lam6 = 0.0123
    g = pkg.compute_yee_lwrq_metrics(*[flat(x) for x in u], lam6, n, h, return_rho=True)
    r = oracle.compute_yee_lwrq_metrics(*u, lam6, h, n, return_rho=True)
    for key in ("lambda_crop", "lambda_3d", "eta_3d", "eta_6d", "delta_lambda", "V", "nu", "cropped_mass"):
        assert abs(g[key] - r[key]) <= 1e-11 * abs(r[key]), (key, g[key], r[key])
    for key in ("delta_part", "delta_identity", "energy_partition_error", "mass_partition_error"):
        assert g[key] < 1e-12 and r[key] < 1e-12, key
    q = 2 * n + 1
    assert np.max(np.abs(unflat(g["rho"], q) - r["rho"])) <= 1e-11 * np.max(np.abs(r["rho"]))


@pytest.mark.parametrize("R", [1, 2, 5, 8])
def test_metrics_vs_oracle_radius(pkg, oracle, R):
    """The fused x/y/z LWRQ pass at other neighbour radii (one compiled instance each), on a
    window whose z extent spans several chunks of the march for the small radii."""
    print("case: Yee metrics n=9 radius %d" % R, flush=True)
    n, h = 9, 0.05
    u = rand_fields(np.random.default_rng(10 + R), n)
    lam6 = 0.0123
    g = pkg.compute_yee_lwrq_metrics(*[flat(x) for x in u], lam6, n, h, radius=R, return_rho=True)
    r = oracle.compute_yee_lwrq_metrics(*u, lam6, h, n, R=R, return_rho=True)
    for key in ("lambda_3d", "eta_3d", "V", "nu", "cropped_mass"):
        assert abs(g[key] - r[key]) <= 1e-11 * abs(r[key]), (key, g[key], r[key])
    q = 2 * n + 1
    assert np.max(np.abs(unflat(g["rho"], q) - r["rho"])) <= 1e-11 * np.max(np.abs(r["rho"]))


def test_metrics_plane_wave_window_n160(pkg):
    """The largest window of the reference's window study (n = 160, h = 0.025,
    exp3/make_config_experiment4.m:89-91) on an exact discrete eigenpair: a transverse
    staggered plane wave in constant eps has lambda = |sigma|^2/epsC, rho = lambda, V = 0 and
    zero residuals, sigma_d = (2/h) sin(k_d h/2)."""
    n, h = 160, 0.025
    print("case: Yee metrics plane wave n=%d" % n, flush=True)
    kvec = np.array([1.3, -0.7, 2.1])
    sigma = 2.0 / h * np.sin(kvec * h / 2)
    p = np.cross(sigma, [0.3, 1.0, -0.4]) * (1.0 + 0.5j)
    p = p / np.linalg.norm(p)
    m = 2 * n + 3
    I = torch.arange(-(n + 1), n + 2, dtype=torch.float64, device="cuda") * h
    u = []
    for c in range(3):
        ax = [I + (h / 2 if d == c else 0.0) for d in range(3)]
        # column-major (x fastest) = row-major [z, y, x]
        ph = kvec[0] * ax[0][None, None, :] + kvec[1] * ax[1][None, :, None] + kvec[2] * ax[2][:, None, None]
        u.append((complex(p[c]) * torch.exp(1j * ph)).reshape(-1))
        del ph
    lam = float(sigma @ sigma) / 4.0
    work = pkg.yee.workspace(n)
    g = pkg.compute_yee_lwrq_metrics(*u, lam, n, h, eps_c=4.0, alpha=0.0, work=work)
    assert u[0].numel() == m ** 3
    # the stencil sums terms of size |u|/h^2 = 1600|u| into a result of size 7|u|: rounding of
    # a few 1e-13 relative is the exact answer here
    assert abs(g["lambda_3d"] - lam) <= 1e-11 * lam
    assert g["eta_3d"] < 1e-11 and g["eta_6d"] < 1e-11 and g["delta_lambda"] < 1e-11
    assert g["V"] < 1e-22 and g["delta_identity"] < 1e-13 and g["delta_part"] < 1e-13


def test_metrics_residual_far_from_6d_eigenvalue(pkg, oracle):
    """lambda6d far from lambda_3d on an exact eigenpair: the residual sums taken about
    lambda6d cancel to nothing at lambda_3d, so the residual is recomputed directly; eta_6d of
    a plane wave with lambda6d = 2 lambda is exactly 1/3 (||Au|| = lambda ||Bu||)."""
    n, h = 8, 0.05
    print("case: Yee metrics residual pass n=%d" % n, flush=True)
    kvec = np.array([0.9, 1.7, -1.1])
    sigma = 2.0 / h * np.sin(kvec * h / 2)
    p = np.cross(sigma, [1.0, -0.2, 0.5])
    I = np.arange(-(n + 1), n + 2) * h
    u = []
    for c in range(3):
        ax = [I + (h / 2 if d == c else 0.0) for d in range(3)]
        ph = kvec[0] * ax[0][:, None, None] + kvec[1] * ax[1][None, :, None] + kvec[2] * ax[2][None, None, :]
        u.append(p[c] * np.exp(1j * ph))
    lam = float(sigma @ sigma) / 4.0
    g = pkg.compute_yee_lwrq_metrics(*[flat(x) for x in u], 2 * lam, n, h, eps_c=4.0, alpha=0.0)
    r = oracle.compute_yee_lwrq_metrics(*u, 2 * lam, h, n, epsC=4.0, alpha=0.0)
    assert g["eta_3d"] < 1e-12 and r["eta_3d"] < 1e-12
    assert abs(g["eta_6d"] - 1.0 / 3.0) < 1e-12 and abs(r["eta_6d"] - 1.0 / 3.0) < 1e-12
    assert abs(g["lambda_3d"] - lam) < 1e-12 * lam


def test_crop_plus_field(pkg, oracle):
    print("case: crop", flush=True)
    rng = np.random.default_rng(10)
    n_large, n_small = 6, 3
    u = rand_fields(rng, n_large)[0]
    c = pkg.crop_plus_field(flat(u), n_large, n_small)
    assert np.array_equal(unflat(c, 2 * n_small + 3), oracle.crop_plus_field(u, n_large, n_small))
