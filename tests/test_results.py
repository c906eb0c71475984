"""Result tables and checkpoints in the reference's formats (host side, no GPU)."""
import math

import numpy as np


def test_case_table_round_trip(pkg, tmp_path):
    R = pkg.results
    m = dict(lambda_crop=0.0119999, lambda_3d=0.01199995658654, eta_3d=1.208691588256517e-05, eta_6d=4.16e-5,
             delta_lambda=7.9e-5, V=1.6e-10, nu=1.29e-5, delta_part=1.9e-16, delta_identity=2.9e-16,
             cropped_mass=3.1, lambda6d=0.012000905502772889, h=0.025, n=20, half_width=0.5, rho=None)
    rows = [R.case_row(m, 11, "mesh", 64, 4, 4.2, 0.011), R.case_row(dict(m, eta_3d=3.4e-5), 114, "mesh")]
    assert list(rows[0]) == list(R.CASE_COLUMNS)
    assert rows[0]["lambda6D"] == m["lambda6d"] and rows[0]["etaUsing6DLambda"] == m["eta_6d"]
    p = tmp_path / "cases.csv"
    R.write_table(p, rows, R.CASE_COLUMNS)
    header = p.read_text().splitlines()[0]
    assert header == ",".join(R.CASE_COLUMNS)
    back = R.read_table(p)
    assert back[0]["study"] == "mesh" and back[0]["mode"] == 11.0
    assert abs(back[0]["eta3D"] - m["eta_3d"]) <= 1e-14 * m["eta_3d"]  # 15 significant digits
    assert math.isnan(back[1]["numCenters"])
    s = R.summarize_case(rows, 0.025, 20, "mesh", 64, 4, 4.2, 0.011)
    assert s["maxEta3D"] == 3.4e-5 and s["halfWidth"] == 0.025 * 20
    out = {k: tmp_path / (k + ".csv") for k in ("detailed", "mesh", "window", "default")}
    R.write_experiment4_tables(rows + [dict(rows[0], h=0.05, n=10)], [s], [s], out["detailed"], out["mesh"],
                               out["window"], out["default"])
    assert len(R.read_table(out["default"])) == 2 and len(R.read_table(out["detailed"])) == 3
    assert R.read_table(out["mesh"])[0]["maxNuLWRQ"] == 1.29e-5


def test_experiment4_checkpoint_signature(pkg, tmp_path):
    R = pkg.results
    cfg = R.config_experiment4()
    sig = R.experiment4_checkpoint_signature(cfg)
    assert R.experiment4_signature_matches(sig, cfg)
    assert R.experiment4_signature_matches(dict(sig, q=[0.13 + 5e-15, 0.21, 0.17]), cfg)
    assert not R.experiment4_signature_matches(dict(sig, q=[0.13 + 1e-13, 0.21, 0.17]), cfg)
    assert not R.experiment4_signature_matches(dict(sig, modes=[11, 113]), cfg)
    assert not R.experiment4_signature_matches({k: v for k, v in sig.items() if k != "kScan"}, cfg)
    # MAT round trip of the selected eigenpairs (complex eigenvectors, struct fields)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((3 * 64, 2)) + 1j * rng.standard_normal((3 * 64, 2))
    selected = dict(N=5, NF=64, modes=np.array([11, 114]), lambda6D=np.array([1.2e-2, 4.46e-2]), Xcpu=X)
    path = str(tmp_path / "ck.mat")
    R.save_selected_eigenpairs(path, selected, cfg)
    got = R.load_selected_eigenpairs(path, cfg)
    assert got is not None and np.array_equal(got["Xcpu"], X) and list(got["modes"]) == [11, 114]
    assert R.load_selected_eigenpairs(path, dict(cfg, kScan=120)) is None
    assert R.load_selected_eigenpairs(str(tmp_path / "missing.mat"), cfg) is None


def test_exp1bc_signature(pkg):
    R = pkg.results
    cfg = R.config_exp1bc()
    a, c = R.exp1bc_signature(cfg, "a", 5), R.exp1bc_signature(cfg, "C", 5)
    assert a["version"] == 6 and c["version"] == 4 and a["method"] == "A"
    assert "guardNearToleranceCeiling" in a and "guardNearToleranceCeiling" not in c
    assert R.exp1_signature_equal(c, R.exp1bc_signature(cfg, "C", 5))
    assert not R.exp1_signature_equal(c, R.exp1bc_signature(cfg, "C", 6))
    assert not R.exp1_signature_equal(a, c)
    assert R.exp1_signature_equal(dict(c, sigma=float("nan")), dict(c, sigma=float("nan")))  # isequaln
    cfg2 = R.config_exp1bc()
    cfg2["shift"]["sigma"] = 2e-4
    assert not R.exp1_signature_equal(c, R.exp1bc_signature(cfg2, "C", 5))
    rows = [dict(N=3, NF=46656, status="completed", lin=1.0, wall=2.0, observedLin=1.0, observedWall=2.0, outer=64,
                 setup=0.1, plateau=0, maxPlateau=float("nan"), gpuFree=200.0)]
    assert list(rows[0]) == list(R.FIG1BC_COLUMNS)
