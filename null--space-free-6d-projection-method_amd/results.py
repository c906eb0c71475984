"""Result tables and checkpoints in the reference's formats, for interchange with its MATLAB users.

CSV tables carry the reference's column names and order:
- experiment-4 cases and per-(h, n) summaries (exp3/write_experiment4_outputs.m:23-51,
  exp3/run_experiment4.m:262-276);
- Fig. 1(b,c) rows per grid size (exp1/run_fig1bc_method.m:184-203).
Values are written as MATLAB's writetable writes doubles (15 significant digits, NaN as
"NaN", strings unquoted).

Checkpoints are MAT files at MATLAB level 5 (scipy.io; MATLAB's `load` reads them; the
reference saves `-v7.3`, an HDF5 container this environment has no writer for) together with
the reference's compatibility signatures and comparison rules:
- the selected-eigenpair checkpoint of experiment 4 (exp3/get_selected_eigenpairs_exp4.m:48-62);
- the Fig. 1(b,c) per-method signature (exp1/make_exp1bc_signature.m:1-52) compared strictly
  (exp1/exp1_signature_equal.m:3).
"""
import csv
import math

import numpy as np

CASE_COLUMNS = ("study", "mode", "h", "n", "halfWidth", "lambda6D", "lambda3D", "lambdaCrop", "eta3D", "nuLWRQ",
                "VLWRQ", "deltaLambda", "etaUsing6DLambda", "deltaPart", "deltaLWRQIdentity", "numCenters",
                "blocksPerDimension", "phaseRadiusBound", "reconstructionTimeSeconds")
SUMMARY_COLUMNS = ("study", "h", "n", "halfWidth", "maxEta3D", "maxNuLWRQ", "maxDeltaLambda", "maxEtaUsing6DLambda",
                   "maxPartitionError", "maxLWRQIdentityError", "numCenters", "blocksPerDimension",
                   "phaseRadiusBound", "reconstructionTimeSeconds")
FIG1BC_COLUMNS = ("N", "NF", "status", "lin", "wall", "observedLin", "observedWall", "outer", "setup", "plateau",
                  "maxPlateau", "gpuFree")

# the metric names of compute_yee_lwrq_metrics (this package) -> the reference's struct fields
_METRIC_NAMES = {"lambda_crop": "lambdaCrop", "lambda_3d": "lambda3D", "eta_3d": "eta3D",
                 "eta_6d": "etaUsing6DLambda", "delta_lambda": "deltaLambda", "V": "VLWRQ", "nu": "nuLWRQ",
                 "delta_part": "deltaPart", "delta_identity": "deltaLWRQIdentity", "cropped_mass": "croppedMass",
                 "lambda6d": "lambda6D", "half_width": "halfWidth"}


def metrics_struct(m):
    """compute_yee_lwrq_metrics output with the field names of compute_yee_lwrq_metrics_exp4
    (exp3/compute_yee_lwrq_metrics_exp4.m:54-68)."""
    return {_METRIC_NAMES.get(k, k): v for k, v in m.items() if k != "rho"}


def case_row(metrics, mode, study, num_centers=math.nan, blocks_per_dimension=math.nan,
             phase_radius_bound=math.nan, reconstruction_seconds=math.nan):
    """One row of the case table (exp3/run_experiment4.m:81-89 fills these fields)."""
    s = metrics_struct(metrics)
    s.update(mode=mode, study=study, numCenters=num_centers, blocksPerDimension=blocks_per_dimension,
             phaseRadiusBound=phase_radius_bound, reconstructionTimeSeconds=reconstruction_seconds)
    return {c: s.get(c, math.nan) for c in CASE_COLUMNS}


def summarize_case(rows, h, n, study, num_centers=math.nan, blocks_per_dimension=math.nan,
                   phase_radius_bound=math.nan, reconstruction_seconds=math.nan):
    """summarize_case (exp3/run_experiment4.m:262-276): maxima over the modes of one (h, n)."""
    mx = lambda key: max(r[key] for r in rows)
    return dict(study=study, h=h, n=n, halfWidth=h * n, maxEta3D=mx("eta3D"), maxNuLWRQ=mx("nuLWRQ"),
                maxDeltaLambda=mx("deltaLambda"), maxEtaUsing6DLambda=mx("etaUsing6DLambda"),
                maxPartitionError=mx("deltaPart"), maxLWRQIdentityError=mx("deltaLWRQIdentity"),
                numCenters=num_centers, blocksPerDimension=blocks_per_dimension,
                phaseRadiusBound=phase_radius_bound, reconstructionTimeSeconds=reconstruction_seconds)


def _fmt(v):
    if isinstance(v, str):
        return v
    if isinstance(v, (bool, np.bool_)):
        return "1" if v else "0"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    x = float(v)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    return "%.15g" % x


def write_table(path, rows, columns):
    """writetable(table(...), path) for a list of row dicts."""
    with open(path, "w", newline="") as f:
        w = csv.writer(f, lineterminator="\n")
        w.writerow(columns)
        for r in rows:
            w.writerow([_fmt(r[c]) for c in columns])


def read_table(path):
    """The rows of a table written by write_table (or MATLAB's writetable): strings stay
    strings, numeric text becomes float."""
    out = []
    with open(path, newline="") as f:
        rd = csv.DictReader(f)
        for r in rd:
            row = {}
            for k, v in r.items():
                try:
                    row[k] = float(v)
                except ValueError:
                    row[k] = v
            out.append(row)
    return out


def write_experiment4_tables(cases, mesh_summary, window_summary, detailed_csv, mesh_csv, window_csv,
                             default_csv=None, default_h=0.025, default_n=20):
    """The four CSV files of write_experiment4_outputs (exp3/write_experiment4_outputs.m:8-19)."""
    write_table(detailed_csv, cases, CASE_COLUMNS)
    write_table(mesh_csv, mesh_summary, SUMMARY_COLUMNS)
    write_table(window_csv, window_summary, SUMMARY_COLUMNS)
    if default_csv is not None:
        dflt = [c for c in cases if c["study"] == "mesh" and abs(c["h"] - default_h) < 1e-14 and c["n"] == default_n]
        write_table(default_csv, dflt, CASE_COLUMNS)


# ----------------------------------------------------------------------------- configurations
def config_experiment4():
    """The fields of make_config_experiment4 that its checkpoint signature reads
    (exp3/make_config_experiment4.m:17-61)."""
    return dict(N=5, q=[0.13, 0.21, 0.17], modes=[11, 114], epsC=4.0, alpha=1.2, tolCG=1e-10, tolLanczos=1e-10,
                tolRecoverM=1e-12, kScan=140)


def experiment4_checkpoint_signature(cfg):
    """make_checkpoint_signature (exp3/get_selected_eigenpairs_exp4.m:48-55)."""
    return dict(N=int(cfg["N"]), q=[float(x) for x in cfg["q"]], modes=[int(m) for m in cfg["modes"]],
                epsC=float(cfg["epsC"]), alpha=float(cfg["alpha"]), tolCG=float(cfg["tolCG"]),
                tolLanczos=float(cfg["tolLanczos"]), tolRecoverM=float(cfg["tolRecoverM"]),
                kScan=int(cfg["kScan"]))


def experiment4_signature_matches(sig, cfg):
    """checkpoint_signature_matches (exp3/get_selected_eigenpairs_exp4.m:57-62): exact fields,
    q to 1e-14 in the max norm."""
    ref = experiment4_checkpoint_signature(cfg)
    try:
        return (int(sig["N"]) == ref["N"] and [int(m) for m in np.ravel(sig["modes"])] == ref["modes"]
                and float(np.max(np.abs(np.ravel(sig["q"]) - np.asarray(ref["q"])))) < 1e-14
                and float(sig["epsC"]) == ref["epsC"] and float(sig["alpha"]) == ref["alpha"]
                and float(sig["tolCG"]) == ref["tolCG"] and float(sig["tolLanczos"]) == ref["tolLanczos"]
                and float(sig["tolRecoverM"]) == ref["tolRecoverM"] and int(sig["kScan"]) == ref["kScan"])
    except (KeyError, TypeError, ValueError):
        return False


def config_exp1bc():
    """make_config_exp1bc (exp1/make_config_exp1bc.m:11-99), the fields its signature reads."""
    r2, r3, r5 = math.sqrt(2.0), math.sqrt(3.0), math.sqrt(5.0)
    return dict(
        problem=dict(P=[[1, 0, 0], [0, 1, 0], [0, 0, 1], [r2, 0, 0], [0, r3, 0], [0, 0, r5]], T=[2 * math.pi] * 6),
        material=dict(epsC=4.0, alpha=1.2), mass=dict(backend="rank1-exact-scalar-lowmem"),
        scale=dict(qPhys=[0.05, 0.04, 0.03], NList=[3, 4, 5, 6, 7]), spectrum=dict(numEig=10),
        eigs=dict(tolerance=1e-10, maxIterations=300, subspaceDimension=40, rngSeed=20260906),
        shift=dict(sigma=1e-4, minresTolerance=1e-10, minresMaxit=4000, useSpectralPreconditioner=True,
                   maxRecoveryEventsPerTop=50, maxResidualCorrectionCycles=8, maxZeroProgressEvents=2,
                   minResidualProgressFraction=1e-8, correctionSafetyFactor=0.25,
                   correctionMaxRelativeTolerance=1e-1, correctionMinRelativeTolerance=1e-6,
                   plateauAcceptanceCeiling=1e-9, plateauMinFailedRecoveryCycles=2,
                   plateauMinRelativeReduction=1e-2, guardNearToleranceCeiling=1e-9),
        lanczos=dict(krylovDimension=40, tolerance=1e-10, maxRestartCycles=50, rngSeed=20260905,
                     hybridCPUOuterBasis=True),
        nested=dict(tolKr=1e-10, tolM=1e-12, maxitKr=600, maxitM=1600, useKrPreconditioner=True))


def exp1bc_signature(cfg, method, N):
    """make_exp1bc_signature (exp1/make_exp1bc_signature.m:3-52): the complete compatibility
    signature of one method and grid size."""
    method = str(method).upper()
    s, sh, lz, ns = {}, cfg["shift"], cfg["lanczos"], cfg["nested"]
    s["version"] = 6 if method == "A" else 4
    s["kind"] = "fig1bc-hybrid-lowmem"
    s["method"], s["N"] = method, int(N)
    s["numEig"] = cfg["spectrum"]["numEig"]
    s["P"], s["T"] = cfg["problem"]["P"], cfg["problem"]["T"]
    s["epsC"], s["alpha"] = cfg["material"]["epsC"], cfg["material"]["alpha"]
    s["rank1Backend"] = cfg["mass"]["backend"]
    s["q"], s["NList"] = cfg["scale"]["qPhys"], cfg["scale"]["NList"]
    e = cfg["eigs"]
    s["eigsTol"], s["eigsMaxIterations"], s["eigsP"], s["eigsSeed"] = (e["tolerance"], e["maxIterations"],
                                                                      e["subspaceDimension"], e["rngSeed"])
    s["sigma"], s["minresTol"], s["minresMaxit"] = sh["sigma"], sh["minresTolerance"], sh["minresMaxit"]
    s["useSpectralPreconditioner"] = sh["useSpectralPreconditioner"]
    s["maxRecoveryEvents"] = sh["maxRecoveryEventsPerTop"]
    s["maxCorrectionCycles"] = sh["maxResidualCorrectionCycles"]
    s["maxZeroProgressEvents"] = sh["maxZeroProgressEvents"]
    s["minResidualProgressFraction"] = sh["minResidualProgressFraction"]
    s["correctionSafetyFactor"] = sh["correctionSafetyFactor"]
    s["correctionMaxTol"] = sh["correctionMaxRelativeTolerance"]
    s["correctionMinTol"] = sh["correctionMinRelativeTolerance"]
    s["plateauCeiling"] = sh["plateauAcceptanceCeiling"]
    s["plateauMinCycles"] = sh["plateauMinFailedRecoveryCycles"]
    s["plateauMinReduction"] = sh["plateauMinRelativeReduction"]
    if method == "A":
        s["guardNearToleranceCeiling"] = sh["guardNearToleranceCeiling"]
    s["lanczosTol"], s["lanczosMaxRestart"] = lz["tolerance"], lz["maxRestartCycles"]
    s["lanczosP"], s["lanczosSeed"] = lz["krylovDimension"], lz["rngSeed"]
    s["hybridCPUOuterBasis"] = lz["hybridCPUOuterBasis"]
    s["tolKr"], s["tolM"], s["maxitKr"], s["maxitM"] = ns["tolKr"], ns["tolM"], ns["maxitKr"], ns["maxitM"]
    s["useKrPreconditioner"] = ns["useKrPreconditioner"]
    return s


def exp1_signature_equal(a, b):
    """isequaln(a, b) (exp1/exp1_signature_equal.m:3): same fields, equal values, NaN equal to NaN."""
    def eq(x, y):
        if isinstance(x, dict) or isinstance(y, dict):
            return isinstance(x, dict) and isinstance(y, dict) and x.keys() == y.keys() and all(
                eq(x[k], y[k]) for k in x)
        if isinstance(x, str) or isinstance(y, str):
            return x == y
        ax, ay = np.asarray(x, dtype=float), np.asarray(y, dtype=float)
        return ax.shape == ay.shape and bool(np.all((ax == ay) | (np.isnan(ax) & np.isnan(ay))))
    return eq(a, b)


# ----------------------------------------------------------------------------- checkpoints
def save_checkpoint(path, **variables):
    """save(path, names...) as a level-5 MAT file; dicts become MATLAB structs, complex arrays
    stay complex."""
    import scipy.io
    scipy.io.savemat(path, variables, do_compression=True, long_field_names=True)


def load_checkpoint(path):
    """The variables of a MAT file as plain Python values (structs -> dicts)."""
    import scipy.io
    raw = scipy.io.loadmat(path, squeeze_me=True, struct_as_record=False, simplify_cells=True)
    return {k: v for k, v in raw.items() if not k.startswith("__")}


def save_selected_eigenpairs(path, selected, cfg):
    """The checkpoint of get_selected_eigenpairs_exp4 (exp3/get_selected_eigenpairs_exp4.m:36-42):
    `selected` (modes, lambda6D, Xcpu, ...) and its checkpointSignature."""
    save_checkpoint(path, selected=selected, checkpointSignature=experiment4_checkpoint_signature(cfg))


def load_selected_eigenpairs(path, cfg):
    """The checkpointed `selected` struct if the file exists and its signature matches cfg
    (exp3/get_selected_eigenpairs_exp4.m:8-22), else None."""
    import os
    if not os.path.exists(path):
        return None
    v = load_checkpoint(path)
    if "selected" not in v or "checkpointSignature" not in v:
        return None
    if not experiment4_signature_matches(v["checkpointSignature"], cfg):
        return None
    return v["selected"]
