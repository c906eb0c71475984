"""CPU oracle: a numpy FP64 restatement of the reference MATLAB hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker.  The product
(the HIP library and its host package) never calls it.

Every function cites the reference file:line it restates (paths relative to the
reference repository root).  Conventions (vectorisation, frequency order, FFT
normalisation, stacking) are those of the reference:

- a grid vector is the (2N)^6 array vectorised column-major (first index fastest);
- frequency index k maps to k for k < N and k - 2N otherwise;
- fftn is unnormalised and ifftn carries 1/NF, M_s x = fftn(eps .* ifftn(x));
- multi-component vectors are stacked component-first.

Parity pins (see tests/test_oracle.py): the reference-held eigenvalues of
exp3/make_config_experiment4.m:40, the rank-1 lattice identity of
exp1/smoke_test_rank1_lowmem.m:18-23, dense brute force at N = 1, closed forms
(constant eps, the inverse mass with 1/eps), the toy-operator Lanczos test of
exp1/smoke_test_rank1_lowmem.m:44-54, and action traces of the Method-A recovery schedule
derived by hand from exp1/make_A_eigs_top_preconditioned.m (tests/test_shift_schedule.py).
"""
import math

import numpy as np

# Projection and period of every reference configuration (exp1/make_config_exp1bc.m:17-26).
P_REF = np.array([[1.0, 0.0, 0.0],
                  [0.0, 1.0, 0.0],
                  [0.0, 0.0, 1.0],
                  [math.sqrt(2.0), 0.0, 0.0],
                  [0.0, math.sqrt(3.0), 0.0],
                  [0.0, 0.0, math.sqrt(5.0)]])

EPS = np.finfo(float).eps

# FFT threads of the oracle: None uses numpy's FFT (one thread); an integer uses scipy.fft
# with that many workers (the timed CPU baseline; the results agree to rounding).
_FFT_WORKERS = [None]


def set_fft_workers(workers):
    """Threads for the oracle's n-dimensional FFTs (None: numpy, one thread)."""
    _FFT_WORKERS[0] = workers


def _fftn(a):
    if _FFT_WORKERS[0] is None:
        return np.fft.fftn(a)
    import scipy.fft
    return scipy.fft.fftn(a, workers=_FFT_WORKERS[0])


def _ifftn(a):
    if _FFT_WORKERS[0] is None:
        return np.fft.ifftn(a)
    import scipy.fft
    return scipy.fft.ifftn(a, workers=_FFT_WORKERS[0])


# ----------------------------------------------------------------------------- grid
def fourier_freq(N):
    """freq1 = [0:N-1, -N:-1] (exp1/build_fourier_grid_6d.m:12)."""
    return np.concatenate([np.arange(N), np.arange(-N, 0)]).astype(float)


def grid_shape(N):
    return (2 * N,) * 6


def vec(a):
    """Column-major vectorisation of a 6D grid array."""
    return np.asarray(a).reshape(-1, order="F")


def unvec(x, N):
    return np.asarray(x).reshape(grid_shape(N), order="F")


def coord_grids(N):
    """x_k = 2*pi*j/nGrid on the 6D grid (exp1/build_experiment1_common.m:10-11)."""
    n = 2 * N
    c = np.arange(n) / n * 2 * np.pi
    return np.meshgrid(*([c] * 6), indexing="ij")


def eps_reference(N, epsC=4.0, alpha=1.2):
    """eps = epsC*exp((alpha/6)*sum cos x_k) (exp1/build_smooth_scalar_material_compact.m:3-7)."""
    xs = coord_grids(N)
    phi = (alpha / 6.0) * sum(np.cos(x) for x in xs)
    return vec(epsC * np.exp(phi))


def eps_cosine(N, amp, phases):
    """eps = 1 + (amp/6)*sum cos(2*pi*y_k + phi_k), y_k = j/nGrid: the cosine family of the
    benchmark configurations with the amplitude split over the six terms so that eps > 0."""
    n = 2 * N
    y = np.arange(n) / n
    ys = np.meshgrid(*([y] * 6), indexing="ij")
    s = sum(np.cos(2 * np.pi * ys[k] + phases[k]) for k in range(6))
    return vec(1.0 + (amp / 6.0) * s)


# ----------------------------------------------------------------------------- mass operator
def mass_apply(eps, x, N):
    """M_s x = fftn(eps .* ifftn(x)) per column (exp1/applyMij_fft.m:9-23)."""
    x = np.asarray(x, dtype=complex)
    E = unvec(eps, N)
    if x.ndim == 1:
        return vec(_fftn(E * _ifftn(unvec(x, N))))
    return np.stack([mass_apply(eps, x[:, j], N) for j in range(x.shape[1])], axis=1)


def mass_inverse_apply(eps, x, N):
    """M_s^{-1} x = fftn(ifftn(x) ./ eps): closed form of the inverse circulant operator."""
    E = unvec(eps, N)
    return vec(_fftn(_ifftn(unvec(x, N)) / E))


def rank1_plan(N, dim=6):
    """Exact radix rank-1 plan (exp1/build_radix_exact_rank1_plan_scalar.m:18-63)."""
    n = 2 * N
    B = 2 * n - 1
    NF = n ** dim
    z = B ** np.arange(dim)
    M = n * B ** (dim - 1)
    s = n - 1
    if M >= 2 ** 32 - 1:
        raise ValueError("Rank-1 length exceeds uint32 indexing range")
    idx = np.arange(NF)
    tmp = idx.copy()
    code = np.zeros(NF, dtype=np.int64)
    for j in range(dim):
        kj = tmp % n
        tmp //= n
        code += kj * z[j]
    shift = (s * int(z.sum())) % M
    target = (code + shift) % M
    return dict(N=N, n=n, B=B, NF=NF, M=M, z=z, input=code, target=target)


def rank1_kernel_fft(eps, plan):
    """kernelFFT = fft(h), h(code) = epsHat[residue(code)]/NF
    (exp1/build_exact_rank1_scalar_kernel_fft.m:9-37)."""
    N, n, B, M = plan["N"], plan["n"], plan["B"], plan["M"]
    Ehat = vec(_fftn(unvec(eps, N)))
    code = np.arange(M)
    tmp = code.copy()
    res = np.zeros(M, dtype=np.int64)
    stride = 1
    s = n - 1
    for _ in range(6):
        digit = tmp % B
        tmp //= B
        res += ((digit - s) % n) * stride
        stride *= n
    h = Ehat[res] / plan["NF"]
    return np.fft.fft(h)


def rank1_apply(x, kfft, plan):
    """applyMij_rank1_exact: scatter -> fft -> .*kernelFFT -> ifft -> gather
    (exp1/applyMij_rank1_exact.m:4-5, rank1_exact_fft_scatter.m:10-12, rank1_exact_ifft_gather.m:9-10)."""
    w = np.zeros(plan["M"], dtype=complex)
    w[plan["input"]] = x
    w = np.fft.ifft(kfft * np.fft.fft(w))
    return w[plan["target"]]


# ----------------------------------------------------------------------------- q-system
def klift(q, P=P_REF):
    """k_lift = P*((P.'*P)\\q) (exp1/build_experiment1_q_system_C_compact.m:9)."""
    return P @ np.linalg.solve(P.T @ P, np.asarray(q, float))


def projected_q_modes(N, q, P=P_REF):
    """qModes(:,c) = sum_l P(l,c)*(Xi_l + k_lift_l) (exp3/build_projected_q_modes.m:14-20);
    Xi = freq on the grid with T = 2*pi (exp1/build_fourier_grid_6d.m:14-19)."""
    f = fourier_freq(N)
    Xi = [vec(g) for g in np.meshgrid(*([f] * 6), indexing="ij")]
    kl = klift(q, P)
    Q = np.zeros((len(Xi[0]), 3))
    for l in range(6):
        for c in range(3):
            Q[:, c] += P[l, c] * (Xi[l] + kl[l])
    return Q


def qsystem(N, q, P=P_REF):
    """Diagonals of the null-space-free basis and the compact Method-C system
    (exp1/assemble_K_and_nullspace_free_basis.m:18-105,
     exp1/build_experiment1_q_system_C_compact.m:9-31)."""
    Q = projected_q_modes(N, q, P)
    q1, q2, q3 = Q[:, 0], Q[:, 1], Q[:, 2]
    Qq = q1 ** 2 + q2 ** 2 + q3 ** 2
    if Qq.min() <= 1e-13:
        raise ValueError("Gamma point: zero projected Fourier mode")
    sq = np.sqrt(Qq)
    Qs = q1 + q2 + q3
    den = 3 * Qq - Qs ** 2
    reg = den > 1e-13
    u = np.zeros((3, 2, len(Qq)))
    s1 = 1.0 / np.sqrt(Qq[reg] * den[reg])
    s2 = 1.0 / np.sqrt(den[reg])
    for c, qc in enumerate((q1, q2, q3)):
        u[c, 0, reg] = (Qq[reg] - qc[reg] * Qs[reg]) * s1
    u[0, 1, reg] = (q3[reg] - q2[reg]) * s2
    u[1, 1, reg] = (q1[reg] - q3[reg]) * s2
    u[2, 1, reg] = (q2[reg] - q1[reg]) * s2
    for m in np.nonzero(~reg)[0]:  # singular branch (:75-91)
        qv = np.array([q1[m], q2[m], q3[m]])
        h = qv / np.linalg.norm(qv)
        a = np.zeros(3)
        a[np.argmin(np.abs(h))] = 1.0
        z1 = a - h * (h @ a)
        z1 /= np.linalg.norm(z1)
        z2 = np.cross(h, z1)
        z2 /= np.linalg.norm(z2)
        u[:, 0, m] = z1
        u[:, 1, m] = z2
    V = u * sq
    V11, V21, V31 = V[0, 0], V[1, 0], V[2, 0]
    V12, V22, V32 = V[0, 1], V[1, 1], V[2, 1]
    det = V11 * V22 - V12 * V21
    W1 = (V22 * V31 - V21 * V32) / det
    W2 = (V11 * V32 - V12 * V31) / det
    invC = 1.0 / np.sqrt(1.0 + W1 ** 2 + W2 ** 2)
    return dict(N=N, NF=len(Qq), q1=q1, q2=q2, q3=q3, Qq=Qq, V11=V11, V12=V12, V21=V21, V22=V22,
                V31=V31, V32=V32, det=det, W1=W1, W2=W2, invC=invC, klift=klift(q, P),
                nsing=int((~reg).sum()))


def apply_V(sys, y):
    """V y (exp1/apply_V_blockdiag.m:12-16)."""
    n = sys["NF"]
    y1, y2 = y[:n], y[n:]
    return np.concatenate([sys["V11"] * y1 + sys["V12"] * y2,
                           sys["V21"] * y1 + sys["V22"] * y2,
                           sys["V31"] * y1 + sys["V32"] * y2])


def apply_VH(sys, x):
    """V^* x (exp1/apply_VH_blockdiag.m:13-15)."""
    n = sys["NF"]
    x1, x2, x3 = x[:n], x[n:2 * n], x[2 * n:]
    return np.concatenate([sys["V11"] * x1 + sys["V21"] * x2 + sys["V31"] * x3,
                           sys["V12"] * x1 + sys["V22"] * x2 + sys["V32"] * x3])


def apply_K(sys, x):
    """Kop (exp1/assemble_K_and_nullspace_free_basis.m:110-119)."""
    n = sys["NF"]
    a, b, c = sys["q1"], sys["q2"], sys["q3"]
    x1, x2, x3 = x[:n], x[n:2 * n], x[2 * n:]
    return np.concatenate([(b * b + c * c) * x1 - a * b * x2 - a * c * x3,
                           -a * b * x1 + (a * a + c * c) * x2 - b * c * x3,
                           -a * c * x1 - b * c * x2 + (a * a + b * b) * x3])


def mass3_apply(eps, x, N):
    """mass.Mop: M = I3 (x) M_s on component-stacked vectors."""
    n = (2 * N) ** 6
    return np.concatenate([mass_apply(eps, x[c * n:(c + 1) * n], N) for c in range(3)])


def kr_apply(sys, eps, y, N):
    """K_r y = V^* M^{-1} V y with the closed-form inverse circulant (1/eps)."""
    n = sys["NF"]
    v = apply_V(sys, y)
    return apply_VH(sys, np.concatenate(
        [mass_inverse_apply(eps, v[c * n:(c + 1) * n], N) for c in range(3)]))


def solve_v1h(sys, q):
    """y = V1^{-*} q (exp1/solveV1H_blockdiag.m:7-8)."""
    n = sys["NF"]
    x1, x2 = q[:n], q[n:]
    d = np.conj(sys["det"])
    return np.concatenate([(np.conj(sys["V22"]) * x1 - np.conj(sys["V21"]) * x2) / d,
                           (-np.conj(sys["V12"]) * x1 + np.conj(sys["V11"]) * x2) / d])


def solve_v1(sys, x):
    """y = V1^{-1} x (exp1/solveV1_blockdiag.m:7-8)."""
    n = sys["NF"]
    x1, x2 = x[:n], x[n:]
    d = sys["det"]
    return np.concatenate([(sys["V22"] * x1 - sys["V12"] * x2) / d,
                           (-sys["V21"] * x1 + sys["V11"] * x2) / d])


def mhat_apply(sys, eps, x, N):
    """Mhat x = invC.*(M z + conj(W1).*M(W1.*z) + conj(W2).*M(W2.*z)), z = invC.*x
    (exp1/apply_Mhat_C_scalar_compact.m:8-21)."""
    z = sys["invC"] * x
    s = mass_apply(eps, z, N)
    s = s + np.conj(sys["W1"]) * mass_apply(eps, sys["W1"] * z, N)
    s = s + np.conj(sys["W2"]) * mass_apply(eps, sys["W2"] * z, N)
    return sys["invC"] * s


def pcg(A, b, tol, maxit, Minv=None):
    """MATLAB pcg with optional preconditioner z = Minv(r), x0 = 0 (call sites
    exp1/make_explicit_krinv_compact.m:43, exp2/recover_original_eigenvectors_exp2.m:21).
    Returns (x, flag, relres, iter).  Restates the recurrence, the true-residual
    confirmation and the 'moresteps' continuation; the stagnation test and the
    minimum-residual iterate on failure are not restated."""
    n = b.shape[0]
    n2b = np.linalg.norm(b)
    x = np.zeros_like(b, dtype=complex)
    if n2b == 0:
        return x, 0, 0.0, 0
    tolb = tol * n2b
    r = b.astype(complex).copy()
    normr = n2b
    if normr <= tolb:
        return x, 0, normr / n2b, 0
    rho = 1.0
    p = None
    moresteps = 0
    maxmsteps = min(n // 50, 5, n - maxit)
    flag, it = 1, maxit
    for ii in range(1, maxit + 1):
        z = r if Minv is None else Minv(r)
        rho1 = rho
        rho = np.vdot(r, z).real
        if rho == 0 or np.isinf(rho):
            flag, it = 4, ii - 1
            break
        if ii == 1:
            p = z.copy()
        else:
            beta = rho / rho1
            if beta == 0 or np.isinf(beta):
                flag, it = 4, ii - 1
                break
            p = z + beta * p
        qv = A(p)
        pq = np.vdot(p, qv).real
        if pq <= 0 or np.isinf(pq):
            flag, it = 4, ii - 1
            break
        alpha = rho / pq
        x = x + alpha * p
        r = r - alpha * qv
        normr = np.linalg.norm(r)
        if normr <= tolb or moresteps:
            r = b - A(x)
            normr = np.linalg.norm(r)
            if normr <= tolb:
                flag, it = 0, ii
                break
            moresteps += 1
            if moresteps >= maxmsteps:
                flag, it = 3, ii
                break
    return x, flag, normr / n2b, it


def krinv_apply(sys, eps, q, N, tol=1e-10, maxit=1200):
    """Method-C K_r^{-1} q (exp1/make_explicit_krinv_compact.m:24-64)."""
    n = sys["NF"]
    r = solve_v1h(sys, q)
    mr1 = mass_apply(eps, r[:n], N)
    mr2 = mass_apply(eps, r[n:], N)
    bhat = sys["invC"] * (-(np.conj(sys["W1"]) * mr1 + np.conj(sys["W2"]) * mr2))
    u, flag, relres, it = pcg(lambda v: mhat_apply(sys, eps, v, N), bhat, tol, maxit)
    if not np.all(np.isfinite(u)) or not np.isfinite(relres) or (flag != 0 and relres > 10 * tol):
        raise RuntimeError("Mhat CG failed: flag=%d relres=%.3e iter=%d" % (flag, relres, it))
    z = sys["invC"] * u
    mr1 = mr1 + mass_apply(eps, sys["W1"] * z, N)
    mr2 = mr2 + mass_apply(eps, sys["W2"] * z, N)
    return solve_v1(sys, np.concatenate([mr1, mr2])), dict(flag=flag, relres=relres, iter=it)


def krinv_nested(sys, eps, b, N, tol_kr=1e-10, tol_m=1e-12, maxit_kr=600, maxit_m=1600, epsC=4.0):
    """Method B: PCG on K_r x = b with preconditioner epsC*r./sigma_r, every K_r z =
    V^*(M^{-1}(V z)) by an inner PCG on M (exp1/make_nested_krinv_instrumented.m:33-132)."""
    sig = np.concatenate([sys["Qq"], sys["Qq"]])

    def Kr(z):
        mi, flag, rr, it = pcg(lambda v: mass3_apply(eps, v, N), apply_V(sys, z), tol_m, maxit_m)
        if flag != 0 and rr > 10 * tol_m:
            raise RuntimeError("Nested M solve failed")
        return apply_VH(sys, mi)

    x, flag, relres, it = pcg(Kr, b, tol_kr, maxit_kr, Minv=lambda r: epsC * r / sig)
    if flag != 0 and relres > 10 * tol_kr:
        raise RuntimeError("Nested K_r solve failed")
    return x, dict(flag=flag, relres=relres, iter=it)


# ----------------------------------------------------------------------------- Lanczos
def lanczos_hybrid(A, n, k, tol, max_restart, p, v0):
    """Thick-restart Lanczos for the largest eigenvalues of a Hermitian operator
    (exp1/Lanczos_hybrid_identity_gpuop.m:11-130), with the reference's column-by-column
    Gram-Schmidt (:132-136).  Returns (theta[k] descending, ritz vectors n x k, info)."""
    v = np.asarray(v0, dtype=complex).copy()
    nv = np.linalg.norm(v)
    if not (np.isfinite(nv) and nv > 0):
        raise ValueError("Invalid Lanczos start vector.")
    v /= nv
    Q = np.zeros((n, p), dtype=complex)
    k0 = k
    Alpha, Beta = [], []
    beta = 0.0
    d = np.zeros(0)
    c = np.zeros(0)
    just_restarted = False
    sizeV = 1
    kk = 0
    info = dict(converged=False, num_converged=0, hit_max_restart=False, restarts=0, steps=0)

    def gs(Qm, r, jj):
        for ii in range(jj):
            r = r - np.vdot(Qm[:, ii], r) * Qm[:, ii]
        return r

    Us = ind = None
    for mm in range(1, max_restart + 1):
        for jj in range(sizeV, p + 1):
            Q[:, jj - 1] = v
            Aq = A(Q[:, jj - 1])
            if not np.all(np.isfinite(Aq)):
                raise RuntimeError("Lanczos inverse action returned invalid data.")
            alpha = np.real(np.vdot(Aq, Q[:, jj - 1]))
            if jj == 1:
                r = Aq - alpha * Q[:, jj - 1]
            elif just_restarted:
                r = gs(Q, Aq, jj)
                just_restarted = False
            else:
                r = Aq - alpha * Q[:, jj - 1] - beta * Q[:, jj - 2]
                corr = np.vdot(Q[:, jj - 2], r)
                r = r - corr * Q[:, jj - 2]
            corr = np.vdot(Q[:, jj - 1], r)
            r = r - corr * Q[:, jj - 1]
            r = gs(Q, r, jj)
            r = gs(Q, r, jj)
            beta2 = np.real(np.vdot(r, r))
            if beta2 <= 0 or math.sqrt(beta2) < 1e-14:
                beta = math.sqrt(max(beta2, 0.0))
                break
            beta = math.sqrt(beta2)
            v = r / beta
            Alpha.append(alpha)
            Beta.append(beta)
            kk += 1
        na, nd = len(Alpha), len(d)
        m = nd + na
        if m == 0:
            raise RuntimeError("Lanczos produced no projected data.")
        H = np.zeros((m, m))
        H[:nd, :nd] = np.diag(d)
        H[nd:, nd:] = np.diag(Alpha) + (np.diag(Beta[:-1], 1) + np.diag(Beta[:-1], -1) if na > 1 else 0)
        if nd:
            H[:k, k] = c
            H[k, :k] = c
        Alpha, Beta = [], []
        w, U = np.linalg.eigh(H)
        res = np.abs(beta * U[-1, :])
        ind = np.argsort(-w, kind="stable")
        d = w[ind]
        res = res[ind]
        Us = U
        nw = min(k0, len(res))
        thr = tol * np.maximum(EPS ** (2.0 / 3.0), np.abs(d[:nw]))
        nconv = int(np.sum(res[:nw] < thr))
        info.update(num_converged=nconv, restarts=mm, steps=kk, max_ritz_residual=float(res[:nw].max()))
        if nconv >= k0:
            info["converged"] = True
            break
        if mm == max_restart:
            info["hit_max_restart"] = True
            break
        k = k0 + min(nconv, (p - k0) // 2)
        if k == 1 and p > 3:
            k = p // 2
        Uk = U[:, ind[:k]]
        Q[:, :k] = Q[:, :m] @ Uk
        d = d[:k]
        c = beta * Uk[-1, :]
        just_restarted = True
        sizeV = k + 1
    if ind is None or len(ind) < k0:
        raise RuntimeError("Lanczos ended without k inverse Ritz values.")
    theta = d[:k0].real
    Y = Q[:, :Us.shape[0]] @ Us[:, ind[:k0]]
    return theta, Y, info


def eigs_method_c(N, eps, q, v0, k=10, p=40, tol=1e-10, max_restart=50, tol_cg=1e-10,
                  maxit_cg=1200, P=P_REF):
    """run_method_C_fig1bc (exp1/run_method_C_fig1bc.m:4-21): lambda = sort(1./theta)."""
    sys = qsystem(N, q, P)
    nr = 2 * sys["NF"]
    stats = dict(calls=0, iters=0)

    def A(x):
        y, st = krinv_apply(sys, eps, x, N, tol_cg, maxit_cg)
        stats["calls"] += 1
        stats["iters"] += st["iter"]
        return y

    theta, Y, info = lanczos_hybrid(A, nr, k, tol, max_restart, p, v0)
    if not info["converged"]:
        raise RuntimeError("Method C outer Lanczos did not converge")
    if np.any(theta <= 0):
        raise RuntimeError("Method C invalid inverse Ritz values")
    order = np.argsort(1.0 / theta, kind="stable")
    info.update(stats)
    return (1.0 / theta)[order], Y[:, order], info, sys


# ----------------------------------------------------------------------------- Method A
def shift_precond_apply(sys, tau, x):
    """P_A^{-1} x with P_A = K + tau I, exactly per mode
    (exp1/build_original_shift_preconditioner_A_only.m:30-56):
    (K_q + tau I)^{-1} = I/(Q+tau) + (1/tau - 1/(Q+tau)) q q^*/Q, Q = |q|^2."""
    n = sys["NF"]
    a, b, c = sys["q1"], sys["q2"], sys["q3"]
    Q = a * a + b * b + c * c
    inv_t = 1.0 / (Q + tau)
    corr = (1.0 / tau - inv_t) / Q
    x1, x2, x3 = x[:n], x[n:2 * n], x[2 * n:]
    cqx = corr * (a * x1 + b * x2 + c * x3)
    return np.concatenate([inv_t * x1 + a * cqx, inv_t * x2 + b * cqx, inv_t * x3 + c * cqx])


def minres_call(A, b, tol, maxit, Minv=None, x0=None, checks=5):
    """One MATLAB minres(A, b, tol, maxit, P, [], x0) call as exp1/make_A_eigs_top_preconditioned.m
    makes it (:284-286 from x0, :619-620 from zero), for Hermitian A and HPD Minv = P^{-1}.
    Returns (x, flag, relres, iter) with relres = ||b - A x|| / ||b|| of the returned x.

    MATLAB's minres is closed source; this restates its documented behaviour.  A start that
    already meets tol returns flag 0 after no iterations.  The Paige-Saunders recurrence runs
    until its estimate |phibar| of the P^{-1}-norm residual reaches tol * beta1 (from zero) or
    tol * ||b||_{P^{-1}} (from x0); the explicit residual is then checked, and on a miss the
    iteration continues with a check after every step, `checks` checks in all, before flag 3.
    Flag 1: maxit ran out; flag 4: beta = 0 (breakdown) without meeting the check."""
    Minv = (lambda v: v) if Minv is None else Minv
    nb = np.linalg.norm(b)
    if nb == 0.0:
        return np.zeros_like(b), 0, 0.0, 0
    if x0 is None:
        x = np.zeros_like(b)
        r1 = b.copy()
        tolb = None
    else:
        x = x0.copy()
        r1 = b - A(x)
        rel = np.linalg.norm(r1) / nb
        if rel <= tol:
            return x, 0, rel, 0
        tolb = tol * math.sqrt(max(np.real(np.vdot(b, Minv(b))), 0.0))
    y = Minv(r1)
    beta1 = math.sqrt(max(np.real(np.vdot(r1, y)), 0.0))
    if tolb is None:
        tolb = tol * beta1
    oldb, beta, dbar, epsln, phibar = 0.0, beta1, 0.0, 0.0, beta1
    cs, sn = -1.0, 0.0
    w = np.zeros_like(b)
    w2 = np.zeros_like(b)
    r2 = r1.copy()
    done_checks = 0
    for it in range(1, maxit + 1):
        v = y / beta
        y = A(v)
        if it >= 2:
            y = y - (beta / oldb) * r1
        alfa = np.real(np.vdot(v, y))
        y = y - (alfa / beta) * r2
        r1, r2 = r2, y
        y = Minv(r2)
        oldb, beta = beta, math.sqrt(max(np.real(np.vdot(r2, y)), 0.0))
        oldeps = epsln
        delta = cs * dbar + sn * alfa
        gbar = sn * dbar - cs * alfa
        epsln = sn * beta
        dbar = -cs * beta
        gamma = max(math.hypot(gbar, beta), EPS)
        cs, sn = gbar / gamma, beta / gamma
        phi, phibar = cs * phibar, sn * phibar
        w1, w2 = w2, w
        w = (v - oldeps * w1 - delta * w2) / gamma
        x = x + phi * w
        if done_checks > 0 or abs(phibar) <= tolb or beta == 0.0:
            rel = np.linalg.norm(b - A(x)) / nb
            done_checks += 1
            if rel <= tol:
                return x, 0, rel, it
            if beta == 0.0:
                return x, 4, rel, it
            if done_checks >= checks:
                return x, 3, rel, it
    rel = np.linalg.norm(b - A(x)) / nb
    return x, (0 if rel <= tol else 1), rel, maxit


# cfg.shift recovery settings of exp1/make_config_exp1bc.m:50-72
SHIFT_RECOVERY = dict(max_recovery_events=50, max_corrections=8, max_zero_progress=2, min_progress=1e-8,
                      corr_safety=0.25, corr_max_tol=1e-1, corr_min_tol=1e-6, plateau_ceiling=1e-9,
                      plateau_min_events=2, plateau_reduction=1e-2, guard_ceiling=1e-9)


def shift_recovery_schedule(tol, call, residual, correct, log=None, cfg=None):
    """The solve loop of one Method-A shifted system (exp1/make_A_eigs_top_preconditioned.m:251-560)
    with its residual-correction cycle (:583-660).

    call(from_current) -> (flag, iter, relres): MINRES on A y = rhs from zero or from y;
    residual() -> relres of y, keeping r = rhs - A y; correct(corr_tol) -> (flag, iter, relres):
    MINRES on A d = r from zero, y += d, relres of the new y.  log(action) sees 'call',
    'stagnation', 'maxit' and 'correct'.  Returns a dict with 'end' in {'converged', 'plateau',
    'failed'}, the final 'relres' and the counters of the reference's info struct.  The two
    ceilings are raised to tol when tol exceeds them (the reference rejects that setting)."""
    c = dict(SHIFT_RECOVERY if cfg is None else cfg)
    c["guard_ceiling"] = max(c["guard_ceiling"], tol)
    c["plateau_ceiling"] = max(c["plateau_ceiling"], c["guard_ceiling"])
    log = log or (lambda a: None)
    info = dict(iterations=0, calls=0, stagnation_restarts=0, maxit_restarts=0, corrections=0,
                correction_iterations=0, recovery_events=0, plateau_events=0)
    prev = math.inf                # previousExplicitRelres
    final = math.inf               # finalExplicitRelres
    zero_events = 0                # zeroProgressEvents
    x0_current = False

    def end(kind, rel):
        return dict(info, end=kind, relres=rel)

    def plateau_update(rel, ref):
        info["plateau_events"] = info["plateau_events"] + 1 if rel >= ref * (1 - c["plateau_reduction"]) else 0

    def plateau_detected(rel):
        return info["plateau_events"] >= c["plateau_min_events"] and math.isfinite(rel) \
            and rel <= c["plateau_ceiling"]

    def near_guard(rel):
        return math.isfinite(rel) and rel <= c["guard_ceiling"]

    while True:
        if info["recovery_events"] > c["max_recovery_events"]:
            return end("plateau" if plateau_detected(final) or near_guard(final) else "failed", final)
        log("call")
        flag, it, rel = call(x0_current)
        info["calls"] += 1
        info["iterations"] += it
        final = rel
        if not math.isfinite(rel):
            return end("failed", rel)
        if flag == 0 or rel <= tol:
            return end("converged", rel)
        made_progress = math.isfinite(prev) and rel < prev * (1 - c["min_progress"])
        first_termination = not math.isfinite(prev)
        if flag == 3:
            if it > 0 and (first_termination or made_progress):
                if math.isfinite(prev):
                    plateau_update(rel, prev)
                    if plateau_detected(rel):
                        return end("plateau", rel)
                log("stagnation")
                info["stagnation_restarts"] += 1
                info["recovery_events"] += 1
                zero_events = 0
                prev = rel
                x0_current = True
                continue
            if info["corrections"] >= c["max_corrections"]:
                return end("plateau" if plateau_detected(rel) or near_guard(rel) else "failed", rel)
            log("correct")
            before = residual()
            if not math.isfinite(before):
                return end("failed", before)
            if before <= tol:
                cflag, cit, after = 0, 0, before
            else:
                ctol = c["corr_safety"] * tol / max(before, np.finfo(float).tiny)
                ctol = min(c["corr_max_tol"], max(c["corr_min_tol"], ctol))
                cflag, cit, after = correct(ctol)
            info["corrections"] += 1
            info["correction_iterations"] += cit
            info["calls"] += 1
            info["iterations"] += cit
            info["recovery_events"] += 1
            final = after
            if not math.isfinite(after):
                return end("failed", after)
            if after <= tol:
                return end("converged", after)
            plateau_update(after, rel)
            if plateau_detected(after):
                return end("plateau", after)
            if after < rel * (1 - c["min_progress"]):
                zero_events = 0
                prev = after
                x0_current = True
                continue
            zero_events += 1
            if zero_events >= c["max_zero_progress"]:
                return end("plateau" if plateau_detected(after) or near_guard(after) else "failed", after)
            prev = after
            x0_current = True
            continue
        if flag == 1:
            if it <= 0:
                return end("plateau" if near_guard(rel) else "failed", rel)
            if math.isfinite(prev):
                plateau_update(rel, prev)
                if plateau_detected(rel):
                    return end("plateau", rel)
            log("maxit")
            info["maxit_restarts"] += 1
            info["recovery_events"] += 1
            prev = rel
            x0_current = True
            continue
        return end("plateau" if near_guard(rel) else "failed", rel)


def method_a_top(sys, eps, x, N, sigma=1e-4, tol=1e-10, maxit=4000, info=None):
    """T x = (K - sigma M)^{-1} M x (exp1/make_A_eigs_top_preconditioned.m:59-60,98-113) with
    P_A = K + tau I, tau = sigma * mean(eps), solved under the reference's recovery schedule
    (shift_recovery_schedule).  Returns (y, iterations); `info` (a dict) receives the
    schedule's outcome.  A solve that ends above the acceptance ceilings raises."""
    tau = sigma * float(np.mean(eps))

    def A(v):
        return apply_K(sys, v) - sigma * mass3_apply(eps, v, N)

    def Pinv(v):
        return shift_precond_apply(sys, tau, v)

    b = mass3_apply(eps, x, N)
    nb = np.linalg.norm(b)
    st = {"y": np.zeros_like(b), "r": None}

    def call(from_current):
        st["y"], flag, rel, it = minres_call(A, b, tol, maxit, Pinv, x0=st["y"] if from_current else None)
        return flag, it, rel

    def residual():
        st["r"] = b - A(st["y"])
        return np.linalg.norm(st["r"]) / nb

    def correct(ctol):
        d, flag, _, it = minres_call(A, st["r"], ctol, maxit, Pinv)
        st["y"] = st["y"] + d
        return flag, it, np.linalg.norm(b - A(st["y"])) / nb

    if nb == 0.0:
        return st["y"], 0
    out = shift_recovery_schedule(tol, call, residual, correct)
    if info is not None:
        info.update(out)
    if out["end"] == "failed":
        raise RuntimeError("Method A MINRES did not converge: true relres %.3e" % out["relres"])
    return st["y"], out["iterations"]


def method_a_dense_lambda(sys, eps, N, sigma, k):
    """The k largest eigenvalues theta of the dense T = (K - sigma M)^{-1} M (small N), mapped
    to lambda = sigma + 1/theta ascending (exp1/run_method_A_fig1a.m:47-49)."""
    n3 = 3 * sys["NF"]
    I = np.eye(n3, dtype=complex)
    K = np.stack([apply_K(sys, I[:, j]) for j in range(n3)], axis=1)
    M = np.stack([mass3_apply(eps, I[:, j], N) for j in range(n3)], axis=1)
    theta = np.linalg.eigvals(np.linalg.solve(K - sigma * M, M))
    th = np.sort(np.real(theta))[::-1][:k]
    return np.sort(sigma + 1.0 / th)


def recover_eigenvectors(sys, eps, Y, N, tol=1e-12, maxit=3000):
    """x_j = M^{-1} V y_j by PCG, M-normalised (exp2/recover_original_eigenvectors_exp2.m:18-51)."""
    X, MX, its = [], [], []
    for j in range(Y.shape[1]):
        rhs = apply_V(sys, Y[:, j])
        x, flag, rr, it = pcg(lambda v: mass3_apply(eps, v, N), rhs, tol, maxit)
        if flag != 0 and rr > 10 * tol:
            raise RuntimeError("Recovery CG failed")
        mx = mass3_apply(eps, x, N)
        nj2 = np.real(np.vdot(x, mx))
        if not (np.isfinite(nj2) and nj2 > 0):
            raise RuntimeError("BadMassNorm")
        X.append(x / math.sqrt(nj2))
        MX.append(mx / math.sqrt(nj2))
        its.append(it)
    return np.stack(X, 1), np.stack(MX, 1), its


def fourier_diagnostics(sys, X, MX, lam):
    """GEVP residual, divergence and Rayleigh quotient (exp2/compute_fourier_diagnostics_exp2.m:10-19)."""
    n = sys["NF"]
    res, div, ray = [], [], []
    for j in range(X.shape[1]):
        x, mx = X[:, j], MX[:, j]
        kx = apply_K(sys, x)
        ray.append(np.real(np.vdot(x, kx) / np.vdot(x, mx)))
        den = np.linalg.norm(kx) + abs(lam[j]) * np.linalg.norm(mx)
        res.append(np.linalg.norm(kx - lam[j] * mx) / max(den, EPS))
        g = sys["q1"] * mx[:n] + sys["q2"] * mx[n:2 * n] + sys["q3"] * mx[2 * n:]
        div.append(np.linalg.norm(g) / max(np.linalg.norm(mx), EPS))
    return np.array(res), np.array(div), np.array(ray)


# ----------------------------------------------------------------------------- Taylor
def multi_index_list(p):
    """All alpha with |alpha| <= p, a1 outermost (exp2/build_multi_index_list_3d_exp3.m:3-14)."""
    out = []
    for a1 in range(p + 1):
        for a2 in range(p - a1 + 1):
            for a3 in range(p - a1 - a2 + 1):
                out.append((a1, a2, a3))
    return np.array(out, dtype=int)


def taylor_coeffs(p):
    """coeffAlpha = i^|alpha| * exp(-sum gammaln(alpha+1)) (exp3/prepare_taylor_fourier_basis_exp4.m:27-31)."""
    al = multi_index_list(p)
    return np.array([(1j) ** int(a.sum()) * math.exp(-sum(math.lgamma(v + 1) for v in a)) for a in al])


def taylor_moments(qmodes, coeff, p, centers):
    """moments(alpha, c) = coeffAlpha .* (Qalpha.' * (exp(i qModes*r_c) .* coeff))
    (exp3/reconstruct_component_taylor_batch_exp4.m:14-23, prepare_taylor_fourier_basis_exp4.m:12-24)."""
    al = multi_index_list(p)
    Qa = np.stack([qmodes[:, 0] ** a[0] * qmodes[:, 1] ** a[1] * qmodes[:, 2] ** a[2] for a in al], 1)
    phase = np.exp(1j * (qmodes @ np.asarray(centers).T))
    return taylor_coeffs(p)[:, None] * (Qa.T @ (phase * coeff[:, None]))


def taylor_evaluate(points, center_id, centers, moments, p):
    """u(pt) = sum_alpha (r - r_c)^alpha moments(alpha, c(pt))
    (exp3/reconstruct_component_taylor_batch_exp4.m:33-38, build_taylor_geometry_exp4.m:7-23)."""
    al = multi_index_list(p)
    R = points - np.asarray(centers)[center_id]
    u = np.zeros(len(points), dtype=complex)
    for ia, a in enumerate(al):
        u += R[:, 0] ** a[0] * R[:, 1] ** a[1] * R[:, 2] ** a[2] * moments[ia, center_id]
    return u


def direct_reconstruct(qmodes, coeff, points, block=256):
    """u = exp(i R q^T) c in point blocks (exp2/reconstruct_field_direct_blocked_exp3.m:10-15)."""
    u = np.zeros(len(points), dtype=complex)
    for s in range(0, len(points), block):
        u[s:s + block] = np.exp(1j * (points[s:s + block] @ qmodes.T)) @ coeff
    return u


def compute_field_errors(test, ref):
    """e2 = ||t - r||_2 / max(||r||_2, eps), eInf = max|t - r| / max(max|r|, eps) over the three
    components together (exp2/compute_field_errors_exp3.m:3-10)."""
    d = [np.asarray(t) - np.asarray(r) for t, r in zip(test, ref)]
    num2 = math.sqrt(sum(float(np.sum(np.abs(x) ** 2)) for x in d))
    den2 = math.sqrt(sum(float(np.sum(np.abs(np.asarray(r)) ** 2)) for r in ref))
    num_inf = max(float(np.max(np.abs(x))) for x in d)
    den_inf = max(float(np.max(np.abs(np.asarray(r)))) for r in ref)
    return dict(e2=num2 / max(den2, EPS), eInf=num_inf / max(den_inf, EPS))


def yee_points(n, h, comp):
    """Plus-grid Yee points of component comp (0,1,2): h*(I + 0.5 e_comp), I in -(n+1)..n+1
    (exp2/build_yee_edge_points_plus.m:11-17)."""
    idx = np.arange(-(n + 1), n + 2, dtype=float)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    pts = np.stack([I.reshape(-1, order="F"), J.reshape(-1, order="F"), K.reshape(-1, order="F")], 1)
    pts[:, comp] += 0.5
    return h * pts


def block_partition(m, nblocks):
    """nblocks^3 nearly equal contiguous blocks of an m^3 grid; first mod(m,nblocks) ranges one
    longer; block order i fastest (exp2/build_yee_block_partition_exp3.m:1-31).  Returns the
    block id of every grid point (column-major)."""
    base, rem = divmod(m, nblocks)
    lab = np.zeros(m, dtype=int)
    s = 0
    for b in range(nblocks):
        ln = base + (1 if b < rem else 0)
        lab[s:s + ln] = b
        s += ln
    I, J, K = np.meshgrid(lab, lab, lab, indexing="ij")
    bid = I + nblocks * J + nblocks * nblocks * K
    return bid.reshape(-1, order="F")


def block_centers(points, bid, nb3):
    """Center = midpoint of each block's bounding box (exp3/build_taylor_geometry_exp4.m:10-11)."""
    C = np.zeros((nb3, 3))
    for b in range(nb3):
        R = points[bid == b]
        C[b] = (R.min(0) + R.max(0)) / 2
    return C


# ----------------------------------------------------------------------------- adjoint
def mass_adjoint_eps(eps, x, gy, N):
    """d/d eps of Re<gy, M_s x> = NF * Re(conj(ifftn gy) .* ifftn x) (derivative of exp1/applyMij_fft.m:11)."""
    n = (2 * N) ** 6
    u = _ifftn(unvec(x, N))
    v = _ifftn(unvec(gy, N))
    return vec(n * np.real(np.conj(v) * u))


def eig_grad_eps(X, lam, N):
    """Hellmann-Feynman: d lambda / d eps_j = -lambda * NF * sum_c |ifftn(x_c)_j|^2 for M-normalised x."""
    n = (2 * N) ** 6
    out = []
    for j in range(X.shape[1]):
        g = np.zeros(grid_shape(N))
        for c in range(3):
            g += np.abs(_ifftn(unvec(X[c * n:(c + 1) * n, j], N))) ** 2
        out.append(vec(-lam[j] * n * g))
    return np.stack(out, 1)


# ----------------------------------------------------------------------------- Yee closure
# Fields on the plus grid Omega_n^+ are (2n+3)^3 arrays indexed [x, y, z] (the device
# layout is their column-major vectorisation, x fastest); the interior (2n+1)^3 is [1:-1]^3.
def _shift(u, dx, dy, dz):
    """u at interior point + (dx, dy, dz), each offset in {-1, 0, 1}."""
    sl = {-1: slice(0, -2), 0: slice(1, -1), 1: slice(2, None)}
    return u[sl[dx], sl[dy], sl[dz]]


def apply_cropped_A(u1, u2, u3, h):
    """Cropped Yee curl-curl action (exp3/apply_cropped_A_exp3.m:8-29), with the nine
    difference operators of :32-56 written as offsets of the interior point."""
    c = lambda u: _shift(u, 0, 0, 0)
    Dxx = lambda u: -_shift(u, -1, 0, 0) + 2 * c(u) - _shift(u, 1, 0, 0)
    Dyy = lambda u: -_shift(u, 0, -1, 0) + 2 * c(u) - _shift(u, 0, 1, 0)
    Dzz = lambda u: -_shift(u, 0, 0, -1) + 2 * c(u) - _shift(u, 0, 0, 1)
    DyT_DxP = lambda u: _shift(u, 1, -1, 0) - _shift(u, 0, -1, 0) - _shift(u, 1, 0, 0) + c(u)
    DzT_DxP = lambda u: _shift(u, 1, 0, -1) - _shift(u, 0, 0, -1) - _shift(u, 1, 0, 0) + c(u)
    DxT_DyP = lambda u: -_shift(u, -1, 0, 0) + c(u) + _shift(u, -1, 1, 0) - _shift(u, 0, 1, 0)
    DzT_DyP = lambda u: -_shift(u, 0, 0, -1) + _shift(u, 0, 1, -1) + c(u) - _shift(u, 0, 1, 0)
    DxT_DzP = lambda u: -_shift(u, -1, 0, 0) + c(u) + _shift(u, -1, 0, 1) - _shift(u, 0, 0, 1)
    DyT_DzP = lambda u: -_shift(u, 0, -1, 0) + c(u) + _shift(u, 0, -1, 1) - _shift(u, 0, 0, 1)
    h2inv = (1.0 / h) ** 2
    A1 = h2inv * (((Dyy(u1) + Dzz(u1)) - DyT_DxP(u2)) - DzT_DxP(u3))
    A2 = h2inv * (((Dxx(u2) + Dzz(u2)) - DxT_DyP(u1)) - DzT_DyP(u3))
    A3 = h2inv * (((Dxx(u3) + Dyy(u3)) - DxT_DzP(u1)) - DyT_DzP(u2))
    return A1, A2, A3


def sample_scalar_yee_material(P, epsC, alpha, h, n):
    """Scalar eps on the three Yee edge families of the (2n+1)^3 interior
    (exp3/sample_scalar_yee_material_exp4.m:4-21): eps_c = epsC exp(alpha/6 sum_l cos(P_l . x))
    with x shifted by h/2 along axis c."""
    idx = np.arange(-n, n + 1) * h
    out = []
    for comp in range(3):
        ax = [idx + (h / 2 if comp == d else 0.0) for d in range(3)]
        X, Y, Z = ax[0][:, None, None], ax[1][None, :, None], ax[2][None, None, :]
        S = np.zeros((2 * n + 1,) * 3)
        for l in range(P.shape[0]):
            S = S + np.cos(P[l, 0] * X + P[l, 1] * Y + P[l, 2] * Z)
        out.append(epsC * np.exp(alpha / 6.0 * S))
    return out


def lwrq_kernel(h, R=3, sigma_factor=2.0):
    """k = exp(-(off/sigma)^2), off = (-R:R) h, sigma = sigma_factor h (exp3/compute_lwrq_exp4.m:12-15)."""
    off = np.arange(-R, R + 1) * h
    sigma = sigma_factor * h
    return np.exp(-(off ** 2) / sigma ** 2)


def separable_conv3_same(X, k):
    """Zero-padded separable 3D convolution, same size (exp3/separable_conv3_same_exp4.m:3-6)."""
    R = (len(k) - 1) // 2
    Y = X
    for axis in range(3):
        Z = np.zeros_like(Y)
        m = Y.shape[axis]
        for t in range(-R, R + 1):
            # Z[i] += k[t + R] * Y[i - t] (convn 'same' with the centred kernel)
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            lo, hi = max(0, t), min(m, m + t)
            if lo >= hi:
                continue
            dst[axis] = slice(lo, hi)
            src[axis] = slice(lo - t, hi - t)
            Z[tuple(dst)] += k[t + R] * Y[tuple(src)]
        Y = Z
    return Y


def compute_lwrq(a, b, h, R=3, sigma_factor=2.0, mass_floor_rel=1e-14, return_rho=False):
    """Column-normalised local Gaussian Rayleigh quotient (exp3/compute_lwrq_exp4.m:12-56)."""
    k = lwrq_kernel(h, R, sigma_factor)
    den = separable_conv3_same(np.ones_like(a), k)
    if not (np.all(np.isfinite(den)) and np.all(den > 0)):
        raise ValueError("Invalid LWRQ column-normalization denominator.")
    E = separable_conv3_same(a / den, k)
    M = separable_conv3_same(b / den, k)
    maxM = M.max()
    if not np.all(np.isfinite(M)) or np.any(M <= mass_floor_rel * maxM):
        raise ValueError("Nonpositive or numerically tiny local LWRQ mass encountered.")
    sumE, sumM = E.sum(), M.sum()
    lam = sumE / sumM
    rho = E / M
    V = np.sum(M * np.abs(rho - lam) ** 2) / (sumM * abs(lam) ** 2)
    sumA, sumB = a.sum(), b.sum()
    e_part = abs(sumE - sumA) / max(abs(sumA), EPS)
    m_part = abs(sumM - sumB) / max(abs(sumB), EPS)
    return dict(lambda_=lam, V=float(np.real(V)), nu=math.sqrt(max(float(np.real(V)), 0.0)),
                partition_error=max(e_part, m_part), energy_partition_error=e_part,
                mass_partition_error=m_part, rho=rho if return_rho else None)


def compute_yee_lwrq_metrics(u1, u2, u3, lambda6d, h, n, P=P_REF, epsC=4.0, alpha=1.2, R=3,
                             sigma_factor=2.0, mass_floor_rel=1e-14, return_rho=False):
    """compute_yee_lwrq_metrics_exp4 (exp3/compute_yee_lwrq_metrics_exp4.m:16-70) on plus-grid fields."""
    m = 2 * n + 3
    for u in (u1, u2, u3):
        if u.shape != (m, m, m):
            raise ValueError("fieldPlus size does not match n")
    A1, A2, A3 = apply_cropped_A(u1, u2, u3, h)
    v1, v2, v3 = (_shift(u, 0, 0, 0) for u in (u1, u2, u3))
    e1, e2, e3 = sample_scalar_yee_material(P, epsC, alpha, h, n)
    B1, B2, B3 = e1 * v1, e2 * v2, e3 * v3
    a = np.real(np.conj(v1) * A1 + np.conj(v2) * A2 + np.conj(v3) * A3)
    b = np.real(np.conj(v1) * B1 + np.conj(v2) * B2 + np.conj(v3) * B3)
    sumB = b.sum()
    if sumB <= 0:
        raise ValueError("Cropped Yee mass is not positive.")
    lam_crop = a.sum() / sumB
    L = compute_lwrq(a, b, h, R, sigma_factor, mass_floor_rel, return_rho)
    lam3 = float(np.real(L["lambda_"]))
    d_id = abs(lam3 - lam_crop) / max(abs(lam_crop), EPS)
    nA = math.sqrt(sum(np.sum(np.abs(x) ** 2) for x in (A1, A2, A3)))
    nB = math.sqrt(sum(np.sum(np.abs(x) ** 2) for x in (B1, B2, B3)))
    res = lambda lam: sum(np.sum(np.abs(A - lam * B) ** 2) for A, B in ((A1, B1), (A2, B2), (A3, B3)))
    eta3 = math.sqrt(res(lam3)) / (nA + abs(lam3) * nB)
    eta6 = math.sqrt(res(lambda6d)) / (nA + abs(lambda6d) * nB)
    return dict(lambda_crop=float(lam_crop), lambda_3d=lam3, eta_3d=eta3, eta_6d=eta6,
                delta_lambda=abs(lam3 - lambda6d) / abs(lambda6d), V=L["V"], nu=L["nu"],
                delta_part=L["partition_error"], delta_identity=d_id, cropped_mass=float(sumB),
                energy_partition_error=L["energy_partition_error"],
                mass_partition_error=L["mass_partition_error"], rho=L["rho"])


def crop_plus_field(u, n_large, n_small):
    """Centred crop of a same-h plus-grid field (exp3/crop_plus_field_exp4.m:3-6)."""
    if n_small > n_large:
        raise ValueError("nSmall must not exceed nLarge")
    c = n_large + 1  # 0-based index of physical grid index zero
    r = slice(c - (n_small + 1), c + (n_small + 1) + 1)
    return u[r, r, r]
