"""GPU parity of the mass operator and its wrappers against the CPU oracle."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.fixture(scope="module")
def dev():
    torch.cuda.set_device(0)
    return torch.device("cuda", 0)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_mscalar_fp64_fp32(pkg, oracle, dev, N):
    print("case: MscalarOp N=%d" % N, flush=True)
    rng = np.random.default_rng(100 + N)
    eps = oracle.eps_reference(N)
    m = pkg.MassBackend(N, eps)
    x = rand_c(rng, m.NF)
    ref = oracle.mass_apply(eps, x, N)
    y = m.MscalarOp(torch.as_tensor(x, device=dev)).cpu().numpy()
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13
    y32 = m.MscalarOp(torch.as_tensor(x, device=dev).to(torch.complex64)).cpu().numpy()
    assert np.linalg.norm(y32 - ref) / np.linalg.norm(ref) < 2e-6
    for va in range(8):
        for vb in range(4):
            m.set_tiles("double", va, vb)
            y2 = m.MscalarOp(torch.as_tensor(x, device=dev)).cpu().numpy()
            assert np.linalg.norm(y2 - ref) / np.linalg.norm(ref) < 1e-13, (va, vb)
            m.set_tiles("single", va, vb)
            y3 = m.MscalarOp(torch.as_tensor(x, device=dev).to(torch.complex64)).cpu().numpy()
            assert np.linalg.norm(y3 - ref) / np.linalg.norm(ref) < 2e-6, (va, vb)
    m.close()


def test_wrappers_multi_rhs(pkg, oracle, dev):
    N = 2
    print("case: Mop/Mcols12/Mcol3 N=%d" % N, flush=True)
    rng = np.random.default_rng(7)
    eps = oracle.eps_reference(N)
    m = pkg.MassBackend(N, eps)
    NF = m.NF
    X = rand_c(rng, 4, NF)
    Y = m.MscalarOp(torch.as_tensor(X, device=dev)).cpu().numpy()
    for j in range(4):
        assert np.allclose(Y[j], oracle.mass_apply(eps, X[j], N), atol=1e-12)
    U = rand_c(rng, 2, 3 * NF)
    MU = m.Mop(torch.as_tensor(U, device=dev)).cpu().numpy()
    for j in range(2):
        assert np.allclose(MU[j], oracle.mass3_apply(eps, U[j], N), atol=1e-12)
    x12 = rand_c(rng, 2, 2 * NF)
    Y12 = m.Mcols12op(torch.as_tensor(x12, device=dev)).cpu().numpy()
    for j in range(2):
        ref = np.concatenate([oracle.mass_apply(eps, x12[j, :NF], N), oracle.mass_apply(eps, x12[j, NF:], N),
                              np.zeros(NF)])
        assert np.allclose(Y12[j], ref, atol=1e-12)
    x3 = rand_c(rng, 3, NF)
    Y3 = m.Mcol3op(torch.as_tensor(x3, device=dev)).cpu().numpy()
    for j in range(3):
        ref = np.concatenate([np.zeros(2 * NF), oracle.mass_apply(eps, x3[j], N)])
        assert np.allclose(Y3[j], ref, atol=1e-12)
    with pytest.raises(pkg.Fame6dError):
        m.MscalarOp(torch.zeros(NF + 1, dtype=torch.complex128, device=dev))
    m.close()


def test_constant_eps_identity(pkg, dev):
    N = 3
    print("case: constant eps N=%d" % N, flush=True)
    m = pkg.MassBackend(N, np.full(6 ** 6, 2.5))
    x = torch.randn(m.NF, dtype=torch.complex128, device=dev)
    y = m.MscalarOp(x)
    assert torch.allclose(y, 2.5 * x, atol=1e-13)
    m.close()


def test_adjoint_eps(pkg, oracle, dev):
    N = 2
    print("case: adjoint eps N=%d" % N, flush=True)
    rng = np.random.default_rng(8)
    eps = oracle.eps_reference(N)
    m = pkg.MassBackend(N, eps)
    x = rand_c(rng, 2, m.NF)
    gy = rand_c(rng, 2, m.NF)
    g = m.adjoint_eps(torch.as_tensor(x, device=dev), torch.as_tensor(gy, device=dev)).cpu().numpy()
    ref = oracle.mass_adjoint_eps(eps, x[0], gy[0], N) + oracle.mass_adjoint_eps(eps, x[1], gy[1], N)
    assert np.linalg.norm(g - ref) / np.linalg.norm(ref) < 1e-12
    m.close()


def test_rejects_nonpositive_eps(pkg, dev):
    with pytest.raises(pkg.Fame6dError) as e:
        pkg.MassBackend(1, -np.ones(64))
    assert e.value.name == "EARG"
