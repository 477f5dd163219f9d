"""Ranked-basis products, their adjoint and the differentiable projector (through the C ABI).

Bars: basis bit-exact; H.v within 1e-12 relative of the oracle (different summation order);
adjoint gradients within 1e-6 relative of central finite differences (h = 1e-5); the
Hellmann-Feynman value <psi0|dH/du|psi0> of exact diagonalisation at convergence.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle as O

pytestmark = pytest.mark.gpu


def make(rimu, M, N, us, ts):
    addr = rimu.near_uniform(N, M)
    hs = [rimu.HubbardReal1D(addr, u=u, t=t) for u, t in zip(us, ts)]
    return rimu.DenseHamiltonians(hs), hs


def cpu_ops(M, N, u, t, basis):
    """(H, D, K) as CPU sparse matrices from the oracle: H = u D + t K."""
    H = O.sparse_hamiltonian(O.real1d(M, N, u, t), basis)
    Dm = O.sparse_hamiltonian(O.real1d(M, N, 1.0, 0.0), basis)
    Km = O.sparse_hamiltonian(O.real1d(M, N, 0.0, 1.0), basis)
    return H, Dm, Km


def test_basis_and_apply(gpu):
    M = N = 7
    us, ts = [1.0, 4.5, 9.0], [1.0, 0.7, 1.3]
    d, _ = make(gpu, M, N, us, ts)
    basis = d.basis()
    assert np.array_equal(basis, O.build_basis(M, N))
    assert d.dim == gpu.basis_size(M, N) == len(basis)
    rng = np.random.default_rng(0)
    v = rng.normal(size=(d.dim, 3))
    a, b = np.array([0.3, 1.0, -2.0]), np.array([1.0, -0.01, 0.5])
    w = d.apply(a, b, v)
    models = [O.real1d(M, N, u, t) for u, t in zip(us, ts)]
    wo = O.apply_affine(models, basis, a, b, v)
    np.testing.assert_allclose(w, wo, rtol=1e-12, atol=1e-12 * np.abs(wo).max())


def test_pullback_matches_explicit_matrices(gpu):
    M, N = 6, 5
    us, ts = [2.0, 6.0], [1.0, 0.5]
    d, _ = make(gpu, M, N, us, ts)
    basis = d.basis()
    rng = np.random.default_rng(1)
    v = rng.normal(size=(d.dim, 2))
    ybar = rng.normal(size=(d.dim, 2))
    a, b = np.array([1.01, 0.0]), np.array([-0.01, 1.0])
    vbar, ubar, tbar = d.pullback(a, b, v, ybar)
    for i in range(2):
        H, Dm, Km = cpu_ops(M, N, us[i], ts[i], basis)
        np.testing.assert_allclose(vbar[:, i], a[i] * ybar[:, i] + b[i] * (H.T @ ybar[:, i]), rtol=1e-12, atol=1e-12)
        assert ubar[i] == pytest.approx(b[i] * ybar[:, i] @ (Dm @ v[:, i]), rel=1e-12)
        assert tbar[i] == pytest.approx(b[i] * ybar[:, i] @ (Km @ v[:, i]), rel=1e-12)
    # the rrule of mul!: value and pullback of y = H v
    y, pb = d.rrule_mul(v)
    np.testing.assert_allclose(y, d.apply(0.0, 1.0, v), rtol=0, atol=0)
    vb, ub, tb = pb(ybar)
    np.testing.assert_allclose(vb, d.apply(0.0, 1.0, ybar), rtol=1e-13, atol=1e-13)


def cpu_projector_energy(M, N, u, t, basis, v0, K, dtau, S):
    H = O.sparse_hamiltonian(O.real1d(M, N, u, t), basis)
    v = v0.copy()
    for _ in range(K):
        v = v + dtau * (S * v - H @ v)
    return (v @ (H @ v)) / (v @ v)


@pytest.mark.parametrize("ckpt", [1, 7, 17, None])
def test_projector_gradient_against_finite_differences(gpu, ckpt):
    M = N = 6
    us = np.array([2.0, 4.0, 6.0, 8.0])
    ts = np.ones(4)
    K, dtau = 120, 0.01
    d, hs = make(gpu, M, N, us, ts)
    basis = d.basis()
    v0 = d.unit_vector(hs[0].starting_address())
    shift = np.zeros(4)
    res = d.projector_energy(K, dtau, shift, v0, ckpt=ckpt)
    h = 1e-5
    up, _ = make(gpu, M, N, us + h, ts)
    dn, _ = make(gpu, M, N, us - h, ts)
    fd_u = (up.projector_energy(K, dtau, shift, v0)["energy"] -
            dn.projector_energy(K, dtau, shift, v0)["energy"]) / (2 * h)
    tp, _ = make(gpu, M, N, us, ts + h)
    tn, _ = make(gpu, M, N, us, ts - h)
    fd_t = (tp.projector_energy(K, dtau, shift, v0)["energy"] -
            tn.projector_energy(K, dtau, shift, v0)["energy"]) / (2 * h)
    np.testing.assert_allclose(res["grad_u"], fd_u, rtol=1e-6)
    np.testing.assert_allclose(res["grad_t"], fd_t, rtol=1e-6)
    for i in range(4):
        # energy against an independent CPU projector, and the CPU's own finite difference
        e_cpu = cpu_projector_energy(M, N, us[i], 1.0, basis, v0[:, i], K, dtau, 0.0)
        assert res["energy"][i] == pytest.approx(e_cpu, rel=1e-12)
        if ckpt == 7:
            g_cpu = (cpu_projector_energy(M, N, us[i] + h, 1.0, basis, v0[:, i], K, dtau, 0.0) -
                     cpu_projector_energy(M, N, us[i] - h, 1.0, basis, v0[:, i], K, dtau, 0.0)) / (2 * h)
            assert res["grad_u"][i] == pytest.approx(g_cpu, rel=1e-6)


def test_config5_gradients_full_size(gpu):
    """The differentiable projector at its own size: HubbardReal1D M = N = 12 (dimension
    1,352,078), u in {2, 4, 6, 8}, t = 1, K = 2000 steps of dtau 1e-3 from the near-uniform
    address, checkpoints every 45 steps (45 recomputed segments).  dE/du and dE/dt of the
    Rayleigh quotient (Hamiltonians/abstract.jl:153) against central differences of the
    forward projector at h = 1e-5 (rtol 1e-6); another checkpoint spacing (200 steps) gives
    the same energies and gradients to 1e-12."""
    M = N = 12
    us = np.array([2.0, 4.0, 6.0, 8.0])
    ts = np.ones(4)
    K, dtau, h = 2000, 1e-3, 1e-5
    d, hs = make(gpu, M, N, us, ts)
    assert d.dim == 1352078
    v0 = d.unit_vector(hs[0].starting_address())
    shift = np.zeros(4)
    print("config5 adjoint K=2000 ckpt=45", flush=True)
    res = d.projector_energy(K, dtau, shift, v0, ckpt=45)
    res2 = d.projector_energy(K, dtau, shift, v0, ckpt=200)
    d.close()
    for f in ("energy", "grad_u", "grad_t"):
        np.testing.assert_allclose(res2[f], res[f], rtol=1e-12)

    def energy(u, t):
        e, _ = make(gpu, M, N, u, t)
        out = e.projector_energy(K, dtau, shift, v0, ckpt=45)["energy"]
        e.close()
        return out

    print("config5 central differences", flush=True)
    fd_u = (energy(us + h, ts) - energy(us - h, ts)) / (2 * h)
    fd_t = (energy(us, ts + h) - energy(us, ts - h)) / (2 * h)
    np.testing.assert_allclose(res["grad_u"], fd_u, rtol=1e-6)
    np.testing.assert_allclose(res["grad_t"], fd_t, rtol=1e-6)
    # E = u <D> + t <K> of the final vector (H is linear in u and t)
    np.testing.assert_allclose(res["energy"], us * res["hf_u"] + ts * res["hf_t"], rtol=1e-12)


def test_projector_converges_to_hellmann_feynman(gpu):
    """At convergence dE/du is <psi0|D|psi0> and E is the ED ground state energy."""
    M = N = 5
    us = np.array([1.0, 6.0])
    d, hs = make(gpu, M, N, us, [1.0, 1.0])
    basis = d.basis()
    v0 = d.unit_vector(hs[0].starting_address())
    res = d.projector_energy(4000, 0.01, np.zeros(2), v0)
    for i, u in enumerate(us):
        H, Dm, Km = cpu_ops(M, N, u, 1.0, basis)
        w, V = np.linalg.eigh(H.toarray())
        psi = V[:, 0]
        assert res["energy"][i] == pytest.approx(w[0], abs=1e-9)
        assert res["grad_u"][i] == pytest.approx(psi @ (Dm @ psi), rel=1e-7)
        assert res["grad_t"][i] == pytest.approx(psi @ (Km @ psi), rel=1e-7)
        assert res["hf_u"][i] == pytest.approx(psi @ (Dm @ psi), rel=1e-7)
    # known answer of the reference: HubbardReal1D 5x5 u=1 (test/lomc.jl:546)
    assert res["energy"][0] == pytest.approx(-8.280991746582686, abs=1e-9)


def test_full_size_config_apply(gpu):
    """M = N = 12 (dimension 1,352,078), four u at once: one product against the oracle."""
    M = N = 12
    us = [2.0, 4.0, 6.0, 8.0]
    d, hs = make(gpu, M, N, us, [1.0] * 4)
    assert d.dim == 1352078
    basis = d.basis()
    rng = np.random.default_rng(7)
    v = rng.normal(size=(d.dim, 4))
    a, b = np.full(4, 1.0), np.full(4, -1e-3)
    w = d.apply(a, b, v)
    wo = O.apply_affine([O.real1d(M, N, u, 1.0) for u in us], basis, a, b, v)
    np.testing.assert_allclose(w, wo, rtol=1e-12, atol=1e-12 * np.abs(wo).max())


def test_ground_state_known_answers(gpu):
    """Lanczos on the device against the reference's ED constants: HubbardReal1D 5x5 u=1
    (test/lomc.jl:546), 7x7 u=6 (test/mpi_runtests.jl:160)."""
    for (M, N, u), E0 in [((5, 5, 1.0), -8.280991746582686), ((7, 7, 6.0), -4.628524493494574)]:
        d, _ = make(gpu, M, N, [u], [1.0])
        e, it = d.ground_state()
        assert e[0] == pytest.approx(E0, abs=1e-9), (M, N, u, it)
        d.close()


def test_ground_state_against_scipy(gpu):
    """Four (u, t) at once, and the M = N = 10, u = 6 case (dimension 92,378) against the
    oracle's sparse matrix and scipy eigsh."""
    M, N = 8, 5
    us, ts = [1.0, 3.0, 6.5, 10.0], [1.0, 0.5, 1.0, 2.0]
    d, _ = make(gpu, M, N, us, ts)
    e, it = d.ground_state()
    basis = d.basis()
    for i in range(4):
        ref = O.ed_energies(O.real1d(M, N, us[i], ts[i]), basis, 1)[0]
        assert e[i] == pytest.approx(ref, abs=1e-9), (i, it)
    d.close()
    d, _ = make(gpu, 10, 10, [6.0], [1.0])
    assert d.dim == 92378
    e, it = d.ground_state()
    ref = O.ed_energies(O.real1d(10, 10, 6.0, 1.0), d.basis(), 1)[0]
    assert e[0] == pytest.approx(ref, abs=1e-9), it
    d.close()


def test_ground_state_exhausted_krylov_space(gpu):
    """Dimension 6 (M = 3, N = 2): the symmetric Krylov space of the start address ends
    after a few vectors; the Ritz value is then exact."""
    d, _ = make(gpu, 3, 2, [2.0], [1.0])
    e, it = d.ground_state(max_iters=100)
    w = np.linalg.eigvalsh(O.sparse_hamiltonian(O.real1d(3, 2, 2.0, 1.0), d.basis()).toarray())
    assert e[0] == pytest.approx(w[0], abs=1e-12)
    assert it <= 6
