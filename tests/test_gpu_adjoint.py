"""Adjoint gradient of piecewise-constant controls (``dynamiqs_amd.adjoint``) on
the GPU against central finite differences of the complex128 oracle (h = 1e-4,
tight tolerances) and the oracle's loss. The reference's jax.grad cannot run on
this platform (no JAX); the FD check replaces it."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(n=8, nseg=20, T=2.0, B=3, seed=0):
    from oracle import operators as op

    rng = np.random.default_rng(seed)
    a = op.destroy(n)
    H0 = a.conj().T @ a
    X = a + a.conj().T
    times = np.linspace(0.0, T, nseg + 1)
    u = rng.normal(0.0, 0.5, size=(B, 1, nseg))
    Ls = [np.sqrt(0.05) * a]
    rho0 = op.fock_dm(n, 0)
    O = a.conj().T @ a
    return H0, [X], times, u, Ls, rho0, O


def _oracle_loss(H0, ops, times, ub, Ls, rho0, O):
    from oracle import mesolve as om

    Hop = om.TimeOp(H0, [om.PWCTerm(times, ub[j], ops[j]) for j in range(len(ops))])
    r = om.mesolve(Hop, Ls, rho0, np.array([times[0], times[-1]]), exp_ops=[O], rtol=1e-10, atol=1e-12)
    return float(r['expects'][0, -1].real)


@pytest.mark.parametrize('every', [10, 5])
def test_adjoint_matches_central_fd(every):
    from dynamiqs_amd.adjoint import mesolve_pwc_grad

    H0, ops, times, u, Ls, rho0, O = _problem()
    J, g = mesolve_pwc_grad(H0, ops, times, u, Ls, rho0, O, rtol=1e-8, atol=1e-10, checkpoint_every=every)
    h = 1e-4
    for b in range(u.shape[0]):
        J_ref = _oracle_loss(H0, ops, times, u[b], Ls, rho0, O)
        assert abs(J[b] - J_ref) < 1e-5
        fd = np.zeros(u.shape[2])
        for k in range(u.shape[2]):
            up, um = u[b].copy(), u[b].copy()
            up[0, k] += h
            um[0, k] -= h
            fd[k] = (_oracle_loss(H0, ops, times, up, Ls, rho0, O) - _oracle_loss(H0, ops, times, um, Ls, rho0, O)) / (2 * h)
        err = np.abs(g[b, 0] - fd).max()
        assert err <= 1e-4 * np.abs(fd).max() + 1e-6, (b, err, np.abs(fd).max())


def test_adjoint_closed_form_constant_drive():
    """Undriven cavity from a coherent state: with u = 0 the gradient of
    n(T) = Tr(a^dag a rho(T)) w.r.t. a displacement drive on segment k has the closed form
    dn/du_k = 2 Re( alpha(T)^* d alpha(T)/du_k ), d alpha/du_k = -i int_k e^{-(i w + kappa/2)(T - t)} dt."""
    from dynamiqs_amd.adjoint import mesolve_pwc_grad
    from oracle import operators as op

    n, w, kappa, T, nseg = 24, 1.0, 0.2, 2.0, 8
    a = op.destroy(n)
    times = np.linspace(0, T, nseg + 1)
    al0 = 1.0
    rho0 = op.coherent_dm(n, al0)
    u = np.zeros((1, 1, nseg))
    J, g = mesolve_pwc_grad(w * a.conj().T @ a, [a + a.conj().T], times, u, [np.sqrt(kappa) * a], rho0,
                            a.conj().T @ a, rtol=1e-8, atol=1e-10, checkpoint_every=4)
    gam = 1j * w + kappa / 2
    alT = al0 * np.exp(-gam * T)
    assert abs(J[0] - abs(alT) ** 2) < 1e-5
    want = []
    for k in range(nseg):
        t0, t1 = times[k], times[k + 1]
        # H = w a^dag a + u (a + a^dag): d alpha/dt = -gam alpha - i u
        integ = (np.exp(-gam * (T - t1)) - np.exp(-gam * (T - t0))) / gam
        want.append(2 * np.real(np.conj(alT) * (-1j) * integ))
    np.testing.assert_allclose(g[0, 0], want, atol=2e-5)


def test_readme_gradient_kat():
    """The reference README's gradient example (README.md:85-121): n = 16, omega = 1,
    kappa = 0.1, alpha0 = 1, T = 2 pi, J = <a^dag a>(T). Printed gradients:
    d/domega 0.0000, d/dkappa -3.3520, d/dalpha0 1.0670; closed form
    (0, -alpha0^2 T e^{-kappa T}, 2 alpha0 e^{-kappa T}).

    omega enters as a piecewise-constant control held at 1 over 40 segments (dJ/domega
    is the sum over segments), kappa through the rate gradient of L = sqrt(kappa) a,
    alpha0 through lambda(0) = dJ/drho0 contracted with d rho0 / d alpha0."""
    from dynamiqs_amd.adjoint import mesolve_adjoint
    from oracle import operators as op

    n, omega, kappa, alpha0 = 16, 1.0, 0.1, 1.0
    T = 2 * np.pi
    a = op.destroy(n)
    N = a.conj().T @ a
    times = np.linspace(0.0, T, 41)
    u = np.full((1, 1, 40), omega)
    res = mesolve_adjoint(
        np.zeros((n, n), dtype=complex), [N], times, u, [np.sqrt(kappa) * a], op.coherent_dm(n, alpha0), N,
        rate_ops=[a], checkpoint_every=4,
    )
    h = 1e-5
    drho = (op.coherent_dm(n, alpha0 + h) - op.coherent_dm(n, alpha0 - h)) / (2 * h)
    g_omega = float(res.grad_u[0, 0].sum())
    g_kappa = float(res.grad_rates[0, 0])
    g_alpha = float(np.real(np.trace(res.lambda0[0].astype(complex) @ drho)))
    decay = np.exp(-kappa * T)
    assert abs(res.J[0] - alpha0**2 * decay) <= 1e-5
    assert abs(g_omega) <= 1e-4, g_omega
    assert abs(g_kappa - (-alpha0**2 * T * decay)) <= 2e-4, g_kappa
    assert abs(g_alpha - 2 * alpha0 * decay) <= 2e-4, g_alpha
    # the README's printed values (4 decimals)
    assert abs(g_kappa - (-3.3520)) <= 1.5e-4 and abs(g_alpha - 1.0670) <= 1.5e-4


def test_dia_quadrature_equals_dense_quadrature():
    """``dqh_adjoint_pwc_quadrature_dia`` (control operators by diagonals) and the
    dense ``dqh_adjoint_pwc_quadrature`` give the same gradients on random node
    states, for operators with 1, 2 and 3 diagonals (n = 32, 3 instances, 4
    segments of 5 nodes), and both equal the FP64 NumPy integrand sum."""
    import ctypes

    import torch

    from dynamiqs_amd import _native
    from dynamiqs_amd.adjoint import _sparse_controls

    rng = np.random.default_rng(11)
    n, B, nseg, nodes = 32, 3, 4, 5
    Q = nseg * nodes
    a = np.diag(np.sqrt(np.arange(1, n)), k=1).astype(complex)
    ops = [a + a.conj().T, a.conj().T @ a, 0.3 * a @ a + 0.2 * a.conj().T + np.diag(rng.normal(size=n))]

    def herm(k):
        x = rng.normal(size=(k, n, n)) + 1j * rng.normal(size=(k, n, n))
        return (x + np.conj(np.swapaxes(x, -1, -2))) / 2

    rho = np.ascontiguousarray(herm(B * Q).reshape(B, Q, n, n), dtype=np.complex64)
    lam = np.ascontiguousarray(herm(B * Q).reshape(B, Q, n, n), dtype=np.complex64)
    w = rng.uniform(0.1, 1.0, size=Q)
    sf = (np.arange(nseg + 1) * nodes).astype(np.int32)
    dev = torch.device('cuda', 0)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    r_t, l_t = t(rho.view(np.float32)), t(lam.view(np.float32))
    w_t, sf_t = t(w), t(sf)
    g_dense = torch.zeros((B, len(ops), nseg), dtype=torch.float64, device=dev)
    g_dia = torch.zeros_like(g_dense)
    lib = _native.lib()
    P = ctypes.c_void_p
    o_t = t(np.stack(ops).astype(np.complex64).view(np.float32))
    _native.check(lib.dqh_adjoint_pwc_quadrature(B, n, Q, len(ops), nseg, P(r_t.data_ptr()), P(l_t.data_ptr()),
                                                 P(o_t.data_ptr()), P(w_t.data_ptr()), P(sf_t.data_ptr()),
                                                 P(g_dense.data_ptr()), None), 'dense')
    dops, keep = _sparse_controls(ops, n)
    _native.check(lib.dqh_adjoint_pwc_quadrature_dia(B, n, Q, nseg, P(r_t.data_ptr()), P(l_t.data_ptr()),
                                                     ctypes.byref(dops), P(w_t.data_ptr()), P(sf_t.data_ptr()),
                                                     P(g_dia.data_ptr()), None), 'dia')
    torch.cuda.synchronize()
    gd, ga = g_dense.cpu().numpy(), g_dia.cpu().numpy()
    ref = np.zeros_like(gd)
    for b in range(B):
        for j, O in enumerate(ops):
            for q in range(Q):
                L, R = lam[b, q].astype(complex), rho[b, q].astype(complex)
                ref[b, j, q // nodes] += w[q] * np.real(np.trace(L @ (-1j * (O @ R - R @ O))))
    scale = np.abs(ref).max()
    assert np.abs(ga - ref).max() <= 1e-5 * scale, np.abs(ga - ref).max() / scale
    assert np.abs(gd - ref).max() <= 1e-5 * scale, np.abs(gd - ref).max() / scale


def test_adjoint_plan_reuse_matches_one_shot():
    """A packed AdjointPlan run on two different control batches gives the same loss
    and gradients as one-shot calls (controls are rewritten on the device, the
    recompute starts from the forward's checkpoints copied on the device)."""
    from dynamiqs_amd.adjoint import AdjointPlan, mesolve_pwc_grad

    H0, ops, times, u, Ls, rho0, O = _problem(n=8, nseg=20, T=2.0, B=3, seed=4)
    plan = AdjointPlan(H0, ops, times, Ls, rho0, O, 3, checkpoint_every=5)
    for k in range(2):
        uk = u * (1.0 + 0.5 * k)
        r = plan.run(uk)
        J, g = mesolve_pwc_grad(H0, ops, times, uk, Ls, rho0, O, checkpoint_every=5)
        np.testing.assert_allclose(r.J, J, rtol=0, atol=1e-12)
        np.testing.assert_allclose(r.grad_u, g, rtol=0, atol=1e-12)


def test_readme_gradient_kat_with_parameter_gradient():
    """The README gradient example (README.md:85-121) with dJ/domega as a parameter
    gradient of the constant term omega a^dag a (no control at all): d/domega 0.0000,
    d/dkappa -3.3520, d/dalpha0 1.0670."""
    from dynamiqs_amd.adjoint import mesolve_adjoint
    from oracle import operators as op

    n, omega, kappa, alpha0 = 16, 1.0, 0.1, 1.0
    T = 2 * np.pi
    a = op.destroy(n)
    N = a.conj().T @ a
    times = np.linspace(0.0, T, 41)
    res = mesolve_adjoint(omega * N, [], times, np.zeros((1, 0, 40)), [np.sqrt(kappa) * a], op.coherent_dm(n, alpha0),
                          N, rate_ops=[a], param_ops=[N], checkpoint_every=4)
    h = 1e-5
    drho = (op.coherent_dm(n, alpha0 + h) - op.coherent_dm(n, alpha0 - h)) / (2 * h)
    g_omega = float(res.grad_params[0, 0])
    g_kappa = float(res.grad_rates[0, 0])
    g_alpha = float(np.real(np.trace(res.lambda0[0].astype(complex) @ drho)))
    assert abs(g_omega) <= 1e-4, g_omega
    assert abs(g_kappa - (-3.3520)) <= 1.5e-4 and abs(g_alpha - 1.0670) <= 1.5e-4, (g_kappa, g_alpha)


def test_loss_over_several_times_and_modulated_parameter_vs_fd():
    """J = Tr(N rho(t_a)) + Tr(X rho(T)) with a PWC drive, a modulated term
    theta g(t) X and a constant term phi Y: dJ/du, dJ/dtheta and dJ/dphi against
    central differences (h = 1e-4) of the complex128 oracle at tight tolerances."""
    from dynamiqs_amd.adjoint import mesolve_adjoint
    from oracle import mesolve as om
    from oracle import operators as op

    rng = np.random.default_rng(3)
    n, nseg, T = 8, 10, 2.0
    a = op.destroy(n)
    N, X = a.conj().T @ a, a + a.conj().T
    Y = 1j * (a.conj().T - a)
    times = np.linspace(0.0, T, nseg + 1)
    u = rng.normal(0.0, 0.5, size=(2, 1, nseg))
    g = lambda t: 0.3 * np.cos(2.0 * t)  # noqa: E731
    phi = 0.2
    H0 = N + phi * Y
    Ls = [np.sqrt(0.1) * a]
    rho0 = op.fock_dm(n, 0)
    losses = [(times[4], N), (T, X)]

    def loss(ub, theta=1.0, ph=phi):
        Hop = om.TimeOp(N + ph * Y, [om.PWCTerm(times, ub[0], X), om.ModTerm(lambda t: theta * g(t), X)])
        r = om.mesolve(Hop, Ls, rho0, times, exp_ops=[N, X], rtol=1e-10, atol=1e-12)
        return float(r['expects'][0, 4].real + r['expects'][1, -1].real)

    res = mesolve_adjoint(H0, [X], times, u, Ls, rho0, losses, mod_terms=[(g, X)], param_ops=[(g, X), Y],
                          rtol=1e-8, atol=1e-10, checkpoint_every=5)
    hh = 1e-4
    for b in range(2):
        assert abs(res.J[b] - loss(u[b])) < 1e-5
        fd = np.zeros(nseg)
        for k in range(nseg):
            up, um = u[b].copy(), u[b].copy()
            up[0, k] += hh
            um[0, k] -= hh
            fd[k] = (loss(up) - loss(um)) / (2 * hh)
        err = np.abs(res.grad_u[b, 0] - fd).max()
        assert err <= 1e-4 * np.abs(fd).max() + 1e-6, (b, err)
        fd_theta = (loss(u[b], theta=1 + hh) - loss(u[b], theta=1 - hh)) / (2 * hh)
        fd_phi = (loss(u[b], ph=phi + hh) - loss(u[b], ph=phi - hh)) / (2 * hh)
        assert abs(res.grad_params[b, 0] - fd_theta) <= 1e-4 * max(1.0, abs(fd_theta)), (res.grad_params[b], fd_theta)
        assert abs(res.grad_params[b, 1] - fd_phi) <= 1e-4 * max(1.0, abs(fd_phi)), (res.grad_params[b], fd_phi)
