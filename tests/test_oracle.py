"""Pins of the CPU oracle (``oracle/``) against the reference's KATs, closed forms,
library routines and order conditions. CPU only."""

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from oracle import dia, lindblad, operators as op, rouchon, tsit5
from oracle import mesolve as om

# ----------------------------------------------------------------- operators


def test_operator_docstring_kats(kats):
    k = kats['coherent_4_0.5']
    np.testing.assert_allclose(op.coherent(4, 0.5)[:, 0].real, k['ket_real'], atol=k['atol'])
    k = kats['coherent_dm_4_0.5']
    np.testing.assert_allclose(op.coherent_dm(4, 0.5).real, k['real'], atol=k['atol'])
    k = kats['displace_4_0.5']
    np.testing.assert_allclose(op.displace(4, 0.5).real, k['real'], atol=k['atol'])
    k = kats['position_3']
    np.testing.assert_allclose(op.position(3).real, k['real'], atol=k['atol'])
    k = kats['momentum_3']
    np.testing.assert_allclose(op.momentum(3).imag, k['imag'], atol=k['atol'])


def test_coherent_poisson_statistics():
    # |<m|alpha>|^2 -> Poisson(|alpha|^2) as the truncation grows (closed form)
    from math import factorial

    alpha = 1.3 + 0.4j
    psi = op.coherent(60, alpha)[:, 0]
    m = np.arange(12)
    lam = abs(alpha) ** 2
    poisson = np.exp(-lam) * lam**m / np.array([factorial(int(i)) for i in m])
    np.testing.assert_allclose(np.abs(psi[:12]) ** 2, poisson, atol=1e-12)


def test_multimode_ordering():
    a, b = op.destroy(3, 4)
    np.testing.assert_allclose(a, np.kron(op.destroy(3), np.eye(4)))
    np.testing.assert_allclose(b, np.kron(np.eye(3), op.destroy(4)))
    # DIA offsets of the two-mode config: a -> +n_b, b -> +1, a^dag b -> -(n_b - 1)
    offs, _ = dia.dia_from_dense(a)
    assert offs == (4,)
    offs, _ = dia.dia_from_dense(b)
    assert offs == (1,)
    offs, _ = dia.dia_from_dense(a.conj().T @ b)
    assert offs == (-3,)


# ----------------------------------------------------------------------- DIA


def _rand_dia(rng, offsets, n, batch=()):
    d = rng.normal(size=(*batch, len(offsets), n)) + 1j * rng.normal(size=(*batch, len(offsets), n))
    # zero the out-of-bounds entries as the reference fixture does
    for i, o in enumerate(offsets):
        if o > 0:
            d[..., i, :o] = 0
        elif o < 0:
            d[..., i, n + o :] = 0
    return tuple(offsets), d


@pytest.mark.parametrize('offsets', [(-2, 0, 2), (-2, -1, 1, 3)])
def test_dia_matches_scipy_dia_matrix(kats, offsets):
    n = kats['sparse_dia_fixture_offsets']['N']
    rng = np.random.default_rng(0)
    offs, d = _rand_dia(rng, offsets, n)
    ref = sp.dia_matrix((d, np.array(offs)), shape=(n, n)).toarray()
    np.testing.assert_array_equal(dia.dia_to_dense(offs, d), ref)
    offs2, d2 = dia.dia_from_dense(ref, offs)
    np.testing.assert_array_equal(d2, d)


@pytest.mark.parametrize('lo,ro', [((-2, 0, 2), (-2, -1, 1, 3)), ((1,), (-3, 0)), ((0,), (0,))])
def test_dia_products_vs_dense(lo, ro):
    n = 10
    rng = np.random.default_rng(1)
    lo, ld = _rand_dia(rng, lo, n)
    ro, rd = _rand_dia(rng, ro, n)
    L = dia.dia_to_dense(lo, ld)
    Rm = dia.dia_to_dense(ro, rd)
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    np.testing.assert_allclose(dia.dia_matmul_dense(lo, ld, x), L @ x, atol=1e-12)
    np.testing.assert_allclose(dia.dense_matmul_dia(x, lo, ld), x @ L, atol=1e-12)
    o, d = dia.dia_matmul_dia(lo, ld, ro, rd)
    np.testing.assert_allclose(dia.dia_to_dense(o, d), L @ Rm, atol=1e-12)
    o, d = dia.dia_dag(lo, ld)
    np.testing.assert_allclose(dia.dia_to_dense(o, d), L.conj().T, atol=0)


# ------------------------------------------------------------------ Lindblad


def test_dissipator_lindbladian_kats(kats):
    a = op.destroy(4)
    out = lindblad.dissipator(a, op.fock_dm(4, 2))
    np.testing.assert_allclose(out, np.diag(kats['dissipator_destroy4_fockdm4_2']['diag_real']), atol=1e-12)
    N = a.conj().T @ a
    out = lindblad.lindbladian(N, [a, N], op.fock_dm(4, 1))
    np.testing.assert_allclose(out, np.diag(kats['lindbladian_num4_[a,num]_fockdm4_1']['diag_real']), atol=1e-12)


def _rand_problem(rng, n, K):
    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = H + H.conj().T
    Ls = [rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) for _ in range(K)]
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    rho = psi @ psi.conj().T
    return H, Ls, rho / np.trace(rho)


def test_rhs_forms_agree_and_invariants():
    rng = np.random.default_rng(2)
    H, Ls, rho = _rand_problem(rng, 5, 2)
    a = lindblad.rhs_standard(H, Ls, rho)
    b = lindblad.rhs_hermitian(H, Ls, rho)
    c = lindblad.unvectorize(lindblad.slindbladian(H, Ls) @ lindblad.vectorize(rho))
    d = lindblad.lindbladian(H, Ls, rho)
    for x in (b, c, d):
        np.testing.assert_allclose(x, a, atol=1e-10)
    assert abs(np.trace(a)) < 1e-10
    np.testing.assert_allclose(a, a.conj().T, atol=1e-10)


def test_rhs_brute_force_index_sum():
    rng = np.random.default_rng(3)
    n = 3
    H, Ls, rho = _rand_problem(rng, n, 1)
    L = Ls[0]
    out = np.zeros((n, n), dtype=complex)
    for i in range(n):
        for j in range(n):
            s = 0
            for k in range(n):
                s += -1j * (H[i, k] * rho[k, j] - rho[i, k] * H[k, j])
                for m in range(n):
                    s += L[i, k] * rho[k, m] * np.conj(L[j, m])
                    s -= 0.5 * np.conj(L[k, i]) * L[k, m] * rho[m, j]
                    s -= 0.5 * rho[i, k] * np.conj(L[m, k]) * L[m, j]
            out[i, j] = s
    np.testing.assert_allclose(lindblad.rhs_hermitian(H, Ls, rho), out, atol=1e-12)


def test_adjoint_lindbladian_is_adjoint():
    rng = np.random.default_rng(4)
    H, Ls, rho = _rand_problem(rng, 4, 2)
    O = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    lhs = np.trace(O.conj().T @ lindblad.lindbladian(H, Ls, rho))
    rhs = np.trace(lindblad.adjoint_lindbladian(H, Ls, O).conj().T @ rho)
    assert abs(lhs - rhs) < 1e-10


# --------------------------------------------------------------------- Tsit5


def _order_conditions(b, A, c):
    Ac, Ac2, Ac3 = A @ c, A @ c**2, A @ c**3
    AAc = A @ Ac
    return {
        (1, 'b'): (b.sum(), 1),
        (2, 'bc'): (b @ c, 1 / 2),
        (3, 'bc2'): (b @ c**2, 1 / 3),
        (3, 'bAc'): (b @ Ac, 1 / 6),
        (4, 'bc3'): (b @ c**3, 1 / 4),
        (4, 'bcAc'): (b @ (c * Ac), 1 / 8),
        (4, 'bAc2'): (b @ Ac2, 1 / 12),
        (4, 'bAAc'): (b @ AAc, 1 / 24),
        (5, 'bc4'): (b @ c**4, 1 / 5),
        (5, 'bc2Ac'): (b @ (c**2 * Ac), 1 / 10),
        (5, 'b(Ac)^2'): (b @ Ac**2, 1 / 20),
        (5, 'bcAc2'): (b @ (c * Ac2), 1 / 15),
        (5, 'bcAAc'): (b @ (c * AAc), 1 / 30),
        (5, 'bAc3'): (b @ Ac3, 1 / 20),
        (5, 'bA(cAc)'): (b @ (A @ (c * Ac)), 1 / 40),
        (5, 'bAAc2'): (b @ (A @ Ac2), 1 / 60),
        (5, 'bAAAc'): (b @ (A @ AAc), 1 / 120),
    }


def test_tsit5_tableau_order_conditions():
    A, b, c, e = tsit5.A, tsit5.B, tsit5.C, tsit5.E
    np.testing.assert_allclose(A.sum(axis=1), c, atol=1e-14)
    conds = _order_conditions(b, A, c)
    assert len(conds) == 17
    for key, (val, want) in conds.items():
        assert abs(val - want) < 1e-12, key
    bhat = b - e
    for key, (val, want) in _order_conditions(bhat, A, c).items():
        if key[0] <= 4:
            assert abs(val - want) < 1e-12, key
    # the embedded pair is genuinely of order 4 (not 5)
    assert any(abs(v - w) > 1e-6 for k, (v, w) in _order_conditions(bhat, A, c).items() if k[0] == 5)


@pytest.mark.parametrize('theta', [0.1, 0.37, 0.5, 0.8, 1.0])
def test_tsit5_interpolant_conditions(theta):
    A, c = tsit5.A, tsit5.C
    w = tsit5.interp_weights(theta)
    Ac = A @ c
    checks = [
        (w.sum(), theta),
        (w @ c, theta**2 / 2),
        (w @ c**2, theta**3 / 3),
        (w @ Ac, theta**3 / 6),
        (w @ c**3, theta**4 / 4),
        (w @ (c * Ac), theta**4 / 8),
        (w @ (A @ c**2), theta**4 / 12),
        (w @ (A @ Ac), theta**4 / 24),
    ]
    for val, want in checks:
        assert abs(val - want) < 1e-12
    if theta == 1.0:
        np.testing.assert_allclose(w, tsit5.B, atol=1e-13)


def test_tsit5_fifth_order_convergence():
    # fixed steps on a linear test problem: error ratio ~ 2^5 when halving dt
    lam = -1.0 + 3.0j

    def run(h):
        y = 1.0 + 0j
        for _ in range(int(round(1.0 / h))):
            ks = [h * lam * y]
            for s in range(1, 7):
                ks.append(h * lam * (y + sum(tsit5.A[s, j] * ks[j] for j in range(s))))
            y = y + sum(tsit5.B[j] * ks[j] for j in range(6))
        return abs(y - np.exp(lam))

    r = run(0.05) / run(0.025)
    assert 24 < r < 40


# --------------------------------------------------------------------- Dopri5


def test_dopri5_tableau_order_conditions():
    from oracle import dopri5

    A, b, c, bhat = dopri5.A, dopri5.B, dopri5.C, dopri5.BHAT
    np.testing.assert_allclose(A.sum(axis=1), c, atol=1e-14)
    for key, (val, want) in _order_conditions(b, A, c).items():
        assert abs(val - want) < 1e-12, key
    for key, (val, want) in _order_conditions(bhat, A, c).items():
        if key[0] <= 4:
            assert abs(val - want) < 1e-12, key
    assert any(abs(v - w) > 1e-6 for k, (v, w) in _order_conditions(bhat, A, c).items() if k[0] == 5)
    np.testing.assert_allclose(dopri5.E, b - bhat, atol=1e-16)
    assert b[6] == 0.0  # FSAL: the last stage is f(y1)


def test_dopri5_midpoint_and_interpolant_conditions():
    """Shampine's midpoint weights are a 4th-order value at theta = 1/2; the quartic
    dense output is of order 4 at every theta, passes through y1 at theta = 1, has
    the end derivatives f_1 and f_7, and reproduces the midpoint value."""
    from oracle import dopri5

    A, c = dopri5.A, dopri5.C
    Ac = A @ c

    def conds(w, theta):
        return [(w.sum(), theta), (w @ c, theta**2 / 2), (w @ c**2, theta**3 / 3), (w @ Ac, theta**3 / 6),
                (w @ c**3, theta**4 / 4), (w @ (c * Ac), theta**4 / 8), (w @ (A @ c**2), theta**4 / 12),
                (w @ (A @ Ac), theta**4 / 24)]

    for val, want in conds(dopri5.CMID, 0.5):
        assert abs(val - want) < 1e-12
    tab = tsit5.tableau('dopri5')
    for theta in (0.1, 0.37, 0.5, 0.8, 1.0):
        w = tsit5.interp_weights(theta, tab)
        for val, want in conds(w, theta):
            assert abs(val - want) < 1e-12
    np.testing.assert_allclose(tsit5.interp_weights(1.0, tab), dopri5.B, atol=1e-13)
    np.testing.assert_allclose(tsit5.interp_weights(0.5, tab), dopri5.CMID, atol=1e-13)
    # derivative of b_i(theta) at 0 and 1: e_1 and e_7
    dR = dopri5.R @ np.array([1.0, 2.0, 3.0, 4.0])
    np.testing.assert_allclose(dopri5.R[:, 0], np.eye(7)[0], atol=1e-15)
    np.testing.assert_allclose(dR, np.eye(7)[6], atol=1e-12)
    # the theta^4 column is the dense-output constants of DOPRI5 (Hairer's code)
    d = [-12715105075 / 11282082432, 0.0, 87487479700 / 32700410799, -10690763975 / 1880347072,
         701980252875 / 199316789632, -1453857185 / 822651844, 69997945 / 29380423]
    np.testing.assert_allclose(dopri5.R[:, 3], d, rtol=1e-12)


def test_dopri5_fifth_order_convergence_and_closed_form():
    from oracle import dopri5

    lam = -1.0 + 3.0j

    def run(h):
        y = 1.0 + 0j
        for _ in range(int(round(1.0 / h))):
            ks = [h * lam * y]
            for s in range(1, 7):
                ks.append(h * lam * (y + sum(dopri5.A[s, j] * ks[j] for j in range(s))))
            y = y + sum(dopri5.B[j] * ks[j] for j in range(6))
        return abs(y - np.exp(lam))

    r = run(0.05) / run(0.025)
    assert 24 < r < 40
    # adaptive Dopri5 against the exact superoperator solution
    rng = np.random.default_rng(5)
    H, Ls, rho0 = _rand_problem(rng, 6, 2)
    H /= np.linalg.norm(H, 2)
    Ls = [0.3 * L / np.linalg.norm(L, 2) for L in Ls]
    tsave = np.linspace(0, 3, 7)
    res = om.mesolve(H, Ls, rho0, tsave, rtol=1e-10, atol=1e-12, method='dopri5')
    Lsup = lindblad.slindbladian(H, Ls)
    for i, t in enumerate(tsave):
        ex = lindblad.unvectorize(sla.expm(t * Lsup) @ lindblad.vectorize(rho0))
        assert np.linalg.norm(res['states'][i] - ex) < 1e-8
    # it is a different method: the step sequences differ from Tsit5's
    r5 = om.mesolve(H, Ls, rho0, tsave, rtol=1e-6, atol=1e-6, method='dopri5')
    rt = om.mesolve(H, Ls, rho0, tsave, rtol=1e-6, atol=1e-6)
    assert r5['infos']['nsteps'] != rt['infos']['nsteps'] or not np.allclose(r5['states'], rt['states'], atol=1e-12)


def _ocavity(n, delta, kappa, alpha0):
    a = op.destroy(n)
    return delta * a.conj().T @ a, [np.sqrt(kappa) * a], op.coherent(n, alpha0)


def test_tsit5_closed_form_ocavity(kats):
    k = kats['ocavity']
    H, Ls, psi0 = _ocavity(k['n'], k['delta'], k['kappa'], k['alpha0'])
    tsave = np.linspace(0, k['t_end'], k['ntsave'])
    Es = [op.position(k['n']), op.momentum(k['n'])]
    r = om.mesolve(H, Ls, psi0, tsave, exp_ops=Es)
    for i, t in enumerate(tsave):
        al = k['alpha0'] * np.exp(-1j * k['delta'] * t - 0.5 * k['kappa'] * t)
        err = np.linalg.norm(r['states'][i] - op.coherent_dm(k['n'], al))
        assert err <= k['ysave_atol']
        np.testing.assert_allclose(r['expects'][:, i], [al.real, al.imag], rtol=k['esave_rtol'], atol=k['esave_atol'])


def test_tsit5_otdqubit_closed_form(kats):
    k = kats['otdqubit']
    eps, w, g = k['eps'], k['omega'], k['gamma']
    sx = op.sigmax()
    # H(t) = eps cos(w t) sigma_x as a modulated term (evaluated at every stage time)
    Hop = om.TimeOp(np.zeros((2, 2), dtype=complex), [om.ModTerm(lambda t: eps * np.cos(w * t), sx)])
    np.testing.assert_allclose(Hop(0.3), eps * np.cos(w * 0.3) * sx)
    assert Hop.jump_ts.size == 0
    tsave = np.linspace(0, k['t_end'], k['ntsave'])
    r = om.mesolve(Hop, [np.sqrt(g) * sx], op.fock(2, 0), tsave)
    for i, t in enumerate(tsave):
        th = 2 * eps / w * np.sin(w * t)
        eta = np.exp(-2 * g * t)
        want = np.array(
            [[0.5 * (1 + eta * np.cos(th)), 0.5j * eta * np.sin(th)], [-0.5j * eta * np.sin(th), 0.5 * (1 - eta * np.cos(th))]]
        )
        assert np.linalg.norm(r['states'][i] - want) < 1e-4


def test_tsit5_vs_superoperator_expm():
    # constant generator: exact solution expm(t L) vec(rho0) (the reference's Expm method)
    rng = np.random.default_rng(5)
    H, Ls, rho0 = _rand_problem(rng, 6, 2)
    H /= np.linalg.norm(H, 2)
    Ls = [0.3 * L / np.linalg.norm(L, 2) for L in Ls]
    tsave = np.linspace(0, 3, 7)
    r = om.mesolve(H, Ls, rho0, tsave, rtol=1e-10, atol=1e-12)
    Lsup = lindblad.slindbladian(H, Ls)
    for i, t in enumerate(tsave):
        ex = lindblad.unvectorize(sla.expm(t * Lsup) @ lindblad.vectorize(rho0))
        assert np.linalg.norm(r['states'][i] - ex) < 1e-8


def test_tsit5_readme_kat(kats):
    k = kats['readme_lossy_oscillator']
    n = k['n']
    a = op.destroy(n)
    tsave = np.linspace(0, k['T'], k['ntsave'])
    r = om.mesolve(k['omega'] * a.conj().T @ a, [np.sqrt(k['kappa']) * a], op.coherent(n, k['alpha0']), tsave)
    # the state stays coherent: alpha(T) = alpha0 exp(-kappa T / 2) exp(-i omega T)
    al = k['alpha0'] * np.exp(-k['kappa'] * k['T'] / 2 - 1j * k['omega'] * k['T'])
    # Tsit5 global error at rtol = atol = 1e-6 over 40 steps is ~1e-4 in Frobenius norm
    assert np.linalg.norm(r['final_state'] - op.coherent_dm(n, al)) < 2e-4
    # the printed step counts, exactly
    assert (r['infos']['nsteps'], r['infos']['naccepted'], r['infos']['nrejected']) == (
        k['nsteps'], k['naccepted'], k['nrejected'])
    d = kats['mesolve_docstring']
    r = om.mesolve(a.conj().T @ a, [a], op.coherent(n, d['alpha0']), np.linspace(0, d['T'], d['ntsave']))
    assert (r['infos']['nsteps'], r['infos']['naccepted'], r['infos']['nrejected']) == (
        d['nsteps'], d['naccepted'], d['nrejected'])


def _counts(st):
    return st.nsteps, st.naccepted, st.nrejected


def test_printed_step_counts_of_sesolve_floquet_and_dopri5(kats):
    """The step counts printed by the reference for its sesolve and floquet docstring
    examples (Tsit5) and its Rabi workflow (Dopri5 at atol = 1e-8, with rejected
    steps). Together with the two mesolve examples they fix the initial-step
    exponent (1/4) and Dopri5's embedded weights (Shampine's). With 1/5, or with the
    1980 embedded weights, these counts are missed: checked below by planting each
    change in a copy of the oracle's tableau and constants."""
    from oracle import dopri5
    from oracle import sesolve as ose

    d = kats['sesolve_docstring']
    a = op.destroy(d['n'])
    r = ose.sesolve(a.conj().T @ a, op.coherent(d['n'], d['alpha0']), np.linspace(0, d['T'], d['ntsave']))
    assert (r['infos']['nsteps'], r['infos']['naccepted'], r['infos']['nrejected']) == (
        d['nsteps'], d['naccepted'], d['nrejected'])

    # floquet: the propagator from 0 to T, saved at tsave and at t0 + T
    q = kats['floquet_docstring']
    sx = op.sigmax()
    ts = np.append(np.linspace(0, q['T'], q['ntsave']), q['T'])
    fq = lambda t, y, tseg: -1j * np.cos(2 * np.pi * t) * (sx @ y)  # noqa: E731
    _, _, st = tsit5.integrate(fq, np.eye(2, dtype=complex), ts)
    assert _counts(st) == (q['nsteps'], q['naccepted'], q['nrejected'])

    w = kats['workflow_rabi_dopri5']
    H = w['delta'] * op.sigmaz() + w['Omega'] * sx
    tsw = np.linspace(0.0, w['sim_time'], w['num_save'])
    r = ose.sesolve(H, op.fock(2, 0), tsw, method='dopri5', rtol=w['rtol'], atol=w['atol'])
    want_w = (w['nsteps'], w['naccepted'], w['nrejected'])
    assert (r['infos']['nsteps'], r['infos']['naccepted'], r['infos']['nrejected']) == want_w

    # planted: HINIT's own exponent 1/5 drops one step of the sesolve example, and the
    # 1980 embedded weights take more accepted Dopri5 steps
    import unittest.mock as um

    with um.patch.object(tsit5, 'INIT_ORDER', 5):
        r = ose.sesolve(a.conj().T @ a, op.coherent(d['n'], d['alpha0']), np.linspace(0, d['T'], d['ntsave']))
        assert r['infos']['nsteps'] == d['nsteps'] - 1
    bhat80 = np.array([5179 / 57600, 0.0, 7571 / 16695, 393 / 640, -92097 / 339200, 187 / 2100, 1 / 40])
    with um.patch.object(dopri5, 'E', dopri5.B - bhat80):
        r = ose.sesolve(H, op.fock(2, 0), tsw, method='dopri5', rtol=w['rtol'], atol=w['atol'])
        assert r['infos']['naccepted'] > w['naccepted'] + 1


def test_qubit_decay_doc_kat(kats):
    r = om.mesolve(op.sigmaz(), [op.sigmam()], op.fock(2, 0), np.linspace(0, 1, 11))
    np.testing.assert_allclose(r['final_state'].real, kats['qubit_decay']['final_real'], atol=kats['qubit_decay']['atol'])


def test_max_steps_error():
    a = op.destroy(4)
    with pytest.raises(tsit5.MaxStepsError):
        om.mesolve(a.conj().T @ a, [a], op.fock(4, 2), np.linspace(0, 10, 3), max_steps=3)


def test_pwc_semantics_and_jumps():
    # left-closed intervals, zero outside [times[0], times[-1]) (tests/core/test_time_qarray.py:285-292)
    term = om.PWCTerm(np.array([0.0, 1.0, 2.0]), np.array([3.0, 5.0]), np.eye(2))
    assert term.coeff(-0.1) == 0 and term.coeff(0.0) == 3 and term.coeff(0.999) == 3
    assert term.coeff(1.0) == 5 and term.coeff(1.999) == 5 and term.coeff(2.0) == 0
    # piecewise-constant drive: exact solution by products of segment propagators
    n = 5
    a = op.destroy(n)
    times = np.array([0.0, 0.4, 0.7, 1.5])
    vals = np.array([0.3, -0.8, 0.5])
    H = om.TimeOp(a.conj().T @ a, [om.PWCTerm(times, vals, a + a.conj().T)])
    Ls = [np.sqrt(0.2) * a]
    rho0 = op.fock_dm(n, 1)
    tsave = np.array([0.0, 0.5, 1.5])
    r = om.mesolve(H, Ls, rho0, tsave, rtol=1e-10, atol=1e-12)
    v = lindblad.vectorize(rho0)
    out = [v]
    bounds = [0.0, 0.4, 0.5, 0.7, 1.5]
    for t0_, t1_ in zip(bounds[:-1], bounds[1:]):
        Hs = H(t0_)
        v = sla.expm((t1_ - t0_) * lindblad.slindbladian(Hs, Ls)) @ v
        out.append(v)
    np.testing.assert_allclose(r['states'][1], lindblad.unvectorize(out[2]), atol=1e-8)
    np.testing.assert_allclose(r['states'][2], lindblad.unvectorize(out[4]), atol=1e-8)


# ------------------------------------------------------------------ Rouchon1


def test_rouchon_banded_matches_dense():
    rng = np.random.default_rng(6)
    n = 24
    a, b = op.destroy(4, 6)
    H = 1.3 * a.conj().T @ a + 0.7 * b.conj().T @ b + 0.2 * (a.conj().T @ b + a @ b.conj().T)
    Ls = [np.sqrt(0.1) * a, np.sqrt(0.3) * b]
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    rho = psi @ psi.conj().T / np.vdot(psi, psi)
    dt = 0.05
    M0, Ms, S = rouchon.kraus_ops(H, Ls, dt)
    offs, _ = dia.dia_from_dense(S)
    w = max(abs(o) for o in offs)
    x = rouchon.kraus_step_dense(H, Ls, rho, dt)
    y = rouchon.kraus_step_banded(H, Ls, rho, dt, w)
    np.testing.assert_allclose(y, x, atol=1e-13)
    # folded form sum A rho A^dag, and completeness sum A^dag A = I
    As = rouchon.normalized_kraus(H, Ls, dt)
    z = sum(A @ rho @ A.conj().T for A in As)
    np.testing.assert_allclose(z, x, atol=1e-13)
    np.testing.assert_allclose(sum(A.conj().T @ A for A in As), np.eye(n), atol=1e-13)
    assert abs(np.trace(x) - 1) < 1e-13


def test_rouchon_block_normalisation_equals_dense_step():
    """kraus_step_dia_blocks (Cholesky of S block by block on the index sets S never
    couples, DIA products) equals the dense LAPACK step on non-Hermitian inputs; the
    blocks of the two-mode problem are the sets of fixed excitation number; and a
    partition that splits a coupled set (a plausible mistake) is caught."""
    rng = np.random.default_rng(8)
    a, b = op.destroy(6, 5)
    ad, bd = a.conj().T, b.conj().T
    H = 5 * ad @ a + 6 * bd @ b + 0.1 * (ad @ b + a @ bd) - 0.01 * ad @ a @ bd @ b
    Ls = [np.sqrt(0.02) * a, np.sqrt(0.05) * b]
    n = H.shape[0]
    dt = 0.02
    M0, Ms, S = rouchon.kraus_ops(H, Ls, dt)
    blocks = rouchon.coupled_blocks(S)
    Ntot = np.real(np.diag(ad @ a + bd @ b)).round().astype(int)
    assert len(blocks) == len(np.unique(Ntot))
    for idx in blocks:
        assert len(np.unique(Ntot[idx])) == 1 and np.all(np.diff(idx) > 0)
    fac = rouchon.block_factor(S, blocks)
    dM = dia.dia_from_dense(M0 - np.eye(n))
    Ld = [dia.dia_from_dense(M) for M in Ms]
    rho = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    x, y = rho.copy(), rho.copy()
    for _ in range(5):
        x = rouchon.kraus_step_dense(H, Ls, x, dt)
        y = rouchon.kraus_step_dia_blocks(dM, Ld, y, fac)
    np.testing.assert_allclose(y, x, atol=1e-13 * np.abs(x).max())
    big = max(blocks, key=len)
    split = [blk for blk in blocks if blk is not big] + [big[:1], big[1:]]
    z = rouchon.kraus_step_dia_blocks(dM, Ld, rho, rouchon.block_factor(S, split))
    assert np.abs(z - rouchon.kraus_step_dense(H, Ls, rho, dt)).max() > 1e-6


def test_c3_fixture_is_a_trace_one_state_consistent_with_its_expectations():
    """The committed full-length DIA Rouchon1 fixture (alpha = 2, 5000 steps): the
    normalised scheme keeps Tr(rho) = 1 at every save to rounding, <a^dag a>(0) =
    |alpha|^2, purity in (0, 1], and the stored final block is Hermitian with
    <a^dag a> equal to the last saved expectation value."""
    import os

    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'c3_dia_rouchon1_alpha2.npz'))
    assert int(fx['nsteps']) == 5000 and len(fx['tsave']) == 51
    assert np.abs(fx['trace'] - 1).max() < 1e-10
    assert abs(fx['expect_a'][0] - 4.0) < 1e-10 and abs(fx['expect_b'][0]) < 1e-12
    assert np.all(fx['purity'] <= 1 + 1e-12) and np.all(fx['purity'] > 0.9)
    blk = fx['final_block'].astype(complex)
    idx = fx['final_index']
    assert np.abs(blk - blk.conj().T).max() < 1e-7
    na = idx // 32  # a^dag a of the basis state (i_a, i_b) with index 32 i_a + i_b
    assert abs(np.real(np.diag(blk)) @ na - fx['expect_a'][-1]) < 1e-6
    assert abs(np.real(np.trace(blk)) - 1) < 1e-6 and float(fx['final_tail_frobenius']) < 1e-6


def test_c5_fixture_fd_matches_exact_derivative():
    """The committed full-size adjoint fixture: its central differences agree with
    the exact (Frechet) derivative to O(h^2), and the loss of the first control
    vector equals the oracle's tight-tolerance Tsit5 solve."""
    import os

    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'c5_adjoint_n32_fd.npz'))
    fd, ex = fx['grad_fd'], fx['grad_exact']
    assert fd.shape == (2, 1, 100)
    assert np.abs(fd - ex).max() <= 1e-8 * np.abs(ex).max()
    a = op.destroy(32)
    X = a + a.conj().T
    Hop = om.TimeOp(a.conj().T @ a, [om.PWCTerm(fx['times'], fx['u'][0, 0], X)])
    r = om.mesolve(Hop, [np.sqrt(0.05) * a], op.fock_dm(32, 0), np.array([0.0, 10.0]), exp_ops=[a.conj().T @ a],
                   rtol=1e-10, atol=1e-12)
    assert abs(r['expects'][0, -1].real - fx['J'][0]) < 1e-8


def test_rouchon_cholesky_matches_reference_formula():
    # cholesky_normalize(S, rho) == T^{-dag} rho T^{-1} with S = T T^dag
    rng = np.random.default_rng(7)
    n = 6
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    S = np.eye(n) + 0.1 * X.conj().T @ X
    rho = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    T = np.linalg.cholesky(S)
    Ti = np.linalg.inv(T)
    np.testing.assert_allclose(rouchon.cholesky_normalize(S, rho), Ti.conj().T @ rho @ Ti, atol=1e-12)


def test_rouchon1_first_order_and_reference_tolerance(kats):
    k = kats['ocavity']
    n = k['n']
    H, Ls, psi0 = _ocavity(n, k['delta'], k['kappa'], k['alpha0'])
    tsave = np.linspace(0, k['t_end'], k['ntsave'])
    errs = []
    for dt in (4e-3, 2e-3, 1e-3):
        r = om.mesolve(H, Ls, psi0, tsave, method='rouchon1', dt=dt)
        Lsup = lindblad.slindbladian(H, Ls)
        ex = lindblad.unvectorize(sla.expm(tsave[-1] * Lsup) @ lindblad.vectorize(op.todm(psi0)))
        errs.append(np.linalg.norm(r['final_state'] - ex))
        assert abs(np.trace(r['final_state']) - 1) < 1e-12
        assert r['infos']['nsteps'] == int(round(tsave[-1] / dt))
    assert 1.7 < errs[0] / errs[1] < 2.3 and 1.7 < errs[1] / errs[2] < 2.3
    # the reference's own test step dt = 1e-4 meets its ysave_atol = 1e-3
    dt = kats['rouchon1_test_dt']['dt']
    r = om.mesolve(H, Ls, psi0, tsave, method='rouchon1', dt=dt)
    for i, t in enumerate(tsave):
        al = k['alpha0'] * np.exp(-1j * k['delta'] * t - 0.5 * k['kappa'] * t)
        assert np.linalg.norm(r['states'][i] - op.coherent_dm(n, al)) < kats['rouchon1_test_dt']['ysave_atol']


def test_tsit5_no_sliver_steps_at_breakpoints():
    """A step proposed equal to the previous segment length can stop an ulp short of
    the next linspace breakpoint; the controller then ends it on the breakpoint
    instead of taking a ~1e-15 step. Seeded driven oscillator (n = 32, 100 control
    segments on [0, 10]) that produced such a sliver at t = 7.8."""
    from oracle import lindblad, tsit5

    n, nseg = 32, 100
    a = np.diag(np.sqrt(np.arange(1, n, dtype=np.float64)), k=1).astype(np.complex128)
    H0, X = a.conj().T @ a, a + a.conj().T
    times = np.linspace(0.0, 10.0, nseg + 1)
    u = np.random.default_rng(4).normal(0.0, 0.5, size=(64, 1, nseg))[34, 0]
    Ls = [np.sqrt(0.05) * a]
    rho0 = np.zeros((n, n), dtype=np.complex128)
    rho0[0, 0] = 1

    def f(t, y, tseg):
        k = np.searchsorted(times, tseg, side='right') - 1
        return lindblad.rhs_hermitian(H0 + (u[k] if 0 <= k < nseg else 0.0) * X, Ls, y)

    tsave = np.array([0.0, 8.0])
    _, _, st = tsit5.integrate(f, rho0, tsave, rtol=1e-6, atol=1e-6, jump_ts=times)
    assert min(st.dts) > 1e-6, min(st.dts)
    assert st.nrejected == 0 and st.naccepted <= 82


def test_tsit5_replay_of_own_steps_reproduces_adaptive_solve():
    """integrate_steps on the accepted step ends that integrate recorded gives the
    adaptive solution (same stages, weights and interpolant; only the FSAL stage
    is recomputed instead of reused), on a driven lossy cavity with PWC jumps."""
    from oracle import lindblad, operators as op, tsit5

    n = 10
    a = op.destroy(n)
    times = np.linspace(0.0, 2.0, 9)
    u = np.random.default_rng(2).normal(size=8)
    Ls = [np.sqrt(0.2) * a]

    def f(t, y, tseg):
        k = np.searchsorted(times, tseg, side='right') - 1
        c = u[k] if 0 <= k < 8 else 0.0
        return lindblad.rhs_hermitian(a.conj().T @ a + c * (a + a.conj().T), Ls, y)

    tsave = np.linspace(0.0, 2.0, 7)
    rho0 = op.coherent_dm(n, 0.7)
    saved, ylast, st = tsit5.integrate(f, rho0, tsave, rtol=1e-7, atol=1e-9, jump_ts=times)
    replay, yrep = tsit5.integrate_steps(f, rho0, tsave, st.ends)
    assert len(st.ends) == st.naccepted
    np.testing.assert_allclose(np.stack(replay), np.stack(saved), atol=1e-12)
    np.testing.assert_allclose(yrep, ylast, atol=1e-12)


def test_rouchon2_kraus_form_equals_reference_stage_form():
    """The host's Kraus operators of the Rouchon2 step (U, sqrt(dt/2) U L_k,
    sqrt(dt/2) L_k U, dt/sqrt(2) L_k U L_j, normalised by S = sum M^dag M) applied as
    sum A rho A^dag equal the oracle's restatement of the reference's stage form
    (KrausMapRK with the Heun tableau, including its adjoint-pass S), with and
    without normalisation; the normalised operators satisfy sum A^dag A = I."""
    from dynamiqs_amd.integrators.dense import _dag, _kraus_sets

    rng = np.random.default_rng(21)
    H, Ls, rho = _rand_problem(rng, 6, 2)
    for normalize in (True, False):
        D, C = _kraus_sets(H, Ls, 0.05, normalize, order=2)
        A = [np.eye(6) + D[0], *D[1:]]
        got = sum(a @ rho @ _dag(a) for a in A)
        want = rouchon.rouchon2_step(H, Ls, rho, 0.05, normalize)
        np.testing.assert_allclose(got, want, atol=1e-13)
        # the kernel's form: rho + tmp + tmp^dag with tmp = sum_k D_k rho C_k
        tmp = sum(d @ rho @ c for d, c in zip(D, C))
        np.testing.assert_allclose(rho + tmp + _dag(tmp), want, atol=1e-13)
        if normalize:
            np.testing.assert_allclose(sum(_dag(a) @ a for a in A), np.eye(6), atol=1e-13)
    assert len(D) == 1 + 2 * 2 + 2 * 2


def test_rouchon2_second_order_and_trace():
    n = 5
    a = op.destroy(n)
    H = a.conj().T @ a + 0.2 * (a + a.conj().T)
    Ls = [0.5 * a, 0.3 * a.conj().T @ a]
    rho0 = op.coherent_dm(n, 0.7)
    ex = lindblad.unvectorize(sla.expm(lindblad.slindbladian(H, Ls)) @ lindblad.vectorize(rho0))

    def run(dt):
        r = rho0.copy()
        for _ in range(int(round(1.0 / dt))):
            r = rouchon.rouchon2_step(H, Ls, r, dt)
        return r

    e1, e2 = np.linalg.norm(run(0.02) - ex), np.linalg.norm(run(0.01) - ex)
    assert 3.6 < e1 / e2 < 4.4
    assert abs(np.trace(run(0.05)) - 1) < 1e-13


# ----------------------------------------------------------------- sesolve


def test_sesolve_cavity_closed_form():
    """The reference's Cavity system (tests/systems/closed_system.py:46-104,
    n=8, delta=2 pi, alpha0=0.5): H = delta a^dag a is diagonal, so every Fock
    amplitude rotates as exp(-1j delta k t); the truncated <x>, <p> follow
    alpha(t) = alpha0 exp(-1j delta t) within the reference's own tolerances."""
    from oracle import sesolve as ose

    n, delta, a0 = 8, 2 * np.pi, 0.5
    ts = np.linspace(0.0, 0.3, 11)
    psi0 = op.coherent(n, a0)
    out = ose.sesolve(delta * op.number(n), psi0, ts, exp_ops=[op.position(n), op.momentum(n)], rtol=1e-10, atol=1e-10)
    k = np.arange(n)
    exact = np.stack([np.exp(-1j * delta * k * t) * psi0[:, 0] for t in ts])
    np.testing.assert_allclose(out['states'][..., 0], exact, atol=1e-8)
    alpha = a0 * np.exp(-1j * delta * ts)
    np.testing.assert_allclose(out['expects'][0], alpha.real, rtol=1e-3, atol=1e-4)
    np.testing.assert_allclose(out['expects'][1], alpha.imag, rtol=1e-3, atol=1e-4)
    assert out['infos']['naccepted'] > 0


def test_sesolve_tdqubit_closed_form():
    """The reference's TDQubit (closed_system.py:107-176, eps=3, omega=10): H(t) =
    eps cos(omega t) sigma_x from |0> gives cos(theta/2)|0> - 1j sin(theta/2)|1>,
    theta = 2 eps/omega sin(omega t); <sigma_y> = -sin theta, <sigma_z> = cos theta."""
    from oracle import sesolve as ose

    eps, w = 3.0, 10.0
    ts = np.linspace(0.0, 1.0, 11)
    H = om.TimeOp(np.zeros((2, 2), dtype=complex), [om.ModTerm(lambda t: eps * np.cos(w * t), op.sigmax())])
    for method in ('tsit5', 'dopri5'):
        out = ose.sesolve(H, op.fock(2, 0), ts, exp_ops=[op.sigmax(), op.sigmay(), op.sigmaz()], method=method,
                          rtol=1e-10, atol=1e-10)
        th = 2 * eps / w * np.sin(w * ts)
        exact = np.stack([np.cos(th / 2), -1j * np.sin(th / 2)], axis=-1)
        np.testing.assert_allclose(out['states'][..., 0], exact, atol=1e-8)
        np.testing.assert_allclose(out['expects'], np.stack([0 * th, -np.sin(th), np.cos(th)]), atol=1e-8)


def test_sesolve_pwc_matches_product_of_exponentials():
    """H(t) = H0 + v(t) O with v piecewise constant on [0, 0.4), [0.4, 1.1) and zero
    after 1.1 (PWCTimeQArray, time_qarray.py:572-666): psi(T) is the product of
    the segment propagators expm(-1j H_k dt_k) (scipy.linalg.expm); the step
    sequence lands on both breakpoints."""
    from oracle import sesolve as ose

    rng = np.random.default_rng(11)
    n = 5
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H0 = x + x.conj().T
    y = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    O = y + y.conj().T
    v = np.array([0.7, -1.3 + 0.0j])
    H = om.TimeOp(H0, [om.PWCTerm(np.array([0.0, 0.4, 1.1]), v, O)])
    psi0 = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    psi0 /= np.linalg.norm(psi0)
    ts = np.array([0.0, 0.2, 0.4, 0.9, 1.1, 1.5])
    out = ose.sesolve(H, psi0, ts, rtol=1e-11, atol=1e-11)
    expected = [
        psi0,
        sla.expm(-1j * (H0 + v[0] * O) * 0.2) @ psi0,
        sla.expm(-1j * (H0 + v[0] * O) * 0.4) @ psi0,
        sla.expm(-1j * (H0 + v[1] * O) * 0.5) @ sla.expm(-1j * (H0 + v[0] * O) * 0.4) @ psi0,
    ]
    p11 = sla.expm(-1j * (H0 + v[1] * O) * 0.7) @ sla.expm(-1j * (H0 + v[0] * O) * 0.4) @ psi0
    expected += [p11, sla.expm(-1j * H0 * 0.4) @ p11]
    np.testing.assert_allclose(out['states'], np.stack(expected), atol=2e-8)
    ends = np.asarray(out['infos']['ends'])
    assert np.any(ends == 0.4) and np.any(ends == 1.1)


def test_sesolve_matches_mesolve_without_jump_operators():
    """With no jump operators the master equation on |psi><psi| is the
    Schrodinger equation: the two oracle drivers agree on states and <E>."""
    from oracle import sesolve as ose

    rng = np.random.default_rng(3)
    n = 6
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = x + x.conj().T
    E = np.diag(np.arange(n, dtype=complex))
    psi0 = op.coherent(n, 0.8 + 0.3j)
    ts = np.linspace(0.0, 0.7, 8)
    se = ose.sesolve(H, psi0, ts, exp_ops=[E], rtol=1e-11, atol=1e-11)
    me = om.mesolve(H, [], psi0, ts, exp_ops=[E], rtol=1e-11, atol=1e-11)
    rho = np.einsum('tia,tja->tij', se['states'], se['states'].conj())
    np.testing.assert_allclose(rho, me['states'], atol=1e-8)
    np.testing.assert_allclose(se['expects'], me['expects'], atol=1e-8)


# ----------------------------------------------------------------- propagators


def test_propagators_match_matrix_exponentials():
    """For constant generators the oracle propagators (one solve per basis state)
    equal expm(-i H t) and expm(t slindbladian(H, Ls)) in the column-stacking
    convention; a PWC Hamiltonian gives the product of segment exponentials."""
    from oracle import propagators as op_

    rng = np.random.default_rng(4)
    n = 3
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = x + x.conj().T
    L = 0.4 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    ts = np.array([0.0, 0.3, 0.8])
    U = op_.sepropagator(H, ts, rtol=1e-11, atol=1e-11)
    for k, t in enumerate(ts):
        np.testing.assert_allclose(U[k], sla.expm(-1j * H * t), atol=1e-8)
    S = lindblad.slindbladian(H, [L])
    P = op_.mepropagator(H, [L], ts, rtol=1e-11, atol=1e-11)
    for k, t in enumerate(ts):
        np.testing.assert_allclose(P[k], sla.expm(S * t), atol=1e-8)
    # vec(P rho) = P vec(rho) on a random state
    rho = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    np.testing.assert_allclose(lindblad.unvectorize(P[2] @ lindblad.vectorize(rho)),
                               om.mesolve(H, [L], rho, ts, form='standard', rtol=1e-11, atol=1e-11)['states'][2],
                               atol=1e-8)
    O = np.diag([1.0, -1.0, 0.5]).astype(complex)
    Hp = om.TimeOp(H, [om.PWCTerm(np.array([0.0, 0.5, 2.0]), np.array([1.0, -2.0]), O)])
    Up = op_.sepropagator(Hp, ts, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(Up[2], sla.expm(-1j * (H - 2 * O) * 0.3) @ sla.expm(-1j * (H + O) * 0.5), atol=1e-8)


# ----------------------------------------------------------------- Rouchon3


def test_rouchon3_kraus_form_equals_reference_stage_form():
    """The host's 1 + 2K + K^2 + K^3 Kraus operators of the Rouchon3 step, applied
    as sum A rho A^dag (and in the kernel's form rho + tmp + tmp^dag), equal the
    oracle's stage form (KrausMapRK with the Heun3 tableau, Bosh3 no-jump
    propagator and its cubic Hermite dense output, adjoint-pass S), with and without
    normalisation."""
    from dynamiqs_amd.integrators.dense import _dag, _kraus_sets

    rng = np.random.default_rng(23)
    H, Ls, rho = _rand_problem(rng, 5, 2)
    for normalize in (True, False):
        D, C = _kraus_sets(H, Ls, 0.05, normalize, order=3)
        A = [np.eye(5) + D[0], *D[1:]]
        want = rouchon.rouchon3_step(H, Ls, rho, 0.05, normalize)
        np.testing.assert_allclose(sum(a @ rho @ _dag(a) for a in A), want, atol=1e-13)
        tmp = sum(d @ rho @ c for d, c in zip(D, C))
        np.testing.assert_allclose(rho + tmp + _dag(tmp), want, atol=1e-13)
        if normalize:
            np.testing.assert_allclose(sum(_dag(a) @ a for a in A), np.eye(5), atol=1e-13)
    assert len(D) == 1 + 2 * 2 + 2 * 2 + 2**3


def test_rouchon3_third_order_trace_and_propagators():
    """Third-order convergence to expm of the superoperator (error ratio ~8 when dt
    halves, with and without normalisation); trace preserved with normalisation; the
    Bosh3 no-jump propagator matches expm(dt G) to third order and its Hermite
    dense output interpolates U(c) = expm(c dt G) to third order."""
    n = 5
    a = op.destroy(n)
    H = a.conj().T @ a + 0.2 * (a + a.conj().T)
    Ls = [0.5 * a, 0.3 * a.conj().T @ a]
    rho0 = op.coherent_dm(n, 0.7)
    ex = lindblad.unvectorize(sla.expm(lindblad.slindbladian(H, Ls)) @ lindblad.vectorize(rho0))

    def run(dt, normalize=True):
        r = rho0.copy()
        for _ in range(int(round(1.0 / dt))):
            r = rouchon.rouchon3_step(H, Ls, r, dt, normalize)
        return r

    for normalize in (True, False):
        e1, e2 = np.linalg.norm(run(0.04, normalize) - ex), np.linalg.norm(run(0.02, normalize) - ex)
        assert 6.5 < e1 / e2 < 9.5, (normalize, e1, e2)
    assert abs(np.trace(run(0.05)) - 1) < 1e-13
    G = -1j * H - 0.5 * sum(L.conj().T @ L for L in Ls)
    errs = []
    for dt in (0.02, 0.01):
        U = rouchon.rouchon3_propagators(H, Ls, dt)
        errs.append([np.linalg.norm(U[c] - sla.expm(c * dt * G)) for c in (1.0 / 3.0, 2.0 / 3.0, 1.0)])
    ratios = np.array(errs[0]) / np.array(errs[1])
    assert np.all((ratios > 12) & (ratios < 20)), ratios  # local error O(dt^4)


# ----------------------------------------------------------------- Euler


def test_euler_step_identity_order_and_host_sets():
    """Euler (diffrax Euler on the Lindblad equation): rho + h L(rho) equals the
    unnormalised Rouchon1 map minus h^2 G rho G^dag (M0 = I + h G); first-order
    convergence to expm; exact trace preservation; the host's kernel sets (order 0:
    D = [h G, sqrt(h) L_k], C = [I, sqrt(h) L_k^dag / 2]) give the same step."""
    from dynamiqs_amd.integrators.dense import _dag, _kraus_sets

    rng = np.random.default_rng(31)
    H, Ls, rho = _rand_problem(rng, 5, 2)
    h = 0.03
    G = -1j * H - 0.5 * sum(L.conj().T @ L for L in Ls)
    ts = np.array([0.0, h])
    eul = om.mesolve(H, Ls, rho, ts, method='euler', dt=h)['states'][-1]
    r1 = rouchon.kraus_step_dense(H, Ls, rho, h, False)
    np.testing.assert_allclose(eul, r1 - h * h * G @ rho @ G.conj().T, atol=1e-13)
    D, C = _kraus_sets(H, Ls, h, False, order=0)
    tmp = sum(d @ rho @ c for d, c in zip(D, C))
    np.testing.assert_allclose(rho + tmp + _dag(tmp), eul, atol=1e-13)
    assert abs(np.trace(eul) - np.trace(rho)) < 1e-13
    n = 5
    a = op.destroy(n)
    H2 = a.conj().T @ a + 0.2 * (a + a.conj().T)
    L2 = [0.5 * a]
    rho0 = op.coherent_dm(n, 0.7)
    ex = lindblad.unvectorize(sla.expm(lindblad.slindbladian(H2, L2)) @ lindblad.vectorize(rho0))
    e = [np.linalg.norm(om.mesolve(H2, L2, rho0, np.array([0.0, 1.0]), method='euler', dt=dt)['states'][-1] - ex)
         for dt in (0.004, 0.002)]
    assert 1.8 < e[0] / e[1] < 2.2


def test_sesolve_euler_first_order():
    """Euler for sesolve converges at first order to exp(-i H t) psi0 and saves by
    linear interpolation between steps (a save time inside a step)."""
    from oracle import sesolve as ose

    rng = np.random.default_rng(12)
    n = 4
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = (x + x.conj().T) / 4
    psi0 = np.eye(n)[:, :1].astype(complex)
    ex = sla.expm(-1j * H) @ psi0
    e = [np.linalg.norm(ose.sesolve(H, psi0, np.array([0.0, 1.0]), method='euler', dt=dt)['states'][-1] - ex)
         for dt in (0.002, 0.001)]
    assert 1.8 < e[0] / e[1] < 2.2
    out = ose.sesolve(H, psi0, np.array([0.0, 0.25, 0.3]), method='euler', dt=0.1)
    y = [psi0]
    for _ in range(3):
        y.append(y[-1] - 0.1j * H @ y[-1])
    np.testing.assert_allclose(out['states'][1], y[2] + 0.5 * (y[3] - y[2]), atol=1e-14)
    assert out['infos']['nsteps'] == 3


def _rooted_trees(n):
    """Rooted trees with n nodes as sorted tuples of subtrees."""
    import itertools

    if n == 1:
        return [()]
    out = set()

    def parts(m, cap):
        if m == 0:
            yield []
            return
        for k in range(min(m, cap), 0, -1):
            for rest in parts(m - k, k):
                yield [k] + rest

    for p in parts(n - 1, n - 1):
        for combo in itertools.product(*[_rooted_trees(k) for k in p]):
            out.add(tuple(sorted(combo)))
    return sorted(out)


def _tree_order(t):
    return 1 + sum(_tree_order(c) for c in t)


def _tree_gamma(t):
    g = _tree_order(t)
    for c in t:
        g *= _tree_gamma(c)
    return g


def test_dopri8_tableau_order_conditions():
    """RK8(7)13M in exact rationals: every elementary-weight condition
    b . Phi(t) = 1 / gamma(t) for the 200 rooted trees up to order 8 and, for the
    embedded weights, the 85 up to order 7 (and not all of order 8); row sums c."""
    from fractions import Fraction as Fr

    from oracle import dopri8

    A, c, b, bh = dopri8.A_FR, dopri8.C_FR, dopri8.B_FR, dopri8.BHAT_FR
    s = dopri8.NSTAGES
    # the published rationals satisfy the conditions to about 1e-16 (they round
    # irrational optimal values)
    assert all(abs(float(sum(A[i]) - c[i])) < 1e-15 for i in range(s))
    memo = {}

    def phi(t):
        if t not in memo:
            v = [Fr(1)] * s
            for ch in t:
                pc = phi(ch)
                w = [sum(A[i][j] * pc[j] for j in range(s)) for i in range(s)]
                v = [v[i] * w[i] for i in range(s)]
            memo[t] = v
        return memo[t]

    def residual(w, t):
        return abs(float(sum(wi * pi for wi, pi in zip(w, phi(t))) - Fr(1, _tree_gamma(t))))

    n8 = n7 = 0
    for q in range(1, 9):
        for t in _rooted_trees(q):
            assert residual(b, t) < 1e-14, (q, t)
            n8 += 1
            if q <= 7:
                assert residual(bh, t) < 1e-14, (q, t)
                n7 += 1
    assert n8 == 200 and n7 == 85
    assert max(residual(bh, t) for t in _rooted_trees(8)) > 1e-6  # bhat is a 7th-order pair member
    np.testing.assert_allclose(dopri8.E, dopri8.B - np.array([float(x) for x in bh]), atol=1e-17)


def test_dopri8_eighth_order_convergence_and_closed_form():
    from oracle import dopri8

    lam = -1.0 + 3.0j

    def run(h):
        y = 1.0 + 0j
        for _ in range(int(round(1.0 / h))):
            ks = [lam * y]
            for s in range(1, dopri8.NSTAGES):
                ks.append(lam * (y + h * sum(dopri8.A[s, j] * ks[j] for j in range(s))))
            y = y + h * sum(dopri8.B[j] * ks[j] for j in range(dopri8.NSTAGES))
        return abs(y - np.exp(lam))

    # at least 8th order (2^8 = 256); on the linear test equation the error falls
    # like h^9 (its stability polynomial matches exp(z) one order further)
    r = run(0.2) / run(0.1)
    assert 200 < r < 700, r
    # adaptive Dopri8 on the lossy cavity against the closed form, and against the
    # superoperator exponential on a random problem
    H, Ls, psi0 = _ocavity(16, 1.0, 0.1, 1.0)
    tsave = np.linspace(0, 2 * np.pi, 11)
    res = om.mesolve(H, Ls, psi0, tsave, rtol=1e-10, atol=1e-12, method='dopri8')
    for i, t in enumerate(tsave):
        al = np.exp(-1j * t - 0.05 * t)
        assert np.linalg.norm(res['states'][i] - op.coherent_dm(16, al)) < 1e-6
    rng = np.random.default_rng(8)
    H, Ls, rho0 = _rand_problem(rng, 6, 2)
    H /= np.linalg.norm(H, 2)
    Ls = [0.3 * L / np.linalg.norm(L, 2) for L in Ls]
    tsave = np.linspace(0, 3, 7)
    res = om.mesolve(H, Ls, rho0, tsave, rtol=1e-10, atol=1e-12, method='dopri8')
    Lsup = lindblad.slindbladian(H, Ls)
    for i, t in enumerate(tsave):
        ex = lindblad.unvectorize(sla.expm(t * Lsup) @ lindblad.vectorize(rho0))
        assert np.linalg.norm(res['states'][i] - ex) < 1e-9
    assert res['infos']['nrhs'] >= 13 * res['infos']['naccepted']


# ----------------------------------------------------------------- adaptive Rouchon


def _ar_problem():
    n = 5
    a = op.destroy(n)
    H = a.conj().T @ a + 0.2 * (a + a.conj().T)
    Ls = [0.5 * a, 0.3 * a.conj().T @ a]
    return H, Ls, op.coherent_dm(n, 0.7)


def test_adaptive_rouchon_pair_reduces_to_fixed_step_maps():
    """Constant operators: the pair's high maps are the fixed-step Rouchon2/3 steps
    (pinned above), the order-2 low map is Rouchon1, and the embedded Bosh3 solution
    has the closed form I + hG + h^2/2 G^2 + 3h^3/16 G^3 + h^4/48 G^4, i.e. second
    order (its h^3 coefficient is not 1/6)."""
    from oracle import rouchon_adaptive as ora

    H, Ls, rho0 = _ar_problem()
    h, n = 0.07, H.shape[0]
    G = -1j * H - 0.5 * sum(L.conj().T @ L for L in Ls)
    U2 = ora.propagators(ora._as_hf(H), Ls, 0.3, h, 2)
    np.testing.assert_allclose(U2['high'][1.0], rouchon.rouchon2_propagator(H, Ls, h), atol=1e-14)
    np.testing.assert_allclose(U2['low'], np.eye(n) + h * G, atol=1e-14)
    U3 = ora.propagators(ora._as_hf(H), Ls, 0.3, h, 3)
    ref3 = rouchon.rouchon3_propagators(H, Ls, h)
    for c, U in ref3.items():
        np.testing.assert_allclose(U3['high'][c], U, atol=1e-14)
    G2 = G @ G
    closed = np.eye(n) + h * G + h**2 / 2 * G2 + 3 * h**3 / 16 * G2 @ G + h**4 / 48 * G2 @ G2
    np.testing.assert_allclose(U3['low'], closed, atol=1e-13)
    for normalize in (True, False):
        hi, err = ora.pair_step(H, Ls, rho0, h, 2, normalize)
        np.testing.assert_allclose(hi, rouchon.rouchon2_step(H, Ls, rho0, h, normalize), atol=1e-14)
        np.testing.assert_allclose(hi - 2 * err, rouchon.kraus_step_dense(H, Ls, rho0, h, normalize), atol=1e-14)
        hi3, _ = ora.pair_step(H, Ls, rho0, h, 3, normalize)
        np.testing.assert_allclose(hi3, rouchon.rouchon3_step(H, Ls, rho0, h, normalize), atol=1e-14)


@pytest.mark.parametrize('order', [2, 3])
def test_adaptive_rouchon_local_error_orders(order):
    """Over one step of h: the error estimate scales as h^order (the low map's local
    error) and the high map's distance to the exact step as h^(order+1)."""
    from oracle import rouchon_adaptive as ora

    H, Ls, rho0 = _ar_problem()
    Lsup = lindblad.slindbladian(H, Ls)
    est, loc = [], []
    for h in (0.04, 0.02):
        hi, err = ora.pair_step(H, Ls, rho0, h, order)
        ex = lindblad.unvectorize(sla.expm(h * Lsup) @ lindblad.vectorize(rho0))
        est.append(np.linalg.norm(err))
        loc.append(np.linalg.norm(hi - ex))
    r_est, r_loc = est[0] / est[1], loc[0] / loc[1]
    assert 0.75 * 2**order < r_est < 1.3 * 2**order, r_est
    assert 0.75 * 2 ** (order + 1) < r_loc < 1.3 * 2 ** (order + 1), r_loc


@pytest.mark.parametrize('order', [2, 3])
def test_adaptive_rouchon_converges_to_expm(order):
    """Lossy driven oscillator: every saved state has unit trace; the error against
    expm of the Lindbladian falls with the tolerance; steps end on the save times."""
    from oracle import rouchon_adaptive as ora

    H, Ls, rho0 = _ar_problem()
    Lsup = lindblad.slindbladian(H, Ls)
    tsave = np.linspace(0.0, 1.0, 6)
    exact = [lindblad.unvectorize(sla.expm(t * Lsup) @ lindblad.vectorize(rho0)) for t in tsave]
    errs = []
    for tol in (1e-4, 1e-6):
        states, st = ora.integrate(H, Ls, rho0, tsave, order=order, rtol=tol, atol=tol)
        assert states.shape == (6, 5, 5)
        assert np.max(np.abs(np.trace(states, axis1=1, axis2=2) - 1)) < 1e-12
        assert abs(sum(st.dts) - 1.0) < 1e-12 and st.naccepted + st.nrejected == st.nsteps
        errs.append(max(np.linalg.norm(s - e) for s, e in zip(states, exact)))
    assert errs[1] < errs[0] / 5, errs
    assert errs[1] < 1e-4, errs


def test_adaptive_rouchon_otdqubit_closed_form(kats):
    """The reference's OTDQubit (H = eps cos(w t) sigma_x, one jump operator
    sqrt(gamma) sigma_x): both pairs against the closed form at the default
    tolerances, the Hamiltonian taken at every stage time."""
    from oracle import rouchon_adaptive as ora

    k = kats['otdqubit']
    eps, w, g = k['eps'], k['omega'], k['gamma']
    sx = op.sigmax()
    Hf = lambda tau, tseg: eps * np.cos(w * tau) * sx  # noqa: E731
    tsave = np.linspace(0, k['t_end'], k['ntsave'])
    for order in (2, 3):
        states, _ = ora.integrate(Hf, [np.sqrt(g) * sx], op.fock_dm(2, 0), tsave, order=order)
        for i, t in enumerate(tsave):
            th = 2 * eps / w * np.sin(w * t)
            eta = np.exp(-2 * g * t)
            want = np.array([[0.5 * (1 + eta * np.cos(th)), 0.5j * eta * np.sin(th)],
                             [-0.5j * eta * np.sin(th), 0.5 * (1 - eta * np.cos(th))]])
            assert np.linalg.norm(states[i] - want) < 1e-4, (order, i)


def test_adaptive_rouchon_pwc_steps_land_on_breakpoints():
    """A piecewise-constant drive: steps end on its breakpoints and the solution
    matches the product of exact segment propagators."""
    from oracle import rouchon_adaptive as ora

    H, Ls, rho0 = _ar_problem()
    a = op.destroy(5)
    X = a + a.conj().T
    bps = np.array([0.0, 0.33, 0.71, 1.0])
    vals = np.array([0.4, -0.8, 0.2])

    def Hf(tau, tseg):
        i = int(np.searchsorted(bps, tseg, side='right')) - 1
        return H + (vals[i] if 0 <= i < len(vals) else 0.0) * X

    states, st = ora.integrate(Hf, Ls, rho0, [0.0, 1.0], order=3, rtol=1e-8, atol=1e-8, jump_ts=bps)
    ends = np.cumsum(st.dts)
    for b in bps[1:-1]:
        assert np.min(np.abs(ends - b)) < 1e-12
    r = lindblad.vectorize(rho0)
    for i in range(3):
        r = sla.expm((bps[i + 1] - bps[i]) * lindblad.slindbladian(H + vals[i] * X, Ls)) @ r
    assert np.linalg.norm(states[-1] - lindblad.unvectorize(r)) < 1e-6
