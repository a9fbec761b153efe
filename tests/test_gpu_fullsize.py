"""Parity at the full sizes of the benchmark workloads, against committed FP64
fixtures written by the scripts in tests/golden/:

* sparse-DIA Rouchon1 on the 1024-dimensional two-mode problem over its whole
  length (dt = 1e-3, 5000 steps on [0, 5], 51 save times), against the complex128
  oracle's solve of the same scheme (``make_c3_fixture.py``). Tolerance 1e-5 on
  <a^dag a>, <b^dag b>, Tr(rho), Tr(rho^2) at every save time and in Frobenius
  norm on the final state.
* adjoint gradients of the driven oscillator at n = 32, 100 control segments on
  [0, 10], against complex128 central finite differences of the exact piecewise-
  constant solution (``make_c5_fixture.py``). Tolerance 1e-4 relative to the
  largest gradient entry at rtol = 1e-8, and 1e-3 at the default rtol = atol = 1e-6.
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _two_mode():
    from oracle import operators as op

    a, b = op.destroy(32, 32)
    ad, bd = a.conj().T, b.conj().T
    H = 5 * ad @ a + 6 * bd @ b + 0.1 * (ad @ b + a @ bd) - 0.01 * ad @ a @ bd @ b
    return H, [np.sqrt(0.02) * a, np.sqrt(0.05) * b], a, b


def test_dia_rouchon1_full_length_matches_fp64_fixture():
    from dynamiqs_amd.integrators import dia
    from oracle import operators as op

    fx = np.load(os.path.join(GOLDEN, 'c3_dia_rouchon1_alpha2.npz'))
    H, Ls, a, b = _two_mode()
    tsave = fx['tsave']
    alphas = [float(fx['alpha']), 0.5, 1.25, 1.9]
    rho0 = np.stack([op.coherent_dm((32, 32), (al, 0.0)) for al in alphas])
    Na, Nb = a.conj().T @ a, b.conj().T @ b
    plan = dia.build_plan(H, Ls, rho0, tsave, float(fx['dt']), exp_ops=[Na, Nb])
    assert plan.nsteps == int(fx['nsteps']) == 5000
    print(f'C3 full length: B={len(alphas)}, {plan.nsteps} steps', flush=True)
    plan.run()
    out = plan.results()
    ex = out['expects'][0]
    err_a = np.abs(ex[0] - fx['expect_a']).max()
    err_b = np.abs(ex[1] - fx['expect_b']).max()
    st = out['states'][0].astype(np.complex128)
    tr = np.real(np.trace(st, axis1=-2, axis2=-1))
    pur = np.sum(np.abs(st) ** 2, axis=(-2, -1))
    err_tr = np.abs(tr - fx['trace']).max()
    err_pur = np.abs(pur - fx['purity']).max()
    # final state: on the fixture's block exactly, outside it by both tails' norms
    fin = out['final_state'][0]
    idx = fx['final_index']
    blk = fin[np.ix_(idx, idx)]
    tail = fin.copy()
    tail[np.ix_(idx, idx)] = 0
    fro = np.linalg.norm(blk - fx['final_block']) + np.linalg.norm(tail) + float(fx['final_tail_frobenius'])
    print(f'C3 full length vs FP64: <a^dag a> {err_a:.2e}, <b^dag b> {err_b:.2e}, trace {err_tr:.2e}, '
          f'purity {err_pur:.2e}, final Frobenius <= {fro:.2e}', flush=True)
    assert err_a <= 1e-5 and err_b <= 1e-5
    assert err_tr <= 1e-5 and err_pur <= 1e-5
    assert fro <= 1e-5
    # every instance of the batch stays a density matrix
    for i in range(len(alphas)):
        f = out['final_state'][i]
        assert abs(np.trace(f) - 1) < 1e-6
        assert np.abs(f - f.conj().T).max() < 1e-7


def _c5_problem():
    from oracle import operators as op

    a = op.destroy(32)
    return a.conj().T @ a, a + a.conj().T, [np.sqrt(0.05) * a], op.fock_dm(32, 0), a.conj().T @ a


@pytest.mark.parametrize('tol_kw,bound', [({'rtol': 1e-8, 'atol': 1e-10}, 1e-4), ({}, 1e-3)])
def test_adjoint_full_size_matches_fd_fixture(tol_kw, bound):
    from dynamiqs_amd.adjoint import mesolve_pwc_grad

    fx = np.load(os.path.join(GOLDEN, 'c5_adjoint_n32_fd.npz'))
    H0, X, Ls, rho0, O = _c5_problem()
    u, times = fx['u'], fx['times']
    J, g = mesolve_pwc_grad(H0, [X], times, u, Ls, rho0, O, checkpoint_every=10, **tol_kw)
    fd = fx['grad_fd']
    scale = np.abs(fd).max()
    err = np.abs(g - fd).max()
    errJ = np.abs(J - fx['J']).max()
    print(f'C5 n=32 100 segments {tol_kw or "default tolerances"}: |dJ - FD| {err:.2e} '
          f'({err / scale:.2e} of max |FD|), |J - J_exact| {errJ:.2e}', flush=True)
    assert errJ <= 1e-5 if tol_kw else errJ <= 1e-4
    assert err <= bound * scale
