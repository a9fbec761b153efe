"""GPU parity of the grid-wide dense Tsit5 (``dqh_mesolve_tsit5_dense_grid``):
against the complex128 oracle (1e-5 Frobenius on every saved state), against the
persistent one-workgroup kernel on the same descriptors, and the captured
hipGraph against plain launches (bit-identical)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _random_instances(rng, n, K, B):
    from dynamiqs_amd.integrators.dense import Instance

    insts = []
    for _ in range(B):
        A = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(2)
        H = (A + A.conj().T) / 2
        H /= np.max(np.abs(np.linalg.eigvalsh(H)))
        Ls = [np.sqrt(rng.uniform(0.01, 0.5)) * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(2 * n)
              for _ in range(K)]
        psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
        psi /= np.linalg.norm(psi)
        insts.append(Instance(H=H, Ls=Ls, rho0=psi @ psi.conj().T))
    return insts


@pytest.mark.parametrize('n,K,B', [(32, 1, 3), (144, 2, 2), (200, 3, 1)])
def test_grid_tsit5_matches_oracle(n, K, B):
    from dynamiqs_amd.integrators import dense
    from oracle import mesolve as om

    rng = np.random.default_rng(n)
    insts = _random_instances(rng, n, K, B)
    tsave = np.linspace(0.0, 2.0, 6)
    E = np.diag(np.arange(n, dtype=float)).astype(complex)
    plan = dense.build_plan('tsit5', insts, tsave, exp_ops=[E], rtol=1e-6, atol=1e-6, grid_tsit5=True)
    assert plan.grid is not None
    plan.run()
    out = plan.results()
    assert np.all(out['status'] == 0)
    for b, inst in enumerate(insts):
        ref = om.mesolve(inst.H, inst.Ls, inst.rho0, tsave, exp_ops=[E], method='tsit5', rtol=1e-6, atol=1e-6)
        err = np.linalg.norm(out['states'][b] - ref['states'], axis=(-2, -1)).max()
        assert err <= TOL, (n, b, err)
        np.testing.assert_allclose(out['final_state'][b], out['states'][b][-1], atol=0)
        assert np.abs(out['expects'][b, 0] - ref['expects'][0]).max() <= 1e-3
        nref = int(ref['infos']['nsteps'])
        assert abs(int(out['nsteps'][b]) - nref) <= max(2, nref // 20), (out['nsteps'][b], nref)


def test_grid_equals_persistent_kernel_and_graph_equals_plain():
    from dynamiqs_amd.integrators import dense

    rng = np.random.default_rng(5)
    insts = _random_instances(rng, 96, 2, 2)
    tsave = np.linspace(0.0, 1.0, 4)
    outs = {}
    for name, kw in (('persistent', dict(grid_tsit5=False)), ('graph', dict(grid_tsit5=True, graph_steps=4)),
                     ('plain', dict(grid_tsit5=True, graph_steps=0))):
        plan = dense.build_plan('tsit5', insts, tsave, **kw)
        plan.run()
        outs[name] = plan.results()
    assert np.array_equal(outs['graph']['states'], outs['plain']['states'])
    assert np.array_equal(outs['graph']['nsteps'], outs['plain']['nsteps'])
    err = np.abs(outs['graph']['states'].astype(complex) - outs['persistent']['states']).max()
    assert err <= TOL, err


def test_dense_above_256_through_mesolve():
    """Dense operators above n = 256 (up to np = 1024): the plan takes the grid-wide
    Tsit5 whatever the batch size, and dq.mesolve matches the complex128 oracle."""
    import dynamiqs_amd as dq
    from dynamiqs_amd.integrators import dense
    from oracle import mesolve as om

    rng = np.random.default_rng(300)
    n = 300
    insts = _random_instances(rng, n, 2, 1)
    tsave = np.linspace(0.0, 1.0, 5)
    plan = dense.build_plan('tsit5', insts, tsave, rtol=1e-6, atol=1e-6)
    assert plan.grid is not None and plan.n == n
    assert dense.use_grid('auto', 600, 128, 0, 0, True)  # large n: grid-wide for any batch
    inst = insts[0]
    print(f'dense n={n} through dq.mesolve', flush=True)
    r = dq.mesolve(inst.H, inst.Ls, inst.rho0, tsave, method=dq.method.Tsit5(rtol=1e-6, atol=1e-6))
    ref = om.mesolve(inst.H, inst.Ls, inst.rho0, tsave, method='tsit5', rtol=1e-6, atol=1e-6)
    err = np.linalg.norm(r.states.to_numpy() - ref['states'], axis=(-2, -1)).max()
    assert err <= TOL, err
