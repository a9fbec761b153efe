"""Continuous-batching server (``dynamiqs_amd.serving``) on the GPU: uneven
requests packed into shared launches return the same result as when solved
alone, match the complex128 oracle to 1e-5, and requests submitted while others
run are admitted as memory frees up."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _random_request(rng, n, K):
    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = H + H.conj().T
    H /= np.linalg.norm(H, 2)
    Ls = [np.sqrt(rng.uniform(0.01, 0.5)) * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(2 * n) for _ in range(K)]
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    psi /= np.linalg.norm(psi)
    T = rng.uniform(1.0, 3.0)
    return H, Ls, psi, np.linspace(0.0, T, 9)


def test_mix_equals_alone_and_oracle():
    from dynamiqs_amd.serving import Server
    from oracle import mesolve as om

    rng = np.random.default_rng(11)
    shapes = [(8, 1), (20, 2), (37, 3), (64, 2), (100, 1), (16, 3), (48, 1), (130, 2)]
    reqs = [_random_request(rng, n, K) for n, K in shapes]
    with Server(memory_budget=1 << 30) as srv:
        futs = [srv.submit(H, Ls, psi, ts, rtol=1e-6, atol=1e-8) for (H, Ls, psi, ts) in reqs]
        mix = [f.result(timeout=120) for f in futs]
        launches_mix = srv.stats['launches']
    assert launches_mix >= 1
    with Server(memory_budget=1 << 30) as srv:
        alone = []
        for (H, Ls, psi, ts) in reqs:
            alone.append(srv.submit(H, Ls, psi, ts, rtol=1e-6, atol=1e-8).result(timeout=120))
    for i, ((H, Ls, psi, ts), a, m) in enumerate(zip(reqs, alone, mix)):
        # batch invariance: bit-identical alone and in the mix
        np.testing.assert_array_equal(a.states, m.states)
        assert a.nsteps == m.nsteps
        if i < 5:
            ref = om.mesolve(H, Ls, psi, ts, rtol=1e-6, atol=1e-8)
            err = np.linalg.norm(m.states - ref['states'], axis=(-2, -1)).max()
            assert err <= TOL, (i, err)


def test_continuous_admission_under_memory_budget():
    from dynamiqs_amd.serving import Server, footprint_bytes

    rng = np.random.default_rng(5)
    reqs = [_random_request(rng, 32, 1) for _ in range(12)]
    budget = 3 * footprint_bytes(32, 1, 9) + 1024  # three requests in flight at most
    with Server(memory_budget=budget, max_batch=2) as srv:
        futs = [srv.submit(H, Ls, psi, ts) for (H, Ls, psi, ts) in reqs]
        res = [f.result(timeout=120) for f in futs]
        stats = dict(srv.stats)
    assert stats['launches'] >= 6
    assert stats['max_inflight_bytes'] <= budget
    assert all(r.status == 0 for r in res)
