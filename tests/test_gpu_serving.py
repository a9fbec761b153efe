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


def _serving_request(rng, n, K, T, nt=50):
    """A request of the serving mix: GUE H of spectral norm 1, K jump operators
    sqrt(gamma) G with G complex Gaussian of entry variance 1/n, random pure state."""
    A = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(2)
    H = (A + A.conj().T) / 2
    H /= np.max(np.abs(np.linalg.eigvalsh(H)))
    Ls = [np.sqrt(rng.uniform(0.01, 0.5)) * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(2 * n)
          for _ in range(K)]
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    psi /= np.linalg.norm(psi)
    return H, Ls, psi, np.linspace(0.0, T, nt)


def test_serving_mix_full_range_vs_oracle_replay():
    """The serving workload's range: n from 8 to 256 (the largest on the grid-wide
    Tsit5), T up to 20, 50 save times, rtol 1e-5. Every request matches the complex128
    oracle replaying its own accepted steps to 1e-5, and its solve alone."""
    from dynamiqs_amd.serving import Server
    from oracle import tsit5

    rng = np.random.default_rng(21)
    shapes = [(8, 1, 20.0), (24, 3, 15.0), (44, 2, 20.0), (64, 1, 12.0), (90, 3, 8.0), (128, 2, 5.0), (200, 1, 3.0),
              (256, 3, 2.0)]
    reqs = [_serving_request(rng, n, K, T) for n, K, T in shapes]
    with Server(memory_budget=4 << 30) as srv:
        futs = [srv.submit(H, Ls, psi, ts, rtol=1e-5, step_log_cap=20_000) for (H, Ls, psi, ts) in reqs]
        mix = [f.result(timeout=300) for f in futs]
    for i, ((H, Ls, psi, ts), m) in enumerate(zip(reqs, mix)):
        print(f'serving request n={H.shape[0]} K={len(Ls)} T={ts[-1]}: {m.nsteps} steps', flush=True)
        assert m.status == 0 and m.step_ends is not None and len(m.step_ends) == m.naccepted
        LdL = sum(L.conj().T @ L for L in Ls)
        Hnh = -1j * H - 0.5 * LdL
        Cs = [0.5 * L.conj().T for L in Ls]

        def f(t, y, tseg, Hnh=Hnh, Ls=Ls, Cs=Cs):
            tmp = Hnh @ y
            for L, C in zip(Ls, Cs):
                tmp = tmp + (L @ y) @ C
            return tmp + tmp.conj().T

        saved, _ = tsit5.integrate_steps(f, psi @ psi.conj().T, ts, m.step_ends)
        err = np.linalg.norm(m.states - np.stack(saved), axis=(-2, -1)).max()
        assert err <= TOL, (i, err)
    with Server(memory_budget=4 << 30) as srv:
        for i in (0, 4, 7):
            H, Ls, psi, ts = reqs[i]
            a = srv.submit(H, Ls, psi, ts, rtol=1e-5).result(timeout=300)
            np.testing.assert_array_equal(a.states, mix[i].states)


def test_requests_retire_before_their_launch_ends():
    """A short request packed into one launch with a long one is answered when its
    own solve is done (its completion word), long before the launch ends."""
    from dynamiqs_amd.serving import Server

    rng = np.random.default_rng(8)
    short = _serving_request(rng, 40, 1, 0.5, nt=5)
    long_ = _serving_request(rng, 64, 3, 200.0)
    with Server(memory_budget=1 << 30) as srv:
        f_long = srv.submit(*long_, rtol=1e-8, atol=1e-10)
        f_short = srv.submit(*short, rtol=1e-5)
        r_short = f_short.result(timeout=120)
        r_long = f_long.result(timeout=120)
        stats = dict(srv.stats)
    print(f'short {r_short.latency_s * 1e3:.2f} ms ({r_short.nsteps} steps), long {r_long.latency_s * 1e3:.2f} ms '
          f'({r_long.nsteps} steps), launches {stats["launches"]}', flush=True)
    if stats['launches'] == 1:  # both in one launch (the usual case): the short one left early
        assert stats['answered_before_launch_end'] >= 1
    assert r_short.latency_s < 0.5 * r_long.latency_s


def test_device_operator_preparation():
    """dqh_dense_prepare_ops: Hnh = -iH - 1/2 sum L^dag L and C_k = L_k^dag / 2 on
    the device, from -iH and the complex64 jump operators, for requests of several
    padded sizes and jump counts (zero padding preserved)."""
    import torch

    from dynamiqs_amd import _native, serving
    from dynamiqs_amd.integrators.dense import pad16

    rng = np.random.default_rng(3)
    reqs = []
    for n, K in [(8, 1), (37, 3), (100, 2), (256, 1), (20, 0)]:
        H, Ls, psi, ts = _random_request(rng, n, K)
        reqs.append(serving.Request(H=H, Ls=Ls, rho0=psi, tsave=ts))
    stream = torch.cuda.Stream()
    keep, outs, host_out, ev, _ = serving.pack_and_launch(reqs, stream)
    ev.synchronize()
    arena = keep[1].cpu().numpy()
    lays = [serving._layout(r.n, len(r.Ls), len(r.tsave), len(r.exp_ops), r.step_log_cap) for r in reqs]
    in_off = np.cumsum([0] + [lay[1] for lay in lays])
    for b, r in enumerate(reqs):
        n, K, npd = r.n, len(r.Ls), pad16(r.n)
        nn = npd * npd
        li = lays[b][0]
        ib = int(in_off[b])
        Hnh = arena[ib + li['H'] : ib + li['H'] + nn * 8].view(np.complex64).reshape(npd, npd)
        L64 = [L.astype(np.complex64).astype(np.complex128) for L in r.Ls]
        want = -1j * r.H.astype(np.complex64).astype(np.complex128) - 0.5 * sum((L.conj().T @ L for L in L64), np.zeros((n, n)))
        scale = max(1.0, np.abs(want).max())
        assert np.abs(Hnh[:n, :n] - want).max() <= 2e-7 * scale, (n, K)
        assert not np.any(Hnh[n:, :]) and not np.any(Hnh[:, n:])
        for k in range(K):
            C = arena[ib + li['C'] + k * nn * 8 : ib + li['C'] + (k + 1) * nn * 8].view(np.complex64).reshape(npd, npd)
            np.testing.assert_array_equal(C[:n, :n], (0.5 * L64[k].conj().T).astype(np.complex64))
            assert not np.any(C[n:, :]) and not np.any(C[:, n:])
    del keep, outs, host_out


def test_work_queue_relaunches_and_matches_batch_kernel():
    """Small requests go to the persistent work-queue kernel. With a short workgroup
    lifetime and idle time the grid exits and is relaunched many times while requests
    arrive in waves; every result equals the same request solved by the batch kernel
    (dqh_mesolve_tsit5_dense, the same bodies) and is answered exactly once."""
    import time

    import torch

    from dynamiqs_amd import serving

    rng = np.random.default_rng(17)
    shapes = [(int(n), int(K)) for n, K in zip(rng.integers(4, 90, size=24), rng.integers(1, 4, size=24))]
    reqs = [_random_request(rng, n, K) for n, K in shapes]
    got = []
    with serving.Server(memory_budget=1 << 30, max_batch=5, queue_workgroups=32, idle_ms=0.2, life_ms=0.5) as srv:
        for w in range(4):
            futs = [srv.submit(H, Ls, psi, ts, rtol=1e-6, atol=1e-8) for (H, Ls, psi, ts) in reqs[6 * w : 6 * w + 6]]
            got += [f.result(timeout=120) for f in futs]
            time.sleep(0.01)  # every workgroup leaves (idle) before the next wave
    stats = dict(srv.stats)
    print(f"queue launches {stats['queue_launches']}, batches {stats['launches']}, "
          f"queue busy {stats['queue_busy_s'] * 1e3:.2f} ms", flush=True)
    assert stats['queue_launches'] >= 4 and stats['requests'] == len(reqs)
    assert stats['queue_busy_s'] > 0
    sreqs = [serving.Request(H=H, Ls=Ls, rho0=psi, tsave=ts, rtol=1e-6, atol=1e-8) for (H, Ls, psi, ts) in reqs]
    stream = torch.cuda.Stream()
    keep, outs, host_out, ev, _ = serving.pack_and_launch(sreqs, stream)
    ev.synchronize()
    for b, (o, g) in enumerate(zip(outs, got)):
        st, ex, ns, na, nr, status, _ = serving._read_outputs(host_out, o)
        assert status == 0 and g.status == 0
        assert (g.nsteps, g.naccepted, g.nrejected) == (ns, na, nr), b
        np.testing.assert_allclose(g.states, st, rtol=0, atol=1e-6)
    del keep
