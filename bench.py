"""Benchmark of the MI355X mesolve hot path.

Default workload (one "step" = one batched solve of the whole batch, inputs
resident in HBM): batched dense Tsit5 on the two-photon driven Kerr oscillator,
n = 64, H_b = -K a^dag^2 a^2 + eps2_b (a^dag^2 + a^2) with K = 1 and
eps2_b ~ U[0.5, 4] (seeded), jump operators sqrt(0.01) a and sqrt(1.0) a^2,
rho0 = vacuum, 201 save times on [0, 10], Tsit5 rtol = atol = 1e-6, complex64
operators and states, B = 256 instances.

Printed JSON line: solves/s of the batched path (``value``), the plain
one-solve-per-call GPU path (the faster of its two single-solve forms) and their
ratio, p50/p99 latencies, the FP32-MFMA roofline of the persistent Tsit5 kernel
(HIP-event timing on the launch stream), the error of 8 instances (the largest
eps2 among them) replayed by the complex128 oracle on the GPU's accepted steps,
and the CPU restatement in complex64 and complex128 timed on the host cores as
the baseline. A ``secondary`` block measures the other workloads in the same run:
the plain n = 16 solve (C1), sparse-DIA Rouchon1 at n = 1024 (C3), the serving mix
at offered rates of 300, 500 and 700 requests/s and at saturation with sampled errors
(C4), and the adjoint
gradient against a forward solve (C5).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

# The serving measurement runs four streams of its own beside a persistent kernel; with
# HIP's default of 4 hardware queues a grid-wide launch can share one with that kernel
# and wait for its end. Set before anything initialises HIP (every GPU use is below).
if int(os.environ.get('GPU_MAX_HW_QUEUES', '4')) < 16:  # raised, never lowered
    os.environ['GPU_MAX_HW_QUEUES'] = '16'

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP32_MFMA_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0

# HBM bytes per batched launch of the resident Tsit5 kernel on the default workload
# (Kerr n=64, B=256), from two rocprofv3 --pmc runs of this script (FETCH_SIZE,
# then WRITE_SIZE) scaled by the same counters on a device copy of known size
# (--config copy: FETCH_SIZE x 2, WRITE_SIZE x 1): 258 134 603 KiB fetched,
# 294 650 707 KiB written (packed FP64 state writes and register spills of the
# FP64-diagonal body, 292 B/lane of scratch).
KERR_TRAFFIC_BYTES = 258_134_603 * 1024 * 2 + 294_650_707 * 1024
KERR_TRAFFIC_SOURCE = 'rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this script, build of 3abd21b'


def kerr_instances(B: int, n: int = 64, seed: int = 0):
    """Two-photon driven Kerr oscillators (complex128 host operators)."""
    from dynamiqs_amd.integrators.dense import Instance

    rng = np.random.default_rng(seed)
    eps2 = rng.uniform(0.5, 4.0, size=B)
    a = np.diag(np.sqrt(np.arange(1, n, dtype=np.float64)), k=1).astype(np.complex128)
    ad = a.conj().T
    a2, ad2 = a @ a, ad @ ad
    Ls = [np.sqrt(0.01) * a, np.sqrt(1.0) * a2]
    rho0 = np.zeros((n, n), dtype=np.complex128)
    rho0[0, 0] = 1.0
    insts = []
    for b in range(B):
        H = -1.0 * ad2 @ a2 + eps2[b] * (ad2 + a2)
        insts.append(Instance(H=H, Ls=Ls, rho0=rho0, key_H=(b,), key_L=(), key_rho=()))
    return insts, eps2


def trace_distance(a, b):
    d = a - b
    return 0.5 * np.abs(np.linalg.eigvalsh(0.5 * (d + d.conj().T))).sum()


class EventTimer:
    """HIP events on the stream the kernels are launched on."""

    def __init__(self, stream):
        import torch

        self.stream = stream
        self.pairs = []
        self.torch = torch

    def start(self):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record(self.stream)
        return e

    def stop(self, e0):
        e1 = self.torch.cuda.Event(enable_timing=True)
        e1.record(self.stream)
        self.pairs.append((e0, e1))

    def times_ms(self):
        self.torch.cuda.synchronize()
        return [a.elapsed_time(b) for a, b in self.pairs]


def run_plan_timed(plan, reps, stream, timer=None):
    import torch

    walls = []
    for _ in range(reps):
        plan.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0 = timer.start() if timer else None
        plan.run(stream)
        if timer:
            timer.stop(e0)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
    return walls


def two_mode_dia(na: int = 32, nb: int = 32):
    a = np.kron(np.diag(np.sqrt(np.arange(1, na, dtype=np.float64)), k=1), np.eye(nb)).astype(np.complex128)
    b = np.kron(np.eye(na), np.diag(np.sqrt(np.arange(1, nb, dtype=np.float64)), k=1)).astype(np.complex128)
    ad, bd = a.conj().T, b.conj().T
    H = 5 * ad @ a + 6 * bd @ b + 0.1 * (ad @ b + a @ bd) - 0.01 * ad @ a @ bd @ b
    return H, [np.sqrt(0.02) * a, np.sqrt(0.05) * b], a, b


def coherent_ket(n: int, alpha: float) -> np.ndarray:
    from scipy.linalg import expm

    a = np.diag(np.sqrt(np.arange(1, n, dtype=np.float64)), k=1).astype(np.complex128)
    v = np.zeros(n, dtype=np.complex128)
    v[0] = 1
    return expm(alpha * a.conj().T - np.conj(alpha) * a) @ v


def bench_dia(args):
    """Sparse-DIA Rouchon1: two coupled 32-level modes (n = 1024), B = 16 coherent
    initial states alpha ~ U[0.5, 2], dt = 1e-3 on [0, 5], 51 save times."""
    import torch

    from dynamiqs_amd.integrators import dia

    H, Ls, a, b = two_mode_dia()
    rng = np.random.default_rng(1)
    B = args.batch if args.batch != 256 else 16
    alphas = rng.uniform(0.5, 2.0, size=B)
    rho0 = []
    for al in alphas:
        v = np.kron(coherent_ket(32, al), np.eye(32)[:, 0])
        rho0.append(np.outer(v, v.conj()))
    rho0 = np.stack(rho0)
    tsave = np.linspace(0.0, args.dia_T, int(round(args.dia_T / 0.1)) + 1)
    variant = {'auto': dia.SWEEP_AUTO, 'global': dia.SWEEP_GLOBAL, 'tiled': dia.SWEEP_TILED, 'split': dia.SWEEP_SPLIT}[
        args.dia_variant
    ]
    if args.dia_method == 'tsit5':
        return bench_dia_tsit5(args, H, Ls, rho0, tsave, variant)
    plan = dia.build_plan(H, Ls, rho0, tsave, 1e-3, variant=variant)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        plan.reset()
        plan.run(stream)
    torch.cuda.synchronize()
    timer = EventTimer(stream)
    walls = []
    for _ in range(args.steps):
        plan.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0 = timer.start()
        plan.run(stream)
        timer.stop(e0)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
    kern = timer.times_ms()
    res = plan.results()
    tr = np.abs(np.trace(res['final_state'], axis1=-2, axis2=-1) - 1).max()
    out = {
        'metric': 'DIA Rouchon1 solves/s',
        'value': B / float(np.mean(walls)),
        'unit': 'solves/s',
        'config': {'workload': f'sparse-DIA Rouchon1, 32x32 two-mode n=1024, B={B}, dt=1e-3, T={args.dia_T}'},
        'ms_per_solve_batch': 1e3 * float(np.mean(walls)),
        'gpu_ms': float(np.mean(kern)),
        'nsteps': plan.nsteps,
        'us_per_step': 1e3 * float(np.mean(kern)) / plan.nsteps,
        'sweep_algorithmic_GBps': plan.sweep_bytes_per_step * plan.nsteps / (1e-3 * float(np.mean(kern))) / 1e9,
        'solve_TFLOPs': plan.solve_flops_per_step * plan.nsteps / (1e-3 * float(np.mean(kern))) / 1e12,
        'max_trace_error': float(tr),
        'sweep_variant': args.dia_variant,
        'rhs_sweep': dia_rhs_sweeps(H, Ls, rho0, stream),
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_dia_tsit5(args, H, Ls, rho0, tsave, variant):
    """The same two-mode batch with adaptive Tsit5 on the DIA sweeps (rtol = atol =
    1e-6): solves/s and the algorithmic bandwidth of its vector fields."""
    import torch

    from dynamiqs_amd.integrators import dia

    B = rho0.shape[0]
    plan = dia.tsit5_plan(H, Ls, rho0, tsave, variant=variant)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        plan.reset()
        plan.run(stream)
    timer = EventTimer(stream)
    walls = run_plan_timed(plan, args.steps, stream, timer)
    kern = float(np.mean(timer.times_ms()))
    res = plan.results()
    nrhs = float(np.sum(res['nrhs']))
    tr = np.abs(np.trace(res['final_state'], axis1=-2, axis2=-1) - 1).max()
    out = {
        'metric': 'DIA Tsit5 solves/s',
        'value': B / float(np.mean(walls)),
        'unit': 'solves/s',
        'config': {'workload': f'sparse-DIA Tsit5, 32x32 two-mode n=1024, B={B}, rtol=atol=1e-6, T={args.dia_T}'},
        'gpu_ms': kern,
        'nsteps_mean': float(np.mean(res['nsteps'])),
        'nrejected_mean': float(np.mean(res['nrejected'])),
        'rhs_algorithmic_GBps': plan.bytes_per_rhs * nrhs / (1e-3 * kern) / 1e9,
        'max_trace_error': float(tr),
        'status': [int(x) for x in np.unique(res['status'])],
    }
    print(json.dumps(out), flush=True)
    return 0


def dia_rhs_sweeps(H, Ls, rho0, stream, reps: int = 20):
    """The DIA Lindblad vector field alone (dqh_lindblad_rhs_dia) on the same batch,
    both sweep forms: mean kernel time and algorithmic GB/s (read rho + write drho)."""
    from dynamiqs_amd.integrators import dia

    out = {}
    for name, v in (('tiled', dia.SWEEP_TILED), ('global', dia.SWEEP_GLOBAL)):
        rp = dia.rhs_plan(H, Ls, rho0, variant=v)
        rp.run(stream)
        timer = EventTimer(stream)
        for _ in range(reps):
            e0 = timer.start()
            rp.run(stream)
            timer.stop(e0)
        ms = float(np.mean(timer.times_ms()))
        out[name] = {'us': 1e3 * ms, 'GBps': rp.bytes_per_call / (1e-3 * ms) / 1e9}
    return out


def serving_requests(N: int, seed: int = 2):
    """Random serving mix: n log-uniform in {8..256}, GUE H scaled to spectral norm 1,
    1-3 jump ops sqrt(gamma) G (G complex Gaussian, entry variance 1/n, gamma ~ U[0.01, 0.5]),
    random pure rho0, 50 save times on [0, T], T ~ U[1, 20], Tsit5 rtol = 1e-5."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(N):
        n = int(round(np.exp(rng.uniform(np.log(8), np.log(256)))))
        A = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(2)
        H = (A + A.conj().T) / 2
        H /= np.max(np.abs(np.linalg.eigvalsh(H)))
        K = int(rng.integers(1, 4))
        Ls = [np.sqrt(rng.uniform(0.01, 0.5)) * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(2 * n) for _ in range(K)]
        psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
        psi /= np.linalg.norm(psi)
        T = rng.uniform(1.0, 20.0)
        reqs.append((H, Ls, psi, np.linspace(0.0, T, 50)))
    return reqs


def bench_serving(args):
    """Poisson arrivals into the continuous-batching server at several rates."""
    import threading

    from dynamiqs_amd.serving import Server

    reqs = serving_requests(args.requests)
    ns = np.array([r[0].shape[0] for r in reqs])
    rows = []
    for rate in args.rates:
        rng = np.random.default_rng(3)
        gaps = rng.exponential(1.0 / rate, size=len(reqs)) if rate > 0 else np.zeros(len(reqs))
        with Server(memory_budget=int(args.budget_gb * (1 << 30))) as srv:
            futs = [None] * len(reqs)

            def client():
                t_next = time.perf_counter()
                for i, (H, Ls, psi, ts) in enumerate(reqs):
                    t_next += gaps[i]
                    dt = t_next - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                    futs[i] = srv.submit(H, Ls, psi, ts, rtol=1e-5)

            t0 = time.perf_counter()
            th = threading.Thread(target=client)
            th.start()
            th.join()
            res = [f.result() for f in futs]
            wall = time.perf_counter() - t0
            st = dict(srv.stats)
        lat = np.array([r.latency_s for r in res])
        rows.append({
            'offered_rate': rate if rate > 0 else 'all-at-once',
            'solves_per_s': len(res) / wall,
            'p50_latency_s': float(np.percentile(lat, 50)),
            'p99_latency_s': float(np.percentile(lat, 99)),
            'launches': st['launches'],
            'max_concurrent_launches': st['max_concurrent_launches'],
            'mean_steps': float(np.mean([r.nsteps for r in res])),
        })
        print(json.dumps(rows[-1]), flush=True)
    out = {
        'metric': 'serving solves/s and p50/p99 latency',
        'config': {'workload': f'serving mix, {len(reqs)} requests, n log-uniform 8..256 (median {int(np.median(ns))}), K 1-3, T U[1,20], 50 tsave, rtol 1e-5'},
        'rates': rows,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_adjoint(args):
    """Adjoint gradient vs forward cost: n = 32, H = a^dag a + u_k (a + a^dag) with 100 PWC
    segments on [0, 10], u ~ N(0, 0.5), L = sqrt(0.05) a, rho0 = vacuum, loss Tr(a^dag a rho(T)),
    B = 64 control vectors, checkpoints every 10 segments."""
    import torch

    from dynamiqs_amd.adjoint import AdjointPlan, mesolve_pwc_grad
    from dynamiqs_amd.integrators import dense

    n, nseg, B = 32, 100, 64
    a = np.diag(np.sqrt(np.arange(1, n, dtype=np.float64)), k=1).astype(np.complex128)
    H0 = a.conj().T @ a
    X = a + a.conj().T
    times = np.linspace(0.0, 10.0, nseg + 1)
    u = np.random.default_rng(4).normal(0.0, 0.5, size=(B, 1, nseg))
    Ls = [np.sqrt(0.05) * a]
    rho0 = np.zeros((n, n), dtype=np.complex128)
    rho0[0, 0] = 1
    O = a.conj().T @ a
    # forward alone (same solve the gradient starts with), timed as one launch
    insts = [dense.Instance(H=H0, Ls=Ls, rho0=rho0, pwc=[(times, u[b, 0], X)], key_L=(), key_rho=()) for b in range(B)]
    fwd = dense.build_plan('tsit5', insts, np.array([0.0, 10.0]), exp_ops=[O])
    fwd.run()
    torch.cuda.synchronize()
    tf = []
    for _ in range(max(args.steps, 1)):
        fwd.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fwd.run()
        torch.cuda.synchronize()
        tf.append(time.perf_counter() - t0)
    # one-shot gradient (packs the three plans, then runs)
    t0 = time.perf_counter()
    J1, g1, tim1 = mesolve_pwc_grad(H0, [X], times, u, Ls, rho0, O, checkpoint_every=10, return_timing=True)
    one_shot = time.perf_counter() - t0
    # repeated gradients on a packed plan (new controls each call, as in an optimisation loop)
    plan = AdjointPlan(H0, [X], times, Ls, rho0, O, B, checkpoint_every=10)
    plan.run(u)  # warm
    rng = np.random.default_rng(5)
    tg, tim = [], None
    for _ in range(max(args.steps, 1)):
        uk = u + 0.01 * rng.normal(size=u.shape)
        t0 = time.perf_counter()
        res = plan.run(uk)
        tg.append(time.perf_counter() - t0)
        tim = res.timing
    res = plan.run(u)
    # sum of the kernel times (the backward overlaps the forward on its own stream)
    gpu_grad = tim['forward_run_s'] + tim['recompute_run_s'] + tim['backward_run_s'] + tim['quadrature_s']
    out = {
        'metric': 'adjoint gradient cost / forward solve',
        'config': {'workload': 'PWC-driven oscillator n=32, 100 segments on [0,10], B=64, loss Tr(a^dag a rho(T)), checkpoints every 10 segments'},
        'forward_s': float(np.median(tf)),
        'gradient_wall_s': float(np.median(tg)),
        'gradient_gpu_s': gpu_grad,
        'gpu_ratio': gpu_grad / float(np.median(tf)),
        'wall_ratio': float(np.median(tg)) / float(np.median(tf)),
        'one_shot_wall_s': one_shot,
        'plan_pack_s': plan.pack_timing['plan_total_s'],
        'phases_s': tim,
        'loss_mean': float(np.mean(res.J)),
        'one_shot_matches_plan': bool(np.allclose(res.grad_u, g1, rtol=0, atol=1e-9) and np.allclose(res.J, J1)),
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_grid(args):
    """One large dense request (n = 256, 3 jump operators, GUE H, T = 10, rtol = 1e-5):
    latency of the persistent one-workgroup kernel against the grid-wide Tsit5."""
    import torch

    from dynamiqs_amd.integrators import dense

    reqs = serving_requests(64, seed=7)
    H, Ls, psi, tsave = next(r for r in reqs if r[0].shape[0] >= 200 and len(r[1]) == 3)
    n = H.shape[0]
    inst = dense.Instance(H=H, Ls=Ls, rho0=psi @ psi.conj().T)
    stream = torch.cuda.current_stream()
    out = {'metric': 'large-request latency', 'config': {'workload': f'dense Tsit5 n={n}, K=3, T={tsave[-1]:.2f}, rtol=1e-5'}}
    states = {}
    for name, g in (('persistent', False), ('grid', True)):
        plan = dense.build_plan('tsit5', [inst], tsave, rtol=1e-5, atol=1e-6, grid_tsit5=g)
        run_plan_timed(plan, 1, stream)
        walls = run_plan_timed(plan, max(args.steps, 1), stream)
        res = plan.results()
        states[name] = res['states'][0]
        out[name] = {'latency_s': float(np.median(walls)), 'nsteps': int(res['nsteps'][0]), 'status': int(res['status'][0])}
    out['speedup'] = out['persistent']['latency_s'] / out['grid']['latency_s']
    out['max_diff'] = float(np.abs(states['grid'].astype(complex) - states['persistent']).max())
    print(json.dumps(out), flush=True)
    return 0


def bench_copy(args):
    """A plain device-to-device copy of known size (512 MiB read, 512 MiB written):
    the calibration run for the FETCH_SIZE / WRITE_SIZE counters, and the HBM
    bandwidth a streaming kernel reaches on this device."""
    import torch

    nbytes = 512 << 20
    x = torch.ones(nbytes // 4, dtype=torch.float32, device='cuda')
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    y.copy_(x)
    timer = EventTimer(stream)
    for _ in range(max(args.steps, 1)):
        e0 = timer.start()
        y.copy_(x)
        timer.stop(e0)
    ms = float(np.median(timer.times_ms()))
    print(json.dumps({'metric': 'device copy', 'bytes_read': nbytes, 'bytes_written': nbytes, 'ms': ms,
                      'GBps': 2 * nbytes / (1e-3 * ms) / 1e9, 'launches': max(args.steps, 1) + 1}), flush=True)
    return 0


def bench_c1(args):
    """The plain single-solve configuration: lossy oscillator n=16, H = a^dag a,
    L = sqrt(0.1) a, rho0 = coherent(1.0), 101 save times on [0, 2 pi], one instance,
    through ``dq.mesolve`` end to end (host packing, launch, read-back). Measured at
    the configuration's tolerances (rtol 1e-6, atol 1e-8) and at the dynamiqs README
    example's (atol 1e-6; README.md:48-77 prints 6.30 ms on unstated hardware)."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om
    from oracle import operators as op

    n = 16
    a = dq.destroy(n)
    H, Ls, psi = a.dag() @ a, [np.sqrt(0.1) * a], dq.coherent(n, 1.0)
    tsave = np.linspace(0.0, 2 * np.pi, 101)
    rows = {}
    for name, atol in (('rtol 1e-6, atol 1e-8', 1e-8), ('rtol 1e-6, atol 1e-6 (README example)', 1e-6)):
        method = dq.method.Tsit5(rtol=1e-6, atol=atol)
        for _ in range(max(args.warmup, 1)):
            dq.mesolve(H, Ls, psi, tsave, method=method)
        lat = []
        for _ in range(max(args.steps, 5)):
            t0 = time.perf_counter()
            res = dq.mesolve(H, Ls, psi, tsave, method=method)
            lat.append(time.perf_counter() - t0)
        ao = op.destroy(n)
        ref = om.mesolve(ao.conj().T @ ao, [np.sqrt(0.1) * ao], op.coherent(n, 1.0), tsave, rtol=1e-6, atol=atol)
        st = res.states.to_numpy()
        rows[name] = {
            'p50_latency_ms': 1e3 * float(np.median(lat)),
            'min_latency_ms': 1e3 * float(np.min(lat)),
            'nsteps': int(res.infos.nsteps),
            'oracle_nsteps': int(ref['infos']['nsteps']),
            'max_frobenius_vs_fp64': float(np.linalg.norm(st - ref['states'], axis=(-2, -1)).max()),
        }
    print(json.dumps({'metric': 'plain mesolve latency (one solve per call)', 'unit': 'ms',
                      'config': {'workload': 'lossy oscillator n=16, 101 tsave on [0, 2 pi], Tsit5, one instance'},
                      'runs': rows}), flush=True)
    return 0


def two_level_exact(H: np.ndarray, psi0: np.ndarray, ts: np.ndarray) -> np.ndarray:
    """exp(-i H t) psi0 for traceless 2x2 Hermitian H (H^2 = w^2 I):
    cos(w t) psi0 - i sin(w t) / w H psi0, with sin(w t) / w -> t at w = 0.
    H: (..., 2, 2), psi0: (2,), ts: (nt,) -> (..., nt, 2)."""
    w = np.sqrt(np.abs(np.einsum('...ij,...ji->...', H, H).real) / 2)[..., None]
    c = np.cos(w * ts)
    sinc = np.where(w > 0, np.sin(w * ts) / np.where(w > 0, w, 1.0), ts)
    Hp = H @ psi0
    return c[..., None] * psi0 - 1j * sinc[..., None] * Hp[..., None, :]


def bench_sesolve(args):
    """sesolve workloads the dynamiqs documentation times (unstated hardware):
    the batching benchmark (docs/documentation/basics/batching-simulations.md:218-260:
    3000 two-level Hamiltonians omega sigma_z + epsilon sigma_x, omega in
    linspace(0, 10, 100), epsilon in linspace(0, 1, 30), psi0 = |0>, 50 save times
    on [0, 1], Tsit5 defaults; 44.1 ms batched) and the workflow example
    (docs/documentation/basics/workflow.md:24-90: 0.3 sigma_z + sigma_x, 101 save
    times on [0, 10], Dopri5(rtol=1e-6, atol=1e-8), <sigma_z>; 4.57 ms, 56 steps
    (48 accepted, 8 rejected)). Both through ``dq.sesolve`` end to end (host
    packing, launch, read-back), plus the kernel alone on the resident batch (HIP
    events); errors against the closed form exp(-i H t) psi0."""
    import torch

    import dynamiqs_amd as dq
    from dynamiqs_amd.integrators.sesolve import build_plan_constant
    from oracle import operators as op
    from oracle import sesolve as ose

    omega = np.linspace(0.0, 10.0, 100)[:, None, None]
    epsilon = np.linspace(0.0, 1.0, 30)[:, None, None, None]
    H = omega * dq.sigmaz() + epsilon * dq.sigmax()  # (30, 100, 2, 2)
    psi0 = dq.fock(2, 0)
    tsave = np.linspace(0.0, 1.0, 50)
    for _ in range(max(args.warmup, 1)):
        dq.sesolve(H, psi0, tsave)
    lat = []
    for _ in range(max(args.steps, 5)):
        t0 = time.perf_counter()
        res = dq.sesolve(H, psi0, tsave)
        lat.append(time.perf_counter() - t0)
    Hn = H.to_numpy()
    exact = two_level_exact(Hn, psi0.to_numpy()[:, 0], tsave)
    err = float(np.max(np.abs(res.states.to_numpy()[..., 0] - exact)))
    # the kernel alone: the same batch packed once, resident in HBM
    plan = build_plan_constant(Hn.reshape(-1, 2, 2), np.arange(3000), psi0.to_numpy().reshape(1, 2),
                               np.zeros(3000, dtype=np.int64), tsave)
    stream = torch.cuda.current_stream()
    timer = EventTimer(stream)
    run_plan_timed(plan, 2, stream)
    run_plan_timed(plan, max(args.steps, 5), stream, timer)
    kern = timer.times_ms()
    st = plan.results()
    # CPU oracle on a bounded sample of the batch (one solve per Hamiltonian)
    ns = max(args.cpu_sample * 300, 1)
    sample = list(range(0, 3000, max(3000 // ns, 1)))[:ns]
    Hf = Hn.reshape(-1, 2, 2)
    t0 = time.perf_counter()
    for i in sample:
        ose.sesolve(Hf[i], op.fock(2, 0), tsave)
    cpu_per = (time.perf_counter() - t0) / len(sample)
    batching = {
        'p50_ms': 1e3 * float(np.median(lat)),
        'min_ms': 1e3 * float(np.min(lat)),
        'kernel_p50_ms': float(np.median(kern)),
        'reference_published_ms': 44.1,
        'max_abs_error_vs_closed_form': err,
        'mean_steps': float(np.mean(st['nsteps'])),
        'max_steps': int(np.max(st['nsteps'])),
        'cpu_baseline': {'value_ms_for_3000': 1e3 * cpu_per * 3000, 'kind': 'port', 'cores': 1,
                         'sample': f'{len(sample)} of the 3000 Hamiltonians, one oracle solve each, scaled'},
    }
    # workflow example: one solve, Dopri5
    Hw = 0.3 * dq.sigmaz() + 1.0 * dq.sigmax()
    tw = np.linspace(0.0, 10.0, 101)
    m = dq.method.Dopri5(rtol=1e-6, atol=1e-8)
    for _ in range(max(args.warmup, 1)):
        dq.sesolve(Hw, psi0, tw, exp_ops=[dq.sigmaz()], method=m)
    lat = []
    for _ in range(max(args.steps, 5)):
        t0 = time.perf_counter()
        rw = dq.sesolve(Hw, psi0, tw, exp_ops=[dq.sigmaz()], method=m)
        lat.append(time.perf_counter() - t0)
    ew = two_level_exact(Hw.to_numpy(), psi0.to_numpy()[:, 0], tw)
    sz_exact = np.abs(ew[:, 0]) ** 2 - np.abs(ew[:, 1]) ** 2
    ref = ose.sesolve(Hw.to_numpy(), op.fock(2, 0), tw, method='dopri5', rtol=1e-6, atol=1e-8)
    workflow = {
        'p50_ms': 1e3 * float(np.median(lat)),
        'min_ms': 1e3 * float(np.min(lat)),
        'reference_published_ms': 4.57,
        'steps': [int(rw.infos.nsteps), int(rw.infos.naccepted), int(rw.infos.nrejected)],
        'reference_published_steps': [56, 48, 8],
        'oracle_fp64_steps': [ref['infos']['nsteps'], ref['infos']['naccepted'], ref['infos']['nrejected']],
        'max_abs_error_sigma_z': float(np.max(np.abs(rw.expects[0].real - sz_exact))),
    }
    # mepropagator docstring example (mepropagator.py:143-165): Expm by default
    from oracle import lindblad
    import scipy.linalg as sla

    n = 16
    a = dq.destroy(n)
    tp = np.linspace(0.0, 1.0, 11)
    dq.mepropagator(a.dag() @ a, [a], tp)
    lat = []
    for _ in range(max(args.steps, 5)):
        t0 = time.perf_counter()
        rp = dq.mepropagator(a.dag() @ a, [a], tp)
        lat.append(time.perf_counter() - t0)
    ao = op.destroy(n)
    S = lindblad.slindbladian(ao.conj().T @ ao, [ao])
    P = rp.propagators.to_numpy()
    propagator = {
        'p50_ms': 1e3 * float(np.median(lat)),
        'method': type(rp.method).__name__,
        'steps': int(rp.infos.nsteps),
        'max_abs_error_vs_scipy_expm': float(max(np.max(np.abs(P[k] - sla.expm(S * t))) for k, t in enumerate(tp))),
    }
    print(json.dumps({'metric': 'sesolve / propagator wall time (end to end)', 'unit': 'ms',
                      'config': {'workload': 'dynamiqs documentation sesolve and mepropagator examples'},
                      'batching_3000_two_level': batching, 'workflow_dopri5': workflow,
                      'mepropagator_expm_n16': propagator}), flush=True)
    return 0


# ------------------------------------------------------------------ CPU pool
# The CPU restatement runs in worker processes forked before the GPU is touched
# (one BLAS thread each), so replays and baseline timings of many instances run
# side by side on the host cores after the GPU measurements.


def _worker_init():
    os.environ['OMP_NUM_THREADS'] = '1'
    try:
        from threadpoolctl import threadpool_limits

        threadpool_limits(limits=1)
    except ImportError:  # pragma: no cover
        pass


def make_cpu_pool(nworkers: int):
    import multiprocessing as mp

    return mp.get_context('fork').Pool(nworkers, initializer=_worker_init)


def cpu_workers(requested: int) -> int:
    cores = len(os.sched_getaffinity(0))
    return max(1, min(requested if requested > 0 else 16, cores, 16))


def _lindblad_f(H, Ls, dtype):
    """The Hermitian-form vector field (diffrax_integrator.py:319-323) in ``dtype``."""
    LdL = sum(L.conj().T @ L for L in Ls)
    Hnh = (-1j * H - 0.5 * LdL).astype(dtype)
    Ds = [L.astype(dtype) for L in Ls]
    Cs = [(0.5 * L.conj().T).astype(dtype) for L in Ls]

    def f(t, y, tseg):
        tmp = Hnh @ y
        for D, C in zip(Ds, Cs):
            tmp = tmp + (D @ y) @ C
        return tmp + tmp.conj().T

    return f


def _task_replay(job):
    """FP64 oracle replay of the GPU's accepted steps; error of the GPU states."""
    from oracle import tsit5

    H, Ls, rho0, tsave, step_ends, got = job
    saved, _ = tsit5.integrate_steps(_lindblad_f(H, Ls, np.complex128), rho0, tsave, step_ends)
    ref = np.stack(saved)
    got = got.astype(np.complex128)
    fro = float(np.linalg.norm(got - ref, axis=(-2, -1)).max())
    td = float(max(trace_distance(got[j], ref[j]) for j in range(len(tsave))))
    return fro, td


def _task_adaptive(job):
    """The oracle's own adaptive Tsit5 in ``dtype`` on one instance, timed (one
    core); with ``got``, the largest Frobenius distance of the GPU states to it."""
    from oracle import tsit5

    H, Ls, rho0, tsave, rtol, atol, dtype, got = job
    f = _lindblad_f(H, Ls, dtype)
    t0 = time.perf_counter()
    saved, _, st = tsit5.integrate(f, rho0.astype(dtype), tsave, rtol=rtol, atol=atol, dtype=dtype)
    sec = time.perf_counter() - t0
    fro = None
    if got is not None:
        fro = float(np.linalg.norm(got.astype(np.complex128) - np.stack(saved), axis=(-2, -1)).max())
    return sec, st.nsteps, fro


# ------------------------------------------------------------------ headline


def bench_kerr(args, pool):
    """The headline: batched dense Tsit5 on the Kerr batch (see module docstring).
    Returns the JSON dict and the CPU jobs to run once every GPU measurement is done."""
    import torch

    from dynamiqs_amd.integrators import dense

    n, K = 64, 2
    rtol = atol = 1e-6
    tsave = np.linspace(0.0, 10.0, 201)
    insts, eps2 = kerr_instances(args.batch, n)
    # the accepted step ends of each instance are logged (one store per accepted
    # step) so that the FP64 oracle can replay the identical step sequence
    plan = dense.build_plan('tsit5', insts, tsave, rtol=rtol, atol=atol, step_log_cap=40_000)
    stream = torch.cuda.current_stream()

    # warmup + timed batched solves (one step = one batched solve of B instances)
    run_plan_timed(plan, args.warmup, stream)
    torch.cuda.synchronize()
    timer = EventTimer(stream)
    walls = run_plan_timed(plan, args.steps, stream, timer)
    torch.cuda.synchronize()
    kern_ms = timer.times_ms()
    res = plan.results()
    if np.any(res['status'] != 0):
        raise RuntimeError(f'solver status {np.unique(res["status"])}')
    solves_per_s = args.batch * args.steps / sum(walls)
    ms_per_step = 1e3 * sum(walls) / args.steps
    f_rhs = (1 + 2 * K) * 8.0 * n**3
    flops = float(np.sum(res['nrhs'])) * f_rhs
    kern_avg_s = 1e-3 * float(np.mean(kern_ms))
    achieved = flops / kern_avg_s / 1e12

    # plain path: one solve per call on the same device-resident inputs, in the
    # faster of the two single-solve forms (one persistent workgroup, or every
    # stage spread grid-wide with the step captured as a hipGraph)
    plain = {}
    for name, g in (('persistent', False), ('grid', True)) if args.plain > 0 else ():
        lat = []
        for i in range(args.plain):
            p1 = dense.build_plan('tsit5', [insts[i]], tsave, rtol=rtol, atol=atol, grid_tsit5=g)
            run_plan_timed(p1, 1, stream)
            lat.append(run_plan_timed(p1, 1, stream)[0])
        plain[name] = lat
    best = min(plain, key=lambda k: float(np.median(plain[k]))) if plain else None
    plain_lat = sorted(plain[best]) if plain else []
    plain_sps = len(plain_lat) / sum(plain_lat) if plain_lat else None

    out = {
        'metric': 'solves/s + p50/p99 latency, batched/served vs plain GPU mesolve at <=1e-5 err',
        'value': solves_per_s,
        'unit': 'solves/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_per_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'c64 (fp32 MFMA), fp64 state/norms',
        'data': 'synthetic',
        'config': {
            'workload': 'batched dense Tsit5, two-photon driven Kerr oscillator n=64, 2 jump ops, B=256, 201 tsave on [0,10], rtol=atol=1e-6',
            'global_batch': args.batch,
            'n': n,
            'jump_ops': K,
            'parallelism': 'one persistent workgroup per instance',
            'timed': 'device-resident inputs and outputs (packing and the read-back of the saved states excluded, '
                     'for the batched and the plain path alike)',
        },
        'plain_solves_per_s': plain_sps,
        'plain_form': best,
        'plain_latency_s_by_form': {k: float(np.median(v)) for k, v in plain.items()},
        'batched_vs_plain': solves_per_s / plain_sps if plain_sps else None,
        'latency_s': {
            'batched_p50': float(np.percentile(walls, 50)),
            'batched_p99': float(np.percentile(walls, 99)),
            'plain_p50': float(np.percentile(plain_lat, 50)) if plain_lat else None,
            'plain_p99': float(np.percentile(plain_lat, 99)) if plain_lat else None,
        },
        'steps_per_instance': {
            'mean': float(np.mean(res['nsteps'])),
            'max': int(np.max(res['nsteps'])),
            'rejected_mean': float(np.mean(res['nrejected'])),
        },
        'roofline': {
            'bound': 'mfma',
            'achieved': achieved,
            'peak': FP32_MFMA_PEAK_TFLOPS,
            'unit': 'TFLOP/s',
            'frac': achieved / FP32_MFMA_PEAK_TFLOPS,
            'traffic': KERR_TRAFFIC_BYTES if (args.batch == 256 and plan.uniform_np == 64) else None,
            'traffic_source': KERR_TRAFFIC_SOURCE,
            'kernel': 'dqh::k_tsit5_res<64>' if plan.uniform_np == 64 else 'dqh::k_tsit5',
            'kernel_avg_ms': float(np.mean(kern_ms)),
            'flops_per_launch': flops,
        },
        'error_vs_fp64': None,
        'cpu_baseline': None,
    }

    def finish(pool):
        if pool is None or args.cpu_sample <= 0:
            return
        # replay >= 8 instances, the largest eps2 among them, on the GPU's steps
        order = list(range(min(args.replay, args.batch)))
        big = int(np.argmax(eps2))
        if big not in order:
            order[-1] = big
        jobs_r = [(insts[i].H, insts[i].Ls, insts[i].rho0, tsave, res['step_ends'][i], res['states'][i])
                  for i in order]
        ns = min(args.cpu_sample, args.batch)
        jobs_128 = [(insts[i].H, insts[i].Ls, insts[i].rho0, tsave, rtol, atol, np.complex128, res['states'][i])
                    for i in range(ns)]
        jobs_64 = [(insts[i].H, insts[i].Ls, insts[i].rho0, tsave, rtol, atol, np.complex64, None)
                   for i in range(ns)]
        a_r = pool.map_async(_task_replay, jobs_r, chunksize=1)
        a_128 = pool.map_async(_task_adaptive, jobs_128, chunksize=1)
        a_64 = pool.map_async(_task_adaptive, jobs_64, chunksize=1)
        rr, r128, r64 = a_r.get(), a_128.get(), a_64.get()
        med128 = float(np.median([x[0] for x in r128]))
        med64 = float(np.median([x[0] for x in r64]))
        cores = cpu_workers(args.cpu_workers)
        out['error_vs_fp64'] = {
            'max_frobenius': max(x[0] for x in rr),
            'max_trace_distance': max(x[1] for x in rr),
            'frobenius_by_instance': [x[0] for x in rr],
            'eps2_by_instance': [float(eps2[i]) for i in order],
            'instances_checked': len(rr),
            'instances': order,
            'largest_eps2_instance': big,
            'reference': 'complex128 Tsit5 oracle on the identical accepted step sequence',
            'vs_adaptive_fp64_max_frobenius': max(x[2] for x in r128),
            'vs_adaptive_fp64_instances': ns,
            'note': 'the FP64 oracle with its own adaptive steps differs from a tight-tolerance solve by the '
            'same order at rtol=atol=1e-6 (stiff, stability-limited steps)',
        }
        out['cpu_baseline'] = {
            'value': cores / med64,
            'unit': 'solves/s',
            'cores': cores,
            'kind': 'port',
            'sample': f'{ns} of {args.batch} instances per precision, one instance per worker process on '
                      f'{cores} host cores (1 BLAS thread each, numpy {np.__version__}); value = cores / median '
                      f'seconds per instance of the complex64 restatement',
            'median_s_per_instance': {'c64': med64, 'c128': med128},
            'c128_solves_per_s': cores / med128,
            'runs': len(r64) + len(r128),
        }

    return out, finish


# ------------------------------------------------------------------ secondary


def secondary_c1():
    """Plain single solve of the lossy oscillator through dq.mesolve end to end."""
    import dynamiqs_amd as dq

    n = 16
    a = dq.destroy(n)
    H, Ls, psi = a.dag() @ a, [np.sqrt(0.1) * a], dq.coherent(n, 1.0)
    tsave = np.linspace(0.0, 2 * np.pi, 101)
    method = dq.method.Tsit5(rtol=1e-6, atol=1e-8)
    for _ in range(2):
        dq.mesolve(H, Ls, psi, tsave, method=method)
    lat = []
    for _ in range(9):
        t0 = time.perf_counter()
        res = dq.mesolve(H, Ls, psi, tsave, method=method)
        lat.append(time.perf_counter() - t0)
    return {'workload': 'lossy oscillator n=16, 101 tsave on [0, 2 pi], Tsit5 rtol 1e-6 atol 1e-8, one solve per '
                        'call through dq.mesolve (packing, launch, read-back)',
            'p50_ms': 1e3 * float(np.median(lat)), 'nsteps': int(res.infos.nsteps)}


def secondary_c3(T: float = 0.5, reps: int = 3):
    """DIA Rouchon1 on the 1024-dimensional two-mode batch of 16, over [0, T]."""
    import torch

    from dynamiqs_amd.integrators import dia

    H, Ls, a, b = two_mode_dia()
    rng = np.random.default_rng(1)
    B = 16
    rho0 = []
    for al in rng.uniform(0.5, 2.0, size=B):
        v = np.kron(coherent_ket(32, al), np.eye(32)[:, 0])
        rho0.append(np.outer(v, v.conj()))
    rho0 = np.stack(rho0)
    tsave = np.linspace(0.0, T, int(round(T / 0.1)) + 1)
    plan = dia.build_plan(H, Ls, rho0, tsave, 1e-3)
    stream = torch.cuda.current_stream()
    plan.reset()
    plan.run(stream)
    timer = EventTimer(stream)
    walls = run_plan_timed(plan, reps, stream, timer)
    kern_s = 1e-3 * float(np.median(timer.times_ms()))
    res = plan.results()
    tr = float(np.abs(np.trace(res['final_state'], axis1=-2, axis2=-1) - 1).max())
    step_s = kern_s / plan.nsteps
    out = {
        'workload': f'sparse-DIA Rouchon1, 32x32 two-mode n=1024, B={B}, dt=1e-3, T={T} ({plan.nsteps} steps)',
        'solves_per_s': B / float(np.median(walls)),
        'us_per_step': 1e6 * step_s,
        'sweep_algorithmic_GBps_over_step': plan.sweep_bytes_per_step / step_s / 1e9,
        'solve_TFLOPs_over_step': plan.solve_flops_per_step / step_s / 1e12,
        'max_trace_error': tr,
    }
    phases = getattr(plan, 'phase_us', None)
    if callable(phases):
        ph = phases(stream)
        out['phase_us_per_step'] = ph
        out['sweep_algorithmic_GBps'] = plan.sweep_bytes_per_step / (1e-6 * ph['sweep']) / 1e9
        out['solve_TFLOPs'] = plan.solve_flops_per_step / (1e-6 * ph['solve']) / 1e12
    return out


def secondary_c4(pool_jobs, nreq: int = 1000, rate: float = 700.0, nsample: int = 32, sweep=(300.0, 500.0)):
    """Serving: Poisson arrivals swept over ``sweep`` and ``rate``, then all at once
    (saturation); the requests sampled at ``rate`` log their accepted steps for the FP64
    replay."""
    import threading

    from dynamiqs_amd.serving import Server

    reqs = serving_requests(nreq)
    ns = np.array([r[0].shape[0] for r in reqs])
    sample = set(np.argsort(ns, kind='stable')[np.linspace(0, nreq - 1, nsample).astype(int)].tolist())
    rows = []
    sampled = {}
    edges = np.percentile(ns, [0, 25, 50, 75, 100])
    quart = [(ns >= edges[q]) & ((ns < edges[q + 1]) if q < 3 else (ns <= edges[q + 1])) for q in range(4)]
    for r in (*sweep, rate, 0.0):
        rng = np.random.default_rng(3)
        gaps = rng.exponential(1.0 / r, size=nreq) if r > 0 else np.zeros(nreq)
        with Server(memory_budget=32 << 30) as srv:
            futs = [None] * nreq

            def client():
                t_next = time.perf_counter()
                for i, (H, Ls, psi, ts) in enumerate(reqs):
                    t_next += gaps[i]
                    dt = t_next - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                    futs[i] = srv.submit(H, Ls, psi, ts, rtol=1e-5, step_log_cap=20_000 if i in sample else 0)

            t0 = time.perf_counter()
            th = threading.Thread(target=client)
            th.start()
            th.join()
            got = [f.result() for f in futs]
            wall = time.perf_counter() - t0
        st = dict(srv.stats)
        lat = np.array([g.latency_s for g in got])
        rows.append({
            'offered_rate': r if r > 0 else 'all-at-once',
            'solves_per_s': nreq / wall,
            'p50_latency_s': float(np.percentile(lat, 50)),
            'p99_latency_s': float(np.percentile(lat, 99)),
            'gpu_busy_fraction': st.get('busy_s', 0.0) / wall,
            'launches': st['launches'],
            'p99_latency_s_by_n_quartile': [float(np.percentile(lat[m], 99)) for m in quart],
            'scheduler_host_s': {k: round(v, 4) for k, v in srv.host_s.items()},
        })
        if r == rate:
            sampled = {i: got[i] for i in sample}
            loaded_lat = lat
    for i, g in sampled.items():
        H, Ls, psi, ts = reqs[i]
        pool_jobs.append((H, Ls, psi @ psi.conj().T, ts, g.step_ends, g.states))
    # every request solved alone (one at a time, nothing else on the device): the
    # latency floor the loaded latencies are compared with, per quartile of n
    iso = np.zeros(nreq)
    with Server(memory_budget=32 << 30) as srv:
        for i, (H, Ls, psi, ts) in enumerate(reqs):
            iso[i] = srv.submit(H, Ls, psi, ts, rtol=1e-5).result().latency_s
    buckets = []
    for q in range(4):
        lo, hi = edges[q], edges[q + 1]
        inb = quart[q]
        p50, p99 = float(np.percentile(loaded_lat[inb], 50)), float(np.percentile(loaded_lat[inb], 99))
        i50, i99 = float(np.percentile(iso[inb], 50)), float(np.percentile(iso[inb], 99))
        buckets.append({'n_range': [int(lo), int(hi)], 'p50_latency_s_at_rate': p50, 'p99_latency_s_at_rate': p99,
                        'isolated_p50_latency_s': i50, 'isolated_p99_latency_s': i99,
                        'p99_over_isolated_p99': p99 / i99, 'p99_over_isolated_p50': p99 / i50})
    return {'workload': f'serving mix, {nreq} requests, n log-uniform 8..256 (median {int(np.median(ns))}), K 1-3, '
                        f'T U[1,20], 50 tsave, rtol 1e-5', 'rates': rows,
            'latency_vs_isolated': {'rate': rate, 'by_n_quartile': buckets,
                                    'note': 'latencies of each quartile of n at the offered rate against the same '
                                            'requests solved one at a time (T spans 1-20, so the isolated p99 is '
                                            'the like-for-like floor of the loaded p99)'},
            'sampled_requests': sorted(int(i) for i in sample),
            'sampled_n': sorted(int(ns[i]) for i in sample)}


def secondary_c5(reps: int = 5):
    """Adjoint gradient of 64 PWC control vectors (n = 32, 100 segments) vs a
    forward solve of the same batch."""
    import torch

    from dynamiqs_amd.adjoint import AdjointPlan
    from dynamiqs_amd.integrators import dense

    n, nseg, B = 32, 100, 64
    a = np.diag(np.sqrt(np.arange(1, n, dtype=np.float64)), k=1).astype(np.complex128)
    H0, X = a.conj().T @ a, a + a.conj().T
    times = np.linspace(0.0, 10.0, nseg + 1)
    u = np.random.default_rng(4).normal(0.0, 0.5, size=(B, 1, nseg))
    Ls = [np.sqrt(0.05) * a]
    rho0 = np.zeros((n, n), dtype=np.complex128)
    rho0[0, 0] = 1
    insts = [dense.Instance(H=H0, Ls=Ls, rho0=rho0, pwc=[(times, u[b, 0], X)], key_L=(), key_rho=()) for b in range(B)]
    fwd = dense.build_plan('tsit5', insts, np.array([0.0, 10.0]), exp_ops=[H0])
    fwd.run()
    torch.cuda.synchronize()
    tf = []
    for _ in range(reps):
        fwd.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fwd.run()
        torch.cuda.synchronize()
        tf.append(time.perf_counter() - t0)
    plan = AdjointPlan(H0, [X], times, Ls, rho0, H0, B, checkpoint_every=10)
    plan.run(u)
    tg = []
    for k in range(reps):
        t0 = time.perf_counter()
        plan.run(u * (1.0 + 0.01 * k))
        tg.append(time.perf_counter() - t0)
    return {'workload': 'PWC-driven oscillator n=32, 100 segments on [0,10], B=64, loss Tr(a^dag a rho(T)), '
                        'checkpoints every 10 segments, packed AdjointPlan',
            'forward_s': float(np.median(tf)), 'gradient_s': float(np.median(tg)),
            'gradient_vs_forward': float(np.median(tg)) / float(np.median(tf))}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--config', default='kerr', choices=['kerr', 'c1', 'dia', 'serving', 'adjoint', 'copy', 'grid', 'sesolve'])
    ap.add_argument('--requests', type=int, default=1000)
    ap.add_argument('--rates', type=float, nargs='+', default=[100.0, 400.0, 0.0])
    ap.add_argument('--budget-gb', type=float, default=32.0)
    ap.add_argument('--dia-T', type=float, default=5.0)
    ap.add_argument('--dia-variant', default='auto', choices=['auto', 'global', 'tiled', 'split'])
    ap.add_argument('--dia-method', default='rouchon1', choices=['rouchon1', 'tsit5'])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--batch', type=int, default=256)
    ap.add_argument('--plain', type=int, default=2, help='one-solve-per-call solves timed per single-solve form')
    ap.add_argument('--cpu-sample', type=int, default=5, help='instances the CPU restatement times per precision (0: skip)')
    ap.add_argument('--replay', type=int, default=8, help='instances replayed by the FP64 oracle on the GPU steps')
    ap.add_argument('--cpu-workers', type=int, default=0, help='host worker processes (0: min(16, cores))')
    ap.add_argument('--no-secondary', action='store_true', help='only the headline (for profiling runs)')
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print(
            'bench.py measures one MI355X: batched mesolve instances are independent, so a multi-GPU run '
            'would shard instances across GPUs with no exchange; only --gpus 1 is measured here.',
            file=sys.stderr,
        )
        return 2

    if args.config == 'copy':
        return bench_copy(args)
    if args.config == 'c1':
        return bench_c1(args)
    if args.config == 'sesolve':
        return bench_sesolve(args)
    if args.config == 'grid':
        return bench_grid(args)
    if args.config == 'dia':
        return bench_dia(args)
    if args.config == 'serving':
        return bench_serving(args)
    if args.config == 'adjoint':
        return bench_adjoint(args)

    # worker processes for the CPU restatement, forked before the GPU is touched
    pool = make_cpu_pool(cpu_workers(args.cpu_workers)) if args.cpu_sample > 0 else None
    try:
        out, finish = bench_kerr(args, pool)
        serving_jobs = []
        if not args.no_secondary:
            sec = {}
            for name, fn in (('c1_plain_latency', secondary_c1), ('c3_dia_rouchon1', secondary_c3),
                             ('c4_serving', lambda: secondary_c4(serving_jobs)), ('c5_adjoint', secondary_c5)):
                t0 = time.perf_counter()
                sec[name] = fn()
                sec[name]['bench_wall_s'] = time.perf_counter() - t0
                print(f'# {name}: {json.dumps(sec[name])}', file=sys.stderr, flush=True)
            out['secondary'] = sec
        # CPU work after every GPU measurement: replays, baseline timings, serving errors
        finish(pool)
        if serving_jobs and pool is not None:
            errs = pool.map(_task_replay, serving_jobs, chunksize=1)
            out['secondary']['c4_serving']['error_vs_fp64'] = {
                'max_frobenius': max(e[0] for e in errs),
                'max_trace_distance': max(e[1] for e in errs),
                'requests_checked': len(errs),
                'reference': 'complex128 Tsit5 oracle on each sampled request\'s accepted steps',
            }
    finally:
        if pool is not None:
            pool.terminate()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
