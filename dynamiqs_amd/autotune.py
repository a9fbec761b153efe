"""Autotuner for the kernel forms of the MI355X solvers.

The sweep kernels have several admissible forms for one problem, and which is
fastest depends on the problem's shape, so they are timed once per key and the
fastest is kept:

* DIA sweeps, keyed on (sweep kind, n, number of operators, total DIA diagonal
  count, largest |offset|): the global form (one thread per element, neighbours
  through the caches) and the LDS-tiled forms with halo 8, 16 and 32 that the
  offsets admit (64 x 64 output tiles; a smaller halo stages less and fits more
  workgroups per CU). For the Lindblad vector field (DIA Tsit5) a candidate is
  timed on ``dqh_lindblad_rhs_dia`` of the actual operators over a synthetic
  batch; for the Rouchon1 Kraus sweep on three steps of
  ``dqh_mesolve_rouchon1_dia``, where the two-pass split form is a candidate too.
* Dense kernels, keyed on (padded n, number of jump operators) where the
  LDS-resident form applies (n <= 64, at most 3 jump operators): resident against
  global, timed on ``dqh_lindblad_rhs_dense`` over a batch of random instances.

Choices live in memory and in a JSON file (``DQH_AUTOTUNE_CACHE``, default
``~/.cache/dynamiqs_amd/autotune.json``), so a key is timed once per machine.
``DQH_AUTOTUNE=0`` turns tuning off (the fixed rule of the kernels applies: the
smallest admissible halo, the resident form where it applies).
"""

from __future__ import annotations

import json
import os
import threading

import numpy as np

_lock = threading.Lock()
_memo: dict[str, dict] = {}
_loaded = False


def enabled() -> bool:
    return os.environ.get('DQH_AUTOTUNE', '1') != '0'


def cache_path() -> str:
    return os.environ.get('DQH_AUTOTUNE_CACHE', os.path.join(os.path.expanduser('~'), '.cache', 'dynamiqs_amd', 'autotune.json'))


def _load() -> None:
    global _loaded
    if _loaded:
        return
    _loaded = True
    try:
        with open(cache_path()) as f:
            data = json.load(f)
        if isinstance(data, dict):
            _memo.update({k: v for k, v in data.items() if isinstance(v, dict)})
    except (OSError, ValueError):
        pass


def _store() -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = f'{path}.{os.getpid()}.tmp'
        with open(tmp, 'w') as f:
            json.dump(_memo, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs a re-tune in the next process


def _device_key() -> str:
    import torch

    return torch.cuda.get_device_name(torch.cuda.current_device()).replace(' ', '_')


def _time_us(run, reps: int = 5) -> float:
    """Median kernel time of ``run(stream)`` over ``reps`` launches (HIP events on
    the launch stream), after one warm-up launch."""
    import torch

    s = torch.cuda.current_stream()
    run(s)
    ts = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(s)
        run(s)
        b.record(s)
        b.synchronize()
        ts.append(1e3 * a.elapsed_time(b))
    return float(np.median(ts))


def lookup(key: str):
    with _lock:
        _load()
        return _memo.get(key)


def record(key: str, entry: dict) -> None:
    with _lock:
        _load()
        _memo[key] = entry
        _store()


# --------------------------------------------------------------------- DIA
def dia_key(n_pad: int, ops) -> str:
    """``ops``: [(offsets, diags)] as packed for the kernels (Hnh first)."""
    ndiag = sum(len(o) for o, _ in ops)
    maxoff = max((abs(int(x)) for o, _ in ops for x in o), default=0)
    return f'dia|{_device_key()}|n={n_pad}|nops={len(ops)}|ndiag={ndiag}|maxoff={maxoff}'


def dia_candidates(n_pad: int, maxoff: int) -> list[int]:
    from .integrators import dia

    out = [dia.SWEEP_GLOBAL]
    if n_pad % dia.NB == 0:
        out += [v for v, h in dia.SWEEP_HALO.items() if maxoff <= h]
    return out


def dia_sweep(H: np.ndarray, Ls, batch: int, kind: str = 'rhs', dt: float = 1e-3) -> int:
    """Fastest sweep form (a ``SWEEP_*`` value) for the DIA operators of (H, Ls):
    ``kind`` 'rhs' times the Lindblad vector field (DIA Tsit5), 'kraus' times three
    Rouchon1 steps of length ``dt``."""
    from .integrators import dia

    if not enabled():
        return dia.SWEEP_AUTO
    n = H.shape[-1]
    n_pad = dia.NB * ((n + dia.NB - 1) // dia.NB)
    Hnh = -1j * H - 0.5 * sum((L.conj().T @ L for L in Ls), np.zeros_like(H))
    ops = [dia._dia_of(Hnh)] + [dia._dia_of(L) for L in Ls]
    key = f'{kind}|' + dia_key(n_pad, ops)
    hit = lookup(key)
    if hit is not None:
        return int(hit['variant'])
    maxoff = max((abs(int(x)) for o, _ in ops for x in o), default=0)
    cands = dia_candidates(n_pad, maxoff)
    if len(cands) == 1:
        record(key, {'variant': cands[0], 'us': {}})
        return cands[0]
    rng = np.random.default_rng(0)
    B = max(1, min(int(batch), 4))
    times = {}
    if kind == 'kraus':
        x = rng.normal(size=(B, n, 1)) + 1j * rng.normal(size=(B, n, 1))
        rho = x @ np.conj(np.swapaxes(x, -1, -2))
        rho /= np.trace(rho, axis1=-2, axis2=-1)[:, None, None]
        for v in cands:
            plan = dia.build_plan(H, Ls, rho, np.array([0.0, 3 * dt]), dt, variant=v)

            def run(s, plan=plan):
                plan.reset()
                plan.run(s)

            times[v] = _time_us(run, reps=3)
    else:
        rho = rng.normal(size=(B, n, n)) + 1j * rng.normal(size=(B, n, n))
        for v in cands:
            plan = dia.rhs_plan(H, Ls, rho, variant=v)
            times[v] = _time_us(plan.run)
    best = min(times, key=times.get)
    record(key, {'variant': int(best), 'us': {str(k): t for k, t in times.items()}, 'batch': B})
    return int(best)


# ------------------------------------------------------------------- dense
def dense_key(np_: int, K: int) -> str:
    return f'dense|{_device_key()}|np={np_}|K={K}'


def dense_global(np_: int, K: int, batch: int) -> bool:
    """True when the global-memory dense form beats the LDS-resident one for
    (np, K); False where only one form applies or tuning is off."""
    from .integrators import dense

    if not enabled() or np_ not in (16, 32, 48, 64) or K > 3:
        return False
    key = dense_key(np_, K)
    hit = lookup(key)
    if hit is not None:
        return bool(hit['global'])
    rng = np.random.default_rng(0)
    B = max(1, min(int(batch), 256))
    insts = []
    for _ in range(B):
        H = rng.normal(size=(np_, np_)) + 1j * rng.normal(size=(np_, np_))
        Ls = [0.1 * (rng.normal(size=(np_, np_)) + 1j * rng.normal(size=(np_, np_))) for _ in range(K)]
        insts.append(dense.Instance(H=H + H.conj().T, Ls=Ls, rho0=np.eye(np_) / np_))
    times = {}
    for g in (False, True):
        plan = dense.build_plan('rhs', insts, np.array([0.0, 1.0]), form='global' if g else 'resident')
        times[g] = _time_us(plan.run)
    best = times[True] < times[False]
    record(key, {'global': bool(best), 'us': {'resident': times[False], 'global': times[True]}, 'batch': B})
    return bool(best)
