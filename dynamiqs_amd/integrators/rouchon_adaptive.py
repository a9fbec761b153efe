"""Adaptive Rouchon2 / Rouchon3 on the MI355X (``dqh_mesolve_rouchon_adaptive_dense``).

The reference runs ``Rouchon2()`` / ``Rouchon3()`` without ``dt`` through
``MESolveAdaptiveRouchonIntegrator`` (``dynamiqs/integrators/core/
rouchon_integrator.py:455-549, 585-651``): per trial step the no-jump propagators
of an embedded Runge-Kutta pair on dU/dt = G U, a high- and a low-order Kraus map
of the same rho (each Cholesky-normalised), and diffrax's PID controller on
``(rho_high - rho_low) / 2`` with ``step_ts = tsave``. This module packs a batch of
dense instances (n <= 64, constant jump operators; constant, piecewise-constant and
modulated Hamiltonian terms) for the FP64 kernel, one workgroup per instance.
"""

from __future__ import annotations

import ctypes

import numpy as np

from .. import _native
from .dense import Instance, _Arena, _dag, _device, _pad, pad16
from .operators import spline_table

AR_INST_FIELDS = [
    ('n', np.int64),
    ('np', np.int64),
    ('nL', np.int64),
    ('nE', np.int64),
    ('ntsave', np.int64),
    ('order', np.int64),
    ('normalize', np.int64),
    ('flags', np.int64),
    ('max_steps', np.int64),
    ('npwc', np.int64),
    ('nseg', np.int64),
    ('nmod', np.int64),
    ('mod_nint', np.int64),
    ('rtol', np.float64),
    ('atol', np.float64),
    ('safety', np.float64),
    ('factormin', np.float64),
    ('factormax', np.float64),
    ('t0', np.float64),
    ('t1', np.float64),
    ('dt0', np.float64),
    ('mod_t0', np.float64),
    ('mod_dt', np.float64),
    ('ops', np.uint64),
    ('pwc_times', np.uint64),
    ('pwc_vals', np.uint64),
    ('mod_coef', np.uint64),
    ('EopsT', np.uint64),
    ('tsave', np.uint64),
    ('rho', np.uint64),
    ('work', np.uint64),
    ('states', np.uint64),
    ('expects', np.uint64),
    ('stats', np.uint64),
]
AR_INST_DTYPE = np.dtype(AR_INST_FIELDS)
ARENA_FIELDS = ('ops', 'pwc_times', 'pwc_vals', 'mod_coef', 'EopsT', 'tsave')

MAX_N = 64
# the reference passes no dt0 for adaptive methods; diffrax's initial-step selection
# returns 0.01 for a term that is not an ODE term, such as the Rouchon term
DT0 = 0.01


def _check(descs: np.ndarray, tensors, work_elems: int) -> None:
    """Host-side guard before the launch: every buffer a descriptor names lies
    inside one device allocation of the plan, with the extent the kernel reads or
    writes."""
    spans = sorted((t.data_ptr(), t.data_ptr() + t.numel() * t.element_size()) for t in tensors)
    for d in descs:
        npd, n, nt, nE = int(d['np']), int(d['n']), int(d['ntsave']), int(d['nE'])
        nn16 = npd * npd * 16
        nops = 2 + int(d['nL']) + int(d['npwc']) + int(d['nmod'])
        need = {
            'ops': nops * nn16,
            'pwc_times': int(d['nseg']) * 8,
            'pwc_vals': int(d['npwc']) * max(int(d['nseg']) - 1, 0) * 16,
            'mod_coef': int(d['nmod']) * int(d['mod_nint']) * 64,
            'EopsT': nE * nn16,
            'tsave': nt * 8,
            'rho': nn16,
            'work': work_elems * 16,
            'states': nt * n * n * 16 if int(d['flags']) & _native.FLAG_SAVE_STATES else 0,
            'expects': nE * nt * 16,
            'stats': 64,
        }
        for f, nbytes in need.items():
            p = int(d[f])
            if nbytes == 0:
                continue
            if p == 0 or not any(lo <= p and p + nbytes <= hi for lo, hi in spans):
                raise AssertionError(f'adaptive Rouchon descriptor field {f} does not cover {nbytes} bytes')


class ARPlan:
    """A packed batch of adaptive Rouchon solves."""

    def __init__(self, n: int, ntsave: int, nE: int, ninst: int, save_states: bool):
        self.n, self.ntsave, self.nE, self.ninst, self.save_states = n, ntsave, nE, ninst, save_states

    def run(self, stream=None):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        lib = _native.lib()
        _native.check(
            lib.dqh_mesolve_rouchon_adaptive_dense(
                ctypes.c_void_p(self.d_descs.data_ptr()), self.ninst, ctypes.c_void_p(s.cuda_stream)
            ),
            'dqh_mesolve_rouchon_adaptive_dense',
        )

    def results(self) -> dict:
        import torch

        torch.cuda.synchronize()
        B, n, nt = self.ninst, self.n, self.ntsave
        stats = self.stats.cpu().numpy()
        rho = self.rho.cpu().numpy().view(np.complex128).reshape(B, self.np, self.np)[:, :n, :n]
        out = {
            'final_state': rho,
            'nsteps': stats[:, 0],
            'naccepted': stats[:, 1],
            'nrejected': stats[:, 2],
            'status': stats[:, 3],
        }
        if self.save_states:
            out['states'] = self.states.cpu().numpy().view(np.complex128).reshape(B, nt, n, n)
        else:
            out['states'] = rho.reshape(B, 1, n, n).copy()
        if self.nE:
            out['expects'] = self.expects.cpu().numpy().view(np.complex128).reshape(B, self.nE, nt)
        return out


def build_plan(instances: list, tsave, *, order: int, exp_ops=None, t0: float | None = None, rtol: float = 1e-6,
               atol: float = 1e-6, safety: float = 0.9, factormin: float = 0.2, factormax: float = 5.0,
               max_steps: int = 100_000, normalize: bool = True, save_states: bool = True,
               dt0: float = DT0) -> ARPlan:
    """Pack ``instances`` (:class:`dense.Instance`, same Hilbert dimension) for the
    adaptive Rouchon kernel of ``order`` 2 (``Rouchon2``) or 3 (``Rouchon3``)."""
    import torch

    if order not in (2, 3):
        raise ValueError(f'adaptive Rouchon order must be 2 or 3, got {order}')
    n = instances[0].H.shape[-1]
    if any(inst.H.shape[-1] != n for inst in instances):
        raise ValueError('all instances of an adaptive Rouchon plan must share the Hilbert dimension')
    if n > MAX_N:
        raise NotImplementedError(
            f'the adaptive Rouchon methods run on the MI355X for n <= {MAX_N} (got n={n}); set `dt` for the'
            ' fixed-step method'
        )
    for inst in instances:
        if any(lm is not None for lm in list(inst.Lmod or []) + list(inst.Lpwc or []) + list(inst.Lsum or [])):
            raise NotImplementedError('the adaptive Rouchon methods take constant jump operators only')
        if inst.Hnh is not None or inst.tsave is not None:
            raise NotImplementedError('explicit generators and per-instance save times are not supported here')
        if len(inst.mod) > _native.MAX_MOD or len(inst.pwc) > _native.MAX_MOD:
            raise NotImplementedError(f'at most {_native.MAX_MOD} modulated and {_native.MAX_MOD} PWC terms')
    dev = _device()
    lib = _native.lib()
    if lib.dqh_sizeof_ar_inst() != AR_INST_DTYPE.itemsize:
        raise RuntimeError('dqh_ar_inst layout mismatch between include/dqh.h and the host mirror')
    npd = pad16(n)
    nn = npd * npd
    tsave = np.asarray(tsave, dtype=np.float64)
    nt = len(tsave)
    B = len(instances)
    Es = [np.asarray(E, dtype=np.complex128) for E in (exp_ops or [])]
    nE = len(Es)
    t_start = float(tsave[0] if t0 is None else t0)
    t_stop = float(tsave[-1])
    plan = ARPlan(n, nt, nE, B, save_states)
    plan.np = npd
    arena = _Arena(dev)
    descs = np.zeros(B, dtype=AR_INST_DTYPE)
    descs['n'], descs['np'], descs['nE'], descs['ntsave'] = n, npd, nE, nt
    descs['order'], descs['normalize'] = order, int(bool(normalize))
    descs['flags'] = _native.FLAG_SAVE_STATES if save_states else 0
    descs['max_steps'] = max_steps
    descs['rtol'], descs['atol'], descs['safety'] = rtol, atol, safety
    descs['factormin'], descs['factormax'] = factormin, factormax
    descs['t0'], descs['t1'], descs['dt0'] = t_start, t_stop, dt0
    descs['tsave'] = arena.put(tsave)
    descs['EopsT'] = arena.put(_pad(np.stack([E.T for E in Es]), npd)) if nE else 0
    tables = {}
    for b, inst in enumerate(instances):
        d = descs[b]
        Ls = [np.asarray(L, dtype=np.complex128) for L in inst.Ls]
        LdL = sum((_dag(L) @ L for L in Ls), np.zeros((n, n), dtype=np.complex128))
        G0 = -1j * np.asarray(inst.H, dtype=np.complex128) - 0.5 * LdL
        mats = [G0, LdL, *Ls]
        mats += [-1j * np.asarray(o, dtype=np.complex128) for (_, _, o) in inst.pwc]
        mats += [-1j * np.asarray(o, dtype=np.complex128) for (_, o, _) in inst.mod]
        d['nL'] = len(Ls)
        d['ops'] = arena.put(_pad(np.stack(mats), npd))
        if inst.pwc:
            grid = np.unique(np.concatenate([np.asarray(t, dtype=np.float64) for (t, _, _) in inst.pwc]))
            left = grid[:-1]
            vals = []
            for t, v, _ in inst.pwc:
                t = np.asarray(t, dtype=np.float64)
                idx = np.clip(np.searchsorted(t, left, side='right') - 1, 0, len(v) - 1)
                inside = (left >= t[0]) & (left < t[-1])
                vals.append(np.where(inside, np.asarray(v, dtype=np.complex128)[idx], 0.0))
            d['npwc'], d['nseg'] = len(inst.pwc), len(grid)
            d['pwc_times'] = arena.put(grid)
            d['pwc_vals'] = arena.put(np.ascontiguousarray(np.stack(vals)))
        if inst.mod:
            tabs = []
            for g, _, key in inst.mod:
                k = (key, t_start, t_stop) if key is not None else None
                if k is None or k not in tables:
                    tab = spline_table(g, t_start, t_stop)
                    if k is not None:
                        tables[k] = tab
                else:
                    tab = tables[k]
                tabs.append((g, tab))
            nint = max(len(tab[2]) for _, tab in tabs)
            tabs = [tab if len(tab[2]) == nint else spline_table(g, t_start, t_stop, n0=nint, nmax=nint) for g, tab in tabs]
            d['nmod'] = len(inst.mod)
            d['mod_t0'], d['mod_dt'], d['mod_nint'] = tabs[0][0], tabs[0][1], nint
            d['mod_coef'] = arena.put(np.ascontiguousarray(np.stack([c for (_, _, c) in tabs])))
    arena.upload()
    for f in ARENA_FIELDS:
        descs[f] = [arena.patch(int(v)) for v in descs[f]]

    welems = int(lib.dqh_ar_work_elems(npd))
    work = arena.empty(B * welems * 16)
    rho0 = np.zeros((B, npd, npd), dtype=np.complex128)
    for b, inst in enumerate(instances):
        rho0[b, :n, :n] = inst.rho0
    rho = torch.from_numpy(rho0.view(np.float64)).to(dev)
    states = torch.zeros((B, (nt if save_states else 1) * n * n * 2), dtype=torch.float64, device=dev)
    expects = torch.zeros((B, max(nE, 1) * nt * 2), dtype=torch.float64, device=dev)
    stats = torch.zeros((B, 8), dtype=torch.int64, device=dev)
    for b in range(B):
        d = descs[b]
        d['rho'] = rho.data_ptr() + b * nn * 16
        d['work'] = work.data_ptr() + b * welems * 16
        d['states'] = states.data_ptr() + b * states.shape[1] * 8
        d['expects'] = expects.data_ptr() + b * expects.shape[1] * 8
        d['stats'] = stats.data_ptr() + b * 64
    _check(descs, [*arena.tensors, rho, states, expects, stats], welems)
    plan.rho, plan.states, plan.expects, plan.stats = rho, states, expects, stats
    plan.arena, plan.work, plan.descs = arena, work, descs
    plan.d_descs = torch.from_numpy(descs.view(np.uint8).copy()).to(dev)
    return plan


def solve(instances: list, tsave, **kw) -> dict:
    """Pack, run and read back one batch (see :func:`build_plan`)."""
    plan = build_plan(instances, tsave, **kw)
    plan.run()
    return plan.results()


__all__ = ['ARPlan', 'build_plan', 'solve', 'Instance', 'AR_INST_DTYPE']
