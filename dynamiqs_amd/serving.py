"""Continuous-batching mesolve server on one MI355X.

Requests (one dense Tsit5 solve each, any Hilbert dimension n <= 256, any number
of jump operators, own save times and tolerances) are admitted while earlier
batches are still on the GPU: a scheduler thread packs whatever is pending into
one launch of the persistent Tsit5 kernel (every workgroup integrates one request
with its own adaptive step, so uneven requests share a launch), bounded by a
device-memory budget, and launches it on the next free HIP stream. Requests of
padded dimension above GRID_MIN_NP run on the grid-wide Tsit5 instead (each stage
spread over many CUs, driven by a worker thread), so that one large request does
not hold a single CU for the whole solve. Launches on
different streams overlap on the device; when a launch's completion event fires,
its requests are answered and their memory budget is released, which admits the
next pending requests. Each request's state, stages and outputs stay resident in
HBM for the whole solve.

The result of a request does not depend on what it was batched with: the kernel
keeps no cross-instance state and each workgroup's arithmetic is fixed by its own
descriptor (see ``tests/test_gpu_serving.py``).

There is no counterpart in the reference (``dq.mesolve`` is one synchronous call
per batch, ``dynamiqs/integrators/apis/mesolve.py:62-304``); the per-request solve
is the same Tsit5 path as :func:`dynamiqs_amd.mesolve`.
"""

from __future__ import annotations

import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .integrators.dense import MAX_NP, pad16


@dataclass
class Request:
    H: np.ndarray  # (n, n)
    Ls: list  # [(n, n)]
    rho0: np.ndarray  # (n, n) or (n, 1)
    tsave: np.ndarray
    rtol: float = 1e-6
    atol: float = 1e-6
    max_steps: int = 100_000
    exp_ops: list = field(default_factory=list)
    t_submit: float = 0.0
    future: Future | None = None

    @property
    def n(self) -> int:
        return self.H.shape[-1]


@dataclass
class Response:
    states: np.ndarray  # (ntsave, n, n) complex64
    expects: np.ndarray | None  # (nE, ntsave) complex64
    nsteps: int
    naccepted: int
    nrejected: int
    status: int
    latency_s: float
    queue_s: float


def _needs_work(npd: int, K: int) -> bool:
    # the register/LDS-resident kernel (np <= 64, K <= 3) keeps no per-instance scratch
    return npd > 64 or K > 3


def _layout(n: int, K: int, nt: int, nE: int):
    """Byte offsets of one request's input and output buffers (256-byte aligned)."""
    npd = pad16(n)
    nn = npd * npd
    lay_in, off = {}, 0
    for name, nbytes in (
        ('H', nn * 8),
        ('D', max(K, 1) * nn * 8),
        ('C', max(K, 1) * nn * 8),
        ('E', max(nE, 1) * nn * 8),
        ('y', 2 * nn * 16),
        ('w', _native.work_elems(npd, K, 0) * 8 if _needs_work(npd, K) else 16),
        ('ts', nt * 8),
    ):
        lay_in[name] = off
        off += (nbytes + 255) // 256 * 256
    size_in = off
    lay_out, off = {}, 0
    for name, nbytes in (('st', nt * n * n * 8), ('fin', n * n * 8), ('ex', max(nE, 1) * nt * 8), ('stats', 64)):
        lay_out[name] = off
        off += (nbytes + 255) // 256 * 256
    return lay_in, size_in, lay_out, off


def footprint_bytes(n: int, K: int, ntsave: int, nE: int = 0) -> int:
    """Device bytes one request holds while in flight."""
    _, a, _, b = _layout(n, K, ntsave, nE)
    return a + b


class _Launch:
    def __init__(self, reqs, stream, event, tensors, outs, nbytes, t_launch, ready=None):
        self.reqs = reqs
        self.stream = stream
        self.event = event
        self.tensors = tensors
        self.outs = outs
        self.nbytes = nbytes
        self.t_launch = t_launch
        self.ready = ready  # set by the worker thread of a grid-wide launch
        self.error = None

    def done(self) -> bool:
        if self.ready is not None and not self.ready.is_set():
            return False
        return self.event is not None and self.event.query()


# Requests with a padded dimension above this run on the grid-wide Tsit5 (every
# stage spread over many CUs) instead of one workgroup per request.
GRID_MIN_NP = 96


def pack_and_launch(reqs: list, stream, grid: bool = False):
    """Pack uneven requests into one descriptor array and launch the persistent
    Tsit5 kernel on ``stream`` (``grid``: return a callable that runs the
    grid-wide Tsit5 on them instead; requests of one grid launch share their padded
    dimension). Returns (tensors kept alive, per-request output layout, completion
    event or the callable, device bytes held)."""
    import torch

    dev = torch.device('cuda', torch.cuda.current_device())
    B = len(reqs)
    descs = np.zeros(B, dtype=_native.ME_INST_DTYPE)
    lays = [_layout(r.n, len(r.Ls), len(r.tsave), len(r.exp_ops)) for r in reqs]
    in_off = np.cumsum([0] + [lay[1] for lay in lays])
    out_off = np.cumsum([0] + [lay[3] for lay in lays])
    with torch.cuda.stream(stream):
        arena = torch.zeros(int(in_off[-1]), dtype=torch.uint8, device=dev)
        oarena = torch.zeros(int(out_off[-1]), dtype=torch.uint8, device=dev)
        base, obase = arena.data_ptr(), oarena.data_ptr()

        def cview(a, npd):
            return arena[a : a + npd * npd * 8].view(torch.complex64).view(npd, npd)

        outs = []
        for b, r in enumerate(reqs):
            n, K, nt, nE = r.n, len(r.Ls), len(r.tsave), len(r.exp_ops)
            npd = pad16(n)
            nn = npd * npd
            li, _, lo, _ = lays[b]
            ib, ob = int(in_off[b]), int(out_off[b])
            H = torch.from_numpy(np.ascontiguousarray(r.H, dtype=np.complex128)).to(dev)
            Ls = [torch.from_numpy(np.ascontiguousarray(L, dtype=np.complex128)).to(dev) for L in r.Ls]
            Hnh = -1j * H
            for L in Ls:
                Hnh = Hnh - 0.5 * (L.conj().T @ L)
            cview(ib + li['H'], npd)[:n, :n] = Hnh.to(torch.complex64)
            for k, L in enumerate(Ls):
                cview(ib + li['D'] + k * nn * 8, npd)[:n, :n] = L.to(torch.complex64)
                cview(ib + li['C'] + k * nn * 8, npd)[:n, :n] = (0.5 * L.conj().T).to(torch.complex64)
            for q, E in enumerate(r.exp_ops):
                ET = torch.from_numpy(np.ascontiguousarray(np.asarray(E).T, dtype=np.complex64)).to(dev)
                cview(ib + li['E'] + q * nn * 8, npd)[:n, :n] = ET
            rho = np.asarray(r.rho0, dtype=np.complex128)
            if rho.shape[-1] == 1:
                rho = rho @ rho.conj().T
            a = ib + li['y']
            arena[a : a + nn * 16].view(torch.complex128).view(npd, npd)[:n, :n] = torch.from_numpy(rho).to(dev)
            a = ib + li['ts']
            arena[a : a + nt * 8].view(torch.float64).copy_(torch.from_numpy(np.asarray(r.tsave, dtype=np.float64)))
            d = descs[b]
            d['n'], d['np'], d['nL'], d['nE'], d['ntsave'] = n, npd, K, nE, nt
            d['flags'] = _native.FLAG_SAVE_STATES | _native.FLAG_HAS_HNH
            d['rtol'], d['atol'], d['safety'], d['factormin'], d['factormax'] = r.rtol, r.atol, 0.9, 0.2, 5.0
            d['max_steps'] = r.max_steps
            d['t0'], d['t1'] = float(r.tsave[0]), float(r.tsave[-1])
            d['tsave'] = base + ib + li['ts']
            d['Hnh'], d['Dops'] = base + ib + li['H'], base + ib + li['D']
            d['Cops'], d['EopsT'] = base + ib + li['C'], base + ib + li['E']
            d['y'], d['work'] = base + ib + li['y'], base + ib + li['w']
            d['states'], d['final_state'] = obase + ob + lo['st'], obase + ob + lo['fin']
            d['expects'], d['stats'] = obase + ob + lo['ex'], obase + ob + lo['stats']
            outs.append((ob + lo['st'], ob + lo['ex'], ob + lo['stats'], n, nt, nE))
        d_descs = torch.from_numpy(descs.view(np.uint8).copy()).to(dev)
        lib = _native.lib()
        c_p = _native.ctypes.c_void_p
        if grid:
            npd = int(descs['np'][0])
            if np.any(descs['np'] != npd):
                raise ValueError('a grid-wide launch needs requests of one padded dimension')
            nb = int(lib.dqh_tsit5_grid_scratch_bytes(npd, B))
            scratch = torch.empty((nb,), dtype=torch.uint8, device=dev)
            flags = torch.zeros((2,), dtype=torch.int64, pin_memory=True)

            def run():
                # blocks this (worker) thread until the batch is done; ctypes releases the GIL
                with torch.cuda.stream(stream):
                    _native.check(
                        lib.dqh_mesolve_tsit5_dense_grid(
                            c_p(d_descs.data_ptr()), descs.ctypes.data_as(c_p), B, c_p(scratch.data_ptr()), nb, 16,
                            c_p(flags.data_ptr()), c_p(stream.cuda_stream),
                        ),
                        'dqh_mesolve_tsit5_dense_grid',
                    )
                    ev = torch.cuda.Event()
                    ev.record(stream)
                return ev

            return [oarena, arena, d_descs, scratch, flags, descs], outs, run, int(in_off[-1] + out_off[-1] + nb)
        _native.check(
            lib.dqh_mesolve_tsit5_dense(c_p(d_descs.data_ptr()), B, c_p(stream.cuda_stream)),
            'dqh_mesolve_tsit5_dense',
        )
        ev = torch.cuda.Event()
        ev.record(stream)
    return [oarena, arena, d_descs], outs, ev, int(in_off[-1] + out_off[-1])


def collect(oarena, outs) -> list:
    """Copy the outputs of a finished launch to host numpy arrays."""
    host = oarena.cpu().numpy()
    res = []
    for a_st, a_ex, a_stats, n, nt, nE in outs:
        st = host[a_st : a_st + nt * n * n * 8].view(np.complex64).reshape(nt, n, n).copy()
        ex = host[a_ex : a_ex + nE * nt * 8].view(np.complex64).reshape(nE, nt).copy() if nE else None
        stats = host[a_stats : a_stats + 64].view(np.int64)
        res.append((st, ex, int(stats[0]), int(stats[1]), int(stats[2]), int(stats[3])))
    return res


def _size_class(r) -> int:
    """Launch grouping: small requests by size band (a persistent launch lasts as
    long as its slowest request), grid-wide requests by exact padded dimension."""
    npd = pad16(r.n)
    if npd > GRID_MIN_NP:
        return 1000 + npd
    return 0 if npd <= 32 else (1 if npd <= 64 else 2)


class Server:
    """Continuous-batching request loop (see module docstring).

    Args:
        memory_budget: device bytes that in-flight requests may hold together.
        max_batch: maximum requests per launch.
        nstreams: HIP streams launches rotate over (default priority).
        max_inflight: maximum concurrent launches.
    """

    def __init__(self, memory_budget: int = 8 << 30, max_batch: int = 64, nstreams: int = 4, max_inflight: int = 8):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError('dynamiqs_amd.serving needs an MI355X GPU')
        _native.lib()
        self.memory_budget = int(memory_budget)
        self.max_batch = int(max_batch)
        self.max_inflight = int(max_inflight)
        self.streams = [torch.cuda.Stream() for _ in range(nstreams)]
        self._q: queue.Queue = queue.Queue()
        self._pending: list = []
        self._inflight: list = []
        self._used = 0
        self._stop = threading.Event()
        self._next_stream = 0
        self.stats = {'launches': 0, 'requests': 0, 'max_inflight_bytes': 0, 'max_concurrent_launches': 0}
        self._thread = threading.Thread(target=self._loop, name='dqh-server', daemon=True)
        self._thread.start()

    # ---------------------------------------------------------------- client
    def submit(self, H, Ls, rho0, tsave, *, rtol=1e-6, atol=1e-6, max_steps=100_000, exp_ops=None) -> Future:
        H = np.asarray(H)
        n = H.shape[-1]
        if H.shape != (n, n) or pad16(n) > MAX_NP:
            raise ValueError(f'serving accepts dense (n, n) Hamiltonians with n <= {MAX_NP}, got {H.shape}')
        tsave = np.asarray(tsave, dtype=np.float64)
        if tsave.ndim != 1 or len(tsave) == 0 or np.any(np.diff(tsave) < 0):
            raise ValueError('Argument tsave must be sorted in strictly ascending order.')
        req = Request(H=H, Ls=[np.asarray(L) for L in Ls], rho0=np.asarray(rho0), tsave=tsave, rtol=rtol, atol=atol,
                      max_steps=max_steps, exp_ops=[np.asarray(E) for E in (exp_ops or [])])
        if footprint_bytes(n, len(req.Ls), len(tsave), len(req.exp_ops)) > self.memory_budget:
            raise ValueError('request does not fit in the server memory budget')
        req.future = Future()
        req.t_submit = time.perf_counter()
        self._q.put(req)
        return req.future

    def close(self) -> None:
        self._stop.set()
        self._thread.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------- scheduler
    def _admit(self) -> None:
        if not self._pending or len(self._inflight) >= self.max_inflight:
            return
        # a launch lasts as long as its slowest request: fill it with requests of
        # the same size class as the oldest pending one (FIFO inside the class)
        cls = _size_class(self._pending[0])
        batch, nbytes, keep = [], 0, []
        for r in self._pending:
            if len(batch) >= self.max_batch or _size_class(r) != cls:
                keep.append(r)
                continue
            fb = footprint_bytes(r.n, len(r.Ls), len(r.tsave), len(r.exp_ops))
            if self._used + nbytes + fb > self.memory_budget:
                keep.append(r)
                continue
            batch.append(r)
            nbytes += fb
        self._pending = keep
        if not batch:
            return
        stream = self.streams[self._next_stream % len(self.streams)]
        self._next_stream += 1
        t0 = time.perf_counter()
        grid = cls >= 1000
        try:
            tensors, outs, ev, total = pack_and_launch(batch, stream, grid=grid)
        except Exception as e:  # answer the batch with the error, keep serving
            for r in batch:
                r.future.set_exception(e)
            return
        self._used += total
        if grid:
            launch = _Launch(batch, stream, None, tensors, outs, total, t0, ready=threading.Event())

            def work(launch=launch, run=ev):
                try:
                    launch.event = run()
                except Exception as e:  # reported when the launch retires
                    launch.error = e
                finally:
                    launch.ready.set()

            threading.Thread(target=work, name='dqh-grid', daemon=True).start()
            self._inflight.append(launch)
        else:
            self._inflight.append(_Launch(batch, stream, ev, tensors, outs, total, t0))
        self.stats['launches'] += 1
        self.stats['max_inflight_bytes'] = max(self.stats['max_inflight_bytes'], self._used)
        self.stats['max_concurrent_launches'] = max(self.stats['max_concurrent_launches'], len(self._inflight))

    def _retire(self) -> bool:
        done = [L for L in self._inflight if (L.ready is not None and L.ready.is_set() and L.error) or L.done()]
        for L in done:
            self._inflight.remove(L)
            try:
                if L.error is not None:
                    raise L.error
                res = collect(L.tensors[0], L.outs)
            except Exception as e:
                for r in L.reqs:
                    r.future.set_exception(e)
                self._used -= L.nbytes
                continue
            t = time.perf_counter()
            for r, (st, ex, ns, na, nr, status) in zip(L.reqs, res):
                self.stats['requests'] += 1
                if status == -3:
                    r.future.set_exception(RuntimeError('The maximum number of method steps has been reached.'))
                    continue
                r.future.set_result(Response(st, ex, ns, na, nr, status, t - r.t_submit, L.t_launch - r.t_submit))
            self._used -= L.nbytes
        return bool(done)

    def _loop(self) -> None:
        while not (self._stop.is_set() and self._q.empty() and not self._pending and not self._inflight):
            got = False
            try:
                while True:
                    self._pending.append(self._q.get_nowait())
                    got = True
            except queue.Empty:
                pass
            retired = self._retire()
            self._admit()
            if not (got or retired):
                time.sleep(0.0005)
