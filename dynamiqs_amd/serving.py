"""Continuous-batching mesolve server on one MI355X.

Requests (one dense Tsit5 solve each, any Hilbert dimension n <= 256, any number
of jump operators, own save times and tolerances) are admitted while earlier
batches are still on the GPU: a scheduler thread packs whatever is pending into
one launch of the persistent Tsit5 kernel (every workgroup integrates one request
with its own adaptive step, so uneven requests share a launch), bounded by a
device-memory budget, and launches it on the next free HIP stream. Requests of
padded dimension above GRID_MIN_NP run on the grid-wide Tsit5 instead (each stage
spread over many CUs, driven by a worker thread), so that one large request does
not hold a single CU for the whole solve. Launches on different streams overlap on
the device.

Packing: the inputs of a launch (-iH, jump operators, initial states, save
times) and its descriptor array are written on the host into one pinned staging
buffer and cross to the device in one copy; a small kernel then completes the
non-Hermitian Hamiltonian -iH - 1/2 sum L^dag L and the right factors L^dag / 2
on the device (FP64 products), so no O(n^3) work is left on the host. The FP64
state, the stage scratch and every input stay resident in HBM for the whole solve.

Retirement is per request: the outputs of a persistent launch (saved states,
expectation values, final state, statistics) live in pinned host memory that the
kernel writes directly, and each workgroup publishes its request's completion in
stats[7] with a system-scope release (``DQH_FLAG_SIGNAL``). The scheduler polls
those words and answers a request as soon as its own solve is done, whatever the
other requests of its launch are still doing; the launch's device memory returns
to the budget when its last request is done.

The result of a request does not depend on what it was batched with: the kernel
keeps no cross-instance state and each workgroup's arithmetic is fixed by its own
descriptor (see ``tests/test_gpu_serving.py``).

There is no counterpart in the reference (``dq.mesolve`` is one synchronous call
per batch, ``dynamiqs/integrators/apis/mesolve.py:62-304``); the per-request solve
is the same Tsit5 path as :func:`dynamiqs_amd.mesolve`.
"""

from __future__ import annotations

import os
import queue
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor
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
    step_log_cap: int = 0  # > 0: record the end time of each accepted step (the first step_log_cap)
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
    step_ends: np.ndarray | None = None  # end times of the accepted steps (when requested)


def _needs_work(npd: int, K: int) -> bool:
    # the register/LDS-resident kernel (np <= 64, K <= 3) keeps no per-instance scratch
    return npd > 64 or K > 3


def _align(nbytes: int) -> int:
    return (nbytes + 255) // 256 * 256


def _layout(n: int, K: int, nt: int, nE: int, nlog: int = 0):
    """Byte offsets of one request's buffers, 256-byte aligned: the staged inputs
    (copied from the host, the FP64 state buffers last), the device scratch of the
    stages (never copied) and the outputs. Returns (input offsets, input bytes,
    scratch bytes, output offsets, output bytes)."""
    npd = pad16(n)
    nn = npd * npd
    lay_in, off = {}, 0
    for name, nbytes in (('H', nn * 8), ('D', max(K, 1) * nn * 8), ('C', max(K, 1) * nn * 8),
                         ('E', max(nE, 1) * nn * 8), ('ts', nt * 8), ('y', 2 * nn * 16)):
        lay_in[name] = off
        off += _align(nbytes)
    size_in = off
    work = _align(_native.work_elems(npd, K, 0) * 8) if _needs_work(npd, K) else 256
    lay_out, off = {}, 0
    for name, nbytes in (('st', nt * n * n * 8), ('fin', n * n * 8), ('ex', max(nE, 1) * nt * 8), ('stats', 64),
                         ('log', nlog * 8)):
        lay_out[name] = off
        off += _align(nbytes)
    return lay_in, size_in, work, lay_out, off


def footprint_bytes(n: int, K: int, ntsave: int, nE: int = 0, nlog: int = 0) -> int:
    """Device bytes one request holds while in flight (inputs, state and scratch,
    and its outputs when they are on the device)."""
    _, size_in, work, _, size_out = _layout(n, K, ntsave, nE, nlog)
    return size_in + work + size_out


# Requests with a padded dimension above this run on the grid-wide Tsit5 (every
# stage spread over many CUs) instead of one workgroup per request.
GRID_MIN_NP = 96


class _Launch:
    """One launch of packed requests: its device and pinned buffers, its requests
    and which of them have been answered."""

    def __init__(self, reqs, stream, tensors, outs, nbytes, t_launch, host_out):
        self.reqs = reqs
        self.stream = stream
        self.tensors = tensors
        self.outs = outs
        self.nbytes = nbytes
        self.t_launch = t_launch
        self.host_out = host_out  # numpy view of the pinned outputs (persistent launches)
        self.answered = [False] * len(reqs)
        if host_out is not None:
            # the completion words stats[7] of every request, polled in one gather
            # (set by the persistent kernel; a grid-wide launch answers at its end)
            self.words = host_out.view(np.int64)
            self.done_at = np.array([(o[2] + 56) // 8 for o in outs], dtype=np.int64)
            self.pending = np.ones(len(reqs), dtype=bool)
        self.event = None  # completion event (recorded after the kernel)
        self.ev0 = None  # timing event recorded before the launch
        self.ready = None  # set by the worker thread of a grid-wide launch
        self.error = None

    def done(self) -> bool:
        if self.ready is not None and not self.ready.is_set():
            return False
        return self.event is not None and self.event.query()


# Host threads that copy finished requests out of pinned memory and answer them
# (numpy releases the GIL in the copies), off the scheduler thread.
HOST_THREADS = max(1, min(8, os.cpu_count() or 1))
_pool_lock = threading.Lock()
_pool_obj = None


def _pool() -> ThreadPoolExecutor:
    global _pool_obj
    with _pool_lock:
        if _pool_obj is None:
            _pool_obj = ThreadPoolExecutor(max_workers=HOST_THREADS, thread_name_prefix='dqh-host')
        return _pool_obj


def _pack_host(reqs: list, base: int, staging: np.ndarray, descs: np.ndarray, lays, in_off, out_base: int, out_off,
               scratch_base: int, scratch_off, signal: bool) -> None:
    """Fill the staging buffer (inputs) and the descriptors of ``reqs``; ``base`` /
    ``out_base`` / ``scratch_base`` are the device addresses the staged inputs, the
    outputs and the scratch will have."""
    for b, r in enumerate(reqs):
        n, K, nt, nE = r.n, len(r.Ls), len(r.tsave), len(r.exp_ops)
        npd = pad16(n)
        nn = npd * npd
        li, _, _, lo, _ = lays[b]
        ib, ob = int(in_off[b]), int(out_off[b])

        def cview(a, count=1):
            return staging[ib + a : ib + a + count * nn * 8].view(np.complex64).reshape(count, npd, npd)

        # -iH and the jump operators; dqh_dense_prepare_ops completes Hnh and the
        # right factors C_k = L_k^dag / 2 on the device
        cview(li['H'])[0, :n, :n] = np.asarray(r.H) * -1j
        for k, L in enumerate(r.Ls):
            cview(li['D'] + k * nn * 8)[0, :n, :n] = L
        if nE:
            cview(li['E'], nE)[:, :n, :n] = np.swapaxes(np.stack([np.asarray(E) for E in r.exp_ops]), -1, -2)
        rho = np.asarray(r.rho0, dtype=np.complex128)
        if rho.shape[-1] == 1:
            rho = rho @ rho.conj().T
        a = ib + li['y']
        staging[a : a + nn * 16].view(np.complex128).reshape(npd, npd)[:n, :n] = rho
        a = ib + li['ts']
        staging[a : a + nt * 8].view(np.float64)[:] = np.asarray(r.tsave, dtype=np.float64)
        d = descs[b]
        d['n'], d['np'], d['nL'], d['nE'], d['ntsave'] = n, npd, K, nE, nt
        d['flags'] = _native.FLAG_SAVE_STATES | _native.FLAG_HAS_HNH | (_native.FLAG_SIGNAL if signal else 0)
        d['rtol'], d['atol'], d['safety'], d['factormin'], d['factormax'] = r.rtol, r.atol, 0.9, 0.2, 5.0
        d['max_steps'] = r.max_steps
        d['t0'], d['t1'] = float(r.tsave[0]), float(r.tsave[-1])
        d['tsave'] = base + ib + li['ts']
        d['Hnh'], d['Dops'] = base + ib + li['H'], base + ib + li['D']
        d['Cops'], d['EopsT'] = base + ib + li['C'], base + ib + li['E']
        d['y'], d['work'] = base + ib + li['y'], scratch_base + int(scratch_off[b])
        d['states'], d['final_state'] = out_base + ob + lo['st'], out_base + ob + lo['fin']
        d['expects'], d['stats'] = out_base + ob + lo['ex'], out_base + ob + lo['stats']
        if r.step_log_cap > 0:
            d['step_log'], d['step_log_cap'] = out_base + ob + lo['log'], r.step_log_cap


def pack_and_launch(reqs: list, stream, grid: bool = False, times: dict | None = None):
    """Pack uneven requests and launch them on ``stream``: the persistent Tsit5 kernel
    with per-request completion words, or (``grid``) a callable that runs the
    grid-wide Tsit5 on them (requests of one grid launch share their padded
    dimension); outputs in pinned host memory either way. Returns (tensors kept
    alive, per-request output layout, pinned output view, completion event or the
    callable, device bytes held). ``times`` (optional) accumulates the host seconds
    spent on buffers ('alloc'), packing ('pack') and the copy and launches ('launch')."""
    import torch

    t_start = time.perf_counter()
    dev = torch.device('cuda', torch.cuda.current_device())
    B = len(reqs)
    lays = [_layout(r.n, len(r.Ls), len(r.tsave), len(r.exp_ops), r.step_log_cap) for r in reqs]
    in_off = np.cumsum([0] + [lay[1] for lay in lays])
    w_off = np.cumsum([0] + [lay[2] for lay in lays])
    out_off = np.cumsum([0] + [lay[4] for lay in lays])
    dsz = B * _native.ME_INST_DTYPE.itemsize
    n_in = int(in_off[-1])
    # staging: every request's inputs, then the descriptor array: one copy
    staging_t = torch.zeros(n_in + _align(dsz), dtype=torch.uint8, pin_memory=True)
    staging = staging_t.numpy()
    # device buffers belong to the launch stream: the allocator's reuse and the
    # zero fill of device outputs are then ordered with the launch itself
    with torch.cuda.stream(stream):
        arena = torch.empty(n_in + _align(dsz), dtype=torch.uint8, device=dev)
        scratch = torch.empty(max(int(w_off[-1]), 256), dtype=torch.uint8, device=dev)
    descs = staging[n_in : n_in + dsz].view(_native.ME_INST_DTYPE)
    lib = _native.lib()
    c_p = _native.ctypes.c_void_p
    # outputs in pinned host memory that the kernels write through its device
    # address: a finished request is read on the host with no device copy and no
    # synchronisation. The kernels write every output; only the stats words (the
    # completion word of the persistent kernel among them) must start at zero.
    oarena = torch.empty(int(out_off[-1]), dtype=torch.uint8, pin_memory=True)
    out_base, host_out = _native.host_device_pointer(oarena.data_ptr()), oarena.numpy()
    for b in range(B):
        a = int(out_off[b]) + lays[b][3]['stats']
        host_out[a : a + 64] = 0
    t_alloc = time.perf_counter()
    _pack_host(reqs, arena.data_ptr(), staging, descs, lays, in_off, out_base, out_off, scratch.data_ptr(), w_off,
               signal=not grid)
    t_pack = time.perf_counter()
    if times is not None:
        times['alloc'] = times.get('alloc', 0.0) + (t_alloc - t_start)
        times['pack'] = times.get('pack', 0.0) + (t_pack - t_alloc)
    npd = int(descs['np'][0])
    if grid and np.any(descs['np'] != npd):
        raise ValueError('a grid-wide launch needs requests of one padded dimension')
    h_descs = descs.copy()
    d_descs_ptr = arena.data_ptr() + n_in
    outs = [(int(out_off[b]) + lays[b][3]['st'], int(out_off[b]) + lays[b][3]['ex'],
             int(out_off[b]) + lays[b][3]['stats'], r.n, len(r.tsave), len(r.exp_ops),
             int(out_off[b]) + lays[b][3]['log'], r.step_log_cap) for b, r in enumerate(reqs)]
    nbytes = int(arena.numel() + scratch.numel() + (oarena.numel() if grid else 0))
    with torch.cuda.stream(stream):
        arena.copy_(staging_t, non_blocking=True)
        _native.check(lib.dqh_dense_prepare_ops(c_p(d_descs_ptr), B, c_p(stream.cuda_stream)), 'dqh_dense_prepare_ops')
        if grid:
            nb = int(lib.dqh_tsit5_grid_scratch_bytes(npd, B))
            gscratch = torch.empty((nb,), dtype=torch.uint8, device=dev)
            flags = torch.zeros((2,), dtype=torch.int64, pin_memory=True)

            def run():
                # blocks this (worker) thread until the batch is done; ctypes releases the GIL
                with torch.cuda.stream(stream):
                    _native.check(
                        lib.dqh_mesolve_tsit5_dense_grid(
                            c_p(d_descs_ptr), h_descs.ctypes.data_as(c_p), B, c_p(gscratch.data_ptr()), nb, 16,
                            c_p(flags.data_ptr()), c_p(stream.cuda_stream),
                        ),
                        'dqh_mesolve_tsit5_dense_grid',
                    )
                    ev = torch.cuda.Event(enable_timing=True)
                    ev.record(stream)
                return ev

            keep = [oarena, arena, scratch, staging_t, gscratch, flags, h_descs]
            if times is not None:
                times['launch'] = times.get('launch', 0.0) + (time.perf_counter() - t_pack)
            return keep, outs, host_out, run, nbytes + nb
        _native.check(lib.dqh_mesolve_tsit5_dense(c_p(d_descs_ptr), B, c_p(stream.cuda_stream)),
                      'dqh_mesolve_tsit5_dense')
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(stream)
    if times is not None:
        times['launch'] = times.get('launch', 0.0) + (time.perf_counter() - t_pack)
    return [oarena, arena, scratch, staging_t], outs, host_out, ev, nbytes


def _read_outputs(host: np.ndarray, out) -> tuple:
    """(states, expects, nsteps, naccepted, nrejected, status, step_ends) of one
    request from its output region (a numpy byte array on the host)."""
    a_st, a_ex, a_stats, n, nt, nE, a_log, nlog = out
    st = host[a_st : a_st + nt * n * n * 8].view(np.complex64).reshape(nt, n, n).copy()
    ex = host[a_ex : a_ex + nE * nt * 8].view(np.complex64).reshape(nE, nt).copy() if nE else None
    stats = host[a_stats : a_stats + 64].view(np.int64)
    log = host[a_log : a_log + nlog * 8].view(np.float64)[: min(int(stats[1]), nlog)].copy() if nlog else None
    return st, ex, int(stats[0]), int(stats[1]), int(stats[2]), int(stats[3]), log


def collect(oarena, outs) -> list:
    """Outputs of a finished launch as host numpy arrays."""
    host = oarena.numpy() if oarena.device.type == 'cpu' else oarena.cpu().numpy()
    return [_read_outputs(host, o) for o in outs]


def status_error(status: int, nsteps: int, naccepted: int, nrejected: int) -> Exception:
    """The exception a request answers with when its solve stopped early (the same
    messages as ``dq.mesolve``: integrators/_utils.py:66-72 for max_steps)."""
    if status == _native.STATUS_MAX_STEPS:
        return RuntimeError('The maximum number of method steps has been reached.')
    if status == _native.STATUS_NONFINITE:
        return RuntimeError(f'solve stopped: non-finite error estimate or vanishing step size (after {nsteps} steps, '
                            f'{naccepted} accepted, {nrejected} rejected)')
    return RuntimeError(f'solve rejected by the kernel with status {status}')


def _request_bytes(r) -> int:
    """Device bytes a request holds while in flight, with its share of the grid-wide
    Tsit5 scratch when it runs grid-wide."""
    nb = footprint_bytes(r.n, len(r.Ls), len(r.tsave), len(r.exp_ops), r.step_log_cap)
    npd = pad16(r.n)
    if npd > GRID_MIN_NP:
        nb += int(_native.lib().dqh_tsit5_grid_scratch_bytes(npd, 1))
    return nb


def _size_class(r) -> int:
    """Launch grouping: small requests by size band (a persistent launch holds its
    memory until its slowest request is done), grid-wide requests by exact padded
    dimension."""
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
        self.stats = {'launches': 0, 'requests': 0, 'max_inflight_bytes': 0, 'max_concurrent_launches': 0,
                      'answered_before_launch_end': 0}
        # host seconds of the scheduler thread: buffers, packing, copy + launches, polling
        # and retirement
        self.host_s = {'alloc': 0.0, 'pack': 0.0, 'launch': 0.0, 'retire': 0.0}
        # device-busy accounting: a timing event before and after every launch,
        # measured from a reference event recorded now
        self._t_ref = torch.cuda.Event(enable_timing=True)
        self._t_ref.record(torch.cuda.current_stream())
        self._busy = []
        self._thread = threading.Thread(target=self._loop, name='dqh-server', daemon=True)
        self._thread.start()

    # ---------------------------------------------------------------- client
    def submit(self, H, Ls, rho0, tsave, *, rtol=1e-6, atol=1e-6, max_steps=100_000, exp_ops=None,
               step_log_cap: int = 0) -> Future:
        H = np.asarray(H)
        n = H.shape[-1]
        if H.shape != (n, n) or pad16(n) > MAX_NP:
            raise ValueError(f'serving accepts dense (n, n) Hamiltonians with n <= {MAX_NP}, got {H.shape}')
        tsave = np.asarray(tsave, dtype=np.float64)
        if tsave.ndim != 1 or len(tsave) == 0 or np.any(np.diff(tsave) < 0):
            raise ValueError('Argument tsave must be sorted in strictly ascending order.')
        req = Request(H=H, Ls=[np.asarray(L) for L in Ls], rho0=np.asarray(rho0), tsave=tsave, rtol=rtol, atol=atol,
                      max_steps=max_steps, exp_ops=[np.asarray(E) for E in (exp_ops or [])],
                      step_log_cap=int(step_log_cap))
        if _request_bytes(req) > self.memory_budget:
            raise ValueError('request does not fit in the server memory budget')
        req.future = Future()
        req.t_submit = time.perf_counter()
        self._q.put(req)
        return req.future

    def close(self) -> None:
        self._stop.set()
        self._thread.join()
        self.stats['busy_s'] = self.busy_seconds()

    def busy_seconds(self) -> float:
        """Time during which at least one launch of this server was on the device
        (union of the [start, end] intervals of the retired launches, HIP event
        timestamps)."""
        iv = sorted((self._t_ref.elapsed_time(a), self._t_ref.elapsed_time(b)) for a, b in self._busy)
        total, cur0, cur1 = 0.0, None, None
        for a, b in iv:
            if cur1 is None or a > cur1:
                if cur1 is not None:
                    total += cur1 - cur0
                cur0, cur1 = a, b
            else:
                cur1 = max(cur1, b)
        if cur1 is not None:
            total += cur1 - cur0
        return 1e-3 * total

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------- scheduler
    def _admit(self) -> None:
        if not self._pending or len(self._inflight) >= self.max_inflight:
            return
        # fill a launch with requests of the same size class as the oldest pending one
        # (FIFO inside the class)
        cls = _size_class(self._pending[0])
        batch, nbytes, keep = [], 0, []
        for r in self._pending:
            if len(batch) >= self.max_batch or _size_class(r) != cls:
                keep.append(r)
                continue
            fb = _request_bytes(r)
            if self._used + nbytes + fb > self.memory_budget:
                keep.append(r)
                continue
            batch.append(r)
            nbytes += fb
        self._pending = keep
        if not batch:
            return
        import torch

        stream = self.streams[self._next_stream % len(self.streams)]
        self._next_stream += 1
        t0 = time.perf_counter()
        grid = cls >= 1000
        ev0 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        try:
            tensors, outs, host_out, ev, total = pack_and_launch(batch, stream, grid=grid, times=self.host_s)
        except Exception as e:  # answer the batch with the error, keep serving
            for r in batch:
                r.future.set_exception(e)
            return
        self._used += total
        launch = _Launch(batch, stream, tensors, outs, total, t0, host_out)
        launch.ev0 = ev0
        if grid:
            launch.ready = threading.Event()

            def work(launch=launch, run=ev):
                try:
                    launch.event = run()
                except Exception as e:  # reported when the launch retires
                    launch.error = e
                finally:
                    launch.ready.set()

            threading.Thread(target=work, name='dqh-grid', daemon=True).start()
        else:
            launch.event = ev
        self._inflight.append(launch)
        self.stats['launches'] += 1
        self.stats['max_inflight_bytes'] = max(self.stats['max_inflight_bytes'], self._used)
        self.stats['max_concurrent_launches'] = max(self.stats['max_concurrent_launches'], len(self._inflight))

    def _answer(self, L: _Launch, b: int, res, t: float) -> None:
        """Complete request b of launch L with its outputs (any thread; the scheduler
        has already marked it answered and counted it)."""
        r = L.reqs[b]
        st, ex, ns, na, nr, status, log = res
        if status != 0:
            r.future.set_exception(status_error(status, ns, na, nr))
            return
        r.future.set_result(Response(st, ex, ns, na, nr, status, t - r.t_submit, L.t_launch - r.t_submit,
                                     step_ends=log))

    def _answer_pinned(self, L: _Launch, b: int, out) -> None:
        try:
            res = _read_outputs(L.host_out, out)
        except Exception as e:  # pragma: no cover - answered with the error
            L.reqs[b].future.set_exception(e)
            return
        self._answer(L, b, res, time.perf_counter())

    def _retire(self) -> bool:
        progress = False
        for L in list(self._inflight):
            if L.host_out is not None and L.ready is None:
                # per request: stats[7] is set by the kernel once that request is done
                # (all words of the launch read in one gather)
                newly = np.flatnonzero(L.pending & (L.words[L.done_at] == 1))
                if len(newly):
                    before_end = not L.done()
                    for b in newly.tolist():
                        L.pending[b] = False
                        L.answered[b] = True
                        self.stats['requests'] += 1
                        # the copy out of pinned memory runs on a host thread (the
                        # launch object keeps the pinned buffer alive until then)
                        _pool().submit(self._answer_pinned, L, b, L.outs[b])
                        if before_end:
                            self.stats['answered_before_launch_end'] += 1
                    progress = True
            finished = (L.ready is not None and L.ready.is_set() and L.error is not None) or L.done()
            if not finished:
                continue
            self._inflight.remove(L)
            progress = True
            if L.event is not None:
                self._busy.append((L.ev0, L.event))
            rest = [b for b in range(len(L.reqs)) if not L.answered[b]]
            for b in rest:
                L.answered[b] = True
            self.stats['requests'] += len(rest)
            if rest:
                if L.error is not None:
                    for b in rest:
                        L.reqs[b].future.set_exception(L.error)
                else:
                    # read out of the pinned outputs on the host threads
                    for b in rest:
                        _pool().submit(self._answer_pinned, L, b, L.outs[b])
            self._used -= L.nbytes
        return progress

    def _loop(self) -> None:
        while not (self._stop.is_set() and self._q.empty() and not self._pending and not self._inflight):
            got = False
            try:
                while True:
                    self._pending.append(self._q.get_nowait())
                    got = True
            except queue.Empty:
                pass
            t0 = time.perf_counter()
            retired = self._retire()
            self.host_s['retire'] += time.perf_counter() - t0
            self._admit()
            if not (got or retired):
                time.sleep(0.0002)
