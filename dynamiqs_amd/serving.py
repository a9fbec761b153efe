"""Continuous-batching mesolve server on one MI355X.

Requests (one dense Tsit5 solve each, any Hilbert dimension n <= 256, any number
of jump operators, own save times and tolerances) are admitted one by one while
others are on the GPU. A scheduler thread sorts them by padded dimension:

* Small requests (padded dimension below GRID_MIN_NP) go to a work queue. A
  persistent grid of ``queue_workgroups`` workgroups
  (``dqh_mesolve_tsit5_dense_queue``) keeps taking requests from a ring in pinned
  host memory: the host stages a request's inputs (-iH, jump operators, initial
  state, save times) and its descriptor in coherent pinned memory, writes a ring
  slot and advances the tail; a free workgroup claims it, copies the inputs to HBM
  itself, completes -iH - 1/2 sum L^dag L and L^dag / 2 (FP64 products), solves it
  and signals it. No copy, launch or synchronisation is needed to start a request,
  and a workgroup takes the next request as soon as it is done, so a request never
  waits for the slowest member of a batch. The grid leaves when no request has
  been claimed for ``idle_ms``, and its workgroups stop claiming after ``life_ms``
  (the scheduler relaunches it while requests are outstanding).
* Large requests run on the grid-wide Tsit5 (each stage spread over many CUs,
  driven by a worker thread), batched by buckets of padded dimension, on two further streams.

The three streams are created together once per process, so that they get
distinct hardware queues and the persistent grid never holds a grid-wide launch
behind it. Device memory (inputs, FP64 state, stage scratch) is bounded by a
memory budget.

Retirement is per request: outputs (saved states, expectation values, final
state, statistics) live in pinned host memory that the kernels write directly, and
each workgroup publishes its request's completion in stats[7] with a system-scope
release (``DQH_FLAG_SIGNAL``). The scheduler polls those words and host threads
copy each answer out as soon as its own solve is done.

The result of a request does not depend on what it was served with: the kernels
keep no cross-instance state and each workgroup's arithmetic is fixed by its own
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


def _layout(n: int, K: int, nt: int, nE: int, nlog: int = 0, npd: int | None = None):
    """Byte offsets of one request's buffers, 256-byte aligned: the staged inputs
    (copied from the host, the FP64 state buffers last), the device scratch of the
    stages (never copied) and the outputs. ``npd``: the padded dimension the request
    runs at (default pad16(n)). Returns (input offsets, input bytes, scratch bytes,
    output offsets, output bytes)."""
    npd = npd or pad16(n)
    nn = npd * npd
    lay_in, off = {}, 0
    # what the host fills comes first (up to the first FP64 state buffer: 'copy'
    # bytes); C is written on the device
    for name, nbytes in (('H', nn * 8), ('D', max(K, 1) * nn * 8), ('E', max(nE, 1) * nn * 8), ('ts', nt * 8),
                         ('y', 2 * nn * 16), ('C', max(K, 1) * nn * 8)):
        lay_in[name] = off
        off += _align(nbytes)
    lay_in['copy'] = lay_in['y'] + nn * 16
    size_in = off
    work = _align(_native.work_elems(npd, K, 0) * 8) if _needs_work(npd, K) else 256
    lay_out, off = {}, 0
    for name, nbytes in (('st', nt * n * n * 8), ('fin', n * n * 8), ('ex', max(nE, 1) * nt * 8), ('stats', 64),
                         ('log', nlog * 8)):
        lay_out[name] = off
        off += _align(nbytes)
    return lay_in, size_in, work, lay_out, off


def footprint_bytes(n: int, K: int, ntsave: int, nE: int = 0, nlog: int = 0, npd: int | None = None) -> int:
    """Device bytes one request holds while in flight (inputs, state and scratch,
    and its outputs when they are on the device), at padded dimension ``npd``."""
    _, size_in, work, _, size_out = _layout(n, K, ntsave, nE, nlog, npd)
    return size_in + work + size_out


# Requests with a padded dimension above this run on the grid-wide Tsit5 (every
# stage spread over many CUs) instead of one workgroup per request.
GRID_MIN_NP = 96
# Under a deep backlog, a grid-wide launch takes the requests of one exact padded
# dimension when at least this many are pending; otherwise those of its bucket
# (_grid_bucket), padded to it.
EXACT_CLASS_MIN = 8


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
        self.queued = False  # requests published to the work queue (no launch of their own)
        self.slots = None  # their ring slots
        self.blocks = None  # their coherent pinned blocks (inputs, outputs)
        self.unread = 0  # answers still to be read out of the output block

    def done(self) -> bool:
        if self.queued:
            return not self.pending.any()
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
        npd = _launch_np(r)
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


def pack_and_launch(reqs: list, stream, grid: bool = False, times: dict | None = None, pool=None, out_pool=None,
                    blocks_out: list | None = None):
    """Pack uneven requests and launch them on ``stream``: the persistent Tsit5 kernel
    with per-request completion words, or (``grid``) a callable that runs the
    grid-wide Tsit5 on them (requests of one grid launch share their padded
    dimension); outputs in pinned host memory either way. Returns (tensors kept
    alive, per-request output layout, pinned output view, completion event or the
    callable, device bytes held). ``times`` (optional) accumulates the host seconds
    spent on buffers ('alloc'), packing ('pack') and the copy and launches ('launch').
    With ``pool`` and ``out_pool`` (pinned pools), the staging and output buffers are
    blocks of them (appended to ``blocks_out``: inputs, outputs) instead of new pinned
    allocations, which can take milliseconds each."""
    import torch

    t_start = time.perf_counter()
    dev = torch.device('cuda', torch.cuda.current_device())
    B = len(reqs)
    lays = [_layout(r.n, len(r.Ls), len(r.tsave), len(r.exp_ops), r.step_log_cap, _launch_np(r)) for r in reqs]
    in_off = np.cumsum([0] + [lay[1] for lay in lays])
    w_off = np.cumsum([0] + [lay[2] for lay in lays])
    out_off = np.cumsum([0] + [lay[4] for lay in lays])
    dsz = B * _native.ME_INST_DTYPE.itemsize
    n_in = int(in_off[-1])
    # staging: every request's inputs, then the descriptor array: one copy
    if pool is not None:
        in_blk = pool.alloc(n_in + _align(dsz))
        staging_t = None
        staging = pool.view(in_blk)
        staging[:] = 0
    else:
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
    if pool is not None:
        out_blk = out_pool.alloc(int(out_off[-1]))
        oarena, out_base, host_out = None, out_pool.device(out_blk), out_pool.view(out_blk)
        if blocks_out is not None:
            blocks_out += [in_blk, out_blk]
    else:
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
    nbytes = int(arena.numel() + scratch.numel())  # device bytes (the outputs are in host memory)
    with torch.cuda.stream(stream):
        if staging_t is not None:
            arena.copy_(staging_t, non_blocking=True)
        else:
            _native.check(lib.dqh_memcpy_h2d_async(c_p(arena.data_ptr()), c_p(staging.ctypes.data), int(staging.nbytes),
                                                   c_p(stream.cuda_stream)), 'dqh_memcpy_h2d_async')
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


# ------------------------------------------------------------------ work queue
SLOT_WORDS = 8  # DQH_QUEUE_SLOT_WORDS


class _DQueue(_native.ctypes.Structure):
    """Host mirror of ``dqh_queue`` (include/dqh.h)."""

    _fields_ = [('ring', _native.ctypes.c_uint64), ('capacity', _native.ctypes.c_int64),
                ('tail', _native.ctypes.c_uint64), ('head', _native.ctypes.c_uint64),
                ('stop', _native.ctypes.c_uint64), ('busy_ticks', _native.ctypes.c_uint64),
                ('idle_ticks', _native.ctypes.c_int64), ('life_ticks', _native.ctypes.c_int64)]


class _PinnedPool:
    """Pinned, device-mapped host memory (``dqh_host_alloc``) in chunks of CHUNK bytes,
    sub-allocated first fit (256-byte blocks) and kept for the life of the process,
    since freeing pinned memory can wait for the device. The coherent (fine-grained)
    pool holds what the persistent queue kernel reads while the host writes it
    (staged inputs, ring and control words): the device never reads a stale cached
    copy. The non-coherent pool holds kernel outputs, which the kernels publish with
    a system-scope release and whose stores the device can combine."""

    CHUNK = 256 << 20

    def __init__(self, coherent: bool):
        self.coherent = bool(coherent)
        self._chunks = []  # [device address, uint8 view, free list of (offset, bytes)]
        self._lock = threading.Lock()

    def alloc(self, nbytes: int):
        nbytes = _align(max(int(nbytes), 16))
        with self._lock:
            for ci, ch in enumerate(self._chunks):
                free = ch[2]
                for k, (off, n) in enumerate(free):
                    if n >= nbytes:
                        if n == nbytes:
                            free.pop(k)
                        else:
                            free[k] = (off + nbytes, n - nbytes)
                        return (ci, off, nbytes)
            size = max(self.CHUNK, nbytes)
            ptr = _native.ctypes.c_void_p()
            _native.check(_native.lib().dqh_host_alloc(size, int(self.coherent), _native.ctypes.byref(ptr)),
                          'dqh_host_alloc')
            view = np.ctypeslib.as_array((_native.ctypes.c_uint8 * size).from_address(ptr.value))
            self._chunks.append([_native.host_device_pointer(ptr.value), view, [(nbytes, size - nbytes)] if size > nbytes else []])
            return (len(self._chunks) - 1, 0, nbytes)

    def view(self, blk) -> np.ndarray:
        ci, off, n = blk
        return self._chunks[ci][1][off : off + n]

    def device(self, blk) -> int:
        return self._chunks[blk[0]][0] + blk[1]

    def free(self, blk) -> None:
        ci, off, n = blk
        with self._lock:
            free = sorted(self._chunks[ci][2] + [(off, n)])
            merged = []
            for o, m in free:
                if merged and merged[-1][0] + merged[-1][1] == o:
                    merged[-1] = (merged[-1][0], merged[-1][1] + m)
                else:
                    merged.append((o, m))
            self._chunks[ci][2] = merged


_pools = {}
_streams = None


def _pinned_pool(coherent: bool) -> _PinnedPool:
    with _pool_lock:
        if coherent not in _pools:
            _pools[coherent] = _PinnedPool(coherent)
        return _pools[coherent]


def _serving_streams():
    """The queue grid's two streams (consecutive launches alternate, so that one can
    start while the previous drains) and the grid-wide launches' two streams,
    created together once per process and shared by every Server: streams created
    in a row get distinct hardware queues, so the persistent queue grid shares none
    with a grid-wide launch (a kernel waits for the earlier ones of its hardware
    queue)."""
    global _streams
    import torch

    with _pool_lock:
        if _streams is None:
            _streams = ([torch.cuda.Stream(), torch.cuda.Stream()], [torch.cuda.Stream(), torch.cuda.Stream()])
        return _streams


class _WorkQueue:
    """The ring of ``dqh_mesolve_tsit5_dense_queue`` and its persistent launches.
    ``publish`` fills ring slots and then advances tail. The scheduler keeps a launch
    of the grid running while published requests are outstanding: a launch's
    workgroups leave when no claim was made for idle_ms or stop claiming after
    life_ms (an idle workgroup then leaves at once, a busy one after its request), so
    the next launch is started on the other stream once the newest one is life_ms
    old: it claims from the same head and takes the CUs the old launch frees while
    that one drains its last requests."""

    def __init__(self, capacity: int, workgroups: int, idle_ms: float, life_ms: float):
        import torch

        lib = _native.lib()
        if capacity & (capacity - 1):
            raise ValueError('the ring capacity must be a power of two')
        self.capacity, self.workgroups = capacity, workgroups
        self.pool = _pinned_pool(True)
        self.out_pool = _pinned_pool(False)
        nwords = capacity * SLOT_WORDS + 64
        self.blk = self.pool.alloc(nwords * 8)
        words = self.pool.view(self.blk).view(np.uint64)
        words[:] = 0
        self.ring = words[: capacity * SLOT_WORDS].reshape(capacity, SLOT_WORDS)
        self.ctrl = words[capacity * SLOT_WORDS :]  # ctrl[0]: tail, ctrl[32]: stop (other cache lines)
        base = self.pool.device(self.blk)
        self.counters = torch.zeros(64, dtype=torch.int64, device=torch.device('cuda', torch.cuda.current_device()))
        self.khz = int(lib.dqh_wall_clock_khz()) or 100_000
        self.q = _DQueue(ring=base, capacity=capacity, tail=base + capacity * SLOT_WORDS * 8,
                         head=self.counters.data_ptr(), stop=base + (capacity * SLOT_WORDS + 32) * 8,
                         busy_ticks=self.counters.data_ptr() + 256, idle_ticks=int(idle_ms * self.khz),
                         life_ticks=int(life_ms * self.khz))
        torch.cuda.synchronize()  # the counters are zero before the first launch
        self.tail = 0
        self.outstanding = 0  # published and not yet completed
        self.slot_busy = np.zeros(capacity, dtype=bool)
        self.streams = _serving_streams()[0]
        self.life_s = float(life_ms) * 1e-3
        self.gens = []  # (stream, start event, end event, host launch time) of launches not known to be done
        self.event = None  # end event of the newest launch
        self.launches = 0
        self._lock = threading.Lock()

    def free_slots(self) -> int:
        return self.capacity - self.outstanding

    def publish(self, entries) -> list:
        """Write one slot per entry (src, dst, bytes, src_desc, dst_desc) and advance
        tail; returns the slots used."""
        slots = []
        for e in entries:
            sl = self.tail & (self.capacity - 1)
            if self.slot_busy[sl]:
                raise RuntimeError('work-queue slot still in use')
            self.ring[sl, :5] = e
            self.slot_busy[sl] = True
            slots.append(sl)
            self.tail += 1
        self.outstanding += len(entries)
        # the lock's atomic operations order the staged inputs and the slots (also
        # write-combined stores) before the tail store that publishes them
        with self._lock:
            self.ctrl[0] = self.tail
        return slots

    def complete(self, slot: int) -> None:
        self.slot_busy[slot] = False
        self.outstanding -= 1

    def running(self) -> bool:
        return self.event is not None and not self.event.query()

    def retire_done(self) -> list:
        """(start, end) events of the launches that have ended since the last call."""
        done = [g for g in self.gens if g[2].query()]
        self.gens = [g for g in self.gens if g not in done]
        return [(g[1], g[2]) for g in done]

    def needs_launch(self) -> bool:
        """Requests are outstanding and no running launch will keep claiming for long."""
        if self.outstanding == 0:
            return False
        live = [g for g in self.gens if not g[2].query()]
        return not live or time.perf_counter() - live[-1][3] > self.life_s

    def launch(self) -> None:
        import torch

        stream = self.streams[self.launches % len(self.streams)]
        with torch.cuda.stream(stream):
            ev0 = torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            _native.check(_native.lib().dqh_mesolve_tsit5_dense_queue(_native.ctypes.byref(self.q), self.workgroups,
                                                                      _native.ctypes.c_void_p(stream.cuda_stream)),
                          'dqh_mesolve_tsit5_dense_queue')
            self.event = torch.cuda.Event(enable_timing=True)
            self.event.record(stream)
        self.gens.append((stream, ev0, self.event, time.perf_counter()))
        self.launches += 1

    def busy_seconds(self) -> float:
        return float(self.counters[32].item()) / (self.khz * 1e3)

    def close(self) -> list:
        """Stop the grid; returns the (start, end) events of the launches not yet retired."""
        with self._lock:
            self.ctrl[32] = 1
        for g in self.gens:
            g[2].synchronize()
        rest = [(g[1], g[2]) for g in self.gens]
        self.gens = []
        self.pool.free(self.blk)
        return rest


def pack_queued(reqs: list, pool: _PinnedPool, out_pool: _PinnedPool, times: dict | None = None):
    """Stage ``reqs`` for the work queue: inputs and descriptors in coherent pinned
    memory (``pool``), device buffers for their copies and scratch (allocated, not
    written: the queue kernel copies the inputs itself), outputs in non-coherent
    pinned memory (``out_pool``).
    Returns (device tensors, [input block, output block], per-request output layout,
    output view, ring entries, device bytes)."""
    import torch

    t_start = time.perf_counter()
    dev = torch.device('cuda', torch.cuda.current_device())
    B = len(reqs)
    lays = [_layout(r.n, len(r.Ls), len(r.tsave), len(r.exp_ops), r.step_log_cap, _launch_np(r)) for r in reqs]
    in_off = np.cumsum([0] + [lay[1] for lay in lays])
    w_off = np.cumsum([0] + [lay[2] for lay in lays])
    out_off = np.cumsum([0] + [lay[4] for lay in lays])
    isz = _native.ME_INST_DTYPE.itemsize
    n_in = int(in_off[-1])
    in_blk = pool.alloc(n_in + _align(B * isz))
    out_blk = out_pool.alloc(int(out_off[-1]))
    staging = pool.view(in_blk)
    staging[:] = 0
    host_out = out_pool.view(out_blk)
    arena = torch.empty(n_in + _align(B * isz), dtype=torch.uint8, device=dev)
    scratch = torch.empty(max(int(w_off[-1]), 256), dtype=torch.uint8, device=dev)
    for b in range(B):
        a = int(out_off[b]) + lays[b][3]['stats']
        host_out[a : a + 64] = 0
    t_alloc = time.perf_counter()
    descs = staging[n_in : n_in + B * isz].view(_native.ME_INST_DTYPE)
    _pack_host(reqs, arena.data_ptr(), staging, descs, lays, in_off, out_pool.device(out_blk), out_off,
               scratch.data_ptr(), w_off, signal=True)
    src, dst = pool.device(in_blk), arena.data_ptr()
    entries = [(src + int(in_off[b]), dst + int(in_off[b]), lays[b][0]['copy'], src + n_in + b * isz,
                dst + n_in + b * isz) for b in range(B)]
    outs = [(int(out_off[b]) + lays[b][3]['st'], int(out_off[b]) + lays[b][3]['ex'],
             int(out_off[b]) + lays[b][3]['stats'], r.n, len(r.tsave), len(r.exp_ops),
             int(out_off[b]) + lays[b][3]['log'], r.step_log_cap) for b, r in enumerate(reqs)]
    if times is not None:
        times['queue_alloc'] = times.get('queue_alloc', 0.0) + (t_alloc - t_start)
        times['queue_pack'] = times.get('queue_pack', 0.0) + (time.perf_counter() - t_alloc)
    return [arena, scratch], [in_blk, out_blk], outs, host_out, entries, int(arena.numel() + scratch.numel())


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
    npd = pad16(r.n)
    if npd > GRID_MIN_NP:
        npd = _grid_bucket(npd)
        nb = footprint_bytes(r.n, len(r.Ls), len(r.tsave), len(r.exp_ops), r.step_log_cap, npd=npd)
        return nb + int(_native.lib().dqh_tsit5_grid_scratch_bytes(npd, 1))
    return footprint_bytes(r.n, len(r.Ls), len(r.tsave), len(r.exp_ops), r.step_log_cap)


def _grid_bucket(npd: int) -> int:
    """Padded dimension a grid-wide request runs at: 128, 160, 192, 224 or 256, then
    multiples of 64. Requests of one bucket share a launch (the grid-wide kernels take one padded
    dimension per launch); the extra padding holds exact zeros, and each request keeps
    its own n for the error norm and the outputs."""
    for b in (128, 160, 192, 224, 256):
        if npd <= b:
            return b
    return min((npd + 63) // 64 * 64, MAX_NP)


def _launch_np(r) -> int:
    """The padded dimension of a request in its launch (set when a grid-wide batch is formed)."""
    return getattr(r, 'np_launch', None) or pad16(r.n)


def _size_class(r) -> int:
    """Launch grouping: small requests by size band (a persistent launch holds its
    memory until its slowest request is done), grid-wide requests by their bucket of
    padded dimensions."""
    npd = pad16(r.n)
    if npd > GRID_MIN_NP:
        return 1000 + _grid_bucket(npd)
    return 0 if npd <= 32 else (1 if npd <= 64 else 2)


class Server:
    """Continuous-batching request loop (see module docstring).

    Args:
        memory_budget: device bytes that in-flight requests may hold together.
        max_batch: maximum requests staged together (one work-queue batch or one
            grid-wide launch).
        nstreams: streams (1 or 2, default priority) the grid-wide launches rotate over.
        max_inflight: maximum concurrent grid-wide launches.
        queue_workgroups: workgroups of the persistent work-queue grid (one per CU;
            the other CUs stay free for grid-wide launches).
        queue_capacity: ring slots (a power of two): requests queued or in service.
        idle_ms: the work-queue grid leaves after this long without a claim.
        life_ms: its workgroups stop claiming after this long (then it is relaunched).
            Work queued behind the persistent kernel in the same hardware queue waits
            up to this long, so it is kept short.

    The server's four streams (two for the work queue, two grid-wide) and the
    process's other streams should each have a hardware queue of their own: with
    HIP's default of 4 (``GPU_MAX_HW_QUEUES``) a grid-wide launch can share one with
    the persistent kernel and wait for its end. Set ``GPU_MAX_HW_QUEUES=16`` before the
    process initialises HIP (``bench.py`` does); a warning is issued otherwise.
    """

    def __init__(self, memory_budget: int = 8 << 30, max_batch: int = 64, nstreams: int = 2, max_inflight: int = 8,
                 queue_workgroups: int = 128, queue_capacity: int = 4096, idle_ms: float = 20.0, life_ms: float = 50.0):
        import os
        import warnings

        import torch

        if not torch.cuda.is_available():
            raise RuntimeError('dynamiqs_amd.serving needs an MI355X GPU')
        if int(os.environ.get('GPU_MAX_HW_QUEUES', '4')) < 8:
            warnings.warn('dynamiqs_amd.serving: fewer than 8 hardware queues (GPU_MAX_HW_QUEUES); grid-wide launches '
                          'may wait behind the persistent work-queue kernel. Set GPU_MAX_HW_QUEUES=16 before HIP starts.',
                          stacklevel=2)
        _native.lib()
        self.memory_budget = int(memory_budget)
        self.max_batch = int(max_batch)
        self.max_inflight = int(max_inflight)
        self.streams = _serving_streams()[1][: max(1, int(nstreams))]
        # small requests go to the persistent work-queue kernel (its own stream)
        self._wq = _WorkQueue(int(queue_capacity), int(queue_workgroups), float(idle_ms), float(life_ms))
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
        self.host_s = {'alloc': 0.0, 'pack': 0.0, 'queue_alloc': 0.0, 'queue_pack': 0.0, 'launch': 0.0, 'retire': 0.0}
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
        self._busy += self._wq.close()
        self.stats['busy_s'] = self.busy_seconds()
        self.stats['queue_launches'] = self._wq.launches
        self.stats['queue_busy_s'] = self._wq.busy_seconds()
        self.stats['queue_workgroups'] = self._wq.workgroups

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
        """Small requests (padded dimension below GRID_MIN_NP) go to the work queue as
        soon as they are pending (ring slots and the memory budget permitting): they
        never wait behind large ones. Large requests are batched by bucket (_grid_bucket), the
        oldest first, into a grid-wide launch when one of max_inflight is free."""
        if not self._pending:
            return
        import torch

        small, large = [], []
        for r in self._pending:
            (small if _size_class(r) < 1000 else large).append(r)
        keep = []
        if small:
            batch, nbytes = [], 0
            room = min(self.max_batch, self._wq.free_slots())
            for r in small:
                fb = _request_bytes(r)
                if len(batch) < room and self._used + nbytes + fb <= self.memory_budget:
                    batch.append(r)
                    nbytes += fb
                else:
                    keep.append(r)
            if batch:
                t0 = time.perf_counter()
                try:
                    tensors, blocks, outs, host_out, entries, total = pack_queued(batch, self._wq.pool, self._wq.out_pool,
                                                                                  times=self.host_s)
                except Exception as e:  # answer the batch with the error, keep serving
                    for r in batch:
                        r.future.set_exception(e)
                else:
                    launch = _Launch(batch, None, tensors, outs, total, t0, host_out)
                    launch.queued, launch.blocks, launch.unread = True, blocks, len(batch)
                    t1 = time.perf_counter()
                    launch.slots = self._wq.publish(entries)
                    self._used += total
                    self._inflight.append(launch)
                    self._ensure_queue()
                    self.host_s['launch'] += time.perf_counter() - t1
                    self._count_launch()
        if large and sum(1 for L in self._inflight if not L.queued) < self.max_inflight:
            # with a deep backlog (more than two full launches), the oldest request's exact
            # padded dimension, which wastes no work on padding; otherwise its bucket:
            # fewer, fuller launches that keep the latency of a steady stream low
            exact = pad16(large[0].n)
            if len(large) > 2 * self.max_batch and sum(1 for r in large if pad16(r.n) == exact) >= EXACT_CLASS_MIN:
                launch_np, member = exact, (lambda r: pad16(r.n) == exact)
            else:
                launch_np = _grid_bucket(exact)
                member = lambda r: _grid_bucket(pad16(r.n)) == launch_np  # noqa: E731
            batch, nbytes = [], 0
            for r in large:
                fb = _request_bytes(r)
                if member(r) and len(batch) < self.max_batch and self._used + nbytes + fb <= self.memory_budget:
                    batch.append(r)
                    nbytes += fb
                else:
                    keep.append(r)
            if batch:
                for r in batch:  # one padded dimension per grid-wide launch
                    r.np_launch = launch_np
                t0 = time.perf_counter()
                stream = self.streams[self._next_stream % len(self.streams)]
                self._next_stream += 1
                ev0 = torch.cuda.Event(enable_timing=True)
                ev0.record(stream)
                blocks = []
                try:
                    tensors, outs, host_out, ev, total = pack_and_launch(batch, stream, grid=True, times=self.host_s,
                                                                         pool=self._wq.pool, out_pool=self._wq.out_pool,
                                                                         blocks_out=blocks)
                except Exception as e:  # answer the batch with the error, keep serving
                    for blk, pl in zip(blocks, (self._wq.pool, self._wq.out_pool)):
                        pl.free(blk)
                    for r in batch:
                        r.future.set_exception(e)
                else:
                    self._used += total
                    launch = _Launch(batch, stream, tensors, outs, total, t0, host_out)
                    launch.blocks, launch.unread = blocks, len(batch)
                    launch.ev0 = ev0
                    launch.ready = threading.Event()

                    def work(launch=launch, run=ev):
                        try:
                            launch.event = run()
                        except Exception as e:  # reported when the launch retires
                            launch.error = e
                        finally:
                            launch.ready.set()

                    threading.Thread(target=work, name='dqh-grid', daemon=True).start()
                    self._inflight.append(launch)
                    self._count_launch()
        else:
            keep += large
        # FIFO order of what stays pending
        kept = set(id(r) for r in keep)
        self._pending = [r for r in self._pending if id(r) in kept]

    def _count_launch(self) -> None:
        self.stats['launches'] += 1
        self.stats['max_inflight_bytes'] = max(self.stats['max_inflight_bytes'], self._used)
        self.stats['max_concurrent_launches'] = max(self.stats['max_concurrent_launches'], len(self._inflight))

    def _ensure_queue(self) -> None:
        """Keep a launch of the work-queue grid running while requests are outstanding
        (a launch that has ended has completed every request it took)."""
        self._busy += self._wq.retire_done()
        if self._wq.needs_launch():
            self._wq.launch()

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
        finally:
            if L.blocks is not None:
                with _pool_lock:
                    L.unread -= 1
                    last = L.unread == 0
                if last:  # every answer of the batch is read out: its output block is free
                    self._wq.out_pool.free(L.blocks[1])
        self._answer(L, b, res, time.perf_counter())

    def _retire(self) -> bool:
        progress = False
        for L in list(self._inflight):
            if L.host_out is not None and L.ready is None:
                # per request: stats[7] is set by the kernel once that request is done
                # (all words of the launch read in one gather)
                newly = np.flatnonzero(L.pending & (L.words[L.done_at] == 1))
                if len(newly):
                    before_end = not (L.pending.sum() == len(newly)) if L.queued else not L.done()
                    for b in newly.tolist():
                        L.pending[b] = False
                        L.answered[b] = True
                        self.stats['requests'] += 1
                        if L.queued:
                            self._wq.complete(L.slots[b])
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
            if L.blocks is not None:  # the inputs were copied to the device: their staging block is free
                self._wq.pool.free(L.blocks[0])
            if L.event is not None:
                self._busy.append((L.ev0, L.event))
            rest = [b for b in range(len(L.reqs)) if not L.answered[b]]
            for b in rest:
                L.answered[b] = True
            self.stats['requests'] += len(rest)
            if rest:
                if L.error is not None:
                    if L.blocks is not None:
                        self._wq.out_pool.free(L.blocks[1])
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
            self._ensure_queue()
            if not (got or retired):
                time.sleep(0.0002)
