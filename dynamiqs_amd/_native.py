"""ctypes binding of ``libdqh.so`` (the HIP C-ABI declared in ``include/dqh.h``).

The product path has no CPU fallback: if the library is missing or cannot be
loaded, :func:`lib` raises ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libdqh.so')

# numpy mirror of ``dqh_me_inst`` (all fields 8 bytes, no padding)
ME_INST_FIELDS = [
    ('n', np.int64),
    ('np', np.int64),
    ('nL', np.int64),
    ('nE', np.int64),
    ('ntsave', np.int64),
    ('npwc', np.int64),
    ('nseg', np.int64),
    ('flags', np.int64),
    ('rtol', np.float64),
    ('atol', np.float64),
    ('safety', np.float64),
    ('factormin', np.float64),
    ('factormax', np.float64),
    ('max_steps', np.int64),
    ('t0', np.float64),
    ('t1', np.float64),
    ('tsave', np.uint64),
    ('Hnh', np.uint64),
    ('Dops', np.uint64),
    ('Cops', np.uint64),
    ('EopsT', np.uint64),
    ('Hpwc', np.uint64),
    ('pwc_times', np.uint64),
    ('pwc_vals', np.uint64),
    ('step_t', np.uint64),
    ('step_set', np.uint64),
    ('y', np.uint64),
    ('work', np.uint64),
    ('states', np.uint64),
    ('expects', np.uint64),
    ('final_state', np.uint64),
    ('stats', np.uint64),
    ('step_log', np.uint64),
    ('step_log_cap', np.int64),
    ('nmod', np.int64),
    ('Hmod', np.uint64),
    ('mod_coef', np.uint64),
    ('mod_t0', np.float64),
    ('mod_dt', np.float64),
    ('mod_nint', np.int64),
    ('hdiag', np.uint64),
]
MAX_MOD = 4
ME_INST_DTYPE = np.dtype(ME_INST_FIELDS)

FLAG_SAVE_STATES = 1
FLAG_HAS_HNH = 2
FLAG_GLOBAL = 4
FLAG_DOPRI5 = 8
FLAG_EULER = 16
FLAG_LMOD = 32
FLAG_SIGNAL = 64
FLAG_DOPRI8 = 128
FLAG_LPWC = 256
FLAG_LPAIR = 512

STATUS_OK = 0
STATUS_MAX_STEPS = -3
STATUS_NONFINITE = -4

# exported symbols (name, restype, argtypes); kept in sync with include/dqh.h
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
SYMBOLS = {
    'dqh_sizeof_me_inst': (_I64, []),
    'dqh_me_work_elems': (_I64, [_I64, _I64, _I64]),
    'dqh_lindblad_rhs_dense': (ctypes.c_int, [_P, _I64, _P]),
    'dqh_dense_prepare_ops': (ctypes.c_int, [_P, _I64, _P]),
    'dqh_mesolve_tsit5_dense': (ctypes.c_int, [_P, _I64, _P]),
    'dqh_mesolve_rouchon1_dense': (ctypes.c_int, [_P, _I64, _P]),
    'dqh_sizeof_ar_inst': (_I64, []),
    'dqh_ar_work_elems': (_I64, [_I64]),
    'dqh_mesolve_rouchon_adaptive_dense': (ctypes.c_int, [_P, _I64, _P]),
    'dqh_zgemm_dense': (ctypes.c_int, [_P, _P, _P, _I64, _I64, _I64, _I64, _P]),
    'dqh_mesolve_tsit5_dense_uniform': (ctypes.c_int, [_P, _I64, _I64, _P]),
    'dqh_mesolve_tsit5_dense_uniform_fp64diag': (ctypes.c_int, [_P, _I64, _I64, _P]),
    'dqh_tsit5_grid_scratch_bytes': (_I64, [_I64, _I64]),
    'dqh_mesolve_tsit5_dense_grid': (ctypes.c_int, [_P, _P, _I64, _P, _I64, _I64, _P, _P]),
    'dqh_sizeof_dia_rouchon1': (_I64, []),
    'dqh_mesolve_rouchon1_dia': (ctypes.c_int, [_P, _P]),
    'dqh_sizeof_dia_rhs': (_I64, []),
    'dqh_lindblad_rhs_dia': (ctypes.c_int, [_P, _P]),
    'dqh_sizeof_dia_tsit5': (_I64, []),
    'dqh_dia_tsit5_work_bytes': (_I64, [_I64, _I64]),
    'dqh_mesolve_tsit5_dia': (ctypes.c_int, [_P, _P]),
    'dqh_sesolve_tsit5': (ctypes.c_int, [_P, _I64, _I64, _P]),
    'dqh_expm_work_bytes': (_I64, [_I64, _I64, _I64]),
    'dqh_expm_propagators': (ctypes.c_int, [_P, _P, _P, _I64, _I64, _I64, _I64, _I64, _P]),
    'dqh_expm_propagators_noalloc': (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _I64, _I64, _P]),
    'dqh_cgemm_batched': (ctypes.c_int, [_P, _I64, _P, _I64, _I64, _P, _I64, _I64, _I64, _P]),
    'dqh_adjoint_pwc_quadrature': (ctypes.c_int, [_I64, _I64, _I64, _I64, _I64, _P, _P, _P, _P, _P, _P, _P]),
    'dqh_adjoint_pwc_quadrature_dia': (ctypes.c_int, [_I64, _I64, _I64, _I64, _P, _P, _P, _P, _P, _P, _P]),
    'dqh_adjoint_rate_quadrature': (ctypes.c_int, [_I64, _I64, _I64, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P]),
    'dqh_host_device_pointer': (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    'dqh_host_alloc': (ctypes.c_int, [_I64, _I64, ctypes.POINTER(_P)]),
    'dqh_host_free': (ctypes.c_int, [_P]),
    'dqh_mesolve_tsit5_dense_queue': (ctypes.c_int, [_P, _I64, _P]),
    'dqh_wall_clock_khz': (_I64, []),
    'dqh_memcpy_h2d_async': (ctypes.c_int, [_P, _P, _I64, _P]),
    'dqh_last_error': (ctypes.c_char_p, []),
    'dqh_device_info': (ctypes.c_int, [ctypes.c_char_p, _I64, ctypes.POINTER(_I64)]),
}

_lib = None


def lib():
    """Load ``libdqh.so`` once; raise if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f'{LIB_PATH} is missing: build the HIP extension first '
                '(python -c "import __graft_entry__ as g; g.build()")'
            )
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SYMBOLS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        if handle.dqh_sizeof_me_inst() != ME_INST_DTYPE.itemsize:
            raise RuntimeError('dqh_me_inst layout mismatch between include/dqh.h and dynamiqs_amd/_native.py')
        from .integrators.dia import DiaRhs, DiaRouchon1, DiaTsit5

        if handle.dqh_sizeof_dia_rouchon1() != ctypes.sizeof(DiaRouchon1):
            raise RuntimeError('dqh_dia_rouchon1 layout mismatch between include/dqh.h and dynamiqs_amd/integrators/dia.py')
        if handle.dqh_sizeof_dia_rhs() != ctypes.sizeof(DiaRhs):
            raise RuntimeError('dqh_dia_rhs layout mismatch between include/dqh.h and dynamiqs_amd/integrators/dia.py')
        if handle.dqh_sizeof_dia_tsit5() != ctypes.sizeof(DiaTsit5):
            raise RuntimeError('dqh_dia_tsit5 layout mismatch between include/dqh.h and dynamiqs_amd/integrators/dia.py')
        _lib = handle
    return _lib


def check(code: int, what: str) -> None:
    if code != 0:
        msg = lib().dqh_last_error().decode(errors='replace')
        raise RuntimeError(f'{what} failed with status {code}: {msg}')


def work_elems(np_: int, nL: int, npwc: int) -> int:
    return int(lib().dqh_me_work_elems(np_, nL, npwc))


def host_device_pointer(host_ptr: int) -> int:
    """Device address of pinned host memory (raises if it is not pinned and mapped)."""
    out = ctypes.c_void_p()
    check(lib().dqh_host_device_pointer(ctypes.c_void_p(host_ptr), ctypes.byref(out)), 'dqh_host_device_pointer')
    return int(out.value)


def current_stream_ptr():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
