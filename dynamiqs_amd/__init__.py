"""dynamiqs_amd -- MI355X-native (gfx950) drop-in for ``dynamiqs.mesolve``.

``import dynamiqs_amd as dq`` exposes the subset of the dynamiqs API that the
``mesolve`` hot path uses: ``dq.mesolve``, ``dq.sesolve`` and the propagators
``dq.sepropagator`` / ``dq.mepropagator``, quantum arrays
(dense and sparse DIA), time-dependent operators (``constant``, ``pwc``,
``modulated``, ``timecallable``), operator/state builders, ``Options``, the
methods ``Tsit5`` / ``Dopri5`` / ``Dopri8`` / ``Euler`` / ``Rouchon1`` / ``Rouchon2`` /
``Rouchon3`` (fixed step, or adaptive without ``dt``) / ``Expm`` and the result classes.
The solve runs in hand-written HIP kernels (``libdqh.so``, C-ABI in
``include/dqh.h``) loaded with ctypes; there is no CPU fallback.
"""

from . import gradient, method
from .integrators import mepropagator, mesolve, sepropagator, sesolve
from .options import Options
from .qarrays import QArray, asqarray, dense, dia, set_layout, set_precision
from .result import MEPropagatorResult, MESolveResult, SEPropagatorResult, SESolveResult
from .time_qarray import TimeQArray, constant, modulated, pwc, timecallable
from .utils import (
    coherent,
    coherent_dm,
    create,
    dag,
    destroy,
    displace,
    dissipator,
    excited,
    excited_dm,
    expect,
    eye,
    fock,
    fock_dm,
    ground,
    ground_dm,
    isket,
    lindbladian,
    momentum,
    number,
    position,
    sigmam,
    sigmap,
    sigmax,
    sigmay,
    sigmaz,
    tensor,
    todm,
    tracemm,
)

__version__ = '0.1.0'
