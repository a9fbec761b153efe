"""Integration methods, mirroring ``dynamiqs/method.py``.

``Tsit5`` (:438-476, defaults of ``_DEAdaptiveStep`` :121-127), ``Dopri5``,
``Dopri8`` (mesolve), ``Euler``, ``Rouchon1`` (:207-237), ``Rouchon2`` / ``Rouchon3``
(fixed step with ``dt``, the adaptive pairs without it) and ``Expm`` run on the GPU. The other method names of the reference exist so that the
solvers can raise the reference's ``TypeError`` for the methods they do not
support.
"""

from __future__ import annotations

from .gradient import BackwardCheckpointed, Direct, Forward, Gradient


class Method:
    SUPPORTED_GRADIENT: tuple = ()

    @classmethod
    def supports_gradient(cls, gradient: Gradient | None) -> bool:
        return isinstance(gradient, cls.SUPPORTED_GRADIENT)

    @classmethod
    def assert_supports_gradient(cls, gradient: Gradient | None) -> None:
        if gradient is not None and not cls.supports_gradient(gradient):
            support_str = ', '.join(f'`{x.__name__}`' for x in cls.SUPPORTED_GRADIENT)
            raise ValueError(
                f'Method `{cls.__name__}` does not support gradient'
                f' `{type(gradient).__name__}` (supported gradient types:'
                f' {support_str}).'
            )

    def _fields(self) -> dict:
        return dict(self.__dict__)

    def __str__(self) -> str:
        inner = ', '.join(f'{k}={v}' for k, v in self._fields().items())
        return f'{type(self).__name__}({inner})'

    __repr__ = __str__

    def __eq__(self, other) -> bool:
        return type(self) is type(other) and self.__dict__ == other.__dict__

    def __hash__(self) -> int:
        return hash((type(self), tuple(sorted(self.__dict__.items()))))


class _DEAdaptiveStep(Method):
    max_steps = 100_000

    def __init__(
        self,
        rtol: float = 1e-6,
        atol: float = 1e-6,
        safety_factor: float = 0.9,
        min_factor: float = 0.2,
        max_factor: float = 5.0,
        max_steps: int = 100_000,
    ):
        self.rtol = rtol
        self.atol = atol
        self.safety_factor = safety_factor
        self.min_factor = min_factor
        self.max_factor = max_factor
        self.max_steps = max_steps


class Tsit5(_DEAdaptiveStep):
    """Tsitouras 5(4) adaptive ODE method (FSAL, PID controller)."""

    SUPPORTED_GRADIENT = (Direct, BackwardCheckpointed, Forward)


class Dopri5(_DEAdaptiveStep):
    """Dormand-Prince 5(4) adaptive ODE method (FSAL, PID controller): the Tsit5
    kernels with the Dopri5 tableau and dense output (``DQH_FLAG_DOPRI5``)."""

    SUPPORTED_GRADIENT = (Direct, BackwardCheckpointed, Forward)


class Dopri8(_DEAdaptiveStep):
    """Dormand-Prince 8(7) adaptive ODE method (Prince & Dormand 1981, RK8(7)13M):
    13 stages on the dense global-memory kernel (``DQH_FLAG_DOPRI8``), the same
    controller with exponent 1/8, steps landing on the save times (the saved states
    are step ends, not a continuous extension)."""

    SUPPORTED_GRADIENT = (Direct, BackwardCheckpointed, Forward)


class Kvaerno3(_DEAdaptiveStep):
    SUPPORTED_GRADIENT = (Direct, BackwardCheckpointed, Forward)


class Kvaerno5(_DEAdaptiveStep):
    SUPPORTED_GRADIENT = (Direct, BackwardCheckpointed, Forward)


class Euler(Method):
    SUPPORTED_GRADIENT = (Direct, BackwardCheckpointed, Forward)

    def __init__(self, dt: float):
        self.dt = dt


class Rouchon1(Method):
    """First-order Rouchon method (fixed step, Kraus map with Cholesky normalisation)."""

    SUPPORTED_GRADIENT = (Direct, BackwardCheckpointed, Forward)

    def __init__(self, dt: float, normalize: bool = True):
        self.dt = dt
        self.normalize = normalize


class _RouchonHigherOrder(Method):
    """Fields of ``Rouchon2`` / ``Rouchon3`` (``dynamiqs/method.py:240-360``)."""

    SUPPORTED_GRADIENT = (Direct, BackwardCheckpointed, Forward)

    def __init__(
        self,
        rtol: float = 1e-6,
        atol: float = 1e-6,
        safety_factor: float = 0.9,
        min_factor: float = 0.2,
        max_factor: float = 5.0,
        max_steps: int = 100_000,
        dt: float | None = None,
        normalize: bool = True,
    ):
        self.rtol = rtol
        self.atol = atol
        self.safety_factor = safety_factor
        self.min_factor = min_factor
        self.max_factor = max_factor
        self.max_steps = max_steps
        self.dt = dt
        self.normalize = normalize


class Rouchon2(_RouchonHigherOrder):
    """Second-order Rouchon method. With ``dt`` it is the fixed-step Kraus-Heun map
    (run on the dense Rouchon kernel with its Kraus operators folded on the host);
    the adaptive form (``dt=None``) is not supported by the MI355X solver."""


class Rouchon3(_RouchonHigherOrder):
    """Third-order Rouchon method. With ``dt`` it is the fixed-step Kraus-Heun3 map
    (1 + 2K + K^2 + K^3 Kraus operators folded on the host, on the dense Rouchon
    kernel); the adaptive form (``dt=None``) is not supported by the MI355X solver."""


class Expm(Method):
    SUPPORTED_GRADIENT = (Direct,)

    def __init__(self):
        pass
