"""Gradient selectors, mirroring ``dynamiqs/gradient.py:10-110``.

They are accepted and validated against the method as in the reference
(``method.assert_supports_gradient``), but the solvers of this package are not
differentiated through an autodiff framework (there is no JAX): passing one to
``mesolve`` / ``sesolve`` / the propagators raises ``NotImplementedError`` after
that validation. Gradients come from the hand-written continuous adjoint,
``dynamiqs_amd.adjoint.mesolve_adjoint`` and ``AdjointPlan``.
"""

from __future__ import annotations


def reject_gradient(gradient, solver: str) -> None:
    """Raise for a gradient request the solvers cannot serve (after the reference's
    method / gradient compatibility check)."""
    if gradient is not None:
        raise NotImplementedError(
            f'`gradient={gradient}` asks {solver} for autodiff gradients, which are not available here (no JAX): '
            'use dynamiqs_amd.adjoint.mesolve_adjoint or dynamiqs_amd.adjoint.AdjointPlan for adjoint gradients '
            'of mesolve losses.'
        )


class Gradient:
    def __str__(self) -> str:
        return f'{type(self).__name__}()'

    def __eq__(self, other) -> bool:
        return type(self) is type(other) and self.__dict__ == other.__dict__

    def __hash__(self) -> int:
        return hash((type(self), tuple(sorted(self.__dict__.items()))))


class Direct(Gradient):
    pass


class BackwardCheckpointed(Gradient):
    def __init__(self, ncheckpoints: int | None = None):
        self.ncheckpoints = ncheckpoints

    def __str__(self) -> str:
        return f'BackwardCheckpointed(ncheckpoints={self.ncheckpoints})'


class Forward(Gradient):
    pass
