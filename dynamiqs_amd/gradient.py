"""Gradient selectors, mirroring ``dynamiqs/gradient.py:10-110``.

On the GPU path a gradient request is served by the hand-written adjoint
(``dynamiqs_amd.adjoint``) instead of diffrax's checkpointed reverse mode.
"""

from __future__ import annotations


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
