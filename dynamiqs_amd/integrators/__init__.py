from .mesolve import mesolve
from .propagators import mepropagator, sepropagator
from .sesolve import sesolve

__all__ = ['mesolve', 'mepropagator', 'sepropagator', 'sesolve']
