from .mesolve import mesolve
from .sesolve import sesolve

__all__ = ['mesolve', 'sesolve']
