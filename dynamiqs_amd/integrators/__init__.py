from .mesolve import mesolve

__all__ = ['mesolve']
