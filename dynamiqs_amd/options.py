"""Solver options, mirroring ``dynamiqs/options.py:17-123``."""

from __future__ import annotations


class Options:
    _defaults = dict(
        save_states=True,
        save_propagators=True,
        cartesian_batching=True,
        progress_meter=None,
        t0=None,
        save_extra=None,
        nmaxclick=10_000,
        vectorized=False,
        assume_hermitian=True,
    )

    def __init__(
        self,
        save_states: bool = True,
        save_propagators: bool = True,
        cartesian_batching: bool = True,
        progress_meter=None,
        t0=None,
        save_extra=None,
        nmaxclick: int = 10_000,
        vectorized: bool = False,
        assume_hermitian: bool = True,
    ):
        self.save_states = save_states
        self.save_propagators = save_propagators
        self.cartesian_batching = cartesian_batching
        self.progress_meter = progress_meter
        self.t0 = t0
        self.save_extra = save_extra
        self.nmaxclick = nmaxclick
        self.vectorized = vectorized
        self.assume_hermitian = assume_hermitian

    def initialise(self) -> Options:
        return Options(**self.__dict__)

    def __str__(self) -> str:
        inner = ', '.join(f'{k}={v}' for k, v in self.__dict__.items())
        return f'Options({inner})'

    def __eq__(self, other) -> bool:
        return isinstance(other, Options) and self.__dict__ == other.__dict__

    def __hash__(self) -> int:
        return hash(tuple((k, id(v) if callable(v) else v) for k, v in self.__dict__.items()))


_SUPPORTED = {
    'sesolve': ('save_states', 'cartesian_batching', 'progress_meter', 't0', 'save_extra'),
    'mesolve': (
        'save_states',
        'cartesian_batching',
        'progress_meter',
        't0',
        'save_extra',
        'vectorized',
        'assume_hermitian',
    ),
    'sepropagator': ('save_propagators', 'progress_meter', 't0', 'save_extra'),
    'mepropagator': ('save_propagators', 'cartesian_batching', 't0', 'save_extra'),
}


def check_options(options: Options, solver_name: str) -> None:
    """``options.py:80-123``: options outside the solver's whitelist must keep their defaults."""
    valid_options = _SUPPORTED[solver_name]
    for key, value in options.__dict__.items():
        if key not in valid_options and value != Options._defaults[key]:
            valid_options_str = ', '.join(f'`{x}`' for x in valid_options)
            raise ValueError(
                f'Option `{key}` was set to `{value}` but is not used by '
                f'the solver `dq.{solver_name}()` (valid options: '
                f'{valid_options_str}).'
            )
