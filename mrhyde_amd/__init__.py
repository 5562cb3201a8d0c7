"""mhx — MI355X-native element assembly and Newton/Krylov solve engine.

`mrhyde_amd.abi` (constants, ctypes structures, the Problem description) loads
without the GPU library.  Everything else lives in libmhx.so; accessing
`Mhx`, `krylov_opts`, `rng` or `MhxError` loads it and raises if it is absent.
"""
from .abi import Problem, rng_fill  # noqa: F401

_LAZY = {"Mhx", "MhxError", "krylov_opts", "rng", "SubgridBatch", "Server", "function_eval"}


def __getattr__(name):
    if name in _LAZY:
        from . import engine
        return getattr(engine, name)
    raise AttributeError(name)
