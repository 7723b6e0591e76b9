"""thewalrus_amd -- MI355X-native power-trace hafnian and loop hafnian.

Drop-in replacements for ``thewalrus.hafnian``, ``thewalrus.hafnian_repeated``,
``thewalrus.mtl`` / ``lmtl`` (and the helpers they are built from), the batched loop hafnians of
``thewalrus.loop_hafnian_batch`` / ``loop_hafnian_batch_gamma`` and the Gaussian-state
probability requests of ``thewalrus.quantum`` (``thewalrus_amd.quantum``), whose sieves run in
hand-written HIP kernels for gfx950, reached through the C ABI of ``libwalrus_hip.so``.
"""
from ._hafnian import (
    hafnian,
    hafnian_repeated,
    loop_hafnian,
    input_validation,
    matched_reps,
    reduction,
    _haf,
)
from ._montrealer import mtl, lmtl

__version__ = "0.1.0"


def version():
    """Version string of this package."""
    return __version__


__all__ = [
    "hafnian",
    "hafnian_repeated",
    "loop_hafnian",
    "input_validation",
    "matched_reps",
    "reduction",
    "mtl",
    "lmtl",
    "version",
]
