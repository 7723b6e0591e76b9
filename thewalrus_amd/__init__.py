"""thewalrus_amd -- MI355X-native power-trace hafnian and loop hafnian.

Drop-in replacements for ``thewalrus.hafnian`` and ``thewalrus.hafnian_repeated`` (and the
helpers they are built from) whose sieve runs in hand-written HIP kernels for gfx950,
reached through the C ABI of ``libwalrus_hip.so``.
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
    "version",
]
