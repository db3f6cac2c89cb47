"""MI355X-native Hawkes likelihood engine (drop-in for hpHawkes' AbstractHawkes engine).

The compute path is libhph.so: hand-written HIP kernels for gfx950 behind the C ABI of
include/hph.h. This package holds its Python mirror (``engine``), the synthetic workload
generators (``synth``) and the torch autograd wrapper (``autograd``).
"""
from . import _lib  # noqa: F401
from .engine import (Engine, HphError, createEngine, getLocationGradient, getLogLikeGradient,  # noqa: F401
                     getLogLikelihood, getProbsSelfExcite, setParameters, setTimesData, updateLocations)

PARAMETER_NAMES = ("sigmaXprec", "tauXprec", "tauTprec", "omega", "theta", "mu0")
P0 = (1.0, 1.0 / 1.6, 1.0 / 336.0, 1.0, 1.0, 1.0)
