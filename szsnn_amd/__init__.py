"""szsnn_amd -- MI355X-native SZSNN simulation core (pyszsnn host interface).

The compute path is the HIP library libszsnn_hip.so (gfx950), reached through
the C ABI in include/szsnn_hip.h.  Importing the simulation classes requires
the library to be built (python -m szsnn_amd.build); nothing here falls back
to a CPU path.
"""
from ._lib import (MODEL_IF_COND_EXP, MODEL_IF_CURR_EXP, MODEL_IZHIKEVICH, SOURCE_CONST_RATE,
                   SOURCE_NONE, SOURCE_POISSON, SzsnnError)
from .network import Network

__all__ = ["Network", "SzsnnError", "MODEL_IF_CURR_EXP", "MODEL_IZHIKEVICH", "MODEL_IF_COND_EXP",
           "SOURCE_NONE", "SOURCE_CONST_RATE", "SOURCE_POISSON", "Context", "Simulation"]


def __getattr__(name):
    if name in ("Context", "Simulation"):
        from . import simulation
        return getattr(simulation, name)
    raise AttributeError(name)
