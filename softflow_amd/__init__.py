"""SoftFlow per-timestep hot path on MI355X (gfx950).

``softflow_core`` is the pybind engine (the reference's Simulation API) over
``lib/libsoftflow_hip.so``, under the module name the reference's own Python
front end imports; ``capi`` binds the C-ABI directly; ``ensemble`` builds
channel ensembles for the batched stepper (``softflow_core.Ensemble``).
Importing the package fails if the native libraries have not been built
(``python -m softflow_amd.build``).
"""
from . import softflow_core  # noqa: F401  (fails loudly when not built)
from . import capi  # noqa: F401

__all__ = ["softflow_core", "capi"]
