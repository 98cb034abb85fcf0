"""atmmetaforce -- MI355X-native ATMMetaForce.

Mirrors the Python surface of the reference's SWIG module `atmmetaforce`
(python/atmmetaforceplugin.i): ATMMetaForce (constructor, particles, the nine
global-parameter names as static methods, default-value getters,
getPerturbationEnergy, updateParametersInContext, cast/isinstance) and
ATMMETAFORCE_VERSION.  Values are plain floats in OpenMM's units (nm, kJ/mol);
openmm.unit is not installed, so no Quantity wrapping is applied.

`atmmetaforce.openmm` holds the minimal OpenMM-compatible classes the force is
used with (System, NonbondedForce, Context, State, Platform, VerletIntegrator,
XmlSerializer); the only platform is "MI355X".

ATMMetaForceUtils (python/ATMMetaForceUtils.py's setup helpers: centroid and
position restraints, non-bonded force group, zero-LJ fix) builds
CentroidRestraintForce / PositionRestraintForce, evaluated on the MI355X.

Beyond the reference: ATMBatch (many systems/replicas per launch through the C
ABI), ATMEnergyFunction (torch.autograd), and the synthetic workloads.
"""
ATMMETAFORCE_VERSION = "0.3.1"

# An installed torch brings its own HIP runtime, with the same library names
# as the /opt/rocm one this package's native libraries link.  The first one
# loaded serves the whole process, and torch sees no GPU when it is not its
# own; so torch (when present) is loaded before the native libraries are.
try:
    import torch as _torch  # noqa: F401
except ImportError:
    pass

from . import workloads  # noqa: E402
from .batch import ATMBatch  # noqa: E402

try:
    from ._atmmetaforce import (ATMMetaForce, CentroidRestraintForce, OpenMMException,  # noqa: E402,F401
                                PositionRestraintForce)
    from . import openmm  # noqa: E402,F401
    from .utils import ATMMetaForceUtils  # noqa: E402,F401
except ImportError as _e:  # the C++ layer is not built: fail when it is used
    _import_error = _e
else:
    _import_error = None


def __getattr__(name):
    if name in ("ATMEnergyFunction", "atm_energy"):  # imports torch on first use
        from . import autograd
        return getattr(autograd, name)
    if name in ("ATMMetaForce", "OpenMMException", "openmm", "ATMMetaForceUtils", "CentroidRestraintForce",
                "PositionRestraintForce") and _import_error is not None:
        raise ImportError(f"atmmetaforce C++ module not built ({_import_error}); run __graft_entry__.build()")
    raise AttributeError(name)

__all__ = ["ATMMETAFORCE_VERSION", "ATMMetaForce", "ATMMetaForceUtils", "ATMBatch", "ATMEnergyFunction",
           "CentroidRestraintForce", "OpenMMException", "PositionRestraintForce", "atm_energy", "workloads"]
