"""atmmetaforce -- MI355X-native ATMMetaForce.

Mirrors the Python surface of the reference's SWIG module `atmmetaforce`
(python/atmmetaforceplugin.i): ATMMetaForce, ATMMETAFORCE_VERSION.  The batched
evaluation API (ATMBatch) and the autograd entry point (ATMEnergyFunction) have
no reference counterpart.
"""
ATMMETAFORCE_VERSION = "0.3.1"

from . import workloads  # noqa: E402
from .batch import ATMBatch  # noqa: E402

__all__ = ["ATMMETAFORCE_VERSION", "ATMBatch", "workloads"]
