"""pyfastflow_amd.graphflood -- the pyfastflow.graphflood call surface
(pyfastflow/graphflood/__init__.py:35-49)."""

from .gf_fields import *  # noqa: F401,F403
from .gf_hydrodynamics import *  # noqa: F401,F403
from . import gf_fields, gf_hydrodynamics  # noqa: F401

__all__ = ["Flooder", "diffuse_Q_constant_prec", "graphflood_core_cte_mannings", "gf_fields", "gf_hydrodynamics"]
