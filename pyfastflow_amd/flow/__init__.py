"""pyfastflow_amd.flow -- the pyfastflow.flow call surface
(pyfastflow/flow/__init__.py:26-35) on gfx950 HIP kernels."""

from ..constants import *  # noqa: F401,F403
from .neighbourer_flat import *  # noqa: F401,F403
from .receivers import *  # noqa: F401,F403
from .downstream_propag import *  # noqa: F401,F403
from .lakeflow import *  # noqa: F401,F403
from .f32_i32_struct import *  # noqa: F401,F403
from .util_taichi import *  # noqa: F401,F403
from .flowfields import *  # noqa: F401,F403
from .fill_topo import *  # noqa: F401,F403
from .environment import *  # noqa: F401,F403
from .batch import *  # noqa: F401,F403
from . import neighbourer_flat, receivers, downstream_propag, lakeflow, f32_i32_struct, util_taichi  # noqa: F401
from . import flowfields, fill_topo, environment, batch  # noqa: F401
from .. import constants  # noqa: F401

__all__ = [
    "neighbourer_flat", "receivers", "downstream_propag", "lakeflow", "f32_i32_struct", "util_taichi",
    "flowfields", "fill_topo", "environment", "batch",
]
