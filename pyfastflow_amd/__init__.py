"""pyfastflow_amd -- MI355X (gfx950) implementation of PyFastFlow's flow
routing and GraphFlood hot path.

Drop-in surface: ``pyfastflow_amd.flow`` and ``pyfastflow_amd.graphflood``
mirror ``pyfastflow.flow`` / ``pyfastflow.graphflood``; the compute runs in
hand-written HIP kernels of libpff_hip.so (C ABI: include/pff.h), reached
through ctypes with torch tensors as device storage.

    import pyfastflow_amd as pf
    router = pf.flow.FlowRouter(nx, ny, dx, boundary_mode="periodic_EW")
    router.set_z(dem)
    router.compute_receivers()
    router.reroute_flow()
    router.accumulate_constant_Q(1e-5, area=True)
    Q = router.get_Q()
"""

__version__ = "0.1.0"

from . import constants  # noqa: F401
from . import flow  # noqa: F401
from . import graphflood  # noqa: F401

__all__ = ["constants", "flow", "graphflood"]
