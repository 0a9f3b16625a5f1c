"""initialise / reboot (pyfastflow/flow/environment.py).

The reference compiles its boundary dispatch once per process and refuses a
second initialise (environment.py:28-29).  Here the boundary mode is a
runtime argument of every kernel, so FlowRouter does not depend on this
flag; the functions are kept for callers that use them.
"""

from __future__ import annotations

from .. import constants as cte

__all__ = ["initialise", "reboot"]


def initialise():
    if cte.INITIALISED:
        raise RuntimeError("pyfastflow_amd already initialised")
    cte.INITIALISED = True


def reboot():
    import torch

    if torch.cuda.is_available():
        torch.cuda.synchronize()
    cte.INITIALISED = False
