"""Loader of tests/golden/routing_cases.npz (written by
tests/golden/gen_fixtures.py)."""

import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "routing_cases.npz")
MODE_NAMES = {0: "normal", 1: "periodic_EW", 2: "periodic_NS"}


def load():
    d = np.load(PATH)  # plain arrays only (no pickles)
    names = sorted({k.split("/")[0] for k in d.files})
    cases = {}
    for n in names:
        c = {k.split("/", 1)[1]: d[k] for k in d.files if k.startswith(n + "/")}
        ny, nx, mode = (int(v) for v in c["meta"])
        c.update(ny=ny, nx=nx, mode=mode, mode_name=MODE_NAMES[mode], dx=float(c["dx"][0]))
        off = np.concatenate([[0], np.cumsum(c["trip_len"])]).astype(np.int64)
        c["trips"] = [(c["trip_basin"][off[t]:off[t + 1]], c["trip_outlet"][off[t]:off[t + 1]],
                       c["trip_snode"][off[t]:off[t + 1]]) for t in range(len(c["trip_len"]))]
        cases[n] = c
    return cases
