"""Writes tests/golden/reference_api.json: the Python-callable surface of
pyfastflow.flow and pyfastflow.graphflood, and the names pyfastflow.constants
defines, read from the reference's source
text with the ast module (nothing of the reference is imported or run).

    python tests/golden/gen_reference_api.py [/root/reference]

Callable from Python in the reference: @ti.kernel functions, undecorated
functions, classes and their public methods, each with its parameter
names in order.  @ti.func functions are device
helpers that Taichi refuses to call from Python scope, so they are listed
apart, as is the boundary-dispatch code that compile_neighbourer rebinds."""

from __future__ import annotations

import ast
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def surface(root: str) -> dict:
    out = {}
    for pkg in ("flow", "graphflood"):
        d = os.path.join(root, "pyfastflow", pkg)
        kernels, functions, device, classes = [], [], [], {}
        for fn in sorted(os.listdir(d)):
            if not fn.endswith(".py") or fn == "__init__.py":
                continue
            tree = ast.parse(open(os.path.join(d, fn)).read())
            for n in tree.body:
                if isinstance(n, ast.FunctionDef) and not n.name.startswith("_"):
                    decs = [ast.unparse(x) for x in n.decorator_list]
                    where = f"{pkg}/{fn}:{n.lineno}"
                    args = [a.arg for a in n.args.args]
                    if "ti.kernel" in decs:
                        kernels.append([n.name, where, args])
                    elif "ti.func" in decs:
                        device.append([n.name, where, args])
                    else:
                        functions.append([n.name, where, args])
                elif isinstance(n, ast.ClassDef):
                    classes[n.name] = {m.name: [a.arg for a in m.args.args] for m in n.body
                                       if isinstance(m, ast.FunctionDef)
                                       and (not m.name.startswith("_") or m.name == "__init__")}
        out[pkg] = {"kernels": kernels, "functions": functions, "device_only": device, "classes": classes}
    # the module globals and functions of pyfastflow.constants
    tree = ast.parse(open(os.path.join(root, "pyfastflow", "constants.py")).read())
    consts = []
    for n in tree.body:
        if isinstance(n, ast.Assign):
            consts += [t.id for t in n.targets if isinstance(t, ast.Name) and not t.id.startswith("_")]
        elif isinstance(n, ast.FunctionDef) and not n.name.startswith("_"):
            consts.append(n.name)
    out["constants"] = consts
    return out


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(HERE, "reference_api.json"), "w") as f:
        json.dump(surface(root), f, indent=1)
        f.write("\n")
