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


def _cells_per_grid_cell(expr: ast.AST) -> int:
    """A field extent written as products of nx, ny and integers, divided by
    nx*ny (arithmetic read from the tree, nothing evaluated)."""
    if isinstance(expr, ast.Name) and expr.id in ("nx", "ny"):
        return 1
    if isinstance(expr, ast.Attribute) and ast.unparse(expr) in ("self.nx", "self.ny"):
        return 1
    if isinstance(expr, ast.Constant) and isinstance(expr.value, int):
        return expr.value
    if isinstance(expr, ast.BinOp) and isinstance(expr.op, ast.Mult):
        return _cells_per_grid_cell(expr.left) * _cells_per_grid_cell(expr.right)
    raise ValueError(f"unexpected field extent {ast.unparse(expr)}")


def class_fields(path: str, cname: str) -> dict:
    """Fields the class constructor creates: {name: [taichi dtype, elements
    per grid cell, or 0 for a 0-d field]} from ti.field(...) and
    fb.dense(ti.i, extent).place(self.name)."""
    tree = ast.parse(open(path).read())
    cls = next(n for n in tree.body if isinstance(n, ast.ClassDef) and n.name == cname)
    init = next(m for m in cls.body if isinstance(m, ast.FunctionDef) and m.name == "__init__")
    fields = {}
    for node in ast.walk(init):
        if (isinstance(node, ast.Assign) and len(node.targets) == 1 and isinstance(node.targets[0], ast.Attribute)
                and isinstance(node.value, ast.Call) and ast.unparse(node.value.func) == "ti.field"):
            shape0 = any(k.arg == "shape" for k in node.value.keywords)
            fields[node.targets[0].attr] = [ast.unparse(node.value.args[0]), 0 if shape0 else None]
    for node in ast.walk(init):
        if (isinstance(node, ast.Call) and isinstance(node.func, ast.Attribute) and node.func.attr == "place"
                and isinstance(node.func.value, ast.Call) and isinstance(node.func.value.func, ast.Attribute)
                and node.func.value.func.attr == "dense"):
            per = _cells_per_grid_cell(node.func.value.args[1])
            for a in node.args:
                if isinstance(a, ast.Attribute) and a.attr in fields:
                    fields[a.attr][1] = per
    return fields


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
    for pkg, fn, cname in (("flow", "flowfields.py", "FlowRouter"), ("graphflood", "gf_fields.py", "Flooder")):
        out[pkg]["fields"] = {cname: class_fields(os.path.join(root, "pyfastflow", pkg, fn), cname)}
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
    api = surface(root)
    with open(os.path.join(HERE, "reference_api.json"), "w") as f:
        json.dump(api, f, indent=1)
        f.write("\n")
