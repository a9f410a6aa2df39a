"""Write tests/golden/surface/setup_py_names.json: every softflow_core name the
reference's Python front end (python/pysoftflow/setup.py) uses, found by
walking its syntax tree.  Runs only where the reference tree exists; the
committed list is data (names), read by tests/test_surface_cpu.py.

    python tests/golden/surface/make_surface_names.py

Names are recorded as
    module:<Name>[.<Member>]        attributes of the module (classes, enums)
    Simulation:<method>             self._core.<method>
    <Class>:<attr>[.<attr>]         attributes read or written on an object whose
                                    class is known: built by _sc.<Class>(...) or
                                    returned by a Simulation accessor.
"""
import ast
import json
import os
import sys

SRC = "/root/reference/python/pysoftflow/setup.py"
HERE = os.path.dirname(os.path.abspath(__file__))
# classes of what the Simulation accessors return (python/bindings.cpp)
ACCESSOR = {"params": "SimulationParams", "field": "LatticeField", "capsules": "CapsuleSystem",
            "advectionDiffusion": "AdvectionDiffusion", "lbmSolver": "LBMSolver",
            "adhesion": "AdhesionModel"}


def chain(node):
    """['a', 'b', 'c'] for the expression a.b.c, else None."""
    out = []
    while isinstance(node, ast.Attribute):
        out.append(node.attr)
        node = node.value
    if isinstance(node, ast.Name):
        out.append(node.id)
        return out[::-1]
    return None


def origin(v, typed):
    """(class, attribute path) of what expression v evaluates to, when known:
    _sc.Cls(...), self._core.<accessor>(...), or attributes below either or
    below a variable already known."""
    path = []
    while isinstance(v, ast.Attribute):
        path.append(v.attr)
        v = v.value
    path = path[::-1]
    if isinstance(v, ast.Call):
        c = chain(v.func)
        if c and c[0] == "_sc" and len(c) == 2:
            return c[1], path
        if c and c[:2] == ["self", "_core"] and len(c) == 3 and c[2] in ACCESSOR:
            return ACCESSOR[c[2]], path
        return None
    if isinstance(v, ast.Name) and v.id in typed and path:
        cls, base = typed[v.id]
        return cls, base + path
    return None


def main():
    if not os.path.exists(SRC):
        sys.exit(f"{SRC} is absent")
    tree = ast.parse(open(SRC).read())
    names = set()
    typed = {}   # variable name -> class (flow-insensitive, per module)
    for _ in range(3):           # variables typed from other typed variables
        for node in ast.walk(tree):
            if isinstance(node, ast.Assign) and len(node.targets) == 1 and isinstance(node.targets[0], ast.Name):
                o = origin(node.value, typed)
                if o:
                    typed[node.targets[0].id] = o
    for node in ast.walk(tree):
        if not isinstance(node, ast.Attribute):
            continue
        c = chain(node)
        if c:
            if c[0] == "_sc" and len(c) >= 2:
                names.add("module:" + ".".join(c[1:3]))
            elif c[:2] == ["self", "_core"] and len(c) >= 3:
                names.add("Simulation:" + c[2])
            elif c[0] in typed and len(c) >= 2:
                cls, base = typed[c[0]]
                names.add(cls + ":" + ".".join(base + c[1:]))
            continue
        o = origin(node, typed)   # attribute chains below an accessor call
        if o and o[1]:
            names.add(o[0] + ":" + ".".join(o[1]))
    # drop prefixes of longer chains (a.b is implied by a.b.c)
    keep = sorted(n for n in names if not any(m != n and m.startswith(n + ".") for m in names))
    with open(os.path.join(HERE, "setup_py_names.json"), "w") as fh:
        json.dump({"source": "python/pysoftflow/setup.py", "names": keep}, fh, indent=1)
    print(len(keep), "names")


if __name__ == "__main__":
    main()
