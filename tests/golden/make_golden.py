"""Write the golden fixtures used by the parity tests.

Run once in a container where the reference checkout is mounted read-only at
/root/reference.  Only data travels: the final displacement vectors ``U.ref``
that the reference's own regression tests compare against
(testfiles/<Name>/U.ref, compared with max|dU| < 1e-6 by
edelweissfe/_cli/_run_tests_edelweissfe.py:95-106), and the node/element tables of
the mesh-file test CantileverBeamHexa8 (testfiles/CantileverBeamHexa8/test.inp).

Nothing of the reference is imported or executed.
"""

import os
import re

import numpy as np

REF = "/root/reference/testfiles"
HERE = os.path.dirname(os.path.abspath(__file__))

UREF_TESTS = [
    "TensionBarHexa8",
    "SimpleBeamHexa8",
    "WallShearHexa8",
    "WallShearHexa20",
    "CantileverBeamHexa8",
    "CantileverBeamVonMisesPE",  # plane strain CPE8, von Mises with nonlinear hardening
    "WallShearHexa8GMRES",  # C3D8N (an alias of C3D8, elements/library.py:541) solved by GMRES
]


def _parse_abaqus_mesh(path):
    """Extract *Node / *Element / *nSet / *elSet data lines of an input deck."""
    nodes, elements, nsets = [], [], {}
    mode, cur = None, None
    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if not line or line.startswith("**"):
                continue
            if line.startswith("*"):
                key = line.lower()
                mode, cur = None, None
                if key.startswith("*node"):
                    mode = "node"
                elif key.startswith("*element"):
                    mode = "element"
                elif key.startswith("*nset"):
                    mode = "nset"
                    cur = re.search(r"nset\s*=\s*([^,\s]+)", line, re.I).group(1)
                    nsets[cur] = {"generate": "generate" in key, "vals": []}
                continue
            vals = [v for v in line.replace(",", " ").split() if v]
            if mode == "node":
                nodes.append([float(v) for v in vals])
            elif mode == "element":
                elements.append([int(v) for v in vals])
            elif mode == "nset":
                nsets[cur]["vals"].extend(int(v) for v in vals)
    out = {}
    for name, d in nsets.items():
        v = d["vals"]
        if d["generate"]:
            v = list(range(v[0], v[1] + 1, v[2] if len(v) > 2 else 1))
        out[name] = np.asarray(v, dtype=np.int64)
    return np.asarray(nodes), np.asarray(elements, dtype=np.int64), out


def main():
    arrays = {}
    for t in UREF_TESTS:
        arrays["uref_" + t] = np.loadtxt(os.path.join(REF, t, "U.ref"))
    nodes, elements, nsets = _parse_abaqus_mesh(os.path.join(REF, "CantileverBeamHexa8", "test.inp"))
    arrays["cantilever_nodes"] = nodes
    arrays["cantilever_elements"] = elements
    for k, v in nsets.items():
        arrays["cantilever_nset_" + k] = v
    np.savez_compressed(os.path.join(HERE, "reference_golden.npz"), **arrays)
    for k, v in arrays.items():
        print(k, v.shape)


if __name__ == "__main__":
    main()
