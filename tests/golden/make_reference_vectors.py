"""Extract the reference's pair-potential test vectors into JSON fixtures.

Reads the reference test modules as text and walks their syntax trees with
``ast`` (nothing from the reference is imported or executed), collecting every
``PotentialTestCase(...)`` literal:

* ``src/pytest/test_pair.py``        -> ``pair_vectors.json``  (26 cases)
* ``src/pytest/test_pair_aniso.py``  -> ``aniso_vectors.json`` (5 cases)

Usage: python tests/golden/make_reference_vectors.py [/root/reference]
"""

import ast
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _literal(node):
    if isinstance(node, ast.Attribute):
        # hoomd.azplugins.pair.<Name>
        return node.attr
    if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Mult):
        # numeric products written in the test, e.g. -0.20567 * 2
        return _literal(node.left) * _literal(node.right)
    return ast.literal_eval(node)


def extract(path, fields):
    with open(path) as fh:
        tree = ast.parse(fh.read(), filename=path)
    cases = []
    for node in ast.walk(tree):
        if (
            isinstance(node, ast.Call)
            and isinstance(node.func, ast.Name)
            and node.func.id == "PotentialTestCase"
        ):
            vals = [_literal(a) for a in node.args]
            case = dict(zip(fields, vals))
            case["source_line"] = node.lineno
            cases.append(case)
    cases.sort(key=lambda c: c["source_line"])
    return cases


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    pair = extract(
        os.path.join(ref, "src/pytest/test_pair.py"),
        ["potential", "params", "r_cut", "shift", "distance", "energy", "force"],
    )
    aniso = extract(
        os.path.join(ref, "src/pytest/test_pair_aniso.py"),
        ["potential", "params", "r_cut", "shift", "energy", "force", "torque"],
    )
    meta = {
        "source": "azplugins v1.3.0 src/pytest",
        "tolerance": "numpy.testing.assert_array_almost_equal(decimal=4)",
    }
    with open(os.path.join(HERE, "pair_vectors.json"), "w") as fh:
        json.dump({"meta": meta | {"file": "src/pytest/test_pair.py"}, "cases": pair}, fh, indent=1)
    with open(os.path.join(HERE, "aniso_vectors.json"), "w") as fh:
        json.dump(
            {
                "meta": meta
                | {
                    "file": "src/pytest/test_pair_aniso.py",
                    "positions": [[-0.5, -0.10, -0.15], [0.5, 0.10, 0.15]],
                    "orientations": [[1, 0, 0, 0], [1, 0, 0, 0]],
                    "r_buff": 0.4,
                },
                "cases": aniso,
            },
            fh,
            indent=1,
        )
    print(f"{len(pair)} pair cases, {len(aniso)} aniso cases")


if __name__ == "__main__":
    main()
