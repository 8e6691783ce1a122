"""Writes rational_tables.json from the reference's approximations.py.

The reference file is parsed as text with ``ast`` (literal lists only); it is
never imported or executed.  Run in the build container, where the
reference is available:

    python tests/golden/extract_tables.py /root/reference/approximations.py
"""
import ast
import json
import os
import sys


def literal_list(node):
    """np.array([...]) -> the literal list argument."""
    if isinstance(node, ast.Call) and node.args:
        return ast.literal_eval(node.args[0])
    return ast.literal_eval(node)


def main(path):
    tree = ast.parse(open(path).read())
    out = {"source": "approximations.py", "rational": {}, "square_root_polynomial": {}}
    for node in tree.body:
        if isinstance(node, ast.Assign) and len(node.targets) == 1:
            name = node.targets[0].id
            if name.endswith("_rational_approximation_coefficients"):
                out["rational"][name] = [float(v) for v in literal_list(node.value)]
        if isinstance(node, ast.FunctionDef) and node.name == "_get_square_root_polynomial_approximation_coefficients":
            for stmt in ast.walk(node):
                if isinstance(stmt, ast.If):
                    n = ast.literal_eval(stmt.test.comparators[0])
                    ret = stmt.body[0]
                    pairs = literal_list(ret.value)
                    out["square_root_polynomial"][str(n)] = [[float(a), float(b)] for a, b in pairs]
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "tfleonardym_amd", "data",
                       "rational_tables.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print(dst, {k: len(v) for k, v in out["rational"].items()}, list(out["square_root_polynomial"]))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/approximations.py")
