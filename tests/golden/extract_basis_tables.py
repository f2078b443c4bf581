"""Extract the committed basis/quadrature constants of the reference as a JSON fixture.

Reads, as text only, the generated headers
  dynamics/src/include/codeGenerationDGinGauss.hpp
  dynamics/src/include/codeGenerationCGinGauss.hpp
of the reference checkout and writes tests/golden/basis_tables.json: for every
``template <> struct XImpl<a, b[, c]>`` the list of numbers in its initialiser,
in the header's (row-major) order. Nothing of the reference is imported or run.

Usage: python tests/golden/extract_basis_tables.py [/root/reference]
"""
import json
import os
import re
import sys

HEADERS = [
    "dynamics/src/include/codeGenerationDGinGauss.hpp",
    "dynamics/src/include/codeGenerationCGinGauss.hpp",
]
STRUCT = re.compile(r"template\s*<>\s*struct\s+(\w+)Impl\s*<\s*([\d,\s]+)>\s*\{(.*?)\};", re.S)
NUMBER = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def extract(root):
    tables = {}
    for rel in HEADERS:
        with open(os.path.join(root, rel)) as f:
            text = f.read()
        for name, args, body in STRUCT.findall(text):
            key = name + "<" + ",".join(a.strip() for a in args.split(",")) + ">"
            # numbers follow the "<<" of the Eigen comma initialiser
            init = body.split("<<", 1)[1]
            tables[key] = [float(x) for x in NUMBER.findall(init)]
    return tables


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    tables = extract(root)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_tables.json")
    with open(out, "w") as f:
        json.dump(tables, f, indent=0, sort_keys=True)
    print(f"wrote {len(tables)} tables to {out}")


if __name__ == "__main__":
    main()
