"""Extract the reference's embedded element-matrix integer tables as a golden vector.

The reference ships no test fixtures; the only embedded numeric data on the solve path are the
Lame tables kelam72[24*24] and kemu72[24*24] in homogenization/templateMatrix.cpp:8-61
(72 * KE for (lambda, mu) = (1, 0) and (0, 1)).  This script reads those two integer arrays out of
the reference source (run in the container where /root/reference exists) and writes them as
data to tests/golden/ke_tables.json.  Only the numbers are kept.
"""
import json
import os
import re
import sys

REF = "/root/reference/homogenization/templateMatrix.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ke_tables.json")


def extract(src, name):
    m = re.search(r"char\s+%s\s*\[\s*\]\s*=\s*\{([^}]*)\}" % name, src)
    if not m:
        raise SystemExit("table %s not found" % name)
    vals = [int(t) for t in re.findall(r"-?\d+", m.group(1))]
    if len(vals) != 576:
        raise SystemExit("table %s has %d entries" % (name, len(vals)))
    return vals


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else REF
    with open(path) as fh:
        src = fh.read()
    data = {
        "source": "homogenization/templateMatrix.cpp:8-61",
        "kelam72": extract(src, "kelam72"),
        "kemu72": extract(src, "kemu72"),
    }
    with open(OUT, "w") as fh:
        json.dump(data, fh)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
