"""Extracts the numeric constants of the reference's default kernel
configuration (ocl/kernel.cl:18-88: LN=10, K=4, PCK=2, LBASE=28) into
tests/golden/kernel_cl_constants.json.  Run in the build container, where the
reference tree is mounted; the JSON (numbers only) is the committed fixture."""
import json
import os
import re
import sys

REF = "/root/reference/ocl/kernel.cl"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    text = open(REF).read()
    text = text[text.index("#if !defined(N_SZ)"):]
    text = text[:text.index("#endif")]
    out = {}
    for name, val in re.findall(r"#define\s+(\w+)\s+(\d+)u?\s*$", text, re.M):
        out[name] = int(val)
    for name, vals in re.findall(r"#define\s+(\w+)\s+const uint \w+\[\d+\] = \{([^}]*)\}", text):
        out[name] = [int(v.strip().rstrip("u")) for v in vals.split(",")]
    json.dump(out, open(os.path.join(HERE, "kernel_cl_constants.json"), "w"), indent=1, sort_keys=True)
    print(f"{len(out)} constants", file=sys.stderr)


if __name__ == "__main__":
    main()
