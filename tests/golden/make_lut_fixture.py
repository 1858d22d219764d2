#!/usr/bin/env python3
"""Write tests/golden/line_lut_ref_samples.json from the reference's tabulated
line-integrated kernels (src/julia/KernelInterpolation/table/
line_integrated_kernel_tables.jl:1,515,1029,1544,...,8744).

The fixture keeps every 64th entry (k = 0, 64, ..., 4032, plus k = 4095) of the
Q, I and Inorm tables of all six kernels as float64 hex strings (exact), and a
comparison report of the full tables against this project's own quadrature
(partia.jl_amd/csrc/gen_line_lut.py).  It only reads the reference as data; it
runs in the development container where the reference is mounted, never on the
GPU machine.
"""
import json
import os
import re
import sys

import numpy as np

REF = "/root/reference/src/julia/KernelInterpolation/table/line_integrated_kernel_tables.jl"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "partia.jl_amd", "csrc"))
import gen_line_lut  # noqa: E402


def parse(path):
    src = open(path).read()
    out = {}
    for m in re.finditer(r"const (_\w+) = SVector\{4096,Float64\}\((.*?)\)", src, re.S):
        out[m.group(1)] = np.array([float(v) for v in m.group(2).split(",") if v.strip()])
    return out


def main():
    tabs = parse(REF)
    idx = list(range(0, 4096, 64)) + [4095]
    fixture = {"source": "src/julia/KernelInterpolation/table/line_integrated_kernel_tables.jl",
               "indices": idx, "kernels": {}, "report": {}}
    for name, K, knots in gen_line_lut.KERNELS:
        Q, I, Inorm = (tabs["_%s_%s" % (name, s)] for s in ("Q", "I", "Inorm"))
        fixture["kernels"][name] = {
            "Q": [float(Q[i]).hex() for i in idx],
            "I": [float(I[i]).hex() for i in idx],
            "Inorm": [float(Inorm[i]).hex() for i in idx],
        }
        q, i_ours, n_ours = gen_line_lut.tables(name, K, knots)
        d = np.abs(n_ours - Inorm)
        fixture["report"][name] = {
            "Q_bitwise_equal": bool(np.array_equal(q, Q)),
            "max_abs_Inorm_diff_over_Inorm0": float(d.max() / Inorm[0]),
            "median_abs_Inorm_diff_over_Inorm0": float(np.median(d) / Inorm[0]),
            "argmax_k": int(np.argmax(d)),
        }
    with open(os.path.join(HERE, "line_lut_ref_samples.json"), "w") as f:
        json.dump(fixture, f, indent=1)
    print(json.dumps(fixture["report"], indent=1))


if __name__ == "__main__":
    main()
