"""Writes tests/golden/reference_structs.json: the field names of the
reference's structs that the Julia binding (julia/ext/*.jl) reads, the names it
imports from Partia, and the positional parameter types of the driver methods
it extends, each with the file:line it was read from.  The reference is read
as text (it is Julia; nothing is run).  Run in a checkout that has
/root/reference:

    python tests/golden/make_reference_structs.py
"""
import json
import os
import re

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_structs.json")
SRC = "src/julia"

STRUCTS = {
    "InterpolationInput": "KernelInterpolation/interpolation_setup/InterpolationInput.jl",
    "InterpolationCatalog": "KernelInterpolation/interpolation_setup/InterpolationCatalog.jl",
    "InterpolationCatalogConcise": "KernelInterpolation/interpolation_setup/InterpolationCatalog.jl",
    "PointSamples": "Grids/grids/PointSamples.jl",
    "LineSamples": "Grids/grids/LineSamples.jl",
    "StructuredGrid": "Grids/grids/StructuredGrid.jl",
    "LinearBVH": "LinearBoundingVolumeHierarchy/LinearBVH/LinearBVH.jl",
    "GridBundle": "Grids/griddataset/GridBundle.jl",
}
# names the binding imports from Partia: where each is defined
NAMES = {
    "AbstractInterpolationStrategy": "KernelInterpolation/interpolation_setup/InterpolationStrategy.jl",
    "itpGather": "KernelInterpolation/interpolation_setup/InterpolationStrategy.jl",
    "itpScatter": "KernelInterpolation/interpolation_setup/InterpolationStrategy.jl",
    "itpSymmetric": "KernelInterpolation/interpolation_setup/InterpolationStrategy.jl",
    "M4_spline": "KernelInterpolation/kernel_function/kernels/M4_spline.jl",
    "M5_spline": "KernelInterpolation/kernel_function/kernels/M5_spline.jl",
    "M6_spline": "KernelInterpolation/kernel_function/kernels/M6_spline.jl",
    "C2_Wendland": "KernelInterpolation/kernel_function/kernels/C2_Wendland.jl",
    "C4_Wendland": "KernelInterpolation/kernel_function/kernels/C4_Wendland.jl",
    "C6_Wendland": "KernelInterpolation/kernel_function/kernels/C6_Wendland.jl",
    "Cartesian": "Grids/grids/coordinate.jl",
    "Cylindrical": "Grids/grids/coordinate.jl",
    "Spherical": "Grids/grids/coordinate.jl",
    "CPUComputeBackend": "KernelInterpolation/ExecutionBackend/AbstractExecutionBackend.jl",
    "AbstractExecutionBackend": "KernelInterpolation/ExecutionBackend/AbstractExecutionBackend.jl",
}
# driver functions the binding adds methods to, and helpers it calls
FUNCTIONS = {
    "PointSamples_interpolation": "KernelInterpolation/grid_interpolation/drivers/PointSamples_driver.jl",
    "LineSamples_interpolation": "KernelInterpolation/grid_interpolation/drivers/LineSamples_driver.jl",
    "StructuredGrid_interpolation": "KernelInterpolation/grid_interpolation/drivers/StructuredGrid_driver.jl",
    "initialize_interpolation": "KernelInterpolation/grid_interpolation/setup/initialize_interpolation.jl",
    "matches_lbvh_leaf_order": "KernelInterpolation/interpolation_setup/InterpolationInput.jl",
    "to_concise_catalog": "KernelInterpolation/interpolation_setup/InterpolationCatalog.jl",
}


def read(rel):
    with open(os.path.join(REF, SRC, rel)) as f:
        return f.read().split("\n")


def struct_fields(name, rel):
    lines = read(rel)
    for i, ln in enumerate(lines):
        if re.match(r"\s*struct\s+%s\b" % name, ln):
            fields = []
            j = i + 1
            while not re.match(r"\s*end\b", lines[j]):
                m = re.match(r"\s*([^\s:#]+)\s*::", lines[j])
                if m:
                    fields.append(m.group(1))
                if re.match(r"\s*function\b", lines[j]):  # inner constructor: fields end here
                    break
                j += 1
            return {"fields": fields, "source": "%s/%s:%d" % (SRC, rel, i + 1)}
    raise SystemExit("struct %s not found in %s" % (name, rel))


def defined(name, rel):
    for i, ln in enumerate(read(rel)):
        if re.match(r"\s*(abstract\s+type|struct)\s+%s\b" % name, ln):
            return "%s/%s:%d" % (SRC, rel, i + 1)
    raise SystemExit("%s not defined in %s" % (name, rel))


def split_top(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def methods(name, rel):
    """Positional parameter type heads of every method `function name(...)`."""
    found = []
    for i, ln in enumerate(read(rel)):
        m = re.match(r"\s*(?:@inline\s+)?function\s+%s\s*\((.*)$" % name, ln)
        if not m:
            continue
        depth, args = 1, ""
        for ch in m.group(1):
            if ch == "(":
                depth += 1
            elif ch == ")":
                depth -= 1
                if depth == 0:
                    break
            args += ch
        heads = []
        for a in split_top(args):
            a = a.split("=")[0]
            t = a.split("::")[-1].strip() if "::" in a else "Any"
            heads.append(re.match(r"[A-Za-z_][A-Za-z0-9_]*", t).group(0))
        found.append({"param_types": heads, "source": "%s/%s:%d" % (SRC, rel, i + 1)})
    if not found:
        raise SystemExit("no method of %s in %s" % (name, rel))
    return found


def main():
    data = {
        "comment": "Field names, definitions and method signatures read from the reference's Julia source "
                   "(text only) by tests/golden/make_reference_structs.py.",
        "structs": {k: struct_fields(k, v) for k, v in STRUCTS.items()},
        "names": {k: defined(k, v) for k, v in NAMES.items()},
        "functions": {k: methods(k, v) for k, v in FUNCTIONS.items()},
    }
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
