"""Writes frame_fixtures.json: the frames and axis parameters of the reference's
frame-plane tests (test/structured_grid_interpolation_tests.jl:110-135 PointSamples
Cartesian plane / box, :137-197 Polar plane, :202-239 LineSamples ParallelBeam,
:241-254 invalid parameters), with each frame's basis vectors.

A Frame(x, f0, u0) in its initial orientation has forward f0 / |f0|, up u0 / |u0|
and right = forward x up, normalised (Frames/struct/Frame.jl:42-78, 145-186);
here numpy's cross product and norm compute them.  Both reference frames look
down -z with +y up, so right is +x."""
import json
import math
import os

import numpy as np


def basis(x, f0, u0):
    f = np.asarray(f0, dtype=np.float64) / np.linalg.norm(f0)
    u = np.asarray(u0, dtype=np.float64) / np.linalg.norm(u0)
    r = np.cross(f, u)
    r = r / np.linalg.norm(r)
    return dict(x=list(map(float, x)), f0=list(map(float, f0)), u0=list(map(float, u0)),
                right=[float(v) + 0.0 for v in r], up=[float(v) for v in u], forward=[float(v) for v in f])


def main():
    out = {
        "source": "test/structured_grid_interpolation_tests.jl:110-254",
        "frames": {
            "origin": basis((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0)),   # :111, :243
            "offset": basis((0.25, -0.5, 1.25), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0)),  # :203
        },
        "xparams": [-1.0, 1.0, 5],
        "yparams": [-0.75, 0.75, 4],
        "zparams": [-0.5, 0.5, 3],
        "sparams": [0.0, 1.0, 4],
        "phiparams": [0.0, 2 * math.pi, 8],
        "partial_phiparams": [math.pi / 4, 3 * math.pi / 4, 5],
        "invalid_cartesian": [[[0.0, 0.0, 2], [0.0, 1.0, 2]], [[0.0, 1.0, 1], [0.0, 1.0, 2]],
                              [[0.0, 1.0, 2], [0.0, 0.0, 2]], [[0.0, 1.0, 2], [0.0, 1.0, 1]]],
        "invalid_polar": [[[-1.0, 1.0, 2], [0.0, 2 * math.pi, 3]], [[1.0, 1.0, 2], [0.0, 2 * math.pi, 3]],
                          [[0.0, 1.0, 1], [0.0, 2 * math.pi, 3]], [[0.0, 1.0, 2], [0.0, 2 * math.pi, 0]],
                          [[0.0, 1.0, 2], [math.pi, 0.0, 3]]],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "frame_fixtures.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
