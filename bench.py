#!/usr/bin/env python3
"""Benchmark of the SPH interpolation hot path on one MI355X.

Workload (the Plummer grid configuration): 2e6 particles drawn from a Plummer
sphere (scale radius 0.1, truncated at 10 a, seed 2), h from the analytic local
density with eta = 1.2, v = solid-body rotation about z plus N(0, 0.01^2) noise,
A = sin(2 pi x) cos(2 pi y); Wendland C4 kernel, symmetric strategy; outputs
rho, A (Shepard), grad A, div v, curl v (L = 9) on a 256^3 Cartesian grid over
[-1, 1]^3.  One step = LinearBVH! on the device (Morton codes, sort, tree,
refit; the particle columns are already in HBM) + the StructuredGrid
interpolation (grid coordinates generated on the device, outputs left in HBM).

Prints one JSON line.  `value` is samples/s of whole steps (build included) in
the FP32-compensated precision (float64 input, float32 kernel evaluation,
float64 accumulation; <= 1e-6 normwise error against float64, checked by the
GPU tests); the same steps in float64 are timed beside it for the speedup.  The
interpolation-only rate, the build time, the dominant kernel's roofline and a
CPU baseline (the C++ oracle on this host's cores, on a bounded subset of the
samples) are reported too.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "partia.jl_amd"))

from partia import _native as N  # noqa: E402
from partia import synthetic  # noqa: E402

METRIC = "GPU samples/s FP32-compensated vs FP64 at ≤1e-6 rel err; LBVH build ms"
HBM_PEAK_GBS = 8000.0
NPART = 2_000_000
NGRID = 256
KERNEL_C4 = 4
# C2 catalog over columns (rho, A, vx, vy, vz): scalars rho (no Shepard), A; grad A; div v; curl v.
CATALOG = dict(sslots=[1, 2], sflags=[0, 1], gslots=[2], dslots=[3, 4, 5], cslots=[3, 4, 5])
# Algorithmic bytes: each accepted neighbour reads x,y,z,h,m,rho,A,vx,vy,vz; each
# sample writes L outputs (its 3 coordinates are generated on the device from the
# axes, but are counted as read as for an explicit sample list).
FIELDS_PER_NEIGHBOUR = 10
L_OUT = 9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "fp32c"], default="fp32c")
    ap.add_argument("--no-fp64-compare", action="store_true", help="skip timing the float64 path beside FP32C")
    ap.add_argument("--traversal", choices=["auto", "lane", "bucket"], default="auto")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU time of the baseline sample")
    return ap.parse_args()


class Dev:
    """Device buffer through the library's allocator."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.p = C.c_void_p()
        ctx.check(N.lib().ph_device_alloc(ctx.handle, int(nbytes), C.byref(self.p)))

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        self.ctx.check(N.lib().ph_memcpy_htod(self.ctx.handle, self.p, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def free(self):
        if self.p:
            N.lib().ph_device_free(self.ctx.handle, self.p)
            self.p = C.c_void_p()


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-probe
    calibration); None when no summary exists."""
    path = os.path.join(ROOT, "profiles", "%s_plummer2e6_c4_symmetric_grid256_hbm.md" % kernel)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline(data, axes, target_seconds):
    """Oracle (C++ restatement of the reference CPU path) on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    cols = [data["rho"], data["A"], data["vx"], data["vy"], data["vz"]]
    t0 = time.perf_counter()
    o = O.OracleSet(data["x"], data["y"], data["z"], data["m"], data["h"], data["rho"], cols,
                    kernel="C4_Wendland", code_bits=63)
    t_build = time.perf_counter() - t0
    nthreads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    nthreads = max(1, min(int(nthreads), 16))
    S = NGRID ** 3
    ax = axes
    # strided subset of the flattened grid: spread over the whole volume
    done, t_interp, stride = 0, 0.0, 4096
    while True:
        idx = np.arange(0, S, stride)
        ix, iy, iz = idx % NGRID, (idx // NGRID) % NGRID, idx // (NGRID * NGRID)
        t0 = time.perf_counter()
        o.point_samples(CATALOG, O.SYMMETRIC, ax[ix], ax[iy], ax[iz], nthreads=nthreads)
        dt = time.perf_counter() - t0
        done, t_interp = len(idx), dt
        if dt * 4 > target_seconds or stride <= 16:
            break
        stride //= 4
    rate_interp = done / t_interp
    total = t_build + S / rate_interp
    return {
        "value": S / total,
        "unit": "samples/s",
        "cores": nthreads,
        "kind": "port",
        "sample": ("C++ oracle (reference algorithm restated) on %d host threads: LinearBVH! on all %d particles "
                   "(%.2f s, serial sort as in the reference) + %d of the %d grid samples (every %d-th, %.2f s); "
                   "whole-grid time extrapolated from that rate" % (nthreads, NPART, t_build, done, S, stride,
                                                                   t_interp)),
        "build_s": t_build,
        "interp_samples_per_s": rate_interp,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (got --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    prec = N.PH_PREC_FP64 if args.precision == "fp64" else N.PH_PREC_FP32C
    print("bench: generating %d Plummer particles" % NPART, file=sys.stderr, flush=True)
    data = synthetic.plummer_sphere(NPART, seed=2)
    axes = synthetic.linrange(-1.0, 1.0, NGRID)
    S = NGRID ** 3

    ctx = N.Context(0)
    ctx.set_traversal({"auto": N.PH_TRAVERSAL_AUTO, "lane": N.PH_TRAVERSAL_LANE,
                       "bucket": N.PH_TRAVERSAL_BUCKET}[args.traversal])
    L = N.lib()
    names = ["x", "y", "z", "m", "h", "rho", "rho", "A", "vx", "vy", "vz"]
    dcols = []
    for nm in names:
        d = Dev(ctx, 8 * NPART)
        d.upload(data[nm].astype(np.float64))
        dcols.append(d)
    dax = Dev(ctx, 8 * NGRID)
    dax.upload(axes)
    out_bytes = 8 * L_OUT * S  # room for float64 outputs (the float64 comparison)
    dout = Dev(ctx, out_bytes)
    cat = N.CatalogArgs(**CATALOG)
    qcols = (C.c_void_p * 5)(*[d.p.value for d in dcols[6:]])

    wall = {"build": [], "interp": [], "destroy": []}

    def step(prec=prec):
        h = C.c_void_p()
        t0 = time.perf_counter()
        ctx.check(L.ph_lbvh_build_device(ctx.handle, NPART, *[d.p for d in dcols[:6]], 5, qcols, KERNEL_C4, 63,
                                         N.PH_PREC_FP64, C.byref(h)))
        t1 = time.perf_counter()
        build_ms = ctx.stats()["build_ms"]
        ctx.check(L.ph_structured_grid_device(ctx.handle, h, cat.ref(), N.PH_SYMMETRIC, prec, dax.p, NGRID, dax.p,
                                              NGRID, dax.p, NGRID, dout.p))
        t2 = time.perf_counter()
        st = ctx.stats()
        L.ph_lbvh_destroy(h)
        t3 = time.perf_counter()
        for k, v in (("build", t1 - t0), ("interp", t2 - t1), ("destroy", t3 - t2)):
            wall[k].append(1e3 * v)
        return build_ms, st["interp_ms"], st["kernel_ms"]

    # one counting pass for the algorithmic bytes (untimed)
    ctx.set_count_neighbours(True)
    step()
    n_acc = ctx.stats()["accepted"]
    n_rechecks = ctx.stats()["rechecks"]
    ctx.set_count_neighbours(False)
    print("bench: accepted neighbours per grid pass = %d (%.2f per sample)" % (n_acc, n_acc / S), file=sys.stderr,
          flush=True)
    for _ in range(args.warmup):
        step()
    ctx.check(L.ph_synchronize(ctx.handle))
    builds, interps, kernels = [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        b, it, k = step()
        builds.append(b)
        interps.append(it)
        kernels.append(k)
        print("bench: step %d build %.2f ms interp %.2f ms (accumulation kernel %.2f ms)" % (i, b, it, k),
              file=sys.stderr, flush=True)
    ctx.check(L.ph_synchronize(ctx.handle))
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    print("bench: host wall ms per call (median of all steps): " +
          ", ".join("%s %.2f" % (k, float(np.median(v))) for k, v in wall.items()), file=sys.stderr, flush=True)
    kernel_ms = float(np.mean(kernels))
    interp_ms = float(np.mean(interps))
    build_ms = float(np.mean(builds))

    fp64 = None
    if prec != N.PH_PREC_FP64 and not args.no_fp64_compare:
        step(N.PH_PREC_FP64)  # warm-up
        ctx.check(L.ph_synchronize(ctx.handle))
        t1 = time.perf_counter()
        k64 = [step(N.PH_PREC_FP64)[2] for _ in range(args.steps)]
        ctx.check(L.ph_synchronize(ctx.handle))
        ms64 = 1e3 * (time.perf_counter() - t1) / args.steps
        fp64 = {"ms_per_step": ms64, "samples_per_s": S / (ms64 * 1e-3), "accumulation_kernel_ms": float(np.mean(k64))}
        print("bench: float64 %.2f ms per step (accumulation kernel %.2f ms)" % (ms64, fp64["accumulation_kernel_ms"]),
              file=sys.stderr, flush=True)

    fbytes = 8 if prec == N.PH_PREC_FP64 else 4
    alg_bytes = n_acc * FIELDS_PER_NEIGHBOUR * fbytes + S * (3 + L_OUT) * fbytes
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic("k_bucket_f64" if prec == N.PH_PREC_FP64 else "k_bucket_c32")
    result = {
        "metric": METRIC,
        "value": S / (ms_per_step * 1e-3),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == N.PH_PREC_FP64 else "f32c",
        "data": "synthetic (seeded Plummer sphere, SplitMix64)",
        "config": {"workload": "plummer_2e6_C4_symmetric_grid256", "particles": NPART, "grid": [NGRID] * 3,
                   "kernel": "C4_Wendland", "strategy": "symmetric", "outputs": L_OUT, "code_bits": 63,
                   "precision": args.precision},
        "lbvh_build_ms": build_ms,
        "interp_ms": interp_ms,
        "interp_samples_per_s": S / (interp_ms * 1e-3),
        "accumulation_kernel_ms": kernel_ms,
        "accepted_neighbours": int(n_acc),
        "guard_band_rechecks": int(n_rechecks),
        "traversal": args.traversal,
        "fp64": fp64,
        "speedup_vs_fp64": (fp64["ms_per_step"] / ms_per_step) if fp64 else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes": int(alg_bytes)},
    }
    for d in dcols + [dax, dout]:
        d.free()
    ctx.close()
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (oracle)", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(data, axes, args.cpu_seconds)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
