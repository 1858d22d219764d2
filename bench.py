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

Other workloads (`--workload`), each printing one line of the same form:
  column2048  8e6 particles in a log-normal density field (Gaussian random field
              with P(k) ~ k^-11/3 on 128^3, sigma_ln = 1.5, seed 3), tabulated
              line-integrated M4 kernel, Float32; column density (:rho,) on a
              2048^2 image of parallel rays along z over [0,1]^2.  One step =
              Float32 LinearBVH! + the projection; value = rays/s.
  grid64      1e5 uniform particles (seed 1), M4, gather, Float64; density and
              A (Shepard) on a 64^3 grid over [0,1]^3.
  adjoint1024 1e6 Plummer particles (the grid workload's recipe, seed 5),
              Float32 column density on a 1024^2 image along z over [-1,1]^2,
              L2 loss against the map of masses m exp(N(0, 0.1^2)); one step =
              the forward projection + its adjoint d/d(m, A, h, x, y, z);
              value = rays/s of the adjoint, with the backward/forward time
              ratio, and a float64 central-difference check of the adjoint on
              64 random particles (relative step 1e-4).
  batch512    4e6 particles (half uniform, half a Plummer clump, seed 4) and 512
              sampling requests of uneven size (point lists of 1e3-1e6 random
              points, oriented planes 256^2-1024^2, oriented boxes 32^3-128^3)
              served by one graph launch of the persistent work-queue kernel
              (C4, symmetric, FP32-compensated, the 9 outputs above); value =
              samples/s of the whole batch, with per-request p50/p99 latency and
              the same requests issued as 512 separate calls beside it.
"""
import argparse
import ctypes as C
import json
import platform
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
    ap.add_argument("--autotune", choices=["on", "off"], default="on",
                    help="tune the launch configuration on the first (untimed) call of each workload key")
    ap.add_argument("--launch", default=None, metavar="OCC,XCD[,ORDER[,LEAVES[,BLOCK]]]",
                    help="fixed launch configuration (occupancy, block order, tile order, leaves per bucket and "
                         "threads per block of the grid tile walks) of the point / grid kernels; turns the autotuner "
                         "off")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU time of the baseline sample")
    ap.add_argument("--workload", choices=["grid256", "column2048", "grid64", "batch512", "adjoint1024", "gridadjoint256"],
                    default="grid256")
    ap.add_argument("--output", choices=list(ADJ_OUTPUTS), default="scalar",
                    help="output group of gridadjoint256: the Shepard scalar A, grad A, div v or curl v")
    return ap.parse_args()


# gridadjoint256's output groups: catalog over the columns (rho, A, vx, vy, vz),
# values per sample, field columns (gradient rows between m and h) and their names
ADJ_OUTPUTS = {
    "scalar": (dict(sslots=[2], sflags=[1]), 1, ["A"], "A (Shepard)"),
    "gradient": (dict(gslots=[2]), 3, ["A"], "grad A"),
    "divergence": (dict(dslots=[3, 4, 5]), 1, ["vx", "vy", "vz"], "div v"),
    "curl": (dict(cslots=[3, 4, 5]), 3, ["vx", "vy", "vz"], "curl v"),
}


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


def apply_launch(ctx, args):
    """--launch OCC,XCD fixes the launch configuration; otherwise --autotune decides."""
    if args.launch:
        v = args.launch.split(",")
        ctx.set_autotune(False)
        ctx.set_launch(float(v[0]), int(v[1]), int(v[2]) if len(v) > 2 else 0)
        ctx.set_tile_shape(int(v[3]) if len(v) > 3 else 0, int(v[4]) if len(v) > 4 else 0)
    else:
        ctx.set_autotune(args.autotune == "on")


def measured_traffic(kernel, workload="plummer2e6_c4_symmetric_grid256"):
    """HBM bytes per launch of `kernel` on a workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-probe
    calibration); None when no summary exists."""
    path = os.path.join(ROOT, "profiles", "%s_%s_hbm.md" % (kernel, workload))
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def oracle_rate(run, S, target_seconds, stride0):
    """Samples/s of run(idx) on a strided subset of S samples, the subset grown
    until it takes about a quarter of target_seconds; returns (rate, count, stride, seconds)."""
    stride = stride0
    while True:
        idx = np.arange(0, S, stride)
        t0 = time.perf_counter()
        run(idx)
        dt = time.perf_counter() - t0
        if dt * 4 > target_seconds or stride <= 16:
            return len(idx) / dt, len(idx), stride, dt
        stride //= 4


def cpu_baseline(data, axes, target_seconds):
    """Oracle (C++ restatement of the reference CPU path) on this host: all the
    threads this process may use, and one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    cols = [data["rho"], data["A"], data["vx"], data["vy"], data["vz"]]
    t0 = time.perf_counter()
    o = O.OracleSet(data["x"], data["y"], data["z"], data["m"], data["h"], data["rho"], cols,
                    kernel="C4_Wendland", code_bits=63)
    t_build = time.perf_counter() - t0  # serial, as the reference's sort and permutation
    nthreads = threads()
    S = NGRID ** 3
    ax = axes

    def run(nt):
        def f(idx):
            ix, iy, iz = idx % NGRID, (idx // NGRID) % NGRID, idx // (NGRID * NGRID)
            o.point_samples(CATALOG, O.SYMMETRIC, ax[ix], ax[iy], ax[iz], nthreads=nt)
        return f

    rate, done, stride, t_interp = oracle_rate(run(nthreads), S, target_seconds, 4096)
    rate1, done1, stride1, t1 = oracle_rate(run(1), S, target_seconds / 2, 65536)
    return {
        "value": S / (t_build + S / rate),
        "unit": "samples/s",
        "cores": nthreads,
        "kind": "port",
        "sample": ("C++ oracle (reference algorithm restated) on %d host threads of a %s: LinearBVH! on all %d "
                   "particles (%.2f s, serial sort as in the reference) + %d of the %d grid samples (every %d-th, "
                   "%.2f s); whole-grid time extrapolated from that rate" % (nthreads, cpu_model(), NPART, t_build,
                                                                           done, S, stride, t_interp)),
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
        "threads_available": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
        "build_s": t_build,
        "interp_samples_per_s": rate,
        "one_core": {"value": S / (t_build + S / rate1), "unit": "samples/s", "interp_samples_per_s": rate1,
                     "sample": "%d of the %d grid samples (every %d-th, %.2f s) on 1 thread" % (done1, S, stride1,
                                                                                             t1)},
    }


FIELDS = (("rho", [0]), ("A", [1]), ("grad_A", [2, 3, 4]), ("div_v", [5]), ("curl_v", [6, 7, 8]))


def field_errors(got, ref):
    """Per output field of the C2 catalog (a vector field counted whole), over
    the samples with neighbours: max and RMS of |got - ref|, the RMS of ref,
    the normwise error ||got - ref|| / ||ref|| and the largest pointwise
    |got - ref| / max(|ref|, 1e-3 rms ref); NaN masks must agree."""
    out = {}
    for name, rows in FIELDS:
        g, r = got[rows].astype(np.float64), ref[rows]
        same_nan = bool(np.array_equal(np.isnan(g), np.isnan(r)))
        f = ~np.isnan(r[0])
        d = np.abs(g[:, f] - r[:, f])
        rr = r[:, f]
        rms_ref = float(np.sqrt(np.mean(rr * rr))) if rr.size else 0.0
        out[name] = {
            "nan_masks_equal": same_nan,
            "max_abs": float(d.max()) if d.size else 0.0,
            "rms_abs": float(np.sqrt(np.mean(d * d))) if d.size else 0.0,
            "rms_ref": rms_ref,
            "normwise": float(np.linalg.norm(d) / max(np.linalg.norm(rr), 1e-300)),
            "max_rel_pointwise": float(np.max(d / np.maximum(np.abs(rr), max(1e-3 * rms_ref, 1e-300))))
            if d.size else 0.0,
        }
    return out


def threads():
    """Threads of the CPU baseline: the CPUs this process may use, at most the
    16 CPUs a one-GPU job of the GPU box is given (OMP_NUM_THREADS there)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(int(n), int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16))


def timed_steps(ctx, step, steps, warmup, label):
    """Untimed warm-up, then `steps` timed steps between two synchronisations;
    returns (ms per step, list of per-step results)."""
    L = N.lib()
    for _ in range(warmup):
        step()
    ctx.check(L.ph_synchronize(ctx.handle))
    res = []
    t0 = time.perf_counter()
    for i in range(steps):
        r = step()
        res.append(r)
        print("bench: %s step %d %s" % (label, i, " ".join("%s %.2f ms" % kv for kv in r.items())),
              file=sys.stderr, flush=True)
    ctx.check(L.ph_synchronize(ctx.handle))
    return 1e3 * (time.perf_counter() - t0) / steps, res


def upload_columns(ctx, data, names):
    out = []
    for nm in names:
        d = Dev(ctx, 8 * len(data[nm]))
        d.upload(np.asarray(data[nm], dtype=np.float64))
        out.append(d)
    return out


def column2048(args):
    """8e6 log-normal particles, M4 line-integrated column density on 2048^2 rays, Float32."""
    npart, nimg = 8_000_000, 2048
    kernel_m4 = 0
    print("bench: generating %d log-normal particles" % npart, file=sys.stderr, flush=True)
    data = synthetic.lognormal_field(npart, seed=3)
    ax = synthetic.linrange(0.0, 1.0, nimg)
    S = nimg * nimg
    ctx = N.Context(0)
    apply_launch(ctx, args)
    if args.traversal == "lane":
        ctx.set_traversal(N.PH_TRAVERSAL_LANE)  # one DFS per ray instead of one per 8 x 8 tile
    L = N.lib()
    dcols = upload_columns(ctx, data, ["x", "y", "z", "m", "h", "rho", "rho"])
    dax = Dev(ctx, 8 * nimg)
    dax.upload(ax)
    dout = Dev(ctx, 4 * S)
    cat = N.CatalogArgs(sslots=[1], sflags=[0])
    qcols = (C.c_void_p * 1)(dcols[6].p.value)
    frame = [np.ascontiguousarray(np.array(v, dtype=np.float64)) for v in
             ([0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0])]

    def step():
        h = C.c_void_p()
        ctx.check(L.ph_lbvh_build_device(ctx.handle, npart, *[d.p for d in dcols[:6]], 1, qcols, kernel_m4, 63,
                                         N.PH_PREC_FP32, C.byref(h)))
        b = ctx.stats()["build_ms"]
        ctx.check(L.ph_line_grid_device(ctx.handle, h, cat.ref(), N.PH_SCATTER, N.PH_PREC_FP32,
                                        *[N.ptr(f) for f in frame], dax.p, nimg, dax.p, nimg, dout.p))
        st = ctx.stats()
        L.ph_lbvh_destroy(h)
        return {"build": b, "interp": st["interp_ms"], "kernel": st["kernel_ms"]}

    ctx.set_count_neighbours(True)
    step()
    n_acc = ctx.stats()["accepted"]
    ctx.set_count_neighbours(False)
    print("bench: accepted neighbours per image = %d (%.1f per ray)" % (n_acc, n_acc / S), file=sys.stderr,
          flush=True)
    ms, res = timed_steps(ctx, step, args.steps, args.warmup, "column2048")
    kernel_ms = float(np.mean([r["kernel"] for r in res]))
    # per accepted neighbour x, y, z, h, m, rho and the quantity column (7 float32);
    # per ray 3 origin + 3 direction floats in, one output out
    alg_bytes = n_acc * 7 * 4 + S * 7 * 4
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic("k_lines_wave", "lognormal8e6_m4_column2048")
    result = {
        "metric": "GPU rays/s, line-integrated column density (Float32); LBVH build ms",
        "value": S / (ms * 1e-3), "unit": "rays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded log-normal Gaussian random field, SplitMix64)",
        "config": {"workload": "lognormal_8e6_M4_column_density_2048", "particles": npart, "image": [nimg, nimg],
                   "kernel": "M4_spline", "strategy": "scatter", "outputs": 1, "code_bits": 63,
                   "precision": "fp32"},
        "lbvh_build_ms": float(np.mean([r["build"] for r in res])),
        "interp_ms": float(np.mean([r["interp"] for r in res])),
        "accumulation_kernel_ms": kernel_ms,
        "accepted_neighbours": int(n_acc),
        "traversal": "per-ray" if args.traversal == "lane" else "wave per 8x8 tile",
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes": int(alg_bytes)},
    }
    for d in dcols + [dax, dout]:
        d.free()
    ctx.close()
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (oracle)", file=sys.stderr, flush=True)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        t0 = time.perf_counter()
        o = O.OracleSet(data["x"], data["y"], data["z"], data["m"], data["h"], data["rho"], [data["rho"]],
                        kernel="M4_spline", code_bits=63, precision=32)
        t_build = time.perf_counter() - t0
        nt = threads()
        stride = 1 << 14
        while True:
            idx = np.arange(0, S, stride)
            i, j = idx % nimg, idx // nimg
            z = np.zeros(len(idx))
            t0 = time.perf_counter()
            o.line_samples({"sslots": [1], "sflags": [0]}, O.SCATTER,
                           [ax[i].astype(np.float32).astype(np.float64), ax[j].astype(np.float32).astype(np.float64),
                            z], [z, z, z + 1.0], nthreads=nt)
            dt = time.perf_counter() - t0
            if dt * 4 > args.cpu_seconds or stride <= 16:
                break
            stride //= 4
        rate = len(idx) / dt
        result["cpu_baseline"] = {
            "value": S / (t_build + S / rate), "unit": "rays/s", "cores": nt, "kind": "port",
            "sample": ("C++ oracle (reference algorithm restated, Float32) on %d host threads: LinearBVH! on all %d "
                       "particles (%.2f s) + %d of the %d rays (every %d-th, %.2f s); whole-image time extrapolated"
                       % (nt, npart, t_build, len(idx), S, stride, dt)),
            "build_s": t_build, "interp_rays_per_s": rate}
    print(json.dumps(result))


def grid64(args):
    """1e5 uniform particles, M4 gather, density + A on 64^3, Float64."""
    npart, ng = 100_000, 64
    data = synthetic.uniform_box(npart, seed=1)
    ax = synthetic.linrange(0.0, 1.0, ng)
    S = ng ** 3
    ctx = N.Context(0)
    apply_launch(ctx, args)
    L = N.lib()
    dcols = upload_columns(ctx, data, ["x", "y", "z", "m", "h", "rho", "rho", "A"])
    dax = Dev(ctx, 8 * ng)
    dax.upload(ax)
    dout = Dev(ctx, 8 * 2 * S)
    cat = N.CatalogArgs(sslots=[1, 2], sflags=[0, 1])
    qcols = (C.c_void_p * 2)(dcols[6].p.value, dcols[7].p.value)

    def step():
        h = C.c_void_p()
        ctx.check(L.ph_lbvh_build_device(ctx.handle, npart, *[d.p for d in dcols[:6]], 2, qcols, 0, 63,
                                         N.PH_PREC_FP64, C.byref(h)))
        b = ctx.stats()["build_ms"]
        ctx.check(L.ph_structured_grid_device(ctx.handle, h, cat.ref(), N.PH_GATHER, N.PH_PREC_FP64, dax.p, ng,
                                              dax.p, ng, dax.p, ng, dout.p))
        st = ctx.stats()
        L.ph_lbvh_destroy(h)
        return {"build": b, "interp": st["interp_ms"], "kernel": st["kernel_ms"]}

    ctx.set_count_neighbours(True)
    step()
    n_acc = ctx.stats()["accepted"]
    ctx.set_count_neighbours(False)
    ms, res = timed_steps(ctx, step, args.steps, args.warmup, "grid64")
    kernel_ms = float(np.mean([r["kernel"] for r in res]))
    alg_bytes = n_acc * 8 * 8 + S * 5 * 8  # x,y,z,h,m,rho,rho,A per neighbour; 3 in + 2 out per sample
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    result = {
        "metric": "GPU samples/s (Float64); LBVH build ms",
        "value": S / (ms * 1e-3), "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded uniform box, SplitMix64)",
        "config": {"workload": "uniform_1e5_M4_gather_grid64", "particles": npart, "grid": [ng] * 3,
                   "kernel": "M4_spline", "strategy": "gather", "outputs": 2, "code_bits": 63,
                   "precision": "fp64"},
        "lbvh_build_ms": float(np.mean([r["build"] for r in res])),
        "interp_ms": float(np.mean([r["interp"] for r in res])),
        "accumulation_kernel_ms": kernel_ms,
        "accepted_neighbours": int(n_acc),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": int(alg_bytes)},
    }
    for d in dcols + [dax, dout]:
        d.free()
    ctx.close()
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        t0 = time.perf_counter()
        o = O.OracleSet(data["x"], data["y"], data["z"], data["m"], data["h"], data["rho"], [data["rho"], data["A"]],
                        kernel="M4_spline", code_bits=63)
        t_build = time.perf_counter() - t0
        nt = threads()
        g = np.meshgrid(ax, ax, ax, indexing="ij")
        P = [a.ravel(order="F") for a in g]
        t0 = time.perf_counter()
        o.point_samples({"sslots": [1, 2], "sflags": [0, 1]}, O.GATHER, *P, nthreads=nt)
        dt = time.perf_counter() - t0
        result["cpu_baseline"] = {
            "value": S / (t_build + dt), "unit": "samples/s", "cores": nt, "kind": "port",
            "sample": "C++ oracle on %d host threads: LinearBVH! (%.2f s) + the whole 64^3 grid (%.2f s)"
                      % (nt, t_build, dt), "build_s": t_build}
    print(json.dumps(result))


def batch_cpu_baseline(data, mix, host_pts, sizes, target_seconds):
    """Oracle on every 32nd request, strided samples, FP64 (the reference's
    arithmetic); whole-batch time extrapolated from the sample rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    cols = [data["rho"], data["A"], data["vx"], data["vy"], data["vz"]]
    t0 = time.perf_counter()
    o = O.OracleSet(data["x"], data["y"], data["z"], data["m"], data["h"], data["rho"], cols,
                    kernel="C4_Wendland", code_bits=63)
    t_build = time.perf_counter() - t0
    nt = threads()
    picks = list(range(0, len(mix), 32))

    def run(stride):
        done = 0
        for r in picks:
            P = host_pts[r]
            k = np.arange(0, sizes[r], stride)
            o.point_samples(CATALOG, O.SYMMETRIC, P[0][k], P[1][k], P[2][k], nthreads=nt)
            done += len(k)
        return done

    stride = 4096
    while True:
        t0 = time.perf_counter()
        done = run(stride)
        dt = time.perf_counter() - t0
        if dt * 4 > target_seconds or stride <= 16:
            break
        stride //= 4
    rate = done / dt
    S_tot = int(sum(sizes))
    return {"value": S_tot / (t_build + S_tot / rate), "unit": "samples/s", "cores": nt, "kind": "port",
            "sample": ("C++ oracle (reference algorithm restated, float64) on %d host threads of a %s: LinearBVH! on "
                       "all %d particles (%.2f s) + %d samples of %d requests (every %d-th sample of every 32nd "
                       "request, %.2f s); whole-batch time extrapolated" % (nt, cpu_model(), len(data["x"]), t_build,
                                                                           done, len(picks), stride, dt)),
            "build_s": t_build, "interp_samples_per_s": rate}


def batch512(args):
    """512 uneven requests over a 4e6-particle uniform-plus-clump set in one graph launch."""
    from partia import batch as PB
    from partia import grids
    npart, nreq = 4_000_000, 512
    print("bench: generating %d particles and %d requests" % (npart, nreq), file=sys.stderr, flush=True)
    data = synthetic.uniform_plus_clump(npart, seed=4)
    mix = synthetic.request_mix(nreq, seed=4)
    ctx = N.Context(0)
    L = N.lib()
    names = ["x", "y", "z", "m", "h", "rho", "rho", "A", "vx", "vy", "vz"]
    dcols = upload_columns(ctx, data, names)
    qcols = (C.c_void_p * 5)(*[d.p.value for d in dcols[6:]])
    h = C.c_void_p()
    ctx.check(L.ph_lbvh_build_device(ctx.handle, npart, *[d.p for d in dcols[:6]], 5, qcols, KERNEL_C4, 63,
                                     N.PH_PREC_FP64, C.byref(h)))
    build_ms = ctx.stats()["build_ms"]
    cat = N.CatalogArgs(**CATALOG)
    reqs, sizes, dev, host_pts = [], [], [], []
    for q in mix:
        if q["kind"] == "points":
            st = synthetic.Stream(q["seed"])
            P = [st.uniform(q["count"]) for _ in range(3)]
        elif q["kind"] == "plane":
            P = grids.cartesian_plane_coordinates(q["position"], q["right"], q["up"], q["axis"], q["axis"])
        else:
            P = grids.cartesian_box_coordinates(q["position"], q["right"], q["up"], q["forward"], q["axis"],
                                                q["axis"], q["axis"])
        S = len(P[0])
        sizes.append(S)
        host_pts.append(P)
        out = Dev(ctx, 4 * L_OUT * max(S, 1))
        dev.append(out)
        if q["kind"] == "points":
            dp = [Dev(ctx, 8 * S) for _ in range(3)]
            for dd, a in zip(dp, P):
                dd.upload(a)
            dev += dp
            reqs.append(PB.points(dp[0].p, dp[1].p, dp[2].p, S, out.p))
        elif q["kind"] == "plane":
            reqs.append(PB.plane(q["position"], q["right"], q["up"], q["axis"], q["axis"], out.p))
        else:
            reqs.append(PB.box(q["position"], q["right"], q["up"], q["forward"], q["axis"], q["axis"], q["axis"],
                               out.p))
    S_tot = int(sum(sizes))
    print("bench: %d samples in %d requests" % (S_tot, nreq), file=sys.stderr, flush=True)
    b = PB.Batch(ctx, h, cat, N.PH_SYMMETRIC, N.PH_PREC_FP32C, reqs)

    def step():
        b.run()
        return {"batch": ctx.stats()["interp_ms"]}

    ms, res = timed_steps(ctx, step, args.steps, args.warmup, "batch512")
    t_first, t_last = b.times()
    lat = t_last / 1e3  # completion after the launch began, ms
    span = (t_last - t_first) / 1e3
    # the same requests as separate calls (plane / box coordinates uploaded beforehand)
    sep_ms = None
    n_acc = 0
    if not args.no_fp64_compare:
        pre = []
        for q, P, S, r in zip(mix, host_pts, sizes, reqs):
            if q["kind"] == "points":
                pre.append((r.px, r.py, r.pz))
            else:
                dp = [Dev(ctx, 8 * S) for _ in range(3)]
                for dd, a in zip(dp, P):
                    dd.upload(a)
                dev += dp
                pre.append(tuple(dd.p.value for dd in dp))
        ctx.check(L.ph_synchronize(ctx.handle))
        t0 = time.perf_counter()
        for (px, py, pz), S, r in zip(pre, sizes, reqs):
            ctx.check(L.ph_point_samples_device(ctx.handle, h, cat.ref(), N.PH_SYMMETRIC, N.PH_PREC_FP32C, S,
                                                C.c_void_p(px), C.c_void_p(py), C.c_void_p(pz), C.c_void_p(r.out)))
        ctx.check(L.ph_synchronize(ctx.handle))
        sep_ms = 1e3 * (time.perf_counter() - t0)
        print("bench: the same requests as %d separate calls: %.1f ms" % (nreq, sep_ms), file=sys.stderr, flush=True)
        # accepted neighbours of every request (the batch's are the same), untimed
        ctx.set_count_neighbours(True)
        for (px, py, pz), S, r in zip(pre, sizes, reqs):
            ctx.check(L.ph_point_samples_device(ctx.handle, h, cat.ref(), N.PH_SYMMETRIC, N.PH_PREC_FP32C, S,
                                                C.c_void_p(px), C.c_void_p(py), C.c_void_p(pz), C.c_void_p(r.out)))
            n_acc += ctx.stats()["accepted"]
        ctx.set_count_neighbours(False)
    result = {
        "metric": "GPU samples/s over 512 uneven requests in one graph launch; per-request p50/p99 latency",
        "value": S_tot / (ms * 1e-3), "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32c",
        "data": "synthetic (seeded uniform + Plummer clump, seeded request mix, SplitMix64)",
        "config": {"workload": "uniform_plus_clump_4e6_512_requests", "particles": npart, "requests": nreq,
                   "samples": S_tot, "kernel": "C4_Wendland", "strategy": "symmetric", "outputs": L_OUT,
                   "precision": "fp32c"},
        "lbvh_build_ms": build_ms,
        "batch_ms": float(np.mean([r["batch"] for r in res])),
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(np.max(lat))},
        "service_span_ms": {"p50": float(np.percentile(span, 50)), "p99": float(np.percentile(span, 99))},
        "separate_calls_ms": sep_ms,
        "speedup_vs_separate_calls": (sep_ms / ms) if sep_ms else None,
    }
    if n_acc:
        # per accepted neighbour the 10 float32 fields of the C2 catalog; per
        # sample 9 float32 outputs, plus 3 float64 coordinates read for point lists
        n_pts = int(sum(S for q, S in zip(mix, sizes) if q["kind"] == "points"))
        alg_bytes = n_acc * FIELDS_PER_NEIGHBOUR * 4 + S_tot * L_OUT * 4 + n_pts * 24
        batch_ms = result["batch_ms"]
        achieved = alg_bytes / (batch_ms * 1e-3) / 1e9
        traffic, traffic_src = measured_traffic("k_batch", "uniform_clump4e6_512_requests")
        result["accepted_neighbours"] = int(n_acc)
        result["roofline"] = {"bound": "hbm", "kernel": "k_batch", "achieved": achieved, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                              "traffic_source": traffic_src, "algorithmic_bytes": int(alg_bytes)}
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = batch_cpu_baseline(data, mix, host_pts, sizes, args.cpu_seconds)
    b.close()
    L.ph_lbvh_destroy(h)
    for d in dcols + dev:
        d.free()
    ctx.close()
    print(json.dumps(result))


def adjoint1024(args):
    """Forward + adjoint of a 1024^2 Float32 column-density map of 1e6 Plummer particles."""
    npart, nimg = 1_000_000, 1024
    print("bench: generating %d Plummer particles" % npart, file=sys.stderr, flush=True)
    data = synthetic.plummer_sphere(npart, seed=5)
    st = synthetic.Stream(55)
    m_target = data["m"] * np.exp(0.1 * st.normal(npart))
    ax = synthetic.linrange(-1.0, 1.0, nimg)
    S = nimg * nimg
    ctx = N.Context(0)
    apply_launch(ctx, args)
    L = N.lib()
    cat = N.CatalogArgs(sslots=[1], sflags=[0])
    frame = [np.ascontiguousarray(np.array(v, dtype=np.float64)) for v in
             ([0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0])]
    dax = Dev(ctx, 8 * nimg)
    dax.upload(ax)

    def build(cols_dict, prec):
        dc = upload_columns(ctx, cols_dict, ["x", "y", "z", "m", "h", "rho", "rho"])
        q = (C.c_void_p * 1)(dc[6].p.value)
        h = C.c_void_p()
        ctx.check(L.ph_lbvh_build_device(ctx.handle, npart, *[d.p for d in dc[:6]], 1, q, 0, 63, prec, C.byref(h)))
        for d in dc:
            d.free()
        return h

    def forward(h, prec, dout):
        ctx.check(L.ph_line_grid_device(ctx.handle, h, cat.ref(), N.PH_SCATTER, prec, *[N.ptr(f) for f in frame],
                                        dax.p, nimg, dax.p, nimg, dout.p))
        return ctx.stats()["interp_ms"]

    def backward(h, prec, dcot, dgrad, kernel=False):
        ctx.check(L.ph_line_grid_adjoint_device(ctx.handle, h, cat.ref(), N.PH_SCATTER, prec,
                                                *[N.ptr(f) for f in frame], dax.p, nimg, dax.p, nimg, dcot.p,
                                                dgrad.p))
        st = ctx.stats()
        return (st["interp_ms"], st["kernel_ms"]) if kernel else st["interp_ms"]

    def get(dev, dtype, count):
        out = np.zeros(count, dtype)
        ctx.check(L.ph_memcpy_dtoh(ctx.handle, out.ctypes.data_as(C.c_void_p), dev.p, out.nbytes))
        return out

    # Float32 forward, target, cotangent
    h32 = build(data, N.PH_PREC_FP32)
    hT = build(dict(data, m=m_target), N.PH_PREC_FP32)
    dmap, dT = Dev(ctx, 4 * S), Dev(ctx, 4 * S)
    forward(hT, N.PH_PREC_FP32, dT)
    L.ph_lbvh_destroy(hT)
    forward(h32, N.PH_PREC_FP32, dmap)
    sigma, target = get(dmap, np.float32, S), get(dT, np.float32, S)
    # rays without neighbours give NaN (as in the reference) in both maps: they
    # are left out of the loss, i.e. their cotangent is 0
    cot = np.nan_to_num(sigma - target, nan=0.0).astype(np.float32)
    dcot = Dev(ctx, 4 * S)
    dcot.upload(cot)
    dgrad = Dev(ctx, 8 * 6 * npart)

    def step():
        f = forward(h32, N.PH_PREC_FP32, dmap)
        b, bk = backward(h32, N.PH_PREC_FP32, dcot, dgrad, kernel=True)
        return {"forward": f, "backward": b, "backward_kernel": bk}

    ctx.set_count_neighbours(True)
    forward(h32, N.PH_PREC_FP32, dmap)
    n_acc = ctx.stats()["accepted"]
    ctx.set_count_neighbours(False)
    ms, res = timed_steps(ctx, step, args.steps, args.warmup, "adjoint1024")
    fwd = float(np.mean([r["forward"] for r in res]))
    bwd = float(np.mean([r["backward"] for r in res]))
    bwd_kernel = float(np.mean([r["backward_kernel"] for r in res]))
    g32 = get(dgrad, np.float64, 6 * npart).reshape(6, npart)
    loss = 0.5 * float(np.sum(cot.astype(np.float64) ** 2))

    # float64 check: adjoint vs central differences of the float64 loss
    h64 = build(data, N.PH_PREC_FP64)
    hT64 = build(dict(data, m=m_target), N.PH_PREC_FP64)
    dm64, dT64 = Dev(ctx, 8 * S), Dev(ctx, 8 * S)
    forward(hT64, N.PH_PREC_FP64, dT64)
    L.ph_lbvh_destroy(hT64)
    forward(h64, N.PH_PREC_FP64, dm64)
    T64 = get(dT64, np.float64, S)
    cot64 = np.nan_to_num(get(dm64, np.float64, S) - T64, nan=0.0)
    dcot64 = Dev(ctx, 8 * S)
    dcot64.upload(cot64)
    backward(h64, N.PH_PREC_FP64, dcot64, dgrad)
    g64 = get(dgrad, np.float64, 6 * npart).reshape(6, npart)
    L.ph_lbvh_destroy(h64)

    def loss64(dct):
        hh = build(dct, N.PH_PREC_FP64)
        forward(hh, N.PH_PREC_FP64, dm64)
        L.ph_lbvh_destroy(hh)
        return 0.5 * float(np.nansum((get(dm64, np.float64, S) - T64) ** 2))

    fd_err = {}
    fd_checks = 0
    if not args.no_fp64_compare:
        rng = np.random.default_rng(64)
        inside = np.nonzero((np.abs(data["x"]) < 0.9) & (np.abs(data["y"]) < 0.9))[0]
        picks = rng.choice(inside, 64, replace=False)
        for row, key in ((0, "m"), (2, "h"), (3, "x")):
            scale = float(np.sqrt(np.mean(g64[row] ** 2)))
            worst = 0.0
            for b in picks:
                eps = 1e-4 * (abs(data[key][b]) if key != "x" else data["h"][b])
                dp, dmn = dict(data), dict(data)
                dp[key] = data[key].copy()
                dmn[key] = data[key].copy()
                dp[key][b] += eps
                dmn[key][b] -= eps
                fd = (loss64(dp) - loss64(dmn)) / (2 * eps)
                err = abs(fd - g64[row, b]) / scale
                worst = max(worst, err) if np.isfinite(err) else float("inf")
                fd_checks += 1
            fd_err[key] = worst
        print("bench: adjoint vs central differences (max |fd - adj| / rms adj): %s" % fd_err, file=sys.stderr,
              flush=True)
    rel32 = {k: float(np.linalg.norm(g32[r] - g64[r]) / max(np.linalg.norm(g64[r]), 1e-300))
             for r, k in enumerate(("m", "A", "h", "x", "y", "z"))}
    result = {
        "metric": "GPU adjoint rays/s, Float32 column density 1024^2 (d/dm, dA, dh, dr); backward/forward time ratio",
        "value": S / (bwd * 1e-3), "unit": "rays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Plummer sphere, seeded mass perturbation, SplitMix64)",
        "config": {"workload": "plummer_1e6_M4_column_density_1024_adjoint", "particles": npart,
                   "image": [nimg, nimg], "kernel": "M4_spline", "precision": "fp32"},
        "forward_ms": fwd, "backward_ms": bwd, "backward_over_forward": bwd / fwd, "loss": loss,
        "fd_check": {"particles": 64, "checks": fd_checks, "step": "1e-4 relative (positions: 1e-4 h)",
                     "max_err_over_rms": fd_err},
        "fp32_vs_fp64_adjoint_normwise": rel32,
    }
    # per accepted (ray, particle) pair the 7 float32 fields read (x, y, z, h, m,
    # rho, A); per ray the float32 cotangent; per particle 6 float64 gradients
    alg_bytes = n_acc * 7 * 4 + S * 4 + npart * 6 * 8
    achieved = alg_bytes / (bwd_kernel * 1e-3) / 1e9
    result["accepted_neighbours"] = int(n_acc)
    result["backward_kernel_ms"] = bwd_kernel
    result["roofline"] = {"bound": "hbm", "kernel": "k_lines_adjoint", "achieved": achieved, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                          "algorithmic_bytes": int(alg_bytes)}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        t0 = time.perf_counter()
        o = O.OracleSet(data["x"], data["y"], data["z"], data["m"], data["h"], data["rho"], [data["rho"]],
                        kernel="M4_spline", code_bits=63, precision=32)
        t_build = time.perf_counter() - t0
        stride = 1 << 14
        while True:
            idx = np.arange(0, S, stride)
            i, j = idx % nimg, idx // nimg
            z = np.zeros(len(idx))
            org = [ax[i].astype(np.float32).astype(np.float64), ax[j].astype(np.float32).astype(np.float64), z]
            t0 = time.perf_counter()
            o.line_adjoint(1, False, org, [z, z, z + 1.0], cot[idx].astype(np.float64))
            dt = time.perf_counter() - t0
            if dt * 4 > args.cpu_seconds or stride <= 16:
                break
            stride //= 4
        rate = len(idx) / dt
        result["cpu_baseline"] = {
            "value": S / (t_build + S / rate), "unit": "rays/s", "cores": 1, "kind": "port",
            "sample": ("C++ oracle adjoint (serial: the rays share particles) on 1 host thread of a %s: LinearBVH! on "
                       "all %d particles (%.2f s) + the adjoint of %d of the %d rays (every %d-th, %.2f s); "
                       "whole-map time extrapolated" % (cpu_model(), npart, t_build, len(idx), S, stride, dt)),
            "build_s": t_build, "adjoint_rays_per_s": rate}
    L.ph_lbvh_destroy(h32)
    for d in (dax, dmap, dT, dcot, dgrad, dm64, dT64, dcot64):
        d.free()
    ctx.close()
    print(json.dumps(result))


def gridadjoint256(args):
    """Forward + adjoint of one output group (--output: the Shepard-normalised
    scalar A, grad A, div v or curl v) of the 2e6-particle Plummer sphere on the
    256^3 grid (C4, symmetric), float64."""
    npart, ng = NPART, NGRID
    print("bench: generating %d Plummer particles" % npart, file=sys.stderr, flush=True)
    data = synthetic.plummer_sphere(npart, seed=2)
    axes = synthetic.linrange(-1.0, 1.0, ng)
    S = ng ** 3
    ctx = N.Context(0)
    apply_launch(ctx, args)
    L = N.lib()
    catkw, width, fields, oname = ADJ_OUTPUTS[args.output]
    cat = N.CatalogArgs(**catkw)
    rows = 5 + len(fields)
    dax = Dev(ctx, 8 * ng)
    dax.upload(axes)
    prec = N.PH_PREC_FP64

    def build(dct):
        dc = upload_columns(ctx, dct, ["x", "y", "z", "m", "h", "rho", "rho", "A", "vx", "vy", "vz"])
        q = (C.c_void_p * 5)(*[d.p.value for d in dc[6:]])
        h = C.c_void_p()
        ctx.check(L.ph_lbvh_build_device(ctx.handle, npart, *[d.p for d in dc[:6]], 5, q, KERNEL_C4, 63, prec,
                                         C.byref(h)))
        for d in dc:
            d.free()
        return h

    dout = Dev(ctx, 8 * width * S)
    cot = synthetic.Stream(77).normal(width * S)  # component-major, as the forward's output rows
    dcot = Dev(ctx, 8 * width * S)
    dcot.upload(cot)
    dgrad = Dev(ctx, 8 * rows * npart)

    def forward(h):
        ctx.check(L.ph_structured_grid_device(ctx.handle, h, cat.ref(), N.PH_SYMMETRIC, prec, dax.p, ng, dax.p, ng,
                                              dax.p, ng, dout.p))
        return ctx.stats()["interp_ms"]

    def backward(h):
        ctx.check(L.ph_structured_grid_adjoint_device(ctx.handle, h, cat.ref(), N.PH_SYMMETRIC, prec, dax.p, ng,
                                                      dax.p, ng, dax.p, ng, dcot.p, dgrad.p))
        return ctx.stats()["interp_ms"]

    def get(dev, count):
        out = np.zeros(count)
        ctx.check(L.ph_memcpy_dtoh(ctx.handle, out.ctypes.data_as(C.c_void_p), dev.p, out.nbytes))
        return out

    h = build(data)

    def step():
        return {"forward": forward(h), "backward": backward(h)}

    ms, res = timed_steps(ctx, step, args.steps, args.warmup, "gridadjoint256")
    fwd = float(np.mean([r["forward"] for r in res]))
    bwd = float(np.mean([r["backward"] for r in res]))
    g = get(dgrad, rows * npart).reshape(rows, npart)
    L.ph_lbvh_destroy(h)

    def objective(dct):  # sum_s cot_s out_s over samples with neighbours
        hh = build(dct)
        forward(hh)
        L.ph_lbvh_destroy(hh)
        return float(np.nansum(cot * get(dout, width * S)))

    fd_err, fd_checks = {}, 0
    if not args.no_fp64_compare:
        rng = np.random.default_rng(256)
        picks = rng.choice(np.flatnonzero(np.abs(g[0]) > 0), 12, replace=False)
        for row, key in [(0, "m")] + [(1 + j, f) for j, f in enumerate(fields)] + [(rows - 4, "h"), (rows - 3, "x")]:
            scale = float(np.sqrt(np.mean(g[row] ** 2)))
            worst = 0.0
            for b in picks:
                eps = 1e-5 * (abs(data[key][b]) if key not in ("x", "vx", "vy", "vz") else data["h"][b])
                dp, dmn = dict(data), dict(data)
                dp[key] = data[key].copy()
                dmn[key] = data[key].copy()
                dp[key][b] += eps
                dmn[key][b] -= eps
                fd = (objective(dp) - objective(dmn)) / (2 * eps)
                err = abs(fd - g[row, b]) / scale
                worst = max(worst, err) if np.isfinite(err) else float("inf")
                fd_checks += 1
            fd_err[key] = worst
        print("bench: adjoint vs central differences (max |fd - adj| / rms adj): %s" % fd_err, file=sys.stderr,
              flush=True)
    result = {
        "metric": "GPU adjoint samples/s, float64 grid %s 256^3 (d/dm, dA, dh, dr); backward/forward time ratio" %
                  args.output,
        "value": S / (bwd * 1e-3), "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Plummer sphere, seeded cotangent, SplitMix64)",
        "config": {"workload": "plummer_2e6_C4_symmetric_grid256_adjoint", "particles": npart, "grid": [ng] * 3,
                   "kernel": "C4_Wendland", "strategy": "symmetric", "output": oname, "precision": "fp64"},
        "forward_ms": fwd, "backward_ms": bwd, "backward_over_forward": bwd / fwd,
        "fd_check": {"particles": 12, "checks": fd_checks, "step": "1e-5 relative (positions: 1e-5 h)",
                     "max_err_over_rms": fd_err},
    }
    for d in (dax, dout, dcot, dgrad):
        d.free()
    ctx.close()
    print(json.dumps(result))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (got --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "column2048":
        return column2048(args)
    if args.workload == "grid64":
        return grid64(args)
    if args.workload == "batch512":
        return batch512(args)
    if args.workload == "adjoint1024":
        return adjoint1024(args)
    if args.workload == "gridadjoint256":
        return gridadjoint256(args)
    prec = N.PH_PREC_FP64 if args.precision == "fp64" else N.PH_PREC_FP32C
    print("bench: generating %d Plummer particles" % NPART, file=sys.stderr, flush=True)
    data = synthetic.plummer_sphere(NPART, seed=2)
    axes = synthetic.linrange(-1.0, 1.0, NGRID)
    S = NGRID ** 3

    ctx = N.Context(0)
    ctx.set_traversal({"auto": N.PH_TRAVERSAL_AUTO, "lane": N.PH_TRAVERSAL_LANE,
                       "bucket": N.PH_TRAVERSAL_BUCKET}[args.traversal])
    apply_launch(ctx, args)
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
    fp32c_error = None
    if prec != N.PH_PREC_FP64 and not args.no_fp64_compare:
        got32 = np.zeros(L_OUT * S, np.float32)
        ctx.check(L.ph_memcpy_dtoh(ctx.handle, got32.ctypes.data_as(C.c_void_p), dout.p, got32.nbytes))
        step(N.PH_PREC_FP64)  # warm-up
        ctx.check(L.ph_synchronize(ctx.handle))
        t1 = time.perf_counter()
        k64 = [step(N.PH_PREC_FP64)[2] for _ in range(args.steps)]
        ctx.check(L.ph_synchronize(ctx.handle))
        ms64 = 1e3 * (time.perf_counter() - t1) / args.steps
        fp64 = {"ms_per_step": ms64, "samples_per_s": S / (ms64 * 1e-3), "accumulation_kernel_ms": float(np.mean(k64))}
        print("bench: float64 %.2f ms per step (accumulation kernel %.2f ms)" % (ms64, fp64["accumulation_kernel_ms"]),
              file=sys.stderr, flush=True)
        ref64 = np.zeros(L_OUT * S)
        ctx.check(L.ph_memcpy_dtoh(ctx.handle, ref64.ctypes.data_as(C.c_void_p), dout.p, ref64.nbytes))
        fp32c_error = field_errors(got32.reshape(L_OUT, S), ref64.reshape(L_OUT, S))
        del got32, ref64
        print("bench: FP32C vs float64 per field: %s" % ", ".join(
            "%s normwise %.2e max %.2e" % (k, v["normwise"], v["max_rel_pointwise"]) for k, v in fp32c_error.items()),
            file=sys.stderr, flush=True)
        # the same float64 steps with one bucket walk per sample (k_bucket_f64)
        ctx.set_traversal(N.PH_TRAVERSAL_BUCKET)
        step(N.PH_PREC_FP64)
        ctx.check(L.ph_synchronize(ctx.handle))
        t1 = time.perf_counter()
        kb = [step(N.PH_PREC_FP64)[2] for _ in range(args.steps)]
        ctx.check(L.ph_synchronize(ctx.handle))
        msb = 1e3 * (time.perf_counter() - t1) / args.steps
        fp64["per_lane_bucket_walk"] = {"ms_per_step": msb, "samples_per_s": S / (msb * 1e-3),
                                        "accumulation_kernel_ms": float(np.mean(kb))}
        # the same float64 steps with the reference-structured walk: one DFS per
        # sample over the LBVH, leaf by leaf (what a one-thread-per-sample port of
        # the reference's GPU kernel does)
        ctx.set_traversal(N.PH_TRAVERSAL_LANE)
        step(N.PH_PREC_FP64)
        ctx.check(L.ph_synchronize(ctx.handle))
        t1 = time.perf_counter()
        kl = [step(N.PH_PREC_FP64)[2] for _ in range(args.steps)]
        ctx.check(L.ph_synchronize(ctx.handle))
        msl = 1e3 * (time.perf_counter() - t1) / args.steps
        ctx.set_traversal({"auto": N.PH_TRAVERSAL_AUTO, "lane": N.PH_TRAVERSAL_LANE,
                           "bucket": N.PH_TRAVERSAL_BUCKET}[args.traversal])
        fp64["per_leaf_walk"] = {"ms_per_step": msl, "samples_per_s": S / (msl * 1e-3),
                                 "accumulation_kernel_ms": float(np.mean(kl))}
        print("bench: float64, per-leaf walk %.2f ms per step (kernel %.2f ms)" % (msl, float(np.mean(kl))),
              file=sys.stderr, flush=True)

    fbytes = 8 if prec == N.PH_PREC_FP64 else 4
    alg_bytes = n_acc * FIELDS_PER_NEIGHBOUR * fbytes + S * (3 + L_OUT) * fbytes
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    # the accumulation kernel the call ran: FP32C grids walk one bucket-tree
    # union per wave (k_tile_c32) unless the per-lane bucket walk is forced
    if prec == N.PH_PREC_FP64:
        kname = {"lane": "k_points", "bucket": "k_bucket_f64"}.get(args.traversal, "k_tile_f64")
    else:
        kname = "k_bucket_c32" if args.traversal == "bucket" else "k_tile_c32"
    traffic, traffic_src = measured_traffic(kname)
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
        "launch": args.launch,
        "autotune": [{k: e[k] for k in ("occupancy", "xcd_map", "tile_order", "bucket_leaves", "block_size", "best_ms",
                                         "default_ms")}
                     for e in ctx.autotune_report()],
        "fp64": fp64,
        "speedup_vs_fp64": (fp64["ms_per_step"] / ms_per_step) if fp64 else None,
        "speedup_vs_fp64_kernel": (fp64["accumulation_kernel_ms"] / kernel_ms) if fp64 else None,
        "fp32c_error_vs_fp64": fp32c_error,
        "speedup_vs_fp64_per_leaf_walk": (fp64["per_leaf_walk"]["ms_per_step"] / ms_per_step) if fp64 else None,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
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
