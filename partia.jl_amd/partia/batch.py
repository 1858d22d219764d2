"""Batched sampling: many PointSamples requests of uneven size over one tree in
one graph-recorded launch of a persistent work-queue kernel (ph_batch_*).

Each request computes what one external-LBVH PointSamples_interpolation call
(KI/grid_interpolation/drivers/PointSamples_driver.jl:98-118) computes, with
identical outputs.  Samples are a device point list, the frame-based
Cartesian plane / box of PointSamples(Cartesian, frame, xparams, yparams[,
zparams]) or the polar plane / cylinder of PointSamples(Polar, frame, sparams,
phiparams) / PointSamples(Cylindrical, frame, sparams, phiparams, zparams)
(Grids/grids/PointSamples.jl:328-397), generated on the device.  Outputs stay
on the device (L x S per request).
"""
import ctypes as C

import numpy as np

from . import _native as N


def points(px, py, pz, count, out):
    """Request over `count` device points (float64 device pointers)."""
    r = N.ph_request()
    r.kind = N.PH_REQ_POINTS
    r.count = int(count)
    r.px, r.py, r.pz = _p(px), _p(py), _p(pz)
    r.out = _p(out)
    return r


def plane(position, right, up, xparams, yparams, out):
    """PointSamples(Cartesian, frame, xparams, yparams): nx * ny samples, x fastest."""
    return _frame(N.PH_REQ_PLANE, position, right, up, (0.0, 0.0, 0.0), (xparams, yparams, (0.0, 1.0, 2)), out)


def box(position, right, up, forward, xparams, yparams, zparams, out):
    """PointSamples(Cartesian, frame, xparams, yparams, zparams): nx * ny * nz samples."""
    return _frame(N.PH_REQ_BOX, position, right, up, forward, (xparams, yparams, zparams), out)


def polar(position, right, up, sparams, phiparams, out, trig=None):
    """PointSamples(Polar, frame, sparams, phiparams): ns * nphi samples, s fastest,
    phi half-open.  trig: optional (nphi, 2) array of (sin phi_j, cos phi_j), e.g.
    the caller's own sincos values; by default the library uses the C library's."""
    r = _frame(N.PH_REQ_POLAR, position, right, up, (0.0, 0.0, 0.0), (sparams, phiparams, (0.0, 1.0, 2)), out)
    return _trig(r, trig)


def cylinder(position, right, up, forward, sparams, phiparams, zparams, out, trig=None):
    """PointSamples(Cylindrical, frame, sparams, phiparams, zparams): ns * nphi * nz samples."""
    r = _frame(N.PH_REQ_CYLINDER, position, right, up, forward, (sparams, phiparams, zparams), out)
    return _trig(r, trig)


def sample_count(r):
    if r.kind == N.PH_REQ_POINTS:
        return int(r.count)
    if r.kind in (N.PH_REQ_PLANE, N.PH_REQ_POLAR):
        return int(r.n[0]) * int(r.n[1])
    return int(r.n[0]) * int(r.n[1]) * int(r.n[2])


class Batch:
    """ph_batch: create once, run many times."""

    def __init__(self, ctx, tree_handle, catalog, strategy, precision, requests):
        self.ctx = ctx
        self.nreq = len(requests)
        self._reqs = (N.ph_request * max(1, self.nreq))(*requests)
        self._cat = catalog
        self.h = C.c_void_p()
        N.register_tree(self)
        ctx.check(N.lib().ph_batch_create(ctx.handle, tree_handle, catalog.ref(), int(strategy), int(precision),
                                          self.nreq, self._reqs, C.byref(self.h)))

    def run(self):
        self.ctx.check(N.lib().ph_batch_run(self.h))

    def times(self):
        """(first_us, last_us) per request of the last run, after the launch began."""
        a, b = np.zeros(max(1, self.nreq)), np.zeros(max(1, self.nreq))
        self.ctx.check(N.lib().ph_batch_times(self.h, N.ptr(a), N.ptr(b)))
        return a[:self.nreq], b[:self.nreq]

    def close(self):
        if getattr(self, "h", None):
            N.lib().ph_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _p(v):
    if v is None:
        return None
    if isinstance(v, C.c_void_p):
        return v.value
    return int(v)


def _trig(r, trig):
    """The table is read by ph_batch_create; the request object keeps it alive."""
    if trig is not None:
        t = np.ascontiguousarray(np.asarray(trig, dtype=np.float64).reshape(-1, 2))
        r.trig = t.ctypes.data
        r.trig_table = t
    return r


def _frame(kind, position, right, up, forward, params, out):
    r = N.ph_request()
    r.kind = kind
    for name, v in (("position", position), ("right", right), ("up", up), ("forward", forward)):
        getattr(r, name)[:] = [float(x) for x in v]
    for a, (lo, hi, n) in enumerate(params):
        r.amin[a], r.amax[a], r.n[a] = float(lo), float(hi), int(n)
    r.out = _p(out)
    return r
