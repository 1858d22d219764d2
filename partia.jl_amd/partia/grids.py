"""Sample containers: PointSamples, LineSamples, StructuredGrid, GridBundle.

Mirrors Grids/grids/PointSamples.jl:25-65, LineSamples.jl:31-75,
StructuredGrid.jl:19-23,74-76,140-222,285-291 and transform.jl:49-112, and
Grids/griddataset/GridBundle.jl:1-4.  Arrays follow Julia's column-major layout:
a StructuredGrid's values are an order='F' numpy array and flattening is x
fastest.  Non-Cartesian coordinate systems are converted on the host.
"""
import math

import numpy as np

from .errors import ArgumentError
from .synthetic import linrange


class AbstractCoordinateSystem:
    pass


class Cartesian(AbstractCoordinateSystem):
    pass


class Polar(AbstractCoordinateSystem):
    pass


class Cylindrical(AbstractCoordinateSystem):
    pass


class Spherical(AbstractCoordinateSystem):
    pass


class ParallelBeam:
    """Beam-model tag of the frame-plane LineSamples constructors: every line
    along the frame's forward direction (Grids/grids/LineSamples.jl:441-537)."""


class Frame:
    """Frame(x, f0, u0) (Frames/struct/Frame.jl:42-78) in its initial
    orientation (identity quaternion): forward and up normalised (hypot), their
    dot product within sqrt(eps) of 0, right = forward x up normalised.  The
    accessors frame_position / frame_right / frame_up / frame_forward
    (Frame.jl:145-186) return the basis vectors as tuples."""

    def __init__(self, x, f0, u0):
        x, f0, u0 = (tuple(float(v) for v in a) for a in (x, f0, u0))
        nf0, nu0 = math.hypot(*f0), math.hypot(*u0)
        if nf0 == 0.0:
            raise ArgumentError("f0 must be nonzero.")
        if nu0 == 0.0:
            raise ArgumentError("u0 must be nonzero.")
        inf0, inu0 = 1.0 / nf0, 1.0 / nu0
        f = tuple(v * inf0 for v in f0)
        u = tuple(v * inu0 for v in u0)
        if abs(f[0] * u[0] + f[1] * u[1] + f[2] * u[2]) > math.sqrt(np.finfo(np.float64).eps):
            raise ArgumentError("f0 and u0 must be orthogonal.")
        r = (f[1] * u[2] - f[2] * u[1], f[2] * u[0] - f[0] * u[2], f[0] * u[1] - f[1] * u[0])
        nr = math.sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])
        self.x, self.f0, self.u0 = x, f, u
        self.r0 = tuple(v / nr for v in r)

    def position(self):
        return self.x

    def right(self):
        return self.r0

    def up(self):
        return self.u0

    def forward(self):
        return self.f0


def frame_position(frame):
    return frame.position()


def frame_right(frame):
    return frame.right()


def frame_up(frame):
    return frame.up()


def frame_forward(frame):
    return frame.forward()


class AbstractGrid:
    pass


class AbstractSamples(AbstractGrid):
    pass


class PointSamples(AbstractSamples):
    def __init__(self, grid, coor):
        coor = tuple(coor)
        n = len(grid)
        for d, c in enumerate(coor):
            if len(c) != n:
                raise ArgumentError("coor[%d] length mismatch" % (d + 1))
        self.grid = grid
        self.coor = coor

    @classmethod
    def from_frame(cls, coord, frame, *params):
        """PointSamples(COORD, frame, params...) (Grids/grids/PointSamples.jl:328-418):
        Cartesian plane (xparams, yparams) or box (+ zparams), Polar plane
        (sparams, phiparams) or Cylindrical (+ zparams) in the frame's basis;
        values zero."""
        pos, r, u, f = frame.position(), frame.right(), frame.up(), frame.forward()
        if coord is Cartesian and len(params) == 2:
            c = cartesian_plane_coordinates(pos, r, u, *params)
        elif coord is Cartesian and len(params) == 3:
            c = cartesian_box_coordinates(pos, r, u, f, *params)
        elif coord is Polar and len(params) == 2:
            c = polar_plane_coordinates(pos, r, u, *params)
        elif coord is Cylindrical and len(params) == 3:
            c = cylindrical_coordinates(pos, r, u, f, *params)
        else:
            raise ArgumentError("no frame-based PointSamples constructor for these arguments")
        return cls(np.zeros(len(c[0])), c)

    def similar(self, dtype=None):
        """similar(grid) shares the coordinates; similar(grid, T) copies them as T."""
        if dtype is None or np.dtype(dtype) == self.grid.dtype:
            return PointSamples(np.empty_like(self.grid), self.coor)
        return PointSamples(np.empty(len(self.grid), dtype=dtype), tuple(np.asarray(c, dtype=dtype) for c in self.coor))

    @property
    def dtype(self):
        return self.grid.dtype

    def __len__(self):
        return len(self.grid)


def angle_table(phiparams):
    """(nphi, 2) array of (sin phi_j, cos phi_j), phi_j = phimin + (j - 1) dphi,
    dphi = (phimax - phimin) / nphi (half-open), with Python's math.sin /
    math.cos (the C library's, as ph_batch_create uses without a table)."""
    pmin, pmax, nphi = float(phiparams[0]), float(phiparams[1]), int(phiparams[2])
    dphi = (pmax - pmin) / float(nphi)
    phi = [pmin + float(j) * dphi for j in range(nphi)]
    return np.array([[math.sin(a), math.cos(a)] for a in phi], dtype=np.float64).reshape(nphi, 2)


def _check_polar(sparams, phiparams):
    smin, smax, ns = float(sparams[0]), float(sparams[1]), int(sparams[2])
    pmin, pmax, nphi = float(phiparams[0]), float(phiparams[1]), int(phiparams[2])
    two_pi = 2.0 * math.pi
    if not smin >= 0.0:
        raise ArgumentError("smin must be nonnegative.")
    if not smax > smin:
        raise ArgumentError("smax must be greater than smin.")
    near = abs(pmax - two_pi) <= math.sqrt(np.finfo(np.float64).eps) * max(abs(pmax), two_pi)
    if not (pmin >= 0.0 and (pmax <= two_pi or near) and pmax > pmin):
        raise ArgumentError("angular range must satisfy 0 ≤ ϕmin < ϕmax ≤ 2π.")
    if ns < 2:
        raise ArgumentError("ns must be at least 2.")
    if nphi < 1:
        raise ArgumentError("nϕ must be at least 1.")
    return smin, smax, ns, nphi


def polar_plane_coordinates(position, right, up, sparams, phiparams, trig=None):
    """_polar_plane_coordinates (Grids/grids/coordinate.jl:153-206): s = smin +
    (i - 1) ds, ds = (smax - smin) / (ns - 1); xi = s cos phi_j, eta = s sin phi_j;
    x = x0 + xi rx + eta ux; sample k = i + (j - 1) ns.  trig: the (sin, cos)
    table (angle_table by default)."""
    x0, r, u = (np.asarray(v, dtype=np.float64) for v in (position, right, up))
    smin, smax, ns, nphi = _check_polar(sparams, phiparams)
    t = angle_table(phiparams) if trig is None else np.asarray(trig, dtype=np.float64).reshape(nphi, 2)
    s = smin + np.arange(ns, dtype=np.float64) * ((smax - smin) / float(ns - 1))
    XI = (s[None, :] * t[:, 1:2]).ravel()  # s fastest, then phi
    ETA = (s[None, :] * t[:, 0:1]).ravel()
    return tuple((x0[a] + XI * r[a]) + ETA * u[a] for a in range(3))


def cylindrical_coordinates(position, right, up, forward, sparams, phiparams, zparams, trig=None):
    """_cylindrical_coordinates (Grids/grids/coordinate.jl:208-276): the polar
    plane plus zeta fx, zeta = zmin + (k3 - 1) dz; s fastest, then phi, z."""
    x0, r, u, f = (np.asarray(v, dtype=np.float64) for v in (position, right, up, forward))
    smin, smax, ns, nphi = _check_polar(sparams, phiparams)
    zmin, zmax, nz = float(zparams[0]), float(zparams[1]), int(zparams[2])
    if not zmax > zmin:
        raise ArgumentError("zmax must be greater than zmin.")
    if nz < 2:
        raise ArgumentError("nz must be at least 2.")
    t = angle_table(phiparams) if trig is None else np.asarray(trig, dtype=np.float64).reshape(nphi, 2)
    s = smin + np.arange(ns, dtype=np.float64) * ((smax - smin) / float(ns - 1))
    zeta = zmin + np.arange(nz, dtype=np.float64) * ((zmax - zmin) / float(nz - 1))
    XI = np.tile((s[None, :] * t[:, 1:2]).ravel(), nz)
    ETA = np.tile((s[None, :] * t[:, 0:1]).ravel(), nz)
    ZE = np.repeat(zeta, ns * nphi)
    return tuple(((x0[a] + XI * r[a]) + ETA * u[a]) + ZE * f[a] for a in range(3))


def _check_axis(name, params):
    lo, hi, n = params
    if not hi > lo:
        raise ArgumentError("%smax must be greater than %smin." % (name, name))
    if int(n) < 2:
        raise ArgumentError("n%s must be at least 2." % name)
    return float(lo), float(hi), int(n)


def cartesian_plane_coordinates(position, right, up, xparams, yparams):
    """_cartesian_plane_coordinates (Grids/grids/coordinate.jl:44-86) from the
    frame's position / right / up vectors (Frames/struct/Frame.jl:145-186):
    xi = xmin + (i - 1) dx, dx = (xmax - xmin) / (nx - 1), x = x0 + xi rx + eta ux,
    sample k = i + (j - 1) nx.  numpy evaluates each operation rounded (no fused
    multiply-add), as Julia does."""
    x0, r, u = (np.asarray(v, dtype=np.float64) for v in (position, right, up))
    xmin, xmax, nx = _check_axis("x", xparams)
    ymin, ymax, ny = _check_axis("y", yparams)
    dx, dy = (xmax - xmin) / float(nx - 1), (ymax - ymin) / float(ny - 1)
    xi = xmin + np.arange(nx, dtype=np.float64) * dx
    eta = ymin + np.arange(ny, dtype=np.float64) * dy
    XI, ETA = np.meshgrid(xi, eta, indexing="xy")  # x fastest
    XI, ETA = XI.ravel(), ETA.ravel()
    return tuple((x0[a] + XI * r[a]) + ETA * u[a] for a in range(3))


def cartesian_box_coordinates(position, right, up, forward, xparams, yparams, zparams):
    """_cartesian_box_coordinates (Grids/grids/coordinate.jl:88-143): as the
    plane plus zeta fx along the frame's forward vector; x fastest, then y, z."""
    x0, r, u, f = (np.asarray(v, dtype=np.float64) for v in (position, right, up, forward))
    xmin, xmax, nx = _check_axis("x", xparams)
    ymin, ymax, ny = _check_axis("y", yparams)
    zmin, zmax, nz = _check_axis("z", zparams)
    dx, dy, dz = (xmax - xmin) / float(nx - 1), (ymax - ymin) / float(ny - 1), (zmax - zmin) / float(nz - 1)
    xi = xmin + np.arange(nx, dtype=np.float64) * dx
    eta = ymin + np.arange(ny, dtype=np.float64) * dy
    zeta = zmin + np.arange(nz, dtype=np.float64) * dz
    ZE, ET, XI = np.meshgrid(zeta, eta, xi, indexing="ij")  # x fastest
    XI, ET, ZE = XI.ravel(), ET.ravel(), ZE.ravel()
    return tuple(((x0[a] + XI * r[a]) + ET * u[a]) + ZE * f[a] for a in range(3))


class LineSamples(AbstractSamples):
    def __init__(self, grid, origin, direction):
        origin, direction = tuple(origin), tuple(direction)
        n = len(grid)
        for d in range(len(origin)):
            if len(origin[d]) != n:
                raise ArgumentError("origin[%d] length mismatch" % (d + 1))
            if len(direction[d]) != n:
                raise ArgumentError("direction[%d] length mismatch" % (d + 1))
        self.grid = grid
        self.origin = origin
        self.direction = direction

    @classmethod
    def from_frame(cls, coord, beam, frame, *params):
        """LineSamples(COORD, ParallelBeam, frame, params...)
        (Grids/grids/LineSamples.jl:462-537): origins on the frame's Cartesian
        (xparams, yparams) or Polar (sparams, phiparams) plane, every direction
        the frame's forward vector; values zero."""
        if beam is not ParallelBeam:
            raise ArgumentError("only the ParallelBeam model is defined")
        if not (coord in (Cartesian, Polar) and len(params) == 2):
            raise ArgumentError("no frame-based LineSamples constructor for these arguments")
        o = PointSamples.from_frame(coord, frame, *params).coor
        n = len(o[0])
        d = tuple(np.full(n, v) for v in frame.forward())
        return cls(np.zeros(n), o, d)

    def similar(self, dtype=None):
        if dtype is None or np.dtype(dtype) == self.grid.dtype:
            return LineSamples(np.empty_like(self.grid), self.origin, self.direction)
        return LineSamples(np.empty(len(self.grid), dtype=dtype), tuple(np.asarray(c, dtype=dtype) for c in self.origin),
                           tuple(np.asarray(c, dtype=dtype) for c in self.direction))

    @property
    def dtype(self):
        return self.grid.dtype

    def __len__(self):
        return len(self.grid)


class StructuredGrid(AbstractGrid):
    """StructuredGrid{D, TF}: values on the tensor product of D axes."""

    def __init__(self, *args, dtype=None):
        if len(args) >= 1 and isinstance(args[0], type) and issubclass(args[0], AbstractCoordinateSystem):
            coord, params = args[0], args[1:]
            self._from_params(coord, params, dtype)
        else:
            grid, axes, size = args
            self.grid = grid
            self.axes = tuple(axes)
            self.size = tuple(size)

    def _from_params(self, coord, params, dtype):
        if dtype is None:
            dtype = np.float64
            for p in params:
                if isinstance(p, tuple) and isinstance(p[0], np.float32):
                    dtype = np.float32
        T = np.dtype(dtype).type
        if coord is Cartesian:
            axes = tuple(linrange(T(p[0]), T(p[1]), int(p[2]), dtype) for p in params)
        elif coord in (Polar, Cylindrical):
            (smin, smax, ns), (pmin, pmax, npp) = params[0], params[1]
            if not (ns >= 1 and npp >= 1):
                raise ArgumentError("ns and nϕ must be ≥ 1")
            if not (smin >= 0 and smax > smin):
                raise ArgumentError("radial range must satisfy 0 ≤ smin < smax")
            if not (pmin >= 0 and (pmax <= 2 * np.pi or np.isclose(pmax, 2 * np.pi)) and pmax > pmin):
                raise ArgumentError("angular range must satisfy 0 ≤ ϕmin < ϕmax ≤ 2π")
            saxes = linrange(T(smin), T(smax), ns, dtype)
            dphi = (T(pmax) - T(pmin)) / T(npp)
            paxes = (T(pmin) + np.arange(npp, dtype=dtype) * dphi).astype(dtype)
            axes = (saxes, paxes)
            if coord is Cylindrical:
                zp = params[2]
                if isinstance(zp, tuple):
                    zmin, zmax, nz = zp
                    if not zmax > zmin:
                        raise ArgumentError("axial range must satisfy zmin < zmax")
                    axes = axes + (linrange(T(zmin), T(zmax), nz, dtype),)
                else:
                    axes = axes + (np.array([T(zp)], dtype=dtype),)
        elif coord is Spherical:
            (rmin, rmax, nr), (pmin, pmax, npp), (tmin, tmax, nt) = params
            axes = (linrange(T(rmin), T(rmax), nr, dtype), linrange(T(pmin), T(pmax), npp, dtype),
                    linrange(T(tmin), T(tmax), nt, dtype))
        else:
            raise ArgumentError("unknown coordinate system")
        self.axes = axes
        self.size = tuple(len(a) for a in axes)
        self.grid = np.zeros(self.size, dtype=dtype, order="F")

    @property
    def dtype(self):
        return self.grid.dtype

    def similar(self, dtype=None):
        if dtype is None or np.dtype(dtype) == self.grid.dtype:
            return StructuredGrid(np.zeros(self.size, dtype=self.grid.dtype, order="F"), self.axes, self.size)
        return StructuredGrid(np.zeros(self.size, dtype=dtype, order="F"),
                              tuple(np.asarray(a, dtype=dtype) for a in self.axes), self.size)

    def __len__(self):
        return int(np.prod(self.size))


def _mesh(axes):
    return np.meshgrid(*axes, indexing="ij")


def trig_pairs(values):
    """(sin v, cos v) pairs of axis values, flattened (math.sin / math.cos)."""
    return np.array([f(float(v)) for v in values for f in (math.sin, math.cos)], dtype=np.float64)


def device_axes(coord, grid):
    """Axes in the layout of ph_structured_grid_coords: angle axes as
    (sin, cos) pairs.  Returns (coords id, a1, a2, a3)."""
    from . import _native as N
    a = [np.asarray(x, dtype=np.float64) for x in grid.axes]
    if coord is Cylindrical:
        return N.PH_COORDS_CYLINDRICAL, a[0], trig_pairs(a[1]), a[2]
    if coord is Spherical:
        return N.PH_COORDS_SPHERICAL, a[0], trig_pairs(a[1]), trig_pairs(a[2])
    return N.PH_COORDS_CARTESIAN, a[0], a[1], a[2]


def coordinates_from_pairs(coord, grid):
    """coordinate_grid with the sin / cos of device_axes, in the reference's
    order of operations (the values the device generator produces)."""
    c, a1, a2, a3 = device_axes(coord, grid)
    n1, n2, n3 = grid.size
    if c == 0:
        return coordinate_grid(coord, grid)
    I, J, K = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    I, J, K = I.ravel(order="F"), J.ravel(order="F"), K.ravel(order="F")
    sp, cp = a2[2 * J], a2[2 * J + 1]
    if coord is Cylindrical:
        s = a1[I]
        return (s * cp, s * sp, a3[K])
    r, st, ct = a1[I], a3[2 * K], a3[2 * K + 1]
    return ((r * st) * cp, (r * st) * sp, r * ct)


def coordinate_grid(coord, grid):
    """Flattened (column-major) Cartesian coordinates of every grid point
    (StructuredGrid.jl:140-222)."""
    m = [a.ravel(order="F") for a in _mesh(grid.axes)]
    if coord is Cartesian:
        return tuple(m)
    if coord is Polar:
        s, phi = m
        return (s * np.cos(phi), s * np.sin(phi))
    if coord is Cylindrical:
        s, phi, z = m
        return (s * np.cos(phi), s * np.sin(phi), z)
    if coord is Spherical:
        r, phi, theta = m
        return (r * np.sin(theta) * np.cos(phi), r * np.sin(theta) * np.sin(phi), r * np.cos(theta))
    raise ArgumentError("unknown coordinate system")


def flatten(coord, grid):
    """flatten(COORD, grid) -> PointSamples sharing the grid values (transform.jl:49-67)."""
    return PointSamples(grid.grid.ravel(order="F"), coordinate_grid(coord, grid))


def restore_struct(coord, points, axes, atol=1e-8, rtol=1e-8):
    """restore_struct(COORD, points, axes) (transform.jl:103-112): checks the
    coordinates against the axes, then reshapes the values."""
    size = tuple(len(a) for a in axes)
    expected = coordinate_grid(coord, StructuredGrid(np.zeros(size, order="F"), axes, size))
    for got, exp in zip(points.coor, expected):
        if got.shape != exp.shape or not np.allclose(got, exp, atol=atol, rtol=rtol):
            raise ArgumentError("PointSamples coordinates do not match the provided axes under %s dispatch"
                                % coord.__name__)
    return StructuredGrid(np.reshape(points.grid, size, order="F"), axes, size)


class GridBundle:
    """GridBundle{L, G}: the grids returned by every driver, with their names."""

    def __init__(self, grids, names):
        self.grids = tuple(grids)
        self.names = tuple(names)

    def __len__(self):
        return len(self.grids)

    def __getitem__(self, key):
        if isinstance(key, str):
            return self.grids[self.names.index(key)]
        return self.grids[key]

    def datatype(self):
        return self.grids[0].dtype if self.grids else np.dtype(np.float64)
