"""GridDataset files (IO/gridsIO/write_grids.jl, IO/gridsIO/read_grids.jl):
the grids an interpolation call returns, their names and metadata, in HDF5.

Layout of a file, as the reference writes it:

    params/<key>           one dataset per metadata entry
    data/names             the grid names (strings)
    data/coord/<x|y|z>     PointSamples: the shared coordinate vectors, or
    data/axes/<x|y|z>      StructuredGrid: the shared axes
    data/grids/<i>         grid i (1-based), a StructuredGrid's values as a
                           Julia array (first axis fastest)

The HDF5 encoding is this package's own (partia.hdf5): neither libhdf5 nor
h5py is available to it.  Differences from HDF5.jl's files: strings are
stored as fixed-length UTF-8 (HDF5.jl writes vectors of strings as
variable-length strings) and booleans as an 8-bit FALSE/TRUE enumeration.
"""
import datetime
import platform

import numpy as np

from . import hdf5
from .errors import ArgumentError, MethodError
from .grids import GridBundle, PointSamples, StructuredGrid

_COORDINATE = ("x", "y", "z")
# the reference package's version (Project.toml:3), recorded in file identifiers
PARTIA_VERSION = "0.1.3"


class GridDataset:
    """GridDataset{L, TF, G} (Grids/griddataset/GridDataset.jl:3-6): a
    GridBundle and a dictionary of metadata (values: str, int, bool or float)."""

    def __init__(self, data, params):
        self.data = data
        self.params = dict(params)


def file_identifier(operation_name):
    """file_identifier (Tools/logging.jl:153-157):
    "FT:Partia:<version> (<operation name or CommonOutput>): dd/mm/yyyy HH:MM:SS.s"."""
    t = datetime.datetime.now()
    stamp = t.strftime("%d/%m/%Y %H:%M:%S") + ".%d" % (t.microsecond // 1000)
    return "FT:Partia:%s (%s): %s" % (PARTIA_VERSION, operation_name or "CommonOutput", stamp)


def _dtype_name(dt):
    dt = np.dtype(dt)
    if dt == np.float64:
        return "Float64"
    if dt == np.float32:
        return "Float32"
    if dt == np.float16:
        return "Float16"
    raise ArgumentError("grids must hold Float16, Float32 or Float64 values (got %s)" % dt)


def _write_params(root, params):
    pg = root.create_group("params")
    for k, v in params.items():
        pg.write(str(k), v)


def write_GridDataset(gd, filename="PartiaGridDataset.h5"):
    """write_GridDataset (IO/gridsIO/write_grids.jl:29-70 for PointSamples,
    :98-142 for StructuredGrid): every grid must share the first grid's
    coordinate (axes) vectors, by identity, or ArgumentError."""
    grids = gd.data.grids
    if not grids:
        raise ArgumentError("GridDataset holds no grids")
    g1 = grids[0]
    root = hdf5.Group()
    if all(isinstance(g, PointSamples) for g in grids):
        for i, gi in enumerate(grids[1:], start=2):
            for d in range(len(g1.coor)):
                if gi.coor[d] is not g1.coor[d]:
                    raise ArgumentError("PointSamples coor not shared: grid=%d dim=%d" % (i, d + 1))
        _write_params(root, gd.params)
        dg = root.create_group("data")
        dg.write("names", [str(n) for n in gd.data.names])
        cg = dg.create_group("coord")
        for d, c in enumerate(g1.coor):
            cg.write(_COORDINATE[d], np.asarray(c))
        gg = dg.create_group("grids")
        for i, gi in enumerate(grids, start=1):
            gg.write(str(i), np.asarray(gi.grid))
    elif all(isinstance(g, StructuredGrid) for g in grids):
        for i, gi in enumerate(grids[1:], start=2):
            for d in range(len(g1.axes)):
                if gi.axes[d] is not g1.axes[d]:
                    raise ArgumentError("StructuredGrid axes not shared: grid=%d dim=%d" % (i, d + 1))
        _write_params(root, gd.params)
        dg = root.create_group("data")
        dg.write("names", [str(n) for n in gd.data.names])
        ag = dg.create_group("axes")
        for d, a in enumerate(g1.axes):
            ag.write(_COORDINATE[d], np.asarray(a))
        gg = dg.create_group("grids")
        for i, gi in enumerate(grids, start=1):
            gg.write(str(i), np.asarray(gi.grid).reshape(gi.size, order="F"), column_major=True)
    else:
        raise MethodError("write_GridDataset: no method for a GridDataset of %s"
                          % ", ".join(sorted({type(g).__name__ for g in grids})))
    hdf5.write(filename, root)


def write_GridBundle(gb, filename="PartiaGridDataset.h5", *, code_units, operation_name="", params=None):
    """write_GridBundle (IO/gridsIO/write_grids.jl:179-215): adds the schema,
    system, data type, grid type, code units and file identifier to the
    metadata and writes the GridDataset.  code_units must hold :umass, :udist,
    :utime and :umagfd, or ArgumentError."""
    p = {} if params is None else dict(params)
    required = ("umass", "udist", "utime", "umagfd")
    keys = {str(k) for k in code_units}
    notfound = tuple(k for k in required if k not in keys)
    if notfound:  # Julia prints the tuple of symbols, (:umass, :udist) or (:umagfd,)
        inner = ", ".join(":" + k for k in notfound) + ("," if len(notfound) == 1 else "")
        raise ArgumentError("code_units missing required keys: (%s)" % inner)
    p["schema_name"] = "GridDataset"
    p["schema_id"] = "GridDataset.v1"
    p["system_kernel"] = platform.system()
    p["dtype"] = _dtype_name(gb.datatype())
    p["grid_type"] = type(gb.grids[0]).__name__ if gb.grids else "AbstractGrid"
    for k, v in code_units.items():
        p[str(k)] = v
    p["file_identifier"] = file_identifier(operation_name)
    write_GridDataset(GridDataset(gb, p), filename)


def _parse_float_type(s):
    s = s.strip()
    table = {"Float16": np.float16, "Float32": np.float32, "Float64": np.float64}
    if s not in table:
        raise ArgumentError("Unsupported dtype string in params[:dtype]: %s" % s)
    return table[s]


def _ordered_keys(g):
    ks = g.keys()
    if all(k in ks for k in _COORDINATE):
        return list(_COORDINATE[:len(ks)])
    return sorted(ks)


def _read_shared_vectors(dg, groupname):
    if groupname not in dg:
        raise ArgumentError("Schema error: missing data/%s" % groupname)
    vg = dg[groupname]
    vecs = tuple(np.asarray(vg[k]) for k in _ordered_keys(vg))
    return vecs, len(vecs)


def _read_dense_grids(dg, TF):
    if "grids" not in dg:
        raise ArgumentError("Schema error: missing data/grids")
    gg = dg["grids"]
    L = len(gg.keys())
    # a stored array's dataspace lists the Julia dimensions reversed
    return [np.asarray(gg[str(i)], dtype=TF).T for i in range(1, L + 1)], L


def read_GridDataset(filename="PRGridDataset.h5"):
    """read_GridDataset (IO/gridsIO/read_grids.jl:30-81): the GridDataset of a
    file written by write_GridDataset, with the shared coordinate / axes vectors
    read once and shared by every grid."""
    f = hdf5.read(filename)
    if "params" not in f:
        raise ArgumentError("Schema error: missing /params")
    if "data" not in f:
        raise ArgumentError("Schema error: missing /data")
    params = {k: f["params"][k] for k in f["params"].keys()}
    dg = f["data"]
    if "dtype" not in params:
        raise ArgumentError("params missing :dtype")
    if "grid_type" not in params:
        raise ArgumentError("params missing :grid_type")
    TF = _parse_float_type(str(params["dtype"]))
    base = str(params["grid_type"])
    names = tuple(dg["names"]) if isinstance(dg["names"], list) else (str(dg["names"]),)
    if base == "StructuredGrid":
        shared, D = _read_shared_vectors(dg, "axes")
        size = tuple(len(v) for v in shared)
        arrays, L = _read_dense_grids(dg, TF)
        if len(names) != L:
            raise ArgumentError("names length != number of grids")
        grids = tuple(StructuredGrid(np.asfortranarray(a), shared, size) for a in arrays)
    elif base == "PointSamples":
        shared, D = _read_shared_vectors(dg, "coord")
        arrays, L = _read_dense_grids(dg, TF)
        if len(names) != L:
            raise ArgumentError("names length != number of grids")
        grids = tuple(PointSamples(a, shared) for a in arrays)
    else:
        raise ArgumentError("Unsupported params[:grid_type] = %s" % base)
    return GridDataset(GridBundle(grids, names), params)
