"""GridDataset files (IO/gridsIO/write_grids.jl, read_grids.jl) through the
package's own HDF5 encoder.  No HDF5 library is available here, so the files
are checked against the fixed layout of the HDF5 file format specification
(superblock version 0, version 1 object headers, local heaps, version 1
B-tree group nodes, symbol table nodes) by reading bytes at the offsets the
specification defines, independently of partia.hdf5's decoder; and round
trips through the decoder.  Whether libhdf5 reads these files is parity
unpinned."""
import struct

import numpy as np
import pytest

from partia import (ArgumentError, Cartesian, GridBundle, GridDataset, MethodError, PointSamples, StructuredGrid,
                    read_GridDataset, write_GridBundle, write_GridDataset)
from partia import hdf5

UNITS = dict(umass=1.989e33, udist=3.086e18, utime=3.156e13, umagfd=1e-6)


def structured_bundle(dtype=np.float64, n=(5, 4, 3)):
    g = StructuredGrid(Cartesian, (0.0, 1.0, n[0]), (0.0, 2.0, n[1]), (0.0, 3.0, n[2]), dtype=dtype)
    rng = np.random.default_rng(3)
    out = []
    for _ in range(3):
        s = g.similar()
        s.grid[:] = rng.random(g.size)
        out.append(s)
    return GridBundle(out, ("rho", "A", "div_v"))


def point_bundle(n=17):
    rng = np.random.default_rng(4)
    coor = tuple(rng.random(n) for _ in range(3))
    grids = [PointSamples(rng.random(n), coor) for _ in range(2)]
    return GridBundle(grids, ("rho", "A"))


# ---- the file against the specification's layout ----

def u(fmt, b, off):
    return struct.unpack_from("<" + fmt, b, off)


def group_members(b, oh):
    """Name -> object header address of the group at oh, by the specification's
    offsets: object header v1 (16-byte prefix, 8-byte message headers), symbol
    table message (B-tree, heap), B-tree node (24-byte header, key/child pairs),
    symbol table node (8-byte header, 40-byte entries), local heap (32 bytes)."""
    ver, _, nmsg, refc, size = u("BBHII", b, oh)
    assert (ver, refc) == (1, 1)
    p, msgs = oh + 16, {}
    for _ in range(nmsg):
        t, sz, _ = u("HHB", b, p)
        assert sz % 8 == 0
        msgs[t] = b[p + 8:p + 8 + sz]
        p += 8 + sz
    assert p == oh + 16 + size
    tree, heap = u("QQ", msgs[0x0011], 0)
    assert b[heap:heap + 4] == b"HEAP" and b[heap + 4] == 0
    hsize, free, hdata = u("QQQ", b, heap + 8)
    assert free == 1 and hsize % 8 == 0 and b[hdata:hdata + 8] == b"\0" * 8  # "" at offset 0
    assert b[tree:tree + 4] == b"TREE" and b[tree + 4:tree + 8] == bytes([0, 0, 1, 0])
    left, right, key0, child, key1 = u("QQQQQ", b, tree + 8)
    assert left == right == hdf5.UNDEF and key0 == 0
    assert b[child:child + 6] == b"SNOD\x01\x00"
    (nsym,) = u("H", b, child + 6)
    out, names = {}, []
    for k in range(nsym):
        off, addr, cache, _ = u("QQII", b, child + 8 + 40 * k)
        name = b[hdata + off:b.index(b"\0", hdata + off)].decode()
        names.append(name)
        out[name] = (addr, cache)
    assert names == sorted(names)
    assert key1 == u("Q", b, child + 8 + 40 * (nsym - 1))[0]  # the heap offset of the largest name
    return out


def dataset_raw(b, oh):
    ver, _, nmsg, _, size = u("BBHII", b, oh)
    p, msgs = oh + 16, {}
    for _ in range(nmsg):
        t, sz, _ = u("HHB", b, p)
        msgs[t] = b[p + 8:p + 8 + sz]
        p += 8 + sz
    sp, dt, lay = msgs[0x0001], msgs[0x0003], msgs[0x0008]
    rank = sp[1]
    dims = u("%dQ" % rank, sp, 8) if rank else ()
    assert lay[0] == 3 and lay[1] == 1  # version 3, contiguous
    addr, size = u("QQ", lay, 2)
    assert msgs[0x0005][:1] == b"\x02"  # fill value message, version 2
    return dims, dt, b[addr:addr + size]


def test_file_layout_follows_the_specification(tmp_path):
    gb = structured_bundle()
    fn = tmp_path / "s.h5"
    write_GridBundle(gb, str(fn), code_units=UNITS, operation_name="layout")
    b = fn.read_bytes()
    assert b[:8] == b"\x89HDF\r\n\x1a\n"
    assert b[8:16] == bytes([0, 0, 0, 0, 0, 8, 8, 0])  # versions 0, 8-byte offsets and lengths
    base, freespace, eof, driver = u("QQQQ", b, 24)
    assert (base, freespace, eof, driver) == (0, hdf5.UNDEF, len(b), hdf5.UNDEF)
    name_off, root, cache, _ = u("QQII", b, 56)
    assert name_off == 0 and cache == 1
    top = group_members(b, root)
    assert set(top) == {"params", "data"} and all(c == 1 for _, c in top.values())
    data = group_members(b, top["data"][0])
    assert set(data) == {"names", "axes", "grids"}
    grids = group_members(b, data["grids"][0])
    assert sorted(grids) == ["1", "2", "3"]
    for i, g in enumerate(gb.grids, start=1):
        dims, dt, raw = dataset_raw(b, grids[str(i)][0])
        # a Julia array of size (nx, ny, nz): dataspace (nz, ny, nx), x fastest
        assert dims == tuple(reversed(g.size))
        assert dt[:8] == struct.pack("<B3BI", 0x11, 0x20, 63, 0, 8)  # IEEE float64, little-endian
        assert raw == g.grid.ravel(order="F").astype("<f8").tobytes()
    axes = group_members(b, data["axes"][0])
    assert sorted(axes) == ["x", "y", "z"]
    for d, k in enumerate("xyz"):
        dims, _, raw = dataset_raw(b, axes[k][0])
        assert dims == (len(gb.grids[0].axes[d]),) and raw == gb.grids[0].axes[d].astype("<f8").tobytes()
    dims, dt, raw = dataset_raw(b, data["names"][0])
    n = max(len(s) for s in gb.names)
    assert dims == (3,) and dt[0] == 0x13 and dt[1] == 0x11 and u("I", dt, 4)[0] == n
    assert raw == b"".join(s.encode().ljust(n, b"\0") for s in gb.names)
    params = group_members(b, top["params"][0])
    dims, dt, raw = dataset_raw(b, params["schema_id"][0])
    assert dims == () and raw.rstrip(b"\0") == b"GridDataset.v1"


def test_decoder_rejects_damaged_files(tmp_path):
    gb = point_bundle()
    fn = tmp_path / "p.h5"
    write_GridBundle(gb, str(fn), code_units=UNITS)
    b = bytearray(fn.read_bytes())
    hdf5.decode(bytes(b))
    for at, what in ((0, "signature"), (24 + 16, "end of file")):
        c = bytearray(b)
        c[at] ^= 0xFF
        with pytest.raises(hdf5.HDF5FormatError):
            hdf5.decode(bytes(c))
    with pytest.raises(hdf5.HDF5FormatError):
        hdf5.decode(bytes(b[:-8]))


# ---- round trips and the reference's behaviour ----

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_structured_grid_round_trip(tmp_path, dtype):
    gb = structured_bundle(dtype)
    fn = str(tmp_path / "s.h5")
    write_GridBundle(gb, fn, code_units=UNITS, operation_name="rt", params={"note": "C1", "step": 12, "ok": True})
    gd = read_GridDataset(fn)
    assert isinstance(gd, GridDataset)
    assert gd.data.names == gb.names
    for a, b in zip(gd.data.grids, gb.grids):
        assert a.size == b.size and a.grid.dtype == np.dtype(dtype)
        assert np.array_equal(a.grid, b.grid)
        for x, y in zip(a.axes, b.axes):
            assert np.array_equal(x, y)
    assert gd.data.grids[0].axes is gd.data.grids[1].axes  # read once, shared
    p = gd.params
    assert p["schema_name"] == "GridDataset" and p["schema_id"] == "GridDataset.v1"
    assert p["dtype"] == ("Float64" if dtype == np.float64 else "Float32") and p["grid_type"] == "StructuredGrid"
    assert p["note"] == "C1" and p["step"] == 12 and p["ok"] is True
    assert all(p[k] == v for k, v in UNITS.items())
    assert p["file_identifier"].startswith("FT:Partia:0.1.3 (rt): ")


def test_point_samples_round_trip(tmp_path):
    gb = point_bundle()
    fn = str(tmp_path / "p.h5")
    write_GridDataset(GridDataset(gb, {"dtype": "Float64", "grid_type": "PointSamples"}), fn)
    gd = read_GridDataset(fn)
    assert gd.data.names == gb.names
    for a, b in zip(gd.data.grids, gb.grids):
        assert np.array_equal(a.grid, b.grid)
        assert all(np.array_equal(x, y) for x, y in zip(a.coor, b.coor))
    assert gd.data.grids[0].coor is gd.data.grids[1].coor


def test_many_grids_and_two_dimensional_axes(tmp_path):
    g = StructuredGrid(Cartesian, (0.0, 1.0, 6), (-1.0, 1.0, 7))
    grids = []
    for k in range(40):
        s = g.similar()
        s.grid[:] = k + np.arange(42).reshape(6, 7, order="F")
        grids.append(s)
    gb = GridBundle(grids, tuple("q%d" % k for k in range(40)))
    fn = str(tmp_path / "m.h5")
    write_GridBundle(gb, fn, code_units=UNITS)
    gd = read_GridDataset(fn)
    assert gd.data.names == gb.names and len(gd.data.grids) == 40
    assert all(np.array_equal(a.grid, b.grid) for a, b in zip(gd.data.grids, gb.grids))
    assert len(gd.data.grids[0].axes) == 2


def test_reference_errors(tmp_path):
    fn = str(tmp_path / "e.h5")
    gb = structured_bundle()
    other = StructuredGrid(Cartesian, (0.0, 1.0, 5), (0.0, 2.0, 4), (0.0, 3.0, 3))
    bad = GridBundle((gb.grids[0], other), ("a", "b"))
    with pytest.raises(ArgumentError, match=r"StructuredGrid axes not shared: grid=2 dim=1"):
        write_GridDataset(GridDataset(bad, {}), fn)
    pb = point_bundle()
    badp = GridBundle((pb.grids[0], PointSamples(pb.grids[1].grid, (pb.grids[0].coor[0], pb.grids[0].coor[1],
                                                                    pb.grids[0].coor[2].copy()))), ("a", "b"))
    with pytest.raises(ArgumentError, match=r"PointSamples coor not shared: grid=2 dim=3"):
        write_GridDataset(GridDataset(badp, {}), fn)
    with pytest.raises(ArgumentError, match=r"code_units missing required keys: \(:utime, :umagfd\)"):
        write_GridBundle(gb, fn, code_units=dict(umass=1.0, udist=1.0))
    with pytest.raises(MethodError):
        write_GridDataset(GridDataset(GridBundle((gb.grids[0], pb.grids[0]), ("a", "b")), {}), fn)
    # read_GridDataset's schema checks
    root = hdf5.Group()
    root.create_group("data")
    hdf5.write(fn, root)
    with pytest.raises(ArgumentError, match="missing /params"):
        read_GridDataset(fn)
    root = hdf5.Group()
    root.create_group("params").write("grid_type", "StructuredGrid")
    root.create_group("data")
    hdf5.write(fn, root)
    with pytest.raises(ArgumentError, match="params missing :dtype"):
        read_GridDataset(fn)
    write_GridDataset(GridDataset(gb, {"dtype": "Float8", "grid_type": "StructuredGrid"}), fn)
    with pytest.raises(ArgumentError, match="Unsupported dtype string"):
        read_GridDataset(fn)
    write_GridDataset(GridDataset(gb, {"dtype": "Float64", "grid_type": "LineSamples"}), fn)
    with pytest.raises(ArgumentError, match="Unsupported params"):
        read_GridDataset(fn)
