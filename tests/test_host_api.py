"""Host-side mirror of the reference API (no GPU needed): build_input, catalogs,
containers, LinRange axes, error types, and that libpartia_hip.so loads and
exports every function include/partia_hip.h declares.  Mirrors
test/interpolation_tests.jl:122-221 and test/structured_grid_interpolation_tests.jl:36-83."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import partia
from partia import (ArgumentError, Cartesian, Cylindrical, DimensionMismatch, GridBundle, HIPComputeBackend,
                    InterpolationCatalog, InterpolationInput, LineSamples, MethodError, PointSamples, Spherical,
                    StructuredGrid, build_input, curl_slots, div_slots, flatten, grad_slot, ordered_quantity_names,
                    restore_struct, scalar_index, to_concise_catalog)
from partia import _native as N
from partia.synthetic import linrange

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bi(fx, dtype_h=np.float32):
    f = {k: np.array(v, dtype=np.float32) for k, v in fx.items() if k != "source"}
    f["rho"] = np.array(fx["rho"], dtype=np.float64)
    return f


def test_build_input_promotes_and_resolves_slots(fixtures):
    f = _bi(fixtures["build_input"])
    inp, cat = build_input(HIPComputeBackend(), f["x"], f["y"], f["z"], f["m"], f["h"], f["rho"],
                           (f["P"], f["vx"], f["vy"], f["vz"], f["Bx"], f["By"], f["Bz"]),
                           column_names=("P", "vx", "vy", "vz", "Bx", "By", "Bz"), scalars=("P",),
                           gradients=("P",), divergences=("v",), curls=("B",))
    assert inp.Npart == 3
    assert partia.get_xcoord(inp).dtype == np.float64
    assert len(inp.quant) == 7
    assert scalar_index(cat, "P") == 1 and grad_slot(cat, "P") == 1
    assert div_slots(cat, "v") == (2, 3, 4) and curl_slots(cat, "B") == (5, 6, 7)
    names = ordered_quantity_names(cat)
    assert names[0] == "P" and len(names) == 8
    assert names == ("P", "∇Pˣ", "∇Pʸ", "∇Pᶻ", "∇⋅v", "∇×Bˣ", "∇×Bʸ", "∇×Bᶻ")
    assert np.all(inp.quant[0] == f["P"].astype(np.float64))
    assert np.all(inp.quant[6] == f["Bz"].astype(np.float64))
    with pytest.raises(KeyError):
        build_input(HIPComputeBackend(), f["x"], f["y"], f["z"], f["m"], f["h"], f["rho"],
                    (f["P"], f["vx"], f["vy"], f["vz"], f["Bx"], f["By"]),
                    column_names=("P", "vx", "vy", "vz", "Bx", "By"), curls=("B",))
    with pytest.raises(KeyError):
        build_input(HIPComputeBackend(), f["x"], f["y"], f["z"], f["m"], f["h"], f["rho"], (f["vx"],),
                    column_names=("vx",), scalars=("P",))


def test_build_input_all_float32_stays_float32(fixtures):
    f = _bi(fixtures["build_input"])
    rho32 = f["rho"].astype(np.float32)
    inp, _ = build_input(HIPComputeBackend(), f["x"], f["y"], f["z"], f["m"], f["h"], rho32, (f["P"],),
                         column_names=("P",), scalars=("P",))
    assert inp.m.dtype == np.float32


def test_input_dimension_mismatch():
    a = np.zeros(3)
    with pytest.raises(DimensionMismatch):
        InterpolationInput((a, a, np.zeros(2)), a, a, a, ())
    with pytest.raises(DimensionMismatch):
        InterpolationInput((a, a, a), a, np.zeros(4), a, ())
    with pytest.raises(DimensionMismatch):
        InterpolationInput((a, a, a), a, a, a, (np.zeros(1),))


def test_catalog_rules():
    cat = InterpolationCatalog.from_columns(("rho", "A", "ρ", "vx", "vy", "vz"), 3, scalars=("rho", "A", "ρ"),
                                            divergences=("v",))
    assert cat.scalar_snormalization == (False, True, False)
    assert cat.div_slots == ((4, 5, 6),)
    conc = to_concise_catalog(cat)
    assert conc.L == 4 and conc.scalar_slots == (1, 2, 3)
    with pytest.raises(ArgumentError):
        InterpolationCatalog.from_columns(("a", "a"), 3, scalars=("a",))
    with pytest.raises(DimensionMismatch):
        InterpolationCatalog(3, scalar_names=("a",), scalar_slots=())


def test_linrange_matches_lerp_definition():
    ax = linrange(-1.0, 1.0, 256)
    assert ax[0] == -1.0 and ax[-1] == 1.0 and len(ax) == 256
    j = np.arange(256) / 255.0
    assert np.array_equal(ax, (1 - j) * -1.0 + j * 1.0)
    assert np.allclose(np.diff(ax), 2.0 / 255, rtol=1e-12)
    a32 = linrange(0.2, 0.7, 3, np.float32)
    assert a32.dtype == np.float32 and a32[1] == np.float32(0.45000000000000001)
    assert len(linrange(0.5, 0.5, 1)) == 1


def test_structured_grid_flatten_restore(fixtures):
    p = fixtures["structured_template"]["params"]
    g = StructuredGrid(Cartesian, *[tuple(q) for q in p])
    assert g.size == (3, 2, 2)
    flat = flatten(Cartesian, g)
    assert len(flat) == 12
    # column-major: x fastest
    assert np.array_equal(flat.coor[0][:3], g.axes[0]) and flat.coor[1][3] == g.axes[1][1]
    back = restore_struct(Cartesian, flat, g.axes)
    assert back.size == g.size
    bad = PointSamples(flat.grid, (flat.coor[0] + 1e-3, flat.coor[1], flat.coor[2]))
    with pytest.raises(ArgumentError):
        restore_struct(Cartesian, bad, g.axes)
    for coord, params in ((Cylindrical, [(0.25, 0.55, 3), (0.30, 1.10, 2), (0.20, 0.50, 2)]),
                          (Spherical, [(0.50, 0.70, 2), (0.35, 0.85, 2), (0.80, 1.10, 2)])):
        sg = StructuredGrid(coord, *params)
        fl = flatten(coord, sg)
        r = np.sqrt(fl.coor[0] ** 2 + fl.coor[1] ** 2 + (0 if coord is Cylindrical else fl.coor[2] ** 2))
        assert np.all(r >= 0.25 - 1e-12)
        assert restore_struct(coord, fl, sg.axes).size == sg.size


def test_samples_validation():
    with pytest.raises(ArgumentError):
        PointSamples(np.zeros(3), (np.zeros(3), np.zeros(2), np.zeros(3)))
    with pytest.raises(ArgumentError):
        LineSamples(np.zeros(2), (np.zeros(2),) * 3, (np.zeros(2), np.zeros(1), np.zeros(2)))
    ps = PointSamples(np.zeros(3), (np.zeros(3),) * 3)
    assert ps.similar().coor is ps.coor
    gb = GridBundle([ps, ps.similar()], ["a", "b"])
    assert gb["b"] is gb.grids[1] and len(gb) == 2


def test_cpu_backend_has_no_methods():
    from partia import CPUComputeBackend, PointSamples_interpolation
    with pytest.raises(MethodError):
        PointSamples_interpolation(CPUComputeBackend(), None, None, None)


def _header_functions():
    src = open(os.path.join(ROOT, "include", "partia_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ph_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = N.lib()  # loads libpartia_hip.so (no GPU call)
    declared = _header_functions()
    assert declared, "no declarations parsed"
    assert sorted(N.EXPORTS) == declared
    for name in declared:
        assert hasattr(L, name), name


def test_line_table_export_matches_generator():
    import gen_line_lut
    out = np.zeros(4096)
    assert N.lib().ph_line_table(0, N.ptr(out)) == 0
    _, _, inorm = gen_line_lut.tables(*gen_line_lut.KERNELS[0])
    assert np.array_equal(out, inorm)
    assert N.lib().ph_line_table(9, N.ptr(out)) == -1


def test_no_device_error_is_reported_not_crashing():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = N.lib().ph_ctx_create(0, C.byref(h))
    assert rc == -9 and not h.value
