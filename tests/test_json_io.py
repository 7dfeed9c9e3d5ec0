"""Detector JSON files (dtr_read_detector_json / dtr_write_detector_json), no
GPU: detray's geometry, surface-grid and homogeneous-material files
(io/include/detray/io/json/*, io/include/detray/io/backend/*).

Pins:
* the reference's io test json_toy_geometry
  (tests/integration_tests/io/io_json_detector_roundtrip.cpp:185-226): the
  default toy detector (4 barrel layers, 3 endcap layers per side) read from
  its geometry file has 22 volumes, 3230 surfaces, 3252 transforms and
  2492 rectangle, 648 trapezoid, 56 concentric-cylinder and 60 ring masks;
* the reference's write -> read -> write round trip
  (io_json_detector_roundtrip.cpp:56-113: files equal after the 15 header
  lines) for the toy detector with homogeneous material and grids and for
  the telescope (1 volume, one slab per plane);
* the tables read back equal the emitted ones surface by surface, and the
  oracle propagates bit-identically through both;
* reader semantics from geometry_reader.hpp / the detector builders on
  edited files: factory order without index_in_coll, the cylinder z-shift
  fold, legacy keys, the material pairing, bins with several surfaces, and
  the errors for what the tables cannot hold;
* the wire chamber (drift cells, tungsten rods, bins of up to 3 wires)
  round trips; the reference's json_wire_chamber_reader reads 11 volumes
  (io_json_detector_roundtrip.cpp:319-342).
"""

import ctypes as C
import json
import math
import random

import numpy as np
import pytest

import detray_amd as da
from detray_amd import capi
from tests import oracle_host

IO_PORTAL_CYLINDER2, IO_RECTANGLE2, IO_RING2, IO_TRAPEZOID2 = 4, 5, 6, 7
INVALID_ARGUMENT, UNSUPPORTED, IO = 1, 5, 6


def read_error(geometry, grids=None, material=None, material_maps=None):
    lib = capi.load()
    d = capi.DetectorDesc()
    msg = C.create_string_buffer(512)
    enc = lambda p: None if p is None else str(p).encode()
    code = lib.dtr_read_detector_json(enc(geometry), enc(grids), enc(material), enc(material_maps), C.byref(d), msg,
                                      len(msg))
    if code == 0:
        lib.dtr_host_detector_free(C.byref(d))
    return code, msg.value.decode()


def dump(path, obj):
    with open(path, "w") as f:
        json.dump(obj, f, indent=4)
    return path


def load(path):
    with open(path) as f:
        return json.load(f)


def surfaces_of(det, v):
    vol = det.desc.volumes[v]
    return [det.desc.surfaces[s] for s in range(vol.sf_begin, vol.sf_end)]


def trf(det, s):
    t = det.desc.transforms[s.transform]
    return tuple(t.r) + tuple(t.t)


def masks(det, s):
    return [(det.desc.masks[m].shape, det.desc.masks[m].volume_link, tuple(det.desc.masks[m].v))
            for m in range(s.mask, s.mask + s.n_masks)]


def material(det, s):
    if s.material == 0xFFFFFFFF:
        return None
    m = det.desc.materials[s.material]
    return tuple(tuple(v) if hasattr(v, "_length_") else v for v in (getattr(m, f) for f, _ in m._fields_))


def bin_entries(det, g):
    cap = max(g.capacity, 1)
    raw = det.table("grid_bins")[g.bins:g.bins + g.n0 * g.n1 * cap]
    return [[s for s in raw[b * cap:(b + 1) * cap] if s != 0xFFFFFFFF] for b in range(g.n0 * g.n1)]


@pytest.fixture(scope="module")
def toy3_files(tmp_path_factory):
    det = da.Detector.toy(4, 3)
    d = tmp_path_factory.mktemp("toy3")
    return det, det.write_json(d, "toy_detector")


def test_toy_geometry_counts(toy3_files):
    _, f = toy3_files
    d = da.Detector.from_json(f["geometry"])
    assert (d.desc.n_volumes, d.desc.n_surfaces, d.desc.n_transforms) == (22, 3230, 3252)
    shapes = np.bincount([m.shape for m in d.table("masks")], minlength=4)
    assert (capi.RECTANGLE2D, capi.TRAPEZOID2D, capi.CONCENTRIC_CYLINDER2D, capi.RING2D) == (0, 1, 2, 3)
    assert list(shapes) == [2492, 648, 56, 60]
    # without the grid file every surface is in the brute-force lists
    assert d.desc.n_grids == 0 and d.desc.n_bf == 3230


def _compare_files(a, b, skip=15):
    la, lb = open(a).read().splitlines(), open(b).read().splitlines()
    assert len(la) == len(lb)
    bad = [i for i in range(skip, len(la)) if la[i] != lb[i]]
    assert not bad, (bad[:3], la[bad[0]], lb[bad[0]])


@pytest.mark.parametrize("which", ["toy3", "toy7", "telescope", "wire_chamber"])
def test_write_read_write_round_trip(tmp_path, which):
    if which == "telescope":
        positions = [1.0, 50.0, 100.0, 150.0, 200.0, 250.0, 300.0, 350.0, 400.0, 450.0, 500.0]
        det = da.Detector.telescope(positions, 100.0, 100.0)
    elif which == "wire_chamber":
        det = da.Detector.wire_chamber()
    else:
        det = da.Detector.toy(4, 3 if which == "toy3" else 7)
    f1 = det.write_json(tmp_path, "det")
    d2 = da.Detector.from_json(f1["geometry"], f1["surface_grids"], f1["homogeneous_material"])
    f2 = d2.write_json(tmp_path, "det_2")
    for k in f1:
        _compare_files(f1[k], f2[k])
    if which == "telescope":
        assert d2.desc.n_volumes == 1 and d2.desc.n_materials == len(positions)
    if which == "wire_chamber":
        # the reference's json_wire_chamber_reader: 11 volumes
        assert d2.desc.n_volumes == 11
        assert all(m.kind == capi.MATERIAL_ROD for m in d2.table("materials"))
        assert load(f1["homogeneous_material"])["header"]["rod_count"] == d2.desc.n_materials


@pytest.mark.parametrize("which", ["toy7", "telescope", "wire_chamber"])
def test_read_tables_equal_emitted(tmp_path, which):
    det = (da.Detector.toy(4, 7) if which == "toy7"
           else da.Detector.wire_chamber(half_z=500.0) if which == "wire_chamber"
           else da.Detector.telescope([100.0 * i for i in range(1, 11)], 500.0, 500.0, 0.15, 60.0))
    f = det.write_json(tmp_path, "det")
    d2 = da.Detector.from_json(f["geometry"], f["surface_grids"], f["homogeneous_material"])
    assert d2.desc.n_volumes == det.desc.n_volumes and d2.desc.n_surfaces == det.desc.n_surfaces
    assert d2.desc.n_transforms == det.desc.n_transforms
    for v in range(det.desc.n_volumes):
        a, b = det.desc.volumes[v], d2.desc.volumes[v]
        assert (a.sf_begin, a.sf_end, a.transform, a.bf_begin, a.bf_end, a.grid) == \
               (b.sf_begin, b.sf_end, b.transform, b.bf_begin, b.bf_end, b.grid)
        for sa, sb in zip(surfaces_of(det, v), surfaces_of(d2, v)):
            assert (sa.barcode, sa.transform, sa.volume, sa.sf_id, sa.shape) == \
                   (sb.barcode, sb.transform, sb.volume, sb.sf_id, sb.shape)
            assert trf(det, sa) == trf(d2, sb)
            assert masks(det, sa) == masks(d2, sb)
            assert material(det, sa) == material(d2, sb)
    assert det.table("bf_surfaces") == d2.table("bf_surfaces")
    # bins: same entries in the same order (the capacity read back is the
    # largest occupancy, the emitter's may be larger)
    for ga, gb in zip(det.table("grids"), d2.table("grids")):
        fields = [f for f, _ in ga._fields_ if f not in ("pad", "capacity", "bins")]
        assert [getattr(ga, f) for f in fields] == [getattr(gb, f) for f in fields]
        assert bin_entries(det, ga) == bin_entries(d2, gb)


def test_oracle_propagates_identically(tmp_path, toy7, solenoid):
    f = toy7.write_json(tmp_path, "toy")
    d2 = da.Detector.from_json(f["geometry"], f["surface_grids"], f["homogeneous_material"])
    fp = da.generate_tracks("toy", 500, 43)
    cfg = da.PropagationConfig(search_window=(3, 3))
    a = oracle_host.propagate(toy7, solenoid, cfg, fp)
    b = oracle_host.propagate(d2, solenoid, cfg, fp)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_factory_order_without_index_in_coll(tmp_path, toy3_files):
    """geometry_reader.hpp:96-150: without index_in_coll, portal factories
    come first, then sensitive/passive factories, each keyed by shape id
    (std::map order); file order inside a factory."""
    _, f = toy3_files
    geo = load(f["geometry"])
    v = 7  # a barrel layer: cylinder and ring portals, rectangle modules
    sfs = geo["data"]["volumes"][v]["surfaces"]
    rng = random.Random(5)
    rng.shuffle(sfs)
    for s in sfs:
        del s["index_in_coll"]
    d = da.Detector.from_json(dump(tmp_path / "g.json", geo))

    def key(s):
        return s["masks"][0]["shape"]

    expect = ([s for k in sorted({key(s) for s in sfs}) for s in sfs if s["type"] == 0 and key(s) == k] +
              [s for k in sorted({key(s) for s in sfs}) for s in sfs if s["type"] != 0 and key(s) == k])
    got = surfaces_of(d, v)
    assert [s["type"] for s in expect] == [s.sf_id for s in got]
    shape_of = {IO_PORTAL_CYLINDER2: 2, IO_RECTANGLE2: 0, IO_RING2: 3, IO_TRAPEZOID2: 1}
    assert [shape_of[key(s)] for s in expect] == [s.shape for s in got]
    assert [tuple(s["transform"]["translation"]) for s in expect] == [tuple(trf(d, s)[9:]) for s in got]
    assert {s.sf_id for s in got[:4]} == {0}


def test_index_in_coll_restores_order(tmp_path, toy3_files):
    det, f = toy3_files
    geo = load(f["geometry"])
    rng = random.Random(6)
    for vol in geo["data"]["volumes"]:
        rng.shuffle(vol["surfaces"])
    d = da.Detector.from_json(dump(tmp_path / "g.json", geo), f["surface_grids"], f["homogeneous_material"])
    for v in range(det.desc.n_volumes):
        for sa, sb in zip(surfaces_of(det, v), surfaces_of(d, v)):
            assert sa.barcode == sb.barcode and trf(det, sa) == trf(d, sb) and masks(det, sa) == masks(d, sb)


def test_cylinder_shift_fold_and_legacy_keys(tmp_path, toy3_files):
    """geometry_reader.hpp:187-203: the z translation of a cylinder moves
    into the mask's z bounds and the transform becomes the identity; single
    "mask" and "barcode" keys of older files are read."""
    _, f = toy3_files
    geo = load(f["geometry"])
    vol = geo["data"]["volumes"][0]
    cyl = next(s for s in vol["surfaces"] if s["masks"][0]["shape"] == IO_PORTAL_CYLINDER2)
    k = cyl["index_in_coll"]
    r, lo, hi = cyl["masks"][0]["boundaries"]
    cyl["transform"]["translation"] = [0.0, 0.0, 125.0]
    cyl["masks"][0]["shape"] = 2  # cylinder2 is read as a portal cylinder
    for s in vol["surfaces"]:
        if len(s["masks"]) == 1:
            s["mask"] = s.pop("masks")[0]
        s["barcode"] = s.pop("identifier")
    assert any("mask" in s for s in vol["surfaces"])
    d = da.Detector.from_json(dump(tmp_path / "g.json", geo))
    s = surfaces_of(d, 0)[k]
    assert s.shape == 2
    assert masks(d, s)[0][2][:3] == (r, lo + 125.0, hi + 125.0)
    assert trf(d, s) == (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0)


def test_world_exit_link(tmp_path, toy3_files):
    """The 16-bit navigation link takes 2^64-1 (the io default) as 0xffff."""
    _, f = toy3_files
    geo = load(f["geometry"])
    for s in geo["data"]["volumes"][0]["surfaces"]:
        for m in s["masks"]:
            if m["volume_link"] == 65535:
                m["volume_link"] = 18446744073709551615
    d = da.Detector.from_json(dump(tmp_path / "g.json", geo))
    links = {m[1] for s in surfaces_of(d, 0) for m in masks(d, s)}
    assert 0xFFFF in links


def test_material_pairing_and_slabs(tmp_path, toy3_files):
    """homogeneous_material_factory.hpp:322-385: surface (min surface_idx + j)
    gets entry j; one slab per material surface, placed at index_in_coll;
    older files name the list "material_slabs"."""
    det, f = toy3_files
    mat = load(f["homogeneous_material"])
    v = next(i for i, mv in enumerate(mat["data"]["volumes"]) if len(mv.get("surface_material", [])) > 3)
    entries = mat["data"]["volumes"][v].pop("surface_material")
    # distinguishable thicknesses, then list the entries in reverse order
    for j, e in enumerate(entries):
        e["thickness"] = 0.1 + 0.01 * j
    first = min(e["surface_idx"] for e in entries)
    mat["data"]["volumes"][v]["material_slabs"] = entries[::-1]
    d = da.Detector.from_json(f["geometry"], f["surface_grids"], dump(tmp_path / "m.json", mat))
    sfs = surfaces_of(d, v)
    rev = entries[::-1]
    for j in range(len(rev)):
        s = sfs[first + j]
        assert d.desc.materials[s.material].thickness == rev[j]["thickness"]
        assert s.material - min(x.material for x in sfs if x.material != 0xFFFFFFFF) == rev[j]["index_in_coll"]
    assert d.desc.n_materials == sum(1 for s in d.table("surfaces") if s.material != 0xFFFFFFFF)


def test_bins_with_several_surfaces(tmp_path, toy3_files):
    """Bins keep every entry in file order; the largest occupancy becomes the
    grid's capacity (the wire chamber's dynamic bins hold up to 3)."""
    _, f = toy3_files
    grids = load(f["surface_grids"])
    gd = grids["data"]["grids"][0]
    g = gd["grid_data"][0]
    full = next(b for b in g["bins"] if b["content"])
    other = next(b["content"][0] for b in g["bins"] if b["content"] and b["content"][0] != full["content"][0])
    full["content"].append(other)
    d = da.Detector.from_json(f["geometry"], dump(tmp_path / "g.json", grids))
    vol = d.desc.volumes[gd["volume_link"]]
    grid = d.desc.grids[vol.grid]
    assert grid.capacity == 2
    i0, i1 = full["loc_index"]
    first = grid.bins + (i0 * grid.n1 + i1) * 2
    assert d.table("grid_bins")[first:first + 2] == [vol.sf_begin + c for c in full["content"]]
    assert d.desc.n_bins == grid.n0 * grid.n1 * 2 + sum(
        g2.n0 * g2.n1 * max(g2.capacity, 1) for k, g2 in enumerate(d.table("grids")) if k != vol.grid)


def test_empty_grid_file(tmp_path, toy3_files):
    """A grid file whose data is null (what the reference writes for a
    detector without grids) adds no grids."""
    _, f = toy3_files
    grids = load(f["surface_grids"])
    grids["data"] = None
    d = da.Detector.from_json(f["geometry"], dump(tmp_path / "g.json", grids))
    assert d.desc.n_grids == 0 and d.desc.n_bf == d.desc.n_surfaces


def test_errors(tmp_path, toy3_files):
    _, f = toy3_files
    assert read_error(tmp_path / "missing.json")[0] == IO
    (tmp_path / "broken.json").write_text('{"header": {}, "data": {"volumes": [}')
    code, msg = read_error(tmp_path / "broken.json")
    assert code == IO and "line 1" in msg

    geo = load(f["geometry"])
    geo["data"]["volumes"][3]["surfaces"][5]["masks"][0]["shape"] = 1  # cuboid3
    code, msg = read_error(dump(tmp_path / "a.json", geo))
    assert code == UNSUPPORTED and "cuboid3" in msg

    geo = load(f["geometry"])
    geo["data"]["volumes"][3]["surfaces"][5]["masks"][0]["shape"] = 0  # annulus2 needs 7 boundaries
    code, msg = read_error(dump(tmp_path / "a2.json", geo))
    assert code == INVALID_ARGUMENT and "7 numbers" in msg

    geo = load(f["geometry"])
    del geo["data"]["volumes"][2]["surfaces"][0]["index_in_coll"]
    code, msg = read_error(dump(tmp_path / "b.json", geo))
    assert code == INVALID_ARGUMENT and "index_in_coll" in msg

    geo = load(f["geometry"])
    geo["data"]["volumes"][2]["surfaces"][0]["index_in_coll"] = geo["data"]["volumes"][2]["surfaces"][1]["index_in_coll"]
    assert read_error(dump(tmp_path / "c.json", geo))[0] == INVALID_ARGUMENT


    grids = load(f["surface_grids"])
    grids["data"]["grids"][0]["grid_data"][0]["axes"][1]["binning"] = 1
    assert read_error(f["geometry"], dump(tmp_path / "g2.json", grids))[0] == UNSUPPORTED

    grids = load(f["surface_grids"])
    grids["data"]["grids"][0]["grid_data"][0]["owner_link"] = 99
    assert read_error(f["geometry"], dump(tmp_path / "g3.json", grids))[0] == INVALID_ARGUMENT

    mat = load(f["homogeneous_material"])
    mv = next(mv for mv in mat["data"]["volumes"] if mv.get("surface_material"))
    mv["surface_material"][0]["type"] = 3  # a material map is not homogeneous material
    assert read_error(f["geometry"], None, dump(tmp_path / "m1.json", mat))[0] == INVALID_ARGUMENT

    mat = load(f["homogeneous_material"])
    mv = next(mv for mv in mat["data"]["volumes"] if len(mv.get("surface_material", [])) > 2)
    mv["surface_material"][1]["surface_idx"] += 1000
    assert read_error(f["geometry"], None, dump(tmp_path / "m2.json", mat))[0] == INVALID_ARGUMENT


@pytest.mark.parametrize("link", [22, 23, 65535 + 1, 2 ** 32, 2 ** 63])
def test_out_of_range_volume_link(tmp_path, toy3_files, link):
    """A mask link must name a volume of the file or be the world exit (the
    invalid 16-, 32- or 64-bit link): a link to volume 22 of the 22-volume
    toy, or one that would wrap to a valid 16-bit index (65536 -> 0), is an
    error, not a silent link."""
    _, f = toy3_files
    geo = load(f["geometry"])
    geo["data"]["volumes"][4]["surfaces"][0]["masks"][0]["volume_link"] = link
    code, msg = read_error(dump(tmp_path / "l.json", geo))
    assert code == INVALID_ARGUMENT and "volume_link" in msg


def test_writer_rejects_inconsistent_tables():
    lib = capi.load()
    msg = C.create_string_buffer(256)
    empty = capi.DetectorDesc()
    assert lib.dtr_write_detector_json(C.byref(empty), b"x", None, b"/dev/null", None, None, None, msg, 256) == \
        INVALID_ARGUMENT
    assert lib.dtr_write_detector_json(None, b"x", None, b"/dev/null", None, None, None, msg, 256) == INVALID_ARGUMENT


def test_volume_names_and_header(tmp_path):
    det = da.Detector.telescope([100.0, 200.0], 50.0, 50.0)
    f = det.write_json(tmp_path, "tel", volume_names=['tele "scope" é'])
    geo = load(f["geometry"])
    assert geo["header"]["common"]["tag"] == "geometry" and geo["header"]["common"]["detector"] == "tel"
    assert geo["header"]["volume_count"] == 1 and geo["header"]["surface_count"] == det.desc.n_surfaces
    assert geo["data"]["volumes"][0]["name"] == 'tele "scope" é'
    assert load(f["homogeneous_material"])["header"]["common"]["tag"] == "homogeneous_material"
    assert load(f["surface_grids"])["header"]["common"]["tag"] == "surface_grids"
    assert load(f["surface_grids"])["data"] is None
    # every number reads back to the same double
    d = da.Detector.from_json(f["geometry"], f["surface_grids"], f["homogeneous_material"])
    assert [tuple(t.r) + tuple(t.t) for t in det.table("transforms")] == \
           [tuple(t.r) + tuple(t.t) for t in d.table("transforms")]
    assert all(not math.isnan(m.thickness_in_X0) for m in d.table("materials"))
