"""Material maps on the CPU: surfaces whose material is a grid of slabs over
their bound local coordinates (material_map.hpp, material_accessor.hpp:
30-48, pointwise_material_interactor.hpp:79-80), the toy detector's maps
(toy_det_config::use_material_maps, material_map_generator.hpp) and the JSON
material-map file (material_map_reader.hpp / material_map_writer.hpp).

Pins of the lookup: a telescope whose planes carry a 4 x 1 map over x, in
a field along x so that tracks starting along z keep their x, must give, per
track, bit for bit the results of the same telescope with the homogeneous
slab of the bin that x falls in (or without material for an empty bin):
x below the axis clamps to bin 0 and above it to the last bin on a closed
axis (axis_bounds.hpp:107-125), and wraps with (n + b % n) % n on a
circular one (axis_bounds.hpp:219-229), with b = int((x - min) / width + 1)
- 1 (axis_binning.hpp:80-82).  Oracle and the batched engine's host build
are held to that; the toy with maps then at the batch contracts."""

import json
import math

import numpy as np
import pytest

import detray_amd as da
from detray_amd import capi
from tests import engine_host, oracle_host

KEYS = ("bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces", "n_steps",
        "n_trials", "status")

POSITIONS = [100.0 * i for i in range(1, 11)]
EDGES = (-200.0, 200.0)


def _copy(items, t):
    return [t.from_buffer_copy(x) for x in items]


def _telescope(plane_material):
    """The 10-plane telescope with every plane's material link replaced by
    plane_material(materials, maps) -> link."""
    base = da.Detector.telescope(POSITIONS, 500.0, 500.0, 0.15, 60.0)
    surfaces = _copy(base.table("surfaces"), capi.Surface)
    materials, maps = [], []
    link = plane_material(materials, maps)
    for s in surfaces:
        if s.sf_id == capi.SENSITIVE:
            s.material = link
    return da.Detector.from_tables(_copy(base.table("volumes"), capi.Volume), surfaces,
                                   _copy(base.table("transforms"), capi.Transform),
                                   _copy(base.table("masks"), capi.Mask), materials,
                                   bf_surfaces=base.table("bf_surfaces"), material_maps=maps)


def _slabs():
    return [capi.MaterialSlab(), da.material("silicon_tml", 0.15), capi.MaterialSlab(),
            da.material("silicon_tml", 0.3)]


def map_telescope(circular: bool):
    def link(materials, maps):
        m = capi.MaterialMap()
        m.n0, m.n1, m.slabs = 4, 1, len(materials)
        m.circular0, m.circular1 = int(circular), 0
        m.min0, m.max0, m.min1, m.max1 = EDGES[0], EDGES[1], -500.0, 500.0
        materials.extend(_slabs())
        maps.append(m)
        return capi.MATERIAL_MAP | 0
    return _telescope(link)


def homogeneous_telescope(k):
    """The telescope with the slab of map bin k on every plane (none for an
    empty bin)."""
    def link(materials, maps):
        slab = _slabs()[k]
        if slab.thickness == 0.0:
            return 0xFFFFFFFF
        materials.append(slab)
        return 0
    return _telescope(link)


def expected_bin(x: float, circular: bool) -> int:
    n, width = 4, (EDGES[1] - EDGES[0]) / 4
    b = int((x - EDGES[0]) / width + 1.0) - 1
    if circular:
        return (n + int(math.fmod(b, n))) % n
    return min(max(b, 0), n - 1)


def beam_tracks(xs):
    """Tracks from (x, 20, 0) along +z, p = 2 GeV, charges alternating; in a
    field along x they bend in y and keep x."""
    n = len(xs)
    fp = np.zeros((8, n))
    fp[0] = xs
    fp[1] = 20.0
    fp[6] = 1.0
    fp[7] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0) / 2.0
    return fp


XS = np.array([-470.0, -430.0, -350.0, -250.0, -205.0, -150.0, -99.0, -50.0, -1.0, 0.0, 30.0, 99.5, 100.0,
               150.0, 199.0, 201.0, 260.0, 330.0, 420.0, 480.0])
FIELD = da.MagneticField.constant(0.5 * da.TESLA, 0.0, 0.0)
CFG = da.PropagationConfig(search_window=(3, 3))


def _assert_same(a, b, cols, label):
    for k in KEYS:
        np.testing.assert_array_equal(a[k][..., cols], b[k][..., cols], err_msg=f"{label}: {k}")


@pytest.mark.parametrize("circular", [False, True])
@pytest.mark.parametrize("run", ["oracle", "plain", "engine_fp64", "engine_mixed"])
def test_map_bins_equal_homogeneous_slabs(circular, run):
    def go(det, fp):
        if run == "oracle":
            return oracle_host.propagate(det, FIELD, CFG, fp)
        return engine_host.propagate(det, FIELD, CFG, fp, mixed=run == "engine_mixed", plain=run == "plain")
    fp = beam_tracks(XS)
    got = go(map_telescope(circular), fp)
    bins = np.array([expected_bin(x, circular) for x in XS])
    # every bin is hit, both clamp/wrap sides included
    assert set(bins) == {0, 1, 2, 3}
    assert (got["n_surfaces"] >= 10).all() and (got["status"] == 0).all()
    # x stays put: the field is along x and the tracks start along z
    np.testing.assert_allclose(got["free_params"][0], XS, atol=1e-9)
    for k in range(4):
        cols = np.flatnonzero(bins == k)
        _assert_same(got, go(homogeneous_telescope(k), fp), cols, f"bin {k}")
    # material makes a difference: the empty and the filled bins differ
    p = np.abs(1.0 / got["bound_params"][4])
    assert p[bins == 3].max() < p[bins == 1].min() < p[bins == 0].min()


def _with(det, surface_link=None, maps=None, materials=None):
    surfaces = _copy(det.table("surfaces"), capi.Surface)
    for s in surfaces:
        if s.sf_id == capi.SENSITIVE and surface_link is not None:
            s.material = surface_link
    return da.Detector.from_tables(_copy(det.table("volumes"), capi.Volume), surfaces,
                                   _copy(det.table("transforms"), capi.Transform), _copy(det.table("masks"), capi.Mask),
                                   materials if materials is not None else _copy(det.table("materials"),
                                                                                 capi.MaterialSlab),
                                   bf_surfaces=det.table("bf_surfaces"),
                                   material_maps=maps if maps is not None else _copy(det.table("material_maps"),
                                                                                     capi.MaterialMap))


def test_map_tables_are_checked():
    """Map links out of range, maps beyond the slab table, empty or inverted
    axes and maps over rods are refused by the table checks that
    dtr_detector_create applies (dtr_check_detector on the host)."""
    det = map_telescope(False)
    det.check()
    da.Detector.toy(4, 7, use_material="maps").check()
    assert det.desc.n_material_maps == 1 and det.table("material_maps")[0].n0 == 4
    with pytest.raises(capi.DetrayError, match="material map out of range"):
        _with(det, surface_link=capi.MATERIAL_MAP | 1).check()

    def bad_map(**kw):
        m = _copy(det.table("material_maps"), capi.MaterialMap)
        for k, v in kw.items():
            setattr(m[0], k, v)
        return _with(det, maps=m)
    with pytest.raises(capi.DetrayError, match="slabs out of range"):
        bad_map(slabs=1).check()
    with pytest.raises(capi.DetrayError, match="empty or inverted axis"):
        bad_map(n1=0).check()
    with pytest.raises(capi.DetrayError, match="empty or inverted axis"):
        bad_map(min0=300.0).check()
    rods = _copy(det.table("materials"), capi.MaterialSlab)
    rods[1].kind = capi.MATERIAL_ROD
    with pytest.raises(capi.DetrayError, match="holds a rod"):
        _with(det, materials=rods).check()


# --- the toy detector's maps -------------------------------------------------

def _f32(x):
    return float(np.float32(x))


def test_toy_maps_follow_the_generator():
    """Every portal and the beampipe carry a map, the modules none; bins and
    thicknesses as material_map_generator.hpp makes them."""
    toy = da.Detector.toy(4, 7, use_material="maps")
    surfaces, masks = toy.table("surfaces"), toy.table("masks")
    mats, maps, vols = toy.table("materials"), toy.table("material_maps"), toy.table("volumes")
    portals = [i for i, s in enumerate(surfaces) if s.sf_id == capi.PORTAL]
    passives = [i for i, s in enumerate(surfaces) if s.sf_id == capi.PASSIVE]
    assert len(passives) == 1 and len(maps) == len(portals) + 1
    assert all(s.material == 0xFFFFFFFF for s in surfaces if s.sf_id == capi.SENSITIVE)
    assert len(mats) == sum(m.n0 * m.n1 for m in maps)
    # mixture.hpp of silicon_tml (0.9) and aluminium (0.1)
    si = (_f32(95.7), _f32(465.2), _f32(28.03), 14.0, 2.32)
    al = (_f32(88.97), _f32(397.0), _f32(26.98), 13.0, 2.699)
    x0 = 1.0 / (0.9 / si[0] + 0.1 / al[0])
    a_mix = si[2] * 0.9 + al[2] * 0.1
    endcap_layers = {v for v, vol in enumerate(vols) if vol.grid >= 0 and toy.table("grids")[vol.grid].type == 1}
    seen_thin = seen_thick = 0
    for i in portals + passives:
        s = surfaces[i]
        m = maps[s.material & ~capi.MATERIAL_MAP]
        assert s.material & capi.MATERIAL_MAP
        ms = [masks[s.mask + k] for k in range(s.n_masks)]
        slab = [mats[m.slabs + b] for b in range(m.n0 * m.n1)]
        if s.shape == capi.CONCENTRIC_CYLINDER2D:
            lo, hi = min(x.v[1] for x in ms), max(x.v[2] for x in ms)
            assert (m.n0, m.n1, m.circular0, m.circular1) == (20, 20, 1, 0)
            assert (m.min0, m.max0, m.min1, m.max1) == (-math.pi, math.pi, lo, hi)
            z = lo + np.arange(20) * (hi - lo) / 19.0
            if s.sf_id == capi.PASSIVE:
                t = np.full(20, _f32(0.8))
                assert slab[0].X0 == _f32(352.8)
            else:
                thin = s.volume in endcap_layers
                seen_thin += thin
                seen_thick += not thin
                t = float(np.float32(1e-8)) * z * z + (_f32(0.15) if thin else 5.0)
                assert slab[0].X0 == pytest.approx(x0, rel=1e-15) and slab[0].A == pytest.approx(a_mix, rel=1e-15)
            got = np.array([[slab[b0 * 20 + b1].thickness for b1 in range(20)] for b0 in range(20)])
            np.testing.assert_allclose(got, np.broadcast_to(t, (20, 20)), rtol=1e-12)
        else:
            lo, hi = min(x.v[0] for x in ms), max(x.v[1] for x in ms)
            assert (m.n0, m.n1, m.circular0, m.circular1) == (5, 20, 0, 1)
            assert (m.min0, m.max0, m.min1, m.max1) == (lo, hi, -math.pi, math.pi)
            r = lo + np.arange(5) * (hi - lo) / 4.0
            t = float(np.float32(1e-4)) * r + 1.0
            got = np.array([[slab[b0 * 20 + b1].thickness for b1 in range(20)] for b0 in range(5)])
            np.testing.assert_allclose(got, np.broadcast_to(t[:, None], (5, 20)), rtol=1e-12)
    assert seen_thin > 0 and seen_thick > 0


@pytest.mark.parametrize("mixed,plain", [(False, False), (True, False), (False, True)])
def test_toy_with_maps_engine_matches_oracle(solenoid, mixed, plain):
    """The batched engine and the plain kernel (host builds) against the
    oracle on the toy with maps in the solenoid map: identical surface
    sequences and the FP64 / mixed contracts on the parameters."""
    from tests.test_engine_host import _rel, _same_sequences
    toy = da.Detector.toy(4, 7, use_material="maps")
    fp = da.generate_tracks("toy", 300, 47)
    ref = oracle_host.propagate(toy, solenoid, CFG, fp)
    got = engine_host.propagate(toy, solenoid, CFG, fp, mixed=mixed, plain=plain)
    _same_sequences(got, ref)
    assert _rel(got["bound_params"], ref["bound_params"]) < (1e-4 if mixed else 1e-9)
    # the maps are crossed: the final momenta differ from the toy without material
    bare = oracle_host.propagate(da.Detector.toy(4, 7, use_material=False), solenoid, CFG, fp)
    assert (np.abs(ref["bound_params"][4]) != np.abs(bare["bound_params"][4])).mean() > 0.9


# --- JSON material-map files ---------------------------------------------------

def _strip_header(path):
    with open(path) as f:
        d = json.load(f)
    d["header"]["common"].pop("date")
    return d


def test_json_round_trip_of_the_toy_with_maps(tmp_path):
    toy = da.Detector.toy(4, 7, use_material="maps")
    files = toy.write_json(tmp_path, "toy")
    assert set(files) == {"geometry", "surface_grids", "homogeneous_material", "material_maps"}
    back = da.Detector.from_json(files["geometry"], files["surface_grids"], files["homogeneous_material"],
                                 material_maps=files["material_maps"])
    for t in ("volumes", "materials", "material_maps"):
        a, b = toy.table(t), back.table(t)
        assert len(a) == len(b), t
        assert all(bytes(x) == bytes(y) for x, y in zip(a, b)), t
    # the builders share module masks, the reader gives each surface its own
    fields = ("barcode", "transform", "material", "volume", "n_masks", "sf_id", "shape")
    a, b = toy.table("surfaces"), back.table("surfaces")
    assert len(a) == len(b)
    assert all(getattr(x, f) == getattr(y, f) for x, y in zip(a, b) for f in fields)
    (tmp_path / "again").mkdir()
    again = back.write_json(tmp_path / "again", "toy")
    for k in files:
        assert _strip_header(files[k]) == _strip_header(again[k]), k
    maps = _strip_header(files["material_maps"])
    assert maps["header"]["common"]["tag"] == "material_maps"
    assert maps["header"]["grid_count"] == toy.desc.n_material_maps
    g = maps["data"]["grids"][0]["grid_data"][0]
    assert [a["bounds"] for a in g["axes"]] in ([2, 1], [1, 2])
    assert len(g["bins"]) == g["axes"][0]["bins"] * g["axes"][1]["bins"]
    assert len(g["bins"][0]["content"]) == 1 and g["bins"][0]["content"][0]["type"] == 6
    # the homogeneous file carries no mapped surface
    hom = _strip_header(files["homogeneous_material"])
    assert hom["header"]["slab_count"] == 0


def _maps_file(tmp_path, mutate):
    toy = da.Detector.toy(4, 7, use_material="maps")
    files = toy.write_json(tmp_path, "toy")
    with open(files["material_maps"]) as f:
        d = json.load(f)
    mutate(d)
    bad = tmp_path / "bad_material_maps.json"
    bad.write_text(json.dumps(d))
    return files, bad


INVALID_ARGUMENT, UNSUPPORTED = 1, 5


@pytest.mark.parametrize("case,code", [
    ("open_axis", UNSUPPORTED), ("irregular", UNSUPPORTED), ("wrong_type", INVALID_ARGUMENT),
    ("bin_out_of_range", INVALID_ARGUMENT), ("rod_content", INVALID_ARGUMENT),
    ("duplicate_bin", INVALID_ARGUMENT), ("owner_out_of_range", INVALID_ARGUMENT),
    ("map_and_slab", INVALID_ARGUMENT),
])
def test_json_map_errors(tmp_path, case, code):
    from tests.test_json_io import read_error
    g0 = lambda d: d["data"]["grids"][0]["grid_data"][0]
    mutate = {
        "open_axis": lambda d: g0(d)["axes"][1].__setitem__("bounds", 0),
        "irregular": lambda d: g0(d)["axes"][0].__setitem__("binning", 1),
        "wrong_type": lambda d: g0(d)["grid_link"].__setitem__("type", 1),
        "bin_out_of_range": lambda d: g0(d)["bins"][0]["loc_index"].__setitem__(0, 99),
        "rod_content": lambda d: g0(d)["bins"][0]["content"][0].__setitem__("type", 7),
        "duplicate_bin": lambda d: g0(d)["bins"][1].__setitem__("loc_index", list(g0(d)["bins"][0]["loc_index"])),
        "owner_out_of_range": lambda d: g0(d).__setitem__("owner_link", 10_000),
        "map_and_slab": lambda d: None,
    }[case]
    files, bad = _maps_file(tmp_path, mutate)
    material = files["homogeneous_material"]
    if case == "map_and_slab":
        hom = json.load(open(material))
        vol = json.load(open(bad))["data"]["grids"][0]
        owner = vol["grid_data"][0]["owner_link"]
        hom["data"]["volumes"][vol["volume_link"]]["surface_material"] = [{
            "type": 6, "surface_idx": owner, "thickness": 1.0, "index_in_coll": 0,
            "material": {"params": [93.7, 465.2, 28.0855, 14.0, 0.002329, 0.0, 0.0]}}]
        material = tmp_path / "hom.json"
        material.write_text(json.dumps(hom))
    rc, msg = read_error(files["geometry"], files["surface_grids"], material, bad)
    assert rc == code, (rc, msg)
    assert "material map" in msg or "axis" in msg


@pytest.mark.parametrize("mixed", [False, True])
def test_direct_navigator_replays_the_toy_with_maps(solenoid, mixed):
    """The direct navigator (host build of the batched engine) replays the
    caching navigator's sequences on the toy with maps, applying the same
    map bins: identical sequences and momenta at the last surface within
    the reference's direct-navigator tolerance (2e-3,
    propagation_direct_navigator.cpp:180-185); the oracle's own replay
    agrees with it."""
    toy = da.Detector.toy(4, 7, use_material="maps")
    fp = da.generate_tracks("toy", 200, 49)
    ref = oracle_host.propagate(toy, solenoid, CFG, fp)
    got = engine_host.propagate_direct(toy, solenoid, CFG, fp, ref["surfaces"], ref["n_surfaces"], mixed=mixed)
    odir = oracle_host.propagate_direct(toy, solenoid, CFG, fp, ref["surfaces"], ref["n_surfaces"])
    np.testing.assert_array_equal(got["n_surfaces"], ref["n_surfaces"])
    has = ref["n_surfaces"] > 0
    for i in np.nonzero(has)[0]:
        k = ref["n_surfaces"][i]
        np.testing.assert_array_equal(got["surfaces"][:k, i], ref["surfaces"][:k, i])
    p = lambda b: np.abs(1.0 / b[4])
    assert np.max(np.abs(p(got["bound_params"][:, has]) - p(ref["bound_params"][:, has])) /
                  p(ref["bound_params"][:, has])) < 2e-3
    tol = 1e-4 if mixed else 1e-9
    assert np.max(np.abs(p(got["bound_params"][:, has]) - p(odir["bound_params"][:, has])) /
                  p(odir["bound_params"][:, has])) < tol
