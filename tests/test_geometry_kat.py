"""Known-answer tests of the detector emitters against the reference's own
toy-detector tests (tests/include/detray/test/cpu/toy_detector_test.hpp) and
the navigation KAT (tests/unit_tests/cpu/navigation/caching_navigator.cpp:
128-292), run through the oracle."""

from collections import Counter

import numpy as np
import pytest

import detray_amd as da
from detray_amd import capi
from tests import oracle_host

EOW = capi.END_OF_WORLD


def test_toy_counts(toy3):
    # toy_detector_test.hpp:148-166 (default toy: 4 barrel, 3 endcap layers)
    d = toy3.desc
    assert d.n_volumes == 22
    assert d.n_surfaces == 3230
    assert d.n_transforms == 3252
    shapes = Counter(m.shape for m in toy3.table("masks"))
    assert shapes[capi.RECTANGLE2D] == 4
    assert shapes[capi.TRAPEZOID2D] == 12
    assert shapes[capi.CONCENTRIC_CYLINDER2D] == 56
    assert shapes[capi.RING2D] == 60
    sfs = toy3.table("surfaces")
    assert sum(1 for s in sfs if s.material != capi.INVALID_INDEX) == 3141
    grids = toy3.table("grids")
    assert sum(1 for g in grids if g.type == 0) == 4
    assert sum(1 for g in grids if g.type == 1) == 6


# (volume, portal range, sensitive range, passive range) from
# toy_detector_test.hpp:399-1315
SF_RANGES = [
    (0, (1, 6), None, (0, 1)),
    (1, (114, 118), (6, 114), None),
    (2, (118, 122), None, None),
    (3, (230, 234), (122, 230), None),
    (4, (234, 238), None, None),
    (5, (346, 350), (238, 346), None),
    (6, (350, 354), None, None),
    (7, (578, 582), (354, 578), None),
    (8, (582, 586), None, None),
    (9, (1034, 1038), (586, 1034), None),
    (10, (1038, 1042), None, None),
    (11, (1770, 1774), (1042, 1770), None),
    (21, (3226, 3230), None, None),
]


@pytest.mark.parametrize("vol,portals,sens,passive", SF_RANGES)
def test_toy_surface_ranges(toy3, vol, portals, sens, passive):
    sfs = toy3.table("surfaces")

    def rng(kind):
        idx = [i for i, s in enumerate(sfs) if s.volume == vol and s.sf_id == kind]
        return (min(idx), max(idx) + 1) if idx else None

    assert rng(capi.PORTAL) == portals
    assert rng(capi.SENSITIVE) == sens
    assert rng(capi.PASSIVE) == passive


def _links(det, sf):
    s = det.table("surfaces")[sf]
    masks = det.table("masks")
    return [masks[s.mask + k].volume_link for k in range(s.n_masks)]


# mask volume links of portals (toy_detector_test.hpp test_portal_links)
PORTAL_LINKS = [
    (1, [EOW]),                      # beampipe: left disc
    (2, [1, 2, 3, 4, 5, 6]),         # beampipe: cylinder, neg. endcap
    (3, [8]),                        # beampipe: central cylinder
    (4, [EOW]),                      # beampipe: right disc
    (5, [16, 17, 18, 19, 20, 21]),   # beampipe: cylinder, pos. endcap
    (114, [0]), (115, [EOW]),        # endcap_1 cylinders
    (116, [4]), (117, [2]),          # endcap_1 discs
    (118, [7, 8, 9, 10, 11, 12, 13, 14, 15]),  # connector: barrel-facing disc
    (119, [1]),                      # connector: outward disc
]


@pytest.mark.parametrize("sf,links", PORTAL_LINKS)
def test_toy_portal_links(toy3, sf, links):
    assert _links(toy3, sf) == links


def test_toy_module_masks(toy3):
    sfs = toy3.table("surfaces")
    masks = toy3.table("masks")
    # barrel layer 1 modules share one rectangle (8.4f, 36f)
    m = masks[sfs[354].mask]
    assert m.shape == capi.RECTANGLE2D
    assert m.v[0] == float(np.float32(8.4)) and m.v[1] == 36.0
    assert sfs[577].mask == sfs[354].mask
    # endcap rings: trapezoids {3, 9.5, 39} and {6, 10, 39}
    inner, outer = masks[sfs[6].mask], masks[sfs[46].mask]
    assert list(inner.v[:3]) == [3.0, 9.5, 39.0] and inner.v[3] == pytest.approx(1 / 78)
    assert list(outer.v[:3]) == [6.0, 10.0, 39.0]
    # beampipe: r = 19 mm, |z| <= last endcap position (820 mm for 3 layers)
    bp = masks[sfs[0].mask]
    assert list(bp.v[:3]) == [19.0, -820.0, 820.0]
    assert sfs[0].sf_id == capi.PASSIVE and sfs[0].material != capi.INVALID_INDEX


def test_barcodes(toy3):
    sfs = toy3.table("surfaces")
    for idx in (0, 475, 2294, 3229):
        b = sfs[idx].barcode
        assert (b >> 52) & 0xFFF == sfs[idx].volume
        assert (b >> 48) & 0xF == sfs[idx].sf_id
        assert (b >> 27) & 0x1FFFFF == idx
        assert (b >> 6) & 0x1FFFFF == sfs[idx].transform


def test_navigation_kat_ray(toy3):
    # caching_navigator.cpp:128-292: straight ray (1,1,0) from the origin
    cfg = da.PropagationConfig(search_window=(3, 3), use_material=False,
                               estimate_scattering_noise=False)
    field = da.MagneticField.constant(0.0, 0.0, 0.0)
    fp = np.zeros((8, 1))
    fp[4, 0] = fp[5, 0] = 1.0 / np.sqrt(2.0)
    fp[7, 0] = -1.0 / np.sqrt(2.0)
    out = oracle_host.propagate(toy3, field, cfg, fp, trace_capacity=128)
    seq = out["surfaces"][: out["n_surfaces"][0], 0].tolist()
    assert seq == [0, 475, 459, 476, 460, 829, 797, 830, 798, 1438, 1386, 2372, 2294]
    n = int(out["trace_n"][0])
    vols, sfs = out["trace_volume"][0, :n], out["trace_surface"][0, :n]
    order = []
    for v in vols.tolist():
        if not order or order[-1] != v:
            order.append(v)
    assert order == [0, 8, 7, 10, 9, 12, 11, 14, 13, 15]
    # portals the navigator stands on after each volume switch
    entry = {v: int(s) for v, s in zip(vols.tolist(), sfs.tolist())
             if s != capi.INVALID_INDEX and v not in (0,)}
    visited = set(int(s) for s in sfs.tolist() if s != capi.INVALID_INDEX)
    for sf in (582, 578, 1038, 1034, 1774, 1770, 2874, 2870, 2878):
        assert sf in visited, sf
    assert out["status"][0] == capi.TRACK_EXIT
    assert out["path_length"][0] == pytest.approx(180.0, abs=1e-6)
    assert entry


def test_telescope_layout(telescope):
    sfs = telescope.table("surfaces")
    trfs = telescope.table("transforms")
    assert telescope.desc.n_volumes == 1
    assert sum(1 for s in sfs if s.sf_id == capi.SENSITIVE) == 10
    assert sum(1 for s in sfs if s.sf_id == capi.PORTAL) == 6
    zs = [trfs[s.transform].t[2] for s in sfs if s.sf_id == capi.SENSITIVE]
    assert zs == [100.0 * i for i in range(1, 11)]
    # the portal box encloses the origin (envelope 60 mm)
    pz = sorted(trfs[s.transform].t[2] for s in sfs if s.sf_id == capi.PORTAL)
    assert pz[0] < 0.0 < pz[-1]
