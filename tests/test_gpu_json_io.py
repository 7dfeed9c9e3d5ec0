"""Device propagation through a detector read from detray JSON files equals
propagation through the emitted tables, bit for bit (same tables, same
kernels), on the plain FP64 kernel and the batched mixed kernel."""

import numpy as np
import pytest

import detray_amd as da

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("launch", [None, da.LaunchConfig("mixed", "persistent", 256, 2, 64, None, 1, "wave")])
def test_json_detector_propagates_identically(tmp_path, toy7, solenoid, launch):
    f = toy7.write_json(tmp_path, "toy")
    det = da.Detector.from_json(f["geometry"], f["surface_grids"], f["homogeneous_material"])
    cfg = da.PropagationConfig(search_window=(3, 3))
    tracks = torch.from_numpy(da.generate_tracks("toy", 4096, 43)).cuda()
    a = da.Propagator(cfg, launch).propagate(toy7, solenoid, tracks)
    b = da.Propagator(cfg, launch).propagate(det, solenoid, tracks)
    torch.cuda.synchronize()
    for k in ("bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces", "status"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    det.close()
