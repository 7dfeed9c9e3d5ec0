"""Device propagation through a detector read from detray JSON files: against
the oracle running on the emitted tables (the plain FP64 kernel at the FP64
contract, the batched mixed kernel at the mixed contract), and bit for bit
equal to device propagation through the emitted tables."""

import numpy as np
import pytest

import detray_amd as da
from tests import oracle_host
from tests.test_gpu_batched import assert_mixed_parity
from tests.test_gpu_parity import assert_parity

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _host(res):
    return {
        "bound_params": res.bound_params.cpu().numpy(),
        "bound_cov": res.bound_cov.cpu().numpy(),
        "free_params": res.free_params.cpu().numpy(),
        "path_length": res.path_length.cpu().numpy(),
        "surfaces": res.surfaces.cpu().numpy().view(np.uint32),
        "n_surfaces": res.n_surfaces.cpu().numpy().view(np.uint32),
        "n_steps": res.n_steps.cpu().numpy().view(np.uint32),
        "n_trials": res.n_trials.cpu().numpy().view(np.uint32),
        "status": res.status.cpu().numpy(),
    }


@pytest.mark.parametrize("launch", [None, da.LaunchConfig("mixed", "persistent", 256, 2, 64, None, 1, "wave")])
def test_json_detector_against_oracle(tmp_path, toy7, solenoid, launch):
    f = toy7.write_json(tmp_path, "toy")
    det = da.Detector.from_json(f["geometry"], f["surface_grids"], f["homogeneous_material"])
    cfg = da.PropagationConfig(search_window=(3, 3))
    fp = da.generate_tracks("toy", 4096, 43)
    tracks = torch.from_numpy(fp).cuda()
    b = da.Propagator(cfg, launch).propagate(det, solenoid, tracks)
    a = da.Propagator(cfg, launch).propagate(toy7, solenoid, tracks)
    torch.cuda.synchronize()
    ref = oracle_host.propagate(toy7, solenoid, cfg, fp)
    if launch is None:
        assert_parity(_host(b), ref)
    else:
        assert_mixed_parity(_host(b), ref)
    for k in ("bound_params", "bound_cov", "free_params", "path_length", "surfaces", "n_surfaces", "status"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    det.close()
