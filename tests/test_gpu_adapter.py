"""propagator<rk_stepper, caching_navigator | direct_navigator,
actor_chain<...>>::propagate of the C++20 adapter on the device, through
the compiled test tests/native/adapter_test.cpp in a subprocess: the plain
FP64 kernel and the batched mixed kernel on the toy detector, the chain with
the momentum aborter and the direct navigator, each against the oracle."""

import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent / "native" / "adapter_test"


def test_adapter_propagates_like_the_oracle():
    assert BIN.exists(), f"{BIN} missing: run __graft_entry__.build()"
    res = subprocess.run([str(BIN), "gpu"], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.count("ok ") >= 8 and "FAIL" not in res.stdout
