"""The C++20 adapter (include/detray_amd.hpp) through its compiled test,
tests/native/adapter_test.cpp, on the CPU: builders, error codes, the
actor-chain config and the detector-view flattener (the reference's
payload structs flattened in memory equal the tables read from the same
files).  The C++ examples of INTEGRATION.md are compiled against the
header."""

import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "tests" / "native" / "adapter_test"


def test_adapter_cpu_checks(tmp_path):
    if not BIN.exists():
        pytest.fail(f"{BIN} missing: run __graft_entry__.build()")
    res = subprocess.run([str(BIN), "cpu", str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "FAIL" not in res.stdout
    assert res.stdout.count("\nok ") + res.stdout.startswith("ok ") >= 15


def _cpp_blocks():
    text = (ROOT / "INTEGRATION.md").read_text()
    return [b for b in re.findall(r"```cpp\n(.*?)```", text, re.S) if "detray_amd.hpp" in b]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_integration_examples_compile(tmp_path):
    blocks = _cpp_blocks()
    assert blocks, "INTEGRATION.md has no C++ adapter example"
    for k, code in enumerate(blocks):
        src = tmp_path / f"example_{k}.cpp"
        src.write_text(code)
        res = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", f"-I{ROOT / 'include'}",
                              "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", str(src)],
                             capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, f"INTEGRATION.md example {k}:\n{res.stderr}"
