"""The C++ drop-in header (include/nufi/gpu_scheduler.hpp) compiles with the
host compiler alone, and its driver (tests/cpp/drop_in_1d.cpp: the reference
driver's step loop with a host round trip) agrees with the device loop."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "drop_in_1d")


def test_header_compiles_with_gxx(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include <nufi/gpu_scheduler.hpp>\n'
                   'int main() { nufi::dim1::config_t<double> c; nufi::dim3::config_t<float> d;\n'
                   '  return int(c.Nx + d.Nx) == 264 ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(tmp_path / "t")], check=True)
    subprocess.run([str(tmp_path / "t")], check=True)


@pytest.mark.gpu
@pytest.mark.parametrize("real", ["double", "float"])
def test_drop_in_driver_runs(real):
    assert os.path.exists(BIN), "build() compiles tests/cpp first"
    r = subprocess.run([BIN] + (["float"] if real == "float" else []), capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
