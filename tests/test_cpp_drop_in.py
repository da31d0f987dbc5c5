"""The C++ drop-in header (include/nufi/gpu_scheduler.hpp) compiles with the
host compiler alone, and its drivers (tests/cpp/drop_in_1d.cpp and
drop_in_3d.cpp: the reference drivers' step loops with a host round trip)
agree with the device loop."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


def test_header_compiles_with_gxx(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include <nufi/gpu_scheduler.hpp>\n'
                   'int main() { nufi::dim1::config_t<double> c; nufi::dim3::config_t<float> d;\n'
                   '  return int(c.Nx + d.Nx) == 264 ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(tmp_path / "t")], check=True)
    subprocess.run([str(tmp_path / "t")], check=True)


@pytest.mark.gpu
@pytest.mark.parametrize("dim", ["1d", "3d"])
@pytest.mark.parametrize("real", ["double", "float"])
def test_drop_in_driver_runs(dim, real):
    binary = os.path.join(BUILD, f"drop_in_{dim}")
    assert os.path.exists(binary), "build() compiles tests/cpp first"
    r = subprocess.run([binary] + (["float"] if real == "float" else []), capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
