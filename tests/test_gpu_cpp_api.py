"""The C++ drop-in classes (include/felt_amd/Surface.hpp, device.hpp) driving the HIP library."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name", ["surface_api", "device_functor"])
def test_cpp_program(name):
    """surface_api: the drop-in classes with host functors and device descriptors;
    device_functor: __device__ functors compiled into the caller's translation unit
    (include/felt_amd/device.hpp), bit for bit against the library's device speeds and the
    host-functor path."""
    exe = os.path.join(HERE, "cpp", name + ".bin")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
