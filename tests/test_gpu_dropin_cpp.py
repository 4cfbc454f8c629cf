"""The reference's sdf_sign_test / voxelize_test, compiled as C++ against include/italy_bake.hpp."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_dropins():
    exe = os.path.join(ROOT, "tests", "cpp", "dropin_test")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.dirname(exe)], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "dropin_test: OK" in r.stdout
