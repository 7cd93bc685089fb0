"""The walks' leaf filter drops a leaf test when a lower bound of the triangle's distance exceeds the
lane's best; it is exact only if that bound never exceeds the leaf test's own result.  This builds
tests/cpp/leaf_bound_fuzz.cpp (a host restatement of the record and bound arithmetic) and checks the
claim on random and adversarial triangle/point pairs."""
import shutil
import subprocess

import pytest

HERE = __import__("os").path.dirname(__file__)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_bound_never_exceeds_leaf_test(tmp_path):
    exe = tmp_path / "leaf_bound_fuzz"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", str(exe), f"{HERE}/cpp/leaf_bound_fuzz.cpp"],
                   check=True, timeout=120)
    for seed in (1, 2):
        r = subprocess.run([str(exe), str(seed), "400000"], capture_output=True, text=True, timeout=120)
        print(r.stdout)
        assert r.returncode == 0, r.stdout
        assert "bad 0" in r.stdout
