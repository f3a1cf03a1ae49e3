"""bench.py's command line, without a GPU: the one-GPU refusal and the
physics options (their flags must be ones the host layer accepts)."""
import os
import subprocess
import sys

import pytest

from conftest import load_package

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

pkg = load_package()


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("name", sorted(bench.PHYSICS))
def test_physics_flags_are_accepted_by_the_host(name):
    run = pkg.Run(["--npx", "24", "--npy", "16", "--npz", "24", "--npey", "2", "--npez", "2"]
                  + bench.PHYSICS[name])
    cfg = run.cfg()
    assert (cfg.itcon == 2) == (name == "heat")
    assert (cfg.id != 0) == (name == "density")
    assert (cfg.rlax > 0) == (name == "relax")
    if name != "plain":
        assert name in bench.PHYSICS_NAMES
