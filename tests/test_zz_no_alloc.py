"""r3d_hip_step_noalloc (include/r3d_hip.h) against r3d_hip_step, through the C ABI.

The entry point allocates nothing: its dt partials live in the library's
static device array, shared by every call on the GPU and holding whatever the
previous call left there.  Two calls in a row, the second on a smaller grid,
must each give the state and scalars of r3d_hip_step bit for bit.
"""
import sys

import numpy as np
import pytest

from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu

X5 = ["ru", "rv", "rw", "ro", "tt"]
Z5 = ["zru", "zrv", "zrw", "zro", "ztt"]
MAG = ["--lmag", "1", "--rm", "50", "--beta", "10", "--ampb", "1", "--bfh", "2", "--bzp", "10"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(args, steps, nstages=3, names=X5 + Z5):
    run = pkg.Run(args)
    out = []
    for noalloc in (False, True):
        sim = run.make_sim()
        if noalloc:
            sim.step_noalloc(steps, nstages)
        else:
            sim.step(steps, nstages)
        out.append((sim.download_global(names), sim.scalars()))
        sim.close()
    run.close()
    return out


def check(out, names):
    (a, sa), (b, sb) = out
    for n in names:
        assert np.array_equal(a[n].view(np.uint64), b[n].view(np.uint64)), n
    assert sa == sb, (sa, sb)


def test_noalloc_step_equals_step_twice_second_smaller():
    # first call: 64x72x128 on 2x2 (432 stage tiles + 10 boundary blocks)
    big = ["--npx", "64", "--npy", "72", "--npz", "128", "--npey", "2", "--npez", "2"]
    say("[gpu] step_noalloc vs step: 64x72x128, 5 steps")
    check(run_both(big, 5), X5 + Z5)
    # second call, smaller grid: the array still holds the first call's partials
    small = ["--npx", "24", "--npy", "16", "--npz", "48", "--npey", "2", "--npez", "2"]
    say("[gpu] step_noalloc vs step: 24x16x48, 5 steps")
    check(run_both(small, 5), X5 + Z5)


def test_noalloc_mhd_and_truncated_steps():
    args = ["--npx", "32", "--npy", "32", "--npz", "64", "--npey", "2", "--npez", "2"] + MAG
    names = X5 + ["bx", "by", "bz"] + Z5 + ["zbx", "zby", "zbz"]
    say("[gpu] step_noalloc vs step: MHD 32x32x64, 3 steps")
    check(run_both(args, 3, names=names), names)
    say("[gpu] step_noalloc vs step: MHD 32x32x64, one step truncated after 2 stages")
    check(run_both(args, 1, nstages=2, names=names), names)


def test_noalloc_refused_while_timing():
    run = pkg.Run(["--npx", "16", "--npy", "16", "--npz", "32", "--npey", "2", "--npez", "2"])
    sim = run.make_sim()
    sim.set_timing(True)
    with pytest.raises(pkg.HipError):
        sim.step_noalloc(1)
    sim.set_timing(False)
    sim.step_noalloc(1)
    sim.close()
