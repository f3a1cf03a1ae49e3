"""The HIP path's LMAG initial state and time step against the reference's own code.

The reference stops in FLUXES with LMAG (src/fluxes.cpp:539-541), but what it
computes before that stop is its own code, recorded in
tests/golden/ref_lmag.json (tests/golden/make_golden.py):

* gstate.start with the magnetic layer (src/static_ic.cpp:94-108): the
  product's host setup uploaded to the device must download as the
  reference's start dump, every rank, ghosts included, at the reference's
  golden geometry and at the rotating-MHD workload size 252x252x512 (2x4);
* the time step dt and umach of STEP (src/step.cpp:121-229) with the fast
  magnetosonic speed and the resistive limit, at rest and with the
  deterministic flow and transverse field of oracle/ref_stepdt.cpp.  dt is
  read after a step truncated after its first stage, which uses the dt of the
  uploaded state.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from conftest import load_package

pkg = load_package()
pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
X = ["ru", "rv", "rw", "ro", "tt", "bx", "by", "bz"]
Z = ["zru", "zrv", "zrw", "zro", "ztt", "zbx", "zby", "zbz"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ref():
    with open(os.path.join(HERE, "golden", "ref_lmag.json")) as f:
        return json.load(f)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:20]


@pytest.mark.parametrize("name", ["m16_2x4", "m252_2x4"])
def test_hip_lmag_start_state_matches_reference_dump(ref, name):
    cfg = ref["start"][name]
    args = cfg["args"] + ["--npey", str(cfg["npey"]), "--npez", str(cfg["npez"])]
    say(f"[gpu] LMAG start state {name}: {' '.join(args)}")
    run = pkg.Run(args)
    sim = run.make_sim()
    run.close()
    bad = []
    for r, rec in enumerate(cfg["dumps"]["0"]):
        got = sim.download_rank(r, X + Z)
        for n in X + Z:
            if digest(got[n]) != rec["fields"][n]:
                bad.append((r, n))
    sim.close()
    assert not bad, bad[:8]


def flowed_state(run):
    """The global initial state with the flow of oracle/ref_stepdt.cpp on the
    interior (global ghost-inclusive indices, as the driver's per-rank ones)."""
    st = {n: a.copy() for n, a in run.state().items() if a is not None}
    npz, npy, npx = (s - 2 for s in st["ro"].shape)
    k = np.arange(1, npz + 1, dtype=np.int64)[:, None, None]
    j = np.arange(1, npy + 1, dtype=np.int64)[None, :, None]
    i = np.arange(1, npx + 1, dtype=np.int64)[None, None, :]
    sl = (slice(1, -1),) * 3
    ro = st["ro"][sl]
    for c, n in enumerate(("ru", "rv", "rw")):
        w = ((i * 7919 + j * 104729 + k * 1299709 + c * 15485863) % 10007) / 10007.0 - 0.5
        st[n][sl] = ro * (w * 0.4)
    if "by" in st:
        for c, n in ((3, "by"), (4, "bz")):
            w = ((i * 7919 + j * 104729 + k * 1299709 + c * 15485863) % 10007) / 10007.0 - 0.5
            st[n][sl] = w * 4.0
    return st


STEPDT = ["sdt_m16_rest", "sdt_m16_flow", "sdt_m16_rm", "sdt_m16_beta", "sdt_m16_adv_rest",
          "sdt_m16_adv_flow", "sdt_m16_adv_beta", "sdt_m24_flow_1x2", "sdt_h24_flow",
          "sdt_m252_flow"]


@pytest.mark.parametrize("name", STEPDT)
def test_hip_time_step_matches_reference(ref, name):
    cfg = ref["stepdt"][name]
    args = cfg["args"] + ["--npey", str(cfg["npey"]), "--npez", str(cfg["npez"])]
    say(f"[gpu] STEP limits {name}: flow={cfg['flow']} {' '.join(args)}")
    run = pkg.Run(args)
    sim = pkg.HipSim(run.cfg())
    sim.set_metrics(run.metrics())
    sim.upload_global(flowed_state(run) if cfg["flow"] else run.state(copy=False))
    sim.set_scalars(**run.scalars())
    sim.step(1, nstages=1)
    sc = sim.scalars()
    sim.close()
    run.close()
    want = cfg["ranks"][0]
    assert sc["dt"].hex() == want["dt"], (sc["dt"], float.fromhex(want["dt"]))
    assert sc["umach"].hex() == want["umach"], (sc["umach"], float.fromhex(want["umach"]))
