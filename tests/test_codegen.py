"""Code-generation guards for the stage kernel's memory path (no GPU needed).

The stage kernel is bound by how its HBM traffic flows.  Three properties of
the generated gfx950 code decide that, and each was once lost by an
innocent-looking source change:

* no FLAT memory instructions: a FLAT access also counts on the LDS counter,
  so the next LDS read would wait until it completed (a pointer selected
  between two buffers compiles to FLAT);
* no scratch (spill) traffic inside the flux waves' z loop: vector memory
  operations complete in issue order, so a spill reload after a plane's
  stores waits for all of them;
* no full drain (s_waitcnt vmcnt(0)) inside that loop, for the same reason.

The kernels are compiled to assembly with the library's own flags and the
largest loop of each full-step stage kernel (the flux waves' z loop) is
checked.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "rast-3dmhd_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None,
                                reason="hipcc not available")


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    out = tmp_path_factory.mktemp("cg") / "r3d_hip.s"
    cc = HIPCC if os.path.exists(HIPCC) else shutil.which("hipcc")
    cmd = [cc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "--cuda-device-only", "-S",
           "-o", str(out), os.path.join(CSRC, "r3d_hip.hip")]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text()


def kernel_body(asm, ty, stage, zmode, mag=False, heat=False, den=False, tw=None):
    """tw: tile columns (default: 32 for the 12-row MHD tile, else 64)."""
    if tw is None:
        tw = 32 if (mag and ty == 12) else 64
    name = (f"_ZN3r3d7k_stageILi{ty}ELi{stage}ELi{zmode}ELb{int(mag)}ELb{int(heat)}"
            f"ELb{int(den)}ELi{tw}EEEvNS_9StageArgsENS_6ConstsENS_7MetricsE")
    i = asm.index(name + ":")
    return asm[i:asm.index(".Lfunc_end", i)].split("\n")


def largest_loop(lines):
    """Lines of the outermost loop with the most lines (blocks labelled as in
    that loop, header included)."""
    best = []
    for h in [n for n, l in enumerate(lines) if "Loop Header" in l and "Depth=1" in l]:
        tag = lines[h].split(":")[0].strip().replace(".L", "")
        cur, rng = False, []
        for n, l in enumerate(lines):
            if re.match(r"^\.LBB|^; %bb", l):
                cur = tag in l
            if cur:
                rng.append(l.strip())
        if len(rng) > len(best):
            best = rng
    return best


FULL_STEP = [(0, 0), (1, 0), (2, 2)]  # (stage, zmode) of the three stages of a full step


@pytest.mark.parametrize("ty", [8, 9])
@pytest.mark.parametrize("stage,zmode", FULL_STEP)
def test_hydro_stage_kernel_memory_path(asm, ty, stage, zmode):
    lines = kernel_body(asm, ty, stage, zmode)
    assert not [l for l in lines if re.match(r"\s+flat_", l)], "FLAT memory instruction"
    loop = largest_loop(lines)
    assert len(loop) > 300, "flux loop not found"
    assert not [l for l in loop if l.startswith("scratch_")], "scratch access in the flux loop"
    assert not [l for l in loop if l.startswith("s_waitcnt") and "vmcnt(0)" in l], \
        "full vmcnt drain in the flux loop"


@pytest.mark.parametrize("stage,zmode", FULL_STEP)
@pytest.mark.parametrize("mag,heat,den", [(True, False, False), (False, True, False),
                                          (True, True, False), (False, False, True),
                                          (True, False, True), (False, True, True)],
                         ids=["mhd", "hydro-heat", "mhd-heat", "hydro-den", "mhd-den",
                              "hydro-heat-den"])
def test_other_stage_kernels_have_no_flat_access(asm, stage, zmode, mag, heat, den):
    lines = kernel_body(asm, 12 if mag else 8, stage, zmode, mag=mag, heat=heat, den=den)
    assert not [l for l in lines if re.match(r"\s+flat_", l)], "FLAT memory instruction"


@pytest.mark.parametrize("ty,mag", [(8, False), (9, False), (12, True), (5, True)],
                         ids=["hydro8", "hydro9", "mhd12x32", "mhd5x64"])
@pytest.mark.parametrize("stage,zmode", FULL_STEP)
def test_no_store_drain_after_the_plane_barrier(asm, ty, mag, stage, zmode):
    """Between a plane's barrier and its first RK-update store, no wait may
    leave all of the plane's own loads in flight: such a wait (vmcnt(N) with
    N >= the loads issued this plane) only drains the previous plane's stores.
    It appears when a load issued before the z loop is still pending at the
    loop head."""
    loop = largest_loop(kernel_body(asm, ty, stage, zmode, mag=mag))
    bar = next(n for n, l in enumerate(loop) if l.startswith("s_barrier"))
    # the loop body in execution order from the barrier round to the barrier
    # (the compiler may place the post-barrier blocks first)
    cyc = loop[bar + 1:] + loop[:bar + 1]
    stores = [n for n, l in enumerate(cyc) if l.startswith("global_store")]
    # loads issued after the plane's last store: in flight at the barrier,
    # behind the stores
    nload = len([l for l in cyc[stores[-1]:] if l.startswith("global_load")])
    drains = [l for l in cyc[:stores[0]]
              if l.startswith("s_waitcnt") and "vmcnt" in l
              and int(re.search(r"vmcnt\((\d+)\)", l).group(1)) >= nload]
    assert nload >= 5, "plane loads not found before the barrier"
    assert not drains, f"store drain after the barrier: {drains} ({nload} loads in flight)"
