"""The C ABI: every entry point declared in include/*.h is exported by the built
libraries, the libraries load without a GPU, configuration errors come back as
codes with a message (no process exit), and the C++ layer include/r3d_gpu.hpp
compiles against the C header.  No call here touches a device.
"""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT, load_package

pkg = load_package()

HEADERS = {
    "r3d_hip.h": os.path.join(ROOT, "rast-3dmhd_amd", "lib", "librast3dmhd_hip.so"),
    "r3d_host.h": os.path.join(ROOT, "rast-3dmhd_amd", "lib", "librast3dmhd_host.so"),
    "r3d_ens.h": os.path.join(ROOT, "rast-3dmhd_amd", "lib", "librast3dmhd_hip.so"),
    "r3d_adj.h": os.path.join(ROOT, "rast-3dmhd_amd", "lib", "librast3dmhd_hip.so"),
}


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(r3d_\w+)\s*\(", text, flags=re.M)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_every_declared_symbol_is_exported(header):
    lib = HEADERS[header]
    assert os.path.exists(lib), f"{lib} not built"
    names = declared(header)
    assert len(names) >= 8, names
    missing = [n for n in names if n not in exported(lib)]
    assert not missing, missing


def test_create_rejects_bad_config_without_a_device():
    L = pkg.load()
    cfg = pkg.Cfg()
    cfg.npx, cfg.npy, cfg.npz, cfg.npey, cfg.npez = 16, 16, 32, 3, 2  # 3 does not divide 16
    h = ctypes.c_void_p()
    rc = L.r3d_hip_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert b"must divide" in L.r3d_hip_last_error()
    cfg.npey = 2
    cfg.variant = 7
    assert L.r3d_hip_create(ctypes.byref(cfg), ctypes.byref(h)) == -5
    assert L.r3d_hip_create(None, ctypes.byref(h)) == -1
    # calls on a null context are errors, not crashes
    assert L.r3d_hip_step(None, 1, 3) == -1
    assert L.r3d_hip_destroy(None) == 0
    assert L.r3d_ens_step(None, 1) == -1
    assert L.r3d_ens_create(0, None, ctypes.byref(h)) == -1
    assert L.r3d_ens_destroy(None) == 0


@pytest.mark.parametrize("bad", [(7, 0, 0), (5, 0, 0), (9, 0, 1), (0, 24, 0), (0, 256, 0)])
def test_create_rejects_unbuilt_tilings_without_a_device(bad):
    """tile_rows / zchunk outside the built kernels are configuration errors."""
    L = pkg.load()
    cfg = pkg.Cfg()
    cfg.npx, cfg.npy, cfg.npz, cfg.npey, cfg.npez = 16, 16, 32, 2, 2
    cfg.tile_rows, cfg.zchunk, cfg.lmag = bad
    h = ctypes.c_void_p()
    assert L.r3d_hip_create(ctypes.byref(cfg), ctypes.byref(h)) == -1 and not h.value
    assert b"tile_rows" in L.r3d_hip_last_error() or b"zchunk" in L.r3d_hip_last_error()


@pytest.mark.parametrize("margin", [-1.0, -2.0, float("nan")])
def test_create_rejects_a_time_step_screen_margin_of_minus_one_or_below(margin):
    """dt_screen scales the previous extrema by 1 +/- margin: it must be > -1."""
    L = pkg.load()
    cfg = pkg.Cfg()
    cfg.npx, cfg.npy, cfg.npz, cfg.npey, cfg.npez = 16, 16, 32, 2, 2
    cfg.dt_screen = margin
    h = ctypes.c_void_p()
    assert L.r3d_hip_create(ctypes.byref(cfg), ctypes.byref(h)) == -1 and not h.value
    assert b"dt_screen" in L.r3d_hip_last_error()


def test_ensemble_rejects_mixed_members_without_a_device():
    L = pkg.load()
    cfgs = (pkg.Cfg * 2)()
    cfgs[1].lmag = 1
    h = ctypes.c_void_p()
    assert L.r3d_ens_create(2, cfgs, ctypes.byref(h)) == -1
    assert b"same device and lmag" in L.r3d_hip_last_error()
    cfgs[1].lmag = 0
    cfgs[0].tc = 1.0
    assert L.r3d_ens_create(2, cfgs, ctypes.byref(h)) == -5


def test_ensemble_members_are_seeded_and_in_range():
    a = pkg.ensemble_members(64)
    assert a == pkg.ensemble_members(64)
    assert len({tuple(x) for x in a}) == 64
    for m, args in enumerate(a):
        d = dict(zip(args[::2], args[1::2]))
        assert d["--npx"] == d["--npy"] and int(d["--npx"]) in (48, 64, 80, 96)
        assert int(d["--npz"]) in (96, 128, 160, 192)
        assert 1.0 <= float(d["--re"]) <= 4.0 and 0.1 <= float(d["--theta"]) <= 0.3
        assert d["--seed"] == str(-(1001 + m))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cpp_layer_compiles(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "r3d_gpu.hpp"\n'
        "int main() {\n"
        "    r3d_hip_cfg cfg{};\n"
        "    try { r3d_amd::Device d(cfg); } catch (const r3d_amd::Error& e) {\n"
        "        return e.code == R3D_EINVAL ? 0 : 1; }\n"
        "    return 2;\n"
        "}\n")
    exe = tmp_path / "use"
    libdir = os.path.join(ROOT, "rast-3dmhd_amd", "lib")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", libdir, "-lrast3dmhd_hip", f"-Wl,-rpath,{libdir}"],
                   check=True)
    # an all-zero configuration is rejected before any device call
    assert subprocess.run([str(exe)]).returncode == 0
