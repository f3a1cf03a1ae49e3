"""rast3dmhd_amd - thin Python binding of the gfx950 Rast-3DMHD time-step library.

The product is the C ABI in include/r3d_hip.h (lib/librast3dmhd_hip.so) and
the C++ host library (lib/librast3dmhd_host.so); this module only marshals
arguments through ctypes.  It never falls back to a CPU path: if the HIP
library is missing, load() raises.

The package directory is named ``rast-3dmhd_amd`` (not an importable
identifier); load it with ``rast3dmhd_amd = load_package()`` from
__graft_entry__, or importlib with a file location.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
HIP_LIB = os.path.join(LIB_DIR, "librast3dmhd_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "librast3dmhd_host.so")

R3D_RU, R3D_RV, R3D_RW, R3D_RO, R3D_TT, R3D_BX, R3D_BY, R3D_BZ = range(8)
FIELD_NAMES = ["ru", "rv", "rw", "ro", "tt", "bx", "by", "bz",
               "zru", "zrv", "zrw", "zro", "ztt", "zbx", "zby", "zbz"]
PRIMARIES = FIELD_NAMES[:8]


class HipError(RuntimeError):
    """A non-zero status from the C ABI; .code holds the R3D_E* value."""

    def __init__(self, code, msg):
        super().__init__(f"r3d_hip error {code}: {msg}")
        self.code = code


class Cfg(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("npx", "npy", "npz", "npey", "npez", "lmag", "lrot", "itcon", "izcon",
                 "ibcon", "variant", "device")] + \
               [(n, ctypes.c_double) for n in
                ("c13", "c23", "c43", "gam1", "gam2", "gam3", "zeta1", "zeta2",
                 "hx", "h2x", "hy", "h2y", "hz", "h2z",
                 "re", "repr", "ore", "cv", "ocv", "gamma", "grav", "theta", "sff",
                 "rm", "orm", "obeta", "omx", "omz",
                 "tp", "tc", "hh", "xp", "yp", "dzf", "tb", "zp", "qfh", "dh", "dv", "zmax")] + \
               [("id", ctypes.c_int)] + \
               [(n, ctypes.c_double) for n in ("rlax", "tstart", "toff")] + \
               [(n, ctypes.c_int) for n in ("tile_rows", "zchunk", "lrem")] + \
               [(n, ctypes.c_double) for n in ("tu", "dztu", "dztb", "dt_screen")]


METRIC_NAMES = ["exx", "dxxdx", "d2xxdx2", "ddx", "wyy", "dyydy", "d2yydy2", "ddy",
                "zee", "dzzdz", "d2zzdz2", "ddz", "rkapa", "dkapa"]


class Metrics(ctypes.Structure):
    _fields_ = [(n, ctypes.POINTER(ctypes.c_double)) for n in METRIC_NAMES]


class Scalars(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("dt", "timc", "timt", "timi", "umach")] + \
               [("nit", ctypes.c_int)]


_hip = None


def load():
    """Load the HIP library (raises if it is not built)."""
    global _hip
    if _hip is not None:
        return _hip
    if not os.path.exists(HIP_LIB):
        raise RuntimeError(f"HIP library not built: {HIP_LIB} (run make -C rast-3dmhd_amd)")
    # torch, when installed, is loaded first: it ships a HIP runtime under the
    # soname this library links (libamdhip64.so.7), the first one loaded
    # serves the whole process, and torch's GPU tensors (the adjoint's torch
    # binding) need torch's own
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(HIP_LIB)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    L.r3d_hip_last_error.restype = ctypes.c_char_p
    sigs = {
        "r3d_hip_create": [ctypes.POINTER(Cfg), ctypes.POINTER(vp)],
        "r3d_hip_destroy": [vp],
        "r3d_hip_set_metrics": [vp, ctypes.POINTER(Metrics)],
        "r3d_hip_upload_rank": [vp, ip, ctypes.POINTER(dp)],
        "r3d_hip_upload_global": [vp, ctypes.POINTER(dp)],
        "r3d_hip_download_rank": [vp, ip, ctypes.POINTER(dp)],
        "r3d_hip_download_global": [vp, ctypes.POINTER(dp)],
        "r3d_hip_set_scalars": [vp, ctypes.POINTER(Scalars)],
        "r3d_hip_get_scalars": [vp, ctypes.POINTER(Scalars)],
        "r3d_hip_snapshot_begin": [vp],
        "r3d_hip_snapshot_rank": [vp, ip, ctypes.POINTER(dp), ctypes.POINTER(Scalars)],
        "r3d_hip_step": [vp, ip, ip],
        "r3d_hip_step_noalloc": [vp, ip, ip],
        "r3d_hip_sync": [vp],
        "r3d_hip_set_timing": [vp, ip],
        "r3d_hip_kernel_stats": [vp, ip, ctypes.POINTER(ctypes.c_long),
                                 ctypes.POINTER(ctypes.c_double)],
        "r3d_hip_memory": [vp, ctypes.POINTER(ctypes.c_ulonglong)],
        "r3d_hip_tiling": [vp, ctypes.POINTER(ip)],
        # batched ensemble (include/r3d_ens.h)
        "r3d_ens_create": [ip, ctypes.POINTER(Cfg), ctypes.POINTER(vp)],
        "r3d_ens_destroy": [vp],
        "r3d_ens_set_metrics": [vp, ip, ctypes.POINTER(Metrics)],
        "r3d_ens_upload_global": [vp, ip, ctypes.POINTER(dp)],
        "r3d_ens_download_global": [vp, ip, ctypes.POINTER(dp)],
        "r3d_ens_download_rank": [vp, ip, ip, ctypes.POINTER(dp)],
        "r3d_ens_set_scalars": [vp, ip, ctypes.POINTER(Scalars)],
        "r3d_ens_get_scalars": [vp, ip, ctypes.POINTER(Scalars)],
        "r3d_ens_set_tiles": [vp, ip, ip],
        "r3d_ens_get_tiles": [vp, ctypes.POINTER(ip), ctypes.POINTER(ip), ctypes.POINTER(ip)],
        "r3d_ens_set_tile_shape": [vp, ip, ip, ip],
        "r3d_ens_get_tile_shape": [vp, ctypes.POINTER(ip), ctypes.POINTER(ip),
                                   ctypes.POINTER(ip), ctypes.POINTER(ip)],
        "r3d_ens_autotune": [vp, ip, ctypes.POINTER(ctypes.c_double)],
        "r3d_ens_set_graph": [vp, ip],
        "r3d_ens_set_persistent": [vp, ip],
        "r3d_ens_step": [vp, ip],
        "r3d_ens_sync": [vp],
        "r3d_ens_memory": [vp, ctypes.POINTER(ctypes.c_ulonglong)],
        # discrete adjoint (include/r3d_adj.h)
        "r3d_adj_create": [ctypes.POINTER(Cfg), ip, ctypes.POINTER(vp)],
        "r3d_adj_destroy": [vp],
        "r3d_adj_set_metrics": [vp, ctypes.POINTER(Metrics)],
        "r3d_adj_set_initial": [vp, ctypes.POINTER(dp)],
        "r3d_adj_forward": [vp, ip, ctypes.POINTER(ctypes.c_double)],
        "r3d_adj_backward": [vp, ctypes.POINTER(dp), ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_double)],
        "r3d_adj_timing": [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
        "r3d_adj_set_fixed_dt": [vp, ctypes.c_double],
        "r3d_adj_dJ_dtb": [vp, ctypes.POINTER(ctypes.c_double)],
        "r3d_adj_download_final": [vp, ctypes.POINTER(dp)],
        "r3d_adj_set_initial_device": [vp, ctypes.POINTER(vp)],
        "r3d_adj_backward_device": [vp, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _hip = L
    return L


def _check(rc):
    if rc != 0:
        raise HipError(rc, load().r3d_hip_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class HipSim:
    """One device context (the whole global domain of a run on one GPU)."""

    def __init__(self, cfg):
        L = load()
        self.cfg = cfg
        h = ctypes.c_void_p()
        _check(L.r3d_hip_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.npey, self.npez = cfg.npey, cfg.npez
        self.nx = cfg.npx + 2
        self.ny = cfg.npy // cfg.npey + 2
        self.nz = cfg.npz // cfg.npez + 2
        self.nf = 8 if cfg.lmag else 5
        self._writer = None  # dump thread reading the snapshot buffers (Run.dump_async)

    def _join_writer(self):
        """Wait for a dump thread still reading this context's snapshot: the
        next snapshot_begin overwrites its pinned buffer and destroy frees it."""
        w, self._writer = getattr(self, "_writer", None), None
        if w is not None and w.is_alive():
            w.join()

    def close(self):
        if self.h:
            self._join_writer()
            load().r3d_hip_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_metrics(self, m):
        """m: dict of global 1-D metric arrays (METRIC_NAMES)."""
        keep = {k: np.ascontiguousarray(m[k], dtype=np.float64) for k in METRIC_NAMES}
        st = Metrics(*[_dptr(keep[k]) for k in METRIC_NAMES])
        _check(load().r3d_hip_set_metrics(self.h, ctypes.byref(st)))

    def upload_rank(self, rank, fields):
        arrs = [np.ascontiguousarray(fields[n], dtype=np.float64) for n in PRIMARIES[:self.nf]]
        for a in arrs:
            assert a.size == self.nx * self.ny * self.nz, "rank array has the wrong size"
        ptrs = (ctypes.POINTER(ctypes.c_double) * 8)(*[_dptr(a) for a in arrs] +
                                                    [None] * (8 - len(arrs)))
        _check(load().r3d_hip_upload_rank(self.h, rank, ptrs))

    def upload_global(self, fields):
        shape = (self.cfg.npz + 2, self.cfg.npy + 2, self.cfg.npx + 2)
        arrs = [np.ascontiguousarray(fields[n], dtype=np.float64) for n in PRIMARIES[:self.nf]]
        for a in arrs:
            assert a.size == np.prod(shape)
        ptrs = (ctypes.POINTER(ctypes.c_double) * 8)(*[_dptr(a) for a in arrs] +
                                                    [None] * (8 - len(arrs)))
        _check(load().r3d_hip_upload_global(self.h, ptrs))

    def download_rank(self, rank, names=None):
        names = names or FIELD_NAMES
        out = {n: np.empty((self.nz, self.ny, self.nx)) for n in names}
        ptrs = (ctypes.POINTER(ctypes.c_double) * 16)()
        for n in names:
            ptrs[FIELD_NAMES.index(n)] = _dptr(out[n])
        _check(load().r3d_hip_download_rank(self.h, rank, ptrs))
        return out

    def download_global(self, names=None):
        names = names or FIELD_NAMES
        shape = (self.cfg.npz + 2, self.cfg.npy + 2, self.cfg.npx + 2)
        out = {n: np.empty(shape) for n in names}
        ptrs = (ctypes.POINTER(ctypes.c_double) * 16)()
        for n in names:
            ptrs[FIELD_NAMES.index(n)] = _dptr(out[n])
        _check(load().r3d_hip_download_global(self.h, ptrs))
        return out

    def set_scalars(self, dt=0.0, timc=0.0, timt=0.0, timi=0.0, umach=0.0, nit=0):
        s = Scalars(dt, timc, timt, timi, umach, int(nit))
        _check(load().r3d_hip_set_scalars(self.h, ctypes.byref(s)))

    def scalars(self):
        s = Scalars()
        _check(load().r3d_hip_get_scalars(self.h, ctypes.byref(s)))
        return {"dt": s.dt, "timc": s.timc, "timt": s.timt, "timi": s.timi,
                "umach": s.umach, "nit": s.nit}

    def step(self, nsteps=1, nstages=3):
        _check(load().r3d_hip_step(self.h, nsteps, nstages))

    def step_noalloc(self, nsteps=1, nstages=3):
        """r3d_hip_step_noalloc: the same steps with no allocation, its scratch
        in the library's static device array."""
        _check(load().r3d_hip_step_noalloc(self.h, nsteps, nstages))

    def snapshot_begin(self):
        """Start an asynchronous snapshot of the current state (include/r3d_hip.h).
        A dump thread still writing the previous snapshot is joined first."""
        self._join_writer()
        _check(load().r3d_hip_snapshot_begin(self.h))

    def snapshot_rank(self, rank, names=None):
        """Wait for the snapshot; one virtual rank's fields and the scalars."""
        names = names or FIELD_NAMES
        out = {n: np.empty((self.nz, self.ny, self.nx)) for n in names}
        ptrs = (ctypes.POINTER(ctypes.c_double) * 16)()
        for n in names:
            ptrs[FIELD_NAMES.index(n)] = _dptr(out[n])
        s = Scalars()
        _check(load().r3d_hip_snapshot_rank(self.h, rank, ptrs, ctypes.byref(s)))
        return out, {"dt": s.dt, "timc": s.timc, "timt": s.timt, "timi": s.timi,
                     "umach": s.umach, "nit": s.nit}

    def sync(self):
        _check(load().r3d_hip_sync(self.h))

    def set_timing(self, on):
        _check(load().r3d_hip_set_timing(self.h, 1 if on else 0))

    def kernel_stats(self, cls):
        n = ctypes.c_long()
        ms = ctypes.c_double()
        _check(load().r3d_hip_kernel_stats(self.h, cls, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def memory_bytes(self):
        b = ctypes.c_ulonglong()
        _check(load().r3d_hip_memory(self.h, ctypes.byref(b)))
        return b.value

    def tiling(self):
        """Stage tiling in use: (tile rows, z chunk, x tiles, y tiles, z chunks)."""
        out = (ctypes.c_int * 5)()
        _check(load().r3d_hip_tiling(self.h, out))
        return tuple(out)


class Ensemble:
    """Many independent runs advanced together on one GPU (include/r3d_ens.h).

    runs: list of Run (host setups); each member gets that run's metrics,
    initial state and zero scalars, exactly as Run.make_sim() does."""

    def __init__(self, runs):
        L = load()
        cfgs = (Cfg * len(runs))(*[r.cfg() for r in runs])
        h = ctypes.c_void_p()
        _check(L.r3d_ens_create(len(runs), cfgs, ctypes.byref(h)))
        self.h = h
        self.n = len(runs)
        self.cfgs = list(cfgs)
        try:
            for m, r in enumerate(runs):
                md = r.metrics()
                keep = {k: np.ascontiguousarray(md[k], dtype=np.float64) for k in METRIC_NAMES}
                st = Metrics(*[_dptr(keep[k]) for k in METRIC_NAMES])
                _check(L.r3d_ens_set_metrics(h, m, ctypes.byref(st)))
                state = r.state(copy=False)
                nf = 8 if self.cfgs[m].lmag else 5
                arrs = [state[nm] for nm in PRIMARIES[:nf]]
                ptrs = (ctypes.POINTER(ctypes.c_double) * 8)(*[_dptr(a) for a in arrs] +
                                                            [None] * (8 - nf))
                _check(L.r3d_ens_upload_global(h, m, ptrs))
                _check(L.r3d_ens_set_scalars(h, m, ctypes.byref(Scalars())))
        except Exception:
            self.close()
            raise

    def close(self):
        if self.h:
            load().r3d_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps=1):
        _check(load().r3d_ens_step(self.h, nsteps))

    def sync(self):
        _check(load().r3d_ens_sync(self.h))

    def set_tiles(self, ty, kc=0, tw=64):
        """ty rows, kc planes (0: default), tw columns (64, or 32 with two
        rows per wave)."""
        _check(load().r3d_ens_set_tile_shape(self.h, ty, kc, tw))

    def tile_shape(self):
        """(rows, planes, columns, tiles per stage) of the current tiling."""
        v = [ctypes.c_int() for _ in range(4)]
        _check(load().r3d_ens_get_tile_shape(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def tiles(self):
        ty, kc, nt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(load().r3d_ens_get_tiles(self.h, ctypes.byref(ty), ctypes.byref(kc),
                                        ctypes.byref(nt)))
        return ty.value, kc.value, nt.value

    def autotune(self, trial_steps=3):
        ms = ctypes.c_double()
        _check(load().r3d_ens_autotune(self.h, trial_steps, ctypes.byref(ms)))
        return ms.value

    def set_graph(self, on):
        _check(load().r3d_ens_set_graph(self.h, 1 if on else 0))

    def set_persistent(self, on):
        """One cooperative launch per step() call for all its steps."""
        _check(load().r3d_ens_set_persistent(self.h, 1 if on else 0))

    def scalars(self, m):
        s = Scalars()
        _check(load().r3d_ens_get_scalars(self.h, m, ctypes.byref(s)))
        return {"dt": s.dt, "timc": s.timc, "timt": s.timt, "timi": s.timi,
                "umach": s.umach, "nit": s.nit}

    def download_global(self, m, names=None):
        names = names or FIELD_NAMES
        c = self.cfgs[m]
        out = {n: np.empty((c.npz + 2, c.npy + 2, c.npx + 2)) for n in names}
        ptrs = (ctypes.POINTER(ctypes.c_double) * 16)()
        for n in names:
            ptrs[FIELD_NAMES.index(n)] = _dptr(out[n])
        _check(load().r3d_ens_download_global(self.h, m, ptrs))
        return out

    def download_rank(self, m, rank, names=None):
        names = names or FIELD_NAMES
        c = self.cfgs[m]
        shape = (c.npz // c.npez + 2, c.npy // c.npey + 2, c.npx + 2)
        out = {n: np.empty(shape) for n in names}
        ptrs = (ctypes.POINTER(ctypes.c_double) * 16)()
        for n in names:
            ptrs[FIELD_NAMES.index(n)] = _dptr(out[n])
        _check(load().r3d_ens_download_rank(self.h, m, rank, ptrs))
        return out

    def memory_bytes(self):
        b = ctypes.c_ulonglong()
        _check(load().r3d_ens_memory(self.h, ctypes.byref(b)))
        return b.value


def cfg_with(cfg, re=None, theta=None, tb=None):
    """A copy of cfg with the Reynolds number, theta and/or the bottom
    temperature tb replaced, the derived coefficients recomputed as the
    reference does (src/params.cpp: ore = 1/Re, repr = Re*Pr).  theta here is
    the heat-flux parameter of the boundary conditions only: the static
    profile (and tb) built from theta come from the host setup."""
    c = Cfg()
    ctypes.pointer(c)[0] = cfg
    if re is not None:
        pr = cfg.repr / cfg.re
        c.re = float(re)
        c.ore = 1.0 / c.re
        c.repr = c.re * pr
    if theta is not None:
        c.theta = float(theta)
    if tb is not None:
        c.tb = float(tb)
    return c


class Adjoint:
    """Discrete adjoint of a hydrodynamic run (include/r3d_adj.h): J = final
    kinetic energy, gradients with respect to the initial primaries, Re and
    theta.  cfg: a Cfg (e.g. cfg_with(run.cfg(), re=...)); metrics: the run's
    global metrics."""

    def __init__(self, cfg, metrics, checkpoint_every=20):
        L = load()
        h = ctypes.c_void_p()
        _check(L.r3d_adj_create(ctypes.byref(cfg), checkpoint_every, ctypes.byref(h)))
        self.h = h
        self.shape = (cfg.npz + 2, cfg.npy + 2, cfg.npx + 2)
        self.names = PRIMARIES[:8 if cfg.lmag else 5]
        try:
            keep = {k: np.ascontiguousarray(metrics[k], dtype=np.float64) for k in METRIC_NAMES}
            st = Metrics(*[_dptr(keep[k]) for k in METRIC_NAMES])
            _check(L.r3d_adj_set_metrics(h, ctypes.byref(st)))
        except Exception:
            self.close()
            raise

    def close(self):
        if self.h:
            load().r3d_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_initial(self, fields):
        arrs = [np.ascontiguousarray(fields[n], dtype=np.float64) for n in self.names]
        for a in arrs:
            assert a.shape == self.shape
        ptrs = (ctypes.POINTER(ctypes.c_double) * len(arrs))(*[_dptr(a) for a in arrs])
        _check(load().r3d_adj_set_initial(self.h, ptrs))

    def set_initial_device(self, ptrs):
        """X0 from device memory: ptrs[name] = address of a dense global array
        (float64, shape self.shape) on the context's GPU, for every name in
        self.names; the caller's writes must be complete."""
        arr = (ctypes.c_void_p * len(self.names))(*[ptrs[n] for n in self.names])
        _check(load().r3d_adj_set_initial_device(self.h, arr))

    def backward_device(self, ptrs):
        """backward() into device memory: ptrs[name] = address of a dense
        global array (float64, shape self.shape) on the context's GPU for each
        gradient wanted (others are skipped).  Returns (dJ/dRe, dJ/dtheta)."""
        arr = (ctypes.c_void_p * len(self.names))(*[ptrs.get(n) for n in self.names])
        dre, dth = ctypes.c_double(), ctypes.c_double()
        _check(load().r3d_adj_backward_device(self.h, arr, ctypes.byref(dre), ctypes.byref(dth)))
        return dre.value, dth.value

    def forward(self, nsteps):
        J = ctypes.c_double()
        _check(load().r3d_adj_forward(self.h, nsteps, ctypes.byref(J)))
        return J.value

    def backward(self):
        out = {n: np.zeros(self.shape) for n in self.names}
        ptrs = (ctypes.POINTER(ctypes.c_double) * len(self.names))(
            *[_dptr(out[n]) for n in self.names])
        dre, dth = ctypes.c_double(), ctypes.c_double()
        _check(load().r3d_adj_backward(self.h, ptrs, ctypes.byref(dre), ctypes.byref(dth)))
        return out, dre.value, dth.value

    def set_fixed_dt(self, dt):
        """dt > 0: later forward runs step with this dt instead of the CFL
        step (J becomes smooth in X0, Re, theta); 0: back to the CFL step."""
        _check(load().r3d_adj_set_fixed_dt(self.h, float(dt)))

    def dJ_dtb(self):
        """dJ/d(tb) of the last backward (the fixed bottom temperature)."""
        v = ctypes.c_double()
        _check(load().r3d_adj_dJ_dtb(self.h, ctypes.byref(v)))
        return v.value

    def timing(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        _check(load().r3d_adj_timing(self.h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    def final_state(self):
        out = {n: np.zeros(self.shape) for n in self.names}
        ptrs = (ctypes.POINTER(ctypes.c_double) * len(self.names))(
            *[_dptr(out[n]) for n in self.names])
        _check(load().r3d_adj_download_final(self.h, ptrs))
        return out


def ensemble_members(n=64, seed=3):
    """Member flags of the seeded ensemble workload: horizontal size in
    {48, 64, 80, 96} (square), vertical size in {96, 128, 160, 192},
    Re uniform in [1, 4], theta uniform in [0.1, 0.3], ran2 seed -(1001 + m),
    virtual 1 x 2 ranks, every other flag at the reference default."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n):
        h = int(rng.choice([48, 64, 80, 96]))
        v = int(rng.choice([96, 128, 160, 192]))
        re = round(float(rng.uniform(1.0, 4.0)), 4)
        th = round(float(rng.uniform(0.1, 0.3)), 4)
        out.append(["--npx", str(h), "--npy", str(h), "--npz", str(v), "--npey", "1",
                    "--npez", "2", "--re", repr(re), "--theta", repr(th),
                    "--seed", str(-(1001 + m))])
    return out


# ----------------------------------------------------------------- host setup --
_host = None


def load_host():
    """Load the host-setup library (raises if it is not built)."""
    global _host
    if _host is not None:
        return _host
    if not os.path.exists(HOST_LIB):
        raise RuntimeError(f"host library not built: {HOST_LIB} (run make -C rast-3dmhd_amd)")
    L = ctypes.CDLL(HOST_LIB)
    vp = ctypes.c_void_p
    L.r3d_host_create.restype = vp
    L.r3d_host_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    L.r3d_host_free.argtypes = [vp]
    L.r3d_host_last_error.restype = ctypes.c_char_p
    L.r3d_host_cfg.argtypes = [vp, ctypes.POINTER(Cfg)]
    L.r3d_host_metrics.argtypes = [vp, ctypes.POINTER(Metrics)]
    L.r3d_host_state.argtypes = [vp, ctypes.POINTER(ctypes.POINTER(ctypes.c_double))]
    L.r3d_host_dims.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.r3d_host_tb.argtypes = [vp]
    L.r3d_host_tb.restype = ctypes.c_double
    L.r3d_host_rank_metrics.argtypes = [vp, ctypes.c_int,
                                        ctypes.POINTER(ctypes.POINTER(ctypes.c_double))]
    L.r3d_host_write_gstate.argtypes = [vp, ctypes.c_int, ctypes.c_char_p,
                                        ctypes.POINTER(ctypes.POINTER(ctypes.c_double)),
                                        ctypes.POINTER(Scalars)]
    L.r3d_host_read_gstate.argtypes = [vp, ctypes.c_int, ctypes.c_char_p,
                                       ctypes.POINTER(ctypes.POINTER(ctypes.c_double)),
                                       ctypes.POINTER(Scalars)]
    L.r3d_host_scalars.argtypes = [vp, ctypes.POINTER(Scalars)]
    _host = L
    return L


class Run:
    """Host setup of one run from reference-style flags (include/r3d_host.h)."""

    def __init__(self, args):
        L = load_host()
        argv = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
        h = L.r3d_host_create(len(args), argv)
        if not h:
            raise HipError(-1, L.r3d_host_last_error().decode())
        self.h = h
        d = (ctypes.c_int * 8)()
        L.r3d_host_dims(h, d)
        (self.npx, self.npy, self.npz, self.npey, self.npez, self.lmag,
         self.ntotal, self.nstep0) = list(d)
        self.shape = (self.npz + 2, self.npy + 2, self.npx + 2)

    def close(self):
        if self.h:
            load_host().r3d_host_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cfg(self):
        c = Cfg()
        if load_host().r3d_host_cfg(self.h, ctypes.byref(c)) != 0:
            raise HipError(-1, load_host().r3d_host_last_error().decode())
        return c

    def metrics(self):
        m = Metrics()
        load_host().r3d_host_metrics(self.h, ctypes.byref(m))
        n = {"x": self.npx + 2, "y": self.npy + 2, "z": self.npz + 2}
        out = {}
        for q, name in enumerate(METRIC_NAMES):
            ln = n["x"] if q < 4 else n["y"] if q < 8 else n["z"]
            out[name] = np.ctypeslib.as_array(getattr(m, name), shape=(ln,)).copy()
        return out

    def state(self, copy=True):
        """Global initial primaries, (npz+2, npy+2, npx+2) each.

        copy=False returns views of the host library's arrays (valid while
        this Run lives), which avoids doubling host memory on large grids."""
        ptrs = (ctypes.POINTER(ctypes.c_double) * 8)()
        load_host().r3d_host_state(self.h, ptrs)
        out = {}
        for q, name in enumerate(PRIMARIES):
            if not ptrs[q]:
                continue  # B fields without LMAG
            a = np.ctypeslib.as_array(ptrs[q], shape=self.shape)
            out[name] = a.copy() if copy else a
        return out

    def tb(self):
        return load_host().r3d_host_tb(self.h)

    def rank_metrics(self, rank):
        nry, nrz = self.npy // self.npey, self.npz // self.npez
        lens = [self.npx + 2] * 4 + [nry + 2] * 4 + [nrz + 2] * 6
        arrs = [np.zeros(n) for n in lens]
        ptrs = (ctypes.POINTER(ctypes.c_double) * 14)(*[_dptr(a) for a in arrs])
        load_host().r3d_host_rank_metrics(self.h, rank, ptrs)
        return dict(zip(METRIC_NAMES, arrs))

    def write_gstate(self, rank, path, fields, scalars):
        ptrs = (ctypes.POINTER(ctypes.c_double) * 16)()
        keep = []
        for q, name in enumerate(FIELD_NAMES):
            if name in fields:
                a = np.ascontiguousarray(fields[name], dtype=np.float64)
                keep.append(a)
                ptrs[q] = _dptr(a)
        s = Scalars(scalars["dt"], scalars["timc"], scalars["timt"], scalars["timi"],
                    scalars["umach"], int(scalars["nit"]))
        if load_host().r3d_host_write_gstate(self.h, rank, path.encode(), ptrs,
                                             ctypes.byref(s)) != 0:
            raise HipError(-1, load_host().r3d_host_last_error().decode())

    def scalars(self):
        """Starting scalars: zero, or those of the dumps on a restart (--nstart)."""
        s = Scalars()
        if load_host().r3d_host_scalars(self.h, ctypes.byref(s)) != 0:
            raise HipError(-1, load_host().r3d_host_last_error().decode())
        return {"dt": s.dt, "timc": s.timc, "timt": s.timt, "timi": s.timi,
                "umach": s.umach, "nit": s.nit}

    def read_gstate(self, rank, path):
        """One virtual rank's gstate dump: (fields by FIELD_NAMES, scalars)."""
        nry, nrz = self.npy // self.npey, self.npz // self.npez
        shape = (nrz + 2, nry + 2, self.npx + 2)
        out = {n: np.zeros(shape) for n in FIELD_NAMES}
        ptrs = (ctypes.POINTER(ctypes.c_double) * 16)(*[_dptr(out[n]) for n in FIELD_NAMES])
        s = Scalars()
        if load_host().r3d_host_read_gstate(self.h, rank, path.encode(), ptrs,
                                            ctypes.byref(s)) != 0:
            raise HipError(-1, load_host().r3d_host_last_error().decode())
        return out, {"dt": s.dt, "timc": s.timc, "timt": s.timt, "timi": s.timi,
                     "umach": s.umach, "nit": s.nit}

    def dump_async(self, sim, directory, tag="step"):
        """Checkpoint without stopping the run: snapshot the device state now
        and write every virtual rank's gstate dump
        (directory/gstate.<tag>.<rank>.<nit>, the reference's names) on a
        background thread while the caller keeps stepping.  Returns the
        started thread.  The next snapshot on this sim (another dump_async or
        snapshot_begin) and sim.close() join it first, so a dump is never
        torn by a later snapshot or by freeing the buffers under it.  A run
        started with --nstart <nit> --dumpdir <directory> continues from it."""
        import threading
        sim.snapshot_begin()
        errors = []

        def write():
            try:
                for r in range(self.npey * self.npez):
                    fields, sc = sim.snapshot_rank(r)
                    self.write_gstate(r, os.path.join(directory, f"gstate.{tag}.{r:03d}.{sc['nit']:04d}"),
                                      fields, sc)
            except Exception as e:  # surfaced by join()
                errors.append(e)

        t = threading.Thread(target=write, daemon=True)
        t.errors = errors
        sim._writer = t
        t.start()
        return t

    def make_sim(self, variant=0, tile_rows=0, zchunk=0, dt_screen=0.0):
        """Create the device context and upload this run's metrics, state and
        starting scalars.  tile_rows / zchunk override the stage-kernel tiling
        and dt_screen the margin of the last stage's time-step screen (0 = the
        library's choice; include/r3d_hip.h)."""
        cfg = self.cfg()
        cfg.variant = variant
        cfg.tile_rows = tile_rows
        cfg.zchunk = zchunk
        cfg.dt_screen = dt_screen
        sim = HipSim(cfg)
        sim.set_metrics(self.metrics())
        sim.upload_global(self.state(copy=False))
        sim.set_scalars(**self.scalars())
        return sim
