"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.  See r3d_oracle.h for what the oracle
restates and how each part is pinned.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libr3d_oracle.so")

FIELDS = ["ru", "rv", "rw", "ro", "tt", "uu", "vv", "ww", "fu", "fv", "fw", "fr", "ft",
          "zru", "zrv", "zrw", "zro", "ztt", "ww1", "ww2", "ww3",
          "bx", "by", "bz", "zbx", "zby", "zbz"]
METRICS = ["exx", "dxxdx", "d2xxdx2", "ddx", "wyy", "dyydy", "d2yydy2", "ddy",
           "zee", "dzzdz", "d2zzdz2", "ddz", "rkapa", "dkapa"]
DERIVED = ["c13", "c23", "c43", "gam1", "gam2", "gam3", "zeta1", "zeta2",
           "hx", "h2x", "hy", "h2y", "hz", "h2z", "repr", "ore", "cv", "ocv",
           "rkapst", "orm", "obeta", "omx", "omz", "pr", "gamma", "ampt", "sf"]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.o_create.restype = ctypes.c_void_p
        L.o_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
        L.o_last_error.restype = ctypes.c_char_p
        L.o_free.argtypes = [ctypes.c_void_p]
        for fn in ("o_init",):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
            getattr(L, fn).restype = ctypes.c_int
        L.o_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.o_step.restype = ctypes.c_int
        L.o_lrem_profile.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_double)] * 5
        L.o_lrem_profile.restype = ctypes.c_int
        L.o_lrem_background.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.o_lrem_background.restype = ctypes.c_int
        L.o_step_dt.argtypes = [ctypes.c_void_p]
        L.o_step_dt.restype = ctypes.c_int
        L.o_fluxes.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.o_fluxes.restype = ctypes.c_int
        L.o_bcon.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.o_bcon.restype = ctypes.c_int
        L.o_density_prep.argtypes = [ctypes.c_void_p]
        L.o_density_prep.restype = ctypes.c_int
        L.o_hmean.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.o_hmean.restype = ctypes.c_int
        L.o_rom.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.o_rom.restype = ctypes.POINTER(ctypes.c_double)
        L.o_relax_prep.argtypes = [ctypes.c_void_p]
        L.o_relax_prep.restype = ctypes.c_int
        L.o_get_relax.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.o_get_relax.restype = None
        L.o_correct.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.o_correct.restype = ctypes.POINTER(ctypes.c_double)
        L.o_communicate.argtypes = [ctypes.c_void_p]
        L.o_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.o_field.restype = ctypes.POINTER(ctypes.c_double)
        L.o_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.o_metric.restype = ctypes.POINTER(ctypes.c_double)
        L.o_metric.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.o_get_scalars.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.o_set_scalars.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.o_get_derived.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.o_write_gstate.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
        L.o_write_gstate.restype = ctypes.c_int
        L.o_ran2_sequence.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.o_static_profile.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
        L.o_static_profile.restype = ctypes.c_int
        L.o_mkgrid.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def lrem_background(z):
    """T, rho of the LREM background at depth z (oracle restatement)."""
    out = (ctypes.c_double * 2)()
    if lib().o_lrem_background(float(z), out) != 0:
        raise OracleError(lib().o_last_error().decode())
    return out[0], out[1]


class OracleError(RuntimeError):
    pass


class Sim:
    """All virtual ranks of one decomposition, reference-style arrays per rank."""

    def __init__(self, args, init=True):
        L = lib()
        argv = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
        h = L.o_create(len(args), argv)
        if not h:
            raise OracleError(L.o_last_error().decode())
        self.h = h
        d = (ctypes.c_int * 10)()
        L.o_dims(h, d)
        (self.nx, self.ny, self.nz, self.npey, self.npez, self.nranks,
         self.npx, self.npy, self.npz, self.lmag) = list(d)
        if init:
            self._check(L.o_init(h))

    def _check(self, rc):
        if rc != 0:
            raise OracleError(lib().o_last_error().decode())

    def close(self):
        if self.h:
            lib().o_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps=1, nstages=3):
        for _ in range(nsteps):
            self._check(lib().o_step(self.h, nstages))

    def lrem_profile(self):
        """LREM STATIC profile: dict of T, R, RK, DRK (npz each) and TU, DZTU, DZTB."""
        n = self.npz
        a = [np.zeros(n) for _ in range(4)]
        b = np.zeros(3)
        self._check(lib().o_lrem_profile(self.h, *[_dptr(x) for x in a], _dptr(b)))
        return dict(zip(["T", "R", "RK", "DRK"], a), TU=b[0], DZTU=b[1], DZTB=b[2])

    def step_dt(self):
        """STEP's time-step limits alone: uu, vv (LMAG), ww1..ww3, umach, dt."""
        self._check(lib().o_step_dt(self.h))

    def fluxes(self, rank):
        self._check(lib().o_fluxes(self.h, rank))

    def bcon(self, rank):
        self._check(lib().o_bcon(self.h, rank))

    def density_prep(self):
        self._check(lib().o_density_prep(self.h))

    def hmean(self, name):
        """Horizontal mean of a field per local plane, into rom()."""
        self._check(lib().o_hmean(self.h, FIELDS.index(name)))

    def rom(self, rank):
        return np.ctypeslib.as_array(lib().o_rom(self.h, rank), shape=(self.nz,))

    def relax_prep(self):
        self._check(lib().o_relax_prep(self.h))

    def relax_state(self):
        v = (ctypes.c_double * 2)()
        lib().o_get_relax(self.h, v)
        return {"tfac": v[0], "ttm_max": v[1]}

    def correct(self, rank):
        return np.ctypeslib.as_array(lib().o_correct(self.h, rank), shape=(self.nz,))

    def communicate(self):
        lib().o_communicate(self.h)

    def field(self, rank, name):
        """Live view (nz, ny, nx) of one rank's field."""
        p = lib().o_field(self.h, rank, FIELDS.index(name))
        return np.ctypeslib.as_array(p, shape=(self.nz, self.ny, self.nx))

    def metric(self, rank, name):
        m = METRICS.index(name)
        n = self.nx if m < 4 else self.ny if m < 8 else self.nz
        return np.ctypeslib.as_array(lib().o_metric(self.h, rank, m), shape=(n,))

    def scalars(self):
        v = (ctypes.c_double * 7)()
        lib().o_get_scalars(self.h, v)
        return dict(zip(["dt", "timc", "timt", "timi", "umach", "tb", "nit"], list(v)))

    def set_scalars(self, sc):
        v = (ctypes.c_double * 7)(*[sc[k] for k in ["dt", "timc", "timt", "timi", "umach", "tb", "nit"]])
        lib().o_set_scalars(self.h, v)

    def derived(self):
        v = (ctypes.c_double * 27)()
        lib().o_get_derived(self.h, v)
        return dict(zip(DERIVED, list(v)))

    def write_gstate(self, rank, path):
        self._check(lib().o_write_gstate(self.h, rank, path.encode()))

    def rank_coords(self, rank):
        return rank % self.npey, rank // self.npey

    def global_interior(self, name):
        """Reassemble the global interior (npz, npy, npx) from all ranks."""
        nry, nrz = self.npy // self.npey, self.npz // self.npez
        out = np.empty((self.npz, self.npy, self.npx))
        for r in range(self.nranks):
            py, pz = self.rank_coords(r)
            out[pz * nrz:(pz + 1) * nrz, py * nry:(py + 1) * nry, :] = \
                self.field(r, name)[1:-1, 1:-1, 1:-1]
        return out

    def static_profile(self):
        t = np.zeros(self.npz)
        r = np.zeros(self.npz)
        self._check(lib().o_static_profile(self.h, _dptr(t), _dptr(r)))
        return t, r

    def mkgrid(self, scode, icoord):
        out = (ctypes.c_double * 3)()
        lib().o_mkgrid(self.h, scode, icoord, out)
        return tuple(out)


def ran2_sequence(n, seed=-62659):
    out = np.zeros(n)
    lib().o_ran2_sequence(seed, n, _dptr(out))
    return out
