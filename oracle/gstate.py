"""Reader for gstate dumps (TEST INFRASTRUCTURE ONLY).

The format is the reference's dump layout (src/main.cpp:44-93):
  int32 magic=1129928788, int32 nfields=27, int32 n1d=14,
  int32[13] nx ny nz mype mypey mypez nit lmag ixcon iycon izcon itcon ibcon,
  f64[27] scalars, 27 x (int64 count, f64[count]) 3-D fields (x fastest),
  14 x (int64 count, f64[count]) 1-D metrics, int64 -1.
"""
import numpy as np

MAGIC = 1129928788
FIELDS = ["ru", "rv", "rw", "ro", "tt", "uu", "vv", "ww", "fu", "fv", "fw", "fr", "ft",
          "zru", "zrv", "zrw", "zro", "ztt", "ww1", "ww2", "ww3",
          "bx", "by", "bz", "zbx", "zby", "zbz"]
METRICS = ["exx", "dxxdx", "d2xxdx2", "ddx", "wyy", "dyydy", "d2yydy2", "ddy",
           "zee", "dzzdz", "d2zzdz2", "ddz", "rkapa", "dkapa"]
HEADER = ["nx", "ny", "nz", "mype", "mypey", "mypez", "nit", "lmag",
          "ixcon", "iycon", "izcon", "itcon", "ibcon"]
SCALARS = ["dt", "timc", "timt", "timi", "umach", "re", "repr", "cv", "ocv", "ore",
           "theta", "grav", "rkapst", "orm", "rm", "obeta", "omx", "omz", "sf", "ampt",
           "dd", "hx", "h2x", "hy", "h2y", "hz", "h2z"]


def read(path):
    buf = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    magic, nf, n1 = take("<i4", 3)
    if magic != MAGIC or nf != 27 or n1 != 14:
        raise ValueError(f"{path}: not a gstate dump")
    hdr = dict(zip(HEADER, (int(v) for v in take("<i4", 13))))
    sc = dict(zip(SCALARS, (float(v) for v in take("<f8", 27))))
    nx, ny, nz = hdr["nx"], hdr["ny"], hdr["nz"]
    f3 = {}
    for name in FIELDS:
        (cnt,) = take("<i8", 1)
        f3[name] = take("<f8", int(cnt)).reshape(nz, ny, nx)
    m1 = {}
    for name in METRICS:
        (cnt,) = take("<i8", 1)
        m1[name] = take("<f8", int(cnt))
    (end,) = take("<i8", 1)
    if end != -1:
        raise ValueError(f"{path}: missing end mark")
    return {"header": hdr, "scalars": sc, "fields": f3, "metrics": m1}
