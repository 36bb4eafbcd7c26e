"""ctypes binding of ``liboracle.so`` (test infrastructure only)."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(lib_path):
            build()
        lib = ctypes.CDLL(lib_path)
        for name in ("oracle_fp", "oracle_bp"):
            f = getattr(lib, name)
            f.argtypes = [ctypes.POINTER(_CGeom), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
            f.restype = None
        lib.oracle_axes.argtypes = [ctypes.POINTER(_CGeom), ctypes.c_void_p]
        lib.oracle_num_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


class _CGeom(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("n_angles", ctypes.c_int32),
        ("det_rows", ctypes.c_int32),
        ("det_cols", ctypes.c_int32),
        ("vectors", ctypes.c_void_p),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("vmin", ctypes.c_double * 3),
        ("vmax", ctypes.c_double * 3),
        ("det_supersampling", ctypes.c_int32),
    ]


class OracleGeometry:
    """ASTRA-form geometry: volume dict + projection dict (vector types)."""

    def __init__(self, vol_geom, proj_geom, det_supersampling=1):
        kinds = {"parallel3d_vec": 0, "cone_vec": 1}
        if proj_geom["type"] not in kinds:
            raise ValueError(f"oracle needs a vector geometry, got {proj_geom['type']}")
        self.kind = kinds[proj_geom["type"]]
        self.vectors = np.ascontiguousarray(proj_geom["Vectors"], dtype=np.float64)
        self.rows = int(proj_geom["DetectorRowCount"])
        self.cols = int(proj_geom["DetectorColCount"])
        self.n_angles = self.vectors.shape[0]
        self.nx, self.ny, self.nz = (int(vol_geom["GridColCount"]), int(vol_geom["GridRowCount"]),
                                     int(vol_geom["GridSliceCount"]))
        o = vol_geom["option"]
        self.vmin = (o["WindowMinX"], o["WindowMinY"], o["WindowMinZ"])
        self.vmax = (o["WindowMaxX"], o["WindowMaxY"], o["WindowMaxZ"])
        self.det_supersampling = int(det_supersampling)

    @property
    def vol_shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def sino_shape(self):
        return (self.rows, self.n_angles, self.cols)

    def _c(self):
        g = _CGeom()
        g.kind = self.kind
        g.n_angles, g.det_rows, g.det_cols = self.n_angles, self.rows, self.cols
        g.vectors = self.vectors.ctypes.data
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.vmin = (ctypes.c_double * 3)(*self.vmin)
        g.vmax = (ctypes.c_double * 3)(*self.vmax)
        g.det_supersampling = self.det_supersampling
        return g


def oracle_geometry(vg, pg, det_supersampling=1):
    """Oracle geometry from tomosipo_amd-style geometry objects (vg axis-aligned)."""
    return OracleGeometry(vg.to_astra(), pg.to_vec().to_astra(), det_supersampling)


def fp(g, vol, out=None):
    lib = _load()
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    assert vol.shape == g.vol_shape, (vol.shape, g.vol_shape)
    additive = out is not None
    if out is None:
        out = np.empty(g.sino_shape, dtype=np.float64)
    assert out.dtype == np.float64 and out.flags.c_contiguous and out.shape == g.sino_shape
    cg = g._c()
    lib.oracle_fp(ctypes.byref(cg), vol.ctypes.data, out.ctypes.data, int(additive))
    return out


def bp(g, sino, out=None):
    lib = _load()
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    assert sino.shape == g.sino_shape, (sino.shape, g.sino_shape)
    additive = out is not None
    if out is None:
        out = np.empty(g.vol_shape, dtype=np.float64)
    assert out.dtype == np.float64 and out.flags.c_contiguous and out.shape == g.vol_shape
    cg = g._c()
    lib.oracle_bp(ctypes.byref(cg), out.ctypes.data, sino.ctypes.data, int(additive))
    return out


def axes(g):
    lib = _load()
    out = np.zeros(g.n_angles, dtype=np.int32)
    cg = g._c()
    lib.oracle_axes(ctypes.byref(cg), out.ctypes.data)
    return out


def num_threads():
    return int(_load().oracle_num_threads())
