"""ctypes binding of ``libtsx.so`` (declared in ``include/tsx.h``).

The projector boundary of tomosipo is ``astra.create_projector`` plus
``astra.experimental.direct_FPBP3D`` (tomosipo ``astra.py:80-153``); this
module is the Python side of its replacement. Device pointers and the HIP
stream come from torch tensors (``tensor.data_ptr()``,
``torch.cuda.current_stream().cuda_stream``) so no data is copied.

There is no CPU fallback: if the library is missing, or a projection is
requested without a HIP device, the call raises.
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "csrc", "libtsx.so")

TSX_PARALLEL3D_VEC = 0
TSX_CONE_VEC = 1
TSX_MODE_ADD = 0
TSX_MODE_SET = 1
TSX_OP_FP = 0
TSX_OP_BP = 1
TSX_EP_SET = 0
TSX_EP_ADD = 1
TSX_EP_RESIDUAL = 2
TSX_EP_SCALED_ADD = 3

_STATUS = {
    -1: "TSX_ERR_INVALID_ARGUMENT",
    -2: "TSX_ERR_UNSUPPORTED",
    -3: "TSX_ERR_HIP",
    -4: "TSX_ERR_NO_DEVICE",
    -5: "TSX_ERR_OUT_OF_MEMORY",
}

EXPORTED = (
    "tsx_plan_create", "tsx_plan_destroy", "tsx_plan_info_get", "tsx_plan_axes",
    "tsx_plan_autotune", "tsx_fdk_workspace_size", "tsx_fdk",
    "tsx_workspace_size",
    "tsx_direct_fpbp3d", "tsx_forward", "tsx_backward", "tsx_forward_batched",
    "tsx_backward_batched", "tsx_multi_workspace_size", "tsx_multi_forward",
    "tsx_multi_backward", "tsx_multi_forward_ep", "tsx_multi_backward_ep",
    "tsx_forward_residual", "tsx_backward_scaled_add",
    "tsx_forward_batched_f64", "tsx_backward_batched_f64",
    "tsx_direct_fpbp3d_noalloc",
    "tsx_last_error", "tsx_version",
)

# Bytes of the static device scratch buffer of tsx_direct_fpbp3d_noalloc
# (TSX_NOALLOC_SCRATCH_BYTES).
NOALLOC_SCRATCH_BYTES = 256 << 20


class TsxError(RuntimeError):
    def __init__(self, code, message):
        self.code = code
        super().__init__(f"{_STATUS.get(code, code)}: {message}")


class TsxGeometry(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("det_row_count", ctypes.c_int32),
        ("det_col_count", ctypes.c_int32),
        ("n_angles", ctypes.c_int32),
        ("vectors", ctypes.c_void_p),
        ("grid_col_count", ctypes.c_int32),
        ("grid_row_count", ctypes.c_int32),
        ("grid_slice_count", ctypes.c_int32),
        ("window_min", ctypes.c_double * 3),
        ("window_max", ctypes.c_double * 3),
        ("det_supersampling", ctypes.c_int32),
        ("vox_supersampling", ctypes.c_int32),
    ]


class TsxEpilogue(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("y", ctypes.c_void_p), ("w", ctypes.c_void_p)]


class TsxPlanInfo(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("n_angles", ctypes.c_int32),
        ("det_rows", ctypes.c_int32),
        ("det_cols", ctypes.c_int32),
        ("vol_shape", ctypes.c_int64 * 3),
        ("plan_bytes", ctypes.c_int64),
        ("axis_count", ctypes.c_int32 * 3),
        ("fp_kernel", ctypes.c_int32),
        ("bp_kernel", ctypes.c_int32),
        ("fp_variant", ctypes.c_int32),
        ("bp_brick_z", ctypes.c_int32),
        ("tuned", ctypes.c_int32),
        ("fp_slabs", ctypes.c_int32),
        ("bp_launch_angles", ctypes.c_int32),
        ("fp_tile_rows", ctypes.c_int32),
        ("bp_axis_swap", ctypes.c_int32),
        ("fp_axis_swap", ctypes.c_int32),
        ("fp_shear_rows", ctypes.c_int32),
        ("fp_column_waves", ctypes.c_int32),
        ("fp_refined_detector", ctypes.c_int32),
        ("bp_refined_detector", ctypes.c_int32),
    ]


_lib = None
_lib_lock = threading.Lock()


def load():
    """Load ``libtsx.so``; raises ``ImportError`` if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m tomosipo_amd._build` "
                "(hipcc, gfx950). tomosipo_amd has no CPU fallback.")
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64, i32 = ctypes.c_int64, ctypes.c_int
        lib.tsx_plan_create.argtypes = [ctypes.POINTER(TsxGeometry), i32, ctypes.POINTER(P)]
        lib.tsx_plan_destroy.argtypes = [P]
        lib.tsx_plan_info_get.argtypes = [P, ctypes.POINTER(TsxPlanInfo)]
        lib.tsx_plan_axes.argtypes = [P, P]
        lib.tsx_plan_autotune.argtypes = [P, P]
        lib.tsx_fdk_workspace_size.argtypes = [P, ctypes.POINTER(i64)]
        lib.tsx_fdk.argtypes = [P, P, P, i32, P, i64, P]
        lib.tsx_direct_fpbp3d.argtypes = [P, P, P, i32, i32, P]
        lib.tsx_forward.argtypes = [P, P, P, i32, P]
        lib.tsx_backward.argtypes = [P, P, P, i32, P]
        lib.tsx_workspace_size.argtypes = [P, i32, i64, ctypes.POINTER(i64)]
        lib.tsx_forward_batched.argtypes = [P, P, i64, P, i64, i64, i32, P, i64, P]
        lib.tsx_backward_batched.argtypes = [P, P, i64, P, i64, i64, i32, P, i64, P]
        lib.tsx_multi_workspace_size.argtypes = [P, i32, i32, ctypes.POINTER(i64)]
        lib.tsx_multi_forward.argtypes = [P, P, P, i32, i32, P, i64, P]
        lib.tsx_multi_backward.argtypes = [P, P, P, i32, i32, P, i64, P]
        lib.tsx_multi_forward_ep.argtypes = [P, P, P, P, i32, P, i64, P]
        lib.tsx_multi_backward_ep.argtypes = [P, P, P, P, i32, P, i64, P]
        lib.tsx_forward_residual.argtypes = [P, P, i64, P, P, P, i64, i64, P, i64, P]
        lib.tsx_backward_scaled_add.argtypes = [P, P, i64, P, i64, P, i64, P, i64, P]
        lib.tsx_forward_batched_f64.argtypes = [P, P, i64, P, i64, i64, i32, P]
        lib.tsx_backward_batched_f64.argtypes = [P, P, i64, P, i64, i64, i32, P]
        lib.tsx_direct_fpbp3d_noalloc.argtypes = [P, P, P, i32, i32, P]
        for name in EXPORTED:
            if name not in ("tsx_last_error", "tsx_version"):
                getattr(lib, name).restype = ctypes.c_int
        lib.tsx_last_error.restype = ctypes.c_char_p
        lib.tsx_version.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        raise TsxError(rc, load().tsx_last_error().decode())


def make_geometry(vol_geom, proj_geom, det_supersampling=1, vox_supersampling=1):
    """``TsxGeometry`` from the ASTRA dictionaries; returns (struct, keep-alive array)."""
    kinds = {"parallel3d_vec": TSX_PARALLEL3D_VEC, "cone_vec": TSX_CONE_VEC}
    if proj_geom["type"] not in kinds:
        raise ValueError(f"projector needs a vector geometry; got '{proj_geom['type']}'")
    vectors = np.ascontiguousarray(proj_geom["Vectors"], dtype=np.float64)
    if vectors.ndim != 2 or vectors.shape[1] != 12:
        raise ValueError(f"Vectors must have shape (n, 12); got {vectors.shape}")
    o = vol_geom["option"]
    g = TsxGeometry()
    g.type = kinds[proj_geom["type"]]
    g.det_row_count = int(proj_geom["DetectorRowCount"])
    g.det_col_count = int(proj_geom["DetectorColCount"])
    g.n_angles = vectors.shape[0]
    g.vectors = vectors.ctypes.data
    g.grid_col_count = int(vol_geom["GridColCount"])
    g.grid_row_count = int(vol_geom["GridRowCount"])
    g.grid_slice_count = int(vol_geom["GridSliceCount"])
    g.window_min = (ctypes.c_double * 3)(o["WindowMinX"], o["WindowMinY"], o["WindowMinZ"])
    g.window_max = (ctypes.c_double * 3)(o["WindowMaxX"], o["WindowMaxY"], o["WindowMaxZ"])
    g.det_supersampling = int(det_supersampling)
    g.vox_supersampling = int(vox_supersampling)
    return g, vectors


class Plan:
    """Owner of one ``tsx_plan`` (immutable geometry tables on one device)."""

    def __init__(self, vol_geom, proj_geom, device=-1, det_supersampling=1, vox_supersampling=1):
        lib = load()
        g, self._vectors = make_geometry(vol_geom, proj_geom, det_supersampling, vox_supersampling)
        h = ctypes.c_void_p()
        rc = lib.tsx_plan_create(ctypes.byref(g), int(device), ctypes.byref(h))
        if rc != 0:
            msg = lib.tsx_last_error().decode()
            if rc in (-1, -2):
                raise ValueError(msg)
            raise TsxError(rc, msg)
        self._h = h
        self.device = int(device)
        self._lib = lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.tsx_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def info(self):
        inf = TsxPlanInfo()
        _check(self._lib.tsx_plan_info_get(self._h, ctypes.byref(inf)))
        return {
            "device": inf.device, "n_angles": inf.n_angles, "det_rows": inf.det_rows,
            "det_cols": inf.det_cols, "vol_shape": tuple(inf.vol_shape),
            "plan_bytes": inf.plan_bytes, "axis_count": tuple(inf.axis_count),
            "fp_kernel": ("general", "padded")[inf.fp_kernel],
            "bp_kernel": ("general", "separable")[inf.bp_kernel],
            "fp_variant": inf.fp_variant, "bp_brick_z": inf.bp_brick_z,
            "tuned": bool(inf.tuned),
            "fp_slabs": inf.fp_slabs, "bp_launch_angles": inf.bp_launch_angles,
            "fp_tile_rows": inf.fp_tile_rows, "bp_axis_swap": inf.bp_axis_swap, "fp_axis_swap": inf.fp_axis_swap,
            "fp_shear_rows": inf.fp_shear_rows,
            "fp_column_waves": bool(inf.fp_column_waves),
            "fp_refined_detector": inf.fp_refined_detector,
            "bp_refined_detector": inf.bp_refined_detector,
        }

    def fdk_workspace_size(self):
        out = ctypes.c_int64()
        _check(self._lib.tsx_fdk_workspace_size(self._h, ctypes.byref(out)))
        return int(out.value)

    def fdk(self, vol_ptr, proj_ptr, additive, stream, workspace=None, workspace_bytes=0):
        """FDK reconstruction of the projections into the volume (``tsx_fdk``)."""
        _check(self._lib.tsx_fdk(self._h, vol_ptr, proj_ptr, int(bool(additive)), workspace,
                                 int(workspace_bytes), stream))

    def autotune(self, stream):
        """Time the kernel variants on this plan's shapes and keep the fastest
        (``tsx_plan_autotune``; synchronises ``stream``)."""
        _check(self._lib.tsx_plan_autotune(self._h, stream))

    def axes(self):
        n = self.info()["n_angles"]
        out = np.zeros(n, dtype=np.int32)
        _check(self._lib.tsx_plan_axes(self._h, out.ctypes.data))
        return out

    def workspace_size(self, op, batch=1):
        out = ctypes.c_int64()
        _check(self._lib.tsx_workspace_size(self._h, int(op), int(batch), ctypes.byref(out)))
        return int(out.value)

    def forward(self, vol_ptr, proj_ptr, additive, stream, batch=1, vol_stride=0, proj_stride=0,
                workspace=None, workspace_bytes=0):
        _check(self._lib.tsx_forward_batched(self._h, vol_ptr, vol_stride, proj_ptr, proj_stride,
                                             batch, int(bool(additive)), workspace,
                                             int(workspace_bytes), stream))

    def backward(self, vol_ptr, proj_ptr, additive, stream, batch=1, vol_stride=0, proj_stride=0,
                 workspace=None, workspace_bytes=0):
        _check(self._lib.tsx_backward_batched(self._h, vol_ptr, vol_stride, proj_ptr, proj_stride,
                                              batch, int(bool(additive)), workspace,
                                              int(workspace_bytes), stream))

    def forward_residual(self, vol_ptr, y_ptr, w_ptr, proj_ptr, stream, batch=1, vol_stride=0,
                         proj_stride=0, workspace=None, workspace_bytes=0):
        _check(self._lib.tsx_forward_residual(self._h, vol_ptr, vol_stride, y_ptr, w_ptr, proj_ptr,
                                              proj_stride, batch, workspace, int(workspace_bytes),
                                              stream))

    def backward_scaled_add(self, vol_ptr, proj_ptr, w_ptr, stream, batch=1, vol_stride=0,
                            proj_stride=0, workspace=None, workspace_bytes=0):
        _check(self._lib.tsx_backward_scaled_add(self._h, vol_ptr, vol_stride, proj_ptr,
                                                 proj_stride, w_ptr, batch, workspace,
                                                 int(workspace_bytes), stream))

    def forward_f64(self, vol_ptr, proj_ptr, additive, stream, batch=1, vol_stride=0, proj_stride=0):
        _check(self._lib.tsx_forward_batched_f64(self._h, vol_ptr, vol_stride, proj_ptr,
                                                 proj_stride, batch, int(bool(additive)), stream))

    def backward_f64(self, vol_ptr, proj_ptr, additive, stream, batch=1, vol_stride=0, proj_stride=0):
        _check(self._lib.tsx_backward_batched_f64(self._h, vol_ptr, vol_stride, proj_ptr,
                                                  proj_stride, batch, int(bool(additive)), stream))


def version():
    return load().tsx_version().decode()


def _ptr_array(values):
    arr = (ctypes.c_void_p * max(len(values), 1))()
    for k, v in enumerate(values):
        arr[k] = v
    return arr


def multi_workspace_size(plans, op):
    lib = load()
    out = ctypes.c_int64()
    hs = _ptr_array([p.handle.value for p in plans])
    _check(lib.tsx_multi_workspace_size(hs, len(plans), int(op), ctypes.byref(out)))
    return int(out.value)


def multi_project(plans, vol_ptrs, proj_ptrs, forward, additive, stream, workspace=None,
                  workspace_bytes=0, epilogues=None):
    """One fused launch over many instances (tsx_multi_forward / tsx_multi_backward).

    ``epilogues``: None (all SET, or all ADD with ``additive``), or one
    ``(mode, y_ptr, w_ptr)`` per instance (tsx_multi_forward_ep / _backward_ep).
    """
    lib = load()
    n = len(plans)
    if len(vol_ptrs) != n or len(proj_ptrs) != n:
        raise ValueError("one volume and one projection pointer per plan are required")
    hs = _ptr_array([p.handle.value for p in plans])
    vs = _ptr_array(vol_ptrs)
    ps = _ptr_array(proj_ptrs)
    if epilogues is None:
        fn = lib.tsx_multi_forward if forward else lib.tsx_multi_backward
        _check(fn(hs, vs, ps, n, int(bool(additive)), workspace, int(workspace_bytes), stream))
        return
    if len(epilogues) != n:
        raise ValueError("one epilogue per plan is required")
    eps = (TsxEpilogue * max(n, 1))()
    for k, (mode, y, w) in enumerate(epilogues):
        eps[k].mode, eps[k].y, eps[k].w = int(mode), y, w
    fn = lib.tsx_multi_forward_ep if forward else lib.tsx_multi_backward_ep
    _check(fn(hs, vs, ps, eps, n, workspace, int(workspace_bytes), stream))
