"""ctypes binding of librtm_hip.so (include/rtm_hip.h).

Loading fails loudly if the library has not been built: there is no CPU
fallback for any entry point.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "librtm_hip.so")

RTM_STRIP_F32, RTM_STRIP_F16, RTM_STRIP_BF16 = 0, 1, 2
FIELD_U, FIELD_U_PREV, FIELD_Q, FIELD_I1, FIELD_I2, FIELD_SU, FIELD_SQ = range(7)
DATA_GATHERS, DATA_RECORDS = 0, 1
(K_FWD_INTERIOR, K_FWD_FRAME, K_SRC_INTERIOR, K_STRIP_RESTORE, K_RCV_INTERIOR, K_RCV_FRAME,
 K_BORN_INTERIOR, K_BORN_FRAME, K_ADJ_INTERIOR, K_ADJ_FRAME, K_SRCI_INTERIOR,
 K_RCVI_INTERIOR) = range(12)
KERNEL_NAMES = {
    K_FWD_INTERIOR: "k_interior<EPI_FWD>", K_FWD_FRAME: "k_frame<FR_FWD>",
    K_SRC_INTERIOR: "k_interior<EPI_SRC>", K_STRIP_RESTORE: "k_strip_restore",
    K_RCV_INTERIOR: "k_interior<EPI_RCV>", K_RCV_FRAME: "k_frame<FR_RCV>",
    K_BORN_INTERIOR: "k_interior<EPI_BORN>", K_BORN_FRAME: "k_frame<FR_BORN>",
    K_ADJ_INTERIOR: "k_interior<EPI_ADJ>", K_ADJ_FRAME: "k_frame<FR_ADJ>",
    K_SRCI_INTERIOR: "k_interior<EPI_SRCI>", K_RCVI_INTERIOR: "k_interior<EPI_RCVI>",
}


class RtmParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nz", "nx", "n2", "nfdmax", "nt", "nt2", "ilste",
                                        "icompen", "s_l", "s_r", "s_z", "ds", "n")] + \
               [(n, C.c_float) for n in ("tao2", "h2", "taoh", "taoh2", "hzx2_1")] + \
               [(n, C.c_int) for n in ("max_shots", "strip_format", "record", "zseg",
                                       "snap_every", "stack_nx", "check_every",
                                       "fuse_frame", "frame_split")]


class RtmShotGeom(C.Structure):
    """One shot of a ragged batch (include/rtm_hip.h: rtm_shot_geom)."""
    _fields_ = [(n, C.c_int) for n in ("nx", "r_u", "r_x", "s_l", "s_r", "n", "x_img")]


class RtmError(RuntimeError):
    pass


_lib = None
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RtmError(f"{LIB_PATH} is missing: run `python -m rtm_gpu_amd.build` first")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.rtm_create.argtypes = [C.POINTER(RtmParams), C.POINTER(vp)]
        L.rtm_destroy.argtypes = [vp]
        L.rtm_last_error.argtypes = [vp]
        L.rtm_last_error.restype = C.c_char_p
        L.rtm_device_bytes.argtypes = [vp]
        L.rtm_device_bytes.restype = C.c_size_t
        L.rtm_set_model.argtypes = [vp, _fp, _fp, _fp, _fp, C.c_int, vp, C.c_int, C.c_float,
                                    C.c_float]
        L.rtm_set_wavelet.argtypes = [vp, _fp]
        L.rtm_set_shots.argtypes = [vp, C.c_int, _ip, _ip]
        L.rtm_set_shot_geometry.argtypes = [vp, C.c_int, C.POINTER(RtmShotGeom)]
        L.rtm_set_shot_model.argtypes = [vp, C.c_int, _fp, _fp]
        L.rtm_set_gathers.argtypes = [vp, _fp]
        L.rtm_gathers_from_records.argtypes = [vp]
        L.rtm_forward_begin.argtypes = [vp, C.c_int]
        L.rtm_forward_advance.argtypes = [vp, C.c_int]
        L.rtm_backward_begin.argtypes = [vp]
        L.rtm_backward_advance.argtypes = [vp, C.c_int]
        L.rtm_migrate.argtypes = [vp]
        L.rtm_set_graph.argtypes = [vp, C.c_int]
        L.rtm_sync.argtypes = [vp]
        L.rtm_read_field.argtypes = [vp, C.c_int, C.c_int, _fp]
        L.rtm_read_records.argtypes = [vp, _fp]
        L.rtm_read_strip.argtypes = [vp, C.c_int, C.c_int, _fp]
        L.rtm_image_post.argtypes = [vp, C.c_float, C.c_float, vp, vp]
        L.rtm_stack_finish.argtypes = [vp, C.c_int, C.c_int, _fp, vp]
        L.rtm_set_timing.argtypes = [vp, C.c_int]
        L.rtm_set_reflectivity.argtypes = [vp, _fp]
        L.rtm_born_begin.argtypes = [vp]
        L.rtm_born_advance.argtypes = [vp, C.c_int]
        L.rtm_adjoint_begin.argtypes = [vp]
        L.rtm_adjoint_advance.argtypes = [vp, C.c_int]
        L.rtm_adjoint_gradient.argtypes = [vp, _fp]
        L.rtm_data_dot.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.rtm_data_axpy.argtypes = [vp, C.c_float, C.c_int, C.c_int]
        L.rtm_data_dot_noalloc.argtypes = [vp, vp, C.c_longlong, C.c_longlong, C.c_int, vp, vp]
        L.rtm_data_dot_noalloc.restype = C.c_int
        L.rtm_stream_copy.argtypes = [vp, vp, C.c_longlong, vp]
        L.rtm_stream_copy.restype = C.c_int
        L.rtm_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
        L.rtm_set_zseg.argtypes = [vp, C.POINTER(C.c_int)]
        L.rtm_set_zseg.restype = C.c_int
        L.rtm_set_frame_runs.argtypes = [vp, C.c_int, C.c_int]
        L.rtm_set_frame_runs.restype = C.c_int
        L.rtm_frame_runs.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.rtm_frame_runs.restype = C.c_int
        L.rtm_kernel_time.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
        for n in ("rtm_create", "rtm_destroy", "rtm_set_model", "rtm_set_wavelet", "rtm_set_shots",
                  "rtm_set_shot_geometry", "rtm_set_shot_model",
                  "rtm_set_gathers", "rtm_gathers_from_records", "rtm_forward_begin", "rtm_forward_advance",
                  "rtm_backward_begin", "rtm_backward_advance", "rtm_migrate", "rtm_set_graph", "rtm_sync",
                  "rtm_read_field", "rtm_read_records", "rtm_read_strip", "rtm_image_post",
                  "rtm_stack_finish", "rtm_set_timing", "rtm_kernel_time", "rtm_autotune"):
            getattr(L, n).restype = C.c_int
        _lib = L
    return _lib


DOT_NOALLOC_MAX_VEC = 64  # RTM_DOT_NOALLOC_MAX_VEC


def data_dot_noalloc(x_ptr, y_ptr, n, nvec=1, stride=0, out_ptr=None, stream=None):
    """FP64 dot products of nvec pairs of device float vectors (addresses
    x_ptr, y_ptr; vector v at v*stride elements) into the device doubles at
    out_ptr, on the given HIP stream handle, with no allocation and no
    synchronisation (rtm_data_dot_noalloc)."""
    rc = lib().rtm_data_dot_noalloc(C.c_void_p(x_ptr), C.c_void_p(y_ptr), int(n), int(stride),
                                    int(nvec), C.c_void_p(out_ptr), C.c_void_p(stream))
    if rc != 0:
        raise RtmError(f"rtm_data_dot_noalloc failed ({rc}): "
                       f"{lib().rtm_last_error(None).decode()}")


class Context:
    """One librtm_hip context (one HIP stream, one batch of shots)."""

    def __init__(self, **params):
        p = RtmParams(**{k: params.get(k, 0) for k, _ in RtmParams._fields_})
        if p.max_shots == 0:
            p.max_shots = 1
        self.p = p
        h = C.c_void_p()
        rc = lib().rtm_create(C.byref(p), C.byref(h))
        if rc != 0:
            raise RtmError(f"rtm_create failed ({rc})")
        self.h = h
        self.nz, self.nx, self.n2 = p.nz, p.nx, p.n2
        self.gw = p.stack_nx if p.stack_nx > 0 else p.nx - 2 * p.n2
        self.nshot = 0

    def _chk(self, rc, what):
        if rc != 0:
            msg = lib().rtm_last_error(self.h).decode()
            raise RtmError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            lib().rtm_destroy(self.h)
            self.h = None

    __del__ = close

    def device_bytes(self):
        return lib().rtm_device_bytes(self.h)

    def set_model(self, v, r1, w, c, index, nvel, vmin, dv):
        self._keep_model = [np.ascontiguousarray(a, np.float32).ravel() for a in (v, r1, w, c)]
        idx = None
        if index is not None:
            self._idx = np.ascontiguousarray(index, np.int32)
            idx = self._idx.ctypes.data
        self._chk(lib().rtm_set_model(self.h, *self._keep_model, len(self._keep_model[3]), idx,
                                      int(nvel), float(vmin), float(dv)), "rtm_set_model")

    def set_wavelet(self, wav):
        self._chk(lib().rtm_set_wavelet(self.h, np.ascontiguousarray(wav, np.float32)),
                  "rtm_set_wavelet")

    def set_shots(self, r_u, r_x):
        r_u = np.ascontiguousarray(r_u, np.int32)
        r_x = np.ascontiguousarray(r_x, np.int32)
        self._chk(lib().rtm_set_shots(self.h, len(r_u), r_u, r_x), "rtm_set_shots")
        self.nshot = len(r_u)

    def set_shot_geometry(self, geoms):
        """geoms: sequence of dicts with nx, r_u, r_x, s_l, s_r, n, x_img."""
        arr = (RtmShotGeom * len(geoms))(*[RtmShotGeom(**g) for g in geoms])
        self._chk(lib().rtm_set_shot_geometry(self.h, len(geoms), arr), "rtm_set_shot_geometry")
        self.nshot = len(geoms)

    def set_shot_model(self, shot, v, r1):
        """Model window of one shot of a ragged batch: v, r1 [nz][nx of the shot]."""
        self._chk(lib().rtm_set_shot_model(self.h, int(shot), np.ascontiguousarray(v, np.float32).ravel(),
                                           np.ascontiguousarray(r1, np.float32).ravel()),
                  "rtm_set_shot_model")

    def set_gathers(self, seis):
        self._chk(lib().rtm_set_gathers(self.h, np.ascontiguousarray(seis, np.float32).ravel()),
                  "rtm_set_gathers")

    def gathers_from_records(self):
        self._chk(lib().rtm_gathers_from_records(self.h), "rtm_gathers_from_records")

    def forward_begin(self, save_strips=True):
        self._chk(lib().rtm_forward_begin(self.h, int(bool(save_strips))), "rtm_forward_begin")

    def forward_advance(self, n):
        self._chk(lib().rtm_forward_advance(self.h, int(n)), "rtm_forward_advance")

    def backward_begin(self):
        self._chk(lib().rtm_backward_begin(self.h), "rtm_backward_begin")

    def backward_advance(self, n):
        self._chk(lib().rtm_backward_advance(self.h, int(n)), "rtm_backward_advance")

    def migrate(self):
        self._chk(lib().rtm_migrate(self.h), "rtm_migrate")

    def set_graph(self, on):
        self._chk(lib().rtm_set_graph(self.h, int(bool(on))), "rtm_set_graph")

    def sync(self):
        self._chk(lib().rtm_sync(self.h), "rtm_sync")

    def field(self, which, shot=0):
        out = np.zeros(self.nz * self.nx, np.float32)
        self._chk(lib().rtm_read_field(self.h, shot, which, out), "rtm_read_field")
        return out.reshape(self.nz, self.nx)

    def records(self):
        out = np.zeros(self.nshot * self.p.n * self.p.nt, np.float32)
        self._chk(lib().rtm_read_records(self.h, out), "rtm_read_records")
        return out.reshape(self.nshot, self.p.n, self.p.nt)

    def strip(self, shot, k):
        ss = 2 * self.p.nfdmax * (self.nz - 2 * self.n2 + self.nx - 2 * self.n2)
        out = np.zeros(ss, np.float32)
        self._chk(lib().rtm_read_strip(self.h, shot, k, out), "rtm_read_strip")
        return out

    def image_post(self, vmax, whitecoe, want=True):
        mnz, mnx = self.nz - 2 * self.n2, self.nx - 2 * self.n2
        up = np.zeros((self.nshot, mnz, mnx), np.float32) if want else None
        down = np.zeros((self.nshot, mnz, mnx), np.float32) if want else None
        self._chk(lib().rtm_image_post(self.h, float(vmax), float(whitecoe),
                                       up.ctypes.data if want else None,
                                       down.ctypes.data if want else None), "rtm_image_post")
        return up, down

    def stack_finish(self, nrec, inorm):
        mnz = self.nz - 2 * self.n2
        img = np.zeros(mnz * self.gw, np.float32)
        ill = np.zeros(mnz * self.gw, np.float32)
        self._chk(lib().rtm_stack_finish(self.h, int(nrec), int(inorm), img, ill.ctypes.data),
                  "rtm_stack_finish")
        return img.reshape(mnz, self.gw), ill.reshape(mnz, self.gw)

    # ---- least-squares migration operators
    def set_reflectivity(self, m):
        """Reflectivity on the stacked image grid [mnz][stack_nx]."""
        mnz = self.nz - 2 * self.n2
        m = np.ascontiguousarray(m, np.float32)
        if m.shape != (mnz, self.gw):
            raise ValueError(f"reflectivity shape {m.shape}, expected {(mnz, self.gw)}")
        self._chk(lib().rtm_set_reflectivity(self.h, m.ravel()), "rtm_set_reflectivity")

    def born_begin(self):
        self._chk(lib().rtm_born_begin(self.h), "rtm_born_begin")

    def born_advance(self, n):
        self._chk(lib().rtm_born_advance(self.h, int(n)), "rtm_born_advance")

    def born(self):
        """Born data of the batch [shot][n][nt] for the current reflectivity."""
        self.born_begin()
        self.born_advance(self.p.nt)
        return self.records()

    def adjoint_begin(self):
        self._chk(lib().rtm_adjoint_begin(self.h), "rtm_adjoint_begin")

    def adjoint_advance(self, n):
        self._chk(lib().rtm_adjoint_advance(self.h, int(n)), "rtm_adjoint_advance")

    def adjoint_gradient(self):
        mnz = self.nz - 2 * self.n2
        out = np.zeros(mnz * self.gw, np.float32)
        self._chk(lib().rtm_adjoint_gradient(self.h, out), "rtm_adjoint_gradient")
        return out.reshape(mnz, self.gw)

    def adjoint(self, resid):
        """Gradient [mnz][stack_nx] of residuals resid[shot][n][nt]; needs a
        complete forward or Born pass with strips just before."""
        self.set_gathers(resid)
        self.adjoint_begin()
        self.adjoint_advance(self.p.nt)
        return self.adjoint_gradient()

    def data_dot(self, x, y, shot0=0, nshot=0):
        """FP64 dot product of device data vectors (DATA_GATHERS / DATA_RECORDS)
        over shots [shot0, shot0 + nshot) (nshot 0: through the last)."""
        out = C.c_double()
        self._chk(lib().rtm_data_dot(self.h, int(x), int(y), int(shot0), int(nshot), C.byref(out)),
                  "rtm_data_dot")
        return out.value

    def data_axpy(self, a, shot0=0, nshot=0):
        """gathers += a * records on the device, over shots [shot0, shot0 + nshot)."""
        self._chk(lib().rtm_data_axpy(self.h, float(a), int(shot0), int(nshot)), "rtm_data_axpy")

    def autotune(self, nsteps=8):
        """Pick the row-segment height of each interior kernel (forward, source
        reconstruction, receiver step) by timing its launches; returns the three."""
        z = (C.c_int * 3)()
        self._chk(lib().rtm_autotune(self.h, int(nsteps), z), "rtm_autotune")
        return list(z)

    def set_zseg(self, fwd, src, rcv):
        """Rows per wave of the forward, reconstruction and receiver-step kernels."""
        z = (C.c_int * 3)(int(fwd), int(src), int(rcv))
        self._chk(lib().rtm_set_zseg(self.h, z), "rtm_set_zseg")

    def set_frame_runs(self, rows_tb, rows_side):
        """Rows per frame-kernel wave: top/bottom bands, side bands."""
        self._chk(lib().rtm_set_frame_runs(self.h, int(rows_tb), int(rows_side)),
                  "rtm_set_frame_runs")

    def frame_runs(self):
        a, b = C.c_int(), C.c_int()
        self._chk(lib().rtm_frame_runs(self.h, C.byref(a), C.byref(b)), "rtm_frame_runs")
        return a.value, b.value

    def set_timing(self, mask):
        """mask: bit set of kernel kinds (1 << K_*), 0 off, -1 all."""
        self._chk(lib().rtm_set_timing(self.h, int(mask)), "rtm_set_timing")

    def kernel_time(self, kind):
        ms = C.c_double()
        n = C.c_longlong()
        self._chk(lib().rtm_kernel_time(self.h, kind, C.byref(ms), C.byref(n)), "rtm_kernel_time")
        return ms.value, n.value
