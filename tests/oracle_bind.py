"""ctypes bindings to the oracle (oracle/liboracle.so) and to the compiled
reference helpers (oracle/_ref/libref.so).  TEST INFRASTRUCTURE ONLY: imported by
tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product package."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libref.so")

fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OrcCfg(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nz", "nx", "n2", "nfdmax", "nt", "nt2", "ilste",
                                        "icompen", "s_l", "s_r", "s_z", "ds", "n", "r_u", "r_x")] + \
               [(n, C.c_float) for n in ("tao2", "h2", "taoh", "taoh2", "hzx2_1", "vmin", "dv")] + \
               [(n, C.c_int) for n in ("record", "nthreads", "strip_slots", "snap_every")]


def _build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "oracle"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            _build_oracle()
        L = C.CDLL(ORACLE_SO)
        L.orc_derive.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                 C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.c_int, ip, fp]
        L.orc_source_row.argtypes = [C.c_float, C.c_float, C.c_int]
        L.orc_source_row.restype = C.c_int
        L.orc_ricker.argtypes = [C.c_float, C.c_float]
        L.orc_ricker.restype = C.c_float
        L.orc_pad_model.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                    fp, fp]
        L.orc_vrange.argtypes = [fp, C.c_int, C.c_float, C.POINTER(C.c_float),
                                 C.POINTER(C.c_float), C.POINTER(C.c_int)]
        L.orc_bins_needed.argtypes = [fp, C.c_int, C.c_float, C.c_float, C.c_int, ip]
        L.orc_blend_weights.argtypes = [C.c_int, fp]
        L.orc_design_ls.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_double] * 7 + \
            [ip, C.c_double, ip, ip, fp, C.c_int]
        L.orc_design_ls.restype = C.c_int
        L.orc_lenfd.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 4 + \
            [C.c_int, C.c_double, C.c_double]
        L.orc_lenfd.restype = C.c_int
        L.orc_lscoef.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, dp]
        L.orc_taylor.argtypes = [C.c_int, fp]
        L.orc_disp_maxerr.argtypes = [dp, C.c_int] + [C.c_double] * 4 + [C.c_int, C.c_int,
                                                                          C.c_double]
        L.orc_disp_maxerr.restype = C.c_double
        L.orc_create.argtypes = [C.POINTER(OrcCfg), C.c_int]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_model.argtypes = [C.c_void_p, fp, fp, fp, fp, C.c_int, C.c_void_p, C.c_int, fp]
        for n in ("orc_forward_begin", "orc_born_begin"):
            getattr(L, n).argtypes = [C.c_void_p]
            getattr(L, n).restype = C.c_int
        for n in ("orc_forward_advance", "orc_backward_advance", "orc_born_advance",
                  "orc_adjoint_advance"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_int]
            getattr(L, n).restype = C.c_int
        for n in ("orc_backward_begin", "orc_adjoint_begin"):
            getattr(L, n).argtypes = [C.c_void_p, fp]
            getattr(L, n).restype = C.c_int
        L.orc_set_reflectivity.argtypes = [C.c_void_p, fp]
        L.orc_read_field.argtypes = [C.c_void_p, C.c_int, fp]
        L.orc_read_field64.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_read_record.argtypes = [C.c_void_p, fp]
        L.orc_read_strip.argtypes = [C.c_void_p, C.c_int, fp]
        L.orc_image_post.argtypes = [fp, fp, fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                     fp, fp]
        L.orc_stack.argtypes = [fp, fp, C.c_int, C.c_int, C.c_int, fp, fp]
        L.orc_time_shots.argtypes = [C.POINTER(OrcCfg), fp, fp, fp, fp, C.c_int, C.c_void_p,
                                     C.c_int, fp, fp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_time_shots.restype = C.c_double
        L.orc_time_rows.argtypes = [C.POINTER(OrcCfg), fp, fp, fp, fp, C.c_int, C.c_void_p,
                                    C.c_int, fp, fp, C.c_int, C.c_int, C.c_int]
        L.orc_time_rows.restype = C.c_double
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_SO)
        R.ref_funMandC.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_double] * 8 + \
            [ip, ip, ip, C.c_double, fp, C.c_int]
        R.ref_funMandC.restype = C.c_int
        R.ref_lenfd.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 4 + \
            [C.c_int, C.c_double, C.c_double]
        R.ref_lenfd.restype = C.c_int
        R.ref_lscoef.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, dp]
        R.ref_taylor.argtypes = [C.c_int, fp]
        R.ref_velocity.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                   C.c_int, fp, fp]
        R.ref_write_sgy.argtypes = [fp, C.c_int, C.c_int, C.c_int, fp, fp, C.c_float, C.c_float,
                                    fp, C.c_char_p]
        R.ref_trace2segy.argtypes = [C.c_char_p, fp, C.c_int, C.c_int]
        R.ref_segy2trace.argtypes = [C.c_char_p, fp, C.c_int, C.c_int]
        R.ref_d2t.argtypes = [C.c_char_p, fp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                              C.c_float, C.c_float]
        R.ref_d2t.restype = C.c_int
        R.ref_t2d.argtypes = [C.c_char_p, fp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_float, C.c_float, C.c_float]
        R.ref_t2d.restype = C.c_int
        R.ref_phase_correction.argtypes = [fp, fp, C.c_int, C.c_int, C.c_float]
        _ref = R
    return _ref


# ---- convenience wrappers -----------------------------------------------------
def derive(mod_nz, mod_nx, n2, nt1, tao1, tao, h, hz, f0, s_l1, s_z1, n, ds, r_x1):
    io = np.zeros(8, np.int32)
    fo = np.zeros(6, np.float32)
    lib().orc_derive(mod_nz, mod_nx, n2, nt1, tao1, tao, h, hz, f0, s_l1, s_z1, n, ds, r_x1, io, fo)
    keys_i = ("NZ", "NX", "NT", "NT2", "s_l", "s_r", "s_z", "r_x")
    keys_f = ("taoh", "tao2", "h2", "taoh2", "hzx", "hzx2_1")
    d = {k: int(v) for k, v in zip(keys_i, io)}
    d.update({k: np.float32(v) for k, v in zip(keys_f, fo)})
    return d


def pad_model(raw_colmajor, mod_nz, mod_nx, n2, ifv, tao, h):
    nz, nx = mod_nz + 2 * n2, mod_nx + 2 * n2
    v = np.zeros(nz * nx, np.float32)
    r1 = np.zeros(nz * nx, np.float32)
    lib().orc_pad_model(np.ascontiguousarray(raw_colmajor, np.float32).ravel(), mod_nz, mod_nx, n2,
                        ifv, tao, h, v, r1)
    return v.reshape(nz, nx), r1.reshape(nz, nx)


def vrange(v, dv):
    vmin, vmax, nvel = C.c_float(), C.c_float(), C.c_int()
    a = np.ascontiguousarray(v, np.float32).ravel()
    lib().orc_vrange(a, a.size, dv, C.byref(vmin), C.byref(vmax), C.byref(nvel))
    return np.float32(vmin.value), np.float32(vmax.value), nvel.value


def bins_needed(v, vmin, dv, nvel):
    a = np.ascontiguousarray(v, np.float32).ravel()
    need = np.zeros(nvel, np.int32)
    lib().orc_bins_needed(a, a.size, vmin, dv, nvel, need)
    return need


def blend_weights(n2):
    w = np.zeros(n2 + 1, np.float32)
    lib().orc_blend_weights(n2, w)
    return w


def design_ls(nthita, nfdmax, nfdmin, nvel, tao, h, df, eps, fmax, vmin, dv, need, hzx):
    M = np.zeros(nvel, np.int32)
    Index = np.zeros(nvel + 1, np.int32)
    cap = nvel * (nfdmax + 1) + 1
    c = np.zeros(cap, np.float32)
    nc = lib().orc_design_ls(nthita, nfdmax, nfdmin, nvel, float(tao), float(h), float(df),
                             float(eps), float(fmax), float(vmin), float(dv),
                             np.ascontiguousarray(need, np.int32), float(hzx), M, Index, c, cap)
    return M, Index, c[:nc].copy()


def ref_design_ls(nthita, nfdmax, nfdmin, nvel, tao, h, df, eps, fmax, vmin, vmax, dv, need, hzx):
    M = np.zeros(nvel, np.int32)
    Index = np.zeros(nvel + 1, np.int32)
    cap = nvel * (nfdmax + 1) + 1
    c = np.zeros(cap, np.float32)
    nc = ref().ref_funMandC(nthita, nfdmax, nfdmin, nvel, float(tao), float(h), float(df),
                            float(eps), float(fmax), float(vmin), float(vmax), float(dv),
                            np.ascontiguousarray(need, np.int32).copy(), M, Index, float(hzx), c,
                            cap)
    return M, Index, c[:nc].copy()


def taylor(nfdmax):
    c = np.zeros(nfdmax + 1, np.float32)
    lib().orc_taylor(nfdmax, c)
    return c


def ricker(t, f0):
    return np.float32(lib().orc_ricker(np.float32(t), np.float32(f0)))


def wavelet_table(nt, tao, f0):
    """w[j] = f(j*tao) with the float product (k-1)*tao of kernel.cu:812."""
    tao = np.float32(tao)
    return np.array([lib().orc_ricker(np.float32(np.float32(j) * tao), np.float32(f0))
                     for j in range(nt + 1)], np.float32)


class Engine:
    """One-shot oracle engine (prec 32 = FP32 replica, 64 = FP64 oracle)."""

    def __init__(self, cfg: dict, model: dict, prec=32):
        self.cfg = OrcCfg(**{k: cfg.get(k, 0) for k, _ in OrcCfg._fields_})
        self.h = lib().orc_create(C.byref(self.cfg), prec)
        self.nz, self.nx = cfg["nz"], cfg["nx"]
        index = model.get("index")
        self._keep = [np.ascontiguousarray(model[k], np.float32).ravel()
                      for k in ("v", "r1", "w", "c", "wav")]
        idx_ptr = None
        if index is not None:
            self._idx = np.ascontiguousarray(index, np.int32)
            idx_ptr = self._idx.ctypes.data
        lib().orc_set_model(self.h, *self._keep[:4], len(self._keep[3]), idx_ptr,
                            int(model.get("nvel", 0)), self._keep[4])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def forward_begin(self):
        return lib().orc_forward_begin(self.h)

    def forward_advance(self, n):
        return lib().orc_forward_advance(self.h, n)

    def backward_begin(self, seis):
        self._seis = np.ascontiguousarray(seis, np.float32).ravel()
        return lib().orc_backward_begin(self.h, self._seis)

    def backward_advance(self, n):
        return lib().orc_backward_advance(self.h, n)

    def set_reflectivity(self, m):
        self._refl = np.ascontiguousarray(m, np.float32).ravel()
        lib().orc_set_reflectivity(self.h, self._refl)

    def born_begin(self):
        return lib().orc_born_begin(self.h)

    def born_advance(self, n):
        return lib().orc_born_advance(self.h, n)

    def adjoint_begin(self, resid):
        self._seis = np.ascontiguousarray(resid, np.float32).ravel()
        return lib().orc_adjoint_begin(self.h, self._seis)

    def adjoint_advance(self, n):
        return lib().orc_adjoint_advance(self.h, n)

    def field(self, which, f64=False):
        if f64:
            out = np.zeros(self.nz * self.nx, np.float64)
            lib().orc_read_field64(self.h, which, out)
        else:
            out = np.zeros(self.nz * self.nx, np.float32)
            lib().orc_read_field(self.h, which, out)
        return out.reshape(self.nz, self.nx)

    def record(self):
        out = np.zeros(self.cfg.n * self.cfg.nt, np.float32)
        lib().orc_read_record(self.h, out)
        return out.reshape(self.cfg.n, self.cfg.nt)

    def strip(self, k):
        n2, nfd = self.cfg.n2, self.cfg.nfdmax
        ss = 2 * nfd * (self.nz - 2 * n2 + self.nx - 2 * n2)
        out = np.zeros(ss, np.float32)
        lib().orc_read_strip(self.h, k, out)
        return out


def image_post(I1, I2, v, n2, vmax, whitecoe):
    nz, nx = I1.shape
    up = np.zeros((nz - 2 * n2) * (nx - 2 * n2), np.float32)
    down = np.zeros_like(up)
    lib().orc_image_post(np.ascontiguousarray(I1, np.float32).ravel(),
                         np.ascontiguousarray(I2, np.float32).ravel(),
                         np.ascontiguousarray(v, np.float32).ravel(), nz, nx, n2,
                         np.float32(vmax), np.float32(whitecoe), up, down)
    return up.reshape(nz - 2 * n2, nx - 2 * n2), down.reshape(nz - 2 * n2, nx - 2 * n2)


def stack(ups, downs, inorm):
    ups = np.ascontiguousarray(ups, np.float32)
    downs = np.ascontiguousarray(downs, np.float32)
    nshot = ups.shape[0]
    nI = ups[0].size
    img = np.zeros(nI, np.float32)
    ill = np.zeros(nI, np.float32)
    lib().orc_stack(ups.ravel(), downs.ravel(), nshot, nI, inorm, img, ill)
    return img.reshape(ups.shape[1:]), ill.reshape(ups.shape[1:])
