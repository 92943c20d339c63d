"""ctypes binding of librtm_host.so (include/rtm_host.h): parameter files,
setup arithmetic, model padding, coefficient design, SEG-Y and the driver."""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "librtm_host.so")

_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class Config(C.Structure):
    _fields_ = [("nfdmax", C.c_int), ("nfdmin", C.c_int), ("n2", C.c_int), ("f0", C.c_float),
                ("fmax", C.c_float), ("df", C.c_float), ("nthita", C.c_int), ("eps", C.c_float),
                ("dv", C.c_float), ("ilste", C.c_int), ("ifv", C.c_int), ("whitecoe", C.c_float),
                ("hz", C.c_float), ("tao", C.c_float), ("inorm", C.c_int), ("icompen", C.c_int),
                ("nsmooth", C.c_int), ("wthite_phase", C.c_float), ("angle", C.c_float),
                ("nx_bg", C.c_int), ("nx_ed", C.c_int), ("nz_bg", C.c_int), ("nz_ed", C.c_int),
                ("out_seis", C.c_char * 512), ("out_vp", C.c_char * 512),
                ("out_dpr", C.c_char * 512), ("out_para", C.c_char * 512),
                ("result", C.c_char * 512), ("h", C.c_float), ("tao1", C.c_float),
                ("mod_nz", C.c_int), ("mod_nx", C.c_int), ("nt1", C.c_int), ("s_l1", C.c_int),
                ("s_z1", C.c_int), ("n", C.c_int), ("ds", C.c_int), ("r_x1", C.c_int),
                ("nrec", C.c_int), ("dr", C.c_int)]


class Derived(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nz", "nx", "nt", "nt2", "s_l", "s_r", "s_z", "r_x")] + \
               [(n, C.c_float) for n in ("taoh", "tao2", "h2", "taoh2", "hzx", "hzx2_1")]


BEFORE_BATCH = C.CFUNCTYPE(C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_void_p)


class RunOpts(C.Structure):
    _fields_ = [("workers", C.c_int), ("batch", C.c_int), ("retries", C.c_int),
                ("strip_format", C.c_int), ("resume", C.c_int), ("verbose", C.c_int),
                ("sgy_template", C.c_char_p), ("before_batch", BEFORE_BATCH),
                ("user", C.c_void_p)]


class HostError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HostError(f"{LIB_PATH} is missing: run `python -m rtm_gpu_amd.build` first")
        L = C.CDLL(LIB_PATH)
        L.rtmh_last_error.restype = C.c_char_p
        L.rtmh_read_config.argtypes = [C.c_char_p, C.POINTER(Config)]
        L.rtmh_read_depths.argtypes = [C.c_char_p, C.c_int, _fp]
        L.rtmh_derive.argtypes = [C.POINTER(Config), C.POINTER(Derived)]
        L.rtmh_derive.restype = None
        L.rtmh_source_row.argtypes = [C.c_float, C.c_float, C.c_int]
        L.rtmh_ricker.argtypes = [C.c_float, C.c_float]
        L.rtmh_ricker.restype = C.c_float
        L.rtmh_wavelet_table.argtypes = [C.c_int, C.c_float, C.c_float, _fp]
        L.rtmh_wavelet_table.restype = None
        L.rtmh_blend_weights.argtypes = [C.c_int, _fp]
        L.rtmh_blend_weights.restype = None
        L.rtmh_pad_model.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                     _fp, _fp]
        L.rtmh_pad_model.restype = None
        L.rtmh_vrange.argtypes = [_fp, C.c_longlong, C.c_float, C.POINTER(C.c_float),
                                  C.POINTER(C.c_float), C.POINTER(C.c_int)]
        L.rtmh_vrange.restype = None
        L.rtmh_bins_needed.argtypes = [_fp, C.c_longlong, C.c_float, C.c_float, C.c_int, _ip]
        L.rtmh_bins_needed.restype = None
        L.rtmh_design_ls.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_double] * 7 + \
            [_ip, C.c_double, _ip, _ip, _fp, C.c_int, C.c_int]
        L.rtmh_taylor.argtypes = [C.c_int, _fp]
        L.rtmh_taylor.restype = None
        L.rtmh_stability.argtypes = [_fp, C.c_longlong, _fp, C.c_void_p, C.c_int, C.c_float,
                                     C.c_float, C.c_int, C.c_float, C.c_float, C.c_float]
        L.rtmh_stability.restype = C.c_double
        L.rtmh_float_to_ibm.argtypes = [_fp, _bp, C.c_longlong]
        L.rtmh_float_to_ibm.restype = None
        L.rtmh_ibm_to_float.argtypes = [_bp, _fp, C.c_longlong]
        L.rtmh_ibm_to_float.restype = None
        L.rtmh_write_segy.argtypes = [C.c_char_p, C.c_char_p, _fp, C.c_int, C.c_int, C.c_int, _fp,
                                      _fp, C.c_float, C.c_float, _fp]
        L.rtmh_read_segy.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.c_void_p, C.c_longlong]
        L.rtmh_stack.argtypes = [_fp, _fp, C.c_int, C.c_longlong, C.c_int, _fp, _fp]
        L.rtmh_stack.restype = None
        L.rtmh_d2t.argtypes = [_fp, _fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                               C.c_void_p, C.c_longlong]
        L.rtmh_t2d.argtypes = [_fp, _fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                               C.c_float, C.c_void_p, C.c_longlong]
        L.rtmh_phase_rotate.argtypes = [_fp, _fp, C.c_int, C.c_int, C.c_float]
        L.rtmh_phase_rotate.restype = None
        L.rtmh_run.argtypes = [C.c_char_p, C.POINTER(RunOpts)]
        _lib = L
    return _lib


def _chk(rc, what):
    if rc < 0:
        raise HostError(f"{what} failed ({rc}): {lib().rtmh_last_error().decode()}")
    return rc


def read_config(path):
    c = Config()
    _chk(lib().rtmh_read_config(str(path).encode(), C.byref(c)), "rtmh_read_config")
    return c


def derive(cfg):
    d = Derived()
    lib().rtmh_derive(C.byref(cfg), C.byref(d))
    return d


def make_config(**kw):
    c = Config()
    for k, v in kw.items():
        if isinstance(v, str):
            v = v.encode()
        setattr(c, k, v)
    return c


def source_row(depth, hz, n2):
    return lib().rtmh_source_row(np.float32(depth), np.float32(hz), int(n2))


def wavelet_table(nt, tao, f0):
    w = np.zeros(nt + 1, np.float32)
    lib().rtmh_wavelet_table(int(nt), np.float32(tao), np.float32(f0), w)
    return w


def blend_weights(n2):
    w = np.zeros(n2 + 1, np.float32)
    lib().rtmh_blend_weights(int(n2), w)
    return w


def pad_model(raw_colmajor, mod_nz, mod_nx, n2, ifv, tao, h):
    nz, nx = mod_nz + 2 * n2, mod_nx + 2 * n2
    v = np.zeros(nz * nx, np.float32)
    r1 = np.zeros(nz * nx, np.float32)
    lib().rtmh_pad_model(np.ascontiguousarray(raw_colmajor, np.float32).ravel(), mod_nz, mod_nx,
                         n2, ifv, np.float32(tao), np.float32(h), v, r1)
    return v.reshape(nz, nx), r1.reshape(nz, nx)


def vrange(v, dv):
    a = np.ascontiguousarray(v, np.float32).ravel()
    vmin, vmax, nvel = C.c_float(), C.c_float(), C.c_int()
    lib().rtmh_vrange(a, a.size, np.float32(dv), C.byref(vmin), C.byref(vmax), C.byref(nvel))
    return np.float32(vmin.value), np.float32(vmax.value), nvel.value


def bins_needed(v, vmin, dv, nvel):
    a = np.ascontiguousarray(v, np.float32).ravel()
    need = np.zeros(nvel, np.int32)
    lib().rtmh_bins_needed(a, a.size, np.float32(vmin), np.float32(dv), nvel, need)
    return need


def design_ls(nthita, nfdmax, nfdmin, nvel, tao, h, df, eps, fmax, vmin, dv, need, hzx,
              nthreads=8):
    M = np.zeros(nvel, np.int32)
    index = np.zeros(nvel + 1, np.int32)
    cap = nvel * (nfdmax + 1) + 1
    c = np.zeros(cap, np.float32)
    nc = _chk(lib().rtmh_design_ls(nthita, nfdmax, nfdmin, nvel, float(tao), float(h), float(df),
                                   float(eps), float(fmax), float(vmin), float(dv),
                                   np.ascontiguousarray(need, np.int32), float(hzx), M, index, c,
                                   cap, int(nthreads)), "rtmh_design_ls")
    return M, index, c[:nc].copy()


def taylor(nfdmax):
    c = np.zeros(nfdmax + 1, np.float32)
    lib().rtmh_taylor(nfdmax, c)
    return c


def stability(v, c, index, nvel, vmin, dv, nfdmax, tao2, h2, hzx2_1):
    """The time-stepping stability number of a model (rtmh_stability): stable
    when <= 1.  index None: one Taylor row c[nfdmax+1]."""
    v = np.ascontiguousarray(v, np.float32).ravel()
    c = np.ascontiguousarray(c, np.float32).ravel()
    idx = None
    if index is not None:
        idx = np.ascontiguousarray(index, np.int32)
    return lib().rtmh_stability(v, v.size, c, None if idx is None else idx.ctypes.data,
                                int(nvel), float(vmin), float(dv), int(nfdmax), float(tao2),
                                float(h2), float(hzx2_1))


def float_to_ibm(x):
    x = np.ascontiguousarray(x, np.float32).ravel()
    out = np.zeros(4 * x.size, np.uint8)
    lib().rtmh_float_to_ibm(x, out, x.size)
    return out


def ibm_to_float(b):
    b = np.ascontiguousarray(b, np.uint8).ravel()
    out = np.zeros(b.size // 4, np.float32)
    lib().rtmh_ibm_to_float(b, out, out.size)
    return out


def write_segy(template, path, data, dt_units, sx, sy, rx, ry, dsr):
    data = np.ascontiguousarray(data, np.float32)
    ntr, ns = data.shape
    _chk(lib().rtmh_write_segy(str(template).encode(), str(path).encode(), data.ravel(), ntr, ns,
                               int(dt_units), np.ascontiguousarray(sx, np.float32),
                               np.ascontiguousarray(sy, np.float32), float(rx), float(ry),
                               np.ascontiguousarray(dsr, np.float32)), "rtmh_write_segy")


def read_segy(path):
    ntr, ns, fmt = C.c_int(), C.c_int(), C.c_int()
    _chk(lib().rtmh_read_segy(str(path).encode(), C.byref(ntr), C.byref(ns), C.byref(fmt), None,
                              0), "rtmh_read_segy")
    out = np.zeros((ntr.value, ns.value), np.float32)
    _chk(lib().rtmh_read_segy(str(path).encode(), C.byref(ntr), C.byref(ns), C.byref(fmt),
                              out.ctypes.data, out.size), "rtmh_read_segy")
    return out, fmt.value


def stack(ups, downs, inorm):
    ups = np.ascontiguousarray(ups, np.float32)
    downs = np.ascontiguousarray(downs, np.float32)
    nI = ups[0].size
    img = np.zeros(nI, np.float32)
    ill = np.zeros(nI, np.float32)
    lib().rtmh_stack(ups.ravel(), downs.ravel(), ups.shape[0], nI, int(inorm), img, ill)
    return img.reshape(ups.shape[1:]), ill.reshape(ups.shape[1:])


def d2t(v, d, dz, dt, x_st=0, x_end=None):
    """Depth -> two-way time of trace-major v, d ([column][depth sample])."""
    v = np.ascontiguousarray(v, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    nx, nz = v.shape
    x_end = nx - 1 if x_end is None else x_end
    nt = lib().rtmh_d2t(v.ravel(), d.ravel(), nx, nz, x_st, x_end, dz, dt, None, 0)
    _chk(min(nt, 0), "rtmh_d2t")
    out = np.zeros((x_end - x_st + 1, nt), np.float32)
    _chk(min(lib().rtmh_d2t(v.ravel(), d.ravel(), nx, nz, x_st, x_end, dz, dt, out.ctypes.data,
                            out.size), 0), "rtmh_d2t")
    return out


def t2d(v, d, dtt, dzz, x_st=0, x_end=None):
    """Two-way time -> depth of trace-major d ([column][time sample]) over the
    depth velocity v ([column][depth sample])."""
    v = np.ascontiguousarray(v, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    nx, nzv = v.shape
    nt = d.shape[1]
    x_end = nx - 1 if x_end is None else x_end
    nzo = lib().rtmh_t2d(v.ravel(), d.ravel(), nx, nt, nzv, x_st, x_end, dtt, dzz, None, 0)
    _chk(min(nzo, 0), "rtmh_t2d")
    out = np.zeros((x_end - x_st + 1, nzo), np.float32)
    _chk(min(lib().rtmh_t2d(v.ravel(), d.ravel(), nx, nt, nzv, x_st, x_end, dtt, dzz,
                            out.ctypes.data, out.size), 0), "rtmh_t2d")
    return out


def phase_rotate(d, angle):
    """Constant-phase rotation of the traces d ([trace][sample]) by angle degrees."""
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros_like(d)
    lib().rtmh_phase_rotate(d.ravel(), out.ravel(), d.shape[0], d.shape[1], float(angle))
    return out


def run(rtm_txt, workers=1, batch=8, retries=1, strip_format=0, resume=1, verbose=0,
        sgy_template=None, before_batch=None):
    """The reference-style driver (rtmh_run).  before_batch(shots, attempt) ->
    bool, if given, is called before each attempt of a batch (shots: 0-based
    source indices); returning False fails that attempt, which is re-queued."""
    cb = BEFORE_BATCH()
    if before_batch is not None:
        def _cb(p, n, attempt, _user):
            return 0 if before_batch([p[i] for i in range(n)], attempt) else 1
        cb = BEFORE_BATCH(_cb)
    o = RunOpts(workers, batch, retries, strip_format, resume, verbose,
                None if sgy_template is None else str(sgy_template).encode(), cb, None)
    return _chk(lib().rtmh_run(str(rtm_txt).encode(), C.byref(o)), "rtmh_run")
