"""Python binding of the deblurring network's C ABI (include/rtm_cnn.h,
librtm_hip.so).  The network runs on the GPU only; there is no CPU path."""
import ctypes as C

import numpy as np

from . import hip

_fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_bound = False


def _lib():
    global _bound
    L = hip.lib()
    if not _bound:
        vp = C.c_void_p
        L.rtm_cnn_create.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.c_longlong, C.c_longlong, C.POINTER(vp)]
        L.rtm_cnn_destroy.argtypes = [vp]
        L.rtm_cnn_last_error.argtypes = [vp]
        L.rtm_cnn_last_error.restype = C.c_char_p
        L.rtm_cnn_num_params.argtypes = [vp]
        L.rtm_cnn_set_params.argtypes = [vp, _fp]
        L.rtm_cnn_get_params.argtypes = [vp, _fp]
        L.rtm_cnn_forward.argtypes = [vp, _fp, C.c_int, C.c_int, C.c_int, _fp]
        L.rtm_cnn_loss_grad.argtypes = [vp, _fp, _fp, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_double), vp]
        L.rtm_cnn_adam_step.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
        L.rtm_cnn_fit.argtypes = [vp, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_float, C.c_uint, vp]
        _bound = True
    return L


class Deblur:
    """y = x[:, 0] + net(x): nlayer 3x3 convolutions (channels[0] inputs, one
    output, optional per-layer dilation), BF16 on the matrix cores with FP32
    accumulation."""

    def __init__(self, channels, max_pixels, max_rows, dilation=None):
        self.channels = list(channels)
        nl = len(self.channels) - 1
        self.dilation = list(dilation) if dilation is not None else [1] * nl
        if len(self.dilation) != nl:
            raise ValueError(f"{len(self.dilation)} dilations for {nl} layers")
        arr = (C.c_int * len(self.channels))(*self.channels)
        dil = (C.c_int * nl)(*self.dilation)
        h = C.c_void_p()
        rc = _lib().rtm_cnn_create(nl, arr, dil, int(max_pixels), int(max_rows), C.byref(h))
        if rc != 0:
            raise hip.RtmError(f"rtm_cnn_create failed ({rc})")
        self.h = h
        self.np = _lib().rtm_cnn_num_params(self.h)

    def close(self):
        if self.h:
            _lib().rtm_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = _lib().rtm_cnn_last_error(self.h)
            raise hip.RtmError(f"{what}: {msg.decode() if msg else rc}")

    def init_params(self, seed=0, scale=1.0, last_zero=True):
        """He-normal weights, zero biases; the last layer zero so the network
        starts as the identity on channel 0 (plain migration as the fallback)."""
        rng = np.random.default_rng(seed)
        p = np.zeros(self.np, np.float32)
        o = 0
        nl = len(self.channels) - 1
        for i in range(nl):
            ci, co = self.channels[i], self.channels[i + 1]
            nw = co * ci * 9
            if not (last_zero and i == nl - 1):
                p[o:o + nw] = rng.standard_normal(nw) * np.sqrt(2.0 / (ci * 9)) * scale
            o += nw + co
        self.set_params(p)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.np:
            raise ValueError(f"{p.size} parameters, expected {self.np}")
        self._chk(_lib().rtm_cnn_set_params(self.h, p), "rtm_cnn_set_params")

    def params(self):
        p = np.zeros(self.np, np.float32)
        self._chk(_lib().rtm_cnn_get_params(self.h, p), "rtm_cnn_get_params")
        return p

    def save(self, path, **extra):
        """Parameters, widths and dilations (and any extra plain arrays) to a .npz."""
        np.savez(path, params=self.params(), channels=np.asarray(self.channels, np.int32),
                 dilation=np.asarray(self.dilation, np.int32), **extra)

    @classmethod
    def load(cls, path, max_pixels, max_rows):
        """A network saved by save(), for batches up to max_pixels / max_rows."""
        with np.load(path, allow_pickle=False) as f:
            net = cls([int(v) for v in f["channels"]], max_pixels, max_rows,
                      dilation=[int(v) for v in f["dilation"]])
            net.set_params(f["params"])
        return net

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        n, _, hh, ww = x.shape
        y = np.zeros((n, hh, ww), np.float32)
        self._chk(_lib().rtm_cnn_forward(self.h, x, n, hh, ww, y), "rtm_cnn_forward")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        n, _, hh, ww = x.shape
        loss = C.c_double()
        g = np.zeros(self.np, np.float32)
        self._chk(_lib().rtm_cnn_loss_grad(self.h, x, t, n, hh, ww, C.byref(loss), g.ctypes.data),
                  "rtm_cnn_loss_grad")
        return loss.value, g

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self._chk(_lib().rtm_cnn_adam_step(self.h, lr, beta1, beta2, eps), "rtm_cnn_adam_step")

    def fit(self, x, t, epochs, batch, lr, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        n, _, hh, ww = x.shape
        hist = np.zeros(max(epochs, 1), np.float64)
        self._chk(_lib().rtm_cnn_fit(self.h, x, t, n, hh, ww, int(epochs), int(batch), float(lr),
                                     int(seed), hist.ctypes.data), "rtm_cnn_fit")
        return hist[:epochs]
