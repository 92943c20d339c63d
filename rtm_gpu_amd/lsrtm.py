"""Least-squares reverse time migration: conjugate gradients on the normal
equations (CGLS) of Born modelling, min_m ||L m - d||^2, with FP64 inner
products.  Not in the reference, which migrates only (kernel.cu:786-1059);
L and L^T are rtm_born_* and rtm_adjoint_* of librtm_hip.so.

The iteration (Hestenes-Stiefel CGLS, e.g. Paige & Saunders 1982, "LSQR",
section 7.1):

    r = d - L m0,  s = L^T r,  p = s,  gamma = <s, s>
    repeat:  q = L p,  alpha = gamma / <q, q>
             m += alpha p,  r -= alpha q
             s = L^T r,  gamma' = <s, s>,  p = s + (gamma' / gamma) p

The data-space vectors (r, q: shots x receivers x time samples) stay on the
GPU: r is the context's gathers, which the adjoint pass reads, and q its
records, which the Born pass writes; their products and the update of r run
there (rtm_data_dot, rtm_data_axpy).  Model-space vectors (m, p, s: the
stacked image grid) are small and kept on the host in float32; every inner
product is accumulated in float64.

One iteration costs one Born pass (background forward with strips + scattered
field) and one adjoint pass, which reuses the Born pass's background."""
import numpy as np

from . import hip


def dot64(a, b):
    """Inner product of two float arrays accumulated in float64."""
    return float(np.dot(np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()))


class BornOperator:
    """L and L^T of one shot batch of a hip.Context (created with record = 1),
    with the data-space vectors resident in the context."""

    def __init__(self, ctx):
        if not ctx.p.record:
            raise hip.RtmError("the Born operator needs a context created with record=1")
        self.ctx = ctx
        self._background = False  # a forward / Born pass has left its end state and strips

    def start(self, d):
        """r = d (observed data [shot][n][nt]); returns <d, d>."""
        self.ctx.set_gathers(d)
        return self.ctx.data_dot(hip.DATA_GATHERS, hip.DATA_GATHERS)

    def forward(self, p):
        """q = L p into the records; returns <q, q>."""
        ctx = self.ctx
        ctx.set_reflectivity(p)
        ctx.born_begin()
        ctx.born_advance(ctx.p.nt)
        self._background = True
        return ctx.data_dot(hip.DATA_RECORDS, hip.DATA_RECORDS)

    def update_residual(self, alpha):
        """r -= alpha q; returns <r, r>."""
        self.ctx.data_axpy(-np.float32(alpha))
        return self.ctx.data_dot(hip.DATA_GATHERS, hip.DATA_GATHERS)

    def adjoint(self):
        """s = L^T r (the stacked gradient)."""
        ctx = self.ctx
        if not self._background:
            ctx.forward_begin(True)
            ctx.forward_advance(ctx.p.nt)
        ctx.adjoint_begin()
        ctx.adjoint_advance(ctx.p.nt)
        self._background = False
        return ctx.adjoint_gradient()

    def model_shape(self):
        return (self.ctx.nz - 2 * self.ctx.n2, self.ctx.gw)


class MatrixOperator:
    """A dense matrix with the BornOperator interface (float32 vectors,
    float64 products): the same CG code runs on it without a GPU."""

    def __init__(self, A):
        self.A = np.asarray(A, np.float32)
        self.r = None
        self.q = None

    def start(self, d):
        self.r = np.asarray(d, np.float32).copy()
        return dot64(self.r, self.r)

    def forward(self, p):
        self.q = (self.A.astype(np.float64) @ np.asarray(p, np.float64).ravel()).astype(np.float32)
        return dot64(self.q, self.q)

    def update_residual(self, alpha):
        self.r = self.r + (-np.float32(alpha)) * self.q
        return dot64(self.r, self.r)

    def adjoint(self):
        return (self.A.T.astype(np.float64) @ self.r.astype(np.float64)).astype(np.float32)

    def model_shape(self):
        return (self.A.shape[1],)


def cgls(op, d, niter, callback=None):
    """niter CGLS iterations from m = 0.  Returns (m, history) with history a
    list of dicts {iter, rr (= ||r||^2), ss (= ||L^T r||^2), alpha, beta}.
    Stops early when L^T r or L p vanishes."""
    rr = op.start(d)
    s = op.adjoint()
    p = s.copy()
    gamma = dot64(s, s)
    m = np.zeros(op.model_shape(), np.float32)
    hist = [dict(iter=0, rr=rr, ss=gamma, alpha=0.0, beta=0.0)]
    for it in range(1, niter + 1):
        if gamma == 0.0:
            break
        qq = op.forward(p)
        if qq == 0.0:
            break
        alpha = gamma / qq
        m = m + np.float32(alpha) * p
        rr = op.update_residual(alpha)
        s = op.adjoint()
        gamma_new = dot64(s, s)
        beta = gamma_new / gamma
        p = s + np.float32(beta) * p
        gamma = gamma_new
        hist.append(dict(iter=it, rr=rr, ss=gamma, alpha=alpha, beta=beta))
        if callback is not None:
            callback(it, m, hist[-1])
    return m, hist
