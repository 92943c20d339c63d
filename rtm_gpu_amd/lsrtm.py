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
    with the data-space vectors resident in the context.

    groups > 1 batches independent problems: the batch's shots split into
    `groups` equal runs, run g imaging columns [g*w, (g+1)*w) of the stacked
    image (w = stack width / groups), and every inner product and update is
    taken per group, so each problem keeps its own step lengths.  Model
    vectors are then [groups][mnz][w]."""

    def __init__(self, ctx, groups=1):
        if not ctx.p.record:
            raise hip.RtmError("the Born operator needs a context created with record=1")
        if groups < 1 or ctx.nshot % groups or ctx.gw % groups:
            raise hip.RtmError(f"{ctx.nshot} shots / {ctx.gw} columns do not split into {groups}")
        self.ctx = ctx
        self.groups = groups
        self.per = ctx.nshot // groups
        self._background = False  # a forward / Born pass has left its end state and strips

    def dots(self, x, y):
        """<x, y> of two data-space vectors (hip.DATA_GATHERS / DATA_RECORDS),
        per group (float64; an array when groups > 1)."""
        if self.groups == 1:
            return self.ctx.data_dot(x, y)
        return np.array([self.ctx.data_dot(x, y, g * self.per, self.per)
                         for g in range(self.groups)])

    def start(self, d):
        """r = d (observed data [shot][n][nt]; None: the context's gathers
        already hold it); returns <d, d>."""
        if d is not None:
            self.ctx.set_gathers(d)
        return self.dots(hip.DATA_GATHERS, hip.DATA_GATHERS)

    def forward(self, p):
        """q = L p into the records; returns <q, q>."""
        ctx = self.ctx
        p = np.asarray(p, np.float32)
        if self.groups > 1:
            G, mnz, w = p.shape
            p = p.transpose(1, 0, 2).reshape(mnz, G * w)
        ctx.set_reflectivity(p)
        ctx.born_begin()
        ctx.born_advance(ctx.p.nt)
        self._background = True
        return self.dots(hip.DATA_RECORDS, hip.DATA_RECORDS)

    def update_residual(self, alpha):
        """r -= alpha q; returns <r, r>."""
        if self.groups == 1:
            self.ctx.data_axpy(-np.float32(alpha))
        else:
            for g in range(self.groups):
                self.ctx.data_axpy(-np.float32(alpha[g]), g * self.per, self.per)
        return self.dots(hip.DATA_GATHERS, hip.DATA_GATHERS)

    def adjoint(self):
        """s = L^T r (the stacked gradient)."""
        ctx = self.ctx
        if not self._background:
            ctx.forward_begin(True)
            ctx.forward_advance(ctx.p.nt)
        ctx.adjoint_begin()
        ctx.adjoint_advance(ctx.p.nt)
        self._background = False
        g = ctx.adjoint_gradient()
        if self.groups > 1:
            mnz = g.shape[0]
            g = np.ascontiguousarray(g.reshape(mnz, self.groups, -1).transpose(1, 0, 2))
        return g

    def model_shape(self):
        mnz = self.ctx.nz - 2 * self.ctx.n2
        if self.groups == 1:
            return (mnz, self.ctx.gw)
        return (self.groups, mnz, self.ctx.gw // self.groups)


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
    An operator with groups > 1 runs one CGLS per group (scalars become
    per-group arrays); a group whose L^T r or L p vanishes stops moving."""
    G = getattr(op, "groups", 1)

    def gdot(a, b):
        a2 = np.asarray(a).reshape(G, -1)
        b2 = np.asarray(b).reshape(G, -1)
        return np.array([dot64(a2[g], b2[g]) for g in range(G)])

    def scale(v, coef):  # float32 per-group multiple of a model vector
        c = np.asarray(coef, np.float64).astype(np.float32).reshape((G,) + (1,) * (v.ndim - 1))
        return c * v if G > 1 else np.float32(coef[0]) * v

    def scalar(v):
        return v if G > 1 else float(v[0])

    rr = np.atleast_1d(op.start(d))
    s = op.adjoint()
    p = s.copy()
    gamma = gdot(s, s)
    m = np.zeros(op.model_shape(), np.float32)
    hist = [dict(iter=0, rr=scalar(rr), ss=scalar(gamma), alpha=scalar(np.zeros(G)),
                 beta=scalar(np.zeros(G)))]
    for it in range(1, niter + 1):
        if not (gamma > 0.0).any():
            break
        qq = np.atleast_1d(op.forward(p))
        ok = (gamma > 0.0) & (qq > 0.0)
        if not ok.any():
            break
        alpha = np.where(ok, gamma / np.where(ok, qq, 1.0), 0.0)
        m = m + scale(p, alpha)
        rr = np.atleast_1d(op.update_residual(scalar(alpha)))
        s = op.adjoint()
        gamma_new = gdot(s, s)
        beta = np.where(gamma > 0.0, gamma_new / np.where(gamma > 0.0, gamma, 1.0), 0.0)
        p = s + scale(p, beta)
        gamma = gamma_new
        hist.append(dict(iter=it, rr=scalar(rr), ss=scalar(gamma), alpha=scalar(alpha),
                         beta=scalar(beta)))
        if callback is not None:
            callback(it, m, hist[-1])
    return m, hist
