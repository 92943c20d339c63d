"""The conjugate-gradient solver of least-squares migration (host logic of
rtm_gpu_amd.lsrtm), without a GPU: on a dense matrix against the textbook
least-squares solution, and on the oracle's Born operator (FP32) for the
monotone decrease of the misfit that CGLS guarantees."""
import numpy as np
from scipy.ndimage import gaussian_filter

import cases
from rtm_gpu_amd import lsrtm


def test_cgls_solves_dense_least_squares():
    rng = np.random.default_rng(4)
    u, _ = np.linalg.qr(rng.standard_normal((80, 30)))
    v, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    A = (u * np.linspace(1.0, 4.0, 30)) @ v.T  # condition number 4
    d = rng.standard_normal(80).astype(np.float32)
    m, hist = lsrtm.cgls(lsrtm.MatrixOperator(A), d, 30)
    exact = np.linalg.lstsq(A.astype(np.float32).astype(np.float64), d.astype(np.float64),
                            rcond=None)[0]
    assert np.linalg.norm(m - exact) <= 1e-4 * np.linalg.norm(exact)
    rr = [h["rr"] for h in hist]
    assert all(b <= a * (1 + 1e-6) for a, b in zip(rr, rr[1:]))
    # the minimum misfit is the part of d outside the range of A
    res = d - A.astype(np.float32) @ exact.astype(np.float32)
    assert abs(rr[-1] - float(res.astype(np.float64) @ res)) <= 1e-4 * rr[0]


def test_cgls_stops_on_zero_data():
    A = np.eye(5, 3)
    m, hist = lsrtm.cgls(lsrtm.MatrixOperator(A), np.zeros(5, np.float32), 4)
    assert not m.any() and len(hist) == 1


def test_cgls_on_the_born_operator_decreases_the_misfit():
    case = cases.build_case(name="cg", ilste=1, nfdmax=4, nt=300, seed=3, mod_nz=44, mod_nx=60)
    op = cases.OracleBornOp(case)
    rng = np.random.default_rng(2)
    mt = np.zeros(op.shape, np.float32)
    mt[10:-4, 6:-6] = (gaussian_filter(rng.standard_normal((op.shape[0] - 14, op.shape[1] - 12)),
                                       1.0) * 0.05)
    d = op.born(mt)
    m, hist = lsrtm.cgls(op, d, 4)
    rr = [h["rr"] / hist[0]["rr"] for h in hist]
    assert all(b < a for a, b in zip(rr, rr[1:])), rr
    assert rr[-1] < 0.8, rr
    # the first step is the scaled migration image, which correlates with the model
    assert float(np.sum(m * mt)) > 0


class _Blocks:
    """G independent dense problems with the grouped operator interface."""

    def __init__(self, As):
        self.ops = [lsrtm.MatrixOperator(A) for A in As]
        self.groups = len(As)

    def start(self, d):
        return np.array([o.start(di) for o, di in zip(self.ops, d)])

    def forward(self, p):
        return np.array([o.forward(pi) for o, pi in zip(self.ops, p)])

    def update_residual(self, alpha):
        return np.array([o.update_residual(a) for o, a in zip(self.ops, alpha)])

    def adjoint(self):
        return np.stack([o.adjoint() for o in self.ops])

    def model_shape(self):
        return (self.groups, self.ops[0].A.shape[1])


def test_grouped_cgls_equals_independent_runs():
    rng = np.random.default_rng(7)
    As = [rng.standard_normal((40, 12)) * (1 + g) for g in range(3)]
    ds = [rng.standard_normal(40).astype(np.float32) for _ in range(3)]
    ds[2][:] = 0.0  # a group with nothing to fit stays at zero
    m, hist = lsrtm.cgls(_Blocks(As), np.stack(ds), 6)
    for g in range(3):
        mg, hg = lsrtm.cgls(lsrtm.MatrixOperator(As[g]), ds[g], 6)
        assert np.array_equal(m[g], mg), g
        for a, b in zip(hist, hg):
            assert a["rr"][g] == b["rr"]
    assert not m[2].any()


def test_deblur_helpers():
    """Scattering strength, network input channels and seeded training
    reflectivities of the deblurring pipeline (host code)."""
    from rtm_gpu_amd import deblur
    raw = np.array([[2000.0, 2500.0], [3000.0, 1500.0]], np.float32)  # [x][z]
    sm = np.full_like(raw, 2000.0)
    m = deblur.reflectivity(raw, sm)
    assert m.shape == (2, 2)  # [z][x]
    assert np.allclose(m, [[0.0, 1.25], [0.5625, -0.4375]])
    S = deblur.build_survey(mod_nz=24, mod_nx=40, nt=200, nshot=2, nfdmax=4, n2=10)
    a = deblur.random_reflectivity(S, 7)
    assert a.shape == (24, 40) and np.array_equal(a, deblur.random_reflectivity(S, 7))
    assert not np.array_equal(a, deblur.random_reflectivity(S, 8))
    illum = np.outer(np.linspace(1.0, 1e-8, 24), np.ones(40))
    x = deblur.features(S, np.ones((3, 24, 40)), 0.5, illum)
    assert x.shape == (3, 3, 24, 40)
    assert np.all(x[:, 0] == 0.5)
    assert x[0, 1].max() == 1.0 and x[0, 1].min() == 0.0  # log illumination, clipped at 1e-6
    assert x[0, 2, 0, 0] == 0.0 and x[0, 2, -1, 0] == 1.0  # depth
    assert S["r_x"] == sorted(S["r_x"]) and len(set(S["r_x"])) == 2


def test_deblur_patches():
    from rtm_gpu_amd import deblur
    rng = np.random.default_rng(0)
    XF = rng.standard_normal((3, 2, 20, 30)).astype(np.float32)
    M = rng.standard_normal((3, 20, 30)).astype(np.float32)
    xs, ms = deblur.patches(XF, M, 8, 50, np.random.default_rng(1))
    assert xs.shape == (50, 2, 8, 8) and ms.shape == (50, 8, 8)
    # every crop is an aligned window of one pair, inputs and target together
    for i in range(50):
        hits = [(k, z, x) for k in range(3) for z in range(13) for x in range(23)
                if np.array_equal(ms[i], M[k, z:z + 8, x:x + 8])]
        assert len(hits) == 1
        k, z, x = hits[0]
        assert np.array_equal(xs[i], XF[k, :, z:z + 8, x:x + 8])
