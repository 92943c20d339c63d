"""Least-squares migration operators on the GPU: Born modelling (L) and its
adjoint (L^T) through the C ABI, against the oracle's FP32 Born / adjoint
engine on the same inputs.  The kernels evaluate the oracle's operation
sequence, so records, fields, the accumulator G and the gradient G/P must be
equal value for value.  Batches (uniform and ragged) must equal the shot-order
sum of single-shot runs, and the pair must pass the dot-product test."""
import numpy as np
import pytest

import cases
from cases import build_case, hip_context, oracle_engine

pytestmark = pytest.mark.gpu

CASES = {
    "ls_nfd4": dict(name="ls_nfd4"),
    "taylor_nfd6": dict(name="taylor_nfd6", ilste=1, nfdmax=6, icompen=0, seed=2),
    # source in the frame, receivers (ds = 3) reaching into the right frame
    "frame_src": dict(name="frame_src", src_depth=5.0, ds=3, s_l1=4, n=26, nfdmax=5, n2=8,
                      seed=3, mod_nx=72),
}


def _refl(cfg, seed=5):
    """Reflectivity on the interior image grid [mnz][mnx] and as a full grid."""
    n2 = cfg["n2"]
    mnz, mnx = cfg["nz"] - 2 * n2, cfg["nx"] - 2 * n2
    rng = np.random.default_rng(seed)
    img = (rng.standard_normal((mnz, mnx)) * 0.02).astype(np.float32)
    full = np.zeros((cfg["nz"], cfg["nx"]), np.float32)
    full[n2:-n2, n2:-n2] = img
    return img, full


def _resid(cfg, nshot=1, seed=7):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((nshot, cfg["n"], cfg["nt"])) * 1e-3).astype(np.float32)


@pytest.mark.parametrize("key", sorted(CASES))
def test_born_matches_oracle(key):
    from rtm_gpu_amd import hip
    case = build_case(**CASES[key])
    cfg = case["cfg"]
    img, full = _refl(cfg)
    print(f"born {key}: grid {cfg['nz']}x{cfg['nx']} nt={cfg['nt']}", flush=True)
    ctx = hip_context(case, nshot=1, record=1)
    ctx.set_reflectivity(img)
    eng = oracle_engine(case, record=1)
    eng.set_reflectivity(full)
    ctx.born_begin()
    eng.born_begin()
    k = 2
    for step in (60, 60, cfg["nt"]):
        ctx.born_advance(step)
        eng.born_advance(step)
        k = min(cfg["nt"], k + step)
        g, o = ctx.field(hip.FIELD_Q), eng.field(7)
        assert np.array_equal(g, o), f"dU after step {k}: max |diff| {np.max(np.abs(g - o))}"
        assert np.array_equal(ctx.field(hip.FIELD_U), eng.field(0)), f"background at {k}"
    rec = ctx.records()[0]
    assert np.array_equal(rec, eng.record())
    assert np.abs(rec).max() > 0


@pytest.mark.parametrize("key", sorted(CASES))
def test_adjoint_matches_oracle(key):
    from rtm_gpu_amd import hip
    case = build_case(**CASES[key])
    cfg = case["cfg"]
    nt, n2 = cfg["nt"], cfg["n2"]
    r = _resid(cfg)
    print(f"adjoint {key}: grid {cfg['nz']}x{cfg['nx']} nt={nt}", flush=True)
    ctx = hip_context(case, nshot=1)
    eng = oracle_engine(case)
    ctx.forward_begin()
    ctx.forward_advance(nt)
    eng.forward_begin()
    eng.forward_advance(nt)
    ctx.set_gathers(r)
    ctx.adjoint_begin()
    eng.adjoint_begin(r[0])
    assert np.array_equal(ctx.field(hip.FIELD_Q), eng.field(2)), "mu(nt-2)"
    for step in (70, nt):
        ctx.adjoint_advance(step)
        eng.adjoint_advance(step)
        assert np.array_equal(ctx.field(hip.FIELD_Q), eng.field(2)), "adjoint field"
        # the reconstructed background is defined on the interior
        inner = (slice(n2, -n2), slice(n2, -n2))
        assert np.array_equal(ctx.field(hip.FIELD_U)[inner], eng.field(0)[inner]), "background"
        assert np.array_equal(ctx.field(hip.FIELD_I1), eng.field(8)), "accumulator G"
    grad = ctx.adjoint_gradient()
    exp = eng.field(9)[n2:-n2, n2:-n2]
    assert np.array_equal(grad, exp)
    assert np.abs(grad).max() > 0


def test_born_then_adjoint_reuses_the_background():
    """A Born pass leaves the background end state and strips that a plain
    forward pass leaves, so L^T may follow L directly (one background pass per
    CG iteration)."""
    case = build_case(**CASES["ls_nfd4"])
    cfg = case["cfg"]
    img, _ = _refl(cfg)
    r = _resid(cfg)
    ctx = hip_context(case, nshot=1, record=1)
    ctx.set_reflectivity(img)
    ctx.born()
    g1 = ctx.adjoint(r)
    ctx.forward_begin()
    ctx.forward_advance(cfg["nt"])
    g2 = ctx.adjoint(r)
    assert np.array_equal(g1, g2)


def test_batch_equals_shot_order_sum():
    case = build_case(name="three", nshot=3, depth_step=25.0, seed=6)
    cfg = case["cfg"]
    n2 = cfg["n2"]
    img, full = _refl(cfg)
    r = _resid(cfg, nshot=3)
    ctx = hip_context(case, nshot=3, record=1)
    ctx.set_reflectivity(img)
    born = ctx.born()
    grad = ctx.adjoint(r)
    acc = None
    for s in range(3):
        e = oracle_engine(case, shot=s, record=1)
        e.set_reflectivity(full)
        e.born_begin()
        e.born_advance(cfg["nt"])
        assert np.array_equal(born[s], e.record()), f"shot {s} Born data"
        e.adjoint_begin(r[s])
        e.adjoint_advance(cfg["nt"])
        g = e.field(9)[n2:-n2, n2:-n2]
        acc = g if acc is None else acc + g
    assert np.array_equal(grad, acc)


def test_ragged_batch_equals_single_shot_runs():
    from rtm_gpu_amd import hip
    from test_gpu_apertures import MOD_NX, MOD_NZ, N2, _keys, _problem
    shots, _ = _problem()
    ns = len(shots)
    wmax = max(s["wd"] for s in shots)
    c0 = shots[0]["case"]
    base = {k: c0["cfg"][k] for k in _keys()}
    nxmax = wmax + 2 * N2
    nt = base["nt"]
    ctx = hip.Context(**base, nx=nxmax, s_l=N2, s_r=N2 + wmax - 1, n=wmax, max_shots=ns,
                      stack_nx=MOD_NX, record=1)
    m = c0["model"]
    vflat = np.full((base["nz"], nxmax), m["vmin"], np.float32)
    ctx.set_model(vflat, vflat * 0, m["w"], m["c"], m["index"], m["nvel"], m["vmin"], m["dv"])
    ctx.set_wavelet(m["wav"])
    ctx.set_shot_geometry([dict(nx=s["case"]["cfg"]["nx"], r_u=s["case"]["r_u"][0],
                                r_x=s["case"]["r_x"][0], s_l=s["case"]["cfg"]["s_l"],
                                s_r=s["case"]["cfg"]["s_r"], n=s["case"]["cfg"]["n"],
                                x_img=s["xb"]) for s in shots])
    for i, s in enumerate(shots):
        ctx.set_shot_model(i, s["case"]["model"]["v"], s["case"]["model"]["r1"])
    rng = np.random.default_rng(3)
    refl = (rng.standard_normal((MOD_NZ, MOD_NX)) * 0.02).astype(np.float32)
    ctx.set_reflectivity(refl)
    born = ctx.born()
    resid = np.zeros((ns, wmax, nt), np.float32)
    for i, s in enumerate(shots):
        resid[i, :s["wd"]] = rng.standard_normal((s["wd"], nt)).astype(np.float32) * 1e-3
    grad = ctx.adjoint(resid)
    exp = np.zeros((MOD_NZ, MOD_NX), np.float32)
    for i, s in enumerate(shots):
        one = cases.hip_context(s["case"], nshot=1, record=1)
        wd, xb = s["wd"], s["xb"]
        one.set_reflectivity(refl[:, xb:xb + wd])
        b1 = one.born()
        assert np.array_equal(born[i, :wd], b1[0]), f"shot {i} Born data"
        assert not born[i, wd:].any()
        g1 = one.adjoint(resid[i:i + 1, :wd])
        exp[:, xb:xb + wd] = exp[:, xb:xb + wd] + g1
    assert np.array_equal(grad, exp)


def test_dot_product_fp32():
    """<L m, r> = <m, L^T r> to the accuracy of the ABC, with one coefficient
    row (Taylor), in FP32 on the GPU."""
    case = build_case(name="bornT", ilste=1, nfdmax=4, nt=300, seed=3)
    cfg = case["cfg"]
    n2 = cfg["n2"]
    rng = np.random.default_rng(1)
    img = np.zeros((cfg["nz"] - 2 * n2, cfg["nx"] - 2 * n2), np.float32)
    img[8:-3, 5:-5] = rng.standard_normal(img[8:-3, 5:-5].shape).astype(np.float32) * 0.01
    r = rng.standard_normal((1, cfg["n"], cfg["nt"])).astype(np.float32)
    ctx = hip_context(case, nshot=1, record=1)
    ctx.set_reflectivity(img)
    b = ctx.born()
    g = ctx.adjoint(r)
    lhs = float(np.sum(b.astype(np.float64) * r))
    rhs = float(np.sum(img.astype(np.float64) * g))
    print(f"<Lm,r> {lhs:.6e} <m,L'r> {rhs:.6e} rel {abs(lhs - rhs) / abs(lhs):.3e}")
    assert abs(lhs - rhs) <= 0.06 * abs(lhs)


def test_state_errors():
    from rtm_gpu_amd import hip
    case = build_case(**CASES["ls_nfd4"])
    ctx = hip_context(case, nshot=1, record=1)
    with pytest.raises(hip.RtmError):
        ctx.born_begin()  # no reflectivity
    with pytest.raises(hip.RtmError):
        ctx.adjoint_begin()  # no forward pass
    ctx.set_reflectivity(_refl(case["cfg"])[0])
    ctx.born()
    with pytest.raises(hip.RtmError):
        ctx.adjoint_begin()  # no residuals
    ctx.set_gathers(_resid(case["cfg"]))
    ctx.adjoint_begin()
    ctx.adjoint_advance(case["cfg"]["nt"])
    ctx.adjoint_gradient()
    with pytest.raises(hip.RtmError):
        ctx.image_post(3000.0, 1e-4)  # the adjoint pass made no migration image
    with pytest.raises(hip.RtmError):
        ctx.adjoint_begin()  # the forward end state is consumed


def test_data_dot_and_axpy():
    from rtm_gpu_amd import hip
    case = build_case(name="three", nshot=3, depth_step=25.0, seed=6)
    cfg = case["cfg"]
    img, _ = _refl(cfg)
    ctx = hip_context(case, nshot=3, record=1)
    ctx.set_reflectivity(img)
    q = ctx.born()
    r = _resid(cfg, nshot=3, seed=9) * 100
    ctx.set_gathers(r)
    f64 = lambda a: a.astype(np.float64).ravel()  # noqa: E731
    for x, y, a, b in ((hip.DATA_GATHERS, hip.DATA_GATHERS, r, r),
                       (hip.DATA_RECORDS, hip.DATA_RECORDS, q, q),
                       (hip.DATA_GATHERS, hip.DATA_RECORDS, r, q)):
        got = ctx.data_dot(x, y)
        exp = float(np.dot(f64(a), f64(b)))
        assert abs(got - exp) <= 1e-12 * abs(exp), (x, y, got, exp)
        assert ctx.data_dot(x, y) == got  # fixed reduction order
    alpha = np.float32(0.37)
    ctx.data_axpy(-alpha)
    r2 = r + (-alpha) * q
    got = ctx.data_dot(hip.DATA_GATHERS, hip.DATA_GATHERS)
    exp = float(np.dot(f64(r2), f64(r2)))
    assert abs(got - exp) <= 1e-12 * abs(exp)


def test_cgls_on_the_gpu_follows_the_oracle():
    """Device-resident CGLS (Born / adjoint passes on the GPU, data-space
    products and updates in device memory) against the same CGLS driven by
    the oracle's FP32 operator: the passes agree bit for bit, the FP64 inner
    products to rounding, so the iterates agree to float32 rounding."""
    from scipy.ndimage import gaussian_filter

    from rtm_gpu_amd import lsrtm
    case = build_case(name="cg", ilste=1, nfdmax=4, nt=300, seed=3, mod_nz=44, mod_nx=60)
    ref = cases.OracleBornOp(case)
    rng = np.random.default_rng(2)
    mt = np.zeros(ref.shape, np.float32)
    mt[10:-4, 6:-6] = (gaussian_filter(rng.standard_normal((ref.shape[0] - 14, ref.shape[1] - 12)),
                                       1.0) * 0.05)
    d = ref.born(mt)
    ctx = hip_context(case, nshot=1, record=1)
    op = lsrtm.BornOperator(ctx)
    m_gpu, h_gpu = lsrtm.cgls(op, d[None], 4)
    m_ref, h_ref = lsrtm.cgls(ref, d, 4)
    for a, b in zip(h_gpu, h_ref):
        print(f"iter {a['iter']}: rr {a['rr']:.9e} / {b['rr']:.9e}")
        assert abs(a["rr"] - b["rr"]) <= 1e-5 * h_ref[0]["rr"]
        assert abs(a["ss"] - b["ss"]) <= 1e-4 * b["ss"]
    assert cases.rel_l2(m_gpu, m_ref) <= 1e-4
    rr = [h["rr"] for h in h_gpu]
    assert all(b < a for a, b in zip(rr, rr[1:]))
