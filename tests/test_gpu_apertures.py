"""Ragged batches (uneven apertures): shots with their own grid width,
receivers, source and model window run in the same launches
(rtm_set_shot_geometry + rtm_set_shot_model).

Every shot of the batch must equal a one-shot context run on its window
alone, bit for bit (and that one equals the FP32 replica); the stack must
equal the shot-order sum of the per-shot images at their columns of the
global image."""
import numpy as np
import pytest

import cases
import oracle_bind as ob

pytestmark = pytest.mark.gpu

MOD_NZ, MOD_NX, N2, NFD, NT = 40, 150, 10, 4, 240
H, TAO, F0 = 10.0, 0.001, 25.0
# (first model column, aperture width, source column within the window, source depth m)
WINDOWS = [(0, 60, 30, 30.0), (35, 47, 10, 50.0), (70, 80, 41, 30.0), (112, 38, 19, 70.0)]


def _problem():
    raw = cases.layered_model(MOD_NZ, MOD_NX, 4, 1500.0, 3000.0, seed=9, dip=0.2)  # [x][z]
    vg, _ = ob.pad_model(raw, MOD_NZ, MOD_NX, N2, 0, TAO, H)
    dv = np.float32(1.0)
    vmin, vmax, nvel = ob.vrange(vg, dv)
    need = ob.bins_needed(vg, vmin, dv, nvel)
    f32 = lambda x: float(np.float32(x))  # noqa: E731
    d0 = ob.derive(MOD_NZ, MOD_NX, N2, NT, TAO, TAO, H, H, F0, 1, 1, MOD_NX, 1, 1)
    M, index, c = ob.design_ls(1000, NFD, 2, nvel, f32(TAO), f32(H), f32(1.0), f32(1e-5),
                               f32(2 * F0 + 1), float(vmin), float(dv), need, float(d0["hzx"]))
    wav = ob.wavelet_table(d0["NT"], TAO, F0)
    w = ob.blend_weights(N2)
    shots = []
    for xb, wd, sx, depth in WINDOWS:
        d = ob.derive(MOD_NZ, wd, N2, NT, TAO, TAO, H, H, F0, 1, 1, wd, 1, sx + 1)
        v, r1 = ob.pad_model(np.ascontiguousarray(raw[xb:xb + wd]), MOD_NZ, wd, N2, 0, TAO, H)
        r_u = ob.lib().orc_source_row(np.float32(depth), np.float32(H), N2)
        cfg = dict(nz=d["NZ"], nx=d["NX"], n2=N2, nfdmax=NFD, nt=d["NT"], nt2=d["NT2"], ilste=0,
                   icompen=1, s_l=d["s_l"], s_r=d["s_r"], s_z=d["s_z"], ds=1, n=wd,
                   tao2=d["tao2"], h2=d["h2"], taoh=d["taoh"], taoh2=d["taoh2"],
                   hzx2_1=d["hzx2_1"], vmin=vmin, dv=dv)
        model = dict(v=v, r1=r1, w=w, c=c, index=index, nvel=nvel, vmin=vmin, vmax=vmax, dv=dv,
                     wav=wav)
        case = dict(cfg=cfg, model=model, r_u=[r_u], r_x=[d["r_x"]])
        shots.append(dict(xb=xb, wd=wd, case=case, seis=cases.synth_gather(case)))
    return shots, vmax


def _keys():
    return ("nz", "n2", "nfdmax", "nt", "nt2", "ilste", "icompen", "s_z", "ds", "tao2", "h2",
            "taoh", "taoh2", "hzx2_1")


def test_ragged_batch_equals_single_shot_runs():
    from rtm_gpu_amd import hip
    shots, vmax = _problem()
    ns = len(shots)
    wmax = max(s["wd"] for s in shots)
    c0 = shots[0]["case"]
    base = {k: c0["cfg"][k] for k in _keys()}
    nxmax = wmax + 2 * N2
    ctx = hip.Context(**base, nx=nxmax, s_l=N2, s_r=N2 + wmax - 1, n=wmax, max_shots=ns,
                      stack_nx=MOD_NX)
    m = c0["model"]
    # the shared model only fixes the coefficient table here; every shot gets a window
    vflat = np.full((base["nz"], nxmax), m["vmin"], np.float32)
    ctx.set_model(vflat, vflat * 0, m["w"], m["c"], m["index"], m["nvel"], m["vmin"], m["dv"])
    ctx.set_wavelet(m["wav"])
    geoms = []
    for s in shots:
        cf = s["case"]["cfg"]
        geoms.append(dict(nx=cf["nx"], r_u=s["case"]["r_u"][0], r_x=s["case"]["r_x"][0],
                          s_l=cf["s_l"], s_r=cf["s_r"], n=cf["n"], x_img=s["xb"]))
    ctx.set_shot_geometry(geoms)
    for i, s in enumerate(shots):
        ctx.set_shot_model(i, s["case"]["model"]["v"], s["case"]["model"]["r1"])
    seis = np.zeros((ns, wmax, base["nt"]), np.float32)
    for i, s in enumerate(shots):
        seis[i, :s["wd"]] = s["seis"]
    ctx.set_gathers(seis)
    ctx.migrate()
    up, down = ctx.image_post(vmax, 1e-4)
    img, ill = ctx.stack_finish(ns, 1)

    exp_s1 = np.zeros((MOD_NZ, MOD_NX), np.float32)
    exp_s2 = np.zeros((MOD_NZ, MOD_NX), np.float32)
    for i, s in enumerate(shots):
        case = s["case"]
        nx_s = case["cfg"]["nx"]
        one = cases.hip_context(case, nshot=1)
        one.set_gathers(s["seis"][None])
        one.migrate()
        for which in (hip.FIELD_I1, hip.FIELD_I2, hip.FIELD_Q):
            a = ctx.field(which, shot=i)
            b = one.field(which)
            assert np.array_equal(a[:, :nx_s], b[:, :nx_s]), f"shot {i} field {which}"
            assert not a[:, nx_s:].any()
        u1, d1 = one.image_post(vmax, 1e-4)
        mnx = s["wd"]
        assert np.array_equal(up[i][:, :mnx], u1[0][:, :mnx]), f"shot {i} up"
        assert np.array_equal(down[i][:, :mnx], d1[0][:, :mnx]), f"shot {i} down"
        xb = s["xb"]
        exp_s1[:, xb:xb + mnx] = exp_s1[:, xb:xb + mnx] + u1[0][:, :mnx]
        exp_s2[:, xb:xb + mnx] = exp_s2[:, xb:xb + mnx] + d1[0][:, :mnx]
        if i == 0:  # and the single-shot run is the FP32 replica
            eng = cases.oracle_engine(case)
            eng.forward_begin()
            eng.forward_advance(case["cfg"]["nt"])
            eng.backward_begin(s["seis"])
            eng.backward_advance(case["cfg"]["nt"])
            assert np.array_equal(one.field(hip.FIELD_I1)[:, :nx_s], eng.field(3))
    a = exp_s1 / np.float32(ns)
    b = exp_s2 / np.float32(ns)
    covered = (exp_s1 != 0) | (exp_s2 != 0)
    exp_img = np.where(covered, a / np.where(covered, b, 1), np.float32(0))
    assert np.array_equal(ill, b)
    assert np.array_equal(img, exp_img.astype(np.float32))
    assert np.isfinite(img).all()


def test_new_geometry_drops_model_windows():
    # A second rtm_set_shot_geometry must not leave the previous batch's model
    # windows in force: the next pass refuses to begin until each shot of the
    # new batch has its window again.
    from rtm_gpu_amd import hip
    shots, _ = _problem()
    wmax = max(s["wd"] for s in shots)
    c0 = shots[0]["case"]
    base = {k: c0["cfg"][k] for k in _keys()}
    nxmax = wmax + 2 * N2
    ctx = hip.Context(**base, nx=nxmax, s_l=N2, s_r=N2 + wmax - 1, n=wmax, max_shots=len(shots),
                      stack_nx=MOD_NX)
    m = c0["model"]
    vflat = np.full((base["nz"], nxmax), m["vmin"], np.float32)
    ctx.set_model(vflat, vflat * 0, m["w"], m["c"], m["index"], m["nvel"], m["vmin"], m["dv"])
    ctx.set_wavelet(m["wav"])
    geoms = []
    for s in shots:
        cf = s["case"]["cfg"]
        geoms.append(dict(nx=cf["nx"], r_u=s["case"]["r_u"][0], r_x=s["case"]["r_x"][0],
                          s_l=cf["s_l"], s_r=cf["s_r"], n=cf["n"], x_img=s["xb"]))
    ctx.set_shot_geometry(geoms)
    for i, s in enumerate(shots):
        ctx.set_shot_model(i, s["case"]["model"]["v"], s["case"]["model"]["r1"])
    ctx.forward_begin()
    ctx.set_shot_geometry(geoms[2:])
    with pytest.raises(hip.RtmError, match=r"\(-4\).*no model window"):
        ctx.forward_begin()
    for i, s in enumerate(shots[2:]):
        ctx.set_shot_model(i, s["case"]["model"]["v"], s["case"]["model"]["r1"])
    ctx.forward_begin()
    ctx.close()


def test_shared_model_windows():
    """Shots whose model windows are equal share one stored window: a batch
    of six shots over two windows (shots 0, 2, 4 on window A and 1, 3, 5 on
    window B, different sources) equals its shots' one-shot runs through a
    Born / adjoint pair with LS rows (the transposed stencil
    reads the shared window's tap planes); re-pointing shot 5 at window A
    then gives shot 5 the results of a one-shot run on A."""
    from rtm_gpu_amd import hip
    shots, _ = _problem()
    A, B = shots[0], shots[2]  # widths 60 and 80
    order = [A, B, A, B, A, B]
    depths = [10, 12, 15, 17, 20, 22]  # source rows
    wmax = max(s["wd"] for s in order)
    base = {k: A["case"]["cfg"][k] for k in _keys()}
    nxmax = wmax + 2 * N2
    nt = base["nt"]

    def batch_ctx():
        ctx = hip.Context(**base, nx=nxmax, s_l=N2, s_r=N2 + wmax - 1, n=wmax, max_shots=6,
                          stack_nx=MOD_NX, record=1)
        m = A["case"]["model"]
        vflat = np.full((base["nz"], nxmax), m["vmin"], np.float32)
        ctx.set_model(vflat, vflat * 0, m["w"], m["c"], m["index"], m["nvel"], m["vmin"], m["dv"])
        ctx.set_wavelet(m["wav"])
        ctx.set_shot_geometry([dict(nx=s["case"]["cfg"]["nx"], r_u=depths[i],
                                    r_x=s["case"]["r_x"][0], s_l=s["case"]["cfg"]["s_l"],
                                    s_r=s["case"]["cfg"]["s_r"], n=s["case"]["cfg"]["n"],
                                    x_img=s["xb"]) for i, s in enumerate(order)])
        for i, s in enumerate(order):
            ctx.set_shot_model(i, s["case"]["model"]["v"], s["case"]["model"]["r1"])
        return ctx

    def one(s, depth):
        case = dict(s["case"])
        case["r_u"] = [depth]
        return cases.hip_context(case, nshot=1, record=1)

    ctx = batch_ctx()
    rng = np.random.default_rng(5)
    seis = np.zeros((6, wmax, nt), np.float32)
    for i, s in enumerate(order):
        seis[i, :s["wd"]] = rng.standard_normal((s["wd"], nt)).astype(np.float32) * 1e-3
    refl = (rng.standard_normal((MOD_NZ, MOD_NX)) * 0.02).astype(np.float32)
    ctx.set_reflectivity(refl)
    born = ctx.born()
    grad = ctx.adjoint(seis)
    exp = np.zeros((MOD_NZ, MOD_NX), np.float32)
    for i, s in enumerate(order):
        wd, xb = s["wd"], s["xb"]
        o = one(s, depths[i])
        o.set_reflectivity(refl[:, xb:xb + wd])
        assert np.array_equal(born[i, :wd], o.born()[0]), f"shot {i} Born data"
        exp[:, xb:xb + wd] = exp[:, xb:xb + wd] + o.adjoint(seis[i:i + 1, :wd])
        o.close()
    assert np.array_equal(grad, exp)
    # shot 5 leaves window B for a window of its own (B 1% faster), then
    # returns: the other shots keep their results throughout
    mB = order[5]["case"]["model"]
    ctx.set_shot_model(5, mB["v"] * np.float32(1.01), mB["r1"])
    born2 = ctx.born()
    assert np.array_equal(born2[:5], born[:5])
    assert not np.array_equal(born2[5], born[5])
    ctx.set_shot_model(5, mB["v"], mB["r1"])
    assert np.array_equal(ctx.born(), born)
    ctx.close()
