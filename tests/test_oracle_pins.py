"""Pins of the oracle (CPU only).

Each oracle function is checked against something other than its own
formula: the reference's own helper sources compiled in oracle/_ref
(bit-exact), closed forms, known answers of the reference's integer
arithmetic, and physical invariants of the wave engine chosen so that a
plausible mistake (dropped term, wrong sign or index, transposed operand)
fails one of them."""
import math
import os
import tempfile

import numpy as np
import pytest

import cases
import oracle_bind as ob

needs_ref = pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built "
                               "(reference tree absent)")


# ---- integer and scalar setup (kernel.cu:602-628, 704-742, 794-795) --------
def test_nt2_known_answers():
    # FP32 product f0*tao (kernel.cu:624): 80 at 25 Hz / 1 ms, 100 at 20 Hz / 1 ms,
    # not the 81 / 101 of double arithmetic.
    d1 = ob.derive(200, 400, 20, 2000, 0.001, 0.001, 10.0, 10.0, 25.0, 1, 1, 400, 1, 201)
    d2 = ob.derive(351, 1701, 30, 4000, 0.001, 0.001, 12.5, 12.5, 20.0, 1, 1, 1701, 1, 851)
    assert d1["NT2"] == 80 and d2["NT2"] == 100
    assert int(2.0 / (25.0 * 0.001)) + 1 == 81  # what double arithmetic would give
    assert (d1["NZ"], d1["NX"], d1["NT"]) == (240, 440, 2000)
    assert (d2["NZ"], d2["NX"], d2["NT"]) == (411, 1761, 4000)
    assert (d2["s_l"], d2["s_r"], d2["s_z"], d2["r_x"]) == (30, 1730, 30, 880)
    # NT from a resampled record length (kernel.cu:613)
    d3 = ob.derive(10, 10, 5, 1001, 0.002, 0.001, 10.0, 10.0, 25.0, 1, 1, 10, 1, 5)
    assert d3["NT"] == 2001


def test_source_rows_cfg2():
    # 128 sources at depths 6.25*(m+1) m, hz = 12.5: 65 distinct rows, the
    # first at N2-1 inside the ABC frame (kernel.cu:794-795).
    rows = [ob.lib().orc_source_row(np.float32(6.25 * (m + 1)), np.float32(12.5), 30)
            for m in range(128)]
    assert len(set(rows)) == 65
    assert rows[0] == 29
    assert rows[-1] == int(800 / 12.5 + 30 - 1)


def test_vrange_snapping():
    v = np.array([1500.3, 2999.9, 2000.0], np.float32)
    vmin, vmax, nvel = ob.vrange(v, np.float32(1.0))
    assert (vmin, vmax, nvel) == (1500.0, 3000.0, 1501)
    v = np.array([1500.0, 3000.0], np.float32)
    vmin, vmax, nvel = ob.vrange(v, np.float32(2.0))
    assert (vmin, vmax, nvel) == (1500.0, 3000.0, 751)
    need = ob.bins_needed(np.array([1500.0, 1500.6, 1502.0], np.float32), np.float32(1500.0),
                          np.float32(1.0), 5)
    assert need.tolist() == [1, 1, 1, 0, 0]


def test_ricker_closed_form():
    for f0 in (15.0, 20.0, 25.0):
        for t in (0.0, 0.013, 1 / f0, 0.05, 0.2):
            a = math.pi * f0 * (np.float32(t) - np.float32(1 / np.float32(f0)))
            ref = (1 - 2 * a * a) * math.exp(-a * a)
            assert abs(ob.ricker(t, f0) - ref) < 1e-6
    assert ob.ricker(1 / 25.0, 25.0) == pytest.approx(1.0, abs=1e-6)  # peak at t0 = 1/f0


def test_blend_weights():
    w = ob.blend_weights(20)
    assert w[0] == 0.0 and w[20] == 1.0 and w[10] == np.float32(0.5)
    assert np.all(np.diff(w) > 0)


# ---- Taylor coefficients (LSMOrCon_rec_2D.cpp:526-551) ----------------------
@pytest.mark.parametrize("M", [1, 2, 4, 6, 8, 10, 12])
def test_taylor_closed_form(M):
    c = ob.taylor(M)
    f = math.factorial
    for l in range(1, M + 1):
        exact = 2 * (-1) ** (l + 1) * f(M) ** 2 / (l * l * f(M - l) * f(M + l))
        assert abs(c[l] - exact) <= 2e-7 * max(1.0, abs(exact))
    assert abs(c[0] + 2 * sum(float(x) for x in c[1:])) < 1e-5
    # second-derivative consistency: sum_l c_l l^2 = 1 (c0 + 2 sum c_l = 0)
    assert abs(sum(float(c[l]) * l * l for l in range(1, M + 1)) - 1.0) < 1e-5


@needs_ref
@pytest.mark.parametrize("M", [1, 4, 7, 10])
def test_taylor_matches_reference(M):
    c = np.zeros(M + 1, np.float32)
    ob.ref().ref_taylor(M, c)
    assert np.array_equal(ob.taylor(M).view(np.uint32), c.view(np.uint32))


# ---- LS coefficient design (LSMOrCon_rec_2D.cpp:22-524) ---------------------
@needs_ref
@pytest.mark.parametrize("M,r,hzx", [(2, 0.2, 1.0), (5, 0.35, 1.0), (10, 0.44, 1.0), (6, 0.3, 2.0),
                                     (8, 0.25, 0.5)])
def test_lscoef_matches_reference(M, r, hzx):
    bb = 2 * 3.1415926535898 * 41.0 * 0.001 / r
    a = np.zeros(M + 1)
    b = np.zeros(M + 1)
    ob.lib().orc_lscoef(r, bb, M, hzx, a)
    ob.ref().ref_lscoef(r, bb, M, hzx, b)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert abs(a[0] + 2 * a[1:].sum()) < 1e-12  # c0 = -2 sum c_l


@needs_ref
@pytest.mark.parametrize("vel,nfdmin,nfdmax,h,tao,fmax,hzx", [
    (1500.0, 2, 10, 12.5, 0.001, 41.0, 1.0), (5500.0, 2, 10, 12.5, 0.001, 41.0, 1.0),
    (2000.0, 2, 4, 10.0, 0.001, 51.0, 1.0), (3000.0, 3, 8, 10.0, 0.0008, 31.0, 2.0)])
def test_lenfd_matches_reference(vel, nfdmin, nfdmax, h, tao, fmax, hzx):
    nfre = int(fmax / 1.0) + 1
    a = ob.lib().orc_lenfd(nfre, nfdmin, nfdmax, vel, tao, h, 1.0, 1000, 1e-5, hzx)
    b = ob.ref().ref_lenfd(nfre, nfdmin, nfdmax, vel, tao, h, 1.0, 1000, 1e-5, hzx)
    assert a == b


def _design_both(vmin, vmax, nfdmax, nfdmin, h, tao, fmax, hzx, need_every):
    nvel = int((vmax - vmin) / 1.0 + 1.5)
    need = np.zeros(nvel, np.int32)
    need[::need_every] = 1
    need[-1] = 1
    a = ob.design_ls(1000, nfdmax, nfdmin, nvel, tao, h, 1.0, 1e-5, fmax, vmin, 1.0, need, hzx)
    b = ob.ref_design_ls(1000, nfdmax, nfdmin, nvel, tao, h, 1.0, 1e-5, fmax, vmin, vmax, 1.0,
                         need, hzx)
    return a, b


@needs_ref
@pytest.mark.parametrize("vmin,vmax,nfdmax,h,tao,fmax,hzx,every", [
    (1500.0, 4500.0, 4, 10.0, 0.001, 51.0, 1.0, 97),     # cfg1-like range
    (1500.0, 5500.0, 10, 12.5, 0.001, 41.0, 1.0, 211),   # cfg2 range
    (1800.0, 2600.0, 6, 10.0, 0.0008, 31.0, 2.0, 13)])
def test_ls_design_matches_reference(vmin, vmax, nfdmax, h, tao, fmax, hzx, every):
    (M, I, c), (Mr, Ir, cr) = _design_both(vmin, vmax, nfdmax, 2, h, tao, fmax, hzx, every)
    assert np.array_equal(M, Mr)
    assert np.array_equal(I, Ir)
    assert np.array_equal(c.view(np.uint32), cr.view(np.uint32))
    assert len(set(M[M > 0].tolist())) >= 1


def test_ls_design_meets_dispersion_bound():
    # For directly evaluated velocities the chosen length meets eps on the
    # (f, theta) grid and one tap less does not (LSMOrCon_rec_2D.cpp:317-343).
    h, tao, fmax, hzx, eps = 12.5, 0.001, 41.0, 1.0, 1e-5
    nfre = int(fmax) + 1
    for vel in (1500.0, 2500.0, 4000.0):
        M = ob.lib().orc_lenfd(nfre, 2, 10, vel, tao, h, 1.0, 1000, eps, hzx)
        r = vel * tao / h
        b = 2 * 3.1415926535898 * 1.0 * (nfre - 1) * tao / r
        c = np.zeros(M + 1)
        ob.lib().orc_lscoef(r, b, M, hzx, c)
        assert ob.lib().orc_disp_maxerr(c, M, vel, tao, h, 1.0, nfre, 1000, hzx) <= eps
        if M > 2:
            c2 = np.zeros(M)
            ob.lib().orc_lscoef(r, b, M - 1, hzx, c2)
            assert ob.lib().orc_disp_maxerr(c2, M - 1, vel, tao, h, 1.0, nfre, 1000, hzx) > eps


# ---- model padding (GPU_velocity_real.cpp:6-118) -----------------------------
@needs_ref
@pytest.mark.parametrize("ifv", [0, 1])
def test_pad_model_matches_reference(ifv):
    mod_nz, mod_nx, n2 = 17, 23, 6
    raw = cases.layered_model(mod_nz, mod_nx, 3, 1500, 3500, seed=7, dip=0.3)
    raw = raw + np.random.default_rng(1).uniform(0, 50, raw.shape).astype(np.float32)
    v, r1 = ob.pad_model(raw, mod_nz, mod_nx, n2, ifv, 0.001, 10.0)
    nz, nx = mod_nz + 2 * n2, mod_nx + 2 * n2
    vr = np.zeros(nz * nx, np.float32)
    rr = np.zeros(nz * nx, np.float32)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "v.dat")
        raw.astype(np.float32).tofile(p)
        ob.ref().ref_velocity(p.encode(), nz, nx, n2, np.float32(0.001), np.float32(10.0), ifv,
                              vr, rr)
    assert np.array_equal(v.ravel(), vr) and np.array_equal(r1.ravel(), rr)


# ---- wave engine invariants ---------------------------------------------------
def _run(case, prec=32, scale=1.0, nsteps=None, record=0):
    c = dict(case)
    if scale != 1.0:
        m = dict(case["model"])
        m["wav"] = (m["wav"] * np.float32(scale)).astype(np.float32)
        c["model"] = m
    e = cases.oracle_engine(c, prec=prec, record=record)
    e.forward_begin()
    e.forward_advance(nsteps or case["cfg"]["nt"])
    return e


def test_linearity_superposition():
    # The update is linear in the field: in FP64 the response to the sum of
    # two wavelets equals the sum of the responses to rounding level.  (Exact
    # power-of-two scaling does not hold in FP32: the numerical precursor
    # ahead of the wavefront is subnormal, where scaling by 2 does not commute
    # with rounding, and the runs then differ by ordinary rounding noise.)
    case = cases.build_case(name="lin")
    m = case["model"]
    # two wavelets with disjoint support, so their float sum is exact
    first = m["wav"].copy()
    first[30:] = 0
    alt = m["wav"] - first

    def run_wav(wav, prec):
        c = dict(case)
        mm = dict(m)
        mm["wav"] = wav.astype(np.float32)
        c["model"] = mm
        e = cases.oracle_engine(c, prec=prec)
        e.forward_begin()
        e.forward_advance(case["cfg"]["nt"])
        return e.field(0, f64=True)

    a = run_wav(first, 64)
    b = run_wav(alt, 64)
    ab = run_wav(m["wav"], 64)
    assert np.array_equal(first + alt, m["wav"])
    assert cases.rel_l2(ab, a + b) < 1e-12
    assert cases.rel_l2(a, ab) > 0.1 and cases.rel_l2(b, ab) > 0.1
    a32 = _run(case).field(0)
    b32 = _run(case, scale=2.0).field(0)
    assert cases.rel_l2(b32, 2 * a32) < 1e-6


def test_mirror_symmetry_fp64():
    # Flat layers and a source on the central column: u(z, x) = u(z, NX-1-x)
    # up to rounding (the left/right ABC velocity quirk is itself symmetric).
    case = cases.build_case(name="sym", mod_nx=61, src_x1=31, dip=0.0, nt=300)
    u = _run(case, prec=64).field(0, f64=True)
    assert np.max(np.abs(u - u[:, ::-1])) <= 1e-11 * np.max(np.abs(u))
    # FP32: the per-tap sum fma(a, h, u[x-l]) + u[x+l] is not mirror-symmetric
    # in its rounding, so the asymmetry is bounded by the FP32 rounding error.
    u32 = _run(case).field(0)
    e32 = cases.rel_l2(u32, u)
    assert cases.rel_l2(u32, u32[:, ::-1]) <= 4 * e32 + 1e-7


def _green_2d(r, c, t, dt, h, f0):
    """Analytic response of u_tt = c^2 lap u + f(t) h^2/dt^2 delta(x) with the
    truncated Ricker f.  With G = H(ct-r)/(2 pi c sqrt(c^2 t^2 - r^2)) and
    s = a cosh(th), a = r/c: (h^2/dt^2)/(2 pi c^2) * int_0^acosh(t/a) f(t - a cosh th) dth."""
    a = r / c
    out = np.zeros_like(t)
    for i, ti in enumerate(t):
        if ti <= a:
            continue
        th = np.linspace(0, math.acosh(ti / a), 4000)
        tt = ti - a * np.cosh(th)
        x = math.pi * f0 * (tt - 1 / f0)
        f = (1 - 2 * x * x) * np.exp(-x * x) * ((tt >= 0) & (tt < 2 / f0 + dt))
        out[i] = np.trapezoid(f, th)
    return out * h * h / (dt * dt) / (2 * math.pi * c * c)


def test_homogeneous_green_function():
    # Direct wave in a homogeneous medium against the 2-D Green's function:
    # pins the v^2 tau^2/h^2 scaling, the stencil signs and the source timing.
    n2, nfd, h, dt, f0, c = 20, 8, 5.0, 0.0005, 25.0, 2000.0
    mod = 241
    rcv_off = 50
    case = cases.build_case(name="green", mod_nz=mod, mod_nx=mod, n2=n2, nfdmax=nfd, ilste=1,
                            nt=520, h=h, hz=h, tao=dt, f0=f0, nlayer=2, vlo=c, vhi=c,
                            src_x1=121, src_depth=120 * h, s_z1=120, s_l1=121 + rcv_off, ds=1,
                            n=1, dip=0.0)
    assert case["r_u"][0] == n2 + 119 and case["cfg"]["s_z"] == case["r_u"][0]
    e = _run(case, prec=64, record=1)
    tr = e.record()[0].astype(np.float64)
    t = np.arange(case["cfg"]["nt"]) * dt
    g = _green_2d(rcv_off * h, c, t, dt, h, f0)
    win = t < 0.23  # before any boundary reflection reaches the receiver
    assert np.argmax(np.abs(tr[win])) == pytest.approx(np.argmax(np.abs(g[win])), abs=3)
    assert cases.rel_l2(tr[win], g[win]) < 0.05


def test_absorbing_frame_small_reflections():
    # The hybrid ABC must absorb.  Reference: the same source in a grid large
    # enough that nothing comes back within the time window.  After the wave
    # has crossed the small grid's frame (step 550), the small grid's interior
    # must match the large grid to a small fraction of the incident wave,
    # while a plain reflecting frame (blend weights w = 0) leaves O(1) energy.
    kw = dict(nfdmax=4, ilste=1, nt=550, nlayer=2, vlo=2000.0, vhi=2000.0, dip=0.0, h=10.0,
              hz=10.0, f0=25.0)
    n2 = 20
    small = cases.build_case(name="abc_small", mod_nz=60, mod_nx=60, n2=n2, src_x1=30,
                             src_depth=300.0, **kw)
    big = cases.build_case(name="abc_big", mod_nz=260, mod_nx=260, n2=n2, src_x1=130,
                           src_depth=1300.0, **kw)
    win = (slice(n2 + 100, n2 + 160), slice(n2 + 100, n2 + 160))
    eb = cases.oracle_engine(big, prec=64)
    eb.forward_begin()
    eb.forward_advance(198)  # u(200): the direct wave fills the window
    incident = np.linalg.norm(eb.field(0, f64=True)[win])
    eb.forward_advance(550)
    ref = eb.field(0, f64=True)[win]
    us = _run(small, prec=64).field(0, f64=True)[n2:-n2, n2:-n2]
    m = dict(small["model"])
    m["w"] = np.zeros_like(m["w"])
    s2 = dict(small)
    s2["model"] = m
    ur = _run(s2, prec=64).field(0, f64=True)[n2:-n2, n2:-n2]
    resid = np.linalg.norm(us - ref)
    resid_plain = np.linalg.norm(ur - ref)
    assert resid < 0.01 * incident, (resid, incident)
    assert resid_plain > 100 * resid, (resid_plain, resid)


def test_time_reversal_reconstruction_fp64():
    # The backward pass rebuilds u(k) from the saved strips: for k >= NT2 (no
    # wavelet terms involved, see kernel.cu:889) it equals the forward field.
    case = cases.build_case(name="tr", nt=300)
    nt = case["cfg"]["nt"]
    kchk = 150
    e = cases.oracle_engine(case, prec=64)
    e.forward_begin()
    e.forward_advance(kchk - 1)       # steps 2..kchk
    fwd = e.field(0, f64=True)
    e.forward_advance(nt)
    seis = np.zeros((case["cfg"]["n"], nt), np.float32)
    e.backward_begin(seis)
    e.backward_advance(nt - 2 - kchk)  # steps nt-3 .. kchk
    rec = e.field(0, f64=True)
    n2 = case["cfg"]["n2"]
    a, b = rec[n2:-n2, n2:-n2], fwd[n2:-n2, n2:-n2]
    assert cases.rel_l2(a, b) < 1e-9


def test_snapshot_checkpoints_fp64():
    # snap_every = 1 checkpoints u(k) for every 2 <= k <= nt-3, so the backward
    # source field is the forward field at every step (u(1) and u(0) are then
    # rebuilt from exact neighbours) and the source illumination
    # I2 = sum_k u(k)^2 follows from the forward fields alone.
    # Without checkpoints the wavelet terms of kernel.cu:889 (f((k+1)tau) for
    # k < NT2) make the early reconstructed fields, and I2, differ.
    case = cases.build_case(name="snap", nt=160)
    nt = case["cfg"]["nt"]
    n2 = case["cfg"]["n2"]
    seis = np.zeros((case["cfg"]["n"], nt), np.float32)
    e = cases.oracle_engine(case, prec=64)
    e.forward_begin()
    fwd = [np.zeros((case["cfg"]["nz"], case["cfg"]["nx"])), e.field(0, f64=True)]
    for _ in range(nt - 2):
        e.forward_advance(1)
        fwd.append(e.field(0, f64=True))
    expect = sum(u * u for u in fwd[1:])[n2:-n2, n2:-n2]

    def i2_of(snap):
        cfg = dict(case["cfg"])
        cfg.update(r_u=case["r_u"][0], r_x=case["r_x"][0], snap_every=snap)
        eng = ob.Engine(cfg, case["model"], prec=64)
        eng.forward_begin()
        eng.forward_advance(nt)
        eng.backward_begin(seis)
        eng.backward_advance(nt)
        return eng.field(4, f64=True)[n2:-n2, n2:-n2]

    err_snap = cases.rel_l2(i2_of(1), expect)
    err_plain = cases.rel_l2(i2_of(0), expect)
    assert err_snap < 1e-12, err_snap
    assert err_plain > 1e3 * max(err_snap, 1e-16), (err_plain, err_snap)
    # a coarser spacing still restores the checkpointed steps exactly
    cfg = dict(case["cfg"])
    cfg.update(r_u=case["r_u"][0], r_x=case["r_x"][0], snap_every=40)
    eng = ob.Engine(cfg, case["model"], prec=64)
    eng.forward_begin()
    eng.forward_advance(nt)
    eng.backward_begin(seis)
    eng.backward_advance(nt - 3 - 40)  # steps nt-3 .. 41: the field is now u(41)
    assert np.array_equal(eng.field(0, f64=True), fwd[41])
    eng.backward_advance(1)            # u(40)
    assert np.array_equal(eng.field(0, f64=True), fwd[40])


def test_flat_reflector_image_depth():
    # One flat interface; data modelled on the sharp model, migrated on it:
    # the Laplacian-filtered image peaks at the interface depth.
    mod_nz, mod_nx, n2 = 80, 80, 15
    case = cases.build_case(name="flat", mod_nz=mod_nz, mod_nx=mod_nx, n2=n2, nfdmax=4, nt=500,
                            nlayer=2, vlo=2000.0, vhi=3000.0, dip=0.0, src_x1=40,
                            src_depth=100.0, seed=5)
    raw = case["raw"]
    iface = int(np.argmax(raw[0] > raw[0][0]))
    seis = cases.synth_gather(case)
    # remove the direct wave: model the same shot in the homogeneous upper layer
    m = dict(case["model"])
    m["v"] = np.full_like(m["v"], m["v"][n2, n2])
    m["r1"] = np.full_like(m["r1"], m["r1"][n2, n2])
    c2 = dict(case)
    c2["model"] = m
    seis = seis - cases.synth_gather(c2)
    e = cases.oracle_engine(case)
    e.forward_begin()
    e.forward_advance(case["cfg"]["nt"])
    e.backward_begin(seis)
    e.backward_advance(case["cfg"]["nt"])
    up, _ = ob.image_post(e.field(3), e.field(4), m["v"] * 0 + case["model"]["v"], n2,
                          case["model"]["vmax"], 1e-4)
    prof = np.abs(up[:, 20:60]).sum(axis=1)
    prof[:8] = 0  # source/receiver neighbourhood
    assert abs(int(np.argmax(prof)) - iface) <= 3


# ---- Born modelling and its adjoint (least-squares migration; not in the reference) ----
def _born_cfg(case, **kw):
    cfg = dict(case["cfg"])
    cfg.update(r_u=case["r_u"][0], r_x=case["r_x"][0], record=1, nthreads=4)
    cfg.update(kw)
    return cfg


def _smooth_perturbation(cfg, seed=1, margin=(8, 3, 5)):
    from scipy.ndimage import gaussian_filter
    nz, nx, n2 = cfg["nz"], cfg["nx"], cfg["n2"]
    rng = np.random.default_rng(seed)
    d = np.zeros((nz, nx))
    d[n2 + margin[0]:nz - n2 - margin[1], n2 + margin[2]:nx - n2 - margin[2]] = rng.standard_normal(
        (nz - 2 * n2 - margin[0] - margin[1], nx - 2 * n2 - 2 * margin[2]))
    d = gaussian_filter(d, 1.5)
    d[:n2] = 0
    d[-n2:] = 0
    d[:, :n2] = 0
    d[:, -n2:] = 0
    d[cfg["r_u"], cfg["r_x"]] = 0
    return d / np.abs(d).max()


def _born_record(cfg, model, refl):
    e = ob.Engine(cfg, model, prec=64)
    e.set_reflectivity(np.asarray(refl, np.float32))
    e.born_begin()
    e.born_advance(cfg["nt"])
    return e.record().astype(np.float64)


def test_born_matches_finite_difference_fp64():
    # With one coefficient row for all cells (Taylor) the first-order change of
    # the discrete forward update under v -> v + dv is exactly the Born source
    # m*((u(k)-2u(k-1))+u(k-2)) with m = 2dv/v (off the source cell): central
    # differences of the recorded data converge to the Born record at second order.
    case = cases.build_case(name="bornT", ilste=1, nfdmax=4, nt=300, seed=3)
    cfg = _born_cfg(case)
    m0 = case["model"]
    v = m0["v"].astype(np.float64)
    dv = _smooth_perturbation(cfg) * 20.0

    def rec(vv):
        m = dict(m0)
        m["v"] = vv.astype(np.float32)
        e = ob.Engine(cfg, m, prec=64)
        e.forward_begin()
        e.forward_advance(cfg["nt"])
        return e.record().astype(np.float64)

    errs = []
    for a in (0.5, 1.0):
        vp, vm = (v + a * dv).astype(np.float32), (v - a * dv).astype(np.float32)
        half = (vp.astype(np.float64) - vm.astype(np.float64)) / 2  # the perturbation as stored
        fd = (rec(vp) - rec(vm)) / 2
        b = _born_record(cfg, m0, 2 * half / v)
        errs.append(cases.rel_l2(fd, b))
    assert errs[1] < 1e-3, errs
    assert 0.1 < errs[0] / errs[1] < 0.45, errs  # second order in the amplitude


def test_born_superposition_fp64():
    case = cases.build_case(name="bornT", ilste=1, nfdmax=4, nt=300, seed=3)
    cfg = _born_cfg(case)
    m1 = (_smooth_perturbation(cfg, seed=1) * 0.01).astype(np.float32)
    m2 = np.zeros_like(m1)
    m2[cfg["n2"] + 20, cfg["n2"] + 30] = 0.1
    b1 = _born_record(cfg, case["model"], m1)
    b2 = _born_record(cfg, case["model"], m2)
    b12 = _born_record(cfg, case["model"], m1 + m2)
    # FP64 arithmetic; the reflectivity (m1 + m2 rounded once) and the record
    # are held in float32, which bounds the agreement near 1e-7
    assert cases.rel_l2(b12, b1 + b2) < 1e-6
    assert cases.rel_l2(b12, b1) > 1e-2


@pytest.mark.parametrize("name,kw,tol", [
    ("born_hom", dict(ilste=1, nfdmax=4, nt=200, seed=3, nlayer=1, vlo=2000, vhi=2000,
                      mod_nz=60, mod_nx=80, n2=20), 0.01),
    ("bornT", dict(ilste=1, nfdmax=4, nt=300, seed=3), 0.06),
    ("bornLS", dict(ilste=0, nfdmax=4, nt=300, seed=3), 0.02)])
def test_born_adjoint_dot_product_fp64(name, kw, tol):
    # <B m, r> = <m, B^T r> up to the hybrid ABC, which is not self-adjoint: the
    # interior recursion is carried as mu = P*lambda (A^T = P^-1 A P), which is
    # exact for one coefficient row (Taylor); per-cell LS rows take the
    # transposed stencil on the interior (each tap with the neighbour's row),
    # which brings them to the same ABC residue (12% with the forward stencil).
    # A wrong time alignment, a missing P scaling or dropped final samples fail this.
    case = cases.build_case(name=name, **kw)
    cfg = _born_cfg(case)
    n2 = cfg["n2"]
    rng = np.random.default_rng(1)
    refl = np.zeros((cfg["nz"], cfg["nx"]), np.float32)
    refl[n2 + 8:cfg["nz"] - n2 - 3, n2 + 5:cfg["nx"] - n2 - 5] = rng.standard_normal(
        (cfg["nz"] - 2 * n2 - 11, cfg["nx"] - 2 * n2 - 10)).astype(np.float32) * 0.01
    b = _born_record(cfg, case["model"], refl)
    r = rng.standard_normal((cfg["n"], cfg["nt"])).astype(np.float32)
    e = ob.Engine(cfg, case["model"], prec=64)
    e.forward_begin()
    e.forward_advance(cfg["nt"])
    e.adjoint_begin(r)
    e.adjoint_advance(cfg["nt"])
    g = e.field(9, f64=True)
    lhs = float(np.sum(b * r.astype(np.float64)))
    rhs = float(np.sum(refl.astype(np.float64) * g))
    print(f"{name}: <Bm, r> = {lhs:.6e}, <m, B^T r> = {rhs:.6e}, mismatch {abs(lhs - rhs) / abs(lhs):.2%}")
    assert abs(lhs - rhs) <= tol * abs(lhs), (lhs, rhs)
