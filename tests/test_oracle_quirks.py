"""Pins of the oracle's image post, stack and the reference's quirks (CPU only).

The image post and stack (kernel.cu:935-1059) are checked against closed
forms: the 5-point Laplacian of a quadratic is a constant, the mirror at the
interior edges reflects about the first and last interior rows and columns,
the whitening constant uses the interior maximum of |I2| only, and stacking
constant images gives their mean (and the quotient when iNorm = 1).

The quirks of the reference's time loop are pinned by known answers on tiny
grids, each chosen so that the test fails if the quirk is "fixed":
  Q1  the bottom, left and right one-way ABC take their v^2 term from top-band
      cells (kernel.cu:128,132,136);
  Q5  the backward wavelet gate k < NT2 with f((k+1)tao) (kernel.cu:889) adds
      two tail samples the forward pass never injected;
  Q6  data replace the receiver field only where the sample is non-zero
      (kernel.cu:349-354);
  Q7  backward step k injects sample k+1 (kernel.cu:905-909);
  Q9  the backward initial state (kernel.cu:859-885).
"""
import numpy as np
import pytest

import cases
import oracle_bind as ob


# ---- image post (kernel.cu:935-990) -------------------------------------------
NZ, NX, N2 = 20, 26, 4
VMAX = np.float32(2000.0)


def _speed_field():
    """v = vmax on even columns and vmax/2 on odd ones: the scaling v^2/vmax^2
    is then exactly 1 or 1/4."""
    v = np.full((NZ, NX), VMAX, np.float32)
    v[:, 1::2] = VMAX / 2
    return v, np.where(np.arange(NX)[None, :] % 2 == 1, 0.25, 1.0) * np.ones((NZ, 1))


def _garbage_frame(a, val=1.0e6):
    a = a.copy()
    a[:N2, :] = val
    a[-N2:, :] = val
    a[:, :N2] = val
    a[:, -N2:] = val
    return a


def test_image_post_quadratic_gives_constant_laplacian():
    # I1 = a z^2 + b x^2 + c z x + d z + e x: the 5-point Laplacian is 2a + 2b
    # everywhere away from the mirrored edges, so up = -(2a + 2b) v^2/vmax^2.
    # Frame cells of I1 are never read (filled with 1e6 here).
    z, x = np.meshgrid(np.arange(NZ, dtype=np.float64), np.arange(NX, dtype=np.float64),
                       indexing="ij")
    a, b, c, d, e = 1.0, 3.0, 2.0, -5.0, 7.0
    I1 = _garbage_frame((a * z * z + b * x * x + c * z * x + d * z + e * x).astype(np.float32))
    I2 = np.ones((NZ, NX), np.float32)
    v, scale = _speed_field()
    up, _ = ob.image_post(I1, I2, v, N2, VMAX, np.float32(0.0))
    want = (-(2 * a + 2 * b) * scale)[N2:-N2, N2:-N2]
    assert np.array_equal(up[1:-1, 1:-1], want[1:-1, 1:-1].astype(np.float32))


@pytest.mark.parametrize("axis", [0, 1])
def test_image_post_mirror_edges(axis):
    # I1 linear along one axis (slope s, zero at the first interior row or
    # column): the interior Laplacian is 0; the mirror about row/column N2
    # (i-1 -> 2N2-(i-1), kernel.cu:939-942) gives 2s on the first interior
    # line, and the mirror about NZ-N2-1 / NX-N2-1 (kernel.cu:943-946) gives
    # -2s on the last.  A mirror about N2-1/2 (edge clamp) would give s.
    s = 3.0
    z, x = np.meshgrid(np.arange(NZ, dtype=np.float64), np.arange(NX, dtype=np.float64),
                       indexing="ij")
    coord = z if axis == 0 else x
    I1 = _garbage_frame((s * (coord - N2)).astype(np.float32))
    v, scale = _speed_field()
    up, _ = ob.image_post(I1, np.ones_like(I1), v, N2, VMAX, np.float32(0.0))
    sc = scale[N2:-N2, N2:-N2]
    lap = np.zeros_like(sc)
    if axis == 0:
        lap[0, :], lap[-1, :] = 2 * s, -2 * s
    else:
        lap[:, 0], lap[:, -1] = 2 * s, -2 * s
    assert np.array_equal(up, (-lap * sc).astype(np.float32))


def test_image_post_whitening_uses_interior_max_abs():
    # down = I2 + whitecoe * max|I2| with the maximum over the interior only
    # (kernel.cu:963-990): the largest |value| inside is a negative one (-8);
    # the frame holds 100, which must not count.
    rng = np.random.default_rng(5)
    I2 = rng.integers(-4, 5, size=(NZ, NX)).astype(np.float32)
    I2[N2 + 3, N2 + 5] = -8.0
    I2 = _garbage_frame(I2, 100.0)
    v, _ = _speed_field()
    _, down = ob.image_post(np.zeros_like(I2), I2, v, N2, VMAX, np.float32(0.25))
    assert np.array_equal(down, I2[N2:-N2, N2:-N2] + np.float32(2.0))


@pytest.mark.parametrize("inorm", [0, 1])
def test_stack_of_constant_images(inorm):
    # kernel.cu:1000-1059: img = sum(up)/nrec, illum = sum(down)/nrec, then
    # img/illum when iNorm = 1.
    nI = (NZ - 2 * N2) * (NX - 2 * N2)
    ups = np.stack([np.full(nI, c, np.float32) for c in (1.0, 2.0, 6.0)])
    downs = np.stack([np.full(nI, c, np.float32) for c in (2.0, 4.0, 6.0)])
    img, ill = ob.stack(ups, downs, inorm)
    assert np.all(ill == 4.0)
    assert np.all(img == (0.75 if inorm else 3.0))


# ---- Q1: mis-indexed ABC velocities (kernel.cu:128,132,136) -------------------
def _q1_fields(mod_nz, mod_nx, src_row, src_col, nsteps=11):
    """Forward fields after nsteps on two models equal except in the top band
    (rows 0..N2-1: 2000 vs 3000 m/s), Taylor coefficients (no velocity bins)."""
    case = cases.build_case(name="q1", mod_nz=mod_nz, mod_nx=mod_nx, n2=10, nfdmax=4, ilste=1,
                            nlayer=1, vlo=2000.0, vhi=2000.0, nt=40)
    cfg = dict(case["cfg"])
    nz, nx, n2 = cfg["nz"], cfg["nx"], cfg["n2"]
    cfg.update(r_u=src_row(nz, n2), r_x=src_col(nx, n2), record=0, nthreads=1)
    mA = dict(case["model"])
    mB = dict(mA)
    vB = mA["v"].reshape(nz, nx).copy()
    vB[:n2, :] = 3000.0
    mB["v"] = vB
    out = []
    for m in (mA, mB):
        e = ob.Engine(cfg, m, prec=64)
        e.forward_begin()
        e.forward_advance(nsteps)
        out.append(e.field(0, f64=True))
    return out[0], out[1], nz, nx, n2


# The numerical domain of influence grows by nfd = 4 cells per step, so after
# 11 steps every cell further than 44 cells from the source is still zero on
# both models.  Each geometry lets the wave reach one kind of ABC band and
# none of the others: a difference there can only come from that band's own
# v^2 term reading the top band.
def test_q1_bottom_abc_reads_top_band_velocity():
    # Dv[(N2-x)*NX+y] (kernel.cu:132): bottom layer x takes the velocity of top
    # row N2-x.  Wide model, source 5 rows above the bottom band, mid-width.
    A, B, nz, nx, n2 = _q1_fields(60, 120, lambda nz, n2: nz - n2 - 5, lambda nx, n2: nx // 2)
    assert np.abs(A[:n2 + 4]).max() == 0.0 and np.abs(A[:, :n2 + 4]).max() == 0.0
    assert np.abs(A[:, nx - n2 - 4:]).max() == 0.0
    assert (A != B)[nz - n2:, :].any()


@pytest.mark.parametrize("side", ["left", "right"])
def test_q1_side_abc_reads_top_band_velocity(side):
    # Dv[(N2-y)*NX+x] (kernel.cu:128,136) with x the row: side layer y at row x
    # takes the velocity of linear index (N2-y)*NX+x, a top-band cell.  Tall
    # model, source at mid-depth 6 columns inside the tested side.
    col = (lambda nx, n2: n2 + 6) if side == "left" else (lambda nx, n2: nx - n2 - 7)
    A, B, nz, nx, n2 = _q1_fields(120, 80, lambda nz, n2: nz // 2, col)
    assert np.abs(A[:n2 + 4]).max() == 0.0 and np.abs(A[nz - n2 - 4:]).max() == 0.0
    band = (slice(None), slice(0, n2)) if side == "left" else (slice(None), slice(nx - n2, nx))
    other = (slice(None), slice(nx - n2 - 4, nx)) if side == "left" else (slice(None), slice(0, n2 + 4))
    assert np.abs(A[other]).max() == 0.0            # the other side is not reached
    diff = A != B
    assert diff[band].any()


# ---- Q5: backward wavelet gate (kernel.cu:889) --------------------------------
def test_q5_backward_gate_adds_two_tail_samples():
    # Reversal of forward step k+2 would inject f((k+1)tao) for k < NT2-2; the
    # reference injects it for k < NT2.  So the reconstructed u(NT2) equals the
    # forward one, and u(NT2-1) exceeds it by exactly f(NT2*tao) at the source.
    case = cases.build_case(name="q5", nt=300)
    cfg = case["cfg"]
    nt, nt2, n2 = cfg["nt"], cfg["nt2"], cfg["n2"]
    e = cases.oracle_engine(case, prec=64)
    e.forward_begin()
    e.forward_advance(nt2 - 2)          # steps 2..nt2-1
    f79 = e.field(0, f64=True)
    e.forward_advance(1)                # step nt2
    f80 = e.field(0, f64=True)
    e.forward_advance(nt)
    e.backward_begin(np.zeros((cfg["n"], nt), np.float32))
    e.backward_advance(nt - 2 - nt2)    # steps nt-3 .. nt2
    r80 = e.field(0, f64=True)
    e.backward_advance(1)               # step nt2-1
    r79 = e.field(0, f64=True)
    inner = (slice(n2, -n2), slice(n2, -n2))
    scale = np.abs(f80[inner]).max()
    assert np.abs(r80[inner] - f80[inner]).max() < 1e-9 * scale
    d = (r79 - f79)[inner]
    src = (case["r_u"][0] - n2, case["r_x"][0] - n2)
    tail = float(case["model"]["wav"][nt2])
    assert abs(tail) > 1e-4
    assert abs(d[src] - tail) < 1e-9 * scale
    d[src] = 0.0
    assert np.abs(d).max() < 1e-9 * scale


# ---- Q6 + Q7: data replacement and the k+1 lag (kernel.cu:349-354, 905-909) ---
def test_q6_q7_replacement_only_for_nonzero_samples_at_k_plus_1():
    # Uncompensated imaging, so the receiver field starts at zero.  Trace s0
    # holds one sample at t0; trace s0+1 one sample at t0-2.
    case = cases.build_case(name="q6", nt=120, icompen=0)
    cfg = case["cfg"]
    nt, n, sz, sl, ds = cfg["nt"], cfg["n"], cfg["s_z"], cfg["s_l"], cfg["ds"]
    s0, t0 = n // 2, nt - 20
    seis = np.zeros((n, nt), np.float32)
    seis[s0, t0] = 0.5
    seis[s0 + 1, t0 - 2] = 0.125
    c0, c1 = sl + s0 * ds, sl + (s0 + 1) * ds
    e = cases.oracle_engine(case)
    e.forward_begin()
    e.forward_advance(nt)
    e.backward_begin(seis)
    e.backward_advance(nt - 2 - t0)     # steps nt-3 .. t0: samples > t0 are all zero
    assert np.abs(e.field(2)).max() == 0.0
    e.backward_advance(1)               # step t0-1 injects sample t0 (Q7)
    q = e.field(2)
    assert q[sz, c0] == np.float32(0.5)
    q[sz, c0] = 0.0
    # the same step's one-way ABC reads the new value of its inner neighbour:
    # only the top-band cells just above the receiver may be non-zero
    nzr, nzc = np.nonzero(q)
    assert sz == cfg["n2"] and np.all(nzr < sz) and np.all(np.abs(nzc - c0) <= 1)
    e.backward_advance(1)               # step t0-2: sample t0-1 is zero -> wave update (Q6)
    q = e.field(2)
    assert q[sz, c0] != 0.0 and q[sz, c0] != np.float32(0.5)
    assert q[sz, c1] != 0.0             # a zero sample does not zero the receiver cell
    e.backward_advance(1)               # step t0-3: sample t0-2 of trace s0+1 replaces
    q = e.field(2)
    assert q[sz, c1] == np.float32(0.125)
    assert q[sz, c0] != 0.0


# ---- Q9: backward initial state (kernel.cu:859-885) ---------------------------
@pytest.mark.parametrize("icompen", [1, 0])
def test_q9_backward_initial_state(icompen):
    # BW0 = u(NT-1), BW1 = u(NT-2), FW0 = 0, FW1 = the source spike f(0)/2.
    #   compensated:   sum U = BW0 + BW1, sum Q = FW1, I1 = FW1*(BW0 + BW1)
    #   uncompensated: sum U = sum Q = 0, I1 = FW1*BW1
    #   I2 = BW1^2 + BW0^2 in both.
    case = cases.build_case(name="q9", nt=200, icompen=icompen)
    cfg = case["cfg"]
    n2 = cfg["n2"]
    e = cases.oracle_engine(case)
    e.forward_begin()
    e.forward_advance(cfg["nt"])
    u1, u2 = e.field(0), e.field(1)      # u(NT-1), u(NT-2)
    e.backward_begin(np.zeros((cfg["n"], cfg["nt"]), np.float32))
    inner = (slice(n2, -n2), slice(n2, -n2))
    spike = np.zeros_like(u1)
    ru, rx = case["r_u"][0], case["r_x"][0]
    spike[ru, rx] = np.float32(np.float64(case["model"]["wav"][0]) / 2.0)
    bsum = (u1 + u2).astype(np.float32)
    I1, I2, SU, SQ = (e.field(w)[inner] for w in (3, 4, 5, 6))
    assert np.array_equal(I2, (u2 * u2 + u1 * u1)[inner])
    if icompen:
        assert np.array_equal(SU, bsum[inner])
        assert np.array_equal(SQ, spike[inner])
        want = spike * bsum
    else:
        assert not SU.any() and not SQ.any()
        want = spike * u2
    assert np.count_nonzero(want) == 1
    assert np.array_equal(I1, want[inner])
