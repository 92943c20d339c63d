"""GPU parity: librtm_hip.so (called through its C ABI) against the FP32
replica in oracle/ on the same seeded inputs.

The bar: per-100-step wavefield relative L2 <= 1e-5 and final image relative
L2 <= 1e-4 against the FP32 replica.  The kernels evaluate the same operation
sequence as the replica, so these tests also report (and require) exact
equality of the values; the FP64 oracle error is printed for information."""
import numpy as np
import pytest

import cases
from cases import build_case, hip_context, oracle_engine, rel_l2, synth_gather

pytestmark = pytest.mark.gpu

WAVE_TOL = 1e-5
IMAGE_TOL = 1e-4

CASES = {
    # adaptive LS coefficients, compensated imaging, source inside the interior
    "ls_nfd4": dict(name="ls_nfd4"),
    # Taylor coefficients (Add_Con / BKAdd_EFF_Con / BKAdd_Con), non-compensated
    "taylor_nfd6": dict(name="taylor_nfd6", ilste=1, nfdmax=6, icompen=0, seed=2),
    # source in the ABC frame (row N2-1, as the first cfg2 source) and a
    # receiver line with ds=3 reaching into the right frame
    "frame_src": dict(name="frame_src", src_depth=5.0, ds=3, s_l1=4, n=26, nfdmax=5, n2=8,
                      seed=3, mod_nx=72),
    # maximum operator length of the reference parameter file (nfdmax = N2 = 10)
    "ls_nfd10_n2_10": dict(name="ls_nfd10_n2_10", nfdmax=10, n2=10, mod_nz=36, mod_nx=48, seed=4,
                           h=12.5, hz=12.5, f0=20.0, vhi=3500.0),
    # the longest operator the kernels take (nfdmax = n2 = 12, Taylor rows), an
    # interior width that is not a multiple of 4
    "taylor_nfd12_n2_12": dict(name="taylor_nfd12_n2_12", ilste=1, nfdmax=12, n2=12, mod_nz=38,
                               mod_nx=53, seed=5, h=12.5, hz=12.5, f0=20.0, src_x1=20),
}


def _case(key):
    return build_case(**CASES[key])


def _fwd_compare(case, every=40):
    ctx = hip_context(case, nshot=1)
    eng = oracle_engine(case)
    nt = case["cfg"]["nt"]
    ctx.forward_begin()
    eng.forward_begin()
    k = 2
    worst = 0.0
    while k < nt:
        step = min(every, nt - k)
        ctx.forward_advance(step)
        eng.forward_advance(step)
        k += step
        g = ctx.field(cases_FIELD_U())
        o = eng.field(0)
        err = rel_l2(g, o)
        worst = max(worst, err)
        assert err <= WAVE_TOL, f"step {k}: rel L2 {err}"
        assert np.array_equal(g, o), f"step {k}: max |diff| {np.max(np.abs(g - o))}"
    return ctx, eng, worst


def cases_FIELD_U():
    from rtm_gpu_amd import hip
    return hip.FIELD_U


@pytest.mark.parametrize("key", list(CASES))
def test_forward_wavefield_parity(key):
    case = _case(key)
    print(f"case {key}: grid {case['cfg']['nz']}x{case['cfg']['nx']} nt={case['cfg']['nt']}",
          flush=True)
    ctx, eng, worst = _fwd_compare(case)
    print(f"  forward worst rel L2 vs FP32 replica: {worst:.3e}")
    # strips of a few slots
    for k in (0, 1, 2, case["cfg"]["nt"] // 2, case["cfg"]["nt"] - 1):
        assert np.array_equal(ctx.strip(0, k), eng.strip(k)), f"strip slot {k}"


@pytest.mark.parametrize("key", ["ls_nfd4", "frame_src"])
def test_receiver_records(key):
    from rtm_gpu_amd import hip  # noqa: F401
    case = _case(key)
    ctx = hip_context(case, nshot=1, record=1)
    ctx.forward_begin(save_strips=False)
    ctx.forward_advance(case["cfg"]["nt"])
    rec = ctx.records()[0]
    ref = synth_gather(case)
    assert rec.shape == ref.shape
    assert np.array_equal(rec, ref)
    assert np.abs(ref).max() > 0


@pytest.mark.parametrize("key", list(CASES))
def test_backward_and_images(key):
    from rtm_gpu_amd import hip
    case = _case(key)
    seis = synth_gather(case)
    ctx = hip_context(case, nshot=1)
    ctx.set_gathers(seis[None])
    eng = oracle_engine(case)
    nt = case["cfg"]["nt"]
    ctx.forward_begin()
    ctx.forward_advance(nt)
    eng.forward_begin()
    eng.forward_advance(nt)
    ctx.backward_begin()
    eng.backward_begin(seis)
    done = 0
    while done < nt - 2:
        step = min(50, nt - 2 - done)
        ctx.backward_advance(step)
        eng.backward_advance(step)
        done += step
        for which, ow in ((hip.FIELD_Q, 2), (hip.FIELD_I1, 3), (hip.FIELD_I2, 4)):
            g = ctx.field(which)
            o = eng.field(ow)
            err = rel_l2(g, o)
            assert err <= WAVE_TOL, f"field {which} after {done} backward steps: {err}"
            assert np.array_equal(g, o), f"field {which} after {done}: {np.max(np.abs(g - o))}"
        # the reconstructed source field on the interior
        n2 = case["cfg"]["n2"]
        g = ctx.field(hip.FIELD_U)[n2:-n2, n2:-n2]
        o = eng.field(0)[n2:-n2, n2:-n2]
        assert np.array_equal(g, o), f"u after {done}"
    i1 = ctx.field(hip.FIELD_I1)
    o1 = eng.field(3)
    assert rel_l2(i1, o1) <= IMAGE_TOL
    e64 = oracle_engine(case, prec=64)
    e64.forward_begin()
    e64.forward_advance(nt)
    e64.backward_begin(seis)
    e64.backward_advance(nt)
    print(f"case {key}: image rel L2 vs FP32 replica {rel_l2(i1, o1):.2e}, "
          f"vs FP64 oracle {rel_l2(i1, e64.field(3, f64=True)):.2e}")


@pytest.mark.parametrize("zseg", [1, 7, 40])
def test_ragged_width_and_segments(zseg):
    """Interior width not a multiple of 4 (the last lane's float4 straddles
    the right frame) and row segments of several lengths (the interior kernel
    rounds them to its queue length; the last segment ends past the interior).
    All fields, frame cells included, must equal the FP32 replica."""
    from rtm_gpu_amd import hip
    case = build_case(name="ragged", mod_nz=37, mod_nx=61, nfdmax=3, n2=9, seed=5, nt=200,
                      src_x1=17)
    assert (case["cfg"]["nx"] - 2 * case["cfg"]["n2"]) % 4 == 1
    seis = synth_gather(case)
    ctx = hip_context(case, nshot=1, zseg=zseg)
    ctx.set_gathers(seis[None])
    eng = oracle_engine(case)
    nt = case["cfg"]["nt"]
    ctx.forward_begin()
    ctx.forward_advance(nt)
    eng.forward_begin()
    eng.forward_advance(nt)
    assert np.array_equal(ctx.field(hip.FIELD_U), eng.field(0))
    ctx.backward_begin()
    eng.backward_begin(seis)
    ctx.backward_advance(nt - 2)
    eng.backward_advance(nt - 2)
    for which, ow in ((hip.FIELD_Q, 2), (hip.FIELD_I1, 3), (hip.FIELD_I2, 4), (hip.FIELD_SU, 5),
                      (hip.FIELD_SQ, 6)):
        g = ctx.field(which)
        o = eng.field(ow)
        assert np.array_equal(g, o), f"field {which}: max |diff| {np.max(np.abs(g - o))}"


def test_many_identical_shots_agree():
    """64 copies of one shot on a 200 x 1100 interior: several hundred blocks
    per launch, so CUs hold more than one block at a time.  Every copy must
    equal the FP32 replica bit for bit (regression: a wide store whose data
    registers were overwritten one instruction later lost dwords only under
    that load)."""
    from rtm_gpu_amd import hip
    case = build_case(name="wide", mod_nz=200, mod_nx=1100, n2=12, nfdmax=4, nt=300, seed=6,
                      src_x1=550, n=1100)
    seis = synth_gather(case)
    ns = 64
    m = case["model"]
    cfg = case["cfg"]
    keys = ("nz", "nx", "n2", "nfdmax", "nt", "nt2", "ilste", "icompen", "s_l", "s_r", "s_z",
            "ds", "n", "tao2", "h2", "taoh", "taoh2", "hzx2_1")
    # the frame split on: the band edges in the four-column kernel (n2 = 12)
    ctx = hip.Context(**{k: cfg[k] for k in keys}, max_shots=ns, frame_split=1)
    ctx.set_model(m["v"], m["r1"], m["w"], m["c"], m["index"], m["nvel"], m["vmin"], m["dv"])
    ctx.set_wavelet(m["wav"])
    ctx.set_shots(case["r_u"] * ns, case["r_x"] * ns)
    ctx.set_gathers(np.repeat(seis[None], ns, axis=0))
    ctx.migrate()
    eng = oracle_engine(case, nthreads=8)
    eng.forward_begin()
    eng.forward_advance(cfg["nt"])
    eng.backward_begin(seis)
    eng.backward_advance(cfg["nt"])
    for which, ow in ((hip.FIELD_I1, 3), (hip.FIELD_I2, 4), (hip.FIELD_Q, 2)):
        ref = eng.field(ow)
        bad = [s for s in range(ns) if not np.array_equal(ctx.field(which, shot=s), ref)]
        assert not bad, f"field {which}: shots {bad[:8]} differ from the replica"


def test_batched_shots_equal_single():
    case = build_case(name="batch3", nshot=3, depth_step=40.0)
    seis = np.stack([synth_gather(case, s) for s in range(3)])
    ctx = hip_context(case, nshot=3)
    ctx.set_gathers(seis)
    ctx.migrate()
    from rtm_gpu_amd import hip
    for s in range(3):
        eng = oracle_engine(case, shot=s)
        eng.forward_begin()
        eng.forward_advance(case["cfg"]["nt"])
        eng.backward_begin(seis[s])
        eng.backward_advance(case["cfg"]["nt"])
        assert np.array_equal(ctx.field(hip.FIELD_I1, shot=s), eng.field(3)), f"shot {s}"
        assert np.array_equal(ctx.field(hip.FIELD_I2, shot=s), eng.field(4)), f"shot {s}"


def test_image_post_and_stack():
    import oracle_bind as ob
    case = build_case(name="batch3", nshot=3, depth_step=40.0)
    seis = np.stack([synth_gather(case, s) for s in range(3)])
    ctx = hip_context(case, nshot=3)
    ctx.set_gathers(seis)
    ctx.migrate()
    m = case["model"]
    up, down = ctx.image_post(m["vmax"], 1e-4)
    img, ill = ctx.stack_finish(3, 1)
    from rtm_gpu_amd import hip
    ups, downs = [], []
    for s in range(3):
        u, d = ob.image_post(ctx.field(hip.FIELD_I1, s), ctx.field(hip.FIELD_I2, s), m["v"],
                             case["cfg"]["n2"], m["vmax"], 1e-4)
        assert np.array_equal(u, up[s]) and np.array_equal(d, down[s]), f"post shot {s}"
        ups.append(u)
        downs.append(d)
    oimg, oill = ob.stack(np.stack(ups), np.stack(downs), 1)
    assert np.array_equal(img, oimg) and np.array_equal(ill, oill)


def test_graph_replay_matches_eager():
    """hipGraph-captured migration: every replay equals the eager result,
    with eager work and host synchronisation between the replays."""
    from rtm_gpu_amd import hip
    case = _case("frame_src")
    seis = synth_gather(case)
    eager = hip_context(case, nshot=1)
    eager.set_gathers(seis[None])
    eager.migrate()
    ref_i1 = eager.field(hip.FIELD_I1)
    ref_i2 = eager.field(hip.FIELD_I2)
    g = hip_context(case, nshot=1)
    g.set_gathers(seis[None])
    g.set_graph(True)
    for rep in range(4):
        g.migrate()
        g.sync()
        assert np.array_equal(g.field(hip.FIELD_I1), ref_i1), f"replay {rep}"
        assert np.array_equal(g.field(hip.FIELD_I2), ref_i2), f"replay {rep}"
        eager.migrate()  # unrelated eager work between replays
        eager.sync()


def test_frame_runs_keep_results():
    """The frame kernel's row runs only re-partition the frame cells among
    waves: every admissible pair of run lengths gives the same fields, strips
    and images bit for bit."""
    from rtm_gpu_amd import hip
    case = build_case(name="batch3", nshot=3, depth_step=40.0)
    seis = np.stack([synth_gather(case, s) for s in range(3)])
    n2 = case["cfg"]["n2"]
    out = []
    for runs in (None, ((n2 + 2) // 3, 10), ((n2 + 1) // 2, 40), (n2, case["cfg"]["nz"])):
        ctx = hip_context(case, nshot=3)
        if runs:
            ctx.set_frame_runs(*runs)
            assert ctx.frame_runs() == runs
        ctx.set_gathers(seis)
        ctx.forward_begin()
        ctx.forward_advance(case["cfg"]["nt"] // 2)
        fwd = [ctx.field(hip.FIELD_U, shot=s) for s in range(3)]
        strip = ctx.strip(1, case["cfg"]["nt"] // 2 - 1)
        ctx.migrate()
        out.append((fwd, strip, [ctx.field(w, shot=s) for s in range(3)
                                 for w in (hip.FIELD_I1, hip.FIELD_I2)]))
        ctx.close()
    for o in out[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(o[0], out[0][0]))
        assert np.array_equal(o[1], out[0][1])
        assert all(np.array_equal(a, b) for a, b in zip(o[2], out[0][2]))
    with pytest.raises(hip.RtmError):
        hip_context(case, nshot=1).set_frame_runs(n2 + 1, 20)


@pytest.mark.parametrize("nfdmax", [2, 4, 6])
def test_frame_split_keeps_results(nfdmax):
    """The band edges of the forward and receiver frames in the four-column
    kernel (frame_split = 1; n2 = 12, several float4 records per band run)
    give the fields, strips and images of the one-column kernel bit for bit,
    with whole-band and shortened frame runs (the receiver frame launch also
    restores the strips of the reconstructed field).  The designs' longest
    operators are 2, 4 and 5: row loops unrolled 12 and 24 times, column halos
    of one and two float4s."""
    from rtm_gpu_amd import hip
    case = build_case(name=f"split{nfdmax}", mod_nz=60, mod_nx=560, n2=12, nfdmax=nfdmax,
                      nfdmin=min(2, nfdmax), nt=200, nshot=3, depth_step=20.0, seed=7)
    seis = np.stack([synth_gather(case, s) for s in range(3)])
    cfg = case["cfg"]
    for runs in (None, ((cfg["n2"] + 2) // 3, 10)):
        out = []
        for split in (0, 1):
            ctx = hip_context(case, nshot=3, frame_split=split)
            if runs:
                ctx.set_frame_runs(*runs)
            ctx.set_gathers(seis)
            ctx.forward_begin()
            ctx.forward_advance(cfg["nt"] // 2)
            fwd = [ctx.field(hip.FIELD_U, shot=s) for s in range(3)]
            strip = ctx.strip(2, cfg["nt"] // 2 - 1)
            ctx.migrate()
            out.append((fwd, strip, [ctx.field(w, shot=s) for s in range(3)
                                     for w in (hip.FIELD_I1, hip.FIELD_I2, hip.FIELD_Q)]))
            ctx.close()
        assert all(np.array_equal(a, b) for a, b in zip(out[1][0], out[0][0]))
        assert np.array_equal(out[1][1], out[0][1])
        assert all(np.array_equal(a, b) for a, b in zip(out[1][2], out[0][2]))


def test_fused_step_launch_keeps_results():
    """A step's interior and frame work in one launch (fuse_frame) gives the
    fields, strips and images of two separate launches bit for bit, for
    several shots (interior blocks of four shots) and for one shot (four row
    segments per block), with the frame runs shortened so that frame units
    spread among the interior units."""
    from rtm_gpu_amd import hip
    for nshot in (3, 5, 1):
        case = build_case(name=f"fuse{nshot}", nshot=nshot, depth_step=20.0)
        seis = np.stack([synth_gather(case, s) for s in range(nshot)])
        keys = ("nz", "nx", "n2", "nfdmax", "nt", "nt2", "ilste", "icompen", "s_l", "s_r",
                "s_z", "ds", "n", "tao2", "h2", "taoh", "taoh2", "hzx2_1")
        out = []
        for fuse in (0, 1):
            cf, m = case["cfg"], case["model"]
            ctx = hip.Context(**{k: cf[k] for k in keys}, max_shots=nshot, fuse_frame=fuse)
            ctx.set_model(m["v"], m["r1"], m["w"], m["c"], m["index"], m["nvel"], m["vmin"], m["dv"])
            ctx.set_wavelet(m["wav"])
            ctx.set_shots(case["r_u"], case["r_x"])
            ctx.set_frame_runs((cf["n2"] + 2) // 3, 10)
            ctx.set_gathers(seis)
            ctx.forward_begin()
            ctx.forward_advance(cf["nt"] // 2)
            fwd = [ctx.field(hip.FIELD_U, shot=s) for s in range(nshot)]
            strip = ctx.strip(nshot - 1, cf["nt"] // 2 - 1)
            ctx.migrate()
            out.append((fwd, strip, [ctx.field(w, shot=s) for s in range(nshot)
                                     for w in (hip.FIELD_I1, hip.FIELD_I2, hip.FIELD_Q)]))
            ctx.close()
        (f0, s0, i0), (f1, s1, i1) = out
        assert all(np.array_equal(a, b) for a, b in zip(f0, f1)), nshot
        assert np.array_equal(s0, s1), nshot
        assert all(np.array_equal(a, b) for a, b in zip(i0, i1)), nshot


def test_autotune_keeps_results():
    """The autotuner only changes the row-segment height of the interior
    kernels and the frame kernels' row runs; images after tuning equal the
    default context's bit for bit."""
    from rtm_gpu_amd import hip
    case = build_case(name="batch3", nshot=3, depth_step=40.0)
    seis = np.stack([synth_gather(case, s) for s in range(3)])
    ref = hip_context(case, nshot=3)
    ref.set_gathers(seis)
    ref.migrate()
    ctx = hip_context(case, nshot=3)
    z = ctx.autotune(3)
    # heights are 1..9 register-queue lengths (2*NFD + 1 + prefetch rows each)
    assert len(z) == 3 and all(4 <= v <= 9 * 27 for v in z), z
    print("frame runs chosen:", ctx.frame_runs())
    ctx.set_gathers(seis)
    ctx.migrate()
    for s in range(3):
        for which in (hip.FIELD_I1, hip.FIELD_I2):
            assert np.array_equal(ctx.field(which, shot=s), ref.field(which, shot=s))
