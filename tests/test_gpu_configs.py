"""GPU runs of the workload configurations at their stated sizes.

  single shot   200 x 400 interior at 10 m, 5 seeded flat layers (1500-4500 m/s),
                8th-order adaptive LS operators (nfdmax 4), 20-point hybrid ABC,
                25 Hz Ricker at x = 200, 30 m deep, dt = 1 ms, 2000 steps, a
                receiver at every surface point; observed data modelled by the
                engine.  The whole run equals the FP32 oracle bit for bit, checked
                every 100 steps in both passes; the image error against the FP64
                oracle is stated; the captured-graph replay equals the eager run.
  apertures     the seeded 4001 x 801 GRF model with 6 reflectors (10 m), surface
                shots with aperture windows of 600..2400 columns, 15 Hz, dt =
                0.8 ms, N2 = 30, nfdmax 4, in two ragged batches at 300 steps:
                every shot equals its one-shot run, and one shot equals the FP32
                oracle.

The reverse-VSP geometry runs at full size in test_gpu_fullsize.py (and with
FP16/BF16 storage in test_gpu_precision.py); the deblurring pipeline at full
size in test_gpu_deblur.py."""
import numpy as np
import pytest

import cases
import oracle_bind as ob

pytestmark = pytest.mark.gpu


def _cfg1():
    return cases.build_case(name="cfg1", kind="layered5", mod_nz=200, mod_nx=400, n2=20,
                            nfdmax=4, nfdmin=2, nt=2000, h=10.0, hz=10.0, tao=0.001, f0=25.0,
                            ilste=0, icompen=1, seed=1, src_x1=201, src_depth=30.0)


def test_single_shot_layered_200x400_nt2000_whole_run():
    from rtm_gpu_amd import hip
    case = _cfg1()
    cfg = case["cfg"]
    nt, n2 = cfg["nt"], cfg["n2"]
    assert (cfg["nz"], cfg["nx"], cfg["nt2"]) == (240, 440, 80)
    print(f"single shot: {cfg['nz']}x{cfg['nx']}, nt={nt}", flush=True)
    # observed data: the engine's forward modelling equals the oracle's
    rec = cases.hip_context(case, record=1)
    rec.forward_begin(save_strips=False)
    rec.forward_advance(nt)
    seis = rec.records()[0]
    rec.close()
    eng = cases.oracle_engine(case, record=1, nthreads=8)
    eng.forward_begin()
    eng.forward_advance(nt)
    assert np.array_equal(seis, eng.record())
    # migration, compared every 100 steps
    ctx = cases.hip_context(case)
    ctx.set_gathers(seis[None])
    e = cases.oracle_engine(case, nthreads=8)
    ctx.forward_begin()
    e.forward_begin()
    done = 2
    while done < nt:
        s = min(100, nt - done)
        ctx.forward_advance(s)
        e.forward_advance(s)
        done += s
        assert np.array_equal(ctx.field(hip.FIELD_U), e.field(0)), f"forward step {done - 1}"
    ctx.backward_begin()
    e.backward_begin(seis)
    inner = (slice(n2, -n2), slice(n2, -n2))
    left = nt - 2
    while left > 0:
        s = min(100, left)
        ctx.backward_advance(s)
        e.backward_advance(s)
        left -= s
        k = left
        assert np.array_equal(ctx.field(hip.FIELD_U)[inner], e.field(0)[inner]), f"u({k})"
        assert np.array_equal(ctx.field(hip.FIELD_Q), e.field(2)), f"q at step {k}"
    for w, o in ((hip.FIELD_I1, 3), (hip.FIELD_I2, 4)):
        assert np.array_equal(ctx.field(w)[inner], e.field(o)[inner])
    img = ctx.field(hip.FIELD_I1)
    # the image error against the FP64 oracle, stated
    e64 = cases.oracle_engine(case, prec=64, nthreads=8)
    e64.forward_begin()
    e64.forward_advance(nt)
    e64.backward_begin(seis)
    e64.backward_advance(nt)
    err = cases.rel_l2(img[inner], e64.field(3, f64=True)[inner])
    print(f"single shot: image relative L2 to the FP64 oracle {err:.2e}")
    assert err < 1e-3
    # the captured time loop replays to the same image
    ctx.set_graph(True)
    for _ in range(2):
        ctx.migrate()
        assert np.array_equal(ctx.field(hip.FIELD_I1), img)
    ctx.close()


def test_grf_apertures_ragged_batches_truncated():
    import bench
    from rtm_gpu_amd import hip
    nshot, nt = 6, 300
    A = bench.build_apertures(nshot, nt=nt)
    d, cfg, n2 = A["d"], A["cfg"], A["n2"]
    print(f"apertures: widths {A['widths'].tolist()}, nt={d.nt}", flush=True)
    common = dict(nz=d.nz, n2=n2, nfdmax=cfg.nfdmax, nt=d.nt, nt2=d.nt2, ilste=0, icompen=1,
                  s_z=d.s_z, ds=1, tao2=d.tao2, h2=d.h2, taoh=d.taoh, taoh2=d.taoh2,
                  hzx2_1=d.hzx2_1)
    nxw = A["wmax"] + 2 * n2
    B = 3
    got = {}
    ctx = hip.Context(**common, nx=nxw, s_l=n2, s_r=n2 + A["wmax"] - 1, n=A["wmax"],
                      max_shots=B, record=1, stack_nx=A["mod_nx"])
    ctx.set_wavelet(A["wav"])
    for b0 in range(0, nshot, B):
        idx = list(range(b0, b0 + B))
        flat = np.full((d.nz, nxw), A["vmin"], np.float32)
        ctx.set_model(flat, flat * 0, A["w"], A["c"], A["index"], A["nvel"], A["vmin"], 1.0)
        ctx.set_shot_geometry([A["geoms"][i] for i in idx])
        for j, i in enumerate(idx):
            ctx.set_shot_model(j, *A["window"](A["true"], i))
        ctx.forward_begin(save_strips=False)
        ctx.forward_advance(d.nt)
        ctx.gathers_from_records()
        recs = ctx.records()
        for j, i in enumerate(idx):
            ctx.set_shot_model(j, *A["window"](A["mig"], i))
        ctx.migrate()
        for j, i in enumerate(idx):
            nxs = A["geoms"][i]["nx"]
            got[i] = (ctx.field(hip.FIELD_I1, shot=j)[:, :nxs], recs[j][:A["geoms"][i]["n"]])
    ctx.close()
    for i in range(nshot):
        gm = A["geoms"][i]
        one = hip.Context(**common, nx=gm["nx"], s_l=gm["s_l"], s_r=gm["s_r"], n=gm["n"],
                          max_shots=1, record=1)
        one.set_model(*A["window"](A["true"], i), A["w"], A["c"], A["index"], A["nvel"],
                      A["vmin"], 1.0)
        one.set_wavelet(A["wav"])
        one.set_shots([gm["r_u"]], [gm["r_x"]])
        one.forward_begin(save_strips=False)
        one.forward_advance(d.nt)
        seis = one.records()[0]
        assert np.array_equal(seis, got[i][1]), f"shot {i} records"
        one.gathers_from_records()
        v, r1 = A["window"](A["mig"], i)
        one.set_model(v, r1, A["w"], A["c"], A["index"], A["nvel"], A["vmin"], 1.0)
        one.migrate()
        img = one.field(hip.FIELD_I1)
        one.close()
        assert np.array_equal(img, got[i][0]), f"shot {i} image"
        if i == 0:  # and the one-shot run is the FP32 replica
            ocfg = dict(nz=d.nz, nx=gm["nx"], n2=n2, nfdmax=cfg.nfdmax, nt=d.nt, nt2=d.nt2,
                        ilste=0, icompen=1, s_l=gm["s_l"], s_r=gm["s_r"], s_z=d.s_z, ds=1,
                        n=gm["n"], r_u=gm["r_u"], r_x=gm["r_x"], tao2=d.tao2, h2=d.h2,
                        taoh=d.taoh, taoh2=d.taoh2, hzx2_1=d.hzx2_1, vmin=A["vmin"], dv=1.0,
                        nthreads=16)
            model = dict(v=v, r1=r1, w=A["w"], c=A["c"], index=A["index"], nvel=A["nvel"],
                         wav=A["wav"])
            e = ob.Engine(ocfg, model)
            e.forward_begin()
            e.forward_advance(d.nt)
            e.backward_begin(seis)
            e.backward_advance(d.nt)
            inner = (slice(n2, -n2), slice(n2, -n2))
            assert np.array_equal(img[inner], e.field(3)[inner])
            print(f"apertures: shot 0 ({gm['n']} columns) equals the FP32 oracle", flush=True)
