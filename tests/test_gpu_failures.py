"""GPU: failure handling.  A time step beyond the stability limit makes the
advance calls return RTM_E_BLOWUP (the reference would run on and write NaN
images); a batch whose attempt fails is re-queued by the driver and the run
still equals the run without the failure."""
import threading

import numpy as np
import pytest

import cases
from test_gpu_driver import _make_run_dir

pytestmark = pytest.mark.gpu


def test_unstable_time_step_reports_blowup():
    from rtm_gpu_amd import hip, host
    kw = dict(ilste=1, nfdmax=4, nlayer=3, vlo=2000.0, vhi=3000.0, nt=400, mod_nz=40,
              mod_nx=50, n2=10, f0=15.0)
    bad = cases.build_case(name="blowup", tao=0.0035, **kw)
    cfg, m = bad["cfg"], bad["model"]
    assert host.stability(m["v"], m["c"], None, 0, m["vmin"], m["dv"], cfg["nfdmax"],
                          cfg["tao2"], cfg["h2"], cfg["hzx2_1"]) > 1.2
    print("blowup: 40x50 interior, tao 3.5 ms, 400 steps", flush=True)
    ctx = cases.hip_context(bad)
    ctx.forward_begin()
    with pytest.raises(hip.RtmError, match=r"\(-5\).*forward pass.*NaN or infinity"):
        ctx.forward_advance(cfg["nt"])
    ctx.close()
    # the same model at a stable step runs through, with the check every 10 steps
    good = cases.build_case(name="blowup_ok", tao=0.001, **kw)
    keys = ("nz", "nx", "n2", "nfdmax", "nt", "nt2", "ilste", "icompen", "s_l", "s_r", "s_z",
            "ds", "n", "tao2", "h2", "taoh", "taoh2", "hzx2_1")
    gc, gm = good["cfg"], good["model"]
    ctx = hip.Context(**{k: gc[k] for k in keys}, max_shots=1, check_every=10)
    ctx.set_model(gm["v"], gm["r1"], gm["w"], gm["c"], gm["index"], gm["nvel"], gm["vmin"],
                  gm["dv"])
    ctx.set_wavelet(gm["wav"])
    ctx.set_shots(good["r_u"], good["r_x"])
    ctx.set_gathers(np.zeros((1, gc["n"], gc["nt"]), np.float32))
    ctx.migrate()
    assert np.isfinite(ctx.field(hip.FIELD_I1)).all()
    ctx.close()


def test_driver_requeues_a_failed_batch(tmp_path):
    from rtm_gpu_amd import host
    _make_run_dir(tmp_path)
    txt = tmp_path / "2D_Real_RVSP_RTM.txt"
    tpl = tmp_path / "SGY_Model.sgy"
    host.run(txt, workers=1, batch=1, resume=0, sgy_template=tpl)
    clean = (tmp_path / "RVSP_Migration_Real_new2.dat").read_bytes()
    ups = [(tmp_path / f"RVSP_RTM_up_{m}.dat").read_bytes() for m in (1, 2, 3)]

    seen = []
    lock = threading.Lock()

    def fail_shot_2_once(shots, attempt):
        with lock:
            seen.append((tuple(shots), attempt))
        return not (shots == [1] and attempt == 0)   # the batch of shot 2 fails once

    host.run(txt, workers=2, batch=1, retries=1, resume=0, sgy_template=tpl,
             before_batch=fail_shot_2_once)
    assert sorted(seen) == [((0,), 0), ((1,), 0), ((1,), 1), ((2,), 0)]
    assert (tmp_path / "RVSP_Migration_Real_new2.dat").read_bytes() == clean
    assert [(tmp_path / f"RVSP_RTM_up_{m}.dat").read_bytes() for m in (1, 2, 3)] == ups

    # without retries the failure is reported, naming the refused batch
    with pytest.raises(host.HostError, match="1 batch\\(es\\) failed.*shot 2 refused"):
        host.run(txt, workers=1, batch=1, retries=0, resume=0, sgy_template=tpl,
                 before_batch=lambda shots, attempt: shots != [1])
