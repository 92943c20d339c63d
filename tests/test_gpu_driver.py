"""GPU: the reference-style driver end to end, and reduced-precision strips.

The driver (librtm_host.so over librtm_hip.so, pthread shot dispatch) must
produce exactly what the oracle pipeline produces from the same files:
per-shot image files (kernel.cu:951-990), the stacked, cropped image
(kernel.cu:1000-1108) and its SEG-Y (kernel.cu:1181-1209)."""
import os
import shutil

import numpy as np
import pytest

import cases
import oracle_bind as ob
from test_host import write_param_files

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _make_run_dir(d, nrec=3, ilste=0):
    vals = write_param_files(d, nrec=nrec, ilste=ilste, nt1=260, mod_nz=34, mod_nx=46, n2=12,
                             nfdmax=4, n=46, r_x1=23, nx_bg=2, nx_ed=44, nz_bg=0, nz_ed=34)
    raw = cases.layered_model(34, 46, 4, 1600, 3200, seed=21, dip=0.2)
    raw.astype(np.float32).tofile(os.path.join(d, "model.dat"))
    depths = [40.0 + 30.0 * m for m in range(nrec)]
    with open(os.path.join(d, "Depth_Of_Receiver.txt"), "w") as f:
        f.write(" ".join(f"{z}" for z in depths))
    case = cases.build_case(name=f"drv{ilste}", mod_nz=34, mod_nx=46, n2=12, nfdmax=4, nt=260,
                            h=10.0, hz=10.0, tao=0.001, f0=25.0, ilste=ilste, icompen=1,
                            src_x1=23, nlayer=4, vlo=1600, vhi=3200, seed=21, dip=0.2, n=46,
                            src_depth=depths[0], depth_step=30.0, nshot=nrec)
    assert np.array_equal(case["raw"], raw)
    seis = []
    for m in range(nrec):
        g = cases.synth_gather(case, m)
        g.astype(np.float32).tofile(os.path.join(d, f"NEW_L10-1932-X_{int(depths[m])}.dat"))
        seis.append(g)
    shutil.copy(os.path.join(GOLDEN, "SGY_Model.sgy"), os.path.join(d, "SGY_Model.sgy"))
    return vals, case, seis


def _oracle_pipeline(case, seis, vals):
    n2 = vals["n2"]
    ups, downs = [], []
    for m in range(len(seis)):
        e = cases.oracle_engine(case, shot=m)
        e.forward_begin()
        e.forward_advance(case["cfg"]["nt"])
        e.backward_begin(seis[m])
        e.backward_advance(case["cfg"]["nt"])
        u, dn = ob.image_post(e.field(3), e.field(4), case["model"]["v"], n2,
                              case["model"]["vmax"], vals["whitecoe"])
        ups.append(u)
        downs.append(dn)
    img, _ = ob.stack(np.stack(ups), np.stack(downs), vals["inorm"])
    return ups, downs, img


@pytest.mark.parametrize("workers,batch", [(1, 3), (2, 1)])
def test_driver_end_to_end(tmp_path, workers, batch):
    from rtm_gpu_amd import host
    vals, case, seis = _make_run_dir(tmp_path)
    host.run(tmp_path / "2D_Real_RVSP_RTM.txt", workers=workers, batch=batch,
             sgy_template=tmp_path / "SGY_Model.sgy")
    ups, downs, img = _oracle_pipeline(case, seis, vals)
    mnz, mnx = 34, 46
    for m in range(3):
        up = np.fromfile(tmp_path / f"RVSP_RTM_up_{m + 1}.dat", np.float32).reshape(mnx, mnz).T
        dn = np.fromfile(tmp_path / f"RVSP_RTM_down_{m + 1}.dat", np.float32).reshape(mnx, mnz).T
        assert np.array_equal(up, ups[m]) and np.array_equal(dn, downs[m]), f"shot {m}"
    crop = np.fromfile(tmp_path / "RVSP_Migration_Real_new2.dat", np.float32)
    exp = img[0:34, 2:44].T.ravel()  # j = NX_BG..NX_ED-1 outer, i = NZ_BG..NZ_ED-1 inner
    assert np.array_equal(crop, exp)
    data, fmt = host.read_segy(tmp_path / "RVSP_migration_Real.sgy")
    assert fmt == 1 and data.shape == (42, 34)
    assert np.allclose(data, exp.reshape(42, 34), rtol=2 ** -20, atol=1e-30)
    # post-stack chain (kernel.cu:1110-1179): both files reread as 42 columns of
    # mod_NZ samples, to time, rotated by `angle`, back to depth
    vw = np.fromfile(tmp_path / "vnew.dat", np.float32)[:42 * 34].reshape(42, 34)
    mig = crop[:42 * 34].reshape(42, 34)
    t = host.d2t(vw, mig, vals["hz"], vals["tao"])
    assert np.array_equal(np.fromfile(tmp_path / "RVSP_Migration_Real_T.dat", np.float32),
                          t.ravel())
    tp = host.phase_rotate(t, vals["angle"])
    assert np.array_equal(np.fromfile(tmp_path / "RVSP_Migration_Real_T_phase.dat", np.float32),
                          tp.ravel())
    dd = host.t2d(vw, tp, vals["tao"], vals["hz"])
    assert np.array_equal(np.fromfile(tmp_path / "RVSP_Migration_Real_D.dat", np.float32),
                          dd.ravel())


def test_driver_resume_skips_done_shots(tmp_path):
    from rtm_gpu_amd import host
    vals, case, seis = _make_run_dir(tmp_path)
    host.run(tmp_path / "2D_Real_RVSP_RTM.txt", workers=1, batch=2,
             sgy_template=tmp_path / "SGY_Model.sgy")
    first = (tmp_path / "RVSP_Migration_Real_new2.dat").read_bytes()
    # a marker in shot 2's file survives a resumed run, then a full rerun restores it
    p = tmp_path / "RVSP_RTM_up_2.dat"
    keep = p.read_bytes()
    p.write_bytes(b"\0" * len(keep))
    host.run(tmp_path / "2D_Real_RVSP_RTM.txt", workers=1, batch=2, resume=1,
             sgy_template=tmp_path / "SGY_Model.sgy")
    assert p.read_bytes() == b"\0" * len(keep)
    host.run(tmp_path / "2D_Real_RVSP_RTM.txt", workers=1, batch=2, resume=0,
             sgy_template=tmp_path / "SGY_Model.sgy")
    assert p.read_bytes() == keep
    assert (tmp_path / "RVSP_Migration_Real_new2.dat").read_bytes() == first


@pytest.mark.parametrize("fmt,tol", [(1, 2e-3), (2, 1e-2)])
def test_reduced_precision_strips(fmt, tol):
    """FP16 / BF16 strip storage with per-step power-of-two max-abs scaling:
    the reconstructed source field and the image stay close to FP32 storage;
    the error against the FP64 oracle is printed."""
    from rtm_gpu_amd import hip
    case = cases.build_case(name="strips16", nt=300)
    seis = cases.synth_gather(case)
    res = {}
    for f in (0, fmt):
        ctx = cases.hip_context(case, nshot=1, strip_format=f)
        ctx.set_gathers(seis[None])
        ctx.migrate()
        res[f] = ctx.field(hip.FIELD_I1)
        if f:
            # decoded strips stay within the format's relative precision of the scaled values
            ctx32 = cases.hip_context(case, nshot=1)
            ctx32.forward_begin()
            ctx32.forward_advance(case["cfg"]["nt"])
            ctx.forward_begin()
            ctx.forward_advance(case["cfg"]["nt"])
            for k in (5, 100, 250):
                a, b = ctx.strip(0, k), ctx32.strip(0, k)
                scale = np.abs(b).max()
                assert np.max(np.abs(a - b)) <= (2 ** -10 if fmt == 1 else 2 ** -7) * scale + 1e-30
    e64 = cases.oracle_engine(case, prec=64)
    e64.forward_begin()
    e64.forward_advance(case["cfg"]["nt"])
    e64.backward_begin(seis)
    e64.backward_advance(case["cfg"]["nt"])
    ref64 = e64.field(3, f64=True)
    err = cases.rel_l2(res[fmt], res[0])
    print(f"strip format {fmt}: image rel L2 vs FP32 strips {err:.2e}, vs FP64 oracle "
          f"{cases.rel_l2(res[fmt], ref64):.2e} (FP32 strips vs FP64: {cases.rel_l2(res[0], ref64):.2e})")
    assert err < tol


def test_snapshot_checkpoints_fp32_exact():
    """Checkpoint snapshots stored in FP32: the backward pass replaces u(k) by
    the saved forward field at the checkpointed steps, exactly as the oracle
    does, so every field equals the FP32 replica run with the same spacing."""
    from rtm_gpu_amd import hip
    case = cases.build_case(name="ls_nfd4")
    seis = cases.synth_gather(case)
    nt = case["cfg"]["nt"]
    ctx = cases.hip_context(case, nshot=1, snap_every=50)
    ctx.set_gathers(seis[None])
    eng = cases.oracle_engine(case, snap_every=50)
    ctx.forward_begin()
    ctx.forward_advance(nt)
    eng.forward_begin()
    eng.forward_advance(nt)
    ctx.backward_begin()
    eng.backward_begin(seis)
    n2 = case["cfg"]["n2"]
    done = 0
    for step in (nt - 3 - 101, 1, 1, 49, nt):  # stop at u(102), u(101), u(100), u(51), end
        ctx.backward_advance(step)
        eng.backward_advance(step)
        done += step
        g = ctx.field(hip.FIELD_U)[n2:-n2, n2:-n2]
        assert np.array_equal(g, eng.field(0)[n2:-n2, n2:-n2]), f"u after {done} steps"
    for which, ow in ((hip.FIELD_I1, 3), (hip.FIELD_I2, 4)):
        assert np.array_equal(ctx.field(which), eng.field(ow))


@pytest.mark.parametrize("fmt", [1, 2])
def test_snapshot_checkpoints_reduced_precision(fmt):
    """FP16 / BF16 checkpoint snapshots (with FP16 / BF16 strips): the image
    stays close to FP32 storage; errors against the FP64 oracle run with the
    same checkpoints are printed."""
    from rtm_gpu_amd import hip
    case = cases.build_case(name="strips16", nt=300)
    seis = cases.synth_gather(case)
    nt = case["cfg"]["nt"]
    res = {}
    for f in (0, fmt):
        ctx = cases.hip_context(case, nshot=1, strip_format=f, snap_every=100)
        ctx.set_gathers(seis[None])
        ctx.migrate()
        res[f] = ctx.field(hip.FIELD_I1)
    e64 = cases.oracle_engine(case, prec=64, snap_every=100)
    e64.forward_begin()
    e64.forward_advance(nt)
    e64.backward_begin(seis)
    e64.backward_advance(nt)
    ref64 = e64.field(3, f64=True)
    err = cases.rel_l2(res[fmt], res[0])
    print(f"snapshots+strips format {fmt}: image rel L2 vs FP32 storage {err:.2e}, vs FP64 "
          f"oracle {cases.rel_l2(res[fmt], ref64):.2e} (FP32 vs FP64: "
          f"{cases.rel_l2(res[0], ref64):.2e})")
    assert err < (2e-3 if fmt == 1 else 1e-2)


@pytest.mark.parametrize("fmt", [1, 2])
def test_reduced_precision_subnormal_slots_stay_finite(fmt):
    """A wavelet scaled by 1e-38 puts every strip and snapshot maximum in the
    subnormal range: the per-slot power-of-two scale must keep 1/scale finite
    (it is clamped at 2^-126), so no stored value becomes inf or NaN."""
    from rtm_gpu_amd import hip
    case = dict(cases.build_case(name="ls_nfd4"))
    m = dict(case["model"])
    m["wav"] = (m["wav"].astype(np.float64) * 1e-38).astype(np.float32)
    case["model"] = m
    seis = cases.synth_gather(case)
    ctx = cases.hip_context(case, nshot=1, strip_format=fmt, snap_every=60)
    ctx.set_gathers(seis[None])
    ctx.forward_begin()
    ctx.forward_advance(case["cfg"]["nt"])
    k = case["cfg"]["nt"] - 1
    strip = ctx.strip(0, k)
    assert np.isfinite(strip).all() and np.abs(strip).max() > 0
    ctx.backward_begin()
    ctx.backward_advance(case["cfg"]["nt"] - 50)
    u = ctx.field(hip.FIELD_U)
    assert np.isfinite(u).all() and np.abs(u).max() > 0
    ctx.backward_advance(case["cfg"]["nt"])
    for which in (hip.FIELD_I1, hip.FIELD_I2, hip.FIELD_Q):
        assert np.isfinite(ctx.field(which)).all(), f"field {which}"
