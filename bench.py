#!/usr/bin/env python3
"""Headline benchmark: 2-D acoustic reverse-time migration on one MI355X.

Workload (the reverse-VSP configuration): 351 x 1701 interior grid at
12.5 m (411 x 1761 with the 30-point hybrid ABC), seeded faulted-layer model
with 8 dipping layers (1500-5500 m/s), 3 normal faults and a 1200 m x 600 m
4500 m/s body; 128 sources down a well at x = 850 (6.25 m apart), 1701
surface receivers, 20 Hz Ricker, dt = 1 ms, 4000 steps, adaptive LS
coefficients (nfdmax 10), compensated imaging, image stacked over shots.

Observed data are modelled on the true model by the same engine (forward
pass with receiver recording, untimed); migration uses a smoothed model.
One step = forward pass + backward pass with imaging + per-shot image post
and stack for the whole batch of sources, with inputs resident on the GPU.

  python bench.py [--steps K] [--warmup W] [--shots S] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Gpt-updates/s (%HBM peak), shots/s; time-to-5%-image-err: CNN vs plain LSRTM")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem(shots, nfdmax=10, n2=30, nt=4000):
    from rtm_gpu_amd import host, synth
    mod_nz, mod_nx, h, tao, f0 = 351, 1701, 12.5, 0.001, 20.0
    cfg = host.make_config(nfdmax=nfdmax, nfdmin=2, n2=n2, f0=f0, fmax=2 * f0 + 1, df=1.0,
                           nthita=1000, eps=1e-5, dv=1.0, ilste=0, ifv=0, whitecoe=1e-4, hz=h,
                           tao=tao, inorm=1, icompen=1, h=h, tao1=tao, mod_nz=mod_nz,
                           mod_nx=mod_nx, nt1=nt, s_l1=1, s_z1=1, n=mod_nx, ds=1, r_x1=851,
                           nrec=shots, dr=1)
    d = host.derive(cfg)
    true_raw = synth.faulted_rvsp(seed=2, mod_nz=mod_nz, mod_nx=mod_nx, h=h)
    mig_raw = synth.smooth(true_raw, 4.0)
    models = {}
    for name, raw in (("true", true_raw), ("mig", mig_raw)):
        t = time.time()
        v, r1 = host.pad_model(raw, mod_nz, mod_nx, n2, 0, tao, h)
        vmin, vmax, nvel = host.vrange(v, 1.0)
        need = host.bins_needed(v, vmin, 1.0, nvel)
        M, index, c = host.design_ls(1000, nfdmax, 2, nvel, tao, h, 1.0, 1e-5, cfg.fmax, vmin,
                                     1.0, need, d.hzx, nthreads=min(16, os.cpu_count() or 1))
        models[name] = dict(v=v, r1=r1, c=c, index=index, nvel=nvel, vmin=vmin, vmax=vmax,
                            nbins=int(need.sum()), design_s=time.time() - t)
        log(f"[bench] {name} model: {int(need.sum())} velocity bins, M in "
            f"{sorted(set(M[M > 0].tolist()))}, design {time.time() - t:.1f}s")
    depths = [6.25 * (m + 1) for m in range(shots)]
    r_u = [host.source_row(z, h, n2) for z in depths]
    r_x = [d.r_x] * shots
    wav = host.wavelet_table(d.nt, tao, f0)
    w = host.blend_weights(n2)
    return dict(cfg=cfg, d=d, models=models, r_u=r_u, r_x=r_x, wav=wav, w=w)


def build_apertures(nshot, seed=3, nt=5000):
    """Surface shots with uneven apertures on the seeded 4001 x 801 model
    (dx = dz = 10 m, Gaussian-random-field velocity 1500-4500 m/s with a
    500 m correlation length plus 6 sharp reflectors).  Shot s covers model
    columns [start, start + width) with width uniform in 600..2400, its source
    10 m deep at the window centre and a receiver at every surface point of
    the window.  15 Hz Ricker, dt = 0.8 ms, N2 = 30, nfdmax = 4."""
    from rtm_gpu_amd import host, synth
    mod_nz, mod_nx, h, tao, f0, n2, nfdmax = 801, 4001, 10.0, 0.0008, 15.0, 30, 4
    true_raw = synth.grf_4001(seed=seed, mod_nz=mod_nz, mod_nx=mod_nx, h=h)
    mig_raw = synth.smooth(true_raw, 4.0)
    rng = np.random.default_rng(seed)
    widths = rng.integers(600, 2401, nshot)
    starts = np.array([int(rng.integers(0, mod_nx - w + 1)) for w in widths])
    wmax = int(widths.max())
    cfg = host.make_config(nfdmax=nfdmax, nfdmin=2, n2=n2, f0=f0, fmax=2 * f0 + 1, df=1.0,
                           nthita=1000, eps=1e-5, dv=1.0, ilste=0, ifv=0, whitecoe=1e-4, hz=h,
                           tao=tao, inorm=1, icompen=1, h=h, tao1=tao, mod_nz=mod_nz,
                           mod_nx=wmax, nt1=nt, s_l1=1, s_z1=1, n=wmax, ds=1, r_x1=1,
                           nrec=nshot, dr=1)
    d = host.derive(cfg)
    # one coefficient table for both models and every window
    vt, _ = host.pad_model(true_raw, mod_nz, mod_nx, n2, 0, tao, h)
    vm, _ = host.pad_model(mig_raw, mod_nz, mod_nx, n2, 0, tao, h)
    vmin, vmax, nvel = host.vrange(np.concatenate([vt.ravel(), vm.ravel()]), 1.0)
    need = host.bins_needed(vt, vmin, 1.0, nvel) | host.bins_needed(vm, vmin, 1.0, nvel)
    M, index, c = host.design_ls(1000, nfdmax, 2, nvel, tao, h, 1.0, 1e-5, cfg.fmax, vmin, 1.0,
                                 need, d.hzx, nthreads=min(16, os.cpu_count() or 1))
    r_u = host.source_row(10.0, h, n2)
    geoms = [dict(nx=int(w) + 2 * n2, r_u=r_u, r_x=n2 + int(w) // 2, s_l=n2, s_r=n2 + int(w) - 1,
                  n=int(w), x_img=int(xb)) for w, xb in zip(widths, starts)]

    def window(raw, i):
        xb, w = int(starts[i]), int(widths[i])
        return host.pad_model(np.ascontiguousarray(raw[xb:xb + w]), mod_nz, w, n2, 0, tao, h)

    return dict(cfg=cfg, d=d, geoms=geoms, widths=widths, starts=starts, true=true_raw,
                mig=mig_raw, window=window, c=c, index=index, nvel=nvel, vmin=vmin, vmax=vmax,
                wav=host.wavelet_table(d.nt, tao, f0), w=host.blend_weights(n2), n2=n2,
                mod_nx=mod_nx, wmax=wmax)


def run_apertures(a):
    """Ragged batches of uneven-aperture shots against one shot per launch on a
    sample of the same shots; one JSON line.  Both run eager: each batch is
    migrated once, so a captured graph (rtm_set_graph, for repeated
    migrations of one batch) would only add its capture time."""
    from rtm_gpu_amd import hip
    t_setup = time.time()
    A = build_apertures(a.shots, nt=a.nt)
    d, cfg, n2 = A["d"], A["cfg"], A["n2"]
    log(f"[apertures] {a.shots} shots, widths {A['widths'].min()}..{A['widths'].max()}, "
        f"{A['nvel']} velocity bins, setup {time.time() - t_setup:.1f}s")
    common = dict(nz=d.nz, n2=n2, nfdmax=cfg.nfdmax, nt=d.nt, nt2=d.nt2, ilste=0, icompen=1,
                  s_z=d.s_z, ds=1, tao2=d.tao2, h2=d.h2, taoh=d.taoh, taoh2=d.taoh2,
                  hzx2_1=d.hzx2_1, strip_format={"f32": 0, "f16": 1, "bf16": 2}[a.strips])

    def table(ctx, nx):
        flat = np.full((d.nz, nx), A["vmin"], np.float32)
        ctx.set_model(flat, flat * 0, A["w"], A["c"], A["index"], A["nvel"], A["vmin"], 1.0)

    B = min(a.batch, a.shots)
    nxw = A["wmax"] + 2 * n2
    ctx = hip.Context(**common, nx=nxw, s_l=n2, s_r=n2 + A["wmax"] - 1, n=A["wmax"],
                      max_shots=B, record=1, stack_nx=A["mod_nx"], zseg=a.zseg)
    log(f"[apertures] batch context: {ctx.device_bytes() / 2**30:.1f} GiB")
    ctx.set_wavelet(A["wav"])
    timed, upd, zseg = 0.0, 0, a.zseg
    for b0 in range(0, a.shots, B):
        idx = list(range(b0, min(a.shots, b0 + B)))
        table(ctx, nxw)
        ctx.set_shot_geometry([A["geoms"][i] for i in idx])
        for j, i in enumerate(idx):
            v, r1 = A["window"](A["true"], i)
            ctx.set_shot_model(j, v, r1)
        ctx.forward_begin(save_strips=False)
        ctx.forward_advance(d.nt)
        ctx.gathers_from_records()
        for j, i in enumerate(idx):
            v, r1 = A["window"](A["mig"], i)
            ctx.set_shot_model(j, v, r1)
        if a.zseg <= 0 and b0 == 0:  # the tuner leaves the uploaded gathers alone
            zseg = ctx.autotune(8)
            log(f"[apertures] autotune: {zseg} rows per wave")
        ctx.sync()
        t0 = time.perf_counter()
        ctx.migrate()
        ctx.sync()
        dt = time.perf_counter() - t0
        timed += dt
        for i in idx:
            w = int(A["widths"][i])
            nxs, nI = w + 2 * n2, (d.nz - 2 * n2) * w
            upd += (d.nt - 2) * (2 * d.nz * nxs + nI)
        ctx.image_post(A["vmax"], cfg.whitecoe, want=False)
        log(f"[apertures] batch {b0 // B + 1}: {len(idx)} shots migrated in {dt:.2f}s")
    img, _ = ctx.stack_finish(a.shots, 1)
    assert np.isfinite(img).all(), "non-finite image"
    ctx.close()
    # one shot per launch: each sampled shot alone on a context of its own width
    sample = sorted(set(np.linspace(0, a.shots - 1, a.sample).astype(int).tolist()))
    single = 0.0
    for i in sample:
        gm = A["geoms"][i]
        one = hip.Context(**common, nx=gm["nx"], s_l=gm["s_l"], s_r=gm["s_r"], n=gm["n"],
                          max_shots=1, record=1, zseg=a.zseg)
        v, r1 = A["window"](A["true"], i)
        one.set_model(v, r1, A["w"], A["c"], A["index"], A["nvel"], A["vmin"], 1.0)
        one.set_wavelet(A["wav"])
        one.set_shots([gm["r_u"]], [gm["r_x"]])
        one.forward_begin(save_strips=False)
        one.forward_advance(d.nt)
        one.gathers_from_records()
        v, r1 = A["window"](A["mig"], i)
        one.set_model(v, r1, A["w"], A["c"], A["index"], A["nvel"], A["vmin"], 1.0)
        if a.zseg <= 0:  # each one-shot context tuned for its own width
            one.autotune(8)
        one.sync()
        t0 = time.perf_counter()
        one.migrate()
        one.sync()
        single += time.perf_counter() - t0
        one.close()
    single_sps = len(sample) / single
    value = a.shots / timed
    out = {"metric": "shots/s, ragged batch vs one shot per launch", "value": round(value, 3),
           "unit": "shots/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": round(timed * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded GRF model with reflectors; gathers modelled by the same engine)",
           "config": {"workload": f"grf_4001x801_apertures_{a.shots}shots", "nt": d.nt,
                      "dt": 0.0008, "f0": 15.0, "n2": n2, "nfdmax": cfg.nfdmax,
                      "aperture_columns": [int(A["widths"].min()), int(A["widths"].max())],
                      "shots_per_launch": B, "rows_per_wave": zseg, "strips": a.strips},
           "gpt_updates_per_s": round(upd / timed / 1e9, 3),
           "one_shot_per_launch": {"shots_per_s": round(single_sps, 3),
                                   "sample_shots": len(sample)},
           "speedup_vs_one_shot_per_launch": round(value / single_sps, 3)}
    print(json.dumps(out), flush=True)


def algorithmic_bytes(kind, d, n2, nfd, nshot, n):
    """Minimum HBM bytes one launch of each kernel must move.  Per interior
    cell and shot: forward / source reconstruction read u1, u0 and write u2
    (12 B); the receiver step reads q1, q0, u(k), sum u, sum q, I1, I2 and
    writes q2, sum u, sum q, I1, I2 (48 B).  The velocity model is shared by
    the shots of a launch (4 B per cell per launch)."""
    from rtm_gpu_amd import hip
    nz, nx = d.nz, d.nx
    nI = (nz - 2 * n2) * (nx - 2 * n2)
    nF = nz * nx - nI
    ss = 2 * nfd * ((nz - 2 * n2) + (nx - 2 * n2))
    if kind == hip.K_FWD_INTERIOR or kind == hip.K_SRC_INTERIOR:
        return nshot * nI * 12 + nI * 4
    if kind == hip.K_FWD_FRAME:
        return nshot * (nF * 12 + ss * 4) + nF * 4
    if kind == hip.K_STRIP_RESTORE:
        return nshot * ss * 8
    if kind == hip.K_RCV_INTERIOR:
        return nshot * (nI * 48 + n * 4) + nI * 4
    if kind == hip.K_RCV_FRAME:
        return nshot * (nF * 12 + n * 4) + nF * 4
    # the split backward step (compensated imaging): the reconstruction also
    # updates sum u and I2 (12 + 16 B); the receiver step reads q1, q0, sum u,
    # sum q, I1 and writes q2, sum q, I1 (32 B)
    if kind == hip.K_SRCI_INTERIOR:
        return nshot * nI * 28 + nI * 4
    if kind == hip.K_RCVI_INTERIOR:
        return nshot * (nI * 32 + n * 4) + nI * 4
    # Born: dU1, dU0, u(k), u(k-1), u(k-2) read, dU2 written (24 B) + v, m shared
    if kind == hip.K_BORN_INTERIOR:
        return nshot * nI * 24 + nI * 8
    # adjoint: mu1, mu0, u(k), u(k+1), u(k+2), G read, mu2, G written (36 B) + v shared
    if kind == hip.K_ADJ_INTERIOR:
        return nshot * (nI * 36 + n * 4) + nI * 4
    if kind in (hip.K_BORN_FRAME, hip.K_ADJ_FRAME):
        return nshot * (nF * 12 + n * 4) + nF * 4
    return 0


def pmc_traffic(kind_name, workload):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, with the
    summary's own record of when and on which kernel build it was measured;
    None when there is none for this kernel."""
    p = os.path.join(ROOT, "profiles", "pmc_summary.json")
    if not os.path.exists(p):
        return None, None
    try:
        with open(p) as f:
            s = json.load(f)
        e = s.get(workload, {}).get(kind_name)
        if not e:
            return None, None
        src = f"profiles/pmc_summary.json ({s.get('measured', 'date unrecorded')})"
        return float(e["hbm_bytes_per_launch"]), src
    except (OSError, ValueError, KeyError):
        return None, None


def host_cpus():
    """Cores this process may use, what the OS reports, and the CPU model."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return allowed, os.cpu_count(), model


def _oracle_cfg(P, shot, nthreads, record=0, strip_slots=0):
    import oracle_bind as ob
    d, m, cfgp = P["d"], P["models"]["mig"], P["cfg"]
    return ob.OrcCfg(nz=d.nz, nx=d.nx, n2=cfgp.n2, nfdmax=cfgp.nfdmax, nt=d.nt, nt2=d.nt2,
                     ilste=0, icompen=1, s_l=d.s_l, s_r=d.s_r, s_z=d.s_z, ds=1, n=cfgp.n,
                     r_u=P["r_u"][shot], r_x=d.r_x, tao2=d.tao2, h2=d.h2, taoh=d.taoh,
                     taoh2=d.taoh2, hzx2_1=d.hzx2_1, vmin=m["vmin"], dv=1.0, record=record,
                     nthreads=nthreads, strip_slots=strip_slots)


def cpu_baseline(P, target_s=12.0):
    """The FP32 CPU restatement (oracle/) on the host cores, in its two forms:
    throughput (one shot per pthread) and latency (one shot whose every step
    is split by rows over the pthreads, a barrier per step), each a bounded
    number of forward and backward steps on the full grid.  The reported
    value is the throughput form.  Threads: the cores this process may use,
    at most 16 (one GPU's share of the box's cores)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes as C
    import oracle_bind as ob
    d, m, cfgp = P["d"], P["models"]["mig"], P["cfg"]
    allowed, nproc, model = host_cpus()
    cores = max(1, min(16, allowed))
    seis = np.zeros(cfgp.n * d.nt, np.float32)
    args = [m["v"].ravel(), m["r1"].ravel(), P["w"], m["c"], len(m["c"]),
            m["index"].ctypes.data, m["nvel"], P["wav"], seis]
    nI = (d.nz - 2 * cfgp.n2) * (d.nx - 2 * cfgp.n2)
    per_pair_upd = 2 * d.nz * d.nx + nI
    # the timing sample keeps 64 strip slots (its backward steps only time the kernels)
    cfg = _oracle_cfg(P, len(P["r_u"]) // 2, 1, strip_slots=64)
    probe_steps = 3
    t_probe = ob.lib().orc_time_shots(C.byref(cfg), *args, 1, 1, probe_steps, probe_steps)
    per_pair = max(t_probe / probe_steps, 1e-3)
    steps = int(max(4, min(d.nt - 3, target_s / per_pair)))
    el = ob.lib().orc_time_shots(C.byref(cfg), *args, cores, cores, steps, steps)
    thr = cores * steps * per_pair_upd / el / 1e9
    # latency form: one shot, rows split over the threads
    steps_r = int(max(4, min(d.nt - 3, target_s * cores / per_pair / 4)))
    el_r = ob.lib().orc_time_rows(C.byref(cfg), *args, cores, steps_r, steps_r)
    lat = steps_r * per_pair_upd / el_r / 1e9
    shot_s = el_r / steps_r * (d.nt - 2)
    return {"value": thr, "unit": "Gpt-updates/s", "cores": cores, "kind": "port",
            "sample": f"FP32 restatement (oracle/), {cores} shots on {cores} pthreads, "
                      f"{steps} forward + {steps} backward steps each on the {d.nz}x{d.nx} "
                      f"grid, {el:.1f} s",
            "row_parallel": {"value": lat, "unit": "Gpt-updates/s", "threads": cores,
                             "s_per_shot": round(shot_s, 2),
                             "sample": f"one shot, {steps_r} forward + {steps_r} backward "
                                       f"steps with rows split over {cores} pthreads, "
                                       f"{el_r:.1f} s"},
            "host": {"threads_allowed": allowed, "nproc": nproc, "cpu_model": model}}


def correctness(P, nt_check=400):
    """One shot of the benched batch (the first source, in the ABC at row
    N2-1) migrated at truncated nt on the full grid, by the GPU and by the
    oracle (the FP32 restatement as the checker, and the FP64 one for the
    error): the GPU's image equals the FP32 oracle's bit for bit, and its
    relative L2 distance to the FP64 image is stated.  Run in the CPU leg."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    from rtm_gpu_amd import hip
    d, cfgp = P["d"], P["cfg"]
    mt, mm = P["models"]["true"], P["models"]["mig"]
    nt = min(nt_check, d.nt)
    nt2 = min(d.nt2, nt)
    ctx = hip.Context(nz=d.nz, nx=d.nx, n2=cfgp.n2, nfdmax=cfgp.nfdmax, nt=nt, nt2=nt2,
                      ilste=0, icompen=1, s_l=d.s_l, s_r=d.s_r, s_z=d.s_z, ds=1, n=cfgp.n,
                      tao2=d.tao2, h2=d.h2, taoh=d.taoh, taoh2=d.taoh2, hzx2_1=d.hzx2_1,
                      max_shots=1, record=1)
    wav = P["wav"][:nt + 1]
    ctx.set_wavelet(wav)
    ctx.set_shots(P["r_u"][:1], P["r_x"][:1])
    ctx.set_model(mt["v"], mt["r1"], P["w"], mt["c"], mt["index"], mt["nvel"], mt["vmin"], 1.0)
    ctx.forward_begin(save_strips=False)
    ctx.forward_advance(nt)
    seis = ctx.records()[0]                       # [n][nt]
    ctx.gathers_from_records()
    ctx.set_model(mm["v"], mm["r1"], P["w"], mm["c"], mm["index"], mm["nvel"], mm["vmin"], 1.0)
    ctx.migrate()
    gpu = ctx.field(hip.FIELD_I1)
    ctx.close()
    cores = max(1, min(16, host_cpus()[0]))
    res = {}
    for prec in (32, 64):
        cfg = _oracle_cfg(P, 0, cores)
        cfg.nt, cfg.nt2 = nt, nt2
        fields = {k: getattr(cfg, k) for k, _ in ob.OrcCfg._fields_}
        model = dict(v=mm["v"], r1=mm["r1"], w=P["w"], c=mm["c"], index=mm["index"],
                     nvel=mm["nvel"], wav=wav)
        e = ob.Engine(fields, model, prec=prec)
        e.forward_begin()
        e.forward_advance(nt)
        e.backward_begin(seis)
        e.backward_advance(nt)
        res[prec] = e.field(3, f64=(prec == 64))
    n2 = cfgp.n2
    g, r32, r64 = (x[n2:-n2, n2:-n2].astype(np.float64) for x in (gpu, res[32], res[64]))
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    return {"check": f"source 1 (row {P['r_u'][0]}), {nt} steps on the {d.nz}x{d.nx} grid, "
                     "image I1 over the interior",
            "image_bitexact_vs_fp32_oracle": bool(np.array_equal(g, r32)),
            "image_rel_l2_vs_fp32_oracle": rel(g, r32),
            "image_rel_l2_vs_fp64_oracle": rel(g, r64),
            "fp32_oracle_rel_l2_vs_fp64": rel(r32, r64)}


def copy_probe(mib):
    """A plain device-to-device copy of known size (read mib MiB, write mib
    MiB), run under rocprofv3 --pmc to calibrate FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = mib * 2 ** 20 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4}))


def run_lsrtm(a):
    """CGLS least-squares migration on the reverse-VSP batch (351 x 1701,
    128 shots, 4000 steps): Born data of the faulted model's reflectivity,
    then timed CGLS iterations (one Born and one adjoint pass each)."""
    from rtm_gpu_amd import deblur, hip, lsrtm, synth
    P = build_problem(a.shots, nt=a.nt)
    d, cfg = P["d"], P["cfg"]
    ctx = hip.Context(nz=d.nz, nx=d.nx, n2=cfg.n2, nfdmax=cfg.nfdmax, nt=d.nt, nt2=d.nt2,
                      ilste=0, icompen=1, s_l=d.s_l, s_r=d.s_r, s_z=d.s_z, ds=1, n=cfg.n,
                      tao2=d.tao2, h2=d.h2, taoh=d.taoh, taoh2=d.taoh2, hzx2_1=d.hzx2_1,
                      max_shots=a.shots, record=1, zseg=a.zseg)
    mm = P["models"]["mig"]
    ctx.set_wavelet(P["wav"])
    ctx.set_shots(P["r_u"], P["r_x"])
    ctx.set_model(mm["v"], mm["r1"], P["w"], mm["c"], mm["index"], mm["nvel"], mm["vmin"], 1.0)
    raw = synth.faulted_rvsp(seed=2, mod_nz=351, mod_nx=1701, h=12.5)
    m_true = deblur.reflectivity(raw, synth.smooth(raw, 4.0))
    zseg = ctx.autotune(8) if a.zseg <= 0 else a.zseg
    log(f"[lsrtm] autotune: {zseg} rows per wave")
    ctx.set_reflectivity(m_true)
    ctx.born_begin()
    ctx.born_advance(d.nt)
    ctx.gathers_from_records()  # observed data: Born data of the true reflectivity
    ctx.sync()
    log("[lsrtm] Born data modelled")
    op = lsrtm.BornOperator(ctx)
    rr0 = ctx.data_dot(hip.DATA_GATHERS, hip.DATA_GATHERS)
    marks = []

    def cb(it, m, h):
        ctx.sync()
        marks.append(time.perf_counter())
        log(f"[lsrtm] iteration {it}: ||r||^2/||d||^2 = {h['rr'] / rr0:.4f}")
    # Each iteration is one Born and one adjoint pass; the timed span runs from
    # the end of the first iteration to the end of the last (kernel timing on).
    niter = max(2, a.steps)
    ctx.set_timing(-1)
    _, hist = lsrtm.cgls(op, None, niter, callback=cb)
    el = (marks[-1] - marks[0]) / (len(marks) - 1)
    a.steps = len(marks) - 1
    nI = (d.nz - 2 * cfg.n2) * (d.nx - 2 * cfg.n2)
    upd_iter = (d.nt - 2) * (2 * d.nz * d.nx + nI + d.nz * d.nx) * a.shots
    value = upd_iter / el / 1e9
    kinds = {}
    for k, name in hip.KERNEL_NAMES.items():
        ms, n = ctx.kernel_time(k)
        if n:
            b = algorithmic_bytes(k, d, cfg.n2, cfg.nfdmax, a.shots, cfg.n)
            if k == hip.K_ADJ_FRAME:  # the adjoint frame launch carries the strip restore
                b += algorithmic_bytes(hip.K_STRIP_RESTORE, d, cfg.n2, cfg.nfdmax, a.shots, cfg.n)
                name = "k_frame<FR_ADJ>+k_strip_restore"
            kinds[name] = dict(avg_ms=round(ms / n, 5), launches=n,
                               GBps=round(b / (ms / n * 1e-3) / 1e9, 1) if b else None)
    ad = kinds.get("k_interior<EPI_ADJ>", {})
    out = {"metric": "least-squares RTM (CGLS): Gpt-updates/s", "value": round(value, 3),
           "unit": "Gpt-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": 1,
           "ms_per_step": round(el * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 (fp64 CG products)",
           "data": "synthetic (Born data of the seeded faulted model's reflectivity)",
           "config": {"workload": f"rvsp_faulted_351x1701_nt{d.nt}_{a.shots}shots_cgls",
                      "shots": a.shots, "rows_per_wave": zseg,
                      "passes_per_iteration": "Born (background + scattered) + adjoint "
                                              "(reconstruction + adjoint field)"},
           "s_per_iteration": round(el, 3),
           "misfit": [round(h["rr"] / rr0, 5) for h in hist],
           "roofline": {"bound": "hbm", "kernel": "k_interior<EPI_ADJ>",
                        "achieved": ad.get("GBps"), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(ad["GBps"] / HBM_PEAK_GBS, 4) if ad.get("GBps") else None,
                        "traffic": None, "avg_launch_ms": ad.get("avg_ms")},
           "kernels": kinds}
    ctx.close()
    print(json.dumps(out), flush=True)


def run_single(a):
    """The plain single-shot configuration: 200 x 400 interior at 10 m, 5
    seeded flat layers (1500-4500 m/s), 8th-order adaptive LS operators
    (nfdmax 4), 20-point hybrid ABC, 25 Hz Ricker at x = 200 and 30 m depth,
    dt = 1 ms, 2000 steps, a receiver at every surface point, observed data
    modelled by the engine.  A migration is ~1.6 ms of launches per time step
    pair at this size, so launch overhead matters: timed eagerly and as a
    replayed HIP graph of the whole time loop (rtm_set_graph)."""
    from rtm_gpu_amd import hip, host, synth
    mod_nz, mod_nx, h, tao, f0, n2, nfdmax, nt = 200, 400, 10.0, 0.001, 25.0, 20, 4, 2000
    cfg = host.make_config(nfdmax=nfdmax, nfdmin=2, n2=n2, f0=f0, fmax=2 * f0 + 1, df=1.0,
                           nthita=1000, eps=1e-5, dv=1.0, ilste=0, ifv=0, whitecoe=1e-4, hz=h,
                           tao=tao, inorm=1, icompen=1, h=h, tao1=tao, mod_nz=mod_nz,
                           mod_nx=mod_nx, nt1=nt, s_l1=1, s_z1=1, n=mod_nx, ds=1, r_x1=201,
                           nrec=1, dr=1)
    d = host.derive(cfg)
    raw = synth.layered_5(seed=1, mod_nz=mod_nz, mod_nx=mod_nx)
    v, r1 = host.pad_model(raw, mod_nz, mod_nx, n2, 0, tao, h)
    vmin, vmax, nvel = host.vrange(v, 1.0)
    need = host.bins_needed(v, vmin, 1.0, nvel)
    _, index, c = host.design_ls(1000, nfdmax, 2, nvel, tao, h, 1.0, 1e-5, cfg.fmax, vmin, 1.0,
                                 need, d.hzx)
    ctx = hip.Context(nz=d.nz, nx=d.nx, n2=n2, nfdmax=nfdmax, nt=d.nt, nt2=d.nt2, ilste=0,
                      icompen=1, s_l=d.s_l, s_r=d.s_r, s_z=d.s_z, ds=1, n=cfg.n, tao2=d.tao2,
                      h2=d.h2, taoh=d.taoh, taoh2=d.taoh2, hzx2_1=d.hzx2_1, max_shots=1,
                      record=1, zseg=a.zseg)
    ctx.set_model(v, r1, host.blend_weights(n2), c, index, nvel, vmin, 1.0)
    ctx.set_wavelet(host.wavelet_table(d.nt, tao, f0))
    ctx.set_shots([host.source_row(30.0, h, n2)], [d.r_x])
    ctx.forward_begin(save_strips=False)
    ctx.forward_advance(d.nt)
    ctx.gathers_from_records()
    zseg = ctx.autotune(8) if a.zseg <= 0 else a.zseg
    log(f"[single] {d.nz}x{d.nx}, nt {d.nt}; autotune: {zseg} rows per wave")
    steps = max(a.steps, 3)

    def timed(graph):
        ctx.set_graph(graph)
        for _ in range(max(a.warmup, 1)):  # the first graph migration captures
            ctx.migrate()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.migrate()
        ctx.sync()
        return (time.perf_counter() - t0) / steps

    t_eager = timed(False)
    img_eager = ctx.field(hip.FIELD_I1)
    t_graph = timed(True)
    same = bool(np.array_equal(ctx.field(hip.FIELD_I1), img_eager))
    ctx.set_graph(False)
    ctx.set_timing(-1)
    ctx.migrate()
    ctx.sync()
    kern = 0.0
    for k in hip.KERNEL_NAMES:
        ms, n = ctx.kernel_time(k)
        kern += ms
    nI = (d.nz - 2 * n2) * (d.nx - 2 * n2)
    upd = (d.nt - 2) * d.nz * d.nx + (d.nt - 2) * (nI + d.nz * d.nx)
    best = min(t_eager, t_graph)
    out = {"metric": METRIC, "value": round(upd / best / 1e9, 3), "unit": "Gpt-updates/s",
           "n_gpus": 1, "steps": steps, "warmup": max(a.warmup, 1),
           "ms_per_step": round(best * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded 5-layer model; gathers modelled by the same engine)",
           "config": {"workload": f"single_shot_layered_{mod_nz}x{mod_nx}_nt{nt}",
                      "grid": [d.nz, d.nx], "nt": d.nt, "shots": 1, "nfdmax": nfdmax, "n2": n2,
                      "rows_per_wave": zseg},
           "ms_per_migration": {"eager": round(t_eager * 1e3, 3),
                                "graph": round(t_graph * 1e3, 3),
                                "sum_of_kernel_times": round(kern, 3)},
           "graph_speedup": round(t_eager / t_graph, 3),
           "graph_image_equals_eager": same,
           "launches_per_migration": int(sum(ctx.kernel_time(k)[1] for k in hip.KERNEL_NAMES))}
    print(json.dumps(out), flush=True)


def run_cnn(a):
    """The deblurring network of the random-reflectivity configuration alone:
    six 3x3 layers of 32 channels, BF16 MFMA with FP32 accumulation, trained
    on random 64 x 64 patches in batches of 256 (Adam, lr 1e-3), and applied
    to a whole 256 x 1024 image.  Synthetic inputs (the kernels' speed does
    not depend on the values)."""
    from rtm_gpu_amd import cnn
    ch = [3] + [32] * 5 + [1]
    B, P, N = 256, 64, 2048
    net = cnn.Deblur(ch, B * P * P, B * P)
    net.init_params(0)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((N, 3, P, P)).astype(np.float32)
    t = rng.standard_normal((N, P, P)).astype(np.float32)
    for _ in range(max(a.warmup, 1)):
        net.fit(x, t, 1, B, 1e-3)
    epochs = max(a.steps, 1)
    t0 = time.perf_counter()
    net.fit(x, t, epochs, B, 1e-3)
    el = time.perf_counter() - t0
    steps = epochs * (N // B)
    mac_f = sum(ch[i] * ch[i + 1] * 9 for i in range(len(ch) - 1))
    mac_b = mac_f + sum(ch[i] * ch[i + 1] * 9 for i in range(1, len(ch) - 1))
    flop = 2.0 * B * P * P * (mac_f + mac_b)  # forward, weight and input gradients
    xi = rng.standard_normal((1, 3, 256, 1024)).astype(np.float32)
    net.forward(xi)
    t0 = time.perf_counter()
    for _ in range(5):
        net.forward(xi)
    t_inf = (time.perf_counter() - t0) / 5
    tf = flop / (el / steps) / 1e12
    out = {"metric": "deblurring network training step: TFLOP/s (BF16 MFMA)",
           "value": round(tf, 2), "unit": "TFLOP/s", "n_gpus": 1, "steps": steps,
           "warmup": max(a.warmup, 1), "ms_per_step": round(el / steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16 (fp32 accumulation)", "data": "synthetic",
           "config": {"workload": "cnn_6x32ch_3x3_patch64_batch256",
                      "channels": ch, "patch": P, "batch": B},
           "roofline": {"bound": "mfma", "achieved": round(tf, 2), "peak": 2500.0,
                        "unit": "TFLOP/s", "frac": round(tf / 2500.0, 4), "traffic": None},
           "inference_256x1024_ms": round(t_inf * 1e3, 3)}
    print(json.dumps(out), flush=True)


def run_deblur(a):
    """The random-reflectivity deblurring configuration (1024 x 256 at 10 m,
    64 shots, 25 Hz, 3000 steps, per-model backgrounds): the trained network
    (rtm_gpu_amd/data/deblur_cfg5.npz, from the two training stages of
    rtm_gpu_amd.deblur on 512 models) on held-out models against 10 CGLS
    iterations, 8 models per launch; time to an image within 5% relative L2
    of the CGLS(10) image, and the speed-up of the network's image."""
    from rtm_gpu_amd import deblur
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rtm_gpu_amd", "data",
                        "deblur_cfg5.npz")
    if not os.path.exists(path):
        sys.exit(f"{path} is missing: train it with python -m rtm_gpu_amd.deblur --cfg5-stage 0 "
                 "--save-net s0.npz, then --cfg5-stage 1 --load-net s0.npz --save-net <path>")
    net, c, norm = deblur.cfg5_load(path)
    with np.load(path, allow_pickle=False) as f:
        stages = int(f["stage"]) + 1
    S = deblur.cfg5_survey()
    neval = a.neval
    res = deblur.cfg5_evaluate(S, net, c, norm, neval=neval, panels=min(8, neval), cg_iters=10)
    t = res["s_per_launch"]
    t_net = t["image"] + t["net"]
    per = res["per_model"]
    within = [q for q in per if q["rel_net_vs_cg"] <= 0.05]
    cg5 = [q["cg_s_to_5pct"] for q in per if q["cg_s_to_5pct"] is not None]
    P = S["P"]
    out = {"metric": "time-to-5%-image-err: deblurring network vs 10-iteration CGLS (speed-up "
                     "of the network's image)",
           "value": round(t["cg"] / t_net, 3), "unit": "x", "n_gpus": 1,
           "steps": res["launches"], "warmup": 0, "ms_per_step": round(t_net * 1e3, 1),
           "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
           "dtype": "f32 stencil, bf16 network (fp32 accumulation), fp64 CG products",
           "data": "synthetic (seeded random-reflectivity models; Born data of each)",
           "config": {"workload": f"reflectors_{P['mod_nz']}x{P['mod_nx']}_{P['nshot']}shots_"
                                  f"nt{P['nt']}_heldout{neval}",
                      "models_per_launch": res["panels"], "cg_iterations": 10,
                      "network": "6 conv3x3 layers, 32 channels, ReLU, residual",
                      "training_models": 256 * stages, "training_stages": stages},
           "rel_l2_network_vs_cgls": round(res["rel_net_vs_cg"], 4),
           "rel_l2_scaled_image_vs_cgls": round(res["rel_rtm_vs_cg"], 4),
           "rel_l2_returned_vs_cgls": round(res["rel_out_vs_cg"], 4),
           "rel_l2_to_true_reflectivity": {"network": round(res["rel_net_vs_true"], 4),
                                           "scaled_image": round(res["rel_rtm_vs_true"], 4),
                                           "cgls10": round(res["rel_cg_vs_true"], 4)},
           "time_to_5pct_of_cgls10_s": {
               "network": round(t_net, 3) if len(within) == len(per) else None,
               "network_models_within_5pct": len(within),
               "cgls_mean": round(float(np.mean(cg5)), 3) if cg5 else None,
               "cgls_iterations": res["cg_iters_to_5pct"]},
           "data_residual": {"best_scaled_image": round(res["resid_rtm"], 4),
                             "network": round(res["resid_net"], 4),
                             "cgls10": round(res["resid_cg"], 4)},
           "fallback_chose_network": res["chose_net"],
           "s_per_launch": {k: round(v, 3) for k, v in t.items()}}
    print(json.dumps(out), flush=True)


def copy_bandwidth(gib=2, reps=5):
    """Device-to-device copy rate (read + write bytes / time) of a gib-GiB
    buffer, timed with HIP events: the bandwidth this GPU delivers to a plain
    streaming kernel at the moment, next to which the stencil's fraction of
    the 8 TB/s peak reads.  Uses the HIP runtime that librtm_hip.so loaded
    (a second runtime in the process would not see the GPU)."""
    import ctypes as C
    from rtm_gpu_amd import hip
    hip.lib()
    path = next(l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64" in l)
    rt = C.CDLL(path)
    vp = C.c_void_p
    rt.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
    rt.hipFree.argtypes = [vp]
    rt.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
    rt.hipEventCreate.argtypes = [C.POINTER(vp)]
    rt.hipEventRecord.argtypes = [vp, vp]
    rt.hipEventSynchronize.argtypes = [vp]
    rt.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
    rt.hipEventDestroy.argtypes = [vp]
    nb = gib * 2 ** 30
    a, b, e0, e1 = vp(), vp(), vp(), vp()
    ok = rt.hipMalloc(C.byref(a), nb) == 0 and rt.hipMalloc(C.byref(b), nb) == 0
    ok = ok and rt.hipEventCreate(C.byref(e0)) == 0 and rt.hipEventCreate(C.byref(e1)) == 0
    ms, ms_k = C.c_float(0.0), C.c_float(0.0)
    if ok:
        rt.hipMemcpy(b, a, nb, 3)  # device to device
        rt.hipEventRecord(e0, None)
        for _ in range(reps):
            rt.hipMemcpy(b, a, nb, 3)
        rt.hipEventRecord(e1, None)
        rt.hipEventSynchronize(e1)
        rt.hipEventElapsedTime(C.byref(ms), e0, e1)
        # the hand-written float4 streaming kernel on the same buffers
        lib = hip.lib()
        ok_k = lib.rtm_stream_copy(a, b, nb // 4, None) == 0
        rt.hipEventRecord(e0, None)
        for _ in range(reps):
            ok_k = ok_k and lib.rtm_stream_copy(a, b, nb // 4, None) == 0
        rt.hipEventRecord(e1, None)
        rt.hipEventSynchronize(e1)
        if ok_k:
            rt.hipEventElapsedTime(C.byref(ms_k), e0, e1)
    for p in (a, b):
        if p.value:
            rt.hipFree(p)
    for e in (e0, e1):
        if e.value:
            rt.hipEventDestroy(e)
    if not ok or ms.value <= 0:
        return None, None
    rate = lambda t: 2 * nb / (t / reps * 1e-3) / 1e9 if t > 0 else None  # noqa: E731
    return rate(ms.value), rate(ms_k.value)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run only a MIB-sized copy kernel (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=128)
    ap.add_argument("--nt", type=int, default=4000)
    ap.add_argument("--strips", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--zseg", type=int, default=0,
                    help="rows per wave of every interior kernel (0: autotune)")
    ap.add_argument("--zsegs", default="",
                    help="rows per wave of the forward, reconstruction and receiver-step "
                         "kernels, comma-separated (instead of the autotuner; for profiles)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fuse-frame", action="store_true",
                    help="run each forward and receiver step's interior and frame kernels as "
                         "one launch (rtm_params.fuse_frame)")
    ap.add_argument("--no-split", action="store_true",
                    help="keep the whole frame in the one-column kernel (rtm_params.frame_split "
                         "= 0; by default the split starts on and the autotuner keeps it only "
                         "where it is faster)")
    ap.add_argument("--host-gathers", action="store_true",
                    help="every timed step uploads the gathers from host memory (the PCIe-"
                         "inclusive rate of a caller that hands over host buffers)")
    ap.add_argument("--workload", default="rvsp",
                    choices=["rvsp", "single", "apertures", "deblur", "lsrtm", "cnn"],
                    help="rvsp: the headline reverse-VSP run; single: the single-shot "
                         "200 x 400 layered run, eager and as a replayed HIP graph; "
                         "apertures: uneven-aperture "
                         "surface shots in ragged batches vs one shot per launch; deblur: the "
                         "trained deblurring network against 10-iteration CGLS least-squares "
                         "migration on held-out random-reflectivity models (about 6 minutes "
                         "per 8 models); lsrtm: "
                         "CGLS iterations on the reverse-VSP batch (--steps iterations); cnn: "
                         "the deblurring network's training step and inference alone")
    ap.add_argument("--batch", type=int, default=128, help="apertures: shots per launch")
    ap.add_argument("--neval", type=int, default=8,
                    help="deblur: held-out models evaluated (8 per launch; 32 = the full set)")
    ap.add_argument("--sample", type=int, default=8,
                    help="apertures: shots timed one per launch")
    a = ap.parse_args()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {a.gpus} requested); shots shard across "
              "GPUs with one image all-reduce at the end, which this build does not run.",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "deblur":
        run_deblur(a)
        return
    if a.workload == "single":
        run_single(a)
        return
    if a.workload == "cnn":
        run_cnn(a)
        return
    if a.workload == "lsrtm":
        run_lsrtm(a)
        return
    if a.workload == "apertures":
        if a.shots == 128 and "--shots" not in sys.argv:
            a.shots = 256
        if a.nt == 4000 and "--nt" not in sys.argv:
            a.nt = 5000
        run_apertures(a)
        return

    from rtm_gpu_amd import hip
    t_setup = time.time()
    P = build_problem(a.shots, nt=a.nt)
    d, cfg = P["d"], P["cfg"]
    strip_fmt = {"f32": 0, "f16": 1, "bf16": 2}[a.strips]
    ctx = hip.Context(nz=d.nz, nx=d.nx, n2=cfg.n2, nfdmax=cfg.nfdmax, nt=d.nt, nt2=d.nt2,
                      ilste=0, icompen=1, s_l=d.s_l, s_r=d.s_r, s_z=d.s_z, ds=1, n=cfg.n,
                      tao2=d.tao2, h2=d.h2, taoh=d.taoh, taoh2=d.taoh2, hzx2_1=d.hzx2_1,
                      max_shots=a.shots, strip_format=strip_fmt, record=1, zseg=a.zseg,
                      fuse_frame=1 if a.fuse_frame else 0, frame_split=0 if a.no_split else 1)
    log(f"[bench] context: {ctx.device_bytes() / 2**30:.1f} GiB on device")
    mt, mm = P["models"]["true"], P["models"]["mig"]
    ctx.set_wavelet(P["wav"])
    ctx.set_shots(P["r_u"], P["r_x"])
    # observed data: forward modelling with receiver recording on the true model
    ctx.set_model(mt["v"], mt["r1"], P["w"], mt["c"], mt["index"], mt["nvel"], mt["vmin"], 1.0)
    ctx.forward_begin(save_strips=False)
    ctx.forward_advance(d.nt)
    ctx.gathers_from_records()
    ctx.sync()
    ctx.set_model(mm["v"], mm["r1"], P["w"], mm["c"], mm["index"], mm["nvel"], mm["vmin"], 1.0)
    zseg = a.zseg
    if a.zsegs:  # forward, reconstruction, receiver step[, frame top/bottom, frame sides]
        zs = [int(v) for v in a.zsegs.split(",")]
        zseg = zs[:3]
        ctx.set_zseg(*zseg)
        if len(zs) == 5:
            ctx.set_frame_runs(zs[3], zs[4])
    elif zseg <= 0:
        t = time.time()
        zseg = ctx.autotune(8)
        log(f"[bench] autotune (forward, reconstruction, receiver step): {zseg} rows per wave; "
            f"frame runs (top/bottom, sides) {ctx.frame_runs()} ({time.time() - t:.1f}s)")
    log(f"[bench] setup + modelling {time.time() - t_setup:.1f}s")

    host_seis = ctx.records() if a.host_gathers else None  # [shot][n][nt], trace-major

    def step():
        if host_seis is not None:
            ctx.set_gathers(host_seis)
        ctx.migrate()
        ctx.image_post(mm["vmax"], cfg.whitecoe, want=False)

    for i in range(a.warmup):
        t = time.time()
        step()
        ctx.sync()
        log(f"[bench] warmup {i + 1}/{a.warmup}: {time.time() - t:.2f}s")
    ctx.set_timing(-1)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(a.steps):
        step()
        ctx.sync()
        log(f"[bench] step {i + 1}/{a.steps} done at {time.perf_counter() - t0:.2f}s")
    ctx.sync()
    el = time.perf_counter() - t0
    copy_gbs, stream_gbs = copy_bandwidth()
    log(f"[bench] device copy right after the timed steps: {copy_gbs or 0:.0f} GB/s (runtime), "
        f"{stream_gbs or 0:.0f} GB/s (float4 streaming kernel)")
    img, _ = ctx.stack_finish(a.shots * (a.warmup + a.steps), 1)
    assert np.isfinite(img).all(), "non-finite image"

    nI = (d.nz - 2 * cfg.n2) * (d.nx - 2 * cfg.n2)
    steps_f = d.nt - 2
    upd_per_shot = steps_f * d.nz * d.nx + steps_f * (nI + d.nz * d.nx)
    value = upd_per_shot * a.shots * a.steps / el / 1e9
    kinds = {}
    # With --fuse-frame a step's interior and frame work run as one launch
    # (k_step): that launch is timed under the interior kind and moves the
    # algorithmic bytes of both.
    fusedk = {} if not a.fuse_frame else {
        hip.K_FWD_INTERIOR: (hip.K_FWD_FRAME, "k_step<EPI_FWD+FR_FWD>"),
        hip.K_RCVI_INTERIOR: (hip.K_RCV_FRAME, "k_step<EPI_RCVI+FR_RCV>")}
    # Without a checkpoint snapshot (none here) and with separate launches, the
    # receiver frame launch of every backward step also restores the strips of
    # the reconstructed field (k_strip_restore's work and bytes).
    restore_in_frame = not a.fuse_frame
    for k, name in hip.KERNEL_NAMES.items():
        try:
            ms, n = ctx.kernel_time(k)
        except hip.RtmError:  # a kind this build of the library does not time
            continue
        if n:
            b = algorithmic_bytes(k, d, cfg.n2, cfg.nfdmax, a.shots, cfg.n)
            if k in fusedk and ctx.kernel_time(fusedk[k][0])[1] == 0:
                b += algorithmic_bytes(fusedk[k][0], d, cfg.n2, cfg.nfdmax, a.shots, cfg.n)
                name = fusedk[k][1]
            if k == hip.K_RCV_FRAME and restore_in_frame:
                b += algorithmic_bytes(hip.K_STRIP_RESTORE, d, cfg.n2, cfg.nfdmax, a.shots, cfg.n)
                name = "k_frame<FR_RCV>+k_strip_restore"
            kinds[name] = dict(total_ms=ms, launches=n, avg_ms=ms / n,
                               gbs=b / (ms / n * 1e-3) / 1e9, bytes_per_launch=b)
    dom = max(kinds, key=lambda k: kinds[k]["total_ms"])
    kd = kinds[dom]
    workload = f"rvsp_faulted_351x1701_nt{d.nt}_{a.shots}shots"
    total_bytes = sum(v["bytes_per_launch"] * v["launches"] for v in kinds.values())
    traffic, traffic_src = pmc_traffic(dom, f"rvsp_faulted_351x1701_{a.shots}shots")
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpt-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(el / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded faulted-layer model; gathers modelled by the same engine)"
                + ("; gathers uploaded from host memory in every timed step" if a.host_gathers
                   else ""),
        "config": {"workload": workload, "grid": [d.nz, d.nx], "interior": [d.nz - 60, d.nx - 60],
                   "nt": d.nt, "shots": a.shots, "nfdmax": cfg.nfdmax, "n2": cfg.n2,
                   "strips": a.strips, "coefficients": "adaptive LS",
                   "velocity_bins": mm["nbins"], "rows_per_wave": zseg,
                   "frame_runs": list(ctx.frame_runs()),
                   "frame_launch": "one launch with the interior" if a.fuse_frame else "separate",
                   "frame_split": ("off" if a.no_split else
                                   "on" if a.zsegs or a.zseg > 0 else "autotuned")},
        "shots_per_s": round(a.shots * a.steps / el, 3),
        "hbm_frac_whole_job": round(total_bytes / el / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(kd["gbs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kd["gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": kd["bytes_per_launch"],
                     "avg_launch_ms": round(kd["avg_ms"], 5),
                     "copy_GBps_same_run": round(copy_gbs, 1) if copy_gbs else None,
                     "stream_kernel_GBps_same_run": round(stream_gbs, 1) if stream_gbs else None,
                     "frac_of_copy": round(kd["gbs"] / copy_gbs, 4) if copy_gbs else None},
        "kernels": {k: {"avg_ms": round(v["avg_ms"], 5), "launches": v["launches"],
                        "GB/s": round(v["gbs"], 1)} for k, v in kinds.items()},
    }
    ctx.close()
    if not a.no_cpu:
        log("[bench] correctness check (one shot, truncated nt, against the oracle) ...")
        out["correctness"] = correctness(P)
        log(f"[bench] correctness: {out['correctness']}")
        log("[bench] CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(P)
    else:
        out["correctness"] = None
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
