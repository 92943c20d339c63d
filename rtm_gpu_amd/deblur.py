"""Learned image-domain deblurring of migrated images, against least-squares
migration.  Not in the reference, which stops at the migrated image
(kernel.cu:935-1059).

The migrated image of Born data is H m = L^T L m: the reflectivity m blurred
by the Hessian of Born modelling.  Least-squares migration (lsrtm.cgls)
undoes the blur iteratively, at two wave-equation passes per iteration.  A
small convolutional network (cnn.Deblur) learns an approximate inverse of H
in the image domain instead, from pairs the engine makes itself:
random reflectivity m_i -> Born data d_i = L m_i -> migrated image
x_i = L^T d_i, many models per launch (each panel of a wide reflectivity is
one group of shots imaging its own columns, with its own background window
and its own CG step lengths).  Output m ~ c x + net(x) (c: the least-squares
scalar between the training images and reflectivities; a zero network is the
scaled migrated image).

Two settings:

  build_survey / run   one background (the smoothed faulted model) for every
                       pair; inputs c x, log source illumination and depth;
                       targets the CGLS images of the pairs or m itself.
  cfg5_*               the random-reflectivity configuration: 1024 x 256 at
                       10 m, every model with its own background (linear
                       1500-4000 m/s gradient plus a Gaussian random field)
                       and 10-40 dipping reflectors (synth.reflectors), 64
                       surface shots, 25 Hz, 3000 steps; inputs c x, the
                       normalised background velocity and depth; targets m;
                       six 3x3 layers of 32 channels trained on 64 x 64
                       patches, batch 256, Adam 1e-3; evaluated on held-out
                       models against 10 CGLS iterations.

guard() is the fallback check: the network's estimate is returned only when
its Born re-modelled data fit the observed data better than both zero and the
best multiple of the migrated image; otherwise that multiple is returned."""
import time

import numpy as np

from . import cnn, hip, host, lsrtm, synth


def log(*a):
    print(*a, flush=True)


def build_survey(mod_nz=200, mod_nx=400, h=10.0, tao=0.001, f0=15.0, nt=2000, nshot=12,
                 nfdmax=6, n2=30, seed=2, src_depth=10.0):
    """Surface survey on the seeded faulted model (column-major [x][z]): nshot
    sources 10 m deep, evenly spaced, a receiver at every surface point, the
    smoothed model as migration background."""
    cfg = host.make_config(nfdmax=nfdmax, nfdmin=2, n2=n2, f0=f0, fmax=2 * f0 + 1, df=1.0,
                           nthita=1000, eps=1e-5, dv=1.0, ilste=0, ifv=0, whitecoe=1e-4, hz=h,
                           tao=tao, inorm=1, icompen=1, h=h, tao1=tao, mod_nz=mod_nz,
                           mod_nx=mod_nx, nt1=nt, s_l1=1, s_z1=1, n=mod_nx, ds=1,
                           r_x1=mod_nx // 2, nrec=nshot, dr=1)
    d = host.derive(cfg)
    true_raw = synth.faulted_rvsp(seed=seed, mod_nz=mod_nz, mod_nx=mod_nx, h=h)
    mig_raw = synth.smooth(true_raw, 4.0)
    v, r1 = host.pad_model(mig_raw, mod_nz, mod_nx, n2, 0, tao, h)
    vmin, vmax, nvel = host.vrange(v, 1.0)
    need = host.bins_needed(v, vmin, 1.0, nvel)
    _, index, c = host.design_ls(1000, nfdmax, 2, nvel, tao, h, 1.0, 1e-5, cfg.fmax, vmin, 1.0,
                                 need, d.hzx)
    cols = np.linspace(0, mod_nx - 1, nshot + 2)[1:-1].round().astype(int)
    r_u = [host.source_row(src_depth, h, n2)] * nshot
    r_x = [int(d.s_l + cx) for cx in cols]
    return dict(cfg=cfg, d=d, v=v, r1=r1, c=c, index=index, nvel=nvel, vmin=vmin, vmax=vmax,
                wav=host.wavelet_table(d.nt, tao, f0), w=host.blend_weights(n2), r_u=r_u,
                r_x=r_x, nshot=nshot, mod_nz=mod_nz, mod_nx=mod_nx, n2=n2, seed=seed,
                m_true=reflectivity(true_raw, mig_raw))


def reflectivity(raw, raw_smooth):
    """Born scattering strength (v^2 - v0^2) / v0^2 on the image grid [z][x]."""
    v = raw.astype(np.float64).T
    v0 = raw_smooth.astype(np.float64).T
    return ((v * v - v0 * v0) / (v0 * v0)).astype(np.float32)


def random_reflectivity(S, seed):
    raw = synth.faulted_rvsp(seed=seed, mod_nz=S["mod_nz"], mod_nx=S["mod_nx"],
                             h=float(S["cfg"].h))
    return reflectivity(raw, synth.smooth(raw, 4.0))


def context(S, panels=1):
    """Context of `panels` copies of the survey's shots; copy p images columns
    [p*mnx, (p+1)*mnx) of a reflectivity panels * mnx wide."""
    d, cfg = S["d"], S["cfg"]
    ns = S["nshot"] * panels
    ctx = hip.Context(nz=d.nz, nx=d.nx, n2=cfg.n2, nfdmax=cfg.nfdmax, nt=d.nt, nt2=d.nt2,
                      ilste=0, icompen=1, s_l=d.s_l, s_r=d.s_r, s_z=d.s_z, ds=1, n=cfg.n,
                      tao2=d.tao2, h2=d.h2, taoh=d.taoh, taoh2=d.taoh2, hzx2_1=d.hzx2_1,
                      max_shots=ns, record=1, stack_nx=S["mod_nx"] * panels)
    ctx.set_model(S["v"], S["r1"], S["w"], S["c"], S["index"], S["nvel"], S["vmin"], 1.0)
    ctx.set_wavelet(S["wav"])
    if panels == 1:
        ctx.set_shots(S["r_u"], S["r_x"])
    else:
        ctx.set_shot_geometry([dict(nx=d.nx, r_u=S["r_u"][s], r_x=S["r_x"][s], s_l=d.s_l,
                                    s_r=d.s_r, n=cfg.n, x_img=p * S["mod_nx"])
                               for p in range(panels) for s in range(S["nshot"])])
    return ctx


def normal_images(ctx, m_panels):
    """H m for each panel (shots of one panel image their own columns)."""
    npan, mnz, mnx = m_panels.shape
    ctx.set_reflectivity(np.ascontiguousarray(m_panels.transpose(1, 0, 2).reshape(mnz, npan * mnx)))
    ctx.born_begin()
    ctx.born_advance(ctx.p.nt)
    ctx.gathers_from_records()
    ctx.adjoint_begin()
    ctx.adjoint_advance(ctx.p.nt)
    g = ctx.adjoint_gradient()
    return np.ascontiguousarray(g.reshape(mnz, npan, mnx).transpose(1, 0, 2))


def illumination(S):
    """Source illumination sum_k u(k)^2 of the survey's shots, stacked [z][x]."""
    ctx = context(S)
    ctx.set_gathers(np.zeros((S["nshot"], S["cfg"].n, S["d"].nt), np.float32))
    ctx.migrate()
    _, down = ctx.image_post(float(S["vmax"]), 0.0)
    return down.sum(axis=0)


def features(S, img, c, illum):
    """Network input [n][3][z][x]: c * image, log illumination in [0, 1], depth."""
    img = np.asarray(img, np.float32)
    if img.ndim == 2:
        img = img[None]
    n, mnz, mnx = img.shape
    li = np.log10(np.maximum(illum / illum.max(), 1e-6)) / 6.0 + 1.0
    depth = np.broadcast_to(np.linspace(0.0, 1.0, mnz, dtype=np.float32)[:, None], (mnz, mnx))
    x = np.empty((n, 3, mnz, mnx), np.float32)
    x[:, 0] = np.float32(c) * img
    x[:, 1] = li
    x[:, 2] = depth
    return x


def make_pairs(S, npairs, panels, seed0=100, target="cg", cg_iters=10):
    """Inputs H m_i and targets (CGLS images of L m_i, or m_i) [npairs][mnz][mnx]."""
    ctx = context(S, panels)
    ms, xs = [], []
    for b in range(0, npairs, panels):
        k = min(panels, npairs - b)
        m = np.stack([random_reflectivity(S, seed0 + b + i) for i in range(panels)])
        if k < panels:
            m[k:] = 0.0
        x = normal_images(ctx, m)
        if target == "cg":
            # the gathers still hold the panels' Born data (normal_images)
            m, _ = lsrtm.cgls(lsrtm.BornOperator(ctx, groups=panels), None, cg_iters,
                              callback=lambda it, *_: log(f"[deblur] pairs {b}..{b + k - 1}: "
                                                          f"CG iteration {it}"))
        ms.append(m[:k])
        xs.append(x[:k])
        log(f"[deblur] pairs {b + k} of {npairs} done")
    return np.concatenate(xs), np.concatenate(ms)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


class _Iterates:
    """CGLS callback keeping each iterate and the time it was reached."""

    def __init__(self, note=None):
        self.m, self.t, self.t0, self.note = [], [], time.time(), note

    def __call__(self, it, m, h):
        self.m.append(np.array(m, copy=True))
        self.t.append(time.time() - self.t0)
        if self.note:
            log(f"{self.note} {it}")

    def first_within(self, final, tol=0.05, axis_models=False):
        """Iteration (1-based) and time at which the iterate first comes within
        tol relative L2 of `final` (per model when axis_models)."""
        if not axis_models:
            for i, m in enumerate(self.m):
                if rel(m, final) <= tol:
                    return i + 1, self.t[i]
            return None, None
        out = []
        for g in range(final.shape[0]):
            hit = next((i for i, m in enumerate(self.m) if rel(m[g], final[g]) <= tol), None)
            out.append((None, None) if hit is None else (hit + 1, self.t[hit]))
        return out


def patches(XF, M, size, count, rng):
    """count random size x size crops (all channels) of the pairs."""
    n, c, hh, ww = XF.shape
    xs = np.empty((count, c, size, size), np.float32)
    ms = np.empty((count, size, size), np.float32)
    k = rng.integers(0, n, count)
    z = rng.integers(0, hh - size + 1, count)
    x = rng.integers(0, ww - size + 1, count)
    for i in range(count):
        xs[i] = XF[k[i], :, z[i]:z[i] + size, x[i]:x[i] + size]
        ms[i] = M[k[i], z[i]:z[i] + size, x[i]:x[i] + size]
    return xs, ms


def born_residual(ctx, m, dobs):
    """Relative misfit ||L m - d|| / ||d|| of the Born re-modelled data of the
    reflectivity m (one Born pass on ctx; its records are overwritten)."""
    ctx.set_reflectivity(m)
    r = ctx.born() - dobs
    return float(np.linalg.norm(r.astype(np.float64)) / np.linalg.norm(dobs.astype(np.float64)))


def data_fit(op, m):
    """One Born pass of m: (<L m, L m>, <L m, d>) per group, d the gathers."""
    qq = np.atleast_1d(op.forward(m)).astype(np.float64)
    qd = np.atleast_1d(op.dots(hip.DATA_RECORDS, hip.DATA_GATHERS)).astype(np.float64)
    return qq, qd


def guard(op, img, m_net):
    """The fallback check, per group of op (lsrtm.BornOperator whose gathers
    hold the observed data d): alpha = <L img, d> / ||L img||^2 scales the
    migrated image img to its best data fit, misfit r_rtm =
    ||alpha L img - d|| / ||d|| (at most 1, the misfit of zero); the network's
    estimate m_net is kept only when its misfit r_net = ||L m_net - d|| / ||d||
    is finite and below both r_rtm and 1.  Two Born passes (one if m_net is
    not finite).  Returns (m, choice, r_net, r_rtm, alpha): arrays over the
    groups when op has groups > 1, scalars otherwise."""
    G = op.groups
    img = np.asarray(img, np.float32)
    m_net = np.asarray(m_net, np.float32)
    dd = np.atleast_1d(op.dots(hip.DATA_GATHERS, hip.DATA_GATHERS)).astype(np.float64)
    qq, qd = data_fit(op, img)
    alpha = np.where(qq > 0, qd / np.where(qq > 0, qq, 1.0), 0.0)
    safe = np.where(dd > 0, dd, 1.0)
    r_rtm = np.sqrt(np.maximum(0.0, 1.0 - alpha * qd / safe))
    fin = np.array([np.isfinite(g).all() for g in m_net.reshape(G, -1)])
    r_net = np.full(G, np.inf)
    if fin.any():
        probe = np.where(fin.reshape((G,) + (1,) * (m_net.ndim - 1)) if G > 1 else fin[0],
                         m_net, np.float32(0))
        nq, nd = data_fit(op, probe)
        r_net = np.where(fin, np.sqrt(np.maximum(0.0, nq - 2.0 * nd + dd) / safe), np.inf)
    use = fin & (dd > 0) & (r_net < np.minimum(r_rtm, 1.0))
    a32 = alpha.astype(np.float32)
    if G == 1:
        m = m_net if use[0] else a32[0] * img
        return m, ("net" if use[0] else "rtm"), float(r_net[0]), float(r_rtm[0]), float(alpha[0])
    m = np.where(use[:, None, None], m_net, a32[:, None, None] * img)
    return m, np.where(use, "net", "rtm"), r_net, r_rtm, alpha


def run(S=None, npairs=128, panels=32, epochs=600, batch=8, lr=1e-3,
        channels=(3,) + (16,) * 11 + (1,), dilation=(1, 2, 4, 8, 16, 32, 16, 8, 4, 2, 1, 1),
        cg_iters=10, seed=0, target="cg", save_net=""):
    """Train on engine-made pairs and compare with CGLS; returns a results dict
    and the images; save_net: write the trained network there (cnn.Deblur.save)."""
    S = S or build_survey()
    res = dict(grid=f"{S['mod_nz']}x{S['mod_nx']}", nt=int(S["d"].nt), shots=S["nshot"],
               npairs=npairs, epochs=epochs, channels=list(channels),
               dilation=list(dilation) if dilation else None, target=target)
    t = time.time()
    X, M = make_pairs(S, npairs, panels, target=target, cg_iters=cg_iters)
    res["pairs_s"] = time.time() - t
    c = float(np.sum(X.astype(np.float64) * M) / np.sum(X.astype(np.float64) ** 2))
    res["scale"] = c
    illum = illumination(S)
    XF = features(S, X, c, illum)
    mnz, mnx = M.shape[1:]
    net = cnn.Deblur(list(channels), batch * mnz * mnx, batch * mnz, dilation=dilation)
    net.init_params(seed=seed)
    t = time.time()
    hist = []
    for e0 in range(0, epochs, 50):  # chunks, so progress shows; Adam state carries over
        hist.extend(net.fit(XF, M, epochs=min(50, epochs - e0), batch=batch, lr=lr,
                            seed=seed + e0))
        log(f"[deblur] epoch {e0 + min(50, epochs - e0)} of {epochs}: loss {hist[-1]:.3e}")
    res["train_s"] = time.time() - t
    res["loss_first"], res["loss_last"] = float(hist[0]), float(hist[-1])
    res["loss_identity"] = float(np.mean((XF[:, 0] - M) ** 2))
    if save_net:
        net.save(save_net)
    log(f"[deblur] pairs {res['pairs_s']:.1f}s, training {res['train_s']:.1f}s, loss "
        f"{res['loss_identity']:.3e} (scaled image) -> {res['loss_last']:.3e}")

    # the evaluated model: Born data of its reflectivity, then three estimates
    ctx = context(S)
    op = lsrtm.BornOperator(ctx)
    m_true = S["m_true"]
    ctx.set_reflectivity(m_true)
    dobs = ctx.born()
    hip_sync = ctx.sync
    t = time.time()
    ctx.forward_begin(True)
    ctx.forward_advance(ctx.p.nt)
    img = ctx.adjoint(dobs)
    hip_sync()
    t_img = time.time() - t
    t = time.time()
    m_net = net.forward(features(S, img, c, illum))[0]
    t_net = time.time() - t
    m_rtm = np.float32(c) * img
    t = time.time()  # the fallback check (two Born passes); the gathers still hold dobs
    _, chosen, r_net, r_rtm, _ = guard(op, img, m_net)
    hip_sync()
    t_guard = time.time() - t
    t = time.time()
    iters = _Iterates()
    m_cg, ch = lsrtm.cgls(op, dobs, cg_iters, callback=iters)
    t_cg = time.time() - t
    k5, t5 = iters.first_within(m_cg)

    res.update(
        t_image_s=t_img, t_net_s=t_net, t_cg_s=t_cg,
        speedup_vs_cg=t_cg / (t_img + t_net),
        rel_net_vs_cg=rel(m_net, m_cg), rel_rtm_vs_cg=rel(m_rtm, m_cg),
        rel_net_vs_true=rel(m_net, m_true), rel_rtm_vs_true=rel(m_rtm, m_true),
        rel_cg_vs_true=rel(m_cg, m_true),
        resid_rtm=r_rtm, resid_net=r_net, resid_cg=born_residual(ctx, m_cg, dobs),
        chosen=chosen, t_guard_s=t_guard, speedup_guarded_vs_cg=t_cg / (t_img + t_net + t_guard),
        cg_rr=[h["rr"] / ch[0]["rr"] for h in ch],
        cg_iters_to_5pct=k5, cg_time_to_5pct_s=t5,
        net_within_5pct=bool(rel(m_net, m_cg) <= 0.05))
    log(f"[deblur] image {t_img:.2f}s + net {t_net:.3f}s vs CGLS({cg_iters}) {t_cg:.2f}s; "
        f"rel L2 to CGLS: net {res['rel_net_vs_cg']:.3f}, scaled RTM {res['rel_rtm_vs_cg']:.3f}; "
        f"data residual: best-scaled RTM {res['resid_rtm']:.3f}, net {res['resid_net']:.3f}, "
        f"CGLS {res['resid_cg']:.3f}; returned: {chosen} (check {t_guard:.2f}s)")
    return res, dict(m_true=m_true, m_rtm=m_rtm, m_net=m_net, m_cg=m_cg)


# ---------------------------------------------------------------------------
# The random-reflectivity configuration: 512 training and 32 held-out models

CFG5 = dict(mod_nz=256, mod_nx=1024, h=10.0, tao=0.001, f0=25.0, nt=3000, nshot=64, nfdmax=4,
            n2=30, vlo=1500.0, vhi=4000.0, src_depth=10.0)
CFG5_CHANNELS = (3,) + (32,) * 5 + (1,)  # six 3x3 layers of 32 channels, residual on channel 0
CFG5_TRAIN_SEED, CFG5_EVAL_SEED = 1000, 5000  # training models 1000.., held-out 5000..


def cfg5_survey(nt=None, nshot=None):
    """Geometry and coefficient table shared by every model: nshot surface
    sources 10 m deep, evenly spaced, a receiver at every surface point; one
    LS coefficient table over every 1 m/s bin of [vlo, vhi], so any
    background's cells find their row."""
    P = dict(CFG5)
    if nt:
        P["nt"] = nt
    if nshot:
        P["nshot"] = nshot
    mod_nz, mod_nx, h, tao, n2 = P["mod_nz"], P["mod_nx"], P["h"], P["tao"], P["n2"]
    cfg = host.make_config(nfdmax=P["nfdmax"], nfdmin=2, n2=n2, f0=P["f0"], fmax=2 * P["f0"] + 1,
                           df=1.0, nthita=1000, eps=1e-5, dv=1.0, ilste=0, ifv=0, whitecoe=1e-4,
                           hz=h, tao=tao, inorm=1, icompen=1, h=h, tao1=tao, mod_nz=mod_nz,
                           mod_nx=mod_nx, nt1=P["nt"], s_l1=1, s_z1=1, n=mod_nx, ds=1,
                           r_x1=mod_nx // 2, nrec=P["nshot"], dr=1)
    d = host.derive(cfg)
    vmin, vmax, nvel = host.vrange(np.array([P["vlo"], P["vhi"]], np.float32), 1.0)
    need = np.ones(nvel, bool)
    _, index, c = host.design_ls(1000, P["nfdmax"], 2, nvel, tao, h, 1.0, 1e-5, cfg.fmax, vmin,
                                 1.0, need, d.hzx, nthreads=16)
    cols = np.linspace(0, mod_nx - 1, P["nshot"] + 2)[1:-1].round().astype(int)
    return dict(P=P, cfg=cfg, d=d, c=c, index=index, nvel=nvel, vmin=vmin, vmax=vmax,
                wav=host.wavelet_table(d.nt, tao, P["f0"]), w=host.blend_weights(n2),
                r_u=host.source_row(P["src_depth"], h, n2),
                r_x=[int(d.s_l + cx) for cx in cols])


def cfg5_context(S, panels):
    """Context of `panels` groups of the survey's shots; group p images
    columns [p*mod_nx, (p+1)*mod_nx) of a wide reflectivity and runs on its
    own background window (cfg5_backgrounds)."""
    P, d, cfg = S["P"], S["d"], S["cfg"]
    ns = P["nshot"]
    ctx = hip.Context(nz=d.nz, nx=d.nx, n2=cfg.n2, nfdmax=cfg.nfdmax, nt=d.nt, nt2=d.nt2,
                      ilste=0, icompen=1, s_l=d.s_l, s_r=d.s_r, s_z=d.s_z, ds=1, n=cfg.n,
                      tao2=d.tao2, h2=d.h2, taoh=d.taoh, taoh2=d.taoh2, hzx2_1=d.hzx2_1,
                      max_shots=ns * panels, record=1, stack_nx=P["mod_nx"] * panels)
    flat = np.full((d.nz, d.nx), S["vmin"], np.float32)
    ctx.set_model(flat, flat * 0, S["w"], S["c"], S["index"], S["nvel"], S["vmin"], 1.0)
    ctx.set_wavelet(S["wav"])
    ctx.set_shot_geometry([dict(nx=d.nx, r_u=S["r_u"], r_x=S["r_x"][s], s_l=d.s_l, s_r=d.s_r,
                                n=cfg.n, x_img=p * P["mod_nx"])
                           for p in range(panels) for s in range(ns)])
    ctx.panels = panels
    return ctx


def cfg5_backgrounds(ctx, S, raws):
    """Background window of each group (raws: column-major [x][z] models)."""
    P = S["P"]
    for p, raw in enumerate(raws):
        v, r1 = host.pad_model(raw, P["mod_nz"], P["mod_nx"], P["n2"], 0, P["tao"], P["h"])
        for s in range(P["nshot"]):
            ctx.set_shot_model(p * P["nshot"] + s, v, r1)


def cfg5_models(seeds, panels):
    """Background and reflectivity of each seed, padded to `panels` with
    copies of the last background and zero reflectivity."""
    bg, ms = [], []
    for s in seeds:
        v, m = synth.reflectors(int(s), CFG5["mod_nz"], CFG5["mod_nx"], CFG5["h"], CFG5["vlo"],
                                CFG5["vhi"])
        bg.append(v)
        ms.append(m)
    while len(bg) < panels:
        bg.append(bg[-1])
        ms.append(np.zeros_like(ms[-1]))
    return bg, np.stack(ms)


def cfg5_features(img, c, bg, norm):
    """Network input [n][3][z][x]: c * image / norm (norm: the RMS of the
    training reflectivities, so that inputs and targets are of order one and
    Adam's steps are small against them), background velocity scaled to
    [0, 1] over [vlo, vhi], depth in [0, 1]."""
    img = np.asarray(img, np.float32)
    n, mnz, mnx = img.shape
    x = np.empty((n, 3, mnz, mnx), np.float32)
    x[:, 0] = np.float32(c / norm) * img
    for i in range(n):
        x[i, 1] = (bg[i].T - CFG5["vlo"]) / (CFG5["vhi"] - CFG5["vlo"])
    x[:, 2] = np.linspace(0.0, 1.0, mnz, dtype=np.float32)[:, None]
    return x


def cfg5_pairs(S, seeds, panels=8):
    """Migrated images H m_i, reflectivities m_i and backgrounds of the seeds,
    `panels` models per launch."""
    ctx = cfg5_context(S, panels)
    X, M, B = [], [], []
    t0 = time.time()
    for b in range(0, len(seeds), panels):
        chunk = list(seeds[b:b + panels])
        bg, m = cfg5_models(chunk, panels)
        cfg5_backgrounds(ctx, S, bg)
        x = normal_images(ctx, m)
        k = len(chunk)
        X.append(x[:k])
        M.append(m[:k])
        B.extend(bg[:k])
        log(f"[cfg5] pairs {b + k} of {len(seeds)} ({time.time() - t0:.0f}s)")
    ctx.close()
    return np.concatenate(X), np.concatenate(M), B


def cfg5_train(X, M, B, c, norm, net=None, epochs=800, patches_per_epoch=2048, patch=64,
               batch=256, lr=1e-3, seed=0):
    """Train (or continue training) the network on random patches of the
    pairs (targets m / norm); returns the network and its per-epoch losses."""
    XF = cfg5_features(X, c, B, norm)
    M = np.asarray(M, np.float32) / np.float32(norm)
    if net is None:
        net = cnn.Deblur(list(CFG5_CHANNELS), batch * patch * patch, batch * patch)
        net.init_params(seed=seed)
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):  # fresh random crops each epoch
        xp, mp = patches(XF, M, patch, patches_per_epoch, rng)
        hist.extend(net.fit(xp, mp, epochs=1, batch=batch, lr=lr, seed=seed + ep))
        if (ep + 1) % 100 == 0:
            log(f"[cfg5] epoch {ep + 1} of {epochs}: loss {np.mean(hist[-100:]):.4e}")
    return net, hist


def cfg5_load(path):
    """A network saved by the training stages, its image scale c and the
    reflectivity normalisation norm."""
    net = cnn.Deblur.load(path, 256 * 64 * 64, 256 * 64)
    with np.load(path, allow_pickle=False) as f:
        c, norm = float(f["scale"]), float(f["norm"])
    return net, c, norm


def cfg5_estimate(net, img, c, norm, bg):
    """The network's reflectivity estimates of the images img [n][z][x]."""
    xe = cfg5_features(img, c, bg, norm)
    return np.stack([net.forward(xe[i:i + 1])[0] for i in range(len(xe))]) * np.float32(norm)


def cfg5_evaluate(S, net, c, norm, neval=32, panels=8, cg_iters=10, seed0=CFG5_EVAL_SEED):
    """Held-out models: Born data of each, then (timed per launch of `panels`
    models) the migrated image, the network, the fallback check and
    cg_iters CGLS iterations; per-model errors and misfits."""
    ctx = cfg5_context(S, panels)
    P = S["P"]
    mnz, mnx = P["mod_nz"], P["mod_nx"]
    per, times = [], dict(image=0.0, net=0.0, guard=0.0, cg=0.0)
    for b in range(0, neval, panels):
        seeds = list(range(seed0 + b, seed0 + min(neval, b + panels)))
        k = len(seeds)
        bg, mt = cfg5_models(seeds, panels)
        cfg5_backgrounds(ctx, S, bg)
        op = lsrtm.BornOperator(ctx, groups=panels)
        op.forward(mt)  # observed data: Born data of the true reflectivities
        ctx.gathers_from_records()
        ctx.sync()
        t = time.time()  # the migrated image: forward pass with strips + adjoint pass
        ctx.forward_begin(True)
        ctx.forward_advance(ctx.p.nt)
        ctx.adjoint_begin()
        ctx.adjoint_advance(ctx.p.nt)
        g = ctx.adjoint_gradient()
        img = np.ascontiguousarray(g.reshape(mnz, panels, mnx).transpose(1, 0, 2))
        times["image"] += time.time() - t
        t = time.time()
        m_net = cfg5_estimate(net, img, c, norm, bg)
        times["net"] += time.time() - t
        t = time.time()
        m_out, choice, r_net, r_rtm, alpha = guard(op, img, m_net)
        ctx.sync()
        times["guard"] += time.time() - t
        t = time.time()
        iters = _Iterates(f"[cfg5] held-out {b}..{b + k - 1}: CGLS iteration")
        m_cg, hist = lsrtm.cgls(op, None, cg_iters, callback=iters)
        times["cg"] += time.time() - t
        reach = iters.first_within(m_cg, axis_models=True)
        resid_cg = np.sqrt(np.maximum(hist[-1]["rr"], 0.0) / hist[0]["rr"])
        m_rtm = np.float32(c) * img
        for i in range(k):
            per.append(dict(seed=seeds[i], rel_net_vs_cg=rel(m_net[i], m_cg[i]),
                            rel_rtm_vs_cg=rel(m_rtm[i], m_cg[i]),
                            rel_out_vs_cg=rel(m_out[i], m_cg[i]),
                            rel_net_vs_true=rel(m_net[i], mt[i]),
                            rel_rtm_vs_true=rel(m_rtm[i], mt[i]),
                            rel_cg_vs_true=rel(m_cg[i], mt[i]),
                            resid_net=float(r_net[i]), resid_rtm=float(r_rtm[i]),
                            resid_cg=float(resid_cg[i]), choice=str(choice[i]),
                            cg_iters_to_5pct=reach[i][0], cg_s_to_5pct=reach[i][1]))
        log(f"[cfg5] held-out {b + k} of {neval}: rel to CGLS net "
            f"{np.mean([q['rel_net_vs_cg'] for q in per[-k:]]):.3f}, scaled image "
            f"{np.mean([q['rel_rtm_vs_cg'] for q in per[-k:]]):.3f}")
    ctx.close()
    nb = (neval + panels - 1) // panels
    res = {key: float(np.mean([q[key] for q in per])) for key in per[0]
           if key not in ("seed", "choice", "cg_iters_to_5pct", "cg_s_to_5pct")}
    res.update(neval=neval, panels=panels, cg_iters=cg_iters, launches=nb,
               chose_net=sum(q["choice"] == "net" for q in per),
               cg_iters_to_5pct=[q["cg_iters_to_5pct"] for q in per],
               s_per_launch={k_: v / nb for k_, v in times.items()}, per_model=per)
    return res


def cfg5_stage(stage, out, load="", nstage=2, ntrain=512, epochs=800, panels=8, seed=0):
    """Training stage `stage` of `nstage`: the pairs of training models
    [stage*ntrain/nstage, (stage+1)*ntrain/nstage), then training from scratch
    (stage 0) or from the network saved at `load`; writes the network and its
    input scale to `out`.  Stages split the 512 pair passes over separate runs
    of bounded length; only the network travels between them."""
    S = cfg5_survey()
    per = ntrain // nstage
    seeds = range(CFG5_TRAIN_SEED + stage * per, CFG5_TRAIN_SEED + (stage + 1) * per)
    t = time.time()
    X, M, B = cfg5_pairs(S, seeds, panels)
    t_pairs = time.time() - t
    if load:
        net, c, norm = cfg5_load(load)
    else:
        c = float(np.sum(X.astype(np.float64) * M) / np.sum(X.astype(np.float64) ** 2))
        norm = float(np.sqrt(np.mean(M.astype(np.float64) ** 2)))
        net = None
    t = time.time()
    net, hist = cfg5_train(X, M, B, c, norm, net=net, epochs=epochs, seed=seed + 7919 * stage)
    t_train = time.time() - t
    loss_identity = float(np.mean(((np.float32(c) * X - M) / np.float32(norm)) ** 2))
    net.save(out, scale=np.float64(c), norm=np.float64(norm), stage=np.int32(stage),
             loss=np.asarray(hist, np.float64))
    res = dict(stage=stage, models=[seeds[0], seeds[-1]], pairs_s=t_pairs, train_s=t_train,
               epochs=epochs, scale=c, norm=norm, loss_identity=loss_identity,
               loss_first=float(np.mean(hist[:20])), loss_last=float(np.mean(hist[-20:])))
    log(f"[cfg5] stage {stage}: pairs {t_pairs:.0f}s, training {t_train:.0f}s, loss "
        f"{res['loss_first']:.4e} -> {res['loss_last']:.4e} (scaled image {loss_identity:.4e})")
    return res


def main():
    import argparse
    import json
    ap = argparse.ArgumentParser(description="deblurring network vs CGLS")
    ap.add_argument("--npairs", type=int, default=128)
    ap.add_argument("--panels", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--cg-iters", type=int, default=10)
    ap.add_argument("--shots", type=int, default=12)
    ap.add_argument("--nt", type=int, default=2000)
    ap.add_argument("--target", default="cg", choices=["cg", "true"],
                    help="train toward the CGLS image of each pair or its reflectivity")
    ap.add_argument("--save", default="", help="write the four images to this .npz")
    ap.add_argument("--save-net", default="", help="write the trained network to this .npz")
    ap.add_argument("--cfg5-stage", type=int, default=-1,
                    help="random-reflectivity configuration: run training stage 0 or 1 "
                         "(256 training models each) and write the network to --save-net")
    ap.add_argument("--load-net", default="", help="stage > 0: continue from this network")
    a = ap.parse_args()
    if a.cfg5_stage >= 0:
        res = cfg5_stage(a.cfg5_stage, a.save_net, load=a.load_net,
                         epochs=a.epochs if a.epochs != 600 else 800)
        print(json.dumps(res))
        return
    S = build_survey(nshot=a.shots, nt=a.nt)
    res, im = run(S, npairs=a.npairs, panels=a.panels, epochs=a.epochs, batch=a.batch, lr=a.lr,
                  cg_iters=a.cg_iters, target=a.target, save_net=a.save_net)
    if a.save:
        np.savez_compressed(a.save, **im)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
