"""Learned image-domain deblurring of migrated images, against least-squares
migration.  Not in the reference, which stops at the migrated image
(kernel.cu:935-1059).

The migrated image of Born data is H m = L^T L m: the reflectivity m blurred
by the Hessian of Born modelling.  Least-squares migration (lsrtm.cgls)
undoes the blur iteratively, at two wave-equation passes per iteration.  A
small convolutional network (cnn.Deblur) learns an approximate inverse of H
in the image domain instead:

  training pairs   random reflectivities m_i (the high-pass velocity
                   contrast (v^2 - v0^2)/v0^2 of faulted layered models drawn
                   with other seeds than the evaluated one) -> Born data
                   d_i = L m_i -> input x_i = L^T d_i = H m_i, target the
                   image of cg_iters CGLS iterations on d_i (or m_i itself
                   with target="true"), all by this engine, many pairs per
                   launch (each panel of a wide reflectivity is one group of
                   shots at its own image columns, with its own CG step
                   lengths);
  input channels   c * x (c: the least-squares scalar between images and
                   reflectivities of the training set), the normalised log
                   source illumination, and depth;
  output           m ~ c x + net(x) (zero net: the scaled migrated image,
                   the plain-migration fallback).

evaluate() compares the network's estimate with 10 CGLS iterations of the
same engine on the evaluated model: relative L2 between the two images, the
Born re-modelled data residual ||L m - d|| / ||d|| of each, and wall times."""
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


def panel_residuals(ctx, panels, dm):
    """||L dm_p|| per panel (dm [panels][mnz][mnx]) from one batched Born pass."""
    op = lsrtm.BornOperator(ctx, groups=panels)
    return np.sqrt(np.maximum(op.forward(dm), 0.0))


def run_models(S, ntrain=96, neval=8, panels=8, epochs=100, patch=64, batch=256,
               patches_per_epoch=2048, lr=1e-3, channels=(3,) + (32,) * 5 + (1,),
               dilation=None, cg_iters=10, target="true", seed=0, save_net=""):
    """Several random models: train on engine-made pairs (patches), evaluate on
    held-out models against CGLS with per-model statistics."""
    res = dict(grid=f"{S['mod_nz']}x{S['mod_nx']}", nt=int(S["d"].nt), shots=S["nshot"],
               ntrain=ntrain, neval=neval, epochs=epochs, patch=patch, batch=batch,
               channels=list(channels), dilation=list(dilation) if dilation else None,
               target=target)
    t = time.time()
    X, M = make_pairs(S, ntrain, panels, seed0=1000, target=target, cg_iters=cg_iters)
    res["pairs_s"] = time.time() - t
    log(f"[deblur] {ntrain} training pairs in {res['pairs_s']:.1f}s")
    c = float(np.sum(X.astype(np.float64) * M) / np.sum(X.astype(np.float64) ** 2))
    res["scale"] = c
    illum = illumination(S)
    XF = features(S, X, c, illum)
    del X
    rng = np.random.default_rng(seed)
    net = cnn.Deblur(list(channels), batch * patch * patch, batch * patch, dilation=dilation)
    net.init_params(seed=seed)
    t = time.time()
    hist = []
    for ep in range(epochs):  # fresh random crops each epoch
        xp, mp = patches(XF, M, patch, patches_per_epoch, rng)
        hist.extend(net.fit(xp, mp, epochs=1, batch=batch, lr=lr, seed=seed + ep))
        if (ep + 1) % 10 == 0:
            log(f"[deblur] epoch {ep + 1} of {epochs}: loss {hist[-1]:.3e}")
    res["train_s"] = time.time() - t
    res["loss_first"], res["loss_last"] = float(hist[0]), float(hist[-1])
    if save_net:
        net.save(save_net)
    del XF, M
    # held-out models: image, network, CGLS per panel
    ctx = context(S, panels)
    per = []
    for b in range(0, neval, panels):
        k = min(panels, neval - b)
        mt = np.stack([random_reflectivity(S, 5000 + b + i) for i in range(panels)])
        if k < panels:
            mt[k:] = 0.0
        normal_images(ctx, mt)  # Born data d of m into the gathers
        t = time.time()  # the migrated image of d: a forward pass and the adjoint pass
        ctx.forward_begin(True)
        ctx.forward_advance(ctx.p.nt)
        ctx.adjoint_begin()
        ctx.adjoint_advance(ctx.p.nt)
        g = ctx.adjoint_gradient()
        img = np.ascontiguousarray(g.reshape(g.shape[0], panels, -1).transpose(1, 0, 2))
        t_img = time.time() - t
        xe = features(S, img, c, illum)
        t = time.time()
        m_net = np.stack([net.forward(xe[i:i + 1])[0] for i in range(panels)])
        t_net = time.time() - t
        t = time.time()
        iters = _Iterates("[deblur] held-out CGLS iteration")
        m_cg, _ = lsrtm.cgls(lsrtm.BornOperator(ctx, groups=panels), None, cg_iters,
                             callback=iters)
        t_cg = time.time() - t
        reach = iters.first_within(m_cg, axis_models=True)
        m_rtm = np.float32(c) * img
        dn = panel_residuals(ctx, panels, mt)
        rn = {name: panel_residuals(ctx, panels, est - mt) / np.where(dn > 0, dn, 1.0)
              for name, est in (("rtm", m_rtm), ("net", m_net), ("cg", m_cg))}
        for i in range(k):
            per.append(dict(cg_iters_to_5pct=reach[i][0],
                            rel_net_vs_cg=rel(m_net[i], m_cg[i]), rel_rtm_vs_cg=rel(m_rtm[i], m_cg[i]),
                            rel_net_vs_true=rel(m_net[i], mt[i]), rel_cg_vs_true=rel(m_cg[i], mt[i]),
                            resid_rtm=float(rn["rtm"][i]), resid_net=float(rn["net"][i]),
                            resid_cg=float(rn["cg"][i])))
        res.setdefault("t_image_s", t_img)
        res.setdefault("t_net_s", t_net)
        res.setdefault("t_cg_s", t_cg)
        log(f"[deblur] held-out models {b}..{b + k - 1}: rel to CGLS net "
            f"{np.mean([q['rel_net_vs_cg'] for q in per[-k:]]):.3f}")
    for key in per[0]:
        if key != "cg_iters_to_5pct":
            res[key] = float(np.mean([q[key] for q in per]))
    res["cg_iters_to_5pct"] = [q["cg_iters_to_5pct"] for q in per]
    res["speedup_vs_cg"] = res["t_cg_s"] / (res["t_image_s"] + res["t_net_s"])
    res["per_model"] = per
    log(f"[deblur] {neval} held-out models: rel L2 to CGLS net {res['rel_net_vs_cg']:.3f}, "
        f"scaled image {res['rel_rtm_vs_cg']:.3f}; data residual RTM {res['resid_rtm']:.3f}, "
        f"net {res['resid_net']:.3f}, CGLS {res['resid_cg']:.3f}; time image+net "
        f"{res['t_image_s'] + res['t_net_s']:.2f}s vs CGLS {res['t_cg_s']:.2f}s")
    return res


def born_residual(ctx, m, dobs):
    """Relative misfit ||L m - d|| / ||d|| of the Born re-modelled data of the
    reflectivity m (one Born pass on ctx; its records are overwritten)."""
    ctx.set_reflectivity(m)
    r = ctx.born() - dobs
    return float(np.linalg.norm(r.astype(np.float64)) / np.linalg.norm(dobs.astype(np.float64)))


def deblur(ctx, net, S, img, c, illum, dobs):
    """The network's reflectivity estimate from the migrated image img, with
    plain migration as the fallback: the network's output is returned only
    when its Born re-modelled data fit the observed data dobs better than
    those of the scaled image c * img, otherwise the scaled image is.  Costs
    the network and two Born passes on ctx.  Returns (m, "net" or "rtm",
    misfit of the network's estimate, misfit of the scaled image)."""
    m_rtm = np.float32(c) * np.asarray(img, np.float32)
    m_net = net.forward(features(S, img, c, illum))[0]
    r_net = born_residual(ctx, m_net, dobs)
    r_rtm = born_residual(ctx, m_rtm, dobs)
    if np.isfinite(r_net) and r_net < r_rtm:
        return m_net, "net", r_net, r_rtm
    return m_rtm, "rtm", r_net, r_rtm


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
    t = time.time()
    iters = _Iterates()
    m_cg, ch = lsrtm.cgls(op, dobs, cg_iters, callback=iters)
    t_cg = time.time() - t
    k5, t5 = iters.first_within(m_cg)
    t = time.time()  # the estimate with its fallback check (network + two Born passes)
    _, chosen, r_net, r_rtm = deblur(ctx, net, S, img, c, illum, dobs)
    hip_sync()
    t_guard = time.time() - t

    res.update(
        t_image_s=t_img, t_net_s=t_net, t_cg_s=t_cg,
        speedup_vs_cg=t_cg / (t_img + t_net),
        rel_net_vs_cg=rel(m_net, m_cg), rel_rtm_vs_cg=rel(m_rtm, m_cg),
        rel_net_vs_true=rel(m_net, m_true), rel_rtm_vs_true=rel(m_rtm, m_true),
        rel_cg_vs_true=rel(m_cg, m_true),
        resid_rtm=r_rtm, resid_net=r_net, resid_cg=born_residual(ctx, m_cg, dobs),
        chosen=chosen, t_guarded_s=t_guard, speedup_guarded_vs_cg=t_cg / (t_img + t_guard),
        cg_rr=[h["rr"] / ch[0]["rr"] for h in ch],
        cg_iters_to_5pct=k5, cg_time_to_5pct_s=t5,
        net_within_5pct=bool(rel(m_net, m_cg) <= 0.05))
    log(f"[deblur] image {t_img:.2f}s + net {t_net:.3f}s vs CGLS({cg_iters}) {t_cg:.2f}s; "
        f"rel L2 to CGLS: net {res['rel_net_vs_cg']:.3f}, scaled RTM {res['rel_rtm_vs_cg']:.3f}; "
        f"data residual: RTM {res['resid_rtm']:.3f}, net {res['resid_net']:.3f}, "
        f"CGLS {res['resid_cg']:.3f}; returned: {chosen} ({t_img + t_guard:.2f}s with the check)")
    return res, dict(m_true=m_true, m_rtm=m_rtm, m_net=m_net, m_cg=m_cg)


def main():
    import argparse
    import json
    ap = argparse.ArgumentParser(description="deblurring network vs CGLS on the faulted model")
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
    ap.add_argument("--models", action="store_true",
                    help="many random models (1024 x 256, 64 shots, 25 Hz, 3000 steps; 6 x 32-"
                         "channel layers on 64 x 64 patches, batch 256), held-out evaluation")
    ap.add_argument("--ntrain", type=int, default=96)
    ap.add_argument("--save-net", default="", help="write the trained network to this .npz")
    ap.add_argument("--neval", type=int, default=8)
    a = ap.parse_args()
    if a.models:
        S = build_survey(mod_nz=256, mod_nx=1024, h=10.0, f0=25.0, nt=3000, nshot=64, nfdmax=4,
                         n2=30)
        res = run_models(S, ntrain=a.ntrain, neval=a.neval, epochs=a.epochs if a.epochs != 600
                         else 100, target=a.target if "--target" in __import__("sys").argv
                         else "true", save_net=a.save_net)
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
