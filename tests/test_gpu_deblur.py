"""Training pairs for the deblurring network, made by the engine: H m = L^T L m
of several reflectivity panels in one batch must equal, bit for bit, the
same computation on the survey alone (the panel's shots form the same sums in
the same order); and the whole pipeline (pairs, training, evaluation against
CGLS) runs at a small size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _survey():
    from rtm_gpu_amd import deblur
    return deblur.build_survey(mod_nz=48, mod_nx=90, nt=500, nshot=3, nfdmax=4, n2=12, h=10.0)


def test_panels_equal_single_survey_runs():
    from rtm_gpu_amd import deblur
    S = _survey()
    m = np.stack([deblur.random_reflectivity(S, 100 + i) for i in range(3)])
    batch = deblur.normal_images(deblur.context(S, panels=3), m)
    one = deblur.context(S)
    for i in range(3):
        single = deblur.normal_images(one, m[i:i + 1])[0]
        assert np.array_equal(batch[i], single), f"panel {i}"
        assert np.abs(single).max() > 0


def test_pipeline_small():
    from rtm_gpu_amd import deblur
    S = _survey()
    res, im = deblur.run(S, npairs=8, panels=4, epochs=20, batch=4, lr=3e-3,
                         channels=(3, 8, 8, 1), dilation=(1, 2, 1), cg_iters=3)
    print({k: v for k, v in res.items() if not isinstance(v, list)})
    for k in ("rel_net_vs_cg", "rel_rtm_vs_cg", "resid_rtm", "resid_net", "resid_cg"):
        assert np.isfinite(res[k]), k
    assert res["loss_last"] < res["loss_identity"]
    assert res["resid_cg"] < res["resid_rtm"]
    assert all(b < a for a, b in zip(res["cg_rr"], res["cg_rr"][1:]))
    assert all(np.isfinite(v).all() for v in im.values())


def test_grouped_cgls_on_panels_equals_single_runs():
    """CGLS of several reflectivity panels in one batch, with per-panel step
    lengths, follows the CGLS of each panel's survey alone."""
    from rtm_gpu_amd import deblur, lsrtm
    S = _survey()
    m = np.stack([deblur.random_reflectivity(S, 200 + i) for i in range(3)])
    ctx = deblur.context(S, panels=3)
    deblur.normal_images(ctx, m)
    mb, hb = lsrtm.cgls(lsrtm.BornOperator(ctx, groups=3), None, 3)
    one = deblur.context(S)
    for i in range(3):
        deblur.normal_images(one, m[i:i + 1])
        mi, hi = lsrtm.cgls(lsrtm.BornOperator(one), None, 3)
        # the same passes bit for bit; FP64 products over the same values
        for a, b in zip(hb, hi):
            assert abs(a["rr"][i] - b["rr"]) <= 1e-12 * hi[0]["rr"], i
        assert float(np.abs(mb[i] - mi).max()) <= 1e-5 * float(np.abs(mi).max()), i


def test_plain_migration_is_the_fallback():
    """guard() returns the network's estimate only when its Born re-modelled
    data fit the observations better than both zero and the best multiple
    alpha = <L img, d> / ||L img||^2 of the migrated image; otherwise that
    multiple, whose misfit is at most 1."""
    from rtm_gpu_amd import cnn, deblur, lsrtm
    S = _survey()
    ctx = deblur.context(S)
    op = lsrtm.BornOperator(ctx)
    m_true = S["m_true"]
    ctx.set_reflectivity(m_true)
    dobs = ctx.born()
    ctx.forward_begin(True)
    ctx.forward_advance(ctx.p.nt)
    img = ctx.adjoint(dobs)  # leaves the gathers holding dobs
    ctx.set_reflectivity(img)
    limg = ctx.born().astype(np.float64)
    d64 = dobs.astype(np.float64)
    alpha = float(np.sum(limg * d64) / np.sum(limg * limg))
    want_rtm = float(np.linalg.norm(alpha * limg - d64) / np.linalg.norm(d64))
    ctx.set_gathers(dobs)

    m, which, r_net, r_rtm, a = deblur.guard(op, img, m_true)  # the exact reflectivity
    assert which == "net" and r_net <= 1e-6 and np.array_equal(m, m_true)
    assert a == pytest.approx(alpha, rel=1e-6) and r_rtm == pytest.approx(want_rtm, rel=1e-5)
    assert 0.0 < r_rtm < 1.0

    # the negated reflectivity fits worse than zero (misfit 2): rejected
    m, which, r_net, _, _ = deblur.guard(op, img, -m_true)
    assert which == "rtm" and r_net == pytest.approx(2.0, rel=1e-5)
    assert np.array_equal(m, np.float32(alpha) * img)

    # an image-domain multiple no better than the best one: rejected
    c = float(np.sum(img.astype(np.float64) * m_true) / np.sum(img.astype(np.float64) ** 2))
    zero = cnn.Deblur([3, 8, 1], img.size, img.shape[0])
    zero.init_params(seed=1)  # last layer zero: the output is the scaled image c * img
    illum = deblur.illumination(S)
    ctx.set_gathers(dobs)
    m_zero = zero.forward(deblur.features(S, img, c, illum))[0]
    assert np.array_equal(m_zero, np.float32(c) * img)
    m, which, r_net, r_rtm, _ = deblur.guard(op, img, m_zero)
    assert which == "rtm" and r_rtm <= r_net + 1e-9

    # a network whose output is not finite is rejected without a Born pass of it
    bad = np.full_like(img, np.nan)
    m, which, r_net, r_rtm, _ = deblur.guard(op, img, bad)
    assert which == "rtm" and r_net == np.inf and np.isfinite(m).all()


def test_cfg5_pipeline_full_size():
    """The random-reflectivity configuration at its full size (1024 x 256,
    64 shots, 3000 steps) for two models in one launch: pairs are finite and
    not zero, and the held-out evaluation orders the data misfits as the
    method fixes them (CGLS(2) below the best-scaled image below zero), with
    the fallback taken for a network that only scales the image."""
    from rtm_gpu_amd import cnn, deblur
    S = deblur.cfg5_survey()
    print("cfg5 full size: two models, 128 shots, 3000 steps", flush=True)
    X, M, B = deblur.cfg5_pairs(S, [deblur.CFG5_TRAIN_SEED, deblur.CFG5_TRAIN_SEED + 1], panels=2)
    assert X.shape == (2, 256, 1024) and np.isfinite(X).all()
    assert all(np.abs(x).max() > 0 for x in X)
    c = float(np.sum(X.astype(np.float64) * M) / np.sum(X.astype(np.float64) ** 2))
    assert c > 0
    net = cnn.Deblur(list(deblur.CFG5_CHANNELS), 256 * 1024, 256)
    net.init_params(seed=0)  # zero last layer: output c * image
    norm = float(np.sqrt(np.mean(M.astype(np.float64) ** 2)))
    res = deblur.cfg5_evaluate(S, net, c, norm, neval=2, panels=2, cg_iters=2)
    print({k: v for k, v in res.items() if k != "per_model"})
    for q in res["per_model"]:
        assert q["choice"] == "rtm"
        assert q["resid_cg"] < q["resid_rtm"] < 1.0
        assert q["rel_cg_vs_true"] < 1.0
        assert np.isfinite(q["rel_net_vs_cg"])
