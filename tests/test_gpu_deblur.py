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
    """deblur() returns the network's estimate only when its Born re-modelled
    data fit the observations better than those of the scaled migrated image."""
    from rtm_gpu_amd import cnn, deblur
    S = _survey()
    ctx = deblur.context(S)
    m_true = S["m_true"]
    ctx.set_reflectivity(m_true)
    dobs = ctx.born()
    ctx.forward_begin(True)
    ctx.forward_advance(ctx.p.nt)
    img = ctx.adjoint(dobs)
    c = float(np.sum(img.astype(np.float64) * m_true) / np.sum(img.astype(np.float64) ** 2))
    illum = deblur.illumination(S)
    mnz, mnx = img.shape

    class Exact:  # stands in for a network that recovers the reflectivity
        def forward(self, x):
            return np.asarray(m_true, np.float32)[None]

    m, which, r_net, r_rtm = deblur.deblur(ctx, Exact(), S, img, c, illum, dobs)
    assert which == "net" and r_net == 0.0 < r_rtm and np.array_equal(m, m_true)

    zero = cnn.Deblur([3, 8, 1], mnz * mnx, mnz)
    zero.init_params(seed=1)  # last layer zero: the output is the scaled image
    m, which, r_net, r_rtm = deblur.deblur(ctx, zero, S, img, c, illum, dobs)
    assert which == "rtm" and r_net == r_rtm
    assert np.array_equal(m, np.float32(c) * img)

    bad = cnn.Deblur([3, 8, 1], mnz * mnx, mnz)
    p = bad.init_params(seed=2, last_zero=False)
    bad.set_params(p * 50)
    m, which, r_net, r_rtm = deblur.deblur(ctx, bad, S, img, c, illum, dobs)
    print(f"misfit: scaled image {r_rtm:.3f}, inflated network {r_net:.3f}")
    assert which == "rtm" and not r_net < r_rtm
    assert np.array_equal(m, np.float32(c) * img)
