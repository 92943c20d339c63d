"""Stability of the time stepping (CPU): rtmh_stability predicts when the FP32
replica of the reference's update blows up, and the driver refuses a time
step beyond the limit before any device work (the reference runs any step
and only produces NaN images)."""
import numpy as np
import pytest

import cases
from rtm_gpu_amd import host
from test_host import write_param_files


def _number(case):
    cfg, m = case["cfg"], case["model"]
    return host.stability(m["v"], m["c"], m["index"], m["nvel"], m["vmin"], m["dv"],
                          cfg["nfdmax"], cfg["tao2"], cfg["h2"], cfg["hzx2_1"])


def _peak(case, nsteps):
    e = cases.oracle_engine(case)
    e.forward_begin()
    peak = 0.0
    for _ in range(nsteps // 50):
        e.forward_advance(50)
        u = e.field(0)
        if not np.isfinite(u).all():
            return np.inf
        peak = max(peak, float(np.abs(u).max()))
    return peak


def test_taylor_number_closed_form():
    # Taylor rows alternate in sign, so max S = S(pi) = 4 (c_1 + c_3 + ...);
    # on a homogeneous model the number is r^2 (1 + rho) S(pi) / 4, r = v tao / h.
    case = cases.build_case(name="stabT", ilste=1, nfdmax=4, nlayer=1, vlo=3000.0, vhi=3000.0,
                            tao=0.001, nt=60)
    c = case["model"]["c"]
    r = 3000.0 * 0.001 / 10.0
    want = r * r * 2.0 * 4.0 * float(c[1] + c[3]) / 4.0
    assert _number(case) == pytest.approx(want, rel=1e-6)


@pytest.mark.parametrize("ilste", [1, 0])
def test_stability_number_predicts_blowup(ilste):
    # The limit is found by bisection on tao (LS rows are designed for the
    # time step, so the number is not exactly proportional to tao^2): 5% above
    # it the replica overflows within 600 steps, 5% below it stays bounded.
    kw = dict(ilste=ilste, nfdmax=4, nlayer=3, vlo=2000.0, vhi=3000.0, nt=600, mod_nz=40,
              mod_nx=50, n2=10, f0=15.0)
    lo, hi = 0.0005, 0.004
    for _ in range(24):
        mid = 0.5 * (lo + hi)
        if _number(cases.build_case(name="stab_probe", tao=float(np.float32(mid)), **kw)) <= 1.0:
            lo = mid
        else:
            hi = mid
    tcrit = 0.5 * (lo + hi)
    below = cases.build_case(name="stab_lo", tao=float(0.95 * tcrit), **kw)
    above = cases.build_case(name="stab_hi", tao=float(1.05 * tcrit), **kw)
    assert _number(below) < 1.0 < _number(above)
    assert _peak(below, 600) < 10.0
    assert not np.isfinite(_peak(above, 600))


def test_driver_refuses_unstable_time_step(tmp_path):
    # the model and depths exist, so the driver gets as far as the coefficient
    # design; tao = 3 ms at 3000 m/s and h = 10 m is far beyond the limit
    vals = write_param_files(tmp_path, tao=0.003, tao1=0.003, ilste=1)
    raw = np.full((vals["mod_nx"], vals["mod_nz"]), 3000.0, np.float32)
    raw.tofile(tmp_path / "model.dat")
    (tmp_path / "Depth_Of_Receiver.txt").write_text("40.0 70.0")
    with pytest.raises(host.HostError, match="stability limit"):
        host.run(tmp_path / "2D_Real_RVSP_RTM.txt")
