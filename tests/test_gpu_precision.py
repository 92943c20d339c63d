"""Reduced-precision storage of the forward/backward interface: boundary
strips (nfdmax = 4 rings per edge per step) and checkpoint snapshots every
200 steps, stored as FP16 or BF16 with a per-slot power-of-two max-abs scale,
against FP32 storage.  FD arithmetic is FP32 throughout.  Errors are relative
L2 over the interior against an FP64 oracle run of the same problem with the
same checkpoints (kept exactly):

  forward u(s) after s = 100, 1000, nt-1 steps; the reconstructed source
  field u(k) of the backward pass after 100 and 1000 backward steps and at
  k = 150 (50 steps below the first checkpoint pair, not itself restored);
  and the final image accumulator I1.

Default size: the reverse-VSP geometry, 351 x 1701 interior, 4000 steps,
nfdmax 4 (about a minute of FP64 on 16 host threads).  RTM_PRECISION_SMALL=1
runs the faulted model at 120 x 400, 1000 steps instead;
RTM_PRECISION_OUT=<path> writes the table as markdown."""
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

FULL = os.environ.get("RTM_PRECISION_SMALL") != "1"
SNAP = 200
FORMATS = {0: "FP32", 1: "FP16", 2: "BF16"}


def _case():
    if FULL:
        return cases.build_case(name="prec_full", kind="faulted", mod_nz=351, mod_nx=1701, n2=30,
                                nfdmax=4, nt=4000, h=12.5, hz=12.5, tao=0.001, f0=20.0, seed=2,
                                src_x1=851, src_depth=400.0)
    return cases.build_case(name="prec_small", kind="faulted", mod_nz=120, mod_nx=400, n2=20,
                            nfdmax=4, nt=1000, h=12.5, hz=12.5, tao=0.001, f0=20.0, seed=2,
                            src_x1=200, src_depth=300.0)


def _fwd_marks(nt):
    return [s for s in (100, 1000) if s < nt - 1] + [nt - 1]


def _bwd_marks(nt):
    """Fields u(k) of the backward pass to compare (descending k)."""
    ks = [nt - 2 - s for s in (100, 1000) if nt - 2 - s > 150]
    return ks + [150]


def test_reduced_precision_storage_errors():
    from rtm_gpu_amd import hip
    case = _case()
    cfg = case["cfg"]
    nt, n2 = cfg["nt"], cfg["n2"]
    print(f"grid {cfg['nz']}x{cfg['nx']} nt={nt} nfdmax={cfg['nfdmax']} snapshots every {SNAP}",
          flush=True)
    # observed data: the engine's own forward modelling (bit-exact to the oracle)
    rec = cases.hip_context(case, nshot=1, record=1)
    rec.forward_begin(save_strips=False)
    rec.forward_advance(nt)
    seis = rec.records()[0]
    del rec

    inner = (slice(n2, -n2), slice(n2, -n2))

    def march(fwd_begin, fwd_adv, bwd_begin, bwd_adv, field_u, field_i1):
        fwd, bwd = {}, {}
        fwd_begin()
        k = 1
        for s in _fwd_marks(nt):
            fwd_adv(s - k)
            k = s
            fwd[s] = field_u()[inner].astype(np.float64)
        bwd_begin()
        k = nt - 2  # the backward state starts at u(nt-2); each step lowers k by one
        for kk in _bwd_marks(nt):
            bwd_adv(k - kk)
            k = kk
            bwd[kk] = field_u()[inner].astype(np.float64)
        bwd_adv(nt)
        return fwd, bwd, field_i1()[inner].astype(np.float64)

    e64 = cases.oracle_engine(case, prec=64, snap_every=SNAP, nthreads=min(16, os.cpu_count() or 1))
    ref = march(e64.forward_begin, e64.forward_advance, lambda: e64.backward_begin(seis),
                e64.backward_advance, lambda: e64.field(0, f64=True),
                lambda: e64.field(3, f64=True))
    del e64

    rows = []
    res = {}
    for fmt, name in FORMATS.items():
        ctx = cases.hip_context(case, nshot=1, strip_format=fmt, snap_every=SNAP)
        ctx.set_gathers(seis[None])
        got = march(ctx.forward_begin, ctx.forward_advance, ctx.backward_begin,
                    ctx.backward_advance, lambda: ctx.field(hip.FIELD_U),
                    lambda: ctx.field(hip.FIELD_I1))
        fe = {s: cases.rel_l2(got[0][s], ref[0][s]) for s in got[0]}
        be = {s: cases.rel_l2(got[1][s], ref[1][s]) for s in got[1]}
        ie = cases.rel_l2(got[2], ref[2])
        res[fmt] = (fe, be, ie)
        rows.append(f"| {name} | " + " | ".join(f"{fe[s]:.2e}" for s in sorted(fe)) + " | " +
                    " | ".join(f"{be[k]:.2e}" for k in sorted(be, reverse=True)) +
                    f" | {ie:.2e} |")
        del ctx
    head = ("| storage | " + " | ".join(f"fwd u({s})" for s in sorted(res[0][0])) + " | " +
            " | ".join(f"bwd u({k})" for k in sorted(res[0][1], reverse=True)) + " | image I1 |")
    sep = "|" + "---|" * (2 + len(res[0][0]) + len(res[0][1]))
    table = "\n".join([head, sep] + rows)
    print(table)
    out = os.environ.get("RTM_PRECISION_OUT")
    if out:
        with open(out, "w") as f:
            f.write(f"grid {cfg['nz']}x{cfg['nx']}, nt={nt}, nfdmax={cfg['nfdmax']}, strips and "
                    f"snapshots (every {SNAP} steps) in the storage format, FD arithmetic FP32; "
                    f"relative L2 over the interior against the FP64 oracle\n\n{table}\n")
    # FP32 storage: only the FP32 arithmetic differs from the FP64 run
    assert res[0][2] < 1e-3
    assert max(res[0][1].values()) < 1e-3
    # the forward pass does not read the stored data
    for fmt in (1, 2):
        assert res[fmt][0] == res[0][0]
    assert res[1][2] < 2e-2 and res[2][2] < 1e-1
    assert res[2][2] >= res[1][2] >= res[0][2] * 0.5
