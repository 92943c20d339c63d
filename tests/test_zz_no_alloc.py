"""rtm_data_dot_noalloc: the FP64 dot products of rtm_data_dot on caller-owned
device vectors, with the partial sums in the library's static device array
(no allocation, no host copy, no synchronisation).  Checked bit for bit
against rtm_data_dot on the same data, on two calls in a row, the second with
a smaller input."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _device(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_noalloc_dot_matches_rtm_data_dot():
    import torch
    from rtm_gpu_amd import hip
    case = cases.build_case(name="noalloc", mod_nz=30, mod_nx=50, n2=8, nfdmax=4, nt=160,
                            nshot=5, depth_step=10.0)
    cfg = case["cfg"]
    nshot, n, nt = 5, cfg["n"], cfg["nt"]
    print(f"noalloc dot: {nshot} shots, {n} receivers, {nt} steps", flush=True)
    ctx = cases.hip_context(case, record=1)
    ctx.forward_begin(save_strips=False)
    ctx.forward_advance(nt)
    rec = ctx.records()                                   # [shot][n][nt]
    rng = np.random.default_rng(7)
    seis = rng.standard_normal((nshot, n, nt)).astype(np.float32)
    ctx.set_gathers(seis)
    want_shot = [ctx.data_dot(hip.DATA_GATHERS, hip.DATA_RECORDS, s, 1) for s in range(nshot)]
    want_all = ctx.data_dot(hip.DATA_GATHERS, hip.DATA_RECORDS)
    want_gg = [ctx.data_dot(hip.DATA_GATHERS, hip.DATA_GATHERS, s, 1) for s in range(nshot)]

    # the context's data layout is time-major [shot][nt][n]
    x = _device(seis.transpose(0, 2, 1))
    y = _device(rec.transpose(0, 2, 1))
    m = nt * n
    out = torch.full((nshot,), float("nan"), dtype=torch.float64, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream

    # call 1: every shot as its own vector
    hip.data_dot_noalloc(x.data_ptr(), y.data_ptr(), m, nvec=nshot, stride=m,
                         out_ptr=out.data_ptr(), stream=stream)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.array_equal(got, np.array(want_shot)), (got, want_shot)

    # call 2, smaller: shots 1 and 2 of x with themselves, into a fresh output
    out2 = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda:0")
    hip.data_dot_noalloc(x.data_ptr() + 4 * m, x.data_ptr() + 4 * m, m, nvec=2, stride=m,
                         out_ptr=out2.data_ptr(), stream=stream)
    torch.cuda.synchronize()
    assert np.array_equal(out2.cpu().numpy(), np.array(want_gg[1:3]))

    # the whole batch as one vector equals rtm_data_dot over all shots
    out1 = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda:0")
    hip.data_dot_noalloc(x.data_ptr(), y.data_ptr(), nshot * m, out_ptr=out1.data_ptr(),
                         stream=stream)
    torch.cuda.synchronize()
    assert out1.item() == want_all

    # sizes beyond the scratch maximum and bad arguments are rejected
    with pytest.raises(hip.RtmError, match="scratch maximum"):
        hip.data_dot_noalloc(x.data_ptr(), y.data_ptr(), 1, nvec=hip.DOT_NOALLOC_MAX_VEC + 1,
                             stride=1, out_ptr=out.data_ptr(), stream=stream)
    with pytest.raises(hip.RtmError, match="n = -1"):
        hip.data_dot_noalloc(x.data_ptr(), y.data_ptr(), -1, out_ptr=out.data_ptr(),
                             stream=stream)
    ctx.close()
