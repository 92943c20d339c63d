"""Seeded synthetic RTM problems shared by the CPU and GPU tests.

A case is built with the oracle's restatement of the reference setup
(kernel.cu:602-753, GPU_velocity_real.cpp, LSMOrCon_rec_2D.cpp) so that the
GPU kernels and the oracle engine receive identical inputs.
"""
import functools

import numpy as np

import oracle_bind as ob


def layered_model(mod_nz, mod_nx, nlayer, vlo, vhi, seed, dip=0.0):
    """Column-major raw model (x outer, z inner) as the reference reads it
    (GPU_velocity_real.cpp:12-18): nlayer layers, seeded interface depths,
    velocities uniform in [vlo, vhi] sorted ascending with depth."""
    rng = np.random.default_rng(seed)
    depths = np.sort(rng.choice(np.arange(3, mod_nz - 3), nlayer - 1, replace=False))
    vels = np.sort(rng.uniform(vlo, vhi, nlayer)).astype(np.float32)
    vz = np.zeros((mod_nz, mod_nx), np.float32)
    xs = np.arange(mod_nx)
    for z in range(mod_nz):
        shift = dip * (xs - mod_nx / 2)
        lay = np.searchsorted(depths, z - shift, side="right")
        vz[z] = vels[lay]
    return np.ascontiguousarray(vz.T)  # [x][z]


@functools.lru_cache(maxsize=None)
def build_case(name="small", mod_nz=40, mod_nx=60, n2=10, nfdmax=4, nfdmin=2, nt=240,
               h=10.0, hz=10.0, tao=0.001, f0=25.0, ilste=0, icompen=1, seed=1, src_x1=30,
               src_depth=30.0, nlayer=4, vlo=1500.0, vhi=2800.0, ds=1, n=None, s_l1=1,
               s_z1=1, dip=0.15, nshot=1, depth_step=20.0, kind="layered", smooth=0.0):
    fmax = 2 * f0 + 1
    if n is None:
        n = (mod_nx - s_l1) // ds + 1
    d = ob.derive(mod_nz, mod_nx, n2, nt, tao, tao, h, hz, f0, s_l1, s_z1, n, ds, src_x1)
    if kind == "layered5":  # the seeded single-shot model: 5 flat layers, 1500-4500 m/s
        from rtm_gpu_amd import synth
        raw = synth.layered_5(seed=seed, mod_nz=mod_nz, mod_nx=mod_nx)
    elif kind == "faulted":  # the seeded reverse-VSP model (8 layers, 3 faults, a body)
        from rtm_gpu_amd import synth
        raw = synth.faulted_rvsp(seed=seed, mod_nz=mod_nz, mod_nx=mod_nx, h=h)
        if smooth:  # Gaussian-smoothed migration velocity: thousands of velocity bins
            raw = synth.smooth(raw, smooth)
    else:
        raw = layered_model(mod_nz, mod_nx, nlayer, vlo, vhi, seed, dip)
    v, r1 = ob.pad_model(raw, mod_nz, mod_nx, n2, 0, tao, h)
    dv = np.float32(1.0)
    vmin, vmax, nvel = ob.vrange(v, dv)
    w = ob.blend_weights(n2)
    if ilste == 0:
        need = ob.bins_needed(v, vmin, dv, nvel)
        # the reference holds tao, h, df, eps, fmax in float variables and
        # promotes them to double for funMandC (kernel.cu:536, 746)
        f32 = lambda x: float(np.float32(x))  # noqa: E731
        M, index, c = ob.design_ls(1000, nfdmax, nfdmin, nvel, f32(tao), f32(h), f32(1.0),
                                   f32(1e-5), f32(fmax), float(vmin), float(dv), need,
                                   float(d["hzx"]))
    else:
        index = None
        c = ob.taylor(nfdmax)
        M = None
    wav = ob.wavelet_table(d["NT"], tao, f0)
    r_u = [ob.lib().orc_source_row(np.float32(src_depth + i * depth_step), np.float32(hz), n2)
           for i in range(nshot)]
    r_x = [d["r_x"]] * nshot
    cfg = dict(nz=d["NZ"], nx=d["NX"], n2=n2, nfdmax=nfdmax, nt=d["NT"], nt2=d["NT2"],
               ilste=ilste, icompen=icompen, s_l=d["s_l"], s_r=d["s_r"], s_z=d["s_z"], ds=ds,
               n=n, tao2=d["tao2"], h2=d["h2"], taoh=d["taoh"], taoh2=d["taoh2"],
               hzx2_1=d["hzx2_1"], vmin=vmin, dv=dv)
    model = dict(v=v, r1=r1, w=w, c=c, index=index, nvel=nvel, vmin=vmin, vmax=vmax, dv=dv,
                 wav=wav, M=M)
    return dict(name=name, cfg=cfg, model=model, r_u=r_u, r_x=r_x, raw=raw, derived=d)


def oracle_engine(case, shot=0, prec=32, record=0, nthreads=1, snap_every=0):
    cfg = dict(case["cfg"])
    cfg.update(r_u=case["r_u"][shot], r_x=case["r_x"][shot], record=record, nthreads=nthreads,
               snap_every=snap_every)
    return ob.Engine(cfg, case["model"], prec=prec)


def hip_context(case, nshot=None, record=0, strip_format=0, zseg=0, snap_every=0, frame_split=0):
    from rtm_gpu_amd import hip
    cfg = case["cfg"]
    nshot = nshot or len(case["r_u"])
    keys = ("nz", "nx", "n2", "nfdmax", "nt", "nt2", "ilste", "icompen", "s_l", "s_r", "s_z",
            "ds", "n", "tao2", "h2", "taoh", "taoh2", "hzx2_1")
    ctx = hip.Context(**{k: cfg[k] for k in keys}, max_shots=nshot, strip_format=strip_format,
                      record=record, zseg=zseg, snap_every=snap_every, frame_split=frame_split)
    m = case["model"]
    ctx.set_model(m["v"], m["r1"], m["w"], m["c"], m["index"], m["nvel"], m["vmin"], m["dv"])
    ctx.set_wavelet(m["wav"])
    ctx.set_shots(case["r_u"][:nshot], case["r_x"][:nshot])
    return ctx


def synth_gather(case, shot=0):
    """Observed data for one shot, modelled by the oracle with receiver
    recording (the forward pass with record=1)."""
    e = oracle_engine(case, shot=shot, record=1)
    e.forward_begin()
    e.forward_advance(case["cfg"]["nt"])
    return e.record()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    if nb == 0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)


class OracleBornOp:
    """Born operator L and its adjoint on the oracle (one shot), with the
    interface of rtm_gpu_amd.lsrtm.BornOperator: float32 vectors, float64
    products.  Test infrastructure."""

    def __init__(self, case, shot=0, prec=32):
        self.case, self.shot, self.prec = case, shot, prec
        cfg = case["cfg"]
        self.n2 = cfg["n2"]
        self.shape = (cfg["nz"] - 2 * self.n2, cfg["nx"] - 2 * self.n2)
        self.r = self.q = None

    def _full(self, m):
        cfg = self.case["cfg"]
        f = np.zeros((cfg["nz"], cfg["nx"]), np.float32)
        f[self.n2:-self.n2, self.n2:-self.n2] = m
        return f

    def born(self, m):
        e = oracle_engine(self.case, shot=self.shot, prec=self.prec, record=1)
        e.set_reflectivity(self._full(m))
        e.born_begin()
        e.born_advance(self.case["cfg"]["nt"])
        return e.record()

    def _dot(self, a, b):
        return float(np.dot(a.astype(np.float64).ravel(), b.astype(np.float64).ravel()))

    def start(self, d):
        self.r = np.asarray(d, np.float32).reshape(self.case["cfg"]["n"], -1).copy()
        return self._dot(self.r, self.r)

    def forward(self, p):
        self.q = self.born(p)
        return self._dot(self.q, self.q)

    def update_residual(self, alpha):
        self.r = self.r + (-np.float32(alpha)) * self.q
        return self._dot(self.r, self.r)

    def adjoint(self):
        nt = self.case["cfg"]["nt"]
        e = oracle_engine(self.case, shot=self.shot, prec=self.prec)
        e.forward_begin()
        e.forward_advance(nt)
        e.adjoint_begin(self.r)
        e.adjoint_advance(nt)
        return e.field(9)[self.n2:-self.n2, self.n2:-self.n2].astype(np.float32)

    def model_shape(self):
        return self.shape
