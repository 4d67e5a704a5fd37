"""Synthetic survey configurations and data generators.

The reference ships no data generator and its example inputs are absent, so
every workload is generated here from a fixed seed (0x5EED0000 + config
number):

* config 1 -- 101^3 image at 10 m, v(z) = 1500 + 0.6 z, Q = 100, 16 shots on a
  4x4 grid at 200 m with 32x32 receiver patches at 25 m, 1001 samples at 2 ms;
  25 Hz Ricker diffractions from 20 point scatterers with straight-ray v(z)
  traveltimes, constant-Q attenuation and 1/(Ls Lr) spreading; 20 m tables;
  K = 16, gain cap 40 dB, aperture 45 degrees.
* config 2 -- 401 x 401 x 301 image at 10 m, six 500 m layers (v 1600..4200,
  Q 40..200) plus a 500 m-sigma Gaussian anomaly (-20 % v, x0.5 Q) at the
  model centre, 200 shots x 2000 receivers jittered over 4 x 4 km, 2001
  samples at 2 ms, band-limited reflectivity traces with 5 % noise; 25 m
  tables; K = 32.

Coordinates are image-relative metres (the image origin is (0, 0, 0)).
"""
import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Config:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    nt: int
    dt: float
    K: int
    f1: float = 1.0
    f3: float = 130.0
    f4: float = 140.0
    fc: float = 40.0
    threshold: float = 100.0     # gain cap G (40 dB)
    aperture_deg: float = 45.0
    contract_deg: float = 3.0
    taper_deg: float = 8.0
    tmin: float = 1e-4
    model: dict = field(default_factory=dict)
    tgeom: dict = field(default_factory=dict)
    seed: int = 0x5EED0001

    def dip_bounds(self):
        a = math.radians(self.aperture_deg - self.contract_deg)
        return -math.tan(a), math.tan(a), -math.tan(a), math.tan(a), math.radians(self.taper_deg)


def _vz_model(v0, k, q):
    return dict(kind=0, v0=v0, kgrad=k, q0=q, nlayer=0, layer_dz=1.0, vl=[], ql=[],
                ax=0.0, ay=0.0, az=0.0, asig=1.0, av=0.0, aq=0.0)


def _layer_model(vl, ql, dz, anomaly):
    ax, ay, az, sig, av, aq = anomaly
    return dict(kind=1, v0=vl[0], kgrad=0.0, q0=ql[0], nlayer=len(vl), layer_dz=dz, vl=list(vl),
                ql=list(ql), ax=ax, ay=ay, az=az, asig=sig, av=av, aq=aq)


def config1():
    c = Config("config1", 101, 101, 101, 10.0, 10.0, 10.0, 1001, 0.002, 16, seed=0x5EED0001)
    c.model = _vz_model(1500.0, 0.6, 100.0)
    # v(z) is laterally invariant: one relative-coordinate table serves every
    # surface position; offsets up to 1187.5 m, depth to 1000 m, 20 m grid
    c.tgeom = dict(nnx=1, nny=1, n0x=500.0, n0y=500.0, ndx=1000.0, ndy=1000.0,
                   nu=121, nv=121, nw=51, u0=-1200.0, v0=-1200.0, du=20.0, dv=20.0, dw=20.0)
    return c


def config2():
    c = Config("config2", 401, 401, 301, 10.0, 10.0, 10.0, 2001, 0.002, 32, seed=0x5EED0002)
    c.model = _layer_model([1600, 2120, 2640, 3160, 3680, 4200], [40, 72, 104, 136, 168, 200],
                           500.0, (2000.0, 2000.0, 1500.0, 500.0, 0.2, 0.5))
    # 5 x 5 surface nodes at 1 km; relative offsets +-4 km; depth 3 km; 25 m
    c.tgeom = dict(nnx=5, nny=5, n0x=0.0, n0y=0.0, ndx=1000.0, ndy=1000.0,
                   nu=321, nv=321, nw=121, u0=-4000.0, v0=-4000.0, du=25.0, dv=25.0, dw=25.0)
    return c


def config5(seed=0x5EED0005):
    """Config 2's layers and anomaly with each layer's Q drawn from U[30, 200]
    (seeded): the model the learned surrogate is trained on and images with."""
    c = config2()
    c.name = "config5"
    c.seed = seed
    q = np.random.default_rng(seed).uniform(30.0, 200.0, len(c.model["vl"]))
    c.model = dict(c.model, ql=[float(v) for v in q], q0=float(q[0]))
    return c


def tiny():
    """Small v(z) case that the FP64 oracle images in well under a second."""
    c = Config("tiny", 24, 20, 30, 10.0, 10.0, 10.0, 501, 0.002, 8, f1=3.0, seed=0x5EED0000)
    c.model = _vz_model(1500.0, 0.6, 60.0)
    c.tgeom = dict(nnx=1, nny=1, n0x=100.0, n0y=100.0, ndx=500.0, ndy=500.0,
                   nu=41, nv=41, nw=17, u0=-400.0, v0=-400.0, du=20.0, dv=20.0, dw=20.0)
    return c


def tiny_layers():
    """Small laterally varying case (layers + anomaly, 3 x 3 nodes)."""
    c = Config("tiny_layers", 20, 16, 24, 10.0, 10.0, 10.0, 501, 0.002, 8, f1=3.0, seed=0x5EED0005)
    c.model = _layer_model([1600, 2000, 2400], [50, 80, 120], 80.0, (100.0, 80.0, 120.0, 80.0, 0.2, 0.5))
    c.tgeom = dict(nnx=3, nny=3, n0x=0.0, n0y=0.0, ndx=100.0, ndy=100.0,
                   nu=31, nv=31, nw=13, u0=-300.0, v0=-300.0, du=20.0, dv=20.0, dw=20.0)
    return c


# ---------------------------------------------------------------- geometry
def shots_config1():
    """16 shots on a 4x4 grid at 200 m; 32x32 receivers at 25 m per shot."""
    off = -387.5 + 25.0 * np.arange(32)
    gx, gy = np.meshgrid(off, off, indexing="xy")
    out = []
    for sy in (200.0, 400.0, 600.0, 800.0):
        for sx in (200.0, 400.0, 600.0, 800.0):
            out.append((sx, sy, np.stack([sx + gx.ravel(), sy + gy.ravel()], 1)))
    return out


def shots_config2(nshots=200, nrec=2000, seed=0x5EED0002):
    """Shots on a jittered 20x10 grid, receivers on a jittered 50x40 grid,
    both over 4 x 4 km with +-40 % cell jitter; receivers redrawn per shot."""
    rng = np.random.default_rng(seed)
    sxg, syg = np.meshgrid((np.arange(20) + 0.5) * 200.0, (np.arange(10) + 0.5) * 400.0,
                           indexing="xy")
    sxs = sxg.ravel() + rng.uniform(-0.4, 0.4, sxg.size) * 200.0
    sys_ = syg.ravel() + rng.uniform(-0.4, 0.4, syg.size) * 400.0
    gxg, gyg = np.meshgrid((np.arange(50) + 0.5) * 80.0, (np.arange(40) + 0.5) * 100.0,
                           indexing="xy")
    out = []
    for s in range(nshots):
        r = np.random.default_rng((seed, s))
        gx = gxg.ravel() + r.uniform(-0.4, 0.4, gxg.size) * 80.0
        gy = gyg.ravel() + r.uniform(-0.4, 0.4, gyg.size) * 100.0
        out.append((float(sxs[s % sxs.size]), float(sys_[s % sys_.size]),
                    np.stack([gx[:nrec], gy[:nrec]], 1)))
    return out


def shots_tiny(nshots=2, nrec=6, seed=0x5EED0000, extent=(240.0, 200.0)):
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nshots):
        sx, sy = rng.uniform(0, extent[0]), rng.uniform(0, extent[1])
        g = np.stack([rng.uniform(-60, extent[0] + 60, nrec), rng.uniform(-60, extent[1] + 60, nrec)], 1)
        out.append((float(sx), float(sy), g))
    return out


def dip_tans(aperture_deg, contract_deg=3.0):
    """Dip bounds (tanx1, tanx2, tany1, tany2) of a symmetric opening angle."""
    t = math.tan(math.radians(aperture_deg - contract_deg))
    return (-t, t, -t, t)


def shots_config4(nshots=500, seed=0x5EED0004, extent=(4000.0, 4000.0)):
    """Uneven shot gathers: trace counts floor(exp(U[ln 200, ln 4000])) and a
    per-shot opening angle U[25, 60] deg; shots and receivers uniformly
    random over the 4 x 4 km patch.  Returns [(sx, sy, g[n, 2], aperture)]."""
    rng = np.random.default_rng(seed)
    n = np.floor(np.exp(rng.uniform(math.log(200.0), math.log(4000.0), nshots))).astype(int)
    ap = rng.uniform(25.0, 60.0, nshots)
    out = []
    for s in range(nshots):
        r = np.random.default_rng((seed, s))
        sx, sy = r.uniform(0, extent[0]), r.uniform(0, extent[1])
        g = np.stack([r.uniform(0, extent[0], n[s]), r.uniform(0, extent[1], n[s])], 1)
        out.append((float(sx), float(sy), g, float(ap[s])))
    return out


def geometry(shots):
    """Per-trace (sx, sy, gx, gy) float32 [ntr, 4] and shot ranges [(tr0, n)]."""
    rows, ranges, t0 = [], [], 0
    for sx, sy, g, *_ in shots:
        n = g.shape[0]
        a = np.empty((n, 4), np.float32)
        a[:, 0], a[:, 1], a[:, 2], a[:, 3] = sx, sy, g[:, 0], g[:, 1]
        rows.append(a)
        ranges.append((t0, n))
        t0 += n
    return np.concatenate(rows, 0), ranges


# ------------------------------------------------------------ straight rays
def ray_times_vz(v0, k, q, sx, sy, x, y, z):
    """Closed-form straight-ray t and t* = t/Q in v(z) = v0 + k z (FP64)."""
    L = np.sqrt((x - sx) ** 2 + (y - sy) ** 2 + z ** 2)
    kz = k * z
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(kz > 1e-9 * v0, L / np.where(kz > 0, kz, 1.0) * np.log((v0 + kz) / v0), L / v0)
    return t, t / q


# ------------------------------------------------------------ trace data
def ricker_spectrum(f, fp):
    """Fourier transform of a unit Ricker wavelet of peak frequency fp."""
    return 2.0 / math.sqrt(math.pi) * f ** 2 / fp ** 3 * np.exp(-(f / fp) ** 2)


def scatterers(cfg, n=20):
    rng = np.random.default_rng(cfg.seed)
    x0, x1 = 0.1 * (cfg.nx - 1) * cfg.dx, 0.9 * (cfg.nx - 1) * cfg.dx
    y0, y1 = 0.1 * (cfg.ny - 1) * cfg.dy, 0.9 * (cfg.ny - 1) * cfg.dy
    z0, z1 = 0.2 * (cfg.nz - 1) * cfg.dz, 0.9 * (cfg.nz - 1) * cfg.dz
    return np.stack([rng.uniform(x0, x1, n), rng.uniform(y0, y1, n), rng.uniform(z0, z1, n)], 1)


def diffraction_traces(cfg, geom, scat, fpk=25.0, amp=1e6, chunk=512):
    """Point-scatterer data in v(z): for each scatterer a Ricker arriving at
    ts + tg, attenuated by exp(-w t*/2) with Kolsky dispersion about fc and
    scaled by 1/(Ls Lr); synthesised in the frequency domain (FP64)."""
    m = cfg.model
    assert m["kind"] == 0, "diffraction data are defined for the v(z) model"
    nfft = 1 << int(math.ceil(math.log2(2 * cfg.nt)))
    f = np.fft.rfftfreq(nfft, cfg.dt)
    w = 2 * math.pi * f
    wr = 2 * math.pi * cfg.fc
    W = ricker_spectrum(f, fpk)
    lnw = np.log(np.where(w > 0, w, 1.0) / wr)
    out = np.zeros((geom.shape[0], cfg.nt), np.float32)
    for a in range(0, geom.shape[0], chunk):
        g = geom[a:a + chunk].astype(np.float64)
        U = np.zeros((g.shape[0], f.size), np.complex128)
        for sx_, sy_, sz_ in scat:
            ts, qs = ray_times_vz(m["v0"], m["kgrad"], m["q0"], g[:, 0], g[:, 1], sx_, sy_, sz_)
            tg, qg = ray_times_vz(m["v0"], m["kgrad"], m["q0"], g[:, 2], g[:, 3], sx_, sy_, sz_)
            Ls = np.sqrt((g[:, 0] - sx_) ** 2 + (g[:, 1] - sy_) ** 2 + sz_ ** 2)
            Lr = np.sqrt((g[:, 2] - sx_) ** 2 + (g[:, 3] - sy_) ** 2 + sz_ ** 2)
            tau = (ts + tg)[:, None]
            tst = (qs + qg)[:, None]
            ph = -w[None, :] * tau + w[None, :] * tst * lnw[None, :] / math.pi
            U += (amp / (Ls * Lr))[:, None] * W[None, :] * np.exp(-w[None, :] * tst / 2) * np.exp(1j * ph)
        # continuous FT -> samples: x[n] = (1/dt) * irfft-normalised sum
        x = np.fft.irfft(U, nfft, axis=1) / cfg.dt
        out[a:a + chunk] = x[:, :cfg.nt]
    return out


def diffraction_traces_rays(cfg, geom, scat, fpk=25.0, amp=1e6, chunk=2048):
    """Point-scatterer data through any of the table builder's models (the
    layered + anomaly one included): per scatterer and trace, the source and
    receiver legs' t and t* from the library's FP64 straight-ray integrals
    (engine.rays), a Ricker of peak fpk arriving at ts + tg, attenuated by
    exp(-w t*/2) with Kolsky dispersion about fc, scaled by 1/(Ls Lr);
    synthesised in the frequency domain in FP64 on the GPU.  Returns a
    float32 device tensor [ntr, nt]."""
    import torch
    from . import engine
    nfft = 1 << int(math.ceil(math.log2(2 * cfg.nt)))
    f = np.fft.rfftfreq(nfft, cfg.dt)
    w = torch.from_numpy(2 * math.pi * f).cuda()
    W = torch.from_numpy(ricker_spectrum(f, fpk)).cuda()
    lnw = torch.log(torch.where(w > 0, w, torch.ones_like(w)) / (2 * math.pi * cfg.fc))
    g = torch.as_tensor(np.asarray(geom, np.float64)).cuda()
    ntr = g.shape[0]
    legs = []
    for x, y, z in scat:
        p = torch.tensor([x, y, z], dtype=torch.float64, device="cuda").expand(ntr, 3)
        rs = engine.rays(cfg.model, torch.cat([g[:, 0:2], p], 1))
        rg = engine.rays(cfg.model, torch.cat([g[:, 2:4], p], 1))
        ls = torch.sqrt((g[:, 0] - x) ** 2 + (g[:, 1] - y) ** 2 + z * z)
        lr = torch.sqrt((g[:, 2] - x) ** 2 + (g[:, 3] - y) ** 2 + z * z)
        legs.append((rs[:, 0] + rg[:, 0], rs[:, 1] + rg[:, 1], amp / (ls * lr)))
    out = torch.empty((ntr, cfg.nt), dtype=torch.float32, device="cuda")
    for a0 in range(0, ntr, chunk):
        sl = slice(a0, min(a0 + chunk, ntr))
        U = torch.zeros((sl.stop - sl.start, f.size), dtype=torch.complex128, device="cuda")
        for tau, tst, amp_ in legs:
            tau, tst = tau[sl, None], tst[sl, None]
            ph = -w[None, :] * tau + w[None, :] * tst * lnw[None, :] / math.pi
            U += (amp_[sl, None] * W[None, :] * torch.exp(-w[None, :] * tst / 2)) * torch.polar(torch.ones_like(ph), ph)
        out[sl] = (torch.fft.irfft(U, nfft, dim=1) / cfg.dt)[:, :cfg.nt].float()
    return out


def reflectivity_traces(cfg, ntr, seed, qeff=100.0, noise=0.05, nwin=8):
    """Band-limited 5-60 Hz reflectivity, attenuated per time window with
    t* = t_window/qeff and Kolsky dispersion, plus 5 % Gaussian noise."""
    rng = np.random.default_rng(seed)
    nt = cfg.nt
    r = rng.laplace(0.0, 1.0, (ntr, nt)) * (rng.random((ntr, nt)) < 0.05)
    nfft = 1 << int(math.ceil(math.log2(2 * nt)))
    f = np.fft.rfftfreq(nfft, cfg.dt)
    band = np.clip((f - 3.0) / 4.0, 0, 1) * np.clip((65.0 - f) / 10.0, 0, 1)
    w = 2 * math.pi * f
    lnw = np.log(np.where(w > 0, w, 1.0) / (2 * math.pi * cfg.fc))
    out = np.zeros((ntr, nt))
    edges = np.linspace(0, nt, nwin + 1).astype(int)
    for k in range(nwin):
        seg = np.zeros((ntr, nfft))
        seg[:, edges[k]:edges[k + 1]] = r[:, edges[k]:edges[k + 1]]
        tc = 0.5 * (edges[k] + edges[k + 1]) * cfg.dt
        tst = tc / qeff
        H = band * np.exp(-w * tst / 2) * np.exp(1j * w * tst * lnw / math.pi)
        out += np.fft.irfft(np.fft.rfft(seg, axis=1) * H[None, :], nfft, axis=1)[:, :nt]
    rms = np.sqrt(np.mean(out ** 2)) + 1e-30
    out += noise * rms * rng.standard_normal(out.shape)
    return out.astype(np.float32)


def reflectivity_traces_gpu(cfg, ntr, seed, qeff=100.0, noise=0.05, nwin=8, device="cuda"):
    """reflectivity_traces on the GPU with torch's seeded generator (another
    random stream than the numpy version, same construction): sparse Laplace
    reflectivity, band-limited 5-60 Hz, attenuated per time window with
    t* = t_window/qeff and Kolsky dispersion, plus 5 % Gaussian noise.
    float32 [ntr][nt] on the device."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed) & 0x7FFFFFFFFFFFFFFF)
    nt = cfg.nt
    u = torch.rand((ntr, nt), generator=g, device=device, dtype=torch.float64) - 0.5
    lap = -torch.sign(u) * torch.log1p(-2.0 * u.abs())
    keep = torch.rand((ntr, nt), generator=g, device=device, dtype=torch.float64) < 0.05
    r = torch.where(keep, lap, torch.zeros_like(lap))
    nfft = 1 << int(math.ceil(math.log2(2 * nt)))
    f = torch.fft.rfftfreq(nfft, cfg.dt, device=device, dtype=torch.float64)
    band = torch.clamp((f - 3.0) / 4.0, 0, 1) * torch.clamp((65.0 - f) / 10.0, 0, 1)
    w = 2 * math.pi * f
    lnw = torch.log(torch.where(w > 0, w, torch.ones_like(w)) / (2 * math.pi * cfg.fc))
    out = torch.zeros((ntr, nt), dtype=torch.float64, device=device)
    edges = np.linspace(0, nt, nwin + 1).astype(int)
    for k in range(nwin):
        seg = torch.zeros((ntr, nfft), dtype=torch.float64, device=device)
        seg[:, edges[k]:edges[k + 1]] = r[:, edges[k]:edges[k + 1]]
        tst = 0.5 * (edges[k] + edges[k + 1]) * cfg.dt / qeff
        H = band * torch.exp(-w * tst / 2) * torch.exp(1j * w * tst * lnw / math.pi)
        out += torch.fft.irfft(torch.fft.rfft(seg, dim=1) * H[None, :], nfft, dim=1)[:, :nt]
    rms = torch.sqrt(torch.mean(out ** 2)) + 1e-30
    out += noise * rms * torch.randn(out.shape, generator=g, device=device, dtype=torch.float64)
    return out.to(torch.float32).contiguous()
