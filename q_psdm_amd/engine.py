"""Host-side mirror of the reference's imaging operator over libqpsdm.so.

The reference drives one trace at a time through psdm_kernel()
(src/q-psdm.cu:2063-2244) with ~85 scalar arguments; here the same state is
grouped into a Survey (window, bank, image and table parameters derived once
per offset class, as qpsdm() does at src/q-psdm.cu:1593-1665) and the
operators work on batches of device-resident traces.  torch supplies device
memory and the current stream; every computation runs in libqpsdm's HIP
kernels (no CPU or torch fallback).
"""
import ctypes as C
import math

import numpy as np
import torch

from . import lib as L
from .lib import check


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ptr(t):
    assert t.is_cuda and t.is_contiguous(), "device tensors must be contiguous CUDA tensors"
    return C.c_void_p(t.data_ptr())


def window(nt, dt, f1, f3, f4):
    w = L.Window()
    check(L.lib().qpsdm_window_params(nt, dt, f1, f3, f4, C.byref(w)))
    return w


def npfar(n):
    return L.lib().qpsdm_npfar(n)


def gain_coefs(threshold):
    out = (C.c_float * 5)()
    check(L.lib().qpsdm_gain_coefs(threshold, out))
    return list(out)


def iaztab(n=2001):
    t = (C.c_int * n)()
    check(L.lib().qpsdm_iaztab(t, n))
    return np.frombuffer(t, dtype=np.int32).copy()


def model_struct(m):
    s = L.Model()
    s.kind, s.v0, s.kgrad, s.q0 = m["kind"], m["v0"], m["kgrad"], m["q0"]
    s.nlayer, s.layer_dz = m["nlayer"], m["layer_dz"]
    for i, (v, q) in enumerate(zip(m["vl"], m["ql"])):
        s.vl[i], s.ql[i] = v, q
    s.ax, s.ay, s.az, s.asig, s.av, s.aq = m["ax"], m["ay"], m["az"], m["asig"], m["av"], m["aq"]
    return s


def tgeom_struct(g):
    s = L.TGeom()
    for k, v in g.items():
        setattr(s, k, v)
    return s


def clamp_band(dt, f3, f4):
    a, b = C.c_float(f3), C.c_float(f4)
    check(L.lib().qpsdm_clamp_band(dt, C.byref(a), C.byref(b)))
    return a.value, b.value


class Survey:
    """Per-offset-class state of the imaging path for one configuration."""

    def __init__(self, cfg, tstar_max):
        self.cfg = cfg
        f3, f4 = clamp_band(cfg.dt, cfg.f3, cfg.f4)
        self.win = window(cfg.nt, cfg.dt, cfg.f1, f3, f4)
        self.tstar_max = float(tstar_max)
        self.dtstar = self.tstar_max / max(cfg.K - 1, 1)
        bp = L.BankParams()
        bp.nt, bp.dt, bp.win, bp.K = cfg.nt, cfg.dt, self.win, cfg.K
        bp.dtstar, bp.fc, bp.threshold = self.dtstar, cfg.fc, cfg.threshold
        self.bp = bp
        ip = L.ImageParams()
        ip.nx, ip.ny, ip.nz, ip.dx, ip.dy, ip.dz = cfg.nx, cfg.ny, cfg.nz, cfg.dx, cfg.dy, cfg.dz
        ip.nt, ip.dt, ip.K, ip.dtstar = cfg.nt, cfg.dt, cfg.K, self.dtstar
        ip.tanx1, ip.tanx2, ip.tany1, ip.tany2, ip.taperzone = cfg.dip_bounds()
        ip.tmin = cfg.tmin
        ip.tg = tgeom_struct(cfg.tgeom)
        self.ip = ip
        self.tg = ip.tg

    DEFAULT_SPANS = ((0.0, 2), (800.0, 4), (1600.0, 8))

    def set_form(self, form="exact", z_exact=400.0, spans=None):
        """Form of the imaging sum: "exact" (every point evaluates its taper
        and amplitude), "nodes" (below z_exact metres, the taper ramps and
        the amplitude are evaluated at the tables' depth nodes and
        interpolated; times and levels stay exact) or "spans" (the same with
        the ramps and amplitude at the ends of spans of table intervals,
        interpolated in 1/z; spans = [(z, S), ...]: from depth z down, spans
        of S intervals; batched imaging with pair banks only).  Returns self."""
        assert form in ("exact", "nodes", "spans")
        self.ip.weight_nodes = {"exact": 0, "nodes": 1, "spans": 2}[form]
        self.ip.z_exact = float(z_exact)
        if form == "spans":
            sp = tuple(spans if spans is not None else self.DEFAULT_SPANS)
            assert 1 <= len(sp) <= 4
            self.ip.span_n = len(sp)
            for k, (z, S) in enumerate(sp):
                self.ip.span_z[k] = float(z)
                self.ip.span_s[k] = int(S)
        return self

    @property
    def form(self):
        return ("exact", "nodes", "spans")[self.ip.weight_nodes]

    def pair_bank(self):
        """The form's bank layout: pair banks (sample and difference) for the
        span form, plain banks otherwise."""
        return self.form == "spans"

    def spans(self):
        """The span schedule [(z, S), ...] of the span form."""
        return [(float(self.ip.span_z[k]), int(self.ip.span_s[k])) for k in range(self.ip.span_n)]

    # -- tables ----------------------------------------------------------
    def table_shape(self):
        g = self.cfg.tgeom
        return (g["nnx"] * g["nny"], g["nv"], g["nu"], g["nw"], 2)

    def build_tables(self, stream=None):
        tab = torch.empty(self.table_shape(), dtype=torch.float32, device="cuda")
        check(L.lib().qpsdm_tables_build(C.byref(model_struct(self.cfg.model)), C.byref(self.tg),
                                         _ptr(tab), _stream(stream)))
        return tab

    def ref_layout(self, nxi0=16, nzi0=8, line1=1):
        """The reference-layout lattice of this survey's node tables: the
        (du, dv, dw) lattice on every column, line and depth node that some
        node's relative table reaches (the union of their footprints, so the
        files hold every table entry), nxi0 columns x nzi0 depth nodes per
        cell.  The node spacing must be a multiple of du and dv."""
        g = self.cfg.tgeom
        kx, ky = g["ndx"] / g["du"], g["ndy"] / g["dv"]
        assert g["nnx"] == 1 or abs(kx - round(kx)) < 1e-6, "node spacing not on the du lattice"
        assert g["nny"] == 1 or abs(ky - round(ky)) < 1e-6, "node spacing not on the dv lattice"
        ncol = (g["nnx"] - 1) * int(round(kx)) + g["nu"]
        nlines = (g["nny"] - 1) * int(round(ky)) + g["nv"]
        return L.RefLayout(line1, nlines, 0, ncol, g["nw"], nxi0, nzi0, g["n0x"] + g["u0"], g["n0y"] + g["v0"],
                           g["du"], g["dv"], g["dw"])

    def write_ref_tables(self, tab, layout, directory):
        """Write FP32 node tables as Lxxx.tidx/qidx + L<n>.ttt/qqq (the
        reference's table files) in the layout's convention."""
        assert tab.dtype == torch.float32 and tuple(tab.shape) == self.table_shape()
        torch.cuda.synchronize()
        check(L.lib().qpsdm_tables_write_ref(C.byref(self.tg), _ptr(tab), C.byref(layout), directory.encode()))

    def read_ref_tables(self, layout, directory, out=None):
        """FP32 node tables from reference-layout table files."""
        if out is None:
            out = torch.zeros(self.table_shape(), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        check(L.lib().qpsdm_tables_read_ref(C.byref(self.tg), C.byref(layout), directory.encode(), _ptr(out)))
        return out

    def quantize_tables(self, tab, code="f16", stream=None):
        """16-bit brick tables; code "f16" (FP16 codes in 8^3 bricks), "u16"
        (fixed point in 8^3 bricks), "u16x" (fixed point in edge-sharing
        9x9x8 bricks) or "u16q" (the u16x codes stored as per-cell corner
        words).  Returned as a uint8 tensor tagged with .qpsdm_code."""
        c = {"f16": L.QPSDM_F16B, "u16": L.QPSDM_U16B, "u16x": L.QPSDM_U16X, "u16q": L.QPSDM_U16Q}[code]
        nbytes = L.lib().qpsdm_tables_bytes(C.byref(self.tg), c)
        out = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        check(L.lib().qpsdm_tables_quantize(C.byref(self.tg), _ptr(tab), _ptr(out), c, _stream(stream)))
        out.qpsdm_code = c
        return out

    def lookup(self, tab, px, py, xyz, stream=None):
        """(t, t*) [n][2] at points xyz [n][3] from the surface position
        (px, py), read from tab in its storage type as the imaging kernel does."""
        _, tcode = self._codes(torch.empty((1, self.cfg.K, self.cfg.nt), device="meta"), tab)
        assert xyz.dtype == torch.float32 and xyz.dim() == 2 and xyz.shape[1] == 3
        out = torch.empty((xyz.shape[0], 2), dtype=torch.float32, device="cuda")
        check(L.lib().qpsdm_tables_lookup(C.byref(self.ip), _ptr(tab), tcode, px, py, _ptr(xyz), xyz.shape[0],
                                          _ptr(out), _stream(stream)))
        return out

    # -- kernel (a) --------------------------------------------------------
    def trace_spectra(self, traces, stream=None):
        n = traces.shape[0]
        re = torch.empty((n, 2 * self.win.nww), dtype=torch.float32, device="cuda")
        im = torch.empty_like(re)
        check(L.lib().qpsdm_trace_spectra(C.byref(self.bp), _ptr(traces), n, _ptr(re), _ptr(im),
                                          _stream(stream)))
        return re, im

    _BANK_CODES = {torch.float32: L.QPSDM_F32, torch.bfloat16: L.QPSDM_BF16, torch.float16: L.QPSDM_F16}

    def bank_shape(self, n, dtype=torch.float32, pair=False):
        """Tensor shape of an n-trace bank: (n, K, nt) for FP32 and BF16; for
        FP16 (n, row) halves per trace, a 16-byte header {1/scale, scale}
        then the K x nt scaled samples.  Pair banks (pair=True): FP32
        (n, K, nt, 2) of (sample, difference to the next); FP16 (n, row)
        halves, the header then K x nt half pairs."""
        c = self.cfg
        if pair:
            if dtype == torch.float32:
                return (n, c.K, c.nt, 2)
            assert dtype == torch.float16, "pair banks are FP32 or FP16"
            return (n, L.lib().qpsdm_bank_row_bytes(c.K, c.nt, L.QPSDM_F16P) // 2)
        if dtype == torch.float16:
            return (n, L.lib().qpsdm_bank_row_bytes(c.K, c.nt, L.QPSDM_F16) // 2)
        return (n, c.K, c.nt)

    def bank_code(self, bank):
        """Storage code of a bank tensor (pair banks are told by their shape)."""
        if bank.dtype == torch.float32 and bank.dim() == 4:
            return L.QPSDM_F32P
        if bank.dtype == torch.float16 and bank.dim() == 2 and \
                bank.shape[1] == self.bank_shape(1, torch.float16, pair=True)[1]:
            return L.QPSDM_F16P
        return self._BANK_CODES[bank.dtype]

    def _check_bank(self, bank, rows):
        assert bank.dtype in self._BANK_CODES and bank.is_contiguous() and bank.shape[0] >= rows
        pair = self.bank_code(bank) in (L.QPSDM_F32P, L.QPSDM_F16P)
        assert tuple(bank.shape[1:]) == self.bank_shape(1, bank.dtype, pair)[1:]

    def decode_bank(self, bank):
        """Bank in filtered-trace units as FP32 (n, K, nt) (host-side check);
        of a pair bank, the samples."""
        c = self.cfg
        code = self.bank_code(bank)
        if code == L.QPSDM_F32P:
            return bank[..., 0].contiguous()
        if code == L.QPSDM_F16P:
            hdr = bank[:, :8].contiguous().view(torch.float32)
            v = bank[:, 8:8 + 2 * c.K * c.nt].float().reshape(-1, c.K, c.nt, 2)[..., 0]
            return v * hdr[:, 0].reshape(-1, 1, 1)
        if bank.dtype != torch.float16:
            return bank.float()
        hdr = bank[:, :8].contiguous().view(torch.float32)  # [n][4]: 1/scale, scale, 0, 0
        return bank[:, 8:8 + c.K * c.nt].float().reshape(-1, c.K, c.nt) * hdr[:, 0].reshape(-1, 1, 1)

    def qbank(self, traces, dtype=torch.float32, out=None, stream=None, pair=False):
        n = traces.shape[0]
        assert traces.dtype == torch.float32 and traces.dim() == 2 and traces.shape[1] == self.cfg.nt
        if out is None:
            out = torch.empty(self.bank_shape(n, dtype, pair), dtype=dtype, device="cuda")
        self._check_bank(out, n)
        code = self.bank_code(out)
        check(L.lib().qpsdm_qbank(C.byref(self.bp), _ptr(traces), n, _ptr(out), code, _stream(stream)))
        return out

    # -- kernel (b) --------------------------------------------------------
    def new_image(self):
        c = self.cfg
        return torch.zeros((c.ny, c.nx, c.nz), dtype=torch.float32, device="cuda")

    def image_shot(self, img, geom, tr0, ntr, bank, tab, launch=None, stream=None):
        """Accumulate the run geom[tr0:tr0+ntr] into img; bank[0:ntr] are the
        run's filtered traces."""
        c = self.cfg
        assert img.shape == (c.ny, c.nx, c.nz) and img.dtype == torch.float32
        assert geom.dim() == 2 and geom.shape[1] == 4 and 0 <= tr0 and tr0 + ntr <= geom.shape[0]
        self._check_bank(bank, ntr)
        bcode, tcode = self._codes(bank, tab)
        lp = None
        if launch is not None:
            lp = L.Launch(*launch[:3])
        check(L.lib().qpsdm_image_shot(C.byref(self.ip), C.byref(lp) if lp else None, _ptr(img),
                                       _ptr(geom), tr0, ntr, _ptr(bank), bcode, _ptr(tab), tcode,
                                       _stream(stream)))

    _ASPECTS = {"all": 1, "x": 2, "y": 3, "none": 4}

    def gather_params(self, dipx=50.0, dipy=50.0, ddipx=1.0, ddipy=1.0, aspect="all"):
        """The reference's gather options (dipx, dipy, ddipx, ddipy in degrees,
        aspect All / X / Y / None) and the bins per axis they give."""
        g = L.GatherParams(dipx, dipy, ddipx, ddipy, self._ASPECTS[aspect.lower()])
        nbx, nby = C.c_int(), C.c_int()
        check(L.lib().qpsdm_gather_bins(C.byref(g), C.byref(nbx), C.byref(nby)))
        return g, nbx.value, nby.value

    def new_gathers(self, ncol, nb):
        return torch.zeros((ncol, nb, self.cfg.nz), dtype=torch.float32, device="cuda") if nb else None

    def gathers_shot(self, gx, gy, cols, geom, tr0, ntr, bank, tab, gp, stream=None):
        """Accumulate the dip-angle gathers of the run at image columns cols
        (int32 device tensor [ncol][2] of (ix, iy)) into gx [ncol][nbx][nz]
        and gy [ncol][nby][nz] (either None when its axis is not asked for)."""
        assert cols.dtype == torch.int32 and cols.dim() == 2 and cols.shape[1] == 2
        assert geom.dim() == 2 and geom.shape[1] == 4 and 0 <= tr0 and tr0 + ntr <= geom.shape[0]
        self._check_bank(bank, ntr)
        g, nbx, nby = gp
        ncol = cols.shape[0]
        for t, nb in ((gx, nbx), (gy, nby)):
            if nb:
                assert t is not None and tuple(t.shape) == (ncol, nb, self.cfg.nz) and t.dtype == torch.float32
        bcode, tcode = self._codes(bank, tab)
        check(L.lib().qpsdm_gathers_shot(C.byref(self.ip), C.byref(g), _ptr(cols), ncol,
                                         _ptr(gx) if nbx else None, _ptr(gy) if nby else None, _ptr(geom), tr0, ntr,
                                         _ptr(bank), bcode, _ptr(tab), tcode, _stream(stream)))

    def count_shot(self, geom, tr0, ntr, counter=None, stream=None):
        assert geom.dim() == 2 and geom.shape[1] == 4 and 0 <= tr0 and tr0 + ntr <= geom.shape[0]
        own = counter is None
        if own:
            counter = torch.zeros(1, dtype=torch.int64, device="cuda")
        check(L.lib().qpsdm_count_shot(C.byref(self.ip), _ptr(geom), tr0, ntr, _ptr(counter),
                                       _stream(stream)))
        return int(counter.item()) if own else counter

    def scale_image(self, img, s, stream=None):
        check(L.lib().qpsdm_image_scale(_ptr(img), img.numel(), s, _stream(stream)))

    # -- batched imaging -----------------------------------------------------
    def _codes(self, bank, tab):
        if tab.dtype == torch.float32:
            assert tuple(tab.shape) == self.table_shape()
            tcode = L.QPSDM_F32
        else:
            tcode = getattr(tab, "qpsdm_code", L.QPSDM_F16B)
            assert tab.numel() == L.lib().qpsdm_tables_bytes(C.byref(self.tg), tcode)
        return self.bank_code(bank), tcode

    def image_batch(self, img, geom, bank, tab, shots, launch=None, nwg=0, work=None, src=None, stream=None):
        """Accumulate every shot of a batch into img with the persistent
        kernel; shots is a ShotList (bank row = geom row); src: the shots'
        source tables from Surrogate.src_tables (None: node tables)."""
        c = self.cfg
        assert img.shape == (c.ny, c.nx, c.nz) and img.dtype == torch.float32
        assert geom.dim() == 2 and geom.shape[1] == 4 and shots.rows <= geom.shape[0]
        self._check_bank(bank, shots.rows)
        bcode, tcode = self._codes(bank, tab)
        if work is None:
            work = torch.empty(1, dtype=torch.int32, device="cuda")
        lp = L.Launch(*launch[:3]) if launch is not None else None
        if src is not None:
            assert src.numel() >= L.lib().qpsdm_surrogate_src_bytes(C.byref(self.ip), shots.n)
        check(L.lib().qpsdm_image_batch_src(C.byref(self.ip), C.byref(lp) if lp else None, _ptr(img), _ptr(geom),
                                            _ptr(bank), bcode, _ptr(tab), tcode, _ptr(shots.dev), shots.n,
                                            _ptr(src) if src is not None else None, _ptr(work), nwg,
                                            _stream(stream)))

    def src_grid(self):
        """(nus, nvs, nw) of one shot's source table: the relative sub-grid
        that reaches the image."""
        c = self.cfg
        g = c.tgeom
        nus = min(g["nu"], math.ceil((c.nx - 1) * c.dx / g["du"]) + 2)
        nvs = min(g["nv"], math.ceil((c.ny - 1) * c.dy / g["dv"]) + 2)
        return nus, nvs, g["nw"]

    def src_bytes(self, nshot):
        return L.lib().qpsdm_surrogate_src_bytes(C.byref(self.ip), nshot)

    def src_tables_nodes(self, geom, shots, tab, out=None, stream=None):
        """Per-shot source tables from the FP32 node tables alone: each value
        the bilinear blend of the four surface nodes around the shot's
        source (uint8 buffer: headers + tables, as Surrogate.src_tables)."""
        assert tab.dtype == torch.float32 and tuple(tab.shape) == self.table_shape()
        assert geom.dim() == 2 and geom.shape[1] == 4 and shots.rows <= geom.shape[0]
        if out is None:
            out = torch.empty(self.src_bytes(shots.n), dtype=torch.uint8, device="cuda")
        assert out.numel() >= self.src_bytes(shots.n)
        check(L.lib().qpsdm_src_tables_nodes(C.byref(self.ip), _ptr(geom), _ptr(shots.dev), shots.n, _ptr(tab),
                                             _ptr(out), _stream(stream)))
        return out

    def count_batch(self, geom, shots, stream=None):
        counter = torch.zeros(1, dtype=torch.int64, device="cuda")
        check(L.lib().qpsdm_count_batch(C.byref(self.ip), _ptr(geom), _ptr(shots.dev), shots.n, _ptr(counter),
                                        _stream(stream)))
        return int(counter.item())


class ShotList:
    """Device array of qpsdm_shot descriptors: (tr0, ntr) rows of geom/bank
    and the shot's dip bounds as tangents."""

    def __init__(self, shots, device="cuda"):
        arr = (L.Shot * max(len(shots), 1))()
        for i, s in enumerate(shots):
            arr[i] = L.Shot(*s)
        raw = bytes(arr)
        self.n = len(shots)
        self.rows = max((s[0] + s[1] for s in shots), default=0)
        self.host = list(shots)
        self.dev = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)

    def update(self, shots):
        """Overwrite the descriptors in place (same count), for graph replays."""
        assert len(shots) == self.n
        arr = (L.Shot * max(self.n, 1))()
        for i, s in enumerate(shots):
            arr[i] = L.Shot(*s)
        self.dev.copy_(torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8))
        self.rows = max((s[0] + s[1] for s in shots), default=0)
        self.host = list(shots)


class BatchGraph:
    """A captured graph of [Q bank of traces -> (surrogate source tables) ->
    persistent imaging of shots].  The buffers are bound at capture; replays
    read their current contents."""

    def __init__(self, survey, traces, bank, img, geom, tab, shots, launch=None, nwg=0, surrogate=None,
                 tab_f32=None, mask=None, src=None, stages=7):
        c = survey.cfg
        assert traces.dtype == torch.float32 and traces.shape[1] == c.nt and traces.is_contiguous()
        survey._check_bank(bank, traces.shape[0])
        assert img.shape == (c.ny, c.nx, c.nz)
        assert shots.rows <= traces.shape[0] <= geom.shape[0]
        bcode, tcode = survey._codes(bank, tab)
        self.work = torch.empty(1, dtype=torch.int32, device="cuda")
        if surrogate is not None and src is None:
            src = surrogate.new_src(survey, shots.n)
        self.src = src
        self._keep = (traces, bank, img, geom, tab, shots, surrogate, tab_f32, mask, src)
        lp = L.Launch(*launch[:3]) if launch is not None else None
        h = C.c_void_p()
        opt = lambda t: _ptr(t) if t is not None else None
        torch.cuda.synchronize()
        check(L.lib().qpsdm_batch_graph_create(
            C.byref(survey.bp), _ptr(traces), traces.shape[0], _ptr(bank), bcode, C.byref(survey.ip),
            C.byref(lp) if lp else None, _ptr(img), _ptr(geom), _ptr(tab), tcode, _ptr(shots.dev), shots.n,
            surrogate.h if surrogate is not None else None, opt(tab_f32), opt(mask), opt(src),
            _ptr(self.work), nwg, stages, C.byref(h)))
        self.h = h

    def launch(self, stream=None):
        check(L.lib().qpsdm_graph_launch(self.h, _stream(stream)))

    def close(self):
        if getattr(self, "h", None):
            L.lib().qpsdm_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_f32(src, dst, stream=None):
    """Device copy by the library's plain copy kernel (counter calibration)."""
    assert src.dtype == dst.dtype == torch.float32 and src.numel() == dst.numel()
    check(L.lib().qpsdm_copy_f32(_ptr(src), _ptr(dst), src.numel(), _stream(stream)))


def build_tables_into(model, tgeom, out, stream=None):
    """Straight-ray tables of tgeom (dict) into the float32 tensor out."""
    g = tgeom_struct(tgeom)
    n = g.nnx * g.nny * g.nv * g.nu * g.nw * 2
    assert out.dtype == torch.float32 and out.numel() >= n
    check(L.lib().qpsdm_tables_build(C.byref(model_struct(model)), C.byref(g), _ptr(out), _stream(stream)))
    return out


def rays(model, pts, stream=None):
    """FP64 straight-ray integrals (t, t*, sigma) [n][3] of the model from the
    surface points (sx, sy) to (x, y, z), pts [n][5] (numpy or a float64
    device tensor); returned as a float64 device tensor."""
    if not isinstance(pts, torch.Tensor):
        pts = torch.from_numpy(np.ascontiguousarray(pts, np.float64))
    p = pts.to("cuda", torch.float64).contiguous()
    assert p.dim() == 2 and p.shape[1] == 5, "rays: pts is [n][5] (sx, sy, x, y, z)"
    out = torch.empty((p.shape[0], 3), dtype=torch.float64, device="cuda")
    check(L.lib().qpsdm_rays(C.byref(model_struct(model)), _ptr(p), p.shape[0], _ptr(out), _stream(stream)))
    return out


def tstar_max_from_tables(tab):
    """Largest two-leg t* the tables allow (2 x the largest table t*)."""
    return 2.0 * float(tab[..., 1].max())


SEGY_BYTES = 240 + 4 * 65535


def su_encode(samples, geom_xy, scalco=0, fldr=1, cdp0=1):
    """Big-endian SU records (the reference's fputtr byte layout) for traces
    samples [ntr][ns] with integer header coordinates geom_xy [ntr][4]
    (sx, sy, gx, gy), through the library's own codec."""
    samples = np.ascontiguousarray(samples, np.float32)
    ntr, ns = samples.shape
    rec = 240 + 4 * ns
    out = bytearray(ntr * rec)
    tr = (C.c_ubyte * SEGY_BYTES)()
    hdr = np.zeros(60, np.int32)
    buf = (C.c_ubyte * rec)()
    for i in range(ntr):
        C.memset(tr, 0, 240)
        h = np.frombuffer(tr, dtype=np.uint8, count=240)
        hv = h.view(np.int32)
        hv[0] = i + 1                                      # tracl
        hv[2] = fldr                                       # fldr
        hv[5] = cdp0 + i                                   # cdp
        h[70:72] = np.array([scalco], np.int16).view(np.uint8)
        hv[18:22] = geom_xy[i]                             # sx sy gx gy
        h[114:116] = np.array([ns], np.uint16).view(np.uint8)
        h[116:118] = np.array([2000], np.uint16).view(np.uint8)
        C.memmove(C.addressof(tr) + 240, samples[i].ctypes.data, 4 * ns)
        L.lib().qpsdm_su_sputtr(buf, tr)
        out[i * rec:(i + 1) * rec] = bytes(buf)
    del hdr
    return bytes(out)


class Pipeline:
    """The qpsdm_ctx pipeline: pinned staging slots, copy + compute streams,
    SU decode on the GPU, Q bank and imaging into a context-owned image."""

    def __init__(self, survey, geom_params=None, bank_dtype="f32", table_dtype="f32",
                 batch_traces=1024, launch=None):
        p = L.CtxParams()
        p.image, p.bank = survey.ip, survey.bp
        if geom_params is None:
            geom_params = L.GeomParams(1.0, -1e30, 1e30, -1e30, 1e30, 0.0, 0.0, 0.0, 0.0,
                                       survey.cfg.dx, survey.cfg.dy)
        p.geom = geom_params
        p.bank_dtype = {"f32": L.QPSDM_F32, "bf16": L.QPSDM_BF16, "f16": L.QPSDM_F16, "f32p": L.QPSDM_F32P,
                        "f16p": L.QPSDM_F16P}[bank_dtype]
        p.table_dtype = {"f32": L.QPSDM_F32, "f16": L.QPSDM_F16B, "u16": L.QPSDM_U16B,
                         "u16x": L.QPSDM_U16X, "u16q": L.QPSDM_U16Q}[table_dtype]
        p.batch_traces = batch_traces
        if launch is not None:
            p.shape = L.Launch(*launch[:3])
        self.survey = survey
        self.params = p
        self.ctx = C.c_void_p()
        check(L.lib().qpsdm_ctx_create(C.byref(p), C.byref(self.ctx)))

    def close(self):
        if self.ctx:
            L.lib().qpsdm_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        self.close()

    def set_tables(self, tab):
        on_dev = isinstance(tab, torch.Tensor) and tab.is_cuda
        ptr = C.c_void_p(tab.data_ptr()) if isinstance(tab, torch.Tensor) else \
            np.ascontiguousarray(tab, np.float32).ctypes.data_as(C.c_void_p)
        check(L.lib().qpsdm_ctx_set_tables(self.ctx, ptr, 1 if on_dev else 0))

    def build_tables(self):
        check(L.lib().qpsdm_ctx_build_tables(self.ctx, C.byref(model_struct(self.survey.cfg.model))))

    def push_su(self, data):
        check(L.lib().qpsdm_ctx_push_su(self.ctx, data, len(data)))

    def push_segy(self, data, fmt=1, keyidx=None):
        """SEG-Y rev0 trace records (no reel header), format 1 (IBM) or 5
        (IEEE), keys at keyidx (segy.keyindex(); None: standard)."""
        check(L.lib().qpsdm_ctx_push_segy(self.ctx, data, len(data), fmt, keyidx))

    def push_traces(self, samples, geom):
        samples = np.ascontiguousarray(samples, np.float32)
        geom = np.ascontiguousarray(geom, np.float32)
        check(L.lib().qpsdm_ctx_push_traces(self.ctx, samples.ctypes.data_as(C.c_void_p),
                                            geom.ctypes.data_as(C.c_void_p), samples.shape[0]))

    def psdm_kernel(self, segy_record):
        """One decoded trace given as SU record bytes (header + samples)."""
        tr = (C.c_ubyte * SEGY_BYTES)()
        L.lib().qpsdm_su_sgettr(segy_record, tr)
        check(L.lib().qpsdm_ctx_psdm_kernel(self.ctx, tr))

    def set_gathers(self, cols, dipx=50.0, dipy=50.0, ddipx=1.0, ddipy=1.0, aspect="all"):
        """Dip-angle gathers at image columns cols [ncol][2] (ix, iy),
        accumulated beside the image from the next batch on."""
        cols = np.ascontiguousarray(cols, np.int32).reshape(-1, 2)
        g, nbx, nby = self.survey.gather_params(dipx, dipy, ddipx, ddipy, aspect)
        check(L.lib().qpsdm_ctx_set_gathers(self.ctx, C.byref(g), cols.ctypes.data_as(C.c_void_p), cols.shape[0]))
        self._gshape = (cols.shape[0], nbx, nby)

    def gathers(self, normalize=False):
        ncol, nbx, nby = self._gshape
        nz = self.survey.cfg.nz
        gx = np.zeros((ncol, nbx, nz), np.float32) if nbx else None
        gy = np.zeros((ncol, nby, nz), np.float32) if nby else None
        check(L.lib().qpsdm_ctx_gathers(self.ctx, gx.ctypes.data_as(C.c_void_p) if nbx else None,
                                        gy.ctypes.data_as(C.c_void_p) if nby else None, 1 if normalize else 0))
        return gx, gy

    def image(self, normalize=False):
        """The image [ny][nx][nz] in host memory (divided by nsb when
        normalize).  The array lives in pinned memory (torch's pinned-block
        cache), which the device copies into several times faster than into
        pageable memory."""
        c = self.survey.cfg
        out = torch.empty((c.ny, c.nx, c.nz), dtype=torch.float32, pin_memory=True).numpy()
        check(L.lib().qpsdm_ctx_image(self.ctx, out.ctypes.data_as(C.c_void_p), 1 if normalize else 0))
        return out

    def reset_image(self):
        check(L.lib().qpsdm_ctx_reset_image(self.ctx))

    def traces_pushed(self):
        return L.lib().qpsdm_ctx_traces_pushed(self.ctx)


class Surrogate:
    """The learned source leg: MLP 5-128-128-128-128-3 on the device.

    box_lo/box_hi: input box of (sx, sy, x, y, z); out_scale: residual scale
    of (t, t*, log A) over the anomaly-free closed form."""

    def __init__(self, model, box_lo, box_hi, out_scale, seed=0x5EED0005):
        self.nparams = L.lib().qpsdm_surrogate_nparams()
        h = C.c_void_p()
        check(L.lib().qpsdm_surrogate_create(C.byref(model_struct(model)), (C.c_float * 5)(*box_lo),
                                             (C.c_float * 5)(*box_hi), (C.c_float * 3)(*out_scale), seed,
                                             C.byref(h)))
        self.h = h
        self.box_lo, self.box_hi, self.out_scale = list(box_lo), list(box_hi), list(out_scale)

    @staticmethod
    def for_survey(cfg, out_scale=(0.02, 0.004, 0.05), margin=100.0, seed=0x5EED0005):
        """Input box: sources over the image's surface and the table nodes,
        points over the image grid widened by margin, depths over the table
        depth range."""
        g = cfg.tgeom
        # the source tables reach up to two table nodes past the image
        margin = max(margin, 2.5 * max(g["du"], g["dv"]))
        ex, ey = (cfg.nx - 1) * cfg.dx, (cfg.ny - 1) * cfg.dy
        zmax = (g["nw"] - 1) * g["dw"]
        sxl, sxh = min(0.0, g["n0x"]), max(ex, g["n0x"] + (g["nnx"] - 1) * g["ndx"])
        syl, syh = min(0.0, g["n0y"]), max(ey, g["n0y"] + (g["nny"] - 1) * g["ndy"])
        lo = (sxl, syl, -margin, -margin, 0.0)
        hi = (sxh, syh, ex + margin, ey + margin, zmax)
        return Surrogate(cfg.model, lo, hi, out_scale, seed)

    def close(self):
        if getattr(self, "h", None):
            L.lib().qpsdm_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(self.nparams, np.float32)
        check(L.lib().qpsdm_surrogate_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        check(L.lib().qpsdm_surrogate_set_params(self.h, p.ctypes.data))

    def samples(self, n, seed, offset=0, stream=None):
        X = torch.empty((n, 5), dtype=torch.float32, device="cuda")
        T = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        check(L.lib().qpsdm_surrogate_samples(self.h, seed, offset, n, _ptr(X), _ptr(T), _stream(stream)))
        return X, T

    def eval(self, X, stream=None):
        assert X.dtype == torch.float32 and X.dim() == 2 and X.shape[1] == 5
        out = torch.empty((X.shape[0], 3), dtype=torch.float32, device="cuda")
        check(L.lib().qpsdm_surrogate_eval(self.h, _ptr(X), X.shape[0], _ptr(out), _stream(stream)))
        return out

    def grad(self, X, T, stream=None):
        g = torch.empty(self.nparams, dtype=torch.float32, device="cuda")
        loss = C.c_float()
        check(L.lib().qpsdm_surrogate_grad(self.h, _ptr(X), _ptr(T), X.shape[0], _ptr(g), C.byref(loss),
                                           _stream(stream)))
        return g, loss.value

    def train(self, X, T, steps, batch, lr0=2e-3, lr1=2e-5, stream=None):
        hist = np.zeros(max(steps, 1), np.float32)
        check(L.lib().qpsdm_surrogate_train(self.h, _ptr(X), _ptr(T), X.shape[0], steps, batch, lr0, lr1,
                                            hist.ctypes.data, _stream(stream)))
        return hist[:steps]

    def mask(self, tg, tab, tol=0.5e-3, stream=None):
        """Validity per (node, 8^3 brick) against the FP32 node tables:
        (mask uint8, max |dt| per brick float32), both on the device."""
        nb = L.lib().qpsdm_surrogate_mask_bricks(C.byref(tg))
        m = torch.empty(nb, dtype=torch.uint8, device="cuda")
        e = torch.empty(nb, dtype=torch.float32, device="cuda")
        check(L.lib().qpsdm_surrogate_mask(self.h, C.byref(tg), _ptr(tab), tol, _ptr(m), _ptr(e),
                                           _stream(stream)))
        return m, e

    def new_src(self, survey, nshot):
        nb = L.lib().qpsdm_surrogate_src_bytes(C.byref(survey.ip), nshot)
        return torch.empty(nb, dtype=torch.uint8, device="cuda")

    def src_tables(self, survey, geom, shots, tab=None, mask=None, out=None, stream=None):
        """Source tables of the shots (header + tables, one uint8 buffer)."""
        if out is None:
            out = self.new_src(survey, shots.n)
        assert tab is None or tab.dtype == torch.float32
        assert geom.dim() == 2 and geom.shape[1] == 4 and shots.rows <= geom.shape[0]
        assert out.numel() >= L.lib().qpsdm_surrogate_src_bytes(C.byref(survey.ip), shots.n)
        check(L.lib().qpsdm_surrogate_src_tables(self.h, C.byref(survey.ip), _ptr(geom), _ptr(shots.dev),
                                                 shots.n, _ptr(tab) if tab is not None else None,
                                                 _ptr(mask) if mask is not None else None, _ptr(out),
                                                 _stream(stream)))
        return out

    def src_grid(self, survey):
        """(nus, nvs, nw) of one shot's source table."""
        return survey.src_grid()
