"""ctypes wrapper of the CPU oracle (liboracle.so) and of the reference's own
host sources (oracle/_ref/libpsdmref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline.  The product package (q_psdm_amd) never imports this module.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libpsdmref.so")


class Window(C.Structure):
    _fields_ = [("nsb", C.c_int), ("itb1", C.c_int), ("itb2", C.c_int), ("nf1", C.c_int),
                ("nf4", C.c_int), ("nww", C.c_int), ("ncoef1", C.c_int), ("nf3", C.c_int),
                ("tdstart1", C.c_float), ("tdstart2", C.c_float), ("tdmid", C.c_float),
                ("df", C.c_float), ("dw", C.c_float)]


class Model(C.Structure):
    _fields_ = [("kind", C.c_int), ("v0", C.c_double), ("kgrad", C.c_double), ("q0", C.c_double),
                ("nlayer", C.c_int), ("layer_dz", C.c_double), ("vl", C.c_double * 16),
                ("ql", C.c_double * 16), ("ax", C.c_double), ("ay", C.c_double),
                ("az", C.c_double), ("asig", C.c_double), ("av", C.c_double), ("aq", C.c_double)]


class TGeom(C.Structure):
    _fields_ = [("nnx", C.c_int), ("nny", C.c_int), ("n0x", C.c_float), ("n0y", C.c_float),
                ("ndx", C.c_float), ("ndy", C.c_float), ("nu", C.c_int), ("nv", C.c_int),
                ("nw", C.c_int), ("u0", C.c_float), ("v0", C.c_float), ("du", C.c_float),
                ("dv", C.c_float), ("dw", C.c_float)]


class IParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_float),
                ("dy", C.c_float), ("dz", C.c_float), ("nt", C.c_int), ("dt", C.c_float),
                ("K", C.c_int), ("dtstar", C.c_float), ("tanx1", C.c_float),
                ("tanx2", C.c_float), ("tany1", C.c_float), ("tany2", C.c_float),
                ("taperzone", C.c_float), ("tmin", C.c_float), ("tg", TGeom)]


class BParams(C.Structure):
    _fields_ = [("nt", C.c_int), ("dt", C.c_float), ("win", Window), ("K", C.c_int),
                ("dtstar", C.c_float), ("fc", C.c_float), ("threshold", C.c_float)]


class Src(C.Structure):
    _fields_ = [("hdr", C.POINTER(C.c_float)), ("tab", C.POINTER(C.c_float)), ("nus", C.c_int),
                ("nvs", C.c_int)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        fp, ip, dp = C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.or_npfa.restype = C.c_int
        L.or_npfar.restype = C.c_int
        L.or_window_params.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                       C.POINTER(Window)]
        L.or_gain_coefs.argtypes = [C.c_float, fp, fp, fp, fp, fp]
        L.or_gain_limit.restype = C.c_double
        L.or_gain_limit.argtypes = [C.c_double, C.c_double]
        L.or_iaztab.argtypes = [ip, C.c_int]
        L.or_azimuth_index.restype = C.c_int
        L.or_azimuth_index.argtypes = [ip, C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_fetch_and_taper.argtypes = [fp, C.c_int, C.c_int, C.c_int, fp]
        L.or_trace_spectra.argtypes = [fp, C.c_int, C.c_float, C.POINTER(Window), fp, fp]
        L.or_qbank_f64.argtypes = [fp, C.c_int, C.POINTER(BParams), dp]
        L.or_qbank_f32.argtypes = [fp, C.c_int, C.POINTER(BParams), fp]
        L.or_ray_times.argtypes = [C.POINTER(Model)] + [C.c_double] * 5 + [dp, dp]
        L.or_ray_sigma.restype = C.c_double
        L.or_ray_sigma.argtypes = [C.POINTER(Model)] + [C.c_double] * 5
        L.or_build_tables.argtypes = [C.POINTER(Model), C.POINTER(TGeom), fp]
        L.or_lookup.argtypes = [C.POINTER(TGeom), fp] + [C.c_double] * 5 + [dp, dp]
        L.or_image_f64.restype = C.c_int64
        L.or_image_f64.argtypes = [dp, fp, C.c_int, fp, C.POINTER(IParams), dp]
        L.or_image_f32.restype = C.c_int64
        L.or_image_f32.argtypes = [fp, fp, C.c_int, fp, C.POINTER(IParams), fp]
        L.or_image_src_f64.restype = C.c_int64
        L.or_image_src_f64.argtypes = [dp, fp, C.c_int, fp, C.POINTER(Src), C.POINTER(IParams), dp]
        L.or_image_src_f32.restype = C.c_int64
        L.or_image_src_f32.argtypes = [fp, fp, C.c_int, fp, C.POINTER(Src), C.POINTER(IParams), fp]
        L.or_image_nodes_f32.restype = C.c_int64
        L.or_image_nodes_f32.argtypes = [fp, fp, C.c_int, fp, C.POINTER(Src), C.POINTER(IParams), C.c_float, fp]
        L.or_node_range.argtypes = [C.POINTER(IParams), C.c_float, ip, ip]
        L.or_node_range_aligned.argtypes = [C.POINTER(IParams), C.c_float, C.c_int, ip, ip]
        L.or_image_spans_f32.restype = C.c_int64
        L.or_image_spans_f32.argtypes = [fp, fp, C.c_int, fp, C.POINTER(Src), C.POINTER(IParams), C.c_float, ip,
                                         fp]
        L.or_set_level_rule.argtypes = [C.c_int]
        L.or_count.restype = C.c_int64
        L.or_count.argtypes = [fp, C.c_int, C.POINTER(IParams)]
        L.or_num_threads.restype = C.c_int
        L.or_set_num_threads.restype = None
        L.or_set_num_threads.argtypes = [C.c_int]
        L.or_gather_edges.argtypes = [C.c_int, C.c_double, fp]
        L.or_gathers_f32.argtypes = [fp, fp, C.c_int, fp, C.POINTER(IParams), C.c_void_p, C.c_int, fp, C.c_int,
                                     fp, C.c_int, fp, fp]
        L.or_ibm_to_ieee.restype = C.c_uint32
        L.or_ibm_to_ieee.argtypes = [C.c_uint32]
        L.or_segy_key.restype = C.c_longlong
        L.or_segy_key.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int]
        L.or_segy_samples.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own segy.cc/psdmpkg.cc/fft.cc, built by oracle/Makefile."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref/libpsdmref.so not built")
        R = C.CDLL(REF_PATH)
        fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int)
        R.ref_parseint.restype = C.POINTER(C.c_int)
        R.ref_parsetuple.restype = C.POINTER(C.c_float)
        R.ref_getscale.restype = C.c_float
        R.ref_getscale.argtypes = [C.c_short]
        R.ref_time_offset_mute.restype = C.c_float
        R.ref_time_offset_mute.argtypes = [fp, C.c_int, C.c_float]
        R.ref_strrtrim.restype = C.c_char_p
        _ref = R
    return _ref


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def window_params(ntl, dt, f1=1.0, f3=130.0, f4=140.0):
    w = Window()
    lib().or_window_params(ntl, dt, f1, f3, f4, C.byref(w))
    return w


def gain_coefs(threshold):
    v = [C.c_float() for _ in range(5)]
    lib().or_gain_coefs(threshold, *[C.byref(x) for x in v])
    return [x.value for x in v]


def gain_limit(x, G):
    return lib().or_gain_limit(float(x), float(G))


def iaztab(n=2001):
    t = np.zeros(n, np.int32)
    lib().or_iaztab(_i(t), n)
    return t


def azimuth_index(tab, sx, sy, gx, gy):
    return lib().or_azimuth_index(_i(tab), sx, sy, gx, gy)


def fetch_and_taper(data, it1, it2):
    data = np.ascontiguousarray(data, np.float32)
    out = np.zeros(it2 - it1 + 1, np.float32)
    lib().or_fetch_and_taper(_f(data), data.size, it1, it2, _f(out))
    return out


def trace_spectra(trace, dt, win):
    trace = np.ascontiguousarray(trace, np.float32)
    re = np.zeros(2 * win.nww, np.float32)
    im = np.zeros(2 * win.nww, np.float32)
    lib().or_trace_spectra(_f(trace), trace.size, dt, C.byref(win), _f(re), _f(im))
    return re, im


def bparams(nt, dt, win, K, dtstar, fc, threshold):
    return BParams(nt, dt, win, K, dtstar, fc, threshold)


def qbank(traces, bp, dtype=np.float64):
    traces = np.ascontiguousarray(traces, np.float32)
    ntr = traces.shape[0]
    out = np.zeros((ntr, bp.K, bp.nt), dtype)
    if dtype == np.float64:
        lib().or_qbank_f64(_f(traces), ntr, C.byref(bp), _d(out))
    else:
        lib().or_qbank_f32(_f(traces), ntr, C.byref(bp), _f(out))
    return out


def model_vz(v0, k, q):
    m = Model()
    m.kind, m.v0, m.kgrad, m.q0 = 0, v0, k, q
    return m


def model_layers(vl, ql, layer_dz, anomaly=None):
    m = Model()
    m.kind = 1
    m.nlayer = len(vl)
    m.layer_dz = layer_dz
    for i, (v, q) in enumerate(zip(vl, ql)):
        m.vl[i], m.ql[i] = v, q
    if anomaly is None:
        m.ax = m.ay = m.az = 0.0
        m.asig, m.av, m.aq = 1.0, 0.0, 0.0
    else:
        m.ax, m.ay, m.az, m.asig, m.av, m.aq = anomaly
    return m


def ray_times(m, sx, sy, x, y, z):
    t, ts = C.c_double(), C.c_double()
    lib().or_ray_times(C.byref(m), sx, sy, x, y, z, C.byref(t), C.byref(ts))
    return t.value, ts.value


def ray_sigma(m, sx, sy, x, y, z):
    return lib().or_ray_sigma(C.byref(m), sx, sy, x, y, z)


def tgeom(**kw):
    g = TGeom()
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def build_tables(m, g):
    tab = np.zeros((g.nnx * g.nny, g.nv, g.nu, g.nw, 2), np.float32)
    lib().or_build_tables(C.byref(m), C.byref(g), _f(tab))
    return tab


def lookup(g, tab, px, py, x, y, z):
    t, ts = C.c_double(), C.c_double()
    lib().or_lookup(C.byref(g), _f(tab), px, py, x, y, z, C.byref(t), C.byref(ts))
    return t.value, ts.value


def image(bank, geom, tab, ip, fp32=False, src=None):
    """Kirchhoff sum; returns (image[ny,nx,nz], update count).  src: the
    traces' per-shot source table as (hdr[4] = (sx, sy, su0, sv0), table
    [nvs][nus][nw][2]); traces whose source is not hdr's are not imaged."""
    geom = np.ascontiguousarray(geom, np.float32)
    tab = np.ascontiguousarray(tab, np.float32)
    ntr = geom.shape[0]
    sp = None
    if src is not None:
        hdr = np.ascontiguousarray(src[0], np.float32).reshape(4)
        st = np.ascontiguousarray(src[1], np.float32)
        assert st.ndim == 4 and st.shape[2] == ip.tg.nw and st.shape[3] == 2
        sp = Src(_f(hdr), _f(st), st.shape[1], st.shape[0])
        keep = (hdr, st)
    if fp32:
        bank = np.ascontiguousarray(bank, np.float32)
        img = np.zeros((ip.ny, ip.nx, ip.nz), np.float32)
        if sp is None:
            n = lib().or_image_f32(_f(bank), _f(geom), ntr, _f(tab), C.byref(ip), _f(img))
        else:
            n = lib().or_image_src_f32(_f(bank), _f(geom), ntr, _f(tab), C.byref(sp), C.byref(ip), _f(img))
    else:
        bank = np.ascontiguousarray(bank, np.float64)
        img = np.zeros((ip.ny, ip.nx, ip.nz), np.float64)
        if sp is None:
            n = lib().or_image_f64(_d(bank), _f(geom), ntr, _f(tab), C.byref(ip), _d(img))
        else:
            n = lib().or_image_src_f64(_d(bank), _f(geom), ntr, _f(tab), C.byref(sp), C.byref(ip), _d(img))
    if sp is not None:
        del keep
    return img, n


def node_range(ip, z_exact, align=1):
    """Points [ps, pe) the node form (align 1) or the span form (align 8)
    interpolates from depth nodes."""
    a, b = C.c_int(), C.c_int()
    lib().or_node_range_aligned(C.byref(ip), z_exact, align, C.byref(a), C.byref(b))
    return a.value, b.value


def image_nodes(bank, geom, tab, ip, z_exact, src=None):
    """Node form of the FP32 path (weight_nodes = 1): (image, count)."""
    geom = np.ascontiguousarray(geom, np.float32)
    tab = np.ascontiguousarray(tab, np.float32)
    bank = np.ascontiguousarray(bank, np.float32)
    img = np.zeros((ip.ny, ip.nx, ip.nz), np.float32)
    sp = None
    if src is not None:
        hdr = np.ascontiguousarray(src[0], np.float32).reshape(4)
        st = np.ascontiguousarray(src[1], np.float32)
        sp = Src(_f(hdr), _f(st), st.shape[1], st.shape[0])
    n = lib().or_image_nodes_f32(_f(bank), _f(geom), geom.shape[0], _f(tab), C.byref(sp) if sp else None,
                                 C.byref(ip), z_exact, _f(img))
    return img, n


def span_table(ip, z_exact, schedule):
    """Weight nodes (start, end) of every table interval, int [nw - 1][2],
    under the span form's schedule [(z, S), ...]: the node range's intervals
    [ws, we) (ws rounded up to a multiple of 8) split into bands, band k from the first interval at or below z_k
    rounded up to a multiple of 8 intervals past ws (intervals above the
    first band: one-interval spans; a band that a later one starts at the
    same interval is dropped), spans of S_k intervals from the band's start,
    the last cut at the band's end.  Intervals outside the range get
    (w, w + 1) (the oracle does not read them)."""
    nw, dw = ip.tg.nw, ip.tg.dw
    sp = np.array([[w, w + 1] for w in range(nw - 1)], np.int32)
    ps, pe = node_range(ip, z_exact, align=8)
    if ps >= pe:
        return sp
    ws = max(1, int(math.ceil(float(np.float32(z_exact)) / float(np.float32(dw)))))
    ws = (ws + 7) // 8 * 8
    inv_dw = np.float32(1.0) / np.float32(dw)
    fw = np.float32(pe - 1) * np.float32(ip.dz) * inv_dw
    fw = min(fw, np.float32(nw - 1))
    we = min(int(fw), nw - 2) + 1
    start, span = [], []
    for z, S in schedule:
        x = float(np.float32(z)) / float(np.float32(dw)) - ws
        b = min(ws + 8 * int(math.ceil(max(0.0, x) / 8.0)), we)
        if not start and b > ws:
            start.append(ws)
            span.append(1)
        if start and start[-1] == b:
            start.pop()
            span.pop()
        start.append(b)
        span.append(int(S))
    if start and start[-1] == we:
        start.pop()
        span.pop()
    ends = start[1:] + [we]
    for b, e, S in zip(start, ends, span):
        for w in range(b, e):
            w0 = b + (w - b) // S * S
            sp[w] = (w0, min(w0 + S, e))
    return sp


def image_spans(bank, geom, tab, ip, z_exact, spans, src=None):
    """Span form of the FP32 path (weight_nodes = 2): (image, count).
    spans: int [nw - 1][2], the weight nodes (start, end) of each table
    depth interval."""
    geom = np.ascontiguousarray(geom, np.float32)
    tab = np.ascontiguousarray(tab, np.float32)
    bank = np.ascontiguousarray(bank, np.float32)
    spans = np.ascontiguousarray(spans, np.int32)
    assert spans.shape == (ip.tg.nw - 1, 2)
    img = np.zeros((ip.ny, ip.nx, ip.nz), np.float32)
    sp = None
    if src is not None:
        hdr = np.ascontiguousarray(src[0], np.float32).reshape(4)
        st = np.ascontiguousarray(src[1], np.float32)
        sp = Src(_f(hdr), _f(st), st.shape[1], st.shape[0])
    n = lib().or_image_spans_f32(_f(bank), _f(geom), geom.shape[0], _f(tab), C.byref(sp) if sp else None,
                                 C.byref(ip), z_exact, _i(spans), _f(img))
    return img, n


def set_level_rule(rule):
    """Test-only: the FP64 sum's level rule ("round", the product's; "floor";
    "zero": always level 0)."""
    lib().or_set_level_rule({"round": 0, "floor": 1, "zero": 2}[rule])


def count(geom, ip):
    geom = np.ascontiguousarray(geom, np.float32)
    return lib().or_count(_f(geom), geom.shape[0], C.byref(ip))


def num_threads():
    return lib().or_num_threads()


def set_num_threads(n):
    lib().or_set_num_threads(int(n))


def ibm_to_ieee(words):
    """FP32 bits of big-endian-decoded IBM words (uint32 array)."""
    w = np.ascontiguousarray(words, np.uint32).ravel()
    f = lib().or_ibm_to_ieee
    return np.array([f(int(x)) for x in w], np.uint32)


def segy_samples(record, ns, fmt):
    """FP32 samples of one SEG-Y record (bytes: 240-byte header + samples)."""
    out = np.zeros(ns, np.uint32)
    rc = lib().or_segy_samples(bytes(record), ns, fmt, out.ctypes.data_as(C.c_void_p))
    assert rc == 0, fmt
    return out.view(np.float32)


def segy_key(header, start, nbytes, kind):
    return lib().or_segy_key(bytes(header[:240]), start, nbytes, kind)


def gather_edges(dip, ddip):
    """Boundaries of the 2 ndip + 1 bins, ndip = (int)(dip / ddip + 0.5)."""
    ndip = int(float(np.float32(dip) / np.float32(ddip)) + 0.5)
    e = np.zeros(2 * ndip + 2, np.float32)
    lib().or_gather_edges(ndip, float(np.float32(ddip)), _f(e))
    return e


def gathers(bank, geom, tab, ip, cols, ex=None, ey=None):
    """FP32 dip-angle gathers at cols [ncol][2] (ix, iy): (gx, gy)."""
    bank = np.ascontiguousarray(bank, np.float32)
    geom = np.ascontiguousarray(geom, np.float32)
    tab = np.ascontiguousarray(tab, np.float32)
    cols = np.ascontiguousarray(cols, np.int32)
    ncol = cols.shape[0]
    nbx = 0 if ex is None else ex.size - 1
    nby = 0 if ey is None else ey.size - 1
    gx = np.zeros((ncol, max(nbx, 1), ip.nz), np.float32)
    gy = np.zeros((ncol, max(nby, 1), ip.nz), np.float32)
    dummy = np.zeros(1, np.float32)
    lib().or_gathers_f32(_f(bank), _f(geom), geom.shape[0], _f(tab), C.byref(ip), cols.ctypes.data_as(C.c_void_p),
                         ncol, _f(ex if ex is not None else dummy), nbx, _f(ey if ey is not None else dummy), nby,
                         _f(gx), _f(gy))
    return (gx if nbx else None), (gy if nby else None)
