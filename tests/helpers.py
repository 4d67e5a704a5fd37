"""Shared test helpers: conversions from the product's parameter blocks to
the oracle's, and error norms."""
import numpy as np

from oracle import oracle as O


def o_window(win):
    return O.Window(*[getattr(win, f) for f, _ in O.Window._fields_])


def o_bparams(sv):
    c = sv.cfg
    return O.BParams(c.nt, c.dt, o_window(sv.win), c.K, sv.dtstar, c.fc, c.threshold)


def o_tgeom(g):
    return O.tgeom(**g)


def o_iparams(sv):
    ip = O.IParams()
    for f, _ in O.IParams._fields_:
        if f != "tg":
            setattr(ip, f, getattr(sv.ip, f))
    ip.tg = o_tgeom(sv.cfg.tgeom)
    return ip


def o_model(m):
    if m["kind"] == 0:
        return O.model_vz(m["v0"], m["kgrad"], m["q0"])
    return O.model_layers(m["vl"], m["ql"], m["layer_dz"],
                          (m["ax"], m["ay"], m["az"], m["asig"], m["av"], m["aq"]))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---- SEG-Y test data ---------------------------------------------------------
# Standard (1-based start, width) of the header keys the tests write.
STD_KEYS = {"tracl": (1, 4), "fldr": (9, 4), "cdp": (21, 4), "offset": (37, 4), "scalco": (71, 2),
            "sx": (73, 4), "sy": (77, 4), "gx": (81, 4), "gy": (85, 4), "ns": (115, 2), "dt": (117, 2)}


def ibm_encode(values):
    """IBM System/360 words (uint32, host order) of float values, truncating
    the fraction: |v| = F / 2^24 * 16^(E-64) with 16^(E-65) <= |v| < 16^(E-64).
    Exact for values whose significant bits fit the hex-normalised
    fraction (float32 values with their three low mantissa bits clear)."""
    v = np.asarray(values, np.float64).ravel()
    a = np.abs(v)
    nz = a > 0
    _, e = np.frexp(np.where(nz, a, 1.0))                  # |v| = m 2^e, m in [0.5, 1)
    E = -((-e.astype(np.int64)) // 4)                      # 16^E >= 2^e > |v|
    F = np.floor(np.ldexp(a, (24 - 4 * E).astype(np.int32))).astype(np.uint64)
    assert np.all((F[nz] > 0) & (F[nz] < 2 ** 24)) and np.all((E[nz] + 64 >= 0) & (E[nz] + 64 < 128))
    w = (np.where(v < 0, 1, 0).astype(np.uint64) << 31) | ((E + 64).astype(np.uint64) << 24) | F
    return np.where(nz, w, 0).astype(np.uint32)


def ibm_value(words):
    """Value of IBM words from the definition, in float64 (exact)."""
    w = np.asarray(words, np.uint64)
    s = np.where(w >> 31, -1.0, 1.0)
    E = ((w >> 24) & 0x7f).astype(np.int64) - 64
    F = (w & 0xffffff).astype(np.float64)
    return s * F / 2.0 ** 24 * np.power(16.0, E.astype(np.float64))


def segy_records(samples, fields, fmt, keys=None, rng=None):
    """Trace records (240-byte header + big-endian samples) with the given
    header fields {name: int array} at keys {name: (start, width)} (default
    STD_KEYS); other header bytes random when rng is given, else zero.
    fmt 1 writes IBM words (ibm_encode), 5 big-endian IEEE."""
    samples = np.ascontiguousarray(samples, np.float32)
    ntr, ns = samples.shape
    keys = dict(STD_KEYS, **(keys or {}))
    rec = np.zeros((ntr, 240 + 4 * ns), np.uint8)
    if rng is not None:
        rec[:, :240] = rng.integers(0, 256, (ntr, 240), dtype=np.uint8)
    for name, vals in fields.items():
        start, width = keys[name]
        v = np.asarray(vals, np.int64) & ((1 << (8 * width)) - 1)
        for b in range(width):
            rec[:, start - 1 + b] = (v >> (8 * (width - 1 - b))) & 0xff
    words = ibm_encode(samples) if fmt == 1 else samples.view(np.uint32).ravel()
    rec[:, 240:] = words.astype(">u4").view(np.uint8).reshape(ntr, 4 * ns)
    return rec
