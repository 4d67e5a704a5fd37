"""SEG-Y rev0 reader over libqpsdm.so: reel header, key positions, IBM/IEEE
trace records on the host (the reference's fsgethd / initkeyindex /
setkeyindex / fsgettr, src/segy.cc:83-2752) and the batch decode on the
device.  No computation happens in Python: every byte goes through the
library's codec or its HIP decode kernel."""
import ctypes as C
import os

import numpy as np

from . import lib as L
from .lib import check

IBM, IEEE = 1, 5
HDR_DTYPE = np.dtype([("sx", "<i4"), ("sy", "<i4"), ("gx", "<i4"), ("gy", "<i4"), ("fldr", "<i4"),
                      ("cdp", "<i4"), ("offset", "<i4"), ("tracl", "<i4"), ("scalco", "<i2"), ("pad", "<i2"),
                      ("ns", "<u2"), ("dt", "<u2")])
assert HDR_DTYPE.itemsize == C.sizeof(L.TrHdr) == 40


def keyindex(**moves):
    """Standard key positions (initkeyindex), then each name=(start, bytes)
    moved as setkeyindex does (start 1-based)."""
    k = L.KeyIndexArray()
    check(L.lib().qpsdm_segy_initkeyindex(k))
    for name, (start, nbytes) in moves.items():
        check(L.lib().qpsdm_segy_setkeyindex(k, name.encode(), int(start), int(nbytes)))
    return k


def _as_bytes(data):
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data).view(np.uint8)
    return np.frombuffer(bytes(data), np.uint8)


def decode(data, ns, fmt=IBM, keyidx=None):
    """Equal-length trace records (no reel header) -> (samples [ntr][ns]
    float32, headers [ntr] of HDR_DTYPE), on the host."""
    b = _as_bytes(data)
    rec = 240 + 4 * ns
    n = b.size // rec
    samples = np.zeros((n, ns), np.float32)
    hdrs = np.zeros(n, HDR_DTYPE)
    got = L.lib().qpsdm_segy_decode(b.ctypes.data_as(C.c_void_p), b.size, ns, fmt, keyidx,
                                   samples.ctypes.data_as(C.c_void_p), hdrs.ctypes.data_as(C.c_void_p))
    if got < 0:
        check(int(got))
    return samples[:got], hdrs[:got]


def decode_dev(raw, ntr, ns, fmt=IBM, keyidx=None, stream=None, headers=True):
    """The same on the device: raw is a CUDA uint8 tensor of ntr records;
    returns (samples [ntr][ns] float32 tensor, headers [ntr][10] int32
    tensor holding HDR_DTYPE rows, or None)."""
    import torch

    from .engine import _ptr, _stream
    assert raw.dtype == torch.uint8 and raw.numel() >= ntr * (240 + 4 * ns)
    samples = torch.empty((ntr, ns), dtype=torch.float32, device=raw.device)
    hdrs = torch.empty((ntr, 10), dtype=torch.int32, device=raw.device) if headers else None
    check(L.lib().qpsdm_segy_decode_dev(_ptr(raw), ntr, ns, fmt, keyidx, _ptr(samples),
                                        _ptr(hdrs) if headers else None, _stream(stream)))
    return samples, hdrs


def headers_view(hdrs_int32):
    """HDR_DTYPE view of decode_dev's header rows (host numpy copy)."""
    return np.ascontiguousarray(hdrs_int32.cpu().numpy()).view(HDR_DTYPE).reshape(-1)


def read_reel_header(path):
    """fsgethd of a file: (text bytes (ASCII as the reference maps them),
    binary-header dict)."""
    buf = (C.c_ubyte * 3600)()
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    fp = libc.fopen(os.fsencode(path), b"rb")
    if not fp:
        raise OSError(f"cannot open {path}")
    try:
        check(L.lib().qpsdm_segy_fgethd(C.c_void_p(fp), buf))
    finally:
        libc.fclose(C.c_void_p(fp))
    raw = bytes(buf)
    ints = np.frombuffer(raw, "<i4", 3, 3200)
    shorts = np.frombuffer(raw, "<i2", 24, 3212)
    names = ("ntrpr nart hdt dto hns nso format fold tsort vscode hsfs hsfe hslen hstyp schn hstas hstae htatyp "
             "hcorr bgrcv rcvm mfeet polyt vpol").split()
    bh = {"jobid": int(ints[0]), "lino": int(ints[1]), "reno": int(ints[2])}
    bh.update({n: int(v) for n, v in zip(names, shorts)})
    return raw[:3200], bh


def read_file(path, keyidx=None, fmt=None):
    """A whole SEG-Y file (3600-byte reel header + equal-length records):
    format from the binary header unless given; -> (text, binary header,
    samples, headers), decoded on the host."""
    text, bh = read_reel_header(path)
    fmt = fmt or bh["format"]
    with open(path, "rb") as f:
        f.seek(3600)
        body = f.read()
    if not body:
        return text, bh, np.zeros((0, 0), np.float32), np.zeros(0, HDR_DTYPE)
    k = keyidx or keyindex()
    ns_key = k[38]
    o = ns_key.start - 1
    ns = int.from_bytes(body[o:o + 2], "big") if ns_key.bytes == 2 else int.from_bytes(body[o:o + 4], "big") & 0xffff
    samples, hdrs = decode(body, ns, fmt, k)
    return text, bh, samples, hdrs
