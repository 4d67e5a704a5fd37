"""SEG-Y rev0 reader (host side, no GPU): the IBM -> IEEE conversion pinned
by its definition, known words and the reference's own reader (oracle/_ref,
compiled from src/segy.cc); header keys at moved positions, the reel
header's EBCDIC text and binary header, and setkeyindex's lookup, each
against the reference build; files written by the reference's writer."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from q_psdm_amd import lib as L
from q_psdm_amd import segy
from tests.helpers import ibm_encode, ibm_value, segy_records

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")


def _ref_keys(moves=()):
    k = np.zeros(150, np.int32)
    R = O.ref()
    R.ref_initkeyindex(k.ctypes.data_as(C.c_void_p))
    for name, start, nb in moves:
        R.ref_setkeyindex(k.ctypes.data_as(C.c_void_p), name.encode(), start, nb)
    return k


def _ref_read(path, fmt, keys, ntr, ns, skip=0):
    hdrs = np.zeros((ntr, 240), np.uint8)
    data = np.zeros((ntr, ns), np.float32)
    n = O.ref().ref_fsgettr_all(str(path).encode(), skip, fmt, keys.ctypes.data_as(C.c_void_p), ntr, ns,
                                hdrs.ctypes.data_as(C.c_void_p), data.ctypes.data_as(C.c_void_p))
    return n, hdrs, data


@pytest.fixture(scope="module", autouse=True)
def _ref_signatures():
    if O.ref_available():
        R = O.ref()
        vp = C.c_void_p
        R.ref_initkeyindex.argtypes = [vp]
        R.ref_setkeyindex.argtypes = [vp, C.c_char_p, C.c_int, C.c_int]
        R.ref_fsgethd.argtypes = [C.c_char_p, vp]
        R.ref_fsgettr_all.argtypes = [C.c_char_p, C.c_long, C.c_int, vp, C.c_int, C.c_int, vp, vp]
        R.ref_fsput_file.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, vp]


# ---- IBM -> IEEE ----------------------------------------------------------------

KNOWN = [(0x42640000, 100.0), (0xC276A000, -118.625), (0x41100000, 1.0), (0x40800000, 0.5),
         (0x3F800000, 0.03125), (0xC1100000, -1.0), (0x42100000, 16.0), (0x41200000, 2.0)]


@pytest.mark.parametrize("w,v", KNOWN)
def test_ibm_known_words(w, v):
    got = O.ibm_to_ieee(np.array([w], np.uint32)).view(np.float32)[0]
    assert got == np.float32(v)
    assert ibm_value([w])[0] == v


def test_ibm_matches_definition_over_the_float32_range():
    """Random words whose value lies in binary32's normal range: the
    oracle's bits equal float32(value), value from the definition in exact
    float64 arithmetic (24 significant bits fit binary32)."""
    rng = np.random.default_rng(3)
    w = rng.integers(0, 2 ** 32, 200_000, dtype=np.uint64).astype(np.uint32)
    val = ibm_value(w)
    ok = (np.abs(val) >= np.finfo(np.float32).tiny) & (np.abs(val) <= np.finfo(np.float32).max)
    assert ok.sum() > 50_000
    got = O.ibm_to_ieee(w[ok])
    assert np.array_equal(got, val[ok].astype(np.float32).view(np.uint32))


def test_ibm_special_cases():
    z = O.ibm_to_ieee(np.array([0x00000000, 0x80000000, 0x7F000000, 0xC1000000], np.uint32))
    assert np.all(z == 0)  # zero fraction: +0 whatever the sign and exponent
    # 16^32 = 2^128 has biased exponent exactly 255, kept: the infinity bits
    big = O.ibm_to_ieee(np.array([0x61100000, 0x61100001], np.uint32))
    assert big[0] == 0x7F800000 and big[1] == 0x7F800000 | (1 << 3)
    # beyond that, or below 2^-126: exponent field 0, sign and shifted
    # fraction kept
    big = O.ibm_to_ieee(np.array([0x62100000, 0xE2100000, 0x62180000], np.uint32))
    assert big[0] == 0x00000000 and big[1] == 0x80000000 and big[2] == 0x00000000 | (0x180000 << 3) & 0x7fffff
    tiny = O.ibm_to_ieee(np.array([0x00800000], np.uint32))  # 0.5 * 16^-64 = 2^-257
    assert tiny[0] == 0


def test_ibm_encode_roundtrip():
    rng = np.random.default_rng(4)
    x = (rng.standard_normal(5000) * 10.0 ** rng.uniform(-20, 20, 5000)).astype(np.float32)
    x = (x.view(np.uint32) & np.uint32(0xFFFFFFF8)).view(np.float32)  # IBM-exact
    w = ibm_encode(x)
    assert np.array_equal(ibm_value(w), x.astype(np.float64))
    assert np.array_equal(O.ibm_to_ieee(w), x.view(np.uint32))


@needs_ref
def test_ibm_oracle_and_library_match_reference_reader(tmp_path):
    """Every kind of word, out-of-range ones included: the reference's
    fsgettr (compiled from its sources), the oracle and the library's host
    decode give the same bits."""
    rng = np.random.default_rng(5)
    ns, ntr = 257, 40
    w = rng.integers(0, 2 ** 32, ntr * ns, dtype=np.uint64).astype(np.uint32)
    w[:16] = [0, 0x80000000, 0x7FFFFFFF, 0xFFFFFFFF, 0x00000001, 0x80000001, 0x7F000001, 0x00FFFFFF,
              0x41100000, 0xC276A000, 0x20100000, 0x21100000, 0x5F100000, 0x60100000, 0x60FFFFFF, 0x1F000001]
    # records carrying the raw words as their samples
    rec = segy_records(np.zeros((ntr, ns), np.float32), {"ns": [ns] * ntr}, 5)
    rec[:, 240:] = w.reshape(ntr, ns).astype(">u4").view(np.uint8).reshape(ntr, 4 * ns)
    p = tmp_path / "ibm.sgy"
    p.write_bytes(rec.tobytes())
    n, _, ref = _ref_read(p, 1, _ref_keys(), ntr, ns)
    assert n == ntr
    orc = O.ibm_to_ieee(w).reshape(ntr, ns)
    lib_s, hdrs = segy.decode(rec, ns, segy.IBM)
    assert np.array_equal(ref.view(np.uint32), orc)
    assert np.array_equal(lib_s.view(np.uint32), orc)
    assert np.all(hdrs["ns"] == ns)


# ---- header keys, reel header, key lookup ------------------------------------------


MOVES = [("sx", 181, 4), ("sy", 185, 4), ("gx", 189, 2), ("gy", 191, 2), ("cdp", 197, 4),
         ("scalco", 201, 4), ("dt", 205, 4), ("trid", 13, 4), ("offset", 29, 2)]


@needs_ref
@pytest.mark.parametrize("moves", [[], MOVES])
def test_trace_headers_match_reference_reader(tmp_path, moves):
    """Random header bytes everywhere, ns at its position: the 240 decoded
    header bytes and the samples of the library's fsgettr path equal the
    reference's, for standard and moved keys (2-byte keys into int fields,
    4-byte keys into short ones)."""
    rng = np.random.default_rng(6)
    ns, ntr = 64, 25
    x = rng.standard_normal((ntr, ns)).astype(np.float32)
    rec = segy_records(x, {"ns": [ns] * ntr}, 5, rng=rng)
    p = tmp_path / "h.sgy"
    p.write_bytes(rec.tobytes())
    n, ref_h, ref_d = _ref_read(p, 5, _ref_keys(moves), ntr, ns)
    assert n == ntr
    k = segy.keyindex(**{m: (s, b) for m, s, b in moves})
    assert np.array_equal(np.frombuffer(bytes(k), np.int32), _ref_keys(moves))
    # the library's sgettr into a zeroed record, as the shim's calloc'd one
    tr = (C.c_ubyte * (240 + 4 * 65535))()
    for i in range(ntr):
        C.memset(tr, 0, 240)
        got_ns = L.lib().qpsdm_segy_sgettr(rec[i].ctypes.data_as(C.c_void_p), tr, 5, k)
        assert got_ns == ns
        h = np.frombuffer(tr, np.uint8, 240)
        assert np.array_equal(h, ref_h[i]), i
        assert np.array_equal(np.frombuffer(tr, np.float32, ns, 240), ref_d[i])
    s, hd = segy.decode(rec, ns, segy.IEEE, k)
    assert np.array_equal(s, ref_d)
    for f, off, dt in (("sx", 72, "<i4"), ("sy", 76, "<i4"), ("gx", 80, "<i4"), ("gy", 84, "<i4"),
                       ("cdp", 20, "<i4"), ("offset", 36, "<i4"), ("scalco", 70, "<i2"), ("dt", 116, "<u2")):
        assert np.array_equal(hd[f], ref_h[:, off:off + np.dtype(dt).itemsize].copy().view(dt).ravel()), f


@needs_ref
def test_reel_header_matches_reference(tmp_path):
    """Text bytes 0..255 (every EBCDIC code) and a random binary header."""
    rng = np.random.default_rng(7)
    raw = np.concatenate([np.arange(3200) % 256, rng.integers(0, 256, 400)]).astype(np.uint8)
    p = tmp_path / "r.sgy"
    p.write_bytes(raw.tobytes())
    ref = np.zeros(3600, np.uint8)
    assert O.ref().ref_fsgethd(str(p).encode(), ref.ctypes.data_as(C.c_void_p)) == 0
    text, bh = segy.read_reel_header(p)
    assert text == ref[:3200].tobytes()
    assert bh["jobid"] == int(ref[3200:3204].view("<i4")[0])
    assert bh["vpol"] == int(ref[3258:3260].view("<i2")[0]) and bh["format"] == int(ref[3224:3226].view("<i2")[0])
    # printable EBCDIC letters and digits map to ASCII
    assert text[0xC1] == ord("A") and text[0x81] == ord("a") and text[0xF0] == ord("0") and text[0x40] == 0x20


@needs_ref
def test_setkeyindex_lookup_matches_reference():
    """First key whose name starts with the given one; names matching no key
    change nothing."""
    for name in ("s", "cd", "tr", "g", "ns", "d", "xline", "of", "nosuchkey", "sp"):
        k = segy.keyindex()
        try:
            L.lib().qpsdm_segy_setkeyindex(k, name.encode(), 201, 4)
        except Exception:
            pass
        mine = np.frombuffer(bytes(k), np.int32)
        ref = _ref_keys([(name, 201, 4)]) if name not in ("nosuchkey", "sp") else _ref_keys()
        assert np.array_equal(mine, ref), name
    with pytest.raises(L.QpsdmError):
        segy.keyindex(nosuchkey=(1, 4))
    with pytest.raises(L.QpsdmError):
        segy.keyindex(sx=(239, 4))


@needs_ref
def test_reads_files_written_by_reference_writer(tmp_path):
    """inithd3600 + fsputhd + fsputtr of the reference (IBM and IEEE), read
    back by the library's file reader: same samples as the reference's own
    reader, format and sizes from the binary header."""
    rng = np.random.default_rng(8)
    ns, ntr = 301, 12
    x = (rng.standard_normal((ntr, ns)) * 1e3).astype(np.float32)
    hdr = np.zeros((ntr, 240), np.uint8)
    hv = hdr.view(np.int32)
    hv[:, 0] = np.arange(1, ntr + 1)
    hv[:, 18:22] = rng.integers(-10 ** 6, 10 ** 6, (ntr, 4))
    hdr[:, 114:116] = np.array([ns], np.uint16).view(np.uint8)
    hdr[:, 116:118] = np.array([2000], np.uint16).view(np.uint8)
    keys = _ref_keys()
    for fmt in (1, 5):
        p = tmp_path / f"w{fmt}.sgy"
        assert O.ref().ref_fsput_file(str(p).encode(), ns, 2000, fmt, keys.ctypes.data_as(C.c_void_p), ntr,
                                      hdr.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p)) == 0
        n, ref_h, ref_d = _ref_read(p, fmt, keys, ntr, ns, skip=3600)
        assert n == ntr
        text, bh, s, hd = segy.read_file(p)
        assert bh["format"] == fmt and bh["hns"] == ns and bh["hdt"] == 2000
        assert text[:2] == b"C "
        assert np.array_equal(s.view(np.uint32), ref_d.view(np.uint32))
        assert np.array_equal(hd["sx"], hv[:, 18]) and np.array_equal(hd["gy"], hv[:, 21])
        if fmt == 5:
            assert np.array_equal(s, x)
        else:  # the writer truncates to the hex-normalised fraction
            assert np.max(np.abs(s - x) / np.abs(x)) < 2.0 ** -20


def test_decode_rejects_mismatched_ns():
    rec = segy_records(np.zeros((3, 10), np.float32), {"ns": [10, 11, 10]}, 1)
    with pytest.raises(L.QpsdmError):
        segy.decode(rec, 10, segy.IBM)
    with pytest.raises(L.QpsdmError):
        segy.decode(rec[:1], 10, 3)
