"""Writes the golden fixtures under tests/golden/ from the reference's own
example data files and from its compiled host sources (oracle/_ref).

Run in the build container only (it reads /root/reference and needs
oracle/_ref/libpsdmref.so, built by `make -C oracle`).  The fixtures are
data: byte-exact copies of small example files the reference holds, a
truncated copy of its golden output SU file, and input/output vectors of the
reference's pfarc, fetch_and_taper, parseint, parsetuple, readparfile,
getscale and time_offset_mute on inputs chosen here.
"""
import ctypes as C
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/examples"
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def main():
    R = O.ref()
    # 1. example data files (byte copies)
    shutil.copy(f"{REF}/traveltime/Lxxx.tidx", f"{HERE}/Lxxx.tidx")
    shutil.copy(f"{REF}/traveltime/Lxxx.qidx", f"{HERE}/Lxxx.qidx")
    shutil.copy(f"{REF}/0/qmig3dd/par/par.dat", f"{HERE}/par.dat")
    raw = open(f"{REF}/0/qmig3dd/result/node2_0.off0.L350.su", "rb").read()
    rec = 240 + 4 * 376
    assert len(raw) == 201 * rec
    with open(f"{HERE}/L350_first3.su", "wb") as f:
        f.write(raw[:3 * rec])
    # decoded headers (reference sgettr) of traces 0, 1, 100, 200
    hdrs = {}
    names = {"tracl": (0, "i"), "fldr": (8, "i"), "cdp": (20, "i"), "trid": (28, "h"), "nvs": (30, "h"),
             "offset": (36, "i"), "scalco": (70, "h"), "sx": (72, "i"), "sy": (76, "i"), "gx": (80, "i"),
             "gy": (84, "i"), "counit": (88, "h"), "ns": (114, "H"), "dt": (116, "H")}
    for i in (0, 1, 100, 200):
        buf = (C.c_ubyte * rec).from_buffer_copy(raw[i * rec:(i + 1) * rec])
        h = (C.c_ubyte * 240)()
        d = (C.c_float * 376)()
        ns = R.ref_sgettr(buf, h, d)
        hb = bytes(h)
        dec = {k: int(np.frombuffer(hb[o:o + (4 if t == "i" else 2)],
                                    {"i": "<i4", "h": "<i2", "H": "<u2"}[t])[0])
               for k, (o, t) in names.items()}
        dec["ns_returned"] = ns
        dec["max_abs_sample"] = float(np.abs(np.frombuffer(bytes(d), np.float32)).max())
        hdrs[str(i)] = dec
    # 2. reference pfarc on seeded inputs (n = 720, 728, 1232)
    pf = {}
    for n in (720, 728, 1232):
        x = np.random.default_rng(n).standard_normal(n).astype(np.float32)
        z = (C.c_float * (n + 2))()
        xin = (C.c_float * n)(*x)
        R.ref_pfarc(-1, n, xin, z)
        pf[str(n)] = np.frombuffer(bytes(z), np.float32).tolist()
    # 3. fetch_and_taper on a few windows
    x = np.random.default_rng(1).standard_normal(500).astype(np.float32)
    fat = {}
    for it1, it2 in ((0, 499), (-30, 150), (100, 899), (450, 520), (0, 300)):
        o = (C.c_float * (it2 - it1 + 1))()
        R.ref_fetch_and_taper((C.c_float * 500)(*x), 500, it1, it2, o)
        fat[f"{it1},{it2}"] = np.frombuffer(bytes(o), np.float32).tolist()
    # 4. parsers
    pi = {}
    for s in ("350", "1-3,5,10:2:14", "299-305", "1,2,3", "7,9-11,40:10:70", "100:25:200"):
        n = C.c_int()
        p = R.ref_parseint(s.encode(), C.byref(n))
        pi[s] = [p[k] for k in range(n.value)]
    pt = {}
    R.ref_parsetuple.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    for s in ("(1600,500),(3200,800),(4800,2000),(6400,5000)", "(1.5 2.5 3e2)", "(0,0);(1,1)"):
        a, b = C.c_int(), C.c_int()
        p = R.ref_parsetuple(s.encode(), C.byref(a), C.byref(b))
        pt[s] = {"ntuple": a.value, "nelem": b.value,
                 "values": [p[k] for k in range(a.value * b.value)]}
    # 5. readparfile on the example par.dat
    rp = {}
    pd = f"{HERE}/par.dat".encode()
    for key, kind in (("ntl=%d", "i"), ("dt=%f", "f"), ("line1=%d", "i"), ("f4=%f", "f"),
                      ("threshold=%f", "f"), ("mute=%s", "s"), ("missing=%d", "i"), ("taperzone=%f", "f")):
        if kind == "i":
            v = C.c_int(-99)
            rc = R.ref_readparfile_i(pd, key.encode(), C.byref(v))
            rp[key] = [rc, v.value]
        elif kind == "f":
            v = C.c_float(-99)
            rc = R.ref_readparfile_f(pd, key.encode(), C.byref(v))
            rp[key] = [rc, v.value]
        else:
            v = C.create_string_buffer(1024)
            rc = R.ref_readparfile_s(pd, key.encode(), v)
            rp[key] = [rc, v.value.decode()]
    a, b = C.c_int(-1), C.c_int(-1)
    rp["line1=%d,line2=%d"] = [R.ref_readparfile_ii(pd, b"line1=%d,line2=%d", C.byref(a), C.byref(b)),
                               a.value, b.value]
    # 6. getscale, time_offset_mute
    gs = {str(s): R.ref_getscale(s) for s in (0, 1, 10, -10, -100, 100)}
    mute = np.array([1600, 500, 3200, 800, 4800, 2000, 6400, 5000], np.float32)
    tm = {str(o): R.ref_time_offset_mute((C.c_float * 8)(*mute), 4, float(o))
          for o in (100.0, 500.0, 650.0, 1999.0, 2000.0, 4000.0, 7000.0)}
    json.dump({"su_headers": hdrs, "pfarc": pf, "fetch_and_taper_input_seed": 1,
               "fetch_and_taper": fat, "parseint": pi, "parsetuple": pt, "readparfile": rp,
               "getscale": gs, "time_offset_mute": tm}, open(f"{HERE}/reference_vectors.json", "w"))
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
