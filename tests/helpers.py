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
