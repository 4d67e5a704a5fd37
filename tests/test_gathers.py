"""Dip-angle gathers on the CPU: the bin count and boundaries against the
reference's formula and the definition of a dip bin, the oracle's FP32
gathers against its FP32 image (every update binned once) and against bins
computed from the geometry in FP64, and the axis choice.  No GPU."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O
from q_psdm_amd import lib as L
from q_psdm_amd import synth
from tests.test_oracle import _tiny_case


def _bins(dipx, dipy, ddipx, ddipy, aspect):
    g = L.GatherParams(dipx, dipy, ddipx, ddipy, aspect)
    nbx, nby = C.c_int(), C.c_int()
    rc = L.lib().qpsdm_gather_bins(C.byref(g), C.byref(nbx), C.byref(nby))
    return rc, nbx.value, nby.value


@pytest.mark.parametrize("dip,ddip", [(50.0, 1.0), (50.0, 3.0), (30.0, 0.7), (0.0, 1.0), (10.0, 4.0)])
def test_gather_bins_follow_reference_formula(dip, ddip):
    ndip = int(float(np.float32(dip) / np.float32(ddip)) + 0.5)   # src/q-psdm.cu:828-829
    rc, nbx, nby = _bins(dip, dip, ddip, ddip, 1)
    assert rc == 0 and nbx == nby == 2 * ndip + 1
    assert _bins(dip, dip, ddip, ddip, 2)[1:] == (2 * ndip + 1, 0)
    assert _bins(dip, dip, ddip, ddip, 3)[1:] == (0, 2 * ndip + 1)
    assert _bins(dip, dip, ddip, ddip, 4)[1:] == (0, 0)


def test_gather_bins_reject_bad_options():
    assert _bins(89.0, 50.0, 1.0, 1.0, 1)[0] < 0      # 179 + 101 bins: over 250 on both axes
    assert _bins(50.0, 50.0, 0.0, 1.0, 1)[0] < 0      # zero width
    assert _bins(50.0, 50.0, 1.0, 1.0, 5)[0] < 0      # aspect
    assert _bins(89.9, 1.0, 0.5, 1.0, 2)[0] < 0        # > 255 bins on one axis
    assert _bins(80.0, 1.0, 1.0, 1.0, 2)[0] == 0       # 161 bins, x only


def test_gather_edges_are_half_bin_tangents():
    for dip, ddip in ((50.0, 1.0), (30.0, 2.5)):
        e = O.gather_edges(dip, ddip)
        ndip = (e.size - 2) // 2
        want = np.tan(np.radians((np.arange(e.size) - ndip - 0.5) * np.float32(ddip)))
        assert np.array_equal(e, want.astype(np.float32))
        assert np.all(np.diff(e) > 0) and np.array_equal(e, -e[::-1])


def _case():
    cfg = synth.tiny()
    tab, ip = _tiny_case(cfg)
    geom, _ = synth.geometry(synth.shots_tiny(2, 5))
    bank = np.random.default_rng(12).standard_normal((geom.shape[0], cfg.K, cfg.nt)).astype(np.float32)
    cols = np.array([[3, 4], [12, 10], [23, 19], [0, 0], [17, 2]], np.int32)
    tanc = max(abs(v) for v in (ip.tanx1, ip.tanx2, ip.tany1, ip.tany2))
    tanc = tanc + ip.taperzone * (1 + tanc * tanc)
    dmax = math.degrees(math.atan(tanc))
    return cfg, tab, ip, geom, bank, cols, dmax


def test_gathers_sum_to_the_image_columns():
    """With bins spanning the widened aperture every update is binned once
    per axis: summed over bins, each axis's gather is the FP32 image's
    column (the same terms, summed per bin first)."""
    cfg, tab, ip, geom, bank, cols, dmax = _case()
    e = O.gather_edges(math.ceil(dmax) + 2, 1.0)
    gx, gy = O.gathers(bank, geom, tab, ip, cols, e, e)
    img, n = O.image(bank, geom, tab, ip, fp32=True)
    assert n > 0
    for k, (ix, iy) in enumerate(cols):
        ref = img[iy, ix].astype(np.float64)
        for g in (gx, gy):
            s = g[k].astype(np.float64).sum(axis=0)
            assert np.allclose(s, ref, rtol=2e-5, atol=2e-5 * np.abs(ref).max() + 1e-30), (k, ix, iy)
    assert np.abs(gx).sum() > 0


def test_gather_bins_from_geometry():
    """One trace: at every depth of a column the update sits in exactly one
    x bin and one y bin, the bins of the bisector's dip angles computed
    here from the geometry in FP64 (points within 1e-4 degrees of a bin
    boundary are not judged)."""
    cfg, tab, ip, geom, bank, cols, dmax = _case()
    ddip = 1.0
    ndip = math.ceil(dmax) + 2
    e = O.gather_edges(ndip, ddip)
    g1 = geom[3:4]
    gx, gy = O.gathers(bank[3:4], g1, tab, ip, cols, e, e)
    sx, sy, gxx, gyy = g1[0].astype(np.float64)
    judged = 0
    for k, (ix, iy) in enumerate(cols):
        x, y = ix * cfg.dx, iy * cfg.dy
        for iz in range(1, cfg.nz):
            z = iz * cfg.dz
            ds = math.sqrt((sx - x) ** 2 + (sy - y) ** 2 + z * z)
            dg = math.sqrt((gxx - x) ** 2 + (gyy - y) ** 2 + z * z)
            X = (sx - x) / ds + (gxx - x) / dg
            Y = (sy - y) / ds + (gyy - y) / dg
            Z = z / ds + z / dg
            nzx = np.nonzero(gx[k, :, iz])[0]
            nzy = np.nonzero(gy[k, :, iz])[0]
            if nzx.size == 0:
                continue   # outside the aperture or the trace's time range
            ax, ay = math.degrees(math.atan(X / Z)), math.degrees(math.atan(Y / Z))
            if min(abs(ax / ddip - round(ax / ddip) - 0.5), abs(ax / ddip - round(ax / ddip) + 0.5)) < 1e-4 or \
               min(abs(ay / ddip - round(ay / ddip) - 0.5), abs(ay / ddip - round(ay / ddip) + 0.5)) < 1e-4:
                continue
            assert nzx.size == 1 and nzy.size == 1
            assert nzx[0] - ndip == round(ax / ddip) and nzy[0] - ndip == round(ay / ddip), (ix, iy, iz)
            assert gx[k, nzx[0], iz] == gy[k, nzy[0], iz]
            judged += 1
    assert judged > 50


def test_gather_axes_are_independent():
    cfg, tab, ip, geom, bank, cols, dmax = _case()
    e = O.gather_edges(40.0, 2.0)
    gx, gy = O.gathers(bank, geom, tab, ip, cols, e, e)
    gx1, none = O.gathers(bank, geom, tab, ip, cols, e, None)
    none2, gy1 = O.gathers(bank, geom, tab, ip, cols, None, e)
    assert none is None and none2 is None
    assert np.array_equal(gx, gx1) and np.array_equal(gy, gy1)
