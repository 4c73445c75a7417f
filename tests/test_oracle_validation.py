"""Host-side setup of the oracle against the reference's stated rules.

- Sim::uploadMeta's two throws (YarnBall/YBSimulator.cu:326-332), which the
  C ABI returns as KSCR_ERR_DETECTION_SCALER (5) and KSCR_ERR_SEGMENT_TOO_SHORT
  (6).
- Sim::setKStretch and Sim::setKBend (YBSimulator.cu:276-298).  kStretch is
  k times the rest length.  The rest rotor keeps its direction, with norm
  4 k / l.
"""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import default_params
from oracle.pyoracle import OracleSim


def _configured(seg_len=1e-3, **kw):
    sc = scenes.groom(5, 12, seed=3, seg_len=seg_len)
    p = default_params(**kw)
    o = OracleSim(p, sc.positions, sc.strand_offsets, precision=32)
    o.configure(1300.0)
    return o, p, sc


def _status(fn, *args):
    try:
        fn(*args)
    except RuntimeError as e:
        return int(str(e).rsplit(" ", 1)[1])
    return 0


def test_detection_scaler_must_exceed_one(oracle_lib):
    o, p, _ = _configured()
    assert _status(o.upload) == 0
    for bad in (1.0, 0.5):
        p.detection_scaler = bad
        assert _status(o.set_params, p) == 5
    p.detection_scaler = 1.0 + 2 ** -20
    assert _status(o.set_params, p) == 0


def test_segments_must_hold_the_barrier(oracle_lib):
    # 2 minSegLen must be >= 2 radius + barrierThickness; the groom's shortest
    # segment is just under 1 mm (helical curl)
    o, p, sc = _configured(radius=1e-4, barrier_thickness=8e-4)
    seg = np.linalg.norm(np.diff(sc.positions.astype(np.float64), axis=0), axis=1)
    tips = sc.strand_offsets[1:-1] - 1
    seg = np.delete(seg, tips)
    lmin = seg.min()
    assert _status(o.upload) == 0
    # a barrier just over the limit fails, just under passes
    p.barrier_thickness = float(np.float32(2 * lmin - 2 * p.radius) * np.float32(1 + 1e-5))
    assert _status(o.set_params, p) == 6
    p.barrier_thickness = float(np.float32(2 * lmin - 2 * p.radius) * np.float32(1 - 1e-5))
    assert _status(o.set_params, p) == 0
    # 0.4 mm straight segments with the default radius and barrier (2r + b = 1 mm)
    sc2 = scenes.cantilever(8, 4e-4)
    o2 = OracleSim(default_params(), sc2.positions, sc2.strand_offsets, precision=32)
    o2.configure(1300.0)
    assert _status(o2.upload) == 6


def test_set_kstretch_and_kbend_closed_forms(oracle_lib):
    o, p, _ = _configured(radius=2e-6, barrier_thickness=7.6e-5)
    r0 = o.get_rest()
    k = np.float32(0.37)
    o.set_kstretch(float(k))
    r1 = o.get_rest()
    assert np.array_equal(r1["k_stretch"], (k * r0["l_rest"]).astype(np.float32))
    kb = np.float32(2.5e-3)
    o.set_kbend(float(kb))
    r2 = o.get_rest()
    live = r0["l_rest"] > 0
    l = r0["l_rest"][live].astype(np.float64)
    n0 = np.linalg.norm(r0["q_rest"][live].astype(np.float64), axis=1)
    n2 = np.linalg.norm(r2["q_rest"][live].astype(np.float64), axis=1)
    dirs0 = r0["q_rest"][live] / n0[:, None]
    dirs2 = r2["q_rest"][live] / n2[:, None]
    np.testing.assert_allclose(n2, 4 * float(kb) / l, rtol=3e-7)
    np.testing.assert_allclose(dirs2, dirs0, atol=3e-7)
    # setKStretch leaves the rotors alone, setKBend the stiffness
    assert np.array_equal(r2["k_stretch"], r1["k_stretch"])
