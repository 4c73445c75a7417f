"""Host-side solver (kamisimulator_amd/host, C++ over the C ABI) on the CPU:
JSON settings with the reference solver's keys and defaults, DivideHair and
the Catmull-Rom resample of the curve preparation, and the skipping of short
strands (rootVertices / strandRefs).  Pins: the reference's documented
defaults (StableCosseratRodsSolver.h:87-108), closed forms (linear
subdivision; a straight line resamples to exact equal spacing), and
invariants (resampled points lie on the spline, segment lengths agree)."""
import json

import numpy as np
import pytest

from kamisimulator_amd import hairio


def _parse(tmp_path, obj):
    p = tmp_path / "s.json"
    p.write_text(json.dumps(obj) if not isinstance(obj, str) else obj)
    return json.loads(hairio.cli("parse", p).stdout)


def test_defaults_are_the_reference_solver_defaults(tmp_path):
    d = _parse(tmp_path, {})
    assert d["EnableResample"] == 1 and d["EnableSDF"] == 1 and d["EnableHairHairColl"] == 0
    assert d["UseYoungsModule"] == 1 and d["NumItr"] == 8 and d["SagFreeItr"] == 0 and d["SubdivCoeff"] == 3
    for k, v in dict(StretchK=3e3, BendK=6e-4, YoungsModule=5, StiffnessDeviationCoeff=1, CurveRadius=4e-5,
                     Density=1300, ResampleLen=1e-3, VelocityDecay=0.05, ExternalForceMultiplier=1, WindPeak=40,
                     WindYFreq=1, WindZFreq=0.7, WindTimeFreq=0.4, WindSharpness=3).items():
        assert d[k] == pytest.approx(v, rel=1e-7), k
    assert d["HasGravity"] == 0 and d["MaxSubstep"] == pytest.approx(4e-4)
    # the SDF solver's strain limiting (SDFBodyCollisionSolver.h:26-27)
    assert d["UseStrainLimiting"] == 0 and d["StrainError"] == pytest.approx(0.1, rel=1e-7)


def test_every_reference_key_is_read(tmp_path):
    # the keys and value kinds of a KamiSimulator SCR setting file
    s = {"EnableResample": True, "EnableSDF": False, "EnableHairHairColl": True, "UseYoungsModule": False,
         "NumItr": 32, "SagFreeItr": 1, "SubdivCoeff": 2, "StretchK": 2.5e3, "BendK": 7e-4, "YoungsModule": 4,
         "StiffnessDeviationCoeff": 0.5, "CurveRadius": 5e-5, "Density": 1.3e6, "ResampleLen": 2e-3,
         "VelocityDecay": 0.1, "ExternalForceMultiplier": 2, "WindPeak": 10, "WindYFreq": 2, "WindZFreq": 0.5,
         "WindTimeFreq": 0.3, "WindSharpness": 4.0, "YBJsonPath": "none.json",
         "Gravity": [0.0, -9.81, 1.0], "SDFThreshold": 0.002, "StrainError": 0.05, "UseStrainLimiting": True,
         "MaxSubstep": 2e-4, "Schedule": "colored",
         "BlockSize": 256, "DeviceResidentHair": True, "Unrelated": {"x": [1, "a"]}}
    d = _parse(tmp_path, s)
    for k, v in s.items():
        if k in ("Unrelated", "Schedule"):
            continue
        if isinstance(v, bool):
            assert d[k] == int(v), k
        elif isinstance(v, str):
            assert d[k] == v, k
        elif isinstance(v, list):
            assert d[k] == pytest.approx(v, rel=1e-7), k
        else:
            assert d[k] == pytest.approx(v, rel=1e-7), k
    assert d["Schedule"] == 1 and d["HasGravity"] == 1


def test_bad_settings_are_reported(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"NumItr": 8,, }')
    r = hairio.cli("parse", p, check=False)
    assert r.returncode == 1 and "JSON" in r.stderr
    p.write_text('{"Schedule": "gauss"}')
    r = hairio.cli("parse", p, check=False)
    assert r.returncode == 1 and "Schedule" in r.stderr


def _prep(tmp_path, settings, counts, cur):
    (tmp_path / "s.json").write_text(json.dumps(settings))
    hairio.write_hair(tmp_path / "in.bin", counts, cur)
    r = hairio.cli("prep", tmp_path / "s.json", tmp_path / "in.bin", tmp_path / "out.bin")
    return json.loads(r.stdout), hairio.read_hair(tmp_path / "out.bin")


def _padded(strands, mv):
    cur = np.zeros((len(strands), mv, 3), np.float32)
    for s, pts in enumerate(strands):
        cur[s, :len(pts)] = pts
    return np.array([len(p) for p in strands], np.uint32), cur


def test_divide_hair_and_short_strand_skipping(tmp_path):
    rng = np.random.default_rng(0)
    strands = [rng.normal(size=(n, 3)).astype(np.float32) for n in (6, 3, 1, 5)]
    counts, cur = _padded(strands, 6)
    refs, (c2, cur2, _) = _prep(tmp_path, {"SubdivCoeff": 3, "EnableResample": False}, counts, cur)
    assert cur2.shape == (4, 16, 3)
    assert list(c2) == [16, 7, 1, 13]
    # strands with fewer than 4 vertices after subdivision are skipped
    assert refs["strandRefs"] == [0, 1, 3] and refs["rootVertices"] == [0, 16, 23] and refs["numVerts"] == 36
    for s, pts in enumerate(strands):
        n = len(pts)
        for v in range(n - 1):
            for k in range(3):
                want = pts[v] + (k / 3.0) * (pts[v + 1].astype(np.float64) - pts[v])
                np.testing.assert_allclose(cur2[s, 3 * v + k], want, rtol=0, atol=4e-7 * (1 + np.abs(pts).max()))
        np.testing.assert_array_equal(cur2[s, 3 * (n - 1)], pts[n - 1])


def test_resample_straight_line_is_exactly_uniform(tmp_path):
    # vertices 1 mm apart on a line; resampling at 0.5 mm between vertex 1 and
    # vertex n-2 gives exactly equally spaced points on the same line
    n = 12
    pts = np.zeros((n, 3), np.float32)
    pts[:, 0] = np.arange(n) * 1e-3
    counts, cur = _padded([pts], 64)
    refs, (c2, cur2, _) = _prep(tmp_path, {"SubdivCoeff": 1, "EnableResample": True, "ResampleLen": 5e-4}, counts, cur)
    m = int(c2[0])
    assert m == 2 * (n - 3) + 1 and refs["numVerts"] == m
    x = cur2[0, :m].astype(np.float64)
    np.testing.assert_allclose(x[:, 1:], 0, atol=1e-9)
    np.testing.assert_allclose(x[:, 0], 1e-3 + 5e-4 * np.arange(m), rtol=0, atol=2e-9)


def _cmr(pts, t):
    """Uniform Catmull-Rom spline through pts at parameter t (segment si = floor(t))."""
    si = int(np.clip(np.floor(t), 1, len(pts) - 3))
    p0, p1, p2, p3 = (pts[si + k].astype(np.float64) for k in (-1, 0, 1, 2))
    u = t - si
    # polynomial form of the uniform Catmull-Rom segment
    return 0.5 * ((2 * p1) + (-p0 + p2) * u + (2 * p0 - 5 * p1 + 4 * p2 - p3) * u * u + (-p0 + 3 * p1 - 3 * p2 + p3) * u ** 3)


def test_resample_curly_strand_invariants(tmp_path):
    # a helix with 2 mm vertex spacing resampled at 0.7 mm
    j = np.arange(40)
    ang = j * 2 * np.pi / 11
    pts = np.stack([3e-3 * np.cos(ang), j * 1.5e-3, 3e-3 * np.sin(ang)], 1).astype(np.float32)
    counts, cur = _padded([pts], 400)
    tar = 7e-4
    refs, (c2, cur2, _) = _prep(tmp_path, {"SubdivCoeff": 1, "EnableResample": True, "ResampleLen": tar}, counts, cur)
    m = int(c2[0])
    x = cur2[0, :m].astype(np.float64)
    # ends: the spline at t = 1 and t = n - 2 (the original vertices 1 and
    # n - 2); the splice of the two marching directions blends the 8 points
    # after the split point, which may move an end point slightly along the
    # curve when the split falls near it
    np.testing.assert_allclose(x[0], pts[1], atol=0.01 * tar)
    np.testing.assert_allclose(x[-1], pts[-2], atol=0.01 * tar)
    # every point lies on the spline: nearest point of a coarse parameter
    # sweep, refined by a fine sweep around it
    ts = np.linspace(1, len(pts) - 2, 20000)
    curve = np.array([_cmr(pts, t) for t in ts])
    k = np.argmin(np.linalg.norm(curve[None, :, :] - x[:, None, :], axis=2), axis=1)
    dt = ts[1] - ts[0]
    d = []
    for p, t0 in zip(x, ts[k]):
        fine = np.array([_cmr(pts, t) for t in np.linspace(max(1, t0 - dt), min(len(pts) - 2, t0 + dt), 801)])
        d.append(np.linalg.norm(fine - p, axis=1).min())
    assert max(d) < 2e-8
    # segment lengths agree with each other to a small fraction of the target
    seg = np.linalg.norm(np.diff(x, axis=0), axis=1)
    assert abs(seg.mean() - tar) < 0.05 * tar
    assert (np.abs(seg - seg.mean()) < 0.02 * tar).all()


def test_resample_is_capped_at_hair_size(tmp_path):
    n = 30
    pts = np.zeros((n, 3), np.float32)
    pts[:, 1] = np.arange(n) * 1e-3
    counts, cur = _padded([pts], 32)
    refs, (c2, cur2, _) = _prep(tmp_path, {"SubdivCoeff": 1, "EnableResample": True, "ResampleLen": 2.5e-4}, counts, cur)
    assert int(c2[0]) == 32 and refs["numVerts"] == 32
    x = cur2[0].astype(np.float64)
    np.testing.assert_allclose(x[0], pts[1], atol=1e-9)
    np.testing.assert_allclose(x[-1], pts[-2], atol=1e-9)
    # the 30 interior points are samples of the uniform 0.25 mm resampling
    # taken at floor(stride * v), stride = (L - 2) / 30
    L = 4 * (n - 3) + 1
    stride = np.float32(L - 2) / np.float32(30)
    idx = [int(np.float32(stride) * np.float32(v)) for v in range(1, 31)]
    np.testing.assert_allclose(x[1:31, 1], 1e-3 + 2.5e-4 * np.array(idx), atol=2e-9)



def _splice_curve(n=160):
    """Control points along x, spaced widest in the middle, on a gentle wave."""
    k = np.arange(n - 1)
    spacing = 1e-3 * (1 + 0.3 * np.exp(-((k - n / 2) / (n / 8)) ** 2))
    xs = np.concatenate([[0.0], np.cumsum(spacing)])
    return np.stack([xs, 2e-3 * np.sin(xs / 0.013), np.zeros(n)], 1).astype(np.float32)


def _t_of_x(P, x, lo, hi):
    """The parameter where the spline through P (x increasing) reaches abscissa x."""
    a, b = lo, hi
    for _ in range(80):
        m = 0.5 * (a + b)
        if _cmr(P, m)[0] < x:
            a = m
        else:
            b = m
    return 0.5 * (a + b)


def _march(P, t0, t_stop, L, direction):
    """Parameters at successive chord length L from t0 towards t_stop (the
    first crossing, by bisection on the chord)."""
    ts = [t0]
    while True:
        c0 = _cmr(P, ts[-1])
        a = ts[-1]
        b = a + direction * 0.25
        if (b - t_stop) * direction > 0:
            break
        while np.linalg.norm(_cmr(P, b) - c0) < L:
            a, b = b, b + direction * 0.25
            if (b - t_stop) * direction > 0:
                return np.array(ts)
        for _ in range(80):
            m = 0.5 * (a + b)
            if np.linalg.norm(_cmr(P, m) - c0) < L:
                a = m
            else:
                b = m
        ts.append(0.5 * (a + b))
    return np.array(ts)


def test_resample_splice_of_the_two_marching_directions(tmp_path):
    """One resampling pass marches from both ends and splices the two lists
    (resample.h:82-134).  On a long curve whose control points are spaced
    widest in the middle, the spline is fastest there, so the two lists'
    parameters come closest in the middle, away from both ends.  Checked on
    the pass's parameter list (`cmrcoords`), from the definition of the
    method rather than its formulas:
    - away from the splice, consecutive points are exactly one step length
      apart (chord), marching from each end;
    - the splice sits at the pair of forward and backward parameters that
      come closest (every pair tried; parameters found by bisection);
    - around it, each point's parameter is corrected by a linear ramp over
      the 9 segments on each side, from 0 to the midpoint of the two
      parameters at the splice: the gap is spread evenly.
    (The kept pass of the full resampler marches at the first pass's average
    chord, which leaves almost no gap to splice; the pass itself is where a
    splice mistake shows.)"""
    P = _splice_curve()
    n = len(P)
    P.tofile(tmp_path / "pts.f32")
    P64 = P.astype(np.float64)
    for tar in (7e-4, 6.1e-4, 9.3e-4):
        ts = np.array(json.loads(hairio.cli("cmrcoords", tmp_path / "pts.f32", 1.0, n - 2.0, repr(tar)).stdout))
        m = len(ts)
        pos = np.array([_cmr(P64, t) for t in ts])
        chord = np.linalg.norm(np.diff(pos, axis=0), axis=1)
        off = np.flatnonzero(np.abs(chord - tar) > 1e-12)
        # exactly the 18 segments around the splice leave the lattice
        assert len(off) == 18 and off[-1] - off[0] == 17, (tar, off)
        sx = int(off[0]) + 9  # the splice point
        sy = m - 1 - sx       # its index in the backward list
        assert 20 < sx < m - 20
        tf = _march(P64, 1.0, n - 2.0, tar, +1)
        tb = _march(P64, n - 2.0, 1.0, tar, -1)
        # the splice is the closest pair of forward and backward parameters
        gaps = np.abs(tf[8:len(tf) - 8, None] - tb[None, :])
        i0, j0 = np.unravel_index(np.argmin(gaps), gaps.shape)
        assert (int(i0) + 8, int(j0)) == (sx, sy), tar
        # away from the splice the points are the marched ones
        np.testing.assert_allclose(ts[:sx - 8], tf[:sx - 8], rtol=0, atol=1e-9)
        np.testing.assert_allclose(ts[sx + 9:], tb[sy - 9::-1], rtol=0, atol=1e-9)
        # parameter corrections around the splice: a linear ramp on each side
        mid = 0.5 * (tf[sx] + tb[sy])
        np.testing.assert_allclose(ts[sx], mid, rtol=0, atol=1e-9)
        np.testing.assert_allclose(ts[sx - 9:sx] - tf[sx - 9:sx], np.arange(9) / 9.0 * (mid - tf[sx]), atol=1e-9)
        np.testing.assert_allclose(ts[sx + 1:sx + 10] - tb[sy - 1:sy - 10:-1],
                                   (8 - np.arange(9)) / 9.0 * (mid - tb[sy]), atol=1e-9)
        assert abs(mid - tf[sx]) > 1e-3  # the splice had a gap to spread
