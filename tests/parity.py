"""Parity measures shared by the GPU tests: the device state against the FP32
restatement (oracle) on identical inputs.

Both sides evaluate the same FP32 operations in the same order, so the bar is
zero differing 32-bit words.  Two relative errors are printed beside it (and
are 0 when no word differs):

- scale-relative: max |x - x_ref| / max |x_ref|;
- displacement-relative: max |x - x_ref| / max |x_ref - x_0|, where x_0 is
  the state the run started from.  This is the stricter bar: a step that
  never moved fails it.

Quaternions are compared up to sign: the step is invariant under q -> -q.
"""
import numpy as np

TOL = 1e-5


def words(a, b):
    """Number of 32-bit words that differ between two float32 arrays."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape
    return int(np.count_nonzero(a.view(np.uint32) != b.view(np.uint32)))


def scale_rel(a, b):
    b64 = np.asarray(b, np.float64)
    return float(np.abs(np.asarray(a, np.float64) - b64).max() / max(np.abs(b64).max(), 1e-30))


def disp_rel(a, b, start):
    b64 = np.asarray(b, np.float64)
    d = np.abs(b64 - np.asarray(start, np.float64)).max()
    return float(np.abs(np.asarray(a, np.float64) - b64).max() / max(d, 1e-30))


def q_disp_rel(a, b, start):
    """Quaternion error up to sign, against the displacement of the rotors."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = np.asarray(start, np.float64)
    sign = np.where(np.sum(a * b, axis=1, keepdims=True) < 0, -1.0, 1.0)
    err = np.abs(a * sign - b).max()
    sign0 = np.where(np.sum(b * s, axis=1, keepdims=True) < 0, -1.0, 1.0)
    d = np.abs(b - s * sign0).max()
    return float(err / max(d, 1e-30))


def check(label, got, ref, start, tol=TOL):
    """got / ref / start: dicts with any of "x", "v", "q" (float32 arrays of the
    device, the oracle, and the state the run started from; start may lack
    "v").  Asserts zero differing words and prints the relative errors;
    returns them."""
    out = {}
    parts = []
    for k in ("x", "v", "q"):
        if k not in got:
            continue
        w = words(got[k], ref[k])
        r = scale_rel(got[k], ref[k])
        out[k] = {"words": w, "scale_rel": r}
        msg = f"{k}: {w} differing words, rel {r:.3e}"
        if k in start:
            d = q_disp_rel(got[k], ref[k], start[k]) if k == "q" else disp_rel(got[k], ref[k], start[k])
            out[k]["disp_rel"] = d
            msg += f", disp-rel {d:.3e}"
        parts.append(msg)
    print(f"[parity] {label}: " + "; ".join(parts), flush=True)
    for k, r in out.items():
        assert r["words"] == 0, f"{label}: {k} differs in {r['words']} words"
        assert r["scale_rel"] <= tol
        if "disp_rel" in r:
            assert r["disp_rel"] <= tol
    return out
