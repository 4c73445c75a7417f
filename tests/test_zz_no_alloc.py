"""kscr_lbvh_self_query_static (no device allocation) against the entry
points it stands for, kscr_lbvh_compute_device + kscr_lbvh_query_device, on
two calls in a row: the reference's 100k-box recipe, then a smaller set.  The
pair sets must be equal (the order of pairs in the output is not defined)."""
import ctypes as C

import numpy as np
import pytest
import torch

from kamisimulator_amd import _lib
from tests.lbvh_cases import recipe_points

pytestmark = pytest.mark.gpu


def _sorted_pairs(p):
    p = np.sort(p.astype(np.int64), axis=1)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def _reference_pairs(lbvh, d_boxes, n, cap):
    out = torch.empty((cap, 2), dtype=torch.int32, device="cuda:0")
    lbvh.compute_device(d_boxes.data_ptr(), n)
    cnt = lbvh.query_device(out.data_ptr(), cap)
    torch.cuda.synchronize()
    return out[:cnt].cpu().numpy(), cnt


def _static_pairs(lib, d_boxes, n, cap):
    out = torch.empty((cap, 2), dtype=torch.int32, device="cuda:0")
    cnt = C.c_size_t(0)
    err = C.create_string_buffer(256)
    rc = lib.kscr_lbvh_self_query_static(C.c_void_p(d_boxes.data_ptr()), n, C.c_void_p(out.data_ptr()), cap,
                                         C.byref(cnt), None, err, 256)
    assert rc == 0, err.value
    torch.cuda.synchronize()
    return out[:cnt.value].cpu().numpy(), cnt.value


def test_static_self_query_matches_allocating_entry_points():
    lib = _lib.load()
    torch.cuda.set_device(0)
    lbvh = _lib.Lbvh(device=0)
    rng = np.random.default_rng(9)
    lo = rng.random((30_000, 3), np.float32)
    cases = [("recipe 100k", recipe_points(100_000)), ("random 30k", np.concatenate([lo, lo + 0.004], 1))]
    for name, boxes in cases:
        n = len(boxes)
        d_boxes = torch.from_numpy(np.ascontiguousarray(boxes, np.float32)).to("cuda:0")
        cap = 64 * n
        print(f"[no-alloc] {name}: {n} boxes", flush=True)
        ref, cref = _reference_pairs(lbvh, d_boxes, n, cap)
        got, cgot = _static_pairs(lib, d_boxes, n, cap)
        print(f"[no-alloc] {name}: pairs {cgot} (allocating entry points {cref})", flush=True)
        assert cgot == cref and cref > 0
        assert np.array_equal(_sorted_pairs(got), _sorted_pairs(ref))


def test_static_self_query_rejects_too_many_boxes():
    lib = _lib.load()
    err = C.create_string_buffer(256)
    cnt = C.c_size_t(0)
    rc = lib.kscr_lbvh_self_query_static(C.c_void_p(16), (1 << 17) + 1, C.c_void_p(16), 1, C.byref(cnt), None, err,
                                         256)
    assert rc == 1 and b"131072" in err.value
