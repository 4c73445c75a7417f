"""The C ABI library (libkscr.so) loads without a GPU, exports every entry
point include/kscr.h declares, and rejects bad input on the host before any
device work.  No compute call is made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from kamisimulator_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(REPO, "include", "kscr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kscr_[a-z_0-9]+)\s*\(", src)))


def test_header_and_binding_agree(kscr_lib):
    names = _declared()
    assert len(names) >= 25
    assert sorted(_lib.SIGNATURES) == names


def test_every_symbol_exported(kscr_lib):
    for name in _declared():
        assert hasattr(kscr_lib, name), name


def test_params_default_matches_sim_ctor(kscr_lib):
    p = _lib.KscrParams()
    assert kscr_lib.kscr_params_default(C.byref(p)) == 0
    # Sim::Sim, YB/YBSimulator.cu:81-112
    assert list(p.gravity) == [0.0, pytest.approx(-9.8), 0.0]
    assert p.num_itr == 8
    assert p.max_h == pytest.approx(4e-4)
    assert p.velocity_decay == pytest.approx(0.2)
    assert p.damping == pytest.approx(1e-6)
    assert p.radius == pytest.approx(1e-4)
    assert p.barrier_thickness == pytest.approx(8e-4)
    assert p.k_collision == pytest.approx(1e-5)
    assert p.detection_scaler == pytest.approx(1.2)
    assert p.friction_coeff == pytest.approx(0.1)
    assert p.k_friction == pytest.approx(5.0)
    assert p.detection_period == 1
    assert p.use_velocity_radius == 1 and p.use_step_size_limit == 0
    assert p.bvh_rebuild_period == pytest.approx(0.1)
    assert p.wind_sharpness == pytest.approx(1.0)
    assert C.sizeof(p) == 27 * 4


def test_create_validates_on_host(kscr_lib):
    p = _lib.default_params()
    pos = np.zeros((2, 3), np.float32)
    h = C.c_void_p()
    off = np.array([0, 2], np.int32)
    rc = kscr_lib.kscr_create(C.byref(p), 2, _lib.fptr(pos), 1, _lib.iptr(off), 0, C.byref(h))
    assert rc == 2  # KSCR_ERR_TOO_FEW_VERTICES (Sim::Sim, YBSimulator.cu:82)
    assert b"Too little" in kscr_lib.kscr_create_error()
    pos = np.zeros((5, 3), np.float32)
    off = np.array([0, 3, 4], np.int32)  # does not end at num_verts
    rc = kscr_lib.kscr_create(C.byref(p), 5, _lib.fptr(pos), 2, _lib.iptr(off), 0, C.byref(h))
    assert rc == 1
    off = np.array([0, 3, 3, 5], np.int32)  # empty strand
    rc = kscr_lib.kscr_create(C.byref(p), 5, _lib.fptr(pos), 3, _lib.iptr(off), 0, C.byref(h))
    assert rc == 1


def test_null_handle_is_rejected(kscr_lib):
    assert kscr_lib.kscr_advance(None, C.c_float(0.01)) == 1
    assert kscr_lib.kscr_configure(None, C.c_float(1.0)) == 1
    assert kscr_lib.kscr_num_verts(None) == -1
