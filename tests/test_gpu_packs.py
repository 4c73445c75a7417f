"""Automatic workgroup (pack) size and its independence from the results:
a pack holds whole strands and the step's arithmetic per vertex does not
depend on which pack or lane holds the vertex, so every pack size gives the
same words."""
import numpy as np
import pytest

from kamisimulator_amd import scenes
from kamisimulator_amd._lib import Sim, default_params

pytestmark = pytest.mark.gpu
DT = np.float32(1 / 600)


def _sim(sc, contacts=False, **kw):
    p = (scenes.contact_params if contacts else scenes.make_params)(default_params, sc.material, **kw)
    return scenes.setup_sim(Sim(p, sc.positions, sc.strand_offsets), sc)


def test_automatic_pack_size():
    # 20,000 packs of 512 vertices: the default
    assert _sim(scenes.groom(100_000, 100, seed=1)).block_size == 512
    # 2000 x 64 gives 250 packs of 512 vertices, fewer than 1024 workgroups:
    # halved down to 64-vertex packs (2000 workgroups)
    assert _sim(scenes.groom(2000, 64, seed=4)).block_size == 64
    # never below the longest strand: 300 strands of up to 256 vertices
    counts = scenes.lognormal_counts(300, seed=8)
    assert _sim(scenes.groom(300, counts, seed=8)).block_size == 256
    # the contact path starts from 256-vertex packs
    sc = scenes.config_large(20_000, seed=2, sdf_dim=64)
    assert _sim(sc, contacts=True).block_size == 256


@pytest.mark.parametrize("contacts", [False, True])
def test_every_pack_size_gives_the_same_words(contacts):
    sc = scenes.config_large(300, seed=3, sdf_dim=64)
    sims = [_sim(sc, contacts=contacts, block_size=b) for b in (0, 256, 512, 1024)]
    print(f"[gpu-packs] contacts={contacts}: blocks {[s.block_size for s in sims]}", flush=True)
    for _ in range(3):
        for s in sims:
            s.solve_next_step(DT)
    ref_p, ref_v = sims[0].get_pos_vel()
    ref_q = sims[0].get_q()
    for s in sims[1:]:
        p, v = s.get_pos_vel()
        assert np.array_equal(p.view(np.uint32), ref_p.view(np.uint32))
        assert np.array_equal(v.view(np.uint32), ref_v.view(np.uint32))
        assert np.array_equal(s.get_q().view(np.uint32), ref_q.view(np.uint32))
