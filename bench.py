"""Benchmark of the Stable Cosserat Rods step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the synthetic groom generator at 100k strands x 100 vertices
(1e7 vertices), FP32, dt = 1/600 s per step = 5 substeps x 10 VBD iterations,
no SDF and no contacts in the timed sweep (the with-SDF step is reported
beside it).  One step is one launch of the fused step kernel; inputs are
resident in HBM before the timed region starts.

Metric: strand-vertex VBD updates per second, one update being one vertex's
position solve plus orientation solve in one iteration, i.e.
updates/step = vertices x substeps x iterations.

Prints ONE JSON line.  The CPU baseline is the FP32 C++ restatement of the
reference step (oracle/, OpenMP over vertices) run on a bounded sample of the
same workload on this machine's host cores.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
# Compulsory HBM bytes of one fused-step launch per vertex: read x, v, u (12 B
# each), q, q_rest, mat (16 B each) = 84 B; write x, v, u (12 B each), q (16 B)
# = 52 B.
FUSED_BYTES_PER_VERTEX = 136
# VALU issue of the fused step kernel.  rocprofv3 SQ_INSTS_VALU on the 100k x
# 100 groom (block 512: 160,000 waves, 50 iterations per launch) counts
# 2.16016e9 wave instructions per dispatch = 270.02 per wave and iteration
# (the loop has no data-dependent trip counts, so the count is the same for
# every dispatch).  gfx950 issues up to two VALU instructions per SIMD per
# quad-cycle (SQ_ACTIVE_INST_VALU exceeds one per SIMD quad-cycle on this
# kernel): 256 CUs x 4 SIMDs x 0.5 instructions per cycle x 2.4 GHz =
# 1.2288e12 wave64 instructions/s, the published FP32 peak of 157.3 TFLOP/s
# divided by the 128 FLOP of a wave64 FMA.
VALU_INSTS_PER_WAVE_ITERATION = 270.02
VALU_PEAK_WAVE_INSTS_PER_S = 256 * 4 * 0.5 * 2.4e9
# HBM bytes per vertex the fused step kernel actually moves, measured with
# rocprofv3 FETCH_SIZE and WRITE_SIZE on the 100k x 100 groom, each scaled by
# what the counter reads for a plain 1 GiB device copy (FETCH x2.0, WRITE x1.0):
# 833,258 KiB read + 508,750 KiB written per launch of 1e7 vertices.
MEASURED_TRAFFIC_BYTES_PER_VERTEX = (833_258 + 508_750) * 1024 / 1e7


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--strands", type=int, default=100_000)
    ap.add_argument("--verts", type=int, default=100)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-sdf-leg", action="store_true", help="skip the with-SDF step timing")
    ap.add_argument("--no-adjoint-leg", action="store_true", help="skip the adjoint/forward timing")
    ap.add_argument("--no-batched-leg", action="store_true", help="skip the batched-vs-per-groom timing")
    ap.add_argument("--no-contacts-leg", action="store_true", help="skip the large groom with contacts")
    ap.add_argument("--no-colored-leg", action="store_true", help="skip the even/odd coloured schedule timing")
    ap.add_argument("--no-mesh-sdf-leg", action="store_true", help="skip the mesh -> body SDF grid build timing")
    ap.add_argument("--no-lbvh-leg", action="store_true", help="skip the LBVH build/refit/query timing")
    ap.add_argument("--mesh-sdf-only", action="store_true", help="run the mesh -> body SDF grid leg only")
    ap.add_argument("--lbvh-only", action="store_true", help="run the LBVH build/refit/query leg only")
    ap.add_argument("--contacts-only", action="store_true", help="run the contacts leg only")
    ap.add_argument("--contacts-steps", type=int, default=100)
    ap.add_argument("--batched-steps", type=int, default=200)
    ap.add_argument("--only-legs", action="store_true", help="run the adjoint and batched legs only")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (rocprof byte-counter calibration)")
    return ap.parse_args()


def calibrate_copy(mib):
    """A plain copy kernel of known size: reads and writes mib MiB once per
    launch, so FETCH_SIZE / WRITE_SIZE per dispatch can be scaled."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_mib": mib, "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n}))


def pack_count(offsets, capacity):
    """Workgroups of the library's greedy packing of whole strands into
    packs of at most `capacity` vertices (one vertex per lane)."""
    offsets = np.asarray(offsets)
    packs, s, ns = 0, 0, len(offsets) - 1
    while s < ns:
        beg = offsets[s]
        # the last strand index whose end still fits this pack
        e = int(np.searchsorted(offsets, beg + capacity, side="right")) - 1
        s = max(e, s + 1)
        packs += 1
    return packs


def ensure_built():
    from kamisimulator_amd import build
    if not os.path.exists(build.LIB):
        build.build()


def cpu_model():
    """Host CPU model name (lscpu, else /proc/cpuinfo) and logical CPU count."""
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    ids = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("model name", "cpu family", "model", "stepping") and k not in ids:
                ids[k] = v.strip()
            if line.strip() == "":
                break
    except Exception:
        pass
    if not model:
        model = ids.get("model name", "")
    # a virtualised host may report only the product line: add the CPUID numbers
    if all(k in ids for k in ("cpu family", "model", "stepping")):
        model += f" (family {ids['cpu family']}, model {ids['model']}, stepping {ids['stepping']})"
    return model, os.cpu_count() or 1


def cpu_baseline(scene, params, strands=1_000, steps=100, runs=5):
    """FP32 restatement of the reference step (oracle/) on the host cores: the
    first `strands` strands of the same groom, `runs` runs of `steps` steps of
    1/600 s each from the same initial state; the median run is reported.
    Threads: OMP_NUM_THREADS (the host share this machine grants one GPU),
    else every logical CPU."""
    from kamisimulator_amd import scenes
    from oracle import pyoracle

    pyoracle.build()
    ns = min(strands, scene.num_strands)
    nv = int(scene.strand_offsets[ns])
    sub = scenes.Scene("cpu-sample", scene.positions[:nv].copy(), scene.strand_offsets[:ns + 1].copy(),
                       material=scene.material)
    model, nproc = cpu_model()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    dt = np.float32(1 / 600)
    S = int(np.ceil(dt / np.float32(params.max_h)))
    times = []
    for _ in range(runs):
        o = pyoracle.OracleSim(params, sub.positions, sub.strand_offsets, precision=32)
        o.set_threads(threads)
        scenes.setup_sim(o, sub)
        t0 = time.perf_counter()
        for _ in range(steps):
            o.advance(dt)
        times.append(time.perf_counter() - t0)
        o.close()
    med = float(np.median(times))
    ups = nv * S * params.num_itr * steps / med
    return {
        "value": ups,
        "unit": "strand-vertex VBD updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP32 C++ restatement of the reference step (oracle/, OpenMP), the first {ns} strands x "
                  f"{nv // ns} vertices of the same groom ({nv} vertices), median of {runs} runs of {steps} steps of "
                  f"dt=1/600 ({S} substeps x {params.num_itr} iterations; runs {', '.join(f'{t:.2f}' for t in times)} "
                  f"s), {threads} threads of {nproc} logical CPUs, {model}",
    }


def adjoint_leg(stream, nsteps=50, every=10):
    """Differentiation workload: 2000 strands x 64 vertices, FP32, nsteps
    steps of 1/600 s with a checkpoint every `every` steps; loss = squared tip
    distances to targets.  Times the plain forward step, the checkpointed
    forward, and the backward (re-recording each segment + reverse sweep)."""
    import torch

    from kamisimulator_amd import scenes
    from kamisimulator_amd._lib import Sim, default_params

    ns, nv = 2000, 64
    sc = scenes.groom(ns, nv, seed=4)
    sc.material = scenes.Material(num_itr=10)
    p = scenes.make_params(default_params, sc.material)
    sim = Sim(p, sc.positions, sc.strand_offsets, device=0)
    sim.set_stream(stream.cuda_stream)
    scenes.setup_sim(sim, sc)
    dt = np.float32(1 / 600)
    tips = (sc.strand_offsets[1:] - 1).astype(np.int64)

    def seed():
        x, _ = sim.get_pos_vel()
        xb = np.zeros_like(x)
        xb[tips] = 2 * (x[tips] - (sc.positions[tips] + np.float32(1e-3)))
        return xb

    # warm-up: code objects, tape and checkpoint buffers
    for _ in range(3):
        sim.advance(dt)
    sim.adjoint_forward(dt, every, every)
    sim.adjoint_backward(seed())
    torch.cuda.synchronize()
    a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    a.record(stream)
    for _ in range(nsteps):
        sim.advance(dt)
    b.record(stream)
    sim.adjoint_forward(dt, nsteps, every)
    c.record(stream)
    torch.cuda.synchronize()
    fwd_ms = a.elapsed_time(b)
    ckpt_fwd_ms = b.elapsed_time(c)
    xb = seed()
    w0 = time.perf_counter()
    sim.adjoint_backward(xb)
    wall_bwd = (time.perf_counter() - w0) * 1e3
    rec_ms, sweep_ms, bwd_ms = sim.adjoint_time()
    fwd_step = fwd_ms / nsteps
    # the same gradient with the whole tape kept in HBM (one segment: the
    # forward records every step, the backward re-records none)
    sim.adjoint_forward(dt, nsteps, nsteps)
    sim.adjoint_backward(seed())
    torch.cuda.synchronize()
    d, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    d.record(stream)
    sim.adjoint_forward(dt, nsteps, nsteps)
    e.record(stream)
    torch.cuda.synchronize()
    full_fwd_ms = d.elapsed_time(e)
    sim.adjoint_backward(seed())
    _, _, full_bwd_ms = sim.adjoint_time()
    # the checkpointed forward keeps the tape of the last segment; the
    # backward re-records the steps of every other segment, each segment on
    # a second stream while the previous (later) one is swept
    rerecorded = ((nsteps - 1) // every) * every
    S = int(np.ceil(dt / np.float32(p.max_h)))
    return {
        "workload": f"{ns} strands x {nv} vertices FP32, {nsteps} steps of dt=1/600 s ({S} substeps x {p.num_itr} "
                    f"iterations), checkpoints every {every} steps, loss on strand tips",
        "forward_ms_per_step": round(fwd_step, 4),
        "checkpointed_forward_ms_per_step": round(ckpt_fwd_ms / nsteps, 4),
        "record_ms_per_rerecorded_step": round(rec_ms / max(rerecorded, 1), 4),
        "reverse_sweep_ms_per_step": round(sweep_ms / nsteps, 4),
        "backward_ms_per_step": round(bwd_ms / nsteps, 4),
        "adjoint_step_over_forward_step": round(bwd_ms / nsteps / fwd_step, 3),
        "sequential_record_plus_sweep_over_forward_step": round(
            ((rec_ms / max(rerecorded, 1) if rerecorded else 0.0) + sweep_ms / nsteps) / fwd_step, 3),
        "backward_over_forward": round(bwd_ms / fwd_ms, 3),
        "gradient_over_forward": round((ckpt_fwd_ms + bwd_ms) / fwd_ms, 3),
        "backward_wall_ms": round(wall_bwd, 2),
        "whole_tape_recording_forward_ms_per_step": round(full_fwd_ms / nsteps, 4),
        "whole_tape_backward_ms_per_step": round(full_bwd_ms / nsteps, 4),
        "whole_tape_adjoint_step_over_forward_step": round(full_bwd_ms / nsteps / fwd_step, 3),
        "whole_tape_gradient_over_forward": round((full_fwd_ms + full_bwd_ms) / fwd_ms, 3),
        "note": "adjoint step = the whole backward pass's device time per step (each segment re-recorded from its "
                "checkpoint on a second stream while the later segment is swept; "
                f"{rerecorded} of {nsteps} steps re-recorded, the last segment's tape kept by the checkpointed "
                "forward) over the plain forward step; the sequential figure adds one re-recorded step and one "
                "reverse sweep as if they ran one after the other; gradient = checkpointed forward + backward; "
                "whole_tape_*: one checkpoint segment, the forward records the tape of every step "
                f"({nsteps} steps held in HBM at once) and the backward only sweeps",
    }


def gradient_leg():
    """The differentiation workload's gradient on the device: 2k strands x 64
    vertices, 50 steps, checkpoints every 10; targets = the tips after the
    same rollout with E and kappa perturbed by a seeded +-20 %; dL/dE and
    dL/dkappa per strand at E = 3 GPa, kappa = 1, through the torch.autograd
    binding.  Returns what the CPU check needs."""
    import torch

    from kamisimulator_amd import scenes
    from kamisimulator_amd.autograd import Rollout

    sc, e_t, k_t = scenes.config_differentiation()
    ns = sc.num_strands
    dt = np.float32(1 / 600)
    roll = Rollout(sc, dt, 50, 10, device=0)
    with torch.no_grad():
        x_t, _ = roll(torch.tensor(e_t), torch.tensor(k_t))
    tips = torch.from_numpy((sc.strand_offsets[1:] - 1).astype(np.int64))
    targets = x_t[tips].double()
    E = torch.full((ns,), 3.0, dtype=torch.float64, requires_grad=True)
    kap = torch.ones(ns, dtype=torch.float64, requires_grad=True)
    w0 = time.perf_counter()
    x, _ = roll(E, kap)
    loss = ((x[tips].double() - targets) ** 2).sum()
    loss.backward()
    wall = time.perf_counter() - w0
    return {"scene": sc, "targets": targets.numpy(), "gE": E.grad.numpy(), "gk": kap.grad.numpy(),
            "dt": dt, "loss": float(loss), "wall_s": wall}


def gradient_accuracy(g, params):
    """CPU check of the device gradient of gradient_leg against the oracle
    (oracle/gradient_check.py): FP32 and FP64 gradients of the first 16
    strands, FP64 central differences on the first 4."""
    from oracle import gradient_check

    ns = g["scene"].num_strands
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    r = gradient_check.check(g["scene"], params, np.full(ns, 3.0), np.ones(ns), g["dt"], 50, 10, g["targets"],
                             g["gE"], g["gk"], strands=16, fd_strands=4, threads=threads)
    r["check_s"] = round(time.perf_counter() - t0, 1)
    r["note"] = ("the FP64 oracle's adjoint (the same reverse sweep in double) meets 1e-3 against FP64 central "
                 "differences; the FP32 gradient does not: the FP32 rollout's rounding noise is larger than the "
                 "FP64 displacement (fp32_position_noise_over_fp64_displacement), because the stretch strain "
                 "gravity induces (~4e-9) lies below the FP32 resolution of the frame direction (~6e-8), and the "
                 "adaptive initial guess's clamp takes another branch in FP32 at a large share of vertex-substeps")
    return r


def batched_leg(stream, nsteps, strands=(1_000, 50_000), num_grooms=64):
    """Batched form: num_grooms independent grooms (strand counts log-uniform
    in `strands`, 32-256 vertices per strand, own 128^3 sphere SDF and gravity
    direction) stepped nsteps times (a) one launch per groom per step and (b)
    one batched launch per step replayed from a hipGraph, block autotuned."""
    import torch

    from kamisimulator_amd import scenes
    from kamisimulator_amd._lib import BATCH_GRID, BATCH_PERSISTENT, Batch, Sim, default_params

    t0 = time.perf_counter()
    grooms = scenes.config_batched(num_grooms=num_grooms, strands=strands)
    sims = []
    for sc in grooms:
        sim = Sim(scenes.make_params(default_params, sc.material), sc.positions, sc.strand_offsets, device=0)
        sim.set_stream(stream.cuda_stream)
        scenes.setup_sim(sim, sc)
        sims.append(sim)
    verts = sum(sc.num_verts for sc in grooms)
    print(f"[bench] batched leg: {num_grooms} grooms, {verts} vertices, setup {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    dt = np.float32(1 / 600)
    for sim in sims:
        sim.solve_next_step(dt)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(nsteps):
        for sim in sims:
            sim.solve_next_step(dt)
    torch.cuda.synchronize()
    per_groom_ms = (time.perf_counter() - w0) * 1e3
    batch = Batch(sims)
    default_block = batch.block_size
    best = batch.autotune(dt, 3)
    best_mapping = batch.mapping

    def graph_run(mapping):
        batch.set_mapping(mapping)
        batch.set_graph(True)
        batch.step(dt, nsteps, with_sdf=True)  # capture + first replay
        batch.synchronize()
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        batch.step(dt, nsteps, with_sdf=True)
        batch.synchronize()
        return (time.perf_counter() - w0) * 1e3

    other = BATCH_GRID if best_mapping == BATCH_PERSISTENT else BATCH_PERSISTENT
    other_ms = graph_run(other)
    graph_ms = graph_run(best_mapping)
    names = {BATCH_GRID: "grid", BATCH_PERSISTENT: "persistent"}
    batch.set_graph(False)
    w0 = time.perf_counter()
    batch.step(dt, nsteps, with_sdf=True)
    batch.synchronize()
    nograph_ms = (time.perf_counter() - w0) * 1e3
    batch.close()
    for sim in sims:
        sim.close()
    p0 = scenes.make_params(default_params, grooms[0].material)
    S = int(np.ceil(dt / np.float32(p0.max_h)))
    ups = verts * S * p0.num_itr
    return {
        "workload": f"{num_grooms} grooms, strands log-uniform {strands[0]}-{strands[1]} ({verts} vertices), "
                    f"32-256 vertices per strand, own 128^3 sphere SDF and gravity, {nsteps} steps of 1/600 s",
        "per_groom_launches_ms_per_step": round(per_groom_ms / nsteps, 4),
        "batched_graph_ms_per_step": round(graph_ms / nsteps, 4),
        "batched_no_graph_ms_per_step": round(nograph_ms / nsteps, 4),
        "speedup_graph_vs_per_groom": round(per_groom_ms / graph_ms, 3),
        "batched_graph_updates_per_s": ups * nsteps / (graph_ms / 1e3),
        "autotuned_block": best,
        "autotuned_mapping": names[best_mapping],
        f"batched_graph_{names[other]}_ms_per_step": round(other_ms / nsteps, 4),
        "default_block": default_block,
        "timing": "host wall clock between device synchronisations; mappings: grid = one workgroup per pack, "
                  "persistent = as many workgroups as the GPU holds taking packs from a device work counter",
    }


def contacts_leg(stream, nsteps=100, strands=100_000):
    """Large groom with hair-hair contacts: 100k strands with lognormal vertex
    counts (median 96, clipped 32-256) on evenly spaced (spiral) roots with
    0.2-0.6 mm curls, sphere + capsule SDF on 256^3, LBVH broad phase, contacts
    at 2 x CurveRadius, dt = 1/600.  (Uniformly random roots put some roots
    microns apart; their strands then run side by side and exceed the 16
    contacts per segment the step allows, which the reference treats as fatal.)"""
    import torch

    from kamisimulator_amd import scenes
    from kamisimulator_amd._lib import KscrError, Sim, default_params

    t0 = time.perf_counter()
    sc = scenes.config_large(strands, seed=2, sdf_dim=256, curl=(2e-4, 6e-4), roots="spiral")
    p = scenes.contact_params(default_params, sc.material)
    sim = Sim(p, sc.positions, sc.strand_offsets, device=0)
    sim.set_stream(stream.cuda_stream)
    scenes.setup_sim(sim, sc)
    setup_s = time.perf_counter() - t0
    dt = np.float32(1 / 600)
    S = int(np.ceil(dt / np.float32(p.max_h)))
    out = {"workload": f"{strands} strands, lognormal vertex counts ({sc.num_verts} vertices), spiral roots, 0.2-0.6 mm "
                       f"curls, 256^3 sphere + capsule SDF, hair-hair contacts, {nsteps} steps of 1/600 s ({S} substeps x "
                       f"{p.num_itr} iterations)"}
    try:
        sim.solve_next_step(dt)
        sim.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(nsteps):
            sim.solve_next_step(dt)
        b.record(stream)
        sim.synchronize()
        ms = a.elapsed_time(b) / nsteps
        num, _ = sim.get_contacts()
        out.update({"ms_per_step": round(ms, 3), "updates_per_s": sc.num_verts * S * p.num_itr / (ms / 1e3),
                    "contacts": int(num.sum()), "max_contacts_per_segment": int(num.max()),
                    "setup_s": round(setup_s, 1)})
    except KscrError as e:
        out["error"] = str(e)
    sim.close()
    return out


def lbvh_leg(stream, n=100_000, reps=20):
    """The reference's own LBVH benchmark (Kitten::testLBVH,
    KE/KittenGpuLBVH/lbvh.cu:554-565: 100k unit-cube points padded by 0.001,
    srand(1)), with the boxes and the pair buffer (100 n pairs) in device
    memory as LBVH::compute / refit / query take them.  Each call is timed with
    HIP events on the launch stream (median of `reps`); compute and query read
    a count back to the host, as the reference's do (lbvh.cu:393, :425).  The
    reference's README quotes 150 / 70 / 60 us on an RTX 3090
    (KittenGpuLBVH/README.md:18)."""
    import torch

    from kamisimulator_amd import scenes
    from kamisimulator_amd._lib import Lbvh

    boxes = scenes.lbvh_recipe_boxes(n)
    d_boxes = torch.from_numpy(boxes).to("cuda:0")
    cap = 100 * n
    d_pairs = torch.empty((cap, 2), dtype=torch.int32, device="cuda:0")
    t = Lbvh(device=0)
    sp = stream.cuda_stream
    times = {"build": [], "refit": [], "query": []}
    count = 0
    for r in range(reps + 2):  # two warm-up rounds
        for name in ("build", "refit", "query"):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            if name == "build":
                t.compute_device(d_boxes.data_ptr(), n, sp)
            elif name == "refit":
                t.refit_device(d_boxes.data_ptr(), sp)
            else:
                count = t.query_device(d_pairs.data_ptr(), cap, sp)
            b.record(stream)
            b.synchronize()
            if r >= 2:
                times[name].append(a.elapsed_time(b) * 1e3)
    t.close()
    med = {k: round(float(np.median(v)), 1) for k, v in times.items()}
    ref = {"build": 150.0, "refit": 70.0, "query": 60.0}
    return {"boxes": n, "pairs": int(count), "us": med,
            "reference_rtx3090_us": ref,
            "ratio_vs_reference": {k: round(ref[k] / med[k], 2) for k in med},
            "timing": f"HIP events on the launch stream around each call, median of {reps}"}


def colored_leg(stream, scene, dt, updates_per_step, nsteps):
    """The headline groom with the even/odd coloured (Gauss-Seidel) schedule:
    same numItr, same updates per step; no reference counterpart."""
    import torch

    from kamisimulator_amd import scenes
    from kamisimulator_amd._lib import SCHEDULE_COLORED, Sim, default_params

    p = scenes.make_params(default_params, scene.material, schedule=SCHEDULE_COLORED)
    sim = Sim(p, scene.positions, scene.strand_offsets, device=0)
    sim.set_stream(stream.cuda_stream)
    scenes.setup_sim(sim, scene)
    for _ in range(2):
        sim.advance(dt)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(nsteps):
        sim.advance(dt)
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / nsteps
    out = {"ms_per_step": round(ms, 4), "updates_per_s": updates_per_step / (ms / 1e3), "block": sim.block_size,
           "note": "even/odd vertex colouring: each colour solved with the other colour's newest dx and q "
                   "(two vertices per lane); k coloured iterations reach about the accuracy of 2k Jacobi "
                   "iterations (tests/test_oracle_colored.py)"}
    sim.close()
    return out


def mesh_sdf_leg(reps=5, cpu=True):
    """Body SDF grid of an 81,920-triangle sphere mesh on the reference's
    default grid (64 x 128 x 64, PNMesh.cpp:18), built on the GPU
    (kscr_mesh_sdf: host pseudo-normals, triangle BVH, cubic grid node
    distances, body grid sampling, copy back); the CPU figure is the
    reference's own construction (Kami::SDF + Discregrid, built from the
    reference sources into oracle/_ref/discregrid_probe, OpenMP) on the same
    mesh and grid, when that probe is present."""
    from kamisimulator_amd import mesh_sdf, scenes

    v, t = scenes.sphere_mesh(6, radius=0.1)
    dim = (64, 128, 64)
    mesh_sdf(v, t, dim, device=0)
    ws = []
    for _ in range(reps):
        w0 = time.perf_counter()
        mesh_sdf(v, t, dim, device=0)
        ws.append((time.perf_counter() - w0) * 1e3)
    n = [int(x) for x in dim]
    cubic_nodes = (n[0] + 1) * (n[1] + 1) * (n[2] + 1) + 2 * (n[0] * (n[1] + 1) * (n[2] + 1) +
                                                              (n[0] + 1) * n[1] * (n[2] + 1) +
                                                              (n[0] + 1) * (n[1] + 1) * n[2])
    out = {"triangles": int(len(t)), "grid": list(dim), "cubic_grid_nodes": cubic_nodes,
           "gpu_ms_per_build": round(float(np.median(ws)), 2)}
    probe = os.path.join(REPO, "oracle", "_ref", "discregrid_probe")
    if cpu and os.path.exists(probe):
        import tempfile
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        with tempfile.TemporaryDirectory() as td:
            src, dst = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
            with open(src, "wb") as f:
                f.write(np.array([len(v), len(t), *dim, 0], np.int32).tobytes())
                f.write(v.tobytes())
                f.write(t.tobytes())
            w0 = time.perf_counter()
            subprocess.run([probe, src, dst], check=True, stdout=subprocess.DEVNULL)
            cpu_ms = (time.perf_counter() - w0) * 1e3
        out["cpu_reference"] = {"ms_per_build": round(cpu_ms, 1), "cores": threads, "kind": "reference",
                                "sample": "the whole build: Kami::SDF::Init (TriangleMeshDistance + "
                                          "CubicLagrangeDiscreteGrid::addFunction) and the StoreSDFQueries node "
                                          "loop, OpenMP"}
        out["speedup_vs_cpu_reference"] = round(cpu_ms / out["gpu_ms_per_build"], 1)
    return out


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the strand-sharded multi-GPU "
              "layout is described in DESIGN.md but not built", file=sys.stderr)
        sys.exit(2)
    ensure_built()
    import torch

    from kamisimulator_amd import scenes
    from kamisimulator_amd._lib import Sim, default_params

    assert torch.cuda.is_available(), "no GPU visible"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    if args.mesh_sdf_only:
        print(json.dumps({"mesh_sdf": mesh_sdf_leg(cpu=not args.no_cpu)}), flush=True)
        return
    if args.lbvh_only:
        print(json.dumps({"lbvh": lbvh_leg(stream)}), flush=True)
        return
    if args.contacts_only:
        print(json.dumps({"contacts": contacts_leg(stream, args.contacts_steps)}), flush=True)
        return
    if args.only_legs:
        legs = {}
        if not args.no_adjoint_leg:
            legs["adjoint"] = adjoint_leg(stream)
            print(f"[bench] adjoint leg {legs['adjoint']}", file=sys.stderr, flush=True)
        if not args.no_batched_leg:
            legs["batched"] = batched_leg(stream, args.batched_steps)
        if not args.no_contacts_leg:
            legs["contacts"] = contacts_leg(stream, args.contacts_steps)
        print(json.dumps(legs), flush=True)
        return

    t0 = time.perf_counter()
    scene = scenes.config_groom(args.strands, args.verts, seed=1, sdf_dim=128, with_sdf=True)
    params = scenes.make_params(default_params, scene.material, block_size=args.block)
    sim = Sim(params, scene.positions, scene.strand_offsets, device=0)
    sim.set_stream(stream.cuda_stream)
    sdf = scene.sdf
    scene.sdf = None
    scenes.setup_sim(sim, scene)
    block = sim.block_size
    print(f"[bench] setup {time.perf_counter() - t0:.1f}s  verts={scene.num_verts} strands={scene.num_strands} "
          f"block={sim.block_size}", file=sys.stderr, flush=True)

    dt = np.float32(1 / 600)
    S = int(np.ceil(dt / np.float32(params.max_h)))
    updates_per_step = scene.num_verts * S * params.num_itr

    for _ in range(args.warmup):
        sim.advance(dt)
    torch.cuda.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t_start.record(stream)
    for k in range(args.steps):
        ev[k][0].record(stream)
        sim.advance(dt)
        ev[k][1].record(stream)
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = t_start.elapsed_time(t_end)
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    ms_per_step = total_ms / args.steps
    value = updates_per_step * args.steps / (total_ms / 1e3)
    print(f"[bench] {args.steps} steps: {total_ms:.2f} ms (wall {wall * 1e3:.1f} ms), kernel avg {kern_ms:.3f} ms, "
          f"{value:.3e} updates/s", file=sys.stderr, flush=True)

    achieved = FUSED_BYTES_PER_VERTEX * scene.num_verts / (kern_ms / 1e3) / 1e9
    waves = pack_count(scene.strand_offsets, block) * (block // 64)
    valu_insts = VALU_INSTS_PER_WAVE_ITERATION * waves * S * params.num_itr
    valu_rate = valu_insts / (kern_ms / 1e3)
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": round(MEASURED_TRAFFIC_BYTES_PER_VERTEX * scene.num_verts),
        "kernel": "step_fused_kernel",
        "algorithmic_bytes_per_launch": FUSED_BYTES_PER_VERTEX * scene.num_verts,
        "avg_launch_ms": round(kern_ms, 4),
        "note": "one launch = one step (5 substeps x 10 iterations) with vertex state resident in registers/LDS; "
                "algorithmic bytes = compulsory 136 B/vertex per launch: the HBM fraction is small by construction "
                "and the bound is VALU issue (valu)",
        "valu": {
            "bound": "valu",
            "achieved": valu_rate,
            "peak": VALU_PEAK_WAVE_INSTS_PER_S,
            "unit": "wave64 VALU instructions/s",
            "frac": round(valu_rate / VALU_PEAK_WAVE_INSTS_PER_S, 4),
            "insts_per_launch": valu_insts,
            "waves_per_launch": waves,
            "note": "achieved = SQ_INSTS_VALU per launch (270.02 per wave and iteration, rocprofv3) / the launch "
                    "time measured here; peak at the 2.4 GHz peak engine clock, two VALU issues per SIMD per "
                    "quad-cycle.  Under load the clock runs near 1.9 GHz (GRBM_GUI_ACTIVE over the dispatch), where "
                    "SQ_ACTIVE_INST_VALU fills about 0.8 of the issue slots "
                    "(profiles/step_fused_kernel_groom100kx100.md)",
        },
    }

    sdf_leg = None
    if not args.no_sdf_leg:
        sim.set_sdf(sdf, threshold=0.0)
        sim.solve_next_step(dt)
        torch.cuda.synchronize()
        n2 = max(5, args.steps // 5)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(n2):
            sim.solve_next_step(dt)
        b.record(stream)
        torch.cuda.synchronize()
        ms2 = a.elapsed_time(b) / n2
        sdf_leg = {"ms_per_step": round(ms2, 4), "updates_per_s": updates_per_step / (ms2 / 1e3),
                   "sdf_grid": "128^3 sphere SDF, threshold 0"}

    colored = None
    if not args.no_colored_leg:
        colored = colored_leg(stream, scene, dt, updates_per_step, max(5, args.steps // 5))
        print(f"[bench] colored leg: {colored['ms_per_step']} ms/step", file=sys.stderr, flush=True)

    # Host exchange of the plugin interface: SolveNextStep downloads positions
    # and velocities every step (StableCosseratRodsSolver.cpp:190, :223-243).
    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - w0) * 1e3 / reps

    copy_ms = timed(sim.get_pos_vel)
    pinned_ms = timed(sim.download_pos_vel)
    nbytes = int(24 * scene.num_verts)
    exchange = {"download_ms": round(pinned_ms, 3), "bytes": nbytes,
                "link_gb_per_s": round(nbytes / (pinned_ms / 1e3) / 1e9, 2),
                "copy_out_download_ms": round(copy_ms, 3),
                "pcie_inclusive_updates_per_s": updates_per_step / ((ms_per_step + pinned_ms) / 1e3),
                "note": "device-to-host copy of x and v once per step, as the plugin's DownloadHair: "
                        "download_ms into the library's page-locked mirror (kscr_download_pos_vel, what the host "
                        "solver scatters from), copy_out_download_ms also copying into a pageable caller array "
                        "(kscr_get_pos_vel); with DeviceResidentHair the download happens once per draw frame "
                        "instead; not part of value"}
    sim.close()
    adjoint = None
    grad = None
    if not args.no_adjoint_leg:
        grad = gradient_leg()
        adjoint = adjoint_leg(stream)
        adjoint["gradient_workload"] = ("2000 strands x 64 vertices, 50 steps, checkpoints every 10, targets from "
                                        "E and kappa perturbed by a seeded +-20 %, gradient at E = 3 GPa, kappa = 1 "
                                        f"(loss {grad['loss']:.6e}; rollout + backward {grad['wall_s']:.2f} s wall)")
        print(f"[bench] adjoint leg: forward {adjoint['forward_ms_per_step']} ms/step, adjoint step "
              f"{adjoint['adjoint_step_over_forward_step']}x", file=sys.stderr, flush=True)
    batched = None
    if not args.no_batched_leg:
        batched = batched_leg(stream, args.batched_steps)
        print(f"[bench] batched leg: per-groom {batched['per_groom_launches_ms_per_step']} ms/step, graph "
              f"{batched['batched_graph_ms_per_step']} ms/step", file=sys.stderr, flush=True)

    lbvh = None
    if not args.no_lbvh_leg:
        lbvh = lbvh_leg(stream)
        print(f"[bench] lbvh leg: {lbvh}", file=sys.stderr, flush=True)

    contacts = None
    if not args.no_contacts_leg:
        contacts = contacts_leg(stream)
        print(f"[bench] contacts leg: {contacts}", file=sys.stderr, flush=True)

    mesh = None
    if not args.no_mesh_sdf_leg:
        mesh = mesh_sdf_leg(cpu=not args.no_cpu)
        print(f"[bench] mesh sdf leg: {mesh}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        if grad is not None:
            from kamisimulator_amd import scenes as _sc
            adjoint["gradient_accuracy"] = gradient_accuracy(grad, _sc.make_params(default_params, grad["scene"].material))
            print(f"[bench] gradient accuracy {adjoint['gradient_accuracy']}", file=sys.stderr, flush=True)
        cpu = cpu_baseline(scene, params)
        print(f"[bench] cpu baseline {cpu['value']:.3e} updates/s on {cpu['cores']} threads", file=sys.stderr, flush=True)

    line = {
        "metric": "strand-vertex VBD updates/s @100k x 100 FP32",
        "value": value,
        "unit": "strand-vertex VBD updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"synthetic groom {scene.num_strands} strands x {args.verts} vertices on a 0.1 m sphere scalp, "
                        f"FP32, dt=1/600 s = {S} substeps x {params.num_itr} VBD iterations, Jacobi schedule, "
                        "no SDF or contacts in the timed step",
            "strands": scene.num_strands,
            "vertices": scene.num_verts,
            "updates_per_step": updates_per_step,
            "block": block,
        },
        "roofline": roofline,
        "with_sdf_step": sdf_leg,
        "colored_step": colored,
        "host_exchange": exchange,
        "adjoint": adjoint,
        "batched": batched,
        "contacts": contacts,
        "lbvh": lbvh,
        "mesh_sdf": mesh,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
