// kscr_api_adjoint.hip -- differentiation entry points of the C ABI
// (kscr_set_material, kscr_adjoint_forward / _backward / _last_time): the
// checkpointed rollout, the two-slot tape, and the backward pass that
// re-records segment j - 1 on a second stream while segment j is swept
// (kernels in kscr_adjoint.hip and kscr_step.hip).
#include <cstring>

#include "kscr_api_internal.h"

using namespace kscr;
using namespace kscr::api;

namespace {

hipError_t checkpoint_copy(kscr_handle* h, size_t k, bool save, hipStream_t stream) {
    AdjointRun& A = h->adj;
    const size_t n = h->n, n3 = 3 * n;
    const hipMemcpyKind kind = hipMemcpyDeviceToDevice;
    float* x = A.ck_x.p + k * n3;
    float* v = A.ck_v.p + k * n3;
    float* u = A.ck_u.p + k * n3;
    float4* q = A.ck_q.p + k * n;
    hipError_t e;
    if (save) {
        if ((e = hipMemcpyAsync(x, h->d_x.p, n3 * 4, kind, stream)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(v, h->d_v.p, n3 * 4, kind, stream)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(u, h->d_u.p, n3 * 4, kind, stream)) != hipSuccess) return e;
        return hipMemcpyAsync(q, h->d_q.p, n * 16, kind, stream);
    }
    if ((e = hipMemcpyAsync(h->d_x.p, x, n3 * 4, kind, stream)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(h->d_v.p, v, n3 * 4, kind, stream)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(h->d_u.p, u, n3 * 4, kind, stream)) != hipSuccess) return e;
    return hipMemcpyAsync(h->d_q.p, q, n * 16, kind, stream);
}

AdjointRun::Clock clock_of(const kscr_handle* h) { return AdjointRun::Clock{h->h, h->last_h, h->time, h->step_counter}; }
void set_clock(kscr_handle* h, const AdjointRun::Clock& c) {
    h->h = c.h, h->last_h = c.last_h, h->time = c.time, h->step_counter = c.counter;
}

// Tapes of two checkpoint segments (slots 0 and 1), each `every` steps of
// `substeps` substeps.
int reserve_tape(kscr_handle* h) {
    AdjointRun& A = h->adj;
    const size_t n = h->n, S = (size_t)A.substeps, I1 = (size_t)A.num_itr + 1, every = 2 * (size_t)A.every;
    const size_t K = (size_t)A.num_itr;
    const size_t per_sub = S * n * every, per_dx = S * I1 * n * every, per_q = S * I1 * n * every;
    const size_t per_ilw = S * K * n * every;
    hipError_t e;
    if (A.wind && A.t_acc.n < per_sub * 3 && (e = A.t_acc.alloc(per_sub * 3)) != hipSuccess)
        return h->hip_fail(e, "hipMalloc tape");
    if (A.t_dx.n >= per_dx * 3 && A.t_q.n >= per_q && A.t_x.n >= per_sub * 3 && A.t_ilw.n >= per_ilw)
        return KSCR_OK;
    if ((e = A.t_x.alloc(per_sub * 3)) != hipSuccess || (e = A.t_y.alloc(per_sub * 3)) != hipSuccess ||
        (e = A.t_sigma.alloc(per_sub)) != hipSuccess || (e = A.t_dx.alloc(per_dx * 3)) != hipSuccess ||
        (e = A.t_q.alloc(per_q)) != hipSuccess || (e = A.t_ilw.alloc(std::max<size_t>(per_ilw, 1))) != hipSuccess)
        return h->hip_fail(e, "hipMalloc tape");
    return KSCR_OK;
}

// Tape of step j of the segment in tape slot `slot`.
TapeArgs tape_of(kscr_handle* h, size_t slot, size_t j) {
    AdjointRun& A = h->adj;
    const size_t n = h->n, S = (size_t)A.substeps, K = (size_t)A.num_itr;
    j += slot * (size_t)A.every;
    TapeArgs t{};
    t.x = A.t_x.p + j * S * n * 3;
    t.y = A.t_y.p + j * S * n * 3;
    t.sigma_free = A.t_sigma.p + j * S * n;
    t.dx = A.t_dx.p + j * S * (K + 1) * n * 3;
    t.q = A.t_q.p + j * S * (K + 1) * n;
    t.ilw = A.t_ilw.p + j * S * K * n;
    t.acc = A.wind ? A.t_acc.p + j * S * n * 3 : nullptr;
    t.n = (int)n;
    return t;
}

int adjoint_supported(kscr_handle* h) {
    if (h->p.enable_hair_hair_coll || h->p.use_step_size_limit)
        return h->fail(KSCR_ERR_UNSUPPORTED, "the adjoint covers the contact-free step (no contacts, no step size limit)");
    if (h->has_glue) return h->fail(KSCR_ERR_UNSUPPORTED, "the adjoint covers strands without glued vertices");
    if (split_strands(h))
        return h->fail(KSCR_ERR_UNSUPPORTED, "the adjoint covers strands of at most " + std::to_string(kMaxBlock) +
                                                 " vertices");
    if (colored(h)) return h->fail(KSCR_ERR_UNSUPPORTED, "the adjoint covers the Jacobi schedule");
    return KSCR_OK;
}

}  // namespace

extern "C" {

int kscr_set_material(kscr_handle* h, const float* k_stretch, const float* q_rest) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->configured) return h->fail(KSCR_ERR_NOT_CONFIGURED, "No rest length. Must call configure()");
    for (int i = 0; i < h->n; i++) {
        if (k_stretch) h->k_stretch[i] = k_stretch[i];
        if (q_rest) h->q_rest[i] = f4{q_rest[4 * i], q_rest[4 * i + 1], q_rest[4 * i + 2], q_rest[4 * i + 3]};
    }
    if (!h->uploaded) return KSCR_OK;
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    return upload_material(h);
}

// ------------------------------------------------------------------ adjoint
namespace {
int adjoint_forward(kscr_handle* h, float dt, int nsteps, int checkpoint_every, bool with_sdf) {
    if (!h || !(dt > 0) || nsteps <= 0 || checkpoint_every <= 0) return KSCR_ERR_INVALID_ARGUMENT;
    if (!h->uploaded) return h->fail(KSCR_ERR_NOT_CONFIGURED, "upload() has not been called");
    int rc = validate_meta(h);
    if (rc) return rc;
    if ((rc = adjoint_supported(h))) return rc;
    if (with_sdf && h->strain_on && h->sdf_on && h->sdf_set)
        return h->fail(KSCR_ERR_UNSUPPORTED, "the adjoint does not cover strain limiting after the SDF pass");
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    AdjointRun& A = h->adj;
    A.ready = false;
    const size_t n = h->n;
    const size_t slots = (size_t)(nsteps + checkpoint_every - 1) / checkpoint_every + 1;  // + end state
    if (A.ck_q.n < slots * n) {
        if ((e = A.ck_x.alloc(slots * n * 3)) != hipSuccess || (e = A.ck_v.alloc(slots * n * 3)) != hipSuccess ||
            (e = A.ck_u.alloc(slots * n * 3)) != hipSuccess || (e = A.ck_q.alloc(slots * n)) != hipSuccess)
            return h->hip_fail(e, "hipMalloc checkpoints");
    }
    A.dt = dt;
    A.nsteps = nsteps;
    A.every = checkpoint_every;
    A.substeps = std::max(1, (int)std::ceil(dt / h->p.max_h));
    A.num_itr = h->p.num_itr;
    A.last_recorded = false;
    A.with_sdf = with_sdf;
    A.wind = h->p.wind_peak != 0;
    if ((rc = reserve_tape(h)) != KSCR_OK) return rc;
    // steps of the last segment write their tape as they run
    const int last_s0 = ((nsteps - 1) / checkpoint_every) * checkpoint_every;
    A.last_h_of.assign(checkpoint_every, 0.f);
    A.last_lh_of.assign(checkpoint_every, 0.f);
    A.last_t_of.assign(checkpoint_every, 0.f);
    A.clocks.clear();
    for (int s = 0; s < nsteps; s++) {
        if (s % checkpoint_every == 0) {
            if ((e = checkpoint_copy(h, (size_t)(s / checkpoint_every), true, h->stream)) != hipSuccess)
                return h->hip_fail(e, "save checkpoint");
            A.clocks.push_back(clock_of(h));
        }
        if (s < last_s0) {
            if ((rc = launch_step(h, dt, with_sdf)) != KSCR_OK) return rc;
            continue;
        }
        const StepArgs sa = step_args(h, dt, with_sdf);
        A.last_h_of[s - last_s0] = sa.h;
        A.last_lh_of[s - last_s0] = sa.last_h;
        A.last_t_of[s - last_s0] = sa.time;
        if ((e = launch_step_record(sa, tape_of(h, 0, (size_t)(s - last_s0)), h->d_packs.p, h->num_packs, h->block,
                                    h->stream)) != hipSuccess)
            return h->hip_fail(e, "record kernel launch");
    }
    A.last_recorded = true;
    A.ready = true;
    return KSCR_OK;
}
}  // namespace

int kscr_adjoint_forward(kscr_handle* h, float dt, int nsteps, int checkpoint_every) {
    return adjoint_forward(h, dt, nsteps, checkpoint_every, false);
}

int kscr_adjoint_forward_sdf(kscr_handle* h, float dt, int nsteps, int checkpoint_every) {
    return adjoint_forward(h, dt, nsteps, checkpoint_every, true);
}

int kscr_adjoint_backward(kscr_handle* h, const float* x_bar, const float* v_bar, float* k_stretch_bar,
                          float* q_rest_bar) {
    if (!h || !k_stretch_bar || !q_rest_bar) return KSCR_ERR_INVALID_ARGUMENT;
    AdjointRun& A = h->adj;
    if (!A.ready) return h->fail(KSCR_ERR_NOT_CONFIGURED, "kscr_adjoint_forward has not been run");
    if (h->p.num_itr != A.num_itr || std::max(1, (int)std::ceil(A.dt / h->p.max_h)) != A.substeps)
        return h->fail(KSCR_ERR_INVALID_ARGUMENT, "num_itr or max_h changed since the forward pass");
    if ((h->p.wind_peak != 0) != A.wind)
        return h->fail(KSCR_ERR_INVALID_ARGUMENT, "the wind was switched on or off since the forward pass");
    int rc = adjoint_supported(h);
    if (rc) return rc;
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    const size_t n = h->n, n3 = 3 * n;
    if ((rc = reserve_tape(h)) != KSCR_OK) return rc;
    if (A.qb.n != n) {
        if ((e = A.xb.alloc(n3)) != hipSuccess || (e = A.vb.alloc(n3)) != hipSuccess || (e = A.ub.alloc(n3)) != hipSuccess ||
            (e = A.ksb.alloc(n)) != hipSuccess || (e = A.qb.alloc(n)) != hipSuccess || (e = A.qrb.alloc(n)) != hipSuccess)
            return h->hip_fail(e, "hipMalloc adjoint state");
    }
    const hipMemcpyKind h2d = hipMemcpyHostToDevice;
    if ((e = x_bar ? hipMemcpyAsync(A.xb.p, x_bar, n3 * 4, h2d, h->stream) : hipMemsetAsync(A.xb.p, 0, n3 * 4, h->stream)) !=
            hipSuccess ||
        (e = v_bar ? hipMemcpyAsync(A.vb.p, v_bar, n3 * 4, h2d, h->stream) : hipMemsetAsync(A.vb.p, 0, n3 * 4, h->stream)) !=
            hipSuccess ||
        (e = hipMemsetAsync(A.ub.p, 0, n3 * 4, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(A.ksb.p, 0, n * 4, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(A.qb.p, 0, n * 16, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(A.qrb.p, 0, n * 16, h->stream)) != hipSuccess)
        return h->hip_fail(e, "seed adjoint state");
    const int nseg = (int)A.clocks.size();
    const size_t end_slot = (size_t)nseg;
    if ((e = checkpoint_copy(h, end_slot, true, h->stream)) != hipSuccess) return h->hip_fail(e, "save end state");
    const AdjointRun::Clock end_clock = clock_of(h);
    if (!A.rec_stream && (e = hipStreamCreateWithFlags(&A.rec_stream, hipStreamNonBlocking)) != hipSuccess)
        return h->hip_fail(e, "hipStreamCreate");
    hipStream_t ms = h->stream, rs = A.rec_stream;

    AdjointArgs aa{};
    aa.mat = h->d_mat.p;
    aa.q_rest = h->d_qr.p;
    aa.num_substeps = A.substeps;
    aa.num_itr = A.num_itr;
    aa.damping = h->p.damping;
    aa.velocity_decay = h->p.velocity_decay;
    aa.acceleration_ratio = h->p.acceleration_ratio;
    for (int k = 0; k < 3; k++) aa.ext[k] = h->p.gravity[k] * h->p.external_force_multiplier;
    // |ext|^2 as the device dot product forms it
    aa.ext2 = std::fma(aa.ext[2], aa.ext[2], std::fma(aa.ext[1], aa.ext[1], aa.ext[0] * aa.ext[0]));
    for (int k = 0; k < 3; k++) aa.gravity[k] = h->p.gravity[k];
    aa.external_force_multiplier = h->p.external_force_multiplier;
    aa.wind_peak = h->p.wind_peak;
    aa.wind_y_freq = h->p.wind_y_freq;
    aa.wind_z_freq = h->p.wind_z_freq;
    aa.wind_time_freq = h->p.wind_time_freq;
    aa.wind_sharpness = h->p.wind_sharpness;
    // the SDF pass as step_args gives it to the forward's steps
    aa.sdf_on = (A.with_sdf && h->sdf_on && h->sdf_set) ? 1 : 0;
    aa.sdf = h->sdf;
    aa.xb = A.xb.p;
    aa.vb = A.vb.p;
    aa.ub = A.ub.p;
    aa.qb = A.qb.p;
    aa.ksb = A.ksb.p;
    aa.qrb = A.qrb.p;
    // events: per segment record begin/end (record stream) and sweep
    // begin/end (handle stream); then start, record-stream done, end
    std::vector<hipEvent_t> ev(4 * (size_t)nseg + 3, nullptr);
    for (auto& x : ev)
        if ((e = hipEventCreate(&x)) != hipSuccess) {
            rc = h->hip_fail(e, "hipEventCreate");
            break;
        }
    auto rec_begin = [&](int seg) { return ev[4 * seg]; };
    auto rec_end = [&](int seg) { return ev[4 * seg + 1]; };
    auto sweep_begin = [&](int seg) { return ev[4 * seg + 2]; };
    auto sweep_end = [&](int seg) { return ev[4 * seg + 3]; };
    hipEvent_t ev_start = ev[4 * nseg], ev_rec_done = ev[4 * nseg + 1], ev_end = ev[4 * nseg + 2];
    // Segment j's tape lives in slot (nseg - 1 - j) % 2: the last segment in
    // slot 0, where the forward pass recorded it.  Re-recording segment j - 1
    // (record stream) overlaps the sweep of segment j (handle stream); it
    // waits for the sweep of segment j + 1, the last user of its slot.
    auto slot_of = [&](int seg) { return (size_t)((nseg - 1 - seg) % 2); };
    std::vector<std::vector<float>> hs(nseg), lhs(nseg), ts(nseg);
    std::vector<bool> rerecorded(nseg, false);
    if (rc == KSCR_OK) {
        (void)hipEventRecord(ev_start, ms);
        if ((e = hipStreamWaitEvent(rs, ev_start, 0)) != hipSuccess) rc = h->hip_fail(e, "hipStreamWaitEvent");
    }
    for (int seg = nseg - 1; seg >= 0 && rc == KSCR_OK; seg--) {
        const int s0 = seg * A.every, s1 = std::min(A.nsteps, s0 + A.every);
        if (seg == nseg - 1 && A.last_recorded) {
            hs[seg] = A.last_h_of;  // recorded by the forward pass
            lhs[seg] = A.last_lh_of;
            ts[seg] = A.last_t_of;
        } else {
            rerecorded[seg] = true;
            hs[seg].assign(A.every, 0.f);
            lhs[seg].assign(A.every, 0.f);
            ts[seg].assign(A.every, 0.f);
            if (seg + 2 <= nseg - 1 && (e = hipStreamWaitEvent(rs, sweep_end(seg + 2), 0)) != hipSuccess) {
                rc = h->hip_fail(e, "hipStreamWaitEvent");
                break;
            }
            (void)hipEventRecord(rec_begin(seg), rs);
            if ((e = checkpoint_copy(h, (size_t)seg, false, rs)) != hipSuccess) {
                rc = h->hip_fail(e, "restore checkpoint");
                break;
            }
            set_clock(h, A.clocks[seg]);
            for (int s = s0; s < s1; s++) {
                const StepArgs sa = step_args(h, A.dt, A.with_sdf);
                hs[seg][s - s0] = sa.h;
                lhs[seg][s - s0] = sa.last_h;
                ts[seg][s - s0] = sa.time;
                if ((e = launch_step_record(sa, tape_of(h, slot_of(seg), (size_t)(s - s0)), h->d_packs.p,
                                            h->num_packs, h->block, rs)) != hipSuccess) {
                    rc = h->hip_fail(e, "record kernel launch");
                    break;
                }
            }
            (void)hipEventRecord(rec_end(seg), rs);
            if (rc == KSCR_OK && (e = hipStreamWaitEvent(ms, rec_end(seg), 0)) != hipSuccess)
                rc = h->hip_fail(e, "hipStreamWaitEvent");
        }
        (void)hipEventRecord(sweep_begin(seg), ms);
        for (int s = s1 - 1; s >= s0 && rc == KSCR_OK; s--) {
            aa.tape = tape_of(h, slot_of(seg), (size_t)(s - s0));
            aa.h = hs[seg][s - s0];
            aa.last_h = lhs[seg][s - s0];
            aa.time = ts[seg][s - s0];
            if ((e = launch_step_adjoint(aa, h->d_packs.p, h->num_packs, h->block, ms)) != hipSuccess)
                rc = h->hip_fail(e, "adjoint kernel launch");
        }
        (void)hipEventRecord(sweep_end(seg), ms);
    }
    // the record stream's work (it writes the state arrays) ends before the
    // end state is restored
    (void)hipEventRecord(ev_rec_done, rs);
    (void)hipStreamWaitEvent(ms, ev_rec_done, 0);
    if (nseg > 1) A.last_recorded = false;  // earlier segments reused the tape
    if ((e = checkpoint_copy(h, end_slot, false, ms)) != hipSuccess && rc == KSCR_OK) rc = h->hip_fail(e, "restore end state");
    set_clock(h, end_clock);
    (void)hipEventRecord(ev_end, ms);
    if (rc == KSCR_OK &&
        ((e = hipMemcpyAsync(k_stretch_bar, A.ksb.p, n * 4, hipMemcpyDeviceToHost, ms)) != hipSuccess ||
         (e = hipMemcpyAsync(q_rest_bar, A.qrb.p, n * 16, hipMemcpyDeviceToHost, ms)) != hipSuccess))
        rc = h->hip_fail(e, "download gradient");
    if ((e = hipStreamSynchronize(ms)) != hipSuccess && rc == KSCR_OK) rc = h->hip_fail(e, "hipStreamSynchronize");
    if ((e = hipStreamSynchronize(rs)) != hipSuccess && rc == KSCR_OK) rc = h->hip_fail(e, "hipStreamSynchronize");
    A.record_ms = A.sweep_ms = A.backward_ms = 0;
    if (rc == KSCR_OK) {
        for (int seg = 0; seg < nseg; seg++) {
            float r = 0, w = 0;
            if (rerecorded[seg]) (void)hipEventElapsedTime(&r, rec_begin(seg), rec_end(seg));
            (void)hipEventElapsedTime(&w, sweep_begin(seg), sweep_end(seg));
            A.record_ms += r;
            A.sweep_ms += w;
        }
        (void)hipEventElapsedTime(&A.backward_ms, ev_start, ev_end);
    }
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    return rc;
}

int kscr_adjoint_last_time(const kscr_handle* h, float* record_ms, float* sweep_ms, float* backward_ms) {
    if (!h) return KSCR_ERR_INVALID_ARGUMENT;
    if (record_ms) *record_ms = h->adj.record_ms;
    if (sweep_ms) *sweep_ms = h->adj.sweep_ms;
    if (backward_ms) *backward_ms = h->adj.backward_ms;
    return KSCR_OK;
}

}  // extern "C"
