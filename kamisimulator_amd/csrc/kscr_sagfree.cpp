// kscr_sagfree.cpp -- SagFree initialisation of a groom (host code).
//
// Sim::SagFree / SingleSagFree (YB/YBSimulator.cu:480-676) runs once, on the
// host, between setYoung / root pinning and upload
// (SCR/StableCosseratRodsSolver.cpp:203-209).  It is sequential along each
// strand (tip to root, then root to tip) and independent across strands, so
// strands are split over host threads; each strand sees exactly the
// reference's sequence of float and double operations (glm host arithmetic:
// no FMA, glm's dot / length / normalize forms; C library sqrt, acos, sinf,
// cosf; fitAngle's long-double fmod), so the result does not depend on the
// thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "kscr_device.h"
#include "kscr_math.h"

namespace kscr {
namespace {

// glm vector arithmetic as the reference's host code evaluates it
struct HV {
    float x, y, z;
};
inline HV operator+(HV a, HV b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline HV operator-(HV a, HV b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline HV operator*(float s, HV a) { return {s * a.x, s * a.y, s * a.z}; }
inline float hdot(HV a, HV b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline HV hcross(HV a, HV b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline float hlength(HV a) { return std::sqrt(hdot(a, a)); }
inline HV hnormalize(HV a) { return (1.0f / std::sqrt(hdot(a, a))) * a; }

struct HQ {
    HV v;
    float w;
};
inline float hdot4(HQ a, HQ b) { return (a.v.x * b.v.x + a.v.y * b.v.y) + (a.v.z * b.v.z + a.w * b.w); }
// Rotor product (KE/includes/modules/Rotor.h:210-212)
inline HQ hmul(HQ a, HQ b) { return {a.w * b.v + b.w * a.v + hcross(a.v, b.v), a.w * b.w - hdot(a.v, b.v)}; }
// Rotor::rotate (Rotor.h:114-123)
inline HV hrotate(HQ r, HV v) {
    const HV a = r.w * v + hcross(r.v, v);
    const float c = hdot(v, r.v);
    return r.w * a + hcross(r.v, a) + c * r.v;
}
// Rotor::getD3 (Rotor.h:206): the unit segment axis of a frame
inline HV axis_of(HQ r) { return hnormalize(hrotate(r, HV{1.0f, 0.0f, 0.0f})); }
inline HQ to_hq(f4 q) { return {{q.x, q.y, q.z}, q.w}; }
inline f4 to_f4(HQ q) { return f4{q.v.x, q.v.y, q.v.z, q.w}; }

// fitAngle (YBSimulator.cu:59-78): angle wrapped into [-pi, pi)
double wrap_angle(double a) {
    const double pi = 3.141592653589793238462643;
    const double two_pi = 6.2831853071795864769252867665590057683943;
    auto fmod_ld = [](double x, double y) -> double {
        const long double lx = x, ly = y;
        if (ly == 0) return 0;
        return (double)(lx - (long long)(lx / ly) * ly);
    };
    return fmod_ld(fmod_ld(a + pi, two_pi) + two_pi, two_pi) - pi;
}
double sgn(double x) { return (double)((0 < x) - (x < 0)); }

struct Groom {
    const float* pos;
    const float* inv_mass;
    float* l_rest;
    const float* k_stretch;
    f4* q;
    f4* q_rest;
    HV g;
};
static_assert(sizeof(f4) == 4 * sizeof(float), "rotors are 4 packed floats");

// The mass a vertex contributes to the weight carried by a segment; a pinned
// vertex counts with the mass its neighbour would have over its own length
// (YBSimulator.cu:500-504, :552-556, :566-569).
float add_half_mass(float total, const Groom& G, int v, int nb, bool subtract) {
    double m = G.inv_mass[v] != 0 ? 0.5 / (double)G.inv_mass[v]
                                  : 0.5 * (double)(G.l_rest[v] / G.l_rest[nb]) / (double)G.inv_mass[nb];
    return (float)(subtract ? (double)total - m : (double)total + m);
}

void sag_free_strand(const Groom& G, int root, int end) {
    const int len = end - root;
    std::vector<HV> carried(len, HV{0, 0, 0});  // stretch force of the segment below, along its axis

    // ---- 1. rest lengths: the stretch force of each segment carries the
    // weight below it (SingleSagFree, YBSimulator.cu:491-547)
    float weight_mass = 1 / G.inv_mass[end - 1];
    for (int seg = end - 2; seg >= root; --seg) {
        const HV axis = axis_of(to_hq(G.q[seg]));
        const float k = G.k_stretch[seg];
        const float old_len = G.l_rest[seg];
        weight_mass = add_half_mass(weight_mass, G, seg, seg + 1, false);
        const double f_s = -(double)(hdot(axis, G.g) * weight_mass);
        float new_len;
        if (f_s == 0) {
            new_len = old_len;
        } else if (1 - 4 * f_s * (double)old_len / (double)k < 0) {
            new_len = 1.5f * old_len;
        } else {
            // rest length l solving (f_s / k) l^2 - l + old_len = 0, the root
            // nearer old_len
            const double a = f_s / (double)k, inv_a = (double)k / f_s;
            const double root_disc = std::sqrt(1 - (double)(4 * old_len) * a);
            const double l1 = (1 + root_disc) * 0.5 * inv_a;
            const double l2 = (1 - root_disc) * 0.5 * inv_a;
            new_len = (float)(std::fabs((double)old_len - l1) >= std::fabs((double)old_len - l2) ? l2 : l1);
        }
        G.l_rest[seg] = std::min(std::max(new_len, 1e-6f), 1.5f * old_len);
        if (seg > root) carried[seg - root - 1] = (float)-f_s * axis;
        weight_mass = (float)((double)weight_mass + 0.5 / (double)G.inv_mass[seg]);
    }

    // ---- 2. frames and rest rotors: bending carries the torque of the weight
    // beyond each vertex (YBSimulator.cu:549-674)
    const float near_angle = 1e-4f;
    const double half_pi = 1.5707963267948966192313216916397;
    float beyond = 0;
    for (int v = root + 1; v < end; ++v)
        beyond += G.inv_mass[v] != 0 ? 1 / G.inv_mass[v] : (G.l_rest[v] / G.l_rest[v - 1]) / G.inv_mass[v - 1];
    HQ turn{{0, 0, 0}, 1};  // rotation accumulated from the root
    for (int v = root + 1; v < end - 1; ++v) {
        const HQ frame = hmul(turn, to_hq(G.q[v]));
        beyond = add_half_mass(beyond, G, v, v + 1, true);
        const HV axis = axis_of(frame);
        const HV load = carried[v - root] + beyond * G.g;
        if (hdot(load, axis) < 1 - near_angle) {
            const float k_bend = std::sqrt(hdot4(to_hq(G.q_rest[v]), to_hq(G.q_rest[v])));
            const HV normal = hnormalize(hcross(axis, load));
            const HV lever = hcross(normal, axis);
            const HV seg = HV{G.pos[3 * v + 3], G.pos[3 * v + 4], G.pos[3 * v + 5]} - HV{G.pos[3 * v], G.pos[3 * v + 1], G.pos[3 * v + 2]};
            const float t_over_k = hdot(lever, load) * hlength(seg) / k_bend;
            // admissible turn: towards -load, at most the angle to it (within
            // near_angle)
            const float l_load = hlength(load);
            const HV dir{load.x / l_load, load.y / l_load, load.z / l_load};
            const double to_load = std::acos(((double)axis.x * -(double)dir.x + (double)axis.y * -(double)dir.y) +
                                             (double)axis.z * -(double)dir.z);
            const double lim1 = wrap_angle(to_load + (double)near_angle), lim2 = wrap_angle(to_load - (double)near_angle);
            const double lim = std::fabs(lim1) >= std::fabs(lim2) ? lim2 : lim1;
            const float lo = lim >= 0 ? 0.f : (float)lim, hi = lim >= 0 ? (float)lim : 0.f;
            double theta = 0;
            if (std::fabs(t_over_k) > 2) {
                theta = lim;
            } else {
                const double tt = 0.5 * (double)t_over_k * (double)t_over_k;
                const double c1 = wrap_angle(2 * std::acos(1 - tt)), c2 = wrap_angle(2 * std::acos(-1 + tt));
                double pick1 = -sgn(lim) * 100, pick2 = -sgn(lim) * 100;
                if (std::fabs(c1) <= half_pi) pick1 = c1;
                if (std::fabs(c2) > half_pi) pick2 = c2;
                const bool ok1 = sgn(pick1) == sgn(lim), ok2 = sgn(pick2) == sgn(lim);
                if (ok1 && ok2)
                    theta = std::min(pick1, pick2);
                else {
                    if (ok1) theta = pick1;
                    if (ok2) theta = pick2;
                }
            }
            theta = std::min(std::max(theta, (double)lo), (double)hi);
            // Rotor::angleAxis (Rotor.h:51-54), in float
            float half = (float)theta;
            half = (float)(half * 0.5);
            const HQ step{std::sin(half) * normal, std::cos(half)};
            turn = hmul(turn, step);
            const HQ moved = hmul(turn, to_hq(G.q[v]));
            G.q[v] = to_f4(moved);
            const HQ prev = to_hq(G.q[v - 1]);
            const HQ rel = hmul(HQ{HV{-prev.v.x, -prev.v.y, -prev.v.z}, prev.w}, moved);
            const float inv = 1.0f / std::sqrt(hdot4(rel, rel));
            const float stiff = std::sqrt(hdot4(to_hq(G.q_rest[v - 1]), to_hq(G.q_rest[v - 1])));
            G.q_rest[v - 1] = f4{rel.v.x * inv * stiff, rel.v.y * inv * stiff, rel.v.z * inv * stiff, rel.w * inv * stiff};
        }
        beyond = (float)((double)beyond - 0.5 / (double)G.inv_mass[v]);
    }
}

}  // namespace

void sag_free_groom(const float* pos, const float* inv_mass, float* l_rest, const float* k_stretch, float* q,
                    float* q_rest, const float gravity[3], const int* roots, int num_roots, int num_verts, int iterations) {
    const Groom G{pos,  inv_mass, l_rest, k_stretch, reinterpret_cast<f4*>(q), reinterpret_cast<f4*>(q_rest),
                  HV{gravity[0], gravity[1], gravity[2]}};
    const int workers = (int)std::max(1u, std::min(std::thread::hardware_concurrency(), 32u));
    for (int it = 0; it < iterations; ++it) {
        auto run = [&](int w) {
            for (int s = w; s < num_roots; s += workers) sag_free_strand(G, roots[s], s + 1 < num_roots ? roots[s + 1] : num_verts);
        };
        if (num_roots < 64 || workers == 1) {
            run(0);
            for (int w = 1; w < workers; ++w) run(w);
            continue;
        }
        std::vector<std::thread> pool;
        for (int w = 0; w < workers; ++w) pool.emplace_back(run, w);
        for (auto& t : pool) t.join();
    }
}

}  // namespace kscr
