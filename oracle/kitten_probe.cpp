// kitten_probe.cpp -- evaluates the reference's header-only math on seeded
// inputs and writes the input/output vectors to a binary file.
//
// TEST INFRASTRUCTURE ONLY.  Built by `make -C oracle ref` against the
// reference's own headers where they lie (Common.h, Rotor.h, SymMat.h with the
// vendored glm; host-only because cuda_runtime.h is absent), output into
// oracle/_ref/.  tests/golden/make_kitten_vectors.py runs it and stores the
// vectors as tests/golden/kitten_vectors.npz, which pins the oracle's
// primitives (tests/test_oracle_primitives.py).
//
// File format: a sequence of sections, each
//   char name[16]; uint32 count; uint32 in_floats; uint32 out_floats;
//   float in[count*in_floats]; float out[count*out_floats]
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "Common.h"
#include "Rotor.h"
#include "SymMat.h"

using namespace Kitten;

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static uint64_t splitmix64() {
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static float uni(float lo, float hi) { return lo + (hi - lo) * (float)((splitmix64() >> 40) * (1.0 / 16777216.0)); }
static glm::vec3 rvec(float s) { return glm::vec3(uni(-s, s), uni(-s, s), uni(-s, s)); }
static glm::vec3 runit() { return glm::normalize(rvec(1.f)); }
static Rotor runit_rotor() { return Rotor(glm::normalize(glm::vec4(uni(-1, 1), uni(-1, 1), uni(-1, 1), uni(-1, 1)))); }

struct Section {
    char name[16];
    uint32_t count, nin, nout;
    std::vector<float> in, out;
};

static void put(std::vector<float>& v, glm::vec3 a) { v.insert(v.end(), {a.x, a.y, a.z}); }
static void put(std::vector<float>& v, glm::vec4 a) { v.insert(v.end(), {a.x, a.y, a.z, a.w}); }

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: kitten_probe out.bin\n");
        return 2;
    }
    std::vector<Section> secs;
    const int N = 512;

    {  // Rotor product (Rotor.h:210-212), unit and scaled (rest rotors carry a stiffness magnitude)
        Section s{"rotor_mul", N, 8, 4};
        for (int i = 0; i < N; i++) {
            Rotor a = runit_rotor(), b = runit_rotor();
            if (i % 3 == 0) b = Rotor(b.v * uni(1e-6f, 10.f));
            put(s.in, a.v), put(s.in, b.v);
            put(s.out, (a * b).v);
        }
        secs.push_back(s);
    }
    {  // Rotor::rotate (Rotor.h:114-123)
        Section s{"rotor_rotate", N, 7, 3};
        for (int i = 0; i < N; i++) {
            Rotor a = runit_rotor();
            glm::vec3 v = rvec(2.f);
            put(s.in, a.v), put(s.in, v);
            put(s.out, a * v);
        }
        secs.push_back(s);
    }
    {  // Rotor::fromTo (Rotor.h:79-88), including antiparallel inputs
        Section s{"rotor_from_to", N, 6, 4};
        for (int i = 0; i < N; i++) {
            glm::vec3 a = (i % 4 == 0) ? glm::vec3(1, 0, 0) : runit();
            glm::vec3 b = (i % 16 == 1) ? -a : runit();
            put(s.in, a), put(s.in, b);
            put(s.out, Rotor::fromTo(a, b).v);
        }
        secs.push_back(s);
    }
    {  // segmentClosestPoints (Common.h:294-312), including parallel and degenerate segments
        Section s{"seg_closest", N, 12, 2};
        for (int i = 0; i < N; i++) {
            glm::vec3 a0 = rvec(1.f), a1 = rvec(1.f), b0 = rvec(1.f), b1 = rvec(1.f);
            if (i % 8 == 1) b1 = b0 + (a1 - a0) * 0.5f;  // parallel
            if (i % 16 == 3) b1 = b0;                    // degenerate second segment
            put(s.in, a0), put(s.in, a1), put(s.in, b0), put(s.in, b1);
            glm::vec2 uv = segmentClosestPoints(a0, a1, b0, b1);
            s.out.insert(s.out.end(), {uv.x, uv.y});
        }
        secs.push_back(s);
    }
    {  // getMorton (Common.h:566-579); inputs stored as floats holding small integers
        Section s{"morton", N, 3, 1};
        for (int i = 0; i < N; i++) {
            ivec3 c((int)(splitmix64() % 1024), (int)(splitmix64() % 1024), (int)(splitmix64() % 1024));
            if (i == 0) c = ivec3(1023);
            s.in.insert(s.in.end(), {(float)c.x, (float)c.y, (float)c.z});
            uint32_t m = getMorton(c);
            float f;
            std::memcpy(&f, &m, 4);
            s.out.push_back(f);
        }
        secs.push_back(s);
    }
    {  // hess3 -> mat3, glm inverse, mat3*vec3 (SymMat.h:66-75, func_matrix.inl:272-294)
        Section s{"sym_solve", N, 9, 3};
        for (int i = 0; i < N; i++) {
            hess3 H(uni(1e-4f, 1e-3f));  // the reference's inertia form: all six entries equal
            float d = uni(1.f, 100.f);
            H.data.diag += glm::vec3(d);
            if (i % 2) H += hess3::outer(rvec(1.f)) * uni(0.f, 5.f);
            glm::vec3 f = rvec(1.f);
            for (int k = 0; k < 6; k++) s.in.push_back(H.dat[k]);
            put(s.in, f);
            put(s.out, inverse((mat3)H) * f);
        }
        secs.push_back(s);
    }
    {  // The orientation update expression of SCR Eq. (18) in the reference's
       // types: normalize((Rotor(v) * Rotor(b) * Rotor(e1)).v + (|v| + |b|) * b)
        Section s{"ori_update", N, 7, 4};
        for (int i = 0; i < N; i++) {
            glm::vec3 v = rvec(1e-3f);
            glm::vec4 b = runit_rotor().v * uni(1e-5f, 1e-3f);
            put(s.in, v), put(s.in, b);
            float lambda = length(v) + length(b);
            put(s.out, normalize((Rotor(v) * Rotor(b) * Rotor(glm::vec3(1, 0, 0))).v + lambda * b));
        }
        secs.push_back(s);
    }

    FILE* f = fopen(argv[1], "wb");
    if (!f) return 1;
    for (auto& s : secs) {
        fwrite(s.name, 1, 16, f);
        fwrite(&s.count, 4, 1, f);
        fwrite(&s.nin, 4, 1, f);
        fwrite(&s.nout, 4, 1, f);
        fwrite(s.in.data(), 4, s.in.size(), f);
        fwrite(s.out.data(), 4, s.out.size(), f);
    }
    fclose(f);
    printf("wrote %zu sections\n", secs.size());
    return 0;
}
