// gen_pow_tables.hip -- build step of libsph_hip.so (host program, no device
// work): the inputs on which the host's libm pow differs from the kernels' own
// correctly rounded powers, written as device tables for the kernels.
//
// The reference computes Wpoly6 as pow(h^2 - r^2, 3) and WspikyGradient with
// pow(h - r, 2) (src/CCPUParticleSimulator.cpp:14, 21) through the C library.
// glibc's pow is not correctly rounded: its error bound is about 0.52 ulp, so
// where the exact power lies within a few thousandths of an ulp of a rounding
// midpoint it can return the other neighbour. The kernels compute cube_cr(t)
// and hr*hr, which are correctly rounded there; to give the reference's bits
// they need the inputs where pow disagrees. Every accepted pair has
// r^2 = a float in [0, fl(h^2)], so the domain is finite (990,437,133 floats)
// and is enumerated here:
//   pow3: t = (double) fl(h^2) - (double) r^2, pow(t, 3.0) vs cube_cr(t)
//   pow2: hr = (double) h - sqrt((double) r^2), pow(hr, 2.0) vs hr * hr
// An entry is the float bit pattern of r^2 (< 2^30) with bit 31 set when pow's
// result is the next double up (else the next double down). Any other kind of
// difference stops the build.
//
// Output: a header with both tables as __device__ arrays, sorted by r^2, plus
// the C library version and the number of inputs checked.

#include <hip/hip_runtime.h>

#include <gnu/libc-version.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "sph_math.h"

namespace {

// pow through a pointer, so that the compiler cannot rewrite pow(x, 2.0) as x * x.
double (*volatile g_pow)(double, double) = std::pow;

double as_double(uint64_t u)
{
    double d;
    std::memcpy(&d, &u, 8);
    return d;
}

uint64_t bits_of(double d)
{
    uint64_t u;
    std::memcpy(&u, &d, 8);
    return u;
}

// +1 / -1 when `lib` is the next double above / below `own` (both positive
// and finite), 0 when equal, 2 for anything else.
int neighbour(double own, double lib)
{
    if (bits_of(own) == bits_of(lib)) return 0;
    if (!(own > 0.0) || !(lib > 0.0) || !std::isfinite(own) || !std::isfinite(lib)) return 2;
    if (bits_of(lib) == bits_of(own) + 1) return 1;
    if (bits_of(lib) + 1 == bits_of(own)) return -1;
    return 2;
}

struct Found {
    std::vector<uint32_t> p3, p2;
    uint64_t bad = 0;
    uint32_t bad_input = 0;
};

void write_table(FILE *f, const char *name, const std::vector<uint32_t> &v)
{
    std::fprintf(f, "constexpr uint32_t %s_count = %zuu;\n", name, v.size());
    std::fprintf(f, "__device__ const uint32_t %s[%zu] = {", name, v.empty() ? (size_t) 1 : v.size());
    for (size_t i = 0; i < v.size(); ++i) std::fprintf(f, "%s0x%08xu,", i % 8 ? " " : "\n    ", v[i]);
    if (v.empty()) std::fprintf(f, "0u");
    std::fprintf(f, "\n};\n\n");
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s OUTPUT_HEADER\n", argv[0]);
        return 2;
    }
    const float hf = 0.0457f;           // CParticle::h (include/CParticle.h:80)
    volatile float hh = hf;
    const float hsq_f = hh * hh;        // fl(h*h), bits 0x3B08DF0C
    uint32_t hsq_bits;
    std::memcpy(&hsq_bits, &hsq_f, 4);
    const double hsq = (double) hsq_f, h = (double) hf;
    const uint64_t count = (uint64_t) hsq_bits + 1;

    unsigned nthreads = std::thread::hardware_concurrency();
    nthreads = std::max(1u, std::min(nthreads, 64u));
    std::vector<Found> part(nthreads);
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nthreads; ++w) {
        pool.emplace_back([&, w]() {
            Found &o = part[w];
            const uint64_t a0 = count * w / nthreads, a1 = count * (w + 1) / nthreads;
            for (uint64_t q = a0; q < a1; ++q) {
                const uint32_t b = (uint32_t) q;
                float r2f;
                std::memcpy(&r2f, &b, 4);
                const double x = (double) r2f;
                const double t = hsq - x;
                const int n3 = neighbour(sph::cube_cr(t), g_pow(t, 3.0));
                const double hr = h - std::sqrt(x);
                const int n2 = neighbour(hr * hr, g_pow(hr, 2.0));
                if (n3 == 2 || n2 == 2) {
                    if (!o.bad++) o.bad_input = b;
                    continue;
                }
                if (n3) o.p3.push_back(b | (n3 > 0 ? 0x80000000u : 0u));
                if (n2) o.p2.push_back(b | (n2 > 0 ? 0x80000000u : 0u));
            }
        });
    }
    for (auto &t : pool) t.join();
    Found all;
    for (auto &o : part) {
        all.p3.insert(all.p3.end(), o.p3.begin(), o.p3.end());
        all.p2.insert(all.p2.end(), o.p2.begin(), o.p2.end());
        if (o.bad && !all.bad) all.bad_input = o.bad_input;
        all.bad += o.bad;
    }
    if (all.bad) {
        std::fprintf(stderr, "gen_pow_tables: %llu inputs where pow is not a neighbour of the correctly rounded "
                             "power (first r^2 bits 0x%08x)\n", (unsigned long long) all.bad, all.bad_input);
        return 1;
    }
    // the threads cover ascending ranges in order, so the tables are sorted by r^2 bits
    auto key = [](uint32_t e) { return e & 0x7FFFFFFFu; };
    for (auto *v : {&all.p3, &all.p2})
        for (size_t i = 1; i < v->size(); ++i)
            if (key((*v)[i - 1]) >= key((*v)[i])) {
                std::fprintf(stderr, "gen_pow_tables: table not sorted\n");
                return 1;
            }

    const std::string tmp = std::string(argv[1]) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) {
        std::perror(tmp.c_str());
        return 1;
    }
    std::fprintf(f, "// Generated by gen_pow_tables: inputs where this host's pow differs from the\n"
                    "// kernels' correctly rounded powers (C library %s, %llu inputs each).\n"
                    "#pragma once\n#include <stdint.h>\n\nnamespace sph {\n\n",
                 gnu_get_libc_version(), (unsigned long long) count);
    std::fprintf(f, "constexpr const char *kPowTablesLibc = \"glibc %s\";\n", gnu_get_libc_version());
    std::fprintf(f, "constexpr uint32_t kPowTablesInputs = %lluu;\n\n", (unsigned long long) count);
    write_table(f, "g_pow3_exc", all.p3);
    write_table(f, "g_pow2_exc", all.p2);
    std::fprintf(f, "}  // namespace sph\n");
    if (std::fclose(f) != 0 || std::rename(tmp.c_str(), argv[1]) != 0) {
        std::perror(argv[1]);
        return 1;
    }
    std::printf("gen_pow_tables: %zu pow3 and %zu pow2 exceptions in %llu inputs (glibc %s)\n", all.p3.size(),
                all.p2.size(), (unsigned long long) count, gnu_get_libc_version());
    return 0;
}
