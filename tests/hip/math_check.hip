// math_check.hip -- device checks of the exactness-critical arithmetic in
// gmu-water-simulation_amd/csrc/sph_math.h against the IEEE/compiler results it
// replaces. Test infrastructure: built into tests/hip/libsph_mathcheck.so and
// called by tests/test_gpu_math.py.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../gmu-water-simulation_amd/csrc/sph_math.h"
#include "sph_pow_tables.h"   // the library's generated tables (gmu-water-simulation_amd/build)

namespace {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Per-workgroup mismatch count and smallest failing index, then one global
// atomic per workgroup.
__device__ void report(uint32_t local, uint64_t first_local, unsigned long long *bad, unsigned long long *first)
{
    __shared__ uint32_t s_cnt;
    __shared__ unsigned long long s_first;
    if (threadIdx.x == 0) {
        s_cnt = 0;
        s_first = ~0ull;
    }
    __syncthreads();
    if (local) {
        atomicAdd(&s_cnt, local);
        atomicMin(&s_first, (unsigned long long) first_local);
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_cnt) {
        atomicAdd(bad, (unsigned long long) s_cnt);
        atomicMin(first, s_first);
    }
}

// For the float bit patterns lo .. lo+count-1 of r^2: sqrt_of_float((double) f)
// == sqrt((double) f), and (f > 0) the reciprocal of the pressure gradient's
// divisor b = (float) sqrt, formed from the same rsq estimate as in the force
// kernels (rcp_from), is within 2^-52 of 1/b: |1 - b * rb| <= 2^-52.
__global__ __launch_bounds__(256) void k_sqrt(uint32_t lo, uint32_t count, unsigned long long *bad,
                                              unsigned long long *first)
{
    uint32_t local = 0;
    uint64_t fl = ~0ull;
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < count; t += gridDim.x * 256u) {
        const double x = (double) __uint_as_float(lo + t);
        const double y = __builtin_amdgcn_rsq(x);
        const double a = sph::sqrt_with_rsq(x, y), b = sqrt(x);
        bool ok = __double_as_longlong(a) == __double_as_longlong(b) &&
                  __double_as_longlong(sph::sqrt_of_float(x)) == __double_as_longlong(b);
        if (x > 0.0) {
            const double bf = (double) (float) a;
            const double e = __builtin_fma(-bf, sph::rcp_from(bf, y), 1.0);
            ok = ok && __builtin_fabs(e) <= 0x1p-52;
        }
        if (!ok) {
            ++local;
            fl = fl < (uint64_t) (lo + t) ? fl : (uint64_t) (lo + t);
        }
    }
    report(local, fl, bad, first);
}

// Random float with a random sign and a binary exponent in [emin, emax]
// (subnormal below -126).
__device__ __forceinline__ float random_float(uint64_t r, int emin, int emax)
{
    const int e = emin + (int) ((r >> 40) % (uint64_t) (emax - emin + 1));
    const float m = 1.0f + (float) (uint32_t) (r & 0x7FFFFF) * 0x1p-23f;   // [1, 2)
    const float v = ldexpf(m, e);
    return (r >> 63) ? -v : v;
}

// The two quotient helpers against the IEEE float division, on n random cases:
// div3_small_divisor for the divisors of the pressure gradient (b = (float)
// sqrt(r^2) for a random float r^2 in [2^-40, h^2], with the force kernels' rsq
// estimate) and div3_rcp (1/b in double) for divisors in [2^-20, 2^20], with
// numerators from 2^-149 to 2^10 so that many quotients round to subnormals, to
// zero and to the smallest normals.
__global__ __launch_bounds__(256) void k_div(uint64_t seed, uint64_t n, unsigned long long *bad,
                                             unsigned long long *first)
{
    uint32_t local = 0;
    uint64_t fl = ~0ull;
    for (uint64_t t = blockIdx.x * 256ull + threadIdx.x; t < n; t += (uint64_t) gridDim.x * 256ull) {
        const uint64_t s = seed + t * 8ull * 0x9E3779B97F4A7C15ull;
        const float a0 = random_float(mix64(s + 1), -149, 10);
        const float a1 = random_float(mix64(s + 2), -130, -100);
        const float a2 = (mix64(s + 3) & 15) == 0 ? 0.0f : random_float(mix64(s + 4), -149, -120);
        const float r2s = fabsf(random_float(mix64(s + 5), -40, -9));
        const float r2 = r2s < 0.0457f * 0.0457f ? r2s : 0.0457f * 0.0457f;
        const double x = (double) r2, y = __builtin_amdgcn_rsq(x);
        const float b = (float) sph::sqrt_with_rsq(x, y);
        float q0, q1, q2;
        sph::div3_small_divisor(a0, a1, a2, b, q0, q1, q2, y);
        bool ok = __float_as_uint(q0) == __float_as_uint(a0 / b) && __float_as_uint(q1) == __float_as_uint(a1 / b) &&
                  __float_as_uint(q2) == __float_as_uint(a2 / b);
        const float c = fabsf(random_float(mix64(s + 6), -20, 20));
        const double rc = 1.0 / (double) c;
        sph::div3_rcp(a0, a1, a2, rc, [&]() { return c; }, q0, q1, q2);
        ok = ok && __float_as_uint(q0) == __float_as_uint(a0 / c) && __float_as_uint(q1) == __float_as_uint(a1 / c) &&
             __float_as_uint(q2) == __float_as_uint(a2 / c);
        if (!ok) {
            ++local;
            fl = fl < t ? fl : t;
        }
    }
    report(local, fl, bad, first);
}

// Results for the float r^2 of bit pattern b are hashed with their input index,
// so that a sum over all inputs is an order-free fingerprint of every result bit.
__host__ __device__ __forceinline__ uint64_t fingerprint(double v, uint32_t b)
{
    uint64_t u;
    __builtin_memcpy(&u, &v, 8);
    return mix64(u ^ ((uint64_t) b * 0x9E3779B97F4A7C15ull));
}

// Fingerprint sums over float bit patterns lo .. lo+count-1 of the kernels' two
// library powers, computed on the device as the kernels compute them (the
// correctly rounded power, corrected by the generated tables):
// pow(h^2 - r^2, 3) (Wpoly6, src/CCPUParticleSimulator.cpp:14) and pow(h - r, 2)
// with the kernels' sqrt (WspikyGradient, src/CCPUParticleSimulator.cpp:20-21).
__global__ __launch_bounds__(256) void k_pow_fingerprint(uint32_t lo, uint32_t count, double hsq, double h,
                                                         unsigned long long *sums)
{
    unsigned long long a = 0, b = 0;
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < count; t += gridDim.x * 256u) {
        const uint32_t bits = lo + t;
        const double x = (double) __uint_as_float(bits);
        double c = sph::cube_cr(hsq - x);
        const int d3 = sph::pow_exception(sph::g_pow3_exc, sph::g_pow3_exc_count, bits);
        if (d3) c = sph::next_double(c, d3);
        a += fingerprint(c, bits);
        const double hr = h - sph::sqrt_of_float(x);
        double q = hr * hr;
        const int d2 = sph::pow_exception(sph::g_pow2_exc, sph::g_pow2_exc_count, bits);
        if (d2) q = sph::next_double(q, d2);
        b += fingerprint(q, bits);
    }
    __shared__ unsigned long long sa[256], sb[256];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(sums, sa[0]);
        atomicAdd(sums + 1, sb[0]);
    }
}

// The kernels' density term and spiky factor (cheap power, exact path near a
// float rounding boundary) against the exact path everywhere, on n random cases:
// r^2 a random float in [0, fl(h^2)], the density a random float in [300, 65536]
// (the isolated-particle density is 328).
__global__ __launch_bounds__(256) void k_pow_paths(uint64_t seed, uint64_t n, double hsq, double h, double poly6,
                                                   double spiky, uint32_t hsq_bits, unsigned long long *bad,
                                                   unsigned long long *first)
{
    uint32_t local = 0;
    uint64_t fl = ~0ull;
    for (uint64_t t = blockIdx.x * 256ull + threadIdx.x; t < n; t += (uint64_t) gridDim.x * 256ull) {
        const uint64_t s = seed + t * 4ull * 0x9E3779B97F4A7C15ull;
        const uint32_t rb = (uint32_t) (mix64(s + 1) % ((uint64_t) hsq_bits + 1));
        const float r2 = __uint_as_float(rb);
        const float dens = ldexpf(1.0f + (float) (uint32_t) (mix64(s + 2) & 0x7FFFFF) * 0x1p-23f,
                                  8 + (int) (mix64(s + 3) % 8));   // [256, 65536)
        const double tt = hsq - (double) r2;
        const float fast = sph::density_term(dens, r2, hsq, poly6, sph::g_pow3_exc, sph::g_pow3_exc_count);
        const float exact = (float) ((double) dens +
                                     poly6 * sph::pow3_lib(tt, rb, sph::g_pow3_exc, sph::g_pow3_exc_count));
        const double hr = h - sph::sqrt_of_float((double) r2);
        const float sf = sph::spiky_factor(hr, r2, spiky, sph::g_pow2_exc, sph::g_pow2_exc_count);
        const float se = (float) (spiky * sph::pow2_lib(hr, rb, sph::g_pow2_exc, sph::g_pow2_exc_count));
        if (__float_as_uint(fast) != __float_as_uint(exact) || __float_as_uint(sf) != __float_as_uint(se)) {
            ++local;
            fl = fl < t ? fl : t;
        }
    }
    report(local, fl, bad, first);
}

int run(void (*launch)(unsigned long long *, unsigned long long *), unsigned long long *bad_out,
        unsigned long long *first_out)
{
    unsigned long long *d = nullptr;
    if (hipMalloc(&d, 2 * sizeof(unsigned long long)) != hipSuccess) return 1;
    const unsigned long long init[2] = {0ull, ~0ull};
    int rc = hipMemcpy(d, init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess;
    if (!rc) {
        launch(d, d + 1);
        rc = hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess;
    }
    unsigned long long out[2] = {0, 0};
    if (!rc) rc = hipMemcpy(out, d, sizeof(out), hipMemcpyDeviceToHost) != hipSuccess;
    (void) hipFree(d);
    *bad_out = out[0];
    *first_out = out[1];
    return rc;
}

uint32_t g_lo, g_count;
uint64_t g_seed, g_n;
double g_c[4];

// glibc's pow, called through a pointer so that the compiler cannot replace pow(x, 2.0) by x * x.
double (*volatile g_pow)(double, double) = std::pow;

// The double next to a positive finite x (host side of sph::next_double).
double step_double(double x, int dir)
{
    uint64_t u;
    std::memcpy(&u, &x, 8);
    u += (uint64_t) (int64_t) dir;
    std::memcpy(&x, &u, 8);
    return x;
}

}  // namespace

extern "C" {

// Mismatches of sqrt_of_float over float bit patterns lo..hi (inclusive); 0 on success.
int mc_sqrt(uint32_t lo, uint32_t hi, unsigned long long *bad, unsigned long long *first)
{
    if (hi < lo) return 2;
    g_lo = lo;
    g_count = hi - lo + 1;
    return run([](unsigned long long *b, unsigned long long *f) {
        hipLaunchKernelGGL(k_sqrt, dim3(8192), dim3(256), 0, 0, g_lo, g_count, b, f);
    }, bad, first);
}

// Mismatches of the quotient helpers on n random cases; 0 on success.
int mc_div(uint64_t seed, uint64_t n, unsigned long long *bad, unsigned long long *first)
{
    g_seed = seed;
    g_n = n;
    return run([](unsigned long long *b, unsigned long long *f) {
        hipLaunchKernelGGL(k_div, dim3(8192), dim3(256), 0, 0, g_seed, g_n, b, f);
    }, bad, first);
}

// Host check against the C library's pow for the float r^2 of bit patterns
// lo..hi (inclusive), on `threads` host threads, of the kernels' powers corrected
// by the tables tab3 / tab2 (sorted entries: r^2 bits | 2^31 when pow is the next
// double up): out[0] / out[1] = inputs where the corrected cube_cr(h^2 - r^2) /
// (h - sqrt(r^2))^2 differs from pow(., 3.0) / pow(., 2.0), out[2] / out[3] = the
// first of them (or ~0), out[4] / out[5] = fingerprint sums of the pow results
// themselves, out[6] / out[7] = inputs where the uncorrected power differs.
int mc_pow_host(uint32_t lo, uint32_t hi, double hsq, double h, int threads, const uint32_t *tab3, uint32_t n3,
                const uint32_t *tab2, uint32_t n2, unsigned long long *out)
{
    if (hi < lo || threads < 1 || threads > 256) return 2;
    const uint64_t count = (uint64_t) hi - lo + 1;
    std::vector<unsigned long long> part((size_t) threads * 8, 0);
    std::vector<std::thread> pool;
    // first table entry whose r^2 bits are >= b
    auto lower = [](const uint32_t *tab, uint32_t n, uint32_t b) {
        uint32_t l = 0, r = n;
        while (l < r) {
            const uint32_t m = (l + r) / 2;
            if ((tab[m] & 0x7FFFFFFFu) < b) l = m + 1; else r = m;
        }
        return l;
    };
    for (int w = 0; w < threads; ++w) {
        pool.emplace_back([&, w]() {
            unsigned long long *o = &part[(size_t) w * 8];
            o[2] = o[3] = ~0ull;
            const uint64_t a0 = count * (uint64_t) w / (uint64_t) threads, a1 = count * (uint64_t) (w + 1) / (uint64_t) threads;
            uint32_t k3 = lower(tab3, n3, lo + (uint32_t) a0), k2 = lower(tab2, n2, lo + (uint32_t) a0);
            for (uint64_t q = a0; q < a1; ++q) {
                const uint32_t bits = lo + (uint32_t) q;
                float r2f;
                std::memcpy(&r2f, &bits, 4);
                const double x = (double) r2f, t = hsq - x;
                double c = sph::cube_cr(t);
                const double cp = g_pow(t, 3.0);
                const double hr = h - std::sqrt(x);
                double s = hr * hr;
                const double sp = g_pow(hr, 2.0);
                o[6] += std::memcmp(&c, &cp, 8) != 0;
                o[7] += std::memcmp(&s, &sp, 8) != 0;
                if (k3 < n3 && (tab3[k3] & 0x7FFFFFFFu) == bits) c = step_double(c, (tab3[k3++] >> 31) ? 1 : -1);
                if (k2 < n2 && (tab2[k2] & 0x7FFFFFFFu) == bits) s = step_double(s, (tab2[k2++] >> 31) ? 1 : -1);
                if (std::memcmp(&c, &cp, 8) != 0) {
                    ++o[0];
                    if (o[2] == ~0ull) o[2] = bits;
                }
                if (std::memcmp(&s, &sp, 8) != 0) {
                    ++o[1];
                    if (o[3] == ~0ull) o[3] = bits;
                }
                o[4] += fingerprint(cp, bits);
                o[5] += fingerprint(sp, bits);
            }
        });
    }
    for (auto &t : pool) t.join();
    for (int q = 0; q < 8; ++q) out[q] = (q == 2 || q == 3) ? ~0ull : 0ull;
    for (int w = 0; w < threads; ++w) {
        const unsigned long long *o = &part[(size_t) w * 8];
        for (int q : {0, 1, 4, 5, 6, 7}) out[q] += o[q];
        out[2] = o[2] < out[2] ? o[2] : out[2];
        out[3] = o[3] < out[3] ? o[3] : out[3];
    }
    return 0;
}

// Mismatches of the kernels' cheap-power paths against the exact ones on n random
// cases (constants: hsq, h, poly6, spiky); 0 on success.
int mc_pow_paths(uint64_t seed, uint64_t n, double hsq, double h, double poly6, double spiky,
                 unsigned long long *bad, unsigned long long *first)
{
    g_seed = seed;
    g_n = n;
    g_c[0] = hsq;
    g_c[1] = h;
    g_c[2] = poly6;
    g_c[3] = spiky;
    return run([](unsigned long long *b, unsigned long long *f) {
        float hf = (float) g_c[0];
        uint32_t hb;
        std::memcpy(&hb, &hf, 4);
        hipLaunchKernelGGL(k_pow_paths, dim3(8192), dim3(256), 0, 0, g_seed, g_n, g_c[0], g_c[1], g_c[2], g_c[3], hb,
                           b, f);
    }, bad, first);
}

// Device fingerprint sums (out[0], out[1]) of the same results over lo..hi.
int mc_pow_device(uint32_t lo, uint32_t hi, double hsq, double h, unsigned long long *out)
{
    if (hi < lo) return 2;
    unsigned long long *d = nullptr;
    if (hipMalloc(&d, 2 * sizeof(unsigned long long)) != hipSuccess) return 1;
    const unsigned long long init[2] = {0ull, 0ull};
    int rc = hipMemcpy(d, init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess;
    if (!rc) {
        hipLaunchKernelGGL(k_pow_fingerprint, dim3(8192), dim3(256), 0, 0, lo, hi - lo + 1, hsq, h, d);
        rc = hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess;
    }
    if (!rc) rc = hipMemcpy(out, d, sizeof(init), hipMemcpyDeviceToHost) != hipSuccess;
    (void) hipFree(d);
    return rc;
}

}  // extern "C"
