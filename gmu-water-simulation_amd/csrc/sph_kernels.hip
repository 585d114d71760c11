// sph_kernels.hip -- gfx950 kernels of the grid-SPH step.
//
// Every particle-level formula reproduces CCPUParticleSimulator
// (src/CCPUParticleSimulator.cpp) operation for operation: float where the
// reference uses QVector3D/float, double where it uses double, no FMA
// contraction (the library is built with -ffp-contract=off; FMAs appear only
// where written explicitly, in the error-free transformations of cube_cr), and
// IEEE division/sqrt. Neighbour sums run in one fixed canonical order (see
// neighbour loops below), so results are bitwise reproducible and equal to the
// CPU oracle's canonical-order mode.
//
// Data layout: structure-of-arrays FP32 particle buffers in HBM, stable-sorted by
// cell every step, so that the 3 x-adjacent cells of a grid row are one
// contiguous run of slots and a particle's 27-cell neighbourhood is 9 runs.

#include "sph_kernels.h"

#include <climits>

namespace sph {
namespace {

// ---------------------------------------------------------------- helpers

// (int) cast of the reference on x86-64 (cvttsd2si): out-of-range and NaN give INT_MIN.
__device__ __forceinline__ int to_int_x86(double q)
{
    return (q >= -2147483648.0 && q < 2147483648.0) ? (int) q : INT_MIN;
}

__device__ __forceinline__ int clamp_cell(int c, int r)
{
    return c < 0 ? 0 : (c >= r ? r - 1 : c);
}

// Cell hash, src/CCPUParticleSimulator.cpp:46-69 with include/CGrid.h:25-28:
// floor(((double)p + (double)box/2.0) / (double)h), clamped, linearised x-fastest.
__device__ __forceinline__ uint32_t hash_key(float px, float py, float pz, const Scene &s, double h)
{
    const int x = clamp_cell(to_int_x86(floor(((double) px + (double) s.bx / 2.0) / h)), s.rx);
    const int y = clamp_cell(to_int_x86(floor(((double) py + (double) s.by / 2.0) / h)), s.ry);
    const int z = clamp_cell(to_int_x86(floor(((double) pz + (double) s.bz / 2.0) / h)), s.rz);
    return s.coff + (uint32_t) (x + y * s.rx + z * s.rx * s.ry);
}

// t^3 correctly rounded (error-free products via FMA), the value libm's
// pow(t, 3.0) returns (src/CCPUParticleSimulator.cpp:14).
__device__ __forceinline__ double cube_cr(double t)
{
    const double p = t * t;
    const double e = __builtin_fma(t, t, -p);   // t*t = p + e exactly
    const double q = p * t;
    const double f = __builtin_fma(p, t, -q);   // p*t = q + f exactly
    const double c = __builtin_fma(e, t, f);    // f + e*t
    return q + c;
}

// CCollisionGeometry::inverseBoundingBoxBounce (src/CCollisionGeometry.cpp:117-132)
// over the six walls of sBoundingBox in order left, bottom, back, right, top, front
// (include/CCollisionGeometry.h:105-118), positioned at -+box/2.
__device__ __forceinline__ void wall_accel(float px, float py, float pz, float vx, float vy, float vz,
                                           const Scene &s, const Consts &k, float &ax, float &ay,
                                           float &az)
{
    const float mnx = -s.bx / 2.0f, mny = -s.by / 2.0f, mnz = -s.bz / 2.0f;
    const float mxx = s.bx / 2.0f, mxy = s.by / 2.0f, mxz = s.bz / 2.0f;
    const float nrm[6][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const float wp[6][3] = {{mnx, 0, 0}, {0, mny, 0}, {0, 0, mnz}, {mxx, 0, 0}, {0, mxy, 0}, {0, 0, mxz}};
    float accx = 0.0f, accy = 0.0f, accz = 0.0f;
#pragma unroll
    for (int w = 0; w < 6; ++w) {
        const float ix = nrm[w][0] * -1.0f, iy = nrm[w][1] * -1.0f, iz = nrm[w][2] * -1.0f;
        const float dot = (wp[w][0] - px) * ix + (wp[w][1] - py) * iy + (wp[w][2] - pz) * iz;
        const double d = (double) dot + 0.01;
        if (d > 0.0) {
            const float df = (float) d;
            const float t1x = (ix * k.wall_k_f) * df, t1y = (iy * k.wall_k_f) * df, t1z = (iz * k.wall_k_f) * df;
            const float vdot = vx * ix + vy * iy + vz * iz;
            const float fac = (float) (kWallDamping * (double) vdot);
            accx = accx + t1x; accy = accy + t1y; accz = accz + t1z;
            accx = accx + ix * fac; accy = accy + iy * fac; accz = accz + iz * fac;
        }
    }
    ax = ax + accx;
    ay = ay + accy;
    az = az + accz;
}

// Blocks are dealt round-robin to the 8 XCDs; remap so that each XCD (own L2)
// walks one contiguous range of slots, whose neighbourhoods overlap.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t g)
{
    const uint32_t q = g >> 3, r = g & 7u, x = b & 7u, l = b >> 3;
    return x * q + (x < r ? x : r) + l;
}

__device__ __forceinline__ void cell_xyz(uint32_t key, const Scene &s, int &x, int &y, int &z)
{
    const uint32_t c = key - s.coff;
    x = (int) (c % (uint32_t) s.rx);
    const uint32_t rem = c / (uint32_t) s.rx;
    y = (int) (rem % (uint32_t) s.ry);
    z = (int) (rem / (uint32_t) s.ry);
}

// ---------------------------------------------------------------- grid

__global__ void k_hash(const float *__restrict__ px, const float *__restrict__ py,
                       const float *__restrict__ pz, uint32_t *__restrict__ key, Scene s, double h)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    const uint32_t g = s.poff + i;
    key[g] = hash_key(px[g], py[g], pz[g], s, h);
}

// Digit histogram of one tile, stored digit-major: hist[d * nblocks + block].
__global__ __launch_bounds__(kSortThreads) void k_radix_count(const uint32_t *__restrict__ keys, uint32_t n,
                                                              int shift, uint32_t *__restrict__ hist,
                                                              uint32_t nblocks)
{
    __shared__ uint32_t h[kRadixBuckets];
    const int t = threadIdx.x;
    for (int d = t; d < kRadixBuckets; d += kSortThreads) h[d] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * (uint32_t) kSortTile;
#pragma unroll 4
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t idx = base + (uint32_t) (k * kSortThreads + t);
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & (kRadixBuckets - 1)], 1u);
    }
    __syncthreads();
    for (int d = t; d < kRadixBuckets; d += kSortThreads) hist[(uint32_t) d * nblocks + blockIdx.x] = h[d];
}

// In-place exclusive scan of m uint32 by one 1024-thread workgroup, 4096 per tile:
// wave64 shuffle scans, then a scan of the 16 wave totals.
__global__ __launch_bounds__(1024) void k_scan_exclusive(uint32_t *__restrict__ data, uint32_t m)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry_s;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < m; base += 4096u) {
        const uint32_t idx = base + (uint32_t) t * 4u;
        uint32_t v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = (idx + q < m) ? data[idx + q] : 0u;
        const uint32_t s = v[0] + v[1] + v[2] + v[3];
        uint32_t x = s;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        if (w == 0) {
            const uint32_t ws = lane < 16 ? wsum[lane] : 0u;
            uint32_t xs = ws;
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                const uint32_t y = __shfl_up(xs, d, 64);
                if (lane >= d) xs += y;
            }
            if (lane < 16) wsum[lane] = xs - ws;
        }
        __syncthreads();
        uint32_t run = carry + wsum[w] + (x - s);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (idx + q < m) data[idx + q] = run;
            run += v[q];
        }
        if (t == 1023) carry_s = run;
        __syncthreads();
        carry = carry_s;
        __syncthreads();
    }
}

// Stable scatter of one LSD radix pass. Items of a tile are taken in slot order
// (round k covers slots tile + k*256 .. +255); equal digits inside a wave are
// matched with 8 ballots, ranked by popcount of the lower lanes, and offset by
// the counts of lower waves and earlier rounds.
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in, uint32_t *__restrict__ keys_out,
    uint32_t *__restrict__ vals_out, uint32_t n, int shift, const uint32_t *__restrict__ hist, uint32_t nblocks)
{
    constexpr int W = kSortThreads / 64;
    __shared__ uint32_t base_s[kRadixBuckets];
    __shared__ uint32_t wcnt[W][kRadixBuckets];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int d = t; d < kRadixBuckets; d += kSortThreads) base_s[d] = hist[(uint32_t) d * nblocks + blockIdx.x];
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t tile = blockIdx.x * (uint32_t) kSortTile;
    for (int k = 0; k < kSortItems; ++k) {
        for (int d = t; d < kRadixBuckets; d += kSortThreads)
#pragma unroll
            for (int ww = 0; ww < W; ++ww) wcnt[ww][d] = 0;
        __syncthreads();
        const uint32_t idx = tile + (uint32_t) (k * kSortThreads + t);
        const bool valid = idx < n;
        const uint32_t key = valid ? keys_in[idx] : 0u;
        const uint32_t val = valid ? (vals_in ? vals_in[idx] : idx) : 0u;
        const uint32_t dig = (key >> shift) & (kRadixBuckets - 1);
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (dig >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t) __popcll(m & below);
        if (valid && rank == 0) wcnt[w][dig] = (uint32_t) __popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pos = base_s[dig] + rank;
            for (int ww = 0; ww < w; ++ww) pos += wcnt[ww][dig];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        for (int d = t; d < kRadixBuckets; d += kSortThreads) {
            uint32_t add = 0;
#pragma unroll
            for (int ww = 0; ww < W; ++ww) add += wcnt[ww][d];
            base_s[d] += add;
        }
    }
}

__global__ void k_cell_start(const uint32_t *__restrict__ keys, uint32_t n, uint32_t ncells,
                             uint32_t *__restrict__ start)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncells) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < c) lo = mid + 1; else hi = mid;
    }
    start[c] = lo;
}

__global__ void k_reorder(ParticleSoA src, ParticleSoA dst, const uint32_t *__restrict__ perm, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = perm[i];
    dst.px[i] = src.px[j];
    dst.py[i] = src.py[j];
    dst.pz[i] = src.pz[j];
    dst.vx[i] = src.vx[j];
    dst.vy[i] = src.vy[j];
    dst.vz[i] = src.vz[j];
    dst.id[i] = src.id[j];
}

// ---------------------------------------------------------------- neighbours
//
// Canonical neighbour order: rows z-1..z+1 (outer), y-1..y+1, and within a row
// the contiguous run of cells max(x-1,0)..min(x+1,X-1) in slot order. Rows off
// the grid are skipped, as the reference's continue/break bounds tests skip
// them (src/CCPUParticleSimulator.cpp:108-118).

constexpr int kNeighbourThreads = 256;

// updateDensityPressure, src/CCPUParticleSimulator.cpp:93-141.
__global__ __launch_bounds__(kNeighbourThreads) void k_density(ParticleSoA p, const uint32_t *__restrict__ cell,
                                                               const uint32_t *__restrict__ start,
                                                               float *__restrict__ rho, double *__restrict__ prr,
                                                               uint32_t *__restrict__ ncount, Scene s, Consts k)
{
    const uint32_t li = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (li >= s.n) return;
    const uint32_t i = s.poff + li;
    const float xi = p.px[i], yi = p.py[i], zi = p.pz[i];
    int x, y, z;
    cell_xyz(cell[i], s, x, y, z);
    const int xlo = x > 0 ? x - 1 : 0, xhi = x + 1 < s.rx ? x + 1 : s.rx - 1;
    float density = 0.0f;
    uint32_t nc = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        const int zz = z + dz;
        if (zz < 0 || zz >= s.rz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int yy = y + dy;
            if (yy < 0 || yy >= s.ry) continue;
            const uint32_t row = s.coff + (uint32_t) (yy * s.rx + zz * s.rx * s.ry);
            const uint32_t a = start[row + xlo], b = start[row + xhi + 1];
            for (uint32_t j = a; j < b; ++j) {
                const float dx = xi - p.px[j], dyv = yi - p.py[j], dzv = zi - p.pz[j];
                const float r2 = dx * dx + dyv * dyv + dzv * dzv;
                if (r2 <= k.hsq_f) {
                    const double wpoly = k.poly6 * cube_cr(k.hsq - (double) r2);
                    density = (float) ((double) density + wpoly);
                    ++nc;
                }
            }
        }
    }
    density = density * k.mass;
    const float pressure = k.gas_k * (density - k.rho0);
    rho[i] = density;
    // p / pow(rho, 2) in double (src/CCPUParticleSimulator.cpp:181); rho*rho is exact in double.
    prr[i] = (double) pressure / ((double) density * (double) density);
    if (ncount) ncount[i] = nc;
}

// Pressure + viscosity sums of updateForces (src/CCPUParticleSimulator.cpp:155-193)
// followed by the acceleration (f_p + f_v + f_g) / rho.
__device__ __forceinline__ void sph_accel(const ParticleSoA &p, const uint32_t *__restrict__ cell,
                                          const uint32_t *__restrict__ start, const float *__restrict__ rho,
                                          const double *__restrict__ prr, uint32_t i, const Scene &s,
                                          const Consts &k, float xi, float yi, float zi, float vxi, float vyi,
                                          float vzi, float &ax, float &ay, float &az)
{
    int x, y, z;
    cell_xyz(cell[i], s, x, y, z);
    const int xlo = x > 0 ? x - 1 : 0, xhi = x + 1 < s.rx ? x + 1 : s.rx - 1;
    const float rhoi = rho[i];
    const double ai = prr[i];
    float fpx = 0.0f, fpy = 0.0f, fpz = 0.0f, fvx = 0.0f, fvy = 0.0f, fvz = 0.0f;
    for (int dz = -1; dz <= 1; ++dz) {
        const int zz = z + dz;
        if (zz < 0 || zz >= s.rz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int yy = y + dy;
            if (yy < 0 || yy >= s.ry) continue;
            const uint32_t row = s.coff + (uint32_t) (yy * s.rx + zz * s.rx * s.ry);
            const uint32_t a = start[row + xlo], b = start[row + xhi + 1];
            for (uint32_t j = a; j < b; ++j) {
                const float dx = xi - p.px[j], dyv = yi - p.py[j], dzv = zi - p.pz[j];
                const float r2 = dx * dx + dyv * dyv + dzv * dzv;
                if (r2 <= k.hsq_f && j != i) {
                    // WspikyGradient / WviscosityLaplacian, src/CCPUParticleSimulator.cpp:17-30
                    const double r = sqrt((double) r2);
                    const double hr = k.h - r;
                    const float sg = (float) (k.spiky * (hr * hr));
                    const float rf = (float) r;
                    const float gx = (dx * sg) / rf, gy = (dyv * sg) / rf, gz = (dzv * sg) / rf;
                    const float pf = (float) (ai + prr[j]);
                    fpx = fpx + gx * pf;
                    fpy = fpy + gy * pf;
                    fpz = fpz + gz * pf;
                    const float lap = (float) (k.visc * hr);
                    const float rj = rho[j];
                    fvx = fvx + ((p.vx[j] - vxi) * lap) / rj;
                    fvy = fvy + ((p.vy[j] - vyi) * lap) / rj;
                    fvz = fvz + ((p.vz[j] - vzi) * lap) / rj;
                }
            }
        }
    }
    const float fac = k.neg_mass * rhoi;
    fpx = fpx * fac; fpy = fpy * fac; fpz = fpz * fac;
    fvx = fvx * k.mu_mass; fvy = fvy * k.mu_mass; fvz = fvz * k.mu_mass;
    const float fgx = s.gx * rhoi, fgy = s.gy * rhoi, fgz = s.gz * rhoi;
    ax = ((fpx + fvx) + fgx) / rhoi;
    ay = ((fpy + fvy) + fgy) / rhoi;
    az = ((fpz + fvz) + fgz) / rhoi;
}

// integrate, src/CCPUParticleSimulator.cpp:220-224.
__device__ __forceinline__ void integrate_one(float xi, float yi, float zi, float vxi, float vyi, float vzi,
                                              float ax, float ay, float az, float dt, float &nx, float &ny,
                                              float &nz, float &nvx, float &nvy, float &nvz)
{
    nx = (xi + vxi * dt) + (ax * dt) * dt;
    ny = (yi + vyi * dt) + (ay * dt) * dt;
    nz = (zi + vzi * dt) + (az * dt) * dt;
    nvx = (nx - xi) / dt;
    nvy = (ny - yi) / dt;
    nvz = (nz - zi) / dt;
}

__global__ __launch_bounds__(kNeighbourThreads) void k_forces(ParticleSoA p, const uint32_t *__restrict__ cell,
                                                              const uint32_t *__restrict__ start,
                                                              const float *__restrict__ rho,
                                                              const double *__restrict__ prr, float *__restrict__ ax,
                                                              float *__restrict__ ay, float *__restrict__ az,
                                                              Scene s, Consts k)
{
    const uint32_t li = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (li >= s.n) return;
    const uint32_t i = s.poff + li;
    float a0, a1, a2;
    sph_accel(p, cell, start, rho, prr, i, s, k, p.px[i], p.py[i], p.pz[i], p.vx[i], p.vy[i], p.vz[i], a0, a1, a2);
    ax[i] = a0;
    ay[i] = a1;
    az[i] = a2;
}

__global__ void k_collide(ParticleSoA p, float *__restrict__ ax, float *__restrict__ ay, float *__restrict__ az,
                          Scene s, Consts k)
{
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= s.n) return;
    const uint32_t i = s.poff + li;
    float a0 = ax[i], a1 = ay[i], a2 = az[i];
    wall_accel(p.px[i], p.py[i], p.pz[i], p.vx[i], p.vy[i], p.vz[i], s, k, a0, a1, a2);
    ax[i] = a0;
    ay[i] = a1;
    az[i] = a2;
}

__global__ void k_integrate(ParticleSoA p, const float *__restrict__ ax, const float *__restrict__ ay,
                            const float *__restrict__ az, ParticleSoA out, uint32_t *__restrict__ key_next,
                            Scene s, Consts k)
{
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= s.n) return;
    const uint32_t i = s.poff + li;
    float nx, ny, nz, nvx, nvy, nvz;
    integrate_one(p.px[i], p.py[i], p.pz[i], p.vx[i], p.vy[i], p.vz[i], ax[i], ay[i], az[i], s.dt, nx, ny, nz,
                  nvx, nvy, nvz);
    out.px[i] = nx; out.py[i] = ny; out.pz[i] = nz;
    out.vx[i] = nvx; out.vy[i] = nvy; out.vz[i] = nvz;
    key_next[i] = hash_key(nx, ny, nz, s, k.h);
}

// Forces, walls and integration of one particle in one pass, plus the cell key of
// its new position for the next step's sort.
template <bool kStoreAcc>
__global__ __launch_bounds__(kNeighbourThreads) void k_forces_fused(
    ParticleSoA p, const uint32_t *__restrict__ cell, const uint32_t *__restrict__ start,
    const float *__restrict__ rho, const double *__restrict__ prr, ParticleSoA out, uint32_t *__restrict__ key_next,
    float *__restrict__ ax, float *__restrict__ ay, float *__restrict__ az, Scene s, Consts k)
{
    const uint32_t li = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (li >= s.n) return;
    const uint32_t i = s.poff + li;
    const float xi = p.px[i], yi = p.py[i], zi = p.pz[i];
    const float vxi = p.vx[i], vyi = p.vy[i], vzi = p.vz[i];
    float a0, a1, a2;
    sph_accel(p, cell, start, rho, prr, i, s, k, xi, yi, zi, vxi, vyi, vzi, a0, a1, a2);
    wall_accel(xi, yi, zi, vxi, vyi, vzi, s, k, a0, a1, a2);
    if (kStoreAcc) {
        ax[i] = a0;
        ay[i] = a1;
        az[i] = a2;
    }
    float nx, ny, nz, nvx, nvy, nvz;
    integrate_one(xi, yi, zi, vxi, vyi, vzi, a0, a1, a2, s.dt, nx, ny, nz, nvx, nvy, nvz);
    out.px[i] = nx; out.py[i] = ny; out.pz[i] = nz;
    out.vx[i] = nvx; out.vy[i] = nvy; out.vz[i] = nvz;
    key_next[i] = hash_key(nx, ny, nz, s, k.h);
}

inline uint32_t blocks_for(uint32_t n, uint32_t t) { return (n + t - 1) / t; }

}  // namespace

// ---------------------------------------------------------------- launchers

void launch_hash(const ParticleSoA &p, uint32_t *key, const Scene &s, double h, hipStream_t st)
{
    if (s.n == 0) return;
    hipLaunchKernelGGL(k_hash, dim3(blocks_for(s.n, 256)), dim3(256), 0, st, p.px, p.py, p.pz, key, s, h);
}

void launch_radix_pass(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t *keys_out, uint32_t *vals_out,
                       uint32_t n, int shift, uint32_t *hist, hipStream_t st)
{
    if (n == 0) return;
    const uint32_t nb = sort_blocks(n);
    hipLaunchKernelGGL(k_radix_count, dim3(nb), dim3(kSortThreads), 0, st, keys_in, n, shift, hist, nb);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, st, hist, nb * (uint32_t) kRadixBuckets);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kSortThreads), 0, st, keys_in, vals_in, keys_out, vals_out,
                       n, shift, (const uint32_t *) hist, nb);
}

void launch_cell_start(const uint32_t *sorted_keys, uint32_t n, uint32_t ncells, uint32_t *start, hipStream_t st)
{
    hipLaunchKernelGGL(k_cell_start, dim3(blocks_for(ncells + 1, 256)), dim3(256), 0, st, sorted_keys, n, ncells,
                       start);
}

void launch_reorder(const ParticleSoA &src, const ParticleSoA &dst, const uint32_t *perm, uint32_t n,
                    hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_reorder, dim3(blocks_for(n, 256)), dim3(256), 0, st, src, dst, perm, n);
}

void launch_density(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start, float *rho, double *prr,
                    uint32_t *ncount, const Scene &s, const Consts &k, hipStream_t st)
{
    if (s.n == 0) return;
    hipLaunchKernelGGL(k_density, dim3(blocks_for(s.n, kNeighbourThreads)), dim3(kNeighbourThreads), 0, st, p,
                       cell, start, rho, prr, ncount, s, k);
}

void launch_forces(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start, const float *rho,
                   const double *prr, float *ax, float *ay, float *az, const Scene &s, const Consts &k,
                   hipStream_t st)
{
    if (s.n == 0) return;
    hipLaunchKernelGGL(k_forces, dim3(blocks_for(s.n, kNeighbourThreads)), dim3(kNeighbourThreads), 0, st, p, cell,
                       start, rho, prr, ax, ay, az, s, k);
}

void launch_collide(const ParticleSoA &p, float *ax, float *ay, float *az, const Scene &s, const Consts &k,
                    hipStream_t st)
{
    if (s.n == 0) return;
    hipLaunchKernelGGL(k_collide, dim3(blocks_for(s.n, 256)), dim3(256), 0, st, p, ax, ay, az, s, k);
}

void launch_integrate(const ParticleSoA &p, const float *ax, const float *ay, const float *az, const ParticleSoA &out,
                      uint32_t *key_next, const Scene &s, const Consts &k, hipStream_t st)
{
    if (s.n == 0) return;
    hipLaunchKernelGGL(k_integrate, dim3(blocks_for(s.n, 256)), dim3(256), 0, st, p, ax, ay, az, out, key_next, s,
                       k);
}

void launch_forces_fused(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start, const float *rho,
                         const double *prr, const ParticleSoA &out, uint32_t *key_next, float *ax, float *ay,
                         float *az, const Scene &s, const Consts &k, hipStream_t st)
{
    if (s.n == 0) return;
    const dim3 grid(blocks_for(s.n, kNeighbourThreads)), block(kNeighbourThreads);
    if (ax)
        hipLaunchKernelGGL(k_forces_fused<true>, grid, block, 0, st, p, cell, start, rho, prr, out, key_next, ax, ay,
                           az, s, k);
    else
        hipLaunchKernelGGL(k_forces_fused<false>, grid, block, 0, st, p, cell, start, rho, prr, out, key_next, ax,
                           ay, az, s, k);
}

}  // namespace sph
