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
// Several independent scenes share the buffers: each owns a segment of slots and
// a segment of cell ids, and one launch of every stage covers all of them.

#include "sph_kernels.h"
#include "sph_math.h"
#include "sph_pow_tables.h"   // generated at build time (gen_pow_tables.hip)

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

// Local cell number of grid cell (x, y, z): x + y*rx + z*rx*ry (include/CGrid.h:25-28)
// in the canonical order, z + y*rz + x*rz*ry (the reference's traversal order,
// src/CCPUParticleSimulator.cpp:37-39) in the reference order.
template <bool kRef>
__device__ __forceinline__ uint32_t cell_number(int x, int y, int z, const Scene &s)
{
    return (uint32_t) (kRef ? z + y * s.rz + x * s.rz * s.ry : x + y * s.rx + z * s.rx * s.ry);
}

// Cell hash, src/CCPUParticleSimulator.cpp:46-69:
// floor(((double)p + (double)box/2.0) / (double)h), clamped, numbered as cell_number.
template <bool kRef>
__device__ __forceinline__ uint32_t hash_key(float px, float py, float pz, const Scene &s, double h)
{
    const int x = clamp_cell(to_int_x86(floor(((double) px + (double) s.bx / 2.0) / h)), s.rx);
    const int y = clamp_cell(to_int_x86(floor(((double) py + (double) s.by / 2.0) / h)), s.ry);
    const int z = clamp_cell(to_int_x86(floor(((double) pz + (double) s.bz / 2.0) / h)), s.rz);
    return s.coff + cell_number<kRef>(x, y, z, s);
}

// hash_key in the scene's own order (kernels that are not specialised on it).
__device__ __forceinline__ uint32_t hash_key_any(float px, float py, float pz, const Scene &s, double h)
{
    return s.order == kOrderReference ? hash_key<true>(px, py, pz, s, h) : hash_key<false>(px, py, pz, s, h);
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

// The scene of a workgroup. A workgroup never straddles two scenes (segments
// start on kGroupSlots boundaries), so the record is read into scalar registers.
__device__ __forceinline__ Scene block_scene(const SceneTable &t, uint32_t first_slot)
{
    const uint32_t sid =
        t.group_scene ? __builtin_amdgcn_readfirstlane(t.group_scene[first_slot / kGroupSlots]) : 0u;
    const uint32_t *w = reinterpret_cast<const uint32_t *>(t.scenes + sid);
    uint32_t v[sizeof(Scene) / 4];
#pragma unroll
    for (int q = 0; q < (int) (sizeof(Scene) / 4); ++q) v[q] = __builtin_amdgcn_readfirstlane(w[q]);
    Scene s;
    __builtin_memcpy(&s, v, sizeof(Scene));
    return s;
}

// The key of an unused slot: the scene's dead cell, after all of its grid cells.
__device__ __forceinline__ uint32_t dead_key(const Scene &s)
{
    return s.coff + s.ncells;
}

template <bool kRef>
__device__ __forceinline__ void cell_xyz(uint32_t key, const Scene &s, int &x, int &y, int &z)
{
    const uint32_t c = key - s.coff;
    if constexpr (kRef) {
        z = (int) (c % (uint32_t) s.rz);
        const uint32_t rem = c / (uint32_t) s.rz;
        y = (int) (rem % (uint32_t) s.ry);
        x = (int) (rem / (uint32_t) s.ry);
    } else {
        x = (int) (c % (uint32_t) s.rx);
        const uint32_t rem = c / (uint32_t) s.rx;
        y = (int) (rem % (uint32_t) s.ry);
        z = (int) (rem / (uint32_t) s.ry);
    }
}

// ---------------------------------------------------------------- grid

__global__ __launch_bounds__(256) void k_hash(const float *__restrict__ px, const float *__restrict__ py,
                                              const float *__restrict__ pz, uint32_t *__restrict__ key,
                                              SceneTable t, double h)
{
    const uint32_t first = blockIdx.x * 256u, i = first + threadIdx.x;
    const Scene s = block_scene(t, first);
    if (i >= t.nslots) return;
    key[i] = (i - s.poff < s.n) ? hash_key_any(px[i], py[i], pz[i], s, h) : dead_key(s);
}

// SplitMix64 output `k` (k >= 1) of the stream that starts at `state`.
__device__ __forceinline__ uint64_t splitmix64_at(uint64_t state, uint64_t k)
{
    uint64_t z = state + k * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float unit24(uint64_t x)
{
    return (float) (uint32_t) (x >> 40) * 0x1p-24f;
}

// Emitters, one workgroup each, run before the step's grid phase. The fountain
// adds the reference's 7 particles (src/CBaseParticleSimulator.cpp:187-210) while
// n < limit - 7 (unsigned, as the reference compares cl_int with cl_uint); the
// disc adds min(rate, limit - n) particles. Particle g of a disc emitter draws
// from its own SplitMix64 stream (state seed + 32*g*golden): 15 tries of a
// point (2u-1, 2u-1) in the unit square until it falls in the unit disc (else the
// centre), then v_y = speed * (1 + spread*(2u-1)) from draw 31. Every operation
// is a single IEEE float operation, so a CPU restatement gives the same bits.
__global__ __launch_bounds__(256) void k_emit(const Emitter *__restrict__ em, Scene *scenes, ParticleSoA p,
                                              uint32_t *__restrict__ key, uint32_t *__restrict__ ocell, double h)
{
    __shared__ uint32_t sh[2];
    const Emitter e = em[blockIdx.x];
    Scene *sp = scenes + e.scene;
    if (threadIdx.x == 0) {
        const uint32_t n0 = sp->n, cap = sp->cap;
        uint32_t cnt;
        if (e.kind == kEmitFountain)
            cnt = n0 >= e.limit - 7u ? 0u : 7u;
        else
            cnt = e.limit > n0 ? min(e.rate, e.limit - n0) : 0u;
        sh[0] = n0;
        sh[1] = min(cnt, cap - n0);
    }
    __syncthreads();
    const uint32_t n0 = sh[0], cnt = sh[1];
    const Scene s = *sp;
    const float floor_y = -s.by / 2.0f;
    for (uint32_t q = threadIdx.x; q < cnt; q += blockDim.x) {
        const uint32_t g = n0 + q, slot = s.poff + g;
        float x, z, vy;
        if (e.kind == kEmitFountain) {
            const float half = kH / 2.0f, quarter = kH / 4.0f;
            const float xs[7] = {0.0f, -half, half, -quarter, quarter, -quarter, quarter};
            const float zs[7] = {0.0f, 0.0f, 0.0f, -half, -half, half, half};
            x = xs[q];
            z = zs[q];
            vy = s.by * 3.2f;
        } else {
            const uint64_t st = ((uint64_t) e.seed_hi << 32 | e.seed_lo) + 32ull * g * 0x9E3779B97F4A7C15ull;
            float a = 0.0f, b = 0.0f;
            for (uint32_t tr = 0; tr < 15; ++tr) {
                const float u = 2.0f * unit24(splitmix64_at(st, 2 * tr + 1)) - 1.0f;
                const float w = 2.0f * unit24(splitmix64_at(st, 2 * tr + 2)) - 1.0f;
                if (u * u + w * w <= 1.0f) {
                    a = u;
                    b = w;
                    break;
                }
            }
            x = e.cx + e.radius * a;
            z = e.cz + e.radius * b;
            vy = e.speed * (1.0f + e.spread * (2.0f * unit24(splitmix64_at(st, 31)) - 1.0f));
        }
        p.px[slot] = x;
        p.py[slot] = floor_y;
        p.pz[slot] = z;
        p.vx[slot] = 0.0f;
        p.vy[slot] = vy;
        p.vz[slot] = 0.0f;
        p.id[slot] = g;
        key[slot] = hash_key_any(x, floor_y, z, s, h);
        if (ocell) ocell[slot] = 0u;   // addParticle puts it at the end of cell (0,0,0)'s bucket
    }
    __syncthreads();
    if (threadIdx.x == 0) sp->n = n0 + cnt;
}

// Digit histogram of one tile (2^BITS digits), stored digit-major (SortTile).
template <int BITS>
__global__ __launch_bounds__(kSortThreads) void k_radix_count(const uint32_t *__restrict__ keys,
                                                              const SortTile *__restrict__ tiles, int shift,
                                                              uint32_t *__restrict__ hist,
                                                              unsigned long long *__restrict__ scan_state,
                                                              uint32_t nscan_state)
{
    constexpr uint32_t B = 1u << BITS;
    __shared__ uint32_t h[B];
    const int t = threadIdx.x;
    // the next launch's single-pass scan starts from a zero ticket counter and no published tiles
    if (blockIdx.x == 0)
        for (uint32_t q = t; q < nscan_state; q += kSortThreads) scan_state[q] = 0ull;
    for (uint32_t d = t; d < B; d += kSortThreads) h[d] = 0;
    __syncthreads();
    const SortTile tl = tiles[blockIdx.x];
    // The keys arrive nearly sorted (the previous step's order), so neighbouring
    // lanes mostly share a digit: each run of equal digits in a wave adds its
    // length with one atomic, from its first lane.
    const uint32_t lane = (uint32_t) t & 63u;
    uint32_t kv[kSortItems];   // all loads in flight together (the kernel is latency-bound)
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t q = (uint32_t) (k * kSortThreads + t);
        kv[k] = q < tl.count ? keys[tl.start + q] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t q = (uint32_t) (k * kSortThreads + t);
        const bool valid = q < tl.count;
        const uint32_t dig = valid ? ((kv[k] - tl.coff) >> shift) & (B - 1) : B;
        const uint32_t prev = (uint32_t) __shfl_up((int) dig, 1, 64);
        const bool head = lane == 0 || dig != prev;
        const uint64_t heads = __ballot(head);
        // the next run's first lane, or 64
        const uint64_t after = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
        const uint32_t len = (after ? (uint32_t) __builtin_ctzll(after) : 64u) - lane;
        if (head && valid) atomicAdd(&h[dig], len);
    }
    __syncthreads();
    for (uint32_t d = t; d < B; d += kSortThreads) hist[tl.first * B + d * tl.ntiles + tl.index] = h[d];
}

// Exclusive scan of a 4096-element tile held 4 per thread by a 1024-thread block:
// wave64 shuffle scans, then a scan of the 16 wave totals. Returns the thread's
// exclusive prefix within the tile and the tile total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t s, uint32_t *wsum, uint32_t &total)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
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
        if (lane == 15) wsum[16] = xs;
    }
    __syncthreads();
    total = wsum[16];
    const uint32_t r = wsum[w] + (x - s);
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint32_t load4(const uint32_t *__restrict__ data, uint32_t idx, uint32_t m, uint32_t v[4])
{
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = (idx + q < m) ? data[idx + q] : 0u;
    return v[0] + v[1] + v[2] + v[3];
}

// Single-block in-place exclusive scan of m values (histograms of one scan tile).
__global__ __launch_bounds__(1024) void k_scan_exclusive(uint32_t *__restrict__ data, uint32_t m)
{
    __shared__ uint32_t wsum[17];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < m; base += kScanTile) {
        const uint32_t idx = base + (uint32_t) threadIdx.x * 4u;
        uint32_t v[4], total;
        uint32_t run = carry + block_exclusive_scan(load4(data, idx, m, v), wsum, total);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (idx + q < m) data[idx + q] = run;
            run += v[q];
        }
        carry += total;
    }
}

// Multi-block scan, step 1: the total of every 4096-element tile.
__global__ __launch_bounds__(1024) void k_scan_reduce(const uint32_t *__restrict__ data, uint32_t m,
                                                      uint32_t *__restrict__ partials)
{
    __shared__ uint32_t wsum[17];
    const uint32_t idx = blockIdx.x * (uint32_t) kScanTile + (uint32_t) threadIdx.x * 4u;
    uint32_t v[4], total;
    block_exclusive_scan(load4(data, idx, m, v), wsum, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Multi-block scan, step 2: exclusive scan of every tile, offset by the totals of
// the tiles before it (each block sums those itself: a few thousand at most).
__global__ __launch_bounds__(1024) void k_scan_apply(uint32_t *__restrict__ data, uint32_t m,
                                                     const uint32_t *__restrict__ partials)
{
    __shared__ uint32_t wsum[17];
    uint32_t before = 0, carry;
    for (uint32_t q = threadIdx.x; q < blockIdx.x; q += 1024u) before += partials[q];
    block_exclusive_scan(before, wsum, carry);
    const uint32_t idx = blockIdx.x * (uint32_t) kScanTile + (uint32_t) threadIdx.x * 4u;
    uint32_t v[4], total;
    uint32_t run = carry + block_exclusive_scan(load4(data, idx, m, v), wsum, total);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (idx + q < m) data[idx + q] = run;
        run += v[q];
    }
}

// Single-pass exclusive scan of the m values `in` into `out`, in tiles of
// kScanTile, with decoupled look-back: each block takes the next tile by ticket
// (so a tile only waits on tiles whose blocks have started), scans it, publishes
// its total, then its first wave reads the published words of up to 64
// predecessors at a time, back to the nearest one that holds an inclusive prefix
// (tile 0 publishes one at once), publishes its own inclusive prefix and writes
// its values. state[0] is the ticket counter, state[1 + k] tile k's word: kind in
// the high half (0 nothing yet, 1 total, 2 inclusive prefix), value in the low
// half, so one relaxed 64-bit atomic carries both and no other memory is shared
// between the blocks; the radix count kernel zeroes the words before every pass.
// The wait is bounded: a block that has polled poll_limit times (0: at once)
// adds up the raw values of the predecessors it has not summed yet from `in`,
// which no block writes, so the result is the same either way.
__global__ __launch_bounds__(1024) void k_scan_chained(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                       uint32_t m, unsigned long long *__restrict__ state,
                                                       uint32_t poll_limit)
{
    __shared__ uint32_t wsum[17];
    __shared__ uint32_t s_tile, s_prefix;
    if (threadIdx.x == 0) s_tile = (uint32_t) atomicAdd(&state[0], 1ull);
    __syncthreads();
    const uint32_t tile = s_tile;
    unsigned long long *word = state + 1;
    const uint32_t idx = tile * (uint32_t) kScanTile + (uint32_t) threadIdx.x * 4u;
    uint32_t v[4], total;
    const uint32_t local = block_exclusive_scan(load4(in, idx, m, v), wsum, total);
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        uint32_t prefix = 0;
        if (tile == 0) {
            if (lane == 0)
                __hip_atomic_store(&word[0], (2ull << 32) | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&word[tile], (1ull << 32) | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // window: predecessors top - lane, top = the nearest one not yet summed
            int top = (int) tile - 1;
            uint32_t polls = 0;
            bool raw = poll_limit == 0;
            while (!raw && top >= 0) {
                const int k = top - lane;
                const unsigned long long w =
                    k >= 0 ? __hip_atomic_load(&word[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : (2ull << 32);
                const uint32_t kind = (uint32_t) (w >> 32);
                const unsigned long long incl = __ballot(kind == 2), none = __ballot(kind == 0);
                // lanes up to the first inclusive prefix (all 64 when there is none in the window)
                const int stop = incl ? __builtin_ctzll(incl) : 63;
                const unsigned long long need = stop == 63 ? ~0ull : ((2ull << stop) - 1ull);
                if (none & need) {   // a predecessor in the window has published nothing yet
                    if (++polls >= poll_limit) {
                        raw = true;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                uint32_t x = lane <= stop ? (uint32_t) w : 0u;
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
                prefix += x;
                if (incl) break;
                top -= 64;
            }
            if (raw) {   // tiles 0..top, from their raw values
                uint32_t x = 0;
                const uint32_t end = (uint32_t) (top + 1) * (uint32_t) kScanTile;
                for (uint32_t q = (uint32_t) lane; q < end; q += 64u) x += in[q];
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
                prefix += x;
            }
            if (lane == 0)
                __hip_atomic_store(&word[tile], (2ull << 32) | (prefix + total), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    uint32_t run = s_prefix + local;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (idx + q < m) out[idx + q] = run;
        run += v[q];
    }
}

// Stable scatter of one LSD radix pass. Wave w of a tile owns its items
// w*64*kSortItems .. (w+1)*64*kSortItems - 1 (slot order), all loaded up front.
// The waves first count their digits into LDS (runs of equal digits in a round
// add their length with one atomic); the tile's offset of each digit (the scanned
// histogram) plus the counts of the waves before gives each wave its first
// offset per digit. Then every wave places its items round by round on its own,
// with no barrier: equal digits inside a round are matched with BITS ballots and
// ranked by popcount of the lower lanes; all lanes of a digit read its running
// offset, store at offset + rank, and the first of them advances the offset.
// Three barriers a tile; 2^BITS * 16 bytes of LDS.
template <int BITS>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in, uint32_t *__restrict__ keys_out,
    uint32_t *__restrict__ vals_out, const SortTile *__restrict__ tiles, int shift,
    const uint32_t *__restrict__ hist)
{
    constexpr uint32_t B = 1u << BITS;
    constexpr int W = kSortThreads / 64;
    __shared__ uint32_t base[W][B];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const SortTile tl = tiles[blockIdx.x];
    for (uint32_t d = t; d < W * B; d += kSortThreads) (&base[0][0])[d] = 0u;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t q0 = (uint32_t) (w * 64 * kSortItems + lane);
    uint32_t keys[kSortItems], vals[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t q = q0 + 64u * k, idx = tl.start + q;
        keys[k] = q < tl.count ? keys_in[idx] : 0u;
        vals[k] = q < tl.count ? (vals_in ? vals_in[idx] : idx) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const bool valid = q0 + 64u * k < tl.count;
        const uint32_t dig = valid ? ((keys[k] - tl.coff) >> shift) & (B - 1) : B;
        const uint32_t prev = (uint32_t) __shfl_up((int) dig, 1, 64);
        const bool head = lane == 0 || dig != prev;
        const uint64_t heads = __ballot(head);
        const uint64_t after = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
        const uint32_t len = (after ? (uint32_t) __builtin_ctzll(after) : 64u) - (uint32_t) lane;
        if (head && valid) atomicAdd(&base[w][dig], len);
    }
    __syncthreads();
    for (uint32_t d = t; d < B; d += kSortThreads) {
        uint32_t run = hist[tl.first * B + d * tl.ntiles + tl.index];
#pragma unroll
        for (int q = 0; q < W; ++q) {
            const uint32_t c = base[q][d];
            base[q][d] = run;
            run += c;
        }
    }
    __syncthreads();
    uint32_t *mine = base[w];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const bool valid = q0 + 64u * k < tl.count;
        const uint32_t dig = ((keys[k] - tl.coff) >> shift) & (B - 1);
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            const bool bit = (dig >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t) __popcll(m & below);
        const uint32_t off = mine[dig];   // the wave's lanes read before its first lane writes
        if (valid && rank == 0) mine[dig] = off + (uint32_t) __popcll(m);
        if (valid) {
            keys_out[off + rank] = keys[k];
            vals_out[off + rank] = vals[k];
        }
    }
}

// start[c] = the first sorted slot whose key is >= c (lower bound), for every
// c in [0, ncells]. Slot i (0 <= i < n) writes i into every cell in
// (keys[i - 1], keys[i]] (from cell 0 for i = 0), so each cell is written once
// and nothing is searched. A lane whose range is longer than kFillShort cells
// (empty regions of a scene) has it filled by the whole wave, 64 cells a store.
// The cells after the last key get n from every thread of the grid, a few
// each: that range is the largest (the empty top of a single scene).
constexpr uint32_t kFillShort = 8;

__global__ __launch_bounds__(256) void k_cell_start(const uint32_t *__restrict__ keys, uint32_t n,
                                                    uint32_t ncells, uint32_t *__restrict__ start)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x, lane = threadIdx.x & 63u;
    uint32_t lo = 1, len = 0;   // cells [lo, lo + len) get i
    if (i < n) {
        lo = i == 0 ? 0u : keys[i - 1] + 1u;
        const uint32_t hi = min(keys[i], ncells);   // keys are cells <= ncells
        len = lo <= hi ? hi - lo + 1u : 0u;
    }
    if (len <= kFillShort) {
        for (uint32_t q = 0; q < len; ++q) start[lo + q] = i;
    }
    uint64_t longs = __ballot(len > kFillShort);
    while (longs) {
        const uint32_t l = (uint32_t) __builtin_ctzll(longs);
        longs &= longs - 1u;
        const uint32_t llo = (uint32_t) __shfl((int) lo, (int) l, 64), llen = (uint32_t) __shfl((int) len, (int) l, 64);
        const uint32_t li = (uint32_t) __shfl((int) i, (int) l, 64);
        for (uint64_t q = lane; q < llen; q += 64u) start[llo + (uint32_t) q] = li;
    }
    // the tail: cells (keys[n - 1], ncells], all of them for n = 0
    const uint32_t tail = n == 0 ? 0u : min(keys[n - 1], ncells) + 1u;
    const uint32_t threads = gridDim.x * 256u;
    for (uint32_t c = tail + i; c <= ncells && c >= tail; c += threads) start[c] = n;
}

// Each thread moves S slots (blockDim apart), their loads in flight together:
// 8 below kReorderWideSlots slots (more bytes in flight per CU), 4 above (the
// gathers of a larger, more scattered permutation measured faster with fewer).
constexpr uint32_t kReorderWideSlots = 4u << 20;

template <uint32_t kReorderSlots>
__global__ void k_reorder(ParticleSoA src, ParticleSoA dst, const uint32_t *__restrict__ perm, uint32_t n)
{
    const uint32_t i0 = blockIdx.x * blockDim.x * kReorderSlots + threadIdx.x;
    uint32_t j[kReorderSlots];
#pragma unroll
    for (uint32_t q = 0; q < kReorderSlots; ++q) {
        const uint32_t i = i0 + q * blockDim.x;
        j[q] = i < n ? perm[i] : 0u;
    }
    float px[kReorderSlots], py[kReorderSlots], pz[kReorderSlots], vx[kReorderSlots], vy[kReorderSlots],
        vz[kReorderSlots];
    uint32_t id[kReorderSlots];
#pragma unroll
    for (uint32_t q = 0; q < kReorderSlots; ++q) {
        if (i0 + q * blockDim.x < n) {
            px[q] = src.px[j[q]];
            py[q] = src.py[j[q]];
            pz[q] = src.pz[j[q]];
            vx[q] = src.vx[j[q]];
            vy[q] = src.vy[j[q]];
            vz[q] = src.vz[j[q]];
            id[q] = src.id[j[q]];
        }
    }
#pragma unroll
    for (uint32_t q = 0; q < kReorderSlots; ++q) {
        const uint32_t i = i0 + q * blockDim.x;
        if (i < n) {
            dst.px[i] = px[q];
            dst.py[i] = py[q];
            dst.pz[i] = pz[q];
            dst.vx[i] = vx[q];
            dst.vy[i] = vy[q];
            dst.vz[i] = vz[q];
            dst.id[i] = id[q];
        }
    }
}

// Particles first .. first+n-1 of the current order written by id into xyz
// triples (render/readback order, as CGPUBaseParticleSimulator::integrate
// un-permutes by id, src/CGPUBaseParticleSimulator.cpp:83-91). vel may be null.
// An id outside 0..n-1 is not written; it raises *bad instead.
__global__ __launch_bounds__(256) void k_scatter_by_id(ParticleSoA p, uint32_t first, uint32_t n,
                                                       float *__restrict__ pos, float *__restrict__ vel,
                                                       uint32_t *__restrict__ bad)
{
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = first + k, id = p.id[i];
    if (id >= n) {
        atomicOr(bad, 1u);
        return;
    }
    pos[3 * (size_t) id + 0] = p.px[i];
    pos[3 * (size_t) id + 1] = p.py[i];
    pos[3 * (size_t) id + 2] = p.pz[i];
    if (vel) {
        vel[3 * (size_t) id + 0] = p.vx[i];
        vel[3 * (size_t) id + 1] = p.vy[i];
        vel[3 * (size_t) id + 2] = p.vz[i];
    }
}

// ---------------------------------------------------------------- reference bucket order
//
// CCPUParticleSimulator::updateGrid (src/CCPUParticleSimulator.cpp:37-88) keeps
// each cell's particles in a bucket and visits the cells in its traversal order
// T (x outer, z inner). In a visited bucket a particle whose cell changed is
// pushed to the back of its new cell's bucket and replaced by the bucket's last
// element (swap-remove), which is then examined in its place. New particles are
// pushed to the back of cell (0,0,0)'s bucket (src/CBaseParticleSimulator.cpp:67-74).
// Cells are numbered by T (cell_number), so the array order -- cells in T order,
// each cell's bucket in order -- is the reference's whole state.
//
// The kernels below compute the buckets after one updateGrid in parallel:
//  * The bucket a cell holds when it is visited is its old bucket followed by
//    its early arrivals: movers from cells visited before it (smaller T), in
//    push order. Early arrivals never move again, so the swap-remove scan of a
//    cell depends on them only through their count, which is order-free.
//  * Per cell, one thread runs the scan on (old bucket, early-arrival
//    placeholders): the final position of every stayer and placeholder, and the
//    departure rank of every mover.
//  * Pushes happen in order of (source T, departure rank). Movers sorted by
//    (destination, that order) give each cell its arrivals: the first n_early
//    fill the placeholders, the rest (from later cells, pushed after the scan)
//    go to the back in order.
// Every slot of the pool takes part (dead and free slots in their segment's
// dead cell, where they never move). ocell holds each live slot's cell from the
// previous grid phase as a local T number (0 for particles added since).

// Old cell of every slot as a global key (the dead cell for unused slots).
__global__ __launch_bounds__(256) void k_ref_prepare(const uint32_t *__restrict__ ocell, uint32_t *__restrict__ gold,
                                                     SceneTable t)
{
    const uint32_t first = blockIdx.x * 256u, i = first + threadIdx.x;
    const Scene s = block_scene(t, first);
    if (i >= t.nslots) return;
    gold[i] = (i - s.poff < s.n) ? s.coff + ocell[i] : dead_key(s);
}

__global__ __launch_bounds__(256) void k_ref_zero(uint32_t *__restrict__ a, uint32_t *__restrict__ b,
                                                  uint32_t *__restrict__ c, uint32_t *__restrict__ d, uint32_t m)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= m) return;
    a[i] = 0u;
    b[i] = 0u;
    c[i] = 0u;
    d[i] = 0u;
}

// Per cell: arrivals, departures and early arrivals (movers from smaller T).
__global__ __launch_bounds__(256) void k_ref_count(const uint32_t *__restrict__ gold, const uint32_t *__restrict__ gnew,
                                                   uint32_t n, uint32_t *__restrict__ n_early,
                                                   uint32_t *__restrict__ n_arrive, uint32_t *__restrict__ n_depart)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = gold[i], c = gnew[i];
    if (o == c) return;
    atomicAdd(&n_arrive[c], 1u);
    atomicAdd(&n_depart[o], 1u);
    if (o < c) atomicAdd(&n_early[c], 1u);
}

constexpr uint32_t kPlaceholder = 0x80000000u;

// The swap-remove scan of cell c (one thread per cell) over its old bucket
// (slots perm[ostart[c] .. ostart[c+1]), in order) followed by its early-arrival
// placeholders, in work[ostart[c] + eoff[c] ...]. rk gets each stayer's final
// position and each mover's departure rank; epos[eoff[c] + k] the final position
// of placeholder k; nstay[c] the bucket's length after the scan.
__global__ __launch_bounds__(256) void k_ref_bucket(uint32_t ncells, const uint32_t *__restrict__ ostart,
                                                    const uint32_t *__restrict__ perm,
                                                    const uint32_t *__restrict__ gnew,
                                                    const uint32_t *__restrict__ eoff, uint32_t *__restrict__ work,
                                                    uint32_t *__restrict__ rk, uint32_t *__restrict__ epos,
                                                    uint32_t *__restrict__ nstay)
{
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= ncells) return;
    const uint32_t a = ostart[c], n = ostart[c + 1] - a;
    const uint32_t e0 = eoff[c], ne = eoff[c + 1] - e0;
    uint32_t *w = work + a + e0;
    for (uint32_t q = 0; q < n; ++q) w[q] = perm[a + q];
    for (uint32_t k = 0; k < ne; ++k) w[n + k] = kPlaceholder | k;
    uint32_t len = n + ne, p = 0, dep = 0;
    while (p < len) {
        const uint32_t e = w[p];
        if (!(e & kPlaceholder) && gnew[e] != c) {
            rk[e] = dep++;
            w[p] = w[len - 1];
            --len;
        } else {
            ++p;
        }
    }
    for (uint32_t q = 0; q < len; ++q) {
        const uint32_t e = w[q];
        if (e & kPlaceholder)
            epos[e0 + (e & ~kPlaceholder)] = q;
        else
            rk[e] = q;
    }
    nstay[c] = len;
}

// Movers in push order: mover i goes to position moff[old cell] + departure rank,
// with its destination as key. Positions from the total mover count on are
// padding (key ncells, after every cell).
__global__ __launch_bounds__(256) void k_ref_movers(const uint32_t *__restrict__ gold, const uint32_t *__restrict__ gnew,
                                                    const uint32_t *__restrict__ rk, const uint32_t *__restrict__ moff,
                                                    uint32_t n, uint32_t ncells, uint32_t *__restrict__ mkey,
                                                    uint32_t *__restrict__ mval)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = gold[i], c = gnew[i];
    if (o != c) {
        const uint32_t d = moff[o] + rk[i];   // < the mover count <= n by construction
        if (d < n) {
            mkey[d] = c;
            mval[d] = i;
        }
    }
    if (i >= moff[ncells]) {
        mkey[i] = ncells;
        mval[i] = i;
    }
}

// Position q of the movers sorted by destination (push order within one):
// arrival k = q - astart[c] of cell c fills placeholder k if it is an early
// arrival, else goes to the back of the bucket.
__global__ __launch_bounds__(256) void k_ref_place(const uint32_t *__restrict__ skey, const uint32_t *__restrict__ sval,
                                                   const uint32_t *__restrict__ astart, const uint32_t *__restrict__ eoff,
                                                   const uint32_t *__restrict__ epos, const uint32_t *__restrict__ nstay,
                                                   uint32_t n, uint32_t ncells, uint32_t *__restrict__ rk)
{
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= n) return;
    const uint32_t c = skey[q];
    if (c >= ncells) return;
    const uint32_t k = q - astart[c], e0 = eoff[c], ne = eoff[c + 1] - e0;
    rk[sval[q]] = k < ne ? epos[e0 + k] : nstay[c] + (k - ne);
}

// Bucket lengths after updateGrid: the scanned bucket plus the late arrivals.
__global__ __launch_bounds__(256) void k_ref_counts(const uint32_t *__restrict__ nstay,
                                                    const uint32_t *__restrict__ n_arrive,
                                                    const uint32_t *__restrict__ eoff, uint32_t ncells,
                                                    uint32_t *__restrict__ count)
{
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c > ncells) return;
    count[c] = c < ncells ? nstay[c] + n_arrive[c] - (eoff[c + 1] - eoff[c]) : 0u;
}

// Slot of every particle in the new array order: its cell's start plus its
// bucket position. Writes the permutation, the sorted keys and the cells the
// next grid phase starts from.
__global__ __launch_bounds__(256) void k_ref_final(const uint32_t *__restrict__ gnew, const uint32_t *__restrict__ rk,
                                                   const uint32_t *__restrict__ start, uint32_t *__restrict__ skey,
                                                   uint32_t *__restrict__ perm, uint32_t *__restrict__ ocell,
                                                   SceneTable t)
{
    const uint32_t first = blockIdx.x * 256u, i = first + threadIdx.x;
    const Scene s = block_scene(t, first);
    if (i >= t.nslots) return;
    const uint32_t c = gnew[i], f = start[c] + rk[i];   // < nslots by construction
    if (f >= t.nslots) return;
    skey[f] = c;
    perm[f] = i;
    ocell[f] = c - s.coff;
}

// ---------------------------------------------------------------- neighbours
//
// Canonical neighbour order: rows z-1..z+1 (outer), y-1..y+1, and within a row
// the contiguous run of cells max(x-1,0)..min(x+1,X-1) in slot order. Rows off
// the grid are skipped, as the reference's continue/break bounds tests skip
// them (src/CCPUParticleSimulator.cpp:108-118).
//
// A wave walks the 9 rows together; each lane tests its own candidates and
// appends the accepted ones, in order, to a private list in LDS. When any lane's
// list is full every lane drains its list through the pair body, so the costly
// pair arithmetic runs on accepted pairs only (about 1 in 6 candidates) and the
// summation order of every particle is unchanged.

constexpr int kListCap = 16;
// The density pass's list: longer for 8-candidate trips (fewer, fuller flushes;
// measured faster with 8 candidates, slower with 4).
template <int U>
constexpr int density_list_cap() { return U >= 8 ? 24 : kListCap; }
constexpr uint32_t kBufWord3 = 0x00020000;  // descriptor word 3 of a gfx9 raw buffer

// Raw buffer views of one scene's per-particle arrays: loads take 32-bit byte
// offsets relative to the scene's first slot (constant parts fold into the
// instruction) instead of 64-bit addresses, and reads outside the scene's
// segment return 0 instead of faulting.
struct BufView {
    __amdgpu_buffer_rsrc_t px, py, pz, rho, rec;
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *) base, (short) 0, (int) bytes, (int) kBufWord3);
}

__device__ __forceinline__ BufView buf_view(const ParticleSoA &p, const DensityFields &f, const Scene &s)
{
    BufView v;
    v.px = rsrc(p.px + s.poff, s.cap * 4u);
    v.py = rsrc(p.py + s.poff, s.cap * 4u);
    v.pz = rsrc(p.pz + s.poff, s.cap * 4u);
    v.rho = rsrc(f.rho + s.poff, s.cap * 4u);
    v.rec = rsrc(f.rec + kRecWords * (size_t) s.poff, s.cap * (4u * kRecWords));
    return v;
}

__device__ __forceinline__ float ldf(__amdgpu_buffer_rsrc_t r, uint32_t byte_off)
{
    return __uint_as_float((uint32_t) __builtin_amdgcn_raw_buffer_load_b32(r, (int) byte_off, 0, 0));
}

__device__ __forceinline__ float4 ld4(__amdgpu_buffer_rsrc_t r, uint32_t byte_off)
{
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, (int) byte_off, 0, 0));
}

__device__ __forceinline__ uint2 ld2u(__amdgpu_buffer_rsrc_t r, uint32_t byte_off)
{
    return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, (int) byte_off, 0, 0));
}

// Squared lower bounds on the distance from a particle to any particle of the
// neighbour cells at offset -1 / +1 along one axis: the particle's distance to
// its own cell's lower / upper face (0 when it sits outside its clamped cell).
struct FaceBounds {
    double lo2, hi2;
};

__device__ __forceinline__ FaceBounds face_bounds(float p, float box, int c, double h)
{
    const double q = (double) p + (double) box / 2.0;
    const double dlo = q - (double) c * h, dhi = (double) (c + 1) * h - q;
    FaceBounds b;
    b.lo2 = dlo > 0.0 ? dlo * dlo : 0.0;
    b.hi2 = dhi > 0.0 ? dhi * dhi : 0.0;
    return b;
}

// Row r (0..8) of a particle's neighbourhood. Canonical order: z-1..z+1 (outer),
// y-1..y+1, and in it the contiguous run of cells x-1..x+1 in slot order.
// Reference order: x-1..x+1 (outer), y-1..y+1, and the run of cells z-1..z+1,
// contiguous in the traversal numbering -- the offsetX, offsetY, offsetZ loops of
// src/CCPUParticleSimulator.cpp:108-121 with each cell's bucket in slot order.
// Rows and end cells of a run whose distance lower bound exceeds h (with a 1e-5
// relative margin, far above the float rounding of r^2) are left out: every
// candidate they hold would fail the r^2 <= h^2 test, so the accepted neighbours
// and their order are unchanged. The density and force passes compute the same
// runs from the same inputs.
struct RowBounds {
    FaceBounds bx, by, bz;
    double lim;
};

__device__ __forceinline__ RowBounds row_bounds(const Scene &s, const Consts &k, int x, int y, int z, float xi,
                                                float yi, float zi)
{
    RowBounds rb;
    rb.bx = face_bounds(xi, s.bx, x, k.h);
    rb.by = face_bounds(yi, s.by, y, k.h);
    rb.bz = face_bounds(zi, s.bz, z, k.h);
    rb.lim = k.hsq * (1.0 + 1e-5);
    return rb;
}

// [j, jend): the candidate slots of row r, relative to the scene's first slot
// (empty when the row is off the grid or culled).
template <bool kRef>
__device__ __forceinline__ void row_run(const RowBounds &rb, const Scene &s, const uint32_t *__restrict__ start, int x,
                                        int y, int z, int r, uint32_t &j, uint32_t &jend)
{
    j = 0;
    jend = 0;
    const int oy = r % 3 - 1, yy = y + oy;
    const double ly = oy < 0 ? rb.by.lo2 : oy > 0 ? rb.by.hi2 : 0.0;
    if constexpr (kRef) {
        const int ox = r / 3 - 1, xx = x + ox;
        const double lxy = ly + (ox < 0 ? rb.bx.lo2 : ox > 0 ? rb.bx.hi2 : 0.0);
        if (xx >= 0 && xx < s.rx && yy >= 0 && yy < s.ry && lxy <= rb.lim) {
            const uint32_t row = s.coff + (uint32_t) (yy * s.rz + xx * s.rz * s.ry);
            const int z0 = (z > 0 && lxy + rb.bz.lo2 <= rb.lim) ? z - 1 : z;
            const int z1 = (z + 1 < s.rz && lxy + rb.bz.hi2 <= rb.lim) ? z + 1 : z;
            j = start[row + z0] - s.poff;
            jend = start[row + z1 + 1] - s.poff;
        }
        return;
    }
    const int oz = r / 3 - 1, zz = z + oz;
    const double lyz = ly + (oz < 0 ? rb.bz.lo2 : oz > 0 ? rb.bz.hi2 : 0.0);
    if (zz >= 0 && zz < s.rz && yy >= 0 && yy < s.ry && lyz <= rb.lim) {
        const uint32_t row = s.coff + (uint32_t) (yy * s.rx + zz * s.rx * s.ry);
        const int x0 = (x > 0 && lyz + rb.bx.lo2 <= rb.lim) ? x - 1 : x;
        const int x1 = (x + 1 < s.rx && lyz + rb.bx.hi2 <= rb.lim) ? x + 1 : x;
        j = start[row + x0] - s.poff;
        jend = start[row + x1 + 1] - s.poff;
    }
}

// Hooks of walk_neighbours that record nothing.
struct NoRowHooks {
    __device__ __forceinline__ void row_begin(int, uint32_t, uint32_t) {}
    __device__ __forceinline__ void trip(uint32_t, uint32_t) {}
    __device__ __forceinline__ void row_end(int) {}
};

// Walks the particle's candidates in canonical order, row by row (row_run).
// test(j, valid, xj, yj, zj) sees U candidates per trip in slot order (j
// relative to the scene's first slot), returns whether it accepted the
// candidate, and appends without branching (writes slot cnt, then adds the
// accept bit), so a list keeps U slots of headroom; when any lane's list is that
// full, flush() drains every lane's list. hooks.trip(j0, bits) gets the accept
// bits of each trip (bit u: candidate j0 + u). kWholeTrips: a trip in which
// every lane has U candidates left skips the per-candidate run-end test.
template <bool kRef, int U, bool kWholeTrips = false, int kCap = kListCap, typename Test, typename Flush,
          typename Hooks = NoRowHooks>
__device__ __forceinline__ void walk_neighbours(const Scene &s, const Consts &k, const uint32_t *__restrict__ start,
                                                const BufView &bv, int x, int y, int z, float xi, float yi,
                                                float zi, int &cnt, Test test, Flush flush, Hooks &&hooks = Hooks())
{
    static_assert(U < kCap, "a list must hold one trip of candidates");
    const RowBounds rb = row_bounds(s, k, x, y, z, xi, yi, zi);
    for (int r = 0; r < 9; ++r) {
        uint32_t j, jend;
        row_run<kRef>(rb, s, start, x, y, z, r, j, jend);
        hooks.row_begin(r, j, jend);
        while (__any(j < jend)) {
            if (j < jend) {
                const uint32_t off = j * 4u, rem = jend - j;
                float cx[U], cy[U], cz[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    cx[u] = ldf(bv.px, off + 4u * u);
                    cy[u] = ldf(bv.py, off + 4u * u);
                    cz[u] = ldf(bv.pz, off + 4u * u);
                }
                uint32_t bits = 0;
                if (kWholeTrips && __all(rem >= (uint32_t) U)) {
                    // every trip of the wave is whole: no per-candidate run-end test
#pragma unroll
                    for (int u = 0; u < U; ++u) bits |= (uint32_t) test(j + u, true, cx[u], cy[u], cz[u]) << u;
                } else {
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        bits |= (uint32_t) test(j + u, (uint32_t) u < rem, cx[u], cy[u], cz[u]) << u;
                }
                hooks.trip(j, bits);
                j += U;
            }
            if (__any(cnt > kCap - U)) {
                flush();
                cnt = 0;
            }
        }
        hooks.row_end(r);
    }
    flush();
    cnt = 0;
}

// Accepted-candidate masks of the density pass, for the force pass: bit b of
// row r's mask is candidate (run start + b) of that row, kept in 32-bit words
// (a run of up to 32 candidates, the common case on sparse scenes, costs 4
// bytes). A row holds up to kMaskBits candidates; a particle with a longer run
// has no masks (its force pass walks). Stored word-major per wave: word w of row
// r of slot i is masks[(((i / 64) * 9 + r) * kMaskWords + w) * 64 + i % 64], so
// the lanes of a wave write and read each word with one coalesced access.
__device__ __forceinline__ size_t mask_word(uint32_t i, int r, uint32_t w)
{
    return ((size_t) ((i >> 6) * 9u + (uint32_t) r) * kMaskWords + w) * 64u + (i & 63u);
}

// Records the accept bits of each trip into the row's current word (a trip of
// U candidates never straddles two words: U divides 32), storing a word when the
// run moves past it and at the end of the row.
struct MaskRecorder {
    uint32_t *masks;
    uint32_t i;
    uint32_t jrow = 0, len = 0, word = 0;
    uint32_t cur = 0;
    bool overflow = false;
    int row = 0;

    __device__ __forceinline__ void row_begin(int r, uint32_t j, uint32_t jend)
    {
        row = r;
        jrow = j;
        len = jend - j;
        word = 0;
        cur = 0;
        overflow |= len > kMaskBits;
    }
    __device__ __forceinline__ void trip(uint32_t j0, uint32_t bits)
    {
        const uint32_t b = j0 - jrow, w = b / kMaskWordBits;
        if (w != word) {   // runs visit their candidates in order: word only grows
            if (word < kMaskWords) masks[mask_word(i, row, word)] = cur;
            word = w;
            cur = 0;
        }
        cur |= bits << (b % kMaskWordBits);
    }
    __device__ __forceinline__ void row_end(int)
    {
        if (len > 0 && word < kMaskWords) masks[mask_word(i, row, word)] = cur;
    }
};

// One term of updateDensityPressure and the spiky gradient's factor, with the
// C library's pow (see density_term, spiky_factor).
__device__ __forceinline__ float density_add(float density, float r2, const Consts &k)
{
    return density_term(density, r2, k.hsq, k.poly6, g_pow3_exc, g_pow3_exc_count);
}

__device__ __forceinline__ float spiky_scale(double hr, float r2, const Consts &k)
{
    return spiky_factor(hr, r2, k.spiky, g_pow2_exc, g_pow2_exc_count);
}

// updateDensityPressure, src/CCPUParticleSimulator.cpp:93-141. Also stores
// p/rho^2 (double, as src/CCPUParticleSimulator.cpp:181 forms it), 1/rho
// (double) and, for the force pass, with kLists the particle's accepted
// neighbours (minus itself), with kMasks its rows' accepted-candidate masks
// (masks, and nl.cnt[i] = 1 when a run is too long for them); with kDiag, the
// accepted count. T threads per workgroup, U candidates per lane per trip.
template <int T, int U, bool kLists, bool kMasks, bool kDiag, bool kRef>
__global__ __launch_bounds__(T) void k_density(ParticleSoA p, const uint32_t *__restrict__ cell,
                                               const uint32_t *__restrict__ start, DensityFields f,
                                               NeighbourList nl, uint32_t *__restrict__ masks, SceneTable t,
                                               Consts k)
{
    static_assert(kMaskWordBits % U == 0, "a trip must not straddle two mask words");
    static_assert(kGroupSlots % T == 0, "a workgroup must lie in one slot group");
    constexpr int kCap = density_list_cap<U>();
    __shared__ float lst[kCap * T];
    const uint32_t first = xcd_block(blockIdx.x, gridDim.x) * (uint32_t) T, i = first + threadIdx.x;
    const Scene s = block_scene(t, first);
    const uint32_t li = i - s.poff;
    if (i >= t.nslots || li >= s.n) return;
    float *my = lst + threadIdx.x;
    uint32_t *nidx = kLists ? nl.idx + list_slot(nl, i) : nullptr;
    const BufView bv = buf_view(p, f, s);
    const float xi = p.px[i], yi = p.py[i], zi = p.pz[i];
    int x, y, z;
    cell_xyz<kRef>(cell[i], s, x, y, z);
    float density = 0.0f;
    uint32_t nc = 0, nn = 0;
    int cnt = 0;
    MaskRecorder rec_masks;
    rec_masks.masks = masks;
    rec_masks.i = i;
    auto test = [&](uint32_t j, bool valid, float xj, float yj, float zj) {
        const float dx = xi - xj, dyv = yi - yj, dzv = zi - zj;
        const float r2 = dx * dx + dyv * dyv + dzv * dzv;
        const bool ok = valid && r2 <= k.hsq_f;
        my[cnt * T] = r2;
        cnt += ok;
        if (kDiag) nc += ok;
        if (kLists && ok && j != li) {
            if (nn < nl.kmax) nidx[(size_t) nn * 64] = j;
            ++nn;
        }
        return ok;
    };
    // the terms of two pairs are formed before their in-order adds, so that the
    // loads and FP64 work of one pair overlap the other's
    auto flush = [&]() {
        int q = 0;
        for (; q + 1 < cnt; q += 2) {
            const float ra = my[q * T], rb = my[(q + 1) * T];
            const double ta = k.hsq - (double) ra, tb = k.hsq - (double) rb;
            const double pa = density_poly(ta, k.poly6), pb = density_poly(tb, k.poly6);
            density = density_accumulate(density, pa, ta, ra, k.poly6, g_pow3_exc, g_pow3_exc_count);
            density = density_accumulate(density, pb, tb, rb, k.poly6, g_pow3_exc, g_pow3_exc_count);
        }
        if (q < cnt) density = density_add(density, my[q * T], k);
    };
    if constexpr (kMasks)
        walk_neighbours<kRef, U, false, kCap>(s, k, start, bv, x, y, z, xi, yi, zi, cnt, test, flush, rec_masks);
    else
        walk_neighbours<kRef, U, false, kCap>(s, k, start, bv, x, y, z, xi, yi, zi, cnt, test, flush);
    density = density * k.mass;
    const float pressure = k.gas_k * (density - k.rho0);
    f.rho[i] = density;
    // rho*rho is exact in double, so this is p / pow(rho, 2)
    const double prr = (double) pressure / ((double) density * (double) density);
    const double rinv = 1.0 / (double) density;
    // the record, as five 8-byte stores (records are 8-byte aligned)
    uint2 *rec = reinterpret_cast<uint2 *>(f.rec + kRecWords * (size_t) i);
    const uint64_t pb = __builtin_bit_cast(uint64_t, prr), rb = __builtin_bit_cast(uint64_t, rinv);
    rec[0] = make_uint2(__float_as_uint(xi), __float_as_uint(yi));
    rec[1] = make_uint2(__float_as_uint(zi), __float_as_uint(p.vx[i]));
    rec[2] = make_uint2(__float_as_uint(p.vy[i]), __float_as_uint(p.vz[i]));
    rec[3] = make_uint2((uint32_t) pb, (uint32_t) (pb >> 32));
    rec[4] = make_uint2((uint32_t) rb, (uint32_t) (rb >> 32));
    if (kLists) nl.cnt[i] = nn;
    if (kMasks) nl.cnt[i] = rec_masks.overflow ? 1u : 0u;
    if (kDiag) f.ncount[i] = nc;
}

struct ForceSums {
    float px, py, pz, vx, vy, vz;
};

// The neighbour data one force pair reads, loaded one pair ahead of its use so
// that the loads of pair q+1 are in flight while pair q is computed.
struct PairData {
    uint32_t off;   // byte offset of the neighbour's record: 40 * its slot
    float x, y, z, vx, vy, vz;
    double prr, rinv;
};

// p/rho^2 of slot i from its record.
__device__ __forceinline__ double own_prr(const DensityFields &f, uint32_t i)
{
    const uint2 w = *reinterpret_cast<const uint2 *>(f.rec + kRecWords * (size_t) i + 6);
    return __builtin_bit_cast(double, (uint64_t) w.x | ((uint64_t) w.y << 32));
}

// The record at byte offset o (40 * slot).
__device__ __forceinline__ PairData load_pair_at(const BufView &bv, uint32_t o)
{
    const float4 a = ld4(bv.rec, o), b = ld4(bv.rec, o + 16u);
    const uint2 c = ld2u(bv.rec, o + 32u);
    PairData d;
    d.off = o;
    d.x = a.x;
    d.y = a.y;
    d.z = a.z;
    d.vx = a.w;
    d.vy = b.x;
    d.vz = b.y;
    d.prr = __builtin_bit_cast(double, (uint64_t) __float_as_uint(b.z) | ((uint64_t) __float_as_uint(b.w) << 32));
    d.rinv = __builtin_bit_cast(double, (uint64_t) c.x | ((uint64_t) c.y << 32));
    return d;
}

__device__ __forceinline__ PairData load_pair(const BufView &bv, uint32_t j)
{
    return load_pair_at(bv, j * (4u * kRecWords));
}

// The six float terms one accepted pair adds to the force sums
// (src/CCPUParticleSimulator.cpp:174-183, with WspikyGradient /
// WviscosityLaplacian, src/CCPUParticleSimulator.cpp:17-30): the pressure term
// grad * pf and the viscosity term (v_j - v_i) * lap / rho_j, each rounded to
// float before it is added, as the reference's QVector3D arithmetic does.
struct PairTerms {
    float px, py, pz, vx, vy, vz;
};

__device__ __forceinline__ PairTerms pair_terms(const BufView &bv, const Consts &k, const PairData &n, float xi,
                                                float yi, float zi, float vxi, float vyi, float vzi, double ai)
{
    const float dx = xi - n.x, dyv = yi - n.y, dzv = zi - n.z;
    const float r2 = dx * dx + dyv * dyv + dzv * dzv;
    const double x2 = (double) r2, y2 = __builtin_amdgcn_rsq(x2);
    const double r = sqrt_with_rsq(x2, y2);
    const double hr = k.h - r;
    const float sg = spiky_scale(hr, r2, k);
    float gx, gy, gz;
    div3_small_divisor(dx * sg, dyv * sg, dzv * sg, (float) r, gx, gy, gz, y2);
    const float pf = (float) (ai + n.prr);
    PairTerms t;
    t.px = gx * pf;
    t.py = gy * pf;
    t.pz = gz * pf;
    const float lap = (float) (k.visc * hr);
    div3_rcp((n.vx - vxi) * lap, (n.vy - vyi) * lap, (n.vz - vzi) * lap, n.rinv,
             [&]() { return ldf(bv.rho, n.off / 10u); }, t.vx, t.vy, t.vz);
    return t;
}

__device__ __forceinline__ void add_terms(ForceSums &a, const PairTerms &t)
{
    a.px = a.px + t.px;
    a.py = a.py + t.py;
    a.pz = a.pz + t.pz;
    a.vx = a.vx + t.vx;
    a.vy = a.vy + t.vy;
    a.vz = a.vz + t.vz;
}

// One accepted pair of updateForces, added to particle i's sums.
__device__ __forceinline__ void force_pair(const BufView &bv, const Consts &k, const PairData &n, float xi, float yi,
                                           float zi, float vxi, float vyi, float vzi, double ai, ForceSums &a)
{
    add_terms(a, pair_terms(bv, k, n, xi, yi, zi, vxi, vyi, vzi, ai));
}

// Drains a lane's LDS pair list `my` (stride T, cnt <= C entries) into its force
// sums, in list order. Two pair buffers alternate, each reloaded as soon as it
// has been used, so the loads of one pair overlap the arithmetic of the other.
// The loads are unconditional (indices past the end are clamped into the list
// and their pairs not added), which keeps the buffers in place instead of
// merging masked loads into copies.
template <int T, int C>
__device__ __forceinline__ void drain_pairs(const BufView &bv, const Consts &k, const uint32_t *my, int cnt, float xi,
                                            float yi, float zi, float vxi, float vyi, float vzi, double ai,
                                            ForceSums &a)
{
    if (cnt == 0) return;
    PairData pa = load_pair(bv, my[0]);
    for (int q = 0; q < cnt; q += 2) {
        const PairData pb = load_pair(bv, my[(q + 1 < C ? q + 1 : C - 1) * T]);
        force_pair(bv, k, pa, xi, yi, zi, vxi, vyi, vzi, ai, a);
        pa = load_pair(bv, my[(q + 2 < C ? q + 2 : C - 1) * T]);
        if (q + 1 < cnt) force_pair(bv, k, pb, xi, yi, zi, vxi, vyi, vzi, ai, a);
    }
}

// The force sums of particle i (slot li of its scene) from walking its cells,
// in canonical order: each lane appends its accepted neighbours to its LDS list
// `my` (stride T) and drains it through the pair body.
template <bool kRef, int T, int U>
__device__ __forceinline__ ForceSums walk_sums(const BufView &bv, const uint32_t *__restrict__ cell,
                                               const uint32_t *__restrict__ start, uint32_t *my, uint32_t i,
                                               uint32_t li, const Scene &s, const Consts &k, float xi, float yi,
                                               float zi, float vxi, float vyi, float vzi, double ai)
{
    ForceSums a = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    int x, y, z;
    cell_xyz<kRef>(cell[i], s, x, y, z);
    int cnt = 0;
    walk_neighbours<kRef, U, true>(
        s, k, start, bv, x, y, z, xi, yi, zi, cnt,
        [&](uint32_t j, bool valid, float xj, float yj, float zj) {
            const float dx = xi - xj, dyv = yi - yj, dzv = zi - zj;
            const float r2 = dx * dx + dyv * dyv + dzv * dzv;
            my[cnt * T] = j;
            const bool ok = valid && r2 <= k.hsq_f && j != li;
            cnt += ok;
            return ok;
        },
        [&]() { drain_pairs<T, kListCap>(bv, k, my, cnt, xi, yi, zi, vxi, vyi, vzi, ai, a); });
    return a;
}

// The force sums of particle i from the density pass's row masks: the same rows
// and runs (row_run), and in each run only the accepted candidates, the set
// bits in ascending order, so the pairs and their order are the walk's. The
// lane's pair list `my` holds C entries (stride T); a longer list means fewer,
// fuller drains. kStall == 0 drains all lists when one is full; kStall == P
// lets full lanes wait (keeping their bits) until P percent of the wave waits,
// or until every lane with bits left waits, so that drains run with most lanes
// busy. Each lane's pairs keep their order either way.
template <bool kRef, int T, int C, int kStall>
__device__ __forceinline__ ForceSums mask_sums(const BufView &bv, const uint32_t *__restrict__ cell,
                                               const uint32_t *__restrict__ start,
                                               const uint32_t *__restrict__ masks, uint32_t *my, uint32_t i,
                                               uint32_t li, const Scene &s, const Consts &k, float xi, float yi,
                                               float zi, float vxi, float vyi, float vzi, double ai)
{
    ForceSums a = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    int x, y, z;
    cell_xyz<kRef>(cell[i], s, x, y, z);
    const RowBounds rb = row_bounds(s, k, x, y, z, xi, yi, zi);
    // rows and words advance together for the wave (r, w uniform); m holds the
    // lane's unread bits of word w of row r, whose first candidate is jw
    int r = -1;
    uint32_t w = 0, nw = 0, jrun = 0, jw = 0;
    uint32_t m = 0;
    bool done = false;
    int cnt = 0;
    for (;;) {
        // fill the pair lists until a drain is due; the sums are not touched here
        for (;;) {
            if (!__any(m != 0u)) {
                ++w;
                if (!__any(w < nw)) {
                    if (++r == 9) {
                        done = true;
                        break;
                    }
                    uint32_t jend;
                    row_run<kRef>(rb, s, start, x, y, z, r, jrun, jend);
                    nw = (jend - jrun + kMaskWordBits - 1) / kMaskWordBits;
                    w = 0;
                }
                m = w < nw ? masks[mask_word(i, r, w)] : 0u;
                jw = jrun + kMaskWordBits * w;
                continue;
            }
            // two bits per round; kStall: a lane whose list is full keeps its
            // bits until the drain (without kStall a list has room for both)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                if (m != 0u && (kStall == 0 || cnt < C)) {
                    const uint32_t jj = jw + (uint32_t) __builtin_ctz(m);
                    m &= m - 1u;
                    my[cnt * T] = jj;
                    cnt += jj != li;
                }
            }
            bool drain;
            if (kStall == 0) {
                drain = __any(cnt > C - 2);
            } else {
                const bool full = cnt == C;
                const uint32_t waiting = (uint32_t) __popcll(__ballot(full));
                const uint32_t lanes = (uint32_t) __popcll(__ballot(true));
                drain = waiting * 100u >= (uint32_t) kStall * lanes || (waiting && !__any(m != 0u && !full));
            }
            if (drain) break;
        }
        drain_pairs<T, C>(bv, k, my, cnt, xi, yi, zi, vxi, vyi, vzi, ai, a);
        cnt = 0;
        if (done) break;
    }
    return a;
}

// Drains a lane's queue of mask words (nq entries of two 32-bit words at
// stride T: the bits and the record offset of bit 0, 40 * its slot) into its
// force sums: the set bits of each word in ascending order, word after word. Two
// pairs are in flight, as in drain_pairs; a lane with no pair left loads the
// record at `fallback` (an offset) and adds nothing. A bit's record offset is
// the word's plus a 24-bit product, so no full 32-bit multiply per pair.
template <int T>
__device__ __forceinline__ void drain_words(const BufView &bv, const Consts &k, const uint32_t *qw, int nq,
                                            uint32_t fallback, float xi, float yi, float zi, float vxi, float vyi,
                                            float vzi, double ai, ForceSums &a)
{
    int q = 0;
    uint32_t m = 0, o0 = 0;
    auto next = [&](uint32_t &o) {
        if (m == 0u && q < nq) {
            m = qw[(2 * q) * T];
            o0 = qw[(2 * q + 1) * T];
            ++q;
        }
        const bool ok = m != 0u;
        o = ok ? o0 + __umul24((uint32_t) __builtin_ctz(m), 4u * kRecWords) : fallback;
        m &= m - 1u;
        return ok;
    };
    uint32_t oa, ob;
    bool va = next(oa);
    PairData pa = load_pair_at(bv, oa);
    while (__any(va)) {
        const bool vb = next(ob);
        const PairData pb = load_pair_at(bv, ob);
        if (va) force_pair(bv, k, pa, xi, yi, zi, vxi, vyi, vzi, ai, a);
        va = next(oa);
        pa = load_pair_at(bv, oa);
        if (vb) force_pair(bv, k, pb, xi, yi, zi, vxi, vyi, vzi, ai, a);
    }
}

// s_waitcnt vmcnt(0) (gfx9 encoding; expcnt 7 and lgkmcnt 15: not waited on)
constexpr unsigned kWaitVmcnt0 = 0x0F70;

// The force sums of particle i from the row masks, as mask_sums, with another
// division of the work: the wave reads the rows and words together (coalesced)
// and each lane queues its non-empty words, without its own bit, in LDS (`qw`,
// W entries at stride T); drain_words runs through every lane's queue when one
// is full and at the end. With W large enough for a particle's words there is
// one drain per wave and no per-bit fill rounds, at the price of more LDS.
template <bool kRef, int T, int W>
__device__ __forceinline__ ForceSums mask_sums_queue(const BufView &bv, const uint32_t *__restrict__ cell,
                                                     const uint32_t *__restrict__ start,
                                                     const uint32_t *__restrict__ masks, uint32_t *qw, uint32_t i,
                                                     uint32_t li, const Scene &s, const Consts &k, float xi,
                                                     float yi, float zi, float vxi, float vyi, float vzi, double ai)
{
    ForceSums a = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    int x, y, z;
    cell_xyz<kRef>(cell[i], s, x, y, z);
    const RowBounds rb = row_bounds(s, k, x, y, z, xi, yi, zi);
    // loaded by idle lanes, never added (a record offset)
    const uint32_t fallback = (li + 1 < s.n ? li + 1 : li - 1) * (4u * kRecWords);
    // the first word of every row, loaded together up front (unused where a row is empty)
    uint32_t m0[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) m0[r] = masks[mask_word(i, r, 0)];
    int nq = 0;
    for (int r = 0; r < 9; ++r) {
        uint32_t jrun, jend;
        row_run<kRef>(rb, s, start, x, y, z, r, jrun, jend);
        const uint32_t nw = (jend - jrun + kMaskWordBits - 1) / kMaskWordBits;
        for (uint32_t w = 0; __any(w < nw); ++w) {
            uint32_t m = w >= nw ? 0u : w == 0 ? m0[r] : masks[mask_word(i, r, w)];
            const uint32_t j0 = jrun + kMaskWordBits * w, self = li - j0;
            if (self < kMaskWordBits) m &= ~(1u << self);
            if (m != 0u) {
                qw[(2 * nq) * T] = m;
                qw[(2 * nq + 1) * T] = j0 * (4u * kRecWords);
                ++nq;
            }
            if (__any(nq == W)) {
                drain_words<T>(bv, k, qw, nq, fallback, xi, yi, zi, vxi, vyi, vzi, ai, a);
                __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);   // the drain's last, unused loads
                nq = 0;
            }
        }
    }
    drain_words<T>(bv, k, qw, nq, fallback, xi, yi, zi, vxi, vyi, vzi, ai, a);
    return a;
}

// (f_p + f_v + f_g) / rho from the force sums (src/CCPUParticleSimulator.cpp:155-193).
__device__ __forceinline__ void accel_from_sums(const ForceSums &a, float rhoi, const Scene &s, const Consts &k,
                                                float &ax, float &ay, float &az)
{
    const float fac = k.neg_mass * rhoi;
    const float fpx = a.px * fac, fpy = a.py * fac, fpz = a.pz * fac;
    const float fvx = a.vx * k.mu_mass, fvy = a.vy * k.mu_mass, fvz = a.vz * k.mu_mass;
    const float fgx = s.gx * rhoi, fgy = s.gy * rhoi, fgz = s.gz * rhoi;
    ax = ((fpx + fvx) + fgx) / rhoi;
    ay = ((fpy + fvy) + fgy) / rhoi;
    az = ((fpz + fvz) + fgz) / rhoi;
}

// The acceleration of particle i. Neighbours come from the density pass's list
// when one was kept (each lane running through its own list); otherwise (or on
// overflow) the cells are walked again in the same order.
template <bool kRef, int T, int U>
__device__ __forceinline__ void sph_accel(const ParticleSoA &p, const uint32_t *__restrict__ cell,
                                          const uint32_t *__restrict__ start, const DensityFields &f,
                                          const NeighbourList &nl, uint32_t *my, uint32_t i, uint32_t li,
                                          const Scene &s, const Consts &k, float xi, float yi, float zi, float vxi,
                                          float vyi, float vzi, float &ax, float &ay, float &az)
{
    const BufView bv = buf_view(p, f, s);
    const float rhoi = f.rho[i];
    const double ai = own_prr(f, i);
    ForceSums a = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t cnt_i = nl.mode == kNbLists ? nl.cnt[i] : 0u;
    if (nl.mode == kNbLists && cnt_i <= nl.kmax) {
        const uint32_t *nidx = nl.idx + list_slot(nl, i);
        uint32_t j2 = cnt_i > 1 ? nidx[64] : 0u;
        PairData nxt = load_pair(bv, cnt_i ? nidx[0] : 0u);
        for (uint32_t q = 0; q < cnt_i; ++q) {
            const PairData cur = nxt;
            if (q + 1 < cnt_i) {
                nxt = load_pair(bv, j2);
                if (q + 2 < cnt_i) j2 = nidx[(size_t) (q + 2) * 64];
            }
            force_pair(bv, k, cur, xi, yi, zi, vxi, vyi, vzi, ai, a);
        }
    } else if (nl.mode == kNbMasks && nl.cnt[i] == 0) {
        a = mask_sums<kRef, T, kListCap, 0>(bv, cell, start, nl.masks, my, i, li, s, k, xi, yi, zi, vxi, vyi, vzi, ai);
    } else {
        a = walk_sums<kRef, T, U>(bv, cell, start, my, i, li, s, k, xi, yi, zi, vxi, vyi, vzi, ai);
    }
    accel_from_sums(a, rhoi, s, k, ax, ay, az);
}

// Force sums of a wave's 64 particles from their neighbour lists, with the pair
// arithmetic spread over the whole wave. The pairs are taken level by level (the
// k-th neighbour of every particle that has one, in lane order, then level k+1,
// ...), 16 levels at a time: their (level, owner) items go to LDS, and each
// round the active lanes compute the terms of the next items, one each,
// whatever their owners. The terms are added to the owners' sums in LDS one level after the
// other; an owner has at most one item per level, so no two lanes add to the
// same sums at once, and every particle adds its terms in its own list order,
// as the per-lane loop would. Lanes with ncnt = 0 take no part.
constexpr uint32_t kLevelWindow = 16;

__device__ __forceinline__ ForceSums wave_list_sums(const BufView &bv, const Consts &k, const NeighbourList &nl,
                                                    uint32_t i, uint32_t ncnt, float xi, float yi, float zi,
                                                    float vxi, float vyi, float vzi, double ai, float4 *s_owner,
                                                    float4 *s_acc, uint16_t *s_items)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t ab = __builtin_bit_cast(uint64_t, ai);
    s_owner[2 * lane] = make_float4(xi, yi, zi, vxi);
    s_owner[2 * lane + 1] = make_float4(vyi, vzi, __uint_as_float((uint32_t) ab), __uint_as_float((uint32_t) (ab >> 32)));
    s_acc[2 * lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    s_acc[2 * lane + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t *wl = nl.idx + (size_t) (i >> 6) * nl.kmax * 64;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    // lanes past the end of the slots have left: items are dealt to the active ones
    const uint64_t active = __ballot(true);
    const uint32_t nactive = (uint32_t) __popcll(active), rank = (uint32_t) __popcll(active & below);
    for (uint32_t k0 = 0;; k0 += kLevelWindow) {
        uint32_t nitems = 0;
        for (uint32_t kk = k0; kk < k0 + kLevelWindow; ++kk) {
            const uint64_t m = __ballot(ncnt > kk);
            if (m == 0) break;
            if (ncnt > kk) s_items[nitems + (uint32_t) __popcll(m & below)] = (uint16_t) (((kk - k0) << 6) | lane);
            nitems += (uint32_t) __popcll(m);
        }
        if (nitems == 0) break;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t c = 0; c < nitems; c += nactive) {
            const uint32_t q = c + rank;
            const bool val = q < nitems;
            const uint32_t item = val ? s_items[q] : 0u;
            const uint32_t lev = item >> 6, o = item & 63u;
            PairTerms tt = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (val) {
                const float4 o0 = s_owner[2 * o], o1 = s_owner[2 * o + 1];
                const double ao = __builtin_bit_cast(double, (uint64_t) __float_as_uint(o1.z) |
                                                                 ((uint64_t) __float_as_uint(o1.w) << 32));
                const PairData n = load_pair(bv, wl[(size_t) (k0 + lev) * 64 + o]);
                tt = pair_terms(bv, k, n, o0.x, o0.y, o0.z, o0.w, o1.x, o1.y, ao);
            }
            const uint32_t first_lev = __builtin_amdgcn_readfirstlane((uint32_t) s_items[c] >> 6);
            const uint32_t last_q = c + nactive - 1 < nitems - 1 ? c + nactive - 1 : nitems - 1;
            const uint32_t last_lev = __builtin_amdgcn_readfirstlane((uint32_t) s_items[last_q] >> 6);
            for (uint32_t l = first_lev; l <= last_lev; ++l) {
                if (val && lev == l) {
                    float4 a0 = s_acc[2 * o], a1 = s_acc[2 * o + 1];
                    a0.x = a0.x + tt.px;
                    a0.y = a0.y + tt.py;
                    a0.z = a0.z + tt.pz;
                    a0.w = a0.w + tt.vx;
                    a1.x = a1.x + tt.vy;
                    a1.y = a1.y + tt.vz;
                    s_acc[2 * o] = a0;
                    s_acc[2 * o + 1] = a1;
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    const float4 a0 = s_acc[2 * lane], a1 = s_acc[2 * lane + 1];
    ForceSums a;
    a.px = a0.x;
    a.py = a0.y;
    a.pz = a0.z;
    a.vx = a0.w;
    a.vy = a1.x;
    a.vz = a1.y;
    return a;
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

constexpr int kPhaseThreads = 256;   // workgroup of the phase-by-phase kernels
constexpr int kPhaseUnroll = 4;

template <bool kRef>
__global__ __launch_bounds__(kPhaseThreads) void k_forces(ParticleSoA p, const uint32_t *__restrict__ cell,
                                                          const uint32_t *__restrict__ start, DensityFields f,
                                                          NeighbourList nl, float *__restrict__ ax,
                                                          float *__restrict__ ay, float *__restrict__ az,
                                                          SceneTable t, Consts k)
{
    __shared__ uint32_t lst[kListCap * kPhaseThreads];
    const uint32_t first = xcd_block(blockIdx.x, gridDim.x) * (uint32_t) kPhaseThreads, i = first + threadIdx.x;
    const Scene s = block_scene(t, first);
    const uint32_t li = i - s.poff;
    if (i >= t.nslots || li >= s.n) return;
    float a0, a1, a2;
    sph_accel<kRef, kPhaseThreads, kPhaseUnroll>(p, cell, start, f, nl, lst + threadIdx.x, i, li, s, k, p.px[i], p.py[i],
                                           p.pz[i], p.vx[i], p.vy[i], p.vz[i], a0, a1, a2);
    ax[i] = a0;
    ay[i] = a1;
    az[i] = a2;
}

__global__ __launch_bounds__(256) void k_collide(ParticleSoA p, float *__restrict__ ax, float *__restrict__ ay,
                                                 float *__restrict__ az, SceneTable t, Consts k)
{
    const uint32_t first = blockIdx.x * 256u, i = first + threadIdx.x;
    const Scene s = block_scene(t, first);
    if (i >= t.nslots || i - s.poff >= s.n) return;
    float a0 = ax[i], a1 = ay[i], a2 = az[i];
    wall_accel(p.px[i], p.py[i], p.pz[i], p.vx[i], p.vy[i], p.vz[i], s, k, a0, a1, a2);
    ax[i] = a0;
    ay[i] = a1;
    az[i] = a2;
}

__global__ __launch_bounds__(256) void k_integrate(ParticleSoA p, const float *__restrict__ ax,
                                                   const float *__restrict__ ay, const float *__restrict__ az,
                                                   ParticleSoA out, uint32_t *__restrict__ key_next, SceneTable t,
                                                   Consts k)
{
    const uint32_t first = blockIdx.x * 256u, i = first + threadIdx.x;
    const Scene s = block_scene(t, first);
    if (i >= t.nslots) return;
    if (i - s.poff >= s.n) {
        key_next[i] = dead_key(s);
        return;
    }
    float nx, ny, nz, nvx, nvy, nvz;
    integrate_one(p.px[i], p.py[i], p.pz[i], p.vx[i], p.vy[i], p.vz[i], ax[i], ay[i], az[i], s.dt, nx, ny, nz,
                  nvx, nvy, nvz);
    out.px[i] = nx; out.py[i] = ny; out.pz[i] = nz;
    out.vx[i] = nvx; out.vy[i] = nvy; out.vz[i] = nvz;
    key_next[i] = hash_key_any(nx, ny, nz, s, k.h);
}

// Forces, walls and integration of one particle in one pass, plus the cell key of
// its new position for the next step's sort (the dead cell for unused slots).
// Neighbours: kNbWalk walks the cells; kNbLists takes the density pass's lists,
// a wave at a time (wave_list_sums); kNbMasks iterates the density pass's row
// masks (mask_sums). Particles whose list or masks overflowed walk their cells.
template <int T, int U, int C, int kStall, bool kQueue, int kMode, bool kStoreAcc, bool kRef>
__global__ __launch_bounds__(T) void k_forces_fused(ParticleSoA p, const uint32_t *__restrict__ cell,
                                                    const uint32_t *__restrict__ start, DensityFields f,
                                                    NeighbourList nl, ParticleSoA out,
                                                    uint32_t *__restrict__ key_next, float *__restrict__ ax,
                                                    float *__restrict__ ay, float *__restrict__ az, SceneTable t,
                                                    Consts k)
{
    static_assert(kGroupSlots % T == 0, "a workgroup must lie in one slot group");
    constexpr bool kLists = kMode == kNbLists;
    __shared__ uint32_t lst[(kMode == kNbMasks && C > kListCap ? C : kListCap) * T];
    __shared__ float4 s_owner[kLists ? 2 * T : 1];
    __shared__ float4 s_acc[kLists ? 2 * T : 1];
    __shared__ uint16_t s_items[kLists ? kLevelWindow * T : 1];
    const uint32_t first = xcd_block(blockIdx.x, gridDim.x) * (uint32_t) T, i = first + threadIdx.x;
    const Scene s = block_scene(t, first);
    const uint32_t li = i - s.poff;
    if (i >= t.nslots) return;
    if (li >= s.n) {
        key_next[i] = dead_key(s);
        return;
    }
    const float xi = p.px[i], yi = p.py[i], zi = p.pz[i];
    const float vxi = p.vx[i], vyi = p.vy[i], vzi = p.vz[i];
    float a0, a1, a2;
    if (kMode == kNbWalk) {
        sph_accel<kRef, T, U>(p, cell, start, f, nl, lst + threadIdx.x, i, li, s, k, xi, yi, zi, vxi, vyi, vzi, a0, a1,
                        a2);
    } else {
        const BufView bv = buf_view(p, f, s);
        const float rhoi = f.rho[i];
        const double ai = own_prr(f, i);
        const uint32_t cnt_i = nl.cnt[i];
        ForceSums a;
        bool overflow;
        if (kLists) {
            overflow = cnt_i > nl.kmax;
            const uint32_t w = threadIdx.x >> 6;
            a = wave_list_sums(bv, k, nl, i, overflow ? 0u : cnt_i, xi, yi, zi, vxi, vyi, vzi, ai, s_owner + 128 * w,
                               s_acc + 128 * w, s_items + kLevelWindow * 64 * w);
        } else {
            overflow = cnt_i != 0;
            if (!overflow && kQueue)
                a = mask_sums_queue<kRef, T, C / 2>(bv, cell, start, nl.masks, lst + threadIdx.x, i, li, s, k, xi, yi, zi,
                                              vxi, vyi, vzi, ai);
            else if (!overflow)
                a = mask_sums<kRef, T, C, kStall>(bv, cell, start, nl.masks, lst + threadIdx.x, i, li, s, k, xi, yi, zi,
                                            vxi, vyi, vzi, ai);
        }
        if (overflow)
            a = walk_sums<kRef, T, U>(bv, cell, start, lst + threadIdx.x, i, li, s, k, xi, yi, zi, vxi, vyi, vzi, ai);
        accel_from_sums(a, rhoi, s, k, a0, a1, a2);
    }
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
    key_next[i] = hash_key<kRef>(nx, ny, nz, s, k.h);
}

inline uint32_t blocks_for(uint32_t n, uint32_t t) { return (n + t - 1) / t; }

// The compiled shapes of the two neighbour kernels, indexed by variant.
template <int T, int U, int C, int kStall, bool kQueue, bool kRef = false>
struct NeighbourKernels {
    // the density pass of the word-queue variants keeps 64-thread workgroups
    // (its fastest shape); `T` then shapes the force pass only
    static constexpr int TD = kQueue ? 64 : T;
    template <bool kLists, bool kMasks>
    static void density_mode(hipStream_t st, const ParticleSoA &p, const uint32_t *cell, const uint32_t *start,
                             const DensityFields &f, const NeighbourList &nl, const SceneTable &t, const Consts &k)
    {
        const dim3 g(blocks_for(t.nslots, TD));
        if (f.ncount)
            hipLaunchKernelGGL((k_density<TD, U, kLists, kMasks, true, kRef>), g, dim3(TD), 0, st, p, cell, start, f,
                               nl, nl.masks, t, k);
        else
            hipLaunchKernelGGL((k_density<TD, U, kLists, kMasks, false, kRef>), g, dim3(TD), 0, st, p, cell, start, f,
                               nl, nl.masks, t, k);
    }
    static void density(hipStream_t st, const ParticleSoA &p, const uint32_t *cell, const uint32_t *start,
                        const DensityFields &f, const NeighbourList &nl, const SceneTable &t, const Consts &k)
    {
        if (nl.mode == kNbLists)
            density_mode<true, false>(st, p, cell, start, f, nl, t, k);
        else if (nl.mode == kNbMasks)
            density_mode<false, true>(st, p, cell, start, f, nl, t, k);
        else
            density_mode<false, false>(st, p, cell, start, f, nl, t, k);
    }
    template <int kMode>
    static void fused_mode(hipStream_t st, const ParticleSoA &p, const uint32_t *cell, const uint32_t *start,
                           const DensityFields &f, const NeighbourList &nl, const ParticleSoA &out,
                           uint32_t *key_next, float *ax, float *ay, float *az, const SceneTable &t, const Consts &k)
    {
        const dim3 g(blocks_for(t.nslots, T));
        if (ax)
            hipLaunchKernelGGL((k_forces_fused<T, U, C, kStall, kQueue, kMode, true, kRef>), g, dim3(T), 0, st, p, cell, start, f,
                               nl, out, key_next, ax, ay, az, t, k);
        else
            hipLaunchKernelGGL((k_forces_fused<T, U, C, kStall, kQueue, kMode, false, kRef>), g, dim3(T), 0, st, p, cell, start, f,
                               nl, out, key_next, ax, ay, az, t, k);
    }
    static void fused(hipStream_t st, const ParticleSoA &p, const uint32_t *cell, const uint32_t *start,
                      const DensityFields &f, const NeighbourList &nl, const ParticleSoA &out, uint32_t *key_next,
                      float *ax, float *ay, float *az, const SceneTable &t, const Consts &k)
    {
        if (nl.mode == kNbLists)
            fused_mode<kNbLists>(st, p, cell, start, f, nl, out, key_next, ax, ay, az, t, k);
        else if (nl.mode == kNbMasks)
            fused_mode<kNbMasks>(st, p, cell, start, f, nl, out, key_next, ax, ay, az, t, k);
        else
            fused_mode<kNbWalk>(st, p, cell, start, f, nl, out, key_next, ax, ay, az, t, k);
    }
};

// X(index, threads, unroll, mask mode: LDS words per lane, stall drain policy, word queue)
#define SPH_VARIANTS(X)         \
    X(0, 256, 4, 16, 0, false)  \
    X(1, 128, 4, 16, 0, false)  \
    X(2, 64, 4, 16, 0, false)   \
    X(3, 256, 2, 16, 0, false)  \
    X(4, 64, 8, 16, 0, false)   \
    X(5, 64, 4, 32, 0, false)   \
    X(6, 64, 8, 32, 0, false)   \
    X(7, 64, 8, 32, 50, false)  \
    X(8, 64, 8, 32, 75, false)  \
    X(9, 64, 4, 32, 75, false)  \
    X(10, 64, 8, 24, 75, false) \
    X(11, 64, 8, 64, 0, false)  \
    X(12, 64, 8, 48, 0, true)   \
    X(13, 64, 8, 64, 0, true)   \
    X(14, 64, 4, 48, 0, true)   \
    X(15, 64, 8, 32, 0, true)   \
    X(16, 64, 4, 32, 0, true)   \
    X(17, 128, 4, 32, 0, true)  \
    X(18, 256, 4, 32, 0, true)  \
    X(19, 256, 8, 48, 0, true)  \
    X(20, 256, 4, 48, 0, true)

}  // namespace

#define SPH_VARIANT_ROW(i, t, u, c, st, q) {t, u, c, st, q ? 1 : 0},
const NeighbourVariant kNeighbourVariants[kNumVariants] = {SPH_VARIANTS(SPH_VARIANT_ROW)};
#undef SPH_VARIANT_ROW

// ---------------------------------------------------------------- launchers

void launch_emit(const Emitter *em, uint32_t nem, Scene *scenes, const ParticleSoA &p, uint32_t *key, uint32_t *ocell,
                 double h, hipStream_t st)
{
    if (nem == 0) return;
    hipLaunchKernelGGL(k_emit, dim3(nem), dim3(256), 0, st, em, scenes, p, key, ocell, h);
}

void launch_scatter_by_id(const ParticleSoA &p, uint32_t first, uint32_t n, float *pos, float *vel, uint32_t *bad,
                          hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_scatter_by_id, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, first, n, pos, vel, bad);
}

void launch_hash(const ParticleSoA &p, uint32_t *key, const SceneTable &t, double h, hipStream_t st)
{
    if (t.nslots == 0) return;
    hipLaunchKernelGGL(k_hash, dim3(blocks_for(t.nslots, 256)), dim3(256), 0, st, p.px, p.py, p.pz, key, t, h);
}

void launch_radix_pass(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t *keys_out, uint32_t *vals_out,
                       const SortTile *tiles, uint32_t ntiles, int shift, int bits, uint32_t *hist, uint32_t *partials,
                       hipStream_t st, uint32_t poll_limit)
{
    if (ntiles == 0) return;
    const uint32_t m = ntiles << bits;
    const uint32_t sb = scan_blocks(m);
    // more than one scan tile: one single-pass scan from hist into partials[0, m), its state after that (zeroed
    // by the count kernel; m is a multiple of 256, so the state is 8-byte aligned)
    uint32_t *offsets = sb > 1 ? partials : hist;
    unsigned long long *state = sb > 1 ? reinterpret_cast<unsigned long long *>(partials + m) : nullptr;
    const uint32_t nstate = sb > 1 ? sb + 1 : 0u;
    if (bits == kRadixBitsWide)
        hipLaunchKernelGGL(k_radix_count<kRadixBitsWide>, dim3(ntiles), dim3(kSortThreads), 0, st, keys_in, tiles,
                           shift, hist, state, nstate);
    else
        hipLaunchKernelGGL(k_radix_count<kRadixBits>, dim3(ntiles), dim3(kSortThreads), 0, st, keys_in, tiles, shift,
                           hist, state, nstate);
    if (sb == 1)
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, st, hist, m);
    else
        hipLaunchKernelGGL(k_scan_chained, dim3(sb), dim3(1024), 0, st, (const uint32_t *) hist, offsets, m, state,
                           poll_limit);
    if (bits == kRadixBitsWide)
        hipLaunchKernelGGL(k_radix_scatter<kRadixBitsWide>, dim3(ntiles), dim3(kSortThreads), 0, st, keys_in, vals_in,
                           keys_out, vals_out, tiles, shift, (const uint32_t *) offsets);
    else
        hipLaunchKernelGGL(k_radix_scatter<kRadixBits>, dim3(ntiles), dim3(kSortThreads), 0, st, keys_in, vals_in,
                           keys_out, vals_out, tiles, shift, (const uint32_t *) offsets);
}

void launch_exclusive_scan(uint32_t *data, uint32_t m, uint32_t *partials, hipStream_t st)
{
    if (m == 0) return;
    const uint32_t sb = scan_blocks(m);
    if (sb == 1) {
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, st, data, m);
    } else {
        hipLaunchKernelGGL(k_scan_reduce, dim3(sb), dim3(1024), 0, st, (const uint32_t *) data, m, partials);
        hipLaunchKernelGGL(k_scan_apply, dim3(sb), dim3(1024), 0, st, data, m, (const uint32_t *) partials);
    }
}

void launch_ref_prepare(const uint32_t *ocell, uint32_t *gold, const SceneTable &t, hipStream_t st)
{
    if (t.nslots == 0) return;
    hipLaunchKernelGGL(k_ref_prepare, dim3(blocks_for(t.nslots, 256)), dim3(256), 0, st, ocell, gold, t);
}

void launch_ref_counts(const RefBuffers &r, const uint32_t *gold, const uint32_t *gnew, uint32_t n, uint32_t ncells,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_ref_zero, dim3(blocks_for(ncells + 1, 256)), dim3(256), 0, st, r.n_early, r.n_arrive,
                       r.n_depart, r.nstay, ncells + 1);
    if (n) hipLaunchKernelGGL(k_ref_count, dim3(blocks_for(n, 256)), dim3(256), 0, st, gold, gnew, n, r.n_early,
                              r.n_arrive, r.n_depart);
    launch_exclusive_scan(r.n_early, ncells + 1, r.partials, st);
    launch_exclusive_scan(r.n_depart, ncells + 1, r.partials, st);
}

void launch_ref_bucket(const RefBuffers &r, uint32_t ncells, const uint32_t *ostart, const uint32_t *perm,
                       const uint32_t *gnew, hipStream_t st)
{
    hipLaunchKernelGGL(k_ref_bucket, dim3(blocks_for(ncells, 256)), dim3(256), 0, st, ncells, ostart, perm, gnew,
                       (const uint32_t *) r.n_early, r.work, r.rk, r.epos, r.nstay);
}

void launch_ref_movers(const RefBuffers &r, const uint32_t *gold, const uint32_t *gnew, uint32_t n, uint32_t ncells,
                       hipStream_t st)
{
    if (n) hipLaunchKernelGGL(k_ref_movers, dim3(blocks_for(n, 256)), dim3(256), 0, st, gold, gnew,
                              (const uint32_t *) r.rk, (const uint32_t *) r.n_depart, n, ncells, r.mkey, r.mval);
}

void launch_ref_place(const RefBuffers &r, const uint32_t *skey, const uint32_t *sval, uint32_t n, uint32_t ncells,
                      uint32_t *start, hipStream_t st)
{
    // r.n_depart (scanned departures) is free again: it holds the arrival starts
    launch_cell_start(skey, n, ncells, r.n_depart, st);
    if (n) hipLaunchKernelGGL(k_ref_place, dim3(blocks_for(n, 256)), dim3(256), 0, st, skey, sval,
                              (const uint32_t *) r.n_depart, (const uint32_t *) r.n_early, (const uint32_t *) r.epos,
                              (const uint32_t *) r.nstay, n, ncells, r.rk);
    hipLaunchKernelGGL(k_ref_counts, dim3(blocks_for(ncells + 1, 256)), dim3(256), 0, st, (const uint32_t *) r.nstay,
                       (const uint32_t *) r.n_arrive, (const uint32_t *) r.n_early, ncells, start);
    launch_exclusive_scan(start, ncells + 1, r.partials, st);
}

void launch_ref_final(const RefBuffers &r, const uint32_t *gnew, const uint32_t *start, uint32_t *skey, uint32_t *perm,
                      uint32_t *ocell, const SceneTable &t, hipStream_t st)
{
    if (t.nslots == 0) return;
    hipLaunchKernelGGL(k_ref_final, dim3(blocks_for(t.nslots, 256)), dim3(256), 0, st, gnew, (const uint32_t *) r.rk,
                       start, skey, perm, ocell, t);
}

void launch_cell_start(const uint32_t *sorted_keys, uint32_t n, uint32_t ncells, uint32_t *start, hipStream_t st)
{
    hipLaunchKernelGGL(k_cell_start, dim3(blocks_for(n + 1, 256)), dim3(256), 0, st, sorted_keys, n, ncells,
                       start);
}

void launch_reorder(const ParticleSoA &src, const ParticleSoA &dst, const uint32_t *perm, uint32_t n,
                    hipStream_t st)
{
    if (n == 0) return;
    if (n < kReorderWideSlots)
        hipLaunchKernelGGL(k_reorder<8>, dim3(blocks_for(n, 256 * 8)), dim3(256), 0, st, src, dst, perm, n);
    else
        hipLaunchKernelGGL(k_reorder<4>, dim3(blocks_for(n, 256 * 4)), dim3(256), 0, st, src, dst, perm, n);
}

// The reference order (kRef) is compiled in one shape, variant 12's.
using RefKernels = NeighbourKernels<64, 8, 48, 0, true, true>;

void launch_density(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start, const DensityFields &f,
                    const NeighbourList &nl, const SceneTable &t, const Consts &k, int variant, bool ref,
                    hipStream_t st)
{
    if (t.nslots == 0) return;
    if (ref) {
        RefKernels::density(st, p, cell, start, f, nl, t, k);
        return;
    }
    switch (variant) {
#define SPH_DENSITY_CASE(i, t_, u, c, stl, q) \
    case i: NeighbourKernels<t_, u, c, stl, q>::density(st, p, cell, start, f, nl, t, k); break;
        SPH_VARIANTS(SPH_DENSITY_CASE)
#undef SPH_DENSITY_CASE
    default: NeighbourKernels<64, 8, 32, 0, false>::density(st, p, cell, start, f, nl, t, k); break;
    }
}

void launch_forces(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start, const DensityFields &f,
                   const NeighbourList &nl, float *ax, float *ay, float *az, const SceneTable &t, const Consts &k,
                   bool ref, hipStream_t st)
{
    if (t.nslots == 0) return;
    if (ref)
        hipLaunchKernelGGL(k_forces<true>, dim3(blocks_for(t.nslots, kPhaseThreads)), dim3(kPhaseThreads), 0, st, p,
                           cell, start, f, nl, ax, ay, az, t, k);
    else
        hipLaunchKernelGGL(k_forces<false>, dim3(blocks_for(t.nslots, kPhaseThreads)), dim3(kPhaseThreads), 0, st, p,
                           cell, start, f, nl, ax, ay, az, t, k);
}

void launch_collide(const ParticleSoA &p, float *ax, float *ay, float *az, const SceneTable &t, const Consts &k,
                    hipStream_t st)
{
    if (t.nslots == 0) return;
    hipLaunchKernelGGL(k_collide, dim3(blocks_for(t.nslots, 256)), dim3(256), 0, st, p, ax, ay, az, t, k);
}

void launch_integrate(const ParticleSoA &p, const float *ax, const float *ay, const float *az, const ParticleSoA &out,
                      uint32_t *key_next, const SceneTable &t, const Consts &k, hipStream_t st)
{
    if (t.nslots == 0) return;
    hipLaunchKernelGGL(k_integrate, dim3(blocks_for(t.nslots, 256)), dim3(256), 0, st, p, ax, ay, az, out, key_next,
                       t, k);
}

void launch_forces_fused(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start, const DensityFields &f,
                         const NeighbourList &nl, const ParticleSoA &out, uint32_t *key_next, float *ax, float *ay,
                         float *az, const SceneTable &t, const Consts &k, int variant, bool ref, hipStream_t st)
{
    if (t.nslots == 0) return;
    if (ref) {
        RefKernels::fused(st, p, cell, start, f, nl, out, key_next, ax, ay, az, t, k);
        return;
    }
    switch (variant) {
#define SPH_FUSED_CASE(i, t_, u, c, stl, q) \
    case i: NeighbourKernels<t_, u, c, stl, q>::fused(st, p, cell, start, f, nl, out, key_next, ax, ay, az, t, k); break;
        SPH_VARIANTS(SPH_FUSED_CASE)
#undef SPH_FUSED_CASE
    default: NeighbourKernels<64, 8, 32, 0, false>::fused(st, p, cell, start, f, nl, out, key_next, ax, ay, az, t, k); break;
    }
}

}  // namespace sph
