// sph_kernels.h -- launch wrappers of the gfx950 grid-SPH kernels (sph_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_common.h"

namespace sph {

// Structure-of-arrays particle buffers in HBM, in the current array order.
struct ParticleSoA {
    float *px, *py, *pz;
    float *vx, *vy, *vz;
    uint32_t *id;
};

// Per-particle results of the density pass that the force pass reads.
struct DensityFields {
    float *rho;       // density
    uint32_t *ncount; // accepted neighbours incl. self (diagnostics; may be null)
    // Everything the force pass reads about a neighbour, one 40-byte record per
    // slot (kRecWords 32-bit words) read with three loads (16 + 16 + 8 bytes):
    // x, y, z, vx, vy, vz, then as doubles p/rho^2 (src/CCPUParticleSimulator.cpp:181)
    // and 1/rho (the exact-division helper of the force pass).
    uint32_t *rec;
};

// Accepted neighbours of every particle except itself, in canonical order, as
// slots relative to the first slot of the particle's scene. Blocked by wave:
// the k-th neighbour of slot i is idx[((i / 64) * kmax + k) * 64 + i % 64], so
// that the 64 particles of a wave keep their lists in one kmax x 64 block and a
// wave reads one level of all its lists with one coalesced load. cnt[i] is the
// true count; particles with cnt > kmax have no usable list and the force pass
// re-walks their cells.
//
// Or (mode kNbMasks) the density pass's per-row accepted-candidate masks
// (`masks`, see the kernels), with cnt[i] = 1 for a particle they cannot hold.
enum NeighbourMode : int { kNbWalk = 0, kNbLists = 1, kNbMasks = 2 };

constexpr uint32_t kRecWords = 10;

// Row masks: each of a particle's 9 rows holds up to kMaskBits candidates in
// kMaskWords words of kMaskWordBits bits. 512 candidates cover the densest rows
// measured (about 325 in a splash of the 1M dam break, 312 in the settled
// fountain pool); a longer run makes the particle walk its cells in the force
// pass, and the wave carrying it finishes last.
constexpr uint32_t kMaskWordBits = 32;
constexpr uint32_t kMaskWords = 16;
constexpr uint32_t kMaskBits = kMaskWordBits * kMaskWords;

struct NeighbourList {
    uint32_t *idx;
    uint32_t *cnt;
    uint32_t kmax;
    int mode;            // NeighbourMode
    uint32_t *masks;     // mask mode: 32-bit words of accepted-candidate bits
};

__host__ __device__ inline size_t list_slot(const NeighbourList &nl, uint32_t i)
{
    return (size_t) (i >> 6) * nl.kmax * 64 + (i & 63u);
}

// Radix digits of 8 bits, or of 10 bits where that saves a pass (keys of 17-20
// bits: two passes instead of three).
constexpr int kRadixBits = 8;
constexpr int kRadixBitsWide = 10;
constexpr int kRadixBucketsMax = 1 << kRadixBitsWide;
// Sort tiles of 4,096 keys, 4 per thread of a 1,024-thread block: the passes are
// latency-bound (one tile per CU at 1M keys), and 16 waves a tile ranking 4
// rounds each measured faster than 4 waves ranking 16.
constexpr int kSortThreads = 1024;
constexpr int kSortItems = 4;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kScanTile = 4096;  // elements per 1024-thread scan block

inline uint32_t sort_blocks(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }
inline uint32_t scan_blocks(uint32_t m) { return (m + kScanTile - 1) / kScanTile; }

// The tunable shape of the two neighbour kernels: workgroup size, candidates
// tested per lane per trip, and how the force pass reads the row masks: with
// `pairs` 32-bit LDS words per lane either as a pair list with a drain policy
// (stall 0: drain when any list is full; P: full lanes wait until P percent of
// the wave waits) or, with `queue`, as a queue of pairs / 2 mask words.
// Results do not depend on the variant.
struct NeighbourVariant {
    int threads;
    int unroll;
    int pairs;
    int stall;
    int queue;
};
constexpr int kNumVariants = 21;
constexpr int kDefaultVariant = 12;  // 64 threads, 8 candidates, a 24-entry word queue: fastest measured on the 1M dam break
extern const NeighbourVariant kNeighbourVariants[kNumVariants];

// Emitters (one workgroup each) add particles at the start of a step and update
// the live counts in the device scene table.
void launch_emit(const Emitter *em, uint32_t nem, Scene *scenes, const ParticleSoA &p, uint32_t *key, uint32_t *ocell,
                 double h, hipStream_t st);

// Cell key of every slot below t.nslots: the hash of src/CCPUParticleSimulator.cpp:46-69
// for live particles, the scene's dead cell for unused slots.
void launch_hash(const ParticleSoA &p, uint32_t *key, const SceneTable &t, double h, hipStream_t st);

// A tile of the radix sort: up to kSortTile slots [start, start + count) of one
// segment (a scene or a free range) whose cells start at coff. The digits are
// those of the segment-local key (key - coff). The segment's tiles are first ..
// first + ntiles - 1 of the launch, this one `index` among them; in a pass of B
// digits the tile's count of digit d goes to hist[first * B + d * ntiles +
// index], so the histograms are ordered segment, digit, tile: their exclusive
// scan places each segment's items in the segment's own slots, sorted by local
// key, which is the order of a sort of the global keys (segments never share a
// slot, and the segments' cells come in slot order).
struct SortTile {
    uint32_t start, count, coff, first, ntiles, index, pad[2];
};

// Digit widths of a sort of `bits`-bit keys: all kRadixBits, or all
// kRadixBitsWide where that takes fewer passes. Returns the pass count.
inline int radix_plan(int bits, int *digit_bits)
{
    const int narrow = (bits + kRadixBits - 1) / kRadixBits, wide = (bits + kRadixBitsWide - 1) / kRadixBitsWide;
    const int passes = wide < narrow ? wide : narrow, w = wide < narrow ? kRadixBitsWide : kRadixBits;
    for (int p = 0; p < passes; ++p) digit_bits[p] = w;
    return passes;
}

// Tiles of kSortTile slots over segments; returns the tile count (at most
// sort_blocks(nslots) + the number of segments).
inline uint32_t max_sort_tiles(uint32_t pool, uint32_t max_segments) { return sort_blocks(pool) + max_segments; }

// Polls of the histogram scan's look-back before a block adds up its
// predecessors' raw counts itself (the result is the same either way).
constexpr uint32_t kScanPollLimit = 1u << 22;

// One stable LSD radix pass over `bits` (kRadixBits or kRadixBitsWide) bits of
// the segment-local keys starting at `shift`. vals_in == nullptr means the
// identity (val = slot). hist holds m = 2^bits * ntiles entries. When that is
// more than one scan block, partials (8-byte aligned) receives the scanned
// offsets (m words) followed by the single-pass scan's state (2 *
// (scan_blocks(m) + 1) words); otherwise it is not used. poll_limit: see
// kScanPollLimit (0: no look-back).
void launch_radix_pass(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t *keys_out,
                       uint32_t *vals_out, const SortTile *tiles, uint32_t ntiles, int shift, int bits,
                       uint32_t *hist, uint32_t *partials, hipStream_t st, uint32_t poll_limit = kScanPollLimit);

// In-place exclusive scan of m values; partials holds scan_blocks(m) entries.
void launch_exclusive_scan(uint32_t *data, uint32_t m, uint32_t *partials, hipStream_t st);

// Buffers of the reference-order grid phase (per slot: rk, epos, mkey, mval; 2 per
// slot: work; per cell + 1: n_early, n_arrive, n_depart, nstay; partials for scans
// of cell arrays).
struct RefBuffers {
    uint32_t *rk, *epos, *mkey, *mval, *work;
    uint32_t *n_early, *n_arrive, *n_depart, *nstay;
    uint32_t *partials;
};

// The reference-order grid phase in steps (see sph_kernels.hip), over n slots and
// ncells global cells: old cells as keys (gold), the counts and their scans, the
// per-cell swap-remove scans over the old buckets (sorted by old cell: ostart,
// perm), the movers in push order (keys mkey, values mval, to be sorted by key),
// their placing and the new cell table (start), the permutation and keys.
void launch_ref_prepare(const uint32_t *ocell, uint32_t *gold, const SceneTable &t, hipStream_t st);
void launch_ref_counts(const RefBuffers &r, const uint32_t *gold, const uint32_t *gnew, uint32_t n, uint32_t ncells,
                       hipStream_t st);
void launch_ref_bucket(const RefBuffers &r, uint32_t ncells, const uint32_t *ostart, const uint32_t *perm,
                       const uint32_t *gnew, hipStream_t st);
void launch_ref_movers(const RefBuffers &r, const uint32_t *gold, const uint32_t *gnew, uint32_t n, uint32_t ncells,
                       hipStream_t st);
void launch_ref_place(const RefBuffers &r, const uint32_t *skey, const uint32_t *sval, uint32_t n, uint32_t ncells,
                      uint32_t *start, hipStream_t st);
void launch_ref_final(const RefBuffers &r, const uint32_t *gnew, const uint32_t *start, uint32_t *skey, uint32_t *perm,
                      uint32_t *ocell, const SceneTable &t, hipStream_t st);

// start[c] = first slot whose sorted key is >= c, c in [0, ncells].
void launch_cell_start(const uint32_t *sorted_keys, uint32_t n, uint32_t ncells, uint32_t *start,
                       hipStream_t st);

// dst[k] = src[perm[k]] for positions, velocities and ids.
void launch_reorder(const ParticleSoA &src, const ParticleSoA &dst, const uint32_t *perm, uint32_t n,
                    hipStream_t st);

// pos[3*id + c] (and vel) for slots first .. first+n-1; ids outside 0..n-1 set *bad.
void launch_scatter_by_id(const ParticleSoA &p, uint32_t first, uint32_t n, float *pos, float *vel, uint32_t *bad,
                          hipStream_t st);

// Density pass: density, p/rho^2, 1/rho and the neighbour lists. `ref`: the
// scenes use the reference order (kOrderReference; one kernel shape).
void launch_density(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start, const DensityFields &f,
                    const NeighbourList &nl, const SceneTable &t, const Consts &k, int variant, bool ref,
                    hipStream_t st);

// Pressure + viscosity + gravity acceleration into acc (the updateForces phase).
void launch_forces(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start, const DensityFields &f,
                   const NeighbourList &nl, float *ax, float *ay, float *az, const SceneTable &t, const Consts &k,
                   bool ref, hipStream_t st);

// acc += wall response (the updateCollisions phase).
void launch_collide(const ParticleSoA &p, float *ax, float *ay, float *az, const SceneTable &t, const Consts &k,
                    hipStream_t st);

// Integrate from acc into out and hash the new positions into key_next (the integrate phase).
void launch_integrate(const ParticleSoA &p, const float *ax, const float *ay, const float *az,
                      const ParticleSoA &out, uint32_t *key_next, const SceneTable &t, const Consts &k,
                      hipStream_t st);

// Fused forces + walls + integrate + next hash. acc pointers may be null (then not stored).
void launch_forces_fused(const ParticleSoA &p, const uint32_t *cell, const uint32_t *start,
                         const DensityFields &f, const NeighbourList &nl, const ParticleSoA &out,
                         uint32_t *key_next, float *ax, float *ay, float *az, const SceneTable &t, const Consts &k,
                         int variant, bool ref, hipStream_t st);

}  // namespace sph
