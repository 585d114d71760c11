// sph_noalloc.hip -- sph_step_noalloc: the fused step of sph_step over particle
// arrays the caller owns in device memory, for callers that cannot allocate
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc). Every buffer beyond the caller's arrays -- the second particle
// buffer, sort keys and permutation, radix histograms, the cell table, the
// density records -- lives in one file-scope device array, `scratch`, of a fixed
// size per GPU. The kernels are the ones sph_step launches, so the results are
// bitwise sph_step's.
//
// `scratch` is a host-side stand-in in host code: its device address is taken
// once per device with hipGetSymbolAddress (the first call on that device) and
// only that address is ever passed to kernels or copies. Every call on one GPU
// shares its array and finds the previous call's contents; nothing in it has to
// start at zero (each kernel writes what the next one reads: the hash or the
// previous step's fused tail the keys, the count kernel every histogram entry it
// covers, the cell-table kernel every entry up to the dead cell, the density pass
// every live record). Two calls that overlap on one GPU race on the array.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/sph_hip.h"
#include "sph_common.h"
#include "sph_kernels.h"

namespace sph {
sph_status report(sph_status code, const std::string &msg);   // sets sph_last_error() (sph_capi.hip)
sph_status host_exception() noexcept;
}  // namespace sph

namespace {

using namespace sph;

#define SPH_NOALLOC_TRY(expr)              \
    do {                                   \
        const sph_status s_ = (expr);      \
        if (s_ != SPH_OK) return s_;       \
    } while (0)

constexpr uint32_t kMaxN = SPH_NOALLOC_MAX_PARTICLES;
constexpr uint32_t kMaxGridCells = SPH_NOALLOC_MAX_CELLS;
constexpr uint32_t kMaxTiles = (kMaxN + kSortTile - 1) / kSortTile;
constexpr uint32_t kHistEntries = (1u << kRadixBits) * kMaxTiles;   // 8-bit passes only
static_assert(kHistEntries <= kScanTile, "the histograms of one call are scanned by one block");

constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Byte offsets of the sub-buffers inside `scratch`, each 256-byte aligned.
constexpr size_t kOffScene = 0;
constexpr size_t kOffTiles = kOffScene + align256(sizeof(Scene));
constexpr size_t kOffHist = kOffTiles + align256(kMaxTiles * sizeof(SortTile));
constexpr size_t kOffPartials = kOffHist + align256((size_t) kHistEntries * 4);
constexpr size_t kOffSoA = kOffPartials + 256;                 // px py pz vx vy vz id, kMaxN each
constexpr size_t kOffKeys = kOffSoA + 7 * align256((size_t) kMaxN * 4);   // key_h key_t key_s val_t val_s
constexpr size_t kOffRho = kOffKeys + 5 * align256((size_t) kMaxN * 4);
constexpr size_t kOffCnt = kOffRho + align256((size_t) kMaxN * 4);
constexpr size_t kOffRec = kOffCnt + align256((size_t) kMaxN * 4);
constexpr size_t kOffStart = kOffRec + align256((size_t) kMaxN * 4 * kRecWords);
constexpr size_t kScratchBytes = kOffStart + align256(((size_t) kMaxGridCells + 2) * 4);
static_assert(kScratchBytes == SPH_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of `scratch` on each device, looked up once per device.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch[kMaxDevices] = {};

sph_status scratch_for_device(unsigned char **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return report(SPH_ERR_NO_DEVICE, std::string("sph_step_noalloc: hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return report(SPH_ERR_NO_DEVICE, "sph_step_noalloc: device index " + std::to_string(dev) + " out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, dev);
        if (e != hipSuccess)
            return report(SPH_ERR_NO_DEVICE, std::string("sph_step_noalloc: hipGetDeviceProperties: ") + hipGetErrorString(e));
        if (std::string(prop.gcnArchName).compare(0, 6, "gfx950") != 0)
            return report(SPH_ERR_NO_DEVICE, std::string("libsph_hip.so is built for gfx950, device is ") + prop.gcnArchName);
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return report(SPH_ERR_HIP, std::string("sph_step_noalloc: hipGetSymbolAddress(scratch): ") +
                                           hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = static_cast<unsigned char *>(g_scratch[dev]);
    return SPH_OK;
}

// Writes the scene record and the sort tiles of one call into the scratch
// array on the stream (no host-to-device copy, so the call can be captured).
__global__ __launch_bounds__(64) void k_noalloc_setup(Scene s, uint32_t ntiles, Scene *scene, SortTile *tiles)
{
    const uint32_t t = threadIdx.x;
    if (t == 0) *scene = s;
    for (uint32_t k = t; k < ntiles; k += 64u) {
        SortTile tl{};
        tl.start = k * (uint32_t) kSortTile;
        tl.count = min((uint32_t) kSortTile, s.n - tl.start);
        tl.coff = 0;
        tl.first = 0;
        tl.ntiles = ntiles;
        tl.index = k;
        tiles[k] = tl;
    }
}

int key_bits(uint32_t maxkey)
{
    int b = 1;
    while (b < 32 && (maxkey >> b) != 0) ++b;
    return b;
}

sph_status check_launch(const char *what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return report(SPH_ERR_HIP, std::string("sph_step_noalloc: ") + what + ": " + hipGetErrorString(e));
    return SPH_OK;
}

sph_status step_noalloc(const sph_params *prm, const sph_device_state *st, uint32_t n, uint32_t nsteps, hipStream_t stream)
{
    if (!prm || !st) return report(SPH_ERR_INVALID, "sph_step_noalloc: null argument");
    if (!st->px || !st->py || !st->pz || !st->vx || !st->vy || !st->vz || !st->id)
        return report(SPH_ERR_INVALID, "sph_step_noalloc: null particle array");
    for (int a = 0; a < 3; ++a)
        if (!(prm->box[a] > 0.0f) || !(prm->box[a] < 1e30f))
            return report(SPH_ERR_INVALID, "sph_step_noalloc: box extents must be positive and finite");
    if (!(prm->dt > 0.0f) || !(prm->dt < 1e30f)) return report(SPH_ERR_INVALID, "sph_step_noalloc: dt must be > 0");
    if (n > kMaxN)
        return report(SPH_ERR_CAPACITY, "sph_step_noalloc: " + std::to_string(n) + " particles exceed the maximum of " +
                                            std::to_string(kMaxN));
    Scene s{};
    s.rx = sph_grid_res(prm->box[0]);
    s.ry = sph_grid_res(prm->box[1]);
    s.rz = sph_grid_res(prm->box[2]);
    const uint64_t cells = (uint64_t) s.rx * (uint64_t) s.ry * (uint64_t) s.rz;
    if (s.rx <= 0 || s.ry <= 0 || s.rz <= 0 || cells > kMaxGridCells)
        return report(SPH_ERR_CAPACITY, "sph_step_noalloc: a grid of " + std::to_string(cells) +
                                            " cells exceeds the maximum of " + std::to_string(kMaxGridCells));
    if (n == 0 || nsteps == 0) return SPH_OK;
    s.ncells = (uint32_t) cells;
    s.coff = 0;
    s.poff = 0;
    s.cap = n;
    s.n = n;
    s.bx = prm->box[0];
    s.by = prm->box[1];
    s.bz = prm->box[2];
    s.gx = prm->gravity[0];
    s.gy = prm->gravity[1];
    s.gz = prm->gravity[2];
    s.dt = prm->dt;

    unsigned char *base = nullptr;
    SPH_NOALLOC_TRY(scratch_for_device(&base));
    auto u32 = [&](size_t off) { return reinterpret_cast<uint32_t *>(base + off); };
    auto f32 = [&](size_t off) { return reinterpret_cast<float *>(base + off); };
    Scene *scene_d = reinterpret_cast<Scene *>(base + kOffScene);
    SortTile *tiles_d = reinterpret_cast<SortTile *>(base + kOffTiles);
    uint32_t *hist = u32(kOffHist), *partials = u32(kOffPartials);
    const size_t arr = align256((size_t) kMaxN * 4);
    uint32_t *key_h = u32(kOffKeys), *key_t = u32(kOffKeys + arr), *key_s = u32(kOffKeys + 2 * arr);
    uint32_t *val_t = u32(kOffKeys + 3 * arr), *val_s = u32(kOffKeys + 4 * arr);
    uint32_t *cstart = u32(kOffStart);

    // the caller's arrays (buffer 0) and the second buffer in scratch (buffer 1)
    ParticleSoA buf[2];
    buf[0] = {st->px, st->py, st->pz, st->vx, st->vy, st->vz, st->id};
    buf[1] = {f32(kOffSoA), f32(kOffSoA + arr), f32(kOffSoA + 2 * arr), f32(kOffSoA + 3 * arr),
              f32(kOffSoA + 4 * arr), f32(kOffSoA + 5 * arr), u32(kOffSoA + 6 * arr)};
    auto soa = [&](int pv, int idb) {
        ParticleSoA p = buf[pv];
        p.id = buf[idb].id;
        return p;
    };
    DensityFields f{};
    f.rho = f32(kOffRho);
    f.ncount = nullptr;
    f.rec = u32(kOffRec);
    NeighbourList nl{};
    nl.idx = nullptr;
    nl.cnt = u32(kOffCnt);
    nl.kmax = 0;
    nl.mode = kNbWalk;
    nl.masks = nullptr;
    SceneTable tab{scene_d, nullptr, n};
    const Consts k = make_consts();

    const uint32_t ntiles = (n + kSortTile - 1) / kSortTile;
    hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(64), 0, stream, s, ntiles, scene_d, tiles_d);
    SPH_NOALLOC_TRY(check_launch("set-up launch"));
    const int passes = (key_bits(s.ncells) + kRadixBits - 1) / kRadixBits;   // keys up to the dead cell ncells
    int cur = 0, idc = 0;
    launch_hash(soa(cur, idc), key_h, tab, k.h, stream);
    for (uint32_t step = 0; step < nsteps; ++step) {
        // grid phase: stable radix sort of (cell, slot), cell table, gather (as run_grid)
        const uint32_t *kin = key_h, *vin = nullptr;
        for (int p = 0; p < passes; ++p) {
            const bool to_s = ((passes - 1 - p) % 2) == 0;
            uint32_t *kout = to_s ? key_s : key_t, *vout = to_s ? val_s : val_t;
            launch_radix_pass(kin, vin, kout, vout, tiles_d, ntiles, p * kRadixBits, kRadixBits, hist, partials,
                              stream);
            kin = kout;
            vin = vout;
        }
        launch_cell_start(key_s, n, s.ncells + 1, cstart, stream);
        launch_reorder(soa(cur, idc), soa(1 - cur, 1 - idc), val_s, n, stream);
        cur = 1 - cur;
        idc = 1 - idc;
        // density, then forces + walls + integrate + the next step's keys
        launch_density(soa(cur, idc), key_s, cstart, f, nl, tab, k, kDefaultVariant, false, stream);
        launch_forces_fused(soa(cur, idc), key_s, cstart, f, nl, soa(1 - cur, idc), key_h, nullptr, nullptr, nullptr,
                            tab, k, kDefaultVariant, false, stream);
        cur = 1 - cur;
        SPH_NOALLOC_TRY(check_launch("step launch"));
    }
    // positions and velocities end in the caller's arrays (two flips a step); the
    // ids flip once a step
    if (idc != 0) {
        const hipError_t e = hipMemcpyAsync(st->id, buf[1].id, (size_t) n * 4, hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) return report(SPH_ERR_HIP, std::string("sph_step_noalloc: id copy: ") + hipGetErrorString(e));
    }
    return SPH_OK;
}

}  // namespace

extern "C" sph_status sph_step_noalloc(const sph_params *params, const sph_device_state *state, uint32_t n,
                                       uint32_t nsteps, void *stream)
{
    try {
        return step_noalloc(params, state, n, nsteps, static_cast<hipStream_t>(stream));
    } catch (...) {
        return sph::host_exception();
    }
}
