// sph_capi.hip -- the C ABI of libsph_hip.so (declared in include/sph_hip.h).
//
// A context owns a pool of particle slots shared by one or more independent
// scenes, its device buffers (SoA particle state ping-pong, sort buffers, the
// global cell table, density fields) and one HIP stream. Each scene holds a
// segment of slots that starts on a multiple of kGroupSlots and a segment of the
// global cell ids (its grid plus one dead cell for unused slots); free ranges of
// the pool are scenes without cells. The device reads the segment table, so one
// launch per stage advances every scene. Every entry point reports failures as a
// status plus a thread-local message and never throws.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/sph_hip.h"
#include "sph_common.h"
#include "sph_kernels.h"

namespace {

thread_local std::string g_last_error = "";

sph_status fail(sph_status code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define SPH_HIP(call)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return fail(SPH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));                 \
    } while (0)

#define SPH_TRY(expr)                                                                                    \
    do {                                                                                                 \
        sph_status s_ = (expr);                                                                          \
        if (s_ != SPH_OK) return s_;                                                                     \
    } while (0)

constexpr int kProfileSlots = 256;          // event pairs buffered before a forced drain
constexpr uint32_t kMaxSceneSlots = 1u << 26;  // 40-byte records of a scene stay below 4 GiB (buffer offsets)
constexpr uint32_t kMaxPool = 1u << 31;
constexpr uint32_t kNoScene = 0xFFFFFFFFu;

uint32_t round_up(uint32_t v, uint32_t m)
{
    return (uint32_t) (((uint64_t) v + m - 1) / m * m);
}

struct SceneRec {
    bool used = false;
    uint32_t req_cap = 0;      // capacity asked for; the segment may be longer (group rounding)
    sph::Scene s{};            // host copy; for a scene with an emitter, n is refreshed from the device
    bool has_emitter = false;
    sph::Emitter em{};
};

}  // namespace

struct sph_ctx {
    uint32_t flags = 0;
    sph::Consts k{};
    int device = 0;
    hipStream_t stream = nullptr;

    // scenes and the device segment table
    uint32_t pool = 0;
    std::vector<SceneRec> scenes;            // by scene id
    std::vector<sph::Scene> table_h;         // segments in slot order (scenes and free ranges)
    std::vector<uint32_t> table_sid;         // scene id of each segment, kNoScene for a free range
    std::vector<uint32_t> group_h;           // segment of each group of kGroupSlots slots
    std::vector<sph::Emitter> emit_h;        // emitters with their segment index
    sph::Scene *table_d = nullptr;
    uint32_t *group_d = nullptr;
    // radix-sort tiles over the segments below nslots, and the bits of the
    // largest segment-local key
    std::vector<sph::SortTile> tiles_h;
    sph::SortTile *tiles_d = nullptr;
    uint32_t tiles_cap = 0;
    int sort_bits = 1;
    sph::Emitter *emit_d = nullptr;
    uint32_t total_cells = 0;                // cells of all segments (grid cells + one dead cell each)
    uint32_t cstart_cap = 0;                 // entries allocated in cstart
    uint32_t nslots = 0;                     // slots the step covers
    bool layout_dirty = true;                // table_h/device table must be rebuilt before the next launch
    bool counts_stale = false;               // emitters ran since the host copies of n were refreshed

    // particle state: pos/vel ping-pong and ids, in the current array order
    float *pos[2][3] = {};
    float *vel[2][3] = {};
    uint32_t *ids[2] = {};
    int cur = 0;      // which pos/vel buffer is current
    int id_cur = 0;   // which id buffer is current

    // sort: hash keys (pre-sort order), ping-pong keys/vals, sorted result
    uint32_t *key_h = nullptr, *key_t = nullptr, *key_s = nullptr;
    uint32_t *val_t = nullptr, *val_s = nullptr;
    uint32_t *hist = nullptr;
    uint32_t *partials = nullptr;
    uint32_t *cstart = nullptr;

    // reference order (SPH_FLAG_REFERENCE_ORDER): each live slot's cell from the
    // last grid phase (local traversal number), the old cells as global keys, the
    // grid phase's work buffers and a one-segment tiling of the slots
    uint32_t *ocell = nullptr, *gold = nullptr;
    sph::RefBuffers ref{};
    uint32_t ref_cells_cap = 0;
    std::vector<sph::SortTile> tiles_flat_h;
    sph::SortTile *tiles_flat_d = nullptr;
    int flat_bits = 1;

    float *rho = nullptr;
    uint32_t *rec = nullptr;
    float *acc[3] = {};
    uint32_t *ncount = nullptr;
    uint32_t *nlist = nullptr;   // neighbour lists, max_neighbours per slot, wave-blocked
    uint32_t *nlist_cnt = nullptr;
    uint32_t kmax = 0;           // list capacity allocated
    uint32_t *masks = nullptr;   // row masks (mode SPH_NEIGHBOURS_MASKS), allocated on first use
    int nb_mode = SPH_NEIGHBOURS_WALK;
    int variant = sph::kDefaultVariant;

    bool keys_valid = false;   // key_h holds the cells of the current positions
    int next_phase = SPH_PHASE_GRID;
    uint64_t steps_done = 0;
    long fault_at_step = -1;
    uint32_t scan_poll_limit = sph::kScanPollLimit;   // SPH_SCAN_POLL_LIMIT (tests)

    // graphs: one executable step per (cur, id_cur) parity
    bool use_graphs = false;
    hipGraphExec_t graph[2][2] = {};

    // readback: pinned host buffers the device writes by id
    float *rb_pos = nullptr, *rb_vel = nullptr;
    size_t rb_cap = 0;
    uint32_t *rb_bad = nullptr;

    // profiling
    hipEvent_t ev[kProfileSlots][2] = {};
    int ev_stage[kProfileSlots] = {};
    uint64_t ev_step[kProfileSlots] = {};
    uint32_t event_stride = 10;
    std::vector<sph_profile_event> events;
    int ev_used = 0;
    float stage_ms[SPH_STAGE_COUNT] = {};
    uint32_t stage_launches[SPH_STAGE_COUNT] = {};
    hipEvent_t span[2] = {};

    std::vector<void *> allocations;

    sph::DensityFields fields(bool diag) const
    {
        sph::DensityFields f;
        f.rho = rho;
        f.ncount = diag ? ncount : nullptr;
        f.rec = rec;
        return f;
    }

    sph::NeighbourList nlists() const
    {
        sph::NeighbourList nl;
        nl.idx = nlist;
        nl.cnt = nlist_cnt;
        nl.kmax = kmax;
        nl.mode = nb_mode;
        nl.masks = masks;
        return nl;
    }

    sph::ParticleSoA soa(int which, int idw) const
    {
        sph::ParticleSoA s;
        s.px = pos[which][0]; s.py = pos[which][1]; s.pz = pos[which][2];
        s.vx = vel[which][0]; s.vy = vel[which][1]; s.vz = vel[which][2];
        s.id = ids[idw];
        return s;
    }

    bool ref_order() const { return (flags & SPH_FLAG_REFERENCE_ORDER) != 0; }

    sph::SceneTable table() const
    {
        sph::SceneTable t;
        t.scenes = table_d;
        t.group_scene = table_h.size() > 1 ? group_d : nullptr;
        t.nslots = nslots;
        return t;
    }
};

namespace sph {

Consts make_consts()
{
    // src/CCPUParticleSimulator.cpp:11-12, 19, 26 (h is float; h*h is a float product)
    Consts k;
    const double h = (double) kH;
    k.poly6 = 315.0 / (64.0 * M_PI * std::pow(h, 9.0));
    k.spiky = -45.0 / (M_PI * std::pow(h, 6.0));
    k.visc = 45.0 / (M_PI * std::pow(h, 6.0));
    volatile float hh = kH * kH;
    k.hsq_f = hh;
    k.hsq = (double) k.hsq_f;
    k.h = h;
    k.mass = kMass;
    k.neg_mass = -kMass;
    volatile float mm = kViscosity * kMass;
    k.mu_mass = mm;
    k.gas_k = kGasStiffness;
    k.rho0 = kRestDensity;
    k.wall_k_f = (float) kWallK;
    return k;
}

sph_status report(sph_status code, const std::string &msg)
{
    return fail(code, msg);
}

// The status of the C++ exception being handled (called in a catch block of an
// entry point, which never lets one cross the C ABI).
sph_status host_exception() noexcept
{
    sph_status code = SPH_ERR_INVALID;
    const char *msg = "host error";
    try {
        throw;
    } catch (const std::bad_alloc &) {
        code = SPH_ERR_OOM;
        msg = "host memory allocation failed";
    } catch (const std::exception &e) {
        msg = e.what();
    } catch (...) {
    }
    try {
        g_last_error = msg;
    } catch (...) {
        g_last_error.clear();
    }
    return code;
}

}  // namespace sph

namespace {

void *dalloc_raw(sph_ctx *c, size_t bytes, sph_status *st)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(bytes, 1));
    if (e != hipSuccess) {
        *st = fail(SPH_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + " B) failed: " + hipGetErrorString(e));
        return nullptr;
    }
    c->allocations.push_back(q);
    *st = SPH_OK;
    return q;
}

template <typename T>
sph_status dalloc(sph_ctx *c, T **p, size_t count)
{
    sph_status st;
    *p = (T *) dalloc_raw(c, count * sizeof(T), &st);
    return st;
}

void dfree(sph_ctx *c, void *p)
{
    if (!p) return;
    auto it = std::find(c->allocations.begin(), c->allocations.end(), p);
    if (it != c->allocations.end()) c->allocations.erase(it);
    (void) hipFree(p);
}

void drop_graphs(sph_ctx *c)
{
    for (auto &row : c->graph)
        for (auto &g : row)
            if (g) {
                (void) hipGraphExecDestroy(g);
                g = nullptr;
            }
}

void release(sph_ctx *c)
{
    if (!c) return;
    if (c->stream) (void) hipStreamSynchronize(c->stream);
    drop_graphs(c);
    if (c->rb_pos) (void) hipHostFree(c->rb_pos);
    if (c->rb_vel) (void) hipHostFree(c->rb_vel);
    for (void *p : c->allocations) (void) hipFree(p);
    c->allocations.clear();
    for (auto &pair : c->ev)
        for (auto &e : pair)
            if (e) (void) hipEventDestroy(e);
    for (auto &e : c->span)
        if (e) (void) hipEventDestroy(e);
    if (c->stream) (void) hipStreamDestroy(c->stream);
    delete c;
}

bool check_ctx(const sph_ctx *c)
{
    if (!c) {
        g_last_error = "null context";
        return false;
    }
    return true;
}

sph_status check_scene(const sph_ctx *c, uint32_t scene, const char *who)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    if (scene >= c->scenes.size() || !c->scenes[scene].used)
        return fail(SPH_ERR_INVALID, std::string(who) + ": no scene " + std::to_string(scene));
    return SPH_OK;
}

sph_status between_steps(const sph_ctx *c, const char *who)
{
    if (c->next_phase != SPH_PHASE_GRID)
        return fail(SPH_ERR_STATE, std::string(who) + ": only between steps (before the grid phase)");
    return SPH_OK;
}

// --- profiling: an event pair around every stage when SPH_FLAG_PROFILE is set.

sph_status drain_profile(sph_ctx *c)
{
    if (c->ev_used == 0) return SPH_OK;
    SPH_HIP(hipEventSynchronize(c->ev[c->ev_used - 1][1]));
    for (int i = 0; i < c->ev_used; ++i) {
        float ms = 0.0f;
        SPH_HIP(hipEventElapsedTime(&ms, c->ev[i][0], c->ev[i][1]));
        const int stage = c->ev_stage[i];
        c->stage_ms[stage] += ms;
        c->stage_launches[stage] += 1;
        const uint64_t step = c->ev_step[i];
        if (c->event_stride && step % c->event_stride == 0) {
            if (c->events.empty() || c->events.back().iteration != step) {
                sph_profile_event e{};
                e.iteration = step;
                c->events.push_back(e);
            }
            sph_profile_event &e = c->events.back();
            switch (stage) {
            case SPH_STAGE_DENSITY: e.update_density_pressure += ms; break;
            case SPH_STAGE_FORCES: e.update_forces += ms; break;
            case SPH_STAGE_COLLISIONS: e.update_collisions += ms; break;
            case SPH_STAGE_INTEGRATE: e.integrate += ms; break;
            default: e.update_grid += ms; break;   // hash, sort, cells, reorder
            }
        }
    }
    c->ev_used = 0;
    return SPH_OK;
}

struct StageTimer {
    sph_ctx *c;
    int slot = -1;
    sph_status st = SPH_OK;
    StageTimer(sph_ctx *ctx, int stage) : c(ctx)
    {
        if (!(c->flags & SPH_FLAG_PROFILE)) return;
        if (c->ev_used == kProfileSlots) st = drain_profile(c);
        if (st != SPH_OK) return;
        slot = c->ev_used++;
        c->ev_stage[slot] = stage;
        c->ev_step[slot] = c->steps_done;
        (void) hipEventRecord(c->ev[slot][0], c->stream);
    }
    ~StageTimer()
    {
        if (slot >= 0) (void) hipEventRecord(c->ev[slot][1], c->stream);
    }
};

sph_status launch_error(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SPH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return SPH_OK;
}

int key_bits(uint32_t ncells)
{
    // keys are in [0, ncells); bits to represent ncells - 1 (at least 1)
    int b = 1;
    while (b < 32 && ((ncells - 1) >> b) != 0) ++b;
    return b;
}

// --- scene layout

// Refreshes the host copies of the live counts that emitters change on the device.
sph_status sync_counts(sph_ctx *c)
{
    if (!c->counts_stale) return SPH_OK;
    SPH_HIP(hipStreamSynchronize(c->stream));
    std::vector<sph::Scene> dev(c->table_h.size());
    SPH_HIP(hipMemcpy(dev.data(), c->table_d, dev.size() * sizeof(sph::Scene), hipMemcpyDeviceToHost));
    for (size_t t = 0; t < dev.size(); ++t) {
        const uint32_t sid = c->table_sid[t];
        if (sid != kNoScene) c->scenes[sid].s.n = dev[t].n;
    }
    c->counts_stale = false;
    return SPH_OK;
}

uint32_t segment_end(const sph_ctx *c, const SceneRec &r)
{
    return r.s.poff + r.s.cap;
}

std::vector<uint32_t> scenes_by_slot(const sph_ctx *c)
{
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < c->scenes.size(); ++i)
        if (c->scenes[i].used) order.push_back(i);
    std::sort(order.begin(), order.end(),
              [&](uint32_t a, uint32_t b) { return c->scenes[a].s.poff < c->scenes[b].s.poff; });
    return order;
}

// Rebuilds the segment table from the scenes (host counts must be current) and
// uploads it with the group map and the emitters.
sph_status rebuild_layout(sph_ctx *c)
{
    SPH_TRY(sync_counts(c));
    SPH_HIP(hipStreamSynchronize(c->stream));
    c->table_h.clear();
    c->table_sid.clear();
    c->emit_h.clear();
    uint64_t coff = 0;
    uint32_t pos = 0;
    auto free_range = [&](uint32_t a, uint32_t b) {
        sph::Scene h{};
        h.coff = (uint32_t) coff;
        h.poff = a;
        h.cap = b - a;
        c->table_h.push_back(h);
        c->table_sid.push_back(kNoScene);
        coff += 1;
    };
    uint64_t active = 0;
    for (uint32_t sid : scenes_by_slot(c)) {
        SceneRec &r = c->scenes[sid];
        if (r.s.poff > pos) free_range(pos, r.s.poff);
        r.s.coff = (uint32_t) coff;
        if (r.has_emitter) {
            r.em.scene = (uint32_t) c->table_h.size();
            c->emit_h.push_back(r.em);
        }
        c->table_h.push_back(r.s);
        c->table_sid.push_back(sid);
        coff += (uint64_t) r.s.ncells + 1;
        pos = segment_end(c, r);
        active = std::max<uint64_t>(active, (uint64_t) r.s.poff + (r.has_emitter ? r.s.cap : r.s.n));
    }
    if (pos < c->pool) free_range(pos, c->pool);
    if (coff >= 0xFFFFFFFFull) return fail(SPH_ERR_CAPACITY, "the scenes' grids exceed 2^32 - 1 cells");
    c->total_cells = (uint32_t) coff;
    c->nslots = (uint32_t) active;

    // only the groups below the active slots are ever read
    const uint32_t ngroups = (uint32_t) (((uint64_t) c->nslots + sph::kGroupSlots - 1) / sph::kGroupSlots);
    c->group_h.assign(ngroups, 0);
    size_t t = 0;
    for (uint32_t g = 0; g < ngroups; ++g) {
        const uint32_t slot = g * sph::kGroupSlots;
        while (t + 1 < c->table_h.size() && c->table_h[t].poff + c->table_h[t].cap <= slot) ++t;
        c->group_h[g] = (uint32_t) t;
    }
    // sort tiles: runs of groups of one segment, cut every kSortTile slots
    c->tiles_h.clear();
    c->sort_bits = 1;
    for (uint32_t g0 = 0; g0 < ngroups;) {
        uint32_t g1 = g0 + 1;
        while (g1 < ngroups && c->group_h[g1] == c->group_h[g0]) ++g1;
        const sph::Scene &seg = c->table_h[c->group_h[g0]];
        const uint32_t a = g0 * sph::kGroupSlots;
        const uint32_t b = std::min<uint64_t>((uint64_t) g1 * sph::kGroupSlots, c->nslots);
        const uint32_t first = (uint32_t) c->tiles_h.size(), count = (b - a + sph::kSortTile - 1) / sph::kSortTile;
        for (uint32_t k = 0; k < count; ++k) {
            sph::SortTile tl{};
            tl.start = a + k * (uint32_t) sph::kSortTile;
            tl.count = std::min<uint32_t>((uint32_t) sph::kSortTile, b - tl.start);
            tl.coff = seg.coff;
            tl.first = first;
            tl.ntiles = count;
            tl.index = k;
            c->tiles_h.push_back(tl);
        }
        c->sort_bits = std::max(c->sort_bits, key_bits(seg.ncells + 1));
        g0 = g1;
    }
    if (c->tiles_h.size() > c->tiles_cap) return fail(SPH_ERR_CAPACITY, "sort tiles exceed their buffer");
    if (!c->tiles_h.empty())
        SPH_HIP(hipMemcpy(c->tiles_d, c->tiles_h.data(), c->tiles_h.size() * sizeof(sph::SortTile),
                          hipMemcpyHostToDevice));
    if (c->total_cells + 1 > c->cstart_cap) {
        dfree(c, c->cstart);
        c->cstart = nullptr;
        const uint32_t want = std::max<uint32_t>(c->total_cells + 1, c->cstart_cap + c->cstart_cap / 2);
        SPH_TRY(dalloc(c, &c->cstart, want));
        c->cstart_cap = want;
    }
    if (c->ref_order()) {
        if (c->total_cells + 1 > c->ref_cells_cap) {
            for (uint32_t **a : {&c->ref.n_early, &c->ref.n_arrive, &c->ref.n_depart, &c->ref.nstay, &c->ref.partials}) {
                dfree(c, *a);
                *a = nullptr;
            }
            const uint32_t want = c->cstart_cap;
            SPH_TRY(dalloc(c, &c->ref.n_early, want));
            SPH_TRY(dalloc(c, &c->ref.n_arrive, want));
            SPH_TRY(dalloc(c, &c->ref.n_depart, want));
            SPH_TRY(dalloc(c, &c->ref.nstay, want));
            SPH_TRY(dalloc(c, &c->ref.partials, sph::scan_blocks(want)));
            c->ref_cells_cap = want;
        }
        // one segment over the active slots, for the sort of the movers by global destination
        c->tiles_flat_h.clear();
        const uint32_t nt = (c->nslots + sph::kSortTile - 1) / sph::kSortTile;
        for (uint32_t k = 0; k < nt; ++k) {
            sph::SortTile tl{};
            tl.start = k * (uint32_t) sph::kSortTile;
            tl.count = std::min<uint32_t>((uint32_t) sph::kSortTile, c->nslots - tl.start);
            tl.coff = 0;
            tl.first = 0;
            tl.ntiles = nt;
            tl.index = k;
            c->tiles_flat_h.push_back(tl);
        }
        if (nt > c->tiles_cap) return fail(SPH_ERR_CAPACITY, "sort tiles exceed their buffer");
        if (nt)
            SPH_HIP(hipMemcpy(c->tiles_flat_d, c->tiles_flat_h.data(), nt * sizeof(sph::SortTile),
                              hipMemcpyHostToDevice));
        c->flat_bits = key_bits(c->total_cells + 1);
    }
    SPH_HIP(hipMemcpy(c->table_d, c->table_h.data(), c->table_h.size() * sizeof(sph::Scene), hipMemcpyHostToDevice));
    if (!c->group_h.empty())
        SPH_HIP(hipMemcpy(c->group_d, c->group_h.data(), c->group_h.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice));
    if (!c->emit_h.empty())
        SPH_HIP(hipMemcpy(c->emit_d, c->emit_h.data(), c->emit_h.size() * sizeof(sph::Emitter),
                          hipMemcpyHostToDevice));
    c->layout_dirty = false;
    c->keys_valid = false;
    drop_graphs(c);
    return SPH_OK;
}

sph_status ensure_layout(sph_ctx *c)
{
    return c->layout_dirty ? rebuild_layout(c) : SPH_OK;
}

// Marks the layout for a rebuild. Callers have refreshed the counts first.
void touch_layout(sph_ctx *c)
{
    c->layout_dirty = true;
    c->keys_valid = false;
    drop_graphs(c);
}

sph_status init_scene(const sph_scene_desc *d, sph::Scene &s, const char *who)
{
    for (int a = 0; a < 3; ++a)
        if (!(d->box[a] > 0.0f) || !std::isfinite(d->box[a]))
            return fail(SPH_ERR_INVALID, std::string(who) + ": box extents must be positive and finite");
    if (d->capacity == 0) return fail(SPH_ERR_INVALID, std::string(who) + ": capacity must be > 0");
    if (d->capacity > kMaxSceneSlots)
        return fail(SPH_ERR_INVALID, std::string(who) + ": capacity above " + std::to_string(kMaxSceneSlots));
    if (!(d->dt > 0.0f) || !std::isfinite(d->dt)) return fail(SPH_ERR_INVALID, std::string(who) + ": dt must be > 0");
    const int rx = sph_grid_res(d->box[0]), ry = sph_grid_res(d->box[1]), rz = sph_grid_res(d->box[2]);
    const uint64_t ncells = (uint64_t) rx * (uint64_t) ry * (uint64_t) rz;
    if (rx <= 0 || ry <= 0 || rz <= 0 || ncells >= (1ull << 31))
        return fail(SPH_ERR_INVALID, std::string(who) + ": grid of " + std::to_string(ncells) + " cells is out of range");
    s = sph::Scene{};
    s.rx = rx;
    s.ry = ry;
    s.rz = rz;
    s.ncells = (uint32_t) ncells;
    s.bx = d->box[0];
    s.by = d->box[1];
    s.bz = d->box[2];
    s.gx = d->gravity[0];
    s.gy = d->gravity[1];
    s.gz = d->gravity[2];
    s.dt = d->dt;
    return SPH_OK;
}

// First free range of the pool, on a group boundary, that holds `cap` slots.
bool find_range(const sph_ctx *c, uint32_t cap, uint32_t &poff, uint32_t &seg)
{
    uint32_t pos = 0;
    auto fits = [&](uint32_t a, uint32_t b) {
        if ((uint64_t) a + cap > b) return false;
        poff = a;
        seg = std::min<uint32_t>(round_up(cap, sph::kGroupSlots), b - a);
        return true;
    };
    for (uint32_t sid : scenes_by_slot(c)) {
        const SceneRec &r = c->scenes[sid];
        if (fits(pos, r.s.poff)) return true;
        pos = segment_end(c, r);
    }
    return fits(pos, c->pool);
}

// --- stages

sph_status run_emit(sph_ctx *c)
{
    if (c->emit_h.empty()) return SPH_OK;
    sph::launch_emit(c->emit_d, (uint32_t) c->emit_h.size(), c->table_d, c->soa(c->cur, c->id_cur), c->key_h,
                     c->ref_order() ? c->ocell : nullptr, c->k.h, c->stream);
    c->counts_stale = true;
    return launch_error("emitter launch");
}

// Stable LSD radix sort of (keys, vals) over the given tiles into key_s / val_s
// (vals == nullptr: the slot index), in passes of sph::radix_plan's digits up to `bits`.
void radix_sort(sph_ctx *c, const uint32_t *keys, const uint32_t *vals, const sph::SortTile *tiles, uint32_t ntiles,
                int bits)
{
    int digit[8];
    const int passes = sph::radix_plan(bits, digit);
    const uint32_t *kin = keys;
    const uint32_t *vin = vals;
    int shift = 0;
    for (int p = 0; p < passes; ++p) {
        const bool to_s = ((passes - 1 - p) % 2) == 0;
        uint32_t *kout = to_s ? c->key_s : c->key_t;
        uint32_t *vout = to_s ? c->val_s : c->val_t;
        sph::launch_radix_pass(kin, vin, kout, vout, tiles, ntiles, shift, digit[p], c->hist, c->partials, c->stream,
                               c->scan_poll_limit);
        shift += digit[p];
        kin = kout;
        vin = vout;
    }
}

// Grid phase in the reference's bucket order (see the reference-order kernels):
// the old buckets (a stable sort by old cell, so that particles added since the
// last grid phase follow cell (0,0,0)'s bucket), per-cell counts, the per-cell
// swap-remove scans, the movers sorted by destination in push order, the new
// cell table, then the gather.
sph_status run_grid_reference(sph_ctx *c)
{
    const uint32_t n = c->nslots, C = c->total_cells;
    const sph::SceneTable tab = c->table();
    {
        StageTimer t(c, SPH_STAGE_SORT);
        if (t.st) return t.st;
        sph::launch_ref_prepare(c->ocell, c->gold, tab, c->stream);
        radix_sort(c, c->gold, nullptr, c->tiles_d, (uint32_t) c->tiles_h.size(), c->sort_bits);
        sph::launch_cell_start(c->key_s, n, C, c->cstart, c->stream);
        sph::launch_ref_counts(c->ref, c->gold, c->key_h, n, C, c->stream);
        sph::launch_ref_bucket(c->ref, C, c->cstart, c->val_s, c->key_h, c->stream);
        sph::launch_ref_movers(c->ref, c->gold, c->key_h, n, C, c->stream);
        radix_sort(c, c->ref.mkey, c->ref.mval, c->tiles_flat_d, (uint32_t) c->tiles_flat_h.size(), c->flat_bits);
    }
    {
        StageTimer t(c, SPH_STAGE_CELLS);
        if (t.st) return t.st;
        sph::launch_ref_place(c->ref, c->key_s, c->val_s, n, C, c->cstart, c->stream);
        sph::launch_ref_final(c->ref, c->key_h, c->cstart, c->key_s, c->val_s, c->ocell, tab, c->stream);
    }
    return SPH_OK;
}

// Grid phase: hash (unless the previous integrate already did), stable LSD radix
// sort of (cell, slot) -- or the reference's buckets -- cell start table, and
// gather into sorted order.
sph_status run_grid(sph_ctx *c)
{
    const uint32_t n = c->nslots;
    const sph::SceneTable tab = c->table();
    sph::ParticleSoA cur = c->soa(c->cur, c->id_cur);
    if (!c->keys_valid) {
        StageTimer t(c, SPH_STAGE_HASH);
        if (t.st) return t.st;
        sph::launch_hash(cur, c->key_h, tab, c->k.h, c->stream);
    }
    if (c->ref_order()) {
        SPH_TRY(run_grid_reference(c));
    } else {
        {
            StageTimer t(c, SPH_STAGE_SORT);
            if (t.st) return t.st;
            radix_sort(c, c->key_h, nullptr, c->tiles_d, (uint32_t) c->tiles_h.size(), c->sort_bits);
        }
        StageTimer t(c, SPH_STAGE_CELLS);
        if (t.st) return t.st;
        sph::launch_cell_start(c->key_s, n, c->total_cells, c->cstart, c->stream);
    }
    {
        StageTimer t(c, SPH_STAGE_REORDER);
        if (t.st) return t.st;
        sph::ParticleSoA dst = c->soa(1 - c->cur, 1 - c->id_cur);
        sph::launch_reorder(cur, dst, c->val_s, n, c->stream);
    }
    SPH_TRY(launch_error("grid phase launch"));
    c->cur = 1 - c->cur;
    c->id_cur = 1 - c->id_cur;
    c->keys_valid = false;  // key_h still holds the pre-sort cells (kept for sph_debug)
    return SPH_OK;
}

sph_status run_density(sph_ctx *c)
{
    StageTimer t(c, SPH_STAGE_DENSITY);
    if (t.st) return t.st;
    const bool diag = (c->flags & SPH_FLAG_DIAGNOSTICS) != 0;
    sph::launch_density(c->soa(c->cur, c->id_cur), c->key_s, c->cstart, c->fields(diag), c->nlists(), c->table(),
                        c->k, c->variant, c->ref_order(), c->stream);
    return launch_error("density launch");
}

sph_status run_fused_tail(sph_ctx *c)
{
    StageTimer t(c, SPH_STAGE_FORCES);
    if (t.st) return t.st;
    const bool diag = (c->flags & SPH_FLAG_DIAGNOSTICS) != 0;
    sph::ParticleSoA in = c->soa(c->cur, c->id_cur);
    sph::ParticleSoA out = c->soa(1 - c->cur, c->id_cur);
    sph::launch_forces_fused(in, c->key_s, c->cstart, c->fields(diag), c->nlists(), out, c->key_h,
                             diag ? c->acc[0] : nullptr, diag ? c->acc[1] : nullptr, diag ? c->acc[2] : nullptr,
                             c->table(), c->k, c->variant, c->ref_order(), c->stream);
    SPH_TRY(launch_error("fused forces launch"));
    c->cur = 1 - c->cur;
    c->keys_valid = true;
    return SPH_OK;
}

sph_status check_fault(sph_ctx *c)
{
    if (c->fault_at_step >= 0 && (long) c->steps_done == c->fault_at_step)
        return fail(SPH_ERR_FAULT, "injected failure at step " + std::to_string(c->steps_done) +
                                       " (SPH_FAULT_AT_STEP)");
    return SPH_OK;
}

sph_status check_runnable(sph_ctx *c, const char *who)
{
    SPH_TRY(ensure_layout(c));
    if (c->nslots == 0) return fail(SPH_ERR_STATE, std::string(who) + ": no particles uploaded");
    return SPH_OK;
}

// One whole step on the stream: emitters, grid, density, fused tail.
sph_status run_step_eager(sph_ctx *c)
{
    SPH_TRY(run_emit(c));
    SPH_TRY(run_grid(c));
    SPH_TRY(run_density(c));
    return run_fused_tail(c);
}

// One step, replayed from a graph captured for the current buffer parity. A step
// starts from valid keys (the previous fused tail wrote them), so the graph holds
// no hash; the first step after a layout change runs eagerly.
sph_status run_step(sph_ctx *c)
{
    const bool graphs = c->use_graphs && !(c->flags & SPH_FLAG_PROFILE) && c->keys_valid;
    if (!graphs) return run_step_eager(c);
    hipGraphExec_t &g = c->graph[c->cur][c->id_cur];
    const int cur0 = c->cur, id0 = c->id_cur;
    if (!g) {
        SPH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        sph_status st = run_step_eager(c);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        // the capture only recorded the launches: put the host state back
        c->cur = cur0;
        c->id_cur = id0;
        c->keys_valid = true;
        if (st) {
            if (graph) (void) hipGraphDestroy(graph);
            return st;
        }
        if (e != hipSuccess) return fail(SPH_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
        (void) hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g = nullptr;
            return fail(SPH_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }
    SPH_HIP(hipGraphLaunch(g, c->stream));
    if (!c->emit_h.empty()) c->counts_stale = true;
    c->id_cur = 1 - c->id_cur;   // grid flips both, the fused tail flips cur back
    c->keys_valid = true;
    return SPH_OK;
}

sph_status run_phase(sph_ctx *c, int phase)
{
    if (phase < SPH_PHASE_GRID || phase > SPH_PHASE_INTEGRATE)
        return fail(SPH_ERR_INVALID, "unknown phase " + std::to_string(phase));
    if (phase != c->next_phase)
        return fail(SPH_ERR_STATE, "phase " + std::to_string(phase) + " called out of order (expected " +
                                       std::to_string(c->next_phase) + ")");
    if (phase == SPH_PHASE_GRID) SPH_TRY(check_runnable(c, "sph_phase"));
    sph_status st = SPH_OK;
    switch (phase) {
    case SPH_PHASE_GRID:
        st = check_fault(c);
        if (!st) st = run_emit(c);
        if (!st) st = run_grid(c);
        break;
    case SPH_PHASE_DENSITY:
        st = run_density(c);
        break;
    case SPH_PHASE_FORCES: {
        StageTimer t(c, SPH_STAGE_FORCES);
        if (t.st) return t.st;
        sph::launch_forces(c->soa(c->cur, c->id_cur), c->key_s, c->cstart, c->fields(false), c->nlists(), c->acc[0],
                           c->acc[1], c->acc[2], c->table(), c->k, c->ref_order(), c->stream);
        st = launch_error("forces launch");
        break;
    }
    case SPH_PHASE_COLLISIONS: {
        StageTimer t(c, SPH_STAGE_COLLISIONS);
        if (t.st) return t.st;
        sph::launch_collide(c->soa(c->cur, c->id_cur), c->acc[0], c->acc[1], c->acc[2], c->table(), c->k,
                            c->stream);
        st = launch_error("collisions launch");
        break;
    }
    case SPH_PHASE_INTEGRATE: {
        StageTimer t(c, SPH_STAGE_INTEGRATE);
        if (t.st) return t.st;
        sph::ParticleSoA in = c->soa(c->cur, c->id_cur);
        sph::ParticleSoA out = c->soa(1 - c->cur, c->id_cur);
        sph::launch_integrate(in, c->acc[0], c->acc[1], c->acc[2], out, c->key_h, c->table(), c->k, c->stream);
        st = launch_error("integrate launch");
        if (!st) {
            c->cur = 1 - c->cur;
            c->keys_valid = true;
            c->steps_done++;
        }
        break;
    }
    }
    if (st) return st;
    c->next_phase = (phase + 1) % 5;
    return SPH_OK;
}

sph_status copy_d2h(sph_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return SPH_OK;
    SPH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(hipStreamSynchronize(c->stream));
    return SPH_OK;
}

sph_status copy_h2d(sph_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return SPH_OK;
    SPH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    SPH_HIP(hipStreamSynchronize(c->stream));
    return SPH_OK;
}

// Writes n particles into slots first.. of the current buffers; ids default to
// id_base + index.
sph_status write_slots(sph_ctx *c, const float *pos3, const float *vel3, const uint32_t *ids, uint32_t first,
                       uint32_t n, uint32_t id_base)
{
    if (n == 0) return SPH_OK;
    std::vector<float> buf((size_t) n);
    std::vector<uint32_t> ib((size_t) n);
    for (int comp = 0; comp < 3; ++comp) {
        for (uint32_t i = 0; i < n; ++i) buf[i] = pos3[3 * (size_t) i + comp];
        SPH_TRY(copy_h2d(c, c->pos[c->cur][comp] + first, buf.data(), n * sizeof(float)));
        for (uint32_t i = 0; i < n; ++i) buf[i] = vel3 ? vel3[3 * (size_t) i + comp] : 0.0f;
        SPH_TRY(copy_h2d(c, c->vel[c->cur][comp] + first, buf.data(), n * sizeof(float)));
    }
    for (uint32_t i = 0; i < n; ++i) ib[i] = ids ? ids[i] : id_base + i;
    SPH_TRY(copy_h2d(c, c->ids[c->id_cur] + first, ib.data(), n * sizeof(uint32_t)));
    if (c->ref_order()) {
        // uploaded and appended particles join the back of cell (0,0,0)'s bucket (addParticle)
        std::fill(ib.begin(), ib.end(), 0u);
        SPH_TRY(copy_h2d(c, c->ocell + first, ib.data(), n * sizeof(uint32_t)));
    }
    return SPH_OK;
}

sph_status scene_write(sph_ctx *c, uint32_t scene, const float *pos3, const float *vel3, const uint32_t *ids,
                       uint32_t n, bool append, const char *who)
{
    SPH_TRY(check_scene(c, scene, who));
    if (n > 0 && !pos3) return fail(SPH_ERR_INVALID, std::string(who) + ": null positions");
    SPH_TRY(sync_counts(c));
    SceneRec &r = c->scenes[scene];
    const uint32_t base = append ? r.s.n : 0u;
    if ((uint64_t) base + n > r.req_cap)
        return fail(SPH_ERR_CAPACITY, std::string(who) + ": " + std::to_string((uint64_t) base + n) +
                                          " particles exceed capacity " + std::to_string(r.req_cap));
    SPH_HIP(hipStreamSynchronize(c->stream));
    SPH_TRY(write_slots(c, pos3, vel3, ids, r.s.poff + base, n, base));
    r.s.n = base + n;
    touch_layout(c);
    return SPH_OK;
}

sph_status scene_download(sph_ctx *c, uint32_t scene, float *pos3, float *vel3, uint32_t *ids, uint32_t n,
                          int by_id, const char *who)
{
    SPH_TRY(check_scene(c, scene, who));
    SPH_TRY(sync_counts(c));
    const sph::Scene &s = c->scenes[scene].s;
    if (n != s.n)
        return fail(SPH_ERR_INVALID, std::string(who) + ": n = " + std::to_string(n) + " but the scene holds " +
                                         std::to_string(s.n));
    if (n == 0) return SPH_OK;
    std::vector<uint32_t> id((size_t) n);
    SPH_TRY(copy_d2h(c, id.data(), c->ids[c->id_cur] + s.poff, n * sizeof(uint32_t)));
    if (by_id)
        for (uint32_t i = 0; i < n; ++i)
            if (id[i] >= n)
                return fail(SPH_ERR_STATE, std::string(who) + ": by_id needs ids 0..n-1, found " +
                                               std::to_string(id[i]));
    if (ids)
        for (uint32_t i = 0; i < n; ++i) ids[by_id ? id[i] : i] = id[i];
    std::vector<float> buf((size_t) n);
    for (int comp = 0; comp < 3; ++comp) {
        if (pos3) {
            SPH_TRY(copy_d2h(c, buf.data(), c->pos[c->cur][comp] + s.poff, n * sizeof(float)));
            for (uint32_t i = 0; i < n; ++i) pos3[3 * (size_t) (by_id ? id[i] : i) + comp] = buf[i];
        }
        if (vel3) {
            SPH_TRY(copy_d2h(c, buf.data(), c->vel[c->cur][comp] + s.poff, n * sizeof(float)));
            for (uint32_t i = 0; i < n; ++i) vel3[3 * (size_t) (by_id ? id[i] : i) + comp] = buf[i];
        }
    }
    return SPH_OK;
}

// Copies n uint32 values from dev + first and subtracts `bias` from each.
sph_status debug_u32(sph_ctx *c, void *out, const uint32_t *dev, uint32_t first, size_t n, uint32_t bias)
{
    SPH_TRY(copy_d2h(c, out, dev + first, n * 4));
    if (bias)
        for (size_t i = 0; i < n; ++i) ((uint32_t *) out)[i] -= bias;
    return SPH_OK;
}

sph_status scene_debug(sph_ctx *c, uint32_t scene, int kind, void *out, size_t bytes)
{
    SPH_TRY(check_scene(c, scene, "sph_debug"));
    if (!out) return fail(SPH_ERR_INVALID, "sph_debug: null output");
    SPH_TRY(sync_counts(c));
    if (c->layout_dirty && (kind == SPH_DEBUG_CELL_START))
        return fail(SPH_ERR_STATE, "sph_debug: the cell table is out of date (no grid phase since the last change)");
    const sph::Scene &s = c->scenes[scene].s;
    const uint32_t n = s.n, p0 = s.poff;
    const size_t need_n4 = (size_t) n * 4;
    auto need = [&](size_t b) -> sph_status {
        if (bytes < b)
            return fail(SPH_ERR_INVALID, "sph_debug: buffer of " + std::to_string(bytes) + " B, need " +
                                             std::to_string(b));
        return SPH_OK;
    };
    switch (kind) {
    case SPH_DEBUG_CELL_PRESORT:
        SPH_TRY(need(need_n4));
        return debug_u32(c, out, c->key_h, p0, n, s.coff);
    case SPH_DEBUG_PERM:
        SPH_TRY(need(need_n4));
        return debug_u32(c, out, c->val_s, p0, n, p0);
    case SPH_DEBUG_CELL:
        SPH_TRY(need(need_n4));
        return debug_u32(c, out, c->key_s, p0, n, s.coff);
    case SPH_DEBUG_CELL_START:
        SPH_TRY(need(((size_t) s.ncells + 1) * 4));
        return debug_u32(c, out, c->cstart, s.coff, (size_t) s.ncells + 1, p0);
    case SPH_DEBUG_NEIGHBOURS:
        if (!(c->flags & SPH_FLAG_DIAGNOSTICS))
            return fail(SPH_ERR_STATE, "sph_debug: neighbour counts need SPH_FLAG_DIAGNOSTICS");
        SPH_TRY(need(need_n4));
        return copy_d2h(c, out, c->ncount + p0, need_n4);
    case SPH_DEBUG_DENSITY:
        SPH_TRY(need(need_n4));
        return copy_d2h(c, out, c->rho + p0, need_n4);
    case SPH_DEBUG_PRESSURE: {
        SPH_TRY(need(need_n4));
        SPH_TRY(copy_d2h(c, out, c->rho + p0, need_n4));
        float *f = (float *) out;
        for (uint32_t i = 0; i < n; ++i) f[i] = sph::kGasStiffness * (f[i] - sph::kRestDensity);
        return SPH_OK;
    }
    case SPH_DEBUG_ACCEL: {
        SPH_TRY(need(need_n4 * 3));
        std::vector<float> buf((size_t) n);
        float *f = (float *) out;
        for (int a = 0; a < 3; ++a) {
            SPH_TRY(copy_d2h(c, buf.data(), c->acc[a] + p0, need_n4));
            for (uint32_t i = 0; i < n; ++i) f[3 * (size_t) i + a] = buf[i];
        }
        return SPH_OK;
    }
    case SPH_DEBUG_IDS:
        SPH_TRY(need(need_n4));
        return copy_d2h(c, out, c->ids[c->id_cur] + p0, need_n4);
    default:
        return fail(SPH_ERR_INVALID, "sph_debug: unknown kind " + std::to_string(kind));
    }
}

sph_status check_device(int *dev_out)
{
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(SPH_ERR_NO_DEVICE, "no HIP device visible");
    int dev = 0;
    SPH_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SPH_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SPH_ERR_NO_DEVICE, std::string("libsph_hip.so is built for gfx950, device is ") + prop.gcnArchName);
    *dev_out = dev;
    return SPH_OK;
}

sph_status alloc_buffers(sph_ctx *c, uint32_t max_segments)
{
    const size_t P = c->pool;
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 3; ++a) {
            SPH_TRY(dalloc(c, &c->pos[b][a], P));
            SPH_TRY(dalloc(c, &c->vel[b][a], P));
        }
    for (int b = 0; b < 2; ++b) SPH_TRY(dalloc(c, &c->ids[b], P));
    SPH_TRY(dalloc(c, &c->key_h, P));
    SPH_TRY(dalloc(c, &c->key_t, P));
    SPH_TRY(dalloc(c, &c->key_s, P));
    SPH_TRY(dalloc(c, &c->val_t, P));
    SPH_TRY(dalloc(c, &c->val_s, P));
    c->tiles_cap = sph::max_sort_tiles(c->pool, max_segments);
    const size_t hist_n = (size_t) sph::kRadixBucketsMax * c->tiles_cap;
    SPH_TRY(dalloc(c, &c->hist, hist_n));
    SPH_TRY(dalloc(c, &c->partials, hist_n + 2 * ((size_t) sph::scan_blocks((uint32_t) hist_n) + 1)));
    SPH_TRY(dalloc(c, &c->tiles_d, c->tiles_cap));
    if (c->ref_order()) {
        SPH_TRY(dalloc(c, &c->ocell, P));
        SPH_TRY(dalloc(c, &c->gold, P));
        SPH_TRY(dalloc(c, &c->ref.rk, P));
        SPH_TRY(dalloc(c, &c->ref.epos, P));
        SPH_TRY(dalloc(c, &c->ref.mkey, P));
        SPH_TRY(dalloc(c, &c->ref.mval, P));
        SPH_TRY(dalloc(c, &c->ref.work, 2 * P));
        SPH_TRY(dalloc(c, &c->tiles_flat_d, c->tiles_cap));
    }
    SPH_TRY(dalloc(c, &c->rho, P));
    SPH_TRY(dalloc(c, &c->rec, (size_t) sph::kRecWords * P));
    for (int a = 0; a < 3; ++a) SPH_TRY(dalloc(c, &c->acc[a], P));
    SPH_TRY(dalloc(c, &c->ncount, P));
    SPH_TRY(dalloc(c, &c->nlist_cnt, P));
    SPH_TRY(dalloc(c, &c->nlist, (size_t) c->kmax * round_up(c->pool, 64)));
    SPH_TRY(dalloc(c, &c->table_d, max_segments));
    SPH_TRY(dalloc(c, &c->group_d, (P + sph::kGroupSlots - 1) / sph::kGroupSlots));
    SPH_TRY(dalloc(c, &c->emit_d, c->scenes.size()));
    if (c->flags & SPH_FLAG_PROFILE)
        for (auto &pair : c->ev)
            for (auto &ev : pair)
                if (hipEventCreate(&ev) != hipSuccess) return fail(SPH_ERR_HIP, "hipEventCreate failed");
    for (auto &ev : c->span)
        if (hipEventCreate(&ev) != hipSuccess) return fail(SPH_ERR_HIP, "hipEventCreate failed");
    return SPH_OK;
}

// Creates a context over `pool` slots with up to max_scenes scene ids and places
// the given scenes one after another.
sph_status create_ctx(uint32_t pool, uint32_t max_scenes, uint32_t flags, uint32_t max_neighbours,
                      const sph_scene_desc *descs, uint32_t nscenes, sph_ctx **out, const char *who)
{
    if (!out) return fail(SPH_ERR_INVALID, std::string(who) + ": null output");
    *out = nullptr;
    if (nscenes > 0 && !descs) return fail(SPH_ERR_INVALID, std::string(who) + ": null scenes");
    if (max_neighbours > 4096 && max_neighbours != SPH_NEIGHBOUR_LISTS_OFF)
        return fail(SPH_ERR_INVALID, std::string(who) + ": max_neighbours above 4096");
    if (max_scenes == 0 || max_scenes < nscenes || max_scenes > (1u << 20))
        return fail(SPH_ERR_INVALID, std::string(who) + ": max_scenes must be in [max(nscenes, 1), 2^20]");
    // scenes one after another, each from a group boundary
    std::vector<sph::Scene> placed(nscenes);
    uint64_t end = 0;
    for (uint32_t i = 0; i < nscenes; ++i) {
        SPH_TRY(init_scene(&descs[i], placed[i], who));
        if (end > kMaxPool)
            return fail(SPH_ERR_CAPACITY, std::string(who) + ": the scenes need more than " +
                                              std::to_string(kMaxPool) + " slots");
        placed[i].poff = (uint32_t) end;
        end += round_up(descs[i].capacity, sph::kGroupSlots);
    }
    if (pool == 0) pool = (uint32_t) std::min<uint64_t>(end, (uint64_t) kMaxPool + 1);
    if (pool == 0 || pool > kMaxPool)
        return fail(SPH_ERR_INVALID, std::string(who) + ": pool of " + std::to_string(pool) + " slots out of range");
    for (uint32_t i = 0; i < nscenes; ++i) {
        if ((uint64_t) placed[i].poff + descs[i].capacity > pool)
            return fail(SPH_ERR_CAPACITY, std::string(who) + ": the scenes do not fit the pool");
        placed[i].cap = std::min<uint32_t>(round_up(descs[i].capacity, sph::kGroupSlots), pool - placed[i].poff);
    }
    int dev = 0;
    SPH_TRY(check_device(&dev));

    sph_ctx *c = new sph_ctx();
    c->flags = flags;
    c->device = dev;
    c->pool = pool;
    c->kmax = max_neighbours == SPH_NEIGHBOUR_LISTS_OFF ? 0u : max_neighbours;
    // lists when asked for, else the row masks (the fastest measured)
    c->nb_mode = c->kmax > 0 ? SPH_NEIGHBOURS_LISTS : SPH_NEIGHBOURS_MASKS;
    c->k = sph::make_consts();
    c->scenes.resize(max_scenes);
    for (uint32_t i = 0; i < nscenes; ++i) {
        c->scenes[i].used = true;
        c->scenes[i].req_cap = descs[i].capacity;
        c->scenes[i].s = placed[i];
        c->scenes[i].s.order = c->ref_order() ? sph::kOrderReference : sph::kOrderCanonical;
    }
    if (const char *f = std::getenv("SPH_FAULT_AT_STEP")) c->fault_at_step = std::atol(f);
    if (const char *f = std::getenv("SPH_SCAN_POLL_LIMIT")) c->scan_poll_limit = (uint32_t) std::strtoul(f, nullptr, 10);

    hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        release(c);
        return fail(SPH_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    sph_status st = alloc_buffers(c, 2 * max_scenes + 1);
    if (!st && c->nb_mode == SPH_NEIGHBOURS_MASKS) st = dalloc(c, &c->masks, (size_t) 9 * sph::kMaskWords * round_up(c->pool, 64));
    if (!st) st = rebuild_layout(c);
    if (st) {
        std::string msg = g_last_error;
        release(c);
        g_last_error = msg;
        return st;
    }
    *out = c;
    return SPH_OK;
}

// Saved state of the autotuner: the current particle buffers, keys and counts.
struct Snapshot {
    sph_ctx *c;
    void *mem = nullptr;
    std::vector<sph::Scene> table;
    int cur = 0, id_cur = 0;
    bool keys_valid = false;
    uint64_t steps_done = 0;

    sph_status save()
    {
        const size_t P = c->pool;
        sph_status st;
        mem = dalloc_raw(c, P * 36, &st);
        if (st) return st;
        char *m = (char *) mem;
        for (int a = 0; a < 3; ++a) {
            SPH_HIP(hipMemcpyAsync(m + a * P * 4, c->pos[c->cur][a], P * 4, hipMemcpyDeviceToDevice, c->stream));
            SPH_HIP(hipMemcpyAsync(m + (3 + a) * P * 4, c->vel[c->cur][a], P * 4, hipMemcpyDeviceToDevice, c->stream));
        }
        SPH_HIP(hipMemcpyAsync(m + 6 * P * 4, c->ids[c->id_cur], P * 4, hipMemcpyDeviceToDevice, c->stream));
        SPH_HIP(hipMemcpyAsync(m + 7 * P * 4, c->key_h, P * 4, hipMemcpyDeviceToDevice, c->stream));
        if (c->ocell) SPH_HIP(hipMemcpyAsync(m + 8 * P * 4, c->ocell, P * 4, hipMemcpyDeviceToDevice, c->stream));
        SPH_HIP(hipStreamSynchronize(c->stream));
        table = c->table_h;
        SPH_HIP(hipMemcpy(table.data(), c->table_d, table.size() * sizeof(sph::Scene), hipMemcpyDeviceToHost));
        cur = c->cur;
        id_cur = c->id_cur;
        keys_valid = c->keys_valid;
        steps_done = c->steps_done;
        return SPH_OK;
    }

    sph_status restore()
    {
        const size_t P = c->pool;
        SPH_HIP(hipStreamSynchronize(c->stream));
        c->cur = cur;
        c->id_cur = id_cur;
        const char *m = (const char *) mem;
        for (int a = 0; a < 3; ++a) {
            SPH_HIP(hipMemcpyAsync(c->pos[cur][a], m + a * P * 4, P * 4, hipMemcpyDeviceToDevice, c->stream));
            SPH_HIP(hipMemcpyAsync(c->vel[cur][a], m + (3 + a) * P * 4, P * 4, hipMemcpyDeviceToDevice, c->stream));
        }
        SPH_HIP(hipMemcpyAsync(c->ids[id_cur], m + 6 * P * 4, P * 4, hipMemcpyDeviceToDevice, c->stream));
        SPH_HIP(hipMemcpyAsync(c->key_h, m + 7 * P * 4, P * 4, hipMemcpyDeviceToDevice, c->stream));
        if (c->ocell) SPH_HIP(hipMemcpyAsync(c->ocell, m + 8 * P * 4, P * 4, hipMemcpyDeviceToDevice, c->stream));
        SPH_HIP(hipMemcpyAsync(c->table_d, table.data(), table.size() * sizeof(sph::Scene), hipMemcpyHostToDevice,
                               c->stream));
        SPH_HIP(hipStreamSynchronize(c->stream));
        c->keys_valid = keys_valid;
        c->steps_done = steps_done;
        c->counts_stale = !c->emit_h.empty();
        return SPH_OK;
    }

    ~Snapshot()
    {
        if (mem) dfree(c, mem);
    }
};

}  // namespace

extern "C" {

void sph_default_params(sph_params *p, float box, uint32_t capacity)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->box[0] = p->box[1] = p->box[2] = box;
    p->gravity[0] = 0.0f;
    p->gravity[1] = sph::kGravityY;
    p->gravity[2] = 0.0f;
    p->dt = sph::kDt;
    p->capacity = capacity;
    p->flags = 0;
    p->max_neighbours = 0;
}

void sph_default_scene(sph_scene_desc *d, float box, uint32_t capacity)
{
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->box[0] = d->box[1] = d->box[2] = box;
    d->gravity[1] = sph::kGravityY;
    d->dt = sph::kDt;
    d->capacity = capacity;
}

uint64_t sph_slot_bytes(uint32_t max_neighbours)
{
    const uint64_t lists = max_neighbours == SPH_NEIGHBOUR_LISTS_OFF ? 0 : max_neighbours;
    // pos/vel ping-pong 48, ids 8, keys and values 20, rho 4, records 40,
    // accelerations 12, counts 8, row masks (the default mode without
    // lists), list entries, and 4 for the histograms, the group map and the
    // scene's share of the cell table
    return 48 + 8 + 20 + 4 + 4 * sph::kRecWords + 12 + 8 + (lists ? 4 * lists : 9 * 4 * sph::kMaskWords) + 4;
}

int sph_grid_res(float box)
{
    volatile float q = box / sph::kH;
    return (int) std::ceil(q);
}

sph_status sph_create(const sph_params *params, sph_ctx **out)
{
    try {
        if (!params || !out) return fail(SPH_ERR_INVALID, "sph_create: null argument");
        sph_scene_desc d;
        std::memcpy(d.box, params->box, sizeof(d.box));
        std::memcpy(d.gravity, params->gravity, sizeof(d.gravity));
        d.dt = params->dt;
        d.capacity = params->capacity;
        if (params->capacity == 0) return fail(SPH_ERR_INVALID, "sph_create: capacity must be > 0");
        return create_ctx(params->capacity, 1, params->flags, params->max_neighbours, &d, 1, out, "sph_create");
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_batch_create(const sph_batch_params *bp, const sph_scene_desc *scenes, uint32_t nscenes,
                            sph_ctx **out)
{
    try {
        if (!bp) return fail(SPH_ERR_INVALID, "sph_batch_create: null parameters");
        const uint32_t max_scenes = bp->max_scenes ? bp->max_scenes : nscenes;
        return create_ctx(bp->pool, max_scenes, bp->flags, bp->max_neighbours, scenes, nscenes, out, "sph_batch_create");
    } catch (...) {
        return sph::host_exception();
    }
}

void sph_destroy(sph_ctx *ctx)
{
    try {
        release(ctx);
    } catch (...) {
        (void) sph::host_exception();
    }
}

sph_status sph_scene_add(sph_ctx *c, const sph_scene_desc *desc, uint32_t *scene)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        if (!desc || !scene) return fail(SPH_ERR_INVALID, "sph_scene_add: null argument");
        SPH_TRY(between_steps(c, "sph_scene_add"));
        sph::Scene s;
        SPH_TRY(init_scene(desc, s, "sph_scene_add"));
        uint32_t id = 0;
        while (id < c->scenes.size() && c->scenes[id].used) ++id;
        if (id == c->scenes.size()) return fail(SPH_ERR_CAPACITY, "sph_scene_add: all scene ids are in use");
        SPH_TRY(sync_counts(c));
        uint32_t poff = 0, seg = 0;
        if (!find_range(c, desc->capacity, poff, seg))
            return fail(SPH_ERR_CAPACITY, "sph_scene_add: no free range of " + std::to_string(desc->capacity) +
                                              " slots in the pool");
        s.poff = poff;
        s.cap = seg;
        s.order = c->ref_order() ? sph::kOrderReference : sph::kOrderCanonical;
        SceneRec &r = c->scenes[id];
        r = SceneRec{};
        r.used = true;
        r.req_cap = desc->capacity;
        r.s = s;
        touch_layout(c);
        *scene = id;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_remove(sph_ctx *c, uint32_t scene)
{
    try {
        SPH_TRY(check_scene(c, scene, "sph_scene_remove"));
        SPH_TRY(between_steps(c, "sph_scene_remove"));
        SPH_TRY(sync_counts(c));
        c->scenes[scene] = SceneRec{};
        touch_layout(c);
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_upload(sph_ctx *c, const float *pos3, const float *vel3, const uint32_t *ids, uint32_t n)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        SPH_TRY(scene_write(c, 0, pos3, vel3, ids, n, false, "sph_upload"));
        c->next_phase = SPH_PHASE_GRID;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_append(sph_ctx *c, const float *pos3, const float *vel3, const uint32_t *ids, uint32_t n)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        if (n == 0) return SPH_OK;
        SPH_TRY(between_steps(c, "sph_append"));
        return scene_write(c, 0, pos3, vel3, ids, n, true, "sph_append");
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_upload(sph_ctx *c, uint32_t scene, const float *pos3, const float *vel3, const uint32_t *ids,
                            uint32_t n)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        SPH_TRY(between_steps(c, "sph_scene_upload"));
        return scene_write(c, scene, pos3, vel3, ids, n, false, "sph_scene_upload");
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_append(sph_ctx *c, uint32_t scene, const float *pos3, const float *vel3, const uint32_t *ids,
                            uint32_t n)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        if (n == 0) return SPH_OK;
        SPH_TRY(between_steps(c, "sph_scene_append"));
        return scene_write(c, scene, pos3, vel3, ids, n, true, "sph_scene_append");
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_set_gravity(sph_ctx *c, uint32_t scene, float gx, float gy, float gz)
{
    try {
        SPH_TRY(check_scene(c, scene, "sph_set_gravity"));
        SPH_TRY(sync_counts(c));
        sph::Scene &s = c->scenes[scene].s;
        s.gx = gx;
        s.gy = gy;
        s.gz = gz;
        // only the gravity of one table entry changes: patch it in place so that the
        // layout, the keys and captured graphs stay valid
        if (!c->layout_dirty) {
            for (size_t t = 0; t < c->table_sid.size(); ++t)
                if (c->table_sid[t] == scene) {
                    c->table_h[t].gx = gx;
                    c->table_h[t].gy = gy;
                    c->table_h[t].gz = gz;
                    SPH_HIP(hipStreamSynchronize(c->stream));
                    SPH_HIP(hipMemcpy(&c->table_d[t].gx, &c->table_h[t].gx, 3 * sizeof(float), hipMemcpyHostToDevice));
                }
        }
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_set_gravity(sph_ctx *c, float gx, float gy, float gz)
{
    try {
        return sph_scene_set_gravity(c, 0, gx, gy, gz);
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_set_emitter(sph_ctx *c, uint32_t scene, const sph_emitter *e)
{
    try {
        SPH_TRY(check_scene(c, scene, "sph_scene_set_emitter"));
        SPH_TRY(between_steps(c, "sph_scene_set_emitter"));
        SPH_TRY(sync_counts(c));
        SceneRec &r = c->scenes[scene];
        if (!e) {
            r.has_emitter = false;
            touch_layout(c);
            return SPH_OK;
        }
        if (e->kind != SPH_EMIT_FOUNTAIN && e->kind != SPH_EMIT_DISC)
            return fail(SPH_ERR_INVALID, "sph_scene_set_emitter: unknown kind " + std::to_string(e->kind));
        if (e->kind == SPH_EMIT_DISC) {
            if (e->rate == 0) return fail(SPH_ERR_INVALID, "sph_scene_set_emitter: rate must be > 0");
            if (!(e->radius >= 0.0f) || !std::isfinite(e->radius) || !std::isfinite(e->speed) ||
                !(e->spread >= 0.0f) || !std::isfinite(e->spread) || !std::isfinite(e->centre[0]) ||
                !std::isfinite(e->centre[1]))
                return fail(SPH_ERR_INVALID, "sph_scene_set_emitter: disc parameters must be finite, radius >= 0");
        }
        sph::Emitter &m = r.em;
        m = sph::Emitter{};
        m.kind = e->kind;
        m.rate = e->rate;
        // never beyond the capacity the caller reserved
        m.limit = e->kind == SPH_EMIT_DISC ? std::min(e->limit, r.req_cap) : e->limit;
        if (e->kind == SPH_EMIT_FOUNTAIN && (e->limit > r.req_cap))
            return fail(SPH_ERR_CAPACITY, "sph_scene_set_emitter: fountain limit above the scene's capacity");
        m.cx = e->centre[0];
        m.cz = e->centre[1];
        m.radius = e->radius;
        m.speed = e->speed;
        m.spread = e->spread;
        m.seed_lo = (uint32_t) e->seed;
        m.seed_hi = (uint32_t) (e->seed >> 32);
        r.has_emitter = true;
        touch_layout(c);
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_phase(sph_ctx *c, int phase, float *ms_out)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        if (ms_out) SPH_HIP(hipEventRecord(c->span[0], c->stream));
        SPH_TRY(run_phase(c, phase));
        if (ms_out) {
            SPH_HIP(hipEventRecord(c->span[1], c->stream));
            SPH_HIP(hipEventSynchronize(c->span[1]));
            SPH_HIP(hipEventElapsedTime(ms_out, c->span[0], c->span[1]));
        }
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_step(sph_ctx *c, uint32_t nsteps, float *ms_out)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        if (c->next_phase != SPH_PHASE_GRID) return fail(SPH_ERR_STATE, "sph_step: a phase sequence is in progress");
        SPH_TRY(check_runnable(c, "sph_step"));
        if (ms_out) SPH_HIP(hipEventRecord(c->span[0], c->stream));
        for (uint32_t s = 0; s < nsteps; ++s) {
            SPH_TRY(check_fault(c));
            SPH_TRY(run_step(c));
            c->steps_done++;
        }
        if (ms_out) {
            SPH_HIP(hipEventRecord(c->span[1], c->stream));
            SPH_HIP(hipEventSynchronize(c->span[1]));
            SPH_HIP(hipEventElapsedTime(ms_out, c->span[0], c->span[1]));
        }
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_download(sph_ctx *c, float *pos3, float *vel3, uint32_t *ids, uint32_t n, int by_id)
{
    try {
        return scene_download(c, 0, pos3, vel3, ids, n, by_id, "sph_download");
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_download(sph_ctx *c, uint32_t scene, float *pos3, float *vel3, uint32_t *ids, uint32_t n,
                              int by_id)
{
    try {
        return scene_download(c, scene, pos3, vel3, ids, n, by_id, "sph_scene_download");
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_debug(sph_ctx *c, int kind, void *out, size_t bytes)
{
    try {
        return scene_debug(c, 0, kind, out, bytes);
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_debug(sph_ctx *c, uint32_t scene, int kind, void *out, size_t bytes)
{
    try {
        return scene_debug(c, scene, kind, out, bytes);
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_count(const sph_ctx *cc, uint32_t *n)
{
    try {
        return sph_scene_count(const_cast<sph_ctx *>(cc), 0, n);
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_count(sph_ctx *c, uint32_t scene, uint32_t *n)
{
    try {
        SPH_TRY(check_scene(c, scene, "sph_count"));
        if (!n) return fail(SPH_ERR_INVALID, "sph_count: null output");
        SPH_TRY(sync_counts(c));
        *n = c->scenes[scene].s.n;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_grid_dims(const sph_ctx *c, int *xyz, uint32_t *ncells)
{
    try {
        return sph_scene_grid(const_cast<sph_ctx *>(c), 0, xyz, ncells, nullptr, nullptr);
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_grid(sph_ctx *c, uint32_t scene, int *xyz, uint32_t *ncells, uint32_t *first_slot,
                          uint32_t *slots)
{
    try {
        SPH_TRY(check_scene(c, scene, "sph_grid_dims"));
        const sph::Scene &s = c->scenes[scene].s;
        if (xyz) {
            xyz[0] = s.rx;
            xyz[1] = s.ry;
            xyz[2] = s.rz;
        }
        if (ncells) *ncells = s.ncells;
        if (first_slot) *first_slot = s.poff;
        if (slots) *slots = s.cap;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_batch_info(sph_ctx *c, uint32_t *scenes, uint32_t *pool, uint32_t *free_slots, uint32_t *active_slots)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        SPH_TRY(ensure_layout(c));
        uint32_t used = 0;
        uint64_t taken = 0;
        for (const SceneRec &r : c->scenes)
            if (r.used) {
                ++used;
                taken += r.s.cap;
            }
        if (scenes) *scenes = used;
        if (pool) *pool = c->pool;
        if (free_slots) *free_slots = (uint32_t) (c->pool - taken);
        if (active_slots) *active_slots = c->nslots;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_set_graphs(sph_ctx *c, int enable)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        SPH_HIP(hipStreamSynchronize(c->stream));
        drop_graphs(c);
        c->use_graphs = enable != 0;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_variant_shape(int variant, int *threads, int *unroll, int *pairs, int *stall, int *queue)
{
    try {
        if (variant < 0 || variant >= sph::kNumVariants)
            return fail(SPH_ERR_INVALID, "sph_variant_shape: no variant " + std::to_string(variant));
        if (threads) *threads = sph::kNeighbourVariants[variant].threads;
        if (unroll) *unroll = sph::kNeighbourVariants[variant].unroll;
        if (pairs) *pairs = sph::kNeighbourVariants[variant].pairs;
        if (stall) *stall = sph::kNeighbourVariants[variant].stall;
        if (queue) *queue = sph::kNeighbourVariants[variant].queue;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

// Allocates the row masks the first time the mode needs them.
sph_status ensure_masks(sph_ctx *c)
{
    if (c->masks) return SPH_OK;
    SPH_HIP(hipStreamSynchronize(c->stream));
    return dalloc(c, &c->masks, (size_t) 9 * sph::kMaskWords * round_up(c->pool, 64));
}

sph_status sph_set_variant(sph_ctx *c, int variant, int mode)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        if (variant < 0 || variant >= sph::kNumVariants)
            return fail(SPH_ERR_INVALID, "sph_set_variant: no variant " + std::to_string(variant));
        if (mode < SPH_NEIGHBOURS_WALK || mode > SPH_NEIGHBOURS_MASKS)
            return fail(SPH_ERR_INVALID, "sph_set_variant: no neighbour mode " + std::to_string(mode));
        if (mode == SPH_NEIGHBOURS_LISTS && c->kmax == 0)
            return fail(SPH_ERR_STATE, "sph_set_variant: the context was created without neighbour lists");
        SPH_TRY(between_steps(c, "sph_set_variant"));
        if (mode == SPH_NEIGHBOURS_MASKS) SPH_TRY(ensure_masks(c));
        SPH_HIP(hipStreamSynchronize(c->stream));
        c->variant = variant;
        c->nb_mode = mode;
        drop_graphs(c);
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_get_variant(const sph_ctx *c, int *variant, int *mode)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        if (variant) *variant = c->variant;
        if (mode) *mode = c->nb_mode;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_autotune(sph_ctx *c, uint32_t steps, float *ms_out)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        if (steps == 0) return fail(SPH_ERR_INVALID, "sph_autotune: steps must be > 0");
        SPH_TRY(between_steps(c, "sph_autotune"));
        SPH_TRY(check_runnable(c, "sph_autotune"));
        SPH_TRY(ensure_masks(c));
        constexpr int kModes = 3;
        if (ms_out)
            for (int i = 0; i < kModes * sph::kNumVariants; ++i) ms_out[i] = -1.0f;
        const uint32_t saved_flags = c->flags;
        const long saved_fault = c->fault_at_step;
        c->flags &= ~(uint32_t) SPH_FLAG_PROFILE;
        c->fault_at_step = -1;
        Snapshot snap{c};
        sph_status st = snap.save();
        int best_v = c->variant, best_m = c->nb_mode;
        float best = -1.0f;
        for (int mode = 0; mode < kModes && !st; ++mode) {
            if (mode == SPH_NEIGHBOURS_LISTS && c->kmax == 0) continue;
            for (int v = 0; v < sph::kNumVariants && !st; ++v) {
                st = snap.restore();
                if (st) break;
                c->variant = v;
                c->nb_mode = mode;
                drop_graphs(c);
                st = run_step(c);   // warm-up (and graph capture)
                if (!st) st = run_step(c);
                float ms = 0.0f;
                if (!st && hipEventRecord(c->span[0], c->stream) != hipSuccess) st = fail(SPH_ERR_HIP, "hipEventRecord");
                for (uint32_t s = 0; s < steps && !st; ++s) st = run_step(c);
                if (!st && hipEventRecord(c->span[1], c->stream) != hipSuccess) st = fail(SPH_ERR_HIP, "hipEventRecord");
                if (!st && hipEventSynchronize(c->span[1]) != hipSuccess) st = fail(SPH_ERR_HIP, "hipEventSynchronize");
                if (!st && hipEventElapsedTime(&ms, c->span[0], c->span[1]) != hipSuccess)
                    st = fail(SPH_ERR_HIP, "hipEventElapsedTime");
                if (st) break;
                ms /= (float) steps;
                if (ms_out) ms_out[kModes * v + mode] = ms;
                if (best < 0.0f || ms < best) {
                    best = ms;
                    best_v = v;
                    best_m = mode;
                }
            }
        }
        sph_status st2 = snap.mem ? snap.restore() : SPH_OK;
        c->flags = saved_flags;
        c->fault_at_step = saved_fault;
        c->variant = best_v;
        c->nb_mode = best_m;
        drop_graphs(c);
        return st ? st : st2;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_scene_readback(sph_ctx *c, uint32_t scene, int with_velocity, const float **pos_xyz,
                              const float **vel_xyz, uint32_t *count)
{
    try {
        SPH_TRY(check_scene(c, scene, "sph_scene_readback"));
        if (!pos_xyz || !count) return fail(SPH_ERR_INVALID, "sph_scene_readback: null output");
        SPH_TRY(sync_counts(c));
        const sph::Scene &s = c->scenes[scene].s;
        const uint32_t n = s.n;
        if (n > c->rb_cap || (with_velocity && !c->rb_vel)) {
            SPH_HIP(hipStreamSynchronize(c->stream));
            const size_t cap = std::max<size_t>(n, c->rb_cap);
            if (c->rb_pos) (void) hipHostFree(c->rb_pos);
            if (c->rb_vel) (void) hipHostFree(c->rb_vel);
            c->rb_pos = c->rb_vel = nullptr;
            c->rb_cap = 0;
            SPH_HIP(hipHostMalloc((void **) &c->rb_pos, std::max<size_t>(cap, 1) * 12, hipHostMallocDefault));
            if (with_velocity) SPH_HIP(hipHostMalloc((void **) &c->rb_vel, std::max<size_t>(cap, 1) * 12, hipHostMallocDefault));
            c->rb_cap = cap;
        }
        if (!c->rb_bad) SPH_TRY(dalloc(c, &c->rb_bad, 1));
        SPH_HIP(hipMemsetAsync(c->rb_bad, 0, 4, c->stream));
        sph::launch_scatter_by_id(c->soa(c->cur, c->id_cur), s.poff, n, c->rb_pos, with_velocity ? c->rb_vel : nullptr,
                                  c->rb_bad, c->stream);
        SPH_TRY(launch_error("readback launch"));
        uint32_t bad = 0;
        SPH_TRY(copy_d2h(c, &bad, c->rb_bad, 4));
        if (bad) return fail(SPH_ERR_STATE, "sph_scene_readback: particle ids are not 0..count-1");
        *pos_xyz = c->rb_pos;
        if (vel_xyz) *vel_xyz = with_velocity ? c->rb_vel : nullptr;
        *count = n;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_set_event_stride(sph_ctx *c, uint32_t stride)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        c->event_stride = stride;
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_profile_events(sph_ctx *c, sph_profile_event *out, uint32_t cap, uint32_t *n, int reset)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        if (!(c->flags & SPH_FLAG_PROFILE)) return fail(SPH_ERR_STATE, "sph_profile_events: needs SPH_FLAG_PROFILE");
        SPH_TRY(drain_profile(c));
        if (out)
            for (size_t i = 0; i < c->events.size() && i < cap; ++i) out[i] = c->events[i];
        if (n) *n = (uint32_t) c->events.size();
        if (reset) c->events.clear();
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_write_logs(const sph_profile_event *ev, uint32_t n, const char *stem, const char *label)
{
    try {
        if (!stem || !label || (n && !ev)) return fail(SPH_ERR_INVALID, "sph_write_logs: null argument");
        const std::string base(stem);
        FILE *sum = std::fopen((base + ".csv").c_str(), "a");
        if (!sum) return fail(SPH_ERR_INVALID, "sph_write_logs: cannot open " + base + ".csv");
        FILE *det = std::fopen((base + "_detail.csv").c_str(), "a");
        if (!det) {
            std::fclose(sum);
            return fail(SPH_ERR_INVALID, "sph_write_logs: cannot open " + base + "_detail.csv");
        }
        // summary: the label, then every event's together() (src/mainwindow.cpp:329-345)
        std::fprintf(sum, "%s;", label);
        for (uint32_t i = 0; i < n; ++i)
            std::fprintf(sum, "%g;", ev[i].update_grid + ev[i].update_density_pressure + ev[i].update_forces +
                                          ev[i].update_collisions + ev[i].integrate);
        std::fprintf(sum, "\n");
        // detail: the label, one line per phase, two empty lines (src/mainwindow.cpp:347-362)
        std::fprintf(det, "%s\n", label);
        const char *names[5] = {"Grid", "Density + pressure", "Forces", "Collisions", "Integrate"};
        for (int ph = 0; ph < 5; ++ph) {
            std::fprintf(det, "%s;", names[ph]);
            for (uint32_t i = 0; i < n; ++i) {
                const double v = ph == 0 ? ev[i].update_grid : ph == 1 ? ev[i].update_density_pressure
                               : ph == 2 ? ev[i].update_forces : ph == 3 ? ev[i].update_collisions : ev[i].integrate;
                std::fprintf(det, "%g;", v);
            }
            std::fprintf(det, "\n");
        }
        std::fprintf(det, "\n\n");
        const bool ok = std::fclose(sum) == 0 && std::fclose(det) == 0;
        return ok ? SPH_OK : fail(SPH_ERR_INVALID, "sph_write_logs: write failed");
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_export_logs(sph_ctx *c, const char *stem, const char *label)
{
    try {
        uint32_t n = 0;
        SPH_TRY(sph_profile_events(c, nullptr, 0, &n, 0));
        return sph_write_logs(c->events.data(), (uint32_t) c->events.size(), stem, label);
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_device_name(char *buf, size_t len)
{
    try {
        if (!buf || len == 0) return fail(SPH_ERR_INVALID, "sph_device_name: empty buffer");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SPH_ERR_NO_DEVICE, "no HIP device visible");
        int dev = 0;
        SPH_HIP(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        SPH_HIP(hipGetDeviceProperties(&prop, dev));
        std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_sync(sph_ctx *c)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        SPH_HIP(hipStreamSynchronize(c->stream));
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

sph_status sph_stage_times(sph_ctx *c, float *ms, uint32_t *launches, int reset)
{
    try {
        if (!check_ctx(c)) return SPH_ERR_INVALID;
        if (!(c->flags & SPH_FLAG_PROFILE)) return fail(SPH_ERR_STATE, "sph_stage_times: needs SPH_FLAG_PROFILE");
        SPH_TRY(drain_profile(c));
        for (int i = 0; i < SPH_STAGE_COUNT; ++i) {
            if (ms) ms[i] = c->stage_ms[i];
            if (launches) launches[i] = c->stage_launches[i];
            if (reset) {
                c->stage_ms[i] = 0.0f;
                c->stage_launches[i] = 0;
            }
        }
        return SPH_OK;
    } catch (...) {
        return sph::host_exception();
    }
}

const char *sph_last_error(void)
{
    return g_last_error.c_str();
}

}  // extern "C"
