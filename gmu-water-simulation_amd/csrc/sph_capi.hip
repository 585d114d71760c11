// sph_capi.hip -- the C ABI of libsph_hip.so (declared in include/sph_hip.h).
//
// A context owns one scene's device state: SoA particle buffers (ping-pong), the
// sort buffers, the cell table and one HIP stream. Every entry point reports
// failures as a status plus a thread-local message and never throws.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

constexpr int kProfileSlots = 256;  // event pairs buffered before a forced drain

}  // namespace

struct sph_ctx {
    sph_params prm{};
    sph::Scene scene{};
    sph::Consts k{};
    uint32_t ncells = 0;
    uint32_t cap = 0;
    int device = 0;
    hipStream_t stream = nullptr;

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

    float *rho = nullptr;
    double *prr = nullptr;
    double *rinv = nullptr;
    float4 *rec = nullptr;
    float *acc[3] = {};
    uint32_t *ncount = nullptr;
    uint32_t *nlist = nullptr;   // neighbour lists, max_neighbours x capacity
    uint32_t *nlist_cnt = nullptr;
    uint32_t kmax = 0;

    bool keys_valid = false;   // key_h holds the cells of the current positions
    bool grid_valid = false;   // sorted order / cell table match the current positions
    int next_phase = SPH_PHASE_GRID;
    uint64_t steps_done = 0;
    long fault_at_step = -1;

    // profiling
    hipEvent_t ev[kProfileSlots][2] = {};
    int ev_stage[kProfileSlots] = {};
    int ev_used = 0;
    float stage_ms[SPH_STAGE_COUNT] = {};
    uint32_t stage_launches[SPH_STAGE_COUNT] = {};
    hipEvent_t span[2] = {};

    std::vector<void *> allocations;

    sph::DensityFields fields(bool diag) const
    {
        sph::DensityFields f;
        f.rho = rho;
        f.prr = prr;
        f.rinv = rinv;
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
        nl.stride = cap;
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

}  // namespace sph

namespace {

template <typename T>
sph_status dalloc(sph_ctx *c, T **p, size_t count)
{
    void *q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
        return fail(SPH_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + " B) failed: " + hipGetErrorString(e));
    c->allocations.push_back(q);
    *p = (T *) q;
    return SPH_OK;
}

void release(sph_ctx *c)
{
    if (!c) return;
    if (c->stream) (void) hipStreamSynchronize(c->stream);
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

// --- profiling: an event pair around every stage when SPH_FLAG_PROFILE is set.

sph_status drain_profile(sph_ctx *c)
{
    if (c->ev_used == 0) return SPH_OK;
    SPH_HIP(hipEventSynchronize(c->ev[c->ev_used - 1][1]));
    for (int i = 0; i < c->ev_used; ++i) {
        float ms = 0.0f;
        SPH_HIP(hipEventElapsedTime(&ms, c->ev[i][0], c->ev[i][1]));
        c->stage_ms[c->ev_stage[i]] += ms;
        c->stage_launches[c->ev_stage[i]] += 1;
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
        if (!(c->prm.flags & SPH_FLAG_PROFILE)) return;
        if (c->ev_used == kProfileSlots) st = drain_profile(c);
        if (st != SPH_OK) return;
        slot = c->ev_used++;
        c->ev_stage[slot] = stage;
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

// Grid phase: hash (unless the previous integrate already did), stable LSD radix
// sort of (cell, slot), cell start table, and gather into sorted order.
sph_status run_grid(sph_ctx *c)
{
    const uint32_t n = c->scene.n;
    sph::ParticleSoA cur = c->soa(c->cur, c->id_cur);
    if (!c->keys_valid) {
        StageTimer t(c, SPH_STAGE_HASH);
        if (t.st) return t.st;
        sph::launch_hash(cur, c->key_h, c->scene, c->k.h, c->stream);
    }
    {
        StageTimer t(c, SPH_STAGE_SORT);
        if (t.st) return t.st;
        const int bits = key_bits(c->ncells);
        const int passes = (bits + sph::kRadixBits - 1) / sph::kRadixBits;
        const uint32_t *kin = c->key_h;
        const uint32_t *vin = nullptr;
        for (int p = 0; p < passes; ++p) {
            const bool to_s = ((passes - 1 - p) % 2) == 0;
            uint32_t *kout = to_s ? c->key_s : c->key_t;
            uint32_t *vout = to_s ? c->val_s : c->val_t;
            sph::launch_radix_pass(kin, vin, kout, vout, n, p * sph::kRadixBits, c->hist, c->partials, c->stream);
            kin = kout;
            vin = vout;
        }
    }
    {
        StageTimer t(c, SPH_STAGE_CELLS);
        if (t.st) return t.st;
        sph::launch_cell_start(c->key_s, n, c->ncells, c->cstart, c->stream);
    }
    {
        StageTimer t(c, SPH_STAGE_REORDER);
        if (t.st) return t.st;
        sph::ParticleSoA dst = c->soa(1 - c->cur, 1 - c->id_cur);
        sph::launch_reorder(cur, dst, c->val_s, n, c->stream);
    }
    sph_status st = launch_error("grid phase launch");
    if (st) return st;
    c->cur = 1 - c->cur;
    c->id_cur = 1 - c->id_cur;
    c->keys_valid = false;  // key_h still holds the pre-sort cells (kept for sph_debug)
    c->grid_valid = true;
    return SPH_OK;
}

sph_status run_density(sph_ctx *c)
{
    StageTimer t(c, SPH_STAGE_DENSITY);
    if (t.st) return t.st;
    const bool diag = (c->prm.flags & SPH_FLAG_DIAGNOSTICS) != 0;
    sph::launch_density(c->soa(c->cur, c->id_cur), c->key_s, c->cstart, c->fields(diag), c->nlists(), c->scene,
                        c->k, c->stream);
    return launch_error("density launch");
}

sph_status run_fused_tail(sph_ctx *c)
{
    StageTimer t(c, SPH_STAGE_FORCES);
    if (t.st) return t.st;
    const bool diag = (c->prm.flags & SPH_FLAG_DIAGNOSTICS) != 0;
    sph::ParticleSoA in = c->soa(c->cur, c->id_cur);
    sph::ParticleSoA out = c->soa(1 - c->cur, c->id_cur);
    sph::launch_forces_fused(in, c->key_s, c->cstart, c->fields(diag), c->nlists(), out, c->key_h,
                             diag ? c->acc[0] : nullptr,
                             diag ? c->acc[1] : nullptr, diag ? c->acc[2] : nullptr, c->scene, c->k, c->stream);
    sph_status st = launch_error("fused forces launch");
    if (st) return st;
    c->cur = 1 - c->cur;
    c->keys_valid = true;
    c->grid_valid = false;
    return SPH_OK;
}

sph_status check_fault(sph_ctx *c)
{
    if (c->fault_at_step >= 0 && (long) c->steps_done == c->fault_at_step)
        return fail(SPH_ERR_FAULT, "injected failure at step " + std::to_string(c->steps_done) +
                                       " (SPH_FAULT_AT_STEP)");
    return SPH_OK;
}

sph_status run_phase(sph_ctx *c, int phase)
{
    if (c->scene.n == 0) return fail(SPH_ERR_STATE, "no particles uploaded");
    if (phase != c->next_phase)
        return fail(SPH_ERR_STATE, "phase " + std::to_string(phase) + " called out of order (expected " +
                                       std::to_string(c->next_phase) + ")");
    sph_status st = SPH_OK;
    switch (phase) {
    case SPH_PHASE_GRID:
        st = check_fault(c);
        if (!st) st = run_grid(c);
        break;
    case SPH_PHASE_DENSITY:
        st = run_density(c);
        break;
    case SPH_PHASE_FORCES: {
        StageTimer t(c, SPH_STAGE_FORCES);
        if (t.st) return t.st;
        sph::launch_forces(c->soa(c->cur, c->id_cur), c->key_s, c->cstart, c->fields(false), c->nlists(), c->acc[0],
                           c->acc[1], c->acc[2], c->scene, c->k, c->stream);
        st = launch_error("forces launch");
        break;
    }
    case SPH_PHASE_COLLISIONS: {
        StageTimer t(c, SPH_STAGE_COLLISIONS);
        if (t.st) return t.st;
        sph::launch_collide(c->soa(c->cur, c->id_cur), c->acc[0], c->acc[1], c->acc[2], c->scene, c->k, c->stream);
        st = launch_error("collisions launch");
        break;
    }
    case SPH_PHASE_INTEGRATE: {
        StageTimer t(c, SPH_STAGE_INTEGRATE);
        if (t.st) return t.st;
        sph::ParticleSoA in = c->soa(c->cur, c->id_cur);
        sph::ParticleSoA out = c->soa(1 - c->cur, c->id_cur);
        sph::launch_integrate(in, c->acc[0], c->acc[1], c->acc[2], out, c->key_h, c->scene, c->k, c->stream);
        st = launch_error("integrate launch");
        if (!st) {
            c->cur = 1 - c->cur;
            c->keys_valid = true;
            c->grid_valid = false;
            c->steps_done++;
        }
        break;
    }
    default:
        return fail(SPH_ERR_INVALID, "unknown phase " + std::to_string(phase));
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

sph_status write_tail(sph_ctx *c, const float *pos3, const float *vel3, const uint32_t *ids, uint32_t first,
                      uint32_t n)
{
    if (n == 0) return SPH_OK;
    std::vector<float> buf((size_t) n);
    std::vector<uint32_t> ib((size_t) n);
    for (int comp = 0; comp < 3; ++comp) {
        for (uint32_t i = 0; i < n; ++i) buf[i] = pos3[3 * (size_t) i + comp];
        SPH_HIP(hipMemcpyAsync(c->pos[c->cur][comp] + first, buf.data(), n * sizeof(float), hipMemcpyHostToDevice,
                               c->stream));
        SPH_HIP(hipStreamSynchronize(c->stream));
        for (uint32_t i = 0; i < n; ++i) buf[i] = vel3 ? vel3[3 * (size_t) i + comp] : 0.0f;
        SPH_HIP(hipMemcpyAsync(c->vel[c->cur][comp] + first, buf.data(), n * sizeof(float), hipMemcpyHostToDevice,
                               c->stream));
        SPH_HIP(hipStreamSynchronize(c->stream));
    }
    for (uint32_t i = 0; i < n; ++i) ib[i] = ids ? ids[i] : first + i;
    SPH_HIP(hipMemcpyAsync(c->ids[c->id_cur] + first, ib.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice,
                           c->stream));
    SPH_HIP(hipStreamSynchronize(c->stream));
    return SPH_OK;
}

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

int sph_grid_res(float box)
{
    volatile float q = box / sph::kH;
    return (int) std::ceil(q);
}

sph_status sph_create(const sph_params *params, sph_ctx **out)
{
    if (!params || !out) return fail(SPH_ERR_INVALID, "sph_create: null argument");
    *out = nullptr;
    for (int a = 0; a < 3; ++a)
        if (!(params->box[a] > 0.0f) || !std::isfinite(params->box[a]))
            return fail(SPH_ERR_INVALID, "sph_create: box extents must be positive and finite");
    if (params->capacity == 0) return fail(SPH_ERR_INVALID, "sph_create: capacity must be > 0");
    if (!(params->dt > 0.0f)) return fail(SPH_ERR_INVALID, "sph_create: dt must be > 0");
    if (params->max_neighbours > 4096 && params->max_neighbours != SPH_NEIGHBOUR_LISTS_OFF)
        return fail(SPH_ERR_INVALID, "sph_create: max_neighbours above 4096");
    const int rx = sph_grid_res(params->box[0]), ry = sph_grid_res(params->box[1]), rz = sph_grid_res(params->box[2]);
    const uint64_t ncells = (uint64_t) rx * (uint64_t) ry * (uint64_t) rz;
    if (ncells == 0 || ncells >= (1ull << 31))
        return fail(SPH_ERR_INVALID, "sph_create: grid of " + std::to_string(ncells) + " cells is out of range");

    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(SPH_ERR_NO_DEVICE, "no HIP device visible");
    int dev = 0;
    SPH_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SPH_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SPH_ERR_NO_DEVICE, std::string("libsph_hip.so is built for gfx950, device is ") + prop.gcnArchName);

    sph_ctx *c = new sph_ctx();
    c->prm = *params;
    c->device = dev;
    c->cap = params->capacity;
    c->kmax = params->max_neighbours == SPH_NEIGHBOUR_LISTS_OFF ? 0u : params->max_neighbours;
    c->ncells = (uint32_t) ncells;
    c->k = sph::make_consts();
    c->scene.rx = rx;
    c->scene.ry = ry;
    c->scene.rz = rz;
    c->scene.coff = 0;
    c->scene.poff = 0;
    c->scene.n = 0;
    c->scene.bx = params->box[0];
    c->scene.by = params->box[1];
    c->scene.bz = params->box[2];
    c->scene.gx = params->gravity[0];
    c->scene.gy = params->gravity[1];
    c->scene.gz = params->gravity[2];
    c->scene.dt = params->dt;
    if (const char *f = std::getenv("SPH_FAULT_AT_STEP")) c->fault_at_step = std::atol(f);

    sph_status st = SPH_OK;
    hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        release(c);
        return fail(SPH_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    const size_t P = c->cap;
    for (int b = 0; b < 2 && !st; ++b)
        for (int a = 0; a < 3 && !st; ++a) {
            st = dalloc(c, &c->pos[b][a], P);
            if (!st) st = dalloc(c, &c->vel[b][a], P);
        }
    for (int b = 0; b < 2 && !st; ++b) st = dalloc(c, &c->ids[b], P);
    if (!st) st = dalloc(c, &c->key_h, P);
    if (!st) st = dalloc(c, &c->key_t, P);
    if (!st) st = dalloc(c, &c->key_s, P);
    if (!st) st = dalloc(c, &c->val_t, P);
    if (!st) st = dalloc(c, &c->val_s, P);
    const size_t hist_n = (size_t) sph::kRadixBuckets * sph::sort_blocks(c->cap);
    if (!st) st = dalloc(c, &c->hist, hist_n);
    if (!st) st = dalloc(c, &c->partials, sph::scan_blocks((uint32_t) hist_n));
    if (!st) st = dalloc(c, &c->cstart, (size_t) c->ncells + 1);
    if (!st) st = dalloc(c, &c->rho, P);
    if (!st) st = dalloc(c, &c->prr, P);
    if (!st) st = dalloc(c, &c->rinv, P);
    if (!st) st = dalloc(c, &c->rec, 4 * P);
    for (int a = 0; a < 3 && !st; ++a) st = dalloc(c, &c->acc[a], P);
    if (!st) st = dalloc(c, &c->ncount, P);
    if (!st) st = dalloc(c, &c->nlist_cnt, P);
    if (!st) st = dalloc(c, &c->nlist, (size_t) c->kmax * P);
    if (!st && (params->flags & SPH_FLAG_PROFILE)) {
        for (auto &pair : c->ev)
            for (auto &ev : pair)
                if (!st && hipEventCreate(&ev) != hipSuccess) st = fail(SPH_ERR_HIP, "hipEventCreate failed");
    }
    for (auto &ev : c->span)
        if (!st && hipEventCreate(&ev) != hipSuccess) st = fail(SPH_ERR_HIP, "hipEventCreate failed");
    if (st) {
        std::string msg = g_last_error;
        release(c);
        g_last_error = msg;
        return st;
    }
    *out = c;
    return SPH_OK;
}

void sph_destroy(sph_ctx *ctx)
{
    release(ctx);
}

sph_status sph_upload(sph_ctx *c, const float *pos3, const float *vel3, const uint32_t *ids, uint32_t n)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    if (n > 0 && !pos3) return fail(SPH_ERR_INVALID, "sph_upload: null positions");
    if (n > c->cap)
        return fail(SPH_ERR_CAPACITY, "sph_upload: " + std::to_string(n) + " particles exceed capacity " +
                                          std::to_string(c->cap));
    SPH_HIP(hipStreamSynchronize(c->stream));
    sph_status st = write_tail(c, pos3, vel3, ids, 0, n);
    if (st) return st;
    c->scene.n = n;
    c->keys_valid = false;
    c->grid_valid = false;
    c->next_phase = SPH_PHASE_GRID;
    return SPH_OK;
}

sph_status sph_append(sph_ctx *c, const float *pos3, const float *vel3, const uint32_t *ids, uint32_t n)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    if (n == 0) return SPH_OK;
    if (!pos3) return fail(SPH_ERR_INVALID, "sph_append: null positions");
    if (c->next_phase != SPH_PHASE_GRID)
        return fail(SPH_ERR_STATE, "sph_append: only between steps (before the grid phase)");
    if ((uint64_t) c->scene.n + n > c->cap)
        return fail(SPH_ERR_CAPACITY, "sph_append: " + std::to_string(c->scene.n + (uint64_t) n) +
                                          " particles exceed capacity " + std::to_string(c->cap));
    SPH_HIP(hipStreamSynchronize(c->stream));
    sph_status st = write_tail(c, pos3, vel3, ids, c->scene.n, n);
    if (st) return st;
    c->scene.n += n;
    c->keys_valid = false;
    c->grid_valid = false;
    return SPH_OK;
}

sph_status sph_set_gravity(sph_ctx *c, float gx, float gy, float gz)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    c->scene.gx = gx;
    c->scene.gy = gy;
    c->scene.gz = gz;
    return SPH_OK;
}

sph_status sph_phase(sph_ctx *c, int phase, float *ms_out)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    if (ms_out) SPH_HIP(hipEventRecord(c->span[0], c->stream));
    sph_status st = run_phase(c, phase);
    if (st) return st;
    if (ms_out) {
        SPH_HIP(hipEventRecord(c->span[1], c->stream));
        SPH_HIP(hipEventSynchronize(c->span[1]));
        SPH_HIP(hipEventElapsedTime(ms_out, c->span[0], c->span[1]));
    }
    return SPH_OK;
}

sph_status sph_step(sph_ctx *c, uint32_t nsteps, float *ms_out)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    if (c->scene.n == 0) return fail(SPH_ERR_STATE, "sph_step: no particles uploaded");
    if (c->next_phase != SPH_PHASE_GRID) return fail(SPH_ERR_STATE, "sph_step: a phase sequence is in progress");
    if (ms_out) SPH_HIP(hipEventRecord(c->span[0], c->stream));
    for (uint32_t s = 0; s < nsteps; ++s) {
        sph_status st = check_fault(c);
        if (!st) st = run_grid(c);
        if (!st) st = run_density(c);
        if (!st) st = run_fused_tail(c);
        if (st) return st;
        c->steps_done++;
    }
    if (ms_out) {
        SPH_HIP(hipEventRecord(c->span[1], c->stream));
        SPH_HIP(hipEventSynchronize(c->span[1]));
        SPH_HIP(hipEventElapsedTime(ms_out, c->span[0], c->span[1]));
    }
    return SPH_OK;
}

sph_status sph_download(sph_ctx *c, float *pos3, float *vel3, uint32_t *ids, uint32_t n, int by_id)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    if (n != c->scene.n)
        return fail(SPH_ERR_INVALID, "sph_download: n = " + std::to_string(n) + " but the context holds " +
                                         std::to_string(c->scene.n));
    if (n == 0) return SPH_OK;
    std::vector<uint32_t> id((size_t) n);
    sph_status st = copy_d2h(c, id.data(), c->ids[c->id_cur], n * sizeof(uint32_t));
    if (st) return st;
    if (by_id)
        for (uint32_t i = 0; i < n; ++i)
            if (id[i] >= n)
                return fail(SPH_ERR_STATE, "sph_download: by_id needs ids 0..n-1, found " + std::to_string(id[i]));
    if (ids) {
        for (uint32_t i = 0; i < n; ++i) ids[by_id ? id[i] : i] = id[i];
    }
    std::vector<float> buf((size_t) n);
    for (int comp = 0; comp < 3; ++comp) {
        if (pos3) {
            st = copy_d2h(c, buf.data(), c->pos[c->cur][comp], n * sizeof(float));
            if (st) return st;
            for (uint32_t i = 0; i < n; ++i) pos3[3 * (size_t) (by_id ? id[i] : i) + comp] = buf[i];
        }
        if (vel3) {
            st = copy_d2h(c, buf.data(), c->vel[c->cur][comp], n * sizeof(float));
            if (st) return st;
            for (uint32_t i = 0; i < n; ++i) vel3[3 * (size_t) (by_id ? id[i] : i) + comp] = buf[i];
        }
    }
    return SPH_OK;
}

sph_status sph_debug(sph_ctx *c, int kind, void *out, size_t bytes)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    if (!out) return fail(SPH_ERR_INVALID, "sph_debug: null output");
    const uint32_t n = c->scene.n;
    const size_t need_n4 = (size_t) n * 4;
    auto need = [&](size_t b) -> sph_status {
        if (bytes < b)
            return fail(SPH_ERR_INVALID, "sph_debug: buffer of " + std::to_string(bytes) + " B, need " +
                                             std::to_string(b));
        return SPH_OK;
    };
    sph_status st = SPH_OK;
    switch (kind) {
    case SPH_DEBUG_CELL_PRESORT:
        if ((st = need(need_n4))) return st;
        return copy_d2h(c, out, c->key_h, need_n4);
    case SPH_DEBUG_PERM:
        if ((st = need(need_n4))) return st;
        return copy_d2h(c, out, c->val_s, need_n4);
    case SPH_DEBUG_CELL:
        if ((st = need(need_n4))) return st;
        return copy_d2h(c, out, c->key_s, need_n4);
    case SPH_DEBUG_CELL_START:
        if ((st = need(((size_t) c->ncells + 1) * 4))) return st;
        return copy_d2h(c, out, c->cstart, ((size_t) c->ncells + 1) * 4);
    case SPH_DEBUG_NEIGHBOURS:
        if (!(c->prm.flags & SPH_FLAG_DIAGNOSTICS))
            return fail(SPH_ERR_STATE, "sph_debug: neighbour counts need SPH_FLAG_DIAGNOSTICS");
        if ((st = need(need_n4))) return st;
        return copy_d2h(c, out, c->ncount, need_n4);
    case SPH_DEBUG_DENSITY:
        if ((st = need(need_n4))) return st;
        return copy_d2h(c, out, c->rho, need_n4);
    case SPH_DEBUG_PRESSURE: {
        if ((st = need(need_n4))) return st;
        st = copy_d2h(c, out, c->rho, need_n4);
        if (st) return st;
        float *f = (float *) out;
        for (uint32_t i = 0; i < n; ++i) f[i] = sph::kGasStiffness * (f[i] - sph::kRestDensity);
        return SPH_OK;
    }
    case SPH_DEBUG_ACCEL: {
        if ((st = need(need_n4 * 3))) return st;
        std::vector<float> buf((size_t) n);
        float *f = (float *) out;
        for (int a = 0; a < 3; ++a) {
            st = copy_d2h(c, buf.data(), c->acc[a], need_n4);
            if (st) return st;
            for (uint32_t i = 0; i < n; ++i) f[3 * (size_t) i + a] = buf[i];
        }
        return SPH_OK;
    }
    case SPH_DEBUG_IDS:
        if ((st = need(need_n4))) return st;
        return copy_d2h(c, out, c->ids[c->id_cur], need_n4);
    default:
        return fail(SPH_ERR_INVALID, "sph_debug: unknown kind " + std::to_string(kind));
    }
}

sph_status sph_count(const sph_ctx *c, uint32_t *n)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    if (!n) return fail(SPH_ERR_INVALID, "sph_count: null output");
    *n = c->scene.n;
    return SPH_OK;
}

sph_status sph_grid_dims(const sph_ctx *c, int *xyz, uint32_t *ncells)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    if (xyz) {
        xyz[0] = c->scene.rx;
        xyz[1] = c->scene.ry;
        xyz[2] = c->scene.rz;
    }
    if (ncells) *ncells = c->ncells;
    return SPH_OK;
}

sph_status sph_device_name(char *buf, size_t len)
{
    if (!buf || len == 0) return fail(SPH_ERR_INVALID, "sph_device_name: empty buffer");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SPH_ERR_NO_DEVICE, "no HIP device visible");
    int dev = 0;
    SPH_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SPH_HIP(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return SPH_OK;
}

sph_status sph_sync(sph_ctx *c)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    SPH_HIP(hipStreamSynchronize(c->stream));
    return SPH_OK;
}

sph_status sph_stage_times(sph_ctx *c, float *ms, uint32_t *launches, int reset)
{
    if (!check_ctx(c)) return SPH_ERR_INVALID;
    if (!(c->prm.flags & SPH_FLAG_PROFILE)) return fail(SPH_ERR_STATE, "sph_stage_times: needs SPH_FLAG_PROFILE");
    sph_status st = drain_profile(c);
    if (st) return st;
    for (int i = 0; i < SPH_STAGE_COUNT; ++i) {
        if (ms) ms[i] = c->stage_ms[i];
        if (launches) launches[i] = c->stage_launches[i];
        if (reset) {
            c->stage_ms[i] = 0.0f;
            c->stage_launches[i] = 0;
        }
    }
    return SPH_OK;
}

const char *sph_last_error(void)
{
    return g_last_error.c_str();
}

}  // extern "C"
