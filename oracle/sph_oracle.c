/*
 * sph_oracle.c -- CPU restatement of the reference's CCPUParticleSimulator.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product; see sph_oracle.h).
 *
 * Arithmetic follows the reference expression by expression: QVector3D
 * operations are float (Qt 5.9 qvector3d.h inline operators, and the
 * out-of-line lengthSquared/dotProduct which sum x*x + y*y + z*z left to right
 * in float), the SPH kernels are double (src/CCPUParticleSimulator.cpp:9-30),
 * and a double value meeting a QVector3D is narrowed to float first because
 * QVector3D's scalar operators take `float`. Built with -ffp-contract=off and
 * without builtin pow folding so that every pow() is the libm call the
 * reference's default (-O0) build makes.
 */
#include "sph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* include/CParticle.h:80-84 */
static const float P_H = 0.0457f;
static const float P_VISCOSITY = 3.5f;
static const float P_MASS = 0.02f;
static const float P_GAS_STIFFNESS = 3.0f;
static const float P_REST_DENSITY = 998.29f;
/* include/CBaseParticleSimulator.h:19 */
#define GRAVITY_ACCELERATION (-9.80665f)
/* include/CCollisionGeometry.h:20-21 (double literals) */
#define WALL_K 10000.0
#define WALL_DAMPING (-0.9)

typedef struct { float x, y, z; } v3;

static inline v3 v3_make(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 v3_sub(v3 a, v3 b) { return v3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_add(v3 a, v3 b) { return v3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
/* operator*(QVector3D, float) and operator*(float, QVector3D): vector.xp * factor */
static inline v3 v3_mul(v3 a, float f) { return v3_make(a.x * f, a.y * f, a.z * f); }
/* operator/(QVector3D, float) */
static inline v3 v3_div(v3 a, float f) { return v3_make(a.x / f, a.y / f, a.z / f); }
/* QVector3D::lengthSquared / dotProduct: left-to-right float sums */
static inline float v3_len2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline float v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

typedef struct {
    uint32_t *items;
    uint32_t size, cap;
} bucket;

static void bucket_push(bucket *b, uint32_t v)
{
    if (b->size == b->cap) {
        uint32_t nc = b->cap ? b->cap * 2 : 8;
        uint32_t *ni = (uint32_t *) realloc(b->items, (size_t) nc * sizeof(uint32_t));
        if (!ni) abort();
        b->items = ni;
        b->cap = nc;
    }
    b->items[b->size++] = v;
}

struct orc_sim {
    v3 box;            /* m_boxSize */
    v3 gravity;        /* gravity */
    float dt;          /* dt = 0.01f, src/CBaseParticleSimulator.cpp:7 */
    int rx, ry, rz;    /* CGrid resolution */
    int ncells;
    int scenario;
    int order_mode;
    int diag;
    uint32_t count, max_count, cap;
    /* per particle, indexed by id (CParticle members) */
    v3 *pos, *vel, *acc;
    float *density, *pressure;
    uint32_t *cell;    /* cell assigned by the last updateGrid */
    uint32_t *order;   /* canonical mode: array order of ids (concatenated buckets) */
    uint32_t *tmp;     /* canonical mode: scratch for the stable counting sort */
    uint32_t *coff;    /* canonical mode: per-cell offsets (ncells + 1) */
    uint32_t *ncount;  /* diagnostics: accepted neighbours incl. self */
    v3 *fscale;        /* diagnostics: absolute force scale */
    bucket *cells;     /* CGrid::m_data */
};

/* Wpoly6 / WspikyGradient / WviscosityLaplacian coefficients and hSquared,
 * src/CCPUParticleSimulator.cpp:11-12, 19, 26. CParticle::h is float; h*h is a
 * float product widened to double. */
static double c_poly6, c_spiky, c_visc, c_hsq;
static int c_ready;

static void init_constants(void)
{
    if (c_ready) return;
    c_poly6 = 315.0 / (64.0 * M_PI * pow((double) P_H, 9.0));
    c_spiky = -45.0 / (M_PI * pow((double) P_H, 6.0));
    c_visc = 45.0 / (M_PI * pow((double) P_H, 6.0));
    {
        volatile float hh = P_H * P_H;
        c_hsq = (double) hh;
    }
    c_ready = 1;
}

void orc_constants(double *out)
{
    init_constants();
    out[0] = c_poly6;
    out[1] = c_spiky;
    out[2] = c_visc;
    out[3] = c_hsq;
    out[4] = (double) P_H;
    out[5] = (double) P_MASS;
    out[6] = (double) P_REST_DENSITY;
}

/* src/CCPUParticleSimulator.cpp:9-15 */
static inline double w_poly6(double r2) { return c_poly6 * pow(c_hsq - r2, 3.0); }

/* src/CCPUParticleSimulator.cpp:17-22: the double scalar narrows to float when
 * it multiplies the QVector3D, and the division is by (float) radius. */
static inline v3 w_spiky_gradient(v3 diff, double r2)
{
    double radius = sqrt(r2);
    float f = (float) (c_spiky * pow((double) P_H - radius, 2.0));
    return v3_div(v3_mul(diff, f), (float) radius);
}

/* src/CCPUParticleSimulator.cpp:24-30 */
static inline double w_visc_laplacian(double r2)
{
    double radius = sqrt(r2);
    return c_visc * ((double) P_H - radius);
}

/* src/CBaseParticleSimulator.cpp:27-31: (int) ceil(float box / float h) */
int orc_grid_res(float box)
{
    volatile float q = box / P_H;
    return (int) ceilf(q);
}

/* x86-64 cvttsd2si semantics of the reference's (int) cast: values outside the
 * int range and NaN become INT_MIN (which the clamp then maps to cell 0). */
static inline int to_int_x86(double q)
{
    if (q >= -2147483648.0 && q < 2147483648.0) return (int) q;
    return INT_MIN;
}

static inline int clamp_cell(int c, int res)
{
    if (c < 0) return 0;
    if (c >= res) return res - 1;
    return c;
}

/* src/CCPUParticleSimulator.cpp:46-69 */
static inline void hash_xyz(v3 p, v3 box, int rx, int ry, int rz, int *cx, int *cy, int *cz)
{
    int x = to_int_x86(floor(((double) p.x + (double) box.x / 2.0) / (double) P_H));
    int y = to_int_x86(floor(((double) p.y + (double) box.y / 2.0) / (double) P_H));
    int z = to_int_x86(floor(((double) p.z + (double) box.z / 2.0) / (double) P_H));
    *cx = clamp_cell(x, rx);
    *cy = clamp_cell(y, ry);
    *cz = clamp_cell(z, rz);
}

/* include/CGrid.h:25-28 */
static inline uint32_t lin(int x, int y, int z, int rx, int ry)
{
    return (uint32_t) (x + y * rx + z * rx * ry);
}

uint32_t orc_hash(float px, float py, float pz, float bx, float by, float bz, int rx, int ry, int rz)
{
    int cx, cy, cz;
    hash_xyz(v3_make(px, py, pz), v3_make(bx, by, bz), rx, ry, rz, &cx, &cy, &cz);
    return lin(cx, cy, cz, rx, ry);
}

void orc_hash_many(const float *pos3, uint32_t n, float bx, float by, float bz,
                   int rx, int ry, int rz, uint32_t *cell)
{
    for (uint32_t i = 0; i < n; i++)
        cell[i] = orc_hash(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], bx, by, bz, rx, ry, rz);
}

uint32_t orc_dam_break_count(float box)
{
    /* src/CBaseParticleSimulator.cpp:40-41 */
    double half = (double) (P_H / 2.0f);
    return (unsigned) (ceil((double) box / half) * ceil((double) box / half) *
                       ceil((double) (box / 4) / half));
}

static void ensure_cap(orc_sim *s, uint32_t need)
{
    if (need <= s->cap) return;
    uint32_t nc = s->cap ? s->cap : 1024;
    while (nc < need) nc *= 2;
#define GROW(ptr, T) do { T *np = (T *) realloc(s->ptr, (size_t) nc * sizeof(T)); if (!np) abort(); s->ptr = np; } while (0)
    GROW(pos, v3); GROW(vel, v3); GROW(acc, v3);
    GROW(density, float); GROW(pressure, float);
    GROW(cell, uint32_t); GROW(ncount, uint32_t); GROW(fscale, v3);
    GROW(order, uint32_t); GROW(tmp, uint32_t);
#undef GROW
    s->cap = nc;
}

orc_sim *orc_create(float bx, float by, float bz, uint32_t capacity, int scenario)
{
    init_constants();
    orc_sim *s = (orc_sim *) calloc(1, sizeof(orc_sim));
    if (!s) return NULL;
    s->box = v3_make(bx, by, bz);
    s->gravity = v3_make(0.0f, GRAVITY_ACCELERATION, 0.0f);
    s->dt = 0.01f;
    s->rx = orc_grid_res(bx);
    s->ry = orc_grid_res(by);
    s->rz = orc_grid_res(bz);
    s->ncells = s->rx * s->ry * s->rz;
    s->scenario = scenario;
    s->cells = (bucket *) calloc((size_t) s->ncells, sizeof(bucket));
    s->coff = (uint32_t *) calloc((size_t) s->ncells + 1, sizeof(uint32_t));
    if (!s->cells || !s->coff) abort();
    ensure_cap(s, capacity ? capacity : 1024);
    return s;
}

void orc_destroy(orc_sim *s)
{
    if (!s) return;
    for (int c = 0; c < s->ncells; c++) free(s->cells[c].items);
    free(s->cells);
    free(s->pos); free(s->vel); free(s->acc);
    free(s->density); free(s->pressure);
    free(s->cell); free(s->ncount); free(s->fscale);
    free(s->order); free(s->tmp); free(s->coff);
    free(s);
}

uint32_t orc_count(const orc_sim *s) { return s->count; }
uint32_t orc_max_count(const orc_sim *s) { return s->max_count; }
void orc_grid_dims(const orc_sim *s, int *xyz) { xyz[0] = s->rx; xyz[1] = s->ry; xyz[2] = s->rz; }
void orc_set_gravity(orc_sim *s, float gx, float gy, float gz) { s->gravity = v3_make(gx, gy, gz); }
void orc_set_order_mode(orc_sim *s, int mode) { s->order_mode = mode; }
void orc_set_diagnostics(orc_sim *s, int on) { s->diag = on; }

/* src/CBaseParticleSimulator.cpp:67-74 and the CParticle constructor
 * (src/CParticle.cpp:3-8): position, velocity from the initial velocity,
 * acceleration 0; Physics density/pressure 0. */
int orc_add_particle(orc_sim *s, float x, float y, float z, float vx, float vy, float vz)
{
    ensure_cap(s, s->count + 1);
    uint32_t id = s->count;
    s->pos[id] = v3_make(x, y, z);
    s->vel[id] = v3_make(vx, vy, vz);
    s->acc[id] = v3_make(0.0f, 0.0f, 0.0f);
    s->density[id] = 0.0f;
    s->pressure[id] = 0.0f;
    s->cell[id] = 0;
    s->ncount[id] = 0;
    s->fscale[id] = v3_make(0.0f, 0.0f, 0.0f);
    s->order[id] = id;
    bucket_push(&s->cells[0], id);
    s->count++;
    return (int) id;
}

void orc_add_particles(orc_sim *s, const float *pos3, const float *vel3, uint32_t n)
{
    ensure_cap(s, s->count + n);
    for (uint32_t i = 0; i < n; i++)
        orc_add_particle(s, pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], vel3 ? vel3[3 * i] : 0.0f,
                         vel3 ? vel3[3 * i + 1] : 0.0f, vel3 ? vel3[3 * i + 2] : 0.0f);
}

/* src/CBaseParticleSimulator.cpp:38-65 */
int orc_setup_scene(orc_sim *s)
{
    double half = (double) (P_H / 2.0f);
    unsigned calc = (unsigned) (ceil((double) s->box.z / half) * ceil((double) s->box.y / half) *
                                ceil((double) (s->box.x / 4) / half));
    ensure_cap(s, calc);
    if (s->scenario == ORC_DAM_BREAK) {
        v3 off = v3_make(-s->box.x / 2.0f, -s->box.y / 2.0f, -s->box.z / 2.0f);
        for (float y = 0; y < s->box.y; y = (float) ((double) y + half)) {
            for (float x = 0; (double) x < (double) s->box.x / 4.0; x = (float) ((double) x + half)) {
                for (float z = 0; z < s->box.z; z = (float) ((double) z + half)) {
                    orc_add_particle(s, x + off.x, y + off.y, z + off.z, 0.0f, 0.0f, 0.0f);
                }
            }
        }
        s->max_count = s->count;
        return (int) s->count;
    }
    s->max_count = calc;
    return 0;
}

/* src/CBaseParticleSimulator.cpp:187-210 */
static void generate_particles(orc_sim *s)
{
    if (s->scenario != ORC_FOUNTAIN) return;
    const int per_iteration = 7;
    if ((unsigned) s->count >= (unsigned) (s->max_count - (unsigned) per_iteration)) return;
    const float half = P_H / 2.0f;
    float oy = -s->box.y / 2.0f;
    float vy = s->box.y * 3.2f;
    orc_add_particle(s, 0, oy, 0, 0, vy, 0);
    orc_add_particle(s, -half, oy, 0, 0, vy, 0);
    orc_add_particle(s, half, oy, 0, 0, vy, 0);
    orc_add_particle(s, -P_H / 4, oy, -half, 0, vy, 0);
    orc_add_particle(s, P_H / 4, oy, -half, 0, vy, 0);
    orc_add_particle(s, -P_H / 4, oy, half, 0, vy, 0);
    orc_add_particle(s, P_H / 4, oy, half, 0, vy, 0);
}

/* Canonical mode: hash every particle, then replace the array order by its
 * stable counting sort by cell; buckets are the runs of the new order. */
static void update_grid_canonical(orc_sim *s)
{
    const int rx = s->rx, ry = s->ry, rz = s->rz;
    const uint32_t n = s->count;
    for (uint32_t i = 0; i < n; i++) {
        int nx, ny, nz;
        hash_xyz(s->pos[i], s->box, rx, ry, rz, &nx, &ny, &nz);
        s->cell[i] = lin(nx, ny, nz, rx, ry);
    }
    memset(s->coff, 0, ((size_t) s->ncells + 1) * sizeof(uint32_t));
    for (uint32_t k = 0; k < n; k++) s->coff[s->cell[s->order[k]] + 1]++;
    for (int c = 0; c < s->ncells; c++) s->coff[c + 1] += s->coff[c];
    for (int c = 0; c < s->ncells; c++) s->cells[c].size = 0;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t id = s->order[k];
        s->tmp[s->coff[s->cell[id]]++] = id;
        bucket_push(&s->cells[s->cell[id]], id);
    }
    memcpy(s->order, s->tmp, (size_t) n * sizeof(uint32_t));
}

/* src/CCPUParticleSimulator.cpp:32-91 */
void orc_update_grid(orc_sim *s)
{
    const int rx = s->rx, ry = s->ry, rz = s->rz;
    if (s->order_mode == ORC_ORDER_CANONICAL) {
        update_grid_canonical(s);
        return;
    }
    for (int x = 0; x < rx; x++) {
        for (int y = 0; y < ry; y++) {
            for (int z = 0; z < rz; z++) {
                bucket *b = &s->cells[lin(x, y, z, rx, ry)];
                for (long p = 0; p < (long) b->size; p++) {
                    uint32_t id = b->items[p];
                    int nx, ny, nz;
                    hash_xyz(s->pos[id], s->box, rx, ry, rz, &nx, &ny, &nz);
                    s->cell[id] = lin(nx, ny, nz, rx, ry);
                    if (x != nx || y != ny || z != nz) {
                        bucket_push(&s->cells[lin(nx, ny, nz, rx, ry)], id);
                        b->items[p] = b->items[b->size - 1];
                        b->size--;
                        p--;
                    }
                }
            }
        }
    }
}

/* In-range neighbour cells of (x,y,z) in the mode's visiting order. The
 * reference's `continue`/`break` bounds tests (src/CCPUParticleSimulator.cpp:
 * 108-118) skip exactly the out-of-range offsets. */
static int neighbour_cells(const orc_sim *s, int x, int y, int z, uint32_t *out)
{
    const int rx = s->rx, ry = s->ry, rz = s->rz;
    int n = 0;
    if (s->order_mode == ORC_ORDER_CANONICAL) {
        for (int oz = -1; oz <= 1; oz++) {
            if (z + oz < 0 || z + oz >= rz) continue;
            for (int oy = -1; oy <= 1; oy++) {
                if (y + oy < 0 || y + oy >= ry) continue;
                for (int ox = -1; ox <= 1; ox++) {
                    if (x + ox < 0 || x + ox >= rx) continue;
                    out[n++] = lin(x + ox, y + oy, z + oz, rx, ry);
                }
            }
        }
        return n;
    }
    for (int ox = -1; ox <= 1; ox++) {
        if (x + ox < 0) continue;
        if (x + ox >= rx) break;
        for (int oy = -1; oy <= 1; oy++) {
            if (y + oy < 0) continue;
            if (y + oy >= ry) break;
            for (int oz = -1; oz <= 1; oz++) {
                if (z + oz < 0) continue;
                if (z + oz >= rz) break;
                out[n++] = lin(x + ox, y + oy, z + oz, rx, ry);
            }
        }
    }
    return n;
}

/* src/CCPUParticleSimulator.cpp:93-141 */
void orc_update_density_pressure(orc_sim *s)
{
    const int rx = s->rx, ry = s->ry, rz = s->rz;
    for (int x = 0; x < rx; x++) {
        for (int y = 0; y < ry; y++) {
            for (int z = 0; z < rz; z++) {
                const bucket *b = &s->cells[lin(x, y, z, rx, ry)];
                uint32_t ncl[27];
                const int ncell = b->size ? neighbour_cells(s, x, y, z, ncl) : 0;
                for (uint32_t k = 0; k < b->size; k++) {
                    uint32_t i = b->items[k];
                    v3 pi = s->pos[i];
                    float density = 0.0f;
                    uint32_t nc = 0;
                    for (int t = 0; t < ncell; t++) {
                        const bucket *nb = &s->cells[ncl[t]];
                        for (uint32_t q = 0; q < nb->size; q++) {
                            v3 d = v3_sub(pi, s->pos[nb->items[q]]);
                            double r2 = (double) v3_len2(d);
                            if (r2 <= c_hsq) {
                                density = (float) ((double) density + w_poly6(r2));
                                nc++;
                            }
                        }
                    }
                    density = density * P_MASS;
                    s->density[i] = density;
                    s->pressure[i] = P_GAS_STIFFNESS * (density - P_REST_DENSITY);
                    if (s->diag) s->ncount[i] = nc;
                }
            }
        }
    }
}

/* CCollisionGeometry::inverseBoundingBoxBounce, src/CCollisionGeometry.cpp:117-132,
 * with the six walls of sBoundingBox (include/CCollisionGeometry.h:99-119) built
 * from the Qt3D cuboid of extents box, whose vertices lie at +-extent/2. */
static v3 wall_accel(v3 pos, v3 vel, v3 box, v3 *absacc)
{
    const float mnx = -box.x / 2.0f, mny = -box.y / 2.0f, mnz = -box.z / 2.0f;
    const float mxx = box.x / 2.0f, mxy = box.y / 2.0f, mxz = box.z / 2.0f;
    const v3 normal[6] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const v3 wpos[6] = {{mnx, 0, 0}, {0, mny, 0}, {0, 0, mnz}, {mxx, 0, 0}, {0, mxy, 0}, {0, 0, mxz}};
    v3 acc = v3_make(0.0f, 0.0f, 0.0f);
    for (int w = 0; w < 6; w++) {
        v3 inv = v3_mul(normal[w], -1.0f);
        double d = (double) v3_dot(v3_sub(wpos[w], pos), inv) + 0.01;
        if (d > 0.0) {
            v3 t1 = v3_mul(v3_mul(inv, (float) WALL_K), (float) d);
            v3 t2 = v3_mul(inv, (float) (WALL_DAMPING * (double) v3_dot(vel, inv)));
            acc = v3_add(acc, t1);
            acc = v3_add(acc, t2);
            if (absacc) {
                absacc->x += fabsf(t1.x) + fabsf(t2.x);
                absacc->y += fabsf(t1.y) + fabsf(t2.y);
                absacc->z += fabsf(t1.z) + fabsf(t2.z);
            }
        }
    }
    return acc;
}

void orc_wall_accel(const float *pos3, const float *vel3, float bx, float by, float bz, float *acc3)
{
    v3 a = wall_accel(v3_make(pos3[0], pos3[1], pos3[2]), v3_make(vel3[0], vel3[1], vel3[2]),
                      v3_make(bx, by, bz), NULL);
    acc3[0] = a.x; acc3[1] = a.y; acc3[2] = a.z;
}

/* src/CCPUParticleSimulator.cpp:143-203 */
void orc_update_forces(orc_sim *s)
{
    const int rx = s->rx, ry = s->ry, rz = s->rz;
    for (int x = 0; x < rx; x++) {
        for (int y = 0; y < ry; y++) {
            for (int z = 0; z < rz; z++) {
                const bucket *b = &s->cells[lin(x, y, z, rx, ry)];
                uint32_t ncl[27];
                const int ncell = b->size ? neighbour_cells(s, x, y, z, ncl) : 0;
                for (uint32_t k = 0; k < b->size; k++) {
                    uint32_t i = b->items[k];
                    v3 pi = s->pos[i], vi = s->vel[i];
                    float rho_i = s->density[i], p_i = s->pressure[i];
                    v3 f_gravity = v3_mul(s->gravity, rho_i);
                    v3 f_pressure = v3_make(0, 0, 0), f_viscosity = v3_make(0, 0, 0);
                    double ap[3] = {0, 0, 0}, av[3] = {0, 0, 0};
                    for (int t = 0; t < ncell; t++) {
                        {
                            {
                                const bucket *nb = &s->cells[ncl[t]];
                                for (uint32_t q = 0; q < nb->size; q++) {
                                    uint32_t j = nb->items[q];
                                    v3 d = v3_sub(pi, s->pos[j]);
                                    double r2 = (double) v3_len2(d);
                                    if (r2 <= c_hsq && i != j) {
                                        v3 grad = w_spiky_gradient(d, r2);
                                        double lap = w_visc_laplacian(r2);
                                        float pf = (float) ((double) p_i / pow((double) rho_i, 2.0) +
                                                            (double) s->pressure[j] / pow((double) s->density[j], 2.0));
                                        v3 tp = v3_mul(grad, pf);
                                        v3 tv = v3_div(v3_mul(v3_sub(s->vel[j], vi), (float) lap), s->density[j]);
                                        f_pressure = v3_add(f_pressure, tp);
                                        f_viscosity = v3_add(f_viscosity, tv);
                                        if (s->diag) {
                                            ap[0] += fabs((double) tp.x); ap[1] += fabs((double) tp.y); ap[2] += fabs((double) tp.z);
                                            av[0] += fabs((double) tv.x); av[1] += fabs((double) tv.y); av[2] += fabs((double) tv.z);
                                        }
                                    }
                                }
                            }
                        }
                    }
                    f_pressure = v3_mul(f_pressure, -P_MASS * rho_i);
                    f_viscosity = v3_mul(f_viscosity, P_VISCOSITY * P_MASS);
                    v3 a = v3_div(v3_add(v3_add(f_pressure, f_viscosity), f_gravity), rho_i);
                    v3 wabs = v3_make(0, 0, 0);
                    v3 w = wall_accel(pi, vi, s->box, s->diag ? &wabs : NULL);
                    s->acc[i] = v3_add(a, w);
                    if (s->diag) {
                        double gn = sqrt((double) s->gravity.x * s->gravity.x + (double) s->gravity.y * s->gravity.y +
                                         (double) s->gravity.z * s->gravity.z);
                        double ar = fabs((double) rho_i);
                        double sp = (double) P_MASS * ar, sv = (double) P_VISCOSITY * P_MASS;
                        double sc[3];
                        for (int c = 0; c < 3; c++)
                            sc[c] = (sp * ap[c] + sv * av[c] + gn * ar) / (ar > 0 ? ar : 1.0);
                        s->fscale[i] = v3_make((float) (sc[0] + wabs.x), (float) (sc[1] + wabs.y), (float) (sc[2] + wabs.z));
                    }
                }
            }
        }
    }
}

/* src/CCPUParticleSimulator.cpp:211-229 */
void orc_integrate(orc_sim *s)
{
    const float dt = s->dt;
    for (int c = 0; c < s->ncells; c++) {
        const bucket *b = &s->cells[c];
        for (uint32_t k = 0; k < b->size; k++) {
            uint32_t i = b->items[k];
            v3 p = s->pos[i];
            v3 np = v3_add(v3_add(p, v3_mul(s->vel[i], dt)), v3_mul(v3_mul(s->acc[i], dt), dt));
            v3 nv = v3_div(v3_sub(np, p), dt);
            s->pos[i] = np;
            s->vel[i] = nv;
        }
    }
}

/* src/CBaseParticleSimulator.cpp:116-144 */
void orc_step(orc_sim *s, uint32_t nsteps)
{
    for (uint32_t k = 0; k < nsteps; k++) {
        generate_particles(s);
        orc_update_grid(s);
        orc_update_density_pressure(s);
        orc_update_forces(s);
        /* updateCollisions is part of updateForces (src/CCPUParticleSimulator.cpp:205-209) */
        orc_integrate(s);
    }
}

void orc_get_state(const orc_sim *s, float *pos3, float *vel3, float *acc3, float *density, float *pressure)
{
    for (uint32_t i = 0; i < s->count; i++) {
        if (pos3) { pos3[3 * i] = s->pos[i].x; pos3[3 * i + 1] = s->pos[i].y; pos3[3 * i + 2] = s->pos[i].z; }
        if (vel3) { vel3[3 * i] = s->vel[i].x; vel3[3 * i + 1] = s->vel[i].y; vel3[3 * i + 2] = s->vel[i].z; }
        if (acc3) { acc3[3 * i] = s->acc[i].x; acc3[3 * i + 1] = s->acc[i].y; acc3[3 * i + 2] = s->acc[i].z; }
        if (density) density[i] = s->density[i];
        if (pressure) pressure[i] = s->pressure[i];
    }
}

void orc_set_state(orc_sim *s, const float *pos3, const float *vel3)
{
    for (uint32_t i = 0; i < s->count; i++) {
        if (pos3) s->pos[i] = v3_make(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]);
        if (vel3) s->vel[i] = v3_make(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2]);
    }
}

void orc_get_cells(const orc_sim *s, uint32_t *cell_id)
{
    memcpy(cell_id, s->cell, (size_t) s->count * sizeof(uint32_t));
}

void orc_get_buckets(const orc_sim *s, uint32_t *cell_start, uint32_t *ids)
{
    uint32_t off = 0;
    for (int c = 0; c < s->ncells; c++) {
        cell_start[c] = off;
        memcpy(ids + off, s->cells[c].items, (size_t) s->cells[c].size * sizeof(uint32_t));
        off += s->cells[c].size;
    }
    cell_start[s->ncells] = off;
}

void orc_get_neighbour_counts(const orc_sim *s, uint32_t *counts)
{
    memcpy(counts, s->ncount, (size_t) s->count * sizeof(uint32_t));
}

void orc_get_force_scale(const orc_sim *s, float *scale3)
{
    for (uint32_t i = 0; i < s->count; i++) {
        scale3[3 * i] = s->fscale[i].x;
        scale3[3 * i + 1] = s->fscale[i].y;
        scale3[3 * i + 2] = s->fscale[i].z;
    }
}

/* Canonical order of the HIP path: stable by cell over the current array
 * order (counting sort). */
void orc_stable_perm(const uint32_t *keys, uint32_t n, uint32_t ncells, uint32_t *perm)
{
    uint32_t *off = (uint32_t *) calloc((size_t) ncells + 1, sizeof(uint32_t));
    if (!off) abort();
    for (uint32_t i = 0; i < n; i++) off[keys[i] + 1]++;
    for (uint32_t c = 0; c < ncells; c++) off[c + 1] += off[c];
    for (uint32_t i = 0; i < n; i++) perm[off[keys[i]]++] = i;
    free(off);
}

void orc_cell_start(const uint32_t *sorted_keys, uint32_t n, uint32_t ncells, uint32_t *start)
{
    uint32_t i = 0;
    for (uint32_t c = 0; c <= ncells; c++) {
        while (i < n && sorted_keys[i] < c) i++;
        start[c] = i;
    }
}
