/*
 * dpd_oracle.c -- FP64 CPU restatement of the Osprey-DPD step.  TEST INFRASTRUCTURE:
 * only the tests, the smoke check and bench.py's CPU-baseline leg load this library.
 * Reference citations are given per function; see dpd_oracle.h for the overview.
 *
 * Parity status of each part (how it is pinned, in tests/):
 *   lcg            6-digit reference draws (tests/golden/rng.txt)
 *   philox4x32_10  rocRAND's philox4x32_10 engine (host build), tests/golden/philox_kat.txt
 *   step (LCG)     reference restart snapshots after 1 and 10 steps, thermostat on/off
 *   forces         two-bead closed forms, brute-force O(N^2) pair sum, Newton's 3rd law
 *   bonds/angles   reference snapshots (chains, lipids incl. phi0 > 0), closed forms,
 *                  finite differences of the angle potential for phi0 = 0
 *   kT, P          reference dmpchs history rows
 */
#include "dpd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- RNGs -------- */

/* src/CNTCell.cpp:2454-2465: state = 6364136223846793005*state + 1; U = (state>>32)/2^32 */
double orc_lcg_uniform(uint64_t* state)
{
    *state = 6364136223846793005ull * (*state) + 1ull;
    uint32_t top = (uint32_t)(*state >> 32);
    return (double)top * (1.0 / 4294967296.0);
}

/* Philox4x32 with 10 rounds, written from the published description (Salmon, Moraes,
 * Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11): each round
 * multiplies words 0 and 2 by M0 = 0xD2511F53, M1 = 0xCD9E8D57 and mixes
 *   (c0,c1,c2,c3) -> (hi(M1*c2)^c1^k0, lo(M1*c2), hi(M0*c0)^c3^k1, lo(M0*c0)),
 * with the key bumped by the Weyl constants (0x9E3779B9, 0xBB67AE85) between rounds. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* pair uniform of the GPU engine: counter (min id, max id, step lo, step hi), key = seed;
 * symmetric in (i, j), independent across steps; U = (word0 >> 8) * 2^-24 in [0, 1) */
double orc_pair_uniform(uint64_t seed, int64_t step, int32_t id_i, int32_t id_j)
{
    uint32_t a = (uint32_t)id_i, b = (uint32_t)id_j;
    uint32_t ctr[4] = {a < b ? a : b, a < b ? b : a, (uint32_t)(uint64_t)step,
                       (uint32_t)((uint64_t)step >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return (double)(out[0] >> 8) * (1.0 / 16777216.0);
}

/* ---------------------------------------------------------------- system ------ */

struct orc_sys {
    orc_params p;
    int32_t n, ncell;
    double L[3], halfL[3], w[3];
    double halfdt, halfdt2, invrootdt, lamdt;   /* src/CNTCell.cpp:180-203 */
    double *a, *g;
    int32_t *id, *type;
    double *x, *u, *v, *f, *xo, *vo, *fo, *stress; /* per bead, [3n] / [9n] */
    uint8_t* movable;
    /* intrusive doubly linked cell lists (std::list<CAbstractBead*> per cell) */
    int32_t *head, *tail, *next, *prev;
    uint8_t* ext;
    int32_t* nn27;                              /* 27 neighbours per cell */
    uint64_t lcg;
    int64_t step;
    /* topology */
    int32_t nb, np;
    int32_t *bh, *bt, *pf, *ps;
    double *bk, *bl0, *pk, *pphi0, *bdx, *blen;
};

static int cell_index(const orc_sys* s, int i, int j, int k)
{
    return i + s->p.nx * (j + s->p.ny * k);
}

/* CSimBox::MapCelltoIndex (src/SimBox.cpp:888-930): periodic neighbour index */
static int wrap_idx(int i, int n) { i %= n; return i < 0 ? i + n : i; }

static void push_front(orc_sys* s, int c, int b)
{
    s->prev[b] = -1;
    s->next[b] = s->head[c];
    if (s->head[c] >= 0) s->prev[s->head[c]] = b; else s->tail[c] = b;
    s->head[c] = b;
}

static void erase(orc_sys* s, int c, int b)
{
    int pv = s->prev[b], nx = s->next[b];
    if (pv >= 0) s->next[pv] = nx; else s->head[c] = nx;
    if (nx >= 0) s->prev[nx] = pv; else s->tail[c] = pv;
    s->prev[b] = s->next[b] = -1;
}

static int bin_of(const orc_sys* s, const double* x)
{
    /* CSimBox::AssignBeadstoCells: ix = (long)(x / width) */
    long ix = (long)(x[0] / s->w[0]), iy = (long)(x[1] / s->w[1]), iz = (long)(x[2] / s->w[2]);
    return (int)(s->p.nx * (s->p.ny * iz + iy) + ix);
}

orc_sys* orc_create(const orc_params* p, int32_t n, const int32_t* id, const int32_t* type,
                    const double* x, const double* unpbc, const double* v, const double* f,
                    const orc_topology* topo)
{
    orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
    s->p = *p;
    s->n = n;
    s->ncell = p->nx * p->ny * p->nz;
    s->w[0] = p->wx; s->w[1] = p->wy; s->w[2] = p->wz;
    s->L[0] = (double)p->nx * p->wx; s->L[1] = (double)p->ny * p->wy; s->L[2] = (double)p->nz * p->wz;
    for (int d = 0; d < 3; ++d) s->halfL[d] = 0.5 * s->L[d];
    s->halfdt = 0.5 * p->dt;
    s->halfdt2 = 0.5 * p->dt * p->dt;
    s->invrootdt = sqrt(24.0 * p->kT / p->dt);
    s->lamdt = p->lambda * p->dt;
    int nt = p->ntypes;
    s->a = (double*)malloc(sizeof(double) * nt * nt);
    s->g = (double*)malloc(sizeof(double) * nt * nt);
    memcpy(s->a, p->a, sizeof(double) * nt * nt);
    memcpy(s->g, p->g, sizeof(double) * nt * nt);
    s->p.a = s->a; s->p.g = s->g;
    s->id = (int32_t*)malloc(sizeof(int32_t) * n);
    s->type = (int32_t*)malloc(sizeof(int32_t) * n);
    memcpy(s->id, id, sizeof(int32_t) * n);
    memcpy(s->type, type, sizeof(int32_t) * n);
    size_t b3 = sizeof(double) * 3 * (size_t)n;
    s->x = (double*)malloc(b3); s->u = (double*)malloc(b3); s->v = (double*)malloc(b3);
    s->f = (double*)malloc(b3); s->xo = (double*)malloc(b3); s->vo = (double*)malloc(b3);
    s->fo = (double*)malloc(b3);
    s->stress = (double*)calloc(9 * (size_t)n, sizeof(double));
    memcpy(s->x, x, b3);
    memcpy(s->u, unpbc ? unpbc : x, b3);
    memcpy(s->v, v, b3);
    if (f) memcpy(s->f, f, b3); else memset(s->f, 0, b3);
    s->movable = (uint8_t*)malloc(n);
    memset(s->movable, 1, n);
    s->head = (int32_t*)malloc(sizeof(int32_t) * s->ncell);
    s->tail = (int32_t*)malloc(sizeof(int32_t) * s->ncell);
    s->next = (int32_t*)malloc(sizeof(int32_t) * n);
    s->prev = (int32_t*)malloc(sizeof(int32_t) * n);
    s->ext = (uint8_t*)malloc(s->ncell);
    s->nn27 = (int32_t*)malloc(sizeof(int32_t) * 27 * s->ncell);
    for (int c = 0; c < s->ncell; ++c) s->head[c] = s->tail[c] = -1;
    /* CSimBox::MakeCNTCells (src/SimBox.cpp:767-877) */
    for (int k = 0; k < p->nz; ++k)
        for (int j = 0; j < p->ny; ++j)
            for (int i = 0; i < p->nx; ++i) {
                int c = cell_index(s, i, j, k);
                s->ext[c] = (i == 0 || i == p->nx - 1 || j == 0 || j == p->ny - 1 || k == 0 || k == p->nz - 1);
                for (int g = -1; g < 2; ++g)
                    for (int b = -1; b < 2; ++b)
                        for (int a = -1; a < 2; ++a)
                            s->nn27[27 * c + 13 + a + 3 * b + 9 * g] =
                                cell_index(s, wrap_idx(i + a, p->nx), wrap_idx(j + b, p->ny), wrap_idx(k + g, p->nz));
            }
    /* CSimBox::AssignBeadstoCells: bead order, push_front (src/CNTCell.cpp:2339-2342) */
    for (int b = 0; b < n; ++b) {
        int c = bin_of(s, s->x + 3 * b);
        if (c < 0 || c >= s->ncell) c = 0;
        push_front(s, c, b);
    }
    s->lcg = p->seed;
    if (topo) {
        s->nb = topo->nbonds;
        s->np = topo->npairs;
        s->bh = (int32_t*)malloc(sizeof(int32_t) * (s->nb + 1));
        s->bt = (int32_t*)malloc(sizeof(int32_t) * (s->nb + 1));
        s->bk = (double*)malloc(sizeof(double) * (s->nb + 1));
        s->bl0 = (double*)malloc(sizeof(double) * (s->nb + 1));
        s->bdx = (double*)calloc(3 * (size_t)(s->nb + 1), sizeof(double));
        s->blen = (double*)calloc((size_t)(s->nb + 1), sizeof(double));
        if (s->nb) {
            memcpy(s->bh, topo->head, sizeof(int32_t) * s->nb);
            memcpy(s->bt, topo->tail, sizeof(int32_t) * s->nb);
            memcpy(s->bk, topo->k, sizeof(double) * s->nb);
            memcpy(s->bl0, topo->l0, sizeof(double) * s->nb);
        }
        s->pf = (int32_t*)malloc(sizeof(int32_t) * (s->np + 1));
        s->ps = (int32_t*)malloc(sizeof(int32_t) * (s->np + 1));
        s->pk = (double*)malloc(sizeof(double) * (s->np + 1));
        s->pphi0 = (double*)malloc(sizeof(double) * (s->np + 1));
        if (s->np) {
            memcpy(s->pf, topo->first, sizeof(int32_t) * s->np);
            memcpy(s->ps, topo->second, sizeof(int32_t) * s->np);
            memcpy(s->pk, topo->modulus, sizeof(double) * s->np);
            memcpy(s->pphi0, topo->phi0, sizeof(double) * s->np);
        }
    }
    return s;
}

void orc_destroy(orc_sys* s)
{
    if (!s) return;
    void* ptrs[] = {s->a, s->g, s->id, s->type, s->x, s->u, s->v, s->f, s->xo, s->vo, s->fo,
                    s->stress, s->movable, s->head, s->tail, s->next, s->prev, s->ext, s->nn27,
                    s->bh, s->bt, s->bk, s->bl0, s->bdx, s->blen, s->pf, s->ps, s->pk, s->pphi0};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(s);
}

void orc_set_gamma(orc_sys* s, const double* g)
{
    memcpy(s->g, g, sizeof(double) * s->p.ntypes * s->p.ntypes);
}

void orc_set_a(orc_sys* s, const double* a)
{
    memcpy(s->a, a, sizeof(double) * s->p.ntypes * s->p.ntypes);
}

/* CCNTCell::ChangeTimeStepConstants (src/CNTCell.cpp:357-377): every constant that
 * depends on the step, the half-step ones included */
void orc_set_timestep(orc_sys* s, double dt)
{
    if (!(dt > 0.0)) return;
    s->p.dt = dt;
    s->halfdt = 0.5 * dt;
    s->halfdt2 = 0.5 * dt * dt;
    s->invrootdt = sqrt(24.0 * s->p.kT / dt);
    s->lamdt = s->p.lambda * dt;
}

/* CCNTCell::ToggleDPDBeadForces (src/CNTCell.cpp:223-227): the half-step constants only
 * (dt = 0 stops forces moving beads or changing velocities) */
void orc_set_halfstep(orc_sys* s, double dt)
{
    s->halfdt = 0.5 * dt;
    s->halfdt2 = 0.5 * dt * dt;
}

/* CCNTCell::UpdatePos (src/CNTCell.cpp:1704-2262), DPD branch */
static void update_pos(orc_sys* s)
{
    const int32_t nx = s->p.nx, ny = s->p.ny;
    for (int c = 0; c < s->ncell; ++c) {
        int ci = c % nx, cj = (c / nx) % ny, ck = c / (nx * ny);
        double bl[3] = {(double)ci * s->w[0], (double)cj * s->w[1], (double)ck * s->w[2]};
        double tr[3] = {bl[0] + s->w[0], bl[1] + s->w[1], bl[2] + s->w[2]};
        for (int b = s->head[c]; b >= 0;) {
            if (!s->movable[b]) { b = s->next[b]; continue; }
            double* x = s->x + 3 * b;
            double* u = s->u + 3 * b;
            double* v = s->v + 3 * b;
            double* f = s->f + 3 * b;
            double dx[3];
            for (int d = 0; d < 3; ++d) {
                s->xo[3 * b + d] = x[d];
                s->vo[3 * b + d] = v[d];
                s->fo[3 * b + d] = f[d];
            }
            for (int d = 0; d < 3; ++d) dx[d] = s->p.dt * v[d] + s->halfdt2 * f[d];
            for (int d = 0; d < 3; ++d) x[d] += dx[d];
            for (int d = 0; d < 3; ++d) u[d] += dx[d];
            for (int d = 0; d < 3; ++d) v[d] = s->vo[3 * b + d] + s->lamdt * s->fo[3 * b + d];
            for (int d = 0; d < 3; ++d) f[d] = 0.0;
            int sh[3];
            for (int d = 0; d < 3; ++d) sh[d] = x[d] > tr[d] ? 1 : (x[d] < bl[d] ? -1 : 0);
            s->movable[b] = 0;
            if (sh[0] || sh[1] || sh[2]) {
                int nc = s->nn27[27 * c + 13 + sh[0] + 3 * sh[1] + 9 * sh[2]];
                if (s->ext[c] && s->ext[nc]) {
                    for (int d = 0; d < 3; ++d) {
                        if (sh[d] > 0 && x[d] > s->L[d]) x[d] -= s->L[d];
                        if (sh[d] < 0 && x[d] < 0.0) x[d] += s->L[d];
                    }
                }
                int nb = s->next[b];
                erase(s, c, b);
                push_front(s, nc, b);
                b = nb;
            } else {
                b = s->next[b];
            }
        }
    }
}

/* the 13 half-shell neighbours in the order CSimBox::MakeCNTCells stores them
 * (src/SimBox.cpp:830-867): gamma outer, beta, alpha inner */
static const int HALF[13][3] = {{1, -1, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}, {-1, 0, 1}, {0, 0, 1},
                                {1, 0, 1}, {-1, 1, 1}, {0, 1, 1}, {1, 1, 1}};

typedef struct pair_ctx {
    const double *a, *g;
    int nt;
    double invrootdt;
    int rng_mode;
    uint64_t* lcg;
    uint64_t seed;
    int64_t step;
    int thermostat;
    double* pe;
} pair_ctx;

/* body of the pair loop, src/CNTCell.cpp:1191-1301 and 1409-1528 */
static int pair_force_out(const pair_ctx* c, int t1, int t2, int id1, int id2, const double dx[3],
                          const double dv[3], double* f1, double* f2, double* stress1, double* nf_out);

static void pair_force(const pair_ctx* c, int t1, int t2, int id1, int id2, const double dx[3],
                       const double dv[3], double* f1, double* f2, double* stress1)
{
    pair_force_out(c, t1, t2, id1, id2, dx, dv, f1, f2, stress1, NULL);
}

/* as pair_force; returns 1 for an interacting pair and its force on bead 1 in nf_out */
static int pair_force_out(const pair_ctx* c, int t1, int t2, int id1, int id2, const double dx[3],
                          const double dv[3], double* f1, double* f2, double* stress1, double* nf_out)
{
    double dr2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    if (!(dr2 < 1.0)) return 0;
    double dr = sqrt(dr2);
    if (!(dr > 0.000000001)) return 0;
    double wr = (1.0 - dr);
    double wr2 = wr * wr;
    double conForce = c->a[t1 * c->nt + t2] * wr;
    double rdotv = (dx[0] * dv[0] + dx[1] * dv[1] + dx[2] * dv[2]) / dr;
    /* thermostat: 1 = dissipative + random (reference), 0 = neither (gamma = 0, as
     * ToggleDPDBeadThermostat), 2 = dissipative only (closed-form checks) */
    double gammap = c->thermostat ? c->g[t1 * c->nt + t2] * wr2 : 0.0;
    double dissForce = -gammap * rdotv;
    double U;
    if (c->rng_mode == ORC_RNG_LCG) U = orc_lcg_uniform(c->lcg);
    else U = orc_pair_uniform(c->seed, c->step, id1, id2);
    double randForce = c->thermostat == 2 ? 0.0 : sqrt(gammap) * c->invrootdt * (0.5 - U);
    double nf[3];
    nf[0] = (conForce + dissForce + randForce) * dx[0] / dr;
    nf[1] = (conForce + dissForce + randForce) * dx[1] / dr;
    nf[2] = (conForce + dissForce + randForce) * dx[2] / dr;
    f1[0] += nf[0]; f1[1] += nf[1]; f1[2] += nf[2];
    f2[0] -= nf[0]; f2[1] -= nf[1]; f2[2] -= nf[2];
    if (stress1) {
        stress1[0] += dx[0] * nf[0]; stress1[1] += dx[1] * nf[0]; stress1[2] += dx[2] * nf[0];
        stress1[3] += dx[0] * nf[1]; stress1[4] += dx[1] * nf[1]; stress1[5] += dx[2] * nf[1];
        stress1[6] += dx[0] * nf[2]; stress1[7] += dx[1] * nf[2]; stress1[8] += dx[2] * nf[2];
    }
    if (c->pe) *c->pe += 0.5 * c->a[t1 * c->nt + t2] * wr2;
    if (nf_out) { nf_out[0] = nf[0]; nf_out[1] = nf[1]; nf_out[2] = nf[2]; }
    return 1;
}

/* CCNTCell::UpdateForce over all cells in order (src/CNTCell.cpp:1144-1555) */
static void update_force(orc_sys* s, double* pe)
{
    pair_ctx ctx = {s->a, s->g, s->p.ntypes, s->invrootdt, s->p.rng_mode, &s->lcg, s->p.seed,
                    s->step, 1, pe};
    for (int c = 0; c < s->ncell; ++c) {
        int ci = c % s->p.nx, cj = (c / s->p.nx) % s->p.ny, ck = c / (s->p.nx * s->p.ny);
        for (int b1 = s->head[c]; b1 >= 0; b1 = s->next[b1]) {
            double* st = s->stress + 9 * b1;
            for (int q = 0; q < 9; ++q) st[q] = 0.0;
            const double* x1 = s->x + 3 * b1;
            const double* v1 = s->v + 3 * b1;
            for (int b2 = s->tail[c]; s->id[b2] != s->id[b1]; b2 = s->prev[b2]) {
                double dx[3], dv[3];
                for (int d = 0; d < 3; ++d) {
                    dx[d] = x1[d] - s->x[3 * b2 + d];
                    dv[d] = v1[d] - s->v[3 * b2 + d];
                }
                pair_force(&ctx, s->type[b1], s->type[b2], s->id[b1], s->id[b2], dx, dv,
                           s->f + 3 * b1, s->f + 3 * b2, st);
            }
            for (int h = 0; h < 13; ++h) {
                int nc = s->nn27[27 * c + 13 + HALF[h][0] + 3 * HALF[h][1] + 9 * HALF[h][2]];
                int pbc = s->ext[c] && s->ext[nc];
                for (int b2 = s->head[nc]; b2 >= 0; b2 = s->next[b2]) {
                    double dx[3], dv[3];
                    for (int d = 0; d < 3; ++d) {
                        dx[d] = x1[d] - s->x[3 * b2 + d];
                        dv[d] = v1[d] - s->v[3 * b2 + d];
                    }
                    if (pbc) {
                        for (int d = 0; d < 3; ++d) {
                            if (dx[d] > s->halfL[d]) dx[d] = dx[d] - s->L[d];
                            else if (dx[d] < -s->halfL[d]) dx[d] = dx[d] + s->L[d];
                        }
                    }
                    pair_force(&ctx, s->type[b1], s->type[b2], s->id[b1], s->id[b2], dx, dv,
                               s->f + 3 * b1, s->f + 3 * b2, st);
                }
            }
        }
        (void)ci; (void)cj; (void)ck;
    }
}

/* CBond::AddForce (src/Bond.cpp:111-162): unwrapped coordinates, stress on the head */
static void bond_forces(orc_sys* s, double* pe)
{
    for (int i = 0; i < s->nb; ++i) {
        int h = s->bh[i], t = s->bt[i];
        double* d = s->bdx + 3 * i;
        d[0] = s->u[3 * h + 0] - s->u[3 * t + 0];
        d[1] = s->u[3 * h + 1] - s->u[3 * t + 1];
        d[2] = s->u[3 * h + 2] - s->u[3 * t + 2];
        double L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        s->blen[i] = L;
        double k = s->bk[i], l0 = s->bl0[i];
        double fx = k * (l0 - L) * d[0] / L;
        double fy = k * (l0 - L) * d[1] / L;
        double fz = k * (l0 - L) * d[2] / L;
        s->f[3 * h + 0] += fx; s->f[3 * h + 1] += fy; s->f[3 * h + 2] += fz;
        s->f[3 * t + 0] -= fx; s->f[3 * t + 1] -= fy; s->f[3 * t + 2] -= fz;
        double* st = s->stress + 9 * h;
        st[0] += d[0] * fx; st[1] += d[1] * fx; st[2] += d[2] * fx;
        st[3] += d[0] * fy; st[4] += d[1] * fy; st[5] += d[2] * fy;
        st[6] += d[0] * fz; st[7] += d[1] * fz; st[8] += d[2] * fz;
        if (pe) *pe += 0.5 * k * (L - l0) * (L - l0);
    }
}

/* CBondPair::AddForce (src/BondPair.cpp:138-286); the bond vectors are those left by
 * the bond loop of the same step.  first = bond later in the chain, second = earlier. */
static void angle_one(double k, double phi0, const double* first, double l1, const double* second,
                      double l2, double ftail_first[3], double fhead_second[3], double fmid[3])
{
    ftail_first[0] = ftail_first[1] = ftail_first[2] = 0.0;
    fhead_second[0] = fhead_second[1] = fhead_second[2] = 0.0;
    fmid[0] = fmid[1] = fmid[2] = 0.0;
    const double magProduct = l1 * l2;
    if (!(magProduct > 0.0001)) return;
    const double cosPhi0 = cos(phi0), sinPhi0 = sin(phi0);
    const double b1MagSq = l1 * l1;
    const double b2MagSq = l2 * l2;
    const double b1Dotb2 = first[0] * second[0] + first[1] * second[1] + first[2] * second[2];
    const double b1b2Overb1Sq = b1Dotb2 / b1MagSq;
    const double b1b2Overb2Sq = b1Dotb2 / b2MagSq;
    const double cosPhiSq = b1b2Overb1Sq * b1b2Overb2Sq;
    double forceMag;
    if (fabs(b1Dotb2) > 0.000001) {
        double prefactor = sqrt(1.0 / cosPhiSq - 1.0);
        if (phi0 > 0.0) forceMag = k * (cosPhi0 - sinPhi0 / prefactor) / magProduct;
        else forceMag = k / magProduct;
    } else {
        forceMag = k / magProduct;
    }
    for (int d = 0; d < 3; ++d) ftail_first[d] = forceMag * (b1b2Overb1Sq * first[d] - second[d]);
    for (int d = 0; d < 3; ++d) fhead_second[d] = forceMag * (first[d] - b1b2Overb2Sq * second[d]);
    for (int d = 0; d < 3; ++d) fmid[d] = -ftail_first[d] - fhead_second[d];
}

/* CBondPair::AddPotentialEnergy (src/BondPair.cpp:320-378) */
static double angle_pe(double k, double phi0, const double* first, double l1, const double* second, double l2)
{
    const double magProduct = l1 * l2;
    if (!(magProduct > 0.0001)) return 0.0;
    const double cosPhi0 = cos(phi0), sinPhi0 = sin(phi0);
    const double b1Dotb2 = first[0] * second[0] + first[1] * second[1] + first[2] * second[2];
    const double cosPhiSq = (b1Dotb2 * b1Dotb2) / (magProduct * magProduct);
    const double sinPhi = sqrt(1.0 - cosPhiSq);
    if (b1Dotb2 > 0.000001) {
        if (phi0 > 0.0) return k * (1.0 - sqrt(cosPhiSq) * cosPhi0 - sinPhi * sinPhi0);
        return k * (1.0 - sqrt(cosPhiSq));
    } else if (b1Dotb2 < -0.000001) {
        if (phi0 > 0.0) return k * (1.0 + sqrt(cosPhiSq) * cosPhi0 - sinPhi * sinPhi0);
        return k * (1.0 + sqrt(cosPhiSq));
    }
    return k * (1.0 - sinPhi0);
}

static void angle_forces(orc_sys* s, double* pe)
{
    for (int i = 0; i < s->np; ++i) {
        int b1 = s->pf[i], b2 = s->ps[i];
        double ft[3], fh[3], fm[3];
        angle_one(s->pk[i], s->pphi0[i], s->bdx + 3 * b1, s->blen[b1], s->bdx + 3 * b2, s->blen[b2], ft, fh, fm);
        int tail_first = s->bt[b1], head_second = s->bh[b2], head_first = s->bh[b1];
        for (int d = 0; d < 3; ++d) s->f[3 * tail_first + d] += ft[d];
        for (int d = 0; d < 3; ++d) s->f[3 * head_second + d] += fh[d];
        for (int d = 0; d < 3; ++d) s->f[3 * head_first + d] += fm[d];
        if (pe) *pe += angle_pe(s->pk[i], s->pphi0[i], s->bdx + 3 * b1, s->blen[b1], s->bdx + 3 * b2, s->blen[b2]);
    }
}

/* CCNTCell::UpdateMom (src/CNTCell.cpp:2280-2310) */
static void update_mom(orc_sys* s)
{
    for (int b = 0; b < s->n; ++b) {
        s->movable[b] = 1;
        for (int d = 0; d < 3; ++d)
            s->v[3 * b + d] = s->vo[3 * b + d] + s->halfdt * (s->f[3 * b + d] + s->fo[3 * b + d]);
    }
}

void orc_run(orc_sys* s, int32_t nsteps, int64_t step0)
{
    for (int32_t t = 0; t < nsteps; ++t) {
        s->step = step0 + 1 + t;
        update_pos(s);
        update_force(s, NULL);
        bond_forces(s, NULL);
        angle_forces(s, NULL);
        update_mom(s);
    }
}

void orc_get(const orc_sys* s, double* x, double* unpbc, double* v, double* f)
{
    size_t b3 = sizeof(double) * 3 * (size_t)s->n;
    if (x) memcpy(x, s->x, b3);
    if (unpbc) memcpy(unpbc, s->u, b3);
    if (v) memcpy(v, s->v, b3);
    if (f) memcpy(f, s->f, b3);
}

/* CMonitor::CalculateSingleBeadData (src/Monitor.cpp:790-1041) */
void orc_observables(const orc_sys* s, orc_obs* o)
{
    memset(o, 0, sizeof(*o));
    double st[9] = {0};
    for (int b = 0; b < s->n; ++b) {
        for (int d = 0; d < 3; ++d) {
            o->momentum[d] += s->v[3 * b + d];
            o->sum_v2[d] += s->v[3 * b + d] * s->v[3 * b + d];
        }
        for (int q = 0; q < 9; ++q) st[q] += s->stress[9 * b + q];
        int c = bin_of(s, s->x + 3 * b);
        int ok = 0;
        if (c >= 0 && c < s->ncell)
            for (int q = s->head[c]; q >= 0; q = s->next[q]) if (q == b) { ok = 1; break; }
        if (!ok) o->wrong_cell++;
    }
    double V = s->L[0] * s->L[1] * s->L[2];
    double invV = 1.0 / V;
    o->kT = (o->sum_v2[0] + o->sum_v2[1] + o->sum_v2[2]) / (3.0 * s->n);
    o->pressure = 0.3333333333 * invV * (o->sum_v2[0] + o->sum_v2[1] + o->sum_v2[2] + st[0] + st[4] + st[8]);
    for (int q = 0; q < 9; ++q) o->stress[q] = st[q] * invV;
}

/* ------------------------------------------------- single force evaluation -- */

void orc_forces(const orc_params* p, int32_t n, const int32_t* id, const int32_t* type,
                const double* x, const double* unpbc, const double* v,
                const orc_topology* topo, int64_t step, int32_t thermostat,
                double* f, double* virial, double* pe)
{
    orc_params q = *p;
    q.rng_mode = ORC_RNG_PHILOX;
    orc_sys* s = orc_create(&q, n, id, type, x, unpbc, v, NULL, topo);
    s->step = step;
    double pes[3] = {0, 0, 0};
    pair_ctx ctx = {s->a, s->g, s->p.ntypes, s->invrootdt, ORC_RNG_PHILOX, &s->lcg, s->p.seed,
                    step, thermostat, &pes[0]};
    memset(s->f, 0, sizeof(double) * 3 * n);
    for (int c = 0; c < s->ncell; ++c) {
        for (int b1 = s->head[c]; b1 >= 0; b1 = s->next[b1]) {
            double* st = s->stress + 9 * b1;
            const double* x1 = s->x + 3 * b1;
            const double* v1 = s->v + 3 * b1;
            for (int b2 = s->tail[c]; b2 != b1; b2 = s->prev[b2]) {
                double dx[3], dv[3];
                for (int d = 0; d < 3; ++d) { dx[d] = x1[d] - s->x[3 * b2 + d]; dv[d] = v1[d] - s->v[3 * b2 + d]; }
                pair_force(&ctx, s->type[b1], s->type[b2], s->id[b1], s->id[b2], dx, dv, s->f + 3 * b1, s->f + 3 * b2, st);
            }
            for (int h = 0; h < 13; ++h) {
                int nc = s->nn27[27 * c + 13 + HALF[h][0] + 3 * HALF[h][1] + 9 * HALF[h][2]];
                int pbc = s->ext[c] && s->ext[nc];
                for (int b2 = s->head[nc]; b2 >= 0; b2 = s->next[b2]) {
                    double dx[3], dv[3];
                    for (int d = 0; d < 3; ++d) { dx[d] = x1[d] - s->x[3 * b2 + d]; dv[d] = v1[d] - s->v[3 * b2 + d]; }
                    if (pbc)
                        for (int d = 0; d < 3; ++d) {
                            if (dx[d] > s->halfL[d]) dx[d] -= s->L[d];
                            else if (dx[d] < -s->halfL[d]) dx[d] += s->L[d];
                        }
                    pair_force(&ctx, s->type[b1], s->type[b2], s->id[b1], s->id[b2], dx, dv, s->f + 3 * b1, s->f + 3 * b2, st);
                }
            }
        }
    }
    bond_forces(s, &pes[1]);
    angle_forces(s, &pes[2]);
    memcpy(f, s->f, sizeof(double) * 3 * n);
    if (virial) {
        for (int qq = 0; qq < 9; ++qq) virial[qq] = 0.0;
        for (int b = 0; b < n; ++b)
            for (int qq = 0; qq < 9; ++qq) virial[qq] += s->stress[9 * b + qq];
    }
    if (pe) { pe[0] = pes[0]; pe[1] = pes[1]; pe[2] = pes[2]; }
    orc_destroy(s);
}

/* The potential energy the reference writes with ToggleEnergyOutput
 * (CMonitor::ZeroTotalEnergy + CCNTCell::UpdateTotalEnergy, src/Monitor.cpp:2755-2778,
 * src/CNTCell.cpp:2909-3214): pairs of one cell once each (reverse loop to the bead
 * itself), but a pair across cells only from the half-shell side and only when the
 * neighbour's id is larger (src/CNTCell.cpp:3081), so about half of the cross-cell pairs
 * are left out; bonds 1/2 k (L - l0)^2 and bond pairs (src/BondPair.cpp:320-378) in
 * full.  out[0] pair, out[1] bond, out[2] bond-pair energy. */
void orc_energy_output(const orc_params* p, int32_t n, const int32_t* id, const int32_t* type,
                       const double* x, const double* unpbc, const orc_topology* topo, double* out)
{
    orc_params q = *p;
    q.rng_mode = ORC_RNG_PHILOX;
    orc_sys* s = orc_create(&q, n, id, type, x, unpbc, x, NULL, topo);
    const int nt = s->p.ntypes;
    double pair = 0.0;
    for (int c = 0; c < s->ncell; ++c) {
        for (int b1 = s->head[c]; b1 >= 0; b1 = s->next[b1]) {
            const double* x1 = s->x + 3 * b1;
            for (int b2 = s->tail[c]; b2 != b1; b2 = s->prev[b2]) {
                double dr2 = 0.0;
                for (int d = 0; d < 3; ++d) { double dd = x1[d] - s->x[3 * b2 + d]; dr2 += dd * dd; }
                if (dr2 < 1.0) {
                    double dr = sqrt(dr2);
                    if (dr > 0.000000001) pair += 0.5 * s->a[s->type[b1] * nt + s->type[b2]] * (1.0 - dr) * (1.0 - dr);
                }
            }
            for (int h = 0; h < 13; ++h) {
                int nc = s->nn27[27 * c + 13 + HALF[h][0] + 3 * HALF[h][1] + 9 * HALF[h][2]];
                int pbc = s->ext[c] && s->ext[nc];
                for (int b2 = s->head[nc]; b2 >= 0; b2 = s->next[b2]) {
                    if (!(s->id[b2] > s->id[b1])) continue;
                    double dx[3];
                    for (int d = 0; d < 3; ++d) dx[d] = x1[d] - s->x[3 * b2 + d];
                    if (pbc)
                        for (int d = 0; d < 3; ++d) {
                            if (dx[d] > s->halfL[d]) dx[d] -= s->L[d];
                            else if (dx[d] < -s->halfL[d]) dx[d] += s->L[d];
                        }
                    double dr2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
                    if (dr2 < 1.0) {
                        double dr = sqrt(dr2);
                        if (dr > 0.000000001) pair += 0.5 * s->a[s->type[b1] * nt + s->type[b2]] * (1.0 - dr) * (1.0 - dr);
                    }
                }
            }
        }
    }
    double pb = 0.0, pa = 0.0;
    bond_forces(s, &pb);
    angle_forces(s, &pa);
    out[0] = pair;
    out[1] = pb;
    out[2] = pa;
    orc_destroy(s);
}

void orc_pair_forces_brute(const orc_params* p, int32_t n, const int32_t* id,
                           const int32_t* type, const double* x, const double* v,
                           int64_t step, int32_t thermostat, double* f, double* virial)
{
    double L[3] = {p->nx * p->wx, p->ny * p->wy, p->nz * p->wz};
    double invrootdt = sqrt(24.0 * p->kT / p->dt);
    pair_ctx ctx = {p->a, p->g, p->ntypes, invrootdt, ORC_RNG_PHILOX, NULL, p->seed, step, thermostat, NULL};
    memset(f, 0, sizeof(double) * 3 * n);
    double st[9] = {0};
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double dx[3], dv[3];
            for (int d = 0; d < 3; ++d) {
                dx[d] = x[3 * i + d] - x[3 * j + d];
                dx[d] -= L[d] * floor(dx[d] / L[d] + 0.5);
                dv[d] = v[3 * i + d] - v[3 * j + d];
            }
            pair_force(&ctx, type[i], type[j], id[i], id[j], dx, dv, f + 3 * i, f + 3 * j, st);
        }
    if (virial) memcpy(virial, st, sizeof(st));
}

/* ------------------------------------------------------- slice stress -- */

/* CMonitor::AddBeadStress (src/Monitor.cpp:1900-2104), also the split of
 * CMonitor::AddBondStress: fdx = F (x) dx of one interaction is shared between the
 * z-slices spanned by the two beads (wrapped z coordinates z1 of the bead F acts on,
 * z2 of the other; dx = x1 - x2 with the periodic correction).  Same slice: all of it,
 * or half to each neighbour slice when both beads sit on the slice border; otherwise
 * slice_width / |dz| to every slice strictly between and the fractional parts to the
 * end slices, wrapping across the periodic boundary.  Adds wgt * share into
 * tab[(slice * nidx + idx) * 9 + component]. */
void orc_slice_add(double* tab, int32_t nslice, double Lz, int32_t idx, int32_t nidx, double z1, double z2,
                   const double F[3], const double dx[3], double wgt)
{
    const double width = Lz / nslice, inv = nslice / Lz;
    double fdx[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) fdx[3 * a + b] = F[a] * dx[b] * wgt;
    double zd = dx[2], zu, zl;
    if (zd > 0.00001) { zu = z1 * inv; zl = z2 * inv; }
    else if (zd < -0.00001) { zd = -zd; zu = z2 * inv; zl = z1 * inv; }
    else { zu = zl = z1 * inv; }
    long iu = (long)zu, il = (long)zl;
    const double ru = zu - (double)iu, rl = zl - (double)il;
#define SLICE_ADD(sl, share) \
    for (int q_ = 0; q_ < 9; ++q_) tab[((long)(sl) * nidx + idx) * 9 + q_] += fdx[q_] * (share)
    if (iu == il) {
        if (zu != zl || ru > 0.00001) {
            SLICE_ADD(il, 1.0);
        } else {
            SLICE_ADD(iu, 0.5);
            if (--il < 0) il += nslice;
            SLICE_ADD(il, 0.5);
        }
    } else {
        const double inner = width / zd;
        if (iu > il) {
            for (long sl = il + 1; sl <= iu - 1; ++sl) SLICE_ADD(sl, inner);
        } else {
            for (long sp = il + 1; sp <= nslice + iu - 1; ++sp) SLICE_ADD(sp >= nslice ? sp - nslice : sp, inner);
        }
        if (rl > 0.000001) SLICE_ADD(il, (1.0 - rl) * inner);
        else SLICE_ADD(il, inner);
        if (ru > 0.000001) SLICE_ADD(iu, ru * inner);
    }
#undef SLICE_ADD
}

void orc_slice_stress(const orc_params* p, int32_t n, const int32_t* id, const int32_t* type,
                      const double* x, const double* unpbc, const double* v, const orc_topology* topo,
                      int64_t step, int32_t thermostat, int32_t nslice, double* pair_tab, double* bond_tab)
{
    orc_params q = *p;
    q.rng_mode = ORC_RNG_PHILOX;
    orc_sys* s = orc_create(&q, n, id, type, x, unpbc, v, NULL, topo);
    const int nt = s->p.ntypes, nidx = nt * nt;
    const double Lz = s->L[2];
    pair_ctx ctx = {s->a, s->g, nt, s->invrootdt, ORC_RNG_PHILOX, &s->lcg, s->p.seed, step, thermostat, NULL};
    if (pair_tab) memset(pair_tab, 0, sizeof(double) * (size_t)nslice * nidx * 9);
    if (bond_tab) memset(bond_tab, 0, sizeof(double) * (size_t)nslice * nidx * 9);
    memset(s->f, 0, sizeof(double) * 3 * n);
    /* the reference's traversal: bead1 in the current cell, bead2 later in it or in
     * the 13-cell half shell (CCNTCell::UpdateForce) */
    for (int c = 0; c < s->ncell && pair_tab; ++c) {
        for (int b1 = s->head[c]; b1 >= 0; b1 = s->next[b1]) {
            const double* x1 = s->x + 3 * b1;
            const double* v1 = s->v + 3 * b1;
            for (int b2 = s->tail[c]; b2 != b1; b2 = s->prev[b2]) {
                double dx[3], dv[3], nf[3];
                for (int d = 0; d < 3; ++d) { dx[d] = x1[d] - s->x[3 * b2 + d]; dv[d] = v1[d] - s->v[3 * b2 + d]; }
                if (pair_force_out(&ctx, s->type[b1], s->type[b2], s->id[b1], s->id[b2], dx, dv, s->f + 3 * b1,
                                   s->f + 3 * b2, NULL, nf))
                    orc_slice_add(pair_tab, nslice, Lz, s->type[b1] * nt + s->type[b2], nidx, x1[2],
                                  s->x[3 * b2 + 2], nf, dx, 1.0);
            }
            for (int h = 0; h < 13; ++h) {
                int nc = s->nn27[27 * c + 13 + HALF[h][0] + 3 * HALF[h][1] + 9 * HALF[h][2]];
                int pbc = s->ext[c] && s->ext[nc];
                for (int b2 = s->head[nc]; b2 >= 0; b2 = s->next[b2]) {
                    double dx[3], dv[3], nf[3];
                    for (int d = 0; d < 3; ++d) { dx[d] = x1[d] - s->x[3 * b2 + d]; dv[d] = v1[d] - s->v[3 * b2 + d]; }
                    if (pbc)
                        for (int d = 0; d < 3; ++d) {
                            if (dx[d] > s->halfL[d]) dx[d] -= s->L[d];
                            else if (dx[d] < -s->halfL[d]) dx[d] += s->L[d];
                        }
                    if (pair_force_out(&ctx, s->type[b1], s->type[b2], s->id[b1], s->id[b2], dx, dv, s->f + 3 * b1,
                                       s->f + 3 * b2, NULL, nf))
                        orc_slice_add(pair_tab, nslice, Lz, s->type[b1] * nt + s->type[b2], nidx, x1[2],
                                      s->x[3 * b2 + 2], nf, dx, 1.0);
                }
            }
        }
    }
    /* bonds: head - tail (unwrapped) and the force on the head, keyed by the ordered
     * (head type, tail type) pair */
    for (int i = 0; i < s->nb && bond_tab; ++i) {
        const int h = s->bh[i], t = s->bt[i];
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = s->u[3 * h + k] - s->u[3 * t + k];
        const double L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double sc = s->bk[i] * (s->bl0[i] - L) / L;
        const double F[3] = {sc * d[0], sc * d[1], sc * d[2]};
        orc_slice_add(bond_tab, nslice, Lz, s->type[h] * nt + s->type[t], nidx, s->x[3 * h + 2], s->x[3 * t + 2],
                      F, d, 1.0);
    }
    orc_destroy(s);
}

/* CMonitor::AddBondPairStress (src/Monitor.cpp:2520-2739): every bond pair adds, for each
 * of its three bead pairs (first tail -> first head -> second head -> first tail),
 * F (x) dx with dx = u[j+1] - u[j] (unwrapped) and F = 0.33333333 (F[j+1] - F[j]) (the
 * bond pair's forces on the two beads) to the z-slices the pair spans, with the split of
 * orc_slice_add.  Keyed by the bead types (second head, first head, first tail) in place of
 * the reference's bond-pair type index, which a declaration of the three bead types fixes. */
void orc_slice_stress_bondpairs(const orc_params* p, int32_t n, const int32_t* id, const int32_t* type,
                                const double* x, const double* unpbc, const orc_topology* topo, int32_t nslice,
                                double* tab)
{
    double* v0 = (double*)calloc((size_t)3 * (n > 0 ? n : 1), sizeof(double));
    orc_sys* s = orc_create(p, n, id, type, x, unpbc, v0, NULL, topo);
    free(v0);
    const int nt = s->p.ntypes, nidx = nt * nt * nt;
    const double Lz = s->L[2];
    memset(tab, 0, sizeof(double) * (size_t)nslice * nidx * 9);
    memset(s->f, 0, sizeof(double) * 3 * n);
    bond_forces(s, NULL);   /* the bond vectors and lengths the bond pairs use */
    for (int i = 0; i < s->np; ++i) {
        const int b1 = s->pf[i], b2 = s->ps[i];
        double ft[3], fh[3], fm[3];
        angle_one(s->pk[i], s->pphi0[i], s->bdx + 3 * b1, s->blen[b1], s->bdx + 3 * b2, s->blen[b2], ft, fh, fm);
        const int bead[4] = {s->bt[b1], s->bh[b1], s->bh[b2], s->bt[b1]};
        const double* F[4] = {ft, fm, fh, ft};
        const int idx = (s->type[bead[2]] * nt + s->type[bead[1]]) * nt + s->type[bead[0]];
        for (int j = 0; j < 3; ++j) {
            double dx[3], fj[3];
            for (int d = 0; d < 3; ++d) {
                dx[d] = s->u[3 * bead[j + 1] + d] - s->u[3 * bead[j] + d];
                fj[d] = 0.33333333 * (F[j + 1][d] - F[j][d]);
            }
            orc_slice_add(tab, nslice, Lz, idx, nidx, s->x[3 * bead[j + 1] + 2], s->x[3 * bead[j] + 2], fj, dx, 1.0);
        }
    }
    orc_destroy(s);
}
