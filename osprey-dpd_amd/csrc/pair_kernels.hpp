// pair_kernels.hpp -- DPD pair-force kernels (CCNTCell::UpdateForce, src/CNTCell.cpp:877-1572).
//
// Two kernels over the same cell-sorted SoA layout (both produce the same forces):
//   DPD_PAIR_FULLSHELL  one thread per bead visits all 27 neighbour cells and keeps only
//                       its own force (no scatter, no atomics).  Simple and exact; the
//                       cross-check for the hit-list kernel and the fallback beyond 2^26
//                       beads.
//   DPD_PAIR_HITLIST    the full-shell walk in two passes: candidates are tested on 4-byte
//                       scan codes (cell offsets to 8 bits), a neighbour row's three cells
//                       as one range, partners that may be in range are listed in LDS,
//                       then evaluated exactly in one dense loop (below).
// Both accumulate each bead's force in one thread in a fixed order, so forces are
// deterministic for a given layout.
#pragma once
#include <hip/hip_runtime.h>

#include "dpd_engine.h"
#include "engine_common.hpp"

namespace dpd {

struct PairArgs {
    int n;                 // beads (all boxes)
    int nbox;
    int ngroups;           // scan groups of 64 cell keys (all boxes)
    const BoxDev* boxes;   // [nbox]
    const float4* pos;     // offset xyz, type | box << 8 (sorted)
    const float4* vp;      // predicted velocity xyz, id bits (sorted)
    const int* key;        // global cell key per slot
    const uint32_t* code;  // scan code per slot (scan_code), kCodePad slots of padding after n
    const int* cstart;     // [nkeys+1] first slot of each global cell key
    const float* table;    // per box: a | gamma | sqrt(gamma), kTypeStride each
    const int64_t* step;   // simulation time of this force evaluation (device)
    float4* frc;           // out: pair force per slot
    double* acc;           // out (OBS): per box [6..14] virial sum dx_b F_a, [15] pair PE
    const int4* hitblk;    // hit-list workgroups of a batch: (box, first slot, end slot)
    int nhitblk;
    // the epilogue of a step that another follows (hit-list kernels, engines without bonded
    // terms): UpdateMom of this step, UpdatePos of the next, the new cell key and
    // counting-sort rank
    const float4* vel;     // v + dt/2 F of the previous step
    int* nkey;             // new cell key per slot
    int* rank;             // rank inside the new cell
    int* cnt;              // beads per cell key (counting sort)
    int* gcnt;             // beads per scan group
    int* flags;            // [0] non-finite force or velocity, [1] moves of more than one cell
};

constexpr int kObsVir = 6;   // virial slots in the per-box accumulator
constexpr int kObsPEP = 15;

// Full-shell kernel: thread per bead.  Each pair is evaluated from both sides with the
// same operands (dx negated), so the two forces are exact negatives of each other;
// virial and PE are halved.
template <bool OBS, bool MULTI>
__global__ __launch_bounds__(256) void k_pair_fullshell(PairArgs A, BoxDev B0, int nblocks, int per_xcd)
{
    // XCD-aware: workgroup b runs on XCD b % 8 and takes block (b % 8) * per_xcd + b / 8
    const int blk = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (blk >= nblocks) return;
    const int i = blk * 256 + threadIdx.x;
    double vir[10];
    if (OBS)
        for (int q = 0; q < 10; ++q) vir[q] = 0.0;
    int box = -1;
    if (i < A.n) {
        const float4 pi = A.pos[i];
        const int wi = __float_as_int(pi.w);
        box = bead_box(wi);
        const BoxDev& B = MULTI ? A.boxes[box] : B0;
        const RngKey K = rng_key(B, *A.step);
        const float* tab = A.table + (MULTI ? box * kTableStride : 0);
        const float4 vi = A.vp[i];
        const int ti = bead_type(wi);
        const uint32_t idi = (uint32_t)__float_as_int(vi.w);
        int c[3];
        key_cell(A.key[i] - B.koff, B, c);
        float fx = 0.f, fy = 0.f, fz = 0.f;
        float sv[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int dz = -1; dz <= 1; ++dz) {
            int cz = c[2] + dz;
            cz += (cz < 0) ? B.n[2] : ((cz >= B.n[2]) ? -B.n[2] : 0);
            for (int dy = -1; dy <= 1; ++dy) {
                int cy = c[1] + dy;
                cy += (cy < 0) ? B.n[1] : ((cy >= B.n[1]) ? -B.n[1] : 0);
                for (int dx = -1; dx <= 1; ++dx) {
                    int cx = c[0] + dx;
                    cx += (cx < 0) ? B.n[0] : ((cx >= B.n[0]) ? -B.n[0] : 0);
                    const int k = B.koff + cell_key(cx, cy, cz, B);
                    const int j0 = A.cstart[k], j1 = A.cstart[k + 1];
                    const float sx = -(float)dx * B.w[0], sy = -(float)dy * B.w[1], sz = -(float)dz * B.w[2];
                    for (int j = j0; j < j1; ++j) {
                        if (j == i) continue;
                        const float4 pj = A.pos[j];
                        const float ddx = (pi.x - pj.x) + sx;
                        const float ddy = (pi.y - pj.y) + sy;
                        const float ddz = (pi.z - pj.z) + sz;
                        const float dr2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (dr2 < 1.0f && dr2 > 1e-18f) {
                            const float4 vj = A.vp[j];
                            const int tt = ti * kMaxTypes + bead_type(__float_as_int(pj.w));
                            const float dxdv = ddx * (vi.x - vj.x) + ddy * (vi.y - vj.y) + ddz * (vi.z - vj.z);
                            float pe;
                            const float s = pair_scalar(dr2, dxdv, tab[tt], tab[2 * kTypeStride + tt],
                                                        tab[kTypeStride + tt], idi, (uint32_t)__float_as_int(vj.w),
                                                        B.invrootdt, K, OBS ? &pe : nullptr);
                            const float Fx = s * ddx, Fy = s * ddy, Fz = s * ddz;
                            fx += Fx; fy += Fy; fz += Fz;
                            if (OBS) {
                                sv[0] += ddx * Fx; sv[1] += ddy * Fx; sv[2] += ddz * Fx;
                                sv[3] += ddx * Fy; sv[4] += ddy * Fy; sv[5] += ddz * Fy;
                                sv[6] += ddx * Fz; sv[7] += ddy * Fz; sv[8] += ddz * Fz;
                                sv[9] += pe;
                            }
                        }
                    }
                }
            }
        }
        A.frc[i] = make_float4(fx, fy, fz, 0.f);
        if (OBS)
            for (int q = 0; q < 10; ++q) vir[q] = 0.5 * (double)sv[q];
    }
    if (OBS) {
        const int box0 = MULTI ? bead_box(__float_as_int(A.pos[blk * 256].w)) : 0;
        accumulate_boxes<10>(vir, box, box0, kObsVir, A.acc);
    }
}

// Hit-list kernel: thread per bead in two passes.  The scan of the 27 neighbour cells
// appends every partner in range to a per-thread list in LDS (slot | neighbour-cell code
// << 26); the list is then evaluated in one dense loop.  In the one-pass kernel a
// wavefront runs the whole pair evaluation (random stream, force) whenever any of its
// lanes has a partner in range -- on nearly every candidate iteration, with ~8% of the
// lanes busy -- while here the evaluation loop runs max(partners per lane) ~20 times per
// wavefront.  Partners are evaluated in candidate order with the same expressions as
// k_pair_fullshell; the compiler contracts multiply-adds differently in the two code
// shapes, so the forces agree to ~1 ulp per term, not bitwise.  Slots must fit in 26 bits.
// A schedule of the hit-list kernel: threads per workgroup, list entries per thread (4 B of
// LDS each) and the flush threshold (after each row segment the wavefront evaluates its
// lists together once any of them holds this many; FLUSH = CAP never flushes).
// dpd_autotune_schedules picks among these by timing them.
constexpr int kHitThreads = 512;   // the default schedule's
constexpr int kHitCap = 24;
constexpr int kHitSlotBits = 26;
constexpr uint32_t kHitSlotMask = (1u << kHitSlotBits) - 1u;
// slots of padding after the last scan code: the scan reads candidates four (or two) at a
// time, one pass ahead (seven past the end at most)
constexpr int kCodePad = 8;
// 1e-18 < dr2 < 1 as one unsigned compare of the bits: dr2 >= 0, NaN fails
constexpr uint32_t kPairLo = 0x219392EFu + 1u;   // bits of 1e-18f, plus one
__device__ inline bool pair_in_range(float dr2) { return __float_as_uint(dr2) - kPairLo < 0x3F800000u - kPairLo; }
// a 32-bit pointer into LDS (a generic pointer is 64 bits)
typedef __attribute__((address_space(3))) uint32_t lds_u32;
// two scan codes from one 4-byte-aligned 8-byte load
typedef uint32_t code2_t __attribute__((ext_vector_type(2), aligned(4)));
// four scan codes from one 4-byte-aligned 16-byte load
typedef uint32_t code4_t __attribute__((ext_vector_type(4), aligned(4)));
// two 16-bit integers in one register (v_pk_sub_i16, v_dot2_i32_i16)
typedef short short2_t __attribute__((ext_vector_type(2)));

// The candidate test of cubic cells in integer code units (u = w / 256): dx and dy as one
// packed pair, dz apart, |d|^2 by one two-term dot product plus dz^2.  QXY holds the bead's
// x (plus the segment's first cell index times 256, mod 2^16) and y in code units, Qz its z:
// floors of its position seen from the row, as the codes are floors of the candidates'.  Each
// axis is then within one unit of the true difference, so every pair inside the cut-off
// passes |d| < (1 + m) / u with m = 1.01 |w| / 256 + 1e-5; the evaluation drops the rest.
__device__ inline int code_dist2(uint32_t cj, uint32_t QXY, int Qz)
{
    // bytes 0, 1 (x16 of the row segment, before the first cell's offset) and 2 (y) of the code
    const uint32_t p = __builtin_amdgcn_perm(0u, cj, 0x0c020100u);
    const short2_t d = __builtin_bit_cast(short2_t, QXY) - __builtin_bit_cast(short2_t, p);
    const int dz = Qz - (int)(cj >> 24);
    return __builtin_amdgcn_sdot2(d, d, __mul24(dz, dz), false);
}
// beads per workgroup of THREADS threads with SPLIT lanes per bead (21 per wavefront for 3,
// 7 for 9; lane 63 idle)
constexpr int hit_beads(int threads, int split) { return split == 1 ? threads : (threads / 64) * (63 / split); }

// The step's end and the next step's start for slot i, whose force (Fx, Fy, Fz) this launch
// computed: the arithmetic of k_advance (MODE kEndBegin).  Used for engines without bonded
// terms (a bonded bead's force is not complete here); a bead flagged bonded is left alone.
// Every lane of the wavefront calls it (the counting-sort ranks group the wavefront's keys);
// `act` marks the lanes that hold a finished bead.
__device__ inline void pair_epilogue(const PairArgs& A, const BoxDev& B, int i, bool act, float Fx, float Fy,
                                     float Fz)
{
    int k = 0;
    if (act) {
        const float4 p = A.pos[i];
        act = !(__float_as_int(p.w) & kBondedBit);
        if (act) {
            // UpdateMom (k_permute repeats the sum and stores the next half step)
            float4 v = A.vel[i];
            v.x = kick(v.x, Fx, B.hdt);
            v.y = kick(v.y, Fy, B.hdt);
            v.z = kick(v.z, Fz, B.hdt);
            if (!(fabsf(Fx) + fabsf(Fy) + fabsf(Fz) < 1e30f) || !(fabsf(v.x) + fabsf(v.y) + fabsf(v.z) < 1e30f))
                atomicAdd(A.flags + 0, 1);
            int c[3];
            key_cell(A.key[i] - B.koff, B, c);
            const float o[3] = {p.x, p.y, p.z};
            const int img[3] = {0, 0, 0};   // the image change is applied by k_permute
            const float d[3] = {drift(v.x, Fx, B), drift(v.y, Fy, B), drift(v.z, Fz, B)};
            Moved m;
            advance(o, c, img, d, B, m);
            k = B.koff + cell_key(m.c[0], m.c[1], m.c[2], B);
            A.nkey[i] = k;
            if (m.jump) atomicAdd(A.flags + 1, 1);
            if (m.bad) atomicAdd(A.flags + 0, 1);
        }
    }
    const int lane = threadIdx.x & 63;
    int r, size, leader;
    wave_groups(k, act, r, size, leader);
    int base = 0;
    if (act && lane == leader) base = atomicAdd(A.cnt + k, size);
    base = __shfl(base, leader, 64);
    if (act) A.rank[i] = base + r;
    wave_groups(k >> 6, act, r, size, leader);
    if (act && lane == leader) atomicAdd(A.gcnt + (k >> 6), size);
}

// body of the hit-list kernel for workgroup blockIdx.x (< 8 * per_xcd); also called by the
// kernel that runs the bonded terms in extra workgroups after the pair workgroups.  With
// several boxes each workgroup takes slots of one box only (A.hitblk[blk] = box, first
// slot, end slot), so the box, its constants and its table are uniform: scalar loads and
// the table in LDS, as for one box.
template <bool OBS, bool MULTI, int THREADS, int CAP, int FLUSH, int SPLIT, bool EPI>
__device__ __attribute__((always_inline)) inline void hitlist_body(const PairArgs& A, const BoxDev& B0, int nblocks,
                                                                   int per_xcd)
{
    static_assert(SPLIT == 1 || SPLIT == 3 || SPLIT == 9, "one, three or nine lanes per bead");
    // candidates per pass of the scan (one 16- or 8-byte load of scan codes)
    constexpr int PASS = SPLIT == 1 ? 4 : 2;
    static_assert(CAP > PASS, "room for a pass");
    static_assert(!(OBS && EPI), "the observing launch ends a sequence of steps");
    __shared__ uint32_t hits[CAP][THREADS];
    __shared__ float stab[kTableStride];
    const int blk = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (blk >= nblocks) return;   // whole workgroups
    int box = 0, i0 = blk * hit_beads(THREADS, SPLIT), i1 = A.n;
    if (MULTI) {
        const int4 hb = A.hitblk[blk];
        box = hb.x; i0 = hb.y; i1 = hb.z;
    }
    const BoxDev& B = MULTI ? A.boxes[box] : B0;
    // the box's interaction table in LDS, and the shift of each neighbour-cell code
    __shared__ float4 shtab[64];
    const float* gtab = A.table + (size_t)box * kTableStride;
    for (int q = threadIdx.x; q < kTableStride; q += THREADS) stab[q] = gtab[q];
    if (threadIdx.x < 64) {
        const int q = threadIdx.x;
        shtab[q] = make_float4((float)(1 - (q & 3)) * B.w[0], (float)(1 - ((q >> 2) & 3)) * B.w[1],
                               (float)(1 - (q >> 4)) * B.w[2], 0.f);
    }
    __syncthreads();
    const int t = threadIdx.x;
    // SPLIT 3: lanes 3g, 3g+1, 3g+2 of a wavefront share bead g (21 per wavefront, lane 63
    // idle) and scan the neighbour rows of z layers -1, 0, +1 each; SPLIT 9: lanes 9g .. 9g+8
    // share bead g (7 per wavefront) and scan one neighbour row each
    const int lane = t & 63;
    const int grp = SPLIT == 1 ? 0 : lane / SPLIT;
    const int part = SPLIT == 1 ? 0 : lane - SPLIT * grp;
    const int i = SPLIT == 1 ? i0 + t : i0 + (t >> 6) * (63 / SPLIT) + grp;
    const bool valid = i < i1 && (SPLIT == 1 || lane < 63);
    double vir[10];
    if (OBS)
        for (int q = 0; q < 10; ++q) vir[q] = 0.0;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (valid) {
        const float4 pi = A.pos[i];
        const int wi = __float_as_int(pi.w);
        const RngKey K = rng_key(B, *A.step);
        const float* tab = stab;
        const float4 vi = A.vp[i];
        const int ti = bead_type(wi);
        const uint32_t idi = (uint32_t)__float_as_int(vi.w);
        int c[3];
        key_cell(A.key[i] - B.koff, B, c);
        float sv[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        // this thread's list: column t of hits, entry e at hits[e][t]; wp is the next free one
        lds_u32* const lbase = (lds_u32*)&hits[0][t];
        const uint32_t laddr = (uint32_t)(size_t)lbase;   // its LDS byte address
        uint32_t wp = laddr;
        // one listed partner: slot | neighbour-cell code, its position and velocity
        auto eval = [&](uint32_t e, const float4& pj, const float4& vj) {
            const float4 sh = shtab[e >> kHitSlotBits];
            const float ddx = (pi.x + sh.x) - pj.x;
            const float ddy = (pi.y + sh.y) - pj.y;
            const float ddz = (pi.z + sh.z) - pj.z;
            const float dr2 = ddx * ddx + ddy * ddy + ddz * ddz;
            const int tt = ti * kMaxTypes + bead_type(__float_as_int(pj.w));
            const float dxdv = ddx * (vi.x - vj.x) + ddy * (vi.y - vj.y) + ddz * (vi.z - vj.z);
            float pe;
            float s = pair_scalar(dr2, dxdv, tab[tt], tab[2 * kTypeStride + tt], tab[kTypeStride + tt], idi,
                                  (uint32_t)__float_as_int(vj.w), B.invrootdt, K, OBS ? &pe : nullptr);
            // the scan lists a superset: the bead itself and partners just past the cut-off
            const bool in_range = pair_in_range(dr2);
            s = in_range ? s : 0.f;
            if (OBS) pe = in_range ? pe : 0.f;
            const float Fx = s * ddx, Fy = s * ddy, Fz = s * ddz;
            fx += Fx; fy += Fy; fz += Fz;
            if (OBS) {
                sv[0] += ddx * Fx; sv[1] += ddy * Fx; sv[2] += ddz * Fx;
                sv[3] += ddx * Fy; sv[4] += ddy * Fy; sv[5] += ddz * Fy;
                sv[6] += ddx * Fz; sv[7] += ddy * Fz; sv[8] += ddz * Fz;
                sv[9] += pe;
            }
        };
        // evaluate and clear this thread's list, two entries per pass (both loads in
        // flight, two independent random-stream chains), in list order
        auto drain = [&]() {
            const int nh = (int)(wp - laddr) / (4 * THREADS);
            int h = 0;
            for (; h + 1 < nh; h += 2) {
                const uint32_t ea = lbase[h * THREADS], eb = lbase[(h + 1) * THREADS];
                const int ja = (int)(ea & kHitSlotMask), jb = (int)(eb & kHitSlotMask);
                const float4 pa = A.pos[ja], va = A.vp[ja];
                const float4 pb = A.pos[jb], vb = A.vp[jb];
                eval(ea, pa, va);
                eval(eb, pb, vb);
            }
            if (h < nh) {
                const uint32_t ea = lbase[h * THREADS];
                const int ja = (int)(ea & kHitSlotMask);
                eval(ea, A.pos[ja], A.vp[ja]);
            }
            wp = laddr;
        };
        // candidate test on the scan codes in integer code units (code_dist2): each axis of
        // the tested difference is within one unit u_a = w_a / 256 of the true one, so the
        // distance is within m = 1.01 |w| / 256 + 1e-5 of it and, with the sum of squares taken
        // in units of the smallest u_a, every pair inside the cut-off passes
        // sum d_a^2 < ((1 + m) / u_min)^2 (exact for cubic cells, a wider superset for
        // others); the evaluation drops the rest (the bead itself among them).
        const float inv_ux = 256.0f / B.w[0], inv_uy = 256.0f / B.w[1], inv_uz = 256.0f / B.w[2];
        const float m = sqrtf(B.w[0] * B.w[0] + B.w[1] * B.w[1] + B.w[2] * B.w[2]) * (1.01f / 256.0f) + 1e-5f;
        const float tu = (1.0f + m) * fmaxf(inv_ux, fmaxf(inv_uy, inv_uz));
        const uint32_t thr = (uint32_t)(tu * tu) + 1u;
        const char* cbase = reinterpret_cast<const char*>(A.code);
        // a row wraps through the periodic x face only for beads in the first or last cell of
        // their row: wavefronts without such a bead scan one segment per row
        const int nseg = __ballot(c[0] == 0 || c[0] == B.n[0] - 1) != 0 ? 2 : 1;
        // the neighbour rows of this lane: all nine, one z layer (SPLIT 3) or one row (SPLIT 9)
        const int dz0 = SPLIT == 1 ? -1 : (SPLIT == 3 ? part - 1 : part / 3 - 1);
        const int dz1 = SPLIT == 1 ? 1 : dz0;
        const int dy0 = SPLIT == 9 ? part % 3 - 1 : -1, dy1 = SPLIT == 9 ? dy0 : 1;
        for (int dz = dz0; dz <= dz1; ++dz) {
            int cz = c[2] + dz;
            cz += (cz < 0) ? B.n[2] : ((cz >= B.n[2]) ? -B.n[2] : 0);
            for (int dy = dy0; dy <= dy1; ++dy) {
                int cy = c[1] + dy;
                cy += (cy < 0) ? B.n[1] : ((cy >= B.n[1]) ? -B.n[1] : 0);
                // the row's cells are consecutive keys: cells cx-1 .. cx+1 are one range, or two
                // where the row wraps through the periodic x face.  Segment s covers the cells
                // fx .. lx (wrapped indices), the first of them at fu = -1, 0 or +1 cells from
                // the bead's along x (unwrapped).  Candidates come in the order of a cell-by-cell
                // walk (x innermost).
                const int krow = B.koff + B.n[0] * (cy + B.n[1] * cz);
                const uint32_t qrow = (uint32_t)(((dy + 1) << 2) | ((dz + 1) << 4));
                // the bead's y and z seen from this row (the shifts of the evaluation), in code
                // units, floored as the codes are
                const int Qy = (int)floorf((pi.y - (float)dy * B.w[1]) * inv_uy);
                const int Qz = (int)floorf((pi.z - (float)dz * B.w[2]) * inv_uz);
                const int cx = c[0], nx = B.n[0];
                for (int sgm = 0; sgm < nseg; ++sgm) {
                    int fx, lx, fu;
                    if (cx >= 1 && cx <= nx - 2) { fx = cx - 1; lx = sgm ? cx - 2 : cx + 1; fu = -1; }
                    else if (cx == 0) { fx = sgm ? 0 : nx - 1; lx = sgm ? 1 : nx - 1; fu = sgm ? 0 : -1; }
                    else { fx = sgm ? 0 : nx - 2; lx = sgm ? 0 : nx - 1; fu = sgm ? 1 : -1; }
                    const int j0 = A.cstart[krow + fx], j1 = (lx >= fx) ? A.cstart[krow + lx + 1] : j0;
                    const uint32_t fxbits = (uint32_t)(fx & 0xFF) << 8;
                    const uint32_t code = ((uint32_t)(fu + 1) | qrow) << kHitSlotBits;
                    // the bead's x in the segment (code units) plus fxbits, mod 2^16: code bytes
                    // 0-1 less fxbits are r * 256 + q_x, r the candidate's cell counted from the
                    // segment's first one, so the packed difference is x less that
                    const int Qx = (int)floorf((pi.x - (float)fu * B.w[0]) * inv_ux) + (int)fxbits;
                    const uint32_t QXY = ((uint32_t)Qx & 0xFFFFu) | ((uint32_t)Qy << 16);
                    // list entry of slot j: (j | code) + (r << 26), the neighbour-cell code of
                    // the candidate's cell; r = (code byte 1 - fx) mod 64 = ((bits 8-13 of the
                    // code) - fx) mod 64, and e0 - j0 holds the parts that do not depend on the
                    // candidate
                    const uint32_t e0 = ((uint32_t)j0 | code) - ((uint32_t)(fx & 0xFF) << kHitSlotBits);
                    // PASS consecutive candidates per pass from one 16-byte (4) or 8-byte (2)
                    // load, the next pass's loaded before these are tested.  (A wavefront
                    // runs each segment to its lanes' longest count rounded up to the pass
                    // width.)  The loads run up to seven codes past the segment (the next
                    // cells' beads, or the padding after the last slot); none of those is
                    // listed.  Slots are below 2^26, so byte offsets fit 32 bits and the loads
                    // address from the array base.  Every candidate is stored at the list's
                    // write pointer, which advances on a hit only (a miss is overwritten by
                    // the next store).
                    uint32_t e = e0;
                    // one candidate: stored at the write pointer, which advances on a hit
                    auto put = [&](uint32_t cj, uint32_t ej, bool in_seg) {
                        const bool hit = ((uint32_t)code_dist2(cj, QXY, Qz) < thr) & in_seg;
                        *(lds_u32*)(size_t)wp = ej + (((cj >> 8) & 63u) << kHitSlotBits);
                        wp += hit ? 4 * THREADS : 0;
                    };
                    if constexpr (PASS == 4) {
                        code4_t c4 = *reinterpret_cast<const code4_t*>(cbase + ((uint32_t)j0 << 2));
                        for (int j = j0; j < j1; j += 4, e += 4) {
                            const code4_t cn = *reinterpret_cast<const code4_t*>(cbase + ((uint32_t)j << 2) + 16);
                            // room for the pass: a list past CAP - 4 entries is evaluated first
                            // (overflow inside a segment: this lane alone)
                            if (wp > laddr + 4 * (CAP - 4) * THREADS) drain();
                            put(c4.x, e, true);
                            put(c4.y, e + 1u, j + 1 < j1);
                            put(c4.z, e + 2u, j + 2 < j1);
                            put(c4.w, e + 3u, j + 3 < j1);
                            c4 = cn;
                        }
                    } else {
                        code2_t c2 = *reinterpret_cast<const code2_t*>(cbase + ((uint32_t)j0 << 2));
                        for (int j = j0; j < j1; j += 2, e += 2) {
                            const code2_t cn = *reinterpret_cast<const code2_t*>(cbase + ((uint32_t)j << 2) + 8);
                            if (wp > laddr + 4 * (CAP - 2) * THREADS) drain();
                            put(c2.x, e, true);
                            put(c2.y, e + 1u, j + 1 < j1);
                            c2 = cn;
                        }
                    }
                    if (FLUSH < CAP && __ballot(wp >= laddr + 4 * FLUSH * THREADS) != 0) drain();
                }
            }
        }
        drain();
        if (SPLIT == 1) A.frc[i] = make_float4(fx, fy, fz, 0.f);
        if (OBS)
            for (int q = 0; q < 10; ++q) vir[q] = 0.5 * (double)sv[q];
    }
    if (SPLIT > 1) {
        // the bead's force: its lanes' parts (z layers, or rows in z-major order) added in
        // lane order (the whole wavefront takes part in the shuffles)
        const int b = SPLIT * grp;
        float sx = __shfl(fx, b), sy = __shfl(fy, b), sz = __shfl(fz, b);
#pragma unroll
        for (int q = 1; q < SPLIT; ++q) {
            sx += __shfl(fx, b + q);
            sy += __shfl(fy, b + q);
            sz += __shfl(fz, b + q);
        }
        fx = sx;
        fy = sy;
        fz = sz;
        if (valid && part == 0) A.frc[i] = make_float4(fx, fy, fz, 0.f);
    }
    if (EPI) pair_epilogue(A, B, i, valid && part == 0, fx, fy, fz);
    if (OBS) {
        accumulate_boxes<10, THREADS / 64>(vir, valid ? box : -1, box, kObsVir, A.acc);
    }
}

// Six wavefronts per SIMD (three 512-thread or six 256-thread workgroups per CU, which the
// LDS of the 24-entry lists also allows) need at most 80 VGPRs: the force-only kernel is held
// to that; the kernel with observables runs once per batch of steps and is left freer.
// wavefronts per SIMD that the LDS of a schedule allows (160 KB per CU, four SIMDs), at most 6
constexpr int hit_waves(int threads, int cap)
{
    const int lds = cap * threads * 4 + kTableStride * 4 + 64 * 16;
    const int w = (160 * 1024 / lds) * (threads / 64) / 4;
    return w < 6 ? w : 6;
}
// the occupancy a kernel is held to: the split schedules run small boxes, which never fill
// the SIMDs, and get the registers of four wavefronts (no spills)
constexpr int hit_waves_eu(bool obs, int threads, int cap, int split)
{
    return (obs || split > 1) ? 4 : hit_waves(threads, cap);
}

template <bool OBS, bool MULTI, int THREADS, int CAP, int FLUSH, int SPLIT, bool EPI>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(hit_waves_eu(OBS, THREADS, CAP, SPLIT))))
void k_pair_hitlist(PairArgs A, BoxDev B0, int nblocks, int per_xcd)
{
    hitlist_body<OBS, MULTI, THREADS, CAP, FLUSH, SPLIT, EPI>(A, B0, nblocks, per_xcd);
}

// The schedules of the pair launch: hit-list kernels with their parameters, and the
// full-shell kernel.  A schedule id is what the autotuner stores.
struct Schedule {
    int kernel;    // DPD_PAIR_HITLIST or DPD_PAIR_FULLSHELL
    int threads, cap, flush;
    int split;     // lanes per bead (hit list): 1, 3 (one z layer of neighbour rows each) or 9 (one row)
    const char* name;
    constexpr int beads() const { return hit_beads(threads, split); }   // per workgroup
};
constexpr int kSchedules = 7;
constexpr Schedule kSchedule[kSchedules] = {
    {DPD_PAIR_HITLIST, 512, 24, 24, 1, "hitlist-t512-c24"},
    {DPD_PAIR_HITLIST, 512, 24, 16, 1, "hitlist-t512-c24-f16"},
    {DPD_PAIR_HITLIST, 256, 24, 24, 1, "hitlist-t256-c24"},
    {DPD_PAIR_HITLIST, 256, 32, 32, 1, "hitlist-t256-c32"},
    {DPD_PAIR_FULLSHELL, 256, 0, 0, 1, "fullshell-t256"},
    {DPD_PAIR_HITLIST, 256, 12, 12, 3, "hitlist-split3-t256-c12"},
    {DPD_PAIR_HITLIST, 256, 8, 8, 9, "hitlist-split9-t256-c8"},
};
constexpr int kDefaultSchedule = 0;
constexpr int kFullShellSchedule = 4;

// hit-list launch of schedule S on a grid of workgroups of its thread count
template <int S, bool OBS, bool MULTI, bool EPI>
inline void launch_hitlist_s(const PairArgs& a, const BoxDev& B0, int nblocks, hipStream_t s)
{
    constexpr Schedule q = kSchedule[S];
    const int ph = (nblocks + 7) / 8;
    hipLaunchKernelGGL((k_pair_hitlist<OBS, MULTI, q.threads, q.cap, q.flush, q.split, EPI>), dim3(8 * ph),
                       dim3(q.threads), 0, s, a, B0, nblocks, ph);
}

template <bool OBS, bool MULTI, bool EPI>
inline void launch_hitlist(int sched, const PairArgs& a, const BoxDev& B0, int nblocks, hipStream_t s)
{
    switch (sched) {
    case 1: launch_hitlist_s<1, OBS, MULTI, EPI>(a, B0, nblocks, s); break;
    case 2: launch_hitlist_s<2, OBS, MULTI, EPI>(a, B0, nblocks, s); break;
    case 3: launch_hitlist_s<3, OBS, MULTI, EPI>(a, B0, nblocks, s); break;
    case 5: launch_hitlist_s<5, OBS, MULTI, EPI>(a, B0, nblocks, s); break;
    case 6: launch_hitlist_s<6, OBS, MULTI, EPI>(a, B0, nblocks, s); break;
    default: launch_hitlist_s<0, OBS, MULTI, EPI>(a, B0, nblocks, s); break;
    }
}

// whether schedule `sched` can carry the step's epilogue (hit-list kernels only)
inline bool epilogue_available(int sched, int64_t n)
{
    return sched >= 0 && sched < kSchedules && kSchedule[sched].kernel == DPD_PAIR_HITLIST &&
           n <= (int64_t)kHitSlotMask;
}

// The automatic choice: the hit-list kernel for every engine whose slots fit its 26-bit list
// entries, the full-shell kernel beyond.
inline int default_pair_kernel(int64_t nbeads)
{
    return nbeads <= (int64_t)kHitSlotMask ? DPD_PAIR_HITLIST : DPD_PAIR_FULLSHELL;
}

inline bool pair_kernel_available(int which) { return which == DPD_PAIR_FULLSHELL || which == DPD_PAIR_HITLIST; }

// The pair launch of schedule `sched`.  A batch's hit-list workgroups take the slots of one
// box each: hitblk (nhitblk entries) must be the table built for the schedule's thread count.
// `epi`: the launch also carries the step's epilogue (epilogue_available(sched) must hold).
inline hipError_t launch_pair_kernel(int sched, const PairArgs& a, const BoxDev& B0, bool obs, hipStream_t s,
                                     bool epi = false)
{
    const bool multi = a.nbox > 1;
    if (sched < 0 || sched >= kSchedules) sched = kDefaultSchedule;
    if (kSchedule[sched].kernel == DPD_PAIR_HITLIST && a.n <= (int)kHitSlotMask && (!multi || a.hitblk)) {
        const int nb = kSchedule[sched].beads();
        const int gh = multi ? a.nhitblk : (a.n + nb - 1) / nb;
        if (obs) {
            if (multi) launch_hitlist<true, true, false>(sched, a, B0, gh, s);
            else launch_hitlist<true, false, false>(sched, a, B0, gh, s);
        } else if (epi) {
            if (multi) launch_hitlist<false, true, true>(sched, a, B0, gh, s);
            else launch_hitlist<false, false, true>(sched, a, B0, gh, s);
        } else {
            if (multi) launch_hitlist<false, true, false>(sched, a, B0, gh, s);
            else launch_hitlist<false, false, false>(sched, a, B0, gh, s);
        }
        return hipGetLastError();
    }
    if (epi) return hipErrorInvalidValue;   // the full-shell kernel carries no epilogue
    // the full-shell kernel (also for hit lists beyond 2^26 beads)
    const int g = (a.n + 255) / 256;
    const int per_xcd = (g + 7) / 8;
    const dim3 grid(8 * per_xcd);
    if (obs) {
        if (multi) hipLaunchKernelGGL((k_pair_fullshell<true, true>), grid, dim3(256), 0, s, a, B0, g, per_xcd);
        else hipLaunchKernelGGL((k_pair_fullshell<true, false>), grid, dim3(256), 0, s, a, B0, g, per_xcd);
    } else {
        if (multi) hipLaunchKernelGGL((k_pair_fullshell<false, true>), grid, dim3(256), 0, s, a, B0, g, per_xcd);
        else hipLaunchKernelGGL((k_pair_fullshell<false, false>), grid, dim3(256), 0, s, a, B0, g, per_xcd);
    }
    return hipGetLastError();
}

}  // namespace dpd
