// pair_kernels.hpp -- DPD pair-force kernels (CCNTCell::UpdateForce, src/CNTCell.cpp:877-1572).
//
// Pair kernels over the same cell-sorted SoA layout (all produce the same forces):
//   DPD_PAIR_FULLSHELL  one thread per bead visits all 27 neighbour cells and keeps only
//                       its own force (no scatter, no atomics).  Simple and exact; the
//                       cross-check for the brick kernel.
//   DPD_PAIR_BRICK      one workgroup per 4x4x4-cell brick: the brick and its one-cell
//                       halo are staged in LDS, candidate pairs are filtered by r < 1 and
//                       compacted per wavefront, and the surviving pairs are evaluated
//                       64 at a time with both forces accumulated in LDS (pair_brick.hip).
//   DPD_PAIR_STREAM     the brick layout with a flat (bead, candidate) stream per home
//                       cell, so the candidate test runs on full wavefronts (pair_stream.hip).
//   DPD_PAIR_COMPACT    the full-shell walk with pairs in range compacted per wavefront
//                       and evaluated 64 at a time (pair_compact.hip).
//   DPD_PAIR_CELLS      a wavefront per home cell: the 27-cell candidate list is loaded once
//                       per cell and tested against each of its beads (pair_cells.hip).
//   DPD_PAIR_HITLIST    the full-shell walk in two passes: candidates are tested on 4-byte
//                       scan codes (cell offsets to 8 bits), partners that may be in range
//                       are listed in LDS, then evaluated exactly in one dense loop (below).
#pragma once
#include <hip/hip_runtime.h>

#include "dpd_engine.h"
#include "engine_common.hpp"

namespace dpd {

struct PairArgs {
    int n;                 // beads (all boxes)
    int nbox;
    int nbricks;           // bricks (all boxes)
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
constexpr int kHitThreads = 512;
constexpr int kHitCap = 24;     // list entries per thread (96 B of LDS each: 48 KB per workgroup)
constexpr int kHitFlush = 24;   // after each neighbour cell, the wavefront evaluates its
                                // lists together once any of them holds this many
constexpr int kHitSlotBits = 26;
constexpr uint32_t kHitSlotMask = (1u << kHitSlotBits) - 1u;
// slots of padding after the last scan code: the scan reads candidates two at a time,
// one pass ahead (three past the end at most)
constexpr int kCodePad = 8;
// 1e-18 < dr2 < 1 as one unsigned compare of the bits: dr2 >= 0, NaN fails
constexpr uint32_t kPairLo = 0x219392EFu + 1u;   // bits of 1e-18f, plus one
__device__ inline bool pair_in_range(float dr2) { return __float_as_uint(dr2) - kPairLo < 0x3F800000u - kPairLo; }
// a 32-bit pointer into LDS (a generic pointer is 64 bits)
typedef __attribute__((address_space(3))) uint32_t lds_u32;
// two scan codes from one 4-byte-aligned 8-byte load
typedef uint32_t code2_t __attribute__((ext_vector_type(2), aligned(4)));

// body of the hit-list kernel for workgroup blockIdx.x (< 8 * per_xcd); also called by the
// kernel that runs the bonded terms in extra workgroups after the pair workgroups.  With
// several boxes each workgroup takes slots of one box only (A.hitblk[blk] = box, first
// slot, end slot), so the box, its constants and its table are uniform: scalar loads and
// the table in LDS, as for one box.
template <bool OBS, bool MULTI>
__device__ __attribute__((always_inline)) inline void hitlist_body(const PairArgs& A, const BoxDev& B0, int nblocks,
                                                                   int per_xcd)
{
    __shared__ uint32_t hits[kHitCap][kHitThreads];
    __shared__ float stab[kTableStride];
    const int blk = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (blk >= nblocks) return;   // whole workgroups
    int box = 0, i0 = blk * kHitThreads, i1 = A.n;
    if (MULTI) {
        const int4 hb = A.hitblk[blk];
        box = hb.x; i0 = hb.y; i1 = hb.z;
    }
    const BoxDev& B = MULTI ? A.boxes[box] : B0;
    // the box's interaction table in LDS, and the shift of each neighbour-cell code
    __shared__ float4 shtab[64];
    const float* gtab = A.table + (size_t)box * kTableStride;
    for (int q = threadIdx.x; q < kTableStride; q += kHitThreads) stab[q] = gtab[q];
    if (threadIdx.x < 64) {
        const int q = threadIdx.x;
        shtab[q] = make_float4((float)(1 - (q & 3)) * B.w[0], (float)(1 - ((q >> 2) & 3)) * B.w[1],
                               (float)(1 - (q >> 4)) * B.w[2], 0.f);
    }
    __syncthreads();
    const int t = threadIdx.x;
    const int i = i0 + t;
    const bool valid = i < i1;
    double vir[10];
    if (OBS)
        for (int q = 0; q < 10; ++q) vir[q] = 0.0;
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
        float fx = 0.f, fy = 0.f, fz = 0.f;
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
            const int nh = (int)(wp - laddr) / (4 * kHitThreads);
            int h = 0;
            for (; h + 1 < nh; h += 2) {
                const uint32_t ea = lbase[h * kHitThreads], eb = lbase[(h + 1) * kHitThreads];
                const int ja = (int)(ea & kHitSlotMask), jb = (int)(eb & kHitSlotMask);
                const float4 pa = A.pos[ja], va = A.vp[ja];
                const float4 pb = A.pos[jb], vb = A.vp[jb];
                eval(ea, pa, va);
                eval(eb, pb, vb);
            }
            if (h < nh) {
                const uint32_t ea = lbase[h * kHitThreads];
                const int ja = (int)(ea & kHitSlotMask);
                eval(ea, A.pos[ja], A.vp[ja]);
            }
            wp = laddr;
        };
        // candidate test on the scan codes.  A decoded offset is within w/512 of the true
        // one on each axis, so the distance is within m of the true distance and every pair
        // inside the cut-off passes |dr| < 1 + m; the evaluation drops the rest (the bead
        // itself among them).  Test in code units: offset = (q + 1/2) u, u = w / 256.
        const float ux = B.w[0] * (1.0f / 256.0f), uy = B.w[1] * (1.0f / 256.0f), uz = B.w[2] * (1.0f / 256.0f);
        const float m = sqrtf(B.w[0] * B.w[0] + B.w[1] * B.w[1] + B.w[2] * B.w[2]) * (1.01f / 512.0f) + 1e-5f;
        const uint32_t thr = __float_as_uint((1.0f + m) * (1.0f + m));
        auto test = [&](uint32_t jc, uint32_t cj, float qx, float qy, float qz, bool in_cell) {
            const float ddx = qx - (float)(cj & 0xFFu) * ux;
            const float ddy = qy - (float)((cj >> 8) & 0xFFu) * uy;
            const float ddz = qz - (float)((cj >> 16) & 0xFFu) * uz;
            const float dr2 = ddx * ddx + ddy * ddy + ddz * ddz;
            if ((__float_as_uint(dr2) < thr) & in_cell) {
                *(lds_u32*)(size_t)wp = jc;
                wp += 4 * kHitThreads;
            }
        };
        const char* cbase = reinterpret_cast<const char*>(A.code);
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
                    // the bead's position seen from this neighbour cell (same shift table as
                    // the evaluation), less half a code unit
                    const uint32_t q = (uint32_t)((dx + 1) | ((dy + 1) << 2) | ((dz + 1) << 4));
                    const uint32_t code = q << kHitSlotBits;
                    const float4 sh = shtab[q];
                    const float qx = (pi.x + sh.x) - 0.5f * ux, qy = (pi.y + sh.y) - 0.5f * uy,
                                qz = (pi.z + sh.z) - 0.5f * uz;
                    // two consecutive candidates per pass from one 8-byte load, the next
                    // pass's two loaded before these are tested.  (Four per pass waste more
                    // slots: a wavefront runs each cell to its lanes' longest count rounded
                    // up to the pass width, 223 candidate slots per lane against 194 for two
                    // and 82 candidates; one per pass loses the overlap of loads.)  The loads
                    // run up to three codes past the cell (the next cells' beads, or the
                    // padding after the last slot); none of those is listed.  Slots are below
                    // 2^26, so byte offsets fit 32 bits and the loads address from the array
                    // base.
                    code2_t c2 = *reinterpret_cast<const code2_t*>(cbase + ((uint32_t)j0 << 2));
                    for (int j = j0; j < j1; j += 2) {
                        const code2_t cn = *reinterpret_cast<const code2_t*>(cbase + ((uint32_t)j << 2) + 8);
                        const uint32_t jc = (uint32_t)j | code;
                        // room for the two: a list past kHitCap - 2 entries is evaluated
                        // first (overflow inside a cell: this lane alone)
                        if (wp > laddr + 4 * (kHitCap - 2) * kHitThreads) drain();
                        test(jc, c2.x, qx, qy, qz, true);
                        test(jc + 1, c2.y, qx, qy, qz, j + 1 < j1);
                        c2 = cn;
                    }
                    if (kHitFlush < kHitCap && __ballot(wp >= laddr + 4 * kHitFlush * kHitThreads) != 0) drain();
                }
            }
        }
        drain();
        A.frc[i] = make_float4(fx, fy, fz, 0.f);
        if (OBS)
            for (int q = 0; q < 10; ++q) vir[q] = 0.5 * (double)sv[q];
    }
    if (OBS) {
        accumulate_boxes<10, kHitThreads / 64>(vir, valid ? box : -1, box, kObsVir, A.acc);
    }
}

// Six wavefronts per SIMD (three 512-thread workgroups per CU, which the 52 KB of LDS
// also allows) need at most 80 VGPRs: the force-only kernel is held to that; the kernel
// with observables runs once per batch of steps and is left unconstrained.
template <bool OBS, bool MULTI>
__global__ __launch_bounds__(kHitThreads) __attribute__((amdgpu_waves_per_eu(OBS ? 4 : 6)))
void k_pair_hitlist(PairArgs A, BoxDev B0, int nblocks, int per_xcd)
{
    hitlist_body<OBS, MULTI>(A, B0, nblocks, per_xcd);
}

// The automatic choice by the number of beads the engine holds (all boxes).  A thread
// per bead leaves most of the 256 CUs idle for small boxes (24k beads are 375 wavefronts),
// where the LDS stream kernel with 1024-thread workgroups is faster (1.25x at 24k, 1.25x
// at 41k beads); the two break even at ~57k beads, and above that the hit-list kernel is
// fastest (1.11x at 66k, 1.13x at 98k beads; 192k-bead box 1.28x, 786k-bead box 1.48x,
// 4M-bead batch of 256 replicas 1.85x the next best kernel before its scan took codes and
// two candidates per pass; measured on MI355X with graph replay).
constexpr int64_t kHitListMinBeads = 56000;

inline int default_pair_kernel(int64_t nbeads)
{
    if (nbeads < kHitListMinBeads) return DPD_PAIR_STREAM1024;
    return nbeads <= (int64_t)kHitSlotMask ? DPD_PAIR_HITLIST : DPD_PAIR_FULLSHELL;
}

hipError_t launch_pair_brick(const PairArgs& a, bool obs, hipStream_t s, int threads);
hipError_t launch_pair_stream(const PairArgs& a, bool obs, hipStream_t s, int threads);
hipError_t launch_pair_compact(const PairArgs& a, const BoxDev& B0, bool obs, hipStream_t s);
hipError_t launch_pair_cells(const PairArgs& a, bool obs, hipStream_t s);

inline hipError_t launch_pair_kernel(int which, const PairArgs& a, const BoxDev& B0, bool obs, hipStream_t s)
{
    if (which == DPD_PAIR_BRICK) return launch_pair_brick(a, obs, s, 256);
    if (which == DPD_PAIR_BRICK512) return launch_pair_brick(a, obs, s, 512);
    if (which == DPD_PAIR_STREAM) return launch_pair_stream(a, obs, s, 512);
    if (which == DPD_PAIR_STREAM1024) return launch_pair_stream(a, obs, s, 1024);
    if (which == DPD_PAIR_COMPACT) return launch_pair_compact(a, B0, obs, s);
    if (which == DPD_PAIR_CELLS) return launch_pair_cells(a, obs, s);
    const int g = (a.n + 255) / 256;
    const int per_xcd = (g + 7) / 8;
    const dim3 grid(8 * per_xcd);
    const bool multi = a.nbox > 1;
    if (which == DPD_PAIR_HITLIST && a.n <= (int)kHitSlotMask && (!multi || a.hitblk)) {
        const int gh = multi ? a.nhitblk : (a.n + kHitThreads - 1) / kHitThreads;
        const int ph = (gh + 7) / 8;
        const dim3 gr(8 * ph), bl(kHitThreads);
        if (obs) {
            if (multi) hipLaunchKernelGGL((k_pair_hitlist<true, true>), gr, bl, 0, s, a, B0, gh, ph);
            else hipLaunchKernelGGL((k_pair_hitlist<true, false>), gr, bl, 0, s, a, B0, gh, ph);
        } else {
            if (multi) hipLaunchKernelGGL((k_pair_hitlist<false, true>), gr, bl, 0, s, a, B0, gh, ph);
            else hipLaunchKernelGGL((k_pair_hitlist<false, false>), gr, bl, 0, s, a, B0, gh, ph);
        }
        return hipGetLastError();
    }
    // the full-shell kernel (also for hit lists beyond 2^26 beads)
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
