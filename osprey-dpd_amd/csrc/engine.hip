// engine.hip -- the DPD step engine for MI355X (gfx950): kernels and the C ABI of
// include/dpd_engine.h.  One engine holds one box or a batch of independent boxes
// (replicas of uneven size and parameters); every kernel spans all boxes at once.
//
// One step (the reference's CSimBox::Evolve, src/SimBox.cpp:977-1120) runs as:
//   k_advance   UpdateMom of the previous step fused with UpdatePos of this one
//               (src/CNTCell.cpp:2280-2310, :1704-1757); computes each bead's new cell
//               key and its rank in that cell
//   cell starts  counting-sort prefix sums, one launch up to 8192 groups of 64 keys
//               (k_cell_scan_small, k_cell_start_direct), else k_group_scan + k_cell_start
//   k_place / k_permute   counting-sort placement into the new cell-sorted SoA buffers,
//               beads of a cell in ascending bead index (deterministic layout)
//   pair kernel DPD pair forces (src/CNTCell.cpp:1144-1555)
//   k_bonded / k_bonded_sum   bonds and angles (src/Bond.cpp:111-162,
//               src/BondPair.cpp:138-286), each once, summed into the bonded beads' forces
// and a last k_finish closes the final step (UpdateMom + FP64 observables).
// The simulation time lives in device memory (bumped by k_advance), so a captured
// sequence of steps replays as a hipGraph without re-capture.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dpd_engine.h"
#include "engine_common.hpp"
#include "pair_kernels.hpp"
#include "slice_stress.hpp"

using namespace dpd;

namespace {

enum AdvanceMode { kRebin = 0, kBegin = 1, kEndBegin = 2 };

// FP64 accumulator slots per box
enum ObsSlot { O_V2 = 0, O_MOM = 3, O_VIR = 6, O_PEP = 15, O_PEB = 16, O_PEA = 17, O_N = 18 };
static_assert(O_VIR == kObsVir && O_PEP == kObsPEP, "accumulator layout");

struct Topo {
    const int* bond_ptr;      // [n+1] bond entries of each bead index (all boxes)
    const int2* bond_csr;     // per bond: the entries of its head and of its tail
    const int2* bond_bead;    // (head, tail) bead index per bond
    const double2* bond_par;  // (k, l0) per bond
    const int* ang_ptr;       // [n+1] angle entries of each bead index
    const int4* ang_csr;      // per angle: the entries of roles 0 (c), 1 (b), 2 (a)
    const int4* ang_bead;     // (a, b, c, 0) per angle
    const double4* ang_par;   // (k, phi0, cos phi0, sin phi0)
    double4* fbond;           // per bond entry: the bond's force on that entry's bead
    double4* fang;            // per angle entry: the angle's force on that entry's bead
    const int* bonded;        // bead indices with at least one bond or angle
    int nbonds, nangles, nbonded;
    int has_bonds, has_angles;
};

struct Buffers {
    float4* pos;   // offset xyz, type | box << 8 | has bonds << 30
    float4* vel;   // velocity xyz, id bits: v at the state's time, or v + dt/2 F inside a sequence
    float4* fold;  // the uploaded force (re-binning only)
    int* bead;     // bead index (upload order) of each slot
    int4* img;     // periodic image counts xyz by bead index (not sorted; one buffer)
    int* key;      // global cell key
    uint32_t* code;   // scan code (scan_code) of the offset
};

template <bool MULTI>
__device__ inline const BoxDev& box_of(const BoxDev* boxes, const BoxDev& B0, int w)
{
    return MULTI ? boxes[bead_box(w)] : B0;
}

__device__ inline void unwrapped_pos(int slot, int bead, const Buffers& S, const BoxDev& B, double u[3])
{
    int c[3];
    key_cell(S.key[slot] - B.koff, B, c);
    const float4 p = S.pos[slot];
    const int4 a = S.img[bead];
    u[0] = (double)(c[0] + a.x * B.n[0]) * B.wd[0] + (double)p.x;
    u[1] = (double)(c[1] + a.y * B.n[1]) * B.wd[1] + (double)p.y;
    u[2] = (double)(c[2] + a.z * B.n[2]) * B.wd[2] + (double)p.z;
}

// Sum of the bonded forces on bead index q, gathered from the per-bond and per-angle
// forces of this step (each bond and angle is evaluated once, by k_bonded).
__device__ inline void bonded_gather(int q, const Topo& T, double f[3])
{
    f[0] = f[1] = f[2] = 0.0;
    if (T.has_bonds) {
        for (int e = T.bond_ptr[q], e1 = T.bond_ptr[q + 1]; e < e1; ++e) {
            const double4 fb = T.fbond[e];
            f[0] += fb.x; f[1] += fb.y; f[2] += fb.z;
        }
    }
    if (T.has_angles) {
        for (int e = T.ang_ptr[q], e1 = T.ang_ptr[q + 1]; e < e1; ++e) {
            const double4 fa = T.fang[e];
            f[0] += fa.x; f[1] += fa.y; f[2] += fa.z;
        }
    }
}

// ---------------------------------------------------------------- kernels ------

// CBond::AddForce (src/Bond.cpp:111-162) and CBondPair::AddForce for every bond and angle
// of this step, in FP64: thread q < nbonds takes bond q, the rest take angles.  The force
// goes to fbond / fang, from where k_advance / k_finish gather each bead's sum.  OBS adds
// the bond virial (to the head bead's box, as CBond::AddForce adds its stress to the head)
// and the bond and angle potential energies.  Bonds and angles never cross boxes.
// q = this thread's bond or angle, q0 = the workgroup's first (its box receives the
// workgroup's observables); workgroups of up to 64 * WAVES threads
template <bool OBS, bool MULTI, int WAVES>
__device__ __attribute__((always_inline)) inline void bonded_body(const Buffers& S, const int* __restrict__ slot_of,
                                                                  const Topo& T, const BoxDev* boxes, const BoxDev& B0,
                                                                  double* __restrict__ acc, int q, int q0)
{
    double ob[12];
    if (OBS)
        for (int i = 0; i < 12; ++i) ob[i] = 0.0;
    int box = -1;
    if (q < T.nbonds) {
        const int2 hb = T.bond_bead[q];
        const int sh = slot_of[hb.x], st = slot_of[hb.y];
        const int w = __float_as_int(S.pos[sh].w);
        box = bead_box(w);
        const BoxDev& B = box_of<MULTI>(boxes, B0, w);
        double uh[3], ut[3];
        unwrapped_pos(sh, hb.x, S, B, uh);
        unwrapped_pos(st, hb.y, S, B, ut);
        const double2 par = T.bond_par[q];
        const double d[3] = {uh[0] - ut[0], uh[1] - ut[1], uh[2] - ut[2]};
        const double L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double s = par.x * (par.y - L) / L;
        const double fb[3] = {s * d[0], s * d[1], s * d[2]};
        // straight into the two beads' entries (negation is exact: the sums are those of
        // one force per bond added with its sign)
        const int2 ent = T.bond_csr[q];
        T.fbond[ent.x] = make_double4(fb[0], fb[1], fb[2], 0.0);
        T.fbond[ent.y] = make_double4(-fb[0], -fb[1], -fb[2], 0.0);
        if (OBS) {
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) ob[3 * a + b] = d[b] * fb[a];
            ob[10] = 0.5 * par.x * (L - par.y) * (L - par.y);
        }
    } else if (q < T.nbonds + T.nangles) {
        const int g = q - T.nbonds;
        const int4 abc = T.ang_bead[g];
        const int sa = slot_of[abc.x], sb = slot_of[abc.y], sc = slot_of[abc.z];
        const int w = __float_as_int(S.pos[sb].w);
        box = bead_box(w);
        const BoxDev& B = box_of<MULTI>(boxes, B0, w);
        double ua[3], ub[3], uc[3], f[3][3], pe = 0.0;
        unwrapped_pos(sa, abc.x, S, B, ua);
        unwrapped_pos(sb, abc.y, S, B, ub);
        unwrapped_pos(sc, abc.z, S, B, uc);
        angle_forces(ua, ub, uc, T.ang_par[g], f, OBS ? &pe : nullptr);
        const int4 ent = T.ang_csr[g];
        T.fang[ent.x] = make_double4(f[0][0], f[0][1], f[0][2], 0.0);
        T.fang[ent.y] = make_double4(f[1][0], f[1][1], f[1][2], 0.0);
        T.fang[ent.z] = make_double4(f[2][0], f[2][1], f[2][2], 0.0);
        if (OBS) ob[11] = pe;
    }
    if (OBS) {
        // slots O_VIR .. O_PEA: virial (9), pair PE (untouched here), bond PE, angle PE
        int box0 = 0;
        if (MULTI) {
            const int s0 = q0 < T.nbonds ? slot_of[T.bond_bead[q0].x] : slot_of[T.ang_bead[q0 - T.nbonds].y];
            box0 = bead_box(__float_as_int(S.pos[s0].w));
        } else {
            box = box < 0 ? -1 : 0;
        }
        accumulate_boxes<12, WAVES>(ob, box, box0, kObsVir, acc);
    }
}

template <bool OBS, bool MULTI>
__global__ __launch_bounds__(256) void k_bonded(Buffers S, const int* __restrict__ slot_of, Topo T,
                                                const BoxDev* boxes, BoxDev B0, double* __restrict__ acc)
{
    const int q0 = blockIdx.x * blockDim.x;
    bonded_body<OBS, MULTI, 4>(S, slot_of, T, boxes, B0, acc, q0 + threadIdx.x, q0);
}

// The hit-list pair kernel and the bonded terms of the same step in one launch: the pair
// workgroups come first, the bond and angle workgroups after them, so the bonded work fills
// the CUs the pair kernel's last workgroups leave idle (the two write disjoint buffers).
// (one box; held to 80 VGPRs like the hit-list kernel, six wavefronts per SIMD)
template <bool OBS, int THREADS, int CAP, int FLUSH, int SPLIT, bool EPI>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(hit_waves_eu(OBS, THREADS, CAP, SPLIT))))
void k_pair_hitlist_bonded(PairArgs A, BoxDev B0, int nblocks, int per_xcd, Buffers S,
                           const int* __restrict__ slot_of, Topo T)
{
    const int npair = 8 * per_xcd;
    if ((int)blockIdx.x >= npair) {
        const int q0 = ((int)blockIdx.x - npair) * THREADS;
        bonded_body<OBS, false, THREADS / 64>(S, slot_of, T, A.boxes, B0, A.acc, q0 + (int)threadIdx.x, q0);
        return;
    }
    hitlist_body<OBS, false, THREADS, CAP, FLUSH, SPLIT, EPI>(A, B0, nblocks, per_xcd);
}

// the fused launch of hit-list schedule S with `nq` bonds and angles
template <int S, bool OBS, bool EPI>
void launch_hitlist_bonded_s(const PairArgs& a, const BoxDev& B0, int nq, const Buffers& Sb, const int* slot_of,
                             const Topo& T, hipStream_t st)
{
    constexpr Schedule q = kSchedule[S];
    const int gh = (a.n + q.beads() - 1) / q.beads();
    const int ph = (gh + 7) / 8;
    const dim3 gr(8 * ph + (nq + q.threads - 1) / q.threads), bl(q.threads);
    hipLaunchKernelGGL((k_pair_hitlist_bonded<OBS, q.threads, q.cap, q.flush, q.split, EPI>), gr, bl, 0, st,
                       a, B0, gh, ph, Sb, slot_of, T);
}

template <bool OBS, bool EPI>
void launch_hitlist_bonded(int sched, const PairArgs& a, const BoxDev& B0, int nq, const Buffers& Sb,
                           const int* slot_of, const Topo& T, hipStream_t st)
{
    switch (sched) {
    case 1: launch_hitlist_bonded_s<1, OBS, EPI>(a, B0, nq, Sb, slot_of, T, st); break;
    case 2: launch_hitlist_bonded_s<2, OBS, EPI>(a, B0, nq, Sb, slot_of, T, st); break;
    case 3: launch_hitlist_bonded_s<3, OBS, EPI>(a, B0, nq, Sb, slot_of, T, st); break;
    case 5: launch_hitlist_bonded_s<5, OBS, EPI>(a, B0, nq, Sb, slot_of, T, st); break;
    case 6: launch_hitlist_bonded_s<6, OBS, EPI>(a, B0, nq, Sb, slot_of, T, st); break;
    default: launch_hitlist_bonded_s<0, OBS, EPI>(a, B0, nq, Sb, slot_of, T, st); break;
    }
}

// The bonded flag in pos.w: cleared on every slot (and the slot of every bead rewritten,
// since steps store it for flagged beads only), then set on the bonded beads' slots.
__global__ __launch_bounds__(256) void k_clear_bonded_flag(int n, float4* __restrict__ pos,
                                                           const int* __restrict__ bead, int* __restrict__ slot_of)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    pos[s].w = __int_as_float(__float_as_int(pos[s].w) & ~kBondedBit);
    slot_of[bead[s]] = s;
}

__global__ __launch_bounds__(256) void k_set_bonded_flag(int nbonded, const int* __restrict__ bonded,
                                                         const int* __restrict__ slot_of, float4* __restrict__ pos)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbonded) return;
    const int s = slot_of[bonded[t]];
    pos[s].w = __int_as_float(__float_as_int(pos[s].w) | kBondedBit);
}

// Add each bonded bead's bond and angle forces of this step to its pair force, in place
// (one thread per bead that has bonds; its sum in FP64, rounded once).
__global__ __launch_bounds__(256) void k_bonded_sum(Topo T, const int* __restrict__ slot_of,
                                                    float4* __restrict__ frc)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T.nbonded) return;
    const int q = T.bonded[t];
    double fb[3];
    bonded_gather(q, T, fb);
    const int s = slot_of[q];
    float4 f = frc[s];
    f.x += (float)fb[0]; f.y += (float)fb[1]; f.z += (float)fb[2];
    frc[s] = f;
}

// Close the previous step (kEndBegin) and start the next one: new cell key and rank.
// kEndBegin also adds the previous step's bond and angle forces of bonded beads (flagged in
// pos.w) to their pair force, in place: within a sequence of steps the bonded sum of every
// step but the last is left to the next step's k_advance (same arithmetic as k_bonded_sum).
template <int MODE, bool MULTI>
__global__ __launch_bounds__(256) void k_advance(int n, Buffers S, float4* __restrict__ frc, Topo T,
                                                 const BoxDev* boxes,
                                                 BoxDev B0, int* __restrict__ cnt, int* __restrict__ gcnt,
                                                 int* __restrict__ nkey, int* __restrict__ rank,
                                                 int* __restrict__ flags, int64_t* __restrict__ step)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (MODE != kRebin && s == 0) *step += 1;   // simulation time of the coming force evaluation
    const bool active = s < n;
    int k = 0;
    if (active) {
        const float4 p = S.pos[s];
        const BoxDev& B = box_of<MULTI>(boxes, B0, __float_as_int(p.w));
        float4 v = S.vel[s];
        float4 f = (MODE == kRebin) ? S.fold[s] : frc[s];
        if (MODE == kEndBegin && (__float_as_int(p.w) & kBondedBit)) {
            double fb[3];
            bonded_gather(S.bead[s], T, fb);
            f.x += (float)fb[0]; f.y += (float)fb[1]; f.z += (float)fb[2];
            frc[s] = f;
        }
        if (MODE == kEndBegin) {
            // UpdateMom: the stored velocity is v + dt/2 F_old; k_permute repeats this sum
            v.x = kick(v.x, f.x, B.hdt);
            v.y = kick(v.y, f.y, B.hdt);
            v.z = kick(v.z, f.z, B.hdt);
            if (!(fabsf(f.x) + fabsf(f.y) + fabsf(f.z) < 1e30f) || !(fabsf(v.x) + fabsf(v.y) + fabsf(v.z) < 1e30f))
                atomicAdd(flags + 0, 1);
        }
        int c[3];
        key_cell(S.key[s] - B.koff, B, c);
        const float o[3] = {p.x, p.y, p.z};
        const int img[3] = {0, 0, 0};   // the image change is applied by k_permute
        float d[3] = {0.f, 0.f, 0.f};
        if (MODE != kRebin) {
            d[0] = drift(v.x, f.x, B);
            d[1] = drift(v.y, f.y, B);
            d[2] = drift(v.z, f.z, B);
        }
        Moved m;
        advance(o, c, img, d, B, m);
        k = B.koff + cell_key(m.c[0], m.c[1], m.c[2], B);
        nkey[s] = k;
        if (m.jump) atomicAdd(flags + 1, 1);
        if (m.bad) atomicAdd(flags + 0, 1);
    }
    // Counting-sort ranks.  Beads arrive cell-sorted, so a wavefront's keys repeat: one
    // atomic per distinct cell (and per distinct scan group) instead of one per bead, which
    // would serialise ~192 same-address atomics per group counter.
    const int lane = threadIdx.x & 63;
    int r, size, leader;
    wave_groups(k, active, r, size, leader);
    int base = 0;
    if (active && lane == leader) base = atomicAdd(cnt + k, size);
    base = __shfl(base, leader, 64);
    if (active) rank[s] = base + r;
    wave_groups(k >> 6, active, r, size, leader);
    if (active && lane == leader) atomicAdd(gcnt + (k >> 6), size);
}

// Exclusive scan of in[0, m) into out, in[] cleared, by one workgroup of 1024 threads: in
// rounds of kScanRound entries, wavefront w takes up to kScanChunks consecutive chunks of 64
// entries, loads all of them at once (one coalesced load per chunk, all in flight), scans each
// chunk with DPP (wave_incl_scan) in order into LDS, and the 16 wavefront totals are added in
// order through LDS.  ex: LDS of kScanRound ints.
constexpr int kScanChunks = 8;
constexpr int kScanRound = 16 * 64 * kScanChunks;
__device__ inline void scan_block(int* in, int* __restrict__ out, int m, int* ex, int* wsum, int* carry)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t == 0) *carry = 0;
    __syncthreads();
    for (int base = 0; base < m; base += kScanRound) {
        const int nch = (min(m - base, kScanRound) + 63) >> 6;   // chunks this round
        const int per = (nch + 15) >> 4;                          // chunks per wavefront
        const int ch0 = wave * per;
        int v[kScanChunks];
#pragma unroll
        for (int j = 0; j < kScanChunks; ++j) {
            const int k = base + ((ch0 + j) << 6) + lane;
            v[j] = (j < per && k < m) ? in[k] : 0;
        }
        int run = 0;
#pragma unroll
        for (int j = 0; j < kScanChunks; ++j) {
            const int incl = wave_incl_scan(v[j]);
            if (j < per) ex[((ch0 + j) << 6) + lane] = run + incl - v[j];
            run += __builtin_amdgcn_readlane(incl, 63);
        }
        if (lane == 0) wsum[wave] = run;
        __syncthreads();
        int off = *carry;
        for (int w = 0; w < wave; ++w) off += wsum[w];
        for (int j = 0; j < per; ++j) {
            const int k = base + ((ch0 + j) << 6) + lane;
            if (k < m) {
                out[k] = off + ex[((ch0 + j) << 6) + lane];
                in[k] = 0;
            }
        }
        __syncthreads();
        if (t == 0) {
            int tot = 0;
            for (int w = 0; w < 16; ++w) tot += wsum[w];
            *carry += tot;
        }
        __syncthreads();
    }
}

// exclusive scan of the scan groups' counts (one workgroup), clears the counts
// `step` (may be null): the simulation time to advance, when the epilogue of the previous
// pair launch did the work of k_advance
__global__ __launch_bounds__(1024) void k_group_scan(int ngroups, int* __restrict__ gcnt, int* __restrict__ gstart,
                                                     int64_t* __restrict__ step)
{
    __shared__ int ex[kScanRound];
    __shared__ int wsum[16];
    __shared__ int carry;
    if (step && threadIdx.x == 0) *step += 1;
    scan_block(gcnt, gstart, ngroups, ex, wsum, &carry);
}

// Engines of up to kSmallScanKeys cell keys: the cell starts in one launch (scan_block over
// the cell counts), which replaces k_group_scan + k_cell_start; clears the cell and group
// counts.
constexpr int kSmallScanKeys = 16384;
__global__ __launch_bounds__(1024) void k_cell_scan_small(int nkeys, int ngroups, int n, int* __restrict__ cnt,
                                                          int* __restrict__ gcnt, int* __restrict__ cstart,
                                                          int64_t* __restrict__ step)
{
    __shared__ int ex[kScanRound];
    __shared__ int wsum[16];
    __shared__ int carry;
    if (step && threadIdx.x == 0) *step += 1;
    scan_block(cnt, cstart, nkeys, ex, wsum, &carry);
    for (int g = threadIdx.x; g < ngroups; g += 1024) gcnt[g] = 0;
    if (threadIdx.x == 0) cstart[nkeys] = n;
}

// cell starts: one wavefront per scan group scans its 64 cell counts; clears the counts
__global__ __launch_bounds__(256) void k_cell_start(int ngroups, int n, const int* __restrict__ gstart,
                                                    int* __restrict__ cnt, int* __restrict__ cstart)
{
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (g >= ngroups) return;
    const int k = g * 64 + lane;
    const int c = cnt[k];
    const int incl = wave_incl_scan(c);
    cstart[k] = gstart[g] + incl - c;
    cnt[k] = 0;
    if (g == ngroups - 1 && lane == 63) cstart[ngroups * 64] = n;
}

// Engines of up to kDirectGroups scan groups: the cell starts in one launch without a scan of
// the group counts.  Each workgroup (four groups, a wavefront each) sums the counts of every
// group before its first itself, from L2 (at most kDirectGroups ints per workgroup, 8 MB
// in all for 4096 groups), then each wavefront scans its 64 cell counts as k_cell_start does.
// Clears the cell counts; the group counts stay for the other workgroups and k_place clears
// them.  `step` (may be null): as for k_group_scan.
constexpr int kDirectGroups = 8192;
__global__ __launch_bounds__(256) void k_cell_start_direct(int ngroups, int n, const int* __restrict__ gcnt,
                                                           int* __restrict__ cnt, int* __restrict__ cstart,
                                                           int64_t* __restrict__ step)
{
    __shared__ int wsum[4];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int g0 = blockIdx.x * 4;
    if (step && blockIdx.x == 0 && t == 0) *step += 1;
    int p = 0;
#pragma unroll 4
    for (int i = t; i < g0; i += 256) p += gcnt[i];
    p = __builtin_amdgcn_readlane(wave_incl_scan(p), 63);
    if (lane == 0) wsum[w] = p;
    __syncthreads();
    const int g = g0 + w;
    if (g >= ngroups) return;
    int start = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    for (int j = g0; j < g; ++j) start += gcnt[j];
    const int k = g * 64 + lane;
    const int c = cnt[k];
    const int incl = wave_incl_scan(c);
    cstart[k] = start + incl - c;
    cnt[k] = 0;
    if (g == ngroups - 1 && lane == 63) cstart[ngroups * 64] = n;
}

// Counting-sort placement, first half: every source slot s writes (bead index, s) to the
// position its rank gives it inside its new cell.  Ranks come from atomics, so the order
// inside a cell is arbitrary here; k_permute restores a deterministic one.  gclear (may be
// null): the group counts k_cell_start_direct left, cleared by threads s < ngroups (the grid
// covers max(n, ngroups) threads).
__global__ __launch_bounds__(256) void k_place(int n, const int* __restrict__ bead, const int* __restrict__ nkey,
                                               const int* __restrict__ rank, const int* __restrict__ cstart,
                                               int2* __restrict__ tmp, int* __restrict__ gclear, int ngroups)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (gclear && s < ngroups) gclear[s] = 0;
    if (s >= n) return;
    tmp[cstart[nkey[s]] + rank[s]] = make_int2(bead[s], s);
}

// Counting-sort placement, second half: the bead placed at d goes to the start of its cell
// plus the number of the cell's beads with a smaller bead index, so the order inside every
// cell is ascending bead index, a function of the state alone (reruns, and a state
// downloaded and uploaded again, lay the beads out identically).  One pass over the cell's m
// entries per bead, from L1.  Then the move itself (recomputing the same advance as
// k_advance).
template <int MODE, bool MULTI>
__global__ __launch_bounds__(256) void k_permute(int n, Buffers S, Buffers D, float4* __restrict__ frc,
                                                 float4* __restrict__ vp, const BoxDev* boxes, BoxDev B0,
                                                 const int* __restrict__ nkey, const int2* __restrict__ tmp,
                                                 const int* __restrict__ cstart, int* __restrict__ slot_of,
                                                 int write_vp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 t0 = tmp[i];
    const int s = t0.y;
    const int k = nkey[s];
    const int c0 = cstart[k], c1 = cstart[k + 1];
    int dst = c0;
    for (int e = c0; e < c1; ++e) dst += tmp[e].x < t0.x;
    const float4 p = S.pos[s];
    const BoxDev& B = box_of<MULTI>(boxes, B0, __float_as_int(p.w));
    const float4 f = (MODE == kRebin) ? S.fold[s] : frc[s];
    float4 v = S.vel[s];
    if (MODE == kEndBegin) {
        // UpdateMom of the step that ends (the same sum as k_advance)
        v.x = kick(v.x, f.x, B.hdt);
        v.y = kick(v.y, f.y, B.hdt);
        v.z = kick(v.z, f.z, B.hdt);
    }
    const int q = t0.x;   // bead index of the source slot
    int c[3];
    key_cell(S.key[s] - B.koff, B, c);
    const float o[3] = {p.x, p.y, p.z};
    const int img[3] = {0, 0, 0};   // m.img is the change of image this step
    float d[3] = {0.f, 0.f, 0.f};
    if (MODE != kRebin) {
        d[0] = drift(v.x, f.x, B);
        d[1] = drift(v.y, f.y, B);
        d[2] = drift(v.z, f.z, B);
    }
    Moved m;
    advance(o, c, img, d, B, m);
    D.pos[dst] = make_float4(m.o[0], m.o[1], m.o[2], p.w);
    if (MODE == kRebin) {
        D.vel[dst] = v;
        D.fold[dst] = f;
    } else {
        // the first half of the next UpdateMom: v + dt/2 F, so the force need not be kept
        D.vel[dst] = make_float4(kick(v.x, f.x, B.hdt), kick(v.y, f.y, B.hdt), kick(v.z, f.z, B.hdt), v.w);
    }
    D.bead[dst] = q;
    // the image counts live by bead index: only a bead that crosses the box changes them,
    // and only its own thread writes them
    if (m.img[0] | m.img[1] | m.img[2]) {
        int4 im = S.img[q];
        im.x += m.img[0]; im.y += m.img[1]; im.z += m.img[2];
        S.img[q] = im;
    }
    D.key[dst] = k;
    D.code[dst] = scan_code(m.o, m.c[0], B);
    if (MODE == kRebin) {
        vp[dst] = v;
    } else {
        // (equal to the new velocity's bits when lambda dt = dt/2: the steps then read that)
        if (write_vp)
            vp[dst] = make_float4(kick(v.x, f.x, B.lamdt), kick(v.y, f.y, B.lamdt), kick(v.z, f.z, B.lamdt), v.w);
    }
    // the slot of each bead is read for bonded beads only (bond and angle terms); a scattered
    // 4-byte store per bead costs a partial-line write each, so the others are skipped once
    // the topology has flagged the bonded ones (an upload re-binning writes every bead)
    if (MODE == kRebin || (__float_as_int(p.w) & kBondedBit)) slot_of[q] = dst;
}

// REBIN leaves the uploaded force in fold; the canonical state needs it in frc too
__global__ void k_copy_force(int n, const float4* __restrict__ fold, float4* __restrict__ frc)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) frc[s] = fold[s];
}

// Close the last step: UpdateMom (UPDATE_V) and the FP64 sums of v^2 and momentum (the
// force already holds the bonded terms; the virial and energies come from the pair and
// bonded kernels).
template <bool UPDATE_V, bool OBS, bool MULTI>
__global__ __launch_bounds__(256) void k_finish(int n, Buffers S, const float4* __restrict__ frc,
                                                const BoxDev* boxes, BoxDev B0, int* __restrict__ flags,
                                                double* __restrict__ acc, float4* __restrict__ vp_keep)
{
    static_assert(O_V2 == 0 && O_MOM == 3, "kinetic slots first");
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double ob[6];
    if (OBS)
        for (int i = 0; i < 6; ++i) ob[i] = 0.0;
    int box = -1;
    if (s < n) {
        const int w = __float_as_int(S.pos[s].w);
        box = bead_box(w);
        const BoxDev& B = box_of<MULTI>(boxes, B0, w);
        const float4 f = frc[s];
        float4 v = S.vel[s];
        if (UPDATE_V) {
            // the last step's predicted velocity, which the steps read from vel (vp_is_vel)
            if (vp_keep) vp_keep[s] = v;
            // the second half of UpdateMom (the stored velocity is v + dt/2 F_old)
            v.x = kick(v.x, f.x, B.hdt);
            v.y = kick(v.y, f.y, B.hdt);
            v.z = kick(v.z, f.z, B.hdt);
            S.vel[s] = v;
        }
        if (!(fabsf(f.x) + fabsf(f.y) + fabsf(f.z) < 1e30f) || !(fabsf(v.x) + fabsf(v.y) + fabsf(v.z) < 1e30f))
            atomicAdd(flags + 0, 1);
        if (OBS) {
            ob[O_V2 + 0] = (double)v.x * v.x;
            ob[O_V2 + 1] = (double)v.y * v.y;
            ob[O_V2 + 2] = (double)v.z * v.z;
            ob[O_MOM + 0] = v.x;
            ob[O_MOM + 1] = v.y;
            ob[O_MOM + 2] = v.z;
        }
    }
    if (OBS) {
        const int box0 = MULTI ? bead_box(__float_as_int(S.pos[blockIdx.x * blockDim.x].w)) : 0;
        accumulate_boxes<6>(ob, MULTI ? box : (s < n ? 0 : -1), box0, 0, acc);
    }
}

// FP64 sums per box of v and v^2 and the bead count: acc[box][0..4]
template <bool MULTI>
__global__ __launch_bounds__(256) void k_mom_sums(int n, Buffers S, double* __restrict__ acc)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double ob[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    int box = -1;
    if (s < n) {
        box = MULTI ? bead_box(__float_as_int(S.pos[s].w)) : 0;
        const float4 v = S.vel[s];
        ob[0] = v.x; ob[1] = v.y; ob[2] = v.z;
        ob[3] = (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z;
        ob[4] = 1.0;
    }
    const int box0 = MULTI ? bead_box(__float_as_int(S.pos[blockIdx.x * blockDim.x].w)) : 0;
    accumulate_boxes<5>(ob, box, box0, 0, acc);
}

// v <- (v - shift) * scale with adj[box] = (shift x, y, z, scale), in FP64
template <bool MULTI>
__global__ __launch_bounds__(256) void k_mom_apply(int n, Buffers S, const double4* __restrict__ adj)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int box = MULTI ? bead_box(__float_as_int(S.pos[s].w)) : 0;
    const double4 a = adj[box];
    float4 v = S.vel[s];
    v.x = (float)(((double)v.x - a.x) * a.w);
    v.y = (float)(((double)v.y - a.y) * a.w);
    v.z = (float)(((double)v.z - a.z) * a.w);
    S.vel[s] = v;
}

// state back to upload order in FP64
template <bool MULTI>
__global__ __launch_bounds__(256) void k_gather(int n, Buffers S, const float4* __restrict__ frc,
                                                const BoxDev* boxes, BoxDev B0, double* __restrict__ x,
                                                double* __restrict__ u, double* __restrict__ v,
                                                double* __restrict__ f)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float4 p = S.pos[s];
    const BoxDev& B = box_of<MULTI>(boxes, B0, __float_as_int(p.w));
    int c[3];
    key_cell(S.key[s] - B.koff, B, c);
    const int q = S.bead[s];
    const int4 a = S.img[q];
    const float4 vv = S.vel[s];
    const float4 ff = frc[s];
    const double o[3] = {p.x, p.y, p.z};
    const int im[3] = {a.x, a.y, a.z};
    for (int d = 0; d < 3; ++d) {
        const double xa = (double)c[d] * B.wd[d] + o[d];
        x[3 * q + d] = xa;
        u[3 * q + d] = xa + (double)im[d] * (double)B.n[d] * B.wd[d];
    }
    v[3 * q + 0] = vv.x; v[3 * q + 1] = vv.y; v[3 * q + 2] = vv.z;
    f[3 * q + 0] = ff.x; f[3 * q + 1] = ff.y; f[3 * q + 2] = ff.z;
}

// clears the FP64 accumulators and error flags (a kernel rather than a memset, so that
// it can sit inside a captured graph)
__global__ void k_zero(double* acc, int nacc, int* flags, int nflags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nacc) acc[i] = 0.0;
    if (i < nflags) flags[i] = 0;
}

__global__ void k_set_step(int64_t* step, int64_t value)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *step = value;
}

}  // namespace

// ------------------------------------------------------------------ host side ---

struct BoxHost {
    dpd_params p{};
    std::vector<double> a, g;
    bool thermostat_on = true, conservative_on = true;
    bool forces_on = true;   // ToggleDPDBeadForces: off zeroes the half-step constants
    double dt = 0, lambda = 0, kT = 0;
    BoxDev B{};
    // staged beads and topology (box-local indices)
    int64_t n = 0;
    std::vector<int32_t> id, type;
    std::vector<double> x, u, v, f;
    bool has_f = false;
    std::vector<int32_t> bh, bt, pf, ps;
    std::vector<double> bk, bl0, pk, pphi;
    int64_t bead_off = 0;
};

struct dpd_engine {
    std::vector<BoxHost> boxes;
    int device = 0;
    hipStream_t stream = nullptr;
    int ngroups = 0, nkeys = 0;   // scan groups of 64 keys, keys (all boxes)
    int64_t n = 0;
    int64_t step = 0;
    int cur = 0;
    int pair_kernel = DPD_PAIR_FULLSHELL;   // the kernel of schedule `sched`
    int sched = kFullShellSchedule;         // pair-launch schedule (kSchedule)
    bool pair_auto = false;   // choose by bead count at each commit
    bool committed = false;
    // device buffers
    BoxDev* dboxes = nullptr;
    float4 *pos[2] = {}, *vel[2] = {}, *fold[2] = {};
    int* bead[2] = {};
    int4* img = nullptr;   // periodic image counts by bead index
    int* key[2] = {};
    uint32_t* code[2] = {};   // n + kCodePad: the hit-list scan reads up to 7 slots past the end
    float4 *frc = nullptr, *vp = nullptr;
    int *slot_of = nullptr, *nkey = nullptr, *rank = nullptr;
    int2* place = nullptr;    // counting-sort placement: (bead index, source slot) per new slot
    int *cnt = nullptr, *gcnt = nullptr, *gstart = nullptr, *cstart = nullptr;
    int* flags = nullptr;
    int64_t* dstep = nullptr;
    double* acc = nullptr;
    float* table = nullptr;
    double* gather = nullptr;   // 12n doubles for downloads
    double* slice = nullptr;    // slice stress tables (pair | bond)
    double4* madj = nullptr;    // per-box velocity shift and scale
    int4* hitblk[kSchedules] = {};   // hit-list workgroups of a batch (box, first slot, end slot)
    int nhitblk[kSchedules] = {};    // per schedule (its beads per workgroup)
    size_t slice_cap = 0;
    // topology
    int *bond_ptr = nullptr, *ang_ptr = nullptr;
    int2 *bond_csr = nullptr, *bond_bead = nullptr;
    int4* ang_csr = nullptr;
    double2* bond_par = nullptr;
    int4* ang_bead = nullptr;
    double4 *ang_par = nullptr, *fbond = nullptr, *fang = nullptr;
    int* bonded = nullptr;
    int nbonds = 0, nangles = 0, nbonded = 0;
    int has_bonds = 0, has_angles = 0;
    // observables
    std::vector<dpd_obs> obs;
    bool obs_valid = false;
    int64_t jumps = 0;
    // graphs (steps, observe) -> executable graph; valid while buffers stay put
    std::map<std::pair<int64_t, int>, hipGraphExec_t> graphs;
    std::map<std::pair<int64_t, int>, hipGraph_t> graph_src;
    // timing
    hipEvent_t ev[2] = {};
    std::vector<hipEvent_t> pev;
    double t_total = 0, t_pair = 0;
    int64_t n_pair = 0;
    std::string err;
};

namespace {

int fail(dpd_engine* e, int code, const std::string& msg)
{
    if (e) e->err = msg;
    return code;
}

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t _st = (expr);                                                           \
        if (_st != hipSuccess)                                                             \
            return fail(e, DPD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_st)); \
    } while (0)

template <class T>
int dalloc(dpd_engine* e, T** p, size_t count)
{
    if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (count == 0) count = 1;
    hipError_t st = hipMalloc((void**)p, sizeof(T) * count);
    if (st != hipSuccess) return fail(e, DPD_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(st));
    return DPD_OK;
}

// With lambda = 1/2 and the forces on, the predicted velocity v + lambda dt f and the stored
// half step v + dt/2 f are the same FP32 expression (lambda dt and dt/2 round to the same
// float): steps then read the predicted velocities from vel, k_permute does not write vp, and
// k_finish keeps the last step's in vp for the calls between steps.
inline bool vp_is_vel(const dpd_engine* e)
{
    for (const auto& b : e->boxes)
        if (!(b.B.hdt == b.B.lamdt)) return false;
    return true;
}

void drop_graphs(dpd_engine* e)
{
    for (auto& g : e->graphs) (void)hipGraphExecDestroy(g.second);
    for (auto& g : e->graph_src) (void)hipGraphDestroy(g.second);
    e->graphs.clear();
    e->graph_src.clear();
}

// switch the pair launch to schedule `s` (captured graphs hold the old launch)
void set_schedule(dpd_engine* e, int s)
{
    if (s != e->sched) drop_graphs(e);
    e->sched = s;
    e->pair_kernel = kSchedule[s].kernel;
}

// the automatic choice: the hit-list kernel's default schedule for every engine whose slots
// fit its 26-bit list entries, the full-shell kernel beyond
int default_schedule(const dpd_engine* e)
{
    return e->n <= (int64_t)kHitSlotMask ? kDefaultSchedule : kFullShellSchedule;
}

void set_step_constants(BoxHost& b)
{
    // CCNTCell::SetTimeStepConstants / ChangeTimeStepConstants (src/CNTCell.cpp:180-203, :357-377)
    // CCNTCell::ToggleDPDBeadForces (src/CNTCell.cpp:223-227) passes dt = 0 for the two
    // half-step constants: forces no longer move beads or change velocities
    const double hdt = b.forces_on ? b.dt : 0.0;
    b.B.dt = (float)b.dt;
    b.B.hdt = (float)(0.5 * hdt);
    b.B.hdt2 = (float)(0.5 * hdt * hdt);
    b.B.lamdt = (float)(b.lambda * b.dt);
    b.B.invrootdt = (float)std::sqrt(24.0 * b.kT / b.dt);
}

int check_params(dpd_engine* e, const dpd_params* p)
{
    if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fail(e, DPD_ERR_PARAM, "need at least 3 cells per axis");
    if ((int64_t)p->nx * p->ny * p->nz > (int64_t)1 << 24)
        return fail(e, DPD_ERR_PARAM, "more than 2^24 cells in one box");
    if (p->wx < 1.0 || p->wy < 1.0 || p->wz < 1.0) return fail(e, DPD_ERR_PARAM, "cell widths must be >= 1");
    if (p->ntypes < 1 || p->ntypes > DPD_MAX_TYPES) return fail(e, DPD_ERR_PARAM, "ntypes out of range");
    if (!(p->dt > 0.0) || p->lambda < 0.0 || p->lambda > 1.0 || p->kT < 0.0)
        return fail(e, DPD_ERR_PARAM, "invalid dt, lambda or kT");
    if (!p->a || !p->gamma) return fail(e, DPD_ERR_PARAM, "null interaction matrices");
    const int nt = p->ntypes;
    for (int i = 0; i < nt; ++i)
        for (int j = 0; j < nt; ++j)
            if (p->a[i * nt + j] != p->a[j * nt + i] || p->gamma[i * nt + j] != p->gamma[j * nt + i] ||
                p->gamma[i * nt + j] < 0.0)
                return fail(e, DPD_ERR_PARAM, "interaction matrices must be symmetric, gamma >= 0");
    return DPD_OK;
}

int upload_tables(dpd_engine* e)
{
    const int nb = (int)e->boxes.size();
    std::vector<float> t((size_t)nb * kTableStride, 0.0f);
    for (int b = 0; b < nb; ++b) {
        const BoxHost& h = e->boxes[b];
        const int nt = h.p.ntypes;
        float* tb = t.data() + (size_t)b * kTableStride;
        for (int i = 0; i < nt; ++i)
            for (int j = 0; j < nt; ++j) {
                const double a = h.conservative_on ? h.a[i * nt + j] : 0.0;
                const double g = h.thermostat_on ? h.g[i * nt + j] : 0.0;
                tb[i * kMaxTypes + j] = (float)a;
                tb[kTypeStride + i * kMaxTypes + j] = (float)g;
                tb[2 * kTypeStride + i * kMaxTypes + j] = (float)std::sqrt(g);
            }
    }
    HIPCHK(hipMemcpyAsync(e->table, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DPD_OK;
}

int upload_boxes(dpd_engine* e)
{
    std::vector<BoxDev> bd;
    for (auto& b : e->boxes) bd.push_back(b.B);
    HIPCHK(hipMemcpyAsync(e->dboxes, bd.data(), sizeof(BoxDev) * bd.size(), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DPD_OK;
}

Buffers bufs(dpd_engine* e, int i)
{
    return Buffers{e->pos[i], e->vel[i], e->fold[i], e->bead[i], e->img, e->key[i], e->code[i]};
}

Topo topo(dpd_engine* e)
{
    return Topo{e->bond_ptr, e->bond_csr, e->bond_bead, e->bond_par, e->ang_ptr, e->ang_csr, e->ang_bead,
                e->ang_par, e->fbond, e->fang, e->bonded, e->nbonds, e->nangles, e->nbonded, e->has_bonds,
                e->has_angles};
}

inline int grid(int64_t n, int b) { return (int)((n + b - 1) / b); }
inline bool multi(dpd_engine* e) { return e->boxes.size() > 1; }

// counting sort of buffer `cur` into buffer `cur^1`; MODE decides the move.  `advanced`: the
// previous pair launch's epilogue already closed the step and computed the new keys and ranks
// (MODE kEndBegin only), so k_advance is left out and the scan advances the simulation time.
template <int MODE, bool MULTI>
void sort_step_t(dpd_engine* e, bool advanced)
{
    const int n = (int)e->n;
    Buffers S = bufs(e, e->cur), D = bufs(e, e->cur ^ 1);
    const BoxDev B0 = e->boxes[0].B;
    if (!advanced)
        hipLaunchKernelGGL((k_advance<MODE, MULTI>), dim3(grid(n, 256)), dim3(256), 0, e->stream, n, S, e->frc,
                           topo(e), e->dboxes, B0, e->cnt, e->gcnt, e->nkey, e->rank, e->flags, e->dstep);
    int64_t* bump = advanced ? e->dstep : nullptr;
    int* gclear = nullptr;
    if (e->nkeys <= kSmallScanKeys) {
        hipLaunchKernelGGL(k_cell_scan_small, dim3(1), dim3(1024), 0, e->stream, e->nkeys, e->ngroups, n, e->cnt,
                           e->gcnt, e->cstart, bump);
    } else if (e->ngroups <= kDirectGroups) {
        hipLaunchKernelGGL(k_cell_start_direct, dim3(grid(e->ngroups, 4)), dim3(256), 0, e->stream, e->ngroups, n,
                           e->gcnt, e->cnt, e->cstart, bump);
        gclear = e->gcnt;
    } else {
        hipLaunchKernelGGL(k_group_scan, dim3(1), dim3(1024), 0, e->stream, e->ngroups, e->gcnt, e->gstart, bump);
        hipLaunchKernelGGL(k_cell_start, dim3(grid(e->ngroups, 4)), dim3(256), 0, e->stream, e->ngroups, n, e->gstart,
                           e->cnt, e->cstart);
    }
    hipLaunchKernelGGL(k_place, dim3(grid(std::max<int64_t>(n, gclear ? e->ngroups : 0), 256)), dim3(256), 0,
                       e->stream, n, S.bead, e->nkey, e->rank, e->cstart, e->place, gclear, (int)e->ngroups);
    hipLaunchKernelGGL((k_permute<MODE, MULTI>), dim3(grid(n, 256)), dim3(256), 0, e->stream, n, S, D, e->frc,
                       e->vp, e->dboxes, B0, e->nkey, e->place, e->cstart, e->slot_of,
                       (MODE == kRebin || !vp_is_vel(e)) ? 1 : 0);
    e->cur ^= 1;
}

template <int MODE>
int sort_step(dpd_engine* e, bool advanced = false)
{
    if (multi(e)) sort_step_t<MODE, true>(e, advanced);
    else sort_step_t<MODE, false>(e, advanced);
    HIPCHK(hipGetLastError());
    return DPD_OK;
}

// Whether the pair launch of a step that another follows carries the step's epilogue: for
// engines without bonded terms.  (With them, the bonded beads would need their own pass once
// the bond and angle forces are known: measured on the condensate box, the epilogue added
// 14 us to the pair launch and that pass took 20 us, against the 25 us of k_advance it
// replaced.)
inline bool epilogue_ok(const dpd_engine* e) { return e->nbonded == 0 && epilogue_available(e->sched, e->n); }

// Pair forces of this step.  With the hit-list kernel the bond and angle terms run in the
// same launch (k_pair_hitlist_bonded); `bonded` then tells launch_bonded to add only the sums.
// `epi` (epilogue_ok): the launch also closes the step and starts the next one (pair_epilogue).
int launch_pair(dpd_engine* e, bool obs, bool* bonded = nullptr, bool epi = false, bool in_seq = false)
{
    if (bonded) *bonded = false;
    PairArgs a;
    {
        const Buffers S = bufs(e, e->cur);
        a.vel = S.vel;
        a.nkey = e->nkey;
        a.rank = e->rank;
        a.cnt = e->cnt;
        a.gcnt = e->gcnt;
        a.flags = e->flags;
    }
    a.n = (int)e->n;
    a.nbox = (int)e->boxes.size();
    a.ngroups = e->ngroups;
    a.boxes = e->dboxes;
    a.pos = e->pos[e->cur];
    // inside a sequence of steps with lambda dt = dt/2 the predicted velocities are the
    // stored half steps (k_permute did not write vp)
    a.vp = (in_seq && vp_is_vel(e)) ? e->vel[e->cur] : e->vp;
    a.key = e->key[e->cur];
    a.code = e->code[e->cur];
    a.cstart = e->cstart;
    a.table = e->table;
    a.step = e->dstep;
    a.frc = e->frc;
    a.acc = e->acc;
    a.hitblk = e->hitblk[e->sched];
    a.nhitblk = e->nhitblk[e->sched];
    const int nq = e->nbonds + e->nangles;
    // one box only: in a batch the bonded path's per-lane box raises the fused kernel to 87
    // VGPRs and costs more than the launch it saves (3.59e9 against 3.72e9 bead-steps/s for
    // 256 replicas)
    if (bonded && nq > 0 && e->pair_kernel == DPD_PAIR_HITLIST && a.n <= (int)kHitSlotMask && !multi(e)) {
        const BoxDev B0 = e->boxes[0].B;
        const Buffers S = bufs(e, e->cur);
        // (never with the epilogue: epilogue_ok excludes engines with bonded terms)
        if (obs) launch_hitlist_bonded<true, false>(e->sched, a, B0, nq, S, e->slot_of, topo(e), e->stream);
        else launch_hitlist_bonded<false, false>(e->sched, a, B0, nq, S, e->slot_of, topo(e), e->stream);
        *bonded = true;
        return hipGetLastError() == hipSuccess ? DPD_OK : fail(e, DPD_ERR_HIP, "pair kernel launch failed");
    }
    return launch_pair_kernel(e->sched, a, e->boxes[0].B, obs, e->stream, epi) == hipSuccess
               ? DPD_OK
               : fail(e, DPD_ERR_HIP, "pair kernel launch failed");
}

template <bool UPDATE_V, bool OBS>
void launch_finish(dpd_engine* e)
{
    const int n = (int)e->n;
    const BoxDev B0 = e->boxes[0].B;
    float4* keep = (UPDATE_V && vp_is_vel(e)) ? e->vp : nullptr;
    if (multi(e))
        hipLaunchKernelGGL((k_finish<UPDATE_V, OBS, true>), dim3(grid(n, 256)), dim3(256), 0, e->stream, n,
                           bufs(e, e->cur), e->frc, e->dboxes, B0, e->flags, e->acc, keep);
    else
        hipLaunchKernelGGL((k_finish<UPDATE_V, OBS, false>), dim3(grid(n, 256)), dim3(256), 0, e->stream, n,
                           bufs(e, e->cur), e->frc, e->dboxes, B0, e->flags, e->acc, keep);
}

// bonds and angles of this step (unless the pair launch ran them) and each bonded bead's
// sum added to its force
// `sum` false: the bonded sum is left to the next step's k_advance (kEndBegin)
template <bool OBS>
void launch_bonded(dpd_engine* e, bool terms_done = false, bool sum = true)
{
    const int nq = e->nbonds + e->nangles;
    if (nq == 0) return;
    const BoxDev B0 = e->boxes[0].B;
    if (terms_done)
        ;
    else if (multi(e))
        hipLaunchKernelGGL((k_bonded<OBS, true>), dim3(grid(nq, 256)), dim3(256), 0, e->stream, bufs(e, e->cur),
                           e->slot_of, topo(e), e->dboxes, B0, e->acc);
    else
        hipLaunchKernelGGL((k_bonded<OBS, false>), dim3(grid(nq, 256)), dim3(256), 0, e->stream, bufs(e, e->cur),
                           e->slot_of, topo(e), e->dboxes, B0, e->acc);
    if (sum)
        hipLaunchKernelGGL(k_bonded_sum, dim3(grid(e->nbonded, 256)), dim3(256), 0, e->stream, topo(e),
                           e->slot_of, e->frc);
}

// clear the observable accumulators (and, when `flags`, the error flags)
void launch_zero(dpd_engine* e, bool flags)
{
    const int nacc = (int)e->boxes.size() * kObsStride;
    hipLaunchKernelGGL(k_zero, dim3(grid(std::max(nacc, 4), 256)), dim3(256), 0, e->stream, e->acc, nacc, e->flags,
                       flags ? 4 : 0);
}

int read_flags(dpd_engine* e)
{
    int h[4];
    HIPCHK(hipMemcpyAsync(h, e->flags, sizeof(h), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->jumps += h[1];
    if (h[0]) return fail(e, DPD_ERR_NAN, "non-finite force, velocity or position detected");
    return DPD_OK;
}

int finish_obs(dpd_engine* e)
{
    const size_t nb = e->boxes.size();
    std::vector<double> h(nb * kObsStride);
    HIPCHK(hipMemcpyAsync(h.data(), e->acc, sizeof(double) * h.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->obs.assign(nb, dpd_obs{});
    for (size_t b = 0; b < nb; ++b) {
        const double* a = h.data() + b * kObsStride;
        const dpd_params& p = e->boxes[b].p;
        dpd_obs& o = e->obs[b];
        o.step = e->step;
        o.nbeads = e->boxes[b].n;
        const double V = (double)p.nx * p.wx * (double)p.ny * p.wy * (double)p.nz * p.wz;
        for (int d = 0; d < 3; ++d) { o.sum_v2[d] = a[O_V2 + d]; o.momentum[d] = a[O_MOM + d]; }
        for (int i = 0; i < 9; ++i) o.stress[i] = a[O_VIR + i] / V;
        const double v2 = o.sum_v2[0] + o.sum_v2[1] + o.sum_v2[2];
        o.kT = o.nbeads ? v2 / (3.0 * (double)o.nbeads) : 0.0;
        // src/Monitor.cpp:981-982 keeps the 10-digit constant
        o.pressure = 0.3333333333 * (1.0 / V) * (v2 + a[O_VIR + 0] + a[O_VIR + 4] + a[O_VIR + 8]);
        o.pe_pair = a[O_PEP];
        o.pe_bond = a[O_PEB];
        o.pe_angle = a[O_PEA];
        o.jumps = e->jumps;
    }
    e->obs_valid = true;
    return DPD_OK;
}

// the steps of one dpd_step call, enqueued on the stream (capturable)
// steps whose pair launch is bracketed by events: every step, or the last of every
// kTimingStride with DPD_STEP_TIMING_SAMPLED (never the first step of a call, whose
// launch follows the host's enqueue rather than the previous step)
constexpr int64_t kTimingStride = 16;
inline bool timed_step(int64_t t, int timing)
{
    return timing == 1 || (timing == 2 && t % kTimingStride == kTimingStride - 1);
}
inline int64_t timed_steps(int64_t nsteps, int timing)
{
    return timing == 1 ? nsteps : (timing == 2 ? nsteps / kTimingStride : 0);
}

// timing: 0 none, 1 every pair launch, 2 every kTimingStride-th
int enqueue_steps(dpd_engine* e, int64_t nsteps, bool obs, int timing)
{
    int rc;
    launch_zero(e, false);
    int64_t ti = 0;   // timed launches so far (event pair 2 ti, 2 ti + 1)
    bool advanced = false;   // the previous pair launch carried the step's epilogue
    for (int64_t t = 0; t < nsteps; ++t) {
        rc = (t == 0) ? sort_step<kBegin>(e) : sort_step<kEndBegin>(e, advanced);
        if (rc) return rc;
        const bool last = t == nsteps - 1;
        const bool epi = !last && epilogue_ok(e);
        const bool timed = timed_step(t, timing);
        if (timed) HIPCHK(hipEventRecord(e->pev[2 * ti], e->stream));
        bool terms_done = false;
        if ((rc = launch_pair(e, obs && last, &terms_done, epi, true))) return rc;
        if (timed) HIPCHK(hipEventRecord(e->pev[2 * ti++ + 1], e->stream));
        // the bonded sum of every step but the last is added by the next step's k_advance
        if (obs && last) launch_bonded<true>(e, terms_done);
        else launch_bonded<false>(e, terms_done, last);
        advanced = epi;
    }
    if (obs) launch_finish<true, true>(e);
    else launch_finish<true, false>(e);
    HIPCHK(hipGetLastError());
    return DPD_OK;
}

// Assemble the staged boxes into device buffers and sort them.
int commit_beads(dpd_engine* e, int64_t step0)
{
    int64_t n = 0;
    int nbr = 0;
    for (size_t b = 0; b < e->boxes.size(); ++b) {
        BoxHost& h = e->boxes[b];
        if (h.n == 0) return fail(e, DPD_ERR_STATE, "box " + std::to_string(b) + " has no beads");
        h.bead_off = n;
        h.B.koff = nbr * kGroupKeys;
        h.B.goff = nbr;
        h.B.box = (int)b;
        nbr += h.B.ngroups;
        n += h.n;
    }
    if (n > ((int64_t)1 << 30) || (int64_t)nbr * kGroupKeys > ((int64_t)1 << 30))
        return fail(e, DPD_ERR_PARAM, "batch too large");
    drop_graphs(e);
    int rc;
    // hit-list workgroups of a batch: slots of one box each (a box's beads are the slots
    // [bead_off, bead_off + n) since the sort keys are ordered by box)
    if (e->boxes.size() > 1) {
        for (int tb = 0; tb < kSchedules; ++tb) {
            if (kSchedule[tb].kernel != DPD_PAIR_HITLIST) continue;
            const int64_t th = kSchedule[tb].beads();
            std::vector<int4> hb;
            for (size_t b = 0; b < e->boxes.size(); ++b) {
                const int64_t s0 = e->boxes[b].bead_off, s1 = s0 + e->boxes[b].n;
                for (int64_t s = s0; s < s1; s += th)
                    hb.push_back(make_int4((int)b, (int)s, (int)std::min<int64_t>(s + th, s1), 0));
            }
            if ((rc = dalloc(e, &e->hitblk[tb], hb.size()))) return rc;
            HIPCHK(hipMemcpyAsync(e->hitblk[tb], hb.data(), sizeof(int4) * hb.size(), hipMemcpyHostToDevice,
                                  e->stream));
            HIPCHK(hipStreamSynchronize(e->stream));
            e->nhitblk[tb] = (int)hb.size();
        }
    }
    if (nbr != e->ngroups) {
        e->ngroups = nbr;
        e->nkeys = nbr * kGroupKeys;
        if ((rc = dalloc(e, &e->cnt, e->nkeys)) || (rc = dalloc(e, &e->cstart, e->nkeys + 1)) ||
            (rc = dalloc(e, &e->gcnt, e->ngroups)) || (rc = dalloc(e, &e->gstart, e->ngroups)))
            return rc;
        HIPCHK(hipMemsetAsync(e->cnt, 0, sizeof(int) * e->nkeys, e->stream));
        HIPCHK(hipMemsetAsync(e->gcnt, 0, sizeof(int) * e->ngroups, e->stream));
    }
    if (n != e->n) {
        for (int b = 0; b < 2; ++b)
            if ((rc = dalloc(e, &e->pos[b], n)) || (rc = dalloc(e, &e->vel[b], n)) ||
                (rc = dalloc(e, &e->fold[b], n)) || (rc = dalloc(e, &e->bead[b], n)) ||
                (rc = dalloc(e, &e->key[b], n)) || (rc = dalloc(e, &e->code[b], n + kCodePad)))
                return rc;
        if ((rc = dalloc(e, &e->img, n))) return rc;
        for (int b = 0; b < 2; ++b) HIPCHK(hipMemsetAsync(e->code[b] + n, 0, sizeof(uint32_t) * kCodePad, e->stream));
        if ((rc = dalloc(e, &e->frc, n)) || (rc = dalloc(e, &e->vp, n)) || (rc = dalloc(e, &e->slot_of, n)) ||
            (rc = dalloc(e, &e->nkey, n)) || (rc = dalloc(e, &e->rank, n)) || (rc = dalloc(e, &e->place, n)) ||
            (rc = dalloc(e, &e->gather, 12 * n)))
            return rc;
        e->n = n;
    }
    if ((rc = upload_boxes(e))) return rc;
    std::vector<float4> hp(n), hv(n), hf(n);
    std::vector<int> hb(n);
    std::vector<int4> hi(n);
    std::vector<int> hk(n);
    for (size_t b = 0; b < e->boxes.size(); ++b) {
        const BoxHost& h = e->boxes[b];
        const BoxDev& B = h.B;
        const double L[3] = {B.n[0] * B.wd[0], B.n[1] * B.wd[1], B.n[2] * B.wd[2]};
        for (int64_t i = 0; i < h.n; ++i) {
            const int64_t g = h.bead_off + i;
            int c[3], im[3];
            float o[3];
            for (int d = 0; d < 3; ++d) {
                const double xd = h.x[3 * i + d];
                int ci = (int)std::floor(xd / B.wd[d]);
                ci = std::min(std::max(ci, 0), B.n[d] - 1);
                float of = (float)(xd - ci * B.wd[d]);
                if (of >= B.w[d]) of = B.w[d] * 0.99999994f;
                if (of < 0.f) of = 0.f;
                c[d] = ci;
                o[d] = of;
                im[d] = (int)std::lround((h.u[3 * i + d] - xd) / L[d]);
            }
            const int32_t w = h.type[i] | ((int32_t)b << 8);
            float wb, ib;
            std::memcpy(&wb, &w, 4);
            std::memcpy(&ib, &h.id[i], 4);
            hp[g] = make_float4(o[0], o[1], o[2], wb);
            hv[g] = make_float4((float)h.v[3 * i], (float)h.v[3 * i + 1], (float)h.v[3 * i + 2], ib);
            hf[g] = h.has_f ? make_float4((float)h.f[3 * i], (float)h.f[3 * i + 1], (float)h.f[3 * i + 2], 0.f)
                            : make_float4(0.f, 0.f, 0.f, 0.f);
            hb[g] = (int)g;
            hi[g] = make_int4(im[0], im[1], im[2], 0);
            hk[g] = B.koff + cell_key(c[0], c[1], c[2], B);
        }
    }
    const int src = e->cur;
    HIPCHK(hipMemcpyAsync(e->pos[src], hp.data(), sizeof(float4) * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->vel[src], hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->fold[src], hf.data(), sizeof(float4) * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->bead[src], hb.data(), sizeof(int) * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->img, hi.data(), sizeof(int4) * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->key[src], hk.data(), sizeof(int) * n, hipMemcpyHostToDevice, e->stream));
    launch_zero(e, true);
    if ((rc = sort_step<kRebin>(e))) return rc;
    hipLaunchKernelGGL(k_copy_force, dim3(grid(n, 256)), dim3(256), 0, e->stream, (int)n, e->fold[e->cur], e->frc);
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, e->stream, e->dstep, step0);
    HIPCHK(hipGetLastError());
    e->step = step0;
    e->jumps = 0;
    e->obs_valid = false;
    e->committed = true;
    if (e->pair_auto) {
        set_schedule(e, default_schedule(e));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    return read_flags(e);
}

// Global CSR topology from the staged per-box bonds and bond pairs.
int commit_topology(dpd_engine* e)
{
    const int64_t n = e->n;
    std::vector<int> bp(n + 1, 0), ap(n + 1, 0);
    std::vector<int> H, Tl, F, Sd;
    std::vector<double> K, L0, M, P;
    for (const BoxHost& h : e->boxes) {
        const int off = (int)h.bead_off;
        const int boff = (int)H.size();
        for (size_t b = 0; b < h.bh.size(); ++b) {
            H.push_back(h.bh[b] + off); Tl.push_back(h.bt[b] + off);
            K.push_back(h.bk[b]); L0.push_back(h.bl0[b]);
        }
        for (size_t q = 0; q < h.pf.size(); ++q) {
            F.push_back(h.pf[q] + boff); Sd.push_back(h.ps[q] + boff);
            M.push_back(h.pk[q]); P.push_back(h.pphi[q]);
        }
    }
    const int64_t nbonds = (int64_t)H.size(), npairs = (int64_t)F.size();
    for (int64_t b = 0; b < nbonds; ++b) { bp[H[b] + 1]++; bp[Tl[b] + 1]++; }
    std::vector<int4> abead(std::max<int64_t>(npairs, 1));
    std::vector<double4> apar(std::max<int64_t>(npairs, 1));
    for (int64_t q = 0; q < npairs; ++q) {
        const int a = H[Sd[q]], bm = H[F[q]], c = Tl[F[q]];
        abead[q] = make_int4(a, bm, c, 0);
        apar[q] = make_double4(M[q], P[q], std::cos(P[q]), std::sin(P[q]));
        ap[a + 1]++; ap[bm + 1]++; ap[c + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) { bp[i + 1] += bp[i]; ap[i + 1] += ap[i]; }
    // each bead's entries in bond order (then angle order): the bonded kernels write every
    // bond's and angle's forces into its beads' entries, and a bead's sum reads its entries
    std::vector<int2> bcsr(std::max<int64_t>(nbonds, 1)), bbead(std::max<int64_t>(nbonds, 1));
    std::vector<double2> bpar(std::max<int64_t>(nbonds, 1));
    std::vector<int4> acsr(std::max<int64_t>(npairs, 1));
    std::vector<int> fill(bp.begin(), bp.end() - 1);
    for (int64_t b = 0; b < nbonds; ++b) {
        const int h = H[b], t = Tl[b];
        bbead[b] = make_int2(h, t);
        bpar[b] = make_double2(K[b], L0[b]);
        const int eh = fill[h]++;
        const int et = fill[t]++;
        bcsr[b] = make_int2(eh, et);
    }
    std::vector<int> bonded;
    for (int64_t i = 0; i < n; ++i)
        if (bp[i + 1] > bp[i] || ap[i + 1] > ap[i]) bonded.push_back((int)i);
    std::vector<int> afill(ap.begin(), ap.end() - 1);
    for (int64_t q = 0; q < npairs; ++q) {
        const int4 abc = abead[q];
        const int ec = afill[abc.z]++;
        const int eb = afill[abc.y]++;
        const int ea = afill[abc.x]++;
        acsr[q] = make_int4(ec, eb, ea, 0);
    }
    drop_graphs(e);
    int rc;
    if ((rc = dalloc(e, &e->bond_ptr, n + 1)) || (rc = dalloc(e, &e->ang_ptr, n + 1)) ||
        (rc = dalloc(e, &e->bond_csr, bcsr.size())) || (rc = dalloc(e, &e->bond_par, bpar.size())) ||
        (rc = dalloc(e, &e->bond_bead, bbead.size())) ||
        (rc = dalloc(e, &e->fbond, std::max<int64_t>(bp[n], 1))) ||
        (rc = dalloc(e, &e->fang, std::max<int64_t>(ap[n], 1))) || (rc = dalloc(e, &e->bonded, bonded.size())) ||
        (rc = dalloc(e, &e->ang_csr, acsr.size())) || (rc = dalloc(e, &e->ang_bead, abead.size())) ||
        (rc = dalloc(e, &e->ang_par, apar.size())))
        return rc;
    HIPCHK(hipMemcpyAsync(e->bond_ptr, bp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->ang_ptr, ap.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->bond_csr, bcsr.data(), sizeof(int2) * bcsr.size(), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->bond_bead, bbead.data(), sizeof(int2) * bbead.size(), hipMemcpyHostToDevice, e->stream));
    if (!bonded.empty())
        HIPCHK(hipMemcpyAsync(e->bonded, bonded.data(), sizeof(int) * bonded.size(), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->bond_par, bpar.data(), sizeof(double2) * bpar.size(), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->ang_csr, acsr.data(), sizeof(int4) * acsr.size(), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->ang_bead, abead.data(), sizeof(int4) * abead.size(), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->ang_par, apar.data(), sizeof(double4) * apar.size(), hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->nbonds = (int)nbonds;
    e->nangles = (int)npairs;
    e->nbonded = (int)bonded.size();
    e->has_bonds = nbonds > 0;
    e->has_angles = npairs > 0;
    hipLaunchKernelGGL(k_clear_bonded_flag, dim3(grid(n, 256)), dim3(256), 0, e->stream, (int)n, e->pos[e->cur],
                       e->bead[e->cur], e->slot_of);
    if (e->nbonded > 0)
        hipLaunchKernelGGL(k_set_bonded_flag, dim3(grid(e->nbonded, 256)), dim3(256), 0, e->stream, e->nbonded,
                           e->bonded, e->slot_of, e->pos[e->cur]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    return DPD_OK;
}

int stage_beads(dpd_engine* e, int32_t box, int64_t n, const int32_t* id, const int32_t* type, const double* x,
                const double* unpbc, const double* v, const double* f)
{
    if (box < 0 || box >= (int32_t)e->boxes.size()) return fail(e, DPD_ERR_PARAM, "box index out of range");
    if (n < 1 || n > ((int64_t)1 << 30)) return fail(e, DPD_ERR_PARAM, "bead count out of range");
    if (!id || !type || !x || !v) return fail(e, DPD_ERR_PARAM, "null bead array");
    BoxHost& h = e->boxes[box];
    const BoxDev& B = h.B;
    const double L[3] = {B.n[0] * B.wd[0], B.n[1] * B.wd[1], B.n[2] * B.wd[2]};
    for (int64_t i = 0; i < n; ++i) {
        if (type[i] < 0 || type[i] >= h.p.ntypes) return fail(e, DPD_ERR_PARAM, "bead type out of range");
        for (int d = 0; d < 3; ++d)
            if (!(x[3 * i + d] >= 0.0 && x[3 * i + d] < L[d]))
                return fail(e, DPD_ERR_OUT_OF_BOX, "bead " + std::to_string(i) + " outside the box");
    }
    if (h.n != n) { h.bh.clear(); h.bt.clear(); h.pf.clear(); h.ps.clear(); }   // topology refers to indices
    h.n = n;
    h.id.assign(id, id + n);
    h.type.assign(type, type + n);
    h.x.assign(x, x + 3 * n);
    h.u.assign(unpbc ? unpbc : x, (unpbc ? unpbc : x) + 3 * n);
    h.v.assign(v, v + 3 * n);
    h.has_f = f != nullptr;
    h.f.assign(f ? f : x, (f ? f : x) + (f ? 3 * n : 0));
    return DPD_OK;
}

int stage_topology(dpd_engine* e, int32_t box, int64_t nbonds, const int32_t* head, const int32_t* tail,
                   const double* k, const double* l0, int64_t npairs, const int32_t* first, const int32_t* second,
                   const double* modulus, const double* phi0)
{
    if (box < 0 || box >= (int32_t)e->boxes.size()) return fail(e, DPD_ERR_PARAM, "box index out of range");
    BoxHost& h = e->boxes[box];
    if (h.n == 0) return fail(e, DPD_ERR_STATE, "upload beads before topology");
    if (nbonds < 0 || npairs < 0) return fail(e, DPD_ERR_PARAM, "negative counts");
    if (nbonds && (!head || !tail || !k || !l0)) return fail(e, DPD_ERR_PARAM, "null bond array");
    if (npairs && (!first || !second || !modulus || !phi0)) return fail(e, DPD_ERR_PARAM, "null pair array");
    for (int64_t b = 0; b < nbonds; ++b)
        if (head[b] < 0 || head[b] >= h.n || tail[b] < 0 || tail[b] >= h.n || head[b] == tail[b])
            return fail(e, DPD_ERR_PARAM, "bond bead index out of range");
    for (int64_t q = 0; q < npairs; ++q) {
        if (first[q] < 0 || first[q] >= nbonds || second[q] < 0 || second[q] >= nbonds)
            return fail(e, DPD_ERR_PARAM, "bond pair index out of range");
        if (tail[second[q]] != head[first[q]]) return fail(e, DPD_ERR_PARAM, "bond pair bonds do not share a bead");
    }
    h.bh.assign(head, head + nbonds); h.bt.assign(tail, tail + nbonds);
    h.bk.assign(k, k + nbonds); h.bl0.assign(l0, l0 + nbonds);
    h.pf.assign(first, first + npairs); h.ps.assign(second, second + npairs);
    h.pk.assign(modulus, modulus + npairs); h.pphi.assign(phi0, phi0 + npairs);
    return DPD_OK;
}

}  // namespace

// Static scratch of dpd_slice_stress_noalloc: the pair and bond slice tables of one call,
// held by every device the library runs on until the process ends.  Its host-side name is
// a stand-in, not a device address: scratch_ptr looks up each device's own copy once.
constexpr size_t kSliceScratchDoubles = (size_t)1 << 19;   // 4 MiB
static __device__ __align__(256) double scratch[kSliceScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// the device address of this engine's device's copy of `scratch` (looked up on the first
// call on that device, with that device current)
int scratch_ptr(dpd_engine* e, double** out)
{
    if (e->device < 0 || e->device >= kMaxDevices) return fail(e, DPD_ERR_PARAM, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[e->device]) {
        HIPCHK(hipSetDevice(e->device));
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(e, DPD_ERR_HIP, "hipGetSymbolAddress returned no address for the slice scratch");
        g_scratch[e->device] = (double*)p;
    }
    *out = g_scratch[e->device];
    return DPD_OK;
}

int slice_check(dpd_engine* e, int32_t box, int32_t nslices, size_t* ntab)
{
    if (!e->committed) return fail(e, DPD_ERR_STATE, "no beads uploaded");
    if (box < 0 || box >= (int32_t)e->boxes.size()) return fail(e, DPD_ERR_PARAM, "box index out of range");
    if (nslices < 1 || nslices > 65536) return fail(e, DPD_ERR_PARAM, "nslices must be in [1, 65536]");
    HIPCHK(hipSetDevice(e->device));
    const int nt = e->boxes[box].B.ntypes;
    *ntab = (size_t)nslices * nt * nt * 9;
    return DPD_OK;
}

// pair and bond tables of box `box` into the cleared device tables tab[0 .. 2 ntab), then
// to the host outputs (either may be null)
int slice_tables(dpd_engine* e, int32_t box, int32_t nslices, double* tab, size_t ntab, double* pair_out,
                 double* bond_out)
{
    const BoxHost& h = e->boxes[box];
    const int nt = h.B.ntypes;
    SliceArgs a{};
    a.B = h.B;
    a.pos = e->pos[e->cur];
    a.vp = e->vp;
    a.img = e->img;
    a.key = e->key[e->cur];
    a.cstart = e->cstart;
    a.table = e->table + (size_t)box * kTableStride;
    a.step = e->dstep;
    a.slot_of = e->slot_of;
    a.bond_bead = e->bond_bead;
    a.bond_par = e->bond_par;
    a.b0 = 0;
    for (int32_t b = 0; b < box; ++b) a.b0 += (int)e->boxes[b].bh.size();
    a.b1 = a.b0 + (int)h.bh.size();
    a.ang_bead = e->ang_bead;
    a.ang_par = e->ang_par;
    a.a0 = a.a1 = 0;
    a.nslice = nslices;
    a.ntypes = nt;
    a.nidx = nt * nt;
    HIPCHK(launch_slice_stress(a, pair_out ? tab : nullptr, bond_out ? tab + ntab : nullptr, e->stream));
    if (pair_out) HIPCHK(hipMemcpyAsync(pair_out, tab, sizeof(double) * ntab, hipMemcpyDeviceToHost, e->stream));
    if (bond_out) HIPCHK(hipMemcpyAsync(bond_out, tab + ntab, sizeof(double) * ntab, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DPD_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI ------

extern "C" {

const char* dpd_last_error(const dpd_engine* e) { return e ? e->err.c_str() : "null engine"; }

int dpd_create_batch(int32_t nbox, const dpd_params* p, dpd_engine** out)
{
    if (!out) return DPD_ERR_PARAM;
    *out = nullptr;
    if (!p || nbox < 1 || nbox > (1 << 22)) return DPD_ERR_PARAM;
    dpd_engine* e = new dpd_engine();
    *out = e;
    int rc;
    e->boxes.resize(nbox);
    for (int b = 0; b < nbox; ++b) {
        if ((rc = check_params(e, p + b))) return rc;
        BoxHost& h = e->boxes[b];
        h.p = p[b];
        const int nt = p[b].ntypes;
        h.a.assign(p[b].a, p[b].a + nt * nt);
        h.g.assign(p[b].gamma, p[b].gamma + nt * nt);
        h.p.a = h.a.data();
        h.p.gamma = h.g.data();
        h.dt = p[b].dt; h.lambda = p[b].lambda; h.kT = p[b].kT;
        BoxDev& B = h.B;
        B.n[0] = p[b].nx; B.n[1] = p[b].ny; B.n[2] = p[b].nz;
        for (int d = 0; d < 2; ++d) B.inv[d] = 1.0f / (float)B.n[d];
        B.wd[0] = p[b].wx; B.wd[1] = p[b].wy; B.wd[2] = p[b].wz;
        for (int d = 0; d < 3; ++d) B.w[d] = (float)B.wd[d];
        B.ntypes = nt;
        B.seed0 = (uint32_t)p[b].seed;
        B.seed1 = (uint32_t)(p[b].seed >> 32);
        B.ngroups = (int)(((int64_t)B.n[0] * B.n[1] * B.n[2] + kGroupKeys - 1) / kGroupKeys);
        B.box = b;
        set_step_constants(h);
    }
    e->device = p[0].device;
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    for (auto& ev : e->ev) HIPCHK(hipEventCreate(&ev));
    if ((rc = dpd_set_pair_kernel(e, p[0].pair_kernel))) return rc;
    if ((rc = dalloc(e, &e->flags, 4)) || (rc = dalloc(e, &e->dstep, 1)) ||
        (rc = dalloc(e, &e->acc, (size_t)nbox * kObsStride)) || (rc = dalloc(e, &e->table, (size_t)nbox * kTableStride)) ||
        (rc = dalloc(e, &e->dboxes, nbox)))
        return rc;
    HIPCHK(hipMemsetAsync(e->flags, 0, sizeof(int) * 4, e->stream));
    HIPCHK(hipMemsetAsync(e->dstep, 0, sizeof(int64_t), e->stream));
    if ((rc = upload_tables(e))) return rc;
    e->err.clear();
    return DPD_OK;
}

int dpd_create(const dpd_params* p, dpd_engine** out) { return dpd_create_batch(1, p, out); }

void dpd_destroy(dpd_engine* e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    void* ptrs[] = {e->pos[0], e->pos[1], e->vel[0], e->vel[1], e->fold[0], e->fold[1], e->bead[0], e->bead[1], e->img,
                    e->key[0], e->key[1], e->code[0], e->code[1], e->frc, e->vp, e->slot_of, e->nkey, e->rank,
                    e->place, e->cnt, e->gcnt,
                    e->gstart, e->cstart, e->flags, e->dstep, e->acc, e->table, e->gather, e->bond_ptr,
                    e->ang_ptr, e->bond_csr, e->bond_par, e->ang_csr, e->ang_bead, e->ang_par, e->bond_bead,
                    e->fbond, e->fang, e->bonded, e->dboxes, e->slice, e->madj};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    for (void* q : e->hitblk)
        if (q) (void)hipFree(q);
    for (auto& ev : e->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : e->pev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int dpd_upload_beads(dpd_engine* e, int64_t n, const int32_t* id, const int32_t* type, const double* x,
                     const double* unpbc, const double* v, const double* f, int64_t step0)
{
    if (!e) return DPD_ERR_PARAM;
    if (e->boxes.size() != 1) return fail(e, DPD_ERR_STATE, "batch engine: use dpd_upload_box and dpd_commit");
    HIPCHK(hipSetDevice(e->device));
    const bool keep_topo = !e->boxes[0].bh.empty() && e->boxes[0].n == n;
    int rc;
    if ((rc = stage_beads(e, 0, n, id, type, x, unpbc, v, f))) return rc;
    if ((rc = commit_beads(e, step0))) return rc;
    if (keep_topo) return commit_topology(e);
    e->has_bonds = e->has_angles = 0;
    e->nbonds = e->nangles = e->nbonded = 0;
    return DPD_OK;
}

int dpd_upload_topology(dpd_engine* e, int64_t nbonds, const int32_t* head, const int32_t* tail, const double* k,
                        const double* l0, int64_t npairs, const int32_t* first_bond, const int32_t* second_bond,
                        const double* modulus, const double* phi0)
{
    if (!e) return DPD_ERR_PARAM;
    if (e->boxes.size() != 1) return fail(e, DPD_ERR_STATE, "batch engine: use dpd_upload_box_topology");
    if (!e->committed) return fail(e, DPD_ERR_STATE, "upload beads before topology");
    HIPCHK(hipSetDevice(e->device));
    int rc;
    if ((rc = stage_topology(e, 0, nbonds, head, tail, k, l0, npairs, first_bond, second_bond, modulus, phi0)))
        return rc;
    return commit_topology(e);
}

int dpd_upload_box(dpd_engine* e, int32_t box, int64_t n, const int32_t* id, const int32_t* type, const double* x,
                   const double* unpbc, const double* v, const double* f)
{
    if (!e) return DPD_ERR_PARAM;
    return stage_beads(e, box, n, id, type, x, unpbc, v, f);
}

int dpd_upload_box_topology(dpd_engine* e, int32_t box, int64_t nbonds, const int32_t* head, const int32_t* tail,
                            const double* k, const double* l0, int64_t npairs, const int32_t* first_bond,
                            const int32_t* second_bond, const double* modulus, const double* phi0)
{
    if (!e) return DPD_ERR_PARAM;
    return stage_topology(e, box, nbonds, head, tail, k, l0, npairs, first_bond, second_bond, modulus, phi0);
}

int dpd_commit(dpd_engine* e, int64_t step0)
{
    if (!e) return DPD_ERR_PARAM;
    HIPCHK(hipSetDevice(e->device));
    int rc;
    if ((rc = commit_beads(e, step0))) return rc;
    return commit_topology(e);
}

int32_t dpd_box_count(const dpd_engine* e) { return e ? (int32_t)e->boxes.size() : -1; }

int dpd_step(dpd_engine* e, int64_t nsteps, uint32_t flags)
{
    if (!e) return DPD_ERR_PARAM;
    if (!e->committed) return fail(e, DPD_ERR_STATE, "no beads uploaded");
    if (nsteps < 0) return fail(e, DPD_ERR_PARAM, "negative step count");
    if (nsteps == 0) return DPD_OK;
    HIPCHK(hipSetDevice(e->device));
    const bool obs = flags & DPD_STEP_OBSERVE;
    const bool timing = flags & DPD_STEP_TIMING;
    const int timing_mode = timing ? ((flags & DPD_STEP_TIMING_SAMPLED) ? 2 : 1) : 0;
    const bool use_graph = (flags & DPD_STEP_GRAPH) && !timing;
    int rc;
    e->jumps = 0;
    if (timing) {
        while ((int64_t)e->pev.size() < 2 * timed_steps(nsteps, timing_mode)) {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            e->pev.push_back(ev);
        }
    }
    launch_zero(e, true);
    HIPCHK(hipEventRecord(e->ev[0], e->stream));
    if (use_graph) {
        // whole calls of an even step count replay one captured graph: the two SoA buffers
        // swap once per step, so an even count returns to the captured pointers
        int64_t chunk = std::min<int64_t>(nsteps, 64);
        if (chunk & 1) chunk -= 1;
        int64_t done = 0;
        while (chunk >= 2 && nsteps - done >= chunk) {
            const bool last_obs = obs && (nsteps - done == chunk);
            const auto keyg = std::make_pair(chunk, (int)last_obs);
            auto it = e->graphs.find(keyg);
            if (it == e->graphs.end()) {
                const int cur0 = e->cur;
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
                rc = enqueue_steps(e, chunk, last_obs, 0);
                hipError_t st = hipStreamEndCapture(e->stream, &g);
                if (rc) return rc;
                if (st != hipSuccess) return fail(e, DPD_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(st));
                e->cur = cur0;
                hipGraphExec_t ge;
                HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
                e->graphs[keyg] = ge;
                e->graph_src[keyg] = g;
                it = e->graphs.find(keyg);
            }
            HIPCHK(hipGraphLaunch(it->second, e->stream));
            done += chunk;
            e->step += chunk;
        }
        if (done < nsteps) {
            if ((rc = enqueue_steps(e, nsteps - done, obs, 0))) return rc;
            e->step += nsteps - done;
        }
    } else {
        if ((rc = enqueue_steps(e, nsteps, obs, timing_mode))) return rc;
        e->step += nsteps;
    }
    HIPCHK(hipEventRecord(e->ev[1], e->stream));
    if ((rc = read_flags(e))) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e->ev[0], e->ev[1]));
    e->t_total = ms;
    e->t_pair = 0.0;
    e->n_pair = 0;
    if (timing) {
        double tp = 0.0;
        int64_t np = 0;
        for (; np < timed_steps(nsteps, timing_mode); ++np) {
            HIPCHK(hipEventElapsedTime(&ms, e->pev[2 * np], e->pev[2 * np + 1]));
            tp += ms;
        }
        e->t_pair = tp;
        e->n_pair = np;
    }
    if (obs) return finish_obs(e);
    e->obs_valid = false;
    return DPD_OK;
}

int dpd_compute_forces(dpd_engine* e, int64_t step, uint32_t flags, double* f_out, double* virial, double* pe)
{
    if (!e) return DPD_ERR_PARAM;
    if (!e->committed) return fail(e, DPD_ERR_STATE, "no beads uploaded");
    HIPCHK(hipSetDevice(e->device));
    int rc;
    std::vector<bool> saved;
    if (flags & DPD_NO_THERMOSTAT) {
        for (auto& b : e->boxes) { saved.push_back(b.thermostat_on); b.thermostat_on = false; }
        if ((rc = upload_tables(e))) return rc;
    }
    launch_zero(e, true);
    // the force the next dpd_step starts from is kept aside (the inactive SoA buffers are
    // scratch between calls) and restored below, so a diagnostic evaluation does not change
    // the trajectory
    const size_t f4 = sizeof(float4) * (size_t)e->n;
    float4* keep = e->fold[e->cur ^ 1];
    HIPCHK(hipMemcpyAsync(keep, e->frc, f4, hipMemcpyDeviceToDevice, e->stream));
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, e->stream, e->dstep, step);
    // the uploaded state is sorted with vp = v: evaluate at it (after dpd_step, vp is the
    // last step's predicted velocity)
    bool terms_done = false;
    if ((rc = launch_pair(e, true, &terms_done))) return rc;
    launch_bonded<true>(e, terms_done);
    launch_finish<false, true>(e);
    HIPCHK(hipGetLastError());
    if ((rc = read_flags(e))) return rc;
    std::vector<double> h(e->boxes.size() * kObsStride);
    HIPCHK(hipMemcpyAsync(h.data(), e->acc, sizeof(double) * h.size(), hipMemcpyDeviceToHost, e->stream));
    const int n = (int)e->n;
    if (f_out) {
        double* g = e->gather;
        if (multi(e))
            hipLaunchKernelGGL(k_gather<true>, dim3(grid(n, 256)), dim3(256), 0, e->stream, n, bufs(e, e->cur), e->frc,
                               e->dboxes, e->boxes[0].B, g, g + 3 * n, g + 6 * n, g + 9 * n);
        else
            hipLaunchKernelGGL(k_gather<false>, dim3(grid(n, 256)), dim3(256), 0, e->stream, n, bufs(e, e->cur),
                               e->frc, e->dboxes, e->boxes[0].B, g, g + 3 * n, g + 6 * n, g + 9 * n);
        HIPCHK(hipMemcpyAsync(f_out, g + 9 * n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, e->stream));
    }
    HIPCHK(hipMemcpyAsync(e->frc, keep, f4, hipMemcpyDeviceToDevice, e->stream));
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, e->stream, e->dstep, e->step);
    HIPCHK(hipStreamSynchronize(e->stream));
    // totals over all boxes
    if (virial)
        for (int i = 0; i < 9; ++i) {
            virial[i] = 0.0;
            for (size_t b = 0; b < e->boxes.size(); ++b) virial[i] += h[b * kObsStride + O_VIR + i];
        }
    if (pe)
        for (int i = 0; i < 3; ++i) {
            pe[i] = 0.0;
            for (size_t b = 0; b < e->boxes.size(); ++b) pe[i] += h[b * kObsStride + O_PEP + i];
        }
    if (flags & DPD_NO_THERMOSTAT) {
        for (size_t b = 0; b < e->boxes.size(); ++b) e->boxes[b].thermostat_on = saved[b];
        if ((rc = upload_tables(e))) return rc;
    }
    return DPD_OK;
}

int dpd_observables_box(dpd_engine* e, int32_t box, dpd_obs* out)
{
    if (!e || !out) return DPD_ERR_PARAM;
    if (!e->obs_valid) return fail(e, DPD_ERR_STATE, "no observables: call dpd_step with DPD_STEP_OBSERVE");
    if (box < 0 || box >= (int32_t)e->obs.size()) return fail(e, DPD_ERR_PARAM, "box index out of range");
    *out = e->obs[box];
    return DPD_OK;
}

int dpd_observables(dpd_engine* e, dpd_obs* out) { return dpd_observables_box(e, 0, out); }

int dpd_download_box(dpd_engine* e, int32_t box, double* x, double* unpbc, double* v, double* f)
{
    if (!e) return DPD_ERR_PARAM;
    if (!e->committed) return fail(e, DPD_ERR_STATE, "no beads uploaded");
    if (box < -1 || box >= (int32_t)e->boxes.size()) return fail(e, DPD_ERR_PARAM, "box index out of range");
    HIPCHK(hipSetDevice(e->device));
    const int n = (int)e->n;
    double* g = e->gather;
    if (multi(e))
        hipLaunchKernelGGL(k_gather<true>, dim3(grid(n, 256)), dim3(256), 0, e->stream, n, bufs(e, e->cur), e->frc,
                           e->dboxes, e->boxes[0].B, g, g + 3 * n, g + 6 * n, g + 9 * n);
    else
        hipLaunchKernelGGL(k_gather<false>, dim3(grid(n, 256)), dim3(256), 0, e->stream, n, bufs(e, e->cur), e->frc,
                           e->dboxes, e->boxes[0].B, g, g + 3 * n, g + 6 * n, g + 9 * n);
    HIPCHK(hipGetLastError());
    const int64_t off = box < 0 ? 0 : e->boxes[box].bead_off;
    const int64_t cnt = box < 0 ? e->n : e->boxes[box].n;
    const size_t b3 = sizeof(double) * 3 * cnt;
    if (x) HIPCHK(hipMemcpyAsync(x, g + 3 * off, b3, hipMemcpyDeviceToHost, e->stream));
    if (unpbc) HIPCHK(hipMemcpyAsync(unpbc, g + 3 * n + 3 * off, b3, hipMemcpyDeviceToHost, e->stream));
    if (v) HIPCHK(hipMemcpyAsync(v, g + 6 * n + 3 * off, b3, hipMemcpyDeviceToHost, e->stream));
    if (f) HIPCHK(hipMemcpyAsync(f, g + 9 * n + 3 * off, b3, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DPD_OK;
}

int dpd_download(dpd_engine* e, double* x, double* unpbc, double* v, double* f)
{
    return dpd_download_box(e, -1, x, unpbc, v, f);
}

int dpd_set_box_interactions(dpd_engine* e, int32_t box, const double* a, const double* gamma)
{
    if (!e) return DPD_ERR_PARAM;
    const int32_t b0 = box < 0 ? 0 : box, b1 = box < 0 ? (int32_t)e->boxes.size() : box + 1;
    if (b0 >= (int32_t)e->boxes.size() || b1 > (int32_t)e->boxes.size())
        return fail(e, DPD_ERR_PARAM, "box index out of range");
    for (int32_t b = b0; b < b1; ++b) {
        BoxHost& h = e->boxes[b];
        const int nt = h.p.ntypes;
        if (a) h.a.assign(a, a + nt * nt);
        if (gamma) h.g.assign(gamma, gamma + nt * nt);
    }
    return upload_tables(e);
}

int dpd_set_interactions(dpd_engine* e, const double* a, const double* gamma)
{
    return dpd_set_box_interactions(e, -1, a, gamma);
}

int dpd_set_timestep(dpd_engine* e, double dt)
{
    if (!e) return DPD_ERR_PARAM;
    if (!(dt > 0.0)) return fail(e, DPD_ERR_PARAM, "dt must be positive");
    for (auto& b : e->boxes) { b.dt = dt; set_step_constants(b); }
    drop_graphs(e);
    return e->committed ? upload_boxes(e) : DPD_OK;
}

int dpd_toggle(dpd_engine* e, int32_t what, int32_t on)
{
    if (!e) return DPD_ERR_PARAM;
    if (what == DPD_TOGGLE_FORCES) {
        // the step constants are kernel arguments of captured graphs: drop them
        for (auto& b : e->boxes) { b.forces_on = on != 0; set_step_constants(b); }
        drop_graphs(e);
        return e->committed ? upload_boxes(e) : DPD_OK;
    }
    for (auto& b : e->boxes) {
        if (what == DPD_TOGGLE_THERMOSTAT) b.thermostat_on = on != 0;
        else if (what == DPD_TOGGLE_CONSERVATIVE) b.conservative_on = on != 0;
        else return fail(e, DPD_ERR_PARAM, "unknown toggle");
    }
    return upload_tables(e);
}

int64_t dpd_last_timing(dpd_engine* e, double* ms3)
{
    if (!e) return -1;
    if (ms3) { ms3[0] = e->t_total; ms3[1] = e->t_pair; ms3[2] = e->t_total - e->t_pair; }
    return e->n_pair;
}

int64_t dpd_get_step(const dpd_engine* e) { return e ? e->step : -1; }

int dpd_time_pair(dpd_engine* e, int32_t nlaunch, double* ms_per_launch)
{
    if (!e || !ms_per_launch) return DPD_ERR_PARAM;
    if (!e->committed) return fail(e, DPD_ERR_STATE, "no beads uploaded");
    if (nlaunch < 1 || nlaunch > 100000) return fail(e, DPD_ERR_PARAM, "nlaunch must be in [1, 100000]");
    HIPCHK(hipSetDevice(e->device));
    int rc;
    // the launches rewrite this step's pair force: the force the next step starts from is
    // kept in the inactive SoA buffer (scratch between calls) and restored afterwards
    const size_t f4 = sizeof(float4) * (size_t)e->n;
    float4* keep = e->fold[e->cur ^ 1];
    HIPCHK(hipMemcpyAsync(keep, e->frc, f4, hipMemcpyDeviceToDevice, e->stream));
    // the launch a step that another follows uses: with the epilogue where the schedule has
    // one, which also writes new cell keys and ranks, counts the counting sort's cells and
    // may raise the flags (restored below)
    const bool epi = epilogue_ok(e);
    int flags0[2] = {0, 0};
    if (epi) HIPCHK(hipMemcpyAsync(flags0, e->flags, sizeof(flags0), hipMemcpyDeviceToHost, e->stream));
    bool terms_done = false;
    if ((rc = launch_pair(e, false, &terms_done, epi))) return rc;   // untimed first launch
    HIPCHK(hipEventRecord(e->ev[0], e->stream));
    for (int32_t i = 0; i < nlaunch; ++i)
        if ((rc = launch_pair(e, false, &terms_done, epi))) return rc;
    HIPCHK(hipEventRecord(e->ev[1], e->stream));
    HIPCHK(hipMemcpyAsync(e->frc, keep, f4, hipMemcpyDeviceToDevice, e->stream));
    if (epi) {
        HIPCHK(hipMemcpyAsync(e->flags, flags0, sizeof(flags0), hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipMemsetAsync(e->cnt, 0, sizeof(int) * (size_t)e->nkeys, e->stream));
        HIPCHK(hipMemsetAsync(e->gcnt, 0, sizeof(int) * (size_t)e->ngroups, e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e->ev[0], e->ev[1]));
    *ms_per_launch = (double)ms / nlaunch;
    return DPD_OK;
}

int dpd_set_pair_kernel(dpd_engine* e, int32_t kernel)
{
    if (!e) return DPD_ERR_PARAM;
    if (kernel != DPD_PAIR_AUTO && (kernel < 1 || kernel > DPD_PAIR_KINDS))
        return fail(e, DPD_ERR_PARAM, "unknown pair kernel");
    if (kernel != DPD_PAIR_AUTO && !pair_kernel_available(kernel))
        return fail(e, DPD_ERR_UNSUPPORTED, "pair kernel " + std::to_string(kernel) + " is retired");
    e->pair_auto = kernel == DPD_PAIR_AUTO;
    if (e->pair_auto) set_schedule(e, e->committed ? default_schedule(e) : kFullShellSchedule);
    else if (kernel == DPD_PAIR_FULLSHELL) set_schedule(e, kFullShellSchedule);
    else if (kSchedule[e->sched].kernel != DPD_PAIR_HITLIST) set_schedule(e, kDefaultSchedule);
    return DPD_OK;
}

int32_t dpd_get_pair_kernel(const dpd_engine* e) { return e ? e->pair_kernel : -1; }

int32_t dpd_schedule_count(void) { return kSchedules; }

const char* dpd_schedule_name(int32_t sched)
{
    return (sched >= 0 && sched < kSchedules) ? kSchedule[sched].name : nullptr;
}

int dpd_set_schedule(dpd_engine* e, int32_t sched)
{
    if (!e) return DPD_ERR_PARAM;
    if (sched < 0 || sched >= kSchedules) return fail(e, DPD_ERR_PARAM, "unknown schedule");
    e->pair_auto = false;
    set_schedule(e, sched);
    return DPD_OK;
}

int32_t dpd_get_schedule(const dpd_engine* e) { return e ? e->sched : -1; }

int dpd_apply_tuning(dpd_engine* e, const char* path, int32_t* chosen)
{
    if (!e || !path) return DPD_ERR_PARAM;
    if (chosen) *chosen = -1;
    if (!e->committed) return fail(e, DPD_ERR_STATE, "no beads uploaded");
    FILE* fh = std::fopen(path, "r");
    if (!fh) return fail(e, DPD_ERR_IO, std::string("cannot open ") + path);
    // the engine's shape, as dpd_amd.tune keys it
    const int64_t nbox = (int64_t)e->boxes.size();
    const int64_t per = std::max<int64_t>(1, e->n / nbox);
    int64_t lo = 1;
    while (lo * 2 <= per) lo *= 2;
    const int64_t key[5] = {nbox == 1 ? 1 : 2, lo, 2 * lo, e->boxes[0].p.ntypes, (e->nbonds + e->nangles) > 0};
    const char* names[6] = {"\"nbox\":", "\"beads_lo\":", "\"beads_hi\":", "\"ntypes\":", "\"bonded\":", "\"schedule\":"};
    char line[4096];
    int found = -1;
    while (found < 0 && std::fgets(line, sizeof(line), fh)) {
        int64_t v[6];
        bool ok = true;
        for (int k = 0; k < 6 && ok; ++k) {
            const char* p = std::strstr(line, names[k]);
            if (!p) { ok = false; break; }
            v[k] = std::strtoll(p + std::strlen(names[k]), nullptr, 10);
        }
        if (!ok) continue;
        bool match = true;
        for (int k = 0; k < 5; ++k) match = match && v[k] == key[k];
        if (match && v[5] >= 0 && v[5] < kSchedules) found = (int)v[5];
    }
    std::fclose(fh);
    if (found >= 0) {
        if (kSchedule[found].kernel == DPD_PAIR_HITLIST && e->n > (int64_t)kHitSlotMask) found = kFullShellSchedule;
        e->pair_auto = false;
        set_schedule(e, found);
    }
    if (chosen) *chosen = found;
    return DPD_OK;
}

namespace {

// Time `nsteps` steps of each candidate schedule from the current state, restoring the
// state after every trial (the simulation does not advance); ms[k] gets the time per step of
// trial k.  The fastest is left set; *best is its schedule.
int tune_trials(dpd_engine* e, int32_t nsteps, const std::vector<int>& trials, double* ms, int* best)
{
    if (!e->committed) return fail(e, DPD_ERR_STATE, "no beads uploaded");
    if (nsteps < 1) return fail(e, DPD_ERR_PARAM, "nsteps must be positive");
    HIPCHK(hipSetDevice(e->device));
    const int64_t n = e->n;
    // snapshot of the state the trials start from
    float4 *sp = nullptr, *sv = nullptr, *sfo = nullptr, *sf = nullptr, *svp = nullptr;
    int4* sa = nullptr;
    int *sk = nullptr, *sso = nullptr, *scs = nullptr, *sbd = nullptr;
    uint32_t* scd = nullptr;
    int rc;
    if ((rc = dalloc(e, &sp, n)) || (rc = dalloc(e, &sv, n)) || (rc = dalloc(e, &sfo, n)) || (rc = dalloc(e, &sf, n)) ||
        (rc = dalloc(e, &svp, n)) || (rc = dalloc(e, &sa, n)) || (rc = dalloc(e, &sk, n)) ||
        (rc = dalloc(e, &sso, n)) || (rc = dalloc(e, &scs, e->nkeys + 1)) || (rc = dalloc(e, &scd, n)) ||
        (rc = dalloc(e, &sbd, n)))
        return rc;
    const int c0 = e->cur;
    const int64_t step0 = e->step;
    auto copy = [&](bool save) -> int {
        const hipMemcpyKind k = hipMemcpyDeviceToDevice;
        const size_t f4 = sizeof(float4) * n;
        if (save) {
            HIPCHK(hipMemcpyAsync(sp, e->pos[c0], f4, k, e->stream));
            HIPCHK(hipMemcpyAsync(sv, e->vel[c0], f4, k, e->stream));
            HIPCHK(hipMemcpyAsync(sfo, e->fold[c0], f4, k, e->stream));
            HIPCHK(hipMemcpyAsync(sf, e->frc, f4, k, e->stream));
            HIPCHK(hipMemcpyAsync(svp, e->vp, f4, k, e->stream));
            HIPCHK(hipMemcpyAsync(sa, e->img, sizeof(int4) * n, k, e->stream));
            HIPCHK(hipMemcpyAsync(sbd, e->bead[c0], sizeof(int) * n, k, e->stream));
            HIPCHK(hipMemcpyAsync(sk, e->key[c0], sizeof(int) * n, k, e->stream));
            HIPCHK(hipMemcpyAsync(scd, e->code[c0], sizeof(uint32_t) * n, k, e->stream));
            HIPCHK(hipMemcpyAsync(sso, e->slot_of, sizeof(int) * n, k, e->stream));
            HIPCHK(hipMemcpyAsync(scs, e->cstart, sizeof(int) * (e->nkeys + 1), k, e->stream));
        } else {
            HIPCHK(hipMemcpyAsync(e->pos[c0], sp, f4, k, e->stream));
            HIPCHK(hipMemcpyAsync(e->vel[c0], sv, f4, k, e->stream));
            HIPCHK(hipMemcpyAsync(e->fold[c0], sfo, f4, k, e->stream));
            HIPCHK(hipMemcpyAsync(e->frc, sf, f4, k, e->stream));
            HIPCHK(hipMemcpyAsync(e->vp, svp, f4, k, e->stream));
            HIPCHK(hipMemcpyAsync(e->img, sa, sizeof(int4) * n, k, e->stream));
            HIPCHK(hipMemcpyAsync(e->bead[c0], sbd, sizeof(int) * n, k, e->stream));
            HIPCHK(hipMemcpyAsync(e->key[c0], sk, sizeof(int) * n, k, e->stream));
            HIPCHK(hipMemcpyAsync(e->code[c0], scd, sizeof(uint32_t) * n, k, e->stream));
            HIPCHK(hipMemcpyAsync(e->slot_of, sso, sizeof(int) * n, k, e->stream));
            HIPCHK(hipMemcpyAsync(e->cstart, scs, sizeof(int) * (e->nkeys + 1), k, e->stream));
            hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, e->stream, e->dstep, step0);
            e->cur = c0;
            e->step = step0;
        }
        HIPCHK(hipStreamSynchronize(e->stream));
        return DPD_OK;
    };
    if ((rc = copy(true))) return rc;
    double bestms = 1e300;
    int bests = e->sched;
    for (size_t v = 0; v < trials.size(); ++v) {
        set_schedule(e, trials[v]);
        if ((rc = dpd_step(e, 1, 0))) break;          // warm-up (code objects, caches)
        if ((rc = copy(false))) break;
        if ((rc = dpd_step(e, nsteps, 0))) break;
        const double t = e->t_total / nsteps;
        if (ms) ms[v] = t;
        if (t < bestms) { bestms = t; bests = trials[v]; }
        if ((rc = copy(false))) break;
    }
    void* tmp[] = {sp, sv, sfo, sf, svp, sa, sk, sso, scs, scd, sbd};
    for (void* q : tmp) (void)hipFree(q);
    if (rc) return rc;
    e->obs_valid = false;
    e->pair_auto = false;   // the tuned choice stays across commits
    set_schedule(e, bests);
    if (best) *best = bests;
    return DPD_OK;
}

}  // namespace

int dpd_autotune(dpd_engine* e, int32_t nsteps, int32_t* chosen, double* ms_per_step)
{
    if (!e) return DPD_ERR_PARAM;
    const std::vector<int> trials = {kDefaultSchedule, kFullShellSchedule};
    double ms[2] = {0.0, 0.0};
    int best = 0;
    int rc = tune_trials(e, nsteps, trials, ms, &best);
    if (rc) return rc;
    if (ms_per_step) {
        for (int k = 0; k < DPD_PAIR_KINDS; ++k) ms_per_step[k] = 0.0;
        ms_per_step[kSchedule[kDefaultSchedule].kernel - 1] = ms[0];
        ms_per_step[kSchedule[kFullShellSchedule].kernel - 1] = ms[1];
    }
    if (chosen) *chosen = kSchedule[best].kernel;
    return DPD_OK;
}

int dpd_autotune_schedules(dpd_engine* e, int32_t nsteps, int32_t* chosen, double* ms_per_step)
{
    if (!e) return DPD_ERR_PARAM;
    std::vector<int> trials;
    for (int k = 0; k < kSchedules; ++k)
        if (kSchedule[k].kernel != DPD_PAIR_HITLIST || e->n <= (int64_t)kHitSlotMask) trials.push_back(k);
    std::vector<double> ms(trials.size(), 0.0);
    int best = 0;
    int rc = tune_trials(e, nsteps, trials, ms.data(), &best);
    if (rc) return rc;
    if (ms_per_step) {
        for (int k = 0; k < kSchedules; ++k) ms_per_step[k] = 0.0;
        for (size_t v = 0; v < trials.size(); ++v) ms_per_step[trials[v]] = ms[v];
    }
    if (chosen) *chosen = best;
    return DPD_OK;
}

namespace {

// shared body of dpd_zero_total_momentum / dpd_renormalise_momenta
int adjust_momenta(dpd_engine* e, int32_t box, bool renormalise)
{
    if (!e) return DPD_ERR_PARAM;
    if (!e->committed) return fail(e, DPD_ERR_STATE, "no beads uploaded");
    const int nbox = (int)e->boxes.size();
    if (box < -1 || box >= nbox) return fail(e, DPD_ERR_PARAM, "box index out of range");
    HIPCHK(hipSetDevice(e->device));
    const int n = (int)e->n;
    const int nacc = nbox * kObsStride;
    hipLaunchKernelGGL(k_zero, dim3(grid(std::max(nacc, 4), 256)), dim3(256), 0, e->stream, e->acc, nacc, e->flags, 0);
    if (multi(e)) hipLaunchKernelGGL(k_mom_sums<true>, dim3(grid(n, 256)), dim3(256), 0, e->stream, n, bufs(e, e->cur), e->acc);
    else hipLaunchKernelGGL(k_mom_sums<false>, dim3(grid(n, 256)), dim3(256), 0, e->stream, n, bufs(e, e->cur), e->acc);
    HIPCHK(hipGetLastError());
    std::vector<double> h(nacc);
    HIPCHK(hipMemcpyAsync(h.data(), e->acc, sizeof(double) * nacc, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    std::vector<double4> adj(nbox, make_double4(0.0, 0.0, 0.0, 1.0));
    for (int b = 0; b < nbox; ++b) {
        if (box >= 0 && b != box) continue;
        const double* a = h.data() + (size_t)b * kObsStride;
        const double cnt = a[4];
        if (!(cnt > 0.0)) continue;
        if (renormalise) {
            // CSimBox::UpdateRenormalisedMom (src/SimBox.cpp:3177-3215): v / sqrt(sum v^2 / (3 N kT))
            const double norm = std::sqrt(a[3] / (3.0 * cnt * e->boxes[b].kT));
            if (norm > 0.0) adj[b].w = 1.0 / norm;
        } else {
            // CSimBox::ZeroTotalMom (src/SimBox.cpp:1919-1947): v - sum v / N
            adj[b] = make_double4(a[0] / cnt, a[1] / cnt, a[2] / cnt, 1.0);
        }
    }
    int rc;
    if (!e->madj && (rc = dalloc(e, &e->madj, (size_t)nbox))) return rc;
    HIPCHK(hipMemcpyAsync(e->madj, adj.data(), sizeof(double4) * nbox, hipMemcpyHostToDevice, e->stream));
    if (multi(e)) hipLaunchKernelGGL(k_mom_apply<true>, dim3(grid(n, 256)), dim3(256), 0, e->stream, n, bufs(e, e->cur), e->madj);
    else hipLaunchKernelGGL(k_mom_apply<false>, dim3(grid(n, 256)), dim3(256), 0, e->stream, n, bufs(e, e->cur), e->madj);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    e->obs_valid = false;
    return DPD_OK;
}

}  // namespace

int dpd_zero_total_momentum(dpd_engine* e, int32_t box) { return adjust_momenta(e, box, false); }

int dpd_renormalise_momenta(dpd_engine* e, int32_t box) { return adjust_momenta(e, box, true); }

int dpd_slice_stress(dpd_engine* e, int32_t box, int32_t nslices, double* pair_out, double* bond_out)
{
    if (!e) return DPD_ERR_PARAM;
    int rc;
    size_t ntab = 0;
    if ((rc = slice_check(e, box, nslices, &ntab))) return rc;
    if (2 * ntab > e->slice_cap) {
        if ((rc = dalloc(e, &e->slice, 2 * ntab))) { e->slice_cap = 0; return rc; }
        e->slice_cap = 2 * ntab;
    }
    HIPCHK(hipMemsetAsync(e->slice, 0, sizeof(double) * 2 * ntab, e->stream));
    return slice_tables(e, box, nslices, e->slice, ntab, pair_out, bond_out);
}

int dpd_slice_stress_noalloc(dpd_engine* e, int32_t box, int32_t nslices, double* pair_out, double* bond_out)
{
    if (!e) return DPD_ERR_PARAM;
    int rc;
    size_t ntab = 0;
    if ((rc = slice_check(e, box, nslices, &ntab))) return rc;
    if (2 * ntab > kSliceScratchDoubles)
        return fail(e, DPD_ERR_PARAM, "slice tables of " + std::to_string(2 * ntab) + " doubles exceed the " +
                                          std::to_string(kSliceScratchDoubles) + " of the static scratch");
    double* tab = nullptr;
    if ((rc = scratch_ptr(e, &tab))) return rc;
    // the scratch keeps the last call's tables: clear the part this call uses (a kernel,
    // not a memset node, so that the call may be captured)
    const int nclr = (int)(2 * ntab);
    hipLaunchKernelGGL(k_zero, dim3(grid(std::max(nclr, 1), 256)), dim3(256), 0, e->stream, tab, nclr, e->flags, 0);
    HIPCHK(hipGetLastError());
    return slice_tables(e, box, nslices, tab, ntab, pair_out, bond_out);
}

int dpd_energy_output(dpd_engine* e, int32_t box, double* out3)
{
    if (!e || !out3) return DPD_ERR_PARAM;
    if (!e->committed) return fail(e, DPD_ERR_STATE, "no beads uploaded");
    if (box < 0 || box >= (int32_t)e->boxes.size()) return fail(e, DPD_ERR_PARAM, "box index out of range");
    HIPCHK(hipSetDevice(e->device));
    const BoxHost& h = e->boxes[box];
    if (e->slice_cap < 3) {
        int rc;
        if ((rc = dalloc(e, &e->slice, 3))) { e->slice_cap = 0; return rc; }
        e->slice_cap = 3;
    }
    HIPCHK(hipMemsetAsync(e->slice, 0, sizeof(double) * 3, e->stream));
    SliceArgs a{};
    a.B = h.B;
    a.pos = e->pos[e->cur];
    a.vp = e->vp;
    a.img = e->img;
    a.key = e->key[e->cur];
    a.cstart = e->cstart;
    a.table = e->table + (size_t)box * kTableStride;
    a.slot_of = e->slot_of;
    a.bond_bead = e->bond_bead;
    a.bond_par = e->bond_par;
    a.ang_bead = e->ang_bead;
    a.ang_par = e->ang_par;
    a.b0 = a.a0 = 0;
    for (int32_t b = 0; b < box; ++b) { a.b0 += (int)e->boxes[b].bh.size(); a.a0 += (int)e->boxes[b].pf.size(); }
    a.b1 = e->nbonds > 0 ? a.b0 + (int)h.bh.size() : a.b0;
    a.a1 = e->nangles > 0 ? a.a0 + (int)h.pf.size() : a.a0;
    a.ntypes = h.B.ntypes;
    HIPCHK(launch_energy_output(a, e->slice, e->stream));
    HIPCHK(hipMemcpyAsync(out3, e->slice, sizeof(double) * 3, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DPD_OK;
}

int dpd_slice_stress_bondpairs(dpd_engine* e, int32_t box, int32_t nslices, double* out)
{
    if (!e || !out) return DPD_ERR_PARAM;
    if (!e->committed) return fail(e, DPD_ERR_STATE, "no beads uploaded");
    if (box < 0 || box >= (int32_t)e->boxes.size()) return fail(e, DPD_ERR_PARAM, "box index out of range");
    if (nslices < 1 || nslices > 65536) return fail(e, DPD_ERR_PARAM, "nslices must be in [1, 65536]");
    HIPCHK(hipSetDevice(e->device));
    const BoxHost& h = e->boxes[box];
    const int nt = h.B.ntypes;
    const size_t ntab = (size_t)nslices * nt * nt * nt * 9;
    if (ntab > e->slice_cap) {
        int rc;
        if ((rc = dalloc(e, &e->slice, ntab))) { e->slice_cap = 0; return rc; }
        e->slice_cap = ntab;
    }
    HIPCHK(hipMemsetAsync(e->slice, 0, sizeof(double) * ntab, e->stream));
    SliceArgs a{};
    a.B = h.B;
    a.pos = e->pos[e->cur];
    a.img = e->img;
    a.key = e->key[e->cur];
    a.slot_of = e->slot_of;
    a.ang_bead = e->ang_bead;
    a.ang_par = e->ang_par;
    a.a0 = 0;
    for (int32_t b = 0; b < box; ++b) a.a0 += (int)e->boxes[b].pf.size();
    a.a1 = e->nangles > 0 ? a.a0 + (int)h.pf.size() : a.a0;
    a.nslice = nslices;
    a.ntypes = nt;
    a.nidx = nt * nt * nt;
    HIPCHK(launch_slice_bondpairs(a, e->slice, e->stream));
    HIPCHK(hipMemcpyAsync(out, e->slice, sizeof(double) * ntab, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DPD_OK;
}

}  // extern "C"
