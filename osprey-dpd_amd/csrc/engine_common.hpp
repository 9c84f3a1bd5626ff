// engine_common.hpp -- device-side building blocks shared by the DPD step kernels.
//
// Layout in HBM: an engine holds one or more boxes (replicas).  Beads of all boxes are
// sorted together every step by a global cell key.  Within a box the key orders cells
// row-major, x fastest: local key = cx + nx (cy + ny cz); a box's keys start at its key
// offset, a multiple of 64 (its key range is padded to whole groups of 64 keys, the unit
// of the two-level counting-sort scan).  A row of cells along x is therefore one
// contiguous slot range -- the hit-list kernel scans the three cells cx-1..cx+1 of a
// neighbour row as one range -- and a box's beads are one contiguous slot range too.
//
// Positions are FP32 offsets inside the bead's cell, o in [0, w), plus the integer cell
// (from the key) and integer periodic image counts.  Differences of positions are
// (o_i - o_j) + (c_i - c_j) * w with exact integer cell deltas, so FP32 keeps ~1e-7
// absolute precision at any box size (absolute FP32 coordinates would lose 7.6e-6 at
// L = 64).  The absolute FP64 position c*w + o and the unwrapped position
// (c + img*n)*w + o are formed only on download and in the FP64 bonded terms.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dpd {

constexpr int kGroupKeys = 64;   // cell keys per scan group
constexpr int kMaxTypes = 16;
constexpr int kTypeStride = kMaxTypes * kMaxTypes;   // one matrix of a box's table
constexpr int kTableStride = 3 * kTypeStride;        // a | gamma | sqrt(gamma) per box
constexpr int kObsStride = 32;                       // FP64 accumulators per box

// Constants of one box.  Mirrors the CCNTCell statics set by SetSimBoxLengths /
// SetTimeStepConstants (src/CNTCell.cpp:143-203).
struct BoxDev {
    int n[3];          // cells per axis
    float inv[2];      // 1 / n[0], 1 / n[1] (division by reciprocal in key_cell)
    float w[3];        // cell widths
    double wd[3];
    float dt, hdt, hdt2, lamdt;   // dt, dt/2, dt^2/2, lambda*dt
    float invrootdt;              // sqrt(24 kT / dt)
    int ntypes;
    uint32_t seed0, seed1;        // key of the pair stream
    int koff;                     // first global cell key of this box (a multiple of 64)
    int goff;                     // first global scan group of this box
    int ngroups;                  // scan groups of 64 keys (cells, padded)
    int box;                      // index of this box
};

// local key of cell (cx, cy, cz) of a box
__host__ __device__ inline int cell_key(int cx, int cy, int cz, const BoxDev& B)
{
    return cx + B.n[0] * (cy + B.n[1] * cz);
}

// x / d for 0 <= x < 2^24, d >= 1, given inv = 1/d rounded to FP32: the FP32 estimate is
// within one of the quotient, and one correction step makes it exact.
__host__ __device__ inline int div_small(int x, int d, float inv)
{
    int q = (int)((float)x * inv);
    const int r = x - q * d;
    q += (r >= d) ? 1 : ((r < 0) ? -1 : 0);
    return q;
}

// cell of a local key (boxes hold at most 2^24 cells, so div_small is exact)
__host__ __device__ inline void key_cell(int key, const BoxDev& B, int c[3])
{
    const int t = div_small(key, B.n[0], B.inv[0]);
    c[0] = key - t * B.n[0];
    c[2] = div_small(t, B.n[1], B.inv[1]);
    c[1] = t - c[2] * B.n[1];
}

__host__ __device__ inline int floordiv(int a, int n)
{
    return (a >= 0) ? a / n : -((-a + n - 1) / n);
}

// bead word packing: pos.w = type | box << 8
__host__ __device__ inline int bead_type(int w) { return w & 0xFF; }
__host__ __device__ inline int bead_box(int w) { return (w >> 8) & 0x3FFFFF; }
// bit 30 of a position's w: the bead has bonds or angles (set from the topology)
constexpr int kBondedBit = 1 << 30;

// Philox4x32-10 (Salmon et al., SC'11).  Pair uniform: counter (min id, max id, step),
// key = seed; U = (word0 >> 8) * 2^-24, exactly representable in FP32.
__device__ inline uint32_t philox_word0(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                        uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        // three-input xors as one v_bitop3_b32 each (truth table 0x96: a ^ b ^ c), which
        // the compiler does not form from two xors itself
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    return c0;
}

// key of the pair stream of one box at one step
struct RngKey {
    uint32_t seed0, seed1, step0, step1;
};

__device__ inline RngKey rng_key(const BoxDev& B, int64_t step)
{
    return RngKey{B.seed0, B.seed1, (uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32)};
}

__device__ inline float pair_uniform(uint32_t ida, uint32_t idb, const RngKey& K)
{
    const uint32_t lo = ida < idb ? ida : idb;
    const uint32_t hi = ida < idb ? idb : ida;
    const uint32_t r = philox_word0(lo, hi, K.step0, K.step1, K.seed0, K.seed1);
    return (float)(r >> 8) * 0x1p-24f;
}

// DPD pair force on bead i from bead j (src/CNTCell.cpp:1191-1225):
//   F = [a w + (-gamma w^2 (r^.v)) + sqrt(gamma) w sqrt(24kT/dt) (0.5 - U)] r^,  w = 1 - r
// returns the scalar F/r so that F_vec = s * dx.  1/r is the hardware reciprocal square
// root (1 ulp) and r = r^2 / r; callers guarantee r^2 > 1e-18 (no denormals).
__device__ inline float pair_scalar(float dr2, float dxdv, float a, float sg, float g, uint32_t idi,
                                    uint32_t idj, float invrootdt, const RngKey& K, float* pe)
{
    const float inv = __builtin_amdgcn_rsqf(dr2);
    const float dr = dr2 * inv;
    const float wr = 1.0f - dr;
    const float con = a * wr;
    const float diss = -(g * wr * wr) * (dxdv * inv);
    const float u = pair_uniform(idi, idj, K);
    const float rnd = sg * wr * invrootdt * (0.5f - u);
    if (pe) *pe = 0.5f * a * wr * wr;
    return (con + diss + rnd) * inv;
}

// Scan code of a bead: its offsets in the cell to 8 bits each, q = floor(256 o / w), with
// byte 0 the x offset, byte 1 the cell's x index modulo 256, byte 2 the y offset and byte 3
// the z offset.  The hit-list kernel tests candidates on these 4 bytes instead of the
// 16-byte position: bytes 0-1 read as one 16-bit number less 256 times a row segment's
// first cell index give the x position in that segment in units of w/256, so a row of
// three cells is scanned as one range.  q is the floor of the offset in units of w/256, so
// the offset lies in [q, q + 1) units.
__device__ inline uint32_t scan_code(const float o[3], int cx, const BoxDev& B)
{
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) q[a] = (uint32_t)min(max((int)(o[a] * (256.0f / B.w[a])), 0), 255);
    return q[0] | ((uint32_t)(cx & 0xFF) << 8) | (q[1] << 16) | (q[2] << 24);
}

// The integration arithmetic with its rounding fixed by explicit fused multiply-adds, so
// that kernels repeating a sum (k_advance and k_permute, the pair epilogue, the last step's
// k_finish) get the same bits however the compiler contracts the rest: kick is v + h f (the
// halves of UpdateMom, the predicted velocity), drift the displacement dt v + dt^2/2 f.
__device__ inline float kick(float v, float f, float h) { return fmaf(h, f, v); }
__device__ inline float drift(float v, float f, const BoxDev& B) { return fmaf(B.dt, v, B.hdt2 * f); }

// Move a bead by displacement d from offset o in cell c (image img): new offset in
// [0, w), cell and image; `jump` when it crossed more than one cell.
struct Moved {
    float o[3];
    int c[3];
    int img[3];
    int jump;
    int bad;
};

__device__ inline void advance(const float o[3], const int c[3], const int img[3], const float d[3],
                               const BoxDev& B, Moved& m)
{
    m.jump = 0;
    m.bad = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float x = o[a] + d[a];
        if (!(fabsf(x) < 1.0e6f)) { m.bad = 1; x = 0.0f; }
        int s = (int)floorf(x / B.w[a]);
        x = x - (float)s * B.w[a];
        if (x < 0.0f) { x += B.w[a]; s -= 1; }
        if (x >= B.w[a]) x = B.w[a] * 0.99999994f;
        if (x < 0.0f) x = 0.0f;
        m.jump |= (s > 1 || s < -1);
        int cc = c[a] + s;
        const int q = (cc >= 0 && cc < B.n[a]) ? 0 : floordiv(cc, B.n[a]);
        cc -= q * B.n[a];
        m.o[a] = x;
        m.c[a] = cc;
        m.img[a] = img[a] + q;
    }
}

// CBondPair::AddForce (src/BondPair.cpp:138-286) for one angle in FP64.  first = bond
// (b -> c), second = bond (a -> b); f[0] is the force on c (tail of first), f[1] on the
// shared bead b, f[2] on a (head of second); pe (may be null) gets
// CBondPair::AddPotentialEnergy (src/BondPair.cpp:320-378).
__device__ inline void angle_forces(const double ua[3], const double ub[3], const double uc[3], const double4 par,
                                    double f[3][3], double* pe)
{
    double first[3], second[3];
    for (int d = 0; d < 3; ++d) { first[d] = ub[d] - uc[d]; second[d] = ua[d] - ub[d]; }
    const double l1 = sqrt(first[0] * first[0] + first[1] * first[1] + first[2] * first[2]);
    const double l2 = sqrt(second[0] * second[0] + second[1] * second[1] + second[2] * second[2]);
    for (int r = 0; r < 3; ++r) f[r][0] = f[r][1] = f[r][2] = 0.0;
    const double magProduct = l1 * l2;
    if (!(magProduct > 0.0001)) { if (pe) *pe = 0.0; return; }
    const double k = par.x, phi0 = par.y, cosPhi0 = par.z, sinPhi0 = par.w;
    const double b1Dotb2 = first[0] * second[0] + first[1] * second[1] + first[2] * second[2];
    const double b1b2Overb1Sq = b1Dotb2 / (l1 * l1);
    const double b1b2Overb2Sq = b1Dotb2 / (l2 * l2);
    const double cosPhiSq = b1b2Overb1Sq * b1b2Overb2Sq;
    double forceMag;
    if (fabs(b1Dotb2) > 0.000001) {
        const double prefactor = sqrt(1.0 / cosPhiSq - 1.0);
        forceMag = (phi0 > 0.0) ? k * (cosPhi0 - sinPhi0 / prefactor) / magProduct : k / magProduct;
    } else {
        forceMag = k / magProduct;
    }
    for (int d = 0; d < 3; ++d) {
        const double ft = forceMag * (b1b2Overb1Sq * first[d] - second[d]);
        const double fh = forceMag * (first[d] - b1b2Overb2Sq * second[d]);
        f[0][d] = ft;
        f[2][d] = fh;
        f[1][d] = -ft - fh;
    }
    if (pe) {
        const double cps = (b1Dotb2 * b1Dotb2) / (magProduct * magProduct);
        const double sinPhi = sqrt(fmax(0.0, 1.0 - cps));
        double e;
        if (b1Dotb2 > 0.000001) e = phi0 > 0.0 ? k * (1.0 - sqrt(cps) * cosPhi0 - sinPhi * sinPhi0) : k * (1.0 - sqrt(cps));
        else if (b1Dotb2 < -0.000001) e = phi0 > 0.0 ? k * (1.0 + sqrt(cps) * cosPhi0 - sinPhi * sinPhi0) : k * (1.0 + sqrt(cps));
        else e = k * (1.0 - sinPhi0);
        *pe = e;
    }
}

// wave-level helpers (64 lanes)
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// number of lanes below this one whose bit is set in mask
__device__ inline int lane_prefix(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

// Inclusive prefix sum over the 64 lanes of a wavefront with DPP moves: row_shr 1, 2, 4, 8
// scan each row of 16 lanes (lanes shifted in from outside the row read 0), row_bcast 15
// adds row 0's total to row 1 and row 2's to row 3, row_bcast 31 adds the first half's total
// to rows 2 and 3.  Six dependent adds instead of six LDS permutes.
__device__ inline int wave_incl_scan(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);   // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);   // row_bcast:31
    return x;
}

// Active lanes of a wavefront holding equal keys form a group: this lane's rank in its
// group, the group's size and its lowest lane.  Must be reached by the whole wavefront.
// One pass per distinct key, with wave-uniform bookkeeping; a wavefront of cell-sorted
// beads holds ~20 distinct cells.
__device__ inline void wave_groups(int key, bool active, int& rank, int& size, int& leader)
{
    uint64_t todo = __ballot(active);
    rank = 0;
    size = 0;
    leader = 0;
    while (todo) {
        const int first = __builtin_ctzll(todo);
        const int k0 = __builtin_amdgcn_readlane(key, first);
        const bool mine = active && key == k0;
        const uint64_t m = __ballot(mine);
        if (mine) {
            rank = lane_prefix(m);
            size = __popcll(m);
            leader = first;
        }
        todo &= ~m;
    }
}

// Add per-thread FP64 values v[0..N) into acc[box * kObsStride + base + i]: one
// block-wide reduction for the threads of the block's first box, direct atomics for
// threads of another box (only blocks straddling a box boundary have any).
template <int N, int WAVES = 4>
__device__ inline void accumulate_boxes(const double (&v)[N], int box, int box0, int base, double* acc)
{
    __shared__ double red[N][WAVES];   // workgroups of up to 64 * WAVES threads
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const bool mine = box == box0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double s = wave_sum(mine ? v[i] : 0.0);
        if (lane == 0) red[i][wid] = s;
    }
    if (!mine && box >= 0)
        for (int i = 0; i < N; ++i)
            if (v[i] != 0.0) atomicAdd(acc + box * kObsStride + base + i, v[i]);
    __syncthreads();
    if (threadIdx.x < N) {
        const int nw = (blockDim.x + 63) >> 6;
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += red[threadIdx.x][w];
        atomicAdd(acc + box0 * kObsStride + base + threadIdx.x, s);
    }
}

}  // namespace dpd
