// pair_compact.hip -- compacted full-shell DPD pair kernel for gfx950.
//
// Thread per bead over the cell-sorted beads, like the full-shell kernel: each lane walks
// the 27 cells around its bead and tests every candidate j (global memory, L1/L2 hits:
// a wavefront's 64 beads sit in ~20 neighbouring cells).  Unlike the full-shell kernel,
// a candidate with r < 1 is not evaluated on the spot -- where ~15% of the lanes would
// run the pair force and its Philox draw while the rest idle -- but appended to a
// per-wavefront LDS ring (ballot + mbcnt); whenever 64 pairs are queued they are
// evaluated together, one per lane.  Each loop trip tests four candidates per lane
// (four independent loads in flight), and the next cell's slot range is fetched while
// the current cell is tested, so each lane's chain of dependent loads stays short.  The force of a queued pair belongs to
// bead i, the bead of the lane that found it; the evaluating lane adds it to i's LDS
// accumulator.
// Consecutive queue entries come from different lanes, so these LDS atomics rarely
// collide.  Each pair is evaluated from both sides with the same operands (dx negated),
// so the two forces are exact negatives and no force leaves its bead's lane; virial and
// PE count half per side.  LDS use is ~19 KB per 256-thread workgroup (no staging: eight
// workgroups, 32 wavefronts, per CU), and no bead count per cell or brick can overflow.  Workgroups are mapped XCD-aware
// (workgroup b runs on XCD b % 8; it takes bead block (b % 8) * ceil(G/8) + b / 8), so each
// XCD walks one contiguous slab of the box and neighbour beads hit its own L2.
#include "pair_kernels.hpp"

namespace dpd {

namespace {

constexpr int kWaves = 4;   // 256 threads
constexpr int kQueue = 128;   // ring; holds < 64 + 64 entries between drains

struct Wave {
    uint64_t queue[kQueue];   // i lane | neighbour offset << 6 | j << 32
    float4 p[64];             // the wavefront's beads: cell offset xyz, w = type | box << 8
    float4 v[64];             // predicted velocity xyz, w = id
    float fx[64], fy[64], fz[64];
};

template <bool OBS, bool MULTI>
__device__ inline void eval_queued(Wave& W, uint64_t e, const PairArgs& A, const BoxDev& B0, const float4* tab0,
                                   float (*vs)[64])
{
    const int il = (int)(e & 63u);
    const int nb = (int)((e >> 6) & 31u);
    const int j = (int)(e >> 32);
    const float4 pi = W.p[il], vi = W.v[il];
    const float4 pj = A.pos[j], vj = A.vp[j];
    const int wi = __float_as_int(pi.w);
    const int box = MULTI ? bead_box(wi) : 0;
    const BoxDev& B = MULTI ? A.boxes[box] : B0;
    const float sx = (float)(nb % 3 - 1) * B.w[0];
    const float sy = (float)((nb / 3) % 3 - 1) * B.w[1];
    const float sz = (float)(nb / 9 - 1) * B.w[2];
    const float dx = (pi.x - pj.x) - sx, dy = (pi.y - pj.y) - sy, dz = (pi.z - pj.z) - sz;
    const float dr2 = dx * dx + dy * dy + dz * dz;
    const int tt = bead_type(wi) * kMaxTypes + bead_type(__float_as_int(pj.w));
    float a, g, sg;
    if (MULTI) {
        const float* t = A.table + box * kTableStride;
        a = t[tt];
        g = t[kTypeStride + tt];
        sg = t[2 * kTypeStride + tt];
    } else {
        const float4 t = tab0[tt];
        a = t.x;
        g = t.y;
        sg = t.z;
    }
    const RngKey K = rng_key(B, *A.step);
    const float dxdv = dx * (vi.x - vj.x) + dy * (vi.y - vj.y) + dz * (vi.z - vj.z);
    float pe;
    const float s = pair_scalar(dr2, dxdv, a, sg, g, (uint32_t)__float_as_int(vi.w), (uint32_t)__float_as_int(vj.w),
                                B.invrootdt, K, OBS ? &pe : nullptr);
    const float Fx = s * dx, Fy = s * dy, Fz = s * dz;
    atomicAdd(&W.fx[il], Fx);
    atomicAdd(&W.fy[il], Fy);
    atomicAdd(&W.fz[il], Fz);
    if (OBS) {
        // each pair is seen from both sides: half per side
        atomicAdd(&vs[0][il], 0.5f * dx * Fx); atomicAdd(&vs[1][il], 0.5f * dy * Fx);
        atomicAdd(&vs[2][il], 0.5f * dz * Fx); atomicAdd(&vs[3][il], 0.5f * dx * Fy);
        atomicAdd(&vs[4][il], 0.5f * dy * Fy); atomicAdd(&vs[5][il], 0.5f * dz * Fy);
        atomicAdd(&vs[6][il], 0.5f * dx * Fz); atomicAdd(&vs[7][il], 0.5f * dy * Fz);
        atomicAdd(&vs[8][il], 0.5f * dz * Fz); atomicAdd(&vs[9][il], 0.5f * pe);
    }
}

template <bool OBS, bool MULTI>
__global__ __launch_bounds__(256) void k_pair_compact(PairArgs A, BoxDev B0, int nblocks, int per_xcd)
{
    const int blk = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (blk >= nblocks) return;   // whole workgroup exits together
    __shared__ Wave waves[kWaves];
    __shared__ float4 tab[MULTI ? 1 : kTypeStride];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Wave& W = waves[wave];
    if (!MULTI) {
        for (int q = threadIdx.x; q < kTypeStride; q += 256)
            tab[q] = make_float4(A.table[q], A.table[kTypeStride + q], A.table[2 * kTypeStride + q], 0.f);
    }
    const int i = blk * 256 + threadIdx.x;
    const bool active = i < A.n;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f), vi = pi;
    int box = -1;
    int c[3] = {0, 0, 0};
    int n0 = B0.n[0], n1 = B0.n[1], n2 = B0.n[2], nb0 = B0.nb[0], nb1 = B0.nb[1], koff = B0.koff;
    float w0 = B0.w[0], w1 = B0.w[1], w2 = B0.w[2];
    if (active) {
        pi = A.pos[i];
        vi = A.vp[i];
        box = MULTI ? bead_box(__float_as_int(pi.w)) : 0;
        const BoxDev& B = MULTI ? A.boxes[box] : B0;
        if (MULTI) {
            n0 = B.n[0]; n1 = B.n[1]; n2 = B.n[2]; nb0 = B.nb[0]; nb1 = B.nb[1]; koff = B.koff;
            w0 = B.w[0]; w1 = B.w[1]; w2 = B.w[2];
        }
        key_cell(A.key[i] - koff, B, c);
    }
    W.p[lane] = pi;
    W.v[lane] = vi;
    W.fx[lane] = 0.f;
    W.fy[lane] = 0.f;
    W.fz[lane] = 0.f;
    __shared__ float vir_s[OBS ? kWaves : 1][10][64];
    if (OBS)
        for (int q = 0; q < 10; ++q) vir_s[wave][q][lane] = 0.f;
    __syncthreads();

    volatile uint64_t* queue = W.queue;
    int head = 0, tail = 0;   // wave-uniform ring positions
    // the slot range of neighbour cell nb (0..26) of this lane's bead
    auto cell_range = [&](int nb, int& j0, int& j1) {
        j0 = j1 = 0;
        if (!active) return;
        int cx = c[0] + nb % 3 - 1, cy = c[1] + (nb / 3) % 3 - 1, cz = c[2] + nb / 9 - 1;
        cx += (cx < 0) ? n0 : ((cx >= n0) ? -n0 : 0);
        cy += (cy < 0) ? n1 : ((cy >= n1) ? -n1 : 0);
        cz += (cz < 0) ? n2 : ((cz >= n2) ? -n2 : 0);
        const int brick = (cx >> 2) + nb0 * ((cy >> 2) + nb1 * (cz >> 2));
        const int k = koff + brick * kBrickCells + (((cz & 3) * 4 + (cy & 3)) * 4 + (cx & 3));
        j0 = A.cstart[k];
        j1 = A.cstart[k + 1];
    };
    uint64_t tag = 0;
    // append a tested candidate; evaluate whenever 64 pairs are queued
    auto push = [&](bool pass, int jj) {
        const uint64_t mk = __ballot(pass);
        if (pass) queue[(tail + lane_prefix(mk)) & (kQueue - 1)] = tag | ((uint64_t)(uint32_t)jj << 32);
        tail += __popcll(mk);
        if (tail - head >= 64) {
            __builtin_amdgcn_wave_barrier();
            eval_queued<OBS, MULTI>(W, queue[(head + lane) & (kQueue - 1)], A, B0, tab, vir_s[OBS ? wave : 0]);
            head += 64;
            __builtin_amdgcn_wave_barrier();
        }
    };
    int nj0, nj1;
    cell_range(0, nj0, nj1);
    for (int nb = 0; nb < 27; ++nb) {
        int j = nj0;
        const int j1 = nj1;
        if (nb < 26) cell_range(nb + 1, nj0, nj1);   // prefetch the next cell's range
        const float sx = (float)(nb % 3 - 1) * w0, sy = (float)((nb / 3) % 3 - 1) * w1,
                    sz = (float)(nb / 9 - 1) * w2;
        tag = (uint64_t)(lane | (nb << 6));
        // four candidates per trip: four independent loads in flight per lane
        while (__any(j < j1)) {
            float4 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                q[u] = make_float4(1e30f, 1e30f, 1e30f, 0.f);
                if (j + u < j1) q[u] = A.pos[j + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float dx = (pi.x - q[u].x) - sx, dy = (pi.y - q[u].y) - sy, dz = (pi.z - q[u].z) - sz;
                const float dr2 = dx * dx + dy * dy + dz * dz;
                push(j + u < j1 && dr2 < 1.0f && dr2 > 1e-18f && j + u != i, j + u);
            }
            j += 4;
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < tail - head)
        eval_queued<OBS, MULTI>(W, queue[(head + lane) & (kQueue - 1)], A, B0, tab, vir_s[OBS ? wave : 0]);
    __builtin_amdgcn_wave_barrier();
    if (active) A.frc[i] = make_float4(W.fx[lane], W.fy[lane], W.fz[lane], 0.f);
    if (OBS) {
        double vir[10];
        for (int q = 0; q < 10; ++q) vir[q] = active ? (double)vir_s[wave][q][lane] : 0.0;
        const int box0 = MULTI ? bead_box(__float_as_int(A.pos[blk * 256].w)) : 0;
        accumulate_boxes<10>(vir, box, box0, kObsVir, A.acc);
    }
}

}  // namespace

hipError_t launch_pair_compact(const PairArgs& a, const BoxDev& B0, bool obs, hipStream_t s)
{
    const int g = (a.n + 255) / 256;
    const int per_xcd = (g + 7) / 8;
    const dim3 grid(8 * per_xcd);
    const bool multi = a.nbox > 1;
    if (obs) {
        if (multi) hipLaunchKernelGGL((k_pair_compact<true, true>), grid, dim3(256), 0, s, a, B0, g, per_xcd);
        else hipLaunchKernelGGL((k_pair_compact<true, false>), grid, dim3(256), 0, s, a, B0, g, per_xcd);
    } else {
        if (multi) hipLaunchKernelGGL((k_pair_compact<false, true>), grid, dim3(256), 0, s, a, B0, g, per_xcd);
        else hipLaunchKernelGGL((k_pair_compact<false, false>), grid, dim3(256), 0, s, a, B0, g, per_xcd);
    }
    return hipGetLastError();
}

}  // namespace dpd
