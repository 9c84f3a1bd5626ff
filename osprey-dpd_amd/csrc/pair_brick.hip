// pair_brick.hip -- LDS cell-brick DPD pair kernel for gfx950.
//
// One workgroup (4 wavefronts) per 4x4x4-cell brick.  Beads are cell-sorted brick-major,
// so the brick's own ("home") beads are one contiguous slot range.  The kernel:
//   1. stages the home beads and the beads of the one-cell halo (6x6x6 cells, periodic)
//      into LDS as brick-local FP32 positions, predicted velocities, type and id;
//   2. walks tasks (home bead i, one of the 27 neighbour cells), tests every candidate j
//      against r < 1 and appends surviving pairs to a per-wavefront LDS queue
//      (ballot + mbcnt compaction), so the expensive force/RNG code always runs on 64
//      live pairs at once instead of the ~15% of lanes whose candidate is in range;
//   3. evaluates queued pairs 64 at a time and accumulates forces with LDS float atomics:
//      a pair of two home beads is evaluated once (Newton's third law in LDS); a pair
//      with a halo bead is evaluated by both bricks, each keeping its own bead's force,
//      so no force ever leaves the workgroup through global atomics;
//   4. writes the home forces with one coalesced store.
// A brick whose staged beads exceed the LDS capacity falls back to the thread-per-bead
// 27-cell loop over global memory (same arithmetic).
//
// Workgroups are mapped to bricks XCD-aware: the dispatcher places workgroup b on XCD
// b % 8, so brick = (b % 8) * ceil(nb/8) + b / 8 gives each XCD a contiguous slab of
// bricks whose halos mostly hit its own L2.
#include "pair_kernels.hpp"

namespace dpd {

namespace {

constexpr int kMaxWaves = 8;       // up to 512 threads per workgroup
constexpr int kMaxStaged = 1024;   // beads in brick + halo
constexpr int kMaxHome = 384;      // beads in the brick
constexpr int kStaged = 216;       // 6x6x6 staged cells
constexpr int kQueue = 128;        // per-wavefront pair queue

__device__ inline uint64_t ballot64(bool p) { return __ballot(p); }

struct Smem {
    float4 pos[kMaxStaged];     // brick-centred position xyz, w = type | local cell << 8
    float4 vp[kMaxStaged];      // predicted velocity xyz, w = id
    float fx[kMaxHome], fy[kMaxHome], fz[kMaxHome];
    int start[kStaged + 1];     // LDS index of each staged cell's first bead
    int gstart[kStaged];        // global slot of each staged cell's first bead
    uint32_t queue[kMaxWaves][kQueue];
    float table[kTableStride];
    double red[10][kMaxWaves];
    int nstaged;
};

// staged cell index of local coords (lx, ly, lz) in [-1, 4]
__device__ inline int sidx(int lx, int ly, int lz) { return (lx + 1) + 6 * ((ly + 1) + 6 * (lz + 1)); }

// evaluate one queued pair; returns virial/PE contributions through acc (OBS)
template <bool OBS>
__device__ inline void eval_pair(Smem& S, uint32_t packed, int nhome, float invrootdt, const RngKey& K,
                                 float* vir)
{
    const int li = (int)(packed & 0xFFFFu), lj = (int)(packed >> 16);
    const float4 pi = S.pos[li], pj = S.pos[lj];
    const float4 vi = S.vp[li], vj = S.vp[lj];
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float dr2 = dx * dx + dy * dy + dz * dz;
    const int ti = __float_as_int(pi.w) & 0xFF, tj = __float_as_int(pj.w) & 0xFF;
    const int tt = ti * kMaxTypes + tj;
    const float a = S.table[tt], g = S.table[kTypeStride + tt], sg = S.table[2 * kTypeStride + tt];
    const float dxdv = dx * (vi.x - vj.x) + dy * (vi.y - vj.y) + dz * (vi.z - vj.z);
    float pe;
    const float s = pair_scalar(dr2, dxdv, a, sg, g, (uint32_t)__float_as_int(vi.w),
                                (uint32_t)__float_as_int(vj.w), invrootdt, K, OBS ? &pe : nullptr);
    const float Fx = s * dx, Fy = s * dy, Fz = s * dz;
    atomicAdd(&S.fx[li], Fx);
    atomicAdd(&S.fy[li], Fy);
    atomicAdd(&S.fz[li], Fz);
    const bool both = lj < nhome;
    if (both) {
        atomicAdd(&S.fx[lj], -Fx);
        atomicAdd(&S.fy[lj], -Fy);
        atomicAdd(&S.fz[lj], -Fz);
    }
    if (OBS) {
        // a halo pair is evaluated by both bricks: count half from each
        const float wgt = both ? 1.0f : 0.5f;
        vir[0] += wgt * dx * Fx; vir[1] += wgt * dy * Fx; vir[2] += wgt * dz * Fx;
        vir[3] += wgt * dx * Fy; vir[4] += wgt * dy * Fy; vir[5] += wgt * dz * Fy;
        vir[6] += wgt * dx * Fz; vir[7] += wgt * dy * Fz; vir[8] += wgt * dz * Fz;
        vir[9] += wgt * pe;
    }
}

// thread-per-bead 27-cell loop over global memory (overflow fallback); same arithmetic
// as k_pair_fullshell
template <bool OBS>
__device__ inline void fullshell_bead(const PairArgs& A, const BoxDev& B, const RngKey& K, int i, const float* table,
                                      float* vir)
{
    const float4 pi = A.pos[i];
    const float4 vi = A.vp[i];
    const int ti = bead_type(__float_as_int(pi.w));
    const uint32_t idi = (uint32_t)__float_as_int(vi.w);
    int c[3];
    key_cell(A.key[i] - B.koff, B, c);
    float fx = 0.f, fy = 0.f, fz = 0.f;
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
                    const float ddx = (pi.x - pj.x) + sx, ddy = (pi.y - pj.y) + sy, ddz = (pi.z - pj.z) + sz;
                    const float dr2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    if (dr2 < 1.0f && dr2 > 1e-18f) {
                        const float4 vj = A.vp[j];
                        const int tt = ti * kMaxTypes + bead_type(__float_as_int(pj.w));
                        const float dxdv = ddx * (vi.x - vj.x) + ddy * (vi.y - vj.y) + ddz * (vi.z - vj.z);
                        float pe;
                        const float s = pair_scalar(dr2, dxdv, table[tt], table[2 * kTypeStride + tt],
                                                    table[kTypeStride + tt], idi, (uint32_t)__float_as_int(vj.w),
                                                    B.invrootdt, K, OBS ? &pe : nullptr);
                        const float Fx = s * ddx, Fy = s * ddy, Fz = s * ddz;
                        fx += Fx; fy += Fy; fz += Fz;
                        if (OBS) {
                            vir[0] += 0.5f * ddx * Fx; vir[1] += 0.5f * ddy * Fx; vir[2] += 0.5f * ddz * Fx;
                            vir[3] += 0.5f * ddx * Fy; vir[4] += 0.5f * ddy * Fy; vir[5] += 0.5f * ddz * Fy;
                            vir[6] += 0.5f * ddx * Fz; vir[7] += 0.5f * ddy * Fz; vir[8] += 0.5f * ddz * Fz;
                            vir[9] += 0.5f * pe;
                        }
                    }
                }
            }
        }
    }
    A.frc[i] = make_float4(fx, fy, fz, 0.f);
}

template <int kThreads, bool OBS>
__global__ __launch_bounds__(kThreads) void k_pair_brick(PairArgs A, int per_xcd)
{
    constexpr int kWaves = kThreads / 64;
    static_assert(kThreads >= kStaged && kWaves <= kMaxWaves, "one thread per staged cell");
    __shared__ Smem S;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int gbrick = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (gbrick >= A.nbricks) return;   // whole workgroup exits together
    // the box owning this brick (bricks of a box are contiguous)
    int lo = 0, hi = A.nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (A.boxes[mid].boff <= gbrick) lo = mid; else hi = mid - 1;
    }
    const BoxDev B = A.boxes[lo];
    const int box = lo;
    const int brick = gbrick - B.boff;
    const RngKey K = rng_key(B, *A.step);
    const int bx = brick % B.nb[0], byz = brick / B.nb[0];
    const int by = byz % B.nb[1], bz = byz / B.nb[1];
    const int X0 = bx * kBrick, Y0 = by * kBrick, Z0 = bz * kBrick;
    const int kb = B.koff + brick * kBrickCells;
    const int base = A.cstart[kb];
    const int nhome = A.cstart[kb + kBrickCells] - base;
    const float* gtab = A.table + box * kTableStride;
    for (int q = tid; q < kTableStride; q += kThreads) S.table[q] = gtab[q];

    // ---- staged cell table: home cells keep their global order, halo cells follow
    int cnt = 0, gs = 0;
    if (tid < kStaged) {
        const int lx = tid % 6 - 1, ly = (tid / 6) % 6 - 1, lz = tid / 36 - 1;
        const int gx = X0 + lx, gy = Y0 + ly, gz = Z0 + lz;
        const bool home = lx >= 0 && lx < 4 && ly >= 0 && ly < 4 && lz >= 0 && lz < 4 && gx < B.n[0] &&
                          gy < B.n[1] && gz < B.n[2];
        const int wx = gx < 0 ? gx + B.n[0] : (gx >= B.n[0] ? gx - B.n[0] : gx);
        const int wy = gy < 0 ? gy + B.n[1] : (gy >= B.n[1] ? gy - B.n[1] : gy);
        const int wz = gz < 0 ? gz + B.n[2] : (gz >= B.n[2] ? gz - B.n[2] : gz);
        const int k = B.koff + cell_key(wx, wy, wz, B);
        gs = A.cstart[k];
        cnt = A.cstart[k + 1] - gs;
        S.gstart[tid] = gs;
        if (home) {
            S.start[tid] = gs - base;   // home cells: LDS index = slot - base
            cnt = 0;                    // not part of the halo prefix
        } else {
            S.start[tid] = -1;
        }
    }
    // halo prefix (exclusive) over the 216 staged cells, halo beads placed after home beads
    __shared__ int wsum[kWaves];
    {
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        if (tid < kStaged && S.start[tid] < 0) S.start[tid] = nhome + before + incl - cnt;
        if (tid == kThreads - 1) S.nstaged = nhome + before + incl;
    }
    __syncthreads();
    const int nstaged = S.nstaged;
    const bool fits = nhome <= kMaxHome && nstaged <= kMaxStaged;
    if (!fits) {
        float vir[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = tid; i < nhome; i += kThreads) fullshell_bead<OBS>(A, B, K, base + i, S.table, vir);
        if (OBS) {
            for (int q = 0; q < 10; ++q) {
                const double s = wave_sum((double)vir[q]);
                if (lane == 0) S.red[q][wave] = s;
            }
            __syncthreads();
            if (tid < 10) {
                double t = 0.0;
                for (int w = 0; w < kWaves; ++w) t += S.red[tid][w];
                atomicAdd(A.acc + box * kObsStride + kObsVir + tid, t);
            }
        }
        return;
    }

    // ---- stage beads: brick-centred positions (cell - 2) * w + offset
    for (int i = tid; i < nhome; i += kThreads) {
        const int slot = base + i;
        const float4 p = A.pos[slot];
        const int l = (A.key[slot] - B.koff) & 63;
        const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
        const int packed = bead_type(__float_as_int(p.w)) | (sidx(lx, ly, lz) << 8);
        S.pos[i] = make_float4((float)(lx - 2) * B.w[0] + p.x, (float)(ly - 2) * B.w[1] + p.y,
                               (float)(lz - 2) * B.w[2] + p.z, __int_as_float(packed));
        S.vp[i] = A.vp[slot];
        S.fx[i] = 0.f; S.fy[i] = 0.f; S.fz[i] = 0.f;
    }
    for (int c = tid; c < kStaged; c += kThreads) {
        const int lx = c % 6 - 1, ly = (c / 6) % 6 - 1, lz = c / 36 - 1;
        const bool home = lx >= 0 && lx < 4 && ly >= 0 && ly < 4 && lz >= 0 && lz < 4 && X0 + lx < B.n[0] &&
                          Y0 + ly < B.n[1] && Z0 + lz < B.n[2];
        if (home) continue;   // home beads were copied above
        const int s0 = S.start[c];
        const int g0 = S.gstart[c];
        int k;
        {
            const int gx = X0 + lx, gy = Y0 + ly, gz = Z0 + lz;
            const int wx = gx < 0 ? gx + B.n[0] : (gx >= B.n[0] ? gx - B.n[0] : gx);
            const int wy = gy < 0 ? gy + B.n[1] : (gy >= B.n[1] ? gy - B.n[1] : gy);
            const int wz = gz < 0 ? gz + B.n[2] : (gz >= B.n[2] ? gz - B.n[2] : gz);
            k = B.koff + cell_key(wx, wy, wz, B);
        }
        const int cntc = A.cstart[k + 1] - g0;
        for (int q = 0; q < cntc; ++q) {
            const float4 p = A.pos[g0 + q];
            S.pos[s0 + q] = make_float4((float)(lx - 2) * B.w[0] + p.x, (float)(ly - 2) * B.w[1] + p.y,
                                        (float)(lz - 2) * B.w[2] + p.z, p.w);
            S.vp[s0 + q] = A.vp[g0 + q];
        }
    }
    // per-cell counts for the candidate loop: home cells from cstart, halo from staging
    __shared__ int scount[kStaged];
    if (tid < kStaged) {
        const int lx = tid % 6 - 1, ly = (tid / 6) % 6 - 1, lz = tid / 36 - 1;
        const int gx = X0 + lx, gy = Y0 + ly, gz = Z0 + lz;
        const int wx = gx < 0 ? gx + B.n[0] : (gx >= B.n[0] ? gx - B.n[0] : gx);
        const int wy = gy < 0 ? gy + B.n[1] : (gy >= B.n[1] ? gy - B.n[1] : gy);
        const int wz = gz < 0 ? gz + B.n[2] : (gz >= B.n[2] ? gz - B.n[2] : gz);
        const int k = B.koff + cell_key(wx, wy, wz, B);
        scount[tid] = A.cstart[k + 1] - A.cstart[k];
    }
    __syncthreads();

    // ---- candidate filter + per-wavefront compaction + evaluation
    float vir[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    volatile uint32_t* q = S.queue[wave];
    int qn = 0;                         // wave-uniform queue length
    const int ntasks = nhome * 27;
    for (int tb = wave * 64; tb < ntasks; tb += kThreads) {
        const int task = tb + lane;
        int li = 0, j = 0, j1 = 0;
        float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
        if (task < ntasks) {
            li = task / 27;
            const int s = task - li * 27;
            pi = S.pos[li];
            const int hc = __float_as_int(pi.w) >> 8;          // staged index of i's cell
            const int nc = hc + (s % 3 - 1) + 6 * ((s / 3) % 3 - 1) + 36 * (s / 9 - 1);
            j = S.start[nc];
            j1 = j + scount[nc];
            if (j < nhome && j1 <= nhome) {
                // home neighbour cell: each home-home pair once (j > i)
                if (j <= li) j = li + 1;
            }
        }
        while (__any(j < j1)) {
            bool pass = false;
            if (j < j1) {
                const float4 pj = S.pos[j];
                const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                const float dr2 = dx * dx + dy * dy + dz * dz;
                pass = dr2 < 1.0f && dr2 > 1e-18f;
            }
            const uint64_t m = ballot64(pass);
            if (pass) q[qn + lane_prefix(m)] = (uint32_t)li | ((uint32_t)j << 16);
            qn += __popcll(m);
            ++j;
            if (qn >= 64) {
                __builtin_amdgcn_wave_barrier();
                const uint32_t pk = q[lane];
                eval_pair<OBS>(S, pk, nhome, B.invrootdt, K, vir);
                const uint32_t rest = (lane + 64 < qn) ? q[lane + 64] : 0u;
                __builtin_amdgcn_wave_barrier();
                if (lane + 64 < qn) q[lane] = rest;
                qn -= 64;
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < qn) eval_pair<OBS>(S, q[lane], nhome, B.invrootdt, K, vir);
    __syncthreads();

    for (int i = tid; i < nhome; i += kThreads) A.frc[base + i] = make_float4(S.fx[i], S.fy[i], S.fz[i], 0.f);
    if (OBS) {
        for (int c = 0; c < 10; ++c) {
            const double s = wave_sum((double)vir[c]);
            if (lane == 0) S.red[c][wave] = s;
        }
        __syncthreads();
        if (tid < 10) {
            double t = 0.0;
            for (int w = 0; w < kWaves; ++w) t += S.red[tid][w];
            atomicAdd(A.acc + box * kObsStride + kObsVir + tid, t);
        }
    }
}

}  // namespace

hipError_t launch_pair_brick(const PairArgs& a, bool obs, hipStream_t s, int threads)
{
    const int per_xcd = (a.nbricks + 7) / 8;
    const dim3 grid(8 * per_xcd);
    if (threads == 512) {
        if (obs) hipLaunchKernelGGL((k_pair_brick<512, true>), grid, dim3(512), 0, s, a, per_xcd);
        else hipLaunchKernelGGL((k_pair_brick<512, false>), grid, dim3(512), 0, s, a, per_xcd);
    } else {
        if (obs) hipLaunchKernelGGL((k_pair_brick<256, true>), grid, dim3(256), 0, s, a, per_xcd);
        else hipLaunchKernelGGL((k_pair_brick<256, false>), grid, dim3(256), 0, s, a, per_xcd);
    }
    return hipGetLastError();
}

}  // namespace dpd
