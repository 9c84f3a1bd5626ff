// pair_cells.hip -- cell-wavefront DPD pair kernel for gfx950.
//
// Each wavefront owns one home cell at a time and evaluates the full 27-cell shell of its
// beads, reading memory per cell instead of per candidate:
//   1. lanes 0..26 look up the slot ranges of the 27 neighbour cells (periodic) and a
//      wave-level scan lays them end to end; the lanes then write the concatenated
//      candidate list (slot, neighbour offset) into a per-wavefront LDS array;
//   2. the wavefront loads the candidates' positions, 64 per pass, with one coalesced
//      load per pass (consecutive list entries are consecutive slots of a cell);
//   3. for each bead i of the home cell (a wave-uniform loop; i's position is a scalar
//      load) every lane tests its candidate against i, and pairs with r < 1 are compacted
//      into a per-wavefront LDS queue (ballot + mbcnt);
//   4. whenever 64 pairs are queued they are evaluated together, one per lane, and the
//      force is added to bead i's LDS accumulator (the queue holds few distinct i, whose
//      atomics serialise ~10 deep; the pair work dominates).
// Each pair is evaluated from both sides with the same operands (dx negated), so the two
// forces are exact negatives and no force leaves its home brick; virial and PE count half
// per side.  A workgroup (4 wavefronts) owns one 4x4x4-cell brick and writes its forces
// with one coalesced store; workgroups are mapped to bricks XCD-aware (workgroup b runs on
// XCD b % 8 and takes brick (b % 8) * ceil(nb/8) + b / 8).  A home cell whose candidate
// list exceeds kMaxM, or a brick with more than kMaxHome beads, takes the thread-per-bead
// loop over global memory (same arithmetic).
#include "pair_kernels.hpp"

namespace dpd {

namespace {

constexpr int kWaves = 4;       // 256 threads
constexpr int kMaxM = 256;      // candidate-list entries per home cell
constexpr int kQueue = 128;     // per-wavefront pair queue (ring)
constexpr int kMaxHome = 1024;  // beads per brick with LDS force accumulators

struct WaveMem {
    int slot[kMaxM];          // candidate slots
    uint8_t code[kMaxM];      // neighbour offset 0..26 of each candidate's cell
    uint64_t queue[kQueue];   // i - base | code << 16 | j << 32
};

struct Smem {
    WaveMem w[kWaves];
    float fx[kMaxHome], fy[kMaxHome], fz[kMaxHome];
    float4 tab[kTypeStride];
};

template <bool OBS>
__device__ inline void eval_queued(Smem& S, uint64_t e, int base, const PairArgs& A, const BoxDev& B,
                                   const RngKey& K, float* vir)
{
    const int ir = (int)(e & 0xFFFFu);
    const int nb = (int)((e >> 16) & 31u);
    const int j = (int)(e >> 32);
    const int i = base + ir;
    const float4 pi = A.pos[i], vi = A.vp[i];
    const float4 pj = A.pos[j], vj = A.vp[j];
    const float dx = (pi.x - pj.x) - (float)(nb % 3 - 1) * B.w[0];
    const float dy = (pi.y - pj.y) - (float)((nb / 3) % 3 - 1) * B.w[1];
    const float dz = (pi.z - pj.z) - (float)(nb / 9 - 1) * B.w[2];
    const float dr2 = dx * dx + dy * dy + dz * dz;
    const float4 t = S.tab[bead_type(__float_as_int(pi.w)) * kMaxTypes + bead_type(__float_as_int(pj.w))];
    const float dxdv = dx * (vi.x - vj.x) + dy * (vi.y - vj.y) + dz * (vi.z - vj.z);
    float pe;
    const float s = pair_scalar(dr2, dxdv, t.x, t.z, t.y, (uint32_t)__float_as_int(vi.w),
                                (uint32_t)__float_as_int(vj.w), B.invrootdt, K, OBS ? &pe : nullptr);
    const float Fx = s * dx, Fy = s * dy, Fz = s * dz;
    atomicAdd(&S.fx[ir], Fx);
    atomicAdd(&S.fy[ir], Fy);
    atomicAdd(&S.fz[ir], Fz);
    if (OBS) {
        vir[0] += 0.5f * dx * Fx; vir[1] += 0.5f * dy * Fx; vir[2] += 0.5f * dz * Fx;
        vir[3] += 0.5f * dx * Fy; vir[4] += 0.5f * dy * Fy; vir[5] += 0.5f * dz * Fy;
        vir[6] += 0.5f * dx * Fz; vir[7] += 0.5f * dy * Fz; vir[8] += 0.5f * dz * Fz;
        vir[9] += 0.5f * pe;
    }
}

// thread-per-bead 27-cell loop over global memory for bead i (capacity fallback)
template <bool OBS>
__device__ inline void fallback_bead(const PairArgs& A, const BoxDev& B, const RngKey& K, int i, const float4* tab,
                                     float* vir, float f[3])
{
    const float4 pi = A.pos[i];
    const float4 vi = A.vp[i];
    const int ti = bead_type(__float_as_int(pi.w));
    const uint32_t idi = (uint32_t)__float_as_int(vi.w);
    int c[3];
    key_cell(A.key[i] - B.koff, B, c);
    f[0] = f[1] = f[2] = 0.f;
    for (int nb = 0; nb < 27; ++nb) {
        const int ox = nb % 3 - 1, oy = (nb / 3) % 3 - 1, oz = nb / 9 - 1;
        int cx = c[0] + ox, cy = c[1] + oy, cz = c[2] + oz;
        cx += (cx < 0) ? B.n[0] : ((cx >= B.n[0]) ? -B.n[0] : 0);
        cy += (cy < 0) ? B.n[1] : ((cy >= B.n[1]) ? -B.n[1] : 0);
        cz += (cz < 0) ? B.n[2] : ((cz >= B.n[2]) ? -B.n[2] : 0);
        const int k = B.koff + cell_key(cx, cy, cz, B);
        const float sx = (float)ox * B.w[0], sy = (float)oy * B.w[1], sz = (float)oz * B.w[2];
        for (int j = A.cstart[k], j1 = A.cstart[k + 1]; j < j1; ++j) {
            if (j == i) continue;
            const float4 pj = A.pos[j];
            const float dx = (pi.x - pj.x) - sx, dy = (pi.y - pj.y) - sy, dz = (pi.z - pj.z) - sz;
            const float dr2 = dx * dx + dy * dy + dz * dz;
            if (dr2 < 1.0f && dr2 > 1e-18f) {
                const float4 vj = A.vp[j];
                const float4 t = tab[ti * kMaxTypes + bead_type(__float_as_int(pj.w))];
                const float dxdv = dx * (vi.x - vj.x) + dy * (vi.y - vj.y) + dz * (vi.z - vj.z);
                float pe;
                const float s = pair_scalar(dr2, dxdv, t.x, t.z, t.y, idi, (uint32_t)__float_as_int(vj.w),
                                            B.invrootdt, K, OBS ? &pe : nullptr);
                const float Fx = s * dx, Fy = s * dy, Fz = s * dz;
                f[0] += Fx; f[1] += Fy; f[2] += Fz;
                if (OBS) {
                    vir[0] += 0.5f * dx * Fx; vir[1] += 0.5f * dy * Fx; vir[2] += 0.5f * dz * Fx;
                    vir[3] += 0.5f * dx * Fy; vir[4] += 0.5f * dy * Fy; vir[5] += 0.5f * dz * Fy;
                    vir[6] += 0.5f * dx * Fz; vir[7] += 0.5f * dy * Fz; vir[8] += 0.5f * dz * Fz;
                    vir[9] += 0.5f * pe;
                }
            }
        }
    }
}

template <bool OBS>
__global__ __launch_bounds__(256) void k_pair_cells(PairArgs A, int per_xcd)
{
    __shared__ Smem S;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int gbrick = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (gbrick >= A.nbricks) return;   // whole workgroup exits together
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
    {
        const float* g = A.table + box * kTableStride;
        for (int q = tid; q < kTypeStride; q += 256)
            S.tab[q] = make_float4(g[q], g[kTypeStride + q], g[2 * kTypeStride + q], 0.f);
    }
    float vir[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (nhome > kMaxHome) {
        __syncthreads();
        for (int i = tid; i < nhome; i += 256) {
            float f[3];
            fallback_bead<OBS>(A, B, K, base + i, S.tab, vir, f);
            A.frc[base + i] = make_float4(f[0], f[1], f[2], 0.f);
        }
    } else {
        for (int i = tid; i < nhome; i += 256) { S.fx[i] = 0.f; S.fy[i] = 0.f; S.fz[i] = 0.f; }
        __syncthreads();
        WaveMem& W = S.w[wave];
        volatile uint64_t* queue = W.queue;
        int head = 0, tail = 0;   // wave-uniform ring positions
        // wavefront w takes home cells w, w + 4, ... of the brick
        for (int h = wave; h < kBrickCells; h += kWaves) {
            const int cx = X0 + (h & 3), cy = Y0 + ((h >> 2) & 3), cz = Z0 + (h >> 4);
            if (cx >= B.n[0] || cy >= B.n[1] || cz >= B.n[2]) continue;   // beyond the box
            const int i0 = A.cstart[kb + h], i1 = A.cstart[kb + h + 1];
            if (i1 == i0) continue;
            // 1. neighbour cell ranges and their concatenation
            int c0 = 0, len = 0;
            if (lane < 27) {
                int nx = cx + lane % 3 - 1, ny = cy + (lane / 3) % 3 - 1, nz = cz + lane / 9 - 1;
                nx += (nx < 0) ? B.n[0] : ((nx >= B.n[0]) ? -B.n[0] : 0);
                ny += (ny < 0) ? B.n[1] : ((ny >= B.n[1]) ? -B.n[1] : 0);
                nz += (nz < 0) ? B.n[2] : ((nz >= B.n[2]) ? -B.n[2] : 0);
                const int k = B.koff + cell_key(nx, ny, nz, B);
                c0 = A.cstart[k];
                len = A.cstart[k + 1] - c0;
            }
            int pre = len;
#pragma unroll
            for (int off = 1; off < 32; off <<= 1) {
                const int y = __shfl_up(pre, off, 64);
                if (lane >= off) pre += y;
            }
            const int M = __builtin_amdgcn_readlane(pre, 26);
            if (M > kMaxM) {
                // oversized list (dense region): this cell's beads take the global loop
                for (int i = i0 + lane; i < i1; i += 64) {
                    float f[3];
                    fallback_bead<OBS>(A, B, K, i, S.tab, vir, f);
                    S.fx[i - base] = f[0]; S.fy[i - base] = f[1]; S.fz[i - base] = f[2];
                }
                continue;
            }
            __builtin_amdgcn_wave_barrier();
            if (lane < 27) {
                const int d0 = pre - len;
                for (int q = 0; q < len; ++q) {
                    W.slot[d0 + q] = c0 + q;
                    W.code[d0 + q] = (uint8_t)lane;
                }
            }
            __builtin_amdgcn_wave_barrier();
            // 2.-4. candidates 64 at a time against every bead of the cell
            for (int m0 = 0; m0 < M; m0 += 64) {
                const int m = m0 + lane;
                int j = -1, nb = 0;
                float4 pj = make_float4(1e30f, 1e30f, 1e30f, 0.f);
                if (m < M) {
                    j = W.slot[m];
                    nb = W.code[m];
                    pj = A.pos[j];
                }
                // same operand order as the evaluation: (pi - pj) - offset * w is exactly
                // antisymmetric between the two sides of a pair
                const float sx = (float)(nb % 3 - 1) * B.w[0], sy = (float)((nb / 3) % 3 - 1) * B.w[1],
                            sz = (float)(nb / 9 - 1) * B.w[2];
                const uint64_t tag = ((uint64_t)(uint32_t)j << 32) | ((uint64_t)nb << 16);
                for (int i = i0; i < i1; ++i) {
                    const float4 pi = A.pos[i];   // wave-uniform address
                    const float dx = (pi.x - pj.x) - sx, dy = (pi.y - pj.y) - sy, dz = (pi.z - pj.z) - sz;
                    const float dr2 = dx * dx + dy * dy + dz * dz;
                    const bool pass = dr2 < 1.0f && dr2 > 1e-18f && j != i;
                    const uint64_t mk = __ballot(pass);
                    if (pass)
                        queue[(tail + lane_prefix(mk)) & (kQueue - 1)] = tag | (uint64_t)(uint32_t)(i - base);
                    tail += __popcll(mk);
                    if (tail - head >= 64) {
                        __builtin_amdgcn_wave_barrier();
                        eval_queued<OBS>(S, queue[(head + lane) & (kQueue - 1)], base, A, B, K, vir);
                        head += 64;
                        __builtin_amdgcn_wave_barrier();
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < tail - head) eval_queued<OBS>(S, queue[(head + lane) & (kQueue - 1)], base, A, B, K, vir);
        __syncthreads();
        for (int i = tid; i < nhome; i += 256) A.frc[base + i] = make_float4(S.fx[i], S.fy[i], S.fz[i], 0.f);
    }
    if (OBS) {
        __syncthreads();
        __shared__ double red[10][kWaves];
        for (int c = 0; c < 10; ++c) {
            const double s = wave_sum((double)vir[c]);
            if (lane == 0) red[c][wave] = s;
        }
        __syncthreads();
        if (tid < 10) {
            double t = 0.0;
            for (int w = 0; w < kWaves; ++w) t += red[tid][w];
            atomicAdd(A.acc + box * kObsStride + kObsVir + tid, t);
        }
    }
}

}  // namespace

hipError_t launch_pair_cells(const PairArgs& a, bool obs, hipStream_t s)
{
    const int per_xcd = (a.nbricks + 7) / 8;
    const dim3 grid(8 * per_xcd);
    if (obs) hipLaunchKernelGGL((k_pair_cells<true>), grid, dim3(256), 0, s, a, per_xcd);
    else hipLaunchKernelGGL((k_pair_cells<false>), grid, dim3(256), 0, s, a, per_xcd);
    return hipGetLastError();
}

}  // namespace dpd
