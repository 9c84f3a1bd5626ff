// pair_stream.hip -- cell-stream DPD pair kernel for gfx950.
//
// One workgroup per 4x4x4-cell brick.  The brick and its one-cell halo (6x6x6 cells,
// periodic) are staged into LDS in one x-fastest cell order.  The work is organised so
// that the candidate test -- the bulk of the instructions, ~60 candidates per bead for
// ~6 interacting pairs -- costs as little as possible and runs on full wavefronts:
//   * during staging the workgroup builds, for every home cell, the list of the staged
//     beads it must test (u16 indices in LDS): all halo neighbour cells, the forward
//     half of the home neighbour cells (staged cell index above its own) and its own
//     cell; a pair of two home beads is thus listed once (half shell), a pair with a
//     halo bead once per brick owning one of its beads;
//   * a wavefront takes one home cell at a time (dynamic, from an LDS counter), holds up
//     to 64 list entries in registers (one per lane; sequential, conflict-free LDS reads)
//     and tests them against each bead of the cell in turn (its position broadcast), so
//     LDS is read once per (home cell, candidate) rather than once per (bead, candidate):
//     random 16-byte LDS gathers cost ~20 cycles per wavefront from bank conflicts and
//     bound any scheme that gathers per pair;
//   * candidates with r < 1 (and, in the own cell, j above i) are compacted into a
//     per-wavefront LDS queue (ballot + mbcnt) and evaluated 64 at a time, so the pair
//     force and its Philox draw run on full wavefronts;
//   * a pair of two home beads adds both forces in LDS (Newton's third law); a pair with
//     a halo bead keeps only the home bead's force (the halo bead's brick evaluates the
//     same pair for its own bead), so forces never leave the workgroup through global
//     atomics and one coalesced store writes the brick's forces.
// LDS holds up to kCap staged beads (~5 standard deviations above a brick at density
// 4.6, the dense phase of a condensate) and kJList list entries; a brick beyond either
// falls back to the thread-per-bead 27-cell loop over global memory.  ~73 KB of LDS per
// workgroup: two 1024-thread workgroups (8 wavefronts per SIMD, <= 64 VGPRs) or two
// 512-thread workgroups per CU.
//
// Workgroups are mapped to bricks XCD-aware: the dispatcher places workgroup b on XCD
// b % 8, so brick = (b % 8) * ceil(nb/8) + b / 8 gives each XCD a contiguous slab of
// bricks whose halos mostly hit its own L2.
#include "pair_kernels.hpp"

namespace dpd {

namespace {

constexpr int kMaxWaves = 16;   // up to 1024 threads per workgroup
constexpr int kCap = 1152;      // staged beads (brick + halo)
constexpr int kMaxHome = 384;   // beads in the brick
constexpr int kStaged = 216;    // 6x6x6 staged cells
constexpr int kQueue = 128;     // per-wavefront pair queue
constexpr int kJList = 8192;    // candidate-list entries of all home cells

struct Smem {
    float4 pos[kCap];   // brick-centred position xyz; w = type | (home index + 1) << 8, 0 for halo
    float4 vp[kCap];    // predicted velocity xyz, w = id
    float fx[kMaxHome], fy[kMaxHome], fz[kMaxHome];
    float4 tab[kTypeStride];   // (a, gamma, sqrt gamma, 0) per type pair
    int start[kStaged + 1];    // staged index of each staged cell's first bead
    int gstart[kStaged];       // global slot of each staged cell's first bead
    int jstart[kBrickCells + 1];   // first list entry of each home cell
    uint16_t jlist[kJList];        // staged indices to test, per home cell
    uint8_t cell_of[kCap];         // staged cell of each staged bead
    union {
        uint32_t queue[kMaxWaves][kQueue];
        uint16_t off[kBrickCells][32];   // list offset of each (home cell, neighbour) while building
        double red[10][kMaxWaves];
    } u;
    int wsum[kMaxWaves];
    int next;   // next home cell to take
};

// staged cell index of brick-local cell coords (lx, ly, lz) in [-1, 4]
__device__ inline int sidx(int lx, int ly, int lz) { return (lx + 1) + 6 * ((ly + 1) + 6 * (lz + 1)); }

template <bool OBS>
__device__ inline void eval_pair(Smem& S, uint32_t packed, float invrootdt, const RngKey& K, float* vir)
{
    const int li = (int)(packed & 0xFFFFu), lj = (int)(packed >> 16);
    const float4 pi = S.pos[li], pj = S.pos[lj];
    const float4 vi = S.vp[li], vj = S.vp[lj];
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float dr2 = dx * dx + dy * dy + dz * dz;
    const int wi = __float_as_int(pi.w), wj = __float_as_int(pj.w);
    const float4 t = S.tab[(wi & 0xFF) * kMaxTypes + (wj & 0xFF)];
    const float dxdv = dx * (vi.x - vj.x) + dy * (vi.y - vj.y) + dz * (vi.z - vj.z);
    float pe;
    const float s = pair_scalar(dr2, dxdv, t.x, t.z, t.y, (uint32_t)__float_as_int(vi.w),
                                (uint32_t)__float_as_int(vj.w), invrootdt, K, OBS ? &pe : nullptr);
    const float Fx = s * dx, Fy = s * dy, Fz = s * dz;
    const int hi = (wi >> 8) - 1, hj = (wj >> 8) - 1;
    atomicAdd(&S.fx[hi], Fx);
    atomicAdd(&S.fy[hi], Fy);
    atomicAdd(&S.fz[hi], Fz);
    if (hj >= 0) {
        atomicAdd(&S.fx[hj], -Fx);
        atomicAdd(&S.fy[hj], -Fy);
        atomicAdd(&S.fz[hj], -Fz);
    }
    if (OBS) {
        // a halo pair is evaluated by both bricks: count half from each
        const float wgt = hj >= 0 ? 1.0f : 0.5f;
        vir[0] += wgt * dx * Fx; vir[1] += wgt * dy * Fx; vir[2] += wgt * dz * Fx;
        vir[3] += wgt * dx * Fy; vir[4] += wgt * dy * Fy; vir[5] += wgt * dz * Fy;
        vir[6] += wgt * dx * Fz; vir[7] += wgt * dy * Fz; vir[8] += wgt * dz * Fz;
        vir[9] += wgt * pe;
    }
}

// thread-per-bead 27-cell loop over global memory (capacity fallback), same arithmetic
template <bool OBS>
__device__ inline void fallback_bead(const PairArgs& A, const BoxDev& B, const RngKey& K, int i, const Smem& S,
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
                        const float4 t = S.tab[ti * kMaxTypes + bead_type(__float_as_int(pj.w))];
                        const float dxdv = ddx * (vi.x - vj.x) + ddy * (vi.y - vj.y) + ddz * (vi.z - vj.z);
                        float pe;
                        const float s = pair_scalar(dr2, dxdv, t.x, t.z, t.y, idi, (uint32_t)__float_as_int(vj.w),
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

template <int kThreads>
__device__ inline void reduce_virial(Smem& S, const float* vir, double* acc)
{
    constexpr int kWaves = kThreads / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __syncthreads();   // queues are dead: their LDS is reused
    for (int c = 0; c < 10; ++c) {
        const double s = wave_sum((double)vir[c]);
        if (lane == 0) S.u.red[c][wave] = s;
    }
    __syncthreads();
    if (tid < 10) {
        double t = 0.0;
        for (int w = 0; w < kWaves; ++w) t += S.u.red[tid][w];
        atomicAdd(acc + tid, t);
    }
}

// Length of home cell h's list part for neighbour offset c (0..26): halo cells and the
// own cell always, home cells only above h in staged order.
__device__ inline int list_len(const Smem& S, int h, int c, int hx, int hy, int hz)
{
    const int cx = h & 3, cy = (h >> 2) & 3, cz = h >> 4;
    const int nx = cx + c % 3 - 1, ny = cy + (c / 3) % 3 - 1, nz = cz + c / 9 - 1;
    const int hc = sidx(cx, cy, cz), sc = sidx(nx, ny, nz);
    const bool home = nx >= 0 && nx < hx && ny >= 0 && ny < hy && nz >= 0 && nz < hz;
    return (!home || sc >= hc) ? S.start[sc + 1] - S.start[sc] : 0;
}

template <int kThreads, int kWavesPerEU, bool OBS>
__global__ __attribute__((amdgpu_flat_work_group_size(1, kThreads), amdgpu_waves_per_eu(kWavesPerEU)))
void k_pair_stream(PairArgs A, int per_xcd)
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
    // home cells that exist along each axis (partial bricks at the far faces)
    const int hx = min(kBrick, B.n[0] - X0), hy = min(kBrick, B.n[1] - Y0), hz = min(kBrick, B.n[2] - Z0);
    const int kb = B.koff + brick * kBrickCells;
    const int base = A.cstart[kb];
    const int nhome = A.cstart[kb + kBrickCells] - base;
    {
        const float* g = A.table + box * kTableStride;
        for (int q = tid; q < kTypeStride; q += kThreads)
            S.tab[q] = make_float4(g[q], g[kTypeStride + q], g[2 * kTypeStride + q], 0.f);
    }

    // ---- staged cells: global start and count of every cell within one cell of a home cell
    int cnt = 0;
    if (tid < kStaged) {
        const int lx = tid % 6 - 1, ly = (tid / 6) % 6 - 1, lz = tid / 36 - 1;
        int gs = 0;
        if (lx <= hx && ly <= hy && lz <= hz) {
            const int gx = X0 + lx, gy = Y0 + ly, gz = Z0 + lz;
            const int wx = gx < 0 ? gx + B.n[0] : (gx >= B.n[0] ? gx - B.n[0] : gx);
            const int wy = gy < 0 ? gy + B.n[1] : (gy >= B.n[1] ? gy - B.n[1] : gy);
            const int wz = gz < 0 ? gz + B.n[2] : (gz >= B.n[2] ? gz - B.n[2] : gz);
            const int k = B.koff + cell_key(wx, wy, wz, B);
            gs = A.cstart[k];
            cnt = A.cstart[k + 1] - gs;
        }
        S.gstart[tid] = gs;
    }
    // exclusive prefix of the counts in staged order
    {
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        if (lane == 63) S.wsum[wave] = incl;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; ++w) before += S.wsum[w];
        if (tid < kStaged) S.start[tid] = before + incl - cnt;
        if (tid == kStaged - 1) S.start[kStaged] = before + incl;
        if (tid == 0) S.next = 0;
    }
    __syncthreads();
    const int nstaged = S.start[kStaged];

    // ---- candidate lists.  Pass 1: the list length of every (home cell, neighbour cell)
    for (int task = tid; task < kBrickCells * 27; task += kThreads) {
        const int h = task / 27, c = task - h * 27;
        const bool present = (h & 3) < hx && ((h >> 2) & 3) < hy && (h >> 4) < hz;
        S.u.off[h][c] = (uint16_t)(present ? list_len(S, h, c, hx, hy, hz) : 0);
    }
    __syncthreads();
    // pass 2: per home cell, lengths -> offsets (entry 27 = the cell's list length), and
    // the first entry of each home cell's list (one wavefront)
    if (tid < kBrickCells) {
        const int h = tid;
        int len[27];
#pragma unroll
        for (int c = 0; c < 27; ++c) len[c] = S.u.off[h][c];
        int run = 0;
#pragma unroll
        for (int c = 0; c < 27; ++c) {
            S.u.off[h][c] = (uint16_t)run;
            run += len[c];
        }
        S.u.off[h][27] = (uint16_t)run;
        int incl = run;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        S.jstart[h] = incl - run;
        if (h == kBrickCells - 1) S.jstart[kBrickCells] = incl;
    }
    __syncthreads();
    const bool fits = nhome <= kMaxHome && nstaged <= kCap && S.jstart[kBrickCells] <= kJList;
    if (!fits) {
        float vir[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = tid; i < nhome; i += kThreads) fallback_bead<OBS>(A, B, K, base + i, S, vir);
        if (OBS) reduce_virial<kThreads>(S, vir, A.acc + box * kObsStride + kObsVir);
        return;
    }
    // pass 3: fill the lists, one thread per (home cell, neighbour cell)
    for (int task = tid; task < kBrickCells * 27; task += kThreads) {
        const int h = task / 27, c = task - h * 27;
        const int o = S.u.off[h][c];
        const int len = S.u.off[h][c + 1] - o;
        if (len == 0) continue;
        const int sc = sidx((h & 3) + c % 3 - 1, ((h >> 2) & 3) + (c / 3) % 3 - 1, (h >> 4) + c / 9 - 1);
        const int s0 = S.start[sc];
        const int d0 = S.jstart[h] + o;
        for (int k = 0; k < len; ++k) S.jlist[d0 + k] = (uint16_t)(s0 + k);
    }
    if (tid < kStaged) {
        const int s0 = S.start[tid];
        for (int q = 0; q < cnt; ++q) S.cell_of[s0 + q] = (uint8_t)tid;
    }
    for (int i = tid; i < nhome; i += kThreads) { S.fx[i] = 0.f; S.fy[i] = 0.f; S.fz[i] = 0.f; }
    __syncthreads();

    // ---- stage beads: brick-centred positions (cell - 2) * w + offset
    for (int q = tid; q < nstaged; q += kThreads) {
        const int c = S.cell_of[q];
        const int slot = S.gstart[c] + (q - S.start[c]);
        const float4 p = A.pos[slot];
        const int lx = c % 6 - 1, ly = (c / 6) % 6 - 1, lz = c / 36 - 1;
        const bool home = lx >= 0 && lx < hx && ly >= 0 && ly < hy && lz >= 0 && lz < hz;
        const int w = bead_type(__float_as_int(p.w)) | (home ? (slot - base + 1) << 8 : 0);
        S.pos[q] = make_float4((float)(lx - 2) * B.w[0] + p.x, (float)(ly - 2) * B.w[1] + p.y,
                               (float)(lz - 2) * B.w[2] + p.z, __int_as_float(w));
        S.vp[q] = A.vp[slot];
    }
    __syncthreads();   // also retires the list offsets: the queues reuse their LDS

    // ---- candidates: a wavefront takes one home cell at a time (dynamic, from an LDS
    // counter), loads up to 64 entries of its list into registers (one per lane, a
    // sequential read of the list and of the staged beads it names) and then, for each
    // bead i of the cell, tests all of them against i's broadcast position, so each staged
    // position is read once per home cell rather than once per (i, j) candidate; pairs in
    // range are compacted per wavefront
    float vir[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    volatile uint32_t* queue = S.u.queue[wave];
    int qn = 0;   // wave-uniform queue length
    for (;;) {
        int h = 0;
        if (lane == 0) h = atomicAdd(&S.next, 1);
        h = __builtin_amdgcn_readfirstlane(h);
        if (h >= kBrickCells) break;
        const int hc = sidx(h & 3, (h >> 2) & 3, h >> 4);
        const int i0 = S.start[hc];
        const int i1 = S.start[hc + 1];
        const int jb = S.jstart[h];
        const int M = S.jstart[h + 1] - jb;   // 0 for absent home cells
        if (i1 == i0) continue;
        for (int m0 = 0; m0 < M; m0 += 64) {
            const int m = m0 + lane;
            int j = 0;
            float4 pj = make_float4(1e30f, 1e30f, 1e30f, 0.f);
            if (m < M) {
                j = S.jlist[jb + m];
                pj = S.pos[j];
            }
            // own cell: each pair once (j above i); the list's other home cells are forward
            const bool own = j >= i0 && j < i1;
            for (int li = i0; li < i1; ++li) {
                const float4 pi = S.pos[li];
                const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                const float dr2 = dx * dx + dy * dy + dz * dz;
                const bool pass = dr2 < 1.0f && dr2 > 1e-18f && (!own || j > li);
                const uint64_t mk = __ballot(pass);
                if (pass) queue[qn + lane_prefix(mk)] = (uint32_t)li | ((uint32_t)j << 16);
                qn += __popcll(mk);
                if (qn >= 64) {
                    __builtin_amdgcn_wave_barrier();
                    const uint32_t pk = queue[lane];
                    eval_pair<OBS>(S, pk, B.invrootdt, K, vir);
                    const uint32_t rest = (lane + 64 < qn) ? queue[lane + 64] : 0u;
                    __builtin_amdgcn_wave_barrier();
                    if (lane + 64 < qn) queue[lane] = rest;
                    qn -= 64;
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < qn) eval_pair<OBS>(S, queue[lane], B.invrootdt, K, vir);
    __syncthreads();

    for (int i = tid; i < nhome; i += kThreads) A.frc[base + i] = make_float4(S.fx[i], S.fy[i], S.fz[i], 0.f);
    if (OBS) reduce_virial<kThreads>(S, vir, A.acc + box * kObsStride + kObsVir);
}

}  // namespace

hipError_t launch_pair_stream(const PairArgs& a, bool obs, hipStream_t s, int threads)
{
    const int per_xcd = (a.nbricks + 7) / 8;
    const dim3 grid(8 * per_xcd);
    if (threads == 1024) {
        if (obs) hipLaunchKernelGGL((k_pair_stream<1024, 8, true>), grid, dim3(1024), 0, s, a, per_xcd);
        else hipLaunchKernelGGL((k_pair_stream<1024, 8, false>), grid, dim3(1024), 0, s, a, per_xcd);
    } else {
        if (obs) hipLaunchKernelGGL((k_pair_stream<512, 4, true>), grid, dim3(512), 0, s, a, per_xcd);
        else hipLaunchKernelGGL((k_pair_stream<512, 4, false>), grid, dim3(512), 0, s, a, per_xcd);
    }
    return hipGetLastError();
}

}  // namespace dpd
