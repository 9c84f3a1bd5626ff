// pair_kernels.hpp -- DPD pair-force kernels (CCNTCell::UpdateForce, src/CNTCell.cpp:877-1572).
//
// Two kernels over the same cell-sorted SoA layout:
//   DPD_PAIR_FULLSHELL  one thread per bead visits all 27 neighbour cells and keeps only
//                       its own force (no scatter, no atomics).  Simple and exact; the
//                       cross-check for the brick kernel.
//   DPD_PAIR_BRICK      one workgroup per 4x4x4-cell brick: the brick and its one-cell
//                       halo are staged in LDS, candidate pairs are filtered by r < 1 and
//                       compacted per wavefront, and the surviving pairs are evaluated
//                       64 at a time with both forces accumulated in LDS (pair_brick.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include "dpd_engine.h"
#include "engine_common.hpp"

namespace dpd {

struct PairArgs {
    int n;
    const float4* pos;     // offset xyz, type bits (sorted)
    const float4* vp;      // predicted velocity xyz, id bits (sorted)
    const int* key;        // cell key per slot
    const int* cstart;     // [nkeys+1] first slot of each cell key
    const float* table;    // a | gamma | sqrt(gamma), ntypes^2 each
    float4* frc;           // out: pair force per slot
    double* acc;           // out (OBS): [0..8] virial sum dx_b F_a, [9] pair PE
    int nbricks;
};

// Full-shell kernel: thread per bead.  Each pair is evaluated from both sides with the
// same operands (dx negated), so the two forces are exact negatives of each other;
// virial and PE are halved.
template <bool OBS>
__global__ __launch_bounds__(256) void k_pair_fullshell(PairArgs A, BoxDev B)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double vir[10];
    if (OBS)
        for (int q = 0; q < 10; ++q) vir[q] = 0.0;
    if (i < A.n) {
        const int nt = B.ntypes;
        const float4 pi = A.pos[i];
        const float4 vi = A.vp[i];
        const int ti = __float_as_int(pi.w);
        const uint32_t idi = (uint32_t)__float_as_int(vi.w);
        int c[3];
        key_cell(A.key[i], B, c);
        float fx = 0.f, fy = 0.f, fz = 0.f;
        float sv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        float spe = 0.f;
        for (int dz = -1; dz <= 1; ++dz) {
            int cz = c[2] + dz;
            cz += (cz < 0) ? B.n[2] : ((cz >= B.n[2]) ? -B.n[2] : 0);
            for (int dy = -1; dy <= 1; ++dy) {
                int cy = c[1] + dy;
                cy += (cy < 0) ? B.n[1] : ((cy >= B.n[1]) ? -B.n[1] : 0);
                for (int dx = -1; dx <= 1; ++dx) {
                    int cx = c[0] + dx;
                    cx += (cx < 0) ? B.n[0] : ((cx >= B.n[0]) ? -B.n[0] : 0);
                    const int k = cell_key(cx, cy, cz, B);
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
                            const int tj = __float_as_int(pj.w);
                            const int tt = ti * nt + tj;
                            const float a = A.table[tt], g = A.table[nt * nt + tt], sg = A.table[2 * nt * nt + tt];
                            const float dxdv = ddx * (vi.x - vj.x) + ddy * (vi.y - vj.y) + ddz * (vi.z - vj.z);
                            float pe;
                            const float s = pair_scalar(dr2, dxdv, a, sg, g, idi, (uint32_t)__float_as_int(vj.w), B,
                                                        OBS ? &pe : nullptr);
                            const float Fx = s * ddx, Fy = s * ddy, Fz = s * ddz;
                            fx += Fx; fy += Fy; fz += Fz;
                            if (OBS) {
                                sv[0] += ddx * Fx; sv[1] += ddy * Fx; sv[2] += ddz * Fx;
                                sv[3] += ddx * Fy; sv[4] += ddy * Fy; sv[5] += ddz * Fy;
                                sv[6] += ddx * Fz; sv[7] += ddy * Fz; sv[8] += ddz * Fz;
                                spe += pe;
                            }
                        }
                    }
                }
            }
        }
        A.frc[i] = make_float4(fx, fy, fz, 0.f);
        if (OBS) {
            for (int q = 0; q < 9; ++q) vir[q] = 0.5 * (double)sv[q];
            vir[9] = 0.5 * (double)spe;
        }
    }
    if (OBS) {
        __shared__ double red[10][4];
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        for (int q = 0; q < 10; ++q) {
            const double s = wave_sum(vir[q]);
            if (lane == 0) red[q][wid] = s;
        }
        __syncthreads();
        if (threadIdx.x < 10) atomicAdd(A.acc + threadIdx.x, red[threadIdx.x][0] + red[threadIdx.x][1] +
                                                                   red[threadIdx.x][2] + red[threadIdx.x][3]);
    }
}

inline int default_pair_kernel() { return DPD_PAIR_FULLSHELL; }

hipError_t launch_pair_brick(const PairArgs& a, const BoxDev& B, bool obs, hipStream_t s);

inline hipError_t launch_pair_kernel(int which, const PairArgs& a, const BoxDev& B, bool obs, hipStream_t s)
{
    if (which == DPD_PAIR_BRICK) return launch_pair_brick(a, B, obs, s);
    const int g = (a.n + 255) / 256;
    if (obs) hipLaunchKernelGGL((k_pair_fullshell<true>), dim3(g), dim3(256), 0, s, a, B);
    else hipLaunchKernelGGL((k_pair_fullshell<false>), dim3(g), dim3(256), 0, s, a, B);
    return hipGetLastError();
}

}  // namespace dpd
