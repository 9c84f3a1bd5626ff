// slice_stress.hip -- z-slice stress profile of one box (CMonitor::AddBeadStress,
// src/Monitor.cpp:1900-2104, CMonitor::AddBondStress, :2283-2518, and
// CMonitor::AddBondPairStress, :2520-2739).
//
// The reference zeroes its slice tables every step and adds F (x) dx of every pair (and,
// when a stress analysis asks for it, of every bond) into the z-slices spanned by the two
// beads.  Here the tables are produced on demand for the engine's current state: the pair
// forces are re-evaluated with the same arithmetic and the same counter-based random
// draws as the step kernels, so right after dpd_step they are the last step's forces.
// Thread per bead walks the 27 neighbour cells (each pair is seen from both sides and
// adds half of its share under each ordered type pair); thread per bond adds the bond's
// share under (head type, tail type).  FP64 tables are accumulated in LDS per workgroup
// when they fit, and with global FP64 atomics otherwise.
#include "pair_kernels.hpp"
#include "slice_stress.hpp"

namespace dpd {

namespace {

// the z-slice split of one interaction (see orc_slice_add for the rule); tab is
// [nslice][nidx][9]
__device__ inline void slice_add(double* tab, int nslice, double Lz, int idx, int nidx, double z1, double z2,
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
    // z is wrapped into [0, Lz): the slice indices are in range; the clamp only guards
    // the table against a position rounded onto Lz
    int iu = min(max((int)zu, 0), nslice - 1), il = min(max((int)zl, 0), nslice - 1);
    const double ru = zu - (double)iu, rl = zl - (double)il;
    auto add = [&](int sl, double share) {
        double* t = tab + ((size_t)sl * nidx + idx) * 9;
        for (int q = 0; q < 9; ++q) atomicAdd(t + q, fdx[q] * share);
    };
    if (iu == il) {
        if (zu != zl || ru > 0.00001) {
            add(il, 1.0);
        } else {
            add(iu, 0.5);
            if (--il < 0) il += nslice;
            add(il, 0.5);
        }
    } else {
        const double inner = width / zd;
        if (iu > il) {
            for (int sl = il + 1; sl <= iu - 1; ++sl) add(sl, inner);
        } else {
            for (int sp = il + 1; sp <= nslice + iu - 1; ++sp) add(sp >= nslice ? sp - nslice : sp, inner);
        }
        add(il, rl > 0.000001 ? (1.0 - rl) * inner : inner);
        if (ru > 0.000001) add(iu, ru * inner);
    }
}

__device__ inline double wrapped_z(const SliceArgs& A, int slot)
{
    int c[3];
    key_cell(A.key[slot] - A.B.koff, A.B, c);
    return (double)c[2] * A.B.wd[2] + (double)A.pos[slot].z;
}

template <bool SHARED>
__global__ __launch_bounds__(256) void k_slice_pairs(SliceArgs A, double* __restrict__ out)
{
    extern __shared__ double lt[];
    const int ntab = A.nslice * A.nidx * 9;
    double* tab = SHARED ? lt : out;
    if (SHARED) {
        for (int q = threadIdx.x; q < ntab; q += blockDim.x) lt[q] = 0.0;
        __syncthreads();
    }
    const BoxDev& B = A.B;
    const double Lz = (double)B.n[2] * B.wd[2];
    const RngKey K = rng_key(B, *A.step);
    // the box's beads are the slots of its cell keys
    const int s0 = A.cstart[B.koff], s1 = A.cstart[B.koff + B.ngroups * kGroupKeys];
    for (int i = s0 + blockIdx.x * blockDim.x + threadIdx.x; i < s1; i += gridDim.x * blockDim.x) {
        const float4 pi = A.pos[i], vi = A.vp[i];
        const int ti = bead_type(__float_as_int(pi.w));
        const uint32_t idi = (uint32_t)__float_as_int(vi.w);
        int c[3];
        key_cell(A.key[i] - B.koff, B, c);
        const double zi = (double)c[2] * B.wd[2] + (double)pi.z;
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
                if (!(dr2 < 1.0f && dr2 > 1e-18f)) continue;
                const float4 vj = A.vp[j];
                const int tj = bead_type(__float_as_int(pj.w));
                const int tt = ti * kMaxTypes + tj;
                const float dxdv = dx * (vi.x - vj.x) + dy * (vi.y - vj.y) + dz * (vi.z - vj.z);
                const float s = pair_scalar(dr2, dxdv, A.table[tt], A.table[2 * kTypeStride + tt],
                                            A.table[kTypeStride + tt], idi, (uint32_t)__float_as_int(vj.w),
                                            B.invrootdt, K, nullptr);
                const double F[3] = {(double)(s * dx), (double)(s * dy), (double)(s * dz)};
                const double d[3] = {(double)dx, (double)dy, (double)dz};
                const double zj = (double)cz * B.wd[2] + (double)pj.z;
                slice_add(tab, A.nslice, Lz, ti * A.ntypes + tj, A.nidx, zi, zj, F, d, 0.5);
            }
        }
    }
    if (SHARED) {
        __syncthreads();
        for (int q = threadIdx.x; q < ntab; q += blockDim.x)
            if (lt[q] != 0.0) atomicAdd(out + q, lt[q]);
    }
}

// bonds of the box: head - tail (unwrapped), force on the head, FP64
__global__ __launch_bounds__(256) void k_slice_bonds(SliceArgs A, double* __restrict__ out)
{
    const int q = A.b0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= A.b1) return;
    const BoxDev& B = A.B;
    const int2 hb = A.bond_bead[q];
    const int sh = A.slot_of[hb.x], st = A.slot_of[hb.y];
    double uh[3], ut[3];
    for (int e = 0; e < 2; ++e) {
        const int slot = e ? st : sh;
        int c[3];
        key_cell(A.key[slot] - B.koff, B, c);
        const float4 p = A.pos[slot];
        const int4 a = A.img[e ? hb.y : hb.x];
        double* u = e ? ut : uh;
        u[0] = (double)(c[0] + a.x * B.n[0]) * B.wd[0] + (double)p.x;
        u[1] = (double)(c[1] + a.y * B.n[1]) * B.wd[1] + (double)p.y;
        u[2] = (double)(c[2] + a.z * B.n[2]) * B.wd[2] + (double)p.z;
    }
    const double2 par = A.bond_par[q];
    const double d[3] = {uh[0] - ut[0], uh[1] - ut[1], uh[2] - ut[2]};
    const double L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double s = par.x * (par.y - L) / L;
    const double F[3] = {s * d[0], s * d[1], s * d[2]};
    const int th = bead_type(__float_as_int(A.pos[sh].w)), tt = bead_type(__float_as_int(A.pos[st].w));
    slice_add(out, A.nslice, (double)B.n[2] * B.wd[2], th * A.ntypes + tt, A.nidx, wrapped_z(A, sh),
              wrapped_z(A, st), F, d, 1.0);
}

__device__ inline void unwrapped(const SliceArgs& A, int bead, int slot, double u[3])
{
    const BoxDev& B = A.B;
    int c[3];
    key_cell(A.key[slot] - B.koff, B, c);
    const float4 p = A.pos[slot];
    const int4 a = A.img[bead];
    u[0] = (double)(c[0] + a.x * B.n[0]) * B.wd[0] + (double)p.x;
    u[1] = (double)(c[1] + a.y * B.n[1]) * B.wd[1] + (double)p.y;
    u[2] = (double)(c[2] + a.z * B.n[2]) * B.wd[2] + (double)p.z;
}

// CMonitor::AddBondPairStress (src/Monitor.cpp:2520-2739), thread per bond pair: the bond
// pair's forces (FP64, as in the step), then its three bead pairs first tail -> first
// head -> second head -> first tail, each adding 0.33333333 (F[j+1] - F[j]) (x)
// (u[j+1] - u[j]) over the slices it spans, keyed by the bead types (second head, first
// head, first tail) where the reference keys by bond-pair type.
__global__ __launch_bounds__(256) void k_slice_bondpairs(SliceArgs A, double* __restrict__ out)
{
    const int q = A.a0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= A.a1) return;
    const int4 abc = A.ang_bead[q];
    const int sa = A.slot_of[abc.x], sb = A.slot_of[abc.y], sc = A.slot_of[abc.z];
    double ua[3], ub[3], uc[3], f[3][3];
    unwrapped(A, abc.x, sa, ua);
    unwrapped(A, abc.y, sb, ub);
    unwrapped(A, abc.z, sc, uc);
    angle_forces(ua, ub, uc, A.ang_par[q], f, nullptr);   // f[0] on c, f[1] on b, f[2] on a
    const double* u[4] = {uc, ub, ua, uc};
    const double* F[4] = {f[0], f[1], f[2], f[0]};
    const int slot[4] = {sc, sb, sa, sc};
    const int nt = A.ntypes;
    const int idx = (bead_type(__float_as_int(A.pos[sa].w)) * nt + bead_type(__float_as_int(A.pos[sb].w))) * nt +
                    bead_type(__float_as_int(A.pos[sc].w));
    const double Lz = (double)A.B.n[2] * A.B.wd[2];
    for (int j = 0; j < 3; ++j) {
        double dx[3], fj[3];
        for (int d = 0; d < 3; ++d) {
            dx[d] = u[j + 1][d] - u[j][d];
            fj[d] = 0.33333333 * (F[j + 1][d] - F[j][d]);
        }
        slice_add(out, A.nslice, Lz, idx, nt * nt * nt, wrapped_z(A, slot[j + 1]), wrapped_z(A, slot[j]), fj, dx, 1.0);
    }
}

// The 13 half-shell neighbours of the reference's cells (CSimBox::MakeCNTCells,
// src/SimBox.cpp:830-867) as bits of the neighbour code (dx+1) + 3 (dy+1) + 9 (dz+1):
// (1,-1,0), (1,0,0), (0,1,0), (1,1,0) and the nine cells of the layer above.
constexpr uint32_t kHalfShellMask = (1u << 11) | (1u << 14) | (1u << 16) | (1u << 17) | (0x1FFu << 18);

// CCNTCell::UpdateTotalEnergy's pair sum (src/CNTCell.cpp:2909-3229): every pair inside a
// cell once, a pair across cells only from the cell whose half shell holds the other cell
// and only when the neighbour's id is the larger (src/CNTCell.cpp:3081), so the
// reference's energy output leaves out about half of the cross-cell pairs.  Thread per bead
// over its 27 neighbour cells, 1/2 a (1 - r)^2 in FP64; out[0] += the sum.
__global__ __launch_bounds__(256) void k_energy_pairs(SliceArgs A, double* __restrict__ out)
{
    const BoxDev& B = A.B;
    const int s0 = A.cstart[B.koff], s1 = A.cstart[B.koff + B.ngroups * kGroupKeys];
    double e = 0.0;
    for (int i = s0 + blockIdx.x * blockDim.x + threadIdx.x; i < s1; i += gridDim.x * blockDim.x) {
        const float4 pi = A.pos[i];
        const int ti = bead_type(__float_as_int(pi.w));
        const int idi = __float_as_int(A.vp[i].w);
        int c[3];
        key_cell(A.key[i] - B.koff, B, c);
        for (int nb = 0; nb < 27; ++nb) {
            const int ox = nb % 3 - 1, oy = (nb / 3) % 3 - 1, oz = nb / 9 - 1;
            const bool same = nb == 13, half = (kHalfShellMask >> nb) & 1u;
            if (!same && !half) continue;
            int cx = c[0] + ox, cy = c[1] + oy, cz = c[2] + oz;
            cx += (cx < 0) ? B.n[0] : ((cx >= B.n[0]) ? -B.n[0] : 0);
            cy += (cy < 0) ? B.n[1] : ((cy >= B.n[1]) ? -B.n[1] : 0);
            cz += (cz < 0) ? B.n[2] : ((cz >= B.n[2]) ? -B.n[2] : 0);
            const int k = B.koff + cell_key(cx, cy, cz, B);
            const float sx = (float)ox * B.w[0], sy = (float)oy * B.w[1], sz = (float)oz * B.w[2];
            for (int j = A.cstart[k], j1 = A.cstart[k + 1]; j < j1; ++j) {
                if (j == i) continue;
                if (half && !(__float_as_int(A.vp[j].w) > idi)) continue;
                const float4 pj = A.pos[j];
                const float dx = (pi.x - pj.x) - sx, dy = (pi.y - pj.y) - sy, dz = (pi.z - pj.z) - sz;
                const double dr2 = (double)(dx * dx + dy * dy + dz * dz);
                if (!(dr2 < 1.0)) continue;
                const double dr = sqrt(dr2);
                if (!(dr > 0.000000001)) continue;
                const int tt = ti * kMaxTypes + bead_type(__float_as_int(pj.w));
                // a pair inside a cell is seen from both of its beads: half from each
                e += (same ? 0.25 : 0.5) * (double)A.table[tt] * (1.0 - dr) * (1.0 - dr);
            }
        }
    }
    e = wave_sum(e);
    if ((threadIdx.x & 63) == 0 && e != 0.0) atomicAdd(out, e);
}

// bond and bond-pair potential energies of the box (src/Bond.cpp:279-288,
// src/BondPair.cpp:320-378): thread per bond, then per bond pair; out[1], out[2]
__global__ __launch_bounds__(256) void k_energy_bonded(SliceArgs A, double* __restrict__ out)
{
    const int nb = A.b1 - A.b0, q = blockIdx.x * blockDim.x + threadIdx.x;
    double eb = 0.0, ea = 0.0;
    if (q < nb) {
        const int2 hb = A.bond_bead[A.b0 + q];
        double uh[3], ut[3];
        unwrapped(A, hb.x, A.slot_of[hb.x], uh);
        unwrapped(A, hb.y, A.slot_of[hb.y], ut);
        const double2 par = A.bond_par[A.b0 + q];
        const double L = sqrt((uh[0] - ut[0]) * (uh[0] - ut[0]) + (uh[1] - ut[1]) * (uh[1] - ut[1]) +
                              (uh[2] - ut[2]) * (uh[2] - ut[2]));
        eb = 0.5 * par.x * (L - par.y) * (L - par.y);
    } else if (q < nb + A.a1 - A.a0) {
        const int g = A.a0 + q - nb;
        const int4 abc = A.ang_bead[g];
        double ua[3], ub[3], uc[3], f[3][3];
        unwrapped(A, abc.x, A.slot_of[abc.x], ua);
        unwrapped(A, abc.y, A.slot_of[abc.y], ub);
        unwrapped(A, abc.z, A.slot_of[abc.z], uc);
        angle_forces(ua, ub, uc, A.ang_par[g], f, &ea);
    }
    eb = wave_sum(eb);
    ea = wave_sum(ea);
    if ((threadIdx.x & 63) == 0) {
        if (eb != 0.0) atomicAdd(out + 1, eb);
        if (ea != 0.0) atomicAdd(out + 2, ea);
    }
}

}  // namespace

hipError_t launch_energy_output(const SliceArgs& a, double* out3, hipStream_t s)
{
    hipLaunchKernelGGL(k_energy_pairs, dim3(kSliceBlocks), dim3(256), 0, s, a, out3);
    const int nq = (a.b1 - a.b0) + (a.a1 - a.a0);
    if (nq > 0) hipLaunchKernelGGL(k_energy_bonded, dim3((nq + 255) / 256), dim3(256), 0, s, a, out3);
    return hipGetLastError();
}

hipError_t launch_slice_bondpairs(const SliceArgs& a, double* tab, hipStream_t s)
{
    if (a.a1 > a.a0)
        hipLaunchKernelGGL(k_slice_bondpairs, dim3((a.a1 - a.a0 + 255) / 256), dim3(256), 0, s, a, tab);
    return hipGetLastError();
}

hipError_t launch_slice_stress(const SliceArgs& a, double* pair_tab, double* bond_tab, hipStream_t s)
{
    const size_t tab_bytes = sizeof(double) * (size_t)a.nslice * a.nidx * 9;
    if (pair_tab) {
        const int g = kSliceBlocks;
        if (tab_bytes <= kSliceLdsBytes)
            hipLaunchKernelGGL(k_slice_pairs<true>, dim3(g), dim3(256), tab_bytes, s, a, pair_tab);
        else
            hipLaunchKernelGGL(k_slice_pairs<false>, dim3(g), dim3(256), 0, s, a, pair_tab);
    }
    if (bond_tab && a.b1 > a.b0)
        hipLaunchKernelGGL(k_slice_bonds, dim3((a.b1 - a.b0 + 255) / 256), dim3(256), 0, s, a, bond_tab);
    return hipGetLastError();
}

}  // namespace dpd
