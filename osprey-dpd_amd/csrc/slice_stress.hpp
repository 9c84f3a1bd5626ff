// slice_stress.hpp -- arguments of the z-slice stress kernels (slice_stress.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "engine_common.hpp"

namespace dpd {

struct SliceArgs {
    BoxDev B;                 // the box whose profile is taken
    const float4* pos;        // offset xyz, type | box << 8 (sorted)
    const float4* vp;         // predicted velocity xyz, id bits (sorted)
    const int4* img;          // periodic image counts by bead index
    const int* key;           // global cell key per slot
    const int* cstart;        // first slot of each global cell key
    const float* table;       // this box's a | gamma | sqrt(gamma)
    const int64_t* step;      // simulation time of the force evaluation (device)
    const int* slot_of;       // bead index -> slot
    const int2* bond_bead;    // (head, tail) bead index per bond
    const double2* bond_par;  // (k, l0) per bond
    int b0, b1;               // this box's bonds
    const int4* ang_bead;     // (second head, first head, first tail) bead index per bond pair
    const double4* ang_par;   // (k, phi0, cos phi0, sin phi0) per bond pair
    int a0, a1;               // this box's bond pairs
    int nslice;
    int ntypes;
    int nidx;                 // ntypes * ntypes
};

// largest table kept in LDS per workgroup; larger tables go straight to global atomics
constexpr size_t kSliceLdsBytes = 32768;
// workgroups of the pair pass (grid-stride over the box's beads)
constexpr int kSliceBlocks = 1024;

hipError_t launch_slice_stress(const SliceArgs& a, double* pair_tab, double* bond_tab, hipStream_t s);
// the reference's energy-output sums: out3[0] pairs, [1] bonds, [2] bond pairs (added to)
hipError_t launch_energy_output(const SliceArgs& a, double* out3, hipStream_t s);
// bond pairs: tab[nslice][ntypes^3][9] (cleared by the caller)
hipError_t launch_slice_bondpairs(const SliceArgs& a, double* tab, hipStream_t s);

}  // namespace dpd
