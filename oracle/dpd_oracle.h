/*
 * dpd_oracle.h -- FP64 CPU restatement of the Osprey-DPD step (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the MI355X engine.  Only tests, the smoke check and
 * the CPU-baseline leg of bench.py load it; the product library never links it.
 *
 * It restates, in plain C, the algorithm of the reference's per-step loop:
 *   CSimBox::Evolve          src/SimBox.cpp:977-1120   (step order)
 *   CSimBox::MakeCNTCells    src/SimBox.cpp:767-877    (cells, 13-cell half shell)
 *   CSimBox::AssignBeadstoCells src/SimBox.cpp:715-762 (initial binning, push_front)
 *   CCNTCell::UpdatePos      src/CNTCell.cpp:1598-2262 (modified velocity Verlet, migration)
 *   CCNTCell::UpdateForce    src/CNTCell.cpp:1144-1555 (DPD pair law, stress)
 *   CCNTCell::UpdateMom      src/CNTCell.cpp:2280-2310
 *   CCNTCell::Randf / lcg    src/CNTCell.cpp:2454-2465 (64-bit LCG)
 *   CBond::AddForce          src/Bond.cpp:111-162
 *   CBondPair::AddForce      src/BondPair.cpp:138-286
 *   CMonitor::CalculateSingleBeadData src/Monitor.cpp:790-1041 (kT, P, stress)
 *   CMonitor::AddBeadStress / AddBondStress src/Monitor.cpp:1900-2104, :2283-2518 (slice stress)
 *
 * Two random-force modes:
 *   ORC_RNG_LCG     the reference's serial LCG, drawn in the reference's linked-list
 *                   order (cell lists emulated exactly, push_front/erase included), so a
 *                   run reproduces the reference's stream on the same snapshot;
 *   ORC_RNG_PHILOX  the counter-based pair-keyed stream of the GPU engine: Philox4x32-10
 *                   (Salmon, Moraes, Dror, Shaw, SC'11) with counter (min id, max id,
 *                   step) and key = seed.  Same uniform law as the reference (U on [0,1),
 *                   force prefactor sqrt(24 kT / dt)).
 */
#ifndef DPD_ORACLE_H
#define DPD_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_RNG_LCG = 0, ORC_RNG_PHILOX = 1 };

typedef struct orc_params {
    int32_t nx, ny, nz;          /* CNT cells per axis            (dmpci Box line)   */
    double wx, wy, wz;           /* CNT cell widths (>= 1)                            */
    double dt, lambda, kT;       /* Step, Lambda, Temp                                */
    uint64_t seed;               /* |RNGSeed|  (CCNTCell::SetRNGSeed)                 */
    int32_t ntypes;
    const double* a;             /* ntypes*ntypes conservative a_ij, symmetric        */
    const double* g;             /* ntypes*ntypes dissipative gamma_ij, symmetric     */
    int32_t rng_mode;            /* ORC_RNG_LCG or ORC_RNG_PHILOX                     */
} orc_params;

/* topology: bonds by bead index (head, tail as in src/InputData.cpp:1499-1500),
 * bond pairs by bond index (first, second as in src/InputData.cpp:1155-1156), both
 * listed in the reference's per-polymer order. */
typedef struct orc_topology {
    int32_t nbonds;
    const int32_t* head;
    const int32_t* tail;
    const double* k;
    const double* l0;
    int32_t npairs;
    const int32_t* first;
    const int32_t* second;
    const double* modulus;
    const double* phi0;          /* radians */
} orc_topology;

typedef struct orc_obs {
    double kT;                   /* sum v^2 / 3N                                       */
    double pressure;             /* 0.3333333333/V (sum v^2 + trace stress)            */
    double stress[9];            /* sum over beads of m_Stress, / V                    */
    double sum_v2[3];
    double momentum[3];          /* sum v                                              */
    double pe_pair, pe_bond, pe_angle;
    int64_t wrong_cell;          /* beads outside their cell after the step            */
} orc_obs;

typedef struct orc_sys orc_sys;

/* x: wrapped positions, unpbc: unwrapped (NULL = same as x), v: velocities, f: forces,
 * all FP64 [3n] in bead order (the reference's GetInitialStateBeads order). */
orc_sys* orc_create(const orc_params* p, int32_t n, const int32_t* id, const int32_t* type,
                    const double* x, const double* unpbc, const double* v, const double* f,
                    const orc_topology* topo);
void     orc_destroy(orc_sys* s);
/* advance nsteps; the first step has simulation time step0+1 (it keys the Philox
 * stream and, for LCG mode, nothing). */
void     orc_run(orc_sys* s, int32_t nsteps, int64_t step0);
void     orc_get(const orc_sys* s, double* x, double* unpbc, double* v, double* f);
void     orc_observables(const orc_sys* s, orc_obs* out);
void     orc_set_gamma(orc_sys* s, const double* g);   /* thermostat toggles */
void     orc_set_a(orc_sys* s, const double* a);
/* CCNTCell::ChangeTimeStepConstants (SetTimeStepSize): dt and every constant using it */
void     orc_set_timestep(orc_sys* s, double dt);
/* CCNTCell::ToggleDPDBeadForces: the half-step constants from dt (0 = forces off) */
void     orc_set_halfstep(orc_sys* s, double dt);

/* One force evaluation at given positions and (predicted) velocities, Philox stream
 * keyed by `step`.  thermostat: 1 = dissipative + random, 0 = neither, 2 = dissipative
 * only.  Cell-list half shell; forces, virial (9, sum r_ij (x) F_ij incl.
 * bonds, no angles) and potential energies out.  Used for kernel parity. */
void orc_forces(const orc_params* p, int32_t n, const int32_t* id, const int32_t* type,
                const double* x, const double* unpbc, const double* v,
                const orc_topology* topo, int64_t step, int32_t thermostat,
                double* f, double* virial, double* pe);
/* The potential energy of the reference's ToggleEnergyOutput rows at positions x: out[0]
 * pair (pairs across cells only when the neighbour's id is larger, as the reference sums
 * them), out[1] bonds, out[2] bond pairs. */
void orc_energy_output(const orc_params* p, int32_t n, const int32_t* id, const int32_t* type,
                       const double* x, const double* unpbc, const orc_topology* topo, double* out);
/* Same as the pair part of orc_forces but by brute force over all pairs with the
 * minimum image; checks the cell list and the half shell. */
void orc_pair_forces_brute(const orc_params* p, int32_t n, const int32_t* id,
                           const int32_t* type, const double* x, const double* v,
                           int64_t step, int32_t thermostat, double* f, double* virial);

/* z-slice split of one interaction (CMonitor::AddBeadStress, src/Monitor.cpp:1900-2104):
 * adds wgt * share of F (x) dx into tab[(slice * nidx + idx) * 9 + c]; z1 is the wrapped
 * z of the bead F acts on, z2 of the other, dx = x1 - x2 (minimum image). */
void orc_slice_add(double* tab, int32_t nslice, double Lz, int32_t idx, int32_t nidx, double z1, double z2,
                   const double F[3], const double dx[3], double wgt);
/* The per-step slice stress tables of one force evaluation at (x, v, step) (Philox
 * stream): pair_tab[nslice][ntypes*ntypes][9] with the reference's half-shell order
 * (bead1 type first), bond_tab[nslice][ntypes*ntypes][9] keyed by (head type, tail
 * type); either may be NULL. */
void orc_slice_stress(const orc_params* p, int32_t n, const int32_t* id, const int32_t* type,
                      const double* x, const double* unpbc, const double* v, const orc_topology* topo,
                      int64_t step, int32_t thermostat, int32_t nslice, double* pair_tab, double* bond_tab);

/* The bond-pair slice stress table of the state (CMonitor::AddBondPairStress,
 * src/Monitor.cpp:2520-2739): tab[nslice][ntypes^3][9] keyed by the bead types (second
 * head, first head, first tail) of each bond pair; summed over slices it is the bond
 * pairs' virial sum_i F_i (x) u_i times 0.99999999. */
void orc_slice_stress_bondpairs(const orc_params* p, int32_t n, const int32_t* id, const int32_t* type,
                                const double* x, const double* unpbc, const orc_topology* topo, int32_t nslice,
                                double* tab);

/* reference LCG (src/CNTCell.cpp:2454-2465): advances *state, returns U in [0,1) */
double   orc_lcg_uniform(uint64_t* state);
/* Philox4x32-10 block: ctr[4], key[2] -> out[4] */
void     orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* the pair uniform used by the GPU engine */
double   orc_pair_uniform(uint64_t seed, int64_t step, int32_t id_i, int32_t id_j);

#ifdef __cplusplus
}
#endif
#endif
