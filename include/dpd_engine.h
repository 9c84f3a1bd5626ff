/*
 * dpd_engine.h -- C ABI of the MI355X DPD step engine (libospreydpd_amd.so).
 *
 * The engine replaces the body of the reference's per-step loop:
 *   CSimBox::Evolve (src/SimBox.cpp:977-1120) = UpdatePos over all cells
 *   (src/CNTCell.cpp:1598), UpdateForce (src/CNTCell.cpp:877), the bond and
 *   bond-pair loops (src/SimBox.cpp:1382-1439 -> src/Bond.cpp:111,
 *   src/BondPair.cpp:138) and UpdateMom (src/CNTCell.cpp:2280),
 * with hand-written HIP kernels for gfx950.  The reference has no plugin or FFI
 * interface: the seam is the call of Evolve() from CSimBox::Run
 * (src/SimBox.cpp:1320); constants are pushed into the CCNTCell statics by
 * CSimState (src/SimState.cpp:69-81, :112), which dpd_create mirrors.
 *
 * Conventions: plain C types only; every call returns DPD_OK (0) or an error code
 * and leaves a message for dpd_last_error(); no exceptions cross the ABI.  One
 * engine per host thread; each engine owns its device buffers and one HIP stream.
 * Host arrays are copied on upload and never aliased.  Bead arrays are in "upload
 * order" (the order of the arrays given to dpd_upload_beads); topology refers to
 * beads by that 0-based index.  The bead `id` (the reference's CAbstractBead id)
 * keys the pair random stream only.
 */
#ifndef DPD_ENGINE_H
#define DPD_ENGINE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference reports these conditions through bool returns and
 * CLog* messages: src/LogMessage.cpp:33-37, src/CNTCell.cpp:2628-2693) */
#define DPD_OK              0
#define DPD_ERR_PARAM       1   /* invalid argument or parameter                      */
#define DPD_ERR_OOM         2   /* device allocation failed                           */
#define DPD_ERR_HIP         3   /* HIP runtime error                                  */
#define DPD_ERR_NAN         4   /* a force or velocity became NaN/Inf                 */
#define DPD_ERR_OUT_OF_BOX  5   /* uploaded bead outside [0, L)                       */
#define DPD_ERR_STATE       6   /* call out of order (e.g. step before upload)        */
#define DPD_ERR_IO          7   /* file or parse error (host layer)                   */
#define DPD_ERR_UNSUPPORTED 8   /* configuration the engine does not handle           */

typedef struct dpd_engine dpd_engine;

/* Box and integration constants.  Mirrors CCNTCell::SetSimBoxLengths
 * (src/CNTCell.cpp:143-158), SetTimeStepConstants (:180-203), SetRNGSeed
 * (:134-137) and SetDPDBeadIntArrays (src/SimState.cpp:112). */
typedef struct dpd_params {
    int32_t nx, ny, nz;         /* CNT cells per axis (>= 3)                         */
    double wx, wy, wz;          /* cell widths (>= 1: the cutoff is 1)               */
    double dt;                  /* Step                                              */
    double lambda;              /* Lambda (velocity prediction, 0..1)                */
    double kT;                  /* Temp                                              */
    uint64_t seed;              /* |RNGSeed|: key of the pair random stream          */
    int32_t ntypes;             /* bead types (<= DPD_MAX_TYPES)                     */
    const double* a;            /* ntypes*ntypes conservative a_ij, row-major        */
    const double* gamma;        /* ntypes*ntypes dissipative gamma_ij                */
    int32_t device;             /* HIP device ordinal                                */
    int32_t pair_kernel;        /* DPD_PAIR_AUTO or one of the kernels below          */
} dpd_params;

#define DPD_MAX_TYPES 16

#define DPD_PAIR_AUTO      0    /* re-chosen at each commit: the hit-list kernel up   */
                                /* to 2^26 beads, the full-shell kernel beyond       */
#define DPD_PAIR_FULLSHELL 2    /* thread-per-bead 27-cell kernel                    */
#define DPD_PAIR_HITLIST   8    /* thread-per-bead scan of the 9 neighbour rows      */
                                /* (cells x-1..x+1 as one range) into an LDS list of */
                                /* partners in range, then one dense evaluation loop */
/* Ids 1 and 3-7 named LDS brick, cell-stream, compacted and wavefront-per-cell kernels that
 * measured slower than the hit-list kernel at every size; they are retired and refused
 * with DPD_ERR_UNSUPPORTED. */
#define DPD_PAIR_BRICK     1
#define DPD_PAIR_BRICK512  3
#define DPD_PAIR_STREAM    4
#define DPD_PAIR_STREAM1024 5
#define DPD_PAIR_COMPACT   6
#define DPD_PAIR_CELLS     7
#define DPD_PAIR_KINDS     8    /* kernel ids are 1..DPD_PAIR_KINDS                  */

/* FP64 observables of the last completed step (CMonitor::CalculateSingleBeadData,
 * src/Monitor.cpp:790-1041).  stress = (sum over pairs and bonds of r (x) F) / V,
 * component [3a+b] = sum dx_b F_a; bond-angle forces add no stress, as in the
 * reference. */
typedef struct dpd_obs {
    int64_t step;               /* simulation time of the sample                     */
    double kT;                  /* sum v^2 / (3N)                                    */
    double pressure;            /* 0.3333333333/V * (sum v^2 + trace sum r(x)F)     */
    double stress[9];
    double sum_v2[3];
    double momentum[3];         /* sum v (unit mass)                                 */
    double pe_pair;             /* sum 1/2 a (1-r)^2                                 */
    double pe_bond;             /* sum 1/2 k (L-l0)^2                                */
    double pe_angle;            /* CBondPair::AddPotentialEnergy                     */
    int64_t jumps;              /* bead moves of more than one cell in the last call */
    int64_t nbeads;
} dpd_obs;

/* flags of dpd_step / dpd_compute_forces */
#define DPD_STEP_OBSERVE   1u   /* accumulate FP64 observables at the last step      */
#define DPD_NO_THERMOSTAT  2u   /* dpd_compute_forces: drop dissipative + random     */
#define DPD_STEP_TIMING    4u   /* time every pair-kernel launch with HIP events     */
#define DPD_STEP_GRAPH     8u   /* replay captured hipGraphs of up to 64 steps       */
#define DPD_STEP_TIMING_SAMPLED 16u /* with DPD_STEP_TIMING: time every 16th pair launch */
                                /* only (a quarter of the event-marker overhead)     */

/* toggles of dpd_toggle (command handlers of src/CNTCell.cpp:223-346) */
#define DPD_TOGGLE_THERMOSTAT     1   /* ToggleDPDBeadThermostat: on=0 zeroes gamma  */
#define DPD_TOGGLE_CONSERVATIVE   2   /* ToggleDPDBeadConservativeForces              */
#define DPD_TOGGLE_FORCES         3   /* ToggleDPDBeadForces (src/CNTCell.cpp:223-227, */
                                      /* src/SimState.cpp:599-618): off zeroes the     */
                                      /* half-step constants, so no force moves a bead  */
                                      /* or changes a velocity (ballistic motion)       */

/* Create an engine on p->device.  Replaces the CCNTCell static set-up. */
int  dpd_create(const dpd_params* p, dpd_engine** out);
void dpd_destroy(dpd_engine* e);
const char* dpd_last_error(const dpd_engine* e);

/* Beads (FP64, [3n] xyz-interleaved).  x must lie in [0, L); unpbc may be NULL
 * (then unwrapped = wrapped); f may be NULL (zero), as for a fresh state.  Replaces
 * CSimBox::AssignBeadstoCells (src/SimBox.cpp:715-762) and the bead fields the
 * restart reader fills (src/RestartState.cpp:312-330).  step0 is the simulation
 * time of the uploaded state. */
int dpd_upload_beads(dpd_engine* e, int64_t n, const int32_t* id, const int32_t* type,
                     const double* x, const double* unpbc, const double* v, const double* f,
                     int64_t step0);

/* Bonds: head/tail bead indices (CBond m_pHead/m_pTail, src/InputData.cpp:1499-1500),
 * spring constant and unstretched length (src/Bond.cpp:111-162).
 * Bond pairs: indices of the first and second bond (CBondPair m_pFirst/m_pSecond,
 * stored in reverse chain order, src/InputData.cpp:1155-1156), modulus and phi0 in
 * radians (src/BondPairType.cpp:85).  Either count may be 0. */
int dpd_upload_topology(dpd_engine* e, int64_t nbonds, const int32_t* head, const int32_t* tail,
                        const double* k, const double* l0, int64_t npairs,
                        const int32_t* first_bond, const int32_t* second_bond,
                        const double* modulus, const double* phi0);

/* Advance nsteps steps (one CSimBox::Evolve each).  Inputs and outputs stay in
 * device memory; the call returns once the steps are enqueued and checked. */
int dpd_step(dpd_engine* e, int64_t nsteps, uint32_t flags);

/* One force evaluation at the uploaded positions, treating the uploaded velocities
 * as the predicted velocities, with the pair stream of simulation time `step`.
 * f_out [3n] in upload order; virial[9] and pe[3] may be NULL.  A diagnostic: the state
 * the next dpd_step starts from is left unchanged.  After dpd_step it evaluates the last
 * step's positions and predicted velocities (the inputs of that step's force). */
int dpd_compute_forces(dpd_engine* e, int64_t step, uint32_t flags, double* f_out,
                       double* virial, double* pe);

int dpd_observables(dpd_engine* e, dpd_obs* out);

/* Copy the state back in upload order (any pointer may be NULL). */
int dpd_download(dpd_engine* e, double* x, double* unpbc, double* v, double* f);

/* Command hooks (src/CNTCell.cpp:357-377, :434-456, :223-346) */
int dpd_set_interactions(dpd_engine* e, const double* a, const double* gamma);
int dpd_set_timestep(dpd_engine* e, double dt);
int dpd_toggle(dpd_engine* e, int32_t what, int32_t on);

/* ---- batches of independent boxes (replicas of uneven size and parameters) ----
 * No counterpart in the reference, which runs one box per process (src/dmpc.cpp:203-213
 * runs several runIds one after another).  All boxes advance together: every kernel of
 * a step spans all boxes, and per-box constants and interaction tables live in device
 * memory.  Bead and topology indices of dpd_upload_box* are local to the box. */
int dpd_create_batch(int32_t nbox, const dpd_params* p, dpd_engine** out);   /* p[nbox] */
int dpd_upload_box(dpd_engine* e, int32_t box, int64_t n, const int32_t* id, const int32_t* type,
                   const double* x, const double* unpbc, const double* v, const double* f);
int dpd_upload_box_topology(dpd_engine* e, int32_t box, int64_t nbonds, const int32_t* head,
                            const int32_t* tail, const double* k, const double* l0, int64_t npairs,
                            const int32_t* first_bond, const int32_t* second_bond,
                            const double* modulus, const double* phi0);
/* move the staged boxes to the device and sort them; step0 = simulation time */
int dpd_commit(dpd_engine* e, int64_t step0);
int32_t dpd_box_count(const dpd_engine* e);
int dpd_observables_box(dpd_engine* e, int32_t box, dpd_obs* out);
/* box = -1: all beads of all boxes in upload order */
int dpd_download_box(dpd_engine* e, int32_t box, double* x, double* unpbc, double* v, double* f);
/* box = -1: every box */
int dpd_set_box_interactions(dpd_engine* e, int32_t box, const double* a, const double* gamma);

/* Timing of the last dpd_step call (ms, from HIP events on the engine's stream):
 * [0] whole call, [1] the timed pair-kernel launches (with DPD_STEP_TIMING; all of them,
 * or every 16th with DPD_STEP_TIMING_SAMPLED), [2] [0] - [1].  Returns the number of
 * pair-kernel launches timed. */
int64_t dpd_last_timing(dpd_engine* e, double* ms3);
int64_t dpd_get_step(const dpd_engine* e);

/* Time the pair-force launch of the current state without event markers between
 * launches: one untimed launch, then `nlaunch` back-to-back launches of the launch the
 * step uses (the pair kernel, with the bonded terms when they share its launch) between
 * two HIP events on the engine's stream; *ms_per_launch = elapsed / nlaunch.  The state
 * is left as it was (the force the next step starts from is restored).  No counterpart
 * in the reference, which has no timers. */
int dpd_time_pair(dpd_engine* e, int32_t nlaunch, double* ms_per_launch);

/* Pair-kernel choice.  dpd_autotune times `nsteps` steps of every pair kernel on the
 * current state (restoring it after each trial, so the simulation does not advance),
 * keeps the fastest and returns it; ms_per_step[DPD_PAIR_KINDS] (may be NULL) receives
 * the time per step of kernel k at index k - 1 (0 for retired ids). */
int dpd_set_pair_kernel(dpd_engine* e, int32_t kernel);
int32_t dpd_get_pair_kernel(const dpd_engine* e);
int dpd_autotune(dpd_engine* e, int32_t nsteps, int32_t* chosen, double* ms_per_step);

/* Schedules of the pair launch: the hit-list kernel with its workgroup size, per-thread
 * list capacity and flush threshold, and the full-shell kernel; dpd_schedule_name gives
 * each one's name ("hitlist-t512-c24", ...), NULL past the last.  dpd_autotune_schedules
 * times `nsteps` steps of every schedule on the current state as dpd_autotune does (the
 * state is restored after each trial), keeps the fastest (the choice then stays across
 * commits) and returns it; ms_per_step[dpd_schedule_count()] (may be NULL) receives each
 * schedule's time per step.  No counterpart in the reference, which has one loop. */
int32_t dpd_schedule_count(void);
const char* dpd_schedule_name(int32_t sched);
int dpd_set_schedule(dpd_engine* e, int32_t sched);
int32_t dpd_get_schedule(const dpd_engine* e);
int dpd_autotune_schedules(dpd_engine* e, int32_t nsteps, int32_t* chosen, double* ms_per_step);
/* Apply a schedule recorded by the autotuner (a file of one JSON object per line, written
 * by dpd_amd.tune, keyed by the number of boxes (1, or 2 for any batch), beads per box as a
 * power-of-two range [beads_lo, beads_hi), bead types and whether there are bonds) to a
 * committed engine of that shape; *chosen (may be NULL) gets the schedule set, -1 when the
 * file holds none for this shape (the engine is then left as it was). */
int dpd_apply_tuning(dpd_engine* e, const char* path, int32_t* chosen);

/* Momentum commands on the current (closed) state of box `box` (-1: every box, each on
 * its own), executed between dpd_step calls as the reference executes commands before
 * Evolve.  dpd_zero_total_momentum subtracts the mean velocity (CSimBox::ZeroTotalMom,
 * src/SimBox.cpp:1919-1947); dpd_renormalise_momenta scales all velocities so that
 * sum v^2 / (3 N) = kT (CSimBox::UpdateRenormalisedMom, src/SimBox.cpp:3177-3215, which
 * the RenormaliseMomenta command schedules every Period steps, src/SimBox.cpp:1313-1317).
 * Sums in FP64. */
int dpd_zero_total_momentum(dpd_engine* e, int32_t box);
int dpd_renormalise_momenta(dpd_engine* e, int32_t box);

/* Stress profile across `nslices` equal z-slices of box `box` (replaces the per-step
 * tables of CMonitor::AddBeadStress, src/Monitor.cpp:1900-2104, and
 * CMonitor::AddBondStress, src/Monitor.cpp:2283-2518).  Each interaction adds F (x) dx
 * (tensor element 3a + b = F_a dx_b; F on the first bead, dx = first - second) to the
 * slices its two beads span: the whole of it to their slice when both lie in one, else
 * the share width / |dz| to each slice crossed and the fractions of that to the two end
 * slices.  pair_out[nslices][nt][nt][9] (may be NULL) holds the pair forces of the
 * current state -- positions, predicted velocities and pair stream of the last force
 * evaluation, i.e. the last step's forces after dpd_step and the uploaded state at
 * step0 after dpd_commit -- with each pair's share split equally between (ti, tj) and
 * (tj, ti); the reference files a pair under the type of whichever bead its cell
 * traversal meets first, so only the sum of the two entries is defined for ti != tj.
 * bond_out[nslices][nt][nt][9] (may be NULL) holds the bonds under (head type, tail
 * type) with dx = head - tail and F the force on the head.  Sums over slices are the
 * pair and bond virials.  The tables are instantaneous, as the reference's are. */
int dpd_slice_stress(dpd_engine* e, int32_t box, int32_t nslices, double* pair_out,
                     double* bond_out);

/* dpd_slice_stress without device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The device tables live in a static scratch of
 * 2^19 doubles (4 MiB) that each GPU holds until the process ends: sizes with
 * 2 * nslices * ntypes^2 * 9 above 2^19 are refused with DPD_ERR_PARAM.  The scratch is
 * cleared by a kernel on every call (no memset node); all calls on one GPU share it, so
 * calls that overlap on one GPU race.  Same arguments and results as dpd_slice_stress. */
int dpd_slice_stress_noalloc(dpd_engine* e, int32_t box, int32_t nslices, double* pair_out,
                             double* bond_out);

/* The potential energy the reference writes with ToggleEnergyOutput
 * (CCNTCell::UpdateTotalEnergy, src/CNTCell.cpp:2909-3229; src/Bond.cpp:279-288;
 * src/BondPair.cpp:320-378) at the current positions of box `box`, FP64: out3[0] the pair
 * sum in the reference's order -- every pair inside a cell, a pair across cells only from
 * the cell whose 13-cell half shell holds the other and only when the neighbour's id is
 * larger, so about half of the cross-cell pairs are left out, as in the reference --,
 * out3[1] the bonds 1/2 k (L - l0)^2, out3[2] the bond pairs.  (dpd_obs.pe_pair is the
 * true pair energy.) */
int dpd_energy_output(dpd_engine* e, int32_t box, double* out3);

/* Bond-pair stress profile across `nslices` equal z-slices of box `box` (replaces the
 * per-step tables of CMonitor::AddBondPairStress, src/Monitor.cpp:2520-2739, which
 * CPolymer::AddBondPairForces fills for every bond pair, src/Polymer.cpp:703).  Each bond
 * pair's three bead pairs, first tail -> first head -> second head -> first tail, add
 * 0.33333333 (F[j+1] - F[j]) (x) (u[j+1] - u[j]) (F the bond pair's force on each bead, u
 * unwrapped positions) with the slice split of dpd_slice_stress.
 * out[nslices][nt][nt][nt][9], keyed by the bead types (second head, first head, first
 * tail); the reference keys by bond-pair type, which the declaration of the three bead
 * types fixes.  Summed over slices and types: the bond pairs' virial sum F (x) u times
 * 0.99999999.  The positions are those of the last completed step. */
int dpd_slice_stress_bondpairs(dpd_engine* e, int32_t box, int32_t nslices, double* out);

#ifdef __cplusplus
}
#endif
#endif
