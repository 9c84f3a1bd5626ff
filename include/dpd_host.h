/*
 * dpd_host.h -- C ABI of the host layer of libospreydpd_amd.so: Osprey-DPD control
 * files (dmpci.<runId>), polymer topology, starting states, restart snapshots and the
 * step-loop driver that calls the engine of dpd_engine.h.
 *
 * These replace, on the host side of the same seam:
 *   CInputDPDFile::Serialize / CInputData::CreateNewTypes  (src/InputDPDFile.cpp:190-1052,
 *     src/InputData.cpp:417-674, :1028-1188, :1469-1513, :1711-1790)
 *   CInitialState::CreatePolymers                          (src/InitialState.cpp:451-521)
 *   CRestartState::Serialize (bead section)                (src/RestartState.cpp:100-330)
 *   CSimBox::Run (commands -> Evolve -> sample -> save)     (src/SimBox.cpp:1272-1355)
 * All functions return DPD_OK or a DPD_ERR_* code of dpd_engine.h.
 */
#ifndef DPD_HOST_H
#define DPD_HOST_H
#include <stdint.h>

#include "dpd_engine.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct dpd_input dpd_input;

typedef struct dpd_counts {
    int32_t ntypes;
    int64_t nbeads, nbonds, npairs, npolymers;
    int32_t npolymer_types;
    int32_t nx, ny, nz;
    double wx, wy, wz;
    double density, kT, lambda, dt;
    int64_t seed;               /* RNGSeed as written (the engine uses |seed|) */
    int64_t total_time, sample_period, restart_period;
    int32_t ncommands;
    int32_t state;              /* 0 random, 1 restart, 2 lamella, 3 other */
    int64_t restart_state;
} dpd_counts;

/* Parse a dmpci file and expand its polymers.  On failure *out is NULL and the message
 * is written to err (errlen bytes). */
int  dpd_input_read(const char* path, dpd_input** out, char* err, int32_t errlen);
int  dpd_input_parse(const char* text, dpd_input** out, char* err, int32_t errlen);
void dpd_input_free(dpd_input* in);
int  dpd_input_counts(const dpd_input* in, dpd_counts* out);
/* bead names (comma-separated) and the ntypes^2 matrices */
int  dpd_input_types(const dpd_input* in, char* names, int32_t len, double* a, double* gamma);
/* per bead: id (1..N), type, polymer id (1..), polymer type */
int  dpd_input_beads(const dpd_input* in, int32_t* id, int32_t* type, int32_t* poly_id, int32_t* poly_type);
/* bonds (head, tail by bead index) and bond pairs (first, second by bond index, phi0 in
 * radians), in the reference's per-polymer order */
int  dpd_input_topology(const dpd_input* in, int32_t* head, int32_t* tail, double* k, double* l0,
                        int32_t* first, int32_t* second, double* modulus, double* phi0);
/* Bond types in declaration order (CInputData::AddBondType, src/InputData.cpp:556-598):
 * *count of them, with the bead types named at head and tail (arrays of *count entries,
 * may be NULL; call with both NULL to size them).  A bond's type is the declaration whose
 * two names match its beads in either order (src/InputData.cpp:1489-1498); the bond
 * slice-stress tables of the reference are keyed by it. */
int  dpd_input_bond_types(const dpd_input* in, int32_t* count, int32_t* head_type, int32_t* tail_type);
/* Write the initial-state summary dmpcis.<runId> the reference writes before its first
 * step (CInitialState::SerializeAsciiFile, src/InitialState.cpp:1153-1347): numbers of
 * polymers, beads, bonds and bond pairs created, polymers per type, bead types (radius,
 * conservative and dissipative rows), bond types, bond-pair types (modulus, phi0 in
 * radians) and each polymer type's fraction and bead, bond and bond-pair types. */
int  dpd_input_write_summary(const dpd_input* in, const char* path);
/* Engine parameters of this input (pointers stay valid while `in` lives). */
int  dpd_input_params(const dpd_input* in, int32_t device, dpd_params* out);
/* Starting state for State random / lamella: positions (wrapped and unwrapped) and
 * Maxwell velocities at Temp with zero total momentum, rescaled to kT exactly. */
int  dpd_input_build_state(const dpd_input* in, uint64_t seed, double* x, double* unpbc, double* v);
/* command i: name, execution time and arguments (space-separated) */
int  dpd_input_command(const dpd_input* in, int32_t i, char* name, int32_t nlen, int64_t* time,
                       char* args, int32_t alen);

/* Restart snapshot bead section (dmpcrs.<runId>.con.<t>.dat): 17 columns per bead,
 * 6 significant digits, coordinates within 0.001 of a face written as 0.001
 * (src/RestartState.cpp:113-156, :492-540). */
int  dpd_restart_read(const char* path, int64_t n, int32_t* poly_id, int32_t* poly_type, int32_t* id,
                      int32_t* type, double* x, double* unpbc, double* v, double* f);
int  dpd_restart_write(const char* path, int64_t n, const int32_t* poly_id, const int32_t* poly_type,
                       const int32_t* id, const int32_t* type, const double* radius_by_type,
                       const double* L3, const double* x, const double* unpbc, const double* v,
                       const double* f);

typedef struct dpd_run_stats {
    int64_t steps;
    int64_t nbeads;
    double seconds;             /* wall time of the step loop                         */
    double bead_steps_per_s;
    double last_kT, last_pressure;
    int64_t samples;
    int64_t commands_applied;
} dpd_run_stats;

/* Run dmpci.<run_id> found in `dir` to completion on `device`, as `dpd <runId>` does:
 * the dmpcis.<run_id> summary first, commands at their times, a sample every SamplePeriod
 * appended to dmpchs.<run_id> (the reference's history rows: time, kT, P, the energies
 * under ToggleEnergyOutput, MSD per bead type, end-to-end distance per polymer type), a
 * restart snapshot every RestartPeriod.  A pair-launch schedule recorded by the autotuner
 * for the box's shape is applied from the file DPD_TUNE_FILE names, else from
 * dpd_tune.json in `dir` when it exists. */
int  dpd_run(const char* dir, const char* run_id, int32_t device, dpd_run_stats* stats, char* err,
             int32_t errlen);

#ifdef __cplusplus
}
#endif
#endif
