/*
 * hsolve_oracle.h — CPU restatement of MOOSE's HSolve (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, imports or calls
 * this code: only the tests, the smoke check of __graft_entry__.py and the
 * cpu_baseline leg of bench.py use it, and only as the checker / the CPU
 * baseline.  It restates, in plain C with the reference's operation order,
 * the setup and per-step arithmetic of
 *   hsolve/HinesMatrix.cpp, hsolve/HSolvePassive.cpp, hsolve/HSolveActive.cpp,
 *   hsolve/HSolveActiveSetup.cpp, hsolve/RateLookup.cpp, hsolve/HSolveStruct.cpp,
 *   hsolve/HSolveUtils.cpp:205-252, hsolve/HSolveInterface.cpp,
 *   biophysics/HHGate.cpp:258-281, biophysics/SpikeGen.cpp:202-227.
 * Build with -ffp-contract=off so that it matches the reference bit for bit.
 *
 * Parity pins (each checked in tests/):
 *   - step()/reinit()/matrix setup: bit-exact against the reference sources
 *     compiled by oracle/Makefile (oracle/_ref/hsolve_ref), and against the
 *     committed golden fixtures written from that build;
 *   - Hines matrix and elimination: the 9 hand-drawn trees of
 *     hsolve/HinesMatrix.cpp:503-689 checked against a dense matrix and dense
 *     Gaussian elimination (hsolve/testHSolve.cpp:37-108,
 *     hsolve/HSolvePassive.cpp:852-960);
 *   - walkTree numbering: hand-derived orders and structural invariants
 *     (reverse DFS pre-order from a terminal);
 *   - lookup tables / gate update / Ca recurrence / spike rule: closed forms.
 */
#ifndef HSOLVE_ORACLE_H
#define HSOLVE_ORACLE_H

#include "../include/hsolve_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_solver or_solver;

/* HSolve::setPath/setup on one cell.  Returns 0 or an HSG_E_* code. */
int  or_create(const hsg_cell_desc* d, double dt, int caAdvance, or_solver** out);
void or_destroy(or_solver* s);
/* HSolve::reinit with the clock tick dt. */
void or_reinit(or_solver* s, double dt);
/* One HSolve::process tick at clock time currTime (end of the step). */
void or_step(or_solver* s, double dt, double currTime);
/* Runs nsteps ticks; step k (1-based, counted from the last reinit) is at
 * time k*dt.  If trace != NULL, writes V of compartments probe[0..nprobe-1]
 * after each step: trace[k*nprobe + j]. */
void or_run(or_solver* s, double dt, long nsteps, const int* probe, int nprobe,
            double* trace);

int  or_n_compartments(const or_solver* s);
int  or_n_channels(const or_solver* s);
int  or_n_states(const or_solver* s);
int  or_n_pools(const or_solver* s);
/* Hines index of each compartment (desc order). */
void or_hines_order(const or_solver* s, int* out);
/* Zombie field access with the C-ABI field codes and desc indexing. */
int  or_get_field(const or_solver* s, int field, const int* idx, int n, double* out);
int  or_set_field(or_solver* s, int field, const int* idx, int n, const double* v);
int  or_add_inject(or_solver* s, int compartment, double value);
int  or_add_gkek(or_solver* s, int compartment, double Gk, double Ek);
int  or_add_spike(or_solver* s, int synchan, int synapse, double time);
int  or_add_conc(or_solver* s, int channel, double conc);
int  or_add_ica(or_solver* s, int pool, double iCa);
/* Spikes recorded since the last call: (spikegen index, step number). */
int  or_read_spikes(or_solver* s, int* gen, long* step, int cap);
/* Raw Hines-ordered arrays: V, VMid, state (gate order of the reference). */
void or_get_V(const or_solver* s, double* out);
void or_get_VMid(const or_solver* s, double* out);
void or_get_state(const or_solver* s, double* out);
void or_get_ca(const or_solver* s, double* out);
/* Writes the solver's flattened arrays for the reference harness. */
int  or_dump_flat(const or_solver* s, const char* path);

/* ---- HinesMatrix-level access (hsolve/HinesMatrix.cpp) -----------------
 * A bare Hines matrix built from a tree given directly in Hines indices, as
 * testHinesMatrix / testHSolvePassive do. children[cstart[i]..cstart[i+1]) are
 * the children of compartment i. */
typedef struct or_matrix or_matrix;
or_matrix* or_matrix_create(int n, const int* cstart, const int* children,
                            const double* Ra, const double* Rm, const double* Cm,
                            const double* Em, const double* V, double dt);
void   or_matrix_destroy(or_matrix* m);
double or_matrix_getA(const or_matrix* m, int row, int col);
double or_matrix_getB(const or_matrix* m, int row);
double or_matrix_getVMid(const or_matrix* m, int row);
double or_matrix_getV(const or_matrix* m, int row);
/* HSolvePassive::updateMatrix / forwardEliminate / backwardSubstitute */
void   or_matrix_update(or_matrix* m);
void   or_matrix_forward(or_matrix* m);
void   or_matrix_backward(or_matrix* m);

/* ---- LookupTable (hsolve/RateLookup.cpp) --------------------------------
 * row + lookup of one species of a table built like createLookupTables. */
double or_lookup_row(double min, double max, unsigned nDivs, double x,
                     unsigned* index);

#ifdef __cplusplus
}
#endif
#endif
