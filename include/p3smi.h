/*
 * p3smi.h -- C ABI of the MI355X batched polar Newton power-flow engine (libp3smi.so).
 *
 * Plain C types only: pointers, sizes, status codes. Every function returns 0 on success
 * and a nonzero P3SMI_E_* code on failure; p3smi_last_error() then gives the message of
 * the calling thread's last failure. Complex arrays are interleaved (re, im) doubles, the
 * memory layout of numpy complex128. A (rows, cases) array is C-ordered: element
 * (i, c) lives at [i * cases + c], the layout of a numpy (n, B) array.
 *
 * The entry points replace the reference's pybind11 module nr_klu
 * (p3s/cpp/nr_klu.cpp:1315-1376) and the GPU solver's batch calls
 * (p3s/cuda/nr_polar_solver.py:466-557); each declaration cites what it replaces.
 */
#ifndef P3SMI_H
#define P3SMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define P3SMI_OK 0
#define P3SMI_E_ARG 1      /* invalid argument (shape, range, NaN start voltage, ...) */
#define P3SMI_E_DEVICE 2   /* HIP runtime failure or no device */
#define P3SMI_E_MEMORY 3   /* device or host allocation failed */
#define P3SMI_E_INTERNAL 4

/* precision of the Newton-step linear solve */
#define P3SMI_PREC_DOUBLE 0 /* FP64 static-pivot LU: the iteration nr_klu runs */
#define P3SMI_PREC_MIXED 1  /* FP32 LU + FP64 iterative refinement, FP64 LU fallback per case */
/* The reference GPU solver's mixed mode (p3s/cuda/nr_polar_kernels.cu:209-288): FP32 LU (no
 * refinement) while the case's mismatch is above the switching value (1e-3 p.u. unless
 * set), FP64 below it, FP64 for good once an FP32 step fails to halve the mismatch or the
 * line search cuts a step, or FP32 pivots fail (that iteration retried in FP64). An inexact
 * Newton: iteration counts may differ from the FP64 iteration. */
#define P3SMI_PREC_MIXED_SWITCH 2

typedef struct p3smi_topology p3smi_topology; /* one grid topology, resident on the device */
typedef struct p3smi_batch p3smi_batch;       /* device workspace for a batch of cases */

typedef struct {
    int32_t n, m, npv, npq, nnz_y, nnz_j;
    int64_t nnz_lu, n_pairs;
    int32_t n_levels, n_blevels, n_slots, ordering;
    double rcond0;
    int32_t lu_g, lu_ns, lu_few_g, lu_few_ns, lu_few_max, lu_group;
    int32_t eval_rows; /* Ybus rows per tile of the mismatch/Jacobian evaluation, at most */
} p3smi_info;

/* Per-case/aggregate counters of the last solve (the reference's last_stats,
 * p3s/cuda/nr_polar_solver.py:671-681). Kernel times are filled only when timing is on. */
typedef struct {
    int32_t cases, newton_passes, line_search_passes;
    int32_t failed_factorizations;
    int32_t guard_trips; /* factorizations whose pivot ratio fell below 1e-3 x the flat-start one */
    int32_t reruns;      /* cases run again from their start with pivoting after not converging */
    int64_t single_factorizations, double_factorizations, refinement_steps;
    double ms_eval, ms_lu, ms_step, ms_decide, ms_other, ms_refine;
    int32_t n_eval, n_lu, n_step, n_decide, n_refine;
    int32_t pivot_factorizations; /* fresh pivoting factorizations (the frozen plan's fallback) */
    /* FP64 factor + solve launches of the batch LU shape (part of ms_lu / n_lu) and the cases
     * they factored: the dominant kernel of a whole-batch FP64 solve, timed apart */
    double ms_lu_batch;
    int32_t n_lu_batch;
    int64_t lu_batch_cases;
    /* of pivot_factorizations: those whose threshold pivoting chose the plan's diagonal at every
     * column, taken from the static-pivot LU after a check (the case then stays on the plan) */
    int32_t diagonal_factorizations;
} p3smi_stats;

/* ---- library ------------------------------------------------------------------------ */
const char* p3smi_last_error(void);
int p3smi_version(void);
/* Number of visible HIP devices (0 without a GPU); never initialises a context. */
int p3smi_device_count(int32_t* count);
/* Free and total device memory in bytes (no counterpart in the reference; the serving
 * layer sizes its admission budget from it). Fails without a device. */
int p3smi_device_memory(int64_t* free_bytes, int64_t* total_bytes);
/* HBM reference patterns (no counterpart in the reference; the measurement of the kernels'
 * HBM traffic scales the hardware counters by them). Allocates two device buffers of `bytes`
 * (a power of two, 1 MiB .. 64 GiB), fills the source, then times `repeats` launches (after one
 * warm-up) of pattern 0, a linear copy, or 1, a copy of 512-byte rows between permuted row
 * positions (every source byte read once, every destination byte written once), and returns
 * the average milliseconds per launch. */
int p3smi_hbm_probe(int32_t pattern, int64_t bytes, int32_t repeats, double* ms_per_launch);

/* ---- topology (replaces nr_klu.Solver.__init__, p3s/cpp/nr_klu.cpp:966-994, and
 *      PolarNewtonSolverCUDA.__init__, p3s/cuda/nr_polar_solver.py:282-376) ---------------
 * Yp (n+1), Yj (nnz): CSR pattern of Ybus, columns sorted per row; Yx (nnz) complex.
 * pv (npv), pq (npq): bus index sets. ordering: 0 fill-reducing (minimum fill on the buses,
 * each bus's P equation before its Q equation), 2 natural.
 * n_slots: task lists per workgroup of the LU kernel (0 = default).
 * The host plan is built here; device buffers are uploaded on first use. */
int p3smi_topology_create(const int32_t* Yp, const int32_t* Yj, const double* Yx, int32_t n, int32_t nnz,
                          const int32_t* pv, int32_t npv, const int32_t* pq, int32_t npq, int32_t ordering,
                          int32_t n_slots, p3smi_topology** out);
void p3smi_topology_free(p3smi_topology* t);
int p3smi_topology_info(const p3smi_topology* t, p3smi_info* out);
/* Switching mismatch of P3SMI_PREC_MIXED_SWITCH (p.u., default 1e-3 as the reference's
 * mixed_switch, p3s/cuda/nr_polar_solver.py:323). */
int p3smi_topology_set_mixed_switch(p3smi_topology* t, double mismatch);
/* Largest number of cases the one-call entry points solve at once: the reference GPU
 * solver's max_chunk (p3s/cuda/nr_polar_solver.py:311, 619-637). 0, the default: as many as
 * 60% of the free HBM holds; else a positive multiple of 64. Larger calls are split into
 * chunks of nearly equal size, whole multiples of 64, solved one after the other. */
int p3smi_topology_set_max_chunk(p3smi_topology* t, int32_t cases);
/* LU launch shape (no counterpart in the reference; set by the autotuner): a workgroup of
 * 64*ns threads factors g consecutive cases (g in {1, 16, 32, 64}, ns in {8, 16}, and ns = 4
 * with g = 1);
 * (few_g, few_ns) is used instead while at most few_max cases are still iterating. group 0
 * with g = 64 treats 2x2 blocks of paired pivots (the P and Q unknowns of a PQ bus) as one
 * task; group > 0 deals runs of up to `group` entries of one pivot's U row / L column to one
 * task list (1: one entry per task). Every shape gives the same bits. */
int p3smi_topology_set_lu_shape(p3smi_topology* t, int32_t g, int32_t ns, int32_t few_g, int32_t few_ns,
                                int32_t few_max, int32_t group);
/* Row tiles of the mismatch/Jacobian evaluation (no counterpart in the reference; set by the
 * autotuner): a workgroup evaluates consecutive Ybus rows, at most `rows` (1..1024) of them
 * and at most 768 nonzeros, staged in LDS. Default 32. Results do not depend on it. */
int p3smi_topology_set_eval_tile(p3smi_topology* t, int32_t rows);
/* Mixed precision (no counterpart in the reference): refinement of the FP32-LU step stops
 * once the correction is below rtol x the step, at most kmax steps, then falls back to an
 * FP64 factorization of that case. Defaults 1e-12 and 6. */
int p3smi_topology_set_mixed(p3smi_topology* t, double rtol, int32_t kmax);
/* Pivot-ratio guard of the frozen static-pivot plan (p3s/cpp/nr_klu.cpp:564-566, 609-620): a
 * factorization whose min/max |pivot| falls below ratio x the flat-start one (or meets a zero
 * pivot) is rejected, and the case factors with threshold partial pivoting (tolerance 1e-3,
 * diagonal preferred, rows scaled) for the rest of its run, refactoring on that pivot order and
 * pivoting again when a pivot vanishes. A case that ends without converging on the frozen plan
 * is run again from its start with pivoting (p3s/cpp/nr_klu.cpp:844-857). Default 1e-3; 1e30
 * rejects every frozen factorization (the reference's tests/cuda/test_lean_lu.py:224-244). */
int p3smi_topology_set_guard(p3smi_topology* t, double ratio);
/* How a fresh pivoting factorization starts (the klu_factor of p3s/cpp/nr_klu.cpp:642-651).
 * diagonal_first = 1 (default): the static-pivot LU of the plan's order is checked against the
 * threshold rule (every column's diagonal at least 1e-3 of its largest scaled candidate); when
 * every column passes, that is the pivoting factorization (the rule picks the diagonal
 * everywhere) and the case keeps the plan's order; otherwise, and always with 0, the device
 * Gilbert-Peierls factorization chooses the pivots. */
int p3smi_topology_set_pivoting(p3smi_topology* t, int32_t diagonal_first);
/* DC-power-flow start on the device (p3s/timeseries.py:200-244 with the reduced solve of
 * p3s/NewtonPowerflowCpp.py:422-440): solves B'(pvpq, pvpq) theta = P(pvpq) - B'(pvpq, ref)
 * theta_ref with the engine's static-pivot LU, theta_ref the angles of V_flat at the reference
 * buses, and returns V_out (n complex) = V_flat with |V_flat| e^(j theta) at pv and pq buses (a
 * singular system gives angle 0; a non-finite result returns V_flat). Bdc (nnz): the DC
 * susceptances (diagonal + sum 1/x, off-diagonal -1/x) on the Ybus CSR pattern, zero where no
 * branch stamps; P (n): the active injections; ref (nref): reference buses. */
int p3smi_dc_start(p3smi_topology* t, const double* Bdc, const double* P, const double* V_flat, int32_t nref,
                   const int32_t* ref, double* V_out);
/* Refresh Ybus values on the same pattern: nr_klu.Solver.update_Y (p3s/cpp/nr_klu.cpp:999-1011). */
int p3smi_update_y(p3smi_topology* t, const double* Yx, int32_t nnz);

/* ---- host-side checks of the plan (no GPU needed) ------------------------------------
 * Jacobian (CSR by equation: Jp (m+1), Jj/Jx (nnz_j)) and residual R = Sspec - Scalc (m)
 * at V (n complex) with injections S (n complex); the reference's debug_J
 * (p3s/cpp/nr_klu.cpp:1284-1304), in this engine's row/entry order. */
int p3smi_host_jacobian(const p3smi_topology* t, const double* S, const double* V, int32_t* Jp, int32_t* Jj,
                        double* Jx, double* R);
/* Factor J (entry order) with the frozen plan and solve J x = R on the host; x in Jacobian
 * column order. pivots[0..2] = min |pivot|, max |pivot|, sum |pivot|. */
int p3smi_host_factor_solve(const p3smi_topology* t, const double* Jx, const double* R, double* x, double* pivots);
/* Host emulation of the device LU arithmetic (no counterpart in the reference; a check of
 * the schedules): factor J (entry order) and solve J x = R on the one-entry schedule
 * (blocks = 0) or the 2x2-block schedule (blocks = 1) dealt to nslots task lists, in FP64
 * or FP32 factors (single = 1), summing as the kernels do; x in Jacobian column order.
 * pivots[0..1] = min |pivot|, max |pivot|. checks[0] = schedule violations (factor
 * positions written twice or never, read before written, the zero row written, a
 * substitution reading an unsolved value; 0 for a valid schedule), [1] pivot pairs,
 * [2] operand steps. Both schedules give the same x bit for bit. */
int p3smi_host_lu_emulate(const p3smi_topology* t, int32_t blocks, int32_t nslots, int32_t single, const double* Jx,
                          const double* R, double* x, double* pivots, int64_t* checks);
/* Bus -> variable maps and the elimination order (m): xva/xvm (n), perm (m). */
int p3smi_host_maps(const p3smi_topology* t, int32_t* row_p, int32_t* row_q, int32_t* perm);

/* ---- one-call batch solves (host arrays in and out) ----------------------------------
 * nr_klu.Solver.solve_batch (p3s/cpp/nr_klu.cpp:1049-1166):
 *   S (n, T) complex, V0 (n,) complex [v0_per_case = 0] or (n, T) [1].
 *   Outputs V (n, T) complex, iterations (T), converged (T), and, if not NULL, the final
 *   mismatch infinity norm of every case (T), computed on the device in FP64. */
int p3smi_solve_batch(p3smi_topology* t, const double* S, const double* V0, int32_t v0_per_case, int32_t T,
                      int32_t max_iter, double tol, int32_t line_search, int32_t precision, double* V_out,
                      int32_t* iterations, uint8_t* converged, double* mismatch);

/* nr_klu.Solver.solve_batch_contingency (p3s/cpp/nr_klu.cpp:1184-1281) with per-case
 * Ybus magnitude/angle (nnz, L), S (n,) complex, V0 (n, L) complex, pin (n, L) uint8. */
int p3smi_solve_batch_contingency(p3smi_topology* t, const double* Yx_mag, const double* Yx_ang, const double* S,
                                  const double* V0, const uint8_t* pin, int32_t L, int32_t max_iter, double tol,
                                  int32_t line_search, int32_t precision, double* V_out, int32_t* iterations,
                                  uint8_t* converged, double* mismatch);

/* PolarNewtonSolverCUDA.solve_batch_contingency_stamp (p3s/cuda/nr_polar_solver.py:531-557):
 * case c runs on Ybus = base - stamp[:, c]; the stamp is a COO list (k, c, value complex)
 * with each (k, c) at most once. The per-case Ybus is expanded on the device. */
int p3smi_solve_batch_contingency_stamp(p3smi_topology* t, const int32_t* stamp_k, const int32_t* stamp_c,
                                        const double* stamp_val, int64_t n_stamp, const double* S, const double* V0,
                                        const uint8_t* pin, int32_t L, int32_t max_iter, double tol,
                                        int32_t line_search, int32_t precision, double* V_out, int32_t* iterations,
                                        uint8_t* converged, double* mismatch);

/* ---- N-1 outage sets ------------------------------------------------------------------
 * Replaces ContingencyCaseGenerator + solve_contingencies_cpp
 * (p3s/contingency/case_generator.py:488-566, p3s/contingency/solver_cpp.py:64-156).
 * Branch k joins buses from[k], to[k]; ypos (nbr x 4) are the Ybus CSR positions of its
 * (ff, ft, tf, tt) stamp entries and yval (nbr x 4 complex) the stamp values. It couples
 * its ends (is an edge of the island graph) iff its ft or tf stamp is non-zero. Case c
 * removes the branches gbr[gptr[c] .. gptr[c+1]). ref (nref): slack buses; gen_z (n):
 * generator short-circuit |Z| per bus, +inf where a bus has no eligible generator;
 * reslack != 0 gives an island without a slack its lowest-|Z| generator as reference. */
typedef struct p3smi_outages p3smi_outages;
int p3smi_outages_create(p3smi_topology* t, int32_t nbr, const int32_t* from, const int32_t* to, const int32_t* ypos,
                         const double* yval, int32_t n_cases, const int32_t* gptr, const int32_t* gbr, int32_t nref,
                         const int32_t* ref, const double* gen_z, int32_t reslack, p3smi_outages** out);
void p3smi_outages_free(p3smi_outages* o);
/* Islands of every case on the device (union-find): served (n, L) bus has a reference,
 * ref_bus (n, L) the reference bus of its island or -1, pinned (L) number of re-slacked
 * references (_components_after_outage + _resolve_references, case_generator.py:355-430). */
int p3smi_outages_islands(p3smi_outages* o, uint8_t* served, int32_t* ref_bus, int32_t* pinned);
/* p3smi_outages_islands without allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. served (n, L), ref_bus (n, L, may be NULL) and pinned (L) are
 * DEVICE pointers the caller owns; the results equal p3smi_outages_islands' bit for bit. The
 * union-find scratch is a 64 MiB array resident on each GPU for the life of the process
 * (2n ints, plus n 8-byte words with re-slacking, per case; cases run in chunks that fit): a
 * grid needing more than 64 MiB for one case (above ~4 million buses) is rejected with an
 * error. Calls share the buffer: this library serialises its own calls, and calls that
 * overlap on one GPU from elsewhere (another copy of the library) would race. The call
 * returns after its kernels have finished on the topology's stream. */
int p3smi_outages_islands_noalloc(p3smi_outages* o, uint8_t* served, int32_t* ref_bus, int32_t* pinned);
/* Solve every case: S (n) complex injections, V_start (n) complex start voltage. V_out
 * (n, L) complex with NaN at unserved buses; served (n, L) may be NULL. */
int p3smi_outages_solve(p3smi_outages* o, const double* S, const double* V_start, int32_t max_iter, double tol,
                        int32_t line_search, int32_t precision, double* V_out, int32_t* iterations,
                        uint8_t* converged, double* mismatch, uint8_t* served);

/* ---- N-1 post-processing on the device, from the last p3smi_outages_solve of the set
 * (its results stay resident when the set was solved in one chunk) ---------------------
 * Branch results (compute_branch_loading + find_overloads, p3s/contingency/
 * find_overloads.py:128-420): per branch k and case c, S_from = conj(yff V_f + yft V_t) V_f
 * sn_mva, S_to likewise, I = |S| / (vn |V| sqrt 3) in kA, loading = max(I_from, I_to) /
 * rating * 100 (rating = max_i_ka) for a line, max(I vn sqrt 3) / rating * 100 (rating =
 * sn_mva) for a transformer (trafo[k] != 0); 0 on the case's outaged branches, NaN on
 * branches touching an unserved bus and in non-converged cases (there the reactive powers
 * are 0, as the reference's NaN assignment into complex arrays leaves them). Outputs: the (nbr, L)
 * tables asked for in `table` (NULL members are skipped), per branch the largest loading
 * over the cases and the first case reaching it (NaN / -1 when none is defined), and the
 * (branch, case, loading, i_ka) of every loading above `threshold` percent: up to `cap`
 * entries, in no particular order, *n_over = their total number. */
typedef struct p3smi_branch_table {
    double *p_from, *q_from, *p_to, *q_to; /* MW, Mvar */
    double *i_from, *i_to, *i_ka;          /* kA; i_ka = max(i_from, i_to) */
    double* loading;                       /* percent */
} p3smi_branch_table;
int p3smi_outages_branch_flows(p3smi_outages* o, const double* vn_from_kv, const double* vn_to_kv, const double* rating,
                               const uint8_t* trafo, double sn_mva, double threshold, p3smi_branch_table* table,
                               double* max_loading, int32_t* worst_case, int32_t cap, int32_t* ov_branch,
                               int32_t* ov_case, double* ov_loading, double* ov_i_ka, int32_t* n_over);
/* Voltage-band violations (find_voltage_violations, p3s/contingency/voltage_band.py:103-182):
 * |V| < vm_min - tol (kind 0, "low") or > vm_max + tol (kind 1, "high") at served, in-service
 * buses of converged cases; NaN bounds never flag. Up to `cap` entries (bus, case, |V|,
 * kind), in no particular order; *n_viol = their total number. */
int p3smi_outages_voltage_violations(p3smi_outages* o, const double* vm_min, const double* vm_max,
                                     const uint8_t* in_service, double tol, int32_t cap, int32_t* bus, int32_t* case_,
                                     double* vm, uint8_t* kind, int32_t* n_viol);

/* ---- resident batches (inputs uploaded once, solved many times) ---------------------- */
int p3smi_batch_create(p3smi_topology* t, int32_t cases, p3smi_batch** out);
void p3smi_batch_free(p3smi_batch* b);
/* Upload time-series inputs: S (n, cases) complex, V0 (n,) or (n, cases) complex. S = NULL
 * keeps injections built on the device by p3smi_batch_build_sbus. */
int p3smi_batch_upload(p3smi_batch* b, const double* S, const double* V0, int32_t v0_per_case);
/* The time-series injections built on the device, replacing the numba builder of
 * build_sbus_matrix (p3s/timeseries.py:36-141): S[i, t] = s0[i] + sum_k coef[k] *
 * (X[row[k], t] - base[k]) over the contributions k of bus i, [ptr[i], ptr[i+1]), in order,
 * on the imaginary part where imag[k]. s0 (n) complex: base injections; X (nrows, cases)
 * float64: the element series stacked; row/coef/base/imag (nk). The result is the batch's
 * S; then p3smi_batch_upload with S = NULL sets V0 and keeps it. */
int p3smi_batch_build_sbus(p3smi_batch* b, const double* s0, const int32_t* ptr, int32_t nk, const int32_t* row,
                           const double* coef, const double* base, const uint8_t* imag, int32_t nrows, const double* X);
/* The same with X given as nblk blocks of rows, stacked in order: block k is a (blk_rows[k],
 * cols) float64 array at blk[k] (one per element series; no host copy stacks them). cols must
 * equal the batch's case count (the call fails otherwise, before any copy). */
int p3smi_batch_build_sbus_blocks(p3smi_batch* b, const double* s0, const int32_t* ptr, int32_t nk,
                                  const int32_t* row, const double* coef, const double* base, const uint8_t* imag,
                                  int32_t cols, int32_t nblk, const int32_t* blk_rows, const double* const* blk);
/* S (n, cases) complex of the batch, as built or uploaded. */
int p3smi_batch_download_sbus(p3smi_batch* b, double* S_out);
/* Run the whole Newton solve of every case on the device from the uploaded inputs. */
int p3smi_batch_solve(p3smi_batch* b, int32_t max_iter, double tol, int32_t line_search, int32_t precision);
int p3smi_batch_download(p3smi_batch* b, double* V_out, int32_t* iterations, uint8_t* converged, double* mismatch);
/* Free the device workspace the one-call entry points keep between calls (it is
 * allocated again, at the size needed, by the next one-call solve). */
int p3smi_topology_release_cache(p3smi_topology* t);
/* The resident batch the one-call entry points (solve_batch*, outages_solve) of this
 * topology ran on last: owned by the topology, valid until it is freed or a one-call solve
 * of another case count replaces it. For timing (batch_set_timing) and statistics. */
int p3smi_topology_last_batch(p3smi_topology* t, p3smi_batch** out);
/* min |V| and max |V| of every case of the last solve (cases), over the buses a served
 * mask keeps: the voltage-band test of NewtonPowerflowCpp.calculate
 * (p3s/NewtonPowerflowCpp.py:553-584), computed on the device. */
int p3smi_batch_vrange(p3smi_batch* b, double* vmin, double* vmax);
int p3smi_batch_stats(const p3smi_batch* b, p3smi_stats* out);
/* 1 = replay each FP64 Newton iteration as a captured hipGraph (one launch and one
 * synchronisation per phase instead of one launch per kernel); used while kernel timing
 * is off. No counterpart in the reference (its GPU path launches kernel by kernel). */
int p3smi_batch_set_graphs(p3smi_batch* b, int32_t on);
/* The same for the one-call entry points of a topology (solve_batch*, outages_solve). */
int p3smi_topology_set_graphs(p3smi_topology* t, int32_t on);
/* 1 = record HIP events around every kernel of later solves (per-kernel times in stats). */
int p3smi_batch_set_timing(p3smi_batch* b, int32_t on);
/* Algorithmic bytes of one pass of each kernel for one case (eval, lu, step, decide). */
int p3smi_batch_bytes_per_case(const p3smi_batch* b, double* out4);

/* ---- continuous batching (no counterpart in the reference) ---------------------------
 * A resident batch used as a pool of slots: cases are loaded into free slots while other
 * slots iterate, every pass advances all loaded cases by one Newton iteration (cases loaded
 * since the last pass get their start point evaluated first), finished slots are fetched
 * and freed. Each slot iterates exactly as it would alone: no arithmetic mixes slots. */
int p3smi_batch_slots_init(p3smi_batch* b);
/* k cases into slots[0..k): S, V0 (n, k) complex. */
int p3smi_batch_load(p3smi_batch* b, int32_t k, const int32_t* slots, const double* S, const double* V0);
/* One pass; counts (8, may be NULL) = slots per status after it:
 * [active, converged, failed, backtrack, needj, stopped, idle, new]. */
int p3smi_batch_pass(p3smi_batch* b, int32_t max_iter, double tol, int32_t line_search, int32_t precision,
                     int32_t* counts);
/* Slots whose case has finished (converged, failed or stopped), n of them. */
int p3smi_batch_finished(p3smi_batch* b, int32_t* slots, int32_t* n);
/* Results of k finished slots (V (n, k) complex, iterations, converged, mismatch); frees them. */
int p3smi_batch_fetch(p3smi_batch* b, int32_t k, const int32_t* slots, double* V, int32_t* iterations,
                      uint8_t* converged, double* mismatch);

/* ---- batch groups: uneven batching in one launch per kernel ---------------------------
 * Replaces the reference's one-topology-at-a-time batch calls (p3s/cpp/nr_klu.cpp:1049-1166,
 * the same Newton per case) for a set of topologies solved together. Each member is a batch
 * created and uploaded as for p3smi_batch_solve (its topology, S, V0); members may differ in
 * grid size, case count and LU launch shape. p3smi_group_solve runs every member's whole-batch
 * FP64 Newton solve at once: each pass is one launch per kernel (per LU launch shape and
 * evaluation variant) over all members, the passes decided on the device and captured
 * passes_per_graph to a graph launch, so the host waits once per graph launch. A member whose
 * case needs a pivoting factorization takes that pass alone on the host side and rejoins. Every
 * case's result (p3smi_batch_download per member) is bit-identical to p3smi_batch_solve of its
 * batch alone. A group does not own its batches; free the group before them. */
typedef struct p3smi_group p3smi_group;
int p3smi_group_create(p3smi_batch* const* members, int32_t count, p3smi_group** out);
void p3smi_group_free(p3smi_group* g);
int p3smi_group_solve(p3smi_group* g, int32_t max_iter, double tol, int32_t line_search, int32_t precision,
                      int32_t passes_per_graph);
/* Passes of the last group solve: [0] device passes launched (graph launches x passes per
 * graph), [1] passes run on the host side for pivoting factorizations, [2] graph launches. */
int p3smi_group_stats(const p3smi_group* g, int64_t* out3);

#ifdef __cplusplus
}
#endif
#endif /* P3SMI_H */
