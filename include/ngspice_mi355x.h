/*
 * ngspice_mi355x.h -- C ABI of the MI355X transient engine.
 *
 * Plain C types only: host pointers, device pointers (as void* / double*
 * obtained from the engine or from hipMalloc), sizes and mode words.  Every
 * entry point returns OK (0) or one of the reference's error numbers
 * (src/include/ngspice/iferrmsg.h:17-38, sperror.h:15-21) unless stated.
 *
 * The entry points replace, one for one, the reference interfaces cited on
 * each declaration; the batched variants take N circuits in one call.
 */
#ifndef NGSPICE_MI355X_H
#define NGSPICE_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error numbers: iferrmsg.h:17-38, sperror.h:15-21 */
#define NGMI_OK          0
#define NGMI_E_PANIC     1
#define NGMI_E_NOMOD     4
#define NGMI_E_NOTERM    6
#define NGMI_E_BADPARM   7
#define NGMI_E_NOMEM     8
#define NGMI_E_UNSUPP    10
#define NGMI_E_PARMVAL   11
#define NGMI_E_NOTFOUND  14
#define NGMI_E_SINGULAR  102
#define NGMI_E_ITERLIM   103
#define NGMI_E_TIMESTEP  106
#define NGMI_E_DEVICE    120   /* HIP runtime failure (no GPU, launch error) */

/* CKTmode bits, cktdefs.h:171-199 */
#define NGMI_MODETRAN        0x1
#define NGMI_MODEAC          0x2
#define NGMI_MODEDCOP        0x10
#define NGMI_MODETRANOP      0x20
#define NGMI_MODEDCTRANCURVE 0x40
#define NGMI_MODEINITFLOAT   0x100
#define NGMI_MODEINITJCT     0x200
#define NGMI_MODEINITFIX     0x400
#define NGMI_MODEINITSMSIG   0x800
#define NGMI_MODEINITTRAN    0x1000
#define NGMI_MODEINITPRED    0x2000
#define NGMI_MODEUIC         0x10000

/* last error text of the calling thread (never NULL) */
const char *ngmi_last_error(void);

/* Launch-shape knobs (no counterpart in the reference): "eval_waves" and
 * "mos_waves" (0..4) set the occupancy the BSIM4 kernels' register
 * allocation targets for ngmi_bsim4_eval and the batched transient;
 * 0 (the default of both) takes 2, or 1 for a launch whose instances fit
 * one wave per SIMD (a single small circuit: no spills, shorter latency);
 * "streams" (1..4, default 2): with 2, a single-topology transient (with
 * the learned guess only when its network runs fused) alternates the two
 * halves of its circuits on two streams so one half's step kernel overlaps
 * the other half's BSIM4 kernel; 3 and 4 split a plain run (no guess, no
 * profiling) into that many parts on as many streams;
 * "specialise" (0 or 1, default 1): compile the default card's selectors as
 * constants when every instance of a set shares them;
 * "order_slack" (0..4, default 1): the engine's own pivot order may pick a
 * node up to this much above the minimum degree when that shortens the
 * elimination tree (fewer factor and solve levels, possibly more fill);
 * it applies to batches built afterwards;
 * "graph" (0 or 1, default 1): ngmi_batch_tran replays chunks of 32 Newton
 * passes as one captured HIP graph instead of launching each kernel;
 * "guess_fused" (0..2, default 2): the learned guess and ngmi_mlp_forward
 * run the network layer by layer (0), in one fused kernel with the weights
 * in registers and the activations in LDS (1), or in one fused kernel of
 * small workgroups that stream each layer's weights from L2 per 32-row
 * tile (2), which fit beside the other half's BSIM4 kernel on the second
 * stream; all three give the same values bit for bit;
 * "guess_grid" (0..1024, default 64): workgroups of the fused network
 * kernels (guess_fused 1 and 2); 0 = one per compute unit.  Fewer leave
 * compute units to the other half's BSIM4 kernel on the second stream
 * (guess_fused 2 on the adder batch: 0.977x plain at 64, 0.976x at 32,
 * 0.972x at 128, 0.965x at 256, 0.950x at 512).
 * Process-wide; defaults are the measured best on MI355X. */
int ngmi_tune(const char *knob, int value);
/* current value of a knob, or minus NGMI_E_BADPARM for an unknown name */
int ngmi_tune_get(const char *knob);

/* ------------------------------------------------------------------ */
/* BSIM4 model cards                                                   */
/* ------------------------------------------------------------------ */
typedef struct ngmi_bsim4_model ngmi_bsim4_model;

/* replaces creating a BSIM4 GENmodel (DEVmodSize, devdefs.h:84) */
ngmi_bsim4_model *ngmi_bsim4_model_new(void);
void ngmi_bsim4_model_free(ngmi_bsim4_model *m);
/* BSIM4mParam (b4mpar.c:20): netlist name (lower case) -> value; "nmos"/"pmos" set the type.
 * Returns NGMI_E_BADPARM for an unknown name. */
int ngmi_bsim4_model_param(ngmi_bsim4_model *m, const char *name, double value);
/* BSIM4setup model defaults (b4set.c:79-2300) + BSIM4temp model part (b4temp.c:90-417).
 * Temperatures in kelvin.  NGMI_E_UNSUPP for selectors the engine does not build
 * (capMod != 2, rdsMod, rgateMod, rbodyMod, NQS, mtrlMod). */
int ngmi_bsim4_model_setup(ngmi_bsim4_model *m, double tnom_kelvin, double temp_kelvin);
/* a derived model quantity by name ("type", "coxe", "vtm", "vcrit") */
double ngmi_bsim4_model_get(const ngmi_bsim4_model *m, const char *what);

/* ------------------------------------------------------------------ */
/* BSIM4 instance sets: host resolution + device-resident tables       */
/* ------------------------------------------------------------------ */
typedef struct ngmi_bsim4_set ngmi_bsim4_set;

ngmi_bsim4_set *ngmi_bsim4_set_new(void);
void ngmi_bsim4_set_free(ngmi_bsim4_set *s);
/* BSIM4param (b4par.c:22) for n instances at once: instance k uses
 * models[model_idx[k]]; params[j*n + k] is the value of names[j] for
 * instance k.  Appends; returns the index of the first new instance or a
 * negative error number. */
int ngmi_bsim4_set_add(ngmi_bsim4_set *s, ngmi_bsim4_model *const *models, int nmodels,
                       const int *model_idx, int n, const char *const *names, int nparams,
                       const double *params);
/* BSIM4temp instance part (b4temp.c:418-2400) for every instance and upload
 * of the evaluation tables to the current HIP device. */
int ngmi_bsim4_set_finalize(ngmi_bsim4_set *s);
int ngmi_bsim4_set_size(const ngmi_bsim4_set *s);
/* copy of one resolved table field of instance k (field names of the engine's
 * evaluation table, e.g. "tvth0", "leff", "weff", "coxe") */
double ngmi_bsim4_set_field(const ngmi_bsim4_set *s, int k, const char *field);
/* number of per-instance / per-model table columns and table bytes on the device */
int ngmi_bsim4_set_layout(const ngmi_bsim4_set *s, int *inst_cols, int *model_cols,
                          size_t *device_bytes);

/* Per-instance evaluation, BSIM4load (b4ld.c:69-5399) on isolated
 * instances: terminal voltages come from vterm (4 x n, rows d g s b, device
 * memory, circuit polarity), the circuit scalars from the arguments.
 * Device arrays (all n-strided, row r of instance k at [r*n + k]):
 *   state[0..2]   29 rows each (bsim4def.h:534-567 order), read/written;
 *   cache         38 rows, values kept between calls (b4ld.c:444-507);
 *   stamp         16 rows out (Jacobian, row-major over d g s b);
 *   rhs           4 rows out (d g s b);
 *   noncon        n ints out (1 where the device convergence test fails).
 * Asynchronous on stream (a hipStream_t, NULL for the default stream). */
typedef struct {
    int mode, order, method;   /* CKTmode; integration order; 1 trap, 2 Gear */
    double ag[7];              /* NIcomCof coefficients */
    double xfact;              /* CKTdelta / CKTdeltaOld[1] (INITPRED) */
    double gmin, reltol, abstol;
} ngmi_ckt_scalars;

int ngmi_bsim4_eval(const ngmi_bsim4_set *s, const ngmi_ckt_scalars *ck, const double *vterm,
                    double *const state[3], double *cache, double *stamp, double *rhs,
                    int *noncon, void *stream);

/* ------------------------------------------------------------------ */
/* Sparse pattern: COO element list -> CSC + per-element slot map      */
/* ------------------------------------------------------------------ */
/* SMPmakeElt (klusmp.c:417-440) + SMPconvertCOOtoCSC (klusmp.c:137-323):
 * elements (row[e], col[e]) 0-based, n x n.  Writes Ap[n+1], and returns in
 * *nnz the number of distinct entries; Ai must hold ne entries, elt2nz[e]
 * receives the CSC position of element e.  Within a column, rows ascend. */
int ngmi_csc_from_elements(int n, int64_t ne, const int *row, const int *col, int64_t *Ap,
                           int *Ai, int64_t *elt2nz, int64_t *nnz);

/* Deterministic assembly ("CKTload" accumulation): Ax[z] = sum over the
 * elements e mapped to z, in ascending e order, of src[e].  The map is
 * inverted once; the plan lives on the device. */
typedef struct ngmi_gather ngmi_gather;
ngmi_gather *ngmi_gather_new(int64_t nnz, int64_t ne, const int64_t *elt2nz);
void ngmi_gather_free(ngmi_gather *g);
int ngmi_gather_run(const ngmi_gather *g, const double *src, double *Ax, void *stream);

/* ------------------------------------------------------------------ */
/* One large sparse system on the whole GPU (multifrontal LU)          */
/* ------------------------------------------------------------------ */
/* For a single circuit too large for the per-circuit LDS programs (the
 * 1024 x 1024 power-grid mesh, ~1.1M unknowns).  n x n, CSC with 0-based
 * rows ascending per column, as ngmi_csc_from_elements writes it.
 *
 * ngmi_splu_analyze replaces klu_analyze (klu_analyze.c:257-488, called from
 * SMPpreOrder, klusmp.c:1255-1285) and fixes the pivot sequence there, where
 * KLU's klu_factor (klu_factor.c:15-260) picks pivots numerically: a column
 * with no structural diagonal (a voltage-source branch) is paired with a
 * node it is incident to and their rows exchanged; the order is nested
 * dissection.  Host-only: it touches no GPU.  NGMI_E_SINGULAR when an
 * unknown without a diagonal has no node to pair with. */
typedef struct ngmi_splu ngmi_splu;
int ngmi_splu_analyze(int n, const int64_t *Ap, const int *Ai, ngmi_splu **out);
void ngmi_splu_free(ngmi_splu *s);
/* sizes: fronts, tree heights, largest front, front storage in doubles,
 * nonzeros of L + U (dense fronts), factorization flops */
int ngmi_splu_info(const ngmi_splu *s, int64_t *fronts, int64_t *heights, int64_t *max_front,
                   int64_t *front_doubles, int64_t *lu_nonzeros, double *flops);
/* front f (postorder): pivot count, boundary count, parent (-1 at a root),
 * and its ns + nb unknowns (pivots, then boundary) */
int ngmi_splu_front(const ngmi_splu *s, int f, int *ns, int *nb, int *parent, int *idx);
/* row r of the factored matrix is row rowsrc[r] of A (n entries) */
int ngmi_splu_rowsrc(const ngmi_splu *s, int *rowsrc);
/* klu_refactor (klu_refactor.c:18-482, via SMPluFac, klusmp.c:573-624):
 * numeric factorization of the device values Ax (nnz doubles, the pattern's
 * order).  The first call allocates the device fronts.  Synchronizes stream;
 * NGMI_E_SINGULAR on a zero or non-finite pivot. */
int ngmi_splu_factor(ngmi_splu *s, const double *Ax, void *stream);
/* klu_solve (klu_solve.c:14-400, via SMPsolve, klusmp.c:950-1011): b (n
 * doubles, device) is overwritten with the solution.  Asynchronous.
 * NGMI_E_SINGULAR when the last ngmi_splu_factor did not succeed (or none
 * ran), where klu_solve would run on the failed factors. */
int ngmi_splu_solve(ngmi_splu *s, double *b, void *stream);
/* the largest |l| among the multipliers of the last successful
 * ngmi_splu_factor (no counterpart in the reference).  The pivot sequence
 * is fixed by ngmi_splu_analyze, with no threshold test: where KLU's
 * partial pivoting (klu_kernel.c:370-530, tol = pivrel = 1e-3) would have
 * exchanged a pivot, its column holds a multiplier above 1/tol (up to KLU's
 * row scaling), so a value above 1e3 flags a factorization that lost
 * accuracy KLU would have kept.  NGMI_E_SINGULAR when the last factor
 * failed or none ran. */
int ngmi_splu_multiplier(const ngmi_splu *s, double *lmax);

/* ------------------------------------------------------------------ */
/* One large circuit on the whole GPU: transient analysis              */
/* ------------------------------------------------------------------ */
/* For one circuit too large for the batched per-circuit programs (the
 * 1024 x 1024 power-grid mesh with its inverter loads, ~1.1M unknowns).
 * ngmi_large_new takes a circuit built with the ngmi_circuit_* calls (its
 * models must outlive the object): CSC pattern and symbolic LU
 * (ngmi_splu_analyze, the klu_analyze of SMPpreOrder), the signed assembly
 * plans of CKTload (cktload.c:31-180) and the BSIM4 tables; probe lists
 * the 1-based unknowns recorded at each accepted point (NULL: all),
 * max_points bounds the record.  NULL on error (ngmi_last_error).
 * .ic values and Gear orders above 2 are not supported. */
typedef struct ngmi_large ngmi_large;
ngmi_large *ngmi_large_new(const struct ngmi_circuit *c, const int *probe, int nprobe, int64_t max_points);
void ngmi_large_free(ngmi_large *L);
/* unknowns, matrix nonzeros, L+U nonzeros and refactor flops */
int ngmi_large_info(const ngmi_large *L, int64_t *n_unknowns, int64_t *nnz, int64_t *lu_nonzeros, double *flops);
/* .tran tstep tstop [tmax]: CKTop's direct Newton (cktop.c:26-116 without
 * the stepping fallbacks), then DCtran (dctran.c:66-971) with NIiter
 * (niiter.c:28-367) driven from the host, the loads, LU (ngmi_splu_*),
 * NIconvTest and CKTterr on the device of the calling thread, on stream.
 * Returns when the run ends; NGMI_E_TIMESTEP, NGMI_E_ITERLIM (operating
 * point) or NGMI_E_SINGULAR (operating point) on failure. */
int ngmi_large_tran(ngmi_large *L, double tstep, double tstop, double tmax, void *stream);
/* recorded points, accepted / rejected timepoints, Newton iterations */
int ngmi_large_stats(const ngmi_large *L, int64_t *npts, int64_t *accepted, int64_t *rejected, int64_t *iterations);
/* the largest |l| multiplier over the last ngmi_large_tran's
 * factorizations (see ngmi_splu_multiplier; no counterpart in the
 * reference) */
int ngmi_large_multiplier(const ngmi_large *L, double *lmax);
/* the record: t[npts], x[npts][nprobe] */
int ngmi_large_waveform(const ngmi_large *L, double *t, double *x, int64_t maxpts, int64_t *npts);
/* KCL at the last accepted solution x: the circuit equations F(x) =
 * A(x) x - b(x) assembled there (one more load in MODETRAN | INITFLOAT);
 * *scaled = max over rows of |F| / (sum |A x| + |b|), per_kind[0..1] =
 * max |F| over node rows (amperes) and branch rows (volts).  After a run. */
int ngmi_large_kcl(ngmi_large *L, double *scaled, double *per_kind);
/* host wall milliseconds of the last run by phase: loads and assembly,
 * refactor, solve, convergence test */
int ngmi_large_times(const ngmi_large *L, double *load_ms, double *factor_ms, double *solve_ms, double *other_ms);

/* ------------------------------------------------------------------ */
/* One circuit in host memory: the cores a DEVload / SMP binding calls */
/* ------------------------------------------------------------------ */
/* For a reference-shaped binding (DEVload, devdefs.h:57; SMP*,
 * smpdefs.h:61-86) of a single circuit whose CKTcircuit vectors and KLU
 * matrix live in host memory.
 *
 * ngmi_devload_new is the setup half of BSIM4setup + BSIM4bindCSC
 * (b4set.c:2318 state offsets, b4bindCSC.c:12-111 matrix pointers): set
 * holds the circuit's BSIM4 instances (finalized); instance i has terminal
 * unknowns nodes[4i..4i+3] (d g s b, 0 = ground, at most n_unknowns), its
 * 29 states at state_off[i] of each state vector (num_states long), and
 * its 16 Jacobian entries (row-major over d g s b) at positions
 * ax_pos[16i..16i+15] of the caller's matrix array (-1: not stamped, a
 * ground row or column).  Device memory is allocated here. */
typedef struct ngmi_devload ngmi_devload;
ngmi_devload *ngmi_devload_new(const ngmi_bsim4_set *set, int n_unknowns, int num_states, const int *nodes,
                               const int *state_off, const int64_t *ax_pos);
void ngmi_devload_free(ngmi_devload *d);
/* BSIM4load (b4ld.c:69-5399) for those instances: reads rhsOld[0..n] and
 * states[0..2] (host arrays, CKTrhsOld and CKTstate0..2), writes back
 * state0 (and state1, which INITTRAN copies), adds each instance's stamps
 * into Ax at its positions and its RHS values into rhs[node], instance by
 * instance in set order, and adds the device convergence votes to
 * *noncon (CKTnoncon, b4ld.c:4072-4110).  Synchronous. */
int ngmi_devload_run(ngmi_devload *d, const ngmi_ckt_scalars *ck, const double *rhsOld, double *const states[3],
                     double *Ax, double *rhs, int *noncon);

/* SMPpreOrder (klusmp.c:1255-1285): symbolic analysis of an n x n CSC
 * pattern (0-based rows, ascending per column) by the whole-GPU sparse LU
 * (ngmi_splu_analyze); host-only. */
typedef struct ngmi_smp ngmi_smp;
int ngmi_smp_new(int n, const int64_t *Ap, const int *Ai, ngmi_smp **out);
void ngmi_smp_free(ngmi_smp *m);
/* SMPluFac / SMPreorder (klusmp.c:573-624, 752-806): LoadGmin_CSC
 * (klusmp.c:1765-1777, gmin onto every existing diagonal entry) on a device
 * copy of Ax (host, nnz values; the caller's array is not changed), then
 * the numeric factorization on the pivot sequence fixed at analysis.
 * NGMI_E_SINGULAR on a zero or non-finite pivot, as KLU_SINGULAR maps to
 * E_SINGULAR.  Synchronous. */
int ngmi_smp_factor(ngmi_smp *m, const double *Ax, double gmin);
/* SMPsolve (klusmp.c:950-1011): rhs[1..n] (host, 1-based, rhs[0] the
 * ground slot) solved in place; rhs[0] is set to 0.  NGMI_E_SINGULAR when
 * the last factorization failed.  Synchronous. */
int ngmi_smp_solve(ngmi_smp *m, double *rhs);

/* ------------------------------------------------------------------ */
/* Circuits and batched transient analysis                             */
/* ------------------------------------------------------------------ */
typedef struct ngmi_circuit ngmi_circuit;
typedef struct ngmi_batch ngmi_batch;
typedef struct ngmi_mlp ngmi_mlp;

/* A netlist with nodes 1..nnodes (0 = ground): the engine's counterpart of
 * building a CKTcircuit (CKTinit + the device setup calls of CKTsetup,
 * cktsetup.c:33-210).  Element values are per circuit; circuits of equal
 * topology in one batch share their symbolic analysis. */
ngmi_circuit *ngmi_circuit_new(int nnodes);
void ngmi_circuit_free(ngmi_circuit *c);
/* one BSIM4 instance (BSIM4param, b4par.c:22, for its instance parameters) */
int ngmi_circuit_add_bsim4(ngmi_circuit *c, ngmi_bsim4_model *m, int d, int g, int s, int b,
                           int nparams, const char *const *names, const double *values);
/* n instances at once: instance k uses models[model_idx[k]], terminals
 * nodes[4k..4k+3] (d g s b) and parameter j value values[j*n + k] */
int ngmi_circuit_add_bsim4_many(ngmi_circuit *c, ngmi_bsim4_model *const *models,
                                const int *model_idx, const int *nodes, int n, int nparams,
                                const char *const *names, const double *values);
/* capacitor (CAPload, capload.c:19-86), resistor (RESload, resload.c:16-38) */
int ngmi_circuit_add_cap(ngmi_circuit *c, int p, int n, double farad);
int ngmi_circuit_add_res(ngmi_circuit *c, int p, int n, double ohm);
/* independent voltage source (VSRCload, vsrcload.c:40-170): DC when pulse is
 * NULL, else PULSE(v1 v2 td tr tf pw per) with npulse of the 7 values given */
int ngmi_circuit_add_vsrc(ngmi_circuit *c, int p, int n, double dc, const double *pulse,
                          int npulse);
/* .ic v(node)=v for the transient operating point (cktload.c:146-173) */
int ngmi_circuit_set_ic(ngmi_circuit *c, int node, double v);
/* .options reltol abstol vntol chgtol trtol gmin method(1 trap, 2 gear) maxord */
int ngmi_circuit_set_options(ngmi_circuit *c, double reltol, double abstol, double vntol,
                             double chgtol, double trtol, double gmin, int method, int maxord);
/* .options noopiter gminsteps srcsteps gminfactor gshunt itl2 for the
 * transient's operating point, CKTop (cktop.c:26-116): direct Newton unless
 * noopiter, then dynamic_gmin and new_gmin (gminsteps = 1,
 * cktop.c:161-478) or spice3_gmin in gminsteps steps (gminsteps > 1,
 * cktop.c:289-347), then gillespie_src (srcsteps = 1, cktop.c:480-670) or
 * spice3_src in srcsteps steps (srcsteps > 1, cktop.c:672-720); defaults
 * 0, 1, 1, 10, 0, 50 (cktntask.c:94-132).  A batch uses the options of
 * each topology group's first circuit. */
int ngmi_circuit_set_op_options(ngmi_circuit *c, int noopiter, int gminsteps, int srcsteps,
                                double gminfactor, double gshunt, int itl2);
/* Pivot sequence for the static LU: step k eliminates row rows[k] and
 * column cols[k] (1-based MNA unknowns, n = nodes + voltage sources).  This
 * is where KLU's own ordering enters: after SMPpreOrder's klu_analyze
 * (klusmp.c:1255-1285) and SMPreorder's klu_factor (klusmp.c:752-806) the
 * caller passes Numeric->Pnum[k] + 1 and Symbolic->Q[k] + 1, and every later
 * refactor runs on that sequence as klu_refactor does (klu_refactor.c:
 * 18-482).  n = 0 restores the engine's own ordering (source rows paired
 * with their node rows, then minimum degree).  Circuits of one batch with
 * different sequences (or options) run as separate topology groups. */
int ngmi_circuit_set_pivot_order(ngmi_circuit *c, int n, const int *rows, const int *cols);
/* CKTpivotRelTol (.options pivrel, default 1e-3), the tolerance of KLU's
 * threshold partial pivoting (klu_kernel.c:370-530 via SMPreorder,
 * klusmp.c:752-806): in (0, 1].  Where NIiter factors afresh -- INITJCT,
 * the first INITTRAN iteration, the iteration after either, and after a
 * refactor met a zero pivot (niiter.c:107-195, 335-347) -- the engine tests
 * whether klu_factor would keep the static pivot sequence (each diagonal
 * candidate at least pivrel times the column's largest candidate, values
 * scaled by their row maxima, klu_scale.c:19-159); where it would not, or
 * a static pivot is zero, the circuit factors with KLU's pivot search on a
 * dense workspace and refactors on those rows until the next fresh
 * factorization.  A zero pivot in a refactor is factored afresh in the
 * same iteration; a zero pivot there is E_SINGULAR. */
int ngmi_circuit_set_pivrel(ngmi_circuit *c, double pivrel);
/* optran (optran.c:285-851, set by the 'optran' command, com_optran): when
 * CKTop's Newton run and its gmin and source stepping all fail, the
 * operating point is sought as a transient from the state they left, the
 * sources held at their time-0 values, in steps of at most opstep up to
 * opfinal (no output), and the transient starts from where it ends.
 * Defaults: ngspice's own 'optran 1 1 1 100n 10u 0' (init.c:77-93), i.e.
 * opstep 100 ns, opfinal 10 us; a step above opfinal / 50 becomes
 * opfinal / 50; opstep = 0 deselects it (the operating point then fails
 * with the stepping's error).  Supply ramping (opramptime) is not
 * offered. */
int ngmi_circuit_set_optran(ngmi_circuit *c, double opstep, double opfinal);
/* uic = 1: '.tran ... uic' (TRANuic, traninit.c; MODEUIC).  The operating
 * point is skipped: CKTic puts the .ic values into the solution vector (0
 * at the other unknowns, cktic.c:19-80), each BSIM4 instance without its
 * own icvds/icvgs/icvbs takes them from the .ic values of its terminals
 * (BSIM4getic, b4getic.c:20-55) and each capacitor its initial voltage
 * (CAPgetic, capgetic.c:19-35), and CKTop's NIiter is a single load on
 * that vector (niiter.c:41-46).  The transient starts from the state that
 * load leaves, with a breakpoint at tstep (dctran.c:509-512); the time-0
 * point is not output (dctran.c:416-419).  noopiter is not taken under
 * uic.  The .ic values are no longer forced during the operating point
 * (cktload.c:118-150 is skipped under MODEUIC).  Default 0; the large-
 * circuit transient rejects it. */
int ngmi_circuit_set_uic(ngmi_circuit *c, int uic);
/* on = 1: without a pivot sequence from the caller, the static LU takes
 * KLU's own ordering, computed by the engine (ngmi_klu_analyze on the
 * circuit's MNA pattern: P as the pivot rows, Q as the pivot columns, what
 * SMPpreOrder's klu_analyze fixes, klusmp.c:1255-1285); on = 0 (the
 * default): the engine's own ordering. */
int ngmi_circuit_use_klu_order(ngmi_circuit *c, int on);
/* KLU's symbolic analysis on the host, restated (no device work): the
 * permutations klu_analyze computes with KLU's defaults -- a block upper
 * triangular form (maximum transversal, then strongly connected components;
 * btf_order.c:36-132) and AMD on each block larger than 3 x 3
 * (klu_analyze.c:14-250, amd_order.c, amd_2.c) -- for the n x n CSC
 * pattern (Ap, Ai), 0-based.  P[k], Q[k]: row and column of pivot k;
 * R[0..nblocks]: block boundaries (R needs n + 1 entries); nzoff: entries
 * outside the diagonal blocks; structural_rank: the matching size.  Equal
 * to Symbolic->P, Q, R of the reference on the same pattern; replaces
 * SMPpreOrder's klu_analyze call (klusmp.c:1255-1285).  Any of nblocks,
 * nzoff, structural_rank may be NULL. */
int ngmi_klu_analyze(int n, const int *Ap, const int *Ai, int *P, int *Q, int *R, int *nblocks, int *nzoff,
                     int *structural_rank);
/* Host-only symbolic analysis of the circuit's topology (no device work):
 * unknowns, L+U nonzeros, factor terms, strictly-lower L nonzeros, levels
 * [factor, forward, backward] and the pivot sequence used (rows/cols,
 * n_unknowns entries each; any output may be NULL). */
int ngmi_circuit_analyze(const ngmi_circuit *c, int *n_unknowns, int *nslots, int *nterms,
                         int *nnz_l, int *levels, int *rows, int *cols);

/* A batch of C circuits on the current device: SMPpreOrder/SMPreorder
 * (klusmp.c:752-806, 1255-1285) once per distinct topology, device tables,
 * state and waveform buffers (probe: unknowns recorded at every accepted
 * point, NULL = all of each circuit; max_points per circuit).  Circuits may
 * differ in topology and size: circuits with equal topology form a group
 * sharing one symbolic analysis, and every Newton pass is one launch over
 * all groups. */
ngmi_batch *ngmi_batch_new(ngmi_circuit *const *ckts, int C, const int *probe, int nprobe,
                           int64_t max_points);
void ngmi_batch_free(ngmi_batch *b);
/* unknowns, L+U nonzeros, factor terms, and levels [factor, forward,
 * backward] of the group of circuit 0 */
int ngmi_batch_info(const ngmi_batch *b, int *n_unknowns, int *nslots, int *nterms, int *levels);
/* number of topology groups */
int ngmi_batch_groups(const ngmi_batch *b, int *ngroups);
/* circuit c's group, unknowns, recorded unknowns per point and L+U nonzeros */
int ngmi_batch_circuit_info(const ngmi_batch *b, int c, int *group, int *n_unknowns, int *nprobe,
                            int *nslots);
/* .tran tstep tstop [tmax] for every circuit: CKTop then DCtran
 * (dctran.c:66-971) with NIiter (niiter.c:28-367) per circuit, all on the
 * device; returns when every circuit has finished or failed, or after
 * max_passes Newton passes (NGMI_E_ITERLIM).  Asynchronous work on stream,
 * synchronised before return. */
int ngmi_batch_tran(ngmi_batch *b, double tstep, double tstop, double tmax, int64_t max_passes,
                    void *stream, int64_t *passes);
/* per-instance and per-model BSIM4 table columns of the batch */
int ngmi_batch_table_cols(const ngmi_batch *b, int *inst_cols, int *model_cols);
/* record HIP events around every kernel of every pass of ngmi_batch_tran
 * (on its stream) and report the summed kernel milliseconds, the BSIM4
 * evaluations and the per-circuit Newton passes of the last run */
int ngmi_batch_set_profile(ngmi_batch *b, int enable);
int ngmi_batch_kernel_times(const ngmi_batch *b, double *mos_ms, double *step_ms,
                            unsigned long long *mos_evals, unsigned long long *newton_passes);
/* Learned Newton initial guess (no counterpart in the reference; it stands
 * in for the device-level predictor of MODEINITPRED, b4ld.c:322-372 and
 * capload.c:53-66): with a network attached, every Newton run of a step
 * after the first four accepted points starts from node voltages
 * v0 + r (v0 - v1) + net(features) instead of the predictor.  Convergence
 * criteria are unchanged; a run that fails redoes the step from the
 * predictor.  NULL detaches.  The network must outlive the batch's runs. */
int ngmi_batch_set_guess(ngmi_batch *b, ngmi_mlp *net);
/* network rows (nodes given a learned starting value) of the last
 * ngmi_batch_tran, summed over its passes; 0 without a network */
int ngmi_batch_guess_rows(const ngmi_batch *b, int64_t *rows);
/* the guess features' neighbour table: [nnodes+1][8] node numbers, -1 = none */
int ngmi_batch_neighbours(const ngmi_batch *b, int *out);
/* per circuit: steps started from the guess, fallbacks to the predictor
 * (a Newton run from the guess that failed), steps started from the
 * predictor while the circuit's guess was paused (after three guessed
 * steps in a row that needed more than two iterations, for 32 steps, then
 * twice as long each time); guess kernels' total milliseconds of the last
 * profiled run */
int ngmi_batch_guess_stats(const ngmi_batch *b, int c, int64_t *guessed, int64_t *fallbacks,
                           int64_t *paused, double *guess_ms);
/* Live data stream, the batched counterpart of SendData / OUTpData
 * (sharedspice.c:2280, outitf.c:633): while ngmi_batch_tran runs, after
 * every 32 Newton passes and once at its end, cb receives each watched
 * circuit's points recorded since its last call, on the calling thread:
 * circuit, index of the first point, count, t[count] and x[count][nprobe].
 * Its return value is ignored, as SendData's is.  circuits = NULL watches
 * all; cb = NULL turns the stream off.  Points past max_points are not
 * recorded and are not sent. */
typedef int (*ngmi_data_cb)(int circuit, int64_t first, int64_t count, const double *t, const double *x,
                            int nprobe, void *user);
int ngmi_batch_set_data_callback(ngmi_batch *b, ngmi_data_cb cb, const int *circuits, int ncircuits,
                                 void *user);
/* per circuit: status (0 finished, -1 running, else the error number),
 * recorded points, accepted / rejected timepoints, Newton iterations */
int ngmi_batch_stats(const ngmi_batch *b, int c, int *status, int64_t *npts, int64_t *accepted,
                     int64_t *rejected, int64_t *iterations);
/* circuit c's pivoting (ngmi_circuit_set_pivrel): pv = 1 while it
 * refactors on pivot rows of its own, the counts of pivoted factorizations
 * (klu_factor) and of refactors on their rows (klu_refactor), and the pivot
 * sequence in use (rows[k], cols[k]: the unknowns, 1-based, whose row and
 * column are the k-th pivot; Numeric->Pnum and Symbolic->Q on KLU's
 * ordering).  rows, cols may be NULL; otherwise n_unknowns long. */
int ngmi_batch_pivot_stats(const ngmi_batch *b, int c, int *pv, int64_t *factors, int64_t *refactors, int *rows,
                           int *cols);
/* waveform of circuit c (CKTdump rows, cktdump.c:33-44): t[n], x[n][nprobe] */
int ngmi_batch_waveform(const ngmi_batch *b, int c, double *t, double *x, int64_t maxpts,
                        int64_t *npts);
/* Circuit c's waveform as a raw file in the layout of ngspice's batch
 * output (outitf.c:881-1110: header, "Variables:", then "Values:" rows of
 * 15-digit values, or "Binary:" rows of native doubles when binary != 0).
 * names[k] names recorded unknown k (a node name, or "<source>#branch" for
 * a source current, written v(...) / i(...) as guess_type does,
 * outitf.c:928-990); NULL names use the unknowns' indices. */
int ngmi_batch_write_raw(const ngmi_batch *b, int c, const char *path, const char *title,
                         const char *const *names, int binary);

/* Autotuner of the launch knobs (no counterpart in the reference): times
 * trial_passes Newton passes of the batched transient of ckts from t = 0
 * under each candidate value of order_slack, mos_waves, streams and graph
 * (see ngmi_tune), one knob at a time from the current settings, keeps the
 * fastest (a change must win by 2 %), leaves the process-wide knobs set to
 * the choice, appends it to the text file cache_path (NULL: none) under the
 * batch's shape, and writes the timings to report.  Builds and frees its
 * own batches. */
int ngmi_autotune_batch(ngmi_circuit *const *ckts, int C, double tstep, double tstop, int64_t trial_passes,
                        const char *cache_path, char *report, int report_len);
/* applies the newest cached choice for the shape of ckts;
 * NGMI_E_NOTFOUND when the cache has none */
int ngmi_autotune_load(const char *cache_path, ngmi_circuit *const *ckts, int C);

/* ------------------------------------------------------------------ */
/* SPICE decks                                                         */
/* ------------------------------------------------------------------ */
typedef struct ngmi_netlist ngmi_netlist;

/* Parses a deck (the reference's input language, inp_readall /
 * inp_subcktexpand / INPpas1-2, for the subset a BSIM4 transient uses:
 * M C R V X elements, .subckt, .model nmos|pmos level=54, .param with
 * expressions, .tran, .options, .temp, .ic, .global).  Unknowns are
 * numbered as ngspice numbers them: nodes by first appearance in the
 * flattened deck (CKTnewNode, cktnewn.c), then source branches in
 * VSRCsetup order, the reverse of deck order (vsrcset.c:41-44,
 * cktcrte.c:62-64).  seed drives agauss/gauss/aunif/unif.  NULL on error
 * (ngmi_last_error names the line). */
ngmi_netlist *ngmi_netlist_parse(const char *text, uint64_t seed);
void ngmi_netlist_free(ngmi_netlist *nl);
const char *ngmi_netlist_title(const ngmi_netlist *nl);
int ngmi_netlist_size(const ngmi_netlist *nl, int *n_nodes, int *n_unknowns, int *n_elements);
/* name of unknown k (0 = ground "0"; nodes; then "<source>#branch") */
const char *ngmi_netlist_name(const ngmi_netlist *nl, int k);
/* unknown number of a node or "<source>#branch" name, -1 if absent */
int ngmi_netlist_find(const ngmi_netlist *nl, const char *name);
/* .tran tstep tstop [tstart [tmax]] [uic] */
int ngmi_netlist_tran(const ngmi_netlist *nl, double *tstep, double *tstop, double *tstart,
                      double *tmax, int *uic);
/* flattened element k in deck order: kind ('m' 'c' 'r' 'v'), its terminal
 * unknowns (up to 4) and value (C, R, or the source's DC value); returns
 * the terminal count, -1 when k is out of range.  Element names follow
 * ngspice's flattening (translate_inst_name, subckt.c:1130-1143). */
int ngmi_netlist_element(const ngmi_netlist *nl, int k, char *kind, int *nodes, double *value);
const char *ngmi_netlist_element_name(const ngmi_netlist *nl, int k);
/* a new circuit of the deck (free with ngmi_circuit_free); its models
 * belong to the netlist, which must outlive the batches built from it */
ngmi_circuit *ngmi_netlist_circuit(const ngmi_netlist *nl);

/* ------------------------------------------------------------------ */
/* Newton initial-guess network (no counterpart in the reference)      */
/* ------------------------------------------------------------------ */
/* 16 -> 256 -> 256 -> 256 -> 1, SiLU, BF16 weights on the matrix cores
 * with FP32 accumulation; FP32 master weights and Adam.  Parameters are one
 * FP32 array of ngmi_mlp_num_params() values (input layer stored 32 wide). */
typedef struct ngmi_mlp ngmi_mlp;
ngmi_mlp *ngmi_mlp_new(unsigned long long seed);
void ngmi_mlp_free(ngmi_mlp *n);
int64_t ngmi_mlp_num_params(void);
/* copy parameters (grads = 0) or the last gradients (grads = 1) to host */
int ngmi_mlp_get_params(const ngmi_mlp *n, float *host, int grads);
int ngmi_mlp_set_params(ngmi_mlp *n, const float *host);
/* y[M] = net(X[M][16]), device pointers, asynchronous on stream */
int ngmi_mlp_forward(ngmi_mlp *n, const float *X, float *y, int M, void *stream);
/* ngmi_mlp_forward without allocating device memory, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  The work buffers are a fixed per-device
 * scratch of NGMI_MLP_NOALLOC_MAX_ROWS rows (17 MB, held by every GPU the
 * library runs on until the process ends); larger M returns
 * NGMI_E_PARMVAL.  Every call on one GPU shares that scratch, so calls that
 * overlap on one GPU (other streams or threads) race.  Same results as
 * ngmi_mlp_forward. */
#define NGMI_MLP_NOALLOC_MAX_ROWS 16384
int ngmi_mlp_forward_noalloc(const ngmi_mlp *n, const float *X, float *y, int M, void *stream);
/* MSE gradients over M samples; lr > 0 also applies one Adam step; the loss
 * before the step goes to *loss_host when given (synchronises) */
int ngmi_mlp_train_step(ngmi_mlp *n, const float *X, const float *target, int M, float lr,
                        float *loss_host, void *stream);

/* The guess network's feature rows, computed by the same device function
 * the batched transient's features kernel uses, for M caller-supplied
 * windows (device arrays): row i from the last four accepted times
 * t[4i..4i+3] and values v[4i..4i+3] of a node (newest first), the step
 * h[i], neighbour values nbv[8i..8i+7] present where bit k of nbmask[i] is
 * set.  F[16i..16i+15] receives the row and act[i] the node's activity;
 * rows the network cannot matter for get zeros and act[i] <= 0: activity
 * below 10 uV, or 4 x activity (the largest correction the network may
 * make) within the Newton tolerance reltol |v0| + vntol.  For training
 * pipelines that build rows on the host and check them against the
 * engine's (no counterpart in the reference).  Asynchronous on stream. */
int ngmi_guess_features(int M, const double *t, const double *h, const double *v, const double *nbv,
                        const unsigned *nbmask, double reltol, double vntol, float *F, double *act,
                        void *stream);

/* ------------------------------------------------------------------ */
/* Device memory helpers for callers without a HIP binding             */
/* ------------------------------------------------------------------ */
void *ngmi_dev_alloc(size_t bytes);
int ngmi_dev_free(void *p);
int ngmi_copy_h2d(void *dst, const void *src, size_t bytes);
int ngmi_copy_d2h(void *dst, const void *src, size_t bytes);
int ngmi_dev_memset(void *dst, int byte, size_t bytes);
int ngmi_sync(void);
/* device name, compute units, and clock (kHz) of the current device */
int ngmi_device_info(char *name, int name_len, int *cus, int *clock_khz);

/* HIP events around work on a stream, for callers that time kernels on the
 * stream they launch on */
typedef struct ngmi_timer ngmi_timer;
ngmi_timer *ngmi_timer_new(void);
void ngmi_timer_free(ngmi_timer *t);
int ngmi_timer_start(ngmi_timer *t, void *stream);
int ngmi_timer_stop(ngmi_timer *t, void *stream);
/* milliseconds between start and stop (synchronises on the stop event) */
double ngmi_timer_ms(ngmi_timer *t);

#ifdef __cplusplus
}
#endif
#endif
