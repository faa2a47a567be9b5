/*
 * xyce_amd.h -- C ABI of the MI355X device-load library.
 *
 * The library evaluates nonlinear semiconductor device instances (junction
 * diode, Level-1 MOSFET, Gummel-Poon BJT, BSIM4 4.6.1) on one AMD Instinct
 * GPU and stamps their DAE contributions: the F and Q residual vectors, the
 * voltage-limiter vectors dFdx*dVp and dQdx*dVp, lead currents, and the dF/dx
 * and dQ/dx Jacobians held as CSR value arrays.
 *
 * Every entry point takes plain pointers and sizes.  Pointers to circuit
 * vectors and CSR value arrays are DEVICE pointers unless stated otherwise.
 * All GPU work is ordered on the context's stream; calls are not re-entrant
 * for one context.  Errors are returned as XA_ERR_* codes with a message in
 * xa_last_error(); nothing throws across the ABI.
 *
 * Reference interfaces replaced (Xyce 6.2 source tree):
 *   Device::updateState / updateSecondaryState / loadDAEVectors /
 *   loadDAEMatrices           src/DeviceModelPKG/Core/include/N_DEV_Device.h
 *                             (default loops N_DEV_DeviceMaster.h:643-725)
 *   per-model Master overrides, e.g. Diode::Master N_DEV_Diode.C:2051-2171,
 *   MOSFET1::Master N_DEV_MOSFET1.C:4069-4688, BJT::Master N_DEV_BJT.C:3891-4366,
 *   MOSFET_B4::Master N_DEV_MOSFET_B4.C:16180-17268
 *   Instance::registerLIDs / registerStateLIDs / registerStoreLIDs /
 *   jacobianStamp / registerJacLIDs   (e.g. N_DEV_Diode.C:479-687)
 *   Model / Instance constructors + processParams + updateTemperature
 *                             (parameter intake, e.g. N_DEV_Diode.C:1816-1940)
 */
#ifndef XYCE_AMD_H
#define XYCE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XA_ABI_VERSION 2

/* ---- status codes ------------------------------------------------------ */
#define XA_OK              0
#define XA_ERR_ARG        -1   /* bad argument / unknown parameter name      */
#define XA_ERR_STATE      -2   /* call made in the wrong setup phase         */
#define XA_ERR_HIP        -3   /* HIP runtime error                          */
#define XA_ERR_UNSUPPORTED -4  /* model option outside the implemented set   */
#define XA_ERR_NOMEM      -5
#define XA_ERR_DEVICE     -6   /* a device instance reported failure         */

/* ---- model types (device letter, level) -------------------------------- */
#define XA_DIODE    1   /* "d", levels 1,2       N_DEV_Diode.C:2179-2186     */
#define XA_MOSFET1  2   /* "m", level 1          N_DEV_MOSFET1.C:4984-4996   */
#define XA_BJT      3   /* "q", level 1          N_DEV_BJT.C:4368-4380       */
#define XA_BSIM4    4   /* "m", levels 14,54     N_DEV_MOSFET_B4.C:17270-17280 */
#define XA_RESISTOR 5   /* "r", level 1          N_DEV_Resistor.C:1198-1224  */
#define XA_CAPACITOR 6  /* "c", level 1          N_DEV_Capacitor.C:2152-2177 */
#define XA_VSOURCE  7   /* "v", level 1          N_DEV_Vsrc.C:1090-1097      */

/* ---- output kinds of a load ------------------------------------------- */
#define XA_KIND_F        0   /* daeFVector rows                              */
#define XA_KIND_Q        1   /* daeQVector rows                              */
#define XA_KIND_FLIM     2   /* dFdxdVpVector rows (voltage limiter)         */
#define XA_KIND_QLIM     3   /* dQdxdVpVector rows                           */
#define XA_KIND_DFDX     4   /* dFdx CSR value slots                         */
#define XA_KIND_DQDX     5   /* dQdx CSR value slots                         */
#define XA_NKINDS        6

typedef struct xa_context_s* xa_ctx;
typedef struct xa_group_s*   xa_group;

/* Solver-state flags read by the device loads: the subset of
 * N_DEV_SolverState.h:93-190 that the four models consult (filled in the
 * reference by DeviceMgr::setupSolverInfo_, N_DEV_DeviceMgr.C:3218-3379). */
typedef struct xa_solver_state {
  int32_t newtonIter;
  int32_t dcopFlag;
  int32_t initJctFlag;
  int32_t initFixFlag;
  int32_t inputOPFlag;
  int32_t locaEnabledFlag;
  int32_t initTranFlag;
  int32_t beginIntegrationFlag;
  int32_t tranopFlag;
  int32_t acopFlag;
  int32_t transientFlag;
  int32_t artParameterFlag;     /* MOSFET gain/nonlinearity homotopy        */
  int32_t bjtArtParameterFlag;  /* BJT homotopy                             */
  int32_t _pad;
  double  pdt;                  /* alpha/dt                                 */
  double  currTimeStep;
  double  lastTimeStep;
  double  currTime;
  double  gainScale;            /* gainScale[0] of the reference            */
  double  nltermScale;
  /* read by the independent sources (SourceData::initializeSource and
     updateSource, N_DEV_SourceData.C): breakpoint tolerance, the analysis'
     final time and first step (the PULSE/SIN/EXP defaults), and whether a
     harmonic-balance analysis is set up (which selects the transient
     waveform) */
  double  bpTol;
  double  finalTime;
  double  startingTimeStep;
  int32_t HBspecified;
  int32_t _pad2;
} xa_solver_state;

/* Device options read by the loads (N_DEV_DeviceOptions.C:66-140). */
typedef struct xa_device_options {
  double  gmin;                 /* default 1e-12                            */
  double  tnom;                 /* kelvin, default 300.15                   */
  double  temp;                 /* kelvin, default 300.15                   */
  double  defl, defw, defad, defas;  /* MOS defaults 1e-4,1e-4,0,0          */
  double  excessPhaseScalar1, excessPhaseScalar2;  /* 1.0, 1.0              */
  int32_t voltageLimiterFlag;   /* default 1                                */
  int32_t newMeyerFlag;         /* default 0                                */
  int32_t newExcessPhase;       /* default 0                                */
  int32_t lambertWFlag;         /* default 0                                */
} xa_device_options;

/* Circuit vectors for one load call (device pointers; NULL where unused).
 * Field names follow N_DEV_ExternData.h. */
typedef struct xa_vectors {
  const double* nextSol;
  const double* currSol;
  const double* flagSol;        /* optional: inputOP flag vector            */
  double*       nextSta;
  double*       currSta;        /* written on the first initTran iteration  */
  double*       lastSta;        /* written by the BJT excess phase when
                                   integration begins                       */
  const double* nextStaDeriv;   /* updateSecondaryState                     */
  double*       nextSto;
  double*       currSto;
  double*       f;
  double*       q;
  double*       b;
  double*       dFdxdVp;
  double*       dQdxdVp;
  double*       storeLeadF;
  double*       storeLeadQ;
  double*       dFdx;           /* CSR values (nnz)                         */
  double*       dQdx;           /* CSR values (nnz)                         */
} xa_vectors;

/* ---- context ------------------------------------------------------------ */
int         xa_ctx_create(int device_ordinal, xa_ctx* out);
int         xa_ctx_destroy(xa_ctx ctx);
int         xa_ctx_sync(xa_ctx ctx);
void*       xa_ctx_stream(xa_ctx ctx);          /* the hipStream_t used    */
/* Use a caller-owned stream (e.g. the framework's current stream). */
int         xa_ctx_set_stream(xa_ctx ctx, void* stream);
const char* xa_last_error(void);
int         xa_abi_version(void);
void        xa_default_options(xa_device_options* opt);
void        xa_default_solver_state(xa_solver_state* ss);

/* ---- device memory for callers with host vectors ------------------------ */
/* No reference counterpart: plumbing for a caller whose circuit vectors live
 * in host memory (a plain Xyce build keeps them in host Epetra vectors).
 * Device buffers on the context's device, and copies ordered on the
 * context's stream that return once the copy is complete (host memory may be
 * pageable). */
int xa_device_alloc(xa_ctx ctx, size_t bytes, void** out);
int xa_device_free(xa_ctx ctx, void* ptr);
int xa_copy_to_device(xa_ctx ctx, void* dst, const void* src, size_t bytes);
int xa_copy_to_host(xa_ctx ctx, void* dst, const void* src, size_t bytes);

/* ---- group setup (one group = all instances of one model type) ---------- */
/* Replaces Config<Traits>::registerDevice + Traits::factory. */
int xa_group_create(xa_ctx ctx, int model_type, const xa_device_options* opt,
                    xa_group* out);
int xa_group_destroy(xa_group g);

/* .MODEL card: name=value pairs; a pair given here is "given" in the
 * reference sense.  Unknown names -> XA_ERR_ARG.  Returns the model index. */
int xa_group_add_model(xa_group g, int level, int n_params,
                       const char* const* names, const double* values,
                       int* model_index);

/* Instances: n instances of model `model_index[i]`.  Instance parameters are
 * given as n_params columns; values[p][i] is parameter p of instance i. */
int xa_group_add_instances(xa_group g, int n, const int* model_index,
                           int n_params, const char* const* names,
                           const double* const* values);

/* Runs Model::processParams, Instance::processParams / updateTemperature for
 * every instance on the host and fixes the per-instance variable counts. */
int xa_group_setup(xa_group g);

/* Per-instance variable counts after setup (host arrays of length n):
 * external nodes, internal nodes, state vars, store vars (incl. lead
 * current slot when requested). */
int xa_group_counts(xa_group g, int* n_ext, int* n_int, int* n_sta, int* n_sto);
int xa_group_num_instances(xa_group g);

/* Lead currents requested per instance (host array of n flags, may be NULL
 * for none).  Must precede xa_group_setup. */
int xa_group_request_lead_current(xa_group g, const int32_t* flags);

/* Instance::registerLIDs/registerStateLIDs/registerStoreLIDs: flattened host
 * arrays, instance-major, lengths sum(n_ext+n_int), sum(n_sta), sum(n_sto).
 * Ground is -1. */
int xa_group_register_lids(xa_group g, const int32_t* node_lids,
                           const int32_t* sta_lids, const int32_t* sto_lids);

/* jacobianStamp(): the (row LID, col LID) pairs the group stamps into dFdx
 * and dQdx, after node collapse, ground rows/cols dropped, duplicates kept.
 * Query with rows==NULL to get the count. */
int xa_group_jac_pairs(xa_group g, int64_t* count, int32_t* rows, int32_t* cols);

/* registerJacLIDs: the CSR structure of dFdx/dQdx (host arrays); the library
 * resolves every stamp to a value slot (Indexor::matrixGlobalToLocal). */
int xa_group_register_csr(xa_group g, int n_rows, const int32_t* row_ptr,
                          const int32_t* col_idx);

/* Builds the fixed-order scatter plans and uploads everything. */
int xa_group_finalize(xa_group g);

/* ---- loads (Master virtuals) ------------------------------------------- */
/* Master::updateState(solVec, staVec, stoVec) */
int xa_update_state(xa_group g, const xa_solver_state* ss, const xa_vectors* v);
/* Master::updateSecondaryState(staDerivVec, stoVec) */
int xa_update_secondary_state(xa_group g, const xa_solver_state* ss,
                              const xa_vectors* v);
/* Master::loadDAEVectors(solVec, fVec, qVec, bVec, storeLeadF, storeLeadQ) */
int xa_load_vectors(xa_group g, const xa_solver_state* ss, const xa_vectors* v);
/* Master::loadDAEMatrices(dFdx, dQdx) */
int xa_load_matrices(xa_group g, const xa_solver_state* ss, const xa_vectors* v);
/* updateState + updateSecondaryState + loadDAEVectors + loadDAEMatrices in
 * one pass (no reference counterpart as one call; same results as the four
 * calls in sequence). */
int xa_load_all(xa_group g, const xa_solver_state* ss, const xa_vectors* v);

/* xa_load_all without any device allocation: no reference counterpart.  For
 * callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  It runs exactly the
 * route xa_load_all takes, with the same results, but the call's scratch --
 * every output kind's stamp value buffer (entries x instances doubles) and
 * the supply-rail partial sums -- lives in one static device array of
 * 4 Mi doubles (32 MiB) per GPU, held from library load to process end,
 * instead of the group's buffers.  XA_ERR_ARG when the group needs more than
 * that (the message gives the size).  The array is shared by every call on
 * one GPU and keeps the last call's contents; it is cleared by a fill kernel
 * at the start of each call.  Calls that overlap on one GPU (different
 * streams or threads) race on it.  Calls no hipMalloc/hipFree and creates no
 * timing events (group timing is not recorded for this call). */
int xa_load_all_noalloc(xa_group g, const xa_solver_state* ss, const xa_vectors* v);

/* Autotuner of the block shapes (no reference counterpart).  Times, on the
 * caller's vectors and flags, the four-call path with update-kernel blocks of
 * 64, 128 and 256 threads, and the fused one-kernel load with blocks of 64,
 * 128 and 256 instances where its block plans exist and its LDS fits, and
 * the pattern-fused load likewise where every kind has a pattern plan; then
 * keeps the fastest: the update-kernel block for the four calls, and the path
 * (and fused block) for xa_load_all.  Every choice gives the same bits.  It
 * runs up to 9 * (reps + 1) load passes that add into the output vectors (zero them
 * afterwards) and rewrites store/state as a load does.  *route: 1 four calls,
 * 2 fused, 3 pattern-fused (see xa_group_set_path); *block: its block size.
 * Synchronizes. */
int xa_group_autotune(xa_group g, const xa_solver_state* ss, const xa_vectors* v, int reps,
                      int* route, int* block);

/* Pins xa_load_all's path (no reference counterpart; for benchmarks and
 * tests; the autotuner picks the fastest instead): 1 the four calls with
 * update-kernel blocks of `block` threads, 2 the fused load with blocks of
 * `block` instances, 3 the pattern-fused load (every kind a pattern plan: one
 * kernel that stages each block's output spans in LDS and adds the stamps
 * into them in order) with blocks of `block` instances.  block: 64, 128 or
 * 256.  XA_ERR_UNSUPPORTED when the group has no such path.  All paths give
 * the same bits. */
int xa_group_set_path(xa_group g, int path, int block);

/* Persistent form of the stored-mode update (no reference counterpart; for
 * batched runs, where one group holds every circuit of a model type): the
 * update kernel is launched with `blocks` workgroups that stay resident and
 * loop over the group's instance tiles (tile t by workgroup t mod blocks,
 * which deals consecutive tiles round-robin over the XCDs), instead of one
 * workgroup per tile.  blocks = 0 returns to one workgroup per tile;
 * blocks < 0 takes the number that fits the GPU at the kernel's occupancy.
 * Same bits either way.  XA_ERR_UNSUPPORTED for models without a persistent
 * update (BSIM4 has one). */
int xa_group_set_persistent(xa_group g, int blocks);

/* Shape of the scatter plan of one output kind (XA_KIND_*), for diagnostics
 * and tests (no reference counterpart): targets, contributions, and the
 * targets per instance of a pattern plan (0 when the kind uses the general
 * segmented plan). */
int xa_group_plan_info(xa_group g, int kind, int64_t* n_targets, int64_t* n_sources,
                       int* pattern_slots);

/* Layout of the per-instance parameter data the update kernel reads (no
 * reference counterpart; diagnostics and tests): the columns read per
 * instance from HBM, and the columns served from the model cards because
 * every instance of each card holds the same value (BSIM4: size-dependent
 * parameters of binless cards and shared instance values). */
int xa_group_param_layout(xa_group g, int* per_instance_columns, int* card_columns);

/* Device-side status of the last update: number of instances whose
 * evaluation returned false (0 = all succeeded). Synchronizes. */
int xa_group_failures(xa_group g, int64_t* n_failed);

/* Per-instance intermediate read-back for tests (host array of n doubles).
 * `name` is a model-specific intermediate ("Id", "Gd", "Vd", ...). */
int xa_group_read_intermediate(xa_group g, const char* name, double* out);

/* Algorithmic HBM bytes of one load pass of this group (inputs read once,
 * outputs written once), for roofline reporting. */
int xa_group_traffic_bytes(xa_group g, double* bytes_update_state,
                           double* bytes_load_vectors, double* bytes_load_matrices);

/* Kernel timing on the context stream with HIP events.  Enabling (re)starts
 * recording; last_times returns the mean milliseconds per call of each phase
 * (update_state, load_vectors, load_matrices; load_all counts as update)
 * over the calls recorded since, 0 when none.  Synchronizes. */
int xa_group_enable_timing(xa_group g, int on);
int xa_group_last_times(xa_group g, float* ms_update, float* ms_vectors,
                        float* ms_matrices);

/* ---- batched Newton update of independent circuits ---------------------- */
/* No reference counterpart as one call: it forms and solves, per circuit,
 * the backward-Euler Newton system that N_TIA_BackwardDifferentiation15.C:
 * 215-265 (obtainResidual) and the nonlinear solver form for one circuit, for
 * a batch of circuits (Monte Carlo / .STEP copies) loaded together.
 *
 * Circuit c owns the unknowns [c_start[c], c_start[c+1]) (at most
 * max_unknowns); no CSR entry may couple two circuits.  pos[g] is unknown g's
 * position in its circuit's elimination order (NULL: natural order); in that
 * order every assembled entry (i, j) must lie within the band: i - j <= kl
 * below and j - i <= ku above the diagonal.  The band, with
 * kl rows of pivoting fill, is factored in LDS: (2kl+ku+2)*max_unknowns
 * doubles must fit in 159 KB.  One wavefront per circuit.
 * With dc = 0:  q_now = q + c_lin .* x;
 *   rhs = -((q_now - q_prev)/h + f) + dFdxdVp + dQdxdVp/h,
 *   J = dFdx + dQdx/h + diag(c_lin/h);
 * with dc = 1:  rhs = -f + dFdxdVp, J = dFdx.
 * Rows and columns of held nodes (held[i] != 0) are left out.  x += J^-1 rhs
 * on the free nodes (banded LU with partial pivoting); dx_max[c] = the
 * circuit's largest |update| (NaN for a circuit outside the limits, with an
 * entry outside the band, or with a zero pivot -- a singular Jacobian, such
 * as a floating node -- in which case its x is left unchanged; a NaN update
 * also gives NaN, never a finite maximum).  All pointers are device pointers; f, q, the
 * limiter vectors and the CSR values are those of a load of the batch. */
int xa_batch_newton_step(xa_ctx ctx, int n_circ, int max_unknowns, int kl, int ku,
                         const int32_t* c_start, const int32_t* pos,
                         const int32_t* row_ptr, const int32_t* col_idx, const double* dFdx,
                         const double* dQdx, const double* f, const double* q,
                         const double* dFdxdVp, const double* dQdxdVp, const double* q_prev,
                         const double* c_lin, const uint8_t* held, double h, int dc,
                         double* x, double* q_now, double* dx_max);

/* The same step with the integrator named: XA_STEP_DCOP (dc = 1 above),
 * XA_STEP_BE (dc = 0 above) or XA_STEP_TRAP, the trapezoid rule of the
 * reference's one-step integrator at order 2 (N_TIA_OneStep.C:272-335
 * obtainResidual, :463-500 obtainJacobian; f_prev = the previous step's F,
 * the integrator's qHistory[2], N_TIA_OneStep.C:1399-1406):
 *   rhs = -((q_now - q_prev)/h + f/2 + f_prev/2) + dFdxdVp/2 + dQdxdVp/h,
 *   J = dFdx/2 + dQdx/h + diag(c_lin/h).
 * f_prev may be NULL except for XA_STEP_TRAP. */
#define XA_STEP_DCOP 0
#define XA_STEP_BE 1
#define XA_STEP_TRAP 2
int xa_batch_newton_step_ex(xa_ctx ctx, int n_circ, int max_unknowns, int kl, int ku,
                            const int32_t* c_start, const int32_t* pos,
                            const int32_t* row_ptr, const int32_t* col_idx, const double* dFdx,
                            const double* dQdx, const double* f, const double* q,
                            const double* dFdxdVp, const double* dQdxdVp, const double* q_prev,
                            const double* f_prev, const double* c_lin, const uint8_t* held,
                            double h, int method, double* x, double* q_now, double* dx_max);

/* ---- batched dense LU in FP32 on the matrix cores ------------------------ */
/* No reference counterpart as code: LAPACK's blocked dgetrf / dgetrs order
 * (partial pivoting, first largest magnitude) on `batch` matrices at once,
 * the trailing updates on v_mfma_f32_16x16x4f32; the FP32 half of the
 * mixed-precision solves (FP64 refinement around it, as in
 * xa_hb_precond_solve).  Each matrix is ld x ld, row-major, the batch
 * contiguous (matrix b at a + b*ld*ld); ld a multiple of 64, at most 3072;
 * rows and columns [m, ld) must hold the identity.  ipiv: batch x ld, ipiv[i]
 * = the 0-based row interchanged with row i at step i (rows < m rounded up to
 * 64).  info[b] = 0, or the first zero pivot's column + 1 (the factor is then
 * singular and a solve divides by zero).  solve: b holds batch x nrhs
 * right-hand sides of ld floats each (rhs r of matrix q at b + (q*nrhs+r)*ld),
 * overwritten by the solutions.  Device pointers; stream-ordered. */
int xa_lu_f32_factor(xa_ctx ctx, int batch, int m, int ld, float* a, int32_t* ipiv,
                     int32_t* info);
int xa_lu_f32_solve(xa_ctx ctx, int batch, int nrhs, int m, int ld, const float* lu,
                    const int32_t* ipiv, float* b);

/* ---- harmonic balance ----------------------------------------------------- */
/* The HB loader's transforms and batched-sample load (N_LOA_HBLoader.C:
 * 540-775) and its block-Jacobi preconditioner (N_LAS_HBBlockJacobiPrecond.C:
 * 178-425), on the device.  N = n_samples = 2H + 1 (odd, N_ANP_HB.C:1470),
 * n = unknowns of one sample, T = period.
 *   time domain:  xt[s*n + j], sample s (time s*T/N) of unknown j;
 *   frequency:    xf[j*2N + 2p], xf[j*2N + 2p + 1] = (Re, Im) of unknown j at
 *                 position p: harmonic p for p <= H, p - N above (the
 *                 permuted block layout of N_LAS_BlockSystemHelpers.C:512-609).
 * xa_hb_ift:  xt[s][j] = Re_0 + 2 sum_{k=1..H} (Re_k cos(2 pi k s/N) - Im_k sin(..));
 * xa_hb_dft:  (Re, Im)_k = (1/N) sum_s x_s (cos, -sin)(2 pi k s/N), harmonic -k
 *             the conjugate; (computePermutedIFT / computePermutedDFT with
 *             the FFTW packing and scalings, N_UTL_FFTW_Interface.C:66-130);
 * xa_hb_add_jwq: ff += j w k qf for k = -H..H, w = 2 pi / T
 *             (N_LOA_HBLoader.C:733-775).
 * xa_hb_load: one HB load: xt = ift(xf) into vt->nextSol (N*n entries), the
 *             time-domain outputs of vt zeroed (f, q, dFdxdVp, dQdxdVp: N*n;
 *             dFdx, dQdx: nnz_t), xa_load_all(g) over every sample at once --
 *             g is the group of all samples' instances, sample s's unknown j
 *             at LID s*n + j -- then ff = dft(f) + j w k dft(q), qf = dft(q),
 *             dfdxdvp_f = dft(dFdxdVp) + j w k dft(dQdxdVp), dqdxdvp_f =
 *             dft(dQdxdVp) (NULL outputs are skipped).  The per-sample
 *             Jacobians stay in vt->dFdx / vt->dQdx for the preconditioner.
 * Preconditioner: on a CSR pattern (row_ptr[n+1], col_idx[nnz], host arrays)
 * shared by every sample, whose sample s values are dFdx_t[s*nnz + e]:
 *   factor: G = (sum_s dFdx_s) (1/N), C likewise; for k = 0..H the real-
 *           equivalent block [G, -k w C; k w C, G] (2n rows, padded to a
 *           multiple of 64, at most 3072) in FP32, factored by
 *           xa_lu_f32_factor; info[k] per block (host array of H+1 or NULL).
 *           Harmonic -k uses harmonic k's factor: K_-k = P K_k P,
 *           P = diag(I, -I).
 *   solve:  x = K_k^-1 b for every position of bf (frequency layout) into xf:
 *           the FP32 solve, then up to max_refine FP64 iterative-refinement
 *           steps r = b - K_k x (FP64, from G and C), x += K_k^-1 r (FP32),
 *           until every system has max|r| <= tol max|b| or a residual that
 *           no longer halves in a step.  Per position p
 *           (host arrays of N, or NULL): rel_fp32[p] = max|r|/max|b| after the
 *           FP32 solve, rel_final[p] after refinement, iterations[p] = the
 *           refinement steps it needed. */
typedef struct xa_hb_plan_s* xa_hb_plan;
typedef struct xa_hb_precond_s* xa_hb_precond;
int xa_hb_plan_create(xa_ctx ctx, int n_samples, int n_unknowns, double period,
                      xa_hb_plan* out);
int xa_hb_plan_destroy(xa_hb_plan p);
int xa_hb_ift(xa_hb_plan p, const double* xf, double* xt);
int xa_hb_dft(xa_hb_plan p, const double* xt, double* xf);
int xa_hb_add_jwq(xa_hb_plan p, double* ff, const double* qf);
int xa_hb_load(xa_hb_plan p, xa_group g, const xa_solver_state* ss, const double* xf,
               const xa_vectors* vt, int64_t nnz_t, double* ff, double* qf, double* dfdxdvp_f,
               double* dqdxdvp_f);
/* y = A x for a CSR matrix of n_rows rows (device arrays, double values):
 * the Jacobian-vector products of a Newton-Krylov solve -- for HB, the
 * time-domain dF/dx and dQ/dx applied to a direction between the transforms
 * (the matrix-free product of N_LOA_HBLoader.C:350-456).  Each row's sum
 * is formed in a fixed order (lane-strided partial sums, then a fixed tree),
 * the same on every call. */
int xa_csr_spmv(xa_ctx ctx, int n_rows, const int32_t* row_ptr, const int32_t* col_idx,
                const double* values, const double* x, double* y);
/* the same for a matrix of few long rows (one 1024-thread block per row,
 * a fixed summation tree): a supply's KCL row split off a Jacobian */
int xa_csr_spmv_long(xa_ctx ctx, int n_rows, const int32_t* row_ptr, const int32_t* col_idx,
                     const double* values, const double* x, double* y);
int xa_hb_precond_create(xa_hb_plan p, int nnz, const int32_t* row_ptr, const int32_t* col_idx,
                         xa_hb_precond* out);
int xa_hb_precond_destroy(xa_hb_precond pc);
int xa_hb_precond_factor(xa_hb_precond pc, const double* dFdx_t, const double* dQdx_t,
                         int32_t* info);
int xa_hb_precond_solve(xa_hb_precond pc, const double* bf, double* xf, int max_refine,
                        double tol, double* rel_fp32, double* rel_final, int32_t* iterations);

/* ---- sparse Newton systems: ILU(0) preconditioner ------------------------ */
/* Incomplete LU with the matrix's own pattern, factored and applied on the
 * device (rocSPARSE csrilu0 and triangular solves): the preconditioner of
 * the device GMRES that solves a large circuit's transient Newton systems,
 * where the reference uses KLU or AztecOO GMRES with an Ifpack ILU
 * (N_LAS_AztecOOSolver.C, N_LAS_IfpackPrecond.C).  create: the CSR pattern
 * (host arrays; columns sorted, every diagonal present); factor: the
 * values (device, nnz), *zero_pivot = the first zero pivot's row or -1;
 * solve: x = U^-1 L^-1 b (device vectors). */
typedef struct xa_ilu0_s* xa_ilu0;
int xa_ilu0_create(xa_ctx ctx, int n, int nnz, const int32_t* row_ptr, const int32_t* col_idx,
                   xa_ilu0* out);
int xa_ilu0_destroy(xa_ilu0 h);
int xa_ilu0_factor(xa_ilu0 h, const double* values, int* zero_pivot);
int xa_ilu0_solve(xa_ilu0 h, const double* b, double* x);
/* mode 1 (default): the solve runs level-scheduled kernels (rows grouped by
 * the dependency depth of the pattern, one kernel per level); mode 2: the
 * same kernels captured once as a HIP graph and replayed; mode 0: rocSPARSE
 * csrsv; -1 keeps the mode.  Returns the level counts of the L and U
 * solves. */
int xa_ilu0_config(xa_ilu0 h, int mode, int* l_levels, int* u_levels);

/* ---- netlist reader and circuit builder --------------------------------- */
/* The SPICE-netlist subset a device-load circuit needs, read into the
 * groups above.  Replaces, for this subset, the reference's parser and
 * device-instantiation path: IO::CircuitBlock reading the netlist
 * (N_IO_CircuitBlock.C:483-511, numbers by N_UTL_Misc.C:197-236),
 * DeviceMgr::addDeviceInstance creating one device type per model type as
 * it is met (N_DEV_DeviceMgr.C:1446-1470, 1206-1229), Topology's LID
 * assignment and the Jacobian graph (N_TOP_Topology.C, N_LAS_Builder.C).
 *
 * Read: a title line; `*` comment lines, `;` end-of-line comments, `+`
 * continuations; M d g s b model (NMOS/PMOS LEVEL=1 -> MOSFET1, 14/54 ->
 * BSIM4), D a c model [area], Q c b e [s] model [area] (NPN/PNP), R and C
 * n1 n2 [model] [value], V n+ n- [DC] v | SIN(...) | EXP(...) | PULSE(...),
 * X name nodes... subckt with .SUBCKT name ports ... .ENDS (flattened: an
 * element inside X1 is named X1:<name>, its own nodes X1:<node>; no
 * subcircuit parameters),
 * name=value instance parameters, .MODEL, .TRAN tstep tstop [tstart
 * [tmax]] [UIC], .IC V(n)=v or n=v, .STEP [LIN] target:param start stop step |
 * target:param LIST v..., .END; .OPTIONS/.PRINT/.MEASURE lines are
 * skipped.  Anything else -> XA_ERR_ARG naming the line.
 *
 * A part is every element of one device type with one set of instance
 * parameter names (one group when built), in order of first appearance.
 * Node ids are 0.. in order of first appearance, ground -1.  `step` selects
 * a .STEP point (-1: the netlist's own values). */
typedef struct xa_netlist_s* xa_netlist;
typedef struct xa_circuit_s* xa_circuit;
int xa_netlist_parse(const char* text, xa_netlist* out);
int xa_netlist_destroy(xa_netlist h);
int xa_netlist_sizes(xa_netlist h, int* n_nodes, int* n_parts, int* n_steps, int* n_ic);
int xa_netlist_node_name(xa_netlist h, int node, char* buf, int len);
/* tstep, tstop, tstart, tmax; XA_ERR_STATE without .TRAN */
int xa_netlist_tran(xa_netlist h, double* out4);
int xa_netlist_ic(xa_netlist h, int32_t* nodes, double* values);
int xa_netlist_part(xa_netlist h, int part, int* model_type, int* n_models, int* n_inst,
                    int* n_ext, int* n_iparams);
/* model k of a part (the part's models in .MODEL order): level and
 * name=value pairs (names NULL: count only); the names stay valid until the
 * next call on this part */
int xa_netlist_part_model(xa_netlist h, int part, int k, int step, int* level, int* n_params,
                          const char** names, double* values);
/* the part's instances: model index in the part, external node ids
 * (instance-major), parameter names and values[p * n_inst + i] */
int xa_netlist_part_instances(xa_netlist h, int part, int step, int32_t* model_index,
                              int32_t* ext_nodes, const char** names, double* values);

/* Builds every part as a group on ctx, with LIDs, state/store slots and the
 * circuit's CSR, finalized.  step >= 0 or -1 as above; step -2 builds every
 * .STEP point's copy of the circuit into the same groups (the batched
 * form): copy c's unknowns are c*U .. c*U+U-1 (U = netlist nodes, then each
 * part's internal nodes instance by instance), its state and store slots
 * likewise contiguous.  The groups belong to the circuit. */
int xa_circuit_build(xa_netlist h, xa_ctx ctx, const xa_device_options* opt, int step,
                     xa_circuit* out);
int xa_circuit_destroy(xa_circuit h);
int xa_circuit_sizes(xa_circuit h, int* n_unknowns, int* n_sta, int* n_sto, int64_t* nnz,
                     int* n_groups, int* n_copies);
int xa_circuit_csr(xa_circuit h, int32_t* row_ptr, int32_t* col_idx);
int xa_circuit_group(xa_circuit h, int k, xa_group* g, int* model_type);
/* group k's node LIDs, instance-major (each instance's external, then
 * internal nodes; ground -1); node_lids NULL: *count only */
int xa_circuit_group_lids(xa_circuit h, int k, int64_t* count, int32_t* node_lids);
/* DeviceMgr's load over every device type (N_DEV_DeviceMgr.C:2053-2502):
 * the groups in creation order onto the same vectors. */
int xa_circuit_load(xa_circuit h, const xa_solver_state* ss, const xa_vectors* v);

#ifdef __cplusplus
}
#endif
#endif /* XYCE_AMD_H */
