/*
 * dqh -- MI355X-native (gfx950) kernels for the Lindblad master-equation solver
 * behind a dynamiqs-compatible ``mesolve``.
 *
 * C ABI: plain pointers, sizes and POD descriptors. Every pointer inside a
 * descriptor is a DEVICE pointer owned by the caller; the library never allocates
 * or frees caller memory. Entry points return 0 on success or a negative
 * ``dqh_status``; ``dqh_last_error()`` returns a thread-local message.
 *
 * Reference interface replaced (dynamiqs 0.3.5):
 *   The integrator seam ``_mesolve`` -> ``integrator_constructor(...).run()``
 *   (dynamiqs/integrators/apis/mesolve.py:350-384,
 *    dynamiqs/integrators/core/abstract_integrator.py:16-54). The library runs
 *   what ``MESolveDiffraxIntegrator.run`` (diffrax_integrator.py:135-147) and
 *   ``MESolveFixedRouchon1Integrator`` (rouchon_integrator.py:385-432) run, for a
 *   batch of independent instances per call. The same seams of ``_sesolve``
 *   (apis/sesolve.py:263-304) and of the propagators' Expm integrator
 *   (expm_integrator.py:60-90) are served by ``dqh_sesolve_tsit5`` and
 *   ``dqh_expm_propagators`` / ``dqh_cgemm_batched``.
 */
#ifndef DQH_H
#define DQH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum dqh_status {
  DQH_OK = 0,
  DQH_BAD_ARG = -1,
  DQH_BAD_SHAPE = -2,
  DQH_MAX_STEPS = -3,  /* per-instance status: max_steps reached before t1 */
  DQH_NONFINITE = -4,  /* per-instance status: NaN/Inf or dt underflow */
  DQH_HIP_ERROR = -5,
  DQH_UNSUPPORTED = -6,
};

/* Per-instance descriptor of one dense master-equation solve (all fields 8 bytes,
 * no padding). Matrices are complex64 (float re, float im interleaved), row-major,
 * zero-padded to np x np (np = 16 * ceil(n / 16)). */
typedef struct dqh_me_inst {
  int64_t n;          /* Hilbert-space dimension */
  int64_t np;         /* padded dimension, multiple of 16, <= 1024 */
  int64_t nL;         /* jump-operator terms in the RHS (D/C pairs) */
  int64_t nE;         /* expectation operators */
  int64_t ntsave;     /* save times */
  int64_t npwc;       /* piecewise-constant Hamiltonian terms (Tsit5) */
  int64_t nseg;       /* PWC breakpoints (length of pwc_times) */
  int64_t flags;      /* bit0: save states; bit1: generator has an Hnh term */
  double rtol, atol, safety, factormin, factormax;
  int64_t max_steps;  /* Tsit5: max steps; Rouchon1: number of fixed steps */
  double t0;          /* start time (<= tsave[0]) */
  double t1;          /* end time (= tsave[ntsave-1]) */
  /* operators */
  uint64_t tsave;     /* const double[ntsave] */
  uint64_t Hnh;       /* const c64[np*np]   -iH0 - 1/2 sum L^dag L                 */
  uint64_t Dops;      /* const c64[nsets*nL*np*np]  left factors  (L_k)            */
  uint64_t Cops;      /* const c64[nsets*nL*np*np]  right factors (1/2 L_k^dag).
                         Tsit5 descriptors must hold Cops = Dops^dag / 2 exactly:
                         the resident body uses X_k C_k = D_k (D_k rho / 2)^dag
                         for the Hermitian stage inputs and reads Dops only.   */
  uint64_t EopsT;     /* const c64[nE*np*np] transposed exp ops                    */
  uint64_t Hpwc;      /* const c64[npwc*np*np]  -i O_j of each PWC term            */
  uint64_t pwc_times; /* const double[nseg]                                        */
  uint64_t pwc_vals;  /* const double[npwc*(nseg-1)*2] complex coefficients        */
  uint64_t step_t;    /* Rouchon1: const double[max_steps+1] step boundaries       */
  uint64_t step_set;  /* Rouchon1: const int32[max_steps] Kraus-set index per step */
  /* state, scratch and outputs */
  uint64_t y;         /* double2[2*np*np]: FP64 state (buffer 0 holds rho0)        */
  uint64_t work;      /* c64 scratch, dqh_me_work_elems(np, nL, npwc) elements     */
  uint64_t states;    /* c64[ntsave*n*n] (or [n*n] when states are not saved)      */
  uint64_t expects;   /* c64[nE*ntsave]                                             */
  uint64_t final_state; /* c64[n*n]                                                 */
  uint64_t stats;     /* int64[8]: nsteps, naccepted, nrejected, status, nrhs       */
  /* optional (0: off): end time of every accepted Tsit5 step, the first
   * step_log_cap of them (replaying the same steps in another precision) */
  uint64_t step_log;  /* double[step_log_cap]                                        */
  int64_t step_log_cap;
  /* optional smooth time dependence (Tsit5 kernels; 0 terms: off): H(t) gains
   * sum_j f_j(t) O_j with f_j a piecewise cubic on the uniform grid
   * mod_t0 + i mod_dt (i < mod_nint): on [t_i, t_i + dt] f_j = ((c0 x + c1) x + c2) x + c3
   * with x = t - t_i. Evaluated at every stage time (modulated operators,
   * time_qarray.py:669-760). At most DQH_MAX_MOD terms. */
  int64_t nmod;
  uint64_t Hmod;      /* const c64[nmod*np*np]  -i O_j of each term                  */
  uint64_t mod_coef;  /* const double2[nmod*mod_nint*4]: c0..c3 per interval        */
  double mod_t0, mod_dt;
  int64_t mod_nint;
  /* optional (0: off), resident Tsit5 body only (np <= 64, nL <= 3, no
   * DQH_FLAG_GLOBAL / DQH_FLAG_LMOD): the diagonal of the constant generator part
   * in FP64, with Hnh holding its off-diagonal part. Every vector-field evaluation
   * then adds (d_i + conj(d_j)) rho_s[i][j] in FP64 to the FP32 products, so the
   * largest single terms (the diagonal energies and decay rates) carry no FP32
   * rounding. Only dqh_mesolve_tsit5_dense_uniform_fp64diag takes such
   * descriptors: dqh_mesolve_tsit5_dense and dqh_mesolve_tsit5_dense_uniform set
   * their status to DQH_UNSUPPORTED, dqh_mesolve_tsit5_dense_grid returns
   * DQH_BAD_ARG. */
  uint64_t hdiag;     /* const double2[np] */
} dqh_me_inst;

#define DQH_MAX_MOD 4

#define DQH_FLAG_SAVE_STATES 1
#define DQH_FLAG_HAS_HNH 2
#define DQH_FLAG_GLOBAL 4 /* run the global-memory kernel form even where the LDS-resident one applies */
#define DQH_FLAG_DOPRI5 8 /* Tsit5 kernels: integrate with the Dopri5 pair (same controller and saving) */
#define DQH_FLAG_EULER 16 /* dqh_sesolve_tsit5: fixed Euler steps between step_t[0..max_steps] */
/* Tsit5 kernels: jump operator k is f_k(t) L_k with f_k the table nmod + k of mod_coef
 * (a constant 1 for the constant ones); Dops / Cops hold L_k and L_k^dag / 2, and the
 * kernels rebuild f_k L_k and conj(f_k) L_k^dag / 2 at every stage time in work at
 * (10 + nL) np^2 (work needs (10 + 3 nL) np^2 elements). The time-dependent part of
 * -1/2 sum L^dag L is among the nmod Hamiltonian terms. Such instances always run on
 * the global-memory kernel form (the uniform entry point reports DQH_UNSUPPORTED). */
#define DQH_FLAG_LMOD 32
/* Tsit5 kernels (dqh_mesolve_tsit5_dense, _uniform): when an instance is done, its
 * workgroup makes its outputs visible at system scope and then stores 1 to stats[7]
 * (release, system scope). A host that places states / expects / final_state /
 * stats in pinned host memory (device address from dqh_host_device_pointer) can
 * poll stats[7] and take each instance's results while the launch still runs. */
#define DQH_FLAG_SIGNAL 64
/* dqh_mesolve_tsit5_dense: integrate the instance with the Dormand-Prince 8(7) pair
 * (13 stages, error order 8) on the global-memory form, steps landing on the save
 * times; work needs (16 + nL) np^2 elements; not with DQH_FLAG_LMOD, not on the
 * uniform or grid-wide entry points. */
#define DQH_FLAG_DOPRI8 128
/* with DQH_FLAG_LMOD: jump operator k is additionally scaled by the piecewise-constant
 * coefficient in row npwc + k of pwc_vals (pwc_vals then holds npwc + nL rows). The
 * value is taken on the PWC segment of the step start, and it is zero outside
 * [pwc_times[0], pwc_times[nseg-1]). The -1/2 |c_k|^2 L_k^dag L_k part of the
 * generator is among the npwc Hamiltonian terms. This is the reference's
 * pwc(times, values, L) as a jump operator (time_qarray.py:572-666). */
#define DQH_FLAG_LPWC 256
/* with DQH_FLAG_LMOD: the right factors take coefficients of their own, C_k <- conj(g_k) C_k
 * with g_k from table nmod + nL + k (and, with DQH_FLAG_LPWC, row npwc + nL + k of
 * pwc_vals), while D_k keeps f_k. A jump operator that is a sum of terms
 * sum_j c_j(t) L_j expands into pairs (c_j L_j, conj(c_j) L_j^dag / 2) and cross pairs
 * (c_j L_j, conj(c_l) L_l^dag) for j < l, since the vector field adds tmp + tmp^dag
 * (time_qarray.py:836-956 as a jump operator). The -1/2 L^dag L products are generator
 * terms of the descriptor. */
#define DQH_FLAG_LPAIR 512

/* Size checks for FFI mirrors. */
int64_t dqh_sizeof_me_inst(void);
/* Number of complex64 scratch elements one instance needs in ``work``. */
int64_t dqh_me_work_elems(int64_t np, int64_t nL, int64_t npwc);

/* Lindblad RHS in the Hermitian form (diffrax_integrator.py:319-323):
 * final_state <- tmp + tmp^dag, tmp = Hnh rho + sum_k D_k rho C_k, rho = y buffer 0
 * rounded to complex64. One workgroup per instance. */
int dqh_lindblad_rhs_dense(const dqh_me_inst *d_insts, int64_t ninst, void *stream);

/* Device-side operator preparation for a batch packed by a server: each
 * descriptor's Hnh holds -iH and Dops the jump operators on entry; on return
 * Hnh = -iH - 1/2 sum_k D_k^dag D_k (products in FP64) and Cops = D_k^dag / 2.
 * Replaces the host-side construction of the non-Hermitian Hamiltonian of
 * dynamiqs/integrators/core/diffrax_integrator.py:319-323 for served requests.
 * ninst <= 65535; stream-ordered, allocates nothing. */
int dqh_dense_prepare_ops(const dqh_me_inst *d_insts, int64_t ninst, void *stream);

/* Adaptive Tsit5 solve of every instance, one persistent workgroup per instance
 * (replaces dx.diffeqsolve(..., dx.Tsit5(), dx.PIDController(...)) at
 * diffrax_integrator.py:101-147 with the vector field of :319-323). */
int dqh_mesolve_tsit5_dense(const dqh_me_inst *d_insts, int64_t ninst, void *stream);

/* Work queue of the serving loop: requests published by the host into a ring while a
 * persistent grid (dqh_mesolve_tsit5_dense_queue) solves them. A ring slot is
 * DQH_QUEUE_SLOT_WORDS uint64: {src, dst, bytes, src_desc, dst_desc, 0, 0, 0}: a
 * request's inputs staged in pinned host memory (device address src, ``bytes`` a
 * multiple of 16) and where they go in device memory (dst), and its dqh_me_inst in
 * pinned host memory (src_desc) and in device memory (dst_desc). The descriptor's
 * pointers name the device copies; as for dqh_dense_prepare_ops, Hnh holds -iH and
 * Dops the jump operators on entry. */
#define DQH_QUEUE_SLOT_WORDS 8
typedef struct dqh_queue {
  uint64_t ring;        /* device address of HOST pinned coherent uint64[capacity * DQH_QUEUE_SLOT_WORDS] */
  int64_t capacity;     /* ring slots, a power of two */
  uint64_t tail;        /* device address of a HOST pinned coherent int64: requests published */
  uint64_t head;        /* device int64[2]: requests claimed and the wall clock of the latest claim (zero
                           before the first launch, kept across launches) */
  uint64_t stop;        /* device address of a HOST pinned coherent int64: nonzero: exit once nothing is published */
  uint64_t busy_ticks;  /* device int64: wall-clock ticks the workgroups spent on requests */
  int64_t idle_ticks;   /* wall-clock ticks without a published request after which a workgroup exits */
  int64_t life_ticks;   /* wall-clock ticks after its start after which a workgroup claims no more requests */
} dqh_queue;

/* Persistent work-queue form of dqh_mesolve_tsit5_dense for continuous batching
 * (replaces the lockstep vmap of a batch, apis/mesolve.py:307-336 and
 * integrators/_utils.py:144-187, by requests admitted one by one). Each of the
 * ``nworkgroups`` workgroups claims the next published slot (claimed % capacity),
 * copies the request's inputs and descriptor to the device, completes Hnh and Cops,
 * solves it with the bodies of dqh_mesolve_tsit5_dense and signals it
 * (DQH_FLAG_SIGNAL), then claims the next. The host fills slot tail % capacity and
 * then increments tail; a slot may be rewritten once its request has completed.
 * Every workgroup exits when stop is set and nothing is published, when no workgroup
 * has claimed a request for idle_ticks, or at its first claim after life_ticks (the launch then ends
 * and other work queued behind it on the same hardware queue runs); requests
 * published after that need another launch (head carries over). */
int dqh_mesolve_tsit5_dense_queue(const dqh_queue *q, int64_t nworkgroups, void *stream);

/* hipMemcpyAsync host -> device (``src`` pinned) on ``stream``: the staging copy of
 * the serving loop's grid-wide launches, from memory of dqh_host_alloc. No
 * counterpart in the reference (serving plumbing). */
int dqh_memcpy_h2d_async(void *dst, const void *src, int64_t bytes, void *stream);

/* Rate of the device wall clock that idle_ticks and busy_ticks count, in kHz. */
int64_t dqh_wall_clock_khz(void);

/* Pinned, device-mapped host memory (hipHostMalloc with hipHostMallocMapped and
 * hipHostMallocCoherent or hipHostMallocNonCoherent). Coherent (fine-grained) memory
 * is for data the host writes while a kernel reads it (the device never caches it);
 * non-coherent memory for kernel outputs, made visible by a system-scope release.
 * No counterpart in the reference (serving plumbing). */
int dqh_host_alloc(int64_t bytes, int64_t coherent, void **host);
int dqh_host_free(void *host);

/* Fixed-step Rouchon1 with the Cholesky normalisation folded into the Kraus
 * operators (rouchon_integrator.py:226-269, 394-407, 552-582): each step adds
 * tmp + tmp^dag with tmp = sum_k D_k rho C_k to the FP64 state. */
int dqh_mesolve_rouchon1_dense(const dqh_me_inst *d_insts, int64_t ninst, void *stream);

/* ------------------------------------------------------- adaptive Rouchon */
/* One adaptive Rouchon solve of dense operators (n <= np <= 64, np % 16 == 0), all
 * in FP64. Operators in ``ops`` (device complex128 [2 + nL + npwc + nmod][np][np],
 * zero-padded): G0 = -i H0 - 1/2 sum L^dag L, LdL = sum L^dag L, the jump operators
 * L_k, then -i P_j of the piecewise-constant terms and -i O_j of the modulated ones,
 * so that G(tau) = G0 + sum_j c_j(seg) (-i P_j) + sum_j f_j(tau) (-i O_j). */
typedef struct dqh_ar_inst {
  int64_t n;           /* Hilbert dimension */
  int64_t np;          /* padded dimension: 16, 32, 48 or 64 */
  int64_t nL;          /* jump operators */
  int64_t nE;          /* expectation operators */
  int64_t ntsave;
  int64_t order;       /* 2: Rouchon 1-2 pair (Rouchon2()), 3: Rouchon 2-3 pair (Rouchon3()) */
  int64_t normalize;   /* 1: Cholesky normalisation of both maps */
  int64_t flags;       /* DQH_FLAG_SAVE_STATES */
  int64_t max_steps;
  int64_t npwc;        /* piecewise-constant terms; pwc_times has nseg breakpoints */
  int64_t nseg;
  int64_t nmod;        /* modulated terms (<= DQH_MAX_MOD), tables as in dqh_me_inst */
  int64_t mod_nint;
  double rtol, atol, safety, factormin, factormax;
  double t0, t1;       /* start time; end time = tsave[ntsave - 1] */
  double dt0;          /* first trial step */
  double mod_t0, mod_dt;
  uint64_t ops;        /* device complex128, layout above */
  uint64_t pwc_times;  /* device double [nseg] */
  uint64_t pwc_vals;   /* device complex128 [npwc][nseg - 1] */
  uint64_t mod_coef;   /* device complex128 [nmod][mod_nint][4] */
  uint64_t EopsT;      /* device complex128 [nE][np][np]: E^T, zero-padded */
  uint64_t tsave;      /* device double [ntsave], increasing */
  uint64_t rho;        /* device complex128 [np][np]: rho0 on entry, final state on exit */
  uint64_t work;       /* device complex128 [dqh_ar_work_elems(np)] */
  uint64_t states;     /* device complex128 [ntsave][n][n] (DQH_FLAG_SAVE_STATES) */
  uint64_t expects;    /* device complex128 [nE][ntsave] */
  uint64_t stats;      /* device int64 [8]: nsteps, naccepted, nrejected, status */
} dqh_ar_inst;

int64_t dqh_sizeof_ar_inst(void);
int64_t dqh_ar_work_elems(int64_t np);

/* Adaptive Rouchon2 / Rouchon3 (``Rouchon2()`` / ``Rouchon3()`` without dt:
 * MESolveAdaptiveRouchonIntegrator, rouchon_integrator.py:455-549, 585-651, with
 * the PIDController of diffrax_integrator.py:59-72 and step_ts = tsave). Per trial
 * step of h from t: the no-jump propagators of the embedded pair (midpoint / Euler
 * or Bogacki-Shampine 3(2)) on dU/dt = G(tau) U, the high and low Kraus maps
 * (KrausHeun2 / KrausEuler or KrausHeun3 / KrausHeun2, :83-302) each after its own
 * Cholesky normalisation (:552-582), the error (rho_high - rho_low) / 2 in the RMS
 * norm with the integral controller (exponent 1/order). Steps land on the save
 * times and PWC breakpoints; the first trial step is dt0. One workgroup per
 * instance; status in stats[3] (DQH_MAX_STEPS, DQH_NONFINITE, DQH_UNSUPPORTED). */
int dqh_mesolve_rouchon_adaptive_dense(const dqh_ar_inst *d_insts, int64_t ninst, void *stream);

/* complex128 products C[b] = op(A[b]) op(B[b]) for b < nbatch (np x np row-major,
 * np % 16 == 0, np <= 64; op: 0 none, 1 conjugate transpose), on the FP64 MFMA tiles
 * of dqh_mesolve_rouchon_adaptive_dense. No counterpart in the reference: the
 * kernel's product exposed for its unit test. */
int dqh_zgemm_dense(const void *A, const void *B, void *C, int64_t np, int64_t opA, int64_t opB, int64_t nbatch,
                    void *stream);

/* dqh_mesolve_tsit5_dense for a batch in which every instance has padded dimension
 * np (16, 32, 48 or 64), at most 3 jump operators and no DQH_FLAG_GLOBAL: a kernel
 * that holds only the LDS/register-resident body of that size. An instance that
 * does not match gets status DQH_UNSUPPORTED in its stats and is not solved. */
int dqh_mesolve_tsit5_dense_uniform(const dqh_me_inst *d_insts, int64_t ninst, int64_t np, void *stream);

/* dqh_mesolve_tsit5_dense_uniform for descriptors that all carry hdiag (the
 * generator's diagonal in FP64, Hnh its off-diagonal part): every vector-field
 * evaluation adds the diagonal term in FP64, so the largest single terms carry no
 * FP32 rounding. Same solve as the reference's Tsit5 mesolve (diffrax_integrator.py
 * :101-147 with the vector field of :319-323), at lower rounding error. */
int dqh_mesolve_tsit5_dense_uniform_fp64diag(const dqh_me_inst *d_insts, int64_t ninst, int64_t np, void *stream);

/* The same Tsit5 solves spread over the whole grid: every stage of every
 * instance is a set of grid-wide launches (MFMA tiles of the two GEMM phases, an
 * elementwise stage update) and the PID controller runs on the device, one
 * control block per instance; graph_steps steps (even, 0: plain launches) are
 * captured as a hipGraph and replayed until every instance is done. For large n
 * (one workgroup per instance would leave the GPU idle). Same descriptors and
 * outputs as dqh_mesolve_tsit5_dense (d_insts on the device, h_insts the same
 * array on the host); all instances share np; constant generators only (npwc ==
 * 0); expectation values need saved states. scratch: device memory of
 * dqh_tsit5_grid_scratch_bytes(np, ninst) bytes; host_flags: pinned int64[2].
 * Returns when the batch is done; results are ready once ``stream`` is
 * synchronised. */
int64_t dqh_tsit5_grid_scratch_bytes(int64_t np, int64_t ninst);
int dqh_mesolve_tsit5_dense_grid(const dqh_me_inst *d_insts, const dqh_me_inst *h_insts, int64_t ninst,
                                 void *scratch, int64_t scratch_bytes, int64_t graph_steps, int64_t *host_flags,
                                 void *stream);

/* ------------------------------------------------------------------ sparse DIA */
/* Forms of the DIA sweeps. TILED_Hh: 64 x 64 output tiles per workgroup, the input
 * staged in LDS with an h-wide halo (needs n % 64 == 0 and every |offset| <= h).
 * GLOBAL: one thread per element, neighbours read through the caches (any
 * problem). TILED: the smallest halo that fits. AUTO: TILED when it fits, else
 * GLOBAL. */
#define DQH_SWEEP_AUTO 0
#define DQH_SWEEP_GLOBAL 1
#define DQH_SWEEP_TILED 2
#define DQH_SWEEP_TILED_H8 3
#define DQH_SWEEP_TILED_H16 4
#define DQH_SWEEP_TILED_H32 5
/* Accepted for the Rouchon1 Kraus sweep and run as its GLOBAL form (the fused
 * one-pass sweep reads the normalised state X once per neighbour). */
#define DQH_SWEEP_SPLIT 6

#define DQH_DIA_MAXOPS 8
#define DQH_DIA_MAXOFF 16

/* Operators of the Kraus sweep as diagonals. Operator 0 is dM = M0 - I = -i dt Hnh
 * (used one-sided, dM X + X dM^dag, and as a sandwich dM X dM^dag); operators
 * 1.. are sqrt(dt) L_k. diags[s] points to noff[s] device rows of n complex64 with
 * A[i][i + off] stored at column i + off (scipy.sparse.dia_matrix convention,
 * dynamiqs/qarrays/sparsedia_primitives.py:11-16, 232-261). */
typedef struct dqh_dia_ops {
  int32_t nops;
  int32_t noff[DQH_DIA_MAXOPS];
  int32_t off[DQH_DIA_MAXOPS][DQH_DIA_MAXOFF];
  int32_t pad_;
  uint64_t diags[DQH_DIA_MAXOPS];
} dqh_dia_ops;

/* A batch of Rouchon1 solves sharing their (constant, DIA) operators and differing
 * in rho0. Set 0 is used for every step but the last, set nsets-1 for the last one
 * (a shorter final step has its own normalisation). */
typedef struct dqh_dia_rouchon1 {
  int64_t n;       /* padded dimension, multiple of nb */
  int64_t n_true;  /* Hilbert-space dimension */
  int64_t nb;      /* block size of the banded solve (64; half-bandwidth of S <= nb) */
  int64_t ninst;
  int64_t nsteps;
  int64_t nsave;
  int64_t nsets;   /* 1 or 2 */
  int64_t pad_;
  uint64_t Vp[2];  /* device c64[n/nb][nb][nb]: V_i - I, V_i = (T_ii^dag)^-1 */
  uint64_t Wn[2];  /* device c64[n/nb][nb][nb]: -V_i T_{i+1,i}^dag */
  dqh_dia_ops ops[2];
  uint64_t rho64;      /* device double2[ninst][n][n]: rho0 on entry, final state on exit */
  uint64_t rho32;      /* device c64[2][ninst][n][n] scratch: the state as hi + lo planes */
  uint64_t work;       /* device c64[2][ninst][n][n] scratch: the stages X and E */
  uint64_t states;     /* device c64[ninst][nsave][n_true][n_true] */
  uint64_t save_step;  /* HOST int64[nsave]: step index whose end state is saved (-1: initial) */
  uint64_t save_theta; /* HOST double[nsave]: interpolation weight inside that step */
  int64_t nE;          /* number of expectation operators (0: none)                  */
  uint64_t EopsT;      /* device c64[nE][n_true][n_true]: transposed E_k              */
  uint64_t expects;    /* device c64[ninst][nE][nsave]: Tr(E_k rho(tsave))            */
  int64_t variant;     /* DQH_SWEEP_*: form of the Kraus sweep                        */
  int64_t solve_panel; /* columns per workgroup of the block solves: 0 (auto), 32, 64 */
  uint64_t timing;     /* optional HOST double[2] (0: off): mean ms per step of the block
                          solves and of the Kraus sweep, from HIP events around steps
                          1..8; the call then synchronises the stream before returning */
  /* Block form of the normalisation (nblk = 0: the banded block solves above). When S
   * never couples indices of different sets (each set of at most DQH_DIA_MAXBLK indices,
   * in ascending order), the Cholesky factor of S is block-diagonal on those sets, and
   * T^-dag rho T^-1 is two memory-bound passes: a pass applying T_b^-dag - I to the
   * rows of each set, and a row-strip pass applying T_b^-1 - I to its columns (one
   * pass, a workgroup per set, when n <= 1024 and every set has at most 32
   * indices). Needs n <= 4096. */
  int64_t nblk;        /* number of index sets (0: banded form) */
  int64_t blk_maxd;    /* largest set size (<= DQH_DIA_MAXBLK) */
  uint64_t blk_of;     /* device int32[n]: set of each index */
  uint64_t blk_loc;    /* device int32[n]: position of each index inside its set */
  uint64_t blk_start;  /* device int32[nblk + 1]: first position of each set in blk_idx */
  uint64_t blk_idx;    /* device int32[n]: the indices of every set, ascending within a set */
  uint64_t blk_mat;    /* device int64[nblk + 1]: first element of each set's d x d matrices */
  uint64_t Lm[2];      /* device c64[sum d^2] per step set: T_b^-dag - I, row-major (upper) */
  uint64_t Rm[2];      /* device c64[sum d^2] per step set: (T_b^-1 - I)^T, row-major, so that
                          column l of T_b^-1 - I (nonzero from row l on) is contiguous */
} dqh_dia_rouchon1;

#define DQH_DIA_MAXBLK 64

int64_t dqh_sizeof_dia_rouchon1(void);

/* Fixed-step Rouchon1 with Cholesky normalisation for DIA operators
 * (replaces MESolveFixedRouchon1Integrator, rouchon_integrator.py:385-432, for
 * large sparse problems): per step two banded block solves on MFMA and one
 * bandwidth-bound stencil sweep; the state is held as two FP32 planes
 * (hi + lo) updated with FP32 increments in FP64 arithmetic. Launches on
 * ``stream`` and returns without synchronising (unless ``timing`` is set). */
int dqh_mesolve_rouchon1_dia(const dqh_dia_rouchon1 *problem, void *stream);

/* A batch of Lindblad vector fields for DIA operators shared by the batch:
 *   drho = Hnh rho + rho Hnh^dag + sum_k L_k rho L_k^dag,  Hnh = -iH - 1/2 sum L^dag L,
 * with ops.diags[0] = Hnh and ops.diags[k] = L_k (k = 1 .. nops-1) by diagonals. */
typedef struct dqh_dia_rhs {
  int64_t n;        /* dimension; the tiled sweep needs n % 64 == 0 */
  int64_t ninst;
  int64_t variant;  /* DQH_SWEEP_* */
  int64_t pad_;
  dqh_dia_ops ops;
  uint64_t rho;     /* device c64[ninst][n][n] */
  uint64_t drho;    /* device c64[ninst][n][n] */
} dqh_dia_rhs;

int64_t dqh_sizeof_dia_rhs(void);

/* The Lindblad vector field for sparse-DIA operators as a bandwidth-bound sweep
 * (replaces vector_field_unvec_hermitian, diffrax_integrator.py:319-323, with the
 * DIA products of sparsedia_primitives.py:172-261). Launches on ``stream``. */
int dqh_lindblad_rhs_dia(const dqh_dia_rhs *problem, void *stream);

/* A batch of adaptive Tsit5 solves for DIA operators shared by the batch (the
 * instances differ in rho0; each keeps its own t, dt and accept/reject). */
typedef struct dqh_dia_tsit5 {
  int64_t n;            /* padded dimension; the tiled sweep needs n % 64 == 0 */
  int64_t n_true;       /* Hilbert-space dimension */
  int64_t ninst;
  int64_t nsave;
  int64_t max_steps;
  int64_t nE;
  int64_t variant;      /* DQH_SWEEP_* */
  int64_t graph_steps;  /* steps per captured hipGraph (even); 0: plain launches */
  double rtol, atol, safety, factormin, factormax, t0, t1;
  int64_t flags;        /* DQH_FLAG_DOPRI5 selects the Dopri5 pair */
  dqh_dia_ops ops;      /* diags[0] = Hnh = -iH - 1/2 sum L^dag L, diags[k] = L_k */
  uint64_t y;           /* device double2[ninst][n][n]: rho0 on entry, final state on exit */
  uint64_t work;        /* device scratch of dqh_dia_tsit5_work_bytes(n, ninst) bytes */
  uint64_t tsave;       /* device double[nsave], increasing */
  uint64_t states;      /* device c64[ninst][nsave][n_true][n_true] */
  uint64_t EopsT;       /* device c64[nE][n_true][n_true] */
  uint64_t expects;     /* device c64[ninst][nE][nsave] */
  uint64_t stats;       /* device int64[ninst][8]: nsteps, naccepted, nrejected, status, nrhs */
  uint64_t host_flags;  /* HOST pinned int64[2]: completion counts read back by the host loop */
  /* optional time-dependent Hamiltonian terms (npwc + nmod = 0: a constant H).
   * H(t) = H0 + sum_j c_j(t) O_j: before every stage the kernels rebuild the diagonals
   * of Hnh(t) = ops.diags[0] + sum_j c_j(t) Hterms[j] per instance into hcur (each
   * term holds -i O_j on the noff[0] offsets of operator 0). PWC coefficients
   * (time_qarray.py:572-666) take their value on the segment of the step start
   * (zero outside [pwc_times[0], pwc_times[nseg-1])), and steps land on the
   * breakpoints, after which the first stage is evaluated anew. Modulated
   * coefficients come from piecewise-cubic tables laid out as in dqh_me_inst,
   * evaluated at the stage time. */
  int64_t npwc;         /* PWC terms (first in Hterms) */
  int64_t nseg;         /* breakpoints in pwc_times */
  int64_t nmod;         /* modulated terms (after the PWC ones in Hterms) */
  uint64_t Hterms;      /* device c64[npwc + nmod][noff[0]][n] */
  uint64_t pwc_times;   /* device double[nseg], increasing */
  uint64_t pwc_vals;    /* device double[npwc][nseg - 1][2]: complex coefficient per segment */
  uint64_t mod_coef;    /* device double2[nmod][mod_nint][4] */
  double mod_t0, mod_dt;
  int64_t mod_nint;
  uint64_t hcur;        /* device c64[ninst][noff[0]][n] scratch */
} dqh_dia_tsit5;

int64_t dqh_sizeof_dia_tsit5(void);
int64_t dqh_dia_tsit5_work_bytes(int64_t n, int64_t ninst);

/* Adaptive Tsit5 with the DIA Lindblad sweep as vector field (replaces
 * MESolveDiffraxIntegrator.run with Tsit5, diffrax_integrator.py:101-147,
 * 373-378, for large sparse problems). Stage sweeps and the device-side
 * controller run as a captured hipGraph of graph_steps steps, replayed until every
 * instance reached t1 or stopped (status in stats). Returns when the batch is
 * done; results are ready once ``stream`` is synchronised. */
int dqh_mesolve_tsit5_dia(const dqh_dia_tsit5 *problem, void *stream);

/* ------------------------------------------------------------------ sesolve */
/* Adaptive Tsit5 (or Dopri5 with DQH_FLAG_DOPRI5) solve of the Schrodinger
 * equation d psi/dt = -i H(t) psi (replaces SESolveDiffraxIntegrator.run,
 * diffrax_integrator.py:225-228, 257-280, with the controller of :101-147), one
 * workgroup per instance; np_max = the largest np of the batch (<= 1024) picks the
 * one-wave kernel (np_max <= 64) or a four-wave one (LDS sized for 256 or 1024). Same descriptor as the mesolve kernels, read
 * for a ket: Hnh = (-i H0)^T, Hpwc / Hmod = the stacked (-i O_k)^T of the
 * piecewise-constant / modulated terms (complex64 [np][np] each), EopsT = E^T,
 * y = double2 [np] holding psi0 (n used), nL = 0; outputs states = complex64
 * [ntsave][n] (or [n] without DQH_FLAG_SAVE_STATES), final_state = complex64 [n],
 * expects[e][i] = <psi(tsave[i])| E_e |psi(tsave[i])>. */
int dqh_sesolve_tsit5(const dqh_me_inst *d_insts, int64_t ninst, int64_t np_max, void *stream);

/* ------------------------------------------------------------------ Expm */
/* Propagators of piecewise-constant generators by explicit exponentiation
 * (replaces ExpmIntegrator.run, expm_integrator.py:60-90: expm of dt_k A_k on every
 * interval, then U_k = E_k U_{k-1}). X: device complex64 [nchain][nint][N][N], the
 * generators already scaled, X = dt_k A_k / 2^s; E = T_m(X)^(2^s) with T_m the
 * degree-m Taylor polynomial (Horner form) and s squarings. U: device complex64
 * [nchain][nint][N][N] receives U[b][k] = E[b][k] ... E[b][0]. work: device memory
 * of dqh_expm_work_bytes(nchain, nint, N) bytes. N % 16 == 0, N <= 4096,
 * 1 <= m <= 30, 0 <= s <= 60. */
int64_t dqh_expm_work_bytes(int64_t nchain, int64_t nint, int64_t N);
int dqh_expm_propagators(const void *X, void *work, void *U, int64_t nchain, int64_t nint, int64_t N, int64_t m,
                         int64_t s, void *stream);

/* The same exponentials and products with no allocation and no work argument:
 * the work memory is a fixed device array of DQH_EXPM_NOALLOC_MAX_BYTES bytes
 * inside the library (one per GPU, held until the process ends), so
 * dqh_expm_work_bytes(nchain, nint, N) must not exceed it (DQH_BAD_SHAPE
 * otherwise). No counterpart in the reference: this entry point is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). It calls no hipMalloc / hipFree. Every call on
 * a GPU uses that GPU's one array: calls that overlap on one GPU (different streams
 * or threads) race and must be serialised by the caller. Results equal
 * dqh_expm_propagators' for the same inputs. */
#define DQH_EXPM_NOALLOC_MAX_BYTES (64ll << 20)
int dqh_expm_propagators_noalloc(const void *X, void *U, int64_t nchain, int64_t nint, int64_t N, int64_t m,
                                 int64_t s, void *stream);

/* Batched complex64 product C[b] = A[b] B[b / bdiv] for b < nbatch (N x N row-major,
 * N % 16 == 0, strides in elements; C must not alias A or B), the GEMM of
 * dqh_expm_propagators. The Expm solves apply the propagators to a block of initial
 * states (vec(rho0) or psi0 as columns): ExpmIntegrator.run's x @ carry with a
 * state as carry (expm_integrator.py:73-78). */
int dqh_cgemm_batched(const void *A, int64_t sA, const void *B, int64_t sB, int64_t bdiv, void *C, int64_t sC,
                      int64_t N, int64_t nbatch, void *stream);

/* ------------------------------------------------------------------ gradients */
/* Gauss-Legendre quadrature of the adjoint gradient of piecewise-constant controls
 * (replaces the reverse-mode path jax.grad -> dx.RecursiveCheckpointAdjoint,
 * diffrax_integrator.py:88-99): for every instance and segment k,
 *   grad[inst][j][k] = sum_{q in segment k} weights[q] Re Tr(lam_q (-i [O_j, rho_q])),
 * rho_nodes / lam_nodes: device complex64 [ninst][nnodes][n][n], ops: device
 * complex64 [nterms][n][n] (nterms <= 8), weights: device double[nnodes],
 * seg_first: device int32[nseg + 1], grad: device double[ninst][nterms][nseg]. */
int dqh_adjoint_pwc_quadrature(int64_t ninst, int64_t n, int64_t nnodes, int64_t nterms, int64_t nseg,
                               const void *rho_nodes, const void *lam_nodes, const void *ops,
                               const double *weights, const int32_t *seg_first, double *grad, void *stream);

/* The same quadrature for control operators stored by diagonals (ops->nops
 * operators, ops->diags[j] device complex64 [noff[j]][n]): 2 d n^2 products per
 * node instead of 2 n^3 for d diagonals. n <= 64. */
int dqh_adjoint_pwc_quadrature_dia(int64_t ninst, int64_t n, int64_t nnodes, int64_t nseg, const void *rho_nodes,
                                   const void *lam_nodes, const dqh_dia_ops *ops, const double *weights,
                                   const int32_t *seg_first, double *grad, void *stream);

/* Gauss-Legendre quadrature of the adjoint gradient with respect to jump rates:
 * for L_r = sqrt(kappa_r) A_r,
 *   grad[inst][r][k] = sum_{q in segment k} weights[q] Tr(lam_q D[A_r] rho_q),
 * D[A] rho = A rho A^dag - (N rho + rho N) / 2 with N = A^dag A. Aops, Nops: device
 * complex64 [nrates][n][n] (n <= 64, nrates <= 8); other arguments as in
 * dqh_adjoint_pwc_quadrature. */
int dqh_adjoint_rate_quadrature(int64_t ninst, int64_t n, int64_t nnodes, int64_t nrates, int64_t nseg,
                                const void *rho_nodes, const void *lam_nodes, const void *Aops, const void *Nops,
                                const double *weights, const int32_t *seg_first, double *grad, void *stream);

/* Device address of pinned host memory (hipHostGetDevicePointer): descriptors that
 * point outputs at host memory use it. Returns DQH_HIP_ERROR when ``host`` is not
 * pinned, mapped memory. No counterpart in the reference (serving plumbing). */
int dqh_host_device_pointer(void *host, void **device);

/* Message of the last failing call on this thread. */
const char *dqh_last_error(void);

/* Device name / compute units of the current device (diagnostics). */
int dqh_device_info(char *name, int64_t len, int64_t *cus);

#ifdef __cplusplus
}
#endif

#endif /* DQH_H */
