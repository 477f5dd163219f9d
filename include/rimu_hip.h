/*
 * rimu_hip.h -- C ABI of the MI355X (gfx950) FCIQMC step engine.
 *
 * The library replaces the vector-matrix product of one FCIQMC step of Rimu.jl
 * (apply_operator!, Interfaces/dictvectors.jl:112-154) together with the reductions that
 * advance! performs right after it (walkernumber_and_length, fciqmc.jl:144; the
 * ProjectedEnergy post-step, poststepstrategy.jl:110-121).  Vectors live in device memory
 * (HBM) between calls; the host keeps the step loop, and the shift strategy unless
 * DoubleLogUpdate runs on the device (rimu_engine_set_shift_update).
 *
 * Conventions: every function returns RIMU_OK (0) or an error code; rimu_last_error()
 * gives a message for the calling thread.  Host pointers are borrowed for the duration of
 * the call and copied.  No function throws or unwinds across the ABI.  An engine is driven
 * by one host thread at a time (its work is ordered on one HIP stream).
 *
 * Addresses are BoseFS bit strings of B = N + M - 1 <= 128 bits, in the layout of
 * BitStringAddresses/bitstring.jl:464-508: mode 1 in the least significant bits, mode i
 * written as n_i one-bits followed by a zero separator.  Through this ABI an address is
 * kw uint64 words: kw = 1 for B <= 64, kw = 2 (low word first, then high word) for
 * 64 < B <= 128.  The order of addresses is the numeric order of the whole bit string,
 * Rimu's isless on BitStrings (bitstring.jl:317-324, chunk 1 most significant).  Arrays
 * of "keys" below hold kw words per address, where kw is that of the engine's or model's
 * M and N.  The batched and dense entry points take one-word addresses (B <= 64).
 */
#ifndef RIMU_HIP_H
#define RIMU_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RIMU_OK = 0,
    RIMU_ERR_ARG = 1,        /* invalid argument (the Julia side throws ArgumentError) */
    RIMU_ERR_HIP = 2,        /* HIP runtime failure */
    RIMU_ERR_CAPACITY = 3,   /* a device buffer would overflow (maps to max_length abort) */
    RIMU_ERR_NODEVICE = 4,   /* no usable gfx950 device */
    RIMU_ERR_INTERNAL = 5
};

/* Hamiltonian kinds: HubbardReal1D (Hamiltonians/HubbardReal1D.jl:22-62),
 * HubbardMom1D (Hamiltonians/HubbardMom1D.jl:43-239, BoseFS branch) and HubbardRealSpace
 * (Hamiltonians/HubbardRealSpace.jl:130-391 with the CubicGrid geometries of
 * Hamiltonians/geometry.jl) on a BoseFS, a FermiFS or a two-component CompositeFS of them. */
enum { RIMU_MODEL_HUBBARD_REAL_1D = 1, RIMU_MODEL_HUBBARD_MOM_1D = 2, RIMU_MODEL_HUBBARD_REAL_SPACE = 3 };

/* Stochastic styles: IsStochasticInteger (StochasticStyles/styles.jl:11-25),
 * IsDeterministic (:76-105), IsDynamicSemistochastic (:175-214). */
enum {
    RIMU_STYLE_STOCHASTIC_INTEGER = 1,
    RIMU_STYLE_DETERMINISTIC = 2,
    RIMU_STYLE_DYNAMIC_SEMISTOCHASTIC = 3
};

/* Plain data of a Hamiltonian: the fields of HubbardReal1D{TT,A,U,T} (type parameters
 * u, t; HubbardReal1D.jl:46-49), HubbardMom1D (u, t and the dispersion table kes,
 * HubbardMom1D.jl:43-47, :80-85; kes[j-1] is epsilon of mode j) or HubbardRealSpace
 * (HubbardRealSpace.jl:163-172: t, u, the precomputed trap potential and the geometry,
 * given here as its neighbour table).
 *
 * HubbardRealSpace keys: a single component is its own bit string (BoseFS: N + M - 1
 * bits as for the 1D models; FermiFS: M bits, site j at bit j-1, fermifs.jl and
 * bitstring.jl:713-733); two components are component 1's bits above component 2's, so
 * numeric order is CompositeFS's isless (multicomponent.jl:79).  At most 128 bits. */
typedef struct rimu_model {
    int32_t kind;
    int32_t M;          /* modes (sites) */
    int32_t N;          /* particles (HubbardRealSpace: of component 1) */
    int32_t reserved;
    double u;           /* HubbardRealSpace: u[1,1] */
    double t;           /* HubbardRealSpace: t[1] */
    double kes[64];     /* HubbardRealSpace: trap potential of component 1 at site j+1 */
    /* HubbardRealSpace only (zero for the 1D models): */
    int32_t ncomp;      /* 1, or 2 for a CompositeFS */
    int32_t N2;         /* particles of component 2 */
    int32_t fermi;      /* bit c set: component c+1 is a FermiFS, else a BoseFS */
    int32_t ndir;       /* 2D neighbour directions per site, D <= 3 */
    int32_t has_u;      /* 0 when u is the zero matrix (HubbardRealSpace.jl:211) */
    int32_t has_pot;    /* 1 when the trap strengths v are nonzero (:214-224) */
    double t2;          /* t[2] */
    double u12;         /* u[1,2] = u[2,1] */
    double u22;         /* u[2,2] */
    double pot2[64];    /* trap potential of component 2 at site j+1 */
    uint8_t nbr[64][6]; /* neighbor_site(geometry, s+1, d+1) - 1 (geometry.jl:236-240), or 255
                           where the move leaves a hard wall; d in Directions order
                           (+e1, .., +eD, -e1, .., -eD, geometry.jl:127-160) */
} rimu_model;

/* Plain data of a StochasticStyle and its CompressionStrategy.
 * compression: 0 NoCompression, 1 ThresholdCompression (compression.jl:8-42).
 * proj_threshold: projection threshold of deposits (spawning.jl:13-31); 0 = exact.
 * rel_threshold / abs_threshold: DynamicSemistochastic (spawning.jl:358-378). */
typedef struct rimu_style {
    int32_t kind;
    int32_t compression;
    double compression_threshold;
    double proj_threshold;
    double rel_threshold;
    double abs_threshold;
} rimu_style;

/* Step results.  The first eight fields are the step_stats of the styles
 * (styles.jl:14-20, :94-96, :203-209) and ThresholdCompression's len_before
 * (compression.jl:16); len and norm are walkernumber_and_length (abstractdvec.jl:270);
 * vproj and hproj are the ProjectedEnergy columns (poststepstrategy.jl:110-121). */
typedef struct rimu_step_stats {
    double spawn_attempts;
    double exact_steps;
    double inexact_steps;
    double spawns;
    double deaths;
    double clones;
    double zombies;
    double len_before;
    double len;
    double norm;
    double vproj;
    double hproj;
    int32_t overflow;   /* 1 when a buffer capacity was hit: the step is void */
    int32_t reserved;
} rimu_step_stats;

typedef struct rimu_engine rimu_engine;

/* Last error message of the calling thread ("" if none). */
const char* rimu_last_error(void);

/* Number of visible devices (0 when there is no GPU). */
int rimu_device_count(int32_t* count);

/* Create an engine: a device-resident vector plus working memory for one FCIQMC state.
 * Replaces the DVec/PDVec + working_memory pair (Interfaces/dictvectors.jl:87,
 * DictVectors/pdworkingmemory.jl:86).  max_entries bounds the vector length (Rimu's
 * max_length, projector_monte_carlo_problem.jl:243-253); max_spawns bounds the spawn list
 * of one step.  seed keys the Philox streams of this engine. */
int rimu_engine_create(const rimu_model* model, const rimu_style* style, uint64_t seed,
                       int64_t max_entries, int64_t max_spawns, rimu_engine** out);
int rimu_engine_destroy(rimu_engine* e);
/* Grow the capacities (never shrinks; the vector is kept).  A step that failed with
 * RIMU_ERR_CAPACITY left the vector unchanged and, because every random draw is a function
 * of (seed, address, attempt, step), gives the same result when repeated. */
int rimu_engine_reserve(rimu_engine* e, int64_t max_entries, int64_t max_spawns);

/* Replace the vector.  keys (n addresses of kw words) must be strictly ascending; vals are int64 for
 * RIMU_STYLE_STOCHASTIC_INTEGER and double otherwise; zero values are dropped.
 * (copyto!/setindex! on the DVec, dvec.jl:129-136.) */
int rimu_vector_upload(rimu_engine* e, const uint64_t* keys, const void* vals, int64_t n);
/* Current length (length(v)). */
int rimu_vector_length(rimu_engine* e, int64_t* n);
/* Copy the vector out, ascending keys of kw words each (pairs(v)).  cap is the capacity of
 * keys/vals in addresses. */
int rimu_vector_download(rimu_engine* e, uint64_t* keys, void* vals, int64_t cap, int64_t* n);

/* Projector of the ProjectedEnergy post-step (poststepstrategy.jl:82-121): vproj is
 * sum_i pc_i * v[pk_i], hproj is sum_i hc_i * v[hk_i] (hproj = freeze(H' * projector)).
 * keys (kw words each) need not be sorted.  n = nh = 0 disables the projection. */
int rimu_set_projector(rimu_engine* e, const uint64_t* pk, const double* pc, int64_t n,
                       const uint64_t* hk, const double* hc, int64_t nh);

/* One step v <- compress(T v) with T = 1 + dtau (shift - H) (fciqmc.jl:78-112, :139),
 * apply_operator! with boost (Interfaces/dictvectors.jl:112), then len, norm and the
 * projections.  step numbers the Philox streams.  On overflow (stats->overflow = 1,
 * return RIMU_ERR_CAPACITY) the vector is left unchanged.  With a device shift update
 * (rimu_engine_set_shift_update) the step uses the engine's shift and ignores the argument. */
int rimu_step(rimu_engine* e, int64_t step, double shift, double dtau, double boost,
              rimu_step_stats* stats);

/* Shift updates: the caller's (the shift argument of rimu_step), or DoubleLogUpdate
 * (shiftstrategy.jl:161-181) on the device. */
enum { RIMU_SHIFT_HOST = 0, RIMU_SHIFT_DOUBLE_LOG = 1 };

/* Keep the shift in device memory and update it there after every step, as
 * update_shift_parameters! of DoubleLogUpdate (shiftstrategy.jl:174-181) does with the
 * step's norm N: shift <- shift - (xi/dtau ln(N/target_walkers) + zeta/dtau ln(N/pnorm)),
 * pnorm <- N, with the dtau of the step (ConstantTimeStep).  shift and pnorm are the
 * initial shift parameters (DefaultShiftParameters, shiftstrategy.jl:32-38; pnorm is the
 * walker number of the start vector); xi = zeta^2/4 is the reference's default.  The next
 * step reads the shift from device memory, so no host value enters the step.  A step that
 * fails or is repeated leaves the shift parameters unchanged, like the vector.
 * RIMU_SHIFT_HOST returns the shift to the caller. */
int rimu_engine_set_shift_update(rimu_engine* e, int32_t mode, double shift, double pnorm,
                                 double target_walkers, double zeta, double xi);
/* Shift parameters after the last step (read back with the step's statistics, no device
 * access here): the shift the next step uses, and pnorm.  RIMU_ERR_ARG without a device
 * shift update. */
int rimu_engine_get_shift(rimu_engine* e, double* shift, double* pnorm);

/* Initiator rules: NonInitiator, Initiator, SimpleInitiator, CoherentInitiator
 * (DictVectors/initiators.jl:132-236). */
enum {
    RIMU_INITIATOR_NONE = 0,
    RIMU_INITIATOR_DEFAULT = 1,
    RIMU_INITIATOR_SIMPLE = 2,
    RIMU_INITIATOR_COHERENT = 3
};

/* Select the initiator rule of the engine's vector: the `initiator` keyword of
 * PDVec / InitiatorDVec (pdvec.jl:168-185, initiatordvec.jl:42-49); threshold is the
 * rule's `threshold` (an entry with |v| > threshold is an initiator).  From the next step
 * on, every deposit is classified by its parent (to_initiator_value, initiators.jl:143-157)
 * and each address keeps the rule's from_initiator_value of its three components
 * (move_and_compress!, pdworkingmemory.jl:262-273).  Under NoCompression an address that
 * received only discarded deposits stays in the vector with value 0, as in the reference.
 * Needs N + M - 1 <= 64 * kw - 1 (one free bit for the deposit's class).
 * RIMU_INITIATOR_NONE restores the plain step. */
int rimu_engine_set_initiator(rimu_engine* e, int32_t rule, double threshold);

/* Operators of the three-way dot: a Hamiltonian (model, off-diagonals included),
 * G2RealCorrelator(d) = (1/M) sum_i n_i (n_{i+d} - delta_{0d}) with |d| < M
 * (Hamiltonians/correlation_functions.jl:68-92), DensityMatrixDiagonal(mode) = n_mode with
 * 1 <= mode <= M (Hamiltonians/DensityMatrixDiagonal.jl:19-33). */
enum { RIMU_OP_HAMILTONIAN = 1, RIMU_OP_G2_REAL = 2, RIMU_OP_DENSITY_DIAGONAL = 3 };
typedef struct rimu_operator {
    int32_t kind;
    int32_t param;      /* G2RealCorrelator: d; DensityMatrixDiagonal: mode */
    rimu_model model;   /* RIMU_OP_HAMILTONIAN: same M and N as the vectors */
} rimu_operator;

/* Overlaps of two engines' vectors (they may be the same engine): out[0] = dot(left, right),
 * out[1 + o] = dot(left, ops[o], right) for nops <= 8 operators -- LinearAlgebra.dot on
 * DVecs and dot_from_right (DictVectors/abstractdvec.jl:313-324), the values AllOverlaps
 * reports per replica pair (replicastrategy.jl:160-183).  Both vectors share M and N.
 * Sums run in a fixed order on the device; integer dots are exact below 2^53. */
int rimu_overlaps(rimu_engine* left, rimu_engine* right, const rimu_operator* ops, int32_t nops,
                  double* out);

/* (No counterpart in the reference; for tests and tuning.)  Addresses of more than 20 bits
 * are stored as a bijective mix of their bits.  Where the bucket merge does not apply (see
 * rimu_engine_set_radix_mode), a step's spawn records with more than `bits` key bits are
 * radix-sorted on the top `bits` bits of the mixed keys only (default 40); runs that agree
 * there are then ordered one by one, or, when such runs are long or many, by the remaining
 * passes.  Any setting gives the same step results; fewer bits exercise the run ordering. */
int rimu_engine_set_sort_bits(rimu_engine* e, int32_t bits);

/* (No counterpart in the reference; for tests and tuning.)  Sort and merge of the spawn
 * records.  2 (default): where eligible (one-word keys of more than 20 bits, at least 4096
 * records), sort on the top bits of the stored keys only and merge bucket by bucket, each
 * bucket's records and vector entries sorted and folded in LDS; other steps as 1.  1: one histogram
 * pass for every digit and scatters that find their offsets by a look-back over the tiles
 * before them, then the merge path.  0: as 1 with counts per pass and a scan between passes.
 * All give the same step results. */
int rimu_engine_set_radix_mode(rimu_engine* e, int32_t mode);

/* (No counterpart in the reference; for tests.)  Every look-back wait of the step kernels
 * (the plan's attempt offsets, the radix scatters' digit offsets) is bounded by spin_limit
 * rounds (default 2^25).  A radix wait that gives up makes the engine repeat the step with
 * the per-pass sort (radix mode 0, no look-back; same result); a plan wait that gives up
 * fails the step with RIMU_ERR_INTERNAL, leaving the vector unchanged.  withhold_first = 1
 * makes the first tile of every radix pass keep its counts to itself, so that every later
 * tile's wait runs into the bound: a host-forced test of the give-up path. */
int rimu_engine_set_lookback(rimu_engine* e, uint32_t spin_limit, int32_t withhold_first);

/* Per-kernel device time accumulated since the last reset, measured with HIP events on
 * the engine's stream when timing is enabled.  names/ms/launches hold up to cap kernels. */
int rimu_engine_set_timing(rimu_engine* e, int32_t enabled);
int rimu_engine_kernel_times(rimu_engine* e, char* names, int32_t name_len, double* ms,
                             int64_t* launches, int32_t cap, int32_t* n);
/* Bytes moved per step by the last step, by the algorithmic byte count of the merge:
 * len_in*(k+v) + 2*S*(k+v) + len_out*(k+v).  Also returns S (records emitted, diagonal
 * survivors included). */
int rimu_engine_last_traffic(rimu_engine* e, double* alg_bytes, double* records);

/* Annihilation of a spawn list on its own (deposit! folded over the list,
 * Interfaces/dictvectors.jl:49-51 with dvec.jl:129-136): equal keys are summed in list
 * order, exact zeros dropped; output ascending.  vtype 0: int64 values, 1: double.  Keys
 * are single uint64 words. */
int rimu_annihilate(int32_t vtype, const uint64_t* keys, const void* vals, int64_t n,
                    uint64_t* out_keys, void* out_vals, int64_t cap, int64_t* n_out);

/* rimu_annihilate without device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (serving loops and harnesses that forbid
 * hipMalloc).  Same arguments, result and errors as rimu_annihilate, for at most 2^20
 * records (larger lists return RIMU_ERR_CAPACITY).  Its working memory is one
 * library-owned device array of about 37 MiB per GPU, held until the process ends; it calls
 * no hipMalloc/hipFree.  Calls that overlap on one GPU share that array and race: the caller
 * serialises them. */
int rimu_annihilate_noalloc(int32_t vtype, const uint64_t* keys, const void* vals, int64_t n,
                            uint64_t* out_keys, void* out_vals, int64_t cap, int64_t* n_out);

/* Hamiltonian interface on the device for a batch of addresses (for tests and
 * ProjectedEnergy set-up): diagonal_element and num_offdiagonals
 * (Interfaces/hamiltonians.jl), and get_offdiagonal for one chosen index per address.
 * addrs and new_addr hold kw words per address (kw of the model's M and N). */
int rimu_hamiltonian_eval(const rimu_model* model, const uint64_t* addrs, const int64_t* chosen,
                          int64_t n, double* diag, int64_t* num_offdiag, uint64_t* new_addr,
                          double* offdiag_value);

/* ---------------------------------------------------------------------------------------
 * Deterministic products on the ranked basis (the IsDeterministic path at full dimension)
 * --------------------------------------------------------------------------------------- */
typedef struct rimu_dense rimu_dense;

/* binomial(N+M-1, N): number of N-boson M-mode addresses (number_conserving_bose_dimension,
 * Hamiltonians/abstract.jl:103-110). */
int rimu_basis_size(int32_t M, int32_t N, int64_t* dim);

/* The ranked basis on the device for nrhs HubbardReal1D Hamiltonians that share M and N
 * (e.g. one per u).  Row r of every vector belongs to the address of colex rank r, which is
 * the r-th address in ascending numeric order of the bit strings (the sorted basis of
 * ExactDiagonalization/basis_fock.jl).  Vectors are row-major [dim][nrhs], nrhs <= 8. */
int rimu_dense_create(const rimu_model* models, int32_t nrhs, rimu_dense** out);
int rimu_dense_destroy(rimu_dense* d);
int rimu_dense_dim(rimu_dense* d, int64_t* dim);
/* The dim addresses, ascending. */
int rimu_dense_basis(rimu_dense* d, uint64_t* keys);

/* w_i = (a_i + b_i H_i) v_i for every right-hand side at once: mul!(w, H, v)
 * (DictVectors/abstractdvec.jl:275-284) for a = 0, b = 1, and one IsDeterministic step with
 * T = 1 + dt (S - H) (styles.jl:101-105, fciqmc.jl:93-112) for a = 1 + dt S, b = -dt. */
int rimu_dense_apply(rimu_dense* d, const double* a, const double* b, const double* v, double* w);

/* Pullback of y_i = (a_i + b_i H_i) v_i, with H_i = u_i D + t_i K (D = sum n(n-1)/2 on the
 * diagonal, K the hopping with unit amplitude): given the cotangent ybar,
 *   vbar_i = (a_i + b_i H_i) ybar_i   (H is real symmetric),
 *   ubar_i = b_i <ybar_i|D|v_i>,   tbar_i = b_i <ybar_i|K|v_i>.
 * No counterpart in the reference: it is the rrule a differentiable mul! needs
 * (ChainRulesCore), for callers that differentiate through the projector. */
int rimu_dense_pullback(rimu_dense* d, const double* a, const double* b, const double* v,
                        const double* ybar, double* vbar, double* ubar, double* tbar);

/* K deterministic steps v <- (1 + dt (S_i - H_i)) v from v0 with fixed shifts S_i, then the
 * Rayleigh quotient E_i = <v|H_i|v>/<v|v> (rayleigh_quotient, Hamiltonians/abstract.jl:153)
 * and dE_i/du_i, dE_i/dt_i by the adjoint of the whole projection: a reverse sweep that
 * recomputes each segment of ckpt steps from a stored checkpoint.  hf_u, hf_t receive
 * <v|D|v>/<v|v> and <v|K|v>/<v|v> of the final vector (Hellmann-Feynman values), v_final the
 * final vector.  Output pointers may be null.  (No counterpart in the reference.) */
int rimu_dense_projector(rimu_dense* d, int64_t K, double dtau, const double* shift, int64_t ckpt,
                         const double* v0, double* energy, double* grad_u, double* grad_t,
                         double* hf_u, double* hf_t, double* v_final);

/* Lowest eigenvalue of every right-hand side's H by Lanczos on the device, started from
 * v0 (host, [dim][R]; its Krylov space must reach the ground state, e.g. the start
 * address, which overlaps the nodeless ground state of HubbardReal1D).  Stops when every
 * lowest Ritz value changes by at most tol * max(1, |E|) over 16 iterations, when a Krylov
 * space is exhausted, or after max_iters.  iters (may be null) receives the iterations run.
 * Replaces the ground-state energy of solve(ExactDiagonalizationProblem(H)) with the
 * KrylovKit Lanczos solver (ext/KrylovKitExt.jl:36-78, ExactDiagonalization/solve.jl),
 * the yardstick Rimu's tests hold FCIQMC energies against. */
int rimu_dense_ground_state(rimu_dense* d, const double* v0, int64_t max_iters, double tol,
                            double* energy, int64_t* iters);

/* ---------------------------------------------------------------------------------------
 * Batched engine: many independent IsStochasticInteger FCIQMC runs of HubbardReal1D
 * advanced together, every run's spawning split over workgroups by its target walker
 * number; a step is five launches covering all runs (spawn slices, per-run sort, merge
 * slices, copy slices, per-run final phase) or one persistent launch that runs the same
 * phases from one work queue per XCD (rimu_batch_set_persistent).  Steps are captured in a
 * HIP graph.
 * Each run is Rimu's lomc!/ProjectorMonteCarloProblem loop with DoubleLogUpdate
 * (shiftstrategy.jl:174-181) and a ProjectedEnergy onto its start address
 * (poststepstrategy.jl:82-121); the shift update happens on the device.
 * --------------------------------------------------------------------------------------- */
typedef struct rimu_batch rimu_batch;

typedef struct rimu_batch_run {
    int32_t M;
    int32_t N;
    double u;
    double t;
    uint64_t seed;
    uint64_t start_address;     /* BoseFS bit pattern, population start_population */
    int64_t start_population;
    double shift;               /* initial shift */
    double target_walkers;      /* DoubleLogUpdate target */
    double zeta;                /* DoubleLogUpdate damping (0.08) */
    double xi;                  /* DoubleLogUpdate restoring force (zeta^2/4) */
    double dtau;
    int64_t max_length;         /* a vector longer than this stops the run */
} rimu_batch_run;

/* Per-step history of every run, row-major [steps][nruns]: shift, norm (walkers), len,
 * vproj, hproj, spawn_attempts.  history_steps bounds the number of recorded steps. */
int rimu_batch_create(const rimu_batch_run* runs, int32_t nruns, int64_t history_steps,
                      rimu_batch** out);
int rimu_batch_destroy(rimu_batch* b);
/* Launch configuration: workgroup size (256, 512 or 1024) and radix digit width (4..8). */
int rimu_batch_set_config(rimu_batch* b, int32_t block_size, int32_t radix_bits);
/* 1: each step is one persistent launch (every run on one XCD; workgroups of the
 * configured size take the items of their XCD's runs from a queue, an item waiting for its
 * run's previous phase); 0: five launches per step.  Results are the same either way. */
int rimu_batch_set_persistent(rimu_batch* b, int32_t persistent);
/* Bound of the persistent step's waits (rounds of one s_sleep; default 2^25) and a
 * host-forced test of its give-up path: with withhold_first, the first run of the first
 * queue never reports its spawn slices, so its sort's wait runs into the bound, every
 * remaining item is skipped and rimu_batch_step returns RIMU_ERR_INTERNAL.  Also clears the
 * persistent step's counters. */
int rimu_batch_set_wait(rimu_batch* b, uint32_t spin_limit, int32_t withhold_first);
/* Advance every live run by nsteps; steps_per_graph launches are captured once in a HIP
 * graph and replayed.  A run stops (its status says why) and the others continue when: its
 * population dies; its vector outgrows max_length (as fciqmc.jl:155-179, that step is
 * completed, recorded and kept, so the run's vector is longer than max_length); or a step's
 * spawns overflow the run's spawn buffer (4 target_walkers + 4096 records: that step is
 * void and leaves no history row). */
int rimu_batch_step(rimu_batch* b, int64_t nsteps, int32_t steps_per_graph);
/* Steps recorded so far (the most rows of any run) and the history arrays (each
 * [steps][nruns], may be null); rows a run has not written (it stopped earlier) hold 0. */
int rimu_batch_history(rimu_batch* b, int64_t* nsteps, double* shift, double* norm, double* len,
                       double* vproj, double* hproj, double* attempts);
/* Status of every run: 0 running, 1 population dead, 2 vector overflow (max_length),
 * 3 spawn buffer overflow; and the current length of each vector. */
int rimu_batch_status(rimu_batch* b, int32_t* status, int64_t* len);
/* Time `trial_steps` steps for every (launch kind, block size, radix width) candidate from a
 * snapshot of the current state (restored after each trial; results do not depend on the
 * choice), select the fastest, which the batch keeps.  ms receives the per-step time of each
 * of the 18 candidates in the order persistent in {0, 1} x block in {256, 512, 1024} x radix
 * in {4, 6, 8} (a persistent trial whose wait gave up is given HUGE_VAL and its counters
 * are cleared); best_block/best_radix/best_persistent the choice. */
int rimu_batch_autotune(rimu_batch* b, int64_t trial_steps, int32_t* best_block,
                        int32_t* best_radix, int32_t* best_persistent, double* ms);
/* The vector of one run (ascending keys, int64 values). */
int rimu_batch_vector(rimu_batch* b, int32_t run, uint64_t* keys, int64_t* vals, int64_t cap,
                      int64_t* n);

#ifdef __cplusplus
}
#endif

#endif /* RIMU_HIP_H */
