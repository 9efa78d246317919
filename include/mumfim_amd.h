/*
 * mumfim_amd: batched explicit fiber-network RVE solver for MI355X (gfx950).
 *
 * C ABI of the drop-in boundary.  It replaces the batched RVE analysis that
 * the FE^2 micro scale calls once per macro Newton iteration:
 *
 *   reference interface                                     replaced by
 *   ------------------------------------------------------  ---------------------------
 *   BatchedFiberRVEAnalysisExplicit(networks, strategies)   mf_batch_create
 *     (src/mumfim/microscale/BatchedFiberRVEAnalysisExplicit.h:80-256)
 *   BatchedRVEAnalysis::run(F, sigma, update_coords)        mf_batch_run / mf_batch_run_device
 *     (BatchedRVEAnalysis.h:29-31, BatchedFiberRVEAnalysisExplicit.h:257-320)
 *   BatchedRVEAnalysis::computeMaterialStiffness(C)         mf_batch_tangent
 *     (BatchedRVEAnalysis.h:36-37, BatchedFiberRVEAnalysisExplicit.h:321-405)
 *   compute3DOrientationTensor(omega)                       mf_batch_orientation_3d
 *     (BatchedRVEAnalysis.h:38-43, batched_analysis/BatchedComputeOrientationTensor.h:130-190)
 *   compute2DOrientationTensor(normal, omega)               mf_batch_orientation_2d
 *     (BatchedRVEAnalysis.h:44-50, BatchedComputeOrientationTensor.h:56-128,191-208)
 *   std::exit(EXIT_FAILURE) on size mismatch                 negative return code + mf_last_error
 *     (BatchedFiberRVEAnalysisExplicit.h:85-90,324-330)
 *
 * Array conventions (all host arrays are C row-major, caller owned):
 *   F      n x 9  doubles, F[r*9 + 3*i + j] = F_ij (incremental w.r.t. the
 *          current configuration, as the macro scale sends it); n == 1
 *          broadcasts one F to every RVE (BatchedFiberRVEAnalysisExplicitBase.h:36).
 *   sigma  n x 6  doubles, Voigt order [xx, yy, zz, yz, xz, xy].
 *   C      n x 36 doubles, C[r*36 + 6*j + i] = d sigma_j / d eps_i.
 *   status n int32 per RVE (MF_STATUS_*), may be NULL.
 *
 * Threading: a batch handle is not thread safe; every call on a handle is
 * serialised on the handle's own HIP stream and blocks until results are
 * written unless the name says _async.
 */
#ifndef MUMFIM_AMD_H
#define MUMFIM_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---------------------------------------------------- */
#define MF_OK 0
#define MF_ERR_INVALID (-1)     /* bad argument or size mismatch            */
#define MF_ERR_HIP (-2)         /* HIP runtime error (incl. no GPU)         */
#define MF_ERR_NOMEM (-3)       /* host or device allocation failed         */
#define MF_ERR_STATE (-4)       /* call sequence error (e.g. tangent first) */
#define MF_ERR_UNSUPPORTED (-5) /* network too large for any kernel variant */
#define MF_ERR_IO (-6)          /* file could not be read / parsed          */

/* ---- per-RVE status ---------------------------------------------------- */
#define MF_STATUS_OK 0
#define MF_STATUS_NAN 1          /* NaN/Inf in residual or stress            */
#define MF_STATUS_STEP_CAP 2     /* relax mode hit max_steps before tol      */
#define MF_STATUS_FALLBACK 3     /* surrogate rejected, explicit solve used  */
#define MF_STATUS_ENERGY 4       /* energy balance check failed (diagnostics) */

/* ---- fiber reactions (FiberReactions.h:15-43, FiberNetworkIO_impl.h:17-39) */
#define MF_REACTION_LINEAR 0
#define MF_REACTION_NONLINEAR 1

typedef struct mf_reaction {
  int32_t type;             /* MF_REACTION_*                                  */
  int32_t pad_;
  double radius;            /* fiber area = pi * radius^2                     */
  double E;                 /* Young's modulus                                */
  double density;           /* rho                                           */
  double B;                 /* nonlinear: exponential coefficient             */
  double length_ratio_trns; /* nonlinear: tension transition lambda_t        */
  double cmp_ratio_trns;    /* nonlinear: compression transition (1.0)       */
} mf_reaction;

/* One fiber network (one RVE).  Node coordinates live in the RVE cube
 * [-0.5, 0.5]^3.  `boundary` lists the prescribed nodes in the order the
 * reference collects them (RVE_impl.h:6-20).  fiber_reaction may be NULL
 * (every fiber uses reaction 0). */
typedef struct mf_network {
  int32_t n_nodes;
  int32_t n_fibers;
  int32_t n_boundary;
  int32_t n_reactions;
  const double *coords;          /* n_nodes x 3                               */
  const int32_t *fibers;         /* n_fibers x 2 node indices                 */
  const int32_t *fiber_reaction; /* n_fibers reaction ids, or NULL            */
  const int32_t *boundary;       /* n_boundary node indices                   */
  const mf_reaction *reactions;  /* n_reactions                               */
  double scale_factor;           /* multiscale stress scale (FiberNetwork.cc:67) */
} mf_network;

/* ---- solver configuration ---------------------------------------------- */
/* MF_MODE_RELAX reproduces the reference loop: affine guess on all nodes,
 *   damped central difference until the internal-force residual (refreshed
 *   every dt_update_freq steps) drops to `tolerance`
 *   (BatchedFiberRVEAnalysisExplicitTeamOuterLoop.h:215-372).
 * MF_MODE_RAMP runs exactly n_steps steps from the committed state with the
 *   boundary displacement ramped linearly from 0 to (F - I) x_b. */
#define MF_MODE_RELAX 0
#define MF_MODE_RAMP 1
#define MF_PREC_FP64 0   /* reference arithmetic, FP64 everywhere           */
#define MF_PREC_FP32C 1  /* FP32 force/update, compensated positions, FP64 stress */
/* MF_REACT_BATCHED_REFERENCE: every fiber uses reaction 0's E, A, rho and the
 *   linear law, as the reference batched path does
 *   (BatchedFiberRVEAnalysisExplicit.h:202-218, ...Base.h:155-172).
 * MF_REACT_PER_FIBER: each fiber uses its own reaction (linear or nonlinear
 *   toe law, FiberReactions.cc:5-71) and its own rho*A in the nodal mass. */
#define MF_REACT_BATCHED_REFERENCE 0
#define MF_REACT_PER_FIBER 1
#define MF_TANGENT_FORWARD 0     /* 6 solves, (sigma(h) - sigma)/h          */
#define MF_TANGENT_CENTRAL 1     /* 12 solves, (sigma(+h) - sigma(-h))/(2h) */
/* Ramp mode, tangent_precision MF_PREC_FP32C: the h -> 0 limit of both
 * schemes, the derivative of each perturbed solve along its perturbation
 * (dF = E_i F, E_i the unit Voigt strain of direction i) integrated with the
 * solve itself (tangent-linear), 6 solves per RVE and no difference of
 * nearly equal stresses; fd_step is not used. */
#define MF_TANGENT_LINEAR 2

typedef struct mf_solver_params {
  int32_t mode;                 /* MF_MODE_*                                  */
  int32_t precision;            /* MF_PREC_* of run()                          */
  int32_t reaction_policy;      /* MF_REACT_*                                  */
  int32_t tangent_scheme;       /* MF_TANGENT_*                                */
  int32_t tangent_precision;    /* MF_PREC_* of the perturbed solves           */
  int32_t n_steps;              /* ramp mode: steps per solve                 */
  int32_t max_steps;            /* relax mode: step cap (status STEP_CAP)     */
  int32_t dt_update_freq;       /* steps between dt / residual refresh (100)  */
  int32_t block_size;           /* threads per solve; 0 = planner's choice;
                                 * binds each precision that has kernels of
                                 * that size (FP32c 64-1024, FP64 256, 512,
                                 * 1024)                                      */
  int32_t amplitude;            /* ramp mode: MF_AMP_* boundary amplitude     */
  double visc_damp_coeff;       /* eta, mass-proportional damping             */
  double crit_time_scale_factor;/* dt = factor * min(l) / c (0.8)             */
  double tolerance;             /* relax mode residual tolerance (1e-6)      */
  double fd_step;               /* tangent strain step h                      */
  /* Energy diagnostics of the non-batched explicit solver
   * (FiberRVEAnalysisExplicit_impl.h:178-208, 1101-1138), FP64 only:
   * W_int, W_ext, W_damp, W_kin per solve; with energy_check_eps > 0 a solve
   * whose |W_kin + W_int + W_damp - W_ext| exceeds eps * max|W| at a check
   * (every dt_update_freq steps after step 10, and at the end) stops with
   * status MF_STATUS_ENERGY. */
  int32_t energy_diagnostics;   /* 1: track energies                           */
  int32_t fp32_kernel;          /* MF_FP32_KERNEL_*: FP32c kernel family      */
  double energy_check_eps;      /* 0: no check                                 */
  /* Ramp mode, FP32c register kernels: each solve of run() is split into
   * this many step chunks, queued all first chunks before any second one,
   * so the end of a launch drains in short pieces (results are bitwise the
   * same).  0 = planner's choice from the batch size and residency, 1 = no
   * chunks, up to 16. */
  int32_t chunks;
  int32_t pad_;
} mf_solver_params;

/* FP32c kernel families.  Node-centric kernels gather each free node's
 * fibers (every fiber is evaluated from both ends); fiber-parallel kernels
 * evaluate each fiber once and let nodes sum the published end forces (fewer
 * arithmetic operations, more LDS traffic; RVEs up to ~1800 nodes).  WAVE
 * restricts the node-centric family to its one-RVE-per-wave shapes (each
 * wave solves its own RVE of up to ~180 nodes with no workgroup barrier;
 * block_size 64 W packs W RVEs per workgroup).  AUTO lets the planner and
 * autotuner choose per RVE. */
#define MF_FP32_KERNEL_AUTO 0
#define MF_FP32_KERNEL_NODE 1
#define MF_FP32_KERNEL_FIBER 2
#define MF_FP32_KERNEL_WAVE 3

/* Ramp-mode boundary amplitude alpha(xi), xi = k / n_steps. */
#define MF_AMP_LINEAR 0         /* alpha = xi                                */
#define MF_AMP_SMOOTHSTEP 1     /* xi^3 (10 - 15 xi + 6 xi^2), SmoothAmp,
                                   ExplicitAmplitude.h:26-62                 */

typedef struct mf_batch mf_batch;

/* Defaults reproduce the reference batched solver: relax mode, FP64,
 * batched-reference reactions, forward tangent with h = 1e-5, eta = 100,
 * factor 0.8, tol 1e-6, refresh every 100 steps. */
void mf_default_params(mf_solver_params *p);

const char *mf_last_error(void);
int mf_version(void);

/* Build the packed device batch (boundary-last reorder, nodal masses,
 * incidence lists) and upload it.  Mirrors the reference constructor. */
int mf_batch_create(const mf_network *nets, int32_t n_rve,
                    const mf_solver_params *params, mf_batch **out);
void mf_batch_destroy(mf_batch *b);
int mf_batch_num_rves(const mf_batch *b);
int mf_batch_set_params(mf_batch *b, const mf_solver_params *params);
int mf_batch_get_params(const mf_batch *b, mf_solver_params *params);
/* Per-RVE viscous damping coefficients (n_rve doubles), as the reference
 * keeps one per RVE from its solution strategy
 * (BatchedFiberRVEAnalysisExplicit.h:213-218); NULL restores
 * params.visc_damp_coeff for every RVE. */
int mf_batch_set_damping(mf_batch *b, const double *eta);

/* run(F, sigma, update_coords): host buffers. */
int mf_batch_run(mf_batch *b, const double *F, int32_t nF, double *sigma,
                 int32_t update_coords, int32_t *status);
/* computeMaterialStiffness(C) at the state of the last update run. */
int mf_batch_tangent(mf_batch *b, double *C, int32_t *status);
/* run + tangent with every solve (1 + 6 or 1 + 12 per RVE) in one launch
 * (ramp mode) or in the reference's sequence (relax mode). */
int mf_batch_run_with_tangent(mf_batch *b, const double *F, int32_t nF,
                              double *sigma, double *C, int32_t *status);

/* Device-pointer variants: F_dev (nF x 9), sigma_dev (n x 6), status_dev (n)
 * live in HBM; work is queued on the batch stream and not waited for. */
int mf_batch_run_device_async(mf_batch *b, const double *F_dev, int32_t nF,
                              double *sigma_dev, int32_t update_coords,
                              int32_t *status_dev);
/* Allocation-free run.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It is
 * mf_batch_run_device_async with the solves' scratch (F staging for a
 * broadcast F, per-solve step counters, statuses when status_dev is NULL,
 * and the global state of RVEs beyond the LDS kernels) taken from a
 * library-wide device array of MF_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends; it calls no allocating routine.  Scratch needed:
 * align256(72 n) + 2 align256(4 n) + 8 * (global-state doubles of the batch's
 * plan, 96 bytes per node of each RVE that runs on the global-state kernel).
 * A batch needing more returns MF_ERR_INVALID.  Energy diagnostics are
 * rejected (MF_ERR_UNSUPPORTED).  Nothing in the scratch needs clearing: the
 * kernels write every word they read.  All calls on one GPU share its array,
 * so calls that overlap on one GPU (two batches on different streams, or
 * threads) race: serialise them. */
#define MF_NOALLOC_SCRATCH_BYTES (64 << 20)
int mf_batch_run_device_noalloc(mf_batch *b, const double *F_dev, int32_t nF,
                                double *sigma_dev, int32_t update_coords,
                                int32_t *status_dev);
int mf_batch_synchronize(mf_batch *b);
/* HIP stream (hipStream_t) the batch queues its work on. */
void *mf_batch_stream(mf_batch *b);
/* Kernel time (ms, HIP events on the batch stream) of the explicit-solve
 * launches of the last run, and how many launches that was. */
int mf_batch_last_solve_ms(mf_batch *b, double *ms, int32_t *n_launches);
/* ---- batching layer (no reference counterpart) -------------------------- */
/* Autotuner: RVEs are bucketed by reactions, register eligibility, lanes
 * needed (wavefronts) and node count (units of 256); for every bucket each
 * kernel variant that fits (threads per solve, nodes per thread, incidence
 * lists in registers / LDS / HBM, one or two position buffers) is timed on
 * ramp solves of `steps` steps of the bucket's RVEs, and the fastest is kept
 * for that bucket.  The tuned buckets are kept only if the whole batch then
 * runs faster than with the plan they replace.  For ramp-mode FP32c batches
 * the step chunks per solve of the register-kernel groups (the planner's
 * count, or 1-4) are then timed on the batch's own run path and the fastest
 * is kept (params.chunks, when set, still forces the count; the chunk
 * choice is not cached).  With cache_path, buckets found in the file are
 * taken from it and newly timed ones are appended (text: prec lin regok
 * lanes/64 nodes/256 block npt flags us_per_solve).  n_timed (may be NULL)
 * counts the timed kernel-variant candidates.  The batch state is left
 * untouched. */
int mf_batch_autotune(mf_batch *b, int32_t precision, int32_t steps,
                      const char *cache_path, int32_t *n_timed);
/* hipGraph mode for mf_batch_run_device_async: the launch sequence of a run
 * (volume update, F broadcast, every launch group on the batch and auxiliary
 * streams, state commits) is captured once per (buffers, flags, state
 * parity) and replayed afterwards. */
int mf_batch_set_graph(mf_batch *b, int32_t enable);

/* Energies of the last run() (energy_diagnostics on): n x 4 doubles
 * W_int, W_ext, W_damp, W_kin per RVE. */
int mf_batch_last_energies(mf_batch *b, double *W);

/* Orientation tensors of the current configuration. omega: n x 9. */
int mf_batch_orientation_3d(mf_batch *b, double *omega);
int mf_batch_orientation_2d(mf_batch *b, const double *normal, int32_t n_normal,
                            double *omega);

/* Persistent state, in each network's input node order:
 * x (current coordinates) and u (displacement), n_nodes x 3 per RVE,
 * concatenated over RVEs; volume per RVE. */
int64_t mf_batch_total_nodes(const mf_batch *b);
int mf_batch_get_state(mf_batch *b, double *x, double *u, double *volume);
int mf_batch_set_state(mf_batch *b, const double *x, const double *u,
                       const double *volume);

/* Per-RVE sizes, 4 ints per RVE: nodes, free nodes, boundary nodes, fibers. */
int mf_batch_rve_sizes(const mf_batch *b, int32_t *sizes);

/* Kernel plan of the batch: for each RVE the variant chosen by the planner,
 * 4 ints per RVE: threads per solve, free-node slots per thread, flags, LDS
 * bytes.  Flags: 1 incidence lists in LDS, 2 double-buffered positions,
 * 4 incidence lists in registers, 8 linear-only, 16 energy diagnostics,
 * 32 one linear and one nonlinear reaction, 64 fiber-parallel, 128
 * fiber-streamed (bits 8-11: fibers per thread), 4096 one RVE per wave
 * (the first entry then counts the workgroup's threads, 64 per RVE), 8192
 * FP64 with the state in global memory (RVEs no other variant holds, in
 * either precision), 16384
 * FP64 evaluating each fiber once (bitwise the node-centric FP64 result);
 * bits 24-28: step chunks per solve of run() when more than one
 * (params.chunks).
 * Returns the number of distinct launch groups. */
int mf_batch_plan(mf_batch *b, int32_t *plan);

/* ---- device memory (no reference counterpart) ----------------------------- */
/* Plain HBM buffers for the *_device entry points, allocated by the same HIP
 * runtime the library runs on. */
int mf_device_alloc(void **ptr, int64_t bytes);
int mf_device_free(void *ptr);
int mf_memcpy_to_device(void *dst, const void *src, int64_t bytes);
int mf_memcpy_to_host(void *dst, const void *src, int64_t bytes);
int mf_device_synchronize(void);

/* ---- diagnostics (no reference counterpart) ------------------------------ */
/* Measured ceilings for the roofline: LDS read bandwidth (ds_read_b128 over
 * every CU) and HBM copy bandwidth (read + write bytes), in bytes/s. */
int mf_measure_lds_bandwidth(double *bytes_per_s);
int mf_measure_hbm_bandwidth(int64_t bytes, double *bytes_per_s);

/* ---- synthetic networks (the seeded generator) ------------------------- */
typedef struct mf_gen_params {
  uint64_t seed;
  int32_t n_nodes;            /* nodes uniform in [-0.5, 0.5]^3              */
  int32_t k_neighbors;        /* fibers = deduplicated k-NN graph edges      */
  double boundary_band;       /* boundary: 0.5 - max|X_i| <= band            */
  double radius, E, density;  /* reaction parameters (both reactions)        */
  double nonlinear_fraction;  /* share of fibers on the nonlinear reaction   */
  double B, length_ratio_trns;/* nonlinear reaction parameters               */
} mf_gen_params;

void mf_default_gen_params(mf_gen_params *g);
/* Generates one network.  Call with NULL buffers to obtain the counts; the
 * second call fills coords (n_nodes*3), fibers (n_fibers*2), fiber_reaction
 * (n_fibers), boundary (n_boundary), reactions (2 entries). */
int mf_generate_network(const mf_gen_params *g, int32_t *n_nodes,
                        int32_t *n_fibers, int32_t *n_boundary, double *coords,
                        int32_t *fibers, int32_t *fiber_reaction,
                        int32_t *boundary, mf_reaction *reactions);
/* Generates n_rve networks (seed of RVE i = g->seed + i, node count
 * n_nodes[i] or g->n_nodes when n_nodes is NULL) and creates a batch. */
int mf_batch_create_generated(const mf_gen_params *g, int32_t n_rve,
                              const int32_t *n_nodes,
                              const mf_solver_params *params, mf_batch **out);
/* Uniform doubles in [0,1) from the SplitMix64 stream (seed, stream). */
int mf_uniform(uint64_t seed, uint64_t stream, int64_t n, double *out);

/* ---- reference text formats (FiberNetworkIO.cc:47-106, FiberNetworkIO_impl.h:17-39,68-87) */
/* Reads "<nn> <ne> <np>\n x y z ...\n n0 n1 ..." and the optional
 * "<path>.params" file.  Boundary nodes are those within 1e-8 of a face of
 * [-0.5, 0.5]^3 (RVE.h:86-103).  Two-call protocol as mf_generate_network;
 * reactions holds up to max_reactions entries. */
int mf_load_network(const char *path, int32_t *n_nodes, int32_t *n_fibers,
                    int32_t *n_boundary, int32_t *n_reactions,
                    int32_t max_reactions, double *coords, int32_t *fibers,
                    int32_t *fiber_reaction, int32_t *boundary,
                    mf_reaction *reactions);

/* ---- FE^2 message adapter (MultiscaleRVEAnalysis.cc:197-305) -------------
 * The micro side of the multiscale loop.  Per-RVE records as the macro scale
 * sends them (MultiscaleMicroFOParams.h:19-120), row-major, one per RVE:
 *   hdrs      7 int32  micro_fo_header: RVE_TYPE, RVE_DIR_TYPE, FIELD_ORDER,
 *                      ELEMENT_TYPE, GAUSS_ID, COMPUTE_ORIENTATION_3D/2D
 *   prms      8 double micro_fo_params: FIBER_RADIUS, VOLUME_FRACTION,
 *                      YOUNGS_MODULUS, NONLINEAR_PARAM, LINEAR_TRANSITION,
 *                      ORIENTATION_AXIS_X/Y/Z
 *   slvr      8 double micro_fo_solver: ..., VISCOUS_DAMPING_FACTOR (5),
 *                      CRITICAL_TIME_SCALE_FACTOR (6), ...
 *   slvr_int 10 int32  micro_fo_int_solver: MICRO_SOLVER_TYPE (0), ...
 *   data      9 double micro_fo_data: deformation gradient, row-major
 *   results  45 double micro_fo_result: sigma (6), 0 (3), C (36)
 *   step     18 double micro_fo_step_result: Omega_3D (9), Omega_2D (9)
 * The networks themselves are the caller's (the reference picks a file per
 * RVE from its network library); each gets scale factor
 * mf_fo_scale_conversion(net, FIBER_RADIUS, VOLUME_FRACTION) and damping
 * VISCOUS_DAMPING_FACTOR.  CRITICAL_TIME_SCALE_FACTOR is not used: the
 * reference's batched loop fixes it at 0.8
 * (BatchedFiberRVEAnalysisExplicitTeamOuterLoop.h:312). */
#define MF_FO_HEADER_FIELDS 7
#define MF_FO_PARAM_FIELDS 8
#define MF_FO_SOLVER_FIELDS 8
#define MF_FO_SOLVER_INT_FIELDS 10
#define MF_FO_DATA 9
#define MF_FO_RESULT 45
#define MF_FO_STEP_RESULT 18
#define MF_FO_FIBER_ONLY 1          /* MicroscaleType::FIBER_ONLY          */
#define MF_FO_ISOTROPIC_NEOHOOKEAN 3 /* MicroscaleType::ISOTROPIC_NEOHOOKEAN */
#define MF_FO_SOLVER_EXPLICIT 1     /* SolverType::Explicit                */

/* calcScaleConversion (MultiscaleCoupling.cc:7-22): 1 / rve_dim^2 with
 * rve_dim = sqrt(total fiber length * pi r^2 / volume fraction). */
int mf_fo_scale_conversion(const mf_network *net, double fiber_radius,
                           double volume_fraction, double *scale);

typedef struct mf_fo mf_fo;
/* updateCoupling (MultiscaleRVEAnalysis.cc:102-196) for n RVEs; params may
 * be NULL (reference defaults). */
int mf_fo_create(const mf_network *nets, int32_t n, const int32_t *hdrs,
                 const double *prms, const double *slvr,
                 const int32_t *slvr_int, const mf_solver_params *params,
                 mf_fo **out);
void mf_fo_destroy(mf_fo *f);
/* The batch behind the adapter (owned by it). */
mf_batch *mf_fo_batch(mf_fo *f);
/* One macro iteration: run(F) with update, then the tangent
 * (MultiscaleRVEAnalysis.cc:233-288).  data n x 9, results n x 45. */
int mf_fo_run(mf_fo *f, const double *data, double *results, int32_t *status);
/* End of a macro step: orientation tensors with the normals of
 * ORIENTATION_AXIS_X/Y/Z (MultiscaleCoupling.h:32-52).  step n x 18. */
int mf_fo_step_results(mf_fo *f, double *step_results);

/* ---- learned surrogate (no reference counterpart) -------------------------
 * An ensemble of MLPs 10 -> 256 -> 256 -> 256 -> 6 with SiLU activations.
 * Weights and activations are BF16 on the matrix cores
 * (v_mfma_f32_16x16x32_bf16, FP32 accumulate); biases, master weights and
 * Adam state are FP32.  Inputs x (10 per sample): the strain
 * e = [e_xx, e_yy, e_zz, g_yz, g_xz, g_xy] of sym(F) - I (engineering shear
 * g = 2 e_ij), node count, fibers per unit volume, mean fiber length,
 * nonlinear-fiber fraction.  Outputs y (6): Cauchy stress, Voigt order. */
#define MF_MLP_IN 10
#define MF_MLP_HIDDEN 256
#define MF_MLP_OUT 6
#define MF_MLP_MAX_MEMBERS 8
#define MF_MLP_BATCH_ALIGN 256 /* training mini-batches are multiples    */

typedef struct mf_mlp mf_mlp;
/* n_members independent networks (1..MF_MLP_MAX_MEMBERS), member k
 * initialised from SplitMix64 stream (seed, 1000 + k). */
int mf_mlp_create(int32_t n_members, uint64_t seed, mf_mlp **out);
void mf_mlp_destroy(mf_mlp *m);
/* Parameters per member, packed W1,b1,W2,b2,W3,b3,W4,b4 (W row-major
 * out x in): 135942. */
int32_t mf_mlp_num_params(void);
int32_t mf_mlp_num_members(const mf_mlp *m);
/* Normalisation: x_n = (x - in_mean) * (1/in_std), y = y_n * out_std +
 * out_mean (FP32; a std <= 0 counts as 1). */
int mf_mlp_set_normalization(mf_mlp *m, const float *in_mean,
                             const float *in_std, const float *out_mean,
                             const float *out_std);
int mf_mlp_get_normalization(const mf_mlp *m, float *in_mean, float *in_std,
                             float *out_mean, float *out_std);
/* FP32 master weights of one member (mf_mlp_num_params floats).  Setting
 * them resets that member's Adam moments. */
int mf_mlp_get_weights(mf_mlp *m, int32_t member, float *w);
int mf_mlp_set_weights(mf_mlp *m, int32_t member, const float *w);

/* Training data and mini-batches live in HBM: X (n_data x 10), Y
 * (n_data x 6), FP32.  idx_dev holds n_members x batch row indices (member k
 * trains on rows idx[k*batch ...]); NULL means rows 0..batch-1 for every
 * member.  batch is a positive multiple of MF_MLP_BATCH_ALIGN.  The loss is
 * the mean squared error of the normalised outputs, averaged over members;
 * passing loss != NULL waits for the step. */
int mf_mlp_train_step_device(mf_mlp *m, const float *X_dev, const float *Y_dev,
                             int64_t n_data, const int32_t *idx_dev,
                             int32_t batch, float lr, float *loss);

typedef struct mf_train_params {
  int32_t epochs;
  int32_t batch;         /* multiple of MF_MLP_BATCH_ALIGN                */
  uint64_t seed;         /* shuffling: stream (seed, epoch*members+k)     */
  float lr, lr_final;    /* cosine decay from lr to lr_final              */
  float beta1, beta2, eps;
} mf_train_params;
void mf_default_train_params(mf_train_params *tp);
/* Adam epochs over (X, Y) with a fresh permutation per member and epoch;
 * epoch_loss (epochs x n_members, may be NULL) gets each epoch's mean loss. */
int mf_mlp_fit_device(mf_mlp *m, const float *X_dev, const float *Y_dev,
                      int64_t n, const mf_train_params *tp, float *epoch_loss);
/* Gradient of the loss w.r.t. one member's packed parameters on rows
 * 0..n-1 (n a multiple of MF_MLP_BATCH_ALIGN), no update. */
int mf_mlp_gradient_device(mf_mlp *m, int32_t member, const float *X_dev,
                           const float *Y_dev, int32_t n, float *grad_host,
                           float *loss);

/* Inference on device buffers.  member = -1 averages the ensemble, k >= 0
 * runs member k alone.  Y_dev: n x 6.  J_dev (may be NULL): n x 36,
 * J[s*36 + 6*i + j] = d y_i / d e_j, the input Jacobian w.r.t. the six
 * strain inputs from the tangent-linear pass fused into the same kernel.
 * spread_dev (may be NULL): n x 6 standard deviation of y over members. */
int mf_mlp_predict_device(mf_mlp *m, int32_t member, const float *X_dev,
                          int64_t n, float *Y_dev, float *J_dev,
                          float *spread_dev);
/* Kernel time (ms, HIP events) of the last predict. */
int mf_mlp_last_predict_ms(const mf_mlp *m, double *ms);
void *mf_mlp_stream(mf_mlp *m);

/* ---- surrogate path with explicit fallback (no reference counterpart) ----
 * Error estimator: a sample falls back to the explicit solver when any input
 * lies outside [in_lo - margin*(in_hi-in_lo), in_hi + margin*(in_hi-in_lo)],
 * or when kappa * max_i spread_i / max(max_i |y_i|, sigma_floor) > tol. */
typedef struct mf_estimator {
  float in_lo[MF_MLP_IN], in_hi[MF_MLP_IN];
  float margin, kappa, sigma_floor, tol;
} mf_estimator;
int mf_mlp_set_estimator(mf_mlp *m, const mf_estimator *est);
int mf_mlp_get_estimator(const mf_mlp *m, mf_estimator *est);

/* The 10 surrogate inputs of every RVE of a batch for deformation F
 * (n x 9, total deformation from the reference configuration): X n x 10. */
int mf_surrogate_inputs(mf_batch *b, const double *F, int32_t nF, float *X);

typedef struct mf_surrogate_stats {
  int64_t n, n_fallback, n_out_of_range;
  double mlp_ms;       /* input kernel + ensemble inference + estimator     */
  double fallback_ms;  /* explicit solves of the rejected RVEs              */
  double total_ms;     /* wall time of the call                             */
} mf_surrogate_stats;
/* Stress (and, when C != NULL, tangent) of every RVE of a ramp-mode batch
 * whose state is the reference configuration (no update run): the ensemble
 * predicts, the estimator rejects, rejected RVEs are solved by the explicit
 * solver (trial runs; tangent by the central scheme with fd_step), which is
 * also what the surrogate's outputs are compared against.  sigma n x 6,
 * C n x 36 in the layout of mf_batch_tangent (the surrogate's Jacobian
 * mapped to the perturbations F_i(h) F), status n (MF_STATUS_OK for accepted
 * predictions, MF_STATUS_FALLBACK or the solver's status otherwise). */
int mf_surrogate_run(mf_batch *b, mf_mlp *m, const double *F, int32_t nF,
                     double *sigma, double *C, int32_t *status,
                     mf_surrogate_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* MUMFIM_AMD_H */
