/*
 * olfhip — C ABI of the MI355X AC Newton-Raphson / batched N-1 library (libolfhip.so).
 *
 * The drop-in boundary for the reference's JNI natives: one JNI function per Java native marshals its arrays into
 * these entry points (olf-gpu_amd/jni/olfhip_jni.c). Plain pointers and sizes only. Every entry returns 0
 * on success or a negative OLFHIP_E_* code; olfhip_last_error() then holds the message (per thread). No exception
 * crosses this boundary. Caller-owned input arrays are borrowed read-only for the duration of the call; outputs go
 * into caller-provided buffers.
 *
 * Threading: a context is not re-entrant (one context per thread, as the reference serialises the device under one
 * lock, GpuBatchPrecomputer.java:65). Contexts and limit tables made on one structure may be used on different
 * threads at the same time (the structure is read-only once on the device; its reference count is atomic); a limit
 * table is used by one thread at a time.
 */
#ifndef OLFHIP_H
#define OLFHIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define OLFHIP_OK 0
#define OLFHIP_E_ARGUMENT -1   /* a contract violation (reference: IllegalArgumentException/IllegalStateException) */
#define OLFHIP_E_DEVICE -2     /* a HIP runtime failure (reference: RuntimeException "GPU batched Newton-Raphson failed") */
#define OLFHIP_E_MEMORY -3     /* a device or host allocation failed (the caller shrinks its chunk) */
#define OLFHIP_E_UNSUPPORTED -4 /* an input shape this build does not model (reference: refusal, handed back to OLF) */

/* per-scenario statuses: reference GpuAcSolverStatus (docs/gpu-vs-olf.md:196) */
#define OLFHIP_CONVERGED 0
#define OLFHIP_MAX_ITERATION_REACHED 1
#define OLFHIP_DIVERGED -1
#define OLFHIP_REJECTED -2             /* batch: a clamped (singular) pivot block, a bad solve at the budget, … */
#define OLFHIP_LINEAR_SOLVE_FAILED -3  /* the step guard, or (batch of one) a clamped pivot block */

/* row modes: reference GpuRowMode */
#define OLFHIP_ROW_FIXED_VARIABLE 0
#define OLFHIP_ROW_EQUATION 1

/* linear-solve modes (reference GpuLinearSolveMode; the reference's UBATCH/ROBUST are cuDSS modes) */
#define OLFHIP_LU_FP64 0            /* batched block LU in FP64 */
#define OLFHIP_LU_FP32_REFINED 1    /* FP32 factors and solves, one FP64-residual refinement step */
/* The robust single-scenario mode, what the reference retries rejected scenarios in (GpuLinearSolveMode.ROBUST = 6,
 * GpuLinearSolveMode.java:24-27; native/cuda/common/check.cuh:38-40; GpuNewtonRaphson.java:755-775,
 * GpuBatchPrecomputer.java:635-790): one scenario per tile, FP64 factors with static pivoting (a nearly singular
 * pivot block is perturbed by sqrt(eps) ||J|| instead of clamped) and every Newton step's linear solve finished by
 * restarted GMRES on the exact FP64 Jacobian, preconditioned by those factors, to the relative residual below. */
#define OLFHIP_LU_ROBUST 2
#define OLFHIP_ROBUST_KRYLOV_DIM 30   /* GMRES basis vectors per cycle */
#define OLFHIP_ROBUST_RESTARTS 4      /* cycles after the first */
#define OLFHIP_ROBUST_TOL 1e-10       /* ||J dx + F||_2 / ||F||_2 target */

/* state vector scaling: reference GpuStateVectorScalingMode */
#define OLFHIP_SCALING_NONE 0
#define OLFHIP_SCALING_LINE_SEARCH 1
#define OLFHIP_SCALING_MAX_VOLTAGE_CHANGE 2

typedef struct olfhip_structure olfhip_structure;
typedef struct olfhip_context olfhip_context;

/* Message of the last failed call on this thread ("" if none). Reference: the RuntimeException text thrown by the
 * JNI layer (jni/batch_solve_jni.cu:296-306). */
const char *olfhip_last_error(void);

/* Number of visible MI355X devices (0 without one); never fails. */
int olfhip_device_count(void);

/*
 * Shared structure of one extraction: stable rows, the superset pattern, the fill gather tables, the fill-reducing
 * order and the symbolic block LU plan. Replaces StructureNative.createStructure
 * (src/main/java/com/powsybl/openloadflow/ac/equations/vector/gpu/StructureNative.java:49-59, JNI
 * native/cuda/jni/structure_jni.cu:8) for the families this path models: closed branches (8 doubles
 * y g1 g2 b1 b2 ksi r1 a1 and 12 ints P1 Q1 P2 Q2 ph1 v1 ph2 v2 a1col r1col altQ1 altQ2 each, the reference's
 * packing, StructureNative.java:39-46) and bus shunts (2 doubles g b, 3 ints Prow Qrow vcol). Rows must be stable
 * bus rows (P on the bus's phi row 2i, Q/V on its V row 2i+1); a1col/r1col/altQ must be -1 (no controls):
 * anything else returns OLFHIP_E_UNSUPPORTED. n = 2 * buses.
 */
int olfhip_create_structure(int n, int n_closed_branch, const double *closed_branch_parameters,
                            const int *closed_branch_indices, int n_shunt, const double *shunt_parameters,
                            const int *shunt_indices, olfhip_structure **out);

/* Drops the caller's reference (contexts keep the structure alive). NULL is a no-op. Reference:
 * StructureNative.destroyStructure (StructureNative.java:67). */
void olfhip_destroy_structure(olfhip_structure *structure);

/* Statistics of the symbolic plan: [0] buses, [1] LU block slots per lane, [2] elimination levels,
 * [3] factor block multiply-adds per lane, [4] pattern nonzeros (scalar), [5] LU nonzeros (scalar).
 * Reference: BatchedLuNative.supernodalPlanStatistics / eliminationLevels (BatchedLuNative.java:75,82). */
int olfhip_structure_statistics(const olfhip_structure *structure, long long *stats6);

/* The bus elimination order of the plan (n/2 ints). Reference: BatchedLuNative.minimumDegreeOrder
 * (BatchedLuNative.java:72). */
int olfhip_structure_order(const olfhip_structure *structure, int *bus_order);

/* Scenarios that fit one chunk: 70 % of free device memory over the per-scenario cost, clamped to [1, max_cap].
 * Reference: StructureNative.recommendBatchSize (StructureNative.java:69-81). */
int olfhip_recommend_batch_size(const olfhip_structure *structure, int max_cap, int mode, int *out);

/* Reactive-limit entries of the local PV -> PQ outer loop (the reactive-limit part of
 * BatchSolveNative.createBatchContext, BatchSolveNative.java:41-48). Values in pu; voltage_row is the entry bus's
 * V row, whose P row is voltage_row - 1. */
typedef struct {
    int n_entries;
    const int *voltage_row;
    const double *min_q;
    const double *max_q;
    const double *load_q;
    const double *target_v;
} olfhip_reactive_limits;

/* Per-context tuning (the autotuner's choices; zeros take the defaults). */
typedef struct {
    int lanes_per_tile;   /* scenarios per workgroup tile: 1, 2, 4, 8, 16 or 32 (0: picked from n_scenarios) */
    int threads_per_tile; /* workgroup size: 256, 512 or 1024 (0: 512) */
} olfhip_tuning;

/* A batch context of n_scenarios slots on a structure: per-scenario device state and LU storage, sized once.
 * Reference: BatchSolveNative.createBatchContext (BatchSolveNative.java:41-48; native/cuda/jni/batch_solve_jni.cu:29-81).
 * tuning may be NULL. */
int olfhip_create_batch_context(olfhip_structure *structure, int n_scenarios, int mode,
                                const olfhip_reactive_limits *reactive_limits, const olfhip_tuning *tuning,
                                olfhip_context **out);

/* Frees the context's device memory. NULL is a no-op. Reference: BatchSolveNative.destroyBatchContext
 * (BatchSolveNative.java:51). */
void olfhip_destroy_batch_context(olfhip_context *context);

/* Distributed slack of a context: the participating buses (never the slack bus), their participation factors
 * (1 - their sum is the slack's retained share), each one's P-row target at the anchor the water-fill restarts
 * from and its room up (>= 0) and down (<= 0) from it, pu; the slack bus and its own P target. The reference takes
 * these in createBatchContext (BatchSolveNative.java:44-47: distributedSlackBusPhiRow, distributedSlackBusFactor,
 * slackBusPhiRow, slackBusTargetP, distributedSlackBusBaseTargetP / MaxDelta / MinDelta; factor mode
 * PROPORTIONAL, no margins); the check is native/cuda/outerloop/distributed_slack.cu:131-195,215-600. */
typedef struct {
    int n_buses;
    const int *bus;
    const double *factor;
    const double *base_target_p;
    const double *max_delta;
    const double *min_delta;
    int slack_bus;
    double slack_target_p;
} olfhip_distributed_slack;

/* Sets (or with NULL clears) the context's distributed slack; olfhip_solve_inputs.distributed_slack turns the
 * loop on per solve. */
int olfhip_set_distributed_slack(olfhip_context *context, const olfhip_distributed_slack *distributed_slack);

/* OLF's PerEquationTypeStoppingCriteria as per-row thresholds on the context, the reference's
 * NewtonSettingsNative.setBatchRowCriteria (NewtonSettingsNative.java:29-36; applied in
 * native/cuda/solver/convergence.cu:28-60 and newton_iteration.cu:296-300): fixed_row_max_mismatch [n] for a row
 * held as an identity (FIXED_VARIABLE) and equation_row_max_mismatch [n] for an equation row, in the system's per
 * unit (e.g. maxActivePowerMismatch / base MVA on a P row, maxVoltageMismatch on a V row). While set, a scenario
 * converges when every row's |F| is below its mode's threshold, instead of ||F||inf < tol or the 2-norm criterion.
 * Both NULL clears them. Thresholds must be > 0. */
int olfhip_set_row_criteria(olfhip_context *context, const double *fixed_row_max_mismatch,
                            const double *equation_row_max_mismatch);

/* Sparse start-state seeds of the context's next solves (NewtonSettingsNative.setBatchX0Seeds,
 * native/cuda/jni/newton_settings_jni.cu:43-63): each flat_index = scenario * n + row sets that scenario's start value
 * of that row over the tiled base start (or its per-scenario start), until the next call; n_seeds 0 clears them.
 * Indices must be below capacity x n and values finite; a solve refuses seeds of scenarios beyond its n_scenarios. */
int olfhip_set_x0_seeds(olfhip_context *context, int n_seeds, const long long *flat_index, const double *value);

/* Inputs of one batched solve: the hot-path subset of BatchSolveNative.solveBatchContext
 * (BatchSolveNative.java:106-125), same meaning and order conventions:
 *   x0_base, target_base, base_row_mode: [n] shared by every scenario (tiled on the device);
 *   row_mode_entries: (scenario, row, mode) triples where a scenario's mode differs from the base (islanded buses
 *     frozen, ...), with target_overrides (scenario * n + row -> value) for the targets those rows hold;
 *   outaged_branches: [n_scenarios][max_outages] closed-branch ids, -1 padded (the closed-branch part of the
 *     disabled-element mask; at most 4 per scenario);
 *   x0_per_scenario: NULL or [n_scenarios][n] (the warm-start hook, BatchSolveNative.java:119);
 *   reactive_limits_seed_state: NULL or [n_scenarios][n_entries] reactive-limit types (0 none, 1 min, 2 max)
 *     of entries that start switched to PQ (the reference's reactiveLimitsSeedState);
 *   max_iter, tol (||F||inf), reactive_limit_tol, state_vector_scaling_mode, max_dv, max_dphi,
 *   max_outer_iterations (<= 0: 20), max_pq_pv_switch (< 0: 3), skip_reactive_limits;
 *   states_to_host: 0 keeps the states on the device (BatchSolveNative.setBatchStatesToHost), where
 *     olfhip_read_states reads them later. */
typedef struct {
    int n_scenarios;
    const double *x0_base;
    const double *target_base;
    const int *base_row_mode;
    int n_row_mode_entries;
    const int *row_mode_entries;
    int n_target_overrides;
    const long long *target_override_index;
    const double *target_override_value;
    int max_outages;
    const int *outaged_branches;
    const double *x0_per_scenario;
    const signed char *reactive_limits_seed_state;
    int max_iter;
    double tol;
    double reactive_limit_tol;
    int state_vector_scaling_mode;
    double max_dv;
    double max_dphi;
    int max_outer_iterations;
    int max_pq_pv_switch;
    int skip_reactive_limits;
    int states_to_host;
    /* NULL or a DEVICE pointer to [n_scenarios][n] start states (e.g. olfhip_mlp_predict_warm_start's output);
     * exclusive with x0_per_scenario */
    const double *x0_per_scenario_device;
    /* the distributed-slack outer loop (needs olfhip_set_distributed_slack), run before the reactive limits; its
     * gate slackBusPMaxMismatch in pu (slackDistTol, BatchSolveNative.java:117); reject_slack_distribution_failure:
     * a scenario whose distribution left more than 1e-5 pu unplaced ends REJECTED (the reference's
     * rejectSlackDistributionFailure, slackDistributionFailureBehavior other than LEAVE_ON_SLACK_BUS) */
    int distributed_slack;
    double slack_dist_tol;
    int reject_slack_distribution_failure;
    /* LINE_SEARCH scaling: backtracks per test (<= 0: 10) and step fold (<= 1: 4/3), the reference's
     * NewtonSettingsNative.setBatchLineSearch (NewtonSettingsNative.java:56) */
    int line_search_max_iteration;
    double line_search_step_fold;
    /* > 0: OLF's stopping criterion ||F||_2 < sqrt(conv_eps_per_eq^2 * active rows) instead of ||F||inf < tol
     * (solveBatchContext's convEpsPerEq, BatchSolveNative.java:117); active_row_counts: NULL (every lane n) or
     * [n_scenarios] equations of each scenario's system (NewtonSettingsNative.setBatchActiveRowCounts) */
    double conv_eps_per_eq;
    const int *active_row_counts;
    /* 1: reactive_limits_seed_state is ONE row [n_entries] that every scenario starts from (the base case's types,
     * the usual N-1 seed), uploaded once instead of n_scenarios times */
    int reactive_limits_seed_shared;
    /* > 0: the linear-solve guard's bound on ||J dx + F||inf / ||F||inf (checked at the 2nd iteration of a scenario's
     * first solve, and on the solve before its budget); <= 0: the reference's LIN_SOLVE_FAIL_RATIO 1e-2
     * (native/cuda/solver/iterative_refinement.cuh:23, newton_iteration.cu:612-652) */
    double linear_solve_fail_ratio;
    /* 0: the solve returns when every scenario of the bank has ended; k > 0: it returns as soon as k scenarios have
     * ended during the call, the others left solving between two Newton steps, for the caller to install new
     * scenarios into the slots handed back (olfhip_begin_batch_install) and call again, which goes on with those still
     * solving (solveBatchContext's pauseAtEnded, BatchSolveNative.java:106-125, newton_iteration.cu:488-493) */
    int pause_at_ended;
} olfhip_solve_inputs;
/* tol <= 0 with conv_eps_per_eq <= 0 and no per-row criterion: the fixed-iteration mode (the reference's tol <= 0,
 * native/cuda/solver/outer_loop_driver.cu:165-175, newton_iteration.cu:141-151,336-343): every scenario runs exactly
 * max_iter Newton steps with no convergence test, no linear-solve guard and no outer loop, and ends
 * MAX_ITERATION_REACHED, or DIVERGED when its final mismatch is not finite. */

/* Outputs (any may be NULL): results [n_scenarios][n+3] = (x, iterations of the last solve, ||F||inf, status) in
 * install order, the reference's layout (jni/batch_solve_jni.cu:308-337); status, iterations (last solve),
 * total_iterations (all solves of the outer loop) and norm per scenario. */
typedef struct {
    double *results;
    int *status;
    int *iterations;
    int *total_iterations;
    double *norm;
    /* the count of rows the solve wrote: the scenarios it handed back. Needed by a bank solve (after
     * olfhip_begin_batch_install) and a paused one (pause_at_ended > 0), whose rows are the scenarios that ended and
     * were not handed back before, in the order they were installed (their slots: olfhip_read_handed_back_slots);
     * otherwise every installed scenario, n_scenarios rows */
    int *n_handed_back;
} olfhip_solve_outputs;

/* Runs the batched Newton-Raphson with per-scenario retirement and the reactive-limit outer loop.
 * Reference: BatchSolveNative.solveBatchContext (BatchSolveNative.java:106-125; native/cuda/jni/batch_solve_jni.cu:84-393,
 * native/cuda/solver/outer_loop_driver.cu:80-410, native/cuda/solver/newton_iteration.cu:93-203). */
int olfhip_solve_batch_context(olfhip_context *context, const olfhip_solve_inputs *inputs,
                               olfhip_solve_outputs *outputs);

/* The bank: local scenario j of the next solve (its per-scenario inputs, row_mode_entries, target overrides, start
 * seeds, outages, seeds, active row counts) goes into bank slot bank_slots[j]; every other slot keeps what it holds,
 * a scenario still solving after a paused solve going on from where it stopped. Each slot appears at most once and is
 * below the context's capacity; n_slots 0 (or bank_slots NULL) chooses the whole bank, scenario j into slot j, with
 * every slot emptied first (the default install). The choice holds for the next solve, whose n_scenarios must be
 * n_slots. After a bank or paused solve the readbacks (olfhip_read_states, olfhip_read_outer_loop_state, the violation
 * screens) index scenarios by bank slot, n_scenarios being the capacity; a slot the solve did not hand back reads as
 * not converged. Reference: BatchSolveNative.beginBatchInstall (BatchSolveNative.java:53-59,
 * native/cuda/jni/batch_solve_jni.cu:444-463). */
int olfhip_begin_batch_install(olfhip_context *context, int n_slots, const int *bank_slots);

/* The bank slots of the scenarios the last solve handed back, in the order of its result rows: *n_slots of them
 * (capacity entries are enough). Reference: BatchSolveNative.readBatchHandedBackSlots (BatchSolveNative.java:98-100,
 * native/cuda/jni/batch_solve_jni.cu:428-442). */
int olfhip_read_handed_back_slots(const olfhip_context *context, int *n_slots, int *slots);

/* Device time of the last solve's Newton-step launches, measured with HIP events on the context's stream around each
 * launch (the step kernel, plus the few-microsecond running-tile count and its 4-byte copy), and the launch count.
 * One launch carries every tile to its end unless OLFHIP_STEPS_PER_LAUNCH=k caps it at k Newton steps per tile.
 * No reference counterpart: the reference's phase timers are print-only (OLF_GPU_PHASE_TIME,
 * native/cuda/jni/batch_solve_jni.cu:12-26). */
int olfhip_context_timing(const olfhip_context *context, double *step_ms, int *steps);

/* How the context's solves ran their graphs: a solve captures its install kernels and its Newton-step launch (with
 * the running-tile count and its copy) as two graphs and brings the context's instantiated graphs up to them, in
 * place when only kernel arguments changed (updates) or instantiating again when the sequence changed (the first
 * solve, or another set of install kernels: instantiations). OLFHIP_NO_GRAPH=1 launches eagerly. No reference
 * counterpart (the reference launches per level and per step). */
int olfhip_context_graph_statistics(const olfhip_context *context, int *instantiations, int *updates);

/* Row modes and targets ([n_scenarios][n], stable rows) and reactive-limit types ([n_scenarios][n_entries], may be
 * NULL) the outer loop left in each scenario after the last solve: what the reference writes back into OLF's buses
 * after a device solve (acceptDeviceSolution, GpuNewtonRaphson.java:789-824), and the seed state a later N-1 chunk
 * starts from. */
int olfhip_read_outer_loop_state(olfhip_context *context, int n_scenarios, int *row_mode, double *target,
                                 signed char *reactive_limit_types);

/* States of the last solve (states_to_host = 0): [n_scenarios][n] into x. Reference: the StateHandback readback of
 * states left on the device (BatchSolveNative.setBatchStatesToHost). */
int olfhip_read_states(olfhip_context *context, int n_scenarios, double *x);

/* Buses cut off from slack_bus once the listed closed branches are out (1 = islanded); returns the count in
 * *n_islanded. Reference: the islanding plan of GpuBatchedSecurityAnalysisSolver.solveBatchN1
 * (GpuBatchedSecurityAnalysisSolver.java:196-238) and batch/plan/IslandedBuses.java:50-60. */
int olfhip_islanded_buses(const olfhip_structure *structure, int slack_bus, int n_outaged, const int *outaged,
                          unsigned char *islanded, int *n_islanded);

/* Standalone batched LU on the structure's Jacobian pattern: factorizes the systems of lane_map and solves them.
 * values: [capacity][nnz] CSR values over olfhip_structure_pattern's pattern, scenario-major; rhs, solutions:
 * [capacity][n]. pivot_clamps: per lane_map entry, the diagonal blocks clamped as near-singular (may be NULL).
 * Reference: BatchedLuNative.factorizeAndSolve (BatchedLuNative.java:96-98; jni/batched_lu_jni.cu:309). */
int olfhip_lu_factorize_and_solve(olfhip_structure *structure, int mode, int capacity, const double *values,
                                  const double *rhs, int n_lanes, const int *lane_map, double *solutions,
                                  int *pivot_clamps);

/* The superset Jacobian pattern of the structure in stable-row CSR (rowptr [n+1], colidx [nnz]); pass NULL arrays
 * to read nnz only. */
int olfhip_structure_pattern(const olfhip_structure *structure, int *nnz, int *rowptr, int *colidx);

/* The level schedule of the structure's LU (host data, for tests and tools): which = 0 item records (4 ints: slot,
 * contribution begin, end, node or -1), 1 item buckets (first, end, thread split, diagonal flag), 2 bucket range of
 * each level, 3 backward records (diagonal slot, S begin, S end, node), 4 backward buckets, 5 backward bucket range of
 * each level, 6 item range of each level, 7 node range of each level, 8 diagonal item of each node, 9 contribution
 * records (4 ints), 10 nodes in level order, 11 the block-column structure S(k) of every node (the members of node
 * k's S set at [S begin, S end) of its backward record). values NULL: *count only. */
int olfhip_structure_schedule(const olfhip_structure *structure, int which, int *count, int *values);

/* Picks lanes_per_tile and threads_per_tile for this workload by timing full solves of `inputs` on the device for
 * every candidate tile shape (a project-written autotuner; the reference sizes its chunks from memory only,
 * StructureNative.java:69-81). */
int olfhip_autotune(olfhip_structure *structure, int mode, const olfhip_reactive_limits *reactive_limits,
                    const olfhip_solve_inputs *inputs, olfhip_tuning *best, double *best_ms);

/* ---- Limit violations of the scenarios the last solve of a context converged (reference
 * native/cuda/sa/olf_gpu_violations.cu:1-27,170-300 and its JNI ViolationScreenNative.uploadViolationTable /
 * screenHandedBack, olf_gpu_violations.cu:430-470,650-690; the decisions of GpuLimitScanner.java:20-262, which are OLF's
 * LimitViolationManager.detectViolations). A table lists, per monitored closed branch, its limit groups (a side and a
 * quantity) with their thresholds most severe first, and per bus a high and a low voltage limit (NaN: none). Values
 * are in per unit of the state: current |S|/V, active power P (|P| compared, the signed value reported), apparent
 * power |S|, voltage V. Within a group the first exceeded threshold is reported and the group stops there. Outaged
 * branches, islanded buses and branches with an islanded end are out of service. */
enum { OLFHIP_QUANTITY_CURRENT = 0, OLFHIP_QUANTITY_ACTIVE_POWER = 1, OLFHIP_QUANTITY_APPARENT_POWER = 2 };
enum { OLFHIP_VIOLATION_BRANCH = 0, OLFHIP_VIOLATION_HIGH_VOLTAGE = 1, OLFHIP_VIOLATION_LOW_VOLTAGE = 2,
       OLFHIP_VIOLATION_HIGH_VOLTAGE_ANGLE = 3, OLFHIP_VIOLATION_LOW_VOLTAGE_ANGLE = 4 };
typedef struct olfhip_limit_table olfhip_limit_table;
typedef struct olfhip_limit_table_desc {
    int n_branches;                 /* monitored branches (table order) */
    const int *branch;              /* [n_branches] closed-branch index */
    const int *group_start;         /* [n_branches + 1] groups of table branch b: [group_start[b], group_start[b+1]) */
    const int *group_side;          /* [n_groups] 1 or 2 */
    const int *group_quantity;      /* [n_groups] OLFHIP_QUANTITY_* */
    const int *limit_start;         /* [n_groups + 1] limits of group g: [limit_start[g], limit_start[g+1]) */
    const double *limit_threshold;  /* [n_limits] */
    int n_buses;                    /* buses 0 .. n_buses - 1 carry voltage limits */
    const double *bus_high;         /* [n_buses] pu, NaN: none */
    const double *bus_low;          /* [n_buses] pu, NaN: none */
    /* Voltage-angle limits (OLF's LfVoltageAngleLimit, checked by GpuLimitScanner.detectVoltageAngleViolations,
     * GpuLimitScanner.java:239-267): the angle difference to - from, in radians, between two buses both in service,
     * against a high and a low value (NaN: none); reported after the buses, high before low, per limit. The screen
     * does not cover them: every limit is decided on the host for every converged scenario. */
    int n_angle_limits;
    const int *angle_from;          /* [n_angle_limits] bus index */
    const int *angle_to;            /* [n_angle_limits] bus index */
    const double *angle_high;       /* [n_angle_limits] rad, NaN: none */
    const double *angle_low;        /* [n_angle_limits] rad, NaN: none */
} olfhip_limit_table_desc;
typedef struct olfhip_violation {
    int scenario;                   /* index in the last solve's scenarios */
    int kind;                       /* OLFHIP_VIOLATION_* */
    int element;                    /* table branch index, bus index, or voltage-angle limit index */
    int side;                       /* 1 or 2 (branch), 0 (bus, angle) */
    int quantity;                   /* OLFHIP_QUANTITY_* (branch) */
    int limit;                      /* limit index (branch), -1 (bus, angle) */
    double value;                   /* the value (signed active power; angle difference in rad) */
    double threshold;
} olfhip_violation;
int olfhip_create_limit_table(olfhip_structure *structure, const olfhip_limit_table_desc *desc,
                              olfhip_limit_table **out);
void olfhip_destroy_limit_table(olfhip_limit_table *table);
/* The device screen alone (candidates, may over-emit): counts [K] and records [K][capacity] for the K scenarios of the
 * last solve; record l is branch limit l, n_limits + 2 i / n_limits + 2 i + 1 the high / low voltage of bus i. A
 * count above capacity means the scenario's candidates overflowed. band: relative screen band (e.g. 1e-7). */
int olfhip_screen_violations(olfhip_context *context, olfhip_limit_table *table, double band, int capacity,
                             int *counts, int *records);
/* The scenarios of the context's last solve (n_scenarios must be its count) whose violation decisions a state
 * difference of relative size `band` could change: flags[k] = 1 when scenario k converged and a monitored quantity lies
 * within band x |threshold| of one of its group's thresholds (up to the first it exceeds beyond the band), or a bus
 * voltage within band x |limit| of its high or low limit unless the bus holds its voltage (PV, slack, islanded). A
 * warm-started answer is another Newton iterate inside the same stopping tolerance as the base-start answer (the
 * reference's semantics): flagged scenarios are the ones to solve again from the base start for lists equal to the
 * base start's. No reference counterpart (the reference has no warm start). */
int olfhip_ambiguous_scenarios(olfhip_context *context, olfhip_limit_table *table, double band, int n_scenarios,
                               int *flags);
/* The device screen alone, allocating nothing: the same candidates as olfhip_screen_violations (the record order of a
 * scenario may differ: emission order is not fixed) for the n_scenarios scenarios of the context's last solve
 * (n_scenarios must be that count), with its working memory (scenario flags, counts, records) in a per-device array
 * of OLFHIP_NO_ALLOC_SCRATCH_BYTES the library holds from its load to the process's end. No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops, harnesses that forbid hipMalloc). It
 * calls no hipMalloc or hipFree; 2 x ceil256(4 n_scenarios) + 4 n_scenarios x capacity bytes must fit the array,
 * else OLFHIP_E_ARGUMENT. Calls that overlap on one device share that array and race. */
#define OLFHIP_NO_ALLOC_SCRATCH_BYTES (32u << 20)
int olfhip_screen_violations_no_alloc(olfhip_context *context, olfhip_limit_table *table, double band,
                                      int n_scenarios, int capacity, int *counts, int *records);
/* Screen plus the exact host decisions: the violations of every converged scenario, ordered by scenario, then
 * branches in table order (a group's first exceeded limit), then buses (high before low), then voltage-angle limits
 * (high before low). Fails with
 * OLFHIP_E_ARGUMENT, *n_violations set to the count, when max_violations is too small. */
int olfhip_detect_violations(olfhip_context *context, olfhip_limit_table *table, double band, int capacity,
                             int max_violations, olfhip_violation *violations, int *n_violations);

/* ---- Learned warm start (no reference counterpart: the reference seeds scenarios from the base case or from
 * caller-given states only, BatchSolveNative.java:119). A 3-layer MLP, hidden width `hidden`, GELU, BF16 weights on
 * the matrix cores with FP32 accumulation and FP32 master weights (Adam). It works on each outage's neighbourhood:
 * the n_local buses nearest the outaged branch by the sum of |weight| (e.g. branch reactance) along paths, endpoints
 * first, -1 padded. Input: the neighbourhood's base-case angles relative to endpoint 1 and voltages minus 1
 * (2 n_local rows, normalised), the outaged branch (an embedding row, the one-hot input of the first layer) and
 * n_features endpoint features; output: 2 n_local rows [angle changes ; voltage changes], normalised. The Newton solve
 * it seeds alone decides the answer. */
typedef struct olfhip_mlp olfhip_mlp;

int olfhip_mlp_create(int n_bus, int n_branch, const int *bus1, const int *bus2, const double *weight, int n_local,
                      int n_features, int hidden, unsigned long long seed, olfhip_mlp **out);
void olfhip_mlp_destroy(olfhip_mlp *mlp);
/* the neighbourhoods olfhip_mlp_create builds, into buses [n_branch][n_local] (host computation, no device use):
 * multi-source Dijkstra over |weight| from both endpoints; endpoint 1, endpoint 2, then buses in order of settlement
 * (ties by bus index); -1 padded when the component is smaller than n_local */
int olfhip_neighbourhoods(int n_bus, int n_branch, const int *bus1, const int *bus2, const double *weight,
                          int n_local, int *buses);
/* x_norm = (x_loc - x_mean) / x_std on input; prediction = y * y_std + y_mean (all [2 n_local], FP32) */
int olfhip_mlp_set_normalization(olfhip_mlp *mlp, const float *x_mean, const float *x_std, const float *y_mean,
                                 const float *y_std);
/* the base states the batches refer to: x_base [n_rows][2 n_bus] (stable rows), copied to the device */
int olfhip_mlp_set_base_states(olfhip_mlp *mlp, int n_rows, const double *x_base);
/* one Adam step on a batch: base_row [batch] (NULL: row 0), branch [batch], features [batch][n_features], target
 * [batch][2 n_local] (normalised); returns the batch's mean squared error */
int olfhip_mlp_train_step(olfhip_mlp *mlp, int batch, const int *base_row, const int *branch, const float *features,
                          const float *target, float learning_rate, double *loss);
/* normalised outputs y_out [batch][2 n_local] */
int olfhip_mlp_forward(olfhip_mlp *mlp, int batch, const int *base_row, const int *branch, const float *features,
                       float *y_out);
/* loss gradient of parameter array `which` (0 Win, 1 Emb, 2 b1, 3 W2, 4 b2, 5 W3, 6 b3) for a batch, no update */
int olfhip_mlp_gradients(olfhip_mlp *mlp, int batch, const int *base_row, const int *branch, const float *features,
                         const float *target, int which, float *grad_out, double *loss);
/* read (values_out) and/or write (values_in) parameter array `which` (FP32 masters; BF16 copies refreshed) */
int olfhip_mlp_parameters(olfhip_mlp *mlp, int which, float *values_out, const float *values_in);
/* warm starts for `batch` scenarios into a device buffer owned by the network (valid until its next call), to pass as
 * olfhip_solve_inputs.x0_per_scenario_device: x0 = the base state of base_row[b] (NULL: row 0), plus
 * (y * y_std + y_mean) * scale[b] (scale NULL: 1) on the rows of the outage's neighbourhood */
int olfhip_mlp_predict_warm_start(olfhip_mlp *mlp, int batch, const int *base_row, const int *branch,
                                  const float *features, const float *scale, double **x0_device);
/* C = A B on the matrix-core GEMM of the network (A [M][K], B [K][N] row-major FP32, rounded to BF16), for tests */
int olfhip_mlp_gemm_test(int M, int N, int K, const float *A, const float *B, float *C);

#ifdef __cplusplus
}
#endif
#endif
