/*
 * CPU oracle for the AC Newton-Raphson N-1 path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference's algorithm, used by tests/, by the smoke check and by
 * bench.py's CPU-baseline leg, always as the checker and never as the thing shipped. The product path
 * (libolfhip.so) never links or calls it.
 *
 * Every function cites the reference file:line it follows. Parity pins live in tests/ (see the
 * test names quoted next to each function). Functions without a pin say "parity unpinned".
 */
#ifndef OLF_ORACLE_H
#define OLF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: reference GpuAcSolverStatus (docs/gpu-vs-olf.md:196). */
enum { ORC_CONVERGED = 0, ORC_MAX_ITERATION_REACHED = 1, ORC_DIVERGED = -1, ORC_REJECTED = -2,
       ORC_LINEAR_SOLVE_FAILED = -3 };
/* Row modes: reference GpuRowMode. */
enum { ORC_FIXED_VARIABLE = 0, ORC_EQUATION = 1 };
/* Reactive-limit type: reference GpuQLimitType. */
enum { ORC_QLIM_NONE = 0, ORC_QLIM_MIN = 1, ORC_QLIM_MAX = 2 };

/* One closed branch, 28 outputs in the order of the reference scatter table
 * (native/cuda/equations/generated/closed_branch_scatter.h:19-51):
 * p1 dp1dv1 dp1dv2 dp1dph1 dp1dph2 dp1da1 dp1dr1, q1 ..., p2 ..., q2 ...
 * p8 = (y, g1, g2, b1, b2, ksi, r1, a1).  Pinned by tests/test_oracle_formulas.py. */
void oracle_closed_branch(const double *p8, double v1, double ph1, double v2, double ph2, double *out28);

typedef struct oracle_sys oracle_sys;

/* Stable-row system over n = 2*n_bus rows with the superset CSR pattern (every row carries its bus's
 * and its neighbours' phi/V columns), reference structure/csr_pattern.cu:75-140. */
oracle_sys *oracle_sys_create(int n_bus, int n_branch, const double *p8, const int *idx12);
/* Bus shunts, per bus conductance g and susceptance b in pu (NULL clears): p = g v^2, q = -b v^2 (reference
 * equations/generated/shunt.cuh:47-60). */
void oracle_sys_set_shunts(oracle_sys *s, const double *g, const double *b);
void oracle_sys_free(oracle_sys *s);
int oracle_sys_nnz(const oracle_sys *s);
void oracle_sys_pattern(const oracle_sys *s, int *rowptr, int *colidx);

/* Mismatch F = sum(flows) - target on EQUATION rows, F = x - target and a unit row on FIXED_VARIABLE rows
 * (reference equations/fill_launch.cu:82-126); qinj[bus] = sum of reactive flows leaving the bus (the
 * reference's qBuf). Branches listed in out_br are skipped (the disabled-element mask).
 * J may be NULL (mismatch only). Pinned by tests/test_oracle_formulas.py and tests/test_oracle_nr.py. */
void oracle_fill(const oracle_sys *s, const double *x, const int *mode, const double *target,
                 const int *out_br, int n_out, double *F, double *qinj, double *J);

/* Sparse LU with threshold partial pivoting (Gilbert-Peierls left-looking) and a bus-level minimum degree
 * column order; solves A x = b. Returns 0, or -1 on a structurally/numerically singular matrix.
 * Replaces the reference's linear solvers (lu/batched_lu.cuh:66-75, cuDSS) on the CPU.
 * Pinned by tests/test_oracle_lu.py (dense partial-pivot solves). */
int oracle_lu_solve(int n, const int *rowptr, const int *colidx, const double *vals, const double *b,
                    double *x);

/* Minimum-degree order of a symmetric graph given as CSR adjacency (no self loops needed).
 * Pinned by tests/test_oracle_lu.py (permutation + fill check). */
void oracle_min_degree(int n, const int *adjptr, const int *adj, int *perm);

typedef struct {
    int max_iter;          /* Newton budget per solve (OLF maxNewtonRaphsonIterations) */
    double tol;            /* ||F||inf stopping threshold (newton_iteration.cu:300) */
    int reactive_limits;   /* run the local PV<->PQ outer loop */
    double rl_tol;         /* reactive-limit tolerance, pu */
    int max_outer;         /* productive outer-loop iterations (outer_loop_driver.cu:191) */
    int max_switch;        /* maxPqPvSwitch (outer_loop_driver.cu:194) */
    int scaling_mode;      /* 0 NONE, 1 LINE_SEARCH (newton_iteration.cu:497-530), 2 MAX_VOLTAGE_CHANGE
                              (state_vector_scaling.cu:57-97) */
    double max_dv, max_dphi;
    int batch_of_one;      /* a bad last solve is LINEAR_SOLVE_FAILED (1) or REJECTED (0) */
    int ls_max_iteration;  /* LINE_SEARCH backtracks per test (<= 0: 10) */
    double ls_step_fold;   /* LINE_SEARCH: mu = ls_step_fold^-b (<= 1: 4/3) */
    double conv_eps_per_eq;/* > 0: OLF's criterion ||F||_2 < sqrt(conv_eps_per_eq^2 n_active) instead of ||F||inf < tol
                              (newton_iteration.cu:241-300) */
    int n_active;          /* active equations of the system (<= 0: n; oracle_n1 uses n - 2 islanded buses) */
    /* OLF's PerEquationTypeStoppingCriteria as per-row thresholds (newton_iteration.cu:296-300, convergence.cu:28-
       60): [n] for a row held as an identity (FIXED_VARIABLE) and [n] for an equation row; both set: converged when
       every row's |F| is below its mode's threshold (instead of the two criteria above) */
    const double *row_thr_fixed;
    const double *row_thr_eq;
    double lin_fail_ratio; /* the linear-solve guard's ||J dx + F||inf / ||F||inf bound (<= 0: the reference's 1e-2,
                              solver/iterative_refinement.cuh:23) */
} oracle_nr_params;
/* tol <= 0 with neither conv_eps_per_eq nor row thresholds: the fixed-iteration mode (reference
 * solver/outer_loop_driver.cu:165-175, newton_iteration.cu:141-151,336-343): exactly max_iter Newton steps with no
 * convergence test, no linear-solve guard and no outer loop; MAX_ITERATION_REACHED, or DIVERGED when the final
 * mismatch is not finite. Pinned by tests/test_oracle_fixed_iterations.py. */

typedef struct {
    int n;                 /* entries */
    const int *voltage_row, *bus;
    const double *min_q, *max_q, *load_q, *target_v;
} oracle_rl;

typedef struct {
    int status, iterations, total_iterations, solves, outer_switches;
    double norm;
    int slack_distribution_failed;   /* the last distributed-slack check left more than P_RESIDUE_EPS unplaced */
} oracle_result;

/* Distributed slack (the reference's per-bus water-fill, outerloop/distributed_slack.cu:131-195,215-305,530-600,
 * OLF's DistributedSlackOuterLoop + ActivePowerDistribution): participating buses (never the slack bus) with
 * factors (1 - their sum is the slack's retained share), their P-row targets at the anchor and their room up
 * (>= 0) and down (<= 0) from it; the slack bus and its own P target; the gate (slackBusPMaxMismatch, pu) and
 * whether a failed distribution rejects the case (slackDistributionFailureBehavior != LEAVE_ON_SLACK_BUS). */
typedef struct {
    int n;
    const int *bus;
    const double *factor, *base_target_p, *max_delta, *min_delta;
    int slack_bus;
    double slack_target_p;
    double tol;
    int reject_failure;
} oracle_ds;

/* One AC load flow with its outer loops, in place on x/mode/target/limit_type/switch_count: distributed slack (ds,
 * may be NULL) then the reactive limits, each driven to STABLE with a re-solve after every unstable check, the
 * sweep restarted from the first loop while it changed something (OLF's AcloadFlowEngine.runOuterLoops as the
 * reference's driver runs it, solver/outer_loop_driver.cu:300-363), at most max_outer re-solves.
 * islanded[bus] (may be NULL) marks buses whose entries leave the outer loops.
 * Follows solver/newton_iteration.cu:93-203,276-397,580-660 and outerloop/reactive_limits.cu:321-371.
 * Pinned by tests/test_oracle_nr.py (two-bus closed form, Y-bus mismatch, power balance) and
 * tests/test_oracle_distributed_slack.py (lossless closed form, water-fill invariants). */
void oracle_nr_solve(const oracle_sys *s, const oracle_nr_params *prm, const oracle_rl *rl, const oracle_ds *ds,
                     const int *out_br, int n_out, const unsigned char *islanded,
                     double *x, int *mode, double *target, signed char *limit_type,
                     signed char *switch_count, oracle_result *res);

/* The slack bus's computed active injection (sum of its in-service branches' p at its side), pu. */
double oracle_bus_p(const oracle_sys *s, const double *x, const int *out_br, int n_out, int bus);

/* Buses cut off from the slack's component once the listed branches are out (1 = islanded).
 * Reference batch/plan/IslandedBuses.java:50-60. Returns the islanded count. Pinned by
 * tests/test_oracle_nr.py (bridge outages vs the generator's bridge list). */
int oracle_islands(const oracle_sys *s, int slack_bus, const int *out_br, int n_out, unsigned char *islanded);

/* Batched N-1 on the CPU: every contingency starts from the base state (or from its own row of x0_per [n_ctg][n] when
 * x0_per is not NULL, the per-scenario start of BatchSolveNative.java:119) and the base modes/targets/limit types,
 * islanded buses are frozen at 1<0 (identity rows), OpenMP over contingencies with n_threads threads.
 * ctg is n_ctg x max_out branch ids, -1 padded. Outputs per contingency (target_out, may be NULL: the final
 * targets, n per contingency). A case whose distribution failed under ds->reject_failure ends REJECTED.
 * Returns the thread count used. */
int oracle_n1(const oracle_sys *s, const oracle_nr_params *prm, const oracle_rl *rl, const oracle_ds *ds,
              int slack_bus, const double *x_base, const int *mode_base, const double *target_base,
              const signed char *limit_type_base, int n_ctg, int max_out, const int *ctg,
              int n_threads, double *x_out, int *status, int *iterations, int *total_iterations,
              double *norm, double *target_out, const double *x0_per);

/* Limit table (the reference's GpuLimitTable, flattened): per monitored closed branch its groups (side 1|2,
 * quantity 0 current |S|/V, 1 active power, 2 apparent power) with thresholds most severe first; per bus a high
 * and a low voltage limit (NaN: none). */
typedef struct {
    int n_branches;
    const int *branch, *group_start, *group_side, *group_quantity, *limit_start;
    const double *threshold;
    int n_buses;
    const double *bus_high, *bus_low;
    int n_angle;                         /* voltage-angle limits: to - from (rad) against high / low (NaN: none) */
    const int *angle_from, *angle_to;
    const double *angle_high, *angle_low;
} oracle_limits;

typedef struct {
    int kind;           /* 0 branch, 1 high voltage, 2 low voltage, 3 high / 4 low voltage angle */
    int element;        /* table branch, bus or voltage-angle limit */
    int side, quantity, limit;
    double value, threshold;
} oracle_violation;

/* OLF's full detection scan on one state (reference sa/gpu/GpuLimitScanner.java:95-131,394-410, which follows
 * LimitViolationManager.detectViolations): every monitored branch in service (not outaged, both ends not
 * islanded) in table order, each group's first exceeded threshold (active power compares |p|, reports p;
 * a NaN apparent power is skipped); then every bus in service, high then low (NaN limits never fire); then every
 * voltage-angle limit whose two buses are in service, the angle difference to - from against its high then its low
 * value (GpuLimitScanner.java:239-267, OLF's detectVoltageAngleViolations).
 * Flows from oracle_closed_branch. Writes at most max_out violations, returns the count found.
 * Pinned by tests/test_oracle_violations.py (textbook flows, hand cases, brute-force rescan). */
int oracle_violations(const oracle_sys *s, const double *x, const int *out_br, int n_out,
                      const unsigned char *islanded, const oracle_limits *lim, oracle_violation *out, int max_out);

#ifdef __cplusplus
}
#endif
#endif
