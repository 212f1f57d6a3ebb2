/*
 * JNI binding of the reference's StructureNative and BatchSolveNative natives over the C ABI of libolfhip.so
 * (include/olfhip.h). Built as libolfgpu.so, the name the reference's loader asks for
 * (NativeLoader.loadLibrary("olfgpu"), GpuNativeLoader.java:32-41), only where <jni.h> is available
 * (olf-gpu_amd/build.py looks under $JAVA_HOME and /usr/lib/jvm).
 *
 * One JNI function per Java native, each one call into the library:
 *   StructureNative.createStructure / destroyStructure / recommendBatchSize
 *     (src/main/java/com/powsybl/openloadflow/ac/equations/vector/gpu/StructureNative.java:49-81)
 *   BatchSolveNative.createBatchContext / destroyBatchContext / beginBatchInstall / solveBatchContext /
 *     readBatchHandedBackSlots (src/main/java/com/powsybl/openloadflow/ac/solver/gpu/BatchSolveNative.java:41-125)
 *
 * Java arrays are borrowed read-only (released with JNI_ABORT, never written back, as the reference's
 * jni_common.cuh:93-121); results go into fresh Java arrays. A library error becomes the exception class the reference's
 * JNI throws: IllegalArgumentException for a contract violation (OLFHIP_E_ARGUMENT), UnsupportedOperationException for an
 * input family or option this library does not model (OLFHIP_E_UNSUPPORTED, the reference's refusal that hands the case
 * back to OLF), RuntimeException "GPU batched Newton-Raphson failed: ..." otherwise (batch_solve_jni.cu:296-306). A null
 * Java allocation returns null with the OutOfMemoryError pending.
 */
#include <jni.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "olfhip.h"

/* The reference's GpuLinearSolveMode codes (codes/GpuLinearSolveMode.java:21-34) and the one this library adds. */
enum { MODE_UBATCH = 3, MODE_ROBUST = 6, MODE_BATCHED_LU = 7, MODE_BATCHED_LU_FP32_REFINED = 8 };
/* GpuDistributedSlackFactorMode.STATIC (codes/GpuDistributedSlackFactorMode.java:22) */
enum { FACTOR_MODE_STATIC = 0 };
/* GpuReactiveLimitsState: limit type in the low 3 bits, PV -> PQ switch count above (gpu_reactive_limits_state.cuh) */
enum { RL_STATE_TYPE_MASK = 7, RL_STATE_COUNT_SHIFT = 3 };

/* What the Java handles point at: the library's handle plus the sizes the marshalling needs. */
typedef struct {
    olfhip_structure *structure;
    int n, nClosed, nOpen, nShunt, nHvdc;
} jni_structure;

typedef struct {
    olfhip_context *context;
    int n, nClosed, nElements, nReactiveLimits, capacity, distributedSlack;
} jni_context;

static void throwClass(JNIEnv *env, const char *cls, const char *msg) {
    jclass c = (*env)->FindClass(env, cls);
    if (c) (*env)->ThrowNew(env, c, msg);
}

static void throwFor(JNIEnv *env, int rc) {
    const char *cls = rc == OLFHIP_E_ARGUMENT      ? "java/lang/IllegalArgumentException"
                      : rc == OLFHIP_E_UNSUPPORTED ? "java/lang/UnsupportedOperationException"
                                                   : "java/lang/RuntimeException";
    throwClass(env, cls, olfhip_last_error());
}

static jsize lengthOf(JNIEnv *env, jarray a) { return a ? (*env)->GetArrayLength(env, a) : 0; }

/* Refuses an input the library does not model: a non-empty array of a family or an option it has no counterpart for. */
static int refuseNonEmpty(JNIEnv *env, jarray a, const char *what) {
    if (lengthOf(env, a) == 0) return 0;
    char msg[256];
    strcpy(msg, what);
    strcat(msg, ": not modelled by the MI355X library (hand the case back to OLF)");
    throwClass(env, "java/lang/UnsupportedOperationException", msg);
    return 1;
}

/* Copies of Java arrays (NULL for a null or empty array); *ok is cleared when a copy fails (OOM pending). */
static double *doubles(JNIEnv *env, jdoubleArray a, int *ok) {
    jsize n = lengthOf(env, a);
    if (n == 0) return NULL;
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    if (!p) { throwClass(env, "java/lang/OutOfMemoryError", "olfgpu JNI"); *ok = 0; return NULL; }
    (*env)->GetDoubleArrayRegion(env, a, 0, n, p);
    return p;
}

static int *ints(JNIEnv *env, jintArray a, int *ok) {
    jsize n = lengthOf(env, a);
    if (n == 0) return NULL;
    int *p = (int *)malloc(sizeof(int) * (size_t)n);
    if (!p) { throwClass(env, "java/lang/OutOfMemoryError", "olfgpu JNI"); *ok = 0; return NULL; }
    (*env)->GetIntArrayRegion(env, a, 0, n, (jint *)p);
    return p;
}

JNIEXPORT jlong JNICALL Java_com_powsybl_openloadflow_ac_equations_vector_gpu_StructureNative_createStructure(
    JNIEnv *env, jclass cls, jint n, jdoubleArray closedBranchParameters, jintArray closedBranchIndices,
    jdoubleArray openBranchParameters, jintArray openBranchIndices, jdoubleArray shuntParameters, jintArray shuntIndices,
    jdoubleArray hvdcAcEmulationParameters, jintArray hvdcAcEmulationIndices, jintArray distrQElement, jintArray distrQSide,
    jdoubleArray distrQWeight, jintArray distrQRow, jintArray distrQKind, jdoubleArray zeroImpedanceBranchParameters,
    jintArray zeroImpedanceBranchIndices, jintArray distrRhoRow, jintArray distrRhoColumn, jdoubleArray distrRhoCoefficient,
    jintArray distrShuntBRow, jintArray distrShuntBColumn, jdoubleArray distrShuntBCoefficient,
    jintArray closedBranchPhaseControlSide, jdoubleArray loadModelParameters, jintArray loadModelIndices,
    jintArray reactiveRelocationGroupStart, jintArray reactiveRelocationCandidateRows,
    jintArray openBranchAlternateReactiveDestination, jintArray zeroImpedanceBranchAlternateReactiveDestination,
    jintArray reactiveRelocationDependentBaseGroup, jintArray reactiveRelocationBaseTargetRow,
    jintArray reactiveRelocationDependentGroupOfRow, jintArray identityRows) {
    (void)cls;
    /* the families and options of the reference's structure that the library does not model */
    if (refuseNonEmpty(env, openBranchIndices, "open branches") ||
        refuseNonEmpty(env, hvdcAcEmulationIndices, "HVDC AC emulation") ||
        refuseNonEmpty(env, distrQElement, "distributed reactive terms") ||
        refuseNonEmpty(env, zeroImpedanceBranchIndices, "zero-impedance branches") ||
        refuseNonEmpty(env, distrRhoRow, "transformer ratio distribution") ||
        refuseNonEmpty(env, distrShuntBRow, "shunt susceptance distribution") ||
        refuseNonEmpty(env, loadModelIndices, "load models") ||
        refuseNonEmpty(env, reactiveRelocationGroupStart, "reactive relocation") ||
        refuseNonEmpty(env, reactiveRelocationCandidateRows, "reactive relocation") ||
        refuseNonEmpty(env, openBranchAlternateReactiveDestination, "reactive relocation") ||
        refuseNonEmpty(env, zeroImpedanceBranchAlternateReactiveDestination, "reactive relocation") ||
        refuseNonEmpty(env, reactiveRelocationDependentBaseGroup, "reactive relocation") ||
        refuseNonEmpty(env, reactiveRelocationBaseTargetRow, "reactive relocation") ||
        refuseNonEmpty(env, reactiveRelocationDependentGroupOfRow, "reactive relocation") ||
        refuseNonEmpty(env, identityRows, "identity rows of zero-impedance dummies"))
        return 0;
    (void)openBranchParameters; (void)hvdcAcEmulationParameters; (void)distrQSide; (void)distrQWeight;
    (void)distrQRow; (void)distrQKind; (void)zeroImpedanceBranchParameters; (void)distrRhoColumn;
    (void)distrRhoCoefficient; (void)distrShuntBColumn; (void)distrShuntBCoefficient; (void)loadModelParameters;
    int ok = 1;
    int *phase = ints(env, closedBranchPhaseControlSide, &ok);
    if (!ok) return 0;
    for (jsize e = 0; phase && e < lengthOf(env, closedBranchPhaseControlSide); e++)
        if (phase[e] >= 0) {
            free(phase);
            throwClass(env, "java/lang/UnsupportedOperationException",
                       "phase control on a closed branch: not modelled by the MI355X library");
            return 0;
        }
    free(phase);
    const int nClosed = (int)(lengthOf(env, closedBranchIndices) / 12);
    const int nShunt = (int)(lengthOf(env, shuntIndices) / 3);
    double *p = doubles(env, closedBranchParameters, &ok);
    int *ix = ints(env, closedBranchIndices, &ok);
    double *sp = doubles(env, shuntParameters, &ok);
    int *si = ints(env, shuntIndices, &ok);
    jni_structure *h = ok ? (jni_structure *)calloc(1, sizeof(jni_structure)) : NULL;
    int rc = OLFHIP_OK;
    if (ok && h) {
        rc = olfhip_create_structure(n, nClosed, p, ix, nShunt, sp, si, &h->structure);
        h->n = n;
        h->nClosed = nClosed;
        h->nShunt = nShunt;
    }
    free(p); free(ix); free(sp); free(si);
    if (!ok) { free(h); return 0; }
    if (!h) { throwClass(env, "java/lang/OutOfMemoryError", "olfgpu JNI"); return 0; }
    if (rc != OLFHIP_OK) { free(h); throwFor(env, rc); return 0; }
    return (jlong)(intptr_t)h;
}

/* Idempotent for handle 0 (StructureNative.java:61-67); contexts keep the library's structure alive. */
JNIEXPORT void JNICALL Java_com_powsybl_openloadflow_ac_equations_vector_gpu_StructureNative_destroyStructure(
    JNIEnv *env, jclass cls, jlong handle) {
    (void)env; (void)cls;
    jni_structure *h = (jni_structure *)(intptr_t)handle;
    if (!h) return;
    olfhip_destroy_structure(h->structure);
    free(h);
}

static int libraryMode(JNIEnv *env, jint mode) {
    switch (mode) {
        case MODE_BATCHED_LU: return OLFHIP_LU_FP64;
        case MODE_BATCHED_LU_FP32_REFINED: return OLFHIP_LU_FP32_REFINED;
        case MODE_ROBUST: return OLFHIP_LU_ROBUST;
        default:
            throwClass(env, "java/lang/IllegalArgumentException",
                       mode == MODE_UBATCH ? "linear-solve mode UBATCH is a cuDSS mode: not available on MI355X"
                                           : "unknown linear-solve mode");
            return -1;
    }
}

JNIEXPORT jint JNICALL Java_com_powsybl_openloadflow_ac_equations_vector_gpu_StructureNative_recommendBatchSize(
    JNIEnv *env, jclass cls, jlong handle, jint maxCap, jint mode) {
    (void)cls;
    jni_structure *h = (jni_structure *)(intptr_t)handle;
    if (!h) { throwClass(env, "java/lang/IllegalArgumentException", "null structure handle"); return 0; }
    const int m = libraryMode(env, mode);
    if (m < 0) return 0;
    int out = 0;
    const int rc = olfhip_recommend_batch_size(h->structure, maxCap, m, &out);
    if (rc != OLFHIP_OK) { throwFor(env, rc); return 0; }
    return out;
}

JNIEXPORT jlong JNICALL Java_com_powsybl_openloadflow_ac_solver_gpu_BatchSolveNative_createBatchContext(
    JNIEnv *env, jclass cls, jlong structureHandle, jint nScenarios, jint mode, jintArray reactiveLimitVoltageRow,
    jdoubleArray reactiveLimitMinQ, jdoubleArray reactiveLimitMaxQ, jdoubleArray reactiveLimitLoadQ,
    jdoubleArray reactiveLimitTargetV, jbyteArray reactiveLimitFlags, jdoubleArray rowSlope,
    jintArray distributedSlackBusPhiRow, jdoubleArray distributedSlackBusFactor, jint slackBusPhiRow,
    jdouble slackBusTargetP, jdoubleArray distributedSlackBusBaseTargetP, jdoubleArray distributedSlackBusMaxDelta,
    jdoubleArray distributedSlackBusMinDelta, jint distributedSlackFactorMode, jdoubleArray distributedSlackBusMarginUp,
    jdoubleArray distributedSlackBusMarginDown, jdouble slackBusMarginUp, jdouble slackBusMarginDown) {
    (void)cls;
    jni_structure *s = (jni_structure *)(intptr_t)structureHandle;
    if (!s) { throwClass(env, "java/lang/IllegalArgumentException", "null structure handle"); return 0; }
    const int m = libraryMode(env, mode);
    if (m < 0) return 0;
    if (refuseNonEmpty(env, rowSlope, "voltage-control slopes") ||
        refuseNonEmpty(env, distributedSlackBusMarginUp, "distributed-slack margins") ||
        refuseNonEmpty(env, distributedSlackBusMarginDown, "distributed-slack margins"))
        return 0;
    if (slackBusMarginUp != 0.0 || slackBusMarginDown != 0.0 ||
        (lengthOf(env, distributedSlackBusPhiRow) > 0 && distributedSlackFactorMode != FACTOR_MODE_STATIC)) {
        throwClass(env, "java/lang/UnsupportedOperationException",
                   "distributed slack with margins or REMAINING_MARGIN factors: not modelled by the MI355X library");
        return 0;
    }
    /* an entry flag (stand-by automaton, not OLF's controller element) asks for behaviour the library does not have */
    const jsize nFlags = lengthOf(env, reactiveLimitFlags);
    if (nFlags > 0) {
        jbyte *f = (*env)->GetByteArrayElements(env, reactiveLimitFlags, NULL);
        int any = 0;
        for (jsize i = 0; f && i < nFlags; i++) any |= f[i] != 0;
        if (f) (*env)->ReleaseByteArrayElements(env, reactiveLimitFlags, f, JNI_ABORT);
        if (any) {
            throwClass(env, "java/lang/UnsupportedOperationException",
                       "reactive-limit entry flags: not modelled by the MI355X library");
            return 0;
        }
    }
    int ok = 1;
    olfhip_reactive_limits rl;
    memset(&rl, 0, sizeof rl);
    rl.n_entries = (int)lengthOf(env, reactiveLimitVoltageRow);
    int *vrow = ints(env, reactiveLimitVoltageRow, &ok);
    double *minQ = doubles(env, reactiveLimitMinQ, &ok), *maxQ = doubles(env, reactiveLimitMaxQ, &ok);
    double *loadQ = doubles(env, reactiveLimitLoadQ, &ok), *targetV = doubles(env, reactiveLimitTargetV, &ok);
    rl.voltage_row = vrow; rl.min_q = minQ; rl.max_q = maxQ; rl.load_q = loadQ; rl.target_v = targetV;
    const int nDs = (int)lengthOf(env, distributedSlackBusPhiRow);
    int *dsRow = ints(env, distributedSlackBusPhiRow, &ok);
    double *dsFactor = doubles(env, distributedSlackBusFactor, &ok);
    double *dsBase = doubles(env, distributedSlackBusBaseTargetP, &ok);
    double *dsMax = doubles(env, distributedSlackBusMaxDelta, &ok), *dsMin = doubles(env, distributedSlackBusMinDelta, &ok);
    jni_context *h = ok ? (jni_context *)calloc(1, sizeof(jni_context)) : NULL;
    int rc = OLFHIP_OK;
    if (ok && h) {
        rc = olfhip_create_batch_context(s->structure, nScenarios, m, &rl, NULL, &h->context);
        if (rc == OLFHIP_OK && nDs > 0) {
            /* the reference passes phi rows; the library takes buses (stable rows: bus i's phi row is 2i) */
            int *bus = (int *)malloc(sizeof(int) * (size_t)nDs);
            if (!bus) { rc = OLFHIP_E_MEMORY; }
            else {
                for (int i = 0; i < nDs; i++) bus[i] = dsRow[i] / 2;
                olfhip_distributed_slack ds = {nDs, bus, dsFactor, dsBase, dsMax, dsMin, slackBusPhiRow / 2,
                                               slackBusTargetP};
                rc = olfhip_set_distributed_slack(h->context, &ds);
                free(bus);
            }
        }
        h->n = s->n;
        h->nClosed = s->nClosed;
        h->nElements = s->nClosed + s->nOpen + s->nShunt + s->nHvdc;
        h->nReactiveLimits = rl.n_entries;
        h->capacity = nScenarios;
        h->distributedSlack = nDs > 0;
    }
    free(vrow); free(minQ); free(maxQ); free(loadQ); free(targetV);
    free(dsRow); free(dsFactor); free(dsBase); free(dsMax); free(dsMin);
    if (!ok) { free(h); return 0; }
    if (!h) { throwClass(env, "java/lang/OutOfMemoryError", "olfgpu JNI"); return 0; }
    if (rc != OLFHIP_OK) {
        olfhip_destroy_batch_context(h->context);
        free(h);
        throwFor(env, rc);
        return 0;
    }
    return (jlong)(intptr_t)h;
}

JNIEXPORT void JNICALL Java_com_powsybl_openloadflow_ac_solver_gpu_BatchSolveNative_destroyBatchContext(
    JNIEnv *env, jclass cls, jlong handle) {
    (void)env; (void)cls;
    jni_context *h = (jni_context *)(intptr_t)handle;
    if (!h) return;
    olfhip_destroy_batch_context(h->context);
    free(h);
}

/* The disabled-element mask [nS][nElements] as at most OLFHIP's four outaged closed branches per scenario (closed
 * branches are the first elements, BatchSolveNative.java:81-83); another disabled element is refused. */
static int outagesOf(JNIEnv *env, jobjectArray mask, int nS, const jni_context *h, int *outaged) {
    for (int s = 0; s < nS; s++) {
        for (int q = 0; q < 4; q++) outaged[4 * s + q] = -1;
        jbooleanArray row = (jbooleanArray)(*env)->GetObjectArrayElement(env, mask, s);
        if (!row) continue;
        const jsize len = (*env)->GetArrayLength(env, row);
        jboolean *b = (*env)->GetBooleanArrayElements(env, row, NULL);
        int count = 0, bad = 0;
        for (jsize e = 0; b && e < len; e++) {
            if (!b[e]) continue;
            if (e >= h->nClosed || count == 4) { bad = 1; break; }
            outaged[4 * s + count++] = (int)e;
        }
        if (b) (*env)->ReleaseBooleanArrayElements(env, row, b, JNI_ABORT);
        (*env)->DeleteLocalRef(env, row);
        if (bad) {
            throwClass(env, "java/lang/UnsupportedOperationException",
                       "a scenario disables an element other than a closed branch, or more than four closed "
                       "branches: not modelled by the MI355X library");
            return 0;
        }
    }
    return 1;
}

JNIEXPORT jobjectArray JNICALL Java_com_powsybl_openloadflow_ac_solver_gpu_BatchSolveNative_solveBatchContext(
    JNIEnv *env, jclass cls, jlong handle, jdoubleArray x0Base, jdoubleArray targetBase, jintArray rowModeEntries,
    jintArray baseRowMode, jobjectArray disabledElementMask, jintArray targetOverrideIndex,
    jdoubleArray targetOverrideValue, jintArray distributedSlackOverrideBus,
    jdoubleArray distributedSlackOverrideFactorDelta, jdoubleArray distributedSlackOverrideBaseTargetPDelta,
    jdoubleArray distributedSlackOverrideMaxDelta, jdoubleArray distributedSlackOverrideMinDelta,
    jdoubleArray distributedSlackBusAnchorShift, jdoubleArray slackBusAnchorShift,
    jintArray reactiveLimitTargetVOverrideScenarioEntry, jdoubleArray reactiveLimitTargetVOverrideValue,
    jintArray reactiveLimitDeltaOverrideScenarioEntry, jdoubleArray reactiveLimitDeltaOverrideValues,
    jintArray reactiveCapabilityCurveOverrideScenarioCurve, jdoubleArray reactiveCapabilityCurveOverrideStepScale,
    jint maxIter, jdouble tol, jdouble reactiveLimitTol, jdouble slackDistTol, jdouble convEpsPerEq,
    jint stateVectorScalingMode, jdouble maxDv, jdouble maxDphi, jintArray varType, jdoubleArray x0PerScenario,
    jintArray shuntOvSlotElem, jdoubleArray shuntOvGb, jintArray reactiveLimitsSeedState,
    jintArray closedBranchParameterOverrideScenarioElement, jdoubleArray closedBranchParameterOverrideValues,
    jint maxOuterIterations, jint maxPqPvSwitch, jint rejectSlackDistributionFailure, jint skipReactiveLimitsPhase,
    jint nBaseClosedZeroImpedanceBranches, jintArray zeroImpedanceBranchCloseScenarioElement,
    jintArray zeroImpedanceBranchOpenScenarioElement, jint pauseAtEnded) {
    (void)cls; (void)varType; (void)targetOverrideValue; (void)reactiveLimitTargetVOverrideValue;
    (void)reactiveLimitDeltaOverrideValues; (void)reactiveCapabilityCurveOverrideStepScale; (void)shuntOvGb;
    (void)closedBranchParameterOverrideValues; (void)distributedSlackOverrideFactorDelta;
    (void)distributedSlackOverrideBaseTargetPDelta; (void)distributedSlackOverrideMaxDelta;
    (void)distributedSlackOverrideMinDelta;
    jni_context *h = (jni_context *)(intptr_t)handle;
    if (!h) { throwClass(env, "java/lang/IllegalArgumentException", "null context handle"); return NULL; }
    if (refuseNonEmpty(env, distributedSlackOverrideBus, "per-scenario distributed-slack overrides") ||
        refuseNonEmpty(env, distributedSlackBusAnchorShift, "distributed-slack anchor shifts") ||
        refuseNonEmpty(env, slackBusAnchorShift, "slack anchor shifts") ||
        refuseNonEmpty(env, reactiveLimitTargetVOverrideScenarioEntry, "per-scenario voltage setpoints") ||
        refuseNonEmpty(env, reactiveLimitDeltaOverrideScenarioEntry, "per-scenario reactive limits") ||
        refuseNonEmpty(env, reactiveCapabilityCurveOverrideScenarioCurve, "capability-curve overrides") ||
        refuseNonEmpty(env, shuntOvSlotElem, "per-scenario shunt overrides") ||
        refuseNonEmpty(env, closedBranchParameterOverrideScenarioElement, "per-scenario branch parameters") ||
        refuseNonEmpty(env, zeroImpedanceBranchCloseScenarioElement, "zero-impedance branch switching") ||
        refuseNonEmpty(env, zeroImpedanceBranchOpenScenarioElement, "zero-impedance branch switching"))
        return NULL;
    if (nBaseClosedZeroImpedanceBranches != 0) {
        throwClass(env, "java/lang/UnsupportedOperationException",
                   "zero-impedance branches: not modelled by the MI355X library");
        return NULL;
    }
    const int n = h->n;
    /* the scenarios this call installs (into the slots beginBatchInstall chose, or the whole bank); it hands back
     * at most the bank's capacity of rows */
    const int nS = (int)lengthOf(env, disabledElementMask);
    if (nS > h->capacity || lengthOf(env, x0Base) != n || lengthOf(env, targetBase) != n ||
        lengthOf(env, baseRowMode) != n) {
        throwClass(env, "java/lang/IllegalArgumentException", "solveBatchContext: array sizes do not match the context");
        return NULL;
    }
    int ok = 1;
    double *x0 = doubles(env, x0Base, &ok), *target = doubles(env, targetBase, &ok);
    int *mode = ints(env, baseRowMode, &ok), *entries = ints(env, rowModeEntries, &ok);
    int *ovIdx = ints(env, targetOverrideIndex, &ok);
    double *ovVal = doubles(env, targetOverrideValue, &ok), *x0Per = doubles(env, x0PerScenario, &ok);
    int *seed = ints(env, reactiveLimitsSeedState, &ok);
    int *outaged = (int *)malloc(sizeof(int) * 4 * (size_t)(nS > 0 ? nS : 1));
    const int nOv = (int)lengthOf(env, targetOverrideIndex);
    long long *ovIndex = nOv ? (long long *)malloc(sizeof(long long) * (size_t)nOv) : NULL;
    const int nSeed = (int)lengthOf(env, reactiveLimitsSeedState);
    signed char *seedType = nSeed ? (signed char *)malloc((size_t)nSeed) : NULL;
    double *res = (double *)malloc(sizeof(double) * (size_t)h->capacity * (size_t)(n + 3));
    jobjectArray rows = NULL;
    int rc = OLFHIP_OK;
    if (ok && (!outaged || !res || (nOv && !ovIndex) || (nSeed && !seedType))) {
        throwClass(env, "java/lang/OutOfMemoryError", "olfgpu JNI");
        ok = 0;
    }
    if (ok && !outagesOf(env, disabledElementMask, nS, h, outaged)) ok = 0;
    if (ok && x0Per && lengthOf(env, x0PerScenario) != (jsize)nS * n) {
        throwClass(env, "java/lang/IllegalArgumentException", "x0PerScenario must hold nScenarios x n values");
        ok = 0;
    }
    if (ok && nSeed && nSeed != h->nReactiveLimits && nSeed != nS * h->nReactiveLimits) {
        throwClass(env, "java/lang/IllegalArgumentException",
                   "reactiveLimitsSeedState must hold one row per scenario (or one shared row)");
        ok = 0;
    }
    for (int i = 0; ok && i < nSeed; i++) {
        if ((seed[i] >> RL_STATE_COUNT_SHIFT) != 0) {
            /* switch counts carried over a re-solve are not modelled: the library counts afresh per solve */
            throwClass(env, "java/lang/UnsupportedOperationException",
                       "seeded PV -> PQ switch counts: not modelled by the MI355X library");
            ok = 0;
        }
        seedType[i] = (signed char)(seed[i] & RL_STATE_TYPE_MASK);
    }
    if (ok) {
        for (int i = 0; i < nOv; i++) ovIndex[i] = ovIdx[i];
        olfhip_solve_inputs in;
        memset(&in, 0, sizeof in);
        in.n_scenarios = nS;
        in.x0_base = x0;
        in.target_base = target;
        in.base_row_mode = mode;
        in.n_row_mode_entries = (int)(lengthOf(env, rowModeEntries) / 3);
        in.row_mode_entries = entries;
        in.n_target_overrides = nOv;
        in.target_override_index = ovIndex;
        in.target_override_value = ovVal;
        in.max_outages = 4;
        in.outaged_branches = outaged;
        in.x0_per_scenario = x0Per;
        in.reactive_limits_seed_state = seedType;
        in.reactive_limits_seed_shared = nSeed > 0 && nSeed == h->nReactiveLimits && nS > 1;
        in.max_iter = maxIter;
        in.tol = tol;
        in.reactive_limit_tol = reactiveLimitTol;
        in.state_vector_scaling_mode = stateVectorScalingMode;
        in.max_dv = maxDv;
        in.max_dphi = maxDphi;
        in.max_outer_iterations = maxOuterIterations;
        in.max_pq_pv_switch = maxPqPvSwitch;
        in.skip_reactive_limits = skipReactiveLimitsPhase;
        in.states_to_host = 1;
        in.distributed_slack = h->distributedSlack;
        in.slack_dist_tol = slackDistTol;
        in.reject_slack_distribution_failure = rejectSlackDistributionFailure;
        in.line_search_max_iteration = 0;       /* NewtonSettingsNative.setBatchLineSearch's defaults */
        in.line_search_step_fold = 0.0;
        in.conv_eps_per_eq = convEpsPerEq;
        in.pause_at_ended = pauseAtEnded;
        int handedBack = 0;
        olfhip_solve_outputs out = {res, NULL, NULL, NULL, NULL, &handedBack};
        rc = olfhip_solve_batch_context(h->context, &in, &out);
        if (rc != OLFHIP_OK) {
            throwFor(env, rc);
        } else {
            /* double[handed back][n + 3]: [x(n), iterations, ||F||inf, status] per scenario the solve handed back, in
             * install order (every installed scenario when the solve ran to its end on the whole bank) */
            jclass rowClass = (*env)->FindClass(env, "[D");
            rows = rowClass ? (*env)->NewObjectArray(env, handedBack, rowClass, NULL) : NULL;
            for (int s = 0; rows && s < handedBack; s++) {
                jdoubleArray r = (*env)->NewDoubleArray(env, n + 3);
                if (!r) { rows = NULL; break; }   /* OutOfMemoryError pending */
                (*env)->SetDoubleArrayRegion(env, r, 0, n + 3, res + (size_t)s * (size_t)(n + 3));
                (*env)->SetObjectArrayElement(env, rows, s, r);
                (*env)->DeleteLocalRef(env, r);
            }
        }
    }
    free(x0); free(target); free(mode); free(entries); free(ovIdx); free(ovVal); free(x0Per); free(seed);
    free(outaged); free(ovIndex); free(seedType); free(res);
    return rows;
}

JNIEXPORT void JNICALL Java_com_powsybl_openloadflow_ac_solver_gpu_BatchSolveNative_beginBatchInstall(
    JNIEnv *env, jclass cls, jlong handle, jintArray bankSlots) {
    (void)cls;
    jni_context *h = (jni_context *)(intptr_t)handle;
    if (!h) return;
    int ok = 1;
    int *slots = ints(env, bankSlots, &ok);
    if (!ok) return;
    const int rc = olfhip_begin_batch_install(h->context, (int)lengthOf(env, bankSlots), slots);
    free(slots);
    if (rc != OLFHIP_OK) throwFor(env, rc);
}

JNIEXPORT jintArray JNICALL Java_com_powsybl_openloadflow_ac_solver_gpu_BatchSolveNative_readBatchHandedBackSlots(
    JNIEnv *env, jclass cls, jlong handle) {
    (void)cls;
    jni_context *h = (jni_context *)(intptr_t)handle;
    if (!h) return NULL;
    int *slots = (int *)malloc(sizeof(int) * (size_t)h->capacity);
    if (!slots) { throwClass(env, "java/lang/OutOfMemoryError", "olfgpu JNI"); return NULL; }
    int count = 0;
    const int rc = olfhip_read_handed_back_slots(h->context, &count, slots);
    jintArray arr = NULL;
    if (rc != OLFHIP_OK) {
        throwFor(env, rc);
    } else {
        arr = (*env)->NewIntArray(env, count);
        if (arr) (*env)->SetIntArrayRegion(env, arr, 0, count, (const jint *)slots);
    }
    free(slots);
    return arr;
}
