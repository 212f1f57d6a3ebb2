/*
 * A fake JVM for the JNI shim's tests (TEST INFRASTRUCTURE ONLY): Java arrays and classes as plain heap objects, the
 * JNIEnv function table over them, one pending exception, and helpers that Python (ctypes) uses to make the shim's
 * Java arguments and read its results. Linked with olf-gpu_amd/jni/olfhip_jni.c into libfakejni.so.
 */
#include <stdlib.h>
#include <string.h>

#include "jni.h"

enum { K_CLASS = 0, K_DOUBLES = 1, K_INTS = 2, K_BOOLEANS = 3, K_BYTES = 4, K_OBJECTS = 5 };

struct fake_object {
    int kind;
    jsize len;
    void *data;
    char name[128];
    int pinned;   /* Get*ArrayElements without the matching release */
};

static const size_t ELEM[] = {0, sizeof(jdouble), sizeof(jint), sizeof(jboolean), sizeof(jbyte), sizeof(jobject)};

static int excPending = 0;
static char excClass[128], excMessage[512];
static int violations = 0;   /* out-of-range accesses and unbalanced pins the shim made */

static jobject newObject(int kind, jsize len) {
    jobject o = (jobject)calloc(1, sizeof(struct fake_object));
    o->kind = kind;
    o->len = len;
    if (kind != K_CLASS) o->data = calloc((size_t)(len > 0 ? len : 1), ELEM[kind]);
    return o;
}

static int inRange(jobject a, int kind, jsize start, jsize n) {
    if (!a || a->kind != kind || start < 0 || n < 0 || start + n > a->len) {
        violations++;
        return 0;
    }
    return 1;
}

static jclass FindClass(JNIEnv *env, const char *name) {
    (void)env;
    jobject c = newObject(K_CLASS, 0);
    strncpy(c->name, name, sizeof c->name - 1);
    return c;
}
static jint ThrowNew(JNIEnv *env, jclass c, const char *msg) {
    (void)env;
    excPending = 1;
    strncpy(excClass, c ? c->name : "?", sizeof excClass - 1);
    strncpy(excMessage, msg ? msg : "", sizeof excMessage - 1);
    return 0;
}
static void DeleteLocalRef(JNIEnv *env, jobject o) { (void)env; (void)o; }
static jsize GetArrayLength(JNIEnv *env, jarray a) {
    (void)env;
    if (!a || a->kind == K_CLASS) { violations++; return 0; }
    return a->len;
}
static jobjectArray NewObjectArray(JNIEnv *env, jsize n, jclass c, jobject init) {
    (void)env; (void)c;
    jobject a = newObject(K_OBJECTS, n);
    for (jsize i = 0; i < n; i++) ((jobject *)a->data)[i] = init;
    return a;
}
static jobject GetObjectArrayElement(JNIEnv *env, jobjectArray a, jsize i) {
    (void)env;
    return inRange(a, K_OBJECTS, i, 1) ? ((jobject *)a->data)[i] : NULL;
}
static void SetObjectArrayElement(JNIEnv *env, jobjectArray a, jsize i, jobject v) {
    (void)env;
    if (inRange(a, K_OBJECTS, i, 1)) ((jobject *)a->data)[i] = v;
}
static jintArray NewIntArray(JNIEnv *env, jsize n) { (void)env; return newObject(K_INTS, n); }
static jdoubleArray NewDoubleArray(JNIEnv *env, jsize n) { (void)env; return newObject(K_DOUBLES, n); }
static jboolean *GetBooleanArrayElements(JNIEnv *env, jbooleanArray a, jboolean *isCopy) {
    (void)env;
    if (isCopy) *isCopy = 0;
    if (!inRange(a, K_BOOLEANS, 0, 0)) return NULL;
    a->pinned++;
    return (jboolean *)a->data;
}
static jbyte *GetByteArrayElements(JNIEnv *env, jbyteArray a, jboolean *isCopy) {
    (void)env;
    if (isCopy) *isCopy = 0;
    if (!inRange(a, K_BYTES, 0, 0)) return NULL;
    a->pinned++;
    return (jbyte *)a->data;
}
static void ReleaseBooleanArrayElements(JNIEnv *env, jbooleanArray a, jboolean *p, jint mode) {
    (void)env; (void)mode;
    if (!a || p != a->data || a->pinned-- <= 0) violations++;
}
static void ReleaseByteArrayElements(JNIEnv *env, jbyteArray a, jbyte *p, jint mode) {
    (void)env; (void)mode;
    if (!a || p != a->data || a->pinned-- <= 0) violations++;
}
static void GetIntArrayRegion(JNIEnv *env, jintArray a, jsize s, jsize n, jint *buf) {
    (void)env;
    if (inRange(a, K_INTS, s, n)) memcpy(buf, (jint *)a->data + s, sizeof(jint) * (size_t)n);
}
static void GetDoubleArrayRegion(JNIEnv *env, jdoubleArray a, jsize s, jsize n, jdouble *buf) {
    (void)env;
    if (inRange(a, K_DOUBLES, s, n)) memcpy(buf, (jdouble *)a->data + s, sizeof(jdouble) * (size_t)n);
}
static void SetIntArrayRegion(JNIEnv *env, jintArray a, jsize s, jsize n, const jint *buf) {
    (void)env;
    if (inRange(a, K_INTS, s, n)) memcpy((jint *)a->data + s, buf, sizeof(jint) * (size_t)n);
}
static void SetDoubleArrayRegion(JNIEnv *env, jdoubleArray a, jsize s, jsize n, const jdouble *buf) {
    (void)env;
    if (inRange(a, K_DOUBLES, s, n)) memcpy((jdouble *)a->data + s, buf, sizeof(jdouble) * (size_t)n);
}

static const struct JNINativeInterface_ TABLE = {
    FindClass, ThrowNew, DeleteLocalRef, GetArrayLength, NewObjectArray, GetObjectArrayElement, SetObjectArrayElement,
    NewIntArray, NewDoubleArray, GetBooleanArrayElements, GetByteArrayElements, ReleaseBooleanArrayElements,
    ReleaseByteArrayElements, GetIntArrayRegion, GetDoubleArrayRegion, SetIntArrayRegion, SetDoubleArrayRegion};
static JNIEnv ENV = &TABLE;

/* ---- helpers for the tests (ctypes) */
JNIEXPORT JNIEnv *fj_env(void) { return &ENV; }
JNIEXPORT jobject fj_doubles(jsize n, const jdouble *v) {
    jobject a = newObject(K_DOUBLES, n);
    if (n > 0 && v) memcpy(a->data, v, sizeof(jdouble) * (size_t)n);
    return a;
}
JNIEXPORT jobject fj_ints(jsize n, const jint *v) {
    jobject a = newObject(K_INTS, n);
    if (n > 0 && v) memcpy(a->data, v, sizeof(jint) * (size_t)n);
    return a;
}
JNIEXPORT jobject fj_booleans(jsize n, const jboolean *v) {
    jobject a = newObject(K_BOOLEANS, n);
    if (n > 0 && v) memcpy(a->data, v, (size_t)n);
    return a;
}
JNIEXPORT jobject fj_bytes(jsize n, const jbyte *v) {
    jobject a = newObject(K_BYTES, n);
    if (n > 0 && v) memcpy(a->data, v, (size_t)n);
    return a;
}
JNIEXPORT jobject fj_objects(jsize n) { return newObject(K_OBJECTS, n); }
JNIEXPORT void fj_set(jobject a, jsize i, jobject v) { SetObjectArrayElement(&ENV, a, i, v); }
JNIEXPORT jobject fj_get(jobject a, jsize i) { return GetObjectArrayElement(&ENV, a, i); }
JNIEXPORT jsize fj_length(jobject a) { return a ? a->len : -1; }
JNIEXPORT int fj_kind(jobject a) { return a ? a->kind : -1; }
JNIEXPORT int fj_read_doubles(jobject a, jdouble *out) {
    if (!a || a->kind != K_DOUBLES) return -1;
    memcpy(out, a->data, sizeof(jdouble) * (size_t)a->len);
    return a->len;
}
JNIEXPORT int fj_read_ints(jobject a, jint *out) {
    if (!a || a->kind != K_INTS) return -1;
    memcpy(out, a->data, sizeof(jint) * (size_t)a->len);
    return a->len;
}
JNIEXPORT int fj_pinned(jobject a) { return a ? a->pinned : 0; }
JNIEXPORT int fj_exception(char *cls, int ncls, char *msg, int nmsg) {
    if (!excPending) return 0;
    strncpy(cls, excClass, (size_t)ncls - 1);
    cls[ncls - 1] = 0;
    strncpy(msg, excMessage, (size_t)nmsg - 1);
    msg[nmsg - 1] = 0;
    return 1;
}
JNIEXPORT void fj_clear_exception(void) { excPending = 0; excClass[0] = excMessage[0] = 0; }
JNIEXPORT int fj_violations(void) { return violations; }
