/*
 * A test double of <jni.h> for the JNI shim's tests (TEST INFRASTRUCTURE ONLY: the product build of libolfgpu.so uses
 * a JDK's own jni.h): the JNI types and the JNIEnv function-table entries the shim calls, with the names and
 * signatures of the JNI specification. The table's layout is this file's own; the fake JVM in fake_jvm.c fills it.
 */
#ifndef OLFHIP_TEST_JNI_H
#define OLFHIP_TEST_JNI_H

#include <stdint.h>

typedef int32_t jint;
typedef int64_t jlong;
typedef int8_t jbyte;
typedef uint8_t jboolean;
typedef double jdouble;
typedef jint jsize;

typedef struct fake_object *jobject;
typedef jobject jclass;
typedef jobject jarray;
typedef jobject jdoubleArray;
typedef jobject jintArray;
typedef jobject jbooleanArray;
typedef jobject jbyteArray;
typedef jobject jobjectArray;

#define JNIEXPORT __attribute__((visibility("default")))
#define JNICALL
#define JNI_ABORT 2

struct JNINativeInterface_;
typedef const struct JNINativeInterface_ *JNIEnv;

struct JNINativeInterface_ {
    jclass (*FindClass)(JNIEnv *, const char *);
    jint (*ThrowNew)(JNIEnv *, jclass, const char *);
    void (*DeleteLocalRef)(JNIEnv *, jobject);
    jsize (*GetArrayLength)(JNIEnv *, jarray);
    jobjectArray (*NewObjectArray)(JNIEnv *, jsize, jclass, jobject);
    jobject (*GetObjectArrayElement)(JNIEnv *, jobjectArray, jsize);
    void (*SetObjectArrayElement)(JNIEnv *, jobjectArray, jsize, jobject);
    jintArray (*NewIntArray)(JNIEnv *, jsize);
    jdoubleArray (*NewDoubleArray)(JNIEnv *, jsize);
    jboolean *(*GetBooleanArrayElements)(JNIEnv *, jbooleanArray, jboolean *);
    jbyte *(*GetByteArrayElements)(JNIEnv *, jbyteArray, jboolean *);
    void (*ReleaseBooleanArrayElements)(JNIEnv *, jbooleanArray, jboolean *, jint);
    void (*ReleaseByteArrayElements)(JNIEnv *, jbyteArray, jbyte *, jint);
    void (*GetIntArrayRegion)(JNIEnv *, jintArray, jsize, jsize, jint *);
    void (*GetDoubleArrayRegion)(JNIEnv *, jdoubleArray, jsize, jsize, jdouble *);
    void (*SetIntArrayRegion)(JNIEnv *, jintArray, jsize, jsize, const jint *);
    void (*SetDoubleArrayRegion)(JNIEnv *, jdoubleArray, jsize, jsize, const jdouble *);
};

#endif
