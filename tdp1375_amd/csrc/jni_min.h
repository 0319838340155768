// Minimal JNI declarations written from the published JNI specification
// ("JNI Functions" chapter: interface function table indices and signatures).
// No jni.h is available where this library is built; only what the
// GPUSortingSearch natives need is declared.  JNIEnv is a pointer to a
// structure whose first member points at the function table.
#pragma once

#include <stdint.h>

typedef int32_t jint;
typedef int64_t jlong;
typedef int8_t jbyte;
typedef uint8_t jboolean;
typedef float jfloat;
typedef jint jsize;

typedef struct _jobject *jobject;
typedef jobject jclass;
typedef jobject jthrowable;
typedef jobject jarray;
typedef jarray jbyteArray;
typedef jarray jintArray;
typedef jarray jfloatArray;
typedef jarray jobjectArray;

#define JNI_ABORT 2
#define JNIEXPORT __attribute__((visibility("default")))
#define JNICALL

struct JNINativeInterface_ {
    void *functions[235];
};

struct JNIEnv_ {
    const struct JNINativeInterface_ *functions;
};
typedef struct JNIEnv_ JNIEnv;

// Function-table indices (JNI specification, Interface Function Table).
enum JniIndex {
    kJniFindClass = 6,
    kJniThrowNew = 14,
    kJniExceptionClear = 17,
    kJniDeleteLocalRef = 23,
    kJniGetArrayLength = 171,
    kJniNewObjectArray = 172,
    kJniSetObjectArrayElement = 174,
    kJniNewByteArray = 176,
    kJniNewIntArray = 179,
    kJniGetByteArrayRegion = 200,
    kJniGetIntArrayRegion = 203,
    kJniSetByteArrayRegion = 208,
    kJniSetIntArrayRegion = 211,
    kJniSetFloatArrayRegion = 213,
    kJniExceptionCheck = 228,
};

namespace jni {

template <typename Fn>
inline Fn fn(JNIEnv *env, int index) {
    return reinterpret_cast<Fn>(env->functions->functions[index]);
}

inline jclass FindClass(JNIEnv *env, const char *name) {
    return fn<jclass (*)(JNIEnv *, const char *)>(env, kJniFindClass)(env, name);
}
inline jint ThrowNew(JNIEnv *env, jclass cls, const char *msg) {
    return fn<jint (*)(JNIEnv *, jclass, const char *)>(env, kJniThrowNew)(env, cls, msg);
}
inline void ExceptionClear(JNIEnv *env) { fn<void (*)(JNIEnv *)>(env, kJniExceptionClear)(env); }
inline jboolean ExceptionCheck(JNIEnv *env) { return fn<jboolean (*)(JNIEnv *)>(env, kJniExceptionCheck)(env); }
inline void DeleteLocalRef(JNIEnv *env, jobject o) {
    fn<void (*)(JNIEnv *, jobject)>(env, kJniDeleteLocalRef)(env, o);
}
inline jsize GetArrayLength(JNIEnv *env, jarray a) {
    return fn<jsize (*)(JNIEnv *, jarray)>(env, kJniGetArrayLength)(env, a);
}
inline jobjectArray NewObjectArray(JNIEnv *env, jsize len, jclass cls, jobject init) {
    return fn<jobjectArray (*)(JNIEnv *, jsize, jclass, jobject)>(env, kJniNewObjectArray)(env, len, cls, init);
}
inline void SetObjectArrayElement(JNIEnv *env, jobjectArray a, jsize i, jobject v) {
    fn<void (*)(JNIEnv *, jobjectArray, jsize, jobject)>(env, kJniSetObjectArrayElement)(env, a, i, v);
}
inline jbyteArray NewByteArray(JNIEnv *env, jsize len) {
    return fn<jbyteArray (*)(JNIEnv *, jsize)>(env, kJniNewByteArray)(env, len);
}
inline jintArray NewIntArray(JNIEnv *env, jsize len) {
    return fn<jintArray (*)(JNIEnv *, jsize)>(env, kJniNewIntArray)(env, len);
}
inline void GetByteArrayRegion(JNIEnv *env, jbyteArray a, jsize start, jsize len, jbyte *buf) {
    fn<void (*)(JNIEnv *, jbyteArray, jsize, jsize, jbyte *)>(env, kJniGetByteArrayRegion)(env, a, start, len, buf);
}
inline void SetByteArrayRegion(JNIEnv *env, jbyteArray a, jsize start, jsize len, const jbyte *buf) {
    fn<void (*)(JNIEnv *, jbyteArray, jsize, jsize, const jbyte *)>(env, kJniSetByteArrayRegion)(env, a, start, len,
                                                                                              buf);
}
inline void GetIntArrayRegion(JNIEnv *env, jintArray a, jsize start, jsize len, jint *buf) {
    fn<void (*)(JNIEnv *, jintArray, jsize, jsize, jint *)>(env, kJniGetIntArrayRegion)(env, a, start, len, buf);
}
inline void SetFloatArrayRegion(JNIEnv *env, jfloatArray a, jsize start, jsize len, const jfloat *buf) {
    fn<void (*)(JNIEnv *, jfloatArray, jsize, jsize, const jfloat *)>(env, kJniSetFloatArrayRegion)(env, a, start,
                                                                                                  len, buf);
}
inline void SetIntArrayRegion(JNIEnv *env, jintArray a, jsize start, jsize len, const jint *buf) {
    fn<void (*)(JNIEnv *, jintArray, jsize, jsize, const jint *)>(env, kJniSetIntArrayRegion)(env, a, start, len, buf);
}

}  // namespace jni
