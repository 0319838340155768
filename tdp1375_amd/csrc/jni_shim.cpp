// JNI natives of br.unb.cic.tdp.base.GPUSortingSearch over the C ABI.
//
// Same five symbols, descriptors and behaviour as the reference's
// src/native/jni/jni_sorter.cu:16-150: inputs are copied, bytes are read
// unsigned, results are fresh Java arrays, and every failure becomes a
// java.lang.RuntimeException carrying the native message while the call
// returns null.  cancelSort never throws.
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/tdp1375.h"
#include "jni_min.h"

namespace {

// jni_sorter.cu:16-23
void throw_runtime(JNIEnv *env, const char *message) {
    if (jni::ExceptionCheck(env)) jni::ExceptionClear(env);
    jclass cls = jni::FindClass(env, "java/lang/RuntimeException");
    if (cls) {
        jni::ThrowNew(env, cls, message);
        jni::DeleteLocalRef(env, cls);
    }
}

bool read_bytes(JNIEnv *env, jbyteArray arr, jsize len, std::vector<uint8_t> &out) {
    out.assign((size_t)(len > 0 ? len : 0), 0);
    if (len <= 0) return true;
    jni::GetByteArrayRegion(env, arr, 0, len, reinterpret_cast<jbyte *>(out.data()));
    return !jni::ExceptionCheck(env);
}

}  // namespace

extern "C" {

JNIEXPORT void JNICALL Java_br_unb_cic_tdp_base_GPUSortingSearch_initSlots(JNIEnv *env, jclass, jint totalSlots,
                                                                           jint devicesCount, jlong queueBytesBudget,
                                                                           jint dedupTableSize) {
    if (tdp_init_slots(totalSlots, devicesCount, queueBytesBudget, dedupTableSize) != TDP_OK)
        throw_runtime(env, tdp_last_error());
}

JNIEXPORT void JNICALL Java_br_unb_cic_tdp_base_GPUSortingSearch_destroySlots(JNIEnv *env, jclass) {
    if (tdp_destroy_slots() != TDP_OK) throw_runtime(env, tdp_last_error());
}

JNIEXPORT jobjectArray JNICALL Java_br_unb_cic_tdp_base_GPUSortingSearch_searchForSortingSeq(
    JNIEnv *env, jclass, jbyteArray javaPi, jbyteArray javaSpi, jint initialEvenCycles, jfloat minRate, jint maxMoves,
    jboolean fullSorting, jint slot) {
    const jsize n = jni::GetArrayLength(env, javaPi);
    if (jni::GetArrayLength(env, javaSpi) < n) {
        throw_runtime(env, "performGpuSortingSearch: inconsistent input sizes");
        return nullptr;
    }
    std::vector<uint8_t> pi, spi;
    if (!read_bytes(env, javaPi, n, pi) || !read_bytes(env, javaSpi, n, spi)) return nullptr;
    int32_t moves[TDP_MAX_DEPTH * 3];
    int32_t depth = 0;
    if (tdp_search(slot, pi.data(), spi.data(), n, initialEvenCycles, minRate, maxMoves, fullSorting ? 1 : 0, moves,
                   &depth) != TDP_OK) {
        throw_runtime(env, tdp_last_error());
        return nullptr;
    }
    // jni_sorter.cu:85-95: int[depth][3] of positions
    jclass intArrayClass = jni::FindClass(env, "[I");
    jobjectArray outer = jni::NewObjectArray(env, depth, intArrayClass, nullptr);
    for (jsize m = 0; m < depth; ++m) {
        jintArray inner = jni::NewIntArray(env, 3);
        jni::SetIntArrayRegion(env, inner, 0, 3, moves + 3 * m);
        jni::SetObjectArrayElement(env, outer, m, inner);
        jni::DeleteLocalRef(env, inner);
    }
    jni::DeleteLocalRef(env, intArrayClass);
    return outer;
}

JNIEXPORT void JNICALL Java_br_unb_cic_tdp_base_GPUSortingSearch_cancelSort(JNIEnv *, jclass, jint slot) {
    tdp_cancel(slot);
}

JNIEXPORT jobjectArray JNICALL Java_br_unb_cic_tdp_base_GPUSortingSearch_filterPiPermutations(
    JNIEnv *env, jclass, jbyteArray javaCycle, jint cycleLen, jbyteArray javaTriple, jint tripleLen, jint n,
    jboolean skip2MoveCheck) {
    std::vector<uint8_t> cycle, triple;
    if (cycleLen < 0 || tripleLen < 0 || jni::GetArrayLength(env, javaCycle) < cycleLen ||
        jni::GetArrayLength(env, javaTriple) < tripleLen) {
        throw_runtime(env, "invalid cycle or triple length");
        return nullptr;
    }
    if (!read_bytes(env, javaCycle, cycleLen, cycle) || !read_bytes(env, javaTriple, tripleLen, triple)) return nullptr;
    int64_t cap = TDP_FILTER_MAX_OUTPUT;
    std::vector<uint8_t> rows;
    if (n >= 1 && n <= TDP_MAX_N) rows.resize((size_t)cap * (size_t)n);
    int64_t total = 0, written = 0;
    if (tdp_filter(cycle.data(), cycleLen, triple.data(), tripleLen, n, skip2MoveCheck ? 1 : 0,
                   rows.empty() ? nullptr : rows.data(), rows.empty() ? 0 : cap, &total, &written) != TDP_OK) {
        throw_runtime(env, tdp_last_error());
        return nullptr;
    }
    // pi_filter.cu:233-259: at most 2^20 rows of n bytes (here the first 2^20 by rank)
    jclass byteArrayClass = jni::FindClass(env, "[B");
    jobjectArray outer = jni::NewObjectArray(env, (jsize)written, byteArrayClass, nullptr);
    for (jsize r = 0; r < (jsize)written; ++r) {
        jbyteArray inner = jni::NewByteArray(env, n);
        jni::SetByteArrayRegion(env, inner, 0, n, reinterpret_cast<const jbyte *>(rows.data() + (size_t)r * n));
        jni::SetObjectArrayElement(env, outer, r, inner);
        jni::DeleteLocalRef(env, inner);
    }
    jni::DeleteLocalRef(env, byteArrayClass);
    return outer;
}

}  // extern "C"
