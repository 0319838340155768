// JNI natives of br.unb.cic.tdp.base.GPUSortingSearch over the C ABI.
//
// Same five symbols, descriptors and behaviour as the reference's
// src/native/jni/jni_sorter.cu:16-150: inputs are copied, bytes are read
// unsigned, results are fresh Java arrays, and every failure becomes a
// java.lang.RuntimeException carrying the native message while the call
// returns null.  cancelSort never throws.  Two more natives (no reference
// counterpart) expose the batched search and the batched canonical forms
// to the same class with the same conventions.
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

bool read_ints(JNIEnv *env, jintArray arr, jsize len, std::vector<int32_t> &out) {
    out.assign((size_t)(len > 0 ? len : 0), 0);
    if (len <= 0) return true;
    jni::GetIntArrayRegion(env, arr, 0, len, reinterpret_cast<jint *>(out.data()));
    return !jni::ExceptionCheck(env);
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

// Many configurations in one persistent launch (tdp_search_batch).  Java:
//   private static native int[] searchBatch(byte[] pis, byte[] spis, int[] ns, int[] initialEvenCycles,
//                                           float minRate, int maxMoves, boolean fullSorting, int flags,
//                                           int slot);
// pis/spis hold count x 64 bytes (configuration c in bytes 64c .. 64c + ns[c]);
// the result holds per configuration found, depth and 15 position triples
// (47 ints), a depth of -1 marking a configuration left undecided by
// TDP_FLAG_SOFT_DEADLINE.
JNIEXPORT jintArray JNICALL Java_br_unb_cic_tdp_base_GPUSortingSearch_searchBatch(
    JNIEnv *env, jclass, jbyteArray javaPis, jbyteArray javaSpis, jintArray javaNs, jintArray javaE0s, jfloat minRate,
    jint maxMoves, jboolean fullSorting, jint flags, jint slot) {
    const jsize count = jni::GetArrayLength(env, javaNs);
    if (jni::GetArrayLength(env, javaE0s) < count || jni::GetArrayLength(env, javaPis) < 64 * count ||
        jni::GetArrayLength(env, javaSpis) < 64 * count) {
        throw_runtime(env, "performGpuSortingSearch: inconsistent input sizes");
        return nullptr;
    }
    std::vector<uint8_t> pis, spis;
    std::vector<int32_t> ns, e0s;
    if (!read_bytes(env, javaPis, 64 * count, pis) || !read_bytes(env, javaSpis, 64 * count, spis) ||
        !read_ints(env, javaNs, count, ns) || !read_ints(env, javaE0s, count, e0s))
        return nullptr;
    std::vector<float> rates((size_t)count, minRate);
    std::vector<int32_t> depths((size_t)count, maxMoves), fulls((size_t)count, fullSorting ? 1 : 0);
    std::vector<tdp_case_result> res((size_t)count);
    if (tdp_search_batch(slot, count, pis.data(), spis.data(), ns.data(), e0s.data(), rates.data(), depths.data(),
                         fulls.data(), (uint32_t)flags, res.data()) != TDP_OK) {
        throw_runtime(env, tdp_last_error());
        return nullptr;
    }
    constexpr int kPer = 2 + 3 * TDP_MAX_DEPTH;
    std::vector<jint> out((size_t)count * kPer, 0);
    for (jsize c = 0; c < count; ++c) {
        jint *o = out.data() + (size_t)c * kPer;
        o[0] = res[c].found;
        o[1] = res[c].depth;
        for (int m = 0; m < res[c].depth && m < TDP_MAX_DEPTH; ++m)
            for (int q = 0; q < 3; ++q) o[2 + 3 * m + q] = res[c].moves[m][q];
    }
    jintArray arr = jni::NewIntArray(env, (jsize)out.size());
    jni::SetIntArrayRegion(env, arr, 0, (jsize)out.size(), out.data());
    return arr;
}

// Canonical forms of many configurations in one launch (tdp_canonical_batch,
// for Configuration.getCanonical / hashCode, Configuration.java:156-177).
// Java:
//   private static native int canonicalBatch(byte[] spis, byte[] pis, int[] ns, float[] signatures,
//                                            int[] hashes, int[] index, byte[] canonicalSpi);
// spis/pis/canonicalSpi hold count x 64 bytes, signatures count x 64 floats;
// returns count.
JNIEXPORT jint JNICALL Java_br_unb_cic_tdp_base_GPUSortingSearch_canonicalBatch(
    JNIEnv *env, jclass, jbyteArray javaSpis, jbyteArray javaPis, jintArray javaNs, jfloatArray javaSignatures,
    jintArray javaHashes, jintArray javaIndex, jbyteArray javaCanonical) {
    const jsize count = jni::GetArrayLength(env, javaNs);
    if (jni::GetArrayLength(env, javaSpis) < 64 * count || jni::GetArrayLength(env, javaPis) < 64 * count ||
        jni::GetArrayLength(env, javaSignatures) < 64 * count || jni::GetArrayLength(env, javaHashes) < count ||
        jni::GetArrayLength(env, javaIndex) < count || jni::GetArrayLength(env, javaCanonical) < 64 * count) {
        throw_runtime(env, "canonicalBatch: inconsistent array sizes");
        return 0;
    }
    std::vector<uint8_t> spis, pis;
    std::vector<int32_t> ns;
    if (!read_bytes(env, javaSpis, 64 * count, spis) || !read_bytes(env, javaPis, 64 * count, pis) ||
        !read_ints(env, javaNs, count, ns))
        return 0;
    std::vector<float> sig((size_t)count * 64);
    std::vector<int32_t> hashes((size_t)count), index((size_t)count);
    std::vector<uint8_t> canon((size_t)count * 64);
    if (count > 0 && tdp_canonical_batch(count, spis.data(), pis.data(), ns.data(), sig.data(), hashes.data(),
                                         index.data(), canon.data()) != TDP_OK) {
        throw_runtime(env, tdp_last_error());
        return 0;
    }
    if (count > 0) {
        jni::SetFloatArrayRegion(env, javaSignatures, 0, 64 * count, sig.data());
        jni::SetIntArrayRegion(env, javaHashes, 0, count, hashes.data());
        jni::SetIntArrayRegion(env, javaIndex, 0, count, index.data());
        jni::SetByteArrayRegion(env, javaCanonical, 0, 64 * count, reinterpret_cast<const jbyte *>(canon.data()));
    }
    return count;
}

}  // extern "C"
