// common.cpp -- error state, launch checks, phase timers and the host random
// streams of the reference.
#include <random>

#include "lshgpu_internal.hpp"

namespace lshgpu {

std::string& last_error() {
    static thread_local std::string msg;
    return msg;
}

void fail(int code, const std::string& msg) { throw Error(code, msg); }

void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(LSH_ERR_DEVICE, std::string("launch of ") + what + " failed: " + hipGetErrorString(e));
}

void set_dynamic_lds(const void* kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    LSH_REQUIRE(bytes <= 160 * 1024, LSH_ERR_UNSUPPORTED, "workgroup LDS request above 160 KiB");
    LSH_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// randomVectorGenerator.cpp:32-53 and :56-74: a default-seeded
// std::default_random_engine feeding std::normal_distribution<double>(0,
// stddev), every draw cast to float.  Calling the same library functions as
// the reference (libstdc++ on Linux) reproduces its hyperplanes bit for bit.
std::vector<float> normal_stream(int64_t n, double stddev) {
    std::vector<float> v((size_t)n);
    std::default_random_engine gen;
    std::normal_distribution<double> dist(0.0, stddev);
    for (int64_t i = 0; i < n; ++i) v[i] = (float)dist(gen);
    return v;
}

// randomVectorGenerator.cpp:76-87: uniform_int_distribution<int>(0, 2),
// 0 -> -1 and 1, 2 -> +1.
std::vector<int32_t> sign_stream(int64_t n) {
    std::vector<int32_t> v((size_t)n);
    std::default_random_engine gen;
    std::uniform_int_distribution<int> dist(0, 2);
    for (int64_t i = 0; i < n; ++i) v[i] = dist(gen) == 0 ? -1 : 1;
    return v;
}

// weightedMinHash.cuh:231-250: uniform_int_distribution<int>(0, 1).
std::vector<uint8_t> bit_stream(int64_t n) {
    std::vector<uint8_t> v((size_t)n);
    std::default_random_engine gen;
    std::uniform_int_distribution<int> dist(0, 1);
    for (int64_t i = 0; i < n; ++i) v[i] = (uint8_t)dist(gen);
    return v;
}

PhaseTimer::~PhaseTimer() {
    for (auto& r : recs_) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto& o : open_) (void)hipEventDestroy(o.second);
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
}

hipEvent_t PhaseTimer::take() {
    if (!pool_.empty()) {
        const hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    LSH_HIP(hipEventCreate(&e));
    return e;
}

void PhaseTimer::enable(uint32_t mask) {
    for (auto& r : recs_) {
        pool_.push_back(r.a);
        pool_.push_back(r.b);
    }
    recs_.clear();
    for (auto& o : open_) pool_.push_back(o.second);
    open_.clear();
    mask_ = mask;
}

void PhaseTimer::begin(int phase, hipStream_t s) {
    if (phase < 0 || phase >= 32 || !((mask_ >> phase) & 1u)) return;
    const hipEvent_t e = take();
    LSH_HIP(hipEventRecord(e, s));
    open_.push_back({phase, e});
}

void PhaseTimer::end(int phase, hipStream_t s) {
    if (phase < 0 || phase >= 32 || !((mask_ >> phase) & 1u)) return;
    for (size_t i = open_.size(); i-- > 0;) {
        if (open_[i].first == phase) {
            const hipEvent_t e = take();
            LSH_HIP(hipEventRecord(e, s));
            recs_.push_back({phase, open_[i].second, e});
            open_.erase(open_.begin() + (long)i);
            return;
        }
    }
}

void PhaseTimer::collect(double* ms, int nphase) {
    for (int p = 0; p < nphase; ++p) ms[p] = 0.0;
    for (auto& r : recs_) {
        LSH_HIP(hipEventSynchronize(r.b));
        float t = 0.0f;
        LSH_HIP(hipEventElapsedTime(&t, r.a, r.b));
        if (r.phase >= 0 && r.phase < nphase) ms[r.phase] += t;
        pool_.push_back(r.a);
        pool_.push_back(r.b);
    }
    recs_.clear();
}

}  // namespace lshgpu
