// ref_shim.cpp -- C entry points onto the reference's own validation and
// random-vector code (LinearScan/validation.cpp, LinearScan/randomVectorGenerator.cpp),
// compiled together with those files into oracle/_ref/libref.so.  Test
// infrastructure only: used to check the oracle's restatements.
#include <cstring>
#include "randomVectorGenerator.h"
#include "validation.h"

extern "C" {

// validation.cpp:201-208 runValidation (reads a truth file, fills
// container[0] = recall, container[1] = distance ratio).
void ref_run_validation(char* truths, float* container, Point* results, int n_queries,
                        int k, int reportK) {
    runValidation(truths, container, results, n_queries, k, reportK);
}

// randomVectorGenerator.cpp:32-53
void ref_random_vectors(float* out, int n) { generateRandomVectors(out, n, false); }

// randomVectorGenerator.cpp:56-74 (function-static buffer sized by the first
// call: call once per process).
void ref_random_vectors_sd(float* out, int n, int sketchDim) {
    float* p = generateRandomVectors(n, sketchDim, false);
    std::memcpy(out, p, sizeof(float) * (size_t)n);
}

// randomVectorGenerator.cpp:76-87
void ref_random_signs(int* out, int n) { generateRandomOnePlusMinusVector(n, out); }
}
