// Host-side pieces of the Monte-Carlo estimator shared by the device path
// (capi.hip) and the CPU path (host_slq.cpp). No HIP types.
#pragma once

#include <cstdint>

#include "slq_hip.h"

namespace slq {

void set_error(const char* fmt, ...);

// ConvergenceTools::check_convergence, convergence_tools.cpp:94-194: error
// and convergence of every inquiry over the samples processed so far (rows
// processed[0 .. num_processed) of samples [*][q]); 1 when all converged.
int check_convergence(const double* samples, int q, int min_num_samples,
                      const int32_t* processed, int num_processed,
                      double confidence, double atol, double rtol,
                      double* error, int32_t* num_used, int32_t* converged);

// ConvergenceTools::average_estimates, convergence_tools.cpp:255-376: mean
// and error of every inquiry after removing outliers.
void average_estimates(const double* samples, int q, double confidence,
                       double outlier, const int32_t* num_used,
                       const int32_t* processed, int32_t* num_outliers,
                       double* trace, double* error);

// Seed drawn from the clock when the caller passes a negative seed.
int64_t clock_seed();

// Thread thread_id's xoshiro256** state for `seed` (SplitMix64 seeding and
// thread_id + 1 jumps), the state `k` steps after `base`, and one step.
void thread_state(int64_t seed, int thread_id, uint64_t out[4]);
void stream_at(const uint64_t base[4], uint64_t k, uint64_t out[4]);
uint64_t stream_next(uint64_t s[4]);

}  // namespace slq

// A valid matrix function code with the coefficients it takes.
bool func_ok(int func, const double* fp);
