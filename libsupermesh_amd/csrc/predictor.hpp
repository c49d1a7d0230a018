// predictor.hpp -- the learned output-size predictor's parameter layout and
// kernel arguments (predictor.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "lsm_runtime.hpp"

namespace lsm {

// FP32 master parameters, in this order: W1 [64][24], b1 [64], W2 [64][64],
// b2 [64], W3 [64][64], b3 [64], W4 [33][64], b4 [33] (W [out][in]).
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + 64 * 24, kOffW2 = kOffB1 + 64, kOffB2 = kOffW2 + 64 * 64,
              kOffW3 = kOffB2 + 64, kOffB3 = kOffW3 + 64 * 64, kOffW4 = kOffB3 + 64,
              kOffB4 = kOffW4 + 33 * 64, kNumParams = kOffB4 + 33;

struct PredictArgs {
  long nwork;
  const int* work;  // pair of each work item (null: item = pair)
  const int* rows;
  const int* cols;
  const double* acoords;
  const double* bcoords;
  const __bf16* weights;  // forward layouts
  const float* bias;      // [4][64]
  float levels[4];           // quantiles, highest first
  unsigned* budget;          // [nwork]: byte k = the budget at levels[k]
  unsigned* spare;           // one word that the lanes past the end store to
  unsigned long long* sums;  // [5]: budget sum per level; 1024 x expected count sum
};

struct TrainArgs {
  long batch;
  const int* index;  // [batch] sample ids of this step
  const __bf16* features;
  const int* labels;
  __bf16* weights;
  float* bias;
  float* master;
  float* adam_m;
  float* adam_v;
  float* grad;
  float* loss;
  __bf16 *ht0, *ht1, *ht2, *ht3, *dt1, *dt2, *dt3, *dt4;
};

void predictor_set(Context& C, const float* params);
void predictor_get(Context& C, float* params);
void predictor_init(Context& C, unsigned seed);
void predictor_collect(Context& C, long nwork, const int* work, const int* rows, const int* cols,
                       const int* counts, const double* acoords, const double* bcoords, long max_samples);
void predictor_add_samples(Context& C, long n, const float* features, const int* labels);
void predictor_train(Context& C, int epochs, int batch, float lr, unsigned seed, float* loss_out);
void predict_budgets(Context& C, const PredictArgs& args);
// The four budget levels for a quantile q: q and, below it, 0.9, 0.7, 0.5.
void budget_levels(float q, float* levels);

}  // namespace lsm
