// Surrogate network (ensemble of identical fully connected GELU networks).
#pragma once

#include "common.h"

namespace imped {

struct Mlp;

// E members, L layers of widths dims[0..L]; parameters per member laid out as
// W_0 [dims[1]][dims[0]], b_0, W_1, b_1, ... (padded to P, a multiple of 16).
Mlp* mlp_create(int E, int L, const int32_t* dims, int max_batch);
void mlp_destroy(Mlp* net);
int64_t mlp_param_count(const Mlp& net);
int64_t mlp_data_count(const Mlp& net);
int mlp_layers(const Mlp& net, int32_t* dims);
void mlp_set_params(Mlp& net, const float* params, hipStream_t s);
void mlp_get_params(Mlp& net, float* params, hipStream_t s);
void mlp_forward(Mlp& net, const float* x_host, int B, float* y_host, hipStream_t s);
// target_host, mean_host and l1_host may be NULL (no misfit; no read-back).
void mlp_predict(Mlp& net, const float* x_host, int B, const double* w_host, double scale, const double* target_host,
                 double* mean_host, double* dis_host, double* l1_host, hipStream_t s);
// mlp_predict in two halves: _begin enqueues (x is consumed before it
// returns), _end waits and reads the results; the caller may work between
// them but must not start another call on the network.
void mlp_predict_begin(Mlp& net, const float* x_host, int B, const double* w_host, double scale,
                       const double* target_host, double* mean_host, hipStream_t s);
void mlp_predict_end(Mlp& net, int B, bool with_target, double* dis_host, double* l1_host, hipStream_t s);
void mlp_gradients(Mlp& net, const float* x_host, const float* y_host, int B, float* g_host, double* loss,
                   hipStream_t s);
void mlp_set_data(Mlp& net, const float* X, const float* Y, int64_t n, hipStream_t s);
void mlp_train(Mlp& net, const int32_t* batches, int steps, int B, float lr, float b1, float b2, float eps,
               double* losses, hipStream_t s);

}  // namespace imped
