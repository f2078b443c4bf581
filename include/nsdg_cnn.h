/* C ABI of the learned warm start of the mEVP solver (gfx950).
 *
 * No counterpart in the reference: neXtSIM_DG runs mEVP from the previous timestep's
 * velocity and stress for a fixed subcycle count (VPCGDynamicsKernel.hpp:63-94). This
 * interface adds a small convolutional network on the element grid that predicts the
 * converged velocity of a timestep from the state and forcing, trained on the GPU from
 * pairs the GPU solver generates, and a residual-gated solve that starts from it. The
 * gated start is opt-in: nsdg_dyn_update never calls it.
 *
 * Network: 4 layers of 3x3 convolutions ('same' zero padding), 32 channels each, ReLU
 * between layers, linear output; inputs and activations BF16 in NHWC layout
 * ([image][row][column][32 channels]), weights BF16 copies of FP32 master values,
 * every product accumulated in FP32 on MFMA (v_mfma_f32_16x16x32_bf16). Outputs: 8 FP32
 * channels per element (the velocity change at the element's 4 owned CG nodes).
 *
 * Parameters (FP32, host view, nsdg_cnn_get_params): per layer l = 0..3, the weights
 * W[co][ci][ky][kx] (32 x 32 x 3 x 3, PyTorch's conv2d order, rows co >= 8 of the last
 * layer unused and zero), then the bias b[co] (32). All device pointers below are plain
 * device addresses; every call is ordered on the library's own stream and returns after
 * the work is done.
 */
#ifndef NSDG_CNN_H
#define NSDG_CNN_H
#include <stddef.h>
#include <stdint.h>

#include "nsdg_dyn.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NSDG_CNN_CHANNELS 32
#define NSDG_CNN_OUTPUTS 8
#define NSDG_CNN_LAYERS 4

typedef struct nsdg_cnn nsdg_cnn;

/* He-normal initial weights from a seeded counter RNG (the last layer scaled by 0.1). */
nsdg_cnn* nsdg_cnn_create(uint64_t seed);
int nsdg_cnn_destroy(nsdg_cnn* net);
int nsdg_cnn_param_count(nsdg_cnn* net, size_t* count);
int nsdg_cnn_get_params(nsdg_cnn* net, float* host);
int nsdg_cnn_set_params(nsdg_cnn* net, const float* host);
/* y[p][8] (FP32) = net(x) for n images of ny x nx (nx a multiple of 16); x BF16 NHWC. */
int nsdg_cnn_forward(nsdg_cnn* net, const uint16_t* x, int n, int ny, int nx, float* y);
/* Loss L = sum_p w_p sum_{c<8} (y_pc - t_pc)^2 / (8 n ny nx) and its gradient with respect
 * to every parameter (kept for nsdg_cnn_adam; copied to grad_host when that is not NULL,
 * in the parameter order above). w may be NULL (all ones). */
int nsdg_cnn_loss_grad(nsdg_cnn* net, const uint16_t* x, const float* t, const float* w, int n, int ny, int nx,
    double* loss, float* grad_host);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the last gradient. */
int nsdg_cnn_adam(nsdg_cnn* net, float lr);
/* Per-kernel timing with HIP events, as nsdg_dyn_set_profiling / nsdg_dyn_kernel_stats. */
int nsdg_cnn_set_profiling(nsdg_cnn* net, int on);
int nsdg_cnn_kernel_stats(nsdg_cnn* net, const char* kernel, int* count, double* total_ms);

/* ---- device buffers for the training data (plain device memory of this library) --- */
int nsdg_device_alloc(size_t bytes, void** ptr);
int nsdg_device_free(void* ptr);
int nsdg_device_upload(void* dst, const void* src, size_t bytes);
int nsdg_device_download(void* dst, const void* src, size_t bytes);
/* dst item i <- src item index[i] (items of item_bytes; index on the host), for batches. */
int nsdg_device_gather(void* dst, const void* src, const int64_t* index, int n, size_t item_bytes);

/* ---- coupling to the dynamics kernel (element grid nx x ny of an initialised CG2 kernel,
 * after begin_iteration) ---------------------------------------------------------------
 * Network input for the current state: per element h and A (cell means), the ocean
 * mask, wind and ocean current at the element's centre node, u0 and v0 at its 4 owned CG
 * nodes (ix, iy) = (2i + a, 2j + b), a, b in {0, 1}, the stress means and the ice
 * strength, scaled to order one; 13 channels zero. x: BF16 [ny][nx][32] (device). */
int nsdg_dyn_features(nsdg_dyn* d, uint16_t* x);
/* Training target (u - u0, v - v0) / 0.1 m/s at the 4 owned nodes, [ny][nx][8] FP32,
 * and the loss weight per element (1 ocean, 0 land), [ny][nx] FP32 (device). */
int nsdg_dyn_velocity_target(nsdg_dyn* d, float* t, float* w);
/* u, v <- u0, v0 + 0.1 m/s * y at the owned non-Dirichlet nodes (other nodes keep u0). */
int nsdg_dyn_apply_velocity(nsdg_dyn* d, const float* y);
/* Warm start of the current timestep: predicted velocity, its VP stress, and the
 * residual gate: when r(predicted) > r(cold start: u0 and the previous stress), the cold
 * start is restored. *used = 1 when the prediction was kept. */
int nsdg_dyn_warm_start(nsdg_dyn* d, nsdg_cnn* net, double* r_predicted, double* r_cold, int* used);
/* Subcycles until r <= r_target or max_subcycles have run, in chunks of at most
 * check_every: once r has fallen over a chunk, the next chunk ends where its rate of
 * decrease reaches r_target (at least one subcycle); an r within 1% of r_target that did
 * not fall over the last chunk gets chunks of at most 5. r_start: the residual of the
 * current state when the caller has it (after nsdg_dyn_warm_start: r_predicted if the
 * prediction was kept, else r_cold), which saves one residual evaluation; a negative
 * value has it computed. */
int nsdg_dyn_subcycles_to_residual(nsdg_dyn* d, double r_target, int max_subcycles, int check_every, double r_start,
    int* done, double* r_final);

#ifdef __cplusplus
}
#endif
#endif /* NSDG_CNN_H */
