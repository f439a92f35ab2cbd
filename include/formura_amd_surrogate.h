/* formura_amd_surrogate.h -- the learned surrogate for periodic 2D
 * Gray-Scott (libsurrogate.so): a residual CNN of five circular 3x3
 * convolutions (4 -> 48 -> 48 -> 48 -> 48 -> 2, GELU) on BF16 MFMA with FP32
 * accumulation that advances a 256^2 field by 32 stencil steps per pass,
 * with its forward and backward kernels, loss and Adam.
 *
 * No counterpart in the reference: the reference generates the plain stencil
 * only.  Callers: formura_amd/surrogate.py (training, checked rollout,
 * evaluation).  All functions return 0 (or a value) on success and -1 on a
 * bad argument; HIP errors abort with a message.  One process-global network
 * per process; not re-entrant.
 *
 * Layouts: state [B][2][H][W] FP32 (u, v planes), fk [B][2] FP32 (F, k of
 * each instance), device pointers unless stated.
 */
#pragma once
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* number of FP32 parameters (weights and biases of the five layers) */
int Sur_Num_Params(void);
/* allocate the network for images H x W (W a multiple of 64, H a multiple of
 * 16, H * W * 48 < 2^31) and batches of up to maxB instances */
int Sur_Init(int H, int W, int maxB);
/* parameters and gradients, host arrays of Sur_Num_Params() floats */
int Sur_Set_Params(const float *host);
int Sur_Get_Params(float *host);
int Sur_Get_Grads(float *host);
/* per-channel output scale s (2 floats, host): out = state + s * z5 */
int Sur_Set_Scale(const float *s2);
/* convolution kernel (all give the same bits): 0 fragments from global
 * memory, 1 LDS tile of 4 rows, 2 LDS tile of 8 rows (default) */
int Sur_Set_Conv(int kind);
/* one pass (32 steps) for B instances: out = net(state, fk) */
int Sur_Forward(const float *state, const float *fk, float *out, int B);
/* forward, squared-error loss against target and all gradients (no update) */
int Sur_Gradients(const float *state, const float *fk, const float *target, int B, double *loss);
/* forward, loss, gradients and one Adam step with learning rate lr */
int Sur_Train_Step(const float *state, const float *fk, const float *target, int B, float lr, double *loss);
/* device memory and copies (kind 0 host to device, 1 device to host, 2
 * device to device) */
void *Sur_Alloc(size_t bytes);
int Sur_Free(void *p);
int Sur_Copy(void *dst, const void *src, size_t bytes, int kind);
/* gather nptr device arrays of n_each floats (addresses in host_ptrs) into dst */
int Sur_Gather(const unsigned long long *host_ptrs, float *dst, int nptr, size_t n_each);
/* the inverse: scatter src into nptr device arrays of n_each floats; returns
 * after the copies are done (they are visible to any stream) */
int Sur_Scatter(const unsigned long long *host_ptrs, const float *src, int nptr, size_t n_each);
/* x += N(0, sigma^2) noise (counter-based, from seed) */
int Sur_Add_Noise(float *x, size_t n, float sigma, unsigned long long seed);
/* relative L2 distance per instance of B arrays of n_each floats (device
 * pointers a, b; host output rel[B]): ||a_i - b_i|| / ||b_i||, FP64 sums */
int Sur_Rel_L2(const float *a, const float *b, int B, size_t n_each, double *rel);
/* milliseconds of `reps` forward passes (HIP events on the network's stream) */
float Sur_Time_Forward(const float *state, const float *fk, float *out, int B, int reps);

#ifdef __cplusplus
}
#endif
