/*
 * mpifdtd_surrogate.h -- the reflectance surrogate's MLP on the GPU.
 *
 * No counterpart in the reference: it sweeps geometries one FDTD run at a
 * time (main.c:114-138, models_isFinish).  The surrogate maps 12 morpho-scale
 * geometry parameters plus the wavelength (each scaled to [0, 1]) to the
 * far-field reflectance at 180 back-hemisphere angles; FDTD spectra train it
 * and measure its error, and FDTD answers the queries on which an ensemble of such
 * networks disagrees.
 *
 * The network is a stack of dense layers, GELU between them, linear output.
 * Its products run in BF16 on the matrix cores with FP32 accumulation; master
 * weights, Adam moments, biases and the loss are FP32.  Widths and the batch
 * are padded to multiples of 64 inside the library; the caller passes real
 * sizes.  Data arguments marked "device" are HIP device pointers (from
 * fdtd_hip_alloc); the others are host memory.  All calls are on the null
 * stream and return when their work is done.
 */
#ifndef MPIFDTD_SURROGATE_H
#define MPIFDTD_SURROGATE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlp_net mlp_net;

/* nlayers dense layers of widths dims[0] -> dims[1] -> ... -> dims[nlayers]
 * (at most 8 layers, widths at most 4096), batches of up to batch_max rows;
 * weights U(-1/sqrt(fan_in), 1/sqrt(fan_in)) and zero biases from a
 * SplitMix64 stream of `seed` */
int mlp_create(int nlayers, const int *dims, int batch_max, unsigned long long seed, mlp_net **out);
int mlp_destroy(mlp_net *net);
/* y[B][dims[nlayers]] = net(x[B][dims[0]]) (device, FP32) */
int mlp_forward(mlp_net *net, const float *x, int B, float *y);
/* gradients of loss = mean over the B x dims[nlayers] outputs of (net(x) - t)^2
 * (x, t device FP32) into the library's gradient buffers; *loss (host) if not NULL */
int mlp_grads(mlp_net *net, const float *x, const float *t, int B, float *loss);
/* mlp_grads followed by one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) */
int mlp_train_step(mlp_net *net, const float *x, const float *t, int B, float lr, float *loss);
/* one pass over rows x rows of x[rows][dims[0]], t[rows][dims[L]] (device,
 * FP32) in shuffled batches of `batch` (a Fisher-Yates permutation from a
 * SplitMix64 stream of `seed`), one Adam step per batch, the steps queued
 * without a host round trip; *loss (host, if not NULL) = the epoch's mean batch
 * loss weighted by batch size */
int mlp_train_epoch(mlp_net *net, const float *x, const float *t, int rows, int batch, unsigned long long seed,
                    float lr, float *loss);
/* layer l's weights w[dims[l+1]][dims[l]] and bias b[dims[l+1]] (host, FP32) */
int mlp_get_params(mlp_net *net, int l, float *w, float *b);
int mlp_set_params(mlp_net *net, int l, const float *w, const float *b);
/* layer l's gradients from the last mlp_grads / mlp_train_step (host, FP32) */
int mlp_get_grads(mlp_net *net, int l, float *dw, float *db);
/* C[M][N] = A B in BF16 with FP32 accumulation, for the layout tests: A is
 * a[M][K] (ta = 0) or a[K][M] (ta = 1), B is b[N][K] (tb = 0) or b[K][N]
 * (tb = 1), all row-major device arrays of bf16 bit patterns; M, N multiples
 * of 64, K a multiple of 32 */
int mlp_gemm_bf16(const unsigned short *a, const unsigned short *b, float *c, int M, int N, int K, int ta, int tb);

#ifdef __cplusplus
}
#endif
#endif
