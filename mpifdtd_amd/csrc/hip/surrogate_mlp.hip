// surrogate_mlp.hip -- the reflectance surrogate's MLP: BF16 matrix-core
// kernels (forward, backward, Adam) and their host driver.
//
// No counterpart in the reference.  Layout: activations are row-major
// [batch][width] with the batch and the widths padded to multiples of 64 (the
// pads hold zeros and stay zero: padded weights and biases are zero and
// receive zero gradients).  Layer l keeps FP32 master weights
// W[N][K] (N = out, K = in), their BF16 copy for the products, the FP32 bias,
// Adam moments, and from the forward pass the FP32 pre-activation Z and the
// BF16 input A of the layer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "fdtd_hip.h"
#include "fdtd_internal.h"
#include "mpifdtd_surrogate.h"

#define HIPCHK(x)                                                                                    \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) return fdtd_fail(FDTD_EHIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

typedef unsigned short bf16_t;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16_t v) { return __uint_as_float((unsigned)v << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ bf16_t f2bf(float f)
{
  const unsigned u = __float_as_uint(f);
  return (bf16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// GELU(z) = z Phi(z), Phi the standard normal CDF; GELU'(z) = Phi(z) + z phi(z)
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z)
{
  const float Phi = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float phi = 0.39894228040143268f * expf(-0.5f * z * z);
  return Phi + z * phi;
}

enum { EPI_F32 = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_GELU_BWD = 3 };

struct GemmOut {
  float *c;           // FP32 result [M][ldc]
  bf16_t *h;          // EPI_BIAS_GELU: BF16 GELU(z); EPI_GELU_BWD: BF16 gradient
  const float *bias;  // EPI_BIAS, EPI_BIAS_GELU: [N]
  const float *zp;    // EPI_GELU_BWD: pre-activation of the layer below [M][ldc]
  int ldc;
};

// C[M][N] = sum_k A(m, k) B(k, n): A(m, k) = a[m lda + k] (TA 0) or a[k lda + m]
// (TA 1); B(k, n) = b[n ldb + k] (TB 0) or b[k ldb + n] (TB 1).  One workgroup
// per 64 x 64 tile of C, four wavefronts of 32 x 32, each 2 x 2 products of
// v_mfma_f32_16x16x16 bf16 per 16-deep slice; the operands pass through LDS
// as [row][k] tiles 32 deep.  M, N multiples of 64, K of 32 (the caller pads).
template <int TA, int TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const bf16_t *__restrict__ a, int lda, const bf16_t *__restrict__ b,
                                              int ldb, int K, GemmOut o)
{
  __shared__ __attribute__((aligned(16))) bf16_t sa[64][40];  // 32 k + 8 pad
  __shared__ __attribute__((aligned(16))) bf16_t sb[64][40];
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    if (TA == 0) {  // 64 rows of 32 k: 4 chunks of 8 per row
      const int r = tid >> 2, c = (tid & 3) * 8;
      *(uint4 *)&sa[r][c] = *(const uint4 *)&a[(size_t)(m0 + r) * lda + k0 + c];
    } else {  // 32 rows of k, 64 m each: 8 chunks of 8 per row, transposed into LDS
      const int kr = tid >> 3, c = (tid & 7) * 8;
      const uint4 v = *(const uint4 *)&a[(size_t)(k0 + kr) * lda + m0 + c];
      const bf16_t *pv = (const bf16_t *)&v;
#pragma unroll
      for (int j = 0; j < 8; j++) sa[c + j][kr] = pv[j];
    }
    if (TB == 0) {
      const int r = tid >> 2, c = (tid & 3) * 8;
      *(uint4 *)&sb[r][c] = *(const uint4 *)&b[(size_t)(n0 + r) * ldb + k0 + c];
    } else {
      const int kr = tid >> 3, c = (tid & 7) * 8;
      const uint4 v = *(const uint4 *)&b[(size_t)(k0 + kr) * ldb + n0 + c];
      const bf16_t *pv = (const bf16_t *)&v;
#pragma unroll
      for (int j = 0; j < 8; j++) sb[c + j][kr] = pv[j];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 32; kk += 16) {
      // 16x16x16 operands: lane l holds row (l mod 16), k = 4 (l / 16) + 0..3
      bf16x4 af[2], bq[2];
#pragma unroll
      for (int i = 0; i < 2; i++) af[i] = *(const bf16x4 *)&sa[wm + i * 16 + (lane & 15)][kk + 4 * (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; j++) bq[j] = *(const bf16x4 *)&sb[wn + j * 16 + (lane & 15)][kk + 4 * (lane >> 4)];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bq[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // result lane l, element r: row 4 (l / 16) + r, column l mod 16 of each 16 x 16 block
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + r, n = n0 + wn + j * 16 + (lane & 15);
        const size_t idx = (size_t)m * o.ldc + n;
        float v = acc[i][j][r];
        if (EPI == EPI_F32) {
          o.c[idx] = v;
        } else if (EPI == EPI_BIAS) {
          o.c[idx] = v + o.bias[n];
        } else if (EPI == EPI_BIAS_GELU) {
          v += o.bias[n];
          o.c[idx] = v;
          o.h[idx] = f2bf(gelu(v));
        } else {  // EPI_GELU_BWD
          v *= gelu_d(o.zp[idx]);
          o.c[idx] = v;
          o.h[idx] = f2bf(v);
        }
      }
}

// x[B][d] (FP32, real sizes) -> X[Bp][dp] (BF16, zero pads)
__global__ __launch_bounds__(256) void k_pack(const float *x, int B, int d, bf16_t *X, int Bp, int dp)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)Bp * dp) return;
  const int r = (int)(i / dp), c = (int)(i % dp);
  X[i] = (r < B && c < d) ? f2bf(x[(size_t)r * d + c]) : (bf16_t)0;
}

// y[B][d] (FP32, real sizes) from Y[Bp][dp]
__global__ __launch_bounds__(256) void k_unpack(const float *Y, int dp, int B, int d, float *y)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)B * d) return;
  y[i] = Y[(i / d) * dp + (i % d)];
}

// rows perm[0..B) of x[.][d] -> xs[B][d]
__global__ __launch_bounds__(256) void k_gather_rows(const float *x, int d, const int *perm, int B, float *xs)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)B * d) return;
  xs[i] = x[(size_t)perm[i / d] * d + (i % d)];
}

// dL/dY of the mean squared error over the B x d real outputs, FP32 and BF16
// (zero on the pads); the loss itself into *loss (zeroed by the caller)
__global__ __launch_bounds__(256) void k_mse_grad(const float *Y, int dp, const float *t, int B, int d, int Bp,
                                                  float *G, bf16_t *Gh, float *loss)
{
  __shared__ float red[4];
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  float l = 0.f;
  if (i < (size_t)Bp * dp) {
    const int r = (int)(i / dp), c = (int)(i % dp);
    float g = 0.f;
    if (r < B && c < d) {
      const float e = Y[i] - t[(size_t)r * d + c];
      l = e * e;
      g = 2.f * e / ((float)B * (float)d);
    }
    G[i] = g;
    Gh[i] = f2bf(g);
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_down(l, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) / ((float)B * (float)d));
}

// db[n] = sum over the batch of G[.][n]
__global__ __launch_bounds__(256) void k_colsum(const float *G, int Bp, int Np, float *db)
{
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= Np) return;
  float s = 0.f;
  for (int r = 0; r < Bp; r++) s += G[(size_t)r * Np + n];
  db[n] = s;
}

// Adam (Kingma and Ba 2015) with bias correction c1 = 1 - beta1^t, c2 = 1 - beta2^t;
// refreshes the BF16 copy when wh is not null
__global__ __launch_bounds__(256) void k_adam(float *w, const float *g, float *m, float *v, bf16_t *wh, size_t n, float lr,
                                              float c1, float c2)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi, vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
  w[i] = wi;
  if (wh) wh[i] = f2bf(wi);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *w, bf16_t *wh, size_t n)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) wh[i] = f2bf(w[i]);
}

// ------------------------------------------------------------------ host

#define MLP_MAX_LAYERS 8

struct MlpLayer {
  int K, N, Kp, Np;  // real and padded (multiples of 64) in / out widths
  float *w, *b, *dw, *db, *mw, *vw, *mb, *vb;
  bf16_t *wh;
  float *z;   // [Bp][Np] pre-activation (FP32)
  bf16_t *a;  // [Bp][Kp] input of the layer (BF16)
};

struct mlp_net {
  int L, Bmax;
  MlpLayer l[MLP_MAX_LAYERS];
  float *g;      // [Bp][maxNp] FP32 gradient w.r.t. a layer's output
  bf16_t *gh;    // its BF16 copy
  float *gtmp;   // FP32 gradient of the layer below
  bf16_t *ghtmp;
  float *loss;   // device losses, one per batch of the current epoch
  int nloss;
  float *xs, *ts;  // staging of a shuffled batch [Bmax][dims[0]], [Bmax][dims[L]]
  int *perm;       // device permutation of an epoch
  int nperm;
  long step;
};

static inline int pad_to(int v, int q) { return (v + q - 1) / q * q; }

template <int TA, int TB, int EPI>
static void gemm(const bf16_t *a, int lda, const bf16_t *b, int ldb, int M, int N, int K, const GemmOut &o)
{
  hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), dim3(N / 64, M / 64), dim3(256), 0, 0, a, lda, b, ldb, K, o);
}

static uint64_t sm64(uint64_t &s)
{
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
static int dalloc(T **p, size_t n)
{
  HIPCHK(hipMalloc(p, std::max<size_t>(1, n) * sizeof(T)));
  HIPCHK(hipMemset(*p, 0, std::max<size_t>(1, n) * sizeof(T)));
  return FDTD_OK;
}

static void free_net(mlp_net *n)
{
  for (int i = 0; i < n->L; i++) {
    MlpLayer &y = n->l[i];
    for (float *p : {y.w, y.b, y.dw, y.db, y.mw, y.vw, y.mb, y.vb, y.z}) (void)hipFree(p);
    (void)hipFree(y.wh);
    (void)hipFree(y.a);
  }
  (void)hipFree(n->g);
  (void)hipFree(n->gh);
  (void)hipFree(n->gtmp);
  (void)hipFree(n->ghtmp);
  (void)hipFree(n->loss);
  (void)hipFree(n->xs);
  (void)hipFree(n->ts);
  (void)hipFree(n->perm);
  delete n;
}

extern "C" int mlp_create(int nlayers, const int *dims, int batch_max, unsigned long long seed, mlp_net **out)
{
  if (!dims || !out || nlayers < 1 || nlayers > MLP_MAX_LAYERS || batch_max < 1)
    return fdtd_fail(FDTD_EINVAL, "mlp_create: bad argument");
  for (int i = 0; i <= nlayers; i++)
    if (dims[i] < 1 || dims[i] > 4096) return fdtd_fail(FDTD_EINVAL, "mlp_create: width %d out of range", dims[i]);
  mlp_net *n = new mlp_net();
  n->L = nlayers;
  n->Bmax = pad_to(batch_max, 64);
  // every width padded to 64: a layer's padded output is the next layer's padded input
  int maxW = 0;
  uint64_t s = seed;
  int rc = FDTD_OK;
  for (int i = 0; i < nlayers && rc == FDTD_OK; i++) {
    MlpLayer &y = n->l[i];
    y.K = dims[i];
    y.N = dims[i + 1];
    y.Kp = pad_to(y.K, 64);
    y.Np = pad_to(y.N, 64);
    maxW = std::max(maxW, std::max(y.Np, y.Kp));
    const size_t nw = (size_t)y.Np * y.Kp;
    if ((rc = dalloc(&y.w, nw)) || (rc = dalloc(&y.dw, nw)) || (rc = dalloc(&y.mw, nw)) || (rc = dalloc(&y.vw, nw)) ||
        (rc = dalloc(&y.wh, nw)) || (rc = dalloc(&y.b, y.Np)) || (rc = dalloc(&y.db, y.Np)) ||
        (rc = dalloc(&y.mb, y.Np)) || (rc = dalloc(&y.vb, y.Np)) || (rc = dalloc(&y.z, (size_t)n->Bmax * y.Np)) ||
        (rc = dalloc(&y.a, (size_t)n->Bmax * y.Kp)))
      break;
    std::vector<float> hw(nw, 0.f);
    const double bound = 1.0 / std::sqrt((double)y.K);
    for (int r = 0; r < y.N; r++)
      for (int c = 0; c < y.K; c++)
        hw[(size_t)r * y.Kp + c] = (float)((2.0 * (double)(sm64(s) >> 11) / 9007199254740992.0 - 1.0) * bound);
    if (hipMemcpy(y.w, hw.data(), nw * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
      rc = fdtd_fail(FDTD_EHIP, "mlp_create: upload failed");
      break;
    }
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, 0, y.w, y.wh, nw);
  }
  if (rc == FDTD_OK) {
    const size_t ng = (size_t)n->Bmax * maxW;
    n->nloss = 1;
    if (!((rc = dalloc(&n->g, ng)) || (rc = dalloc(&n->gh, ng)) || (rc = dalloc(&n->gtmp, ng)) ||
          (rc = dalloc(&n->ghtmp, ng)) || (rc = dalloc(&n->loss, 1)) ||
          (rc = dalloc(&n->xs, (size_t)n->Bmax * dims[0])) || (rc = dalloc(&n->ts, (size_t)n->Bmax * dims[nlayers]))) &&
        hipDeviceSynchronize() != hipSuccess)
      rc = fdtd_fail(FDTD_EHIP, "mlp_create: init failed");
  }
  if (rc) {
    free_net(n);
    return rc;
  }
  *out = n;
  return FDTD_OK;
}

extern "C" int mlp_destroy(mlp_net *net)
{
  if (!net) return FDTD_EINVAL;
  (void)hipDeviceSynchronize();
  free_net(net);
  return FDTD_OK;
}

// forward over the padded batch Bp; layer i's input is l[i].a, its output the
// next layer's input (BF16 GELU) or l[L-1].z (the linear output)
static int forward(mlp_net *n, const float *x, int B, int Bp)
{
  MlpLayer &l0 = n->l[0];
  const size_t np = (size_t)Bp * l0.Kp;
  hipLaunchKernelGGL(k_pack, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, 0, x, B, l0.K, l0.a, Bp, l0.Kp);
  for (int i = 0; i < n->L; i++) {
    MlpLayer &y = n->l[i];
    GemmOut o = {};
    o.c = y.z;
    o.bias = y.b;
    o.ldc = y.Np;
    if (i + 1 < n->L) {
      o.h = n->l[i + 1].a;
      gemm<0, 0, EPI_BIAS_GELU>(y.a, y.Kp, y.wh, y.Kp, Bp, y.Np, y.Kp, o);
    } else {
      gemm<0, 0, EPI_BIAS>(y.a, y.Kp, y.wh, y.Kp, Bp, y.Np, y.Kp, o);
    }
  }
  HIPCHK(hipGetLastError());
  return FDTD_OK;
}

extern "C" int mlp_forward(mlp_net *n, const float *x, int B, float *y)
{
  if (!n || !x || !y || B < 1 || B > n->Bmax) return fdtd_fail(FDTD_EINVAL, "mlp_forward: bad argument");
  const int Bp = pad_to(B, 64);
  int rc = forward(n, x, B, Bp);
  if (rc) return rc;
  MlpLayer &o = n->l[n->L - 1];
  const size_t ny = (size_t)B * o.N;
  hipLaunchKernelGGL(k_unpack, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, 0, o.z, o.Np, B, o.N, y);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return FDTD_OK;
}

// gradients of the batch into the layers' dW / db, the batch's loss into
// *loss_dev (accumulated: the caller zeroes it); no synchronisation
static int grads_async(mlp_net *n, const float *x, const float *t, int B, float *loss_dev)
{
  const int Bp = pad_to(B, 64);
  int rc = forward(n, x, B, Bp);
  if (rc) return rc;
  MlpLayer &lo = n->l[n->L - 1];
  const size_t ng = (size_t)Bp * lo.Np;
  hipLaunchKernelGGL(k_mse_grad, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, 0, lo.z, lo.Np, t, B, lo.N, Bp, n->g,
                     n->gh, loss_dev);
  float *G = n->g, *Gt = n->gtmp;
  bf16_t *Gh = n->gh, *Ght = n->ghtmp;
  for (int i = n->L - 1; i >= 0; i--) {
    MlpLayer &y = n->l[i];
    // dW[N][K] = G^T A: A(m = out n, k = batch) = G[batch][n] (TA 1), B(k = batch, n = in) = A[batch][in] (TB 1)
    GemmOut o = {};
    o.c = y.dw;
    o.ldc = y.Kp;
    gemm<1, 1, EPI_F32>(Gh, y.Np, y.a, y.Kp, y.Np, y.Kp, Bp, o);
    hipLaunchKernelGGL(k_colsum, dim3((y.Np + 255) / 256), dim3(256), 0, 0, G, Bp, y.Np, y.db);
    if (i > 0) {
      // dA[batch][in] = G W: A = G[batch][out] (TA 0, k = out), B(k = out, n = in) = W[out][in] (TB 1);
      // times GELU'(z) of the layer below gives that layer's output gradient
      MlpLayer &yb = n->l[i - 1];
      GemmOut ob = {};
      ob.c = Gt;
      ob.h = Ght;
      ob.zp = yb.z;
      ob.ldc = y.Kp;
      gemm<0, 1, EPI_GELU_BWD>(Gh, y.Np, y.wh, y.Kp, Bp, y.Kp, y.Np, ob);
      std::swap(G, Gt);
      std::swap(Gh, Ght);
    }
  }
  HIPCHK(hipGetLastError());
  return FDTD_OK;
}

static int adam_async(mlp_net *n, float lr)
{
  n->step++;
  const float c1 = (float)(1.0 - std::pow(0.9, (double)n->step)), c2 = (float)(1.0 - std::pow(0.999, (double)n->step));
  for (int i = 0; i < n->L; i++) {
    MlpLayer &y = n->l[i];
    const size_t nw = (size_t)y.Np * y.Kp;
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, 0, y.w, y.dw, y.mw, y.vw, y.wh, nw, lr,
                       c1, c2);
    hipLaunchKernelGGL(k_adam, dim3((y.Np + 255) / 256), dim3(256), 0, 0, y.b, y.db, y.mb, y.vb, (bf16_t *)nullptr,
                       (size_t)y.Np, lr, c1, c2);
  }
  HIPCHK(hipGetLastError());
  return FDTD_OK;
}

extern "C" int mlp_grads(mlp_net *n, const float *x, const float *t, int B, float *loss)
{
  if (!n || !x || !t || B < 1 || B > n->Bmax) return fdtd_fail(FDTD_EINVAL, "mlp_grads: bad argument");
  HIPCHK(hipMemset(n->loss, 0, sizeof(float)));
  int rc = grads_async(n, x, t, B, n->loss);
  if (rc) return rc;
  if (loss) HIPCHK(hipMemcpy(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipDeviceSynchronize());
  return FDTD_OK;
}

extern "C" int mlp_train_step(mlp_net *n, const float *x, const float *t, int B, float lr, float *loss)
{
  if (!n || !x || !t || B < 1 || B > n->Bmax) return fdtd_fail(FDTD_EINVAL, "mlp_train_step: bad argument");
  HIPCHK(hipMemset(n->loss, 0, sizeof(float)));
  int rc = grads_async(n, x, t, B, n->loss);
  if (rc || (rc = adam_async(n, lr))) return rc;
  if (loss) HIPCHK(hipMemcpy(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipDeviceSynchronize());
  return FDTD_OK;
}

extern "C" int mlp_train_epoch(mlp_net *n, const float *x, const float *t, int rows, int batch, unsigned long long seed,
                               float lr, float *loss)
{
  if (!n || !x || !t || rows < 1 || batch < 1 || batch > n->Bmax) return fdtd_fail(FDTD_EINVAL, "mlp_train_epoch: bad argument");
  // Fisher-Yates shuffle from a SplitMix64 stream of `seed`
  std::vector<int> perm(rows);
  for (int i = 0; i < rows; i++) perm[i] = i;
  uint64_t s = seed;
  for (int i = rows - 1; i > 0; i--) std::swap(perm[i], perm[(int)(sm64(s) % (uint64_t)(i + 1))]);
  if (n->nperm < rows) {
    (void)hipFree(n->perm);
    n->perm = nullptr;
    n->nperm = 0;
    HIPCHK(hipMalloc(&n->perm, (size_t)rows * sizeof(int)));
    n->nperm = rows;
  }
  const int nb = (rows + batch - 1) / batch;
  if (n->nloss < nb) {
    (void)hipFree(n->loss);
    n->loss = nullptr;
    n->nloss = 0;
    HIPCHK(hipMalloc(&n->loss, (size_t)nb * sizeof(float)));
    n->nloss = nb;
  }
  HIPCHK(hipMemcpy(n->perm, perm.data(), (size_t)rows * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(n->loss, 0, (size_t)nb * sizeof(float)));
  const int d0 = n->l[0].K, dL = n->l[n->L - 1].N;
  for (int b = 0; b < nb; b++) {
    const int B = std::min(batch, rows - b * batch);
    const size_t nx = (size_t)B * d0, nt = (size_t)B * dL;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, 0, x, d0, n->perm + b * batch, B,
                       n->xs);
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, 0, t, dL, n->perm + b * batch, B,
                       n->ts);
    int rc = grads_async(n, n->xs, n->ts, B, n->loss + b);
    if (rc || (rc = adam_async(n, lr))) return rc;
  }
  std::vector<float> lb(nb);
  HIPCHK(hipMemcpy(lb.data(), n->loss, (size_t)nb * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipDeviceSynchronize());
  if (loss) {
    double acc = 0.0;
    for (int b = 0; b < nb; b++) acc += (double)lb[b] * std::min(batch, rows - b * batch);
    *loss = (float)(acc / rows);
  }
  return FDTD_OK;
}

static int copy_layer(mlp_net *n, int l, float *w, float *b, const float *dw_src, const float *db_src)
{
  MlpLayer &y = n->l[l];
  std::vector<float> h((size_t)y.Np * y.Kp), hb(y.Np);
  HIPCHK(hipMemcpy(h.data(), dw_src, h.size() * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(hb.data(), db_src, hb.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int r = 0; r < y.N; r++)
    for (int c = 0; c < y.K; c++) w[(size_t)r * y.K + c] = h[(size_t)r * y.Kp + c];
  for (int r = 0; r < y.N; r++) b[r] = hb[r];
  return FDTD_OK;
}

extern "C" int mlp_get_params(mlp_net *n, int l, float *w, float *b)
{
  if (!n || l < 0 || l >= n->L || !w || !b) return fdtd_fail(FDTD_EINVAL, "mlp_get_params: bad argument");
  return copy_layer(n, l, w, b, n->l[l].w, n->l[l].b);
}

extern "C" int mlp_get_grads(mlp_net *n, int l, float *dw, float *db)
{
  if (!n || l < 0 || l >= n->L || !dw || !db) return fdtd_fail(FDTD_EINVAL, "mlp_get_grads: bad argument");
  return copy_layer(n, l, dw, db, n->l[l].dw, n->l[l].db);
}

extern "C" int mlp_set_params(mlp_net *n, int l, const float *w, const float *b)
{
  if (!n || l < 0 || l >= n->L || !w || !b) return fdtd_fail(FDTD_EINVAL, "mlp_set_params: bad argument");
  MlpLayer &y = n->l[l];
  std::vector<float> h((size_t)y.Np * y.Kp, 0.f), hb(y.Np, 0.f);
  for (int r = 0; r < y.N; r++)
    for (int c = 0; c < y.K; c++) h[(size_t)r * y.Kp + c] = w[(size_t)r * y.K + c];
  for (int r = 0; r < y.N; r++) hb[r] = b[r];
  HIPCHK(hipMemcpy(y.w, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(y.b, hb.data(), hb.size() * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((h.size() + 255) / 256)), dim3(256), 0, 0, y.w, y.wh, h.size());
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return FDTD_OK;
}

extern "C" int mlp_gemm_bf16(const unsigned short *a, const unsigned short *b, float *c, int M, int N, int K, int ta, int tb)
{
  if (!a || !b || !c || M <= 0 || N <= 0 || K <= 0 || M % 64 || N % 64 || K % 32 || (ta & ~1) || (tb & ~1))
    return fdtd_fail(FDTD_EINVAL, "mlp_gemm_bf16: sizes must be multiples of 64 (M, N) and 32 (K)");
  GemmOut o = {};
  o.c = c;
  o.ldc = N;
  const int lda = ta ? M : K, ldb = tb ? N : K;
  if (!ta && !tb)
    gemm<0, 0, EPI_F32>(a, lda, b, ldb, M, N, K, o);
  else if (!ta && tb)
    gemm<0, 1, EPI_F32>(a, lda, b, ldb, M, N, K, o);
  else if (ta && !tb)
    gemm<1, 0, EPI_F32>(a, lda, b, ldb, M, N, K, o);
  else
    gemm<1, 1, EPI_F32>(a, lda, b, ldb, M, N, K, o);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return FDTD_OK;
}
