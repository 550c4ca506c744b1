// fdtd_grid.hip -- FP64 reference-order UPML path for gfx950.
//
// One fdtd_grid is one simulation instance of one of the reference's four UPML
// solvers (MPI TM/TE single rank, plain TM/TE).  Per reference update():
//   MPI   : E half (calcJD + calcE + scatteredWave [+ point source]) -> H half
//           (calcMB + calcH) -> per-step NTFF          (mpiTM_UPML.c:196-217, mpiTE_UPML.c:290-304)
//   plain : H half -> E half (+ Gaussian pulse) -> ntff*_TimeCalc
//                                                       (fdtdTM_upml.c:54-66, fdtdTE_upml.c:168-192)
// The arithmetic of every update keeps the reference's operation order and is
// compiled without FMA contraction, so results agree with the CPU solver to
// rounding of the device sin/cos/exp only.
//
// HBM layout: each complex field is an array of double2 (re, im), row-major
// k = i*NPYa + j (j = y contiguous), exactly the reference's layout, so host
// mirrors are plain memcpy targets.  UPML coefficients live in 1-D tables per
// axis; only eps is a 2-D map.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <climits>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>
#include <algorithm>

#include "fdtd_hip.h"
#include "fdtd_model_eps.h"
#include "fdtd_internal.h"

static thread_local std::string g_err;

extern "C" const char *fdtd_hip_last_error(void) { return g_err.c_str(); }

int fdtd_fail(int code, const char *fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

extern "C" int fdtd_hip_device_count(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

namespace {
std::mutex g_amu;
std::unordered_map<void *, size_t> g_live;   // block -> rounded size
std::multimap<size_t, void *> g_kept;        // freed blocks by size
size_t g_kept_bytes = 0;
const size_t KEEP_MAX = (size_t)64 << 30;
}  // namespace

hipError_t dev_malloc_bytes(void **p, size_t bytes)
{
  const size_t b = ((bytes ? bytes : 1) + 255) & ~(size_t)255;
  std::lock_guard<std::mutex> lk(g_amu);
  auto it = g_kept.find(b);
  if (it != g_kept.end()) {
    *p = it->second;
    g_kept.erase(it);
    g_kept_bytes -= b;
    g_live[*p] = b;
    return hipSuccess;
  }
  hipError_t e = hipMalloc(p, b);
  if (e == hipErrorOutOfMemory && !g_kept.empty()) {  // return the kept blocks and try again
    (void)hipGetLastError();
    for (auto &kv : g_kept) (void)hipFree(kv.second);
    g_kept.clear();
    g_kept_bytes = 0;
    e = hipMalloc(p, b);
  }
  if (e == hipSuccess) g_live[*p] = b;
  return e;
}

hipError_t dev_free(void *p)
{
  if (!p) return hipSuccess;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_amu);
  auto it = g_live.find(p);
  if (it == g_live.end()) return hipFree(p);
  const size_t b = it->second;
  g_live.erase(it);
  if (g_kept_bytes + b > KEEP_MAX) return hipFree(p);
  g_kept.emplace(b, p);
  g_kept_bytes += b;
  return hipSuccess;
}

extern "C" int fdtd_hip_trim(void)
{
  hipError_t e = hipDeviceSynchronize();
  std::lock_guard<std::mutex> lk(g_amu);
  for (auto &kv : g_kept) (void)hipFree(kv.second);
  g_kept.clear();
  g_kept_bytes = 0;
  return e == hipSuccess ? FDTD_OK : fdtd_fail(FDTD_EHIP, "hipDeviceSynchronize failed: %s", hipGetErrorString(e));
}

extern "C" int fdtd_hip_alloc(size_t bytes, void **out)
{
  if (!out) return fdtd_fail(FDTD_EINVAL, "null argument");
  *out = nullptr;
  hipError_t e = dev_malloc(out, bytes ? bytes : 1);
  if (e != hipSuccess) return fdtd_fail(FDTD_ENOMEM, "dev_malloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return FDTD_OK;
}

extern "C" int fdtd_hip_free(void *p)
{
  hipError_t e = hipFree(p);
  return e == hipSuccess ? FDTD_OK : fdtd_fail(FDTD_EHIP, "hipFree failed: %s", hipGetErrorString(e));
}

extern "C" int fdtd_hip_upload(void *dst, const void *src, size_t bytes)
{
  if ((!dst || !src) && bytes) return fdtd_fail(FDTD_EINVAL, "null argument");
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? FDTD_OK : fdtd_fail(FDTD_EHIP, "upload failed: %s", hipGetErrorString(e));
}

extern "C" int fdtd_hip_download(void *dst, const void *src, size_t bytes)
{
  if ((!dst || !src) && bytes) return fdtd_fail(FDTD_EINVAL, "null argument");
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? FDTD_OK : fdtd_fail(FDTD_EHIP, "download failed: %s", hipGetErrorString(e));
}

extern "C" int fdtd_hip_synchronize(void)
{
  hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? FDTD_OK : fdtd_fail(FDTD_EHIP, "hipDeviceSynchronize failed: %s", hipGetErrorString(e));
}

// Streaming copy of n 16-byte elements: each thread moves 4 consecutive
// float4 per pass of a grid-stride loop with non-temporal loads and stores
// (the copy is read once and written once, so it should not displace the
// caches), one 1 KiB-per-wavefront-instruction contiguous segment per load.
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_stream_copy(const v4f *__restrict__ src, v4f *__restrict__ dst, size_t n)
{
  const size_t stride = (size_t)gridDim.x * 1024;
  for (size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x; base < n; base += stride) {
    v4f v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const size_t k = base + 256 * u;
      if (k < n) v[u] = __builtin_nontemporal_load(src + k);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const size_t k = base + 256 * u;
      if (k < n) __builtin_nontemporal_store(v[u], dst + k);
    }
  }
}

// The achievable HBM rate: a device-to-device copy of `bytes` (rounded down to
// 16 B) repeated `reps` times, each launch timed with events.  out[0] best and
// out[1] mean rate in GB/s of bytes read plus bytes written; out[2] mean
// launch time in microseconds; out[3] bytes moved per launch.
extern "C" int fdtd_hip_copy_rate(size_t bytes, int reps, double *out)
{
  if (!out || reps < 1 || bytes < 16) return fdtd_fail(FDTD_EINVAL, "fdtd_hip_copy_rate: bad argument");
  const size_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FDTD_OK;
  double best = 0, sum_ms = 0;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int nblk = (int)std::min<size_t>((size_t)cus * 16, (n + 1023) / 1024);
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
    rc = fdtd_fail(FDTD_ENOMEM, "fdtd_hip_copy_rate: hipMalloc of 2 x %zu bytes failed", n * 16);
  } else if (hipMemset(a, 0, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
             hipEventCreate(&e1) != hipSuccess) {
    rc = fdtd_fail(FDTD_EHIP, "fdtd_hip_copy_rate: set-up failed");
  } else {
    hipLaunchKernelGGL(k_stream_copy, dim3(nblk), dim3(256), 0, 0, (const v4f *)a, (v4f *)b, n);  // warm
    for (int r = 0; r < reps && rc == FDTD_OK; r++) {
      float ms = 0;
      if (hipEventRecord(e0, 0) != hipSuccess) rc = fdtd_fail(FDTD_EHIP, "event record failed");
      hipLaunchKernelGGL(k_stream_copy, dim3(nblk), dim3(256), 0, 0, (const v4f *)a, (v4f *)b, n);
      if (rc == FDTD_OK && (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = fdtd_fail(FDTD_EHIP, "copy timing failed: %s", hipGetErrorString(hipGetLastError()));
      sum_ms += ms;
      if (ms > 0) best = std::max(best, 2.0 * n * 16 / (ms * 1e-3) / 1e9);
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (rc == FDTD_OK) {
    const double mean_ms = sum_ms / reps;
    out[0] = best;
    out[1] = mean_ms > 0 ? 2.0 * n * 16 / (mean_ms * 1e-3) / 1e9 : 0.0;
    out[2] = mean_ms * 1e3;
    out[3] = 2.0 * n * 16;
  }
  return rc;
}

// ---------------------------------------------------------------- kernels

struct GridPtrs {
  double2 *F[9];
  const double *eps[3];
  int NPXa, NPYa, off;
  double mu0;
};

struct Tabs {
  const double *x[5];
  const double *sx;
  const double *y[4];
  const double *sy;
};

struct StepArgs {
  double w_t;             // omega * time, as the reference forms it
  double ray_coef;
  double ks_cos, ks_sin;  // MPI scattered wave
  double ps_re, ps_im;    // point source value
  int ps_i, ps_j, ps_on;
  // plain Gaussian pulse (field.c:224-256), component 0 = Ez / Ex, 1 = Ey
  double time, w_s, cpc, spc;
  double t0[2], dot[2], gx[2], gy[2];
  int on[2];
};

#define FOR_CELL()                                                   \
  const int j = 1 + blockIdx.x * 64 + threadIdx.x;                   \
  const int i = 1 + blockIdx.y * 4 + threadIdx.y;                    \
  if (i > g.NPXa - 2 || j > g.NPYa - 2) return;                      \
  const int NPY = g.NPYa;                                            \
  const size_t k = (size_t)i * NPY + j;                              \
  const int x = i - g.off, y = j - g.off;

// Gaussian pulse term dot*g*(1/eps-1)*cexp(i r w) of field_scatteredPulse
__device__ __forceinline__ void pulse_add(double2 &v, const StepArgs &s, int c, int i, int j, double e)
{
  const double r = (i + s.gx[c]) * s.cpc + (j + s.gy[c]) * s.spc - (s.time - s.t0[c]);
  const double q = r / 50.0;
  const double gauss = exp(-(q * q));
  const double a = s.dot[c] * gauss * (1.0 / e - 1);
  double sn, cs;
  sincos(r * s.w_s, &sn, &cs);
  v.x += a * cs;
  v.y += a * sn;
}

// TM E half: calcJD + calcE (+ source).  x tables: 0 C_JZ 1 C_JZHXHY; y: 0 C_DZ 1 C_DZJZ
template <int PLAIN>
__global__ __launch_bounds__(256) void k_tm_e(GridPtrs g, Tabs t, StepArgs s)
{
  FOR_CELL();
  double2 *Ez = g.F[0], *Jz = g.F[1], *Dz = g.F[2];
  const double2 *Hx = g.F[3], *Hy = g.F[6];
  const double2 jz = Jz[k], dz = Dz[k];
  const double2 hy = Hy[k], hyl = Hy[k - NPY], hx = Hx[k], hxb = Hx[k - 1];
  const double cr = ((hy.x - hyl.x) - hx.x) + hxb.x;
  const double ci = ((hy.y - hyl.y) - hx.y) + hxb.y;
  const double a = t.x[0][x], b = t.x[1][x];
  double2 nj;
  nj.x = a * jz.x + b * cr;
  nj.y = a * jz.y + b * ci;
  const double c = t.y[0][y], d = t.y[1][y];
  double2 nd;
  nd.x = (c * dz.x + d * nj.x) - d * jz.x;
  nd.y = (c * dz.y + d * nj.y) - d * jz.y;
  Jz[k] = nj;
  Dz[k] = nd;
  const double e = g.eps[0][k];
  double2 ez;
  ez.x = nd.x / e;
  ez.y = nd.y / e;
  if (!PLAIN) {
    if (e != 1.0) {  // vacuum cells add an exact zero (mpiTM_UPML.c:370)
      const double kr = x * s.ks_cos + y * s.ks_sin;
      const double amp = s.ray_coef * (1.0 / e - 1.0);
      double sn, cs;
      sincos(kr - s.w_t, &sn, &cs);
      ez.x += amp * cs;
      ez.y += amp * sn;
    }
    if (s.ps_on && i == s.ps_i && j == s.ps_j) {
      ez.x += s.ps_re;
      ez.y += s.ps_im;
    }
  } else {
    if (s.on[0] && e != 1.0 && i <= g.NPXa - 2 && j <= g.NPYa - 2) pulse_add(ez, s, 0, i, j, e);
  }
  Ez[k] = ez;
}

// TM H half: calcMB + calcH.  x: 2 C_BXMX1 3 C_BXMX0 4 C_BY, sx; y: 2 C_MX 3 C_MXEZ, sy
__global__ __launch_bounds__(256) void k_tm_h(GridPtrs g, Tabs t)
{
  FOR_CELL();
  const double2 *Ez = g.F[0];
  double2 *Hx = g.F[3], *Mx = g.F[4], *Bx = g.F[5], *Hy = g.F[6], *My = g.F[7], *By = g.F[8];
  const double2 ez = Ez[k], ezt = Ez[k + 1], ezr = Ez[k + NPY];
  const double2 mx = Mx[k], bx = Bx[k], my = My[k], by = By[k];
  const double cmx = t.y[2][y], cmxe = t.y[3][y];
  double2 nmx;
  nmx.x = cmx * mx.x - cmxe * (ezt.x - ez.x);
  nmx.y = cmx * mx.y - cmxe * (ezt.y - ez.y);
  const double b1 = t.x[2][x], b0 = t.x[3][x];
  double2 nbx;  // C_BX == 1
  nbx.x = (bx.x + b1 * nmx.x) - b0 * mx.x;
  nbx.y = (bx.y + b1 * nmx.y) - b0 * mx.y;
  double2 nmy;  // C_MY == C_MYEZ == 1
  nmy.x = my.x - ((-ezr.x) + ez.x);
  nmy.y = my.y - ((-ezr.y) + ez.y);
  const double cby = t.x[4][x];
  const double sxv = t.sx[x], syv = t.sy[y];
  double c1 = 1.0, c0 = 1.0;
  if (sxv != 0.0 || syv != 0.0) {
    c1 = (2.0 + syv) / (2.0 + sxv);
    c0 = (2.0 - syv) / (2.0 + sxv);
  }
  double2 nby;
  nby.x = (cby * by.x + c1 * nmy.x) - c0 * my.x;
  nby.y = (cby * by.y + c1 * nmy.y) - c0 * my.y;
  Mx[k] = nmx;
  Bx[k] = nbx;
  My[k] = nmy;
  By[k] = nby;
  double2 hx, hy;
  hx.x = nbx.x / g.mu0;
  hx.y = nbx.y / g.mu0;
  hy.x = nby.x / g.mu0;
  hy.y = nby.y / g.mu0;
  Hx[k] = hx;
  Hy[k] = hy;
}

// Split-field PML TM (fdtdTM.c): coefficients of field_pmlCoef / field_pmlCoef_LXY
// (field.c:288-295), formed per cell in the reference's operation order
__device__ __forceinline__ double pml_coef(double ep_mu, double sig) { return (1.0 - sig / ep_mu) / (1.0 + sig / ep_mu); }
__device__ __forceinline__ double pml_coef_lxy(double ep_mu, double sig) { return 1.0 / (ep_mu + sig); }

// calcH (fdtdTM.c:313-327).  x: 1 sig_hy_xx; y: 1 sig_hx_yy
__global__ __launch_bounds__(256) void k_split_tm_h(GridPtrs g, Tabs t)
{
  FOR_CELL();
  const double2 *Ezx = g.F[1], *Ezy = g.F[2];
  double2 *Hx = g.F[3], *Hy = g.F[6];
  const double2 zx = Ezx[k], zy = Ezy[k];
  const double2 zxt = Ezx[k + 1], zyt = Ezy[k + 1], zxr = Ezx[k + NPY], zyr = Ezy[k + NPY];
  const double shx = t.y[1][y], shy = t.x[1][x];
  const double chx = pml_coef(g.mu0, shx), chxly = pml_coef_lxy(g.mu0, shx);
  const double chy = pml_coef(g.mu0, shy), chylx = pml_coef_lxy(g.mu0, shy);
  const double2 hx = Hx[k], hy = Hy[k];
  // EZX(i,j+1)-EZX(i,j) + EZY(i,j+1)-EZY(i,j), left to right
  const double dxr = ((zxt.x - zx.x) + zyt.x) - zy.x, dxi = ((zxt.y - zx.y) + zyt.y) - zy.y;
  const double dyr = ((zxr.x - zx.x) + zyr.x) - zy.x, dyi = ((zxr.y - zx.y) + zyr.y) - zy.y;
  Hx[k] = make_double2(chx * hx.x - chxly * dxr, chx * hx.y - chxly * dxi);
  Hy[k] = make_double2(chy * hy.x + chylx * dyr, chy * hy.y + chylx * dyi);
}

// calcE (fdtdTM.c:294-311) and the scattered wave on Ezx after Ez is formed
// (fdtdTM.c:283-288, field.c:179-199 with gap 0).  x: 0 sig_ez_x; y: 0 sig_ez_y
__global__ __launch_bounds__(256) void k_split_tm_e(GridPtrs g, Tabs t, StepArgs s)
{
  FOR_CELL();
  double2 *Ez = g.F[0], *Ezx = g.F[1], *Ezy = g.F[2];
  const double2 *Hx = g.F[3], *Hy = g.F[6];
  const double e = g.eps[0][k];
  const double sx = t.x[0][x], sy = t.y[0][y];
  const double cezx = pml_coef(e, sx), cezxlx = pml_coef_lxy(e, sx);
  const double cezy = pml_coef(e, sy), cezyly = pml_coef_lxy(e, sy);
  const double2 zx = Ezx[k], zy = Ezy[k];
  const double2 hy = Hy[k], hyl = Hy[k - NPY], hx = Hx[k], hxb = Hx[k - 1];
  double2 nzx, nzy;
  nzx.x = cezx * zx.x + cezxlx * (hy.x - hyl.x);
  nzx.y = cezx * zx.y + cezxlx * (hy.y - hyl.y);
  nzy.x = cezy * zy.x - cezyly * (hx.x - hxb.x);
  nzy.y = cezy * zy.y - cezyly * (hx.y - hxb.y);
  Ez[k] = make_double2(nzx.x + nzy.x, nzx.y + nzy.y);
  if (e != 1.0) {  // vacuum cells add an exact zero
    const double kr = (i + 0.0) * s.ks_cos + (j + 0.0) * s.ks_sin;
    const double amp = s.ray_coef * (1.0 / e - 1.0);
    double s1, c1, s0, c0;
    sincos(kr - s.w_s * (s.time + 0.5), &s1, &c1);
    sincos(kr - s.w_s * (s.time - 0.5), &s0, &c0);
    nzx.x += amp * (c1 - c0);
    nzx.y += amp * (s1 - s0);
  }
  Ezx[k] = nzx;
  Ezy[k] = nzy;
}

// Split-field PML TE (fdtdTE.c): calcE (fdtdTE.c:290-303), then the scattered wave
// on Ey (fdtdTE.c:283-288, field.c:179-199 with gap (0, 0.5)).  x: 0 sig_ey_x; y: 0 sig_ex_y
__global__ __launch_bounds__(256) void k_split_te_e(GridPtrs g, Tabs t, StepArgs s)
{
  FOR_CELL();
  double2 *Ex = g.F[0], *Ey = g.F[3];
  const double2 *Hzx = g.F[7], *Hzy = g.F[8];
  const double eex = g.eps[0][k], eey = g.eps[1][k];
  const double sy = t.y[0][y], sx = t.x[0][x];
  const double cex = pml_coef(eex, sy), cexly = pml_coef_lxy(eex, sy);
  const double cey = pml_coef(eey, sx), ceylx = pml_coef_lxy(eey, sx);
  const double2 zx = Hzx[k], zy = Hzy[k];
  const double2 zxb = Hzx[k - 1], zyb = Hzy[k - 1], zxl = Hzx[k - NPY], zyl = Hzy[k - NPY];
  const double2 ex = Ex[k], ey = Ey[k];
  // HZX(i,j) - HZX(i,j-1) + HZY(i,j) - HZY(i,j-1), left to right (and i-1 for Ey)
  const double ar = ((zx.x - zxb.x) + zy.x) - zyb.x, ai = ((zx.y - zxb.y) + zy.y) - zyb.y;
  const double br = ((zx.x - zxl.x) + zy.x) - zyl.x, bi = ((zx.y - zxl.y) + zy.y) - zyl.y;
  double2 nex, ney;
  nex.x = cex * ex.x + cexly * ar;
  nex.y = cex * ex.y + cexly * ai;
  ney.x = cey * ey.x - ceylx * br;
  ney.y = cey * ey.y - ceylx * bi;
  if (eey != 1.0) {  // vacuum cells add an exact zero
    const double kr = (i + 0.0) * s.ks_cos + (j + 0.5) * s.ks_sin;
    const double amp = s.ray_coef * (1.0 / eey - 1.0);
    double s1, c1, s0, c0;
    sincos(kr - s.w_s * (s.time + 0.5), &s1, &c1);
    sincos(kr - s.w_s * (s.time - 0.5), &s0, &c0);
    ney.x += amp * (c1 - c0);
    ney.y += amp * (s1 - s0);
  }
  Ex[k] = nex;
  Ey[k] = ney;
}

// calcH (fdtdTE.c:305-321): Hzx, Hzy, then Hz = Hzx + Hzy.  x: 1 sig_hz_xx; y: 1 sig_hz_yy
__global__ __launch_bounds__(256) void k_split_te_h(GridPtrs g, Tabs t)
{
  FOR_CELL();
  const double2 *Ex = g.F[0], *Ey = g.F[3];
  double2 *Hz = g.F[6], *Hzx = g.F[7], *Hzy = g.F[8];
  const double shx = t.x[1][x], shy = t.y[1][y];
  const double chzx = pml_coef(g.mu0, shx), chzxlx = pml_coef_lxy(g.mu0, shx);
  const double chzy = pml_coef(g.mu0, shy), chzyly = pml_coef_lxy(g.mu0, shy);
  const double2 ey = Ey[k], eyr = Ey[k + NPY], ex = Ex[k], ext = Ex[k + 1];
  const double2 zx = Hzx[k], zy = Hzy[k];
  double2 nzx, nzy;
  nzx.x = chzx * zx.x - chzxlx * (eyr.x - ey.x);
  nzx.y = chzx * zx.y - chzxlx * (eyr.y - ey.y);
  nzy.x = chzy * zy.x + chzyly * (ext.x - ex.x);
  nzy.y = chzy * zy.y + chzyly * (ext.y - ex.y);
  Hzx[k] = nzx;
  Hzy[k] = nzy;
  Hz[k] = make_double2(nzx.x + nzy.x, nzx.y + nzy.y);
}

// Non-standard FDTD TM (nsFdtdTM.c).  Per-cell coefficient maps nsc[q * ncell + k]:
// 0 C_EZXLX, 1 C_EZYLY, 2 C_HXLY, 3 C_HYLX, 4 the NS scattered-wave factor.
// calcH (nsFdtdTM.c:111-151): the Yee difference plus r/2 times the difference of
// the second differences across the other axis.  x: 1 C_HY; y: 1 C_HX
__global__ __launch_bounds__(256) void k_ns_tm_h(GridPtrs g, Tabs t, const double *__restrict__ nsc, double r2)
{
  FOR_CELL();
  const size_t n = (size_t)g.NPXa * NPY;
  const double2 *Ez = g.F[0];
  double2 *Hx = g.F[3], *Hy = g.F[6];
  // c (i,j), u (i,j+1), d (i,j-1), r (i+1,j), l (i-1,j), ur (i+1,j+1), ul (i-1,j+1), dr (i+1,j-1)
  const double2 c = Ez[k], u = Ez[k + 1], d = Ez[k - 1], r = Ez[k + NPY], l = Ez[k - NPY];
  const double2 ur = Ez[k + NPY + 1], ul = Ez[k - NPY + 1], dr = Ez[k + NPY - 1];
  const double chx = t.y[1][y], chy = t.x[1][x];
  const double chxly = nsc[2 * n + k], chylx = nsc[3 * n + k];
  const double2 hx = Hx[k], hy = Hy[k];
  // (Ez[k+dy+dx] + Ez[k+dy-dx] - 2 Ez[k+dy]) - (Ez[k+dx] + Ez[k-dx] - 2 Ez[k])
  const double nxr = r2 * (((ur.x + ul.x) - 2 * u.x) - ((r.x + l.x) - 2 * c.x));
  const double nxi = r2 * (((ur.y + ul.y) - 2 * u.y) - ((r.y + l.y) - 2 * c.y));
  // (Ez[k+dx+dy] + Ez[k+dx-dy] - 2 Ez[k+dx]) - (Ez[k+dy] + Ez[k-dy] - 2 Ez[k])
  const double nyr = r2 * (((ur.x + dr.x) - 2 * r.x) - ((u.x + d.x) - 2 * c.x));
  const double nyi = r2 * (((ur.y + dr.y) - 2 * r.y) - ((u.y + d.y) - 2 * c.y));
  Hx[k] = make_double2(chx * hx.x - chxly * ((u.x - c.x) + nxr), chx * hx.y - chxly * ((u.y - c.y) + nxi));
  Hy[k] = make_double2(chy * hy.x + chylx * ((r.x - c.x) + nyr), chy * hy.y + chylx * ((r.y - c.y) + nyi));
}

// calcE (nsFdtdTM.c:85-109), the NS scattered wave on Ezy (field.c:155-177 with gap
// 0 and dot 1: phases w (t+1) and w t), then Ez = Ezx + Ezy (nsFdtdTM.c:68-80).
// x: 0 C_EZX; y: 0 C_EZY
__global__ __launch_bounds__(256) void k_ns_tm_e(GridPtrs g, Tabs t, const double *__restrict__ nsc, StepArgs s)
{
  FOR_CELL();
  const size_t n = (size_t)g.NPXa * NPY;
  double2 *Ez = g.F[0], *Ezx = g.F[1], *Ezy = g.F[2];
  const double2 *Hx = g.F[3], *Hy = g.F[6];
  const double cezx = t.x[0][x], cezy = t.y[0][y];
  const double cezxlx = nsc[k], cezyly = nsc[n + k], fac = nsc[4 * n + k];
  const double2 zx = Ezx[k], zy = Ezy[k];
  const double2 hy = Hy[k], hyl = Hy[k - NPY], hx = Hx[k], hxb = Hx[k - 1];
  double2 nzx, nzy;
  nzx.x = cezx * zx.x + cezxlx * (hy.x - hyl.x);
  nzx.y = cezx * zx.y + cezxlx * (hy.y - hyl.y);
  nzy.x = cezy * zy.x - cezyly * (hx.x - hxb.x);
  nzy.y = cezy * zy.y - cezyly * (hx.y - hxb.y);
  if (fac != 0.0) {  // vacuum cells (n = 1) have a factor of exactly zero
    const double kr = (i + 0.0) * s.ks_cos + (j + 0.0) * s.ks_sin;
    const double amp = s.ray_coef * fac;
    double s1, c1, s0, c0;
    sincos(kr - s.w_s * (s.time + 1.0), &s1, &c1);
    sincos(kr - s.w_s * s.time, &s0, &c0);
    nzy.x += amp * (c1 - c0);
    nzy.y += amp * (s1 - s0);
  }
  Ezx[k] = nzx;
  Ezy[k] = nzy;
  Ez[k] = make_double2(nzx.x + nzy.x, nzx.y + nzy.y);
}

// Non-standard FDTD TE (nsFdtdTE.c).  Per-cell coefficient maps nsc[q * ncell + k]:
// 0 C_EX, 1 C_EXLY, 2 C_EY, 3 C_EYLX, 4 C_HZX, 5 C_HZXLX, 6 C_HZY, 7 C_HZYLY,
// 8 / 9 the NS scattered-wave factors of EPS_EX / EPS_EY.
// calcH (nsFdtdTE.c:291-308), then Hz = Hzx + Hzy (nsFdtdTE.c:243-248)
__global__ __launch_bounds__(256) void k_ns_te_h(GridPtrs g, const double *__restrict__ nsc)
{
  FOR_CELL();
  const size_t n = (size_t)g.NPXa * NPY;
  const double2 *Ex = g.F[0], *Ey = g.F[3];
  double2 *Hz = g.F[6], *Hzx = g.F[7], *Hzy = g.F[8];
  const double chzx = nsc[4 * n + k], chzxlx = nsc[5 * n + k], chzy = nsc[6 * n + k], chzyly = nsc[7 * n + k];
  const double2 ey = Ey[k], eyr = Ey[k + NPY], ex = Ex[k], ext = Ex[k + 1];
  const double2 zx = Hzx[k], zy = Hzy[k];
  double2 nzx, nzy;
  nzx.x = chzx * zx.x - chzxlx * (eyr.x - ey.x);
  nzx.y = chzx * zx.y - chzxlx * (eyr.y - ey.y);
  nzy.x = chzy * zy.x + chzyly * (ext.x - ex.x);
  nzy.y = chzy * zy.y + chzyly * (ext.y - ex.y);
  Hzx[k] = nzx;
  Hzy[k] = nzy;
  Hz[k] = make_double2(nzx.x + nzy.x, nzx.y + nzy.y);
}

// the NS scattered-wave term (field.c:155-177) (ray_coef dot) fac (e^{i(kr - w(t+1))} - e^{i(kr - w t)})
__device__ __forceinline__ void ns_wave_add(double2 &v, const StepArgs &s, double dot, double fac, double kr)
{
  const double amp = (s.ray_coef * dot) * fac;
  double s1, c1, s0, c0;
  sincos(kr - s.w_s * (s.time + 1.0), &s1, &c1);
  sincos(kr - s.w_s * s.time, &s0, &c0);
  v.x += amp * (c1 - c0);
  v.y += amp * (s1 - s0);
}

// calcE (nsFdtdTE.c:251-289) with the non-standard second differences of Hz, then
// the NS scattered wave on Ex (dot cos(angle + 90)) and Ey (dot sin(angle + 90)),
// both at gap (0, 1/2) (nsFdtdTE.c:250-256)
__global__ __launch_bounds__(256) void k_ns_te_e(GridPtrs g, const double *__restrict__ nsc, double r2, StepArgs s)
{
  FOR_CELL();
  const size_t n = (size_t)g.NPXa * NPY;
  double2 *Ex = g.F[0], *Ey = g.F[3];
  const double2 *Hz = g.F[6];
  // c (i,j), r (i+1,j), l (i-1,j), u (i,j+1), d (i,j-1), dr (i+1,j-1), dl (i-1,j-1), ul (i-1,j+1)
  const double2 c = Hz[k], r = Hz[k + NPY], l = Hz[k - NPY], u = Hz[k + 1], d = Hz[k - 1];
  const double2 dr = Hz[k + NPY - 1], dl = Hz[k - NPY - 1], ul = Hz[k - NPY + 1];
  const double cex = nsc[k], cexly = nsc[n + k], cey = nsc[2 * n + k], ceylx = nsc[3 * n + k];
  const double2 ex = Ex[k], ey = Ey[k];
  // (Hz[k+dx] + Hz[k-dx] - 2 Hz[k]) - (Hz[k-dy+dx] + Hz[k-dy-dx] - 2 Hz[k-dy])
  const double nxr = r2 * (((r.x + l.x) - 2 * c.x) - ((dr.x + dl.x) - 2 * d.x));
  const double nxi = r2 * (((r.y + l.y) - 2 * c.y) - ((dr.y + dl.y) - 2 * d.y));
  // (Hz[k+dy] + Hz[k-dy] - 2 Hz[k]) - (Hz[k-dx+dy] + Hz[k-dx-dy] - 2 Hz[k-dx])
  const double nyr = r2 * (((u.x + d.x) - 2 * c.x) - ((ul.x + dl.x) - 2 * l.x));
  const double nyi = r2 * (((u.y + d.y) - 2 * c.y) - ((ul.y + dl.y) - 2 * l.y));
  double2 nex = make_double2(cex * ex.x + cexly * ((c.x - d.x) + nxr), cex * ex.y + cexly * ((c.y - d.y) + nxi));
  double2 ney = make_double2(cey * ey.x - ceylx * ((c.x - l.x) + nyr), cey * ey.y - ceylx * ((c.y - l.y) + nyi));
  const double kr = (i + 0.0) * s.ks_cos + (j + 0.5) * s.ks_sin;
  const double fx = nsc[8 * n + k], fy = nsc[9 * n + k];
  if (s.on[0] && fx != 0.0) ns_wave_add(nex, s, s.dot[0], fx, kr);  // vacuum cells: factor exactly zero
  if (s.on[1] && fy != 0.0) ns_wave_add(ney, s, s.dot[1], fy, kr);
  Ex[k] = nex;
  Ey[k] = ney;
}

// TE E half.  x: 0 C_DXJX1 1 C_DXJX0 2 C_DY, sx; y: 0 C_JX 1 C_JXHZ, sy
template <int PLAIN>
__global__ __launch_bounds__(256) void k_te_e(GridPtrs g, Tabs t, StepArgs s)
{
  FOR_CELL();
  double2 *Ex = g.F[0], *Jx = g.F[1], *Dx = g.F[2], *Ey = g.F[3], *Jy = g.F[4], *Dy = g.F[5];
  const double2 *Hz = g.F[6];
  const double2 hz = Hz[k], hzb = Hz[k - 1], hzl = Hz[k - NPY];
  const double2 jx = Jx[k], dx = Dx[k], jy = Jy[k], dy = Dy[k];
  const double cjx = t.y[0][y], cjxh = t.y[1][y];
  double2 njx;
  njx.x = cjx * jx.x + cjxh * (hz.x - hzb.x);
  njx.y = cjx * jx.y + cjxh * (hz.y - hzb.y);
  const double d1 = t.x[0][x], d0 = t.x[1][x];
  double2 ndx;  // C_DX == 1
  ndx.x = (dx.x + d1 * njx.x) - d0 * jx.x;
  ndx.y = (dx.y + d1 * njx.y) - d0 * jx.y;
  double2 njy;  // C_JY == C_JYHZ == 1
  njy.x = jy.x + ((-hz.x) + hzl.x);
  njy.y = jy.y + ((-hz.y) + hzl.y);
  const double cdy = t.x[2][x];
  const double sxv = t.sx[x], syv = t.sy[y];
  double c1 = 1.0, c0 = 1.0;
  if (sxv != 0.0 || syv != 0.0) {
    c1 = (2.0 + syv) / (2.0 + sxv);
    c0 = (2.0 - syv) / (2.0 + sxv);
  }
  double2 ndy;
  ndy.x = (cdy * dy.x + c1 * njy.x) - c0 * jy.x;
  ndy.y = (cdy * dy.y + c1 * njy.y) - c0 * jy.y;
  Jx[k] = njx;
  Dx[k] = ndx;
  Jy[k] = njy;
  Dy[k] = ndy;
  const double ex_e = g.eps[0][k], ey_e = g.eps[1][k];
  double2 ex, ey;
  ex.x = ndx.x / ex_e;
  ex.y = ndx.y / ex_e;
  ey.x = ndy.x / ey_e;
  ey.y = ndy.y / ey_e;
  if (!PLAIN) {
    if (ey_e != 1.0) {  // mpiTE_UPML.c:278, on Ey only
      const double ikx = x * s.ks_cos + y * s.ks_sin;
      const double amp = s.ray_coef * (1.0 / ey_e - 1);
      double sn, cs;
      sincos(ikx - s.w_t, &sn, &cs);
      ey.x += amp * cs;
      ey.y += amp * sn;
    }
  } else {
    if (s.on[0] && ex_e != 1.0) pulse_add(ex, s, 0, i, j, ex_e);
    if (s.on[1] && ey_e != 1.0) pulse_add(ey, s, 1, i, j, ey_e);
  }
  Ex[k] = ex;
  Ey[k] = ey;
}

// TE H half.  x: 3 C_MZ 4 C_MZEXEY; y: 2 C_BZ 3 C_BZMZ
__global__ __launch_bounds__(256) void k_te_h(GridPtrs g, Tabs t)
{
  FOR_CELL();
  const double2 *Ex = g.F[0], *Ey = g.F[3];
  double2 *Hz = g.F[6], *Mz = g.F[7], *Bz = g.F[8];
  const double2 ey = Ey[k], eyr = Ey[k + NPY], ex = Ex[k], ext = Ex[k + 1];
  const double2 mz = Mz[k], bz = Bz[k];
  const double cmz = t.x[3][x], cmze = t.x[4][x];
  double2 nmz;
  nmz.x = cmz * mz.x - cmze * (((eyr.x - ey.x) - ext.x) + ex.x);
  nmz.y = cmz * mz.y - cmze * (((eyr.y - ey.y) - ext.y) + ex.y);
  const double cbz = t.y[2][y], cbm = t.y[3][y];
  double2 nbz;
  nbz.x = (cbz * bz.x + cbm * nmz.x) - cbm * mz.x;
  nbz.y = (cbz * bz.y + cbm * nmz.y) - cbm * mz.y;
  Mz[k] = nmz;
  Bz[k] = nbz;
  double2 hz;
  hz.x = nbz.x / g.mu0;
  hz.y = nbz.y / g.mu0;
  Hz[k] = hz;
}

// ---------------------------------------------------------------- NTFF (time domain)

// gather the contour samples of this step: E value (signed) and the averaged,
// signed H value of every contour cell
__global__ __launch_bounds__(256) void k_ntff_gather(const double2 *const *F, Contour ct, double2 *cE, double2 *cH)
{
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= ct.total) return;
  int s = 0;
  while (s < 3 && p >= ct.s[s + 1].off) s++;
  const Seg &sg = ct.s[s];
  const int c = p - sg.off;
  const long k = sg.k0 + (long)c * sg.dk;
  const double2 e = F[sg.ef][k];
  const double2 h0 = F[sg.hf][k], h1 = F[sg.hf][k + sg.hnb];
  double2 ev, hv;
  ev.x = sg.es * e.x;
  ev.y = sg.es * e.y;
  hv.x = sg.hs * (h0.x + h1.x);
  hv.y = sg.hs * (h0.y + h1.y);
  cE[p] = ev;
  cH[p] = hv;
}

#define NT_GUARD_DEV 4  // guard bins allocated before and after the accumulators
#define NTFF_WIN 1024  // bin offsets per LDS window of k_ntff_bin
#define NTFF_GRP 32    // sample groups per workgroup of k_ntff_bin
#define NTFF_RUN 16    // samples per group and part
#define NTFF_AHEAD 4   // records of a group's run loaded together

// Binning of n consecutive steps per launch.
//
// For one angle and contour cell the time shift ts is fixed, so the E sample of
// the step with base time tb = time - 1 lands on bins tb + dE + (-1, 0, +1) and
// the H sample on tb + dH + (-1, 0, +1), with dE = floor(ts + 0.5), dH =
// floor(ts) + 1 and 3-tap weights (b, a - b, -a), b = 1 - a, that do not depend
// on the step (a = 0.5 + ts - dE for E, ts - (dH - 1) for H): the bins and
// weights of the reference's calc() (ntffTM.c:279-288, mpiTM_UPML.c:994-1036).
//
// The host sorts the samples of each (angle, target array) by offset d and cuts
// the sorted list into parts of at most NTFF_GRP x NTFF_RUN samples spanning at
// most NTFF_WIN offsets, one workgroup each.  A workgroup is NTFF_GRP groups of
// NS >= n lanes; group g takes a contiguous run of the part's sorted samples
// (records stored [iteration][group], so that their loads coalesce) and lane s
// of the group handles step s: the n steps of a sample are contiguous in the
// step-minor sample buffer ([cell][nbuf], E then H), so one load instruction
// reads whole cache lines; a record holds the element of its sample's first
// step, so a value's address is one add.  Lane s keeps the three bins d + s - 1 + k of its step in
// registers; as the run's offsets grow the window shifts, and only the bins it
// leaves go to the LDS window with an atomic.  The LDS window goes to the
// accumulators once per launch: its first and last n + 2 bins with atomics (the
// neighbouring parts of the same row reach them), the others with a plain
// read-modify-write.
struct BinRec {
  int p;     // element of the sample's first step in the sample buffer, or -1 (padding at the end of a group's run)
  int r;     // d minus the part's smallest offset
  double a;  // tap fraction
};
struct BinPart {
  int rec0, per;  // first record, iterations (records per group)
  int dlo, nd;    // smallest offset, offsets spanned
  int row, arr;   // angle, target array
};

template <int NS>
__global__ __launch_bounds__(NTFF_GRP * NS) void k_ntff_bin(const double2 *cS, int n,
                                                             const BinPart *parts, const BinRec *rec, double2 *U0,
                                                             double2 *U1, double2 *U2, int arraySize, long tb0,
                                                             double coef, int useCoef)
{
  constexpr int NT = NTFF_GRP * NS;
  __shared__ double lds[NTFF_WIN + NS + 1][2];
  const BinPart W = parts[blockIdx.x];
  double2 *U = W.arr == 0 ? U0 : (W.arr == 1 ? U1 : U2);
  const long stp = (long)W.row * arraySize;
  const double cf = useCoef ? coef : 1.0;
  const int nb = W.nd + n + 1;
  for (int q = threadIdx.x; q < nb; q += NT) lds[q][0] = lds[q][1] = 0.0;
  __syncthreads();
  const int grp = threadIdx.x / NS, s = threadIdx.x % NS;
  const bool act = s < n;
  double xr0 = 0.0, xr1 = 0.0, xr2 = 0.0, xi0 = 0.0, xi1 = 0.0, xi2 = 0.0;
  int rc = -1;  // offset (window-relative) of the run's current sample; xr0 is bin rc + s
  const BinRec *rr = rec + W.rec0 + grp;
  auto value = [&](int p) {
    // p: the element of the sample's first step; padding (p < 0) and lanes past
    // n read nothing
    return (p >= 0 && act) ? cS[p + s] : make_double2(0.0, 0.0);
  };
  // the run is taken NTFF_AHEAD records at a time: all their record loads, then
  // all their value loads, are in flight together (a record's value load
  // depends on the record); padding records (p < 0) sit only at the run's end
  for (int i0 = 0; i0 < W.per; i0 += NTFF_AHEAD) {
    BinRec R[NTFF_AHEAD];
    double2 v[NTFF_AHEAD];
#pragma unroll
    for (int k = 0; k < NTFF_AHEAD; k++) {
      R[k] = rr[(size_t)min(i0 + k, W.per - 1) * NTFF_GRP];
      if (i0 + k >= W.per) R[k].p = -1;
    }
#pragma unroll
    for (int k = 0; k < NTFF_AHEAD; k++) v[k] = value(R[k].p);
#pragma unroll
    for (int k = 0; k < NTFF_AHEAD; k++) {
      if (R[k].p < 0) break;
      if (R[k].r != rc) {
        if (rc >= 0) {
          const int delta = R[k].r - rc;
          if (act) {
            if (delta >= 3) {
              atomicAdd(&lds[rc + s][0], xr0);
              atomicAdd(&lds[rc + s][1], xi0);
              atomicAdd(&lds[rc + s + 1][0], xr1);
              atomicAdd(&lds[rc + s + 1][1], xi1);
              atomicAdd(&lds[rc + s + 2][0], xr2);
              atomicAdd(&lds[rc + s + 2][1], xi2);
              xr0 = xr1 = xr2 = xi0 = xi1 = xi2 = 0.0;
            } else if (delta == 2) {
              atomicAdd(&lds[rc + s][0], xr0);
              atomicAdd(&lds[rc + s][1], xi0);
              atomicAdd(&lds[rc + s + 1][0], xr1);
              atomicAdd(&lds[rc + s + 1][1], xi1);
              xr0 = xr2;
              xi0 = xi2;
              xr1 = xr2 = xi1 = xi2 = 0.0;
            } else {
              atomicAdd(&lds[rc + s][0], xr0);
              atomicAdd(&lds[rc + s][1], xi0);
              xr0 = xr1;
              xi0 = xi1;
              xr1 = xr2;
              xi1 = xi2;
              xr2 = xi2 = 0.0;
            }
          }
        }
        rc = R[k].r;
      }
      const double a = R[k].a, b = 1.0 - a;
      const double vr = v[k].x, vi = v[k].y;  // the coefficient is applied to the window's bins
      xr0 += vr * b;
      xi0 += vi * b;
      xr1 += vr * (a - b);
      xi1 += vi * (a - b);
      xr2 -= vr * a;
      xi2 -= vi * a;
    }
    if (R[NTFF_AHEAD - 1].p < 0) break;
  }
  if (rc >= 0 && act) {
    atomicAdd(&lds[rc + s][0], xr0);
    atomicAdd(&lds[rc + s][1], xi0);
    atomicAdd(&lds[rc + s + 1][0], xr1);
    atomicAdd(&lds[rc + s + 1][1], xi1);
    atomicAdd(&lds[rc + s + 2][0], xr2);
    atomicAdd(&lds[rc + s + 2][1], xi2);
  }
  __syncthreads();
  // window bins [tb0 + dlo - 1, + nb) into the accumulators
  for (int q = threadIdx.x; q < nb; q += NT) {
    const double re = lds[q][0] * cf, im = lds[q][1] * cf;
    if (re == 0.0 && im == 0.0) continue;
    const long m = tb0 + W.dlo - 1 + q;
    // taps outside the [360][arraySize] allocation (+ guard) are dropped; the
    // reference writes them out of bounds on strongly non-square grids
    const long gidx = stp + m;
    if (gidx < -NT_GUARD_DEV || gidx >= 360L * arraySize + NT_GUARD_DEV) continue;
    double2 *dst = U + gidx;
    // bins near a row end are shared with the neighbouring angle (the flat
    // [360][arraySize] layout lets m-1 / m+1 spill over, Q10), the window's
    // edge bins with the neighbouring parts: atomics there
    if (m < 3 || m >= arraySize - 3 || q < n + 2 || q >= nb - (n + 2)) {
      atomicAdd(&dst->x, re);
      atomicAdd(&dst->y, im);
    } else {
      double2 u = *dst;
      u.x += re;
      u.y += im;
      *dst = u;
    }
  }
}

// ---------------------------------------------------------------- running DFT of the contour

// V_f[c] += sum_s val_s[c] e^{i w_f tb_s}, tb_s = tb0 + s the binning's base time
// of step s, for n steps of samples val_s[c] at src[s * ss + c * sp] (the
// staging buffers [step][cell], or the binning's [cell][n]).  A workgroup covers
// 64 cells, one per lane, and DFT_WAVES * NFL frequencies: wavefront w keeps the
// sums of its lane's cell at frequencies w, w + DFT_WAVES, ... in registers over
// all n steps and adds them to the FP64 accumulators ([f][cell]) once.  A
// frequency is the same for the whole wavefront, so its phase factor
// (ph[s][f] = e^{i w_f tb_s}, from k_dft_phase) is a scalar load, and each
// complex product-sum is four fused multiply-adds per accumulator pair.
#define DFT_WAVES 4
#define DFT_NFL_MAX 8
#define DFT_STAGE_STEPS 128  // steps staged before a DFT launch
template <int NFL>
__global__ __launch_bounds__(64 * DFT_WAVES) void k_ctr_dft(const double2 *cE, const double2 *cH, long ss, long sp,
                                                           int n, int P, const double2 *ph, int nf, double2 *VE,
                                                           double2 *VH)
{
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int p = blockIdx.x * 64 + lane;
  const int fb = blockIdx.y * (DFT_WAVES * NFL) + wv;  // this wavefront's frequencies: fb + DFT_WAVES k
  const size_t pc = (size_t)min(p, P - 1) * (size_t)sp;  // clamped cell: every load stays in bounds
  double ar[NFL], ai[NFL], br[NFL], bi[NFL];
#pragma unroll
  for (int k = 0; k < NFL; k++) ar[k] = ai[k] = br[k] = bi[k] = 0.0;
  for (int s = 0; s < n; s++) {
    const double2 e = cE[(size_t)s * ss + pc], h = cH[(size_t)s * ss + pc];
    const double2 *row = ph + (size_t)s * nf;
#pragma unroll
    for (int k = 0; k < NFL; k++) {
      const double2 x = row[min(fb + DFT_WAVES * k, nf - 1)];  // clamped: sums past nf are never stored
      ar[k] = fma(e.x, x.x, ar[k]);
      ar[k] = fma(-e.y, x.y, ar[k]);
      ai[k] = fma(e.x, x.y, ai[k]);
      ai[k] = fma(e.y, x.x, ai[k]);
      br[k] = fma(h.x, x.x, br[k]);
      br[k] = fma(-h.y, x.y, br[k]);
      bi[k] = fma(h.x, x.y, bi[k]);
      bi[k] = fma(h.y, x.x, bi[k]);
    }
  }
  if (p >= P) return;
#pragma unroll
  for (int k = 0; k < NFL; k++) {
    const int f = fb + DFT_WAVES * k;
    if (f >= nf) continue;
    const size_t q = (size_t)f * P + p;
    double2 a = VE[q], b = VH[q];
    a.x += ar[k];
    a.y += ai[k];
    b.x += br[k];
    b.y += bi[k];
    VE[q] = a;
    VH[q] = b;
  }
}

// ph[s][f] = e^{i w_f (tb0 + s)} (FP64 sincos), s < n
__global__ __launch_bounds__(256) void k_dft_phase(const double *omega, int nf, long tb0, int n, double2 *ph)
{
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n * nf) return;
  const int s = q / nf, f = q % nf;
  double sn, cs;
  sincos(omega[f] * (double)(tb0 + s), &sn, &cs);
  ph[q] = make_double2(cs, sn);
}

// The DFT at w_f of the three time-domain accumulator rows of every angle,
// sum_m U[m] e^{i w m}, formed from the running DFT of the contour: a sample
// binned at tb + d + (-1, 0, 1) with weights (b, a - b, -a) contributes
// V e^{i w (d - 1)} (b + (a - b) e^{i w} - a e^{2 i w}), with d and a from the
// same time shifts as the binning.  One workgroup per (angle, frequency);
// out[f][array][angle].
__global__ __launch_bounds__(256) void k_dft_far(const double2 *VE, const double2 *VH, Contour ct, const double *ts,
                                                 const double *omega, double cf, double2 *out)
{
  __shared__ double red[6][4];
  const int ang = blockIdx.x, f = blockIdx.y, P = ct.total;
  const double w = omega[f];
  double sn1, cs1, sn2, cs2;
  sincos(w, &sn1, &cs1);
  sincos(2.0 * w, &sn2, &cs2);
  double acc[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  for (int p = threadIdx.x; p < P; p += 256) {
    int q = 0;
    while (q < 3 && p >= ct.s[q + 1].off) q++;
    const double t = ts[(size_t)ang * P + p];
#pragma unroll
    for (int hh = 0; hh < 2; hh++) {
      const int d = hh ? (int)floor(t) + 1 : (int)floor(t + 0.5);
      const double a = hh ? t - (double)(d - 1) : (0.5 + t) - (double)d, b = 1.0 - a;
      // (b + (a - b) e^{iw} - a e^{2iw}) e^{i w (d - 1)}
      const double kr = b + (a - b) * cs1 - a * cs2, ki = (a - b) * sn1 - a * sn2;
      double sd, cd;
      sincos(w * (double)(d - 1), &sd, &cd);
      const double Kr = kr * cd - ki * sd, Ki = kr * sd + ki * cd;
      const double2 v = hh ? VH[(size_t)f * P + p] : VE[(size_t)f * P + p];
      const int arr = hh ? ct.s[q].wa : ct.s[q].ua;
      const double re = v.x * Kr - v.y * Ki, im = v.x * Ki + v.y * Kr;
#pragma unroll
      for (int aa = 0; aa < 3; aa++)
        if (aa == arr) {
          acc[aa][0] += re;
          acc[aa][1] += im;
        }
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    double v = acc[k >> 1][k & 1];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[k][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int aa = threadIdx.x;
    double re = 0.0, im = 0.0;
    for (int k = 0; k < 4; k++) {
      re += red[2 * aa][k];
      im += red[2 * aa + 1][k];
    }
    out[((size_t)f * 3 + aa) * 360 + ang] = make_double2(cf * re, cf * im);
  }
}

// ---------------------------------------------------------------- far field

__device__ __forceinline__ void cmul_acc(double2 &acc, const double2 a, double cs, double sn, double sgn)
{
  // acc += sgn * (a * (cs + i sn)), complex product (ac - bd, ad + bc)
  const double re = a.x * cs - a.y * sn;
  const double im = a.x * sn + a.y * cs;
  acc.x += sgn * re;
  acc.y += sgn * im;
}

__device__ __forceinline__ double wave_sum(double v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__device__ double block_sum(double v, double *red)
{
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += red[w];
  return r;
}

struct FFArgs {
  int N_PX, N_PY, NPXa, NPYa, off;
  int top, bottom, left, right;
  double cx, cy, k, z0, coef_re, coef_im, scale;
  const double *cosA, *sinA;
};

// A field of either precision read in FP64 at physical cell (x, y): stored at
// (x + off, y + off) of an nx x ny array with row pitch `pitch` (FP64 grids: the
// solver layout, off 1 for the padded MPI arrays; FP32 grids: the physical
// layout, off 0).  at() reads 0 outside the stored extent (the MPI ghost ring).
template <class V>
struct FieldView {
  const V *p;
  int nx, ny, off, pitch;
  __device__ __forceinline__ double2 atu(int x, int y) const
  {
    const V v = p[(size_t)(x + off) * pitch + (y + off)];
    return make_double2((double)v.x, (double)v.y);
  }
  __device__ __forceinline__ double2 at(int x, int y) const
  {
    const int ax = x + off, ay = y + off;
    if (ax < 0 || ay < 0 || ax >= nx || ay >= ny) return make_double2(0.0, 0.0);
    return atu(x, y);
  }
};

// contour cells of ntffTM_Frequency (ntffTM.c:72-158): bottom, right, top, left
__host__ __device__ __forceinline__ int far_tm_cells(const FFArgs &a)
{
  return 2 * (a.right - a.left) + 2 * (a.top - a.bottom);
}

// contour cell c of ntffTM_Frequency: its terms added to Nz, Lx, Ly
template <class V>
__device__ __forceinline__ void far_tm_cell(int c, const FieldView<V> &Ez, const FieldView<V> &Hx, const FieldView<V> &Hy,
                                            const FFArgs &a, double rx, double ry, double2 &Nz, double2 &Lx, double2 &Ly)
{
  const int nx = a.right - a.left, ny = a.top - a.bottom;
  int side, px, py;
  if (c < nx) { side = 0; px = a.left + c; py = a.bottom; }
  else if (c < nx + ny) { side = 1; px = a.right; py = a.bottom + (c - nx); }
  else if (c < 2 * nx + ny) { side = 2; px = a.left + (c - nx - ny); py = a.top; }
  else { side = 3; px = a.left; py = a.bottom + (c - 2 * nx - ny); }
  const double r2x = px - a.cx, r2y = py - a.cy;
  const double ip = rx * r2x + ry * r2y;
  double sn, cs;
  sincos(a.k * ip, &sn, &cs);
  const double2 ez = Ez.atu(px, py);
  double2 hv;
  if (side == 0 || side == 2) {
    const double2 h0 = Hx.atu(px, py), h1 = Hx.atu(px, py - 1);
    hv.x = 0.5 * (h0.x + h1.x);
    hv.y = 0.5 * (h0.y + h1.y);
  } else {
    const double2 h0 = Hy.atu(px, py), h1 = Hy.atu(px - 1, py);
    hv.x = 0.5 * (h0.x + h1.x);
    hv.y = 0.5 * (h0.y + h1.y);
  }
  const double sg = (side < 2) ? 1.0 : -1.0;
  cmul_acc(Nz, hv, cs, sn, sg);
  if (side == 0 || side == 2) cmul_acc(Lx, ez, cs, sn, sg);
  else cmul_acc(Ly, ez, cs, sn, sg);
}

// v = {Nz, Lx, Ly}: Lphi = -Lx sin + Ly cos; E = coef (Z0 Nz + Lphi) sqrt(h)
__device__ __forceinline__ double2 far_tm_result(const double v[6], const FFArgs &a, double rx, double ry)
{
  const double sn = ry, cs = rx;
  const double lpr = -v[2] * sn + v[4] * cs, lpi = -v[3] * sn + v[5] * cs;
  const double tr = a.z0 * v[0] + lpr, ti = a.z0 * v[1] + lpi;
  double2 r;
  r.x = (a.coef_re * tr - a.coef_im * ti) * a.scale;
  r.y = (a.coef_re * ti + a.coef_im * tr) * a.scale;
  return r;
}

// side extents of single-rank ntffTE_FrequencySplit (ntffTE.c:341-470) in padded coordinates
struct TeSides {
  int L, B, nb, nr, nt, nl;
};
__host__ __device__ __forceinline__ TeSides far_te_sides(const FFArgs &a)
{
  const int SPX = a.N_PX + 2, SPY = a.N_PY + 2;
  const int L = a.left > 1 ? a.left : 1, Rx = a.right < SPX - 1 ? a.right : SPX - 1;
  const int B = a.bottom > 1 ? a.bottom : 1, T = a.top < SPY - 1 ? a.top : SPY - 1;
  const int w = Rx - L > 0 ? Rx - L : 0, h = T - B > 0 ? T - B : 0;
  TeSides t;
  t.L = L;
  t.B = B;
  t.nb = (0 < a.bottom && a.bottom < SPY - 1) ? w : 0;
  t.nr = (0 < a.right && a.right < SPX - 1) ? h : 0;
  t.nt = (0 < a.top && a.top < SPY - 1) ? w : 0;
  t.nl = (0 < a.left && a.left < SPX - 1) ? h : 0;
  return t;
}

// contour cell c of the TE far field: its terms added to Nx, Ny, Lz
template <class V>
__device__ __forceinline__ void far_te_cell(int c, const TeSides &t, const FieldView<V> &Ex, const FieldView<V> &Ey,
                                            const FieldView<V> &Hz, const FFArgs &a, double rx, double ry, double2 &Nx,
                                            double2 &Ny, double2 &Lz)
{
  int side, pi, pj;
  double r2x, r2y;
  if (c < t.nb) { side = 0; pi = t.L + c; pj = a.bottom; r2x = pi - a.cx + 0.5; r2y = pj - a.cy; }
  else if (c < t.nb + t.nr) { side = 1; pi = a.right; pj = t.B + (c - t.nb); r2x = pi - a.cx; r2y = pj - a.cy + 0.5; }
  else if (c < t.nb + t.nr + t.nt) { side = 2; pi = t.L + (c - t.nb - t.nr); pj = a.top; r2x = pi - a.cx + 0.5; r2y = pj - a.cy; }
  else { side = 3; pi = a.left; pj = t.B + (c - t.nb - t.nr - t.nt); r2x = pi - a.cx; r2y = pj - a.cy + 0.5; }
  const double ip = rx * r2x + ry * r2y;
  double sn, cs;
  sincos(a.k * ip, &sn, &cs);
  // padded (pi, pj) is physical (pi - 1, pj - 1); the ghost ring reads 0
  const double2 h0 = Hz.at(pi - 1, pj - 1);
  const double2 h1 = (side == 0 || side == 2) ? Hz.at(pi - 1, pj - 2) : Hz.at(pi - 2, pj - 1);
  double2 hv;
  hv.x = 0.5 * (h0.x + h1.x);
  hv.y = 0.5 * (h0.y + h1.y);
  const double2 ev = (side == 0 || side == 2) ? Ex.at(pi - 1, pj - 1) : Ey.at(pi - 1, pj - 1);
  const double sg = (side < 2) ? -1.0 : 1.0;
  if (side == 0 || side == 2) cmul_acc(Nx, hv, cs, sn, sg);
  else cmul_acc(Ny, hv, cs, sn, sg);
  cmul_acc(Lz, ev, cs, sn, sg);
}

// v = {Nx, Ny, Lz}: N_phi = -Nx sin + Ny cos; E = coef (Z0 N_phi - Lz) h
__device__ __forceinline__ double2 far_te_result(const double v[6], const FFArgs &a, double rx, double ry)
{
  const double sn = ry, cs = rx;
  const double npr = -v[0] * sn + v[2] * cs, npi = -v[1] * sn + v[3] * cs;
  const double tr = a.z0 * npr - v[4], ti = a.z0 * npi - v[5];
  double2 r;
  r.x = (a.coef_re * tr - a.coef_im * ti) * a.scale;
  r.y = (a.coef_re * ti + a.coef_im * tr) * a.scale;
  return r;
}

__host__ __device__ __forceinline__ int far_cells(int mode, const FFArgs &a)
{
  if (mode == FDTD_MODE_TM) return far_tm_cells(a);
  const TeSides t = far_te_sides(a);
  return t.nb + t.nr + t.nt + t.nl;
}

// sums over contour cells [c0, c1) of angle `ang`: {Nz, Lx, Ly} (TM) or {Nx, Ny, Lz} (TE), block-reduced
template <int MODE, class V>
__device__ __forceinline__ void far_sums(const FieldView<V> &f0, const FieldView<V> &f1, const FieldView<V> &f2,
                                         const FFArgs &a, int ang, int c0, int c1, double v[6], double *red)
{
  const double rx = a.cosA[ang], ry = a.sinA[ang];
  double2 s0 = {0, 0}, s1 = {0, 0}, s2 = {0, 0};
  const TeSides t = far_te_sides(a);
  for (int c = c0 + (int)threadIdx.x; c < c1; c += blockDim.x) {
    if (MODE == FDTD_MODE_TM) far_tm_cell(c, f0, f1, f2, a, rx, ry, s0, s1, s2);
    else far_te_cell(c, t, f0, f1, f2, a, rx, ry, s0, s1, s2);
  }
  v[0] = s0.x; v[1] = s0.y; v[2] = s1.x; v[3] = s1.y; v[4] = s2.x; v[5] = s2.y;
  for (int q = 0; q < 6; q++) v[q] = block_sum(v[q], red);
}

// ntffTM_Frequency (ntffTM.c:72-158) / single-rank ntffTE_FrequencySplit
// (ntffTE.c:341-470) on the physical cells of either layout: one workgroup per angle
template <int MODE>
__global__ __launch_bounds__(256) void k_far(const double2 *F0, const double2 *F1, const double2 *F2, FFArgs a, double2 *out)
{
  __shared__ double red[8];
  const int ang = blockIdx.x;
  const FieldView<double2> f0 = {F0, a.NPXa, a.NPYa, a.off, a.NPYa}, f1 = {F1, a.NPXa, a.NPYa, a.off, a.NPYa},
                           f2 = {F2, a.NPXa, a.NPYa, a.off, a.NPYa};
  double v[6];
  far_sums<MODE>(f0, f1, f2, a, ang, 0, far_cells(MODE, a), v, red);
  if (threadIdx.x == 0) out[ang] = MODE == FDTD_MODE_TM ? far_tm_result(v, a, a.cosA[ang], a.sinA[ang])
                                                        : far_te_result(v, a, a.cosA[ang], a.sinA[ang]);
}

// ---------------------------------------------------------------- allocation-free far field
//
// The same far field without any device allocation: the fields are read in
// place (FP32 grids in FP32, converted per value, instead of exported to FP64
// arrays first), the contour of each angle is cut into parts of FF_NA_CHUNK
// cells, one workgroup per (part, angle), and the parts' sums and the 360
// results live in the file-scope array `scratch` below.  Every call on a GPU
// shares that GPU's copy: calls that overlap on one GPU race.
#define FF_NA_CHUNK 1024
#define FF_NA_PARTS 64  // at most 65536 contour cells
#define FF_NA_SCRATCH (360 * FF_NA_PARTS * 3 + 360)
static __device__ __align__(256) double2 scratch[FF_NA_SCRATCH];  // 1.11 MB per GPU

template <int MODE, class V>
__global__ __launch_bounds__(256) void k_far_part(FieldView<V> f0, FieldView<V> f1, FieldView<V> f2, FFArgs a, int n,
                                                  double2 *part)
{
  __shared__ double red[8];
  const int p = blockIdx.x, ang = blockIdx.y;
  double v[6];
  far_sums<MODE>(f0, f1, f2, a, ang, p * FF_NA_CHUNK, min(n, (p + 1) * FF_NA_CHUNK), v, red);
  if (threadIdx.x == 0) {
    double2 *o = part + ((size_t)ang * FF_NA_PARTS + p) * 3;
    o[0] = make_double2(v[0], v[1]);
    o[1] = make_double2(v[2], v[3]);
    o[2] = make_double2(v[4], v[5]);
  }
}

// the parts of each angle summed in order, then the angle's far-field value
template <int MODE>
__global__ __launch_bounds__(64) void k_far_join(const double2 *part, int nparts, FFArgs a, double2 *out)
{
  const int ang = blockIdx.x * 64 + threadIdx.x;
  if (ang >= 360) return;
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int p = 0; p < nparts; p++) {
    const double2 *o = part + ((size_t)ang * FF_NA_PARTS + p) * 3;
    for (int q = 0; q < 3; q++) {
      v[2 * q] += o[q].x;
      v[2 * q + 1] += o[q].y;
    }
  }
  out[ang] = MODE == FDTD_MODE_TM ? far_tm_result(v, a, a.cosA[ang], a.sinA[ang])
                                  : far_te_result(v, a, a.cosA[ang], a.sinA[ang]);
}

// ---------------------------------------------------------------- time-domain far field + spectra

#define NTFF_NUM 8192

__device__ __forceinline__ int bitrev13(int v) { return (int)(__brev((unsigned)v) >> 19); }

// TimeTranslate + 8192-point FFT (fft1x, flag 0) + |X|^2 interpolation, one
// workgroup per angle, the transform held in LDS (128 KiB)
__global__ __launch_bounds__(1024) void k_time_spectra(const double2 *U0, const double2 *U1, const double2 *U2, int arraySize,
                                                       int maxTime, int tm, double z0, double coef_re, double coef_im,
                                                       const double *cosA, const double *sinA, const double2 *tw,
                                                       double c0h, double *out)
{
  extern __shared__ double2 buf[];
  const int ang = blockIdx.x;
  const long k0 = (long)ang * arraySize;
  const double cphi = cosA[ang], sphi = sinA[ang];
  const int ncopy = min(maxTime, NTFF_NUM);
  for (int i = threadIdx.x; i < NTFF_NUM; i += blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (i < ncopy) {
      double tr, ti;
      if (tm) {
        // Eth = coef (-Z0 WTH - UPH), WTH = -Wz, UPH = -Ux sin + Uy cos
        const double2 ux = U0[k0 + i], uy = U1[k0 + i], wz = U2[k0 + i];
        const double wthr = wz.x * -1.0, wthi = wz.y * -1.0;
        const double uphr = ux.x * -sphi + uy.x * cphi, uphi = ux.y * -sphi + uy.y * cphi;
        tr = -z0 * wthr - uphr;
        ti = -z0 * wthi - uphi;
      } else {
        // Eph = coef (-Z0 WPH + UTH), WPH = -Wx sin + Wy cos, UTH = -Uz
        const double2 wx = U0[k0 + i], wy = U1[k0 + i], uz = U2[k0 + i];
        const double wphr = wx.x * -sphi + wy.x * cphi, wphi = wx.y * -sphi + wy.y * cphi;
        const double uthr = uz.x * -1.0, uthi = uz.y * -1.0;
        tr = -z0 * wphr + uthr;
        ti = -z0 * wphi + uthi;
      }
      v.x = coef_re * tr - coef_im * ti;
      v.y = coef_re * ti + coef_im * tr;
    }
    buf[i] = v;
  }
  __syncthreads();
  // radix-2 decimation in frequency, twiddle e^{+i pi k / xp2} = tw[k * 4096/xp2]
  for (int xp2 = NTFF_NUM / 2, sh = 0; xp2 >= 1; xp2 >>= 1, sh++) {
    const int xp = xp2 * 2;
    for (int bI = threadIdx.x; bI < NTFF_NUM / 2; bI += blockDim.x) {
      const int kk = bI & (xp2 - 1);
      const int base = (bI / xp2) * xp + kk;
      const double2 p = buf[base], q = buf[base + xp2];
      const double2 w = tw[kk << sh];
      const double tr = p.x - q.x, ti = p.y - q.y;
      buf[base] = make_double2(p.x + q.x, p.y + q.y);
      buf[base + xp2] = make_double2(tr * w.x - ti * w.y, tr * w.y + ti * w.x);
    }
    __syncthreads();
  }
  for (int l = threadIdx.x; l <= 320; l += blockDim.x) {
    const int lambda_nm = 380 + l;
    double p = c0h * NTFF_NUM / lambda_nm;
    const int index = (int)floor(p);
    p = p - index;
    if (index < 0 || index + 1 >= NTFF_NUM) {  // beyond the transform (the reference reads past its buffer)
      out[(size_t)l * 360 + ang] = __builtin_nan("");
      continue;
    }
    const double2 e0 = buf[bitrev13(index)], e1 = buf[bitrev13(index + 1)];
    const double n0 = e0.x * e0.x + e0.y * e0.y, n1 = e1.x * e1.x + e1.y * e1.y;
    out[(size_t)l * 360 + ang] = ((1 - p) * n0 + p * n1) / NTFF_NUM;
  }
}

// ---------------------------------------------------------------- eps fill

// variant: 0 MPI / split / NS TM, 1 plain UPML, 2 NS TE (Hz sampled at D_XY and a zero
// border, nsFdtdTE.c:116-121)
__global__ __launch_bounds__(256) void k_fill_eps(const fdtd_model_state *m, double *e0, double *e1, double *e2,
                                                  int NPXa, int NPYa, int off, int tm, int variant)
{
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y * 4 + threadIdx.y;
  if (i >= NPXa || j >= NPYa) return;
  const size_t k = (size_t)i * NPYa + j;
  const int plain = variant == 1;
  if ((off || variant == 2) && (i == 0 || j == 0 || i == NPXa - 1 || j == NPYa - 1)) {
    e0[k] = 0.0; e1[k] = 0.0; e2[k] = 0.0;   // ghost ring / unset border (never read by the updates)
    return;
  }
  const double x = i - off, y = j - off;
  if (tm) {
    e0[k] = fdtd_model_eps(m, x, y, 1, 1);
    if (plain) {  // fdtdTM_upml.c:238-239 samples D_XY (Q19)
      e1[k] = fdtd_model_eps(m, x, y + 0.5, 1, 1);
      e2[k] = fdtd_model_eps(m, x + 0.5, y, 1, 1);
    } else {
      e1[k] = fdtd_model_eps(m, x, y + 0.5, 0, 1);
      e2[k] = fdtd_model_eps(m, x + 0.5, y, 1, 0);
    }
  } else {
    e0[k] = fdtd_model_eps(m, x + 0.5, y, 0, 1);
    e1[k] = fdtd_model_eps(m, x, y + 0.5, 1, 0);
    e2[k] = variant == 2 ? fdtd_model_eps(m, x + 0.5, y + 0.5, 1, 1)
                         : 0.5 * (fdtd_model_eps(m, x + 0.5, y + 0.5, 1, 0) + fdtd_model_eps(m, x + 0.5, y + 0.5, 0, 1));
  }
}

// ---------------------------------------------------------------- host side

#define HIPCHK(x)                                                                            \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) return fdtd_fail(FDTD_EHIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

struct fdtd_grid {
  fdtd_grid_desc d;
  int NPXa, NPYa, off;
  size_t ncell;
  double2 *F[9];
  double2 **Fdev;  // device copy of F[] for the gather kernel
  double *eps[3];
  double *tab;
  Tabs tabs;
  // NTFF
  Contour ct;
  double2 *ntb[3];
  double2 *nt[3];
  size_t ntn;
  double2 *cE, *cH;
  BinPart *binPart;  // records of k_ntff_bin: parts, samples
  BinRec *binRec;
  int nparts;
  int nbuf;  // steps the sample buffers cE / cH hold
  // running DFT of the contour samples (fdtd_hip_grid_set_dft)
  int nf;
  double *dftOmega;      // [nf] angular frequencies (rad per step)
  double2 *dftE, *dftH;  // [nf][total] FP64 accumulators
  double *dftTs;         // [360][total] the binning's time shifts
  double2 *dftOut;       // [nf][3][360]
  // samples waiting for the DFT: columns [0, sFill) of the [sCap][total]
  // buffers hold consecutive steps from base time sTb0 (running DFT without the
  // time-domain binning); dftPh [sCap][nf] the phase factors of a launch
  double2 *sE, *sH, *dftPh;
  int sCap, sFill;
  long sTb0;
  int ctrSet;            // FP32: the fused kernel samples the contour
  // FP32: stream of the sample packing, binning and running DFT, beside the next
  // block's step kernel; ev_k orders it after the block that took the samples
  hipStream_t ns;
  hipEvent_t ev_k;
  int overlap;  // -1: by grid size (on up to NTFF_OVERLAP_CELLS cells), 0: in order on the step stream, 1: on ns
  double *cosA, *sinA, *cosB, *sinB;
  double coef;
  double2 *ffout;
  unsigned long long *nfcnt;  // [3] non-finite values per physical field (fdtd_hip_grid_count_nonfinite)
  double *spec;
  double2 *tw;
  fdtd_model_state *mstate;
  FastState *fs;  // FP32 fused path (precision FDTD_FP32); F[] then serve as FP64 export buffers
  double *nsc;    // NS flavour: [nns][ncell] per-cell coefficients (fdtd_hip_grid_set_ns_coefs)
  int nns;        // NS flavour: TM 5, TE 10 maps
  double ns_r2;   // NS flavour: r/2 of the second differences
};

#define NT_GUARD NT_GUARD_DEV

// contour segments of the per-step NTFF in array coordinates
static void build_contour(fdtd_grid *g)
{
  const fdtd_grid_desc &d = g->d;
  const int NPY = g->NPYa;
  const bool tm = d.mode == FDTD_MODE_TM, mpi = d.flavour == FDTD_FLAVOUR_MPI;
  Contour &ct = g->ct;
  memset(&ct, 0, sizeof ct);
  const int tp = d.ntff_top, bm = d.ntff_bottom, rt = d.ntff_right, lt = d.ntff_left;
  // field slots: TM Ez 0, Hx 3, Hy 6; TE Ex 0, Ey 3, Hz 6
  if (mpi) {
    const int SPX = g->NPXa, SPY = g->NPYa;
    const int cx = d.N_PX / 2, cy = d.N_PY / 2;
    const int L = std::max(1, lt), R = std::min(SPX, rt), B = std::max(1, bm), T = std::min(SPY, tp);
    const double hx0 = tm ? 0.0 : 0.5, hy0 = tm ? 0.0 : 0.5;
    Seg s[4];
    memset(s, 0, sizeof s);
    // bottom (i = L..R-1, j = bm)
    s[0].n = (0 < bm && bm < SPY - 1) ? std::max(0, R - L) : 0;
    s[0].k0 = (long)L * NPY + bm; s[0].dk = NPY;
    s[0].r2x0 = L - cx + hx0; s[0].r2y0 = bm - cy; s[0].dr2x = 1; s[0].dr2y = 0;
    // right (i = rt, j = B..T-1)
    s[1].n = (0 < rt && rt < SPX - 1) ? std::max(0, T - B) : 0;
    s[1].k0 = (long)rt * NPY + B; s[1].dk = 1;
    s[1].r2x0 = rt - cx; s[1].r2y0 = B - cy + hy0; s[1].dr2x = 0; s[1].dr2y = 1;
    // top
    s[2].n = (0 < tp && tp < SPY - 1) ? std::max(0, R - L) : 0;
    s[2].k0 = (long)L * NPY + tp; s[2].dk = NPY;
    s[2].r2x0 = L - cx + hx0; s[2].r2y0 = tp - cy; s[2].dr2x = 1; s[2].dr2y = 0;
    // left
    s[3].n = (0 < lt && lt < SPX) ? std::max(0, T - B) : 0;
    s[3].k0 = (long)lt * NPY + B; s[3].dk = 1;
    s[3].r2x0 = lt - cx; s[3].r2y0 = B - cy + hy0; s[3].dr2x = 0; s[3].dr2y = 1;
    for (int q = 0; q < 4; q++) ct.s[q] = s[q];
  } else {
    const int cx = d.ntff_cx, cy = d.ntff_cy;
    Seg s[4];
    memset(s, 0, sizeof s);
    const double hx0 = tm ? 0.0 : 0.5, hy0 = tm ? 0.0 : 0.5;
    s[0].n = rt - lt; s[0].k0 = (long)lt * NPY + bm; s[0].dk = NPY;
    s[0].r2x0 = lt - cx + hx0; s[0].r2y0 = bm - cy; s[0].dr2x = 1; s[0].dr2y = 0;
    s[1].n = tp - bm; s[1].k0 = (long)rt * NPY + bm; s[1].dk = 1;
    s[1].r2x0 = rt - cx; s[1].r2y0 = bm - cy + hy0; s[1].dr2x = 0; s[1].dr2y = 1;
    s[2].n = rt - lt; s[2].k0 = (long)lt * NPY + tp; s[2].dk = NPY;
    s[2].r2x0 = lt - cx + hx0; s[2].r2y0 = tp - cy; s[2].dr2x = 1; s[2].dr2y = 0;
    s[3].n = tp - bm; s[3].k0 = (long)lt * NPY + bm; s[3].dk = 1;
    s[3].r2x0 = lt - cx; s[3].r2y0 = bm - cy + hy0; s[3].dr2x = 0; s[3].dr2y = 1;
    for (int q = 0; q < 4; q++) ct.s[q] = s[q];
  }
  // field roles and signs
  for (int q = 0; q < 4; q++) {
    Seg &s = ct.s[q];
    const bool horiz = (q == 0 || q == 2);
    const double sgn = tm ? (q < 2 ? 1.0 : -1.0) : (q < 2 ? -1.0 : 1.0);
    if (tm) {
      s.ef = 0;
      s.hf = horiz ? 3 : 6;
      s.ua = horiz ? 0 : 1;  // Ux / Uy
      s.wa = 2;              // Wz
    } else {
      s.ef = horiz ? 0 : 3;
      s.hf = 6;
      s.ua = 2;              // Uz
      s.wa = horiz ? 0 : 1;  // Wx / Wy
    }
    s.hnb = horiz ? -1 : -NPY;
    s.es = sgn;
    s.hs = 0.5 * sgn;
  }
  int off = 0;
  for (int q = 0; q < 4; q++) {
    ct.s[q].off = off;
    off += ct.s[q].n;
  }
  ct.total = off;
}

// retarded-time shift of every (angle, contour cell), formed on the host with
// the reference's own arithmetic: direct per cell for the MPI solvers
// (mpiTM_UPML.c:894), accumulated along each side for the plain ones (ntffTM.c:329-334)
// The reference forms the angle as ang*M_PI/180.0 in the MPI NTFF, the far
// field and ntffTE_TimeTranslate (angle table A), but as ang*ToRad with
// ToRad = M_PI/180.0 in ntffTM_TimeCalc / ntffTE_TimeCalc / ntffTM_TimeTranslate
// (table B, ntffTM.c:314-318, ntffTE.c:89-92, ntffTM.c:169-173); the two differ in the last bit.
// the per-(angle, contour cell) time shifts (tab, only when with_tab) and the
// angle tables
static void build_time_shift(const fdtd_grid *g, std::vector<double> &tab, std::vector<double> &mn, std::vector<double> &mx,
                             std::vector<double> &ca, std::vector<double> &sa, std::vector<double> &cb, std::vector<double> &sb,
                             bool with_tab = true)
{
  const fdtd_grid_desc &d = g->d;
  const Contour &ct = g->ct;
  const bool mpi = d.flavour == FDTD_FLAVOUR_MPI;
  tab.assign(with_tab ? (size_t)360 * ct.total : 0, 0.0);
  mn.assign(360, 0.0);
  mx.assign(360, 0.0);
  ca.resize(360);
  sa.resize(360);
  cb.resize(360);
  sb.resize(360);
  const double ToRad = M_PI / 180.0;
  for (int ang = 0; ang < 360; ang++) {
    const double radA = ang * M_PI / 180.0, radB = ang * ToRad;
    ca[ang] = cos(radA);
    sa[ang] = sin(radA);
    cb[ang] = cos(radB);
    sb[ang] = sin(radB);
    if (!with_tab) continue;
    const double r1x = mpi ? ca[ang] : cb[ang], r1y = mpi ? sa[ang] : sb[ang];
    double *t = &tab[(size_t)ang * ct.total];
    for (int q = 0; q < 4; q++) {
      const Seg &s = ct.s[q];
      if (mpi) {
        for (int c = 0; c < s.n; c++) {
          const double r2x = s.r2x0 + c * s.dr2x, r2y = s.r2y0 + c * s.dr2y;
          t[s.off + c] = -(r1x * r2x + r1y * r2y) / d.c0 + d.rf_per_c;
        }
      } else {
        const double r1x_per_c = r1x / d.c0, r1y_per_c = r1y / d.c0;
        double ts = -(r1x_per_c * s.r2x0 + r1y_per_c * s.r2y0) + d.rf_per_c;
        const double step = (q == 0 || q == 2) ? r1x_per_c : r1y_per_c;
        for (int c = 0; c < s.n; c++) {
          t[s.off + c] = ts;
          ts -= step;
        }
      }
    }
    double lo = 1e300, hi = -1e300;
    for (int p = 0; p < ct.total; p++) {
      lo = std::min(lo, t[p]);
      hi = std::max(hi, t[p]);
    }
    mn[ang] = ct.total ? lo : 0.0;
    mx[ang] = ct.total ? hi : 0.0;
  }
}

// The records of k_ntff_bin: the samples of each (angle, target array) sorted by
// offset d, cut into parts of at most NTFF_GRP x NTFF_RUN samples and NTFF_WIN
// offsets, each part split into NTFF_GRP contiguous runs stored [iteration][group].
static int build_bin_groups(fdtd_grid *g, const std::vector<double> &tab)
{
  const Contour &ct = g->ct;
  const int P = ct.total;
  // a record addresses its sample's first step directly: cell p of E at p * nb,
  // of H at (P + p) * nb (the sample buffer holds E then H, nb steps per cell)
  const int nb = g->nbuf;
  if ((size_t)2 * P * nb > (size_t)INT_MAX) return fdtd_fail(FDTD_ENOMEM, "NTFF contour of %d cells too long", P);
  std::vector<BinPart> parts;
  std::vector<BinRec> rec;
  std::vector<int> side(P);
  for (int q = 0; q < 4; q++)
    for (int c = 0; c < ct.s[q].n; c++) side[ct.s[q].off + c] = q;
  struct Smp {
    int d, p;
    double a;
  };
  std::vector<Smp> by[3];
  for (int ang = 0; ang < 360; ang++) {
    const double *ts = &tab[(size_t)ang * P];
    for (int arr = 0; arr < 3; arr++) by[arr].clear();
    for (int p = 0; p < P; p++) {
      const Seg &sg = ct.s[side[p]];
      const int dE = (int)floor(ts[p] + 0.5), dH = (int)floor(ts[p]) + 1;
      by[sg.ua].push_back({dE, p * nb, (0.5 + ts[p]) - (double)dE});
      by[sg.wa].push_back({dH, (P + p) * nb, ts[p] - (double)(dH - 1)});
    }
    for (int arr = 0; arr < 3; arr++) {
      std::vector<Smp> &v = by[arr];
      std::stable_sort(v.begin(), v.end(), [](const Smp &x, const Smp &y) { return x.d < y.d; });
      size_t x0 = 0;
      while (x0 < v.size()) {
        const int dlo = v[x0].d;
        size_t x1 = x0;
        while (x1 < v.size() && x1 - x0 < NTFF_GRP * NTFF_RUN && v[x1].d < dlo + NTFF_WIN) x1++;
        BinPart W;
        W.rec0 = (int)rec.size();
        W.dlo = dlo;
        W.nd = v[x1 - 1].d - dlo + 1;
        W.row = ang;
        W.arr = arr;
        const int cnt = (int)(x1 - x0);
        W.per = (cnt + NTFF_GRP - 1) / NTFF_GRP;
        rec.resize(rec.size() + (size_t)W.per * NTFF_GRP, BinRec{-1, 0, 0.0});
        for (int t = 0; t < NTFF_GRP; t++)
          for (int i = 0; i < W.per; i++) {
            const size_t x = x0 + (size_t)t * W.per + i;
            if (x >= x1) break;
            rec[(size_t)W.rec0 + (size_t)i * NTFF_GRP + t] = BinRec{v[x].p, v[x].d - dlo, v[x].a};
          }
        parts.push_back(W);
        x0 = x1;
      }
    }
  }
  if (rec.size() > (size_t)INT_MAX) return fdtd_fail(FDTD_ENOMEM, "NTFF binning records exceed 2^31");
  g->nparts = (int)parts.size();
  HIPCHK(dev_malloc(&g->binPart, std::max<size_t>(1, parts.size()) * sizeof(BinPart)));
  HIPCHK(dev_malloc(&g->binRec, std::max<size_t>(1, rec.size()) * sizeof(BinRec)));
  if (!parts.empty()) HIPCHK(hipMemcpy(g->binPart, parts.data(), parts.size() * sizeof(BinPart), hipMemcpyHostToDevice));
  if (!rec.empty()) HIPCHK(hipMemcpy(g->binRec, rec.data(), rec.size() * sizeof(BinRec), hipMemcpyHostToDevice));
  return FDTD_OK;
}

// the first n steps of the samples in cE / cH ([cell][nbuf]) into the accumulators
static void launch_bin(fdtd_grid *g, int n, double time0, hipStream_t st)
{
  const fdtd_grid_desc &d = g->d;
  const long tb0 = (long)llround(time0) - 1;
  const int useCoef = d.flavour == FDTD_FLAVOUR_PLAIN ? 0 : 1;
  if (g->nparts <= 0) return;
#define NTFF_BIN(NS)                                                                                              \
  hipLaunchKernelGGL(k_ntff_bin<NS>, dim3(g->nparts), dim3(NTFF_GRP * NS), 0, st, g->cE, n, g->binPart, g->binRec,        \
                     g->nt[0], g->nt[1], g->nt[2], d.array_size, tb0, g->coef, useCoef)
  if (n <= 2)
    NTFF_BIN(2);
  else if (n <= 4)
    NTFF_BIN(4);
  else if (n <= 8)
    NTFF_BIN(8);
  else
    NTFF_BIN(16);
#undef NTFF_BIN
}

// n steps of samples at E[s * ss + c * sp], H[...] taken at base times
// tb0 + s into the running DFT
static void dft_add(fdtd_grid *g, const double2 *E, const double2 *H, long ss, long sp, int n, long tb0, hipStream_t st)
{
  const int nf = g->nf, P = g->ct.total;
  if (n <= 0 || P <= 0 || nf <= 0) return;
  hipLaunchKernelGGL(k_dft_phase, dim3((n * nf + 255) / 256), dim3(256), 0, st, g->dftOmega, nf, tb0, n, g->dftPh);
  // frequencies per wavefront: the NFL (<= DFT_NFL_MAX) that leaves the fewest
  // unused frequency slots in the launch's blocks of DFT_WAVES * NFL
  int nfl = 1, best = INT_MAX;
  for (int c = DFT_NFL_MAX; c >= 1; c--) {
    const int slots = (nf + DFT_WAVES * c - 1) / (DFT_WAVES * c) * DFT_WAVES * c;
    if (slots < best) best = slots, nfl = c;
  }
#define CTR_DFT(NFL)                                                                                              \
  hipLaunchKernelGGL(k_ctr_dft<NFL>, dim3((P + 63) / 64, (nf + DFT_WAVES * NFL - 1) / (DFT_WAVES * NFL)),          \
                     dim3(64 * DFT_WAVES), 0, st, E, H, ss, sp, n, P, g->dftPh, nf, g->dftE, g->dftH)
  switch (nfl) {
    case 1: CTR_DFT(1); break;
    case 2: CTR_DFT(2); break;
    case 3: CTR_DFT(3); break;
    case 4: CTR_DFT(4); break;
    case 5: CTR_DFT(5); break;
    case 6: CTR_DFT(6); break;
    case 7: CTR_DFT(7); break;
    default: CTR_DFT(8); break;
  }
#undef CTR_DFT
}

// the staged samples into the running DFT
static void dft_flush(fdtd_grid *g, hipStream_t st)
{
  if (g->sFill) dft_add(g, g->sE, g->sH, g->ct.total, 1, g->sFill, g->sTb0, st);
  g->sFill = 0;
}

// the staging column of n steps from time0 (the first at base time
// llround(time0) - 1): the staged steps go to the DFT first when these do not
// follow them or do not fit
static int dft_stage(fdtd_grid *g, int n, double time0, hipStream_t st)
{
  const long tb0 = (long)llround(time0) - 1;
  if (g->sFill && (tb0 != g->sTb0 + g->sFill || g->sFill + n > g->sCap)) dft_flush(g, st);
  if (!g->sFill) g->sTb0 = tb0;
  const int col = g->sFill;
  g->sFill += n;
  return col;
}

// contour samples of n inner steps of the FP32 kernel ([n][3][total] float2:
// signed E, H on the side, H one cell back) -> the packing of k_ntff_gather,
// step s of cell p at cE[s * ss + p * sp] (the binning's [cell][n], or the
// DFT's staging [step][cell])
__global__ __launch_bounds__(256) void k_ctr_combine(const float2 *smp, Contour ct, long ss, long sp, double2 *cE,
                                                     double2 *cH)
{
  const int p = blockIdx.x * 256 + threadIdx.x, s = blockIdx.y;
  if (p >= ct.total) return;
  int q = 0;
  while (q < 3 && p >= ct.s[q + 1].off) q++;
  const double hs = ct.s[q].hs;
  const float2 *sm = smp + (size_t)s * 3 * ct.total;
  const float2 e = sm[p], h0 = sm[ct.total + p], h1 = sm[2 * ct.total + p];
  cE[(size_t)s * ss + (size_t)p * sp] = make_double2((double)e.x, (double)e.y);
  cH[(size_t)s * ss + (size_t)p * sp] = make_double2(hs * ((double)h0.x + (double)h1.x), hs * ((double)h0.y + (double)h1.y));
}

#define NTFF_OVERLAP_CELLS (8L << 20)

// after a launch of n steps of an FP32 grid: pack the samples, bin every step at
// its time and add them to the running DFT, on the grid's NTFF stream (the
// next block's step kernel runs meanwhile)
static int grid_bin_samples(void *ctx, const float2 *samples, int n, const double *times, hipEvent_t consumed)
{
  fdtd_grid *g = (fdtd_grid *)ctx;
  if (n > g->nbuf) return fdtd_fail(FDTD_EINVAL, "%d steps exceed the %d-step sample buffers", n, g->nbuf);
  for (int s = 1; s < n; s++)
    if (times[s] != times[0] + s) return fdtd_fail(FDTD_ESTATE, "steps of a block are not consecutive");
  // beside the next block on small grids, whose step kernel leaves the GPU
  // partly idle; in order on large ones, whose step kernel fills it (there the
  // overlap only stretches both)
  const bool ovl = g->overlap < 0 ? g->ncell <= NTFF_OVERLAP_CELLS : g->overlap != 0;
  hipStream_t st = 0;
  if (ovl) {
    if (!g->ns) {
      HIPCHK(hipStreamCreateWithFlags(&g->ns, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&g->ev_k, hipEventDisableTiming));
    }
    HIPCHK(hipEventRecord(g->ev_k, 0));
    HIPCHK(hipStreamWaitEvent(g->ns, g->ev_k, 0));
    st = g->ns;
  }
  const dim3 cgrid((g->ct.total + 255) / 256, n);
  if (g->ntn) {
    // binned step by step: [cell][nbuf], the first n steps of each cell; the DFT reads the same buffers
    hipLaunchKernelGGL(k_ctr_combine, cgrid, dim3(256), 0, st, samples, g->ct, 1L, (long)g->nbuf, g->cE, g->cH);
    HIPCHK(hipEventRecord(consumed, st));
    launch_bin(g, n, times[0], st);
    if (g->nf) dft_add(g, g->cE, g->cH, 1, g->nbuf, n, (long)llround(times[0]) - 1, st);
  } else if (g->nf) {
    // DFT only: staged, the DFT runs over up to sCap steps at once
    const size_t col = (size_t)dft_stage(g, n, times[0], st) * g->ct.total;
    hipLaunchKernelGGL(k_ctr_combine, cgrid, dim3(256), 0, st, samples, g->ct, (long)g->ct.total, 1L, g->sE + col,
                       g->sH + col);
    HIPCHK(hipEventRecord(consumed, st));
  } else {
    HIPCHK(hipEventRecord(consumed, st));
  }
  HIPCHK(hipGetLastError());
  return FDTD_OK;
}

// the NTFF stream's queued work done (before the host or the null stream touches
// the accumulators, the DFT sums or the sample buffers)
static int ntff_sync(fdtd_grid *g)
{
  if (g->ns) HIPCHK(hipStreamSynchronize(g->ns));
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_create(const fdtd_grid_desc *d, fdtd_grid **out)
{
  if (!d || !out) return fdtd_fail(FDTD_EINVAL, "null argument");
  if (d->N_PX < 4 || d->N_PY < 4) return fdtd_fail(FDTD_EINVAL, "grid %dx%d too small", d->N_PX, d->N_PY);
  if (d->precision != FDTD_FP64 && d->precision != FDTD_FP32) return fdtd_fail(FDTD_EINVAL, "bad precision %d", d->precision);
  if (d->mode != FDTD_MODE_TM && d->mode != FDTD_MODE_TE) return fdtd_fail(FDTD_EINVAL, "bad mode %d", d->mode);
  if (d->flavour != FDTD_FLAVOUR_MPI && d->flavour != FDTD_FLAVOUR_PLAIN && d->flavour != FDTD_FLAVOUR_SPLIT &&
      d->flavour != FDTD_FLAVOUR_NS)
    return fdtd_fail(FDTD_EINVAL, "bad flavour");
  if (d->flavour == FDTD_FLAVOUR_SPLIT && (d->precision != FDTD_FP64 || d->ntff_time))
    return fdtd_fail(FDTD_EINVAL, "the split-field PML solvers run in FP64, without the per-step NTFF");
  if (d->flavour == FDTD_FLAVOUR_NS && (d->precision != FDTD_FP64 || d->ntff_time))
    return fdtd_fail(FDTD_EINVAL, "the NS-FDTD solvers run in FP64, without the per-step NTFF");
  if (d->ntff_left < 1 || d->ntff_bottom < 1 || d->ntff_right >= d->N_PX || d->ntff_top >= d->N_PY ||
      d->ntff_left >= d->ntff_right || d->ntff_bottom >= d->ntff_top)
    return fdtd_fail(FDTD_EINVAL, "NTFF contour (%d..%d x %d..%d) outside the %dx%d grid", d->ntff_left, d->ntff_right,
                     d->ntff_bottom, d->ntff_top, d->N_PX, d->N_PY);
  if (d->array_size < d->max_time || d->max_time < 0) return fdtd_fail(FDTD_EINVAL, "bad NTFF array size");
  if (fdtd_hip_device_count() < 1) return fdtd_fail(FDTD_EHIP, "no HIP device");
  fdtd_grid *g = new fdtd_grid();
  memset((void *)g, 0, sizeof(*g));
  g->d = *d;
  g->off = d->flavour == FDTD_FLAVOUR_MPI ? 1 : 0;
  g->overlap = -1;
  g->NPXa = d->N_PX + 2 * g->off;
  g->NPYa = d->N_PY + 2 * g->off;
  g->ncell = (size_t)g->NPXa * g->NPYa;
  *out = g;
  // the FP64 state; an FP32 grid uses these only as export buffers, made on first use
  for (int f = 0; f < 9 && d->precision == FDTD_FP64; f++) {
    HIPCHK(dev_malloc(&g->F[f], g->ncell * sizeof(double2)));
    HIPCHK(hipMemset(g->F[f], 0, g->ncell * sizeof(double2)));
  }
  HIPCHK(dev_malloc(&g->Fdev, 9 * sizeof(double2 *)));
  HIPCHK(hipMemcpy(g->Fdev, g->F, 9 * sizeof(double2 *), hipMemcpyHostToDevice));
  for (int e = 0; e < 3; e++) {
    HIPCHK(dev_malloc(&g->eps[e], g->ncell * sizeof(double)));
    HIPCHK(hipMemset(g->eps[e], 0, g->ncell * sizeof(double)));
  }
  if (d->flavour == FDTD_FLAVOUR_NS) {
    g->nns = d->mode == FDTD_MODE_TM ? 5 : 10;
    HIPCHK(dev_malloc(&g->nsc, g->nns * g->ncell * sizeof(double)));
    HIPCHK(hipMemset(g->nsc, 0, g->nns * g->ncell * sizeof(double)));
  }
  const size_t ntab = 6 * (size_t)d->N_PX + 5 * (size_t)d->N_PY;
  HIPCHK(dev_malloc(&g->tab, ntab * sizeof(double)));
  HIPCHK(hipMemset(g->tab, 0, ntab * sizeof(double)));
  for (int q = 0; q < 5; q++) g->tabs.x[q] = g->tab + q * (size_t)d->N_PX;
  g->tabs.sx = g->tab + 5 * (size_t)d->N_PX;
  for (int q = 0; q < 4; q++) g->tabs.y[q] = g->tab + 6 * (size_t)d->N_PX + q * (size_t)d->N_PY;
  g->tabs.sy = g->tab + 6 * (size_t)d->N_PX + 4 * (size_t)d->N_PY;
  // NTFF
  build_contour(g);
  // time-domain accumulators only when the per-step NTFF runs (3 x 360 x arraySize
  // complex: 240 MB for a 10000-step run of a 2080-cell grid)
  g->ntn = d->ntff_time ? (size_t)360 * d->array_size + 2 * NT_GUARD : 0;
  for (int a = 0; a < 3 && g->ntn; a++) {
    HIPCHK(dev_malloc(&g->ntb[a], g->ntn * sizeof(double2)));
    HIPCHK(hipMemset(g->ntb[a], 0, g->ntn * sizeof(double2)));
    g->nt[a] = g->ntb[a] + NT_GUARD;
  }
  std::vector<double> tab, mn, mx, ca, sa, cb, sb;
  build_time_shift(g, tab, mn, mx, ca, sa, cb, sb, d->ntff_time != 0);
  const int P = std::max(1, g->ct.total);
  // FP32 grids bin a whole block of steps at once: sample buffers for FAST_MAX_T steps
  g->nbuf = d->precision == FDTD_FP32 ? FAST_MAX_T : 1;
  // the sample buffer: E samples of every cell, then H, nbuf steps per cell
  // (the binning's records address both from cE)
  HIPCHK(dev_malloc(&g->cE, (size_t)2 * g->nbuf * P * sizeof(double2)));
  g->cH = g->cE + (size_t)g->nbuf * P;
  if (d->ntff_time && g->ct.total > 0) {
    int rc = build_bin_groups(g, tab);
    if (rc) return rc;
  }
  HIPCHK(dev_malloc(&g->cosA, 360 * sizeof(double)));
  HIPCHK(dev_malloc(&g->sinA, 360 * sizeof(double)));
  HIPCHK(hipMemcpy(g->cosA, ca.data(), 360 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g->sinA, sa.data(), 360 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(dev_malloc(&g->cosB, 360 * sizeof(double)));
  HIPCHK(dev_malloc(&g->sinB, 360 * sizeof(double)));
  HIPCHK(hipMemcpy(g->cosB, cb.data(), 360 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g->sinB, sb.data(), 360 * sizeof(double), hipMemcpyHostToDevice));
  // 1/(4 pi c R), R = 1e6, applied per step by the MPI solvers (mpiTM_UPML.c:852-854)
  g->coef = 1.0 / (4 * M_PI * d->c0 * 1.0e6);
  HIPCHK(dev_malloc(&g->ffout, 360 * sizeof(double2)));
  HIPCHK(dev_malloc(&g->nfcnt, 3 * sizeof(unsigned long long)));
  HIPCHK(dev_malloc(&g->spec, 321 * 360 * sizeof(double)));
  HIPCHK(dev_malloc(&g->tw, 4096 * sizeof(double2)));
  HIPCHK(dev_malloc(&g->mstate, sizeof(fdtd_model_state)));
  if (d->precision == FDTD_FP32) {
    int rc = fast_state_create(*d, &g->fs);
    if (rc) return rc;
    // the per-step time-domain NTFF: the fused kernel samples the contour after
    // every inner step, and each block's samples are binned step by step
    if (d->ntff_time && g->ct.total > 0) {
      if ((rc = fast_state_set_contour(g->fs, g->ct, g->off, grid_bin_samples, g))) return rc;
      g->ctrSet = 1;
    }
  }
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_set_blocking(fdtd_grid *g, int T, int R)
{
  if (!g || !g->fs) return fdtd_fail(FDTD_EINVAL, "temporal blocking applies to FP32 grids only");
  return fast_state_set_blocking(g->fs, T, R);
}

extern "C" int fdtd_hip_grid_get_blocking(const fdtd_grid *g, int *T, int *R)
{
  if (!g || !g->fs) return fdtd_fail(FDTD_EINVAL, "temporal blocking applies to FP32 grids only");
  fast_state_get_blocking(g->fs, T, R);
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_set_ntff_overlap(fdtd_grid *g, int on)
{
  if (!g || on < -1 || on > 1) return fdtd_fail(FDTD_EINVAL, "bad argument");
  int rc;
  if ((g->fs && (rc = fast_state_flush(g->fs))) || (rc = ntff_sync(g))) return rc;
  g->overlap = on;
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_set_launch_mode(fdtd_grid *g, int mode)
{
  if (!g || !g->fs) return fdtd_fail(FDTD_EINVAL, "launch modes apply to FP32 grids only");
  return fast_state_set_launch_mode(g->fs, mode);
}

extern "C" int fdtd_hip_grid_get_launch_mode(const fdtd_grid *g, int *mode)
{
  if (!g || !g->fs || !mode) return fdtd_fail(FDTD_EINVAL, "launch modes apply to FP32 grids only");
  *mode = fast_state_get_launch_mode(g->fs);
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_flush(fdtd_grid *g)
{
  if (!g) return fdtd_fail(FDTD_EINVAL, "null argument");
  if (g->fs) return fast_state_flush(g->fs);
  return FDTD_OK;
}

extern "C" void fdtd_hip_profile_enable(int on) { fast_profile_enable(on); }

extern "C" void fdtd_hip_profile_read(double out[6]) { fast_profile_read6(out); }

extern "C" int fdtd_hip_grid_destroy(fdtd_grid *g)
{
  if (!g) return FDTD_OK;
  (void)hipDeviceSynchronize();
  if (g->ns) {
    (void)hipStreamDestroy(g->ns);
    (void)hipEventDestroy(g->ev_k);
    g->ns = nullptr;
  }
  for (int f = 0; f < 9; f++) (void)dev_free(g->F[f]);
  (void)dev_free(g->Fdev);
  for (int e = 0; e < 3; e++) (void)dev_free(g->eps[e]);
  (void)dev_free(g->nsc);
  (void)dev_free(g->tab);
  for (int a = 0; a < 3; a++) (void)dev_free(g->ntb[a]);
  (void)dev_free(g->cE);  // cH lies in the same block
  (void)dev_free(g->binPart);
  (void)dev_free(g->binRec);
  (void)dev_free(g->dftOmega);
  (void)dev_free(g->dftE);
  (void)dev_free(g->dftH);
  (void)dev_free(g->dftTs);
  (void)dev_free(g->dftOut);
  (void)dev_free(g->sE);
  (void)dev_free(g->sH);
  (void)dev_free(g->dftPh);
  (void)dev_free(g->cosA);
  (void)dev_free(g->sinA);
  (void)dev_free(g->cosB);
  (void)dev_free(g->sinB);
  (void)dev_free(g->ffout);
  (void)dev_free(g->nfcnt);
  (void)dev_free(g->spec);
  (void)dev_free(g->tw);
  (void)dev_free(g->mstate);
  fast_state_destroy(g->fs);
  delete g;
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_set_coefs(fdtd_grid *g, const fdtd_coef_tables *t)
{
  if (!g || !t) return fdtd_fail(FDTD_EINVAL, "null argument");
  const size_t nx = g->d.N_PX, ny = g->d.N_PY;
  for (int q = 0; q < 5; q++) {
    if (!t->x[q]) return fdtd_fail(FDTD_EINVAL, "missing x table %d", q);
    HIPCHK(hipMemcpy((void *)g->tabs.x[q], t->x[q], nx * sizeof(double), hipMemcpyHostToDevice));
  }
  for (int q = 0; q < 4; q++) {
    if (!t->y[q]) return fdtd_fail(FDTD_EINVAL, "missing y table %d", q);
    HIPCHK(hipMemcpy((void *)g->tabs.y[q], t->y[q], ny * sizeof(double), hipMemcpyHostToDevice));
  }
  if (!t->sx || !t->sy) return fdtd_fail(FDTD_EINVAL, "missing sigma tables");
  HIPCHK(hipMemcpy((void *)g->tabs.sx, t->sx, nx * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)g->tabs.sy, t->sy, ny * sizeof(double), hipMemcpyHostToDevice));
  if (g->fs) return fast_state_set_coefs(g->fs, t);
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_set_ns_coefs(fdtd_grid *g, const double *cells, double r_half)
{
  if (!g || !cells) return fdtd_fail(FDTD_EINVAL, "null argument");
  if (g->d.flavour != FDTD_FLAVOUR_NS) return fdtd_fail(FDTD_EINVAL, "per-cell NS coefficients need the NS flavour");
  HIPCHK(hipMemcpy(g->nsc, cells, g->nns * g->ncell * sizeof(double), hipMemcpyHostToDevice));
  g->ns_r2 = r_half;
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_fill_eps(fdtd_grid *g, const fdtd_model_state *m)
{
  if (!g || !m) return fdtd_fail(FDTD_EINVAL, "null argument");
  HIPCHK(hipMemcpy(g->mstate, m, sizeof(*m), hipMemcpyHostToDevice));
  dim3 blk(64, 4), grd((g->NPYa + 63) / 64, (g->NPXa + 3) / 4);
  hipLaunchKernelGGL(k_fill_eps, grd, blk, 0, 0, g->mstate, g->eps[0], g->eps[1], g->eps[2], g->NPXa, g->NPYa, g->off,
                     g->d.mode == FDTD_MODE_TM ? 1 : 0,
                     g->d.flavour == FDTD_FLAVOUR_PLAIN ? 1
                     : (g->d.flavour == FDTD_FLAVOUR_NS && g->d.mode == FDTD_MODE_TE ? 2 : 0));
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  if (g->fs) return fast_state_set_inv_eps(g->fs, g->eps, g->NPYa, g->off);
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_set_eps(fdtd_grid *g, int which, const double *host)
{
  if (!g || !host || which < 0 || which > 2) return fdtd_fail(FDTD_EINVAL, "bad argument");
  HIPCHK(hipMemcpy(g->eps[which], host, g->ncell * sizeof(double), hipMemcpyHostToDevice));
  if (g->fs) return fast_state_set_inv_eps(g->fs, g->eps, g->NPYa, g->off);
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_get_eps(fdtd_grid *g, int which, double *host)
{
  if (!g || !host || which < 0 || which > 2) return fdtd_fail(FDTD_EINVAL, "bad argument");
  HIPCHK(hipMemcpy(host, g->eps[which], g->ncell * sizeof(double), hipMemcpyDeviceToHost));
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_step(fdtd_grid *g, const fdtd_step_params *p)
{
  if (!g || !p) return fdtd_fail(FDTD_EINVAL, "null argument");
  const fdtd_grid_desc &d = g->d;
  if (g->fs) {
    int launched = 0;
    return fast_state_step(g->fs, p, &launched);
  }
  GridPtrs gp;
  for (int f = 0; f < 9; f++) gp.F[f] = g->F[f];
  for (int e = 0; e < 3; e++) gp.eps[e] = g->eps[e];
  gp.NPXa = g->NPXa;
  gp.NPYa = g->NPYa;
  gp.off = g->off;
  gp.mu0 = d.mu0;
  StepArgs s;
  memset(&s, 0, sizeof s);
  const bool plain = d.flavour == FDTD_FLAVOUR_PLAIN;
  const double rad = p->angle_deg * M_PI / 180.0;
  s.w_t = p->omega * p->time;
  s.ray_coef = p->ray_coef;
  s.ks_cos = cos(rad) * p->k;
  s.ks_sin = sin(rad) * p->k;
  s.ps_on = d.point_source && d.mode == FDTD_MODE_TM && !plain;
  s.ps_i = d.N_PX / 2 + g->off;
  s.ps_j = d.N_PY / 2 + g->off;
  s.ps_re = p->ps_re;
  s.ps_im = p->ps_im;
  if (plain) {
    // field_scatteredPulse constants (field.c:226-241)
    s.time = p->time;
    s.w_s = p->omega;
    s.cpc = cos(rad) / d.c0;
    s.spc = sin(rad) / d.c0;
    const double gx[2] = {d.mode == FDTD_MODE_TM ? 0.0 : 0.5, 0.0};
    const double gy[2] = {0.0, d.mode == FDTD_MODE_TM ? 0.0 : 0.5};
    for (int c = 0; c < 2; c++) {
      s.gx[c] = gx[c];
      s.gy[c] = gy[c];
      const double center_peak = (d.N_PX / 2.0 + gx[c]) * s.cpc + (d.N_PY / 2 + gy[c]) * s.spc;
      s.t0[c] = -center_peak + 500;
    }
    if (d.mode == FDTD_MODE_TM) {
      s.dot[0] = 1.0;
      s.on[0] = 1;
    } else {
      s.dot[0] = p->dot_x;
      s.dot[1] = p->dot_y;
      s.on[0] = p->use_x;
      s.on[1] = p->use_y;
    }
  }
  dim3 blk(64, 4), grd((g->NPYa - 2 + 63) / 64, (g->NPXa - 2 + 3) / 4);
  if (d.flavour == FDTD_FLAVOUR_SPLIT) {
    s.ps_on = 0;
    s.time = p->time;
    s.w_s = p->omega;
    if (d.mode == FDTD_MODE_TM) {  // update (fdtdTM.c:283-288): calcH, calcE, scattered wave
      hipLaunchKernelGGL(k_split_tm_h, grd, blk, 0, 0, gp, g->tabs);
      hipLaunchKernelGGL(k_split_tm_e, grd, blk, 0, 0, gp, g->tabs, s);
    } else {  // update (fdtdTE.c:283-288): calcE, scattered wave, calcH
      hipLaunchKernelGGL(k_split_te_e, grd, blk, 0, 0, gp, g->tabs, s);
      hipLaunchKernelGGL(k_split_te_h, grd, blk, 0, 0, gp, g->tabs);
    }
  } else if (d.flavour == FDTD_FLAVOUR_NS) {
    s.ps_on = 0;
    s.time = p->time;
    s.w_s = p->omega;
    if (d.mode == FDTD_MODE_TM) {  // update (nsFdtdTM.c:68-80): calcH, calcE, NS wave, Ez
      hipLaunchKernelGGL(k_ns_tm_h, grd, blk, 0, 0, gp, g->tabs, (const double *)g->nsc, g->ns_r2);
      hipLaunchKernelGGL(k_ns_tm_e, grd, blk, 0, 0, gp, g->tabs, (const double *)g->nsc, s);
    } else {  // update (nsFdtdTE.c:242-257): calcH, Hz, calcE, NS waves
      s.dot[0] = p->dot_x;
      s.dot[1] = p->dot_y;
      s.on[0] = p->use_x;
      s.on[1] = p->use_y;
      hipLaunchKernelGGL(k_ns_te_h, grd, blk, 0, 0, gp, (const double *)g->nsc);
      hipLaunchKernelGGL(k_ns_te_e, grd, blk, 0, 0, gp, (const double *)g->nsc, g->ns_r2, s);
    }
  } else if (d.mode == FDTD_MODE_TM) {
    if (plain) {
      hipLaunchKernelGGL(k_tm_h, grd, blk, 0, 0, gp, g->tabs);
      hipLaunchKernelGGL(k_tm_e<1>, grd, blk, 0, 0, gp, g->tabs, s);
    } else {
      hipLaunchKernelGGL(k_tm_e<0>, grd, blk, 0, 0, gp, g->tabs, s);
      hipLaunchKernelGGL(k_tm_h, grd, blk, 0, 0, gp, g->tabs);
    }
  } else {
    if (plain) {
      hipLaunchKernelGGL(k_te_h, grd, blk, 0, 0, gp, g->tabs);
      hipLaunchKernelGGL(k_te_e<1>, grd, blk, 0, 0, gp, g->tabs, s);
    } else {
      hipLaunchKernelGGL(k_te_e<0>, grd, blk, 0, 0, gp, g->tabs, s);
      hipLaunchKernelGGL(k_te_h, grd, blk, 0, 0, gp, g->tabs);
    }
  }
  HIPCHK(hipGetLastError());
  if (g->ntn && g->ct.total > 0) {
    hipLaunchKernelGGL(k_ntff_gather, dim3((g->ct.total + 255) / 256), dim3(256), 0, 0, (const double2 *const *)g->Fdev,
                       g->ct, g->cE, g->cH);
    launch_bin(g, 1, p->time, 0);
    if (g->nf) dft_add(g, g->cE, g->cH, 1, 1, 1, (long)llround(p->time) - 1, 0);
    HIPCHK(hipGetLastError());
  } else if (g->nf && g->ct.total > 0) {
    const size_t col = (size_t)dft_stage(g, 1, p->time, 0) * g->ct.total;
    hipLaunchKernelGGL(k_ntff_gather, dim3((g->ct.total + 255) / 256), dim3(256), 0, 0, (const double2 *const *)g->Fdev,
                       g->ct, g->sE + col, g->sH + col);
    HIPCHK(hipGetLastError());
  }
  return FDTD_OK;
}

// an FP32 grid's FP64 export buffer for slot f, zeroed when first made (the
// ghost ring and the cells the export does not write stay zero)
static int ensure_F(fdtd_grid *g, int f)
{
  if (g->F[f]) return FDTD_OK;
  HIPCHK(dev_malloc(&g->F[f], g->ncell * sizeof(double2)));
  HIPCHK(hipMemset(g->F[f], 0, g->ncell * sizeof(double2)));
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_get_field(fdtd_grid *g, int which, double *host)
{
  if (!g || !host || which < 0 || which > 8) return fdtd_fail(FDTD_EINVAL, "bad argument");
  int rc0 = ensure_F(g, which);
  if (rc0) return rc0;
  if (g->fs) {
    int rc = fast_state_export(g->fs, which, g->F[which], g->NPYa, g->off);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpy(host, g->F[which], g->ncell * sizeof(double2), hipMemcpyDeviceToHost));
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_set_field(fdtd_grid *g, int which, const double *host)
{
  if (!g || !host || which < 0 || which > 8) return fdtd_fail(FDTD_EINVAL, "bad argument");
  int rc0 = ensure_F(g, which);
  if (rc0) return rc0;
  HIPCHK(hipMemcpy(g->F[which], host, g->ncell * sizeof(double2), hipMemcpyHostToDevice));
  if (g->fs) return fast_state_import(g->fs, which, g->F[which], g->NPYa, g->off);
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_get_ntff_time(fdtd_grid *g, int which, double *host)
{
  if (!g || !host || which < 0 || which > 2) return fdtd_fail(FDTD_EINVAL, "bad argument");
  if (!g->ntn) return fdtd_fail(FDTD_ESTATE, "the grid was created without the per-step time-domain NTFF");
  int rc;
  if (g->fs && (rc = fast_state_flush(g->fs))) return rc;  // bins the queued steps
  if ((rc = ntff_sync(g))) return rc;
  HIPCHK(hipMemcpy(host, g->nt[which], (size_t)360 * g->d.array_size * sizeof(double2), hipMemcpyDeviceToHost));
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_reset(fdtd_grid *g)
{
  if (!g) return fdtd_fail(FDTD_EINVAL, "null argument");
  int rc0 = ntff_sync(g);
  if (rc0) return rc0;
  for (int f = 0; f < 9; f++)
    if (g->F[f]) HIPCHK(hipMemset(g->F[f], 0, g->ncell * sizeof(double2)));
  for (int a = 0; a < 3 && g->ntn; a++) HIPCHK(hipMemset(g->ntb[a], 0, g->ntn * sizeof(double2)));
  if (g->nf) {
    HIPCHK(hipMemset(g->dftE, 0, (size_t)g->nf * g->ct.total * sizeof(double2)));
    HIPCHK(hipMemset(g->dftH, 0, (size_t)g->nf * g->ct.total * sizeof(double2)));
  }
  g->sFill = 0;
  HIPCHK(hipDeviceSynchronize());
  if (g->fs) return fast_state_reset(g->fs);
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_far_field(fdtd_grid *g, double k, double coef_re, double coef_im, double scale, double *out360)
{
  if (!g || !out360) return fdtd_fail(FDTD_EINVAL, "null argument");
  const fdtd_grid_desc &d = g->d;
  if (g->fs) {  // the far field of the FP32 fields is formed in FP64
    for (int slot : {0, 3, 6}) {
      int rc = ensure_F(g, slot);
      if (rc || (rc = fast_state_export(g->fs, slot, g->F[slot], g->NPYa, g->off))) return rc;
    }
  }
  FFArgs a;
  a.N_PX = d.N_PX; a.N_PY = d.N_PY; a.NPXa = g->NPXa; a.NPYa = g->NPYa; a.off = g->off;
  a.top = d.ntff_top; a.bottom = d.ntff_bottom; a.left = d.ntff_left; a.right = d.ntff_right;
  a.cx = d.ntff_cx; a.cy = d.ntff_cy; a.k = k; a.z0 = d.z0;
  a.coef_re = coef_re; a.coef_im = coef_im; a.scale = scale;
  a.cosA = g->cosA; a.sinA = g->sinA;
  if (d.mode == FDTD_MODE_TM)
    hipLaunchKernelGGL(k_far<FDTD_MODE_TM>, dim3(360), dim3(256), 0, 0, g->F[0], g->F[3], g->F[6], a, g->ffout);
  else {
    a.cx = d.N_PX / 2; a.cy = d.N_PY / 2;
    hipLaunchKernelGGL(k_far<FDTD_MODE_TE>, dim3(360), dim3(256), 0, 0, g->F[0], g->F[3], g->F[6], a, g->ffout);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out360, g->ffout, 360 * sizeof(double2), hipMemcpyDeviceToHost));
  return FDTD_OK;
}

// Device address of `scratch` on each GPU, looked up on the first call there
// (the array's name is a host-side stand-in that kernels cannot use).
static std::mutex g_scratch_mu;
static std::map<int, double2 *> g_scratch_ptr;

static int scratch_ptr(double2 **out)
{
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    it = g_scratch_ptr.emplace(dev, (double2 *)p).first;
  }
  *out = it->second;
  return FDTD_OK;
}

// fdtd_hip_grid_far_field without any allocation (see the scratch array above):
// FDTD_ESTATE while FP32 steps are queued (launching them may build tiles),
// FDTD_EINVAL for a contour longer than FF_NA_PARTS * FF_NA_CHUNK cells.
extern "C" __attribute__((visibility("hidden"))) int fdtd_hip_grid_far_field_noalloc(fdtd_grid *g, double k, double coef_re, double coef_im,
                                                         double scale, double *out360)
{
  if (!g || !out360) return fdtd_fail(FDTD_EINVAL, "null argument");
  const fdtd_grid_desc &d = g->d;
  FFArgs a;
  a.N_PX = d.N_PX; a.N_PY = d.N_PY; a.NPXa = g->NPXa; a.NPYa = g->NPYa; a.off = g->off;
  a.top = d.ntff_top; a.bottom = d.ntff_bottom; a.left = d.ntff_left; a.right = d.ntff_right;
  a.cx = d.ntff_cx; a.cy = d.ntff_cy; a.k = k; a.z0 = d.z0;
  a.coef_re = coef_re; a.coef_im = coef_im; a.scale = scale;
  a.cosA = g->cosA; a.sinA = g->sinA;
  if (d.mode == FDTD_MODE_TE) { a.cx = d.N_PX / 2; a.cy = d.N_PY / 2; }
  const int n = far_cells(d.mode, a);
  const int nparts = (n + FF_NA_CHUNK - 1) / FF_NA_CHUNK;
  if (nparts > FF_NA_PARTS)
    return fdtd_fail(FDTD_EINVAL, "allocation-free far field: contour of %d cells exceeds the limit of %d", n,
                     FF_NA_PARTS * FF_NA_CHUNK);
  double2 *scr;
  int rc;
  if ((rc = scratch_ptr(&scr))) return rc;
  double2 *part = scr, *res = scr + (size_t)360 * FF_NA_PARTS * 3;
  const dim3 grid(std::max(1, nparts), 360);
  if (g->fs) {
    if (fast_state_pending(g->fs))
      return fdtd_fail(FDTD_ESTATE, "allocation-free far field: FP32 steps are queued; launch them first "
                                    "(fdtd_hip_grid_flush)");
    const float2 *f[3];
    size_t nc = 0;
    int pitch = 0;
    if ((rc = fast_state_fields(g->fs, f, &nc)) || (rc = fast_state_pitch(g->fs, &pitch))) return rc;
    const FieldView<float2> v0 = {f[0], d.N_PX, d.N_PY, 0, pitch}, v1 = {f[1], d.N_PX, d.N_PY, 0, pitch},
                            v2 = {f[2], d.N_PX, d.N_PY, 0, pitch};
    if (d.mode == FDTD_MODE_TM) hipLaunchKernelGGL((k_far_part<FDTD_MODE_TM, float2>), grid, dim3(256), 0, 0, v0, v1, v2, a, n, part);
    else hipLaunchKernelGGL((k_far_part<FDTD_MODE_TE, float2>), grid, dim3(256), 0, 0, v0, v1, v2, a, n, part);
  } else {
    const FieldView<double2> v0 = {g->F[0], g->NPXa, g->NPYa, g->off, g->NPYa},
                             v1 = {g->F[3], g->NPXa, g->NPYa, g->off, g->NPYa},
                             v2 = {g->F[6], g->NPXa, g->NPYa, g->off, g->NPYa};
    if (d.mode == FDTD_MODE_TM) hipLaunchKernelGGL((k_far_part<FDTD_MODE_TM, double2>), grid, dim3(256), 0, 0, v0, v1, v2, a, n, part);
    else hipLaunchKernelGGL((k_far_part<FDTD_MODE_TE, double2>), grid, dim3(256), 0, 0, v0, v1, v2, a, n, part);
  }
  if (d.mode == FDTD_MODE_TM) hipLaunchKernelGGL(k_far_join<FDTD_MODE_TM>, dim3(6), dim3(64), 0, 0, part, nparts, a, res);
  else hipLaunchKernelGGL(k_far_join<FDTD_MODE_TE>, dim3(6), dim3(64), 0, 0, part, nparts, a, res);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out360, res, 360 * sizeof(double2), hipMemcpyDeviceToHost));
  return FDTD_OK;
}

// Non-finite values (NaN or +-Inf in the real or the imaginary part) of a field
// array: each thread counts over a grid-stride loop and adds its count to *cnt
// only when it found one, so a healthy field costs one read of the array.
template <class V>
__global__ __launch_bounds__(256) void k_count_nonfinite(const V *a, size_t n, unsigned long long *cnt)
{
  unsigned long long c = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const V v = a[i];
    c += (isfinite(v.x) && isfinite(v.y)) ? 0 : 1;
  }
  if (c) atomicAdd(cnt, c);
}

extern "C" int fdtd_hip_grid_count_nonfinite(fdtd_grid *g, long long out[3])
{
  if (!g || !out) return fdtd_fail(FDTD_EINVAL, "null argument");
  HIPCHK(hipMemset(g->nfcnt, 0, 3 * sizeof(unsigned long long)));
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  auto blocks = [&](size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((size_t)cus * 8, (n + 255) / 256)); };
  if (g->fs) {
    const float2 *f[3];
    size_t n = 0;
    int rc = fast_state_fields(g->fs, f, &n);
    if (rc) return rc;
    for (int q = 0; q < 3; q++)
      hipLaunchKernelGGL(k_count_nonfinite<float2>, dim3(blocks(n)), dim3(256), 0, 0, f[q], n, g->nfcnt + q);
  } else {
    for (int q = 0; q < 3; q++)
      hipLaunchKernelGGL(k_count_nonfinite<double2>, dim3(blocks(g->ncell)), dim3(256), 0, 0, g->F[3 * q], g->ncell,
                         g->nfcnt + q);
  }
  HIPCHK(hipGetLastError());
  unsigned long long h[3];
  HIPCHK(hipMemcpy(h, g->nfcnt, sizeof h, hipMemcpyDeviceToHost));
  for (int q = 0; q < 3; q++) out[q] = (long long)h[q];
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_time_spectra(fdtd_grid *g, double coef_re, double coef_im, const double *twiddles, int h_u_nm,
                                          double *out)
{
  if (!g || !twiddles || !out) return fdtd_fail(FDTD_EINVAL, "null argument");
  if (!g->ntn) return fdtd_fail(FDTD_ESTATE, "the grid was created without the per-step time-domain NTFF");
  int rc;
  if (g->fs && (rc = fast_state_flush(g->fs))) return rc;  // bins the queued steps
  if ((rc = ntff_sync(g))) return rc;
  const fdtd_grid_desc &d = g->d;
  HIPCHK(hipMemcpy(g->tw, twiddles, 4096 * sizeof(double2), hipMemcpyHostToDevice));
  const size_t lds = NTFF_NUM * sizeof(double2);
  const bool tmPlain = d.mode == FDTD_MODE_TM && d.flavour == FDTD_FLAVOUR_PLAIN;
  HIPCHK(hipFuncSetAttribute((const void *)k_time_spectra, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_time_spectra, dim3(360), dim3(1024), lds, 0, g->nt[0], g->nt[1], g->nt[2], d.array_size, d.max_time,
                     d.mode == FDTD_MODE_TM ? 1 : 0, d.z0, coef_re, coef_im, tmPlain ? g->cosB : g->cosA,
                     tmPlain ? g->sinB : g->sinA, g->tw, d.c0 * h_u_nm, g->spec);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out, g->spec, 321 * 360 * sizeof(double), hipMemcpyDeviceToHost));
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_set_dft(fdtd_grid *g, const double *omega, int nf)
{
  if (!g || (nf > 0 && !omega) || nf < 0) return fdtd_fail(FDTD_EINVAL, "bad argument");
  if (nf > 0 && g->ct.total <= 0) return fdtd_fail(FDTD_ESTATE, "the grid has no NTFF contour");
  int rc;
  if (g->fs && (rc = fast_state_flush(g->fs))) return rc;
  HIPCHK(hipDeviceSynchronize());
  (void)dev_free(g->dftOmega);
  (void)dev_free(g->dftE);
  (void)dev_free(g->dftH);
  (void)dev_free(g->dftTs);
  (void)dev_free(g->dftOut);
  (void)dev_free(g->sE);
  (void)dev_free(g->sH);
  (void)dev_free(g->dftPh);
  g->dftOmega = nullptr;
  g->dftE = g->dftH = g->dftOut = nullptr;
  g->sE = g->sH = g->dftPh = nullptr;
  g->dftTs = nullptr;
  g->nf = 0;
  g->sFill = 0;
  if (nf == 0) return FDTD_OK;
  const int P = g->ct.total;
  // staging for DFT_STAGE_STEPS steps (at least one block of the FP32 path)
  g->sCap = std::max(DFT_STAGE_STEPS, g->nbuf);
  HIPCHK(dev_malloc(&g->sE, (size_t)g->sCap * P * sizeof(double2)));
  HIPCHK(dev_malloc(&g->sH, (size_t)g->sCap * P * sizeof(double2)));
  HIPCHK(dev_malloc(&g->dftPh, (size_t)g->sCap * nf * sizeof(double2)));
  HIPCHK(dev_malloc(&g->dftOmega, nf * sizeof(double)));
  HIPCHK(hipMemcpy(g->dftOmega, omega, nf * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(dev_malloc(&g->dftE, (size_t)nf * P * sizeof(double2)));
  HIPCHK(dev_malloc(&g->dftH, (size_t)nf * P * sizeof(double2)));
  HIPCHK(hipMemset(g->dftE, 0, (size_t)nf * P * sizeof(double2)));
  HIPCHK(hipMemset(g->dftH, 0, (size_t)nf * P * sizeof(double2)));
  HIPCHK(dev_malloc(&g->dftOut, (size_t)nf * 3 * 360 * sizeof(double2)));
  std::vector<double> tab, mn, mx, ca, sa, cb, sb;
  build_time_shift(g, tab, mn, mx, ca, sa, cb, sb);
  HIPCHK(dev_malloc(&g->dftTs, tab.size() * sizeof(double)));
  HIPCHK(hipMemcpy(g->dftTs, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  g->nf = nf;
  // the FP32 fused kernel samples the contour after every inner step
  if (g->fs && !g->ctrSet) {
    if ((rc = fast_state_set_contour(g->fs, g->ct, g->off, grid_bin_samples, g))) return rc;
    g->ctrSet = 1;
  }
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_get_dft(fdtd_grid *g, double *out)
{
  if (!g || !out) return fdtd_fail(FDTD_EINVAL, "null argument");
  if (!g->nf) return fdtd_fail(FDTD_ESTATE, "no running DFT on this grid (fdtd_hip_grid_set_dft)");
  int rc;
  if (g->fs && (rc = fast_state_flush(g->fs))) return rc;
  if ((rc = ntff_sync(g))) return rc;
  dft_flush(g, 0);
  const double cf = g->d.flavour == FDTD_FLAVOUR_PLAIN ? 1.0 : g->coef;
  hipLaunchKernelGGL(k_dft_far, dim3(360, g->nf), dim3(256), 0, 0, g->dftE, g->dftH, g->ct, g->dftTs, g->dftOmega, cf,
                     g->dftOut);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out, g->dftOut, (size_t)g->nf * 3 * 360 * sizeof(double2), hipMemcpyDeviceToHost));
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_dims(const fdtd_grid *g, int out[4])
{
  if (!g || !out) return fdtd_fail(FDTD_EINVAL, "null argument");
  out[0] = g->NPXa;
  out[1] = g->NPYa;
  out[2] = g->off;
  out[3] = g->d.array_size;
  return FDTD_OK;
}

extern "C" int fdtd_hip_grid_sync(fdtd_grid *g)
{
  (void)g;
  HIPCHK(hipDeviceSynchronize());
  return FDTD_OK;
}

// ---------------------------------------------------------------- batched instances

struct fdtd_hip_batch {
  FastBatch *fb;
  std::vector<fdtd_grid *> grids;
};

extern "C" int fdtd_hip_batch_create(fdtd_grid *const *grids, const fdtd_step_params *p0, const int *steps, int n,
                                     int T, int R, fdtd_hip_batch **out)
{
  if (!grids || !p0 || !steps || !out || n < 1) return fdtd_fail(FDTD_EINVAL, "bad argument");
  *out = nullptr;
  std::vector<FastState *> st(n);
  for (int i = 0; i < n; i++) {
    if (!grids[i] || !grids[i]->fs) return fdtd_fail(FDTD_EINVAL, "instance %d is not an FP32 grid", i);
    if (grids[i]->d.ntff_time) return fdtd_fail(FDTD_EINVAL, "instance %d runs the per-step NTFF (one step per launch)", i);
    st[i] = grids[i]->fs;
  }
  FastBatch *fb = nullptr;
  int rc = fast_batch_create(st.data(), p0, steps, n, T, R, &fb);
  if (rc) return rc;
  fdtd_hip_batch *b = new fdtd_hip_batch();
  b->fb = fb;
  b->grids.assign(grids, grids + n);
  *out = b;
  return FDTD_OK;
}

extern "C" int fdtd_hip_batch_run(fdtd_hip_batch *b, int blocks)
{
  if (!b) return fdtd_fail(FDTD_EINVAL, "null batch");
  int rc;
  if (blocks < 0) {
    if ((rc = fast_batch_sync(b->fb, nullptr))) return rc;
    if ((rc = fast_batch_run(b->fb, fast_batch_blocks_left(b->fb), 1))) return rc;
    return fast_batch_sync(b->fb, nullptr);
  }
  return fast_batch_run(b->fb, blocks, 1);
}

extern "C" int fdtd_hip_batch_set_launch_mode(fdtd_hip_batch *b, int mode)
{
  if (!b) return fdtd_fail(FDTD_EINVAL, "null batch");
  return fast_batch_set_launch(b->fb, mode);  // checks the mode
}

extern "C" int fdtd_hip_batch_sync(fdtd_hip_batch *b, int *steps_done)
{
  if (!b) return fdtd_fail(FDTD_EINVAL, "null batch");
  return fast_batch_sync(b->fb, steps_done);
}

extern "C" int fdtd_hip_batch_destroy(fdtd_hip_batch *b)
{
  if (!b) return FDTD_OK;
  int rc = fast_batch_sync(b->fb, nullptr);
  fast_batch_destroy(b->fb);
  delete b;
  return rc;
}
