// fdtd_fast.hip -- FP32 fused, temporally blocked UPML step for gfx950.
//
// One launch advances every tile of every instance by T time steps.  A
// workgroup (4 wavefronts) owns a 64-column x 4R-row block of cells: lane l
// holds column j0-(T+1)+l, wave w holds R consecutive rows, and each thread
// keeps its R cells' E and H in registers for all T steps.  Neighbours along
// j (the contiguous axis) come from the adjacent lane through cross-lane
// shuffles; neighbours along i are in the same thread's registers except at
// wave boundaries, which are exchanged through a 2 x 4 x 64-entry LDS row
// buffer.  The block reads its cells once, runs T E/H half-step pairs, and
// writes back the core cells [i0, i0+ni) x [j0, j0+nj) that stay valid after
// T steps (halo T+1 below, T above).  Inputs and outputs are ping-pong
// buffers, so the halos read by one tile are never written by another in the
// same launch.  This replaces the reference's per-step MPI halo exchange
// (mpiTM_UPML.c:252-296) with on-chip halos.
//
// Physics per reference update (mpiTM_UPML.c:196-217, mpiTE_UPML.c:290-304,
// fdtdTM_upml.c:54-66, fdtdTE_upml.c:168-192).  Where every UPML coefficient
// is 1 (the interior) the J/D and M/B recurrences collapse to
//   E_n = E_{n-1} + curl(H)/eps + (s_n - s_{n-1}),   H += c0^2 * curl(E)
// where E holds the total field the reference stores (D/eps plus the source
// s_n of that step); UPML cells keep the full recurrence with their
// auxiliaries.  All arithmetic is FP32.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <vector>

#include "fdtd_hip.h"
#include "fdtd_internal.h"

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// complex product with its roundings spelled out (one product of each part fused
// into its sum, the other rounded on its own), so that every kernel that inlines
// it rounds alike whatever the compiler's contraction choices elsewhere
__device__ __forceinline__ float2 c_mul(float2 a, float2 b)
{
  return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
// b + a s and b - a s per part, each a single rounding (fused)
__device__ __forceinline__ float2 c_fma(float2 a, float s, float2 b) { return make_float2(fmaf(s, a.x, b.x), fmaf(s, a.y, b.y)); }
__device__ __forceinline__ float2 c_fnma(float2 a, float s, float2 b) { return make_float2(fmaf(-s, a.x, b.x), fmaf(-s, a.y, b.y)); }
__device__ __forceinline__ float2 sel(bool c, float2 a, float2 b) { return make_float2(c ? a.x : b.x, c ? a.y : b.y); }
// Neighbours along j (the lane axis) through DPP wave shifts, a VALU operand
// modifier with no LDS round trip (the compiler folds the move into the
// subtraction that consumes it): wave_shr:1 gives lane l the value of lane
// l-1, wave_shl:1 that of lane l+1, and the lane without a source reads 0.
// Lane 0 and lane 63 hold halo columns: a garbage value entering there moves
// one lane per step, and the core starts T+1 lanes in, so after T steps no core
// cell has seen it.
__device__ __forceinline__ float2 shfl_up1(float2 v)
{
  return make_float2(__builtin_amdgcn_mov_dpp(v.x, 0x138, 0xf, 0xf, true),
                     __builtin_amdgcn_mov_dpp(v.y, 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float2 shfl_dn1(float2 v)
{
  return make_float2(__builtin_amdgcn_mov_dpp(v.x, 0x130, 0xf, 0xf, true),
                     __builtin_amdgcn_mov_dpp(v.y, 0x130, 0xf, 0xf, true));
}

// Wave-uniform values read once into scalar registers (every lane holds the same value).
__device__ __forceinline__ int uni_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni_f(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
// A device-memory pointer held in scalar registers and typed as global memory,
// so that loads and stores use the global instructions rather than flat ones
// (the host pass of hipcc has no address spaces and sees a plain pointer).
#ifdef __HIP_DEVICE_COMPILE__
#define GPTR(T) __attribute__((address_space(1))) T *
#else
#define GPTR(T) T *
#endif
// Element access at a 32-bit byte offset from a scalar base: the global_load /
// global_store "scalar base + vector offset" form, one offset register for all
// arrays of the same shape instead of a 64-bit address per array and row.
template <class T>
__device__ __forceinline__ T ldo(GPTR(const T) base, unsigned boff)
{
  return *(GPTR(const T))((GPTR(const char))base + boff);
}
template <class T>
__device__ __forceinline__ void sto(GPTR(T) base, unsigned boff, T v)
{
  *(GPTR(T))((GPTR(char))base + boff) = v;
}
template <class P>
__device__ __forceinline__ GPTR(P) uni_p(P *p)
{
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return (GPTR(P))(((unsigned long long)hi << 32) | lo);
}

// Cell index of (i, j) in the UPML auxiliaries, which store only the cells
// outside the box [axlo, axhi) x [aylo, ayhi) that holds no UPML cell: first the
// rows below and above the box (whole rows of `pitch` cells), then each row of
// the box's row range as its acw cells left and right of the box.  (i, j) must
// lie outside the box.  On the 8256 x 4160 grid with its 32-cell UPML this is
// 2.3 % of the full layout.
__host__ __device__ __forceinline__ unsigned aux_cell(int i, int j, int NPX, int pitch, int axlo, int axhi, int aylo,
                                                      int ayhi, int acw)
{
  if (i < axlo) return (unsigned)(i * pitch + j);
  if (i >= axhi) return (unsigned)((i - axhi + axlo) * pitch + j);
  return (unsigned)((axlo + NPX - axhi) * pitch + (i - axlo) * acw + (j < aylo ? j : j - ayhi + aylo));
}

// Constants of the plain flavour's Gaussian pulse, held in registers for a tile.
struct PulseC {
  float cpc, spc, w;
  float gx[2], gy[2], dot[2];
};

// Gaussian pulse of field_scatteredPulse (field.c:224-256) for delay tn = n - t0
__device__ __forceinline__ float2 pulse(const PulseC &pc, int c, int i, int j, float ie, float tn)
{
  const float r = (i + pc.gx[c]) * pc.cpc + (j + pc.gy[c]) * pc.spc - tn;
  const float q = r * (1.0f / 50.0f);
  const float a = pc.dot[c] * __expf(-q * q) * (ie - 1.0f);
  float sn, cs;
  __sincosf(r * pc.w, &sn, &cs);
  return make_float2(a * cs, a * sn);
}

// Every value the step loop needs beyond the cell state is staged once per
// tile: the step scalars and the per-row tables in LDS (rows are uniform across
// a wavefront, so each read is a broadcast), the per-column tables and the
// pulse constants in registers.  The loop then issues no global load: the
// kernel writes global memory, so loads from it cannot go through the scalar
// cache and would otherwise be dependent vector round trips on every step.
// step scalars of a batched instance at step t, formed on the device in FP64 as
// the single-instance host path forms them (field.c ray_coef ramp, e^{-i w t}):
// c_t = ray_coef(t) e^{-i w t}, dc = c_t - c_{t-1} with c_{-1} = 0
__device__ __forceinline__ void batch_c(double omega, int t, double &re, double &im)
{
  const double x = 0.01 * t;
  const double rc = 1.0 - exp(-(x * x));
  double sn, cs;
  sincos(omega * t, &sn, &cs);
  re = rc * cs;
  im = -rc * sn;
}
__device__ __forceinline__ FastStep batch_step(const FastBatchInst &B, int s)
{
  const int t = B.n + s;
  double re, im, pre = 0.0, pim = 0.0;
  batch_c(B.omega, t, re, im);
  if (t >= 1) batch_c(B.omega, t - 1, pre, pim);
  FastStep st;
  st.c = make_float2((float)re, (float)im);
  st.dc = make_float2((float)(re - pre), (float)(im - pim));
  st.tn = st.tn1 = st.tp = st.tp1 = 0.0f;
  st.hasprev = t >= 1 ? 1.0f : 0.0f;
  st.pad = 0.0f;
  return st;
}

// One tile: load the block, run its T steps, write back the core.  BATCH: the
// instance's depth, ping-pong index and step scalars come from its device
// progress record instead of the launch arguments.
template <int MODE, int FLAV, int R, int PML, int BATCH, int SMP, int W = 4>
__device__ __forceinline__ void fast_tile(const FastArgs &a, int tid)
{
  constexpr int NR = W * R;  // rows of the loaded block (W wavefronts of R rows)
  // Rows exchanged between wavefronts: wave w writes its first E row to xE[w]
  // and its last H row to xH[w + 1]; xE[W] and xH[0] stay zero (the block's
  // outer halo rows), so every read is unconditional and the step code has no
  // branch that would split it.
  __shared__ float2 xE[W + 1][64];
  __shared__ float2 xH[W + 1][64];
  __shared__ FastStep sst[FAST_MAX_T];
  __shared__ float2 sex[NR];  // e^{i kx x} of the block rows (MPI scattered wave)
  // UPML x tables of the block rows: tx0..tx4, 1/(2 + sigma_x)
  __shared__ float srow[6][PML ? NR : 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  {
    const FastTile t = a.tiles[tid];
    const FastInst &I = a.inst[t.inst];
    int T = a.T, cur = a.cur;
    if (BATCH) {
      T = uni_i(a.binst[t.inst].Ti);
      cur = uni_i(a.binst[t.inst].cur);
    }
    const int nxt = cur ^ 1;
    __syncthreads();  // the previous tile's last reads of the shared arrays are done
    if (BATCH && T == 0) return;  // instance finished (uniform)
    if (w == 0) {
      xE[W][lane] = make_float2(0.f, 0.f);
      xH[0][lane] = make_float2(0.f, 0.f);
    }
    // Instance scalars and pointers, read once into scalar registers.  The
    // readfirstlane is convergent, so the compiler cannot sink these loads into
    // the per-row branches below (where each row would wait on its own round trip).
    const int NPX = uni_i(I.NPX), NPY = uni_i(I.NPY), P = uni_i(I.pitch);
    GPTR(const float2) g0 = uni_p(I.f0[cur]);
    GPTR(const float2) g1 = uni_p(I.f1[cur]);
    GPTR(const float2) g2 = uni_p(I.f2[cur]);
    GPTR(const float) gie0 = uni_p(I.ie0);
    GPTR(const float) gie1 = uni_p(I.ie1);
    GPTR(const float2) ga[6];
    GPTR(const float2) gex = nullptr;
    GPTR(const float2) gey = nullptr;
    GPTR(const float) gtx[5];
    GPTR(const float) gty[4];
    GPTR(const float) gtsx = nullptr;
    GPTR(const float) gtsy = nullptr;
    GPTR(const unsigned char) gpx = nullptr;
    GPTR(const unsigned char) gpy = nullptr;
    if (FLAV == FDTD_FLAVOUR_MPI) {
      gex = uni_p(I.ex);
      gey = uni_p(I.ey);
    }
    if (PML) {
#pragma unroll
      for (int qq = 0; qq < 6; qq++) ga[qq] = uni_p(I.aux[qq][cur]);
#pragma unroll
      for (int qq = 0; qq < 5; qq++) gtx[qq] = uni_p(I.tx[qq]);
#pragma unroll
      for (int qq = 0; qq < 4; qq++) gty[qq] = uni_p(I.ty[qq]);
      gtsx = uni_p(I.tsx);
      gtsy = uni_p(I.tsy);
      gpx = uni_p(I.pmlx);
      gpy = uni_p(I.pmly);
    }
    const int jj = t.bj + lane;
    const int ib0 = t.bi;  // first row of the block
    const int ib = ib0 + w * R;
    const bool jin = jj >= 0 && jj < NPY;
    const int jc = min(max(jj, 0), NPY - 1);  // clamped column: every load below stays in bounds
    const float c2 = uni_f(I.c2);
    if ((int)threadIdx.x < T) sst[threadIdx.x] = BATCH ? a.bsteps[t.inst * FAST_MAX_T + threadIdx.x] : a.st[threadIdx.x];
    for (int r = threadIdx.x; r < NR; r += 64 * W) {
      const int i = ib0 + r;
      const bool ok = i >= 0 && i < NPX;
      const int ic = min(max(i, 0), NPX - 1);
      if (FLAV == FDTD_FLAVOUR_MPI) {
        const float2 e = gex[ic];
        sex[r] = ok ? e : make_float2(0.f, 0.f);
      }
      if (PML) {
        float v[5];
#pragma unroll
        for (int qq = 0; qq < 5; qq++) v[qq] = gtx[qq][ic];
        const float sx = gtsx[ic];
#pragma unroll
        for (int qq = 0; qq < 5; qq++) srow[qq][r] = ok ? v[qq] : 1.0f;
        srow[5][r] = 1.0f / (2.0f + (ok ? sx : 0.0f));
      }
    }
    float2 eyl = make_float2(0.f, 0.f);
    if (FLAV == FDTD_FLAVOUR_MPI) {
      const float2 e = gey[jc];
      if (jin) eyl = e;
    }
    PulseC pc;
    if (FLAV == FDTD_FLAVOUR_PLAIN) {
      pc.cpc = uni_f(I.cpc); pc.spc = uni_f(I.spc); pc.w = uni_f(I.w);
      pc.gx[0] = uni_f(I.gx[0]); pc.gx[1] = uni_f(I.gx[1]); pc.gy[0] = uni_f(I.gy[0]); pc.gy[1] = uni_f(I.gy[1]);
      pc.dot[0] = uni_f(I.dot[0]); pc.dot[1] = uni_f(I.dot[1]);
    }
    // per-lane (column) coefficient tables of the UPML cells
    float cy0 = 1.f, cy1 = 1.f, cy2 = 1.f, cy3 = 1.f, sy = 0.f;
    unsigned char pmly = 0;
    if (PML) {
      const float v0 = gty[0][jc], v1 = gty[1][jc], v2 = gty[2][jc], v3 = gty[3][jc], vs = gtsy[jc];
      const unsigned char fl = gpy[jc];
      if (jin) { cy0 = v0; cy1 = v1; cy2 = v2; cy3 = v3; sy = vs; pmly = fl; }
    }

    // the block's cells: every load is issued before any value is used
    float2 F0[R], F1[R], F2[R];
    float ie0[R], ie1[R];
    unsigned inb = 0, pml = 0, scat0 = 0, scat1 = 0;
    float2 A0[PML ? R : 1], A1[PML ? R : 1], A2[PML ? R : 1], A3[PML ? R : 1], A4[PML ? R : 1], A5[PML ? R : 1];
    unsigned char pmlx[PML ? R : 1];
    // UPML tiles: auxiliaries are loaded only outside the box that holds no UPML
    // cell (half the lanes of a side-band tile, most rows of an end-band tile's
    // inner part); inside it they are never read
    const int axlo = PML ? uni_i(I.axlo) : 0, axhi = PML ? uni_i(I.axhi) : 0;
    const int aylo = PML ? uni_i(I.aylo) : 0, ayhi = PML ? uni_i(I.ayhi) : 0, acw = PML ? uni_i(I.acw) : 0;
    const bool colbox = PML && jc >= aylo && jc < ayhi;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = ib + r;
      const int ic = min(max(i, 0), NPX - 1);
      const unsigned k = (unsigned)(ic * P + jc);  // cells < 2^28: checked at allocation
      F0[r] = ldo(g0, 8u * k);
      F1[r] = ldo(g1, 8u * k);
      F2[r] = ldo(g2, 8u * k);
      ie0[r] = ldo(gie0, 4u * k);
      ie1[r] = MODE == FDTD_MODE_TE ? ldo(gie1, 4u * k) : 1.0f;
      if (PML) {
        pmlx[r] = gpx[ic];
        A0[r] = A1[r] = A2[r] = A3[r] = A4[r] = A5[r] = make_float2(0.f, 0.f);
        // (ic, jc) lies outside the box, so its auxiliary cell exists
        if (!(colbox && ic >= axlo && ic < axhi)) {
          const unsigned ka = 8u * aux_cell(ic, jc, NPX, P, axlo, axhi, aylo, ayhi, acw);
          A0[r] = ldo(ga[0], ka);
          A1[r] = ldo(ga[1], ka);
          A2[r] = ldo(ga[2], ka);
          A3[r] = ldo(ga[3], ka);
          A4[r] = ldo(ga[4], ka);
          A5[r] = ldo(ga[5], ka);
        }
      }
      if (jin && i >= 0 && i < NPX) inb |= 1u << r;
    }
    unsigned upd = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = ib + r;
      const bool in = inb & (1u << r);
      if (!in) {
        // outside the grid: the reference's ghost ring, zero and never updated
        F0[r] = F1[r] = F2[r] = make_float2(0.f, 0.f);
        ie0[r] = ie1[r] = 1.0f;
      }
      const bool u = in && ((FLAV == FDTD_FLAVOUR_MPI) || (i >= 1 && i <= NPX - 2 && jj >= 1 && jj <= NPY - 2));
      if (u) upd |= 1u << r;
      if (BATCH || a.src) {
        if (ie0[r] != 1.0f) scat0 |= 1u << r;
        if (MODE == FDTD_MODE_TE && ie1[r] != 1.0f) scat1 |= 1u << r;
      }
      if (PML) {
        if (in && (pmlx[r] | pmly)) pml |= 1u << r;
        if (!(pml & (1u << r))) A0[r] = A1[r] = A2[r] = A3[r] = A4[r] = A5[r] = make_float2(0.f, 0.f);
      }
    }
    // UPML tiles: a wavefront none of whose cells uses the UPML recurrence (the
    // margin rows of the end bands' tiles) runs the regular update under the
    // update mask only; its auxiliaries stay zero and are never stored
    const bool wpm = PML && __ballot(pml != 0) != 0;
    // TM: the plain Gaussian pulse and the MPI scattered wave act on Ez (component 0)
    // TE: MPI scattered wave on Ey only (mpiTE_UPML.c:296); plain pulse on Ex and Ey
    if (MODE == FDTD_MODE_TE && FLAV == FDTD_FLAVOUR_MPI) scat0 = 0;
    if (FLAV == FDTD_FLAVOUR_PLAIN) {
      if (!uni_i(I.on[0])) scat0 = 0;
      if (!uni_i(I.on[1])) scat1 = 0;
    }

    // Interior tiles of the TE and MPI bodies: the 1/eps of each cell are kept in
    // LDS rather than registers (each wavefront writes and reads its own rows
    // only), which keeps these bodies within the register budget of two
    // workgroups per compute unit (the plain TM body fits with them in registers)
    constexpr bool IEL = !PML && (MODE == FDTD_MODE_TE || FLAV == FDTD_FLAVOUR_MPI);
    __shared__ float sie[IEL ? 2 : 1][IEL ? NR : 1][64];
    if constexpr (IEL) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        sie[0][w * R + r][lane] = ie0[r];
        sie[1][w * R + r][lane] = ie1[r];
      }
    }
    // b + a s, b - a s of the regular updates: the bodies with 1/eps in LDS spell
    // the fused rounding out (their merged and split instantiations otherwise
    // contract it differently, and the launch shapes must agree bit for bit);
    // the other bodies leave the contraction to the compiler (spelled out, the
    // plain TM interior body is scheduled 4 % slower)
    auto fma2 = [&](float2 a, float sc, float2 b) -> float2 {
      if constexpr (IEL) return c_fma(a, sc, b);
      else return c_add(b, c_scale(a, sc));
    };
    auto fnma2 = [&](float2 a, float sc, float2 b) -> float2 {
      if constexpr (IEL) return c_fnma(a, sc, b);
      else return c_sub(b, c_scale(a, sc));
    };
    auto IE0 = [&](int r) -> float {
      if constexpr (IEL) return sie[0][w * R + r][lane];
      else return ie0[r];
    };
    auto IE1 = [&](int r) -> float {
      if constexpr (IEL) return sie[1][w * R + r][lane];
      else return ie1[r];
    };

    // source term of component c in row r of a scatterer cell: (value now, change since the previous step)
    auto source = [&](int c, int r, float ie, const FastStep &st, float2 &now, float2 &delta) {
      if (FLAV == FDTD_FLAVOUR_MPI) {
        const float2 g = c_scale(c_mul(sex[w * R + r], eyl), ie - 1.0f);
        now = c_mul(g, st.c);
        delta = c_mul(g, st.dc);
      } else {
        const int i = ib + r;
        const float tn = c == 0 ? st.tn : st.tn1;
        const float tp = c == 0 ? st.tp : st.tp1;
        now = pulse(pc, c, i, jj, ie, tn);
        const float2 prev = (st.hasprev != 0.0f) ? pulse(pc, c, i, jj, ie, tp) : make_float2(0.f, 0.f);
        delta = c_sub(now, prev);
      }
    };

    // ------------------------------------------------------------ half steps
    // Interior tiles (PML == 0) hold only regular cells within the dependency
    // cone of their core (the tiling keeps UPML, ghost and fixed-edge cells
    // outside it), so every row is updated without masks; UPML tiles keep the
    // per-cell update / UPML bits.  In the MPI flavour the scattered-wave change
    // (s_n - s_{n-1}) = g * dc with g = e^{i k.r} (1/eps - 1) is added to every
    // row of a wavefront that holds a scatterer cell (g = 0 in vacuum).
    const unsigned smask = MODE == FDTD_MODE_TM ? scat0 : (scat0 | scat1);
    const bool wsrc = __ballot(smask != 0) != 0;  // wavefront-uniform
    auto src_g = [&](int r, float ie) -> float2 { return c_scale(c_mul(sex[w * R + r], eyl), ie - 1.0f); };
    // MPI flavour, interior tiles without sampling: g of each cell, formed once
    // per tile in wavefronts holding a scatterer cell (TM: on Ez with 1/eps_z;
    // TE: on Ey with 1/eps_y) and kept in LDS (each wavefront writes and reads
    // only its own rows; registers would cost the second wavefront per SIMD);
    // the other tiles, and the 8-wavefront TM interior bodies, form it at each
    // step (the same roundings).  Those bodies fit their 128 registers that way:
    // the two step loops of the LDS form (with and without the source) cost them
    // 8 spilled registers at 8 rows per wavefront, and the 9-row body fits
    // without spilling; the 4-wavefront bodies have registers to spare and keep
    // the LDS form, which saves them the per-step work.  The change g * dc is
    // four FMAs.
    constexpr bool GROWS = FLAV == FDTD_FLAVOUR_MPI && !PML && !SMP && !(MODE == FDTD_MODE_TM && W == 8);
    __shared__ float2 sg[GROWS ? NR : 1][64];
    auto add_src = [&](float2 e, int r, float ie, const FastStep &st) -> float2 {
      float2 g;
      if constexpr (GROWS) g = sg[w * R + r][lane];
      else g = src_g(r, ie);
      e.x = fmaf(g.x, st.dc.x, e.x);
      e.x = fmaf(-g.y, st.dc.y, e.x);
      e.y = fmaf(g.x, st.dc.y, e.y);
      e.y = fmaf(g.y, st.dc.x, e.y);
      return e;
    };
    // the source of component c in row r: value now (sn) and change (sd); zero in vacuum
    auto src_any = [&](int c, int r, float ie, const FastStep &st, auto SRCT, float2 &sn, float2 &sd) {
      constexpr bool SRC = decltype(SRCT)::value;
      if (FLAV == FDTD_FLAVOUR_MPI) {
        if (SRC) {
          const float2 g = src_g(r, ie);
          sn = c_mul(g, st.c);
          sd = c_mul(g, st.dc);
        }
      } else if (SRC && ((c == 0 ? scat0 : scat1) & (1u << r))) {
        source(c, r, ie, st, sn, sd);
      }
    };
    // UPML tiles run both recurrences on every cell of a wavefront that holds a
    // UPML cell and keep the right one with per-cell selects (no divergent
    // branches), the regular one alone in the other wavefronts (a uniform
    // branch); an auxiliary is written back only where the cell is updated, so
    // the registers of the other cells may hold anything.
    const float s2p = 2.0f + sy, s2m = 2.0f - sy;  // C_BYMY1/0, C_DYJY1/0 numerators of this column

    auto e_rows = [&](const FastStep &st, auto SRCT) {
      constexpr bool SRC = decltype(SRCT)::value;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int rr = w * R + r;  // row within the block (LDS tables)
        const unsigned bit = 1u << r;
        if (MODE == FDTD_MODE_TM) {
          const float2 hym = r > 0 ? F2[r - 1] : xH[w][lane];
          const float2 hxm = shfl_up1(F1[r]);
          const float2 curl = c_sub(c_sub(F2[r], hym), c_sub(F1[r], hxm));
          if (!PML) {
            float2 e = fma2(curl, IE0(r), F0[r]);
            if (FLAV == FDTD_FLAVOUR_MPI) {
              if (SRC) e = add_src(e, r, IE0(r), st);
            }
            F0[r] = e;  // plain: the pulse is added below, after every row's stencil
          } else {
            // -0 is the additive identity (x + -0 == x for every x, -0 included), so
            // the source additions fold away where no source is formed
            float2 sn = make_float2(-0.f, -0.f), sd = sn;
            src_any(0, r, IE0(r), st, SRCT, sn, sd);
            const float2 reg = c_add(fma2(curl, IE0(r), F0[r]), sd);
            if (wpm) {
              const float cjz = srow[0][rr], cjzh = srow[1][rr];
              const float2 nj = c_add(c_scale(A0[r], cjz), c_scale(curl, cjzh));
              const float2 nd = c_add(c_scale(A1[r], cy0), c_scale(c_sub(nj, A0[r]), cy1));
              A0[r] = nj;
              A1[r] = nd;
              const float2 e = sel(pml & bit, c_add(c_scale(nd, IE0(r)), sn), reg);
              F0[r] = sel(upd & bit, e, F0[r]);
            } else {
              F0[r] = sel(upd & bit, reg, F0[r]);
            }
          }
        } else {
          const float2 hzj = shfl_up1(F2[r]);
          const float2 hzi = r > 0 ? F2[r - 1] : xH[w][lane];
          const float2 cx = c_sub(F2[r], hzj);   // Hz - Hz(j-1)
          const float2 cyv = c_sub(hzi, F2[r]);  // Hz(i-1) - Hz
          if (!PML) {
            float2 ex = fma2(cx, IE0(r), F0[r]);
            float2 ey = fma2(cyv, IE1(r), F1[r]);
            if (FLAV == FDTD_FLAVOUR_MPI) {
              if (SRC) ey = add_src(ey, r, IE1(r), st);
            }
            F0[r] = ex;  // plain: the pulse is added below, after every row's stencil
            F1[r] = ey;
            // 10 rows per wavefront: a scheduling fence after each row keeps the
            // rows' 1/eps reads from LDS next to their use; hoisted ahead of the
            // whole E phase, their 20 live registers took the body past the 128
            // of four wavefronts per SIMD (6 spilled, 124 with the fence)
            if constexpr (IEL && R == 10) __builtin_amdgcn_sched_barrier(0);
          } else {
            float2 sn0 = make_float2(-0.f, -0.f), sd0 = sn0, sn1 = sn0, sd1 = sn0;  // -0: as in TM
            // TE MPI: the scattered wave acts on Ey only
            if (FLAV == FDTD_FLAVOUR_PLAIN) src_any(0, r, IE0(r), st, SRCT, sn0, sd0);
            src_any(1, r, IE1(r), st, SRCT, sn1, sd1);
            const float2 rex = c_add(fma2(cx, IE0(r), F0[r]), sd0);
            const float2 rey = c_add(fma2(cyv, IE1(r), F1[r]), sd1);
            if (!wpm) {
              F0[r] = sel(upd & bit, rex, F0[r]);
              F1[r] = sel(upd & bit, rey, F1[r]);
            } else {
              // Jx/Dx: x tables 0 C_DXJX1 1 C_DXJX0; y: 0 C_JX 1 C_JXHZ
              const float2 njx = c_add(c_scale(A0[r], cy0), c_scale(cx, cy1));
              const float2 ndx = c_sub(c_add(A1[r], c_scale(njx, srow[0][rr])), c_scale(A0[r], srow[1][rr]));
              // Jy/Dy: x table 2 C_DY; C_DYJY1/0 = (2 +- sigma_y) / (2 + sigma_x)
              const float2 njy = c_add(A2[r], cyv);
              const float rsx = srow[5][rr];
              const float c1 = s2p * rsx, c0 = s2m * rsx;
              const float2 ndy = c_sub(c_add(c_scale(A3[r], srow[2][rr]), c_scale(njy, c1)), c_scale(A2[r], c0));
              A0[r] = njx; A1[r] = ndx; A2[r] = njy; A3[r] = ndy;
              const bool p = pml & bit, u = upd & bit;
              const float2 ex = sel(p, c_add(c_scale(ndx, IE0(r)), sn0), rex);
              const float2 ey = sel(p, c_add(c_scale(ndy, IE1(r)), sn1), rey);
              F0[r] = sel(u, ex, F0[r]);
              F1[r] = sel(u, ey, F1[r]);
            }
          }
        }
      }
      // Plain interior tiles: the pulse's change in the scatterer cells, added to
      // every row after all rows' stencils (the same two roundings as adding it
      // right after each stencil: it depends on the cell and step only), so the
      // stencils and the next step's H rows form one block without branches,
      // whose lane shifts fold into the subtractions that use them
      if (FLAV == FDTD_FLAVOUR_PLAIN && !PML && wsrc) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const unsigned bit = 1u << r;
          if (scat0 & bit) {
            float2 sn, sd;
            source(0, r, IE0(r), st, sn, sd);
            F0[r] = c_add(F0[r], sd);
          }
          if (MODE == FDTD_MODE_TE && (scat1 & bit)) {
            float2 sn, sd;
            source(1, r, IE1(r), st, sn, sd);
            F1[r] = c_add(F1[r], sd);
          }
        }
      }
      // first row of E (TM Ez, TE Ey) for wave w - 1
      xE[w][lane] = MODE == FDTD_MODE_TM ? F0[0] : F1[0];
    };

    auto h_phase = [&]() {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int rr = w * R + r;
        const unsigned bit = 1u << r;
        if (MODE == FDTD_MODE_TM) {
          const float2 ejp = shfl_dn1(F0[r]);
          const float2 eip = r < R - 1 ? F0[r + 1] : xE[w + 1][lane];
          const float2 dEy = c_sub(ejp, F0[r]);  // Ez(j+1) - Ez
          const float2 dEx = c_sub(eip, F0[r]);  // Ez(i+1) - Ez
          if (!PML) {
            F1[r] = fnma2(dEy, c2, F1[r]);
            F2[r] = fma2(dEx, c2, F2[r]);
          } else {
            const float2 rhx = fnma2(dEy, c2, F1[r]), rhy = fma2(dEx, c2, F2[r]);
            if (!wpm) {
              F1[r] = sel(upd & bit, rhx, F1[r]);
              F2[r] = sel(upd & bit, rhy, F2[r]);
            } else {
              // Mx/Bx: y 2 C_MX 3 C_MXEZ, x 2 C_BXMX1 3 C_BXMX0 ; My/By: x 4 C_BY, C_BYMY1/0
              // the old Mx / My terms are taken first, so that the new values can
              // take the old ones' registers (4 fewer register copies per row and step)
              const float rsx = srow[5][rr];
              const float c1 = s2p * rsx, c0 = s2m * rsx;
              const float2 bx0 = c_sub(A3[r], c_scale(A2[r], srow[3][rr]));
              const float2 by0 = c_sub(c_scale(A5[r], srow[4][rr]), c_scale(A4[r], c0));
              const float2 nmx = c_sub(c_scale(A2[r], cy2), c_scale(dEy, cy3));
              const float2 nmy = c_add(A4[r], dEx);
              const float2 nbx = c_add(bx0, c_scale(nmx, srow[2][rr]));
              const float2 nby = c_add(by0, c_scale(nmy, c1));
              A2[r] = nmx; A3[r] = nbx; A4[r] = nmy; A5[r] = nby;
              const bool p = pml & bit, u = upd & bit;
              const float2 hx = sel(p, c_scale(nbx, c2), rhx);
              const float2 hy = sel(p, c_scale(nby, c2), rhy);
              F1[r] = sel(u, hx, F1[r]);
              F2[r] = sel(u, hy, F2[r]);
            }
          }
        } else {
          const float2 exj = shfl_dn1(F0[r]);
          const float2 eyi = r < R - 1 ? F1[r + 1] : xE[w + 1][lane];
          // ((Ey(i+1) - Ey) - Ex(j+1)) + Ex
          const float2 ce = c_add(c_sub(c_sub(eyi, F1[r]), exj), F0[r]);
          if (!PML) {
            F2[r] = fnma2(ce, c2, F2[r]);
          } else {
            const float2 rhz = fnma2(ce, c2, F2[r]);
            if (!wpm) {
              F2[r] = sel(upd & bit, rhz, F2[r]);
            } else {
              // Mz/Bz: x 3 C_MZ 4 C_MZEXEY ; y 2 C_BZ 3 C_BZMZ
              const float2 nmz = c_sub(c_scale(A4[r], srow[3][rr]), c_scale(ce, srow[4][rr]));
              const float2 nbz = c_sub(c_add(c_scale(A5[r], cy2), c_scale(nmz, cy3)), c_scale(A4[r], cy3));
              A4[r] = nmz; A5[r] = nbz;
              const float2 hz = sel(pml & bit, c_scale(nbz, c2), rhz);
              F2[r] = sel(upd & bit, hz, F2[r]);
            }
          }
        }
      }
      xH[w + 1][lane] = F2[R - 1];  // last row of Hy (TM) / Hz (TE) for wave w + 1
    };

    // per-step NTFF samples of the contour cells in this tile's core (the state
    // after each full step, as the reference's per-step NTFF reads it).  The
    // plan is fixed for the tile.  A side with j fixed lies in one lane: that
    // lane stores its rows `lmask` (packed index lc + r) with a store per row
    // under its own mask; the host keeps the two such sides' columns apart, so
    // a lane has one role at most (1: E and H of the side, 2: H one column
    // back).  A side with i fixed lies in one row of one wave: the wave keeps
    // the rows holding such a side (evr), with the side (2 bits per row, evq)
    // and whether the row is the side itself (E and H) or its H neighbour (evk);
    // the host keeps the two sides' rows apart, so a row has one event at most.
    bool smpL = false;
    unsigned lmask = 0, evr = 0, evq = 0, evk = 0;
    int lrole = 0, lc = 0;
    bool lef = false, lhf = false;
    float les = 1.f;
    const bool jcore = jj >= t.j0 && jj < t.j0 + t.nj;
    if (SMP) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const FastCtrSide &g = a.ct.s[q];
        if (g.n <= 0) continue;
        if (g.horiz) {
          const int role = !jcore ? 0 : (jj == g.fi ? 1 : (jj == g.fi - 1 ? 2 : 0));
          if (role) {
            unsigned m = 0;
            for (int r = 0; r < R; r++) {
              const int i = ib + r;
              if (i >= t.i0 && i < t.i0 + t.ni && i >= g.lo && i < g.lo + g.n) m |= 1u << r;
            }
            lmask = m;
            lrole = role;
            lc = g.off + (ib - g.lo);
            lef = g.ef != 0;
            lhf = g.hf == 1;
            les = g.es;
          }
        } else {
          const int r1 = g.fi - ib, r2 = g.fi - 1 - ib;
          if (g.fi >= t.i0 && g.fi < t.i0 + t.ni && r1 >= 0 && r1 < R) {
            evr |= 1u << r1;
            evq |= (unsigned)q << (2 * r1);
            evk |= 1u << r1;
          }
          if (g.fi - 1 >= t.i0 && g.fi - 1 < t.i0 + t.ni && r2 >= 0 && r2 < R) {
            evr |= 1u << r2;
            evq |= (unsigned)q << (2 * r2);
          }
        }
      }
      smpL = __ballot(lmask != 0) != 0;
      evr = (unsigned)uni_i((int)evr);
      evq = (unsigned)uni_i((int)evq);
      evk = (unsigned)uni_i((int)evk);
    }
    auto sample = [&](int s) {
      float2 *base = a.csamp + (size_t)s * 3 * a.ct.total;
      const int tot = a.ct.total;
      if (smpL) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          if (lmask & (1u << r)) {
            const float2 h = lhf ? F1[r] : F2[r];
            const float2 e = c_scale(lef ? F1[r] : F0[r], les);
            const int c = lc + r;
            base[lrole == 1 ? c : 2 * tot + c] = lrole == 1 ? e : h;
            if (lrole == 1) base[tot + c] = h;
          }
        }
      }
      if (evr) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          if (!(evr & (1u << r))) continue;
          const FastCtrSide &g = a.ct.s[(evq >> (2 * r)) & 3];
          if (jcore && jj >= g.lo && jj < g.lo + g.n) {
            const int c = g.off + (jj - g.lo);
            const float2 h = g.hf == 1 ? F1[r] : F2[r];
            if (evk & (1u << r)) {
              base[c] = c_scale(g.ef == 0 ? F0[r] : F1[r], g.es);
              base[tot + c] = h;
            } else {
              base[2 * tot + c] = h;
            }
          }
        }
      }
    };

    if (FLAV == FDTD_FLAVOUR_MPI) {
      xH[w + 1][lane] = F2[R - 1];
      __syncthreads();
      if constexpr (GROWS) {
        // interior tiles: two copies of the step loop, chosen per wavefront, with
        // the scattered wave (g formed once here) and without it (vacuum)
        auto steps = [&](auto SRCT) {
          constexpr bool SRC = decltype(SRCT)::value;
          if (SRC) {
#pragma unroll
            for (int r = 0; r < R; r++) sg[w * R + r][lane] = src_g(r, MODE == FDTD_MODE_TM ? IE0(r) : IE1(r));
          }
          for (int s = 0; s < T; s++) {
            e_rows(sst[s], SRCT);
            __syncthreads();
            h_phase();
            __syncthreads();
            if (SMP) sample(s);
          }
        };
        if (wsrc) steps(std::true_type());
        else steps(std::false_type());
      } else {
        // UPML and sampling tiles (their registers are full with the auxiliaries
        // or the sampling plan): one loop
        for (int s = 0; s < T; s++) {
          if (wsrc) e_rows(sst[s], std::true_type());
          else e_rows(sst[s], std::false_type());
          __syncthreads();
          h_phase();
          __syncthreads();
          if (SMP) sample(s);
        }
      }
    } else {
      xE[w][lane] = MODE == FDTD_MODE_TM ? F0[0] : F1[0];
      __syncthreads();
      for (int s = 0; s < T; s++) {
        h_phase();
        __syncthreads();
        // TM UPML tiles of 8 or more wavefronts: the E rows without the pulse in
        // wavefronts that hold no scatterer cell (a second copy of them costs
        // the TE bodies, and the 4-wavefront kernels that also run interior
        // tiles, spilled registers); the interior tiles add the pulse after
        // their stencils
        if constexpr (PML && MODE == FDTD_MODE_TM && W >= 8) {
          if (wsrc) e_rows(sst[s], std::true_type());
          else e_rows(sst[s], std::false_type());
        } else if constexpr (PML) {
          e_rows(sst[s], std::true_type());
        } else {
          e_rows(sst[s], std::false_type());
        }
        __syncthreads();
        if (SMP) sample(s);
      }
    }

    // ------------------------------------------------------------ write back the core
    GPTR(float2) o0 = uni_p(I.f0[nxt]);
    GPTR(float2) o1 = uni_p(I.f1[nxt]);
    GPTR(float2) o2 = uni_p(I.f2[nxt]);
    GPTR(float2) oa[6];
    if (PML) {
#pragma unroll
      for (int qq = 0; qq < 6; qq++) oa[qq] = uni_p(I.aux[qq][nxt]);
    }
    if (jj >= t.j0 && jj < t.j0 + t.nj) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = ib + r;
        if (i < t.i0 || i >= t.i0 + t.ni) continue;
        const unsigned bo = 8u * (unsigned)(i * P + jj);
        sto(o0, bo, F0[r]);
        sto(o1, bo, F1[r]);
        sto(o2, bo, F2[r]);
        if (PML && (pml & upd & (1u << r))) {
          // a UPML cell lies outside the box
          const unsigned ba = 8u * aux_cell(i, jj, NPX, P, axlo, axhi, aylo, ayhi, acw);
          sto(oa[0], ba, A0[r]);
          sto(oa[1], ba, A1[r]);
          sto(oa[2], ba, A2[r]);
          sto(oa[3], ba, A3[r]);
          sto(oa[4], ba, A4[r]);
          sto(oa[5], ba, A5[r]);
        }
      }
    }
  }
}

// XCD-aware static schedule: workgroups are dispatched round-robin over the 8
// XCDs, so workgroup b (of nblk, both offsets that are multiples of 8) runs on
// XCD b % 8 and walks that XCD's spatially contiguous partition of the tile
// list [qbeg[q], qbeg[q+1]) with stride nblk/8.  Neighbouring tiles (which read
// each other's halos) share an L2, and no device-scope atomic is needed.
template <int MODE, int FLAV, int R, int PML, int SMP = 0, int W = 4>
__device__ __forceinline__ void fast_tiles(const FastArgs &a, const int *qbeg, int b, int nblk)
{
  const int q = b & 7;
  const int nb = nblk >> 3;
  for (int tid = qbeg[q] + (b >> 3); tid < qbeg[q + 1]; tid += nb) fast_tile<MODE, FLAV, R, PML, 0, SMP, W>(a, tid);
}

// KIND 0: one launch per block of T steps, workgroups [0, nb_pml) take the UPML
// tiles (R = 8, full recurrence) and the rest the interior tiles, so the two
// kinds run side by side.  Its register budget is that of the UPML body, which
// limits a compute unit to two workgroups; on large grids, where UPML tiles are
// a small share and the interior runs in many dispatch rounds, the interior
// (KIND 1) and UPML (KIND 2) tiles are launched separately instead, so the
// interior kernel keeps the occupancy of its own, smaller register budget.
// KIND 3: KIND 0 with the per-step NTFF samples (a separate instantiation, so
// that the other kinds do not carry the sampling code's registers).
// Interior tile bodies are held to the occupancy their registers allow (R = 16:
// two wavefronts per SIMD, R = 8: four, R = 4: six): the step loop is
// straight-line code, and unconstrained scheduling spends registers on overlap
// past the point where a further wavefront would no longer fit.  The TM kinds
// that run UPML tiles are held to two wavefronts per SIMD (256 registers): the
// sampling ones needed 262-264 and ran at one, for a spill of 6-8 registers;
// the TE bodies would spill far more and keep their one.
__host__ __device__ constexpr int fast_interior_waves(int R) { return R >= 16 ? 2 : (R >= 8 ? 4 : 6); }

template <int MODE, int FLAV, int R, int KIND>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu((KIND == 1 || KIND == 4) ? fast_interior_waves(R) : (MODE == FDTD_MODE_TM ? 2 : 1)))) void k_fast(FastArgs a)
{
  if (KIND == 3) {
    if ((int)blockIdx.x < a.nb_pml) fast_tiles<MODE, FLAV, 8, 1, 1>(a, a.qb[0], blockIdx.x, a.nb_pml);
    else fast_tiles<MODE, FLAV, R, 0, 1>(a, a.qb[1], (int)blockIdx.x - a.nb_pml, (int)gridDim.x - a.nb_pml);
  } else if (KIND == 1) {
    fast_tiles<MODE, FLAV, R, 0>(a, a.qb[1], blockIdx.x, gridDim.x);
  } else if (KIND == 2) {
    fast_tiles<MODE, FLAV, 8, 1>(a, a.qb[0], blockIdx.x, gridDim.x);
  } else if (KIND == 4) {  // KIND 1 with the per-step NTFF samples
    fast_tiles<MODE, FLAV, R, 0, 1>(a, a.qb[1], blockIdx.x, gridDim.x);
  } else if (KIND == 5) {  // KIND 2 with the per-step NTFF samples
    fast_tiles<MODE, FLAV, 8, 1, 1>(a, a.qb[0], blockIdx.x, gridDim.x);
  } else {
    if ((int)blockIdx.x < a.nb_pml) fast_tiles<MODE, FLAV, 8, 1>(a, a.qb[0], blockIdx.x, a.nb_pml);
    else fast_tiles<MODE, FLAV, R, 0>(a, a.qb[1], (int)blockIdx.x - a.nb_pml, (int)gridDim.x - a.nb_pml);
  }
}

// The wide UPML tiles (FAST_UPML_ROWS_WIDE = 48 rows).  TM: 12 wavefronts of 4
// rows (768 threads, 156-162 VGPRs: three wavefronts per SIMD, one workgroup
// per compute unit).  TE, whose body holds a second eps and source and needs
// 168+ VGPRs at 4 rows: 8 wavefronts of 6 rows (512 threads, two per SIMD).  A
// 48-row tile keeps 48 - 2T - 1 core rows (27 at T = 10) where a 32-row tile
// keeps 11: the halo rows the UPML tiles compute again fall by 2.5x.
__host__ __device__ constexpr int pml48_waves(int mode) { return mode == FDTD_MODE_TM ? 12 : 8; }
template <int MODE, int FLAV, int SMP>
__global__ __launch_bounds__(64 * pml48_waves(MODE)) void k_fast_pml48(FastArgs a)
{
  fast_tiles<MODE, FLAV, FAST_UPML_ROWS_WIDE / pml48_waves(MODE), 1, SMP, pml48_waves(MODE)>(a, a.qb[0], blockIdx.x,
                                                                                            gridDim.x);
}

// The wide UPML tiles as 16 wavefronts of 3 rows (1024 threads): fewer
// registers per wavefront, four wavefronts per SIMD, for grids whose UPML band
// has fewer tiles than the chip has CUs (the band's launch is then latency-bound).
template <int MODE, int FLAV, int SMP>
__global__ __launch_bounds__(1024) void k_fast_pml16(FastArgs a)
{
  fast_tiles<MODE, FLAV, 3, 1, SMP, 16>(a, a.qb[0], blockIdx.x, gridDim.x);
}

// Interior tiles (4R rows) as 8 wavefronts of R/2 rows: the same tiles and
// per-cell code as KIND 1 / 4, with half the registers per wavefront, so that
// twice as many wavefronts share each SIMD (R = 16: four instead of two).
template <int MODE, int FLAV, int RH, int SMP>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(fast_interior_waves(RH)))) void k_fast_interior8(FastArgs a)
{
  fast_tiles<MODE, FLAV, RH, 0, SMP, 8>(a, a.qb[1], blockIdx.x, gridDim.x);
}

// Batched instances: every workgroup pulls tiles from the launch's two device
// queues, first the UPML tiles (the longest per tile), then the interior tiles,
// each queue ordered by decreasing instance size and cut into 8 parts, one per
// XCD (a workgroup takes its own XCD's part first, then the others'), until both
// are empty; every workgroup then exits.  One launch advances every unfinished instance by its
// own depth; k_batch_advance between launches moves the progress records.
// KQ: the queues pulled, 0 UPML tiles, 1 interior tiles, 2 both (UPML first).
// A one-queue instantiation carries one tile body only, so it runs at that
// body's register budget.
// W8 (one-queue instantiations): the tiles as 8 wavefronts (512 threads), UPML
// tiles of 32 rows as 8 x 4 rows, interior tiles of 4R rows as 8 x R/2 rows.
template <int MODE, int R, int KQ, int W8 = 0>
__global__ __launch_bounds__(W8 ? 512 : 256)
__attribute__((amdgpu_waves_per_eu(KQ == 1 ? fast_interior_waves(W8 ? R / 2 : R) : 1))) void k_fast_batch(FastArgs a)
{
  static_assert(!W8 || (KQ != 2 && R >= 8), "8-wavefront bodies: one queue, R >= 8");
  __shared__ int s_tid;
  for (int k = (KQ == 1 ? 1 : 0); k < (KQ == 0 ? 1 : 2); k++) {
    // the queue in 8 contiguous parts, one per XCD (workgroup b runs on XCD b % 8):
    // a workgroup pulls from its XCD's part, so neighbouring tiles of an instance
    // meet in one L2, then from the other parts once its own is empty
    const int xcd = blockIdx.x & 7;
    for (int pp = 0; pp < 8;) {
      __syncthreads();  // every thread has read s_tid of the previous pull
      if (threadIdx.x == 0) {
        const int q = (xcd + pp) & 7;
        const int lo = (int)((long)a.qn[k] * q / 8), hi = (int)((long)a.qn[k] * (q + 1) / 8);
        const int t = lo + atomicAdd(&a.qctr[8 * k + q], 1);
        s_tid = t < hi ? t : -1;
      }
      __syncthreads();
      const int tid = s_tid;
      if (tid < 0) {
        pp++;  // this part is empty (uniform: s_tid is shared)
        continue;
      }
      if constexpr (W8) {
        if (KQ == 0) fast_tile<MODE, FDTD_FLAVOUR_MPI, 4, 1, 1, 0, 8>(a, a.qoff[0] + tid);
        else fast_tile<MODE, FDTD_FLAVOUR_MPI, R / 2, 0, 1, 0, 8>(a, a.qoff[1] + tid);
      } else {
        if (KQ != 1 && k == 0) fast_tile<MODE, FDTD_FLAVOUR_MPI, 8, 1, 1, 0>(a, a.qoff[0] + tid);
        else if (KQ != 0) fast_tile<MODE, FDTD_FLAVOUR_MPI, R, 0, 1, 0>(a, a.qoff[1] + tid);
      }
    }
  }
}

// After a batched launch: count its steps into every instance, flip its
// ping-pong index, set the depth of the next launch (0 once finished) and clear
// the queue counters.  With n0 = 1 it only sets the depths (first launch).
__global__ __launch_bounds__(256) void k_batch_advance(FastBatchInst *b, int n, int T, int *qctr, int nq, int first,
                                                       FastStep *steps)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nq) qctr[i] = 0;
  if (i < n) {
    FastBatchInst x = b[i];
    if (!first && x.Ti > 0) {
      x.n += x.Ti;
      x.cur ^= 1;
    }
    const int left = x.steps - x.n;
    x.Ti = left <= 0 ? 0 : (left < T ? left : T);
    b[i] = x;
    // the next launch's step scalars, formed here once per instance (in FP64)
    // rather than by every tile, whose registers the FP64 code would weigh down
    for (int s = 0; s < x.Ti; s++) steps[i * FAST_MAX_T + s] = batch_step(x, s);
  }
}

// FP32 physical layout -> FP64 solver layout (MPI: padded, ghost ring untouched)
__global__ __launch_bounds__(256) void k_fast_export(const float2 *src, int pitch, int NPX, int NPY, double2 *dst, int NPYa,
                                                     int off)
{
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y * 4 + threadIdx.y;
  if (i >= NPX || j >= NPY) return;
  const float2 v = src[(size_t)i * pitch + j];
  dst[(size_t)(i + off) * NPYa + (j + off)] = make_double2((double)v.x, (double)v.y);
}

__global__ __launch_bounds__(256) void k_fast_import(const double2 *src, int NPYa, int off, int NPX, int NPY, float2 *dst,
                                                     int pitch)
{
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y * 4 + threadIdx.y;
  if (i >= NPX || j >= NPY) return;
  const double2 v = src[(size_t)(i + off) * NPYa + (j + off)];
  dst[(size_t)i * pitch + j] = make_float2((float)v.x, (float)v.y);
}

// The same for a UPML auxiliary in its compact layout (aux_cell): the cells of
// the box, which has no storage, export as zero (no update ever writes them)
// and import nothing.  src == nullptr (auxiliaries not allocated yet): zeros.
__global__ __launch_bounds__(256) void k_fast_aux_export(const float2 *src, FastInst h, double2 *dst, int NPYa, int off)
{
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y * 4 + threadIdx.y;
  if (i >= h.NPX || j >= h.NPY) return;
  const bool box = i >= h.axlo && i < h.axhi && j >= h.aylo && j < h.ayhi;
  float2 v = make_float2(0.f, 0.f);
  if (src && !box) v = src[aux_cell(i, j, h.NPX, h.pitch, h.axlo, h.axhi, h.aylo, h.ayhi, h.acw)];
  dst[(size_t)(i + off) * NPYa + (j + off)] = make_double2((double)v.x, (double)v.y);
}

__global__ __launch_bounds__(256) void k_fast_aux_import(const double2 *src, int NPYa, int off, FastInst h, float2 *dst)
{
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y * 4 + threadIdx.y;
  if (i >= h.NPX || j >= h.NPY) return;
  if (i >= h.axlo && i < h.axhi && j >= h.aylo && j < h.ayhi) return;
  const double2 v = src[(size_t)(i + off) * NPYa + (j + off)];
  dst[aux_cell(i, j, h.NPX, h.pitch, h.axlo, h.axhi, h.aylo, h.ayhi, h.acw)] = make_float2((float)v.x, (float)v.y);
}

// 1/eps in FP32 from the FP64 eps map (solver layout)
__global__ __launch_bounds__(256) void k_fast_inv_eps(const double *eps, int NPYa, int off, int NPX, int NPY, float *ie, int pitch)
{
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int i = blockIdx.y * 4 + threadIdx.y;
  if (i >= NPX || j >= NPY) return;
  const double e = eps[(size_t)(i + off) * NPYa + (j + off)];
  ie[(size_t)i * pitch + j] = (e == 1.0) ? 1.0f : (float)(1.0 / e);
}

// contour samples of one step from the FP32 fields, in the packing of k_ntff_gather
__global__ __launch_bounds__(256) void k_fast_gather(const FastInst *inst, int cur, Contour ct, int off, int mode,
                                                     double2 *cE, double2 *cH)
{
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= ct.total) return;
  const FastInst &I = inst[0];
  int s = 0;
  while (s < 3 && p >= ct.s[s + 1].off) s++;
  const Seg &sg = ct.s[s];
  const int c = p - sg.off;
  const long ka = sg.k0 + (long)c * sg.dk;  // solver-layout index
  const int NPYa = I.NPY + 2 * off;
  const int ia = (int)(ka / NPYa), ja = (int)(ka % NPYa);
  const int i = ia - off, j = ja - off;
  auto get = [&](int slot, int di, int dj) -> float2 {
    const int ii = i + di, jv = j + dj;
    if (ii < 0 || jv < 0 || ii >= I.NPX || jv >= I.NPY) return make_float2(0.f, 0.f);
    const size_t k = (size_t)ii * I.pitch + jv;
    return slot == 0 ? I.f0[cur][k] : (slot == 3 ? I.f1[cur][k] : I.f2[cur][k]);
  };
  // slots in the FP64 numbering: TM Ez 0 Hx 3 Hy 6 ; TE Ex 0 Ey 3 Hz 6
  const float2 e = get(sg.ef, 0, 0);
  const int di = sg.hnb == -1 ? 0 : -1, dj = sg.hnb == -1 ? -1 : 0;
  const float2 h0 = get(sg.hf, 0, 0), h1 = get(sg.hf, di, dj);
  cE[p] = make_double2(sg.es * (double)e.x, sg.es * (double)e.y);
  cH[p] = make_double2(sg.hs * ((double)h0.x + (double)h1.x), sg.hs * ((double)h0.y + (double)h1.y));
  (void)mode;
}

// ---------------------------------------------------------------- host side

#define HIPCHK(x)                                                                                    \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) return fdtd_fail(FDTD_EHIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

template <int MODE, int FLAV, int R, int KIND>
static void launch_one(const FastArgs &a, int nblocks, hipStream_t st)
{
  hipLaunchKernelGGL((k_fast<MODE, FLAV, R, KIND>), dim3(nblocks), dim3(256), 0, st, a);
}

template <int MODE, int FLAV, int SMP>
static void launch_pml48(const FastArgs &a, int nblocks, hipStream_t st)
{
  hipLaunchKernelGGL((k_fast_pml48<MODE, FLAV, SMP>), dim3(nblocks), dim3(64 * pml48_waves(MODE)), 0, st, a);
}

template <int MODE, int FLAV, int SMP>
static void launch_pml16(const FastArgs &a, int nblocks, hipStream_t st)
{
  hipLaunchKernelGGL((k_fast_pml16<MODE, FLAV, SMP>), dim3(nblocks), dim3(1024), 0, st, a);
}

template <int MODE, int FLAV, int RH, int SMP>
static void launch_interior8(const FastArgs &a, int nblocks, hipStream_t st)
{
  hipLaunchKernelGGL((k_fast_interior8<MODE, FLAV, RH, SMP>), dim3(nblocks), dim3(512), 0, st, a);
}

typedef void (*fast_launcher)(const FastArgs &, int, hipStream_t);

template <int MODE, int FLAV>
static fast_launcher pick(int R, int kind)
{
  if (kind == 2) return launch_one<MODE, FLAV, 8, 2>;
  if (kind == 5) return launch_one<MODE, FLAV, 8, 5>;
  if (kind == 6) return launch_pml48<MODE, FLAV, 0>;
  if (kind == 7) return launch_pml48<MODE, FLAV, 1>;
  if (kind == 10) return launch_pml16<MODE, FLAV, 0>;
  if (kind == 11) return launch_pml16<MODE, FLAV, 1>;
  // interior tiles of 4R rows as 8 wavefronts of R/2 rows (R = 8, 16, 18, 20;
  // R = 18 and 20 have only these bodies: the launch refuses them for the other shapes)
  if (kind == 8 || kind == 9) {
    const bool smp = kind == 9;
    if (R == 8) return smp ? launch_interior8<MODE, FLAV, 4, 1> : launch_interior8<MODE, FLAV, 4, 0>;
    if (R == 18) return smp ? launch_interior8<MODE, FLAV, 9, 1> : launch_interior8<MODE, FLAV, 9, 0>;
    if (R == 20) return smp ? launch_interior8<MODE, FLAV, 10, 1> : launch_interior8<MODE, FLAV, 10, 0>;
    return smp ? launch_interior8<MODE, FLAV, 8, 1> : launch_interior8<MODE, FLAV, 8, 0>;
  }
  if (kind == 4) {
    if (R == 4) return launch_one<MODE, FLAV, 4, 4>;
    if (R == 8) return launch_one<MODE, FLAV, 8, 4>;
    return launch_one<MODE, FLAV, 16, 4>;
  }
  if (kind == 3) {
    if (R == 4) return launch_one<MODE, FLAV, 4, 3>;
    if (R == 8) return launch_one<MODE, FLAV, 8, 3>;
    return launch_one<MODE, FLAV, 16, 3>;
  }
  if (kind == 1) {
    if (R == 4) return launch_one<MODE, FLAV, 4, 1>;
    if (R == 8) return launch_one<MODE, FLAV, 8, 1>;
    return launch_one<MODE, FLAV, 16, 1>;
  }
  if (R == 4) return launch_one<MODE, FLAV, 4, 0>;
  if (R == 8) return launch_one<MODE, FLAV, 8, 0>;
  return launch_one<MODE, FLAV, 16, 0>;
}

// kind 0 (3 with NTFF samples): nblocks = a.nb_pml UPML workgroups + the interior workgroups;
// kind 1 (4 with samples): interior workgroups only; kind 2 (5 with samples): UPML workgroups only;
// kind 6 (7 with samples): wide UPML tiles (TM 12 wavefronts x 4 rows, TE 8 x 6);
// kind 10 (11 with samples): wide UPML tiles as 16 wavefronts x 3 rows (1024 threads)
int fast_launch(int mode, int flavour, int R, int kind, const FastArgs &a, int nblocks, hipStream_t st)
{
  fast_launcher f;
  if (mode == FDTD_MODE_TM) f = flavour == FDTD_FLAVOUR_MPI ? pick<FDTD_MODE_TM, FDTD_FLAVOUR_MPI>(R, kind)
                                                              : pick<FDTD_MODE_TM, FDTD_FLAVOUR_PLAIN>(R, kind);
  else f = flavour == FDTD_FLAVOUR_MPI ? pick<FDTD_MODE_TE, FDTD_FLAVOUR_MPI>(R, kind)
                                       : pick<FDTD_MODE_TE, FDTD_FLAVOUR_PLAIN>(R, kind);
  f(a, nblocks, st);
  return FDTD_OK;
}

int fast_export(const float2 *src, int pitch, int NPX, int NPY, double2 *dst, int NPYa, int off, hipStream_t st)
{
  dim3 blk(64, 4), grd((NPY + 63) / 64, (NPX + 3) / 4);
  hipLaunchKernelGGL(k_fast_export, grd, blk, 0, st, src, pitch, NPX, NPY, dst, NPYa, off);
  return FDTD_OK;
}

int fast_import(const double2 *src, int NPYa, int off, int NPX, int NPY, float2 *dst, int pitch, hipStream_t st)
{
  dim3 blk(64, 4), grd((NPY + 63) / 64, (NPX + 3) / 4);
  hipLaunchKernelGGL(k_fast_import, grd, blk, 0, st, src, NPYa, off, NPX, NPY, dst, pitch);
  return FDTD_OK;
}

int fast_inv_eps(const double *eps, int NPYa, int off, int NPX, int NPY, float *ie, int pitch, hipStream_t st)
{
  dim3 blk(64, 4), grd((NPY + 63) / 64, (NPX + 3) / 4);
  hipLaunchKernelGGL(k_fast_inv_eps, grd, blk, 0, st, eps, NPYa, off, NPX, NPY, ie, pitch);
  return FDTD_OK;
}

int fast_gather(const FastInst *inst, int cur, const Contour &ct, int off, int mode, double2 *cE, double2 *cH, hipStream_t st)
{
  if (ct.total <= 0) return FDTD_OK;
  hipLaunchKernelGGL(k_fast_gather, dim3((ct.total + 255) / 256), dim3(256), 0, st, inst, cur, ct, off, mode, cE, cH);
  return FDTD_OK;
}

// ---------------------------------------------------------------- FP32 instance state (host)

struct FastState {
  fdtd_grid_desc d;
  FastInst h;            // host copy of the device instance struct
  FastInst *dinst = nullptr;
  std::vector<void *> allocs;
  float *ftab = nullptr;
  // the 12 UPML auxiliary arrays (6 x ping-pong) in one block of 12 x auxn
  // cells, allocated when the coefficients fix the box without UPML cells
  float2 *auxbuf = nullptr;
  size_t auxn = 0;
  unsigned char *pmlflags = nullptr;
  std::vector<unsigned char> pmlx, pmly;
  float2 *exy = nullptr;
  // tiles: [0, n_int) interior (no UPML cell within the loaded block), [n_int, ntiles) UPML
  FastTile *dtiles = nullptr;
  int n_int = 0, n_pml = 0;
  int qint[9], qpml[9];
  // with contour samples each list is [tiles off the contour, tiles on it]: the
  // sampling instantiations run on the second part only
  int n_int_c = 0, n_pml_c = 0;
  int qint_p[9], qint_c[9], qpml_p[9], qpml_c[9];
  FastStep hsteps[FAST_MAX_T];
  int T = 4, R = 16, tilesT = -1, tilesR = -1;
  // rows of the UPML tiles (FAST_UPML_ROWS or FAST_UPML_ROWS_WIDE, by launch
  // shape: each UPML body covers one height) and the height the lists hold
  int upmlRows = FAST_UPML_ROWS, tilesU = -1;
  bool batch = false;  // an instance of a batch: its UPML bodies are the 32-row ones
  int cur = 0, pending = 0;
  int src_on = 0;
  double prev_c_re = 0, prev_c_im = 0;
  int steps_since_reset = 0;
  double last_angle = 1e300, last_k = 0;
  double t0[2] = {0, 0};
  double bytes_per_step = 0, cells_per_step = 0;
  double bytes_int_step = 0, bytes_pml_step = 0;  // algorithmic bytes of the two tile lists
  std::vector<FastTile> htp, hti;  // host copies of the UPML and interior tiles (batch builder)
  // per-step NTFF samples (null when the contour is not sampled): two buffers
  // used in turn, so that a block's samples are consumed (on the consumer's
  // stream, which records cfree) while the next block runs
  float2 *csamp = nullptr;
  float2 *csampb[2] = {nullptr, nullptr};
  hipEvent_t cfree[2] = {nullptr, nullptr};
  bool cused[2] = {false, false};
  int cbuf = 0;
  FastContour hct;
  fast_post_fn post = nullptr;
  void *post_ctx = nullptr;
  double times[FAST_MAX_T];
  // launch shape (FAST_LAUNCH_*); the second stream and its events exist once the
  // two-stream shape has been used
  int lmode = 0, lmode_last = 0;  // setting; shape of the last launch
  hipStream_t s2 = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

// Event timing of the fused launches.  Events are recorded around every launch
// and resolved when the totals are read, so timing adds no host-device
// synchronisation to the step loop.
struct ProfRec {
  hipEvent_t e0, e1;
  double bytes_int, bytes_pml, cells;
};
static struct {
  int on = 0;     // time every on-th launch (0: off)
  long tick = 0;  // launches since the enable
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;  // created once and reused: creating events per launch stalls the host
  double ms = 0, bytes_int = 0, bytes_pml = 0, cells = 0;
  long launches = 0;
} g_prof;

static int prof_event(hipEvent_t *e)
{
  if (!g_prof.pool.empty()) {
    *e = g_prof.pool.back();
    g_prof.pool.pop_back();
    return FDTD_OK;
  }
  HIPCHK(hipEventCreate(e));
  return FDTD_OK;
}

static void prof_resolve()
{
  for (ProfRec &r : g_prof.recs) {
    float a = 0;
    (void)hipEventSynchronize(r.e1);
    (void)hipEventElapsedTime(&a, r.e0, r.e1);
    g_prof.ms += a;
    g_prof.bytes_int += r.bytes_int;
    g_prof.bytes_pml += r.bytes_pml;
    g_prof.cells += r.cells;
    g_prof.launches += 1;
    g_prof.pool.push_back(r.e0);
    g_prof.pool.push_back(r.e1);
  }
  g_prof.recs.clear();
}

void fast_profile_enable(int on)
{
  prof_resolve();
  g_prof.on = on > 0 ? on : 0;
  g_prof.tick = 0;
  g_prof.ms = g_prof.bytes_int = g_prof.bytes_pml = g_prof.cells = 0;
  g_prof.launches = 0;
}

// out: ms of the fused launches, 0 (reserved), launches, algorithmic bytes of the
// interior and of the UPML tiles, cell updates
void fast_profile_read6(double out[6])
{
  prof_resolve();
  out[0] = g_prof.ms;
  out[1] = 0.0;
  out[2] = (double)g_prof.launches;
  out[3] = g_prof.bytes_int;
  out[4] = g_prof.bytes_pml;
  out[5] = g_prof.cells;
}

void fast_profile_read(double *ms, long *launches, double *bytes, double *cells)
{
  double o[6];
  fast_profile_read6(o);
  if (ms) *ms = o[0];
  if (launches) *launches = (long)o[2];
  if (bytes) *bytes = o[3] + o[4];
  if (cells) *cells = o[5];
}

// The box without UPML cells from the row and column flags: [axlo, axhi) the
// rows between the leading and trailing UPML rows, [aylo, ayhi) the same for the
// columns, empty in a direction whose UPML rows (columns) do not all lie at the
// grid's ends; acw the auxiliary cells of a row of the box's row range.
static void upml_box(FastState *s)
{
  const int NPX = s->h.NPX, NPY = s->h.NPY;
  int xlo = 0, xhi = NPX, ylo = 0, yhi = NPY;
  while (xlo < NPX && s->pmlx[xlo]) xlo++;
  while (xhi > xlo && s->pmlx[xhi - 1]) xhi--;
  while (ylo < NPY && s->pmly[ylo]) ylo++;
  while (yhi > ylo && s->pmly[yhi - 1]) yhi--;
  for (int x = xlo; x < xhi; x++)
    if (s->pmlx[x]) xhi = xlo;  // UPML cells inside: treat everything as UPML
  for (int y = ylo; y < yhi; y++)
    if (s->pmly[y]) yhi = ylo;
  s->h.axlo = xlo;
  s->h.axhi = xhi;
  s->h.aylo = ylo;
  s->h.ayhi = yhi;
  s->h.acw = ylo + NPY - yhi;
}

// cells of one auxiliary array (aux_cell's layout), rounded up to whole 64-cell lines
static size_t aux_cells(const FastInst &h)
{
  const size_t n = (size_t)(h.axlo + h.NPX - h.axhi) * h.pitch + (size_t)(h.axhi - h.axlo) * h.acw;
  return std::max<size_t>(64, (n + 63) / 64 * 64);
}

template <class T>
static int dev_alloc(FastState *s, T **p, size_t n)
{
  HIPCHK(dev_malloc((void **)p, std::max<size_t>(1, n) * sizeof(T)));
  HIPCHK(hipMemset(*p, 0, std::max<size_t>(1, n) * sizeof(T)));
  s->allocs.push_back((void *)*p);
  return FDTD_OK;
}

int fast_state_create(const fdtd_grid_desc &d, FastState **out)
{
  FastState *s = new FastState();
  *out = s;
  s->d = d;
  FastInst &h = s->h;
  memset(&h, 0, sizeof h);
  h.NPX = d.N_PX;
  h.NPY = d.N_PY;
  h.pitch = (d.N_PY + 63) / 64 * 64;
  h.flavour = d.flavour;
  h.c2 = (float)(1.0 / d.mu0);
  const size_t n = (size_t)h.NPX * h.pitch;
  // the kernels index a field with a 32-bit cell offset
  if (n >= ((size_t)1 << 28)) return fdtd_fail(FDTD_EINVAL, "grid of %zu cells exceeds the FP32 path's 2^28 limit", n);
  int rc;
  for (int b = 0; b < 2; b++) {
    if ((rc = dev_alloc(s, &h.f0[b], n))) return rc;
    if ((rc = dev_alloc(s, &h.f1[b], n))) return rc;
    if ((rc = dev_alloc(s, &h.f2[b], n))) return rc;
  }
  float *ie0, *ie1;
  if ((rc = dev_alloc(s, &ie0, n))) return rc;
  if ((rc = dev_alloc(s, &ie1, n))) return rc;
  h.ie0 = ie0;
  h.ie1 = ie1;
  if ((rc = dev_alloc(s, &s->ftab, 6 * (size_t)h.NPX + 5 * (size_t)h.NPY))) return rc;
  for (int q = 0; q < 5; q++) h.tx[q] = s->ftab + q * (size_t)h.NPX;
  h.tsx = s->ftab + 5 * (size_t)h.NPX;
  for (int q = 0; q < 4; q++) h.ty[q] = s->ftab + 6 * (size_t)h.NPX + q * (size_t)h.NPY;
  h.tsy = s->ftab + 6 * (size_t)h.NPX + 4 * (size_t)h.NPY;
  if ((rc = dev_alloc(s, &s->pmlflags, (size_t)h.NPX + h.NPY))) return rc;
  h.pmlx = s->pmlflags;
  h.pmly = s->pmlflags + h.NPX;
  if ((rc = dev_alloc(s, &s->exy, (size_t)h.NPX + h.NPY))) return rc;
  h.ex = s->exy;
  h.ey = s->exy + h.NPX;
  if ((rc = dev_alloc(s, &s->dinst, 1))) return rc;
  s->pmlx.assign(h.NPX, 1);
  s->pmly.assign(h.NPY, 1);
  upml_box(s);
  HIPCHK(hipMemcpy(s->dinst, &h, sizeof h, hipMemcpyHostToDevice));
  return FDTD_OK;
}

void fast_state_destroy(FastState *s)
{
  if (!s) return;
  if (s->s2) {
    (void)hipStreamSynchronize(s->s2);
    (void)hipStreamDestroy(s->s2);
    (void)hipEventDestroy(s->ev_in);
    (void)hipEventDestroy(s->ev_out);
  }
  for (void *p : s->allocs) (void)dev_free(p);
  if (s->auxbuf) (void)dev_free(s->auxbuf);
  if (s->dtiles) (void)dev_free(s->dtiles);
  for (int b = 0; b < 2; b++) {
    if (s->csampb[b]) (void)dev_free(s->csampb[b]);
    if (s->cfree[b]) (void)hipEventDestroy(s->cfree[b]);
  }
  delete s;
}

int fast_state_set_coefs(FastState *s, const fdtd_coef_tables *t)
{
  const int nx = s->h.NPX, ny = s->h.NPY;
  std::vector<float> f(6 * (size_t)nx + 5 * (size_t)ny);
  for (int q = 0; q < 5; q++)
    for (int x = 0; x < nx; x++) f[q * (size_t)nx + x] = (float)t->x[q][x];
  for (int x = 0; x < nx; x++) f[5 * (size_t)nx + x] = (float)t->sx[x];
  for (int q = 0; q < 4; q++)
    for (int y = 0; y < ny; y++) f[6 * (size_t)nx + q * (size_t)ny + y] = (float)t->y[q][y];
  for (int y = 0; y < ny; y++) f[6 * (size_t)nx + 4 * (size_t)ny + y] = (float)t->sy[y];
  HIPCHK(hipMemcpy(s->ftab, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
  // a cell runs the UPML recurrence when any coefficient of its row or column differs from
  // the sigma = 0 value (x tables at sigma 0: 1, 1, 1, 1, 1 ; y tables: 1, 1, 1, 1)
  for (int x = 0; x < nx; x++) {
    bool p = t->sx[x] != 0.0;
    for (int q = 0; q < 5; q++) p = p || t->x[q][x] != 1.0;
    s->pmlx[x] = p;
  }
  for (int y = 0; y < ny; y++) {
    bool p = t->sy[y] != 0.0;
    for (int q = 0; q < 4; q++) p = p || t->y[q][y] != 1.0;
    s->pmly[y] = p;
  }
  std::vector<unsigned char> fl(nx + ny);
  std::copy(s->pmlx.begin(), s->pmlx.end(), fl.begin());
  std::copy(s->pmly.begin(), s->pmly.end(), fl.begin() + nx);
  HIPCHK(hipMemcpy(s->pmlflags, fl.data(), fl.size(), hipMemcpyHostToDevice));
  // the auxiliaries' compact layout follows the box: a new box gets new, zeroed
  // arrays (the coefficients are set before the first step)
  const int box0[5] = {s->h.axlo, s->h.axhi, s->h.aylo, s->h.ayhi, s->h.acw};
  upml_box(s);
  const int box1[5] = {s->h.axlo, s->h.axhi, s->h.aylo, s->h.ayhi, s->h.acw};
  if (!s->auxbuf || memcmp(box0, box1, sizeof box0) != 0) {
    if (s->auxbuf) HIPCHK(dev_free(s->auxbuf));
    s->auxbuf = nullptr;
    s->auxn = aux_cells(s->h);
    HIPCHK(dev_malloc(&s->auxbuf, 12 * s->auxn * sizeof(float2)));
    HIPCHK(hipMemset(s->auxbuf, 0, 12 * s->auxn * sizeof(float2)));
    for (int b = 0; b < 2; b++)
      for (int q = 0; q < 6; q++) s->h.aux[q][b] = s->auxbuf + (size_t)(6 * b + q) * s->auxn;
  }
  HIPCHK(hipMemcpy(s->dinst, &s->h, sizeof s->h, hipMemcpyHostToDevice));
  s->tilesT = -1;  // rebuild tiles
  return FDTD_OK;
}

int fast_state_set_inv_eps(FastState *s, double *const *eps_dev, int NPYa, int off)
{
  fast_inv_eps(eps_dev[0], NPYa, off, s->h.NPX, s->h.NPY, (float *)s->h.ie0, s->h.pitch, 0);
  if (s->d.mode == FDTD_MODE_TE) fast_inv_eps(eps_dev[1], NPYa, off, s->h.NPX, s->h.NPY, (float *)s->h.ie1, s->h.pitch, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return FDTD_OK;
}

// Cover [lo, hi) of an axis of n cells with tile cores for blocks of B cells and
// T steps: a core needs T + 1 cells of halo below and T above, except at the
// grid's edges, whose far side is the zero ghost ring (or a fixed edge) and
// needs none: a core starting at 0 loads its block from 0 and keeps B - T
// cells; a block reaching n keeps every core cell up to n.  {core start,
// core length, block start} per segment.
struct Seg1 {
  int c0, n, b0;
};
static std::vector<Seg1> cover(int lo, int hi, int n, int B, int T)
{
  std::vector<Seg1> v;
  for (int c = lo; c < hi;) {
    const int b0 = c == 0 ? 0 : c - (T + 1);
    const int end = b0 + B >= n ? n : b0 + B - T;
    const int e = std::min(hi, end);
    v.push_back({c, e - c, b0});
    c = e;
  }
  return v;
}

static void tile_band(std::vector<FastTile> &v, int i_lo, int i_hi, int j_lo, int j_hi, int NPX, int NPY, int Brow,
                      int T)
{
  const std::vector<Seg1> rows = cover(i_lo, i_hi, NPX, Brow, T), cols = cover(j_lo, j_hi, NPY, 64, T);
  for (const Seg1 &r : rows)
    for (const Seg1 &c : cols) {
      FastTile t;
      t.inst = 0;
      t.i0 = r.c0;
      t.ni = r.n;
      t.bi = r.b0;
      t.j0 = c.c0;
      t.nj = c.n;
      t.bj = c.b0;
      v.push_back(t);
    }
}

static void split_queues(int *q, int base, int n)
{
  for (int k = 0; k <= 8; k++) q[k] = base + (int)((long)n * k / 8);
}

// vertical stripes of the interior tile list (see build_tiles)
#define FAST_STRIPES 4
static int build_tiles(FastState *s)
{
  const int T = s->T, R = s->R, H0 = T + 1;
  const int NPX = s->h.NPX, NPY = s->h.NPY;
  const int UR = s->upmlRows;  // rows of a UPML block
  const int CW = 64 - 2 * T - 1, CHi = 4 * R - 2 * T - 1, CHp = UR - 2 * T - 1;
  if (CW < 1 || CHi < 1 || CHp < 1) return fdtd_fail(FDTD_EINVAL, "temporal block depth %d too deep for R=%d", T, R);
  if (!s->auxbuf) return fdtd_fail(FDTD_ESTATE, "the FP32 solver's UPML coefficients are not set");
  // extent of the cells without UPML recurrence: the box of the auxiliaries' layout
  int xlo = s->h.axlo, xhi = s->h.axhi, ylo = s->h.aylo, yhi = s->h.ayhi;
  HIPCHK(hipMemcpy(s->dinst, &s->h, sizeof s->h, hipMemcpyHostToDevice));
  if (s->d.flavour == FDTD_FLAVOUR_PLAIN) {
    // the plain solvers never update their edge cells (fdtdTM_upml.c:158-159): keep
    // them out of the interior tiles' dependency cones, which the interior kernel
    // updates without masks
    xlo = std::max(xlo, 1);
    xhi = std::min(xhi, NPX - 1);
    ylo = std::max(ylo, 1);
    yhi = std::min(yhi, NPY - 1);
  }
  // Interior tile cores [I_lo, I_hi) x [J_lo, J_hi): with the block loaded from
  // T+1 rows below to T rows above the core, every cell whose update a core
  // cell depends on lies in [xlo+1, xhi-1] x [ylo+1, yhi-1].
  int I_lo = xlo + H0, I_hi = xhi - T, J_lo = ylo + H0, J_hi = yhi - T;
  // An interior extent that is not a whole number of tile cores ends in a
  // partial row (column) of tiles, which costs a full one.  The frame's tiles
  // often have room to spare (a 43-row end band in two 48-row UPML blocks keeps
  // up to 65 rows), so the remainder moves into the frame wherever its tile
  // count stays the same (8256 x 4160, T = 10: 138 rows of 80-row interior
  // tiles instead of 139, the last of which kept 29 of its 59 core rows).
  auto trim = [&](int &lo, int &hi, int n, int core, int Bf) {
    const int span = hi - lo, rem = span % core;
    if (span <= core || rem == 0) return;
    const size_t f0 = cover(0, lo, n, Bf, T).size(), f1 = cover(hi, n, n, Bf, T).size();
    for (int a = 0; a <= rem; a++)
      if (cover(0, lo + a, n, Bf, T).size() == f0 && cover(hi - (rem - a), n, n, Bf, T).size() == f1) {
        lo += a;
        hi -= rem - a;
        return;
      }
  };
  if (I_lo < I_hi && J_lo < J_hi) {
    trim(I_lo, I_hi, NPX, CHi, UR);
    trim(J_lo, J_hi, NPY, CW, 64);
  }
  std::vector<FastTile> ti, tp;
  if (I_lo < I_hi && J_lo < J_hi) {
    tile_band(ti, I_lo, I_hi, J_lo, J_hi, NPX, NPY, 4 * R, T);
    tile_band(tp, 0, I_lo, 0, NPY, NPX, NPY, UR, T);
    tile_band(tp, I_hi, NPX, 0, NPY, NPX, NPY, UR, T);
    tile_band(tp, I_lo, I_hi, 0, J_lo, NPX, NPY, UR, T);
    tile_band(tp, I_lo, I_hi, J_hi, NPY, NPX, NPY, UR, T);
  } else {
    tile_band(tp, 0, NPX, 0, NPY, NPX, NPY, UR, T);
  }
  auto byrow = [](const FastTile &a, const FastTile &b) { return a.i0 != b.i0 ? a.i0 < b.i0 : a.j0 < b.j0; };
  std::sort(tp.begin(), tp.end(), byrow);
  // interior tiles: FAST_STRIPES vertical stripes of tile columns, each walked
  // row by row, so that the eighth of the list an XCD takes is (half) a stripe
  // and the tiles it runs at a time are neighbours in both directions, their
  // halos overlapping in its L2 (8256 x 4160 grid, 95 tile columns: 4 stripes
  // of 24 columns 1 % faster than 8 of 12 and than 2 of 48, 2 % faster than
  // rows across the whole grid, 4 % faster than 24 stripes of 4)
  {
    std::vector<int> cols;
    for (const FastTile &t : ti) cols.push_back(t.j0);
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    const int nc = (int)cols.size();
    auto stripe = [&](const FastTile &t) {
      const int c = (int)(std::lower_bound(cols.begin(), cols.end(), t.j0) - cols.begin());
      return (int)((long)c * FAST_STRIPES / std::max(1, nc));
    };
    std::sort(ti.begin(), ti.end(), [&](const FastTile &a, const FastTile &b) {
      const int sa = stripe(a), sb = stripe(b);
      return sa != sb ? sa < sb : byrow(a, b);
    });
  }
  // tiles whose core holds a contour cell the kernel samples (a side's line fi or
  // its neighbour line fi - 1, within the side's range) go last in their list
  auto on_contour = [&](const FastTile &t) {
    if (!s->csamp) return false;
    for (int q = 0; q < 4; q++) {
      const FastCtrSide &g = s->hct.s[q];
      if (g.n <= 0) continue;
      const int a0 = g.horiz ? t.j0 : t.i0, a1 = a0 + (g.horiz ? t.nj : t.ni);  // across the side
      const int b0 = g.horiz ? t.i0 : t.j0, b1 = b0 + (g.horiz ? t.ni : t.nj);  // along it
      const bool line = (g.fi >= a0 && g.fi < a1) || (g.fi - 1 >= a0 && g.fi - 1 < a1);
      if (line && g.lo < b1 && g.lo + g.n > b0) return true;
    }
    return false;
  };
  auto off_first = [&](std::vector<FastTile> &v) {
    std::stable_partition(v.begin(), v.end(), [&](const FastTile &t) { return !on_contour(t); });
    int c = 0;
    for (const FastTile &t : v) c += on_contour(t) ? 1 : 0;
    return c;
  };
  s->n_int_c = off_first(ti);
  s->n_pml_c = off_first(tp);
  s->htp = tp;
  s->hti = ti;
  s->n_int = (int)ti.size();
  s->n_pml = (int)tp.size();
  std::vector<FastTile> all(ti);
  all.insert(all.end(), tp.begin(), tp.end());
  if (s->dtiles) (void)dev_free(s->dtiles);
  s->dtiles = nullptr;
  HIPCHK(dev_malloc(&s->dtiles, std::max<size_t>(1, all.size()) * sizeof(FastTile)));
  if (!all.empty()) HIPCHK(hipMemcpy(s->dtiles, all.data(), all.size() * sizeof(FastTile), hipMemcpyHostToDevice));
  split_queues(s->qint, 0, s->n_int);
  split_queues(s->qpml, s->n_int, s->n_pml);
  split_queues(s->qint_p, 0, s->n_int - s->n_int_c);
  split_queues(s->qint_c, s->n_int - s->n_int_c, s->n_int_c);
  split_queues(s->qpml_p, s->n_int, s->n_pml - s->n_pml_c);
  split_queues(s->qpml_c, s->n_int + s->n_pml - s->n_pml_c, s->n_pml_c);
  // algorithmic bytes per step (4-byte reals): complex E,H read+write everywhere, UPML
  // auxiliaries read+write in UPML cells, 1/eps read (TM 1, TE 2)
  // updated cells (the plain solvers skip their edge cells) and how many of them
  // are UPML cells: a cell is UPML when its row or its column is
  const bool plain = s->d.flavour == FDTD_FLAVOUR_PLAIN;
  const int x0 = plain ? 1 : 0, x1 = plain ? NPX - 1 : NPX, y0 = plain ? 1 : 0, y1 = plain ? NPY - 1 : NPY;
  double nx_in = 0, ny_in = 0;
  for (int x = x0; x < x1; x++) nx_in += s->pmlx[x] ? 0 : 1;
  for (int y = y0; y < y1; y++) ny_in += s->pmly[y] ? 0 : 1;
  const double ncell = (double)std::max(0, x1 - x0) * std::max(0, y1 - y0);
  const double npml = ncell - nx_in * ny_in;
  const double neps = s->d.mode == FDTD_MODE_TM ? 1 : 2;
  s->bytes_per_step = 4.0 * 2 * 2 * (3 * ncell + 6 * npml) + 4.0 * neps * ncell;
  // split by tile list: the interior tiles hold no UPML cell
  double cint = 0;
  for (const FastTile &t : ti) cint += (double)t.ni * t.nj;
  s->bytes_int_step = (4.0 * 2 * 2 * 3 + 4.0 * neps) * cint;
  s->bytes_pml_step = s->bytes_per_step - s->bytes_int_step;
  s->cells_per_step = ncell;
  s->tilesT = T;
  s->tilesR = R;
  s->tilesU = UR;
  return FDTD_OK;
}

int fast_state_set_blocking(FastState *s, int T, int R)
{
  if (T < 1 || T >= FAST_MAX_T) return fdtd_fail(FDTD_EINVAL, "T=%d out of range", T);
  if (R != 4 && R != 8 && R != 16 && R != 18 && R != 20) return fdtd_fail(FDTD_EINVAL, "R=%d not in {4,8,16,18,20}", R);
  int rc;
  if ((rc = fast_state_flush(s))) return rc;
  s->T = T;
  s->R = R;  // the tiles are rebuilt at the next launch when T or R changed
  return FDTD_OK;
}

void fast_state_get_blocking(const FastState *s, int *T, int *R)
{
  if (T) *T = s->T;
  if (R) *R = s->R;
}

// wave tables e^{i x k cos}, e^{i y k sin} and pulse constants for the current angle
static int update_wave(FastState *s, const fdtd_step_params &p)
{
  if (p.angle_deg == s->last_angle && p.k == s->last_k) return FDTD_OK;
  const fdtd_grid_desc &d = s->d;
  const double rad = p.angle_deg * M_PI / 180.0;
  const double ks_cos = cos(rad) * p.k, ks_sin = sin(rad) * p.k;
  std::vector<float2> t(s->h.NPX + s->h.NPY);
  for (int x = 0; x < s->h.NPX; x++) t[x] = make_float2((float)cos(x * ks_cos), (float)sin(x * ks_cos));
  for (int y = 0; y < s->h.NPY; y++) t[s->h.NPX + y] = make_float2((float)cos(y * ks_sin), (float)sin(y * ks_sin));
  HIPCHK(hipMemcpy(s->exy, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice));
  FastInst &h = s->h;
  const double cpc = cos(rad) / d.c0, spc = sin(rad) / d.c0;
  h.cpc = (float)cpc;
  h.spc = (float)spc;
  h.w = (float)p.omega;
  const double gx[2] = {d.mode == FDTD_MODE_TM ? 0.0 : 0.5, 0.0};
  const double gy[2] = {0.0, d.mode == FDTD_MODE_TM ? 0.0 : 0.5};
  for (int c = 0; c < 2; c++) {
    h.gx[c] = (float)gx[c];
    h.gy[c] = (float)gy[c];
    const double center_peak = (d.N_PX / 2.0 + gx[c]) * cpc + (d.N_PY / 2 + gy[c]) * spc;
    s->t0[c] = -center_peak + 500;  // field.c:236-241
  }
  if (d.mode == FDTD_MODE_TM) {
    h.dot[0] = 1.0f;
    h.on[0] = 1;
    h.on[1] = 0;
  } else {
    h.dot[0] = (float)p.dot_x;
    h.dot[1] = (float)p.dot_y;
    h.on[0] = p.use_x;
    h.on[1] = p.use_y;
  }
  HIPCHK(hipMemcpy(s->dinst, &h, sizeof h, hipMemcpyHostToDevice));
  s->last_angle = p.angle_deg;
  s->last_k = p.k;
  return FDTD_OK;
}

int fast_state_flush(FastState *s)
{
  if (s->pending == 0) return FDTD_OK;
  int rc;
  const int Tsave = s->T;
  // the UPML bodies of the *_PML8, *_W8 and *_P16 shapes cover the wide tiles
  {
    const int lm = s->lmode;
    const bool wide = !s->batch && (lm == FAST_LAUNCH_SPLIT_PML8 || lm == FAST_LAUNCH_TWO_STREAMS_PML8 ||
                                    lm == FAST_LAUNCH_SPLIT_W8 || lm == FAST_LAUNCH_TWO_STREAMS_W8 ||
                                    lm == FAST_LAUNCH_SPLIT_P16 || lm == FAST_LAUNCH_TWO_STREAMS_P16);
    s->upmlRows = wide ? FAST_UPML_ROWS_WIDE : FAST_UPML_ROWS;
    // R = 18 and 20 (72- and 80-row interior tiles) exist only as 8 wavefronts of 9 and 10 rows
    const bool i8 = lm == FAST_LAUNCH_SPLIT_W8 || lm == FAST_LAUNCH_TWO_STREAMS_W8 || lm == FAST_LAUNCH_SPLIT_P16 ||
                    lm == FAST_LAUNCH_TWO_STREAMS_P16;
    if ((s->R == 18 || s->R == 20) && !i8)
      return fdtd_fail(FDTD_EINVAL, "R=%d needs a launch shape with 8-wavefront interior tiles (*_W8, *_P16), not %d",
                       s->R, lm);
  }
  // a partial block runs with T = pending; the tiles of the full depth stay valid
  // for any smaller depth, so only the block depth passed to the kernel changes
  if (s->tilesT != s->T || s->tilesR != s->R || s->tilesU != s->upmlRows)
    if ((rc = build_tiles(s))) return rc;
  FastArgs a;
  memset(&a, 0, sizeof a);
  a.inst = s->dinst;
  a.T = s->pending;
  a.src = s->src_on;
  a.cur = s->cur;
  memcpy(a.st, s->hsteps, sizeof a.st);
  const int sb = s->cbuf;
  if (s->csamp) {
    // this block's sample buffer was last read by the consumer two blocks ago
    if (s->cused[sb]) HIPCHK(hipStreamWaitEvent(0, s->cfree[sb], 0));
    a.csamp = s->csampb[sb];
    a.ct = s->hct;
  }
  ProfRec pr;
  const bool timed = g_prof.on && (g_prof.tick++ % g_prof.on) == 0;
  if (timed) {
    if ((rc = prof_event(&pr.e0)) || (rc = prof_event(&pr.e1))) return rc;
    HIPCHK(hipEventRecord(pr.e0, 0));
  }
  // one launch: the UPML tiles' workgroups first, then the interior tiles', each
  // split into 8 XCD partitions of ceil(n/8) workgroups
  a.tiles = s->dtiles;
  memcpy(a.qb[0], s->qpml, sizeof a.qb[0]);
  memcpy(a.qb[1], s->qint, sizeof a.qb[1]);
  const int nbp = 8 * ((s->n_pml + 7) / 8), nbi = 8 * ((s->n_int + 7) / 8);
  int mode = s->lmode;
  if (mode == FAST_LAUNCH_AUTO)
    // large grid: interior and UPML tiles as two launches, each at its own
    // occupancy (the interior body's register budget is below the UPML body's)
    mode = (nbi >= 8 * nbp && nbi > 4096) ? FAST_LAUNCH_SPLIT : FAST_LAUNCH_MERGED;
  s->lmode_last = mode;
  // the interior (UPML) tiles: those off the contour with the plain kind, those
  // on it with the sampling kind (without samples, all of a list is off it)
  // UPML tiles as 8 wavefronts of 4 rows (512 threads) in the *_PML8 shapes
  const bool w8 = mode == FAST_LAUNCH_SPLIT_PML8 || mode == FAST_LAUNCH_TWO_STREAMS_PML8 ||
                  mode == FAST_LAUNCH_SPLIT_W8 || mode == FAST_LAUNCH_TWO_STREAMS_W8;
  // or, in the *_P16 shapes, as 16 wavefronts of 2 rows (1024 threads)
  const bool p16 = mode == FAST_LAUNCH_SPLIT_P16 || mode == FAST_LAUNCH_TWO_STREAMS_P16;
  // and in the *_W8 and *_P16 shapes the interior tiles as 8 wavefronts of R/2 rows (R >= 8)
  const bool i8 = (mode == FAST_LAUNCH_SPLIT_W8 || mode == FAST_LAUNCH_TWO_STREAMS_W8 || p16) && s->R >= 8;
  auto launch_list = [&](bool pml, hipStream_t st) {
    const int nt = pml ? s->n_pml : s->n_int, nc = pml ? s->n_pml_c : s->n_int_c;
    FastArgs b = a;
    b.nb_pml = 0;
    if (pml && nc > 0) {
      // UPML tiles with contour samples: all of them in one launch of the sampling
      // kind (a tile without contour cells samples nothing).  The UPML tiles are
      // few and long, one workgroup per compute unit, so two launches would each
      // end on a part-filled dispatch round; the interior tiles, most of which
      // hold no contour cell, keep the plain kind for those
      memcpy(b.qb[0], s->qpml, sizeof b.qb[0]);
      fast_launch(s->d.mode, s->d.flavour, s->R, p16 ? 11 : w8 ? 7 : 5, b, 8 * ((nt + 7) / 8), st);
      return;
    }
    if (nt - nc > 0) {
      memcpy(pml ? b.qb[0] : b.qb[1], pml ? s->qpml_p : s->qint_p, sizeof b.qb[0]);
      fast_launch(s->d.mode, s->d.flavour, s->R, pml ? (p16 ? 10 : w8 ? 6 : 2) : (i8 ? 8 : 1), b, 8 * ((nt - nc + 7) / 8), st);
    }
    if (nc > 0) {
      memcpy(pml ? b.qb[0] : b.qb[1], pml ? s->qpml_c : s->qint_c, sizeof b.qb[0]);
      fast_launch(s->d.mode, s->d.flavour, s->R, pml ? (p16 ? 11 : w8 ? 7 : 5) : (i8 ? 9 : 4), b, 8 * ((nc + 7) / 8), st);
    }
  };
  const int km = s->csamp ? 3 : 0;
  if (mode == FAST_LAUNCH_SPLIT || mode == FAST_LAUNCH_SPLIT_PML8 || mode == FAST_LAUNCH_SPLIT_W8 ||
      mode == FAST_LAUNCH_SPLIT_P16) {
    launch_list(false, 0);
    launch_list(true, 0);
  } else if (mode == FAST_LAUNCH_TWO_STREAMS || mode == FAST_LAUNCH_TWO_STREAMS_PML8 ||
             mode == FAST_LAUNCH_TWO_STREAMS_W8 || mode == FAST_LAUNCH_TWO_STREAMS_P16) {
    // UPML tiles on a second stream beside the interior tiles: both read the same
    // buffer and write disjoint cores, so the two launches may overlap, each at its
    // own occupancy; events order them after the previous block and before the next
    if (!s->s2) {
      HIPCHK(hipStreamCreateWithFlags(&s->s2, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&s->ev_in, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming));
    }
    HIPCHK(hipEventRecord(s->ev_in, 0));
    HIPCHK(hipStreamWaitEvent(s->s2, s->ev_in, 0));
    launch_list(true, s->s2);
    launch_list(false, 0);
    HIPCHK(hipEventRecord(s->ev_out, s->s2));
    HIPCHK(hipStreamWaitEvent(0, s->ev_out, 0));
  } else {
    a.nb_pml = nbp;
    if (nbp + nbi > 0) fast_launch(s->d.mode, s->d.flavour, s->R, km, a, nbp + nbi, 0);
  }
  HIPCHK(hipGetLastError());
  if (timed) {
    HIPCHK(hipEventRecord(pr.e1, 0));
    pr.bytes_int = s->bytes_int_step * s->pending;
    pr.bytes_pml = s->bytes_pml_step * s->pending;
    pr.cells = s->cells_per_step * s->pending;
    g_prof.recs.push_back(pr);
    if (g_prof.recs.size() > 4096) prof_resolve();
  }
  s->cur ^= 1;
  const int n = s->pending;
  s->pending = 0;
  s->T = Tsave;
  if (s->csamp && s->post) {
    s->cused[sb] = true;
    s->cbuf ^= 1;
    return s->post(s->post_ctx, s->csampb[sb], n, s->times, s->cfree[sb]);
  }
  return FDTD_OK;
}

int fast_state_set_launch_mode(FastState *s, int mode)
{
  if (mode < FAST_LAUNCH_AUTO || mode > FAST_LAUNCH_TWO_STREAMS_P16) return fdtd_fail(FDTD_EINVAL, "bad launch mode %d", mode);
  int rc = fast_state_flush(s);
  if (rc) return rc;
  s->lmode = mode;
  return FDTD_OK;
}

// the setting, or with AUTO the shape of the last launch (AUTO before any)
int fast_state_get_launch_mode(const FastState *s) { return s->lmode != FAST_LAUNCH_AUTO ? s->lmode : s->lmode_last; }

int fast_state_step(FastState *s, const fdtd_step_params *p, int *launched)
{
  int rc;
  if ((rc = update_wave(s, *p))) return rc;
  FastStep &st = s->hsteps[s->pending];
  memset(&st, 0, sizeof st);
  s->times[s->pending] = p->time;
  // MPI scattered wave: ray_coef * e^{-i w t}, w*t formed as the reference does
  const double wt = p->omega * p->time;
  const double cre = p->ray_coef * cos(wt), cim = -p->ray_coef * sin(wt);
  st.c = make_float2((float)cre, (float)cim);
  st.dc = make_float2((float)(cre - s->prev_c_re), (float)(cim - s->prev_c_im));
  s->prev_c_re = cre;
  s->prev_c_im = cim;
  // plain Gaussian pulse delays
  st.tn = (float)(p->time - s->t0[0]);
  st.tn1 = (float)(p->time - s->t0[1]);
  st.tp = (float)((p->time - 1) - s->t0[0]);
  st.tp1 = (float)((p->time - 1) - s->t0[1]);
  st.hasprev = s->steps_since_reset > 0 ? 1.0f : 0.0f;
  s->src_on = 1;
  s->steps_since_reset++;
  s->pending++;
  if (launched) *launched = 0;
  if (s->pending >= s->T) {
    if ((rc = fast_state_flush(s))) return rc;
    if (launched) *launched = 1;
  }
  return FDTD_OK;
}

// slot numbering of the FP64 grid: TM Ez Jz Dz Hx Mx Bx Hy My By / TE Ex Jx Dx Ey Jy Dy Hz Mz Bz
static float2 *slot_ptr(FastState *s, int slot, int buf)
{
  FastInst &h = s->h;
  switch (slot) {
  case 0: return h.f0[buf];
  case 3: return s->d.mode == FDTD_MODE_TM ? h.f1[buf] : h.f1[buf];
  case 6: return h.f2[buf];
  case 1: return h.aux[0][buf];
  case 2: return h.aux[1][buf];
  case 4: return h.aux[2][buf];
  case 5: return h.aux[3][buf];
  case 7: return h.aux[4][buf];
  case 8: return h.aux[5][buf];
  default: return nullptr;
  }
}
static bool aux_slot(int slot) { return slot == 1 || slot == 2 || slot == 4 || slot == 5 || slot == 7 || slot == 8; }

int fast_state_export(FastState *s, int slot, double2 *dst, int NPYa, int off)
{
  int rc;
  if ((rc = fast_state_flush(s))) return rc;
  if (slot < 0 || slot > 8) return fdtd_fail(FDTD_EINVAL, "bad slot %d", slot);
  if (aux_slot(slot)) {
    dim3 blk(64, 4), grd((s->h.NPY + 63) / 64, (s->h.NPX + 3) / 4);
    hipLaunchKernelGGL(k_fast_aux_export, grd, blk, 0, 0, s->auxbuf ? slot_ptr(s, slot, s->cur) : nullptr, s->h, dst,
                       NPYa, off);
  } else {
    fast_export(slot_ptr(s, slot, s->cur), s->h.pitch, s->h.NPX, s->h.NPY, dst, NPYa, off, 0);
  }
  HIPCHK(hipGetLastError());
  return FDTD_OK;
}

int fast_state_import(FastState *s, int slot, const double2 *src, int NPYa, int off)
{
  int rc;
  if ((rc = fast_state_flush(s))) return rc;
  if (slot < 0 || slot > 8) return fdtd_fail(FDTD_EINVAL, "bad slot %d", slot);
  if (aux_slot(slot)) {
    if (!s->auxbuf) return fdtd_fail(FDTD_ESTATE, "the FP32 solver's UPML coefficients are not set");
    dim3 blk(64, 4), grd((s->h.NPY + 63) / 64, (s->h.NPX + 3) / 4);
    hipLaunchKernelGGL(k_fast_aux_import, grd, blk, 0, 0, src, NPYa, off, s->h, slot_ptr(s, slot, s->cur));
  } else {
    fast_import(src, NPYa, off, s->h.NPX, s->h.NPY, slot_ptr(s, slot, s->cur), s->h.pitch, 0);
  }
  HIPCHK(hipGetLastError());
  return FDTD_OK;
}

int fast_state_reset(FastState *s)
{
  s->pending = 0;
  const size_t n = (size_t)s->h.NPX * s->h.pitch * sizeof(float2);
  for (int b = 0; b < 2; b++) {
    HIPCHK(hipMemset(s->h.f0[b], 0, n));
    HIPCHK(hipMemset(s->h.f1[b], 0, n));
    HIPCHK(hipMemset(s->h.f2[b], 0, n));
  }
  if (s->auxbuf) HIPCHK(hipMemset(s->auxbuf, 0, 12 * s->auxn * sizeof(float2)));
  s->cur = 0;
  s->prev_c_re = s->prev_c_im = 0;
  s->steps_since_reset = 0;
  HIPCHK(hipDeviceSynchronize());
  return FDTD_OK;
}

int fast_state_fields(FastState *s, const float2 *f[3], size_t *n)
{
  int rc;
  if ((rc = fast_state_flush(s))) return rc;
  f[0] = s->h.f0[s->cur];
  f[1] = s->h.f1[s->cur];
  f[2] = s->h.f2[s->cur];
  *n = (size_t)s->h.NPX * s->h.pitch;
  return FDTD_OK;
}

int fast_state_pending(const FastState *s) { return s->pending; }

int fast_state_pitch(const FastState *s, int *pitch)
{
  *pitch = s->h.pitch;
  return FDTD_OK;
}

int fast_state_gather(FastState *s, const Contour &ct, int off, double2 *cE, double2 *cH)
{
  int rc;
  if ((rc = fast_state_flush(s))) return rc;
  fast_gather(s->dinst, s->cur, ct, off, s->d.mode, cE, cH, 0);
  HIPCHK(hipGetLastError());
  return FDTD_OK;
}

// ---------------------------------------------------------------- batched instances
//
// Several FP32 MPI-flavour instances of uneven size advance together: per T-step
// block and mode, one k_fast_batch launch pulls the interior tiles of every
// unfinished instance from a device queue and a second the UPML tiles (each at
// its own register budget: two workgroups per compute unit for the interior
// body at R = 16, one for the UPML body), and
// k_batch_advance moves the per-instance progress records on the device.  No
// host value changes between blocks, so the loop of blocks is captured once as
// a hipGraph and replayed.

struct FastBatch {
  int n = 0, T = 4, R = 16;
  std::vector<FastState *> st;
  std::vector<double> omega;
  std::vector<int> steps;
  FastInst *dinst = nullptr;
  FastBatchInst *dbinst = nullptr;
  FastStep *dsteps = nullptr;       // [instance][FAST_MAX_T]
  FastTile *dtiles = nullptr;
  int *dqctr = nullptr;             // [2 modes][2 queues][8 XCD parts]
  int nt[2][2] = {{0, 0}, {0, 0}};  // tiles per mode and queue
  int off[2][2] = {{0, 0}, {0, 0}};
  int nblk[2][2] = {{0, 0}, {0, 0}};  // workgroups per mode and queue
  hipStream_t stream = nullptr;
  // two_streams: the UPML launches of a block on s2 beside the interior launches
  // (fork and join by events, captured into the graph as dependencies);
  // pml8 / inner8: the UPML / interior tiles as 8 wavefronts (the *_PML8, *_W8 shapes)
  // Default: TWO_STREAMS_PML8, the fastest on the 64-instance sweep (260.7 G cell
  // updates/s against 243.5 for SPLIT; every shape gives the same bits)
  int two_streams = 1, pml8 = 1, inner8 = 0;
  hipStream_t s2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int glevels = 0;
};

template <int MODE, int KQ>
static void batch_launch_q(const FastArgs &a, int R, int nblk, bool w8, hipStream_t st)
{
  if (w8 && KQ == 0) hipLaunchKernelGGL((k_fast_batch<MODE, 8, 0, 1>), dim3(nblk), dim3(512), 0, st, a);
  else if (w8 && R == 8) hipLaunchKernelGGL((k_fast_batch<MODE, 8, KQ, 1>), dim3(nblk), dim3(512), 0, st, a);
  else if (w8 && R == 16) hipLaunchKernelGGL((k_fast_batch<MODE, 16, KQ, 1>), dim3(nblk), dim3(512), 0, st, a);
  else if (R == 4) hipLaunchKernelGGL((k_fast_batch<MODE, 4, KQ>), dim3(nblk), dim3(256), 0, st, a);
  else if (R == 8) hipLaunchKernelGGL((k_fast_batch<MODE, 8, KQ>), dim3(nblk), dim3(256), 0, st, a);
  else hipLaunchKernelGGL((k_fast_batch<MODE, 16, KQ>), dim3(nblk), dim3(256), 0, st, a);
}

template <int MODE>
static void batch_launch_mode(const FastBatch *b, const FastArgs &a, const int nblk[2], hipStream_t st,
                              hipStream_t st_pml)
{
  if (nblk[1] > 0) batch_launch_q<MODE, 1>(a, b->R, nblk[1], b->inner8 != 0, st);
  if (nblk[0] > 0) batch_launch_q<MODE, 0>(a, b->R, nblk[0], b->pml8 != 0, st_pml);
}

static FastArgs batch_args(const FastBatch *b, int m)
{
  FastArgs a;
  memset(&a, 0, sizeof a);
  a.inst = b->dinst;
  a.tiles = b->dtiles;
  a.T = b->T;
  a.src = 1;
  a.binst = b->dbinst;
  a.bsteps = b->dsteps;
  a.qctr = b->dqctr + 16 * m;
  a.qn[0] = b->nt[m][0];
  a.qn[1] = b->nt[m][1];
  a.qoff[0] = b->off[m][0];
  a.qoff[1] = b->off[m][1];
  return a;
}

// one block of T steps for every unfinished instance, then the progress update
static int batch_level(FastBatch *b)
{
  hipStream_t sp = b->stream;
  if (b->two_streams) {
    HIPCHK(hipEventRecord(b->ev_fork, b->stream));
    HIPCHK(hipStreamWaitEvent(b->s2, b->ev_fork, 0));
    sp = b->s2;
  }
  for (int m = 0; m < 2; m++) {
    if (b->nblk[m][0] + b->nblk[m][1] == 0) continue;
    const FastArgs a = batch_args(b, m);
    if (m == FDTD_MODE_TM) batch_launch_mode<FDTD_MODE_TM>(b, a, b->nblk[m], b->stream, sp);
    else batch_launch_mode<FDTD_MODE_TE>(b, a, b->nblk[m], b->stream, sp);
  }
  if (b->two_streams) {
    HIPCHK(hipEventRecord(b->ev_join, b->s2));
    HIPCHK(hipStreamWaitEvent(b->stream, b->ev_join, 0));
  }
  hipLaunchKernelGGL(k_batch_advance, dim3((std::max(b->n, 4) + 255) / 256), dim3(256), 0, b->stream, b->dbinst, b->n,
                     b->T, b->dqctr, 32, 0, b->dsteps);
  HIPCHK(hipGetLastError());
  return FDTD_OK;
}

void fast_batch_destroy(FastBatch *b)
{
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->dinst) (void)dev_free(b->dinst);
  if (b->dbinst) (void)dev_free(b->dbinst);
  if (b->dsteps) (void)dev_free(b->dsteps);
  if (b->dtiles) (void)dev_free(b->dtiles);
  if (b->dqctr) (void)dev_free(b->dqctr);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->s2) {
    (void)hipStreamSynchronize(b->s2);
    (void)hipStreamDestroy(b->s2);
    (void)hipEventDestroy(b->ev_fork);
    (void)hipEventDestroy(b->ev_join);
  }
  delete b;
}

int fast_batch_create(FastState *const *st, const fdtd_step_params *p0, const int *steps, int n, int T, int R,
                      FastBatch **out)
{
  *out = nullptr;
  if (n < 1) return fdtd_fail(FDTD_EINVAL, "empty batch");
  if (T < 1 || T >= FAST_MAX_T) return fdtd_fail(FDTD_EINVAL, "T=%d out of range", T);
  if (R != 4 && R != 8 && R != 16) return fdtd_fail(FDTD_EINVAL, "R=%d not in {4,8,16}", R);
  FastBatch *b = new FastBatch();
  b->n = n;
  b->T = T;
  b->R = R;
  int rc;
  std::vector<FastInst> hinst(n);
  std::vector<FastBatchInst> hb(n);
  // instance order of the queues: largest first (longest-processing-time order)
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return (long)st[x]->h.NPX * st[x]->h.NPY > (long)st[y]->h.NPX * st[y]->h.NPY;
  });
  for (int i = 0; i < n; i++) {
    FastState *s = st[i];
    if (s->d.flavour != FDTD_FLAVOUR_MPI) {
      fast_batch_destroy(b);
      return fdtd_fail(FDTD_EINVAL, "batched instances use the MPI (CW source) solvers; instance %d does not", i);
    }
    if (steps[i] < 0) {
      fast_batch_destroy(b);
      return fdtd_fail(FDTD_EINVAL, "instance %d: negative step count", i);
    }
    // the source tables of the instance's wave (angle, k), as its first step would set them
    if ((rc = fast_state_flush(s)) || (rc = fast_state_set_blocking(s, T, R)) || (rc = update_wave(s, p0[i]))) {
      fast_batch_destroy(b);
      return rc;
    }
    s->batch = true;
    s->upmlRows = FAST_UPML_ROWS;
    if (s->tilesT != s->T || s->tilesR != s->R || s->tilesU != s->upmlRows) {
      if ((rc = build_tiles(s))) {
        fast_batch_destroy(b);
        return rc;
      }
    }
    b->st.push_back(s);
    b->omega.push_back(p0[i].omega);
    b->steps.push_back(steps[i]);
    hinst[i] = s->h;
    hb[i].omega = p0[i].omega;
    hb[i].n = s->steps_since_reset;
    hb[i].steps = steps[i];
    hb[i].Ti = 0;
    hb[i].cur = s->cur;
  }
  // tiles: per mode, the UPML queue then the interior queue, instances largest first
  std::vector<FastTile> all;
  for (int m = 0; m < 2; m++) {
    for (int k = 0; k < 2; k++) {
      b->off[m][k] = (int)all.size();
      for (int i : order) {
        FastState *s = st[i];
        if (s->d.mode != m) continue;
        for (FastTile t : (k == 0 ? s->htp : s->hti)) {
          t.inst = i;
          all.push_back(t);
        }
      }
      b->nt[m][k] = (int)all.size() - b->off[m][k];
    }
    // resident workgroups: at most one per compute unit for the UPML bodies,
    // two for the interior body at R = 16
    b->nblk[m][0] = std::min(b->nt[m][0], 256);
    b->nblk[m][1] = std::min(b->nt[m][1], 512);
  }
  HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&b->s2, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
  HIPCHK(dev_malloc(&b->dinst, n * sizeof(FastInst)));
  HIPCHK(dev_malloc(&b->dbinst, n * sizeof(FastBatchInst)));
  HIPCHK(dev_malloc(&b->dsteps, (size_t)n * FAST_MAX_T * sizeof(FastStep)));
  HIPCHK(dev_malloc(&b->dtiles, std::max<size_t>(1, all.size()) * sizeof(FastTile)));
  HIPCHK(dev_malloc(&b->dqctr, 32 * sizeof(int)));
  HIPCHK(hipMemcpy(b->dinst, hinst.data(), n * sizeof(FastInst), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->dbinst, hb.data(), n * sizeof(FastBatchInst), hipMemcpyHostToDevice));
  if (!all.empty()) HIPCHK(hipMemcpy(b->dtiles, all.data(), all.size() * sizeof(FastTile), hipMemcpyHostToDevice));
  HIPCHK(hipDeviceSynchronize());
  // depths of the first block, queue counters cleared
  hipLaunchKernelGGL(k_batch_advance, dim3((std::max(n, 4) + 255) / 256), dim3(256), 0, b->stream, b->dbinst, n, T,
                     b->dqctr, 32, 1, b->dsteps);
  HIPCHK(hipGetLastError());
  *out = b;
  return FDTD_OK;
}

// advance every instance by `blocks` blocks of T steps (instances stop at their own step count)
int fast_batch_run(FastBatch *b, int blocks, int use_graph)
{
  int rc;
  const int G = 16;
  if (use_graph && blocks >= G && !b->gexec) {
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    for (int l = 0; l < G; l++)
      if ((rc = batch_level(b))) {
        hipGraph_t g;
        (void)hipStreamEndCapture(b->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
    HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
    HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    b->glevels = G;
  }
  int done = 0;
  if (use_graph && b->gexec)
    for (; done + b->glevels <= blocks; done += b->glevels) HIPCHK(hipGraphLaunch(b->gexec, b->stream));
  for (; done < blocks; done++)
    if ((rc = batch_level(b))) return rc;
  return FDTD_OK;
}

// launch shape (FAST_LAUNCH_*): the UPML launches beside the interior launches on a
// second stream or after them, each tile kind as 4 or 8 wavefronts (interior tiles
// as 8 wavefronts only for R >= 8); applies to blocks queued afterwards (the
// captured graph is rebuilt)
int fast_batch_set_launch(FastBatch *b, int mode)
{
  if (mode < FAST_LAUNCH_AUTO || mode > FAST_LAUNCH_TWO_STREAMS_P16) return fdtd_fail(FDTD_EINVAL, "bad launch mode %d", mode);
  // the 16-wavefront UPML tiles exist for single grids only (the batch kernel's
  // tile bodies are 4 or 8 wavefronts)
  if (mode == FAST_LAUNCH_SPLIT_P16 || mode == FAST_LAUNCH_TWO_STREAMS_P16)
    return fdtd_fail(FDTD_EINVAL, "launch mode %d (16-wavefront UPML tiles) is not available for batches", mode);
  HIPCHK(hipStreamSynchronize(b->stream));
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->two_streams = mode == FAST_LAUNCH_TWO_STREAMS || mode == FAST_LAUNCH_TWO_STREAMS_PML8 ||
                   mode == FAST_LAUNCH_TWO_STREAMS_W8;
  b->pml8 = mode >= FAST_LAUNCH_SPLIT_PML8;
  b->inner8 = (mode == FAST_LAUNCH_SPLIT_W8 || mode == FAST_LAUNCH_TWO_STREAMS_W8) && b->R >= 8;
  return FDTD_OK;
}

// blocks still needed by the slowest instance
int fast_batch_blocks_left(const FastBatch *b)
{
  int most = 0;
  for (int i = 0; i < b->n; i++) {
    const int left = b->steps[i] - b->st[i]->steps_since_reset;
    most = std::max(most, (left + b->T - 1) / b->T);
  }
  return most;
}

// wait for the queued blocks and hand every instance's progress back to its state
int fast_batch_sync(FastBatch *b, int *steps_done)
{
  HIPCHK(hipStreamSynchronize(b->stream));
  std::vector<FastBatchInst> hb(b->n);
  HIPCHK(hipMemcpy(hb.data(), b->dbinst, b->n * sizeof(FastBatchInst), hipMemcpyDeviceToHost));
  for (int i = 0; i < b->n; i++) {
    FastState *s = b->st[i];
    s->cur = hb[i].cur;
    s->pending = 0;
    s->steps_since_reset = hb[i].n;
    s->src_on = 1;
    // the single-instance path continues from c_{n-1}
    const int t = hb[i].n - 1;
    if (t >= 1) {
      const double x = 0.01 * t, rc = 1.0 - exp(-(x * x));
      s->prev_c_re = rc * cos(b->omega[i] * t);
      s->prev_c_im = -rc * sin(b->omega[i] * t);
    } else {
      s->prev_c_re = s->prev_c_im = 0.0;
    }
    if (steps_done) steps_done[i] = hb[i].n;
  }
  return FDTD_OK;
}

// Contour of the per-step NTFF in the physical layout: the sides of `ct`
// (solver-layout indices, `off` = 1 for the padded MPI layout) converted for
// the fused kernel, and a sample buffer for FAST_MAX_T steps.
int fast_state_set_contour(FastState *s, const Contour &ct, int off, fast_post_fn post, void *ctx)
{
  int rc;
  if ((rc = fast_state_flush(s))) return rc;
  const long NPYa = s->h.NPY + 2L * off;
  FastContour &c = s->hct;
  memset(&c, 0, sizeof c);
  for (int q = 0; q < 4; q++) {
    const Seg &g = ct.s[q];
    FastCtrSide &f = c.s[q];
    const int ia = (int)(g.k0 / NPYa) - off, ja = (int)(g.k0 % NPYa) - off;
    f.horiz = g.dk == 1 ? 0 : 1;
    f.fi = f.horiz ? ja : ia;
    f.lo = f.horiz ? ia : ja;
    f.n = g.n;
    f.off = g.off;
    f.ef = g.ef == 0 ? 0 : 1;
    f.hf = g.hf == 3 ? 1 : 2;
    f.es = (float)g.es;
  }
  c.total = ct.total;
  // the kernel gives a lane one role at most on the sides with j fixed, and a
  // row one event at most on the sides with i fixed: the lines (fi and fi - 1)
  // of the two sides of each kind must not meet
  for (int hz = 0; hz < 2; hz++) {
    int fl[2], nc = 0;
    for (int q = 0; q < 4; q++)
      if (c.s[q].horiz == hz && c.s[q].n > 0 && nc < 2) fl[nc++] = c.s[q].fi;
    if (nc == 2 && abs(fl[0] - fl[1]) < 2)
      return fdtd_fail(FDTD_EINVAL, "NTFF contour %d cells across: too small for the FP32 path", abs(fl[0] - fl[1]));
  }
  HIPCHK(hipDeviceSynchronize());
  for (int b = 0; b < 2; b++) {
    if (s->csampb[b]) (void)dev_free(s->csampb[b]);
    s->csampb[b] = nullptr;
    s->cused[b] = false;
    if (!s->cfree[b]) HIPCHK(hipEventCreateWithFlags(&s->cfree[b], hipEventDisableTiming));
  }
  s->csamp = nullptr;
  s->cbuf = 0;
  s->tilesT = -1;  // the tile lists are split by the contour when next built
  if (ct.total > 0) {
    for (int b = 0; b < 2; b++) HIPCHK(dev_malloc(&s->csampb[b], (size_t)FAST_MAX_T * 3 * ct.total * sizeof(float2)));
    s->csamp = s->csampb[0];
  }
  s->post = post;
  s->post_ctx = ctx;
  return FDTD_OK;
}
