// fdtd_internal.h -- types shared by the HIP sources of libmpifdtd_amd (not part of the ABI)
#ifndef FDTD_INTERNAL_H
#define FDTD_INTERNAL_H

// One side of the closed NTFF contour, in array coordinates of the solver layout.
struct Seg {
  long k0;        // array index of the side's first cell
  int dk;         // array stride between consecutive cells
  int n;          // number of cells
  int off;        // offset of the side in the packed contour arrays
  int ef;         // field slot of the tangential E component
  int hf;         // field slot of the H component
  int hnb;        // array offset of the H neighbour averaged with it
  int ua, wa;     // accumulator slots receiving the E (U) and H (W) taps
  double es, hs;  // sign of the E sample, signed 0.5 of the H average
  double r2x0, r2y0, dr2x, dr2y;  // position of cell c relative to the centre
};

struct Contour {
  Seg s[4];  // bottom, right, top, left
  int total;
};

int fdtd_fail(int code, const char *fmt, ...);

// Device memory of the solver objects.  Freed blocks are kept per exact size
// (up to 64 GiB) and handed to the next request of that size, so that sweeps
// creating and destroying many equal grids skip hipMalloc / hipFree; dev_free
// synchronises the device first, so no queued kernel still uses a block that
// is handed out again.  fdtd_hip_trim() returns the kept blocks.
hipError_t dev_malloc_bytes(void **p, size_t bytes);
hipError_t dev_free(void *p);
template <typename T>
static inline hipError_t dev_malloc(T **p, size_t bytes)
{
  return dev_malloc_bytes((void **)p, bytes);
}

// ---------------------------------------------------------------- FP32 fast path
// A tile of instance `inst`: core cells [i0, i0 + ni) x [j0, j0 + nj) written after the block's
// steps, computed from the block of cells loaded from row bi and column bj
// (bi = i0 - (T + 1) and bj = j0 - (T + 1) inside the grid; a core that
// starts at the grid's first row or column needs no halo on that side).
struct FastTile {
  int inst, i0, j0, ni, nj, bi, bj;
};

// Per inner step scalars of the sources (host-computed in FP64, rounded to FP32).
struct FastStep {
  float2 c;        // MPI: ray_coef * e^{-i w n}
  float2 dc;       // c_n - c_{n-1}
  float tn, tn1;   // plain: n - t0 of pulse component 0 / 1
  float tp, tp1;   // plain: (n - 1) - t0
  float hasprev;   // 0 on the first step after a reset (no previous source)
  float pad;
};

#define FAST_MAX_T 16
// rows of a UPML tile's block: the 4-wavefront UPML body (4 x 8 rows, the
// merged, split and two-stream shapes, and batches) and the wide bodies of the
// *_PML8, *_W8 and *_P16 shapes (12 x 4 or 16 x 3 rows)
#define FAST_UPML_ROWS 32
#define FAST_UPML_ROWS_WIDE 48

// Device view of one FP32 instance (arrays in the physical N_PX x N_PY layout, row pitch `pitch`).
struct FastInst {
  int NPX, NPY, pitch, flavour;
  float c2;                 // 1/mu0 = c0^2
  // fields, ping-pong [2]: TM: f0 = Ez, f1 = Hx, f2 = Hy ; TE: f0 = Ex, f1 = Ey, f2 = Hz
  float2 *f0[2], *f1[2], *f2[2];
  // UPML auxiliaries (touched only in PML cells), ping-pong, in the compact
  // layout of aux_cell (the cells outside the box [axlo, axhi) x [aylo, ayhi))
  //   TM: a0 Jz a1 Dz a2 Mx a3 Bx a4 My a5 By ; TE: a0 Jx a1 Dx a2 Jy a3 Dy a4 Mz a5 Bz
  float2 *aux[6][2];
  const float *ie0, *ie1;   // 1/eps of the E components (TM: Ez ; TE: Ex, Ey)
  // 1-D coefficient tables (FP32), same slot meaning as fdtd_coef_tables
  const float *tx[5], *tsx, *ty[4], *tsy;
  const unsigned char *pmlx, *pmly;  // 1 where the cell uses the UPML recurrence
  // the box [axlo, axhi) x [aylo, ayhi) holds no UPML cell: the auxiliaries have
  // no storage there and the UPML tiles load none (an empty row or column range
  // when UPML rows or columns lie inside the grid); acw = aylo + NPY - ayhi, the
  // auxiliary cells of one row of the box's row range
  int axlo, axhi, aylo, ayhi, acw;
  const float2 *ex, *ey;    // e^{i kx x}, e^{i ky y} (MPI scattered wave)
  float cpc, spc, w;        // plain pulse: cos/c0, sin/c0, omega
  float gx[2], gy[2], dot[2];
  int on[2];
};

// One side of the NTFF contour in the FP32 physical layout, for sampling inside
// the fused kernel: cells (fi, lo..lo+n-1) of a vertical side (i fixed) or
// (lo..lo+n-1, fi) of a horizontal one (j fixed), packed from `off`; the H
// average of a contour cell takes its neighbour one cell back across the side.
struct FastCtrSide {
  int horiz;      // 1: j fixed (bottom, top), 0: i fixed (right, left)
  int fi, lo, n;  // fixed coordinate, first varying coordinate, cells
  int off;        // offset in the packed contour
  int ef;         // E component: 0 -> f0, 1 -> f1
  int hf;         // H component: 1 -> f1, 2 -> f2
  float es;       // sign of the E sample
};
struct FastContour {
  FastCtrSide s[4];
  int total;
};

// Progress of one batched instance (device memory, advanced between launches).
struct FastBatchInst {
  double omega;  // angular frequency of the scattered wave (MPI flavour)
  int n;         // steps done
  int steps;     // steps to run in total
  int Ti;        // steps of the current launch (0: finished)
  int cur;       // ping-pong index of the input buffers
};

struct FastArgs {
  const FastInst *inst;     // device array of instances
  const FastTile *tiles;    // tile list of this launch
  int qb[2][9];             // 8 XCD partitions of the UPML (0) and interior (1) tiles:
                            // tiles [qb[k][q], qb[k][q+1]) run on XCD q
  int nb_pml;               // workgroups [0, nb_pml) run the UPML tiles (a multiple of 8)
  int T;                    // steps per launch
  int src;                  // 1: sources active
  int cur;                  // ping-pong index of the input buffers
  // batched runs (k_fast_batch): per-instance progress and the launch's two tile
  // queues, UPML (0) and interior (1), at tiles[qoff[k], qoff[k] + qn[k])
  FastBatchInst *binst;
  const FastStep *bsteps;   // [instance][FAST_MAX_T] the step scalars of the launch's steps
  int *qctr;
  int qn[2], qoff[2];
  // per-step NTFF samples: after every inner step s, each contour cell of a tile's
  // core writes csamp[s][0][c] = es * E, csamp[s][1][c] = H, and each cell one
  // back across a side writes csamp[s][2][c] = H of that neighbour (null: off)
  float2 *csamp;
  FastContour ct;
  FastStep st[FAST_MAX_T];  // single-instance runs: the block's steps, passed by value
};

struct FastState;
int fast_state_create(const fdtd_grid_desc &d, FastState **out);
void fast_state_destroy(FastState *s);
int fast_state_set_coefs(FastState *s, const fdtd_coef_tables *t);
int fast_state_set_inv_eps(FastState *s, double *const *eps_dev, int NPYa, int off);
int fast_state_step(FastState *s, const fdtd_step_params *p, int *launched);
int fast_state_flush(FastState *s);
int fast_state_export(FastState *s, int slot, double2 *dst_dev, int NPYa, int off);
int fast_state_import(FastState *s, int slot, const double2 *src_dev, int NPYa, int off);
int fast_state_reset(FastState *s);
int fast_state_gather(FastState *s, const Contour &ct, int off, double2 *cE, double2 *cH);
// the three physical fields of the current buffer (after launching the queued
// steps): TM Ez Hx Hy / TE Ex Ey Hz, each NPX x pitch float2; *n = NPX * pitch
int fast_state_fields(FastState *s, const float2 *f[3], size_t *n);
// steps queued and not yet launched; the row pitch of the physical layout (cells)
int fast_state_pending(const FastState *s);
int fast_state_pitch(const FastState *s, int *pitch);
// Per-step contour sampling inside the fused kernel: after every launch of n
// steps, post(ctx, samples, n, times) receives the device buffer
// [n][3][total] float2 and the time of each step.
// after each block of n steps with contour samples: the consumer queues its work
// on its own stream after the block, and records `consumed` there once it has
// read `samples` (the buffer is written again two blocks later)
typedef int (*fast_post_fn)(void *ctx, const float2 *samples, int n, const double *times, hipEvent_t consumed);
int fast_state_set_contour(FastState *s, const Contour &ct, int off, fast_post_fn post, void *ctx);
int fast_state_set_blocking(FastState *s, int T, int R);
int fast_state_set_launch_mode(FastState *s, int mode);
int fast_state_get_launch_mode(const FastState *s);

struct FastBatch;
int fast_batch_create(FastState *const *st, const fdtd_step_params *p0, const int *steps, int n, int T, int R,
                      FastBatch **out);
int fast_batch_run(FastBatch *b, int blocks, int use_graph);
int fast_batch_set_launch(FastBatch *b, int mode);
int fast_batch_blocks_left(const FastBatch *b);
int fast_batch_sync(FastBatch *b, int *steps_done);
void fast_batch_destroy(FastBatch *b);
void fast_state_get_blocking(const FastState *s, int *T, int *R);
// profiling of the fused launches: total ms, launches, algorithmic bytes moved
void fast_profile_enable(int on);
void fast_profile_read(double *ms, long *launches, double *bytes, double *cell_updates);
void fast_profile_read6(double out[6]);

#endif
