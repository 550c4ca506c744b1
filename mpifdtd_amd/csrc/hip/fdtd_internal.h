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

#endif
