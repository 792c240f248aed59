// Corner-quirk spill of the border fold, shared by the deposition and the
// tiled adjoint.
//
// remap_grid_tcl (dens.py:101-107) sends a padded contribution whose three
// sides are all off the core C ("corner") to the z slot chosen by its *y*
// side: with y on H and z on L it lands on z = u_z + Nb, with y on L and z on
// H on z = u_z - Nb, instead of u_z mod N.  Those destinations lie in the
// bands z in [0, Nb) (band 0) and [N - Nb, N) (band 1), up to Nb cells from
// the atom's own window, so they can fall in a z tile that the window
// [ir - A, ir + A) does not reach.  The tiled passes visit only reached tiles
// and evaluate the quirk there; the "spill" is the rest: quirk contributions
// whose destination z tile the window does not reach.
#pragma once

#include "mdsf_internal.h"
#include "tiles.h"

namespace mdsf {

// Cells [lo, hi) of a window image range on one axis.
struct SpillRange {
  int lo, hi;
  __device__ int len() const { return hi > lo ? hi - lo : 0; }
};

// The window's images off the core on one axis: u < 0 (L) and u >= N (H).
__device__ __forceinline__ SpillRange spill_side_l(int ir, int A) {
  return {ir - A, min(ir + A, 0)};
}
__device__ __forceinline__ SpillRange spill_side_h(int ir, int A, int N) {
  return {max(ir - A, N), ir + A};
}

// Does the quirk destination z tile [t0, t1) lie outside the tiles the
// atom's z window reaches (so that only the spill pass covers it)?
__device__ __forceinline__ bool spill_tile_unreached(int irz, int Az, int t0, int t1, int Nz) {
  return !window_reaches(irz, Az, t0, t1, Nz);
}

// Bands in which an atom spills (bit b for band b): it has off-core x images,
// y images on the band's y side, z images on its z side, and one of their
// destination z tiles lies outside the tiles its z window reaches.
__device__ __forceinline__ int spill_bands(const int4& q, const TypeDev& ty, const SysDev& sys) {
  const int N0 = sys.n[0], N1 = sys.n[1], N2 = sys.n[2], nb = sys.nb;
  const int Ax = ty.half[0], Ay = ty.half[1], Az = ty.half[2];
  if (Ax == 0 || Ay == 0 || Az == 0) return 0;
  if (!(q.x - Ax < 0 || q.x + Ax > N0)) return 0;
  int bands = 0;
#pragma unroll
  for (int band = 0; band < 2; ++band) {
    const SpillRange yr = band == 0 ? spill_side_h(q.y, Ay, N1) : spill_side_l(q.y, Ay);
    const SpillRange zr = band == 0 ? spill_side_l(q.z, Az) : spill_side_h(q.z, Az, N2);
    if (yr.len() == 0 || zr.len() == 0) continue;
    const int lo = band == 0 ? zr.lo + nb : zr.lo - nb, hi = band == 0 ? zr.hi + nb : zr.hi - nb;
    for (int t0 = (lo / kTile) * kTile; t0 < hi; t0 += kTile)
      if (spill_tile_unreached(q.z, Az, t0, min(t0 + kTile, N2), N2)) {
        bands |= 1 << band;
        break;
      }
  }
  return bands;
}

// Grid point contribution of one padded image (ux, uy, uz) of an atom at r:
// amp exp(-(b^T G b) / 2 sigma^2), b = r - u dr (dens.py:297-308), with the
// cell metric G = U U^T (xy coupling only).  With `grad` also d/dr of it.
__device__ __forceinline__ double spill_value(const SysDev& sys, const TypeDev& ty,
                                              const double r[3], int ux, int uy, int uz,
                                              double* grad) {
  const double bx = r[0] - (double)ux * sys.dr[0];
  const double by = r[1] - (double)uy * sys.dr[1];
  const double bz = r[2] - (double)uz * sys.dr[2];
  const double q = sys.g00 * bx * bx + 2.0 * sys.g01 * bx * by + sys.g11 * by * by +
                   sys.g22 * bz * bz;
  const double v = ty.amp * exp(-q / ty.two_sig2);
  if (grad) {
    const double k = -2.0 * v / ty.two_sig2;
    grad[0] = k * (sys.g00 * bx + sys.g01 * by);
    grad[1] = k * (sys.g01 * bx + sys.g11 * by);
    grad[2] = k * (sys.g22 * bz);
  }
  return v;
}

// Spill tiles per axis: tiles meeting [0, Nb) or [N - Nb, N), where folded
// images of off-core windows land (x, y) and the quirk bands lie (z): the
// low tiles [0, k0), k0 = ceil(Nb / 16), and the high tiles [fh, nt),
// fh = floor((N - Nb) / 16); all tiles when the two ranges meet.
__host__ __device__ inline void spill_axis(int N, int nb, int* k0, int* fh, int* cnt) {
  const int nt = (N + kTile - 1) / kTile;
  if (nb <= 0) {
    *k0 = 0;
    *fh = nt;
    *cnt = 0;
    return;
  }
  *k0 = (nb + kTile - 1) / kTile < nt ? (nb + kTile - 1) / kTile : nt;
  *fh = (N - nb) / kTile;
  *cnt = *fh <= *k0 ? nt : *k0 + nt - *fh;
  if (*fh <= *k0) *fh = *k0;
}
// position of tile i among the spill tiles of axis d, or -1
__device__ __forceinline__ int spill_index(const SysDev& sys, int d, int i) {
  const int k0 = sys.sp_k0[d], fh = sys.sp_fh[d];
  return i < k0 ? i : (i >= fh ? k0 + (i - fh) : -1);
}
// the j-th spill tile of axis d
__device__ __forceinline__ int spill_tile(const SysDev& sys, int d, int j) {
  const int k0 = sys.sp_k0[d];
  return j < k0 ? j : sys.sp_fh[d] + (j - k0);
}

// Marks the spill tiles an atom spills into on the given bands (its
// destination cells: off-core x images, y images on the band's side, z
// destinations in tiles its z window does not reach).
__device__ inline void spill_mark(const SysDev& sys, const int4& q, const TypeDev& ty, int bands,
                                  int* __restrict__ flag) {
  const int N0 = sys.n[0], N1 = sys.n[1], N2 = sys.n[2], nb = sys.nb;
  const int Ax = ty.half[0], Ay = ty.half[1], Az = ty.half[2];
  const SpillRange xl = spill_side_l(q.x, Ax), xh = spill_side_h(q.x, Ax, N0);
  for (int band = 0; band < 2; ++band) {
    if (!((bands >> band) & 1)) continue;
    const SpillRange yr = band == 0 ? spill_side_h(q.y, Ay, N1) : spill_side_l(q.y, Ay);
    const SpillRange zr = band == 0 ? spill_side_l(q.z, Az) : spill_side_h(q.z, Az, N2);
    // destination cell ranges: x and y wrapped into [0, N), z shifted by -/+ Nb
    const int ylo = band == 0 ? yr.lo - N1 : yr.lo + N1, yhi = band == 0 ? yr.hi - N1 : yr.hi + N1;
    const int zlo = band == 0 ? zr.lo + nb : zr.lo - nb, zhi = band == 0 ? zr.hi + nb : zr.hi - nb;
    for (int part = 0; part < 2; ++part) {
      const SpillRange xr = part == 0 ? xl : xh;
      if (xr.len() == 0) continue;
      const int xlo = part == 0 ? xr.lo + N0 : xr.lo - N0, xhi = part == 0 ? xr.hi + N0 : xr.hi - N0;
      for (int tx = xlo / kTile; tx * kTile < xhi; ++tx)
        for (int ty_ = ylo / kTile; ty_ * kTile < yhi; ++ty_)
          for (int tz = zlo / kTile; tz * kTile < zhi; ++tz) {
            if (!spill_tile_unreached(q.z, Az, tz * kTile, min(tz * kTile + kTile, N2), N2)) continue;
            const int jx = spill_index(sys, 0, tx), jy = spill_index(sys, 1, ty_),
                      jz = spill_index(sys, 2, tz);
            if (jx >= 0 && jy >= 0 && jz >= 0)
              flag[(jx * sys.sp_cnt[1] + jy) * sys.sp_cnt[2] + jz] = 1;
          }
    }
  }
}

}  // namespace mdsf
