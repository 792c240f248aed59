// Parked atoms: windows that leave the padded grid on some axis.
//
// The reference adds an atom's window into d0[x0:x1, y0:y1, z0:z1] and takes
// the coordinates from the same slices of its grid (dens.py:292-308), so
// numpy clips both alike: per axis the padded cells p that
// numpy_slice(x0, x1) keeps, each at grid coordinate (p - Nb) dr.  A window
// across the low edge or past the high edge keeps nothing, one cut by the
// high edge keeps its inner part, one wholly at negative indices is counted
// from the far end.  k_prep parks such atoms in their type's last bucket,
// which the tiled passes never enumerate; a pass after the deposition adds
// their clipped windows through the remap_grid_tcl fold (dens.py:86-108,
// corner z slot included), and the adjoint reductions sum their gradient
// over the same cells.
#pragma once

#include "mdsf_internal.h"

namespace mdsf {

// Clipped window of a parked atom: padded cells [lo_d, lo_d + n_d) per axis.
struct ClipWin {
  long long lo0, lo1, lo2, n0, n1, n2;
  bool empty;
};

__device__ __forceinline__ void clip_axis(int ir, int A, int N, int nb, long long* lo,
                                          long long* n) {
  long long s0, s1;
  numpy_slice((long long)ir - A + nb, (long long)ir + A + nb, (long long)N + 2LL * nb, &s0, &s1);
  *lo = s0;
  *n = A == 0 ? 0 : s1 - s0;
}

__device__ __forceinline__ ClipWin clip_window(const SysDev& sys, const TypeDev& ty,
                                               const int4& q) {
  ClipWin w;
  clip_axis(q.x, ty.half[0], sys.n[0], sys.nb, &w.lo0, &w.n0);
  clip_axis(q.y, ty.half[1], sys.n[1], sys.nb, &w.lo1, &w.n1);
  clip_axis(q.z, ty.half[2], sys.n[2], sys.nb, &w.lo2, &w.n2);
  w.empty = w.n0 <= 0 || w.n1 <= 0 || w.n2 <= 0;
  return w;
}

// amp exp(-(b^T G b) / 2 s^2) at padded cell (px, py, pz), b = r - (p - Nb) dr;
// G b into (gb0, gb1, gb2).
__device__ __forceinline__ double clip_value(const SysDev& sys, const TypeDev& ty, double rx,
                                             double ry, double rz, long long px, long long py,
                                             long long pz, double& gb0, double& gb1,
                                             double& gb2) {
  const double bx = rx - (double)(px - sys.nb) * sys.dr[0];
  const double by = ry - (double)(py - sys.nb) * sys.dr[1];
  const double bz = rz - (double)(pz - sys.nb) * sys.dr[2];
  gb0 = sys.g00 * bx + sys.g01 * by + sys.g02 * bz;
  gb1 = sys.g01 * bx + sys.g11 * by + sys.g12 * bz;
  gb2 = sys.g02 * bx + sys.g12 * by + sys.g22 * bz;
  return ty.amp * exp(-(bx * gb0 + by * gb1 + bz * gb2) / ty.two_sig2);
}

__device__ __forceinline__ int clip_side(long long p, int N, int nb) {
  return p < nb ? SIDE_L : (p < nb + N ? SIDE_C : SIDE_H);
}

// Folded destination of padded cell p (remap_grid_tcl, dens.py:86-108): per
// axis p - Nb mod N, except that a corner copy (no axis on the core) takes
// its z slot from the y side (dens.py:107).
__device__ __forceinline__ long long clip_dest(const SysDev& sys, long long px, long long py,
                                               long long pz) {
  const int nb = sys.nb, N0 = sys.n[0], N1 = sys.n[1], N2 = sys.n[2];
  const int sx = clip_side(px, N0, nb), sy = clip_side(py, N1, nb), sz = clip_side(pz, N2, nb);
  const long long dx = sx == SIDE_L ? px - nb + N0 : (sx == SIDE_C ? px - nb : px - nb - N0);
  const long long dy = sy == SIDE_L ? py - nb + N1 : (sy == SIDE_C ? py - nb : py - nb - N1);
  long long dz = sz == SIDE_L ? pz - nb + N2 : (sz == SIDE_C ? pz - nb : pz - nb - N2);
  if (sx != SIDE_C && sy != SIDE_C && sz != SIDE_C)
    dz = (sz == SIDE_L ? pz : pz - nb - N2) + (sy == SIDE_L ? N2 - nb : 0);
  return (dx * N1 + dy) * N2 + dz;
}

// Lane li of nl lanes: its share of sum_p dL/drho(dest(p)) dv/dr over the
// clipped window of a parked atom (q, position r), added into (g0, g1, g2).
template <typename R>
__device__ __forceinline__ void parked_grad(const SysDev& sys, const TypeDev& ty, const int4& q,
                                            double rx, double ry, double rz,
                                            const R* __restrict__ grho, int li, int nl,
                                            double& g0, double& g1, double& g2) {
  const ClipWin w = clip_window(sys, ty, q);
  if (w.empty) return;
  const long long total = w.n0 * w.n1 * w.n2;
  for (long long i = li; i < total; i += nl) {
    const long long px = w.lo0 + i / (w.n1 * w.n2), py = w.lo1 + (i / w.n2) % w.n1,
                    pz = w.lo2 + i % w.n2;
    double gb0, gb1, gb2;
    const double v = clip_value(sys, ty, rx, ry, rz, px, py, pz, gb0, gb1, gb2);
    const double f = (double)grho[clip_dest(sys, px, py, pz)] * (-2.0 * v / ty.two_sig2);
    g0 += f * gb0;
    g1 += f * gb1;
    g2 += f * gb2;
  }
}

}  // namespace mdsf
