// Output assembly on the GPU:
//   sf in the reference layout                 dens.py:241,318
//   sfplt = get_dplot(sf)                      dens.py:142-163,323
//   kgrid, kgridplt                            dens.py:325-344
//   Ewald-sphere sections of kgridplt          plot2d.py:218-263
//   fold of a padded grid (remap_grid[_tcl])   dens.py:65-108
#include <cmath>

#include "mdsf_internal.h"

namespace mdsf {

namespace {

// sf (Nx, Ny, Nz/2+1) from the accumulators sf_main[kz][ky][kx] (kz < Nz/2)
// and sf_nyq[kx][ky]
__global__ void k_sf_read(const double* __restrict__ sfm, const double* __restrict__ sfn,
                          double* __restrict__ out, int nx, int ny, int m) {
  const long long h = m + 1;
  const long long total = (long long)nx * ny * h;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long row = i / h;
    const int kz = (int)(i - row * h);
    const long long x = row / ny, y = row - x * ny;
    out[i] = kz < m ? sfm[((long long)kz * ny + y) * nx + x] : sfn[row];
  }
}

__device__ __forceinline__ int pm(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// value of k * 2 pi / L with the reference's operation order; `num` is the
// Python-float numerator, the division is float32 when L is float32 (NEP 50).
__device__ __forceinline__ double kdiv(double num, double L, int f32) {
#pragma clang fp contract(off)
  if (f32) return (double)((float)num / (float)L);
  return num / L;
}

__global__ void k_outputs(const double* __restrict__ sf, int nx, int ny, int nz, double L0,
                          double L1, double L2, int f32, double* __restrict__ sfplt,
                          double* __restrict__ kgrid, double* __restrict__ kgridplt) {
#pragma clang fp contract(off)
  const double twopi = 6.283185307179586;  // 2.0 * math.pi
  const double pi = 3.141592653589793;
  const int h = nz / 2 + 1;
  const long long nkg = (long long)nx * ny * h;
  const int px = nx - 2, py = ny - 2, pz = nz - 1;
  const long long nplt = (px > 0 && py > 0 && pz > 0) ? (long long)px * py * pz : 0;
  const long long total = nkg > nplt ? nkg : nplt;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    if (kgrid && i < nkg) {
      const int iz = (int)(i % h);
      const int iy = (int)((i / h) % ny);
      const int ix = (int)(i / ((long long)h * ny));
      double kx, ky;
      if (ix < nx / 2) kx = kdiv((double)ix * 2.0 * pi, L0, f32);
      else kx = kdiv(-((double)(nx - 1 - ix) + 0.5) * 2.0 * pi, L0, f32);
      if (iy < ny / 2) ky = kdiv((double)iy * 2.0 * pi, L1, f32);
      else ky = kdiv(-((double)(ny - 1 - iy) + 0.5) * 2.0 * pi, L1, f32);
      const double kz = kdiv((double)iz * 2.0 * pi, L2, f32);
      double* o = kgrid + 4 * i;
      o[0] = kx;
      o[1] = ky;
      o[2] = kz;
      o[3] = 0.0;
    }
    if (i < nplt) {
      const int c = (int)(i % pz);
      const int b = (int)((i / pz) % py);
      const int a = (int)(i / ((long long)pz * py));
      const int kx = a + 1 - nx / 2, ky = b + 1 - ny / 2, kz = c + 1 - nz / 2;
      double v;
      if (kx == 0 && ky == 0 && kz == 0) {
        const double s100 = sf[((long long)1 * ny + 0) * h + 0];
        const double s010 = sf[((long long)0 * ny + 1) * h + 0];
        const double s001 = sf[((long long)0 * ny + 0) * h + 1];
        v = 1.0 / 3.0 * (s100 + s010 + s001);
      } else if (kz >= 0) {
        v = sf[((long long)pm(kx, nx) * ny + pm(ky, ny)) * h + kz];
      } else {
        v = sf[((long long)pm(-kx, nx) * ny + pm(-ky, ny)) * h + (-kz)];
      }
      if (sfplt) sfplt[i] = v;
      if (kgridplt) {
        double* o = kgridplt + 4 * i;
        o[0] = kdiv(((double)a - (double)px / 2.0) * 2.0 * pi, L0, f32);
        o[1] = kdiv(((double)b - (double)py / 2.0) * 2.0 * pi, L1, f32);
        o[2] = kdiv(((double)c - (double)pz / 2.0) * 2.0 * pi, L2, f32);
        o[3] = v;
      }
    }
  }
  (void)twopi;
}

// --- Ewald sections: RegularGridInterpolator(method="linear", bounds_error=False)
__device__ int find_cell(const double* __restrict__ kp, long long stride, int n, double x,
                         bool* oob) {
  // grid values g[i] = kp[i * stride]; i = searchsorted(g, x, 'left') - 1,
  // clipped to [0, n - 2], i.e. g[i] < x <= g[i + 1] inside the grid.  The
  // k-grid axes are evenly spaced, so the cell is guessed from the spacing
  // and then moved until the comparisons against the stored values hold
  // (the same cell a binary search finds, in one or two loads).
  const double g0 = kp[0], gl = kp[(long long)(n - 1) * stride];
  if (x < g0 || x > gl || x != x) {
    *oob = true;
    return 0;
  }
  const double h = (gl - g0) / (double)(n - 1);
  int i = h > 0.0 ? (int)floor((x - g0) / h) : 0;
  i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
  while (i > 0 && kp[(long long)i * stride] >= x) --i;
  while (i < n - 2 && kp[(long long)(i + 1) * stride] < x) ++i;
  return i;
}

__device__ double interp3(const double* __restrict__ kp, int nx, int ny, int nz, double x,
                          double y, double z) {
#pragma clang fp contract(off)
  const long long sx = (long long)ny * nz * 4, sy = (long long)nz * 4, sz = 4;
  bool oob = false;
  const int i = find_cell(kp + 0, sx, nx, x, &oob);
  const int j = find_cell(kp + 1, sy, ny, y, &oob);
  const int k = find_cell(kp + 2, sz, nz, z, &oob);
  if (oob) return nan("");
  const double tx = (x - kp[i * sx]) / (kp[(i + 1) * sx] - kp[i * sx]);
  const double ty = (y - kp[1 + j * sy]) / (kp[1 + (j + 1) * sy] - kp[1 + j * sy]);
  const double tz = (z - kp[2 + k * sz]) / (kp[2 + (k + 1) * sz] - kp[2 + k * sz]);
  const double wx[2] = {1.0 - tx, tx}, wy[2] = {1.0 - ty, ty}, wz[2] = {1.0 - tz, tz};
  double value = 0.0;
  // corner order of itertools.product over ((i, 1-t), (i+1, t)) per axis
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        double w = 1.0;
        w = w * wx[a];
        w = w * wy[b];
        w = w * wz[c];
        const double v = kp[(i + a) * sx + (j + b) * sy + (k + c) * sz + 3];
        value = value + v * w;
      }
  return value;
}

__global__ void k_ewald(const double* __restrict__ kp, int nx, int ny, int nz, double K,
                        double* __restrict__ xy, double* __restrict__ xz, double* __restrict__ yz) {
#pragma clang fp contract(off)
  const long long sx = (long long)ny * nz * 4, sy = (long long)nz * 4;
  const long long nxy = (long long)nx * ny, nxz = (long long)nx * nz, nyz = (long long)ny * nz;
  const long long total = nxy + nxz + nyz;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < nxy) {
      const int ix = (int)(i / ny), iy = (int)(i % ny);
      const double X = kp[ix * sx], Y = kp[1 + iy * sy];
      const double th = atan(sqrt(pow(X, 2.0) + pow(Y, 2.0)) / K);
      const double c = cos(th);
      if (xy) xy[i] = interp3(kp, nx, ny, nz, X * c, Y * c, K * (1.0 - c));
    } else if (i < nxy + nxz) {
      const long long j = i - nxy;
      const int ix = (int)(j / nz), iz = (int)(j % nz);
      const double X = kp[ix * sx], Z = kp[2 + iz * 4];
      const double th = atan(sqrt(pow(X, 2.0) + pow(Z, 2.0)) / K);
      const double c = cos(th);
      if (xz) xz[j] = interp3(kp, nx, ny, nz, X * c, K * (1.0 - c), Z * c);
    } else {
      const long long j = i - nxy - nxz;
      const int iy = (int)(j / nz), iz = (int)(j % nz);
      const double Y = kp[1 + iy * sy], Z = kp[2 + iz * 4];
      const double th = atan(sqrt(pow(Y, 2.0) + pow(Z, 2.0)) / K);
      const double c = cos(th);
      if (yz) yz[j] = interp3(kp, nx, ny, nz, K * (1.0 - c), Y * c, Z * c);
    }
  }
}

// --- cylindrical (r, z) average: the r-theta-z interpolation of PLOT_RAD_NEW
// (plot2d.py:595-627).  oa[ir][iz] = mean over k < nth[ir] of the
// interpolant at (r cos t, r sin t, z) @ binv, t = k (2 pi / nth[ir]),
// r = rarr[ir], z = zar[iz]; NaN when any sample leaves the grid (the
// reference's np.average propagates the interpolator's NaN fill).  The
// theta sum runs in sample order, as numpy's reduction over that axis does.
struct Mat3 {
  double m[9];  // row-major
};

__global__ void k_rz_average(const double* __restrict__ kp, int nx, int ny, int nz,
                             const double* __restrict__ rarr, const int* __restrict__ nth, int nr,
                             const double* __restrict__ zar, int nzb, Mat3 b,
                             double* __restrict__ oa) {
#pragma clang fp contract(off)
  const long long total = (long long)nr * nzb;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int ir = (int)(i / nzb), iz = (int)(i - (long long)ir * nzb);
    const double r = rarr[ir], z = zar[iz];
    const int nt = nth[ir];
    const double step = (2.0 * 3.141592653589793) / (double)nt;  // np.linspace(0, 2 pi, nt, endpoint=False)
    double sum = 0.0;
    for (int k = 0; k < nt; ++k) {
      const double t = (double)k * step;
      const double x = r * cos(t), y = r * sin(t);
      // (x, y, z) @ binv
      const double u = x * b.m[0] + y * b.m[3] + z * b.m[6];
      const double v = x * b.m[1] + y * b.m[4] + z * b.m[7];
      const double w = x * b.m[2] + y * b.m[5] + z * b.m[8];
      sum = sum + interp3(kp, nx, ny, nz, u, v, w);
    }
    oa[i] = sum / (double)nt;
  }
}

// Interpolant at arbitrary points (n, 3): the RegularGridInterpolator calls
// on plain grids (plot2d.py:588, 800-840).
__global__ void k_interp_points(const double* __restrict__ kp, int nx, int ny, int nz,
                                const double* __restrict__ pts, long long n,
                                double* __restrict__ out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = interp3(kp, nx, ny, nz, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

unsigned grid_for(long long total) {
  long long b = (total + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Fold of a padded FP64 grid onto the periodic cell with the reference's
// tables des = [0, Nb, N - Nb, N], ori = [0, Nb, Nb + N, 2 Nb + N] per axis
// (dens.py:86-108; remap_grid, dens.py:65-83, is the same with one table for
// every axis).  One thread per destination cell adds, in the reference's
// order of slab additions, the sources that land on it: the core copy, the 6
// sides (z, y, x normal for the L side, then for the H side inside the
// same io1 loop), the 12 edges and the 8 corners, the corners' z destination
// taken from the y side's range (dens.py:107).  Same operations in the same
// order, so the sums are bitwise the reference's.
struct FoldAxis {
  int n, nb;
};

// Source index on one axis for destination c and source side io (0 = L,
// 2 = H), or -1 when c lies outside that side's destination range
// des[2 - io] .. des[3 - io].
__device__ __forceinline__ int fold_src(int c, int io, const FoldAxis& a) {
  if (io == 0) return c >= a.n - a.nb ? c - (a.n - a.nb) : -1;   // ori [0, Nb) -> des [N-Nb, N)
  return c < a.nb ? a.nb + a.n + c : -1;                          // ori [Nb+N, 2Nb+N) -> des [0, Nb)
}

__global__ void k_fold_padded(const double* __restrict__ d0, double* __restrict__ d1,
                              FoldAxis ax, FoldAxis ay, FoldAxis az) {
#pragma clang fp contract(off)
  const long long total = (long long)ax.n * ay.n * az.n;
  const long long py = ay.n + 2 * ay.nb, pz = az.n + 2 * az.nb;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int cz = (int)(i % az.n);
    const int cy = (int)((i / az.n) % ay.n);
    const int cx = (int)(i / ((long long)az.n * ay.n));
    auto at = [&](int px, int qy, int qz) { return d0[((long long)px * py + qy) * pz + qz]; };
    const int kx = cx + ax.nb, ky = cy + ay.nb, kz = cz + az.nb;   // core sources
    double v = at(kx, ky, kz);
    for (int io1 = 0; io1 < 4; io1 += 2) {
      const int sx1 = fold_src(cx, io1, ax), sy1 = fold_src(cy, io1, ay),
                sz1 = fold_src(cz, io1, az);
      if (sz1 >= 0) v += at(kx, ky, sz1);   // sides normal to z
      if (sy1 >= 0) v += at(kx, sy1, kz);   // sides normal to y
      if (sx1 >= 0) v += at(sx1, ky, kz);   // sides normal to x
      for (int io2 = 0; io2 < 4; io2 += 2) {
        const int sy2 = fold_src(cy, io2, ay), sz2 = fold_src(cz, io2, az);
        if (sy1 >= 0 && sz2 >= 0) v += at(kx, sy1, sz2);   // edges parallel to x
        if (sx1 >= 0 && sz2 >= 0) v += at(sx1, ky, sz2);   // edges parallel to y
        if (sx1 >= 0 && sy2 >= 0) v += at(sx1, sy2, kz);   // edges parallel to z
        // corners: z destination range of the y side io2, z source side io3
        const bool zdst = io2 == 0 ? cz >= az.n - az.nb : cz < az.nb;
        if (sx1 >= 0 && sy2 >= 0 && zdst) {
          const int rel = io2 == 0 ? cz - (az.n - az.nb) : cz;   // offset in that range
          for (int io3 = 0; io3 < 4; io3 += 2) {
            const int sz3 = io3 == 0 ? rel : az.nb + az.n + rel;
            v += at(sx1, sy2, sz3);
          }
        }
      }
    }
    d1[i] = v;
  }
}

}  // namespace

hipError_t launch_fold_padded(const double* d0, const int n[3], const int nb[3], double* d1,
                              hipStream_t s) {
  const long long total = (long long)n[0] * n[1] * n[2];
  if (total == 0) return hipSuccess;
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fold_padded, dim3((unsigned)blocks), dim3(256), 0, s, d0, d1,
                     FoldAxis{n[0], nb[0]}, FoldAxis{n[1], nb[1]}, FoldAxis{n[2], nb[2]});
  return hipGetLastError();
}

hipError_t launch_sf_read(const Plan& p, double* sf, hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  hipLaunchKernelGGL(k_sf_read, dim3(grid_for((long long)nx * ny * (m + 1))), dim3(256), 0, s,
                     p.sf_main, p.sf_nyq, sf, nx, ny, m);
  return hipGetLastError();
}

hipError_t launch_outputs(const int n[3], const double* sf, const double box[3], int box_f32,
                          double* sfplt, double* kgrid, double* kgridplt, hipStream_t s) {
  const long long nkg = (long long)n[0] * n[1] * (n[2] / 2 + 1);
  hipLaunchKernelGGL(k_outputs, dim3(grid_for(nkg)), dim3(256), 0, s, sf, n[0], n[1], n[2],
                     box[0], box[1], box[2], box_f32, sfplt, kgrid, kgridplt);
  return hipGetLastError();
}

hipError_t launch_ewald(const double* kp, int nx, int ny, int nz, double wl, double* xy,
                        double* xz, double* yz, hipStream_t s) {
  const double K = 2.0 * 3.141592653589793 / wl;
  const long long total = (long long)nx * ny + (long long)nx * nz + (long long)ny * nz;
  hipLaunchKernelGGL(k_ewald, dim3(grid_for(total)), dim3(256), 0, s, kp, nx, ny, nz, K, xy, xz,
                     yz);
  return hipGetLastError();
}

hipError_t launch_rz_average(const double* kp, int nx, int ny, int nz, const double* rarr,
                             const int* nth, int nr, const double* zar, int nzb,
                             const double binv[9], double* oa, hipStream_t s) {
  Mat3 b;
  for (int k = 0; k < 9; ++k) b.m[k] = binv[k];
  hipLaunchKernelGGL(k_rz_average, dim3(grid_for((long long)nr * nzb)), dim3(256), 0, s, kp, nx, ny,
                     nz, rarr, nth, nr, zar, nzb, b, oa);
  return hipGetLastError();
}

hipError_t launch_interp_points(const double* kp, int nx, int ny, int nz, const double* pts,
                                long long n, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_interp_points, dim3(grid_for(n)), dim3(256), 0, s, kp, nx, ny, nz, pts, n,
                     out);
  return hipGetLastError();
}

}  // namespace mdsf
