// Adjoint of the per-frame S(q) path: dL/d(coords) from dL/dS.
//
// Forward (dens.py:283-318):
//   d0(w)  = sum_a amp_a exp(-b^T G b / 2 s_a^2),  b = r_a - x_w,  G = U U^T   (window of a)
//   rho(v) = sum_{w : fold(w) = v} d0(w)                                       (dens.py:86-108)
//   S(q)  += |F(q)|^2,  F = rfftn(rho)
// Adjoint, per frame, with g = dL/dS on the stored half spectrum:
//   dL/drho(v) = 2 Re sum_q g(q) F(q) e^{+2 pi i q.v / N}      (inverse FFT with weights 2 g)
//   dL/dr_a    = sum_{w in window(a)} dL/drho(fold(w)) * amp exp(..) * (-2 G b / 2 s^2)
// The gather below is the exact transpose of the deposition + fold (including
// the reference's corner z mis-fold), so it differentiates the reference's
// function, not an idealised periodic one.
//
// Work decomposition: an item is (atom, x-chunk of <= 64 window planes,
// chunk of (y, z) pairs).  One wavefront per item: lane l owns x entry l of
// the chunk (factor, factor * b_x, destination plane), and each lane walks
// its (y, z) pairs with an inner loop over the chunk's x entries, whose
// values are broadcast with readlane.  Items write 3 FP64 partial sums; a
// second kernel adds each atom's partials in a fixed order (deterministic).
#include "clip.h"
#include "mdsf_internal.h"
#include "spill.h"
#include "tiles.h"

namespace mdsf {

namespace {

constexpr int kAdjThreads = 256;
constexpr double kLog2e = 1.4426950408889634;

__device__ __forceinline__ int wrap_idx(int u, int n) { return u < 0 ? u + n : (u >= n ? u - n : u); }

__device__ __forceinline__ int side_of(int u, int n) { return u < 0 ? SIDE_L : (u >= n ? SIDE_H : SIDE_C); }

__device__ __forceinline__ int rl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rl_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rl_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <typename R>
__device__ __forceinline__ R rl(R v, int l);
template <>
__device__ __forceinline__ float rl<float>(float v, int l) { return rl_f(v, l); }
template <>
__device__ __forceinline__ double rl<double>(double v, int l) { return rl_d(v, l); }

// exp(-q) of a non-negative FP64 exponent, evaluated in the plan precision.
template <typename R>
__device__ __forceinline__ double pair_exp(double q);
template <>
__device__ __forceinline__ double pair_exp<float>(double q) { return (double)expf(-(float)q); }
template <>
__device__ __forceinline__ double pair_exp<double>(double q) { return exp(-q); }

// exp of the xy cross term: exp2 with a pre-scaled slope in FP32, exp in FP64.
__device__ __forceinline__ float cross_exp(float k, float by) { return exp2f(k * by); }
__device__ __forceinline__ double cross_exp(double k, double by) { return exp(k * by); }

template <int MODE, typename R>
__global__ __launch_bounds__(kAdjThreads) void k_deposit_adjoint(
    SysDev sys, const TypeDev* __restrict__ td, const int4* __restrict__ s_ir,
    const double* __restrict__ s_r, const int* __restrict__ nvalid, const R* __restrict__ grho,
    double* __restrict__ part, long long nitems) {
  const int lane = threadIdx.x & 63;
  // XCD-aware order: workgroups go round-robin over the 8 XCDs, so block b
  // takes the b/8-th group of XCD (b mod 8)'s contiguous share of the items.
  // Items follow the atoms' (type, home bucket) order, so each XCD's L2 sees
  // one spatial band of the gradient grid at a time.
  const int nblk = gridDim.x, bid = blockIdx.x;
  const int q8 = nblk / 8;
  const int blk = bid < 8 * q8 ? (bid % 8) * q8 + bid / 8 : bid;
  const long long item = (long long)blk * (kAdjThreads / 64) + (threadIdx.x >> 6);
  if (item >= nitems) return;  // uniform per wavefront
  int t = -1;
  for (int u = 0; u < sys.ntypes; ++u) {
    const long long b = td[u].adj_item_base;
    const long long n = (long long)td[u].count * td[u].adj_nxc * td[u].adj_npc;
    if (item >= b && item < b + n) t = u;
  }
  double gx = 0.0, gy = 0.0, gz = 0.0;
  if (t >= 0) {
    const TypeDev& ty = td[t];
    const int per_atom = ty.adj_nxc * ty.adj_npc;
    const long long local = item - ty.adj_item_base;
    const int slot = (int)(local / per_atom);
    const int rem = (int)(local - (long long)slot * per_atom);
    const int xc = rem / ty.adj_npc, pc = rem - xc * ty.adj_npc;
    const int s = ty.type_start + slot;
    const int4 q = s < *nvalid ? s_ir[s] : make_int4(0, 0, 0, -1);
    if ((q.w & 0xffff) == t && !(q.w & kAtomParked)) {  // parked: k_adjoint_reduce
      const int n0 = sys.n[0], n1 = sys.n[1], n2 = sys.n[2], nb = sys.nb;
      const int Ax = ty.half[0], Ay = ty.half[1], Az = ty.half[2];
      const int Wz = 2 * Az;
      const double rx = s_r[3 * (size_t)s + 0], ry = s_r[3 * (size_t)s + 1],
                   rz = s_r[3 * (size_t)s + 2];
      const double ts2 = ty.two_sig2;
      const int x0 = xc * 64;
      const int xw = min(64, 2 * Ax - x0);
      // lane-owned x entry: factor amp*exp(-g00 bx^2/2s^2), factor*bx, plane, side
      int xd = 0, xnc = 0;
      R xv = (R)0, xb = (R)0, xk = (R)0;
      if (lane < xw) {
        const int u = q.x - Ax + x0 + lane;
        const double b = rx - (double)u * sys.dr[0];
        const double f = ty.amp * exp(-(sys.g00 * (b * b)) / ts2);
        xv = (R)f;
        xb = (R)(f * b);
        if (MODE == MODE_XY) {
          const double k = -2.0 * sys.g01 * b / ts2;
          xk = (R)(sizeof(R) == 4 ? k * kLog2e : k);
        }
        xd = wrap_idx(u, n0);
        xnc = side_of(u, n0) != SIDE_C;
      }
      const long long plane = (long long)n1 * n2;
      const int npairs = 2 * Ay * Wz;
      const int p_end = min(npairs, (pc + 1) * ty.adj_ppi);
      // The pair loop runs the same trip count on every lane (lanes past the
      // end work on a valid duplicate pair and drop the result), so the
      // readlane broadcasts below always execute with the full wavefront.
      for (int p0 = pc * ty.adj_ppi; p0 < p_end; p0 += 64) {
        const bool valid = p0 + lane < p_end;
        const int p = valid ? p0 + lane : p0;
        const int j = p / Wz, k = p - j * Wz;
        const int uy = q.y - Ay + j, uz = q.z - Az + k;
        const double by = ry - (double)uy * sys.dr[1];
        const double bz = rz - (double)uz * sys.dr[2];
        // pair factors: the exponent in FP64, exp in the plan precision
        const double ey = pair_exp<R>((sys.g11 * (by * by)) / ts2);
        const double ez = pair_exp<R>((sys.g22 * (bz * bz)) / ts2);
        const int sy = side_of(uy, n1), sz = side_of(uz, n2);
        const int yd = wrap_idx(uy, n1), zs = wrap_idx(uz, n2);
        // corner images (all three sides off the core) take the reference's
        // z slot: uz + Nb for (y high, z low), uz - Nb for (y low, z high)
        int za = zs;
        if (sy == SIDE_H && sz == SIDE_L) za = uz + nb;
        if (sy == SIDE_L && sz == SIDE_H) za = uz - nb;
        const bool cp = sy != SIDE_C && sz != SIDE_C;
        const long long ostd = (long long)yd * n2 + zs, oalt = (long long)yd * n2 + za;
        const R byr = (R)by;
        R a0 = (R)0, a1 = (R)0;
#pragma unroll 8
        for (int i = 0; i < xw; ++i) {
          const int xdi = rl_i(xd, i);
          const int xni = rl_i(xnc, i);
          const R xvi = rl<R>(xv, i), xbi = rl<R>(xb, i);
          const long long off = (long long)xdi * plane + ((xni && cp) ? oalt : ostd);
          R tv = grho[off];
          if (MODE == MODE_XY) tv *= cross_exp(rl<R>(xk, i), byr);
          a0 += xvi * tv;
          a1 += xbi * tv;
        }
        if (valid) {
          const double inv = -2.0 / ts2;
          const double yz = ey * ez;
          const double d0 = (double)a0, d1 = (double)a1;
          gx += inv * yz * (sys.g00 * d1 + sys.g01 * by * d0);
          gy += inv * yz * (sys.g01 * d1 + sys.g11 * by * d0);
          gz += inv * yz * (sys.g22 * bz * d0);
        }
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    gx += __shfl_down(gx, o, 64);
    gy += __shfl_down(gy, o, 64);
    gz += __shfl_down(gz, o, 64);
  }
  if (lane == 0) {
    part[3 * item + 0] = gx;
    part[3 * item + 1] = gy;
    part[3 * item + 2] = gz;
  }
}

// grad[atom] = scale * (sum of the atom's item partials, in item order);
// parked atoms sum their clipped window here (one thread each).
template <typename R>
__global__ void k_adjoint_reduce(SysDev sys, const TypeDev* __restrict__ td,
                                 const int4* __restrict__ s_ir, const int* __restrict__ s_idx,
                                 const int* __restrict__ nvalid, const double* __restrict__ part,
                                 const double* __restrict__ scale3, const double* __restrict__ s_r,
                                 const R* __restrict__ grho, double* __restrict__ grad) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= *nvalid) return;
  const int4 q = s_ir[s];
  const int t = q.w & 0xffff;
  if (t >= sys.ntypes) return;
  const TypeDev& ty = td[t];
  const int slot = s - ty.type_start;
  if (slot < 0 || slot >= ty.count) return;
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;
  if (q.w & kAtomParked) {
    // window clipped by the padded grid: the gradient over its kept part
    parked_grad<R>(sys, ty, q, s_r[3 * (size_t)s + 0], s_r[3 * (size_t)s + 1],
                   s_r[3 * (size_t)s + 2], grho, 0, 1, g0, g1, g2);
  } else {
    const int per = ty.adj_nxc * ty.adj_npc;
    const long long b = ty.adj_item_base + (long long)slot * per;
    for (int it = 0; it < per; ++it) {
      g0 += part[3 * (b + it) + 0];
      g1 += part[3 * (b + it) + 1];
      g2 += part[3 * (b + it) + 2];
    }
  }
  const double c0 = scale3 ? scale3[0] : 1.0, c1 = scale3 ? scale3[1] : 1.0,
               c2 = scale3 ? scale3[2] : 1.0;
  const int a = s_idx[s];
  grad[3 * (size_t)a + 0] = c0 * g0;
  grad[3 * (size_t)a + 1] = c1 * g1;
  grad[3 * (size_t)a + 2] = c2 * g2;
}

// Tiled adjoint: grad[atom] = scale * (sum of the atom's tile slots).  16
// lanes per atom read its slots coalesced (lane i takes slots i, i + 16, ...
// in order) and a fixed xor-shuffle tree adds the 16 partial sums, so the
// result is deterministic.  Only the slots of tiles the atom's window reaches
// on all three axes are read: exactly the slots the tiled kernel wrote (its
// candidate test is the same window_reaches), so the buffer needs no clearing.
constexpr int kRedLanes = 16;

__global__ void k_adjoint_tile_reduce(SysDev sys, const TypeDev* __restrict__ td,
                                      const int4* __restrict__ s_ir, const int* __restrict__ s_idx,
                                      const int* __restrict__ nvalid,
                                      const double* __restrict__ tpart,
                                      const double* __restrict__ scale3,
                                      const double* __restrict__ s_r,
                                      const float* __restrict__ grho, double* __restrict__ grad) {
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const int s = (int)(g / kRedLanes), li = (int)(g % kRedLanes);
  // every lane of a 16-lane group takes the same branch, so the shuffles
  // below run with the whole group
  bool act = s < *nvalid;
  int t = 0, slot = 0;
  bool parked = false;
  if (act) {
    t = s_ir[s].w & 0xffff;
    act = t < sys.ntypes;
  }
  if (act) {
    slot = s - td[t].type_start;
    act = slot >= 0 && slot < td[t].count;
    parked = (s_ir[s].w & kAtomParked) != 0;
  }
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;
  if (act && parked) {
    // window clipped by the padded grid: the 16 lanes share its kept part
    parked_grad<float>(sys, td[t], s_ir[s], s_r[3 * (size_t)s + 0], s_r[3 * (size_t)s + 1],
                       s_r[3 * (size_t)s + 2], grho, li, kRedLanes, g0, g1, g2);
  } else if (act) {
    const TypeDev& ty = td[t];
    const int4 q = s_ir[s];
    const int per = ty.nsl[0] * ty.nsl[1] * ty.nsl[2];
    const double* src = tpart + 3 * (ty.adjt_base + (long long)slot * per);
    // first tile of the window per axis (the slice numbering of k_slices)
    int first[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int ir = d == 0 ? q.x : (d == 1 ? q.y : q.z);
      int lo = (ir - ty.half[d]) % sys.n[d];
      lo = lo < 0 ? lo + sys.n[d] : lo;
      first[d] = lo / kTile;
    }
    for (int it = li; it < per; it += kRedLanes) {
      const int j2 = it % ty.nsl[2], j01 = it / ty.nsl[2];
      const int j1 = j01 % ty.nsl[1], j0 = j01 / ty.nsl[1];
      bool reached = true;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int ir = d == 0 ? q.x : (d == 1 ? q.y : q.z);
        const int j = d == 0 ? j0 : (d == 1 ? j1 : j2);
        const int tile = (first[d] + j) % sys.ntiles[d];
        const int t0 = tile * kTile, t1 = min(t0 + kTile, sys.n[d]);
        reached = reached && window_reaches(ir, ty.half[d], t0, t1, sys.n[d]);
      }
      if (!reached) continue;
      g0 += src[3 * it + 0];
      g1 += src[3 * it + 1];
      g2 += src[3 * it + 2];
    }
    const int bands = ((q.w >> 16) & 1) ? spill_bands(q, ty, sys) : 0;
    if (bands) {
      // corner atom: the quirk images whose destination z tile the window
      // does not reach (the deposition's spill pass), dL/dr = sum grho dv/dr
      const int N0 = sys.n[0], N1 = sys.n[1], N2 = sys.n[2], nb = sys.nb;
      const double r[3] = {s_r[3 * (size_t)s + 0], s_r[3 * (size_t)s + 1],
                           s_r[3 * (size_t)s + 2]};
      const SpillRange xl = spill_side_l(q.x, ty.half[0]), xh = spill_side_h(q.x, ty.half[0], N0);
      const int nxl = xl.len(), nx = nxl + xh.len();
      for (int band = 0; band < 2; ++band) {
        if (!((bands >> band) & 1)) continue;
        // band 0: y on H, z on L -> z + Nb; band 1: y on L, z on H -> z - Nb
        const SpillRange yr = band == 0 ? spill_side_h(q.y, ty.half[1], N1)
                                        : spill_side_l(q.y, ty.half[1]);
        const SpillRange zr = band == 0 ? spill_side_l(q.z, ty.half[2])
                                        : spill_side_h(q.z, ty.half[2], N2);
        const int ny = yr.len(), nz = zr.len();
        const long long total = (long long)nx * ny * nz;
        for (long long it = li; it < total; it += kRedLanes) {
          const int kz = (int)(it % nz), kxy = (int)(it / nz);
          const int ky = kxy % ny, kx = kxy / ny;
          const int ux = kx < nxl ? xl.lo + kx : xh.lo + (kx - nxl);
          const int uy = yr.lo + ky, uz = zr.lo + kz;
          const int zd = band == 0 ? uz + nb : uz - nb;
          const int t0 = (zd / kTile) * kTile, t1 = min(t0 + kTile, N2);
          if (!spill_tile_unreached(q.z, ty.half[2], t0, t1, N2)) continue;
          const int xd = ux < 0 ? ux + N0 : ux - N0, yd = band == 0 ? uy - N1 : uy + N1;
          const double w = (double)grho[((size_t)xd * N1 + yd) * N2 + zd];
          double dv[3];
          spill_value(sys, ty, r, ux, uy, uz, dv);
          g0 += w * dv[0];
          g1 += w * dv[1];
          g2 += w * dv[2];
        }
      }
    }
  }
#pragma unroll
  for (int o = kRedLanes / 2; o > 0; o >>= 1) {
    g0 += __shfl_xor(g0, o, kRedLanes);
    g1 += __shfl_xor(g1, o, kRedLanes);
    g2 += __shfl_xor(g2, o, kRedLanes);
  }
  if (act && li == 0) {
    const double c0 = scale3 ? scale3[0] : 1.0, c1 = scale3 ? scale3[1] : 1.0,
                 c2 = scale3 ? scale3[2] : 1.0;
    const int a = s_idx[s];
    grad[3 * (size_t)a + 0] = c0 * g0;
    grad[3 * (size_t)a + 1] = c1 * g1;
    grad[3 * (size_t)a + 2] = c2 * g2;
  }
}

__global__ void k_fill_zero(double* __restrict__ p, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// w_main[kz][y][x] = 2 g[x, y, kz] (kz < m), w_nyq[x][y] = 2 g[x, y, m]: the
// layouts the FFT keeps its spectrum and S(q) in
__global__ void k_adjoint_weights(const double* __restrict__ g, double* __restrict__ wmain,
                                  double* __restrict__ wnyq, int nx, int ny, int m) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long nxy = (long long)nx * ny;
  if (i >= nxy * (m + 1)) return;
  const long long xy = i / (m + 1);
  const int kz = (int)(i - xy * (m + 1));
  if (kz < m) {
    const long long x = xy / ny, y = xy - x * ny;
    wmain[((long long)kz * ny + y) * nx + x] = 2.0 * g[i];
  } else {
    wnyq[xy] = 2.0 * g[i];
  }
}

}  // namespace

hipError_t launch_fill_zero(double* p, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_adjoint_weights(Plan& p, const double* dsf, hipStream_t s) {
  const long long nxy = (long long)p.sys.n[0] * p.sys.n[1];
  const int m = p.sys.n[2] / 2;
  const long long n = nxy * (m + 1);
  hipLaunchKernelGGL(k_adjoint_weights, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dsf,
                     p.adj_w, p.adj_w + nxy * m, p.sys.n[0], p.sys.n[1], m);
  return hipGetLastError();
}

hipError_t launch_deposit_adjoint(Plan& p, const void* grho, const double* scale3, double* grad,
                                  hipStream_t s, bool derivs_ready, bool grad_zeroed) {
  const int n = p.sys.natoms;
  if (n == 0) return hipSuccess;
  hipError_t e = hipSuccess;
  if (!grad_zeroed && (e = launch_fill_zero(grad, 3LL * n, s)) != hipSuccess) return e;
  const int* nvalid = p.bucket_off + p.sys.nbuckets;
  if (p.adj_tslots > 0 && p.adj_tpart && p.dslices) {
    // FP32 orthogonal cells: MFMA transpose of the tiled deposition
    if (!derivs_ready && (e = launch_slice_derivs(p, s)) != hipSuccess) return e;
    if ((e = launch_adjoint_tiles(p, static_cast<const float*>(grho), s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_adjoint_tile_reduce,
                       dim3((unsigned)(((long long)n * kRedLanes + 255) / 256)), dim3(256), 0, s,
                       p.sys, p.types, p.s_ir, p.s_idx, nvalid, p.adj_tpart, scale3, p.s_r,
                       static_cast<const float*>(grho), grad);
    return hipGetLastError();
  }
  if (p.adj_items > 0) {
    const dim3 grid((unsigned)((p.adj_items + 3) / 4)), block(kAdjThreads);
    if (p.fft64) {
      if (p.sys.mode == MODE_XY)
        hipLaunchKernelGGL((k_deposit_adjoint<MODE_XY, double>), grid, block, 0, s, p.sys,
                           p.types, p.s_ir, p.s_r, nvalid, static_cast<const double*>(grho),
                           p.adj_part, p.adj_items);
      else
        hipLaunchKernelGGL((k_deposit_adjoint<MODE_DIAG, double>), grid, block, 0, s, p.sys,
                           p.types, p.s_ir, p.s_r, nvalid, static_cast<const double*>(grho),
                           p.adj_part, p.adj_items);
    } else {
      if (p.sys.mode == MODE_XY)
        hipLaunchKernelGGL((k_deposit_adjoint<MODE_XY, float>), grid, block, 0, s, p.sys, p.types,
                           p.s_ir, p.s_r, nvalid, static_cast<const float*>(grho), p.adj_part,
                           p.adj_items);
      else
        hipLaunchKernelGGL((k_deposit_adjoint<MODE_DIAG, float>), grid, block, 0, s, p.sys,
                           p.types, p.s_ir, p.s_r, nvalid, static_cast<const float*>(grho),
                           p.adj_part, p.adj_items);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (p.fft64)
    hipLaunchKernelGGL(k_adjoint_reduce<double>, dim3((n + 255) / 256), dim3(256), 0, s, p.sys,
                       p.types, p.s_ir, p.s_idx, nvalid, p.adj_part, scale3, p.s_r,
                       static_cast<const double*>(grho), grad);
  else
    hipLaunchKernelGGL(k_adjoint_reduce<float>, dim3((n + 255) / 256), dim3(256), 0, s, p.sys,
                       p.types, p.s_ir, p.s_idx, nvalid, p.adj_part, scale3, p.s_r,
                       static_cast<const float*>(grho), grad);
  return hipGetLastError();
}

}  // namespace mdsf
