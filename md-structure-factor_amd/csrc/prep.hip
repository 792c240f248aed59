// Per-frame atom preparation: rescale, periodic wrap, home cells, a
// deterministic counting sort of atoms by (type, home bucket), and the
// per-atom 1-D Gaussian factor tables used by the deposition gather.
//
// Reference behaviour restated here:
//   rescale        dens.py:46-62   r[t,:,i] *= avg(dims)[i] / dims[t,i]   (in place)
//   wrap           dens.py:206-221 r<L ? r : r-L ; r>0 ? r : r+L          (in place)
//   home cell      dens.py:285     ir = int(r / dr)   (FP64 IEEE division, truncation)
//   window         dens.py:287-297 padded indices [ir-A+Nb, ir+A+Nb)
//   Gaussian       dens.py:299-308 (Nel/sig^3) exp(-|(r - x_v) ucell|^2 / (2 sig^2))
#include <climits>

#include "mdsf_internal.h"
#include "spill.h"

namespace mdsf {

namespace {

__device__ __forceinline__ int pos_mod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// T: coordinate storage type; AR: arithmetic type numpy uses for the
// in-place rescale and wrap (float only when coords and dims are float32).
template <typename T, typename AR>
__global__ void k_prep(T* __restrict__ coords, SysDev sys, const TypeDev* __restrict__ td,
                       const int32_t* __restrict__ types, const double* __restrict__ scale,
                       double L0, double L1, double L2, long long wlo, long long whi,
                       int4* __restrict__ ir4, int* __restrict__ key, double* __restrict__ rpos,
                       int* __restrict__ bcount, unsigned* __restrict__ occ,
                       int* __restrict__ home, int* __restrict__ err, int coords_only,
                       int write_back) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sys.natoms) return;
  T x[3];
  x[0] = coords[3 * (size_t)i + 0];
  x[1] = coords[3 * (size_t)i + 1];
  x[2] = coords[3 * (size_t)i + 2];
  const AR box[3] = {(AR)L0, (AR)L1, (AR)L2};
#pragma unroll
  for (int d = 0; d < 3; ++d) x[d] = (T)((AR)x[d] * (AR)scale[d]);
  if (i >= wlo && i < whi) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const AR v = (AR)x[d];
      if (!(v < box[d])) x[d] = (T)(v - box[d]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const AR v = (AR)x[d];
      if (!(v > (AR)0)) x[d] = (T)(v + box[d]);
    }
  }
  if (write_back) {
    coords[3 * (size_t)i + 0] = x[0];
    coords[3 * (size_t)i + 1] = x[1];
    coords[3 * (size_t)i + 2] = x[2];
  }
  if (coords_only) return;

  const int t = types[i];
  if (t < 0 || t >= sys.ntypes) {
    atomicOr(err, ERRBIT_TYPE);
    key[i] = -1;
    return;
  }
  const TypeDev& ty = td[t];
  int ir[3], h[3];
  bool finite = true, parked = false;
  bool nonc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double q = (double)x[d] / sys.dr[d];
    if (q > -1.0e9 && q < 1.0e9) {
      ir[d] = (int)q;  // truncation toward zero, as ndarray.astype(int)
    } else {
      // non-finite or beyond any window: the reference's slices come out
      // empty; this home cell makes the recomputed slice empty too
      finite = false;
      ir[d] = -(1 << 30);
    }
    // The reference adds the window into d0[x0:x1] on each axis, padded
    // indices x0 = ir - A + Nb, x1 = ir + A + Nb of a grid of len = N + 2 Nb,
    // at grid[x0:x1] (dens.py:292-308).  Both are the same numpy slice, so a
    // window that leaves the padded grid is clipped (negative bounds count
    // from the end, then clamp), never an error.  Such atoms are parked:
    // their own bucket, which the tiled passes never enumerate, and the
    // clipped-window passes add them (or nothing, when a slice is empty).
    const int A = ty.half[d];
    const long long len = (long long)sys.n[d] + 2LL * sys.nb;
    const long long x0 = (long long)ir[d] - A + sys.nb, x1 = (long long)ir[d] + A + sys.nb;
    if (A == 0 || x0 < 0 || x1 > len) parked = true;
    nonc[d] = A > 0 && (ir[d] - A < 0 || ir[d] + A > sys.n[d]);
    h[d] = pos_mod(ir[d], sys.n[d]);
    if (home) home[3 * (size_t)i + d] = finite ? ir[d] : INT_MIN;
  }
  const int corner = (!parked && nonc[0] && nonc[1] && nonc[2]) ? kAtomCorner : 0;
  const int b = parked ? ty.park
                       : ty.bucket_base + ((h[0] / ty.bsz[0]) * ty.nbk[1] + h[1] / ty.bsz[1]) *
                                              ty.nbk[2] + h[2] / ty.bsz[2];
  key[i] = b;
  ir4[i] = make_int4(ir[0], ir[1], ir[2], t | corner | (parked ? kAtomParked : 0));
  rpos[3 * (size_t)i + 0] = (double)x[0];
  rpos[3 * (size_t)i + 1] = (double)x[1];
  rpos[3 * (size_t)i + 2] = (double)x[2];
  atomicAdd(&bcount[b], 1);
  if (occ && finite) atomicAdd(&occ[((size_t)h[0] * sys.n[1] + h[1]) * sys.n[2] + h[2]], 1u);
}

// Exclusive scan of the bucket counts in two coalesced passes: per-block
// scans (k_scan_blocks), then block offsets (k_scan_add).  The counts are
// cleared for the next frame and the fill cursors seeded.
constexpr int kScanThreads = 1024;

__device__ __forceinline__ int block_excl_scan(int v, int* wsum, int* total) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  int off = 0, tot = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    const int s = wsum[w];
    off += w < wave ? s : 0;
    tot += s;
  }
  *total = tot;
  return off + x - v;
}

// clear: zero the counts for the slot's next frame (plans with tile lists
// leave them to k_tile_lists, after k_tile_count has read them)
__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(int* __restrict__ bcount,
                                                              int* __restrict__ boff,
                                                              int* __restrict__ aux, int nbk,
                                                              int clear) {
  __shared__ int wsum[kScanThreads / 64];
  const int i = blockIdx.x * kScanThreads + threadIdx.x;
  const int v = i < nbk ? bcount[i] : 0;
  int total;
  const int ex = block_excl_scan(v, wsum, &total);
  if (i < nbk) {
    boff[i] = ex;
    if (clear) bcount[i] = 0;
  }
  if (threadIdx.x == 0) aux[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int* __restrict__ boff,
                                                           int* __restrict__ bcur,
                                                           const int* __restrict__ aux, int nbk,
                                                           int nblocks, int* __restrict__ corners) {
  __shared__ int base_sh;
  if (blockIdx.x == 0 && threadIdx.x == 0) corners[0] = 0;  // refilled by k_rank_gather
  if (threadIdx.x < 64) {
    int s = 0;
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += 64) s += aux[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (threadIdx.x == 0) base_sh = s;
  }
  __syncthreads();
  const int base = base_sh;
  const int i = blockIdx.x * kScanThreads + threadIdx.x;
  if (i < nbk) {
    const int v = boff[i] + base;
    boff[i] = v;
    bcur[i] = v;
  }
  if (blockIdx.x == nblocks - 1 && threadIdx.x == 0) boff[nbk] = base + aux[blockIdx.x];
}

__global__ void k_fill(const int* __restrict__ key, int* __restrict__ bcur,
                       int* __restrict__ s_tmp, int natoms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const int b = key[i];
  if (b < 0) return;
  const int pos = atomicAdd(&bcur[b], 1);
  s_tmp[pos] = i;
}

// Restores atom-index order inside each bucket, so that the gather visits
// atoms in a canonical order (bitwise-reproducible sums): the entry at
// position p of bucket b moves to boff[b] + (number of entries of b with a
// smaller atom index), together with its atom's home cell / type word and
// position; corner atoms are listed (in arrival order; k_corner_order sorts
// the list).  One thread per entry; the entries of a bucket are read by the
// threads of that bucket together (mostly the same wavefront).
__global__ void k_rank_gather(const int* __restrict__ key, const int* __restrict__ boff,
                              const int* __restrict__ s_tmp, int* __restrict__ s_idx,
                              const int4* __restrict__ ir4, const double* __restrict__ rpos,
                              int4* __restrict__ s_ir, double* __restrict__ s_r,
                              const int* __restrict__ nvalid, int* __restrict__ corners) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= *nvalid) return;
  const int i = s_tmp[p];
  const int b = key[i];
  const int lo = boff[b], hi = boff[b + 1];
  int r = 0;
#pragma unroll 4
  for (int q = lo; q < hi; ++q) r += s_tmp[q] < i ? 1 : 0;
  const int s = lo + r;
  s_idx[s] = i;
  const int4 q = ir4[i];
  s_ir[s] = q;
  if ((q.w >> 16) & 1) {
    const int k = atomicAdd(corners, 1);
    if (k < kCornerCap) corners[1 + k] = s;
  }
  s_r[3 * (size_t)s + 0] = rpos[3 * (size_t)i + 0];
  s_r[3 * (size_t)s + 1] = rpos[3 * (size_t)i + 1];
  s_r[3 * (size_t)s + 2] = rpos[3 * (size_t)i + 2];
}

// The corner atoms that spill (spill_bands != 0), in sorted-index order, one
// block: entry v goes to its rank among them, so the spill pass visits them
// in a fixed order whatever order k_rank_gather's atomics gave.  Their count goes
// to corners[1 + 2 kCornerCap].
constexpr int kOrderThreads = 256;
__global__ __launch_bounds__(kOrderThreads) void k_corner_order(SysDev sys,
                                                                const TypeDev* __restrict__ td,
                                                                const int4* __restrict__ s_ir,
                                                                int* __restrict__ corners,
                                                                int* __restrict__ flag,
                                                                int nflag) {
  __shared__ int lst[kCornerCap];
  __shared__ int nrel;
  const int n = min(corners[0], kCornerCap);
  if (threadIdx.x == 0) nrel = 0;
  // spill tile flags (staged spill): cleared, then set per spilling atom below;
  // past kCornerCap corner atoms every tile is flagged and the spill pass
  // enumerates its candidates per tile instead of walking the list
  const bool over = corners[0] > kCornerCap;
  if (flag)
    for (int i = threadIdx.x; i < nflag; i += kOrderThreads) flag[i] = over ? 1 : 0;
  if (over) return;  // uniform
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kOrderThreads) {
    const int c = corners[1 + i];
    const int4 q = s_ir[c];
    const bool spills = spill_bands(q, td[q.w & 0xffff], sys) != 0;
    lst[i] = spills ? c : -1;
    if (spills) atomicAdd(&nrel, 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kOrderThreads) {
    const int v = lst[i];
    if (v < 0) continue;
    int r = 0;
    for (int j = 0; j < n; ++j) r += (lst[j] >= 0 && lst[j] < v) ? 1 : 0;
    corners[1 + kCornerCap + r] = v;
    if (flag) {
      const int4 q = s_ir[v];
      const TypeDev& ty = td[q.w & 0xffff];
      spill_mark(sys, q, ty, spill_bands(q, ty, sys), flag);
    }
  }
  if (threadIdx.x == 0) corners[1 + 2 * kCornerCap] = nrel;
}

// Per-atom 1-D factor tables, one thread per entry:
//   x: amp * exp(-g00 bx^2 / 2s^2), y: exp(-g11 by^2 / 2s^2), z: exp(-g22 bz^2 / 2s^2)
// with b = r - (ir - A + k) dr in FP64 exactly as buf = r - grid (dens.py:252-256,299),
// k in [0, 2A).  Values are rounded once to the plan's precision (FP32 or FP64).
template <typename TT>
__global__ void k_tables(SysDev sys, const TypeDev* __restrict__ td, const int4* __restrict__ s_ir,
                         const double* __restrict__ s_r, TT* __restrict__ tables,
                         long long total) {
#pragma clang fp contract(off)
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (g >= total) return;
  int t = 0;
  while (t + 1 < sys.ntypes && td[t + 1].table_base <= g) ++t;
  const TypeDev& ty = td[t];
  const long long local = g - ty.table_base;
  const int slot = (int)(local / ty.stride);
  const int e = (int)(local - (long long)slot * ty.stride);
  const int s = ty.type_start + slot;
  const int4 q = s_ir[s];
  if (q.w & kAtomParked) return;  // parked atoms have no tables
  const int len0 = 2 * ty.half[0], len1 = 2 * ty.half[1];
  int d, k;
  if (e < len0) {
    d = 0;
    k = e;
  } else if (e < len0 + len1) {
    d = 1;
    k = e - len0;
  } else {
    d = 2;
    k = e - len0 - len1;
  }
  const int ir = d == 0 ? q.x : (d == 1 ? q.y : q.z);
  const int A = d == 0 ? ty.half[0] : (d == 1 ? ty.half[1] : ty.half[2]);
  const double dr = d == 0 ? sys.dr[0] : (d == 1 ? sys.dr[1] : sys.dr[2]);
  const double gm = d == 0 ? sys.g00 : (d == 1 ? sys.g11 : sys.g22);
  const double r = s_r[3 * (size_t)s + d];
  const int u = ir - A + k;
  const double gridv = (double)u * dr;
  const double b = r - gridv;
  const double qv = (gm * (b * b)) / ty.two_sig2;
  double v = exp(-qv);
  if (d == 0) v = ty.amp * v;
  tables[g] = (TT)v;
}

// Tile slices for the FP32 deposition.  16 lanes per (sorted atom, axis),
// lane i owning cell i of every slice, so that each slice row is written as
// one contiguous 64-byte segment.  Slice j holds cell c = 16 T + i of tile
// T = (T0 + j) mod ntiles (T0 = tile of the window start lo mod N); its value
// is the sum over the periodic images u = c + sN (s = -1, 0, 1) inside the
// window [lo, lo + 2A) of amp^[d==0] exp(-g_dd (r - u dr)^2 / 2 s^2): the
// exponent in FP64, exp in FP32 (dens.py:252-256, 299-308).  Cells past N
// are zero.
//
// xy-coupled cells (FP32 plans whose windows meet each x and y tile through
// one image, so at most one of the three terms is nonzero): the factor
// amp exp(-(g00 b_x^2 + 2 g01 b_x b_y + g11 b_y^2) / 2 s^2) of a cell pair is
// X(x) 2^(LY(y) + CX(x) BY(y)), with X as above, and in the y slices (regular
// and corner components) LY = log2 of the y factor (-inf outside the window)
// instead of the factor; bslices holds CX = -2 g01 log2(e) / 2 s^2 b_x at the
// regular x slice cells and BY = b_y at the regular y slice cells (0 outside
// the window).
__global__ void k_slices(SysDev sys, const TypeDev* __restrict__ td, const int4* __restrict__ s_ir,
                         const double* __restrict__ s_r, const int* __restrict__ nvalid,
                         float* __restrict__ slices, float* __restrict__ dslices,
                         float* __restrict__ bslices, int mode) {
#pragma clang fp contract(off)
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const int pair = (int)(g / kTile), i = (int)(g - (long long)pair * kTile);
  const int s = pair / 3, d = pair - 3 * (pair / 3);
  if (s >= *nvalid) return;
  const int4 q = s_ir[s];
  const int t = q.w & 0xffff;
  if (t >= sys.ntypes || (q.w & kAtomParked)) return;  // parked atoms have no slices
  const TypeDev& ty = td[t];
  const int slot = s - ty.type_start;
  if (slot < 0 || slot >= ty.count) return;  // only after a reported error
  const int N = d == 0 ? sys.n[0] : (d == 1 ? sys.n[1] : sys.n[2]);
  const int nt = d == 0 ? sys.ntiles[0] : (d == 1 ? sys.ntiles[1] : sys.ntiles[2]);
  const int A = d == 0 ? ty.half[0] : (d == 1 ? ty.half[1] : ty.half[2]);
  const int nsl = d == 0 ? ty.nsl[0] : (d == 1 ? ty.nsl[1] : ty.nsl[2]);
  const int ir = d == 0 ? q.x : (d == 1 ? q.y : q.z);
  const double dr = d == 0 ? sys.dr[0] : (d == 1 ? sys.dr[1] : sys.dr[2]);
  const double gm = d == 0 ? sys.g00 : (d == 1 ? sys.g11 : sys.g22);
  const double gk = gm / ty.two_sig2;
  const float am = d == 0 ? ty.amp_f : 1.f;
  const double r = s_r[3 * (size_t)s + d];
  const int lo = ir - A, hi = ir + A;
  const int off_d = d == 0 ? 0 : (d == 1 ? kTile * ty.nsl[0] : kTile * (ty.nsl[0] + ty.nsl[1]));
  const long long aoff = ty.slice_base + (long long)slot * ty.slice_stride + off_d + i;
  const bool corner = ((q.w >> 16) & 1) != 0;
  const int coff_d = d == 0 ? 0 : (d == 1 ? 2 * kTile * ty.nsl[0]
                                          : 2 * kTile * ty.nsl[0] + 3 * kTile * ty.nsl[1]);
  const long long acoff = ty.corner_base + (long long)slot * ty.corner_stride + coff_d + i;
  // mode 0: slices, 1: slices and derivatives, 2: derivatives only
  const bool wv = mode != 2, wd = mode != 0;
  // factor of the image at grid coordinate u, 0 outside the window, and its
  // derivative in r: am exp(-gk b^2) and am exp(-gk b^2) (-2 gk b), b = r - u dr
  auto f = [&](int u, float& df) -> float {
    df = 0.f;
    if (u < lo || u >= hi) return 0.f;
    const double b = r - (double)u * dr;
    const float v = am * expf(-(float)(gk * (b * b)));
    df = v * (float)(-2.0 * gk * b);
    return v;
  };
  auto put = [&](long long o, float v, float dv) {
    if (wv) slices[o] = v;
    if (wd) dslices[o] = dv;
  };
  // xy-coupled cells (the branch is uniform across the grid; within it only
  // the bslices store depends on the axis)
  const bool xy = sys.mode == MODE_XY;
  const bool ylog = xy && d == 1;
  constexpr double kLog2e = 1.4426950408889634;
  const double cxy = -2.0 * sys.g01 * kLog2e / ty.two_sig2;
  auto inw = [&](int u) { return u >= lo && u < hi; };
  // log2 of the y factor of the image at u, -inf outside the window
  auto lg = [&](int u) -> float {
    if (!inw(u)) return -__builtin_inff();
    const double b = r - (double)u * dr;
    return (float)(-(gk * (b * b)) * kLog2e);
  };
  int T = pos_mod(lo, N) / kTile;
  for (int j = 0; j < nsl; ++j) {
    const int c = T * kTile + i;
    const bool in = c < N;
    float dL = 0.f, dC = 0.f, dH = 0.f;
    float vL = in ? f(c - N, dL) : 0.f, vC = in ? f(c, dC) : 0.f, vH = in ? f(c + N, dH) : 0.f;
    if (xy) {
      // at most one image in the window (the plan's condition): its b gives
      // CX (x) or BY (y); y values become log2 (then the regular value, the
      // one image's, is the largest of the three)
      const int us = inw(c) ? c : (inw(c - N) ? c - N : c + N);
      const bool any = in && inw(us);
      const double bs = r - (double)us * dr;
      if (d < 2) bslices[aoff + j * kTile] = any ? (float)(d == 0 ? cxy * bs : bs) : 0.f;
      if (ylog) {
        vL = in ? lg(c - N) : -__builtin_inff();
        vC = in ? lg(c) : -__builtin_inff();
        vH = in ? lg(c + N) : -__builtin_inff();
      }
    }
    put(aoff + j * kTile, ylog ? fmaxf(fmaxf(vL, vC), vH) : vC + vL + vH, dC + dL + dH);
    if (corner) {
      // the components the corner virtual atoms take (dens.py:101-107): on z
      // the slots where the reference's corner copy lands for y on L / on H
      if (d == 0) {
        put(acoff + j * kTile, vC, dC);
        put(acoff + (nsl + j) * kTile, vL + vH, dL + dH);
      } else if (d == 1) {
        put(acoff + j * kTile, vC, dC);
        put(acoff + (nsl + j) * kTile, vL, dL);
        put(acoff + (2 * nsl + j) * kTile, vH, dH);
      } else {
        const int u1 = c + sys.nb, u2 = c - sys.nb;
        float d1 = 0.f, d2 = 0.f;
        const float v1 = in && u1 >= N ? f(u1, d1) : 0.f;
        const float v2 = in && u2 < 0 ? f(u2, d2) : 0.f;
        put(acoff + j * kTile, vC + vL + v1, dC + dL + d1);
        put(acoff + (nsl + j) * kTile, vC + vH + v2, dC + dH + d2);
      }
    }
    T = T + 1 == nt ? 0 : T + 1;
  }
}

}  // namespace

hipError_t launch_prep(Plan& p, void* coords, int dtype, int arith, const int32_t* types,
                       const double* scale3, const double box[3], int64_t wrap_lo,
                       int64_t wrap_hi, uint32_t* occ, int32_t* home, hipStream_t s,
                       bool coords_only, bool write_back) {
  const int n = p.sys.natoms;
  const int co = coords_only ? 1 : 0;
  const int wb = write_back ? 1 : 0;
  if (n == 0) return hipSuccess;
  const dim3 grid((n + 255) / 256), block(256);
  if (dtype == MDSF_F32 && arith == MDSF_F32) {
    hipLaunchKernelGGL((k_prep<float, float>), grid, block, 0, s, (float*)coords, p.sys, p.types,
                       types, scale3, box[0], box[1], box[2], (long long)wrap_lo,
                       (long long)wrap_hi, p.ir4, p.key, p.rpos, p.bucket_count, occ, home, p.err, co, wb);
  } else if (dtype == MDSF_F32) {
    hipLaunchKernelGGL((k_prep<float, double>), grid, block, 0, s, (float*)coords, p.sys, p.types,
                       types, scale3, box[0], box[1], box[2], (long long)wrap_lo,
                       (long long)wrap_hi, p.ir4, p.key, p.rpos, p.bucket_count, occ, home, p.err, co, wb);
  } else {
    hipLaunchKernelGGL((k_prep<double, double>), grid, block, 0, s, (double*)coords, p.sys,
                       p.types, types, scale3, box[0], box[1], box[2], (long long)wrap_lo,
                       (long long)wrap_hi, p.ir4, p.key, p.rpos, p.bucket_count, occ, home, p.err, co, wb);
  }
  return hipGetLastError();
}

hipError_t launch_rescale(void* coords, int dtype, int arith, long long nframes, int natoms,
                          const double* scale, hipStream_t s) {
  if (natoms == 0 || nframes == 0) return hipSuccess;
  SysDev sys{};
  sys.natoms = natoms;
  const dim3 grid((natoms + 255) / 256), block(256);
  const size_t fstride = (size_t)natoms * 3;
  for (long long t = 0; t < nframes; ++t) {
    // wrap range [0, 0): the coordinates are only rescaled (dens.py:54-58)
    if (dtype == MDSF_F32 && arith == MDSF_F32)
      hipLaunchKernelGGL((k_prep<float, float>), grid, block, 0, s, (float*)coords + t * fstride,
                         sys, nullptr, nullptr, scale + 3 * t, 0.0, 0.0, 0.0, 0LL, 0LL, nullptr,
                         nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 1, 1);
    else if (dtype == MDSF_F32)
      hipLaunchKernelGGL((k_prep<float, double>), grid, block, 0, s, (float*)coords + t * fstride,
                         sys, nullptr, nullptr, scale + 3 * t, 0.0, 0.0, 0.0, 0LL, 0LL, nullptr,
                         nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 1, 1);
    else
      hipLaunchKernelGGL((k_prep<double, double>), grid, block, 0, s,
                         (double*)coords + t * fstride, sys, nullptr, nullptr, scale + 3 * t, 0.0,
                         0.0, 0.0, 0LL, 0LL, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr, 1, 1);
  }
  return hipGetLastError();
}

hipError_t launch_slice_derivs(Plan& p, hipStream_t s) {
  const int n = p.sys.natoms;
  if (n == 0 || p.slice_floats == 0 || !p.dslices) return hipSuccess;
  hipLaunchKernelGGL(k_slices, dim3((unsigned)((3LL * kTile * n + 255) / 256)), dim3(256), 0, s,
                     p.sys, p.types, p.s_ir, p.s_r, p.bucket_off + p.sys.nbuckets, p.slices,
                     p.dslices, p.bslices, 2);
  return hipGetLastError();
}

hipError_t launch_binning(Plan& p, hipStream_t s, bool derivs, hipStream_t aux,
                          hipStream_t side, bool join, hipEvent_t aux_end) {
  const int n = p.sys.natoms;
  const int nbk = p.sys.nbuckets;
  const int nblk = (nbk + kScanThreads - 1) / kScanThreads;
  if (nblk == 0) return hipSuccess;
  // Branches (when the streams are given): the tile lists on `aux` (their
  // regions need only the bucket counts, from here on; their entries the
  // sorted atoms), the corner list and the spill staging on `side` (the
  // stream whose deposition reads them), the bucket scan, the sort and the
  // slices here.
  const bool lists = p.tl_cap > 0 && n > 0;
  const bool split = aux && lists;
  hipError_t e = hipSuccess;
  if (split) {
    if ((e = hipEventRecord(p.ev_lfork, s)) != hipSuccess ||
        (e = hipStreamWaitEvent(aux, p.ev_lfork, 0)) != hipSuccess ||
        (e = launch_tile_regions(p, aux)) != hipSuccess)
      return e;
  }
  hipLaunchKernelGGL(k_scan_blocks, dim3(nblk), dim3(kScanThreads), 0, s, p.bucket_count,
                     p.bucket_off, p.bucket_aux, nbk, lists ? 0 : 1);
  hipLaunchKernelGGL(k_scan_add, dim3(nblk), dim3(kScanThreads), 0, s, p.bucket_off,
                     p.bucket_cur, p.bucket_aux, nbk, nblk, p.corners);
  if (n == 0) return hipGetLastError();
  hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, p.key, p.bucket_cur,
                     p.s_tmp, n);
  hipLaunchKernelGGL(k_rank_gather, dim3((n + 255) / 256), dim3(256), 0, s, p.key,
                     p.bucket_off, p.s_tmp, p.s_idx, p.ir4, p.rpos, p.s_ir, p.s_r,
                     p.bucket_off + nbk, p.corners);
  if (split || side) {
    if ((e = hipEventRecord(p.ev_lgath, s)) != hipSuccess) return e;
  }
  if (split) {
    if ((e = hipStreamWaitEvent(aux, p.ev_lgath, 0)) != hipSuccess ||
        (e = launch_tile_lists(p, aux)) != hipSuccess ||
        (e = hipEventRecord(p.ev_ljoin, aux)) != hipSuccess)
      return e;
    if (aux_end && (e = hipEventRecord(aux_end, aux)) != hipSuccess) return e;
  } else {
    if ((e = launch_tile_regions(p, s)) != hipSuccess || (e = launch_tile_lists(p, s)) != hipSuccess)
      return e;
  }
  const hipStream_t cs = side ? side : s;
  if (side && (e = hipStreamWaitEvent(side, p.ev_lgath, 0)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_corner_order, dim3(1), dim3(kOrderThreads), 0, cs, p.sys, p.types, p.s_ir,
                     p.corners, p.spill ? p.spill_flag : nullptr,
                     (int)(p.spill ? p.spill_tiles : 0));
  // the corner-fold spill, staged for the deposition of this frame
  if ((e = launch_spill_stage(p, cs)) != hipSuccess) return e;
  if (p.slice_floats > 0)
    hipLaunchKernelGGL(k_slices, dim3((unsigned)((3LL * kTile * n + 255) / 256)), dim3(256), 0, s,
                       p.sys, p.types, p.s_ir, p.s_r, p.bucket_off + nbk, p.slices, p.dslices,
                       p.bslices, (derivs && p.dslices) ? 1 : 0);
  // factor tables feed the FP64 gather
  const long long total = p.fft64 ? p.table_floats : 0;
  if (total > 0) {
    const dim3 grid((unsigned)((total + 255) / 256)), block(256);
    if (p.fft64)
      hipLaunchKernelGGL(k_tables<double>, grid, block, 0, s, p.sys, p.types, p.s_ir, p.s_r,
                         static_cast<double*>(p.tables), total);
    else
      hipLaunchKernelGGL(k_tables<float>, grid, block, 0, s, p.sys, p.types, p.s_ir, p.s_r,
                         static_cast<float*>(p.tables), total);
  }
  if (split && join && (e = hipStreamWaitEvent(s, p.ev_ljoin, 0)) != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace mdsf
