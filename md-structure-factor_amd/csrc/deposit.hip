// Gaussian deposition + periodic border fold as one output-stationary gather.
//
// Replaces dens.py:283-311: the per-atom window sums into the padded grid
// d0 (dens.py:292-308) followed by remap_grid_tcl (dens.py:86-108).  Instead
// of a padded grid, every padded index p = u + Nb (u in [ir-A, ir+A)) is sent
// straight to its folded destination:
//   standard      (u mod N) on each axis, side(u) = L (u<0), C, H (u>=N);
//   corner quirk  when no axis is C, dens.py:107 takes the z slot from the y
//                 side: z lands on u_z - Nb (y on L, z on H) or u_z + Nb
//                 (y on H, z on L) instead of u_z mod N.
// A tile cell c receives, per axis, at most one image per side:
// u = c - N (L), c (C), c + N (H).
//
// One workgroup (4 waves) owns a 16x16x16 tile of the periodic grid.  Each
// wave owns an 8x8 block of (x, y) columns, each lane one column with 16 z
// accumulators in registers, so the tile is written once with plain stores:
// no global atomics.  Candidate atoms come from a per-type bucket grid
// (sorted by (type, home bucket), atom order inside a bucket): all bucket
// columns that can reach the tile are resolved in one parallel step (one
// load per thread + a block scan), candidates are streamed 256 at a time,
// tested exactly and compacted in order, and their 1-D factor slices for the
// tile are staged in LDS.  The visit order is fixed, so results are bitwise
// reproducible.  ACC is the arithmetic and output type (float or double).
#include <type_traits>

#include "mdsf_internal.h"

namespace mdsf {

namespace {

constexpr int T = kTile;
constexpr int NB = kBatch;
constexpr int kThreads = 256;

__device__ __forceinline__ int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// Bucket index ranges along one axis whose atoms can reach tile [t0, t1).
__device__ __forceinline__ int bucket_ranges(int t0, int t1, int A, int N, int B, int nbk,
                                             int* lo, int* hi) {
  const int len = t1 - t0 + 2 * A - 1;
  if (len >= N) {
    lo[0] = 0;
    hi[0] = nbk - 1;
    return 1;
  }
  const int a = pmod(t0 - A + 1, N), b = a + len - 1;
  if (b < N) {
    lo[0] = a / B;
    hi[0] = b / B;
    return 1;
  }
  const int b0 = a / B, e1 = (b - N) / B;
  if (e1 >= b0) {
    lo[0] = 0;
    hi[0] = nbk - 1;
    return 1;
  }
  lo[0] = b0;
  hi[0] = nbk - 1;
  lo[1] = 0;
  hi[1] = e1;
  return 2;
}

__device__ __forceinline__ bool overlap1(int ir, int A, int t0, int t1, int N) {
  const int lo = ir - A, hi = ir + A;
#pragma unroll
  for (int j = -1; j <= 1; ++j) {
    const int a = max(lo, t0 + j * N), b = min(hi, t1 + j * N);
    if (a < b) return true;
  }
  return false;
}

// Per-type bucket ranges for one tile.
struct TypeRanges {
  int lo[3][2], hi[3][2], rn[3];
  int nx, ny, ncols;
};

__device__ __forceinline__ void type_ranges(const TypeDev& ty, const SysDev& sys, const int* t0,
                                            const int* t1, TypeRanges& r) {
  if (ty.count == 0 || ty.half[0] == 0 || ty.half[1] == 0 || ty.half[2] == 0) {
    r.ncols = 0;
    r.nx = r.ny = 0;
    r.rn[2] = 0;
    return;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
    r.rn[d] = bucket_ranges(t0[d], t1[d], ty.half[d], sys.n[d], ty.bsz[d], ty.nbk[d], r.lo[d],
                            r.hi[d]);
  r.nx = (r.hi[0][0] - r.lo[0][0] + 1) + (r.rn[0] == 2 ? r.hi[0][1] - r.lo[0][1] + 1 : 0);
  r.ny = (r.hi[1][0] - r.lo[1][0] + 1) + (r.rn[1] == 2 ? r.hi[1][1] - r.lo[1][1] + 1 : 0);
  r.ncols = r.nx * r.ny * r.rn[2];
}

__device__ __forceinline__ int range_item(const TypeRanges& r, int d, int i) {
  const int len0 = r.hi[d][0] - r.lo[d][0] + 1;
  return i < len0 ? r.lo[d][0] + i : r.lo[d][1] + (i - len0);
}

template <typename ACC>
struct StageDiag {
  // x0 = C part (or all images), x1 = L+H part; y0 = C (or all), y1 = L, y2 = H;
  // z0 = standard slots, z1 / z2 = slots for y on L / y on H (corner quirk).
  __align__(16) ACC x0[NB][T];
  __align__(16) ACC x1[NB][T];
  __align__(16) ACC y0[NB][T];
  __align__(16) ACC y1[NB][T];
  __align__(16) ACC y2[NB][T];
  __align__(16) ACC z0[NB][T];
  __align__(16) ACC z1[NB][T];
  __align__(16) ACC z2[NB][T];
  int xm[NB], ym[NB], zm[NB], cf[NB];
};

template <typename ACC>
struct StageXY {
  // b[side][slot][cell] = r - u * dr of the image on that side (valid bits in bv)
  __align__(16) ACC bx[3][NB][T];
  __align__(16) ACC by[3][NB][T];
  __align__(16) ACC z0[NB][T];
  __align__(16) ACC z1[NB][T];
  __align__(16) ACC z2[NB][T];
  int xv[NB][T], yv[NB][T];
  int xm[NB], ym[NB], zm[NB], cf[NB];
  ACC amp[NB], k2[NB];
};

struct Tile {
  int t0[3], t1[3];
};

// value of a 3-array at a runtime axis without dynamic private indexing
template <typename V>
__device__ __forceinline__ V pick3(const V* v, int d) {
  return d == 0 ? v[0] : (d == 1 ? v[1] : v[2]);
}

template <typename ACC>
__device__ __forceinline__ const ACC* atom_tables(const void* tables, const TypeDev& ty, int slot) {
  return static_cast<const ACC*>(tables) + ty.table_base + (long long)slot * ty.stride;
}

// Stage one atom's slices for the tile (DIAG).  part 0: x, 1: y, 2: z, 3: flag.
template <typename ACC>
__device__ void stage_diag(StageDiag<ACC>& st, int a, int part, int c, const SysDev& sys,
                           const TypeDev* __restrict__ td, const int4* __restrict__ s_ir,
                           const void* __restrict__ tables, const Tile& tl) {
  const int4 q = s_ir[c];
  const int t = q.w & 0xffff;
  const int corner = q.w >> 16;
  const TypeDev& ty = td[t];
  const int slot = c - ty.type_start;
  if (slot < 0 || slot >= ty.count) {  // only after a reported range / type error
    if (part == 0) st.xm[a] = 0;
    if (part == 1) st.ym[a] = 0;
    if (part == 2) st.zm[a] = 0;
    if (part == 3) st.cf[a] = 0;
    return;
  }
  if (part == 3) {
    st.cf[a] = corner;
    return;
  }
  const int d = part;
  const int N = pick3(sys.n, d);
  const int A = pick3(ty.half, d);
  const int ir = d == 0 ? q.x : (d == 1 ? q.y : q.z);
  const int lo = ir - A, len = 2 * A;
  const ACC* tb = atom_tables<ACC>(tables, ty, slot) +
                  (d == 0 ? 0 : (d == 1 ? 2 * ty.half[0] : 2 * ty.half[0] + 2 * ty.half[1]));
  const int t0 = pick3(tl.t0, d), tn = pick3(tl.t1, d) - t0;
  // k of cell i for the L / C / H image
  const int kL = t0 - N - lo, kC = t0 - lo, kH = t0 + N - lo;
  int mask = 0;
  if (d < 2) {
    ACC* o0 = d == 0 ? st.x0[a] : st.y0[a];
#pragma unroll
    for (int i = 0; i < T; ++i) {
      const bool in = i < tn;
      const ACC vL = (in && kL + i >= 0 && kL + i < len) ? tb[kL + i] : (ACC)0;
      const ACC vC = (in && kC + i >= 0 && kC + i < len) ? tb[kC + i] : (ACC)0;
      const ACC vH = (in && kH + i >= 0 && kH + i < len) ? tb[kH + i] : (ACC)0;
      const bool hit = in && ((kL + i >= 0 && kL + i < len) || (kC + i >= 0 && kC + i < len) ||
                              (kH + i >= 0 && kH + i < len));
      if (hit) mask |= 1 << i;
      if (!corner) {
        o0[i] = vC + vL + vH;
      } else if (d == 0) {
        st.x0[a][i] = vC;
        st.x1[a][i] = vL + vH;
      } else {
        st.y0[a][i] = vC;
        st.y1[a][i] = vL;
        st.y2[a][i] = vH;
      }
    }
    if (d == 0) st.xm[a] = mask;
    else st.ym[a] = mask;
    return;
  }
  const int nb = sys.nb;
  const int kSL = t0 + nb - lo;  // z on H, y on L: cell c holds u = c + Nb (u >= N)
  const int kSH = t0 - nb - lo;  // z on L, y on H: cell c holds u = c - Nb (u < 0)
#pragma unroll
  for (int i = 0; i < T; ++i) {
    const bool in = i < tn;
    const bool bL = in && kL + i >= 0 && kL + i < len;
    const bool bC = in && kC + i >= 0 && kC + i < len;
    const bool bH = in && kH + i >= 0 && kH + i < len;
    const ACC vL = bL ? tb[kL + i] : (ACC)0;
    const ACC vC = bC ? tb[kC + i] : (ACC)0;
    const ACC vH = bH ? tb[kH + i] : (ACC)0;
    bool hit = bL || bC || bH;
    st.z0[a][i] = vC + vL + vH;
    if (corner) {
      const int u1 = t0 + i + nb, u2 = t0 + i - nb;
      const bool s1 = in && u1 >= N && kSL + i >= 0 && kSL + i < len;
      const bool s2 = in && u2 < 0 && kSH + i >= 0 && kSH + i < len;
      const ACC v1 = s1 ? tb[kSL + i] : (ACC)0;
      const ACC v2 = s2 ? tb[kSH + i] : (ACC)0;
      st.z1[a][i] = vC + vL + v1;
      st.z2[a][i] = vC + vH + v2;
      hit = hit || s1 || s2;
    }
    if (hit) mask |= 1 << (i / 4);
  }
  st.zm[a] = mask;
}

// Stage one atom for the xy-coupled cell (XY): b values of each image per side.
template <typename ACC>
__device__ void stage_xy(StageXY<ACC>& st, int a, int part, int c, const SysDev& sys,
                         const TypeDev* __restrict__ td, const int4* __restrict__ s_ir,
                         const double* __restrict__ s_r, const void* __restrict__ tables,
                         const Tile& tl) {
  const int4 q = s_ir[c];
  const int t = q.w & 0xffff;
  const int corner = q.w >> 16;
  const TypeDev& ty = td[t];
  const int slot = c - ty.type_start;
  if (slot < 0 || slot >= ty.count) {
    if (part == 0) st.xm[a] = 0;
    if (part == 1) st.ym[a] = 0;
    if (part == 2) st.zm[a] = 0;
    if (part == 3) st.cf[a] = 0;
    return;
  }
  if (part == 3) {
    st.cf[a] = corner;
    st.amp[a] = (ACC)ty.amp;
    // exponent scale: exp(-q / 2s^2) in FP64, exp2(-q * log2(e) / 2s^2) in FP32
    st.k2[a] = std::is_same<ACC, double>::value ? (ACC)(1.0 / ty.two_sig2) : (ACC)ty.c2;
    return;
  }
  if (part == 2) {
    // z factor exp(-g22 bz^2 / 2s^2) from the tables; same slots as DIAG
    const int N = sys.n[2];
    const int A = ty.half[2];
    const int lo = q.z - A, len = 2 * A;
    const ACC* tb = atom_tables<ACC>(tables, ty, slot) + 2 * ty.half[0] + 2 * ty.half[1];
    const int t0 = tl.t0[2], tn = tl.t1[2] - tl.t0[2];
    const int kL = t0 - N - lo, kC = t0 - lo, kH = t0 + N - lo;
    const int nb = sys.nb;
    const int kSL = t0 + nb - lo, kSH = t0 - nb - lo;
    int mask = 0;
#pragma unroll
    for (int i = 0; i < T; ++i) {
      const bool in = i < tn;
      const bool bL = in && kL + i >= 0 && kL + i < len;
      const bool bC = in && kC + i >= 0 && kC + i < len;
      const bool bH = in && kH + i >= 0 && kH + i < len;
      const ACC vL = bL ? tb[kL + i] : (ACC)0;
      const ACC vC = bC ? tb[kC + i] : (ACC)0;
      const ACC vH = bH ? tb[kH + i] : (ACC)0;
      bool hit = bL || bC || bH;
      st.z0[a][i] = vC + vL + vH;
      if (corner) {
        const int u1 = t0 + i + nb, u2 = t0 + i - nb;
        const bool s1 = in && u1 >= N && kSL + i >= 0 && kSL + i < len;
        const bool s2 = in && u2 < 0 && kSH + i >= 0 && kSH + i < len;
        st.z1[a][i] = vC + vL + (s1 ? tb[kSL + i] : (ACC)0);
        st.z2[a][i] = vC + vH + (s2 ? tb[kSH + i] : (ACC)0);
        hit = hit || s1 || s2;
      }
      if (hit) mask |= 1 << (i / 4);
    }
    st.zm[a] = mask;
    return;
  }
  const int d = part;
  const int N = d == 0 ? sys.n[0] : sys.n[1];
  const int A = d == 0 ? ty.half[0] : ty.half[1];
  const int ir = d == 0 ? q.x : q.y;
  const int lo = ir - A, len = 2 * A;
  const double r = s_r[3 * (size_t)c + d];
  const double dr = d == 0 ? sys.dr[0] : sys.dr[1];
  const int t0 = d == 0 ? tl.t0[0] : tl.t0[1];
  const int tn = (d == 0 ? tl.t1[0] : tl.t1[1]) - t0;
  ACC(*b)[NB][T] = d == 0 ? st.bx : st.by;
  int(*v)[T] = d == 0 ? st.xv : st.yv;
  int mask = 0;
#pragma unroll
  for (int i = 0; i < T; ++i) {
    int bits = 0;
#pragma unroll
    for (int sd = 0; sd < 3; ++sd) {
      const int u = t0 + i + (sd - 1) * N;
      const int k = u - lo;
      const bool ok = i < tn && k >= 0 && k < len;
      b[sd][a][i] = ok ? (ACC)(r - (double)u * dr) : (ACC)0;
      bits |= ok ? (1 << sd) : 0;
    }
    v[a][i] = bits;
    if (bits) mask |= 1 << i;
  }
  if (d == 0) st.xm[a] = mask;
  else st.ym[a] = mask;
}

template <typename ACC>
__device__ __forceinline__ void fma4(ACC* acc, ACC w, const ACC* z) {
#pragma unroll
  for (int k = 0; k < 4; ++k) acc[k] = fma(w, z[k], acc[k]);
}

template <typename ACC>
__device__ __forceinline__ void load4(const ACC* src, ACC* z) {
  if constexpr (std::is_same<ACC, float>::value) {
    const float4 v = *reinterpret_cast<const float4*>(src);
    z[0] = v.x;
    z[1] = v.y;
    z[2] = v.z;
    z[3] = v.w;
  } else {
    const double2 v0 = *reinterpret_cast<const double2*>(src);
    const double2 v1 = *reinterpret_cast<const double2*>(src + 2);
    z[0] = v0.x;
    z[1] = v0.y;
    z[2] = v1.x;
    z[3] = v1.y;
  }
}

template <typename ACC>
__device__ __forceinline__ void zapply(ACC* acc, int zm, int corner, ACC ws, ACC wl, ACC wh,
                                       const ACC* z0, const ACC* z1, const ACC* z2) {
#pragma unroll
  for (int g = 0; g < T / 4; ++g) {
    if ((zm >> g) & 1) {
      ACC z[4];
      load4(z0 + 4 * g, z);
      fma4(acc + 4 * g, ws, z);
      if (corner) {
        load4(z1 + 4 * g, z);
        fma4(acc + 4 * g, wl, z);
        load4(z2 + 4 * g, z);
        fma4(acc + 4 * g, wh, z);
      }
    }
  }
}

template <typename ACC>
__device__ void accumulate_diag(const StageDiag<ACC>& st, int nbatch, int wave, int ci, int cj,
                                ACC* acc) {
  const int wsh_x = (wave & 1) * 8, wsh_y = (wave >> 1) * 8;
  for (int a = 0; a < nbatch; ++a) {
    const int xm = __builtin_amdgcn_readfirstlane(st.xm[a]);
    const int ym = __builtin_amdgcn_readfirstlane(st.ym[a]);
    if (((xm >> wsh_x) & 0xff) == 0 || ((ym >> wsh_y) & 0xff) == 0) continue;
    const int zm = __builtin_amdgcn_readfirstlane(st.zm[a]);
    const int corner = __builtin_amdgcn_readfirstlane(st.cf[a]);
    ACC ws, wl = 0, wh = 0;
    if (!corner) {
      ws = st.x0[a][ci] * st.y0[a][cj];
    } else {
      const ACC xc = st.x0[a][ci], xn = st.x1[a][ci];
      const ACC yc = st.y0[a][cj], yl = st.y1[a][cj], yh = st.y2[a][cj];
      ws = xc * (yc + yl + yh) + xn * yc;
      wl = xn * yl;
      wh = xn * yh;
    }
    zapply(acc, zm, corner, ws, wl, wh, st.z0[a], st.z1[a], st.z2[a]);
  }
}

template <typename ACC>
__device__ void accumulate_xy(const StageXY<ACC>& st, int nbatch, int wave, int ci, int cj,
                              ACC* acc, ACC g00, ACC g01x2, ACC g11) {
  const int wsh_x = (wave & 1) * 8, wsh_y = (wave >> 1) * 8;
  for (int a = 0; a < nbatch; ++a) {
    const int xm = __builtin_amdgcn_readfirstlane(st.xm[a]);
    const int ym = __builtin_amdgcn_readfirstlane(st.ym[a]);
    if (((xm >> wsh_x) & 0xff) == 0 || ((ym >> wsh_y) & 0xff) == 0) continue;
    const int zm = __builtin_amdgcn_readfirstlane(st.zm[a]);
    const int corner = __builtin_amdgcn_readfirstlane(st.cf[a]);
    const int xv = st.xv[a][ci], yv = st.yv[a][cj];
    const ACC amp = st.amp[a], k2 = st.k2[a];
    ACC ws = 0, wl = 0, wh = 0;
#pragma unroll
    for (int sx = 0; sx < 3; ++sx) {
      if (!((xv >> sx) & 1)) continue;
      const ACC bx = st.bx[sx][a][ci];
#pragma unroll
      for (int sy = 0; sy < 3; ++sy) {
        if (!((yv >> sy) & 1)) continue;
        const ACC by = st.by[sy][a][cj];
        const ACC qf = fma(g00 * bx, bx, fma(g01x2 * bx, by, g11 * by * by));
        ACC v;
        if constexpr (std::is_same<ACC, double>::value) v = amp * exp(-qf * k2);
        else v = amp * exp2f(-qf * k2);
        if (!corner || sx == SIDE_C || sy == SIDE_C) ws += v;
        else if (sy == SIDE_L) wl += v;
        else wh += v;
      }
    }
    zapply(acc, zm, corner, ws, wl, wh, st.z0[a], st.z1[a], st.z2[a]);
  }
}

template <int MODE, typename ACC>
__global__ __launch_bounds__(kThreads) void k_deposit(SysDev sys, const TypeDev* __restrict__ td,
                                                      const int* __restrict__ boff,
                                                      const int4* __restrict__ s_ir,
                                                      const double* __restrict__ s_r,
                                                      const void* __restrict__ tables,
                                                      ACC* __restrict__ rho) {
  using Stage = typename std::conditional<MODE == MODE_DIAG, StageDiag<ACC>, StageXY<ACC>>::type;
  __shared__ Stage st;
  __shared__ int list[NB + kThreads];
  __shared__ int wsum[kThreads / 64];
  __shared__ int col_s[kThreads], col_pre[kThreads + 1], col_t[kThreads];
  __shared__ int sh_half[kMaxTypes][3];

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  // XCD-aware order: consecutive tiles (neighbours in z) share an XCD's L2.
  const int nblk = gridDim.x, bid = blockIdx.x;
  const int q8 = nblk / 8;
  const int tile = bid < 8 * q8 ? (bid % 8) * q8 + bid / 8 : bid;
  const int ntz = sys.ntiles[2], nty = sys.ntiles[1];
  const int tz = tile % ntz, ty_ = (tile / ntz) % nty, tx = tile / (ntz * nty);
  Tile tl;
  tl.t0[0] = tx * T;
  tl.t0[1] = ty_ * T;
  tl.t0[2] = tz * T;
#pragma unroll
  for (int d = 0; d < 3; ++d) tl.t1[d] = min(tl.t0[d] + T, sys.n[d]);
  for (int i = tid; i < sys.ntypes * 3; i += kThreads) sh_half[i / 3][i % 3] = td[i / 3].half[i % 3];

  const int ci = (wave & 1) * 8 + (lane & 7);
  const int cj = (wave >> 1) * 8 + (lane >> 3);
  const ACC g00 = (ACC)sys.g00, g01x2 = (ACC)(2.0 * sys.g01), g11 = (ACC)sys.g11;
  ACC acc[T];
#pragma unroll
  for (int k = 0; k < T; ++k) acc[k] = 0;
  int count = 0;

  // block-wide exclusive scan of one int per thread; returns the total
  auto block_scan = [&](int v, int* excl) -> int {
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
      const int s = wsum[w];
      off += w < wave ? s : 0;
      tot += s;
    }
    *excl = off + x - v;
    __syncthreads();
    return tot;
  };

  auto process = [&](int nbatch) {
    __syncthreads();
    const int a = tid >> 2, part = tid & 3;
    if (a < nbatch) {
      if constexpr (MODE == MODE_DIAG) stage_diag<ACC>(st, a, part, list[a], sys, td, s_ir, tables, tl);
      else stage_xy<ACC>(st, a, part, list[a], sys, td, s_ir, s_r, tables, tl);
    }
    __syncthreads();
    if constexpr (MODE == MODE_DIAG) accumulate_diag<ACC>(st, nbatch, wave, ci, cj, acc);
    else accumulate_xy<ACC>(st, nbatch, wave, ci, cj, acc, g00, g01x2, g11);
  };

  // total bucket columns over all types (uniform)
  int total_cols = 0;
  for (int t = 0; t < sys.ntypes; ++t) {
    TypeRanges r;
    type_ranges(td[t], sys, tl.t0, tl.t1, r);
    total_cols += r.ncols;
  }

  for (int cbase = 0; cbase < total_cols; cbase += kThreads) {
    // phase A: one bucket column per thread -> (start, length, type)
    int my_s = 0, my_len = 0, my_t = 0;
    int cidx = cbase + tid;
    if (cidx < total_cols) {
      for (int t = 0; t < sys.ntypes; ++t) {
        TypeRanges r;
        type_ranges(td[t], sys, tl.t0, tl.t1, r);
        if (cidx < r.ncols) {
          const int zr = cidx % r.rn[2];
          const int byi = (cidx / r.rn[2]) % r.ny;
          const int bxi = cidx / (r.rn[2] * r.ny);
          const TypeDev& ty = td[t];
          const int bx = range_item(r, 0, bxi), by = range_item(r, 1, byi);
          const int col = ty.bucket_base + (bx * ty.nbk[1] + by) * ty.nbk[2];
          const int zlo = zr == 0 ? r.lo[2][0] : r.lo[2][1];
          const int zhi = zr == 0 ? r.hi[2][0] : r.hi[2][1];
          my_s = boff[col + zlo];
          my_len = boff[col + zhi + 1] - my_s;
          my_t = t;
          break;
        }
        cidx -= r.ncols;
      }
    }
    int excl;
    const int ncand = block_scan(my_len, &excl);
    col_s[tid] = my_s;
    col_t[tid] = my_t;
    col_pre[tid] = excl;
    if (tid == 0) col_pre[kThreads] = ncand;
    __syncthreads();
    // phase B: stream the candidates, exact test, ordered compaction
    for (int g0 = 0; g0 < ncand; g0 += kThreads) {
      const int g = g0 + tid;
      bool hit = false;
      int c = 0;
      if (g < ncand) {
        int lo = 0, hi = kThreads;  // last r with col_pre[r] <= g
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (col_pre[mid] <= g) lo = mid;
          else hi = mid;
        }
        c = col_s[lo] + (g - col_pre[lo]);
        const int t = col_t[lo];
        const int4 q = s_ir[c];
        hit = overlap1(q.x, sh_half[t][0], tl.t0[0], tl.t1[0], sys.n[0]) &&
              overlap1(q.y, sh_half[t][1], tl.t0[1], tl.t1[1], sys.n[1]) &&
              overlap1(q.z, sh_half[t][2], tl.t0[2], tl.t1[2], sys.n[2]);
      }
      int pos;
      const int nhit = block_scan(hit ? 1 : 0, &pos);
      if (hit) list[count + pos] = c;
      count += nhit;
      __syncthreads();
      while (count >= NB) {
        process(NB);
        __syncthreads();
        const int rem = count - NB;
        const int v = tid < rem ? list[NB + tid] : 0;
        __syncthreads();
        if (tid < rem) list[tid] = v;
        count = rem;
        __syncthreads();
      }
    }
    __syncthreads();
  }
  if (count > 0) process(count);

  const int gx = tl.t0[0] + ci, gy = tl.t0[1] + cj;
  if (gx < sys.n[0] && gy < sys.n[1]) {
    ACC* dst = rho + ((size_t)gx * sys.n[1] + gy) * sys.n[2] + tl.t0[2];
    const int nz = tl.t1[2] - tl.t0[2];
    if (nz == T && (sys.n[2] & 3) == 0) {
#pragma unroll
      for (int g = 0; g < T / 4; ++g) {
        if constexpr (std::is_same<ACC, float>::value) {
          reinterpret_cast<float4*>(dst)[g] =
              make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
        } else {
          reinterpret_cast<double2*>(dst)[2 * g] = make_double2(acc[4 * g], acc[4 * g + 1]);
          reinterpret_cast<double2*>(dst)[2 * g + 1] = make_double2(acc[4 * g + 2], acc[4 * g + 3]);
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < T; ++k)
        if (k < nz) dst[k] = acc[k];
    }
  }
}

template <typename ACC>
hipError_t launch_impl(Plan& p, ACC* out, hipStream_t s) {
  const SysDev& sys = p.sys;
  const long long ntiles = (long long)sys.ntiles[0] * sys.ntiles[1] * sys.ntiles[2];
  if (ntiles == 0) return hipSuccess;
  if (sys.mode == MODE_DIAG)
    hipLaunchKernelGGL((k_deposit<MODE_DIAG, ACC>), dim3((unsigned)ntiles), dim3(kThreads), 0, s,
                       sys, p.types, p.bucket_off, p.s_ir, p.s_r, p.tables, out);
  else
    hipLaunchKernelGGL((k_deposit<MODE_XY, ACC>), dim3((unsigned)ntiles), dim3(kThreads), 0, s,
                       sys, p.types, p.bucket_off, p.s_ir, p.s_r, p.tables, out);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_deposit(Plan& p, void* out, hipStream_t s) {
  if (p.fft64) return launch_impl<double>(p, static_cast<double*>(out), s);
  return launch_impl<float>(p, static_cast<float*>(out), s);
}

}  // namespace mdsf
