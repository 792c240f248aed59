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

#include <hip/hip_ext.h>

#include "clip.h"
#include "mdsf_internal.h"
#include "spill.h"
#include "tiles.h"

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
  return window_reaches(ir, A, t0, t1, N);
}

// Per-type bucket ranges for one tile.
struct TypeRanges {
  int lo[3][2], hi[3][2], rn[3];
  int nx, ny, ncols;
  int bucket_base, nbk1, nbk2;  // the type's bucket numbering
};

__device__ __forceinline__ void type_ranges(const TypeDev& ty, const SysDev& sys, const int* t0,
                                            const int* t1, TypeRanges& r) {
  // the type's fields as one batch of loads (the empty asm keeps them ahead
  // of the branches below, which would otherwise wait for them one by one)
  const int count = ty.count, bucket_base = ty.bucket_base;
  int half[3], bsz[3], nbk[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    half[d] = ty.half[d];
    bsz[d] = ty.bsz[d];
    nbk[d] = ty.nbk[d];
  }
  asm volatile("" ::"v"(count), "v"(half[0]), "v"(half[1]), "v"(half[2]), "v"(bsz[0]),
               "v"(bsz[1]), "v"(bsz[2]), "v"(nbk[0]), "v"(nbk[1]), "v"(nbk[2]), "v"(bucket_base));
  r.bucket_base = bucket_base;
  r.nbk1 = nbk[1];
  r.nbk2 = nbk[2];
  if (count == 0 || half[0] == 0 || half[1] == 0 || half[2] == 0) {
    r.ncols = 0;
    r.nx = r.ny = 0;
    r.rn[2] = 0;
    return;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
    r.rn[d] = bucket_ranges(t0[d], t1[d], half[d], sys.n[d], bsz[d], nbk[d], r.lo[d], r.hi[d]);
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
  int idx;  // (tx * ntiles_y + ty) * ntiles_z + tz
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
  const int corner = (q.w >> 16) & 1;
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
  const int corner = (q.w >> 16) & 1;
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
#pragma unroll 2
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
#pragma unroll 1
    for (int sx = 0; sx < 3; ++sx) {
      if (!((xv >> sx) & 1)) continue;
      const ACC bx = st.bx[sx][a][ci];
#pragma unroll 1
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

// ---------------------------------------------------------------------------
// Candidate enumeration shared by both accumulation paths.
// Candidate ring buffer: at most trigger (<= 64) - 1 + one 256-candidate chunk
// entries are live; a power of two makes the wrap a mask.
constexpr int kListCap = 512;
constexpr int kListMask = kListCap - 1;

// MT: the most atom types the block handles (LDS is sized for it)
template <int MT>
struct EnumShared {
  int list[kListCap];  // candidate sorted index | corner << 31
  int wsum[kThreads / 64];
  int col_s[kThreads], col_pre[kThreads + 1], col_t[kThreads];
  int half[MT][3];
  TypeRanges tr[MT];
  int col_base[MT + 1];
};

// block-wide exclusive scan of one int per thread; returns the total
template <class SH>
__device__ __forceinline__ int block_scan(SH& sh, int v, int* excl) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh.wsum[wave] = x;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) {
    const int s = sh.wsum[w];
    off += w < wave ? s : 0;
    tot += s;
  }
  *excl = off + x - v;
  __syncthreads();
  return tot;
}

// The tile's bucket columns: per-type half-widths and bucket ranges, and the
// columns numbered type by type (col_base).  Returns the column count.
template <class SH>
__device__ int enum_setup(const SysDev& sys, const TypeDev* __restrict__ td, const Tile& tl,
                          SH& sh) {
  const int tid = threadIdx.x;
  for (int i = tid; i < sys.ntypes * 3; i += kThreads) sh.half[i / 3][i % 3] = td[i / 3].half[i % 3];
  for (int t = tid; t < sys.ntypes; t += kThreads) type_ranges(td[t], sys, tl.t0, tl.t1, sh.tr[t]);
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int t = 0; t < sys.ntypes; ++t) {
      sh.col_base[t] = acc;
      acc += sh.tr[t].ncols;
    }
    sh.col_base[sys.ntypes] = acc;
  }
  __syncthreads();
  return sh.col_base[sys.ntypes];
}

// Bucket column cidx of the tile: its first sorted index, its length (the
// atoms of the column's z bucket range) and its type.
template <class SH>
__device__ __forceinline__ void column_at(const int* __restrict__ boff, const SH& sh, int cidx,
                                          int& s, int& len, int& type) {
  int t = 0;
  while (sh.col_base[t + 1] <= cidx) ++t;
  const TypeRanges& r = sh.tr[t];
  const int ci = cidx - sh.col_base[t];
  const int zr = r.rn[2] == 2 ? (ci & 1) : 0;
  const int cxy = r.rn[2] == 2 ? (ci >> 1) : ci;
  const int bxi = cxy / r.ny, byi = cxy - bxi * r.ny;
  const int bx = range_item(r, 0, bxi), by = range_item(r, 1, byi);
  const int col = r.bucket_base + (bx * r.nbk1 + by) * r.nbk2;
  const int zlo = zr == 0 ? r.lo[2][0] : r.lo[2][1];
  const int zhi = zr == 0 ? r.hi[2][0] : r.hi[2][1];
  s = boff[col + zlo];
  len = boff[col + zhi + 1] - s;
  type = t;
}

// Bucket column cidx's length from the bucket counts themselves (before the
// bucket scan has turned them into offsets): the sum over its z buckets.
template <class SH>
__device__ __forceinline__ int column_count(const int* __restrict__ bcount, const SH& sh,
                                            int cidx) {
  int t = 0;
  while (sh.col_base[t + 1] <= cidx) ++t;
  const TypeRanges& r = sh.tr[t];
  const int ci = cidx - sh.col_base[t];
  const int zr = r.rn[2] == 2 ? (ci & 1) : 0;
  const int cxy = r.rn[2] == 2 ? (ci >> 1) : ci;
  const int bxi = cxy / r.ny, byi = cxy - bxi * r.ny;
  const int bx = range_item(r, 0, bxi), by = range_item(r, 1, byi);
  const int col = r.bucket_base + (bx * r.nbk1 + by) * r.nbk2;
  const int zlo = zr == 0 ? r.lo[2][0] : r.lo[2][1];
  const int zhi = zr == 0 ? r.hi[2][0] : r.hi[2][1];
  int len = 0;
  for (int z = zlo; z <= zhi; ++z) len += bcount[col + z];
  return len;
}

// Candidate g of the current column round (col_s / col_pre / col_t): its
// sorted index and type.
template <class SH>
__device__ __forceinline__ void candidate_at(const SH& sh, int g, int& c, int& t) {
  // last column r with col_pre[r] <= g (col_pre[0] = 0 <= g): a fixed number
  // of steps, so that several searches of one thread interleave
  static_assert(kThreads == 256, "8 search steps");
  int lo = 0;
#pragma unroll
  for (int step = kThreads / 2; step > 0; step >>= 1)
    lo += sh.col_pre[lo + step] <= g ? step : 0;
  c = sh.col_s[lo] + (g - sh.col_pre[lo]);
  t = sh.col_t[lo];
}

// Streams every atom whose window reaches the tile into the ring sh.list in a
// fixed order; `process(count, head)` consumes entries head, head+1, ...
// (mod kListCap) and returns how many it took.
template <class SH, typename Process>
__device__ void enumerate_tile(const SysDev& sys, const TypeDev* __restrict__ td,
                               const int* __restrict__ boff, const int4* __restrict__ s_ir,
                               const Tile& tl, SH& sh, int trigger, Process&& process) {
  const int tid = threadIdx.x;
  const int total_cols = enum_setup(sys, td, tl, sh);
  int count = 0, head = 0;
  for (int cbase = 0; cbase < total_cols; cbase += kThreads) {
    // phase A: one bucket column per thread -> (start, length, type)
    int my_s = 0, my_len = 0, my_t = 0;
    if (cbase + tid < total_cols) column_at(boff, sh, cbase + tid, my_s, my_len, my_t);
    int excl;
    const int ncand = block_scan(sh, my_len, &excl);
    sh.col_s[tid] = my_s;
    sh.col_t[tid] = my_t;
    sh.col_pre[tid] = excl;
    if (tid == 0) sh.col_pre[kThreads] = ncand;
    __syncthreads();
    // phase B: stream the candidates, exact test, ordered compaction
    for (int g0 = 0; g0 < ncand; g0 += kThreads) {
      const int g = g0 + tid;
      bool hit = false;
      int entry = 0;
      if (g < ncand) {
        int c, t;
        candidate_at(sh, g, c, t);
        const int4 q = s_ir[c];
        hit = overlap1(q.x, sh.half[t][0], tl.t0[0], tl.t1[0], sys.n[0]) &&
              overlap1(q.y, sh.half[t][1], tl.t0[1], tl.t1[1], sys.n[1]) &&
              overlap1(q.z, sh.half[t][2], tl.t0[2], tl.t1[2], sys.n[2]);
        entry = c | (((q.w >> 16) & 1) ? (int)0x80000000u : 0);
      }
      int pos;
      const int nhit = block_scan(sh, hit ? 1 : 0, &pos);
      if (hit) sh.list[(head + count + pos) & kListMask] = entry;
      count += nhit;
      __syncthreads();
      while (count >= trigger) {
        const int n = process(count, head);
        head = (head + n) & kListMask;
        count -= n;
      }
    }
    __syncthreads();
  }
  while (count > 0) {
    const int n = process(count, head);
    head = (head + n) & kListMask;
    count -= n;
  }
}

__device__ __forceinline__ void tile_coords(const SysDev& sys, Tile& tl) {
  // XCD-aware order: consecutive tiles (neighbours in z) share an XCD's L2.
  const int nblk = gridDim.x, bid = blockIdx.x;
  const int q8 = nblk / 8;
  const int tile = bid < 8 * q8 ? (bid % 8) * q8 + bid / 8 : bid;
  const int ntz = sys.ntiles[2], nty = sys.ntiles[1];
  const int tz = tile % ntz, ty_ = (tile / ntz) % nty, tx = tile / (ntz * nty);
  tl.idx = tile;
  tl.t0[0] = tx * T;
  tl.t0[1] = ty_ * T;
  tl.t0[2] = tz * T;
#pragma unroll
  for (int d = 0; d < 3; ++d) tl.t1[d] = min(tl.t0[d] + T, sys.n[d]);
}

// The staged spill tile of tile tl (or null): added by the tiled deposition
// as it stores the tile.
template <typename ACC>
__device__ __forceinline__ const ACC* spill_stage_tile(const SysDev& sys, const ACC* stage,
                                                       const int* flag, const Tile& tl) {
  if (!stage) return nullptr;
  const int jx = spill_index(sys, 0, tl.t0[0] / T), jy = spill_index(sys, 1, tl.t0[1] / T),
            jz = spill_index(sys, 2, tl.t0[2] / T);
  if (jx < 0 || jy < 0 || jz < 0) return nullptr;
  const int j = (jx * sys.sp_cnt[1] + jy) * sys.sp_cnt[2] + jz;
  return flag[j] ? stage + (size_t)j * (T * T * T) : nullptr;
}

// ---------------------------------------------------------------------------
// VALU path (FP64 plans): lane = one (x, y) column, 16 z accumulators.
template <int MODE, typename ACC>
__global__ __launch_bounds__(kThreads) void k_deposit(SysDev sys, const TypeDev* __restrict__ td,
                                                      const int* __restrict__ boff,
                                                      const int4* __restrict__ s_ir,
                                                      const double* __restrict__ s_r,
                                                      const void* __restrict__ tables,
                                                      const ACC* __restrict__ spill,
                                                      const int* __restrict__ spill_flag,
                                                      ACC* __restrict__ rho) {
  using Stage = typename std::conditional<MODE == MODE_DIAG, StageDiag<ACC>, StageXY<ACC>>::type;
  __shared__ Stage st;
  __shared__ EnumShared<kMaxTypes> sh;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  Tile tl;
  tile_coords(sys, tl);
  const int ci = (wave & 1) * 8 + (lane & 7);
  const int cj = (wave >> 1) * 8 + (lane >> 3);
  const ACC g00 = (ACC)sys.g00, g01x2 = (ACC)(2.0 * sys.g01), g11 = (ACC)sys.g11;
  ACC acc[T];
#pragma unroll
  for (int k = 0; k < T; ++k) acc[k] = 0;

  auto process = [&](int count, int head) -> int {
    const int nbatch = min(count, NB);
    __syncthreads();
    const int a = tid >> 2, part = tid & 3;
    if (a < nbatch) {
      const int c = sh.list[(head + a) & kListMask] & 0x7fffffff;
      if constexpr (MODE == MODE_DIAG) stage_diag<ACC>(st, a, part, c, sys, td, s_ir, tables, tl);
      else stage_xy<ACC>(st, a, part, c, sys, td, s_ir, s_r, tables, tl);
    }
    __syncthreads();
    if constexpr (MODE == MODE_DIAG) accumulate_diag<ACC>(st, nbatch, wave, ci, cj, acc);
    else accumulate_xy<ACC>(st, nbatch, wave, ci, cj, acc, g00, g01x2, g11);
    return nbatch;
  };
  enumerate_tile(sys, td, boff, s_ir, tl, sh, NB, process);

  const int gx = tl.t0[0] + ci, gy = tl.t0[1] + cj;
  const ACC* sp = spill_stage_tile(sys, spill, spill_flag, tl);
  if (gx < sys.n[0] && gy < sys.n[1]) {
    ACC* dst = rho + ((size_t)gx * sys.n[1] + gy) * sys.n[2] + tl.t0[2];
    const int nz = tl.t1[2] - tl.t0[2];

    const ACC* sc = sp ? sp + (size_t)(cj * T + ci) * T : nullptr;
#pragma unroll
    for (int k = 0; k < T; ++k)
      if (k < nz) dst[k] = sc ? acc[k] + sc[k] : acc[k];
  }
}

// ---------------------------------------------------------------------------
// General cell (the metric G = U U^T couples z with x or y, dens.py:301 with
// any ucell): nothing separates, so every image is evaluated per voxel in
// FP64, amp exp(-(b^T G b) / 2 s^2), b = r - u dr, and rounded once to the
// plan precision.  Lane = one (x, y) column with 16 z accumulators; the
// enumeration and visit order are the other kernels'.  Per (x, y) image pair
// the z sources of tile cell c are u = c - N, c, c + N, except where the
// corner slot applies (x and y off the core): for y on L the H source is
// u = c + Nb (u >= N), for y on H the L source is u = c - Nb (u < 0).
struct StageGen {
  double r[NB][3];
  double amp[NB], k2[NB];  // Nel / sigma^3, 1 / (2 sigma^2)
  int lo[NB][3], len[NB][3];
};

template <typename ACC>
__global__ __launch_bounds__(kThreads) void k_deposit_gen(SysDev sys,
                                                          const TypeDev* __restrict__ td,
                                                          const int* __restrict__ boff,
                                                          const int4* __restrict__ s_ir,
                                                          const double* __restrict__ s_r,
                                                          const ACC* __restrict__ spill,
                                                          const int* __restrict__ spill_flag,
                                                          ACC* __restrict__ rho) {
  __shared__ StageGen st;
  __shared__ EnumShared<kMaxTypes> sh;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  Tile tl;
  tile_coords(sys, tl);
  const int ci = (wave & 1) * 8 + (lane & 7);
  const int cj = (wave >> 1) * 8 + (lane >> 3);
  const int gx = tl.t0[0] + ci, gy = tl.t0[1] + cj;
  const int nz = tl.t1[2] - tl.t0[2];
  const int N0 = sys.n[0], N1 = sys.n[1], N2 = sys.n[2], nb = sys.nb;
  const double g00 = sys.g00, g11 = sys.g11, g22 = sys.g22;
  const double g01x2 = 2.0 * sys.g01, g02x2 = 2.0 * sys.g02, g12x2 = 2.0 * sys.g12;
  double acc[T];
#pragma unroll
  for (int k = 0; k < T; ++k) acc[k] = 0.0;

  auto process = [&](int count, int head) -> int {
    const int nbatch = min(count, NB);
    __syncthreads();
    if (tid < nbatch) {
      const int c = sh.list[(head + tid) & kListMask] & 0x7fffffff;
      const int4 q = s_ir[c];
      const TypeDev& ty = td[q.w & 0xffff];
      const int slot = c - ty.type_start;
      const bool ok = slot >= 0 && slot < ty.count;  // false only after a reported error
      const int ir[3] = {q.x, q.y, q.z};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        st.r[tid][d] = s_r[3 * (size_t)c + d];
        st.lo[tid][d] = ir[d] - ty.half[d];
        st.len[tid][d] = ok ? 2 * ty.half[d] : 0;
      }
      st.amp[tid] = ty.amp;
      st.k2[tid] = 1.0 / ty.two_sig2;
    }
    __syncthreads();
    if (gx < N0 && gy < N1) {
      for (int a = 0; a < nbatch; ++a) {
        const int lox = st.lo[a][0], loy = st.lo[a][1], loz = st.lo[a][2];
        const int lnx = st.len[a][0], lny = st.len[a][1], lnz = st.len[a][2];
        const double rx = st.r[a][0], ry = st.r[a][1], rz = st.r[a][2];
        const double amp = st.amp[a], k2 = st.k2[a];
#pragma unroll 1
        for (int sx = 0; sx < 3; ++sx) {
          const int ux = gx + (sx - 1) * N0;
          if (ux < lox || ux >= lox + lnx) continue;
          const double bx = rx - (double)ux * sys.dr[0];
#pragma unroll 1
          for (int sy = 0; sy < 3; ++sy) {
            const int uy = gy + (sy - 1) * N1;
            if (uy < loy || uy >= loy + lny) continue;
            const double by = ry - (double)uy * sys.dr[1];
            const double qxy = g00 * bx * bx + g01x2 * bx * by + g11 * by * by;
            const double cz = g02x2 * bx + g12x2 * by;
            const bool slot_rule = sx != SIDE_C && sy != SIDE_C;
            // z offsets of the L and H sources (the C source is u = c)
            const int offL = (slot_rule && sy == SIDE_H) ? -nb : -N2;
            const int offH = (slot_rule && sy == SIDE_L) ? nb : N2;
#pragma unroll
            for (int k = 0; k < T; ++k) {
              if (k >= nz) continue;
              const int c = tl.t0[2] + k;
              const int us[3] = {c + offL, c, c + offH};
#pragma unroll
              for (int j = 0; j < 3; ++j) {
                const int uz = us[j];
                // the moved slots take only off-core sources (u < 0 / u >= N)
                const bool side_ok = j == 1 || (j == 0 ? uz < 0 : uz >= N2);
                if (!side_ok || uz < loz || uz >= loz + lnz) continue;
                const double bz = rz - (double)uz * sys.dr[2];
                const double qv = qxy + g22 * bz * bz + cz * bz;
                acc[k] += amp * exp(-qv * k2);
              }
            }
          }
        }
      }
    }
    return nbatch;
  };
  enumerate_tile(sys, td, boff, s_ir, tl, sh, NB, process);

  const ACC* sp = spill_stage_tile(sys, spill, spill_flag, tl);
  if (gx < N0 && gy < N1) {
    ACC* dst = rho + ((size_t)gx * N1 + gy) * N2 + tl.t0[2];

    const ACC* sc = sp ? sp + (size_t)(cj * T + ci) * T : nullptr;
#pragma unroll
    for (int k = 0; k < T; ++k)
      if (k < nz) dst[k] = sc ? (ACC)acc[k] + sc[k] : (ACC)acc[k];
  }
}

// ---------------------------------------------------------------------------
// MFMA path (FP32 plans).  Each atom becomes one or more separable "virtual
// atoms" (X[16], Y[16], Z[16] over the tile), and the tile is the sum of their
// rank-1 products: D[(x, y), z] += sum_k X_k[x] Y_k[y] Z_k[z].  Wave w owns the
// x rows 4w..4w+3, i.e. four 16(y) x 16(z) blocks, accumulated with
// v_mfma_f32_16x16x4f32 (A[y][k] = X_k[x] Y_k[y], B[k][z] = Z_k[z], 4 virtual
// atoms per instruction).  DIAG corner-quirk atoms expand into 4 separable
// terms: (Xc, Yall, Zs), (Xn, Yc, Zs), (Xn, Yl, ZL), (Xn, Yh, ZH); XY atoms
// stage image offsets and build A per lane (classes all / std / L / H).
constexpr int NV = 64;  // virtual atom slots per batch

typedef float v4f __attribute__((ext_vector_type(4)));

// Row NV is kept zero: the padding entry of the per-wave lists.
struct StageM {
  __align__(16) float vx[NV + 1][T];
  __align__(16) float vy[NV + 1][T];
  __align__(16) float vz[NV + 1][T];
  // per virtual atom and wave w: nonzero X on x rows 4w..4w+3 (read as one int)
  __align__(4) unsigned char vq[NV][4];
  // per wave: the virtual atoms it accumulates, padded to a multiple of 4 with NV
  unsigned char wl[kThreads / 64][NV + 4];
};

// xy-coupled cells: per virtual atom and cell the offsets b of the three
// images per axis and a 3-bit mask of the valid ones (a byte per cell)
template <int NS>
struct StageMXY {
  __align__(16) float bx[3][NS][T];
  __align__(16) float by[3][NS][T];
  __align__(16) float vz[NS][T];
  unsigned char xv[NS][T], yv[NS][T];
  float amp[NS], k2[NS];
  int vreal[NS], vcls[NS], vxm[NS];
};

// The gather is latency-bound (batch staging, candidate streaming).  The
// orthogonal-cell kernel fits 7 blocks per CU (23.4 KB LDS, 72 VGPRs) but is
// held to 6 waves per SIMD: alone it runs as fast as with 7, and in the frame
// pipeline the LDS left free takes a block of the concurrent z pass.
template <int MODE, int MT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(MODE == MODE_DIAG ? 6 : 4, MODE == MODE_DIAG ? 6 : 8)))
void k_deposit_mfma(SysDev sys, const TypeDev* __restrict__ td,
                                                           const int* __restrict__ boff,
                                                           const int4* __restrict__ s_ir,
                                                           const double* __restrict__ s_r,
                                                           const float* __restrict__ slices,
                                                           const float* __restrict__ spill,
                                                           const int* __restrict__ spill_flag,
                                                           float* __restrict__ rho) {
  // virtual atom slots per batch: 64 for orthogonal cells, 48 for xy-coupled
  // cells (whose staging area then fits 4 blocks per CU)
  constexpr int NVK = MODE == MODE_DIAG ? NV : 48;
  using Stage = typename std::conditional<MODE == MODE_DIAG, StageM, StageMXY<NVK>>::type;
  constexpr int EXTRA = MODE == MODE_DIAG ? 3 : 2;  // extra virtual slots of a corner atom
  __shared__ Stage st;
  __shared__ EnumShared<MT> sh;
  constexpr int NXY = MODE == MODE_XY ? NVK : 1;         // arrays only the xy path reads
  constexpr int TXY = MODE == MODE_XY ? MT : 1;
  __shared__ struct {
    double r[NXY][3];    // wrapped position (FP64), xy-coupled cells
    int vbase[NVK], entry[NXY];
    int4 q[NVK];          // home cell, type | corner << 16
    int ok[NXY];         // 0 only after a reported error
    // tile slice offset per axis and corner component 0 of that slice (floats
    // into the slice buffer, which the plan keeps below 2^31 floats), -1: none
    int so[MODE == MODE_DIAG ? NVK : 1][3];
    int co[MODE == MODE_DIAG ? NVK : 1][3];
    int clist[NVK];       // batch indices of the corner atoms
  } bt;
  // per-type constants: g_dd / (2 sigma^2) per axis and Nel / sigma^3
  __shared__ double tgk[TXY][3];
  __shared__ float tamp[TXY];
  // grid size, spacing and tile extent per axis
  __shared__ struct {
    double dr[3];
    int n[3], t0[3], t1[3];
  } axs;
  // per-type slice layout (orthogonal cells), read by the batch set-up
  constexpr int TD = MODE == MODE_DIAG ? MT : 1;
  __shared__ struct {
    long long slice_base[TD], corner_base[TD];
    int slice_stride[TD], corner_stride[TD], type_start[TD], count[TD], nsl[TD][3];
  } tyl;
  if constexpr (MODE == MODE_DIAG) {
    for (int i = threadIdx.x; i < sys.ntypes; i += kThreads) {
      tyl.slice_base[i] = td[i].slice_base;
      tyl.corner_base[i] = td[i].corner_base;
      tyl.slice_stride[i] = td[i].slice_stride;
      tyl.corner_stride[i] = td[i].corner_stride;
      tyl.type_start[i] = td[i].type_start;
      tyl.count[i] = td[i].count;
      tyl.nsl[i][0] = td[i].nsl[0];
      tyl.nsl[i][1] = td[i].nsl[1];
      tyl.nsl[i][2] = td[i].nsl[2];
    }
  }
  if constexpr (MODE == MODE_XY) {
    for (int i = threadIdx.x; i < sys.ntypes; i += kThreads) {
      tgk[i][0] = sys.g00 / td[i].two_sig2;
      tgk[i][1] = sys.g11 / td[i].two_sig2;
      tgk[i][2] = sys.g22 / td[i].two_sig2;
      tamp[i] = td[i].amp_f;
    }
  }
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  Tile tl;
  tile_coords(sys, tl);
  if (tid == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      axs.dr[d] = sys.dr[d];
      axs.n[d] = sys.n[d];
      axs.t0[d] = tl.t0[d];
      axs.t1[d] = tl.t1[d];
    }
  }
  const float g00 = (float)sys.g00, g01x2 = (float)(2.0 * sys.g01), g11 = (float)sys.g11;
  // acc: this batch's MFMA sums; tot: the tile's running sum.  The MFMA
  // accumulation rounds its partial sums with a bias toward zero (measured:
  // 99.5% of a dense tile's voxels come out low, by 9e-7 relative after a few
  // hundred rank-1 terms), so each batch starts from zero and is added into
  // tot with an ordinary round-to-nearest add: the biased roundings then act
  // on batch-sized partial sums only.
  v4f acc[4], tot[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) acc[b] = tot[b] = v4f{0.f, 0.f, 0.f, 0.f};
  if constexpr (MODE == MODE_DIAG) {
    // (ordered before the first batch by the enumeration's barriers)
    if (tid < T) st.vx[NVK][tid] = st.vy[NVK][tid] = st.vz[NVK][tid] = 0.f;
  }

  auto process = [&](int count, int head) -> int {
    // separates the previous batch's MFMA reads of the staging area from
    // this batch's writes
    __syncthreads();
    // how many real atoms fit into NVK virtual slots (same result in every wave)
    const int e = lane < count ? sh.list[(head + lane) & kListMask] : 0;
    const bool isc = lane < count && (e < 0);
    const unsigned long long cm = __ballot(isc);
    const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
    const int vupto = (lane + 1) + EXTRA * __popcll(cm & le);
    const bool fits = lane < count && vupto <= NVK;
    const int nreal = __popcll(__ballot(fits));
    const int ncorner = __popcll(__ballot(fits && isc));
    const int vtot = nreal > 0 ? __shfl(vupto, nreal - 1, 64) : 0;
    const int vpad = (vtot + 3) & ~3;
    if (wave == 0 && lane < nreal) {
      // per-atom lookups once per batch, so that the staging loop below only
      // issues independent table loads
      const int c = e & 0x7fffffff;
      const int4 q = s_ir[c];
      bt.vbase[lane] = lane + EXTRA * __popcll(cm & (le >> 1));
      bt.q[lane] = q;
      if constexpr (MODE == MODE_XY) {
        const TypeDev& ty = td[q.w & 0xffff];
        const int slot = c - ty.type_start;
        bt.entry[lane] = c;
        bt.ok[lane] = slot >= 0 && slot < ty.count;  // false only after a reported error
        bt.r[lane][0] = s_r[3 * (size_t)c + 0];
        bt.r[lane][1] = s_r[3 * (size_t)c + 1];
        bt.r[lane][2] = s_r[3 * (size_t)c + 2];
      }
      if constexpr (MODE == MODE_DIAG) {
        // this tile's slice per axis: slice j = (tile - T0) mod ntiles, T0 the
        // tile of the window start (same numbering as k_slices); the window
        // start lies in [-Nb, N + Nb), Nb <= N, so one wrap suffices
        const int t = q.w & 0xffff;
        const int slot = c - tyl.type_start[t];
        const bool ok = slot >= 0 && slot < tyl.count[t];  // false only after a reported error
        long long off = tyl.slice_base[t] + (long long)slot * tyl.slice_stride[t];
        long long coff = tyl.corner_base[t] + (long long)slot * tyl.corner_stride[t];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int ir = d == 0 ? q.x : (d == 1 ? q.y : q.z);
          const int N = sys.n[d], nt = sys.ntiles[d], ns = tyl.nsl[t][d];
          int lo = ir - sh.half[t][d];
          lo = lo < 0 ? lo + N : (lo >= N ? lo - N : lo);
          int j = tl.t0[d] / T - lo / T;
          if (j < 0) j += nt;
          const bool has = ok && j < ns;
          bt.so[lane][d] = has ? (int)(off + (long long)T * j) : -1;
          bt.co[lane][d] = has ? (int)(coff + (long long)T * j) : -1;
          off += (long long)T * ns;
          coff += (long long)(d == 1 ? 3 : 2) * T * ns;
        }
        if (isc) bt.clist[__popcll(cm & ((1ull << lane) - 1))] = lane;
      }
    }
    __syncthreads();
    if constexpr (MODE == MODE_DIAG) {
      // atoms off the corner regions: four lanes per atom (NVK * 4 = the
      // block), each copying a quarter of the atom's three tile slices; the
      // three loads are issued before the first store, so the batch waits
      // for one memory round trip
      {
        static_assert(NVK * 4 == kThreads, "one quarter-slice lane per atom and quarter");
        const int a = tid >> 2, q4 = tid & 3;
        if (a < nreal && ((bt.q[a].w >> 16) & 1) == 0) {
          // an axis without a slice (-1) loads the buffer's first quarter
          // slice and zeroes it: unconditional loads, no branch between them
          const int o0 = bt.so[a][0], o1 = bt.so[a][1], o2 = bt.so[a][2];
          float4 v0 = *reinterpret_cast<const float4*>(slices + max(o0, 0) + 4 * q4);
          float4 v1 = *reinterpret_cast<const float4*>(slices + max(o1, 0) + 4 * q4);
          float4 v2 = *reinterpret_cast<const float4*>(slices + max(o2, 0) + 4 * q4);
          const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
          if (o0 < 0) v0 = z4;
          if (o1 < 0) v1 = z4;
          if (o2 < 0) v2 = z4;
          const int vb = bt.vbase[a];
          *reinterpret_cast<float4*>(&st.vx[vb][4 * q4]) = v0;
          *reinterpret_cast<float4*>(&st.vy[vb][4 * q4]) = v1;
          *reinterpret_cast<float4*>(&st.vz[vb][4 * q4]) = v2;
          st.vq[vb][q4] = (v0.x != 0.f || v0.y != 0.f || v0.z != 0.f || v0.w != 0.f) ? 1 : 0;
        }
      }
      for (int j = tid; j < (vpad - vtot) * T; j += kThreads) {
        const int v = vtot + j / T, cc = j % T;
        st.vx[v][cc] = 0.f;
        st.vy[v][cc] = 0.f;
        st.vz[v][cc] = 0.f;
        if (cc < 4) st.vq[v][cc] = 0;
      }
      // corner atoms: 4 virtual atoms (Xc, Yall, Zs), (Xn, Yc, Zs),
      // (Xn, Yl, ZL), (Xn, Yh, ZH) from the regular slices and the corner
      // components (x {C, L+H}, y {C, L, H}, z {ZL, ZH})
      for (int it = tid; it < ncorner * 48; it += kThreads) {
        const int ca = it / 48, r48 = it - 48 * ca;
        const int v = r48 / 12, r12 = r48 - 12 * v, d = r12 >> 2, q4 = r12 & 3;
        const int a = bt.clist[ca];
        const int cb = bt.co[a][d];
        const int cs = T * tyl.nsl[bt.q[a].w & 0xffff][d];  // corner component stride
        int off;
        if (d == 0) off = cb >= 0 ? cb + (v == 0 ? 0 : cs) : -1;
        else if (d == 1) off = v == 0 ? bt.so[a][1] : (cb >= 0 ? cb + (v - 1) * cs : -1);
        else off = v < 2 ? bt.so[a][2] : (cb >= 0 ? cb + (v - 2) * cs : -1);
        float4 val = make_float4(0.f, 0.f, 0.f, 0.f);
        if (off >= 0) val = *reinterpret_cast<const float4*>(slices + off + 4 * q4);
        const int vb = bt.vbase[a] + v;
        float* row = d == 0 ? st.vx[vb] : (d == 1 ? st.vy[vb] : st.vz[vb]);
        *reinterpret_cast<float4*>(row + 4 * q4) = val;
        if (d == 0)
          st.vq[vb][q4] = (val.x != 0.f || val.y != 0.f || val.z != 0.f || val.w != 0.f) ? 1 : 0;
      }
    }
    // staging by evaluation (xy-coupled cells): wave d < 3 builds axis d for
    // every (atom, cell) of the batch
    if constexpr (MODE == MODE_XY) if (wave < 3) {
      const int d = wave;
      // this wave's axis (from LDS: a runtime index into the tile's arrays
      // would otherwise go through private memory)
      const int Nd = axs.n[d], t0d = axs.t0[d], tnd = axs.t1[d] - axs.t0[d];
      const double drd = axs.dr[d];
#pragma unroll 1
      for (int base = 0; base < nreal * T; base += 64) {
        const int idx = base + lane;
        const bool act = idx < nreal * T;
        const int a = act ? idx >> 4 : 0, i = idx & 15;
        const int vb = bt.vbase[a];
        const int4 q = bt.q[a];
        const int t = q.w & 0xffff;
        const bool corner = ((q.w >> 16) & 1) != 0;
        const bool valid = act && bt.ok[a];
        const int nv = corner ? EXTRA + 1 : 1;
        bool hC = false, hN = false;
        if (MODE == MODE_DIAG && !corner) {
          // staged from the slices above
        } else if (act && !valid) {
          for (int v = 0; v < nv; ++v) {
            if constexpr (MODE == MODE_DIAG) {
              if (d == 0) st.vx[vb + v][i] = 0.f;
              if (d == 1) st.vy[vb + v][i] = 0.f;
            } else {
              if (d == 0) st.xv[vb + v][i] = 0;
              if (d == 1) st.yv[vb + v][i] = 0;
            }
            if (d == 2) st.vz[vb + v][i] = 0.f;
          }
        } else if (valid) {
          if (MODE == MODE_XY && d < 2) {
            const int N = d == 0 ? sys.n[0] : sys.n[1];
            const int A = sh.half[t][d];
            const int lo = (d == 0 ? q.x : q.y) - A, len = 2 * A;
            const double r = bt.r[a][d];
            const double dr = d == 0 ? sys.dr[0] : sys.dr[1];
            const int t0 = d == 0 ? tl.t0[0] : tl.t0[1];
            const int tn = (d == 0 ? tl.t1[0] : tl.t1[1]) - t0;
            int bits = 0;
#pragma unroll
            for (int sd = 0; sd < 3; ++sd) {
              const int u = t0 + i + (sd - 1) * N;
              const int k = u - lo;
              const bool ok = i < tn && k >= 0 && k < len;
              const float bv = ok ? (float)(r - (double)u * dr) : 0.f;
              if constexpr (MODE == MODE_XY) {
                if (d == 0) st.bx[sd][vb][i] = bv;
                else st.by[sd][vb][i] = bv;
              }
              bits |= ok ? (1 << sd) : 0;
            }
            if constexpr (MODE == MODE_XY) {
              if (d == 0) st.xv[vb][i] = bits;
              else st.yv[vb][i] = bits;
            }
            hC = bits != 0;
          } else {
            // 1-D factor of the image at grid coordinate u (window offset
            // k = u - lo): amp^[d==0] exp(-g_dd b^2 / 2 s^2), b = r - u dr,
            // the exponent in FP64 (dens.py:252-256, 299-308)
            const int A = sh.half[t][d];
            const int lo = (d == 0 ? q.x : (d == 1 ? q.y : q.z)) - A, len = 2 * A;
            const double r = bt.r[a][d], gk = tgk[t][d];
            const float am = d == 0 ? tamp[t] : 1.f;
            const bool in = i < tnd;
            const int uC = t0d + i;
            auto inwin = [&](int u) { return in && u - lo >= 0 && u - lo < len; };
            auto val = [&](int u, bool ok) -> float {
              if (!ok) return 0.f;
              const double b = r - (double)u * drd;
              return am * expf(-(float)(gk * (b * b)));
            };
            const bool bL = inwin(uC - Nd), bC = inwin(uC), bH = inwin(uC + Nd);
            const float vL = val(uC - Nd, bL), vC = val(uC, bC), vH = val(uC + Nd, bH);
            const float all = vC + vL + vH;
            if (d == 2) {
              st.vz[vb][i] = all;
              if (corner) {
                const int u1 = uC + sys.nb, u2 = uC - sys.nb;
                const bool s1 = u1 >= Nd && inwin(u1);
                const bool s2 = u2 < 0 && inwin(u2);
                const float zl = vC + vL + val(u1, s1);
                const float zh = vC + vH + val(u2, s2);
                if constexpr (MODE == MODE_DIAG) {
                  st.vz[vb + 1][i] = all;
                  st.vz[vb + 2][i] = zl;
                  st.vz[vb + 3][i] = zh;
                } else {
                  st.vz[vb + 1][i] = zl;
                  st.vz[vb + 2][i] = zh;
                }
              }
            } else if constexpr (MODE == MODE_DIAG) {
              if (d == 0) {
                st.vx[vb][i] = corner ? vC : all;
                if (corner) {
                  const float xn = vL + vH;
                  st.vx[vb + 1][i] = xn;
                  st.vx[vb + 2][i] = xn;
                  st.vx[vb + 3][i] = xn;
                }
                hC = corner ? bC : (bL || bC || bH);
                hN = bL || bH;
              } else {
                st.vy[vb][i] = all;
                if (corner) {
                  st.vy[vb + 1][i] = vC;
                  st.vy[vb + 2][i] = vL;
                  st.vy[vb + 3][i] = vH;
                }
              }
            }
          }
        }
        // x-cell masks of the virtual atoms: 16 lanes per atom in the ballot
        if (d == 0) {
          const unsigned long long bC = __ballot(hC), bN = __ballot(hN);
          if (act && i == 0) {
            const int sh16 = lane & ~15;
            const int mC = (int)((bC >> sh16) & 0xffff), mN = (int)((bN >> sh16) & 0xffff);
            if (!valid) {
              for (int v = 0; v < nv; ++v) st.vxm[vb + v] = 0;
            } else if (MODE == MODE_DIAG && corner) {
              st.vxm[vb] = mC;
              st.vxm[vb + 1] = st.vxm[vb + 2] = st.vxm[vb + 3] = mN;
            } else {
              for (int v = 0; v < nv; ++v) st.vxm[vb + v] = mC;
            }
          }
        }
      }
    } else if (wave == 3 && MODE == MODE_XY) {
      // wave 3: per-atom scalars (XY) and the zero padding of the last group
      if constexpr (MODE == MODE_XY) {
        for (int a = lane; a < nreal; a += 64) {
          const int vb = bt.vbase[a];
          const int4 q = bt.q[a];
          const TypeDev& ty = td[q.w & 0xffff];
          const bool corner = ((q.w >> 16) & 1) != 0;
          st.amp[vb] = ty.amp_f;
          st.k2[vb] = ty.c2;
          const int nv = corner ? EXTRA + 1 : 1;
          for (int v = 0; v < nv; ++v) {
            st.vreal[vb + v] = vb;
            st.vcls[vb + v] = corner ? v + 1 : 0;
          }
        }
      }
      for (int j = lane; j < (vpad - vtot) * T; j += 64) {
        const int v = vtot + j / T, cc = j % T;
        if constexpr (MODE == MODE_DIAG) {
          st.vx[v][cc] = 0.f;
          st.vy[v][cc] = 0.f;
        } else {
          st.xv[v][cc] = 0;
          st.yv[v][cc] = 0;
          if (cc == 0) {
            st.vreal[v] = v;
            st.vcls[v] = 0;
            st.amp[v] = 0.f;
            st.k2[v] = 0.f;
          }
        }
        st.vz[v][cc] = 0.f;
        if (cc == 0) st.vxm[v] = 0;
      }
    }
    __syncthreads();
    // MFMA accumulation over groups of 4 virtual atoms
    const int kk = lane >> 4, jj = lane & 15;
    int kend = vpad;
    if constexpr (MODE == MODE_DIAG) {
      // this wave's list: the virtual atoms with nonzero X on its 4 x rows
      // (about 2/3 of them), padded with the zero row to whole groups of 4;
      // the list is written and read by this wave only
      const bool on = lane < vtot && st.vq[lane][wave] != 0;
      const unsigned long long bm = __ballot(on);
      const int pos = __popcll(bm & ((1ull << lane) - 1));
      if (on) st.wl[wave][pos] = (unsigned char)lane;
      const int cnt = __popcll(bm);
      kend = (cnt + 3) & ~3;
      if (lane >= cnt && lane < kend) st.wl[wave][lane] = (unsigned char)NVK;
      __builtin_amdgcn_wave_barrier();
      // every x row of the wave's block: testing single rows (LDS read +
      // ballot per group) costs more than the MFMAs it would skip
      for (int k0 = 0; k0 < kend; k0 += 4) {
        const int k = st.wl[wave][k0 + kk];
        const float zv = st.vz[k][jj], yv = st.vy[k][jj];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const float av = st.vx[k][4 * wave + b] * yv;
          acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, zv, acc[b], 0, 0, 0);
        }
      }
    }
    if constexpr (MODE == MODE_XY) for (int k0 = 0; k0 < kend; k0 += 4) {
      const int m = ((st.vxm[k0] | st.vxm[k0 + 1] | st.vxm[k0 + 2] | st.vxm[k0 + 3]) >> (4 * wave)) & 0xf;
      const int mu = __builtin_amdgcn_readfirstlane(m);
      if (!mu) continue;
      const int k = k0 + kk;
      const float zv = st.vz[k][jj];
      {
        const int ra = st.vreal[k], cls = st.vcls[k];
        const float amp = st.amp[ra], k2 = st.k2[ra];
        const int yvb = st.yv[ra][jj];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          if ((mu >> b) & 1) {
            const int xc = 4 * wave + b;
            const int xvb = st.xv[ra][xc];
            float w = 0.f;
            if (cls == 0 && ((xvb & (xvb - 1)) | (yvb & (yvb - 1))) == 0) {
              // at most one image per axis (every cell but wrap-arounds of
              // windows wider than the grid): one exponential, the same
              // value the loop below would sum
              if (xvb != 0 && yvb != 0) {
                const float bx = st.bx[__builtin_ctz(xvb)][ra][xc];
                const float by = st.by[__builtin_ctz(yvb)][ra][jj];
                const float qf = fmaf(g00 * bx, bx, fmaf(g01x2 * bx, by, g11 * by * by));
                w = 0.f + amp * exp2f(-qf * k2);
              }
            } else {
#pragma unroll 1
              for (int sx = 0; sx < 3; ++sx) {
                if (!((xvb >> sx) & 1)) continue;
                const float bx = st.bx[sx][ra][xc];
#pragma unroll 1
                for (int sy = 0; sy < 3; ++sy) {
                  if (!((yvb >> sy) & 1)) continue;
                  const bool take = cls == 0 || (cls == 1 && (sx == SIDE_C || sy == SIDE_C)) ||
                                    (cls == 2 && sx != SIDE_C && sy == SIDE_L) ||
                                    (cls == 3 && sx != SIDE_C && sy == SIDE_H);
                  if (!take) continue;
                  const float by = st.by[sy][ra][jj];
                  const float qf = fmaf(g00 * bx, bx, fmaf(g01x2 * bx, by, g11 * by * by));
                  w += amp * exp2f(-qf * k2);
                }
              }
            }
            acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(w, zv, acc[b], 0, 0, 0);
          }
        }
      }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      tot[b] += acc[b];
      acc[b] = v4f{0.f, 0.f, 0.f, 0.f};
    }
    return nreal;
  };
  enumerate_tile(sys, td, boff, s_ir, tl, sh, NVK, process);

  // D[y][z] of block b: lane l holds y = 4 (l / 16) + r, z = l % 16
  const int gz = tl.t0[2] + (lane & 15);
  const float* sp = spill_stage_tile(sys, spill, spill_flag, tl);

#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int gx = tl.t0[0] + 4 * wave + b;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gy = tl.t0[1] + 4 * (lane >> 4) + r;
      if (gx < tl.t1[0] && gy < tl.t1[1] && gz < tl.t1[2]) {
        float v = tot[b][r];
        if (sp) v += sp[((gy - tl.t0[1]) * T + (gx - tl.t0[0])) * T + (gz - tl.t0[2])];
        rho[((size_t)gx * sys.n[1] + gy) * sys.n[2] + gz] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Tile lists (FP32 orthogonal cells).  The candidate enumeration of the tiled
// deposition runs in the prepare stage as kernels of its own.  Each tile
// gets a region of the slot's pool with one entry per candidate (the atoms of
// the bucket columns that can reach it): k_tile_count counts them,
// k_tile_scan places the regions in tile order (no atomics, which one
// counter shared by every tile would serialise).  k_tile_lists then streams
// the tile's candidates in the enumeration's fixed order, tests them exactly
// and writes the hits into the region, atoms off the corner regions upward
// from its start and corner atoms downward from its end.  The tile's head is
// {region start, regular atoms, corner atoms, region size}.  An entry holds
// what the deposition needs to find the atom's tile slices:
//   x  regular atom: float offset of its slices; corner atom: sorted index
//   y  slice index j_d of the tile on axis d in byte d (kListNoSlice: none),
//      type in bits 24..29
// The counts come from the bucket counts of k_prep, so that the regions can
// be placed while the bucket scan runs.
template <int MT>
__global__ __launch_bounds__(kThreads) void k_tile_count(SysDev sys, const TypeDev* __restrict__ td,
                                                         const int* __restrict__ bcount,
                                                         int4* __restrict__ heads) {
  __shared__ EnumShared<MT> sh;
  const int tid = threadIdx.x;
  Tile tl;
  tile_coords(sys, tl);
  const int total_cols = enum_setup(sys, td, tl, sh);
  int nc = 0;
  for (int cbase = 0; cbase < total_cols; cbase += kThreads) {
    const int len = cbase + tid < total_cols ? column_count(bcount, sh, cbase + tid) : 0;
    int excl;
    nc += block_scan(sh, len, &excl);
  }
  if (tid == 0) heads[tl.idx] = make_int4(0, 0, 0, nc);
}

// Region starts: exclusive scan of the region sizes in tile order (one
// block, 4 tiles per thread and round).
constexpr int kTileScanThreads = 1024;
__global__ __launch_bounds__(kTileScanThreads) void k_tile_scan(int4* __restrict__ heads,
                                                                int ntiles) {
  __shared__ int wsum[kTileScanThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int carry = 0;
  for (int r0 = 0; r0 < ntiles; r0 += 4 * kTileScanThreads) {
    const int i0 = r0 + 4 * tid;
    int v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = i0 + k < ntiles ? heads[i0 + k].w : 0;
    const int sum = v[0] + v[1] + v[2] + v[3];
    int x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kTileScanThreads / 64; ++w) {
      const int s = wsum[w];
      off += w < wave ? s : 0;
      tot += s;
    }
    int run = carry + off + x - sum;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (i0 + k < ntiles) heads[i0 + k].x = run;
      run += v[k];
    }
    carry += tot;
    __syncthreads();  // wsum is rewritten by the next round
  }
}

// Candidates per thread and phase-B round of k_tile_lists (consecutive ones,
// their searches and atom loads in flight together).  Measured at 256^3,
// 50 k atoms, pipelined: 1, 2, 3, 4, 8 gave 4,202, 4,216, 4,173, 4,140,
// 3,922 frames/s (fewer registers, more blocks per CU beside the deposition)
constexpr int kListUnroll = 2;

template <int MT>
__global__ __launch_bounds__(kThreads) void k_tile_lists(SysDev sys, const TypeDev* __restrict__ td,
                                                         const int* __restrict__ boff,
                                                         const int4* __restrict__ s_ir,
                                                         int2* __restrict__ pool, int cap,
                                                         int4* __restrict__ heads,
                                                         int* __restrict__ err,
                                                         int* __restrict__ bcount, int nbk) {
  __shared__ EnumShared<MT> sh;
  // the bucket counts, read by k_tile_count and the bucket scan before this
  // kernel, start the slot's next frame at zero
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < nbk; i += gridDim.x * kThreads)
    bcount[i] = 0;
  __shared__ struct {
    int slice_base[MT], slice_stride[MT], type_start[MT], count[MT], nsl[MT][3];
  } tyl;
  const int tid = threadIdx.x;
  for (int i = tid; i < sys.ntypes; i += kThreads) {
    tyl.slice_base[i] = (int)td[i].slice_base;
    tyl.slice_stride[i] = td[i].slice_stride;
    tyl.type_start[i] = td[i].type_start;
    tyl.count[i] = td[i].count;
    tyl.nsl[i][0] = td[i].nsl[0];
    tyl.nsl[i][1] = td[i].nsl[1];
    tyl.nsl[i][2] = td[i].nsl[2];
  }
  Tile tl;
  tile_coords(sys, tl);
  const int4 hd = heads[tl.idx];
  const int base = hd.x, nc = hd.w;
  const int total_cols = enum_setup(sys, td, tl, sh);  // (its barriers order tyl too)
  if (base > cap - nc) {  // the plan's size bound did not hold (uniform over the block)
    if (tid == 0) {
      atomicOr(err, ERRBIT_LIST);
      heads[tl.idx] = make_int4(0, 0, 0, 0);
    }
    return;
  }
  int nreg = 0, ncor = 0;
  for (int cbase = 0; cbase < total_cols; cbase += kThreads) {
    int my_s = 0, my_len = 0, my_t = 0;
    if (cbase + tid < total_cols) column_at(boff, sh, cbase + tid, my_s, my_len, my_t);
    int excl;
    const int ncand = block_scan(sh, my_len, &excl);
    sh.col_s[tid] = my_s;
    sh.col_t[tid] = my_t;
    sh.col_pre[tid] = excl;
    if (tid == 0) sh.col_pre[kThreads] = ncand;
    __syncthreads();
    for (int g0 = 0; g0 < ncand; g0 += kListUnroll * kThreads) {
      // candidates gb .. gb + kListUnroll - 1 of this thread; those past the
      // end read the last one, so that the loads go out unconditionally
      const int gb = g0 + kListUnroll * tid;
      int cc[kListUnroll], ct[kListUnroll];
      int4 q[kListUnroll];
#pragma unroll
      for (int k = 0; k < kListUnroll; ++k) {
        candidate_at(sh, min(gb + k, ncand - 1), cc[k], ct[k]);
        q[k] = s_ir[cc[k]];
      }
      int2 e[kListUnroll];
      int fl[kListUnroll];  // 1: regular hit, 1 << 16: corner hit
      int h = 0;
#pragma unroll
      for (int k = 0; k < kListUnroll; ++k) {
        fl[k] = 0;
        e[k] = make_int2(0, 0);
        const int t = ct[k];
        if (gb + k < ncand &&
            overlap1(q[k].x, sh.half[t][0], tl.t0[0], tl.t1[0], sys.n[0]) &&
            overlap1(q[k].y, sh.half[t][1], tl.t0[1], tl.t1[1], sys.n[1]) &&
            overlap1(q[k].z, sh.half[t][2], tl.t0[2], tl.t1[2], sys.n[2])) {
          const bool corner = (q[k].w & kAtomCorner) != 0;
          const int slot = cc[k] - tyl.type_start[t];
          const bool ok = slot >= 0 && slot < tyl.count[t];  // false only after a reported error
          // this tile's slice per axis: j = (tile - T0) mod ntiles, T0 the
          // tile of the window start (k_slices' numbering); the start lies
          // in [-Nb, N + Nb), Nb <= N, so one wrap suffices
          int packed = t << 24;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const int ir = d == 0 ? q[k].x : (d == 1 ? q[k].y : q[k].z);
            const int N = sys.n[d];
            int lo = ir - sh.half[t][d];
            lo = lo < 0 ? lo + N : (lo >= N ? lo - N : lo);
            int j = (tl.t0[d] >> 4) - (lo >> 4);
            if (j < 0) j += sys.ntiles[d];
            packed |= ((ok && j < tyl.nsl[t][d]) ? j : kListNoSlice) << (8 * d);
          }
          e[k] = make_int2(corner ? cc[k] : tyl.slice_base[t] + slot * tyl.slice_stride[t],
                           packed);
          fl[k] = corner ? (1 << 16) : 1;
        }
        h += fl[k];
      }
      int pos;
      const int nhit = block_scan(sh, h, &pos);
      int pr = nreg + (pos & 0xffff), pc = ncor + (pos >> 16);
#pragma unroll
      for (int k = 0; k < kListUnroll; ++k) {
        if (fl[k] == 1) pool[base + pr++] = e[k];
        else if (fl[k]) pool[base + nc - 1 - pc++] = e[k];
      }
      nreg += nhit & 0xffff;
      ncor += nhit >> 16;
    }
    __syncthreads();  // the column arrays are rewritten by the next round
  }
  if (tid == 0) heads[tl.idx] = make_int4(base, nreg, ncor, nc);
}

// Tiled deposition from the tile lists (FP32 orthogonal cells): the rank-1
// MFMA sums of k_deposit_mfma without the enumeration.  Regular atoms come in
// batches of 64, four lanes per atom, each loading a quarter of the atom's
// three tile slices; corner atoms in batches of 16, four lanes per virtual
// atom.  The staging rows are double-buffered: batch i + 1 is written to the
// other buffer right after batch i's MFMAs are issued (one barrier per
// batch), and its slices and the entries of batch i + 2 were loaded during
// the batch before.  Sum order: the regular atoms in list order, then the
// corner atoms in list order (fixed, so results are bitwise reproducible).
//
// XY: xy-coupled cells whose windows meet each x and y tile through one
// image.  The y rows hold LY = log2 Y and two more rows per virtual atom hold
// CX (per x cell) and BY (per y cell) from bslices, so that the A operand is
// X[x] 2^(LY[y] + CX[x] BY[y]): the cross factor exp(-2 g01 b_x b_y / 2 s^2)
// costs one fma and one v_exp_f32 per operand element.
// xy-coupled cells stage into one buffer (two barriers per batch): with two
// their 43 KB of LDS held the kernel to 3 blocks per CU, and one buffer at 5
// waves per SIMD ran 13% faster (126 -> 109 us at 256^3, 50 k atoms)
constexpr int kXYStageBuffers = 1;

template <bool XY>
struct StageRows {
  __align__(16) float vx[NV + 1][T];  // row NV stays zero: the lists' padding row
  __align__(16) float vy[NV + 1][T];
  __align__(16) float vz[NV + 1][T];
  __align__(16) float cx[XY ? NV + 1 : 1][T];
  __align__(16) float by[XY ? NV + 1 : 1][T];
  // per row and wave w: nonzero X on x rows 4w..4w+3
  __align__(4) unsigned char vq[NV][4];
};

template <int MT, bool XY>
__device__ __forceinline__ void deposit_lists(SysDev sys, const TypeDev* __restrict__ td,
                                              const int2* __restrict__ pool,
                                              const int4* __restrict__ heads,
                                              const float* __restrict__ slices,
                                              const float* __restrict__ bslices,
                                              const float* __restrict__ spill,
                                              const int* __restrict__ spill_flag,
                                              float* __restrict__ rho) {
  // staging buffers: two (the next batch is staged while this one's MFMAs
  // run) or, for xy-coupled cells, kXYStageBuffers
  constexpr int NB = XY ? kXYStageBuffers : 2;
  __shared__ StageRows<XY> st[NB];
  // per wave: the rows it accumulates, padded to a multiple of 4 with NV
  __shared__ unsigned char wl[kThreads / 64][NV + 4];
  // per type: slice axis offsets (in slices) of the regular and the corner
  // layout, corner component stride per axis (floats), layout bases
  __shared__ struct {
    int pre[MT][3], cpre[MT][3], cs[MT][3];
    int slice_base[MT], slice_stride[MT], corner_base[MT], corner_stride[MT], type_start[MT];
  } ty;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  for (int i = tid; i < sys.ntypes; i += kThreads) {
    const TypeDev& d = td[i];
    ty.pre[i][0] = 0;
    ty.pre[i][1] = d.nsl[0];
    ty.pre[i][2] = d.nsl[0] + d.nsl[1];
    ty.cpre[i][0] = 0;
    ty.cpre[i][1] = 2 * d.nsl[0];
    ty.cpre[i][2] = 2 * d.nsl[0] + 3 * d.nsl[1];
    ty.cs[i][0] = T * d.nsl[0];
    ty.cs[i][1] = T * d.nsl[1];
    ty.cs[i][2] = T * d.nsl[2];
    ty.slice_base[i] = (int)d.slice_base;
    ty.slice_stride[i] = d.slice_stride;
    ty.corner_base[i] = (int)d.corner_base;
    ty.corner_stride[i] = d.corner_stride;
    ty.type_start[i] = d.type_start;
  }
  if (tid < T) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      st[b].vx[NV][tid] = st[b].vy[NV][tid] = st[b].vz[NV][tid] = 0.f;
      if constexpr (XY) st[b].cx[NV][tid] = st[b].by[NV][tid] = 0.f;
    }
  }
  Tile tl;
  tile_coords(sys, tl);
  const int4 hd = heads[tl.idx];
  const int base = hd.x, nreg = hd.y, ncor = hd.z, nc = hd.w;
  __syncthreads();

  // tot: the tile's running sum.  Each batch's MFMA chain starts from zero
  // and is added into tot with round-to-nearest adds: the MFMA accumulation
  // rounds its partial sums toward zero (measured: 99.5% of a dense tile's
  // voxels came out low, by 9e-7 relative after a few hundred rank-1 terms),
  // so the biased roundings act on batch-sized partial sums only.
  v4f tot[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) tot[b] = v4f{0.f, 0.f, 0.f, 0.f};
  const int q4 = tid & 3;
  // a virtual atom's quarter rows: X, Y (XY: LY), Z and (XY) CX, BY
  struct Rows {
    float4 x, y, z, cx, by;
  };
  // quarter slices at float offsets o0..o2 (-1: none, zero; XY: a missing LY
  // is -inf) and (XY) the cross rows at ox, oy (-1: none, zero); a missing one
  // loads the buffer's first quarter slice and replaces it, so the loads go
  // out unconditionally
  auto load = [&](int o0, int o1, int o2, int ox, int oy, Rows& v) {
    v.x = *reinterpret_cast<const float4*>(slices + max(o0, 0) + 4 * q4);
    v.y = *reinterpret_cast<const float4*>(slices + max(o1, 0) + 4 * q4);
    v.z = *reinterpret_cast<const float4*>(slices + max(o2, 0) + 4 * q4);
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (XY) {
      v.cx = *reinterpret_cast<const float4*>(bslices + max(ox, 0) + 4 * q4);
      v.by = *reinterpret_cast<const float4*>(bslices + max(oy, 0) + 4 * q4);
      if (ox < 0) v.cx = z4;
      if (oy < 0) v.by = z4;
    }
    const float ly0 = XY ? -__builtin_inff() : 0.f;
    if (o0 < 0) v.x = z4;
    if (o1 < 0) v.y = make_float4(ly0, ly0, ly0, ly0);
    if (o2 < 0) v.z = z4;
  };
  auto stage = [&](StageRows<XY>& sb, int row, const Rows& v) {
    *reinterpret_cast<float4*>(&sb.vx[row][4 * q4]) = v.x;
    *reinterpret_cast<float4*>(&sb.vy[row][4 * q4]) = v.y;
    *reinterpret_cast<float4*>(&sb.vz[row][4 * q4]) = v.z;
    if constexpr (XY) {
      *reinterpret_cast<float4*>(&sb.cx[row][4 * q4]) = v.cx;
      *reinterpret_cast<float4*>(&sb.by[row][4 * q4]) = v.by;
    }
    sb.vq[row][q4] = (v.x.x != 0.f || v.x.y != 0.f || v.x.z != 0.f || v.x.w != 0.f) ? 1 : 0;
  };
  // A operand of virtual atom k at x row xr, y = jj
  auto aop = [&](const StageRows<XY>& sb, int k, int xr, float yv, float bv) -> float {
    if constexpr (XY)
      return sb.vx[k][xr] * __builtin_amdgcn_exp2f(fmaf(sb.cx[k][xr], bv, yv));
    else
      return sb.vx[k][xr] * yv;
  };
  // MFMA accumulation of staged rows [0, vtot) of sb: the wave's list holds
  // the rows with nonzero X on its 4 x rows, padded with the zero row NV;
  // the first group starts from a zero accumulator
  auto accumulate = [&](const StageRows<XY>& sb, int vtot) {
    const int kk = lane >> 4, jj = lane & 15;
    const bool on = lane < vtot && sb.vq[lane][wave] != 0;
    const unsigned long long bm = __ballot(on);
    const int pos = __popcll(bm & ((1ull << lane) - 1));
    if (on) wl[wave][pos] = (unsigned char)lane;
    const int cnt = __popcll(bm);
    const int kend = (cnt + 3) & ~3;
    if (lane >= cnt && lane < kend) wl[wave][lane] = (unsigned char)NV;
    __builtin_amdgcn_wave_barrier();
    if (kend == 0) return;
    v4f acc[4];
    {
      const int k = wl[wave][kk];
      const float zv = sb.vz[k][jj], yv = sb.vy[k][jj], bv = XY ? sb.by[k][jj] : 0.f;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop(sb, k, 4 * wave + b, yv, bv), zv,
                                                      v4f{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    }
    for (int k0 = 4; k0 < kend; k0 += 4) {
      const int k = wl[wave][k0 + kk];
      const float zv = sb.vz[k][jj], yv = sb.vy[k][jj], bv = XY ? sb.by[k][jj] : 0.f;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop(sb, k, 4 * wave + b, yv, bv), zv, acc[b],
                                                      0, 0, 0);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) tot[b] += acc[b];
  };

  // regular atoms: lanes 4a .. 4a + 3 take atom a of the batch
  {
    const int a = tid >> 2;
    // the atom's regular slice offsets (XY: the cross rows at the x and y
    // slice offsets)
    auto fetch = [&](const int2& e, bool valid, Rows& v) {
      const int t = (int)((unsigned)e.y >> 24);
      const int j0 = e.y & 255, j1 = (e.y >> 8) & 255, j2 = (e.y >> 16) & 255;
      const int o0 = valid && j0 != kListNoSlice ? e.x + T * (ty.pre[t][0] + j0) : -1;
      const int o1 = valid && j1 != kListNoSlice ? e.x + T * (ty.pre[t][1] + j1) : -1;
      const int o2 = valid && j2 != kListNoSlice ? e.x + T * (ty.pre[t][2] + j2) : -1;
      load(o0, o1, o2, o0, o1, v);
    };
    if (nreg > 0) {
      Rows v;
      int2 en = make_int2(0, 0);
      {
        const int2 e = a < nreg ? pool[base + a] : make_int2(0, 0);
        fetch(e, a < nreg, v);
        if (NV + a < nreg) en = pool[base + NV + a];
      }
      stage(st[0], a, v);  // rows past the batch's atoms get zeros
      if (NV < nreg) {  // batch 1 into the registers, the entries of batch 2
        fetch(en, NV + a < nreg, v);
        if (2 * NV + a < nreg) en = pool[base + 2 * NV + a];
      }
      __syncthreads();
      int buf = 0;
      for (int b0 = 0; b0 < nreg; b0 += NV, buf ^= NB - 1) {
        accumulate(st[buf], min(NV, nreg - b0));
        if (b0 + NV < nreg) {
          // batch i + 1 into the other buffer (last read by batch i - 1's
          // MFMAs, which every wave finished before the barrier below it);
          // with one buffer, once every wave's MFMAs of batch i have read it
          if constexpr (NB == 1) __syncthreads();
          stage(st[buf ^ (NB - 1)], a, v);
          if (b0 + 2 * NV < nreg) {
            fetch(en, b0 + 2 * NV + a < nreg, v);
            if (b0 + 3 * NV + a < nreg) en = pool[base + b0 + 3 * NV + a];
          }
        }
        __syncthreads();
      }
    }
  }
  // corner atoms: 4 virtual atoms each, (Xc, Yall, Zs), (Xn, Yc, Zs),
  // (Xn, Yl, ZL), (Xn, Yh, ZH), from the regular slices and the corner
  // components (x {C, L+H}, y {C, L, H}, z {ZL, ZH}); lanes 16 ca .. 16 ca +
  // 15 take corner atom ca, four lanes per virtual atom v (buffer 0, whose
  // last MFMA reads ended before the regular loop's last barrier)
  for (int c0 = 0; c0 < ncor; c0 += NV / 4) {
    const int nca = min(NV / 4, ncor - c0);
    const int ca = tid >> 4, v = (tid >> 2) & 3;
    int o0 = -1, o1 = -1, o2 = -1, ox = -1, oy = -1;
    if (ca < nca) {
      const int2 e = pool[base + nc - 1 - (c0 + ca)];
      const int t = (int)((unsigned)e.y >> 24), slot = e.x - ty.type_start[t];
      const int ab = ty.slice_base[t] + slot * ty.slice_stride[t];
      const int cb = ty.corner_base[t] + slot * ty.corner_stride[t];
      const int j0 = e.y & 255, j1 = (e.y >> 8) & 255, j2 = (e.y >> 16) & 255;
      const int so0 = j0 != kListNoSlice ? ab + T * (ty.pre[t][0] + j0) : -1;
      const int so1 = j1 != kListNoSlice ? ab + T * (ty.pre[t][1] + j1) : -1;
      const int so2 = j2 != kListNoSlice ? ab + T * (ty.pre[t][2] + j2) : -1;
      const int cb0 = j0 != kListNoSlice ? cb + T * (ty.cpre[t][0] + j0) : -1;
      const int cb1 = j1 != kListNoSlice ? cb + T * (ty.cpre[t][1] + j1) : -1;
      const int cb2 = j2 != kListNoSlice ? cb + T * (ty.cpre[t][2] + j2) : -1;
      o0 = cb0 >= 0 ? cb0 + (v == 0 ? 0 : ty.cs[t][0]) : -1;
      o1 = v == 0 ? so1 : (cb1 >= 0 ? cb1 + (v - 1) * ty.cs[t][1] : -1);
      o2 = v < 2 ? so2 : (cb2 >= 0 ? cb2 + (v - 2) * ty.cs[t][2] : -1);
      // XY: the cross rows at the regular x and y slices (one image per
      // tile, so the 4 virtual atoms share them)
      ox = so0;
      oy = so1;
    }
    Rows w;
    load(o0, o1, o2, ox, oy, w);
    __syncthreads();
    stage(st[0], tid >> 2, w);  // row 4 ca + v
    __syncthreads();
    accumulate(st[0], 4 * nca);
  }

  // D[y][z] of block b: lane l holds y = 4 (l / 16) + r, z = l % 16
  const int gz = tl.t0[2] + (lane & 15);
  const float* sp = spill_stage_tile(sys, spill, spill_flag, tl);
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int gx = tl.t0[0] + 4 * wave + b;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gy = tl.t0[1] + 4 * (lane >> 4) + r;
      if (gx < tl.t1[0] && gy < tl.t1[1] && gz < tl.t1[2]) {
        float v = tot[b][r];
        if (sp) v += sp[((gy - tl.t0[1]) * T + (gx - tl.t0[0])) * T + (gz - tl.t0[2])];
        rho[((size_t)gx * sys.n[1] + gy) * sys.n[2] + gz] = v;
      }
    }
  }
}

// Orthogonal cells: 26.6 KB LDS, 6 waves per SIMD.
template <int MT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6, 6)))
void k_deposit_lists(SysDev sys, const TypeDev* __restrict__ td, const int2* __restrict__ pool,
                     const int4* __restrict__ heads, const float* __restrict__ slices,
                     const float* __restrict__ bslices, const float* __restrict__ spill,
                     const int* __restrict__ spill_flag, float* __restrict__ rho) {
  deposit_lists<MT, false>(sys, td, pool, heads, slices, bslices, spill, spill_flag, rho);
}

// xy-coupled cells: five rows per virtual atom in one staging buffer, 22 KB
// LDS, 5 waves per SIMD (92 VGPRs).
template <int MT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5, 8)))
void k_deposit_lists_xy(SysDev sys, const TypeDev* __restrict__ td, const int2* __restrict__ pool,
                        const int4* __restrict__ heads, const float* __restrict__ slices,
                        const float* __restrict__ bslices, const float* __restrict__ spill,
                        const int* __restrict__ spill_flag, float* __restrict__ rho) {
  deposit_lists<MT, true>(sys, td, pool, heads, slices, bslices, spill, spill_flag, rho);
}

// ---------------------------------------------------------------------------
// Adjoint of the MFMA deposition (FP32 orthogonal cells).  With G = dL/drho
// on the folded grid and a virtual atom's tile slices X, Y, Z and their
// derivatives in the atom position dX, dY, dZ, the atom's gradient from one
// tile is
//   ( sum G dX Y Z,  sum G X dY Z,  sum G X Y dZ )   over the tile's cells,
// the exact transpose of the forward rank-1 sum (corner atoms: the sum over
// their 4 virtual atoms).  The tile of G stays in registers as the A operand
// (wave w: x rows 4w..4w+3; lane l: y = l % 16, z = 4 (l / 16) + c for
// c < 4), and per 16 virtual atoms and x row two chains of
// v_mfma_f32_16x16x4f32 contract z with Z and with dZ:
//   U[y][k] = sum_z G[x][y][z] Z_k[z],   V[y][k] = sum_z G[x][y][z] dZ_k[z].
// Each lane then accumulates W1 += dX_k[x] U, W2 += X_k[x] U, W3 += X_k[x] V
// over the wave's rows, contracts y with Y, dY, Y, and the lanes and waves
// are summed in a fixed order.  A real atom's result for the tile goes to
// its own slot (atom, tile slice j0, j1, j2): no atomics, and a second kernel
// adds the slots in a fixed order.
struct StageAdj {
  __align__(16) float vx[NV][T];
  __align__(16) float vy[NV][T];
  __align__(16) float vz[NV][T];
  __align__(16) float dx[NV][T];
  __align__(16) float dy[NV][T];
  __align__(16) float dz[NV][T];
  float red[kThreads / 64][NV][3];  // per-wave sums of each virtual atom
};

template <int MT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_adjoint_mfma(
    SysDev sys, const TypeDev* __restrict__ td, const int* __restrict__ boff,
    const int4* __restrict__ s_ir, const float* __restrict__ slices,
    const float* __restrict__ dslices, const float* __restrict__ grho,
    double* __restrict__ tpart) {
  constexpr int EXTRA = 3;  // extra virtual slots of a corner atom
  __shared__ StageAdj st;
  __shared__ EnumShared<MT> sh;
  __shared__ struct {
    int vbase[NV], nv[NV];
    int4 q[NV];
    // tile slice offset per axis and corner component 0 of that slice
    // (floats, below 2^31), -1: none
    int so[NV][3];
    int co[NV][3];
    int clist[NV];        // batch indices of the corner atoms
    long long pidx[NV];   // result slot, -1: none
  } bt;
  __shared__ struct {
    long long slice_base[MT], corner_base[MT], part_base[MT];
    int slice_stride[MT], corner_stride[MT], type_start[MT], count[MT], nsl[MT][3];
  } tyl;
  for (int i = threadIdx.x; i < sys.ntypes; i += kThreads) {
    tyl.slice_base[i] = td[i].slice_base;
    tyl.corner_base[i] = td[i].corner_base;
    tyl.part_base[i] = td[i].adjt_base;
    tyl.slice_stride[i] = td[i].slice_stride;
    tyl.corner_stride[i] = td[i].corner_stride;
    tyl.type_start[i] = td[i].type_start;
    tyl.count[i] = td[i].count;
    tyl.nsl[i][0] = td[i].nsl[0];
    tyl.nsl[i][1] = td[i].nsl[1];
    tyl.nsl[i][2] = td[i].nsl[2];
  }
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int n16 = lane & 15, q4l = lane >> 4;
  Tile tl;
  tile_coords(sys, tl);
  // this wave's share of the tile, zero outside the grid
  float g[4][4];
  {
    const int gy = tl.t0[1] + n16;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int gx = tl.t0[0] + 4 * wave + b;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int gz = tl.t0[2] + 4 * q4l + c;
        g[b][c] = (gx < tl.t1[0] && gy < tl.t1[1] && gz < tl.t1[2])
                      ? grho[((size_t)gx * sys.n[1] + gy) * sys.n[2] + gz]
                      : 0.f;
      }
    }
  }

  auto process = [&](int count, int head) -> int {
    // separates the previous batch's reads of the staging area and of bt
    // from this batch's writes
    __syncthreads();
    const int e = lane < count ? sh.list[(head + lane) & kListMask] : 0;
    const bool isc = lane < count && (e < 0);
    const unsigned long long cm = __ballot(isc);
    const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
    const int vupto = (lane + 1) + EXTRA * __popcll(cm & le);
    const bool fits = lane < count && vupto <= NV;
    const int nreal = __popcll(__ballot(fits));
    const int ncorner = __popcll(__ballot(fits && isc));
    const int vtot = nreal > 0 ? __shfl(vupto, nreal - 1, 64) : 0;
    const int vpad = (vtot + 15) & ~15;
    if (wave == 0 && lane < nreal) {
      const int c = e & 0x7fffffff;
      const int4 q = s_ir[c];
      bt.vbase[lane] = lane + EXTRA * __popcll(cm & (le >> 1));
      bt.nv[lane] = isc ? EXTRA + 1 : 1;
      bt.q[lane] = q;
      const int t0 = q.w & 0xffff;
      const bool tv = t0 < sys.ntypes;
      const int t = tv ? t0 : 0;
      const int slot = c - tyl.type_start[t];
      const bool ok = tv && slot >= 0 && slot < tyl.count[t];  // false only after a reported error
      long long off = tyl.slice_base[t] + (long long)slot * tyl.slice_stride[t];
      long long coff = tyl.corner_base[t] + (long long)slot * tyl.corner_stride[t];
      long long jj = 0;
      bool all = ok;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        // same slice numbering as the forward: j = (tile - T0) mod ntiles
        const int ir = d == 0 ? q.x : (d == 1 ? q.y : q.z);
        const int N = sys.n[d], nt = sys.ntiles[d], ns = tyl.nsl[t][d];
        int lo = ir - sh.half[t][d];
        lo = lo < 0 ? lo + N : (lo >= N ? lo - N : lo);
        int j = tl.t0[d] / T - lo / T;
        if (j < 0) j += nt;
        const bool has = ok && j < ns;
        bt.so[lane][d] = has ? (int)(off + (long long)T * j) : -1;
        bt.co[lane][d] = has ? (int)(coff + (long long)T * j) : -1;
        off += (long long)T * ns;
        coff += (long long)(d == 1 ? 3 : 2) * T * ns;
        jj = jj * ns + j;
        all = all && has;
      }
      const long long nslp = (long long)tyl.nsl[t][0] * tyl.nsl[t][1] * tyl.nsl[t][2];
      bt.pidx[lane] = all ? tyl.part_base[t] + (long long)slot * nslp + jj : -1LL;
      if (isc) bt.clist[__popcll(cm & ((1ull << lane) - 1))] = lane;
    }
    __syncthreads();
    // atoms off the corner regions: values and derivatives of their three
    // tile slices (float4 per thread)
    for (int it = tid; it < nreal * 12; it += kThreads) {
      const int a = it / 12, r12 = it - 12 * (it / 12), d = r12 >> 2, qq = r12 & 3;
      if (((bt.q[a].w >> 16) & 1) != 0) continue;  // corner atom: below
      const int off = bt.so[a][d];
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f), w = v;
      if (off >= 0) {
        v = *reinterpret_cast<const float4*>(slices + off + 4 * qq);
        w = *reinterpret_cast<const float4*>(dslices + off + 4 * qq);
      }
      const int vb = bt.vbase[a];
      float* row = d == 0 ? st.vx[vb] : (d == 1 ? st.vy[vb] : st.vz[vb]);
      float* drow = d == 0 ? st.dx[vb] : (d == 1 ? st.dy[vb] : st.dz[vb]);
      *reinterpret_cast<float4*>(row + 4 * qq) = v;
      *reinterpret_cast<float4*>(drow + 4 * qq) = w;
    }
    for (int j = tid; j < (vpad - vtot) * T; j += kThreads) {
      const int v = vtot + j / T, cc = j % T;
      st.vx[v][cc] = st.vy[v][cc] = st.vz[v][cc] = 0.f;
      st.dx[v][cc] = st.dy[v][cc] = st.dz[v][cc] = 0.f;
    }
    // corner atoms: (Xc, Yall, Zs), (Xn, Yc, Zs), (Xn, Yl, ZL), (Xn, Yh, ZH)
    for (int it = tid; it < ncorner * 48; it += kThreads) {
      const int ca = it / 48, r48 = it - 48 * ca;
      const int v = r48 / 12, r12 = r48 - 12 * v, d = r12 >> 2, qq = r12 & 3;
      const int a = bt.clist[ca];
      const int cb = bt.co[a][d];
      const int tq = bt.q[a].w & 0xffff;
      const int cs = T * tyl.nsl[tq < sys.ntypes ? tq : 0][d];  // corner component stride
      int off;
      if (d == 0) off = cb >= 0 ? cb + (v == 0 ? 0 : cs) : -1;
      else if (d == 1) off = v == 0 ? bt.so[a][1] : (cb >= 0 ? cb + (v - 1) * cs : -1);
      else off = v < 2 ? bt.so[a][2] : (cb >= 0 ? cb + (v - 2) * cs : -1);
      float4 val = make_float4(0.f, 0.f, 0.f, 0.f), dval = val;
      if (off >= 0) {
        val = *reinterpret_cast<const float4*>(slices + off + 4 * qq);
        dval = *reinterpret_cast<const float4*>(dslices + off + 4 * qq);
      }
      const int vb = bt.vbase[a] + v;
      float* row = d == 0 ? st.vx[vb] : (d == 1 ? st.vy[vb] : st.vz[vb]);
      float* drow = d == 0 ? st.dx[vb] : (d == 1 ? st.dy[vb] : st.dz[vb]);
      *reinterpret_cast<float4*>(row + 4 * qq) = val;
      *reinterpret_cast<float4*>(drow + 4 * qq) = dval;
    }
    __syncthreads();
    for (int k0 = 0; k0 < vpad; k0 += 16) {
      const int k = k0 + n16;
      const float4 zf = *reinterpret_cast<const float4*>(&st.vz[k][4 * q4l]);
      const float4 dzf = *reinterpret_cast<const float4*>(&st.dz[k][4 * q4l]);
      v4f w1 = v4f{0.f, 0.f, 0.f, 0.f}, w2 = w1, w3 = w1;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        v4f u = v4f{0.f, 0.f, 0.f, 0.f}, vv = u;
        u = __builtin_amdgcn_mfma_f32_16x16x4f32(g[b][0], zf.x, u, 0, 0, 0);
        vv = __builtin_amdgcn_mfma_f32_16x16x4f32(g[b][0], dzf.x, vv, 0, 0, 0);
        u = __builtin_amdgcn_mfma_f32_16x16x4f32(g[b][1], zf.y, u, 0, 0, 0);
        vv = __builtin_amdgcn_mfma_f32_16x16x4f32(g[b][1], dzf.y, vv, 0, 0, 0);
        u = __builtin_amdgcn_mfma_f32_16x16x4f32(g[b][2], zf.z, u, 0, 0, 0);
        vv = __builtin_amdgcn_mfma_f32_16x16x4f32(g[b][2], dzf.z, vv, 0, 0, 0);
        u = __builtin_amdgcn_mfma_f32_16x16x4f32(g[b][3], zf.w, u, 0, 0, 0);
        vv = __builtin_amdgcn_mfma_f32_16x16x4f32(g[b][3], dzf.w, vv, 0, 0, 0);
        const float ax = st.vx[k][4 * wave + b], adx = st.dx[k][4 * wave + b];
        w1 += adx * u;
        w2 += ax * u;
        w3 += ax * vv;
      }
      // U, V rows held by this lane: y = 4 (lane / 16) + r
      const float4 yf = *reinterpret_cast<const float4*>(&st.vy[k][4 * q4l]);
      const float4 dyf = *reinterpret_cast<const float4*>(&st.dy[k][4 * q4l]);
      float sx = yf.x * w1[0] + yf.y * w1[1] + yf.z * w1[2] + yf.w * w1[3];
      float sy = dyf.x * w2[0] + dyf.y * w2[1] + dyf.z * w2[2] + dyf.w * w2[3];
      float sz = yf.x * w3[0] + yf.y * w3[1] + yf.z * w3[2] + yf.w * w3[3];
      sx += __shfl_xor(sx, 16, 64);
      sy += __shfl_xor(sy, 16, 64);
      sz += __shfl_xor(sz, 16, 64);
      sx += __shfl_xor(sx, 32, 64);
      sy += __shfl_xor(sy, 32, 64);
      sz += __shfl_xor(sz, 32, 64);
      if (q4l == 0) {
        st.red[wave][k][0] = sx;
        st.red[wave][k][1] = sy;
        st.red[wave][k][2] = sz;
      }
    }
    __syncthreads();
    // per real atom: its virtual atoms over the waves, in a fixed order
    if (tid < nreal) {
      const long long pi = bt.pidx[tid];
      if (pi >= 0) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        const int vb = bt.vbase[tid], nv = bt.nv[tid];
        for (int v = vb; v < vb + nv; ++v)
#pragma unroll
          for (int w = 0; w < kThreads / 64; ++w) {
            ax += (double)st.red[w][v][0];
            ay += (double)st.red[w][v][1];
            az += (double)st.red[w][v][2];
          }
        tpart[3 * pi + 0] = ax;
        tpart[3 * pi + 1] = ay;
        tpart[3 * pi + 2] = az;
      }
    }
    return nreal;
  };
  enumerate_tile(sys, td, boff, s_ir, tl, sh, NV, process);
}

// ---------------------------------------------------------------------------
// Corner-quirk spill (dens.py:107): one workgroup per destination tile of the
// spill grid.  The candidates are the frame's corner atoms in sorted-index
// order (the list k_rank_gather and k_corner_order build); past kCornerCap of them,
// per band, the atoms whose windows reach the source tile (the destination
// tile shifted in z by -Nb for band 0, +Nb for band 1) in the deposition's
// enumeration order.  A corner atom whose z window does not reach the
// destination z tile adds its quirk images (x off-core, y on H / L, z on L /
// H) to the tile.  Each lane owns one (x, y) column and adds to the density
// after the tiled deposition has written it: one writer per cell, a fixed
// order per cell, bitwise reproducible.
constexpr int kSpillTrigger = 128;
constexpr int kSpillChunk = 64;  // listed atoms staged per step

// exp(-q) of an FP64 exponent in the plan precision (FP32 exp for FP32 plans,
// as the tile slices take it)
template <typename ACC>
__device__ __forceinline__ ACC spill_exp(double q);
template <>
__device__ __forceinline__ float spill_exp<float>(double q) { return expf(-(float)q); }
template <>
__device__ __forceinline__ double spill_exp<double>(double q) { return exp(-q); }

// With `stage` the block writes its whole tile (zeros included) to
// stage[block * 16^3 + (ly * 16 + lx) * 16 + lz] (l = cell within the tile)
// for the tiled deposition to add as it stores; otherwise it adds to rho after
// the tiled deposition.
template <typename ACC, bool GEN>
__global__ __launch_bounds__(kThreads) void k_deposit_spill(SysDev sys,
                                                            const TypeDev* __restrict__ td,
                                                            const int* __restrict__ boff,
                                                            const int4* __restrict__ s_ir,
                                                            const double* __restrict__ s_r,
                                                            const int* __restrict__ corners,
                                                            const int* __restrict__ flag,
                                                            ACC* __restrict__ rho,
                                                            ACC* __restrict__ stage) {
  __shared__ EnumShared<kMaxTypes> sh;
  __shared__ int4 sq[kSpillChunk];
  __shared__ double sr[kSpillChunk][3];
  __shared__ ACC sez[kSpillChunk][T];
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  const int ncorner = corners[0], nspill = corners[1 + 2 * kCornerCap];
  const bool listed = ncorner <= kCornerCap;
  // uniform over the block; a staged tile nothing spills into is not read
  if (ncorner == 0 || (listed && nspill == 0) || (stage && !flag[b])) return;
  const int cz = sys.sp_cnt[2], cy = sys.sp_cnt[1];
  const int iz = b % cz, iy = (b / cz) % cy, ix = b / (cz * cy);
  Tile dst;
  dst.t0[0] = spill_tile(sys, 0, ix) * T;
  dst.t0[1] = spill_tile(sys, 1, iy) * T;
  dst.t0[2] = spill_tile(sys, 2, iz) * T;
#pragma unroll
  for (int d = 0; d < 3; ++d) dst.t1[d] = min(dst.t0[d] + T, sys.n[d]);
  const int N0 = sys.n[0], N1 = sys.n[1], N2 = sys.n[2], nb = sys.nb;
  const int xd = dst.t0[0] + (tid & 15), yd = dst.t0[1] + (tid >> 4);
  ACC acc[T];
#pragma unroll
  for (int k = 0; k < T; ++k) acc[k] = (ACC)0;

  for (int band = 0; band < 2; ++band) {
    const int zlo = band == 0 ? 0 : N2 - nb, zhi = band == 0 ? nb : N2;
    if (dst.t1[2] <= zlo || dst.t0[2] >= zhi) continue;  // uniform over the block
    const int shift = band == 0 ? -nb : nb;               // u_z = z + shift
    // does atom (q) spill into this destination tile on this band?
    auto reaches = [&](const int4& q, const TypeDev& ty) __attribute__((always_inline)) -> bool {
      return ((spill_bands(q, ty, sys) >> band) & 1) &&
             spill_tile_unreached(q.z, ty.half[2], dst.t0[2], dst.t1[2], N2) &&
             window_reaches(q.x, ty.half[0], dst.t0[0], dst.t1[0], N0) &&
             window_reaches(q.y, ty.half[1], dst.t0[1], dst.t1[1], N1);
    };
    // z factor exp(-g22 bz^2 / 2 s^2) of tile z cell k (0 off the band or the window)
    auto zfactor = [&](const int4& q, const TypeDev& ty, double rz, int k)
                       __attribute__((always_inline)) -> ACC {
      const int z = dst.t0[2] + k;
      const int uz = z + shift;  // < 0 on band 0, >= N on band 1
      if (z >= dst.t1[2] || z < zlo || z >= zhi || uz < q.z - ty.half[2] || uz >= q.z + ty.half[2])
        return (ACC)0;
      const double bz = rz - (double)uz * sys.dr[2];
      return spill_exp<ACC>((sys.g22 * bz * bz) / ty.two_sig2);
    };
    // this lane's cells of the tile from one reaching atom; ez: its z factors
    // (null: evaluated here)
    auto add = [&](const int4& q, const double* r, const ACC* ez) __attribute__((always_inline)) {
      if (xd >= N0 || yd >= N1) return;
      const TypeDev& ty = td[q.w & 0xffff];
      const int Ax = ty.half[0], Ay = ty.half[1];
      // y image on this band's side: H (u = y + N) for band 0, L (u = y - N) for band 1
      const int uy = band == 0 ? yd + N1 : yd - N1;
      if (uy < q.y - Ay || uy >= q.y + Ay) return;
      const double by = r[1] - (double)uy * sys.dr[1];
#pragma unroll
      for (int sx = 0; sx < 2; ++sx) {
        const int ux = sx == 0 ? xd - N0 : xd + N0;  // the off-core x images
        if (ux < q.x - Ax || ux >= q.x + Ax) continue;
        const double bx = r[0] - (double)ux * sys.dr[0];
        // amp exp(-(b^T G b) / 2 s^2) split into its xy and z factors (dens.py:297-308)
        const ACC exy = (ACC)ty.amp *
                        spill_exp<ACC>((sys.g00 * bx * bx + 2.0 * sys.g01 * bx * by + sys.g11 * by * by) /
                                       ty.two_sig2);
        if constexpr (GEN) {
          // general cells: the z couplings add exp(-2 (g02 bx + g12 by) bz / 2 s^2)
          // (a kernel of its own: the extra exponential per cell would cost the
          // other cells' kernel its registers)
          const double cz = 2.0 * (sys.g02 * bx + sys.g12 * by);
#pragma unroll
          for (int k = 0; k < T; ++k) {
            const ACC f = ez ? ez[k] : zfactor(q, ty, r[2], k);
            acc[k] += exy * f *
                      spill_exp<ACC>(cz * (r[2] - (double)(dst.t0[2] + k + shift) * sys.dr[2]) /
                                     ty.two_sig2);
          }
        } else {
#pragma unroll
          for (int k = 0; k < T; ++k) acc[k] += exy * (ez ? ez[k] : zfactor(q, ty, r[2], k));
        }
      }
    };
    // One step over up to kSpillChunk candidates (lanes below kSpillChunk hold
    // one each, c < 0 for none): the ones reaching this tile are compacted in
    // candidate order into LDS with their z factors, then every lane walks them.
    // Called by the whole block.
    auto step = [&](int c) __attribute__((always_inline)) {
      int4 q = make_int4(0, 0, 0, 0);
      bool hit = false;
      if (c >= 0) {
        q = s_ir[c];
        hit = reaches(q, td[q.w & 0xffff]);
      }
      int pos;
      const int m = block_scan(sh, hit ? 1 : 0, &pos);
      if (m == 0) return;  // uniform
      if (hit) {
        sq[pos] = q;
        sr[pos][0] = s_r[3 * (size_t)c + 0];
        sr[pos][1] = s_r[3 * (size_t)c + 1];
        sr[pos][2] = s_r[3 * (size_t)c + 2];
      }
      __syncthreads();
      for (int i = tid; i < m * T; i += kThreads) {
        const int j = i / T, k = i - (i / T) * T;
        sez[j][k] = zfactor(sq[j], td[sq[j].w & 0xffff], sr[j][2], k);
      }
      __syncthreads();
      for (int j = 0; j < m; ++j) {
        const double r[3] = {sr[j][0], sr[j][1], sr[j][2]};
        add(sq[j], r, sez[j]);
      }
      __syncthreads();
    };
    if (listed) {
      for (int base = 0; base < nspill; base += kSpillChunk) {
        const int a = base + tid;
        step(tid < kSpillChunk && a < nspill ? corners[1 + kCornerCap + a] : -1);
      }
    } else {
      // the atoms whose windows reach the source tile, in enumeration order
      Tile src = dst;
      src.t0[2] += shift;
      src.t1[2] += shift;
      auto process = [&](int count, int head) __attribute__((always_inline)) -> int {
        for (int base = 0; base < count; base += kSpillChunk) {
          int c = -1;
          if (tid < kSpillChunk && base + tid < count) {
            const int e = sh.list[(head + base + tid) & kListMask];
            if (e < 0) c = e & 0x7fffffff;  // corner flag in bit 31
          }
          step(c);
        }
        return count;
      };
      enumerate_tile(sys, td, boff, s_ir, src, sh, kSpillTrigger, process);
      __syncthreads();
    }
  }
  if (stage) {
    ACC* o = stage + (size_t)b * (T * T * T) + (size_t)tid * T;
#pragma unroll
    for (int k = 0; k < T; ++k) o[k] = (ACC)acc[k];
  } else if (xd < N0 && yd < N1) {
    ACC* dcol = rho + ((size_t)xd * N1 + yd) * N2 + dst.t0[2];
    const int nz = dst.t1[2] - dst.t0[2];
#pragma unroll
    for (int k = 0; k < T; ++k)
      if (k < nz && acc[k] != (ACC)0) dcol[k] += acc[k];
  }
}


// ---------------------------------------------------------------------------
// Parked atoms (windows that leave the padded grid, clipped as the
// reference's numpy slices clip them; clip.h): a block per parked atom adds
// its clipped window through the fold's closed form, after the tiled
// deposition, with atomics (only parked atoms take this path, so the order
// of their few additions is the one thing left unfixed).  Blocks of types
// without parked atoms return at once.
template <typename ACC>
__global__ __launch_bounds__(kThreads) void k_parked(SysDev sys, const TypeDev* __restrict__ td,
                                                     const int* __restrict__ boff,
                                                     const int4* __restrict__ s_ir,
                                                     const double* __restrict__ s_r,
                                                     ACC* __restrict__ rho) {
  const TypeDev& ty = td[blockIdx.y];
  const int first = boff[ty.park], last = boff[ty.park + 1];
  for (int s = first + (int)blockIdx.x; s < last; s += (int)gridDim.x) {
    const ClipWin w = clip_window(sys, ty, s_ir[s]);
    if (w.empty) continue;  // uniform over the block
    const double rx = s_r[3 * (size_t)s + 0], ry = s_r[3 * (size_t)s + 1],
                 rz = s_r[3 * (size_t)s + 2];
    const long long total = w.n0 * w.n1 * w.n2;
    for (long long i = threadIdx.x; i < total; i += kThreads) {
      const long long px = w.lo0 + i / (w.n1 * w.n2), py = w.lo1 + (i / w.n2) % w.n1,
                      pz = w.lo2 + i % w.n2;
      double g0, g1, g2;
      const double v = clip_value(sys, ty, rx, ry, rz, px, py, pz, g0, g1, g2);
      if (v != 0.0) atomicAdd(&rho[clip_dest(sys, px, py, pz)], (ACC)v);
    }
  }
}

// Corner-fold spill grid: staged into the slot's tiles (stage) or added to the
// density after the tiled deposition (stage null, out the density).
template <typename ACC>
hipError_t launch_spill(Plan& p, ACC* out, ACC* stage, hipStream_t s) {
  const SysDev& sys = p.sys;
  if (p.spill_tiles == 0 || sys.natoms == 0) return hipSuccess;
  if (sys.mode == MODE_GEN)
    hipLaunchKernelGGL((k_deposit_spill<ACC, true>), dim3((unsigned)p.spill_tiles), dim3(kThreads),
                       0, s, sys, p.types, p.bucket_off, p.s_ir, p.s_r, p.corners, p.spill_flag,
                       out, stage);
  else
    hipLaunchKernelGGL((k_deposit_spill<ACC, false>), dim3((unsigned)p.spill_tiles),
                       dim3(kThreads), 0, s, sys, p.types, p.bucket_off, p.s_ir, p.s_r, p.corners,
                       p.spill_flag, out, stage);
  return hipGetLastError();
}

template <typename ACC>
hipError_t launch_impl(Plan& p, ACC* out, const ACC* stage, hipStream_t s) {
  const SysDev& sys = p.sys;
  const long long ntiles = (long long)sys.ntiles[0] * sys.ntiles[1] * sys.ntiles[2];
  if (ntiles == 0) return hipSuccess;
  const dim3 grid((unsigned)ntiles), block(kThreads);
  // with the plan's deposition events set (stage timing), the events are
  // recorded with the kernel's own dispatch (the kernel's begin and end)
#define DEP_LAUNCH(K, ...)                                                              \
  do {                                                                                  \
    if (p.dep_ev[0])                                                                    \
      hipExtLaunchKernelGGL(K, grid, block, 0, s, p.dep_ev[0], p.dep_ev[1], 0, __VA_ARGS__); \
    else                                                                                \
      hipLaunchKernelGGL(K, grid, block, 0, s, __VA_ARGS__);                            \
  } while (0)
  if (sys.mode == MODE_GEN) {
    DEP_LAUNCH((k_deposit_gen<ACC>), sys, p.types, p.bucket_off,
                       p.s_ir, p.s_r, stage, p.spill_flag, out);
  } else if constexpr (std::is_same<ACC, float>::value) {
    // blocks size their per-type LDS for 16 types when the system has no more
    const bool few = sys.ntypes <= 16;
    if (sys.mode == MODE_DIAG && p.tl_cap > 0 && few)
      DEP_LAUNCH((k_deposit_lists<16>), sys, p.types, p.tl_pool, p.tl_head, p.slices,
                 p.bslices, stage, p.spill_flag, out);
    else if (sys.mode == MODE_DIAG && p.tl_cap > 0)
      DEP_LAUNCH((k_deposit_lists<kMaxTypes>), sys, p.types, p.tl_pool, p.tl_head,
                 p.slices, p.bslices, stage, p.spill_flag, out);
    else if (sys.mode == MODE_XY && p.tl_cap > 0 && few)
      DEP_LAUNCH((k_deposit_lists_xy<16>), sys, p.types, p.tl_pool, p.tl_head, p.slices,
                 p.bslices, stage, p.spill_flag, out);
    else if (sys.mode == MODE_XY && p.tl_cap > 0)
      DEP_LAUNCH((k_deposit_lists_xy<kMaxTypes>), sys, p.types, p.tl_pool, p.tl_head,
                 p.slices, p.bslices, stage, p.spill_flag, out);
    else if (sys.mode == MODE_DIAG && few)
      DEP_LAUNCH((k_deposit_mfma<MODE_DIAG, 16>), sys, p.types,
                         p.bucket_off, p.s_ir, p.s_r, p.slices, stage, p.spill_flag, out);
    else if (sys.mode == MODE_DIAG)
      DEP_LAUNCH((k_deposit_mfma<MODE_DIAG, kMaxTypes>), sys, p.types,
                         p.bucket_off, p.s_ir, p.s_r, p.slices, stage, p.spill_flag, out);
    else if (few)
      DEP_LAUNCH((k_deposit_mfma<MODE_XY, 16>), sys, p.types,
                         p.bucket_off, p.s_ir, p.s_r, p.slices, stage, p.spill_flag, out);
    else
      DEP_LAUNCH((k_deposit_mfma<MODE_XY, kMaxTypes>), sys, p.types,
                         p.bucket_off, p.s_ir, p.s_r, p.slices, stage, p.spill_flag, out);
  } else {
    if (sys.mode == MODE_DIAG)
      DEP_LAUNCH((k_deposit<MODE_DIAG, ACC>), sys, p.types,
                         p.bucket_off, p.s_ir, p.s_r, p.tables, stage, p.spill_flag, out);
    else
      DEP_LAUNCH((k_deposit<MODE_XY, ACC>), sys, p.types,
                         p.bucket_off, p.s_ir, p.s_r, p.tables, stage, p.spill_flag, out);
  }
#undef DEP_LAUNCH
  return hipGetLastError();
}

}  // namespace

hipError_t launch_deposit(Plan& p, void* out, hipStream_t s) {
  // staged spill tiles (launch_spill_stage, run with the frame's binning) are
  // added as the tiles are stored; otherwise the spill pass follows in place
  hipError_t e;
  if (p.fft64) {
    double* o = static_cast<double*>(out);
    const double* st = static_cast<const double*>(p.spill);
    if ((e = launch_impl<double>(p, o, st, s)) != hipSuccess || st) return e;
    return launch_spill<double>(p, o, nullptr, s);
  }
  float* o = static_cast<float*>(out);
  const float* st = static_cast<const float*>(p.spill);
  if ((e = launch_impl<float>(p, o, st, s)) != hipSuccess || st) return e;
  return launch_spill<float>(p, o, nullptr, s);
}

hipError_t launch_parked(Plan& p, void* out, hipStream_t s) {
  const SysDev& sys = p.sys;
  if (sys.natoms == 0 || sys.ntypes == 0) return hipSuccess;
  const dim3 grid(8, (unsigned)sys.ntypes), block(kThreads);
  if (p.fft64)
    hipLaunchKernelGGL(k_parked<double>, grid, block, 0, s, sys, p.types, p.bucket_off, p.s_ir,
                       p.s_r, static_cast<double*>(out));
  else
    hipLaunchKernelGGL(k_parked<float>, grid, block, 0, s, sys, p.types, p.bucket_off, p.s_ir,
                       p.s_r, static_cast<float*>(out));
  return hipGetLastError();
}

hipError_t launch_spill_stage(Plan& p, hipStream_t s) {
  if (!p.spill) return hipSuccess;
  if (p.fft64) return launch_spill<double>(p, nullptr, static_cast<double*>(p.spill), s);
  return launch_spill<float>(p, nullptr, static_cast<float*>(p.spill), s);
}

hipError_t launch_tile_regions(Plan& p, hipStream_t s) {
  const SysDev& sys = p.sys;
  const long long ntiles = (long long)sys.ntiles[0] * sys.ntiles[1] * sys.ntiles[2];
  if (p.tl_cap == 0 || ntiles == 0 || sys.natoms == 0) return hipSuccess;
  const dim3 grid((unsigned)ntiles), block(kThreads);
  if (sys.ntypes <= 16)
    hipLaunchKernelGGL(k_tile_count<16>, grid, block, 0, s, sys, p.types, p.bucket_count,
                       p.tl_head);
  else
    hipLaunchKernelGGL(k_tile_count<kMaxTypes>, grid, block, 0, s, sys, p.types, p.bucket_count,
                       p.tl_head);
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kTileScanThreads), 0, s, p.tl_head, (int)ntiles);
  return hipGetLastError();
}

hipError_t launch_tile_lists(Plan& p, hipStream_t s) {
  const SysDev& sys = p.sys;
  const long long ntiles = (long long)sys.ntiles[0] * sys.ntiles[1] * sys.ntiles[2];
  if (p.tl_cap == 0 || ntiles == 0 || sys.natoms == 0) return hipSuccess;
  const dim3 grid((unsigned)ntiles), block(kThreads);
  if (sys.ntypes <= 16)
    hipLaunchKernelGGL(k_tile_lists<16>, grid, block, 0, s, sys, p.types, p.bucket_off, p.s_ir,
                       p.tl_pool, (int)p.tl_cap, p.tl_head, p.err, p.bucket_count,
                       sys.nbuckets);
  else
    hipLaunchKernelGGL(k_tile_lists<kMaxTypes>, grid, block, 0, s, sys, p.types, p.bucket_off,
                       p.s_ir, p.tl_pool, (int)p.tl_cap, p.tl_head, p.err, p.bucket_count,
                       sys.nbuckets);
  return hipGetLastError();
}

hipError_t launch_adjoint_tiles(Plan& p, const float* grho, hipStream_t s) {
  const SysDev& sys = p.sys;
  const long long ntiles = (long long)sys.ntiles[0] * sys.ntiles[1] * sys.ntiles[2];
  if (ntiles == 0) return hipSuccess;
  const dim3 grid((unsigned)ntiles), block(kThreads);
  if (sys.ntypes <= 16)
    hipLaunchKernelGGL(k_adjoint_mfma<16>, grid, block, 0, s, sys, p.types, p.bucket_off, p.s_ir,
                       p.slices, p.dslices, grho, p.adj_tpart);
  else
    hipLaunchKernelGGL(k_adjoint_mfma<kMaxTypes>, grid, block, 0, s, sys, p.types, p.bucket_off,
                       p.s_ir, p.slices, p.dslices, grho, p.adj_tpart);
  return hipGetLastError();
}

}  // namespace mdsf
