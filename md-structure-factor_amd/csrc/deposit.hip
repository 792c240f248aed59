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
//
// One workgroup (4 waves) owns a 16x16x16 tile of the periodic grid.  Each
// wave owns an 8x8 block of (x, y) columns, each lane one column with 16 z
// accumulators in registers, so the tile is written once with plain stores:
// no global atomics.  Candidate atoms come from a per-type bucket grid
// (sorted by (type, home bucket), atom order inside a bucket), are tested
// exactly and compacted in candidate order, and their 1-D factor slices for
// the tile (all periodic images, split by side) are staged in LDS.  The
// visit order is fixed, so results are bitwise reproducible.
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

__device__ __forceinline__ int side_of(int u, int N) { return u < 0 ? SIDE_L : (u < N ? SIDE_C : SIDE_H); }

struct Stage {
  // DIAG: x0 = ExC (or Ex), x1 = ExN; y0 = EyC (or Ey), y1 = EyL, y2 = EyH.
  // XY:   x0/x1 = bx of image 0/1, y0/y1 = by of image 0/1, codes in xs/ys.
  float x0[NB][T], x1[NB][T];
  float y0[NB][T], y1[NB][T], y2[NB][T];
  float4 z0[NB][T / 4], z1[NB][T / 4], z2[NB][T / 4];
  int xs[NB][T], ys[NB][T];
  int xm[NB], ym[NB], zm[NB], cf[NB];
  float amp[NB], c2[NB];
};

struct Tile {
  int t0[3], t1[3];
};

// Stage one atom's slices for the tile.  part 0: x, 1: y, 2: z, 3: scalars.
template <int MODE>
__device__ void stage_atom(Stage& st, int a, int part, int c, const SysDev& sys,
                           const TypeDev* __restrict__ td, const int4* __restrict__ s_ir,
                           const double* __restrict__ s_r, const float* __restrict__ tables,
                           const Tile& tl) {
  const int4 q = s_ir[c];
  const int t = q.w & 0xffff;
  const int corner = q.w >> 16;
  const TypeDev& ty = td[t];
  const int slot = c - ty.type_start;
  const bool valid = slot >= 0 && slot < ty.count;  // false only after a reported error
  const float* tab = tables + ty.table_base + (long long)(valid ? slot : 0) * ty.stride;
  if (!valid) {
    if (part == 0) st.xm[a] = 0;
    if (part == 1) st.ym[a] = 0;
    if (part == 2) st.zm[a] = 0;
    if (part == 3) st.cf[a] = 0;
    return;
  }
  if (part == 3) {
    st.cf[a] = corner;
    st.amp[a] = ty.amp_f;
    st.c2[a] = ty.c2;
    return;
  }
  const int d = part;
  const int N = sys.n[d];
  const int A = ty.half[d];
  const int ir = d == 0 ? q.x : (d == 1 ? q.y : q.z);
  const int lo = ir - A;
  const int len = 2 * A;
  int mask = 0;
  if (d < 2 && MODE == MODE_XY) {
    const double r = s_r[3 * (size_t)c + d];
    const double dr = sys.dr[d];
    float* b0 = d == 0 ? st.x0[a] : st.y0[a];
    float* b1 = d == 0 ? st.x1[a] : st.y1[a];
    int* code = d == 0 ? st.xs[a] : st.ys[a];
    for (int i = 0; i < T; ++i) {
      const int cc = tl.t0[d] + i;
      int cd = 0, nimg = 0;
      float v0 = 0.f, v1 = 0.f;
      if (cc < tl.t1[d]) {
        for (int j = -1; j <= 1; ++j) {
          const int u = cc + j * N;
          const int k = u - lo;
          if (k >= 0 && k < len) {
            const float bv = (float)(r - (double)u * dr);
            const int sd = side_of(u, N);
            if (nimg == 0) {
              v0 = bv;
              cd |= 1 | (sd << 1);
            } else {
              v1 = bv;
              cd |= 8 | (sd << 4);
            }
            ++nimg;
          }
        }
      }
      b0[i] = v0;
      b1[i] = v1;
      code[i] = cd;
      if (nimg) mask |= 1 << i;
    }
    if (d == 0) st.xm[a] = mask; else st.ym[a] = mask;
    return;
  }
  const float* tb = tab + (d == 0 ? 0 : (d == 1 ? 2 * ty.half[0] : 2 * ty.half[0] + 2 * ty.half[1]));
  if (d == 0) {
    for (int i = 0; i < T; ++i) {
      const int cc = tl.t0[0] + i;
      float vc = 0.f, vn = 0.f;
      bool hit = false;
      if (cc < tl.t1[0]) {
        for (int j = -1; j <= 1; ++j) {
          const int u = cc + j * N, k = u - lo;
          if (k >= 0 && k < len) {
            const float v = tb[k];
            if (u >= 0 && u < N) vc += v; else vn += v;
            hit = true;
          }
        }
      }
      if (corner) {
        st.x0[a][i] = vc;
        st.x1[a][i] = vn;
      } else {
        st.x0[a][i] = vc + vn;
      }
      if (hit) mask |= 1 << i;
    }
    st.xm[a] = mask;
  } else if (d == 1) {
    for (int i = 0; i < T; ++i) {
      const int cc = tl.t0[1] + i;
      float vc = 0.f, vl = 0.f, vh = 0.f;
      bool hit = false;
      if (cc < tl.t1[1]) {
        for (int j = -1; j <= 1; ++j) {
          const int u = cc + j * N, k = u - lo;
          if (k >= 0 && k < len) {
            const float v = tb[k];
            if (u < 0) vl += v; else if (u < N) vc += v; else vh += v;
            hit = true;
          }
        }
      }
      if (corner) {
        st.y0[a][i] = vc;
        st.y1[a][i] = vl;
        st.y2[a][i] = vh;
      } else {
        st.y0[a][i] = vc + vl + vh;
      }
      if (hit) mask |= 1 << i;
    }
    st.ym[a] = mask;
  } else {
    const int nb = sys.nb;
    float zs[T], zl[T], zh[T];
    for (int i = 0; i < T; ++i) {
      const int cc = tl.t0[2] + i;
      float s = 0.f, l = 0.f, h = 0.f;
      bool hit = false;
      if (cc < tl.t1[2]) {
        for (int j = -1; j <= 1; ++j) {
          const int u = cc + j * N, k = u - lo;
          if (k >= 0 && k < len) {
            const float v = tb[k];
            s += v;
            if (u < N) l += v;   // C and L images keep the standard slot (y on L)
            if (u >= 0) h += v;  // C and H images keep the standard slot (y on H)
            hit = true;
          }
        }
        if (corner) {
          const int uL = cc + nb;  // z on H, y on L: lands on u - Nb
          const int kL = uL - lo;
          if (uL >= N && kL >= 0 && kL < len) {
            l += tb[kL];
            hit = true;
          }
          const int uH = cc - nb;  // z on L, y on H: lands on u + Nb
          const int kH = uH - lo;
          if (uH < 0 && kH >= 0 && kH < len) {
            h += tb[kH];
            hit = true;
          }
        }
      }
      zs[i] = s;
      zl[i] = l;
      zh[i] = h;
      if (hit) mask |= 1 << (i / 4);
    }
#pragma unroll
    for (int g = 0; g < T / 4; ++g) {
      st.z0[a][g] = make_float4(zs[4 * g], zs[4 * g + 1], zs[4 * g + 2], zs[4 * g + 3]);
      if (corner) {
        st.z1[a][g] = make_float4(zl[4 * g], zl[4 * g + 1], zl[4 * g + 2], zl[4 * g + 3]);
        st.z2[a][g] = make_float4(zh[4 * g], zh[4 * g + 1], zh[4 * g + 2], zh[4 * g + 3]);
      }
    }
    st.zm[a] = mask;
  }
}

__device__ __forceinline__ void fma4(float* acc, float w, float4 z) {
  acc[0] = fmaf(w, z.x, acc[0]);
  acc[1] = fmaf(w, z.y, acc[1]);
  acc[2] = fmaf(w, z.z, acc[2]);
  acc[3] = fmaf(w, z.w, acc[3]);
}

template <int MODE>
__device__ void accumulate(const Stage& st, int nbatch, int wave, int ci, int cj, float* acc,
                           float g00, float g01x2, float g11) {
  const int wsh_x = (wave & 1) * 8, wsh_y = (wave >> 1) * 8;
  for (int a = 0; a < nbatch; ++a) {
    const int xm = __builtin_amdgcn_readfirstlane(st.xm[a]);
    const int ym = __builtin_amdgcn_readfirstlane(st.ym[a]);
    if (((xm >> wsh_x) & 0xff) == 0 || ((ym >> wsh_y) & 0xff) == 0) continue;
    const int zm = __builtin_amdgcn_readfirstlane(st.zm[a]);
    const int corner = __builtin_amdgcn_readfirstlane(st.cf[a]);
    float ws, wl = 0.f, wh = 0.f;
    if (MODE == MODE_DIAG) {
      if (!corner) {
        ws = st.x0[a][ci] * st.y0[a][cj];
      } else {
        const float xc = st.x0[a][ci], xn = st.x1[a][ci];
        const float yc = st.y0[a][cj], yl = st.y1[a][cj], yh = st.y2[a][cj];
        ws = xc * (yc + yl + yh) + xn * yc;
        wl = xn * yl;
        wh = xn * yh;
      }
    } else {
      ws = 0.f;
      const int xs = st.xs[a][ci], ys = st.ys[a][cj];
      const float amp = st.amp[a], c2 = st.c2[a];
#pragma unroll
      for (int ix = 0; ix < 2; ++ix) {
        const int xcode = (xs >> (3 * ix)) & 7;
        if (!(xcode & 1)) continue;
        const float bx = ix ? st.x1[a][ci] : st.x0[a][ci];
        const int sx = xcode >> 1;
#pragma unroll
        for (int iy = 0; iy < 2; ++iy) {
          const int ycode = (ys >> (3 * iy)) & 7;
          if (!(ycode & 1)) continue;
          const float by = iy ? st.y1[a][cj] : st.y0[a][cj];
          const int sy = ycode >> 1;
          const float qf = fmaf(g00 * bx, bx, fmaf(g01x2 * bx, by, g11 * by * by));
          const float v = amp * exp2f(-qf * c2);
          if (!corner || sx == SIDE_C || sy == SIDE_C) ws += v;
          else if (sy == SIDE_L) wl += v;
          else wh += v;
        }
      }
    }
    if (!corner) {
#pragma unroll
      for (int g = 0; g < T / 4; ++g)
        if ((zm >> g) & 1) fma4(acc + 4 * g, ws, st.z0[a][g]);
    } else {
#pragma unroll
      for (int g = 0; g < T / 4; ++g) {
        if ((zm >> g) & 1) {
          fma4(acc + 4 * g, ws, st.z0[a][g]);
          fma4(acc + 4 * g, wl, st.z1[a][g]);
          fma4(acc + 4 * g, wh, st.z2[a][g]);
        }
      }
    }
  }
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_deposit(SysDev sys, const TypeDev* __restrict__ td,
                                                      const int* __restrict__ boff,
                                                      const int4* __restrict__ s_ir,
                                                      const double* __restrict__ s_r,
                                                      const float* __restrict__ tables,
                                                      float* __restrict__ rho) {
  __shared__ Stage st;
  __shared__ int list[NB + kThreads];
  __shared__ int wcnt[kThreads / 64];

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

  const int ci = (wave & 1) * 8 + (lane & 7);
  const int cj = (wave >> 1) * 8 + (lane >> 3);
  const float g00 = (float)sys.g00, g01x2 = (float)(2.0 * sys.g01), g11 = (float)sys.g11;
  float acc[T];
#pragma unroll
  for (int k = 0; k < T; ++k) acc[k] = 0.f;
  int count = 0;

  auto process = [&](int nbatch) {
    __syncthreads();
    const int a = tid >> 2, part = tid & 3;
    if (a < nbatch) stage_atom<MODE>(st, a, part, list[a], sys, td, s_ir, s_r, tables, tl);
    __syncthreads();
    accumulate<MODE>(st, nbatch, wave, ci, cj, acc, g00, g01x2, g11);
  };

  for (int t = 0; t < sys.ntypes; ++t) {
    const TypeDev& ty = td[t];
    if (ty.count == 0 || ty.half[0] == 0 || ty.half[1] == 0 || ty.half[2] == 0) continue;
    int rlo[3][2], rhi[3][2], rn[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      rn[d] = bucket_ranges(tl.t0[d], tl.t1[d], ty.half[d], sys.n[d], ty.bsz[d], ty.nbk[d],
                            rlo[d], rhi[d]);
    for (int ix = 0; ix < rn[0]; ++ix)
      for (int bx = rlo[0][ix]; bx <= rhi[0][ix]; ++bx)
        for (int iy = 0; iy < rn[1]; ++iy)
          for (int by = rlo[1][iy]; by <= rhi[1][iy]; ++by) {
            const int col = ty.bucket_base + (bx * ty.nbk[1] + by) * ty.nbk[2];
            for (int iz = 0; iz < rn[2]; ++iz) {
              const int s = boff[col + rlo[2][iz]];
              const int e = boff[col + rhi[2][iz] + 1];
              for (int c0 = s; c0 < e; c0 += kThreads) {
                const int c = c0 + tid;
                bool hit = false;
                if (c < e) {
                  const int4 q = s_ir[c];
                  hit = overlap1(q.x, ty.half[0], tl.t0[0], tl.t1[0], sys.n[0]) &&
                        overlap1(q.y, ty.half[1], tl.t0[1], tl.t1[1], sys.n[1]) &&
                        overlap1(q.z, ty.half[2], tl.t0[2], tl.t1[2], sys.n[2]);
                }
                const unsigned long long m = __ballot(hit);
                if (lane == 0) wcnt[wave] = __popcll(m);
                __syncthreads();
                int off = 0, tot = 0;
#pragma unroll
                for (int w = 0; w < kThreads / 64; ++w) {
                  const int v = wcnt[w];
                  off += w < wave ? v : 0;
                  tot += v;
                }
                if (hit) {
                  const int rank = __builtin_amdgcn_mbcnt_hi(
                      (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                  list[count + off + rank] = c;
                }
                count += tot;
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
            }
          }
  }
  if (count > 0) process(count);

  const int gx = tl.t0[0] + ci, gy = tl.t0[1] + cj;
  if (gx < sys.n[0] && gy < sys.n[1]) {
    float* dst = rho + ((size_t)gx * sys.n[1] + gy) * sys.n[2] + tl.t0[2];
    const int nz = tl.t1[2] - tl.t0[2];
    if (nz == T && (sys.n[2] & 3) == 0) {
      float4* d4 = reinterpret_cast<float4*>(dst);
#pragma unroll
      for (int g = 0; g < T / 4; ++g)
        d4[g] = make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
    } else {
#pragma unroll
      for (int k = 0; k < T; ++k)
        if (k < nz) dst[k] = acc[k];
    }
  }
}

}  // namespace

hipError_t launch_deposit(Plan& p, float* out, hipStream_t s) {
  const SysDev& sys = p.sys;
  const long long ntiles = (long long)sys.ntiles[0] * sys.ntiles[1] * sys.ntiles[2];
  if (ntiles == 0) return hipSuccess;
  if (sys.mode == MODE_DIAG)
    hipLaunchKernelGGL(k_deposit<MODE_DIAG>, dim3((unsigned)ntiles), dim3(kThreads), 0, s, sys,
                       p.types, p.bucket_off, p.s_ir, p.s_r, p.tables, out);
  else
    hipLaunchKernelGGL(k_deposit<MODE_XY>, dim3((unsigned)ntiles), dim3(kThreads), 0, s, sys,
                       p.types, p.bucket_off, p.s_ir, p.s_r, p.tables, out);
  return hipGetLastError();
}

}  // namespace mdsf
