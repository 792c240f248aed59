// 3-D real-to-complex FFT for gfx950: Stockham autosort stages, FP64 arithmetic.
//
// Replaces np.fft.rfftn (dens.py:313) and the |F|^2 frame sum
// (dens.py:315-318): unnormalised forward transform, exponent -2 pi i k n / N.
//
// Precision.  Storage is the plan precision R (float or double); every
// butterfly, twiddle product and the real-to-complex split is evaluated in
// FP64, with FP64 twiddles generated on the device.  An FP32 plan rounds to
// FP32 only where values are stored between passes, and the y pass's output,
// whose rounding error the x pass would spread along whole rows of the
// spectrum (the error that dominates faint bins next to bright ones), keeps a
// 4-bit residual per value: the rounding error in units of 1/16 of the FP32
// ulp, so that the x pass reads it back 16 times more precisely.
//
// Layout.  Three passes, each reading whole contiguous rows and writing P
// pencils side by side (runs of P values):
//   z  rho[x][y][0..Nz) (density, R)  ->  spec[x][kz][y] (kz < Nz/2), nyq[x][y] (kz = Nz/2)
//   y  spec[x][kz][0..Ny)             ->  rho[kz][ky][x] (the density buffer, free by then)
//                                         + res[kz][ky][x] (FP32 plans)
//   x  rho[kz][ky][0..Nx)             ->  F in place, or |F|^2 added into sf[kz][ky][kx] (FP64)
// The Nyquist plane nyq (FP64 complex) is transformed in FP64 by a few extra
// blocks of the y and x launches (sf_nyq[x][y]).  The inverse passes of the
// adjoint run the same kernels backwards: x in place (weights applied on
// load), y from rho[kz][ky][x] back to spec[x][kz][y], the complex-to-real z
// pass from spec to density rows.
//
// Register-staged passes (n = R1 R2 or R1 R2 R3, n a power of two up to
// 1024): a block holds 4096 values (16 per thread); the first stage reads its
// inputs straight from global memory, each later stage takes its inputs from
// the previous one through LDS (real and imaginary parts in turn, one FP64
// plane), and the last stage stores straight to global memory or folds |X|^2
// into the accumulator.  Any other length takes the LDS-staged pass (every
// stage in LDS, radices 2/3/4/5/8/16 and generic odd primes).
#include <map>
#include <mutex>

#include <hip/hip_ext.h>

#include "mdsf_internal.h"

namespace mdsf {

namespace {

constexpr int kFftThreads = 256;
constexpr int kRegElems = 4096;  // values per block of the register-staged passes
constexpr int kZRegElems = 2048; // of the register-staged z pass (room beside the deposition)

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };

struct Cx {
  double x, y;
};

__device__ __forceinline__ Cx mk(double a, double b) {
  Cx c;
  c.x = a;
  c.y = b;
  return c;
}
__device__ __forceinline__ Cx cmul(Cx a, Cx b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ Cx cmulc(Cx a, Cx b) { return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
__device__ __forceinline__ Cx cadd(Cx a, Cx b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ Cx csub(Cx a, Cx b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ Cx cscale(double s, Cx a) { return mk(s * a.x, s * a.y); }
__device__ __forceinline__ Cx csel(bool c, Cx a, Cx b) { return mk(c ? a.x : b.x, c ? a.y : b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ Cx mul_mi(Cx a) {
  return INV ? mk(-a.y, a.x) : mk(a.y, -a.x);
}
__device__ __forceinline__ Cx ld_tw(const double2* tw, int k) {
  const double2 v = tw[k];
  return mk(v.x, v.y);
}

// --- small DFTs in registers -------------------------------------------------
template <bool INV>
__device__ __forceinline__ void dft2(Cx* v) {
  const Cx a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void dft3(Cx* v) {
  const double s = 0.86602540378443864676;  // sin(2 pi / 3)
  const Cx t1 = cadd(v[1], v[2]);
  const Cx t2 = mk(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
  const Cx t3 = mul_mi<INV>(cscale(s, csub(v[1], v[2])));
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

template <bool INV>
__device__ __forceinline__ void dft4(Cx* v) {
  const Cx t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  const Cx t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft5(Cx* v) {
  const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
  const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
  const Cx t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const Cx t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const Cx x0 = v[0];
  const Cx a1 = mk(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
  const Cx a2 = mk(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
  const Cx b1 = mul_mi<INV>(mk(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  const Cx b2 = mul_mi<INV>(mk(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  v[0] = mk(x0.x + t1.x + t2.x, x0.y + t1.y + t2.y);
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

template <bool INV>
__device__ __forceinline__ void dft8(Cx* v) {
  const double h = 0.70710678118654752440;
  Cx e[4] = {v[0], v[2], v[4], v[6]};
  Cx o[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(e);
  dft4<INV>(o);
  const Cx w1 = mk(h, INV ? h : -h);
  const Cx w3 = mk(-h, INV ? h : -h);
  const Cx o1 = cmul(o[1], w1);
  const Cx o2 = mul_mi<INV>(o[2]);
  const Cx o3 = cmul(o[3], w3);
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <bool INV>
__device__ __forceinline__ void dft16(Cx* v) {
  Cx e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    e[i] = v[2 * i];
    o[i] = v[2 * i + 1];
  }
  dft8<INV>(e);
  dft8<INV>(o);
  const double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173;
  const double h = 0.70710678118654752440;
  const double sg = INV ? 1.0 : -1.0;
  const Cx w[8] = {mk(1, 0),        mk(c1, sg * s1),  mk(h, sg * h),  mk(s1, sg * c1),
                   mk(0, sg),       mk(-s1, sg * c1), mk(-h, sg * h), mk(-c1, sg * s1)};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const Cx t = k == 0 ? o[0] : cmul(o[k], w[k]);
    v[k] = cadd(e[k], t);
    v[k + 8] = csub(e[k], t);
  }
}

template <int RAD, bool INV>
__device__ __forceinline__ void dft(Cx* v) {
  if constexpr (RAD == 2) dft2<INV>(v);
  else if constexpr (RAD == 3) dft3<INV>(v);
  else if constexpr (RAD == 4) dft4<INV>(v);
  else if constexpr (RAD == 5) dft5<INV>(v);
  else if constexpr (RAD == 8) dft8<INV>(v);
  else dft16<INV>(v);
}

// inputs r >= 1 of a butterfly at k times W^(k r step)
template <int RAD, bool INV>
__device__ __forceinline__ void twiddle_in(Cx* v, int k, int step, const double2* tw) {
  // k = 0 multiplies by the table's exact 1 (no divergent branch)
#pragma unroll
  for (int r = 1; r < RAD; ++r) {
    const Cx w = ld_tw(tw, k * r * step);
    v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
  }
}

// --- storage conversions ---------------------------------------------------------
template <typename R>
__device__ __forceinline__ Cx load_cx(const typename V2<R>::T* p, long long i) {
  const auto v = p[i];
  return mk((double)v.x, (double)v.y);
}
template <typename R>
__device__ __forceinline__ void store_cx(typename V2<R>::T* p, long long i, Cx v) {
  typename V2<R>::T o;
  o.x = (R)v.x;
  o.y = (R)v.y;
  p[i] = o;
}

// 4-bit residual of an FP32 rounding: round((v - (float)v) * 16 / ulp) in
// [-8, 7], ulp = 2^(max(e, 1) - 150) the FP32 spacing at (float)v (biased
// exponent e).  Both scalings are exact powers of two (ldexp), so encoding
// and decoding take a few instructions and no division.
__device__ __forceinline__ int res_exp(float h) {
  const int e = (int)((__float_as_uint(h) >> 23) & 0xffu);
  return e > 1 ? e : 1;
}
__device__ __forceinline__ unsigned res_nibble(double v, float h) {
  const double q = rint(ldexp(v - (double)h, 154 - res_exp(h)));
  const double c = fmin(fmax(q, -8.0), 7.0);
  return (unsigned)(int)c & 15u;
}
__device__ __forceinline__ double res_value(unsigned nib, float h) {
  const int qi = (int)((nib ^ 8u) & 15u) - 8;
  return ldexp((double)qi, res_exp(h) - 154);
}

// --- LDS-staged stages (any radix), pencil-fastest, FP64 planes --------------------
template <int RAD, bool INV>
__device__ void stage_fixed(const double* __restrict__ ire, const double* __restrict__ iim,
                            double* __restrict__ ore, double* __restrict__ oim, int n, int P,
                            int pitch, int ns, const double2* __restrict__ tw) {
  const int nq = n / RAD;
  const int nbf = P * nq;
  const int twstep = n / (ns * RAD);
  for (int b = threadIdx.x; b < nbf; b += kFftThreads) {
    const int p = b % P, j = b / P;
    const int k = j % ns;
    const int row = p * pitch;
    Cx v[RAD];
#pragma unroll
    for (int r = 0; r < RAD; ++r) {
      const int a = row + j + r * nq;
      v[r] = mk(ire[a], iim[a]);
    }
    twiddle_in<RAD, INV>(v, k, twstep, tw);
    dft<RAD, INV>(v);
    const int d0 = row + (j / ns) * ns * RAD + k;
#pragma unroll
    for (int r = 0; r < RAD; ++r) {
      ore[d0 + r * ns] = v[r].x;
      oim[d0 + r * ns] = v[r].y;
    }
  }
  __syncthreads();
}

// Any other radix (odd primes): one output element per thread iteration,
// direct DFT with W_RAD^m = W_n^(m n / RAD) from the twiddle table.
template <bool INV>
__device__ void stage_generic(const double* __restrict__ ire, const double* __restrict__ iim,
                              double* __restrict__ ore, double* __restrict__ oim, int n, int P,
                              int pitch, int ns, int RAD, const double2* __restrict__ tw) {
  const int nq = n / RAD;
  const int twstep = n / (ns * RAD);
  const int step = n / RAD;
  for (int o = threadIdx.x; o < P * n; o += kFftThreads) {
    const int p = o % P, rest = o / P;
    const int j = rest % nq, m = rest / nq;
    const int k = j % ns;
    const int row = p * pitch;
    Cx acc = mk(0, 0);
    int wm = 0;  // (r * m) mod RAD
    for (int r = 0; r < RAD; ++r) {
      const int a = row + j + r * nq;
      Cx x = mk(ire[a], iim[a]);
      if (r > 0 && k > 0) {
        const Cx w = ld_tw(tw, k * r * twstep);
        x = INV ? cmulc(x, w) : cmul(x, w);
      }
      const Cx w = ld_tw(tw, wm * step);
      acc = cadd(acc, INV ? cmulc(x, w) : cmul(x, w));
      wm += m;
      if (wm >= RAD) wm -= RAD;
    }
    const int d = row + (j / ns) * ns * RAD + k + m * ns;
    ore[d] = acc.x;
    oim[d] = acc.y;
  }
  __syncthreads();
}

// Every stage of `ax` on P pencils in LDS, ping-ponging between plane pairs
// 0 and 1 of `lds` ([re0 | im0 | re1 | im1], P * pitch values each).
// Returns the plane pair holding the result.  A radix-16 stage runs as two
// radix-4 stages (the same factorisation of n): a 16-point butterfly held
// with its twiddles would set the register budget of every kernel that
// contains this loop.
template <bool INV>
__device__ int run_stages(double* lds, int n, int P, int pitch, const FftAxis& ax) {
  const int plane = P * pitch;
  int ns = 1, cur = 0;
  for (int s = 0; s < ax.nstages; ++s) {
    const int RAD0 = ax.radix[s];
    const int sub = RAD0 == 16 ? 2 : 1;
    const int RAD = RAD0 == 16 ? 4 : RAD0;
    for (int h = 0; h < sub; ++h) {
      const double* ire = lds + (2 * cur) * plane;
      const double* iim = ire + plane;
      double* ore = lds + (2 * (cur ^ 1)) * plane;
      double* oim = ore + plane;
      switch (RAD) {
        case 2: stage_fixed<2, INV>(ire, iim, ore, oim, n, P, pitch, ns, ax.tw64); break;
        case 3: stage_fixed<3, INV>(ire, iim, ore, oim, n, P, pitch, ns, ax.tw64); break;
        case 4: stage_fixed<4, INV>(ire, iim, ore, oim, n, P, pitch, ns, ax.tw64); break;
        case 5: stage_fixed<5, INV>(ire, iim, ore, oim, n, P, pitch, ns, ax.tw64); break;
        case 8: stage_fixed<8, INV>(ire, iim, ore, oim, n, P, pitch, ns, ax.tw64); break;
        default: stage_generic<INV>(ire, iim, ore, oim, n, P, pitch, ns, RAD, ax.tw64); break;
      }
      ns *= RAD;
      cur ^= 1;
    }
  }
  return cur;
}

__host__ __device__ inline int lds_pitch(int n) { return (n % 2 == 0) ? n + 1 : n; }

// --- twiddles: W_n^k = exp(-2 pi i k / n), generated in FP64 ----------------------
__global__ void k_twiddles(float2* __restrict__ tw, double2* __restrict__ tw64, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double s, c;
  sincospi(-2.0 * (double)k / (double)n, &s, &c);
  if (tw) tw[k] = make_float2((float)c, (float)s);
  tw64[k] = make_double2(c, s);
}

// --- pencil passes -----------------------------------------------------------------
// Pencil p of a block at slow index s and first fast index f0 reads its element
// e at in + s*in_s + (f0+p)*in_f + e*in_e and writes its element k to
// out + s*out_s + (f0+p)*out_f + k*out_e.  Blocks run over (s, f-tile).
struct Geom {
  long long in_s, in_f, in_e, out_s, out_f, out_e;
  int nf;            // fast indices per slow index
  int tiles_f;       // ceil(nf / P)
  long long blocks;  // slow count * tiles_f
};

enum Epilogue { EPI_STORE = 0, EPI_ACC = 1 };
// residual nibbles read with the input (RES_IN) and / or written with the output (RES_OUT)
enum Residual { RES_NONE = 0, RES_OUT = 1, RES_IN = 2 };

template <typename R>
struct PassArgs {
  const typename V2<R>::T* in;
  typename V2<R>::T* out;
  double* acc;                  // EPI_ACC: |X|^2 added at the output index
  const double* w;              // X *= w[input index] before the transform (inverse x pass)
  const unsigned char* res_in;  // RES_IN: residual nibbles at the input index
  unsigned char* res_out;       // RES_OUT: residual nibbles at the output index
  Geom g;
  int P;
  int n;
  FftAxis ax;
  // Nyquist plane (FP64, in place): blocks [g.blocks, g.blocks + nq_pencils)
  // take one pencil each: element e at nyq + pencil*nq_p + e*nq_e
  double2* nyq;
  double* nyq_acc;
  const double* nyq_w;
  long long nq_p, nq_e;
  int nq_pencils;
  int tpb;                      // tiles per block of the register-staged pass
  // EPI_ACC, register-staged pass: a second spectrum (the next frame's, with
  // its residuals and Nyquist plane) transformed by the same blocks, its
  // |X|^2 added after the first one's, so that the FP64 sums are read and
  // written once for two frames (null: one spectrum)
  const typename V2<R>::T* in2;
  const unsigned char* res_in2;
  double2* nyq2;
};

// One Nyquist-plane pencil per block, every stage in LDS (FP64).
template <bool INV, int EPI, typename A>
__device__ void nyq_pencil(const A& a, int pencil, double* lds) {
  const int n = a.n, pitch = lds_pitch(n);
  double* re = lds;
  double* im = lds + pitch;
  const long long base = (long long)pencil * a.nq_p;
  __syncthreads();  // the LDS may still be read for the previous pencil or tile
  for (int e = threadIdx.x; e < n; e += kFftThreads) {
    const long long g = base + e * a.nq_e;
    const double2 v = a.nyq[g];
    const double sc = a.nyq_w ? a.nyq_w[g] : 1.0;
    re[e] = v.x * sc;
    im[e] = v.y * sc;
  }
  __syncthreads();
  const int cur = run_stages<INV>(lds, n, 1, pitch, a.ax);
  re = lds + 2 * cur * pitch;
  im = re + pitch;
  for (int e = threadIdx.x; e < n; e += kFftThreads) {
    const long long g = base + e * a.nq_e;
    if (EPI == EPI_STORE) a.nyq[g] = make_double2(re[e], im[e]);
    else a.nyq_acc[g] += re[e] * re[e] + im[e] * im[e];
  }
}

// Global element access of a pass: load with the optional weight and
// residual, store / accumulate with the optional residual.
template <typename R, int RES>
__device__ __forceinline__ Cx pass_load(const PassArgs<R>& a, long long gi) {
  Cx x = load_cx<R>(a.in, gi);
  if constexpr ((RES & RES_IN) != 0) {
    const unsigned b = a.res_in[gi];
    x.x += res_value(b & 15u, (float)x.x);
    x.y += res_value(b >> 4, (float)x.y);
  }
  if (a.w) x = cscale(a.w[gi], x);
  return x;
}
template <typename R, int EPI, int RES>
__device__ __forceinline__ void pass_store(const PassArgs<R>& a, long long go, Cx v) {
  if constexpr (EPI == EPI_ACC) {
    a.acc[go] += v.x * v.x + v.y * v.y;
  } else {
    store_cx<R>(a.out, go, v);
    if constexpr ((RES & RES_OUT) != 0)
      a.res_out[go] = (unsigned char)(res_nibble(v.x, (float)v.x) | (res_nibble(v.y, (float)v.y) << 4));
  }
}

// The register-staged pass for n = R1 R2 R3 (R3 = 1: two stages).  A tile is
// P = 4096 / n pencils; each thread holds 16 values per stage (16 / R
// butterflies of radix R).  IN_E / OUT_E: the input / output elements of a
// pencil are adjacent in memory (lanes element-fastest, and in_e / out_e = 1),
// else the pencils are (lanes pencil-fastest).
template <int n>
constexpr int reg_pencils() { return kRegElems / n; }

template <int n>
__host__ __device__ constexpr int reg_pitch() { return n + n / 16 + 1; }
__device__ __forceinline__ int reg_pos(int i) { return i + (i >> 4); }

// Workgroup barrier that orders LDS accesses only: global loads issued
// before it (the next tile's inputs) stay in flight across it.  The empty
// asm statements keep the compiler from moving memory accesses (twiddle
// loads in particular) across it, which would hold them in registers
// through the exchange.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  asm volatile("" ::: "memory");
}

// x, hidden from the optimiser: per-lane LDS positions derived from it are
// recomputed in each tile iteration instead of being hoisted out of the tile
// loop and held in registers for the whole kernel.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// Element `off` (a 32-bit offset) of an array at a wave-uniform base.
template <typename T>
__device__ __forceinline__ const T& at(const T* base, unsigned off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T& at(T* base, unsigned off) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off * (unsigned)sizeof(T));
}

// Moves each thread's values v[i] (slot i at LDS position pos_out(i) of the
// pencil plane, negative: unused) to its next-stage inputs, v[i] read back
// from pos_in(i), real and imaginary parts in turn through one FP64 plane.
// Positions are recomputed per slot rather than kept in registers.
template <int CNT, typename PO, typename PI>
__device__ __forceinline__ void lds_exchange(double* plane, Cx* v, PO pos_out, PI pos_in) {
  lds_barrier();
#pragma unroll
  for (int i = 0; i < CNT; ++i) {
    const int q = pos_out(i);
    if (q >= 0) plane[q] = v[i].x;
  }
  lds_barrier();
#pragma unroll
  for (int i = 0; i < CNT; ++i) {
    const int q = pos_in(i);
    if (q >= 0) v[i].x = plane[q];
  }
  lds_barrier();
#pragma unroll
  for (int i = 0; i < CNT; ++i) {
    const int q = pos_out(i);
    if (q >= 0) plane[q] = v[i].y;
  }
  lds_barrier();
#pragma unroll
  for (int i = 0; i < CNT; ++i) {
    const int q = pos_in(i);
    if (q >= 0) v[i].y = plane[q];
  }
}

// Tiles per block of the register-staged passes: about `rounds` blocks per
// resident slot on the device (blocks resident at once: occupancy of this
// kernel times the CU count, cached per kernel and device), at most 8.  A
// block transforms its tiles in turn and loads the next one's inputs while it
// transforms the current one; the block scheduler still balances the load
// when the pass shares the GPU with the deposition.
inline int stream_tpb(const void* kernel, size_t lds, long long ntiles, int rounds = 2) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, long long> cap;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  long long c;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cap.find({kernel, dev});
    if (it != cap.end()) {
      c = it->second;
    } else {
      int nb = 0, cus = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kFftThreads, lds) != hipSuccess ||
          nb <= 0)
        nb = 1;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          cus <= 0)
        cus = 1;
      c = (long long)nb * cus;
      cap[{kernel, dev}] = c;
    }
  }
  const long long t = ntiles / (rounds * c);
  return (int)(t < 1 ? 1 : (t > 8 ? 8 : t));
}

template <typename R, int R1, int R2, int R3, bool INV, bool IN_E, bool OUT_E, int EPI, int RES,
          bool WGT>
__global__ __launch_bounds__(kFftThreads) void k_pass_stream(PassArgs<R> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double* plane = reinterpret_cast<double*>(lds_raw);
  constexpr int n = R1 * R2 * R3;
  constexpr int P = reg_pencils<n>();
  constexpr int pitch = reg_pitch<n>();
  constexpr int RL = R3 > 1 ? R3 : R2;  // radix of the last stage
  constexpr int nq1 = n / R1, nsL = n / RL;
  constexpr bool PF = sizeof(R) == 4;   // prefetch (FP32 storage)
  using T2 = typename V2<R>::T;
  const Geom& g = a.g;
  int tid = threadIdx.x;
  const unsigned in_e = IN_E ? 1u : (unsigned)g.in_e, out_e = OUT_E ? 1u : (unsigned)g.out_e;
  const unsigned in_f = (unsigned)g.in_f, out_f = (unsigned)g.out_f;
  // the axis twiddles, copied to LDS after the exchange plane (read after an
  // exchange barrier, at LDS rather than L2 latency), then the tile's 4096
  // residual bytes: they move between global memory and LDS 16 contiguous
  // bytes per lane, and between LDS and the lanes' strided elements a byte
  // at a time
  double2* twl = reinterpret_cast<double2*>(plane + ((P * pitch + 1) & ~1));
  const double2* tw = twl;
  unsigned char* resl = reinterpret_cast<unsigned char*>(twl + n);
  static_assert((RES & RES_IN) == 0 || IN_E, "input residuals: tiles of whole contiguous rows");
  static_assert((RES & RES_OUT) == 0 || !OUT_E, "output residuals: pencil-fastest stores");
  const long long ntiles = g.blocks;
  const unsigned tiles_f = (unsigned)g.tiles_f;
  // stage-1 slot i = t R1 + r: pencil p, element j + r nq1
  auto in_pos = [&](int t, int pc, unsigned* o) __attribute__((always_inline)) {
    const int b = tid + t * kFftThreads;
    int p = IN_E ? b / nq1 : b % P;
    const int j = IN_E ? b % nq1 : b / P;
    p = p < pc ? p : pc - 1;  // pencils past a partial tile read the last one
    *o = (unsigned)p * in_f + (unsigned)j * in_e;
  };
  T2 raw[16];
  uint4 rq;  // RES_IN: bytes 16 tid .. 16 tid + 15 of the tile's residuals (P whole rows)
  // EPI_ACC with a second spectrum: each tile is transformed twice (sub-tile
  // h = 0: a.in, h = 1: a.in2), into the same accumulator values
  const int nsub = (EPI == EPI_ACC && a.in2) ? 2 : 1;
  // sub-tile vt = nsub tile + h
  auto load = [&](long long vt) __attribute__((always_inline)) {
    const long long tile = nsub == 2 ? vt >> 1 : vt;
    const bool second = nsub == 2 && (vt & 1);
    const unsigned s = (unsigned)(tile / tiles_f);
    const int f0 = (int)((unsigned)tile - s * tiles_f) * P;
    const int pc = min(P, g.nf - f0);
    const long long ib = (long long)s * g.in_s + (long long)f0 * g.in_f;
    const T2* src = (second ? a.in2 : a.in) + ib;
#pragma unroll
    for (int t = 0; t < 16 / R1; ++t) {
      unsigned o;
      in_pos(t, pc, &o);
#pragma unroll
      for (int r = 0; r < R1; ++r) raw[t * R1 + r] = at(src, o + (unsigned)(r * nq1) * in_e);
    }
    if constexpr ((RES & RES_IN) != 0) {
      // past a partial tile's rows: the tile's first bytes (never used)
      const unsigned q = 16u * (unsigned)tid < (unsigned)(pc * n) ? (unsigned)tid : 0u;
      rq = at(reinterpret_cast<const uint4*>((second ? a.res_in2 : a.res_in) + ib), q);
    }
  };
  const long long nblk = (ntiles + a.tpb - 1) / a.tpb;
  if ((long long)blockIdx.x >= nblk) {
    // the Nyquist-plane pencils: one per block after the tile blocks (the
    // second spectrum's after the first's, into the same sums)
    if (a.nyq) nyq_pencil<INV, EPI>(a, (int)((long long)blockIdx.x - nblk), plane);
    if constexpr (EPI == EPI_ACC) {
      if (a.nyq2) {
        PassArgs<R> b = a;
        b.nyq = a.nyq2;
        nyq_pencil<INV, EPI>(b, (int)((long long)blockIdx.x - nblk), plane);
      }
    }
    return;
  }
  const long long vend = min(ntiles, (long long)(blockIdx.x + 1) * a.tpb) * nsub;
  long long vt = (long long)blockIdx.x * a.tpb * nsub;
  if (PF) load(vt);
  for (int i = threadIdx.x; i < n; i += kFftThreads) twl[i] = a.ax.tw64[i];
  double accv[EPI == EPI_ACC ? 16 : 1];  // this tile's accumulator values (EPI_ACC)
  for (; vt < vend; ++vt) {
    tid = opaque(tid);
    if (!PF) load(vt);
    const long long tile = nsub == 2 ? vt >> 1 : vt;
    const bool first = nsub == 2 && !(vt & 1);  // a first spectrum's sub-tile (accumulator kept)
    const unsigned s = (unsigned)(tile / tiles_f);
    const int f0 = (int)((unsigned)tile - s * tiles_f) * P;
    const int pc = min(P, g.nf - f0);
    const long long ib = (long long)s * g.in_s + (long long)f0 * g.in_f;
    const long long ob = (long long)s * g.out_s + (long long)f0 * g.out_f;
    Cx v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = mk((double)raw[i].x, (double)raw[i].y);
    if constexpr ((RES & RES_IN) != 0) {
      // the previous tile's last reads of resl preceded its exchange barriers
      reinterpret_cast<uint4*>(resl)[tid] = rq;
      lds_barrier();
#pragma unroll
      for (int t = 0; t < 16 / R1; ++t) {
        const int b = tid + t * kFftThreads;
        const int q = (b / nq1) * n + b % nq1;  // row p, element j (in_f = n, in_e = 1)
#pragma unroll
        for (int r = 0; r < R1; ++r) {
          const unsigned rb = resl[q + r * nq1];
          const int i = t * R1 + r;
          v[i].x += res_value(rb & 15u, raw[i].x);
          v[i].y += res_value(rb >> 4, raw[i].y);
        }
      }
    }
    if constexpr (WGT) {
#pragma unroll
      for (int t = 0; t < 16 / R1; ++t) {
        unsigned o;
        in_pos(t, pc, &o);
#pragma unroll
        for (int r = 0; r < R1; ++r) v[t * R1 + r] = cscale(at(a.w + ib, o + (unsigned)(r * nq1) * in_e), v[t * R1 + r]);
      }
    }
    if (PF && vt + 1 < vend) load(vt + 1);
    // the accumulator values this tile adds to, loaded now so that their
    // latency passes during the transform (pencils past a partial tile read
    // the last one and are not stored); with two spectra, loaded with the
    // first and kept in registers through the second
    if constexpr (EPI == EPI_ACC) {
      if (!(nsub == 2 && (vt & 1))) {
#pragma unroll
        for (int t = 0; t < 16 / RL; ++t) {
          const int b = tid + t * kFftThreads;
          int p = OUT_E ? b / nsL : b % P;
          const int k = OUT_E ? b % nsL : b / P;
          p = p < pc ? p : pc - 1;
          const unsigned o = (unsigned)p * out_f + (unsigned)k * out_e;
#pragma unroll
          for (int r = 0; r < RL; ++r) accv[t * RL + r] = at(a.acc + ob, o + (unsigned)(r * nsL) * out_e);
        }
      }
    }
    // stage 1 (span 1): DFT_R1; output position j R1 + r
#pragma unroll
    for (int t = 0; t < 16 / R1; ++t) dft<R1, INV>(v + t * R1);
    auto out1 = [&](int i) __attribute__((always_inline)) {
      const int b = tid + (i / R1) * kFftThreads, r = i % R1;
      const int p = IN_E ? b / nq1 : b % P, j = IN_E ? b % nq1 : b / P;
      return p * pitch + reg_pos(j * R1 + r);
    };
    // last stage inputs: element k + r nsL of pencil p
    auto inL = [&](int i) __attribute__((always_inline)) {
      const int b = tid + (i / RL) * kFftThreads, r = i % RL;
      const int p = OUT_E ? b / nsL : b % P, k = OUT_E ? b % nsL : b / P;
      return p * pitch + reg_pos(k + r * nsL);
    };
    if constexpr (R3 > 1) {
      // middle stage (span R1, radix R2): LDS -> DFT_R2 -> LDS
      constexpr int nq2 = n / R2, ns2 = R1;
      auto in2 = [&](int i) __attribute__((always_inline)) {
        const int b = tid + (i / R2) * kFftThreads, r = i % R2;
        return (b % P) * pitch + reg_pos(b / P + r * nq2);
      };
      auto out2 = [&](int i) __attribute__((always_inline)) {
        const int b = tid + (i / R2) * kFftThreads, r = i % R2;
        const int j = b / P, k = j % ns2;
        return (b % P) * pitch + reg_pos((j / ns2) * ns2 * R2 + k + r * ns2);
      };
      lds_exchange<16>(plane, v, out1, in2);
#pragma unroll
      for (int t = 0; t < 16 / R2; ++t) {
        const int b = tid + t * kFftThreads;
        twiddle_in<R2, INV>(v + t * R2, (b / P) % ns2, n / (ns2 * R2), tw);
        dft<R2, INV>(v + t * R2);
      }
      lds_exchange<16>(plane, v, out2, inL);
    } else {
      lds_exchange<16>(plane, v, out1, inL);
    }
#pragma unroll
    for (int t = 0; t < 16 / RL; ++t) {
      const int b = tid + t * kFftThreads;
      const int p = OUT_E ? b / nsL : b % P, k = OUT_E ? b % nsL : b / P;
      Cx* x = v + t * RL;
      twiddle_in<RL, INV>(x, k, 1, tw);
      dft<RL, INV>(x);
      if (p < pc) {
        const unsigned o = (unsigned)p * out_f + (unsigned)k * out_e;
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const unsigned oe = o + (unsigned)(r * nsL) * out_e;
          if constexpr (EPI == EPI_ACC) {
            const double sum = accv[t * RL + r] + (x[r].x * x[r].x + x[r].y * x[r].y);
            if (first) accv[t * RL + r] = sum;
            else at(a.acc + ob, oe) = sum;
          } else {
            T2 st;
            st.x = (R)x[r].x;
            st.y = (R)x[r].y;
            at(a.out + ob, oe) = st;
            if constexpr ((RES & RES_OUT) != 0)  // element k + r nsL of pencil p
              resl[(k + r * nsL) * P + p] =
                  (unsigned char)(res_nibble(x[r].x, st.x) | (res_nibble(x[r].y, st.y) << 4));
          }
        }
      }
    }
    if constexpr ((RES & RES_OUT) != 0) {
      // the tile's residuals: n rows (elements) of P bytes (pencils), stored
      // at ob + element out_e + pencil (out_f = 1); the next tile writes resl
      // only after its exchange barriers
      lds_barrier();
      const unsigned q = 16u * (unsigned)tid;
      const int row = (int)(q / P), col = (int)(q % P);
      unsigned char* dst = a.res_out + ob;
      if (P >= 16 && (out_e & 15u) == 0 && col + 16 <= pc) {
        at(reinterpret_cast<uint4*>(dst + (size_t)row * out_e + col), 0u) =
            *reinterpret_cast<const uint4*>(resl + q);
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int rw = (int)((q + i) / P), cl = (int)((q + i) % P);
          if (cl < pc) dst[(size_t)rw * out_e + cl] = resl[q + i];
        }
      }
    }
  }
}

// The LDS-staged pass (any n): P pencils loaded into LDS, every stage there.
template <typename R, bool INV, int EPI, int RES>
__global__ __launch_bounds__(kFftThreads) void k_pass_lds(PassArgs<R> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double* lds = reinterpret_cast<double*>(lds_raw);
  const long long blk = blockIdx.x;
  if (blk >= a.g.blocks) {
    nyq_pencil<INV, EPI>(a, (int)(blk - a.g.blocks), lds);
    return;
  }
  const Geom& g = a.g;
  const int n = a.n, P = a.P, pitch = lds_pitch(n);
  const long long s = blk / g.tiles_f;
  const int f0 = (int)(blk - s * g.tiles_f) * P;
  const int pc = min(P, g.nf - f0);
  const long long ib = s * g.in_s + (long long)f0 * g.in_f;
  const long long ob = s * g.out_s + (long long)f0 * g.out_f;
  double* re = lds;
  double* im = lds + P * pitch;
  const bool in_e = g.in_e == 1, out_e = g.out_e == 1;
  for (int idx = threadIdx.x; idx < pc * n; idx += kFftThreads) {
    const int p = in_e ? idx / n : idx % pc, e = in_e ? idx % n : idx / pc;
    const Cx x = pass_load<R, RES>(a, ib + (long long)p * g.in_f + (long long)e * g.in_e);
    re[p * pitch + e] = x.x;
    im[p * pitch + e] = x.y;
  }
  __syncthreads();
  const int cur = run_stages<INV>(lds, n, P, pitch, a.ax);
  re = lds + 2 * cur * P * pitch;
  im = re + P * pitch;
  for (int idx = threadIdx.x; idx < pc * n; idx += kFftThreads) {
    const int p = out_e ? idx / n : idx % pc, k = out_e ? idx % n : idx / pc;
    pass_store<R, EPI, RES>(a, ob + (long long)p * g.out_f + (long long)k * g.out_e,
                            mk(re[p * pitch + k], im[p * pitch + k]));
  }
}

// --- z pass: real rows -> half spectrum --------------------------------------------
// Rows (x, y) of nz reals are read as m = nz/2 complex values Z and
// transformed; the half spectrum follows from the real-pair split
//   X_k = A_k + W_nz^k B_k,  A = (Z_k + conj Z_{m-k}) / 2,  B = (Z_k - conj Z_{m-k}) / 2i,
//   X_m = Re Z_0 - Im Z_0.
// Block = P rows y0..y0+P-1 of one x; output spec[(x m + k) Ny + y] in runs
// of P, nyq[x Ny + y] (FP64).
template <typename R>
struct ZArgs {
  const R* in;                       // density rows (x, y) of nz reals
  typename V2<R>::T* out;            // spec[x][kz][y]
  double2* nyq;                      // nyq[x][y]
  int nx, ny, m;
  int P, tiles_y;
  FftAxis ax;                        // length m
  const double2* tw2;                // W_nz^k, k < nz
  int tpb;                           // tiles per block of the register-staged pass
};

__device__ __forceinline__ Cx split_bin(Cx z, Cx zc_conj_src, const double2* tw2, int k) {
  // zc_conj_src = Z_{m-k} (Z_0 for k = 0)
  const Cx zc = mk(zc_conj_src.x, -zc_conj_src.y);
  const Cx A = cscale(0.5, cadd(z, zc));
  const Cx d = cscale(0.5, csub(z, zc));
  const Cx B = mk(d.y, -d.x);  // d / i
  return cadd(A, cmul(B, ld_tw(tw2, k)));
}

template <typename R>
__device__ __forceinline__ void z_store(const ZArgs<R>& a, int x, int y, int k, Cx v) {
  const long long o = ((long long)x * a.m + k) * a.ny + y;
  store_cx<R>(a.out, o, v);
}

// Register-staged z pass (c2r form below keeps 2048 values per block).
template <int m>
constexpr int zreg_rows() { return kZRegElems / m; }

// Persistent register-staged z pass, m = R1 R2 R3: a tile is P = 4096 / m rows
// y0..y0+P-1 of one x, 16 values per thread and stage.  The last stage pairs
// its butterflies k and ns - k (0 with ns/2) in one thread, which then holds
// Z_k and Z_{m-k} for each of its bins and applies the real-pair split in
// registers.  Block b takes tiles b, b + gridDim.x, ... and loads the next
// tile's rows while it transforms the current one (FP32 storage).
template <typename R, int R1, int R2, int R3>
__global__ __launch_bounds__(kFftThreads) void k_r2c_z_stream(ZArgs<R> a) {
  constexpr int m = R1 * R2 * R3;
  constexpr int P = kRegElems / m;
  constexpr int pitch = reg_pitch<m>();
  constexpr int RL = R3 > 1 ? R3 : R2;
  constexpr int nq1 = m / R1, ns = m / RL, NB = 16 / RL;
  static_assert(NB % 2 == 0, "the last stage pairs butterflies within a thread");
  constexpr bool PF = sizeof(R) == 4;
  using T2 = typename V2<R>::T;
  __shared__ double plane[P * pitch];
  // the axis twiddles W_m^k and the split's W_nz^k (k < m) in LDS, read after
  // an exchange barrier
  __shared__ double2 twl[m], tw2l[m];
  int tid = threadIdx.x;
  const double2* tw = twl;
  const long long ntiles = (long long)a.nx * a.tiles_y;
  T2 raw[16];
  auto load = [&](long long tile) __attribute__((always_inline)) {
    const int x = (int)(tile / a.tiles_y);
    const int y0 = (int)(tile - (long long)x * a.tiles_y) * P;
    const int pc = min(P, a.ny - y0);
    const T2* src = reinterpret_cast<const T2*>(a.in) + ((long long)x * a.ny + y0) * m;
#pragma unroll
    for (int t = 0; t < 16 / R1; ++t) {
      const int b = tid + t * kFftThreads;
      int p = b / nq1;
      const int j = b % nq1;
      p = p < pc ? p : pc - 1;
#pragma unroll
      for (int r = 0; r < R1; ++r) raw[t * R1 + r] = at(src, (unsigned)(p * m + j + r * nq1));
    }
  };
  long long tile = (long long)blockIdx.x * a.tpb;
  const long long tend = min(ntiles, tile + a.tpb);
  if (PF && tile < tend) load(tile);
  for (int i = threadIdx.x; i < m; i += kFftThreads) {
    twl[i] = a.ax.tw64[i];
    tw2l[i] = a.tw2[i];
  }
  for (; tile < tend; ++tile) {
    tid = opaque(tid);
    if (!PF) load(tile);
    const int x = (int)(tile / a.tiles_y);
    const int y0 = (int)(tile - (long long)x * a.tiles_y) * P;
    const int pc = min(P, a.ny - y0);
    Cx v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = mk((double)raw[i].x, (double)raw[i].y);
    if (PF && tile + 1 < tend) load(tile + 1);
    // stage 1: rows along the lanes; output position j R1 + r
#pragma unroll
    for (int t = 0; t < 16 / R1; ++t) dft<R1, false>(v + t * R1);
    auto out1 = [&](int i) __attribute__((always_inline)) {
      const int b = tid + (i / R1) * kFftThreads, r = i % R1;
      return (b / nq1) * pitch + reg_pos((b % nq1) * R1 + r);
    };
    // last stage: slot i = (2u + h) RL + r, pair c = tid + u 256 of row c % P,
    // butterfly kA = c / P (h = 0) or its partner kB (h = 1), bin k + r ns
    auto kpair = [&](int u, int h) __attribute__((always_inline)) {
      const int kp = (tid + u * kFftThreads) / P;
      return h == 0 ? kp : (kp == 0 ? ns / 2 : ns - kp);
    };
    auto inL = [&](int i) __attribute__((always_inline)) {
      const int r = i % RL, h = (i / RL) & 1, u = i / (2 * RL);
      const int p = (tid + u * kFftThreads) % P;
      return p * pitch + reg_pos(kpair(u, h) + r * ns);
    };
    if constexpr (R3 > 1) {
      constexpr int nq2 = m / R2, ns2 = R1;
      auto in2 = [&](int i) __attribute__((always_inline)) {
        const int b = tid + (i / R2) * kFftThreads, r = i % R2;
        return (b % P) * pitch + reg_pos(b / P + r * nq2);
      };
      auto out2 = [&](int i) __attribute__((always_inline)) {
        const int b = tid + (i / R2) * kFftThreads, r = i % R2;
        const int j = b / P, k = j % ns2;
        return (b % P) * pitch + reg_pos((j / ns2) * ns2 * R2 + k + r * ns2);
      };
      lds_exchange<16>(plane, v, out1, in2);
#pragma unroll
      for (int t = 0; t < 16 / R2; ++t) {
        const int b = tid + t * kFftThreads;
        twiddle_in<R2, false>(v + t * R2, (b / P) % ns2, m / (ns2 * R2), tw);
        dft<R2, false>(v + t * R2);
      }
      lds_exchange<16>(plane, v, out2, inL);
    } else {
      lds_exchange<16>(plane, v, out1, inL);
    }
#pragma unroll
    for (int u = 0; u < NB / 2; ++u) {
      const int c = tid + u * kFftThreads;
      const int p = c % P, kA = kpair(u, 0), kB = kpair(u, 1);
      Cx* za = v + (2 * u) * RL;
      Cx* zb = v + (2 * u + 1) * RL;
      twiddle_in<RL, false>(za, kA, 1, tw);
      dft<RL, false>(za);
      twiddle_in<RL, false>(zb, kB, 1, tw);
      dft<RL, false>(zb);
      if (p >= pc) continue;
      T2* dst = a.out + (long long)x * m * a.ny + y0 + p;
      // bin k of this row at dst[k Ny]; Z_{m-k} of bin kA + r ns is za[RL - r]
      // (kA = 0, r > 0), za[0] (kA = r = 0) or zb[RL - 1 - r]; of bin kB + r ns
      // it is zb[RL - 1 - r] (kA = 0) or za[RL - 1 - r]
      const bool self = kA == 0;
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int k1 = kA + r * ns, k2 = kB + r * ns;
        const Cx m1 = csel(self, za[r == 0 ? 0 : RL - r], zb[RL - 1 - r]);
        const Cx m2 = csel(self, zb[RL - 1 - r], za[RL - 1 - r]);
        const Cx x1 = split_bin(za[r], m1, tw2l, k1);
        const Cx x2 = split_bin(zb[r], m2, tw2l, k2);
        T2 s1, s2;
        s1.x = (R)x1.x;
        s1.y = (R)x1.y;
        s2.x = (R)x2.x;
        s2.y = (R)x2.y;
        at(dst, (unsigned)(k1 * a.ny)) = s1;
        at(dst, (unsigned)(k2 * a.ny)) = s2;
      }
      if (self) a.nyq[(long long)x * a.ny + y0 + p] = make_double2(za[0].x - za[0].y, 0.0);
    }
  }
}

// LDS-staged z pass (any m): P rows per block, every stage in LDS.
template <typename R>
__global__ __launch_bounds__(kFftThreads) void k_r2c_z(ZArgs<R> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double* lds = reinterpret_cast<double*>(lds_raw);
  const int m = a.m, P = a.P, pitch = lds_pitch(m);
  const int x = blockIdx.x / a.tiles_y;
  const int y0 = (blockIdx.x - x * a.tiles_y) * P;
  const int pc = min(P, a.ny - y0);
  const auto* src = reinterpret_cast<const typename V2<R>::T*>(a.in) + ((long long)x * a.ny + y0) * m;
  double* re = lds;
  double* im = lds + P * pitch;
  for (int idx = threadIdx.x; idx < pc * m; idx += kFftThreads) {
    const int p = idx / m, e = idx - p * m;
    const Cx v = load_cx<R>(src, idx);
    re[p * pitch + e] = v.x;
    im[p * pitch + e] = v.y;
  }
  __syncthreads();
  const int cur = run_stages<false>(lds, m, P, pitch, a.ax);
  re = lds + 2 * cur * P * pitch;
  im = re + P * pitch;
  for (int idx = threadIdx.x; idx < pc * m; idx += kFftThreads) {
    const int p = idx % pc, k = idx / pc;
    const int kk = k == 0 ? 0 : m - k;
    const Cx z = mk(re[p * pitch + k], im[p * pitch + k]);
    const Cx zm = mk(re[p * pitch + kk], im[p * pitch + kk]);
    z_store<R>(a, x, y0 + p, k, split_bin(z, zm, a.tw2, k));
  }
  for (int p = threadIdx.x; p < pc; p += kFftThreads)
    a.nyq[(long long)x * a.ny + y0 + p] = make_double2(re[p * pitch] - im[p * pitch], 0.0);
}

// --- inverse z pass: half spectrum -> real rows ---------------------------------------
// Real adjoint of the R2C along z: x_n = Re sum_{k=0..m} X_k e^{+2 pi i k n / nz}.
// With X'_0 = Re X_0, X'_m = Re X_m and X'_k = X_k / 2 inside, this is the
// Hermitian inverse of X': Z_k = (X'_k + conj X'_{m-k}) + i W^{-k} (X'_k - conj X'_{m-k})
// and x_{2j} + i x_{2j+1} = sum_k Z_k e^{+2 pi i jk/m}.  Reads spec[x][kz][y]
// in runs of P (pencils y), writes density rows.
template <typename R>
struct CArgs {
  const typename V2<R>::T* spec;  // spec[x][kz][y]
  const double2* nyq;              // nyq[x][y]
  R* out;                          // rows (x, y) of nz reals
  int nx, ny, m;
  int P, tiles_y;
  FftAxis ax;
  const double2* tw2;
};

template <typename R>
__device__ __forceinline__ Cx c2r_input(const CArgs<R>& a, int x, int y, int k) {
  const long long base = (long long)x * a.m * a.ny + y;
  Cx xk, xc;
  if (k == 0) {
    xk = mk(load_cx<R>(a.spec, base).x, 0.0);
    xc = mk(a.nyq[(long long)x * a.ny + y].x, 0.0);
  } else {
    const Cx u = load_cx<R>(a.spec, base + (long long)k * a.ny);
    const Cx w = load_cx<R>(a.spec, base + (long long)(a.m - k) * a.ny);
    xk = cscale(0.5, u);
    xc = mk(0.5 * w.x, -0.5 * w.y);
  }
  const Cx e = cadd(xk, xc);
  const Cx o = cmulc(csub(xk, xc), ld_tw(a.tw2, k));  // (X - conj X') W^{-k}
  return mk(e.x - o.y, e.y + o.x);                    // e + i o
}

template <typename R, int R1, int R2>
__global__ __launch_bounds__(kFftThreads) void k_c2r_z_reg(CArgs<R> a) {
  constexpr int m = R1 * R2;
  constexpr int P = zreg_rows<m>();
  constexpr int pitch = reg_pitch<m>();
  constexpr int nq1 = m / R1, ns = m / R2;
  constexpr int T1 = (P * nq1 + kFftThreads - 1) / kFftThreads;
  constexpr int T2 = (P * ns + kFftThreads - 1) / kFftThreads;
  constexpr int VPT = T1 * R1 > T2 * R2 ? T1 * R1 : T2 * R2;
  __shared__ double plane[P * pitch];
  const int x = blockIdx.x / a.tiles_y;
  const int y0 = (blockIdx.x - x * a.tiles_y) * P;
  const int pc = min(P, a.ny - y0);
  const double2* tw = a.ax.tw64;
  const int tid = threadIdx.x;
  Cx v[VPT];
  // stage 1: inputs Z_k built from spec runs (pencil-fastest lanes)
  auto out1 = [&](int i) __attribute__((always_inline)) {
    const int b = tid + (i / R1) * kFftThreads, r = i % R1;
    if (i >= T1 * R1 || b >= P * nq1) return -1;
    return (b % P) * pitch + reg_pos((b / P) * R1 + r);
  };
  auto in2 = [&](int i) __attribute__((always_inline)) {
    const int b = tid + (i / R2) * kFftThreads, r = i % R2;
    if (i >= T2 * R2 || b >= P * ns) return -1;
    return (b / ns) * pitch + reg_pos(b % ns + r * ns);
  };
#pragma unroll
  for (int t = 0; t < T1; ++t) {
    const int b = tid + t * kFftThreads;
    if (b >= P * nq1) break;
    const int p = b % P, j = b / P;
    Cx* xv = v + t * R1;
#pragma unroll
    for (int r = 0; r < R1; ++r) xv[r] = p < pc ? c2r_input<R>(a, x, y0 + p, j + r * nq1) : mk(0, 0);
    dft<R1, true>(xv);
  }
  lds_exchange<VPT>(plane, v, out1, in2);
  R* dst = a.out + ((long long)x * a.ny + y0) * (2LL * m);
#pragma unroll
  for (int t = 0; t < T2; ++t) {
    const int b = tid + t * kFftThreads;
    if (b >= P * ns) break;
    const int p = b / ns, k = b % ns;
    Cx* xv = v + t * R2;
    twiddle_in<R2, true>(xv, k, 1, tw);
    dft<R2, true>(xv);
    if (p < pc) {
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const long long o = 2 * ((long long)p * m + k + r * ns);
        dst[o] = (R)xv[r].x;
        dst[o + 1] = (R)xv[r].y;
      }
    }
  }
}

template <typename R>
__global__ __launch_bounds__(kFftThreads) void k_c2r_z(CArgs<R> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double* lds = reinterpret_cast<double*>(lds_raw);
  const int m = a.m, P = a.P, pitch = lds_pitch(m);
  const int x = blockIdx.x / a.tiles_y;
  const int y0 = (blockIdx.x - x * a.tiles_y) * P;
  const int pc = min(P, a.ny - y0);
  double* re = lds;
  double* im = lds + P * pitch;
  for (int idx = threadIdx.x; idx < pc * m; idx += kFftThreads) {
    const int p = idx % pc, k = idx / pc;
    const Cx z = c2r_input<R>(a, x, y0 + p, k);
    re[p * pitch + k] = z.x;
    im[p * pitch + k] = z.y;
  }
  __syncthreads();
  const int cur = run_stages<true>(lds, m, P, pitch, a.ax);
  re = lds + 2 * cur * P * pitch;
  im = re + P * pitch;
  R* dst = a.out + ((long long)x * a.ny + y0) * (2LL * m);
  for (int idx = threadIdx.x; idx < pc * m; idx += kFftThreads) {
    const int p = idx / m, j = idx - p * m;
    dst[2 * (long long)idx] = (R)re[p * pitch + j];
    dst[2 * (long long)idx + 1] = (R)im[p * pitch + j];
  }
}

// --- launch helpers --------------------------------------------------------------------
// LDS bytes one workgroup may allocate on the current device (queried once per device).
inline size_t max_lds_per_block() {
  static int cached[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  if (cached[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0)
      v = 65536;
    cached[dev] = v;
  }
  return (size_t)cached[dev];
}

// pencils per block of the LDS-staged passes (4 FP64 planes of P * pitch)
inline int lds_pencils(int n) {
  int P = 2048 / (n > 0 ? n : 1);
  if (P > 16) P = 16;
  if (P < 1) P = 1;
  return P;
}
inline size_t lds_bytes(int n, int P) { return (size_t)P * lds_pitch(n) * 4 * sizeof(double); }

// Events the next stream-kernel launch of a pass records with its own
// dispatch (its begin and end: the frame loop's stage timing), or null.
thread_local hipEvent_t t_pass_ev[2] = {nullptr, nullptr};
thread_local bool t_pass_ev_used = false;

template <typename K, typename A>
void pass_launch(K kernel, dim3 grid, size_t lds, hipStream_t s, const A& a) {
  if (t_pass_ev[0]) {
    hipExtLaunchKernelGGL(kernel, grid, dim3(kFftThreads), lds, s, t_pass_ev[0], t_pass_ev[1], 0,
                          a);
    t_pass_ev_used = true;
  } else {
    hipLaunchKernelGGL(kernel, grid, dim3(kFftThreads), lds, s, a);
  }
}

// Radix split of the register-staged passes for n, or 0.
inline int reg_split(int n, int* r1, int* r2, int* r3) {
  switch (n) {
    case 16: *r1 = 4; *r2 = 4; *r3 = 1; return 1;
    case 32: *r1 = 8; *r2 = 4; *r3 = 1; return 1;
    case 64: *r1 = 8; *r2 = 8; *r3 = 1; return 1;
    case 128: *r1 = 16; *r2 = 8; *r3 = 1; return 1;
    case 256: *r1 = 16; *r2 = 16; *r3 = 1; return 1;
    case 512: *r1 = 8; *r2 = 8; *r3 = 8; return 1;
    case 1024: *r1 = 16; *r2 = 8; *r3 = 8; return 1;
    default: return 0;
  }
}

template <typename R, int R1, int R2, int R3, bool INV, bool IN_E, bool OUT_E, int EPI, int RES,
          bool WGT>
void launch_stream(size_t lds, hipStream_t s, PassArgs<R> a) {
  const void* k =
      reinterpret_cast<const void*>(&k_pass_stream<R, R1, R2, R3, INV, IN_E, OUT_E, EPI, RES, WGT>);
  a.tpb = stream_tpb(k, lds, a.g.blocks);
  const long long grid = (a.g.blocks + a.tpb - 1) / a.tpb + (a.nyq ? a.nq_pencils : 0);
  if (grid == 0) return;
  pass_launch(k_pass_stream<R, R1, R2, R3, INV, IN_E, OUT_E, EPI, RES, WGT>, dim3((unsigned)grid),
              lds, s, a);
}

template <typename R, bool INV, bool IN_E, bool OUT_E, int EPI, int RES, bool WGT>
void launch_reg_n(int n, size_t lds, hipStream_t s, const PassArgs<R>& a) {
  switch (n) {
    case 16: launch_stream<R, 4, 4, 1, INV, IN_E, OUT_E, EPI, RES, WGT>(lds, s, a); break;
    case 32: launch_stream<R, 8, 4, 1, INV, IN_E, OUT_E, EPI, RES, WGT>(lds, s, a); break;
    case 64: launch_stream<R, 8, 8, 1, INV, IN_E, OUT_E, EPI, RES, WGT>(lds, s, a); break;
    case 128: launch_stream<R, 16, 8, 1, INV, IN_E, OUT_E, EPI, RES, WGT>(lds, s, a); break;
    case 256: launch_stream<R, 16, 16, 1, INV, IN_E, OUT_E, EPI, RES, WGT>(lds, s, a); break;
    case 512: launch_stream<R, 8, 8, 8, INV, IN_E, OUT_E, EPI, RES, WGT>(lds, s, a); break;
    default: launch_stream<R, 16, 8, 8, INV, IN_E, OUT_E, EPI, RES, WGT>(lds, s, a); break;
  }
}

// The pass kinds the transform uses (forward z is separate):
enum PassKind {
  PK_Y = 0,       // forward y: spec[x][kz][:] -> rho[kz][:][x]
  PK_X_ACC,       // forward x, |X|^2 into sf
  PK_X_STORE,     // forward x, X in place
  PK_X_INV,       // inverse x (weights), in place
  PK_Y_INV,       // inverse y: rho[kz][:][x] -> spec[x][kz][:]
};

// choice: 0 = default (register-staged when n has a split), kTuneReg = that,
// 1..16 = the LDS-staged pass with that many pencils.
template <typename R>
hipError_t run_pass(PassArgs<R> a, PassKind kind, hipStream_t s, int choice) {
  int r1, r2, r3;
  const int n = a.n;
  const size_t nyq_lds = (size_t)4 * lds_pitch(n) * sizeof(double);
  bool reg = (choice == 0 || choice == kTuneReg) && reg_split(n, &r1, &r2, &r3);
  size_t lds = 0;
  if (reg) {
    a.P = kRegElems / n;
    // exchange plane (rounded to whole double2), the n axis twiddles, the
    // tile's residual bytes
    lds = (((size_t)a.P * (n + n / 16 + 1) + 1) & ~(size_t)1) * sizeof(double) +
          (size_t)n * sizeof(double2) + (size_t)kRegElems;
  } else {
    a.P = (choice >= 1 && choice <= 16) ? choice : lds_pencils(n);
    while (a.P > 1 && lds_bytes(n, a.P) > max_lds_per_block()) a.P /= 2;
    lds = lds_bytes(n, a.P);
  }
  if (nyq_lds > lds) lds = nyq_lds;
  if (lds > max_lds_per_block()) return hipErrorInvalidValue;
  a.g.tiles_f = (a.g.nf + a.P - 1) / a.P;
  a.g.blocks *= a.g.tiles_f;   // caller set the slow count
  const long long total = a.g.blocks + (a.nyq ? a.nq_pencils : 0);
  if (total == 0) return hipSuccess;
  const dim3 g((unsigned)total);
  const bool f32 = sizeof(R) == 4;
  switch (kind) {
    case PK_Y:
      if (reg) {
        if (f32) launch_reg_n<R, false, true, false, EPI_STORE, RES_OUT, false>(n, lds, s, a);
        else launch_reg_n<R, false, true, false, EPI_STORE, RES_NONE, false>(n, lds, s, a);
      } else {
        if (f32) hipLaunchKernelGGL((k_pass_lds<R, false, EPI_STORE, RES_OUT>), g, dim3(kFftThreads), lds, s, a);
        else hipLaunchKernelGGL((k_pass_lds<R, false, EPI_STORE, RES_NONE>), g, dim3(kFftThreads), lds, s, a);
      }
      break;
    case PK_X_ACC:
      if (reg) {
        if (f32) launch_reg_n<R, false, true, true, EPI_ACC, RES_IN, false>(n, lds, s, a);
        else launch_reg_n<R, false, true, true, EPI_ACC, RES_NONE, false>(n, lds, s, a);
      } else {
        if (f32) hipLaunchKernelGGL((k_pass_lds<R, false, EPI_ACC, RES_IN>), g, dim3(kFftThreads), lds, s, a);
        else hipLaunchKernelGGL((k_pass_lds<R, false, EPI_ACC, RES_NONE>), g, dim3(kFftThreads), lds, s, a);
      }
      break;
    case PK_X_STORE:
      if (reg) {
        if (f32) launch_reg_n<R, false, true, true, EPI_STORE, RES_IN, false>(n, lds, s, a);
        else launch_reg_n<R, false, true, true, EPI_STORE, RES_NONE, false>(n, lds, s, a);
      } else {
        if (f32) hipLaunchKernelGGL((k_pass_lds<R, false, EPI_STORE, RES_IN>), g, dim3(kFftThreads), lds, s, a);
        else hipLaunchKernelGGL((k_pass_lds<R, false, EPI_STORE, RES_NONE>), g, dim3(kFftThreads), lds, s, a);
      }
      break;
    case PK_X_INV:
      if (reg) launch_reg_n<R, true, true, true, EPI_STORE, RES_NONE, true>(n, lds, s, a);
      else hipLaunchKernelGGL((k_pass_lds<R, true, EPI_STORE, RES_NONE>), g, dim3(kFftThreads), lds, s, a);
      break;
    default:  // PK_Y_INV
      if (reg) launch_reg_n<R, true, false, true, EPI_STORE, RES_NONE, false>(n, lds, s, a);
      else hipLaunchKernelGGL((k_pass_lds<R, true, EPI_STORE, RES_NONE>), g, dim3(kFftThreads), lds, s, a);
      break;
  }
  return hipGetLastError();
}

template <typename R>
typename V2<R>::T* cx(void* p) {
  return static_cast<typename V2<R>::T*>(p);
}

// Register-staged z pass rows per block for half-length m, 0 without a split.
inline int z_reg_rows(int m) {
  switch (m) {
    case 32: return zreg_rows<32>();
    case 64: return zreg_rows<64>();
    case 128: return zreg_rows<128>();
    case 256: return zreg_rows<256>();
    default: return 0;
  }
}

// Half lengths m the persistent z pass handles (the others take k_r2c_z).
inline bool z_stream_ok(int m) {
  return m == 16 || m == 32 || m == 64 || m == 128 || m == 256 || m == 512 || m == 1024;
}

template <typename R, int R1, int R2, int R3>
void launch_z_stream(long long ntiles, hipStream_t s, ZArgs<R> a) {
  // one round of blocks: at 256^3 the z pass alone takes 27.8 us against
  // 30.0 with two, and the pipelined frame loop is unchanged (the y and x
  // passes lose in the frame loop with one round, so they keep two)
  a.tpb = stream_tpb(reinterpret_cast<const void*>(&k_r2c_z_stream<R, R1, R2, R3>), 0, ntiles, 1);
  const long long grid = (ntiles + a.tpb - 1) / a.tpb;
  if (grid == 0) return;
  pass_launch(k_r2c_z_stream<R, R1, R2, R3>, dim3((unsigned)grid), 0, s, a);
}

template <typename R>
hipError_t pass_z(Plan& p, const void* density, hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  ZArgs<R> a;
  a.in = static_cast<const R*>(density);
  a.out = cx<R>(p.spec);
  a.nyq = static_cast<double2*>(p.spec_nyq);
  a.nx = nx;
  a.ny = ny;
  a.m = m;
  a.ax = p.ax_zc;
  a.tw2 = p.tw_r2c64;
  if (z_stream_ok(m) && (p.tune_z == 0 || p.tune_z == kTuneReg)) {
    a.P = kRegElems / m;
    a.tiles_y = (ny + a.P - 1) / a.P;
    const long long ntiles = (long long)nx * a.tiles_y;
    switch (m) {
      case 16: launch_z_stream<R, 4, 4, 1>(ntiles, s, a); break;
      case 32: launch_z_stream<R, 8, 4, 1>(ntiles, s, a); break;
      case 64: launch_z_stream<R, 8, 8, 1>(ntiles, s, a); break;
      case 128: launch_z_stream<R, 16, 8, 1>(ntiles, s, a); break;
      case 256: launch_z_stream<R, 16, 4, 4>(ntiles, s, a); break;
      case 512: launch_z_stream<R, 16, 8, 4>(ntiles, s, a); break;
      default: launch_z_stream<R, 16, 8, 8>(ntiles, s, a); break;
    }
    return hipGetLastError();
  }
  a.P = (p.tune_z >= 1 && p.tune_z <= 16) ? p.tune_z : lds_pencils(m);
  while (a.P > 1 && lds_bytes(m, a.P) > max_lds_per_block()) a.P /= 2;
  a.tiles_y = (ny + a.P - 1) / a.P;
  hipLaunchKernelGGL((k_r2c_z<R>), dim3((unsigned)((long long)nx * a.tiles_y)), dim3(kFftThreads),
                     lds_bytes(m, a.P), s, a);
  return hipGetLastError();
}

template <typename R>
PassArgs<R> base_args(Plan& p, int n, const FftAxis& ax) {
  PassArgs<R> a{};
  a.n = n;
  a.ax = ax;
  return a;
}

// y pass: spec[x][kz][0..Ny) -> rho[kz][ky][x]; Nyquist plane nyq[x][:] in place
template <typename R>
hipError_t pass_y(Plan& p, hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  PassArgs<R> a = base_args<R>(p, ny, p.ax_y);
  a.in = cx<R>(p.spec);
  a.out = cx<R>(p.rho);
  a.res_out = static_cast<unsigned char*>(p.res);
  a.g.in_s = ny;                         // slow = kz
  a.g.in_f = (long long)m * ny;          // fast = x
  a.g.in_e = 1;
  a.g.out_s = (long long)ny * nx;
  a.g.out_f = 1;
  a.g.out_e = nx;
  a.g.nf = nx;
  a.g.blocks = m;
  a.nyq = static_cast<double2*>(p.spec_nyq);
  a.nq_p = ny;
  a.nq_e = 1;
  a.nq_pencils = nx;
  return run_pass<R>(a, PK_Y, s, p.tune_y);
}

// x pass: rho[kz][ky][0..Nx) in place (F) or |F|^2 into sf_main; the Nyquist
// plane nyq[:][y] in place or into sf_nyq
template <typename R>
PassArgs<R> x_args(Plan& p, void* rho, void* res, void* nyq) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  PassArgs<R> a = base_args<R>(p, nx, p.ax_x);
  a.in = cx<R>(rho);
  a.out = cx<R>(rho);
  a.acc = p.sf_main;
  a.res_in = static_cast<const unsigned char*>(res);
  a.g.in_s = 0;
  a.g.in_f = nx;                         // fast = row kz * Ny + ky
  a.g.in_e = 1;
  a.g.out_s = 0;
  a.g.out_f = nx;
  a.g.out_e = 1;
  a.g.nf = m * ny;
  a.g.blocks = 1;
  a.nyq = static_cast<double2*>(nyq);
  a.nyq_acc = p.sf_nyq;
  a.nq_p = 1;
  a.nq_e = ny;
  a.nq_pencils = ny;
  return a;
}

template <typename R>
hipError_t pass_x(Plan& p, bool accumulate, hipStream_t s) {
  return run_pass<R>(x_args<R>(p, p.rho, p.res, p.spec_nyq), accumulate ? PK_X_ACC : PK_X_STORE, s,
                     p.tune_x);
}

// Both frames' x passes with |F|^2 into sf_main / sf_nyq, the frame in slot
// sa first: one read and one write of the FP64 sums for the two frames, the
// same additions in the same order as two single passes.
template <typename R>
hipError_t pass_x_pair(Plan& p, int sa, int sb, hipStream_t s) {
  PassArgs<R> a = x_args<R>(p, p.rho_buf[sa], p.res_buf[sa], p.nyq_buf[sa]);
  a.in2 = cx<R>(p.rho_buf[sb]);
  a.res_in2 = static_cast<const unsigned char*>(p.res_buf[sb]);
  a.nyq2 = static_cast<double2*>(p.nyq_buf[sb]);
  return run_pass<R>(a, PK_X_ACC, s, p.tune_x);
}

template <typename R>
hipError_t forward_impl(Plan& p, const void* density, bool accumulate, hipStream_t s) {
  hipError_t e = pass_z<R>(p, density, s);
  if (e != hipSuccess) return e;
  e = pass_y<R>(p, s);
  if (e != hipSuccess) return e;
  return pass_x<R>(p, accumulate, s);
}

// Real adjoint of the forward transform applied to the stored spectrum
// (rho[kz][ky][kx] + nyq after a storing forward pass), each bin first
// multiplied by the FP64 weight w (w_main[kz][ky][kx], w_nyq[x][y]):
//   out(r) = Re sum_{kx, ky, kz in [0, Nz/2]} w(q) X(q) e^{+i q.r}.
template <typename R>
hipError_t inverse_impl(Plan& p, const double* w_main, const double* w_nyq, void* out,
                        hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  {  // x: in place, weights on load
    PassArgs<R> a = base_args<R>(p, nx, p.ax_x);
    a.in = cx<R>(p.rho);
    a.out = cx<R>(p.rho);
    a.w = w_main;
    a.g.in_s = 0;
    a.g.in_f = nx;
    a.g.in_e = 1;
    a.g.out_s = 0;
    a.g.out_f = nx;
    a.g.out_e = 1;
    a.g.nf = m * ny;
    a.g.blocks = 1;
    a.nyq = static_cast<double2*>(p.spec_nyq);
    a.nyq_w = w_nyq;
    a.nq_p = 1;
    a.nq_e = ny;
    a.nq_pencils = ny;
    const hipError_t e = run_pass<R>(a, PK_X_INV, s, 0);
    if (e != hipSuccess) return e;
  }
  {  // y: rho[kz][:][x] -> spec[x][kz][:]
    PassArgs<R> a = base_args<R>(p, ny, p.ax_y);
    a.in = cx<R>(p.rho);
    a.out = cx<R>(p.spec);
    a.g.in_s = (long long)ny * nx;       // slow = kz
    a.g.in_f = 1;                        // fast = x
    a.g.in_e = nx;
    a.g.out_s = ny;
    a.g.out_f = (long long)m * ny;
    a.g.out_e = 1;
    a.g.nf = nx;
    a.g.blocks = m;
    a.nyq = static_cast<double2*>(p.spec_nyq);
    a.nq_p = ny;
    a.nq_e = 1;
    a.nq_pencils = nx;
    const hipError_t e = run_pass<R>(a, PK_Y_INV, s, 0);
    if (e != hipSuccess) return e;
  }
  CArgs<R> c;
  c.spec = cx<R>(p.spec);
  c.nyq = static_cast<const double2*>(p.spec_nyq);
  c.out = static_cast<R*>(out);
  c.nx = nx;
  c.ny = ny;
  c.m = m;
  c.ax = p.ax_zc;
  c.tw2 = p.tw_r2c64;
  if (z_reg_rows(m) > 0) {
    c.P = z_reg_rows(m);
    c.tiles_y = (ny + c.P - 1) / c.P;
    const dim3 g((unsigned)((long long)nx * c.tiles_y)), b(kFftThreads);
    switch (m) {
      case 32: hipLaunchKernelGGL((k_c2r_z_reg<R, 8, 4>), g, b, 0, s, c); break;
      case 64: hipLaunchKernelGGL((k_c2r_z_reg<R, 8, 8>), g, b, 0, s, c); break;
      case 128: hipLaunchKernelGGL((k_c2r_z_reg<R, 16, 8>), g, b, 0, s, c); break;
      default: hipLaunchKernelGGL((k_c2r_z_reg<R, 16, 16>), g, b, 0, s, c); break;
    }
    return hipGetLastError();
  }
  c.P = lds_pencils(m);
  while (c.P > 1 && lds_bytes(m, c.P) > max_lds_per_block()) c.P /= 2;
  c.tiles_y = (ny + c.P - 1) / c.P;
  hipLaunchKernelGGL((k_c2r_z<R>), dim3((unsigned)((long long)nx * c.tiles_y)), dim3(kFftThreads),
                     lds_bytes(m, c.P), s, c);
  return hipGetLastError();
}

// stored spectrum (rho[kz][ky][kx] complex, nyq[x][y] FP64) -> complex128 (Nx, Ny, Nz/2+1)
template <typename R>
__global__ void k_spec_read(const typename V2<R>::T* __restrict__ main,
                            const double2* __restrict__ nyq, double2* __restrict__ out, int nx,
                            int ny, int m) {
  const long long h = m + 1;
  const long long total = (long long)nx * ny * h;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long row = i / h;
    const int kz = (int)(i - row * h);
    const int y = (int)(row % ny), x = (int)(row / ny);
    if (kz < m) {
      const auto v = main[((long long)kz * ny + y) * nx + x];
      out[i] = make_double2((double)v.x, (double)v.y);
    } else {
      out[i] = nyq[row];
    }
  }
}

}  // namespace

bool fft_factor(int n, int* radix, int* nstages) {
  int k = 0;
  int m = n;
  static const int pref[] = {16, 8, 4, 2, 3, 5};
  for (int r : pref) {
    while (m % r == 0 && m > 1) {
      if (k >= kMaxStages) return false;
      radix[k++] = r;
      m /= r;
    }
  }
  for (int r = 7; m > 1; r += 2) {  // remaining odd primes: generic stage
    while (m % r == 0) {
      if (k >= kMaxStages) return false;
      radix[k++] = r;
      m /= r;
    }
  }
  *nstages = k;
  return true;
}

hipError_t fft_make_twiddles(float2* tw, double2* tw64, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_twiddles, dim3((n + 255) / 256), dim3(256), 0, s, tw, tw64, n);
  return hipGetLastError();
}

hipError_t fft_forward(Plan& p, const void* density, bool accumulate, hipStream_t s) {
  return p.fft64 ? forward_impl<double>(p, density, accumulate, s)
                 : forward_impl<float>(p, density, accumulate, s);
}

namespace {

struct TuneKey {
  int n[3];
  int fft64;
  bool operator<(const TuneKey& o) const {
    for (int d = 0; d < 3; ++d)
      if (n[d] != o.n[d]) return n[d] < o.n[d];
    return fft64 < o.fft64;
  }
};
struct TuneVal {
  int z, y, x;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune_cache;

// Best of 3 timed launches (after one warm-up) of `run` on stream s, in ms.
template <typename F>
float time_best(F&& run, hipStream_t s, hipEvent_t e0, hipEvent_t e1, hipError_t* err) {
  float best = 1e30f;
  for (int it = 0; it < 4; ++it) {
    (void)hipEventRecord(e0, s);
    const hipError_t e = run();
    (void)hipEventRecord(e1, s);
    if (e != hipSuccess) {
      *err = e;
      return best;
    }
    if ((*err = hipEventSynchronize(e1)) != hipSuccess) return best;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (it > 0 && ms < best) best = ms;
  }
  return best;
}

template <typename R>
hipError_t autotune_impl(Plan& p, hipStream_t s) {
  const int m = p.sys.n[2] / 2;
  hipEvent_t e0, e1;
  hipError_t err = hipEventCreate(&e0);
  if (err != hipSuccess) return err;
  if ((err = hipEventCreate(&e1)) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return err;
  }
  const int cand[5] = {kTuneReg, 2, 4, 8, 16};
  // z pass: the register-staged kernel (m with a two-stage split) or the
  // LDS-staged pass with that many rows
  float best = 1e30f;
  int pick = 0;
  for (int c : cand) {
    if (c == kTuneReg && !z_stream_ok(m)) continue;
    if (c != kTuneReg && lds_bytes(m, c) > max_lds_per_block()) continue;
    p.tune_z = c;
    const float t = time_best([&] { return pass_z<R>(p, p.rho, s); }, s, e0, e1, &err);
    if (err != hipSuccess) break;
    if (t < best) {
      best = t;
      pick = c;
    }
  }
  p.tune_z = pick;
  // y and x passes: the register-staged pass or the LDS-staged one
  for (int axis = 0; axis < 2 && err == hipSuccess; ++axis) {
    int& slot = axis == 0 ? p.tune_y : p.tune_x;
    const int n = p.sys.n[axis == 0 ? 1 : 0];
    int r1, r2, r3;
    best = 1e30f;
    pick = 0;
    for (int c : cand) {
      if (c == kTuneReg && !reg_split(n, &r1, &r2, &r3)) continue;
      if (c != kTuneReg && lds_bytes(n, c) > max_lds_per_block()) continue;
      slot = c;
      const float t = time_best(
          [&] { return axis == 0 ? pass_y<R>(p, s) : pass_x<R>(p, true, s); }, s, e0, e1, &err);
      if (err != hipSuccess) break;
      if (t < best) {
        best = t;
        pick = c;
      }
    }
    slot = pick;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return err;
}

}  // namespace

// Picks the z tile height and the y / x pass kernels for this grid by timing
// the candidates on the plan's own buffers (results cached per grid shape and
// precision).  Clobbers the density, spectrum and S(q) accumulators.
hipError_t fft_autotune(Plan& p, hipStream_t s) {
  const TuneKey key{{p.sys.n[0], p.sys.n[1], p.sys.n[2]}, p.fft64};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      p.tune_z = it->second.z;
      p.tune_y = it->second.y;
      p.tune_x = it->second.x;
      return hipSuccess;
    }
  }
  const hipError_t e = p.fft64 ? autotune_impl<double>(p, s) : autotune_impl<float>(p, s);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = TuneVal{p.tune_z, p.tune_y, p.tune_x};
  return hipSuccess;
}

hipError_t fft_inverse(Plan& p, const double* w_main, const double* w_nyq, void* out,
                       hipStream_t s) {
  return p.fft64 ? inverse_impl<double>(p, w_main, w_nyq, out, s)
                 : inverse_impl<float>(p, w_main, w_nyq, out, s);
}

void fft_pass_events(hipEvent_t begin, hipEvent_t end) {
  t_pass_ev[0] = begin;
  t_pass_ev[1] = end;
  t_pass_ev_used = false;
}

bool fft_pass_events_used() { return t_pass_ev_used; }

bool fft_x_pairs(const Plan& p) {
  int r1, r2, r3;
  return (p.tune_x == 0 || p.tune_x == kTuneReg) && reg_split(p.sys.n[0], &r1, &r2, &r3);
}

hipError_t fft_x_pair(Plan& p, int slot_a, int slot_b, hipStream_t s) {
  return p.fft64 ? pass_x_pair<double>(p, slot_a, slot_b, s)
                 : pass_x_pair<float>(p, slot_a, slot_b, s);
}

hipError_t fft_pass(Plan& p, int pass, const void* density, hipStream_t s) {
  switch (pass) {
    case 0: return p.fft64 ? pass_z<double>(p, density, s) : pass_z<float>(p, density, s);
    case 1: return p.fft64 ? pass_y<double>(p, s) : pass_y<float>(p, s);
    default: return p.fft64 ? pass_x<double>(p, true, s) : pass_x<float>(p, true, s);
  }
}

hipError_t fft_spectrum_read(Plan& p, double* out, hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  long long total = (long long)nx * ny * (m + 1);
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (p.fft64)
    hipLaunchKernelGGL(k_spec_read<double>, dim3((unsigned)blocks), dim3(256), 0, s,
                       cx<double>(p.rho), static_cast<const double2*>(p.spec_nyq),
                       reinterpret_cast<double2*>(out), nx, ny, m);
  else
    hipLaunchKernelGGL(k_spec_read<float>, dim3((unsigned)blocks), dim3(256), 0, s,
                       cx<float>(p.rho), static_cast<const double2*>(p.spec_nyq),
                       reinterpret_cast<double2*>(out), nx, ny, m);
  return hipGetLastError();
}

size_t fft_spec_bytes(const int n[3], int fft64) {
  return (size_t)n[0] * n[1] * (n[2] / 2) * (fft64 ? 16 : 8);
}

size_t fft_res_bytes(const int n[3], int fft64) {
  return fft64 ? 0 : (size_t)n[0] * n[1] * (n[2] / 2);
}

// --- plan-free transform for callers that cannot allocate -----------------------
// The working set of mdsf_rfft3_noalloc (twiddles, the z and y pass outputs,
// the y pass residuals, Nyquist plane) lives in this fixed device array, one copy per GPU, held until the
// process ends.  Host code never passes the array's name to a kernel: it is a
// host-side stand-in.  The device address of each GPU's copy comes from
// hipGetSymbolAddress on the first call on that device and is kept per device.
namespace {

constexpr size_t kScratchBytes = (size_t)144 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array (resolved once per device).
hipError_t scratch_address(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

}  // namespace mdsf

extern "C" int mdsf_rfft3_noalloc(const int32_t n[3], int precision, const void* d_density,
                                  double* d_spectrum, mdsf_stream stream) {
  using namespace mdsf;
  if (!n || !d_density || !d_spectrum) return fail(MDSF_ERR_ARG, "null argument");
  if (precision != MDSF_F32 && precision != MDSF_F64)
    return fail(MDSF_ERR_ARG, "precision must be MDSF_F32 or MDSF_F64");
  for (int d = 0; d < 3; ++d)
    if (n[d] < 2 || n[d] % 2 != 0 || n[d] > 4096)
      return fail(MDSF_ERR_ARG, "grid sizes must be even and in [2, 4096]");
  const int fft64 = precision == MDSF_F64 ? 1 : 0;
  const int nn[3] = {n[0], n[1], n[2]};
  const size_t nx = n[0], ny = n[1], nz = n[2], m = nz / 2;
  // layout: the spectrum buffer (the y pass output, transformed in place by
  // the x pass), the z pass output, the y pass residuals (FP32), the Nyquist
  // plane (FP64 complex), then the FP64 twiddles of the x, y, z/2 axes and
  // of the R2C split (length nz)
  const size_t lens[4] = {nx, ny, m, nz};
  size_t off = 0;
  const size_t off_rho = off;
  off = align256(off + fft_spec_bytes(nn, fft64));
  const size_t off_spec = off;
  off = align256(off + fft_spec_bytes(nn, fft64));
  const size_t off_res = off;
  off = align256(off + fft_res_bytes(nn, fft64));
  const size_t off_nyq = off;
  off = align256(off + sizeof(double2) * nx * ny);
  size_t off_tw64[4];
  for (int k = 0; k < 4; ++k) {
    off_tw64[k] = off;
    off = align256(off + sizeof(double2) * lens[k]);
  }
  if (off > kScratchBytes)
    return fail(MDSF_ERR_ARG, "grid too large for mdsf_rfft3_noalloc: it needs " +
                                  std::to_string(off) + " bytes of its " +
                                  std::to_string(kScratchBytes) +
                                  "-byte device scratch array (use a plan and mdsf_rfft3)");
  unsigned char* base = nullptr;
  MDSF_HIP(scratch_address(&base));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  Plan p;  // host-side description only: every device pointer is into `base`
  for (int d = 0; d < 3; ++d) p.sys.n[d] = n[d];
  p.fft64 = fft64;
  p.rho = base + off_rho;
  p.spec = base + off_spec;
  p.res = fft_res_bytes(nn, fft64) > 0 ? base + off_res : nullptr;
  p.spec_nyq = base + off_nyq;
  FftAxis* axes[3] = {&p.ax_x, &p.ax_y, &p.ax_zc};
  for (int k = 0; k < 4; ++k) {
    double2* tw64 = reinterpret_cast<double2*>(base + off_tw64[k]);
    if (k < 3) {
      FftAxis& ax = *axes[k];
      ax.n = (int)lens[k];
      if (!fft_factor(ax.n, ax.radix, &ax.nstages))
        return fail(MDSF_ERR_UNSUPPORTED,
                    "FFT length " + std::to_string(ax.n) + " needs too many stages");
      ax.tw = nullptr;
      ax.tw64 = tw64;
    } else {
      p.tw_r2c = nullptr;
      p.tw_r2c64 = tw64;
    }
    MDSF_HIP(fft_make_twiddles(nullptr, tw64, (int)lens[k], s));
  }
  MDSF_HIP(fft_forward(p, d_density, false, s));
  MDSF_HIP(fft_spectrum_read(p, d_spectrum, s));
  return MDSF_OK;
}
