// 3-D real-to-complex FFT for gfx950: Stockham autosort stages in LDS.
//
// Replaces np.fft.rfftn (dens.py:313) and the |F|^2 frame sum
// (dens.py:315-318).  Unnormalised forward transform, exponent -2*pi*i*k*n/N.
// Element type R = float (FP32 storage and arithmetic) or double (FP64); the
// twiddles of both are generated in FP64 on the device.
//
// Layout between passes (internal): the z pass packs each real row of Nz
// floats as Nz/2 complex values and writes the half spectrum,
// kz in [0, Nz/2) -> main[x][y][kz], and the kz = Nz/2 plane -> nyq[x][y].
// The y and x passes are strided pencil transforms over those arrays; the
// x pass can fold |F|^2 into FP64 accumulators of the same layout.
//
// Each workgroup owns a tile of P pencils.  LDS holds re/im planes of
// P x pitch values (pitch odd) twice (ping-pong between stages), and threads
// are mapped pencil-fastest, so every Stockham read and write is
// conflict-free for any radix and span.
#include <map>
#include <mutex>

#include "mdsf_internal.h"

namespace mdsf {

namespace {

constexpr int kFftThreads = 256;

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };

template <typename R>
struct Cx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ Cx<R> mk(R a, R b) {
  Cx<R> c;
  c.x = a;
  c.y = b;
  return c;
}
template <typename R>
__device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b) {
  return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R>
__device__ __forceinline__ Cx<R> cmulc(Cx<R> a, Cx<R> b) {  // a * conj(b)
  return mk<R>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename R>
__device__ __forceinline__ Cx<R> cadd(Cx<R> a, Cx<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <typename R>
__device__ __forceinline__ Cx<R> csub(Cx<R> a, Cx<R> b) { return mk<R>(a.x - b.x, a.y - b.y); }
template <typename R>
__device__ __forceinline__ Cx<R> cscale(R s, Cx<R> a) { return mk<R>(s * a.x, s * a.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename R>
__device__ __forceinline__ Cx<R> mul_mi(Cx<R> a) {
  return INV ? mk<R>(-a.y, a.x) : mk<R>(a.y, -a.x);
}
template <typename R>
__device__ __forceinline__ Cx<R> ld_tw(const void* tw, int k) {
  const auto v = static_cast<const typename V2<R>::T*>(tw)[k];
  return mk<R>(v.x, v.y);
}

// --- small DFT kernels (in registers) ---------------------------------------
template <bool INV, typename R>
__device__ __forceinline__ void dft2(Cx<R>* v) {
  const Cx<R> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <bool INV, typename R>
__device__ __forceinline__ void dft3(Cx<R>* v) {
  const R s = (R)0.86602540378443864676;  // sin(2 pi / 3)
  const Cx<R> t1 = cadd(v[1], v[2]);
  const Cx<R> t2 = mk<R>(v[0].x - (R)0.5 * t1.x, v[0].y - (R)0.5 * t1.y);
  const Cx<R> t3 = mul_mi<INV>(cscale(s, csub(v[1], v[2])));
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

template <bool INV, typename R>
__device__ __forceinline__ void dft4(Cx<R>* v) {
  const Cx<R> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  const Cx<R> t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <bool INV, typename R>
__device__ __forceinline__ void dft5(Cx<R>* v) {
  const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
  const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
  const Cx<R> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const Cx<R> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const Cx<R> x0 = v[0];
  const Cx<R> a1 = mk<R>(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
  const Cx<R> a2 = mk<R>(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
  const Cx<R> b1 = mul_mi<INV>(mk<R>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  const Cx<R> b2 = mul_mi<INV>(mk<R>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  v[0] = mk<R>(x0.x + t1.x + t2.x, x0.y + t1.y + t2.y);
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

template <bool INV, typename R>
__device__ __forceinline__ void dft8(Cx<R>* v) {
  const R h = (R)0.70710678118654752440;
  Cx<R> e[4] = {v[0], v[2], v[4], v[6]};
  Cx<R> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(e);
  dft4<INV>(o);
  const Cx<R> w1 = mk<R>(h, INV ? h : -h);
  const Cx<R> w3 = mk<R>(-h, INV ? h : -h);
  const Cx<R> o1 = cmul(o[1], w1);
  const Cx<R> o2 = mul_mi<INV>(o[2]);
  const Cx<R> o3 = cmul(o[3], w3);
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <bool INV, typename R>
__device__ __forceinline__ void dft16(Cx<R>* v) {
  Cx<R> e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    e[i] = v[2 * i];
    o[i] = v[2 * i + 1];
  }
  dft8<INV>(e);
  dft8<INV>(o);
  const R c1 = (R)0.92387953251128675613, s1 = (R)0.38268343236508977173;
  const R h = (R)0.70710678118654752440;
  const R sg = INV ? (R)1 : (R)-1;
  const Cx<R> w[8] = {mk<R>(1, 0),        mk<R>(c1, sg * s1),  mk<R>(h, sg * h),
                      mk<R>(s1, sg * c1), mk<R>(0, sg),        mk<R>(-s1, sg * c1),
                      mk<R>(-h, sg * h),  mk<R>(-c1, sg * s1)};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const Cx<R> t = k == 0 ? o[0] : cmul(o[k], w[k]);
    v[k] = cadd(e[k], t);
    v[k + 8] = csub(e[k], t);
  }
}

template <int RAD, bool INV, typename R>
__device__ __forceinline__ void dft(Cx<R>* v) {
  if constexpr (RAD == 2) dft2<INV>(v);
  else if constexpr (RAD == 3) dft3<INV>(v);
  else if constexpr (RAD == 4) dft4<INV>(v);
  else if constexpr (RAD == 5) dft5<INV>(v);
  else if constexpr (RAD == 8) dft8<INV>(v);
  else dft16<INV>(v);
}

// One Stockham stage, out of place (in -> out planes), one butterfly per
// thread iteration: X[(j/ns)*ns*RAD + k + r*ns] = DFT(x[j + r*n/RAD] * W^(k r n/(ns RAD))).
template <int RAD, bool INV, typename R>
__device__ void stage_fixed(const R* __restrict__ ire, const R* __restrict__ iim,
                            R* __restrict__ ore, R* __restrict__ oim, int n, int P, int pitch,
                            int ns, const void* __restrict__ tw) {
  const int nq = n / RAD;
  const int nbf = P * nq;
  const int twstep = n / (ns * RAD);
  for (int b = threadIdx.x; b < nbf; b += kFftThreads) {
    const int p = b % P, j = b / P;
    const int k = j % ns;
    const int row = p * pitch;
    Cx<R> v[RAD];
#pragma unroll
    for (int r = 0; r < RAD; ++r) {
      const int a = row + j + r * nq;
      Cx<R> x = mk<R>(ire[a], iim[a]);
      if (r > 0 && k > 0) {
        const Cx<R> w = ld_tw<R>(tw, k * r * twstep);
        x = INV ? cmulc(x, w) : cmul(x, w);
      }
      v[r] = x;
    }
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
template <bool INV, typename R>
__device__ void stage_generic(const R* __restrict__ ire, const R* __restrict__ iim,
                              R* __restrict__ ore, R* __restrict__ oim, int n, int P, int pitch,
                              int ns, int RAD, const void* __restrict__ tw) {
  const int nq = n / RAD;
  const int twstep = n / (ns * RAD);
  const int step = n / RAD;
  for (int o = threadIdx.x; o < P * n; o += kFftThreads) {
    const int p = o % P, rest = o / P;
    const int j = rest % nq, m = rest / nq;
    const int k = j % ns;
    const int row = p * pitch;
    Cx<R> acc = mk<R>(0, 0);
    int wm = 0;  // (r * m) mod RAD
    for (int r = 0; r < RAD; ++r) {
      const int a = row + j + r * nq;
      Cx<R> x = mk<R>(ire[a], iim[a]);
      if (r > 0 && k > 0) {
        const Cx<R> w = ld_tw<R>(tw, k * r * twstep);
        x = INV ? cmulc(x, w) : cmul(x, w);
      }
      const Cx<R> w = ld_tw<R>(tw, wm * step);
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

// Runs every stage of `ax`, ping-ponging between plane pairs 0 and 1 of
// `lds` ([re0 | im0 | re1 | im1], P*pitch values each).  Returns the plane
// pair holding the result.
template <bool INV, typename R>
__device__ int run_stages(R* lds, int n, int P, int pitch, const FftAxis& ax) {
  const int plane = P * pitch;
  const void* tw = sizeof(R) == 8 ? (const void*)ax.tw64 : (const void*)ax.tw;
  int ns = 1, cur = 0;
  for (int s = 0; s < ax.nstages; ++s) {
    const int RAD = ax.radix[s];
    const R* ire = lds + (2 * cur) * plane;
    const R* iim = ire + plane;
    R* ore = lds + (2 * (cur ^ 1)) * plane;
    R* oim = ore + plane;
    switch (RAD) {
      case 2: stage_fixed<2, INV>(ire, iim, ore, oim, n, P, pitch, ns, tw); break;
      case 3: stage_fixed<3, INV>(ire, iim, ore, oim, n, P, pitch, ns, tw); break;
      case 4: stage_fixed<4, INV>(ire, iim, ore, oim, n, P, pitch, ns, tw); break;
      case 5: stage_fixed<5, INV>(ire, iim, ore, oim, n, P, pitch, ns, tw); break;
      case 8: stage_fixed<8, INV>(ire, iim, ore, oim, n, P, pitch, ns, tw); break;
      case 16: stage_fixed<16, INV>(ire, iim, ore, oim, n, P, pitch, ns, tw); break;
      default: stage_generic<INV>(ire, iim, ore, oim, n, P, pitch, ns, RAD, tw); break;
    }
    ns *= RAD;
    cur ^= 1;
  }
  return cur;
}

__host__ __device__ inline int lds_pitch(int n) { return (n % 2 == 0) ? n + 1 : n; }

// --- twiddles: W_n^k = exp(-2 pi i k / n), generated in FP64 ------------------
__global__ void k_twiddles(float2* __restrict__ tw, double2* __restrict__ tw64, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double s, c;
  sincospi(-2.0 * (double)k / (double)n, &s, &c);
  tw[k] = make_float2((float)c, (float)s);
  tw64[k] = make_double2(c, s);
}

// --- z pass: real rows -> half spectrum + Nyquist plane ------------------------
template <typename R>
struct R2CArgs {
  const R* in;                      // rows of nz reals (plan precision)
  typename V2<R>::T* main;          // rows of m = nz/2 complex (may alias `in` for R = float)
  typename V2<R>::T* nyq;           // one complex per row
  long long rows;
  int m;                            // nz / 2
  int P;                            // rows per tile
  FftAxis ax;                       // length m
  const void* tw2;                  // W_nz^k, k < nz
};

template <typename R>
__global__ __launch_bounds__(kFftThreads) void k_r2c_z(R2CArgs<R> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  R* lds = reinterpret_cast<R*>(lds_raw);
  const int m = a.m, P = a.P, pitch = lds_pitch(m);
  R* re = lds;
  R* im = lds + P * pitch;
  const long long r0 = (long long)blockIdx.x * P;
  const int pc = (int)min((long long)P, a.rows - r0);
  const auto* src = reinterpret_cast<const typename V2<R>::T*>(a.in) + r0 * m;
  for (int idx = threadIdx.x; idx < pc * m; idx += kFftThreads) {
    const int p = idx / m, e = idx - p * m;
    const auto v = src[idx];
    re[p * pitch + e] = (R)v.x;
    im[p * pitch + e] = (R)v.y;
  }
  __syncthreads();
  {
    const int cur = run_stages<false, R>(lds, m, P, pitch, a.ax);
    re = lds + 2 * cur * P * pitch;
    im = re + P * pitch;
  }
  // split: X_k = A_k + W_nz^k B_k, A = (Z_k + conj Z_{m-k})/2, B = (Z_k - conj Z_{m-k})/(2i)
  auto* dst = a.main + r0 * m;
  for (int idx = threadIdx.x; idx < pc * m; idx += kFftThreads) {
    const int p = idx / m, k = idx - p * m;
    const int kk = (k == 0) ? 0 : m - k;
    const Cx<R> z = mk<R>(re[p * pitch + k], im[p * pitch + k]);
    const Cx<R> zc = mk<R>(re[p * pitch + kk], -im[p * pitch + kk]);
    const Cx<R> A = cscale((R)0.5, cadd(z, zc));
    const Cx<R> d = cscale((R)0.5, csub(z, zc));
    const Cx<R> B = mk<R>(d.y, -d.x);  // d / i
    const Cx<R> X = cadd(A, cmul(B, ld_tw<R>(a.tw2, k)));
    dst[idx].x = X.x;
    dst[idx].y = X.y;
  }
  for (int p = threadIdx.x; p < pc; p += kFftThreads) {
    // X_m = A_0 - B_0 = Re z0 - Im z0 (real)
    a.nyq[r0 + p].x = re[p * pitch] - im[p * pitch];
    a.nyq[r0 + p].y = (R)0;
  }
}

// Rows per block of the register-staged z passes: half the threads own one
// stage-1 butterfly (m / R1 per row).
constexpr int z_rows_per_block(int m, int r1) { return kFftThreads / (m / r1) / 2; }

// Register-staged z pass (FP32, m = R1 * R2, two stages).  Stage 1 reads its
// R1 inputs per butterfly straight from global memory (row-contiguous across
// lanes) into registers; stage 2 runs in place in LDS (each thread reads and
// writes the same R2 slots); the split into the half spectrum then reads the
// pairs (k, m - k) from LDS and stores whole rows contiguously.  A block holds
// P rows (half the threads take a stage-1 butterfly) in one plane pair of
// LDS (16.5 KB at m = 128): small enough to fit beside the deposition's
// blocks when the frame pipeline runs both, and faster alone than 2P rows.
template <typename R, int R1, int R2>
__global__ __launch_bounds__(kFftThreads) void k_r2c_z_reg(R2CArgs<R> a) {
  constexpr int m = R1 * R2;
  constexpr int nq1 = m / R1;             // stage-1 butterflies per row
  constexpr int P = z_rows_per_block(m, R1);
  constexpr int pitch = m + 1;
  constexpr int ns = m / R2;              // stage-2 butterflies per row
  __shared__ R re[P * pitch];
  __shared__ R im[P * pitch];
  const long long r0 = (long long)blockIdx.x * P;
  const int pc = (int)min((long long)P, a.rows - r0);
  const auto* src = reinterpret_cast<const typename V2<R>::T*>(a.in) + r0 * m;
  const void* tw = sizeof(R) == 8 ? (const void*)a.ax.tw64 : (const void*)a.ax.tw;
  if (threadIdx.x < P * nq1) {
    const int j = threadIdx.x % nq1, p = threadIdx.x / nq1;
    Cx<R> v[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      if (p < pc) {
        const auto x = src[p * m + j + r * nq1];
        v[r] = mk<R>(x.x, x.y);
      } else {
        v[r] = mk<R>(0, 0);
      }
    }
    dft<R1, false>(v);
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      re[p * pitch + j * R1 + r] = v[r].x;
      im[p * pitch + j * R1 + r] = v[r].y;
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < P * ns; b += kFftThreads) {
    const int k = b % ns, p = b / ns;
    Cx<R> v[R2];
#pragma unroll
    for (int r = 0; r < R2; ++r) v[r] = mk<R>(re[p * pitch + k + r * ns], im[p * pitch + k + r * ns]);
    twiddle_in<R2, false>(v, k, 1, tw);
    dft<R2, false>(v);
#pragma unroll
    for (int r = 0; r < R2; ++r) {
      re[p * pitch + k + r * ns] = v[r].x;
      im[p * pitch + k + r * ns] = v[r].y;
    }
  }
  __syncthreads();
  // split: X_k = A_k + W_nz^k B_k, A = (Z_k + conj Z_{m-k})/2, B = (Z_k - conj Z_{m-k})/(2i)
  auto* dst = a.main + r0 * m;
  for (int idx = threadIdx.x; idx < pc * m; idx += kFftThreads) {
    const int p = idx / m, k = idx - p * m;
    const int kk = (k == 0) ? 0 : m - k;
    const Cx<R> z = mk<R>(re[p * pitch + k], im[p * pitch + k]);
    const Cx<R> zc = mk<R>(re[p * pitch + kk], -im[p * pitch + kk]);
    const Cx<R> A = cscale((R)0.5, cadd(z, zc));
    const Cx<R> d = cscale((R)0.5, csub(z, zc));
    const Cx<R> B = mk<R>(d.y, -d.x);  // d / i
    const Cx<R> X = cadd(A, cmul(B, ld_tw<R>(a.tw2, k)));
    dst[idx].x = X.x;
    dst[idx].y = X.y;
  }
  for (int p = threadIdx.x; p < pc; p += kFftThreads) {
    // X_m = A_0 - B_0 = Re z0 - Im z0 (real)
    a.nyq[r0 + p].x = re[p * pitch] - im[p * pitch];
    a.nyq[r0 + p].y = (R)0;
  }
}

// --- adjoint z pass: half spectrum -> real rows ----------------------------------
// Real adjoint of the R2C along z:  x_n = Re sum_{k=0..m} X_k e^{+2 pi i k n / nz}.
// With X'_0 = Re X_0, X'_m = Re X_m and X'_k = X_k / 2 inside, this is the
// Hermitian inverse of X': Z_k = (X'_k + conj X'_{m-k}) + i W^{-k} (X'_k - conj X'_{m-k})
// and x_{2j} + i x_{2j+1} = sum_k Z_k e^{+2 pi i jk/m}.
template <typename R, typename OUT>
struct C2RArgs {
  const typename V2<R>::T* main;
  const typename V2<R>::T* nyq;
  OUT* out;                         // rows of nz reals
  long long rows;
  int m;
  int P;
  FftAxis ax;
  const void* tw2;
};

template <typename R, typename OUT>
__global__ __launch_bounds__(kFftThreads) void k_c2r_z(C2RArgs<R, OUT> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  R* lds = reinterpret_cast<R*>(lds_raw);
  const int m = a.m, P = a.P, pitch = lds_pitch(m);
  R* re = lds;
  R* im = lds + P * pitch;
  const long long r0 = (long long)blockIdx.x * P;
  const int pc = (int)min((long long)P, a.rows - r0);
  const auto* src = a.main + r0 * m;
  for (int idx = threadIdx.x; idx < pc * m; idx += kFftThreads) {
    const int p = idx / m, k = idx - p * m;
    Cx<R> x, xc;
    if (k == 0) {
      x = mk<R>(src[idx].x, (R)0);
      xc = mk<R>(a.nyq[r0 + p].x, (R)0);
    } else {
      const auto v = src[p * m + (m - k)];
      x = mk<R>((R)0.5 * src[idx].x, (R)0.5 * src[idx].y);
      xc = mk<R>((R)0.5 * v.x, (R)-0.5 * v.y);
    }
    const Cx<R> e = cadd(x, xc);
    const Cx<R> o = cmulc(csub(x, xc), ld_tw<R>(a.tw2, k));  // (X - conj X') W^{-k}
    re[p * pitch + k] = e.x - o.y;                            // e + i o
    im[p * pitch + k] = e.y + o.x;
  }
  __syncthreads();
  {
    const int cur = run_stages<true, R>(lds, m, P, pitch, a.ax);
    re = lds + 2 * cur * P * pitch;
    im = re + P * pitch;
  }
  OUT* dst = a.out + r0 * (2 * (long long)m);
  for (int idx = threadIdx.x; idx < pc * m; idx += kFftThreads) {
    const int p = idx / m, j = idx - p * m;
    dst[2 * (long long)idx] = (OUT)re[p * pitch + j];
    dst[2 * (long long)idx + 1] = (OUT)im[p * pitch + j];
  }
}

// Register-staged adjoint z pass (FP32, m = R1 * R2): stage 1 forms Z_k from
// the pair (X_k, X_{m-k}) read straight from global memory, the second stage
// runs from LDS into registers and stores the real pairs (x_2j, x_2j+1)
// straight to global memory.  A block reads all of its rows before the
// barrier, so `out` may alias the spectrum.
template <typename R, int R1, int R2>
__global__ __launch_bounds__(kFftThreads) void k_c2r_z_reg(C2RArgs<R, R> a) {
  constexpr int m = R1 * R2;
  constexpr int nq1 = m / R1;
  constexpr int P = z_rows_per_block(m, R1);
  constexpr int pitch = m + 1;
  constexpr int ns = m / R2;
  __shared__ R re[P * pitch];
  __shared__ R im[P * pitch];
  const long long r0 = (long long)blockIdx.x * P;
  const int pc = (int)min((long long)P, a.rows - r0);
  const auto* src = a.main + r0 * m;
  const void* tw = sizeof(R) == 8 ? (const void*)a.ax.tw64 : (const void*)a.ax.tw;
  if (threadIdx.x < P * nq1) {
    const int j = threadIdx.x % nq1, p = threadIdx.x / nq1;
    Cx<R> v[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      const int k = j + r * nq1;
      if (p < pc) {
        Cx<R> x, xc;
        if (k == 0) {
          x = mk<R>(src[p * m].x, (R)0);
          xc = mk<R>(a.nyq[r0 + p].x, (R)0);
        } else {
          const auto u = src[p * m + k];
          const auto w = src[p * m + (m - k)];
          x = mk<R>((R)0.5 * u.x, (R)0.5 * u.y);
          xc = mk<R>((R)0.5 * w.x, (R)-0.5 * w.y);
        }
        const Cx<R> e = cadd(x, xc);
        const Cx<R> o = cmulc(csub(x, xc), ld_tw<R>(a.tw2, k));  // (X - conj X') W^{-k}
        v[r] = mk<R>(e.x - o.y, e.y + o.x);                       // e + i o
      } else {
        v[r] = mk<R>(0, 0);
      }
    }
    dft<R1, true>(v);
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      re[p * pitch + j * R1 + r] = v[r].x;
      im[p * pitch + j * R1 + r] = v[r].y;
    }
  }
  __syncthreads();
  R* dst = a.out + r0 * (2 * (long long)m);
  for (int b = threadIdx.x; b < P * ns; b += kFftThreads) {
    const int k = b % ns, p = b / ns;
    Cx<R> v[R2];
#pragma unroll
    for (int r = 0; r < R2; ++r) v[r] = mk<R>(re[p * pitch + k + r * ns], im[p * pitch + k + r * ns]);
    twiddle_in<R2, true>(v, k, 1, tw);
    dft<R2, true>(v);
    if (p < pc) {
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const long long o = 2 * ((long long)p * m + k + r * ns);
        dst[o] = v[r].x;
        dst[o + 1] = v[r].y;
      }
    }
  }
}

// --- strided pencil pass ------------------------------------------------------
// One launch transforms two jobs: the main half spectrum and the Nyquist
// plane (workgroups [0, blocks0) take job 0, the rest job 1).
template <typename R>
struct PencilJob {
  typename V2<R>::T* data;
  double* acc;              // EPI_ACC: acc += |X|^2 (FP64) at the same index
  const double* scale;      // if set: X *= scale[index] before the transform
  long long estride;        // between elements of a pencil
  long long pstride;        // between adjacent pencils of a tile
  long long outer_stride;   // between outer groups
  int pencils;              // pencils per outer group
  int tiles_inner;          // ceil(pencils / P)
  int P;
  int pencil_major;         // elements contiguous (estride == 1)
};

template <typename R>
struct PencilArgs {
  PencilJob<R> job[2];
  long long blocks0;
  int n;
  FftAxis ax;
};

enum Epilogue { EPI_STORE = 0, EPI_ACC = 1 };

template <typename R>
int tile_pencils(int n);

// Generic tile: load P pencils into LDS, all stages in LDS, store / accumulate.
template <bool INV, typename R, int EPI>
__device__ void generic_tile(const PencilJob<R>& J, long long tile, int n, const FftAxis& ax,
                             R* lds) {
  const int P = J.P, pitch = lds_pitch(n);
  R* re = lds;
  R* im = lds + P * pitch;
  const long long outer = tile / J.tiles_inner;
  const int inner = (int)(tile - outer * J.tiles_inner);
  const int p0 = inner * P;
  const int pc = min(P, J.pencils - p0);
  const long long base = outer * J.outer_stride + (long long)p0 * J.pstride;
  const int total = pc * n;
  for (int idx = threadIdx.x; idx < total; idx += kFftThreads) {
    int p, e;
    if (J.pencil_major) {
      p = idx / n;
      e = idx - p * n;
    } else {
      e = idx / pc;
      p = idx - e * pc;
    }
    const long long g = base + p * J.pstride + e * J.estride;
    const auto v = J.data[g];
    R sc = (R)1;
    if (J.scale) sc = (R)J.scale[g];
    re[p * pitch + e] = v.x * sc;
    im[p * pitch + e] = v.y * sc;
  }
  __syncthreads();
  {
    const int cur = run_stages<INV, R>(lds, n, P, pitch, ax);
    re = lds + 2 * cur * P * pitch;
    im = re + P * pitch;
  }
  for (int idx = threadIdx.x; idx < total; idx += kFftThreads) {
    int p, e;
    if (J.pencil_major) {
      p = idx / n;
      e = idx - p * n;
    } else {
      e = idx / pc;
      p = idx - e * pc;
    }
    const R xr = re[p * pitch + e], xi = im[p * pitch + e];
    const long long g = base + p * J.pstride + e * J.estride;
    if (EPI == EPI_STORE) {
      J.data[g].x = xr;
      J.data[g].y = xi;
    } else {
      J.acc[g] += (double)xr * (double)xr + (double)xi * (double)xi;
    }
  }
}

template <bool INV, typename R, int EPI>
__global__ __launch_bounds__(kFftThreads) void k_pencils(PencilArgs<R> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  R* lds = reinterpret_cast<R*>(lds_raw);
  const bool second = (long long)blockIdx.x >= a.blocks0;
  const long long tile = second ? (long long)blockIdx.x - a.blocks0 : (long long)blockIdx.x;
  generic_tile<INV, R, EPI>(second ? a.job[1] : a.job[0], tile, a.n, a.ax, lds);
}

// Register-staged tile for n = R1*R2*R3 (R3 = 1: two stages) with pencils
// adjacent in memory: the first stage reads its inputs straight from global
// memory (16 pencils = 128-byte rows), the last stage writes its outputs
// straight to global memory (or folds |X|^2 into the FP64 accumulator), so
// a two-stage transform makes a single round trip through LDS.
template <typename R>
constexpr int reg_tile_pencils(int stages) {
  return (sizeof(R) == 4 ? 16 : 8) / (stages == 3 ? 2 : 1);
}

template <int RAD, bool INV, typename R>
__device__ __forceinline__ void twiddle_in(Cx<R>* v, int k, int step, const void* tw) {
#pragma unroll
  for (int r = 1; r < RAD; ++r) {
    if (k > 0) {
      const Cx<R> w = ld_tw<R>(tw, k * r * step);
      v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
    }
  }
}

template <typename R, int R1, int R2, int R3, bool INV, int EPI>
__device__ void reg_tile(const PencilJob<R>& J, long long tile, const FftAxis& ax, R* lds) {
  constexpr int n = R1 * R2 * R3;
  constexpr int P = reg_tile_pencils<R>(R3 == 1 ? 2 : 3);
  constexpr int pitch = n + 1;
  constexpr int plane = P * pitch;
  const void* tw = sizeof(R) == 8 ? (const void*)ax.tw64 : (const void*)ax.tw;
  const long long outer = tile / J.tiles_inner;
  const int inner = (int)(tile - outer * J.tiles_inner);
  const int p0 = inner * P;
  const int pc = min(P, J.pencils - p0);
  const long long base = outer * J.outer_stride + (long long)p0 * J.pstride;
  R* re0 = lds;
  R* im0 = lds + plane;
  // stage 1 (span 1): global -> DFT_R1 -> LDS plane 0 at j*R1 + r
  {
    constexpr int nq = n / R1;
    for (int b = threadIdx.x; b < P * nq; b += kFftThreads) {
      const int p = b % P, j = b / P;
      Cx<R> v[R1];
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        const long long g = base + p * J.pstride + (long long)(j + r * nq) * J.estride;
        if (p < pc) {
          const auto x = J.data[g];
          const R sc = J.scale ? (R)J.scale[g] : (R)1;
          v[r] = mk<R>(x.x * sc, x.y * sc);
        } else {
          v[r] = mk<R>(0, 0);
        }
      }
      dft<R1, INV>(v);
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        re0[p * pitch + j * R1 + r] = v[r].x;
        im0[p * pitch + j * R1 + r] = v[r].y;
      }
    }
    __syncthreads();
  }
  const R* sre = re0;
  const R* sim = im0;
  if constexpr (R3 > 1) {
    // middle stage (span R1): LDS plane 0 -> DFT_R2 -> LDS plane 1
    R* re1 = lds + 2 * plane;
    R* im1 = lds + 3 * plane;
    constexpr int ns = R1, nq = n / R2;
    for (int b = threadIdx.x; b < P * nq; b += kFftThreads) {
      const int p = b % P, j = b / P, k = j % ns;
      Cx<R> v[R2];
#pragma unroll
      for (int r = 0; r < R2; ++r) v[r] = mk<R>(re0[p * pitch + j + r * nq], im0[p * pitch + j + r * nq]);
      twiddle_in<R2, INV>(v, k, n / (ns * R2), tw);
      dft<R2, INV>(v);
      const int d0 = (j / ns) * ns * R2 + k;
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        re1[p * pitch + d0 + r * ns] = v[r].x;
        im1[p * pitch + d0 + r * ns] = v[r].y;
      }
    }
    __syncthreads();
    sre = re1;
    sim = im1;
  }
  // last stage (span n / RL): LDS -> DFT_RL -> global, element k + r * ns
  {
    constexpr int RL = R3 > 1 ? R3 : R2;
    constexpr int ns = n / RL, nq = n / RL;
    for (int b = threadIdx.x; b < P * nq; b += kFftThreads) {
      const int p = b % P, k = b / P;  // k < ns
      Cx<R> v[RL];
#pragma unroll
      for (int r = 0; r < RL; ++r) v[r] = mk<R>(sre[p * pitch + k + r * nq], sim[p * pitch + k + r * nq]);
      twiddle_in<RL, INV>(v, k, 1, tw);
      dft<RL, INV>(v);
      if (p < pc) {
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const long long g = base + p * J.pstride + (long long)(k + r * ns) * J.estride;
          if (EPI == EPI_STORE) {
            J.data[g].x = v[r].x;
            J.data[g].y = v[r].y;
          } else {
            J.acc[g] += (double)v[r].x * (double)v[r].x + (double)v[r].y * (double)v[r].y;
          }
        }
      }
    }
  }
}

template <typename R, int R1, int R2, int R3, bool INV, int EPI>
__global__ __launch_bounds__(kFftThreads) void k_pencils_reg(PencilArgs<R> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  R* lds = reinterpret_cast<R*>(lds_raw);
  if ((long long)blockIdx.x < a.blocks0) {
    reg_tile<R, R1, R2, R3, INV, EPI>(a.job[0], (long long)blockIdx.x, a.ax, lds);
  } else {
    generic_tile<INV, R, EPI>(a.job[1], (long long)blockIdx.x - a.blocks0, a.n, a.ax, lds);
  }
}

template <typename R>
int tile_pencils(int n) {
  int P = (int)(8192 / sizeof(R)) / (n > 0 ? n : 1);
  if (P > 16) P = 16;
  if (P < 1) P = 1;
  return P;
}

template <typename R>
size_t tile_lds(int n, int P) {
  return (size_t)P * lds_pitch(n) * 4 * sizeof(R);
}

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

// Radix split of the register-staged kernel for n (R3 = 1: two stages), or 0.
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

template <typename R, bool INV, int EPI>
void launch_reg(int n, dim3 g, dim3 b, size_t lds, hipStream_t s, const PencilArgs<R>& a) {
  switch (n) {
    case 16: hipLaunchKernelGGL((k_pencils_reg<R, 4, 4, 1, INV, EPI>), g, b, lds, s, a); break;
    case 32: hipLaunchKernelGGL((k_pencils_reg<R, 8, 4, 1, INV, EPI>), g, b, lds, s, a); break;
    case 64: hipLaunchKernelGGL((k_pencils_reg<R, 8, 8, 1, INV, EPI>), g, b, lds, s, a); break;
    case 128: hipLaunchKernelGGL((k_pencils_reg<R, 16, 8, 1, INV, EPI>), g, b, lds, s, a); break;
    case 256: hipLaunchKernelGGL((k_pencils_reg<R, 16, 16, 1, INV, EPI>), g, b, lds, s, a); break;
    case 512: hipLaunchKernelGGL((k_pencils_reg<R, 8, 8, 8, INV, EPI>), g, b, lds, s, a); break;
    default: hipLaunchKernelGGL((k_pencils_reg<R, 16, 8, 8, INV, EPI>), g, b, lds, s, a); break;
  }
}

// job[0] runs over outer0 outer groups, job[1] over outer1.  job[0] uses the
// register-staged kernel when n has a supported split and its pencils are
// adjacent in memory; job[1] (the Nyquist plane) always the generic tile.
// `choice`: 0 = default (register-staged when available), kTuneReg = the
// register-staged kernel, 1..16 = the generic LDS tile with that many pencils.
template <typename R>
hipError_t run_pencils(PencilArgs<R> a, long long outer0, long long outer1, bool inv, int epi,
                       hipStream_t s, int choice = 0) {
  int r1, r2, r3;
  bool reg = (choice == 0 || choice == kTuneReg) && reg_split(a.n, &r1, &r2, &r3) &&
             !a.job[0].pencil_major;
  if (reg) {
    const size_t need =
        (size_t)reg_tile_pencils<R>(r3 == 1 ? 2 : 3) * (a.n + 1) * (r3 == 1 ? 2 : 4) * sizeof(R);
    reg = need <= max_lds_per_block();
  }
  long long blocks[2];
  const long long outer[2] = {outer0, outer1};
  size_t lds = 0;
  for (int j = 0; j < 2; ++j) {
    PencilJob<R>& J = a.job[j];
    const bool use_reg = reg && j == 0;
    const bool forced = j == 0 && choice >= 1 && choice <= 16 &&
                        tile_lds<R>(a.n, choice) <= max_lds_per_block();
    J.P = use_reg ? reg_tile_pencils<R>(r3 == 1 ? 2 : 3) : (forced ? choice : tile_pencils<R>(a.n));
    J.tiles_inner = (J.pencils + J.P - 1) / J.P;
    blocks[j] = J.pencils > 0 ? outer[j] * J.tiles_inner : 0;
    if (blocks[j] > 0) {
      const size_t need = use_reg ? (size_t)J.P * (a.n + 1) * (r3 == 1 ? 2 : 4) * sizeof(R)
                                  : tile_lds<R>(a.n, J.P);
      lds = need > lds ? need : lds;
    }
  }
  a.blocks0 = blocks[0];
  const long long total = blocks[0] + blocks[1];
  if (total == 0) return hipSuccess;
  const dim3 g((unsigned)total), b(kFftThreads);
  if (reg) {
    if (inv && epi == EPI_STORE) launch_reg<R, true, EPI_STORE>(a.n, g, b, lds, s, a);
    else if (!inv && epi == EPI_STORE) launch_reg<R, false, EPI_STORE>(a.n, g, b, lds, s, a);
    else launch_reg<R, false, EPI_ACC>(a.n, g, b, lds, s, a);
  } else {
    if (inv && epi == EPI_STORE) hipLaunchKernelGGL((k_pencils<true, R, EPI_STORE>), g, b, lds, s, a);
    else if (!inv && epi == EPI_STORE) hipLaunchKernelGGL((k_pencils<false, R, EPI_STORE>), g, b, lds, s, a);
    else hipLaunchKernelGGL((k_pencils<false, R, EPI_ACC>), g, b, lds, s, a);
  }
  return hipGetLastError();
}

template <typename R>
typename V2<R>::T* spec_main(Plan& p) {
  return static_cast<typename V2<R>::T*>(p.spec_main);
}
template <typename R>
typename V2<R>::T* spec_nyq(Plan& p) {
  return static_cast<typename V2<R>::T*>(p.spec_nyq);
}

// Rows per block of the register-staged z pass for half-length m (FP32), 0
// when m has no two-stage split.
inline int z_reg_rows(int m) {
  switch (m) {
    case 64: return z_rows_per_block(64, 8);
    case 128: return z_rows_per_block(128, 16);
    case 256: return z_rows_per_block(256, 16);
    default: return 0;
  }
}

template <typename R>
hipError_t pass_z(Plan& p, const void* density, hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  R2CArgs<R> a;
  a.in = static_cast<const R*>(density);
  a.main = spec_main<R>(p);
  a.nyq = spec_nyq<R>(p);
  a.rows = (long long)nx * ny;
  a.m = m;
  a.P = (p.tune_z >= 1 && p.tune_z <= 16 && tile_lds<R>(m, p.tune_z) <= max_lds_per_block())
            ? p.tune_z
            : tile_pencils<R>(m);
  a.ax = p.ax_zc;
  a.tw2 = sizeof(R) == 8 ? (const void*)p.tw_r2c64 : (const void*)p.tw_r2c;
  if constexpr (sizeof(R) == 4) {
    if (z_reg_rows(m) > 0 && (p.tune_z == 0 || p.tune_z == kTuneReg)) {
      // register-staged kernel: the default where m has a two-stage split
      const int P = z_reg_rows(m);
      const dim3 g((unsigned)((a.rows + P - 1) / P)), b(kFftThreads);
      switch (m) {
        case 64: hipLaunchKernelGGL((k_r2c_z_reg<R, 8, 8>), g, b, 0, s, a); break;
        case 128: hipLaunchKernelGGL((k_r2c_z_reg<R, 16, 8>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((k_r2c_z_reg<R, 16, 16>), g, b, 0, s, a); break;
      }
      return hipGetLastError();
    }
  }
  const long long blocks = (a.rows + a.P - 1) / a.P;
  hipLaunchKernelGGL((k_r2c_z<R>), dim3((unsigned)blocks), dim3(kFftThreads),
                     tile_lds<R>(m, a.P), s, a);
  return hipGetLastError();
}

// y pass: main (pencils = kz, elements = y) + nyq (pencils = x, elements contiguous)
template <typename R>
hipError_t pass_y(Plan& p, bool inv, hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  PencilArgs<R> a{};
  a.n = ny;
  a.ax = p.ax_y;
  PencilJob<R>& J = a.job[0];
  J.data = spec_main<R>(p);
  J.estride = m;
  J.pstride = 1;
  J.outer_stride = (long long)ny * m;
  J.pencils = m;
  PencilJob<R>& K = a.job[1];
  K.data = spec_nyq<R>(p);
  K.estride = 1;
  K.pstride = ny;
  K.pencils = nx;
  K.pencil_major = 1;
  return run_pencils(a, nx, 1, inv, EPI_STORE, s, p.tune_y);
}

// x pass: main (pencils = (y, kz), elements = x) + nyq (pencils = y).  Forward
// with `accumulate` adds |F|^2 into sf_main / sf_nyq; the inverse applies the
// per-bin FP64 weights w first.
template <typename R>
hipError_t pass_x(Plan& p, bool inv, bool accumulate, const double* w_main, const double* w_nyq,
                  hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  PencilArgs<R> a{};
  a.n = nx;
  a.ax = p.ax_x;
  PencilJob<R>& J = a.job[0];
  J.data = spec_main<R>(p);
  J.acc = p.sf_main;
  J.scale = w_main;
  J.estride = (long long)ny * m;
  J.pstride = 1;
  J.pencils = ny * m;
  PencilJob<R>& K = a.job[1];
  K.data = spec_nyq<R>(p);
  K.acc = p.sf_nyq;
  K.scale = w_nyq;
  K.estride = ny;
  K.pstride = 1;
  K.pencils = ny;
  return run_pencils(a, 1, 1, inv, accumulate ? EPI_ACC : EPI_STORE, s, p.tune_x);
}

template <typename R>
hipError_t forward_impl(Plan& p, const void* density, bool accumulate, hipStream_t s) {
  hipError_t e = pass_z<R>(p, density, s);
  if (e != hipSuccess) return e;
  e = pass_y<R>(p, false, s);
  if (e != hipSuccess) return e;
  return pass_x<R>(p, false, accumulate, nullptr, nullptr, s);
}

// Real adjoint of the forward transform applied to the stored spectrum, each
// bin first multiplied by the FP64 weight w (internal layout):
//   out(r) = Re sum_{kx, ky, kz in [0, Nz/2]} w(q) X(q) e^{+i q.r}.
template <typename R>
hipError_t inverse_impl(Plan& p, const double* w_main, const double* w_nyq, void* out,
                        hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  hipError_t e = pass_x<R>(p, true, false, w_main, w_nyq, s);
  if (e != hipSuccess) return e;
  e = pass_y<R>(p, true, s);
  if (e != hipSuccess) return e;
  C2RArgs<R, R> a;
  a.main = spec_main<R>(p);
  a.nyq = spec_nyq<R>(p);
  a.out = static_cast<R*>(out);
  a.rows = (long long)nx * ny;
  a.m = m;
  a.P = tile_pencils<R>(m);
  a.ax = p.ax_zc;
  a.tw2 = sizeof(R) == 8 ? (const void*)p.tw_r2c64 : (const void*)p.tw_r2c;
  if constexpr (sizeof(R) == 4) {
    if (z_reg_rows(m) > 0) {
      const int P = z_reg_rows(m);
      const dim3 g((unsigned)((a.rows + P - 1) / P)), b(kFftThreads);
      switch (m) {
        case 64: hipLaunchKernelGGL((k_c2r_z_reg<R, 8, 8>), g, b, 0, s, a); break;
        case 128: hipLaunchKernelGGL((k_c2r_z_reg<R, 16, 8>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((k_c2r_z_reg<R, 16, 16>), g, b, 0, s, a); break;
      }
      return hipGetLastError();
    }
  }
  const long long blocks = (a.rows + a.P - 1) / a.P;
  hipLaunchKernelGGL((k_c2r_z<R, R>), dim3((unsigned)blocks), dim3(kFftThreads),
                     tile_lds<R>(m, a.P), s, a);
  return hipGetLastError();
}

template <typename R>
__global__ void k_spec_read(const typename V2<R>::T* __restrict__ main,
                            const typename V2<R>::T* __restrict__ nyq, double2* __restrict__ out,
                            int nx, int ny, int m) {
  const long long h = m + 1;
  const long long total = (long long)nx * ny * h;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long row = i / h;
    const int kz = (int)(i - row * h);
    const auto v = kz < m ? main[row * m + kz] : nyq[row];
    out[i] = make_double2((double)v.x, (double)v.y);
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
  const int cand_p[5] = {kTuneReg, 2, 4, 8, 16};
  // z pass: the register-staged kernel (FP32, m with a two-stage split) or
  // the generic tile with that many rows
  float best = 1e30f;
  int pick = 0;
  for (int P : cand_p) {
    if (P == kTuneReg && !(sizeof(R) == 4 && z_reg_rows(m) > 0)) continue;
    if (P != kTuneReg && tile_lds<R>(m, P) > max_lds_per_block()) continue;
    p.tune_z = P;
    const float t = time_best([&] { return pass_z<R>(p, p.rho, s); }, s, e0, e1, &err);
    if (err != hipSuccess) break;
    if (t < best) {
      best = t;
      pick = P;
    }
  }
  p.tune_z = pick;
  // y and x passes: the register-staged kernel or the generic tile
  const int cand_c[5] = {kTuneReg, 2, 4, 8, 16};
  for (int axis = 0; axis < 2 && err == hipSuccess; ++axis) {
    int& slot = axis == 0 ? p.tune_y : p.tune_x;
    best = 1e30f;
    pick = 0;
    for (int c : cand_c) {
      slot = c;
      const float t = time_best(
          [&] { return axis == 0 ? pass_y<R>(p, false, s) : pass_x<R>(p, false, true, nullptr, nullptr, s); },
          s, e0, e1, &err);
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

// Picks the z tile height and the y / x pencil kernels for this grid by timing
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

hipError_t fft_pass(Plan& p, int pass, const void* density, hipStream_t s) {
  switch (pass) {
    case 0: return p.fft64 ? pass_z<double>(p, density, s) : pass_z<float>(p, density, s);
    case 1: return p.fft64 ? pass_y<double>(p, false, s) : pass_y<float>(p, false, s);
    default:
      return p.fft64 ? pass_x<double>(p, false, true, nullptr, nullptr, s)
                     : pass_x<float>(p, false, true, nullptr, nullptr, s);
  }
}

hipError_t fft_spectrum_read(Plan& p, double* out, hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  long long total = (long long)nx * ny * (m + 1);
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (p.fft64)
    hipLaunchKernelGGL(k_spec_read<double>, dim3((unsigned)blocks), dim3(256), 0, s,
                       spec_main<double>(p), spec_nyq<double>(p), reinterpret_cast<double2*>(out),
                       nx, ny, m);
  else
    hipLaunchKernelGGL(k_spec_read<float>, dim3((unsigned)blocks), dim3(256), 0, s,
                       spec_main<float>(p), spec_nyq<float>(p), reinterpret_cast<double2*>(out),
                       nx, ny, m);
  return hipGetLastError();
}

// --- plan-free transform for callers that cannot allocate -----------------------
// The working set of mdsf_rfft3_noalloc (twiddles, half spectrum, Nyquist
// plane) lives in this fixed device array, one copy per GPU, held until the
// process ends.  Host code never passes the array's name to a kernel: it is a
// host-side stand-in.  The device address of each GPU's copy comes from
// hipGetSymbolAddress on the first call on that device and is kept per device.
namespace {

constexpr size_t kScratchBytes = (size_t)80 << 20;
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
  const size_t esz = precision == MDSF_F64 ? 8 : 4;
  const size_t nx = n[0], ny = n[1], nz = n[2], m = nz / 2;
  // layout: main half spectrum (nx ny m complex), Nyquist plane (nx ny
  // complex), then the FP32 and FP64 twiddles of the x, y, z/2 axes and of
  // the R2C split (length nz)
  const size_t lens[4] = {nx, ny, m, nz};
  size_t off = 0;
  const size_t off_main = off;
  off = align256(off + 2 * esz * nx * ny * m);
  const size_t off_nyq = off;
  off = align256(off + 2 * esz * nx * ny);
  size_t off_tw[4], off_tw64[4];
  for (int k = 0; k < 4; ++k) {
    off_tw[k] = off;
    off = align256(off + sizeof(float2) * lens[k]);
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
  p.fft64 = precision == MDSF_F64 ? 1 : 0;
  p.spec_main = p.rho = base + off_main;
  p.spec_nyq = base + off_nyq;
  FftAxis* axes[3] = {&p.ax_x, &p.ax_y, &p.ax_zc};
  for (int k = 0; k < 4; ++k) {
    float2* tw = reinterpret_cast<float2*>(base + off_tw[k]);
    double2* tw64 = reinterpret_cast<double2*>(base + off_tw64[k]);
    if (k < 3) {
      FftAxis& ax = *axes[k];
      ax.n = (int)lens[k];
      if (!fft_factor(ax.n, ax.radix, &ax.nstages))
        return fail(MDSF_ERR_UNSUPPORTED,
                    "FFT length " + std::to_string(ax.n) + " needs too many stages");
      ax.tw = tw;
      ax.tw64 = tw64;
    } else {
      p.tw_r2c = tw;
      p.tw_r2c64 = tw64;
    }
    MDSF_HIP(fft_make_twiddles(tw, tw64, (int)lens[k], s));
  }
  MDSF_HIP(fft_forward(p, d_density, false, s));
  MDSF_HIP(fft_spectrum_read(p, d_spectrum, s));
  return MDSF_OK;
}
