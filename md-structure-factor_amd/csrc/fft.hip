// 3-D real-to-complex FFT for gfx950: Stockham autosort stages in LDS.
//
// Replaces np.fft.rfftn (dens.py:313) and the |F|^2 frame sum
// (dens.py:315-318).  Unnormalised forward transform, exponent -2*pi*i*k*n/N.
//
// Layout between passes (internal): the z pass packs each real row of Nz
// floats as Nz/2 complex values and writes the half spectrum in place,
// kz in [0, Nz/2) -> main[x][y][kz], and the kz = Nz/2 plane -> nyq[x][y].
// The y and x passes are strided pencil transforms over those arrays; the
// x pass can fold |F|^2 into FP64 accumulators of the same layout.
//
// Each workgroup owns a tile of P pencils.  LDS holds re/im planes of
// P x pitch floats (pitch odd), and threads are mapped pencil-fastest, so
// every Stockham read and write is conflict-free for any radix and span.
#include "mdsf_internal.h"

namespace mdsf {

namespace {

constexpr int kFftThreads = 256;
constexpr int kRegBudget = 16;  // complex values held per thread per stage

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// --- small DFT kernels (in registers) ---------------------------------------
template <bool INV>
__device__ __forceinline__ void dft2(float2* v) {
  const float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void dft3(float2* v) {
  const float s = 0.86602540378443864676f;  // sin(2 pi / 3)
  const float2 t1 = cadd(v[1], v[2]);
  const float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
  const float2 d = csub(v[1], v[2]);
  const float2 t3 = mul_mi<INV>(make_float2(s * d.x, s * d.y));
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2* v) {
  const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  const float2 t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft5(float2* v) {
  const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
  const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
  const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const float2 x0 = v[0];
  const float2 a1 = make_float2(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
  const float2 a2 = make_float2(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
  const float2 b1 = mul_mi<INV>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  const float2 b2 = mul_mi<INV>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  v[0] = make_float2(x0.x + t1.x + t2.x, x0.y + t1.y + t2.y);
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2* v) {
  const float h = 0.70710678118654752440f;
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(e);
  dft4<INV>(o);
  // W8^k for k = 0..3 (forward: exp(-i pi k / 4))
  const float2 w1 = make_float2(h, INV ? h : -h);
  const float2 w3 = make_float2(-h, INV ? h : -h);
  const float2 o1 = cmul(o[1], w1);
  const float2 o2 = mul_mi<INV>(o[2]);
  const float2 o3 = cmul(o[3], w3);
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
__device__ __forceinline__ void dft16(float2* v) {
  float2 e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    e[i] = v[2 * i];
    o[i] = v[2 * i + 1];
  }
  dft8<INV>(e);
  dft8<INV>(o);
  // W16^k, k = 0..7, forward exp(-2 pi i k / 16)
  const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
  const float h = 0.70710678118654752440f;
  const float sg = INV ? 1.0f : -1.0f;
  const float2 w[8] = {make_float2(1.f, 0.f),   make_float2(c1, sg * s1), make_float2(h, sg * h),
                       make_float2(s1, sg * c1), make_float2(0.f, sg),     make_float2(-s1, sg * c1),
                       make_float2(-h, sg * h),  make_float2(-c1, sg * s1)};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float2 t = k == 0 ? o[0] : cmul(o[k], w[k]);
    v[k] = cadd(e[k], t);
    v[k + 8] = csub(e[k], t);
  }
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 2) dft2<INV>(v);
  else if constexpr (R == 3) dft3<INV>(v);
  else if constexpr (R == 4) dft4<INV>(v);
  else if constexpr (R == 5) dft5<INV>(v);
  else if constexpr (R == 8) dft8<INV>(v);
  else dft16<INV>(v);
}

// One Stockham stage, out of place (in -> out planes), one butterfly per
// thread iteration: X[(j/ns)*ns*R + k + r*ns] = DFT_R(x[j + r*n/R] * W^(k r n/(ns R))).
template <int R, bool INV>
__device__ void stage_fixed(const float* __restrict__ ire, const float* __restrict__ iim,
                            float* __restrict__ ore, float* __restrict__ oim, int n, int P,
                            int pitch, int ns, const float2* __restrict__ tw) {
  const int nq = n / R;
  const int nbf = P * nq;
  const int twstep = n / (ns * R);
  for (int b = threadIdx.x; b < nbf; b += kFftThreads) {
    const int p = b % P, j = b / P;
    const int k = j % ns;
    const int row = p * pitch;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int a = row + j + r * nq;
      float2 x = make_float2(ire[a], iim[a]);
      if (r > 0 && k > 0) {
        const float2 w = tw[k * r * twstep];
        x = INV ? cmulc(x, w) : cmul(x, w);
      }
      v[r] = x;
    }
    dft<R, INV>(v);
    const int d0 = row + (j / ns) * ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      ore[d0 + r * ns] = v[r].x;
      oim[d0 + r * ns] = v[r].y;
    }
  }
  __syncthreads();
}

// Any other radix (odd primes): one output element per thread iteration,
// direct DFT with W_R^m = W_n^(m n / R) from the twiddle table.
template <bool INV>
__device__ void stage_generic(const float* __restrict__ ire, const float* __restrict__ iim,
                              float* __restrict__ ore, float* __restrict__ oim, int n, int P,
                              int pitch, int ns, int R, const float2* __restrict__ tw) {
  const int nq = n / R;
  const int twstep = n / (ns * R);
  const int step = n / R;
  for (int o = threadIdx.x; o < P * n; o += kFftThreads) {
    const int p = o % P, rest = o / P;
    const int j = rest % nq, m = rest / nq;
    const int k = j % ns;
    const int row = p * pitch;
    float2 acc = make_float2(0.f, 0.f);
    int wm = 0;  // (r * m) mod R
    for (int r = 0; r < R; ++r) {
      const int a = row + j + r * nq;
      float2 x = make_float2(ire[a], iim[a]);
      if (r > 0 && k > 0) {
        const float2 w = tw[k * r * twstep];
        x = INV ? cmulc(x, w) : cmul(x, w);
      }
      const float2 w = tw[wm * step];
      acc = cadd(acc, INV ? cmulc(x, w) : cmul(x, w));
      wm += m;
      if (wm >= R) wm -= R;
    }
    const int d = row + (j / ns) * ns * R + k + m * ns;
    ore[d] = acc.x;
    oim[d] = acc.y;
  }
  __syncthreads();
}

// Runs every stage of `ax`, ping-ponging between plane pairs 0 and 1 of
// `lds` ([re0 | im0 | re1 | im1], P*pitch floats each).  Returns the plane
// pair holding the result.
template <bool INV>
__device__ int run_stages(float* lds, int n, int P, int pitch, const FftAxis& ax) {
  const int plane = P * pitch;
  int ns = 1, cur = 0;
  for (int s = 0; s < ax.nstages; ++s) {
    const int R = ax.radix[s];
    const float* ire = lds + (2 * cur) * plane;
    const float* iim = ire + plane;
    float* ore = lds + (2 * (cur ^ 1)) * plane;
    float* oim = ore + plane;
    switch (R) {
      case 2: stage_fixed<2, INV>(ire, iim, ore, oim, n, P, pitch, ns, ax.tw); break;
      case 3: stage_fixed<3, INV>(ire, iim, ore, oim, n, P, pitch, ns, ax.tw); break;
      case 4: stage_fixed<4, INV>(ire, iim, ore, oim, n, P, pitch, ns, ax.tw); break;
      case 5: stage_fixed<5, INV>(ire, iim, ore, oim, n, P, pitch, ns, ax.tw); break;
      case 8: stage_fixed<8, INV>(ire, iim, ore, oim, n, P, pitch, ns, ax.tw); break;
      case 16: stage_fixed<16, INV>(ire, iim, ore, oim, n, P, pitch, ns, ax.tw); break;
      default: stage_generic<INV>(ire, iim, ore, oim, n, P, pitch, ns, R, ax.tw); break;
    }
    ns *= R;
    cur ^= 1;
  }
  return cur;
}

__host__ __device__ inline int lds_pitch(int n) { return (n % 2 == 0) ? n + 1 : n; }

// --- twiddles: W_n^k = exp(-2 pi i k / n), generated in FP64 ------------------
__global__ void k_twiddles(float2* __restrict__ tw, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double s, c;
  sincospi(-2.0 * (double)k / (double)n, &s, &c);
  tw[k] = make_float2((float)c, (float)s);
}

// --- z pass: real rows -> half spectrum (in place) + Nyquist plane ------------
struct R2CArgs {
  const float* in;     // rows of nz reals (may alias main)
  float2* main;        // rows of m = nz/2 complex
  float2* nyq;         // one complex per row
  long long rows;
  int m;               // nz / 2
  int P;               // rows per tile
  FftAxis ax;          // length m
  const float2* tw2;   // W_nz^k, k < nz
};

__global__ __launch_bounds__(kFftThreads) void k_r2c_z(R2CArgs a) {
  extern __shared__ float lds[];
  const int m = a.m, P = a.P, pitch = lds_pitch(m);
  float* re = lds;
  float* im = lds + P * pitch;
  const long long r0 = (long long)blockIdx.x * P;
  const int pc = (int)min((long long)P, a.rows - r0);
  const float2* src = reinterpret_cast<const float2*>(a.in) + r0 * m;
  for (int idx = threadIdx.x; idx < pc * m; idx += kFftThreads) {
    const int p = idx / m, e = idx - p * m;
    const float2 v = src[idx];
    re[p * pitch + e] = v.x;
    im[p * pitch + e] = v.y;
  }
  __syncthreads();
  {
    const int cur = run_stages<false>(lds, m, P, pitch, a.ax);
    re = lds + 2 * cur * P * pitch;
    im = re + P * pitch;
  }
  // split: X_k = A_k + W_nz^k B_k, A = (Z_k + conj Z_{m-k})/2, B = (Z_k - conj Z_{m-k})/(2i)
  float2* dst = a.main + r0 * m;
  for (int idx = threadIdx.x; idx < pc * m; idx += kFftThreads) {
    const int p = idx / m, k = idx - p * m;
    const int kk = (k == 0) ? 0 : m - k;
    const float2 z = make_float2(re[p * pitch + k], im[p * pitch + k]);
    const float2 zc = make_float2(re[p * pitch + kk], -im[p * pitch + kk]);
    const float2 A = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
    const float2 d = make_float2(0.5f * (z.x - zc.x), 0.5f * (z.y - zc.y));
    const float2 B = make_float2(d.y, -d.x);  // d / i
    const float2 w = a.tw2[k];
    dst[idx] = cadd(A, cmul(B, w));
  }
  for (int p = threadIdx.x; p < pc; p += kFftThreads) {
    const float2 z = make_float2(re[p * pitch], im[p * pitch]);
    // X_{m} = A_0 - B_0 = Re z - Im z (both real)
    a.nyq[r0 + p] = make_float2(z.x - z.y, 0.0f);
  }
}

// --- strided pencil pass ------------------------------------------------------
struct PencilArgs {
  float2* data;
  double* acc;              // if set: acc += |X|^2 (FP64) at the same index
  int write_back;           // store the transform into data
  long long estride;        // between elements of a pencil
  long long pstride;        // between adjacent pencils of a tile
  long long outer_stride;   // between outer groups
  int pencils;              // pencils per outer group
  int tiles_inner;          // ceil(pencils / P)
  int P;
  int n;
  int pencil_major;         // elements contiguous (estride == 1)
  FftAxis ax;
};

template <bool INV>
__global__ __launch_bounds__(kFftThreads) void k_pencils(PencilArgs a) {
  extern __shared__ float lds[];
  const int n = a.n, P = a.P, pitch = lds_pitch(n);
  float* re = lds;
  float* im = lds + P * pitch;
  const long long tile = blockIdx.x;
  const long long outer = tile / a.tiles_inner;
  const int inner = (int)(tile - outer * a.tiles_inner);
  const int p0 = inner * P;
  const int pc = min(P, a.pencils - p0);
  const long long base = outer * a.outer_stride + (long long)p0 * a.pstride;
  const int total = pc * n;
  for (int idx = threadIdx.x; idx < total; idx += kFftThreads) {
    int p, e;
    if (a.pencil_major) {
      p = idx / n;
      e = idx - p * n;
    } else {
      e = idx / pc;
      p = idx - e * pc;
    }
    const float2 v = a.data[base + p * a.pstride + e * a.estride];
    re[p * pitch + e] = v.x;
    im[p * pitch + e] = v.y;
  }
  __syncthreads();
  {
    const int cur = run_stages<INV>(lds, n, P, pitch, a.ax);
    re = lds + 2 * cur * P * pitch;
    im = re + P * pitch;
  }
  for (int idx = threadIdx.x; idx < total; idx += kFftThreads) {
    int p, e;
    if (a.pencil_major) {
      p = idx / n;
      e = idx - p * n;
    } else {
      e = idx / pc;
      p = idx - e * pc;
    }
    const float xr = re[p * pitch + e], xi = im[p * pitch + e];
    const long long g = base + p * a.pstride + e * a.estride;
    if (a.write_back) a.data[g] = make_float2(xr, xi);
    if (a.acc) a.acc[g] += (double)xr * (double)xr + (double)xi * (double)xi;
  }
}

int tile_pencils(int n) {
  int P = 2048 / (n > 0 ? n : 1);
  if (P > 16) P = 16;
  if (P < 1) P = 1;
  return P;
}

hipError_t run_pencils(const PencilArgs& a0, long long outer_count, hipStream_t s, bool inv) {
  PencilArgs a = a0;
  a.P = tile_pencils(a.n);
  a.tiles_inner = (a.pencils + a.P - 1) / a.P;
  const long long blocks = outer_count * a.tiles_inner;
  if (blocks == 0) return hipSuccess;
  const size_t lds = (size_t)a.P * lds_pitch(a.n) * 4 * sizeof(float);
  if (inv)
    hipLaunchKernelGGL(k_pencils<true>, dim3((unsigned)blocks), dim3(kFftThreads), lds, s, a);
  else
    hipLaunchKernelGGL(k_pencils<false>, dim3((unsigned)blocks), dim3(kFftThreads), lds, s, a);
  return hipGetLastError();
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

hipError_t fft_make_twiddles(float2* tw, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_twiddles, dim3((n + 255) / 256), dim3(256), 0, s, tw, n);
  return hipGetLastError();
}

hipError_t fft_r2c_z(Plan& p, const float* in, float2* main, float2* nyq, hipStream_t s) {
  R2CArgs a;
  a.in = in;
  a.main = main;
  a.nyq = nyq;
  a.rows = (long long)p.sys.n[0] * p.sys.n[1];
  a.m = p.sys.n[2] / 2;
  a.P = tile_pencils(a.m);
  a.ax = p.ax_zc;
  a.tw2 = p.tw_r2c;
  const long long blocks = (a.rows + a.P - 1) / a.P;
  const size_t lds = (size_t)a.P * lds_pitch(a.m) * 4 * sizeof(float);
  hipLaunchKernelGGL(k_r2c_z, dim3((unsigned)blocks), dim3(kFftThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t fft_y(Plan& p, float2* main, float2* nyq, hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  PencilArgs a{};
  a.data = main;
  a.acc = nullptr;
  a.write_back = 1;
  a.estride = m;
  a.pstride = 1;
  a.outer_stride = (long long)ny * m;
  a.pencils = m;
  a.n = ny;
  a.pencil_major = 0;
  a.ax = p.ax_y;
  hipError_t e = run_pencils(a, nx, s, false);
  if (e != hipSuccess) return e;
  PencilArgs b{};
  b.data = nyq;
  b.acc = nullptr;
  b.write_back = 1;
  b.estride = 1;
  b.pstride = ny;
  b.outer_stride = 0;
  b.pencils = nx;
  b.n = ny;
  b.pencil_major = 1;
  b.ax = p.ax_y;
  return run_pencils(b, 1, s, false);
}

static hipError_t fft_x_impl(Plan& p, float2* main, float2* nyq, bool accum, hipStream_t s) {
  const int nx = p.sys.n[0], ny = p.sys.n[1], m = p.sys.n[2] / 2;
  PencilArgs a{};
  a.data = main;
  a.acc = accum ? p.sf_main : nullptr;
  a.write_back = accum ? 0 : 1;
  a.estride = (long long)ny * m;
  a.pstride = 1;
  a.outer_stride = 0;
  a.pencils = ny * m;
  a.n = nx;
  a.pencil_major = 0;
  a.ax = p.ax_x;
  hipError_t e = run_pencils(a, 1, s, false);
  if (e != hipSuccess) return e;
  PencilArgs b{};
  b.data = nyq;
  b.acc = accum ? p.sf_nyq : nullptr;
  b.write_back = accum ? 0 : 1;
  b.estride = ny;
  b.pstride = 1;
  b.outer_stride = 0;
  b.pencils = ny;
  b.n = nx;
  b.pencil_major = 0;
  b.ax = p.ax_x;
  return run_pencils(b, 1, s, false);
}

hipError_t fft_x_accum(Plan& p, float2* main, float2* nyq, hipStream_t s) {
  return fft_x_impl(p, main, nyq, true, s);
}
hipError_t fft_x(Plan& p, float2* main, float2* nyq, hipStream_t s) {
  return fft_x_impl(p, main, nyq, false, s);
}

}  // namespace mdsf
