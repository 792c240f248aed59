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
#include "mdsf_internal.h"

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
                       int* __restrict__ home, int* __restrict__ err, int coords_only) {
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
  coords[3 * (size_t)i + 0] = x[0];
  coords[3 * (size_t)i + 1] = x[1];
  coords[3 * (size_t)i + 2] = x[2];
  if (coords_only) return;

  const int t = types[i];
  if (t < 0 || t >= sys.ntypes) {
    atomicOr(err, ERRBIT_TYPE);
    key[i] = -1;
    return;
  }
  const TypeDev& ty = td[t];
  int ir[3], h[3];
  bool ok = true;
  bool nonc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double q = (double)x[d] / sys.dr[d];
    if (!(q > -1.0e9 && q < 1.0e9)) {
      ok = false;
      ir[d] = 0;
    } else {
      ir[d] = (int)q;  // truncation toward zero, as ndarray.astype(int)
    }
    const int A = ty.half[d];
    if (ir[d] - A < -sys.nb || ir[d] + A > sys.n[d] + sys.nb) ok = false;
    nonc[d] = A > 0 && (ir[d] - A < 0 || ir[d] + A > sys.n[d]);
    h[d] = pos_mod(ir[d], sys.n[d]);
    if (home) home[3 * (size_t)i + d] = ir[d];
  }
  if (!ok) {
    atomicOr(err, ERRBIT_RANGE);
    key[i] = -1;
    return;
  }
  const int corner = (nonc[0] && nonc[1] && nonc[2]) ? 1 : 0;
  const int b = ty.bucket_base +
                ((h[0] / ty.bsz[0]) * ty.nbk[1] + h[1] / ty.bsz[1]) * ty.nbk[2] + h[2] / ty.bsz[2];
  key[i] = b;
  ir4[i] = make_int4(ir[0], ir[1], ir[2], t | (corner << 16));
  rpos[3 * (size_t)i + 0] = (double)x[0];
  rpos[3 * (size_t)i + 1] = (double)x[1];
  rpos[3 * (size_t)i + 2] = (double)x[2];
  atomicAdd(&bcount[b], 1);
  if (occ) atomicAdd(&occ[((size_t)h[0] * sys.n[1] + h[1]) * sys.n[2] + h[2]], 1u);
}

// Exclusive scan of the bucket counts (one workgroup), also seeds the fill
// cursors and clears the counts for the next frame.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_scan(int* __restrict__ bcount,
                                                       int* __restrict__ boff,
                                                       int* __restrict__ bcur, int nbk) {
  __shared__ int part[kScanThreads];
  const int tid = threadIdx.x;
  const int per = (nbk + kScanThreads - 1) / kScanThreads;
  const int lo = min(nbk, tid * per), hi = min(nbk, lo + per);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += bcount[i];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - s;
  for (int i = lo; i < hi; ++i) {
    const int c = bcount[i];
    boff[i] = run;
    bcur[i] = run;
    bcount[i] = 0;
    run += c;
  }
  if (tid == kScanThreads - 1) boff[nbk] = part[tid];
}

__global__ void k_fill(const int* __restrict__ key, int* __restrict__ bcur,
                       int* __restrict__ s_idx, int natoms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const int b = key[i];
  if (b < 0) return;
  const int pos = atomicAdd(&bcur[b], 1);
  s_idx[pos] = i;
}

// Restores atom-index order inside each bucket so that the gather visits
// atoms in a canonical order (bitwise-reproducible sums).
__global__ void k_bucket_sort(const int* __restrict__ boff, int* __restrict__ s_idx, int nbk) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbk) return;
  const int lo = boff[b], hi = boff[b + 1];
  for (int i = lo + 1; i < hi; ++i) {
    const int v = s_idx[i];
    int j = i - 1;
    while (j >= lo && s_idx[j] > v) {
      s_idx[j + 1] = s_idx[j];
      --j;
    }
    s_idx[j + 1] = v;
  }
}

// Gathers atoms into sorted order and writes their 1-D factor tables:
//   x: amp * exp(-g00 bx^2 / 2s^2), y: exp(-g11 by^2 / 2s^2), z: exp(-g22 bz^2 / 2s^2)
// with b = r - (ir - A + k) dr in FP64 exactly as buf = r - grid (dens.py:252-256,299),
// k in [0, 2A).  Values are rounded once to the plan's precision (FP32 or FP64).
constexpr int kTabThreads = 64;
template <typename TT>
__global__ __launch_bounds__(256) void k_gather_tables(SysDev sys, const TypeDev* __restrict__ td,
                                                       const int* __restrict__ s_idx,
                                                       const int4* __restrict__ ir4,
                                                       const double* __restrict__ rpos,
                                                       int4* __restrict__ s_ir,
                                                       double* __restrict__ s_r,
                                                       TT* __restrict__ tables,
                                                       const int* __restrict__ nvalid) {
#pragma clang fp contract(off)
  const int s = blockIdx.x * (blockDim.x / kTabThreads) + threadIdx.x / kTabThreads;
  const int lane = threadIdx.x % kTabThreads;
  if (s >= *nvalid) return;
  const int i = s_idx[s];
  const int4 q = ir4[i];
  const double r[3] = {rpos[3 * (size_t)i], rpos[3 * (size_t)i + 1], rpos[3 * (size_t)i + 2]};
  if (lane == 0) {
    s_ir[s] = q;
    s_r[3 * (size_t)s + 0] = r[0];
    s_r[3 * (size_t)s + 1] = r[1];
    s_r[3 * (size_t)s + 2] = r[2];
  }
  const int t = q.w & 0xffff;
  const TypeDev& ty = td[t];
  // sorted positions match the per-type table slots only when every atom was
  // valid; after a range / type error (reported to the host) write nothing
  const int slot = s - ty.type_start;
  if (slot < 0 || slot >= ty.count) return;
  TT* tab = tables + ty.table_base + (long long)slot * ty.stride;
  const int ir[3] = {q.x, q.y, q.z};
  const double g[3] = {sys.g00, sys.g11, sys.g22};
  const int len0 = 2 * ty.half[0], len1 = 2 * ty.half[1], len2 = 2 * ty.half[2];
  const int total = len0 + len1 + len2;
  for (int e = lane; e < total; e += kTabThreads) {
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
    const int u = ir[d] - ty.half[d] + k;
    const double gridv = (double)u * sys.dr[d];
    const double b = r[d] - gridv;
    const double qv = (g[d] * (b * b)) / ty.two_sig2;
    double v = exp(-qv);
    if (d == 0) v = ty.amp * v;
    tab[e] = (TT)v;
  }
}

}  // namespace

hipError_t launch_prep(Plan& p, void* coords, int dtype, int arith, const int32_t* types,
                       const double* scale3, const double box[3], int64_t wrap_lo,
                       int64_t wrap_hi, uint32_t* occ, int32_t* home, hipStream_t s,
                       bool coords_only) {
  const int n = p.sys.natoms;
  const int co = coords_only ? 1 : 0;
  if (n == 0) return hipSuccess;
  const dim3 grid((n + 255) / 256), block(256);
  if (dtype == MDSF_F32 && arith == MDSF_F32) {
    hipLaunchKernelGGL((k_prep<float, float>), grid, block, 0, s, (float*)coords, p.sys, p.types,
                       types, scale3, box[0], box[1], box[2], (long long)wrap_lo,
                       (long long)wrap_hi, p.ir4, p.key, p.rpos, p.bucket_count, occ, home, p.err, co);
  } else if (dtype == MDSF_F32) {
    hipLaunchKernelGGL((k_prep<float, double>), grid, block, 0, s, (float*)coords, p.sys, p.types,
                       types, scale3, box[0], box[1], box[2], (long long)wrap_lo,
                       (long long)wrap_hi, p.ir4, p.key, p.rpos, p.bucket_count, occ, home, p.err, co);
  } else {
    hipLaunchKernelGGL((k_prep<double, double>), grid, block, 0, s, (double*)coords, p.sys,
                       p.types, types, scale3, box[0], box[1], box[2], (long long)wrap_lo,
                       (long long)wrap_hi, p.ir4, p.key, p.rpos, p.bucket_count, occ, home, p.err, co);
  }
  return hipGetLastError();
}

hipError_t launch_binning(Plan& p, hipStream_t s) {
  const int n = p.sys.natoms;
  const int nbk = p.sys.nbuckets;
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, s, p.bucket_count, p.bucket_off,
                     p.bucket_cur, nbk);
  if (n > 0) {
    hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, s, p.key, p.bucket_cur,
                       p.s_idx, n);
    hipLaunchKernelGGL(k_bucket_sort, dim3((nbk + 255) / 256), dim3(256), 0, s, p.bucket_off,
                       p.s_idx, nbk);
    const int per_block = 256 / kTabThreads;
    const dim3 grid((n + per_block - 1) / per_block), block(256);
    if (p.fft64)
      hipLaunchKernelGGL(k_gather_tables<double>, grid, block, 0, s, p.sys, p.types, p.s_idx,
                         p.ir4, p.rpos, p.s_ir, p.s_r, static_cast<double*>(p.tables),
                         p.bucket_off + nbk);
    else
      hipLaunchKernelGGL(k_gather_tables<float>, grid, block, 0, s, p.sys, p.types, p.s_idx,
                         p.ir4, p.rpos, p.s_ir, p.s_r, static_cast<float*>(p.tables),
                         p.bucket_off + nbk);
  }
  return hipGetLastError();
}

}  // namespace mdsf
