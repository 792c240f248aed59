// C ABI of libmdsf: plan lifecycle and the per-frame pipeline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>

#include "mdsf_internal.h"
#include "spill.h"

namespace mdsf {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(MDSF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

// Bucket edge (cells) of a type's binning grid: wide enough that a tile's
// candidate columns stay few, small enough that a bucket holds few atoms
// (each bucket is sorted by one thread to fix the gather order).
int bucket_edge(int half) { return half < 24 ? 16 : 32; }

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

// The frame pipeline's three streams (deposition, FFT, and the tile-list
// branch of the prepare stage), one set per device shared by every plan and
// kept until the process ends.  HIP gives a process few hardware queues
// (GPU_MAX_HW_QUEUES, 4 by default: these three and the caller's) and hands
// them to streams in turn, so private streams per plan would soon share
// queues, and a deposition could no longer overlap an FFT queued behind it.
// Each frame loop starts from and joins back into the caller's stream through
// its own plan's events, so plans sharing the streams stay correctly ordered.
struct PipelineStreams {
  hipStream_t dep = nullptr, fft = nullptr, aux = nullptr;
};
std::mutex g_stream_mu;
std::map<int, PipelineStreams> g_streams;

hipError_t pipeline_streams(int device, hipStream_t* dep, hipStream_t* fft, hipStream_t* aux) {
  std::lock_guard<std::mutex> lock(g_stream_mu);
  auto it = g_streams.find(device);
  if (it == g_streams.end()) {
    PipelineStreams ps;
    hipError_t e = hipStreamCreateWithFlags(&ps.dep, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ps.fft, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ps.aux, hipStreamNonBlocking);
    if (e != hipSuccess) {
      for (hipStream_t q : {ps.dep, ps.fft, ps.aux})
        if (q) (void)hipStreamDestroy(q);
      return e;
    }
    it = g_streams.emplace(device, ps).first;
  }
  *dep = it->second.dep;
  *fft = it->second.fft;
  *aux = it->second.aux;
  return hipSuccess;
}

void free_plan(Plan* p) {
  if (!p) return;
  for (hipEvent_t e : {p->ev_fork, p->ev_lfork, p->ev_lgath, p->ev_ljoin})
    if (e) (void)hipEventDestroy(e);
  for (int b = 0; b < kSlots; ++b)
    for (hipEvent_t e : {p->ev_prep[b], p->ev_dep[b], p->ev_fft[b]})
      if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : p->tev) (void)hipEventDestroy(e);
  for (FrameState& f : p->slot) {
    void* fs[] = {f.ir4,   f.key,   f.rpos, f.bucket_count, f.bucket_off, f.bucket_cur,
                  f.bucket_aux, f.s_tmp, f.s_idx, f.s_ir, f.s_r, f.tables, f.slices,
                  f.dslices, f.bslices, f.corners, f.spill, f.spill_flag, f.tl_pool,
                  f.tl_head};
    for (void* q : fs)
      if (q) (void)hipFree(q);
  }
  for (int b = 0; b < kSlots; ++b)
    for (void* q : {p->rho_buf[b], p->nyq_buf[b], p->spec_buf[b], p->res_buf[b]})
      if (q) (void)hipFree(q);
  void* ptrs[] = {p->types,     p->err,       p->sf_main,
                  p->sf_nyq,    p->ax_x.tw,   p->ax_y.tw,    p->ax_zc.tw,   p->ax_x.tw64,
                  p->ax_y.tw64, p->ax_zc.tw64, p->tw_r2c,    p->tw_r2c64,   p->adj_part,
                  p->adj_w,     p->adj_tpart};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete p;
}

int make_axis(FftAxis& ax, int n, hipStream_t s) {
  ax.n = n;
  if (!fft_factor(n, ax.radix, &ax.nstages))
    return fail(MDSF_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " needs too many stages");
  MDSF_HIP(dalloc(&ax.tw, (size_t)n));
  MDSF_HIP(dalloc(&ax.tw64, (size_t)n));
  MDSF_HIP(fft_make_twiddles(ax.tw, ax.tw64, n, s));
  return MDSF_OK;
}

int ensure_adjoint(Plan& p) {
  const size_t nxy = (size_t)p.sys.n[0] * p.sys.n[1];
  const size_t m = (size_t)p.sys.n[2] / 2;
  if (!p.adj_part) MDSF_HIP(dalloc(&p.adj_part, 3 * (size_t)std::max(p.adj_items, 1LL)));
  if (!p.adj_w) MDSF_HIP(dalloc(&p.adj_w, nxy * (m + 1)));
  if (p.adj_tslots > 0 && !p.adj_tpart) {
    // tiled adjoint: slice derivatives per frame slot and the result slots
    for (FrameState& f : p.slot)
      if (!f.dslices) MDSF_HIP(dalloc(&f.dslices, (size_t)p.slice_floats));
    MDSF_HIP(dalloc(&p.adj_tpart, 3 * (size_t)p.adj_tslots));
    p.use_slot(p.cur_slot);
  }
  return MDSF_OK;
}

constexpr const char* kNoGenAdjoint =
    "the adjoint supports orthogonal and xy-coupled cells (no ucell coupling z with x or y)";

int run_fft_accumulate(Plan& p, const void* density, hipStream_t s) {
  MDSF_HIP(fft_forward(p, density, true, s));
  return MDSF_OK;
}

int check_frame_args(Plan* p, void* coords, int dtype, const int32_t* types, const double* scale,
                     const double* box, int arith) {
  if (!p) return fail(MDSF_ERR_ARG, "null plan");
  if (p->sys.natoms > 0 && (!coords || !types || !scale))
    return fail(MDSF_ERR_ARG, "null coordinate / type / scale pointer");
  if (!box) return fail(MDSF_ERR_ARG, "null box");
  if (dtype != MDSF_F32 && dtype != MDSF_F64) return fail(MDSF_ERR_ARG, "dtype must be F32 or F64");
  if (arith != MDSF_F32 && arith != MDSF_F64) return fail(MDSF_ERR_ARG, "arith must be F32 or F64");
  if (dtype == MDSF_F64 && arith == MDSF_F32)
    return fail(MDSF_ERR_ARG, "float64 coordinates are rescaled in float64");
  return MDSF_OK;
}

}  // namespace
}  // namespace mdsf

using namespace mdsf;

extern "C" {

const char* mdsf_last_error(void) { return g_last_error.c_str(); }

int mdsf_version(void) { return 10000; }

int mdsf_plan_create(const mdsf_grid_desc* desc, mdsf_plan** out) {
  if (!desc || !out) return fail(MDSF_ERR_ARG, "null descriptor or output");
  *out = nullptr;
  for (int d = 0; d < 3; ++d) {
    const int n = desc->n[d];
    if (n < 2 || (n % 2) != 0) return fail(MDSF_ERR_ARG, "grid sizes must be even and >= 2");
    if (n > 4096) return fail(MDSF_ERR_UNSUPPORTED, "grid size above 4096");
    if (desc->nborder > n)
      return fail(MDSF_ERR_ARG, "Nborder exceeds the grid (the reference's fold needs Nborder <= N)");
    if (!(desc->dr[d] > 0.0)) return fail(MDSF_ERR_ARG, "dr must be positive");
  }
  if (desc->nborder < 0) return fail(MDSF_ERR_ARG, "negative Nborder");
  if (desc->ntypes < 0 || desc->ntypes > kMaxTypes)
    return fail(MDSF_ERR_ARG, "between 0 and 64 atom types are supported");
  if (desc->natoms < 0) return fail(MDSF_ERR_ARG, "negative atom count");
  if (desc->ntypes > 0 && !desc->types) return fail(MDSF_ERR_ARG, "null type table");

  const double* U = desc->ucell;
  double G[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) G[i][j] = U[3 * i + 0] * U[3 * j + 0] + U[3 * i + 1] * U[3 * j + 1] + U[3 * i + 2] * U[3 * j + 2];

  Plan* p = new Plan();
  SysDev& s = p->sys;
  for (int d = 0; d < 3; ++d) {
    s.n[d] = desc->n[d];
    s.dr[d] = desc->dr[d];
    s.ntiles[d] = (desc->n[d] + kTile - 1) / kTile;
  }
  s.nb = desc->nborder;
  p->spill_tiles = 1;
  for (int d = 0; d < 3; ++d) {
    spill_axis(s.n[d], s.nb, &s.sp_k0[d], &s.sp_fh[d], &s.sp_cnt[d]);
    p->spill_tiles *= s.sp_cnt[d];
  }
  s.g00 = G[0][0];
  s.g01 = G[0][1];
  s.g11 = G[1][1];
  s.g22 = G[2][2];
  s.g02 = G[0][2];
  s.g12 = G[1][2];
  // z coupled to x or y: the general gather (dens.py:301 with any 3x3 ucell)
  s.mode = (G[0][2] != 0.0 || G[1][2] != 0.0) ? MODE_GEN
                                               : ((G[0][1] == 0.0) ? MODE_DIAG : MODE_XY);
  s.ntypes = desc->ntypes;
  s.natoms = desc->natoms;
  p->fft64 = (desc->flags & MDSF_FFT_FP64) ? 1 : 0;

  long long table_floats = 0, slice_floats = 0;
  int start = 0, bucket_base = 0;
  long long counted = 0;
  for (int t = 0; t < desc->ntypes; ++t) {
    const mdsf_type_desc& in = desc->types[t];
    TypeDev& ty = p->types_host[t];
    ty.amp = in.amp;
    ty.two_sig2 = in.two_sig2;
    ty.amp_f = (float)in.amp;
    ty.c2 = (float)(1.4426950408889634 / in.two_sig2);
    if (!(in.two_sig2 > 0.0)) {
      free_plan(p);
      return fail(MDSF_ERR_ARG, "sigma must be positive");
    }
    int stride = 0;
    for (int d = 0; d < 3; ++d) {
      if (in.half[d] < 0 || in.half[d] > desc->nborder) {
        free_plan(p);
        return fail(MDSF_ERR_ARG, "type window half-width outside [0, Nborder]");
      }
      ty.half[d] = in.half[d];
      stride += 2 * in.half[d];
      ty.bsz[d] = bucket_edge(in.half[d]);
      ty.nbk[d] = (desc->n[d] + ty.bsz[d] - 1) / ty.bsz[d];
    }
    // tile slices: a cyclic window of 2A cells covers at most ceil(2A/16)+1
    // tiles, one more when the last tile along the axis is partial
    int sstride = 0;
    for (int d = 0; d < 3; ++d) {
      const int nt = (desc->n[d] + kTile - 1) / kTile;
      const int extra = (desc->n[d] % kTile) ? 2 : 1;
      ty.nsl[d] = std::min(nt, (2 * in.half[d] + kTile - 1) / kTile + extra);
      sstride += kTile * ty.nsl[d];
    }
    ty.slice_stride = sstride;
    ty.slice_base = slice_floats;
    slice_floats += (long long)in.count * sstride;
    ty.corner_stride = kTile * (2 * ty.nsl[0] + 3 * ty.nsl[1] + 2 * ty.nsl[2]);
    ty.count = in.count;
    ty.type_start = start;
    ty.stride = stride;
    ty.table_base = table_floats;
    ty.bucket_base = bucket_base;
    ty.park = bucket_base + ty.nbk[0] * ty.nbk[1] * ty.nbk[2];
    table_floats += (long long)in.count * stride;
    start += in.count;
    counted += in.count;
    bucket_base = ty.park + 1;
  }
  if (counted != desc->natoms) {
    free_plan(p);
    return fail(MDSF_ERR_ARG, "type counts do not add up to natoms");
  }
  s.nbuckets = bucket_base;
  p->table_floats = table_floats;
  const long long regular_floats = slice_floats;
  for (int t = 0; t < desc->ntypes; ++t) {  // corner components follow the regular slices
    TypeDev& ty = p->types_host[t];
    ty.corner_base = slice_floats;
    slice_floats += (long long)ty.count * ty.corner_stride;
  }
  // FP32 xy-coupled cells deposit from the tile slices and lists too when
  // every type's window meets each x and each y tile through one periodic
  // image (2A <= N - 16: two images in one tile need 2A >= N - 14), so that
  // an atom's factor over a tile is X(x) Y(y) 2^(c b_x b_y), and when the
  // cross exponent |c b_x b_y| <= c (A_x + 1) dr_x (A_y + 1) dr_y stays
  // within 100 (2^(...) finite in FP32); otherwise they keep the enumerating
  // kernel, which sums the images of each cell pair.
  bool xy_lists = !p->fft64 && s.mode == MODE_XY && slice_floats < (1LL << 31);
  for (int t = 0; t < desc->ntypes && xy_lists; ++t) {
    const TypeDev& ty = p->types_host[t];
    for (int d = 0; d < 2; ++d)
      if (2 * ty.half[d] > desc->n[d] - kTile) xy_lists = false;
    const double c = 2.0 * std::fabs(s.g01) * 1.4426950408889634 / ty.two_sig2;
    if (c * (ty.half[0] + 1) * s.dr[0] * (ty.half[1] + 1) * s.dr[1] > 100.0) xy_lists = false;
  }
  p->slice_floats = (!p->fft64 && (s.mode == MODE_DIAG || xy_lists)) ? slice_floats : 0;
  p->bslice_floats = xy_lists ? regular_floats : 0;
  if (p->slice_floats >= (1LL << 31)) {  // the deposition addresses slices with 32-bit offsets
    free_plan(p);
    return fail(MDSF_ERR_UNSUPPORTED, "too many atoms for the FP32 tile-slice layout (2^31 floats)");
  }
  // Tile lists: a tile's pool region holds its candidates (atoms of the
  // bucket columns that can reach it), so the pool is sized by the most
  // tiles whose candidate range includes one bucket: per axis the tiles whose
  // cells within A - 1 of the tile (t0 - A + 1 .. t1 + A - 1) meet the B
  // cells of the bucket, at most (B + 2A + 13) / 16 + 1 of them, plus one
  // for a partial last tile or bucket at the periodic seam, and at most all
  // tiles.  Entries carry slice indices and the type in bytes.
  p->tl_cap = 0;
  if (p->slice_floats > 0) {
    long long cap = 0;
    bool ok = desc->ntypes <= kMaxTypes;
    for (int t = 0; t < desc->ntypes && ok; ++t) {
      const TypeDev& ty = p->types_host[t];
      long long per_atom = 1;
      for (int d = 0; d < 3; ++d) {
        if (ty.nsl[d] >= kListNoSlice) ok = false;
        const long long nt = (desc->n[d] + kTile - 1) / kTile;
        per_atom *= std::min(nt, (long long)(ty.bsz[d] + 2 * ty.half[d] + 13) / kTile + 2);
      }
      cap += (long long)ty.count * per_atom;
    }
    if (ok && cap < (1LL << 31)) p->tl_cap = std::max(cap, 1LL);
  }
  {  // tiled adjoint result slots: per atom one per tile slice triple
    long long slots = 0;
    for (int t = 0; t < desc->ntypes; ++t) {
      TypeDev& ty = p->types_host[t];
      ty.adjt_base = slots;
      slots += (long long)ty.count * ty.nsl[0] * ty.nsl[1] * ty.nsl[2];
    }
    p->adj_tslots = (p->slice_floats > 0 && s.mode == MODE_DIAG) ? slots : 0;
  }
  // adjoint work items (atom, x-chunk of <= 64 planes, chunk of (y, z)
  // pairs sized for ~16k window voxels), heaviest windows numbered first so
  // that their workgroups start first
  {
    int order[kMaxTypes];
    for (int t = 0; t < desc->ntypes; ++t) order[t] = t;
    auto vol = [&](int t) {
      const TypeDev& y = p->types_host[t];
      return (double)y.half[0] * y.half[1] * y.half[2];
    };
    std::stable_sort(order, order + desc->ntypes, [&](int a, int b) { return vol(a) > vol(b); });
    long long items = 0;
    for (int k = 0; k < desc->ntypes; ++k) {
      TypeDev& y = p->types_host[order[k]];
      const int wx = 2 * y.half[0];
      const long long npairs = 4LL * y.half[1] * y.half[2];
      y.adj_item_base = items;
      y.adj_ppi = 64;
      y.adj_nxc = y.adj_npc = 0;
      if (wx == 0 || npairs == 0) continue;
      const int xw = std::min(64, wx);
      y.adj_nxc = (wx + 63) / 64;
      y.adj_ppi = 64 * ((256 + xw - 1) / xw);
      y.adj_npc = (int)((npairs + y.adj_ppi - 1) / y.adj_ppi);
      items += (long long)y.count * y.adj_nxc * y.adj_npc;
    }
    p->adj_items = items;
  }
  (void)hipGetDevice(&p->device);

  const size_t nat = (size_t)desc->natoms;
  const size_t nx = desc->n[0], ny = desc->n[1], nz = desc->n[2];
  const size_t m = nz / 2;
  hipError_t e = hipSuccess;
#define TRY(x)                          \
  do {                                  \
    e = (x);                            \
    if (e != hipSuccess) {              \
      int rc = hip_fail(e, #x);         \
      free_plan(p);                     \
      return rc;                        \
    }                                   \
  } while (0)
  TRY(dalloc(&p->types, (size_t)kMaxTypes));
  TRY(hipMemcpy(p->types, p->types_host, sizeof(TypeDev) * kMaxTypes, hipMemcpyHostToDevice));
  for (FrameState& f : p->slot) {
    TRY(dalloc(&f.ir4, nat));
    TRY(dalloc(&f.key, nat));
    TRY(dalloc(&f.rpos, 3 * nat));
    TRY(dalloc(&f.bucket_count, (size_t)s.nbuckets + 1));
    TRY(dalloc(&f.bucket_off, (size_t)s.nbuckets + 1));
    TRY(dalloc(&f.bucket_cur, (size_t)s.nbuckets + 1));
    TRY(dalloc(&f.bucket_aux, (size_t)s.nbuckets / 1024 + 2));
    TRY(hipMemset(f.bucket_count, 0, sizeof(int) * ((size_t)s.nbuckets + 1)));
    TRY(hipMemset(f.bucket_off, 0, sizeof(int) * ((size_t)s.nbuckets + 1)));
    TRY(dalloc(&f.s_idx, nat));
    TRY(dalloc(&f.s_tmp, nat));
    TRY(dalloc(&f.s_ir, nat));
    TRY(dalloc(&f.s_r, 3 * nat));
    if (p->fft64) {
      double* tb = nullptr;
      TRY(dalloc(&tb, (size_t)table_floats));
      f.tables = tb;
    } else {
      float* tb = nullptr;  // unused: the FP32 path stages tile slices instead
      TRY(dalloc(&tb, (size_t)1));
      f.tables = tb;
    }
    TRY(dalloc(&f.slices, (size_t)p->slice_floats));
    if (p->bslice_floats > 0) TRY(dalloc(&f.bslices, (size_t)p->bslice_floats));
    TRY(dalloc(&f.corners, (size_t)2 + 2 * kCornerCap));
    TRY(hipMemset(f.corners, 0, sizeof(int) * ((size_t)2 + 2 * kCornerCap)));
    if (p->spill_tiles > 0 && p->spill_tiles <= kMaxSpillBufTiles) {
      unsigned char* sb = nullptr;
      TRY(dalloc(&sb, (size_t)p->spill_tiles * kTile * kTile * kTile * (p->fft64 ? 8 : 4)));
      f.spill = sb;
      TRY(dalloc(&f.spill_flag, (size_t)p->spill_tiles));
      // no tile is read before a frame's binning flags it (a frame without atoms
      // never runs the binning)
      TRY(hipMemset(f.spill_flag, 0, sizeof(int) * (size_t)p->spill_tiles));
    }
    if (p->tl_cap > 0) {
      TRY(dalloc(&f.tl_pool, (size_t)p->tl_cap));
      TRY(dalloc(&f.tl_head, (size_t)s.ntiles[0] * s.ntiles[1] * s.ntiles[2]));
    }
  }
  TRY(dalloc(&p->err, (size_t)1));
  TRY(hipMemset(p->err, 0, sizeof(int)));
  for (int b = 0; b < kSlots; ++b) {
    unsigned char* rho = nullptr;
    unsigned char* spec = nullptr;
    double2* nyq = nullptr;
    TRY(dalloc(&rho, nx * ny * nz * (p->fft64 ? 8 : 4)));
    TRY(dalloc(&spec, fft_spec_bytes(s.n, p->fft64)));
    TRY(dalloc(&nyq, nx * ny));
    p->rho_buf[b] = rho;
    p->spec_buf[b] = spec;
    p->nyq_buf[b] = nyq;
    if (fft_res_bytes(s.n, p->fft64) > 0) {
      unsigned char* res = nullptr;
      TRY(dalloc(&res, fft_res_bytes(s.n, p->fft64)));
      p->res_buf[b] = res;
    }
  }
  p->use_slot(0);
  TRY(dalloc(&p->sf_main, nx * ny * m));
  TRY(dalloc(&p->sf_nyq, nx * ny));
  TRY(hipMemset(p->sf_main, 0, sizeof(double) * nx * ny * m));
  TRY(hipMemset(p->sf_nyq, 0, sizeof(double) * nx * ny));
#undef TRY
  int rc;
  if ((rc = make_axis(p->ax_x, desc->n[0], nullptr)) != MDSF_OK ||
      (rc = make_axis(p->ax_y, desc->n[1], nullptr)) != MDSF_OK ||
      (rc = make_axis(p->ax_zc, desc->n[2] / 2, nullptr)) != MDSF_OK) {
    free_plan(p);
    return rc;
  }
  if ((e = dalloc(&p->tw_r2c, nz)) != hipSuccess || (e = dalloc(&p->tw_r2c64, nz)) != hipSuccess ||
      (e = fft_make_twiddles(p->tw_r2c, p->tw_r2c64, (int)nz, nullptr)) != hipSuccess) {
    rc = hip_fail(e, "r2c twiddles");
    free_plan(p);
    return rc;
  }
  if ((e = hipDeviceSynchronize()) != hipSuccess) {
    rc = hip_fail(e, "plan set-up");
    free_plan(p);
    return rc;
  }
  e = pipeline_streams(p->device, &p->dep_stream, &p->fft_stream, &p->aux_stream);
  for (hipEvent_t* ev : {&p->ev_fork, &p->ev_lfork, &p->ev_lgath, &p->ev_ljoin})
    if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
  for (int b = 0; b < kSlots && e == hipSuccess; ++b)
    for (hipEvent_t* ev : {&p->ev_prep[b], &p->ev_dep[b], &p->ev_fft[b]})
      if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    rc = hip_fail(e, "frame pipeline stream / events");
    free_plan(p);
    return rc;
  }
  if (desc->flags & MDSF_AUTOTUNE) {
    if ((e = fft_autotune(*p, nullptr)) != hipSuccess ||
        (e = hipMemset(p->sf_main, 0, sizeof(double) * nx * ny * m)) != hipSuccess ||
        (e = hipMemset(p->sf_nyq, 0, sizeof(double) * nx * ny)) != hipSuccess ||
        (e = hipDeviceSynchronize()) != hipSuccess) {
      rc = hip_fail(e, "FFT autotuning");
      free_plan(p);
      return rc;
    }
  }
  *out = reinterpret_cast<mdsf_plan*>(p);
  return MDSF_OK;
}

int mdsf_plan_destroy(mdsf_plan* plan) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p) return MDSF_OK;
  (void)hipDeviceSynchronize();
  free_plan(p);
  return MDSF_OK;
}

int mdsf_sf_reset(mdsf_plan* plan, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p) return fail(MDSF_ERR_ARG, "null plan");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const long long nxy = (long long)p->sys.n[0] * p->sys.n[1];
  // a fill kernel rather than a memset node keeps this capturable in graphs
  MDSF_HIP(launch_fill_zero(p->sf_main, nxy * (p->sys.n[2] / 2), s));
  MDSF_HIP(launch_fill_zero(p->sf_nyq, nxy, s));
  return MDSF_OK;
}

int mdsf_check_errors(mdsf_plan* plan, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p) return fail(MDSF_ERR_ARG, "null plan");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  int flag = 0;
  MDSF_HIP(hipMemcpyAsync(&flag, p->err, sizeof(int), hipMemcpyDeviceToHost, s));
  MDSF_HIP(hipStreamSynchronize(s));
  if (flag) MDSF_HIP(hipMemsetAsync(p->err, 0, sizeof(int), s));
  if (flag & ERRBIT_TYPE) return fail(MDSF_ERR_TYPE, "atom type index out of range");
  if (flag & ERRBIT_LIST)
    return fail(MDSF_ERR_UNSUPPORTED, "tile list pool overflow (its size bound was exceeded)");
  if (flag & ERRBIT_RANGE)
    return fail(MDSF_ERR_RANGE,
                "an atom's window leaves the padded grid (atom more than the border outside the "
                "cell)");
  return MDSF_OK;
}

int mdsf_accumulate_frames(mdsf_plan* plan, void* d_coords, int dtype, int64_t nframes,
                           const int32_t* d_types, const double* d_scale, const double box[3],
                           int arith, int64_t wrap_lo, int64_t wrap_hi, uint32_t* d_occupancy,
                           int32_t* d_home_cells, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  int rc = check_frame_args(p, d_coords, dtype, d_types, d_scale, box, arith);
  if (rc != MDSF_OK) return rc;
  if (nframes < 0) return fail(MDSF_ERR_ARG, "negative frame count");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (nframes == 0) return MDSF_OK;
  const size_t esz = dtype == MDSF_F32 ? 4 : 8;
  const size_t fstride = (size_t)p->sys.natoms * 3 * esz;
  // Three-stage pipeline over the kSlots frame slots (atom state + density
  // buffer): frame t is prepared on `s` (its tile lists on aux_stream, its
  // corner list and spill staging on dep_stream ahead of its deposition),
  // deposited on dep_stream and transformed on fft_stream, so that while
  // frame t is deposited frame t + 1 is prepared and earlier frames
  // transformed.  A slot is reused kSlots frames later: preparing frame t
  // waits until frame t - kSlots's deposition (tiled and parked passes) has
  // read its atom state, depositing it until frame t - kSlots's FFT has
  // released its density buffer.  With three slots the deposition does not
  // wait for the FFT of the frame just before the previous one (the FFT
  // shares the GPU with the deposition and finishes later than it).  The
  // S(q) sum stays in frame order on one stream.
  hipStream_t d = p->dep_stream, f = p->fft_stream;
  // the prepare stage's branches go on their own streams on large grids; on
  // small ones the cross-stream waits cost more than the branches overlap
  const long long ntiles = (long long)p->sys.ntiles[0] * p->sys.ntiles[1] * p->sys.ntiles[2];
  const bool split = ntiles >= kSplitPrepTiles;
  const bool pair = nframes >= 2 && fft_x_pairs(*p);
  // timed frames accumulate over calls until mdsf_stage_times reads them
  // (frame t of this call is timed frame tb + t), so that one-frame calls
  // can be queued back to back
  int64_t tb = 0;
  if (p->timing) {
    tb = p->timed_frames < kMaxTimedFrames ? p->timed_frames : 0;
    while ((int64_t)p->tev.size() < kTimedEvents * (tb + nframes)) {
      hipEvent_t e;
      MDSF_HIP(hipEventCreate(&e));
      p->tev.push_back(e);
    }
    p->timed_frames = tb + nframes;
  }
  // timing mode 1: every stage; mode 2: only the deposition (two events per
  // frame on its stream, the least perturbation of the overlapped loop)
  auto mark = [&](int64_t t, int k, hipStream_t st) -> hipError_t {
    if (!p->timing || (p->timing == 2 && k != 2 && k != 3)) return hipSuccess;
    return hipEventRecord(p->tev[kTimedEvents * (tb + t) + k], st);
  };
  MDSF_HIP(hipEventRecord(p->ev_fork, s));
  MDSF_HIP(hipStreamWaitEvent(d, p->ev_fork, 0));
  MDSF_HIP(hipStreamWaitEvent(f, p->ev_fork, 0));
  for (int b = 0; b < kSlots; ++b) {
    MDSF_HIP(hipEventRecord(p->ev_dep[b], d));
    MDSF_HIP(hipEventRecord(p->ev_fft[b], f));
  }
  for (int64_t t = 0; t < nframes; ++t) {
    const int b = (int)(t % kSlots);
    char* c = static_cast<char*>(d_coords) + (size_t)t * fstride;
    int32_t* home = (t == nframes - 1) ? d_home_cells : nullptr;
    p->use_slot(b);
    MDSF_HIP(hipStreamWaitEvent(s, p->ev_dep[b], 0));
    MDSF_HIP(mark(t, 0, s));
    MDSF_HIP(launch_prep(*p, c, dtype, arith, d_types, d_scale + 3 * t, box, wrap_lo, wrap_hi,
                         d_occupancy, home, s));
    // the deposition waits for the tile-list branch itself (one cross-stream
    // wait instead of two in a row through `s`, so the binning does not join
    // the branch back into `s`); under stage timing the branch records its
    // own end (event 10), and the prepare interval ends with the later of the
    // two
    const bool lsplit = split && p->tl_cap > 0 && p->sys.natoms > 0;  // lists on aux_stream
    hipEvent_t aux_end = (p->timing == 1 && lsplit) ? p->tev[kTimedEvents * (tb + t) + 10] : nullptr;
    MDSF_HIP(launch_binning(*p, s, false, split ? p->aux_stream : nullptr, split ? d : nullptr,
                            /*join=*/false, aux_end));
    MDSF_HIP(mark(t, 1, s));
    if (p->timing == 1 && !aux_end) MDSF_HIP(mark(t, 10, s));
    MDSF_HIP(hipEventRecord(p->ev_prep[b], s));
    MDSF_HIP(hipStreamWaitEvent(d, p->ev_prep[b], 0));
    if (lsplit) MDSF_HIP(hipStreamWaitEvent(d, p->ev_ljoin, 0));
    MDSF_HIP(hipStreamWaitEvent(d, p->ev_fft[b], 0));
    // the deposition's events are recorded with the tiled kernel's own
    // dispatch (its begin and end, without the queue's dispatch latency)
    if (p->timing) {
      p->dep_ev[0] = p->tev[kTimedEvents * (tb + t) + 2];
      p->dep_ev[1] = p->tev[kTimedEvents * (tb + t) + 3];
    }
    const hipError_t de = launch_deposit(*p, p->rho, d);
    p->dep_ev[0] = p->dep_ev[1] = nullptr;
    MDSF_HIP(de);
    // parked atoms' clipped windows: on the deposition's stream, before
    // ev_dep[b], because they read the slot's atom state, which preparing
    // frame t + kSlots overwrites as soon as ev_dep[b] has passed
    MDSF_HIP(launch_parked(*p, p->rho, d));
    MDSF_HIP(hipEventRecord(p->ev_dep[b], d));
    MDSF_HIP(hipStreamWaitEvent(f, p->ev_dep[b], 0));
    // x passes in pairs: frame t (even) waits for frame t + 1, and one launch
    // adds both frames' |F|^2, reading and writing the FP64 sums once; frame
    // t's slot is released with frame t + 1's (preparing and depositing frame
    // t + kSlots wait for it).  The last two frames of a longer call are not
    // paired: both x passes would wait for the last y pass and lengthen the
    // call's tail by about one x pass.
    auto defers = [&](int64_t u) {
      return pair && (u % 2) == 0 && (u + 2 < nframes || nframes == 2);
    };
    const bool defer = defers(t);
    const bool paired = t > 0 && defers(t - 1);
    for (int pass = 0; pass < 3; ++pass) {
      // with stage timing, the pass's stream kernel records the pass events
      // with its own dispatch (the marks stand when another kernel runs it);
      // a deferred x pass is timed as zero, its pair's launch as frame t + 1's
      MDSF_HIP(mark(t, 4 + 2 * pass, f));
      if (pass == 2 && defer) {
        MDSF_HIP(mark(t, 5 + 2 * pass, f));
        continue;
      }
      if (p->timing == 1)
        fft_pass_events(p->tev[kTimedEvents * (tb + t) + 4 + 2 * pass],
                        p->tev[kTimedEvents * (tb + t) + 5 + 2 * pass]);
      const hipError_t fe = (pass == 2 && paired)
                                ? fft_x_pair(*p, (int)((t - 1) % kSlots), b, f)
                                : fft_pass(*p, pass, p->rho, f);
      const bool took = fft_pass_events_used();
      fft_pass_events(nullptr, nullptr);
      MDSF_HIP(fe);
      if (!took) MDSF_HIP(mark(t, 5 + 2 * pass, f));
    }
    if (!defer) {
      if (paired) MDSF_HIP(hipEventRecord(p->ev_fft[(t - 1) % kSlots], f));
      MDSF_HIP(hipEventRecord(p->ev_fft[b], f));
    }
  }
  MDSF_HIP(hipStreamWaitEvent(s, p->ev_fft[(nframes - 1) % kSlots], 0));
  return MDSF_OK;
}

int mdsf_rescale_wrap(mdsf_plan* plan, void* d_coords, int dtype, int64_t nframes,
                      const double* d_scale, const double box[3], int arith, int64_t wrap_lo,
                      int64_t wrap_hi, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p) return fail(MDSF_ERR_ARG, "null plan");
  if (p->sys.natoms > 0 && (!d_coords || !d_scale)) return fail(MDSF_ERR_ARG, "null coordinates");
  if (!box) return fail(MDSF_ERR_ARG, "null box");
  if (dtype != MDSF_F32 && dtype != MDSF_F64) return fail(MDSF_ERR_ARG, "dtype must be F32 or F64");
  if (dtype == MDSF_F64 && arith == MDSF_F32)
    return fail(MDSF_ERR_ARG, "float64 coordinates are rescaled in float64");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const size_t esz = dtype == MDSF_F32 ? 4 : 8;
  const size_t fstride = (size_t)p->sys.natoms * 3 * esz;
  for (int64_t t = 0; t < nframes; ++t) {
    char* c = static_cast<char*>(d_coords) + (size_t)t * fstride;
    MDSF_HIP(launch_prep(*p, c, dtype, arith, nullptr, d_scale + 3 * t, box, wrap_lo, wrap_hi,
                         nullptr, nullptr, s, true));
  }
  return MDSF_OK;
}

int mdsf_deposit_fold(mdsf_plan* plan, void* d_coords, int dtype, const int32_t* d_types,
                      const double* d_scale, const double box[3], int arith, int64_t wrap_lo,
                      int64_t wrap_hi, void* d_density, int32_t* d_home_cells,
                      mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  int rc = check_frame_args(p, d_coords, dtype, d_types, d_scale, box, arith);
  if (rc != MDSF_OK) return rc;
  if (!d_density) return fail(MDSF_ERR_ARG, "null density");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  MDSF_HIP(launch_prep(*p, d_coords, dtype, arith, d_types, d_scale, box, wrap_lo, wrap_hi,
                       nullptr, d_home_cells, s));
  MDSF_HIP(launch_binning(*p, s));
  MDSF_HIP(launch_deposit(*p, d_density, s));
  MDSF_HIP(launch_parked(*p, d_density, s));
  return MDSF_OK;
}

int mdsf_prepare_frame(mdsf_plan* plan, void* d_coords, int dtype, const int32_t* d_types,
                       const double* d_scale, const double box[3], int arith, int64_t wrap_lo,
                       int64_t wrap_hi, uint32_t* d_occupancy, int32_t* d_home_cells,
                       mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  int rc = check_frame_args(p, d_coords, dtype, d_types, d_scale, box, arith);
  if (rc != MDSF_OK) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  MDSF_HIP(launch_prep(*p, d_coords, dtype, arith, d_types, d_scale, box, wrap_lo, wrap_hi,
                       d_occupancy, d_home_cells, s));
  MDSF_HIP(launch_binning(*p, s));
  return MDSF_OK;
}

int mdsf_set_stage_timing(mdsf_plan* plan, int32_t enable) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p) return fail(MDSF_ERR_ARG, "null plan");
  if (enable < 0 || enable > 2) return fail(MDSF_ERR_ARG, "stage timing mode must be 0, 1 or 2");
  p->timing = enable;
  p->timed_frames = 0;
  return MDSF_OK;
}

int mdsf_stage_times(mdsf_plan* plan, double* avg_ms) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p || !avg_ms) return fail(MDSF_ERR_ARG, "null argument");
  const int64_t n = p->timed_frames;
  for (int k = 0; k < 5; ++k) avg_ms[k] = 0.0;
  if (n == 0) return fail(MDSF_ERR_ARG, "no timed mdsf_accumulate_frames call");
  const bool all = p->timing != 2;
  if (all) MDSF_HIP(hipEventSynchronize(p->tev[kTimedEvents * (n - 1) + 9]));
  MDSF_HIP(hipEventSynchronize(p->tev[kTimedEvents * (n - 1) + 3]));
  for (int64_t t = 0; t < n; ++t) {
    for (int k = 0; k < 5; ++k) {
      if (!all && k != 1) continue;
      float ms = 0.f;
      MDSF_HIP(hipEventElapsedTime(&ms, p->tev[kTimedEvents * t + 2 * k],
                                   p->tev[kTimedEvents * t + 2 * k + 1]));
      if (k == 0) {
        // the prepare stage ends with the later of its caller-stream part and
        // its tile-list branch
        float ms_aux = 0.f;
        MDSF_HIP(hipEventSynchronize(p->tev[kTimedEvents * t + 10]));
        MDSF_HIP(hipEventElapsedTime(&ms_aux, p->tev[kTimedEvents * t], p->tev[kTimedEvents * t + 10]));
        ms = ms_aux > ms ? ms_aux : ms;
      }
      avg_ms[k] += ms;
    }
  }
  for (int k = 0; k < 5; ++k) avg_ms[k] /= (double)n;
  p->timed_frames = 0;
  return MDSF_OK;
}

int mdsf_set_frame_buffer(mdsf_plan* plan, int32_t which) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p) return fail(MDSF_ERR_ARG, "null plan");
  if (which < 0 || which >= kSlots)
    return fail(MDSF_ERR_ARG, "frame buffer must be 0, 1 or 2");
  p->use_slot(which);
  return MDSF_OK;
}

int mdsf_deposit(mdsf_plan* plan, void* d_density, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p) return fail(MDSF_ERR_ARG, "null plan");
  void* out = d_density ? d_density : p->rho;
  MDSF_HIP(launch_deposit(*p, out, reinterpret_cast<hipStream_t>(stream)));
  MDSF_HIP(launch_parked(*p, out, reinterpret_cast<hipStream_t>(stream)));
  return MDSF_OK;
}

int mdsf_fft_pass(mdsf_plan* plan, int32_t pass, const void* d_density, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p) return fail(MDSF_ERR_ARG, "null plan");
  if (pass < 0 || pass > 2) return fail(MDSF_ERR_ARG, "pass must be 0 (z), 1 (y) or 2 (x)");
  MDSF_HIP(fft_pass(*p, pass, d_density ? d_density : p->rho,
                    reinterpret_cast<hipStream_t>(stream)));
  return MDSF_OK;
}

int mdsf_accumulate_density(mdsf_plan* plan, const void* d_density, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p || !d_density) return fail(MDSF_ERR_ARG, "null plan or density");
  return run_fft_accumulate(*p, d_density, reinterpret_cast<hipStream_t>(stream));
}

int mdsf_rfft3(mdsf_plan* plan, const void* d_density, double* d_spectrum, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p || !d_density || !d_spectrum) return fail(MDSF_ERR_ARG, "null argument");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  MDSF_HIP(fft_forward(*p, d_density, false, s));
  MDSF_HIP(fft_spectrum_read(*p, d_spectrum, s));
  return MDSF_OK;
}

int mdsf_density_adjoint(mdsf_plan* plan, const void* d_density, const double* d_dsf,
                         void* d_grad_density, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p || !d_density || !d_dsf || !d_grad_density) return fail(MDSF_ERR_ARG, "null argument");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  int rc = ensure_adjoint(*p);
  if (rc != MDSF_OK) return rc;
  const long long nxy = (long long)p->sys.n[0] * p->sys.n[1];
  MDSF_HIP(launch_adjoint_weights(*p, d_dsf, s));
  MDSF_HIP(fft_forward(*p, d_density, false, s));
  MDSF_HIP(fft_inverse(*p, p->adj_w, p->adj_w + nxy * (p->sys.n[2] / 2), d_grad_density, s));
  return MDSF_OK;
}

int mdsf_deposit_adjoint(mdsf_plan* plan, const void* d_grad_density, double* d_grad,
                         mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p || !d_grad_density) return fail(MDSF_ERR_ARG, "null plan or density gradient");
  if (p->sys.natoms > 0 && !d_grad) return fail(MDSF_ERR_ARG, "null coordinate gradient");
  if (p->sys.mode == MODE_GEN) return fail(MDSF_ERR_UNSUPPORTED, kNoGenAdjoint);
  int rc = ensure_adjoint(*p);
  if (rc != MDSF_OK) return rc;
  MDSF_HIP(launch_deposit_adjoint(*p, d_grad_density, nullptr, d_grad,
                                  reinterpret_cast<hipStream_t>(stream)));
  return MDSF_OK;
}

int mdsf_adjoint_frames(mdsf_plan* plan, void* d_coords, int dtype, int64_t nframes,
                        const int32_t* d_types, const double* d_scale, const double box[3],
                        int arith, int64_t wrap_lo, int64_t wrap_hi, const double* d_dsf,
                        double* d_grad, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  int rc = check_frame_args(p, d_coords, dtype, d_types, d_scale, box, arith);
  if (rc != MDSF_OK) return rc;
  if (nframes < 0) return fail(MDSF_ERR_ARG, "negative frame count");
  if (!d_dsf) return fail(MDSF_ERR_ARG, "null dL/dS");
  if (p->sys.natoms > 0 && nframes > 0 && !d_grad) return fail(MDSF_ERR_ARG, "null gradient");
  if (p->sys.mode == MODE_GEN) return fail(MDSF_ERR_UNSUPPORTED, kNoGenAdjoint);
  if ((rc = ensure_adjoint(*p)) != MDSF_OK) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const long long nxy = (long long)p->sys.n[0] * p->sys.n[1];
  const double* w_nyq = p->adj_w + nxy * (p->sys.n[2] / 2);
  MDSF_HIP(launch_adjoint_weights(*p, d_dsf, s));
  if (nframes == 0) return MDSF_OK;
  const size_t esz = dtype == MDSF_F32 ? 4 : 8;
  const size_t fstride = (size_t)p->sys.natoms * 3 * esz;
  // Same slot pipeline as the forward frame loop: frame t is prepared on
  // `s`, deposited on dep_stream, and transformed forward, transformed back
  // with the weights and gathered on fft_stream.  Slot b is reused by frame
  // t + kSlots only after frame t's gather (the last reader of its atom state
  // and density buffer) has finished.
  hipStream_t d = p->dep_stream, f = p->fft_stream;
  MDSF_HIP(hipEventRecord(p->ev_fork, s));
  MDSF_HIP(hipStreamWaitEvent(d, p->ev_fork, 0));
  MDSF_HIP(hipStreamWaitEvent(f, p->ev_fork, 0));
  for (int b = 0; b < kSlots; ++b) MDSF_HIP(hipEventRecord(p->ev_fft[b], f));
  for (int64_t t = 0; t < nframes; ++t) {
    const int b = (int)(t % kSlots);
    char* c = static_cast<char*>(d_coords) + (size_t)t * fstride;
    double* grad_t = d_grad + (size_t)t * p->sys.natoms * 3;
    p->use_slot(b);
    MDSF_HIP(hipStreamWaitEvent(s, p->ev_fft[b], 0));
    MDSF_HIP(launch_fill_zero(grad_t, 3LL * p->sys.natoms, s));  // off the gather's stream
    MDSF_HIP(launch_prep(*p, c, dtype, arith, d_types, d_scale + 3 * t, box, wrap_lo, wrap_hi,
                         nullptr, nullptr, s));
    MDSF_HIP(launch_binning(*p, s, /*derivs=*/true));
    MDSF_HIP(hipEventRecord(p->ev_prep[b], s));
    MDSF_HIP(hipStreamWaitEvent(d, p->ev_prep[b], 0));
    MDSF_HIP(launch_deposit(*p, p->rho, d));
    MDSF_HIP(hipEventRecord(p->ev_dep[b], d));
    MDSF_HIP(hipStreamWaitEvent(f, p->ev_dep[b], 0));
    MDSF_HIP(launch_parked(*p, p->rho, f));
    MDSF_HIP(fft_forward(*p, p->rho, false, f));
    MDSF_HIP(fft_inverse(*p, p->adj_w, w_nyq, p->rho, f));
    MDSF_HIP(launch_deposit_adjoint(*p, p->rho, d_scale + 3 * t, grad_t, f,
                                    /*derivs_ready=*/true, /*grad_zeroed=*/true));
    MDSF_HIP(hipEventRecord(p->ev_fft[b], f));
  }
  MDSF_HIP(hipStreamWaitEvent(s, p->ev_fft[(nframes - 1) % kSlots], 0));
  return MDSF_OK;
}

int mdsf_sf_read(mdsf_plan* plan, double* d_sf, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p || !d_sf) return fail(MDSF_ERR_ARG, "null argument");
  MDSF_HIP(launch_sf_read(*p, d_sf, reinterpret_cast<hipStream_t>(stream)));
  return MDSF_OK;
}

int mdsf_outputs(mdsf_plan* plan, const double* d_sf, const double box[3], int box_f32,
                 double* d_sfplt, double* d_kgrid, double* d_kgridplt, mdsf_stream stream) {
  Plan* p = reinterpret_cast<Plan*>(plan);
  if (!p || !d_sf || !box) return fail(MDSF_ERR_ARG, "null argument");
  MDSF_HIP(launch_outputs(p->sys.n, d_sf, box, box_f32, d_sfplt, d_kgrid, d_kgridplt,
                          reinterpret_cast<hipStream_t>(stream)));
  return MDSF_OK;
}

int mdsf_rescale(void* d_coords, int dtype, int64_t nframes, int64_t natoms,
                 const double* d_scale, int arith, mdsf_stream stream) {
  if (nframes < 0 || natoms < 0) return fail(MDSF_ERR_ARG, "negative frame or atom count");
  if (natoms > 2147483647LL) return fail(MDSF_ERR_UNSUPPORTED, "more than 2^31 - 1 atoms");
  if (nframes == 0 || natoms == 0) return MDSF_OK;
  if (!d_coords || !d_scale) return fail(MDSF_ERR_ARG, "null coordinates or scale");
  if (dtype != MDSF_F32 && dtype != MDSF_F64) return fail(MDSF_ERR_ARG, "dtype must be F32 or F64");
  if (arith != MDSF_F32 && arith != MDSF_F64) return fail(MDSF_ERR_ARG, "arith must be F32 or F64");
  if (dtype == MDSF_F64 && arith == MDSF_F32)
    return fail(MDSF_ERR_ARG, "float64 coordinates are rescaled in float64");
  MDSF_HIP(launch_rescale(d_coords, dtype, arith, nframes, (int)natoms, d_scale,
                          reinterpret_cast<hipStream_t>(stream)));
  return MDSF_OK;
}

int mdsf_fold_padded(const double* d_padded, const int32_t n[3], const int32_t nborder[3],
                     double* d_out, mdsf_stream stream) {
  if (!n || !nborder) return fail(MDSF_ERR_ARG, "null grid sizes");
  int nn[3], nb[3];
  for (int d = 0; d < 3; ++d) {
    nn[d] = n[d];
    nb[d] = nborder[d];
    if (nn[d] < 1 || nb[d] < 0) return fail(MDSF_ERR_ARG, "grid sizes must be positive");
    if (nb[d] > nn[d])
      return fail(MDSF_ERR_ARG, "the border exceeds the grid (the fold needs Nborder <= N)");
  }
  if (!d_padded || !d_out) return fail(MDSF_ERR_ARG, "null grid");
  MDSF_HIP(launch_fold_padded(d_padded, nn, nb, d_out, reinterpret_cast<hipStream_t>(stream)));
  return MDSF_OK;
}

int mdsf_ewald_slices(const double* d_kgridplt, int32_t nx, int32_t ny, int32_t nz,
                      double wavelength, double* d_xy, double* d_xz, double* d_yz,
                      mdsf_stream stream) {
  if (!d_kgridplt) return fail(MDSF_ERR_ARG, "null kgridplt");
  if (nx < 2 || ny < 2 || nz < 2) return fail(MDSF_ERR_ARG, "each axis needs at least 2 points");
  if (!(wavelength > 0.0)) return fail(MDSF_ERR_ARG, "wavelength must be positive");
  MDSF_HIP(launch_ewald(d_kgridplt, nx, ny, nz, wavelength, d_xy, d_xz, d_yz,
                        reinterpret_cast<hipStream_t>(stream)));
  return MDSF_OK;
}

int mdsf_rz_average(const double* d_kgridplt, int32_t nx, int32_t ny, int32_t nz,
                    const double* d_rarr, const int32_t* d_ntheta, int32_t nr,
                    const double* d_zar, int32_t nzb, const double binv[9], double* d_oa,
                    mdsf_stream stream) {
  if (!d_kgridplt || !binv) return fail(MDSF_ERR_ARG, "null kgridplt or binv");
  if (nx < 2 || ny < 2 || nz < 2) return fail(MDSF_ERR_ARG, "each axis needs at least 2 points");
  if (nr < 0 || nzb < 0) return fail(MDSF_ERR_ARG, "negative bin count");
  if ((long long)nr * nzb == 0) return MDSF_OK;
  if (!d_rarr || !d_ntheta || !d_zar || !d_oa) return fail(MDSF_ERR_ARG, "null bin array");
  MDSF_HIP(launch_rz_average(d_kgridplt, nx, ny, nz, d_rarr, d_ntheta, nr, d_zar, nzb, binv, d_oa,
                             reinterpret_cast<hipStream_t>(stream)));
  return MDSF_OK;
}

int mdsf_interp_points(const double* d_kgridplt, int32_t nx, int32_t ny, int32_t nz,
                       const double* d_pts, int64_t npts, double* d_out, mdsf_stream stream) {
  if (!d_kgridplt) return fail(MDSF_ERR_ARG, "null kgridplt");
  if (nx < 2 || ny < 2 || nz < 2) return fail(MDSF_ERR_ARG, "each axis needs at least 2 points");
  if (npts < 0) return fail(MDSF_ERR_ARG, "negative point count");
  if (npts == 0) return MDSF_OK;
  if (!d_pts || !d_out) return fail(MDSF_ERR_ARG, "null points or output");
  MDSF_HIP(launch_interp_points(d_kgridplt, nx, ny, nz, d_pts, npts, d_out,
                                reinterpret_cast<hipStream_t>(stream)));
  return MDSF_OK;
}

}  // extern "C"
