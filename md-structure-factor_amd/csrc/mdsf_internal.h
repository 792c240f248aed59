// Internal declarations shared by the HIP translation units of libmdsf.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mdsf.h"

namespace mdsf {

constexpr int kMaxTypes = 64;
constexpr int kTile = 16;       // deposition tile edge along x, y and z
constexpr int kBatch = 32;      // atoms staged per deposition batch
constexpr int kMaxStages = 16;  // FFT stages per axis
constexpr int kTuneReg = 100;   // FFT pencil choice: register-staged kernel
constexpr int kMaxSpillBufTiles = 4096;  // spill tiles staged per frame slot (else added in place)
constexpr int kCornerCap = 512;     // corner atoms listed per frame for the spill pass
constexpr int kListNoSlice = 255;  // tile-list entry: no slice on that axis (a byte per axis)
constexpr long long kSplitPrepTiles = 2048;  // tiles from which the frame loop splits its prepare stage
// stage-timing events per frame: prepare start / end (caller's stream), the
// deposition kernel's begin / end, the three FFT passes' begin / end, and the
// end of the prepare stage's tile-list branch
constexpr int kTimedEvents = 11;
constexpr int64_t kMaxTimedFrames = 1 << 16;  // stage-timing events kept before they start over

enum Side { SIDE_L = 0, SIDE_C = 1, SIDE_H = 2 };
// Cell metric G = U U^T of the plan: diagonal, xy-coupled (g01 != 0), or
// general (z coupled to x or y: g02 or g12 != 0).
enum Mode { MODE_DIAG = 0, MODE_XY = 1, MODE_GEN = 2 };

// Flags in the w component of an atom's int4 (home cell x, y, z; w = type |
// flags): the atom is a corner atom (its window leaves the core on all three
// axes), or it is parked -- its window leaves the padded grid on some axis,
// so the tiled passes skip it and the clipped-window passes handle it.
constexpr int kAtomCorner = 1 << 16;
constexpr int kAtomParked = 1 << 17;

// numpy's slice [i0:i1] of an axis of length len: negative bounds count from
// the end, then both clamp to [0, len]; empty when *s1 <= *s0.  The reference
// slices its padded grids this way (dens.py:292-308).
__host__ __device__ inline void numpy_slice(long long i0, long long i1, long long len,
                                            long long* s0, long long* s1) {
  if (i0 < 0) i0 = i0 + len < 0 ? 0 : i0 + len;
  else if (i0 > len) i0 = len;
  if (i1 < 0) i1 = i1 + len < 0 ? 0 : i1 + len;
  else if (i1 > len) i1 = len;
  *s0 = i0;
  *s1 = i1;
}

// Error bits raised by kernels into the plan's device flag word.
enum ErrBit { ERRBIT_RANGE = 1, ERRBIT_TYPE = 2, ERRBIT_LIST = 4 };

// Per-type constants on the device.
struct TypeDev {
  double amp;        // Nel / sigma^3
  double two_sig2;   // 2 sigma^2
  float c2;          // log2(e) / (2 sigma^2), for exp2 in the XY path
  float amp_f;
  int half[3];       // window half-widths A
  int count;         // atoms of this type
  int type_start;    // first sorted index of this type
  int stride;        // floats per atom: 2A_x + 2A_y + 2A_z
  long long table_base;  // float offset of this type's per-atom tables
  int bsz[3];        // bucket edge per axis
  int nbk[3];        // buckets per axis
  int bucket_base;   // first bucket id of this type
  int park;          // bucket id of its parked atoms (after its grid buckets)
  // tile slices (FP32 DIAG deposition): per atom and axis nsl[d] slices of
  // kTile floats, the image-summed 1-D factor over one tile along that axis
  int nsl[3];
  int slice_stride;        // floats per atom: 16 * (nsl[0] + nsl[1] + nsl[2])
  long long slice_base;    // float offset of this type's slices
  // corner components (atoms whose window leaves the cell on all three axes):
  // x {C, L+H}, y {C, L, H}, z {z slot for y on L, z slot for y on H}, each
  // nsl[d] slices per component
  int corner_stride;       // floats per atom: 16 * (2 nsl[0] + 3 nsl[1] + 2 nsl[2])
  long long corner_base;   // float offset of this type's corner components
  // adjoint work items: per atom adj_nxc x-chunks times adj_npc chunks of
  // adj_ppi (y, z) pairs, numbered from adj_item_base
  int adj_nxc;
  int adj_ppi;
  int adj_npc;
  long long adj_item_base;
  // tiled adjoint (FP32 DIAG): one result slot of 3 FP64 per (atom, tile
  // slice triple j0, j1, j2), nsl[0] nsl[1] nsl[2] per atom, from adjt_base
  long long adjt_base;
};

// Per-system constants passed by value to kernels.
struct SysDev {
  int n[3];
  int nb;
  double dr[3];
  double g00, g01, g11, g22;  // cell metric G = U U^T
  double g02, g12;            // its z couplings (nonzero only in MODE_GEN)
  int ntypes;
  int natoms;
  int nbuckets;
  int ntiles[3];
  int mode;
  // corner-fold spill tiles per axis: [0, sp_k0) and [sp_fh, ntiles), sp_cnt
  // of them (all tiles when the two ranges meet); none when nb = 0
  int sp_k0[3], sp_fh[3], sp_cnt[3];
};

// One FFT axis: length, radix sequence and twiddles W_n^k (k < n), FP32
// values generated in FP64 on the device.
struct FftAxis {
  int n = 0;
  int nstages = 0;
  int radix[kMaxStages] = {};
  float2* tw = nullptr;
  double2* tw64 = nullptr;
};

// Per-frame atom state (binning, sorted atoms, factor slices / tables): one
// per frame slot, so that one frame can be prepared while the previous one is
// deposited.
struct FrameState {
  int4* ir4 = nullptr;
  int* key = nullptr;
  double* rpos = nullptr;
  int* bucket_count = nullptr;
  int* bucket_off = nullptr;
  int* bucket_cur = nullptr;
  int* bucket_aux = nullptr;
  int* s_tmp = nullptr;
  int* s_idx = nullptr;
  int4* s_ir = nullptr;
  double* s_r = nullptr;
  void* tables = nullptr;
  float* slices = nullptr;
  float* dslices = nullptr;  // derivatives of the slices in the atom position (adjoint)
  // xy-coupled cells on tile lists: per regular x slice cell c_xy b_x, per
  // regular y slice cell b_y (same offsets as the slices), for the cross factor
  float* bslices = nullptr;
  // corner atoms (window off the core on all three axes): [0] count, then
  // kCornerCap sorted indices in arrival order, then those that spill in index
  // order, then their count
  int* corners = nullptr;
  // corner-fold spill of the frame, one 16^3 tile per spill tile (plan precision),
  // and per spill tile whether any atom spills into it
  void* spill = nullptr;
  int* spill_flag = nullptr;
  // tile lists (FP32 orthogonal-cell plans): entry pool and per-tile heads
  // {region start, regular atoms, corner atoms, region size}
  int2* tl_pool = nullptr;
  int4* tl_head = nullptr;
};

// Frame slots of a plan (atom state + density / spectrum buffers): frame t
// uses slot t mod kSlots, so that a frame can be prepared, the previous one
// deposited and the ones before it transformed without waiting for a buffer.
constexpr int kSlots = 3;

struct Plan {
  SysDev sys{};
  TypeDev types_host[kMaxTypes]{};
  TypeDev* types = nullptr;
  int device = 0;
  long long table_floats = 0;
  // per-atom scratch (unsorted)
  int4* ir4 = nullptr;
  int* key = nullptr;
  double* rpos = nullptr;
  // binning
  int* bucket_count = nullptr;
  int* bucket_off = nullptr;
  int* bucket_cur = nullptr;
  int* bucket_aux = nullptr;     // per-1024-block totals of the bucket scan
  int* s_tmp = nullptr;          // bucket fill (arbitrary order within a bucket)
  int* s_idx = nullptr;          // atom indices sorted by (type, bucket, index)
  int4* s_ir = nullptr;
  double* s_r = nullptr;
  void* tables = nullptr;         // per-atom factor tables (FP64 plans)
  float* slices = nullptr;        // per-atom tile slices (FP32 plans, orthogonal-axis cells)
  float* dslices = nullptr;       // their derivatives (adjoint only)
  float* bslices = nullptr;       // cross-term rows (xy-coupled cells on tile lists)
  int* corners = nullptr;         // corner-atom list of the current slot
  void* spill = nullptr;          // staged spill tiles of the current slot (or null)
  int* spill_flag = nullptr;      // their occupancy flags
  long long spill_tiles = 0;      // sp_cnt[0] sp_cnt[1] sp_cnt[2]
  long long slice_floats = 0;     // regular slices, then the corner components
  long long bslice_floats = 0;    // the regular part, when bslices are used (else 0)
  // tile lists of the current slot (null when the plan does not use them)
  int2* tl_pool = nullptr;
  int4* tl_head = nullptr;
  long long tl_cap = 0;           // pool entries per slot (0: no tile lists)
  int* err = nullptr;
  // grids: one set of density / spectrum buffers per frame slot, so that a
  // frame's FFT can run while the next frames are deposited; rho / spec / res /
  // spec_nyq point at the current one
  void* rho_buf[kSlots] = {};
  void* spec_buf[kSlots] = {};
  void* res_buf[kSlots] = {};
  void* nyq_buf[kSlots] = {};
  // rho: Nx*Ny*Nz folded density (plan precision), which the y pass then
  // overwrites with the spectrum rho[kz][ky][x] (Nx*Ny*Nz/2 complex)
  void* rho = nullptr;
  void* spec = nullptr;          // z pass output spec[x][kz][y], Nx*Ny*Nz/2 complex
  void* res = nullptr;           // y pass rounding residuals, a byte per complex (FP32 plans)
  void* spec_nyq = nullptr;      // Nx*Ny FP64 complex: the kz = Nz/2 plane, nyq[x][y]
  FrameState slot[kSlots];
  // Frame slot b: its atom state and density / spectrum buffer become the
  // ones the launches below use (kernel arguments are taken at launch).
  int cur_slot = 0;
  void use_slot(int b) {
    b %= kSlots;
    cur_slot = b;
    const FrameState& f = slot[b];
    ir4 = f.ir4;
    key = f.key;
    rpos = f.rpos;
    bucket_count = f.bucket_count;
    bucket_off = f.bucket_off;
    bucket_cur = f.bucket_cur;
    bucket_aux = f.bucket_aux;
    s_tmp = f.s_tmp;
    s_idx = f.s_idx;
    s_ir = f.s_ir;
    s_r = f.s_r;
    tables = f.tables;
    slices = f.slices;
    dslices = f.dslices;
    bslices = f.bslices;
    corners = f.corners;
    spill = f.spill;
    spill_flag = f.spill_flag;
    tl_pool = f.tl_pool;
    tl_head = f.tl_head;
    rho = rho_buf[b];
    spec = spec_buf[b];
    res = res_buf[b];
    spec_nyq = nyq_buf[b];
  }
  double* sf_main = nullptr;     // FP64 accumulator sf[kz][ky][kx], kz < Nz/2
  double* sf_nyq = nullptr;      // FP64 accumulator of the kz = Nz/2 plane, [kx][ky]
  // FFT
  int fft64 = 0;                 // deposition, density and FFT in FP64 (else FP32)
  FftAxis ax_x, ax_y, ax_zc;     // zc: complex length Nz/2
  // FFT launch choices (0 = default): z rows per tile; y / x pencil kernel
  // (kTuneReg = register-staged, 1..16 = generic LDS tile with that many pencils)
  int tune_z = 0, tune_y = 0, tune_x = 0;
  float2* tw_r2c = nullptr;      // W_Nz^k, k < Nz
  double2* tw_r2c64 = nullptr;
  // frame pipeline: frame t + 1 is prepared on the caller's stream while frame
  // t is deposited on dep_stream and frame t - 1 transformed on fft_stream
  // (both shared by all plans on the device, never destroyed)
  hipStream_t dep_stream = nullptr, fft_stream = nullptr;
  // the prepare stage's tile-list branch (frame loop; shared like the two above)
  hipStream_t aux_stream = nullptr;
  hipEvent_t ev_lfork = nullptr, ev_lgath = nullptr, ev_ljoin = nullptr;
  // optional stage timing of mdsf_accumulate_frames: 10 timing events per
  // frame (prepare, deposit, z, y, x: start / end, each on its stream)
  int timing = 0;  // 0 off, 1 every stage, 2 the deposition only
  // events the next tiled deposition records with its own dispatch (begin,
  // end), or null: set by the timed frame loop around launch_deposit
  hipEvent_t dep_ev[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> tev;
  int64_t timed_frames = 0;  // frames timed since the last mdsf_stage_times
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_prep[kSlots] = {}, ev_dep[kSlots] = {}, ev_fft[kSlots] = {};
  // adjoint (allocated on first use)
  long long adj_items = 0;
  double* adj_part = nullptr;    // 3 FP64 partial sums per work item
  long long adj_tslots = 0;      // tiled adjoint result slots
  double* adj_tpart = nullptr;   // 3 FP64 per slot
  double* adj_w = nullptr;       // 2 dL/dS in the sf layouts: main [kz][ky][kx] + Nyquist [kx][ky]
};

// error handling -------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define MDSF_HIP(call)                                        \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return ::mdsf::hip_fail(e_, #call); \
  } while (0)

// kernel launch wrappers (defined in the .hip units) --------------------------
hipError_t launch_prep(Plan& p, void* coords, int dtype, int arith, const int32_t* types,
                       const double* scale3, const double box[3], int64_t wrap_lo,
                       int64_t wrap_hi, uint32_t* occ, int32_t* home, hipStream_t s,
                       bool coords_only = false, bool write_back = true);
// sorts the prepared atoms and builds their tile slices; with `derivs` (and
// dslices allocated) also the slices' derivatives for the tiled adjoint.
// With `aux`, the tile lists are built on that stream, in parallel with the
// rest, and joined back into `s` (without `join`, the caller waits for
// p.ev_ljoin on the stream that reads them; `aux_end`, if given, is recorded
// on `aux` after the lists); with `side`, the corner list and the spill
// staging go on that stream (the one the frame's deposition then runs on),
// after the sorted atoms are ready.
hipError_t launch_binning(Plan& p, hipStream_t s, bool derivs = false, hipStream_t aux = nullptr,
                          hipStream_t side = nullptr, bool join = true,
                          hipEvent_t aux_end = nullptr);
// only the slice derivatives of the last sorted frame
hipError_t launch_slice_derivs(Plan& p, hipStream_t s);
hipError_t launch_deposit(Plan& p, void* out, hipStream_t s);
// the clipped windows of the frame's parked atoms (windows leaving the padded
// grid), added into the deposited density `out`; after launch_deposit
hipError_t launch_parked(Plan& p, void* out, hipStream_t s);
// corner-fold spill of the sorted frame into the slot's staging tiles (when the
// plan stages them; launch_deposit adds them as it stores the tiles)
hipError_t launch_spill_stage(Plan& p, hipStream_t s);
// per-tile atom lists (plans with tile lists): the tiles' pool regions (after
// the bucket scan), then their entries (after the atoms are sorted)
hipError_t launch_tile_regions(Plan& p, hipStream_t s);
hipError_t launch_tile_lists(Plan& p, hipStream_t s);
// tiled adjoint: adj_tpart slots <- per (atom, tile) gradient sums from
// dL/drho (FP32 DIAG plans; slots of tiles an atom does not reach are left
// untouched)
hipError_t launch_adjoint_tiles(Plan& p, const float* grho, hipStream_t s);

hipError_t fft_make_twiddles(float2* tw, double2* tw64, int n, hipStream_t s);
// density (FP32, Nx*Ny*Nz) -> spectrum; with `accumulate` the x pass adds
// |F|^2 into sf_main / sf_nyq instead of storing F.
hipError_t fft_forward(Plan& p, const void* density, bool accumulate, hipStream_t s);
// out = Re sum_q w(q) F(q) e^{+i q.r} over the stored spectrum (consumes it);
// out has the plan precision and may be the plan's density buffer.
hipError_t fft_inverse(Plan& p, const double* w_main, const double* w_nyq, void* out,
                       hipStream_t s);
// one forward pass: 0 = z (R2C from `density`), 1 = y, 2 = x with |F|^2 accumulation
hipError_t fft_pass(Plan& p, int pass, const void* density, hipStream_t s);
// Whether the plan's x pass can take two frames in one launch, and that
// launch: |F|^2 of the frame in slot_a, then of the one in slot_b, into the
// S(q) sums (the same FP64 additions in the same order as two x passes).
bool fft_x_pairs(const Plan& p);
hipError_t fft_x_pair(Plan& p, int slot_a, int slot_b, hipStream_t s);
// events (or null) that the next pass's stream kernel records with its own
// dispatch on this thread, and whether a launch took them
void fft_pass_events(hipEvent_t begin, hipEvent_t end);
bool fft_pass_events_used();
// stored spectrum -> complex128 (Nx, Ny, Nz/2+1)
hipError_t fft_spectrum_read(Plan& p, double* out, hipStream_t s);
bool fft_factor(int n, int* radix, int* nstages);
// bytes of the z pass output (spec) and of the y pass residuals (res) of a grid
size_t fft_spec_bytes(const int n[3], int fft64);
size_t fft_res_bytes(const int n[3], int fft64);
hipError_t fft_autotune(Plan& p, hipStream_t s);

hipError_t launch_fill_zero(double* p, long long n, hipStream_t s);
// adj_w <- 2 dsf split into main / Nyquist parts
hipError_t launch_adjoint_weights(Plan& p, const double* dsf, hipStream_t s);
// grad[atom] = scale3 * dL/dr for the last prepared frame from dL/drho (grho,
// plan precision); scale3 may be null (unit scale).  `derivs_ready`: the
// binning of this frame already built the slice derivatives; `grad_zeroed`:
// the caller cleared grad (needed only for atoms the last frame failed to bin).
hipError_t launch_deposit_adjoint(Plan& p, const void* grho, const double* scale3, double* grad,
                                  hipStream_t s, bool derivs_ready = false,
                                  bool grad_zeroed = false);

hipError_t launch_sf_read(const Plan& p, double* sf, hipStream_t s);
// fold of a padded FP64 grid (n + 2 nb per axis) onto n (remap_grid_tcl)
hipError_t launch_fold_padded(const double* d0, const int n[3], const int nb[3], double* d1,
                              hipStream_t s);
// in-place rescale of `nframes` frames of `natoms` coordinates (no wrap)
hipError_t launch_rescale(void* coords, int dtype, int arith, long long nframes, int natoms,
                          const double* scale, hipStream_t s);
hipError_t launch_outputs(const int n[3], const double* sf, const double box[3], int box_f32,
                          double* sfplt, double* kgrid, double* kgridplt, hipStream_t s);
hipError_t launch_ewald(const double* kp, int nx, int ny, int nz, double wl, double* xy,
                        double* xz, double* yz, hipStream_t s);
hipError_t launch_rz_average(const double* kp, int nx, int ny, int nz, const double* rarr,
                             const int* nth, int nr, const double* zar, int nzb,
                             const double binv[9], double* oa, hipStream_t s);
hipError_t launch_interp_points(const double* kp, int nx, int ny, int nz, const double* pts,
                                long long n, double* out, hipStream_t s);

}  // namespace mdsf
