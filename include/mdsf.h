/*
 * mdsf.h - C ABI of the MI355X (gfx950) structure-factor hot path.
 *
 * Drop-in replacement for the compute core of joeyelk/MD-Structure-Factor
 * `dens.compute_sf` (dens.py:166-346): per-frame Gaussian deposition onto a
 * periodic grid with the reference's border fold, a 3-D real-to-complex FFT,
 * and the frame sum of |F(q)|^2 into an FP64 structure factor.
 *
 * Conventions
 *   - Plain C types only.  Pointers named `d_*` are device pointers owned by
 *     the caller; the plan owns its workspaces and twiddle tables.
 *   - Every entry point returns MDSF_OK or an error code; the message of the
 *     last error on the calling thread is returned by mdsf_last_error().
 *   - Work is enqueued on the given HIP stream (NULL = legacy default stream);
 *     nothing synchronises the device except where stated.
 *   - A plan is not thread-safe; use one plan per host thread.
 *   - Grid arrays are C-order [x][y][z]; the spectrum keeps the reference's
 *     half-spectrum shape (Nx, Ny, Nz/2+1).
 *   - A plan computes in FP32 (default) or FP64 (flag MDSF_FFT_FP64): the
 *     density arrays passed in or out (`d_density`) have that element type.
 *     The S(q) frame sum is FP64 in both.
 */
#ifndef MDSF_H
#define MDSF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDSF_OK 0
#define MDSF_ERR_ARG 1        /* invalid argument (the reference raises ValueError)   */
#define MDSF_ERR_HIP 2        /* HIP runtime error                                    */
#define MDSF_ERR_RANGE 3      /* reserved: windows leaving the padded grid are clipped
                                 as the reference's numpy slices clip them           */
#define MDSF_ERR_TYPE 4       /* an atom type index is out of range (KeyError)        */
#define MDSF_ERR_UNSUPPORTED 5

#define MDSF_F32 0
#define MDSF_F64 1

/* mdsf_grid_desc.flags */
#define MDSF_FFT_FP64 1       /* deposition, density and FFT in FP64 (default FP32) */
#define MDSF_AUTOTUNE 2       /* time the FFT tile / kernel choices for this grid at
                                 plan creation (cached per grid shape and precision) */

typedef struct mdsf_plan mdsf_plan;
typedef void* mdsf_stream; /* a hipStream_t */

/* One atom type: replaces the per-type values `compute_sf` reads from the
 * radii dictionary and `get_borders` (dens.py:38-43, 287, 305-308).  A plan
 * takes at most 64.  An atom enters the density only through these values,
 * so names that share them (force-field names C1, C2, H17 ... of one element)
 * can and should be passed as one type. */
typedef struct mdsf_type_desc {
  double amp;        /* Nel / sigma**3                               dens.py:308 */
  double two_sig2;   /* 2 * sigma**2                                 dens.py:308 */
  int32_t half[3];   /* A = int(sigma*sqrt(ln(Nel/1e-24))/dr) per axis  dens.py:42,287 */
  int32_t count;     /* atoms of this type in the system             */
} mdsf_type_desc;

/* One system: grid, border, cell and types (dens.py:181-263). */
typedef struct mdsf_grid_desc {
  int32_t n[3];          /* Nspatialgrid, even                        dens.py:181-189 */
  int32_t nborder;       /* Nborder                                   dens.py:231     */
  double dr[3];          /* L / N                                     dens.py:202     */
  double ucell[9];       /* row-major 3x3; buf2 = buf @ ucell          dens.py:301     */
  int32_t ntypes;
  const mdsf_type_desc* types;
  int32_t natoms;
  int32_t flags;         /* MDSF_FFT_FP64 or 0                                          */
} mdsf_grid_desc;

/* Plan lifecycle.  Validates the descriptor (Nborder <= N, even N),
 * allocates workspaces on the current device and generates the FFT twiddles
 * in FP64 on the GPU.  Any 3x3 ucell is accepted (dens.py:301): cells whose
 * metric couples z with x or y deposit through an FP64 per-voxel gather (any
 * plan precision); the adjoint entry points reject such plans with
 * MDSF_ERR_UNSUPPORTED. */
int mdsf_plan_create(const mdsf_grid_desc* desc, mdsf_plan** out);
int mdsf_plan_destroy(mdsf_plan* plan);
const char* mdsf_last_error(void);
int mdsf_version(void);

/* Reset the plan's FP64 S(q) accumulator to zero (the `sf = np.zeros` of
 * dens.py:241). */
int mdsf_sf_reset(mdsf_plan* plan, mdsf_stream stream);

/* Synchronises `stream` and reports (then clears) the error flags the
 * kernels raised since the last check: MDSF_ERR_TYPE for a type index out of
 * range (the reference's KeyError, dens.py:287).  An atom whose window leaves
 * the padded grid is no error: the reference slices its grid and d0 with the
 * same indices (dens.py:292-308), which numpy clips alike, and so does the
 * deposition (a window across the low edge or past the high edge adds
 * nothing, one cut by the high edge adds its kept part, one wholly at
 * negative indices is counted from the far end of the padded grid).
 * MDSF_ERR_UNSUPPORTED reports an internal size bound the frame exceeded
 * (the tile-list pool of FP32 orthogonal-cell plans). */
int mdsf_check_errors(mdsf_plan* plan, mdsf_stream stream);

/* The frame loop of dens.py:277-321 for `nframes` frames.
 *   d_coords   (nframes, natoms, 3) of dtype MDSF_F32 / MDSF_F64; rescaled and
 *              wrapped IN PLACE exactly as dens.py:54-58 and 209-221 do.
 *   d_types    int32[natoms] type index of each atom (the `typ[im]` lookup).
 *   d_scale    double[nframes*3] per-frame rescale factors avg(dims)/dims[t]
 *              (dens.py:53).
 *   box        host double[3], the mean box L used by the wrap (dens.py:211-212).
 *   arith      MDSF_F32 when coords and dims are both float32 (numpy then does
 *              the rescale and wrap in float32), else MDSF_F64.
 *   wrap_lo/hi atom index range the wrap touches (all atoms below 1e6 atoms;
 *              see dens.py:213-221 for the buffered branch).
 *   d_occupancy optional uint32[Nx*Ny*Nz]: += 1 at (ir mod N) per atom and frame.
 *   d_home_cells optional int32[natoms*3]: ir of the LAST frame (dens.py:285);
 *              INT32_MIN for a non-finite coordinate or |r / dr| >= 1e9.
 * Adds sum_frames |rfftn(folded density)|^2 into the plan accumulator, in
 * frame order (the x passes of frames 2k and 2k + 1 may share one launch,
 * which adds the two frames' terms one after the other: the same result as
 * frame by frame). */
int mdsf_accumulate_frames(mdsf_plan* plan, void* d_coords, int dtype,
                           int64_t nframes, const int32_t* d_types,
                           const double* d_scale, const double box[3], int arith,
                           int64_t wrap_lo, int64_t wrap_hi,
                           uint32_t* d_occupancy, int32_t* d_home_cells,
                           mdsf_stream stream);

/* Only the in-place rescale and wrap of dens.py:179,206-221 for `nframes`
 * frames (arguments as for mdsf_accumulate_frames).  The RANDOM_NOISE branch
 * of the reference replaces the density but still mutates the coordinates. */
int mdsf_rescale_wrap(mdsf_plan* plan, void* d_coords, int dtype, int64_t nframes,
                      const double* d_scale, const double box[3], int arith, int64_t wrap_lo,
                      int64_t wrap_hi, mdsf_stream stream);

/* The in-place rescale of dens.py:46-62 (`rescale(coords, dims)`) alone, no
 * plan: d_coords (nframes, natoms, 3) of dtype MDSF_F32 / MDSF_F64 is
 * multiplied per frame and axis by d_scale[3 t + i] = avg(dims)[i] / dims[t, i]
 * in the arithmetic `arith` (as for mdsf_accumulate_frames). */
int mdsf_rescale(void* d_coords, int dtype, int64_t nframes, int64_t natoms,
                 const double* d_scale, int arith, mdsf_stream stream);

/* The border fold of a padded FP64 grid onto the periodic cell, with the
 * tables compute_sf builds (des = [0, Nb, N - Nb, N], ori = [0, Nb, Nb + N,
 * 2 Nb + N] per axis, dens.py:258-263): remap_grid_tcl (dens.py:86-108),
 * corner z slot included, and remap_grid (dens.py:65-83) with one table for
 * every axis.  d_padded (n + 2 nborder per axis) -> d_out (n); the additions
 * run in the reference's order, so the result is bitwise the reference's.
 * Requires nborder[d] <= n[d]. */
int mdsf_fold_padded(const double* d_padded, const int32_t n[3], const int32_t nborder[3],
                     double* d_out, mdsf_stream stream);

/* One frame of deposition + periodic fold (dens.py:283-311), no FFT:
 * writes the folded density d_density[Nx*Ny*Nz] (plan precision).  Coordinates are
 * processed as in mdsf_accumulate_frames (rescale with d_scale[0..2], wrap). */
int mdsf_deposit_fold(mdsf_plan* plan, void* d_coords, int dtype,
                      const int32_t* d_types, const double* d_scale,
                      const double box[3], int arith, int64_t wrap_lo,
                      int64_t wrap_hi, void* d_density, int32_t* d_home_cells,
                      mdsf_stream stream);

/* The frame loop of mdsf_accumulate_frames split into its launches, for
 * callers that interleave their own work or time each stage:
 *   mdsf_prepare_frame  rescale + wrap in place, home cells, binning and the
 *                       per-atom factor tables of one frame (dens.py:179-231,285-287);
 *   mdsf_deposit        deposition + fold of the prepared frame into d_density
 *                       (NULL: the plan's own density buffer) (dens.py:283-311);
 *   mdsf_fft_pass       one forward pass over the plan's spectrum: 0 = z (R2C of
 *                       d_density, NULL: the plan's density), 1 = y, 2 = x with
 *                       |F|^2 summed into the accumulator (dens.py:313-318). */
int mdsf_prepare_frame(mdsf_plan* plan, void* d_coords, int dtype, const int32_t* d_types,
                       const double* d_scale, const double box[3], int arith, int64_t wrap_lo,
                       int64_t wrap_hi, uint32_t* d_occupancy, int32_t* d_home_cells,
                       mdsf_stream stream);
int mdsf_deposit(mdsf_plan* plan, void* d_density, mdsf_stream stream);
int mdsf_fft_pass(mdsf_plan* plan, int32_t pass, const void* d_density, mdsf_stream stream);

/* The plan keeps three frame slots (per-frame atom state and density /
 * spectrum buffer) so that mdsf_accumulate_frames can prepare one frame,
 * deposit the previous one and transform the ones before on three streams.
 * This selects the slot (0, 1 or 2) that mdsf_prepare_frame, mdsf_deposit,
 * mdsf_fft_pass and the other single-frame entry points use from now on; a
 * caller pipelining the stages itself rotates them.  No counterpart in the
 * reference (its frame loop is sequential, dens.py:277-321). */
int mdsf_set_frame_buffer(mdsf_plan* plan, int32_t which);

/* Stage timing of mdsf_accumulate_frames (off by default): with enable = 1
 * the frame loop records HIP events for each frame's preparation (from its
 * start on the caller's stream to the end of the later of its caller-stream
 * part and its tile-list branch), deposition and three FFT passes (with the tiled
 * deposition kernel's and each pass kernel's own dispatch); with enable = 2
 * only for the deposition (the least perturbation of the overlapped loop).
 * Timed frames accumulate over calls; mdsf_stage_times waits for them and
 * returns the average milliseconds per frame of {prepare, deposit, z pass,
 * y pass, x pass + |F|^2 sum} in avg_ms[5] (0 for the stages not timed),
 * then starts the count over.  An x pass shared by two frames is timed as the
 * second frame's and the first frame's as zero, so the average is the cost
 * per frame.  No counterpart in the reference. */
int mdsf_set_stage_timing(mdsf_plan* plan, int32_t enable);
int mdsf_stage_times(mdsf_plan* plan, double* avg_ms);

/* |rfftn(density)|^2 summed into the accumulator for a caller-supplied
 * periodic density in plan precision (the RANDOM_NOISE branch,
 * dens.py:279-280, 313-318). */
int mdsf_accumulate_density(mdsf_plan* plan, const void* d_density,
                            mdsf_stream stream);

/* Forward R2C transform of one density in plan precision (np.fft.rfftn,
 * dens.py:313): d_spectrum receives complex128 (Nx, Ny, Nz/2+1) as
 * interleaved re/im doubles. */
int mdsf_rfft3(mdsf_plan* plan, const void* d_density, double* d_spectrum,
               mdsf_stream stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the transform of mdsf_rfft3 without a plan, on the grid
 * n[3] (even sizes) in `precision` (MDSF_F32 / MDSF_F64, the element type of
 * d_density).  d_spectrum receives complex128 (Nx, Ny, Nz/2+1), bit for bit
 * what mdsf_rfft3 gives on a plan of this grid and precision created without
 * MDSF_AUTOTUNE.  Allocates nothing: its twiddles, pass outputs and Nyquist
 * plane live in one fixed 144 MiB device array inside the library (one copy
 * per GPU, held until the process ends), which fits every FP32 grid up to
 * 256^3 and every FP64 grid up to 208^3; larger grids are rejected with
 * MDSF_ERR_ARG.  Every call on a GPU uses that GPU's one array, so calls that
 * overlap in time on one GPU (other streams, other host threads) race. */
int mdsf_rfft3_noalloc(const int32_t n[3], int precision, const void* d_density,
                       double* d_spectrum, mdsf_stream stream);

/* The accumulated S(q) in the reference layout: d_sf double (Nx, Ny, Nz/2+1)
 * (the `sf` key of the .npz, dens.py:318,346). */
int mdsf_sf_read(mdsf_plan* plan, double* d_sf, mdsf_stream stream);

/* The derived outputs of dens.py:323-344 from an S(q) in reference layout:
 *   d_sfplt    (Nx-2, Ny-2, Nz-1)        get_dplot(sf), dens.py:142-163
 *   d_kgrid    (Nx, Ny, Nz/2+1, 4)       dens.py:325-334
 *   d_kgridplt (Nx-2, Ny-2, Nz-1, 4)     dens.py:336-344
 * box_f32 != 0 when L is float32 (numpy then divides in float32). Any output
 * pointer may be NULL. */
int mdsf_outputs(mdsf_plan* plan, const double* d_sf, const double box[3],
                 int box_f32, double* d_sfplt, double* d_kgrid,
                 double* d_kgridplt, mdsf_stream stream);

/* ---- adjoint (dL/dS -> dL/dx) ------------------------------------------
 * The reference has no gradient; these differentiate the function the
 * reference computes (dens.py:283-318, including the corner fold of
 * dens.py:107) for a loss L(S).  d_dsf is dL/dS in the reference layout,
 * double (Nx, Ny, Nz/2+1).  Nothing per frame is stored by the forward pass:
 * each frame's density and spectrum are recomputed. */

/* dL/d(coords) for `nframes` frames laid out as for mdsf_accumulate_frames
 * (plans whose cell couples z with x or y: MDSF_ERR_UNSUPPORTED, as for
 * mdsf_deposit_adjoint).
 * The coordinates are rescaled and wrapped in place exactly as the forward
 * pass does (pass a copy of the original coordinates); d_grad (nframes,
 * natoms, 3) double receives dL/d(original coordinates) = scale * dL/dr.
 * Atoms whose windows the reference's slices empty get a zero gradient;
 * windows cut by the padded grid's high edge differentiate their kept part. */
int mdsf_adjoint_frames(mdsf_plan* plan, void* d_coords, int dtype,
                        int64_t nframes, const int32_t* d_types,
                        const double* d_scale, const double box[3], int arith,
                        int64_t wrap_lo, int64_t wrap_hi, const double* d_dsf,
                        double* d_grad, mdsf_stream stream);

/* dL/d(density) = 2 Re sum_q dsf(q) F(q) exp(+2 pi i q.r/N) for one periodic
 * density (plan precision in and out; the two may be the same buffer):
 * the transpose of density -> |rfftn(density)|^2 (dens.py:313-318). */
int mdsf_density_adjoint(mdsf_plan* plan, const void* d_density,
                         const double* d_dsf, void* d_grad_density,
                         mdsf_stream stream);

/* dL/dr (natoms, 3) double, for the wrapped coordinates of the frame last
 * given to mdsf_prepare_frame / mdsf_deposit_fold, from dL/d(folded density)
 * in plan precision: the transpose of deposition + fold (dens.py:283-311). */
int mdsf_deposit_adjoint(mdsf_plan* plan, const void* d_grad_density,
                         double* d_grad, mdsf_stream stream);

/* ---- batches of independent systems ---------------------------------------
 * No counterpart in the reference, whose driver runs one system per call
 * (main_gromacs.py:209-212).  A batch owns one plan per system (created with
 * MDSF_AUTOTUNE); mdsf_batch_run runs every system's frame loop -- S(q)
 * reset, then per frame the work of mdsf_accumulate_frames -- from one
 * instantiated HIP graph, captured on the first run after the frames change
 * and replayed afterwards.  Unlike mdsf_accumulate_frames the batch reads the
 * coordinates without rescaling / wrapping them in place, so every replay
 * sees the same input.  The device buffers passed to mdsf_batch_set_frames
 * must stay valid while the batch may run. */
typedef struct mdsf_batch mdsf_batch;

int mdsf_batch_create(const mdsf_grid_desc* descs, int32_t nsys,
                      mdsf_batch** out);
int mdsf_batch_destroy(mdsf_batch* batch);

/* The plan of system `sys` (for mdsf_sf_read, mdsf_outputs, ...); NULL if out
 * of range. */
mdsf_plan* mdsf_batch_plan(mdsf_batch* batch, int32_t sys);

/* Frames of system `sys`, with the meaning of mdsf_accumulate_frames'
 * arguments. */
int mdsf_batch_set_frames(mdsf_batch* batch, int32_t sys,
                          const void* d_coords, int dtype, int64_t nframes,
                          const int32_t* d_types, const double* d_scale,
                          const double box[3], int arith, int64_t wrap_lo,
                          int64_t wrap_hi);

/* Enqueue the whole batch on `stream` (one graph launch). */
int mdsf_batch_run(mdsf_batch* batch, mdsf_stream stream);

/* mdsf_check_errors over every system (synchronises the stream). */
int mdsf_batch_check_errors(mdsf_batch* batch, mdsf_stream stream);

/* Ewald-sphere sections of a kgridplt volume (plot2d.py:218-263): trilinear
 * interpolation of channel 3 at the points of the xy, xz and yz sections for
 * wavelength `wavelength` (Angstrom); NaN outside the grid
 * (RegularGridInterpolator(..., bounds_error=False)).
 *   d_kgridplt (nx, ny, nz, 4) double; outputs (nx,ny), (nx,nz), (ny,nz). */
int mdsf_ewald_slices(const double* d_kgridplt, int32_t nx, int32_t ny,
                      int32_t nz, double wavelength, double* d_xy, double* d_xz,
                      double* d_yz, mdsf_stream stream);

/* Cylindrical average over theta, the r-theta-z interpolation of
 * PLOT_RAD_NEW (plot2d.py:595-627, run by main_gromacs.py:231 through
 * Plot_Ewald_triclinic, plot2d.py:1010):
 *   d_oa[ir * nzb + iz] = mean over k < d_ntheta[ir] of S at
 *   (r cos t_k, r sin t_k, z) @ binv,  t_k = k * 2 pi / d_ntheta[ir],
 *   r = d_rarr[ir], z = d_zar[iz], S the trilinear interpolant of channel 3
 *   of kgridplt (RegularGridInterpolator, linear); NaN when a sample leaves
 *   the grid.  binv: row-major 3x3 (the inverse of the reciprocal basis).
 *   Requires d_ntheta[ir] >= 1. */
int mdsf_rz_average(const double* d_kgridplt, int32_t nx, int32_t ny, int32_t nz,
                    const double* d_rarr, const int32_t* d_ntheta, int32_t nr,
                    const double* d_zar, int32_t nzb, const double binv[9], double* d_oa,
                    mdsf_stream stream);

/* The same interpolant at npts points d_pts (npts, 3) -> d_out (npts), NaN
 * outside the grid (the plain-grid sections of plot2d.py:800-840). */
int mdsf_interp_points(const double* d_kgridplt, int32_t nx, int32_t ny, int32_t nz,
                       const double* d_pts, int64_t npts, double* d_out, mdsf_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* MDSF_H */
