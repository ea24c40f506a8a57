/*
 * cbs_hip.h -- C ABI of the MI355X convergent Born series (CBS) solver.
 *
 * Drop-in native backend for the CBS forward solver of BMOLKAIST/Inverse_solver.
 * Each entry point replaces one action of the reference's CudaBorn MEX protocol
 * (Codes/CUDA_CODE/main.cpp:53-213) or one step of the MATLAB class
 * FORWARD_SOLVER_CONVERGENT_BORN that the MEX serves.  Plain pointers and
 * sizes only; no exceptions cross this boundary.  Every function returns a
 * CBS_* status code; on failure cbs_last_error() returns a message (thread local).
 *
 * Arrays are complex64 (interleaved re/im float pairs), column-major with x
 * fastest, exactly as MATLAB passes single complex arrays with -R2018a
 * (Codes/CUDA_CODE/compile_CudaBorn.m:4).  A batch of n_rhs arrays is stored
 * back to back (the right-hand-side index is the slowest).
 *
 * Threading: one handle serialises its own calls on its own HIP stream;
 * distinct handles are independent.  Calls on one handle from several threads
 * at once are not allowed (as with the reference MEX).
 */
#ifndef CBS_HIP_H
#define CBS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CBS_ABI_VERSION 1

typedef struct cbs_solver* cbs_handle;

enum {
  CBS_OK = 0,
  CBS_ERR_ARG = 1,          /* invalid argument or size mismatch */
  CBS_ERR_HANDLE = 2,       /* unknown or destroyed handle */
  CBS_ERR_HIP = 3,          /* HIP runtime error */
  CBS_ERR_UNSUPPORTED = 4,  /* grid shape without a kernel plan */
  CBS_ERR_STATE = 5,        /* call out of order (e.g. solve before set_potential) */
  CBS_ERR_MEMORY = 6        /* device allocation failed or over the admission limit */
};

enum { CBS_MEM_HOST = 0, CBS_MEM_DEVICE = 1 };
enum { CBS_FORM_FUSED = 0, CBS_FORM_PLAIN = 1, CBS_FORM_ROWS = 2, CBS_FORM_AXES = 3 };

/* Geometry of a solver instance.
 * Replaces the struct of CudaBorn action 0 (main.cpp:57-101: simulation_size,
 * RI_size, field_size, conv_size, green_size, used_gpu).  green_size is not
 * needed because the I/O kernels are passed whole (cbs_set_io_kernels). */
typedef struct {
  uint32_t simulation_size[3]; /* RI grid + 2 * boundary pixels */
  uint32_t RI_size[3];
  uint32_t field_size[3];      /* xy size of the 2D fields; [2] ignored */
  uint32_t conv_size[3];       /* xy convolution grid (field + RI); [2] ignored */
  int32_t device;              /* HIP device ordinal */
  int32_t RI_center[2];        /* xy offset of the RI grid in the field (RI_center, :40-41); the
                                  I/O kernels honour it (:77-106), but cbs_set_io_kernels refuses
                                  values other than 0, as the reference's crop_field2RI (:173-179)
                                  stops every 2D solve with such a value */
  int32_t pole_num;            /* field components: 1 scalar, 3 vector with the dyadic Green's
                                  function (FORWARD_SOLVER_CONVERGENT_BORN.m:294-297; 0 means 1).
                                  With 3, every per-rhs array holds the x, y, z components in turn:
                                  [rhs][component][z][y][x]. */
} cbs_geometry;

/* Scalars set with the potential.
 * Replaces the struct of CudaBorn action 1 (main.cpp:102-138: k0_nm, eps_imag,
 * fourier_res1..3, ROI, Bornmax, dV); the remaining fields carry what the
 * MATLAB class derives from its parameters (FORWARD_SOLVER_CONVERGENT_BORN.m:
 * boundary_sharpness :37, acyclic :39, the shifted-axis choice :300-315 and the
 * phase-ramp axes :355-377). */
typedef struct {
  double k0_nm;               /* RI_bg / wavelength */
  double eps_imag;            /* 1.01 * max |Vt| (> 0) */
  double fourier_res[3];      /* 1 / (resolution * simulation_size) */
  uint32_t ROI[6];            /* 1-based inclusive [x0 x1 y0 y1 z0 z1] of the RI grid */
  uint32_t Bornmax;           /* number of Born iterations (FFT pairs) */
  double dV;                  /* voxel volume */
  double boundary_sharpness;  /* attenuation-mask sharpness (reference default 1) */
  uint32_t acyclic;           /* 1: alternate +-1/4-bin shifted Green's function */
  uint32_t shift_axes;        /* bit j set: axis j carries the 1/4-bin shift */
  uint32_t ramp_axes;         /* bit j set: axis j carries the phase ramp */
} cbs_potential_params;

typedef struct {
  int32_t form;         /* CBS_FORM_FUSED (two passes per iteration; grids without such a plan fall
                           back to CBS_FORM_ROWS, then CBS_FORM_PLAIN), CBS_FORM_ROWS (four passes:
                           row blocks in real space, y transforms, z transforms with the Green's
                           multiply), CBS_FORM_PLAIN (per-axis transforms over the whole batch:
                           for one field component the x inverse, the pointwise update and the x
                           forward run as one launch and the z forward, the Green's multiply and
                           the z inverse as another, 4 launches per iteration; grids whose line
                           lengths have no compiled kernel run it on run-time-planned line FFTs)
                           or CBS_FORM_AXES (the per-axis form with every transform, the Green's
                           multiply and the update in a launch of its own: 8 per iteration) */
  int32_t split_a;      /* fused form: force ny = a*b split (0 = planner's choice) */
  int32_t tile_t;       /* fused form: force the k-space x tile (0 = planner's choice) */
  int32_t max_batch;    /* right-hand sides per launch (0 = all of the call) */
  int32_t track_norms;  /* 1: accumulate FP64 |f_n|^2 per rhs and iteration */
  int32_t use_graph;    /* 1 (the default): capture the iteration loop in a hipGraph on first use
                           and replay it (not while time_kernels is set); 0: launch every kernel.
                           The library keeps its own device-wide operations (allocations, frees,
                           stream and graph teardown) from overlapping a capture on another
                           thread; a caller that frees device memory or synchronises the device
                           from one thread while another thread's handle solves must hold off
                           the same way (or set use_graph = 0) */
  int32_t time_kernels; /* 1: time each pass kind with HIP events (cbs_get_kernel_times) */
  int32_t threads_real;   /* fused form: force threads per block of the real-space pass (0 = planner) */
  int32_t threads_kspace; /* fused form: force threads per block of the k-space pass (0 = planner) */
  int32_t block_order;    /* fused form, 0..15 (default 4; cbs_autotune picks): bit 0 real-space
                             pass, bit 1 k-space pass, 0 = rhs-fastest block numbering, 1 = rows
                             of one rhs together; bit 2 real-space tiles in storage order, rhs
                             slowest (overrides bit 0); bit 3 nontemporal k-space loads (large
                             batches) */
} cbs_options;

/* Creates a solver on geo->device.  Replaces CudaBorn action 0 (main.cpp:57-101). */
int cbs_create(const cbs_geometry* geo, cbs_handle* out);

/* Replaces CudaBorn action 3 (main.cpp:164-174): frees every device buffer. */
int cbs_destroy(cbs_handle h);

/* Options (no counterpart in the reference, which has a fixed cuFFT plan). */
int cbs_set_options(cbs_handle h, const cbs_options* opt);
int cbs_get_options(cbs_handle h, cbs_options* opt);

/* Uploads V = Vt - i*eps (complex64, simulation grid) and the scalars.
 * Replaces CudaBorn action 1 (main.cpp:102-138), i.e. the device half of
 * FORWARD_SOLVER_CONVERGENT_BORN.set_RI (:180-198).  Resets the alternating
 * shift state, as set_RI's init() does (:280-396). */
int cbs_set_potential(cbs_handle h, const cbs_potential_params* prm, const void* V, int mem);

/* The host half of set_RI moved to the device.  The RI grid (RI_size, x fastest)
 * is FP64: real values (ri_complex = 0) or interleaved complex (1), on the host or
 * the device (mem).  Vt = k^2 (RI^2 / RI_bg^2 - 1), k = 2 pi RI_bg / wavelength
 * (Codes/sub_code/RI2potential.m:1-4), evaluated in FP64 with the operation order
 * of the host formula, so real RI gives the same V bit for bit.
 *
 * cbs_potential_max: max |Vt| over the grid (eps_imag = 1.01 of it,
 * FORWARD_SOLVER_CONVERGENT_BORN.m:205; the caller also derives Bornmax from it,
 * :326-336).  The same for conj(RI).
 * cbs_set_potential_ri: as cbs_set_potential, with V = pad(Vt) - i eps_imag built
 * on the device (:212-215, create_boundary_RI :238-241) from the RI grid instead of
 * uploaded; flip_conj = 1 uses conj(RI) reversed along z, the adjoint potential of
 * BACKWARD_SOLVER_MULTI.m:48, so one resident RI serves both solves. */
int cbs_potential_max(cbs_handle h, const void* ri, int ri_complex, double wavelength, double RI_bg, int mem,
                      double* vmax);
int cbs_set_potential_ri(cbs_handle h, const cbs_potential_params* prm, const void* ri, int ri_complex, int flip_conj,
                         double wavelength, double RI_bg, int mem);

/* Born loop from the incident 3D field: u_out = Field[ROI] + u_in for n_rhs
 * right-hand sides (RI grid each).  Replaces CudaBorn action 2 (main.cpp:139-163)
 * and the MATLAB loop FORWARD_SOLVER_CONVERGENT_BORN.solve_raw :564-647.
 * u_in and u_out are both host or both device pointers (mem); they may alias.
 * With host pointers and enough right-hand sides (about six chunks of at least 8),
 * the copies of the neighbouring chunks run on a second stream while a chunk solves. */
int cbs_solve_3d(cbs_handle h, const void* u_in, uint32_t n_rhs, void* u_out, int mem);

/* cbs_solve_3d on device buffers without allocating: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The batch's state (Field, f, the
 * spectrum, flip parities, FP64 norm accumulators) lives in a fixed workspace of
 * CBS_NOALLOC_WORKSPACE_BYTES that the library holds on every GPU it runs on until the
 * process ends; a batch that needs more (about 3 * 8 B * simulation voxels * field
 * components per rhs, per batch of max_batch rhs) is refused with CBS_ERR_MEMORY.  The
 * workspace is shared by every handle on a GPU: calls that overlap on one GPU race.
 * The kernel plan, V and the tables are set up beforehand (cbs_set_potential*,
 * cbs_autotune); this call creates no graph and no timing events (use_graph and
 * time_kernels are ignored).  u_in and u_out are device pointers; they may alias. */
#define CBS_NOALLOC_WORKSPACE_BYTES (64ull << 20)
int cbs_solve_3d_noalloc(cbs_handle h, const void* u_in, uint32_t n_rhs, void* u_out);

/* Uploads the refocusing and transmission kernels for the 2D input/output
 * projections (each conv_x * conv_y * RI_z complex64, centred (kx, ky) order,
 * z in real space, as FORWARD_SOLVER_CONVERGENT_BORN's refocusing_util and
 * kernel_trans :80-115).  kernel_ref may be null (then reflection is not
 * available).  Replaces the green octants of CudaBorn action 0 (main.cpp:93-94). */
int cbs_set_io_kernels(cbs_handle h, const void* refocus, const void* kernel_trans, const void* kernel_ref, int mem);

/* The same from the first-octant kernels that CudaBorn action 0 receives
 * (main.cpp:93-94; FORWARD_SOLVER_CONVERGENT_BORN_CUDA.m:28-38 passes
 * refocusing_util(1:gx, 1:gy, 1:gz) and kernel_trans(1:gx, 1:gy, 1:gz), gx = floor(conv_x/2)+1,
 * gy = floor(conv_y/2)+1, gz = floor(RI_z/2)+1).  They are expanded with the symmetry the
 * reference's field_in/field_out kernels apply (execution2.cu:195-275); the reflection
 * kernel is the conjugated transmission kernel, as there. */
int cbs_set_io_kernels_octant(cbs_handle h, const void* refocus_octant, const void* trans_octant, int mem);

/* Full 2D-in / 2D-out solve for n_rhs angles.  source_2d: field grid, real
 * space, already NA/Abbe filtered (what FORWARD_SOLVER_CONVERGENT_BORN_CUDA.m:122
 * passes).  Outputs (any may be null): trans_2d and ref_2d on the field grid
 * before the host adds the input spectrum and applies transform_field_2D
 * (:133-145), field_3d on the RI grid.  Replaces CudaBorn action 4
 * (main.cpp:175-210: execute_2D_in, execute, execute_2D_out). */
int cbs_solve_2d(cbs_handle h, const void* source_2d, uint32_t n_rhs, void* trans_2d, void* ref_2d, void* field_3d,
                 int mem);

/* Isotropic 3D TV proximal step of the reconstruction loop, on the handle's
 * device and stream: replaces Codes/sub_code/TV_FISTA_inner_v2.m for a real
 * input (kappamax = 0), in single precision as the reference runs it.
 * in/out: dims[0]*dims[1]*dims[2] floats (x fastest), host or device (mem);
 * they may alias.  non_neg clamps to [vmin, vmax]; dirichlet zeroes the faces. */
int cbs_tv_prox(cbs_handle h, const float* in, float* out, const uint32_t dims[3], double lambda, double vmin,
                double vmax, int non_neg, int dirichlet, uint32_t inner_itt, int mem);

/* The same proximal step for a complex potential: Codes/sub_code/TV_FISTA_inner_v2.m
 * when Vimag_max != 0, which includes the reference defaults (nmax = Inf and
 * kappamax = 0 give Vimag_max = 0*Inf = NaN at BACKWARD_SOLVER_MULTI.m:135, so
 * :10-12 keeps the imaginary part).  in/out: complex64 (re, im interleaved).
 * project_non_neg (:57-75): negative imaginary parts become 0, and when
 * vimag_max > 0 the imaginary part is capped at it (NaN means no cap); the dual
 * norm is sqrt(sum_k |P_k|^2) over complex components (project_TV :79-91). */
int cbs_tv_prox_complex(cbs_handle h, const void* in, void* out, const uint32_t dims[3], double lambda, double vmin,
                        double vmax, double vimag_max, int non_neg, int dirichlet, uint32_t inner_itt, int mem);

/* Misfit gradient of BACKWARD_SOLVER_MULTI.get_gradiant_static (:50-53) from the
 * 3D fields the forward (u) and adjoint (w) solves return through
 * cbs_solve_2d's field_3d: grad = -i/n_rhs * sum_r conj(u_r) * flip_z(conj(w_r)),
 * summed in FP64.  u, w: n_rhs blocks of dims[0]*dims[1]*dims[2] complex64
 * (x fastest); grad: dims complex128.  The reference forms it from gpuArrays in
 * MATLAB; this keeps both fields in device memory. */
int cbs_field_gradient(cbs_handle h, const void* u, const void* w, uint32_t n_rhs, const uint32_t dims[3], void* grad,
                       int mem);

/* cbs_solve_2d for scalar fields from the incident field itself: the source is
 * cbs_field_pupil(in_field, pupil_in), and the outputs are
 * trans = cbs_field_pupil(t + source, pupil_out) and refl = cbs_field_pupil(r, pupil_out)
 * (fft2(t) + U = fft2(t + source), U the filtered input spectrum), t and r being what
 * cbs_solve_2d returns; field_3d as there.  The whole of
 * FORWARD_SOLVER_CONVERGENT_BORN.solve (FORWARD_SOLVER_CONVERGENT_BORN.m:397-520) for
 * pole_num 1 in one call, with no host transform between the MEX actions; pupils as in
 * cbs_field_pupil; mem applies to in_field, trans, refl and field_3d. */
int cbs_solve_field(cbs_handle h, const void* in_field, uint32_t n_rhs, const float* pupil_in, const float* pupil_out,
                    void* trans, void* refl, void* field_3d, int mem);

/* Centred pupil filter of scalar fields on the field grid (field_size x, y):
 *   out_f = fftshift(ifft2(ifftshift(P .* fftshift(fft2(ifftshift(in_f + add_f))))))
 * for n_fields fields of complex64 (x fastest); add may be null; out may be in.  P:
 * field_x * field_y floats on the host, x fastest, in centred frequency order.  The
 * scalar case of FORWARD_SOLVER.transform_field_3D / transform_field_2D
 * (FORWARD_SOLVER.m:38-97) between the centred transforms of
 * FORWARD_SOLVER_CONVERGENT_BORN.m:415-417 and :481-508, which the reference runs in
 * MATLAB around the MEX: P is the NA circle times 1/sqrt(cos theta) (input) or
 * sqrt(cos theta) (output) under use_abbe_sine.  mem applies to in, add and out. */
int cbs_field_pupil(cbs_handle h, const void* in, const void* add, uint32_t n_fields, const float* pupil, void* out,
                    int mem);

/* The outer FISTA iteration of BACKWARD_SOLVER_MULTI (:152-160) on device arrays of n
 * values, x fastest (the reference runs it on gpuArrays in MATLAB), so the iterates, the
 * gradient and the RI stay in device memory between gradients.  Both run on the
 * handle's stream and return when their work is done.
 * cbs_fista_prox_input: out = u - inv_alpha * grad (u, grad complex128), rounded to
 * complex64 when cplx = 1 (the input of cbs_tv_prox_complex) or its real part to float32
 * when cplx = 0 (cbs_tv_prox).
 * cbs_fista_momentum: with s the proximal output (complex64 when cplx = 1, else float32),
 * u = s + c (s - x), x = s and ri = n_m sqrt(u + 1) (principal branch), all complex128. */
int cbs_fista_prox_input(cbs_handle h, const void* u, const void* grad, double inv_alpha, uint64_t n, int cplx,
                         void* out);
int cbs_fista_momentum(cbs_handle h, const void* s, int cplx, void* x, void* u, void* ri, uint64_t n, double c,
                       double n_m);

/* Phase unwrapping for the Rytov initialiser of BACKWARD_SOLVER_MULTI (handle-free;
 * synchronous on the default stream of `device`).  Arrays are column-major n1 x n2
 * phase maps, n3 of them (one per angle), single precision.
 *
 * Residue map of every 2x2 loop, (n1-1) x (n2-1) x n3 values in {-1, 0, 1}.
 * Replaces get_residue_gpu (Codes/sub_code/goldstein_GPU/get_residue_mex.cpp,
 * get_residue_cuda.cu:35-66). */
int cbs_phase_residues(int device, const float* wrapped, float* residues, const uint32_t dims[3], int mem);

/* Raster unwrap: the wrapped differences are integrated down the first column
 * (minus 2 pi step2[n2 x n3]) and then along every row (minus 2 pi
 * step1[n1 x n2 x n3]); zero steps unwrap a residue-free map.  Replaces
 * unwrapp_phase_goldstein_gpu (goldstein_GPU/unwrapp_phase_goldstein_mex.cpp,
 * unwrapp_phase_goldstein_cuda.cu:60-113).  unwrapped may alias wrapped. */
int cbs_phase_integrate(int device, const float* wrapped, const int32_t* step1, const int32_t* step2, float* unwrapped,
                        const uint32_t dims[3], int mem);

/* Goldstein branch cuts: from the residue map (as cbs_phase_residues gives it, not
 * modified), the step maps that make cbs_phase_integrate unwrap around the cuts:
 * step1[n1 x n2 x n3] and step2[n2 x n3] are overwritten.  Dipoles (opposite residues in
 * adjacent loops) are cancelled first; the remaining residues are joined to their nearest
 * residues or to the border until every tree balances, with the reference's search order,
 * and the branches are rastered into the step maps.  The dipole pass, the nearest-residue
 * lists and the raster run on the device, the branch search on the host.  Replaces
 * remove_dipoles_gpu and goldstein_branch_cut (Codes/sub_code/goldstein_GPU/
 * remove_dipoles_cuda.cu:50-130, goldstein_branch_cut.m:1-87, get_shortest_sort_cuda.cu,
 * get_branches_goldstein_compute_v2.cpp:16-203, raster_branches_cuda.cu:65-135). */
int cbs_phase_branch_cuts(int device, const float* residues, int32_t* step1, int32_t* step2, const uint32_t dims[3],
                          int mem);

/* The whole unwrap of n3 column-major n1 x n2 wrapped phase maps: residues, branch cuts,
 * raster integration.  Replaces unwrapp2_gpu (goldstein_GPU/unwrapp2_gpu.m:1-5), as
 * called by BACKWARD_SOLVER_RYTOV.m:49.  unwrapped may alias wrapped. */
int cbs_phase_unwrap(int device, const float* wrapped, float* unwrapped, const uint32_t dims[3], int mem);

/* FP64 per-iteration |f_n|^2 of the last solve (track_norms=1, or a convergence test
 * set): out[r * n_iter + j], summed over the field components. */
int cbs_get_norms(cbs_handle h, double* out, uint32_t n_rhs, uint32_t n_iter);

/* FP64 convergence test (no counterpart in the reference, whose loop always runs
 * Bornmax iterations, FORWARD_SOLVER_CONVERGENT_BORN.m:577): with tol > 0, rhs r stops
 * after the first iteration n >= 3 at which |f_n|_2 <= tol |F_1 + F_2|_2, the norm of
 * the series term against the first-order field (both summed in FP64 on the device
 * over the simulation grid and the field components).  The test is made on the device
 * by the next launches (a stopped rhs leaves Field and f unchanged), so no launch
 * waits for the host and the iteration loop stays one captured graph.  tol = 0 (the
 * default) turns it off.  Norms are then tracked as with track_norms. */
int cbs_set_convergence(cbs_handle h, double tol);

/* Iterations each rhs of the last solve ran (Bornmax unless the convergence test
 * stopped it earlier): out[r], r < n_rhs, n_rhs equal to the last solve's. */
int cbs_get_iterations(cbs_handle h, uint32_t* out, uint32_t n_rhs);

/* Per-kind kernel time of the last solve (time_kernels=1):
 * out_ms[k], out_count[k] for k = 0 pass_real, 1 pass_kspace, 2 other. */
int cbs_get_kernel_times(cbs_handle h, double* out_ms, uint64_t* out_count, uint32_t n);

/* The HIP stream (hipStream_t) the handle launches on. */
int cbs_get_stream(cbs_handle h, void** stream);

/* Device bytes a solve of n_rhs right-hand sides needs (state + workspace). */
int cbs_footprint(cbs_handle h, uint32_t n_rhs, uint64_t* bytes);

/* Kernel plan chosen for the current options: form, a, b, tile_t. */
int cbs_get_plan(cbs_handle h, int32_t* form, int32_t* a, int32_t* b, int32_t* tile_t);

/* Autotuner (no counterpart in the reference, whose cuFFT plan is fixed): times
 * every fused kernel plan available for the grid on the current potential with
 * n_rhs right-hand sides and `iterations` Born iterations each, and keeps the
 * fastest as the handle's plan (overriding split_a/tile_t/threads options); then
 * times that plan with n_rhs/2, n_rhs/4, n_rhs/8, 2 and 1 rhs per launch and sets
 * max_batch to a narrower width when it solves the n_rhs at least 3 % faster.
 * Leaves the alternating-shift state untouched.  out_ms (may be null) receives
 * the best time per solve of n_rhs of the trial. */
int cbs_autotune(cbs_handle h, uint32_t n_rhs, uint32_t iterations, double* out_ms);

const char* cbs_last_error(void);
int cbs_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CBS_HIP_H */
