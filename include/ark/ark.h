/* ark.h -- C-ABI of the MI355X-native CABARET/ARK step.
 *
 * One context = one rank's subdomain of the reference solver
 * (albanul/ark-cpp-mpi), resident in HBM on one GPU, advanced by
 * hand-written gfx950 kernels on one HIP stream.  The reference's MPI halo
 * exchange becomes an on-device periodic ghost fill.
 *
 * Every entry point replaces one reference function or global; the citation
 * on each names it (paths relative to the reference checkout).  Fields are
 * exchanged with the caller in the reference's own layout: a full x-fastest
 * array with the reference's extents (source/Arr3d.cpp:19-21,
 * source/main.cpp:198-277), i.e. exactly what Arr3d::ref() points to.
 *
 * Errors: every call returns an ark_status; no exception crosses the ABI.
 * The context keeps a sticky message for the last failure (ark_last_error).
 * Threading: one host thread per context.  Calls are asynchronous on the
 * context's stream except where marked "synchronises".
 */
#ifndef ARK_ARK_H
#define ARK_ARK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ARK_OK = 0,
    ARK_ERR_INVALID = 1,      /* bad argument, extent or field id */
    ARK_ERR_HIP = 2,          /* a HIP runtime call failed */
    ARK_ERR_UNSUPPORTED = 3,  /* valid input this build does not implement */
    ARK_ERR_NONFINITE = 4,    /* a NaN/Inf was found by a diagnostic */
    ARK_ERR_STATE = 5         /* call out of order for the context's state */
} ark_status;

/* Which CABARET variant to run.  REFERENCE reproduces the reference bit for
 * bit, including its X1/X3 sweeps taking gq's time derivative from the
 * R invariant (source/main.cpp:1804, 2321; X2 uses Q, 2083).  On a periodic
 * Taylor-Green vortex that variant grows spurious velocities after some
 * hundred steps.  CORRECTED takes gq from Q on every axis, as the reference's
 * X2 sweep does; it has no counterpart in the reference. */
typedef enum {
    ARK_SCHEME_REFERENCE = 0,
    ARK_SCHEME_CORRECTED = 1
} ark_scheme;

typedef enum {
    ARK_PREC_FP64 = 0,   /* every operation in IEEE binary64, bitwise equal to the reference */
    ARK_PREC_MIXED = 1   /* FP32 fluxes/invariants/stress, FP64 closing update and sums */
} ark_precision;

/* The reference's arrays, in the reference's naming (headers/bulk.h,
 * headers/forces.h, headers/stress.h).  Cell fields have extents
 * (N1+2)(N2+2)(N3+2), face and stress fields (N1+3)(N2+3)(N3+3). */
typedef enum {
    ARK_roCon, ARK_u1Con, ARK_u2Con, ARK_u3Con, ARK_tCon,
    ARK_ronCon, ARK_u1nCon, ARK_u2nCon, ARK_u3nCon, ARK_tnCon,
    ARK_f1, ARK_f2, ARK_f3,
    ARK_condition,
    ARK_ro1, ARK_t1, ARK_u11, ARK_u21, ARK_u31, ARK_p1,
    ARK_ro2, ARK_t2, ARK_u12, ARK_u22, ARK_u32, ARK_p2,
    ARK_ro3, ARK_t3, ARK_u13, ARK_u23, ARK_u33, ARK_p3,
    ARK_sigm11, ARK_sigm21, ARK_sigm31,
    ARK_sigm12, ARK_sigm22, ARK_sigm32,
    ARK_sigm13, ARK_sigm23, ARK_sigm33,
    ARK_NUM_FIELDS
} ark_field;

/* Replaces the reference's scalar globals (headers/constants.h:20,
 * headers/bulk.h:17-34) for one rank: N_d = n_d - 1 interior cells per axis,
 * grid steps, CFL, VIS, sound, ro0_g, t0.  periodic[d] = 1: the reference's
 * MPI exchange with periods[d] = 1 on one rank (no physical-boundary branch
 * on that axis).  periodic[d] = 0: one rank of the reference's own set-up,
 * periods={0,0,0} (source/main.cpp:55-61): the exchange along d delivers
 * zeros (MPI_PROC_NULL neighbours; the reference copies receive buffers that
 * nothing wrote) and the physical-boundary branches of d run (X1: zero-gradient
 * ends and the inflow/outflow invariants, source/main.cpp:803-823, 1633-1658,
 * 1729-1745; X3: zero-gradient ends, 825-845, 1660-1698; X2 has none: walls
 * are solid cells in `condition`).  Non-periodic axes and solid cells run on
 * the FP64 phase path; the fused and mixed-precision steps need a fluid,
 * triply periodic box. */
typedef struct {
    int N1, N2, N3;
    double dx1, dx2, dx3;
    double CFL, VIS, sound, ro0, t0;
    int periodic[3];
    int precision;   /* ark_precision */
    int device;      /* HIP device ordinal */
    int scheme;      /* ark_scheme */
} ark_params;

typedef struct ark_ctx ark_ctx;

/* Diagnostics over interior cells of the level-n fields (nCon), FP64 sums
 * reduced in a fixed order (deterministic). */
typedef struct {
    double ke_reference;   /* sum 1/2|u|^2 (not rho-weighted, no volume): source/main.cpp:2853-2861 */
    double ke_physical;    /* sum 1/2 rho|u|^2 dV */
    double mass;           /* sum rho dV */
    double max_abs_u;      /* max over cells of max_d |u_d| */
    int64_t nonfinite;     /* cells with a NaN/Inf in rho, u1, u2, u3 or t */
    double dt;             /* dt of the last step taken */
    double time;           /* TIME: += dt/2 twice per step (source/main.cpp:84,91) */
    int64_t nstep;         /* nStep (source/main.cpp:83) */
} ark_diag;

/* Allocation done by the reference's Input() (source/main.cpp:198-277):
 * every array on the device, zero-filled. */
ark_status ark_create(const ark_params* params, ark_ctx** out);
/* FreeMemory() (source/main.cpp:2893-2961). */
void ark_destroy(ark_ctx* ctx);
const char* ark_last_error(const ark_ctx* ctx);

/* Copy one Arr3d (host pointer = Arr3d::ref(), extents as constructed) into
 * or out of the context.  Extents must be the reference's for that field.
 * ark_download synchronises.  In a mixed-precision context that has stepped,
 * the ghost cells of the level-n cell fields read as the periodic images of
 * the interior (the mixed step itself reads no ghost). */
ark_status ark_upload(ark_ctx* ctx, int field, const double* host, int e1, int e2, int e3);
ark_status ark_download(ark_ctx* ctx, int field, double* host, int e1, int e2, int e3);

/* The global dt (headers/constants.h:20) and TIME/nStep bookkeeping. */
ark_status ark_set_dt(ark_ctx* ctx, double dt);
ark_status ark_get_dt(ark_ctx* ctx, double* dt);   /* synchronises */
/* The bookkeeping statements of the reference's loop, for a caller that runs
 * the phase functions itself: ark_next_step is nStep++ (source/main.cpp:83),
 * ark_half_time is TIME += 0.5 * dt (84 and 91), on the device in stream
 * order.  ark_step and ark_step_phases do both themselves. */
ark_status ark_next_step(ark_ctx* ctx);
ark_status ark_half_time(ark_ctx* ctx);

/* The step clock of a checkpoint: the reference's TIME and nStep
 * (source/main.cpp:83-84, 91) and dt (headers/constants.h:20), plus the next
 * step's dt when a fused step has already reduced it (then the next step does
 * not read Con).  A checkpoint of a fluid, periodic box is the persistent
 * state -- the level-n cells ronCon..tnCon and, per axis, the faces u1*, u2*,
 * u3*, p*, t* (downloaded) -- and this clock; a restart uploads them into a
 * context of the same parameters and calls ark_restart.  No counterpart in
 * the reference, which cannot restart (its VTK output holds no faces). */
typedef struct {
    double time;          /* TIME */
    int64_t nstep;        /* nStep */
    double dt;            /* dt of the last step taken */
    double dt_next;       /* the next step's dt, when dt_next_known */
    int dt_next_known;    /* 1: reduced by a fused step; 0: the next step's
                             TimeStepSize scans u1Con..u3Con */
} ark_clock;
/* Synchronises. */
ark_status ark_get_clock(ark_ctx* ctx, ark_clock* out);
/* Continue from uploaded arrays: sets the clock; on a fluid, triply periodic
 * FP64 box the face densities become ro0 + p/c^2 of the uploaded pressures
 * (source/main.cpp:1930: what every step leaves there), elsewhere they stay
 * as uploaded.  Clears the non-finite record.  With dt_next_known = 0 the
 * next step computes dt from u1Con..u3Con, which must be uploaded too.
 * Synchronises. */
ark_status ark_restart(ark_ctx* ctx, const ark_clock* clock);

/* TimeStepSize(), source/main.cpp:374-415: dt = min over interior cells of
 * the CFL bounds, on the device; the MPI MIN allreduce is the identity on one
 * GPU.  dt_out may be NULL (then nothing synchronises). */
ark_status ark_time_step_size(ark_ctx* ctx, double* dt_out);
/* Phase1(), source/main.cpp:417-575. */
ark_status ark_phase1(ark_ctx* ctx);
/* StressTensor(), source/main.cpp:577-1024 (the u*Con halo exchange at
 * 603-801 becomes the on-device periodic ghost fill). */
ark_status ark_stress_tensor(ark_ctx* ctx);
/* UseForces(), source/main.cpp:1026-1045. */
ark_status ark_use_forces(ark_ctx* ctx);
/* Phase2(), source/main.cpp:1047-2495 (the 12-field halo exchange at
 * 1049-1631 becomes the on-device periodic ghost fill). */
ark_status ark_phase2(ark_ctx* ctx);

/* nsteps iterations of the reference's driver loop body (source/main.cpp:
 * 82-91: TimeStepSize, Phase1, StressTensor, UseForces, Phase2, Phase1,
 * UseForces) through the fused kernels.  Results are bitwise equal to the
 * phase-by-phase calls in FP64 mode.  Nothing synchronises: dt stays on the
 * device between steps.  Non-finite values: every step counts the cells whose
 * half-level velocity or density is NaN or Inf in its CFL reduction (the
 * reference checks nothing); ark_synchronize, and any stepping call that
 * finds the count of an earlier call already copied back, then return
 * ARK_ERR_NONFINITE, and stepping is refused until a new initial state.  In mixed precision the first step of a context also
 * times its two tile shapes (a few launches that leave the state unchanged)
 * and keeps the faster; results do not depend on the shape. */
ark_status ark_step(ark_ctx* ctx, int nsteps);

/* Tile shape of the mixed-precision step: -1 = chosen by timing at the next
 * step (the default), 0 = 30x14 owned columns, 1 = 32x13.  Every shape gives
 * the same results bit for bit; only the speed differs.  No counterpart in
 * the reference. */
ark_status ark_set_tile_shape(ark_ctx* ctx, int shape);
/* The tile shape in use (-1 before the first mixed-precision step chose it). */
ark_status ark_get_tile_shape(const ark_ctx* ctx, int* shape);

/* nsteps through the fused kernels only, maintaining the persistent state:
 * the level-n cell fields (ronCon, u1nCon, u2nCon, u3nCon, tnCon) on the
 * interior cells, the face fields u*, p*, t* on the faces 1..N_d+1 along
 * their axis, dt, TIME and nStep -- bitwise as the reference.  The
 * reference's scratch arrays (Con, f, sigma), the ghost copies and the face
 * densities ro1..ro3 are NOT brought up to date (the next ark_step or phase
 * call works from the persistent state and refreshes them).  No counterpart in
 * the reference: it is ark_step without the final phase-by-phase step, for
 * callers that only read the persistent fields. */
ark_status ark_advance(ark_ctx* ctx, int nsteps);

/* The same nsteps through the phase-by-phase kernels (one launch per
 * reference phase): the unfused baseline that ark_step is checked against. */
ark_status ark_step_phases(ark_ctx* ctx, int nsteps);

/* WriteEnergy(), source/main.cpp:2851-2876 (without the file write), plus
 * mass and a non-finite scan.  Synchronises. */
ark_status ark_diagnostics(ark_ctx* ctx, ark_diag* out);

/* ark_diagnostics without a context and without allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (work under HIP graph capture, tight serving loops) and hold the level-n
 * cell fields themselves.  cells = device pointers to ronCon, u1nCon,
 * u2nCon, u3nCon, tnCon in the reference's Arr3d layout (extents
 * (N1+2)(N2+2)(N3+2), x fastest); from params only N1..N3, dx1..dx3 and
 * device are read.  out (device, 5 doubles) receives ke_reference,
 * ke_physical, mass, max_abs_u and the non-finite count, the same sums bit
 * for bit as ark_diagnostics on the same fields.  Two kernel launches on
 * `stream` (a hipStream_t, or NULL), nothing synchronises, no hipMalloc: the
 * partial sums use a 40 KiB device array of the library, one per GPU, shared
 * by every call there, so calls that overlap on one GPU race.  Grids with
 * 2^31 or more elements per array are refused (ARK_ERR_UNSUPPORTED). */
ark_status ark_diagnostics_noalloc(const ark_params* params, const double* const cells[5], double* out,
                                   void* stream);

/* Taylor-Green vortex initial state built on the device from 1-D trig tables
 * that are evaluated on the host with the C library's sin/cos, so the fields
 * are bitwise equal to a host evaluation of the same formula.  Replaces the
 * reference's InitializeData() (source/main.cpp:280-372) for the periodic
 * box; Con := nCon, ghosts are periodic copies, f and sigma are zero. */
ark_status ark_init_tgv(ark_ctx* ctx, double U0);
/* The reference's own problem: its InitializeData() (source/main.cpp:280-372)
 * on one rank of its Input() set-up (118-195), for a context with N+1 nodes
 * per axis and the grid steps Input implies (20/200, 6/60, 0.1/1 for the
 * reference's 201x61x2 nodes) and periodic = {0,0,0}: the channel with a
 * backward-facing step (solid where x <= 5 and y <= 1), walls as solid ghost
 * rows in y, u1 = u10 above y = 1; every element InitializeData leaves unset
 * is 0; Con := nCon; TIME and nStep 0. */
ark_status ark_init_channel(ark_ctx* ctx, double u10);

/* Block until all work on the context's stream is done.  Returns
 * ARK_ERR_NONFINITE when a step has met a non-finite value (see ark_step). */
ark_status ark_synchronize(ark_ctx* ctx);
/* The context's hipStream_t, for callers that order their own work on it. */
ark_status ark_get_stream(ark_ctx* ctx, void** stream);

/* Per-kernel timing with HIP events on the context's stream.  When enabled,
 * every launch is bracketed by events; ark_profile_read synchronises and
 * returns, per kernel name, the number of launches and the summed duration. */
ark_status ark_profile_enable(ark_ctx* ctx, int on);
ark_status ark_profile_read(ark_ctx* ctx, int max_entries, char (*names)[64],
                            int64_t* launches, double* total_ms, int* n_entries);
ark_status ark_profile_reset(ark_ctx* ctx);

/* Halo planes, as the reference packs them for MPI_Sendrecv
 * (source/main.cpp:619-653, 1101-1189 and the X2/X3 analogues): plane q
 * normal to `axis` (0..N_d+1 for cell fields, 0..N_d+2 for face fields),
 * the interior 1..N of the two other axes, in the Arr2d layout
 * buf[b*na + a] (source/Arr2d.cpp:14-16) with a the lower of the two other
 * axes; na, nb must equal their interior extents.  A multi-rank host packs
 * its send buffers with ark_get_plane and unpacks the received ones with
 * ark_set_plane.  Both synchronise.  Face fields are FP64-only. */
ark_status ark_get_plane(ark_ctx* ctx, int field, int axis, int q, double* host, int na, int nb);
ark_status ark_set_plane(ark_ctx* ctx, int field, int axis, int q, const double* host, int na, int nb);

/* BoundaryConditions::alongX1/2/3() (source/BoundaryConditions.cpp:22-85) on
 * five cell fields given as field ids {u1, u2, u3, ro, t}: ghost planes 0
 * and N_d+1 of axis `axis` over the interior of the other two axes.  Along
 * x1 the neighbouring interior plane is copied whatever `periodic` says;
 * along x2 and x3 the planes wrap when periodic, otherwise u = 0, ro is
 * zero-gradient and t = t0; the periodic top plane along x3 takes t from
 * ro(i, j, 1), as the reference does.  Enqueued on the context's stream. */
ark_status ark_boundary_conditions(ark_ctx* ctx, int axis, const int fields[5], int periodic, double t0);

/* Batched steps (no counterpart in the reference): many independent boxes,
 * each an ordinary mixed-precision context of its own size and parameters
 * (same device and scheme), advanced together by ONE kernel launch per step
 * whose workgroups cover the tiles of every box.  The steps are replayed from
 * HIP graphs captured at the first ark_batch_step (and again whenever a box
 * has stepped on its own since).  Every box evolves exactly as ark_step on
 * its own context would evolve it.  The boxes stay usable through their own
 * contexts between batch calls; the batch orders its work after theirs and
 * theirs after its own.  A context must outlive the batches that hold it. */
typedef struct ark_batch ark_batch;
ark_status ark_batch_create(ark_ctx* const* boxes, int nboxes, ark_batch** out);
void ark_batch_destroy(ark_batch* batch);
const char* ark_batch_last_error(const ark_batch* batch);
ark_status ark_batch_step(ark_batch* batch, int nsteps);
ark_status ark_batch_synchronize(ark_batch* batch);
/* Times every (tile shape, chunk length) candidate on the batch's own boxes,
 * reps launches each, and keeps the fastest.  Tile shapes: 0 = 30x6 owned
 * (256 threads), 1 = 30x14 (512), 2 = 62x6 (512), 3 = 32x13 (510); chunk
 * lengths 16, 32, 64 planes.  ms_out[shape*3 + lz_index] receives each candidate's ms per step
 * (max_out entries).  The boxes' state is not changed. */
ark_status ark_batch_autotune(ark_batch* batch, int reps, double* ms_out, int max_out);
ark_status ark_batch_set_config(ark_batch* batch, int shape, int lz);
ark_status ark_batch_get_config(const ark_batch* batch, int* shape, int* lz, int* tiles, int* threads);

/* Homogeneous turbulence (no counterpart in the reference).
 *
 * ark_init_turbulence: a seeded solenoidal velocity field with the shell
 * spectrum E(k) ~ k^4 exp(-2 (k/k0)^2) and per-component rms `urms`, uniform
 * density rho0, zero pressure, t = t0; face values are the same Fourier
 * series evaluated on the faces (exactly periodic: face N+1 is face 1).  The
 * field is Gaussian white noise from Philox4x32-10 (key = seed, counter =
 * (cell index, 0x7a3c51, 0)), transformed, shaped by |k| exp(-(|k|/k0)^2)
 * per mode, projected normal to k, zero mean and Nyquist modes, scaled.
 * Needs even N.  TIME and nStep restart at 0.
 *
 * ark_spectrum: E[b] = sum over integer-index shells b = round(|m|) of
 * 1/2 |u_hat|^2 / N^6, b < nbins; sum_b E[b] = 1/2 <u.u> over the cells.
 *
 * ark_dissipation: nu <sum_ij (d_j u_i)^2> with central differences.
 *
 * ark_set_forcing (mixed precision): random-phase forcing of the 16 pairs of
 * wavevectors with 1 <= |m|^2 <= 4.  With tau = 0 every step gives every
 * mode a fresh random amplitude vector normal to k (Philox, key = seed,
 * counter = (mode, nStep, 0xf0c3, 0|1)) of norm sqrt(eps/(16 dt)); with
 * tau > 0 the amplitudes are redrawn once per interval of length tau
 * (counter = the interval of the step's mid time) with norm
 * sqrt(eps/(16 tau)), so runs of different resolution see the same force
 * history.  Either way the mean injection rate is eps per unit mass.
 * eps = 0 switches it off.
 *
 * ark_filter_box: the box filter of width w cells (w divides N) of the
 * level-n velocity: ubar (3 x Nc), the subgrid stresses tau = <u_i u_j> -
 * ubar_i ubar_j (6 x Nc, ij = 11 22 33 12 13 23) and the central-difference
 * gradients of ubar on the coarse grid (9 x Nc, component 3i + j = d_j ubar_i),
 * Nc = (N1/w)(N2/w)(N3/w) values each, x fastest.  All synchronise. */
ark_status ark_init_turbulence(ark_ctx* ctx, unsigned long long seed, double k0, double urms);
ark_status ark_spectrum(ark_ctx* ctx, int nbins, double* E);
ark_status ark_dissipation(ark_ctx* ctx, double* eps);
ark_status ark_set_forcing(ark_ctx* ctx, unsigned long long seed, double eps, double tau);
ark_status ark_filter_box(ark_ctx* ctx, int width, double* ubar, double* tau, double* grad);

/* Learned subgrid closure (no counterpart in the reference): an MLP per cell,
 * 10 inputs (the 9 filtered velocity-gradient components d_j u_i, index
 * 3i + j, and the filter width) -> 64 -> 64 -> 6 subgrid stresses
 * (11 22 33 12 13 23), SiLU, BF16 weights and activations with FP32
 * accumulation on the matrix cores; FP32 master weights and Adam
 * (beta 0.9, 0.999, eps 1e-8) for training.  Inputs and targets are in raw
 * units; ark_mlp_set_normalization gives the per-feature shift and scale the
 * network sees.  Parameters (ark_mlp_param_count() floats): W1[64][10] b1[64]
 * W2[64][64] b2[64] W3[6][64] b3[6].  The loss is the mean squared error of
 * the normalised outputs.  Training batches are drawn with replacement from
 * the data set (Philox, key = seed, counter = (position, step)). */
typedef struct ark_mlp ark_mlp;
int ark_mlp_param_count(void);
ark_status ark_mlp_create(int device, unsigned long long seed, ark_mlp** out);
void ark_mlp_destroy(ark_mlp* mlp);
const char* ark_mlp_last_error(const ark_mlp* mlp);
ark_status ark_mlp_set_params(ark_mlp* mlp, const float* params);
ark_status ark_mlp_get_params(ark_mlp* mlp, float* params);
ark_status ark_mlp_set_normalization(ark_mlp* mlp, const float* in_mean, const float* in_std,
                                     const float* out_mean, const float* out_std);
ark_status ark_mlp_set_data(ark_mlp* mlp, const float* X, const float* T, long long n);
/* Gradient of the loss over all n samples (X: n x 10, T: n x 6, host). */
ark_status ark_mlp_gradient(ark_mlp* mlp, const float* X, const float* T, long long n, double* grad,
                            double* loss);
ark_status ark_mlp_train(ark_mlp* mlp, int nsteps, long long batch, float lr, double* last_loss);
ark_status ark_mlp_predict(ark_mlp* mlp, const float* X, float* Y, long long n);

/* Large-eddy simulation with the learned closure (mixed precision).  With a
 * network attached, every step computes on the level-n cells the central-
 * difference velocity gradient (inputs 0..8) and `delta` (input 9, the
 * filter width the network was trained for), runs the network per cell and
 * adds the acceleration -d_j tau_ij (central differences) to the body force
 * (with the forcing, if any).  NULL detaches.  The network must outlive the
 * attachment; a context with a closure cannot join a batch. */
ark_status ark_set_closure(ark_ctx* ctx, ark_mlp* mlp, double delta);
/* Scale-selective closure (variational multiscale, Hughes, Mazzei & Jansen,
 * Comput. Visual. Sci. 3 (2000) 47-59): with kcut > 0 only the Fourier modes
 * of the closure's acceleration in shells round(|m|) >= kcut (m the integer
 * wavevector) are added to the body force, so the network acts on the small
 * resolved scales only; kcut = 0 (the default) adds all of it.  FP32
 * transforms on the context's stream, three each way per step. */
ark_status ark_set_closure_cutoff(ark_ctx* ctx, int kcut);
/* Initial state from host cell velocities u (3 x N1*N2*N3 interior values,
 * component-major, x fastest): the cells take them, the faces the same
 * Fourier series half a cell down (as ark_init_turbulence); uniform rho0 and
 * t0, zero pressure.  Needs even N.  TIME and nStep restart at 0. */
ark_status ark_init_velocity(ark_ctx* ctx, const double* u);
/* ark_init_velocity with a density: rho (N1*N2*N3 interior values, x
 * fastest) on the cells and, by the same half-cell shift of rho - rho0, on
 * the faces, with the face pressures c^2 (rho - rho0) (the equation of state,
 * source/main.cpp:1774-1775, 1930).  NULL rho is ark_init_velocity. */
ark_status ark_init_state(ark_ctx* ctx, const double* u, const double* rho);
/* The box filter of width w (dividing N) of the level-n cell density:
 * (N1/w)(N2/w)(N3/w) block means, x fastest.  Synchronises. */
ark_status ark_filter_density(ark_ctx* ctx, int width, double* rhobar);

/* The reference's output path (source/main.cpp:2577-2891), for one rank.
 * ark_write_paraview replaces WriteDataParaView (2577-2745): the file
 * <dir_path>out_<nStep>.vtk, binary legacy VTK on the rectilinear grid of the
 * N+1 nodes per axis ((i-1) dx), with the level-n velocity (U), density (PC)
 * and the vorticity from the face velocities (R, the reference's expression,
 * 2712-2733) of every interior cell; need_swap writes big-endian as the
 * reference's swap8 with needSwap = true (120, 2975-2988).
 * ark_average_values replaces AverageValues (2878-2891): averageU += u * dt
 * on the device, the accumulators zero at the first call.
 * ark_write_average_values replaces WriteAverageValues (2747-2849): the file
 * <dir_path>average.vtk with averageU / (TIME - start_time).
 * The files are byte for byte the reference's for the same state; they are
 * created or truncated. */
ark_status ark_write_paraview(ark_ctx* ctx, const char* dir_path, int need_swap);
ark_status ark_average_values(ark_ctx* ctx);
ark_status ark_write_average_values(ark_ctx* ctx, const char* dir_path, double start_time, int need_swap);

/* Bytes of device memory held by the context. */
ark_status ark_device_bytes(const ark_ctx* ctx, int64_t* bytes);

#ifdef __cplusplus
}
#endif
#endif /* ARK_ARK_H */
