/*
 * eup.h -- C-ABI of the MI355X D3Q19 BGK hot path (libeup.so).
 *
 * Drop-in for the per-time-step routines of the sponge-flow solver's serial/OpenMP path.  In the
 * reference these are argument-less Fortran subroutines sharing module state (`use storage`,
 * CODE/SRC/storage.F90:26-437); here the state lives on the GPU behind an opaque handle and every
 * call takes plain pointers and sizes, so a Fortran host binds it with ISO_C_BINDING (see
 * euplectella_hpc_amd/fortran/eup_mod.F90) and Python with ctypes.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative EUP_E* code; the message is available
 *     from eup_last_error(handle) (or eup_last_error(NULL) for failures without a handle).  This
 *     replaces the reference's `write(6,*) "ERROR..."; stop`.
 *   - One handle = one GPU + one HIP stream; calls on a handle are synchronous with respect to the
 *     host arrays they take (the host may reuse a buffer when a call returns).  A handle is not
 *     thread-safe; use one handle per host thread.
 *   - Host arrays are owned by the caller and never retained.  Device memory is owned by the
 *     handle and released by eup_destroy.
 *   - Lattice indices are 1-based like the reference: interior sites i = 1..l, j = 1..m,
 *     k = 1..n; ghost layers 0 and l+1 (m+1, n+1).  Ghosted host arrays have
 *     (l+2)(m+2)(n+2) elements with i fastest (Fortran order of a01..a19, alloca.F90:94-133).
 *   - Population q is 1..19 with the reference's direction table (bgk3d.F90:31-49).
 */
#ifndef EUP_H
#define EUP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EUP_ABI_VERSION 1

/* error codes */
#define EUP_OK 0
#define EUP_EINVAL (-1)   /* invalid argument or configuration */
#define EUP_EHIP (-2)     /* HIP runtime error (no device, out of memory, launch failure) */
#define EUP_ESTATE (-3)   /* call order violated (e.g. step before mask and populations are set) */
#define EUP_ERANGE (-4)   /* index outside the lattice */

/* population storage on the device */
#define EUP_STORE_F64 0   /* full f in FP64 (parity mode requires this) */
#define EUP_STORE_F32 1   /* full f in FP32, FP32 collision */
#define EUP_STORE_F16 2   /* deviation f_q - w_q in IEEE binary16, FP32 collision */
#define EUP_STORE_BF16 3  /* deviation f_q - w_q in bfloat16, FP32 collision */
#define EUP_STORE_E3M12 4 /* deviation (f_q - w_q) / w_q as a 16-bit float with 3 exponent and 12
                             mantissa bits (range below 4, step 2^-17 near zero; out of range
                             stores NaN), FP32 collision */

/* time-step semantics */
#define EUP_MODE_PARITY 0  /* the reference's serial step, bit for bit (FP64, ghosted arrays) */
#define EUP_MODE_PHYSICS 1 /* fused pull collide+stream: y/z periodic by index wrap, bounce-back
                              from the site's own opposite population, solid sites frozen */
#define EUP_MODE_GREY 2    /* physics semantics on a solid-fraction field s in [0,1] (partial
                              bounce-back per link), with the discrete adjoint of the drag;
                              storage EUP_STORE_F64 or EUP_STORE_F32 */

/* face codes of build.bcond.F90 (0 periodic, 4 inflow, 5 outflow) */
#define EUP_FACE_PERIODIC 0
#define EUP_FACE_INFLOW 4
#define EUP_FACE_OUTFLOW 5

/* flags */
#define EUP_FLAG_NO_TILE_SKIP 1u   /* physics mode: process every tile (disables the all-solid
                                      tile skip and the pure-fluid fast path) */
#define EUP_FLAG_SHIFTED_DEVIATION 2u /* the 16-bit storages (FP16, BF16, e3m12) hold
                                         f_q - f_eq_q(1, (u_inflow, 0, 0)) instead of f_q - w_q
                                         (e3m12: divided by w_q): smaller stored values in the
                                         free stream, smaller rounding error */

typedef struct eup_handle eup_handle;

typedef struct eup_config {
    int32_t l, m, n;          /* interior lattice sizes (lx, ly, lz of the serial deck) */
    int32_t mode;             /* EUP_MODE_* */
    int32_t storage;          /* EUP_STORE_* */
    int32_t front, rear;      /* x = l+1 and x = 0 face codes (storage.F90 front(1), rear(1)) */
    int32_t up, down, left, right; /* z and y faces: must be EUP_FACE_PERIODIC */
    int32_t device;           /* HIP device ordinal */
    uint32_t flags;           /* EUP_FLAG_* */
    double omega;             /* 2/(6*svisc + 1), computed by the caller as hencol.F90:30 */
    double fgrad;             /* x body-force coefficient (hencol.F90:37; 0 for CYLINDER_INF); grey
                                 mode has no body force and rejects fgrad != 0 */
    double u_inflow;          /* inlet velocity (build.bcond.F90:324) */
} eup_config;

typedef struct eup_info {
    int64_t sites;            /* l*m*n */
    int64_t nsolid;           /* solid sites */
    int64_t tiles, tiles_fluid, tiles_mixed, tiles_solid; /* 64-site tiles (physics mode) */
    int32_t bbox[6];          /* imin, imax, jmin, jmax, kmin, kmax (build.obs.F90:155-163) */
    int32_t bytes_per_site;   /* algorithmic HBM bytes per updated site per step */
    int64_t device_bytes;     /* device memory allocated by eup_create (two state buffers, staging,
                                 solid flags, outlet planes); per-mask tables and diagnostic
                                 scratch allocated later are not counted */
} eup_info;

/* --- lifetime ------------------------------------------------------------------------------ */

/* Replaces setup/alloca/build_bcond/hencol bookkeeping (setup.F90:22, alloca.F90:91-133,
 * build.bcond.F90:319-339): validates the configuration and allocates the device state. */
int eup_create(const eup_config *cfg, eup_handle **out);

/* Releases every device allocation of the handle (dealloca.F90). */
void eup_destroy(eup_handle *h);

/* Message of the last error on this handle (h == NULL: last error of the calling thread). */
const char *eup_last_error(const eup_handle *h);

/* --- state --------------------------------------------------------------------------------- */

/* Solid mask obs(1:l,1:m,1:n), int32, i fastest, 1 = solid -- the array build_obs fills
 * (build.obs.F90:147-186).  Computes the obstacle bounding box and the tile classes. */
int eup_set_mask(eup_handle *h, const int32_t *obs);

/* The obstacle list itself, streamed in chunks, without a dense host array: the file
 * obs.RRRR.TTTTTTTT.bin holds nobs records of (i, j, k), 1-based (read by build.obs.F90:147-186,
 * which fills obs and the bounding box while reading).  eup_add_obstacles marks count sites
 * ijk[3*t .. 3*t+2] solid; the first call after eup_create, eup_set_mask or a commit starts from
 * an all-fluid lattice, later calls add to it (sites may repeat).  eup_commit_obstacles then counts
 * the solid sites and their bounding box on the GPU and prepares the mode's mask data, as
 * eup_set_mask does; a commit without any add gives a lattice without obstacles.  Until the
 * commit the handle has no mask (eup_step fails with EUP_ESTATE).  A list with a site outside the
 * lattice is rejected (EUP_ERANGE) before anything changes.  Not for grey mode. */
int eup_add_obstacles(eup_handle *h, const int32_t *ijk, int64_t count);
int eup_commit_obstacles(eup_handle *h);

/* Population q (1..19) as a ghosted FP64 host array: the record restore_raw reads
 * (restore_raw.F90:43-63).  In physics mode only the interior is used. */
int eup_set_pops(eup_handle *h, int32_t q, const double *a);

/* Population q (1..19) into a ghosted FP64 host array: the record save_raw writes
 * (save_raw.F90:54-74).  Physics mode writes 0 into the ghost layer. */
int eup_get_pops(eup_handle *h, int32_t q, double *a);

/* Every site (ghosts included in parity mode) set to the equilibrium of rho = 1 and
 * u = (ux, uy + noise * r, uz), r uniform in [-1, 1) from a counter-based hash of (seed, site).
 * With ux = uy = uz = noise = 0 this is the reference's rest start (init.F90:128-236, opt 0);
 * otherwise it is the device-side counterpart of a restore file holding an inlet-velocity start.
 * Equilibrium form w*(1 + (3 c.u + 1.5 (3 (c.u)^2 - u^2))). */
int eup_init_equilibrium(eup_handle *h, double ux, double uy, double uz, double noise, uint64_t seed);

/* All 19 populations at once: a[(q-1)*(l+2)(m+2)(n+2) + offset]. */
int eup_set_state(eup_handle *h, const double *a);
int eup_get_state(eup_handle *h, double *a);

/* --- time stepping ------------------------------------------------------------------------- */

/* New relaxation frequency for the following steps: the reference reaches low viscosity by
 * restarting from the dump of a higher-viscosity run with a new svisc in the deck (the Re chain
 * of CODE/RUN/README:15-21, omega recomputed by hencol.F90:30); this keeps the state on the GPU. */
int eup_set_omega(eup_handle *h, double omega);

/* nsteps iterations of `boundaries; propagation; collision` (bgk3d.F90:128-138). */
int eup_step(eup_handle *h, int32_t nsteps);

/* As eup_step, and reports the GPU time of the nsteps (HIP events on the handle's stream). */
int eup_step_timed(eup_handle *h, int32_t nsteps, double *gpu_ms);

/* Blocks until all work on the handle's stream is done. */
int eup_sync(eup_handle *h);

/* --- diagnostics (FP64) -------------------------------------------------------------------- */

/* Control-volume stress drag of flux_X/flux_Y (flux_X.F90:50-301, flux_Y.F90:51-282) on the box
 * {xstart, xstop, ystart, ystop, zstart, zstop} (global 1-based indices, drag.F90:48-58).  Per-site
 * stresses are computed on the GPU and summed in the reference's loop order, so the result equals
 * the reference's to the last bit on the same populations.  A box whose x (or y) range ends at 0
 * puts a face on plane 0, the reference's ghost plane: parity mode reads it there, the physics and
 * grey layouts have no plane 0 and return EUP_ERANGE. */
int eup_drag_cv(eup_handle *h, const int32_t box[6], double *fluxX, double *fluxY);

/* eup_drag_cv without any device allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same box, same results bit for bit; the per-site stresses go to a
 * fixed device array of the library (2 x 1 048 576 doubles = 16 MiB per GPU, kept until the
 * process ends) instead of the handle's scratch, and are summed on the host through a fixed
 * buffer.  A box whose clipped x (or y) face holds more than 1 048 576 sites is rejected with
 * EUP_EINVAL before anything runs.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (several host threads, several handles on the same device) race. */
int eup_drag_cv_noalloc(eup_handle *h, const int32_t box[6], double *fluxX, double *fluxY);

/* Momentum-exchange force on the solid, F = sum over fluid sites x and links p with x + c_p solid
 * of 2 c_p f_p(x), from the current (post-collision) populations; wavefront-reduced in FP64.
 * No reference counterpart (the reference only has eup_drag_cv's control-volume drag). */
int eup_drag_mem(eup_handle *h, double F[3]);

/* Velocity probe of probe_global.F90:41-77 at interior site (i, j, k): out = {u, v, w, rho}. */
int eup_probe(eup_handle *h, int32_t i, int32_t j, int32_t k, double out[4]);

/* Global means of diagno.F90:53-131: out = {<rho>, <u>, <v>, <w>, <u^2>}, solid sites counted as
 * rho = 1, u = 0.  FP64 reduction on the GPU (fixed order; not the reference's sequential order). */
int eup_diagno(eup_handle *h, double out[5]);

/* Fields of the reference's binary VTK writers, for a host that writes those files (header and
 * big-endian data layout: vtk_xy.binary.F90:43-72).  Velocity of a plane or of the lattice, three
 * floats (u_x, u_y, u_z) per site in the writer's loop order:
 *   EUP_VTK_YZ  x = index, loops k, j      (vtk_yz.binary.F90:90-118, vtk_yz_bin(itime, i0))
 *   EUP_VTK_XZ  y = index, loops k, i      (vtk_xz.binary.F90:90-118, vtk_xz_bin(itime, j0))
 *   EUP_VTK_XY  z = index, loops j, i      (vtk_xy.binary.F90:90-118, vtk_xy_bin(itime, k0))
 *   EUP_VTK_3D  whole lattice, loops k, j, i; index ignored (vtk_3d.binary.F90:78-104)
 * computed as the writers do (momenta and density summed in FP64 in their order, rounded to
 * float, divided in float), so a parity-mode state gives the reference's files bit for bit.
 * out holds 3 * (sites of the plane) floats; 1 <= index <= the plane's normal extent. */
#define EUP_VTK_YZ 0
#define EUP_VTK_XZ 1
#define EUP_VTK_XY 2
#define EUP_VTK_3D 3
int eup_vtk_velocity(eup_handle *h, int32_t plane, int32_t index, float *out);

/* Vorticity of the xy plane z = k0 as vtk_om.binary.F90:88-131 computes it, including the
 * writer's evaluation order: (l-2)*(m-2) floats, i = 2..l-1 fastest, j = 2..m-1.  l, m >= 3. */
int eup_vtk_vorticity(eup_handle *h, int32_t k0, float *out);

/* --- grey lattice and adjoint (EUP_MODE_GREY; no reference counterpart) --------------------- */
/* The link between x and its source y = x - c_q carries sigma = 1 - (1 - s(x))(1 - s(y)) and the
 * pulled population is (1 - sigma) a_q(y) + sigma a_qbar(x); s in {0, 1} is the physics mode's
 * bounce-back.  Force F = sum over interior links of sigma c_q (a_q(y) - a_qbar(x)).             */

/* Solid fraction s(1:l,1:m,1:n), i fastest, values in [0, 1] (host FP64 array). */
int eup_set_grey(eup_handle *h, const double *s);

/* As eup_set_grey from device memory of the handle's device; dtype EUP_STORE_F64 or _F32. */
int eup_set_grey_device(eup_handle *h, const void *s, int32_t dtype);

/* Force on the grey body from the current state. */
int eup_grey_force(eup_handle *h, double F[3]);

/* J = sum_{t=t0..t1} F_x(a^t) / ((t1 - t0 + 1) * norm), a^0 = the current state (left unchanged),
 * and, if grad is not NULL, dJ/ds(1:l,1:m,1:n) by the discrete adjoint; forward states are
 * checkpointed every `every` steps and recomputed per segment.  Host FP64 gradient. */
int eup_grey_objective(eup_handle *h, int32_t t0, int32_t t1, int32_t every, double norm, double *J,
                       double *grad);

/* As eup_grey_objective with the gradient written to device memory (dtype EUP_STORE_F64/_F32). */
int eup_grey_objective_device(eup_handle *h, int32_t t0, int32_t t1, int32_t every, double norm,
                              double *J, void *grad, int32_t dtype);

/* --- introspection ------------------------------------------------------------------------- */

int eup_get_info(eup_handle *h, eup_info *info);

/* The handle's HIP stream (hipStream_t) as an opaque pointer, for callers that order their own
 * work (timing events, PyTorch) with the library's. */
void *eup_stream(eup_handle *h);

/* HBM bandwidth probe on `device`: GB/s of a copy (read + write), a read-only and a write-only
 * stream over `bytes` per buffer, `reps` repetitions each: out = {copy, read, write}.
 * Measurement utility (no reference counterpart). */
int eup_membw(int32_t device, int64_t bytes, int32_t reps, double out[3]);

/* ABI version of the loaded library (EUP_ABI_VERSION at build time). */
int32_t eup_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* EUP_H */
