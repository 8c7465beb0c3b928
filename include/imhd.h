/* imhd: MI355X-native isothermal-MHD cell-update library (C ABI).
 *
 * Drop-in boundary for the hot path of tsultan16/Isothermal_MHD (3D_MPI).  The reference
 * exposes Fortran module procedures that act on module arrays (q_3d, bface_3d,
 * qpass_3d; 3D_MPI/src/grid_data_mod.f90:9-14); each entry point below replaces one of
 * them and is bound from Fortran through iso_c_binding (fortran/imhd_shim.f90).
 *
 * Host arrays use the reference layout: FP64, Fortran column-major, SoA by field,
 * ghost-padded with nb layers:
 *   q_3d(1-nb:nx+nb, 1-nb:ny+nb, 1-nb:nz+nb, 7)  rho, rho vx, rho vy, rho vz, Bx, By, Bz (cell)
 *   bface_3d(same, 3)                              Bx(i+1/2), By(j+1/2), Bz(k+1/2)
 *   qpass_3d(3-nb:nx+nb-2, ..., npass)             passive scalars
 * All functions return 0 on success and a negative IMHD_E* code on error;
 * imhd_last_error() then describes it (the reference's PRINT + STOP becomes a return code).
 * A context owns one HIP stream and its device buffers; it is not thread-safe.
 */
#ifndef IMHD_H
#define IMHD_H
#ifdef __cplusplus
extern "C" {
#endif

#define IMHD_OK 0
#define IMHD_EINVAL (-1)   /* bad argument / configuration */
#define IMHD_EHIP (-2)     /* HIP runtime error (allocation, launch, copy) */
#define IMHD_ESTATE (-3)   /* call out of order (e.g. no state set) */
#define IMHD_ENOSPEED (-4) /* max speed below 1e-20 (vanleer_mod.f90:3328-3331) */
#define IMHD_EIO (-5)      /* file: cannot open, short read/write, or records that do not fit the context */

/* integrator: reference mhd_solver_type (constants_mod.f90:58) */
#define IMHD_TVD 1
#define IMHD_VL 2
/* riemann: reference riemann_solver_type as read by the VL module (constants_mod.f90:57).
 * With IMHD_TVD: IMHD_ROE selects the K98 Roe-TVD flux (the TVD module's type 2) and
 * IMHD_HLLE its HLLE flux (type 1) (isotvd_mod.f90:400-428); IMHD_HLLD is rejected. */
#define IMHD_ROE 1
#define IMHD_HLLD 2
#define IMHD_HLLE 3
/* recon: corrector-stage reconstruction of the VL integrator */
#define IMHD_PLM_PRIM 0   /* primitive PLM with MC-beta slope (vanleer_mod.f90:1160-1215) */
#define IMHD_PPM_PRIM 1   /* primitive PPM as compiled (vanleer_mod.f90:1217-1334) */
#define IMHD_PPM_CHAR 2   /* characteristic PPM, the reference default (vanleer_mod.f90:941-1141, 1488-1818) */
/* bc: reference boundary_type (constants_mod.f90:36) */
#define IMHD_BC_OPEN 1
#define IMHD_BC_PERIODIC 2

typedef struct imhd_config {
    int nx, ny, nz;     /* interior cells (per box) */
    int nb;             /* ghost layers; the reference uses 6 (constants_mod.f90:12) */
    int npass;          /* passive scalars, 0..4 (constants_mod.f90:63) */
    int integrator;     /* IMHD_VL (vanleer_mod.f90) or IMHD_TVD (isotvd_mod.f90) */
    int riemann;        /* IMHD_ROE / IMHD_HLLD / IMHD_HLLE */
    int recon;          /* IMHD_PLM_PRIM / IMHD_PPM_PRIM / IMHD_PPM_CHAR (VL only) */
    int bc;             /* IMHD_BC_OPEN / IMHD_BC_PERIODIC */
    int precision;      /* 64: FP64 throughout; 32: FP32 bulk, FP64 face B, corner EMF, CFL, diagnostics */
    int compensated;    /* precision 32 only: Kahan-compensated accumulation of rho, rho v */
    int flux_protection; /* VL + Roe (precision 64 or 32): the reference's flux_protection
                            (vanleer_mod.f90:81, off there): cells whose updated density falls
                            below MIN_DENS take HLLE fluxes and corner EMFs (:2084-2200) */
    double cs;          /* isothermal sound speed (constants_mod.f90:35) */
    double mc_beta;     /* MC limiter beta (vanleer_mod.f90:77), 1.01 in the reference */
    double cour;        /* Courant number for imhd_run (constants_mod.f90:16) */
} imhd_config;

typedef struct imhd_diag {
    double ke;          /* 0.5 sum rho v^2 dx^3 (driver.f90:518-553, same clamps) */
    double vrms_sq;     /* sum v^2 dx^3 */
    double momx, momy, momz;
    double divb_sum;    /* signed raw sum of face differences (driver.f90:492-515) */
    double divb_max;    /* max over cells of |raw face-difference sum| */
    double mass;        /* sum rho dx^3 */
} imhd_diag;

typedef struct imhd_ctx imhd_ctx;

/* Context for one box, or an ensemble of nbox boxes (cfg[b] per box; all boxes share the
 * solver options of cfg[0]).  Replaces init_solver_vanleer (vanleer_mod.f90:127-158) or
 * init_solver_tvd (isotvd_mod.f90:122-133, 144-202), init_passive_solver (passive_mod.f90:44-57)
 * and create_grid_arrays (grid_data_mod.f90:33-67) for the device side. */
int imhd_create(const imhd_config* cfg, imhd_ctx** out);
int imhd_create_ensemble(const imhd_config* cfg, int nbox, imhd_ctx** out);
/* destroy_solver_vanleer / destroy_passive_solver (vanleer_mod.f90:161-165) */
int imhd_destroy(imhd_ctx* ctx);

/* Upload / download the reference-layout host arrays of box `box` (qpass may be NULL
 * when npass == 0).  get_state returns the state as the last call left it: call
 * imhd_fill_ghosts() first for valid ghost cells. */
int imhd_set_state(imhd_ctx* ctx, int box, const double* q_3d, const double* bface_3d, const double* qpass_3d);
int imhd_get_state(imhd_ctx* ctx, int box, double* q_3d, double* bface_3d, double* qpass_3d);
/* One field of box `box` into a host array: field 0..6 = q_3d(…,field+1), 7..9 =
 * bface_3d(…,field-6), both on the padded grid; 10 + p = qpass_3d(…,p+1) on its own bounds
 * (no counterpart in the reference, which reads its module arrays directly; lets a native
 * caller fetch a 1024^3 box field by field instead of as one 89 GB host state). */
int imhd_get_field(imhd_ctx* ctx, int box, int field, double* out);

/* Seeded random-field start of box `box`, generated on the device (no counterpart in the
 * reference, whose initial conditions are fixed test problems, init_domain_mod.f90; this
 * is the benchmark workloads' start, built where the state lives so that a 1024^3 box needs
 * neither host FFTs nor an 89 GB host upload):
 *   rho   = rho0 exp(sigma_ln g - sigma_ln^2 / 2), g standard normal per interior cell
 *           (Philox4x32-10, key = seed, counter = (cell, 0, 2, 0) with cell = i + nx (j +
 *           ny k) from 0; Box-Muller cosine branch), uniform rho0 when sigma_ln = 0;
 *   v     = sum over integer wave vectors n in one half space (n_z > 0, or n_z = 0 and
 *           n_y > 0, or n_z = n_y = 0 and n_x > 0) with kmin <= |n| <= kmax of
 *           2 Re(v_n e^{2 pi i (n_x i / nx + n_y j / ny + n_z k / nz)}), cell (i, j, k)
 *           from 0.  Mode m (its place in that enumeration: n_z, then n_y, then n_x
 *           ascending) draws v_n,d = |n|^{-(slope+2)/2} (g1 + i g2) from Philox counter
 *           (m, d, 1, 0); solenoidal != 0 projects each mode normal to its physical
 *           wave vector (n_x / nx, n_y / ny, n_z / nz; normal to n in a cube).  The amplitudes
 *           are scaled so that the mean of |v|^2 over the interior is exactly (mach cs)^2
 *           (Parseval; needs floor(kmax) < n_d / 2 on every axis);
 *   rho v = rho * v (FP64, rounded once in FP32 contexts); B = b0 in every cell and face;
 *   passive scalars (passive != 0) U[0,1) per cell from counter (cell, 0, 3 + p, 0), else 0.
 * Ghost cells are filled (imhd_fill_ghosts); the state replaces any earlier one as
 * imhd_set_state does.  A temporary device buffer of 48 (floor(kmax)+1) nx ny bytes is
 * allocated and freed during the call. */
typedef struct imhd_turb_ic {
    double rho0, sigma_ln;
    double b0[3];
    double kmin, kmax, slope, mach;
    unsigned long long seed;
    int solenoidal;
    int passive;
} imhd_turb_ic;
int imhd_init_turbulence(imhd_ctx* ctx, int box, const imhd_turb_ic* ic);

/* One integrator step of every box with the given dt:
 * vanleer_solve_3d(dt_in, drive_next) (vanleer_mod.f90:279-426) or tvd_solve_3d(dt_in)
 * (isotvd_mod.f90:226-364).  drive_next != 0 (VL only) applies the turbulence force set
 * by imhd_set_force as the reference does: get_rho_turb on the t^n state before the step
 * and add_random_force after it (vanleer_mod.f90:304-305, 419-421); its delv_max enters
 * imhd_max_speed.  Cells the step does not update keep their t^n values, as the
 * reference's restore of q_3d / bface_3d leaves them; the driver fills the ghost cells
 * afterwards. */
int imhd_step(imhd_ctx* ctx, double dt, int drive_next);

/* The driver's force field for box `box`: delv_re(…,3) of grid_data_mod.f90:50 on the
 * padded grid (same layout as q_3d, 3 components), as generate_rand_force
 * (turb_driver_mod.f90:197-391) leaves it, with Edot and dt_turb of constants_mod
 * (constant energy input Edot * dt_turb per kick, turb_driver_mod.f90:428-442).
 * Replaces the delv argument of add_random_force (turb_driver_mod.f90:395). */
int imhd_set_force(imhd_ctx* ctx, int box, const double* delv_3d, double edot, double dt_turb);
/* == generate_rand_force (turb_driver_mod.f90:197-391) on the device, for box `box`: the
 * next force field, generated from the driver's taus88 state seeds = (s1, s2, s3)
 * (grid_data_mod.f90:24; advanced in place by 6 k_max^3 draws) with modes
 * 0 <= kx,ky,kz < k_max, pure solenoidal projection, and the 1-D direct inverse DFTs
 * along x, y, z over the padded box.  The modes the reference leaves uninitialised
 * (mixed-sign wave vectors, turb_driver_mod.f90:199-200) are zero here.  Takes the place
 * of imhd_set_force for the next drive step; k_max must satisfy
 * 1 <= k_max < min(nx,ny,nz)/2 (turb_driver_mod.f90:31-34). */
int imhd_generate_force(imhd_ctx* ctx, int box, int k_max, double edot, double dt_turb, int seeds[3]);
/* The current force field of box `box` (layout of imhd_set_force). */
int imhd_get_force(imhd_ctx* ctx, int box, double* delv_3d);

/* Ornstein-Uhlenbeck turbulence forcing of the native loop (imhd_run); no counterpart in
 * the reference, whose driver kicks impulsively (turb_driver_mod.f90).  Modes: integer
 * wave vectors n with kmin <= |n| <= kmax (one of each +/- pair; k = 2 pi n_d / L_d),
 * amplitudes a_n (3 complex components) following
 *     a <- f a + sigma sqrt(1 - f^2) P(k) xi,   f = exp(-dt / t_corr),
 * xi complex standard normal from Philox4x32-10 keyed by (seed, box) and counted by
 * (OU step, mode, component); P(k) projects onto the plane normal to k when solenoidal.
 * Each step after the MHD update: acceleration a(x) = sum 2 Re(a_n e^{i k.x}) and
 * delta(rho v) = dt (rho a - <rho a>) on interior cells (no net momentum injection), before
 * the passive step, CFL reduction and ghost fill.  Amplitudes start from the stationary
 * distribution.  sigma <= 0 (or ou == NULL) switches the forcing off. */
typedef struct imhd_ou {
    double kmin, kmax;  /* mode-number shell, e.g. 1 and 2 */
    double t_corr;      /* correlation time */
    double sigma;       /* stationary rms of each real amplitude component (acceleration units) */
    unsigned long long seed;
    int solenoidal;     /* 1: project the forcing onto solenoidal modes */
    int _reserved;
} imhd_ou;
int imhd_set_ou_forcing(imhd_ctx* ctx, const imhd_ou* ou);
/* Number of OU modes; with non-NULL arrays also their integer wave vectors (3 per mode)
 * and the current amplitudes of box `box` (per mode: 3 real parts, then 3 imaginary). */
int imhd_get_ou_modes(imhd_ctx* ctx, int box, int* nmode, int* kvec, double* amp);

/* boundary_conditions() on one rank (boundary_conditions_mod.f90:20-234) */
int imhd_fill_ghosts(imhd_ctx* ctx);

/* get_max_speed_vanleer (vanleer_mod.f90:3266-3334): max over interior cells of
 * |v| + c_f, combined with the interface maximum smax of the last step; with IMHD_TVD
 * get_max_speed_tvd (isotvd_mod.f90:2607-2673): max of |v| + c_f of the interface-mean
 * states over i, j, k in [0, n] (reads ghost layer 0: the driver calls it before the
 * ghost fill, driver.f90:131-160).  cmax receives one value per box. */
int imhd_max_speed(imhd_ctx* ctx, double* cmax);

/* copy_velocity(predictor_stage) (passive_mod.f90:204-244) and passive_solve_3d(dt)
 * (passive_mod.f90:118-201) */
int imhd_copy_velocity(imhd_ctx* ctx, int stage);
int imhd_passive_step(imhd_ctx* ctx, double dt);

/* compute_KE + compute_divB (driver.f90:492-553) and max|div B|; out[box].  Sums are
 * reduced in a fixed order (per-workgroup partial rows, then one fixed-order pass per box),
 * so repeated calls on the same state return the same bits.  The first call allocates the
 * partial rows: 1024 x 8 doubles per box (64 KiB), counted in imhd_device_bytes. */
int imhd_diagnostics(imhd_ctx* ctx, imhd_diag* out);

/* imhd_diagnostics without any allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc), where imhd_diagnostics's first call would allocate its
 * partial rows.  Same kernels and fixed reduction order, so the same bits.  The partial
 * rows live in a fixed 4.2 MB device array per GPU, kept until the process ends, whose
 * address imhd_create looks up: contexts of at most 64 boxes (IMHD_EINVAL above that).
 * Every call on a GPU uses that GPU's one array: calls that overlap on one GPU (other
 * threads, other contexts' streams) race and must be serialised by the caller. */
int imhd_diagnostics_static(imhd_ctx* ctx, imhd_diag* out);

/* Number of cells the density floor reset since creation (final_check_unphysical,
 * vanleer_mod.f90:3243-3262, which prints instead). */
int imhd_floor_count(imhd_ctx* ctx, unsigned int* count);

/* Device memory held by the context's field and work arrays, in bytes (no counterpart in
 * the reference, whose arrays live in module memory).  Work arrays a configuration never
 * uses are not allocated: with both VL stages fused (PLM) the flux, face-EMF and
 * corner-EMF arrays are absent, and the drop-in passive path's velocity and flux arrays
 * appear on the first imhd_copy_velocity / imhd_passive_step call. */
int imhd_device_bytes(imhd_ctx* ctx, unsigned long long* bytes);

/* Native, device-resident time loop of the driver (driver.f90:35-186) for nsteps steps:
 * copy_velocity, step, passive, CFL dt = COUR dx / cmax with dt = MIN(t_end - t, dt),
 * ghost fill.  dt stays on the device; the first call starts from dt = 1e-15
 * (driver.f90:249) unless imhd_set_time set another.  With use_graph != 0 one step is
 * captured once in a hipGraph and replayed. */
int imhd_run(imhd_ctx* ctx, int nsteps, double t_end, int use_graph);

/* The reference driver's impulsive turbulence forcing inside imhd_run (driver.f90:57-70,
 * 102-111, 336-344; replaces the driver's init_turb_driver + generate_rand_force set-up
 * and its drive_next schedule).  Draws the first field now (as imhd_generate_force, from
 * the taus88 state seeds = s1, s2, s3) and sets the next drive time to t + dt_turb.  Each
 * imhd_run step then kicks when that time falls in (t, t + dt]: the step runs with dt
 * halved and the force applied (get_rho_turb + add_random_force), the next drive time
 * becomes MAX(t_turb + dt_turb, t + 0.01 dt), the next field is drawn, and the following
 * dt is taken with COUR halved.  The host reads t and dt once per step.  Single-box VL
 * contexts, exclusive with OU forcing; k_max <= 0 switches it off.  Errors: IMHD_EINVAL
 * (k_max >= min(n)/2, TVD, ensemble, OU on, dt_turb <= 0), IMHD_ESTATE (no state). */
int imhd_set_impulsive_forcing(imhd_ctx* ctx, int k_max, double edot, double dt_turb, const int seeds[3]);
/* The current taus88 state (write_restart's s1, s2, s3) and next drive time. */
int imhd_get_impulsive_forcing(imhd_ctx* ctx, int seeds[3], double* t_turb);
int imhd_set_time(imhd_ctx* ctx, const double* t, const double* dt); /* per box; NULL keeps */
int imhd_get_time(imhd_ctx* ctx, double* t, double* dt);    /* per box */

/* Per-kernel device time of the last imhd_run, from HIP events on the context's stream
 * (enable with imhd_set_profiling before the run).  Returns a ';'-separated
 * "name=ms_total/launches/work_items_total" list. */
int imhd_set_profiling(imhd_ctx* ctx, int on);
const char* imhd_profile_report(imhd_ctx* ctx);

/* Launch-shape autotuner (no counterpart in the reference): times `steps` native-loop steps
 * (imhd_run without a graph) from the current state for each candidate z-segment length of
 * the column update kernels and y-strip height of the work enumeration, keeps the fastest,
 * and restores state, time and forcing state exactly -- also when a trial fails, in which
 * case the launch shape in use before the call is kept and the first error returned.  `report` (may be NULL) receives the
 * per-candidate timings and the choice. */
int imhd_autotune(imhd_ctx* ctx, int steps, char* report, int report_len);

/* Synchronise the context's stream. */
int imhd_sync(imhd_ctx* ctx);

/* In-situ diagnostics of the reference's analysis program (3D_MPI/Analysis/src/
 * analysis_main.f90), formed from the device state instead of snapshot files, in the
 * program's REAL*4 arithmetic:
 * imhd_spectrum == compute_pk (:543-633) of field 0 density, 1 velocity (rho v / rho),
 *   2 cell-centred B, 3 velocity parallel to B, 4 velocity perpendicular to B (:105-226):
 *   unnormalised 3-D DFT of each component (hipFFT), |F|^2 of the octant
 *   0 <= kx,ky,kz < n/2 summed over components, doubled (zero mode single), binned into
 *   n shells of width dk = sqrt(3) 2 pi (n/2) / n, times dx^3.  k[b] = b dk, pk[b]: n
 *   values each.  Cubic boxes with even n; bins accumulate in FP64 with device atomics, so
 *   the order of the additions varies from call to call: a bin of m modes may differ
 *   between two calls on the same state by at most m 2^-53 of its value (in practice a few
 *   ulp).  The first call on a cube edge n creates the hipFFT plan and keeps its buffers
 *   until imhd_destroy: 4 n^3 (packed field) + n^3 / 2 (octant power) + 8 n^2 (n/2+1)
 *   (transform) + 64 n + 24 (bins) bytes plus hipFFT's work area, and 16 n^2 (n/2+1) more
 *   once imhd_mode_spectra ran; all of it counts in imhd_device_bytes.  A failed set-up
 *   releases what it made, so the next call plans again.
 * imhd_density_pdf == compute_density_pdf (:637-698): nbins bins on [0, rho_max) (the
 *   reference: n bins on [0, 4)), normalised to unit integral over the counted cells;
 *   rho[b] is the bin centre.  Cells at or above rho_max are not counted (the reference
 *   writes past its array there). */
int imhd_spectrum(imhd_ctx* ctx, int box, int field, double* k, double* pk);
int imhd_density_pdf(imhd_ctx* ctx, int box, int nbins, double rho_max, double* rho, double* pdf);
/* == mode_decomposition (:305-540, Cho & Lazarian 2003): per octant mode k != 0 the
 * velocity transforms are projected onto the Alfven, fast and slow displacement vectors
 * about the mean field B0; P[m*n + b], m = 0..7: v_Alfven, v_fast, v_slow, b_Alfven,
 * b_fast, b_slow, rho_fast, rho_slow, shells as imhd_spectrum.  As the analysis program
 * calls it (:249), the transforms projected are those of v_perp.  Geometry in FP64; slow
 * amplitudes with a zero slow speed (0/0 in the reference) are not counted. */
int imhd_mode_spectra(imhd_ctx* ctx, int box, double* k, double* P);

/* Restart file of one box == write_restart / read_restart (io_mod.f90:440-517): Fortran
 * sequential-unformatted records t_sim, q_3d, bface_3d, qpass_3d (only when npass > 0),
 * dump_count, s1, s2, s3, framed as gfortran frames them (the reference's compiler,
 * 3D_MPI/build/Makefile:3): 4-byte length markers around each record, records above
 * 2147483639 bytes split into subrecords with negated markers.  The state moves between
 * the device and the file in bounded chunks (no whole-array host copy).  The reader also
 * accepts single records whose markers hold the length modulo 2^32 (flang), checks every
 * marker against this context's sizes before it changes any state, and rejects trailing
 * data; it sets the state as imhd_set_state does and returns t_sim, dump_count and the
 * seeds without applying them (the driver owns its clock and RNG, driver.f90:311). */
int imhd_write_restart(imhd_ctx* ctx, int box, const char* path, double t_sim, int dump_count,
                       const int seeds[3]);
int imhd_read_restart(imhd_ctx* ctx, int box, const char* path, double* t_sim, int* dump_count,
                      int seeds[3]);
/* Snapshot cube of one box, packed on the device: interior cells, i fastest, per cell the 7
 * cell fields then passive scalars 1..npass_out, REAL*8 (double_prec != 0) or REAL*4
 * rounded to nearest (SNGL), no record markers.  npass_out = min(npass, 1) is the xyz cube
 * of writetofile_unformatted (io_mod.f90:249-277); npass_out = npass is the layout of the
 * parallel dump on one rank (load_parallel_file_buffer + writetofile_unformatted_parallel,
 * io_mod.f90:302-436). */
int imhd_write_cube(imhd_ctx* ctx, int box, const char* path, int double_prec, int npass_out);

const char* imhd_last_error(const imhd_ctx* ctx);
const char* imhd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* IMHD_H */
