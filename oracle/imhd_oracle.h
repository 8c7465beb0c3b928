/* CPU FP64 restatement of the reference's isothermal-MHD cell-update path
 * (tsultan16/Isothermal_MHD, 3D_MPI/src).  TEST INFRASTRUCTURE ONLY: the parity tests,
 * smoke() and bench.py's cpu_baseline leg call it as the checker; the product never
 * links it.
 *
 * Arrays use the reference's layout (grid_data_mod.f90:38-42): Fortran column-major,
 * SoA by field, ghost-padded with nb layers, i.e. element (i,j,k,f) of q_3d with
 * i in [1-nb, nx+nb] lives at ((f*NZ + k+nb-1)*NY + j+nb-1)*NX + i+nb-1.
 * qpass_3d has bounds [3-nb, n+nb-2] per axis (grid_data_mod.f90:42).
 */
#ifndef IMHD_ORACLE_H
#define IMHD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_cfg {
    int nx, ny, nz;   /* interior cells */
    int nb;           /* ghost layers (reference: 6) */
    int npass;        /* passive scalars */
    int mhd;          /* integrator, reference numbering: 1 TVD (K98 CTU+CT), 2 VL2+CT */
    int riemann;      /* reference numbering inside the integrator:
                         VL: 1 ROE, 2 HLLD, 3 HLLE;  TVD: 1 HLLE, 2 K98 Roe-TVD */
    int recon;        /* VL corrector reconstruction: 0 PLM primitive (MC-beta),
                         1 PPM primitive, 2 PPM characteristic (reference default) */
    int bc;           /* 1 open (zero-gradient), 2 periodic */
    double cs;        /* isothermal sound speed */
    double mc_beta;   /* MC limiter beta (reference 1.01) */
    int protection;   /* VL + Roe: flux_protection (vanleer_mod.f90:81, off in the reference) */
    int _pad;
    double prot_dens; /* density below which a cell is protected (0: MIN_DENS = 1e-20) */
} orc_cfg;

typedef struct orc_ctx orc_ctx;

typedef struct orc_diag {
    double ke, vrms_sq, momx, momy, momz;   /* driver.f90:518-553 (incl. dx^3, 0.5 on ke) */
    double divb_sum;                         /* driver.f90:492-515, signed raw sum */
    double divb_max;                         /* max_cell |sum of face differences| (raw) */
    double mass;                             /* sum rho * dx^3 */
} orc_diag;

orc_ctx *orc_create(const orc_cfg *cfg);
void orc_destroy(orc_ctx *c);
void orc_set_state(orc_ctx *c, const double *q, const double *bface, const double *qpass);
void orc_get_state(const orc_ctx *c, double *q, double *bface, double *qpass, double *emf_corner);
void orc_bc(orc_ctx *c);
void orc_step(orc_ctx *c, double dt);
/* Forcing application of the reference's turbulence driver: delv on the padded grid
 * ((1-nb:n+nb)^3, 3; the driver's delv_re), Edot and dt_turb (constants_mod.f90).
 * orc_step_drive = vanleer_solve_3d(dt, drive_next=.TRUE.): get_rho_turb before the step,
 * add_random_force after it; returns -1 on the reference's bad-normalisation STOP. */
void orc_set_force(orc_ctx *c, const double *delv, double edot, double dt_turb);
/* Ornstein-Uhlenbeck forcing of the library's native loop (imhd_set_ou_forcing):
 * setup returns the number of modes; orc_ou_step updates the amplitudes with dt and kicks
 * the current state; orc_ou_modes_step only updates the amplitudes. */
int orc_ou_setup(orc_ctx *c, double kmin, double kmax, double t_corr, double sigma,
                 unsigned long long seed, int solenoidal, int box);
void orc_ou_step(orc_ctx *c, double dt);
void orc_ou_modes_step(orc_ctx *c, double dt);
void orc_ou_get(const orc_ctx *c, int *kvec, double *amp);
void orc_ou_field(const orc_ctx *c, double *out);
void orc_philox4x32(const unsigned int *ctr, const unsigned int *key, unsigned int *out);
/* The reference's random force (generate_rand_force, turb_driver_mod.f90:197-391) with its
 * unassigned modes zeroed: taus88 state seeds[3] in/out, K = k_max; delv: 3 padded fields;
 * modes (may be NULL): the Fourier coefficients, re then im, [3][2K+1]^3. */
void orc_taus88(int *seeds, int n, double *out);
void orc_idft_1d(const double *in_re, const double *in_im, double *out_re, double *out_im, int ilow, int ihi,
                 int nnx, int K);
void orc_rand_force(int nx, int ny, int nz, int nb, int K, int *seeds, double *delv, double *modes);
int orc_step_drive(orc_ctx *c, double dt);
double orc_max_speed(orc_ctx *c);
double orc_smax(const orc_ctx *c);
void orc_copy_velocity(orc_ctx *c, int stage);
void orc_passive_step(orc_ctx *c, double dt);
void orc_diagnostics(const orc_ctx *c, orc_diag *out);
int orc_floor_count(const orc_ctx *c);
/* one hydro + CT update of the current state with given Fs, Fs_prot, emf_corner and
 * emf_corner_prot (protection contexts only; -1 otherwise) */
int orc_update_test(orc_ctx *c, double dtdx, int off, const double *Fs, const double *Fsp, const double *ec,
                    const double *ecp);

/* Per-interface pins: ni x nj planes (Fortran order, component slowest). */
void orc_riemann_plane(int kind, int ni, int nj, double cs, const double *qL,
                       const double *qR, double *F, double *smax);
void orc_recon_plane(int kind, int ni, int nj, int nb, double cs, double mc_beta,
                     const double *w, const double *wb, double *qL, double *qR);
/* TVD-module interface fluxes along i of cell states q (slot 4 = the cell's own face
 * field): kind 2 K98 Roe-TVD with dtdx = dt/dx (x 0.5 when predictor), kind 1 HLLE. */
void orc_tvd_flux_plane(int kind, int ni, int nj, double cs, double dt, double dx, int predictor,
                        const double *q, double *F);

#ifdef __cplusplus
}
#endif
#endif
