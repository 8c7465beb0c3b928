! bind(C) harness around the reference Fortran solver modules (tsultan16/Isothermal_MHD,
! 3D_MPI/src).  TEST INFRASTRUCTURE ONLY: it is compiled with the reference modules by
! oracle/ref/build_ref.py into oracle/_ref/libimhd_ref_<cfg>.so and loaded by the CPU
! parity tests and the golden-vector generator.  It never ships in the product.
!
! Every entry point only moves data into/out of the reference's own module arrays and
! calls the reference's own procedures; no solver arithmetic lives here.
MODULE ref_harness
  USE ISO_C_BINDING
  USE constants_mod
  USE grid_data_mod
  IMPLICIT NONE
  LOGICAL, SAVE :: is_setup = .FALSE.
CONTAINS

  SUBROUTINE ref_dims(d) BIND(C, name='ref_dims')
    INTEGER(C_INT), INTENT(OUT) :: d(8)
    d = (/ nx, ny, nz, nb, npass, mhd_solver_type, riemann_solver_type, boundary_type /)
  END SUBROUTINE ref_dims

  ! driver.f90:237-266 (main_init) minus MPI, I/O and the built-in IC
  SUBROUTINE ref_setup() BIND(C, name='ref_setup')
    USE VanLeer_Solver_mod, ONLY: init_solver_vanleer
    USE Iso_TVD_Solver_mod, ONLY: init_solver_tvd
    USE passive_solver_mod, ONLY: init_passive_solver
    USE turb_driver_mod, ONLY: init_turb_driver
    USE mpi_domain_mod, ONLY: mpi_comm3d_init
    INTEGER :: ierr
    IF (is_setup) RETURN
    ! the forcing path reduces over comm3d (get_rho_turb, turb_driver_mod.f90:184-188):
    ! one-rank MPI and the Cartesian communicator of mpi_domain_mod, without its RMA set-up
    IF (drive_turbulence) THEN
      CALL MPI_INIT(ierr)
      CALL mpi_comm3d_init()
    END IF
    myrank = 0
    my_coord = 0
    CALL create_grid_arrays()
    emf_corner = 0.d0
    my_dt = 1.d-15
    my_dx = 1.d0 / DBLE(MAX(nx, ny, nz))
    IF (mhd_solver_type .EQ. 1) CALL init_solver_tvd()
    IF (mhd_solver_type .EQ. 2) CALL init_solver_vanleer()
    IF (npass .GT. 0) CALL init_passive_solver()
    IF (drive_turbulence) CALL init_turb_driver(s1_def, s2_def, s3_def)
    is_setup = .TRUE.
  END SUBROUTINE ref_setup

  SUBROUTINE ref_set_sound_speed(cs) BIND(C, name='ref_set_sound_speed')
    REAL(C_DOUBLE), VALUE :: cs
    sound_speed = cs
  END SUBROUTINE ref_set_sound_speed

  SUBROUTINE ref_set_state(q, bf, qp) BIND(C, name='ref_set_state')
    REAL(C_DOUBLE), INTENT(IN) :: q(1-nb:nx+nb, 1-nb:ny+nb, 1-nb:nz+nb, 7)
    REAL(C_DOUBLE), INTENT(IN) :: bf(1-nb:nx+nb, 1-nb:ny+nb, 1-nb:nz+nb, 3)
    REAL(C_DOUBLE), INTENT(IN) :: qp(3-nb:nx+nb-2, 3-nb:ny+nb-2, 3-nb:nz+nb-2, MAX(npass,1))
    q_3d = q
    bface_3d = bf
    IF (npass .GT. 0) qpass_3d = qp(:,:,:,1:npass)
  END SUBROUTINE ref_set_state

  SUBROUTINE ref_get_state(q, bf, qp, ec) BIND(C, name='ref_get_state')
    REAL(C_DOUBLE), INTENT(OUT) :: q(1-nb:nx+nb, 1-nb:ny+nb, 1-nb:nz+nb, 7)
    REAL(C_DOUBLE), INTENT(OUT) :: bf(1-nb:nx+nb, 1-nb:ny+nb, 1-nb:nz+nb, 3)
    REAL(C_DOUBLE), INTENT(OUT) :: qp(3-nb:nx+nb-2, 3-nb:ny+nb-2, 3-nb:nz+nb-2, MAX(npass,1))
    REAL(C_DOUBLE), INTENT(OUT) :: ec(1-nb:nx+nb, 1-nb:ny+nb, 1-nb:nz+nb, 3)
    q = q_3d
    bf = bface_3d
    qp = 0.d0
    IF (npass .GT. 0) qp(:,:,:,1:npass) = qpass_3d
    ec = emf_corner
  END SUBROUTINE ref_get_state

  SUBROUTINE ref_bc() BIND(C, name='ref_bc')
    USE boundary_conditions_mod, ONLY: boundary_conditions
    CALL boundary_conditions()
  END SUBROUTINE ref_bc

  ! one call of the reference integrator (driver.f90:378-389)
  SUBROUTINE ref_step(dt) BIND(C, name='ref_step')
    USE VanLeer_Solver_mod, ONLY: vanleer_solve_3d
    USE Iso_TVD_Solver_mod, ONLY: tvd_solve_3d
    REAL(C_DOUBLE), VALUE :: dt
    REAL*8 :: dtl
    dtl = dt
    IF (mhd_solver_type .EQ. 1) CALL tvd_solve_3d(dtl)
    IF (mhd_solver_type .EQ. 2) CALL vanleer_solve_3d(dtl, .FALSE.)
  END SUBROUTINE ref_step

  ! driver.f90:432-448 (compute_new_time_step) returns cmax
  SUBROUTINE ref_max_speed(cmax) BIND(C, name='ref_max_speed')
    USE VanLeer_Solver_mod, ONLY: get_max_speed_vanleer
    USE Iso_TVD_Solver_mod, ONLY: get_max_speed_tvd
    REAL(C_DOUBLE), INTENT(OUT) :: cmax
    REAL*8 :: c
    c = 0.d0
    IF (mhd_solver_type .EQ. 1) CALL get_max_speed_tvd(c)
    IF (mhd_solver_type .EQ. 2) CALL get_max_speed_vanleer(c)
    cmax = c
  END SUBROUTINE ref_max_speed

  SUBROUTINE ref_vl_smax(s) BIND(C, name='ref_vl_smax')
    USE VanLeer_Solver_mod, ONLY: smax
    REAL(C_DOUBLE), INTENT(OUT) :: s
    s = smax
  END SUBROUTINE ref_vl_smax

  SUBROUTINE ref_copy_velocity(flag) BIND(C, name='ref_copy_velocity')
    USE passive_solver_mod, ONLY: copy_velocity
    INTEGER(C_INT), VALUE :: flag
    CALL copy_velocity(flag .NE. 0)
  END SUBROUTINE ref_copy_velocity

  SUBROUTINE ref_passive_step(dt) BIND(C, name='ref_passive_step')
    USE passive_solver_mod, ONLY: passive_solve_3d
    REAL(C_DOUBLE), VALUE :: dt
    REAL*8 :: dtl
    dtl = dt
    CALL passive_solve_3d(dtl)
  END SUBROUTINE ref_passive_step

  ! Per-interface Riemann fluxes of the VL module on an ni x nj plane of interface states.
  ! kind: 1 ROE chain, 2 HLLD, 3 HLLE (vanleer_mod.f90:2264-3239).  Returns smax.
  SUBROUTINE ref_vl_riemann(kind, ni, nj, qL, qR, F, sm) BIND(C, name='ref_vl_riemann')
    USE VanLeer_Solver_mod, ONLY: qintL, qintR, Fs_2d, smax, compute_HLLD_fluxes_1d, &
         compute_HLLE_fluxes_1d, compute_roe_avg_state_1d, compute_roe_eigenvalues_1d, &
         compute_roe_eigenvectors_1d, compute_ROE_fluxes_1d
    INTEGER(C_INT), VALUE :: kind, ni, nj
    REAL(C_DOUBLE), INTENT(IN) :: qL(ni, nj, 7), qR(ni, nj, 7)
    REAL(C_DOUBLE), INTENT(OUT) :: F(ni, nj, 7), sm
    INTEGER :: ilow, ihi, jlow, jhi
    ilow = 1
    ihi = ni - 2*nb + 1
    jlow = 1
    jhi = nj - 2*nb
    qintL = 0.d0
    qintR = 0.d0
    Fs_2d = 0.d0
    qintL(1-nb:ni-nb, 1-nb:nj-nb, :) = qL
    qintR(1-nb:ni-nb, 1-nb:nj-nb, :) = qR
    smax = 0.d0
    IF (kind .EQ. 2) CALL compute_HLLD_fluxes_1d(Fs_2d, ilow, ihi, jlow, jhi)
    IF (kind .EQ. 3) CALL compute_HLLE_fluxes_1d(Fs_2d, ilow, ihi, jlow, jhi)
    IF (kind .EQ. 1) THEN
      CALL compute_roe_avg_state_1d(ilow, ihi, jlow, jhi)
      CALL compute_roe_eigenvalues_1d(ilow, ihi, jlow, jhi)
      CALL compute_roe_eigenvectors_1d(ilow, ihi, jlow, jhi)
      CALL compute_ROE_fluxes_1d(Fs_2d, ilow, ihi, jlow, jhi)
    END IF
    F = Fs_2d(1-nb:ni-nb, 1-nb:nj-nb, :)
    sm = smax
  END SUBROUTINE ref_vl_riemann

  ! Turbulence forcing (builds with drive_turbulence = .TRUE.): the driver's force field
  ! delv_re on the padded grid (grid_data_mod.f90:50; turb_driver_mod.f90:379 index shift),
  ! and vanleer_solve_3d(dt, drive_next=.TRUE.) (get_rho_turb + add_random_force).
  SUBROUTINE ref_set_force(delv) BIND(C, name='ref_set_force')
    REAL(C_DOUBLE), INTENT(IN) :: delv(1-nb:nx+nb, 1-nb:ny+nb, 1-nb:nz+nb, 3)
    delv_re(-nb-nx/2:nb-1+nx/2, -nb-ny/2:nb-1+ny/2, -nb-nz/2:nb-1+nz/2, :) = delv
  END SUBROUTINE ref_set_force

  SUBROUTINE ref_step_drive(dt) BIND(C, name='ref_step_drive')
    USE VanLeer_Solver_mod, ONLY: vanleer_solve_3d
    REAL(C_DOUBLE), VALUE :: dt
    REAL*8 :: dtl
    dtl = dt
    CALL vanleer_solve_3d(dtl, .TRUE.)
  END SUBROUTINE ref_step_drive

  ! The reference's 1-D inverse DFT by direct summation (turb_driver_mod.f90:461-486) on
  ! in(-k_max : k_max-1); out holds ihi - ilow + 1 values.  Builds with drive_turbulence
  ! only (nnmax is set by init_turb_driver).
  SUBROUTINE ref_idft_1d(in_re, in_im, out_re, out_im, ilow, ihi, nnx) BIND(C, name='ref_idft_1d')
    USE turb_driver_mod, ONLY: direct_idft_1d, nnmax
    REAL(C_DOUBLE), INTENT(IN) :: in_re(2*k_max), in_im(2*k_max)
    REAL(C_DOUBLE), INTENT(OUT) :: out_re(*), out_im(*)
    INTEGER(C_INT), VALUE :: ilow, ihi, nnx
    REAL*8 :: ore(0:nnmax-1+2*nb), oim(0:nnmax-1+2*nb)
    CALL direct_idft_1d(in_re, in_im, ore, oim, ilow, ihi, nnx)
    out_re(1:ihi-ilow+1) = ore(0:ihi-ilow)
    out_im(1:ihi-ilow+1) = oim(0:ihi-ilow)
  END SUBROUTINE ref_idft_1d

  ! Restart and snapshot files through the reference's own io_mod (io_mod.f90:249-277,
  ! 440-517): write_restart / read_restart of the module arrays, and
  ! writetofile_unformatted, whose only enabled output in these builds is the xyz cube.
  ! The files land under Output/Restart and Output/Snapshots of the working directory.
  SUBROUTINE ref_write_restart(t_sim, dcount, seeds) BIND(C, name='ref_write_restart')
    USE io_mod, ONLY: write_restart
    REAL(C_DOUBLE), VALUE :: t_sim
    INTEGER(C_INT), VALUE :: dcount
    INTEGER(C_INT), INTENT(IN) :: seeds(3)
    dump_count = dcount
    s1 = seeds(1)
    s2 = seeds(2)
    s3 = seeds(3)
    CALL write_restart(t_sim)
  END SUBROUTINE ref_write_restart

  SUBROUTINE ref_read_restart(t_sim, dcount, seeds) BIND(C, name='ref_read_restart')
    USE io_mod, ONLY: read_restart
    REAL(C_DOUBLE), INTENT(OUT) :: t_sim
    INTEGER(C_INT), INTENT(OUT) :: dcount
    INTEGER(C_INT), INTENT(OUT) :: seeds(3)
    REAL*8 :: t
    CALL read_restart(t)
    t_sim = t
    dcount = dump_count
    seeds = (/ s1, s2, s3 /)
  END SUBROUTINE ref_read_restart

  SUBROUTINE ref_write_snapshot(dnum) BIND(C, name='ref_write_snapshot')
    USE io_mod, ONLY: writetofile_unformatted
    INTEGER(C_INT), VALUE :: dnum
    CALL writetofile_unformatted(dnum)
  END SUBROUTINE ref_write_snapshot

  ! Parallel-I/O buffer fill (io_mod.f90:368-436), read back element by element.  Needs a
  ! build with parallel_io = .TRUE. (the buffer is allocated by create_grid_arrays).
  SUBROUTINE ref_parallel_buffer(n, out) BIND(C, name='ref_parallel_buffer')
    USE io_mod, ONLY: load_parallel_file_buffer
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(OUT) :: out(n)
    CALL load_parallel_file_buffer(0.d0)
    IF (double_prec) THEN
      out = file_buffer_dble(1:n)
    ELSE
      out = DBLE(file_buffer_sngl(1:n))
    END IF
  END SUBROUTINE ref_parallel_buffer

  ! taus88 stream of the driver (turb_driver_mod.f90:95-126)
  SUBROUTINE ref_rand_taus(seeds, n, out) BIND(C, name='ref_rand_taus')
    USE turb_driver_mod, ONLY: set_taus_seed, rand_taus
    INTEGER(C_INT), INTENT(IN) :: seeds(3)
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(OUT) :: out(n)
    INTEGER :: i
    CALL set_taus_seed(seeds(1), seeds(2), seeds(3))
    DO i = 1, n
      out(i) = rand_taus()
    END DO
  END SUBROUTINE ref_rand_taus

  ! Interface fluxes of the TVD module on an ni x nj plane of cell states q(7) (slot 5 =
  ! the cell's own face field).  kind 2: compute_roe_avg_state_1d .. compute_TVD_fluxes_1d
  ! (isotvd_mod.f90:1856-2468); kind 1: compute_HLLE_fluxes_1d (isotvd_mod.f90:2473-2578).
  ! The module's dt is set from dt_in; dtdx as in compute_xfluxes_3d (isotvd_mod.f90:376-378).
  SUBROUTINE ref_tvd_flux(kind, ni, nj, dt_in, predictor, qin, F) BIND(C, name='ref_tvd_flux')
    USE Iso_TVD_Solver_mod, ONLY: tq => q, tFs_2d => Fs_2d, tdt => dt, tdx => dx, &
         t_roe_avg => compute_roe_avg_state_1d, compute_eigenvalues_1d, compute_eigenvectors_1d, &
         compute_TVD_fluxes_1d, t_hlle => compute_HLLE_fluxes_1d
    INTEGER(C_INT), VALUE :: kind, ni, nj, predictor
    REAL(C_DOUBLE), VALUE :: dt_in
    REAL(C_DOUBLE), INTENT(IN) :: qin(ni, nj, 7)
    REAL(C_DOUBLE), INTENT(OUT) :: F(ni, nj, 7)
    INTEGER :: ilow, ihi, jlow, jhi
    REAL*8 :: dtdx
    ilow = 1
    ihi = ni - 2*nb
    jlow = 1
    jhi = nj - 2*nb
    tdt = dt_in
    dtdx = tdt / tdx
    IF (predictor .NE. 0) dtdx = 0.5d0 * dtdx
    tq = 0.d0
    tFs_2d = 0.d0
    tq(1-nb:ni-nb, 1-nb:nj-nb, :) = qin
    IF (kind .EQ. 1) CALL t_hlle(tFs_2d, ilow, ihi, jlow, jhi)
    IF (kind .EQ. 2) THEN
      CALL t_roe_avg(ilow, ihi, jlow, jhi)
      CALL compute_eigenvalues_1d(ilow, ihi, jlow, jhi)
      CALL compute_eigenvectors_1d(ilow, ihi, jlow, jhi)
      CALL compute_TVD_fluxes_1d(tFs_2d, ilow, ihi, jlow, jhi, dtdx)
    END IF
    F = tFs_2d(1-nb:ni-nb, 1-nb:nj-nb, :)
  END SUBROUTINE ref_tvd_flux

  ! Interface-state reconstruction of the VL module on an ni x nj plane of primitive
  ! cell values w(6) and face normal field wb.  kind 1: compute_interface_states_prim as
  ! compiled (vanleer_mod.f90:1144-1485); kind 2: compute_primitive_eigenvectors +
  ! compute_interface_states_char (vanleer_mod.f90:941-1818).
  SUBROUTINE ref_vl_recon(kind, ni, nj, wi, wbi, qL, qR) BIND(C, name='ref_vl_recon')
    USE VanLeer_Solver_mod, ONLY: qintL, qintR, vw => w, wb, compute_interface_states_prim, &
         compute_interface_states_char, compute_primitive_eigenvectors
    INTEGER(C_INT), VALUE :: kind, ni, nj
    REAL(C_DOUBLE), INTENT(IN) :: wi(ni, nj, 6), wbi(ni, nj)
    REAL(C_DOUBLE), INTENT(OUT) :: qL(ni, nj, 7), qR(ni, nj, 7)
    INTEGER :: ilow, ihi, jlow, jhi
    ilow = 1
    ihi = ni - 2*nb
    jlow = 1
    jhi = nj - 2*nb
    qintL = 0.d0
    qintR = 0.d0
    vw = 0.d0
    wb = 0.d0
    vw(1-nb:ni-nb, 1-nb:nj-nb, :) = wi
    wb(1-nb:ni-nb, 1-nb:nj-nb) = wbi
    IF (kind .EQ. 1) CALL compute_interface_states_prim(ilow, ihi, jlow, jhi)
    IF (kind .EQ. 2) THEN
      CALL compute_primitive_eigenvectors(ilow, ihi, jlow, jhi)
      CALL compute_interface_states_char(ilow, ihi, jlow, jhi)
    END IF
    qL = qintL(1-nb:ni-nb, 1-nb:nj-nb, :)
    qR = qintR(1-nb:ni-nb, 1-nb:nj-nb, :)
  END SUBROUTINE ref_vl_recon

END MODULE ref_harness
