! Drop-in replacements for the reference's solver modules (tsultan16/Isothermal_MHD,
! 3D_MPI/src), under the reference's own module and procedure names, so a driver written
! against them (driver.f90) compiles and links unchanged against the MI355X library:
!
!   VanLeer_Solver_mod       vanleer_mod.f90:   init_solver_vanleer, destroy_solver_vanleer,
!                                               vanleer_solve_3d, get_max_speed_vanleer
!   Iso_TVD_Solver_mod       isotvd_mod.f90:    init_solver_tvd, destroy_solver_tvd,
!                                               tvd_solve_3d, get_max_speed_tvd
!   passive_solver_mod       passive_mod.f90:   init_passive_solver, destroy_passive_solver,
!                                               copy_velocity, passive_solve_3d
!   boundary_conditions_mod  boundary_conditions_mod.f90: boundary_conditions
!
! Like the reference modules, these USE the application's constants_mod (grid size, nb,
! npass, solver and boundary types, sound_speed, COUR_def, drive_turbulence, Edot,
! dt_turb) and grid_data_mod (q_3d, bface_3d, qpass_3d, delv_re), so this file is
! compiled together with them.  The state lives on
! the GPU; the host arrays are uploaded when the host has changed them (after init, or
! after imhd_shim_host_modified) and downloaded after every call that changes the state,
! because the driver reads them between calls (compute_KE, file dumps).  A failing call
! prints the library's message and STOPs, as the reference does on errors.
MODULE imhd_shim_core
  USE ISO_C_BINDING
  USE imhd_api
  USE constants_mod
  USE grid_data_mod
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: shim_create, shim_destroy, shim_upload, shim_download, shim_check, shim_ctx, &
            imhd_shim_host_modified, imhd_shim_recon, imhd_shim_precision, shim_force_on_device, &
            imhd_shim_flux_protection

  TYPE(C_PTR), SAVE :: shim_ctx = C_NULL_PTR
  LOGICAL, SAVE :: host_dirty = .TRUE.
  ! corrector reconstruction of the VL integrator: the reference as compiled uses
  ! characteristic PPM (vanleer_mod.f90:83-84); IMHD_PLM_PRIM selects the MC-beta PLM
  INTEGER(C_INT), SAVE :: imhd_shim_recon = IMHD_PPM_CHAR
  ! 64: FP64 throughout (bitwise with the reference); 32: FP32 bulk with FP64 face B,
  ! corner EMF, CFL and diagnostics, Kahan-compensated rho and rho v
  INTEGER(C_INT), SAVE :: imhd_shim_precision = 64
  ! 1: the reference's flux_protection (vanleer_mod.f90:81, a PARAMETER that is .FALSE. there)
  INTEGER(C_INT), SAVE :: imhd_shim_flux_protection = 0
  ! set by the turb_driver_mod shim (imhd_shim_turb.f90) once the force field is generated
  ! on the device; the host delv_re is then not uploaded
  LOGICAL, SAVE :: shim_force_on_device = .FALSE.
  REAL(C_DOUBLE), ALLOCATABLE, SAVE :: nopass(:)

CONTAINS

  SUBROUTINE shim_check(rc, where)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: where
    IF (rc .NE. IMHD_OK) THEN
      PRINT*, 'imhd: ', where, ' failed (code ', rc, '): ', TRIM(imhd_last_error(shim_ctx))
      STOP
    END IF
  END SUBROUTINE shim_check

  SUBROUTINE shim_create(integrator, riemann)
    INTEGER(C_INT), INTENT(IN) :: integrator, riemann
    TYPE(imhd_config) :: cfg
    IF (C_ASSOCIATED(shim_ctx)) RETURN
    cfg%nx = nx
    cfg%ny = ny
    cfg%nz = nz
    cfg%nb = nb
    cfg%npass = npass
    cfg%integrator = integrator
    cfg%riemann = riemann
    cfg%recon = imhd_shim_recon
    cfg%bc = boundary_type
    cfg%precision = imhd_shim_precision
    cfg%compensated = MERGE(1, 0, imhd_shim_precision .EQ. 32)
    cfg%flux_protection = imhd_shim_flux_protection
    cfg%cs = sound_speed
    cfg%mc_beta = 1.01d0
    cfg%cour = COUR_def
    CALL shim_check(imhd_create(cfg, shim_ctx), 'imhd_create')
    IF (.NOT. ALLOCATED(nopass)) ALLOCATE(nopass(1))
    host_dirty = .TRUE.
  END SUBROUTINE shim_create

  SUBROUTINE shim_destroy()
    INTEGER(C_INT) :: rc
    IF (.NOT. C_ASSOCIATED(shim_ctx)) RETURN
    rc = imhd_destroy(shim_ctx)
    shim_ctx = C_NULL_PTR
  END SUBROUTINE shim_destroy

  ! Call after changing q_3d, bface_3d or qpass_3d on the host between solver calls.
  SUBROUTINE imhd_shim_host_modified()
    host_dirty = .TRUE.
  END SUBROUTINE imhd_shim_host_modified

  SUBROUTINE shim_upload()
    IF (.NOT. C_ASSOCIATED(shim_ctx)) THEN
      PRINT*, 'imhd: solver used before init_solver_vanleer / init_solver_tvd'
      STOP
    END IF
    IF (.NOT. host_dirty) RETURN
    IF (npass .GT. 0) THEN
      CALL shim_check(imhd_set_state(shim_ctx, 0_C_INT, q_3d, bface_3d, qpass_3d), 'imhd_set_state')
    ELSE
      CALL shim_check(imhd_set_state(shim_ctx, 0_C_INT, q_3d, bface_3d, nopass), 'imhd_set_state')
    END IF
    host_dirty = .FALSE.
  END SUBROUTINE shim_upload

  SUBROUTINE shim_download()
    IF (npass .GT. 0) THEN
      CALL shim_check(imhd_get_state(shim_ctx, 0_C_INT, q_3d, bface_3d, qpass_3d), 'imhd_get_state')
    ELSE
      CALL shim_check(imhd_get_state(shim_ctx, 0_C_INT, q_3d, bface_3d, nopass), 'imhd_get_state')
    END IF
  END SUBROUTINE shim_download

END MODULE imhd_shim_core


MODULE VanLeer_Solver_mod
  USE ISO_C_BINDING
  USE imhd_api
  USE imhd_shim_core
  USE constants_mod
  USE grid_data_mod, ONLY: delv_re
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: init_solver_vanleer, destroy_solver_vanleer, vanleer_solve_3d, get_max_speed_vanleer

CONTAINS

  ! vanleer_mod.f90:127-158; riemann_solver_type keeps the VL numbering (1 ROE, 2 HLLD, 3 HLLE)
  SUBROUTINE init_solver_vanleer()
    CALL shim_create(IMHD_VL, INT(riemann_solver_type, C_INT))
  END SUBROUTINE init_solver_vanleer

  SUBROUTINE destroy_solver_vanleer()
    CALL shim_destroy()
  END SUBROUTINE destroy_solver_vanleer

  ! vanleer_mod.f90:279-426.  With drive_turbulence and drive_next, the force field (the
  ! device-generated one when the turb_driver_mod shim is linked, else the one the
  ! driver's generate_rand_force left in delv_re) is applied on the GPU as the reference
  ! applies it (get_rho_turb before the step, add_random_force after it).
  SUBROUTINE vanleer_solve_3d(dt_in, drive_next)
    REAL*8, INTENT(INOUT) :: dt_in
    LOGICAL, INTENT(IN) :: drive_next
    INTEGER(C_INT) :: drive
    CALL shim_upload()
    drive = 0
    IF (drive_turbulence .AND. drive_next) THEN
      IF (.NOT. shim_force_on_device) &
        CALL shim_check(imhd_set_force(shim_ctx, 0_C_INT, delv_re, Edot, dt_turb), 'imhd_set_force')
      drive = 1
    END IF
    CALL shim_check(imhd_step(shim_ctx, dt_in, drive), 'vanleer_solve_3d')
    CALL shim_download()
  END SUBROUTINE vanleer_solve_3d

  ! vanleer_mod.f90:3266-3334
  SUBROUTINE get_max_speed_vanleer(max_speed)
    REAL*8, INTENT(INOUT) :: max_speed
    REAL(C_DOUBLE) :: c(1)
    CALL shim_upload()
    CALL shim_check(imhd_max_speed(shim_ctx, c), 'get_max_speed_vanleer')
    max_speed = c(1)
  END SUBROUTINE get_max_speed_vanleer

END MODULE VanLeer_Solver_mod


MODULE Iso_TVD_Solver_mod
  USE ISO_C_BINDING
  USE imhd_api
  USE imhd_shim_core
  USE constants_mod
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: init_solver_tvd, destroy_solver_tvd, tvd_solve_3d, get_max_speed_tvd

CONTAINS

  ! isotvd_mod.f90:122-133; the TVD module numbers its fluxes 1 HLLE, 2 K98 Roe-TVD
  ! (isotvd_mod.f90:400-428)
  SUBROUTINE init_solver_tvd()
    INTEGER(C_INT) :: rs
    rs = MERGE(IMHD_ROE, IMHD_HLLE, riemann_solver_type .EQ. 2)
    CALL shim_create(IMHD_TVD, rs)
  END SUBROUTINE init_solver_tvd

  SUBROUTINE destroy_solver_tvd()
    CALL shim_destroy()
  END SUBROUTINE destroy_solver_tvd

  ! isotvd_mod.f90:226-364
  SUBROUTINE tvd_solve_3d(dt_in)
    REAL*8, INTENT(INOUT) :: dt_in
    CALL shim_upload()
    CALL shim_check(imhd_step(shim_ctx, dt_in, 0_C_INT), 'tvd_solve_3d')
    CALL shim_download()
  END SUBROUTINE tvd_solve_3d

  ! isotvd_mod.f90:2607-2673
  SUBROUTINE get_max_speed_tvd(max_speed)
    REAL*8, INTENT(INOUT) :: max_speed
    REAL(C_DOUBLE) :: c(1)
    CALL shim_upload()
    CALL shim_check(imhd_max_speed(shim_ctx, c), 'get_max_speed_tvd')
    max_speed = c(1)
  END SUBROUTINE get_max_speed_tvd

END MODULE Iso_TVD_Solver_mod


MODULE passive_solver_mod
  USE ISO_C_BINDING
  USE imhd_api
  USE imhd_shim_core
  USE constants_mod
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: init_passive_solver, destroy_passive_solver, copy_velocity, passive_solve_3d

CONTAINS

  ! passive_mod.f90:44-57: the scalars are part of the solver context (npass)
  SUBROUTINE init_passive_solver()
  END SUBROUTINE init_passive_solver

  SUBROUTINE destroy_passive_solver()
  END SUBROUTINE destroy_passive_solver

  ! passive_mod.f90:204-244
  SUBROUTINE copy_velocity(predictor_stage)
    LOGICAL, INTENT(IN) :: predictor_stage
    CALL shim_upload()
    CALL shim_check(imhd_copy_velocity(shim_ctx, MERGE(1_C_INT, 0_C_INT, predictor_stage)), 'copy_velocity')
  END SUBROUTINE copy_velocity

  ! passive_mod.f90:118-201
  SUBROUTINE passive_solve_3d(dt_in)
    REAL*8, INTENT(INOUT) :: dt_in
    CALL shim_upload()
    CALL shim_check(imhd_passive_step(shim_ctx, dt_in), 'passive_solve_3d')
    CALL shim_download()
  END SUBROUTINE passive_solve_3d

END MODULE passive_solver_mod


MODULE boundary_conditions_mod
  USE ISO_C_BINDING
  USE imhd_api
  USE imhd_shim_core
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: boundary_conditions

CONTAINS

  ! boundary_conditions_mod.f90:20-234 on one rank (the whole halo)
  SUBROUTINE boundary_conditions()
    CALL shim_upload()
    CALL shim_check(imhd_fill_ghosts(shim_ctx), 'boundary_conditions')
    CALL shim_download()
  END SUBROUTINE boundary_conditions

END MODULE boundary_conditions_mod
