! Drop-in replacement for the reference's turbulence driver module (turb_driver_mod.f90)
! as the driver uses it (driver.f90:104, 224, 336-339): init_turb_driver,
! generate_rand_force and destroy_turb_driver under the reference's names.  The force
! field is generated on the GPU by imhd_generate_force from the driver's taus88 state
! s1, s2, s3 (grid_data_mod.f90:24), with the Fourier modes the reference leaves
! uninitialised set to zero, and stays on the device for the next
! vanleer_solve_3d(dt, drive_next=.TRUE.) of the solver shim (imhd_shim.f90), which then
! applies it as get_rho_turb + add_random_force do.  Compile in place of the reference's
! turb_driver_mod.f90, after imhd_shim.f90.
MODULE turb_driver_mod
  USE ISO_C_BINDING
  USE imhd_api
  USE imhd_shim_core
  USE constants_mod
  USE grid_data_mod
  IMPLICIT NONE

CONTAINS

  ! turb_driver_mod.f90:22-83: the k_max check and the taus88 seeds
  SUBROUTINE init_turb_driver(s1_in, s2_in, s3_in)
    INTEGER, INTENT(IN) :: s1_in, s2_in, s3_in
    IF (k_max .GE. MIN(nx*nranks_x, ny*nranks_y, nz*nranks_z)/2) THEN
      PRINT*,'ERROR!! Invalid k_max value.'
      STOP
    END IF
    s1 = s1_in
    s2 = s2_in
    s3 = s3_in
  END SUBROUTINE init_turb_driver

  SUBROUTINE destroy_turb_driver()
  END SUBROUTINE destroy_turb_driver

  ! turb_driver_mod.f90:197-391 on the device; advances s1, s2, s3 like rand_taus does
  SUBROUTINE generate_rand_force()
    INTEGER(C_INT) :: seeds(3)
    seeds = (/ INT(s1, C_INT), INT(s2, C_INT), INT(s3, C_INT) /)
    CALL shim_check(imhd_generate_force(shim_ctx, 0_C_INT, INT(k_max, C_INT), Edot, dt_turb, seeds), &
                    'generate_rand_force')
    s1 = seeds(1)
    s2 = seeds(2)
    s3 = seeds(3)
    shim_force_on_device = .TRUE.
  END SUBROUTINE generate_rand_force

END MODULE turb_driver_mod
