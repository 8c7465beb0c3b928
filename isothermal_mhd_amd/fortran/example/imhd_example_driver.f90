! Example host driver: the loop body of the reference's driver (driver.f90:35-160 --
! copy_velocity, mhd_update, copy_velocity, passive_solve_3d, compute_new_time_step,
! boundary_conditions, and with drive_turbulence the impulsive forcing schedule of
! driver.f90:49-70, 102-111, 336-344) written against the solver and turbulence-driver
! modules' reference names, which resolve to imhd_shim.f90 / imhd_shim_turb.f90 and run
! on the GPU.
!
! usage: imhd_example_<cfg> INPUT OUTPUT NSTEPS
!   INPUT  raw FP64 stream: q_3d, bface_3d, qpass_3d (reference layout, column-major)
!   OUTPUT raw FP64 stream: q_3d, bface_3d, qpass_3d, t, dt, t_turb, s1, s2, s3 after
!          NSTEPS steps
PROGRAM imhd_example_driver
  USE constants_mod
  USE grid_data_mod
  USE VanLeer_Solver_mod
  USE Iso_TVD_Solver_mod
  USE passive_solver_mod
  USE boundary_conditions_mod
  USE turb_driver_mod
  IMPLICIT NONE
  CHARACTER(LEN=512) :: fin, fout, arg
  INTEGER :: nsteps, step
  REAL*8 :: t_sim, cmax, t_old, t_turb, COUR
  LOGICAL :: found, drive_next

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  CALL GET_COMMAND_ARGUMENT(3, arg)
  READ(arg, *) nsteps
  INQUIRE(FILE=TRIM(fin), EXIST=found)
  IF (.NOT. found) THEN
    PRINT*, 'imhd example driver: input file not found: ', TRIM(fin)
    STOP 1
  END IF

  CALL create_grid_arrays()
  OPEN(UNIT=11, FILE=TRIM(fin), FORM='UNFORMATTED', ACCESS='STREAM', STATUS='OLD')
  READ(11) q_3d
  READ(11) bface_3d
  READ(11) qpass_3d
  CLOSE(11)

  my_dt = 1.d-15                                   ! driver.f90:249
  my_dx = 1.d0 / DBLE(MAX(nx, ny, nz))
  t_sim = 0.d0
  IF (mhd_solver_type .EQ. 1) THEN
    CALL init_solver_tvd()
  ELSE
    CALL init_solver_vanleer()
  END IF
  IF (npass .GT. 0) CALL init_passive_solver()
  CALL boundary_conditions()
  IF (drive_turbulence) CALL init_turb_driver(s1_def, s2_def, s3_def)   ! driver.f90:303-344
  IF (drive_turbulence) CALL generate_rand_force()
  drive_next = .FALSE.
  t_turb = t_sim + dt_turb

  DO step = 1, nsteps
    t_old = t_sim
    COUR = COUR_def
    IF (npass .GT. 0) CALL copy_velocity(.FALSE.)
    IF (drive_turbulence .AND. (t_turb .GT. t_old) .AND. (t_turb .LE. t_old + my_dt)) THEN
      drive_next = .TRUE.                           ! driver.f90:57-70
      t_turb = MAX(t_turb + dt_turb, t_sim + 1d-2 * my_dt)
      my_dt = 0.5d0 * my_dt
    END IF
    IF (mhd_solver_type .EQ. 1) THEN
      CALL tvd_solve_3d(my_dt)
    ELSE
      CALL vanleer_solve_3d(my_dt, drive_next)
    END IF
    IF (npass .GT. 0) THEN
      CALL copy_velocity(.TRUE.)
      CALL passive_solve_3d(my_dt)
    END IF
    t_sim = t_sim + my_dt
    IF (drive_turbulence .AND. drive_next) THEN    ! driver.f90:102-111
      CALL generate_rand_force()
      drive_next = .FALSE.
      COUR = 0.5d0 * COUR_def
    END IF
    cmax = 0.d0                                     ! compute_new_time_step, driver.f90:432-448
    IF (mhd_solver_type .EQ. 1) THEN
      CALL get_max_speed_tvd(cmax)
    ELSE
      CALL get_max_speed_vanleer(cmax)
    END IF
    my_dt = COUR * my_dx / cmax
    CALL boundary_conditions()
  END DO

  OPEN(UNIT=12, FILE=TRIM(fout), FORM='UNFORMATTED', ACCESS='STREAM', STATUS='REPLACE')
  WRITE(12) q_3d
  WRITE(12) bface_3d
  WRITE(12) qpass_3d
  WRITE(12) t_sim
  WRITE(12) my_dt
  WRITE(12) t_turb
  WRITE(12) DBLE(s1), DBLE(s2), DBLE(s3)
  CLOSE(12)

  IF (mhd_solver_type .EQ. 1) THEN
    CALL destroy_solver_tvd()
  ELSE
    CALL destroy_solver_vanleer()
  END IF
  PRINT*, 'imhd example driver: ', nsteps, ' steps, t = ', t_sim
END PROGRAM imhd_example_driver
