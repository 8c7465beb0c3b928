! Example host driver: the loop body of the reference's driver (driver.f90:35-160 --
! copy_velocity, mhd_update, copy_velocity, passive_solve_3d, compute_new_time_step,
! boundary_conditions) written against the solver modules' reference names, which
! resolve to imhd_shim.f90 and run on the GPU.
!
! usage: imhd_example_<cfg> INPUT OUTPUT NSTEPS
!   INPUT  raw FP64 stream: q_3d, bface_3d, qpass_3d (reference layout, column-major)
!   OUTPUT raw FP64 stream: q_3d, bface_3d, qpass_3d, t, dt after NSTEPS steps
PROGRAM imhd_example_driver
  USE constants_mod
  USE grid_data_mod
  USE VanLeer_Solver_mod
  USE Iso_TVD_Solver_mod
  USE passive_solver_mod
  USE boundary_conditions_mod
  IMPLICIT NONE
  CHARACTER(LEN=512) :: fin, fout, arg
  INTEGER :: nsteps, step
  REAL*8 :: t_sim, cmax
  LOGICAL :: found

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

  DO step = 1, nsteps
    IF (npass .GT. 0) CALL copy_velocity(.FALSE.)
    IF (mhd_solver_type .EQ. 1) THEN
      CALL tvd_solve_3d(my_dt)
    ELSE
      CALL vanleer_solve_3d(my_dt, .FALSE.)
    END IF
    IF (npass .GT. 0) THEN
      CALL copy_velocity(.TRUE.)
      CALL passive_solve_3d(my_dt)
    END IF
    t_sim = t_sim + my_dt
    cmax = 0.d0                                     ! compute_new_time_step, driver.f90:432-448
    IF (mhd_solver_type .EQ. 1) THEN
      CALL get_max_speed_tvd(cmax)
    ELSE
      CALL get_max_speed_vanleer(cmax)
    END IF
    my_dt = COUR_def * my_dx / cmax
    CALL boundary_conditions()
  END DO

  OPEN(UNIT=12, FILE=TRIM(fout), FORM='UNFORMATTED', ACCESS='STREAM', STATUS='REPLACE')
  WRITE(12) q_3d
  WRITE(12) bface_3d
  WRITE(12) qpass_3d
  WRITE(12) t_sim
  WRITE(12) my_dt
  CLOSE(12)

  IF (mhd_solver_type .EQ. 1) THEN
    CALL destroy_solver_tvd()
  ELSE
    CALL destroy_solver_vanleer()
  END IF
  PRINT*, 'imhd example driver: ', nsteps, ' steps, t = ', t_sim
END PROGRAM imhd_example_driver
