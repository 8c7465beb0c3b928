! Configuration module of the example driver (VL2 + CT, HLLD, one passive scalar): the
! names the solver modules read from an application's constants_mod.
MODULE constants_mod
  IMPLICIT NONE
  INTEGER, PARAMETER :: nx = 16, ny = 12, nz = 10
  INTEGER, PARAMETER :: nb = 6
  INTEGER, PARAMETER :: npass = 1
  INTEGER, PARAMETER :: boundary_type = 2          ! 1 open, 2 periodic
  INTEGER, PARAMETER :: mhd_solver_type = 2        ! 1 TVD, 2 VL
  INTEGER, PARAMETER :: riemann_solver_type = 2    ! VL: 1 ROE, 2 HLLD, 3 HLLE
  REAL*8, PARAMETER :: COUR_def = 0.35d0
  LOGICAL, PARAMETER :: drive_turbulence = .FALSE.
  REAL*8, PARAMETER :: Edot = 1.d-3, dt_turb = 0.05d0   ! forcing (unused while not driven)
  INTEGER, PARAMETER :: k_max = 2
  INTEGER, PARAMETER :: nranks_x = 1, nranks_y = 1, nranks_z = 1
  INTEGER, PARAMETER :: s1_def = 1234, s2_def = 5678, s3_def = 9123
  REAL*8 :: sound_speed = 1.d0
  INTEGER :: myrank = 0
END MODULE constants_mod
