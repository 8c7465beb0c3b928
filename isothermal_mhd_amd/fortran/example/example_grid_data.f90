! State arrays of the example driver, in the layout the solver modules expect of an
! application's grid_data_mod: q_3d(7) and bface_3d(3) padded by nb, qpass_3d by nb-2,
! the turbulence force field delv_re (allocated only when driven) and the driver's taus88
! state.
MODULE grid_data_mod
  USE constants_mod
  IMPLICIT NONE
  REAL*8, ALLOCATABLE :: q_3d(:,:,:,:), bface_3d(:,:,:,:), qpass_3d(:,:,:,:), delv_re(:,:,:,:)
  REAL*8 :: my_dt, my_dx
  INTEGER :: s1, s2, s3                            ! taus88 state of the turbulence driver
CONTAINS
  SUBROUTINE create_grid_arrays()
    ALLOCATE(q_3d(1-nb:nx+nb, 1-nb:ny+nb, 1-nb:nz+nb, 7))
    ALLOCATE(bface_3d(1-nb:nx+nb, 1-nb:ny+nb, 1-nb:nz+nb, 3))
    ALLOCATE(qpass_3d(3-nb:nx+nb-2, 3-nb:ny+nb-2, 3-nb:nz+nb-2, MAX(npass, 1)))
    q_3d = 0.d0
    bface_3d = 0.d0
    qpass_3d = 0.d0
    IF (drive_turbulence) ALLOCATE(delv_re(1-nb:nx+nb, 1-nb:ny+nb, 1-nb:nz+nb, 3))
  END SUBROUTINE create_grid_arrays
END MODULE grid_data_mod
