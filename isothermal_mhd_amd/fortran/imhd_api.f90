! Fortran (iso_c_binding) interface to the imhd C ABI declared in include/imhd.h.
! Generic: depends on no application module, so it can be compiled once.
MODULE imhd_api
  USE ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT), PARAMETER :: IMHD_OK = 0, IMHD_EINVAL = -1, IMHD_EHIP = -2, IMHD_ESTATE = -3, &
                               IMHD_ENOSPEED = -4, IMHD_EIO = -5
  INTEGER(C_INT), PARAMETER :: IMHD_TVD = 1, IMHD_VL = 2
  INTEGER(C_INT), PARAMETER :: IMHD_ROE = 1, IMHD_HLLD = 2, IMHD_HLLE = 3
  INTEGER(C_INT), PARAMETER :: IMHD_PLM_PRIM = 0, IMHD_PPM_PRIM = 1, IMHD_PPM_CHAR = 2
  INTEGER(C_INT), PARAMETER :: IMHD_BC_OPEN = 1, IMHD_BC_PERIODIC = 2

  TYPE, BIND(C) :: imhd_config
    INTEGER(C_INT) :: nx, ny, nz, nb, npass, integrator, riemann, recon, bc, precision, compensated, &
                      flux_protection
    REAL(C_DOUBLE) :: cs, mc_beta, cour
  END TYPE imhd_config

  TYPE, BIND(C) :: imhd_diag
    REAL(C_DOUBLE) :: ke, vrms_sq, momx, momy, momz, divb_sum, divb_max, mass
  END TYPE imhd_diag

  INTERFACE
    INTEGER(C_INT) FUNCTION imhd_create(cfg, ctx) BIND(C, name='imhd_create')
      IMPORT :: C_INT, C_PTR, imhd_config
      TYPE(imhd_config), INTENT(IN) :: cfg
      TYPE(C_PTR), INTENT(OUT) :: ctx
    END FUNCTION imhd_create

    INTEGER(C_INT) FUNCTION imhd_create_ensemble(cfg, nbox, ctx) BIND(C, name='imhd_create_ensemble')
      IMPORT :: C_INT, C_PTR, imhd_config
      TYPE(imhd_config), INTENT(IN) :: cfg(*)
      INTEGER(C_INT), VALUE :: nbox
      TYPE(C_PTR), INTENT(OUT) :: ctx
    END FUNCTION imhd_create_ensemble

    INTEGER(C_INT) FUNCTION imhd_destroy(ctx) BIND(C, name='imhd_destroy')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION imhd_destroy

    INTEGER(C_INT) FUNCTION imhd_set_state(ctx, box, q, bf, qp) BIND(C, name='imhd_set_state')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: box
      REAL(C_DOUBLE), INTENT(IN) :: q(*), bf(*), qp(*)
    END FUNCTION imhd_set_state

    INTEGER(C_INT) FUNCTION imhd_get_state(ctx, box, q, bf, qp) BIND(C, name='imhd_get_state')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: box
      REAL(C_DOUBLE), INTENT(INOUT) :: q(*), bf(*), qp(*)
    END FUNCTION imhd_get_state

    INTEGER(C_INT) FUNCTION imhd_step(ctx, dt, drive_next) BIND(C, name='imhd_step')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
      INTEGER(C_INT), VALUE :: drive_next
    END FUNCTION imhd_step

    INTEGER(C_INT) FUNCTION imhd_set_force(ctx, box, delv, edot, dt_turb) BIND(C, name='imhd_set_force')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: box
      REAL(C_DOUBLE), INTENT(IN) :: delv(*)
      REAL(C_DOUBLE), VALUE :: edot, dt_turb
    END FUNCTION imhd_set_force

    INTEGER(C_INT) FUNCTION imhd_generate_force(ctx, box, k_max, edot, dt_turb, seeds) &
        BIND(C, name='imhd_generate_force')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: box, k_max
      REAL(C_DOUBLE), VALUE :: edot, dt_turb
      INTEGER(C_INT), INTENT(INOUT) :: seeds(3)
    END FUNCTION imhd_generate_force

    INTEGER(C_INT) FUNCTION imhd_get_force(ctx, box, delv) BIND(C, name='imhd_get_force')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: box
      REAL(C_DOUBLE), INTENT(OUT) :: delv(*)
    END FUNCTION imhd_get_force

    INTEGER(C_INT) FUNCTION imhd_fill_ghosts(ctx) BIND(C, name='imhd_fill_ghosts')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION imhd_fill_ghosts

    INTEGER(C_INT) FUNCTION imhd_max_speed(ctx, cmax) BIND(C, name='imhd_max_speed')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: cmax(*)
    END FUNCTION imhd_max_speed

    INTEGER(C_INT) FUNCTION imhd_copy_velocity(ctx, stage) BIND(C, name='imhd_copy_velocity')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: stage
    END FUNCTION imhd_copy_velocity

    INTEGER(C_INT) FUNCTION imhd_passive_step(ctx, dt) BIND(C, name='imhd_passive_step')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
    END FUNCTION imhd_passive_step

    INTEGER(C_INT) FUNCTION imhd_diagnostics(ctx, out) BIND(C, name='imhd_diagnostics')
      IMPORT :: C_INT, C_PTR, imhd_diag
      TYPE(C_PTR), VALUE :: ctx
      TYPE(imhd_diag), INTENT(OUT) :: out(*)
    END FUNCTION imhd_diagnostics

    INTEGER(C_INT) FUNCTION imhd_floor_count(ctx, count) BIND(C, name='imhd_floor_count')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), INTENT(OUT) :: count
    END FUNCTION imhd_floor_count

    INTEGER(C_INT) FUNCTION imhd_device_bytes(ctx, bytes) BIND(C, name='imhd_device_bytes')
      IMPORT :: C_INT, C_PTR, C_LONG_LONG
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_LONG_LONG), INTENT(OUT) :: bytes
    END FUNCTION imhd_device_bytes

    INTEGER(C_INT) FUNCTION imhd_run(ctx, nsteps, t_end, use_graph) BIND(C, name='imhd_run')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: nsteps, use_graph
      REAL(C_DOUBLE), VALUE :: t_end
    END FUNCTION imhd_run

    INTEGER(C_INT) FUNCTION imhd_set_time(ctx, t, dt) BIND(C, name='imhd_set_time')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(IN) :: t(*), dt(*)
    END FUNCTION imhd_set_time

    INTEGER(C_INT) FUNCTION imhd_get_time(ctx, t, dt) BIND(C, name='imhd_get_time')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: t(*), dt(*)
    END FUNCTION imhd_get_time

    ! path: a NUL-terminated C_CHAR array, e.g. TRIM(name)//C_NULL_CHAR
    INTEGER(C_INT) FUNCTION imhd_write_restart(ctx, box, path, t_sim, dump_count, seeds) &
        BIND(C, name='imhd_write_restart')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_CHAR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: box, dump_count
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
      REAL(C_DOUBLE), VALUE :: t_sim
      INTEGER(C_INT), INTENT(IN) :: seeds(3)
    END FUNCTION imhd_write_restart

    INTEGER(C_INT) FUNCTION imhd_read_restart(ctx, box, path, t_sim, dump_count, seeds) &
        BIND(C, name='imhd_read_restart')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_CHAR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: box
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
      REAL(C_DOUBLE), INTENT(OUT) :: t_sim
      INTEGER(C_INT), INTENT(OUT) :: dump_count, seeds(3)
    END FUNCTION imhd_read_restart

    INTEGER(C_INT) FUNCTION imhd_write_cube(ctx, box, path, double_prec, npass_out) &
        BIND(C, name='imhd_write_cube')
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: box, double_prec, npass_out
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    END FUNCTION imhd_write_cube

    INTEGER(C_INT) FUNCTION imhd_sync(ctx) BIND(C, name='imhd_sync')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION imhd_sync

    TYPE(C_PTR) FUNCTION imhd_last_error_c(ctx) BIND(C, name='imhd_last_error')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION imhd_last_error_c

    INTEGER(C_SIZE_T) FUNCTION c_strlen(s) BIND(C, name='strlen')
      IMPORT :: C_SIZE_T, C_PTR
      TYPE(C_PTR), VALUE :: s
    END FUNCTION c_strlen
  END INTERFACE

CONTAINS

  ! imhd_last_error as a Fortran string (at most 512 characters)
  FUNCTION imhd_last_error(ctx) RESULT(msg)
    TYPE(C_PTR), INTENT(IN) :: ctx
    CHARACTER(LEN=512) :: msg
    TYPE(C_PTR) :: p
    CHARACTER(KIND=C_CHAR), POINTER :: chars(:)
    INTEGER :: n, i
    msg = ''
    p = imhd_last_error_c(ctx)
    IF (.NOT. C_ASSOCIATED(p)) RETURN
    n = INT(MIN(c_strlen(p), INT(512, C_SIZE_T)))
    CALL C_F_POINTER(p, chars, [n])
    DO i = 1, n
      msg(i:i) = chars(i)
    END DO
  END FUNCTION imhd_last_error

END MODULE imhd_api
