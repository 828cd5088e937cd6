! cylpic.f90 -- ISO_C_BINDING interface to the gfx950 quasi-3D PIC library
! (the C ABI declared in cylpic.h).  A Fortran host (cylindrical EPOCH's main
! loop, src/epoch2d.F90:213-265) USEs this module and calls the functions in
! place of update_eb_fields_half, push_particles, current_finish and
! update_eb_fields_final.  Every function returns 0 on success; a nonzero
! value maps to abort_code(c_err_other) (src/housekeeping/utilities.f90).
MODULE cylpic

  USE, INTRINSIC :: ISO_C_BINDING

  IMPLICIT NONE

  INTEGER(C_INT), PARAMETER :: CYLPIC_FP64 = 0, CYLPIC_MIXED = 1
  INTEGER(C_INT), PARAMETER :: CYLPIC_SORT_RADIX = 0, CYLPIC_SORT_MIGRATE = 1, &
                              CYLPIC_SORT_CELLS = 2
  ! distribution function directions (c_dir_* of constants.F90)
  INTEGER(C_INT), PARAMETER :: CYLPIC_DIR_X = 1, CYLPIC_DIR_Y = 2, CYLPIC_DIR_PX = 3, &
      CYLPIC_DIR_PY = 4, CYLPIC_DIR_PZ = 5, CYLPIC_DIR_EN = 6, CYLPIC_DIR_GAMMA_M1 = 7, &
      CYLPIC_DIR_XY_ANGLE = 8, CYLPIC_DIR_YZ_ANGLE = 9, CYLPIC_DIR_ZX_ANGLE = 10, &
      CYLPIC_DIR_MOD_P = 11

  ! field identifiers, in the order of src/shared_data.F90:474-478
  INTEGER(C_INT), PARAMETER :: CYLPIC_EXM = 0, CYLPIC_ERM = 1, CYLPIC_ETM = 2
  INTEGER(C_INT), PARAMETER :: CYLPIC_BXM = 3, CYLPIC_BRM = 4, CYLPIC_BTM = 5
  INTEGER(C_INT), PARAMETER :: CYLPIC_JXM = 6, CYLPIC_JRM = 7, CYLPIC_JTM = 8
  INTEGER(C_INT), PARAMETER :: CYLPIC_BXM_OLD = 9, CYLPIC_BRM_OLD = 10
  INTEGER(C_INT), PARAMETER :: CYLPIC_BTM_OLD = 11, CYLPIC_JXM_OLD = 12
  INTEGER(C_INT), PARAMETER :: CYLPIC_JRM_OLD = 13, CYLPIC_JTM_OLD = 14

  INTERFACE

    FUNCTION cylpic_init(h, nx, ny, n_mode, ng, dx, dy, dt, x_min, x_max, &
        y_min, y_max, bc_field, bc_particle, precision, device) &
        BIND(C, NAME='cylpic_init') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), INTENT(OUT) :: h
      INTEGER(C_INT), VALUE :: nx, ny, n_mode, ng
      REAL(C_DOUBLE), VALUE :: dx, dy, dt, x_min, x_max, y_min, y_max
      INTEGER(C_INT), INTENT(IN) :: bc_field(4), bc_particle(4)
      INTEGER(C_INT), VALUE :: precision, device
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_init

    FUNCTION cylpic_finalize(h) BIND(C, NAME='cylpic_finalize') RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_finalize

    FUNCTION cylpic_last_error() BIND(C, NAME='cylpic_last_error') RESULT(p)
      IMPORT :: C_PTR
      TYPE(C_PTR) :: p
    END FUNCTION cylpic_last_error

    FUNCTION cylpic_set_dt(h, dt) BIND(C, NAME='cylpic_set_dt') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), VALUE :: dt
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_set_dt

    FUNCTION cylpic_set_time(h, time) BIND(C, NAME='cylpic_set_time') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), VALUE :: time
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_set_time

    FUNCTION cylpic_get_restart_state(h, time, x_min, x_grid_min, push_count, tile_shift) &
        BIND(C, NAME='cylpic_get_restart_state') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE, C_INT64_T
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: time, x_min, x_grid_min
      INTEGER(C_INT64_T), INTENT(OUT) :: push_count
      INTEGER(C_INT), INTENT(OUT) :: tile_shift
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_get_restart_state

    FUNCTION cylpic_set_restart_state(h, time, x_min, x_grid_min, push_count, tile_shift) &
        BIND(C, NAME='cylpic_set_restart_state') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE, C_INT64_T
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), VALUE :: time, x_min, x_grid_min
      INTEGER(C_INT64_T), VALUE :: push_count
      INTEGER(C_INT), VALUE :: tile_shift
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_set_restart_state

    FUNCTION cylpic_set_species(h, id, charge, mass, immobile, capacity) &
        BIND(C, NAME='cylpic_set_species') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE, C_INT64_T
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: id
      REAL(C_DOUBLE), VALUE :: charge, mass
      INTEGER(C_INT), VALUE :: immobile
      INTEGER(C_INT64_T), VALUE :: capacity
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_set_species

    ! x, y, z, px, py, pz, w: the SoA staging arrays a loop over the species'
    ! linked list fills (pos, p, weight: src/housekeeping/partlist.F90:414-488)
    FUNCTION cylpic_upload_particles(h, species, n, x, y, z, px, py, pz, w, &
        ids) BIND(C, NAME='cylpic_upload_particles') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE, C_INT64_T
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: species
      INTEGER(C_INT64_T), VALUE :: n
      REAL(C_DOUBLE), INTENT(IN) :: x(*), y(*), z(*), px(*), py(*), pz(*), w(*)
      TYPE(C_PTR), VALUE :: ids
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_upload_particles

    FUNCTION cylpic_download_particles(h, species, capacity, n_out, x, y, z, &
        px, py, pz, w, ids) BIND(C, NAME='cylpic_download_particles') &
        RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE, C_INT64_T
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: species
      INTEGER(C_INT64_T), VALUE :: capacity
      INTEGER(C_INT64_T), INTENT(OUT) :: n_out
      REAL(C_DOUBLE), INTENT(OUT) :: x(*), y(*), z(*), px(*), py(*), pz(*), w(*)
      TYPE(C_PTR), VALUE :: ids
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_download_particles

    FUNCTION cylpic_particle_count(h, species, n_out) &
        BIND(C, NAME='cylpic_particle_count') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_INT64_T
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: species
      INTEGER(C_INT64_T), INTENT(OUT) :: n_out
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_particle_count

    ! f: COMPLEX(8) f(1-ng:nx+ng, 1-ng:ny+ng, 0:n_mode-1), e.g. exm
    FUNCTION cylpic_upload_field(h, which, f) &
        BIND(C, NAME='cylpic_upload_field') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE_COMPLEX
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: which
      COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: f(*)
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_upload_field

    FUNCTION cylpic_download_field(h, which, f) &
        BIND(C, NAME='cylpic_download_field') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE_COMPLEX
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: which
      COMPLEX(C_DOUBLE_COMPLEX), INTENT(OUT) :: f(*)
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_download_field

    FUNCTION cylpic_setup_field_boundaries(h) &
        BIND(C, NAME='cylpic_setup_field_boundaries') RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_setup_field_boundaries

    FUNCTION cylpic_update_eb_fields_half(h) &
        BIND(C, NAME='cylpic_update_eb_fields_half') RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_update_eb_fields_half

    FUNCTION cylpic_push_particles(h) BIND(C, NAME='cylpic_push_particles') &
        RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_push_particles

    FUNCTION cylpic_current_finish(h) BIND(C, NAME='cylpic_current_finish') &
        RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_current_finish

    ! src_xmin / src_xmax: source1(0:ny), source2(0:ny) from the deck's laser
    ! blocks (src/laser.f90:444-461), or C_NULL_PTR
    FUNCTION cylpic_update_eb_fields_final(h, src_xmin, src_xmax) &
        BIND(C, NAME='cylpic_update_eb_fields_final') RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h, src_xmin, src_xmax
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_update_eb_fields_final

    FUNCTION cylpic_step(h, src_xmin, src_xmax) BIND(C, NAME='cylpic_step') &
        RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h, src_xmin, src_xmax
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_step

    FUNCTION cylpic_sort_particles(h) BIND(C, NAME='cylpic_sort_particles') &
        RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_sort_particles

    FUNCTION cylpic_set_sort_interval(h, interval) &
        BIND(C, NAME='cylpic_set_sort_interval') RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: interval
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_set_sort_interval

    FUNCTION cylpic_inject_particles(h, species, n, x, y, z, px, py, pz, w, ids) &
        BIND(C, NAME='cylpic_inject_particles') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_INT64_T, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: species
      INTEGER(C_INT64_T), VALUE :: n
      REAL(C_DOUBLE), INTENT(IN) :: x(*), y(*), z(*), px(*), py(*), pz(*), w(*)
      TYPE(C_PTR), VALUE :: ids
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_inject_particles

    FUNCTION cylpic_move_window(h, ncells) BIND(C, NAME='cylpic_move_window') &
        RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: ncells
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_move_window

    FUNCTION cylpic_autotune(h, pushes, tile_x, tile_r, gather_lds, ms_per_push) &
        BIND(C, NAME='cylpic_autotune') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: pushes
      INTEGER(C_INT), INTENT(OUT) :: tile_x, tile_r, gather_lds
      REAL(C_DOUBLE), INTENT(OUT) :: ms_per_push
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_autotune

    FUNCTION cylpic_continuity_begin(h) BIND(C, NAME='cylpic_continuity_begin') RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_continuity_begin

    FUNCTION cylpic_continuity_end(h, residual, dq, q) &
        BIND(C, NAME='cylpic_continuity_end') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: residual, dq, q
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_continuity_end

    FUNCTION cylpic_continuity_modes(h, residual, dq, jtheta, z) &
        BIND(C, NAME='cylpic_continuity_modes') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: residual(*), dq(*), jtheta(*), z(*)
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_continuity_modes

    FUNCTION cylpic_gauss_residual_m0(h, residual, q, flux) &
        BIND(C, NAME='cylpic_gauss_residual_m0') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: residual, q, flux
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_gauss_residual_m0

    FUNCTION cylpic_gauss_residual_m0_noalloc(h, residual, q, flux) &
        BIND(C, NAME='cylpic_gauss_residual_m0_noalloc') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: residual, q, flux
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_gauss_residual_m0_noalloc

    FUNCTION cylpic_number_density_modes(h, species, out) &
        BIND(C, NAME='cylpic_number_density_modes') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE_COMPLEX
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: species
      COMPLEX(C_DOUBLE_COMPLEX), INTENT(OUT) :: out(*)
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_number_density_modes

    FUNCTION cylpic_dist_fn(h, species, ndims, direction, ranges, resolution, &
        restrictions, use_restrictions, out, capacity, nbins) &
        BIND(C, NAME='cylpic_dist_fn') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_INT64_T, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: species, ndims
      INTEGER(C_INT), INTENT(IN) :: direction(*)
      REAL(C_DOUBLE), INTENT(INOUT) :: ranges(*)
      INTEGER(C_INT), INTENT(INOUT) :: resolution(*)
      REAL(C_DOUBLE), INTENT(IN) :: restrictions(*)
      INTEGER(C_INT), INTENT(IN) :: use_restrictions(*)
      REAL(C_DOUBLE), INTENT(OUT) :: out(*)
      INTEGER(C_INT64_T), VALUE :: capacity
      INTEGER(C_INT64_T), INTENT(OUT) :: nbins
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_dist_fn

    FUNCTION cylpic_set_smoothing(h, its, comp_its, strides, nstrides) &
        BIND(C, NAME='cylpic_set_smoothing') RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: its, comp_its, nstrides
      INTEGER(C_INT), INTENT(IN) :: strides(*)
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_set_smoothing

    FUNCTION cylpic_set_sort_method(h, method) &
        BIND(C, NAME='cylpic_set_sort_method') RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: method
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_set_sort_method

    FUNCTION cylpic_sort_counts(h, radix, migrate) &
        BIND(C, NAME='cylpic_sort_counts') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_INT64_T
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT64_T), INTENT(OUT) :: radix, migrate
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_sort_counts

    FUNCTION cylpic_set_tile(h, tile_x, tile_r) BIND(C, NAME='cylpic_set_tile') &
        RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: tile_x, tile_r
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_set_tile

    FUNCTION cylpic_tune(h, acc_mode, gather_lds) BIND(C, NAME='cylpic_tune') &
        RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: acc_mode, gather_lds
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_tune

    FUNCTION cylpic_mode_energy(h, w_e, w_b) BIND(C, NAME='cylpic_mode_energy') &
        RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: w_e(*), w_b(*)
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_mode_energy

    FUNCTION cylpic_synchronize(h) BIND(C, NAME='cylpic_synchronize') &
        RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_synchronize

    FUNCTION cylpic_profile(h, enable) BIND(C, NAME='cylpic_profile') &
        RESULT(ierr)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: enable
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_profile

    FUNCTION cylpic_kernel_time(h, name, ms, count) &
        BIND(C, NAME='cylpic_kernel_time') RESULT(ierr)
      IMPORT :: C_PTR, C_INT, C_DOUBLE, C_INT64_T, C_CHAR
      TYPE(C_PTR), VALUE :: h
      CHARACTER(KIND=C_CHAR), INTENT(IN) :: name(*)
      REAL(C_DOUBLE), INTENT(OUT) :: ms
      INTEGER(C_INT64_T), INTENT(OUT) :: count
      INTEGER(C_INT) :: ierr
    END FUNCTION cylpic_kernel_time

  END INTERFACE

CONTAINS

  ! The library's last error message as a Fortran string.
  FUNCTION cylpic_error_message() RESULT(msg)
    CHARACTER(LEN=256) :: msg
    CHARACTER(KIND=C_CHAR), POINTER :: c(:)
    TYPE(C_PTR) :: p
    INTEGER :: i
    msg = ''
    p = cylpic_last_error()
    IF (.NOT. C_ASSOCIATED(p)) RETURN
    CALL C_F_POINTER(p, c, [256])
    DO i = 1, 256
      IF (c(i) == C_NULL_CHAR) EXIT
      msg(i:i) = c(i)
    END DO
  END FUNCTION cylpic_error_message

END MODULE cylpic
