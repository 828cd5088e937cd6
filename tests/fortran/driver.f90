! Fortran host driving the HIP library through include/cylpic.f90, in the
! call order of cylindrical EPOCH's main loop (src/epoch2d.F90:213-265):
! update_eb_fields_half, push_particles, current_finish,
! update_eb_fields_final.  Reads a run description and particles from a
! stream file, writes particles, mode energies and two mode arrays back.
! Built as a shared library; cylpic_fortran_driver(fin, fout) returns 0 on
! success, else the failing call's status after printing its message.
FUNCTION cylpic_fortran_driver(cfin, cfout) BIND(C, NAME='cylpic_fortran_driver') RESULT(status)

  USE cylpic
  IMPLICIT NONE

  CHARACTER(KIND=C_CHAR), INTENT(IN) :: cfin(*), cfout(*)
  INTEGER(C_INT) :: status
  CHARACTER(LEN=512) :: fin, fout
  INTEGER(C_INT) :: nx, ny, n_mode, nsteps, prec, bcf(4), bcp(4), tx, tr, ierr
  REAL(C_DOUBLE) :: dx, dy, dt, x_min, charge, mass
  INTEGER(C_INT64_T) :: n, n_out
  REAL(C_DOUBLE), ALLOCATABLE :: x(:), y(:), z(:), px(:), py(:), pz(:), w(:)
  REAL(C_DOUBLE), ALLOCATABLE :: w_e(:), w_b(:)
  COMPLEX(C_DOUBLE_COMPLEX), ALLOCATABLE :: f(:,:,:)
  TYPE(C_PTR) :: h
  INTEGER :: istep, u

  status = 0
  CALL c_to_f(cfin, fin)
  CALL c_to_f(cfout, fout)
  OPEN(NEWUNIT=u, FILE=TRIM(fin), ACCESS='stream', FORM='unformatted', STATUS='old')
  READ(u) nx, ny, n_mode, nsteps, prec, bcf, bcp, tx, tr
  READ(u) dx, dy, dt, x_min, charge, mass
  READ(u) n
  ALLOCATE(x(n), y(n), z(n), px(n), py(n), pz(n), w(n))
  READ(u) x, y, z, px, py, pz, w
  CLOSE(u)

  ierr = cylpic_init(h, nx, ny, n_mode, 5_C_INT, dx, dy, dt, x_min, &
      x_min + nx * dx, 0.0_C_DOUBLE, ny * dy, bcf, bcp, prec, 0_C_INT)
  IF (failed('cylpic_init')) RETURN
  ierr = cylpic_set_tile(h, tx, tr)
  IF (failed('cylpic_set_tile')) RETURN
  ierr = cylpic_set_species(h, 0_C_INT, charge, mass, 0_C_INT, n)
  IF (failed('cylpic_set_species')) RETURN
  ierr = cylpic_upload_particles(h, 0_C_INT, n, x, y, z, px, py, pz, w, C_NULL_PTR)
  IF (failed('cylpic_upload_particles')) RETURN
  ierr = cylpic_setup_field_boundaries(h)
  IF (failed('cylpic_setup_field_boundaries')) RETURN

  DO istep = 1, nsteps
    ierr = cylpic_update_eb_fields_half(h)
    IF (failed('cylpic_update_eb_fields_half')) RETURN
    ierr = cylpic_push_particles(h)
    IF (failed('cylpic_push_particles')) RETURN
    ierr = cylpic_current_finish(h)
    IF (failed('cylpic_current_finish')) RETURN
    ierr = cylpic_update_eb_fields_final(h, C_NULL_PTR, C_NULL_PTR)
    IF (failed('cylpic_update_eb_fields_final')) RETURN
  END DO

  ierr = cylpic_download_particles(h, 0_C_INT, n, n_out, x, y, z, px, py, pz, w, C_NULL_PTR)
  IF (failed('cylpic_download_particles')) RETURN
  ALLOCATE(w_e(n_mode), w_b(n_mode))
  ierr = cylpic_mode_energy(h, w_e, w_b)
  IF (failed('cylpic_mode_energy')) RETURN
  ALLOCATE(f(1-5:nx+5, 1-5:ny+5, 0:n_mode-1))

  OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS='stream', FORM='unformatted', STATUS='replace')
  WRITE(u) n_out
  WRITE(u) x(1:n_out), y(1:n_out), z(1:n_out), px(1:n_out), py(1:n_out), pz(1:n_out), w(1:n_out)
  WRITE(u) w_e, w_b
  ierr = cylpic_download_field(h, CYLPIC_EXM, f)
  IF (failed('cylpic_download_field')) RETURN
  WRITE(u) f
  ierr = cylpic_download_field(h, CYLPIC_JRM, f)
  IF (failed('cylpic_download_field')) RETURN
  WRITE(u) f
  CLOSE(u)
  ierr = cylpic_finalize(h)
  IF (failed('cylpic_finalize')) RETURN
  PRINT '(A)', 'driver ok'

CONTAINS

  LOGICAL FUNCTION failed(what)
    CHARACTER(LEN=*), INTENT(IN) :: what
    failed = ierr /= 0
    IF (failed) THEN
      PRINT '(A,A,I0,A,A)', what, ' failed with ', ierr, ': ', TRIM(cylpic_error_message())
      FLUSH(6)
      status = ierr
    END IF
  END FUNCTION failed

  SUBROUTINE c_to_f(c, f)
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: c(*)
    CHARACTER(LEN=*), INTENT(OUT) :: f
    INTEGER :: i
    f = ''
    DO i = 1, LEN(f)
      IF (c(i) == C_NULL_CHAR) EXIT
      f(i:i) = c(i)
    END DO
  END SUBROUTINE c_to_f

END FUNCTION cylpic_fortran_driver
