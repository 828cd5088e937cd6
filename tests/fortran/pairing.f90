! Which element of a (0:ny) array a Fortran compiler pairs with row ir of a
! (1:ny) section in a non-conformable elemental expression, the form of
! outflow_bcs_x_min's "bxm(1,1:ny,im) / r_d_vals" and outflow_bcs_x_max's
! "source_t" terms (src/laser.f90:490-493, 599-605): allocatable arrays
! whose extents are known only at run time, as in the reference.  Prints
! "ir k" per row, k the index of the r_d_vals element that row ir used.
PROGRAM pairing
  IMPLICIT NONE
  INTEGER, PARAMETER :: num = KIND(1.d0)
  INTEGER :: nx, ny, ng, n_mode, ir, im
  CHARACTER(LEN=16) :: arg
  COMPLEX(num), ALLOCATABLE :: bxm(:,:,:), btm(:,:,:), source_t(:)
  REAL(num), ALLOCATABLE :: r_d_vals(:)
  COMPLEX(num), PARAMETER :: imagi = (0.0_num, 1.0_num)

  CALL GET_COMMAND_ARGUMENT(1, arg)
  READ(arg, *) ny
  nx = 4
  ng = 5
  n_mode = 2
  im = 1
  ALLOCATE(bxm(1-ng:nx+ng, 1-ng:ny+ng, 0:n_mode-1), btm(1-ng:nx+ng, 1-ng:ny+ng, 0:n_mode-1))
  ALLOCATE(r_d_vals(0:ny), source_t(0:ny))
  bxm = 1.0_num
  btm = 0.0_num
  DO ir = 0, ny
    r_d_vals(ir) = 1000.0_num + REAL(ir, num)
    source_t(ir) = REAL(ir, num)
  END DO
  ! the x_min form: a (1:ny) section over a (0:ny) array
  btm(1,1:ny,im) = - imagi * im * bxm(1,1:ny,im) / r_d_vals
  DO ir = 1, ny
    WRITE(*, '(A,1X,I0,1X,I0)') 'rd', ir, NINT(-1.0_num / AIMAG(btm(1,ir,im)) - 1000.0_num)
  END DO
  ! the x_max form: a whole (0:ny) array added to (1:ny) sections
  btm(nx,1:ny,im) = 2.0_num * source_t + 0.0_num * bxm(nx-1,1:ny,im)
  DO ir = 1, ny
    WRITE(*, '(A,1X,I0,1X,I0)') 'src', ir, NINT(REAL(btm(nx,ir,im)) / 2.0_num)
  END DO
END PROGRAM pairing
