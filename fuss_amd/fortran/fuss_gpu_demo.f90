!> Fortran driver of the C-ABI: the steady 2-D plate of the parity tests
!> (n x n cells on [0,1]^2, k=50, rho=7800, cp=460, 302 walls T=500/300 on
!> faces 1/2, 301 q=0 on faces 3/4, null faces 5/6, IC T=300 K uniform),
!> built in Fortran arrays with the reference's layouts and stepped with
!> fuss_step, then fuss_run, then the defect-correction steady solve
!> (fuss_refine).  Prints one line per step, the run and refine outcomes and
!> the final state checksum; tests/test_fortran_driver.py compares it with
!> the Python binding on the same case.
program fuss_gpu_demo
  use iso_c_binding
  use fuss_gpu_m
  implicit none
  integer, parameter :: n = 32, nT = 2000
  type(c_ptr) :: ctx
  real(c_double), allocatable :: nodes(:,:,:,:), T(:,:,:), matID(:,:,:), qvol(:,:,:)
  real(c_double) :: cp(nT), rho(nT), kap(nT), h(nT), hist_r(200), hist_t(200)
  integer(c_int), allocatable :: ib(:,:), sid(:)
  real(c_double), allocatable :: rb(:,:)
  type(fuss_step_params) :: p
  type(fuss_step_result) :: r
  type(fuss_refine_params) :: rp
  type(fuss_refine_result) :: rr
  integer(c_int) :: st, nb, e, i, j, k, f, it, ndone
  real(c_double) :: dx, s

  dx = 1d0 / n
  allocate(nodes(3, 0:n, 0:n, 0:1), T(n, n, 1), matID(n, n, 1), qvol(n, n, 1))
  do k = 0, 1; do j = 0, n; do i = 0, n
    nodes(:, i, j, k) = [i * dx, j * dx, k * dx]
  end do; end do; end do
  T = 300d0
  matID = 1d0
  qvol = 0d0
  do i = 1, nT
    cp(i) = 460d0; rho(i) = 7800d0; kap(i) = 50d0
  end do
  h(1) = rho(1) * cp(1)
  do i = 2, nT
    h(i) = h(i-1) + rho(i) * cp(i)
  end do

  ! BC list grouped by (block, face): faces 1..6
  nb = 2 * n + 2 * n + 2 * n * n
  allocate(ib(15, nb), rb(6, nb), sid(nb))
  ib = 0; rb = 0d0; sid = -1
  e = 0
  do f = 1, 6
    select case (f)
    case (1, 2)
      do j = 1, n
        e = e + 1
        ib(1:6, e) = [1, merge(1, n, f == 1), j, 1, f, 302]
        rb(2, e) = merge(500d0, 300d0, f == 1)
      end do
    case (3, 4)
      do i = 1, n
        e = e + 1
        ib(1:6, e) = [1, i, merge(1, n, f == 3), 1, f, 301]
      end do
    case (5, 6)
      do j = 1, n; do i = 1, n
        e = e + 1
        ib(1:6, e) = [1, i, j, 1, f, 0]
      end do; end do
    end select
  end do

  st = fuss_create(1, 1, FUSS_FP64, ctx);                                call check(st, 'fuss_create')
  st = fuss_set_table(ctx, 1, 1, nT, cp, rho, kap, h);                   call check(st, 'fuss_set_table')
  st = fuss_set_block(ctx, 1, n, n, 1, nodes, T, matID, qvol);           call check(st, 'fuss_set_block')
  st = fuss_set_bc(ctx, 1, nb, ib, rb, sid);                             call check(st, 'fuss_set_bc')
  st = fuss_setup(ctx, -1d0);                                            call check(st, 'fuss_setup')

  p%n_rk = 3
  p%vnn = 0.8d0
  p%integration = FUSS_PRIM
  p%compute_residual = 1
  do it = 1, 10
    st = fuss_step(ctx, 1, p, r);                                        call check(st, 'fuss_step')
    write(*, '(A,I6,ES24.16)') 'step', r%iter, sqrt(r%res_sq)
  end do
  st = fuss_run(ctx, 1, p, 200, 200, 1d-10, 1d30, ndone, hist_r, hist_t, r); call check(st, 'fuss_run')
  write(*, '(A,I6,ES24.16)') 'run', r%iter, sqrt(r%res_sq)
  st = fuss_refine_setup(ctx, rp);                                       call check(st, 'fuss_refine_setup')
  st = fuss_refine(ctx, p, rr, c_null_ptr);                              call check(st, 'fuss_refine')
  write(*, '(A,2I6,ES24.16)') 'refine', rr%outer, rr%converged, rr%res
  st = fuss_get_T(ctx, 1, 1, T);                                         call check(st, 'fuss_get_T')
  s = sum(T)
  write(*, '(A,ES24.16)') 'sumT', s
  call fuss_destroy(ctx)

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(*), intent(in) :: what
    if (code /= FUSS_OK) then
      write(*, '(A,A,I3)') 'error in ', what, code
      stop 1
    end if
  end subroutine check

end program fuss_gpu_demo
