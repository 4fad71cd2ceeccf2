!> iso_c_binding interface of libfuss_gpu.so (include/fuss_gpu.h).
!>
!> This is the module a FUSS Fortran driver uses to replace the OpenMP loops
!> of Explicit_Step (src/lib/numerics/time/explicit/Mod_Explicit.f90:10-111)
!> by the MI355X kernels: FUSS_setup hands the flattened blocks, BC lists and
!> tables over once, FUSS_solve calls fuss_step / fuss_run per iteration and
!> fuss_prolong on a multigrid level switch, FUSS_postprocess reads T back
!> with fuss_get_T on its output cadence.
module fuss_gpu_m
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: FUSS_OK = 0, FUSS_EARG = 1, FUSS_ESTATE = 2, FUSS_ETABLE = 3, &
                                       FUSS_EMETRIC = 4, FUSS_EUNSUPPORTED = 5, FUSS_EHIP = 6, FUSS_EBC = 7
  integer(c_int), parameter, public :: FUSS_FP64 = 0, FUSS_FP32 = 1
  integer(c_int), parameter, public :: FUSS_CONS = 0, FUSS_PRIM = 1
  integer(c_int), parameter, public :: FUSS_IRS_JACOBI2 = 0, FUSS_IRS_TRIDIAG = 1

  !> numerics of one step (fuss_step_params)
  type, bind(c), public :: fuss_step_params
    integer(c_int) :: n_rk = 1
    real(c_double) :: vnn = 0.3d0
    integer(c_int) :: rampa_iter = 0
    integer(c_int) :: time_accurate = 0
    integer(c_int) :: integration = FUSS_CONS
    real(c_double) :: irs_beta = 0d0
    integer(c_int) :: irs_mode = FUSS_IRS_JACOBI2
    integer(c_int) :: compute_residual = 1
  end type

  !> outputs of one step (fuss_step_result)
  type, bind(c), public :: fuss_step_result
    real(c_double) :: res_sq, dtglobal, time
    integer(c_int) :: iter, status
    integer(c_int) :: bijk(4)
  end type

  !> defect-correction steady solve (fuss_refine_params)
  type, bind(c), public :: fuss_refine_params
    integer(c_int) :: precision = FUSS_FP32
    integer(c_int) :: n_stage = 4
    real(c_double) :: lam_lo = 0.3d0, lam_hi = 2.0d0
    real(c_double) :: irs_beta = 0d0
    integer(c_int) :: nu1 = 1, nu2 = 1
    integer(c_int) :: vcycles = 2
    integer(c_int) :: max_levels = 0
    integer(c_int) :: ncoarse = 20
    integer(c_int) :: max_outer = 200
    real(c_double) :: res_threshold = 1d-10
    integer(c_int) :: per_block = 0
    integer(c_int) :: coarse_cg = 0
  end type

  !> outcome of fuss_refine (fuss_refine_result)
  type, bind(c), public :: fuss_refine_result
    integer(c_int) :: outer, vcycles, levels, converged
    real(c_double) :: res, defect
  end type

  public :: fuss_create, fuss_set_table, fuss_set_block, fuss_add_series, fuss_set_bc, fuss_setup
  public :: fuss_set_chimera
  public :: fuss_step, fuss_step_hook, fuss_run, fuss_restrict, fuss_prolong
  public :: fuss_refine_setup, fuss_refine
  public :: fuss_get_T, fuss_set_T, fuss_probe, fuss_get_clock, fuss_set_clock, fuss_ndir, fuss_destroy
  public :: fuss_wall

  interface
    integer(c_int) function fuss_create(nlevels, nblocks, precision, ctx) bind(c, name='fuss_create')
      import :: c_int, c_ptr
      integer(c_int), value :: nlevels, nblocks, precision
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function fuss_set_table(ctx, nmat, Ti1, nT, cp, rho, kappa, h) bind(c, name='fuss_set_table')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nmat, Ti1, nT
      real(c_double), intent(in) :: cp(*), rho(*), kappa(*), h(*)
    end function

    integer(c_int) function fuss_set_block(ctx, b, ni, nj, nk, nodes, T, matID, qvol) bind(c, name='fuss_set_block')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: b, ni, nj, nk
      real(c_double), intent(in) :: nodes(*), T(*), matID(*), qvol(*)
    end function

    integer(c_int) function fuss_add_series(ctx, periodic, n, t, v, id) bind(c, name='fuss_add_series')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: periodic, n
      real(c_double), intent(in) :: t(*), v(*)
      integer(c_int), intent(out) :: id
    end function

    integer(c_int) function fuss_set_bc(ctx, level, nbound, ib, rb, series) bind(c, name='fuss_set_bc')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level, nbound
      integer(c_int), intent(in) :: ib(*)
      real(c_double), intent(in) :: rb(*)
      integer(c_int), intent(in) :: series(*)
    end function

    ! chimera (102) donor lists: bc(i)%ni(1:2), %donorID(:,1:4), %volume_fraction
    integer(c_int) function fuss_set_chimera(ctx, level, n, entry, n1, n2, donors, vf) &
        bind(c, name='fuss_set_chimera')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level, n
      integer(c_int), intent(in) :: entry(*), n1(*), n2(*), donors(*)
      real(c_double), intent(in) :: vf(*)
    end function

    integer(c_int) function fuss_setup(ctx, initial_time) bind(c, name='fuss_setup')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: initial_time
    end function

    integer(c_int) function fuss_step(ctx, level, p, r) bind(c, name='fuss_step')
      import :: c_int, c_ptr, fuss_step_params, fuss_step_result
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
      type(fuss_step_params), intent(in) :: p
      type(fuss_step_result), intent(out) :: r
    end function

    !> fuss_step with the External_Function slot: fn(user, level, irk, block,
    !> ni, nj, nk, T, R) receives the block's stage input T and raw residual R
    !> (ni*nj*nk each, i fastest) between BC_Fluxes and RK_Newstate; it may
    !> change R.  fn is a bind(c) procedure (c_funloc), user is passed through.
    integer(c_int) function fuss_step_hook(ctx, level, p, fn, user, r) bind(c, name='fuss_step_hook')
      import :: c_int, c_ptr, c_funptr, fuss_step_params, fuss_step_result
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
      type(fuss_step_params), intent(in) :: p
      type(c_funptr), value :: fn
      type(c_ptr), value :: user
      type(fuss_step_result), intent(out) :: r
    end function

    integer(c_int) function fuss_run(ctx, level, p, max_iters, itermax_left, res_threshold, time_threshold, &
                                     n_done, hist_res, hist_time, last) bind(c, name='fuss_run')
      import :: c_int, c_ptr, c_double, fuss_step_params, fuss_step_result
      type(c_ptr), value :: ctx
      integer(c_int), value :: level, max_iters, itermax_left
      type(fuss_step_params), intent(in) :: p
      real(c_double), value :: res_threshold, time_threshold
      integer(c_int), intent(out) :: n_done
      real(c_double), intent(out) :: hist_res(*), hist_time(*)
      type(fuss_step_result), intent(out) :: last
    end function

    integer(c_int) function fuss_refine_setup(ctx, rp) bind(c, name='fuss_refine_setup')
      import :: c_int, c_ptr, fuss_refine_params
      type(c_ptr), value :: ctx
      type(fuss_refine_params), intent(in) :: rp
    end function

    integer(c_int) function fuss_refine(ctx, monitor, res, hist_res) bind(c, name='fuss_refine')
      import :: c_int, c_ptr, fuss_step_params, fuss_refine_result
      type(c_ptr), value :: ctx
      type(fuss_step_params), intent(in) :: monitor
      type(fuss_refine_result), intent(out) :: res
      type(c_ptr), value :: hist_res
    end function

    integer(c_int) function fuss_restrict(ctx, fine_level) bind(c, name='fuss_restrict')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: fine_level
    end function

    integer(c_int) function fuss_prolong(ctx, fine_level) bind(c, name='fuss_prolong')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: fine_level
    end function

    integer(c_int) function fuss_get_T(ctx, level, b, T) bind(c, name='fuss_get_T')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level, b
      real(c_double), intent(out) :: T(*)
    end function

    integer(c_int) function fuss_set_T(ctx, level, b, T) bind(c, name='fuss_set_T')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level, b
      real(c_double), intent(in) :: T(*)
    end function

    integer(c_int) function fuss_probe(ctx, n, bijk, out) bind(c, name='fuss_probe')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      integer(c_int), intent(in) :: bijk(*)
      real(c_double), intent(out) :: out(*)
    end function

    integer(c_int) function fuss_get_clock(ctx, level, time, iter) bind(c, name='fuss_get_clock')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
      real(c_double), intent(out) :: time
      integer(c_int), intent(out) :: iter
    end function

    integer(c_int) function fuss_set_clock(ctx, level, time, iter) bind(c, name='fuss_set_clock')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level, iter
      real(c_double), value :: time
    end function

    integer(c_int) function fuss_ndir(ctx, ndir) bind(c, name='fuss_ndir')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: ndir
    end function

    !> Extract_Properties_From_BC: Tw, qw of each BC entry (nbound each)
    integer(c_int) function fuss_wall(ctx, level, Tw, qw) bind(c, name='fuss_wall')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
      real(c_double), intent(out) :: Tw(*), qw(*)
    end function

    !> the same without device allocation (levels of at most FUSS_WALL_NOALLOC_MAX entries)
    integer(c_int) function fuss_wall_noalloc(ctx, level, Tw, qw) bind(c, name='fuss_wall_noalloc')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
      real(c_double), intent(out) :: Tw(*), qw(*)
    end function

    !> how fuss_run issues its steps: captured graphs, fused 2-D stages,
    !> persistent cooperative launches (enable = 0 or 1)
    integer(c_int) function fuss_graphs(ctx, enable) bind(c, name='fuss_graphs')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function

    integer(c_int) function fuss_fused(ctx, enable) bind(c, name='fuss_fused')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function

    integer(c_int) function fuss_persistent(ctx, enable) bind(c, name='fuss_persistent')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function

    !> column tiles of a 3-D level (kc planes, row march, stage / IRS rows)
    integer(c_int) function fuss_set_tiles(ctx, level, kc, row_march, stage_rows, irs_rows) &
        bind(c, name='fuss_set_tiles')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: level, kc, row_march, stage_rows, irs_rows
    end function

    integer(c_int) function fuss_autotune(ctx, level, p, steps, kc, row_march, ms_per_step) &
        bind(c, name='fuss_autotune')
      import :: c_int, c_ptr, c_double, fuss_step_params
      type(c_ptr), value :: ctx
      integer(c_int), value :: level, steps
      type(fuss_step_params), intent(in) :: p
      integer(c_int), intent(out) :: kc, row_march
      real(c_double), intent(out) :: ms_per_step
    end function

    subroutine fuss_destroy(ctx) bind(c, name='fuss_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
  end interface

end module fuss_gpu_m
