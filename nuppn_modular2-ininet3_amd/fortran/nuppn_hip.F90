! nuppn_hip - ISO_C_BINDING interfaces of the MI355X network-integration library
! (libnuppn_hip.so, C ABI in include/nuppn_hip.h) for NuPPN's Fortran solver.
!
! The derived types are bind(c) mirrors of the C structs; every interface below binds one C
! entry point by name.  Paths in the comments are relative to the NuPPN checkout.
!
! What each entry point replaces:
!   nuppn_ctx_create / nuppn_ctx_destroy     solver_init + makemap + the network broadcast
!                                            (solver/source/solver.F90:29,
!                                             solver/source/jac_rhs.F90:461-540,565-580)
!   nuppn_integrate_zone                     evaluate_all_rates at the cycle mid-point +
!                                            integrate_network (frames/ppn/source/ppn.F90:289-301,
!                                            solver/source/solver.F90:53-238)
!   nuppn_integrate_batch[_mixed]            mppnp's per-zone worker loop
!                                            (frames/mppnp/source/mppnp.F90:713-753)
!   nuppn_rates_eval                         reaclib_calculate_rates
!                                            (physics/source/reaclib.F90:431-582)
!   slu_solve                                solver/source/slu_f90/slu.f90:6-11 (same interface)
!   c_fortran_dgssv                          SuperLU_5.0/FORTRAN/c_fortran_dgssv.c:35 (the BD
!                                            step calls it with an implicit interface,
!                                            solver/source/bader_deuflhard.F90:438-464, so it
!                                            needs no declaration here)
module nuppn_hip
  use iso_c_binding
  implicit none
  private :: c_int, c_double, c_long_long, c_ptr

  ! integration codes (utils/source/errors.F90:5-16) and library errors
  integer(c_int), parameter, public :: NUPPN_OK = 0, NUPPN_SLV_CNVRG = 1, &
       NUPPN_MAX_SUBSTEPS = 2, NUPPN_BAD_STEP = 3, NUPPN_NEG_ABUNDS = 4, &
       NUPPN_HUGE_ABUNDS = 5, NUPPN_SUM_X = 6, NUPPN_GOT_NAN = 7, NUPPN_SMALL_DT = 11
  integer(c_int), parameter, public :: NUPPN_ERR_ARG = -1, NUPPN_ERR_HIP = -2, &
       NUPPN_ERR_PATTERN = -3, NUPPN_ERR_SINGULAR = -4, NUPPN_ERR_UNSUPPORTED = -5, &
       NUPPN_ERR_MIXING = -6

  ! reaction network: k(nre, 8) holds k1..k8 (0-based species indices and counts), rtype 0
  ! two-body, 1 photodisintegration, 2 decay; coef(nre, 7) the REACLIB a0..a6; aw / zw mass
  ! number and charge; igamma the 0-based gamma pseudo-species
  type, bind(c), public :: nuppn_network
     integer(c_int) :: nsp, nre
     type(c_ptr)    :: k, rtype, coef, aw, zw
     integer(c_int) :: igamma
  end type

  ! solver knobs (solver/source/solver_knobs.F90:36-81, same names)
  type, bind(c), public :: nuppn_knobs
     integer(c_int) :: ittd, irdn, max_sub_steps, integration_method
     integer(c_int) :: interpolate_temp, interpolate_dens
     real(c_double) :: dgrd, grdthreshold, cyminformal, min_rate_ov_dt
     real(c_double) :: min_t9_for_slv, delta_state_max
  end type

  ! per-zone diagnostics (solver/source/solver_diagnostics.F90:7-17 plus work counters)
  type, bind(c), public :: nuppn_zone_stats
     integer(c_int)       :: nsubt, titers, iter, nvar1, nsucc, nfail
     integer(c_long_long) :: n_newton, n_factor, n_solve, lu_nnz, n_refine, n_factor_fp64
  end type

  ! options of the batched path (nuppn_integrate_batch_mixed)
  type, bind(c), public :: nuppn_batch_opts
     integer(c_int) :: factor_precision, max_refine
     real(c_double) :: refine_tol, accept_tol
     integer(c_int) :: team_zones, team_waves, queue_passes, use_graph
  end type

  ! options of the mppnp cycle (nuppn_mppnp_cycle)
  type, bind(c), public :: nuppn_mppnp_opts
     real(c_double) :: t9threshold, sig_term_limit
     integer(c_int) :: code_source, path
  end type

  ! rate physics beyond the bare REACLIB sum (nuppn_set_rate_physics)
  type, bind(c), public :: nuppn_rate_physics
     integer(c_int) :: pf_n
     type(c_ptr)    :: pf_t9, pf_tab, rev_ch, rev_part
     integer(c_int) :: detailed_balance
     type(c_ptr)    :: g_spin, mass_g, q_mev
     integer(c_int) :: screen
  end type

  ! the autotuner's report (nuppn_autotune)
  type, bind(c), public :: nuppn_tune_result
     integer(c_int) :: block_size, team_zones, team_waves, from_cache, ncand, consistent
     integer(c_int) :: cand_nb(8), cand_tz(8), cand_tw(8)
     real(c_double) :: ms(8)
  end type

  public :: nuppn_default_knobs, nuppn_ctx_create, nuppn_ctx_destroy, nuppn_set_column_order, &
       nuppn_get_column_order, nuppn_column_order, nuppn_reduced_size, nuppn_reduced_pattern, &
       nuppn_rates_eval, nuppn_set_rate_cache, nuppn_rate_cache_table, nuppn_dxdt, &
       nuppn_jacobian_reduced, nuppn_lu_solve_reduced, nuppn_integrate_zone, &
       nuppn_integrate_batch, nuppn_integrate_batch_dev, nuppn_default_batch_opts, &
       nuppn_batch_block_info, nuppn_set_block_form, nuppn_integrate_batch_mixed, &
       nuppn_integrate_batch_mixed_dev, &
       nuppn_last_timing, slu_solve, slu_solve_noalloc, nuppn_default_mppnp_opts, &
       nuppn_mppnp_cycle, nuppn_mppnp_cycle_dev, nuppn_set_rate_physics, nuppn_reverse_pairs, &
       nuppn_evaluate_all_rates, nuppn_autotune

  interface
     subroutine nuppn_default_knobs(k) bind(c, name="nuppn_default_knobs")
       import :: nuppn_knobs
       type(nuppn_knobs), intent(out) :: k
     end subroutine

     type(c_ptr) function nuppn_ctx_create(net, knobs, status) bind(c, name="nuppn_ctx_create")
       import :: c_ptr, c_int, nuppn_network, nuppn_knobs
       type(nuppn_network), intent(in) :: net
       type(nuppn_knobs), intent(in)   :: knobs
       integer(c_int), intent(out)     :: status
     end function

     subroutine nuppn_ctx_destroy(ctx) bind(c, name="nuppn_ctx_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine

     integer(c_int) function nuppn_set_column_order(ctx, perm_c) &
          bind(c, name="nuppn_set_column_order")
       import :: c_ptr, c_int
       type(c_ptr), value         :: ctx
       integer(c_int), intent(in) :: perm_c(*)
     end function

     integer(c_int) function nuppn_get_column_order(ctx, perm_c) &
          bind(c, name="nuppn_get_column_order")
       import :: c_ptr, c_int
       type(c_ptr), value          :: ctx
       integer(c_int), intent(out) :: perm_c(*)
     end function

     integer(c_int) function nuppn_column_order(n, colptr, rowind, perm_c, etree) &
          bind(c, name="nuppn_column_order")
       import :: c_int
       integer(c_int), value       :: n
       integer(c_int), intent(in)  :: colptr(*), rowind(*)
       integer(c_int), intent(out) :: perm_c(*), etree(*)
     end function

     integer(c_int) function nuppn_reduced_size(ctx, n, nnz) bind(c, name="nuppn_reduced_size")
       import :: c_ptr, c_int
       type(c_ptr), value          :: ctx
       integer(c_int), intent(out) :: n, nnz
     end function

     integer(c_int) function nuppn_reduced_pattern(ctx, colptr, rowind, red2sp) &
          bind(c, name="nuppn_reduced_pattern")
       import :: c_ptr, c_int
       type(c_ptr), value          :: ctx
       integer(c_int), intent(out) :: colptr(*), rowind(*), red2sp(*)
     end function

     integer(c_int) function nuppn_rates_eval(ctx, nz, t9, rho, ye, v) &
          bind(c, name="nuppn_rates_eval")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value          :: ctx
       integer(c_int), value       :: nz
       real(c_double), intent(in)  :: t9(*), rho(*)
       type(c_ptr), value          :: ye        ! c_null_ptr (the synthetic networks have no ec)
       real(c_double), intent(out) :: v(*)
     end function

     integer(c_int) function nuppn_set_rate_cache(ctx, use_cache) &
          bind(c, name="nuppn_set_rate_cache")
       import :: c_ptr, c_int
       type(c_ptr), value    :: ctx
       integer(c_int), value :: use_cache
     end function

     integer(c_int) function nuppn_rate_cache_table(ctx, table) &
          bind(c, name="nuppn_rate_cache_table")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value          :: ctx
       real(c_double), intent(out) :: table(*)
     end function

     integer(c_int) function nuppn_dxdt(ctx, x, v, dxdt) bind(c, name="nuppn_dxdt")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value          :: ctx
       real(c_double), intent(in)  :: x(*), v(*)
       real(c_double), intent(out) :: dxdt(*)
     end function

     integer(c_int) function nuppn_jacobian_reduced(ctx, dt, x, v, csc_val) &
          bind(c, name="nuppn_jacobian_reduced")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value          :: ctx
       real(c_double), value       :: dt
       real(c_double), intent(in)  :: x(*), v(*)
       real(c_double), intent(out) :: csc_val(*)
     end function

     integer(c_int) function nuppn_lu_solve_reduced(ctx, csc_val, nrhs, b, perm_r) &
          bind(c, name="nuppn_lu_solve_reduced")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value            :: ctx
       real(c_double), intent(in)    :: csc_val(*)
       integer(c_int), value         :: nrhs
       real(c_double), intent(inout) :: b(*)
       integer(c_int), intent(out)   :: perm_r(*)
     end function

     integer(c_int) function nuppn_integrate_zone(ctx, yps, t9_0, t9_1, rho_0, rho_1, ye, &
                                                  dtwant, ierr, stats) &
          bind(c, name="nuppn_integrate_zone")
       import :: c_ptr, c_int, c_double, nuppn_zone_stats
       type(c_ptr), value                  :: ctx
       real(c_double), intent(inout)       :: yps(*), ye
       real(c_double), value               :: t9_0, t9_1, rho_0, rho_1, dtwant
       integer(c_int), intent(out)         :: ierr
       type(nuppn_zone_stats), intent(out) :: stats
     end function

     integer(c_int) function nuppn_integrate_batch(ctx, nz, yps, t9_0, t9_1, rho_0, rho_1, ye, &
                                                   dtwant, ierr, stats) &
          bind(c, name="nuppn_integrate_batch")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value            :: ctx
       integer(c_int), value         :: nz
       real(c_double), intent(inout) :: yps(*), ye(*)
       real(c_double), intent(in)    :: t9_0(*), t9_1(*), rho_0(*), rho_1(*), dtwant(*)
       integer(c_int), intent(out)   :: ierr(*)
       type(c_ptr), value            :: stats   ! c_loc(stats(1)) or c_null_ptr
     end function

     integer(c_int) function nuppn_integrate_batch_dev(ctx, nz, yps, t9_0, t9_1, rho_0, rho_1, &
                                                       ye, dtwant, ierr, stats, stream) &
          bind(c, name="nuppn_integrate_batch_dev")
       import :: c_ptr, c_int
       type(c_ptr), value    :: ctx
       integer(c_int), value :: nz
       type(c_ptr), value    :: yps, t9_0, t9_1, rho_0, rho_1, ye, dtwant, ierr, stats, stream
     end function

     subroutine nuppn_default_batch_opts(o) bind(c, name="nuppn_default_batch_opts")
       import :: nuppn_batch_opts
       type(nuppn_batch_opts), intent(out) :: o
     end subroutine

     integer(c_int) function nuppn_batch_block_info(ctx, info, ninfo) &
          bind(c, name="nuppn_batch_block_info")
       import :: c_ptr, c_int
       type(c_ptr), value          :: ctx
       integer(c_int), intent(out) :: info(*)
       integer(c_int), value       :: ninfo
     end function

     integer(c_int) function nuppn_set_block_form(ctx, nb) bind(c, name="nuppn_set_block_form")
       import :: c_ptr, c_int
       type(c_ptr), value    :: ctx
       integer(c_int), value :: nb
     end function

     integer(c_int) function nuppn_integrate_batch_mixed(ctx, nz, yps, t9_0, t9_1, rho_0, &
                                                         rho_1, ye, dtwant, ierr, stats, opts) &
          bind(c, name="nuppn_integrate_batch_mixed")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value            :: ctx
       integer(c_int), value         :: nz
       real(c_double), intent(inout) :: yps(*), ye(*)
       real(c_double), intent(in)    :: t9_0(*), t9_1(*), rho_0(*), rho_1(*), dtwant(*)
       integer(c_int), intent(out)   :: ierr(*)
       type(c_ptr), value            :: stats, opts   ! c_null_ptr: no stats / default options
     end function

     integer(c_int) function nuppn_integrate_batch_mixed_dev(ctx, nz, yps, t9_0, t9_1, rho_0, &
                                                             rho_1, ye, dtwant, ierr, stats, &
                                                             opts, stream) &
          bind(c, name="nuppn_integrate_batch_mixed_dev")
       import :: c_ptr, c_int
       type(c_ptr), value    :: ctx
       integer(c_int), value :: nz
       type(c_ptr), value    :: yps, t9_0, t9_1, rho_0, rho_1, ye, dtwant, ierr, stats, opts, &
                                stream
     end function

     integer(c_int) function nuppn_set_rate_physics(ctx, rp) bind(c, name="nuppn_set_rate_physics")
       import :: c_ptr, c_int, nuppn_rate_physics
       type(c_ptr), value                   :: ctx
       type(nuppn_rate_physics), intent(in) :: rp
     end function

     integer(c_int) function nuppn_reverse_pairs(ctx, irev) bind(c, name="nuppn_reverse_pairs")
       import :: c_ptr, c_int
       type(c_ptr), value          :: ctx
       integer(c_int), intent(out) :: irev(*)
     end function

     ! evaluate_all_rates (physics/source/evaluate_rates.F90:14-172) at abundances x(nsp, nz)
     integer(c_int) function nuppn_evaluate_all_rates(ctx, nz, t9, rho, x, v) &
          bind(c, name="nuppn_evaluate_all_rates")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value          :: ctx
       integer(c_int), value       :: nz
       real(c_double), intent(in)  :: t9(*), rho(*), x(*)
       real(c_double), intent(out) :: v(*)
     end function

     ! batched-path shapes chosen by timing a sample batch (no counterpart in the reference)
     integer(c_int) function nuppn_autotune(ctx, nz, yps, t9_0, t9_1, rho_0, rho_1, dtwant, &
                                            cache_path, res) bind(c, name="nuppn_autotune")
       import :: c_ptr, c_int, c_double, nuppn_tune_result
       type(c_ptr), value                   :: ctx
       integer(c_int), value                :: nz
       real(c_double), intent(in)           :: yps(*), t9_0(*), t9_1(*), rho_0(*), rho_1(*)
       real(c_double), intent(in)           :: dtwant(*)
       type(c_ptr), value                   :: cache_path   ! c_loc of a NUL-terminated name
       type(nuppn_tune_result), intent(out) :: res
     end function

     subroutine nuppn_default_mppnp_opts(o) bind(c, name="nuppn_default_mppnp_opts")
       import :: nuppn_mppnp_opts
       type(nuppn_mppnp_opts), intent(out) :: o
     end subroutine

     ! one mppnp cycle (frames/mppnp/source/mppnp.F90:637-781); yps is zone-major, yps(nsp, m)
     integer(c_int) function nuppn_mppnp_cycle(ctx, m, yps, t9, rho, xm, dq, rad, d_eul, dzeit, &
                                               opts, bopts, ierr, stats) &
          bind(c, name="nuppn_mppnp_cycle")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value            :: ctx
       integer(c_int), value         :: m
       real(c_double), intent(inout) :: yps(*)
       real(c_double), intent(in)    :: t9(*), rho(*), xm(*), dq(*), rad(*), d_eul(*)
       real(c_double), value         :: dzeit
       type(c_ptr), value            :: opts, bopts   ! c_null_ptr: defaults
       integer(c_int), intent(out)   :: ierr(*)
       type(c_ptr), value            :: stats         ! c_loc(stats(1)) or c_null_ptr
     end function

     integer(c_int) function nuppn_mppnp_cycle_dev(ctx, m, yps, t9, rho, xm, dq, rad, d_eul, &
                                                   dzeit, opts, bopts, ierr, stats, stream) &
          bind(c, name="nuppn_mppnp_cycle_dev")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: ctx
       integer(c_int), value :: m
       type(c_ptr), value    :: yps, t9, rho, xm, dq, rad, d_eul
       real(c_double), value :: dzeit
       type(c_ptr), value    :: opts, bopts, ierr, stats, stream
     end function

     integer(c_int) function nuppn_last_timing(ctx, out, nout) bind(c, name="nuppn_last_timing")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value          :: ctx
       real(c_double), intent(out) :: out(*)
       integer(c_int), value       :: nout
     end function

     ! solver/source/slu_f90/slu.f90:6-11: the same interface; colind / rowptr are shifted
     ! to 0-based in place when shiftindex is 1 (the caller restores them, linalg.F90:117-118)
     subroutine slu_solve(n, nnz, isrow, shiftindex, val, colind, rowptr, rhs, ierr) &
          bind(c, name="slu_solve")
       import :: c_int, c_double
       integer(c_int), value         :: n, nnz, isrow, shiftindex
       real(c_double), intent(in)    :: val(*)
       integer(c_int), intent(inout) :: colind(*), rowptr(*)
       real(c_double), intent(inout) :: rhs(*)
       integer(c_int), intent(out)   :: ierr
     end subroutine

     ! slu_solve for callers that cannot allocate device memory (no counterpart in the
     ! reference; n up to about 2,000)
     subroutine slu_solve_noalloc(n, nnz, isrow, shiftindex, val, colind, rowptr, rhs, ierr) &
          bind(c, name="slu_solve_noalloc")
       import :: c_int, c_double
       integer(c_int), value         :: n, nnz, isrow, shiftindex
       real(c_double), intent(in)    :: val(*)
       integer(c_int), intent(inout) :: colind(*), rowptr(*)
       real(c_double), intent(inout) :: rhs(*)
       integer(c_int), intent(out)   :: ierr
     end subroutine
  end interface
end module nuppn_hip
