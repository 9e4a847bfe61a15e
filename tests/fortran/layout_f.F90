! Prints the size and field offsets of every bind(c) type of the nuppn_hip module (one line per
! type: name, c_sizeof, offsets in declaration order) and takes the address of every interface,
! so that linking against libnuppn_hip.so resolves each bound symbol.
program layout_f
  use iso_c_binding
  use nuppn_hip
  implicit none
  type(nuppn_network), target :: n
  type(nuppn_knobs), target :: k
  type(nuppn_zone_stats), target :: s
  type(nuppn_batch_opts), target :: o
  type(nuppn_mppnp_opts), target :: q
  type(nuppn_rate_physics), target :: rp
  type(nuppn_tune_result), target :: tr
  type(c_funptr) :: f(32)
  integer(c_intptr_t) :: b

  b = transfer(c_loc(n), b)
  write(*, '(a, 9(1x, i0))') 'nuppn_network', c_sizeof(n), off(c_loc(n%nsp)), &
       off(c_loc(n%nre)), off(c_loc(n%k)), off(c_loc(n%rtype)), off(c_loc(n%coef)), &
       off(c_loc(n%aw)), off(c_loc(n%zw)), off(c_loc(n%igamma))
  b = transfer(c_loc(k), b)
  write(*, '(a, 13(1x, i0))') 'nuppn_knobs', c_sizeof(k), off(c_loc(k%ittd)), &
       off(c_loc(k%irdn)), off(c_loc(k%max_sub_steps)), off(c_loc(k%integration_method)), &
       off(c_loc(k%interpolate_temp)), off(c_loc(k%interpolate_dens)), off(c_loc(k%dgrd)), &
       off(c_loc(k%grdthreshold)), off(c_loc(k%cyminformal)), off(c_loc(k%min_rate_ov_dt)), &
       off(c_loc(k%min_t9_for_slv)), off(c_loc(k%delta_state_max))
  b = transfer(c_loc(s), b)
  write(*, '(a, 13(1x, i0))') 'nuppn_zone_stats', c_sizeof(s), off(c_loc(s%nsubt)), &
       off(c_loc(s%titers)), off(c_loc(s%iter)), off(c_loc(s%nvar1)), off(c_loc(s%nsucc)), &
       off(c_loc(s%nfail)), off(c_loc(s%n_newton)), off(c_loc(s%n_factor)), &
       off(c_loc(s%n_solve)), off(c_loc(s%lu_nnz)), off(c_loc(s%n_refine)), &
       off(c_loc(s%n_factor_fp64))
  b = transfer(c_loc(o), b)
  write(*, '(a, 9(1x, i0))') 'nuppn_batch_opts', c_sizeof(o), off(c_loc(o%factor_precision)), &
       off(c_loc(o%max_refine)), off(c_loc(o%refine_tol)), off(c_loc(o%accept_tol)), &
       off(c_loc(o%team_zones)), off(c_loc(o%team_waves)), off(c_loc(o%queue_passes)), &
       off(c_loc(o%use_graph))
  b = transfer(c_loc(q), b)
  write(*, '(a, 5(1x, i0))') 'nuppn_mppnp_opts', c_sizeof(q), off(c_loc(q%t9threshold)), &
       off(c_loc(q%sig_term_limit)), off(c_loc(q%code_source)), off(c_loc(q%path))
  b = transfer(c_loc(rp), b)
  write(*, '(a, 11(1x, i0))') 'nuppn_rate_physics', c_sizeof(rp), off(c_loc(rp%pf_n)), &
       off(c_loc(rp%pf_t9)), off(c_loc(rp%pf_tab)), off(c_loc(rp%rev_ch)), &
       off(c_loc(rp%rev_part)), off(c_loc(rp%detailed_balance)), off(c_loc(rp%g_spin)), &
       off(c_loc(rp%mass_g)), off(c_loc(rp%q_mev)), off(c_loc(rp%screen))
  b = transfer(c_loc(tr), b)
  write(*, '(a, 11(1x, i0))') 'nuppn_tune_result', c_sizeof(tr), off(c_loc(tr%block_size)), &
       off(c_loc(tr%team_zones)), off(c_loc(tr%team_waves)), off(c_loc(tr%from_cache)), &
       off(c_loc(tr%ncand)), off(c_loc(tr%consistent)), off(c_loc(tr%cand_nb)), &
       off(c_loc(tr%cand_tz)), off(c_loc(tr%cand_tw)), off(c_loc(tr%ms))

  f(1) = c_funloc(nuppn_default_knobs)
  f(2) = c_funloc(nuppn_ctx_create)
  f(3) = c_funloc(nuppn_ctx_destroy)
  f(4) = c_funloc(nuppn_set_column_order)
  f(5) = c_funloc(nuppn_get_column_order)
  f(6) = c_funloc(nuppn_column_order)
  f(7) = c_funloc(nuppn_reduced_size)
  f(8) = c_funloc(nuppn_reduced_pattern)
  f(9) = c_funloc(nuppn_rates_eval)
  f(10) = c_funloc(nuppn_set_rate_cache)
  f(11) = c_funloc(nuppn_rate_cache_table)
  f(12) = c_funloc(nuppn_dxdt)
  f(13) = c_funloc(nuppn_jacobian_reduced)
  f(14) = c_funloc(nuppn_lu_solve_reduced)
  f(15) = c_funloc(nuppn_integrate_zone)
  f(16) = c_funloc(nuppn_integrate_batch)
  f(17) = c_funloc(nuppn_integrate_batch_dev)
  f(18) = c_funloc(nuppn_default_batch_opts)
  f(19) = c_funloc(nuppn_batch_block_info)
  f(20) = c_funloc(nuppn_integrate_batch_mixed)
  f(21) = c_funloc(nuppn_integrate_batch_mixed_dev)
  f(22) = c_funloc(nuppn_last_timing)
  f(23) = c_funloc(slu_solve)
  f(24) = c_funloc(slu_solve_noalloc)
  f(25) = c_funloc(nuppn_default_mppnp_opts)
  f(26) = c_funloc(nuppn_mppnp_cycle)
  f(27) = c_funloc(nuppn_mppnp_cycle_dev)
  f(28) = c_funloc(nuppn_set_rate_physics)
  f(29) = c_funloc(nuppn_reverse_pairs)
  f(30) = c_funloc(nuppn_evaluate_all_rates)
  f(31) = c_funloc(nuppn_autotune)
  f(32) = c_funloc(nuppn_set_block_form)
  write(*, '(a, 1x, i0)') 'bound_procedures', count([(c_associated(f(b)), b = 1, 32)])
contains
  integer(c_intptr_t) function off(p)
    type(c_ptr), intent(in) :: p
    off = transfer(p, off) - b
  end function
end program layout_f
