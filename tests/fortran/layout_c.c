/* Prints the size and field offsets of every struct of include/nuppn_hip.h in the format of
 * layout_f.F90 (one line per struct: name, sizeof, offsets in declaration order). */
#include <stddef.h>
#include <stdio.h>

#include "nuppn_hip.h"

#define O(T, f) (int) offsetof(T, f)

int main(void) {
    printf("nuppn_network %d %d %d %d %d %d %d %d %d\n", (int) sizeof(nuppn_network),
           O(nuppn_network, nsp), O(nuppn_network, nre), O(nuppn_network, k),
           O(nuppn_network, rtype), O(nuppn_network, coef), O(nuppn_network, aw),
           O(nuppn_network, zw), O(nuppn_network, igamma));
    printf("nuppn_knobs %d %d %d %d %d %d %d %d %d %d %d %d %d\n", (int) sizeof(nuppn_knobs),
           O(nuppn_knobs, ittd), O(nuppn_knobs, irdn), O(nuppn_knobs, max_sub_steps),
           O(nuppn_knobs, integration_method), O(nuppn_knobs, interpolate_temp),
           O(nuppn_knobs, interpolate_dens), O(nuppn_knobs, dgrd), O(nuppn_knobs, grdthreshold),
           O(nuppn_knobs, cyminformal), O(nuppn_knobs, min_rate_ov_dt),
           O(nuppn_knobs, min_t9_for_slv), O(nuppn_knobs, delta_state_max));
    printf("nuppn_zone_stats %d %d %d %d %d %d %d %d %d %d %d %d %d\n",
           (int) sizeof(nuppn_zone_stats), O(nuppn_zone_stats, nsubt),
           O(nuppn_zone_stats, titers), O(nuppn_zone_stats, iter), O(nuppn_zone_stats, nvar1),
           O(nuppn_zone_stats, nsucc), O(nuppn_zone_stats, nfail), O(nuppn_zone_stats, n_newton),
           O(nuppn_zone_stats, n_factor), O(nuppn_zone_stats, n_solve),
           O(nuppn_zone_stats, lu_nnz), O(nuppn_zone_stats, n_refine),
           O(nuppn_zone_stats, n_factor_fp64));
    printf("nuppn_batch_opts %d %d %d %d %d %d %d %d %d\n", (int) sizeof(nuppn_batch_opts),
           O(nuppn_batch_opts, factor_precision), O(nuppn_batch_opts, max_refine),
           O(nuppn_batch_opts, refine_tol), O(nuppn_batch_opts, accept_tol),
           O(nuppn_batch_opts, team_zones), O(nuppn_batch_opts, team_waves),
           O(nuppn_batch_opts, queue_passes), O(nuppn_batch_opts, use_graph));
    printf("nuppn_mppnp_opts %d %d %d %d %d\n", (int) sizeof(nuppn_mppnp_opts),
           O(nuppn_mppnp_opts, t9threshold), O(nuppn_mppnp_opts, sig_term_limit),
           O(nuppn_mppnp_opts, code_source), O(nuppn_mppnp_opts, path));
    printf("nuppn_rate_physics %d %d %d %d %d %d %d %d %d %d %d\n",
           (int) sizeof(nuppn_rate_physics), O(nuppn_rate_physics, pf_n),
           O(nuppn_rate_physics, pf_t9), O(nuppn_rate_physics, pf_tab),
           O(nuppn_rate_physics, rev_ch), O(nuppn_rate_physics, rev_part),
           O(nuppn_rate_physics, detailed_balance), O(nuppn_rate_physics, g_spin),
           O(nuppn_rate_physics, mass_g), O(nuppn_rate_physics, q_mev),
           O(nuppn_rate_physics, screen));
    printf("nuppn_tune_result %d %d %d %d %d %d %d %d %d %d %d\n",
           (int) sizeof(nuppn_tune_result), O(nuppn_tune_result, block_size),
           O(nuppn_tune_result, team_zones), O(nuppn_tune_result, team_waves),
           O(nuppn_tune_result, from_cache), O(nuppn_tune_result, ncand),
           O(nuppn_tune_result, consistent), O(nuppn_tune_result, cand_nb),
           O(nuppn_tune_result, cand_tz), O(nuppn_tune_result, cand_tw),
           O(nuppn_tune_result, ms));
    return 0;
}
