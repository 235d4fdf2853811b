// ppr.hpp -- ppr() of KnightKing src/apps/ppr.hpp:30-64: continue with
// probability 1 - terminate_prob at each step; walker w starts at w % |V| or at
// (*start_vertices)[w % n].  The device caps a path at kkw's default length
// (probability of reaching it below 1e-12) and flags walkers that hit it.
#pragma once

#include "static_comp.hpp"

template <typename edge_data_t>
void ppr(WalkEngine<edge_data_t, EmptyData> *graph, walker_id_t walker_num, real_t terminate_prob,
         std::vector<vertex_id_t> *start_vertices = nullptr, WalkConfig *walk_conf = nullptr)
{
    real_t extension_comp = 1 - terminate_prob;
    WalkerConfig<edge_data_t, EmptyData> walker_conf(walker_num);
    TransitionConfig<edge_data_t, EmptyData> tr_conf(
        [extension_comp](Walker<EmptyData> &, vertex_id_t) { return extension_comp; },
        get_trivial_static_comp(graph));
    auto &d = tr_conf.device;
    d.present = true;
    d.job.kind = KKW_PPR;
    d.job.use_weights = knightking_detail::edge_traits<edge_data_t>::weighted ? 1 : 0;
    d.job.terminate_prob = terminate_prob;
    d.job.walk_length = 0;
    if (start_vertices != nullptr) {
        d.starts.assign(start_vertices->begin(), start_vertices->end());
        d.job.start_mode = KKW_START_LIST;
    } else {
        d.job.start_mode = KKW_START_EQUAL;
    }
    graph->random_walk(&walker_conf, &tr_conf, walk_conf);
}
