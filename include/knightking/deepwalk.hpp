// deepwalk.hpp -- deepwalk() of KnightKing src/apps/deepwalk.hpp:30-47:
// truncated walk of walk_length steps with the trivial static component.
#pragma once

#include "static_comp.hpp"

template <typename edge_data_t>
void deepwalk(WalkEngine<edge_data_t, EmptyData> *graph, walker_id_t walker_num, step_t walk_length,
              WalkConfig *walk_conf = nullptr)
{
    WalkerConfig<edge_data_t, EmptyData> walker_conf(walker_num);
    TransitionConfig<edge_data_t, EmptyData> tr_conf(
        [walk_length](Walker<EmptyData> &w, vertex_id_t) { return w.step >= walk_length ? 0.0f : 1.0f; },
        get_trivial_static_comp(graph));
    tr_conf.device.present = true;
    tr_conf.device.job.kind = KKW_DEEPWALK;
    tr_conf.device.job.use_weights = knightking_detail::edge_traits<edge_data_t>::weighted ? 1 : 0;
    tr_conf.device.job.walk_length = walk_length;
    tr_conf.device.job.start_mode = KKW_START_EQUAL;
    graph->random_walk(&walker_conf, &tr_conf, walk_conf);
}
