// node2vec.hpp -- node2vec() of KnightKing src/apps/node2vec.hpp:108-202:
// second-order walk with P_d = 1/p for the return edge, 1 for a neighbour of
// the previous vertex, 1/q otherwise, upper bound max(1, 1/q), lower bound
// min(1/p, 1, 1/q), the outlier appendix when 1/p exceeds the upper bound, and
// adjacency lists sorted before walking.
#pragma once

#include "static_comp.hpp"

struct Node2vecState {
    vertex_id_t previous_vertex;
};

struct Node2vecConf {
    real_t p;
    real_t q;
    walker_id_t walker_num;
    step_t walk_length;
};

template <typename edge_data_t>
void node2vec(WalkEngine<edge_data_t, Node2vecState> *graph, Node2vecConf conf, WalkConfig *walk_conf = nullptr)
{
    graph->sort_adjacency();
    WalkerConfig<edge_data_t, Node2vecState> walker_conf(
        conf.walker_num, nullptr,
        [](Walker<Node2vecState> &w, vertex_id_t cur, AdjUnit<edge_data_t> *) { w.data.previous_vertex = cur; });
    SecondOrderTransitionConfig<edge_data_t, Node2vecState, vertex_id_t, bool> tr_conf;
    auto &d = tr_conf.device;
    d.present = true;
    d.job.kind = KKW_NODE2VEC;
    d.job.use_weights = knightking_detail::edge_traits<edge_data_t>::weighted ? 1 : 0;
    d.job.walk_length = conf.walk_length;
    d.job.p = conf.p;
    d.job.q = conf.q;
    d.job.start_mode = KKW_START_EQUAL;
    graph->random_walk(&walker_conf, &tr_conf, walk_conf);
}
