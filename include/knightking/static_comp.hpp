// static_comp.hpp -- get_trivial_static_comp (KnightKing
// src/apps/static_comp.hpp:29-42): no static component for unweighted graphs,
// the edge weight for weighted ones.  On the device this selects uniform
// candidates or the alias tables.
#pragma once

#include "walk.hpp"

template <typename walker_state_t>
std::function<real_t(vertex_id_t, AdjUnit<EmptyData> *)> get_trivial_static_comp(
    WalkEngine<EmptyData, walker_state_t> *)
{
    return nullptr;
}

template <typename walker_state_t>
std::function<real_t(vertex_id_t, AdjUnit<real_t> *)> get_trivial_static_comp(WalkEngine<real_t, walker_state_t> *)
{
    return [](vertex_id_t, AdjUnit<real_t> *e) { return e->data; };
}
