// type.hpp -- ids, reals and the binary edge record, as KnightKing
// include/type.hpp:30-85 defines them (same widths and layouts, so existing
// binary graph files load unchanged).
#pragma once

#include <stdint.h>
#include <utility>

typedef uint32_t vertex_id_t;
typedef uint64_t edge_id_t;
typedef uint8_t partition_id_t;
typedef uint32_t walker_id_t;
typedef uint32_t step_t;
typedef float real_t;

struct EmptyData {};

template <typename edge_data_t>
struct Edge {
    vertex_id_t src;
    vertex_id_t dst;
    edge_data_t data;
    Edge() {}
    Edge(vertex_id_t s, vertex_id_t d, edge_data_t x) : src(s), dst(d), data(x) {}
    void transpose() { std::swap(src, dst); }
};

template <>
struct Edge<EmptyData> {
    vertex_id_t src;
    union {
        vertex_id_t dst;
        EmptyData data;
    };
    Edge() {}
    Edge(vertex_id_t s, vertex_id_t d) : src(s), dst(d) {}
    void transpose() { std::swap(src, dst); }
};

enum GraphFormat { GF_Binary, GF_Edgelist };
