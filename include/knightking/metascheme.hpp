// metascheme.hpp -- metapath edge data and scheme files (KnightKing
// src/apps/metascheme.hpp:36-116).
#pragma once

#include <assert.h>
#include <stdio.h>

#include <vector>

#include "walk.hpp"

typedef int scheme_id_t;
typedef int meta_state_t;

struct MetapathState {
    scheme_id_t scheme_id;
    meta_state_t state;
};

struct WeightedMetaData {
    real_t weight;
    meta_state_t meta_info;
    meta_state_t get_meta() { return meta_info; }
    real_t get_weight() { return weight; }
};

struct UnweightedMetaData {
    meta_state_t meta_info;
    meta_state_t get_meta() { return meta_info; }
    real_t get_weight() { return 1.0; }
};

namespace knightking_detail {
template <>
struct edge_traits<WeightedMetaData> {
    static const bool weighted = true, typed = true;
    static float weight(const WeightedMetaData &x) { return x.weight; }
    static int meta(const WeightedMetaData &x) { return x.meta_info; }
    static void set(WeightedMetaData &x, float w, int m)
    {
        x.weight = w;
        x.meta_info = m;
    }
};
template <>
struct edge_traits<UnweightedMetaData> {
    static const bool weighted = false, typed = true;
    static float weight(const UnweightedMetaData &) { return 1.0f; }
    static int meta(const UnweightedMetaData &x) { return x.meta_info; }
    static void set(UnweightedMetaData &x, float, int m) { x.meta_info = m; }
};
}  // namespace knightking_detail

// "scheme_num state_num" then per scheme "length" and length rows of
// state_num 0/1 flags (metascheme.hpp:69-94).
inline std::vector<std::vector<std::vector<bool> > > read_metapath_schemes(const char *path)
{
    FILE *f = fopen(path, "r");
    assert(f != NULL);
    int scheme_num = 0, state_num = 0;
    if (fscanf(f, "%d %d", &scheme_num, &state_num) != 2) scheme_num = 0;
    std::vector<std::vector<std::vector<bool> > > schemes(scheme_num);
    for (int s = 0; s < scheme_num; s++) {
        int len = 0;
        if (fscanf(f, "%d", &len) != 1) len = 0;
        schemes[s].resize(len);
        for (int l = 0; l < len; l++) {
            schemes[s][l].resize(state_num);
            for (int t = 0; t < state_num; t++) {
                int x = 0;
                if (fscanf(f, "%d", &x) != 1) x = 0;
                schemes[s][l][t] = x != 0;
            }
        }
    }
    fclose(f);
    return schemes;
}

inline void write_metapath_schemes(std::vector<std::vector<std::vector<bool> > > schemes, const char *path)
{
    FILE *f = fopen(path, "w");
    assert(f != NULL);
    fprintf(f, "%d %d\n", (int)schemes.size(), (int)schemes[0][0].size());
    for (auto &sc : schemes) {
        fprintf(f, "%d\n", (int)sc.size());
        for (auto &row : sc) {
            for (bool b : row) fprintf(f, "%d ", (int)b);
            fprintf(f, "\n");
        }
    }
    fclose(f);
}
