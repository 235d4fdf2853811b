// metapath.hpp -- metapath() of KnightKing src/apps/metapath.hpp:32-128:
// each walker draws a scheme at start; at state s it may only take edges whose
// type the scheme allows (P_d in {0,1}, upper bound 1) and stops at
// walk_length or where no out-edge type is allowed.
#pragma once

#include "metascheme.hpp"

typedef uint8_t scheme_mask_t;

inline std::vector<std::vector<scheme_mask_t> > get_scheme_mask(std::vector<std::vector<std::vector<bool> > > &schemes)
{
    std::vector<std::vector<scheme_mask_t> > ret(schemes.size());
    for (size_t i = 0; i < schemes.size(); i++) {
        ret[i].resize(schemes[i].size());
        for (size_t j = 0; j < schemes[i].size(); j++) {
            scheme_mask_t v = 0;
            for (size_t k = 0; k < schemes[i][j].size(); k++)
                if (schemes[i][j][k]) v |= (scheme_mask_t)(1u << k);
            ret[i][j] = v;
        }
    }
    return ret;
}

template <typename edge_data_t>
void metapath(WalkEngine<edge_data_t, MetapathState> *graph, std::vector<std::vector<std::vector<bool> > > schemes,
              walker_id_t walker_num, step_t walk_length, WalkConfig *walk_conf = nullptr)
{
    auto masks = get_scheme_mask(schemes);
    WalkerConfig<edge_data_t, MetapathState> walker_conf(walker_num);
    TransitionConfig<edge_data_t, MetapathState> tr_conf;
    auto &d = tr_conf.device;
    d.present = true;
    d.job.kind = KKW_METAPATH;
    d.job.use_weights = knightking_detail::edge_traits<edge_data_t>::weighted ? 1 : 0;
    d.job.walk_length = walk_length;
    d.job.start_mode = KKW_START_EQUAL;
    d.job.n_schemes = (uint32_t)masks.size();
    for (auto &m : masks) {
        if (m.size() < 1 || m.size() > 8) throw std::invalid_argument("metapath schemes hold 1 to 8 states");
        d.scheme_len.push_back((uint32_t)m.size());
        for (size_t k = 0; k < 8; k++) d.scheme_mask.push_back(k < m.size() ? m[k] : 0);
    }
    graph->random_walk(&walker_conf, &tr_conf, walk_conf);
}
