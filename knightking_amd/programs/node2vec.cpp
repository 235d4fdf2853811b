// node2vec as a walk program (src/apps/node2vec.hpp:108-197), outlier pair
// included (node2vec.hpp:47-106; skipped at step 0, where the reference reads
// an unset previous vertex).  Lists must be sorted (node2vec.hpp:122-125).
// The adjacency test uses the engine's neighbour hash sets (kkw_adjacent).
// param: 0 walk length, 1 upper, 2 lower, 3 1/p, 4 1/q, 5 overflow
// (1/p - upper), 6 weighted (1) or not (0).
// Reads no edge index: draws come from the engine's tables of vertex ids
// and rows (KKW_EDGE_INDEX 0).  A -DKKW_EDGE_INDEX=1 build uses the
// local-index tables instead; paths are the same.
#ifndef KKW_EDGE_INDEX
#define KKW_EDGE_INDEX 0
#endif
#define KKW_DYNAMIC 1
#define KKW_LOWER 1
#define KKW_OUTLIER 1

struct WalkerState {
    uint32_t prev;
};

KKW_FN void init_state(WalkerState &s, const kkw_walker &, uint32_t, const kkw_ctx &) { s.prev = 0; }
KKW_FN void update_state(WalkerState &s, const kkw_walker &, uint32_t cur, const kkw_edge &, const kkw_ctx &) { s.prev = cur; }
KKW_FN float extension(const WalkerState &, const kkw_walker &w, uint32_t, const kkw_ctx &c)
{
    return w.step >= (uint32_t)c.param[0] ? 0.0f : 1.0f;
}

// std::lower_bound of x in the sorted list of v: global position (may be the end)
KKW_FN uint64_t lower_pos(uint32_t v, uint32_t x, const kkw_ctx &c)
{
    uint64_t lo = c.row_ptr[v], len = c.row_ptr[v + 1] - lo;
    while (len > 0) {
        const uint64_t half = len >> 1;
        if (c.col[lo + half] < x) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

KKW_FN float dynamic_comp(const WalkerState &s, const kkw_walker &w, uint32_t, const kkw_edge &e, const kkw_ctx &c)
{
    if (w.step == 0) return c.param[1];
    const uint32_t cand = e.neighbour;
    if (cand == s.prev) return c.param[3];
    return kkw_adjacent(s.prev, cand, c) ? 1.0f : c.param[4];   // respond_query (node2vec.hpp:158-166)
}
KKW_FN float upper_bound(uint32_t, const kkw_ctx &c) { return c.param[1]; }
KKW_FN float lower_bound(uint32_t, const kkw_ctx &c) { return c.param[2]; }

KKW_FN void outlier_bound(const WalkerState &s, const kkw_walker &w, uint32_t cur, float &overflow, uint32_t &num,
                          const kkw_ctx &c)
{
    overflow = 0.0f;
    num = 0;
    if (!(c.param[5] > 0.0f) || w.step == 0) return;
    const uint64_t pos = lower_pos(cur, s.prev, c);
    if (pos >= c.row_ptr[cur + 1] || c.col[pos] != s.prev) return;
    overflow = c.param[6] != 0.0f ? c.weight[pos] * c.param[5] : c.param[5];
    num = 1;
}
KKW_FN int64_t outlier_search(const WalkerState &s, const kkw_walker &, uint32_t cur, uint32_t, const kkw_ctx &c)
{
    const uint64_t pos = lower_pos(cur, s.prev, c);
    return (pos < c.row_ptr[cur + 1] && c.col[pos] == s.prev) ? (int64_t)pos : -1;
}
