// node2vec through the second-order query / response functions, as
// KnightKing's node2vec() states it (src/apps/node2vec.hpp:141-196): a trial
// the lower bound does not settle posts a query to the previous vertex
// carrying the candidate (none at step 0); the previous vertex answers
// whether the candidate is among its neighbours (std::binary_search there,
// kkw_adjacent here); dynamic_comp turns the answer into 1/p, 1 or 1/q.
// The outlier pair is node2vec.hpp:47-106 (skipped at step 0).  Lists must
// be sorted (node2vec.hpp:122-125).
// param: 0 walk length, 1 upper, 2 lower, 3 1/p, 4 1/q, 5 overflow
// (1/p - upper), 6 weighted (1) or not (0).  Reads no edge index of a
// static draw (KKW_EDGE_INDEX 0).
#ifndef KKW_EDGE_INDEX
#define KKW_EDGE_INDEX 0
#endif
#define KKW_DYNAMIC 1
#define KKW_LOWER 1
#define KKW_OUTLIER 1
#define KKW_SECOND_ORDER 1

struct WalkerState {
    uint32_t prev;   // Node2vecState::previous_vertex
};
struct Query {
    uint32_t cand;   // stateQuery<vertex_id_t>::data: the candidate's id
};
struct Response {
    uint32_t found;  // stateResponse<bool>::data
};

KKW_FN void init_state(WalkerState &s, const kkw_walker &, uint32_t, const kkw_ctx &) { s.prev = 0; }
KKW_FN void update_state(WalkerState &s, const kkw_walker &, uint32_t cur, const kkw_edge &, const kkw_ctx &) { s.prev = cur; }
KKW_FN float extension(const WalkerState &, const kkw_walker &w, uint32_t, const kkw_ctx &c)
{
    return w.step >= (uint32_t)c.param[0] ? 0.0f : 1.0f;
}

KKW_FN bool post_query(const WalkerState &s, const kkw_walker &w, uint32_t, const kkw_edge &e, uint32_t &target,
                       Query &q, const kkw_ctx &)
{
    if (w.step == 0) return false;
    target = s.prev;
    q.cand = e.neighbour;
    return true;
}
KKW_FN void respond_query(uint32_t v, const Query &q, Response &r, const kkw_ctx &c)
{
    r.found = (uint32_t)kkw_adjacent(v, q.cand, c);
}
KKW_FN float dynamic_comp(const WalkerState &s, const kkw_walker &w, const Response &r, uint32_t,
                          const kkw_edge &e, const kkw_ctx &c)
{
    if (w.step == 0) return c.param[1];
    if (s.prev == e.neighbour) return c.param[3];
    return r.found ? 1.0f : c.param[4];
}
KKW_FN float upper_bound(uint32_t, const kkw_ctx &c) { return c.param[1]; }
KKW_FN float lower_bound(uint32_t, const kkw_ctx &c) { return c.param[2]; }

// std::lower_bound of x in the sorted list of v: global position (may be the end)
KKW_FN uint64_t n2v_lower_pos(uint32_t v, uint32_t x, const kkw_ctx &c)
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
KKW_FN void outlier_bound(const WalkerState &s, const kkw_walker &w, uint32_t cur, float &overflow, uint32_t &num,
                          const kkw_ctx &c)
{
    overflow = 0.0f;
    num = 0;
    if (!(c.param[5] > 0.0f) || w.step == 0) return;
    const uint64_t pos = n2v_lower_pos(cur, s.prev, c);
    if (pos >= c.row_ptr[cur + 1] || c.col[pos] != s.prev) return;
    overflow = c.param[6] != 0.0f ? c.weight[pos] * c.param[5] : c.param[5];
    num = 1;
}
KKW_FN int64_t outlier_search(const WalkerState &s, const kkw_walker &, uint32_t cur, uint32_t, const kkw_ctx &c)
{
    const uint64_t pos = n2v_lower_pos(cur, s.prev, c);
    return (pos < c.row_ptr[cur + 1] && c.col[pos] == s.prev) ? (int64_t)pos : -1;
}
