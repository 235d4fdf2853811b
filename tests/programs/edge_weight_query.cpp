// A second-order walk whose response carries data, and whose queries are
// sent on some trials only: on even steps from 2 on, a trial asks the
// previous vertex for the weight of its edge to the candidate (0 when there
// is none); dynamic_comp = 0.25 + 0.35 * that weight against an upper bound
// of 2.  Odd steps send nothing and see the walker's previous answer.  Used
// to check that the device loop and the oracle's loop pass queries, answers
// and the kept answer alike.  param: 0 walk length.
#define KKW_DYNAMIC 1
#define KKW_SECOND_ORDER 1

struct WalkerState {
    uint32_t prev;
};
struct Query {
    uint32_t cand;
};
struct Response {
    float weight;
    uint32_t answered;
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
    if (w.step == 0 || (w.step & 1u)) return false;
    target = s.prev;
    q.cand = e.neighbour;
    return true;
}
KKW_FN void respond_query(uint32_t v, const Query &q, Response &r, const kkw_ctx &c)
{
    uint64_t lo = c.row_ptr[v], len = c.row_ptr[v + 1] - lo;
    while (len > 0) {   // lists sorted by neighbour
        const uint64_t half = len >> 1;
        if (c.col[lo + half] < q.cand) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    const bool found = lo < c.row_ptr[v + 1] && c.col[lo] == q.cand;
    r.weight = found ? (c.weight ? c.weight[lo] : 1.0f) : 0.0f;
    r.answered += 1;
}
KKW_FN float dynamic_comp(const WalkerState &, const kkw_walker &, const Response &r, uint32_t, const kkw_edge &,
                          const kkw_ctx &)
{
    return 0.25f + 0.35f * r.weight;
}
KKW_FN float upper_bound(uint32_t, const kkw_ctx &) { return 2.0f; }
