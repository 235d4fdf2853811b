// The dynamic walk of the reference's example (src/examples/dynamic_walk.cpp:
// 53-99): returning to the previous vertex is twice as likely; upper bound 2.
// param[0] = walk length.
#define KKW_DYNAMIC 1

struct WalkerState {
    uint32_t last;
};

KKW_FN void init_state(WalkerState &s, const kkw_walker &, uint32_t, const kkw_ctx &) { s.last = 0xFFFFFFFFu; }
KKW_FN void update_state(WalkerState &s, const kkw_walker &, uint32_t cur, const kkw_edge &, const kkw_ctx &) { s.last = cur; }
KKW_FN float extension(const WalkerState &, const kkw_walker &w, uint32_t, const kkw_ctx &c)
{
    return w.step >= (uint32_t)c.param[0] ? 0.0f : 1.0f;
}
KKW_FN float dynamic_comp(const WalkerState &s, const kkw_walker &w, uint32_t, const kkw_edge &e, const kkw_ctx &)
{
    if (w.step == 0) return 1.0f;
    return e.neighbour == s.last ? 2.0f : 1.0f;
}
KKW_FN float upper_bound(uint32_t, const kkw_ctx &) { return 2.0f; }
