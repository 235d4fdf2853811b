// The tag walk of the reference's bound test (src/tests/test_bound.cpp:107-160,
// first order): walker tag advances by one per step; the dynamic component
// mixes the previous vertex's tag, the walker's and the edge's, plus a
// per-vertex term that also sets the lower and upper bounds.
// array[0]: int32 vertex tag per vertex; etype: edge tag.
// param: 0 walk length, 1 vertex tag weight.
#define KKW_DYNAMIC 1
#define KKW_LOWER 1
#define TAGS 4

struct WalkerState {
    int32_t tag;
    uint32_t prev;
};

KKW_FN const int32_t *vtag(const kkw_ctx &c) { return (const int32_t *)c.array[0]; }

KKW_FN void init_state(WalkerState &s, const kkw_walker &w, uint32_t, const kkw_ctx &)
{
    s.tag = (int32_t)(w.id % TAGS);
    s.prev = 0;
}
KKW_FN void update_state(WalkerState &s, const kkw_walker &, uint32_t cur, const kkw_edge &, const kkw_ctx &)
{
    s.tag = (s.tag + 1) % TAGS;
    s.prev = cur;
}
KKW_FN float extension(const WalkerState &, const kkw_walker &w, uint32_t, const kkw_ctx &c)
{
    return w.step >= (uint32_t)c.param[0] ? 0.0f : 1.0f;
}
KKW_FN float dynamic_comp(const WalkerState &s, const kkw_walker &w, uint32_t cur, const kkw_edge &e, const kkw_ctx &c)
{
    if (w.step == 0) return (float)TAGS;
    float temp = (float)1 + (float)((vtag(c)[s.prev] + s.tag + (int32_t)c.etype[e.index]) % TAGS);
    temp += (float)vtag(c)[cur] * c.param[1];
    return temp;
}
KKW_FN float upper_bound(uint32_t v, const kkw_ctx &c) { return (float)TAGS + (float)vtag(c)[v] * c.param[1]; }
KKW_FN float lower_bound(uint32_t v, const kkw_ctx &c) { return (float)vtag(c)[v] * c.param[1]; }
