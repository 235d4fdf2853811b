// The tag walk of the reference's outlier test (src/tests/test_outlier.cpp:
// 105-210, first order): edges whose tag equals the walker's get the dynamic
// component times an amplifier, above the upper bound -- they are served by
// the outlier appendix.  Lists sorted by edge tag; tag ranges per vertex in
// array[1] (begin) and array[2] (end): uint64 global positions, [v * TAGS + t].
// array[0]: int32 vertex tag; etype: edge tag.
// param: 0 walk length, 1 outlier amplifier, 2 weighted (1) or not (0).
#define KKW_DYNAMIC 1
#define KKW_LOWER 1
#define KKW_OUTLIER 1
#define TAGS 4

struct WalkerState {
    int32_t tag;
    uint32_t prev;
};

KKW_FN const int32_t *vtag(const kkw_ctx &c) { return (const int32_t *)c.array[0]; }
KKW_FN uint64_t tag_begin(uint32_t v, int32_t t, const kkw_ctx &c) { return ((const uint64_t *)c.array[1])[(uint64_t)v * TAGS + t]; }
KKW_FN uint64_t tag_end(uint32_t v, int32_t t, const kkw_ctx &c) { return ((const uint64_t *)c.array[2])[(uint64_t)v * TAGS + t]; }

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
KKW_FN float dynamic_comp(const WalkerState &s, const kkw_walker &w, uint32_t, const kkw_edge &e, const kkw_ctx &c)
{
    if (w.step == 0) return (float)TAGS;
    const int32_t et = (int32_t)c.etype[e.index];
    float temp = (float)1 + (float)((vtag(c)[s.prev] + s.tag + et) % TAGS);
    if (s.tag == et) temp *= c.param[1];
    return temp;
}
KKW_FN float upper_bound(uint32_t, const kkw_ctx &) { return (float)TAGS; }
KKW_FN float lower_bound(uint32_t, const kkw_ctx &) { return 1.0f; }

KKW_FN void outlier_bound(const WalkerState &s, const kkw_walker &w, uint32_t cur, float &overflow, uint32_t &num,
                          const kkw_ctx &c)
{
    overflow = 0.0f;
    num = 0;
    if (w.step == 0) return;
    const uint64_t b = tag_begin(cur, s.tag, c), e = tag_end(cur, s.tag, c);
    num = (uint32_t)(e - b);
    for (uint64_t p = b; p < e; p++) {
        const float wt = c.param[2] != 0.0f ? c.weight[p] : 1.0f;
        const float x = (float)TAGS * (c.param[1] - 1.0f) * wt;
        overflow = overflow > x ? overflow : x;
    }
}
KKW_FN int64_t outlier_search(const WalkerState &s, const kkw_walker &, uint32_t cur, uint32_t idx, const kkw_ctx &c)
{
    return (int64_t)(tag_begin(cur, s.tag, c) + idx);
}
