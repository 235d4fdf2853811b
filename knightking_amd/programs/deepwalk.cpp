// DeepWalk as a walk program (src/apps/deepwalk.hpp:30-43): truncated static
// walk, no dynamic component.  param[0] = walk length.
// Reads no edge index: draws come from the engine's tables of vertex ids
// and rows (KKW_EDGE_INDEX 0).  A -DKKW_EDGE_INDEX=1 build uses the
// local-index tables instead; paths are the same.
#ifndef KKW_EDGE_INDEX
#define KKW_EDGE_INDEX 0
#endif
struct WalkerState {
    uint32_t unused;
};

KKW_FN void init_state(WalkerState &s, const kkw_walker &, uint32_t, const kkw_ctx &) { s.unused = 0; }
KKW_FN void update_state(WalkerState &, const kkw_walker &, uint32_t, const kkw_edge &, const kkw_ctx &) {}
KKW_FN float extension(const WalkerState &, const kkw_walker &w, uint32_t, const kkw_ctx &c)
{
    return w.step >= (uint32_t)c.param[0] ? 0.0f : 1.0f;
}
