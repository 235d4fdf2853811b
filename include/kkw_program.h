/* kkw_program.h -- what a walk program sees.
 *
 * A walk program is the device form of a first-order TransitionConfig plus its
 * WalkerConfig (include/walk.hpp:147-271 of KnightKing): the parts a user
 * writes as closures there -- walker state init/update, extension_comp,
 * dynamic_comp, its upper/lower bounds and the outlier pair -- are written
 * here as functions of fixed names, compiled for the GPU at run time
 * (kkw_program_create) together with the library's walk loop.  The static
 * component is the graph's edge weights (alias tables) or uniform.
 *
 * The same source compiles on the host (KKW_FN becomes static inline), which
 * is how a CPU replay of a program is built.
 *
 * A program defines:
 *
 *   struct WalkerState { ... };          // trivially copyable, per walker
 *   KKW_FN void init_state(WalkerState &s, const kkw_walker &w, uint32_t start,
 *                          const kkw_ctx &c);                // walker_init_state_func
 *   KKW_FN void update_state(WalkerState &s, const kkw_walker &w, uint32_t cur,
 *                            uint64_t edge, const kkw_ctx &c); // walker_update_state_func
 *   KKW_FN float extension(const WalkerState &s, const kkw_walker &w, uint32_t cur,
 *                          const kkw_ctx &c);                // extension_comp
 *
 * and, after `#define KKW_DYNAMIC 1`:
 *
 *   KKW_FN float dynamic_comp(const WalkerState &s, const kkw_walker &w, uint32_t cur,
 *                             uint64_t edge, const kkw_ctx &c);
 *   KKW_FN float upper_bound(uint32_t v, const kkw_ctx &c);  // dcomp_upperbound_func
 *
 * optionally `#define KKW_LOWER 1` with
 *
 *   KKW_FN float lower_bound(uint32_t v, const kkw_ctx &c);  // dcomp_lowerbound_func
 *
 * and optionally `#define KKW_OUTLIER 1` with
 *
 *   KKW_FN void outlier_bound(const WalkerState &s, const kkw_walker &w, uint32_t cur,
 *                             float &overflow, uint32_t &num, const kkw_ctx &c);
 *   KKW_FN int64_t outlier_search(const WalkerState &s, const kkw_walker &w, uint32_t cur,
 *                                 uint32_t idx, const kkw_ctx &c);  // edge or -1
 *
 * `edge` is a global CSR position: the candidate is c.col[edge], its weight
 * c.weight[edge] (when the graph has weights), its type c.etype[edge].  In
 * update_state, w.step is already the new step (walk.hpp:1039-1043); in the
 * other functions it is the step being taken.
 */
#ifndef KKW_PROGRAM_H
#define KKW_PROGRAM_H

#if defined(__HIPCC_RTC__)
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::int16_t int16_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
#define KKW_FN __device__ __forceinline__
#elif defined(__HIPCC__)
#include <stdint.h>
#define KKW_FN __device__ __forceinline__
#else
#include <stdint.h>
#define KKW_FN static inline
#endif

#define KKW_PROGRAM_ARRAYS 8
#define KKW_PROGRAM_PARAMS 8

/* The graph and the caller's data, as the program's functions read them.  On
 * the device every pointer is a device address. */
typedef struct kkw_ctx {
    uint32_t v_num;
    uint32_t reserved;
    uint64_t e_num;
    const uint64_t *row_ptr;   /* [v_num + 1] */
    const uint32_t *col;       /* [e_num] */
    const float *weight;       /* [e_num], or null for an unweighted graph */
    const uint8_t *etype;      /* [e_num], or null */
    const void *array[KKW_PROGRAM_ARRAYS];   /* kkw_program_set_array slots (null if unset) */
    float param[KKW_PROGRAM_PARAMS];         /* kkw_prog_job.param */
} kkw_ctx;

/* Walker<T>::id and ::step (include/walk.hpp:31-50). */
typedef struct kkw_walker {
    uint32_t id;
    uint32_t step;
} kkw_walker;

#endif
