// ss_device.h -- data layout shared by the host runtime and the HIP kernels.
//
// One subgame ("slot") lives in one device arena. Its hand axis is padded to
// Mp (a multiple of 64, one wavefront), so every (node x hand) sweep is a
// coalesced stream of 64-float rows.
//
//   nodes     DevNode[N]            tree, struct per node (64 B)
//   children  int[C]                child ids, reference layout
//   dec_lvl   int[D]                decision nodes grouped by BFS depth,
//   dec_off   int[L+1]                level l = dec_lvl[dec_off[l] .. dec_off[l+1])
//   term      int[T]                fold / showdown / leaf nodes
//   hinfo     uint32[NP][Mp]        packed hand: c0 | c1<<6 | combo<<12 | live<<23
//                                   (c0, c1 in the table numbering: cards off
//                                   the board first, board cards last; combo
//                                   is the real combo id)
//   weights   float[NP][Mp]         root reach (raw input weights)
//   R, SS     float[A][Mp]          regrets / strategy sums, one row per action
//                                   of each decision node (rows act_base..+nchild)
//   cfv       float[N][Mp]          counterfactual values of the current pass
//   reach     float[NP][N][Mp]      reach of every player at every node
//   rsig      float[8][Mp]          strategy at node 0 taken before the last pass
//   river tables (showdown closed forms, see ss_kernels.hip):
//   cstr      uint32[1326]          7-card rank of each combo (0 = blocked)
//   gsort     uint16[ng]            live combos by ascending rank (cards a | b << 8)
//   clist     uint8[52][52]         per card: partner cards by ascending rank
//   clen      uint8[52]
//   bnd       uint8[ncls+1][52]     row t, card c: 4 x # entries of c's list ranked
//                                   below class t (row ncls: the list length), so
//                                   for combo k the entries ranked <, <= rank(k)
//                                   are rows kcls[k], kcls[k]+1 (contiguous)
//   cpos      uint8[52][52]         4 x position of partner d in card c's list (255: none)
//   hmeta     uint4[NP][Mp]         each player's hands by ascending strength (dead
//                                   last), one 16-byte record per hand:
//                                   x = h | c0 << 11 | c1 << 17 | ranked << 23,
//                                   y = strength class t (distinct live ranks,
//                                   ascending), z, w = # live combos ranked <, <=.
//                                   Showdown sweeps walk hands in this order so
//                                   a wavefront shares class boundaries
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SS_NCOMBO 1326
#define SS_BLOCK 256          // threads per workgroup of the (node x hand) sweeps
#define SS_BND_BYTES (1082 * 52)   // <= 1081 live combos on a 5-card board, + 1 row
#define SS_LIVE_BIT (1u << 23)

namespace ss {

struct DevNode {
    int type, player, nchild, first_child;
    int pot, leaf_idx, act_base, active_mask;
    int bets[6];
    int depth, pad;
};

// One decision-node task of the phased sweeps, filled on the host whenever
// the task lists are built: every per-task value a sweep workgroup needs
// besides the slot's progress, so that the workgroup starts with one load.
struct alignas(16) SweepRec {
    int slot, node, Mp, N, NP, actor, na, act_base, fa;   // fa: frozen action or -1
    int ch[8];                                            // children, the last repeated past na
    int nh[6];
    int arow;                                             // row of the actor's reach at the node
    float *R, *SS, *cfv, *reach;
};
static_assert(sizeof(SweepRec) == 128, "SweepRec is two 64-byte scalar loads");

struct SlotDesc {
    // shape
    int N, NP, Mp, maxh;
    int A, L, T, D;
    int is_river, has_leaf, num_leaves, root_player;
    int nh[6];
    int ng, odometer;          // odometer: k>=3 terminals present (slow path)
    int uniq;                  // live combos unique per player: scatters are plain stores
    // progress (device-updated; iteration-dependent data never travels as a kernel argument)
    int iter;                  // last completed iteration
    int iter_end;              // the slot stops advancing at this iteration
    int final_iter;            // iteration whose last pass snapshots rsig
    int snap_every;            // also snapshot rsig at iterations divisible by this (0 = off)
    int active;
    // arrays
    const DevNode *nodes;
    const int *children;
    const int *dec_lvl;
    const int *dec_off;
    const int *term;
    const uint32_t *hinfo;
    const float *weights;
    const float *leaf_values;
    const int *frozen;
    const int *rrow;           // [NP][N] row of player p's reach at node n (reach is not copied down the tree)
    float *R;
    float *SS;
    float *cfv;
    float *reach;
    float *rsig;
    const uint32_t *cstr;
    const uint16_t *gsort;
    const uint8_t *clist;
    const uint8_t *clen;
    const uint8_t *bnd;
    const uint8_t *cpos;
    const uint4 *hmeta;        // [NP][Mp] river only
    // precision form (2-player rivers under SS_PREC_BF16X2), zero otherwise
    int prec_S, prec_NC;       // showdown nodes; 16-column tiles of the 2 S hi / lo columns
    int prec_MpK;              // row length of prec_Bt: Mp rounded up to 256
    const int *prec_sd;        // [prec_S] showdown node ids
    const uint32_t *prec_W[2]; // per traverser: W as 2-bit FP4 fragments, lane-ordered
    uint16_t *prec_Bt;         // [prec_MpK / 8][16 prec_NC][8] bf16 reach: column b hi, column S + b lo
    float *prec_Z;             // [prec_S][Mp] card-removal mass z per node and traverser hand, -1 for value 0
    const int2 *prec_zi[2];    // per traverser [Mp] hand: x = c0 | c1 << 6 | valid << 12, y = opponent hand with its cards or -1
    const float4 *prec_C;      // [2][prec_S] per traverser and node: P/2, traverser bet, node id, opponent reach row (bits)
    // exploitability scratch (FP64), allocated on demand
    float *avg;                // [A][Mp]
    double *xreach;            // [NP][N][Mp]
    double *xev;               // [N][Mp]
    double *xbr;               // [N][Mp]
    double *xout;              // [2*NP + 1]: gains, values, joint mass
};

__host__ __device__ inline int combo_id(int a, int b) {
    if (a > b) { int t = a; a = b; b = t; }
    return 51 * a - (a * (a - 1)) / 2 + (b - a - 1);
}

// load through the global address space (the pointers come out of a
// SlotDesc, which the compiler would otherwise treat as generic)
template <typename P>
__device__ __forceinline__ P gld(const P *p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(1))) P *)p;
#else
    return *p;   // host pass of the device code: never executed
#endif
}
// store through the global address space (as gld)
template <typename P>
__device__ __forceinline__ void gst(P *p, P v) {
#if defined(__HIP_DEVICE_COMPILE__)
    *(__attribute__((address_space(1))) P *)p = v;
#else
    *p = v;   // host pass of the device code: never executed
#endif
}
// access at a wave-uniform base plus a 32-bit byte offset: a scalar base per
// row and one shared offset register, instead of a 64-bit address register
// pair per access
template <typename P>
__device__ __forceinline__ P gldo(const P *base, unsigned off) {
#if defined(__HIP_DEVICE_COMPILE__)
    const __attribute__((address_space(1))) char *b = (const __attribute__((address_space(1))) char *)base;
    return *(const __attribute__((address_space(1))) P *)(b + off);
#else
    return *(const P *)((const char *)base + off);   // host pass: never executed
#endif
}
template <typename P>
__device__ __forceinline__ void gsto(P *base, unsigned off, P v) {
#if defined(__HIP_DEVICE_COMPILE__)
    __attribute__((address_space(1))) char *b = (__attribute__((address_space(1))) char *)base;
    *(__attribute__((address_space(1))) P *)(b + off) = v;
#else
    *(P *)((char *)base + off) = v;   // host pass: never executed
#endif
}

}  // namespace ss
