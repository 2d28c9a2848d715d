// ss_precision.hip -- precision form: the 2-player river showdown as a BF16
// hi/lo MFMA contraction with FP32 accumulation.
//
// For traverser t, opponent o and showdown node n (pot P, traverser bet b):
//     v(i, n) = P/2 * sum_j W(i, j) r_o(n, j) + (P/2 - b) * Z(i, n)
// with W(i, j) = sign(rank_i - rank_j) for card-disjoint live hands (else 0)
// and Z the opponent reach disjoint from hand i (the card-removal mass, O(1)
// per hand from per-card totals). This is the reference's 2-player showdown
// (street_solve.cu:886-903: pot * (P_win + P_tie / 2) - bet * Z) rewritten with
// P_win + P_tie / 2 = (Z + W r) / 2.
//
// The contraction over opponent hands runs on v_mfma_f32_16x16x32_bf16 for
// all S showdown nodes at once: Y[Mp x 2S] = W[Mp x Mp] * [R_hi | R_lo], where
// R_hi = bf16(r) and R_lo = bf16(r - R_hi), so the reach carries 16
// significant bits and W (entries -1, 0, +1) is exact. The 2S hi / lo columns
// are packed into ceil(2S / 16) column tiles. W travels at 2 bits per entry,
// already in the order of the MFMA's A operand: one AND (and a shift for odd
// steps) gives a lane's 8 entries as FP4 (E2M1) nibbles, and four
// v_cvt_scalef32_pk_bf16_fp4 turn them into the bf16 fragment. Z, the mass
// floor and the epilogue are FP32 as in the closed-form path.
#include "ss_internal.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <utility>

namespace ss {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;
typedef __attribute__((__vector_size__(2 * sizeof(uint32_t)))) uint32_t u32x2;

#define PREC_MAX_TILES 16   // ceil(2 S / 16) <= 16: at most 128 showdown nodes
#define DN2_P (52 * 52)

struct PrecArgs {
    int S, NC, MpK;           // showdown nodes, column tiles, Bt row length
    const int *sd_nodes;      // [S]
    const uint32_t *W;        // FP4 fragments of W (k_prec_w)
    uint16_t *Bt;             // [MpK / 8][16 NC][8] bf16 bits: column b hi, column S + b lo, others 0
    float *Z;                 // [S][Mp]: card-removal mass per node and traverser hand, -1 where the value is 0
    const int2 *zi;           // [Mp] traverser hand: x = c0 | c1 << 6 | valid << 12, y = same-cards opponent hand or -1
    const floatx4 *C;         // [S] per node: P/2, the traverser's bet, node id, opponent's reach row (bits)
};

// round-to-nearest-even float -> bf16 bits (finite inputs)
__device__ __forceinline__ uint16_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// W(i, j) for opponent o given hand i's card mask and strength: +1 / -1 by
// rank, 0 for dead, blocked or card-sharing pairs and ties
__device__ __forceinline__ int prec_w_entry(const SlotDesc *__restrict__ s, int o, uint64_t mi, uint32_t si, int j) {
    if (j >= s->nh[o] || !si) return 0;
    const uint32_t hj = s->hinfo[(size_t)o * s->Mp + j];
    if (!(hj & SS_LIVE_BIT)) return 0;
    const uint64_t mj = (1ull << (hj & 63)) | (1ull << ((hj >> 6) & 63));
    const uint32_t sj = s->cstr[(hj >> 12) & 2047];
    if ((mi & mj) || !sj) return 0;
    return si > sj ? 1 : (si < sj ? -1 : 0);
}

// W for traverser t as the GEMM's A fragments. MFMA 16x16x32 lane l holds
// A[l % 16][8 (l / 16) + 0..7]: for row tile rt and 32-k step q, lane l's 8
// entries are W(16 rt + l % 16, 32 q + 8 (l / 16) + e), e = 0..7, as FP4
// nibbles (+1 = 0x2, -1 = 0xA, 0 = 0x0; nibble e is element e, which is the
// order v_cvt_scalef32_pk_bf16_fp4 reads them in). Two steps share one 32-bit
// word: step 2m keeps bits 1 and 3 of each nibble, step 2m + 1 bits 0 and 2,
// so `word & 0xAAAAAAAA` and `(word << 1) & 0xAAAAAAAA` are the two FP4 words.
// Layout chunk-major, [nch][Mp / 16][64 lanes][PREC_KS / 2 words] (chunks of
// PREC_KS steps), so the words of one chunk for consecutive row tiles are one
// contiguous range; steps past Mp / 32 are 0, and PREC_W_PAD zero words follow
// the last chunk (a wavefront past the last row tile reads them).
#define PREC_KS 4                       // 32-k steps per staged chunk
#define PREC_W_PAD (4 * 64 * PREC_KS / 2)   // words: 4 row tiles of one chunk
__global__ void __launch_bounds__(SS_BLOCK) k_prec_w(const SlotDesc *__restrict__ s, int t, uint32_t *W) {
    constexpr int WPC = PREC_KS / 2;
    const int Mp = s->Mp, o = 1 - t, nq = Mp / 32, ntile = Mp / 16, nch = (nq + PREC_KS - 1) / PREC_KS;
    const size_t total = (size_t)nch * ntile * 64;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(e % 64), rt = (int)((e / 64) % ntile), c = (int)(e / 64 / ntile);
        const int i = 16 * rt + (l & 15), g = l >> 4;
        uint32_t word[WPC] = {};
        const uint32_t hi = i < s->nh[t] ? s->hinfo[(size_t)t * Mp + i] : 0u;
        if (hi & SS_LIVE_BIT) {
            const uint64_t mi = (1ull << (hi & 63)) | (1ull << ((hi >> 6) & 63));
            const uint32_t si = s->cstr[(hi >> 12) & 2047];
            for (int m = 0; m < WPC; ++m)
                for (int odd = 0; odd < 2; ++odd) {
                    const int q = PREC_KS * c + 2 * m + odd;
                    if (q >= nq) continue;
                    for (int x = 0; x < 8; ++x) {
                        const int v = prec_w_entry(s, o, mi, si, 32 * q + 8 * g + x);
                        if (v) word[m] |= ((v > 0 ? 0x2u : 0xAu) << (4 * x)) >> odd;
                    }
                }
        }
        for (int m = 0; m < WPC; ++m) W[WPC * e + m] = word[m];
    }
}

__device__ __forceinline__ PrecArgs prec_args(const SlotDesc *__restrict__ s, int t) {
    PrecArgs a;
    a.S = s->prec_S;
    a.NC = s->prec_NC;
    a.MpK = s->prec_MpK;
    a.sd_nodes = s->prec_sd;
    a.W = s->prec_W[t];
    a.Bt = s->prec_Bt;
    a.Z = s->prec_Z;
    a.zi = s->prec_zi[t];
    a.C = (const floatx4 *)(s->prec_C + (size_t)t * s->prec_S);
    return a;
}

// A traverser hand's static inputs of the card-removal mass, packed at setup:
// its cards and whether it takes part (live, not blocked), and the opponent
// hand holding the same two cards. zi comes in holding {0, same-cards hand}.
__global__ void __launch_bounds__(SS_BLOCK) k_prec_zinfo(const SlotDesc *__restrict__ s, int t, int2 *zi) {
    const int Mp = s->Mp;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < Mp; i += gridDim.x * blockDim.x) {
        const uint32_t hv = i < s->nh[t] ? s->hinfo[(size_t)t * Mp + i] : 0u;
        const bool valid = (hv & SS_LIVE_BIT) && s->cstr[min((int)((hv >> 12) & 2047), 1325)];
        zi[i].x = valid ? (int)((hv & 63) | ((hv >> 6) & 63) << 6 | 1u << 12) : 0;
    }
}

// Showdown node b of traverser t's pass: the opponent's reach as bf16 hi / lo
// rows of Bt, its per-card totals (dense scatter + wave64 sums, fixed order)
// and from them, per traverser hand i, the opponent reach disjoint from i:
// z = S - C[x] - C[y] + r(same(i)) (S the total, C the per-card totals, the
// hand holding i's cards counted twice), written as -1 where the showdown
// value is 0 (dead or blocked hands, :861, the mass floor, and the padding
// hands past the traverser's count). One workgroup; M and C are LDS. Both
// passes over the hands keep SPLIT_U hands' loads per thread in flight.
#define SPLIT_U 6
__device__ void prec_split_node(const SlotDesc *__restrict__ s, const PrecArgs &a, int t, int b, float *M,
                                float *C) {
    const int o = 1 - t, Mp = s->Mp, nh = s->nh[o], bd = blockDim.x;
    const int rrow = __float_as_int(gld(a.C + b)[3]);   // the opponent's reach row at the node
    const float *r = s->reach + ((size_t)o * s->N + rrow) * Mp;
    const uint32_t *ho = s->hinfo + (size_t)o * Mp;
    for (int i = threadIdx.x; i < DN2_P; i += bd) M[i] = 0.f;
    __syncthreads();
    // one pass over the opponent's hands: hi / lo split and the card-pair scatter
    for (int j0 = threadIdx.x; j0 < Mp; j0 += SPLIT_U * bd) {
        uint32_t hj[SPLIT_U];
        float xr[SPLIT_U];
#pragma unroll
        for (int u = 0; u < SPLIT_U; ++u) {
            const int j = j0 + u * bd;
            hj[u] = j < nh ? gld(ho + j) : 0u;
            xr[u] = j < Mp ? gld(r + j) : 0.f;
        }
#pragma unroll
        for (int u = 0; u < SPLIT_U; ++u) {
            const int j = j0 + u * bd;
            if (j >= Mp) break;
            const bool ok = hj[u] & SS_LIVE_BIT;
            const float x = ok ? xr[u] : 0.f;
            const uint16_t h = bf16_rne(x);
            uint16_t *piece = a.Bt + (size_t)(j >> 3) * 16 * a.NC * 8 + (j & 7);   // k group j / 8, element j % 8
            piece[8 * b] = h;
            piece[8 * (a.S + b)] = bf16_rne(x - bf16_to_f(h));
            if (!ok) continue;
            const int c0 = hj[u] & 63, c1 = (hj[u] >> 6) & 63;
            if (s->uniq) {   // one writer per cell
                M[c0 * 52 + c1] = x;
                M[c1 * 52 + c0] = x;
            } else {
                atomicAdd(&M[c0 * 52 + c1], x);
                atomicAdd(&M[c1 * 52 + c0], x);
            }
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int c = threadIdx.x >> 6; c < 52; c += bd >> 6) {
        float v = lane < 52 ? M[c * 52 + lane] : 0.f;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) C[c] = v;
    }
    __syncthreads();
    if (threadIdx.x < 64) {   // wavefront 0: half the sum of the card totals (each hand holds two cards)
        float tot = lane < 52 ? C[lane] : 0.f;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off, 64);
        if (lane == 0) C[52] = 0.5f * tot;
    }
    __syncthreads();
    const float Sb = C[52];
    float *Zb = a.Z + (size_t)b * Mp;
    for (int i0 = threadIdx.x; i0 < Mp; i0 += SPLIT_U * bd) {   // padding hands have valid = 0: -1
        int2 zi[SPLIT_U];
        float rv[SPLIT_U];
#pragma unroll
        for (int u = 0; u < SPLIT_U; ++u) {
            const int i = i0 + u * bd;
            zi[u] = i < Mp ? gld(a.zi + i) : make_int2(0, -1);
        }
#pragma unroll
        for (int u = 0; u < SPLIT_U; ++u) rv[u] = zi[u].y >= 0 ? gld(r + zi[u].y) : 0.f;
#pragma unroll
        for (int u = 0; u < SPLIT_U; ++u) {
            const int i = i0 + u * bd;
            if (i >= Mp) break;
            const int x = zi[u].x & 63, y = (zi[u].x >> 6) & 63;
            const float z = Sb - C[x] - C[y] + rv[u];
            const bool zero = !(zi[u].x >> 12) || z < 1e-5f * Sb;
            gst(Zb + i, zero ? -1.f : z);
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(SS_BLOCK) k_prec_split(const SlotDesc *__restrict__ s, int t) {
    __shared__ float M[DN2_P];
    __shared__ float C[64];
    prec_split_node(s, prec_args(s, t), t, blockIdx.x, M, C);
}

// phased (batched) forms: tasks are slot << 20 | showdown index / row block
__global__ void __launch_bounds__(SS_BLOCK) k_prec_split_phased(SlotDesc *__restrict__ slots, const uint32_t *tasks,
                                                                int ntasks, int t) {
    __shared__ float M[DN2_P];
    __shared__ float C[64];
    for (int j = blockIdx.x; j < ntasks; j += gridDim.x) {
        const uint32_t tk = tasks[j];
        const SlotDesc *s = slots + (tk >> 20);
        if (!(s->active && s->iter < s->iter_end && t < s->NP)) continue;
        prec_split_node(s, prec_args(s, t), t, (int)(tk & 0xFFFFF), M, C);
    }
}

// Y = W * [R_hi | R_lo] for a block of rows and all column tiles, then the
// showdown values written into cfv. A workgroup owns 64 RT rows in 4 row
// slots of RT row tiles (16 hands each) and has KG wavefronts per slot, one
// per K group: K (opponent hands) is walked in chunks of PREC_KS 32-k steps,
// and K group g takes chunks g, g + KG, g + 2 KG, ... into its own
// accumulators. Chunks travel through a ring of NBUF LDS stages, stage k
// holding chunks KG k .. KG k + KG - 1 (one per K group): the chunk's packed
// reach for all columns and each wavefront's W words for its row tiles, both
// copied by global_load_lds_dwordx4 (global memory straight into LDS, no
// registers) by the K group that uses the chunk, NBUF - 1 stages ahead. Each
// stage costs one barrier, after each wavefront has waited for its own copies
// of that stage (s_waitcnt vmcnt counts them: the loop issues no other vector
// memory instruction). With two K groups each SIMD runs two wavefronts, one
// computing while the other waits. The K groups' partial sums are added in a
// fixed order (group 1's, then group 0's), so every form rounds alike. Bt is
// stored in 16-byte pieces of 8 consecutive k of one column, all columns of
// one 8-k group together, so a chunk is one contiguous range and a lane's B
// fragment one piece, the 16 lanes of a column tile reading 256 contiguous
// bytes. MFMA 16x16x32 lane layout: lane l holds A[l % 16][8 (l / 16) + 0..7],
// B[8 (l / 16) + 0..7][l % 16], D[4 (l / 16) + 0..3][l % 16].
constexpr int prec_kg(int nc) { return nc <= 4 ? 2 : 1; }   // K groups (two stages of reach fit LDS)
template <int NC, int RT_> struct PrecShape {
    static constexpr int RT = RT_;                // row tiles per wavefront
    static constexpr int KG = prec_kg(NC);        // K groups
    static constexpr int WV = 4 * KG;             // wavefronts per workgroup
    static constexpr int NTHR = 64 * WV;
    static constexpr int ROWS = 64 * RT;          // rows per workgroup / task
    static constexpr int KS = PREC_KS;            // 32-k steps per chunk
    static constexpr int WPC = KS / 2;            // W words per lane and row tile of a chunk
    static constexpr int NCOL = 16 * NC;          // columns (hi and lo)
    static constexpr int BP = 4 * KS * NCOL;      // 16-byte reach pieces per chunk
    static constexpr int BPT = BP / 256;          // of them per thread of a K group
    static constexpr int WQ = RT * 64 * WPC / 4;  // 16-byte W pieces per wavefront and chunk
    static constexpr int WDMA = (WQ + 63) / 64;   // W copies per wavefront and chunk
    static constexpr int STAGE = 16 * (KG * BP + WV * WDMA * 64);   // bytes of one stage
    static constexpr int NBUF = KG == 2 ? (4 * STAGE <= 155648 ? 4 : (3 * STAGE <= 155648 ? 3 : 2))
                                        : (4 * STAGE <= 106496 ? 4 : (3 * STAGE <= 122880 ? 3 : 2));
    static constexpr int SMAX = 8 * NC;           // showdown nodes: 2 S <= 16 NC
    static constexpr int RW = 16 * RT;            // rows per row slot
    static constexpr int YSTR = RW + 4;           // epilogue: column of accumulators, 16 lanes' 4-row writes a bank apart
    static constexpr int RQ = RW / 4;             // epilogue: quads of 4 rows per row slot
    static constexpr int QPL = (RQ * SMAX + 64 * KG - 1) / (64 * KG);   // epilogue: (node, quad) outputs per lane at most
    static_assert(BP % 256 == 0, "a chunk's reach splits evenly over a K group");
};
// row tiles per wavefront: the batched form picks one of 1 .. prec_rt_max(NC)
// per launch (the more column tiles, the fewer rows, so the accumulators and
// fragments stay within the 256 registers of two wavefronts per SIMD); the
// plain form uses one row tile, for more workgroups over a single subgame.
// The K order is the same for any row tiling, so both forms round alike.
constexpr int prec_rt_max(int nc) { return nc <= 2 ? 6 : (nc == 3 ? 5 : (nc == 4 ? 3 : (nc <= 8 ? 2 : 1))); }
constexpr int prec_threads(int nc) { return 256 * prec_kg(nc); }

template <int NC, int RT>
struct PrecLds {
    using SH = PrecShape<NC, RT>;
    union {
        struct {
            u32x4 B[SH::NBUF][SH::KG][SH::BP];                  // reach pieces of the stage's chunks
            u32x4 W[SH::NBUF][SH::WV][SH::WDMA * 64];           // each wavefront's W words of its chunk
        } st;
        struct {
            float Ys[SH::KG][4][SH::NCOL][SH::YSTR];            // each K group's and row slot's sums, column-major
            alignas(16) float Cs[SH::SMAX][4];                  // per node: P/2, traverser bet, node id
        } ep;
    } u;
};

// s_waitcnt vmcnt(N): at most N vector memory instructions of this
// wavefront still in flight. The "memory" clobber keeps the compiler from
// moving loads or copies across it, so N counts what the source issues after
// the awaited copies.
template <int N> __device__ __forceinline__ void prec_wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}
// workgroup barrier after this wavefront's LDS accesses, without the memory
// fence of __syncthreads (which would also wait for every global-to-LDS copy
// in flight). The "memory" clobber keeps the compiler from moving LDS
// accesses across it.
__device__ __forceinline__ void prec_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// one wavefront's 1 KB (64 lanes x 16 bytes) global-to-LDS copy; lds is the
// wavefront's destination (uniform), src the lane's 16 bytes
__device__ __forceinline__ void prec_dma(const void *src, void *lds) {
    __builtin_amdgcn_global_load_lds(src, lds, 16, 0, 0);
}

// the bf16 A fragment of one step from its FP4 word
__device__ __forceinline__ bf16x8 prec_expand(uint32_t f) {
    const bf16x2 p0 = __builtin_amdgcn_cvt_scalef32_pk_bf16_fp4(f, 1.0f, 0);
    const bf16x2 p1 = __builtin_amdgcn_cvt_scalef32_pk_bf16_fp4(f, 1.0f, 1);
    const bf16x2 p2 = __builtin_amdgcn_cvt_scalef32_pk_bf16_fp4(f, 1.0f, 2);
    const bf16x2 p3 = __builtin_amdgcn_cvt_scalef32_pk_bf16_fp4(f, 1.0f, 3);
    return __builtin_shufflevector(__builtin_shufflevector(p0, p1, 0, 1, 2, 3),
                                   __builtin_shufflevector(p2, p3, 0, 1, 2, 3), 0, 1, 2, 3, 4, 5, 6, 7);
}

// one chunk's fragments in registers: the lane's reach fragments of every
// step and column tile, and its W words of every row tile
template <int NC, int RT> struct PrecFrag {
    bf16x8 vb[PREC_KS][NC];
    uint32_t wd[RT][PREC_KS / 2];
};
template <int NC, int RT>
__device__ __forceinline__ void prec_load_frag(PrecFrag<NC, RT> &F, const u32x4 *Bc, const uint32_t *Wc, int l) {
    using SH = PrecShape<NC, RT>;
    const int lane_piece = (l >> 4) * SH::NCOL + (l & 15);
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        const u32x2 v = *(const u32x2 *)(Wc + (r * 64 + l) * SH::WPC);
        F.wd[r][0] = v[0], F.wd[r][1] = v[1];
    }
#pragma unroll
    for (int ks = 0; ks < SH::KS; ++ks)
#pragma unroll
        for (int q = 0; q < NC; ++q)
            F.vb[ks][q] = __builtin_bit_cast(bf16x8, Bc[4 * ks * SH::NCOL + 16 * q + lane_piece]);
}
// the steps of one chunk (the first `steps` of them when GUARD: the last
// chunk, whose later steps lie past the subgame's hands)
template <int NC, int RT, bool GUARD>
__device__ __forceinline__ void prec_chunk(const PrecFrag<NC, RT> &F, floatx4 (&acc)[RT][NC], int steps) {
#pragma unroll
    for (int ks = 0; ks < PREC_KS; ++ks) {
        if (GUARD && ks >= steps) break;
        bf16x8 va[RT];
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const uint32_t word = F.wd[r][ks >> 1];
            va[r] = prec_expand(((ks & 1) ? word << 1 : word) & 0xAAAAAAAAu);
        }
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int q = 0; q < NC; ++q)
                acc[r][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va[r], F.vb[ks][q], acc[r][q], 0, 0, 0);
    }
}

template <int NC, int RT>
__device__ void prec_gemm_block(const SlotDesc *__restrict__ s, const PrecArgs &a, int t, int row0,
                                PrecLds<NC, RT> &L) {
    using SH = PrecShape<NC, RT>;
    constexpr int KS = SH::KS, KG = SH::KG, BP = SH::BP, BPT = SH::BPT, WPC = SH::WPC, NBUF = SH::NBUF;
    constexpr int RQ = SH::RQ, QPL = SH::QPL, WQ = SH::WQ, WDMA = SH::WDMA;
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6, g = l >> 4, Mp = s->Mp;
    const int slot = w & 3, kg = w >> 2;    // row slot and K group of this wavefront
    const int nq = Mp / 32, nch = (nq + KS - 1) / KS, nst = (nch + KG - 1) / KG, ntile = Mp / 16;
    const int rt0 = (row0 >> 4) + RT * slot;   // this wavefront's first row tile
    const bool live = rt0 < ntile;             // past the subgame's rows: the wavefront only stages reach
    const u32x4 *Bg = (const u32x4 *)a.Bt;
    const u32x4 *Wg = (const u32x4 *)a.W;
    auto mine = [&](int k) { return KG * k + kg < nch; };   // stage k holds a chunk of this K group
    // stage k % NBUF gets chunk KG k + kg from this wavefront's K group: the
    // reach pieces (each wavefront of the group 1 KB blocks), then a live
    // wavefront's W words for its row tiles (contiguous: RT row tiles of the
    // chunk; a partly covered wavefront reads on into the next chunk or the
    // padding, rows it never writes)
    auto issue = [&](int k) {
        if (!mine(k)) return;
        const int c = KG * k + kg;
        u32x4 *Bs = L.u.st.B[k % NBUF][kg];
#pragma unroll
        for (int u = 0; u < BPT; ++u) prec_dma(Bg + (size_t)c * BP + u * 256 + 64 * slot + l, Bs + u * 256 + 64 * slot);
        if (live) {
            const u32x4 *src = Wg + ((size_t)c * ntile + rt0) * 64 * WPC / 4;
            u32x4 *dst = L.u.st.W[k % NBUF][w];
#pragma unroll
            for (int u = 0; u < WDMA; ++u)
                if (WQ % 64 == 0 || u * 64 + l < WQ) prec_dma(src + u * 64 + l, dst + u * 64);
        }
    };
    // the wait for stage k leaves this wavefront's copies of the stages
    // issued after it in flight (vector memory instructions per stage: D)
    constexpr int D_LIVE = BPT + WDMA, D_DEAD = BPT;
    auto wait_stage = [&](int k, int last) {   // last: the latest stage issued
        int after = 0;
        for (int j = k + 1; j <= last && j < nst; ++j) after += mine(j) ? 1 : 0;
        if (live) {
            if (after >= 3) prec_wait_vm<3 * D_LIVE>();
            else if (after == 2) prec_wait_vm<2 * D_LIVE>();
            else if (after == 1) prec_wait_vm<D_LIVE>();
            else prec_wait_vm<0>();
        } else {
            if (after >= 3) prec_wait_vm<3 * D_DEAD>();
            else if (after == 2) prec_wait_vm<2 * D_DEAD>();
            else if (after == 1) prec_wait_vm<D_DEAD>();
            else prec_wait_vm<0>();
        }
    };
    static_assert(NBUF <= 4 && 3 * D_LIVE < 64, "vmcnt waits are 6 bits");
    // the epilogue's inputs, issued after the last stage's wait, in flight
    // during the last chunk: per node its half pot, the traverser's bet and
    // its id (one load per thread, clamped), and the split kernel's
    // card-removal masses. Outputs go by quads of 4 consecutive rows of one
    // node (16-byte loads and stores); of a row slot's quads, the wavefront of
    // K group kg takes quad KG 64 u + 64 kg + l, node-major.
    floatx4 cnode, zr[QPL];
    auto epilogue_loads = [&]() {
        cnode = gld(a.C + min(tid, a.S - 1));
        if (live) {
#pragma unroll
            for (int u = 0; u < QPL; ++u) {
                const int o = 64 * (KG * u + kg) + l, b = min(o / RQ, a.S - 1),
                          i = min(16 * rt0 + 4 * (o % RQ), Mp - 4);
                zr[u] = gld((const floatx4 *)(a.Z + (size_t)b * Mp + i));
            }
        }
    };
    floatx4 acc[RT][NC];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[r][q] = floatx4{0.f, 0.f, 0.f, 0.f};
    // Every stage before the last holds full chunks for both K groups (the
    // last chunk, possibly short, lies in the last stage), so the loop has one
    // compute body and the last stage its own.
    for (int k = 0; k < NBUF - 1 && k < nst; ++k) issue(k);
    for (int k = 0; k + 1 < nst; ++k) {
        wait_stage(k, k + NBUF - 2);
        prec_barrier();   // stage k in LDS; stage k - 1 read by every wavefront
        if (k + NBUF - 1 < nst) issue(k + NBUF - 1);
        if (live) {
            PrecFrag<NC, RT> F;
            prec_load_frag(F, L.u.st.B[k % NBUF][kg], (const uint32_t *)L.u.st.W[k % NBUF][w], l);
            prec_chunk<NC, RT, false>(F, acc, KS);
        }
    }
    {
        const int k = nst - 1, c = KG * k + kg;
        wait_stage(k, k);
        prec_barrier();
        epilogue_loads();
        if (live && c < nch) {
            PrecFrag<NC, RT> F;
            prec_load_frag(F, L.u.st.B[k % NBUF][kg], (const uint32_t *)L.u.st.W[k % NBUF][w], l);
            prec_chunk<NC, RT, true>(F, acc, nq - c * KS);
        }
    }
    prec_barrier();   // every wavefront is done with the stages
    // epilogue: each K group's sums go to LDS column-major over the stages (a
    // lane's 4 rows in one write), then the lanes walk the (node, quad)
    // outputs node-major, so the cfv stores run along hands. The column sums
    // are group 1's + group 0's (one group: its own), and Y = hi + lo, in that
    // order in every form. Quads of padding hands store 0 (their z is -1);
    // quads past Mp (clamped row tiles) store nothing.
    if (live) {
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int q = 0; q < NC; ++q)
                *(floatx4 *)&L.u.ep.Ys[kg][slot][16 * q + (l & 15)][16 * r + 4 * g] = acc[r][q];
    }
    if (tid < a.S) *(floatx4 *)L.u.ep.Cs[tid] = cnode;
    __syncthreads();
    if (live) {
        // every input of the lane's outputs read first (indices clamped) and
        // held together, so the LDS reads go out back to back with one wait;
        // then the values, and the stores of the outputs that exist
        float *cfv = s->cfv;   // read once: the stores below could otherwise alias the descriptor
        floatx4 yh[QPL], yl[QPL], cb[QPL];
#pragma unroll
        for (int u = 0; u < QPL; ++u) {
            const int o = 64 * (KG * u + kg) + l, b = min(o / RQ, a.S - 1), r4 = 4 * (o % RQ);
            yh[u] = *(const floatx4 *)&L.u.ep.Ys[KG - 1][slot][b][r4];
            yl[u] = *(const floatx4 *)&L.u.ep.Ys[KG - 1][slot][a.S + b][r4];
            if (KG == 2) {
                yh[u] = yh[u] + *(const floatx4 *)&L.u.ep.Ys[0][slot][b][r4];
                yl[u] = yl[u] + *(const floatx4 *)&L.u.ep.Ys[0][slot][a.S + b][r4];
            }
            cb[u] = *(const floatx4 *)L.u.ep.Cs[b];
        }
#pragma unroll
        for (int u = 0; u < QPL; ++u) asm volatile("" ::"v"(yh[u]), "v"(yl[u]), "v"(cb[u]));
#pragma unroll
        for (int u = 0; u < QPL; ++u) {
            const int o = 64 * (KG * u + kg) + l, i = 16 * rt0 + 4 * (o % RQ);
            const float half = cb[u][0], bet = cb[u][1];
            floatx4 val;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float Y = yh[u][e] + yl[u][e], z = zr[u][e];
                val[e] = z < 0.f ? 0.f : half * Y + (half - bet) * z;
            }
            if (o / RQ < a.S && i < Mp) gst((floatx4 *)(cfv + (size_t)__float_as_int(cb[u][2]) * Mp + i), val);
        }
    }
    prec_barrier();   // LDS is reused by the workgroup's next task (the stores need no wait)
}

// plain form: workgroup b covers rows [64 b, 64 (b + 1))
template <int NC>
__global__ void __launch_bounds__(prec_threads(NC)) k_prec_gemm(const SlotDesc *__restrict__ s, int t) {
    __shared__ PrecLds<NC, 1> L;
    prec_gemm_block<NC, 1>(s, prec_args(s, t), t, PrecShape<NC, 1>::ROWS * blockIdx.x, L);
}

template <int NC, int RT>
__global__ void __launch_bounds__(prec_threads(NC)) k_prec_gemm_phased(SlotDesc *__restrict__ slots,
                                                                      const uint32_t *tasks, int ntasks, int t) {
    __shared__ PrecLds<NC, RT> L;
    for (int j = blockIdx.x; j < ntasks; j += gridDim.x) {
        const uint32_t tk = tasks[j];
        const SlotDesc *s = slots + (tk >> 20);
        if (!(s->active && s->iter < s->iter_end && t < s->NP)) continue;
        prec_gemm_block<NC, RT>(s, prec_args(s, t), t, PrecShape<NC, RT>::ROWS * (int)(tk & 0xFFFFF), L);
    }
}

typedef void (*PrecPlain)(const SlotDesc *, int);
typedef void (*PrecPhased)(SlotDesc *, const uint32_t *, int, int);
#define PREC_RT_SLOTS 6   // table columns: row tiles 1 .. 6
template <int NC, int RT> static PrecPhased prec_entry() {
    if constexpr (RT <= prec_rt_max(NC)) return k_prec_gemm_phased<NC, RT>;
    else return nullptr;
}
template <int... I> static PrecPhased prec_phased(int nc, int rt, std::integer_sequence<int, I...>) {
    static const PrecPhased k[] = {prec_entry<I / PREC_RT_SLOTS + 1, I % PREC_RT_SLOTS + 1>()...};
    return k[(nc - 1) * PREC_RT_SLOTS + rt - 1];
}
static PrecPhased prec_phased(int nc, int rt) {
    return prec_phased(nc, rt, std::make_integer_sequence<int, PREC_MAX_TILES * PREC_RT_SLOTS>());
}
template <int... I> static PrecPlain prec_plain(int nc, std::integer_sequence<int, I...>) {
    static const PrecPlain k[] = {k_prec_gemm<I + 1>...};
    return k[nc - 1];
}
static PrecPlain prec_plain(int nc) { return prec_plain(nc, std::make_integer_sequence<int, PREC_MAX_TILES>()); }

int prec_choose_rt(int nc, const std::vector<int> &tiles) {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipDeviceProp_t prop;
        cus = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
                  ? prop.multiProcessorCount
                  : 256;
    }
    const int rmax = prec_rt_max(nc);
    if (const char *e = getenv("SS_PREC_RT")) {   // forced (measurement)
        const int r = atoi(e);
        if (r >= 1 && r <= rmax) return r;
    }
    // A workgroup takes 4 rt row tiles of one subgame (4 row slots, each
    // with one wavefront per K group); as many as registers and LDS allow
    // share a compute unit, whose SIMDs then run one row slot per resident
    // workgroup. Cost: rt x resident workgroups x dispatch rounds (row-tile
    // contractions on the busiest SIMD).
    int best = 1;
    long best_cost = -1;
    for (int rt = 1; rt <= rmax; ++rt) {
        long g = 0;
        for (int n : tiles) g += (n + 4 * rt - 1) / (4 * rt);
        int cap = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&cap, prec_phased(nc, rt), prec_threads(nc), 0) != hipSuccess)
            cap = 1;
        cap = std::max(1, cap);
        const long rounds = (g + (long)cus * cap - 1) / ((long)cus * cap);
        const long cost = (long)rt * cap * rounds;
        if (best_cost < 0 || cost <= best_cost) best = rt, best_cost = cost;
    }
    return best;
}

// ---- host ------------------------------------------------------------------
PrecState::~PrecState() {
    if (block) (void)hipFree(block);
}

int prec_setup(Slot &s, hipStream_t st) {
    const Plan &pl = s.plan;
    if (pl.NP != 2 || !s.desc.is_river)
        return set_error(SS_ERR_UNSUPPORTED, "the BF16 precision form covers 2-player river subgames");
    std::vector<int> sd;
    for (int j = 0; j < pl.T; ++j)
        if (pl.nodes[pl.term[j]].type == SS_NODE_SHOWDOWN) sd.push_back(pl.term[j]);
    if (sd.empty() || sd.size() > 8 * PREC_MAX_TILES)
        return set_error(SS_ERR_UNSUPPORTED, "precision form: %zu showdown nodes (1..%d supported)", sd.size(),
                         8 * PREC_MAX_TILES);
    std::unique_ptr<PrecState> p(new PrecState());
    const int Mp = pl.Mp;
    p->S = (int)sd.size();
    p->NC = (2 * p->S + 15) / 16;
    p->MpK = (Mp + 255) / 256 * 256;
    // every buffer of the form in one allocation, rounded to 2 MiB, so the
    // contraction's streams stay within a few large pages
    const int nch = (Mp / 32 + PREC_KS - 1) / PREC_KS;
    const size_t w_words = (size_t)nch * (Mp / 16) * 64 * (PREC_KS / 2);
    const size_t bt_bytes = sizeof(uint16_t) * 16 * p->NC * (size_t)p->MpK;
    const size_t w_bytes = sizeof(uint32_t) * (w_words + PREC_W_PAD);
    auto up = [](size_t n) { return (n + 255) / 256 * 256; };
    const size_t off_bt = 0, off_w0 = off_bt + up(bt_bytes), off_w1 = off_w0 + up(w_bytes),
                 off_z = off_w1 + up(w_bytes), off_zi = off_z + up(sizeof(float) * (size_t)p->S * Mp),
                 off_sd = off_zi + 2 * up(sizeof(int2) * Mp), off_c = off_sd + up(sizeof(int) * p->S),
                 total = off_c + up(sizeof(float4) * 2 * p->S);
    SS_HIP(hipMalloc(&p->block, (total + (2u << 20) - 1) / (2u << 20) * (2u << 20)));
    p->Bt = (uint16_t *)(p->block + off_bt);
    p->W[0] = (uint32_t *)(p->block + off_w0);
    p->W[1] = (uint32_t *)(p->block + off_w1);
    p->Z = (float *)(p->block + off_z);
    p->zi[0] = (int2 *)(p->block + off_zi);
    p->zi[1] = (int2 *)(p->block + off_zi + up(sizeof(int2) * Mp));
    p->sd_nodes = (int *)(p->block + off_sd);
    p->C = (float4 *)(p->block + off_c);
    std::vector<float4> cn(2 * p->S);
    for (int t = 0; t < 2; ++t)
        for (int b = 0; b < p->S; ++b) {
            const DevNode &nd = pl.nodes[sd[b]];
            int bits = sd[b];
            float id;
            memcpy(&id, &bits, sizeof(id));
            int rbits = pl.rrow[(size_t)(1 - t) * pl.N + sd[b]];
            float rr;
            memcpy(&rr, &rbits, sizeof(rr));
            cn[(size_t)t * p->S + b] = make_float4(0.5f * (float)nd.pot, (float)nd.bets[t], id, rr);
        }
    SS_HIP(hipMemcpyAsync(p->C, cn.data(), sizeof(float4) * cn.size(), hipMemcpyHostToDevice, st));
    SS_HIP(hipMemcpyAsync(p->sd_nodes, sd.data(), sizeof(int) * p->S, hipMemcpyHostToDevice, st));
    SS_HIP(hipMemsetAsync(p->Bt, 0, bt_bytes, st));   // padding columns and k past Mp stay 0
    // opponent hand holding the same two cards as each traverser hand
    std::vector<int2> same(Mp);
    for (int t = 0; t < 2; ++t) {
        const int o = 1 - t;
        std::vector<int> at(SS_NCOMBO, -1);
        for (int j = 0; j < pl.nh[o]; ++j) {
            const uint32_t h = pl.hinfo[(size_t)o * Mp + j];
            if (h & SS_LIVE_BIT) at[(h >> 12) & 2047] = j;
        }
        for (int i = 0; i < Mp; ++i) {
            const uint32_t h = i < pl.nh[t] ? pl.hinfo[(size_t)t * Mp + i] : 0u;
            same[i] = make_int2(0, (h & SS_LIVE_BIT) ? at[(h >> 12) & 2047] : -1);
        }
        SS_HIP(hipMemcpyAsync(p->zi[t], same.data(), sizeof(int2) * Mp, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_prec_zinfo, dim3((Mp + SS_BLOCK - 1) / SS_BLOCK), dim3(SS_BLOCK), 0, st, s.d_desc, t,
                           p->zi[t]);
        SS_HIP(hipStreamSynchronize(st));   // `same` is reused
        SS_HIP(hipMemsetAsync(p->W[t] + w_words, 0, sizeof(uint32_t) * PREC_W_PAD, st));
        hipLaunchKernelGGL(k_prec_w, dim3((unsigned)std::min<size_t>(4096, (w_words / 2 + SS_BLOCK - 1) / SS_BLOCK)),
                           dim3(SS_BLOCK), 0, st, s.d_desc, t, p->W[t]);
    }
    SS_HIP(hipGetLastError());
    SlotDesc &d = s.desc;
    d.prec_S = p->S;
    d.prec_NC = p->NC;
    d.prec_MpK = p->MpK;
    d.prec_sd = p->sd_nodes;
    d.prec_W[0] = p->W[0];
    d.prec_W[1] = p->W[1];
    d.prec_Bt = p->Bt;
    d.prec_Z = p->Z;
    d.prec_zi[0] = p->zi[0];
    d.prec_zi[1] = p->zi[1];
    d.prec_C = p->C;
    SS_HIP(hipMemcpyAsync(s.d_desc, &d, sizeof(d), hipMemcpyHostToDevice, st));
    SS_HIP(hipStreamSynchronize(st));
    s.prec = std::move(p);
    return SS_OK;
}

// the showdown nodes of traverser t's pass (after its forward sweep)
int prec_showdowns(Slot &s, int t, hipStream_t st) {
    const PrecState &p = *s.prec;
    hipLaunchKernelGGL(k_prec_split, dim3(p.S), dim3(SS_BLOCK), 0, st, s.d_desc, t);
    hipLaunchKernelGGL(prec_plain(p.NC), dim3((s.plan.Mp + 63) / 64), dim3(prec_threads(p.NC)), 0, st,
                       s.d_desc, t);
    SS_HIP(hipGetLastError());
    return SS_OK;
}

int launch_prec_phased(SlotDesc *d_slots, const PhasedTasks &pt, int t, hipStream_t st) {
    if (pt.npsplit == 0) return SS_OK;
    hipLaunchKernelGGL(k_prec_split_phased, dim3(pt.npsplit), dim3(SS_BLOCK), 0, st, d_slots,
                       pt.d_tasks + pt.psplit_off, pt.npsplit, t);
    int off = pt.pgemm_off;
    for (const auto &g : pt.pgemm) {
        hipLaunchKernelGGL(prec_phased(g.nc, g.rt), dim3(g.count), dim3(prec_threads(g.nc)), 0, st, d_slots,
                           pt.d_tasks + off, g.count, t);
        off += g.count;
    }
    SS_HIP(hipGetLastError());
    return SS_OK;
}

int64_t prec_flops_per_pass(const Slot &s) {
    if (!s.prec) return 0;
    return 2LL * s.plan.Mp * s.plan.Mp * 16 * s.prec->NC;
}

}  // namespace ss
