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
// all S showdown nodes at once: Y[Mp x 2 S16] = W[Mp x Mp] * [R_hi | R_lo],
// where R_hi = bf16(r) and R_lo = bf16(r - R_hi), so the reach carries 16
// significant bits and W (entries -1, 0, +1) is exact. W is stored as two sign
// masks per row and 32-column chunk (2 bits per entry, 8x fewer bytes than
// bf16) and expanded to bf16 in registers through a 256-entry LDS table. Z,
// the mass floor and the epilogue are FP32 as in the closed-form path.
#include "ss_internal.h"

namespace ss {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;
typedef __attribute__((__vector_size__(2 * sizeof(uint32_t)))) uint32_t u32x2;

#define PREC_MAX_TILES 8   // S16 / 16 <= 8: at most 128 showdown nodes
#define DN2_P (52 * 52)
#define PREC_WAVES 4       // wavefronts per GEMM workgroup

struct PrecArgs {
    int S, S16, NT;           // showdown nodes, S rounded up to 16, S16 / 16
    const int *sd_nodes;      // [S]
    const uint2 *W;           // [Mp][Mp / 32] sign masks {+1 bits, -1 bits}, rows = traverser hands
    uint16_t *Bt;             // [2 S16][Mp] bf16 bits: rows b < S16 hi, S16 + b lo
    float *Cn;                // [S][64]: per-card opponent totals, [52] = total
    const int *same;          // [Mp] opponent hand with the traverser hand's cards, or -1
};

// round-to-nearest-even float -> bf16 bits (finite inputs)
__device__ __forceinline__ uint16_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// W for traverser t: +1 / -1 / 0 by rank, 0 for blocked or card-sharing
// pairs, as masks: W[i][q] = {bits j with W(i, 32 q + j) = +1, ... = -1}
__global__ void __launch_bounds__(SS_BLOCK) k_prec_w(const SlotDesc *__restrict__ s, int t, uint2 *W) {
    const int Mp = s->Mp, o = 1 - t, nq = Mp / 32;
    const size_t total = (size_t)Mp * nq;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / nq), q = (int)(e % nq);
        uint32_t pos = 0u, neg = 0u;
        const uint32_t hi = i < s->nh[t] ? s->hinfo[(size_t)t * Mp + i] : 0u;
        if (hi & SS_LIVE_BIT) {
            const uint64_t mi = (1ull << (hi & 63)) | (1ull << ((hi >> 6) & 63));
            const uint32_t si = s->cstr[(hi >> 12) & 2047];
            for (int jb = 0; jb < 32; ++jb) {
                const int j = 32 * q + jb;
                if (j >= s->nh[o] || !si) continue;
                const uint32_t hj = s->hinfo[(size_t)o * Mp + j];
                if (!(hj & SS_LIVE_BIT)) continue;
                const uint64_t mj = (1ull << (hj & 63)) | (1ull << ((hj >> 6) & 63));
                const uint32_t sj = s->cstr[(hj >> 12) & 2047];
                if ((mi & mj) || !sj) continue;
                if (si > sj) pos |= 1u << jb;
                else if (si < sj) neg |= 1u << jb;
            }
        }
        W[e] = make_uint2(pos, neg);
    }
}

__device__ __forceinline__ PrecArgs prec_args(const SlotDesc *__restrict__ s, int t) {
    PrecArgs a;
    a.S = s->prec_S;
    a.S16 = s->prec_S16;
    a.NT = s->prec_S16 / 16;
    a.sd_nodes = s->prec_sd;
    a.W = s->prec_W[t];
    a.Bt = s->prec_Bt;
    a.Cn = s->prec_Cn;
    a.same = s->prec_same[t];
    return a;
}

// Showdown node b of traverser t's pass: the opponent's reach as bf16 hi / lo
// rows of Bt, and its per-card totals (dense scatter + wave64 sums, fixed
// order). One workgroup; M and C are LDS.
__device__ void prec_split_node(const SlotDesc *__restrict__ s, const PrecArgs &a, int t, int b, float *M,
                                float *C) {
    const int n = a.sd_nodes[b], o = 1 - t, Mp = s->Mp, nh = s->nh[o];
    const float *r = s->reach + ((size_t)o * s->N + s->rrow[(size_t)o * s->N + n]) * Mp;
    for (int i = threadIdx.x; i < DN2_P; i += blockDim.x) M[i] = 0.f;
    __syncthreads();
    // one pass over the opponent's hands: hi / lo split and the card-pair scatter
    for (int j = threadIdx.x; j < Mp; j += blockDim.x) {
        const uint32_t hj = j < nh ? gld(s->hinfo + (size_t)o * Mp + j) : 0u;
        const bool ok = hj & SS_LIVE_BIT;
        const float x = ok ? gld(r + j) : 0.f;
        const uint16_t h = bf16_rne(x);
        a.Bt[(size_t)b * Mp + j] = h;
        a.Bt[(size_t)(a.S16 + b) * Mp + j] = bf16_rne(x - bf16_to_f(h));
        if (!ok) continue;
        const int c0 = hj & 63, c1 = (hj >> 6) & 63;
        if (s->uniq) {   // one writer per cell
            M[c0 * 52 + c1] = x;
            M[c1 * 52 + c0] = x;
        } else {
            atomicAdd(&M[c0 * 52 + c1], x);
            atomicAdd(&M[c1 * 52 + c0], x);
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int c = threadIdx.x >> 6; c < 52; c += blockDim.x >> 6) {
        float v = lane < 52 ? M[c * 52 + lane] : 0.f;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) C[c] = v;
    }
    __syncthreads();
    if (threadIdx.x < 64) {   // wavefront 0: the card totals and their half-sum
        const float v = lane < 52 ? C[lane] : 0.f;
        if (lane < 52) a.Cn[(size_t)b * 64 + lane] = v;
        float tot = v;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off, 64);
        if (lane == 0) a.Cn[(size_t)b * 64 + 52] = 0.5f * tot;
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
// showdown values written into cfv. A workgroup of 4 wavefronts owns 64 RT
// rows (each wavefront RT row tiles of 16 hands); K (opponent hands) is walked
// in chunks: the chunk of packed reach (all columns) and of W's sign masks
// (all rows) is staged in LDS with coalesced loads, then every wavefront reads
// its fragments from LDS, so each reach fragment fetched from L2 feeds 4 RT
// MFMAs. K runs in order within a wavefront, so every form rounds alike.
// MFMA 16x16x32 lane layout: lane l holds A[l % 16][8 (l / 16) + 0..7],
// B[8 (l / 16) + 0..7][l % 16], D[4 (l / 16) + 0..3][l % 16].
template <int NT, int RT_> struct PrecShape {
    static constexpr int RT = RT_;                                // row tiles per wavefront
    static constexpr int ROWS = 64 * RT;                          // rows per workgroup / task
    static constexpr int KCH = (512 / NT) / 64 * 64;              // K per staged chunk (multiple of 64)
    static constexpr int BPAD = KCH + 8;                          // LDS row of one column, padded a bank
};
// row tiles per wavefront in the batched form: the more column tiles, the
// fewer rows, so the accumulators stay within the register file; the plain
// form uses one row tile for more workgroups over a single subgame. The K
// order is the same for any row tiling, so both forms round alike.
template <int NT> constexpr int prec_rt() { return NT <= 2 ? 4 : (NT <= 4 ? 2 : 1); }
// lut_one[m]: bf16 1.0 in each of 8 slots whose bit is set in m; lut_sign[m]:
// the sign bit in those slots
__device__ __forceinline__ void prec_fill_luts(uint4 *lut_one, uint4 *lut_sign) {
    for (int m = threadIdx.x; m < 256; m += blockDim.x) {
        uint32_t w1[4], ws[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t b0 = (m >> (2 * k)) & 1u, b1 = (m >> (2 * k + 1)) & 1u;
            w1[k] = b0 * 0x3F80u | (b1 * 0x3F80u) << 16;
            ws[k] = b0 * 0x8000u | (b1 * 0x8000u) << 16;
        }
        lut_one[m] = make_uint4(w1[0], w1[1], w1[2], w1[3]);
        lut_sign[m] = make_uint4(ws[0], ws[1], ws[2], ws[3]);
    }
}

template <int NT, int RT>
struct PrecLds {
    using SH = PrecShape<NT, RT>;
    uint4 lut_one[256], lut_sign[256];
    alignas(16) uint16_t Bs[32 * NT][SH::BPAD];              // reach chunk: [column][k]
    uint2 Ws[SH::ROWS][SH::KCH / 32 + 1];                    // W mask chunk: [row][32-k step], padded a bank
};

template <int NT, int RT>
__device__ void prec_gemm_block(const SlotDesc *__restrict__ s, const PrecArgs &a, int t, int row0,
                                PrecLds<NT, RT> &L) {
    using SH = PrecShape<NT, RT>;
    constexpr int KCH = SH::KCH, NCOL = 32 * NT;
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6, Mp = s->Mp, o = 1 - t;
    const int nq = Mp / 32, sh = 8 * (l >> 4), wrow0 = 16 * RT * w;   // this wavefront's first row in the block
    floatx4 acc[RT][2 * NT];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int q = 0; q < 2 * NT; ++q) acc[r][q] = floatx4{0.f, 0.f, 0.f, 0.f};
    // staging: the chunk of packed reach for every column (hi columns are Bt
    // rows 0 .. 16 NT - 1, lo columns Bt rows S16 ..) and the W masks of the
    // block's rows, coalesced along k. Each thread holds a fixed number of
    // loads in registers; the next chunk's loads are issued before the
    // current chunk's MFMAs so that their latency is hidden. Addresses are
    // clamped into the arrays instead of branched on: a clamped element lands
    // in an LDS column past the chunk's end (never read) or in a row past Mp
    // (never written back).
    constexpr int NTHR = 64 * PREC_WAVES;
    constexpr int BE = NCOL * KCH / 8, BU = BE / NTHR;                 // uint4 of reach per chunk
    constexpr int WE = SH::ROWS * KCH / 32, WU = (WE + NTHR - 1) / NTHR;   // uint2 of masks per chunk
    static_assert(BE % NTHR == 0, "reach chunk must split evenly over the workgroup");
    u32x4 rb[BU];   // native vectors: kept in registers across the MFMA loop
    u32x2 rw[WU];
    // pass k0 = -KCH only fetches chunk 0; pass k0 stores chunk k0 into LDS,
    // fetches chunk k0 + KCH, then runs chunk k0's MFMAs
    for (int k0 = -KCH; k0 < Mp; k0 += KCH) {
        if (k0 >= 0) {
#pragma unroll
            for (int u = 0; u < BU; ++u) {
                const int e = u * NTHR + tid;
                *(u32x4 *)&L.Bs[e / (KCH / 8)][8 * (e % (KCH / 8))] = rb[u];
            }
#pragma unroll
            for (int u = 0; u < WU; ++u) {
                const int e = u * NTHR + tid;
                if (WE % NTHR == 0 || e < WE) *(u32x2 *)&L.Ws[e / (KCH / 32)][e % (KCH / 32)] = rw[u];
            }
            __syncthreads();
        }
        const int kf = k0 + KCH;
        if (kf < Mp) {
            const int k8 = min(KCH, Mp - kf) / 8, k32f = k8 / 4;
#pragma unroll
            for (int u = 0; u < BU; ++u) {
                const int e = u * NTHR + tid, c = e / (KCH / 8), j = min(e % (KCH / 8), k8 - 1);
                const int brow = c < 16 * NT ? c : a.S16 + (c - 16 * NT);
                rb[u] = gld((const u32x4 *)(a.Bt + (size_t)brow * Mp + kf + 8 * j));
            }
#pragma unroll
            for (int u = 0; u < WU; ++u) {
                const int e = min(u * NTHR + tid, WE - 1), r = e / (KCH / 32), j = min(e % (KCH / 32), k32f - 1);
                rw[u] = gld((const u32x2 *)&a.W[(size_t)min(row0 + r, Mp - 1) * nq + (kf >> 5) + j]);
            }
        }
        if (k0 < 0) continue;
        const int k32 = min(KCH, Mp - k0) / 32;   // the chunk's 32-k steps
        for (int ks = 0; ks < k32; ++ks) {
            bf16x8 va[RT], vb[2 * NT];
#pragma unroll
            for (int q = 0; q < 2 * NT; ++q)
                vb[q] = __builtin_bit_cast(bf16x8, *(const uint4 *)&L.Bs[16 * q + (l & 15)][32 * ks + sh]);
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                // lane's 8 entries: bits sh .. sh + 7 of the step's masks
                const uint2 m = L.Ws[wrow0 + 16 * r + (l & 15)][ks];
                const uint32_t p8 = (m.x >> sh) & 0xFFu, n8 = (m.y >> sh) & 0xFFu;
                const uint4 mag = L.lut_one[p8 | n8], sgn = L.lut_sign[n8];
                va[r] = __builtin_bit_cast(bf16x8,
                                           make_uint4(mag.x | sgn.x, mag.y | sgn.y, mag.z | sgn.z, mag.w | sgn.w));
            }
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int q = 0; q < 2 * NT; ++q)
                    acc[r][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va[r], vb[q], acc[r][q], 0, 0, 0);
        }
        __syncthreads();
    }
    // epilogue. Per-column constants go to LDS, over the reach buffer (free
    // after the last chunk's barrier): Cn's 53 values, then half the pot, the
    // traverser's bet, the node and the row of the opponent's reach at the
    // node (reach-row table). Per-row inputs are loaded as one batch,
    // then per column the opponent reach of the hand holding the same cards,
    // so a thread's loads overlap instead of waiting one after another.
    constexpr int CSTRIDE = 65;   // padded a bank: 16 columns read the same card
    float *Cs = reinterpret_cast<float *>(&L.Bs[0][0]);
    static_assert(sizeof(L.Bs) >= sizeof(float) * 16 * NT * CSTRIDE, "column constants must fit the reach buffer");
    for (int e = tid; e < 16 * NT * 64; e += NTHR) {
        const int b = e >> 6, c = e & 63;
        float v = 0.f;   // columns past S: zeros, node 0
        if (b < a.S && c <= 52) v = gld(a.Cn + (size_t)b * 64 + c);
        else if (b < a.S && c <= 56) {
            const int n = gld(a.sd_nodes + b);
            const DevNode &nd = s->nodes[n];
            v = c == 53   ? 0.5f * (float)nd.pot
                : c == 54 ? (float)nd.bets[t]
                : c == 55 ? __int_as_float(n)
                          : __int_as_float(gld(s->rrow + (size_t)o * s->N + n));
        }
        Cs[b * CSTRIDE + c] = v;
    }
    __syncthreads();
    const int nh = s->nh[t], rbase = row0 + wrow0 + 4 * (l >> 4);
    const float *ro = s->reach + (size_t)o * s->N * Mp;
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        uint32_t hv[4], cs[4];
        int sv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int ic = min(rbase + 16 * r + e, Mp - 1);
            hv[e] = gld(s->hinfo + (size_t)t * Mp + ic);
            sv[e] = gld(a.same + ic);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) cs[e] = gld(s->cstr + min((int)((hv[e] >> 12) & 2047), 1325));
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const int b = 16 * q + (l & 15);
            const float *Cb = Cs + b * CSTRIDE;
            const float Sb = Cb[52], half = Cb[53], bet = Cb[54];
            const int n = __float_as_int(Cb[55]), nrow = __float_as_int(Cb[56]);
            float rv[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) rv[e] = gld(ro + (size_t)nrow * Mp + max(sv[e], 0));
            // the values are needed here: keeps the compiler from sinking each
            // load into the guarded store below, one wait per load
#pragma unroll
            for (int e = 0; e < 4; ++e) asm volatile("" ::"v"(rv[e]));
            if (b >= a.S) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = rbase + 16 * r + e;
                if (i >= nh) continue;
                // computed for every lane and selected (no branch, so no
                // load is sunk into one); dead or blocked hands give 0, :861
                const int x = hv[e] & 63, y = (hv[e] >> 6) & 63, sm = sv[e];
                const float z = Sb - Cb[x] - Cb[y] + (sm >= 0 ? rv[e] : 0.f);
                const float Y = acc[r][q][e] + acc[r][NT + q][e];
                const bool zero = !((hv[e] & SS_LIVE_BIT) && cs[e]) || z < 1e-5f * Sb;
                s->cfv[(size_t)n * Mp + i] = zero ? 0.f : half * Y + (half - bet) * z;
            }
        }
    }
}

// plain form: workgroup b covers rows [ROWS b, ROWS (b + 1))
template <int NT>
__global__ void __launch_bounds__(64 * PREC_WAVES) k_prec_gemm(const SlotDesc *__restrict__ s, int t) {
    __shared__ PrecLds<NT, 1> L;
    prec_fill_luts(L.lut_one, L.lut_sign);
    __syncthreads();
    prec_gemm_block<NT, 1>(s, prec_args(s, t), t, PrecShape<NT, 1>::ROWS * blockIdx.x, L);
}

template <int NT>
__global__ void __launch_bounds__(64 * PREC_WAVES) k_prec_gemm_phased(SlotDesc *__restrict__ slots,
                                                                      const uint32_t *tasks, int ntasks, int t) {
    constexpr int RT = prec_rt<NT>();
    __shared__ PrecLds<NT, RT> L;
    prec_fill_luts(L.lut_one, L.lut_sign);
    __syncthreads();
    for (int j = blockIdx.x; j < ntasks; j += gridDim.x) {
        const uint32_t tk = tasks[j];
        const SlotDesc *s = slots + (tk >> 20);
        if (!(s->active && s->iter < s->iter_end && t < s->NP)) continue;
        prec_gemm_block<NT, RT>(s, prec_args(s, t), t, PrecShape<NT, RT>::ROWS * (int)(tk & 0xFFFFF), L);
    }
}

int prec_rows_per_task(int nt) {
    switch (nt) {
    case 1: return 64 * prec_rt<1>();
    case 2: return 64 * prec_rt<2>();
    case 3: return 64 * prec_rt<3>();
    case 4: return 64 * prec_rt<4>();
    case 5: return 64 * prec_rt<5>();
    case 6: return 64 * prec_rt<6>();
    case 7: return 64 * prec_rt<7>();
    default: return 64 * prec_rt<8>();
    }
}

typedef void (*PrecGemmKernel)(const SlotDesc *, int);
static PrecGemmKernel prec_gemm_kernel(int nt) {
    switch (nt) {
    case 1: return k_prec_gemm<1>;
    case 2: return k_prec_gemm<2>;
    case 3: return k_prec_gemm<3>;
    case 4: return k_prec_gemm<4>;
    case 5: return k_prec_gemm<5>;
    case 6: return k_prec_gemm<6>;
    case 7: return k_prec_gemm<7>;
    default: return k_prec_gemm<8>;
    }
}

typedef void (*PrecGemmPhased)(SlotDesc *, const uint32_t *, int, int);
static PrecGemmPhased prec_gemm_phased_kernel(int nt) {
    switch (nt) {
    case 1: return k_prec_gemm_phased<1>;
    case 2: return k_prec_gemm_phased<2>;
    case 3: return k_prec_gemm_phased<3>;
    case 4: return k_prec_gemm_phased<4>;
    case 5: return k_prec_gemm_phased<5>;
    case 6: return k_prec_gemm_phased<6>;
    case 7: return k_prec_gemm_phased<7>;
    default: return k_prec_gemm_phased<8>;
    }
}

// ---- host ------------------------------------------------------------------
PrecState::~PrecState() {
    for (void *p : {(void *)W[0], (void *)W[1], (void *)Bt, (void *)Cn, (void *)same[0], (void *)same[1],
                    (void *)sd_nodes})
        if (p) (void)hipFree(p);
}

int prec_setup(Slot &s, hipStream_t st) {
    const Plan &pl = s.plan;
    if (pl.NP != 2 || !s.desc.is_river)
        return set_error(SS_ERR_UNSUPPORTED, "the BF16 precision form covers 2-player river subgames");
    std::vector<int> sd;
    for (int j = 0; j < pl.T; ++j)
        if (pl.nodes[pl.term[j]].type == SS_NODE_SHOWDOWN) sd.push_back(pl.term[j]);
    if (sd.empty() || sd.size() > 16 * PREC_MAX_TILES)
        return set_error(SS_ERR_UNSUPPORTED, "precision form: %zu showdown nodes (1..%d supported)", sd.size(),
                         16 * PREC_MAX_TILES);
    std::unique_ptr<PrecState> p(new PrecState());
    const int Mp = pl.Mp;
    p->S = (int)sd.size();
    p->S16 = (p->S + 15) / 16 * 16;
    SS_HIP(hipMalloc(&p->sd_nodes, sizeof(int) * p->S));
    SS_HIP(hipMemcpyAsync(p->sd_nodes, sd.data(), sizeof(int) * p->S, hipMemcpyHostToDevice, st));
    SS_HIP(hipMalloc(&p->Bt, sizeof(uint16_t) * 2 * p->S16 * Mp));
    SS_HIP(hipMemsetAsync(p->Bt, 0, sizeof(uint16_t) * 2 * p->S16 * Mp, st));   // padding columns stay 0
    SS_HIP(hipMalloc(&p->Cn, sizeof(float) * 64 * p->S));
    // opponent hand holding the same two cards as each traverser hand
    std::vector<int> same(Mp);
    for (int t = 0; t < 2; ++t) {
        const int o = 1 - t;
        std::vector<int> at(SS_NCOMBO, -1);
        for (int j = 0; j < pl.nh[o]; ++j) {
            const uint32_t h = pl.hinfo[(size_t)o * Mp + j];
            if (h & SS_LIVE_BIT) at[(h >> 12) & 2047] = j;
        }
        for (int i = 0; i < Mp; ++i) {
            const uint32_t h = i < pl.nh[t] ? pl.hinfo[(size_t)t * Mp + i] : 0u;
            same[i] = (h & SS_LIVE_BIT) ? at[(h >> 12) & 2047] : -1;
        }
        SS_HIP(hipMalloc(&p->same[t], sizeof(int) * Mp));
        SS_HIP(hipMemcpyAsync(p->same[t], same.data(), sizeof(int) * Mp, hipMemcpyHostToDevice, st));
        SS_HIP(hipStreamSynchronize(st));   // `same` is reused
        SS_HIP(hipMalloc(&p->W[t], sizeof(uint2) * (size_t)Mp * (Mp / 32)));
        hipLaunchKernelGGL(k_prec_w, dim3(std::min<int64_t>(4096, ((int64_t)Mp * (Mp / 32) + SS_BLOCK - 1) / SS_BLOCK)),
                           dim3(SS_BLOCK), 0, st, s.d_desc, t, p->W[t]);
    }
    SS_HIP(hipGetLastError());
    SlotDesc &d = s.desc;
    d.prec_S = p->S;
    d.prec_S16 = p->S16;
    d.prec_sd = p->sd_nodes;
    d.prec_W[0] = p->W[0];
    d.prec_W[1] = p->W[1];
    d.prec_Bt = p->Bt;
    d.prec_Cn = p->Cn;
    d.prec_same[0] = p->same[0];
    d.prec_same[1] = p->same[1];
    SS_HIP(hipMemcpyAsync(s.d_desc, &d, sizeof(d), hipMemcpyHostToDevice, st));
    SS_HIP(hipStreamSynchronize(st));
    s.prec = std::move(p);
    return SS_OK;
}

// the showdown nodes of traverser t's pass (after its forward sweep)
int prec_showdowns(Slot &s, int t, hipStream_t st) {
    const PrecState &p = *s.prec;
    hipLaunchKernelGGL(k_prec_split, dim3(p.S), dim3(SS_BLOCK), 0, st, s.d_desc, t);
    hipLaunchKernelGGL(prec_gemm_kernel(p.S16 / 16), dim3((s.plan.Mp + 63) / 64), dim3(64 * PREC_WAVES), 0,
                       st, s.d_desc, t);
    SS_HIP(hipGetLastError());
    return SS_OK;
}

int launch_prec_phased(SlotDesc *d_slots, const PhasedTasks &pt, int t, hipStream_t st) {
    if (pt.npsplit == 0) return SS_OK;
    hipLaunchKernelGGL(k_prec_split_phased, dim3(pt.npsplit), dim3(SS_BLOCK), 0, st, d_slots,
                       pt.d_tasks + pt.psplit_off, pt.npsplit, t);
    int off = pt.pgemm_off;
    for (const auto &g : pt.pgemm) {
        hipLaunchKernelGGL(prec_gemm_phased_kernel(g.first), dim3(g.second), dim3(64 * PREC_WAVES), 0, st, d_slots,
                           pt.d_tasks + off, g.second, t);
        off += g.second;
    }
    SS_HIP(hipGetLastError());
    return SS_OK;
}

int64_t prec_flops_per_pass(const Slot &s) {
    if (!s.prec) return 0;
    return 2LL * s.plan.Mp * s.plan.Mp * 2 * s.prec->S16;
}

}  // namespace ss
