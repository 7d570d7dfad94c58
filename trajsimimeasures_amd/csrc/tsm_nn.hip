// tsm_nn.hip -- kernels of the learned kNN filter: a trajectory encoder
// (resampling, Conv1d(2->64, k=3) + ReLU, GRU(64->128), linear head plus a
// fixed linear skip of the resampled points) with forward and backward
// passes, its Adam update, and the embedding search (squared L2 distances
// through an MFMA GEMM, then per-row top-k selection).
//
// GEMMs are BF16 x BF16 -> FP32 on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16): C[M,N] = alpha * A[M,K] . B[N,K]^T + beta * C
// (+ bias[N]), both operands row-major with K contiguous, so every lane's
// fragment (8 consecutive K values of one row) is one 16-byte load.  A wave
// owns a 32x32 output tile (2x2 MFMA tiles), a 256-thread block 64x64.
//
// Gate order and equations of the GRU follow the usual convention:
//   r = sig(Gx_r + Gh_r), z = sig(Gx_z + Gh_z), n = tanh(Gx_n + r * Gh_n),
//   h' = (1 - z) * n + z * h,  Gx = x W_ih^T + b_ih,  Gh = h W_hh^T + b_hh.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "../../include/tsm.h"
#include "../../include/tsm_nn.h"
#include "tsm_host.hpp"

namespace tsm {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float bf2f(uint16_t v) { return __builtin_bit_cast(float, (uint32_t)v << 16); }
// round-to-nearest-even FP32 -> BF16 (inputs are finite)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// ---------------------------------------------------------------- GEMM

struct GemmArgs {
    int M, N, K;
    const uint16_t *A;  // [M, lda] bf16, K contiguous
    int lda;
    const uint16_t *B;  // [N, ldb] bf16, K contiguous
    int ldb;
    float *C;           // [M, ldc] fp32
    int ldc;
    float alpha, beta;
    const float *bias;  // [N] or null
};

__device__ __forceinline__ bf16x8 load_frag(const uint16_t *base, int ld, int row, int rows, int k, int K) {
    // 8 consecutive K values of `row`; zero outside the matrix (K % 8 == 0)
    if (row < rows && k < K) {
        const uint4 v = *reinterpret_cast<const uint4 *>(base + (int64_t)row * ld + k);
        return __builtin_bit_cast(bf16x8, v);
    }
    return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
}

__global__ __launch_bounds__(256) void gemm_nt_bf16_kernel(GemmArgs g) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int r16 = lane & 15, kb = (lane >> 4) * 8;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        const bf16x8 a0 = load_frag(g.A, g.lda, m0 + r16, g.M, k0 + kb, g.K);
        const bf16x8 a1 = load_frag(g.A, g.lda, m0 + 16 + r16, g.M, k0 + kb, g.K);
        const bf16x8 b0 = load_frag(g.B, g.ldb, n0 + r16, g.N, k0 + kb, g.K);
        const bf16x8 b1 = load_frag(g.B, g.ldb, n0 + 16 + r16, g.N, k0 + kb, g.K);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
    }
    // D layout: lane holds column (lane & 15), rows 4 * (lane >> 4) + i
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + 16 * j + r16;
            if (col >= g.N) continue;
            const float bb = g.bias ? g.bias[col] : 0.f;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = m0 + 16 * i + 4 * (lane >> 4) + e;
                if (row >= g.M) continue;
                float *c = g.C + (int64_t)row * g.ldc + col;
                const float v = g.alpha * acc[i][j][e] + bb;
                *c = g.beta == 0.f ? v : v + g.beta * *c;
            }
        }
}

// LDS-tiled variant for large products: a 256-thread block computes a
// 128 x 128 tile of C, each wave a 64 x 64 quarter (4 x 4 MFMA tiles), over
// K in steps of 64 staged through double-buffered LDS tiles (rows padded to
// 72 BF16 so fragment reads spread over banks).  Every A and B fragment
// read from LDS feeds four MFMAs.  With kFilter the product is not stored:
// each value below its row's threshold tau[row] is appended to that row's
// candidate list (cnt[row] counts, entries past `cap` are dropped and
// counted), which is how the embedding search keeps the C nearest database
// rows without writing the full distance matrix.
// LDS row stride of a 64-wide K step: 80 BF16 (40 dwords).  Row strides of
// 8 or 24 dwords mod 32 measured fastest for these MFMA operand reads (16
// rows x 16 bytes per read); 36 dwords (a 72-element row) ran the search
// GEMM 9 % slower.
constexpr int kBM = 128, kBN = 128, kBK = 64, kLdsStride = kBK + 16;

struct FilterArgs {
    const float *tau;  // [M] per-row threshold
    int cap;           // candidate slots per row
    uint32_t *cnt;     // [M] candidates found (may exceed cap)
    float *key;        // [M, cap] candidate values
    int32_t *idx;      // [M, cap] candidate columns
};

// Global -> registers -> LDS staging of a ROWS x kBK BF16 tile by NT
// threads, 16 bytes (8 values of one row) per chunk.
// Branch-free: every lane loads from an index clamped into the operand and
// zeroes what lies outside it, so the loads stay straight-line code and the
// compiler can wait for exactly the ones it needs (a load under a branch
// makes it wait for every load in flight).
template <int ROWS, int NT>
__device__ __forceinline__ void tile_load(const uint16_t *base, int ld, int row0, int rows, int k0, int K,
                                          uint4 (&r)[ROWS * 8 / NT]) {
#pragma unroll
    for (int i = 0; i < ROWS * 8 / NT; ++i) {
        const int q = (int)threadIdx.x + NT * i;
        const int row = row0 + (q >> 3), k = k0 + (q & 7) * 8;
        const bool ok = row < rows && k < K;
        const uint4 v = *reinterpret_cast<const uint4 *>(base + (int64_t)(row < rows ? row : rows - 1) * ld +
                                                         (k < K ? k : 0));
        r[i] = ok ? v : make_uint4(0, 0, 0, 0);
    }
}
// Unconditional form for operands whose K range is whole steps: rows past
// `rows` read row rows - 1 (their products land in columns nobody keeps).
// Native vectors (not the uint4 struct), so the values stay in registers.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// KW: the K width of the tile (KW / 8 chunks of 16 bytes per row).
template <int ROWS, int NT, int KW = kBK>
__device__ __forceinline__ void tile_load_rows(const uint16_t *base, int ld, int row0, int rows, int k0,
                                               u32x4 (&r)[ROWS * (KW / 8) / NT]) {
    constexpr int CH = KW / 8;
#pragma unroll
    for (int i = 0; i < ROWS * CH / NT; ++i) {
        const int q = (int)threadIdx.x + NT * i;
        const int row = min(row0 + q / CH, rows - 1), k = k0 + (q % CH) * 8;
        r[i] = *reinterpret_cast<const u32x4 *>(base + (int64_t)row * ld + k);
    }
}
template <int ROWS, int NT, int ST = kLdsStride, int KW = kBK>
__device__ __forceinline__ void tile_store_v(uint16_t *s, const u32x4 (&r)[ROWS * (KW / 8) / NT]) {
    constexpr int CH = KW / 8;
#pragma unroll
    for (int i = 0; i < ROWS * CH / NT; ++i) {
        const int q = (int)threadIdx.x + NT * i;
        *reinterpret_cast<u32x4 *>(s + (q / CH) * ST + (q % CH) * 8) = r[i];
    }
}
template <int ROWS, int NT> __device__ __forceinline__ void tile_store(uint16_t *s, const uint4 (&r)[ROWS * 8 / NT]) {
#pragma unroll
    for (int i = 0; i < ROWS * 8 / NT; ++i) {
        const int q = (int)threadIdx.x + NT * i;
        *reinterpret_cast<uint4 *>(s + (q >> 3) * kLdsStride + (q & 7) * 8) = r[i];
    }
}

// kMode: 0 store (alpha A B^T + beta C + bias), 1 filter (see above), 2
// split-K accumulate: block z covers K range [z * kchunk, (z + 1) * kchunk)
// and adds alpha * its partial product into C with float atomics.
// BM (128 or 256) rows of C per block, 2 * BM threads: (BM / 64) x 2 waves
// of 64 x 64.  The 256-row form halves how often the B operand (the
// database, in the search) is streamed.
template <int kMode, int BM = kBM>
__global__ __launch_bounds__(2 * BM) void gemm_nt_lds_kernel(GemmArgs g, FilterArgs f, int kchunk) {
    constexpr bool kFilter = kMode == 1;
    constexpr int NT = 2 * BM;
    __shared__ __attribute__((aligned(16))) uint16_t As[2][BM * kLdsStride];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][kBN * kLdsStride];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int bm = blockIdx.y * BM, bn = blockIdx.x * kBN;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int r16 = lane & 15, kb = (lane >> 4) * 8;
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    uint4 ra[BM * 8 / NT], rb[kBN * 8 / NT];
    // K range of this block (the whole K unless split)
    const int kbeg = kMode == 2 ? (int)blockIdx.z * kchunk : 0;
    const int kend = kMode == 2 ? min(g.K, kbeg + kchunk) : g.K;
    tile_load<BM, NT>(g.A, g.lda, bm, g.M, kbeg, kend, ra);
    tile_load<kBN, NT>(g.B, g.ldb, bn, g.N, kbeg, kend, rb);
    tile_store<BM, NT>(As[0], ra);
    tile_store<kBN, NT>(Bs[0], rb);
    __syncthreads();
    const int nk = (kend - kbeg + kBK - 1) / kBK;
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            tile_load<BM, NT>(g.A, g.lda, bm, g.M, kbeg + (kt + 1) * kBK, kend, ra);
            tile_load<kBN, NT>(g.B, g.ldb, bn, g.N, kbeg + (kt + 1) * kBK, kend, rb);
        }
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 32) {
            bf16x8 af[4], bfr[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                af[i] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4 *>(&As[cur][(wm + 16 * i + r16) * kLdsStride + kk + kb]));
#pragma unroll
            for (int j = 0; j < 4; ++j)
                bfr[j] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4 *>(&Bs[cur][(wn + 16 * j + r16) * kLdsStride + kk + kb]));
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            tile_store<BM, NT>(As[cur ^ 1], ra);
            tile_store<kBN, NT>(Bs[cur ^ 1], rb);
        }
        __syncthreads();
    }
    // D layout: lane holds column (lane & 15) of each 16 x 16 tile, rows 4 * (lane >> 4) + e
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = bn + wn + 16 * j + r16;
        if (col >= g.N) continue;
        const float bb = g.bias ? g.bias[col] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = bm + wm + 16 * i + 4 * (lane >> 4) + e;
                if (row >= g.M) continue;
                const float v = g.alpha * acc[i][j][e] + bb;
                if constexpr (kFilter) {
                    if (v < f.tau[row]) {
                        const uint32_t pos = atomicAdd(&f.cnt[row], 1u);
                        if (pos < (uint32_t)f.cap) {
                            f.key[(int64_t)row * f.cap + pos] = v;
                            f.idx[(int64_t)row * f.cap + pos] = col;
                        }
                    }
                } else if constexpr (kMode == 2) {
                    atomicAdd(g.C + (int64_t)row * g.ldc + col, g.alpha * acc[i][j][e]);
                } else {
                    float *c = g.C + (int64_t)row * g.ldc + col;
                    *c = g.beta == 0.f ? v : v + g.beta * *c;
                }
            }
    }
}

// Persistent search GEMM (the embedding search: M queries x N database rows,
// K = 32 * KC, 256 or 384).  A 1024-thread block owns 128 query rows, whose whole A
// operand it stages in LDS once, and streams 128-column database tiles
// through double-buffered LDS in K steps of 64, tile after tile, with no
// per-tile prologue.  The database loads run a full tile ahead: each step
// stores the ring slot of the step after it (loaded one tile earlier) and
// reloads that slot one tile further on, so each load has a whole tile of
// steps to arrive (a load issued one step ahead would stall every step on
// HBM / L2 latency at this occupancy).  The filter epilogue runs after each
// tile's last step and queues its hits in LDS.  The grid has one block per CU-sized share: block b runs on XCD
// b % 8 (round-robin dispatch), and the blocks of one XCD take consecutive
// database chunks with all query blocks of a chunk on that XCD, so a database
// tile comes from HBM once per XCD and from its L2 for the other query
// blocks.  Accumulation order per output equals gemm_nt_lds_kernel's (K
// chunks of 32 in order), so values are bit-identical to the stored GEMM's.
constexpr int kSearchRows = 128;
// LDS row strides of the search kernel (BF16 elements): A rows 32 KC + 48
// (216 dwords at K = 384, 152 at K = 256: 24 mod 32 both), database rows
// KW + 16 for a K step of KW (40 or 72 dwords: 8 mod 32).  Measured on the
// 1,000 x 1M search: strides of 8 or 24 dwords mod 32 ran it 9 % faster than
// 4 or 16 (at 4, SQ_LDS_BANK_CONFLICT was 43 % of the LDS-active cycles).
constexpr int kSearchAPad = 48;

// The one-tile-ahead register ring of the search kernel: six named tiles
// picked at compile time (an array here ends up in scratch memory).  1024
// threads: sixteen waves of 32 x 32 outputs (two column tiles each), 128
// VGPRs, four waves per SIMD to cover each step's barrier and LDS waits.
// Measured at K = 256 on the 1,000 x 1M search against eight waves per
// block at two per SIMD: 32 x 64 outputs per wave 3 - 12 % slower; 64 x 32
// 4 - 13 % slower; 32 x 64 with the query fragments held in registers (half
// the LDS reads) 3 % faster to 7 % slower from box to box.  LDS was busy
// 49 % of the cycles and the matrix cores 38 %: each step's barrier, not
// LDS bandwidth, sets the pace.
constexpr int kSearchThreads = 1024, kSearchColTiles = 2;
// K width of a search step at K = 256 (64 or 128): a 128-wide step halves the
// block barriers per database tile, its database buffers taking the LDS of
// half the hit queue (163,556 of 163,840 bytes, 126 VGPRs).  Measured on the
// 1,000 x 1M search (random embeddings, 731 hits per row, identical hit
// sets): 719 / 715 us mean against 735 / 731 with 64-wide steps.
#ifndef TSM_SEARCH_KW256
#define TSM_SEARCH_KW256 128
#endif
constexpr int kSearchKW256 = TSM_SEARCH_KW256;
template <int RV> struct SearchRingT {
    u32x4 r0[RV], r1[RV], r2[RV], r3[RV], r4[RV], r5[RV];
    template <int S> __device__ __forceinline__ u32x4 (&at())[RV] {
        if constexpr (S == 0) return r0;
        else if constexpr (S == 1) return r1;
        else if constexpr (S == 2) return r2;
        else if constexpr (S == 3) return r3;
        else if constexpr (S == 4) return r4;
        else return r5;
    }
};

// One K step s of tile t (NS steps of KW per tile, KW = 64 or 128).  The
// fragments of step s come from Bs[cur] (stored and made visible in the step
// before), 64 of K at a time; after the first 64's fragment reads, the ring
// slot of the next step goes to Bs[cur ^ 1] (last read in the step before,
// finished at its barrier) and is reloaded from tile tn (tn2 for the next
// tile's step 0); then the products; then the step's one barrier.  No wave
// waits on its own LDS store before reading.  A wave owns 32 rows x 16 JT
// columns (2 x JT MFMA tiles); K chunks of 32 are multiplied in order.
template <int s, int NS, int NT, int JT, int KW>
__device__ __forceinline__ void search_step(const GemmArgs &g, SearchRingT<kBN * (KW / 8) / NT> &ring,
                                           const uint16_t *As, int AS, uint16_t (*Bs)[kBN * (KW + 16)], int &cur,
                                           int tn, int tn2, int wm, int wn, int r16, int kb, f32x4 (&acc)[2][JT]) {
    constexpr int BST = KW + 16;  // database row stride in LDS: 40 or 72 dwords, 8 mod 32
    constexpr int sn = (s + 1) % NS;
#pragma unroll
    for (int h = 0; h < KW / 64; ++h) {
        bf16x8 af[2][2], bfr[2][JT];
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
                af[kk][i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(
                                                           &As[(wm + 16 * i + r16) * AS + KW * s + 64 * h + 32 * kk + kb]));
#pragma unroll
            for (int j = 0; j < JT; ++j)
                bfr[kk][j] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4 *>(&Bs[cur][(wn + 16 * j + r16) * BST + 64 * h + 32 * kk + kb]));
        }
        if (h == 0) {
            tile_store_v<kBN, NT, BST, KW>(Bs[cur ^ 1], ring.template at<sn>());
            tile_load_rows<kBN, NT, KW>(g.B, g.ldb, (s + 1 < NS ? tn : tn2) * kBN, g.N, sn * KW, ring.template at<sn>());
        }
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < JT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[kk][i], bfr[kk][j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
    cur ^= 1;
}

// Candidates found by the search kernel go to a block queue in LDS (row,
// column, value; slots taken with LDS atomics) and reach the caller's lists
// in flushes: when the queue is half full at a tile boundary, and at the
// end.  A global atomic returns through the vector-memory counter, whose
// waits are in order: one in the tile epilogue would also wait for the
// next tile's database loads in flight.  Hits past a full queue take the
// global path at once.
template <int QN> struct SearchQueue {
    int row[QN], col[QN];
    float val[QN];
    unsigned n;
};
template <int QN>
__device__ __forceinline__ void search_flush(const SearchQueue<QN> &q, unsigned n, const FilterArgs &f, int NT) {
    if (n > (unsigned)QN) n = QN;
    for (int e = threadIdx.x; e < (int)n; e += NT) {
        const int row = q.row[e];
        const uint32_t pos = atomicAdd(&f.cnt[row], 1u);
        if (pos < (uint32_t)f.cap) {
            f.key[(int64_t)row * f.cap + pos] = q.val[e];
            f.idx[(int64_t)row * f.cap + pos] = q.col[e];
        }
    }
}

template <int KC, int NT, int JT, int KW>
__global__ __launch_bounds__(NT, NT / 256) void search_filter_kernel(GemmArgs g, FilterArgs f, int qblocks,
                                                                          int chunks, int tiles_per_chunk) {
    static_assert(KC == 8 || KC == 12, "K = 256 or 384");
    static_assert(KW == 64 || (KW == 128 && KC == 8), "K steps of 64, or of 128 at K = 256");
    constexpr int NS = 32 * KC / KW;  // K steps per tile: the register ring holds two to six
    constexpr int AS = 32 * KC + kSearchAPad;  // A row stride in LDS (BF16), padded
    // queue slots within the 160 KB of LDS (A 76 KB at K = 256, B 40 or 72 KB)
    constexpr int QN = KC == 12 ? 512 : KW == 64 ? 2048 : 1000;
    __shared__ __attribute__((aligned(16))) uint16_t As[kSearchRows * AS];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][kBN * (KW + 16)];
    __shared__ SearchQueue<QN> hq;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kb = (lane >> 4) * 8;
    const int xcd = blockIdx.x & 7, jx = blockIdx.x >> 3;
    const int q = jx % qblocks, c = xcd * (chunks / 8) + jx / qblocks;
    const int ntiles = (g.N + kBN - 1) / kBN;
    const int t0 = c * tiles_per_chunk, t1 = min(ntiles, t0 + tiles_per_chunk);
    if (t0 >= t1) return;  // block-uniform
    constexpr int CG = 128 / (16 * JT);  // column groups of waves
    static_assert((NT / 64) * 32 * 16 * JT == kSearchRows * kBN, "waves tile the block");
    const int bm = q * kSearchRows, wm = (wave / CG) * 32, wn = (wave % CG) * (16 * JT);
    if (threadIdx.x == 0) hq.n = 0;  // visible after the first barrier below
    // the block's A rows, all of K, into LDS (16-byte chunks; zero past M / K)
    for (int e = threadIdx.x; e < kSearchRows * 4 * KC; e += NT) {
        const int row = e / (4 * KC), k = (e % (4 * KC)) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (bm + row < g.M && k < g.K) v = *reinterpret_cast<const uint4 *>(g.A + (int64_t)(bm + row) * g.lda + k);
        *reinterpret_cast<uint4 *>(&As[row * AS + k]) = v;
    }
    // the thresholds of this lane's 8 output rows, once per block (-inf past M: nothing passes)
    float taur[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int row = bm + wm + 16 * i + 4 * (lane >> 4) + e;
            const float tv = f.tau[row < g.M ? row : g.M - 1];
            taur[i][e] = row < g.M ? tv : -__builtin_huge_valf();
        }
    SearchRingT<kBN * (KW / 8) / NT> ring;  // step s of the next tile
    tile_load_rows<kBN, NT, KW>(g.B, g.ldb, t0 * kBN, g.N, 0 * KW, ring.template at<0>());
    tile_load_rows<kBN, NT, KW>(g.B, g.ldb, t0 * kBN, g.N, 1 * KW, ring.template at<1>());
    if constexpr (NS > 2) {
        tile_load_rows<kBN, NT, KW>(g.B, g.ldb, t0 * kBN, g.N, 2 * KW, ring.template at<2>());
        tile_load_rows<kBN, NT, KW>(g.B, g.ldb, t0 * kBN, g.N, 3 * KW, ring.template at<3>());
    }
    if constexpr (NS > 4) {
        tile_load_rows<kBN, NT, KW>(g.B, g.ldb, t0 * kBN, g.N, 4 * KW, ring.template at<4>());
        tile_load_rows<kBN, NT, KW>(g.B, g.ldb, t0 * kBN, g.N, 5 * KW, ring.template at<5>());
    }
    // step 0 of tile t0 into Bs[0], its ring slot reloaded with tile t0 + 1
    int cur = 0;
    tile_store_v<kBN, NT, KW + 16, KW>(Bs[0], ring.template at<0>());
    tile_load_rows<kBN, NT, KW>(g.B, g.ldb, (t0 + 1 < t1 ? t0 + 1 : t0) * kBN, g.N, 0, ring.template at<0>());
    __syncthreads();
    for (int t = t0; t < t1; ++t) {
        f32x4 acc[2][JT];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < JT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        // biases of the lane's columns, issued before the tile's database loads
        // (vector-memory waits are in order: loaded in the epilogue, the wait
        // for them would also wait for the next tile's loads in flight)
        float bcol[JT];
#pragma unroll
        for (int j = 0; j < JT; ++j) {
            const int col = t * kBN + wn + 16 * j + r16;
            bcol[j] = g.bias ? g.bias[col < g.N ? col : g.N - 1] : 0.f;
        }
        // the last tile reloads itself (unused)
        const int tn = t + 1 < t1 ? t + 1 : t, tn2 = t + 2 < t1 ? t + 2 : t1 - 1;
        search_step<0, NS, NT, JT, KW>(g, ring, As, AS, Bs, cur, tn, tn2, wm, wn, r16, kb, acc);
        // the queue count as every wave sees it: no wave adds to the queue
        // between step 0's barrier and the last step's
        const unsigned qn = hq.n;
        search_step<1, NS, NT, JT, KW>(g, ring, As, AS, Bs, cur, tn, tn2, wm, wn, r16, kb, acc);
        if constexpr (NS > 2) {
            search_step<2, NS, NT, JT, KW>(g, ring, As, AS, Bs, cur, tn, tn2, wm, wn, r16, kb, acc);
            search_step<3, NS, NT, JT, KW>(g, ring, As, AS, Bs, cur, tn, tn2, wm, wn, r16, kb, acc);
        }
        if constexpr (NS > 4) {
            search_step<4, NS, NT, JT, KW>(g, ring, As, AS, Bs, cur, tn, tn2, wm, wn, r16, kb, acc);
            search_step<5, NS, NT, JT, KW>(g, ring, As, AS, Bs, cur, tn, tn2, wm, wn, r16, kb, acc);
        }
        if (qn >= (unsigned)QN / 2) {  // block-uniform
            search_flush(hq, qn, f, NT);
            __syncthreads();
            if (threadIdx.x == 0) hq.n = 0;
            __syncthreads();
        }
        // filter epilogue of tile t (hits are rare: about 4 C per query over
        // the whole database)
#pragma unroll
        for (int j = 0; j < JT; ++j) {
            const int col = t * kBN + wn + 16 * j + r16;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float v = g.alpha * acc[i][j][e] + bcol[j];
                    if (col < g.N && v < taur[i][e]) {
                        const int row = bm + wm + 16 * i + 4 * (lane >> 4) + e;
                        const unsigned slot = atomicAdd(&hq.n, 1u);
                        if (slot < (unsigned)QN) {
                            hq.row[slot] = row;
                            hq.col[slot] = col;
                            hq.val[slot] = v;
                        } else {
                            const uint32_t pos = atomicAdd(&f.cnt[row], 1u);
                            if (pos < (uint32_t)f.cap) {
                                f.key[(int64_t)row * f.cap + pos] = v;
                                f.idx[(int64_t)row * f.cap + pos] = col;
                            }
                        }
                    }
                }
        }
    }
    __syncthreads();
    search_flush(hq, hq.n, f, NT);
}

// The C smallest candidates of each row, ascending: keys (order-preserving
// bits of the value, then the column; unique within a row, ties broken by
// position all the same).  A row whose list is short (fewer than C) or
// overflowed (more than cap) is flagged for the caller's exact fallback and
// gets column 0 in every slot (a valid index, so work queued on it before
// the fallback reads no garbage).  One 256-thread block per row; cap <= 4096.
// With C <= 256, four 8-bit radix passes over the value bits find the
// value V of the C-th key (three barriers a pass), four more over the
// column bits of the keys of value V when the C-th key splits them, the C
// keys at or below it are gathered, and they are ordered by counting ranks.
// Sorting every key with a bitonic network instead (kept for C > 256) takes
// a barrier per stage, 45 at 512 keys and 66 at 2,048, and the launch lasts
// as long as its longest row: 72 µs against 17-18 µs per 1,000 rows of
// about 400 keys.

// The bin d where the cumulative count of the 256 bins reaches rank k
// (1-based): pick = {d, k minus the counts below d, d's count}.  Wave 0
// scans 4 bins per lane.
__device__ __forceinline__ void radix_pick(const uint32_t *hist, uint32_t k, uint32_t *pick) {
    const int lane = threadIdx.x;
    uint32_t c[4], sum = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        c[i] = hist[4 * lane + i];
        sum += c[i];
    }
    uint32_t incl = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    uint32_t before = incl - sum;
    if (before < k && k <= incl) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (before < k && k <= before + c[i]) {
                pick[0] = (uint32_t)(4 * lane + i);
                pick[1] = k - before;
                pick[2] = c[i];
            }
            before += c[i];
        }
    }
}

// Optional pair output of the selection: slot s = row * C + r of the
// candidate list also gives the pair (row_ids[row], col_ids[column]) that
// an exact re-rank of the list computes (-1 for a column past ncol_ids,
// which the DP planner turns into a NaN result).  pa == nullptr: none.
struct SelectPairs {
    const int32_t *row_ids, *col_ids;
    int64_t ncol_ids;
    int32_t *pa, *pb;
};

__device__ __forceinline__ void select_emit(const SelectPairs &sp, int32_t *out, int64_t row, int64_t s,
                                            int32_t col) {
    out[s] = col;
    if (sp.pa) {
        sp.pa[s] = sp.row_ids[row];
        sp.pb[s] = (col >= 0 && col < sp.ncol_ids) ? sp.col_ids[col] : -1;
    }
}

__global__ __launch_bounds__(256) void select_candidates_kernel(const uint32_t *cnt, const float *key,
                                                                const int32_t *idx, int cap, int C, int32_t *out,
                                                                int32_t *flag, SelectPairs sp) {
    __shared__ uint64_t sk[4096];
    __shared__ uint64_t sel[256];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t pick[3];
    __shared__ uint32_t nsel;
    const int64_t row = blockIdx.x;
    const uint32_t n = cnt[row];
    if (n < (uint32_t)C || n > (uint32_t)cap) {
        if (threadIdx.x == 0) flag[row] = 1;
        for (int t = threadIdx.x; t < C; t += blockDim.x) select_emit(sp, out, row, row * C + t, 0);
        return;
    }
    if (threadIdx.x == 0) flag[row] = 0;
    if (C <= 256) {  // block-uniform
        for (int t = threadIdx.x; t < (int)n; t += blockDim.x) {
            const uint32_t u = __builtin_bit_cast(uint32_t, key[row * cap + t]);
            const uint32_t k = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
            sk[t] = ((uint64_t)k << 32) | (uint32_t)idx[row * cap + t];
        }
        // radix passes over the value bits (hi), then, when the C-th key's
        // value is shared by more keys than the rank needs, over the column
        // bits (lo) of those keys
        uint32_t vhi = 0, vlo = 0, kk = (uint32_t)C, group = 0;
        bool split = false;  // the tie group is split by column
        for (int word = 1; word >= 0; --word) {
            if (word == 0) {
                if (group == kk) break;  // block-uniform: the whole tie group is needed
                split = true;
            }
            uint32_t prefix = 0, pmask = 0;
            for (int shift = 24; shift >= 0; shift -= 8) {
                hist[threadIdx.x] = 0u;
                __syncthreads();  // keys stored (first pass), bins cleared, last pick read
                for (int t = threadIdx.x; t < (int)n; t += blockDim.x) {
                    const uint64_t me = sk[t];
                    const uint32_t h = (uint32_t)(me >> 32);
                    if (word == 1) {
                        if ((h & pmask) == prefix) atomicAdd(&hist[(h >> shift) & 255u], 1u);
                    } else {
                        const uint32_t l = (uint32_t)me;
                        if (h == vhi && (l & pmask) == prefix) atomicAdd(&hist[(l >> shift) & 255u], 1u);
                    }
                }
                __syncthreads();
                if (threadIdx.x < 64) radix_pick(hist, kk, pick);
                __syncthreads();
                prefix |= pick[0] << shift;
                pmask |= 255u << shift;
                kk = pick[1];
                group = pick[2];
            }
            if (word == 1) vhi = prefix;
            else vlo = prefix;
        }
        // the C keys: value below vhi, or value vhi and (the whole group was
        // needed, or column at most vlo)
        if (threadIdx.x == 0) nsel = 0u;
        __syncthreads();
        for (int t = threadIdx.x; t < (int)n; t += blockDim.x) {
            const uint64_t me = sk[t];
            const uint32_t h = (uint32_t)(me >> 32);
            if (h < vhi || (h == vhi && (!split || (uint32_t)me <= vlo))) {
                const uint32_t slot = atomicAdd(&nsel, 1u);  // < C with unique keys; bounded all the same
                if (slot < (uint32_t)C) sel[slot] = me;
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < C; t += blockDim.x) {
            const uint64_t me = sel[t];
            int r = 0;
            for (int j = 0; j < C; ++j) r += (sel[j] < me) | ((sel[j] == me) & (j < t));
            select_emit(sp, out, row, row * C + r, (int32_t)(uint32_t)me);
        }
        return;
    }
    int P = 1;
    while (P < (int)n) P <<= 1;
    for (int t = threadIdx.x; t < P; t += blockDim.x) {
        if (t < (int)n) {
            const uint32_t u = __builtin_bit_cast(uint32_t, key[row * cap + t]);
            const uint32_t k = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
            sk[t] = ((uint64_t)k << 32) | (uint32_t)idx[row * cap + t];
        } else {
            sk[t] = ~0ull;
        }
    }
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < P; t += blockDim.x) {
                const int o = t ^ stride;
                if (o > t) {
                    const bool up = (t & size) == 0;
                    const uint64_t a = sk[t], b = sk[o];
                    if ((a > b) == up) {
                        sk[t] = b;
                        sk[o] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < C; t += blockDim.x) select_emit(sp, out, row, row * C + t, (int32_t)(uint32_t)sk[t]);
}

// FP32 [rows, cols] (ld_src) -> BF16, optionally transposed: dst[c][r].
__global__ void cast_bf16_kernel(const float *src, int rows, int cols, int ld_src, uint16_t *dst, int ld_dst,
                                 int transpose) {
    const int64_t total = (int64_t)rows * cols;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / cols), c = (int)(i % cols);
        const uint16_t v = f2bf(src[(int64_t)r * ld_src + c]);
        if (transpose) dst[(int64_t)c * ld_dst + r] = v;
        else dst[(int64_t)r * ld_dst + c] = v;
    }
}


// ---------------------------------------------------------------- encoder kernels

// Resample every trajectory to T points by linear interpolation along the
// point index, centre to [-1, 1]: X[n][t] = 2 * (p(t / (T-1) * (len-1)) - 0.5).
// lenf[n] = log(len) - 4.
__global__ void resample_kernel(const double *xy, const int64_t *offs, const int32_t *ids, int n, int T, float *X,
                                float *lenf) {
    const int64_t total = (int64_t)n * T;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / T), t = (int)(i % T);
        const int64_t tr = ids ? ids[r] : r;
        const int64_t o = offs[tr];
        const int64_t len = offs[tr + 1] - o;
        const double s = T > 1 ? (double)t / (double)(T - 1) * (double)(len - 1) : 0.0;
        int64_t lo = (int64_t)floor(s);
        if (lo > len - 1) lo = len - 1;
        const int64_t hi = lo + 1 < len ? lo + 1 : len - 1;
        const double f = s - (double)lo;
        const double x = xy[2 * (o + lo)] * (1.0 - f) + xy[2 * (o + hi)] * f;
        const double y = xy[2 * (o + lo) + 1] * (1.0 - f) + xy[2 * (o + hi) + 1] * f;
        X[2 * i] = (float)((x - 0.5) * 2.0);
        X[2 * i + 1] = (float)((y - 0.5) * 2.0);
        if (t == 0) lenf[r] = (float)(log((double)len) - 4.0);
    }
}

// Conv1d(2 -> C, kernel 3, zero padding 1) + ReLU over [n, T] sequences.
// Wc[c][i][k], bc[c]; H (fp32) and Hb (bf16) are [n*T, C].
__global__ void conv_fwd_kernel(const float *X, int n, int T, int C, const float *Wc, const float *bc, float *H,
                                uint16_t *Hb) {
    const int64_t total = (int64_t)n * T * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t nt = i / C;
        const int t = (int)(nt % T);
        float a = bc[c];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int tt = t + k - 1;
            if (tt < 0 || tt >= T) continue;
            const float *xp = X + 2 * (nt - t + tt);
            a = fmaf(Wc[(c * 2 + 0) * 3 + k], xp[0], a);
            a = fmaf(Wc[(c * 2 + 1) * 3 + k], xp[1], a);
        }
        a = a > 0.f ? a : 0.f;
        H[i] = a;
        Hb[i] = f2bf(a);
    }
}

// dWc[c][i][k] = sum dpre[n,t,c] X[n,t+k-1,i], dbc[c] = sum dpre, with
// dpre = dH * (H > 0).  One block per channel; the reduction is in fixed
// order within a block, so results are deterministic.
__global__ __launch_bounds__(256) void conv_bwd_kernel(const float *X, const float *H, const float *dH, int n, int T,
                                                       int C, float *dWc, float *dbc) {
    const int c = blockIdx.x;
    float acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const int64_t total = (int64_t)n * T;
    for (int64_t nt = threadIdx.x; nt < total; nt += blockDim.x) {
        const int64_t i = nt * C + c;
        const float g = H[i] > 0.f ? dH[i] : 0.f;
        if (g == 0.f) continue;
        const int t = (int)(nt % T);
        acc[6] += g;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int tt = t + k - 1;
            if (tt < 0 || tt >= T) continue;
            const float *xp = X + 2 * (nt - t + tt);
            acc[0 * 3 + k] = fmaf(g, xp[0], acc[0 * 3 + k]);
            acc[1 * 3 + k] = fmaf(g, xp[1], acc[1 * 3 + k]);
        }
    }
    __shared__ float sh[7][256];
#pragma unroll
    for (int q = 0; q < 7; ++q) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
#pragma unroll
            for (int q = 0; q < 7; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 6) dWc[c * 6 + threadIdx.x] = sh[threadIdx.x][0];
    if (threadIdx.x == 6) dbc[c] = sh[6][0];
}

__device__ __forceinline__ float sigm(float x) { return 1.f / (1.f + __expf(-x)); }

// h0 = tanh(u * lenf); writes fp32 and bf16 copies.
__global__ void gru_h0_kernel(const float *u, const float *lenf, int n, int Hd, float *h, uint16_t *hb) {
    const int64_t total = (int64_t)n * Hd;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = tanhf(u[i % Hd] * lenf[i / Hd]);
        h[i] = v;
        hb[i] = f2bf(v);
    }
}

// One GRU step: Gx row stride ldgx (the [n, T, 3H] input projections at step
// t), Gh [n, 3H] the recurrent projection of h.  Saves r, z, nn, Gh_n for the
// backward pass and writes h' (fp32 + bf16).
__global__ void gru_step_fwd_kernel(const float *Gx, int64_t ldgx, const float *Gh, const float *h, int n, int Hd,
                                    float *r_s, float *z_s, float *n_s, float *ghn_s, float *hn, uint16_t *hnb) {
    const int64_t total = (int64_t)n * Hd;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / Hd;
        const int j = (int)(i % Hd);
        const float *gx = Gx + row * ldgx;
        const float *gh = Gh + row * 3 * Hd;
        const float r = sigm(gx[j] + gh[j]);
        const float z = sigm(gx[Hd + j] + gh[Hd + j]);
        const float nn = tanhf(gx[2 * Hd + j] + r * gh[2 * Hd + j]);
        const float hv = (1.f - z) * nn + z * h[i];
        r_s[i] = r;
        z_s[i] = z;
        n_s[i] = nn;
        ghn_s[i] = gh[2 * Hd + j];
        hn[i] = hv;
        hnb[i] = f2bf(hv);
    }
}

// The whole T-step GRU forward for inference (no activations kept), H = 128:
// one 256-thread block per 16 trajectories runs every step, so T recurrent
// GEMMs and T gate launches become one launch.  Wave w owns hidden units
// [32 w, 32 w + 32): the three gate columns of unit j are the same lane's
// outputs in three MFMA tiles, so the gates and the new state stay in
// registers; the wave's W_hh fragments (6 tiles x 4 K chunks) stay in VGPRs
// for all steps; only the BF16 state goes through LDS (double-buffered, one
// barrier per step).  The arithmetic is the stepwise path's: each gate sum
// is the same four MFMA K chunks in order plus the bias (alpha = 1), and
// the gates use gru_step_fwd_kernel's expressions, so Hs and Hb are
// bit-identical to it.
constexpr int kGruH = 128, kGruRows = 16, kGruStride = kGruH + 16;  // LDS rows of 72 dwords
__global__ __launch_bounds__(256) void gru_fwd_fused_kernel(const float *Gx, int T, int n, const uint16_t *Whh,
                                                            const float *bhh, float *Hs, uint16_t *Hb) {
    __shared__ __attribute__((aligned(16))) uint16_t hl[2][kGruRows * kGruStride];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, kb = (lane >> 4) * 8;
    const int row0 = blockIdx.x * kGruRows;
    const int64_t nh = (int64_t)n * kGruH;
    // W_hh fragments of the wave's 6 output tiles: gate g, half q -> columns
    // g * 128 + 32 w + 16 q + (0..15)
    bf16x8 bw[6][4];
    float bias[6];
#pragma unroll
    for (int tile = 0; tile < 6; ++tile) {
        const int c = (tile >> 1) * kGruH + 32 * w + 16 * (tile & 1) + r16;
#pragma unroll
        for (int ch = 0; ch < 4; ++ch)
            bw[tile][ch] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(Whh + (int64_t)c * kGruH + 32 * ch + kb));
        bias[tile] = bhh[c];
    }
    // this lane's state: rows 4 * (lane >> 4) + e, units 32 w + 16 q + r16
    float h[2][4];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int row = min(row0 + 4 * (lane >> 4) + e, n - 1);
            h[q][e] = Hs[(int64_t)row * kGruH + 32 * w + 16 * q + r16];
        }
    // initial BF16 state into LDS buffer 0 (rows past n repeat row n - 1)
    for (int i = threadIdx.x; i < kGruRows * kGruH / 8; i += 256) {
        const int r = i / (kGruH / 8), k = (i % (kGruH / 8)) * 8;
        const int row = min(row0 + r, n - 1);
        *reinterpret_cast<uint4 *>(&hl[0][r * kGruStride + k]) = *reinterpret_cast<const uint4 *>(Hb + (int64_t)row * kGruH + k);
    }
    __syncthreads();
    for (int t = 0; t < T; ++t) {
        const int cur = t & 1;
        bf16x8 af[4];
#pragma unroll
        for (int ch = 0; ch < 4; ++ch)
            af[ch] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&hl[cur][r16 * kGruStride + 32 * ch + kb]));
        f32x4 acc[6];
#pragma unroll
        for (int tile = 0; tile < 6; ++tile) acc[tile] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ch = 0; ch < 4; ++ch)
#pragma unroll
            for (int tile = 0; tile < 6; ++tile)
                acc[tile] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[ch], bw[tile][ch], acc[tile], 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int rl = 4 * (lane >> 4) + e, row = row0 + rl;
                const int j = 32 * w + 16 * q + r16;
                const float *gx = Gx + ((int64_t)min(row, n - 1) * T + t) * 3 * kGruH;
                const float ghr = 1.f * acc[q][e] + bias[q];
                const float ghz = 1.f * acc[2 + q][e] + bias[2 + q];
                const float ghn = 1.f * acc[4 + q][e] + bias[4 + q];
                const float r = sigm(gx[j] + ghr);
                const float z = sigm(gx[kGruH + j] + ghz);
                const float nn = tanhf(gx[2 * kGruH + j] + r * ghn);
                const float hv = (1.f - z) * nn + z * h[q][e];
                h[q][e] = hv;
                const uint16_t hb = f2bf(hv);
                hl[cur ^ 1][rl * kGruStride + j] = hb;
                if (row < n) {
                    Hs[(int64_t)(t + 1) * nh + (int64_t)row * kGruH + j] = hv;
                    if (t == T - 1) Hb[(int64_t)row * kGruH + j] = hb;
                }
            }
        __syncthreads();
    }
}

// Backward of one GRU step.  dh: gradient w.r.t. h' (in), h: the step's
// input state.  Writes dGx (pre-activation gradients of the input
// projection, row stride lddgx), dGh [n, 3H] (fp32 + bf16) and
// dhp = dh * z, the direct part of the gradient w.r.t. h.
__global__ void gru_step_bwd_kernel(const float *dh, const float *h, const float *r_s, const float *z_s,
                                    const float *n_s, const float *ghn_s, int n, int Hd, float *dGx, int64_t lddgx,
                                    float *dGh, uint16_t *dGhb, float *dhp) {
    const int64_t total = (int64_t)n * Hd;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / Hd;
        const int j = (int)(i % Hd);
        const float g = dh[i], r = r_s[i], z = z_s[i], nn = n_s[i], ghn = ghn_s[i];
        const float dnn = g * (1.f - z);
        const float dz = g * (h[i] - nn);
        const float dan = dnn * (1.f - nn * nn);
        const float dar = dan * ghn * r * (1.f - r);
        const float daz = dz * z * (1.f - z);
        float *gx = dGx + row * lddgx;
        gx[j] = dar;
        gx[Hd + j] = daz;
        gx[2 * Hd + j] = dan;
        float *gh = dGh + row * 3 * Hd;
        uint16_t *ghb = dGhb + row * 3 * Hd;
        const float dghn = dan * r;
        gh[j] = dar;
        gh[Hd + j] = daz;
        gh[2 * Hd + j] = dghn;
        ghb[j] = f2bf(dar);
        ghb[Hd + j] = f2bf(daz);
        ghb[2 * Hd + j] = f2bf(dghn);
        dhp[i] = g * z;
    }
}

// du[j] = sum_n dh0[n,j] * (1 - h0^2) * lenf[n]  (h0 = tanh(u * lenf)).
__global__ __launch_bounds__(256) void gru_h0_bwd_kernel(const float *dh0, const float *h0, const float *lenf, int n,
                                                         int Hd, float *du) {
    const int j = blockIdx.x;
    float acc = 0.f;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        const float hv = h0[(int64_t)r * Hd + j];
        acc = fmaf(dh0[(int64_t)r * Hd + j] * (1.f - hv * hv), lenf[r], acc);
    }
    __shared__ float sh[256];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) du[j] = sh[0];
}

// out[c] = sum over rows of M[r, c] (row stride ld); one block per column,
// fixed reduction order.
__global__ __launch_bounds__(256) void colsum_kernel(const float *M, int64_t rows, int cols, int64_t ld, float *out,
                                                     float beta) {
    const int c = blockIdx.x;
    float acc = 0.f;
    for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) acc += M[r * ld + c];
    __shared__ float sh[256];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0] + (beta == 0.f ? 0.f : beta * out[c]);
}

// Pairwise ranking loss over anchor lists.  Embeddings E [B*(1+K), D]: row
// b*(1+K) is anchor b, the next K rows its candidates sorted by ascending
// exact distance.  For positives p < P and negatives q >= Q0 (ranks),
//   loss_b = mean softplus((d2_p - d2_q) / (tau * s_b)),  s_b = mean_k d2_k
// (s_b held constant), d2_k = |e_k - e_b|^2.  Writes dE (mean over anchors)
// and the per-anchor losses.  One block of 256 threads per anchor.
__global__ __launch_bounds__(256) void rank_loss_kernel(const float *E, int B, int K, int D, int P, int Q0, float tau,
                                                        float *dE, float *loss) {
    const int b = blockIdx.x;
    const float *ea = E + (int64_t)b * (1 + K) * D;
    __shared__ float d2[1024];
    __shared__ float g[1024];
    __shared__ float red[256];
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const float *ek = ea + (int64_t)(1 + k) * D;
        float s = 0.f;
        for (int q = 0; q < D; ++q) {
            const float t = ek[q] - ea[q];
            s = fmaf(t, t, s);
        }
        d2[k] = s;
        g[k] = 0.f;
    }
    __syncthreads();
    float part = 0.f;
    for (int k = threadIdx.x; k < K; k += blockDim.x) part += d2[k];
    red[threadIdx.x] = part;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const float scale = fmaxf(red[0] / (float)K, 1e-12f) * tau;
    __syncthreads();
    const int nneg = K - Q0;
    const float inv = 1.f / ((float)P * (float)nneg * (float)B);
    float lsum = 0.f;
    // thread handles negatives q; loops over the P positives
    for (int q = Q0 + threadIdx.x; q < K; q += blockDim.x) {
        float gq = 0.f;
        for (int p = 0; p < P; ++p) {
            const float x = (d2[p] - d2[q]) / scale;
            const float sp = x > 20.f ? x : log1pf(__expf(x));
            lsum += sp;
            const float s = sigm(x) * inv / scale;
            gq -= s;
            atomicAdd(&g[p], s);
        }
        g[q] = gq;
    }
    red[threadIdx.x] = lsum;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[b] = red[0] / ((float)P * (float)nneg);
    __syncthreads();
    // dE_k = 2 (e_k - e_a) g_k ; dE_a = -sum_k dE_k
    float *dea = dE + (int64_t)b * (1 + K) * D;
    for (int q = threadIdx.x; q < D; q += blockDim.x) dea[q] = 0.f;
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const float *ek = ea + (int64_t)(1 + k) * D;
        float *dek = dea + (int64_t)(1 + k) * D;
        for (int q = 0; q < D; ++q) dek[q] = 2.f * (ek[q] - ea[q]) * g[k];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < D; q += blockDim.x) {
        float s = 0.f;
        for (int k = 0; k < K; ++k) s += dea[(int64_t)(1 + k) * D + q];
        dea[q] = -s;
    }
}

// Distance regression over B anchor lists of K candidates with exact
// distances Y [B, K] (the reference's trainer fits embedding distances to the
// ground-truth ones, nn/MLP_trajsimi_train.py:78-82), in log space so near
// and far pairs weigh alike.  Per list, with d_k = |e_k - e_a|:
//   loss = (1/K) sum_k (log(d_k + c) - log(s y_k + c))^2,
// s = sum_k d_k / sum_k y_k held constant (the embedding keeps its own
// units), c = 1e-3 * mean d_k.  dE carries the 1/B of the mean over lists.
__global__ __launch_bounds__(256) void dist_loss_kernel(const float *E, const float *Y, int B, int K, int D,
                                                        float *dE, float *loss) {
    const int b = blockIdx.x;
    const float *ea = E + (int64_t)b * (1 + K) * D;
    const float *y = Y + (int64_t)b * K;
    __shared__ float dist[1024];
    __shared__ float g[1024];
    __shared__ float red[2][256];
    float pd = 0.f, py = 0.f;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const float *ek = ea + (int64_t)(1 + k) * D;
        float s = 0.f;
        for (int q = 0; q < D; ++q) {
            const float t = ek[q] - ea[q];
            s = fmaf(t, t, s);
        }
        dist[k] = sqrtf(s);
        pd += dist[k];
        py += y[k];
    }
    red[0][threadIdx.x] = pd;
    red[1][threadIdx.x] = py;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            red[0][threadIdx.x] += red[0][threadIdx.x + o];
            red[1][threadIdx.x] += red[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    const float sd = red[0][0], sy = red[1][0];
    const float sc = sy > 0.f ? sd / sy : 0.f;
    const float c = fmaxf(1e-3f * sd / (float)K, 1e-12f);
    __syncthreads();
    float lsum = 0.f;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const float r = logf(dist[k] + c) - logf(sc * y[k] + c);
        lsum += r * r;
        // dL/dd_k, then / d_k for the direction (e_k - e_a)
        g[k] = dist[k] > 0.f ? 2.f * r / ((dist[k] + c) * (float)K * (float)B * dist[k]) : 0.f;
    }
    red[0][threadIdx.x] = lsum;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[0][threadIdx.x] += red[0][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[b] = red[0][0] / (float)K;
    __syncthreads();
    // dE_k = (e_k - e_a) g_k ; dE_a = -sum_k dE_k
    float *dea = dE + (int64_t)b * (1 + K) * D;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const float *ek = ea + (int64_t)(1 + k) * D;
        float *dek = dea + (int64_t)(1 + k) * D;
        for (int q = 0; q < D; ++q) dek[q] = (ek[q] - ea[q]) * g[k];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < D; q += blockDim.x) {
        float s = 0.f;
        for (int k = 0; k < K; ++k) s += dea[(int64_t)(1 + k) * D + q];
        dea[q] = -s;
    }
}

// Adam on a flat FP32 parameter vector (bias-corrected, no weight decay).
__global__ void adam_kernel(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// Row squared norms of a bf16 [n, D] matrix, in FP32.
__global__ void rownorm_bf16_kernel(const uint16_t *E, int64_t n, int D, float *out) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int q = 0; q < D; ++q) {
            const float v = bf2f(E[r * D + q]);
            s = fmaf(v, v, s);
        }
        out[r] = s;
    }
}


// out[n, Dout] = X[n, K] . W[Dout, K]^T + beta * out, in FP32 (the fixed skip
// map of the encoder: precision matters more than speed for it).
__global__ void linear_f32_kernel(const float *X, int64_t n, int K, const float *W, int Dout, float *out, float beta) {
    const int64_t total = n * Dout;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / Dout;
        const int c = (int)(i % Dout);
        const float *x = X + r * K;
        const float *w = W + (int64_t)c * K;
        float a = 0.f;
        for (int k = 0; k < K; ++k) a = fmaf(x[k], w[k], a);
        out[i] = beta == 0.f ? a : a + beta * out[i];
    }
}

// Split FP32 rows E [n, D] into BF16 hi and lo = bf16(e - hi) parts laid
// out as [hi, lo] (database side) or [hi, hi] (query side) in out [n, 2D],
// and the FP32 squared row norms of E: one K = 2D product then gives
// q_hi . (d_hi + d_lo), the database side to ~16 mantissa bits.  The
// query's own rounding (q_lo, dropped) shifts the distances of two
// candidates d1, d2 by -2 q_lo . (d1 - d2), small exactly where ranking
// matters (near neighbours lie near each other); the database rounding,
// kept, would not be.  On the 1,000 x 1,000,000 kNN workload this gives
// the FP32 product's recall (0.9562 / 0.9565 at 100 candidates), against
// 0.918 with plain BF16 embeddings.
// One wavefront per row: lanes over the D values (coalesced), the row's
// squared norm as a fixed-order wave sum (lane partials over q = lane,
// lane + 64, ..., then a butterfly).
__global__ __launch_bounds__(256) void split_bf16_kernel(const float *E, int64_t n, int D, int query_side,
                                                         uint16_t *out, float *norm) {
    const int lane = threadIdx.x & 63;
    for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; r < n;
         r += ((int64_t)gridDim.x * blockDim.x) >> 6) {
        const float *e = E + r * D;
        uint16_t *o = out + r * 2 * D;
        float s = 0.f;
        for (int q = lane; q < D; q += 64) {
            const float v = e[q];
            s = fmaf(v, v, s);
            const uint16_t hi = f2bf(v);
            const uint16_t lo = f2bf(v - bf2f(hi));
            o[q] = hi;
            o[D + q] = query_side ? hi : lo;
        }
#pragma unroll
        for (int w = 32; w >= 1; w >>= 1) s += __shfl_xor(s, w, 64);
        if (norm && lane == 0) norm[r] = s;
    }
}

// ---------------------------------------------------------------- top-k

// Order-preserving unsigned keys of floating-point values.
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint64_t fkey(double f) {
    const uint64_t u = __builtin_bit_cast(uint64_t, f);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

// k smallest entries of each row of V [rows, cols] (row stride ld): one
// block per row; radix select on 8-bit digits of the order-preserving key
// finds the k-th smallest key, then the row is gathered (all keys below it,
// then keys equal to it up to k).  Output order within a row is unspecified.
template <typename T>
__global__ __launch_bounds__(256) void topk_kernel(const T *V, int64_t rows, int64_t cols, int64_t ld, int k,
                                                   int32_t *out_idx, T *out_val) {
    using K_t = decltype(fkey((T)0));
    constexpr int kBits = sizeof(K_t) * 8;
    __shared__ uint32_t hist[256];
    __shared__ K_t prefix_s;
    __shared__ int remain_s;
    __shared__ int n_less, n_eq;
    const int64_t row = blockIdx.x;
    const T *v = V + row * ld;
    K_t prefix = 0, mask = 0;
    int remain = k;
    for (int shift = kBits - 8; shift >= 0; shift -= 8) {
        hist[threadIdx.x] = 0;
        __syncthreads();
        for (int64_t c = threadIdx.x; c < cols; c += blockDim.x) {
            const K_t key = fkey(v[c]);
            if ((key & mask) == prefix) atomicAdd(&hist[(uint32_t)(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int cum = 0, d = 0;
            for (; d < 256; ++d) {
                if (cum + (int)hist[d] >= remain) break;
                cum += (int)hist[d];
            }
            if (d > 255) d = 255;
            prefix_s = prefix | ((K_t)d << shift);
            remain_s = remain - cum;
        }
        __syncthreads();
        prefix = prefix_s;
        remain = remain_s;
        mask |= ((K_t)255) << shift;
        __syncthreads();
    }
    // prefix is the k-th smallest key; (k - remain) keys are strictly below it
    if (threadIdx.x == 0) {
        n_less = 0;
        n_eq = 0;
    }
    __syncthreads();
    const int below = k - remain;
    int32_t *oi = out_idx + row * (int64_t)k;
    T *ov = out_val ? out_val + row * (int64_t)k : nullptr;
    for (int64_t c = threadIdx.x; c < cols; c += blockDim.x) {
        const K_t key = fkey(v[c]);
        int slot = -1;
        if (key < prefix) slot = atomicAdd(&n_less, 1);
        else if (key == prefix) {
            const int e = atomicAdd(&n_eq, 1);
            if (e < remain) slot = below + e;
        }
        if (slot >= 0 && slot < k) {
            oi[slot] = (int32_t)c;
            if (ov) ov[slot] = v[c];
        }
    }
}

// The k-th smallest value of each row (what the threshold search needs: the
// largest of the k smallest), for rows of at most kKthCols values: the row's
// order-preserving keys sit in LDS, and four radix passes of 8 bits build
// per-wave histograms there (so a pass never re-reads global memory and
// same-bin atomics collide only within a wave), merged by 256 threads and
// scanned by one wave.
constexpr int kKthCols = 16384, kKthThreads = 1024;

__global__ __launch_bounds__(kKthThreads) void kth_smallest_kernel(const float *V, int64_t cols, int64_t ld, int k,
                                                                   float *out) {
    __shared__ uint32_t keys[kKthCols];
    __shared__ uint32_t hist[kKthThreads / 64][256];
    __shared__ uint32_t tot[256];
    __shared__ uint32_t prefix_s, remain_s;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const float *v = V + (int64_t)blockIdx.x * ld;
    for (int64_t c = t; c < cols; c += kKthThreads) keys[c] = fkey(v[c]);
    uint32_t prefix = 0, mask = 0, remain = (uint32_t)k;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = t; i < (kKthThreads / 64) * 256; i += kKthThreads) (&hist[0][0])[i] = 0;
        __syncthreads();
        for (int64_t c = t; c < cols; c += kKthThreads) {
            const uint32_t key = keys[c];
            if ((key & mask) == prefix) atomicAdd(&hist[wave][(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (t < 256) {
            uint32_t sum = 0;
#pragma unroll
            for (int w = 0; w < kKthThreads / 64; ++w) sum += hist[w][t];
            tot[t] = sum;
        }
        __syncthreads();
        if (wave == 0) {  // digit d with cum(d - 1) < remain <= cum(d): lane l scans bins 4l .. 4l + 3
            const uint32_t h0 = tot[4 * lane], h1 = tot[4 * lane + 1], h2 = tot[4 * lane + 2], h3 = tot[4 * lane + 3];
            uint32_t incl = h0 + h1 + h2 + h3;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o, 64);
                if (lane >= o) incl += y;
            }
            const uint32_t before = incl - (h0 + h1 + h2 + h3);
            const bool here = before < remain && remain <= incl;
            const uint64_t m = __builtin_amdgcn_ballot_w64(here);
            const int owner = m ? __builtin_ctzll(m) : 63;
            if (lane == owner) {
                uint32_t cum = before, d = 4 * lane;
                const uint32_t hs[4] = {h0, h1, h2, h3};
                int j = 0;
                while (j < 3 && cum + hs[j] < remain) {
                    cum += hs[j];
                    ++j;
                }
                d += j;
                prefix_s = prefix | (d << shift);
                remain_s = remain - cum;
            }
        }
        __syncthreads();
        prefix = prefix_s;
        remain = remain_s;
        mask |= 255u << shift;
    }
    if (t == 0) {
        const uint32_t u = (prefix & 0x80000000u) ? (prefix & 0x7fffffffu) : ~prefix;
        out[blockIdx.x] = __builtin_bit_cast(float, u);
    }
}

// Top-k of short rows (at most 512 columns, k <= kKthSmallK): one wave per
// row; each lane sorts its 8 (key, column) pairs (columns lane + 64 i), and k
// rounds of a wave minimum pop the k smallest in ascending (value, column)
// order.  Used for the 10 nearest of each query's re-ranked candidates,
// where the radix top-k's passes cost more than the row itself.
constexpr int kTopkSmallCols = 512;
template <typename T>
__global__ __launch_bounds__(256) void topk_small_kernel(const T *V, int64_t rows, int64_t cols, int64_t ld, int k,
                                                         int32_t *out_idx, T *out_val) {
    using K_t = decltype(fkey((T)0));
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;  // wave-uniform
    const T *v = V + row * ld;
    K_t key[8];
    int col[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int c = lane + 64 * i;
        key[i] = c < cols ? fkey(v[c]) : (K_t)~(K_t)0;
        col[i] = c < cols ? c : 0x7fffffff;
    }
    auto less = [](K_t ka, int ca, K_t kb, int cb) { return ka < kb || (ka == kb && ca < cb); };
#pragma unroll
    for (int kk = 2; kk <= 8; kk <<= 1)
#pragma unroll
        for (int j = kk >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & kk) == 0;
                    const bool sw = up ? less(key[l], col[l], key[i], col[i]) : less(key[i], col[i], key[l], col[l]);
                    const K_t ka = key[i];
                    const int ca = col[i];
                    key[i] = sw ? key[l] : key[i];
                    col[i] = sw ? col[l] : col[i];
                    key[l] = sw ? ka : key[l];
                    col[l] = sw ? ca : col[l];
                }
            }
    for (int r = 0; r < k; ++r) {
        K_t mk = key[0];
        int mc = col[0];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const K_t ok = __shfl_xor(mk, o, kWave);
            const int oc = __shfl_xor(mc, o, kWave);
            if (less(ok, oc, mk, mc)) {
                mk = ok;
                mc = oc;
            }
        }
        const bool pop = col[0] == mc;  // columns are unique: one lane
#pragma unroll
        for (int i = 0; i < 7; ++i) {
            key[i] = pop ? key[i + 1] : key[i];
            col[i] = pop ? col[i + 1] : col[i];
        }
        key[7] = pop ? (K_t)~(K_t)0 : key[7];
        col[7] = pop ? 0x7fffffff : col[7];
        if (lane == 0) {
            out_idx[row * k + r] = mc;
            if (out_val) out_val[row * k + r] = v[mc];
        }
    }
}

// The same for small k (<= kKthSmallK), without histograms: each of 1024
// lanes sorts its 16 keys (columns lane + 1024 i) in registers, each wave
// pops its k smallest by k rounds of a wave minimum (the first lane holding
// it pops), and one wave merges the 16 sorted lists the same way.  The
// k-th smallest key is the radix kernel's, so the value is identical.  (In
// the radix kernel's first pass almost every key of a row shares the top
// byte, so its histogram atomics serialise within each wave.)
constexpr int kKthSmallK = 32;

// Minimum over the 64 lanes (all active), uniform: DPP moves instead of LDS
// permutes.  quad_perm xor 1 and xor 2, then the half-row and row mirrors,
// leave every lane of a 16-lane row holding the row's minimum; row_bcast:15
// folds rows 0 into 1 and 2 into 3, row_bcast:31 folds row 1 into rows 2
// and 3, so lane 63 holds the wave's minimum.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_min_u32(uint32_t v) {
    const uint32_t o = (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, ROW_MASK, 0xf, false);
    return min(v, o);
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    v = dpp_min_u32<0xb1>(v);        // quad_perm [1, 0, 3, 2]
    v = dpp_min_u32<0x4e>(v);        // quad_perm [2, 3, 0, 1]
    v = dpp_min_u32<0x141>(v);       // row_half_mirror
    v = dpp_min_u32<0x140>(v);       // row_mirror
    v = dpp_min_u32<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_min_u32<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__global__ __launch_bounds__(kKthThreads) void kth_small_k_kernel(const float *V, int64_t cols, int64_t ld, int k,
                                                                  float *out) {
    __shared__ uint32_t wl[kKthThreads / 64][kKthSmallK];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const float *v = V + (int64_t)blockIdx.x * ld;
    uint32_t a[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int64_t c = t + (int64_t)kKthThreads * i;
        a[i] = c < cols ? fkey(v[c]) : 0xffffffffu;  // past the row: larger than every key
    }
    // bitonic sort of the 16 keys, ascending
#pragma unroll
    for (int kk = 2; kk <= 16; kk <<= 1)
#pragma unroll
        for (int j = kk >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t x = a[i], y = a[l];
                    const uint32_t lo = min(x, y), hi = max(x, y);
                    if ((i & kk) == 0) {
                        a[i] = lo;
                        a[l] = hi;
                    } else {
                        a[i] = hi;
                        a[l] = lo;
                    }
                }
            }
    for (int r = 0; r < k; ++r) {
        const uint32_t m = wave_min_u32(a[0]);
        const uint64_t b = __builtin_amdgcn_ballot_w64(a[0] == m);
        const bool pop = lane == (int)__builtin_ctzll(b);
#pragma unroll
        for (int i = 0; i < 15; ++i) a[i] = pop ? a[i + 1] : a[i];
        a[15] = pop ? 0xffffffffu : a[15];
        if (lane == 0) wl[wave][r] = m;
    }
    __syncthreads();
    if (wave == 0) {
        int p = 0;
        uint32_t m = 0;
        for (int r = 0; r < k; ++r) {
            const uint32_t head = (lane < kKthThreads / 64 && p < k) ? wl[lane][p] : 0xffffffffu;
            m = wave_min_u32(head);
            const uint64_t b = __builtin_amdgcn_ballot_w64(head == m);
            if (lane == (int)__builtin_ctzll(b)) ++p;
        }
        if (lane == 0) {
            const uint32_t u = (m & 0x80000000u) ? (m & 0x7fffffffu) : ~m;
            out[blockIdx.x] = __builtin_bit_cast(float, u);
        }
    }
}

}  // namespace tsm

// ---------------------------------------------------------------- C ABI (part)

namespace {
using namespace tsm;
inline int grid1(int64_t n, int b = 256) {
    int64_t g = (n + b - 1) / b;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}
}  // namespace

extern "C" {

int tsm_nn_gemm_nt(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb, float *C, int ldc,
                   float alpha, float beta, const float *bias, void *stream) {
    if (M < 0 || N < 0 || K < 0 || (K % 8) != 0) return fail(TSM_E_ARG, "gemm: K must be a multiple of 8");
    if (M == 0 || N == 0) return TSM_OK;
    if (!A || !B || !C || lda < K || ldb < K || ldc < N || (lda % 8) || (ldb % 8) ||
        (reinterpret_cast<uintptr_t>(A) % 16) || (reinterpret_cast<uintptr_t>(B) % 16))
        return fail(TSM_E_ARG, "gemm: bad operands (16-byte aligned, leading dimensions multiples of 8)");
    GemmArgs g{M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, bias};
    if (M >= kBM && N >= kBN && K >= kBK && (M + kBM - 1) / kBM <= 65535) {
        dim3 grid((N + kBN - 1) / kBN, (M + kBM - 1) / kBM);
        hipLaunchKernelGGL(gemm_nt_lds_kernel<0>, grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), g,
                           FilterArgs{}, 0);
    } else {
        if ((M + 63) / 64 > 65535) return fail(TSM_E_ARG, "gemm: M too large");
        dim3 grid((N + 63) / 64, (M + 63) / 64);
        hipLaunchKernelGGL(gemm_nt_bf16_kernel, grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), g);
    }
    TSM_HIP(hipGetLastError());
    return TSM_OK;
}

int tsm_nn_gemm_nt_filter(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb, float alpha,
                          const float *bias, const float *tau, int cap, uint32_t *cnt, float *key, int32_t *idx,
                          void *stream) {
    if (M < 0 || N < 0 || K < 0 || (K % 8) != 0) return fail(TSM_E_ARG, "gemm_filter: K must be a multiple of 8");
    if (M == 0 || N == 0) return TSM_OK;
    if (!A || !B || !tau || !cnt || !key || !idx || cap < 1 || lda < K || ldb < K || (lda % 8) || (ldb % 8) ||
        (reinterpret_cast<uintptr_t>(A) % 16) || (reinterpret_cast<uintptr_t>(B) % 16))
        return fail(TSM_E_ARG, "gemm_filter: bad operands (16-byte aligned, leading dimensions multiples of 8)");
    if ((M + 255) / 256 > 65535) return fail(TSM_E_ARG, "gemm_filter: M too large");
    GemmArgs g{M, N, K, A, lda, B, ldb, nullptr, 0, alpha, 0.f, bias};
    const FilterArgs fa{tau, cap, cnt, key, idx};
    if (K == 256 || K == 384) {  // the embedding search: persistent blocks (K in whole steps of 64)
        int dev = 0, ncu = 256;
        TSM_HIP(hipGetDevice(&dev));
        TSM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const int qblocks = (M + kSearchRows - 1) / kSearchRows, ntiles = (N + kBN - 1) / kBN;
        int chunks = (ncu / qblocks) / 8 * 8;
        if (chunks < 8) chunks = 8;
        const int tpc = (ntiles + chunks - 1) / chunks;
        if (K == 256)
            hipLaunchKernelGGL((search_filter_kernel<8, kSearchThreads, kSearchColTiles, kSearchKW256>), dim3(qblocks * chunks),
                               dim3(kSearchThreads), 0, reinterpret_cast<hipStream_t>(stream), g, fa, qblocks, chunks,
                               tpc);
        else
            hipLaunchKernelGGL((search_filter_kernel<12, kSearchThreads, kSearchColTiles, 64>), dim3(qblocks * chunks),
                               dim3(kSearchThreads), 0, reinterpret_cast<hipStream_t>(stream), g, fa, qblocks, chunks,
                               tpc);
    } else if (M > kBM) {  // 256 query rows per block: the database is streamed half as often
        dim3 grid((N + kBN - 1) / kBN, (M + 255) / 256);
        hipLaunchKernelGGL((gemm_nt_lds_kernel<1, 256>), grid, dim3(512), 0, reinterpret_cast<hipStream_t>(stream), g,
                           fa, 0);
    } else {
        dim3 grid((N + kBN - 1) / kBN, (M + kBM - 1) / kBM);
        hipLaunchKernelGGL((gemm_nt_lds_kernel<1, kBM>), grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), g,
                           fa, 0);
    }
    TSM_HIP(hipGetLastError());
    return TSM_OK;
}

int tsm_nn_gemm_nt_acc(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb, float *C,
                       int ldc, float alpha, void *stream) {
    if (M < 0 || N < 0 || K < 0 || (K % 8) != 0) return fail(TSM_E_ARG, "gemm_acc: K must be a multiple of 8");
    if (M == 0 || N == 0 || K == 0) return TSM_OK;
    if (!A || !B || !C || lda < K || ldb < K || ldc < N || (lda % 8) || (ldb % 8) ||
        (reinterpret_cast<uintptr_t>(A) % 16) || (reinterpret_cast<uintptr_t>(B) % 16))
        return fail(TSM_E_ARG, "gemm_acc: bad operands (16-byte aligned, leading dimensions multiples of 8)");
    const int tm = (M + kBM - 1) / kBM, tn = (N + kBN - 1) / kBN;
    if (tm > 65535) return fail(TSM_E_ARG, "gemm_acc: M too large");
    // enough K slices for ~512 blocks, each at least 16 K steps long
    const int ksteps = (K + kBK - 1) / kBK;
    int splits = 512 / (tm * tn);
    if (splits > ksteps / 16) splits = ksteps / 16;
    if (splits < 1) splits = 1;
    if (splits > 65535) splits = 65535;
    const int kchunk = ((ksteps + splits - 1) / splits) * kBK;
    splits = (K + kchunk - 1) / kchunk;
    GemmArgs g{M, N, K, A, lda, B, ldb, C, ldc, alpha, 1.f, nullptr};
    hipLaunchKernelGGL(gemm_nt_lds_kernel<2>, dim3(tn, tm, splits), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       g, FilterArgs{}, kchunk);
    TSM_HIP(hipGetLastError());
    return TSM_OK;
}

int tsm_nn_select_candidates(const uint32_t *cnt, const float *key, const int32_t *idx, int rows, int cap, int C,
                             int32_t *out, int32_t *flag, void *stream) {
    if (rows < 0 || cap < 1 || cap > 4096 || C < 1 || C > cap || !cnt || !key || !idx || !out || !flag)
        return fail(TSM_E_ARG, "select_candidates: need 1 <= C <= cap <= 4096");
    if (rows == 0) return TSM_OK;
    hipLaunchKernelGGL(select_candidates_kernel, dim3(rows), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), cnt,
                       key, idx, cap, C, out, flag, SelectPairs{nullptr, nullptr, 0, nullptr, nullptr});
    TSM_HIP(hipGetLastError());
    return TSM_OK;
}

int tsm_nn_select_candidates_pairs(const uint32_t *cnt, const float *key, const int32_t *idx, int rows, int cap,
                                   int C, int32_t *out, int32_t *flag, const int32_t *row_ids,
                                   const int32_t *col_ids, int64_t ncol_ids, int32_t *pa, int32_t *pb,
                                   void *stream) {
    if (rows < 0 || cap < 1 || cap > 4096 || C < 1 || C > cap || !cnt || !key || !idx || !out || !flag)
        return fail(TSM_E_ARG, "select_candidates_pairs: need 1 <= C <= cap <= 4096");
    if (!row_ids || !col_ids || ncol_ids < 1 || !pa || !pb)
        return fail(TSM_E_ARG, "select_candidates_pairs: need row_ids, col_ids (ncol_ids >= 1), pa and pb");
    if (rows == 0) return TSM_OK;
    hipLaunchKernelGGL(select_candidates_kernel, dim3(rows), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), cnt,
                       key, idx, cap, C, out, flag, SelectPairs{row_ids, col_ids, ncol_ids, pa, pb});
    TSM_HIP(hipGetLastError());
    return TSM_OK;
}

int tsm_nn_cast_bf16(const float *src, int rows, int cols, int ld_src, uint16_t *dst, int ld_dst, int transpose,
                     void *stream) {
    if (rows < 0 || cols < 0 || !src || !dst) return fail(TSM_E_ARG, "cast: bad arguments");
    if (ld_src < cols || ld_dst < (transpose ? rows : cols)) return fail(TSM_E_ARG, "cast: bad leading dimension");
    if ((int64_t)rows * cols == 0) return TSM_OK;
    hipLaunchKernelGGL(cast_bf16_kernel, dim3(grid1((int64_t)rows * cols)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), src, rows, cols, ld_src, dst, ld_dst, transpose);
    TSM_HIP(hipGetLastError());
    return TSM_OK;
}


#define TSM_ST reinterpret_cast<hipStream_t>(stream)
#define TSM_LAUNCHED()            \
    do {                          \
        TSM_HIP(hipGetLastError()); \
        return TSM_OK;            \
    } while (0)

int tsm_nn_resample(const double *xy64, const int64_t *offsets, const int32_t *ids, int n, int T, float *X,
                    float *lenf, void *stream) {
    if (n < 0 || T < 2 || !xy64 || !offsets || !X || !lenf) return fail(TSM_E_ARG, "resample: bad arguments");
    if (n == 0) return TSM_OK;
    hipLaunchKernelGGL(resample_kernel, dim3(grid1((int64_t)n * T)), dim3(256), 0, TSM_ST, xy64, offsets, ids, n, T, X,
                       lenf);
    TSM_LAUNCHED();
}

int tsm_nn_conv_fwd(const float *X, int n, int T, int C, const float *Wc, const float *bc, float *H, uint16_t *Hb,
                    void *stream) {
    if (n < 0 || T < 1 || C < 1 || !X || !Wc || !bc || !H || !Hb) return fail(TSM_E_ARG, "conv_fwd: bad arguments");
    if (n == 0) return TSM_OK;
    hipLaunchKernelGGL(conv_fwd_kernel, dim3(grid1((int64_t)n * T * C)), dim3(256), 0, TSM_ST, X, n, T, C, Wc, bc, H,
                       Hb);
    TSM_LAUNCHED();
}

int tsm_nn_conv_bwd(const float *X, const float *H, const float *dH, int n, int T, int C, float *dWc, float *dbc,
                    void *stream) {
    if (n < 1 || T < 1 || C < 1 || !X || !H || !dH || !dWc || !dbc) return fail(TSM_E_ARG, "conv_bwd: bad arguments");
    hipLaunchKernelGGL(conv_bwd_kernel, dim3(C), dim3(256), 0, TSM_ST, X, H, dH, n, T, C, dWc, dbc);
    TSM_LAUNCHED();
}

int tsm_nn_gru_h0(const float *u, const float *lenf, int n, int Hd, float *h, uint16_t *hb, void *stream) {
    if (n < 0 || Hd < 1 || !u || !lenf || !h || !hb) return fail(TSM_E_ARG, "gru_h0: bad arguments");
    if (n == 0) return TSM_OK;
    hipLaunchKernelGGL(gru_h0_kernel, dim3(grid1((int64_t)n * Hd)), dim3(256), 0, TSM_ST, u, lenf, n, Hd, h, hb);
    TSM_LAUNCHED();
}

int tsm_nn_gru_step_fwd(const float *Gx, int64_t ldgx, const float *Gh, const float *h, int n, int Hd, float *r,
                        float *z, float *nn, float *ghn, float *hn, uint16_t *hnb, void *stream) {
    if (n < 0 || Hd < 1 || ldgx < 3 * Hd) return fail(TSM_E_ARG, "gru_step_fwd: bad arguments");
    if (n == 0) return TSM_OK;
    hipLaunchKernelGGL(gru_step_fwd_kernel, dim3(grid1((int64_t)n * Hd)), dim3(256), 0, TSM_ST, Gx, ldgx, Gh, h, n, Hd,
                       r, z, nn, ghn, hn, hnb);
    TSM_LAUNCHED();
}

int tsm_nn_gru_forward(const float *Gx, int T, int n, int Hd, const uint16_t *Whh, const float *bhh, float *Hs,
                       uint16_t *Hb, float *Gh, float *r, float *z, float *nn, float *ghn, int keep, void *stream) {
    if (T < 1 || n < 0 || Hd < 1 || (Hd % 8) != 0) return fail(TSM_E_ARG, "gru_forward: bad arguments");
    if (n == 0) return TSM_OK;
    if (!keep && Hd == kGruH && !(reinterpret_cast<uintptr_t>(Whh) % 16) && !(reinterpret_cast<uintptr_t>(Hb) % 16)) {
        hipLaunchKernelGGL(gru_fwd_fused_kernel, dim3((n + kGruRows - 1) / kGruRows), dim3(256), 0, TSM_ST, Gx, T, n,
                           Whh, bhh, Hs, Hb);
        TSM_LAUNCHED();
    }
    const int64_t nh = (int64_t)n * Hd;
    for (int t = 0; t < T; ++t) {
        int rc = tsm_nn_gemm_nt(n, 3 * Hd, Hd, Hb, Hd, Whh, Hd, Gh, 3 * Hd, 1.f, 0.f, bhh, stream);
        if (rc) return rc;
        const int64_t k = keep ? t : 0;
        rc = tsm_nn_gru_step_fwd(Gx + (int64_t)t * 3 * Hd, (int64_t)T * 3 * Hd, Gh, Hs + t * nh, n, Hd, r + k * nh,
                                 z + k * nh, nn + k * nh, ghn + k * nh, Hs + (t + 1) * nh, Hb, stream);
        if (rc) return rc;
    }
    return TSM_OK;
}

int tsm_nn_gru_step_bwd(const float *dh, const float *h, const float *r, const float *z, const float *nn,
                        const float *ghn, int n, int Hd, float *dGx, int64_t lddgx, float *dGh, uint16_t *dGhb,
                        float *dhp, void *stream) {
    if (n < 0 || Hd < 1 || lddgx < 3 * Hd) return fail(TSM_E_ARG, "gru_step_bwd: bad arguments");
    if (n == 0) return TSM_OK;
    hipLaunchKernelGGL(gru_step_bwd_kernel, dim3(grid1((int64_t)n * Hd)), dim3(256), 0, TSM_ST, dh, h, r, z, nn, ghn,
                       n, Hd, dGx, lddgx, dGh, dGhb, dhp);
    TSM_LAUNCHED();
}

int tsm_nn_gru_h0_bwd(const float *dh0, const float *h0, const float *lenf, int n, int Hd, float *du, void *stream) {
    if (n < 1 || Hd < 1) return fail(TSM_E_ARG, "gru_h0_bwd: bad arguments");
    hipLaunchKernelGGL(gru_h0_bwd_kernel, dim3(Hd), dim3(256), 0, TSM_ST, dh0, h0, lenf, n, Hd, du);
    TSM_LAUNCHED();
}

int tsm_nn_colsum(const float *M, int64_t rows, int cols, int64_t ld, float *out, float beta, void *stream) {
    if (rows < 0 || cols < 1 || ld < cols || !M || !out) return fail(TSM_E_ARG, "colsum: bad arguments");
    hipLaunchKernelGGL(colsum_kernel, dim3(cols), dim3(256), 0, TSM_ST, M, rows, cols, ld, out, beta);
    TSM_LAUNCHED();
}

int tsm_nn_rank_loss(const float *E, int B, int K, int D, int P, int Q0, float tau, float *dE, float *loss,
                     void *stream) {
    if (B < 1 || K < 2 || K > 1024 || D < 1 || P < 1 || Q0 < P || Q0 >= K || !(tau > 0.f) || !E || !dE || !loss)
        return fail(TSM_E_ARG, "rank_loss: need 1 <= P <= Q0 < K <= 1024, tau > 0");
    hipLaunchKernelGGL(rank_loss_kernel, dim3(B), dim3(256), 0, TSM_ST, E, B, K, D, P, Q0, tau, dE, loss);
    TSM_LAUNCHED();
}

int tsm_nn_dist_loss(const float *E, const float *Y, int B, int K, int D, float *dE, float *loss, void *stream) {
    if (B < 1 || K < 1 || K > 1024 || D < 1 || !E || !Y || !dE || !loss)
        return fail(TSM_E_ARG, "dist_loss: need B >= 1, 1 <= K <= 1024, D >= 1");
    hipLaunchKernelGGL(dist_loss_kernel, dim3(B), dim3(256), 0, TSM_ST, E, Y, B, K, D, dE, loss);
    TSM_LAUNCHED();
}

int tsm_nn_adam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                int64_t step, void *stream) {
    if (n < 0 || step < 1 || !p || !g || !m || !v) return fail(TSM_E_ARG, "adam: bad arguments");
    if (n == 0) return TSM_OK;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid1(n)), dim3(256), 0, TSM_ST, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    TSM_LAUNCHED();
}

int tsm_nn_rownorm_bf16(const uint16_t *E, int64_t n, int D, float *out, void *stream) {
    if (n < 0 || D < 1 || !E || !out) return fail(TSM_E_ARG, "rownorm: bad arguments");
    if (n == 0) return TSM_OK;
    hipLaunchKernelGGL(rownorm_bf16_kernel, dim3(grid1(n)), dim3(256), 0, TSM_ST, E, n, D, out);
    TSM_LAUNCHED();
}

int tsm_nn_topk_f32(const float *V, int64_t rows, int64_t cols, int64_t ld, int k, int32_t *idx, float *val,
                    void *stream) {
    if (rows < 0 || k < 1 || k > cols || ld < cols || !V || !idx) return fail(TSM_E_ARG, "topk: need 1 <= k <= cols");
    if (rows == 0) return TSM_OK;
    if (rows > 0x7fffffff) return fail(TSM_E_ARG, "topk: too many rows");
    if (cols <= kTopkSmallCols && k <= kKthSmallK)
        hipLaunchKernelGGL(topk_small_kernel<float>, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, TSM_ST, V, rows, cols,
                           ld, k, idx, val);
    else
        hipLaunchKernelGGL(topk_kernel<float>, dim3((unsigned)rows), dim3(256), 0, TSM_ST, V, rows, cols, ld, k, idx, val);
    TSM_LAUNCHED();
}

int tsm_nn_kth_smallest_f32(const float *V, int64_t rows, int64_t cols, int64_t ld, int k, float *out, void *stream) {
    if (rows < 0 || k < 1 || k > cols || ld < cols || !V || !out) return fail(TSM_E_ARG, "kth_smallest: need 1 <= k <= cols");
    if (cols > kKthCols) return fail(TSM_E_ARG, "kth_smallest: at most 16384 columns");
    if (rows == 0) return TSM_OK;
    if (rows > 0x7fffffff) return fail(TSM_E_ARG, "kth_smallest: too many rows");
    if (k <= kKthSmallK)
        hipLaunchKernelGGL(kth_small_k_kernel, dim3((unsigned)rows), dim3(kKthThreads), 0, TSM_ST, V, cols, ld, k, out);
    else
        hipLaunchKernelGGL(kth_smallest_kernel, dim3((unsigned)rows), dim3(kKthThreads), 0, TSM_ST, V, cols, ld, k, out);
    TSM_LAUNCHED();
}

int tsm_nn_topk_f64(const double *V, int64_t rows, int64_t cols, int64_t ld, int k, int32_t *idx, double *val,
                    void *stream) {
    if (rows < 0 || k < 1 || k > cols || ld < cols || !V || !idx) return fail(TSM_E_ARG, "topk: need 1 <= k <= cols");
    if (rows == 0) return TSM_OK;
    if (rows > 0x7fffffff) return fail(TSM_E_ARG, "topk: too many rows");
    if (cols <= kTopkSmallCols && k <= kKthSmallK)
        hipLaunchKernelGGL(topk_small_kernel<double>, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, TSM_ST, V, rows, cols,
                           ld, k, idx, val);
    else
        hipLaunchKernelGGL(topk_kernel<double>, dim3((unsigned)rows), dim3(256), 0, TSM_ST, V, rows, cols, ld, k, idx, val);
    TSM_LAUNCHED();
}


int tsm_nn_linear_f32(const float *X, int64_t n, int K, const float *W, int Dout, float *out, float beta,
                      void *stream) {
    if (n < 0 || K < 1 || Dout < 1 || !X || !W || !out) return fail(TSM_E_ARG, "linear: bad arguments");
    if (n == 0) return TSM_OK;
    hipLaunchKernelGGL(linear_f32_kernel, dim3(grid1(n * Dout)), dim3(256), 0, TSM_ST, X, n, K, W, Dout, out, beta);
    TSM_LAUNCHED();
}

int tsm_nn_split_bf16(const float *E, int64_t n, int D, int query_side, uint16_t *out, float *norm, void *stream) {
    if (n < 0 || D < 1 || !E || !out) return fail(TSM_E_ARG, "split: bad arguments");
    if (n == 0) return TSM_OK;
    hipLaunchKernelGGL(split_bf16_kernel, dim3(grid1(n * 64)), dim3(256), 0, TSM_ST, E, n, D, query_side, out, norm);
    TSM_LAUNCHED();
}

}  // extern "C"
