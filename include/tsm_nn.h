/*
 * tsm_nn.h -- C ABI of the learned kNN filter kernels in libtsm.so: the
 * trajectory encoder (resampling, Conv1d(2->64) + ReLU, GRU(64->128), linear
 * head), its backward pass and Adam update, and the embedding search (BF16
 * MFMA GEMM, per-row top-k).  The Python trainer and searcher
 * (trajsimimeasures_amd/filter.py) drive these; every pointer is device
 * memory, every call is stream-ordered on `stream` and returns TSM_OK or a
 * negative TSM_E* code (tsm_last_error()).
 *
 * No reference counterpart for the kernels themselves: the reference's
 * learned measures are PyTorch models (nn/MLP.py and the other nn modules)
 * queried through faiss (test_knn_time.py:138-202).  This filter keeps that
 * embed -> search -> candidates pipeline and adds an exact re-rank of the
 * candidates with the DP kernels (tsm_pairs).
 *
 * Conventions: matrices are row-major; "bf16" buffers hold bfloat16 bit
 * patterns (uint16_t).  The GRU gate order is (r, z, n).
 */
#ifndef TSM_NN_H
#define TSM_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C[M,N] = alpha * A[M,K] . B[N,K]^T + beta * C (+ bias[N]); A, B bf16 with K
 * contiguous (K, lda, ldb multiples of 8, 16-byte aligned), C fp32. */
int tsm_nn_gemm_nt(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb, float *C, int ldc,
                   float alpha, float beta, const float *bias, void *stream);
/* Embedding search without the distance matrix: v = alpha * A[M,K] . B[N,K]^T
 * + bias[N]; every v < tau[row] is appended to row's candidate list (key[row,
 * pos], idx[row, pos] for pos = atomic count < cap).  cnt[M] must be zeroed
 * by the caller; it ends at the number of values below tau (possibly > cap). */
int tsm_nn_gemm_nt_filter(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb, float alpha,
                          const float *bias, const float *tau, int cap, uint32_t *cnt, float *key, int32_t *idx,
                          void *stream);
/* C[M,N] += alpha * A[M,K] . B[N,K]^T with K split across blocks (float
 * atomic accumulation, so the summation order varies run to run): the
 * weight-gradient products, whose outputs are small and K long. */
int tsm_nn_gemm_nt_acc(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb, float *C,
                       int ldc, float alpha, void *stream);
/* The C smallest candidates of each row (sorted ascending; ties by column):
 * out[rows, C].  flag[row] = 1 when the row had fewer than C or more than cap
 * candidates (then its out row holds column 0 in every slot), else 0.
 * cap <= 4096. */
int tsm_nn_select_candidates(const uint32_t *cnt, const float *key, const int32_t *idx, int rows, int cap, int C,
                             int32_t *out, int32_t *flag, void *stream);
/* The same, and the exact re-rank's pairs of each slot s = row * C + r:
 * pa[s] = row_ids[row], pb[s] = col_ids[out[s]] (-1 for a column outside
 * [0, ncol_ids)), so the re-rank needs no separate gather.  The reference's
 * learned query (test_knn_time.py:179-201) stops at the embedding search's
 * ids; these pairs are those ids' exact-distance pairs for the re-rank. */
int tsm_nn_select_candidates_pairs(const uint32_t *cnt, const float *key, const int32_t *idx, int rows, int cap,
                                   int C, int32_t *out, int32_t *flag, const int32_t *row_ids,
                                   const int32_t *col_ids, int64_t ncol_ids, int32_t *pa, int32_t *pb,
                                   void *stream);
/* FP32 [rows, cols] -> BF16 (transpose != 0: dst[c][r]). */
int tsm_nn_cast_bf16(const float *src, int rows, int cols, int ld_src, uint16_t *dst, int ld_dst, int transpose,
                     void *stream);
/* Trajectories ids[0..n) (NULL: 0..n) of a packed set resampled to T points,
 * centred to [-1, 1]: X [n, T, 2]; lenf[n] = log(len) - 4. */
int tsm_nn_resample(const double *xy64, const int64_t *offsets, const int32_t *ids, int n, int T, float *X,
                    float *lenf, void *stream);
/* Conv1d(2 -> C, k = 3, pad 1) + ReLU: Wc [C][2][3], bc [C]; H, Hb [n*T, C]. */
int tsm_nn_conv_fwd(const float *X, int n, int T, int C, const float *Wc, const float *bc, float *H, uint16_t *Hb,
                    void *stream);
int tsm_nn_conv_bwd(const float *X, const float *H, const float *dH, int n, int T, int C, float *dWc, float *dbc,
                    void *stream);
/* h0 = tanh(u * lenf) [n, Hd] (fp32 + bf16) and its backward du[Hd]. */
int tsm_nn_gru_h0(const float *u, const float *lenf, int n, int Hd, float *h, uint16_t *hb, void *stream);
int tsm_nn_gru_h0_bwd(const float *dh0, const float *h0, const float *lenf, int n, int Hd, float *du, void *stream);
/* One GRU step given Gx (row stride ldgx) and Gh = h W_hh^T + b_hh [n, 3Hd]. */
int tsm_nn_gru_step_fwd(const float *Gx, int64_t ldgx, const float *Gh, const float *h, int n, int Hd, float *r,
                        float *z, float *nn, float *ghn, float *hn, uint16_t *hnb, void *stream);
int tsm_nn_gru_step_bwd(const float *dh, const float *h, const float *r, const float *z, const float *nn,
                        const float *ghn, int n, int Hd, float *dGx, int64_t lddgx, float *dGh, uint16_t *dGhb,
                        float *dhp, void *stream);
/* out[c] = sum_r M[r, c] (+ beta * out[c]). */
int tsm_nn_colsum(const float *M, int64_t rows, int cols, int64_t ld, float *out, float beta, void *stream);
/* Pairwise ranking loss over B anchor lists of K candidates sorted by exact
 * distance: positives are ranks < P, negatives ranks >= Q0. */
int tsm_nn_rank_loss(const float *E, int B, int K, int D, int P, int Q0, float tau, float *dE, float *loss,
                     void *stream);
/* Log-space distance regression over B anchor lists of K candidates with
 * exact distances Y [B, K]: mean over each list of (log(d + c) - log(s y +
 * c))^2, d the embedding distance to the anchor, s = sum d / sum y of the
 * list (held constant), c = 1e-3 * mean d; dE includes the 1/B of the mean
 * over lists.  Counterpart of the MSE fit of embedding distances to
 * ground-truth ones in nn/MLP_trajsimi_train.py:78-82 (random-pair labels). */
int tsm_nn_dist_loss(const float *E, const float *Y, int B, int K, int D, float *dE, float *loss, void *stream);
/* Adam step `step` (>= 1) on a flat FP32 parameter vector. */
int tsm_nn_adam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                int64_t step, void *stream);
/* out[n, Dout] = X[n, K] . W[Dout, K]^T + beta * out, FP32. */
int tsm_nn_linear_f32(const float *X, int64_t n, int K, const float *W, int Dout, float *out, float beta,
                      void *stream);
/* BF16 hi/lo split of FP32 rows into [hi, lo] (query_side = 0) or
 * [hi, hi] (query_side = 1), out [n, 2D]; norm[n] (or NULL) = |E_r|^2. */
/* The whole GRU forward over T steps (Gx: [n, T, 3H] input projections; Hs:
 * [T + 1, n, H] with Hs[0] the initial state; Hb: BF16 copy of the current h;
 * Gh: [n, 3H] scratch; r/z/nn/ghn: [T, n, H] when keep, else [1, n, H]):
 * per step the recurrent GEMM and tsm_nn_gru_step_fwd, launched from C.  With
 * keep = 0 and H = 128 one fused kernel runs all T steps instead (bit-identical
 * Hs and Hb; the scratch Gh/r/z/nn/ghn is then not written). */
int tsm_nn_gru_forward(const float *Gx, int T, int n, int Hd, const uint16_t *Whh, const float *bhh, float *Hs,
                       uint16_t *Hb, float *Gh, float *r, float *z, float *nn, float *ghn, int keep, void *stream);
int tsm_nn_split_bf16(const float *E, int64_t n, int D, int query_side, uint16_t *out, float *norm, void *stream);
/* Squared row norms of a bf16 [n, D] matrix. */
int tsm_nn_rownorm_bf16(const uint16_t *E, int64_t n, int D, float *out, void *stream);
/* The k smallest entries of each row (unordered): idx [rows, k], val [rows, k] or NULL. */
int tsm_nn_topk_f32(const float *V, int64_t rows, int64_t cols, int64_t ld, int k, int32_t *idx, float *val,
                    void *stream);
/* The k-th smallest value of each row of V [rows, cols <= 16384] (row stride
 * ld) into out [rows]: the largest of the k smallest, which the threshold
 * search takes as its per-query threshold. */
int tsm_nn_kth_smallest_f32(const float *V, int64_t rows, int64_t cols, int64_t ld, int k, float *out, void *stream);
int tsm_nn_topk_f64(const double *V, int64_t rows, int64_t cols, int64_t ld, int k, int32_t *idx, double *val,
                    void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TSM_NN_H */
