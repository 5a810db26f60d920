// Full -> band reduction (SBR) with ZY two-sided updates, and its back
// transformation.
//
// SBR (the role of my_ZY_ZY_SBR_Vector, api/zy_zy_sy2sb.cu:24-486):
// levels of nb columns; inside a level, panel p (columns c..c+b, rows
// r0 = c+b..n) is factored (panel_qr), then
//   Z_p = A W_p - Y_q (Z_q^T W_p) - Z_q (Y_q^T W_p)      (lazy: A is the
//         level-start trailing matrix, q = the level's earlier panels)
//   Z_p -= 1/2 Y_p (W_p^T Z_p)
// (W_p^T Z_p formed from W_p^T A W_p and the two small products, so the
// three corrections are one product; see sy2sb)
// so that (I - Y W^T) A (I - W Y^T) = A - Y Z^T - Z Y^T, and only the next
// panel's b columns are updated inside the level.  At the level end the
// trailing matrix gets one SYR2K with k = nb on MFMA, mirrored to full
// storage in the epilogue.  The reference keeps two n x n copies (dOriA and
// dA); one matrix suffices because inside a level the lazy column updates
// touch only columns that later panels of the level no longer read.
//
// Back transformation (the SBR part of my_SBRBack_BCBack,
// src/EVD/myEVD.cu:135-245): panels are merged into wide WY blocks by
// doubling (W_b <- W_b - W_a (Y_a^T W_b), batched), then Z <- Q_SBR Z is
// applied block by block from the last block, each as two MFMA GEMMs that
// skip the zero rows above the block.
#include "common.hpp"

namespace evd {

int panel_qr(evd_ctx *ctx, long m, long nc, const float *A, long lda, float *Y, long ldy, float *W, long ldw,
             float *R, long ldr, float *Aout, long ldo, double *tau_out);
int panel_qr_batched(evd_ctx *ctx, int count, long m, long nc, float *A, long lda, long sA, float *Y, long ldy,
                     long sY, float *W, long ldw, long sW);

// GEMM launches carry d problems' products shaped as one problem's (the
// tile and K-split choice, hence the summation order, of a single product)
struct ShapeDiv {
  evd_ctx *c;
  long saved;
  ShapeDiv(evd_ctx *c_, long d) : c(c_), saved(c_->shape_div) { c->shape_div = d; }
  ~ShapeDiv() { c->shape_div = saved; }
};

size_t sy2sb_workspace(long n, long b, long nb) {
  size_t s = 0;
  s += align256((size_t)n * nb * sizeof(float));      // Z
  s += align256((size_t)nb * b * sizeof(float)) * 3;  // small products
  s += (size_t)96 << 20;                              // split-K partials, panel partials
  return s;
}

int sy2sb(evd_ctx *ctx, long n, long b, long nb, float *A, long lda, float *W, long ldw, float *Y, long ldy,
          long wcols) {
  if (n <= 0) return 0;
  EVD_TRY(fill_f32(ctx, W, n, wcols, ldw, 0.f));
  EVD_TRY(fill_f32(ctx, Y, n, wcols, ldy, 0.f));
  ArenaFrame frame(ctx);
  const long ldz = n;
  float *Z = arena_alloc_n<float>(ctx, (size_t)n * nb);
  float *t = arena_alloc_n<float>(ctx, (size_t)2 * nb * b);  // [Z_q^T W_p; Y_q^T W_p]
  float *t3 = arena_alloc_n<float>(ctx, (size_t)b * b);
  if (!Z || !t || !t3) return EVD_E_NOMEM;
  const long P = nb / b;
  long c0 = 0;
  while (n - c0 - b > 1) {
    long p = 0;
    for (;; p++) {
      const long c = c0 + p * b, r0 = c + b, m = n - r0;
      float *Ap = A + r0 + c * lda;
      float *Wp = W + r0 + c * ldw;
      float *Yp = Y + r0 + c * ldy;
      float *Zp = Z + r0 + p * b * ldz;
      EVD_TRY(panel_qr(ctx, m, b, Ap, lda, Yp, ldy, Wp, ldw, nullptr, 0, Ap, lda, nullptr));
      // Z_p = A[r0:, r0:] W_p  (A here is the level-start trailing matrix)
      EVD_TRY(sgemm(ctx, 'N', 'N', m, b, m, 1.f, A + r0 + r0 * lda, lda, Wp, ldw, 0.f, Zp, ldz));
      if (p > 0) {
        // Z_p -= Y_q t_z + Z_q t_y + Y_p t3 / 2 with t_z = Z_q^T W_p,
        // t_y = Y_q^T W_p and t3 = W_p^T Z_p (after the first two terms).
        // [Z_q Z_p]^T W_p and [Y_q Y_p]^T W_p run as one launch (a batch of
        // two, shaped as one product; Z_p and Y_p follow Z_q and Y_q in
        // memory) and give t_z, t_y and X1 = Z_p^T W_p (Z_p = A W_p so far).
        // t3 = W_p^T A' W_p is symmetric in exact arithmetic (A' the updated
        // trailing matrix), so t3^T = X1 - t_y^T t_z - t_z^T t_y stands in for
        // it: formed in place of X1 as a short K-concatenated product (K =
        // 2 pb instead of a product over the m rows), [t_z; t3^T / 2] is then
        // the right operand of Y_q t_z + Y_p t3 / 2, and the update is one
        // K-concatenated product
        const long pb = p * b, pbx = pb + b;
        const float *Yq = Y + r0 + c0 * ldy;
        const float *Zq = Z + r0;
        float *tz = t, *ty = t + pbx * b;
        if (ldy == ldz) {
          ShapeDiv pair(ctx, 2);
          const long dYZ = (long)(reinterpret_cast<intptr_t>(Yq) - reinterpret_cast<intptr_t>(Zq)) / (long)sizeof(float);
          EVD_TRY(sgemm(ctx, 'T', 'N', pbx, b, m, 1.f, Zq, ldz, Wp, ldw, 0.f, tz, pbx, 2, dYZ, 0, pbx * b));
        } else {
          EVD_TRY(sgemm(ctx, 'T', 'N', pbx, b, m, 1.f, Zq, ldz, Wp, ldw, 0.f, tz, pbx));
          EVD_TRY(sgemm(ctx, 'T', 'N', pbx, b, m, 1.f, Yq, ldy, Wp, ldw, 0.f, ty, pbx));
        }
        EVD_TRY(sgemm_kcat(ctx, 'T', 'N', b, b, pb, pb, -0.5f, ty, pbx, tz, pbx, tz, pbx, ty, pbx, 0.5f, tz + pb, pbx));
        EVD_TRY(sgemm_kcat(ctx, 'N', 'N', m, b, pbx, pb, -1.f, Yq, ldy, tz, pbx, Zq, ldz, ty, pbx, 1.f, Zp, ldz));
      } else {
        EVD_TRY(sgemm(ctx, 'T', 'N', b, b, m, 1.f, Wp, ldw, Zp, ldz, 0.f, t3, b));
        EVD_TRY(sgemm(ctx, 'N', 'N', m, b, b, -0.5f, Yp, ldy, t3, b, 1.f, Zp, ldz));
      }
      const long kk = (p + 1) * b;
      const float *Ya = Y + r0 + c0 * ldy;
      const float *Za = Z + r0;
      const bool more = (p + 1 < P) && (n - (r0 + b) > 1);
      if (more) {
        // lazy update of the next panel's b columns inside the level:
        // A_n -= Y_a Z_a^T + Z_a Y_a^T as one K-concatenated product
        float *An = A + r0 + r0 * lda;
        EVD_TRY(sgemm_kcat(ctx, 'N', 'T', m, b, kk, kk, -1.f, Ya, ldy, Za, ldz, Za, ldz, Ya, ldy, 1.f, An, lda));
        continue;
      }
      // level end: two-sided update of the whole trailing matrix
      EVD_TRY(syr2k_lower_mirror(ctx, m, kk, -1.f, Ya, ldy, Za, ldz, 1.f, A + r0 + r0 * lda, lda));
      break;
    }
    c0 += (p + 1) * b;
  }
  return 0;
}

// Same-size SBR of `count` problems (A, W, Y of problem q at q times the
// strides): the panel loop of sy2sb with every kernel launched once for all
// problems.  GEMM shapes and K splits are chosen as for one problem
// (ctx->shape_div), so each problem's result equals its sy2sb bit for bit.
// Split-K partial buffer of one problem's largest product in the batched SBR
// and its back transformation (floats): outputs are at most n x max(n, 1024);
// a K split multiplies the output by at most min(k / 128, 512 / tiles + 1)
// with tiles of at most 128 x 128 (512 = the split target), i.e. by at most
// n / 128 + 1 chunks, and adds at most 512 * 128 * 128 words beyond one copy.
size_t gemm_partial_words(long n) {
  const size_t out = (size_t)n * std::max<long>(n, 1024);
  return std::min((size_t)512 * 128 * 128 + out, (size_t)(n / 128 + 1) * out);
}

size_t sy2sb_batched_workspace(int count, long n, long b, long nb) {
  size_t per = align256((size_t)n * nb * sizeof(float)) + align256((size_t)nb * b * sizeof(float)) * 3;
  size_t part = align256(gemm_partial_words(n) * sizeof(float));
  return (per + part) * (size_t)count + ((size_t)8 << 20);
}

int sy2sb_batched(evd_ctx *ctx, int count, long n, long b, long nb, float *A, long lda, long sA, float *W, long ldw,
                  long sW, float *Y, long ldy, long sY, long wcols) {
  if (n <= 0 || count <= 0) return 0;
  if (count == 1) return sy2sb(ctx, n, b, nb, A, lda, W, ldw, Y, ldy, wcols);
  for (int q = 0; q < count; q++) {
    EVD_TRY(fill_f32(ctx, W + q * sW, n, wcols, ldw, 0.f));
    EVD_TRY(fill_f32(ctx, Y + q * sY, n, wcols, ldy, 0.f));
  }
  ArenaFrame frame(ctx);
  ShapeDiv shape(ctx, count);
  const long ldz = n, sZ = n * nb, sT = 2 * nb * b, sT3 = b * b;
  float *Z = arena_alloc_n<float>(ctx, (size_t)sZ * count);
  float *t = arena_alloc_n<float>(ctx, (size_t)sT * count);
  float *t3 = arena_alloc_n<float>(ctx, (size_t)sT3 * count);
  if (!Z || !t || !t3) return EVD_E_NOMEM;
  const long P = nb / b;
  long c0 = 0;
  while (n - c0 - b > 1) {
    long p = 0;
    for (;; p++) {
      const long c = c0 + p * b, r0 = c + b, m = n - r0;
      float *Ap = A + r0 + c * lda;
      float *Wp = W + r0 + c * ldw;
      float *Yp = Y + r0 + c * ldy;
      float *Zp = Z + r0 + p * b * ldz;
      EVD_TRY(panel_qr_batched(ctx, count, m, b, Ap, lda, sA, Yp, ldy, sY, Wp, ldw, sW));
      EVD_TRY(sgemm(ctx, 'N', 'N', m, b, m, 1.f, A + r0 + r0 * lda, lda, Wp, ldw, 0.f, Zp, ldz, count, sA, sW, sZ));
      if (p > 0) {
        const long pb = p * b;
        const float *Yq = Y + r0 + c0 * ldy;
        const float *Zq = Z + r0;
        const long pbx = pb + b;
        float *tz = t, *ty = t + pbx * b;
        EVD_TRY(sgemm(ctx, 'T', 'N', pbx, b, m, 1.f, Zq, ldz, Wp, ldw, 0.f, tz, pbx, count, sZ, sW, sT));
        EVD_TRY(sgemm(ctx, 'T', 'N', pbx, b, m, 1.f, Yq, ldy, Wp, ldw, 0.f, ty, pbx, count, sY, sW, sT));
        EVD_TRY(sgemm_kcat(ctx, 'T', 'N', b, b, pb, pb, -0.5f, ty, pbx, tz, pbx, tz, pbx, ty, pbx, 0.5f, tz + pb, pbx,
                           count, sT, sT, sT, sT, sT));
        EVD_TRY(sgemm_kcat(ctx, 'N', 'N', m, b, pbx, pb, -1.f, Yq, ldy, tz, pbx, Zq, ldz, ty, pbx, 1.f, Zp, ldz, count,
                           sY, sT, sZ, sT, sZ));
      } else {
        EVD_TRY(sgemm(ctx, 'T', 'N', b, b, m, 1.f, Wp, ldw, Zp, ldz, 0.f, t3, b, count, sW, sZ, sT3));
        EVD_TRY(sgemm(ctx, 'N', 'N', m, b, b, -0.5f, Yp, ldy, t3, b, 1.f, Zp, ldz, count, sY, sT3, sZ));
      }
      const long kk = (p + 1) * b;
      const float *Ya = Y + r0 + c0 * ldy;
      const float *Za = Z + r0;
      const bool more = (p + 1 < P) && (n - (r0 + b) > 1);
      if (more) {
        float *An = A + r0 + r0 * lda;
        EVD_TRY(sgemm_kcat(ctx, 'N', 'T', m, b, kk, kk, -1.f, Ya, ldy, Za, ldz, Za, ldz, Ya, ldy, 1.f, An, lda, count,
                           sY, sZ, sZ, sY, sA));
        continue;
      }
      EVD_TRY(syr2k_lower_mirror_batched(ctx, count, m, kk, -1.f, Ya, ldy, sY, Za, ldz, sZ, 1.f, A + r0 + r0 * lda,
                                         lda, sA));
      break;
    }
    c0 += (p + 1) * b;
  }
  return 0;
}

// ---------------------------------------------------------------- WY merge
// Doubling merge of the panel WY pairs inside each blk-wide block (batched
// over the blocks' pairs).  W and Y must have `wcols` (a multiple of blk)
// zero-padded columns.
int wy_merge_doubling(evd_ctx *ctx, long n, long b, long blk, long wcols, float *W, long ldw, const float *Y,
                      long ldy) {
  ArenaFrame frame(ctx);
  float *T = arena_alloc_n<float>(ctx, (size_t)(wcols / 2) * blk);
  if (!T) return EVD_E_NOMEM;
  for (long cw = b; cw < blk; cw *= 2) {
    const long npairs = wcols / (2 * cw);
    if (npairs <= 0) break;
    // T_q = Y_a^T W_b  (cw x cw), full height (rows above a block are zero)
    EVD_TRY(sgemm(ctx, 'T', 'N', cw, cw, n, 1.f, Y, ldy, W + cw * ldw, ldw, 0.f, T, cw, npairs, 2 * cw * ldy,
                  2 * cw * ldw, cw * cw));
    // W_b -= W_a T_q
    EVD_TRY(sgemm(ctx, 'N', 'N', n, cw, cw, -1.f, W, ldw, T, cw, 1.f, W + cw * ldw, ldw, npairs, 2 * cw * ldw, cw * cw,
                  2 * cw * ldw));
  }
  return 0;
}

// Z <- Q_SBR Z with Q_SBR = prod over the merged blocks (first block leftmost).
int sbr_apply(evd_ctx *ctx, long n, long b, long blk, long npanel_cols, long ncols, float *Z, long ldz,
              const float *W, long ldw, const float *Y, long ldy) {
  ArenaFrame frame(ctx);
  float *T1 = arena_alloc_n<float>(ctx, (size_t)blk * ncols);
  if (!T1) return EVD_E_NOMEM;
  const long nblk = (npanel_cols + blk - 1) / blk;
  for (long k = nblk - 1; k >= 0; k--) {
    const long B0 = k * blk;
    const long bw = std::min(blk, npanel_cols - B0);
    const long r0 = B0 + b;
    if (r0 >= n) continue;
    const long m = n - r0;
    EVD_TRY(sgemm(ctx, 'T', 'N', bw, ncols, m, 1.f, Y + r0 + B0 * ldy, ldy, Z + r0, ldz, 0.f, T1, bw));
    EVD_TRY(sgemm(ctx, 'N', 'N', m, ncols, bw, -1.f, W + r0 + B0 * ldw, ldw, T1, bw, 1.f, Z + r0, ldz));
  }
  return 0;
}

// WY block width of the SBR back transformation: the knob (0 = 1024), made
// b * 2^k (so the doubling merge lands exactly on it) and shrunk until it
// divides the zero-padded column count wcols (a multiple of 1024, so any
// b * 2^k <= 1024 divides it: a knob of 2048 with wcols an odd multiple of
// 1024 becomes 1024).
long sbr_back_blk(long knob, long b, long wcols) {
  long blk = knob > 0 ? knob : 1024;
  while (blk > b && blk > wcols) blk /= 2;
  if (blk < b) blk = b;
  long bb = b;
  while (bb * 2 <= blk) bb *= 2;
  while (bb > b && wcols % bb != 0) bb /= 2;
  return bb;
}

int sbr_back(evd_ctx *ctx, long n, long b, long ncols, float *Z, long ldz, float *W, long ldw, const float *Y,
             long ldy, long wcols) {
  const long blk = sbr_back_blk(ctx->tune.sbr_back_block, b, wcols);
  // a block of b (no merge; b not dividing the padding, e.g. b = 24) is applied panel by panel
  if (blk != b && wcols % blk != 0)
    return set_error(ctx, EVD_E_UNSUPPORTED, "W/Y column padding %ld not a multiple of %ld", wcols, blk);
  EVD_TRY(wy_merge_doubling(ctx, n, b, blk, wcols, W, ldw, Y, ldy));
  return sbr_apply(ctx, n, b, blk, wcols, ncols, Z, ldz, W, ldw, Y, ldy);
}

// Same-size back transformation of `count` problems: W, Y of problem q at
// q (ldw wcols) / q (ldy wcols) (contiguous, so the pairwise merges of all
// problems form one strided batch), Z at q sZ.  Shapes as for one problem.
int sbr_back_batched(evd_ctx *ctx, int count, long n, long b, long ncols, float *Z, long ldz, long sZ, float *W,
                     long ldw, const float *Y, long ldy, long wcols) {
  if (count == 1) return sbr_back(ctx, n, b, ncols, Z, ldz, W, ldw, Y, ldy, wcols);
  const long blk = sbr_back_blk(ctx->tune.sbr_back_block, b, wcols);
  if (blk != b && wcols % blk != 0) return set_error(ctx, EVD_E_UNSUPPORTED, "W/Y column padding %ld not a multiple of %ld", wcols, blk);
  ArenaFrame frame(ctx);
  ShapeDiv shape(ctx, count);
  const long sW = ldw * wcols, sY = ldy * wcols;
  // merges: the pairs of every problem are one batch (pair p of problem q is
  // pair q * npairs + p at stride 2 cw ld)
  float *T = arena_alloc_n<float>(ctx, (size_t)(wcols / 2) * blk * count);
  if (!T) return EVD_E_NOMEM;
  for (long cw = b; cw < blk; cw *= 2) {
    const long npairs = wcols / (2 * cw);
    if (npairs <= 0) break;
    const long nb_all = npairs * count;
    EVD_TRY(sgemm(ctx, 'T', 'N', cw, cw, n, 1.f, Y, ldy, W + cw * ldw, ldw, 0.f, T, cw, nb_all, 2 * cw * ldy,
                  2 * cw * ldw, cw * cw));
    EVD_TRY(sgemm(ctx, 'N', 'N', n, cw, cw, -1.f, W, ldw, T, cw, 1.f, W + cw * ldw, ldw, nb_all, 2 * cw * ldw, cw * cw,
                  2 * cw * ldw));
  }
  // Z <- Q_SBR Z block by block from the last block
  const long sT1 = blk * ncols;
  float *T1 = arena_alloc_n<float>(ctx, (size_t)sT1 * count);
  if (!T1) return EVD_E_NOMEM;
  const long nblk = (wcols + blk - 1) / blk;
  for (long k = nblk - 1; k >= 0; k--) {
    const long B0 = k * blk;
    const long bw = std::min(blk, wcols - B0);
    const long r0 = B0 + b;
    if (r0 >= n) continue;
    const long m = n - r0;
    EVD_TRY(sgemm(ctx, 'T', 'N', bw, ncols, m, 1.f, Y + r0 + B0 * ldy, ldy, Z + r0, ldz, 0.f, T1, bw, count, sY, sZ,
                  sT1));
    EVD_TRY(sgemm(ctx, 'N', 'N', m, ncols, bw, -1.f, W + r0 + B0 * ldw, ldw, T1, bw, 1.f, Z + r0, ldz, count, sW, sT1,
                  sZ));
  }
  return 0;
}

}  // namespace evd

using namespace evd;

extern "C" int evd_syr2k_f32(evd_ctx *ctx, long n, long k, float alpha, const float *A, long lda, const float *B,
                             long ldb, float beta, float *C, long ldc) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, k >= 0, 3);
  EVD_ARG(ctx, lda >= std::max<long>(1, n), 6);
  EVD_ARG(ctx, ldb >= std::max<long>(1, n), 8);
  EVD_ARG(ctx, ldc >= std::max<long>(1, n), 11);
  return syr2k_lower_mirror(ctx, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int evd_sy2sb_f32(evd_ctx *ctx, long n, long b, long nb, float *A, long lda, float *W, long ldw, float *Y,
                             long ldy) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, b >= 1 && b <= 64, 3);
  EVD_ARG(ctx, nb >= b && nb % b == 0, 4);
  EVD_ARG(ctx, lda >= std::max<long>(1, n), 6);
  EVD_ARG(ctx, ldw >= std::max<long>(1, n), 8);
  EVD_ARG(ctx, ldy >= std::max<long>(1, n), 10);
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, sy2sb_workspace(n, b, nb)));
  return sy2sb(ctx, n, b, nb, A, lda, W, ldw, Y, ldy, n);
}

// computerWYFromSlide semantics: sequential left-to-right merge of slide-wide blocks.
extern "C" int evd_wy_from_slide_f32(evd_ctx *ctx, long M, long N, long slide, float *W, long ldw, const float *Y,
                                     long ldy) {
  if (!ctx) return -1;
  EVD_ARG(ctx, M >= 0, 2);
  EVD_ARG(ctx, N >= 0, 3);
  EVD_ARG(ctx, slide >= 1, 4);
  EVD_ARG(ctx, ldw >= std::max<long>(1, M), 6);
  EVD_ARG(ctx, ldy >= std::max<long>(1, M), 8);
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, (size_t)N * slide * sizeof(float) + ((size_t)64 << 20)));
  ArenaFrame frame(ctx);
  float *work = arena_alloc_n<float>(ctx, (size_t)std::max<long>(N, 1) * slide);
  if (!work) return EVD_E_NOMEM;
  for (long i = 2 * slide; i <= N; i += slide) {
    const long w = slide;
    EVD_TRY(sgemm(ctx, 'T', 'N', i - w, w, M, 1.f, Y, ldy, W + (i - w) * ldw, ldw, 0.f, work, i - w));
    EVD_TRY(sgemm(ctx, 'N', 'N', M, w, i - w, -1.f, W, ldw, work, i - w, 1.f, W + (i - w) * ldw, ldw));
  }
  return 0;
}

__global__ void add_identity_kernel(long m, float *Q, long ldq) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < m; i += (long)gridDim.x * blockDim.x)
    Q[i + i * ldq] += 1.f;
}

extern "C" int evd_q_from_wy_f32(evd_ctx *ctx, long M, long N, float *Q, long ldq, const float *W, long ldw,
                                 const float *Y, long ldy) {
  if (!ctx) return -1;
  EVD_ARG(ctx, M >= 0, 2);
  EVD_ARG(ctx, N >= 0, 3);
  EVD_ARG(ctx, ldq >= std::max<long>(1, M), 5);
  EVD_ARG(ctx, ldw >= std::max<long>(1, M), 7);
  EVD_ARG(ctx, ldy >= std::max<long>(1, M), 9);
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, (size_t)64 << 20));
  EVD_TRY(sgemm(ctx, 'N', 'T', M, M, N, -1.f, W, ldw, Y, ldy, 0.f, Q, ldq));
  if (M > 0) {
    add_identity_kernel<<<(unsigned)std::min<long>((M + 255) / 256, 1024), 256, 0, ctx->stream>>>(M, Q, ldq);
    EVD_LAUNCHED(ctx);
  }
  return 0;
}

extern "C" int evd_sbr_back_f32(evd_ctx *ctx, long n, long b, long ncols, float *Z, long ldz, float *W, long ldw,
                                const float *Y, long ldy) {
  if (!ctx) return -1;
  EVD_ARG(ctx, n >= 0, 2);
  EVD_ARG(ctx, b >= 1 && b <= 64, 3);
  EVD_ARG(ctx, ncols >= 0, 4);
  EVD_ARG(ctx, ldz >= std::max<long>(1, n), 6);
  EVD_ARG(ctx, ldw >= std::max<long>(1, n), 8);
  EVD_ARG(ctx, ldy >= std::max<long>(1, n), 10);
  if (ctx->arena_top == 0)
    EVD_TRY(arena_ensure(ctx, (size_t)1024 * (ncols + n) * sizeof(float) + ((size_t)128 << 20)));
  // the caller's W/Y have n columns: merge only whole blocks of b * 2^k that fit
  long blk = ctx->tune.sbr_back_block > 0 ? ctx->tune.sbr_back_block : 1024;
  long bb = b;
  while (bb * 2 <= blk && bb * 2 <= n) bb *= 2;
  while (bb > b && n % bb != 0) bb /= 2;
  if (n % bb != 0) bb = b;
  if (n % bb != 0) {
    // ragged: apply panel by panel
    return sbr_apply(ctx, n, b, b, n, ncols, Z, ldz, W, ldw, Y, ldy);
  }
  EVD_TRY(wy_merge_doubling(ctx, n, b, bb, n, W, ldw, Y, ldy));
  return sbr_apply(ctx, n, b, bb, n, ncols, Z, ldz, W, ldw, Y, ldy);
}

extern "C" int evd_orgqr_f32(evd_ctx *ctx, long m, long n, float *Q, long ldq, const float *W, long ldw, float *work) {
  if (!ctx) return -1;
  EVD_ARG(ctx, m >= 0, 2);
  EVD_ARG(ctx, n >= 0 && n <= m, 3);
  EVD_ARG(ctx, ldq >= std::max<long>(1, m), 5);
  EVD_ARG(ctx, ldw >= std::max<long>(1, m), 7);
  EVD_ARG(ctx, work != nullptr || n == 0, 8);
  if (n == 0) return 0;
  if (ctx->arena_top == 0) EVD_TRY(arena_ensure(ctx, (size_t)64 << 20));
  EVD_TRY(copy_f32(ctx, Q, ldq, work, n, n, n));
  EVD_TRY(sgemm(ctx, 'N', 'T', m, n, n, -1.f, W, ldw, work, n, 0.f, Q, ldq));
  add_identity_kernel<<<(unsigned)std::min<long>((n + 255) / 256, 1024), 256, 0, ctx->stream>>>(n, Q, ldq);
  EVD_LAUNCHED(ctx);
  return 0;
}
