// Internal FP64 dense linear algebra (linalg.hip).
#pragma once
#include "impl.hpp"

namespace chs {

int zgemm(chs_context* ctx, char opa, char opb, int m, int n, int k, z64 alpha, const z64* A,
          int lda, const z64* B, int ldb, z64 beta, z64* C, int ldc);
// C = A^H B (n x n) for a product the caller knows to be Hermitian: only the
// tiles on or above the diagonal are computed, the lower triangle mirrored.
int zgemm_herm(chs_context* ctx, int n, int k, const z64* A, int lda, const z64* B, int ldb, z64* C,
               int ldc);
int cgemm(chs_context* ctx, char opa, char opb, int m, int n, int k, c32 alpha, const c32* A,
          int lda, const c32* B, int ldb, c32 beta, c32* C, int ldc);
// y = alpha op(A) x + beta y (op N, T or C) with A held in complex64 and x,
// y and the accumulation in complex128.
int zgemv_c32(chs_context* ctx, char opa, int m, int k, z64 alpha, const c32* A, int lda,
              const z64* x, z64 beta, z64* y);
int gemm_opcode(char op);
int hermitise(chs_context* ctx, z64* M, int n);
int cholesky(chs_context* ctx, z64* A, int n, int* info_host);
int tri_inv(chs_context* ctx, const z64* L, z64* X, int n);
int heev_jacobi(chs_context* ctx, z64* A, int n, double* w, z64* U, z64* V, int max_sweeps,
                int* sweeps_out);
int heev_tridiag(chs_context* ctx, z64* A, int n, double* w, z64* U);
int zhegv(chs_context* ctx, z64* A, z64* B, int n, double* w);
// w_out (optional, with projectors and no HY_out): beta^H Y of the rotated
// block (P x n_b), from the pre-rotation coefficients times the Ritz vectors.
// Y32 (optional, same case): Y held exactly in complex64 (the complex64
// filter's output); H|Y> and beta^H Y are then formed in complex64 and
// widened, the Gram matrices, the eigensolve and the rotation stay FP64.
int rayleigh_ritz_impl(chs_context* ctx, z64* Y, int n_b, double* eps_host, z64* HY_out,
                       double* eps_dev_out, z64* w_out = nullptr, const c32* Y32 = nullptr);
int cholqr_impl(chs_context* ctx, z64* Y, int n_b);
// fp16x3 complex GEMM (gemm_h3.hip): split planes of op(A) for a c128
// projector block (P x n_pw, lda; cn: op(A) = A^H, else A), per-column
// exponents of a c64 operand, and C = alpha op(A) B + beta C.
int h3_split_planes(chs_context* ctx, const z64* A, int lda, int P, int n_pw, bool cn,
                    H3Planes& out);
int h3_column_exponents(chs_context* ctx, const c32* B, int k, int n, int ldb, int* d_eb);
int h3_gemm(chs_context* ctx, const H3Planes& A, int n, c32 alpha, const c32* B, int ldb,
            const int* d_eb, c32 beta, c32* C, int ldc);

}  // namespace chs
