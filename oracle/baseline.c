/*
 * CPU baseline: a C restatement of mojoBLAS's own CPU algorithms for the
 * level-3 hot path, timed beside the GPU (the reference itself is Mojo and
 * its fp32 kernel emits Apple AMX opcodes, so it cannot run on an x86 host).
 *
 * TEST/BENCH INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's
 * cpu_baseline leg; never linked into the product library.
 *
 *   base_sgemm  dispatch as src/level3/gemm.mojo:104-124:
 *     f32 N,N  -> gemm_v11 blocking (gemm.mojo:2068-2435): TK=1024 k-panels,
 *                 B panel packed once per k-panel in 16-column groups
 *                 (:2230-2240), row blocks of MC=256 (64 when n<=1024) run in
 *                 parallel (:2429-2430), A row block packed with alpha folded
 *                 (:2251-2260), 64x16 C blocks = 4 16x16 tiles (:2286-2322);
 *                 the AMX fma32 16x16 outer product becomes a 16x16 SIMD
 *                 outer-product micro-kernel (16 vector accumulators).
 *     others   -> _gemm_naive (gemm.mojo:127-377): column-parallel (n>=64)
 *                 axpy sweep over l, skipping B(l,j)==0, strided for TN/TT.
 *   base_strsm_left  trsm.mojo:96-168 (side L), serial column loop,
 *                    zero-pivot skip.
 *   base_ssyrk_n     syrk.mojo:89-166 (trans N), column-parallel axpy.
 *   base_ssymm       symm.mojo:88-239, column-parallel fused axpy + dot
 *                    (side L) or axpy sweep (side R), zero skip on B(l,j).
 *   base_strmm       trmm.mojo:90-303, every side / uplo / trans branch
 *                    with the reference's threading (only L,T parallel).
 *
 * Threads: OpenMP, the analogue of Mojo `parallelize`; base_threads() reports
 * how many the runtime uses.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TILE 16
#define NZ 4
#define TK 1024
#define MC 256
#define SMALL_MC 64
#define ROW_PAR_THRESHOLD 1024
#define PAR_THRESHOLD 64

typedef float v16 __attribute__((vector_size(64)));

int base_threads(void) { return omp_get_max_threads(); }

static int is_n(char c) { return c == 'N' || c == 'n'; }

/* Z(16x16, column-major: zc[col]) += A_tile(16 rows x klen) (x) B_group(klen x 16).
 * a_tile[l*16 + r], b_grp[l*16 + col]. */
__attribute__((target_clones("avx512f", "avx2", "default")))
static void micro_16x16(const float* restrict a_tile, const float* restrict b_grp, int64_t klen,
                        float* restrict z /* 16x16, z[col*16+r] */) {
    v16 acc[TILE];
    for (int j = 0; j < TILE; ++j) acc[j] = (v16){0};
    for (int64_t l = 0; l < klen; ++l) {
        v16 av;
        memcpy(&av, a_tile + l * TILE, sizeof(v16));
        const float* bl = b_grp + l * TILE;
        for (int j = 0; j < TILE; ++j) acc[j] += av * bl[j];
    }
    for (int j = 0; j < TILE; ++j) memcpy(z + j * TILE, &acc[j], sizeof(v16));
}

static void sgemm_v11(int64_t m, int64_t n, int64_t k, float alpha, const float* a, int64_t lda,
                      const float* b, int64_t ldb, float beta, float* c, int64_t ldc) {
    if (alpha == 0.0f || k == 0) {
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) c[i + j * ldc] = (beta == 0.0f) ? 0.0f : beta * c[i + j * ldc];
        return;
    }
    int direct_first = (beta == 0.0f) && n <= ROW_PAR_THRESHOLD;
    if (beta == 0.0f && !direct_first) {
        for (int64_t j = 0; j < n; ++j) memset(c + j * ldc, 0, sizeof(float) * (size_t)m);
    } else if (beta != 0.0f && beta != 1.0f) {
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) c[i + j * ldc] *= beta;
    }
    int64_t n_groups = (n + TILE - 1) / TILE;
    int64_t mc_stride = (n <= ROW_PAR_THRESHOLD) ? SMALL_MC : MC;
    int64_t mc_groups = (m + mc_stride - 1) / mc_stride;
    for (int64_t k0 = 0; k0 < k; k0 += TK) {
        int64_t klen = (k - k0 < TK) ? (k - k0) : TK;
        int first_panel = (k0 == 0) && direct_first;
        float* b_panel = (float*)aligned_alloc(64, sizeof(float) * (size_t)(n_groups * klen * TILE));
        float* a_panels = (float*)aligned_alloc(64, sizeof(float) * (size_t)(mc_groups * mc_stride * klen));
        /* Serial B pack (gemm.mojo:2230-2240). */
        for (int64_t g = 0; g < n_groups; ++g) {
            int64_t j0 = g * TILE, jlen = (n - j0 < TILE) ? n - j0 : TILE;
            float* dst = b_panel + g * klen * TILE;
            for (int64_t l = 0; l < klen; ++l) {
                for (int64_t r = 0; r < jlen; ++r) dst[l * TILE + r] = b[(k0 + l) + (j0 + r) * ldb];
                for (int64_t r = jlen; r < TILE; ++r) dst[l * TILE + r] = 0.0f;
            }
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t blk = 0; blk < mc_groups; ++blk) {
            int64_t mc0 = blk * mc_stride;
            int64_t mc_len = (m - mc0 < mc_stride) ? m - mc0 : mc_stride;
            float* a_pack = a_panels + blk * mc_stride * klen;
            for (int64_t t = 0; t < mc_len; t += TILE) {
                int64_t rc = (mc_len - t < TILE) ? mc_len - t : TILE;
                float* dst = a_pack + t * klen;
                for (int64_t l = 0; l < klen; ++l) {
                    for (int64_t r = 0; r < rc; ++r) dst[l * TILE + r] = alpha * a[mc0 + t + r + (k0 + l) * lda];
                    for (int64_t r = rc; r < TILE; ++r) dst[l * TILE + r] = 0.0f;
                }
            }
            float z[TILE * TILE];
            for (int64_t g = 0; g < n_groups; ++g) {
                int64_t j0 = g * TILE, jlen = (n - j0 < TILE) ? n - j0 : TILE;
                const float* bg = b_panel + g * klen * TILE;
                for (int64_t t = 0; t < mc_len; t += TILE) {
                    int64_t rc = (mc_len - t < TILE) ? mc_len - t : TILE;
                    micro_16x16(a_pack + t * klen, bg, klen, z);
                    for (int64_t jj = 0; jj < jlen; ++jj) {
                        float* cc = c + (mc0 + t) + (j0 + jj) * ldc;
                        if (first_panel)
                            for (int64_t r = 0; r < rc; ++r) cc[r] = z[jj * TILE + r];
                        else
                            for (int64_t r = 0; r < rc; ++r) cc[r] += z[jj * TILE + r];
                    }
                }
            }
        }
        free(b_panel);
        free(a_panels);
    }
}

static void sgemm_naive(char ta, char tb, int64_t m, int64_t n, int64_t k, float alpha, const float* a,
                        int64_t lda, const float* b, int64_t ldb, float beta, float* c, int64_t ldc) {
    int nta = is_n(ta), ntb = is_n(tb);
#pragma omp parallel for schedule(static) if (n >= PAR_THRESHOLD)
    for (int64_t j = 0; j < n; ++j) {
        float* cj = c + j * ldc;
        if (beta == 0.0f)
            for (int64_t i = 0; i < m; ++i) cj[i] = 0.0f;
        else if (beta != 1.0f)
            for (int64_t i = 0; i < m; ++i) cj[i] *= beta;
        for (int64_t l = 0; l < k; ++l) {
            float bv = ntb ? b[l + j * ldb] : b[j + l * ldb];
            if (bv == 0.0f) continue;
            float temp = alpha * bv;
            if (nta) {
                const float* al = a + l * lda;
                for (int64_t i = 0; i < m; ++i) cj[i] += temp * al[i];
            } else {
                const float* al = a + l;
                for (int64_t i = 0; i < m; ++i) cj[i] += temp * al[i * lda];
            }
        }
    }
}

void base_sgemm(char ta, char tb, int64_t m, int64_t n, int64_t k, float alpha, const float* a, int64_t lda,
                const float* b, int64_t ldb, float beta, float* c, int64_t ldc) {
    if (m == 0 || n == 0) return;
    if (is_n(ta) && is_n(tb))
        sgemm_v11(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
    else if (alpha == 0.0f || k == 0) {
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) c[i + j * ldc] = (beta == 0.0f) ? 0.0f : beta * c[i + j * ldc];
    } else
        sgemm_naive(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

/* trsm.mojo:96-168 for side L (the config-3 case is L,L,N,N): alpha pre-scale,
 * then per column forward/backward substitution with a zero-pivot skip; no
 * threading (the reference's trsm has none). */
void base_strsm_left(char uplo, char trans, char diag, int64_t m, int64_t n, float alpha, const float* a,
                     int64_t lda, float* b, int64_t ldb) {
    int upper = (uplo == 'U' || uplo == 'u'), notr = is_n(trans), nounit = is_n(diag);
    if (m == 0 || n == 0) return;
    if (alpha != 1.0f)
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) b[i + j * ldb] *= alpha;
    for (int64_t j = 0; j < n; ++j) {
        float* bj = b + j * ldb;
        if (notr && upper) {
            for (int64_t l = m - 1; l >= 0; --l) {
                if (bj[l] == 0.0f) continue;
                if (nounit) bj[l] /= a[l + l * lda];
                float p = bj[l];
                const float* al = a + l * lda;
                for (int64_t i = 0; i < l; ++i) bj[i] -= p * al[i];
            }
        } else if (notr) {
            for (int64_t l = 0; l < m; ++l) {
                if (bj[l] == 0.0f) continue;
                if (nounit) bj[l] /= a[l + l * lda];
                float p = bj[l];
                const float* al = a + l * lda;
                for (int64_t i = l + 1; i < m; ++i) bj[i] -= p * al[i];
            }
        } else if (upper) {
            for (int64_t i = 0; i < m; ++i) {
                for (int64_t l = 0; l < i; ++l) bj[i] -= a[l + i * lda] * bj[l];
                if (nounit) bj[i] /= a[i + i * lda];
            }
        } else {
            for (int64_t i = m - 1; i >= 0; --i) {
                for (int64_t l = i + 1; l < m; ++l) bj[i] -= a[l + i * lda] * bj[l];
                if (nounit) bj[i] /= a[i + i * lda];
            }
        }
    }
}

/* syrk.mojo:89-166, trans N: triangle beta-scale, then column-parallel axpy
 * with the zero skip. */
void base_ssyrk_n(char uplo, int64_t n, int64_t k, float alpha, const float* a, int64_t lda, float beta,
                  float* c, int64_t ldc) {
    int upper = (uplo == 'U' || uplo == 'u');
    if (n == 0) return;
    for (int64_t j = 0; j < n; ++j) {
        int64_t i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
        for (int64_t i = i0; i < i1; ++i) c[i + j * ldc] = (beta == 0.0f) ? 0.0f : beta * c[i + j * ldc];
    }
    if (alpha == 0.0f) return;
#pragma omp parallel for schedule(dynamic, 16) if (n >= PAR_THRESHOLD)
    for (int64_t j = 0; j < n; ++j) {
        float* cj = c + j * ldc;
        int64_t i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
        for (int64_t l = 0; l < k; ++l) {
            float av = a[j + l * lda];
            if (av == 0.0f) continue;
            float temp = alpha * av;
            const float* al = a + l * lda;
            for (int64_t i = i0; i < i1; ++i) cj[i] += temp * al[i];
        }
    }
}

/* symm.mojo:88-239.  Side L: per column j (parallel when n >= 64), C(:,j)
 * scaled by beta, then for each l with B(l,j) != 0 a fused axpy of
 * alpha*B(l,j)*A(:,l) over the stored part of column l and a dot product
 * temp2 = A(:,l).B(:,j) over the same rows, finished on the diagonal
 * (upper: l from m-1 down, rows i < l; lower: l upward, rows i > l).  The
 * reference skips the whole step, temp2 included, when B(l,j) == 0
 * (symm.mojo:114,133); restated as written.  Side R: per column j,
 * C(:,j) += alpha*A(j,j)*B(:,j) + sum_{k != j} alpha*A(k,j)*B(:,k), with
 * A(k,j) read from the stored triangle. */
void base_ssymm(char side, char uplo, int64_t m, int64_t n, float alpha, const float* a, int64_t lda,
                const float* b, int64_t ldb, float beta, float* c, int64_t ldc) {
    int left = (side == 'L' || side == 'l'), upper = (uplo == 'U' || uplo == 'u');
    if (m == 0 || n == 0 || (alpha == 0.0f && beta == 1.0f)) return;
#pragma omp parallel for schedule(static) if (n >= PAR_THRESHOLD)
    for (int64_t j = 0; j < n; ++j) {
        float* cj = c + j * ldc;
        const float* bj = b + j * ldb;
        if (beta == 0.0f)
            for (int64_t i = 0; i < m; ++i) cj[i] = 0.0f;
        else if (beta != 1.0f)
            for (int64_t i = 0; i < m; ++i) cj[i] *= beta;
        if (left) {
            if (upper) {
                for (int64_t l = m - 1; l >= 0; --l) {
                    if (bj[l] == 0.0f) continue;
                    float t1 = alpha * bj[l], t2 = 0.0f;
                    const float* al = a + l * lda;
                    for (int64_t i = 0; i < l; ++i) {
                        cj[i] += t1 * al[i];
                        t2 += al[i] * bj[i];
                    }
                    cj[l] = cj[l] + t1 * al[l] + alpha * t2;
                }
            } else {
                for (int64_t l = 0; l < m; ++l) {
                    if (bj[l] == 0.0f) continue;
                    float t1 = alpha * bj[l], t2 = 0.0f;
                    const float* al = a + l * lda;
                    for (int64_t i = l + 1; i < m; ++i) {
                        cj[i] += t1 * al[i];
                        t2 += al[i] * bj[i];
                    }
                    cj[l] = cj[l] + t1 * al[l] + alpha * t2;
                }
            }
        } else {
            float td = alpha * a[j + j * lda];
            for (int64_t i = 0; i < m; ++i) cj[i] += td * bj[i];
            for (int64_t k = 0; k < n; ++k) {
                if (k == j) continue;
                /* A(k,j) of the symmetric matrix from the stored triangle. */
                int stored_kj = upper ? (k < j) : (k > j);
                float tk = alpha * (stored_kj ? a[k + j * lda] : a[j + k * lda]);
                const float* bk = b + k * ldb;
                for (int64_t i = 0; i < m; ++i) cj[i] += tk * bk[i];
            }
        }
    }
}

static void trmm_scale_col(float* x, int64_t m, float t) {
    for (int64_t i = 0; i < m; ++i) x[i] *= t;
}
static void trmm_axpy_col(float* y, const float* x, int64_t m, float t) {
    for (int64_t i = 0; i < m; ++i) y[i] += t * x[i];
}

/* trmm.mojo:90-303: B := alpha*op(A)*B (side L) or alpha*B*op(A) (side R),
 * loop for loop, with the reference's zero skips (B(k,j) == 0 on the left,
 * A(k,j) == 0 on the right) and its threading: only the left transposed
 * branches run columns in parallel (trmm.mojo:149-185); the others are
 * serial, as the reference leaves them. */
void base_strmm(char side, char uplo, char trans, char diag, int64_t m, int64_t n, float alpha, const float* a,
                int64_t lda, float* b, int64_t ldb) {
    int left = (side == 'L' || side == 'l'), upper = (uplo == 'U' || uplo == 'u');
    int notr = is_n(trans), nounit = is_n(diag);
    if (m == 0 || n == 0) return;
    if (alpha == 0.0f) {
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) b[i + j * ldb] = 0.0f;
        return;
    }
    if (left && notr) {
        for (int64_t j = 0; j < n; ++j) {
            float* bj = b + j * ldb;
            if (upper) {
                for (int64_t k = 0; k < m; ++k) {
                    if (bj[k] == 0.0f) continue;
                    float t = alpha * bj[k];
                    trmm_axpy_col(bj, a + k * lda, k, t);
                    if (nounit) t *= a[k + k * lda];
                    bj[k] = t;
                }
            } else {
                for (int64_t k = m - 1; k >= 0; --k) {
                    if (bj[k] == 0.0f) continue;
                    float t = alpha * bj[k];
                    bj[k] = nounit ? t * a[k + k * lda] : t;
                    trmm_axpy_col(bj + k + 1, a + k * lda + k + 1, m - k - 1, t);
                }
            }
        }
    } else if (left) {
#pragma omp parallel for schedule(static) if (n >= PAR_THRESHOLD)
        for (int64_t j = 0; j < n; ++j) {
            float* bj = b + j * ldb;
            if (upper) {
                for (int64_t i = m - 1; i >= 0; --i) {
                    float t = bj[i];
                    if (nounit) t *= a[i + i * lda];
                    for (int64_t kk = 0; kk < i; ++kk) t += a[kk + i * lda] * bj[kk];
                    bj[i] = alpha * t;
                }
            } else {
                for (int64_t i = 0; i < m; ++i) {
                    float t = bj[i];
                    if (nounit) t *= a[i + i * lda];
                    for (int64_t kk = i + 1; kk < m; ++kk) t += a[kk + i * lda] * bj[kk];
                    bj[i] = alpha * t;
                }
            }
        }
    } else if (notr) {
        if (upper) {
            for (int64_t j = n - 1; j >= 0; --j) {
                float* bj = b + j * ldb;
                trmm_scale_col(bj, m, nounit ? alpha * a[j + j * lda] : alpha);
                for (int64_t kk = 0; kk < j; ++kk)
                    if (a[kk + j * lda] != 0.0f) trmm_axpy_col(bj, b + kk * ldb, m, alpha * a[kk + j * lda]);
            }
        } else {
            for (int64_t j = 0; j < n; ++j) {
                float* bj = b + j * ldb;
                trmm_scale_col(bj, m, nounit ? alpha * a[j + j * lda] : alpha);
                for (int64_t kk = j + 1; kk < n; ++kk)
                    if (a[kk + j * lda] != 0.0f) trmm_axpy_col(bj, b + kk * ldb, m, alpha * a[kk + j * lda]);
            }
        }
    } else {
        if (upper) {
            for (int64_t k = 0; k < n; ++k) {
                float* bk = b + k * ldb;
                for (int64_t j = 0; j < k; ++j)
                    if (a[j + k * lda] != 0.0f) trmm_axpy_col(b + j * ldb, bk, m, alpha * a[j + k * lda]);
                float t = nounit ? alpha * a[k + k * lda] : alpha;
                if (t != 1.0f) trmm_scale_col(bk, m, t);
            }
        } else {
            for (int64_t k = n - 1; k >= 0; --k) {
                float* bk = b + k * ldb;
                for (int64_t j = k + 1; j < n; ++j)
                    if (a[j + k * lda] != 0.0f) trmm_axpy_col(b + j * ldb, bk, m, alpha * a[j + k * lda]);
                float t = nounit ? alpha * a[k + k * lda] : alpha;
                if (t != 1.0f) trmm_scale_col(bk, m, t);
            }
        }
    }
}
