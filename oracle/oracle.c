/*
 * CPU oracle for the quantized vector-sparse SpMM / SDDMM path.
 * TEST INFRASTRUCTURE ONLY (checker for tests/, smoke() and bench.py's
 * cpu_baseline leg).  Citations: path:line in the Magicube reference.
 * Build: oracle/Makefile (gcc -O3, optional -fopenmp for orc_spmm_plain_omp).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Generator: std::default_random_engine is std::minstd_rand0 in libstdc++   */
/* (x <- 16807 x mod 2^31-1, seed 1).  uniform_real_distribution<float>(0,1) */
/* draws generate_canonical<float,24>: one engine call (log2 range = 30 bits  */
/* >= 24), ret = float(x - 1) / float(2147483646.0L) where the divisor rounds */
/* to 2^31, clamped below 1.  bm_test_utils.h:130-164.                       */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t x; } orc_engine;

static uint64_t engine_next(orc_engine *g) {
    g->x = (g->x * 16807ULL) % 2147483647ULL;
    return g->x;
}

static float engine_canonical(orc_engine *g) {
    const float r = (float)2147483646.0L;          /* __tmp *= __r  -> 2^31f */
    float sum = (float)(engine_next(g) - 1ULL);     /* __urng() - __urng.min() */
    float ret = sum / r;
    if (ret >= 1.0f) ret = nextafterf(1.0f, 0.0f);
    return ret;                                     /* * (b - a) + a = * 1 + 0 */
}

void orc_canonical_stream(int64_t count, float *out) {
    orc_engine g = {1};
    for (int64_t i = 0; i < count; ++i) out[i] = engine_canonical(&g);
}

void orc_make_dense_matrix(int type_bytes, int64_t count, void *out, int sddmm_variant) {
    orc_engine g = {1};   /* the engine is passed by value: every call restarts */
    for (int64_t i = 0; i < count; ++i) {
        if (type_bytes == 4) {
            /* float temp = 2147483647.0*distribution(generator);  (draw discarded)
               matrix[i] = 0x11111111 + int(117.0*distribution(generator)); */
            (void)engine_canonical(&g);
            float u = engine_canonical(&g);
            ((int32_t *)out)[i] = 0x11111111 + (int32_t)(117.0 * (double)u);
        } else if (type_bytes == 8) {
            if (sddmm_variant) {
                /* SDDMM/SDDMM/include/bm_test_utils.h:150-153 */
                float u = engine_canonical(&g);
                double t = 9223372036854775807.0 * (double)u;
                ((int64_t *)out)[i] = (int64_t)t;
            } else {
                (void)engine_canonical(&g);
                float u = engine_canonical(&g);
                ((int64_t *)out)[i] = 0x1111111111111111LL + (int64_t)(int32_t)(117.0 * (double)u);
            }
        } else if (type_bytes == 2) {
            float u = engine_canonical(&g);
            float t = (float)(10240.0 * (double)u);
            ((int16_t *)out)[i] = (int16_t)t;
        } else if (type_bytes == 1) {
            float u = engine_canonical(&g);
            float t = (float)(127.0 * (double)u);
            ((int8_t *)out)[i] = (int8_t)t;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Vector-CSR alignment: spmm_benchmark.cpp:177-222                          */
/* ------------------------------------------------------------------------ */

int orc_align_row_offsets(int m_vec, const int *row_offsets, int align, int *ap) {
    int aligned_num_item = 0;
    ap[0] = aligned_num_item;
    for (int i = 1; i < m_vec + 1; i++) {
        int num_item = row_offsets[i] - row_offsets[i - 1];
        aligned_num_item += (num_item + align - 1) / align * align;
        if (i != m_vec) ap[i * 2] = aligned_num_item;
        ap[i * 2 - 1] = ap[i * 2 - 2] + num_item;
    }
    return aligned_num_item;
}

void orc_align_col_indices(int m_vec, const int *row_offsets, const int *col_indices,
                           const int *ap, int aligned_num, int *aligned_cols) {
    for (int i = 0; i < aligned_num; i++) aligned_cols[i] = -1;
    for (int i = 1; i < m_vec + 1; i++) {
        int b = row_offsets[i - 1], e = row_offsets[i];
        for (int j = b; j < e; j++) aligned_cols[ap[(i - 1) * 2] + j - b] = col_indices[j];
    }
}

void orc_shuffle_col_indices(int aligned_num, const int *aligned_cols, int *shuffled) {
    for (int i = 0; i < aligned_num; i++) shuffled[i] = -1;
    for (int i = 0; i < aligned_num / 8; i++)
        for (int j = 0; j < 8; j++)
            shuffled[i * 8 + (j % 2) * 4 + j / 2] = aligned_cols[i * 8 + j];
}

/* bm_test_utils.h:222-237 sorts row ids by length, longest first, with
 * std::sort (ties in unspecified order).  Ties are broken here by row id. */
static const int *g_swz_offsets;
static int swz_cmp(const void *pa, const void *pb) {
    int a = *(const int *)pa, b = *(const int *)pb;
    int la = g_swz_offsets[a + 1] - g_swz_offsets[a];
    int lb = g_swz_offsets[b + 1] - g_swz_offsets[b];
    if (la != lb) return la > lb ? -1 : 1;
    return a < b ? -1 : (a > b);
}
void orc_sorted_row_swizzle(int rows, const int *row_offsets, int *row_indices) {
    for (int i = 0; i < rows; i++) row_indices[i] = i;
    g_swz_offsets = row_offsets;
    qsort(row_indices, (size_t)rows, sizeof(int), swz_cmp);
}

/* ------------------------------------------------------------------------ */
/* Value preparation: spmm_benchmark.cpp:238-377, buffer choice :425-428     */
/* ------------------------------------------------------------------------ */

void orc_prepare_values(const void *values, int type_bytes, int scaleA,
                        int m_vec, const int *row_offsets, const int *ap,
                        int aligned_num, int V, int preA, int preA_cut, int preB,
                        void *out) {
    (void)preA;
    const int mma_k = (preA_cut == 4 || preB == 4) ? 32 : 16;
    const size_t word = (size_t)type_bytes;
    const size_t total = (size_t)aligned_num * (size_t)scaleA * word;
    unsigned char *al = (unsigned char *)calloc(total + 16, 1);
    unsigned char *tr = (unsigned char *)calloc(total + 16, 1);
    unsigned char *de = (unsigned char *)calloc(total + 16, 1);
    const unsigned char *src = (const unsigned char *)values;

    /* :246-251 copy into padded slots (TypeA words) */
    for (int i = 1; i < m_vec + 1; i++) {
        long b = (long)row_offsets[i - 1] * scaleA, e = (long)row_offsets[i] * scaleA;
        for (long j = b; j < e; j++)
            memcpy(al + ((long)ap[(i - 1) * 2] * scaleA + j - b) * word, src + j * word, word);
    }
    unsigned short *al_s = (unsigned short *)al, *tr_s = (unsigned short *)tr, *de_s = (unsigned short *)de;
    const long nV = (long)aligned_num * V;

    if (preA_cut == 8) {                                            /* :255-281 */
        for (long i = 0; i < nV; i += (long)mma_k * V)
            for (int j = 0; j < mma_k; j++)
                for (int v = 0; v < V; v++)
                    tr[i + v * mma_k + j] = al[i + (long)j * V + v];
        if (mma_k == 32) {
            const unsigned char mask = 15;
            for (long i = 0; i < nV; i += (long)mma_k * V)
                for (int j = 0; j < mma_k * V; j++) {
                    int o0 = (j % 2) * 4, o1 = ((j + 1) % 2) * 4;
                    de[i + j / 2] |= (unsigned char)((tr[i + j] & mask) << o0);
                    de[i + mma_k * V / 2 + j / 2] |= (unsigned char)((tr[i + j] & (mask << 4)) >> o1);
                }
        }
    } else if ((preA_cut == 12 || preA_cut == 16) && mma_k == 32) {  /* :282-312, :322-356 */
        const int planes = preA_cut / 4;
        for (long i = 0; i < nV; i += (long)mma_k * V)
            for (int j = 0; j < mma_k; j++)
                for (int v = 0; v < V; v++)
                    tr_s[i + v * mma_k + j] = al_s[i + (long)j * V + v];
        const unsigned short mask = 15;
        const long q = (long)mma_k * V / 4;
        for (long i = 0; i < nV; i += (long)mma_k * V)
            for (int j = 0; j < mma_k * V; j++) {
                unsigned t = tr_s[i + j];
                int pos = (j % 4) * 4;    /* nibble slot within the output short */
                for (int p = 0; p < planes; p++) {
                    unsigned nib = (t >> (4 * p)) & mask;
                    de_s[i + p * q + j / 4] |= (unsigned short)(nib << pos);
                }
            }
    } else if (preA_cut == 12 && mma_k == 16) {                     /* :313-321 */
        const unsigned char mask = 15;
        for (long i = 0; i < nV * 2; i += (long)mma_k * V * 2)
            for (int j = 0; j < mma_k; j++)
                for (int v = 0; v < V * 2; v += 2) {
                    de[i + j + (v / 2) * mma_k] = al[i + (long)j * V * 2 + v];
                    de[i + mma_k * V + j + (v / 2) * mma_k] = al[i + (long)j * V * 2 + v + 1] & mask;
                }
    } else if (preA_cut == 16 && mma_k == 16) {                     /* :357-364 */
        for (long i = 0; i < nV * 2; i += (long)mma_k * V * 2)
            for (int j = 0; j < mma_k; j++)
                for (int v = 0; v < V * 2; v += 2) {
                    de[i + j + (v / 2) * mma_k] = al[i + (long)j * V * 2 + v];
                    de[i + mma_k * V + j + (v / 2) * mma_k] = al[i + (long)j * V * 2 + v + 1];
                }
    } else if (preA_cut == 4) {                                     /* :365-377 */
        const unsigned char mask = 15;
        const int hv = V / 2;
        for (long i = 0; i < (long)aligned_num * hv; i += (long)mma_k * hv)
            for (int j = 0; j < mma_k; j++)
                for (int v = 0; v < hv; v++) {
                    int o0 = (j % 2) * 4, o1 = ((j + 1) % 2) * 4;
                    unsigned char a = al[i + (long)j * hv + v];
                    tr[i + mma_k * v + j / 2] |= (unsigned char)((a & mask) << o0);
                    tr[i + mma_k * v + mma_k / 2 + j / 2] |= (unsigned char)((a & (mask << 4)) >> o1);
                }
    }
    const int use_decomposed = (preA_cut > preB) || (preA_cut == 16 && preB == 16);
    memcpy(out, use_decomposed ? de : tr, total);
    free(al); free(tr); free(de);
}

/* ------------------------------------------------------------------------ */
/* compute_ref_integers, spmm_benchmark.cpp:69-120, with C++ integral        */
/* promotion of TypeA emulated: masks are TypeA values (sign-extended on    */
/* promotion), shifts and ands happen in int (TypeA <= 4 bytes) or          */
/* long long, right shifts are arithmetic.  B lanes use int arithmetic, so  */
/* the top lane of each int32 sign-extends (:106).                          */
/* ------------------------------------------------------------------------ */

static int64_t sext_bits(uint64_t v, int bits) {
    if (bits >= 64) return (int64_t)v;
    uint64_t m = 1ULL << (bits - 1);
    v &= (bits == 64) ? ~0ULL : ((1ULL << bits) - 1);
    return (int64_t)((v ^ m) - m);
}

static int64_t load_typeA(const void *A, int type_bytes, long idx) {
    switch (type_bytes) {
    case 1: return ((const int8_t *)A)[idx];
    case 2: return ((const int16_t *)A)[idx];
    case 4: return ((const int32_t *)A)[idx];
    default: return ((const int64_t *)A)[idx];
    }
}

static int32_t ref_extract_a(int64_t tile, int type_bytes, int preA, int preA_cut, int shift) {
    const int tb = type_bytes * 8;
    int64_t maskA = sext_bits((1ULL << preA) - 1, tb);          /* (TypeA)(power2n(preA)-1) */
    int64_t maskC = sext_bits((1ULL << preA_cut) - 1, tb);
    if (type_bytes <= 4) {   /* promoted to int */
        int32_t m = (int32_t)((uint32_t)(int32_t)maskA << shift);
        int32_t t = (int32_t)tile;
        int32_t r = (m & t) >> shift;
        r &= (int32_t)maskC;
        return r;
    } else {
        int64_t m = (int64_t)((uint64_t)maskA << shift);
        int64_t r = (m & tile) >> shift;
        r &= maskC;
        return (int32_t)r;
    }
}

static int32_t ref_extract_int(int32_t tile, int pre, int shift) {
    int32_t m = (int32_t)((uint32_t)((1u << pre) - 1u) << shift);
    return (m & tile) >> shift;
}

double orc_spmm_ref_check(const void *A, int type_bytes, const int32_t *B, int32_t *C,
                          int M, int K, int N, int preA, int preA_cut, int preB, int V,
                          const int *row_offsets, const int *col_indices, int m_vec,
                          int scaleA) {
    (void)K;
    for (long i = 0; i < (long)M * N; i++) C[i] = 0;
    double flops = 0;
    const int b_tile = 32 / preB;
    for (int i = 0; i < m_vec; i++) {
        for (int j = row_offsets[i]; j < row_offsets[i + 1]; j++) {
            int col_idx = col_indices[j];
            for (int tt = 0; tt < scaleA; tt++) {
                int64_t tile = load_typeA(A, type_bytes, (long)j * scaleA + tt);
                for (int av = 0; av < V / scaleA; av++) {
                    int row_idx = i * V + av + tt * V / scaleA;
                    int a_val = ref_extract_a(tile, type_bytes, preA, preA_cut, av * preA);
                    for (int n = 0; n < N / b_tile; n++) {
                        int32_t Bt = B[(long)col_idx * (N / b_tile) + n];
                        for (int bv = 0; bv < b_tile; bv++) {
                            int b_val = ref_extract_int(Bt, preB, bv * preB);
                            long o = (long)row_idx * N + n * b_tile + bv;
                            C[o] = (int32_t)((uint32_t)C[o] + (uint32_t)(a_val * b_val));
                            flops += 2.0;
                        }
                    }
                }
            }
        }
    }
    return flops;
}

/* Host_sddmm_integers, sddmm_benchmark.cpp:24-60 */
double orc_sddmm_ref_check(const int32_t *lhs, const int32_t *rhs, int32_t *out,
                           int M, int K, int N, int preA, int preB, int V,
                           const int *row_offsets, const int *col_indices, int m_vec,
                           int alignment) {
    (void)M; (void)N;
    const int a_tiles = 32 / preA, b_tiles = 32 / preB;
    double flops = 0;
    for (int i = 0; i < m_vec; i++) {
        for (int j = row_offsets[i * 2]; j < row_offsets[i * 2 + 1]; j++) {
            for (int v = 0; v < V; v++) {
                uint32_t acc = 0;
                long idx_m = (long)i * V + v;
                long idx_n = col_indices[j];
                for (int l = 0; l < K; l += a_tiles) {
                    int32_t at = lhs[idx_m * K / a_tiles + l / a_tiles];
                    int32_t bt = rhs[idx_n * K / b_tiles + l / b_tiles];
                    for (int t = 0; t < a_tiles; t++) {
                        int a = ref_extract_int(at, preA, t * preA);
                        int b = ref_extract_int(bt, preB, t * preB);
                        acc += (uint32_t)(a * b);
                        flops += 2.0;
                    }
                }
                out[(long)(j / alignment) * alignment * V + alignment * v + j % alignment] = (int32_t)acc;
            }
        }
    }
    return flops;
}

/* ------------------------------------------------------------------------ */
/* Plain restatements on unpacked integers.                                  */
/* ------------------------------------------------------------------------ */

void orc_unpack_lanes(const void *packed, int bits, int64_t count, int is_signed, int32_t *out) {
    const unsigned char *p = (const unsigned char *)packed;
    for (int64_t i = 0; i < count; i++) {
        int64_t bit = i * bits;
        uint32_t v = 0;
        for (int b = 0; b < bits; b++) {
            int64_t q = bit + b;
            v |= (uint32_t)((p[q >> 3] >> (q & 7)) & 1u) << b;
        }
        if (is_signed && bits < 32 && (v >> (bits - 1)) & 1u) v |= ~((1u << bits) - 1u);
        out[i] = (int32_t)v;
    }
}

void orc_pack_lanes(const int32_t *vals, int bits, int64_t count, void *packed) {
    unsigned char *p = (unsigned char *)packed;
    int64_t nbytes = (count * bits + 7) / 8;
    memset(p, 0, (size_t)nbytes);
    for (int64_t i = 0; i < count; i++) {
        uint32_t v = (uint32_t)vals[i];
        for (int b = 0; b < bits; b++) {
            int64_t q = i * bits + b;
            p[q >> 3] |= (unsigned char)(((v >> b) & 1u) << (q & 7));
        }
    }
}

void orc_spmm_plain(int m_vec, int V, int N, const int *row_offsets, const int *col_indices,
                    const int32_t *a, const int32_t *b, int32_t *C) {
    for (long i = 0; i < (long)m_vec * V * N; i++) C[i] = 0;
    for (int i = 0; i < m_vec; i++)
        for (int j = row_offsets[i]; j < row_offsets[i + 1]; j++) {
            const int32_t *brow = b + (long)col_indices[j] * N;
            for (int v = 0; v < V; v++) {
                uint32_t av = (uint32_t)a[(long)j * V + v];
                uint32_t *crow = (uint32_t *)C + ((long)i * V + v) * N;
                for (int n = 0; n < N; n++) crow[n] += av * (uint32_t)brow[n];
            }
        }
}

void orc_sddmm_plain(int m_vec, int V, int K, const int *ap, const int *ac,
                     const int32_t *lhs, const int32_t *rhs, int32_t *out, int alignment) {
    for (int i = 0; i < m_vec; i++)
        for (int j = ap[2 * i]; j < ap[2 * i + 1]; j++)
            for (int v = 0; v < V; v++) {
                const int32_t *arow = lhs + ((long)i * V + v) * K;
                const int32_t *brow = rhs + (long)ac[j] * K;
                uint32_t acc = 0;
                for (int k = 0; k < K; k++) acc += (uint32_t)arow[k] * (uint32_t)brow[k];
                out[(long)(j / alignment) * alignment * V + alignment * v + j % alignment] = (int32_t)acc;
            }
}
