/*
 * oracle.c -- CPU restatement of the sparse-DNN inference path, used ONLY as a
 * checker (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline leg).
 * Nothing in the product library links, loads or calls this file.
 *
 * What it restates (reference paths relative to the CoSpDNN source tree):
 *   - layer rule  y' = min(max(x.W + b, 0), 32)         kernels/kernel.cuh:132-136,
 *                                                        kernels/kernel.cu:368-376,
 *                                                        utils/cpu_spmm_fuse.h:46,74
 *     with a float bias (the reference engine takes `float bias`, cospdnn.cu:8;
 *     the CPU checker's `int bias` at cpu_spmm_fuse.h:11 is a truncation bug).
 *   - W is stored by column: column j = output neuron j, rows = input neurons
 *     (COO "T" branch, utils/matrix.h:282-289; Y.W convention of cospdnn_single).
 *   - liveness of a sample = any activation > 0 (kernels/kernel.cu:129-139 uses
 *     > 1e-5; the Graph Challenge definition, used here, is > 0).
 *   - live-row pruning per layer (cospdnn.cu:240-276): dead rows are all-zero and
 *     stay all-zero when b <= 0, so they are skipped.
 *   - the synthetic generator (splitmix64 counter streams) that defines the
 *     benchmark inputs; the product library carries its own implementation and
 *     tests check the two agree bit for bit.
 *
 * Two forwards:
 *   orc_forward_fp64   FP64 oracle. Short-circuits decided rows: all-zero rows
 *                      (exact fixed point when b <= 0) and all-32 rows when the
 *                      layer maps an all-32 input to all-32 exactly.
 *   orc_forward_fp32   "repo-style" FP32 CPU forward: per-sample CSC gather,
 *                      clamp, dead-row pruning only (no saturated-row skip), as
 *                      cpu_spmm_fuse.h:19-75 does per layer.  Timed as the CPU
 *                      baseline.
 *
 * Pinning: tests/test_oracle.py checks both forwards against brute-force dense
 * numpy FP64 on tiny nets, against closed forms (first layer from binary input,
 * banded class sums, fixed points), against the permutation invariance of the
 * category set, and against the committed golden fixtures in tests/golden/.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

/* ------------------------------------------------------------------------- */
/* synthetic generator (restated; product copy lives in cospdnn_amd/csrc)    */
/* ------------------------------------------------------------------------- */
#define ORC_GOLDEN 0x9E3779B97F4A7C15ULL
enum { ORC_TAG_W = 1, ORC_TAG_P = 2, ORC_TAG_Y = 3 };
enum { ORC_GEN_RANDOM = 0, ORC_GEN_BANDED = 1 };

static inline uint64_t orc_fmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
/* n-th output of a splitmix64 generator whose state starts at `key` */
static inline uint64_t orc_draw(uint64_t key, uint64_t n) {
    return orc_fmix64(key + (n + 1) * ORC_GOLDEN);
}
static inline uint64_t orc_key(uint64_t seed, uint64_t tag, uint64_t a, uint64_t b) {
    return orc_fmix64(orc_fmix64(orc_fmix64(seed * ORC_GOLDEN + tag) + a) + b);
}
/* uniform integer in [0, n), n < 2^32: multiply-shift of the top 32 bits */
static inline uint32_t orc_below(uint64_t u, uint32_t n) {
    return (uint32_t)(((u >> 32) * (uint64_t)n) >> 32);
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Rows of column j of layer l, K distinct uniform rows (Floyd's sampling). */
void orc_gen_random_col(uint64_t seed, int N, int K, int l, int j, int32_t *rows) {
    uint64_t key = orc_key(seed, ORC_TAG_W, (uint64_t)l, (uint64_t)j);
    int cnt = 0;
    uint64_t t = 0;
    for (int i = N - K; i < N; ++i) {
        int32_t r = (int32_t)orc_below(orc_draw(key, t++), (uint32_t)(i + 1));
        int seen = 0;
        for (int q = 0; q < cnt; ++q)
            if (rows[q] == r) { seen = 1; break; }
        rows[cnt++] = seen ? i : r;
    }
    qsort(rows, (size_t)K, sizeof(int32_t), cmp_i32);
}

/* Seeded Fisher-Yates permutation of [0, N) for layer l (banded layers). */
void orc_gen_perm(uint64_t seed, int N, int l, int32_t *perm) {
    uint64_t key = orc_key(seed, ORC_TAG_P, (uint64_t)l, 0);
    for (int i = 0; i < N; ++i) perm[i] = i;
    uint64_t t = 0;
    for (int i = N - 1; i >= 1; --i) {
        int r = (int)orc_below(orc_draw(key, t++), (uint32_t)(i + 1));
        int32_t tmp = perm[i]; perm[i] = perm[r]; perm[r] = tmp;
    }
}

/* Whole layer in CSC form: colptr[N+1], rows[N*K] sorted per column, vals. */
void orc_gen_layer(uint64_t seed, int kind, int N, int K, int l, float wval,
                   int32_t *colptr, int32_t *rows, float *vals) {
    if (kind == ORC_GEN_BANDED) {
        int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
        orc_gen_perm(seed, N, l, perm);
        int s = N / K;
        for (int j = 0; j < N; ++j) {
            int32_t *r = rows + (size_t)j * K;
            for (int k = 0; k < K; ++k) r[k] = (int32_t)(((int64_t)perm[j] + (int64_t)k * s) % N);
            qsort(r, (size_t)K, sizeof(int32_t), cmp_i32);
        }
        free(perm);
    } else {
#pragma omp parallel for schedule(static)
        for (int j = 0; j < N; ++j) orc_gen_random_col(seed, N, K, l, j, rows + (size_t)j * K);
    }
    for (int j = 0; j <= N; ++j) colptr[j] = j * K;
    for (size_t q = 0; q < (size_t)N * K; ++q) vals[q] = wval;
}

/* Per-row density threshold (fixed point, /2^32) and the row's binary input. */
static inline uint32_t orc_row_thr(uint64_t key) {
    return 644245094u + (uint32_t)(((orc_draw(key, 0) >> 32) * 1288490189ULL) >> 32);
}
void orc_gen_y0_row(uint64_t seed, int N, int64_t s, uint8_t *out) {
    uint64_t key = orc_key(seed, ORC_TAG_Y, (uint64_t)s, 0);
    uint32_t thr = orc_row_thr(key);
    for (int n = 0; n < N; ++n) out[n] = (uint8_t)((orc_draw(key, (uint64_t)n + 1) >> 32) < thr);
}
uint32_t orc_y0_row_thr(uint64_t seed, int64_t s) {
    return orc_row_thr(orc_key(seed, ORC_TAG_Y, (uint64_t)s, 0));
}

/* ------------------------------------------------------------------------- */
/* network description                                                       */
/* ------------------------------------------------------------------------- */
typedef struct {
    int N, L;
    float bias;
    /* generated weights: gen_kind >= 0; explicit CSC: gen_kind = -1 */
    int gen_kind;
    int K;            /* nonzeros per column for generated layers */
    uint64_t seed;
    float wval;
    const int32_t *const *colptr; /* [L][N+1] */
    const int32_t *const *rows;   /* [L][nnz_l] */
    const float *const *vals;     /* [L][nnz_l] */
    /* input: y0_gen != 0 -> rows generated from seed; else dense uint8 [S][N] */
    int y0_gen;
    const uint8_t *y0;
} orc_net;

typedef struct {
    int32_t *cp;
    int32_t *rw;
    float *vl;
    int ready;
    double colsum_min; /* min_j sum of column j */
    int nonneg;        /* every weight >= 0 */
} orc_layer_cache;

static void layer_get(const orc_net *net, orc_layer_cache *c, int l,
                      const int32_t **cp, const int32_t **rw, const float **vl) {
    if (net->gen_kind < 0) {
        *cp = net->colptr[l]; *rw = net->rows[l]; *vl = net->vals[l];
    } else {
        if (!c[l].ready) {
            size_t nnz = (size_t)net->N * net->K;
            c[l].cp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(net->N + 1));
            c[l].rw = (int32_t *)malloc(sizeof(int32_t) * nnz);
            c[l].vl = (float *)malloc(sizeof(float) * nnz);
            orc_gen_layer(net->seed, net->gen_kind, net->N, net->K, l, net->wval, c[l].cp, c[l].rw, c[l].vl);
            c[l].ready = 1;
        }
        *cp = c[l].cp; *rw = c[l].rw; *vl = c[l].vl;
    }
    if (c[l].ready < 2) {
        /* saturation test data: y_in all 32 -> z_j = 32 * colsum_j (exact in FP64) */
        double mn = INFINITY;
        int nonneg = 1;
        for (int j = 0; j < net->N; ++j) {
            double s = 0;
            for (int32_t q = (*cp)[j]; q < (*cp)[j + 1]; ++q) {
                s += (double)(*vl)[q];
                if ((*vl)[q] < 0) nonneg = 0;
            }
            if (s < mn) mn = s;
        }
        c[l].colsum_min = mn;
        c[l].nonneg = nonneg;
        c[l].ready = 2;
    }
}

static void cache_free(orc_layer_cache *c, int L, int owned) {
    for (int l = 0; l < L; ++l)
        if (owned && c[l].ready) { free(c[l].cp); free(c[l].rw); free(c[l].vl); }
    free(c);
}

static void load_y0_row(const orc_net *net, int64_t s, uint8_t *tmp) {
    if (net->y0_gen) orc_gen_y0_row(net->seed, net->N, s, tmp);
    else memcpy(tmp, net->y0 + (size_t)s * net->N, (size_t)net->N);
}

/* ------------------------------------------------------------------------- */
/* FP64 oracle forward                                                        */
/* ------------------------------------------------------------------------- */
/* state per row: 0 = undecided, 1 = dead (all 0), 2 = saturated (all 32) */
enum { ROW_OPEN = 0, ROW_DEAD = 1, ROW_SAT = 2 };

/*
 * Runs samples [s0, s0+S) through all L layers.
 *   alive[S]        : 1 if the sample has any activation > 0 after layer L-1.
 *   live[L+1]       : (accumulated) number of live samples entering each layer,
 *                     live[L] = number of categories.  May be NULL.
 *   final_ids/final : optional: for the `nfinal` sample offsets in final_ids
 *                     (relative to s0, ascending), the final FP64 activations
 *                     [nfinal][N].  May be NULL.
 * Returns 0 on success.
 */
int orc_forward_fp64(const orc_net *net, int64_t s0, int64_t S, uint8_t *alive, int64_t *live,
                     const int64_t *final_ids, int64_t nfinal, double *final_out, int nthreads) {
    const int N = net->N, L = net->L;
    const double b = (double)net->bias;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    orc_layer_cache *cache = (orc_layer_cache *)calloc((size_t)L, sizeof(orc_layer_cache));
    int64_t chunk = (int64_t)((1u << 28) / ((size_t)N * 8 * 2));
    if (chunk < 64) chunk = 64;
    if (chunk > 4096) chunk = 4096;
    if (live) memset(live, 0, sizeof(int64_t) * (size_t)(L + 1));
    int64_t fpos = 0;
    double *ya = (double *)malloc(sizeof(double) * (size_t)chunk * N);
    double *yb = (double *)malloc(sizeof(double) * (size_t)chunk * N);
    uint8_t *state = (uint8_t *)malloc((size_t)chunk);
    for (int64_t c0 = 0; c0 < S; c0 += chunk) {
        int64_t cn = S - c0 < chunk ? S - c0 : chunk;
        int64_t n_open = 0, n_sat = 0;
#pragma omp parallel
        {
            uint8_t *tmp = (uint8_t *)malloc((size_t)N);
#pragma omp for schedule(static) reduction(+ : n_open)
            for (int64_t r = 0; r < cn; ++r) {
                load_y0_row(net, s0 + c0 + r, tmp);
                int any = 0;
                for (int n = 0; n < N; ++n) { ya[(size_t)r * N + n] = tmp[n]; any |= tmp[n]; }
                /* an all-zero input row stays all-zero only when b <= 0 */
                int open = any || b > 0.0;
                state[r] = open ? ROW_OPEN : ROW_DEAD;
                n_open += open;
            }
            free(tmp);
        }
        if (live) live[0] += cn; /* every sample enters layer 0 */
        for (int l = 0; l < L; ++l) {
            if (n_open == 0 && n_sat == 0) continue;
            const int32_t *cp, *rw;
            const float *vl;
            layer_get(net, cache, l, &cp, &rw, &vl);
            int sat_ok = cache[l].nonneg && (32.0 * cache[l].colsum_min + b >= 32.0);
            int64_t n_live = 0;
            int64_t opened = 0, sat_now = 0;
            if (!sat_ok && n_sat > 0) {
                /* saturated rows are no longer a fixed point: reopen them */
                for (int64_t r = 0; r < cn; ++r)
                    if (state[r] == ROW_SAT) {
                        for (int n = 0; n < N; ++n) ya[(size_t)r * N + n] = 32.0;
                        state[r] = ROW_OPEN;
                        opened++;
                    }
                n_sat -= opened;
                n_open += opened;
            }
            if (n_open == 0) {
                if (live) live[l + 1] += n_sat;
                continue;
            }
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : n_live, sat_now)
            for (int64_t r = 0; r < cn; ++r) {
                if (state[r] != ROW_OPEN) continue;
                const double *yi = ya + (size_t)r * N;
                double *yo = yb + (size_t)r * N;
                int any = 0, all32 = 1;
                for (int j = 0; j < N; ++j) {
                    double acc = 0.0;
                    for (int32_t q = cp[j]; q < cp[j + 1]; ++q) acc += yi[rw[q]] * (double)vl[q];
                    double z = acc + b;
                    double y = z < 0.0 ? 0.0 : (z > 32.0 ? 32.0 : z);
                    yo[j] = y;
                    any |= (y > 0.0);
                    all32 &= (y == 32.0);
                }
                if (!any && b <= 0.0) state[r] = ROW_DEAD;
                else if (all32 && sat_ok) { state[r] = ROW_SAT; sat_now++; n_live++; }
                else n_live++;
                memcpy((double *)yi, yo, sizeof(double) * (size_t)N);
            }
            /* rows reported live entering layer l+1: open-and-live + already saturated */
            n_open = n_live - sat_now;
            n_sat += sat_now;
            if (live) live[l + 1] += n_live - sat_now + n_sat;
        }
        for (int64_t r = 0; r < cn; ++r) {
            int64_t id = c0 + r;
            if (state[r] == ROW_OPEN) {
                int any = 0;
                for (int n = 0; n < N; ++n) any |= (ya[(size_t)r * N + n] > 0.0);
                alive[id] = (uint8_t)any;
            } else {
                alive[id] = (uint8_t)(state[r] == ROW_SAT);
            }
            if (final_out && fpos < nfinal && final_ids[fpos] == id) {
                double *dst = final_out + (size_t)fpos * N;
                if (state[r] == ROW_OPEN) memcpy(dst, ya + (size_t)r * N, sizeof(double) * (size_t)N);
                else for (int n = 0; n < N; ++n) dst[n] = state[r] == ROW_SAT ? 32.0 : 0.0;
                fpos++;
            }
        }
    }
    free(ya); free(yb); free(state);
    cache_free(cache, L, net->gen_kind >= 0);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* "repo-style" FP32 CPU forward (the CPU baseline)                           */
/* ------------------------------------------------------------------------- */
/*
 * Layer-major over the sample set, as the reference engine runs (cospdnn.cu:
 * 172-307): every layer computes all live samples with a per-output CSC gather
 * in FP32 (cpu_spmm_fuse.h:52-75), then prunes the samples with no activation
 * > 0 (cospdnn.cu:249-261).  Saturated samples are NOT skipped.
 * alive[S], live[L+1] as in orc_forward_fp64.
 */
int orc_forward_fp32(const orc_net *net, int64_t s0, int64_t S, uint8_t *alive, int64_t *live, int nthreads) {
    const int N = net->N, L = net->L;
    const float b = net->bias;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    orc_layer_cache *cache = (orc_layer_cache *)calloc((size_t)L, sizeof(orc_layer_cache));
    int64_t chunk = (int64_t)((1u << 28) / ((size_t)N * 4 * 2));
    if (chunk < 64) chunk = 64;
    if (chunk > 8192) chunk = 8192;
    if (live) memset(live, 0, sizeof(int64_t) * (size_t)(L + 1));
    float *ya = (float *)malloc(sizeof(float) * (size_t)chunk * N);
    float *yb = (float *)malloc(sizeof(float) * (size_t)chunk * N);
    int64_t *list = (int64_t *)malloc(sizeof(int64_t) * (size_t)chunk);
    uint8_t *lv = (uint8_t *)malloc((size_t)chunk);
    for (int64_t c0 = 0; c0 < S; c0 += chunk) {
        int64_t cn = S - c0 < chunk ? S - c0 : chunk;
#pragma omp parallel
        {
            uint8_t *tmp = (uint8_t *)malloc((size_t)N);
#pragma omp for schedule(static)
            for (int64_t r = 0; r < cn; ++r) {
                load_y0_row(net, s0 + c0 + r, tmp);
                int any = 0;
                for (int n = 0; n < N; ++n) { ya[(size_t)r * N + n] = (float)tmp[n]; any |= tmp[n]; }
                lv[r] = (uint8_t)(any || b > 0.0f);
            }
            free(tmp);
        }
        int64_t nl = 0;
        for (int64_t r = 0; r < cn; ++r) if (lv[r]) list[nl++] = r;
        if (live) live[0] += cn;
        for (int l = 0; l < L; ++l) {
            if (nl == 0) continue;
            const int32_t *cp, *rw;
            const float *vl;
            if (net->gen_kind < 0) { cp = net->colptr[l]; rw = net->rows[l]; vl = net->vals[l]; }
            else layer_get(net, cache, l, &cp, &rw, &vl);
#pragma omp parallel for schedule(dynamic, 4)
            for (int64_t i = 0; i < nl; ++i) {
                int64_t r = list[i];
                const float *yi = ya + (size_t)r * N;
                float *yo = yb + (size_t)r * N;
                int any = 0;
                for (int j = 0; j < N; ++j) {
                    float acc = 0.0f;
                    for (int32_t q = cp[j]; q < cp[j + 1]; ++q) acc += yi[rw[q]] * vl[q];
                    float z = acc + b;
                    float y = z < 0.0f ? 0.0f : (z > 32.0f ? 32.0f : z);
                    yo[j] = y;
                    any |= (y > 0.0f);
                }
                memcpy((float *)yi, yo, sizeof(float) * (size_t)N);
                lv[r] = (uint8_t)any;
            }
            int64_t k = 0;
            for (int64_t i = 0; i < nl; ++i) if (lv[list[i]]) list[k++] = list[i];
            nl = k;
            if (live) live[l + 1] += nl;
        }
        for (int64_t r = 0; r < cn; ++r) alive[c0 + r] = 0;
        for (int64_t i = 0; i < nl; ++i) alive[c0 + list[i]] = 1;
    }
    free(ya); free(yb); free(list); free(lv);
    cache_free(cache, L, net->gen_kind >= 0);
    return 0;
}

/* max |a - b| between FP64 oracle rows and a caller's float rows [n][N] */
double orc_max_abs_diff(const double *a, const float *bv, int64_t count) {
    double m = 0.0;
#pragma omp parallel for reduction(max : m)
    for (int64_t i = 0; i < count; ++i) {
        double d = fabs(a[i] - (double)bv[i]);
        if (d > m) m = d;
    }
    return m;
}

int orc_num_threads(void) { return omp_get_max_threads(); }
