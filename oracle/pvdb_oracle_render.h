/*
 * Merged-VDB renderer restatement (MGRenderer) -- TEST INFRASTRUCTURE ONLY.
 *
 * Follows plenvdb/lib/vdb/renderer.cu: get_rays (:122-167), get_tminmax
 * (:50-64), the counting march first_look_merged (:222-268) and the gathering
 * march ray_marching_merged (:312-367).  The MLP and the final composite
 * (cuda_rgbnet :83-109, final_render :112-119) are done in numpy by
 * oracle/oracle.py.
 *
 * The merged index grid is given densely: idx[x][y][z] (int32, 0 = inactive /
 * background, otherwise 1 + rank of the voxel, plenvdb/vdb_compression.py:29-33).
 * A voxel is active iff its index is non-zero (copyFromArray activates values
 * different from the background 0).  Coordinates outside [0, reso) read 0.
 *
 * The FP32 variant (_f32) keeps the reference's two passes, including the
 * pass-1 rewrite of tmin/tmax (quirk: pass 2 restarts from t_first - step,
 * renderer.cu:263) and the different density summation order of the two
 * passes (trigetDensity :191-220 vs trigetDensity2 :271-300).  Pass 2 writes
 * at most as many samples as pass 1 counted (the reference would write past
 * the ray's segment, which is undefined).  The FP64 variant (_f64) keeps the
 * reference's FP32 geometry (ray set-up, t accumulation, sample positions, box
 * and active tests, corner indices and fractions) and does the rest in double,
 * so both variants see the same samples.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

#define RENDER_IMPL(REAL, GREAL, SFX)                                                                   \
    static inline int CAT(ridx, SFX)(const int32_t* idx, const int* reso, int i, int j, int k) { \
        if (i < 0 || j < 0 || k < 0 || i >= reso[0] || j >= reso[1] || k >= reso[2]) return 0;   \
        return idx[((long)i * reso[1] + j) * reso[2] + k];                                         \
    }                                                                                              \
    /* Per-pixel ray set-up (renderer.cu:122-167): normalised origin/direction, world step, */     \
    /* 27-d positional encoding of the view direction, and the [0,1]^3 slab test.           */     \
    static void CAT(rsetup, SFX)(long n, const float* c2w, const float* K, const float* mn,        \
                                 const float* mx, int W, int inverse_y, float near_,                \
                                 float stepdist, GREAL* o, GREAL* d, GREAL* step, GREAL* tmin,      \
                                 GREAL* tmax, REAL* pe) {                                           \
        for (int a = 0; a < 3; ++a)                                                                \
            o[a] = ((GREAL)c2w[4 * a + 3] - (GREAL)mn[a]) / ((GREAL)mx[a] - (GREAL)mn[a]);         \
        GREAL pi = (GREAL)((float)(n % W) + 0.5f), pj = (GREAL)((float)(n / W) + 0.5f);            \
        GREAL dir[3];                                                                              \
        dir[0] = (pi - (GREAL)K[2]) / (GREAL)K[0];                                                 \
        if (inverse_y) {                                                                           \
            dir[1] = (pj - (GREAL)K[5]) / (GREAL)K[4];                                             \
            dir[2] = 1;                                                                            \
        } else {                                                                                   \
            dir[1] = -((pj - (GREAL)K[5]) / (GREAL)K[4]);                                          \
            dir[2] = -1;                                                                           \
        }                                                                                          \
        GREAL rd[3];                                                                               \
        for (int a = 0; a < 3; ++a)                                                                \
            rd[a] = dir[0] * (GREAL)c2w[4 * a] + dir[1] * (GREAL)c2w[4 * a + 1] +                  \
                    dir[2] * (GREAL)c2w[4 * a + 2];                                                \
        GREAL len = (GREAL)sqrt(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]);                    \
        *step = (GREAL)stepdist / len;                                                             \
        for (int a = 0; a < 3; ++a) d[a] = rd[a] / ((GREAL)mx[a] - (GREAL)mn[a]);                  \
        REAL inv = (REAL)1 / (REAL)len, vd[3];                                                     \
        for (int a = 0; a < 3; ++a) vd[a] = (REAL)rd[a] * inv;                                     \
        GREAL v[3], lo[3], hi[3];                                                                  \
        for (int a = 0; a < 3; ++a) {                                                              \
            v[a] = d[a] == 0 ? (GREAL)1e-6f : d[a];                                                \
            GREAL A = ((GREAL)1 - o[a]) / v[a], B = (-o[a]) / v[a];                                \
            lo[a] = A < B ? A : B;                                                                 \
            hi[a] = A < B ? B : A;                                                                 \
        }                                                                                          \
        GREAL l = fmax(fmax(lo[0], lo[1]), lo[2]), h = fmin(fmin(hi[0], hi[1]), hi[2]);            \
        *tmin = fmax(fmin(l, (GREAL)1e9f), (GREAL)near_);                                          \
        *tmax = fmax(fmin(h, (GREAL)1e9f), (GREAL)near_);                                          \
        pe[0] = vd[0];                                                                             \
        pe[1] = vd[1];                                                                             \
        pe[2] = vd[2];                                                                             \
        int base = 1;                                                                              \
        for (int q = 3; q < 7; ++q) {                                                              \
            pe[q] = (REAL)sin(vd[0] * (REAL)base);                                                 \
            pe[q + 4] = (REAL)sin(vd[1] * (REAL)base);                                             \
            pe[q + 8] = (REAL)sin(vd[2] * (REAL)base);                                             \
            pe[q + 12] = (REAL)cos(vd[0] * (REAL)base);                                            \
            pe[q + 16] = (REAL)cos(vd[1] * (REAL)base);                                            \
            pe[q + 20] = (REAL)cos(vd[2] * (REAL)base);                                            \
            base *= 2;                                                                             \
        }                                                                                          \
    }                                                                                              \
    /* Trilinear weights and the 8 corner indices with int() truncation (renderer.cu:271-300). */ \
    static void CAT(rcorners, SFX)(const int32_t* idx, const int* reso, const GREAL* xyz,          \
                                   int* link, REAL* sc) {                                          \
        int i = (int)xyz[0], j = (int)xyz[1], k = (int)xyz[2];                                     \
        REAL u = (REAL)(xyz[0] - (GREAL)i), v = (REAL)(xyz[1] - (GREAL)j), w = (REAL)(xyz[2] - (GREAL)k);                     \
        static const int cn[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0},                   \
                                     {1, 1, 0}, {1, 1, 1}, {1, 0, 1}, {1, 0, 0}};                  \
        for (int q = 0; q < 8; ++q) {                                                              \
            link[q] = CAT(ridx, SFX)(idx, reso, i + cn[q][0], j + cn[q][1], k + cn[q][2]);         \
            REAL a = cn[q][0] ? u : (REAL)1 - u;                                                   \
            REAL b = cn[q][1] ? v : (REAL)1 - v;                                                   \
            REAL c = cn[q][2] ? w : (REAL)1 - w;                                                   \
            sc[q] = (a * b) * c;                                                                   \
        }                                                                                          \
    }                                                                                              \
    /* Pass 1 (first_look_merged, renderer.cu:222-268): count kept samples, record the first */  \
    /* kept t (tmin <- t - step) and the early-stop t (tmax <- t).                            */  \
    void CAT(pvdb_render_count, SFX)(                                                              \
        const int32_t* idx, const int* reso, const float* den, const float* c2w, const float* K, \
        const float* mn, const float* mx, int H, int W, int inverse_y, float near_,              \
        float stepdist, float act_shift, float interval, float thres, double t_stop, REAL* tmin_o,\
        REAL* tmax_o, REAL* step_o, REAL* pe_o, int64_t* count) {                                 \
        const long HW = (long)H * W;                                                               \
        _Pragma("omp parallel for schedule(dynamic, 256)") for (long n = 0; n < HW; ++n) {         \
            GREAL o[3], d[3], step, tmin, tmax;                                                    \
            CAT(rsetup, SFX)(n, c2w, K, mn, mx, W, inverse_y, near_, stepdist, o, d, &step, &tmin, \
                             &tmax, pe_o + 27 * n);                                                \
            GREAL ws[3] = {(GREAL)(reso[0] - 1), (GREAL)(reso[1] - 1), (GREAL)(reso[2] - 1)};     \
            REAL T = 1;                                                                            \
            GREAL t = tmin;                                                                        \
            int64_t cnt = 0;                                                                       \
            int first = 0;                                                                         \
            while (t < tmax) {                                                                     \
                t += step;                                                                         \
                GREAL p[3], xyz[3];                                                                \
                int out = 0;                                                                       \
                for (int a = 0; a < 3; ++a) {                                                      \
                    p[a] = o[a] + t * d[a];                                                        \
                    out |= (p[a] < 0) | (p[a] > 1);                                                \
                }                                                                                  \
                if (out) continue;                                                                 \
                for (int a = 0; a < 3; ++a) xyz[a] = p[a] * ws[a];                                 \
                if (!CAT(ridx, SFX)(idx, reso, (int)rint(xyz[0]), (int)rint(xyz[1]),               \
                                    (int)rint(xyz[2])))                                            \
                    continue;                                                                      \
                int link[8];                                                                       \
                REAL sc[8];                                                                        \
                CAT(rcorners, SFX)(idx, reso, xyz, link, sc);                                      \
                int i = (int)xyz[0], j = (int)xyz[1], k = (int)xyz[2];                             \
                REAL u = (REAL)(xyz[0] - (GREAL)i), v = (REAL)(xyz[1] - (GREAL)j), w = (REAL)(xyz[2] - (GREAL)k);             \
                static const int cn[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0},           \
                                             {1, 1, 0}, {1, 1, 1}, {1, 0, 1}, {1, 0, 0}};          \
                REAL sig = 0;                                                                      \
                for (int q = 0; q < 8; ++q) {                                                      \
                    REAL a = cn[q][0] ? u : (REAL)1 - u;                                           \
                    REAL b = cn[q][1] ? v : (REAL)1 - v;                                           \
                    REAL c = cn[q][2] ? w : (REAL)1 - w;                                           \
                    sig = sig + (((REAL)den[link[q]] * a) * b) * c;                                \
                }                                                                                  \
                REAL e = (REAL)exp(sig + (REAL)act_shift);                                         \
                REAL alpha = (REAL)1 - (REAL)pow((REAL)1 + e, -(REAL)interval);                    \
                if (alpha <= (REAL)thres) continue;                                                \
                REAL wt = T * alpha;                                                               \
                T = T * ((REAL)1 - alpha);                                                         \
                if (wt <= (REAL)thres) continue;                                                   \
                cnt++;                                                                             \
                if (!first) {                                                                      \
                    tmin = t - step;                                                               \
                    first = 1;                                                                     \
                }                                                                                  \
                if ((double)T < t_stop) {                                                          \
                    tmax = t;                                                                      \
                    break;                                                                         \
                }                                                                                  \
            }                                                                                      \
            tmin_o[n] = tmin;                                                                      \
            tmax_o[n] = tmax;                                                                      \
            step_o[n] = step;                                                                      \
            count[n] = cnt;                                                                        \
        }                                                                                          \
    }                                                                                              \
    /* Pass 2 (ray_marching_merged, renderer.cu:312-367): re-march from the rewritten tmin   */    \
    /* to tmax, gather 12-ch features and weights of kept samples; T_last is the final T.    */    \
    /* aux (optional, 4 per sample): the sample's index-space position and its alpha.        */    \
    void CAT(pvdb_render_fill, SFX)(                                                               \
        const int32_t* idx, const int* reso, const float* den, const float* col, int dcol,        \
        const float* c2w, const float* K, const float* mn, const float* mx, int H, int W,         \
        int inverse_y, float near_, float stepdist, float act_shift, float interval, float thres, \
        const REAL* tmin_i, const REAL* tmax_i, const int64_t* count, const int64_t* offs,        \
        REAL* feat, REAL* wts, REAL* aux, int64_t* pix, int64_t* written, REAL* t_last) {         \
        const long HW = (long)H * W;                                                               \
        _Pragma("omp parallel for schedule(dynamic, 256)") for (long n = 0; n < HW; ++n) {         \
            GREAL o[3], d[3], step, tmin, tmax;                                                    \
            REAL pe[27];                                                                           \
            CAT(rsetup, SFX)(n, c2w, K, mn, mx, W, inverse_y, near_, stepdist, o, d, &step, &tmin, \
                             &tmax, pe);                                                           \
            written[n] = 0;                                                                        \
            if (count[n] == 0) {                                                                   \
                t_last[n] = 1;                                                                     \
                continue;                                                                          \
            }                                                                                      \
            GREAL ws[3] = {(GREAL)(reso[0] - 1), (GREAL)(reso[1] - 1), (GREAL)(reso[2] - 1)};     \
            REAL T = 1;                                                                            \
            GREAL t = (GREAL)tmin_i[n];                                                            \
            tmax = (GREAL)tmax_i[n];                                                               \
            int64_t r = offs[n], rend = offs[n] + count[n];                                        \
            while (t < tmax) {                                                                     \
                t += step;                                                                         \
                GREAL p[3], xyz[3];                                                                \
                int out = 0;                                                                       \
                for (int a = 0; a < 3; ++a) {                                                      \
                    p[a] = o[a] + t * d[a];                                                        \
                    out |= (p[a] < 0) | (p[a] > 1);                                                \
                }                                                                                  \
                if (out) continue;                                                                 \
                for (int a = 0; a < 3; ++a) xyz[a] = p[a] * ws[a];                                 \
                if (!CAT(ridx, SFX)(idx, reso, (int)rint(xyz[0]), (int)rint(xyz[1]),               \
                                    (int)rint(xyz[2])))                                            \
                    continue;                                                                      \
                int link[8];                                                                       \
                REAL sc[8];                                                                        \
                CAT(rcorners, SFX)(idx, reso, xyz, link, sc);                                      \
                REAL sig = (REAL)den[link[0]] * sc[0];                                             \
                for (int q = 1; q < 8; ++q) sig = sig + (REAL)den[link[q]] * sc[q];                \
                REAL e = (REAL)exp(sig + (REAL)act_shift);                                         \
                REAL alpha = (REAL)1 - (REAL)pow((REAL)1 + e, -(REAL)interval);                    \
                if (alpha <= (REAL)thres) continue;                                                \
                REAL wt = T * alpha;                                                               \
                T = T * ((REAL)1 - alpha);                                                         \
                if (wt <= (REAL)thres) continue;                                                   \
                if (r < rend) {                                                                    \
                    for (int c = 0; c < dcol; ++c) {                                               \
                        REAL f = (REAL)col[(long)link[0] * dcol + c] * sc[0];                      \
                        for (int q = 1; q < 8; ++q)                                                \
                            f = f + (REAL)col[(long)link[q] * dcol + c] * sc[q];                   \
                        feat[r * dcol + c] = f;                                                    \
                    }                                                                              \
                    wts[r] = wt;                                                                   \
                    if (aux) {                                                                     \
                        for (int a = 0; a < 3; ++a) aux[4 * r + a] = (REAL)xyz[a];                 \
                        aux[4 * r + 3] = alpha;                                                    \
                    }                                                                              \
                    pix[r] = n;                                                                    \
                    r++;                                                                           \
                }                                                                                  \
            }                                                                                      \
            written[n] = r - offs[n];                                                              \
            t_last[n] = T;                                                                         \
        }                                                                                          \
    }

RENDER_IMPL(float, float, _f32)
RENDER_IMPL(double, float, _f64)
