/*
 * Oracle (test infrastructure only): C restatement of the reference's single-operator H_eff matvec,
 *   res[s',a',b'] = sum_{s,a,b,w,w'} L[a,a',w] W[w,w',s,s'] theta[s,a,b] R[b,b',w'],
 * following tools::common::contraction::matrix_vector_product (contraction_tblis.cpp:213-292): the same
 * three pairwise contractions in the same order, with the branch on chiL >= chiR (theta.envL first) or
 * chiL < chiR (theta.envR first). Tensors are Eigen column-major (first index fastest), complex128 stored
 * as interleaved (re, im) doubles.
 *
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg load this library, as the checker and
 * as the host-core baseline; the GPU library never calls it. Parity pins: it must agree with the literal
 * 6-index sum on small shapes and with the numpy restatement (oracle/heff.py) on both branches.
 *
 * The two large contractions run as a packed, cache-blocked complex GEMM (split re/im panels, a 4x4 AVX2
 * FMA register tile, OpenMP over output blocks); the MPO contraction in between streams its operand once.
 * Every operand is read through index maps (no transposed copies): a matrix index i = i0 + n0*i1 sits at
 * complex offset i0*s0 + i1*s1.
 *
 * Build: gcc -O3 -mavx2 -mfma -fopenmp -fPIC -shared heff_c.c -o liboracle_heff.so (oracle/build.py).
 */
#include <immintrin.h>
#include <omp.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    long n0; /* extent of the fast sub-index */
    long s0; /* stride of the fast sub-index (complex elements) */
    long s1; /* stride of the slow sub-index */
} imap;

static imap imap1(long n, long s) { imap m = {n > 0 ? n : 1, s, s * (n > 0 ? n : 1)}; return m; }
static imap imap2(long n0, long s0, long s1) { imap m = {n0, s0, s1}; return m; }
static inline long ioff(imap m, long i) { return (i % m.n0) * m.s0 + (i / m.n0) * m.s1; }

enum { MR = 4, NR = 4, MC = 128, NC = 64, KC = 128 };

/* C tile (4x4) += A micro-panel (4 x kc) . B micro-panel (kc x 4); panels hold per k: 4 re then 4 im. */
static void tile4x4(long kc, const double *restrict ap, const double *restrict bp, double *restrict cre,
                    double *restrict cim) {
    __m256d r0 = _mm256_setzero_pd(), r1 = r0, r2 = r0, r3 = r0;
    __m256d i0 = r0, i1 = r0, i2 = r0, i3 = r0;
    for(long k = 0; k < kc; ++k) {
        const __m256d ar = _mm256_load_pd(ap), ai = _mm256_load_pd(ap + 4);
        __m256d br = _mm256_broadcast_sd(bp + 0), bi = _mm256_broadcast_sd(bp + 4);
        r0 = _mm256_fmadd_pd(ar, br, r0); r0 = _mm256_fnmadd_pd(ai, bi, r0);
        i0 = _mm256_fmadd_pd(ar, bi, i0); i0 = _mm256_fmadd_pd(ai, br, i0);
        br = _mm256_broadcast_sd(bp + 1); bi = _mm256_broadcast_sd(bp + 5);
        r1 = _mm256_fmadd_pd(ar, br, r1); r1 = _mm256_fnmadd_pd(ai, bi, r1);
        i1 = _mm256_fmadd_pd(ar, bi, i1); i1 = _mm256_fmadd_pd(ai, br, i1);
        br = _mm256_broadcast_sd(bp + 2); bi = _mm256_broadcast_sd(bp + 6);
        r2 = _mm256_fmadd_pd(ar, br, r2); r2 = _mm256_fnmadd_pd(ai, bi, r2);
        i2 = _mm256_fmadd_pd(ar, bi, i2); i2 = _mm256_fmadd_pd(ai, br, i2);
        br = _mm256_broadcast_sd(bp + 3); bi = _mm256_broadcast_sd(bp + 7);
        r3 = _mm256_fmadd_pd(ar, br, r3); r3 = _mm256_fnmadd_pd(ai, bi, r3);
        i3 = _mm256_fmadd_pd(ar, bi, i3); i3 = _mm256_fmadd_pd(ai, br, i3);
        ap += 8;
        bp += 8;
    }
    _mm256_storeu_pd(cre + 0, r0); _mm256_storeu_pd(cre + 4, r1);
    _mm256_storeu_pd(cre + 8, r2); _mm256_storeu_pd(cre + 12, r3);
    _mm256_storeu_pd(cim + 0, i0); _mm256_storeu_pd(cim + 4, i1);
    _mm256_storeu_pd(cim + 8, i2); _mm256_storeu_pd(cim + 12, i3);
}

/* C[m,n] = sum_k A[m,k] B[k,n] (complex), every operand read and written through index maps. Work is split
 * over (MC x NC) output blocks; each block packs its own panels and owns its part of C. */
static void zgemm(long M, long N, long K, const double *A, imap am, imap ak, const double *B, imap bk, imap bn,
                  int conjb, double *C, imap cm, imap cn) {
    const double sb = conjb ? -1.0 : 1.0; /* B's imaginary parts are read times sb (conj(B) when conjb) */
    const long mb = (M + MC - 1) / MC, nb = (N + NC - 1) / NC;
#pragma omp parallel
    {
        double *ap = aligned_alloc(64, sizeof(double) * 2 * MC * KC);
        double *bp = aligned_alloc(64, sizeof(double) * 2 * NC * KC);
        long *om = malloc(sizeof(long) * MC), *on = malloc(sizeof(long) * NC);
        long *oka = malloc(sizeof(long) * KC), *okb = malloc(sizeof(long) * KC);
        long *ocm = malloc(sizeof(long) * MC), *ocn = malloc(sizeof(long) * NC);
        double cre[MR * NR], cim[MR * NR];
#pragma omp for schedule(dynamic, 1)
        for(long blk = 0; blk < mb * nb; ++blk) {
            const long ib = blk % mb, jb = blk / mb;
            const long m0 = ib * MC, n0 = jb * NC;
            const long mc = (M - m0 < MC) ? M - m0 : MC, nc = (N - n0 < NC) ? N - n0 : NC;
            const long mcp = (mc + MR - 1) / MR * MR, ncp = (nc + NR - 1) / NR * NR;
            for(long i = 0; i < mc; ++i) { om[i] = ioff(am, m0 + i); ocm[i] = ioff(cm, m0 + i); }
            for(long j = 0; j < nc; ++j) { on[j] = ioff(bn, n0 + j); ocn[j] = ioff(cn, n0 + j); }
            for(long k0 = 0; k0 < K; k0 += KC) {
                const long kc = (K - k0 < KC) ? K - k0 : KC;
                for(long k = 0; k < kc; ++k) { oka[k] = ioff(ak, k0 + k); okb[k] = ioff(bk, k0 + k); }
                /* A block -> per micro-row-panel p: [k][4 re | 4 im], zero-padded past mc */
                for(long p = 0; p < mcp / MR; ++p)
                    for(long k = 0; k < kc; ++k) {
                        double *dst = ap + (p * kc + k) * 2 * MR;
                        for(long r = 0; r < MR; ++r) {
                            const long i = p * MR + r;
                            if(i < mc) {
                                const double *src = A + 2 * (om[i] + oka[k]);
                                dst[r] = src[0];
                                dst[MR + r] = src[1];
                            } else {
                                dst[r] = 0.0;
                                dst[MR + r] = 0.0;
                            }
                        }
                    }
                for(long q = 0; q < ncp / NR; ++q)
                    for(long k = 0; k < kc; ++k) {
                        double *dst = bp + (q * kc + k) * 2 * NR;
                        for(long c = 0; c < NR; ++c) {
                            const long j = q * NR + c;
                            if(j < nc) {
                                const double *src = B + 2 * (okb[k] + on[j]);
                                dst[c] = src[0];
                                dst[NR + c] = sb * src[1];
                            } else {
                                dst[c] = 0.0;
                                dst[NR + c] = 0.0;
                            }
                        }
                    }
                for(long q = 0; q < ncp / NR; ++q)
                    for(long p = 0; p < mcp / MR; ++p) {
                        tile4x4(kc, ap + p * kc * 2 * MR, bp + q * kc * 2 * NR, cre, cim);
                        for(long c = 0; c < NR; ++c) {
                            const long j = q * NR + c;
                            if(j >= nc) break;
                            for(long r = 0; r < MR; ++r) {
                                const long i = p * MR + r;
                                if(i >= mc) break;
                                double *dst = C + 2 * (ocm[i] + ocn[j]);
                                if(k0 == 0) {
                                    dst[0] = cre[c * MR + r];
                                    dst[1] = cim[c * MR + r];
                                } else {
                                    dst[0] += cre[c * MR + r];
                                    dst[1] += cim[c * MR + r];
                                }
                            }
                        }
                    }
            }
        }
        free(ap); free(bp); free(om); free(on); free(oka); free(okb); free(ocm); free(ocn);
    }
}

/* The MPO contraction: out[o + no*(s' + d*(x + nx*w_o))] = sum_{w_i, s} W(w_i, w_o, s, s') in[s + d*(o + no*(x +
 * nx*w_i))]; (o, x) are the two bond indices carried through, o the fast one on both sides. wl selects whether
 * the contracted MPO bond is W's left (1) or right (0) index. */
static void mpo_leg(const double *in, const double *W, double *out, long d, long no, long nx, long Din, long Dout,
                    long DL, long DR, int wl) {
#pragma omp parallel for collapse(2) schedule(static)
    for(long wo = 0; wo < Dout; ++wo)
        for(long x = 0; x < nx; ++x)
            for(long sp = 0; sp < d; ++sp) {
                double *y = out + 2 * no * (sp + d * (x + nx * wo));
                memset(y, 0, sizeof(double) * 2 * no);
                for(long wi = 0; wi < Din; ++wi)
                    for(long s = 0; s < d; ++s) {
                        const long w = wl ? wi + DL * (wo + DR * (s + d * sp)) : wo + DL * (wi + DR * (s + d * sp));
                        const double cr = W[2 * w], ci = W[2 * w + 1];
                        if(cr == 0.0 && ci == 0.0) continue;
                        const double *v = in + 2 * (s + d * no * (x + nx * wi));
                        for(long o = 0; o < no; ++o) {
                            const double vr = v[2 * d * o], vi = v[2 * d * o + 1];
                            y[2 * o] += cr * vr - ci * vi;
                            y[2 * o + 1] += cr * vi + ci * vr;
                        }
                    }
            }
}

/* Intermediates live in one workspace that grows and is kept between calls, so a repeated matvec does not
 * page-fault 2 x 40 MB of fresh memory each time (at chi=512 that cost more than the contractions). */
static double *ws;
static size_t ws_doubles;
static pthread_mutex_t ws_lock = PTHREAD_MUTEX_INITIALIZER;

static double *workspace(size_t doubles) {
    if(doubles > ws_doubles) {
        free(ws);
        ws_doubles = 0;
        ws = malloc(sizeof(double) * doubles);
        if(!ws) return NULL;
        const long chunks = 64, per = (long)((doubles + chunks - 1) / chunks);
#pragma omp parallel for schedule(static)
        for(long c = 0; c < chunks; ++c) {
            const long b = c * per, e = (long)doubles < b + per ? (long)doubles : b + per;
            if(b < e) memset(ws + b, 0, sizeof(double) * (size_t)(e - b));
        }
        ws_doubles = doubles;
    }
    return ws;
}

/* res = H_eff . theta. order: 1 contracts theta.envL first, 0 theta.envR first, -1 picks chiL >= chiR as the
 * reference does. nthreads > 0 sets the OpenMP thread count (for later calls too). Returns 0, or 1 on bad sizes, 2 on allocation
 * failure. Calls are serialised (one shared workspace). */
int oracle_heff_matvec(const double *theta, const double *mpo, const double *envL, const double *envR, double *res,
                       long d, long chiL, long chiR, long DL, long DR, int order, int nthreads) {
    if(d <= 0 || chiL <= 0 || chiR <= 0 || DL <= 0 || DR <= 0) return 1;
    if(nthreads > 0) omp_set_num_threads(nthreads);
    if(order < 0) order = chiL >= chiR ? 1 : 0;
    const long n1 = d * chiL * chiR * (order ? DL : DR), n2 = d * chiL * chiR * (order ? DR : DL);
    pthread_mutex_lock(&ws_lock);
    double *t1 = workspace((size_t)(2 * (n1 + n2)));
    if(!t1) {
        pthread_mutex_unlock(&ws_lock);
        return 2;
    }
    double *t2 = t1 + 2 * n1;
    if(order) {
        /* t1[s, b, a', w] = sum_a theta[s,a,b] L[a,a',w]   (mps.contract(envL, idx({1},{0}))) */
        zgemm(d * chiR, chiL * DL, chiL, theta, imap2(d, 1, d * chiL), imap1(chiL, d), envL, imap1(chiL, 1),
              imap1(chiL * DL, chiL), 0, t1, imap1(d * chiR, 1), imap1(chiL * DL, d * chiR));
        /* t2[b, s', a', w'] = sum_{w,s} t1[s, b, a', w] W[w,w',s,s']   (.contract(mpo, idx({3,0},{0,2}))) */
        mpo_leg(t1, mpo, t2, d, chiR, chiL, DL, DR, DL, DR, 1);
        /* res[s',a',b'] = sum_{b,w'} t2[b, s', a', w'] R[b,b',w']   (.contract(envR, idx({0,2},{0,2})), shuffled) */
        zgemm(d * chiL, chiR, chiR * DR, t2, imap1(d * chiL, chiR), imap2(chiR, 1, chiR * d * chiL), envR,
              imap2(chiR, 1, chiR * chiR), imap1(chiR, chiR), 0, res, imap1(d * chiL, 1), imap1(chiR, d * chiL));
    } else {
        /* t1[s, a, b', w'] = sum_b theta[s,a,b] R[b,b',w']   (mps.contract(envR, idx({2},{0}))) */
        zgemm(d * chiL, chiR * DR, chiR, theta, imap1(d * chiL, 1), imap1(chiR, d * chiL), envR, imap1(chiR, 1),
              imap1(chiR * DR, chiR), 0, t1, imap1(d * chiL, 1), imap1(chiR * DR, d * chiL));
        /* t2[a, s', b', w] = sum_{w',s} t1[s, a, b', w'] W[w,w',s,s']   (.contract(mpo, idx({3,0},{1,2}))) */
        mpo_leg(t1, mpo, t2, d, chiL, chiR, DR, DL, DL, DR, 0);
        /* res[s',a',b'] = sum_{a,w} t2[a, s', b', w] L[a,a',w]   (.contract(envL, idx({0,2},{0,2})), shuffled) */
        zgemm(d * chiR, chiL, chiL * DL, t2, imap1(d * chiR, chiL), imap2(chiL, 1, chiL * d * chiR), envL,
              imap2(chiL, 1, chiL * chiL), imap1(chiL, chiL), 0, res, imap2(d, 1, d * chiL), imap1(chiL, d));
    }
    pthread_mutex_unlock(&ws_lock);
    return 0;
}

/* The MPO contraction of the environment update: out[x + nx*(o + no*(wo + Dout*s))] = sum_{wi, s'} W(., ., s, s')
 * in[x + nx*(wi + Din*(s' + d*o))], with W's contracted bond the left one (wl = 1, left block) or the right one
 * (wl = 0, right block); W's third index s stays free and its fourth s' meets the conjugated site tensor. */
static void env_mpo_leg(const double *in, const double *W, double *out, long d, long nx, long no, long Din, long Dout,
                        long DL, long DR, int wl) {
#pragma omp parallel for collapse(2) schedule(static)
    for(long s = 0; s < d; ++s)
        for(long wo = 0; wo < Dout; ++wo)
            for(long o = 0; o < no; ++o) {
                double *y = out + 2 * nx * (o + no * (wo + Dout * s));
                memset(y, 0, sizeof(double) * 2 * nx);
                for(long sp = 0; sp < d; ++sp)
                    for(long wi = 0; wi < Din; ++wi) {
                        const long w = wl ? wi + DL * (wo + DR * (s + d * sp)) : wo + DL * (wi + DR * (s + d * sp));
                        const double cr = W[2 * w], ci = W[2 * w + 1];
                        if(cr == 0.0 && ci == 0.0) continue;
                        const double *v = in + 2 * nx * (wi + Din * (sp + d * o));
                        for(long x = 0; x < nx; ++x) {
                            const double vr = v[2 * x], vi = v[2 * x + 1];
                            y[2 * x] += cr * vr - ci * vi;
                            y[2 * x + 1] += cr * vi + ci * vr;
                        }
                    }
            }
}

/* New environment block of EnvBase::build_block (EnvBase.cpp:67-151) in its contraction order.
 * side 0 (left): out[b, b', w'] (chiR, chiR, DR) from other = L (chiL, chiL, DL):
 *   t1[a, w, s', b'] = sum_a' L[a, a', w] conj(mps[s', a', b'])        (.contract(mps.conjugate(), idx({1},{1})))
 *   t2[a, b', w', s] = sum_{w, s'} t1 W[w, w', s, s']                  (.contract(mpo, idx({1,2},{0,3})))
 *   out[b, b', w']   = sum_{s, a} mps[s, a, b] t2[a, b', w', s]        (mps.contract(., idx({0,1},{3,0})))
 * side 1 (right): out[a, a', w] (chiL, chiL, DL) from other = R (chiR, chiR, DR), symmetric over b.
 * The intermediates keep their first index contiguous (Eigen's output order). Returns 0, 1 on bad sizes, 2 on
 * allocation failure. */
int oracle_env_build_block(int side, const double *other, const double *mps, const double *mpo, double *out, long d,
                           long chiL, long chiR, long DL, long DR, int nthreads) {
    if(d <= 0 || chiL <= 0 || chiR <= 0 || DL <= 0 || DR <= 0 || (side != 0 && side != 1)) return 1;
    if(nthreads > 0) omp_set_num_threads(nthreads);
    const long n1 = side == 0 ? chiL * DL * d * chiR : chiR * DR * d * chiL;
    const long n2 = side == 0 ? chiL * chiR * DR * d : chiR * chiL * DL * d;
    pthread_mutex_lock(&ws_lock);
    double *t1 = workspace((size_t)(2 * (n1 + n2)));
    if(!t1) {
        pthread_mutex_unlock(&ws_lock);
        return 2;
    }
    double *t2 = t1 + 2 * n1;
    if(side == 0) {
        /* t1[a + chiL*(w + DL*(s' + d*b'))]: M = (a, w), K = a', N = (s', b') */
        zgemm(chiL * DL, d * chiR, chiL, other, imap2(chiL, 1, chiL * chiL), imap1(chiL, chiL), mps, imap1(chiL, d),
              imap2(d, 1, d * chiL), 1, t1, imap1(chiL * DL, 1), imap1(d * chiR, chiL * DL));
        /* t2[a + chiL*(b' + chiR*(w' + DR*s))] */
        env_mpo_leg(t1, mpo, t2, d, chiL, chiR, DL, DR, DL, DR, 1);
        /* out[b + chiR*(b' + chiR*w')]: M = b, K = (s, a), N = (b', w') */
        zgemm(chiR, chiR * DR, d * chiL, mps, imap1(chiR, d * chiL), imap1(d * chiL, 1), t2, imap2(d, chiL * chiR * DR, 1),
              imap1(chiR * DR, chiL), 0, out, imap1(chiR, 1), imap1(chiR * DR, chiR));
    } else {
        /* t1[b + chiR*(w' + DR*(s' + d*a'))]: M = (b, w'), K = b', N = (s', a') */
        zgemm(chiR * DR, d * chiL, chiR, other, imap2(chiR, 1, chiR * chiR), imap1(chiR, chiR), mps, imap1(chiR, d * chiL),
              imap1(d * chiL, 1), 1, t1, imap1(chiR * DR, 1), imap1(d * chiL, chiR * DR));
        /* t2[b + chiR*(a' + chiL*(w + DL*s))] */
        env_mpo_leg(t1, mpo, t2, d, chiR, chiL, DR, DL, DL, DR, 0);
        /* out[a + chiL*(a' + chiL*w)]: M = a, K = (s, b), N = (a', w) */
        zgemm(chiL, chiL * DL, d * chiR, mps, imap1(chiL, d), imap2(d, 1, d * chiL), t2, imap2(d, chiR * chiL * DL, 1),
              imap1(chiL * DL, chiR), 0, out, imap1(chiL, 1), imap1(chiL * DL, chiL));
    }
    pthread_mutex_unlock(&ws_lock);
    return 0;
}
