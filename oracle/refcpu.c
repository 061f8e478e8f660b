/*
 * refcpu.c -- CPU restatement of the reference's FP64 force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests, the smoke check
 * and the CPU-baseline leg of bench.py call it to check or time against the
 * HIP product path.  Nothing in akitsushima_amd/ links, loads or calls it.
 *
 * Each function restates one piece of the reference (akitsushima 0.4.0a4)
 * and cites the file:line it follows.  Loop structure and operation order
 * follow the reference's numba kernels; threads use per-thread force buffers
 * exactly as the reference's prange kernels do (engine/tersoff.py:464-509,
 * engine/stillinger_weber.py:258-395).
 *
 * Compiled with -ffp-contract=off so that the neighbour-list predicate is
 * evaluated with the reference's unfused multiply/add order
 * (engine/neighborlist.py:350-363, numba without fastmath).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793 /* math.pi */

static int orc_threads(int want)
{
#ifdef _OPENMP
    if (want <= 0) return omp_get_max_threads();
    return want;
#else
    (void)want;
    return 1;
#endif
}

/* engine/_numba_common.py:12-32 -- exp clamped at +-500 */
static inline double nb_exp(double x)
{
    if (x > 500.0) return exp(500.0);
    if (x < -500.0) return 0.0;
    return exp(x);
}

static inline int pymod(int64_t a, int64_t n)
{
    int64_t r = a % n;
    return (int)(r < 0 ? r + n : r);
}

/* Minimum-image displacement r_j - r_i, operation order of
 * engine/neighborlist.py:350-361 (round = half-to-even = rint). */
static inline void mic(const double *pi, const double *pj, const double *lat,
                       const double *inv, double *d)
{
    double rx = pj[0] - pi[0], ry = pj[1] - pi[1], rz = pj[2] - pi[2];
    double sx = rx * inv[0] + ry * inv[3] + rz * inv[6];
    double sy = rx * inv[1] + ry * inv[4] + rz * inv[7];
    double sz = rx * inv[2] + ry * inv[5] + rz * inv[8];
    sx -= rint(sx);
    sy -= rint(sy);
    sz -= rint(sz);
    d[0] = sx * lat[0] + sy * lat[3] + sz * lat[6];
    d[1] = sx * lat[1] + sy * lat[4] + sz * lat[7];
    d[2] = sx * lat[2] + sy * lat[5] + sz * lat[8];
}

/* ------------------------------------------------------------------------
 * Cell-linked neighbour list: engine/neighborlist.py:935-1024 (build),
 * :312-373 (pair search), :270-309 (canonical offsets), :1166-1181
 * (canonical (i, j) order).  Returns the number of pairs; when out_i is
 * NULL only counts.  Pairs are emitted i ascending, and for each i sorted by
 * j, which is the order _canonical_pair_order produces.
 * ---------------------------------------------------------------------- */
static int axis_offsets(int nc, int *o)
{
    if (nc == 1) { o[0] = 0; return 1; }
    if (nc == 2) { o[0] = 0; o[1] = 1; return 2; }
    o[0] = -1; o[1] = 0; o[2] = 1;
    return 3;
}

int64_t orc_nl_build(int32_t n, const int32_t *nc, const double *pos,
                     const double *lat, const double *inv, double rl2,
                     int32_t *out_i, int32_t *out_j, double *out_rij,
                     double *out_r)
{
    int nc0 = nc[0], nc1 = nc[1], nc2 = nc[2];
    int64_t ncell = (int64_t)nc0 * nc1 * nc2;
    int64_t *head = (int64_t *)malloc(sizeof(int64_t) * (size_t)ncell);
    int64_t *nxt = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int32_t *cidx = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(n > 0 ? n : 1));
    for (int64_t c = 0; c < ncell; ++c) head[c] = -1;
    for (int32_t a = 0; a < n; ++a) {
        /* frac = positions @ inv_lat; frac -= floor(frac); floor(frac*nc), clipped */
        const double *p = pos + 3 * (size_t)a;
        int ncs[3] = {nc0, nc1, nc2};
        for (int b = 0; b < 3; ++b) {
            double f = p[0] * inv[b] + p[1] * inv[3 + b] + p[2] * inv[6 + b];
            f = f - floor(f);
            int64_t k = (int64_t)floor(f * ncs[b]);
            if (k < 0) k = 0;
            if (k > ncs[b] - 1) k = ncs[b] - 1;
            cidx[3 * (size_t)a + b] = (int32_t)k;
        }
    }
    for (int32_t a = n - 1; a >= 0; --a) {
        int64_t c = (int64_t)cidx[3 * (size_t)a] * nc1 * nc2 +
                    (int64_t)cidx[3 * (size_t)a + 1] * nc2 + cidx[3 * (size_t)a + 2];
        nxt[a] = head[c];
        head[c] = a;
    }
    int o0[3], o1[3], o2[3];
    int n0 = axis_offsets(nc0, o0), n1 = axis_offsets(nc1, o1), n2 = axis_offsets(nc2, o2);
    int64_t found = 0;
    for (int32_t i = 0; i < n; ++i) {
        int64_t row0 = found;
        int ci0 = cidx[3 * (size_t)i], ci1 = cidx[3 * (size_t)i + 1], ci2 = cidx[3 * (size_t)i + 2];
        for (int a0 = 0; a0 < n0; ++a0)
            for (int a1 = 0; a1 < n1; ++a1)
                for (int a2 = 0; a2 < n2; ++a2) {
                    int cj0 = pymod(ci0 + o0[a0], nc0);
                    int cj1 = pymod(ci1 + o1[a1], nc1);
                    int cj2 = pymod(ci2 + o2[a2], nc2);
                    int64_t cf = (int64_t)cj0 * nc1 * nc2 + (int64_t)cj1 * nc2 + cj2;
                    for (int64_t j = head[cf]; j >= 0; j = nxt[j]) {
                        if (j <= i) continue;
                        double d[3];
                        mic(pos + 3 * (size_t)i, pos + 3 * (size_t)j, lat, inv, d);
                        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                        if (r2 < rl2 && r2 > 1e-20) {
                            if (out_i) {
                                out_i[found] = i;
                                out_j[found] = (int32_t)j;
                                out_rij[3 * found] = d[0];
                                out_rij[3 * found + 1] = d[1];
                                out_rij[3 * found + 2] = d[2];
                                out_r[found] = sqrt(r2);
                            }
                            ++found;
                        }
                    }
                }
        if (out_i) {
            /* canonical order within the row of i: sort by j (insertion) */
            for (int64_t a = row0 + 1; a < found; ++a) {
                int32_t jj = out_j[a];
                double r = out_r[a], x = out_rij[3 * a], y = out_rij[3 * a + 1], z = out_rij[3 * a + 2];
                int64_t b = a - 1;
                while (b >= row0 && out_j[b] > jj) {
                    out_j[b + 1] = out_j[b];
                    out_r[b + 1] = out_r[b];
                    out_rij[3 * (b + 1)] = out_rij[3 * b];
                    out_rij[3 * (b + 1) + 1] = out_rij[3 * b + 1];
                    out_rij[3 * (b + 1) + 2] = out_rij[3 * b + 2];
                    --b;
                }
                out_j[b + 1] = jj;
                out_r[b + 1] = r;
                out_rij[3 * (b + 1)] = x;
                out_rij[3 * (b + 1) + 1] = y;
                out_rij[3 * (b + 1) + 2] = z;
            }
        }
    }
    free(head);
    free(nxt);
    free(cidx);
    return found;
}

/* ------------------------------------------------------------------------
 * Shared glue of the potentials' compute(): minimum image of the candidate
 * pairs and the r < rc & r > 1e-12 mask (engine/tersoff.py:1276-1291), then
 * the directed-bond CSR built by a stable argsort of [idx_i, idx_j]
 * (engine/tersoff.py:1296-1307): row a holds first the pairs with i == a,
 * then the pairs with j == a, each in pair order.
 * ---------------------------------------------------------------------- */
typedef struct {
    int64_t P;
    int32_t *ii, *jj;
    double *rij, *r;
    int64_t *nb_start; /* n + 1 */
    int32_t *nb_idx;
    double *nb_rij, *nb_r;
} pairs_t;

static void pairs_free(pairs_t *p)
{
    free(p->ii); free(p->jj); free(p->rij); free(p->r);
    free(p->nb_start); free(p->nb_idx); free(p->nb_rij); free(p->nb_r);
}

static void pairs_make(pairs_t *pr, int32_t n, const double *pos, const double *lat,
                       const double *inv, int64_t ncand, const int32_t *ci,
                       const int32_t *cj, double rc)
{
    pr->ii = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ncand + 1));
    pr->jj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ncand + 1));
    pr->rij = (double *)malloc(sizeof(double) * 3 * (size_t)(ncand + 1));
    pr->r = (double *)malloc(sizeof(double) * (size_t)(ncand + 1));
    int64_t P = 0;
    for (int64_t c = 0; c < ncand; ++c) {
        double d[3];
        mic(pos + 3 * (size_t)ci[c], pos + 3 * (size_t)cj[c], lat, inv, d);
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (r < rc && r > 1e-12) {
            pr->ii[P] = ci[c];
            pr->jj[P] = cj[c];
            pr->rij[3 * P] = d[0];
            pr->rij[3 * P + 1] = d[1];
            pr->rij[3 * P + 2] = d[2];
            pr->r[P] = r;
            ++P;
        }
    }
    pr->P = P;
    pr->nb_start = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    pr->nb_idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * P + 1));
    pr->nb_rij = (double *)malloc(sizeof(double) * 3 * (size_t)(2 * P + 1));
    pr->nb_r = (double *)malloc(sizeof(double) * (size_t)(2 * P + 1));
    for (int64_t p = 0; p < P; ++p) {
        pr->nb_start[pr->ii[p] + 1]++;
        pr->nb_start[pr->jj[p] + 1]++;
    }
    for (int32_t a = 0; a < n; ++a) pr->nb_start[a + 1] += pr->nb_start[a];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    memcpy(fill, pr->nb_start, sizeof(int64_t) * ((size_t)n + 1));
    for (int64_t p = 0; p < P; ++p) { /* forward entries first (stable) */
        int64_t s = fill[pr->ii[p]]++;
        pr->nb_idx[s] = pr->jj[p];
        for (int d = 0; d < 3; ++d) pr->nb_rij[3 * s + d] = pr->rij[3 * p + d];
        pr->nb_r[s] = pr->r[p];
    }
    for (int64_t p = 0; p < P; ++p) {
        int64_t s = fill[pr->jj[p]]++;
        pr->nb_idx[s] = pr->ii[p];
        for (int d = 0; d < 3; ++d) pr->nb_rij[3 * s + d] = -pr->rij[3 * p + d];
        pr->nb_r[s] = pr->r[p];
    }
    free(fill);
}

/* ------------------------------------------------------------------------
 * Tersoff scalar helpers: engine/tersoff.py:48-58 (_te_fc), :127-135
 * (_te_dfc_scalar), :161-202 (_te_zterm_grad).
 * ---------------------------------------------------------------------- */
static inline double te_fc(double r, double R, double D)
{
    if (r < R - D) return 1.0;
    if (r > R + D) return 0.0;
    return 0.5 - 0.5 * sin(0.5 * ORC_PI * (r - R) / D);
}

static inline double te_dfc(double r, double R, double D)
{
    if (r < R - D || r > R + D) return 0.0;
    return -0.25 * ORC_PI / D * cos(0.5 * ORC_PI * (r - R) / D);
}

static inline void te_zterm_grad(const double *rij, const double *rik, double R, double D,
                                 double c, double d, double h, double l3, double m,
                                 double gam, double *gj, double *gk)
{
    double r_ij = sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
    double r_ik = sqrt(rik[0] * rik[0] + rik[1] * rik[1] + rik[2] * rik[2]);
    double fc_ik = te_fc(r_ik, R, D);
    if (fc_ik == 0.0) {
        gj[0] = gj[1] = gj[2] = gk[0] = gk[1] = gk[2] = 0.0;
        return;
    }
    double dfc_ik = te_dfc(r_ik, R, D);
    double inv_ij = 1.0 / r_ij, inv_ik = 1.0 / r_ik;
    double cos_t = (rij[0] * rik[0] + rij[1] * rik[1] + rij[2] * rik[2]) * inv_ij * inv_ik;
    double hc = h - cos_t;
    double q = d * d + hc * hc;
    double g = 1.0 + c * c / (d * d) - c * c / q;
    double dg = -2.0 * c * c * hc / (q * q);
    double u = r_ij - r_ik;
    double ex = nb_exp(pow(l3, m) * pow(u, m));
    double dex = ex * m * pow(l3, m) * pow(u, m - 1.0);
    double a_j = gam * fc_ik * dg * ex;
    double b_j = gam * fc_ik * g * dex * inv_ij;
    double b_k = gam * (dfc_ik * g * ex - fc_ik * g * dex) * inv_ik;
    double ii_k = inv_ij * inv_ik;
    double cj = cos_t * inv_ij * inv_ij;
    double ck = cos_t * inv_ik * inv_ik;
    for (int x = 0; x < 3; ++x) {
        gj[x] = a_j * (rik[x] * ii_k - cj * rij[x]) + b_j * rij[x];
        gk[x] = a_j * (rij[x] * ii_k - ck * rik[x]) + b_k * rik[x];
    }
}

/* ------------------------------------------------------------------------
 * Tersoff, legacy flat single-species path: engine/tersoff.py:1233-1513
 * with the numba kernels _te_zeta_numpy (:464-509) and
 * _te_3body_force_numpy (:208-325); m = 3 and gamma = 1 are fixed there.
 * prm = A, B, lambda1, lambda2, lambda3, beta, n, c, d, h, R, D.
 * Candidate pairs (i < j) come from a neighbour list or all pairs.
 * ---------------------------------------------------------------------- */
int orc_tersoff_flat(int32_t n, const double *pos, const double *lat, const double *inv,
                     double volume, int64_t ncand, const int32_t *ci, const int32_t *cj,
                     const double *prm, int nthreads, double *energy, double *forces,
                     double *stress)
{
    double A = prm[0], B = prm[1], l1 = prm[2], l2 = prm[3], l3 = prm[4], bet = prm[5];
    double nn = prm[6], c = prm[7], d = prm[8], h = prm[9], R = prm[10], D = prm[11];
    double rc = R + D;
    memset(forces, 0, sizeof(double) * 3 * (size_t)n);
    memset(stress, 0, sizeof(double) * 9);
    *energy = 0.0;
    pairs_t pr;
    pairs_make(&pr, n, pos, lat, inv, ncand, ci, cj, rc);
    int64_t P = pr.P;
    if (P == 0) { pairs_free(&pr); return 0; }
    int T = orc_threads(nthreads);
    double *zij = (double *)calloc((size_t)P, sizeof(double));
    double *zji = (double *)calloc((size_t)P, sizeof(double));
    double l3c = pow(l3, 3.0);
    /* _te_zeta_numpy */
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t p = 0; p < P; ++p) {
        int32_t ii = pr.ii[p], jj = pr.jj[p];
        double r_ij = pr.r[p];
        const double *v = pr.rij + 3 * p;
        double zs = 0.0;
        for (int64_t nb = pr.nb_start[ii]; nb < pr.nb_start[ii + 1]; ++nb) {
            if (pr.nb_idx[nb] == jj) continue;
            double r_ik = pr.nb_r[nb];
            double fc_ik = te_fc(r_ik, R, D);
            if (fc_ik == 0.0) continue;
            const double *w = pr.nb_rij + 3 * nb;
            double dot = 0.0;
            for (int x = 0; x < 3; ++x) dot += v[x] * w[x];
            double cs = dot / (r_ij * r_ik);
            double g = 1.0 + c * c / (d * d) - c * c / (d * d + (h - cs) * (h - cs));
            zs += fc_ik * g * nb_exp(l3c * pow(r_ij - r_ik, 3.0));
        }
        zij[p] = zs;
        zs = 0.0;
        for (int64_t nb = pr.nb_start[jj]; nb < pr.nb_start[jj + 1]; ++nb) {
            if (pr.nb_idx[nb] == ii) continue;
            double r_jk = pr.nb_r[nb];
            double fc_jk = te_fc(r_jk, R, D);
            if (fc_jk == 0.0) continue;
            const double *w = pr.nb_rij + 3 * nb;
            double dot = 0.0;
            for (int x = 0; x < 3; ++x) dot += (-v[x]) * w[x];
            double cs = dot / (r_ij * r_jk);
            double g = 1.0 + c * c / (d * d) - c * c / (d * d + (h - cs) * (h - cs));
            zs += fc_jk * g * nb_exp(l3c * pow(r_ij - r_jk, 3.0));
        }
        zji[p] = zs;
    }
    /* _bij / _dbij (engine/tersoff.py:1380-1391) and the 2-body term (:1399-1414) */
    double *dbi = (double *)malloc(sizeof(double) * (size_t)P);
    double *dbj = (double *)malloc(sizeof(double) * (size_t)P);
    double *fcv = (double *)malloc(sizeof(double) * (size_t)P);
    double *fav = (double *)malloc(sizeof(double) * (size_t)P);
    double e = 0.0;
    double *F = forces, S[9] = {0};
    for (int64_t p = 0; p < P; ++p) {
        double b1 = 1.0, b2 = 1.0, d1 = 0.0, d2 = 0.0;
        if (zij[p] > 0) {
            double bz = pow(bet * zij[p], nn);
            b1 = pow(1.0 + bz, -0.5 / nn);
            d1 = -0.5 * b1 * bz / (zij[p] * (1.0 + bz));
        }
        if (zji[p] > 0) {
            double bz = pow(bet * zji[p], nn);
            b2 = pow(1.0 + bz, -0.5 / nn);
            d2 = -0.5 * b2 * bz / (zji[p] * (1.0 + bz));
        }
        dbi[p] = d1;
        dbj[p] = d2;
        double r = pr.r[p];
        double bavg = 0.5 * (b1 + b2);
        double fc = te_fc(r, R, D), dfc = te_dfc(r, R, D);
        double fr = A * exp(-l1 * r), fa = -B * exp(-l2 * r);
        fcv[p] = fc;
        fav[p] = fa;
        e += fc * (fr + bavg * fa);
        double dEdr = dfc * (fr + bavg * fa) + fc * (-l1 * fr + bavg * (-l2 * fa));
        double s = dEdr / r;
        const double *v = pr.rij + 3 * p;
        double fij[3] = {s * v[0], s * v[1], s * v[2]};
        for (int x = 0; x < 3; ++x) {
            F[3 * (size_t)pr.ii[p] + x] += fij[x];
            F[3 * (size_t)pr.jj[p] + x] -= fij[x];
        }
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) S[3 * x + y] -= fij[x] * v[y];
    }
    /* _te_3body_force_numpy, per-thread buffers as its prange twin */
    double *Ft = (double *)calloc((size_t)T * 3 * (size_t)n, sizeof(double));
    double *St = (double *)calloc((size_t)T * 9, sizeof(double));
#pragma omp parallel num_threads(T)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *Fb = Ft + (size_t)tid * 3 * (size_t)n;
        double *Sb = St + (size_t)tid * 9;
#pragma omp for schedule(static)
        for (int64_t p = 0; p < P; ++p) {
            int32_t ii = pr.ii[p], jj = pr.jj[p];
            const double *v = pr.rij + 3 * p;
            double fc_ij = fcv[p], fa_ij = fav[p];
            if (fabs(dbi[p]) > 1e-30) {
                double coeff = 0.5 * fc_ij * fa_ij * dbi[p];
                for (int64_t nb = pr.nb_start[ii]; nb < pr.nb_start[ii + 1]; ++nb) {
                    int32_t kk = pr.nb_idx[nb];
                    if (kk == jj) continue;
                    if (te_fc(pr.nb_r[nb], R, D) == 0.0) continue;
                    const double *w = pr.nb_rij + 3 * nb;
                    double gj[3], gk[3];
                    te_zterm_grad(v, w, R, D, c, d, h, l3, 3.0, 1.0, gj, gk);
                    for (int x = 0; x < 3; ++x) {
                        double Fj = -coeff * gj[x], Fk = -coeff * gk[x];
                        Fb[3 * (size_t)ii + x] += coeff * (gj[x] + gk[x]);
                        Fb[3 * (size_t)jj + x] += Fj;
                        Fb[3 * (size_t)kk + x] += Fk;
                        for (int y = 0; y < 3; ++y) Sb[3 * x + y] += Fj * v[y] + Fk * w[y];
                    }
                }
            }
            if (fabs(dbj[p]) > 1e-30) {
                double coeff = 0.5 * fc_ij * fa_ij * dbj[p];
                double vji[3] = {-v[0], -v[1], -v[2]};
                for (int64_t nb = pr.nb_start[jj]; nb < pr.nb_start[jj + 1]; ++nb) {
                    int32_t kk = pr.nb_idx[nb];
                    if (kk == ii) continue;
                    if (te_fc(pr.nb_r[nb], R, D) == 0.0) continue;
                    const double *w = pr.nb_rij + 3 * nb;
                    double gi[3], gk[3];
                    te_zterm_grad(vji, w, R, D, c, d, h, l3, 3.0, 1.0, gi, gk);
                    for (int x = 0; x < 3; ++x) {
                        double Fi = -coeff * gi[x], Fk = -coeff * gk[x];
                        Fb[3 * (size_t)jj + x] += coeff * (gi[x] + gk[x]);
                        Fb[3 * (size_t)ii + x] += Fi;
                        Fb[3 * (size_t)kk + x] += Fk;
                        for (int y = 0; y < 3; ++y) Sb[3 * x + y] += Fi * vji[y] + Fk * w[y];
                    }
                }
            }
        }
    }
    for (int t = 0; t < T; ++t) {
        for (size_t k = 0; k < 3 * (size_t)n; ++k) F[k] += Ft[(size_t)t * 3 * (size_t)n + k];
        for (int k = 0; k < 9; ++k) S[k] += St[(size_t)t * 9 + k];
    }
    for (int k = 0; k < 9; ++k) stress[k] = S[k] / volume;
    *energy = e;
    free(Ft); free(St); free(zij); free(zji); free(dbi); free(dbj); free(fcv); free(fav);
    pairs_free(&pr);
    return 0;
}

/* ------------------------------------------------------------------------
 * Tersoff, table-driven multi-species path: engine/tersoff.py:1518-1951
 * with _te_multi_zeta_numpy (:528-606), the vectorised pass 2
 * (:1673-1702) and _te_multi_pass3_numpy (:685-855).
 * tb: (ns*ns) x 9 rows [valid, A, B, l1, l2, R, D, beta, n];
 * tri: (ns^3) x 7 rows [valid, c, d, h, l3, m, gamma].
 * ---------------------------------------------------------------------- */
#define TB(a, b, k) tb[((size_t)(a) * ns + (b)) * 9 + (k)]
#define TR(a, b, c, k) tri[(((size_t)(a) * ns + (b)) * ns + (c)) * 7 + (k)]

int orc_tersoff_multi(int32_t n, const double *pos, const double *lat, const double *inv,
                      double volume, int64_t ncand, const int32_t *ci, const int32_t *cj,
                      const int32_t *type_ids, int32_t ns, const double *tb,
                      const double *tri, double rc, int nthreads, double *energy,
                      double *forces, double *stress)
{
    memset(forces, 0, sizeof(double) * 3 * (size_t)n);
    memset(stress, 0, sizeof(double) * 9);
    *energy = 0.0;
    pairs_t pr;
    pairs_make(&pr, n, pos, lat, inv, ncand, ci, cj, rc);
    int64_t P = pr.P;
    if (P == 0) { pairs_free(&pr); return 0; }
    int T = orc_threads(nthreads);
    double *zij = (double *)calloc((size_t)P, sizeof(double));
    double *zji = (double *)calloc((size_t)P, sizeof(double));
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t p = 0; p < P; ++p) {
        int32_t ii = pr.ii[p], jj = pr.jj[p];
        int si = type_ids[ii], sj = type_ids[jj];
        if (TB(si, sj, 0) == 0.0) continue;
        double r_ij = pr.r[p];
        const double *v = pr.rij + 3 * p;
        double zs = 0.0;
        for (int64_t nb = pr.nb_start[ii]; nb < pr.nb_start[ii + 1]; ++nb) {
            int32_t kk = pr.nb_idx[nb];
            if (kk == jj) continue;
            int sk = type_ids[kk];
            if (TB(si, sk, 0) == 0.0 || TR(si, sj, sk, 0) == 0.0) continue;
            double r_ik = pr.nb_r[nb];
            double fc_ik = te_fc(r_ik, TB(si, sk, 5), TB(si, sk, 6));
            if (fc_ik == 0.0) continue;
            const double *w = pr.nb_rij + 3 * nb;
            double dot = 0.0;
            for (int x = 0; x < 3; ++x) dot += v[x] * w[x];
            double cs = dot / (r_ij * r_ik);
            double c = TR(si, sj, sk, 1), d = TR(si, sj, sk, 2), h = TR(si, sj, sk, 3);
            double l3 = TR(si, sj, sk, 4), m = TR(si, sj, sk, 5), gam = TR(si, sj, sk, 6);
            double g = 1.0 + c * c / (d * d) - c * c / (d * d + (h - cs) * (h - cs));
            zs += gam * fc_ik * g * nb_exp(pow(l3, m) * pow(r_ij - r_ik, m));
        }
        zij[p] = zs;
        zs = 0.0;
        for (int64_t nb = pr.nb_start[jj]; nb < pr.nb_start[jj + 1]; ++nb) {
            int32_t kk = pr.nb_idx[nb];
            if (kk == ii) continue;
            int sk = type_ids[kk];
            if (TB(sj, sk, 0) == 0.0 || TR(sj, si, sk, 0) == 0.0) continue;
            double r_jk = pr.nb_r[nb];
            double fc_jk = te_fc(r_jk, TB(sj, sk, 5), TB(sj, sk, 6));
            if (fc_jk == 0.0) continue;
            const double *w = pr.nb_rij + 3 * nb;
            double dot = (-v[0]) * w[0] + (-v[1]) * w[1] + (-v[2]) * w[2];
            double cs = dot / (r_ij * r_jk);
            double c = TR(sj, si, sk, 1), d = TR(sj, si, sk, 2), h = TR(sj, si, sk, 3);
            double l3 = TR(sj, si, sk, 4), m = TR(sj, si, sk, 5), gam = TR(sj, si, sk, 6);
            double g = 1.0 + c * c / (d * d) - c * c / (d * d + (h - cs) * (h - cs));
            zs += gam * fc_jk * g * nb_exp(pow(l3, m) * pow(r_ij - r_jk, m));
        }
        zji[p] = zs;
    }
    double *bij = (double *)malloc(sizeof(double) * (size_t)P);
    double *bji = (double *)malloc(sizeof(double) * (size_t)P);
    double *dbi = (double *)malloc(sizeof(double) * (size_t)P);
    double *dbj = (double *)malloc(sizeof(double) * (size_t)P);
    for (int64_t p = 0; p < P; ++p) {
        int si = type_ids[pr.ii[p]], sj = type_ids[pr.jj[p]];
        int valid = TB(si, sj, 0) != 0.0;
        bij[p] = bji[p] = 1.0;
        dbi[p] = dbj[p] = 0.0;
        double be = TB(si, sj, 7), nn = TB(si, sj, 8);
        if (valid && zij[p] > 0 && be != 0.0) {
            double z = zij[p], bz = pow(be * z, nn), b = pow(1.0 + bz, -0.5 / nn);
            bij[p] = b;
            dbi[p] = -0.5 * b * bz / (z * (1.0 + bz));
        }
        be = TB(sj, si, 7);
        nn = TB(sj, si, 8);
        if (valid && zji[p] > 0 && be != 0.0) {
            double z = zji[p], bz = pow(be * z, nn), b = pow(1.0 + bz, -0.5 / nn);
            bji[p] = b;
            dbj[p] = -0.5 * b * bz / (z * (1.0 + bz));
        }
    }
    double *Ft = (double *)calloc((size_t)T * 3 * (size_t)n, sizeof(double));
    double *St = (double *)calloc((size_t)T * 9, sizeof(double));
    double *Et = (double *)calloc((size_t)T, sizeof(double));
#pragma omp parallel num_threads(T)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *Fb = Ft + (size_t)tid * 3 * (size_t)n;
        double *Sb = St + (size_t)tid * 9;
        double eb = 0.0;
#pragma omp for schedule(static)
        for (int64_t p = 0; p < P; ++p) {
            int32_t ii = pr.ii[p], jj = pr.jj[p];
            int si = type_ids[ii], sj = type_ids[jj];
            if (TB(si, sj, 0) == 0.0) continue;
            double A = TB(si, sj, 1), B = TB(si, sj, 2), l1 = TB(si, sj, 3), l2 = TB(si, sj, 4);
            double R = TB(si, sj, 5), D = TB(si, sj, 6);
            double r = pr.r[p];
            const double *v = pr.rij + 3 * p;
            double fc = te_fc(r, R, D), dfc = te_dfc(r, R, D);
            if (fc == 0.0 && dfc == 0.0) continue;
            double fr = A * nb_exp(-l1 * r), fa = -B * nb_exp(-l2 * r);
            double bavg = 0.5 * (bij[p] + bji[p]);
            eb += fc * (fr + bavg * fa);
            double dEdr = dfc * (fr + bavg * fa) + fc * (-l1 * fr + bavg * (-l2 * fa));
            double cx = dEdr / r;
            for (int x = 0; x < 3; ++x) {
                Fb[3 * (size_t)ii + x] += cx * v[x];
                Fb[3 * (size_t)jj + x] += -cx * v[x];
            }
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 3; ++y) Sb[3 * x + y] -= cx * v[x] * v[y];
            if (fabs(dbi[p]) > 1e-30) {
                double coeff = 0.5 * fc * fa * dbi[p];
                for (int64_t nb = pr.nb_start[ii]; nb < pr.nb_start[ii + 1]; ++nb) {
                    int32_t kk = pr.nb_idx[nb];
                    if (kk == jj) continue;
                    int sk = type_ids[kk];
                    if (TB(si, sk, 0) == 0.0 || TR(si, sj, sk, 0) == 0.0) continue;
                    double Rk = TB(si, sk, 5), Dk = TB(si, sk, 6);
                    if (te_fc(pr.nb_r[nb], Rk, Dk) == 0.0) continue;
                    const double *w = pr.nb_rij + 3 * nb;
                    double gj[3], gk[3];
                    te_zterm_grad(v, w, Rk, Dk, TR(si, sj, sk, 1), TR(si, sj, sk, 2),
                                  TR(si, sj, sk, 3), TR(si, sj, sk, 4), TR(si, sj, sk, 5),
                                  TR(si, sj, sk, 6), gj, gk);
                    for (int x = 0; x < 3; ++x) {
                        double Fj = -coeff * gj[x], Fk = -coeff * gk[x];
                        Fb[3 * (size_t)ii + x] += coeff * (gj[x] + gk[x]);
                        Fb[3 * (size_t)jj + x] += Fj;
                        Fb[3 * (size_t)kk + x] += Fk;
                        for (int y = 0; y < 3; ++y) Sb[3 * x + y] += Fj * v[y] + Fk * w[y];
                    }
                }
            }
            if (fabs(dbj[p]) > 1e-30) {
                double coeff = 0.5 * fc * fa * dbj[p];
                double vji[3] = {-v[0], -v[1], -v[2]};
                for (int64_t nb = pr.nb_start[jj]; nb < pr.nb_start[jj + 1]; ++nb) {
                    int32_t kk = pr.nb_idx[nb];
                    if (kk == ii) continue;
                    int sk = type_ids[kk];
                    if (TB(sj, sk, 0) == 0.0 || TR(sj, si, sk, 0) == 0.0) continue;
                    double Rk = TB(sj, sk, 5), Dk = TB(sj, sk, 6);
                    if (te_fc(pr.nb_r[nb], Rk, Dk) == 0.0) continue;
                    const double *w = pr.nb_rij + 3 * nb;
                    double gi[3], gk[3];
                    te_zterm_grad(vji, w, Rk, Dk, TR(sj, si, sk, 1), TR(sj, si, sk, 2),
                                  TR(sj, si, sk, 3), TR(sj, si, sk, 4), TR(sj, si, sk, 5),
                                  TR(sj, si, sk, 6), gi, gk);
                    for (int x = 0; x < 3; ++x) {
                        double Fi = -coeff * gi[x], Fk = -coeff * gk[x];
                        Fb[3 * (size_t)jj + x] += coeff * (gi[x] + gk[x]);
                        Fb[3 * (size_t)ii + x] += Fi;
                        Fb[3 * (size_t)kk + x] += Fk;
                        for (int y = 0; y < 3; ++y) Sb[3 * x + y] += Fi * vji[y] + Fk * w[y];
                    }
                }
            }
        }
        Et[tid] = eb;
    }
    double e = 0.0, S[9] = {0};
    for (int t = 0; t < T; ++t) {
        e += Et[t];
        for (size_t k = 0; k < 3 * (size_t)n; ++k) forces[k] += Ft[(size_t)t * 3 * (size_t)n + k];
        for (int k = 0; k < 9; ++k) S[k] += St[(size_t)t * 9 + k];
    }
    for (int k = 0; k < 9; ++k) stress[k] = S[k] / volume;
    *energy = e;
    free(Ft); free(St); free(Et); free(zij); free(zji); free(bij); free(bji);
    free(dbi); free(dbj);
    pairs_free(&pr);
    return 0;
}
#undef TB
#undef TR

/* ------------------------------------------------------------------------
 * Stillinger-Weber, flat single-species path:
 * engine/stillinger_weber.py:927-1159 with _sw_2body_numpy (:23-52),
 * _sw_h3_grad (:96-154) and _sw_3body_numpy (:157-255).
 * prm = epsilon, sigma, A, B, p, q, a, lambda, gamma, costheta0.
 * ---------------------------------------------------------------------- */
static inline void sw_h3_grad(const double *rij, const double *rik, double sig, double a,
                              double gam, double c3, double costh0, double *h, double *gj,
                              double *gk)
{
    double r_ij = sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
    double r_ik = sqrt(rik[0] * rik[0] + rik[1] * rik[1] + rik[2] * rik[2]);
    double xij = r_ij / sig, xik = r_ik / sig;
    if (xij >= a || xik >= a) {
        *h = 0.0;
        gj[0] = gj[1] = gj[2] = gk[0] = gk[1] = gk[2] = 0.0;
        return;
    }
    double dj = xij - a, dk = xik - a;
    double eij = nb_exp(gam / dj), eik = nb_exp(gam / dk);
    double cs = (rij[0] * rik[0] + rij[1] * rik[1] + rij[2] * rik[2]) / (r_ij * r_ik);
    double dc = cs - costh0;
    *h = c3 * eij * eik * dc * dc;
    double pre = c3 * eij * eik * dc;
    double inv_ij = 1.0 / r_ij, inv_ik = 1.0 / r_ik;
    double uij[3] = {rij[0] * inv_ij, rij[1] * inv_ij, rij[2] * inv_ij};
    double uik[3] = {rik[0] * inv_ik, rik[1] * inv_ik, rik[2] * inv_ik};
    double rad_j = pre * dc * (-gam / (sig * dj * dj));
    double rad_k = pre * dc * (-gam / (sig * dk * dk));
    double ang_j = 2.0 * pre * inv_ij, ang_k = 2.0 * pre * inv_ik;
    for (int x = 0; x < 3; ++x) {
        gj[x] = rad_j * uij[x] + ang_j * (uik[x] - cs * uij[x]);
        gk[x] = rad_k * uik[x] + ang_k * (uij[x] - cs * uik[x]);
    }
}

int orc_sw_flat(int32_t n, const double *pos, const double *lat, const double *inv,
                double volume, int64_t ncand, const int32_t *ci, const int32_t *cj,
                const double *prm, int nthreads, double *energy, double *forces,
                double *stress)
{
    double eps = prm[0], sig = prm[1], Asw = prm[2], Bsw = prm[3], pp = prm[4], qq = prm[5];
    double a = prm[6], lam = prm[7], gam = prm[8], costh0 = prm[9];
    double rc = a * sig;
    memset(forces, 0, sizeof(double) * 3 * (size_t)n);
    memset(stress, 0, sizeof(double) * 9);
    *energy = 0.0;
    pairs_t pr;
    pairs_make(&pr, n, pos, lat, inv, ncand, ci, cj, rc);
    if (pr.P == 0) { pairs_free(&pr); return 0; }
    int T = orc_threads(nthreads);
    double e = 0.0, S[9] = {0};
    /* _sw_2body_numpy */
    for (int64_t p = 0; p < pr.P; ++p) {
        double r = pr.r[p], xi = r / sig;
        if (xi >= a) continue;
        double t = 1.0 / (xi - a);
        double ex = nb_exp(t);
        double V2 = eps * Asw * (Bsw * pow(xi, -pp) - pow(xi, -qq)) * ex;
        double dV2 = eps * Asw * ex * ((-pp * Bsw * pow(xi, -pp - 1) + qq * pow(xi, -qq - 1)) +
                                       (Bsw * pow(xi, -pp) - pow(xi, -qq)) * (-t * t));
        double cx = (dV2 / sig) / r;
        e += V2;
        const double *v = pr.rij + 3 * p;
        for (int x = 0; x < 3; ++x) {
            forces[3 * (size_t)pr.ii[p] + x] += cx * v[x];
            forces[3 * (size_t)pr.jj[p] + x] += -cx * v[x];
        }
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) S[3 * x + y] -= cx * v[x] * v[y];
    }
    /* _sw_3body_numpy over centres, per-thread buffers as its prange twin */
    double c3 = lam * eps;
    double *Ft = (double *)calloc((size_t)T * 3 * (size_t)n, sizeof(double));
    double *St = (double *)calloc((size_t)T * 9, sizeof(double));
    double *Et = (double *)calloc((size_t)T, sizeof(double));
#pragma omp parallel num_threads(T)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *Fb = Ft + (size_t)tid * 3 * (size_t)n;
        double *Sb = St + (size_t)tid * 9;
        double eb = 0.0;
#pragma omp for schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            int64_t s0 = pr.nb_start[i], s1 = pr.nb_start[i + 1];
            if (s1 - s0 < 2) continue;
            for (int64_t pj = s0; pj < s1; ++pj) {
                if (pr.nb_r[pj] / sig >= a) continue;
                const double *vj = pr.nb_rij + 3 * pj;
                for (int64_t pk = pj + 1; pk < s1; ++pk) {
                    if (pr.nb_r[pk] / sig >= a) continue;
                    const double *vk = pr.nb_rij + 3 * pk;
                    double h, gj[3], gk[3];
                    sw_h3_grad(vj, vk, sig, a, gam, c3, costh0, &h, gj, gk);
                    eb += h;
                    int32_t ja = pr.nb_idx[pj], ka = pr.nb_idx[pk];
                    for (int x = 0; x < 3; ++x) {
                        Fb[3 * (size_t)i + x] += gj[x] + gk[x];
                        Fb[3 * (size_t)ja + x] += -gj[x];
                        Fb[3 * (size_t)ka + x] += -gk[x];
                        for (int y = 0; y < 3; ++y) Sb[3 * x + y] += -gj[x] * vj[y] - gk[x] * vk[y];
                    }
                }
            }
        }
        Et[tid] = eb;
    }
    for (int t = 0; t < T; ++t) {
        e += Et[t];
        for (size_t k = 0; k < 3 * (size_t)n; ++k) forces[k] += Ft[(size_t)t * 3 * (size_t)n + k];
        for (int k = 0; k < 9; ++k) S[k] += St[(size_t)t * 9 + k];
    }
    for (int k = 0; k < 9; ++k) stress[k] = S[k] / volume;
    *energy = e;
    free(Ft); free(St); free(Et);
    pairs_free(&pr);
    return 0;
}

/* ------------------------------------------------------------------------
 * Stillinger-Weber, multi-species table path:
 * StillingerWeberPotential._compute_multi_species
 * (engine/stillinger_weber.py:1165-1451) with _sw_multi_2body_numpy
 * (:414-447) and _sw_multi_3body_numpy (:476-560), _sw_h3_grad (:96-154).
 *   tb  (ns, ns, 8): valid, epsilon, sigma, a, A, B, p, q of the two-body
 *       record (i, j, j), valid when it exists and A != 0;
 *   leg (ns, ns, 3): sigma, a, gamma of the leg record (i, j, j);
 *   tri (ns, ns, ns, 4): valid, epsilon, lambda, costheta0 of (i, j, k),
 *       valid when (i, j, k), (i, j, j) and (i, k, k) exist.
 * ---------------------------------------------------------------------- */
static inline void sw_h3_grad_legs(const double *rij, const double *rik, double sig_j,
                                   double a_j, double gam_j, double sig_k, double a_k,
                                   double gam_k, double c3, double costh0, double *h,
                                   double *gj, double *gk)
{
    double r_ij = sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
    double r_ik = sqrt(rik[0] * rik[0] + rik[1] * rik[1] + rik[2] * rik[2]);
    double xij = r_ij / sig_j, xik = r_ik / sig_k;
    if (xij >= a_j || xik >= a_k) {
        *h = 0.0;
        gj[0] = gj[1] = gj[2] = gk[0] = gk[1] = gk[2] = 0.0;
        return;
    }
    double dj = xij - a_j, dk = xik - a_k;
    double eij = nb_exp(gam_j / dj), eik = nb_exp(gam_k / dk);
    double cs = (rij[0] * rik[0] + rij[1] * rik[1] + rij[2] * rik[2]) / (r_ij * r_ik);
    double dc = cs - costh0;
    *h = c3 * eij * eik * dc * dc;
    double pre = c3 * eij * eik * dc;
    double inv_ij = 1.0 / r_ij, inv_ik = 1.0 / r_ik;
    double uij[3] = {rij[0] * inv_ij, rij[1] * inv_ij, rij[2] * inv_ij};
    double uik[3] = {rik[0] * inv_ik, rik[1] * inv_ik, rik[2] * inv_ik};
    double rad_j = pre * dc * (-gam_j / (sig_j * dj * dj));
    double rad_k = pre * dc * (-gam_k / (sig_k * dk * dk));
    double ang_j = 2.0 * pre * inv_ij, ang_k = 2.0 * pre * inv_ik;
    for (int x = 0; x < 3; ++x) {
        gj[x] = rad_j * uij[x] + ang_j * (uik[x] - cs * uij[x]);
        gk[x] = rad_k * uik[x] + ang_k * (uij[x] - cs * uik[x]);
    }
}

int orc_sw_multi(int32_t n, const double *pos, const double *lat, const double *inv,
                 double volume, int64_t ncand, const int32_t *ci, const int32_t *cj,
                 int32_t ns, const int32_t *type_ids, const double *tb, const double *leg,
                 const double *tri, double rc, int nthreads, double *energy, double *forces,
                 double *stress)
{
    memset(forces, 0, sizeof(double) * 3 * (size_t)n);
    memset(stress, 0, sizeof(double) * 9);
    *energy = 0.0;
    pairs_t pr;
    pairs_make(&pr, n, pos, lat, inv, ncand, ci, cj, rc);
    if (pr.P == 0) { pairs_free(&pr); return 0; }
    int T = orc_threads(nthreads);
    double e = 0.0, S[9] = {0};
    /* _sw_multi_2body_numpy: each pair i < j once, record (s_i, s_j) */
    for (int64_t p = 0; p < pr.P; ++p) {
        int si = type_ids[pr.ii[p]], sj = type_ids[pr.jj[p]];
        const double *t2 = tb + 8 * ((size_t)si * ns + sj);
        if (t2[0] == 0.0) continue;
        double eps = t2[1], sig = t2[2], a = t2[3], Asw = t2[4], Bsw = t2[5], pp = t2[6],
               qq = t2[7];
        double r = pr.r[p], xi = r / sig;
        if (xi >= a) continue;
        double t = 1.0 / (xi - a);
        double ex = nb_exp(t);
        double V2 = eps * Asw * (Bsw * pow(xi, -pp) - pow(xi, -qq)) * ex;
        double dV2 = eps * Asw * ex * ((-pp * Bsw * pow(xi, -pp - 1) + qq * pow(xi, -qq - 1)) +
                                       (Bsw * pow(xi, -pp) - pow(xi, -qq)) * (-t * t));
        double cx = (dV2 / sig) / r;
        e += V2;
        const double *v = pr.rij + 3 * p;
        for (int x = 0; x < 3; ++x) {
            forces[3 * (size_t)pr.ii[p] + x] += cx * v[x];
            forces[3 * (size_t)pr.jj[p] + x] += -cx * v[x];
        }
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) S[3 * x + y] -= cx * v[x] * v[y];
    }
    /* _sw_multi_3body_numpy: ordered (j, k), each record weighted 1/2 */
    double *Ft = (double *)calloc((size_t)T * 3 * (size_t)n, sizeof(double));
    double *St = (double *)calloc((size_t)T * 9, sizeof(double));
    double *Et = (double *)calloc((size_t)T, sizeof(double));
#pragma omp parallel num_threads(T)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *Fb = Ft + (size_t)tid * 3 * (size_t)n;
        double *Sb = St + (size_t)tid * 9;
        double eb = 0.0;
#pragma omp for schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            int64_t s0 = pr.nb_start[i], s1 = pr.nb_start[i + 1];
            if (s1 - s0 < 2) continue;
            int isp = type_ids[i];
            for (int64_t pj = s0; pj < s1; ++pj) {
                int32_t ja = pr.nb_idx[pj];
                int jsp = type_ids[ja];
                const double *vj = pr.nb_rij + 3 * pj;
                for (int64_t pk = s0; pk < s1; ++pk) {
                    if (pk == pj) continue;
                    int32_t ka = pr.nb_idx[pk];
                    int ksp = type_ids[ka];
                    const double *t3 = tri + 4 * (((size_t)isp * ns + jsp) * ns + ksp);
                    if (t3[0] == 0.0) continue;
                    const double *lj = leg + 3 * ((size_t)isp * ns + jsp);
                    const double *lk = leg + 3 * ((size_t)isp * ns + ksp);
                    const double *vk = pr.nb_rij + 3 * pk;
                    double eps_t = t3[1], lam_t = 0.5 * t3[2], cos_t = t3[3];
                    double h, gj[3], gk[3];
                    sw_h3_grad_legs(vj, vk, lj[0], lj[1], lj[2], lk[0], lk[1], lk[2],
                                    eps_t * lam_t, cos_t, &h, gj, gk);
                    if (h == 0.0) continue;
                    eb += h;
                    for (int x = 0; x < 3; ++x) {
                        Fb[3 * (size_t)i + x] += gj[x] + gk[x];
                        Fb[3 * (size_t)ja + x] += -gj[x];
                        Fb[3 * (size_t)ka + x] += -gk[x];
                        for (int y = 0; y < 3; ++y) Sb[3 * x + y] += -gj[x] * vj[y] - gk[x] * vk[y];
                    }
                }
            }
        }
        Et[tid] = eb;
    }
    for (int t = 0; t < T; ++t) {
        e += Et[t];
        for (size_t k = 0; k < 3 * (size_t)n; ++k) forces[k] += Ft[(size_t)t * 3 * (size_t)n + k];
        for (int k = 0; k < 9; ++k) S[k] += St[(size_t)t * 9 + k];
    }
    for (int k = 0; k < 9; ++k) stress[k] = S[k] / volume;
    *energy = e;
    free(Ft); free(St); free(Et);
    pairs_free(&pr);
    return 0;
}

/* ------------------------------------------------------------------------
 * Zhou-Johnson-Wadley EAM: engine/_eam_kernels.py:148-223 (evaluators) and
 * :231-307 (_eam_kernel, four phases).  A function is a row of ZW doubles:
 * [code, rc, width, chi, pa[20], pb[20]] with code 3 = density,
 * 4 = pair, 5 = cross, 6 = embedding; param order r_e, f_e, rho_e, rho_s,
 * alpha, beta, A, B, kappa, lambda, Fn0..Fn3, F0..F3, eta, F_e.
 * ---------------------------------------------------------------------- */
#define ZW 44

static inline void z_decay(double r, double re, double pre, double expo, double shift,
                           double *val, double *dval)
{
    double x = r / re;
    double e = pre * exp(-expo * (x - 1.0));
    double u = x - shift;
    double den = 1.0 + pow(u, 20.0);
    *val = e / den;
    *dval = (-expo * e / den - e * 20.0 * pow(u, 19.0) / (den * den)) / re;
}

static inline void z_dens(const double *p, double r, double *v, double *dv)
{
    z_decay(r, p[0], p[1], p[5], p[9], v, dv);
}

static inline void z_pair(const double *p, double r, double *v, double *dv)
{
    double a, da, b, db;
    z_decay(r, p[0], p[6], p[4], p[8], &a, &da);
    z_decay(r, p[0], p[7], p[5], p[9], &b, &db);
    *v = a - b;
    *dv = da - db;
}

static inline void z_embed(const double *p, double rho, double *F, double *dF)
{
    double re = p[2], rs = p[3], eta = p[18];
    double rn = 0.85 * re;
    if (rho < rn) {
        double z = rho / rn - 1.0;
        *F = p[10] + p[11] * z + p[12] * z * z + p[13] * pow(z, 3.0);
        *dF = (p[11] + 2 * p[12] * z + 3 * p[13] * z * z) / rn;
    } else if (rho >= 1.15 * re) {
        double t = rho / rs, lt = log(t);
        *F = p[19] * (1.0 - eta * lt) * pow(t, eta);
        *dF = -p[19] * eta * eta * lt * pow(t, eta - 1.0) / rs;
    } else {
        double z = rho / re - 1.0;
        *F = p[14] + p[15] * z + p[16] * z * z + p[17] * pow(z, 3.0);
        *dF = (p[15] + 2 * p[16] * z + 3 * p[17] * z * z) / re;
    }
}

/* Tabulated functions (engine/_eam_kernels.py:77-121 build_tables): function
 * f with code 1 (spline) or 2 (spline / r) owns knots X[xo[f] .. xo[f] +
 * nx[f]) and the PPoly coefficients C[0..3][same range] (C has NX columns). */
typedef struct {
    const double *fn;
    const int64_t *xo, *nx;
    const double *X, *C;
    int64_t NX;
} eam_tab_t;

/* engine/_eam_kernels.py:128-145 */
static inline void spline_eval(const eam_tab_t *T, int f, double r, double *v, double *dv)
{
    int64_t o = T->xo[f], n = T->nx[f];
    int64_t lo = 0, hi = n;
    while (lo < hi) {  /* searchsorted side='right' */
        int64_t mid = (lo + hi) / 2;
        if (T->X[o + mid] <= r) lo = mid + 1;
        else hi = mid;
    }
    int64_t i = lo - 1;
    if (i < 0) i = 0;
    if (i > n - 2) i = n - 2;
    double dx = r - T->X[o + i];
    double c0 = T->C[o + i], c1 = T->C[T->NX + o + i], c2 = T->C[2 * T->NX + o + i],
           c3 = T->C[3 * T->NX + o + i];
    *v = ((c0 * dx + c1) * dx + c2) * dx + c3;
    *dv = (3.0 * c0 * dx + 2.0 * c1) * dx + c2;
}

static inline void z_feval_row(const double *f, double r, double *v, double *dv)
{
    int code = (int)f[0];
    if (code == 6) { z_embed(f + 4, r, v, dv); return; }
    double val, dval;
    if (code == 3) {
        z_dens(f + 4, r, &val, &dval);
    } else if (code == 4) {
        z_pair(f + 4, r, &val, &dval);
    } else {
        double fa, dfa, fb, dfb, ua, dua, ub, dub;
        z_dens(f + 4, r, &fa, &dfa);
        z_dens(f + 24, r, &fb, &dfb);
        z_pair(f + 4, r, &ua, &dua);
        z_pair(f + 24, r, &ub, &dub);
        double k = f[3];
        double q = fb / fa;
        double dq = (dfb * fa - fb * dfa) / (fa * fa);
        val = 0.5 * k * (q * ua + ub / q);
        dval = 0.5 * k * (dq * ua + q * dua + dub / q - ub * dq / (q * q));
    }
    double rc = f[1], w = f[2];
    double s = (r - (rc - w)) / w;
    if (s < 0.0) s = 0.0;
    else if (s > 1.0) s = 1.0;
    double S = 1.0 - s * s * s * (10.0 - 15.0 * s + 6.0 * s * s);
    double dS = -30.0 * s * s * (1.0 - s) * (1.0 - s) / w;
    *v = val * S;
    *dv = dval * S + val * dS;
}

/* engine/_eam_kernels.py:199-205 (_feval, spline codes) and the Zhou rows */
static inline void z_feval(const eam_tab_t *T, int f, double r, double *v, double *dv)
{
    int code = (int)T->fn[(size_t)ZW * f];
    if (code == 1 || code == 2) {
        double s_, ds;
        spline_eval(T, f, r, &s_, &ds);
        if (code == 1) { *v = s_; *dv = ds; }
        else { *v = s_ / r; *dv = (ds * r - s_) / (r * r); }
        return;
    }
    z_feval_row(T->fn + (size_t)ZW * f, r, v, dv);
}

/* rho_f[s], emb_f[s], phi_f[s*ns+t]: function row indices into fn (ZW each);
 * xo, nx, X, C, NX: the spline tables (unused when no row has code 1 or 2) */
int orc_eam(int32_t n, const double *pos, const double *lat, const double *inv, double volume,
            int64_t ncand, const int32_t *ci, const int32_t *cj, const int32_t *spc,
            int32_t ns, const int32_t *rho_f, const int32_t *emb_f, const int32_t *phi_f,
            const double *fn, const int64_t *xo, const int64_t *nx, const double *X,
            const double *C, int64_t NX, double rc, int nthreads, double *energy,
            double *forces, double *stress)
{
    eam_tab_t TT = {fn, xo, nx, X, C, NX};
    const eam_tab_t *Tb = &TT;
    memset(forces, 0, sizeof(double) * 3 * (size_t)n);
    memset(stress, 0, sizeof(double) * 9);
    *energy = 0.0;
    pairs_t pr;
    pairs_make(&pr, n, pos, lat, inv, ncand, ci, cj, rc);
    int64_t P = pr.P;
    if (P == 0) { pairs_free(&pr); return 0; }
    int T = orc_threads(nthreads);
    double *rho_i = (double *)malloc(sizeof(double) * (size_t)P);
    double *drho_j = (double *)malloc(sizeof(double) * (size_t)P);
    double *rho_j = (double *)malloc(sizeof(double) * (size_t)P);
    double *drho_i = (double *)malloc(sizeof(double) * (size_t)P);
    double *phi = (double *)malloc(sizeof(double) * (size_t)P);
    double *dphi = (double *)malloc(sizeof(double) * (size_t)P);
    /* phase 1: per half pair */
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t p = 0; p < P; ++p) {
        int si = spc[pr.ii[p]], sj = spc[pr.jj[p]];
        double r = pr.r[p];
        z_feval(Tb, rho_f[sj], r, &rho_i[p], &drho_j[p]);
        z_feval(Tb, rho_f[si], r, &rho_j[p], &drho_i[p]);
        z_feval(Tb, phi_f[si * ns + sj], r, &phi[p], &dphi[p]);
    }
    /* two-way CSR: atom -> its pairs with side +1 (a = i) or -1 (a = j) */
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t p = 0; p < P; ++p) { cnt[pr.ii[p] + 1]++; cnt[pr.jj[p] + 1]++; }
    for (int32_t a = 0; a < n; ++a) cnt[a + 1] += cnt[a];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    memcpy(fill, cnt, sizeof(int64_t) * ((size_t)n + 1));
    int64_t *lst = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * P));
    int8_t *side = (int8_t *)malloc((size_t)(2 * P));
    for (int64_t p = 0; p < P; ++p) {
        int64_t s = fill[pr.ii[p]]++;
        lst[s] = p; side[s] = 1;
        s = fill[pr.jj[p]]++;
        lst[s] = p; side[s] = -1;
    }
    double *Fv = (double *)malloc(sizeof(double) * (size_t)n);
    double *dF = (double *)malloc(sizeof(double) * (size_t)n);
    /* phase 2: per atom rho_bar, F, F' */
#pragma omp parallel for num_threads(T) schedule(static)
    for (int32_t a = 0; a < n; ++a) {
        double rb = 0.0;
        for (int64_t k = cnt[a]; k < cnt[a + 1]; ++k) {
            int64_t p = lst[k];
            rb += side[k] == 1 ? rho_i[p] : rho_j[p];
        }
        if (rb < 0.0) rb = 0.0;
        z_feval(Tb, emb_f[spc[a]], rb, &Fv[a], &dF[a]);
    }
    /* phase 3: per pair force */
    double *fij = (double *)malloc(sizeof(double) * 3 * (size_t)P);
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t p = 0; p < P; ++p) {
        double c = dphi[p] + dF[pr.ii[p]] * drho_j[p] + dF[pr.jj[p]] * drho_i[p];
        double s = c / pr.r[p];
        for (int x = 0; x < 3; ++x) fij[3 * p + x] = s * pr.rij[3 * p + x];
    }
    /* phase 4: per atom gather */
#pragma omp parallel for num_threads(T) schedule(static)
    for (int32_t a = 0; a < n; ++a) {
        double f[3] = {0, 0, 0};
        for (int64_t k = cnt[a]; k < cnt[a + 1]; ++k) {
            int64_t p = lst[k];
            for (int x = 0; x < 3; ++x) f[x] += side[k] * fij[3 * p + x];
        }
        for (int x = 0; x < 3; ++x) forces[3 * (size_t)a + x] = f[x];
    }
    double vir[9] = {0};
    for (int64_t p = 0; p < P; ++p)
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) vir[3 * x + y] -= fij[3 * p + x] * pr.rij[3 * p + y];
    double e = 0.0, ephi = 0.0;
    for (int32_t a = 0; a < n; ++a) e += Fv[a];
    for (int64_t p = 0; p < P; ++p) ephi += phi[p];
    *energy = e + ephi;
    for (int k = 0; k < 9; ++k) stress[k] = vir[k] / volume;
    free(rho_i); free(drho_j); free(rho_j); free(drho_i); free(phi); free(dphi);
    free(cnt); free(fill); free(lst); free(side); free(Fv); free(dF); free(fij);
    pairs_free(&pr);
    return 0;
}

/* One function value, for host-side checks of the Zhou construction. */
void orc_zhou_feval(const double *f, int64_t m, const double *r, double *v, double *dv)
{
    for (int64_t k = 0; k < m; ++k) z_feval_row(f, r[k], &v[k], &dv[k]);
}

/* Function f of a table (any code) at m arguments. */
void orc_eam_feval(const double *fn, const int64_t *xo, const int64_t *nx, const double *X,
                   const double *C, int64_t NX, int32_t f, int64_t m, const double *r,
                   double *v, double *dv)
{
    eam_tab_t TT = {fn, xo, nx, X, C, NX};
    for (int64_t k = 0; k < m; ++k) z_feval(&TT, f, r[k], &v[k], &dv[k]);
}
