/*
 * FP64 CPU oracle for the sTDA / sTD-DFT excited-state core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, loop nest by loop nest, the
 * reference's algorithm for the hot path (grimme-lab/stda, std2 v2.0) in
 * double precision.  It is used by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py as the checker and as the CPU baseline.
 * The product (stda_amd, libstda_hip.so) never links, loads or calls it.
 *
 * Differences from the reference, all deliberate:
 *   - every quantity the reference stores in real*4 (gamma, q, p, A, A+-B,
 *     0.5B) is kept in real*8 here, so this is an FP64 restatement;
 *   - the PT2 reduction over secondaries is accumulated per thread and
 *     summed at the end (the reference uses an OpenMP reduction too, so its
 *     summation order is also thread-count dependent).
 *
 * Layouts follow the reference: column-major, atom index fastest for the
 * charge arrays q(ncent, pair); occupied-virtual pair index (i-1)*nv + a;
 * packed pair index lin(i,j) = min + max*(max-1)/2 (1-based, stda.f:2174-2181).
 * All indices at this C interface are 0-based except iaoat and iconf, which
 * keep the reference's 1-based convention (iconf(.,2) holds the absolute
 * virtual MO index no+1..moci).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* hardness table of setrep, stda.f:2060-2167 */
static const double ETA[94] = {
    0.472592880, 0.922033910, 0.174528880, 0.257007330, 0.339490860,
    0.421954120, 0.504381930, 0.586918630, 0.669313510, 0.751916070,
    0.179641050, 0.221572760, 0.263485780, 0.305396450, 0.347340140,
    0.389247250, 0.431156700, 0.473082690, 0.171054690, 0.202762440,
    0.210073220, 0.217396470, 0.224710390, 0.232015010, 0.239339690,
    0.246656380, 0.253982550, 0.261288630, 0.268594760, 0.275925650,
    0.307629990, 0.339315800, 0.372359850, 0.402735490, 0.434457760,
    0.466117080, 0.155850790, 0.186493240, 0.193562100, 0.200633110,
    0.207705220, 0.214772540, 0.221846140, 0.228918720, 0.235986210,
    0.243056120, 0.250130180, 0.257199370, 0.287847800, 0.318486730,
    0.349124310, 0.379765930, 0.410408080, 0.441057770, 0.050193320,
    0.067625700, 0.085044450, 0.102477360, 0.119911050, 0.137327720,
    0.154762970, 0.172182650, 0.189612880, 0.207047600, 0.224467520,
    0.241896450, 0.259325030, 0.276760940, 0.294182310, 0.311595870,
    0.329022740, 0.345922980, 0.363880480, 0.381305860, 0.398774760,
    0.416142980, 0.433645100, 0.451040140, 0.468489860, 0.485845500,
    0.125267300, 0.142686770, 0.160116150, 0.177558890, 0.194975570,
    0.212407780, 0.072635250, 0.094221580, 0.099202950, 0.104186210,
    0.142356330, 0.163942940, 0.185519410, 0.223701390};

static inline int64_t lin(int64_t i1, int64_t i2) { /* 1-based, stda.f:2174 */
    int64_t a = i1 > i2 ? i1 : i2, b = i1 > i2 ? i2 : i1;
    return b + a * (a - 1) / 2;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* gamma^J and gamma^K, stda.f:509-532.  xyz(4,ncent) Bohr + Z. */
void orc_gamma(int ncent, const double *xyz, double ax, double alphak,
               double betaj, double *gamj, double *gamk) {
    for (int i = 0; i < ncent; ++i) {
        int ii = (int)xyz[4 * i + 3];
        for (int j = 0; j <= i; ++j) {
            int jj = (int)xyz[4 * j + 3];
            double xj = 0.5 * (ETA[ii - 1] + ETA[jj - 1]) * ax;
            double xk = 0.5 * (ETA[ii - 1] + ETA[jj - 1]);
            double dx = xyz[4 * i] - xyz[4 * j], dy = xyz[4 * i + 1] - xyz[4 * j + 1],
                   dz = xyz[4 * i + 2] - xyz[4 * j + 2];
            double rabx = sqrt(dx * dx + dy * dy + dz * dz);
            /* 1/xj**betaj is +Inf for ax = 0, which makes gamma^J exactly 0 */
            double gj = 1.0 / pow(pow(rabx, betaj) + 1.0 / pow(xj, betaj), 1.0 / betaj);
            double gk = 1.0 / pow(pow(rabx, alphak) + 1.0 / pow(xk, alphak), 1.0 / alphak);
            gamj[j + (int64_t)ncent * i] = gj;
            gamj[i + (int64_t)ncent * j] = gj;
            gamk[j + (int64_t)ncent * i] = gk;
            gamk[i + (int64_t)ncent * j] = gk;
        }
    }
}

/* lo12pop, stda.f:1727-1744: q(A) = sum_{k on A} ca(k,i) ca(k,j); i, j 0-based */
static inline void lo12pop(int i, int j, int ncent, int nao, const int *iaoat,
                           const double *ca, double *q) {
    for (int a = 0; a < ncent; ++a) q[a] = 0.0;
    const double *ci = ca + (int64_t)i * nao, *cj = ca + (int64_t)j * nao;
    for (int k = 0; k < nao; ++k) q[iaoat[k] - 1] += ci[k] * cj[k];
}

/* q_ii (ncent,no) and q_aa (ncent,nv): stda.f:465-479 */
void orc_charges_diag(int ncent, int nao, int no, int nv, const int *iaoat,
                      const double *clow, double *qii, double *qaa) {
    for (int i = 0; i < no; ++i) lo12pop(i, i, ncent, nao, iaoat, clow, qii + (int64_t)ncent * i);
    for (int a = 0; a < nv; ++a)
        lo12pop(no + a, no + a, ncent, nao, iaoat, clow, qaa + (int64_t)ncent * a);
}

/* q_ia over all no*nv pairs, pair index i*nv + a: stda.f:576-587 */
void orc_charges_ia(int ncent, int nao, int no, int nv, const int *iaoat,
                    const double *clow, double *qia) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < no; ++i)
        for (int a = 0; a < nv; ++a)
            lo12pop(i, no + a, ncent, nao, iaoat, clow,
                    qia + (int64_t)ncent * ((int64_t)i * nv + a));
}

/* q_ij packed over occupied pairs and q_ab packed over virtual pairs:
 * stda.f:728-763.  Packed column lin(i,j)-1. */
void orc_charges_packed(int ncent, int nao, int no, int nv, const int *iaoat,
                        const double *clow, double *qij, double *qab) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < no; ++i)
        for (int j = 0; j <= i; ++j)
            lo12pop(i, j, ncent, nao, iaoat, clow, qij + (int64_t)ncent * (lin(i + 1, j + 1) - 1));
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < nv; ++k)
        for (int l = 0; l <= k; ++l)
            lo12pop(no + k, no + l, ncent, nao, iaoat, clow,
                    qab + (int64_t)ncent * (lin(k + 1, l + 1) - 1));
}

/* p = gam * q for n columns of length ncent (ssymm 'l','l', stda.f:558,589,742) */
void orc_symm(int ncent, int64_t n, const double *gam, const double *q, double *p) {
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        const double *qc = q + (int64_t)ncent * c;
        double *pc = p + (int64_t)ncent * c;
        for (int a = 0; a < ncent; ++a) {
            double s = 0.0;
            const double *g = gam + (int64_t)ncent * a; /* symmetric: row a == column a */
            for (int b = 0; b < ncent; ++b) s += g[b] * qc[b];
            pc[a] = s;
        }
    }
}

static inline double dot(int n, const double *x, const double *y) {
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += x[k] * y[k];
    return s;
}

/* kshift_to_ediag, stda.f:1750-1768 */
static inline double kshift(double de, double ek, double smax, double width, double steep) {
    double sau = smax * 0.036749320, wau = width * 0.036749320;
    double shft = 1.0 + pow(ek / wau, steep);
    return de + sau / shft;
}

/*
 * CSF screening and ordering, stda.f:593-668.
 *   jdiag(nv,no): (ii|aa)^J, the sgemm of stda.f:564 (uci(l,io))
 *   pia, qia (ncent, no*nv)
 * Outputs (capacity no*nv each): iconf(mx,2) / ed for primaries (sorted by
 * the reference's selection sort), kconf(mx,2) / edpt for secondaries.
 */
void orc_screen(int ncent, int no, int nv, const double *epsi, const double *jdiag,
                const double *pia, const double *qia, double ak, double thr, double fthr,
                int dokshift, double smax, double width, double steep, int64_t mx,
                int *iconf, double *ed, int *kconf, double *edpt, int64_t *nex_out,
                int64_t *nexpt_out) {
    int64_t k = 0, j = 0;
    for (int io = 0; io < no; ++io) {
        for (int l = 0; l < nv; ++l) {
            double de = epsi[no + l] - epsi[io];
            de -= jdiag[l + (int64_t)nv * io];
            int64_t i = (int64_t)nv * io + l;
            double ek = dot(ncent, pia + (int64_t)ncent * i, qia + (int64_t)ncent * i);
            de += ak * ek;
            if (dokshift) de = kshift(de, ek, smax, width, steep);
            if (de <= thr) {
                iconf[k] = io + 1;
                iconf[k + mx] = no + l + 1;
                ed[k] = de;
                ++k;
            }
            if (de > thr && de < fthr) {
                kconf[j] = io + 1;
                kconf[j + mx] = no + l + 1;
                edpt[j] = de;
                ++j;
            }
        }
    }
    int64_t nex = k;
    /* selection sort, stda.f:651-668 (first strict minimum, swap) */
    for (int64_t ii = 1; ii < nex; ++ii) {
        int64_t i = ii - 1, kk = i;
        double pp = ed[i];
        for (int64_t jj = ii; jj < nex; ++jj)
            if (ed[jj] < pp) { kk = jj; pp = ed[jj]; }
        if (kk == i) continue;
        ed[kk] = ed[i];
        ed[i] = pp;
        for (int m = 0; m < 2; ++m) {
            int h = iconf[i + m * mx];
            iconf[i + m * mx] = iconf[kk + m * mx];
            iconf[kk + m * mx] = h;
        }
    }
    *nex_out = nex;
    *nexpt_out = j;
}

/*
 * Dual-threshold screening of std2 -dual, stda-rw_dual.f:712-853: the occupied
 * MOs are visited in occ_order (the chromophore MOs first, then the rest, each
 * ascending), each with its own primary threshold thr_i and PT2 window fthr_i;
 * then the same selection sort.
 */
void orc_screen_dual(int ncent, int no, int nv, const double *epsi, const double *jdiag,
                     const double *pia, const double *qia, double ak, const int *occ_order,
                     const double *thr_i, const double *fthr_i, int dokshift, double smax,
                     double width, double steep, int64_t mx, int *iconf, double *ed, int *kconf,
                     double *edpt, int64_t *nex_out, int64_t *nexpt_out) {
    int64_t k = 0, j = 0;
    for (int q = 0; q < no; ++q) {
        const int io = occ_order[q];
        for (int l = 0; l < nv; ++l) {
            double de = epsi[no + l] - epsi[io];
            de -= jdiag[l + (int64_t)nv * io];
            int64_t i = (int64_t)nv * io + l;
            double ek = dot(ncent, pia + (int64_t)ncent * i, qia + (int64_t)ncent * i);
            de += ak * ek;
            if (dokshift) de = kshift(de, ek, smax, width, steep);
            if (de <= thr_i[io]) {
                iconf[k] = io + 1;
                iconf[k + mx] = no + l + 1;
                ed[k] = de;
                ++k;
            }
            if (de > thr_i[io] && de < fthr_i[io]) {
                kconf[j] = io + 1;
                kconf[j + mx] = no + l + 1;
                edpt[j] = de;
                ++j;
            }
        }
    }
    int64_t nex = k;
    for (int64_t ii = 1; ii < nex; ++ii) {   /* stda-rw_dual.f:836-853 */
        int64_t i = ii - 1, kk = i;
        double pp = ed[i];
        for (int64_t jj = ii; jj < nex; ++jj)
            if (ed[jj] < pp) { kk = jj; pp = ed[jj]; }
        if (kk == i) continue;
        ed[kk] = ed[i];
        ed[i] = pp;
        for (int m = 0; m < 2; ++m) {
            int h = iconf[i + m * mx];
            iconf[i + m * mx] = iconf[kk + m * mx];
            iconf[kk + m * mx] = h;
        }
    }
    *nex_out = nex;
    *nexpt_out = j;
}

/*
 * PT2 selection of secondary CSFs, stda.f:1773-1863.  Appends the selected
 * secondaries to iconf/ed (in kconf order), lowers ed of the primaries by the
 * PT2 sum of the rejected secondaries.  Returns the number added.
 */
int64_t orc_ptselect(int64_t nex, int ncent, int no, int nv, int64_t nexpt, int64_t mx,
                     int *iconf, const int *kconf, double ak, double *ed,
                     const double *edpt, const double *pia, const double *qia,
                     const double *pij, const double *qab, double thrp,
                     unsigned char *incl) {
    int nth = orc_num_threads();
    double *pt2_all = (double *)calloc((size_t)nth * (size_t)nex, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *pt2 = pt2_all + (size_t)tid * nex;
        double *pt = (double *)malloc(sizeof(double) * (size_t)(nex > 0 ? nex : 1));
#pragma omp for schedule(dynamic, 16)
        for (int64_t k = 0; k < nexpt; ++k) {
            int io = kconf[k], iv = kconf[k + mx];
            int iiv = iv - no;
            int64_t iwrk = (int64_t)(io - 1) * nv + iiv - 1;
            const double *qk = pia + (int64_t)ncent * iwrk;
            double de = edpt[k];
            double pert = 0.0;
            for (int64_t j = 0; j < nex; ++j) {
                int jo = iconf[j], jv = iconf[j + mx];
                int jjv = jv - no;
                int64_t jwrk = (int64_t)(jo - 1) * nv + jjv - 1;
                double ek = dot(ncent, qk, qia + (int64_t)ncent * jwrk);
                double ej = dot(ncent, pij + (int64_t)ncent * (lin(io, jo) - 1),
                                qab + (int64_t)ncent * (lin(iiv, jjv) - 1));
                double amat = ak * ek - ej;
                pt[j] = amat * amat / (de - ed[j] + 1.0e-10);
            }
            for (int64_t j = 0; j < nex; ++j) pert += pt[j];
            if (pert > thrp) {
                incl[k] = 1;
            } else {
                incl[k] = 0;
                for (int64_t j = 0; j < nex; ++j) pt2[j] += pt[j];
            }
        }
        free(pt);
    }
    int64_t nnew = 0;
    for (int64_t i = 0; i < nexpt; ++i) {
        if (!incl[i]) continue;
        iconf[nex + nnew] = kconf[i];
        iconf[nex + nnew + mx] = kconf[i + mx];
        ed[nex + nnew] = edpt[i];
        ++nnew;
    }
    for (int64_t i = 0; i < nex; ++i) {
        double s = 0.0;
        for (int t = 0; t < nth; ++t) s += pt2_all[(size_t)t * nex + i];
        ed[i] -= s;
    }
    free(pt2_all);
    return nnew;
}

/* explicit TDA matrix A (nci,nci), rtdamat stda.f:2010-2056 */
void orc_rtdamat(int64_t nci, int ncent, int no, int nv, int64_t mx, const int *iconf,
                 double ak, const double *ed, const double *pia, const double *qia,
                 const double *pij, const double *qab, double *hci) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < nci; ++i) {
        int io = iconf[i], iv = iconf[i + mx];
        int iiv = iv - no;
        int64_t iwrk = (int64_t)(io - 1) * nv + iiv - 1;
        const double *qk = pia + (int64_t)ncent * iwrk;
        for (int64_t j = 0; j < i; ++j) {
            int jo = iconf[j], jv = iconf[j + mx];
            int jjv = jv - no;
            int64_t jwrk = (int64_t)(jo - 1) * nv + jjv - 1;
            double ek = dot(ncent, qk, qia + (int64_t)ncent * jwrk);
            double ej = dot(ncent, pij + (int64_t)ncent * (lin(io, jo) - 1),
                            qab + (int64_t)ncent * (lin(iiv, jjv) - 1));
            double v = ak * ek - ej;
            hci[j + nci * i] = v;
            hci[i + nci * j] = v;
        }
        hci[i + nci * i] = ed[i];
    }
}

/*
 * A+B and A-B packed (upper, lin(i,j)), rrpamat stda.f:1870-2004.
 * The ax < 1e-6 branch keeps A-B diagonal; there, the stored A-B diagonal
 * is (ed - ak K) (the resp/aresp/optrota case, stda.f:1916-1920); taking its
 * square root for the eigenvalue path is the caller's job.
 */
void orc_rrpamat(int64_t nci, int ncent, int no, int nv, int64_t mx, const int *iconf,
                 double ak, double ax, const double *ed, const double *pia,
                 const double *qia, const double *pij, const double *qab, double *apb,
                 double *amb) {
    int small_ax = fabs(ax) < 1.0e-6;
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < nci; ++i) {
        int io = iconf[i], iv = iconf[i + mx];
        int iiv = iv - no;
        int64_t iwrk = (int64_t)(io - 1) * nv + iiv - 1;
        const double *qk = pia + (int64_t)ncent * iwrk;
        for (int64_t j = 0; j < i; ++j) {
            int64_t ij = lin(i + 1, j + 1) - 1;
            int jo = iconf[j], jv = iconf[j + mx];
            int jjv = jv - no;
            int64_t jwrk = (int64_t)(jo - 1) * nv + jjv - 1;
            double ek = dot(ncent, qk, qia + (int64_t)ncent * jwrk);
            if (small_ax) {
                apb[ij] = 2.0 * ak * ek;
                amb[ij] = 0.0;
                continue;
            }
            double ej = dot(ncent, pij + (int64_t)ncent * (lin(io, jo) - 1),
                            qab + (int64_t)ncent * (lin(iiv, jjv) - 1));
            int64_t w1 = (int64_t)(io - 1) * nv + jjv - 1;
            int64_t w2 = (int64_t)(jo - 1) * nv + iiv - 1;
            double ex = dot(ncent, pia + (int64_t)ncent * w1, qia + (int64_t)ncent * w2);
            apb[ij] = 2.0 * ak * ek - ax * ex - ej;
            amb[ij] = ax * ex - ej;
        }
        int64_t ii = lin(i + 1, i + 1) - 1;
        double ek = dot(ncent, qk, qia + (int64_t)ncent * iwrk);
        if (small_ax) {
            amb[ii] = ed[i] - ak * ek;
            apb[ii] = ed[i] + ak * ek;
        } else {
            amb[ii] = ed[i] - ak * ek + ax * ek;
            apb[ii] = ed[i] - ax * ek + ak * ek;
        }
    }
}

/* 0.5 B packed (lin8), rtdacorr apbtrafo.f:580-637 */
void orc_rtdacorr(int64_t nci, int ncent, int no, int nv, int64_t mx, const int *iconf,
                  double ak, double ax, const double *pia, const double *qia, double *bmat) {
    const double fact = 0.5;
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < nci; ++i) {
        int io = iconf[i], iv = iconf[i + mx];
        int iiv = iv - no;
        int64_t iwrk = (int64_t)(io - 1) * nv + iiv - 1;
        const double *qk = pia + (int64_t)ncent * iwrk;
        for (int64_t j = 0; j < i; ++j) {
            int64_t ij = lin(i + 1, j + 1) - 1;
            int jo = iconf[j], jv = iconf[j + mx];
            int jjv = jv - no;
            int64_t jwrk = (int64_t)(jo - 1) * nv + jjv - 1;
            double ek = dot(ncent, qk, qia + (int64_t)ncent * jwrk);
            double b = fact * ak * ek;
            int64_t w1 = (int64_t)(io - 1) * nv + jjv - 1;
            int64_t w2 = (int64_t)(jo - 1) * nv + iiv - 1;
            ek = dot(ncent, pia + (int64_t)ncent * w1, qia + (int64_t)ncent * w2);
            bmat[ij] = b - fact * ax * ek;
        }
        int64_t ii = lin(i + 1, i + 1) - 1;
        double ek = dot(ncent, qk, qia + (int64_t)ncent * iwrk);
        bmat[ii] = fact * (ak * ek - ax * ek);
    }
}

/* Rows i0..i1-1 of the TDA matrix (couplings with every j < i), the loop of
 * rtdamat for a slice of rows; used to time the A build on a sample. */
void orc_rtdamat_rows(int64_t i0, int64_t i1, int64_t nci, int ncent, int no, int nv,
                      int64_t mx, const int *iconf, double ak, const double *pia,
                      const double *qia, const double *pij, const double *qab, double *rows) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = i0; i < i1; ++i) {
        int io = iconf[i], iv = iconf[i + mx];
        int iiv = iv - no;
        int64_t iwrk = (int64_t)(io - 1) * nv + iiv - 1;
        const double *qk = pia + (int64_t)ncent * iwrk;
        double *out = rows + (i - i0) * nci;
        for (int64_t j = 0; j < i; ++j) {
            int jo = iconf[j], jv = iconf[j + mx];
            int jjv = jv - no;
            int64_t jwrk = (int64_t)(jo - 1) * nv + jjv - 1;
            double ek = dot(ncent, qk, qia + (int64_t)ncent * jwrk);
            double ej = dot(ncent, pij + (int64_t)ncent * (lin(io, jo) - 1),
                            qab + (int64_t)ncent * (lin(iiv, jjv) - 1));
            out[j] = ak * ek - ej;
        }
    }
}
