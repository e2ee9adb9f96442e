/*
 * qk_oracle.c -- CPU restatement of QUICK's ERI Fock-build path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product path (quick_amd/csrc) never links or calls it.
 *
 * What it restates (reference file:line):
 *   qko_create        shell tables as QUICK's readbasis leaves them (basis.f90:377-599)
 *   g2eshell          primitive-pair data Apri/Ppri/Kpri/Xcoeff (shell.f90:3-88)
 *   qko_fmt           Boys F_m(T), QUICK's piecewise fit + recursion (subs/fmt.f90:12-84)
 *   vrr               Obara-Saika / Head-Gordon-Pople VRR [e0|f0]^(m) with the W
 *                     intermediates QUICK's `vertical` uses (subs/hrr.f90:1018-3423;
 *                     the recurrence, not its unrolled blocks)
 *   hrr_term          HRR as QUICK's lefthrr expands it: (a,b) onto (a+k,0) with
 *                     binomial (A-B) powers, times the cons factors (subs/hrr.f90:6-285)
 *   qko_schwarz       Ycutoff / cutprim (schwarz.f90:12-548)
 *   qko_fock mode 0   CPU enumeration hfoperator->get2e->shell->iclass
 *                     (hfoperator.f90:709-741, shell.f90:463-1012) incl. copySym
 *   qko_fock mode 1   GPU enumeration: Qshell pairs with shell(i)<=shell(j) and
 *                     Ycutoff>pair_cut (gpu.cu:444-647), pair-of-pairs with
 *                     shell(ii)<=shell(kk), the unique-component predicate
 *                     (gpu_get2e_subs.h:166-336,792-801); digestion by the
 *                     mathematically equivalent symmetric J/K update.
 *   qko_grad          four-centre gradient (gradient.f90:261-320, shell.f90:1863-2535)
 *
 * Parity pins: tests/test_oracle.py checks every function above against the
 * reference's own outputs (tests/golden/*.npz), closed forms and brute force.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXPRIM 10
#define LMAX 7                       /* largest |e| or |f| the VRR tables hold */
#define NCART 120                    /* Cartesians with |e| <= 7 */
#define MMAX 16

static const double PI = 3.1415926535897932384626433832795;

/* ---------------- Cartesian index tables ---------------- */
static int cidx[LMAX + 1][LMAX + 1][LMAX + 1];
static int cl[NCART][3];
static int ncart_upto[LMAX + 2];     /* number of Cartesians with |e| < L */
static int tables_ready = 0;

static void init_tables(void) {
    if (tables_ready) return;
    int n = 0;
    for (int L = 0; L <= LMAX; ++L) {
        ncart_upto[L] = n;
        for (int i = L; i >= 0; --i)
            for (int j = L - i; j >= 0; --j) {
                int k = L - i - j;
                cidx[i][j][k] = n;
                cl[n][0] = i; cl[n][1] = j; cl[n][2] = k;
                ++n;
            }
    }
    ncart_upto[LMAX + 1] = n;
    tables_ready = 1;
}

/* ---------------- context ---------------- */
typedef struct {
    int natom, nshell, nbasis, jbasis;
    double *xyz;                          /* 3 x natom */
    int *kstart, *kprim, *katom, *ktype, *Qstart, *Qfinal, *Qsbasis, *Qfbasis, *ksumtype, *KLMN;
    double *cons, *gccoeff, *gcexpo;      /* gccoeff(MAXPRIM, nbasis) */
    /* g2eshell */
    double *Apri, *Ppri, *Kpri, *Xcoeff;  /* jbasis^2, 3*jbasis^2, jbasis^2, jbasis^2*16 */
    double *Ycutoff, *cutprim;            /* nshell^2, jbasis^2 */
    int boys_exact;
} qko_t;

#define IX2(a, b, n) ((a) + (size_t)(n) * (b))

static double *dup_d(const double *s, size_t n) { double *d = malloc(n * sizeof(double)); memcpy(d, s, n * sizeof(double)); return d; }
static int *dup_i(const int *s, size_t n) { int *d = malloc(n * sizeof(int)); memcpy(d, s, n * sizeof(int)); return d; }

static void g2eshell(qko_t *c);

qko_t *qko_create(int natom, int nshell, int nbasis, int jbasis, const double *xyz,
                  const int *kstart, const int *kprim, const int *katom, const int *ktype,
                  const int *Qstart, const int *Qfinal, const int *Qsbasis, const int *Qfbasis,
                  const int *ksumtype, const int *KLMN, const double *cons,
                  const double *gccoeff, const double *gcexpo) {
    init_tables();
    qko_t *c = calloc(1, sizeof(qko_t));
    c->natom = natom; c->nshell = nshell; c->nbasis = nbasis; c->jbasis = jbasis;
    c->xyz = dup_d(xyz, 3 * (size_t)natom);
    c->kstart = dup_i(kstart, nshell); c->kprim = dup_i(kprim, nshell);
    c->katom = dup_i(katom, nshell); c->ktype = dup_i(ktype, nshell);
    c->Qstart = dup_i(Qstart, nshell); c->Qfinal = dup_i(Qfinal, nshell);
    c->Qsbasis = dup_i(Qsbasis, 4 * (size_t)nshell); c->Qfbasis = dup_i(Qfbasis, 4 * (size_t)nshell);
    c->ksumtype = dup_i(ksumtype, nshell + 1);
    c->KLMN = dup_i(KLMN, 3 * (size_t)nbasis);
    c->cons = dup_d(cons, nbasis);
    c->gccoeff = dup_d(gccoeff, MAXPRIM * (size_t)nbasis);
    c->gcexpo = dup_d(gcexpo, MAXPRIM * (size_t)nbasis);
    size_t jb2 = (size_t)jbasis * jbasis;
    c->Apri = calloc(jb2, sizeof(double));
    c->Ppri = calloc(3 * jb2, sizeof(double));
    c->Kpri = calloc(jb2, sizeof(double));
    c->Xcoeff = calloc(16 * jb2, sizeof(double));
    c->Ycutoff = calloc((size_t)nshell * nshell, sizeof(double));
    c->cutprim = calloc(jb2, sizeof(double));
    g2eshell(c);
    return c;
}

void qko_destroy(qko_t *c) {
    if (!c) return;
    free(c->xyz); free(c->kstart); free(c->kprim); free(c->katom); free(c->ktype);
    free(c->Qstart); free(c->Qfinal); free(c->Qsbasis); free(c->Qfbasis); free(c->ksumtype);
    free(c->KLMN); free(c->cons); free(c->gccoeff); free(c->gcexpo);
    free(c->Apri); free(c->Ppri); free(c->Kpri); free(c->Xcoeff); free(c->Ycutoff); free(c->cutprim);
    free(c);
}

void qko_set_boys_exact(qko_t *c, int on) { c->boys_exact = on; }

/* accessors mirroring the Fortran arrays (1-based arguments) */
#define KSTART(s) (c->kstart[(s)])
#define XC(a, b, i, j) c->Xcoeff[(a) + (size_t)c->jbasis * ((b) + (size_t)c->jbasis * ((i) + 4 * (size_t)(j)))]
#define GCC(k, f) c->gccoeff[(k) + MAXPRIM * (size_t)(f)]
#define GCE(k, f) c->gcexpo[(k) + MAXPRIM * (size_t)(f)]
#define ATOMXYZ(m, a) c->xyz[(m) + 3 * (size_t)(a)]

/* shell.f90:3-88 : Apri = a+b, Ppri = (aA+bB)/(a+b), Kpri = exp(-ab/(a+b)|AB|^2)/(a+b),
 * Xcoeff(A,B,I,J) = Kpri * gccoeff(ips, ksumtype+I) * gccoeff(jps, ksumtype+J). */
static void g2eshell(qko_t *c) {
    int jb = c->jbasis;
    for (int ics = 0; ics < c->nshell; ++ics)
        for (int ips = 0; ips < c->kprim[ics]; ++ips)
            for (int jcs = 0; jcs < c->nshell; ++jcs)
                for (int jps = 0; jps < c->kprim[jcs]; ++jps) {
                    int NA = c->kstart[ics] - 1 + ips, NB = c->kstart[jcs] - 1 + jps;
                    double AA = GCE(ips, c->ksumtype[ics] - 1), BB = GCE(jps, c->ksumtype[jcs] - 1);
                    int ia = c->katom[ics] - 1, ib = c->katom[jcs] - 1;
                    c->Apri[IX2(NA, NB, jb)] = AA + BB;
                    double dx = ATOMXYZ(0, ia) - ATOMXYZ(0, ib), dy = ATOMXYZ(1, ia) - ATOMXYZ(1, ib),
                           dz = ATOMXYZ(2, ia) - ATOMXYZ(2, ib);
                    double DAB = sqrt(dx * dx + dy * dy + dz * dz);
                    for (int m = 0; m < 3; ++m)
                        c->Ppri[m + 3 * IX2(NA, NB, jb)] = (ATOMXYZ(m, ia) * AA + ATOMXYZ(m, ib) * BB) / (AA + BB);
                    double K = exp(-AA * BB / (AA + BB) * (DAB * DAB)) / (AA + BB);
                    c->Kpri[IX2(NA, NB, jb)] = K;
                    for (int it = c->Qstart[ics]; it <= c->Qfinal[ics]; ++it)
                        for (int it2 = c->Qstart[jcs]; it2 <= c->Qfinal[jcs]; ++it2)
                            XC(NA, NB, it, it2) = K * GCC(ips, c->ksumtype[ics] - 1 + it) *
                                                  GCC(jps, c->ksumtype[jcs] - 1 + it2);
                }
}

/* ---------------- Boys function ---------------- */
/* subs/fmt.f90:12-84 restated: piecewise F_0 (or F_M near T=0), then upward
 * (T>3e-5, or T>3e-7 when M<4) or downward recursion. */
void qko_fmt(int MaxM, double X, double *vals) {
    const double PIE4 = PI / 4.0;
    double XINV = 1.0 / X, E = exp(-X), WW1, F1, Y;
    if (X > 5.0) {
        if (X > 15.0) {
            if (X > 33.0) WW1 = sqrt(PIE4 * XINV);
            else WW1 = ((1.9623264149430e-01 * XINV - 4.9695241464490e-01) * XINV - 6.0156581186481e-05) * E + sqrt(PIE4 * XINV);
        } else if (X > 10.0) {
            WW1 = (((-1.8784686463512e-01 * XINV + 2.2991849164985e-01) * XINV - 4.9893752514047e-01) * XINV - 2.1916512131607e-05) * E + sqrt(PIE4 * XINV);
        } else {
            WW1 = ((((((4.6897511375022e-01 * XINV - 6.9955602298985e-01) * XINV + 5.3689283271887e-01) * XINV - 3.2883030418398e-01) * XINV +
                     2.4645596956002e-01) * XINV - 4.9984072848436e-01) * XINV - 3.1501078774085e-06) * E + sqrt(PIE4 * XINV);
        }
    } else if (X > 1.0) {
        if (X > 3.0) {
            Y = X - 4.0;
            F1 = ((((((((((-2.62453564772299e-11 * Y + 3.24031041623823e-10) * Y - 3.614965656163e-09) * Y + 3.760256799971e-08) * Y -
                        3.553558319675e-07) * Y + 3.022556449731e-06) * Y - 2.290098979647e-05) * Y + 1.526537461148e-04) * Y -
                    8.81947375894379e-04) * Y + 4.33207949514611e-03) * Y - 1.75257821619926e-02) * Y + 5.28406320615584e-02;
            WW1 = (X + X) * F1 + E;
        } else {
            Y = X - 2.0;
            F1 = ((((((((((-1.61702782425558e-10 * Y + 1.96215250865776e-09) * Y - 2.14234468198419e-08) * Y + 2.17216556336318e-07) * Y -
                        1.98850171329371e-06) * Y + 1.62429321438911e-05) * Y - 1.16740298039895e-04) * Y + 7.24888732052332e-04) * Y -
                    3.79490003707156e-03) * Y + 1.61723488664661e-02) * Y - 5.29428148329736e-02) * Y + 1.15702180856167e-01;
            WW1 = (X + X) * F1 + E;
        }
    } else if (X > 3.0e-05 || (X > 3.0e-07 && MaxM < 4)) {
        F1 = ((((((((-8.36313918003957e-08 * X + 1.21222603512827e-06) * X - 1.15662609053481e-05) * X + 9.25197374512647e-05) * X -
                  6.40994113129432e-04) * X + 3.78787044215009e-03) * X - 1.85185172458485e-02) * X + 7.14285713298222e-02) * X -
              1.99999999997023e-01) * X + 3.33333333333318e-01;
        WW1 = (X + X) * F1 + E;
    } else {
        WW1 = (1.0 - X) / (double)(2 * MaxM + 1);
    }
    if (X > 3.0e-5 || (X > 3.0e-07 && MaxM < 4)) {
        vals[0] = WW1;
        for (int m = 1; m <= MaxM; ++m) vals[m] = (((2 * m - 1) * vals[m - 1]) - E) * 0.5 * XINV;
    } else {
        vals[MaxM] = WW1;
        double twot = X * 2.0;
        for (int m = MaxM - 1; m >= 0; --m) vals[m] = (twot * vals[m + 1] + E) / (double)(m * 2 + 1);
    }
}

/* Accurate Boys values (not the reference's algorithm): series for F_M at
 * small/moderate T then stable downward recursion; asymptotic + upward for
 * large T.  Used only to measure how far the reference's fit is from the
 * mathematical value. */
void qko_fmt_exact(int M, double T, double *vals) {
    double E = exp(-T);
    if (T < 35.0) {
        /* F_M(T) = e^{-T} sum_k (2T)^k / ((2M+1)(2M+3)...(2M+2k+1)) */
        double term = 1.0 / (2 * M + 1), sum = term;
        for (int k = 1; k < 400; ++k) {
            term *= 2.0 * T / (2 * M + 2 * k + 1);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        vals[M] = E * sum;
        for (int m = M - 1; m >= 0; --m) vals[m] = (2.0 * T * vals[m + 1] + E) / (2 * m + 1);
    } else {
        vals[0] = 0.5 * sqrt(PI / T) * erf(sqrt(T));
        for (int m = 1; m <= M; ++m) vals[m] = ((2 * m - 1) * vals[m - 1] - E) / (2.0 * T);
    }
}

/* ---------------- VRR ---------------- */
/* Computes V[e][f] = [e0|f0]^(0) for |e| <= LE, |f| <= LF given F_m (already
 * carrying the primitive prefactor).  Recurrences (Obara & Saika 1986; Head-Gordon
 * & Pople 1988) in the variables QUICK's `vertical` receives:
 *   PA = P-A, WP = W-P, QC = Q-C, WQ = W-Q, 1/2zeta, 1/2eta, 1/2(zeta+eta),
 *   rho/zeta = eta/(zeta+eta), rho/eta = zeta/(zeta+eta). */
typedef struct {
    double PA[3], WP[3], QC[3], WQ[3];
    double oo2z, oo2e, oo2ze, roz, roe;
} vrr_par;

static double Vbuf[MMAX + 1][NCART][NCART];

static void vrr(int LE, int LF, const double *Fm, const vrr_par *p, double out[NCART][NCART]) {
    int MT = LE + LF;
    int ne = ncart_upto[LE + 1], nf = ncart_upto[LF + 1];
    for (int m = 0; m <= MT; ++m) Vbuf[m][0][0] = Fm[m];
    /* bra build-up with f = 0 */
    for (int e = 1; e < ne; ++e) {
        int ex = cl[e][0], ey = cl[e][1], ez = cl[e][2];
        int i = ex > 0 ? 0 : (ey > 0 ? 1 : 2);
        int em[3] = {ex, ey, ez};
        em[i] -= 1;
        int e1 = cidx[em[0]][em[1]][em[2]];
        int n1 = em[i];
        int e2 = -1;
        if (n1 > 0) { em[i] -= 1; e2 = cidx[em[0]][em[1]][em[2]]; }
        int Le = ex + ey + ez;
        for (int m = 0; m <= MT - Le; ++m) {
            double v = p->PA[i] * Vbuf[m][e1][0] + p->WP[i] * Vbuf[m + 1][e1][0];
            if (n1 > 0) v += n1 * p->oo2z * (Vbuf[m][e2][0] - p->roz * Vbuf[m + 1][e2][0]);
            Vbuf[m][e][0] = v;
        }
    }
    /* ket build-up */
    for (int f = 1; f < nf; ++f) {
        int fx = cl[f][0], fy = cl[f][1], fz = cl[f][2];
        int i = fx > 0 ? 0 : (fy > 0 ? 1 : 2);
        int fm[3] = {fx, fy, fz};
        fm[i] -= 1;
        int f1 = cidx[fm[0]][fm[1]][fm[2]];
        int n1 = fm[i];
        int f2 = -1;
        if (n1 > 0) { fm[i] -= 1; f2 = cidx[fm[0]][fm[1]][fm[2]]; }
        int Lf = fx + fy + fz;
        for (int e = 0; e < ne; ++e) {
            int Le = cl[e][0] + cl[e][1] + cl[e][2];
            int ne_i = cl[e][i];
            int e1 = -1;
            if (ne_i > 0) {
                int emm[3] = {cl[e][0], cl[e][1], cl[e][2]};
                emm[i] -= 1;
                e1 = cidx[emm[0]][emm[1]][emm[2]];
            }
            for (int m = 0; m <= MT - Le - Lf; ++m) {
                double v = p->QC[i] * Vbuf[m][e][f1] + p->WQ[i] * Vbuf[m + 1][e][f1];
                if (n1 > 0) v += n1 * p->oo2e * (Vbuf[m][e][f2] - p->roe * Vbuf[m + 1][e][f2]);
                if (ne_i > 0) v += ne_i * p->oo2ze * Vbuf[m + 1][e1][f1];
                Vbuf[m][e][f] = v;
            }
        }
    }
    for (int e = 0; e < ne; ++e)
        for (int f = 0; f < nf; ++f) out[e][f] = Vbuf[0][e][f];
}

/* ---------------- HRR ---------------- */
/* lefthrr: (a b| = sum_k prod_i C(b_i,k_i) (A-B)_i^(b_i-k_i) (a+k 0| */
static int hrr_expand(const int *a, const int *b, const double *AB, double *coef, int *idx) {
    int n = 0;
    static const int binom[4][4] = {{1, 0, 0, 0}, {1, 1, 0, 0}, {1, 2, 1, 0}, {1, 3, 3, 1}};
    for (int kx = 0; kx <= b[0]; ++kx)
        for (int ky = 0; ky <= b[1]; ++ky)
            for (int kz = 0; kz <= b[2]; ++kz) {
                double cf = binom[b[0]][kx] * binom[b[1]][ky] * binom[b[2]][kz];
                for (int t = 0; t < b[0] - kx; ++t) cf *= AB[0];
                for (int t = 0; t < b[1] - ky; ++t) cf *= AB[1];
                for (int t = 0; t < b[2] - kz; ++t) cf *= AB[2];
                coef[n] = cf;
                idx[n] = cidx[a[0] + kx][a[1] + ky][a[2] + kz];
                ++n;
            }
    return n;
}

/* ---------------- contracted shell-quartet store ---------------- */
static double store[NCART][NCART];
static double vtmp[NCART][NCART];

/* Contracted [e0|f0] for Qshell types (I,J,K,L) of shells (II,JJ,KK,LL) (0-based
 * shells), summing primitive quartets that pass the cutprim test when
 * use_primcut is set (shell.f90:514-644 + iclass:718-753).  Mboys is the order
 * passed to FmT (>= LE+LF).  Returns the count of primitive quartets used. */
static long contracted_store(qko_t *c, int II, int JJ, int KK, int LL, int I, int J, int K, int L,
                             double dnmax, double primLimit, int use_primcut, int raise, int Mboys) {
    int jb = c->jbasis;
    int LE = I + J + raise, LF = K + L + raise;
    int ne = ncart_upto[LE + 1], nf = ncart_upto[LF + 1];
    for (int e = 0; e < ne; ++e) for (int f = 0; f < nf; ++f) store[e][f] = 0.0;
    const double X0 = 2.0 * pow(PI, 2.5);
    double RA[3], RC[3];
    for (int m = 0; m < 3; ++m) { RA[m] = ATOMXYZ(m, c->katom[II] - 1); RC[m] = ATOMXYZ(m, c->katom[KK] - 1); }
    long nq = 0;
    double Fm[MMAX + 1];
    for (int jj = 0; jj < c->kprim[JJ]; ++jj)
        for (int ii = 0; ii < c->kprim[II]; ++ii) {
            int Pi = c->kstart[II] - 1 + ii, Pj = c->kstart[JJ] - 1 + jj;
            double AB = c->Apri[IX2(Pi, Pj, jb)];
            const double *P = &c->Ppri[3 * IX2(Pi, Pj, jb)];
            double X2 = X0 * XC(Pi, Pj, I, J);
            double cp1 = dnmax * c->cutprim[IX2(Pi, Pj, jb)];
            for (int ll = 0; ll < c->kprim[LL]; ++ll)
                for (int kk = 0; kk < c->kprim[KK]; ++kk) {
                    int Pk = c->kstart[KK] - 1 + kk, Pl = c->kstart[LL] - 1 + ll;
                    if (use_primcut && !(cp1 * c->cutprim[IX2(Pk, Pl, jb)] > primLimit)) continue;
                    double CD = c->Apri[IX2(Pk, Pl, jb)];
                    const double *Q = &c->Ppri[3 * IX2(Pk, Pl, jb)];
                    double ABCD = AB + CD, ROU = AB * CD / ABCD, RPQ = 0.0;
                    vrr_par vp;
                    for (int m = 0; m < 3; ++m) {
                        double W = (P[m] * AB + Q[m] * CD) / ABCD;
                        double d = P[m] - Q[m];
                        RPQ += d * d;
                        vp.PA[m] = P[m] - RA[m];
                        vp.QC[m] = Q[m] - RC[m];
                        vp.WP[m] = W - P[m];
                        vp.WQ[m] = W - Q[m];
                    }
                    vp.oo2z = 0.5 / AB; vp.oo2e = 0.5 / CD; vp.oo2ze = 0.5 / ABCD;
                    vp.roz = CD / ABCD; vp.roe = AB / ABCD;
                    double T = RPQ * ROU;
                    /* the order the reference asks FmT for: the shell quartet's
                     * Qfinal sum on the CPU path, the Qshell types' sum on the
                     * GPU path; it selects FmT's branch near T = 0 */
                    int M = Mboys;
                    if (c->boys_exact) qko_fmt_exact(M, T, Fm); else qko_fmt(M, T, Fm);
                    double X44 = X2 * XC(Pk, Pl, K, L);
                    double s = X44 / sqrt(ABCD);
                    for (int m = 0; m <= M; ++m) Fm[m] *= s;
                    vrr(LE, LF, Fm, &vp, vtmp);
                    for (int e = 0; e < ne; ++e) for (int f = 0; f < nf; ++f) store[e][f] += vtmp[e][f];
                    ++nq;
                }
        }
    return nq;
}

/* (iii jjj|kkk lll) from the contracted store (functions 0-based). */
static double hrr_value(qko_t *c, int II, int JJ, int KK, int LL, int iii, int jjj, int kkk, int lll) {
    double RA[3], RB[3], RC[3], RD[3], AB[3], CD[3];
    for (int m = 0; m < 3; ++m) {
        RA[m] = ATOMXYZ(m, c->katom[II] - 1); RB[m] = ATOMXYZ(m, c->katom[JJ] - 1);
        RC[m] = ATOMXYZ(m, c->katom[KK] - 1); RD[m] = ATOMXYZ(m, c->katom[LL] - 1);
        AB[m] = RA[m] - RB[m]; CD[m] = RC[m] - RD[m];
    }
    double cl_[16], cr[16];
    int il[16], ir[16];
    int nl = hrr_expand(&c->KLMN[3 * iii], &c->KLMN[3 * jjj], AB, cl_, il);
    int nr = hrr_expand(&c->KLMN[3 * kkk], &c->KLMN[3 * lll], CD, cr, ir);
    double Y = 0.0;
    for (int a = 0; a < nl; ++a)
        for (int b = 0; b < nr; ++b) Y += cl_[a] * cr[b] * store[il[a]][ir[b]];
    return Y * c->cons[iii] * c->cons[jjj] * c->cons[kkk] * c->cons[lll];
}

/* Shell (or Qshell) function range: shell s (0-based), type I -> [f0, f1] 0-based */
static void qrange(qko_t *c, int s, int I, int *f0, int *f1) {
    *f0 = c->ksumtype[s] - 1 + c->Qsbasis[s + c->nshell * I];
    *f1 = c->ksumtype[s] - 1 + c->Qfbasis[s + c->nshell * I];
}

/* ---------------- Schwarz (schwarz.f90:12-548) ---------------- */
void qko_schwarz(qko_t *c, double *Yout, double *cpout) {
    int ns = c->nshell, jb = c->jbasis;
    for (int II = 0; II < ns; ++II)
        for (int JJ = II; JJ < ns; ++JJ) {
            double Ymax = 0.0;
            for (int I = c->Qstart[II]; I <= c->Qfinal[II]; ++I)
                for (int J = c->Qstart[JJ]; J <= c->Qfinal[JJ]; ++J) {
                    contracted_store(c, II, JJ, II, JJ, I, J, I, J, 0.0, 0.0, 0, 0,
                                     2 * (c->Qfinal[II] + c->Qfinal[JJ]));
                    int a0, a1, b0, b1;
                    qrange(c, II, I, &a0, &a1); qrange(c, JJ, J, &b0, &b1);
                    for (int iii = a0; iii <= a1; ++iii)
                        for (int jjj = b0; jjj <= b1; ++jjj) {
                            double y = fabs(hrr_value(c, II, JJ, II, JJ, iii, jjj, iii, jjj));
                            if (y > Ymax) Ymax = y;
                        }
                }
            c->Ycutoff[IX2(II, JJ, ns)] = c->Ycutoff[IX2(JJ, II, ns)] = sqrt(Ymax);
            /* classprim: per primitive pair, [ab|ab] with the same primitive pair
             * on both sides, weighted by (X0*Xcoeff)^2 */
            for (int ii = 0; ii < c->kprim[II]; ++ii)
                for (int jj = 0; jj < c->kprim[JJ]; ++jj) {
                    int Pi = c->kstart[II] - 1 + ii, Pj = c->kstart[JJ] - 1 + jj;
                    double Ymp = 0.0;
                    for (int I = c->Qstart[II]; I <= c->Qfinal[II]; ++I)
                        for (int J = c->Qstart[JJ]; J <= c->Qfinal[JJ]; ++J) {
                            /* primitive quartet (Pi Pj|Pi Pj) with weight X4444^2 */
                            int LE = I + J, ne = ncart_upto[LE + 1];
                            double AB = c->Apri[IX2(Pi, Pj, jb)];
                            const double *P = &c->Ppri[3 * IX2(Pi, Pj, jb)];
                            double RA[3];
                            for (int m = 0; m < 3; ++m) RA[m] = ATOMXYZ(m, c->katom[II] - 1);
                            vrr_par vp;
                            for (int m = 0; m < 3; ++m) {
                                vp.PA[m] = P[m] - RA[m]; vp.QC[m] = P[m] - RA[m];
                                vp.WP[m] = 0.0; vp.WQ[m] = 0.0;
                            }
                            double ABCD = 2 * AB;
                            vp.oo2z = 0.5 / AB; vp.oo2e = 0.5 / AB; vp.oo2ze = 0.5 / ABCD;
                            vp.roz = AB / ABCD; vp.roe = AB / ABCD;
                            double Fm[MMAX + 1];
                            /* W-P and W-Q are computed, not assumed zero, by the reference */
                            for (int m = 0; m < 3; ++m) {
                                double W = (P[m] * AB + P[m] * AB) / ABCD;
                                vp.WP[m] = W - P[m]; vp.WQ[m] = W - P[m];
                            }
                            if (c->boys_exact) qko_fmt_exact(2 * LE, 0.0, Fm); else qko_fmt(2 * LE, 0.0, Fm);
                            const double X0 = 2.0 * pow(PI, 2.5);
                            double x4 = X0 * XC(Pi, Pj, I, J);
                            double s = 1.0 / sqrt(ABCD);
                            for (int m = 0; m <= 2 * LE; ++m) Fm[m] *= s;
                            vrr(LE, LE, Fm, &vp, vtmp);
                            for (int e = 0; e < ne; ++e) for (int f = 0; f < ne; ++f) store[e][f] = vtmp[e][f] * pow(x4, 2.0);
                            int a0, a1, b0, b1;
                            qrange(c, II, I, &a0, &a1); qrange(c, JJ, J, &b0, &b1);
                            for (int iii = a0; iii <= a1; ++iii)
                                for (int jjj = b0; jjj <= b1; ++jjj) {
                                    double y = fabs(hrr_value(c, II, JJ, II, JJ, iii, jjj, iii, jjj));
                                    if (y > Ymp) Ymp = y;
                                }
                        }
                    c->cutprim[IX2(Pi, Pj, jb)] = sqrt(Ymp);
                }
        }
    if (Yout) memcpy(Yout, c->Ycutoff, sizeof(double) * ns * ns);
    if (cpout) memcpy(cpout, c->cutprim, sizeof(double) * (size_t)jb * jb);
}

void qko_set_cutoffs(qko_t *c, const double *Ycut, const double *cutprim) {
    memcpy(c->Ycutoff, Ycut, sizeof(double) * c->nshell * c->nshell);
    memcpy(c->cutprim, cutprim, sizeof(double) * (size_t)c->jbasis * c->jbasis);
}

/* densityCutoff / DNscreen (schwarz.f90:550-599): max |D| over a shell block */
static void density_cutoff(qko_t *c, const double *D, double *cutmatrix) {
    int ns = c->nshell, n = c->nbasis;
    for (int II = 0; II < ns; ++II)
        for (int JJ = II; JJ < ns; ++JJ) {
            int a0 = c->ksumtype[II] - 1 + c->Qsbasis[II + ns * c->Qstart[II]];
            int a1 = c->ksumtype[II] - 1 + c->Qfbasis[II + ns * c->Qfinal[II]];
            int b0 = c->ksumtype[JJ] - 1 + c->Qsbasis[JJ + ns * c->Qstart[JJ]];
            int b1 = c->ksumtype[JJ] - 1 + c->Qfbasis[JJ + ns * c->Qfinal[JJ]];
            double m = 0.0;
            for (int i = a0; i <= a1; ++i)
                for (int j = b0; j <= b1; ++j) { double v = fabs(D[IX2(j, i, n)]); if (v > m) m = v; }
            cutmatrix[IX2(II, JJ, ns)] = cutmatrix[IX2(JJ, II, ns)] = m;
        }
}

void qko_density_cutoff(qko_t *c, const double *D, double *cutmatrix) { density_cutoff(c, D, cutmatrix); }

static double dnmax_of(const double *cm, int ns, int II, int JJ, int KK, int LL) {
    double a = 4.0 * cm[IX2(II, JJ, ns)], b = 4.0 * cm[IX2(KK, LL, ns)];
    double m = a > b ? a : b;
    double v[4] = {cm[IX2(II, LL, ns)], cm[IX2(II, KK, ns)], cm[IX2(JJ, KK, ns)], cm[IX2(JJ, LL, ns)]};
    for (int i = 0; i < 4; ++i) if (v[i] > m) m = v[i];
    return m;
}

/* ---------------- Fock: CPU enumeration and digestion ---------------- */
#define O(i, j) o[IX2((i), (j), n)]
#define DEN(i, j) D[IX2((i), (j), n)]

/* iclass digestion (shell.f90:846-1008, direct branch) for one (I,J,K,L) */
static void iclass_cpu(qko_t *c, int II, int JJ, int KK, int LL, int I, int J, int K, int L,
                       const double *D, double *o) {
    int n = c->nbasis;
    int III1, III2, JJJ1, JJJ2, KKK1, KKK2, LLL1, LLL2;
    qrange(c, II, I, &III1, &III2); qrange(c, JJ, J, &JJJ1, &JJJ2);
    qrange(c, KK, K, &KKK1, &KKK2); qrange(c, LL, L, &LLL1, &LLL2);
    if (II < JJ && II < KK && KK < LL) {
        for (int III = III1; III <= III2; ++III)
            for (int JJJ = JJJ1; JJJ <= JJJ2; ++JJJ)
                for (int KKK = KKK1; KKK <= KKK2; ++KKK)
                    for (int LLL = LLL1; LLL <= LLL2; ++LLL) {
                        double Y = hrr_value(c, II, JJ, KK, LL, III, JJJ, KKK, LLL);
                        double DKI = DEN(KKK, III), DKJ = DEN(KKK, JJJ), DLJ = DEN(LLL, JJJ),
                               DLI = DEN(LLL, III), DLK = DEN(LLL, KKK), DJI = DEN(JJJ, III);
                        O(JJJ, III) += 2.0 * DLK * Y;
                        O(LLL, KKK) += 2.0 * DJI * Y;
                        O(KKK, III) -= 0.5 * DLJ * Y;
                        O(LLL, III) -= 0.5 * DKJ * Y;
                        O(JJJ, KKK) -= 0.5 * DLI * Y;
                        O(JJJ, LLL) -= 0.5 * DKI * Y;
                        O(KKK, JJJ) -= 0.5 * DLI * Y;
                        O(LLL, JJJ) -= 0.5 * DKI * Y;
                    }
        return;
    }
    for (int III = III1; III <= III2; ++III)
        for (int JJJ = (III > JJJ1 ? III : JJJ1); JJJ <= JJJ2; ++JJJ)
            for (int KKK = (III > KKK1 ? III : KKK1); KKK <= KKK2; ++KKK)
                for (int LLL = (KKK > LLL1 ? KKK : LLL1); LLL <= LLL2; ++LLL) {
                    if (III < KKK) {
                        double Y = hrr_value(c, II, JJ, KK, LL, III, JJJ, KKK, LLL);
                        if (III < JJJ && KKK < LLL) {
                            double DKI = DEN(KKK, III), DKJ = DEN(KKK, JJJ), DLJ = DEN(LLL, JJJ),
                                   DLI = DEN(LLL, III), DLK = DEN(LLL, KKK), DJI = DEN(JJJ, III);
                            O(JJJ, III) += 2.0 * DLK * Y;
                            O(LLL, KKK) += 2.0 * DJI * Y;
                            O(KKK, III) -= 0.5 * DLJ * Y;
                            O(LLL, III) -= 0.5 * DKJ * Y;
                            O(JJJ, KKK) -= 0.5 * DLI * Y;
                            O(JJJ, LLL) -= 0.5 * DKI * Y;
                            O(KKK, JJJ) -= 0.5 * DLI * Y;
                            O(LLL, JJJ) -= 0.5 * DKI * Y;
                        } else if (III == JJJ && KKK == LLL) {
                            double DJI = DEN(KKK, III), DJJ = DEN(KKK, KKK), DII = DEN(III, III);
                            O(III, III) += DJJ * Y;
                            O(KKK, KKK) += DII * Y;
                            O(KKK, III) -= 0.5 * DJI * Y;
                        } else if (JJJ == KKK && JJJ == LLL) {
                            double DJI = DEN(JJJ, III), DJJ = DEN(JJJ, JJJ);
                            O(JJJ, III) += 0.5 * DJJ * Y;
                            O(JJJ, JJJ) += DJI * Y;
                        } else if (KKK == LLL && III < JJJ && JJJ != KKK) {
                            double DKI = DEN(KKK, III), DKJ = DEN(KKK, JJJ), DKK = DEN(KKK, KKK), DJI = DEN(JJJ, III);
                            O(JJJ, III) += DKK * Y;
                            O(KKK, KKK) += 2.0 * DJI * Y;
                            O(KKK, III) -= 0.5 * DKJ * Y;
                            O(KKK, JJJ) -= 0.5 * DKI * Y;
                            O(JJJ, KKK) -= 0.5 * DKI * Y;
                        } else if (III == JJJ && KKK < LLL) {
                            double DII = DEN(III, III), DJI = DEN(KKK, III), DKI = DEN(LLL, III), DKJ = DEN(LLL, KKK);
                            O(LLL, KKK) += DII * Y;
                            O(III, III) += 2.0 * DKJ * Y;
                            O(KKK, III) -= 0.5 * DKI * Y;
                            O(LLL, III) -= 0.5 * DJI * Y;
                        }
                    } else if (JJJ <= LLL) {
                        double Y = hrr_value(c, II, JJ, KK, LL, III, JJJ, KKK, LLL);
                        if (III == JJJ && III == KKK && III == LLL) {
                            double DII = DEN(III, III);
                            O(III, III) += 0.5 * DII * Y;
                        } else if (III == JJJ && III == KKK && III < LLL) {
                            double DJI = DEN(LLL, III), DII = DEN(III, III);
                            O(LLL, III) += 0.5 * DII * Y;
                            O(III, III) += DJI * Y;
                        } else if (III == KKK && JJJ == LLL && III < JJJ) {
                            double DJI = DEN(JJJ, III), DJJ = DEN(JJJ, JJJ), DII = DEN(III, III);
                            O(JJJ, III) += 1.5 * DJI * Y;
                            O(JJJ, JJJ) -= 0.5 * DII * Y;
                            O(III, III) -= 0.5 * DJJ * Y;
                        } else if (III == KKK && III < JJJ && JJJ < LLL) {
                            double DKI = DEN(LLL, III), DKJ = DEN(LLL, JJJ), DII = DEN(III, III), DJI = DEN(JJJ, III);
                            O(JJJ, III) += 1.5 * DKI * Y;
                            O(LLL, III) += 1.5 * DJI * Y;
                            O(III, III) -= 1.0 * DKJ * Y;
                            O(LLL, JJJ) -= 0.5 * DII * Y;
                        }
                    }
                }
}

/* Symmetric J/K update for one unique component quartet with weight w on the
 * quartet (used by the GPU-enumeration mode).  X accumulates so that the
 * two-electron operator is (X + X^T)/2. */
static void digest_sym(double *X, int n, const double *D, int a, int b, int cc, int d, double v, double w, double hyb) {
    X[IX2(a, b, n)] += 4.0 * w * D[IX2(cc, d, n)] * v;
    X[IX2(cc, d, n)] += 4.0 * w * D[IX2(a, b, n)] * v;
    double k = hyb * w * v;
    X[IX2(a, cc, n)] -= k * D[IX2(b, d, n)];
    X[IX2(a, d, n)] -= k * D[IX2(b, cc, n)];
    X[IX2(b, cc, n)] -= k * D[IX2(a, d, n)];
    X[IX2(b, d, n)] -= k * D[IX2(a, cc, n)];
}

/* orbit size / 8 of the basis-function quartet (ab|cd) under the 8-fold symmetry */
static double quartet_weight(int a, int b, int cc, int d) {
    int ab = a == b, cd = cc == d, pq = (a == cc && b == d) || (a == d && b == cc);
    double w = 1.0;
    if (ab) w *= 0.5;
    if (cd) w *= 0.5;
    if (pq) w *= 0.5;
    return w;
}

/* mode 0: CPU enumeration (hfoperator get2e), mode 1: GPU enumeration.
 * Returns the number of surviving contracted shell (mode 0) or Qshell (mode 1)
 * quartets; *nprimq (if non-null) receives the primitive-quartet count. */
long qko_fock(qko_t *c, const double *D, double *G, double tau, double primLimit, double pair_cut,
              double hyb, int mode, long *nprimq) {
    /* mode 1: Boys order = sum of the Qshell types (reference GPU, gpu_get2e_subs.h:640);
     * mode 2: same enumeration, Boys order = sum of the parent shells' Qfinal
     * (reference CPU, shell.f90:509) -- the convention the HIP path defaults to. */
    int ns = c->nshell, n = c->nbasis;
    double *cm = calloc((size_t)ns * ns, sizeof(double));
    double *o = calloc((size_t)n * n, sizeof(double));
    density_cutoff(c, D, cm);
    long nq = 0, np = 0;
    const double *Y = c->Ycutoff;
    if (mode == 0) {
        for (int II = 0; II < ns; ++II)
            for (int JJ = II; JJ < ns; ++JJ) {
                double t = Y[IX2(II, JJ, ns)];
                for (int KK = II; KK < ns; ++KK)
                    for (int LL = KK; LL < ns; ++LL) {
                        double ct = t * Y[IX2(KK, LL, ns)];
                        if (!(ct > tau)) continue;
                        double dn = dnmax_of(cm, ns, II, JJ, KK, LL);
                        if (!(ct * dn > tau)) continue;
                        ++nq;
                        for (int I = c->Qstart[II]; I <= c->Qfinal[II]; ++I)
                            for (int J = c->Qstart[JJ]; J <= c->Qfinal[JJ]; ++J)
                                for (int K = c->Qstart[KK]; K <= c->Qfinal[KK]; ++K)
                                    for (int L = c->Qstart[LL]; L <= c->Qfinal[LL]; ++L) {
                                        np += contracted_store(c, II, JJ, KK, LL, I, J, K, L, dn, primLimit, 1, 0,
                                                               c->Qfinal[II] + c->Qfinal[JJ] + c->Qfinal[KK] + c->Qfinal[LL]);
                                        iclass_cpu(c, II, JJ, KK, LL, I, J, K, L, D, o);
                                    }
                    }
            }
        /* copySym: upper <- lower */
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) O(i, j) = O(j, i);
        memcpy(G, o, sizeof(double) * n * n);
    } else {
        /* Qshell list in shell order (the reference's sorted order only
         * changes the visiting order, not the set) */
        int nq_ = 0;
        int *qs = malloc(sizeof(int) * 4 * ns), *qt = malloc(sizeof(int) * 4 * ns);
        for (int s = 0; s < ns; ++s)
            for (int L = c->Qstart[s]; L <= c->Qfinal[s]; ++L) { qs[nq_] = s; qt[nq_] = L; ++nq_; }
        int npair = 0;
        int *pa = malloc(sizeof(int) * nq_ * nq_), *pb = malloc(sizeof(int) * nq_ * nq_);
        for (int i = 0; i < nq_; ++i)
            for (int j = 0; j < nq_; ++j)
                if (qs[i] <= qs[j] && Y[IX2(qs[i], qs[j], ns)] > pair_cut) { pa[npair] = i; pb[npair] = j; ++npair; }
        for (int x = 0; x < npair; ++x)
            for (int yv = 0; yv < npair; ++yv) {
                int QI = pa[x], QJ = pb[x], QK = pa[yv], QL = pb[yv];
                int II = qs[QI], JJ = qs[QJ], KK = qs[QK], LL = qs[QL];
                if (!(II <= KK)) continue;
                double ct = Y[IX2(KK, LL, ns)] * Y[IX2(II, JJ, ns)];
                double dn = dnmax_of(cm, ns, II, JJ, KK, LL);
                if (!(ct > tau && ct * dn > tau)) continue;
                int I = qt[QI], J = qt[QJ], K = qt[QK], L = qt[QL];
                int III1, III2, JJJ1, JJJ2, KKK1, KKK2, LLL1, LLL2;
                qrange(c, II, I, &III1, &III2); qrange(c, JJ, J, &JJJ1, &JJJ2);
                qrange(c, KK, K, &KKK1, &KKK2); qrange(c, LL, L, &LLL1, &LLL2);
                int any = 0;
                for (int III = III1; III <= III2; ++III)
                    for (int JJJ = (III > JJJ1 ? III : JJJ1); JJJ <= JJJ2; ++JJJ)
                        for (int KKK = (III > KKK1 ? III : KKK1); KKK <= KKK2; ++KKK)
                            for (int LLL = (KKK > LLL1 ? KKK : LLL1); LLL <= LLL2; ++LLL) {
                                if (III < KKK || (III == JJJ && III == LLL) || (III == JJJ && III < LLL) ||
                                    (JJJ == LLL && III < JJJ) || (III == KKK && III < JJJ && JJJ < LLL)) {
                                    if (!any) {
                                        int Mb = mode == 2 ? c->Qfinal[II] + c->Qfinal[JJ] + c->Qfinal[KK] + c->Qfinal[LL]
                                                           : I + J + K + L;
                                        np += contracted_store(c, II, JJ, KK, LL, I, J, K, L, dn, primLimit, 1, 0, Mb);
                                        any = 1; ++nq;
                                    }
                                    double v = hrr_value(c, II, JJ, KK, LL, III, JJJ, KKK, LLL);
                                    if (fabs(v) > 0.0)
                                        digest_sym(o, n, D, III, JJJ, KKK, LLL, v, quartet_weight(III, JJJ, KKK, LLL), hyb);
                                }
                            }
            }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                double v = 0.5 * (O(i, j) + O(j, i));
                G[IX2(i, j, n)] = G[IX2(j, i, n)] = v;
            }
        free(qs); free(qt); free(pa); free(pb);
    }
    free(cm); free(o);
    if (nprimq) *nprimq = np;
    return nq;
}

/* ---------------- four-centre gradient ---------------- */
/* Contracted stores for the first derivative (shellopt + classopt,
 * shell.f90:1863-2114): VRR raised by one on both sides, FmT asked for order
 * Mboys, primitive quartets kept when dnmax*cutprim*cutprim > cutoff.
 * gstore[0] carries the plain contraction weights X44, gstore[1..3] the same
 * times 2a, 2b, 2c (exponents of the primitives on A, B, C). */
static double gstore[4][NCART][NCART];

static long contracted_store_grad(qko_t *c, int II, int JJ, int KK, int LL, int I, int J, int K, int L,
                                  double dnmax, double cutoff, int Mboys) {
    int jb = c->jbasis;
    int LE = I + J + 1, LF = K + L + 1;
    int ne = ncart_upto[LE + 1], nf = ncart_upto[LF + 1];
    for (int g = 0; g < 4; ++g)
        for (int e = 0; e < ne; ++e) for (int f = 0; f < nf; ++f) gstore[g][e][f] = 0.0;
    const double X0 = 2.0 * pow(PI, 2.5);
    double RA[3], RC[3];
    for (int m = 0; m < 3; ++m) { RA[m] = ATOMXYZ(m, c->katom[II] - 1); RC[m] = ATOMXYZ(m, c->katom[KK] - 1); }
    long nq = 0;
    double Fm[MMAX + 1];
    for (int jj = 0; jj < c->kprim[JJ]; ++jj)
        for (int ii = 0; ii < c->kprim[II]; ++ii) {
            int Pi = c->kstart[II] - 1 + ii, Pj = c->kstart[JJ] - 1 + jj;
            double AB = c->Apri[IX2(Pi, Pj, jb)];
            const double *P = &c->Ppri[3 * IX2(Pi, Pj, jb)];
            double X2 = X0 * XC(Pi, Pj, I, J);
            double cp1 = dnmax * c->cutprim[IX2(Pi, Pj, jb)];
            double aa = GCE(ii, c->ksumtype[II] - 1), bb = GCE(jj, c->ksumtype[JJ] - 1);
            for (int ll = 0; ll < c->kprim[LL]; ++ll)
                for (int kk = 0; kk < c->kprim[KK]; ++kk) {
                    int Pk = c->kstart[KK] - 1 + kk, Pl = c->kstart[LL] - 1 + ll;
                    if (!(cp1 * c->cutprim[IX2(Pk, Pl, jb)] > cutoff)) continue;
                    double cc = GCE(kk, c->ksumtype[KK] - 1);
                    double CD = c->Apri[IX2(Pk, Pl, jb)];
                    const double *Q = &c->Ppri[3 * IX2(Pk, Pl, jb)];
                    double ABCD = AB + CD, ROU = AB * CD / ABCD, RPQ = 0.0;
                    vrr_par vp;
                    for (int m = 0; m < 3; ++m) {
                        double W = (P[m] * AB + Q[m] * CD) / ABCD;
                        double d = P[m] - Q[m];
                        RPQ += d * d;
                        vp.PA[m] = P[m] - RA[m];
                        vp.QC[m] = Q[m] - RC[m];
                        vp.WP[m] = W - P[m];
                        vp.WQ[m] = W - Q[m];
                    }
                    vp.oo2z = 0.5 / AB; vp.oo2e = 0.5 / CD; vp.oo2ze = 0.5 / ABCD;
                    vp.roz = CD / ABCD; vp.roe = AB / ABCD;
                    double T = RPQ * ROU;
                    if (c->boys_exact) qko_fmt_exact(Mboys, T, Fm); else qko_fmt(Mboys, T, Fm);
                    double s = 1.0 / sqrt(ABCD);
                    for (int m = 0; m <= Mboys; ++m) Fm[m] *= s;
                    vrr(LE, LF, Fm, &vp, vtmp);
                    double X44 = X2 * XC(Pk, Pl, K, L);
                    double w[4] = {X44, X44 * aa * 2.0, X44 * bb * 2.0, X44 * cc * 2.0};
                    for (int g = 0; g < 4; ++g)
                        for (int e = 0; e < ne; ++e)
                            for (int f = 0; f < nf; ++f) gstore[g][e][f] += w[g] * vtmp[e][f];
                    ++nq;
                }
        }
    return nq;
}

/* sum over the HRR expansions of (a b| and |c d) of store[e][f] */
static double hrr_pair(const int *a, const int *b, const int *cc, const int *d, const double *AB, const double *CD,
                       double st[NCART][NCART]) {
    double cl_[32], cr[32];
    int il[32], ir[32];
    int nl = hrr_expand(a, b, AB, cl_, il);
    int nr = hrr_expand(cc, d, CD, cr, ir);
    double Y = 0.0;
    for (int x = 0; x < nl; ++x)
        for (int y = 0; y < nr; ++y) Y += cl_[x] * cr[y] * st[il[x]][ir[y]];
    return Y;
}

/* d(iii jjj|kkk lll)/dA, /dB, /dC (hrrwholeopt, subs/hrr.f90:862-1014):
 * d/dA_x (a b| = 2a (a+1_x b| - a_x (a-1_x b|, likewise for B and C, with the
 * exponent factors carried by gstore[1..3]; times the four cons factors. */
static void hrr_grad(qko_t *c, int II, int JJ, int KK, int LL, int iii, int jjj, int kkk, int lll,
                     double *Yaa, double *Ybb, double *Ycc) {
    double AB[3], CD[3];
    for (int m = 0; m < 3; ++m) {
        AB[m] = ATOMXYZ(m, c->katom[II] - 1) - ATOMXYZ(m, c->katom[JJ] - 1);
        CD[m] = ATOMXYZ(m, c->katom[KK] - 1) - ATOMXYZ(m, c->katom[LL] - 1);
    }
    const int *a = &c->KLMN[3 * iii], *b = &c->KLMN[3 * jjj], *cc = &c->KLMN[3 * kkk], *d = &c->KLMN[3 * lll];
    double cons4 = c->cons[iii] * c->cons[jjj] * c->cons[kkk] * c->cons[lll];
    for (int x = 0; x < 3; ++x) {
        int t[3];
        memcpy(t, a, sizeof t); t[x] += 1;
        double v = hrr_pair(t, b, cc, d, AB, CD, gstore[1]);
        if (a[x] >= 1) { memcpy(t, a, sizeof t); t[x] -= 1; v -= a[x] * hrr_pair(t, b, cc, d, AB, CD, gstore[0]); }
        Yaa[x] = v * cons4;
        memcpy(t, b, sizeof t); t[x] += 1;
        v = hrr_pair(a, t, cc, d, AB, CD, gstore[2]);
        if (b[x] >= 1) { memcpy(t, b, sizeof t); t[x] -= 1; v -= b[x] * hrr_pair(a, t, cc, d, AB, CD, gstore[0]); }
        Ybb[x] = v * cons4;
        memcpy(t, cc, sizeof t); t[x] += 1;
        v = hrr_pair(a, b, t, d, AB, CD, gstore[3]);
        if (cc[x] >= 1) { memcpy(t, cc, sizeof t); t[x] -= 1; v -= cc[x] * hrr_pair(a, b, t, d, AB, CD, gstore[0]); }
        Ycc[x] = v * cons4;
    }
}

/* The density factor of one unique function quartet in classopt's branches
 * (shell.f90:2190-2460), with the exchange part scaled by hyb (1 = HF, as the
 * reference). Returns 0 for quartets the branches skip. */
static double grad_constant(const double *D, int n, int III, int JJJ, int KKK, int LLL, double hyb, int *use) {
    *use = 1;
    if (III < KKK) {
        if (III < JJJ && KKK < LLL)
            return 4.0 * DEN(JJJ, III) * DEN(LLL, KKK) - hyb * (DEN(KKK, III) * DEN(LLL, JJJ) + DEN(LLL, III) * DEN(KKK, JJJ));
        if (III == JJJ && KKK == LLL)
            return DEN(III, III) * DEN(KKK, KKK) - hyb * 0.5 * DEN(KKK, III) * DEN(KKK, III);
        if (JJJ == KKK && JJJ == LLL)
            return (2.0 - hyb) * DEN(JJJ, JJJ) * DEN(JJJ, III);
        if (KKK == LLL && III < JJJ && JJJ != KKK)
            return 2.0 * DEN(JJJ, III) * DEN(KKK, KKK) - hyb * DEN(KKK, III) * DEN(KKK, JJJ);
        if (III == JJJ && KKK < LLL)
            return 2.0 * DEN(LLL, KKK) * DEN(III, III) - hyb * DEN(KKK, III) * DEN(LLL, III);
        *use = 0;
        return 0.0;
    }
    if (JJJ <= LLL) {
        if (III == JJJ && III == KKK && III == LLL) { *use = 0; return 0.0; }
        if (III == JJJ && III == KKK && III < LLL) return (2.0 - hyb) * DEN(LLL, III) * DEN(III, III);
        if (III == KKK && JJJ == LLL && III < JJJ)
            return (2.0 - 0.5 * hyb) * DEN(JJJ, III) * DEN(JJJ, III) - 0.5 * hyb * DEN(JJJ, JJJ) * DEN(III, III);
        if (III == KKK && III < JJJ && JJJ < LLL)
            return (4.0 - hyb) * DEN(JJJ, III) * DEN(LLL, III) - hyb * DEN(LLL, JJJ) * DEN(III, III);
    }
    *use = 0;
    return 0.0;
}

/* Four-centre part of the nuclear gradient, dE2/dR with E2 = 1/2 sum D G(D) at
 * fixed D (hfgrad, gradient.f90:266-320): shell quartets II<=JJ, II<=KK<=LL,
 * one-centre quartets skipped, Y_IJ*Y_KL > cutoff and Y_IJ*Y_KL*DNmax > cutoff,
 * FmT order = sum of the four shells' Qfinal + 2 (shell.f90:1906-1912);
 * dE/dD = -(dE/dA + dE/dB + dE/dC) (classopt:2463-2520).
 * grad is 3 x natom (column-major), accumulated into. */
long qko_grad(qko_t *c, const double *D, double *grad, double cutoff, double hyb, long *nprimq) {
    int ns = c->nshell, n = c->nbasis;
    double *cm = calloc((size_t)ns * ns, sizeof(double));
    density_cutoff(c, D, cm);
    const double *Y = c->Ycutoff;
    long nq = 0, np = 0;
    for (int II = 0; II < ns; ++II)
        for (int JJ = II; JJ < ns; ++JJ)
            for (int KK = II; KK < ns; ++KK)
                for (int LL = KK; LL < ns; ++LL) {
                    int a0 = c->katom[II];
                    if (a0 == c->katom[JJ] && a0 == c->katom[KK] && a0 == c->katom[LL]) continue;
                    double ct = Y[IX2(II, JJ, ns)] * Y[IX2(KK, LL, ns)];
                    if (!(ct > cutoff)) continue;
                    double dn = dnmax_of(cm, ns, II, JJ, KK, LL);
                    if (!(ct * dn > cutoff)) continue;
                    ++nq;
                    int Mb = c->Qfinal[II] + c->Qfinal[JJ] + c->Qfinal[KK] + c->Qfinal[LL] + 2;
                    int iA = c->katom[II] - 1, iB = c->katom[JJ] - 1, iC = c->katom[KK] - 1, iD = c->katom[LL] - 1;
                    for (int I = c->Qstart[II]; I <= c->Qfinal[II]; ++I)
                        for (int J = c->Qstart[JJ]; J <= c->Qfinal[JJ]; ++J)
                            for (int K = c->Qstart[KK]; K <= c->Qfinal[KK]; ++K)
                                for (int L = c->Qstart[LL]; L <= c->Qfinal[LL]; ++L) {
                                    np += contracted_store_grad(c, II, JJ, KK, LL, I, J, K, L, dn, cutoff, Mb);
                                    int III1, III2, JJJ1, JJJ2, KKK1, KKK2, LLL1, LLL2;
                                    qrange(c, II, I, &III1, &III2); qrange(c, JJ, J, &JJJ1, &JJJ2);
                                    qrange(c, KK, K, &KKK1, &KKK2); qrange(c, LL, L, &LLL1, &LLL2);
                                    double Ag[3] = {0, 0, 0}, Bg[3] = {0, 0, 0}, Cg[3] = {0, 0, 0};
                                    double Yaa[3], Ybb[3], Ycc[3];
                                    int fast = II < JJ && II < KK && KK < LL;
                                    for (int III = III1; III <= III2; ++III)
                                        for (int JJJ = fast ? JJJ1 : (III > JJJ1 ? III : JJJ1); JJJ <= JJJ2; ++JJJ)
                                            for (int KKK = fast ? KKK1 : (III > KKK1 ? III : KKK1); KKK <= KKK2; ++KKK)
                                                for (int LLL = fast ? LLL1 : (KKK > LLL1 ? KKK : LLL1); LLL <= LLL2; ++LLL) {
                                                    int use = 1;
                                                    double k;
                                                    if (fast)
                                                        k = 4.0 * DEN(JJJ, III) * DEN(LLL, KKK) -
                                                            hyb * (DEN(KKK, III) * DEN(LLL, JJJ) + DEN(LLL, III) * DEN(KKK, JJJ));
                                                    else
                                                        k = grad_constant(D, n, III, JJJ, KKK, LLL, hyb, &use);
                                                    if (!use) continue;
                                                    hrr_grad(c, II, JJ, KK, LL, III, JJJ, KKK, LLL, Yaa, Ybb, Ycc);
                                                    for (int x = 0; x < 3; ++x) {
                                                        Ag[x] += Yaa[x] * k;
                                                        Bg[x] += Ybb[x] * k;
                                                        Cg[x] += Ycc[x] * k;
                                                    }
                                                }
                                    for (int x = 0; x < 3; ++x) {
                                        grad[3 * iA + x] += Ag[x];
                                        grad[3 * iB + x] += Bg[x];
                                        grad[3 * iC + x] += Cg[x];
                                        grad[3 * iD + x] += -Ag[x] - Bg[x] - Cg[x];
                                    }
                                }
                }
    free(cm);
    if (nprimq) *nprimq = np;
    return nq;
}

/* Contracted (ij|kl) for basis functions (0-based), no screening. */
double qko_eri(qko_t *c, int i, int j, int k, int l) {
    int sh[4], ty[4], f[4] = {i, j, k, l};
    for (int q = 0; q < 4; ++q) {
        for (int s = 0; s < c->nshell; ++s)
            if (f[q] >= c->ksumtype[s] - 1 && f[q] < c->ksumtype[s + 1] - 1) { sh[q] = s; break; }
        for (int L = c->Qstart[sh[q]]; L <= c->Qfinal[sh[q]]; ++L) {
            int a0, a1;
            qrange(c, sh[q], L, &a0, &a1);
            if (f[q] >= a0 && f[q] <= a1) ty[q] = L;
        }
    }
    contracted_store(c, sh[0], sh[1], sh[2], sh[3], ty[0], ty[1], ty[2], ty[3], 0.0, 0.0, 0, 0,
                     ty[0] + ty[1] + ty[2] + ty[3]);
    return hrr_value(c, sh[0], sh[1], sh[2], sh[3], i, j, k, l);
}
