/*
 * qk_oneint_oracle.c -- CPU restatement of QUICK's one-electron integrals and
 * their nuclear-gradient terms.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ and __graft_entry__.smoke() may load
 * this code, and only as the checker of the HIP one-electron kernels; the
 * product path (quick_amd/csrc) never links or calls it.
 *
 * What it restates (reference file:line):
 *   overlap         primitive Cartesian overlap by the binomial expansion of
 *                   the Gaussian product (overlap, quick_one_electron_integral.f90:193-306)
 *   kinetic         -1/2 <a|lap|b> as overlaps with the powers of the first
 *                   function shifted by -2, 0, +2 (ekinetic, :130-190)
 *   attraction      (a|1/r_C|b)^(m) by the Obara-Saika recursion from
 *                   (0|1/r_C|0)^(m) = 2 sqrt(g/pi) (0||0) F_m(g |P-C|^2)
 *                   (attrashell, :1816-1925; nuclearattra, nuclear.f90:4-225);
 *                   F_m is QUICK's fit (qko_fmt, subs/fmt.f90) asked for the
 *                   order Qfinal(I) + Qfinal(J) as attrashell does
 *   qko1_hcore      S and H = T + V over the basis with the per-function
 *                   coefficients dcoeff (get1e/get1eO, :1633-1809); dcoeff of a
 *                   function equals gccoeff(k, ksumtype + L) cons, the
 *                   coefficient attrashell applies (nuclear.f90:196-221)
 *   qko1_grad       the one-electron part of hfgrad (gradient.f90:40-250):
 *                   nuclear repulsion, -2 W dS/dR + 2 D dT/dR over function
 *                   pairs on different atoms, and D dV/dR with
 *                   d/dA_x g_a = 2 alpha g_{a+1x} - a_x g_{a-1x} on A and B and
 *                   the nucleus term by translational invariance, skipping the
 *                   nucleus when both functions sit on it (attrashellopt,
 *                   shell.f90:1747-1860, Boys order Qfinal(I) + Qfinal(J) + 2)
 *
 * Parity pins: tests/test_oneint_oracle.py checks S and H against the
 * reference's own get1e/overlap output (tests/golden/*_scf.npz), and the
 * gradient against the reference's full hfgrad gradient minus its four-centre
 * term; also against central finite differences and translational invariance.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

void qko_fmt(int MaxM, double X, double *vals);

static const double PI1 = 3.1415926535897932384626433832795;

/* ---------------- overlap and kinetic (primitive, unnormalised) ---------------- */

static double fact(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
}

/* overlap of x^i y^j z^k e^{-a r_A^2} and x^ii y^jj z^kk e^{-b r_B^2} */
static double overlap(double a, double b, int i, int j, int k, int ii, int jj, int kk, const double *A,
                      const double *B) {
    if (i < 0 || j < 0 || k < 0 || ii < 0 || jj < 0 || kk < 0) return 0.0;
    const double g = a + b;
    const double P[3] = {(a * A[0] + b * B[0]) / g, (a * A[1] + b * B[1]) / g, (a * A[2] + b * B[2]) / g};
    const double PA[3] = {P[0] - A[0], P[1] - A[1], P[2] - A[2]}, PB[3] = {P[0] - B[0], P[1] - B[1], P[2] - B[2]};
    const int la[3] = {i, j, k}, lb[3] = {ii, jj, kk};
    double s = 1.0;
    for (int d = 0; d < 3; ++d) {
        /* one dimension: sum over the binomial terms of even total power */
        double e = 0.0;
        for (int p = 0; p <= la[d]; ++p)
            for (int q = 0; q <= lb[d]; ++q) {
                const int t = p + q;
                if (t & 1) continue;
                double term = fact(la[d]) / (fact(p) * fact(la[d] - p)) * fact(lb[d]) / (fact(q) * fact(lb[d] - q)) *
                              pow(PA[d], la[d] - p) * pow(PB[d], lb[d] - q);
                /* integral of x^t e^{-g x^2}: Gamma((t+1)/2) / g^{(t+1)/2} */
                double gm = sqrt(PI1);
                for (int r = 1; r <= t / 2; ++r) gm *= ((double)t / 2.0 - (double)r + 0.5);
                e += term * gm * pow(g, -(double)(t + 1) / 2.0);
            }
        s *= e;
    }
    const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
    return s * exp(-a * b * AB2 / g);
}

/* -1/2 <a | laplacian | b>, the derivatives applied to the first function */
static double kinetic(double a, double b, int i, int j, int k, int ii, int jj, int kk, const double *A,
                      const double *B) {
    double t = 0.0;
    const int l[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
        int m[3] = {i, j, k};
        const double x = l[d];
        m[d] = l[d] - 2;
        t += (x - 1.0) * x * overlap(a, b, m[0], m[1], m[2], ii, jj, kk, A, B);
        m[d] = l[d];
        t -= 2.0 * a * (1.0 + 2.0 * x) * overlap(a, b, m[0], m[1], m[2], ii, jj, kk, A, B);
        m[d] = l[d] + 2;
        t += 4.0 * a * a * overlap(a, b, m[0], m[1], m[2], ii, jj, kk, A, B);
    }
    return -0.5 * t;
}

/* ---------------- nuclear attraction (Obara-Saika, memoised per primitive pair and nucleus) ---------------- */

#define NA_MAXL 4      /* per-component power: d + 1 (gradient) + 1 (recursion) */
#define NA_MAXM 12
typedef struct {
    double PA[3], PB[3], PC[3], oo2g;
    double aux[NA_MAXM + 1];
    double val[NA_MAXL * NA_MAXL * NA_MAXL * NA_MAXL * NA_MAXL * NA_MAXL][NA_MAXM + 1];
    unsigned gen[NA_MAXL * NA_MAXL * NA_MAXL * NA_MAXL * NA_MAXL * NA_MAXL][NA_MAXM + 1];
    unsigned cur;
} attra_t;

static double attra(attra_t *s, const int *a, const int *b, int m) {
    for (int d = 0; d < 3; ++d)
        if (a[d] < 0 || b[d] < 0) return 0.0;
    if (a[0] + a[1] + a[2] + b[0] + b[1] + b[2] == 0) return s->aux[m];
    const int key = ((((a[0] * NA_MAXL + a[1]) * NA_MAXL + a[2]) * NA_MAXL + b[0]) * NA_MAXL + b[1]) * NA_MAXL + b[2];
    if (s->gen[key][m] == s->cur) return s->val[key][m];
    double v;
    int d, on_a = (a[0] + a[1] + a[2]) > 0;
    if (on_a) {
        d = a[0] > 0 ? 0 : (a[1] > 0 ? 1 : 2);
        int a1[3] = {a[0], a[1], a[2]}, a2[3], b1[3] = {b[0], b[1], b[2]};
        a1[d] -= 1;
        memcpy(a2, a1, sizeof(a2));
        a2[d] -= 1;
        b1[d] -= 1;
        v = s->PA[d] * attra(s, a1, b, m) - s->PC[d] * attra(s, a1, b, m + 1) +
            a1[d] * s->oo2g * (attra(s, a2, b, m) - attra(s, a2, b, m + 1)) +
            b[d] * s->oo2g * (attra(s, a1, b1, m) - attra(s, a1, b1, m + 1));
    } else {
        d = b[0] > 0 ? 0 : (b[1] > 0 ? 1 : 2);
        int b1[3] = {b[0], b[1], b[2]}, b2[3];
        b1[d] -= 1;
        memcpy(b2, b1, sizeof(b2));
        b2[d] -= 1;
        v = s->PB[d] * attra(s, a, b1, m) - s->PC[d] * attra(s, a, b1, m + 1) +
            b1[d] * s->oo2g * (attra(s, a, b2, m) - attra(s, a, b2, m + 1));
    }
    s->val[key][m] = v;
    s->gen[key][m] = s->cur;
    return v;
}

/* the auxiliary integrals of one primitive pair and one nucleus of charge Z:
 * aux(m) = -Z 2 sqrt(g/pi) (0||0) F_m(g |P-C|^2), F asked for order M */
static void attra_setup(attra_t *s, double a, double b, const double *A, const double *B, const double *C, double Z,
                        int M) {
    const double g = a + b;
    double P[3];
    for (int d = 0; d < 3; ++d) {
        P[d] = (a * A[d] + b * B[d]) / g;
        s->PA[d] = P[d] - A[d];
        s->PB[d] = P[d] - B[d];
        s->PC[d] = P[d] - C[d];
    }
    s->oo2g = 0.5 / g;
    const double ss = overlap(a, b, 0, 0, 0, 0, 0, 0, A, B) * 2.0 * sqrt(g / PI1);
    const double U = g * (s->PC[0] * s->PC[0] + s->PC[1] * s->PC[1] + s->PC[2] * s->PC[2]);
    double F[NA_MAXM + 1];
    qko_fmt(M, U, F);
    for (int m = 0; m <= M; ++m) s->aux[m] = F[m] * ss * (-Z);
    for (int m = M + 1; m <= NA_MAXM; ++m) s->aux[m] = 0.0;
    s->cur += 1;
}

/* ---------------- matrices ---------------- */

typedef struct {
    int natom, nbasis, maxc;
    const double *xyz, *chg;
    const int *ncenter, *itype, *ncontract, *qf;
    const double *aexp, *dcoeff;
} qko1_sys;

/* S and H (n x n, column-major, full symmetric) */
void qko1_hcore(int natom, const double *xyz, const double *chg, int nbasis, int maxc, const int *ncenter,
                const int *itype, const int *ncontract, const int *qf, const double *aexp, const double *dcoeff,
                double *S, double *H) {
    attra_t *st = calloc(1, sizeof(attra_t));
    for (int I = 0; I < nbasis; ++I) {
        const double *A = xyz + 3 * (ncenter[I] - 1);
        const int *li = itype + 3 * I;
        for (int J = I; J < nbasis; ++J) {
            const double *B = xyz + 3 * (ncenter[J] - 1);
            const int *lj = itype + 3 * J;
            double s = 0.0, t = 0.0, v = 0.0;
            for (int ci = 0; ci < ncontract[I]; ++ci)
                for (int cj = 0; cj < ncontract[J]; ++cj) {
                    const double a = aexp[ci + maxc * I], b = aexp[cj + maxc * J];
                    const double cc = dcoeff[ci + maxc * I] * dcoeff[cj + maxc * J];
                    s += cc * overlap(a, b, li[0], li[1], li[2], lj[0], lj[1], lj[2], A, B);
                    t += cc * kinetic(b, a, lj[0], lj[1], lj[2], li[0], li[1], li[2], B, A);
                    for (int C = 0; C < natom; ++C) {
                        attra_setup(st, a, b, A, B, xyz + 3 * C, chg[C], qf[I] + qf[J]);
                        v += cc * attra(st, li, lj, 0);
                    }
                }
            S[I + (long)nbasis * J] = S[J + (long)nbasis * I] = s;
            H[I + (long)nbasis * J] = H[J + (long)nbasis * I] = t + v;
        }
    }
    free(st);
}

/* grad (3 x natom) += the one-electron part of dE/dR for the density D and the
 * energy-weighted density W (both n x n, full): nuclear repulsion, -W dS/dR,
 * D dT/dR and D dV/dR */
void qko1_grad(int natom, const double *xyz, const double *chg, int nbasis, int maxc, const int *ncenter,
               const int *itype, const int *ncontract, const int *qf, const double *aexp, const double *dcoeff,
               const double *D, const double *W, double *grad) {
    for (int I = 0; I < natom; ++I)
        for (int J = I + 1; J < natom; ++J) {
            double d[3], r2 = 0.0;
            for (int x = 0; x < 3; ++x) {
                d[x] = xyz[3 * J + x] - xyz[3 * I + x];
                r2 += d[x] * d[x];
            }
            const double f = chg[I] * chg[J] * pow(r2, -1.5);
            for (int x = 0; x < 3; ++x) {
                grad[3 * I + x] += d[x] * f;
                grad[3 * J + x] -= d[x] * f;
            }
        }
    attra_t *st = calloc(1, sizeof(attra_t));
    for (int I = 0; I < nbasis; ++I) {
        const int atA = ncenter[I] - 1;
        const double *A = xyz + 3 * atA;
        for (int J = I; J < nbasis; ++J) {
            const int atB = ncenter[J] - 1;
            const double *B = xyz + 3 * atB;
            const double wt = (I == J) ? 1.0 : 2.0;   /* (I,J) and (J,I) */
            const double dIJ = D[I + (long)nbasis * J] * wt, wIJ = W[I + (long)nbasis * J] * wt;
            for (int ci = 0; ci < ncontract[I]; ++ci)
                for (int cj = 0; cj < ncontract[J]; ++cj) {
                    const double a = aexp[ci + maxc * I], b = aexp[cj + maxc * J];
                    const double cc = dcoeff[ci + maxc * I] * dcoeff[cj + maxc * J];
                    for (int x = 0; x < 3; ++x) {
                        int la[3] = {itype[3 * I], itype[3 * I + 1], itype[3 * I + 2]};
                        int lb[3] = {itype[3 * J], itype[3 * J + 1], itype[3 * J + 2]};
                        const int ax = la[x], bx = lb[x];
                        /* d/dA_x and d/dB_x of S and T: two-centre, so only pairs on different atoms */
                        if (atA != atB) {
                            double dSA, dTA;
                            la[x] = ax + 1;
                            dSA = 2.0 * a * overlap(a, b, la[0], la[1], la[2], lb[0], lb[1], lb[2], A, B);
                            dTA = 2.0 * a * kinetic(b, a, lb[0], lb[1], lb[2], la[0], la[1], la[2], B, A);
                            la[x] = ax - 1;
                            dSA -= ax * overlap(a, b, la[0], la[1], la[2], lb[0], lb[1], lb[2], A, B);
                            dTA -= ax * kinetic(b, a, lb[0], lb[1], lb[2], la[0], la[1], la[2], B, A);
                            la[x] = ax;
                            grad[3 * atA + x] += cc * (dIJ * dTA - wIJ * dSA);
                            grad[3 * atB + x] -= cc * (dIJ * dTA - wIJ * dSA);
                        }
                        /* d/dA_x, d/dB_x of (a|V_C|b); the nucleus by translational invariance */
                        for (int C = 0; C < natom; ++C) {
                            if (atA == C && atB == C) continue;
                            attra_setup(st, a, b, A, B, xyz + 3 * C, chg[C], qf[I] + qf[J] + 2);
                            double dA, dB;
                            la[x] = ax + 1;
                            dA = 2.0 * a * attra(st, la, lb, 0);
                            la[x] = ax - 1;
                            dA -= ax * attra(st, la, lb, 0);
                            la[x] = ax;
                            lb[x] = bx + 1;
                            dB = 2.0 * b * attra(st, la, lb, 0);
                            lb[x] = bx - 1;
                            dB -= bx * attra(st, la, lb, 0);
                            lb[x] = bx;
                            grad[3 * atA + x] += cc * dIJ * dA;
                            grad[3 * atB + x] += cc * dIJ * dB;
                            grad[3 * C + x] -= cc * dIJ * (dA + dB);
                        }
                    }
                }
        }
    }
    free(st);
}
