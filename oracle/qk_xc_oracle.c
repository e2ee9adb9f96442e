/*
 * qk_xc_oracle.c -- CPU restatement of QUICK's B3LYP exchange-correlation step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ and __graft_entry__.smoke() may load
 * this code, and only as the checker of the HIP XC path; the product path
 * (quick_amd/csrc) never links or calls it.
 *
 * What it restates (reference file:line):
 *   qkx_b3lyp       closed-shell B3LYP energy density f(rho, sigma) with rho the
 *                   total density and sigma = |grad rho|^2 -- what b3lyp_e /
 *                   b3lypf evaluate (dft.f90:108-312) -- written here from the
 *                   published formulas (Slater, Becke 88, VWN-RPA, LYP in the
 *                   Miehlich form; coefficients 0.8, 0.72, 0.19, 0.81) in complex
 *                   arithmetic; df/drho and df/dsigma come from complex-step
 *                   differentiation (exact to rounding for an analytic f)
 *   qkx_sigrad2     radius of significance of each basis function
 *                   (get_sigrad, dft.f90:536-613)
 *   qkx_sg1_grid    the SG-1 grid as dftoperator enumerates it: Euler-Maclaurin
 *                   radial points (gridformSG1, dft.f90:10146-10156), the SG-1
 *                   angular pruning by radial region (gridformnew,
 *                   dft.f90:3106-3190), SG-1 radii (quick_constants_module
 *                   RADII), Stratmann-Scuseria-Frisch weights (subs/ssw.f90),
 *                   points with weight >= cut kept (dft.f90:934-936); the
 *                   Lebedev rules are passed in
 *   qkx_xc          the B3LYP XC loop of dftoperator on a given grid
 *                   (dft.f90:906-1018): basis values and gradients (pteval,
 *                   subs/pteval.f90), density and its gradient with denspt's
 *                   skips (subs/denspt.f90), E_xc, electron counts, and
 *                   V_mn += w (vrho phi_m phi_n + 4 vsigma grad rho_a .
 *                   grad(phi_m phi_n)) over m >= n with the quicktest skip
 *
 * Parity pins: tests/test_xc_oracle.py checks these against the reference's own
 * dftoperator output on its own SG-1 grid (tests/golden/*_xc.npz), against
 * finite differences, the uniform-gas limits of the functional's parts, and
 * the electron count.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define XMAXPRIM 10
static const double XPI = 3.1415926535897932384626433832795;

/* ---------------- functional ---------------- */

static double complex b3lyp_c(double complex rho, double complex sigma) {
    const double third = 1.0 / 3.0;
    /* Slater exchange */
    double complex ex = -0.75 * cpow(3.0 / XPI, third) * cpow(rho, 4.0 / 3.0);
    /* Becke 88 gradient correction, spin densities rho/2, |grad rho_s| = sqrt(sigma)/2 */
    const double beta = 0.0042;
    double complex rs = rho / 2.0;
    double complex gs = csqrt(sigma) / 2.0;
    double complex xs = gs / cpow(rs, 4.0 / 3.0);
    double complex dex = -2.0 * beta * cpow(rs, 4.0 / 3.0) * xs * xs / (1.0 + 6.0 * beta * xs * casinh(xs));
    /* VWN RPA, paramagnetic */
    const double A = 0.0310907, x0 = -0.409286, b = 13.0720, c = 42.7198;
    double complex r_s = cpow(3.0 / (4.0 * XPI * rho), third);
    double complex x = csqrt(r_s);
    double complex Xx = x * x + b * x + c;
    const double Xx0 = x0 * x0 + b * x0 + c;
    const double Q = sqrt(4.0 * c - b * b);
    double complex at = catan(Q / (2.0 * x + b));
    double complex eps = A * (clog(x * x / Xx) + 2.0 * b / Q * at
                              - b * x0 / Xx0 * (clog((x - x0) * (x - x0) / Xx) + 2.0 * (b + 2.0 * x0) / Q * at));
    double complex ec_vwn = rho * eps;
    /* LYP at rho_a = rho_b = rho/2, sigma_aa = sigma_bb = sigma_ab = sigma/4 */
    const double la = 0.04918, lb = 0.132, lc = 0.2533, ld = 0.349;
    const double CF = 0.3 * pow(3.0 * XPI * XPI, 2.0 / 3.0);
    double complex ra = rho / 2.0, rb = rho / 2.0, saa = sigma / 4.0, sbb = sigma / 4.0;
    double complex rm = cpow(rho, -third);
    double complex omega = cexp(-lc * rm) / (1.0 + ld * rm) * cpow(rho, -11.0 / 3.0);
    double complex delta = lc * rm + ld * rm / (1.0 + ld * rm);
    double complex ec_lyp =
        -la * 4.0 / (1.0 + ld * rm) * ra * rb / rho
        - la * lb * omega *
              (ra * rb * (pow(2.0, 11.0 / 3.0) * CF * (cpow(ra, 8.0 / 3.0) + cpow(rb, 8.0 / 3.0))
                          + (47.0 / 18.0 - 7.0 / 18.0 * delta) * sigma - (2.5 - delta / 18.0) * (saa + sbb)
                          - (delta - 11.0) / 9.0 * (ra / rho * saa + rb / rho * sbb))
               - 2.0 / 3.0 * rho * rho * sigma + (2.0 / 3.0 * rho * rho - ra * ra) * sbb
               + (2.0 / 3.0 * rho * rho - rb * rb) * saa);
    return 0.8 * ex + 0.72 * dex + 0.19 * ec_vwn + 0.81 * ec_lyp;
}

/* f(rho, sigma) and its partial derivatives (complex step, h = 1e-20 rho or sigma) */
void qkx_b3lyp(double rho, double sigma, double *f, double *vrho, double *vsigma) {
    const double hr = 1e-20 * rho, hs = 1e-20 * (sigma > 1e-300 ? sigma : 1e-300);
    *f = creal(b3lyp_c(rho, sigma));
    *vrho = cimag(b3lyp_c(rho + I * hr, sigma)) / hr;
    *vsigma = cimag(b3lyp_c(rho, sigma + I * hs)) / hs;
}

/* ---------------- basis values ---------------- */

/* radius of significance squared of each function (get_sigrad) */
void qkx_sigrad2(int nbasis, const int *ncontract, const double *aexp, const int *itype, double cutoff,
                 double *out) {
    for (int ib = 0; ib < nbasis; ++ib) {
        double amin = 10.0e10;
        for (int c = 0; c < ncontract[ib]; ++c) amin = fmin(amin, aexp[c + XMAXPRIM * ib]);
        const int L = itype[3 * ib] + itype[3 * ib + 1] + itype[3 * ib + 2];
        double gamma = 1.0;
        for (int i = 1; i <= L + 1; ++i) gamma *= (double)(L + 1 - i) + 0.5;
        const double gamma2pi = gamma * 11.13665599366341569;
        const double target = cutoff * pow(pow(2.0 * amin, (double)L + 1.5) / gamma2pi, -0.5);
        double step = 1.0, r = 0.0;
        while (step > 1e-4) {
            r += step;
            const double cur = exp(-amin * r * r) * pow(r, (double)L);
            if (cur < target) {
                r -= step;
                step /= 10.0;
            }
        }
        out[ib] = r * r;
    }
}

/* value and gradient of function ib at (x, y, z) (pteval) */
static void pteval(double x, double y, double z, int ib, const double *xyz, const int *ncenter, const int *itype,
                   const int *ncontract, const double *dcoeff, const double *aexp, const double *sigrad2,
                   double *phi, double *dx, double *dy, double *dz) {
    const int a = ncenter[ib] - 1;
    const double x1 = x - xyz[3 * a], y1 = y - xyz[3 * a + 1], z1 = z - xyz[3 * a + 2];
    const double r2 = x1 * x1 + y1 * y1 + z1 * z1;
    *phi = *dx = *dy = *dz = 0.0;
    if (r2 > sigrad2[ib]) return;
    const int l = itype[3 * ib], m = itype[3 * ib + 1], n = itype[3 * ib + 2];
    double xm1, xi, xp1, ym1, yi, yp1, zm1, zi, zp1;
    if (l == 0) { xm1 = 0.0; xi = 1.0; xp1 = x1; } else { xm1 = pow(x1, l - 1); xi = xm1 * x1; xp1 = xi * x1; }
    if (m == 0) { ym1 = 0.0; yi = 1.0; yp1 = y1; } else { ym1 = pow(y1, m - 1); yi = ym1 * y1; yp1 = yi * y1; }
    if (n == 0) { zm1 = 0.0; zi = 1.0; zp1 = z1; } else { zm1 = pow(z1, n - 1); zi = zm1 * z1; zp1 = zi * z1; }
    double p = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (int c = 0; c < ncontract[ib]; ++c) {
        const double al = aexp[c + XMAXPRIM * ib];
        const double t = dcoeff[c + XMAXPRIM * ib] * exp(-al * r2);
        p += t;
        gx += t * (-2.0 * al * xp1 + l * xm1);
        gy += t * (-2.0 * al * yp1 + m * ym1);
        gz += t * (-2.0 * al * zp1 + n * zm1);
    }
    *phi = p * xi * yi * zi;
    *dx = gx * yi * zi;
    *dy = gy * xi * zi;
    *dz = gz * xi * yi;
}

/* ---------------- grid ---------------- */

/* SG-1 atomic radii in bohr by atomic number (Gill, Johnson, Pople, CPL 209,
 * 506 (1993), Table 1; QUICK's RADII table) */
static double sg1_radius(int Z) {
    static const double r[19] = {0.0, 1.0, 0.5882, 3.0769, 2.0513, 1.5385, 1.2308, 1.0256, 0.8791, 0.7692,
                                 0.6838, 4.0909, 3.1579, 2.5714, 2.1687, 1.8750, 1.6514, 1.4754, 1.3333};
    return (Z >= 1 && Z <= 18) ? r[Z] : 2.25;
}

/* Stratmann-Scuseria-Frisch cell weight of point g for parent atom ip (subs/ssw.f90) */
static double ssf_weight(double gx, double gy, double gz, int ip, int natom, const double *xyz) {
    if (natom == 1) return 1.0;
    double P[1024];
    double total = 0.0, wparent = 0.0;
    for (int i = 0; i < natom; ++i) {
        const double rig = sqrt((gx - xyz[3 * i]) * (gx - xyz[3 * i]) + (gy - xyz[3 * i + 1]) * (gy - xyz[3 * i + 1]) +
                                (gz - xyz[3 * i + 2]) * (gz - xyz[3 * i + 2]));
        double w = 1.0;
        for (int j = 0; j < natom && w != 0.0; ++j) {
            if (j == i) continue;
            const double rjg = sqrt((gx - xyz[3 * j]) * (gx - xyz[3 * j]) + (gy - xyz[3 * j + 1]) * (gy - xyz[3 * j + 1]) +
                                    (gz - xyz[3 * j + 2]) * (gz - xyz[3 * j + 2]));
            const double Rij = sqrt((xyz[3 * i] - xyz[3 * j]) * (xyz[3 * i] - xyz[3 * j]) +
                                    (xyz[3 * i + 1] - xyz[3 * j + 1]) * (xyz[3 * i + 1] - xyz[3 * j + 1]) +
                                    (xyz[3 * i + 2] - xyz[3 * j + 2]) * (xyz[3 * i + 2] - xyz[3 * j + 2]));
            const double mu = (rig - rjg) / Rij;
            if (mu >= 0.64) w = 0.0;
            else if (mu >= -0.64) {
                const double f = mu / 0.64, f3 = f * f * f, f5 = f3 * f * f, f7 = f5 * f * f;
                const double g = (35.0 * f - 35.0 * f3 + 21.0 * f5 - 5.0 * f7) / 16.0;
                w *= 0.5 * (1.0 - g);
            }
        }
        if (i == ip) {
            wparent = w;
            if (w == 0.0) return 0.0;
        }
        P[i] = w;
        total += w;
    }
    (void)P;
    return wparent / total;
}

/* The SG-1 grid.  leb_n[k], leb[k] (x, y, z, w with sum w = 4 pi) are the Lebedev
 * rules of 6, 38, 86 and 194 points (k = 0..3).  Writes up to cap points
 * (x, y, z, w, parent atom) and returns the count. */
long qkx_sg1_grid(int natom, const double *xyz, const int *Z, const int *leb_n, const double *const *leb, double cut,
                  double *out, long cap) {
    static const double hpart[4] = {0.25, 0.5, 1.0, 4.5};
    static const double lpart[4] = {0.1667, 0.5, 0.9, 3.5};
    static const double npart[4] = {0.1, 0.4, 0.8, 2.5};
    const int nr = 50;
    long np = 0;
    for (int a = 0; a < natom; ++a) {
        const double *part = Z[a] <= 2 ? hpart : (Z[a] <= 10 ? lpart : npart);
        const double rad = sg1_radius(Z[a]), rad3 = rad * rad * rad;
        for (int i = 1; i <= nr; ++i) {
            const double rg = (double)i * i / ((double)(nr + 1 - i) * (nr + 1 - i));
            const double rw = 2.0 * (nr + 1) * pow((double)i, 5.0) * pow((double)(nr + 1 - i), -7.0);
            int k = rg < part[0] ? 0 : rg < part[1] ? 1 : rg < part[2] ? 2 : rg < part[3] ? 3 : 2;
            for (int j = 0; j < leb_n[k]; ++j) {
                const double *q = leb[k] + 4 * j;
                const double gx = xyz[3 * a] + rad * rg * q[0], gy = xyz[3 * a + 1] + rad * rg * q[1],
                             gz = xyz[3 * a + 2] + rad * rg * q[2];
                const double w = ssf_weight(gx, gy, gz, a, natom, xyz) * q[3] * rw * rad3;
                if (w >= cut) {
                    if (np < cap) {
                        double *o = out + 5 * np;
                        o[0] = gx; o[1] = gy; o[2] = gz; o[3] = w; o[4] = a;
                    }
                    ++np;
                }
            }
        }
    }
    return np;
}

/* ---------------- the XC loop ---------------- */

/* B3LYP XC on a grid of npts points (x, y, z, w): V (n x n, column-major,
 * full symmetric on return), E_xc, alpha and beta electron counts. */
void qkx_xc(int natom, const double *xyz, int nbasis, const int *ncenter, const int *itype, const int *ncontract,
            const double *dcoeff, const double *aexp, const double *sigrad2, const double *D, long npts,
            const double *pts, double dmcut, double *V, double *Exc, double *aelec, double *belec) {
    (void)natom;
    const int n = nbasis;
    double *phi = malloc(sizeof(double) * 4 * n);
    memset(V, 0, sizeof(double) * n * n);
    double exc = 0.0, ae = 0.0;
    for (long p = 0; p < npts; ++p) {
        const double gx = pts[5 * p], gy = pts[5 * p + 1], gz = pts[5 * p + 2], w = pts[5 * p + 3];
        if (w < dmcut) continue;
        for (int i = 0; i < n; ++i)
            pteval(gx, gy, gz, i, xyz, ncenter, itype, ncontract, dcoeff, aexp, sigrad2, phi + 4 * i, phi + 4 * i + 1,
                   phi + 4 * i + 2, phi + 4 * i + 3);
        /* denspt: alpha density and its gradient from D = Da + Db, Da = D/2 */
        double ra = 0.0, gax = 0.0, gay = 0.0, gaz = 0.0;
        for (int i = 0; i < n; ++i) {
            const double dii = D[i + (long)n * i];
            if (fabs(dii) < dmcut) continue;
            const double *pi = phi + 4 * i;
            if (fabs(pi[1] + pi[2] + pi[3] + pi[0]) < dmcut) continue;
            ra += dii * pi[0] * pi[0] / 2.0;
            gax += dii * pi[0] * pi[1];
            gay += dii * pi[0] * pi[2];
            gaz += dii * pi[0] * pi[3];
            for (int j = i + 1; j < n; ++j) {
                const double dij = D[j + (long)n * i];
                const double *pj = phi + 4 * j;
                ra += dij * pi[0] * pj[0];
                gax += dij * (pi[0] * pj[1] + pj[0] * pi[1]);
                gay += dij * (pi[0] * pj[2] + pj[0] * pi[2]);
                gaz += dij * (pi[0] * pj[3] + pj[0] * pi[3]);
            }
        }
        if (ra < dmcut) continue;
        const double rho = 2.0 * ra, sigma = 4.0 * (gax * gax + gay * gay + gaz * gaz);
        double f, vr, vs;
        qkx_b3lyp(rho, sigma, &f, &vr, &vs);
        exc += f * w;
        ae += w * ra;
        /* grad rho . grad(phi_m phi_n) weight: 2 vsigma grad rho = 4 vsigma grad rho_a */
        const double xd = 4.0 * vs * gax, yd = 4.0 * vs * gay, zd = 4.0 * vs * gaz;
        for (int i = 0; i < n; ++i) {
            const double *pi = phi + 4 * i;
            if (fabs(pi[1] + pi[2] + pi[3] + pi[0]) < dmcut) continue;
            for (int j = i; j < n; ++j) {
                const double *pj = phi + 4 * j;
                const double t = pi[0] * pj[0];
                const double tx = pi[0] * pj[1] + pj[0] * pi[1], ty = pi[0] * pj[2] + pj[0] * pi[2],
                             tz = pi[0] * pj[3] + pj[0] * pi[3];
                V[j + (long)n * i] += (t * vr + xd * tx + yd * ty + zd * tz) * w;
            }
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) V[i + (long)n * j] = V[j + (long)n * i];
    *Exc = exc;
    *aelec = ae;
    *belec = ae;
    free(phi);
}
