// qk_oneint.hip -- one-electron integrals and their nuclear-gradient terms on
// the device (gfx950).
//
// Replaces QUICK's get1e (kinetic get1eO and nuclear attraction attrashell,
// quick_one_electron_integral.f90:1633-1925) and the one-electron part of
// hfgrad (gradient.f90:40-250: nuclear repulsion, -2 W dS/dR + 2 D dT/dR, and
// attrashellopt's D dV/dR, shell.f90:1747-1860).
//
// One thread per unordered pair of basis functions (mu <= nu), looping over
// the pair's primitive pairs and the nuclei.  Each integral is built by the
// McMurchie-Davidson scheme: per Cartesian direction the Hermite expansion
// coefficients E^{ij}_t of the Gaussian product (recurrences in i and j from
// E^{00}_0 = 1), overlap and kinetic from E^{ij}_0, and the attraction of a
// nucleus as sum_tuv E^x_t E^y_u E^z_v R_tuv with the Hermite Coulomb integrals
// R_tuv from (-2p)^n F_n(p |P-C|^2).  F_n is QUICK's fit asked for the order
// Qfinal(I) + Qfinal(J) (+ 2 for the gradient), as the reference asks it.  A
// thread owns its matrix elements, so S and H are written without atomics;
// the gradient goes to a per-block int64 fixed-point copy in LDS (2^-44 units)
// and then to the global accumulator, so it is bitwise reproducible.
#include "qk_kernels.h"

namespace qk {

// QUICK's F_0..F_M (M <= 8) with the order asked for M, as boys_fit does for a
// compile-time order: the seven-region fit and upward recursion, or near T = 0
// the Taylor value at M and downward recursion (subs/fmt.f90:12-84)
__device__ void boys_fit_n(double X, int M, double* F) {
    const double E = exp(-X);
    const bool up = (X > 3.0e-5) || (X > 3.0e-7 && M < 4);
    if (up) {
        const double RS = rsqrt_nr(X);
        const double XINV = RS * RS;
        const int r = int(X > 1.0) + int(X > 3.0) + int(X > 5.0) + int(X > 10.0) + int(X > 15.0) + int(X > 33.0);
        const double z = r == 0 ? X : (r == 1 ? X - 2.0 : (r == 2 ? X - 4.0 : XINV));
        const double* c = kBoysFit + kBoysFitTerms * r;
        double v = c[0];
        for (int k = 1; k < kBoysFitTerms; ++k) v = fma(v, z, c[k]);
        const double s = 0.88622692545275801365 * RS;
        F[0] = r <= 2 ? fma(v, X + X, E) : fma(v, E, s);
        for (int m = 1; m <= M; ++m) F[m] = (double(2 * m - 1) * F[m - 1] - E) * 0.5 * XINV;
    } else {
        const double twot = X * 2.0;
        F[M] = (1.0 - X) / double(2 * M + 1);
        for (int m = M - 1; m >= 0; --m) F[m] = (twot * F[m + 1] + E) / double(m * 2 + 1);
    }
}

// E^{ij}_t, t = 0..i+j, of one direction (i, j >= 0; entries beyond i+j are 0)
__device__ __forceinline__ void herm1d(int i, int j, double PA, double PB, double oo2p, double* E) {
    double T[8];
    E[0] = 1.0;
    for (int t = 1; t < 8; ++t) E[t] = 0.0;
    int len = 1;
    for (int s = 0; s < i + j; ++s) {
        const double X = s < i ? PA : PB;
        for (int t = 0; t <= len; ++t)
            T[t] = (t > 0 ? oo2p * E[t - 1] : 0.0) + (t < len ? X * E[t] : 0.0) + (t + 1 < len ? (t + 1) * E[t + 1] : 0.0);
        ++len;
        for (int t = 0; t < len; ++t) E[t] = T[t];
    }
}

// 1D overlap factor E^{ij}_0 (0 for a negative power)
__device__ __forceinline__ double s1d(int i, int j, double PA, double PB, double oo2p) {
    if (i < 0 || j < 0) return 0.0;
    double E[8];
    herm1d(i, j, PA, PB, oo2p, E);
    return E[0];
}

// 1D kinetic factor: -1/2 d^2/dx^2 on the second function, in units of the
// 1D overlap's sqrt(pi/p): b(2j+1) S_ij - 2 b^2 S_i,j+2 - j(j-1)/2 S_i,j-2
__device__ __forceinline__ double t1d(int i, int j, double b, double PA, double PB, double oo2p) {
    if (i < 0 || j < 0) return 0.0;
    double t = b * (2 * j + 1) * s1d(i, j, PA, PB, oo2p) - 2.0 * b * b * s1d(i, j + 2, PA, PB, oo2p);
    if (j >= 2) t -= 0.5 * j * (j - 1) * s1d(i, j - 2, PA, PB, oo2p);
    return t;
}

// Hermite Coulomb integrals R^{(0)}_tuv for t <= TX, u <= TY, v <= TZ from F_n:
// R^{(n)}_000 = (-2p)^n F_n, raised one index at a time; only the entries with
// n + t + u + v <= L are formed (the contractions read no others).  R is laid
// out [n][t][u][v] with extents (L+1, TX+1, TY+1, TZ+1): at most 6 x 36 for the
// gradient of a d-d pair (L = 5, TX + TY + TZ = 7).
constexpr int kRMax = 6 * 36;
__device__ void hermite_r(int L, int TX, int TY, int TZ, double p, const double* PC, const double* F, double* R) {
    const int SY = TZ + 1, SX = (TY + 1) * SY, SN = (TX + 1) * SX;
    double pw = 1.0;
    for (int n = 0; n <= L; ++n) {
        R[n * SN] = pw * F[n];
        pw *= -2.0 * p;
    }
    for (int n = L - 1; n >= 0; --n)
        for (int t = 0; t <= TX; ++t)
            for (int u = 0; u <= TY; ++u)
                for (int v = 0; v <= TZ; ++v) {
                    if (t + u + v == 0 || n + t + u + v > L) continue;
                    const double* R1 = R + (n + 1) * SN;
                    double r;
                    if (t > 0) r = (t > 1 ? (t - 1) * R1[(t - 2) * SX + u * SY + v] : 0.0) + PC[0] * R1[(t - 1) * SX + u * SY + v];
                    else if (u > 0) r = (u > 1 ? (u - 1) * R1[t * SX + (u - 2) * SY + v] : 0.0) + PC[1] * R1[t * SX + (u - 1) * SY + v];
                    else r = (v > 1 ? (v - 1) * R1[t * SX + u * SY + v - 2] : 0.0) + PC[2] * R1[t * SX + u * SY + v - 1];
                    R[n * SN + t * SX + u * SY + v] = r;
                }
}

// sum_tuv Ex_t Ey_u Ez_v R^{(0)}_tuv
__device__ __forceinline__ double hermite_contract(const double* Ex, int nx, const double* Ey, int ny, const double* Ez,
                                                   int nz, const double* R, int TY, int TZ) {
    const int SY = TZ + 1, SX = (TY + 1) * SY;
    double s = 0.0;
    for (int t = 0; t <= nx; ++t) {
        double su = 0.0;
        for (int u = 0; u <= ny; ++u) {
            double sv = 0.0;
            for (int v = 0; v <= nz; ++v) sv += Ez[v] * R[t * SX + u * SY + v];
            su += Ey[u] * sv;
        }
        s += Ex[t] * su;
    }
    return s;
}

// unordered pair k of the triangle (mu <= nu): nu (nu + 1) / 2 + mu = k
__device__ __forceinline__ void tri_pair(long long k, int& mu, int& nu) {
    int j = (int)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while ((long long)(j + 1) * (j + 2) / 2 <= k) ++j;
    while ((long long)j * (j + 1) / 2 > k) --j;
    nu = j;
    mu = (int)(k - (long long)j * (j + 1) / 2);
}

constexpr double kPi = 3.14159265358979323846;
// primitive pairs whose Gaussian-product prefactor lies below this contribute
// less than 1e-24 to any element and are skipped
constexpr double k1eKabCut = 1e-24;

// S and H = T + V (n x n, column-major, both triangles) for every function pair
__global__ void __launch_bounds__(256) oneint_kernel(const OneFuncs f, const OneSys sys, double* __restrict__ S,
                                                     double* __restrict__ H) {
    const long long npair = (long long)sys.n * (sys.n + 1) / 2;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < npair;
         k += (long long)gridDim.x * blockDim.x) {
    int I, J;
    tri_pair(k, I, J);
    const int atA = f.atom[I], atB = f.atom[J];
    const double A[3] = {sys.xyz[3 * atA], sys.xyz[3 * atA + 1], sys.xyz[3 * atA + 2]};
    const double B[3] = {sys.xyz[3 * atB], sys.xyz[3 * atB + 1], sys.xyz[3 * atB + 2]};
    const int li[3] = {f.lmn[I] & 15, (f.lmn[I] >> 4) & 15, (f.lmn[I] >> 8) & 15};
    const int lj[3] = {f.lmn[J] & 15, (f.lmn[J] >> 4) & 15, (f.lmn[J] >> 8) & 15};
    const int L = li[0] + li[1] + li[2] + lj[0] + lj[1] + lj[2];
    const int M = max(f.qf[I] + f.qf[J], L);
    const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
    double s = 0.0, t = 0.0, v = 0.0;
    double R[kRMax], F[9], Ex[8], Ey[8], Ez[8];
    for (int ci = 0; ci < f.np[I]; ++ci)
        for (int cj = 0; cj < f.np[J]; ++cj) {
            const double a = f.a[k1eMaxPrim * I + ci], b = f.a[k1eMaxPrim * J + cj];
            const double p = a + b, oop = 1.0 / p, oo2p = 0.5 * oop;
            const double Kab = exp(-a * b * oop * AB2);
            if (Kab < k1eKabCut) continue;
            const double cc = f.c[k1eMaxPrim * I + ci] * f.c[k1eMaxPrim * J + cj];
            double P[3], PA[3], PB[3];
            for (int d = 0; d < 3; ++d) {
                P[d] = (a * A[d] + b * B[d]) * oop;
                PA[d] = P[d] - A[d];
                PB[d] = P[d] - B[d];
            }
            const double sx = s1d(li[0], lj[0], PA[0], PB[0], oo2p), sy = s1d(li[1], lj[1], PA[1], PB[1], oo2p),
                         sz = s1d(li[2], lj[2], PA[2], PB[2], oo2p);
            const double tx = t1d(li[0], lj[0], b, PA[0], PB[0], oo2p), ty = t1d(li[1], lj[1], b, PA[1], PB[1], oo2p),
                         tz = t1d(li[2], lj[2], b, PA[2], PB[2], oo2p);
            const double s0 = Kab * kPi * oop * sqrt(kPi * oop);   // (pi/p)^{3/2} K_ab
            s += cc * s0 * sx * sy * sz;
            t += cc * s0 * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
            herm1d(li[0], lj[0], PA[0], PB[0], oo2p, Ex);
            herm1d(li[1], lj[1], PA[1], PB[1], oo2p, Ey);
            herm1d(li[2], lj[2], PA[2], PB[2], oo2p, Ez);
            const int TX = li[0] + lj[0], TY = li[1] + lj[1], TZ = li[2] + lj[2];
            const double v0 = 2.0 * kPi * oop * Kab;
            for (int C = 0; C < sys.natom; ++C) {
                const double PC[3] = {P[0] - sys.xyz[3 * C], P[1] - sys.xyz[3 * C + 1], P[2] - sys.xyz[3 * C + 2]};
                boys_fit_n(p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), M, F);
                hermite_r(L, TX, TY, TZ, p, PC, F, R);
                v -= cc * sys.chg[C] * v0 * hermite_contract(Ex, TX, Ey, TY, Ez, TZ, R, TY, TZ);
            }
        }
    const long long o1 = (long long)I * sys.n + J, o2 = (long long)J * sys.n + I;
    if (S) S[o1] = S[o2] = s;
    if (H) H[o1] = H[o2] = t + v;
    }
}

__device__ __forceinline__ void lds_gadd(unsigned long long* g, int i, double v) {
    const long long iv = __double2ll_rn(v * 0x1p44);
    if (iv) atomicAdd(g + i, (unsigned long long)iv);
}

// d/dA of one integral over a primitive pair: 2a g(a + 1x) - a_x g(a - 1x)
// along direction x, for overlap / kinetic factors (1D products)
struct Prim1e {
    double a, b, oo2p, PA[3], PB[3];
};

__device__ __forceinline__ double st_prod(const Prim1e& q, const int* li, const int* lj, bool kin) {
    double sd[3], td[3];
    for (int d = 0; d < 3; ++d) {
        sd[d] = s1d(li[d], lj[d], q.PA[d], q.PB[d], q.oo2p);
        td[d] = kin ? t1d(li[d], lj[d], q.b, q.PA[d], q.PB[d], q.oo2p) : 0.0;
    }
    return kin ? td[0] * sd[1] * sd[2] + sd[0] * td[1] * sd[2] + sd[0] * sd[1] * td[2] : sd[0] * sd[1] * sd[2];
}

// grad += sum over function pairs of D dH/dR - W dS/dR (3 x natom, 2^-44 units,
// accumulated per block in LDS: dynamic LDS of 3 natom int64)
__global__ void __launch_bounds__(256) oneint_grad_kernel(const OneFuncs f, const OneSys sys, const double* __restrict__ Dm,
                                                          const double* __restrict__ Wm,
                                                          unsigned long long* __restrict__ G) {
    extern __shared__ unsigned long long s_g[];
    const int m3 = 3 * sys.natom;
    for (int i = threadIdx.x; i < m3; i += blockDim.x) s_g[i] = 0ull;
    __syncthreads();
    const long long npair = (long long)sys.n * (sys.n + 1) / 2;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < npair;
         k += (long long)gridDim.x * blockDim.x) {
        int I, J;
        tri_pair(k, I, J);
        const int atA = f.atom[I], atB = f.atom[J];
        const double A[3] = {sys.xyz[3 * atA], sys.xyz[3 * atA + 1], sys.xyz[3 * atA + 2]};
        const double B[3] = {sys.xyz[3 * atB], sys.xyz[3 * atB + 1], sys.xyz[3 * atB + 2]};
        const int li[3] = {f.lmn[I] & 15, (f.lmn[I] >> 4) & 15, (f.lmn[I] >> 8) & 15};
        const int lj[3] = {f.lmn[J] & 15, (f.lmn[J] >> 4) & 15, (f.lmn[J] >> 8) & 15};
        const int L = li[0] + li[1] + li[2] + lj[0] + lj[1] + lj[2] + 1;
        const int M = max(f.qf[I] + f.qf[J] + 2, L);
        const double wt = I == J ? 1.0 : 2.0;   // (I, J) and (J, I)
        const double dIJ = wt * Dm[(long long)I * sys.n + J], wIJ = wt * Wm[(long long)I * sys.n + J];
        const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
        double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
        double R[kRMax], F[10];
        for (int ci = 0; ci < f.np[I]; ++ci)
            for (int cj = 0; cj < f.np[J]; ++cj) {
                const double a = f.a[k1eMaxPrim * I + ci], b = f.a[k1eMaxPrim * J + cj];
                const double p = a + b, oop = 1.0 / p;
                const double Kab = exp(-a * b * oop * AB2);
                if (Kab < k1eKabCut) continue;
                const double cc = f.c[k1eMaxPrim * I + ci] * f.c[k1eMaxPrim * J + cj];
                Prim1e q;
                q.a = a; q.b = b; q.oo2p = 0.5 * oop;
                double P[3];
                for (int d = 0; d < 3; ++d) {
                    P[d] = (a * A[d] + b * B[d]) * oop;
                    q.PA[d] = P[d] - A[d];
                    q.PB[d] = P[d] - B[d];
                }
                const double s0 = cc * Kab * kPi * oop * sqrt(kPi * oop);
                // overlap and kinetic: two-centre, d/dB = -d/dA; zero on one atom
                if (atA != atB)
                    for (int x = 0; x < 3; ++x) {
                        int lp[3] = {li[0], li[1], li[2]}, lm[3] = {li[0], li[1], li[2]};
                        lp[x] += 1;
                        lm[x] -= 1;
                        const double dS = 2.0 * a * st_prod(q, lp, lj, false) - li[x] * st_prod(q, lm, lj, false);
                        const double dT = 2.0 * a * st_prod(q, lp, lj, true) - li[x] * st_prod(q, lm, lj, true);
                        const double c = s0 * (dIJ * dT - wIJ * dS);
                        gA[x] += c;
                        gB[x] -= c;
                    }
                // attraction: one raised/lowered index on A or on B per direction
                double E0[3][8], Eap[3][8], Eam[3][8], Ebp[3][8], Ebm[3][8];
                for (int d = 0; d < 3; ++d) {
                    herm1d(li[d], lj[d], q.PA[d], q.PB[d], q.oo2p, E0[d]);
                    herm1d(li[d] + 1, lj[d], q.PA[d], q.PB[d], q.oo2p, Eap[d]);
                    herm1d(li[d], lj[d] + 1, q.PA[d], q.PB[d], q.oo2p, Ebp[d]);
                    if (li[d] > 0) herm1d(li[d] - 1, lj[d], q.PA[d], q.PB[d], q.oo2p, Eam[d]);
                    else for (int t = 0; t < 8; ++t) Eam[d][t] = 0.0;
                    if (lj[d] > 0) herm1d(li[d], lj[d] - 1, q.PA[d], q.PB[d], q.oo2p, Ebm[d]);
                    else for (int t = 0; t < 8; ++t) Ebm[d][t] = 0.0;
                }
                const int TX = li[0] + lj[0] + 1, TY = li[1] + lj[1] + 1, TZ = li[2] + lj[2] + 1;
                const double v0 = cc * dIJ * 2.0 * kPi * oop * Kab;
                for (int C = 0; C < sys.natom; ++C) {
                    if (atA == C && atB == C) continue;
                    const double PC[3] = {P[0] - sys.xyz[3 * C], P[1] - sys.xyz[3 * C + 1], P[2] - sys.xyz[3 * C + 2]};
                    boys_fit_n(p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), M, F);
                    hermite_r(L, TX, TY, TZ, p, PC, F, R);
                    const double z = -sys.chg[C] * v0;
                    double gc[3];
                    for (int x = 0; x < 3; ++x) {
                        const double* ex[3] = {E0[0], E0[1], E0[2]};
                        int nx[3] = {li[0] + lj[0], li[1] + lj[1], li[2] + lj[2]};
                        ex[x] = Eap[x];
                        nx[x] += 1;
                        double dA = 2.0 * a * hermite_contract(ex[0], nx[0], ex[1], nx[1], ex[2], nx[2], R, TY, TZ);
                        if (li[x] > 0) {
                            ex[x] = Eam[x];
                            nx[x] -= 2;
                            dA -= li[x] * hermite_contract(ex[0], nx[0], ex[1], nx[1], ex[2], nx[2], R, TY, TZ);
                            nx[x] += 2;
                        }
                        ex[x] = Ebp[x];
                        double dB = 2.0 * b * hermite_contract(ex[0], nx[0], ex[1], nx[1], ex[2], nx[2], R, TY, TZ);
                        if (lj[x] > 0) {
                            ex[x] = Ebm[x];
                            nx[x] -= 2;
                            dB -= lj[x] * hermite_contract(ex[0], nx[0], ex[1], nx[1], ex[2], nx[2], R, TY, TZ);
                        }
                        gA[x] += z * dA;
                        gB[x] += z * dB;
                        gc[x] = -z * (dA + dB);
                    }
                    for (int x = 0; x < 3; ++x) lds_gadd(s_g, 3 * C + x, gc[x]);
                }
            }
        for (int x = 0; x < 3; ++x) {
            lds_gadd(s_g, 3 * atA + x, gA[x]);
            lds_gadd(s_g, 3 * atB + x, gB[x]);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m3; i += blockDim.x)
        if (s_g[i]) atomicAdd(G + i, s_g[i]);
}

// nuclear repulsion gradient: one thread per atom (2^-44 units)
__global__ void nuclear_grad_kernel(const OneSys sys, unsigned long long* __restrict__ G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sys.natom) return;
    double g[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < sys.natom; ++j) {
        if (j == i) continue;
        double d[3], r2 = 0.0;
        for (int x = 0; x < 3; ++x) {
            d[x] = sys.xyz[3 * j + x] - sys.xyz[3 * i + x];
            r2 += d[x] * d[x];
        }
        const double c = sys.chg[i] * sys.chg[j] / (r2 * sqrt(r2));
        for (int x = 0; x < 3; ++x) g[x] += d[x] * c;
    }
    for (int x = 0; x < 3; ++x) {
        const long long iv = __double2ll_rn(g[x] * 0x1p44);
        if (iv) atomicAdd(G + 3 * i + x, (unsigned long long)iv);
    }
}

// grids bounded by the kernels' scratch backing (their per-thread arrays spill);
// the grid-stride loops cover every pair
void launch_oneint(const OneFuncs& f, const OneSys& sys, double* S, double* H, hipStream_t s) {
    const long long npair = (long long)sys.n * (sys.n + 1) / 2;
    static const long long cap = scratch_grid_cap(kernel_scratch(reinterpret_cast<const void*>(&oneint_kernel)), 4);
    const long long grid = std::min((npair + 255) / 256, std::min(cap, 256LL * 16));
    hipLaunchKernelGGL(oneint_kernel, dim3((unsigned)std::max(grid, 1LL)), dim3(256), 0, s, f, sys, S, H);
}

void launch_oneint_grad(const OneFuncs& f, const OneSys& sys, const double* D, const double* W,
                        unsigned long long* G, hipStream_t s) {
    const long long npair = (long long)sys.n * (sys.n + 1) / 2;
    const size_t lds = sizeof(unsigned long long) * 3 * (size_t)sys.natom;
    static const long long cap =
        scratch_grid_cap(kernel_scratch(reinterpret_cast<const void*>(&oneint_grad_kernel)), 4);
    const long long grid = std::min((npair + 255) / 256, std::min(cap, 256LL * 16));
    hipLaunchKernelGGL(oneint_grad_kernel, dim3((unsigned)std::max(grid, 1LL)), dim3(256), lds, s, f, sys, D, W, G);
    hipLaunchKernelGGL(nuclear_grad_kernel, dim3((sys.natom + 255) / 256), dim3(256), 0, s, sys, G);
}

}  // namespace qk
