// qk_oneint.hip -- one-electron integrals and their nuclear-gradient terms on
// the device (gfx950).
//
// Replaces QUICK's get1e (kinetic get1eO and nuclear attraction attrashell,
// quick_one_electron_integral.f90:1633-1925) and the one-electron part of
// hfgrad (gradient.f90:40-250: nuclear repulsion, -2 W dS/dR + 2 D dT/dR, and
// attrashellopt's D dV/dR, shell.f90:1747-1860).
//
// One thread per unordered pair of basis functions (mu <= nu), looping over
// the pair's primitive pairs and the nuclei; one kernel per total angular
// momentum of the pair (the host lists the pairs of each), so that every
// table below has a compile-time size and compile-time indices and stays in
// registers (a table indexed at run time would live in per-lane scratch
// memory).  Each integral is built by the
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

// QUICK's F_0..F_MMAX with the order M asked for (L <= M <= MMAX): the
// seven-region fit and upward recursion, or near T = 0 the Taylor value at M
// and downward recursion (subs/fmt.f90:12-84).  Every loop runs to the
// compile-time bound with the order as a predicate, so F stays in registers;
// entries above M are not read.
template <int MMAX>
__device__ __forceinline__ void boys_fit_u(double X, int M, double* F) {
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
#pragma unroll
        for (int m = 1; m <= MMAX; ++m) F[m] = (double(2 * m - 1) * F[m - 1] - E) * 0.5 * XINV;
    } else {
        const double twot = X * 2.0;
        double cur = 0.0;
#pragma unroll
        for (int m = MMAX; m >= 0; --m) {
            if (m == M) cur = (1.0 - X) / double(2 * M + 1);
            else if (m < M) cur = (twot * cur + E) / double(m * 2 + 1);
            F[m] = cur;
        }
    }
}

// E^{ij}_t, t = 0..S, of one direction for i + j <= S (entries beyond i + j are
// 0): the recurrences in i and j from E^{00}_0 = 1, unrolled to S steps with
// the step count i + j as a predicate
template <int S>
__device__ __forceinline__ void herm_u(int i, int j, double PA, double PB, double oo2p, double* E) {
    E[0] = 1.0;
#pragma unroll
    for (int t = 1; t <= S; ++t) E[t] = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (s < i + j) {
            const double X = s < i ? PA : PB;
            double T[S + 1];
#pragma unroll
            for (int t = 0; t <= s + 1; ++t)
                T[t] = (t > 0 ? oo2p * E[t - 1] : 0.0) + (t <= s ? X * E[t] : 0.0) +
                       (t + 1 <= s ? (t + 1) * E[t + 1] : 0.0);
#pragma unroll
            for (int t = 0; t <= s + 1; ++t) E[t] = T[t];
        }
    }
}

// 1D overlap factor E^{ij}_0 (0 for a negative power), i + j <= S
template <int S>
__device__ __forceinline__ double s1d_u(int i, int j, double PA, double PB, double oo2p) {
    if (i < 0 || j < 0) return 0.0;
    double E[S + 1];
    herm_u<S>(i, j, PA, PB, oo2p, E);
    return E[0];
}

// 1D kinetic factor: -1/2 d^2/dx^2 on the second function, in units of the
// 1D overlap's sqrt(pi/p): b(2j+1) S_ij - 2 b^2 S_i,j+2 - j(j-1)/2 S_i,j-2
// (i + j + 2 <= S)
template <int S>
__device__ __forceinline__ double t1d_u(int i, int j, double b, double PA, double PB, double oo2p) {
    if (i < 0 || j < 0) return 0.0;
    double t = b * (2 * j + 1) * s1d_u<S>(i, j, PA, PB, oo2p) - 2.0 * b * b * s1d_u<S>(i, j + 2, PA, PB, oo2p);
    if (j >= 2) t -= 0.5 * j * (j - 1) * s1d_u<S>(i, j - 2, PA, PB, oo2p);
    return t;
}

// (t, u, v) with t + u + v <= L in one tetrahedral array: the shells of equal
// t + u + v one after another, within a shell by u + v, then v
constexpr int tet_size(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
constexpr int tet(int t, int u, int v) {
    return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}

// Hermite Coulomb integrals R^{(0)}_tuv for t + u + v <= L: R^{(n)}_000 =
// (-2p)^n F_n, raised one index at a time (t first, then u, then v), level n
// holding t + u + v <= L - n; all indices compile-time, so the table stays in
// registers
template <int L>
__device__ __forceinline__ void hermite_r_u(double p, const double* PC, const double* F, double* R0) {
    double R[L + 1][tet_size(L)];
    double pw = 1.0;
#pragma unroll
    for (int n = 0; n <= L; ++n) {
        R[n][0] = pw * F[n];
        pw *= -2.0 * p;
    }
#pragma unroll
    for (int n = L - 1; n >= 0; --n)
#pragma unroll
        for (int s = 1; s <= L - n; ++s)
#pragma unroll
            for (int r = 0; r <= s; ++r)
#pragma unroll
                for (int v = 0; v <= r; ++v) {
                    const int u = r - v, t = s - r;
                    double x;
                    if (t > 0)
                        x = (t > 1 ? (t - 1) * R[n + 1][tet(t - 2, u, v)] : 0.0) + PC[0] * R[n + 1][tet(t - 1, u, v)];
                    else if (u > 0)
                        x = (u > 1 ? (u - 1) * R[n + 1][tet(t, u - 2, v)] : 0.0) + PC[1] * R[n + 1][tet(t, u - 1, v)];
                    else
                        x = (v > 1 ? (v - 1) * R[n + 1][tet(t, u, v - 2)] : 0.0) + PC[2] * R[n + 1][tet(t, u, v - 1)];
                    R[n][tet(t, u, v)] = x;
                }
#pragma unroll
    for (int k = 0; k < tet_size(L); ++k) R0[k] = R[0][k];
}

// sum_tuv Ex_t Ey_u Ez_v R^{(0)}_tuv over t + u + v <= L (E zero beyond each
// direction's i + j, so the sum is that over the box)
template <int L>
__device__ __forceinline__ double contract_u(const double* Ex, const double* Ey, const double* Ez, const double* R0) {
    double s = 0.0;
#pragma unroll
    for (int t = 0; t <= L; ++t) {
        double su = 0.0;
#pragma unroll
        for (int u = 0; u <= L - t; ++u) {
            double sv = 0.0;
#pragma unroll
            for (int v = 0; v <= L - t - u; ++v) sv += Ez[v] * R0[tet(t, u, v)];
            su += Ey[u] * sv;
        }
        s += Ex[t] * su;
    }
    return s;
}

constexpr double kPi = 3.14159265358979323846;
// primitive pairs whose Gaussian-product prefactor lies below this contribute
// less than 1e-24 to any element and are skipped
constexpr double k1eKabCut = 1e-24;
// Boys orders asked: Qfinal(I) + Qfinal(J) <= 4 (d shells), + 2 for the gradient
constexpr int kBoys1eMax = 4;

__device__ __forceinline__ void unpack_lmn(int v, int* l) {
    l[0] = v & 15;
    l[1] = (v >> 4) & 15;
    l[2] = (v >> 8) & 15;
}

// S and H = T + V (n x n, column-major, both triangles) for the function pairs
// (mu <= nu) of total angular momentum LT
template <int LT>
__global__ void __launch_bounds__(256) oneint_kernel(const OneFuncs f, const OneSys sys, const int2* __restrict__ pairs,
                                                     long long npair, double* __restrict__ S, double* __restrict__ H) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < npair;
         k += (long long)gridDim.x * blockDim.x) {
        const int I = pairs[k].x, J = pairs[k].y;
        const int atA = f.atom[I], atB = f.atom[J];
        const double A[3] = {sys.xyz[3 * atA], sys.xyz[3 * atA + 1], sys.xyz[3 * atA + 2]};
        const double B[3] = {sys.xyz[3 * atB], sys.xyz[3 * atB + 1], sys.xyz[3 * atB + 2]};
        int li[3], lj[3];
        unpack_lmn(f.lmn[I], li);
        unpack_lmn(f.lmn[J], lj);
        const int M = max(f.qf[I] + f.qf[J], LT);
        const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
        double s = 0.0, t = 0.0, v = 0.0;
        for (int ci = 0; ci < f.np[I]; ++ci)
            for (int cj = 0; cj < f.np[J]; ++cj) {
                const double a = f.a[k1eMaxPrim * I + ci], b = f.a[k1eMaxPrim * J + cj];
                const double p = a + b, oop = 1.0 / p, oo2p = 0.5 * oop;
                const double Kab = exp(-a * b * oop * AB2);
                if (Kab < k1eKabCut) continue;
                const double cc = f.c[k1eMaxPrim * I + ci] * f.c[k1eMaxPrim * J + cj];
                double P[3], PA[3], PB[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    P[d] = (a * A[d] + b * B[d]) * oop;
                    PA[d] = P[d] - A[d];
                    PB[d] = P[d] - B[d];
                }
                double E[3][LT + 1];
#pragma unroll
                for (int d = 0; d < 3; ++d) herm_u<LT>(li[d], lj[d], PA[d], PB[d], oo2p, E[d]);
                const double sx = E[0][0], sy = E[1][0], sz = E[2][0];
                const double tx = t1d_u<LT + 2>(li[0], lj[0], b, PA[0], PB[0], oo2p),
                             ty = t1d_u<LT + 2>(li[1], lj[1], b, PA[1], PB[1], oo2p),
                             tz = t1d_u<LT + 2>(li[2], lj[2], b, PA[2], PB[2], oo2p);
                const double s0 = Kab * kPi * oop * sqrt(kPi * oop);   // (pi/p)^{3/2} K_ab
                s += cc * s0 * sx * sy * sz;
                t += cc * s0 * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
                const double v0 = 2.0 * kPi * oop * Kab;
                for (int C = 0; C < sys.natom; ++C) {
                    const double PC[3] = {P[0] - sys.xyz[3 * C], P[1] - sys.xyz[3 * C + 1], P[2] - sys.xyz[3 * C + 2]};
                    double F[kBoys1eMax + 1], R0[tet_size(LT)];
                    boys_fit_u<kBoys1eMax>(p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), M, F);
                    hermite_r_u<LT>(p, PC, F, R0);
                    v -= cc * sys.chg[C] * v0 * contract_u<LT>(E[0], E[1], E[2], R0);
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

template <int S>
__device__ __forceinline__ double st_prod(const Prim1e& q, const int* li, const int* lj, bool kin) {
    double sd[3], td[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        sd[d] = s1d_u<S>(li[d], lj[d], q.PA[d], q.PB[d], q.oo2p);
        td[d] = kin ? t1d_u<S>(li[d], lj[d], q.b, q.PA[d], q.PB[d], q.oo2p) : 0.0;
    }
    return kin ? td[0] * sd[1] * sd[2] + sd[0] * td[1] * sd[2] + sd[0] * sd[1] * td[2] : sd[0] * sd[1] * sd[2];
}

// grad += sum over the function pairs of total angular momentum LT of
// D dH/dR - W dS/dR (3 x natom, 2^-44 units, accumulated per block in LDS:
// dynamic LDS of 3 natom int64).  The derivative integrals raise or lower one
// index, so the Hermite tables run to LG = LT + 1.
template <int LT>
__global__ void __launch_bounds__(256) oneint_grad_kernel(const OneFuncs f, const OneSys sys, const int2* __restrict__ pairs,
                                                          long long npair, const double* __restrict__ Dm,
                                                          const double* __restrict__ Wm,
                                                          unsigned long long* __restrict__ G) {
    constexpr int LG = LT + 1;
    extern __shared__ unsigned long long s_g[];
    const int m3 = 3 * sys.natom;
    for (int i = threadIdx.x; i < m3; i += blockDim.x) s_g[i] = 0ull;
    __syncthreads();
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < npair;
         k += (long long)gridDim.x * blockDim.x) {
        const int I = pairs[k].x, J = pairs[k].y;
        const int atA = f.atom[I], atB = f.atom[J];
        const double A[3] = {sys.xyz[3 * atA], sys.xyz[3 * atA + 1], sys.xyz[3 * atA + 2]};
        const double B[3] = {sys.xyz[3 * atB], sys.xyz[3 * atB + 1], sys.xyz[3 * atB + 2]};
        int li[3], lj[3];
        unpack_lmn(f.lmn[I], li);
        unpack_lmn(f.lmn[J], lj);
        const int M = max(f.qf[I] + f.qf[J] + 2, LG);
        const double wt = I == J ? 1.0 : 2.0;   // (I, J) and (J, I)
        const double dIJ = wt * Dm[(long long)I * sys.n + J], wIJ = wt * Wm[(long long)I * sys.n + J];
        const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
        double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
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
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    P[d] = (a * A[d] + b * B[d]) * oop;
                    q.PA[d] = P[d] - A[d];
                    q.PB[d] = P[d] - B[d];
                }
                const double s0 = cc * Kab * kPi * oop * sqrt(kPi * oop);
                // overlap and kinetic: two-centre, d/dB = -d/dA; zero on one atom
                if (atA != atB)
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        int lp[3] = {li[0], li[1], li[2]}, lm[3] = {li[0], li[1], li[2]};
                        lp[x] += 1;
                        lm[x] -= 1;
                        const double dS = 2.0 * a * st_prod<LG + 2>(q, lp, lj, false) -
                                          li[x] * st_prod<LG + 2>(q, lm, lj, false);
                        const double dT = 2.0 * a * st_prod<LG + 2>(q, lp, lj, true) -
                                          li[x] * st_prod<LG + 2>(q, lm, lj, true);
                        const double c = s0 * (dIJ * dT - wIJ * dS);
                        gA[x] += c;
                        gB[x] -= c;
                    }
                // attraction: one raised/lowered index on A or on B per direction
                double E0[3][LG + 1], Eap[3][LG + 1], Eam[3][LG + 1], Ebp[3][LG + 1], Ebm[3][LG + 1];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    herm_u<LG>(li[d], lj[d], q.PA[d], q.PB[d], q.oo2p, E0[d]);
                    herm_u<LG>(li[d] + 1, lj[d], q.PA[d], q.PB[d], q.oo2p, Eap[d]);
                    herm_u<LG>(li[d], lj[d] + 1, q.PA[d], q.PB[d], q.oo2p, Ebp[d]);
                    if (li[d] > 0) herm_u<LG>(li[d] - 1, lj[d], q.PA[d], q.PB[d], q.oo2p, Eam[d]);
                    else
#pragma unroll
                        for (int t = 0; t <= LG; ++t) Eam[d][t] = 0.0;
                    if (lj[d] > 0) herm_u<LG>(li[d], lj[d] - 1, q.PA[d], q.PB[d], q.oo2p, Ebm[d]);
                    else
#pragma unroll
                        for (int t = 0; t <= LG; ++t) Ebm[d][t] = 0.0;
                }
                const double v0 = cc * dIJ * 2.0 * kPi * oop * Kab;
                for (int C = 0; C < sys.natom; ++C) {
                    if (atA == C && atB == C) continue;
                    const double PC[3] = {P[0] - sys.xyz[3 * C], P[1] - sys.xyz[3 * C + 1], P[2] - sys.xyz[3 * C + 2]};
                    double F[kBoys1eMax + 3], R0[tet_size(LG)];
                    boys_fit_u<kBoys1eMax + 2>(p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), M, F);
                    hermite_r_u<LG>(p, PC, F, R0);
                    const double z = -sys.chg[C] * v0;
                    double gc[3];
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        const double* e0 = x == 0 ? Eap[0] : E0[0];
                        const double* e1 = x == 1 ? Eap[1] : E0[1];
                        const double* e2 = x == 2 ? Eap[2] : E0[2];
                        double dA = 2.0 * a * contract_u<LG>(e0, e1, e2, R0);
                        if (li[x] > 0) {
                            dA -= li[x] * contract_u<LG>(x == 0 ? Eam[0] : E0[0], x == 1 ? Eam[1] : E0[1],
                                                         x == 2 ? Eam[2] : E0[2], R0);
                        }
                        double dB = 2.0 * b * contract_u<LG>(x == 0 ? Ebp[0] : E0[0], x == 1 ? Ebp[1] : E0[1],
                                                             x == 2 ? Ebp[2] : E0[2], R0);
                        if (lj[x] > 0) {
                            dB -= lj[x] * contract_u<LG>(x == 0 ? Ebm[0] : E0[0], x == 1 ? Ebm[1] : E0[1],
                                                         x == 2 ? Ebm[2] : E0[2], R0);
                        }
                        gA[x] += z * dA;
                        gB[x] += z * dB;
                        gc[x] = -z * (dA + dB);
                    }
#pragma unroll
                    for (int x = 0; x < 3; ++x) lds_gadd(s_g, 3 * C + x, gc[x]);
                }
            }
#pragma unroll
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

// one launch per total angular momentum of the function pairs (lists built
// on the host once per system: pairs[L] holds the pairs mu <= nu with
// l(mu) + l(nu) = L); no kernel here spills, so the grids are not bounded by a
// scratch backing
template <int LT>
static void launch_oneint_l(const OneFuncs& f, const OneSys& sys, const int2* pairs, long long np, double* S,
                            double* H, hipStream_t s) {
    if (np <= 0) return;
    const long long grid = std::min((np + 255) / 256, 256LL * 16);
    hipLaunchKernelGGL(oneint_kernel<LT>, dim3((unsigned)grid), dim3(256), 0, s, f, sys, pairs, np, S, H);
}

void launch_oneint(const OneFuncs& f, const OneSys& sys, const int2* const* pairs, const long long* npairs, double* S,
                   double* H, hipStream_t s) {
    launch_oneint_l<0>(f, sys, pairs[0], npairs[0], S, H, s);
    launch_oneint_l<1>(f, sys, pairs[1], npairs[1], S, H, s);
    launch_oneint_l<2>(f, sys, pairs[2], npairs[2], S, H, s);
    launch_oneint_l<3>(f, sys, pairs[3], npairs[3], S, H, s);
    launch_oneint_l<4>(f, sys, pairs[4], npairs[4], S, H, s);
}

template <int LT>
static void launch_oneint_grad_l(const OneFuncs& f, const OneSys& sys, const int2* pairs, long long np,
                                 const double* D, const double* W, unsigned long long* G, size_t lds, hipStream_t s) {
    if (np <= 0) return;
    const long long grid = std::min((np + 255) / 256, 256LL * 16);
    hipLaunchKernelGGL(oneint_grad_kernel<LT>, dim3((unsigned)grid), dim3(256), lds, s, f, sys, pairs, np, D, W, G);
}

void launch_oneint_grad(const OneFuncs& f, const OneSys& sys, const int2* const* pairs, const long long* npairs,
                        const double* D, const double* W, unsigned long long* G, hipStream_t s) {
    const size_t lds = sizeof(unsigned long long) * 3 * (size_t)sys.natom;
    launch_oneint_grad_l<0>(f, sys, pairs[0], npairs[0], D, W, G, lds, s);
    launch_oneint_grad_l<1>(f, sys, pairs[1], npairs[1], D, W, G, lds, s);
    launch_oneint_grad_l<2>(f, sys, pairs[2], npairs[2], D, W, G, lds, s);
    launch_oneint_grad_l<3>(f, sys, pairs[3], npairs[3], D, W, G, lds, s);
    launch_oneint_grad_l<4>(f, sys, pairs[4], npairs[4], D, W, G, lds, s);
    hipLaunchKernelGGL(nuclear_grad_kernel, dim3((sys.natom + 255) / 256), dim3(256), 0, s, sys, G);
}

}  // namespace qk
