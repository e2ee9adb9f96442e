// qk_kernels.h -- device code shared by the per-class translation units.
#pragma once
#include "qk_internal.h"
#include "eri_generated.inc"

namespace qk {

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nrange(int lo, int hi) {
    return hi < lo ? 0 : ncart(hi) + nrange(lo, hi - 1);
}

// QUICK's FmT (subs/fmt.f90:12-84; gpu_get2e_subs.h:1386-1457) restated: the
// piecewise fit of F_0 (or F_Mb near T = 0), then upward recursion for
// T > 3e-5 (T > 3e-7 when Mb < 4), else downward recursion from Mb.  Mb is the
// order the reference asks for; values F_0..F_LT are produced (LT <= Mb).
template <int LT>
__device__ __forceinline__ void boys_ref(double X, int Mb, double* F) {
    const double PIE4 = 3.1415926535897932384626433832795 / 4.0;
    const double E = exp(-X);
    const bool up = (X > 3.0e-5) || (X > 3.0e-7 && Mb < 4);
    if (up) {
        const double XINV = 1.0 / X;
        double WW1;
        if (X > 5.0) {
            const double s = sqrt(PIE4 * XINV);
            if (X > 33.0) {
                WW1 = s;
            } else if (X > 15.0) {
                WW1 = ((1.9623264149430e-01 * XINV - 4.9695241464490e-01) * XINV - 6.0156581186481e-05) * E + s;
            } else if (X > 10.0) {
                WW1 = (((-1.8784686463512e-01 * XINV + 2.2991849164985e-01) * XINV - 4.9893752514047e-01) * XINV -
                       2.1916512131607e-05) * E + s;
            } else {
                WW1 = ((((((4.6897511375022e-01 * XINV - 6.9955602298985e-01) * XINV + 5.3689283271887e-01) * XINV -
                          3.2883030418398e-01) * XINV + 2.4645596956002e-01) * XINV - 4.9984072848436e-01) * XINV -
                       3.1501078774085e-06) * E + s;
            }
        } else if (X > 1.0) {
            double F1;
            if (X > 3.0) {
                const double Y = X - 4.0;
                F1 = ((((((((((-2.62453564772299e-11 * Y + 3.24031041623823e-10) * Y - 3.614965656163e-09) * Y +
                             3.760256799971e-08) * Y - 3.553558319675e-07) * Y + 3.022556449731e-06) * Y -
                          2.290098979647e-05) * Y + 1.526537461148e-04) * Y - 8.81947375894379e-04) * Y +
                       4.33207949514611e-03) * Y - 1.75257821619926e-02) * Y + 5.28406320615584e-02;
            } else {
                const double Y = X - 2.0;
                F1 = ((((((((((-1.61702782425558e-10 * Y + 1.96215250865776e-09) * Y - 2.14234468198419e-08) * Y +
                             2.17216556336318e-07) * Y - 1.98850171329371e-06) * Y + 1.62429321438911e-05) * Y -
                          1.16740298039895e-04) * Y + 7.24888732052332e-04) * Y - 3.79490003707156e-03) * Y +
                       1.61723488664661e-02) * Y - 5.29428148329736e-02) * Y + 1.15702180856167e-01;
            }
            WW1 = (X + X) * F1 + E;
        } else {
            const double F1 = ((((((((-8.36313918003957e-08 * X + 1.21222603512827e-06) * X - 1.15662609053481e-05) * X +
                                    9.25197374512647e-05) * X - 6.40994113129432e-04) * X + 3.78787044215009e-03) * X -
                                 1.85185172458485e-02) * X + 7.14285713298222e-02) * X - 1.99999999997023e-01) * X +
                              3.33333333333318e-01;
            WW1 = (X + X) * F1 + E;
        }
        F[0] = WW1;
#pragma unroll
        for (int m = 1; m <= LT; ++m) F[m] = (double(2 * m - 1) * F[m - 1] - E) * 0.5 * XINV;
    } else {
        double v = (1.0 - X) / double(2 * Mb + 1);
        const double twot = X * 2.0;
        for (int m = Mb - 1; m >= LT; --m) v = (twot * v + E) / double(m * 2 + 1);
        F[LT] = v;
#pragma unroll
        for (int m = LT - 1; m >= 0; --m) F[m] = (twot * F[m + 1] + E) / double(m * 2 + 1);
    }
}

// The same fit without branches on T: the seven regions' polynomials, zero-padded
// to 12 coefficients (highest power first), are evaluated by one Horner chain
// over coefficients read from a per-block copy of kBoysFit; the padding makes
// every region's chain identical to its own shorter one (fma(0, z, c) = c), and
// F_0 = fma(v, u, w) with (z, u, w) = (X, 2X, E), (X-2, 2X, E), (X-4, 2X, E) below
// T = 5 and (1/X, E, sqrt(pi/4X)) above.  Lanes of a wavefront with different T
// then execute one path instead of one per region.
constexpr int kBoysFitRegions = 7, kBoysFitTerms = 12;
__constant__ const double kBoysFit[kBoysFitRegions * kBoysFitTerms] = {
    0, 0, -8.36313918003957e-08, 1.21222603512827e-06, -1.15662609053481e-05, 9.25197374512647e-05,
    -6.40994113129432e-04, 3.78787044215009e-03, -1.85185172458485e-02, 7.14285713298222e-02,
    -1.99999999997023e-01, 3.33333333333318e-01,
    -1.61702782425558e-10, 1.96215250865776e-09, -2.14234468198419e-08, 2.17216556336318e-07,
    -1.98850171329371e-06, 1.62429321438911e-05, -1.16740298039895e-04, 7.24888732052332e-04,
    -3.79490003707156e-03, 1.61723488664661e-02, -5.29428148329736e-02, 1.15702180856167e-01,
    -2.62453564772299e-11, 3.24031041623823e-10, -3.614965656163e-09, 3.760256799971e-08,
    -3.553558319675e-07, 3.022556449731e-06, -2.290098979647e-05, 1.526537461148e-04,
    -8.81947375894379e-04, 4.33207949514611e-03, -1.75257821619926e-02, 5.28406320615584e-02,
    0, 0, 0, 0, 0, 4.6897511375022e-01, -6.9955602298985e-01, 5.3689283271887e-01, -3.2883030418398e-01,
    2.4645596956002e-01, -4.9984072848436e-01, -3.1501078774085e-06,
    0, 0, 0, 0, 0, 0, 0, 0, -1.8784686463512e-01, 2.2991849164985e-01, -4.9893752514047e-01, -2.1916512131607e-05,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 1.9623264149430e-01, -4.9695241464490e-01, -6.0156581186481e-05,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

// 1/x for the positive exponent sums of a primitive quartet: the hardware
// reciprocal refined by two Newton-Raphson steps, r <- r + r (1 - x r), each of
// which squares the relative error, so the result is within an ulp or two of
// 1/x (the IEEE division sequence costs twice the instructions).  FP32: the
// hardware reciprocal (the FP32 path's error budget is 1e-6).
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }

// 1/sqrt(x) for the same sums: the hardware estimate refined by two Newton
// steps, y <- y + y (1 - x y^2) / 2 (FP64); the hardware estimate (FP32).  A
// primitive quartet takes 1/(zeta+eta) as its square (one rounding more).
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    y = fma(y * 0.5, fma(-x * y, y, 1.0), y);
    return fma(y * 0.5, fma(-x * y, y, 1.0), y);
}
__device__ __forceinline__ float rsqrt_nr(float x) { return __builtin_amdgcn_rsqf(x); }

template <int LT>
__device__ __forceinline__ void boys_fit(double X, int Mb, const double* __restrict__ tab, double* F) {
    const double E = exp(-X);
    const bool up = (X > 3.0e-5) || (X > 3.0e-7 && Mb < 4);
    if (up) {
        const double RS = rsqrt_nr(X);           // 1/sqrt(X); 1/X as its square
        const double XINV = RS * RS;
        const int r = int(X > 1.0) + int(X > 3.0) + int(X > 5.0) + int(X > 10.0) + int(X > 15.0) + int(X > 33.0);
        const double z = r == 0 ? X : (r == 1 ? X - 2.0 : (r == 2 ? X - 4.0 : XINV));
        const double* c = tab + kBoysFitTerms * r;
        double v = c[0];
#pragma unroll
        for (int k = 1; k < kBoysFitTerms; ++k) v = fma(v, z, c[k]);
        const double s = 0.88622692545275801365 * RS;   // sqrt(PIE4 / X), sqrt(pi/4) = 0.886...
        F[0] = r <= 2 ? fma(v, X + X, E) : fma(v, E, s);
#pragma unroll
        for (int m = 1; m <= LT; ++m) F[m] = (double(2 * m - 1) * F[m - 1] - E) * 0.5 * XINV;
    } else {
        double v = (1.0 - X) / double(2 * Mb + 1);
        const double twot = X * 2.0;
        for (int m = Mb - 1; m >= LT; --m) v = (twot * v + E) / double(m * 2 + 1);
        F[LT] = v;
#pragma unroll
        for (int m = LT - 1; m >= 0; --m) F[m] = (twot * F[m + 1] + E) / double(m * 2 + 1);
    }
}

// Copies primitive pairs base, base + stride, ... (n <= 64 of them) into the
// wavefront's slab: field f of pair i at slab[f * 64 + i] (zeta, 1/2zeta, P, K,
// cutprim).  Every lane of the wavefront takes part; the fences order the slab's
// earlier reads before these writes and these writes before later reads (LDS
// operations of one wavefront complete in order).
template <class PP, typename R>
__device__ __forceinline__ void stage_bra(const PP& pp, int base, int stride, int n, int lane, R* slab) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < n) {
        const int k = base + lane * stride;
        slab[lane] = pp.zeta[k];
        slab[64 + lane] = pp.oo2z[k];
        slab[2 * 64 + lane] = pp.Px[k];
        slab[3 * 64 + lane] = pp.Py[k];
        slab[4 * 64 + lane] = pp.Pz[k];
        slab[5 * 64 + lane] = pp.K[k];
        slab[6 * 64 + lane] = pp.cut[k];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exchange contribution to X[row r of the bra][col]: into the block's LDS rows
// (xr, n columns per bra function) when the block works on one bra at a time,
// else straight into X.
__device__ __forceinline__ void xk_add(unsigned long long* X, unsigned long long* xr, int r, int grow, int col,
                                       int n, double v) {
    const long long iv = __double2ll_rn(v * 0x1p48);
    if (iv == 0) return;
    if (xr) atomicAdd(xr + (long long)r * n + col, (unsigned long long)iv);
    else atomicAdd(X + (long long)grow * n + col, (unsigned long long)iv);
}

__device__ __forceinline__ void fx_add(unsigned long long* X, long long idx, double v) {
    const long long iv = __double2ll_rn(v * 0x1p48);
    if (iv != 0) atomicAdd(X + idx, (unsigned long long)iv);
}

// round(v * 2^48) for |v| < 2^15 in FP32 arithmetic: with a = v 2^16 (exact),
// its integer part h = trunc(a) and fraction f = a - h (exact: the low bits of
// a, |f| <= 1 - 2^-24), the result is h 2^32 + rint(f 2^32); as h 2^32 is even
// and rint is odd-symmetric, this is the integer __double2ll_rn(double(v) *
// 0x1p48) gives, without FP64 instructions
__device__ __forceinline__ long long fx_round_f(float v) {
    const float a = v * 65536.0f;
    const float h = truncf(a);
    const unsigned lo = (unsigned)rintf(fabsf(a - h) * 4294967296.0f);
    const long long hi = (long long)(int)h * 4294967296LL;
    return a < 0.0f ? hi - (long long)lo : hi + (long long)lo;
}

// xk_add for an FP32 contribution
__device__ __forceinline__ void xk_add_f(unsigned long long* X, unsigned long long* xr, int r, int grow, int col,
                                         int n, float v) {
    const long long iv = fx_round_f(v);
    if (iv == 0) return;
    if (xr) atomicAdd(xr + (long long)r * n + col, (unsigned long long)iv);
    else atomicAdd(X + (long long)grow * n + col, (unsigned long long)iv);
}

__device__ __forceinline__ double dn_max(const double* cm, int ns, int A, int B, int C, int D) {
    double m = fmax(4.0 * cm[A * ns + B], 4.0 * cm[C * ns + D]);
    m = fmax(m, fmax(fmax(cm[A * ns + D], cm[A * ns + C]), fmax(cm[B * ns + C], cm[B * ns + D])));
    return m;
}

template <int LA, int LB, int LC, int LD>
struct ClassOps;

#define QK_CLASS(la, lb, lc, ld, TAG)                                                                    \
    template <>                                                                                          \
    struct ClassOps<la, lb, lc, ld> {                                                                    \
        template <typename R>                                                                            \
        static __device__ __forceinline__ void vrr(const R* F, R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, \
                                                   R QCx, R QCy, R QCz, R WQx, R WQy, R WQz, R oo2z,     \
                                                   R oo2e, R oo2ze, R roz, R roe, R* acc) {              \
            vrr_##TAG<R>(F, PAx, PAy, PAz, WPx, WPy, WPz, QCx, QCy, QCz, WQx, WQy, WQz, oo2z, oo2e,      \
                         oo2ze, roz, roe, acc);                                                          \
        }                                                                                                \
        template <typename R>                                                                            \
        static __device__ __forceinline__ void hrr(const R* acc, R ABx, R ABy, R ABz, R CDx, R CDy,      \
                                                   R CDz, R* out) {                                      \
            hrr_##TAG<R>(acc, ABx, ABy, ABz, CDx, CDy, CDz, out);                                        \
        }                                                                                                \
    };

QK_CLASS(0, 0, 0, 0, ssss)
QK_CLASS(1, 0, 0, 0, psss)
QK_CLASS(1, 0, 1, 0, psps)
QK_CLASS(1, 1, 0, 0, ppss)
QK_CLASS(1, 1, 1, 0, ppps)
QK_CLASS(1, 1, 1, 1, pppp)
QK_CLASS(2, 0, 0, 0, dsss)
QK_CLASS(2, 0, 1, 0, dsps)
QK_CLASS(2, 0, 1, 1, dspp)
QK_CLASS(2, 0, 2, 0, dsds)
QK_CLASS(2, 1, 0, 0, dpss)
QK_CLASS(2, 1, 1, 0, dpps)
QK_CLASS(2, 1, 1, 1, dppp)
QK_CLASS(2, 1, 2, 0, dpds)
QK_CLASS(2, 1, 2, 1, dpdp)
QK_CLASS(2, 2, 0, 0, ddss)
QK_CLASS(2, 2, 1, 0, ddps)
QK_CLASS(2, 2, 1, 1, ddpp)
QK_CLASS(2, 2, 2, 0, ddds)
QK_CLASS(2, 2, 2, 1, dddp)
QK_CLASS(2, 2, 2, 2, dddd)

// Wavefronts per SIMD the light kernels are compiled for: two (256 VGPRs, no
// spills for the small accumulators), one for the larger classes.
template <int LA, int LB, int LC, int LD>
struct Occupancy {
    static constexpr int nacc = nrange(LA, LA + LB) * nrange(LC, LC + LD);
    static constexpr int value = nacc <= 36 ? 2 : 1;
};

// int64 sum over the 64 lanes of a wavefront (exact, so order-free)
__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// STAGE: run mode only (the host launches it with runs): the run's bra rows of
// the density staged in LDS after the bra slab, the exchange contractions read
// there; a separate instantiation, so the unstaged kernel carries no second path
template <int LA, int LB, int LC, int LD, bool STAGE = false>
__global__ void __launch_bounds__(256, (Occupancy<LA, LB, LC, LD>::value)) fock_kernel(const FockArgs a, const ClassWork w) {
    constexpr int NE = nrange(LA, LA + LB), NF = nrange(LC, LC + LD);
    constexpr int NACC = NE * NF;
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NOUT = NA * NB * NC * ND;
    constexpr int LT = LA + LB + LC + LD;
    const double X0 = 34.98683665524972;  // 2 pi^(5/2), quick_constants_module.f90:55
    const int n = a.n;
    const int lane = threadIdx.x & 63;
    unsigned long long* const Xc = a.X + (long long)(blockIdx.x % a.xcopies) * n * n;
    const long long wave0 = __builtin_amdgcn_readfirstlane(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    unsigned long long nsurv = 0, nprimq = 0;
    __shared__ double s_boys[kBoysFitRegions * kBoysFitTerms];
    for (int k = threadIdx.x; k < kBoysFitRegions * kBoysFitTerms; k += blockDim.x) s_boys[k] = kBoysFit[k];
    __syncthreads();
    // one tile per wavefront and iteration: a bra pair shared by all lanes and
    // one ket pair per lane; every lane reaches the reduction below
    // dynamic LDS: in run mode the exchange rows of the run's bra, then that
    // bra's primitive pairs (at most 64; one slab for the block, staged once per
    // run); otherwise one slab per wavefront, staged per tile
    extern __shared__ unsigned long long s_xrows[];
    const bool runs = w.runs != nullptr;
    unsigned long long* const xr = runs ? s_xrows : nullptr;
    const int wib = threadIdx.x >> 6;
    double* const bra = runs ? reinterpret_cast<double*>(s_xrows + (NA + NB) * n)
                             : reinterpret_cast<double*>(s_xrows) + wib * (7 * 64);
    double* const sD = STAGE ? bra + 7 * 64 : nullptr;   // the run's bra rows of D
    const bool compact = runs && w.segs != nullptr;
    int c_seg = 0, c_end = 0, c_step = 0, c_nst = -1, c_gt = 0;   // compacted-run walk
    long long run = blockIdx.x;
    for (;;) {
    long long t0, t1, tstep;
    int rp = 0;
    if (runs) {
        if (run >= w.nruns) break;
        const int2 rg = w.runs[run];
        t0 = rg.x + wib; t1 = rg.y; tstep = 4;
        rp = compact ? w.segs[rg.x].x : w.tiles[rg.x].p;
        c_seg = rg.x; c_end = rg.y; c_step = 0; c_nst = -1; c_gt = 0;
        for (int k = threadIdx.x; k < (NA + NB) * n; k += blockDim.x) s_xrows[k] = 0ull;
        if (wib == 0) stage_bra(a.pp, a.pr.prim0[rp], a.pr.pstride[rp], a.pr.nprim[rp], lane, bra);
        if constexpr (STAGE) {
            const int fa0 = a.q.f0[a.pr.qa[rp]], fb0 = a.q.f0[a.pr.qb[rp]];
            for (int k = threadIdx.x; k < (NA + NB) * n; k += blockDim.x) {
                const int r = k / n, col = k - r * n;
                sD[k] = a.Dp[(long long)(r < NA ? fa0 + r : fb0 + r - NA) * n + col];
            }
        }
        __syncthreads();
    } else {
        t0 = wave0; t1 = w.ntiles; tstep = nwaves;
    }
    for (long long t = t0;; t += tstep) {
        int p, q;
        bool inrange;
        if (compact) {
            // the run's segments in steps of 64 surviving kets; step g of the
            // run goes to wavefront g % 4 (all of this is wavefront-uniform)
            bool found = false;
            while (c_seg < c_end) {
                if (c_nst < 0) c_nst = (w.counts[c_seg] + 63) >> 6;
                if (c_step < c_nst) {
                    const bool mine = (c_gt & 3) == wib;
                    ++c_gt;
                    if (mine) { found = true; break; }
                    ++c_step;
                    continue;
                }
                ++c_seg; c_step = 0; c_nst = -1;
            }
            if (!found) break;
            const int4 sg = w.segs[c_seg];
            const int idx = c_step * 64 + lane;
            p = __builtin_amdgcn_readfirstlane(sg.x);
            inrange = idx < w.counts[c_seg];
            q = inrange ? w.list[sg.w + idx] : 0;
            ++c_step;
        } else {
            if (t >= t1) break;
            const Tile tl = w.tiles[t];
            p = __builtin_amdgcn_readfirstlane(tl.p);
            q = tl.q0 + lane;
            inrange = lane < tl.len;
        }

        const int qa = a.pr.qa[p], qb = a.pr.qb[p];
        const int qc = inrange ? a.pr.qa[q] : 0, qd = inrange ? a.pr.qb[q] : 0;
        bool active = false;
        double dn = 0.0;
        if (inrange) {
            const int sA = a.q.shell[qa], sB = a.q.shell[qb], sC = a.q.shell[qc], sD = a.q.shell[qd];
            dn = dn_max(a.cutmat, a.nshell, sA, sB, sC, sD);
            active = a.pr.Y[q] * a.pr.Y[p] * dn > a.tau;
        }
        const int fa = a.q.f0[qa], fb = a.q.f0[qb];
        const double* Dp = a.Dp;
        double jab[NA * NB];
#pragma unroll
        for (int k = 0; k < NA * NB; ++k) jab[k] = 0.0;

        const int pb0 = a.pr.prim0[p], pbs = a.pr.pstride[p], nbp = a.pr.nprim[p];
        if (!runs) stage_bra(a.pp, pb0, pbs, nbp, lane, bra);   // run mode: p is the run's bra, staged

        if (active) {
            ++nsurv;
            const int Mb = a.boys_qshell_order ? LT : a.q.qfinal[qa] + a.q.qfinal[qb] + a.q.qfinal[qc] + a.q.qfinal[qd];
            const int atA = a.q.atom[qa], atB = a.q.atom[qb], atC = a.q.atom[qc], atD = a.q.atom[qd];
            const double Ax = a.xyz[3 * atA], Ay = a.xyz[3 * atA + 1], Az = a.xyz[3 * atA + 2];
            const double Cx = a.xyz[3 * atC], Cy = a.xyz[3 * atC + 1], Cz = a.xyz[3 * atC + 2];

            double acc[NACC];
#pragma unroll
            for (int k = 0; k < NACC; ++k) acc[k] = 0.0;

            // kets outer (vector loads; the next ket primitive is fetched while the
            // current one runs), bra primitives inner, read from the bra slab in LDS
            // (the same address on every lane: broadcast reads)
            const int pk0 = a.pr.prim0[q], pks = a.pr.pstride[q], nkp = a.pr.nprim[q];
            {
                const int ni = nbp;   // <= 64 primitive pairs per Qshell pair (checked on the host)
                double nz = a.pp.zeta[pk0], no = a.pp.oo2z[pk0], nx = a.pp.Px[pk0], ny = a.pp.Py[pk0],
                       nzz = a.pp.Pz[pk0], nK = a.pp.K[pk0], nc = a.pp.cut[pk0];
                for (int j = 0; j < nkp; ++j) {
                    const double zcd = nz, oo2e = no, Qx = nx, Qy = ny, Qz = nzz, Kcd = nK, ccd = nc;
                    if (j + 1 < nkp) {
                        const int kn = pk0 + (j + 1) * pks;
                        nz = a.pp.zeta[kn]; no = a.pp.oo2z[kn]; nx = a.pp.Px[kn]; ny = a.pp.Py[kn];
                        nzz = a.pp.Pz[kn]; nK = a.pp.K[kn]; nc = a.pp.cut[kn];
                    }
                    const double QCx = Qx - Cx, QCy = Qy - Cy, QCz = Qz - Cz;
                    // two bra primitives interleaved where it measured faster
                    // ((ss|ss): 4.24 -> 3.74 ms on (H2O)32; the five-unit classes ~5%)
                    constexpr int kUnroll = (LT == 0 || LT == 5) ? 2 : 1;
#pragma unroll kUnroll
                    for (int i = 0; i < ni; ++i) {
                        if (!((dn * bra[6 * 64 + i]) * ccd > a.prim_limit)) continue;
                        ++nprimq;
                        const double zab = bra[i], oo2z = bra[64 + i];
                        const double Px = bra[2 * 64 + i], Py = bra[3 * 64 + i], Pz = bra[4 * 64 + i];
                        const double rs = rsqrt_nr(zab + zcd);   // 1/sqrt(zeta+eta): reciprocal and root in one
                        const double inv = rs * rs;
                        const double roz = zcd * inv, roe = zab * inv;
                        const double ROU = zab * roz;
                        const double Wx = (Px * zab + Qx * zcd) * inv;
                        const double Wy = (Py * zab + Qy * zcd) * inv;
                        const double Wz = (Pz * zab + Qz * zcd) * inv;
                        const double dx = Px - Qx, dy = Py - Qy, dz = Pz - Qz;
                        const double T = (dx * dx + dy * dy + dz * dz) * ROU;
                        double F[LT + 1];
                        boys_fit<LT>(T, Mb, s_boys, F);
                        const double s = X0 * bra[5 * 64 + i] * Kcd * rs;
#pragma unroll
                        for (int m = 0; m <= LT; ++m) F[m] *= s;
                        ClassOps<LA, LB, LC, LD>::template vrr<double>(
                            F, Px - Ax, Py - Ay, Pz - Az, Wx - Px, Wy - Py, Wz - Pz, QCx, QCy, QCz,
                            Wx - Qx, Wy - Qy, Wz - Qz, oo2z, oo2e, 0.5 * inv, roz, roe, acc);
                    }
                }
            }

            double out[NOUT];
            {
                const double Bx = a.xyz[3 * atB], By = a.xyz[3 * atB + 1], Bz = a.xyz[3 * atB + 2];
                const double Dx = a.xyz[3 * atD], Dy = a.xyz[3 * atD + 1], Dz = a.xyz[3 * atD + 2];
                ClassOps<LA, LB, LC, LD>::template hrr<double>(acc, Ax - Bx, Ay - By, Az - Bz, Cx - Dx, Cy - Dy,
                                                               Cz - Dz, out);
            }

            // digestion: G = (X + X^T)/2 with X accumulating
            //   X[a][b] += 4w D[c][d] I,  X[c][d] += 4w D[a][b] I,
            //   X[a][c] -= h w D[b][d] I, X[a][d] -= h w D[b][c] I,
            //   X[b][c] -= h w D[a][d] I, X[b][d] -= h w D[a][c] I
            // w = 1/((1+[qa==qb])(1+[qc==qd])(1+[p==q])) is the share of the quartet's
            // 8-fold orbit this component set represents.
            const int fc = a.q.f0[qc], fd = a.q.f0[qd];
            double wgt = 1.0;
            if (qa == qb) wgt *= 0.5;
            if (qc == qd) wgt *= 0.5;
            if (p == q) wgt *= 0.5;
            const double jw = 4.0 * wgt, kw = -a.hybrid * wgt;
#define OUT(ia, ib, ic, id) out[(((ia)*NB + (ib)) * NC + (ic)) * ND + (id)]
            for (int ia = 0; ia < NA; ++ia)
                for (int ib = 0; ib < NB; ++ib) {
                    double s = 0.0;
                    for (int ic = 0; ic < NC; ++ic)
                        for (int id = 0; id < ND; ++id) s += Dp[(long long)(fc + ic) * n + fd + id] * OUT(ia, ib, ic, id);
                    jab[ia * NB + ib] = jw * s;
                }
            for (int ic = 0; ic < NC; ++ic)
                for (int id = 0; id < ND; ++id) {
                    double s = 0.0;
                    for (int ia = 0; ia < NA; ++ia)
                        for (int ib = 0; ib < NB; ++ib) s += Dp[(long long)(fa + ia) * n + fb + ib] * OUT(ia, ib, ic, id);
                    fx_add(a.J, a.jofs[q] + (long long)(ic * ND + id) * pks, jw * s);
                }
            if (a.hybrid != 0.0) {
                // exchange: the bra rows of D from LDS (staged per run) or from memory
                auto exch = [&](auto DG) {
                    for (int ia = 0; ia < NA; ++ia)
                        for (int ic = 0; ic < NC; ++ic) {
                            double s = 0.0;
                            for (int ib = 0; ib < NB; ++ib)
                                for (int id = 0; id < ND; ++id) s += DG(NA + ib, fb + ib, fd + id) * OUT(ia, ib, ic, id);
                            xk_add(Xc, xr, ia, fa + ia, fc + ic, n, (kw * s));
                        }
                    for (int ia = 0; ia < NA; ++ia)
                        for (int id = 0; id < ND; ++id) {
                            double s = 0.0;
                            for (int ib = 0; ib < NB; ++ib)
                                for (int ic = 0; ic < NC; ++ic) s += DG(NA + ib, fb + ib, fc + ic) * OUT(ia, ib, ic, id);
                            xk_add(Xc, xr, ia, fa + ia, fd + id, n, (kw * s));
                        }
                    for (int ib = 0; ib < NB; ++ib)
                        for (int ic = 0; ic < NC; ++ic) {
                            double s = 0.0;
                            for (int ia = 0; ia < NA; ++ia)
                                for (int id = 0; id < ND; ++id) s += DG(ia, fa + ia, fd + id) * OUT(ia, ib, ic, id);
                            xk_add(Xc, xr, NA + ib, fb + ib, fc + ic, n, (kw * s));
                        }
                    for (int ib = 0; ib < NB; ++ib)
                        for (int id = 0; id < ND; ++id) {
                            double s = 0.0;
                            for (int ia = 0; ia < NA; ++ia)
                                for (int ic = 0; ic < NC; ++ic) s += DG(ia, fa + ia, fc + ic) * OUT(ia, ib, ic, id);
                            xk_add(Xc, xr, NA + ib, fb + ib, fd + id, n, (kw * s));
                        }
                };
                if constexpr (STAGE) exch([&](int r, int, int col) { return sD[r * n + col]; });
                else exch([&](int, int grow, int col) { return Dp[(long long)grow * n + col]; });
            }
#undef OUT
        }

        // J on the bra block: converted to fixed point per lane and summed across
        // the wavefront (all lanes share the bra pair): one atomic per element
        // instead of up to 64.
#pragma unroll
        for (int k = 0; k < NA * NB; ++k) {
            const long long v = wave_sum_i64(__double2ll_rn(jab[k] * 0x1p48));
            if (lane == 0 && v != 0) {
                if (xr) atomicAdd(xr + (long long)(k / NB) * n + fb + k % NB, (unsigned long long)v);
                else atomicAdd(a.J + a.jofs[p] + (long long)k * pbs, (unsigned long long)v);
            }
        }
    }
    if (!runs) break;
    // add the run's rows to X once
    __syncthreads();
    {
        const int fa0 = a.q.f0[a.pr.qa[rp]], fb0 = a.q.f0[a.pr.qb[rp]];
        for (int k = threadIdx.x; k < (NA + NB) * n; k += blockDim.x) {
            const unsigned long long v = s_xrows[k];
            if (v == 0ull) continue;
            const int r = k / n, col = k - r * n;
            const int grow = r < NA ? fa0 + r : fb0 + (r - NA);
            atomicAdd(Xc + (long long)grow * n + col, v);
        }
    }
    __syncthreads();
    run += gridDim.x;
    }
    if (nsurv) atomicAdd(a.counters + 0, nsurv);
    if (nprimq) atomicAdd(a.counters + 1, nprimq);
}

// ---------------------------------------------------------------------------
// FP32 path (mixed precision): quartets whose Schwarz bound Y_ij*Y_kl lies
// below the FP32 threshold.  QUICK's Boys fit recurs upward, which is unusable
// in single precision at small T, so this path uses a tabulated Taylor
// expansion of F_LT about the nearest grid point (5 terms, grid 1/8: the
// truncation |d|^5/5! <= 8e-9 relative is below FP32 resolution) and the
// stable downward recursion; beyond T = 36 the asymptotic form and upward
// recursion (the exp(-T) terms are below FP32 resolution there).  The operator
// contributions are promoted to FP64 before the fixed-point accumulation.
//
// The light kernels of the small classes stage the five orders they expand in
// LDS once per block (the per-primitive gather then costs an LDS round trip
// instead of a memory one); where those 5.8 KB would cost resident blocks
// (larger exchange rows, the one-wavefront team kernels) the table is read from
// memory (measured on (H2O)32: (ss|ss) FP32 1.85 -> 1.36 ms with it, (ds|ps)
// 1.28 -> 1.57 ms, dppp gradient 13.5 -> 21.9 ms).
template <int LT>
__device__ __forceinline__ void stage_boys_f32(const float* __restrict__ tab, float* s_tab) {
    for (int i = threadIdx.x; i < kBoysTabRows * kBoysLdsCols; i += blockDim.x) {
        const int k = i / kBoysLdsCols, j = i - k * kBoysLdsCols;
        s_tab[i] = tab[k * kBoysTabM + LT + j];
    }
}

// F_0..F_LT(T) from orders LT..LT+4 of the table: tab points at order LT of
// grid point 0, rows kStride apart (the LDS slice: kBoysLdsCols; the table in
// memory: a.boys_tab + LT and kBoysTabM)
template <int LT, int kStride = kBoysLdsCols>
__device__ __forceinline__ void boys_f32(float T, const float* tab, float* F) {
    if (T < kBoysTabTmax) {
        const int k = __float2int_rn(T * float(kBoysTabPerUnit));
        const float d = float(k) * (1.0f / float(kBoysTabPerUnit)) - T;
        const float* r = tab + k * kStride;
        // F_LT(T) = sum_j F_{LT+j}(T_k) d^j / j!,  d = T_k - T, |d| <= 1/16
        F[LT] = (((r[4] * (d * 0.25f) + r[3]) * (d * (1.0f / 3.0f)) + r[2]) * (d * 0.5f) + r[1]) * d + r[0];
        const float e = __expf(-T);
        const float t2 = T + T;
#pragma unroll
        for (int m = LT - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * (1.0f / float(2 * m + 1));
    } else {
        // T >= kBoysTabTmax: sqrt(pi / 4T) by the hardware reciprocal square root
        // (no denormal or zero operand here)
        F[0] = 0.88622692545275801f * __builtin_amdgcn_rsqf(T);
        const float h = 0.5f * __builtin_amdgcn_rcpf(T);
#pragma unroll
        for (int m = 1; m <= LT; ++m) F[m] = F[m - 1] * (float(2 * m - 1) * h);
    }
}

// the primitive-pair data, exponents and density in the kernel's precision
template <typename R>
__device__ __forceinline__ const auto& grad_prims(const GradArgs& a) {
    if constexpr (sizeof(R) == 8) return a.pp; else return a.ppf;
}
template <typename R>
__device__ __forceinline__ const R* grad_ea(const GradArgs& a) {
    if constexpr (sizeof(R) == 8) return a.ea; else return a.eaf;
}
template <typename R>
__device__ __forceinline__ const R* grad_eb(const GradArgs& a) {
    if constexpr (sizeof(R) == 8) return a.eb; else return a.ebf;
}
template <typename R>
__device__ __forceinline__ const R* grad_density(const GradArgs& a) {
    if constexpr (sizeof(R) == 8) return a.Dp; else return a.Dpf;
}
__device__ __forceinline__ double sqrt_r(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrt_r(float x) { return sqrtf(x); }

// F_0..F_kMaxGradF(T) in FP32 for the gradient's table-driven kernel: the
// expansion of boys_f32 at the highest order the table supports for it (read
// from the table in memory), then downward recursion (stable), so every order
// is valid whatever the class
constexpr int kMaxGradF = 10;
static_assert(kMaxGradF + 4 < kBoysTabM, "Boys table too narrow for the gradient");
__device__ __forceinline__ void boys_f32_all(float T, const float* tab, float* F) {
    boys_f32<kMaxGradF, kBoysTabM>(T, tab + kMaxGradF, F);
}

__device__ __forceinline__ void fx_add_f(unsigned long long* X, long long idx, float v) {
    const long long iv = fx_round_f(v);
    if (iv != 0) atomicAdd(X + idx, (unsigned long long)iv);
}

template <int LA, int LB, int LC, int LD>
struct OccupancyF {
    static constexpr int nacc = nrange(LA, LA + LB) * nrange(LC, LC + LD);
    // six or more exchange rows already hold a run-mode block to three per CU
    // at 800 functions (more functions: fewer): asking for four wavefronts per
    // SIMD only made such kernels spill (measured: ppps 1.34 -> 1.27 ms, dsds
    // 0.44 -> 0.42 ms with three)
    static constexpr int rows = ncart(LA) + ncart(LB);
    static constexpr int value = nacc <= 36 ? (rows >= 6 ? 3 : 4) : 2;
};

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(256, (OccupancyF<LA, LB, LC, LD>::value)) fock_kernel_f32(const FockArgs a, const ClassWork w) {
    constexpr int NE = nrange(LA, LA + LB), NF = nrange(LC, LC + LD);
    constexpr int NACC = NE * NF;
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NOUT = NA * NB * NC * ND;
    constexpr int LT = LA + LB + LC + LD;
    const float X0 = 34.98683665524972f;
    const int n = a.n;
    const int lane = threadIdx.x & 63;
    unsigned long long* const Xc = a.X + (long long)(blockIdx.x % a.xcopies) * n * n;
    const long long wave0 = __builtin_amdgcn_readfirstlane(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    unsigned long long nsurv = 0, nprimq = 0;
    const float plim = float(a.prim_limit);
    constexpr bool kBoysLds = NA + NB <= 4;   // two or four exchange rows
    __shared__ float s_btab[kBoysLds ? kBoysTabRows * kBoysLdsCols : 1];
    if constexpr (kBoysLds) {
        stage_boys_f32<LT>(a.boys_tab, s_btab);
        __syncthreads();
    }

    // dynamic LDS: in run mode the exchange rows of the run's bra, then that
    // bra's primitive pairs (one slab for the block, staged once per run);
    // otherwise one slab per wavefront, staged per tile
    extern __shared__ unsigned long long s_xrows[];
    const bool runs = w.runs != nullptr;
    unsigned long long* const xr = runs ? s_xrows : nullptr;
    const int wib = threadIdx.x >> 6;
    float* const bra = runs ? reinterpret_cast<float*>(s_xrows + (NA + NB) * n)
                            : reinterpret_cast<float*>(s_xrows) + wib * (7 * 64);
    const bool compact = runs && w.segs != nullptr;
    int c_seg = 0, c_end = 0, c_step = 0, c_nst = -1, c_gt = 0;   // compacted-run walk
    long long run = blockIdx.x;
    for (;;) {
    long long t0, t1, tstep;
    int rp = 0;
    if (runs) {
        if (run >= w.nruns) break;
        const int2 rg = w.runs[run];
        t0 = rg.x + wib; t1 = rg.y; tstep = 4;
        rp = compact ? w.segs[rg.x].x : w.tiles[rg.x].p;
        c_seg = rg.x; c_end = rg.y; c_step = 0; c_nst = -1; c_gt = 0;
        for (int k = threadIdx.x; k < (NA + NB) * n; k += blockDim.x) s_xrows[k] = 0ull;
        if (wib == 0) stage_bra(a.ppf, a.pr.prim0[rp], a.pr.pstride[rp], a.pr.nprim[rp], lane, bra);
        __syncthreads();
    } else {
        t0 = wave0; t1 = w.ntiles; tstep = nwaves;
    }
    for (long long t = t0;; t += tstep) {
        int p, q;
        bool inrange;
        if (compact) {
            // the run's segments in steps of 64 surviving kets; step g of the
            // run goes to wavefront g % 4 (all of this is wavefront-uniform)
            bool found = false;
            while (c_seg < c_end) {
                if (c_nst < 0) c_nst = (w.counts[c_seg] + 63) >> 6;
                if (c_step < c_nst) {
                    const bool mine = (c_gt & 3) == wib;
                    ++c_gt;
                    if (mine) { found = true; break; }
                    ++c_step;
                    continue;
                }
                ++c_seg; c_step = 0; c_nst = -1;
            }
            if (!found) break;
            const int4 sg = w.segs[c_seg];
            const int idx = c_step * 64 + lane;
            p = __builtin_amdgcn_readfirstlane(sg.x);
            inrange = idx < w.counts[c_seg];
            q = inrange ? w.list[sg.w + idx] : 0;
            ++c_step;
        } else {
            if (t >= t1) break;
            const Tile tl = w.tiles[t];
            p = __builtin_amdgcn_readfirstlane(tl.p);
            q = tl.q0 + lane;
            inrange = lane < tl.len;
        }
        const int qa = a.pr.qa[p], qb = a.pr.qb[p];
        const int qc = inrange ? a.pr.qa[q] : 0, qd = inrange ? a.pr.qb[q] : 0;
        bool active = false;
        double dn = 0.0;
        if (inrange) {
            const int sA = a.q.shell[qa], sB = a.q.shell[qb], sC = a.q.shell[qc], sD = a.q.shell[qd];
            dn = dn_max(a.cutmat, a.nshell, sA, sB, sC, sD);
            active = a.pr.Y[q] * a.pr.Y[p] * dn > a.tau;
        }
        const int fa = a.q.f0[qa], fb = a.q.f0[qb];
        const float* Dp = a.Dpf;
        float jab[NA * NB];
#pragma unroll
        for (int k = 0; k < NA * NB; ++k) jab[k] = 0.0f;
        const int pb0 = a.pr.prim0[p], pbs = a.pr.pstride[p], nbp = a.pr.nprim[p];
        if (!runs) stage_bra(a.ppf, pb0, pbs, nbp, lane, bra);   // run mode: p is the run's bra, staged

        if (active) {
            ++nsurv;
            const int atA = a.q.atom[qa], atB = a.q.atom[qb], atC = a.q.atom[qc], atD = a.q.atom[qd];
            const float Ax = float(a.xyz[3 * atA]), Ay = float(a.xyz[3 * atA + 1]), Az = float(a.xyz[3 * atA + 2]);
            const float Cx = float(a.xyz[3 * atC]), Cy = float(a.xyz[3 * atC + 1]), Cz = float(a.xyz[3 * atC + 2]);
            float acc[NACC];
#pragma unroll
            for (int k = 0; k < NACC; ++k) acc[k] = 0.0f;
            const int pk0 = a.pr.prim0[q], pks = a.pr.pstride[q], nkp = a.pr.nprim[q];
            const float dnf = float(dn);
            {
                const int ni = nbp;   // <= 64 primitive pairs per Qshell pair (checked on the host)
                float nz = a.ppf.zeta[pk0], no = a.ppf.oo2z[pk0], nx = a.ppf.Px[pk0], ny = a.ppf.Py[pk0],
                      nzz = a.ppf.Pz[pk0], nK = a.ppf.K[pk0], nc = a.ppf.cut[pk0];
                for (int j = 0; j < nkp; ++j) {
                    const float zcd = nz, oo2e = no, Qx = nx, Qy = ny, Qz = nzz, Kcd = nK, ccd = nc;
                    if (j + 1 < nkp) {
                        const int kn = pk0 + (j + 1) * pks;
                        nz = a.ppf.zeta[kn]; no = a.ppf.oo2z[kn]; nx = a.ppf.Px[kn]; ny = a.ppf.Py[kn];
                        nzz = a.ppf.Pz[kn]; nK = a.ppf.K[kn]; nc = a.ppf.cut[kn];
                    }
                    const float QCx = Qx - Cx, QCy = Qy - Cy, QCz = Qz - Cz;
                    for (int i = 0; i < ni; ++i) {
                        if (!((dnf * bra[6 * 64 + i]) * ccd > plim)) continue;
                        ++nprimq;
                        const float zab = bra[i], oo2z = bra[64 + i];
                        const float Px = bra[2 * 64 + i], Py = bra[3 * 64 + i], Pz = bra[4 * 64 + i];
                        const float rs = rsqrt_nr(zab + zcd);
                        const float inv = rs * rs;
                        const float roz = zcd * inv, roe = zab * inv;
                        const float ROU = zab * roz;
                        const float Wx = (Px * zab + Qx * zcd) * inv;
                        const float Wy = (Py * zab + Qy * zcd) * inv;
                        const float Wz = (Pz * zab + Qz * zcd) * inv;
                        const float dx = Px - Qx, dy = Py - Qy, dz = Pz - Qz;
                        const float T = (dx * dx + dy * dy + dz * dz) * ROU;
                        float F[LT + 1];
                        if constexpr (kBoysLds) boys_f32<LT>(T, s_btab, F);
                        else boys_f32<LT, kBoysTabM>(T, a.boys_tab + LT, F);
                        const float s = X0 * bra[5 * 64 + i] * Kcd * rs;
#pragma unroll
                        for (int m = 0; m <= LT; ++m) F[m] *= s;
                        ClassOps<LA, LB, LC, LD>::template vrr<float>(
                            F, Px - Ax, Py - Ay, Pz - Az, Wx - Px, Wy - Py, Wz - Pz, QCx, QCy, QCz,
                            Wx - Qx, Wy - Qy, Wz - Qz, oo2z, oo2e, 0.5f * inv, roz, roe, acc);
                    }
                }
            }
            float out[NOUT];
            {
                const float Bx = float(a.xyz[3 * atB]), By = float(a.xyz[3 * atB + 1]), Bz = float(a.xyz[3 * atB + 2]);
                const float Dx = float(a.xyz[3 * atD]), Dy = float(a.xyz[3 * atD + 1]), Dz = float(a.xyz[3 * atD + 2]);
                ClassOps<LA, LB, LC, LD>::template hrr<float>(acc, Ax - Bx, Ay - By, Az - Bz, Cx - Dx, Cy - Dy,
                                                              Cz - Dz, out);
            }
            // digestion as in fock_kernel, in FP32, promoted to FP64 at the accumulate
            const int fc = a.q.f0[qc], fd = a.q.f0[qd];
            float wgt = 1.0f;
            if (qa == qb) wgt *= 0.5f;
            if (qc == qd) wgt *= 0.5f;
            if (p == q) wgt *= 0.5f;
            const float jw = 4.0f * wgt, kw = -float(a.hybrid) * wgt;
#define OUT(ia, ib, ic, id) out[(((ia)*NB + (ib)) * NC + (ic)) * ND + (id)]
            for (int ia = 0; ia < NA; ++ia)
                for (int ib = 0; ib < NB; ++ib) {
                    float s = 0.0f;
                    for (int ic = 0; ic < NC; ++ic)
                        for (int id = 0; id < ND; ++id) s += Dp[(long long)(fc + ic) * n + fd + id] * OUT(ia, ib, ic, id);
                    jab[ia * NB + ib] = jw * s;
                }
            for (int ic = 0; ic < NC; ++ic)
                for (int id = 0; id < ND; ++id) {
                    float s = 0.0f;
                    for (int ia = 0; ia < NA; ++ia)
                        for (int ib = 0; ib < NB; ++ib) s += Dp[(long long)(fa + ia) * n + fb + ib] * OUT(ia, ib, ic, id);
                    fx_add_f(a.J, a.jofs[q] + (long long)(ic * ND + id) * pks, jw * s);
                }
            if (a.hybrid != 0.0) {
                for (int ia = 0; ia < NA; ++ia)
                    for (int ic = 0; ic < NC; ++ic) {
                        float s = 0.0f;
                        for (int ib = 0; ib < NB; ++ib)
                            for (int id = 0; id < ND; ++id) s += Dp[(long long)(fb + ib) * n + fd + id] * OUT(ia, ib, ic, id);
                        xk_add_f(Xc, xr, ia, fa + ia, fc + ic, n, kw * s);
                    }
                for (int ia = 0; ia < NA; ++ia)
                    for (int id = 0; id < ND; ++id) {
                        float s = 0.0f;
                        for (int ib = 0; ib < NB; ++ib)
                            for (int ic = 0; ic < NC; ++ic) s += Dp[(long long)(fb + ib) * n + fc + ic] * OUT(ia, ib, ic, id);
                        xk_add_f(Xc, xr, ia, fa + ia, fd + id, n, kw * s);
                    }
                for (int ib = 0; ib < NB; ++ib)
                    for (int ic = 0; ic < NC; ++ic) {
                        float s = 0.0f;
                        for (int ia = 0; ia < NA; ++ia)
                            for (int id = 0; id < ND; ++id) s += Dp[(long long)(fa + ia) * n + fd + id] * OUT(ia, ib, ic, id);
                        xk_add_f(Xc, xr, NA + ib, fb + ib, fc + ic, n, kw * s);
                    }
                for (int ib = 0; ib < NB; ++ib)
                    for (int id = 0; id < ND; ++id) {
                        float s = 0.0f;
                        for (int ia = 0; ia < NA; ++ia)
                            for (int ic = 0; ic < NC; ++ic) s += Dp[(long long)(fa + ia) * n + fc + ic] * OUT(ia, ib, ic, id);
                        xk_add_f(Xc, xr, NA + ib, fb + ib, fd + id, n, kw * s);
                    }
            }
#undef OUT
        }
#pragma unroll
        for (int k = 0; k < NA * NB; ++k) {
            const long long v = wave_sum_i64(fx_round_f(jab[k]));
            if (lane == 0 && v != 0) {
                if (xr) atomicAdd(xr + (long long)(k / NB) * n + fb + k % NB, (unsigned long long)v);
                else atomicAdd(a.J + a.jofs[p] + (long long)k * pbs, (unsigned long long)v);
            }
        }
    }
    if (!runs) break;
    // add the run's rows to X once
    __syncthreads();
    {
        const int fa0 = a.q.f0[a.pr.qa[rp]], fb0 = a.q.f0[a.pr.qb[rp]];
        for (int k = threadIdx.x; k < (NA + NB) * n; k += blockDim.x) {
            const unsigned long long v = s_xrows[k];
            if (v == 0ull) continue;
            const int r = k / n, col = k - r * n;
            const int grow = r < NA ? fa0 + r : fb0 + (r - NA);
            atomicAdd(Xc + (long long)grow * n + col, v);
        }
    }
    __syncthreads();
    run += gridDim.x;
    }
    if (nsurv) atomicAdd(a.counters + 0, nsurv);
    if (nprimq) atomicAdd(a.counters + 1, nprimq);
}

// ---------------------------------------------------------------------------
// Schwarz bounds (schwarz.f90:46-548): for every Qshell pair of one pair class,
//   ymax[p]        = max over components of |(ab|ab)| contracted over all
//                    primitive quartets (classcutoff), and
//   pmax[prim i]   = max over components of |[ab|ab]| for the primitive pair i
//                    with itself, weighted by (X0 Xcoeff_i)^2 (classprim).
// The host reduces both over the Qshells of a shell pair and takes sqrt.
template <int LA, int LB>
__global__ void __launch_bounds__(256) schwarz_kernel(const int* __restrict__ pairs, int npairs, const int* __restrict__ pqa,
                                                      const int* __restrict__ pqb, const int* __restrict__ prim0,
                                                      const int* __restrict__ nprim, const DevQshells qs,
                                                      const DevPrims pp, const double* __restrict__ xyz,
                                                      const double* __restrict__ cons,
                                                      double* __restrict__ ymax, double* __restrict__ pmax) {
    constexpr int NE = nrange(LA, LA + LB);
    constexpr int NACC = NE * NE;
    constexpr int NA = ncart(LA), NB = ncart(LB);
    constexpr int NOUT = NA * NB * NA * NB;
    constexpr int LT = 2 * (LA + LB);
    const double X0 = 34.98683665524972;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    const int p = pairs[t];
    const int qa = pqa[p], qb = pqb[p];
    const int Mb = 2 * (qs.qfinal[qa] + qs.qfinal[qb]);
    const int atA = qs.atom[qa], atB = qs.atom[qb];
    const double Ax = xyz[3 * atA], Ay = xyz[3 * atA + 1], Az = xyz[3 * atA + 2];
    const double ABx = Ax - xyz[3 * atB], ABy = Ay - xyz[3 * atB + 1], ABz = Az - xyz[3 * atB + 2];
    const int p0 = prim0[p], np = nprim[p];
    double acc[NACC];
    double out[NOUT];
#pragma unroll
    for (int k = 0; k < NACC; ++k) acc[k] = 0.0;
    for (int i = 0; i < np; ++i) {
        const double zab = pp.zeta[p0 + i];
        const double Px = pp.Px[p0 + i], Py = pp.Py[p0 + i], Pz = pp.Pz[p0 + i];
        const double X2 = X0 * pp.K[p0 + i];
        for (int j = 0; j < np; ++j) {
            const double zcd = pp.zeta[p0 + j];
            const double Qx = pp.Px[p0 + j], Qy = pp.Py[p0 + j], Qz = pp.Pz[p0 + j];
            const double ABCD = zab + zcd;
            const double ROU = zab * zcd / ABCD;
            const double Wx = (Px * zab + Qx * zcd) / ABCD, Wy = (Py * zab + Qy * zcd) / ABCD,
                         Wz = (Pz * zab + Qz * zcd) / ABCD;
            const double dx = Px - Qx, dy = Py - Qy, dz = Pz - Qz;
            const double T = (dx * dx + dy * dy + dz * dz) * ROU;
            double F[LT + 1];
            boys_ref<LT>(T, Mb, F);
            const double s = X2 * pp.K[p0 + j] / sqrt(ABCD);
#pragma unroll
            for (int m = 0; m <= LT; ++m) F[m] *= s;
            ClassOps<LA, LB, LA, LB>::template vrr<double>(F, Px - Ax, Py - Ay, Pz - Az, Wx - Px, Wy - Py, Wz - Pz,
                                                           Qx - Ax, Qy - Ay, Qz - Az, Wx - Qx, Wy - Qy, Wz - Qz,
                                                           0.5 / zab, 0.5 / zcd, 0.5 / ABCD, zcd / ABCD, zab / ABCD, acc);
        }
    }
    ClassOps<LA, LB, LA, LB>::template hrr<double>(acc, ABx, ABy, ABz, ABx, ABy, ABz, out);
    // (ab|ab) carries cons(a)^2 cons(b)^2 (hrrwhole's cons product, subs/hrr.f90:295)
    double cab[NA * NB];
    for (int a = 0; a < NA; ++a)
        for (int b = 0; b < NB; ++b) {
            const double ca = cons[qs.f0[qa] + a], cb = cons[qs.f0[qb] + b];
            cab[a * NB + b] = (ca * ca) * (cb * cb);
        }
    double m = 0.0;
    for (int a = 0; a < NA; ++a)
        for (int b = 0; b < NB; ++b) m = fmax(m, fabs(out[((a * NB + b) * NA + a) * NB + b] * cab[a * NB + b]));
    ymax[p] = m;
    // one primitive pair with itself: P = Q, T = 0
    for (int i = 0; i < np; ++i) {
#pragma unroll
        for (int k = 0; k < NACC; ++k) acc[k] = 0.0;
        const double zab = pp.zeta[p0 + i];
        const double Px = pp.Px[p0 + i], Py = pp.Py[p0 + i], Pz = pp.Pz[p0 + i];
        const double ABCD = zab + zab;
        const double Wx = (Px * zab + Px * zab) / ABCD, Wy = (Py * zab + Py * zab) / ABCD,
                     Wz = (Pz * zab + Pz * zab) / ABCD;
        double F[LT + 1];
        boys_ref<LT>(0.0, Mb, F);
        const double x4 = X0 * pp.K[p0 + i];
        const double s = 1.0 / sqrt(ABCD);
#pragma unroll
        for (int k = 0; k <= LT; ++k) F[k] *= s;
        ClassOps<LA, LB, LA, LB>::template vrr<double>(F, Px - Ax, Py - Ay, Pz - Az, Wx - Px, Wy - Py, Wz - Pz,
                                                       Px - Ax, Py - Ay, Pz - Az, Wx - Px, Wy - Py, Wz - Pz,
                                                       0.5 / zab, 0.5 / zab, 0.5 / ABCD, zab / ABCD, zab / ABCD, acc);
#pragma unroll
        for (int k = 0; k < NACC; ++k) acc[k] *= x4 * x4;
        ClassOps<LA, LB, LA, LB>::template hrr<double>(acc, ABx, ABy, ABz, ABx, ABy, ABz, out);
        double mp = 0.0;
        for (int a = 0; a < NA; ++a)
            for (int b = 0; b < NB; ++b) mp = fmax(mp, fabs(out[((a * NB + b) * NA + a) * NB + b] * cab[a * NB + b]));
        pmax[p0 + i] = mp;
    }
}

}  // namespace qk
