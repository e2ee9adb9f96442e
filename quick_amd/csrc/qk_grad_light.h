// qk_grad_light.h -- nuclear gradient of the light quartet classes, one
// contracted quartet per lane (the generated straight-line code of
// gen_grad.py: vrg_* raises the VRR by one on both sides and accumulates the
// four contracted sets, gcon_* applies the derivative HRR and contracts with
// the density factors).  Same work tiles, screening, Boys order and
// weighting as the table-driven kernel in qk_grad.hip; the derivative sums of
// every quartet are converted to int64 fixed point (2^-44) and added into a
// per-block LDS copy of the gradient, flushed once per block (or straight into
// global memory when the gradient does not fit in LDS).  The work comes either
// as tiles (a bra pair and up to 64 consecutive kets) or, as the host launches
// it, compacted by compact_kernel to the quartets that pass the screen, packed
// 64 to a step in the tiles' places.
#pragma once
#include "qk_kernels.h"
#include "grad_light.inc"

namespace qk {

template <int LA, int LB, int LC, int LD>
struct GradLight;

#define QK_GRAD_LIGHT(la, lb, lc, ld, TAG)                                                                   \
    template <>                                                                                              \
    struct GradLight<la, lb, lc, ld> {                                                                       \
        using Sets = GradSets_##TAG;                                                                         \
        template <int M, typename R>                                                                         \
        static __device__ __forceinline__ void vrg(const R* F, R PAx, R PAy, R PAz, R WPx, R WPy, R WPz,   \
                                                   R QCx, R QCy, R QCz, R WQx, R WQy, R WQz, R oo2z, R oo2e, \
                                                   R oo2ze, R roz, R roe, R wA, R wB, R wC, R* a0, R* aA,   \
                                                   R* aB, R* aC) {                                           \
            vrg_##TAG<M, R>(F, PAx, PAy, PAz, WPx, WPy, WPz, QCx, QCy, QCz, WQx, WQy, WQz, oo2z, oo2e, oo2ze, \
                            roz, roe, wA, wB, wC, a0, aA, aB, aC);                                           \
        }                                                                                                    \
        template <int M, typename R>                                                                         \
        static __device__ __forceinline__ void gcon(const R* a0, const R* aA, const R* aB, const R* aC,     \
                                                    R ABx, R ABy, R ABz, R CDx, R CDy, R CDz, const R* K,   \
                                                    R* g) {                                                  \
            gcon_##TAG<M, R>(a0, aA, aB, aC, ABx, ABy, ABz, CDx, CDy, CDz, K, g);                           \
        }                                                                                                    \
    };
QK_GRAD_LIGHT(0, 0, 0, 0, ssss)
QK_GRAD_LIGHT(1, 0, 0, 0, psss)
QK_GRAD_LIGHT(1, 0, 1, 0, psps)
QK_GRAD_LIGHT(1, 1, 0, 0, ppss)
QK_GRAD_LIGHT(2, 0, 0, 0, dsss)
QK_GRAD_LIGHT(1, 1, 1, 0, ppps)
QK_GRAD_LIGHT(2, 0, 1, 0, dsps)
QK_GRAD_LIGHT(2, 1, 0, 0, dpss)
QK_GRAD_LIGHT(1, 1, 1, 1, pppp)
QK_GRAD_LIGHT(2, 0, 1, 1, dspp)
QK_GRAD_LIGHT(2, 0, 2, 0, dsds)
QK_GRAD_LIGHT(2, 1, 1, 0, dpps)
QK_GRAD_LIGHT(2, 2, 0, 0, ddss)
#undef QK_GRAD_LIGHT

// Two wavefronts per SIMD (256 VGPRs a lane) only while one pass's
// accumulators and density factors leave the VRR room there: up to 104 values
// in FP32, 64 in FP64; otherwise one wavefront per SIMD, with 512 registers,
// spills far less and runs faster (measured per class on (H2O)32: ppps FP64
// 19.8 -> 8.6 ms, dspp FP32 4.3 -> 2.0 ms; dsps FP32 and dsss FP64 prefer two).
template <int LA, int LB, int LC, int LD, typename R>
struct GradLightOcc {
    using S = typename GradLight<LA, LB, LC, LD>::Sets;
    static constexpr int nout = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
    static constexpr int value = S::kWidth + nout <= (sizeof(R) == 8 ? 64 : 104) ? 2 : 1;
};

// primitive pairs of the bra for the gradient: the Fock slab's seven fields
// plus the two exponents
template <typename R>
__device__ __forceinline__ void stage_bra_grad(const GradArgs& a, int base, int stride, int n, int lane, R* slab) {
    const auto& P = grad_prims<R>(a);
    const R* ea = grad_ea<R>(a);
    const R* eb = grad_eb<R>(a);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < n) {
        const int k = base + lane * stride;
        slab[lane] = P.zeta[k];
        slab[64 + lane] = P.oo2z[k];
        slab[2 * 64 + lane] = P.Px[k];
        slab[3 * 64 + lane] = P.Py[k];
        slab[4 * 64 + lane] = P.Pz[k];
        slab[5 * 64 + lane] = P.K[k];
        slab[6 * 64 + lane] = P.cut[k];
        slab[7 * 64 + lane] = ea[k];
        slab[8 * 64 + lane] = eb[k];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void gl_add(unsigned long long* lds, unsigned long long* G, bool use_lds, int idx,
                                       double v) {
    const long long iv = __double2ll_rn(v * 0x1p44);
    if (iv == 0) return;
    if (use_lds) atomicAdd(lds + idx, (unsigned long long)iv);
    else atomicAdd(G + idx, (unsigned long long)iv);
}

// R = double: QUICK's Boys fit (LDS coefficients); R = float (mixed
// precision, weak quartets): the tabulated Boys expansion of boys_f32, FP32
// arithmetic throughout, the nine sums promoted to FP64 at the accumulate.
template <int LA, int LB, int LC, int LD, typename R>
__global__ void __launch_bounds__(256, (GradLightOcc<LA, LB, LC, LD, R>::value)) grad_light_kernel(
    const GradArgs a, const ClassWork w) {
    using Ops = GradLight<LA, LB, LC, LD>;
    constexpr bool f64 = sizeof(R) == 8;
    constexpr int n0 = Ops::Sets::n0, nA = Ops::Sets::nA, nB = Ops::Sets::nB, nC = Ops::Sets::nC;
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NOUT = NA * NB * NC * ND;
    constexpr int LT2 = LA + LB + LC + LD + 2;
    const R X0 = R(34.98683665524972);
    const int n = a.n;
    const int lane = threadIdx.x & 63;
    const long long wave0 = __builtin_amdgcn_readfirstlane(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    unsigned long long nsurv = 0, nprimq = 0;
    extern __shared__ unsigned long long s_grad[];   // 3 x natom when a.lds_grad
    __shared__ double s_boys[f64 ? kBoysFitRegions * kBoysFitTerms : 1];
    __shared__ float s_btab[f64 ? 1 : kBoysTabRows * kBoysLdsCols];
    __shared__ R s_bra[4][9 * 64];
    R* bra = s_bra[threadIdx.x >> 6];
    const bool use_lds = a.lds_grad != 0;
    const auto& PP = grad_prims<R>(a);
    const R* ea = grad_ea<R>(a);
    const R* Dp = grad_density<R>(a);
    const R rcut = R(a.cutoff), hyb = R(a.hybrid);
    if constexpr (f64)
        for (int k = threadIdx.x; k < kBoysFitRegions * kBoysFitTerms; k += blockDim.x) s_boys[k] = kBoysFit[k];
    else
        stage_boys_f32<LT2>(a.boys_tab, s_btab);
    if (use_lds)
        for (int k = threadIdx.x; k < 3 * a.natom; k += blockDim.x) s_grad[k] = 0ull;
    __syncthreads();

    // compacted work (w.steps set): the steps written by compact_kernel in the
    // tiles' places, each a bra pair and up to 64 surviving kets in w.list;
    // otherwise tiles of a bra pair and consecutive kets
    const bool compact = w.steps != nullptr;
    for (long long t = wave0; t < w.ntiles; t += nwaves) {
        const Tile tl = compact ? w.steps[t] : w.tiles[t];
        if (tl.len == 0) continue;   // wavefront-uniform: no survivor in this slot
        const int p = __builtin_amdgcn_readfirstlane(tl.p);
        const bool inrange = lane < tl.len;
        const int q = compact ? (inrange ? w.list[tl.q0 + lane] : 0) : tl.q0 + lane;
        const int qa = a.pr.qa[p], qb = a.pr.qb[p];
        const int qc = inrange ? a.pr.qa[q] : 0, qd = inrange ? a.pr.qb[q] : 0;
        const int atA = a.q.atom[qa], atB = a.q.atom[qb], atC = a.q.atom[qc], atD = a.q.atom[qd];
        bool active = false;
        double dn = 0.0;
        if (inrange && !(atA == atB && atA == atC && atA == atD)) {   // one-centre quartets vanish
            const int sA = a.q.shell[qa], sB = a.q.shell[qb], sC = a.q.shell[qc], sD = a.q.shell[qd];
            dn = dn_max(a.cutmat, a.nshell, sA, sB, sC, sD);
            const double ct = a.pr.Y[q] * a.pr.Y[p];
            active = ct > a.cutoff && ct * dn > a.cutoff;
        }
        const int pb0 = a.pr.prim0[p], pbs = a.pr.pstride[p], nbp = a.pr.nprim[p];
        stage_bra_grad<R>(a, pb0, pbs, nbp, lane, bra);
        if (!active) continue;
        ++nsurv;
        const int Mb = a.q.qfinal[qa] + a.q.qfinal[qb] + a.q.qfinal[qc] + a.q.qfinal[qd] + 2;
        const R Ax = R(a.xyz[3 * atA]), Ay = R(a.xyz[3 * atA + 1]), Az = R(a.xyz[3 * atA + 2]);
        const R Cx = R(a.xyz[3 * atC]), Cy = R(a.xyz[3 * atC + 1]), Cz = R(a.xyz[3 * atC + 2]);
        // density factors K = w (4 D_ab D_cd - h (D_ac D_bd + D_ad D_bc)), cons folded into D
        const int fa = a.q.f0[qa], fb = a.q.f0[qb], fc = a.q.f0[qc], fd = a.q.f0[qd];
        R wq = R(1);
        if (qa == qb) wq *= R(0.5);
        if (qc == qd) wq *= R(0.5);
        if (p == q) wq *= R(0.5);
        // formed (from the six density blocks, re-read from memory) just before
        // each pass's contraction, so that neither K nor D is live in the
        // register-heavy primitive loop
        auto form_K = [&](R* K) {
#pragma unroll
            for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                        for (int id = 0; id < ND; ++id) {
                            const long long ra = (long long)(fa + ia) * n, rb = (long long)(fb + ib) * n;
                            const R Dab = Dp[ra + fb + ib], Dcd = Dp[(long long)(fc + ic) * n + fd + id];
                            const R Dac = Dp[ra + fc + ic], Dbd = Dp[rb + fd + id];
                            const R Dad = Dp[ra + fd + id], Dbc = Dp[rb + fc + ic];
                            K[((ia * NB + ib) * NC + ic) * ND + id] =
                                wq * (R(4) * Dab * Dcd - hyb * (Dac * Dbd + Dad * Dbc));
                        }
        };
        R g[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) g[k] = R(0);
        const R Bx = R(a.xyz[3 * atB]), By = R(a.xyz[3 * atB + 1]), Bz = R(a.xyz[3 * atB + 2]);
        const R Dx = R(a.xyz[3 * atD]), Dy = R(a.xyz[3 * atD + 1]), Dz = R(a.xyz[3 * atD + 2]);
        const R rdn = R(dn);
        const int pk0 = a.pr.prim0[q], pks = a.pr.pstride[q], nkp = a.pr.nprim[q];
        // one pass over the primitive quartets per group of sets (the VRR is
        // recomputed; the sets of one pass are the only accumulators live)
        auto pass = [&](auto mtag) {
            constexpr int M = decltype(mtag)::value;
            R acc0[(M & 1) ? n0 : 1], accA[(M & 2) ? nA : 1], accB[(M & 4) ? nB : 1], accC[(M & 8) ? nC : 1];
#pragma unroll
            for (int k = 0; k < ((M & 1) ? n0 : 1); ++k) acc0[k] = R(0);
#pragma unroll
            for (int k = 0; k < ((M & 2) ? nA : 1); ++k) accA[k] = R(0);
#pragma unroll
            for (int k = 0; k < ((M & 4) ? nB : 1); ++k) accB[k] = R(0);
#pragma unroll
            for (int k = 0; k < ((M & 8) ? nC : 1); ++k) accC[k] = R(0);
            {
                const int ni = nbp;   // <= 64 primitive pairs per Qshell pair (checked on the host)
                for (int j = 0; j < nkp; ++j) {
                    const int kj = pk0 + j * pks;
                    const R zcd = PP.zeta[kj], oo2e = PP.oo2z[kj];
                    const R Qx = PP.Px[kj], Qy = PP.Py[kj], Qz = PP.Pz[kj];
                    const R Kcd = PP.K[kj], ccd = PP.cut[kj], wC = R(2) * ea[kj];
                    const R QCx = Qx - Cx, QCy = Qy - Cy, QCz = Qz - Cz;
                    for (int i = 0; i < ni; ++i) {
                        if (!((rdn * bra[6 * 64 + i]) * ccd > rcut)) continue;
                        if (M == Ops::Sets::kMask[0]) ++nprimq;
                        const R zab = bra[i], oo2z = bra[64 + i];
                        const R Px = bra[2 * 64 + i], Py = bra[3 * 64 + i], Pz = bra[4 * 64 + i];
                        const R rs = rsqrt_nr(zab + zcd);
                        const R inv = rs * rs;
                        const R roz = zcd * inv, roe = zab * inv;
                        const R ROU = zab * roz;
                        const R Wx = (Px * zab + Qx * zcd) * inv;
                        const R Wy = (Py * zab + Qy * zcd) * inv;
                        const R Wz = (Pz * zab + Qz * zcd) * inv;
                        const R dx = Px - Qx, dy = Py - Qy, dz = Pz - Qz;
                        const R T = (dx * dx + dy * dy + dz * dz) * ROU;
                        R F[LT2 + 1];
                        if constexpr (f64) boys_fit<LT2>(T, Mb, s_boys, F);
                        else boys_f32<LT2>(T, s_btab, F);
                        const R s = X0 * bra[5 * 64 + i] * Kcd * rs;
#pragma unroll
                        for (int m = 0; m <= LT2; ++m) F[m] *= s;
                        Ops::template vrg<M, R>(F, Px - Ax, Py - Ay, Pz - Az, Wx - Px, Wy - Py, Wz - Pz, QCx,
                                                QCy, QCz, Wx - Qx, Wy - Qy, Wz - Qz, oo2z, oo2e, R(0.5) * inv,
                                                roz, roe, R(2) * bra[7 * 64 + i], R(2) * bra[8 * 64 + i], wC,
                                                acc0, accA, accB, accC);
                    }
                }
            }
            R K[NOUT];
            form_K(K);
            Ops::template gcon<M, R>(acc0, accA, accB, accC, Ax - Bx, Ay - By, Az - Bz, Cx - Dx, Cy - Dy,
                                     Cz - Dz, K, g);
        };
        pass(std::integral_constant<int, Ops::Sets::kMask[0]>{});
        if constexpr (Ops::Sets::kPasses > 1) {
            pass(std::integral_constant<int, Ops::Sets::kMask[1]>{});
        }
        if constexpr (Ops::Sets::kPasses > 2) {
            pass(std::integral_constant<int, Ops::Sets::kMask[2]>{});
        }
        if constexpr (Ops::Sets::kPasses > 3) {
            pass(std::integral_constant<int, Ops::Sets::kMask[3]>{});
        }
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double gA = double(g[x]), gB = double(g[3 + x]), gC = double(g[6 + x]);
            gl_add(s_grad, a.G, use_lds, 3 * atA + x, gA);
            gl_add(s_grad, a.G, use_lds, 3 * atB + x, gB);
            gl_add(s_grad, a.G, use_lds, 3 * atC + x, gC);
            gl_add(s_grad, a.G, use_lds, 3 * atD + x, -(gA + gB + gC));
        }
    }
    if (nsurv) atomicAdd(a.counters + 0, nsurv);
    if (nprimq) atomicAdd(a.counters + 1, nprimq);
    if (use_lds) {
        __syncthreads();
        for (int k = threadIdx.x; k < 3 * a.natom; k += blockDim.x)
            if (s_grad[k] != 0ull) atomicAdd(a.G + k, s_grad[k]);
    }
}

}  // namespace qk
