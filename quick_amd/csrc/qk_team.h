// qk_team.h -- wavefront-cooperative kernel for the heavy d classes.
//
// For classes whose thread-per-quartet form needs thousands of intermediates
// per lane (e.g. (dd|dd): 961 contracted [e0|f0] values, 1296 integrals and a
// VRR working set of ~2300 values), one wavefront evaluates one quartet at a
// time with every intermediate in LDS:
//   * the VRR is a generated node program (team_tables.inc) executed level by
//     level -- the nodes of one level are independent, the lanes stride
//     through them; level 0 holds F_m(T) (QUICK's Boys fit), and each node is
//         v = c1 V[a] + c2 V[b] + c3 (V[c] - c4 V[d]) + c5 V[x]
//     with (c1..c5) the Obara-Saika coefficients of its direction and kind and
//     absent operands pointing at the zero slot;
//   * targets (|e| in [la,la+lb], |f| in [lc,lc+ld], m = 0) are accumulated
//     over the primitive quartets into acc;
//   * the HRR is applied in two table-driven passes (bra then ket), each entry
//     a binomial sum of <= 4 terms with (A-B) or (C-D) powers;
//   * the six J/K digestion blocks are element-parallel: a lane owns a block
//     element and sums over the two free indices, so every sum has a fixed order
//     and the fixed-point result is reproducible.
#pragma once
#include "qk_kernels.h"

namespace qk {

template <int LA, int LB, int LC, int LD, class TT>
struct TeamLayout {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NAB = NA * NB, NCD = NC * ND, NOUT = NAB * NCD;
    static constexpr int NACC = TT::kNE * TT::kNF;
    static constexpr int LT = LA + LB + LC + LD;
    static constexpr int kV = TT::kSlots > NAB * TT::kNF + NOUT ? TT::kSlots : NAB * TT::kNF + NOUT;
    static constexpr int kDbl = NCD + NAB + NB * ND + NB * NC + NA * ND + NA * NC;
    // V | acc | D blocks | powers of AB, CD (3 axes x 3) | F_m
    static constexpr int oAcc = kV, oD = oAcc + NACC, oPw = oD + kDbl, oF = oPw + 18;
    static constexpr int total = oF + 16;
};

template <int LA, int LB, int LC, int LD, class TT>
constexpr unsigned team_lds_bytes() {
    return sizeof(double) * TeamLayout<LA, LB, LC, LD, TT>::total;
}

template <int LA, int LB, int LC, int LD, class TT>
__global__ void __launch_bounds__(64) team_kernel(const FockArgs a, const ClassWork w) {
    using Lo = TeamLayout<LA, LB, LC, LD, TT>;
    constexpr int NA = Lo::NA, NB = Lo::NB, NC = Lo::NC, ND = Lo::ND;
    constexpr int NAB = Lo::NAB, NCD = Lo::NCD, NOUT = Lo::NOUT, NACC = Lo::NACC, LT = Lo::LT;
    constexpr int NF = TT::kNF;
    extern __shared__ double lds[];
    double* V = lds;
    double* hb = lds;                  // reuses V after the primitive loop
    double* outv = lds + NAB * NF;
    double* acc = lds + Lo::oAcc;
    double* Dcd = lds + Lo::oD;
    double* Dab = Dcd + NCD;
    double* Dbd = Dab + NAB;
    double* Dbc = Dbd + NB * ND;
    double* Dad = Dbc + NB * NC;
    double* Dac = Dad + NA * ND;
    double* pw = lds + Lo::oPw;        // pw[axis*3 + k] = AB_axis^k, pw[9 + axis*3 + k] = CD_axis^k
    double* Fl = lds + Lo::oF;
    const double X0 = 34.98683665524972;
    const int lane = threadIdx.x;
    const int n = a.n;
    unsigned long long* const Xc = a.X + (long long)(blockIdx.x % kXCopies) * n * n;
    const uint4* nodes = TT::nodes();
    const int* lev = TT::levels();
    const unsigned short* tgt = TT::targets();
    unsigned long long nsurv = 0, nprimq = 0;

    for (long long t = blockIdx.x; t < w.ntiles; t += gridDim.x) {
        const Tile tl = w.tiles[t];
        const int p = tl.p;
        const int qa = a.pr.qa[p], qb = a.pr.qb[p];
        const int sA = a.q.shell[qa], sB = a.q.shell[qb];
        for (int qq = 0; qq < tl.len; ++qq) {
            const int q = tl.q0 + qq;
            const int qc = a.pr.qa[q], qd = a.pr.qb[q];
            const int sC = a.q.shell[qc], sD = a.q.shell[qd];
            const double dn = dn_max(a.cutmat, a.nshell, sA, sB, sC, sD);
            if (!(a.pr.Y[q] * a.pr.Y[p] * dn > a.tau)) continue;   // uniform across the wavefront
            ++nsurv;
            const int Mb = a.boys_qshell_order ? LT : a.q.qfinal[qa] + a.q.qfinal[qb] + a.q.qfinal[qc] + a.q.qfinal[qd];
            const int atA = a.q.atom[qa], atB = a.q.atom[qb], atC = a.q.atom[qc], atD = a.q.atom[qd];
            const double Ax = a.xyz[3 * atA], Ay = a.xyz[3 * atA + 1], Az = a.xyz[3 * atA + 2];
            const double Cx = a.xyz[3 * atC], Cy = a.xyz[3 * atC + 1], Cz = a.xyz[3 * atC + 2];
            for (int k = lane; k < NACC; k += 64) acc[k] = 0.0;
            if (lane == 0) V[0] = 0.0;

            const int pb0 = a.pr.prim0[p], pbs = a.pr.pstride[p], nbp = a.pr.nprim[p];
            const int pk0 = a.pr.prim0[q], pks = a.pr.pstride[q], nkp = a.pr.nprim[q];
            for (int i = 0; i < nbp; ++i) {
                const int bi = pb0 + i * pbs;
                const double zab = a.pp.zeta[bi], oo2z = a.pp.oo2z[bi];
                const double Px = a.pp.Px[bi], Py = a.pp.Py[bi], Pz = a.pp.Pz[bi];
                const double X2 = X0 * a.pp.K[bi];
                const double cp1 = dn * a.pp.cut[bi];
                for (int j = 0; j < nkp; ++j) {
                    const int kj = pk0 + j * pks;
                    if (!(cp1 * a.pp.cut[kj] > a.prim_limit)) continue;
                    ++nprimq;
                    const double zcd = a.pp.zeta[kj], oo2e = a.pp.oo2z[kj];
                    const double Qx = a.pp.Px[kj], Qy = a.pp.Py[kj], Qz = a.pp.Pz[kj];
                    const double inv = 1.0 / (zab + zcd);
                    const double roz = zcd * inv, roe = zab * inv;
                    const double ROU = zab * roz;
                    const double Wx = (Px * zab + Qx * zcd) * inv;
                    const double Wy = (Py * zab + Qy * zcd) * inv;
                    const double Wz = (Pz * zab + Qz * zcd) * inv;
                    const double dx = Px - Qx, dy = Py - Qy, dz = Pz - Qz;
                    const double T = (dx * dx + dy * dy + dz * dz) * ROU;
                    double F[LT + 1];
                    boys_ref<LT>(T, Mb, F);
                    const double s = X2 * a.pp.K[kj] * sqrt(inv);
                    double fsel = F[0] * s;
#pragma unroll
                    for (int m = 1; m <= LT; ++m) fsel = (lane == m) ? F[m] * s : fsel;
                    __syncthreads();                       // previous readers of V and Fl are done
                    if (lane <= LT) Fl[lane] = fsel;
                    const double PA[3] = {Px - Ax, Py - Ay, Pz - Az}, WP[3] = {Wx - Px, Wy - Py, Wz - Pz};
                    const double QC[3] = {Qx - Cx, Qy - Cy, Qz - Cz}, WQ[3] = {Wx - Qx, Wy - Qy, Wz - Qz};
                    const double oo2ze = 0.5 * inv;
                    __syncthreads();
                    for (int k = lev[0] + lane; k < lev[1]; k += 64) {
                        const uint4 nd = nodes[k];
                        V[nd.x & 0xffffu] = Fl[(nd.w >> 8) & 0xffu];
                    }
                    for (int L = 1; L < TT::kLevels; ++L) {
                        __syncthreads();
                        for (int k = lev[L] + lane; k < lev[L + 1]; k += 64) {
                            const uint4 nd = nodes[k];
                            const unsigned fl = nd.w;
                            const bool bra = (fl & 3u) == 1u;
                            const int dir = (fl >> 2) & 3;
                            const double n1 = double((fl >> 4) & 15u), ne = double((fl >> 8) & 15u);
                            const double pa = dir == 0 ? (bra ? PA[0] : QC[0]) : (dir == 1 ? (bra ? PA[1] : QC[1]) : (bra ? PA[2] : QC[2]));
                            const double wp = dir == 0 ? (bra ? WP[0] : WQ[0]) : (dir == 1 ? (bra ? WP[1] : WQ[1]) : (bra ? WP[2] : WQ[2]));
                            const double c3 = n1 * (bra ? oo2z : oo2e);
                            const double c4 = bra ? roz : roe;
                            const double v = pa * V[nd.x >> 16] + wp * V[nd.y & 0xffffu] +
                                             c3 * (V[nd.y >> 16] - c4 * V[nd.z & 0xffffu]) + ne * oo2ze * V[nd.z >> 16];
                            V[nd.x & 0xffffu] = v;
                        }
                    }
                    __syncthreads();
                    for (int k = lane; k < NACC; k += 64) acc[k] += V[tgt[k]];
                }
            }
            __syncthreads();

            // HRR: powers of A-B and C-D, density blocks
            if (lane < 18) {
                const double Bx = a.xyz[3 * atB], By = a.xyz[3 * atB + 1], Bz = a.xyz[3 * atB + 2];
                const double Dx = a.xyz[3 * atD], Dy = a.xyz[3 * atD + 1], Dz = a.xyz[3 * atD + 2];
                const int grp = lane / 9, axis = (lane % 9) / 3, k = lane % 3;
                const double d = grp == 0 ? (axis == 0 ? Ax - Bx : (axis == 1 ? Ay - By : Az - Bz))
                                          : (axis == 0 ? Cx - Dx : (axis == 1 ? Cy - Dy : Cz - Dz));
                pw[lane] = k == 0 ? 1.0 : (k == 1 ? d : d * d);
            }
            const int fa = a.q.f0[qa], fb = a.q.f0[qb], fc = a.q.f0[qc], fd = a.q.f0[qd];
            const double* Dp = a.Dp;
            for (int k = lane; k < NCD; k += 64) Dcd[k] = Dp[(long long)(fc + k / ND) * n + fd + k % ND];
            for (int k = lane; k < NAB; k += 64) Dab[k] = Dp[(long long)(fa + k / NB) * n + fb + k % NB];
            for (int k = lane; k < NB * ND; k += 64) Dbd[k] = Dp[(long long)(fb + k / ND) * n + fd + k % ND];
            for (int k = lane; k < NB * NC; k += 64) Dbc[k] = Dp[(long long)(fb + k / NC) * n + fc + k % NC];
            for (int k = lane; k < NA * ND; k += 64) Dad[k] = Dp[(long long)(fa + k / ND) * n + fd + k % ND];
            for (int k = lane; k < NA * NC; k += 64) Dac[k] = Dp[(long long)(fa + k / NC) * n + fc + k % NC];
            __syncthreads();
            const int* bt = TT::bra();
            for (int h = lane; h < NAB * NF; h += 64) {
                const int ab = h / NF, f = h % NF;
                double v = 0.0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int e = bt[ab * 12 + 3 * u], pw3 = bt[ab * 12 + 3 * u + 1], cf = bt[ab * 12 + 3 * u + 2];
                    const double c = double(cf) * pw[pw3 & 3] * pw[3 + ((pw3 >> 2) & 3)] * pw[6 + ((pw3 >> 4) & 3)];
                    v += c * acc[e * NF + f];
                }
                hb[h] = v;
            }
            __syncthreads();
            const int* kt = TT::ket();
            for (int o = lane; o < NOUT; o += 64) {
                const int ab = o / NCD, cd = o % NCD;
                double v = 0.0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int f = kt[cd * 12 + 3 * u], pw3 = kt[cd * 12 + 3 * u + 1], cf = kt[cd * 12 + 3 * u + 2];
                    const double c = double(cf) * pw[9 + (pw3 & 3)] * pw[12 + ((pw3 >> 2) & 3)] * pw[15 + ((pw3 >> 4) & 3)];
                    v += c * hb[ab * NF + f];
                }
                outv[o] = v;
            }
            __syncthreads();

            // digestion (see fock_kernel): G = (X + X^T)/2
            double wgt = 1.0;
            if (qa == qb) wgt *= 0.5;
            if (qc == qd) wgt *= 0.5;
            if (p == q) wgt *= 0.5;
            const double jw = 4.0 * wgt, kw = -a.hybrid * wgt;
#define OUT(ia, ib, ic, id) outv[(((ia)*NB + (ib)) * NC + (ic)) * ND + (id)]
            constexpr int E1 = NAB, E2 = E1 + NCD, E3 = E2 + NA * NC, E4 = E3 + NA * ND, E5 = E4 + NB * NC,
                          E6 = E5 + NB * ND;
            const int eend = a.hybrid != 0.0 ? E6 : E2;
            for (int e = lane; e < eend; e += 64) {
                double s = 0.0;
                long long idx;
                double wt;
                if (e < E1) {
                    const int ia = e / NB, ib = e % NB;
                    for (int ic = 0; ic < NC; ++ic)
                        for (int id = 0; id < ND; ++id) s += Dcd[ic * ND + id] * OUT(ia, ib, ic, id);
                    idx = (long long)(fa + ia) * n + fb + ib;
                    wt = jw;
                } else if (e < E2) {
                    const int ic = (e - E1) / ND, id = (e - E1) % ND;
                    for (int ia = 0; ia < NA; ++ia)
                        for (int ib = 0; ib < NB; ++ib) s += Dab[ia * NB + ib] * OUT(ia, ib, ic, id);
                    idx = (long long)(fc + ic) * n + fd + id;
                    wt = jw;
                } else if (e < E3) {
                    const int ia = (e - E2) / NC, ic = (e - E2) % NC;
                    for (int ib = 0; ib < NB; ++ib)
                        for (int id = 0; id < ND; ++id) s += Dbd[ib * ND + id] * OUT(ia, ib, ic, id);
                    idx = (long long)(fa + ia) * n + fc + ic;
                    wt = kw;
                } else if (e < E4) {
                    const int ia = (e - E3) / ND, id = (e - E3) % ND;
                    for (int ib = 0; ib < NB; ++ib)
                        for (int ic = 0; ic < NC; ++ic) s += Dbc[ib * NC + ic] * OUT(ia, ib, ic, id);
                    idx = (long long)(fa + ia) * n + fd + id;
                    wt = kw;
                } else if (e < E5) {
                    const int ib = (e - E4) / NC, ic = (e - E4) % NC;
                    for (int ia = 0; ia < NA; ++ia)
                        for (int id = 0; id < ND; ++id) s += Dad[ia * ND + id] * OUT(ia, ib, ic, id);
                    idx = (long long)(fb + ib) * n + fc + ic;
                    wt = kw;
                } else {
                    const int ib = (e - E5) / ND, id = (e - E5) % ND;
                    for (int ia = 0; ia < NA; ++ia)
                        for (int ic = 0; ic < NC; ++ic) s += Dac[ia * NC + ic] * OUT(ia, ib, ic, id);
                    idx = (long long)(fb + ib) * n + fd + id;
                    wt = kw;
                }
                fx_add(Xc, idx, wt * s);
            }
#undef OUT
            __syncthreads();   // LDS is reused by the next quartet
        }
    }
    if (lane == 0) {
        if (nsurv) atomicAdd(a.counters + 0, nsurv);
        if (nprimq) atomicAdd(a.counters + 1, nprimq);
    }
}

}  // namespace qk
