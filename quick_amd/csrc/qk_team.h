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
    // V | acc | D blocks | products of AB, CD powers by power code (64 + 64) | F_m |
    // the primitive quartet's VRR factors (16)
    static constexpr int oAcc = kV, oD = oAcc + NACC, oPw = oD + kDbl, oF = oPw + 128, oGeo = oF + 16;
    static constexpr int total = oGeo + 16;
};

template <int LA, int LB, int LC, int LD, class TT, typename R = double>
constexpr unsigned team_lds_bytes() {
    return sizeof(R) * TeamLayout<LA, LB, LC, LD, TT>::total;
}

// The precision-dependent inputs of the cooperative kernel: primitive-pair
// data, density and Boys evaluation (QUICK's fit in FP64; the tabulated Taylor
// expansion in FP32, the mixed-precision path).
template <typename R>
struct TeamIn;
template <>
struct TeamIn<double> {
    const double *zeta, *oo2z, *Px, *Py, *Pz, *K, *cut, *Dp;
    __device__ explicit TeamIn(const FockArgs& a)
        : zeta(a.pp.zeta), oo2z(a.pp.oo2z), Px(a.pp.Px), Py(a.pp.Py), Pz(a.pp.Pz), K(a.pp.K), cut(a.pp.cut), Dp(a.Dp) {}
    template <int LT>
    __device__ __forceinline__ void boys(double T, int Mb, const double* fit, double* F) const {
        boys_fit<LT>(T, Mb, fit, F);
    }
};
template <>
struct TeamIn<float> {
    const float *zeta, *oo2z, *Px, *Py, *Pz, *K, *cut, *Dp;
    __device__ explicit TeamIn(const FockArgs& a)
        : zeta(a.ppf.zeta), oo2z(a.ppf.oo2z), Px(a.ppf.Px), Py(a.ppf.Py), Pz(a.ppf.Pz), K(a.ppf.K), cut(a.ppf.cut),
          Dp(a.Dpf), tab(a.boys_tab) {}
    const float* tab;
    template <int LT>
    __device__ __forceinline__ void boys(float T, int, const double*, float* F) const {
        boys_f32<LT, kBoysTabM>(T, tab + LT, F);
    }
};

// FP32: five wavefronts per SIMD (measured on (H2O)32: dpdp 1.31 -> 1.20 ms,
// ddds 0.86 -> 0.80 ms; asking the FP64 form for three made it slower)
template <int LA, int LB, int LC, int LD, class TT, typename R>
__global__ void __launch_bounds__(64, (sizeof(R) == 4 ? 5 : 1)) team_kernel(const FockArgs a, const ClassWork w) {
    using Lo = TeamLayout<LA, LB, LC, LD, TT>;
    constexpr int NA = Lo::NA, NB = Lo::NB, NC = Lo::NC, ND = Lo::ND;
    constexpr int NAB = Lo::NAB, NCD = Lo::NCD, NOUT = Lo::NOUT, NACC = Lo::NACC, LT = Lo::LT;
    constexpr int NF = TT::kNF;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    R* lds = reinterpret_cast<R*>(lds_raw);
    R* V = lds;
    R* hb = lds;                  // reuses V after the primitive loop
    R* outv = lds + NAB * NF;
    R* acc = lds + Lo::oAcc;
    R* Dcd = lds + Lo::oD;
    R* Dab = Dcd + NCD;
    R* Dbd = Dab + NAB;
    R* Dbc = Dbd + NB * ND;
    R* Dad = Dbc + NB * NC;
    R* Dac = Dad + NA * ND;
    // pwc[c] = AB_x^(c&3) AB_y^((c>>2)&3) AB_z^((c>>4)&3) for the power code c of a
    // bra HRR term; pwc[64 + c] the same in C-D for the ket HRR
    R* pwc = lds + Lo::oPw;
    R* Fl = lds + Lo::oF;
    // VRR factors indexed by side (0 bra, 1 ket) and direction: PA|QC (0-5),
    // WP|WQ (6-11), 1/2zeta|1/2eta (12-13), rho/zeta|rho/eta (14-15); read by
    // index, the per-node choice needs no branches
    R* geo = lds + Lo::oGeo;
    // QUICK's fit coefficients (the FP32 path reads its Boys table from memory)
    constexpr bool f64 = sizeof(R) == 8;
    __shared__ double s_boys[f64 ? kBoysFitRegions * kBoysFitTerms : 1];
    if constexpr (f64)
        for (int k = threadIdx.x; k < kBoysFitRegions * kBoysFitTerms; k += blockDim.x) s_boys[k] = kBoysFit[k];
    __syncthreads();
    const TeamIn<R> in(a);
    const R X0 = R(34.98683665524972);
    const R plim = R(a.prim_limit);
    const int lane = threadIdx.x;
    const int n = a.n;
    unsigned long long* const Xc = a.X + (long long)(blockIdx.x % a.xcopies) * n * n;
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
            const R Ax = R(a.xyz[3 * atA]), Ay = R(a.xyz[3 * atA + 1]), Az = R(a.xyz[3 * atA + 2]);
            const R Cx = R(a.xyz[3 * atC]), Cy = R(a.xyz[3 * atC + 1]), Cz = R(a.xyz[3 * atC + 2]);
            for (int k = lane; k < NACC; k += 64) acc[k] = R(0);
            if (lane == 0) V[0] = R(0);

            const int pb0 = a.pr.prim0[p], pbs = a.pr.pstride[p], nbp = a.pr.nprim[p];
            const int pk0 = a.pr.prim0[q], pks = a.pr.pstride[q], nkp = a.pr.nprim[q];
            const R dnr = R(dn);
            for (int i = 0; i < nbp; ++i) {
                const int bi = pb0 + i * pbs;
                const R zab = in.zeta[bi], oo2z = in.oo2z[bi];
                const R Px = in.Px[bi], Py = in.Py[bi], Pz = in.Pz[bi];
                const R X2 = X0 * in.K[bi];
                const R cp1 = dnr * in.cut[bi];
                for (int j = 0; j < nkp; ++j) {
                    const int kj = pk0 + j * pks;
                    if (!(cp1 * in.cut[kj] > plim)) continue;
                    ++nprimq;
                    const R zcd = in.zeta[kj], oo2e = in.oo2z[kj];
                    const R Qx = in.Px[kj], Qy = in.Py[kj], Qz = in.Pz[kj];
                    const R rs = rsqrt_nr(zab + zcd);
                    const R inv = rs * rs;
                    const R roz = zcd * inv, roe = zab * inv;
                    const R ROU = zab * roz;
                    const R Wx = (Px * zab + Qx * zcd) * inv;
                    const R Wy = (Py * zab + Qy * zcd) * inv;
                    const R Wz = (Pz * zab + Qz * zcd) * inv;
                    const R dx = Px - Qx, dy = Py - Qy, dz = Pz - Qz;
                    const R T = (dx * dx + dy * dy + dz * dz) * ROU;
                    R F[LT + 1];
                    in.template boys<LT>(T, Mb, s_boys, F);
                    const R s = X2 * in.K[kj] * rs;
                    R fsel = F[0] * s;
#pragma unroll
                    for (int m = 1; m <= LT; ++m) fsel = (lane == m) ? F[m] * s : fsel;
                    __syncthreads();                       // previous readers of V and Fl are done
                    if (lane <= LT) Fl[lane] = fsel;
                    const R PAx = Px - Ax, PAy = Py - Ay, PAz = Pz - Az, WPx = Wx - Px, WPy = Wy - Py, WPz = Wz - Pz;
                    const R QCx = Qx - Cx, QCy = Qy - Cy, QCz = Qz - Cz, WQx = Wx - Qx, WQy = Wy - Qy, WQz = Wz - Qz;
                    const R oo2ze = R(0.5) * inv;
                    if (lane == 0) {
                        geo[0] = PAx; geo[1] = PAy; geo[2] = PAz;
                        geo[3] = QCx; geo[4] = QCy; geo[5] = QCz;
                        geo[6] = WPx; geo[7] = WPy; geo[8] = WPz;
                        geo[9] = WQx; geo[10] = WQy; geo[11] = WQz;
                        geo[12] = oo2z; geo[13] = oo2e; geo[14] = roz; geo[15] = roe;
                    }
                    __syncthreads();
                    for (int k = lev[0] + lane; k < lev[1]; k += 64) {
                        const uint4 nd = nodes[k];
                        V[nd.x & 0xffffu] = Fl[(nd.w >> 8) & 0xffu];
                    }
                    for (int L = 1; L < TT::kLevels; ++L) {
                        __syncthreads();
                        const int k1 = lev[L + 1];
                        // the descriptors of up to four of the lane's nodes are
                        // loaded before any is evaluated (one memory latency)
                        for (int kb = lev[L] + lane; kb < k1; kb += 4 * 64) {
                          uint4 ndb[4];
#pragma unroll
                          for (int u = 0; u < 4; ++u) {
                              const int k = kb + u * 64;
                              ndb[u] = k < k1 ? nodes[k] : make_uint4(0u, 0u, 0u, 0u);
                          }
#pragma unroll
                          for (int u = 0; u < 4; ++u) {
                            if (kb + u * 64 >= k1) break;
                            const uint4 nd = ndb[u];
                            const unsigned fl = nd.w;
                            const int side = (fl & 3u) == 1u ? 0 : 1;
                            const int gi = side * 3 + int((fl >> 2) & 3u);
                            const R n1 = R((fl >> 4) & 15u), ne = R((fl >> 8) & 15u);
                            const R pa = geo[gi], wp = geo[6 + gi];
                            const R c3 = n1 * geo[12 + side];
                            const R c4 = geo[14 + side];
                            const R v = pa * V[nd.x >> 16] + wp * V[nd.y & 0xffffu] +
                                        c3 * (V[nd.y >> 16] - c4 * V[nd.z & 0xffffu]) + ne * oo2ze * V[nd.z >> 16];
                            V[nd.x & 0xffffu] = v;
                          }
                        }
                    }
                    __syncthreads();
                    for (int k = lane; k < NACC; k += 64) acc[k] += V[tgt[k]];
                }
            }
            __syncthreads();

            // HRR: products of powers of A-B and C-D by power code, density blocks
            for (int c = lane; c < 128; c += 64) {
                const int u = c < 64 ? atA : atC, v = c < 64 ? atB : atD;
                R px[3];
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    const R d = R(a.xyz[3 * u + ax]) - R(a.xyz[3 * v + ax]);
                    const int e = ((c & 63) >> (2 * ax)) & 3;
                    R r = R(1);
                    for (int k = 0; k < e; ++k) r *= d;
                    px[ax] = r;
                }
                pwc[c] = px[0] * px[1] * px[2];
            }
            const int fa = a.q.f0[qa], fb = a.q.f0[qb], fc = a.q.f0[qc], fd = a.q.f0[qd];
            const auto* Dp = in.Dp;
            for (int k = lane; k < NCD; k += 64) Dcd[k] = Dp[(long long)(fc + k / ND) * n + fd + k % ND];
            for (int k = lane; k < NAB; k += 64) Dab[k] = Dp[(long long)(fa + k / NB) * n + fb + k % NB];
            for (int k = lane; k < NB * ND; k += 64) Dbd[k] = Dp[(long long)(fb + k / ND) * n + fd + k % ND];
            for (int k = lane; k < NB * NC; k += 64) Dbc[k] = Dp[(long long)(fb + k / NC) * n + fc + k % NC];
            for (int k = lane; k < NA * ND; k += 64) Dad[k] = Dp[(long long)(fa + k / ND) * n + fd + k % ND];
            for (int k = lane; k < NA * NC; k += 64) Dac[k] = Dp[(long long)(fa + k / NC) * n + fc + k % NC];
            __syncthreads();
            // HRR tables: per output row four terms packed as (index | power code << 12 | binomial << 18)
            const int4* bt = reinterpret_cast<const int4*>(TT::bra());
            for (int h = lane; h < NAB * NF; h += 64) {
                const int ab = h / NF, f = h % NF;
                const int4 row = bt[ab];
                const int tt[4] = {row.x, row.y, row.z, row.w};
                R v = R(0);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int e = tt[u] & 0xfff, pw3 = (tt[u] >> 12) & 63, cf = (tt[u] >> 18) & 15;
                    v += R(cf) * pwc[pw3] * acc[e * NF + f];
                }
                hb[h] = v;
            }
            __syncthreads();
            const int4* kt = reinterpret_cast<const int4*>(TT::ket());
            for (int o = lane; o < NOUT; o += 64) {
                const int ab = o / NCD, cd = o % NCD;
                const int4 row = kt[cd];
                const int tt[4] = {row.x, row.y, row.z, row.w};
                R v = R(0);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int f = tt[u] & 0xfff, pw3 = (tt[u] >> 12) & 63, cf = (tt[u] >> 18) & 15;
                    v += R(cf) * pwc[64 + pw3] * hb[ab * NF + f];
                }
                outv[o] = v;
            }
            __syncthreads();

            // digestion (see fock_kernel): G = (X + X^T)/2
            R wgt = R(1);
            if (qa == qb) wgt *= R(0.5);
            if (qc == qd) wgt *= R(0.5);
            if (p == q) wgt *= R(0.5);
            const R jw = R(4) * wgt, kw = -R(a.hybrid) * wgt;
#define OUT(ia, ib, ic, id) outv[(((ia)*NB + (ib)) * NC + (ic)) * ND + (id)]
            constexpr int E1 = NAB, E2 = E1 + NCD, E3 = E2 + NA * NC, E4 = E3 + NA * ND, E5 = E4 + NB * NC,
                          E6 = E5 + NB * ND;
            const int eend = a.hybrid != 0.0 ? E6 : E2;
            for (int e = lane; e < eend; e += 64) {
                R s = R(0);
                long long idx;
                R wt;
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
                fx_add(Xc, idx, double(wt * s));
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
