// qk_grad.hip -- four-centre part of the nuclear gradient on MI355X (gfx950).
//
// dE2/dR at fixed density, E2 = 1/2 sum D G(D), over the same class- and
// bound-sorted Qshell pair lists as the Fock build.  Per surviving quartet
// (Y_ij Y_kl > cutoff and Y_ij Y_kl DNmax > cutoff, one-centre quartets
// skipped; gradient.f90:289-316) one block of four wavefronts:
//   1. runs the primitive loop: QUICK's Boys fit asked for order
//      sum(Qfinal) + 2 (shell.f90:1906-1912), the VRR raised by one on both
//      sides as a levelled node program in LDS, and accumulates four
//      contracted sets: S0 (weight 1), SA/SB/SC (weights 2 alpha, 2 beta,
//      2 gamma; classopt, shell.f90:2056-2114);
//   2. forms the density factor of every function quartet,
//      K = w (4 D_ab D_cd - h (D_ac D_bd + D_ad D_bc)), with the quartet's
//      symmetry weight w and exchange weight h;
//   3. for each derivative term (A+, A-, B+, B-, C+, C-) applies the bra and
//      ket HRR tables to the source set and contracts the result with K:
//      d/dA_x (ab|cd) = 2 alpha (a+1x b|cd) - a_x (a-1x b|cd), likewise for B
//      and C (hrrwholeopt, subs/hrr.f90:862-1014);
//   4. reduces the nine sums over the wavefront in a fixed order and adds
//      them, and D = -(A + B + C) (classopt, shell.f90:2463-2520), to the
//      int64 fixed-point gradient with integer atomics (reproducible).
// The class tables come from gen_grad.py (layout GradHdr in qk_internal.h).
#include "qk_kernels.h"

namespace qk {

// QUICK's FmT (subs/fmt.f90:12-84) with the number of values a runtime count
// nf <= kMaxF: the same branches and recursions as boys_ref, F_0..F_nf
// produced (loops unrolled to kMaxF so that F stays in registers).
constexpr int kMaxF = 10;
static_assert(kMaxF == kMaxGradF, "FP32 Boys orders of the gradient");
__device__ __forceinline__ void boys_ref_n(double X, int Mb, int nf, double* F) {
    const double E = exp(-X);
    const bool up = (X > 3.0e-5) || (X > 3.0e-7 && Mb < 4);
    if (up) {
        double G[1];
        boys_ref<0>(X, Mb, G);
        F[0] = G[0];
        const double XINV = 1.0 / X;
#pragma unroll
        for (int m = 1; m <= kMaxF; ++m) F[m] = (double(2 * m - 1) * F[m - 1] - E) * 0.5 * XINV;
    } else {
        double v = (1.0 - X) / double(2 * Mb + 1);
        const double twot = X * 2.0;
        for (int m = Mb - 1; m >= nf; --m) v = (twot * v + E) / double(m * 2 + 1);
        double cur = v;
#pragma unroll
        for (int m = kMaxF; m >= 0; --m) {
            if (m < nf) cur = (twot * cur + E) / double(m * 2 + 1);
            F[m] = cur;
        }
    }
}


__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void gx_add(unsigned long long* G, long long idx, double v) {
    const long long iv = __double2ll_rn(v * 0x1p44);
    if (iv != 0) atomicAdd(G + idx, (unsigned long long)iv);
}

// kGradTeam threads (one or four wavefronts) work on one quartet.  R = double:
// QUICK's Boys fit; R = float (mixed precision, weak quartets): the tabulated
// FP32 Boys expansion, FP32 LDS buffers and arithmetic, the nine sums promoted
// to FP64 before the wavefront reduction.
// At least three wavefronts per SIMD in FP64, five in FP32: the FP64 form then
// fits 168 VGPRs (76 bytes of spills) instead of 205 at two, the FP32 form 96
// (no spills) instead of 100 at four.  Measured on (H2O)32: FP64 gradient
// 174.5 -> 172.4 ms (dpdp 14.3 -> 11.0 ms); the mixed gradient's FP32 classes
// gain 0.3-0.9 ms each at five (dpdp 6.7 -> 5.7 ms), except dppp (16.0 ms at
// five), which runs a form built for four (kMinWaves 4).
template <int kGradTeam, typename R, int kMinWaves = (sizeof(R) == 4 ? 5 : 3)>
__global__ void __launch_bounds__(kGradTeam, kMinWaves) grad_kernel(const GradArgs a, const ClassWork w, const int* __restrict__ tab) {
    constexpr bool f64 = sizeof(R) == 8;
    extern __shared__ double lds[];
    const int lane = threadIdx.x;          // thread of the block (kGradTeam threads share one quartet)
    const int n = a.n;
    const int nSlots = tab[kGhSlots], nLevels = tab[kGhLevels], LT2 = tab[kGhLT] + 2;
    const int n0 = tab[kGhN0], nA = tab[kGhNA], nB = tab[kGhNB], nC = tab[kGhNC];
    const int nS = n0 + nA + nB + nC;
    const int NOUT = tab[kGhNOut], CB = tab[kGhCB], CC = tab[kGhCC], CD = tab[kGhCD];
    const int nStages = tab[kGhStages];
    const int RV = nSlots > tab[kGhMaxStage] ? nSlots : tab[kGhMaxStage];
    const uint4* nodes = reinterpret_cast<const uint4*>(tab + tab[kGhOffNodes]);
    const int* lev = tab + tab[kGhOffLev];
    const int* tgt = tab + tab[kGhOffTgt];
    double* gred = lds;           // per-wavefront partial sums of the nine derivatives (9 x 4)
    R* V = reinterpret_cast<R*>(lds + 36);   // VRR slots; after the primitive loop the stage buffers
    R* S = V + RV;                // S0 | SA | SB | SC
    R* Kv = S + nS;               // density factor per function quartet
    // pwc[c] = AB_x^(c&3) AB_y^((c>>2)&3) AB_z^((c>>4)&3) for the power code c of
    // a bra HRR term; pwc[64 + c] the same in C-D for the ket HRR
    R* pwc = Kv + NOUT;
    R* Fl = pwc + 128;
    // the primitive quartet's VRR factors, indexed by side (0 bra, 1 ket) and
    // direction: PA|QC (0-5), WP|WQ (6-11), 1/2zeta|1/2eta (12-13), rho/zeta|rho/eta (14-15)
    R* geo = Fl + 16;
    const auto& PP = grad_prims<R>(a);
    const R* ea = grad_ea<R>(a);
    const R* eb = grad_eb<R>(a);
    const R* Dp = grad_density<R>(a);
    const R X0 = R(34.98683665524972);
    const R rcut = R(a.cutoff), hyb = R(a.hybrid);
    unsigned long long nsurv = 0, nprimq = 0;

    for (long long t = blockIdx.x; t < w.ntiles; t += gridDim.x) {
        const Tile tl = w.tiles[t];
        const int p = tl.p;
        const int qa = a.pr.qa[p], qb = a.pr.qb[p];
        const int sA = a.q.shell[qa], sB = a.q.shell[qb];
        for (int qq = 0; qq < tl.len; ++qq) {
            const int q = tl.q0 + qq;
            const int qc = a.pr.qa[q], qd = a.pr.qb[q];
            const int atA = a.q.atom[qa], atB = a.q.atom[qb], atC = a.q.atom[qc], atD = a.q.atom[qd];
            if (atA == atB && atA == atC && atA == atD) continue;   // translational invariance
            const int sC = a.q.shell[qc], sD = a.q.shell[qd];
            const double ct = a.pr.Y[q] * a.pr.Y[p];
            const double dn = dn_max(a.cutmat, a.nshell, sA, sB, sC, sD);
            if (!(ct > a.cutoff && ct * dn > a.cutoff)) continue;   // uniform across the wavefront
            ++nsurv;
            const int Mb = a.q.qfinal[qa] + a.q.qfinal[qb] + a.q.qfinal[qc] + a.q.qfinal[qd] + 2;
            const R Ax = R(a.xyz[3 * atA]), Ay = R(a.xyz[3 * atA + 1]), Az = R(a.xyz[3 * atA + 2]);
            const R Cx = R(a.xyz[3 * atC]), Cy = R(a.xyz[3 * atC + 1]), Cz = R(a.xyz[3 * atC + 2]);
            __syncthreads();                       // previous quartet's readers of S/V are done
            for (int k = lane; k < nS; k += kGradTeam) S[k] = R(0);
            if (lane == 0) V[0] = R(0);

            const int pb0 = a.pr.prim0[p], pbs = a.pr.pstride[p], nbp = a.pr.nprim[p];
            const int pk0 = a.pr.prim0[q], pks = a.pr.pstride[q], nkp = a.pr.nprim[q];
            for (int i = 0; i < nbp; ++i) {
                const int bi = pb0 + i * pbs;
                const R zab = PP.zeta[bi], oo2z = PP.oo2z[bi];
                const R Px = PP.Px[bi], Py = PP.Py[bi], Pz = PP.Pz[bi];
                const R X2 = X0 * PP.K[bi];
                const R cp1 = R(dn) * PP.cut[bi];
                const R wA = R(2) * ea[bi], wB = R(2) * eb[bi];
                for (int j = 0; j < nkp; ++j) {
                    const int kj = pk0 + j * pks;
                    if (!(cp1 * PP.cut[kj] > rcut)) continue;
                    ++nprimq;
                    const R zcd = PP.zeta[kj], oo2e = PP.oo2z[kj];
                    const R wC = R(2) * ea[kj];
                    const R Qx = PP.Px[kj], Qy = PP.Py[kj], Qz = PP.Pz[kj];
                    const R rs = rsqrt_nr(zab + zcd);
                    const R inv = rs * rs;
                    const R roz = zcd * inv, roe = zab * inv;
                    const R ROU = zab * roz;
                    const R Wx = (Px * zab + Qx * zcd) * inv;
                    const R Wy = (Py * zab + Qy * zcd) * inv;
                    const R Wz = (Pz * zab + Qz * zcd) * inv;
                    const R dx = Px - Qx, dy = Py - Qy, dz = Pz - Qz;
                    const R T = (dx * dx + dy * dy + dz * dz) * ROU;
                    R F[kMaxF + 1];
                    if constexpr (f64) boys_ref_n(T, Mb, LT2, F);
                    else boys_f32_all(T, a.boys_tab, F);
                    const R s = X2 * PP.K[kj] * rs;
                    R fsel = F[0] * s;
#pragma unroll
                    for (int m = 1; m <= kMaxF; ++m) fsel = (lane == m) ? F[m] * s : fsel;
                    __syncthreads();
                    if (lane <= LT2) Fl[lane] = fsel;
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
                    for (int k = lev[0] + lane; k < lev[1]; k += kGradTeam) {
                        const uint4 nd = nodes[k];
                        V[nd.x & 0xffffu] = Fl[(nd.w >> 8) & 0xffu];
                    }
                    for (int L = 1; L < nLevels; ++L) {
                        __syncthreads();
                        const int k1 = lev[L + 1];
                        // the descriptors of up to four of the lane's nodes are
                        // loaded before any is evaluated (one memory latency)
                        for (int kb = lev[L] + lane; kb < k1; kb += 4 * kGradTeam) {
                            uint4 ndb[4];
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                const int k = kb + u * kGradTeam;
                                ndb[u] = k < k1 ? nodes[k] : make_uint4(0u, 0u, 0u, 0u);
                            }
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                if (kb + u * kGradTeam >= k1) break;
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
                    for (int kb = lane; kb < nS; kb += 4 * kGradTeam) {
                        int tg[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int k = kb + u * kGradTeam;
                            tg[u] = k < nS ? tgt[k] : 0;
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int k = kb + u * kGradTeam;
                            if (k >= nS) break;
                            const R wk = k < n0 ? R(1) : (k < n0 + nA ? wA : (k < n0 + nA + nB ? wB : wC));
                            S[k] += wk * V[tg[u]];
                        }
                    }
                }
            }
            __syncthreads();

            // products of powers of A-B and C-D by power code; density factors
            for (int c = lane; c < 128; c += kGradTeam) {
                const int u = c < 64 ? atA : atC, v = c < 64 ? atB : atD;
                R px[3];
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    const R d = R(a.xyz[3 * u + ax] - a.xyz[3 * v + ax]);
                    const int e = ((c & 63) >> (2 * ax)) & 3;
                    R r = R(1);
                    for (int k = 0; k < e; ++k) r *= d;
                    px[ax] = r;
                }
                pwc[c] = px[0] * px[1] * px[2];
            }
            {
                const int fa = a.q.f0[qa], fb = a.q.f0[qb], fc = a.q.f0[qc], fd = a.q.f0[qd];
                R wq = R(1);
                if (qa == qb) wq *= R(0.5);
                if (qc == qd) wq *= R(0.5);
                if (p == q) wq *= R(0.5);
                for (int o = lane; o < NOUT; o += kGradTeam) {
                    const int id = o % CD, ic = (o / CD) % CC, ib = (o / (CD * CC)) % CB, ia = o / (CD * CC * CB);
                    const long long ra = (long long)(fa + ia) * n, rb = (long long)(fb + ib) * n;
                    const R Dab = Dp[ra + fb + ib], Dcd = Dp[(long long)(fc + ic) * n + fd + id];
                    const R Dac = Dp[ra + fc + ic], Dbd = Dp[rb + fd + id];
                    const R Dad = Dp[ra + fd + id], Dbc = Dp[rb + fc + ic];
                    Kv[o] = wq * (R(4) * Dab * Dcd - hyb * (Dac * Dbd + Dad * Dbc));
                }
            }
            R g[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) g[k] = R(0);
            for (int st = 0; st < nStages; ++st) {
                const int* rec = tab + kGhSize + st * kGradStageFields;
                const int src = rec[1], nR1 = rec[2], nCol = rec[3], nR2 = rec[4];
                // direction; the terms a bra / ket table row has at most (the
                // remaining slots of a row are padding and are not read)
                const int dir = rec[0] & 0xff, nbt = (rec[0] >> 8) & 0xff, nkt = (rec[0] >> 16) & 0xff;
                // idx / nCol and idx / nR2 by a float reciprocal: exact for
                // idx < 2^12 and divisors < 64 (the rounding error stays far
                // below the 0.5 / divisor margin of idx + 0.5)
                const float rcol = 1.0f / float(nCol), rr2 = 1.0f / float(nR2);
                const int* bt = tab + rec[5];
                const int* kt = tab + rec[6];
                const int* con = tab + rec[7];
                const R* Ss = S + (src == 0 ? 0 : (src == 1 ? n0 : (src == 2 ? n0 + nA : n0 + nA + nB)));
                R* o1 = V;
                R* o2 = V + nR1 * nCol;
                __syncthreads();                   // pw, Kv, previous stage's readers
                for (int idx = lane; idx < nR1 * nCol; idx += kGradTeam) {
                    const int r1 = __float2int_rz((float(idx) + 0.5f) * rcol), col = idx - r1 * nCol;
                    const int4 t0 = reinterpret_cast<const int4*>(bt)[2 * r1];
                    const int4 t1 = nbt > 4 ? reinterpret_cast<const int4*>(bt)[2 * r1 + 1] : make_int4(0, 0, 0, 0);
                    const int tt[8] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w};
                    R v = R(0);
#pragma unroll
                    for (int u = 0; u < kGradHrrTerms; ++u) {
                        if (u >= nbt) break;
                        const int e = tt[u];
                        const int row = e & 0xfff, pc = (e >> 12) & 63, cf = (e >> 18) & 15;
                        v += R(cf) * pwc[pc] * Ss[row * nCol + col];
                    }
                    o1[idx] = v;
                }
                __syncthreads();
                for (int idx = lane; idx < nR1 * nR2; idx += kGradTeam) {
                    const int r1 = __float2int_rz((float(idx) + 0.5f) * rr2), r2 = idx - r1 * nR2;
                    const int4 t0 = reinterpret_cast<const int4*>(kt)[2 * r2];
                    const int4 t1 = nkt > 4 ? reinterpret_cast<const int4*>(kt)[2 * r2 + 1] : make_int4(0, 0, 0, 0);
                    const int tt[8] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w};
                    R v = R(0);
#pragma unroll
                    for (int u = 0; u < kGradHrrTerms; ++u) {
                        if (u >= nkt) break;
                        const int e = tt[u];
                        const int colk = e & 0xfff, pc = (e >> 12) & 63, cf = (e >> 18) & 15;
                        v += R(cf) * pwc[64 + pc] * o1[r1 * nCol + colk];
                    }
                    o2[idx] = v;
                }
                __syncthreads();
                for (int o = lane; o < NOUT; o += kGradTeam) {
                    const R K = Kv[o];
                    const int4 cr = reinterpret_cast<const int4*>(con)[o];
                    const int ce[3] = {cr.x, cr.y, cr.z};
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        const int ix = ce[x] & 0xffff;
                        if (ix != 0xffff) {
                            const R v = R(ce[x] >> 16) * K * o2[ix];
                            if (dir == 0) g[x] += v; else if (dir == 1) g[3 + x] += v; else g[6 + x] += v;
                        }
                    }
                }
            }
            // nine sums: over each wavefront (fixed order), then over the block's
            // wavefronts in LDS (fixed order): reproducible
            double gs[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) gs[k] = wave_sum_f64(double(g[k]));
            if ((lane & 63) == 0)
#pragma unroll
                for (int k = 0; k < 9; ++k) gred[(lane >> 6) * 9 + k] = gs[k];
            __syncthreads();
            if (lane < 3) {
                double gA = 0.0, gB = 0.0, gC = 0.0;
                for (int wv = 0; wv < kGradTeam / 64; ++wv) {
                    gA += gred[wv * 9 + lane];
                    gB += gred[wv * 9 + 3 + lane];
                    gC += gred[wv * 9 + 6 + lane];
                }
                gx_add(a.G, 3LL * atA + lane, gA);
                gx_add(a.G, 3LL * atB + lane, gB);
                gx_add(a.G, 3LL * atC + lane, gC);
                gx_add(a.G, 3LL * atD + lane, -(gA + gB + gC));
            }
        }
    }
    if (lane == 0) {
        if (nsurv) atomicAdd(a.counters + 0, nsurv);
        if (nprimq) atomicAdd(a.counters + 1, nprimq);
    }
}

// grad <- grad + G * 2^-44
// grad += G (2^-44 units); a sum within a factor of two of the int64 range
// (|dE/dR| of 2^18 Eh/bohr and more) raises the range flag
__global__ void grad_finalize(const unsigned long long* __restrict__ G, double* __restrict__ grad, int m,
                              unsigned long long* __restrict__ flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const long long s = (long long)G[k];
    if (s >= (1LL << 62) || s <= -(1LL << 62)) atomicOr(flag, 1ull);
    grad[k] += (double)s * 0x1p-44;
}

// four wavefronts per quartet for the classes with the largest node programs
// (dpdp, dddp, dddd); measured on (H2O)32, ddds and ddpp run as fast or faster
// with one (ddds FP64 9.4 -> 7.5 ms) and dppp-sized classes twice as fast
static int grad_team_of(const int* hdr) { return hdr[kGhSlots] >= 1800 ? 256 : 64; }

// dynamic LDS of grad_kernel: the FP64 partial sums (9 x 4), then the VRR /
// stage buffers, sets, density factors, powers, Boys values and the VRR factors in the
// kernel's precision
size_t grad_lds_bytes(const int* hdr, bool fp32) {
    const int RV = hdr[kGhSlots] > hdr[kGhMaxStage] ? hdr[kGhSlots] : hdr[kGhMaxStage];
    const int nS = hdr[kGhN0] + hdr[kGhNA] + hdr[kGhNB] + hdr[kGhNC];
    return sizeof(double) * 36 + (fp32 ? sizeof(float) : sizeof(double)) * (size_t)(RV + nS + hdr[kGhNOut] + 128 + 16 + 16);
}

template <int TEAM, typename R, int W = (sizeof(R) == 4 ? 5 : 3)>
static int launch_team(const GradArgs& a, const ClassWork& w, const int* dtab, size_t lds, int grid, hipStream_t s) {
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&grad_kernel<TEAM, R, W>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return -1;
    static const long long cap =
        scratch_grid_cap(kernel_scratch(reinterpret_cast<const void*>(&grad_kernel<TEAM, R, W>)), TEAM / 64);
    hipLaunchKernelGGL((grad_kernel<TEAM, R, W>), dim3((int)std::min<long long>(grid, cap)), dim3(TEAM), lds, s, a, w,
                       dtab);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// hdr: the class's table header (host copy); dtab: the same table on the device
// four_waves (FP32, one-wavefront teams): the form built for four wavefronts
// per SIMD (100 VGPRs) instead of five (96), for a class that runs slower at five
int launch_grad_class(const GradArgs& a, const ClassWork& w, const int* hdr, const int* dtab, size_t lds, int grid,
                      bool fp32, bool four_waves, hipStream_t s) {
    if (fp32 && four_waves && grad_team_of(hdr) == 64) return launch_team<64, float, 4>(a, w, dtab, lds, grid, s);
    if (fp32)
        return grad_team_of(hdr) == 256 ? launch_team<256, float>(a, w, dtab, lds, grid, s)
                                        : launch_team<64, float>(a, w, dtab, lds, grid, s);
    return grad_team_of(hdr) == 256 ? launch_team<256, double>(a, w, dtab, lds, grid, s)
                                    : launch_team<64, double>(a, w, dtab, lds, grid, s);
}

#define QK_DECL(T)                                                                                 \
    void launch_grad_light_##T(const GradArgs&, const ClassWork&, int, size_t, hipStream_t);      \
    void launch_grad_light_f32_##T(const GradArgs&, const ClassWork&, int, size_t, hipStream_t);
QK_DECL(ssss) QK_DECL(psss) QK_DECL(psps) QK_DECL(ppss) QK_DECL(dsss) QK_DECL(ppps) QK_DECL(dsps)
QK_DECL(dpss) QK_DECL(pppp) QK_DECL(dspp) QK_DECL(dsds) QK_DECL(dpps) QK_DECL(ddss)
#undef QK_DECL
using GradLightLauncher = void (*)(const GradArgs&, const ClassWork&, int, size_t, hipStream_t);
// classes with a one-quartet-per-lane kernel (order of gen_eri.quartet_classes())
static const GradLightLauncher kGradLight[kNumQuartetClasses] = {
    launch_grad_light_ssss, launch_grad_light_psss, launch_grad_light_psps, launch_grad_light_ppss,
    launch_grad_light_ppps, launch_grad_light_pppp, launch_grad_light_dsss, launch_grad_light_dsps,
    launch_grad_light_dspp, launch_grad_light_dsds, launch_grad_light_dpss, launch_grad_light_dpps, nullptr, nullptr,
    nullptr, launch_grad_light_ddss, nullptr, nullptr, nullptr, nullptr, nullptr};
static const GradLightLauncher kGradLightF32[kNumQuartetClasses] = {
    launch_grad_light_f32_ssss, launch_grad_light_f32_psss, launch_grad_light_f32_psps, launch_grad_light_f32_ppss,
    launch_grad_light_f32_ppps, launch_grad_light_f32_pppp, launch_grad_light_f32_dsss, launch_grad_light_f32_dsps,
    launch_grad_light_f32_dspp, launch_grad_light_f32_dsds, launch_grad_light_f32_dpss, launch_grad_light_f32_dpps,
    nullptr, nullptr, nullptr, launch_grad_light_f32_ddss, nullptr, nullptr, nullptr, nullptr, nullptr};
bool grad_class_is_light(int k) { return kGradLight[k] != nullptr; }
int launch_grad_light(int k, const GradArgs& a, const ClassWork& w, int grid, bool fp32, hipStream_t s) {
    const size_t lds = a.lds_grad ? sizeof(unsigned long long) * 3 * (size_t)a.natom : 0;
    (fp32 ? kGradLightF32 : kGradLight)[k](a, w, grid, lds, s);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

void launch_grad_finalize(const unsigned long long* G, double* grad, int m, unsigned long long* flag, hipStream_t s) {
    hipLaunchKernelGGL(grad_finalize, dim3((m + 255) / 256), dim3(256), 0, s, G, grad, m, flag);
}

}  // namespace qk
