// qk_xc.hip -- closed-shell B3LYP exchange-correlation on the device (gfx950).
//
// Replaces QUICK's XC step (dftoperator's B3LYP branch, dft.f90:894-1018, and
// the GPU getxc_kernel, gpu_getxc.cu:92-267), which walks one grid point per
// thread and adds V_mn for every basis pair with an atomic.  Here the work is
// shaped for the matrix cores:
//
//   per geometry (qk_xc_build_grid / qk_xc_set_grid)
//     * grid points: radial shells x Lebedev rules (SG-1 with its angular
//       pruning, or a product grid), Stratmann-Scuseria-Frisch cell weights
//       (subs/ssw.f90), points whose weight is below the cut or that lie
//       outside every basis function's sphere of significance dropped (they
//       contribute exactly nothing), kept in generation order;
//     * the points sorted along a Morton curve and cut into batches of 256,
//       each batch with the list of functions whose sphere of significance
//       meets its bounding sphere (m functions);
//   per density (qk_xc_device), batch by batch
//     * Phi (256 x m) and its gradient: basis values (pteval);
//     * T = Phi D_sub on FP64 MFMA (v_mfma_f64_16x16x4), fused with the row
//       dots rho_p = sum_j T_pj Phi_pj and grad rho_p = 2 sum_j T_pj grad Phi_pj;
//     * per point: B3LYP f, df/drho, df/dsigma; Z = w (df/drho Phi / 2 +
//       2 df/dsigma grad rho . grad Phi);
//     * V_sub = Phi^T Z + Z^T Phi on FP64 MFMA (lower-triangle tiles only),
//       added to the int64 fixed-point accumulator V (2^-48 units) with integer
//       atomics, so the result is bitwise reproducible;
//     * E_xc = sum w f and the electron count in a fixed-order reduction.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "quick_mi355x.h"
#include "qk_internal.h"

#define QK_HD __host__ __device__
#include "xc_b3lyp.inc"
#include "lebedev.inc"

namespace qk {

constexpr int kXcMaxPrim = 10;     // QUICK's MAXPRIM
constexpr int kXcBatch = 256;      // grid points per batch
constexpr int kXcTile = 64;        // GEMM tile (rows and columns)
constexpr int kXcShellThreads = 320;
constexpr double kXcFix = 0x1p48;  // V accumulator units

typedef double v4d __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// grid

struct XcAtoms {
    const double* xyz;     // 3 x natom
    const double* reach2;  // per atom: largest sigrad2 of its functions
    int natom;
};

// Stratmann-Scuseria-Frisch cell weight of point g for its parent atom ip:
// P_i = prod_{j != i} s(mu_ij), s(mu) = (1 - g(mu / a)) / 2 with a = 0.64 and
// g the SSF polynomial, 0 beyond mu >= a; weight = P_ip / sum_i P_i
// (subs/ssw.f90).
__device__ double ssf_weight(double gx, double gy, double gz, int ip, const XcAtoms at) {
    const int na = at.natom;
    if (na == 1) return 1.0;
    auto dist = [&](int i) {
        const double dx = gx - at.xyz[3 * i], dy = gy - at.xyz[3 * i + 1], dz = gz - at.xyz[3 * i + 2];
        return sqrt(dx * dx + dy * dy + dz * dz);
    };
    auto cellP = [&](int i, double ri, int first) {
        double w = 1.0;
        for (int jj = -1; jj < na && w != 0.0; ++jj) {
            const int j = jj < 0 ? first : jj;   // the most likely zero first
            if (j == i || (jj >= 0 && j == first)) continue;
            const double rj = dist(j);
            const double dx = at.xyz[3 * i] - at.xyz[3 * j], dy = at.xyz[3 * i + 1] - at.xyz[3 * j + 1],
                         dz = at.xyz[3 * i + 2] - at.xyz[3 * j + 2];
            const double mu = (ri - rj) / sqrt(dx * dx + dy * dy + dz * dz);
            if (mu >= 0.64) w = 0.0;
            else if (mu >= -0.64) {
                // the polynomial in the monomial form of subs/ssw.f90: 1 - g cancels
                // near |mu| = a, so the form decides the last digits of such weights
                const double f = mu / 0.64, f3 = f * f * f, f5 = f3 * f * f, f7 = f5 * f * f;
                const double g = (35.0 * f - 35.0 * f3 + 21.0 * f5 - 5.0 * f7) / 16.0;
                w *= 0.5 * (1.0 - g);
            }
        }
        return w;
    };
    const double rp = dist(ip);
    const double wp = cellP(ip, rp, ip == 0 ? 1 : 0);
    if (wp == 0.0) return 0.0;
    double total = wp;
    for (int i = 0; i < na; ++i)
        if (i != ip) total += cellP(i, dist(i), ip);
    return wp / total;
}

// One block per radial shell (atom, radius): the shell's Lebedev points, their
// weight rw * w_ang * SSF, and a keep flag (weight >= cut and inside some
// function's sphere of significance).  raw: 4 doubles per slot (x, y, z, w with
// w = 0 for dropped points); count: kept points of the shell.
__global__ void __launch_bounds__(kXcShellThreads) xc_grid_shells(const int* __restrict__ sh_atom,
                                                                  const int* __restrict__ sh_rule,
                                                                  const double* __restrict__ sh_r,
                                                                  const double* __restrict__ sh_rw, const XcAtoms at,
                                                                  double cut, double* __restrict__ raw,
                                                                  int* __restrict__ count) {
    const int s = blockIdx.x, j = threadIdx.x;
    const int a = sh_atom[s], rule = sh_rule[s];
    bool keep = false;
    double px = 0, py = 0, pz = 0, w = 0;
    if (j < kLebSize[rule]) {
        const double* q = kLebedev + 4 * (kLebOffset[rule] + j);
        const double r = sh_r[s];
        px = at.xyz[3 * a] + r * q[0];
        py = at.xyz[3 * a + 1] + r * q[1];
        pz = at.xyz[3 * a + 2] + r * q[2];
        bool reach = false;
        for (int i = 0; i < at.natom && !reach; ++i) {
            const double dx = px - at.xyz[3 * i], dy = py - at.xyz[3 * i + 1], dz = pz - at.xyz[3 * i + 2];
            reach = dx * dx + dy * dy + dz * dz <= at.reach2[i];
        }
        if (reach) {
            w = ssf_weight(px, py, pz, a, at) * q[3] * sh_rw[s];
            keep = w > 0.0 && w >= cut;
        }
    }
    double* o = raw + 4 * ((long long)s * kXcShellThreads + j);
    o[0] = px;
    o[1] = py;
    o[2] = pz;
    o[3] = keep ? w : 0.0;
    __shared__ int s_cnt;
    if (j == 0) s_cnt = 0;
    __syncthreads();
    if (keep) atomicAdd(&s_cnt, 1);
    __syncthreads();
    if (j == 0) count[s] = s_cnt;
}

// Kept points in shell order, within a shell in Lebedev order.
__global__ void __launch_bounds__(kXcShellThreads) xc_grid_compact(const double* __restrict__ raw,
                                                                   const long long* __restrict__ off,
                                                                   const int* __restrict__ sh_atom,
                                                                   double* __restrict__ pts, int* __restrict__ patom) {
    const int s = blockIdx.x, j = threadIdx.x;
    const double* r = raw + 4 * ((long long)s * kXcShellThreads + j);
    const bool keep = r[3] > 0.0;
    __shared__ int s_flag[kXcShellThreads];
    s_flag[j] = keep ? 1 : 0;
    __syncthreads();
    if (!keep) return;
    int rank = 0;
    for (int i = 0; i < j; ++i) rank += s_flag[i];
    const long long o = off[s] + rank;
    for (int c = 0; c < 4; ++c) pts[4 * o + c] = r[c];
    patom[o] = sh_atom[s];
}

// caller-provided points: drop those outside every sphere of significance
__global__ void xc_points_reach(const double* __restrict__ in, long long n, const XcAtoms at,
                                int* __restrict__ keep) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double px = in[4 * p], py = in[4 * p + 1], pz = in[4 * p + 2];
    bool reach = false;
    for (int i = 0; i < at.natom && !reach; ++i) {
        const double dx = px - at.xyz[3 * i], dy = py - at.xyz[3 * i + 1], dz = pz - at.xyz[3 * i + 2];
        reach = dx * dx + dy * dy + dz * dz <= at.reach2[i];
    }
    keep[p] = reach && in[4 * p + 3] > 0.0 ? 1 : 0;
}

__device__ __forceinline__ unsigned spread3(unsigned v) {
    v &= 0x3ff;
    v = (v | (v << 16)) & 0x030000ff;
    v = (v | (v << 8)) & 0x0300f00f;
    v = (v | (v << 4)) & 0x030c30c3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

// 30-bit Morton key of each point in a box (lo, 1024 cells of size h per axis)
__global__ void xc_morton(const double* __restrict__ pts, long long n, double lx, double ly, double lz, double inv_h,
                          unsigned* __restrict__ key, int* __restrict__ idx) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    auto q = [&](double v, double lo) {
        const double c = (v - lo) * inv_h;
        return (unsigned)(c < 0.0 ? 0.0 : (c > 1023.0 ? 1023.0 : c));
    };
    key[p] = spread3(q(pts[4 * p], lx)) | (spread3(q(pts[4 * p + 1], ly)) << 1) | (spread3(q(pts[4 * p + 2], lz)) << 2);
    idx[p] = (int)p;
}

__global__ void xc_gather_points(const double* __restrict__ in, const int* __restrict__ in_atom,
                                 const int* __restrict__ perm, long long n, double* __restrict__ out,
                                 int* __restrict__ out_atom) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long long s = perm[p];
    for (int c = 0; c < 4; ++c) out[4 * p + c] = in[4 * s + c];
    out_atom[p] = in_atom ? in_atom[s] : -1;
}

// Per batch of kXcBatch consecutive points: bounding sphere (centre of the
// box, radius to the farthest point) and the functions whose sphere of
// significance meets it, in ascending order.
__global__ void __launch_bounds__(256) xc_batch_lists(const double* __restrict__ pts, long long npts,
                                                      const int* __restrict__ f_atom,
                                                      const double* __restrict__ f_r2, const double* __restrict__ xyz,
                                                      int nbasis, int* __restrict__ sig, int* __restrict__ nsig) {
    const int b = blockIdx.x, t = threadIdx.x;
    const long long p0 = (long long)b * kXcBatch;
    const int cnt = (int)min<long long>(kXcBatch, npts - p0);
    __shared__ double s_lo[3][256], s_hi[3][256];
    __shared__ double s_c[4];
    for (int c = 0; c < 3; ++c) {
        s_lo[c][t] = t < cnt ? pts[4 * (p0 + t) + c] : 1e300;
        s_hi[c][t] = t < cnt ? pts[4 * (p0 + t) + c] : -1e300;
    }
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st)
            for (int c = 0; c < 3; ++c) {
                s_lo[c][t] = fmin(s_lo[c][t], s_lo[c][t + st]);
                s_hi[c][t] = fmax(s_hi[c][t], s_hi[c][t + st]);
            }
        __syncthreads();
    }
    if (t == 0)
        for (int c = 0; c < 3; ++c) s_c[c] = 0.5 * (s_lo[c][0] + s_hi[c][0]);
    __syncthreads();
    const double cx = s_c[0], cy = s_c[1], cz = s_c[2];
    {
        double r2 = 0.0;
        if (t < cnt) {
            const double dx = pts[4 * (p0 + t)] - cx, dy = pts[4 * (p0 + t) + 1] - cy, dz = pts[4 * (p0 + t) + 2] - cz;
            r2 = dx * dx + dy * dy + dz * dz;
        }
        s_lo[0][t] = r2;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (t < st) s_lo[0][t] = fmax(s_lo[0][t], s_lo[0][t + st]);
            __syncthreads();
        }
    }
    const double rb = sqrt(s_lo[0][0]);
    __shared__ int s_base, s_wcnt[4];
    if (t == 0) s_base = 0;
    __syncthreads();
    int* out = sig + (long long)b * nbasis;
    const int lane = t & 63, wv = t >> 6;
    for (int f0 = 0; f0 < nbasis; f0 += 256) {
        const int f = f0 + t;
        bool on = false;
        if (f < nbasis) {
            const int a = f_atom[f];
            const double dx = xyz[3 * a] - cx, dy = xyz[3 * a + 1] - cy, dz = xyz[3 * a + 2] - cz;
            const double d = sqrt(dx * dx + dy * dy + dz * dz), lim = rb + sqrt(f_r2[f]);
            on = d <= lim;
        }
        const unsigned long long m = __ballot(on);
        if (lane == 0) s_wcnt[wv] = __popcll(m);
        __syncthreads();
        int before = s_base;
        for (int i = 0; i < wv; ++i) before += s_wcnt[i];
        if (on) out[before + __popcll(m & ((1ull << lane) - 1))] = f;
        __syncthreads();
        if (t == 0) s_base += s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
        __syncthreads();
    }
    if (t == 0) nsig[b] = s_base;
}

// ---------------------------------------------------------------------------
// per density

struct XcFuncs {
    const int* atom;     // 0-based
    const int* lmn;      // l | m << 4 | n << 8
    const int* np;
    const double* c;     // kXcMaxPrim per function
    const double* a;
    const double* r2;    // sigrad2
    const double* xyz;
};

// Per-chunk batch descriptor: first point, points, functions, padded width
// (a multiple of kXcTile), offset of its planes in the workspace (Phi, dPhi/dx,
// dy, dz, Z: 5 planes of kXcBatch x width doubles) and of its row-dot partials.
struct XcBatch {
    long long p0;
    long long ws;
    long long part;
    int cnt;
    int m;
    int w;
    int batch;    // global batch index (its function list)
};

// Phi and its gradient for kXcBatch points x kXcTile functions of a batch.
__global__ void __launch_bounds__(256) xc_values(const XcBatch* __restrict__ bt, const double* __restrict__ pts,
                                                 const int* __restrict__ sig, int nbasis, const XcFuncs fn,
                                                 double* __restrict__ ws) {
    const XcBatch B = bt[blockIdx.y];
    const int f0 = blockIdx.x * kXcTile;
    if (f0 >= B.w) return;
    const int* sl = sig + (long long)B.batch * nbasis;
    double* P0 = ws + B.ws;
    const long long plane = (long long)kXcBatch * B.w;
    for (int q = 0; q < kXcBatch * kXcTile / 256; ++q) {
        const int e = threadIdx.x + 256 * q;
        const int p = e / kXcTile, k = f0 + (e % kXcTile);
        double v = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
        if (p < B.cnt && k < B.m) {
            const int mu = sl[k];
            const int at = fn.atom[mu];
            const double* P = pts + 4 * (B.p0 + p);
            const double x = P[0] - fn.xyz[3 * at], y = P[1] - fn.xyz[3 * at + 1], z = P[2] - fn.xyz[3 * at + 2];
            const double r2 = x * x + y * y + z * z;
            if (r2 <= fn.r2[mu]) {
                double rad = 0.0, drad = 0.0;   // sum c e^{-a r^2}, sum -2 a c e^{-a r^2}
                const int np = fn.np[mu];
                for (int c = 0; c < np; ++c) {
                    const double al = fn.a[kXcMaxPrim * mu + c];
                    const double t = fn.c[kXcMaxPrim * mu + c] * exp(-al * r2);
                    rad += t;
                    drad -= 2.0 * al * t;
                }
                const int code = fn.lmn[mu];
                const int l = code & 15, m = (code >> 4) & 15, n = (code >> 8) & 15;
                auto ipow = [](double b, int k) { return k == 0 ? 1.0 : (k == 1 ? b : (k == 2 ? b * b : b * b * b)); };
                const double xl = ipow(x, l), ym = ipow(y, m), zn = ipow(z, n);
                const double xl1 = l ? ipow(x, l - 1) : 0.0, ym1 = m ? ipow(y, m - 1) : 0.0,
                             zn1 = n ? ipow(z, n - 1) : 0.0;
                v = rad * xl * ym * zn;
                gx = (drad * xl * x + l * rad * xl1) * ym * zn;
                gy = (drad * ym * y + m * rad * ym1) * xl * zn;
                gz = (drad * zn * z + n * rad * zn1) * xl * ym;
            }
        }
        const long long o = (long long)p * B.w + k;
        P0[o] = v;
        P0[plane + o] = gx;
        P0[2 * plane + o] = gy;
        P0[3 * plane + o] = gz;
    }
}

// ---- FP64 MFMA tile: C(64 x 64) += sum_k A(i, k) B(k, j), operands staged
// through LDS in 16 x 64 slabs (k-major).  Block of four wavefronts; wavefront
// v owns rows 32 (v >> 1) .. +32 and columns 32 (v & 1) .. +32 as 2 x 2 tiles
// of v_mfma_f64_16x16x4.  Lane l of a 16x16x4 FP64 MFMA supplies A[l % 16][l / 16]
// and B[l / 16][l % 16] and holds C[4 r + l / 16][l % 16] in accumulator r (the
// FP64 form interleaves the four lane groups by row; the FP32 16x16x4 form
// instead gives each lane group four consecutive rows).
__device__ __forceinline__ void mfma_slab(v4d acc[2][2], const double* As, const double* Bs) {
    const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
    const int r0 = 32 * (v >> 1), c0 = 32 * (v & 1);
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        const int kk = 4 * ks + (lane >> 4);
        const double a0 = As[kk * kXcTile + r0 + (lane & 15)], a1 = As[kk * kXcTile + r0 + 16 + (lane & 15)];
        const double b0 = Bs[kk * kXcTile + c0 + (lane & 15)], b1 = Bs[kk * kXcTile + c0 + 16 + (lane & 15)];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
}

// rho GEMM task: (chunk batch, row tile of points, column tile of functions)
// T = Phi D_sub over all of the batch's functions; the epilogue forms the row
// dots of the tile with Phi and grad Phi and writes them as partials
// part[(col tile) x kXcBatch + row][4] (fixed order, no atomics).
__global__ void __launch_bounds__(256) xc_rho_gemm(const int4* __restrict__ tasks, const XcBatch* __restrict__ bt,
                                                   const int* __restrict__ sig, int nbasis,
                                                   const double* __restrict__ D, double* __restrict__ ws,
                                                   double* __restrict__ part) {
    const int4 tk = tasks[blockIdx.x];
    const XcBatch B = bt[tk.x];
    const int pr0 = tk.y * kXcTile, jc0 = tk.z * kXcTile;
    const int* sl = sig + (long long)B.batch * nbasis;
    const double* P0 = ws + B.ws;
    const long long plane = (long long)kXcBatch * B.w;
    // two LDS slabs: the MFMAs read one while the next is stored into the other
    __shared__ double As[2][16 * kXcTile], Bs[2][16 * kXcTile];
    __shared__ int s_sj[kXcTile];
    __shared__ double s_red[2][kXcTile][4];
    const int t = threadIdx.x;
    if (t < kXcTile) s_sj[t] = jc0 + t < B.m ? sl[jc0 + t] : -1;
    v4d acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    __syncthreads();
    // the slab at k0 into registers: A = Phi rows (points) x k, k contiguous in
    // memory; B = D[sig k][sig j]
    double ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            const int ia = e >> 4, ka = e & 15;
            ra[q] = P0[(long long)(pr0 + ia) * B.w + k0 + ka];
            const int kb = e >> 6, jb = e & 63;
            const int gk = k0 + kb < B.m ? sl[k0 + kb] : -1, gj = s_sj[jb];
            rb[q] = (gk >= 0 && gj >= 0) ? D[(long long)gk * nbasis + gj] : 0.0;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            As[buf][(e & 15) * kXcTile + (e >> 4)] = ra[q];
            Bs[buf][(e >> 6) * kXcTile + (e & 63)] = rb[q];
        }
    };
    load(0);
    store(0);
    __syncthreads();
    // slab it is in buffer it & 1; the loads of slab it + 1 are in flight
    // during its MFMAs, and the barrier at the end of step it orders both the
    // reads of buffer it & 1 before its next stores and the stores of slab
    // it + 1 before its reads
    for (int k0 = 0, it = 0; k0 < B.w; k0 += 16, ++it) {
        const bool more = k0 + 16 < B.w;
        if (more) load(k0 + 16);
        mfma_slab(acc, As[it & 1], Bs[it & 1]);
        if (more) store((it & 1) ^ 1);
        __syncthreads();
    }
    // epilogue: row dots with Phi, dPhi/dx, dPhi/dy, dPhi/dz over the tile's columns
    const int lane = t & 63, v = t >> 6;
    const int r0 = 32 * (v >> 1), c0 = 32 * (v & 1);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + 16 * i + 4 * r + (lane >> 4);
            double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int col = c0 + 16 * j + (lane & 15);
                const double tv = acc[i][j][r];
                const long long o = (long long)(pr0 + row) * B.w + jc0 + col;
#pragma unroll
                for (int c = 0; c < 4; ++c) s[c] += tv * P0[c * plane + o];
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
                for (int off = 8; off >= 1; off >>= 1) s[c] += __shfl_xor(s[c], off, 64);
            if ((lane & 15) == 0)
                for (int c = 0; c < 4; ++c) s_red[v & 1][row][c] = s[c];
        }
    __syncthreads();
    {
        const int row = t >> 2, c = t & 3;
        part[B.part + ((long long)tk.z * kXcBatch + pr0 + row) * 4 + c] = s_red[0][row][c] + s_red[1][row][c];
    }
}

// Per point: rho and grad rho from the partials, B3LYP, the point's energy and
// electron count, then Z = w (vrho/2 Phi + 2 vsigma grad rho . grad Phi) for the
// whole batch (one block per batch).
__global__ void __launch_bounds__(256) xc_point(const XcBatch* __restrict__ bt, const double* __restrict__ pts,
                                                double* __restrict__ ws, const double* __restrict__ part,
                                                double dcut, double* __restrict__ e_pt, double* __restrict__ n_pt) {
    const XcBatch B = bt[blockIdx.x];
    const int p = threadIdx.x;
    const int ntile = B.w / kXcTile;
    __shared__ double s_c[kXcBatch][4];
    double c0 = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    if (p < B.cnt) {
        double rho = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
        for (int jt = 0; jt < ntile; ++jt) {
            const double* q = part + B.part + ((long long)jt * kXcBatch + p) * 4;
            rho += q[0];
            gx += q[1];
            gy += q[2];
            gz += q[3];
        }
        gx *= 2.0;
        gy *= 2.0;
        gz *= 2.0;
        const double w = pts[4 * (B.p0 + p) + 3];
        double e = 0.0, ne = 0.0;
        if (0.5 * rho >= dcut) {
            const double sigma = gx * gx + gy * gy + gz * gz;
            double f, vr, vs;
            b3lyp_closed<double>(rho, sigma, &f, &vr, &vs);
            e = w * f;
            ne = w * rho;
            c0 = 0.5 * w * vr;
            const double g = 2.0 * w * vs;
            cx = g * gx;
            cy = g * gy;
            cz = g * gz;
        }
        e_pt[B.p0 + p] = e;
        n_pt[B.p0 + p] = ne;
    }
    s_c[p][0] = c0;
    s_c[p][1] = cx;
    s_c[p][2] = cy;
    s_c[p][3] = cz;
    __syncthreads();
    double* P0 = ws + B.ws;
    const long long plane = (long long)kXcBatch * B.w;
    // element e = pp w + k: point pp, column k, advanced by 256 elements per
    // step without a division (k + dr < 2 w)
    const int w = B.w, dq = 256 / w, dr = 256 % w;
    int pp = threadIdx.x / w, k = threadIdx.x % w;
    for (long long e = threadIdx.x; e < plane; e += 256) {
        const double* c = s_c[pp];
        pp += dq;
        k += dr;
        if (k >= w) {
            k -= w;
            ++pp;
        }
        P0[4 * plane + e] = c[0] * P0[e] + c[1] * P0[plane + e] + c[2] * P0[2 * plane + e] + c[3] * P0[3 * plane + e];
    }
}

// V GEMM task: (chunk batch, row tile I >= column tile J) of
// V_sub = Phi^T Z + Z^T Phi over the batch's points, added to the fixed-point
// accumulator (packed lower triangle over basis functions); in a diagonal tile
// only the elements on or below its diagonal.
__global__ void __launch_bounds__(256) xc_v_gemm(const int4* __restrict__ tasks, const XcBatch* __restrict__ bt,
                                                 const int* __restrict__ sig, int nbasis,
                                                 const double* __restrict__ ws, unsigned long long* __restrict__ V) {
    const int4 tk = tasks[blockIdx.x];
    const XcBatch B = bt[tk.x];
    const int I0 = tk.y * kXcTile, J0 = tk.z * kXcTile;
    const int* sl = sig + (long long)B.batch * nbasis;
    const double* P0 = ws + B.ws;
    const double* Z = P0 + 4 * (long long)kXcBatch * B.w;
    __shared__ double As[2][16 * kXcTile], Bs[2][16 * kXcTile];
    const int t = threadIdx.x;
    v4d acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    // step st: pass st / nk (0: A(i, k) = Phi[k][I0 + i], B(k, j) = Z[k][J0 + j];
    // 1: the operands swapped), points 16 (st % nk) ..; loaded into registers
    // one step ahead and stored into the LDS slab the MFMAs are not reading,
    // as in xc_rho_gemm
    constexpr int nk = kXcBatch / 16;
    double ra[4], rb[4];
    auto load = [&](int st) {
        const double* L = st < nk ? P0 : Z;
        const double* R = st < nk ? Z : P0;
        const int k0 = 16 * (st % nk);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            const long long row = (long long)(k0 + (e >> 6)) * B.w;
            ra[q] = L[row + I0 + (e & 63)];
            rb[q] = R[row + J0 + (e & 63)];
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            As[buf][e] = ra[q];   // (e >> 6) * kXcTile + (e & 63)
            Bs[buf][e] = rb[q];
        }
    };
    load(0);
    store(0);
    __syncthreads();
    for (int st = 0; st < 2 * nk; ++st) {
        const bool more = st + 1 < 2 * nk;
        if (more) load(st + 1);
        mfma_slab(acc, As[st & 1], Bs[st & 1]);
        if (more) store((st & 1) ^ 1);
        __syncthreads();
    }
    const int lane = t & 63, v = t >> 6;
    const int r0 = 32 * (v >> 1), c0 = 32 * (v & 1);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = I0 + r0 + 16 * i + 4 * r + (lane >> 4), col = J0 + c0 + 16 * j + (lane & 15);
                if (row >= B.m || col >= B.m) continue;
                if (I0 == J0 && row < col) continue;
                const long long iv = __double2ll_rn(acc[i][j][r] * kXcFix);
                if (iv != 0) atomicAdd(V + xtri(sl[row], sl[col]), (unsigned long long)iv);
            }
}

// V (n x n, full symmetric) from the accumulator; flags a sum near the range
__global__ void xc_finalize(const unsigned long long* __restrict__ V, double* __restrict__ out, int n,
                            unsigned long long* __restrict__ flag) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)n * n) return;
    const int i = int(k / n), j = int(k % n);
    const long long s = (long long)V[xtri(i, j)];
    if (s >= (1LL << 62) || s <= -(1LL << 62)) atomicOr(flag, 1ull);
    out[k] = (double)s * 0x1p-48;
}

// Sums of e_pt and n_pt in a fixed order, in two stages: block b of
// kXcSumBlocks sums the points of its contiguous chunk (thread-strided, then a
// tree), writing part_e[b], part_n[b]; xc_sums then adds the partials.  (One
// block over every point streamed them at one CU's bandwidth: 1.37 ms for
// 2.9M points.)
constexpr int kXcSumBlocks = 1024;
__global__ void __launch_bounds__(256) xc_sums_part(const double* __restrict__ e, const double* __restrict__ ne,
                                                    long long n, long long chunk, double* __restrict__ part_e,
                                                    double* __restrict__ part_n) {
    __shared__ double s0[256], s1[256];
    const long long p0 = (long long)blockIdx.x * chunk, p1 = p0 + chunk < n ? p0 + chunk : n;
    double a = 0.0, b = 0.0;
    for (long long p = p0 + threadIdx.x; p < p1; p += 256) {
        a += e[p];
        b += ne[p];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            s0[threadIdx.x] += s0[threadIdx.x + st];
            s1[threadIdx.x] += s1[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_e[blockIdx.x] = s0[0];
        part_n[blockIdx.x] = s1[0];
    }
}

// sums of e and ne in a fixed order (one block of 1024 threads)
__global__ void __launch_bounds__(1024) xc_sums(const double* __restrict__ e, const double* __restrict__ ne,
                                                long long n, double* __restrict__ out) {
    __shared__ double s0[1024], s1[1024];
    double a = 0.0, b = 0.0;
    for (long long p = threadIdx.x; p < n; p += 1024) {
        a += e[p];
        b += ne[p];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            s0[threadIdx.x] += s0[threadIdx.x + st];
            s1[threadIdx.x] += s1[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s0[0];
        out[1] = s1[0];
    }
}

__global__ void xc_fill_u64(unsigned long long* p, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0ull;
}

}  // namespace qk

// ===========================================================================
// host side

using namespace qk;

namespace {

std::string g_xc_err;

void xerr(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    qk_set_error_text(buf);
}

#define XCHK(call)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            xerr("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return -1;                                                                      \
        }                                                                                   \
    } while (0)

template <typename T>
struct XBuf {
    T* p = nullptr;
    size_t cap = 0;
    int alloc(size_t n) {
        if (n == 0) n = 1;
        if (p && n <= cap) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) {
            p = nullptr;
            xerr("hipMalloc of %zu bytes failed", n * sizeof(T));
            return -1;
        }
        cap = n;
        return 0;
    }
    int upload(const std::vector<T>& v) {
        if (alloc(v.size())) return -1;
        if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
            xerr("hipMemcpy H2D failed");
            return -1;
        }
        return 0;
    }
    ~XBuf() {
        if (p) (void)hipFree(p);
    }
};

// SG-1 atomic radii in bohr by atomic number (Gill, Johnson, Pople, CPL 209,
// 506 (1993), Table 1; the table QUICK's SG-1 grid uses)
double sg1_radius(int Z) {
    static const double r[19] = {0.0, 1.0, 0.5882, 3.0769, 2.0513, 1.5385, 1.2308, 1.0256, 0.8791, 0.7692,
                                 0.6838, 4.0909, 3.1579, 2.5714, 2.1687, 1.8750, 1.6514, 1.4754, 1.3333};
    return (Z >= 1 && Z <= 18) ? r[Z] : 2.25;
}

// Lebedev rule index of the SG-1 pruning region of radial coordinate x = r / R
// (SG-1 partition boundaries; dft.f90:3106-3190): 6, 38, 86, 194, then 86
int sg1_rule(int Z, double x) {
    static const double h[4] = {0.25, 0.5, 1.0, 4.5}, l[4] = {0.1667, 0.5, 0.9, 3.5}, n[4] = {0.1, 0.4, 0.8, 2.5};
    const double* b = Z <= 2 ? h : (Z <= 10 ? l : n);
    return x < b[0] ? 0 : x < b[1] ? 1 : x < b[2] ? 2 : x < b[3] ? 3 : 2;
}

int leb_rule_of(int npoints) {
    for (int r = 0; r < kLebRules; ++r)
        if (kLebSize[r] == npoints) return r;
    return -1;
}

}  // namespace

extern "C" {

struct qk_xc {
    int device = 0;
    int natom = 0, nbasis = 0;
    std::vector<double> xyz;
    std::vector<int> Z;
    std::vector<double> sig2;
    std::vector<double> reach2;
    XBuf<double> d_xyz, d_reach2, d_fc, d_fa, d_fr2;
    XBuf<int> d_fatom, d_flmn, d_fnp;
    // grid (Morton order)
    long long npts = 0;
    XBuf<double> d_pts;
    XBuf<int> d_patom;
    // batches
    int nbatch = 0;
    XBuf<int> d_sig, d_nsig;
    std::vector<int> h_nsig;
    // chunks of batches and their GEMM tasks
    struct Chunk {
        int b0 = 0, nb = 0;
        long long ws = 0, part = 0;
        long long rho0 = 0, nrho = 0, v0 = 0, nv = 0;
    };
    std::vector<Chunk> chunks;
    XBuf<XcBatch> d_bt;
    XBuf<int4> d_rho_tasks, d_v_tasks;
    XBuf<double> d_ws, d_part, d_ept, d_npt, d_sums, d_D, d_V;
    XBuf<unsigned long long> d_Vacc, d_flag;
    double weight_cut = 1e-10, basis_cut = 1e-10, density_cut = 1e-10;
    double flops = 0.0;   // algorithmic FP64 MFMA flops of one evaluation
};

// Workspace of the batches one launch covers (values, gradients and Z of each
// point and function): an eighth of the device memory free when the grid is
// built, at least 1 GiB.  Fewer, larger chunks fill the GPU better: on (H2O)64
// (75 x 302 grid, about 22 GB of workspace in all) one evaluation took 45.8 ms
// in 1 GiB chunks, 35.9 ms in 4 GiB and 33.3 ms in 8 GiB (smaller chunks, to
// stay in the 256 MB last-level cache, were slower still: 69 ms at 256 MiB).
static long long xc_workspace_bytes() {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 1LL << 30;
    return std::max(1LL << 30, (long long)(free_b / 8));
}

void qk_xc_default_opts(qk_xc_opts* o) {
    o->grid = QK_GRID_SG1;
    o->nrad = 50;
    o->nang = 194;
    o->weight_cut = 1e-10;
    o->basis_cut = 1e-10;
    o->density_cut = 1e-10;
}

qk_xc* qk_xc_create(int device) {
    if (hipSetDevice(device) != hipSuccess) {
        xerr("qk_xc_create: hipSetDevice(%d) failed", device);
        return nullptr;
    }
    qk_xc* x = new qk_xc();
    x->device = device;
    return x;
}

void qk_xc_destroy(qk_xc* x) {
    if (!x) return;
    (void)hipSetDevice(x->device);
    (void)hipDeviceSynchronize();
    delete x;
}

// radius of significance squared of each function: the r where the spherical
// average of its most diffuse primitive, r^L exp(-a r^2), falls to the basis
// cut scaled by that primitive's normalisation (get_sigrad, dft.f90:536-613),
// found by the same decimal step search
static double sigrad2_of(double amin, int L, double cut) {
    double gamma = 1.0;
    for (int i = 1; i <= L + 1; ++i) gamma *= (double)(L + 1 - i) + 0.5;
    const double gamma2pi = gamma * 11.13665599366341569;
    const double target = cut * std::pow(std::pow(2.0 * amin, (double)L + 1.5) / gamma2pi, -0.5);
    double step = 1.0, r = 0.0;
    while (step > 1e-4) {
        r += step;
        if (std::exp(-amin * r * r) * std::pow(r, (double)L) < target) {
            r -= step;
            step /= 10.0;
        }
    }
    return r * r;
}

int qk_xc_set_system(qk_xc* x, int natom, const double* xyz, const int* Z, int nbasis, int maxcontract,
                     const int* ncontract, const int* itype, const double* aexp, const double* dcoeff,
                     const int* ncenter, const double* sigrad2_or_null, double basis_cut) {
    if (!x) { xerr("null XC context"); return -1; }
    if (maxcontract > kXcMaxPrim) { xerr("qk_xc_set_system: more than %d primitives per function", kXcMaxPrim); return -1; }
    (void)hipSetDevice(x->device);
    x->natom = natom;
    x->nbasis = nbasis;
    x->xyz.assign(xyz, xyz + 3 * (size_t)natom);
    x->Z.assign(Z, Z + natom);
    x->basis_cut = basis_cut;
    std::vector<int> fatom(nbasis), flmn(nbasis), fnp(nbasis);
    std::vector<double> fc((size_t)kXcMaxPrim * nbasis, 0.0), fa((size_t)kXcMaxPrim * nbasis, 0.0);
    x->sig2.resize(nbasis);
    x->reach2.assign(natom, 0.0);
    for (int f = 0; f < nbasis; ++f) {
        const int a = ncenter[f] - 1;
        if (a < 0 || a >= natom) { xerr("qk_xc_set_system: function %d on atom %d of %d", f, a + 1, natom); return -1; }
        const int l = itype[3 * f], m = itype[3 * f + 1], n = itype[3 * f + 2];
        if (l + m + n > 3 || ncontract[f] < 1 || ncontract[f] > maxcontract) {
            xerr("qk_xc_set_system: function %d: unsupported shape", f);
            return -1;
        }
        fatom[f] = a;
        flmn[f] = l | (m << 4) | (n << 8);
        fnp[f] = ncontract[f];
        double amin = 1e300;
        for (int c = 0; c < ncontract[f]; ++c) {
            fc[(size_t)kXcMaxPrim * f + c] = dcoeff[(size_t)maxcontract * f + c];
            fa[(size_t)kXcMaxPrim * f + c] = aexp[(size_t)maxcontract * f + c];
            amin = std::min(amin, aexp[(size_t)maxcontract * f + c]);
        }
        x->sig2[f] = sigrad2_or_null ? sigrad2_or_null[f] : sigrad2_of(amin, l + m + n, basis_cut);
        x->reach2[a] = std::max(x->reach2[a], x->sig2[f]);
    }
    if (x->d_xyz.upload(x->xyz) || x->d_reach2.upload(x->reach2) || x->d_fatom.upload(fatom) ||
        x->d_flmn.upload(flmn) || x->d_fnp.upload(fnp) || x->d_fc.upload(fc) || x->d_fa.upload(fa) ||
        x->d_fr2.upload(x->sig2))
        return -1;
    x->npts = 0;
    x->nbatch = 0;
    x->chunks.clear();
    return 0;
}

int qk_xc_sigrad2(qk_xc* x, double* out) {
    if (!x || x->sig2.empty()) { xerr("qk_xc_sigrad2: no system"); return -1; }
    std::copy(x->sig2.begin(), x->sig2.end(), out);
    return 0;
}

// Sort the points along a Morton curve, cut batches, build their function
// lists, chunk them and build the GEMM task lists (once per grid).
static int xc_finish_grid(qk_xc* x, XBuf<double>& pts_in, XBuf<int>& atom_in, long long n) {
    x->npts = n;
    if (x->d_sums.alloc(2 + 2 * kXcSumBlocks) || x->d_flag.alloc(1) || x->d_Vacc.alloc((size_t)xtri_size(x->nbasis)) ||
        x->d_ept.alloc(std::max(1LL, n)) || x->d_npt.alloc(std::max(1LL, n)))
        return -1;
    if (n == 0) { x->nbatch = 0; x->chunks.clear(); return 0; }
    // Morton order in the box of the atoms widened by the largest reach
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, reach = 0.0;
    for (int a = 0; a < x->natom; ++a) {
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], x->xyz[3 * a + c]);
            hi[c] = std::max(hi[c], x->xyz[3 * a + c]);
        }
        reach = std::max(reach, std::sqrt(x->reach2[a]));
    }
    double side = 0.0;
    for (int c = 0; c < 3; ++c) {
        lo[c] -= reach;
        hi[c] += reach;
        side = std::max(side, hi[c] - lo[c]);
    }
    XBuf<unsigned> key, key2;
    XBuf<int> idx, idx2;
    if (key.alloc(n) || key2.alloc(n) || idx.alloc(n) || idx2.alloc(n)) return -1;
    const int nb256 = (int)((n + 255) / 256);
    hipLaunchKernelGGL(xc_morton, dim3(nb256), dim3(256), 0, 0, pts_in.p, n, lo[0], lo[1], lo[2],
                       1024.0 / std::max(side, 1e-6), key.p, idx.p);
    XCHK(hipGetLastError());
    size_t tmp_bytes = 0;
    XCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key.p, key2.p, idx.p, idx2.p, (int)n, 0, 30));
    XBuf<char> tmp;
    if (tmp.alloc(tmp_bytes)) return -1;
    XCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, key.p, key2.p, idx.p, idx2.p, (int)n, 0, 30));
    if (x->d_pts.alloc(4 * (size_t)n) || x->d_patom.alloc(n)) return -1;
    hipLaunchKernelGGL(xc_gather_points, dim3(nb256), dim3(256), 0, 0, pts_in.p, atom_in.p, idx2.p, n, x->d_pts.p,
                       x->d_patom.p);
    XCHK(hipGetLastError());
    // batches and their function lists
    x->nbatch = (int)((n + kXcBatch - 1) / kXcBatch);
    if (x->d_sig.alloc((size_t)x->nbatch * x->nbasis) || x->d_nsig.alloc(x->nbatch)) return -1;
    hipLaunchKernelGGL(xc_batch_lists, dim3(x->nbatch), dim3(256), 0, 0, x->d_pts.p, n, x->d_fatom.p, x->d_fr2.p,
                       x->d_xyz.p, x->nbasis, x->d_sig.p, x->d_nsig.p);
    XCHK(hipGetLastError());
    x->h_nsig.resize(x->nbatch);
    XCHK(hipMemcpy(x->h_nsig.data(), x->d_nsig.p, sizeof(int) * x->nbatch, hipMemcpyDeviceToHost));
    // chunks within the workspace; tasks per chunk
    std::vector<XcBatch> bts;
    std::vector<int4> rho_t, v_t;
    x->chunks.clear();
    qk_xc::Chunk ch;
    long long ws = 0, part = 0, maxws = 0, maxpart = 0;
    const long long kXcWorkspaceBytes = xc_workspace_bytes();
    x->flops = 0.0;
    for (int b = 0; b < x->nbatch; ++b) {
        const int m = x->h_nsig[b];
        const int w = std::max(kXcTile, (m + kXcTile - 1) / kXcTile * kXcTile);
        const long long need = 5LL * kXcBatch * w;
        if (ch.nb > 0 && (ws + need) * 8 > kXcWorkspaceBytes) {
            x->chunks.push_back(ch);
            ch = qk_xc::Chunk();
            ch.b0 = b;
            ch.rho0 = (long long)rho_t.size();
            ch.v0 = (long long)v_t.size();
            ws = part = 0;
        }
        XcBatch B;
        B.p0 = (long long)b * kXcBatch;
        B.cnt = (int)std::min<long long>(kXcBatch, n - B.p0);
        B.m = m;
        B.w = w;
        B.ws = ws;
        B.part = part;
        B.batch = b;
        const int local = ch.nb;
        bts.push_back(B);
        ws += need;
        part += 4LL * kXcBatch * (w / kXcTile);
        maxws = std::max(maxws, ws);
        maxpart = std::max(maxpart, part);
        ch.nb += 1;
        if (m == 0) continue;
        const int T = w / kXcTile;
        for (int r = 0; r < kXcBatch / kXcTile; ++r)
            for (int c = 0; c < T; ++c) rho_t.push_back(int4{local, r, c, 0});
        for (int I = 0; I < T; ++I)
            for (int J = 0; J <= I; ++J) v_t.push_back(int4{local, I, J, 0});
        ch.nrho = (long long)rho_t.size() - ch.rho0;
        ch.nv = (long long)v_t.size() - ch.v0;
        x->flops += 2.0 * kXcBatch * (double)w * w + 2.0 * 2.0 * kXcBatch * kXcTile * kXcTile * (T * (T + 1) / 2);
    }
    if (ch.nb > 0) x->chunks.push_back(ch);
    // batch descriptors are per chunk: local index within the chunk
    if (x->d_bt.upload(bts) || x->d_rho_tasks.upload(rho_t) || x->d_v_tasks.upload(v_t)) return -1;
    if (x->d_ws.alloc(std::max(1LL, maxws)) || x->d_part.alloc(std::max(1LL, maxpart)) || x->d_ept.alloc(n) ||
        x->d_npt.alloc(n) || x->d_sums.alloc(2 + 2 * kXcSumBlocks) || x->d_flag.alloc(1) ||
        x->d_Vacc.alloc((size_t)xtri_size(x->nbasis)))
        return -1;
    XCHK(hipDeviceSynchronize());
    return 0;
}

int qk_xc_build_grid(qk_xc* x, const qk_xc_opts* o) {
    if (!x || x->natom == 0) { xerr("qk_xc_build_grid: no system"); return -1; }
    qk_xc_opts d;
    if (!o) { qk_xc_default_opts(&d); o = &d; }
    (void)hipSetDevice(x->device);
    x->weight_cut = o->weight_cut;
    x->density_cut = o->density_cut;
    const int nrad = o->grid == QK_GRID_SG1 ? 50 : o->nrad;
    int prod_rule = -1;
    if (o->grid == QK_GRID_PRODUCT) {
        prod_rule = leb_rule_of(o->nang);
        if (prod_rule < 0) { xerr("qk_xc_build_grid: no Lebedev rule of %d points (6, 38, 86, 194, 302)", o->nang); return -1; }
    } else if (o->grid != QK_GRID_SG1) {
        xerr("qk_xc_build_grid: unknown grid kind %d", o->grid);
        return -1;
    }
    if (nrad < 1 || nrad > 1000) { xerr("qk_xc_build_grid: %d radial points", nrad); return -1; }
    // radial shells: Euler-Maclaurin points r = R i^2 / (N + 1 - i)^2 with weights
    // 2 R^3 (N + 1) i^5 / (N + 1 - i)^7 (the r^2 of the volume element included)
    std::vector<int> sh_atom, sh_rule;
    std::vector<double> sh_r, sh_rw;
    for (int a = 0; a < x->natom; ++a) {
        const double R = sg1_radius(x->Z[a]), R3 = R * R * R;
        for (int i = 1; i <= nrad; ++i) {
            const double xr = (double)i * i / ((double)(nrad + 1 - i) * (nrad + 1 - i));
            const double rw = 2.0 * (nrad + 1) * std::pow((double)i, 5.0) * std::pow((double)(nrad + 1 - i), -7.0);
            sh_atom.push_back(a);
            sh_rule.push_back(o->grid == QK_GRID_SG1 ? sg1_rule(x->Z[a], xr) : prod_rule);
            sh_r.push_back(R * xr);
            sh_rw.push_back(rw * R3);
        }
    }
    const int ns = (int)sh_atom.size();
    XBuf<int> d_sa, d_sr, d_cnt;
    XBuf<double> d_r, d_rw, raw;
    XBuf<long long> d_off;
    if (d_sa.upload(sh_atom) || d_sr.upload(sh_rule) || d_r.upload(sh_r) || d_rw.upload(sh_rw) ||
        raw.alloc(4 * (size_t)ns * kXcShellThreads) || d_cnt.alloc(ns))
        return -1;
    XcAtoms at{x->d_xyz.p, x->d_reach2.p, x->natom};
    hipLaunchKernelGGL(xc_grid_shells, dim3(ns), dim3(kXcShellThreads), 0, 0, d_sa.p, d_sr.p, d_r.p, d_rw.p, at,
                       o->weight_cut, raw.p, d_cnt.p);
    XCHK(hipGetLastError());
    std::vector<int> cnt(ns);
    XCHK(hipMemcpy(cnt.data(), d_cnt.p, sizeof(int) * ns, hipMemcpyDeviceToHost));
    std::vector<long long> off(ns);
    long long tot = 0;
    for (int s = 0; s < ns; ++s) {
        off[s] = tot;
        tot += cnt[s];
    }
    XBuf<double> pts;
    XBuf<int> patom;
    if (d_off.upload(off) || pts.alloc(4 * (size_t)std::max(1LL, tot)) || patom.alloc(std::max(1LL, tot))) return -1;
    hipLaunchKernelGGL(xc_grid_compact, dim3(ns), dim3(kXcShellThreads), 0, 0, raw.p, d_off.p, d_sa.p, pts.p, patom.p);
    XCHK(hipGetLastError());
    return xc_finish_grid(x, pts, patom, tot);
}

int qk_xc_set_grid(qk_xc* x, long long npts, const double* pts_host) {
    if (!x || x->natom == 0) { xerr("qk_xc_set_grid: no system"); return -1; }
    (void)hipSetDevice(x->device);
    XBuf<double> in;
    XBuf<int> keep;
    if (in.alloc(4 * (size_t)std::max(1LL, npts)) || keep.alloc(std::max(1LL, npts))) return -1;
    if (npts > 0) XCHK(hipMemcpy(in.p, pts_host, sizeof(double) * 4 * npts, hipMemcpyHostToDevice));
    XcAtoms at{x->d_xyz.p, x->d_reach2.p, x->natom};
    if (npts > 0)
        hipLaunchKernelGGL(xc_points_reach, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, 0, in.p, npts, at,
                           keep.p);
    std::vector<int> hk(npts);
    if (npts > 0) XCHK(hipMemcpy(hk.data(), keep.p, sizeof(int) * npts, hipMemcpyDeviceToHost));
    std::vector<double> hin(pts_host, pts_host + 4 * npts), sel;
    std::vector<int> satom;
    for (long long p = 0; p < npts; ++p)
        if (hk[p]) {
            sel.insert(sel.end(), hin.begin() + 4 * p, hin.begin() + 4 * p + 4);
            satom.push_back(-1);
        }
    XBuf<double> pts;
    XBuf<int> patom;
    if (pts.upload(sel) || patom.upload(satom)) return -1;
    return xc_finish_grid(x, pts, patom, (long long)satom.size());
}

long long qk_xc_grid_size(qk_xc* x) { return x ? x->npts : -1; }

int qk_xc_get_grid(qk_xc* x, double* pts_out, int* atom_out) {
    if (!x) { xerr("null XC context"); return -1; }
    (void)hipSetDevice(x->device);
    if (x->npts == 0) return 0;
    if (pts_out) XCHK(hipMemcpy(pts_out, x->d_pts.p, sizeof(double) * 4 * x->npts, hipMemcpyDeviceToHost));
    if (atom_out) XCHK(hipMemcpy(atom_out, x->d_patom.p, sizeof(int) * x->npts, hipMemcpyDeviceToHost));
    return 0;
}

int qk_xc_batch_info(qk_xc* x, int* nbatch, double* mean_functions, double* flops) {
    if (!x) { xerr("null XC context"); return -1; }
    *nbatch = x->nbatch;
    double s = 0.0;
    for (int m : x->h_nsig) s += m;
    *mean_functions = x->nbatch ? s / x->nbatch : 0.0;
    *flops = x->flops;
    return 0;
}

int qk_xc_device(qk_xc* x, const double* dD, double* dV, double* Exc, double* nelec, void* stream) {
    if (!x || x->natom == 0) { xerr("qk_xc_device: no system"); return -1; }
    if (!x->d_Vacc.p || !x->d_sums.p || !x->d_flag.p || !x->d_ept.p) {
        xerr("qk_xc_device: no grid (call qk_xc_build_grid or qk_xc_set_grid first)");
        return -1;
    }
    (void)hipSetDevice(x->device);
    hipStream_t s = (hipStream_t)stream;
    const int n = x->nbasis;
    const long long tri = xtri_size(n);
    hipLaunchKernelGGL(xc_fill_u64, dim3((unsigned)((tri + 255) / 256)), dim3(256), 0, s, x->d_Vacc.p, tri);
    hipLaunchKernelGGL(xc_fill_u64, dim3(1), dim3(64), 0, s, x->d_flag.p, 1LL);
    XCHK(hipGetLastError());
    XcFuncs fn{x->d_fatom.p, x->d_flmn.p, x->d_fnp.p, x->d_fc.p, x->d_fa.p, x->d_fr2.p, x->d_xyz.p};
    for (const auto& ch : x->chunks) {
        const XcBatch* bt = x->d_bt.p + ch.b0;
        int wmax = 0;
        for (int b = ch.b0; b < ch.b0 + ch.nb; ++b)
            wmax = std::max(wmax, std::max(kXcTile, (x->h_nsig[b] + kXcTile - 1) / kXcTile * kXcTile));
        hipLaunchKernelGGL(xc_values, dim3(wmax / kXcTile, ch.nb), dim3(256), 0, s, bt, x->d_pts.p, x->d_sig.p, n, fn,
                           x->d_ws.p);
        if (ch.nrho > 0)
            hipLaunchKernelGGL(xc_rho_gemm, dim3((unsigned)ch.nrho), dim3(256), 0, s, x->d_rho_tasks.p + ch.rho0, bt,
                               x->d_sig.p, n, dD, x->d_ws.p, x->d_part.p);
        hipLaunchKernelGGL(xc_point, dim3(ch.nb), dim3(256), 0, s, bt, x->d_pts.p, x->d_ws.p, x->d_part.p,
                           x->density_cut, x->d_ept.p, x->d_npt.p);
        if (ch.nv > 0)
            hipLaunchKernelGGL(xc_v_gemm, dim3((unsigned)ch.nv), dim3(256), 0, s, x->d_v_tasks.p + ch.v0, bt,
                               x->d_sig.p, n, x->d_ws.p, x->d_Vacc.p);
        XCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(xc_finalize, dim3((unsigned)(((long long)n * n + 255) / 256)), dim3(256), 0, s, x->d_Vacc.p,
                       dV, n, x->d_flag.p);
    {
        const long long chunk = (x->npts + kXcSumBlocks - 1) / kXcSumBlocks;
        double* pe = x->d_sums.p + 2;
        double* pn = pe + kXcSumBlocks;
        hipLaunchKernelGGL(xc_sums_part, dim3(kXcSumBlocks), dim3(256), 0, s, x->d_ept.p, x->d_npt.p, x->npts, chunk,
                           pe, pn);
        hipLaunchKernelGGL(xc_sums, dim3(1), dim3(1024), 0, s, pe, pn, (long long)kXcSumBlocks, x->d_sums.p);
    }
    XCHK(hipGetLastError());
    if (Exc || nelec) {
        double h[2];
        XCHK(hipMemcpyAsync(h, x->d_sums.p, sizeof(h), hipMemcpyDeviceToHost, s));
        unsigned long long f = 0;
        XCHK(hipMemcpyAsync(&f, x->d_flag.p, sizeof(f), hipMemcpyDeviceToHost, s));
        XCHK(hipStreamSynchronize(s));
        if (f) { xerr("qk_xc_device: V_xc accumulator near its range"); return -5; }
        if (Exc) *Exc = h[0];
        if (nelec) *nelec = h[1];
    }
    return 0;
}

int qk_xc_host(qk_xc* x, const double* D, double* V, double* Exc, double* nelec) {
    if (!x || x->natom == 0) { xerr("qk_xc_host: no system"); return -1; }
    (void)hipSetDevice(x->device);
    const size_t nn = (size_t)x->nbasis * x->nbasis;
    if (x->d_D.alloc(nn) || x->d_V.alloc(nn)) return -1;
    XCHK(hipMemcpy(x->d_D.p, D, nn * sizeof(double), hipMemcpyHostToDevice));
    const int rc = qk_xc_device(x, x->d_D.p, x->d_V.p, Exc, nelec, nullptr);
    if (rc) return rc;
    XCHK(hipMemcpy(V, x->d_V.p, nn * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
