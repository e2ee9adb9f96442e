// qk_lists.hip -- the per-geometry work lists of the Fock build and of the
// four-centre gradient, built on the device.
//
// Replaces the host-side list construction of the reference GPU path
// (gpu_upload_cutoff_matrix_, gpu.cu:401-1003: sorted pair list by class and
// bound) for this library's list layout.  The host sorts the significant
// Qshell pairs into groups (pair class, primitive-pair count, Schwarz bound
// descending) and describes every (bra group, ket groups) block of a quartet
// class as a ListEntry.  One thread per item -- a bra pair against one ket
// group -- then finds by binary search the prefix of kets passing
// Y_ij * Y_kl > tau (the kets are sorted by Y) and, in mixed precision, the
// split at Y_ij * Y_kl >= fp32_threshold; exclusive scans over the items give
// every item its place in the class's tile, segment and candidate lists; a
// second pass writes them.  For the Fock build one thread per bra pair then
// cuts the bra's tiles and segments into runs of at most `cap` candidate kets.
// The lists are those of a sequential walk over (bra group, bra, ket group)
// (Fock) or (bra group, ket group, bra) (gradient), element for element.
#include <hipcub/hipcub.hpp>

#include "qk_internal.h"

namespace qk {

namespace {

__device__ __forceinline__ int find_entry(const ListEntry* E, int ne, long long item, bool by_bra) {
    int lo = 0, hi = ne - 1;
    while (lo < hi) {   // last entry whose first item (bra) is <= item
        const int mid = (lo + hi + 1) >> 1;
        const long long f = by_bra ? E[mid].bra0 : E[mid].item0;
        if (f <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

struct Item {
    int p, s2, qb, lo, lo64;
};

// the bra pair, the ket group and the ket ranges of item `it` of entry `e`
__device__ __forceinline__ Item decode(const ListEntry& E, const KetGroup* KG, const ListClass& C, const double* Y,
                                       double tau, double thr, long long r) {
    const int i = (int)(r / E.m), t = (int)(r - (long long)i * E.m);
    const KetGroup K = KG[E.kg0 + t];
    Item it;
    it.p = E.s1 + i;
    it.s2 = K.s2;
    it.qb = K.same ? i : 0;
    const double yp = Y[it.p];
    int lo = 0, hi = K.n2;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (yp * Y[K.s2 + mid] > tau) lo = mid + 1; else hi = mid;
    }
    it.lo = lo;
    it.lo64 = lo;
    if (C.split) {
        int a0 = 0, a1 = lo;
        while (a0 < a1) {
            const int mid = (a0 + a1) >> 1;
            if (yp * Y[K.s2 + mid] >= thr) a0 = mid + 1; else a1 = mid;
        }
        it.lo64 = a0 > it.qb ? a0 : it.qb;
    }
    return it;
}

__device__ __forceinline__ long long ceil_div(long long a, long long b) { return a > 0 ? (a + b - 1) / b : 0; }

// per item: FP64 / FP32 tiles, segments and candidates (kListCounts values)
__global__ void list_count_kernel(const ListEntry* __restrict__ E, int ne, const KetGroup* __restrict__ KG,
                                  const ListClass* __restrict__ LC, const double* __restrict__ Y, double tau,
                                  double thr, long long N, int4* __restrict__ items, long long* __restrict__ cnt) {
    for (long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x; it <= N;
         it += (long long)gridDim.x * blockDim.x) {
        if (it == N) {   // the trailing zero of each count array: its exclusive scan is the total
            for (int a = 0; a < kListCounts; ++a) cnt[a * (N + 1) + N] = 0;
            continue;
        }
        const int e = find_entry(E, ne, it, false);
        const ListClass C = LC[E[e].cls];
        const Item x = decode(E[e], KG, C, Y, tau, thr, it - E[e].item0);
        items[it] = make_int4(x.qb, x.lo, x.lo64, 0);
        const long long c64 = x.lo64 > x.qb ? x.lo64 - x.qb : 0;
        const long long c32 = x.lo > x.lo64 ? x.lo - x.lo64 : 0;
        long long v[kListCounts];
        v[0] = ceil_div(c64, C.chunk);
        v[1] = ceil_div(c32, C.chunk);
        if (!C.segs) v[2] = v[3] = 0;
        else if (C.grad) { v[2] = c64 > 0; v[3] = c32 > 0; }
        else { v[2] = ceil_div(c64, C.cap64); v[3] = ceil_div(c32, C.cap32); }
        v[4] = c64;
        v[5] = c32;
        for (int a = 0; a < kListCounts; ++a) cnt[a * (N + 1) + it] = v[a];
    }
}

// per class: the sums of each count (from the exclusive scans at the class's bounds)
__global__ void list_sums_kernel(const long long* __restrict__ off, long long N, const long long* __restrict__ cls_item,
                                 int ncls, int ncnt, long long* __restrict__ sums) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncls) return;
    for (int a = 0; a < ncnt; ++a)
        sums[k * ncnt + a] = off[a * (N + 1) + cls_item[k + 1]] - off[a * (N + 1) + cls_item[k]];
}

__global__ void list_write_kernel(const ListEntry* __restrict__ E, int ne, const KetGroup* __restrict__ KG,
                                  const ListClass* __restrict__ LC, long long N, const int4* __restrict__ items,
                                  const long long* __restrict__ off, const long long* __restrict__ cls_item,
                                  const ListDst* __restrict__ dst) {
    for (long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x; it < N;
         it += (long long)gridDim.x * blockDim.x) {
        const int e = find_entry(E, ne, it, false);
        const int k = E[e].cls;
        const ListClass C = LC[k];
        const long long r = it - E[e].item0;
        const int i = (int)(r / E[e].m), t = (int)(r - (long long)i * E[e].m);
        const int p = E[e].s1 + i, s2 = KG[E[e].kg0 + t].s2;
        const int4 x = items[it];
        const int qb = x.x, lo = x.y, lo64 = x.z;
        const long long c0 = cls_item[k];
        long long o[kListCounts];
        for (int a = 0; a < kListCounts; ++a) o[a] = off[a * (N + 1) + it] - off[a * (N + 1) + c0];
        const ListDst D = dst[k];
        long long j = o[0];
        for (int q = qb; q < lo64; q += C.chunk) D.t64[j++] = Tile{p, s2 + q, min(C.chunk, lo64 - q), 0};
        j = o[1];
        for (int q = lo64; q < lo; q += C.chunk) D.t32[j++] = Tile{p, s2 + q, min(C.chunk, lo - q), 0};
        if (!C.segs) continue;
        if (C.grad) {
            if (lo64 > qb) {
                D.s64[o[2]] = make_int4(p, s2 + qb, s2 + lo64, (int)o[4]);
                D.slot64[o[2]] = (int)o[0];
            }
            if (lo > lo64) {
                D.s32[o[3]] = make_int4(p, s2 + lo64, s2 + lo, (int)o[5]);
                D.slot32[o[3]] = (int)o[1];
            }
        } else {
            j = o[2];
            long long c = o[4];
            for (int q = qb; q < lo64;) {
                const int en = lo64 - q > C.cap64 ? q + C.cap64 : lo64;
                D.s64[j++] = make_int4(p, s2 + q, s2 + en, (int)c);
                c += en - q;
                q = en;
            }
            j = o[3];
            c = o[5];
            for (int q = lo64; q < lo;) {
                const int en = lo - q > C.cap32 ? q + C.cap32 : lo;
                D.s32[j++] = make_int4(p, s2 + q, s2 + en, (int)c);
                c += en - q;
                q = en;
            }
        }
    }
}

// Greedy runs over one bra's pieces (tiles of `chunk` kets, or segments of
// `cap` kets) in list order: a run ends before a piece that would take its
// candidates past cap.  emit(first, end) gets piece indices relative to `base`.
template <class Emit>
__device__ __forceinline__ long long bra_runs(const ListEntry& E, const int4* items, int i, bool fp32, int piece,
                                              int cap, Emit&& emit) {
    long long nrun = 0, idx = 0, first = -1;
    int cands = 0;
    for (int t = 0; t < E.m; ++t) {
        const int4 x = items[E.item0 + (long long)i * E.m + t];
        const int a = fp32 ? x.z : x.x, b = fp32 ? x.y : x.z;
        for (int q = a; q < b;) {
            const int len = min(piece, b - q);
            q += len;
            if (first < 0 || cands + len > cap) {
                if (first >= 0) emit(first, idx);
                first = idx;
                ++nrun;
                cands = 0;
            }
            cands += len;
            ++idx;
        }
    }
    if (first >= 0) emit(first, idx);
    return nrun;
}

// per bra pair of the Fock lists: runs of FP64 / FP32 tiles and segments
__global__ void list_runs_kernel(const ListEntry* __restrict__ E, int ne, const ListClass* __restrict__ LC,
                                 long long NB, const int4* __restrict__ items, const long long* __restrict__ off,
                                 long long N, const long long* __restrict__ cls_item, const long long* __restrict__ roff,
                                 const long long* __restrict__ cls_bra, const ListDst* __restrict__ dst,
                                 long long* __restrict__ rcnt) {
    for (long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x; b <= NB;
         b += (long long)gridDim.x * blockDim.x) {
        if (b == NB) {
            if (!roff)
                for (int a = 0; a < kRunCounts; ++a) rcnt[a * (NB + 1) + NB] = 0;
            continue;
        }
        const int e = find_entry(E, ne, b, true);
        const int k = E[e].cls;
        const ListClass C = LC[k];
        const int i = (int)(b - E[e].bra0);
        if (!C.runs) {
            if (!roff)
                for (int a = 0; a < kRunCounts; ++a) rcnt[a * (NB + 1) + b] = 0;
            continue;
        }
        if (!roff) {   // counting pass
            auto none = [](long long, long long) {};
            rcnt[0 * (NB + 1) + b] = bra_runs(E[e], items, i, false, C.chunk, C.cap64, none);
            rcnt[1 * (NB + 1) + b] = bra_runs(E[e], items, i, true, C.chunk, C.cap32, none);
            rcnt[2 * (NB + 1) + b] = C.segs ? bra_runs(E[e], items, i, false, C.cap64, C.cap64, none) : 0;
            rcnt[3 * (NB + 1) + b] = C.segs ? bra_runs(E[e], items, i, true, C.cap32, C.cap32, none) : 0;
            continue;
        }
        // writing pass: pieces numbered from the bra's first item's list offsets
        const long long it0 = E[e].item0 + (long long)i * E[e].m;
        const ListDst D = dst[k];
        const long long c0 = cls_item[k], b0 = cls_bra[k];
        for (int a = 0; a < kRunCounts; ++a) {
            if (a >= 2 && !C.segs) break;
            const long long base = off[a * (N + 1) + it0] - off[a * (N + 1) + c0];
            long long j = roff[a * (NB + 1) + b] - roff[a * (NB + 1) + b0];
            int2* out = a == 0 ? D.r64 : a == 1 ? D.r32 : a == 2 ? D.sr64 : D.sr32;
            auto put = [&](long long f, long long en) { out[j++] = make_int2((int)(base + f), (int)(base + en)); };
            const bool fp32 = a & 1;
            const int piece = a < 2 ? C.chunk : (fp32 ? C.cap32 : C.cap64);
            bra_runs(E[e], items, i, fp32, piece, fp32 ? C.cap32 : C.cap64, put);
        }
    }
}

int grid_for(long long n) {
    const long long g = (n + 255) / 256;
    return (int)(g < 1 ? 1 : (g > 256LL * 64 ? 256LL * 64 : g));
}

}  // namespace

size_t list_scan_bytes(long long n) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const long long*)nullptr, (long long*)nullptr, (int)n);
    return bytes;
}

int lists_count(const ListEntry* E, int ne, const KetGroup* KG, const ListClass* LC, const double* Y, double tau,
                double thr, long long N, int4* items, long long* cnt, long long* off, void* tmp, size_t tmp_bytes,
                const long long* cls_item, int ncls, long long* sums, hipStream_t s) {
    hipLaunchKernelGGL(list_count_kernel, dim3(grid_for(N + 1)), dim3(256), 0, s, E, ne, KG, LC, Y, tau, thr, N, items,
                       cnt);
    for (int a = 0; a < kListCounts; ++a) {
        size_t b = tmp_bytes;
        if (hipcub::DeviceScan::ExclusiveSum(tmp, b, cnt + a * (N + 1), off + a * (N + 1), (int)(N + 1), s) !=
            hipSuccess)
            return -1;
    }
    hipLaunchKernelGGL(list_sums_kernel, dim3((ncls + 63) / 64), dim3(64), 0, s, off, N, cls_item, ncls, kListCounts,
                       sums);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

void lists_write(const ListEntry* E, int ne, const KetGroup* KG, const ListClass* LC, long long N, const int4* items,
                 const long long* off, const long long* cls_item, const ListDst* dst, hipStream_t s) {
    if (N > 0)
        hipLaunchKernelGGL(list_write_kernel, dim3(grid_for(N)), dim3(256), 0, s, E, ne, KG, LC, N, items, off,
                           cls_item, dst);
}

int lists_runs_count(const ListEntry* E, int ne, const ListClass* LC, long long NB, const int4* items,
                     long long* rcnt, long long* roff, void* tmp, size_t tmp_bytes, const long long* cls_bra, int ncls,
                     long long* sums, hipStream_t s) {
    hipLaunchKernelGGL(list_runs_kernel, dim3(grid_for(NB + 1)), dim3(256), 0, s, E, ne, LC, NB, items,
                       (const long long*)nullptr, 0LL, (const long long*)nullptr, (const long long*)nullptr,
                       (const long long*)nullptr, (const ListDst*)nullptr, rcnt);
    for (int a = 0; a < kRunCounts; ++a) {
        size_t b = tmp_bytes;
        if (hipcub::DeviceScan::ExclusiveSum(tmp, b, rcnt + a * (NB + 1), roff + a * (NB + 1), (int)(NB + 1), s) !=
            hipSuccess)
            return -1;
    }
    hipLaunchKernelGGL(list_sums_kernel, dim3((ncls + 63) / 64), dim3(64), 0, s, roff, NB, cls_bra, ncls, kRunCounts,
                       sums);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

void lists_runs_write(const ListEntry* E, int ne, const ListClass* LC, long long NB, const int4* items,
                      const long long* off, long long N, const long long* cls_item, const long long* roff,
                      const long long* cls_bra, const ListDst* dst, hipStream_t s) {
    if (NB > 0)
        hipLaunchKernelGGL(list_runs_kernel, dim3(grid_for(NB + 1)), dim3(256), 0, s, E, ne, LC, NB, items, off, N,
                           cls_item, roff, cls_bra, dst, (long long*)nullptr);
}

// primitive-major copy of the significant pairs' primitive data: entry j of
// sorted pair k at prim0[k] + j * pstride[k], from the all-pairs arrays at
// src0[sel[k]] + j; FP32 copies (round to nearest) when f32 is given
__global__ void gather_prims_kernel(const int* __restrict__ sel, const int* __restrict__ src0,
                                    const int* __restrict__ nprim, const int* __restrict__ prim0,
                                    const int* __restrict__ pstride, int npair, PrimArrays src, PrimArrays dst,
                                    PrimArraysF f32) {
    const long long total = (long long)npair * 64;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(g >> 6), j = (int)(g & 63);
        if (j >= nprim[k]) continue;
        const long long sidx = (long long)src0[sel[k]] + j, didx = (long long)prim0[k] + (long long)j * pstride[k];
        for (int a = 0; a < kPrimArrays; ++a) {
            const double v = src.a[a][sidx];
            dst.a[a][didx] = v;
            if (f32.a[0]) f32.a[a][didx] = (float)v;
        }
    }
}

void launch_gather_prims(const int* sel, const int* src0, const int* nprim, const int* prim0, const int* pstride,
                         int npair, const PrimArrays& src, const PrimArrays& dst, const PrimArraysF& f32,
                         hipStream_t s) {
    if (npair > 0)
        hipLaunchKernelGGL(gather_prims_kernel, dim3(grid_for((long long)npair * 64)), dim3(256), 0, s, sel, src0,
                           nprim, prim0, pstride, npair, src, dst, f32);
}

}  // namespace qk
