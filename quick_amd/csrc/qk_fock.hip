// qk_fock.hip -- Fock-build kernels for MI355X (gfx950).
//
// One kernel instantiation per quartet class (la lb|lc ld).  Each thread owns
// one contracted Qshell quartet at a time: it screens it (Schwarz x density,
// hfoperator.f90:725-736 / gpu_get2e_subs.h:323-328), runs the primitive loop
// (Boys F_m(T) with QUICK's fit, VRR accumulated into the contracted [e0|f0]),
// applies the HRR, and digests the quartet's integrals into the Fock
// accumulator with the 8-fold-symmetry weights.  The accumulator is int64
// fixed point (2^-48 units); integer atomics make the result independent of
// the order quartets finish in, so every build is bitwise reproducible.
//
// Work layout: the host cuts each class's quartet space into tiles {bra pair,
// first ket, length <= 64} (bra and kets from the class- and bound-sorted pair
// lists, so the kets of a tile are consecutive and all pass the Schwarz test).
// A wavefront takes one tile at a time: the bra's pair and primitive data are
// broadcast loads and the kets' primitive-major data are coalesced.  Where a
// block takes one bra pair at a time (exchange rows in LDS), the density
// screen runs first in compact_kernel, which packs each segment's surviving
// kets, so that the class kernel's lanes are all busy (about 40% of the Schwarz
// candidates fail the density test, scattered over the tiles).  The heavy d
// classes use the wavefront-cooperative kernel (qk_team.h) instead.
#include "qk_kernels.h"

#include <mutex>

namespace qk {

// ---------------------------------------------------------------------------
// helper kernels

// Density screen of the candidate quartets, compacted: one wavefront per
// segment (a bra pair and a ket range of one group) tests its candidate kets
// 64 at a time and appends the survivors in order (ballot + prefix count) at
// list[segs[s].w ...]; counts[s] gets their number.  The class kernels then
// fill all 64 lanes with surviving quartets.  The test is the class kernels'
// own, same operation order: Fock Y_p Y_q DNmax > tau; gradient (one-centre
// quartets skipped) Y_p Y_q > tau and Y_p Y_q DNmax > tau.  With a.steps set,
// the survivors are also cut into steps of 64, written to the slots of the
// segment's tiles (same order as the tiles; the gradient's light kernels
// stride over them).
__global__ void __launch_bounds__(256) compact_kernel(const CompactArgs a, const int4* __restrict__ segs,
                                                      long long nsegs, int* __restrict__ list, int* __restrict__ counts,
                                                      int stage_rows) {
    extern __shared__ double s_cmrows[];
    const long long s = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (s >= nsegs) return;   // the whole wavefront
    const int4 sg = segs[s];
    const int shp = a.psh[sg.x], atp = a.pat[sg.x];
    const int sA = shp & 0xffff, sB = shp >> 16;
    const int atA = atp & 0xffff, atB = atp >> 16;
    // with stage_rows, the segment's bra shell rows A and B of the density
    // maxima in LDS (2 nshell values per wavefront): the four cross terms of
    // each candidate's test read them there instead of gathering from memory
    const int ns = a.nshell;
    double* rA = s_cmrows + (size_t)(threadIdx.x >> 6) * 2 * ns;
    double* rB = rA + ns;
    if (stage_rows) {
        for (int j = lane; j < ns; j += 64) {
            rA[j] = a.cutmat[(long long)sA * ns + j];
            rB[j] = a.cutmat[(long long)sB * ns + j];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const double m_ab = 4.0 * a.cutmat[(long long)sA * ns + sB];
    const double yp = a.Y[sg.x];
    int cnt = 0;
    for (int q0 = sg.y; q0 < sg.z; q0 += 64) {
        const int q = q0 + lane;
        bool keep = false;
        if (q < sg.z) {
            const int shq = a.psh[q];
            const int sC = shq & 0xffff, sD = shq >> 16;
            // dn_max's value (fmax picks one of its operands, so the grouping
            // does not change it)
            const double dn = stage_rows
                                  ? fmax(fmax(m_ab, 4.0 * a.cutmat[(long long)sC * ns + sD]),
                                         fmax(fmax(rA[sD], rA[sC]), fmax(rB[sC], rB[sD])))
                                  : dn_max(a.cutmat, ns, sA, sB, sC, sD);
            if (a.grad) {
                const int atq = a.pat[q];
                const int atC = atq & 0xffff, atD = atq >> 16;
                const double ct = a.Y[q] * yp;
                keep = !(atA == atB && atA == atC && atA == atD) && ct > a.tau && ct * dn > a.tau;
            } else {
                keep = a.Y[q] * yp * dn > a.tau;
            }
        }
        const unsigned long long bal = __ballot(keep);
        if (keep) list[sg.w + cnt + __popcll(bal & ((1ull << lane) - 1ull))] = q;
        cnt += __popcll(bal);
    }
    if (lane == 0) counts[s] = cnt;
    if (a.steps != nullptr) {
        // this segment's slots: its steps of 64 survivors, then empty ones
        const int nslot = (sg.z - sg.y + 63) >> 6, base = a.slot0[s];
        for (int j = lane; j < nslot; j += 64)
            a.steps[base + j] = Tile{sg.x, sg.w + 64 * j, max(0, min(64, cnt - 64 * j)), 0};
    }
}

// Dp = diag(cons) D diag(cons): folds QUICK's per-function cons factors
// (applied inside hrrwhole, subs/hrr.f90:261-295) into the density; the
// operator is scaled back by the same factors in fock_finalize.
// Dpf (optional) receives the same matrix rounded to FP32 for the FP32 kernels.
__global__ void scale_density(const double* __restrict__ D, const double* __restrict__ cons, double* __restrict__ Dp,
                              float* __restrict__ Dpf, int n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)n * n) return;
    const int i = int(k / n), j = int(k % n);
    const double v = cons[i] * cons[j] * D[k];
    Dp[k] = v;
    if (Dpf) Dpf[k] = float(v);
}

// cutmatrix(II,JJ) = max |D| over the (II,JJ) shell block (schwarz.f90:550-599)
__global__ void density_cutoff(const double* __restrict__ D, const int* __restrict__ sh_f0,
                               const int* __restrict__ sh_nf, double* __restrict__ cm, int n, int ns) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)ns * ns) return;
    const int I = int(k / ns), J = int(k % ns);
    double m = 0.0;
    for (int i = sh_f0[I]; i < sh_f0[I] + sh_nf[I]; ++i)
        for (int j = sh_f0[J]; j < sh_f0[J] + sh_nf[J]; ++j) m = fmax(m, fabs(D[(long long)j * n + i]));
    cm[k] = m;
}

__global__ void fill_u64(unsigned long long* p, long long count) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) p[k] = 0ull;
}

// G = diag(cons) (X + X^T)/2 diag(cons) in FP64, X the sum of the copies.
// A sum within a factor of two of the int64 range (|T| >= 2^62, i.e. |G| of
// about 8192 and more) raises the range flag: the accumulator would wrap at
// twice that.
__global__ void fock_finalize(const unsigned long long* __restrict__ X, int copies, const double* __restrict__ cons,
                              double* __restrict__ G, int n, unsigned long long* __restrict__ flag) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)n * n) return;
    const int i = int(k / n), j = int(k % n);
    const long long nn = (long long)n * n, t = (long long)i * n + j, tt = (long long)j * n + i;
    unsigned long long u = 0;
    for (int c = 0; c < copies; ++c) u += X[c * nn + t] + X[c * nn + tt];
    const long long s = (long long)u;
    if (s >= (1LL << 62) || s <= -(1LL << 62)) atomicOr(flag, 1ull);
    G[k] = cons[i] * cons[j] * ((double)s * 0x1p-49);
}

// The pair-indexed Coulomb buffer into X (copy 0): one thread per significant
// pair k, element (ia, ib) of its block at J[jofs[k] + (ia NB + ib) pstride[k]]
// to X[f0(qa) + ia][f0(qb) + ib].  Each function pair belongs to one Qshell
// pair, so the adds of different threads never meet.
__global__ void jpair_scatter(const unsigned long long* __restrict__ J, const int* __restrict__ jofs,
                              const int* __restrict__ pstride, const int* __restrict__ qa,
                              const int* __restrict__ qb, const DevQshells q, int npair,
                              unsigned long long* __restrict__ X, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npair) return;
    const int a = qa[k], b = qb[k];
    const int fa = q.f0[a], fb = q.f0[b], na = q.nc[a], nb = q.nc[b];
    const long long o = jofs[k], st = pstride[k];
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) {
            const unsigned long long v = J[o + (ia * nb + ib) * st];
            if (v) atomicAdd(X + (long long)(fa + ia) * n + fb + ib, v);
        }
}

// ---------------------------------------------------------------------------
// launchers

using FockLauncher = void (*)(const FockArgs&, const ClassWork&, int, size_t, hipStream_t);
using TeamLauncher = void (*)(const FockArgs&, const ClassWork&, int, hipStream_t);
using SchwarzLauncher = void (*)(const int*, int, const int*, const int*, const int*, const int*, const DevQshells&,
                                 const DevPrims&, const double*, const double*, double*, double*, hipStream_t);
#define QK_DECL(T) void launch_fock_##T(const FockArgs&, const ClassWork&, int, size_t, hipStream_t);
QK_DECL(ssss) QK_DECL(psss) QK_DECL(psps) QK_DECL(ppss) QK_DECL(ppps) QK_DECL(pppp) QK_DECL(dsss)
QK_DECL(dsps) QK_DECL(dspp) QK_DECL(dsds) QK_DECL(dpss) QK_DECL(dpps) QK_DECL(dppp) QK_DECL(dpds)
QK_DECL(dpdp) QK_DECL(ddss) QK_DECL(ddps) QK_DECL(ddpp) QK_DECL(ddds) QK_DECL(dddp) QK_DECL(dddd)
#undef QK_DECL
#define QK_DECL(T) size_t fock_scratch_##T();
QK_DECL(ssss) QK_DECL(psss) QK_DECL(psps) QK_DECL(ppss) QK_DECL(ppps) QK_DECL(pppp) QK_DECL(dsss)
QK_DECL(dsps) QK_DECL(dspp) QK_DECL(dsds) QK_DECL(dpss) QK_DECL(dpps) QK_DECL(dppp) QK_DECL(dpds)
QK_DECL(dpdp) QK_DECL(ddss) QK_DECL(ddps) QK_DECL(ddpp) QK_DECL(ddds) QK_DECL(dddp) QK_DECL(dddd)
#undef QK_DECL
#define QK_DECL(T)                                                                                              \
    void launch_schwarz_##T(const int*, int, const int*, const int*, const int*, const int*, const DevQshells&, \
                            const DevPrims&, const double*, const double*, double*, double*, hipStream_t);
QK_DECL(ss) QK_DECL(ps) QK_DECL(pp) QK_DECL(ds) QK_DECL(dp) QK_DECL(dd)
#undef QK_DECL

// order matches gen_eri.quartet_classes()
static const FockLauncher kFockLaunchers[kNumQuartetClasses] = {
    launch_fock_ssss, launch_fock_psss, launch_fock_psps, launch_fock_ppss, launch_fock_ppps, launch_fock_pppp,
    launch_fock_dsss, launch_fock_dsps, launch_fock_dspp, launch_fock_dsds, launch_fock_dpss, launch_fock_dpps,
    launch_fock_dppp, launch_fock_dpds, launch_fock_dpdp, launch_fock_ddss, launch_fock_ddps, launch_fock_ddpp,
    launch_fock_ddds, launch_fock_dddp, launch_fock_dddd};
static const SchwarzLauncher kSchwarzLaunchers[kNumPairClasses] = {
    launch_schwarz_ss, launch_schwarz_ps, launch_schwarz_pp, launch_schwarz_ds, launch_schwarz_dp, launch_schwarz_dd};

#define QK_DECL(T)                                                                   \
    void launch_team_##T(const FockArgs&, const ClassWork&, int, hipStream_t); \
    void launch_team_f32_##T(const FockArgs&, const ClassWork&, int, hipStream_t);
QK_DECL(dpdp) QK_DECL(ddpp) QK_DECL(ddds) QK_DECL(dddp) QK_DECL(dddd)
#undef QK_DECL
// classes evaluated by the wavefront-cooperative kernel (one quartet per wavefront)
static const TeamLauncher kTeamLaunchers[kNumQuartetClasses] = {
    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
    nullptr, nullptr, launch_team_dpdp, nullptr, nullptr, launch_team_ddpp, launch_team_ddds, launch_team_dddp,
    launch_team_dddd};
static const TeamLauncher kTeamF32Launchers[kNumQuartetClasses] = {
    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
    nullptr, nullptr, launch_team_f32_dpdp, nullptr, nullptr, launch_team_f32_ddpp, launch_team_f32_ddds,
    launch_team_f32_dddp, launch_team_f32_dddd};
bool fock_class_is_team(int k) { return kTeamLaunchers[k] != nullptr; }
void launch_team_class(int k, const FockArgs& a, const ClassWork& w, int grid, hipStream_t s) {
    kTeamLaunchers[k](a, w, grid, s);
}
void launch_team_f32_class(int k, const FockArgs& a, const ClassWork& w, int grid, hipStream_t s) {
    kTeamF32Launchers[k](a, w, grid, s);
}

#define QK_DECL(T) int fock_blocks_##T(size_t, bool);
QK_DECL(ssss) QK_DECL(psss) QK_DECL(psps) QK_DECL(ppss) QK_DECL(ppps) QK_DECL(pppp) QK_DECL(dsss)
QK_DECL(dsps) QK_DECL(dspp) QK_DECL(dsds) QK_DECL(dpss) QK_DECL(dpps) QK_DECL(dppp) QK_DECL(dpds)
QK_DECL(dpdp) QK_DECL(ddss) QK_DECL(ddps) QK_DECL(ddpp) QK_DECL(ddds) QK_DECL(dddp) QK_DECL(dddd)
#undef QK_DECL
using BlocksQuery = int (*)(size_t, bool);
static const BlocksQuery kFockBlocks[kNumQuartetClasses] = {
    fock_blocks_ssss, fock_blocks_psss, fock_blocks_psps, fock_blocks_ppss, fock_blocks_ppps, fock_blocks_pppp,
    fock_blocks_dsss, fock_blocks_dsps, fock_blocks_dspp, fock_blocks_dsds, fock_blocks_dpss, fock_blocks_dpps,
    fock_blocks_dppp, fock_blocks_dpds, fock_blocks_dpdp, fock_blocks_ddss, fock_blocks_ddps, fock_blocks_ddpp,
    fock_blocks_ddds, fock_blocks_dddp, fock_blocks_dddd};
// resident blocks of 256 threads per CU of the FP64 thread-per-quartet kernel
// of class k (its staged form if staged) with lds bytes of dynamic LDS
// (registers, LDS and wave limits)
int fock_class_blocks(int k, size_t lds, bool staged) { return kFockBlocks[k](lds, staged); }

using ScratchQuery = size_t (*)();
static const ScratchQuery kFockScratch[kNumQuartetClasses] = {
    fock_scratch_ssss, fock_scratch_psss, fock_scratch_psps, fock_scratch_ppss, fock_scratch_ppps, fock_scratch_pppp,
    fock_scratch_dsss, fock_scratch_dsps, fock_scratch_dspp, fock_scratch_dsds, fock_scratch_dpss, fock_scratch_dpps,
    fock_scratch_dppp, fock_scratch_dpds, fock_scratch_dpdp, fock_scratch_ddss, fock_scratch_ddps, fock_scratch_ddpp,
    fock_scratch_ddds, fock_scratch_dddp, fock_scratch_dddd};

#define QK_DECL(T) void launch_fock_f32_##T(const FockArgs&, const ClassWork&, int, size_t, hipStream_t); size_t fock_scratch_f32_##T();
QK_DECL(ssss) QK_DECL(psss) QK_DECL(psps) QK_DECL(ppss) QK_DECL(ppps) QK_DECL(pppp) QK_DECL(dsss)
QK_DECL(dsps) QK_DECL(dspp) QK_DECL(dsds) QK_DECL(dpss) QK_DECL(dpps) QK_DECL(dppp) QK_DECL(dpds)
QK_DECL(ddss) QK_DECL(ddps)
#undef QK_DECL
// FP32 kernels of the mixed-precision mode; the cooperative classes stay FP64
static const FockLauncher kFockF32Launchers[kNumQuartetClasses] = {
    launch_fock_f32_ssss, launch_fock_f32_psss, launch_fock_f32_psps, launch_fock_f32_ppss, launch_fock_f32_ppps,
    launch_fock_f32_pppp, launch_fock_f32_dsss, launch_fock_f32_dsps, launch_fock_f32_dspp, launch_fock_f32_dsds,
    launch_fock_f32_dpss, launch_fock_f32_dpps, launch_fock_f32_dppp, launch_fock_f32_dpds, nullptr,
    launch_fock_f32_ddss, launch_fock_f32_ddps, nullptr, nullptr, nullptr, nullptr};
static const ScratchQuery kFockF32Scratch[kNumQuartetClasses] = {
    fock_scratch_f32_ssss, fock_scratch_f32_psss, fock_scratch_f32_psps, fock_scratch_f32_ppss, fock_scratch_f32_ppps,
    fock_scratch_f32_pppp, fock_scratch_f32_dsss, fock_scratch_f32_dsps, fock_scratch_f32_dspp, fock_scratch_f32_dsds,
    fock_scratch_f32_dpss, fock_scratch_f32_dpps, fock_scratch_f32_dppp, fock_scratch_f32_dpds, nullptr,
    fock_scratch_f32_ddss, fock_scratch_f32_ddps, nullptr, nullptr, nullptr, nullptr};
bool fock_class_has_f32(int k) { return kFockF32Launchers[k] != nullptr || kTeamF32Launchers[k] != nullptr; }
void launch_fock_f32_class(int k, const FockArgs& a, const ClassWork& w, int grid, size_t lds, hipStream_t s) {
    kFockF32Launchers[k](a, w, grid, lds, s);
}
size_t fock_class_scratch_f32(int k) { return kFockF32Scratch[k] ? kFockF32Scratch[k]() : 0; }

void launch_fock_class(int k, const FockArgs& a, const ClassWork& w, int grid, size_t lds, hipStream_t s) {
    kFockLaunchers[k](a, w, grid, lds, s);
}
size_t fock_class_scratch(int k) { return kFockScratch[k](); }
void launch_schwarz(int pair_class, const int* pairs, int np, const int* qa, const int* qb, const int* prim0,
                    const int* nprim, const DevQshells& qs, const DevPrims& pp, const double* xyz, const double* cons,
                    double* ymax, double* pmax, hipStream_t s) {
    if (np > 0) kSchwarzLaunchers[pair_class](pairs, np, qa, qb, prim0, nprim, qs, pp, xyz, cons, ymax, pmax, s);
}

static inline int blocks_for(long long count, int bs) { return int((count + bs - 1) / bs); }

void launch_scale_density(const double* D, const double* cons, double* Dp, float* Dpf, int n, hipStream_t s) {
    hipLaunchKernelGGL(scale_density, dim3(blocks_for((long long)n * n, 256)), dim3(256), 0, s, D, cons, Dp, Dpf, n);
}
void launch_density_cutoff(const double* D, const int* sh_f0, const int* sh_nf, double* cm, int n, int ns,
                           hipStream_t s) {
    hipLaunchKernelGGL(density_cutoff, dim3(blocks_for((long long)ns * ns, 256)), dim3(256), 0, s, D, sh_f0, sh_nf,
                       cm, n, ns);
}
void launch_compact(const CompactArgs& a, const int4* segs, long long nsegs, int* list, int* counts, hipStream_t s) {
    if (nsegs <= 0) return;
    // rows staged while the block's four wavefronts need at most 64 KB
    // (nshell <= 1024)
    const size_t rows = sizeof(double) * 2 * 4 * (size_t)a.nshell;
    const int stage = rows <= 64 * 1024 ? 1 : 0;
    hipLaunchKernelGGL(compact_kernel, dim3((unsigned)((nsegs + 3) / 4)), dim3(256), stage ? rows : 0, s, a, segs,
                       nsegs, list, counts, stage);
}
void launch_fill_u64(unsigned long long* p, long long count, hipStream_t s) {
    if (count > 0) hipLaunchKernelGGL(fill_u64, dim3(blocks_for(count, 256)), dim3(256), 0, s, p, count);
}
void launch_jpair_scatter(const unsigned long long* J, const int* jofs, const int* pstride, const int* qa,
                          const int* qb, const DevQshells& q, int npair, unsigned long long* X, int n,
                          hipStream_t s) {
    if (npair > 0)
        hipLaunchKernelGGL(jpair_scatter, dim3(blocks_for(npair, 256)), dim3(256), 0, s, J, jofs, pstride, qa, qb, q,
                           npair, X, n);
}
void launch_fock_finalize(const unsigned long long* X, int copies, const double* cons, double* G, int n,
                          unsigned long long* flag, hipStream_t s) {
    hipLaunchKernelGGL(fock_finalize, dim3(blocks_for((long long)n * n, 256)), dim3(256), 0, s, X, copies, cons, G, n,
                       flag);
}

// ---------------------------------------------------------------------------
// the no-allocation entry point's scratch

// The device array itself.  In host code `scratch` names a host-side stand-in,
// not device memory: host code uses only the address noalloc_scratch_init
// looked up for the device (kernel arguments, fills, copies).
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};
int g_scratch_rc[kMaxScratchDevices] = {};
bool g_scratch_done[kMaxScratchDevices] = {};
}  // namespace

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_done[device]) return g_scratch_rc[device];
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
    g_scratch_done[device] = true;
    g_scratch_rc[device] = (int)e;
    g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    return (int)e;
}

void* noalloc_scratch_ptr(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace qk
