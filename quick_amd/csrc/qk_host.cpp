// qk_host.cpp -- host side of the MI355X Fock-build library: the context, the
// class/bound-sorted pair lists (the device work queue), the launches, and the
// C-ABI (reference gpu_*_ entry points and the explicit qk_* API).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/quick_mi355x.h"
#include "qk_internal.h"
#include "eri_flops.h"
#include "grad_tables.inc"

namespace qk {
void launch_fock_class(int k, const FockArgs& a, const ClassWork& w, int grid, size_t lds, hipStream_t s);
size_t fock_class_scratch(int k);
int fock_class_blocks(int k, size_t lds, bool staged);
bool fock_class_is_team(int k);
bool fock_class_has_f32(int k);
void launch_fock_f32_class(int k, const FockArgs& a, const ClassWork& w, int grid, size_t lds, hipStream_t s);
size_t fock_class_scratch_f32(int k);
void launch_team_class(int k, const FockArgs& a, const ClassWork& w, int grid, hipStream_t s);
void launch_team_f32_class(int k, const FockArgs& a, const ClassWork& w, int grid, hipStream_t s);
void launch_scale_density(const double* D, const double* cons, double* Dp, float* Dpf, int n, hipStream_t s);
size_t grad_lds_bytes(const int* hdr, bool fp32);
int launch_grad_class(const GradArgs& a, const ClassWork& w, const int* hdr, const int* dtab, size_t lds, int grid,
                      bool fp32, bool four_waves, hipStream_t s);
void launch_grad_finalize(const unsigned long long* G, double* grad, int m, unsigned long long* flag, hipStream_t s);
bool grad_class_is_light(int k);
int launch_grad_light(int k, const GradArgs& a, const ClassWork& w, int grid, bool fp32, hipStream_t s);
void launch_density_cutoff(const double* D, const int* sh_f0, const int* sh_nf, double* cm, int n, int ns,
                           hipStream_t s);
void launch_fill_u64(unsigned long long* p, long long count, hipStream_t s);
void launch_compact(const CompactArgs& a, const int4* segs, long long nsegs, int* list, int* counts, hipStream_t s);
void launch_oneint(const OneFuncs& f, const OneSys& sys, const int2* const* pairs, const long long* npairs, double* S,
                   double* H, hipStream_t s);
void launch_oneint_grad(const OneFuncs& f, const OneSys& sys, const int2* const* pairs, const long long* npairs,
                        const double* D, const double* W,
                        unsigned long long* G, hipStream_t s);
void launch_jpair_scatter(const unsigned long long* J, const int* jofs, const int* pstride, const int* qa,
                          const int* qb, const DevQshells& q, int npair, unsigned long long* X, int n,
                          hipStream_t s);
void launch_fock_finalize(const unsigned long long* X, int copies, const double* cons, double* G, int n,
                          unsigned long long* flag, hipStream_t s);
size_t list_scan_bytes(long long n);
int lists_count(const ListEntry* E, int ne, const KetGroup* KG, const ListClass* LC, const double* Y, double tau,
                double thr, long long N, int4* items, long long* cnt, long long* off, void* tmp, size_t tmp_bytes,
                const long long* cls_item, int ncls, long long* sums, hipStream_t s);
void lists_write(const ListEntry* E, int ne, const KetGroup* KG, const ListClass* LC, long long N, const int4* items,
                 const long long* off, const long long* cls_item, const ListDst* dst, hipStream_t s);
int lists_runs_count(const ListEntry* E, int ne, const ListClass* LC, long long NB, const int4* items,
                     long long* rcnt, long long* roff, void* tmp, size_t tmp_bytes, const long long* cls_bra, int ncls,
                     long long* sums, hipStream_t s);
void lists_runs_write(const ListEntry* E, int ne, const ListClass* LC, long long NB, const int4* items,
                      const long long* off, long long N, const long long* cls_item, const long long* roff,
                      const long long* cls_bra, const ListDst* dst, hipStream_t s);
void launch_gather_prims(const int* sel, const int* src0, const int* nprim, const int* prim0, const int* pstride,
                         int npair, const PrimArrays& src, const PrimArrays& dst, const PrimArraysF& f32,
                         hipStream_t s);
void launch_schwarz(int pair_class, const int* pairs, int np, const int* qa, const int* qb, const int* prim0,
                    const int* nprim, const DevQshells& qs, const DevPrims& pp, const double* xyz, const double* cons,
                    double* ymax, double* pmax, hipStream_t s);
}  // namespace qk

using namespace qk;

// ---------------------------------------------------------------------------
// errors

static thread_local std::string g_err;
static void set_err(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}
extern "C" const char* qk_last_error(void) { return g_err.c_str(); }
// the error text of the other translation units (qk_xc.hip)
void qk_set_error_text(const char* msg) { g_err = msg; }

#define HIPCHK(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            set_err("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

// ---------------------------------------------------------------------------
// device buffer helper

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;     // elements in use
    size_t cap = 0;   // elements allocated: a smaller count reuses the buffer
                      // (work lists are rebuilt per option set; hipFree and
                      // hipMalloc of hundreds of buffers each time cost more)
    bool owned = true;   // false: a view into memory owned elsewhere (an arena)
    // a view of count elements at q (not freed here)
    void view(T* q, size_t count) {
        release();
        p = count ? q : nullptr;
        n = count;
        owned = false;
    }
    int alloc(size_t count) {
        if (count == 0) { release(); return 0; }
        if (p && count <= cap) { n = count; return 0; }
        release();
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            set_err("hipMalloc of %zu bytes failed", count * sizeof(T));
            return -1;
        }
        n = cap = count;
        return 0;
    }
    int upload(const std::vector<T>& v) {
        if (alloc(v.size())) return -1;
        if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
            set_err("hipMemcpy H2D failed");
            return -1;
        }
        return 0;
    }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = cap = 0;
        owned = true;
    }
    ~DBuf() { release(); }
};

// Carves views out of one device allocation (sizes first, then place): the work
// lists of all classes in one hipMalloc, reused while large enough.
struct Arena {
    DBuf<unsigned char> mem;
    std::vector<size_t> off;
    size_t total = 0;
    void reset() { off.clear(); total = 0; }
    // reserve count elements of T; returns the slot
    template <typename T>
    int reserve(size_t count) {
        off.push_back(total);
        total += (count * sizeof(T) + 255) / 256 * 256;
        return (int)off.size() - 1;
    }
    int commit() { return mem.alloc(std::max<size_t>(total, 1)); }
    template <typename T>
    void place(DBuf<T>& b, int slot, size_t count) { b.view(reinterpret_cast<T*>(mem.p + off[slot]), count); }
};

// ---------------------------------------------------------------------------
// context

static const int kClassL[kNumQuartetClasses][4] = {
    {0, 0, 0, 0}, {1, 0, 0, 0}, {1, 0, 1, 0}, {1, 1, 0, 0}, {1, 1, 1, 0}, {1, 1, 1, 1}, {2, 0, 0, 0},
    {2, 0, 1, 0}, {2, 0, 1, 1}, {2, 0, 2, 0}, {2, 1, 0, 0}, {2, 1, 1, 0}, {2, 1, 1, 1}, {2, 1, 2, 0},
    {2, 1, 2, 1}, {2, 2, 0, 0}, {2, 2, 1, 0}, {2, 2, 1, 1}, {2, 2, 2, 0}, {2, 2, 2, 1}, {2, 2, 2, 2}};
static const int kPairL[kNumPairClasses][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};

static int ncart_of(int l) { return (l + 1) * (l + 2) / 2; }

static int pair_class_of(int la, int lb) {
    for (int k = 0; k < kNumPairClasses; ++k)
        if (kPairL[k][0] == la && kPairL[k][1] == lb) return k;
    return -1;
}

struct ClassHost {
    // FP64 tiles (the tiles of one bra pair consecutive) and FP32 tiles (mixed
    // precision only); runs: [first, end) tile of each bra pair
    long long ntiles = 0, ntiles32 = 0, nruns = 0, nruns32 = 0;
    long long total = 0;         // quartets passing the Schwarz test
    long long total32 = 0;       // of which in the FP32 lists
    DBuf<Tile> d_tiles, d_tiles32;
    DBuf<int2> d_runs, d_runs32;
    // the same candidates as segments {bra pair, first ket, end ket, offset}
    // (one bra pair and one ket group each), for the per-build density
    // compaction; segruns: [first, end) segment of each bra pair
    long long nsegs = 0, nsegs32 = 0, nsegruns = 0, nsegruns32 = 0;
    DBuf<int4> d_segs, d_segs32;
    DBuf<int2> d_segruns, d_segruns32;
    DBuf<int> d_list, d_list32, d_counts, d_counts32;
    // compaction pays where segments span several 64-ket steps (with short
    // segments every step is already the whole segment)
    bool compact = false, compact32 = false;
};

// Boys values F_m(k / kBoysTabPerUnit) for the FP32 kernels, from the series
// F_m(T) = exp(-T) sum_i (2T)^i / ((2m+1)(2m+3)...(2m+2i+1)) in FP64 (all terms
// positive, so it is accurate to a few ulp over the whole table).
static void boys_table(std::vector<float>& tab) {
    tab.assign((size_t)kBoysTabRows * kBoysTabM, 0.0f);
    for (int k = 0; k < kBoysTabRows; ++k) {
        const double T = double(k) / kBoysTabPerUnit;
        for (int m = 0; m < kBoysTabM; ++m) {
            double term = 1.0 / (2 * m + 1), sum = term;
            for (int i = 1; i < 1000; ++i) {
                term *= 2.0 * T / (2 * m + 2 * i + 1);
                sum += term;
                if (term < 1e-17 * sum) break;
            }
            tab[(size_t)k * kBoysTabM + m] = float(std::exp(-T) * sum);
        }
    }
}

struct qk_ctx {
    int device = 0;
    // system (host copies, QUICK layouts)
    int natom = 0, nshell = 0, nbasis = 0, jbasis = 0;
    std::vector<double> xyz;  // 3 x natom
    std::vector<int> kstart, kprim, katom, ktype, Qstart, Qfinal, Qsbasis, Qfbasis, ksumtype, KLMN;
    std::vector<double> cons, gccoeff, gcexpo;
    bool have_system = false;
    std::vector<double> Ycut, cutprim;
    bool have_schwarz = false;

    // Qshells
    std::vector<int> q_shell, q_L, q_f0, q_atom, q_qf;

    // every Qshell pair (oriented: L(a) >= L(b), equal L: a <= b) and its
    // primitive pairs; built once per geometry
    std::vector<int> ap_qa, ap_qb, ap_cls, ap_prim0, ap_nprim;
    std::vector<int> ap_Pi, ap_Pj;   // global primitive indices, lower shell first (cutprim orientation)
    std::vector<double> ap_Y;        // Schwarz bound of the parent shell pair
    std::vector<double> ap_zeta, ap_oo2z, ap_Px, ap_Py, ap_Pz, ap_K, ap_cut;  // per primitive pair
    std::vector<double> ap_ea, ap_eb;   // exponents of the qa and qb primitives
    DBuf<int> dap_qa, dap_qb, dap_prim0, dap_nprim;
    // the kernels' primitive-major copy of the significant pairs' primitive pairs
    DBuf<int> dp_pstride;
    DBuf<double> dk_zeta, dk_oo2z, dk_Px, dk_Py, dk_Pz, dk_K, dk_cut;
    DBuf<float> dkf_zeta, dkf_oo2z, dkf_Px, dkf_Py, dkf_Pz, dkf_K, dkf_cut;   // FP32 copy
    DBuf<float> d_boys_tab, d_Dpf;
    DBuf<double> dk_ea, dk_eb;   // exponents, primitive-major like dk_*
    DBuf<float> dkf_ea, dkf_eb;   // FP32 copy (mixed-precision gradient)

    // prepared work lists and the options they were built for
    bool prepared = false;
    double prep_tau = -1, prep_pair_cut = -1, prep_fp32_thr = -1;
    int prep_precision = -1;
    int nclass_pairs = 0;
    std::vector<ClassHost> classes;
    // the pair groups of the prepared lists (class, first pair, count) and the
    // pairs' Schwarz bounds, kept for the gradient's work lists
    std::vector<int> grp_cls, grp_start, grp_count;
    std::vector<double> pair_Y;

    // gradient: per-class tables on the device, work lists, accumulators
    DBuf<int> d_gtab[kNumQuartetClasses];
    long long ngtiles[kNumQuartetClasses] = {}, ngtiles32[kNumQuartetClasses] = {};   // FP64 / FP32 work
    DBuf<Tile> d_gtiles[kNumQuartetClasses], d_gtiles32[kNumQuartetClasses];
    // light classes: segments {bra pair, first ket, end ket, list offset} of the
    // per-gradient density compaction (index 0 FP64, 1 FP32) with the first
    // tile of each (d_gslot0)
    long long ngsegs[2][kNumQuartetClasses] = {};
    DBuf<int4> d_gsegs[2][kNumQuartetClasses];
    DBuf<Tile> d_gsteps[2][kNumQuartetClasses];
    DBuf<int> d_gslot0[2][kNumQuartetClasses];
    DBuf<int> d_glist[2][kNumQuartetClasses], d_gcounts[2][kNumQuartetClasses];
    double grad_prep_cutoff = -1, grad_prep_thr = -1;
    int grad_prep_precision = -1;
    bool grad_prepared = false;
    DBuf<unsigned long long> d_gacc, d_gcounters;
    long long gcls_quartets[kNumQuartetClasses] = {};
    long long gcls_prims[kNumQuartetClasses] = {};
    DBuf<double> d_grad;

    // device data
    DBuf<double> d_xyz, d_cons;
    DBuf<int> dq_atom, dq_f0, dq_shell, dq_qf, dq_nc;
    DBuf<int> dp_qa, dp_qb, dp_prim0, dp_nprim;
    DBuf<int> dp_sh, dp_at;   // per pair: shells and atoms of its two Qshells, 16 bits each (< 2^15)
    DBuf<double> dp_Y;
    DBuf<double> dpp_zeta, dpp_oo2z, dpp_Px, dpp_Py, dpp_Pz, dpp_K, dpp_cut, dpp_ea, dpp_eb;
    // work-list construction (qk_lists.hip): the significant pairs' indices into
    // the all-pairs arrays, and the scratch of the list passes
    DBuf<int> d_sel;
    DBuf<ListEntry> dl_E;
    DBuf<KetGroup> dl_KG;
    DBuf<ListClass> dl_LC;
    DBuf<ListDst> dl_dst;
    DBuf<int4> dl_items;
    DBuf<long long> dl_ci, dl_cb, dl_cnt, dl_off, dl_sums, dl_rcnt, dl_roff, dl_rsums;
    DBuf<unsigned char> dl_tmp;
    Arena fock_arena, fock_run_arena, grad_arena;   // the lists themselves
    // per class, whether the FP64 run-mode kernel stages its bra rows of D in
    // LDS (-1: not yet decided for stage_n functions)
    int stage_n = -1;
    signed char stage_ok[kNumQuartetClasses] = {};
    DBuf<int> d_sh_f0, d_sh_nf;
    DBuf<double> d_Dp, d_cm, d_D, d_G;
    DBuf<unsigned long long> d_X, d_counters;
    // one-electron integrals: per basis function its atom, (l, m, n), primitive
    // count, shell Qfinal, exponents and coefficients (10 per function), and
    // the nuclear charges (qk_set_charges)
    DBuf<int> d1_atom, d1_lmn, d1_np, d1_qf;
    // function pairs mu <= nu by total angular momentum l(mu) + l(nu) (0..4)
    DBuf<int2> d1_pairs[5];
    long long n1_pairs[5] = {};
    DBuf<double> d1_a, d1_c, d_chg;
    std::vector<double> chg;
    DBuf<unsigned long long> d1_acc;
    // pair-indexed Coulomb buffer of the prepared pair list (FockArgs::J)
    DBuf<int> dp_jofs;
    DBuf<unsigned long long> d_J;
    long long jsize = 0;

    hipStream_t streams[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ev_start = nullptr, ev_cls[4] = {nullptr, nullptr, nullptr, nullptr};
    qk_stats last{};
    long long cls_quartets[kNumQuartetClasses] = {};
    long long cls_prims[kNumQuartetClasses] = {};
    double cls_flops[kNumQuartetClasses] = {};
    long long cls_quartets32[kNumQuartetClasses] = {};
    long long cls_prims32[kNumQuartetClasses] = {};
    double cls_flops32[kNumQuartetClasses] = {};
    // optional per-class kernel timing (HIP events on the launch stream)
    bool timing = false;
    // (slot k: the FP64 launch of class k; slot kNumQuartetClasses + k: its FP32
    // launch; slot 2 kNumQuartetClasses + k: its gradient launch; slot
    // 3 kNumQuartetClasses + k: its FP32 gradient launch)
    static constexpr int kSlots = 4 * kNumQuartetClasses;
    hipEvent_t ev_t0[kSlots] = {}, ev_t1[kSlots] = {};
    bool cls_timed[kSlots] = {};
    double cls_ms_sum[kSlots] = {};
    int cls_ms_n[kSlots] = {};

    // reference-ABI state
    double abi_tau = 1e-11, abi_prim_limit = 1e-20;
    DBuf<double> d_abi_cm;   // cutMatrix of the last gpu_upload_cutoff_
    bool abi_have_cutmatrix = false;
    bool abi_have_dense = false;
    double abi_grad_cutoff = 1e-11;
    std::vector<double> abi_grad_seed;
    bool abi_have_grad_seed = false;
};

extern "C" long long qk_scratch_grid_cap(long long lane_bytes, int waves_per_block) {
    return scratch_grid_cap(lane_bytes, waves_per_block);
}

extern "C" void qk_default_opts(qk_opts* o) {
    o->tau = 1e-11;
    o->prim_limit = 1e-20;
    o->pair_cut = 0.0;
    o->hybrid = 1.0;
    o->boys_order = QK_BOYS_ORDER_SHELL;
    o->precision = QK_FP64;
    o->fp32_threshold = 1e-5;
}

extern "C" qk_ctx* qk_create(int device) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err("no HIP device available");
        return nullptr;
    }
    if (device < 0) device = 0;
    if (device >= ndev) {
        set_err("device %d out of range (%d devices)", device, ndev);
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        set_err("hipSetDevice(%d) failed", device);
        return nullptr;
    }
    qk_ctx* c = new qk_ctx();
    c->device = device;
    // the no-allocation entry point's scratch address on this device (a failure
    // is kept and reported by that entry point; the rest of the library works)
    (void)noalloc_scratch_init(device);
    for (auto& s : c->streams)
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            set_err("stream creation failed");
            delete c;
            return nullptr;
        }
    (void)hipEventCreateWithFlags(&c->ev_start, hipEventDisableTiming);
    for (auto& e : c->ev_cls) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (c->d_counters.alloc(kNumCounters)) {
        delete c;
        return nullptr;
    }
    return c;
}

extern "C" void qk_destroy(qk_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    for (auto& s : c->streams)
        if (s) (void)hipStreamDestroy(s);
    if (c->ev_start) (void)hipEventDestroy(c->ev_start);
    for (int k = 0; k < qk_ctx::kSlots; ++k) {
        if (c->ev_t0[k]) (void)hipEventDestroy(c->ev_t0[k]);
        if (c->ev_t1[k]) (void)hipEventDestroy(c->ev_t1[k]);
    }
    for (auto& e : c->ev_cls)
        if (e) (void)hipEventDestroy(e);
    delete c;
}


// Every Qshell pair and its primitive-pair data (shell.f90:3-88 restated per
// Qshell pair): zeta = a+b, P = (aA+bB)/(a+b), K = Kpri c_a c_b with
// Kpri = exp(-ab/(a+b) |AB|^2)/(a+b) and the contraction coefficients QUICK's
// Xcoeff takes (gccoeff column ksumtype+L, shell.f90:74-80).
static int build_all_pairs(qk_ctx* c) {
    const int nq = (int)c->q_shell.size();
    c->ap_qa.clear(); c->ap_qb.clear(); c->ap_cls.clear(); c->ap_prim0.clear(); c->ap_nprim.clear();
    c->ap_Pi.clear(); c->ap_Pj.clear();
    std::vector<double> zeta, Px, Py, Pz, K, ea, eb;
    for (int i = 0; i < nq; ++i)
        for (int j = i; j < nq; ++j) {
            int qa = i, qb = j;
            if (c->q_L[qa] < c->q_L[qb]) std::swap(qa, qb);
            const int sa = c->q_shell[qa], sb = c->q_shell[qb];
            const int La = c->q_L[qa], Lb = c->q_L[qb];
            c->ap_qa.push_back(qa);
            c->ap_qb.push_back(qb);
            c->ap_cls.push_back(pair_class_of(La, Lb));
            c->ap_prim0.push_back((int)zeta.size());
            const double* A = &c->xyz[3 * (size_t)(c->katom[sa] - 1)];
            const double* B = &c->xyz[3 * (size_t)(c->katom[sb] - 1)];
            const double dx = A[0] - B[0], dy = A[1] - B[1], dz = A[2] - B[2];
            const double DAB = std::sqrt(dx * dx + dy * dy + dz * dz);
            const int ca = c->ksumtype[sa] - 1 + La, cb = c->ksumtype[sb] - 1 + Lb;
            for (int u = 0; u < c->kprim[sa]; ++u)
                for (int v = 0; v < c->kprim[sb]; ++v) {
                    const double a = c->gcexpo[u + 10 * (size_t)ca], b = c->gcexpo[v + 10 * (size_t)cb];
                    const double kp = std::exp(-a * b / (a + b) * (DAB * DAB)) / (a + b);
                    zeta.push_back(a + b);
                    ea.push_back(a);
                    eb.push_back(b);
                    Px.push_back((A[0] * a + B[0] * b) / (a + b));
                    Py.push_back((A[1] * a + B[1] * b) / (a + b));
                    Pz.push_back((A[2] * a + B[2] * b) / (a + b));
                    K.push_back(kp * c->gccoeff[u + 10 * (size_t)ca] * c->gccoeff[v + 10 * (size_t)cb]);
                    const int Pu = c->kstart[sa] - 1 + u, Pv = c->kstart[sb] - 1 + v;
                    if (sa <= sb) { c->ap_Pi.push_back(Pu); c->ap_Pj.push_back(Pv); }
                    else { c->ap_Pi.push_back(Pv); c->ap_Pj.push_back(Pu); }
                }
            c->ap_nprim.push_back((int)zeta.size() - c->ap_prim0.back());
        }
    c->ap_Y.assign(c->ap_qa.size(), 0.0);
    std::vector<double> oo2z(zeta.size());
    for (size_t k = 0; k < zeta.size(); ++k) oo2z[k] = 0.5 / zeta[k];
    c->ap_zeta = zeta; c->ap_oo2z = oo2z; c->ap_ea = ea; c->ap_eb = eb; c->ap_Px = Px; c->ap_Py = Py; c->ap_Pz = Pz; c->ap_K = K;
    c->ap_cut.assign(zeta.size(), 0.0);
    (void)hipSetDevice(c->device);
    if (c->dpp_zeta.upload(zeta) || c->dpp_oo2z.upload(oo2z) || c->dpp_Px.upload(Px) || c->dpp_Py.upload(Py) || c->dpp_Pz.upload(Pz) ||
        c->dpp_K.upload(K) || c->dpp_ea.upload(ea) || c->dpp_eb.upload(eb) || c->dpp_cut.alloc(zeta.size()) ||
        c->dap_qa.upload(c->ap_qa) ||
        c->dap_qb.upload(c->ap_qb) || c->dap_prim0.upload(c->ap_prim0) || c->dap_nprim.upload(c->ap_nprim))
        return -1;
    return 0;
}

extern "C" int qk_set_system(qk_ctx* c, int natom, const double* xyz, int nshell, int nbasis, int jbasis,
                             const int* kstart, const int* kprim, const int* katom, const int* ktype,
                             const int* Qstart, const int* Qfinal, const int* Qsbasis, const int* Qfbasis,
                             const int* ksumtype, const int* KLMN, const double* cons, const double* gccoeff,
                             const double* gcexpo) {
    if (!c) { set_err("null context"); return -1; }
    if (natom <= 0 || nshell <= 0 || nbasis <= 0 || jbasis <= 0) {
        set_err("invalid sizes natom=%d nshell=%d nbasis=%d jbasis=%d", natom, nshell, nbasis, jbasis);
        return -2;
    }
    if (nbasis > kMaxBasis) {
        set_err("%d basis functions: at most %d are supported", nbasis, kMaxBasis);
        return -2;
    }
    c->natom = natom; c->nshell = nshell; c->nbasis = nbasis; c->jbasis = jbasis;
    c->xyz.assign(xyz, xyz + 3 * (size_t)natom);
    c->kstart.assign(kstart, kstart + nshell);
    c->kprim.assign(kprim, kprim + nshell);
    c->katom.assign(katom, katom + nshell);
    c->ktype.assign(ktype, ktype + nshell);
    c->Qstart.assign(Qstart, Qstart + nshell);
    c->Qfinal.assign(Qfinal, Qfinal + nshell);
    c->Qsbasis.assign(Qsbasis, Qsbasis + 4 * (size_t)nshell);
    c->Qfbasis.assign(Qfbasis, Qfbasis + 4 * (size_t)nshell);
    c->ksumtype.assign(ksumtype, ksumtype + nshell + 1);
    c->KLMN.assign(KLMN, KLMN + 3 * (size_t)nbasis);
    c->cons.assign(cons, cons + nbasis);
    c->gccoeff.assign(gccoeff, gccoeff + 10 * (size_t)nbasis);
    c->gcexpo.assign(gcexpo, gcexpo + 10 * (size_t)nbasis);
    // validate what the kernels assume
    for (int s = 0; s < nshell; ++s) {
        if (Qstart[s] < 0 || Qfinal[s] > 2 || Qstart[s] > Qfinal[s]) {
            set_err("shell %d: angular momentum %d..%d outside s/p/d", s + 1, Qstart[s], Qfinal[s]);
            return -3;
        }
        if (kprim[s] < 1 || kprim[s] > 10 || kstart[s] < 1 || kstart[s] + kprim[s] - 1 > jbasis) {
            set_err("shell %d: primitive range %d+%d outside 1..%d", s + 1, kstart[s], kprim[s], jbasis);
            return -3;
        }
        if (kprim[s] > 8) {   // a wavefront holds one bra's primitive pairs, one per lane
            set_err("shell %d: %d primitives; this build supports at most 8 per shell (64 primitive pairs)",
                    s + 1, kprim[s]);
            return -3;
        }
        if (katom[s] < 1 || katom[s] > natom) {
            set_err("shell %d: atom %d outside 1..%d", s + 1, katom[s], natom);
            return -3;
        }
        if (ksumtype[s] < 1 || ksumtype[s + 1] - 1 > nbasis) {
            set_err("shell %d: function range outside 1..%d", s + 1, nbasis);
            return -3;
        }
    }
    // Qshells
    c->q_shell.clear(); c->q_L.clear(); c->q_f0.clear(); c->q_atom.clear(); c->q_qf.clear();
    static const int nf_of[3] = {1, 3, 6};
    for (int s = 0; s < nshell; ++s)
        for (int L = Qstart[s]; L <= Qfinal[s]; ++L) {
            const int f0 = ksumtype[s] - 1 + Qsbasis[s + nshell * L];
            const int f1 = ksumtype[s] - 1 + Qfbasis[s + nshell * L];
            if (f1 - f0 + 1 != nf_of[L] || f0 < 0 || f1 >= nbasis) {
                set_err("shell %d type %d: function range %d..%d does not hold %d functions", s + 1, L, f0 + 1,
                        f1 + 1, nf_of[L]);
                return -3;
            }
            c->q_shell.push_back(s);
            c->q_L.push_back(L);
            c->q_f0.push_back(f0);
            c->q_atom.push_back(katom[s] - 1);
            c->q_qf.push_back(Qfinal[s]);
        }
    (void)hipSetDevice(c->device);
    std::vector<int> sh_f0(nshell), sh_nf(nshell);
    for (int s = 0; s < nshell; ++s) {
        sh_f0[s] = ksumtype[s] - 1 + Qsbasis[s + nshell * Qstart[s]];
        sh_nf[s] = ksumtype[s] - 1 + Qfbasis[s + nshell * Qfinal[s]] - sh_f0[s] + 1;
    }
    if (c->d_xyz.upload(c->xyz) || c->d_cons.upload(c->cons) || c->dq_atom.upload(c->q_atom) ||
        c->dq_f0.upload(c->q_f0) || c->dq_shell.upload(c->q_shell) || c->dq_qf.upload(c->q_qf) ||
        c->d_sh_f0.upload(sh_f0) || c->d_sh_nf.upload(sh_nf))
        return -1;
    {
        std::vector<int> qnc(c->q_L.size());
        for (size_t k = 0; k < qnc.size(); ++k) qnc[k] = ncart_of(c->q_L[k]);
        if (c->dq_nc.upload(qnc)) return -1;
        // per basis function, for the one-electron integrals: the shell's
        // exponents and the coefficients QUICK's attrashell applies,
        // gccoeff(k, ksumtype + L) cons(f) (nuclear.f90:196-221)
        std::vector<int> fat(nbasis), flmn(nbasis), fnp(nbasis), fqf(nbasis);
        std::vector<double> fa(10 * (size_t)nbasis, 0.0), fc(10 * (size_t)nbasis, 0.0);
        for (size_t q = 0; q < c->q_L.size(); ++q) {
            const int sh = c->q_shell[q], L = c->q_L[q];
            const int col = ksumtype[sh] - 1 + L;
            for (int i = 0; i < ncart_of(L); ++i) {
                const int fn = c->q_f0[q] + i;
                fat[fn] = katom[sh] - 1;
                flmn[fn] = KLMN[3 * fn] | (KLMN[3 * fn + 1] << 4) | (KLMN[3 * fn + 2] << 8);
                fnp[fn] = kprim[sh];
                fqf[fn] = Qfinal[sh];
                for (int u = 0; u < kprim[sh]; ++u) {
                    fa[10 * (size_t)fn + u] = gcexpo[u + 10 * (size_t)col];
                    fc[10 * (size_t)fn + u] = gccoeff[u + 10 * (size_t)col] * cons[fn];
                }
            }
        }
        if (c->d1_atom.upload(fat) || c->d1_lmn.upload(flmn) || c->d1_np.upload(fnp) || c->d1_qf.upload(fqf) ||
            c->d1_a.upload(fa) || c->d1_c.upload(fc))
            return -1;
        std::vector<int> fl(nbasis);
        for (int fn = 0; fn < nbasis; ++fn) fl[fn] = KLMN[3 * fn] + KLMN[3 * fn + 1] + KLMN[3 * fn + 2];
        std::vector<int2> pl[5];
        for (int nu = 0; nu < nbasis; ++nu)
            for (int mu = 0; mu <= nu; ++mu) {
                const int L = fl[mu] + fl[nu];
                if (L > 4) { set_err("one-electron integrals: functions above d"); return -1; }
                pl[L].push_back(make_int2(mu, nu));
            }
        for (int L = 0; L < 5; ++L) {
            if (c->d1_pairs[L].upload(pl[L])) return -1;
            c->n1_pairs[L] = (long long)pl[L].size();
        }
        c->chg.clear();
    }
    size_t nn = (size_t)nbasis * nbasis;
    if (c->d_Dp.alloc(nn) || c->d_X.alloc(nn * kXCopies) || c->d_cm.alloc((size_t)nshell * nshell)) return -1;
    if (build_all_pairs(c)) return -1;
    c->have_system = true;
    c->have_schwarz = false;
    c->prepared = false;
    return 0;
}

extern "C" int qk_set_schwarz(qk_ctx* c, const double* Ycutoff, const double* cutprim) {
    if (!c || !c->have_system) { set_err("qk_set_schwarz: no system uploaded"); return -1; }
    c->Ycut.assign(Ycutoff, Ycutoff + (size_t)c->nshell * c->nshell);
    c->cutprim.assign(cutprim, cutprim + (size_t)c->jbasis * c->jbasis);
    const size_t npp = c->ap_Pi.size();
    std::vector<double> cut(npp);
    for (size_t k = 0; k < npp; ++k) cut[k] = c->cutprim[c->ap_Pi[k] + (size_t)c->jbasis * c->ap_Pj[k]];
    c->ap_cut = cut;
    for (size_t k = 0; k < c->ap_qa.size(); ++k)
        c->ap_Y[k] = c->Ycut[c->q_shell[c->ap_qa[k]] + (size_t)c->nshell * c->q_shell[c->ap_qb[k]]];
    (void)hipSetDevice(c->device);
    if (c->dpp_cut.upload(cut)) return -1;
    c->have_schwarz = true;
    c->prepared = false;
    return 0;
}

// ---------------------------------------------------------------------------
// pair lists and the per-class work queue (once per geometry/options)

using ListSumsFn = std::function<int(const std::vector<long long>&, std::vector<ListDst>&)>;

// The work lists of the entries E (qk_lists.hip): per-item counts and their
// scans on the device, the per-class sums to the host, where dst_f allocates
// the lists and says where they go; then the lists are written.  With runs_f
// (the Fock lists) the per-bra runs follow in the same way.  Synchronous.
static int build_lists(qk_ctx* c, const std::vector<ListEntry>& E, const std::vector<KetGroup>& KG,
                       const std::vector<ListClass>& LC, const std::vector<long long>& cls_item,
                       const std::vector<long long>& cls_bra, double tau, double thr, const ListSumsFn& dst_f,
                       const ListSumsFn* runs_f) {
    const int ncls = kNumQuartetClasses;
    const long long N = cls_item[ncls], NB = cls_bra[ncls];
    std::vector<ListDst> dst(ncls, ListDst{});
    std::vector<long long> sums((size_t)ncls * kListCounts, 0), rsums((size_t)ncls * kRunCounts, 0);
    if (N == 0 || E.empty()) {
        if (dst_f(sums, dst)) return -1;
        return runs_f ? (*runs_f)(rsums, dst) : 0;
    }
    if (N + 1 > INT32_MAX || NB + 1 > INT32_MAX) { set_err("more than 2^31 work-list items"); return -1; }
    if (c->dl_E.upload(E) || c->dl_KG.upload(KG) || c->dl_LC.upload(LC) || c->dl_ci.upload(cls_item) ||
        c->dl_cb.upload(cls_bra) || c->dl_items.alloc((size_t)N) || c->dl_cnt.alloc((size_t)kListCounts * (N + 1)) ||
        c->dl_off.alloc((size_t)kListCounts * (N + 1)) || c->dl_sums.alloc(sums.size()))
        return -1;
    const size_t tb = std::max(list_scan_bytes(N + 1), list_scan_bytes(NB + 1));
    if (c->dl_tmp.alloc(std::max<size_t>(tb, 1))) return -1;
    const int ne = (int)E.size();
    if (lists_count(c->dl_E.p, ne, c->dl_KG.p, c->dl_LC.p, c->dp_Y.p, tau, thr, N, c->dl_items.p, c->dl_cnt.p,
                    c->dl_off.p, c->dl_tmp.p, tb, c->dl_ci.p, ncls, c->dl_sums.p, nullptr)) {
        set_err("work-list count pass failed: %s", hipGetErrorString(hipGetLastError()));
        return -1;
    }
    HIPCHK(hipMemcpy(sums.data(), c->dl_sums.p, sums.size() * sizeof(long long), hipMemcpyDeviceToHost));
    if (dst_f(sums, dst) || c->dl_dst.upload(dst)) return -1;
    lists_write(c->dl_E.p, ne, c->dl_KG.p, c->dl_LC.p, N, c->dl_items.p, c->dl_off.p, c->dl_ci.p, c->dl_dst.p, nullptr);
    HIPCHK(hipGetLastError());
    if (runs_f) {
        if (c->dl_rcnt.alloc((size_t)kRunCounts * (NB + 1)) || c->dl_roff.alloc((size_t)kRunCounts * (NB + 1)) ||
            c->dl_rsums.alloc(rsums.size()))
            return -1;
        if (lists_runs_count(c->dl_E.p, ne, c->dl_LC.p, NB, c->dl_items.p, c->dl_rcnt.p, c->dl_roff.p, c->dl_tmp.p, tb,
                             c->dl_cb.p, ncls, c->dl_rsums.p, nullptr)) {
            set_err("work-list run pass failed: %s", hipGetErrorString(hipGetLastError()));
            return -1;
        }
        HIPCHK(hipMemcpy(rsums.data(), c->dl_rsums.p, rsums.size() * sizeof(long long), hipMemcpyDeviceToHost));
        for (long long v : rsums)
            if (v > INT32_MAX) { set_err("more than 2^31 runs in a class"); return -1; }
        if ((*runs_f)(rsums, dst) || c->dl_dst.upload(dst)) return -1;
        lists_runs_write(c->dl_E.p, ne, c->dl_LC.p, NB, c->dl_items.p, c->dl_off.p, N, c->dl_ci.p, c->dl_roff.p,
                         c->dl_cb.p, c->dl_dst.p, nullptr);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(nullptr));
    return 0;
}

static int prepare(qk_ctx* c, const qk_opts* o) {
    if (!c->have_system || !c->have_schwarz) {
        set_err("prepare: system and Schwarz bounds must be uploaded first");
        return -1;
    }
    if (o->precision != QK_FP64 && o->precision != QK_MIXED) {
        set_err("precision must be QK_FP64 or QK_MIXED (got %d)", o->precision);
        return -4;
    }
    const bool mixed = o->precision == QK_MIXED;
    if (mixed && !(o->fp32_threshold >= 0.0)) {
        set_err("fp32_threshold must be >= 0 (got %g)", o->fp32_threshold);
        return -4;
    }
    if (c->prepared && c->prep_tau == o->tau && c->prep_pair_cut == o->pair_cut &&
        c->prep_precision == o->precision && (!mixed || c->prep_fp32_thr == o->fp32_threshold))
        return 0;
    double Ymax = 0.0;
    for (double y : c->Ycut) Ymax = std::max(Ymax, y);

    // significant pairs: grouped by pair class, then primitive-pair count
    // (descending), then Schwarz bound (descending)
    std::vector<int> sel;
    for (size_t k = 0; k < c->ap_qa.size(); ++k) {
        const double y = c->ap_Y[k];
        if (!(y > o->pair_cut)) continue;
        if (!(y * Ymax > o->tau)) continue;  // no partner can pass Y_ij*Y_kl > tau
        sel.push_back((int)k);
    }
    // the order of a stable sort by (class, primitive-pair count descending, Y
    // descending): a stable LSD radix sort on Y descending (the bits of a
    // positive double order as the value; complemented for descending), then a
    // stable counting sort by (class, count) bucket
    {
        const size_t m = sel.size();
        std::vector<uint64_t> key(m), key2(m);
        std::vector<int> idx2(m);
        for (size_t t = 0; t < m; ++t) {
            uint64_t b;
            memcpy(&b, &c->ap_Y[sel[t]], sizeof b);
            key[t] = ~b;
        }
        std::vector<int> hist(1 << 16);
        for (int sh = 0; sh < 64; sh += 16) {
            std::fill(hist.begin(), hist.end(), 0);
            for (size_t t = 0; t < m; ++t) ++hist[(key[t] >> sh) & 0xffff];
            if (m && hist[(key[0] >> sh) & 0xffff] == (int)m) continue;   // one digit value: order unchanged
            int run = 0;
            for (int& h : hist) { const int v = h; h = run; run += v; }
            for (size_t t = 0; t < m; ++t) {
                const int d = hist[(key[t] >> sh) & 0xffff]++;
                key2[d] = key[t];
                idx2[d] = sel[t];
            }
            key.swap(key2);
            sel.swap(idx2);
        }
        constexpr int kNb = kNumPairClasses * 65;   // (class, 64 - primitive pairs)
        std::vector<int> cnt(kNb + 1, 0);
        for (int e : sel) {
            if (c->ap_nprim[e] > 64) { set_err("more than 64 primitive pairs in a Qshell pair"); return -1; }
            ++cnt[c->ap_cls[e] * 65 + (64 - c->ap_nprim[e]) + 1];
        }
        for (int b = 0; b < kNb; ++b) cnt[b + 1] += cnt[b];
        for (int e : sel) idx2[cnt[c->ap_cls[e] * 65 + (64 - c->ap_nprim[e])]++] = e;
        sel.swap(idx2);
    }
    const int npair = (int)sel.size();
    struct Group { int cls, start, count; };
    std::vector<Group> groups;
    for (int k = 0; k < npair; ++k) {
        const int e = sel[k];
        if (groups.empty() || groups.back().cls != c->ap_cls[e] || c->ap_nprim[sel[groups.back().start]] != c->ap_nprim[e])
            groups.push_back({c->ap_cls[e], k, 0});
        groups.back().count++;
    }
    std::vector<int> qa(npair), qb(npair), prim0(npair), pstride(npair), nprim(npair), psh(npair), pat(npair);
    std::vector<double> Yp(npair);
    if (c->nshell > 0x7fff || c->natom > 0x7fff) { set_err("more than 32767 shells or atoms"); return -1; }
    for (int k = 0; k < npair; ++k) {
        const int e = sel[k];
        qa[k] = c->ap_qa[e]; qb[k] = c->ap_qb[e]; nprim[k] = c->ap_nprim[e];
        Yp[k] = c->ap_Y[e];
        psh[k] = c->q_shell[qa[k]] | (c->q_shell[qb[k]] << 16);
        pat[k] = c->q_atom[qa[k]] | (c->q_atom[qb[k]] << 16);
    }
    // primitive-major layout per group: entry (pair k, primitive j) at
    // base(group) + j * count(group) + (k - start(group)); the copy itself runs
    // on the device (gather_prims_kernel)
    long long nprim_total = 0;
    for (const Group& g : groups) {
        const int np = c->ap_nprim[sel[g.start]];
        for (int k = g.start; k < g.start + g.count; ++k) {
            prim0[k] = (int)(nprim_total + (k - g.start));
            pstride[k] = g.count;
        }
        nprim_total += (long long)np * g.count;
        if (nprim_total > INT32_MAX) { set_err("more than 2^31 primitive pairs"); return -1; }
    }
    // the pair-indexed Coulomb buffer: per group, element-major blocks
    std::vector<int> jofs(npair);
    long long jsize = 0;
    for (const Group& g : groups) {
        const int ne = ncart_of(kPairL[g.cls][0]) * ncart_of(kPairL[g.cls][1]);
        for (int k = g.start; k < g.start + g.count; ++k) jofs[k] = (int)(jsize + (k - g.start));
        jsize += (long long)ne * g.count;
        if (jsize > INT32_MAX) { set_err("pair Coulomb buffer beyond 2^31 entries"); return -1; }
    }
    (void)hipSetDevice(c->device);
    if (c->dp_qa.upload(qa) || c->dp_qb.upload(qb) || c->dp_prim0.upload(prim0) || c->dp_pstride.upload(pstride) ||
        c->dp_nprim.upload(nprim) || c->dp_Y.upload(Yp) || c->dp_sh.upload(psh) || c->dp_at.upload(pat) ||
        c->dp_jofs.upload(jofs) || c->d_sel.upload(sel) || c->d_J.alloc((size_t)std::max(jsize, 1LL)))
        return -1;
    {
        const size_t np = (size_t)std::max(nprim_total, 1LL);
        DBuf<double>* dd[kPrimArrays] = {&c->dk_zeta, &c->dk_oo2z, &c->dk_Px, &c->dk_Py, &c->dk_Pz,
                                         &c->dk_K, &c->dk_cut, &c->dk_ea, &c->dk_eb};
        DBuf<float>* df[kPrimArrays] = {&c->dkf_zeta, &c->dkf_oo2z, &c->dkf_Px, &c->dkf_Py, &c->dkf_Pz,
                                        &c->dkf_K, &c->dkf_cut, &c->dkf_ea, &c->dkf_eb};
        DBuf<double>* src[kPrimArrays] = {&c->dpp_zeta, &c->dpp_oo2z, &c->dpp_Px, &c->dpp_Py, &c->dpp_Pz,
                                          &c->dpp_K, &c->dpp_cut, &c->dpp_ea, &c->dpp_eb};
        PrimArrays ps{}, pd{};
        PrimArraysF pf{};
        for (int a = 0; a < kPrimArrays; ++a) {
            if (dd[a]->alloc(np) || (mixed && df[a]->alloc(np))) return -1;
            ps.a[a] = src[a]->p;
            pd.a[a] = dd[a]->p;
            pf.a[a] = mixed ? df[a]->p : nullptr;
        }
        launch_gather_prims(c->d_sel.p, c->dap_prim0.p, c->dp_nprim.p, c->dp_prim0.p, c->dp_pstride.p, npair, ps, pd,
                            pf, nullptr);
        HIPCHK(hipGetLastError());
    }
    c->jsize = jsize;
    c->grp_cls.clear(); c->grp_start.clear(); c->grp_count.clear();
    for (const Group& g : groups) {
        c->grp_cls.push_back(g.cls);
        c->grp_start.push_back(g.start);
        c->grp_count.push_back(g.count);
    }
    c->pair_Y = Yp;
    c->grad_prepared = false;

    // per-class work lists on the device: for every (bra group, bra, ket
    // group) the prefix of kets passing Y_ij*Y_kl > tau, in tiles of <= 64
    // kets (2 for the cooperative kernel, whose wavefront works through its
    // tile's quartets one after another), compaction segments and runs.
    // Runs (one bra pair per block) hold at most 2048 kets where one strong bra
    // pair's kets would otherwise hold a single block for much of the kernel:
    // (ss|ss), and the FP64 list of a mixed build (the strongest pairs: few
    // bras, long ket ranges).  Measured on (H2O)32: (ss|ss) FP64 3.40 -> 2.49 ms,
    // FP32 1.41 -> 1.22 ms; mixed FP64 lists 0.4 ms less in total (psps 0.61 ->
    // 0.50 ms).  Elsewhere the cap was slower (psss FP64 4.40 -> 5.15 ms): every
    // run zeroes and adds its exchange rows.
    std::vector<ListEntry> E;
    std::vector<KetGroup> KG;
    std::vector<ListClass> LC(kNumQuartetClasses);
    std::vector<long long> cls_item(kNumQuartetClasses + 1), cls_bra(kNumQuartetClasses + 1);
    long long nitem = 0, nbra = 0;
    for (int k = 0; k < kNumQuartetClasses; ++k) {
        cls_item[k] = nitem;
        cls_bra[k] = nbra;
        const int cb = pair_class_of(kClassL[k][0], kClassL[k][1]);
        const int ck = pair_class_of(kClassL[k][2], kClassL[k][3]);
        const bool team = fock_class_is_team(k);
        const bool cap_ss = kClassL[k][0] + kClassL[k][1] == 0;
        LC[k] = ListClass{team ? 2 : 64, cap_ss || mixed ? 2048 : INT_MAX, cap_ss ? 2048 : INT_MAX,
                          mixed && fock_class_has_f32(k) ? 1 : 0, team ? 0 : 1, 0, team ? 0 : 1, 0};
        for (size_t gb = 0; gb < groups.size(); ++gb) {
            if (groups[gb].cls != cb) continue;
            const int kg0 = (int)KG.size();
            for (size_t gk = 0; gk < groups.size(); ++gk) {
                if (groups[gk].cls != ck || (cb == ck && gk < gb)) continue;
                KG.push_back(KetGroup{groups[gk].start, groups[gk].count, gb == gk ? 1 : 0, 0});
            }
            const int m = (int)KG.size() - kg0;
            if (m == 0) continue;
            E.push_back(ListEntry{groups[gb].start, groups[gb].count, m, kg0, nitem, nbra, k, 0});
            nitem += (long long)groups[gb].count * m;
            nbra += groups[gb].count;
        }
    }
    cls_item[kNumQuartetClasses] = nitem;
    cls_bra[kNumQuartetClasses] = nbra;
    if (c->classes.size() != (size_t)kNumQuartetClasses) c->classes.resize(kNumQuartetClasses);
    const ListSumsFn fock_dst = [&](const std::vector<long long>& sm, std::vector<ListDst>& dst) -> int {
        for (int k = 0; k < kNumQuartetClasses; ++k) {
            ClassHost& CH = c->classes[k];
            const long long* v = &sm[(size_t)k * kListCounts];
            if (v[0] > INT32_MAX || v[1] > INT32_MAX || v[4] > INT32_MAX || v[5] > INT32_MAX) {
                set_err("class %d: more than 2^31 tiles or candidate quartets", k);
                return -1;
            }
            CH.ntiles = v[0]; CH.ntiles32 = v[1]; CH.nsegs = v[2]; CH.nsegs32 = v[3];
            CH.total = v[4] + v[5];
            CH.total32 = v[5];
        }
        // every class's lists in one allocation
        Arena& A = c->fock_arena;
        A.reset();
        std::vector<int> slot((size_t)kNumQuartetClasses * 8);
        for (int k = 0; k < kNumQuartetClasses; ++k) {
            const long long* v = &sm[(size_t)k * kListCounts];
            int* sl = &slot[(size_t)k * 8];
            sl[0] = A.reserve<Tile>(v[0]); sl[1] = A.reserve<Tile>(v[1]);
            sl[2] = A.reserve<int4>(v[2]); sl[3] = A.reserve<int4>(v[3]);
            sl[4] = A.reserve<int>(v[4]); sl[5] = A.reserve<int>(v[5]);
            sl[6] = A.reserve<int>(v[2]); sl[7] = A.reserve<int>(v[3]);
        }
        if (A.commit()) return -1;
        for (int k = 0; k < kNumQuartetClasses; ++k) {
            ClassHost& CH = c->classes[k];
            const long long* v = &sm[(size_t)k * kListCounts];
            const int* sl = &slot[(size_t)k * 8];
            A.place(CH.d_tiles, sl[0], v[0]); A.place(CH.d_tiles32, sl[1], v[1]);
            A.place(CH.d_segs, sl[2], v[2]); A.place(CH.d_segs32, sl[3], v[3]);
            A.place(CH.d_list, sl[4], v[4]); A.place(CH.d_list32, sl[5], v[5]);
            A.place(CH.d_counts, sl[6], v[2]); A.place(CH.d_counts32, sl[7], v[3]);
            dst[k] = ListDst{CH.d_tiles.p, CH.d_tiles32.p, CH.d_segs.p, CH.d_segs32.p, nullptr, nullptr,
                             nullptr, nullptr, nullptr, nullptr};
        }
        return 0;
    };
    const ListSumsFn fock_runs = [&](const std::vector<long long>& rs, std::vector<ListDst>& dst) -> int {
        for (int k = 0; k < kNumQuartetClasses; ++k) {
            ClassHost& CH = c->classes[k];
            const long long* v = &rs[(size_t)k * kRunCounts];
            CH.nruns = v[0]; CH.nruns32 = v[1]; CH.nsegruns = v[2]; CH.nsegruns32 = v[3];
        }
        Arena& A = c->fock_run_arena;
        A.reset();
        for (int k = 0; k < kNumQuartetClasses; ++k)
            for (int a = 0; a < kRunCounts; ++a) A.reserve<int2>(rs[(size_t)k * kRunCounts + a]);
        if (A.commit()) return -1;
        for (int k = 0; k < kNumQuartetClasses; ++k) {
            ClassHost& CH = c->classes[k];
            const long long* v = &rs[(size_t)k * kRunCounts];
            const int s0 = k * kRunCounts;
            A.place(CH.d_runs, s0, v[0]); A.place(CH.d_runs32, s0 + 1, v[1]);
            A.place(CH.d_segruns, s0 + 2, v[2]); A.place(CH.d_segruns32, s0 + 3, v[3]);
            dst[k].r64 = CH.d_runs.p; dst[k].r32 = CH.d_runs32.p;
            dst[k].sr64 = CH.d_segruns.p; dst[k].sr32 = CH.d_segruns32.p;
        }
        return 0;
    };
    if (build_lists(c, E, KG, LC, cls_item, cls_bra, o->tau, mixed ? o->fp32_threshold : 0.0, fock_dst, &fock_runs))
        return -1;
    for (int k = 0; k < kNumQuartetClasses; ++k) {
        // compaction costs about as much per candidate as an FP32 (ss|ss), (ps|ss)
        // or (ps|ps) quartet saves, and more than an FP64 (ss|ss) one (measured on
        // (H2O)32); it pays for every other class
        ClassHost& CH = c->classes[k];
        const int LT = kClassL[k][0] + kClassL[k][1] + kClassL[k][2] + kClassL[k][3];
        const bool small = kClassL[k][1] == 0 && kClassL[k][3] == 0 && LT <= 2 && kClassL[k][0] <= 1;
        CH.compact = CH.nsegs > 0 && LT >= 1;
        CH.compact32 = CH.nsegs32 > 0 && !small;
    }
    if (mixed && !c->d_boys_tab.p) {
        std::vector<float> tab;
        boys_table(tab);
        if (c->d_boys_tab.upload(tab)) return -1;
    }
    if (mixed && c->d_Dpf.alloc((size_t)c->nbasis * c->nbasis)) return -1;
    c->nclass_pairs = npair;
    c->prepared = true;
    c->prep_tau = o->tau;
    c->prep_pair_cut = o->pair_cut;
    c->prep_precision = o->precision;
    c->prep_fp32_thr = o->fp32_threshold;
    return 0;
}

extern "C" int qk_prepare(qk_ctx* c, const qk_opts* o) {
    if (!c) { set_err("null context"); return -1; }
    qk_opts d;
    if (!o) { qk_default_opts(&d); o = &d; }
    (void)hipSetDevice(c->device);
    return prepare(c, o);
}

// ---------------------------------------------------------------------------
// Fock build

// the per-build scratch of a Fock build: the context's buffers, or the fixed
// scratch array of the no-allocation entry point
struct FockBufs {
    unsigned long long* X;          // xcopies packed triangles
    int xcopies;
    unsigned long long* counters;   // kNumCounters
    unsigned long long* J;          // pair-indexed Coulomb buffer (jsize)
    double* Dp;                     // n x n
    float* Dpf;                     // n x n (mixed precision)
    double* cm;                     // nshell x nshell
    bool compact;                   // density compaction into the prepared ket lists
};

static FockArgs make_args(qk_ctx* c, const qk_opts* o, const FockBufs& fb) {
    FockArgs a;
    a.q = {c->dq_atom.p, c->dq_f0.p, c->dq_shell.p, c->dq_qf.p, c->dq_nc.p};
    a.pr = {c->dp_qa.p, c->dp_qb.p, c->dp_prim0.p, c->dp_pstride.p, c->dp_nprim.p, c->dp_Y.p};
    a.pp = {c->dk_zeta.p, c->dk_oo2z.p, c->dk_Px.p, c->dk_Py.p, c->dk_Pz.p, c->dk_K.p, c->dk_cut.p};
    a.ppf = {c->dkf_zeta.p, c->dkf_oo2z.p, c->dkf_Px.p, c->dkf_Py.p, c->dkf_Pz.p, c->dkf_K.p, c->dkf_cut.p};
    a.Dpf = fb.Dpf;
    a.boys_tab = c->d_boys_tab.p;
    a.xyz = c->d_xyz.p;
    a.Dp = fb.Dp;
    a.cutmat = fb.cm;
    a.X = fb.X;
    a.counters = fb.counters;
    a.J = fb.J;
    a.jofs = c->dp_jofs.p;
    a.xcopies = fb.xcopies;
    a.n = c->nbasis;
    a.nshell = c->nshell;
    a.tau = o->tau;
    a.prim_limit = o->prim_limit;
    a.hybrid = o->hybrid;
    a.boys_qshell_order = o->boys_order == QK_BOYS_ORDER_QSHELL ? 1 : 0;
    a.stage_d = 0;
    return a;
}

// inputs of the density compaction: the Fock test (grad 0) or the gradient's
static CompactArgs compact_args(const qk_ctx* c, const FockArgs& a, int grad, double tau) {
    CompactArgs ca;
    ca.psh = c->dp_sh.p;
    ca.pat = c->dp_at.p;
    ca.Y = a.pr.Y;
    ca.cutmat = a.cutmat;
    ca.nshell = a.nshell;
    ca.tau = tau;
    ca.grad = grad;
    ca.steps = nullptr;
    ca.slot0 = nullptr;
    return ca;
}

// folds the previous launches' kernel times into the running sums (waits for them)
static int fold_timing(qk_ctx* c) {
    for (int k = 0; k < qk_ctx::kSlots; ++k) {
        if (c->cls_timed[k]) {
            float ms = 0.0f;
            HIPCHK(hipEventSynchronize(c->ev_t1[k]));
            HIPCHK(hipEventElapsedTime(&ms, c->ev_t0[k], c->ev_t1[k]));
            c->cls_ms_sum[k] += ms;
            c->cls_ms_n[k] += 1;
        }
        c->cls_timed[k] = false;
    }
    return 0;
}

// enqueue the whole build on `s` (classes fan out to the context's streams and join back)
// (fb: the scratch to use; NULL = the context's own, preparing the work lists
// for these options first)
static int fock_enqueue(qk_ctx* c, const double* dD, const double* dcutmat_or_null, double* dG, const qk_opts* o,
                        hipStream_t s, const FockBufs* fb = nullptr) {
    FockBufs own;
    if (!fb) {
        if (prepare(c, o)) return -1;
        own = {c->d_X.p, kXCopies, c->d_counters.p, c->d_J.p, c->d_Dp.p, c->d_Dpf.p, c->d_cm.p, true};
        fb = &own;
    }
    const int n = c->nbasis, ns = c->nshell;
    const bool mixed = o->precision == QK_MIXED;
    launch_scale_density(dD, c->d_cons.p, fb->Dp, mixed ? fb->Dpf : nullptr, n, s);
    if (dcutmat_or_null)
        HIPCHK(hipMemcpyAsync(fb->cm, dcutmat_or_null, sizeof(double) * ns * ns, hipMemcpyDeviceToDevice, s));
    else
        launch_density_cutoff(dD, c->d_sh_f0.p, c->d_sh_nf.p, fb->cm, n, ns, s);
    launch_fill_u64(fb->X, (long long)n * n * fb->xcopies, s);
    launch_fill_u64(fb->J, c->jsize, s);
    launch_fill_u64(fb->counters, kNumCounters, s);
    HIPCHK(hipGetLastError());
    FockArgs a = make_args(c, o, *fb);
    const CompactArgs ca = compact_args(c, a, 0, a.tau);
    if (fold_timing(c)) return -1;
    // heaviest classes first, round-robin over the streams
    std::vector<int> order(kNumQuartetClasses);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return c->classes[x].total > c->classes[y].total; });
    if (mixed && !c->d_boys_tab.p) { set_err("mixed precision: Boys table missing"); return -1; }
    // Two streams, forked from the caller's and joined with events.  In a mixed
    // build the FP64 lists hold the few strongest pairs (long ket ranges on few
    // bras): those kernels run on the second stream beside the FP32 lists.  In
    // an FP64 build the classes alternate between the streams in this order, so
    // that one class's tail overlaps the next class (measured: (H2O)32 36.0 ->
    // 34.2 ms, n-icosane 20.3 -> 17.3 ms).
    const hipStream_t s64 = c->streams[0];
    HIPCHK(hipEventRecord(c->ev_start, s));
    HIPCHK(hipStreamWaitEvent(s64, c->ev_start, 0));
    // The classes of one stream run one after another, so at most two kernels
    // are in flight.  Kernels that spill get a grid small enough
    // that their private-segment backing stays under kScratchBudget
    // (scratch_grid_cap); the grid-stride loop over tiles covers the rest.
    // LDS for one bra's exchange rows.  gfx950 gives a block up to 160 KB of
    // dynamic LDS (152 KB leave room for the bra slab and the kernels' static
    // LDS, <= 6 KB), but rows beyond 64 KB hold a CU to one or two blocks and
    // every run zeroes and flushes them: measured on (H2O)64 (1,600 functions)
    // against direct atomics into X, run mode with large rows pays for the FP64
    // classes up to a d-p bra (ppps 16.5 -> 10.7 ms, dpps 17.4 -> 9.5 ms, FP64
    // build 235 -> 202 ms) and for the FP32 classes with a p or p-p ket (ppps
    // 10.3 -> 7.6 ms), and costs the FP32 classes with an s-s ket (dsss 6.6 ->
    // 9.1 ms) and every d-d bra (on (H2O)32, ddps 1.04 -> 1.77 ms)
    constexpr size_t kRowsLdsMax = 152 * 1024, kRowsLdsSmall = 64 * 1024;
    constexpr size_t kStageLds = 64 * 1024;
    // With light FP64 lists (under a tenth of the tiles) the second stream
    // ends first, so the FP32 lists of (dd|dp) and (dd|dd) follow the FP64
    // lists there.  Measured, mixed: (H2O)32 at 1e-3 (FP64 lists 7.5% of the
    // tiles) 23.48 -> 22.98 ms, (H2O)64 at exchange weight 0.2 and 1e-3 (5.5%)
    // 145.4 -> 144.0 ms; at 1e-4 (13.1%) the same move cost 23.86 -> 24.33 ms.
    long long tiles64 = 0, tiles32 = 0;
    for (int k = 0; k < kNumQuartetClasses; ++k) {
        tiles64 += c->classes[k].ntiles;
        tiles32 += c->classes[k].ntiles32;
    }
    const bool light64 = mixed && tiles64 * 10 < tiles64 + tiles32;
    for (size_t pos = 0; pos < order.size(); ++pos) {
        const int k = order[pos];
        ClassHost& CH = c->classes[k];
        if (CH.ntiles == 0 && CH.ntiles32 == 0) continue;
        // the stream of this class's FP64 list
        const hipStream_t l64 = !mixed && (pos & 1) ? s : s64;
        // one bra pair per block with its exchange rows in LDS, when they fit
        const int nrow = ncart_of(kClassL[k][0]) + ncart_of(kClassL[k][1]);
        const size_t rows_lds = sizeof(unsigned long long) * (size_t)nrow * n;
        const bool big_ok = nrow <= 9 && rows_lds <= kRowsLdsMax;
        const int nket = ncart_of(kClassL[k][2]) * ncart_of(kClassL[k][3]);
        const bool use_runs = !fock_class_is_team(k) && (rows_lds <= kRowsLdsSmall || big_ok);
        const bool use_runs32 = !fock_class_is_team(k) && (rows_lds <= kRowsLdsSmall || (big_ok && nket >= 3));
        // dynamic LDS of the thread-per-quartet kernels: the exchange rows and
        // one bra slab (run mode), or four bra slabs (7 x 64 values each)
        const size_t slab64 = 7 * 64 * sizeof(double), slab32 = 7 * 64 * sizeof(float);
        // FP64 run mode: the bra rows of D in LDS too, where they fit kStageLds
        // and leave the kernel as many resident blocks per CU as without them
        // (measured on (H2O)32: FP64 build 38.4 -> 37.5 ms, psps 3.14 -> 2.73 ms;
        // on (H2O)64 (ss|ss) staged fell from five blocks per CU to two, 16.6 ->
        // 21.6 ms; in the FP32 kernels the second code path cost more than it saved)
        const size_t drows64 = sizeof(double) * (size_t)nrow * n;
        if (c->stage_n != n) {
            for (int j = 0; j < kNumQuartetClasses; ++j) c->stage_ok[j] = -1;
            c->stage_n = n;
        }
        if (use_runs && c->stage_ok[k] < 0) {
            const size_t plain = rows_lds + slab64, staged = plain + drows64;
            c->stage_ok[k] = staged <= kStageLds && fock_class_blocks(k, staged, true) >= fock_class_blocks(k, plain, false) ? 1 : 0;
        }
        const bool stage64 = use_runs && c->stage_ok[k] == 1;
        const size_t lds64 = use_runs ? rows_lds + slab64 + (stage64 ? drows64 : 0) : 4 * slab64;
        const size_t lds32 = use_runs32 ? rows_lds + slab32 : 4 * slab32;
        if (CH.ntiles) {
            if (c->timing) HIPCHK(hipEventRecord(c->ev_t0[k], l64));
            a.counters = fb->counters + 2 * k;
            a.stage_d = stage64 ? 1 : 0;
            if (fock_class_is_team(k)) {
                ClassWork w{CH.d_tiles.p, CH.ntiles, nullptr, 0};
                // one 64-lane block per tile; up to 16 resident per CU
                const int grid = (int)std::min<long long>(CH.ntiles, 256LL * 16);
                launch_team_class(k, a, w, grid, l64);
            } else {
                const long long cap = std::min(256LL * 16, scratch_grid_cap((long long)fock_class_scratch(k), 4));
                if (use_runs && CH.compact && fb->compact) {
                    // density screen first: every lane of the class kernel then has a quartet
                    launch_compact(ca, CH.d_segs.p, CH.nsegs, CH.d_list.p, CH.d_counts.p, l64);
                    ClassWork w{CH.d_tiles.p, CH.ntiles, CH.d_segruns.p, CH.nsegruns, CH.d_segs.p, CH.d_list.p,
                                CH.d_counts.p};
                    launch_fock_class(k, a, w, (int)std::min<long long>(CH.nsegruns, cap), lds64, l64);
                } else if (use_runs) {
                    ClassWork w{CH.d_tiles.p, CH.ntiles, CH.d_runs.p, CH.nruns};
                    launch_fock_class(k, a, w, (int)std::min<long long>(CH.nruns, cap), lds64, l64);
                } else {
                    ClassWork w{CH.d_tiles.p, CH.ntiles, nullptr, 0};
                    const long long want = (CH.ntiles + 3) / 4;   // 4 wavefronts per block
                    launch_fock_class(k, a, w, (int)std::min<long long>(want, cap), lds64, l64);
                }
            }
            HIPCHK(hipGetLastError());
            if (c->timing) HIPCHK(hipEventRecord(c->ev_t1[k], l64));
            c->cls_timed[k] = c->timing;
        }
        if (mixed && CH.ntiles32) {
            const hipStream_t ls = light64 && (k == 19 || k == 20) ? s64 : s;
            a.counters = fb->counters + 2 * kNumQuartetClasses + 2 * k;
            a.stage_d = 0;
            const long long cap = std::min(256LL * 16, scratch_grid_cap((long long)fock_class_scratch_f32(k), 4));
            const int slot = kNumQuartetClasses + k;
            if (c->timing) HIPCHK(hipEventRecord(c->ev_t0[slot], ls));
            ClassWork w{CH.d_tiles32.p, CH.ntiles32, nullptr, 0};
            int grid = (int)std::min<long long>((CH.ntiles32 + 3) / 4, cap);
            size_t lds = lds32;
            // one 64-lane block per tile; up to 20 resident per CU (five wavefronts per SIMD)
            if (fock_class_is_team(k)) grid = (int)std::min<long long>(CH.ntiles32, 256LL * 20);
            if (use_runs32 && CH.compact32 && fb->compact) {
                launch_compact(ca, CH.d_segs32.p, CH.nsegs32, CH.d_list32.p, CH.d_counts32.p, ls);
                w.runs = CH.d_segruns32.p;
                w.nruns = CH.nsegruns32;
                w.segs = CH.d_segs32.p;
                w.list = CH.d_list32.p;
                w.counts = CH.d_counts32.p;
                grid = (int)std::min<long long>(w.nruns, cap);
                lds = lds32;
            } else if (use_runs32) {
                w.runs = CH.d_runs32.p;
                w.nruns = CH.nruns32;
                grid = (int)std::min<long long>(w.nruns, cap);
                lds = lds32;
            }
            if (fock_class_is_team(k)) launch_team_f32_class(k, a, w, grid, ls);
            else launch_fock_f32_class(k, a, w, grid, lds, ls);
            HIPCHK(hipGetLastError());
            if (c->timing) HIPCHK(hipEventRecord(c->ev_t1[slot], ls));
            c->cls_timed[slot] = c->timing;
        }
    }
    HIPCHK(hipEventRecord(c->ev_cls[0], s64));
    HIPCHK(hipStreamWaitEvent(s, c->ev_cls[0], 0));
    launch_jpair_scatter(fb->J, c->dp_jofs.p, c->dp_pstride.p, c->dp_qa.p, c->dp_qb.p, a.q, c->nclass_pairs, fb->X,
                         n, s);
    launch_fock_finalize(fb->X, fb->xcopies, c->d_cons.p, dG, n, fb->counters + kRangeFlag, s);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int qk_fock_device(qk_ctx* c, const double* dD, double* dG, const qk_opts* o, void* stream) {
    if (!c) { set_err("null context"); return -1; }
    qk_opts d;
    if (!o) { qk_default_opts(&d); o = &d; }
    (void)hipSetDevice(c->device);
    return fock_enqueue(c, dD, nullptr, dG, o, (hipStream_t)stream);
}

// true when the work lists on the device were built for these options
static bool prepared_for(const qk_ctx* c, const qk_opts* o) {
    const bool mixed = o->precision == QK_MIXED;
    return c->prepared && c->prep_tau == o->tau && c->prep_pair_cut == o->pair_cut &&
           c->prep_precision == o->precision && (!mixed || c->prep_fp32_thr == o->fp32_threshold);
}

// The Fock build with its per-build scratch in the fixed device array: one
// accumulator copy, the plain tile runs instead of the density compaction (the
// same quartets contribute the same fixed-point values, so the operator is
// bitwise that of qk_fock_device), no hipMalloc / hipFree / allocating library
// call.  The work lists must already be on the device (qk_prepare, or an
// earlier build with the same options).
extern "C" int qk_fock_device_noalloc(qk_ctx* c, const double* dD, double* dG, const qk_opts* o, void* stream) {
    if (!c) { set_err("null context"); return -1; }
    qk_opts d;
    if (!o) { qk_default_opts(&d); o = &d; }
    (void)hipSetDevice(c->device);
    if (o->precision != QK_FP64 && o->precision != QK_MIXED) {
        set_err("precision must be QK_FP64 or QK_MIXED (got %d)", o->precision);
        return -4;
    }
    if (!prepared_for(c, o)) {
        set_err("qk_fock_device_noalloc: no work lists for these options on the device (call qk_prepare first)");
        return -6;
    }
    unsigned char* base = (unsigned char*)noalloc_scratch_ptr(c->device);
    if (!base) {
        set_err("qk_fock_device_noalloc: the scratch array's address on device %d is unknown (hip error %d)",
                c->device, noalloc_scratch_init(c->device));
        return -1;
    }
    const bool mixed = o->precision == QK_MIXED;
    const long long n = c->nbasis, ns = c->nshell;
    auto up = [](long long b) { return (b + 255) / 256 * 256; };
    const long long oX = 0, oJ = oX + up(8 * n * n), oC = oJ + up(8 * std::max(c->jsize, 1LL)),
                    oD = oC + up(8LL * kNumCounters), oDf = oD + up(8 * n * n),
                    oCm = oDf + (mixed ? up(4 * n * n) : 0), end = oCm + up(8 * ns * ns);
    if (end > kNoAllocScratchBytes) {
        set_err("qk_fock_device_noalloc: %lld functions and %lld shells need %lld bytes of scratch, more than the "
                "%lld of the fixed array",
                n, ns, end, (long long)kNoAllocScratchBytes);
        return -7;
    }
    const FockBufs fb{(unsigned long long*)(base + oX), 1, (unsigned long long*)(base + oC),
                      (unsigned long long*)(base + oJ), (double*)(base + oD), mixed ? (float*)(base + oDf) : nullptr,
                      (double*)(base + oCm), false};
    return fock_enqueue(c, dD, nullptr, dG, o, (hipStream_t)stream, &fb);
}

// Algorithmic operation count of one class (FMA = 2): per primitive quartet the
// set-up of P-Q, W and T (35), QUICK's Boys fit plus recursion (24 + 4 LT, exp,
// sqrt and divisions counted as one each), the prefactor scaling (LT + 1) and the
// generated VRR; per contracted quartet the screening (12), the generated HRR and
// the six J/K digestion contractions (12 per integral).
static double class_flops(int k, long long nq, long long np, bool fp32) {
    const int* t = kEriFlopTable[k];
    const int LT = t[0] + t[1] + t[2] + t[3];
    // FP32 Boys: table Taylor (5 terms), exp, and the downward recursion
    const double boys = fp32 ? 12.0 + 3.0 * LT : 24.0 + 4.0 * LT;
    const double per_prim = 35.0 + boys + (LT + 1) + t[5];
    const double per_quartet = 12.0 + t[6] + 12.0 * t[7];
    return per_prim * double(np) + per_quartet * double(nq);
}

static int read_stats(qk_ctx* c) {
    unsigned long long h[kNumCounters];
    HIPCHK(hipMemcpy(h, c->d_counters.p, sizeof(h), hipMemcpyDeviceToHost));
    c->last = qk_stats{};
    const int o32 = 2 * kNumQuartetClasses;
    for (int k = 0; k < kNumQuartetClasses; ++k) {
        const long long q64 = (long long)h[2 * k], p64 = (long long)h[2 * k + 1];
        const long long q32 = (long long)h[o32 + 2 * k], p32 = (long long)h[o32 + 2 * k + 1];
        const double f64 = class_flops(k, q64, p64, false), f32 = class_flops(k, q32, p32, true);
        c->cls_quartets[k] = q64;
        c->cls_prims[k] = p64;
        c->cls_flops[k] = f64;
        c->cls_quartets32[k] = q32;
        c->cls_prims32[k] = p32;
        c->cls_flops32[k] = f32;
        c->last.quartets += q64 + q32;
        c->last.quartets_fp32 += q32;
        c->last.prim_quartets += p64 + p32;
        c->last.flops_fp64 += f64;
        c->last.flops_fp32 += f32;
    }
    return 0;
}

extern "C" int qk_set_timing(qk_ctx* c, int on) {
    if (!c) { set_err("null context"); return -1; }
    (void)hipSetDevice(c->device);
    if (on && !c->ev_t0[0])
        for (int k = 0; k < qk_ctx::kSlots; ++k) {
            HIPCHK(hipEventCreate(&c->ev_t0[k]));
            HIPCHK(hipEventCreate(&c->ev_t1[k]));
        }
    c->timing = on != 0;
    for (int k = 0; k < qk_ctx::kSlots; ++k) {
        c->cls_timed[k] = false;
        c->cls_ms_sum[k] = 0.0;
        c->cls_ms_n[k] = 0;
    }
    return 0;
}

// average kernel time of a timing slot over every timed build since qk_set_timing(c, 1)
static int slot_ms(qk_ctx* c, int slot, float* ms) {
    double sum = c->cls_ms_sum[slot];
    int cnt = c->cls_ms_n[slot];
    if (c->cls_timed[slot]) {
        float ms1 = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms1, c->ev_t0[slot], c->ev_t1[slot]));
        sum += ms1;
        cnt += 1;
    }
    *ms = cnt ? float(sum / cnt) : 0.0f;
    return 0;
}

// per-class figures of the last build: quartets, primitive quartets,
// algorithmic flops (FP64 and FP32 kernels together), and (when timing was on)
// the FP64 kernel's elapsed ms
extern "C" int qk_class_stats(qk_ctx* c, int k, long long* nq, long long* np, double* flops, float* ms) {
    if (!c || k < 0 || k >= kNumQuartetClasses) { set_err("bad arguments"); return -1; }
    (void)hipSetDevice(c->device);
    HIPCHK(hipDeviceSynchronize());
    if (read_stats(c)) return -1;
    *nq = c->cls_quartets[k];
    *np = c->cls_prims[k];
    *flops = c->cls_flops[k];
    return slot_ms(c, k, ms);
}

// the FP32 share of qk_class_stats' figures (mixed precision) and the FP32 kernel's ms
extern "C" int qk_class_stats_fp32(qk_ctx* c, int k, long long* nq, long long* np, double* flops, float* ms) {
    if (!c || k < 0 || k >= kNumQuartetClasses) { set_err("bad arguments"); return -1; }
    (void)hipSetDevice(c->device);
    HIPCHK(hipDeviceSynchronize());
    if (read_stats(c)) return -1;
    *nq = c->cls_quartets32[k];
    *np = c->cls_prims32[k];
    *flops = c->cls_flops32[k];
    return slot_ms(c, kNumQuartetClasses + k, ms);
}

extern "C" int qk_boys_table_f32(float* out, int rows, int cols) {
    if (rows != kBoysTabRows || cols != kBoysTabM) {
        set_err("Boys table is %d x %d", kBoysTabRows, kBoysTabM);
        return -1;
    }
    std::vector<float> tab;
    boys_table(tab);
    std::copy(tab.begin(), tab.end(), out);
    return 0;
}

extern "C" void qk_boys_table_shape(int* rows, int* cols, int* per_unit) {
    *rows = kBoysTabRows;
    *cols = kBoysTabM;
    *per_unit = kBoysTabPerUnit;
}

extern "C" int qk_last_stats(qk_ctx* c, qk_stats* s) {
    if (!c || !s) { set_err("null argument"); return -1; }
    (void)hipSetDevice(c->device);
    HIPCHK(hipDeviceSynchronize());
    if (read_stats(c)) return -1;
    *s = c->last;
    return 0;
}

// Reads a range flag the finalize kernels raise (waits for the work before it).
static int check_range(const unsigned long long* dflag, const char* what) {
    unsigned long long f = 0;
    if (!dflag) return 0;
    HIPCHK(hipMemcpy(&f, dflag, sizeof(f), hipMemcpyDeviceToHost));
    if (f) {
        set_err("%s: the int64 fixed-point accumulator came within a factor of two of its range "
                "(|G| >= 8192 or |dE/dR| >= 2^18); the result is not trustworthy", what);
        return -5;
    }
    return 0;
}

extern "C" int qk_check_range(qk_ctx* c) {
    if (!c) { set_err("null context"); return -1; }
    (void)hipSetDevice(c->device);
    if (check_range(c->d_counters.p ? c->d_counters.p + kRangeFlag : nullptr, "last Fock build")) return -5;
    if (check_range(c->d_gcounters.p ? c->d_gcounters.p + kGradRangeFlag : nullptr, "last gradient")) return -5;
    return 0;
}

extern "C" int qk_fock_host(qk_ctx* c, const double* D, double* F, const qk_opts* o) {
    if (!c) { set_err("null context"); return -1; }
    qk_opts d;
    if (!o) { qk_default_opts(&d); o = &d; }
    (void)hipSetDevice(c->device);
    const size_t nn = (size_t)c->nbasis * c->nbasis;
    if (c->d_D.alloc(nn) || c->d_G.alloc(nn)) return -1;
    HIPCHK(hipMemcpy(c->d_D.p, D, nn * sizeof(double), hipMemcpyHostToDevice));
    if (qk_fock_device(c, c->d_D.p, c->d_G.p, o, nullptr)) return -1;
    std::vector<double> G(nn);
    HIPCHK(hipMemcpy(G.data(), c->d_G.p, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (check_range(c->d_counters.p + kRangeFlag, "qk_fock_host")) return -5;
    for (size_t k = 0; k < nn; ++k) F[k] += G[k];
    return 0;
}

extern "C" int qk_num_classes(void) { return kNumQuartetClasses; }

extern "C" int qk_class_info(qk_ctx* c, int k, int* l, long long* work) {
    if (k < 0 || k >= kNumQuartetClasses) { set_err("class index out of range"); return -1; }
    for (int i = 0; i < 4; ++i) l[i] = kClassL[k][i];
    *work = (c && c->prepared) ? c->classes[k].total : 0;
    return 0;
}

// lanes occupied by the tiles of class k: quartets / (64 * tiles)
extern "C" double qk_class_lane_fill(qk_ctx* c, int k) {
    if (!c || !c->prepared || k < 0 || k >= kNumQuartetClasses || c->classes[k].ntiles == 0) return 0.0;
    return double(c->classes[k].total) / (64.0 * double(c->classes[k].ntiles));
}

extern "C" int qk_compute_schwarz(qk_ctx* c, double* Ycut_out, double* cutprim_out) {
    if (!c || !c->have_system) { set_err("qk_compute_schwarz: no system uploaded"); return -1; }
    (void)hipSetDevice(c->device);
    const size_t npairs = c->ap_qa.size(), npp = c->ap_Pi.size();
    DBuf<double> d_ymax, d_pmax;
    DBuf<int> d_list;
    if (d_ymax.alloc(npairs) || d_pmax.alloc(npp)) return -1;
    std::vector<int> list;
    std::vector<int> start(kNumPairClasses + 1, 0);
    for (int cls = 0; cls < kNumPairClasses; ++cls) {
        start[cls] = (int)list.size();
        for (size_t k = 0; k < npairs; ++k)
            if (c->ap_cls[k] == cls) list.push_back((int)k);
    }
    start[kNumPairClasses] = (int)list.size();
    if (d_list.upload(list)) return -1;
    DevQshells qs{c->dq_atom.p, c->dq_f0.p, c->dq_shell.p, c->dq_qf.p, c->dq_nc.p};
    DevPrims pp{c->dpp_zeta.p, c->dpp_oo2z.p, c->dpp_Px.p, c->dpp_Py.p, c->dpp_Pz.p, c->dpp_K.p, c->dpp_cut.p};
    for (int cls = 0; cls < kNumPairClasses; ++cls) {
        launch_schwarz(cls, d_list.p + start[cls], start[cls + 1] - start[cls], c->dap_qa.p, c->dap_qb.p,
                       c->dap_prim0.p, c->dap_nprim.p, qs, pp, c->d_xyz.p, c->d_cons.p, d_ymax.p, d_pmax.p, nullptr);
        HIPCHK(hipGetLastError());
    }
    std::vector<double> ymax(npairs), pmax(npp);
    HIPCHK(hipMemcpy(ymax.data(), d_ymax.p, npairs * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(pmax.data(), d_pmax.p, npp * sizeof(double), hipMemcpyDeviceToHost));
    const int ns = c->nshell, jb = c->jbasis;
    std::vector<double> Ymx((size_t)ns * ns, 0.0), Pmx((size_t)jb * jb, 0.0);
    for (size_t k = 0; k < npairs; ++k) {
        const int sa = c->q_shell[c->ap_qa[k]], sb = c->q_shell[c->ap_qb[k]];
        const size_t i1 = sa + (size_t)ns * sb, i2 = sb + (size_t)ns * sa;
        Ymx[i1] = std::max(Ymx[i1], ymax[k]);
        Ymx[i2] = std::max(Ymx[i2], ymax[k]);
    }
    for (size_t k = 0; k < npp; ++k) {
        const size_t i = c->ap_Pi[k] + (size_t)jb * c->ap_Pj[k];
        Pmx[i] = std::max(Pmx[i], pmax[k]);
    }
    // Ycutoff = sqrt(max |(ij|ij)|), cutprim = sqrt(max |[ab|ab]|) (schwarz.f90:26-27,185)
    for (auto& v : Ymx) v = std::sqrt(v);
    // within one shell the reference fills both orientations, and (ab|ab) = (ba|ba):
    // the cross-type entries of an SP shell land transposed here, so symmetrise
    for (int s = 0; s < ns; ++s)
        for (int u = 0; u < c->kprim[s]; ++u)
            for (int v = u + 1; v < c->kprim[s]; ++v) {
                const size_t i1 = (c->kstart[s] - 1 + u) + (size_t)jb * (c->kstart[s] - 1 + v);
                const size_t i2 = (c->kstart[s] - 1 + v) + (size_t)jb * (c->kstart[s] - 1 + u);
                Pmx[i1] = Pmx[i2] = std::max(Pmx[i1], Pmx[i2]);
            }
    std::vector<double> cp((size_t)jb * jb, 0.0);
    for (size_t k = 0; k < Pmx.size(); ++k) cp[k] = std::sqrt(Pmx[k]);
    if (Ycut_out) memcpy(Ycut_out, Ymx.data(), Ymx.size() * sizeof(double));
    if (cutprim_out) memcpy(cutprim_out, cp.data(), cp.size() * sizeof(double));
    return qk_set_schwarz(c, Ymx.data(), cp.data());
}

// ---------------------------------------------------------------------------
// Nuclear gradient (four-centre part)

// Work lists of the gradient: the prepared pair groups, every bra with the
// prefix of kets passing Y_ij*Y_kl > cutoff, in runs of <= chunk quartets (a
// wavefront works through a run one quartet after another).  Mixed precision:
// the kets with Y_ij*Y_kl below the FP32 threshold (a suffix of the prefix, as
// kets are sorted by Y) go to the FP32 list.
static int grad_prepare(qk_ctx* c, const qk_opts* o) {
    const bool mixed = o->precision == QK_MIXED;
    const double cutoff = o->tau;
    if (c->grad_prepared && c->grad_prep_cutoff == cutoff && c->grad_prep_precision == o->precision &&
        (!mixed || c->grad_prep_thr == o->fp32_threshold))
        return 0;
    (void)hipSetDevice(c->device);
    for (int k = 0; k < kNumQuartetClasses; ++k)
        if (!c->d_gtab[k].p) {
            std::vector<int> blob(kGradBlobs[k], kGradBlobs[k] + kGradBlobLen[k]);
            if (c->d_gtab[k].upload(blob)) return -1;
        }
    // the work lists on the device: (bra group, ket group, bra) order, so that
    // neighbouring wavefronts read the same kets (measured faster than
    // bra-major order for the gradient); light classes one quartet per lane in
    // tiles of up to 64 kets, the table-driven kernel one quartet at a time per
    // wavefront in tiles of 16, 4 or 1
    std::vector<ListEntry> E;
    std::vector<KetGroup> KG;
    std::vector<ListClass> LC(kNumQuartetClasses);
    std::vector<long long> cls_item(kNumQuartetClasses + 1), cls_bra(kNumQuartetClasses + 1, 0);
    long long nitem = 0;
    const size_t ng = c->grp_cls.size();
    for (int k = 0; k < kNumQuartetClasses; ++k) {
        cls_item[k] = nitem;
        const int cb = pair_class_of(kClassL[k][0], kClassL[k][1]);
        const int ck = pair_class_of(kClassL[k][2], kClassL[k][3]);
        const int LT = kClassL[k][0] + kClassL[k][1] + kClassL[k][2] + kClassL[k][3];
        const bool light = grad_class_is_light(k);
        const int chunk = light ? 64 : (LT <= 3 ? 16 : (LT <= 5 ? 4 : 1));
        LC[k] = ListClass{chunk, INT_MAX, INT_MAX, mixed ? 1 : 0, light ? 1 : 0, 1, 0, 0};
        for (size_t gb = 0; gb < ng; ++gb) {
            if (c->grp_cls[gb] != cb) continue;
            for (size_t gk = 0; gk < ng; ++gk) {
                if (c->grp_cls[gk] != ck || (cb == ck && gk < gb)) continue;
                KG.push_back(KetGroup{c->grp_start[gk], c->grp_count[gk], gb == gk ? 1 : 0, 0});
                E.push_back(ListEntry{c->grp_start[gb], c->grp_count[gb], 1, (int)KG.size() - 1, nitem, nitem, k, 0});
                nitem += c->grp_count[gb];
            }
        }
    }
    cls_item[kNumQuartetClasses] = nitem;
    const ListSumsFn grad_dst = [&](const std::vector<long long>& sm, std::vector<ListDst>& dst) -> int {
        for (int k = 0; k < kNumQuartetClasses; ++k) {
            const long long* v = &sm[(size_t)k * kListCounts];
            if (v[0] > INT32_MAX || v[1] > INT32_MAX || v[4] > INT32_MAX || v[5] > INT32_MAX) {
                set_err("gradient class %d: more than 2^31 tiles or candidate quartets", k);
                return -1;
            }
            c->ngtiles[k] = v[0];
            c->ngtiles32[k] = v[1];
            c->ngsegs[0][k] = v[2];
            c->ngsegs[1][k] = v[3];
        }
        // every class's lists in one allocation
        Arena& A = c->grad_arena;
        A.reset();
        std::vector<int> slot((size_t)kNumQuartetClasses * 12);
        for (int k = 0; k < kNumQuartetClasses; ++k) {
            const long long* v = &sm[(size_t)k * kListCounts];
            const bool light = grad_class_is_light(k);
            int* sl = &slot[(size_t)k * 12];
            sl[0] = A.reserve<Tile>(v[0]);
            sl[1] = A.reserve<Tile>(v[1]);
            for (int f = 0; f < 2; ++f) {
                const long long nslot = light ? v[f] : 0;   // a light class's segment owns the slots of its tiles
                sl[2 + 5 * f] = A.reserve<int4>(v[2 + f]);
                sl[3 + 5 * f] = A.reserve<int>(v[2 + f]);
                sl[4 + 5 * f] = A.reserve<Tile>(std::max<long long>(nslot, 1));
                sl[5 + 5 * f] = A.reserve<int>(std::max<long long>(v[4 + f], 1));
                sl[6 + 5 * f] = A.reserve<int>(std::max<long long>(v[2 + f], 1));
            }
        }
        if (A.commit()) return -1;
        for (int k = 0; k < kNumQuartetClasses; ++k) {
            const long long* v = &sm[(size_t)k * kListCounts];
            const bool light = grad_class_is_light(k);
            const int* sl = &slot[(size_t)k * 12];
            A.place(c->d_gtiles[k], sl[0], v[0]);
            A.place(c->d_gtiles32[k], sl[1], v[1]);
            for (int f = 0; f < 2; ++f) {
                const long long nslot = light ? v[f] : 0;
                A.place(c->d_gsegs[f][k], sl[2 + 5 * f], v[2 + f]);
                A.place(c->d_gslot0[f][k], sl[3 + 5 * f], v[2 + f]);
                A.place(c->d_gsteps[f][k], sl[4 + 5 * f], std::max<long long>(nslot, 1));
                A.place(c->d_glist[f][k], sl[5 + 5 * f], std::max<long long>(v[4 + f], 1));
                A.place(c->d_gcounts[f][k], sl[6 + 5 * f], std::max<long long>(v[2 + f], 1));
            }
            dst[k] = ListDst{c->d_gtiles[k].p, c->d_gtiles32[k].p, c->d_gsegs[0][k].p, c->d_gsegs[1][k].p,
                             c->d_gslot0[0][k].p, c->d_gslot0[1][k].p, nullptr, nullptr, nullptr, nullptr};
        }
        return 0;
    };
    if (build_lists(c, E, KG, LC, cls_item, cls_bra, cutoff, mixed ? o->fp32_threshold : 0.0, grad_dst, nullptr))
        return -1;
    if (c->d_gacc.alloc(3 * (size_t)c->natom) || c->d_gcounters.alloc(kNumGradCounters)) return -1;
    c->grad_prep_cutoff = cutoff;
    c->grad_prep_precision = o->precision;
    c->grad_prep_thr = o->fp32_threshold;
    c->grad_prepared = true;
    return 0;
}

// Whether the light class k's list is density-compacted before its kernel.
// The compaction pass costs about one gather-heavy test per candidate quartet;
// it pays where lanes are expensive and many candidates fail the density test.
// Measured per class on (H2O)32 (compaction everywhere against nowhere): every
// light class of an FP64 gradient gains or is even (dpps 17.1 -> 12.7 ms, pppp
// 17.6 -> 14.4 ms, ssss even); in a mixed gradient the FP64 list holds the
// short ket ranges of the strongest pairs, where pppp (3.7 vs 3.8 ms) and ssss
// (1.16 -> 0.99 ms) gain and the other light classes lose up to 0.35 ms
// (psss); of the FP32 lists those of total angular momentum >= 3 gain (dpps
// 2.81 -> 2.45 ms) while psss, ssss and ppss lose.
static bool grad_compacts(int k, bool fp32, bool mixed) {
    const int LT = kClassL[k][0] + kClassL[k][1] + kClassL[k][2] + kClassL[k][3];
    if (fp32) return LT >= 3;
    if (!mixed) return true;
    return LT == 0 || ncart_of(kClassL[k][0]) * ncart_of(kClassL[k][1]) * ncart_of(kClassL[k][2]) *
                              ncart_of(kClassL[k][3]) >= 81;
}

// dgrad <- dgrad + dE2/dR for density dD (cutmatrix from dD unless given)
static int grad_enqueue(qk_ctx* c, const double* dD, const double* dcutmat_or_null, double* dgrad,
                        const qk_opts* o, hipStream_t s) {
    if (prepare(c, o)) return -1;
    if (grad_prepare(c, o)) return -1;
    if (fold_timing(c)) return -1;
    const bool mixed = o->precision == QK_MIXED;
    if (mixed && !c->d_boys_tab.p) { set_err("mixed precision: Boys table missing"); return -1; }
    const int n = c->nbasis, ns = c->nshell;
    launch_scale_density(dD, c->d_cons.p, c->d_Dp.p, mixed ? c->d_Dpf.p : nullptr, n, s);
    if (dcutmat_or_null)
        HIPCHK(hipMemcpyAsync(c->d_cm.p, dcutmat_or_null, sizeof(double) * ns * ns, hipMemcpyDeviceToDevice, s));
    else
        launch_density_cutoff(dD, c->d_sh_f0.p, c->d_sh_nf.p, c->d_cm.p, n, ns, s);
    launch_fill_u64(c->d_gacc.p, 3LL * c->natom, s);
    launch_fill_u64(c->d_gcounters.p, kNumGradCounters, s);
    HIPCHK(hipGetLastError());
    GradArgs a;
    a.q = {c->dq_atom.p, c->dq_f0.p, c->dq_shell.p, c->dq_qf.p, c->dq_nc.p};
    a.pr = {c->dp_qa.p, c->dp_qb.p, c->dp_prim0.p, c->dp_pstride.p, c->dp_nprim.p, c->dp_Y.p};
    a.pp = {c->dk_zeta.p, c->dk_oo2z.p, c->dk_Px.p, c->dk_Py.p, c->dk_Pz.p, c->dk_K.p, c->dk_cut.p};
    a.ea = c->dk_ea.p;
    a.eb = c->dk_eb.p;
    a.xyz = c->d_xyz.p;
    a.Dp = c->d_Dp.p;
    a.cutmat = c->d_cm.p;
    a.G = c->d_gacc.p;
    a.n = n;
    a.nshell = ns;
    a.cutoff = o->tau;
    a.hybrid = o->hybrid;
    a.natom = c->natom;
    a.lds_grad = 3 * c->natom <= 8192 ? 1 : 0;   // per-block LDS copy of the gradient up to 64 KB
    a.ppf = {c->dkf_zeta.p, c->dkf_oo2z.p, c->dkf_Px.p, c->dkf_Py.p, c->dkf_Pz.p, c->dkf_K.p, c->dkf_cut.p};
    a.eaf = c->dkf_ea.p;
    a.ebf = c->dkf_eb.p;
    a.Dpf = c->d_Dpf.p;
    a.boys_tab = c->d_boys_tab.p;
    CompactArgs ca;   // the gradient's screen (one-centre skip, Y Y > cutoff, Y Y DNmax > cutoff)
    ca.psh = c->dp_sh.p;
    ca.pat = c->dp_at.p;
    ca.Y = a.pr.Y;
    ca.cutmat = a.cutmat;
    ca.nshell = ns;
    ca.tau = o->tau;
    ca.grad = 1;
    ca.steps = nullptr;
    ca.slot0 = nullptr;
    // Two streams, forked from the caller's and joined with events.  A mixed
    // gradient's FP64 list holds the few strongest pairs: its kernels have
    // little parallelism and long per-quartet chains, so they run on the second
    // stream beside the FP32 list.  In an FP64 gradient the classes alternate
    // between the streams (measured: (H2O)32 188.8 -> 172 ms, n-icosane 113.9
    // -> 102.5 ms).  One kernel per stream is in flight, each within its grid's
    // scratch bound, so at most two bounded backings are live.
    const hipStream_t s64 = c->streams[0];
    HIPCHK(hipEventRecord(c->ev_start, s));
    HIPCHK(hipStreamWaitEvent(s64, c->ev_start, 0));
    // When the FP64 lists are light (under 7.5% of the candidate tiles), the
    // FP32 lists of (dp|ds) and of the table-driven d-d bra classes ((dd|ps),
    // (dd|pp), (dd|ds), (dd|dp), (dd|dd)) follow them on the second stream.
    // Measured on mixed gradients at 1e-2: (H2O)32, FP64 lists 4.0% of the
    // tiles, 82.3 -> 76.2 ms (the second stream had stood idle for half the
    // call); n-icosane, 12.2%, where moving them cost 54.2 -> 56.5 ms (its two
    // streams were already even), so there they stay; moving every
    // table-driven class cost both.
    constexpr unsigned kFp32OnSecond =
        (1u << 13) | (1u << 16) | (1u << 17) | (1u << 18) | (1u << 19) | (1u << 20);
    long long tiles64 = 0, tiles32 = 0;
    for (int k = 0; k < kNumQuartetClasses; ++k) {
        tiles64 += c->ngtiles[k];
        tiles32 += c->ngtiles32[k];
    }
    const unsigned on_second = mixed && tiles64 * 40 < (tiles64 + tiles32) * 3 ? kFp32OnSecond : 0u;
    for (int pass = 0; pass < (mixed ? 2 : 1); ++pass) {
        const bool fp32 = pass == 1;
        for (int k = 0; k < kNumQuartetClasses; ++k) {
            const hipStream_t ls = fp32 ? (((on_second >> k) & 1u) ? s64 : s) : (!mixed && (k & 1) ? s : s64);
            const long long nt = fp32 ? c->ngtiles32[k] : c->ngtiles[k];
            if (nt == 0) continue;
            ClassWork w{(fp32 ? c->d_gtiles32[k] : c->d_gtiles[k]).p, nt};
            a.counters = c->d_gcounters.p + 2 * (fp32 ? kNumQuartetClasses + k : k);
            const int slot = (fp32 ? 3 : 2) * kNumQuartetClasses + k;
            if (c->timing) HIPCHK(hipEventRecord(c->ev_t0[slot], ls));
            if (grad_class_is_light(k) && grad_compacts(k, fp32, mixed)) {
                // density compaction, then the kernel over the queued steps
                const int f = fp32 ? 1 : 0;
                CompactArgs cq = ca;
                cq.steps = c->d_gsteps[f][k].p;
                cq.slot0 = c->d_gslot0[f][k].p;
                launch_compact(cq, c->d_gsegs[f][k].p, c->ngsegs[f][k], c->d_glist[f][k].p,
                               c->d_gcounts[f][k].p, ls);
                w.list = c->d_glist[f][k].p;
                w.steps = cq.steps;
                const int grid = (int)std::min<long long>((nt + 3) / 4, 256LL * 8);
                if (launch_grad_light(k, a, w, grid, fp32, ls)) {
                    set_err("gradient kernel launch failed for class %d", k);
                    return -1;
                }
            } else if (grad_class_is_light(k)) {
                const int grid = (int)std::min<long long>((nt + 3) / 4, 256LL * 8);
                if (launch_grad_light(k, a, w, grid, fp32, ls)) {
                    set_err("gradient kernel launch failed for class %d", k);
                    return -1;
                }
            } else {
                const size_t lds = grad_lds_bytes(kGradBlobs[k], fp32);
                // at most 16 blocks per CU (with one-wavefront blocks: four
                // wavefronts per SIMD), 20 for the FP32 form (five), except for
                // dppp, measured slower at five (12.4 ms at four, 16.0 at five on
                // (H2O)32): it runs the FP32 form built for four
                const bool dppp = kClassL[k][0] == 2 && kClassL[k][1] == 1 && kClassL[k][2] == 1 && kClassL[k][3] == 1;
                const int grid = (int)std::min<long long>(nt, 256LL * (fp32 && !dppp ? 20 : 16));
                if (launch_grad_class(a, w, kGradBlobs[k], c->d_gtab[k].p, lds, grid, fp32, dppp, ls)) {
                    set_err("gradient kernel launch failed for class %d (%zu bytes LDS)", k, lds);
                    return -1;
                }
            }
            if (c->timing) HIPCHK(hipEventRecord(c->ev_t1[slot], ls));
            c->cls_timed[slot] = c->timing;
        }
    }
    HIPCHK(hipEventRecord(c->ev_cls[0], s64));
    HIPCHK(hipStreamWaitEvent(s, c->ev_cls[0], 0));
    launch_grad_finalize(c->d_gacc.p, dgrad, 3 * c->natom, c->d_gcounters.p + kGradRangeFlag, s);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int qk_grad_device(qk_ctx* c, const double* dD, double* dgrad, const qk_opts* o, void* stream) {
    if (!c) { set_err("null context"); return -1; }
    qk_opts d;
    if (!o) { qk_default_opts(&d); o = &d; }
    (void)hipSetDevice(c->device);
    return grad_enqueue(c, dD, nullptr, dgrad, o, (hipStream_t)stream);
}

extern "C" int qk_grad_host(qk_ctx* c, const double* D, double* grad, const qk_opts* o) {
    if (!c) { set_err("null context"); return -1; }
    qk_opts d;
    if (!o) { qk_default_opts(&d); o = &d; }
    (void)hipSetDevice(c->device);
    const size_t nn = (size_t)c->nbasis * c->nbasis, m = 3 * (size_t)c->natom;
    if (c->d_D.alloc(nn) || c->d_grad.alloc(m)) return -1;
    HIPCHK(hipMemcpy(c->d_D.p, D, nn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_grad.p, grad, m * sizeof(double), hipMemcpyHostToDevice));
    if (grad_enqueue(c, c->d_D.p, nullptr, c->d_grad.p, o, nullptr)) return -1;
    HIPCHK(hipMemcpy(grad, c->d_grad.p, m * sizeof(double), hipMemcpyDeviceToHost));
    return check_range(c->d_gcounters.p + kGradRangeFlag, "qk_grad_host");
}

int qk_context_info(qk_ctx* c, int* device, int* nbasis) {
    if (!c || !c->have_system) { set_err("no system"); return -1; }
    *device = c->device;
    *nbasis = c->nbasis;
    return 0;
}

// ---------------------------------------------------------------------------
// one-electron integrals and their gradient terms

extern "C" int qk_set_charges(qk_ctx* c, const double* charges) {
    if (!c || !c->have_system) { set_err("qk_set_charges: no system"); return -1; }
    (void)hipSetDevice(c->device);
    c->chg.assign(charges, charges + c->natom);
    return c->d_chg.upload(c->chg);
}

static int oneint_ready(qk_ctx* c, const char* what) {
    if (!c || !c->have_system) { set_err("%s: no system", what); return -1; }
    if ((int)c->chg.size() != c->natom) { set_err("%s: nuclear charges not set (qk_set_charges)", what); return -1; }
    (void)hipSetDevice(c->device);
    return 0;
}

static OneFuncs one_funcs(const qk_ctx* c) {
    return OneFuncs{c->d1_atom.p, c->d1_lmn.p, c->d1_np.p, c->d1_qf.p, c->d1_a.p, c->d1_c.p};
}

extern "C" int qk_oneint_device(qk_ctx* c, double* dS, double* dH, void* stream) {
    if (oneint_ready(c, "qk_oneint_device")) return -1;
    const OneSys sys{c->d_xyz.p, c->d_chg.p, c->natom, c->nbasis};
    const int2* pl[5] = {c->d1_pairs[0].p, c->d1_pairs[1].p, c->d1_pairs[2].p, c->d1_pairs[3].p, c->d1_pairs[4].p};
    launch_oneint(one_funcs(c), sys, pl, c->n1_pairs, dS, dH, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int qk_nuclear_repulsion(qk_ctx* c, double* E) {
    if (oneint_ready(c, "qk_nuclear_repulsion")) return -1;
    double e = 0.0;
    for (int i = 0; i < c->natom; ++i)
        for (int j = i + 1; j < c->natom; ++j) {
            const double dx = c->xyz[3 * i] - c->xyz[3 * j], dy = c->xyz[3 * i + 1] - c->xyz[3 * j + 1],
                         dz = c->xyz[3 * i + 2] - c->xyz[3 * j + 2];
            e += c->chg[i] * c->chg[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    *E = e;
    return 0;
}

extern "C" int qk_oneint_grad_device(qk_ctx* c, const double* dD, const double* dW, double* dgrad, void* stream) {
    if (oneint_ready(c, "qk_oneint_grad_device")) return -1;
    hipStream_t s = (hipStream_t)stream;
    const int m = 3 * c->natom;
    if (8LL * m > 150 * 1024) {   // the kernel's per-block LDS copy of the gradient
        set_err("qk_oneint_grad_device: %d atoms; at most 6400 are supported", c->natom);
        return -2;
    }
    if (c->d1_acc.alloc(m + 1)) return -1;
    launch_fill_u64(c->d1_acc.p, m + 1, s);
    const OneSys sys{c->d_xyz.p, c->d_chg.p, c->natom, c->nbasis};
    const int2* pl[5] = {c->d1_pairs[0].p, c->d1_pairs[1].p, c->d1_pairs[2].p, c->d1_pairs[3].p, c->d1_pairs[4].p};
    launch_oneint_grad(one_funcs(c), sys, pl, c->n1_pairs, dD, dW, c->d1_acc.p, s);
    launch_grad_finalize(c->d1_acc.p, dgrad, m, c->d1_acc.p + m, s);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int qk_oneint_host(qk_ctx* c, double* S, double* H) {
    if (oneint_ready(c, "qk_oneint_host")) return -1;
    const size_t nn = (size_t)c->nbasis * c->nbasis;
    DBuf<double> dS, dH;
    if (dS.alloc(nn) || dH.alloc(nn)) return -1;
    if (qk_oneint_device(c, dS.p, dH.p, nullptr)) return -1;
    if (S) HIPCHK(hipMemcpy(S, dS.p, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (H) HIPCHK(hipMemcpy(H, dH.p, nn * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int qk_oneint_grad_host(qk_ctx* c, const double* D, const double* W, double* grad) {
    if (oneint_ready(c, "qk_oneint_grad_host")) return -1;
    const size_t nn = (size_t)c->nbasis * c->nbasis, m = 3 * (size_t)c->natom;
    DBuf<double> dD, dW, dg;
    if (dD.alloc(nn) || dW.alloc(nn) || dg.alloc(m)) return -1;
    HIPCHK(hipMemcpy(dD.p, D, nn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dW.p, W, nn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dg.p, grad, m * sizeof(double), hipMemcpyHostToDevice));
    if (qk_oneint_grad_device(c, dD.p, dW.p, dg.p, nullptr)) return -1;
    HIPCHK(hipMemcpy(grad, dg.p, m * sizeof(double), hipMemcpyDeviceToHost));
    return check_range(c->d1_acc.p + m, "qk_oneint_grad_host");
}

// the forward/backward pair of the test harness (same operations as the
// host entry points above)
extern "C" int qk_fock_build(qk_ctx* c, const double* D, double* F, const qk_opts* o) {
    return qk_fock_host(c, D, F, o);
}

// W NULL: the four-centre term only; W given: the complete dE/dR of hfgrad
// (nuclear repulsion, one-electron terms with W, four-centre term), which needs
// the nuclear charges (qk_set_charges)
extern "C" int qk_fock_grad(qk_ctx* c, const double* D, const double* W, double* grad, const qk_opts* o) {
    if (W && oneint_ready(c, "qk_fock_grad")) return -4;
    if (qk_grad_host(c, D, grad, o)) return -1;
    return W ? qk_oneint_grad_host(c, D, W, grad) : 0;
}

// surviving quartets and primitive quartets of class k in the last gradient
// (FP64 part), and its kernel time
extern "C" int qk_grad_class_stats(qk_ctx* c, int k, long long* nq, long long* np, float* ms) {
    if (!c || k < 0 || k >= kNumQuartetClasses) { set_err("bad arguments"); return -1; }
    (void)hipSetDevice(c->device);
    HIPCHK(hipDeviceSynchronize());
    *ms = 0.0f;
    if (!c->d_gcounters.p) { *nq = *np = 0; return 0; }
    unsigned long long h[2];
    HIPCHK(hipMemcpy(h, c->d_gcounters.p + 2 * k, sizeof(h), hipMemcpyDeviceToHost));
    *nq = (long long)h[0];
    *np = (long long)h[1];
    return slot_ms(c, 2 * kNumQuartetClasses + k, ms);
}

// the same for the FP32 part of a mixed-precision gradient
extern "C" int qk_grad_class_stats_fp32(qk_ctx* c, int k, long long* nq, long long* np, float* ms) {
    if (!c || k < 0 || k >= kNumQuartetClasses) { set_err("bad arguments"); return -1; }
    (void)hipSetDevice(c->device);
    HIPCHK(hipDeviceSynchronize());
    *ms = 0.0f;
    if (!c->d_gcounters.p) { *nq = *np = 0; return 0; }
    unsigned long long h[2];
    HIPCHK(hipMemcpy(h, c->d_gcounters.p + 2 * (kNumQuartetClasses + k), sizeof(h), hipMemcpyDeviceToHost));
    *nq = (long long)h[0];
    *np = (long long)h[1];
    return slot_ms(c, 3 * kNumQuartetClasses + k, ms);
}

// ---------------------------------------------------------------------------
// reference boundary (process-global context, gpu.cu)

static qk_ctx* g_ctx = nullptr;
static int g_dev_req = -1;
static bool g_started = false;
static int g_natom = 0, g_nbasis = 0;
static std::vector<double> g_xyz;
static std::vector<double> g_atom_chg;   // nuclear charges (gpu_upload_atom_and_chg_)
static std::vector<int> g_atom_type;     // atomic numbers
// gpu_upload_method_ may come before gpu_init_ (QUICK calls it per operator
// build); the method is kept here and read by every build
static int g_method = 0;

extern "C" qk_ctx* qk_global_context(void) { return g_ctx; }

static void fatal(const char* what) {
    fprintf(stderr, "quick_mi355x: fatal: %s: %s\n", what, g_err.c_str());
    fflush(stderr);
    std::abort();
}

static double hybrid_of_method(int m) { return m == 0 ? 1.0 : (m == 1 ? 0.2 : 0.0); }

extern "C" void gpu_startup_(void) { g_started = true; }
extern "C" void gpu_set_device_(int* id) { g_dev_req = id ? *id : -1; }
extern "C" void gpu_init_(void) {
    if (g_ctx) return;
    g_ctx = qk_create(g_dev_req < 0 ? 0 : g_dev_req);
    if (!g_ctx) fatal("gpu_init");
}
extern "C" void gpu_get_device_info_(int* count, int* id, int* memMB, int* nCU, double* GHz, char* name,
                                     int* namelen, int* major, int* minor) {
    int n = 0;
    (void)hipGetDeviceCount(&n);
    *count = n;
    int dev = g_ctx ? g_ctx->device : (g_dev_req < 0 ? 0 : g_dev_req);
    *id = dev;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) {
        set_err("hipGetDeviceProperties failed");
        *memMB = 0; *nCU = 0; *GHz = 0; *namelen = 0; *major = 0; *minor = 0;
        return;
    }
    *memMB = (int)(p.totalGlobalMem / (1024 * 1024));
    *nCU = p.multiProcessorCount;
    *GHz = p.clockRate * 1e-6;
    const int L = (int)strnlen(p.gcnArchName, sizeof(p.gcnArchName));
    memcpy(name, p.gcnArchName, L);
    *namelen = L;
    *major = p.major;
    *minor = p.minor;
}
extern "C" void gpu_setup_(int* natom, int* nbasis, int*, int*, int*, int*) {
    g_natom = *natom;
    g_nbasis = *nbasis;
}
extern "C" void gpu_upload_xyz_(double* xyz) { g_xyz.assign(xyz, xyz + 3 * (size_t)g_natom); }
extern "C" void gpu_upload_atom_and_chg_(int* atom_type, double* chg) {
    g_atom_chg.assign(chg, chg + (size_t)g_natom);
    g_atom_type.assign(atom_type, atom_type + (size_t)g_natom);
}

// the per-function arrays the XC step reads (gpu_getxc_), kept from gpu_upload_basis_
static int g_maxcontract = 0;
static std::vector<int> g_ncontract, g_itype, g_ncenter;
static std::vector<double> g_aexp, g_dcoeff;
static long long g_geometry = 0;        // bumped by every basis upload
static qk_xc* g_xc = nullptr;
static long long g_xc_geometry = -1;

extern "C" void gpu_upload_basis_(int* nshell, int*, int*, int* jbasis, int* maxcontract, int* ncontract, int* itype,
                                  double* aexp, double* dcoeff, int*, int*, int*, int*, int* ncenter, int* kstart,
                                  int* katom, int* ktype, int* kprim, int*, int* Ksumtype, int*, int* Qstart,
                                  int* Qfinal, int* Qsbasis, int* Qfbasis, double* gccoeff, double* cons,
                                  double* gcexpo, int* KLMN) {
    if (!g_ctx) gpu_init_();
    if (qk_set_system(g_ctx, g_natom, g_xyz.data(), *nshell, g_nbasis, *jbasis, kstart, kprim, katom, ktype, Qstart,
                      Qfinal, Qsbasis, Qfbasis, Ksumtype, KLMN, cons, gccoeff, gcexpo))
        fatal("gpu_upload_basis");
    const size_t n = (size_t)g_nbasis, mc = (size_t)*maxcontract;
    g_maxcontract = *maxcontract;
    g_ncontract.assign(ncontract, ncontract + n);
    g_itype.assign(itype, itype + 3 * n);
    g_ncenter.assign(ncenter, ncenter + n);
    g_aexp.assign(aexp, aexp + mc * n);
    g_dcoeff.assign(dcoeff, dcoeff + mc * n);
    ++g_geometry;
}
extern "C" void gpu_upload_cutoff_matrix_(double* Y, double* cutPrim) {
    if (!g_ctx || qk_set_schwarz(g_ctx, Y, cutPrim)) fatal("gpu_upload_cutoff_matrix");
}
extern "C" void gpu_upload_method_(int* m) {
    g_method = *m;
}
extern "C" void gpu_upload_calculated_(double*, double*, double*, double* dense) {
    if (!g_ctx) fatal("gpu_upload_calculated (no context)");
    const size_t nn = (size_t)g_ctx->nbasis * g_ctx->nbasis;
    if (g_ctx->d_D.alloc(nn) || g_ctx->d_G.alloc(nn)) fatal("gpu_upload_calculated");
    if (hipMemcpy(g_ctx->d_D.p, dense, nn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        set_err("density upload failed");
        fatal("gpu_upload_calculated");
    }
    g_ctx->abi_have_dense = true;
}
// The caller's cutMatrix goes into a device buffer kept with the context (one
// allocation per system size); the next gpu_get2e_ / gpu_grad_ / gpu_getxc_
// consumes it, as the reference frees it after the build (gpu.cu:1857-1861).
extern "C" void gpu_upload_cutoff_(double* cutMatrix, double* integralCutoff, double* primLimit, double*) {
    if (!g_ctx) fatal("gpu_upload_cutoff (no context)");
    qk_ctx* c = g_ctx;
    c->abi_tau = *integralCutoff;
    c->abi_prim_limit = *primLimit;
    const size_t m = (size_t)c->nshell * c->nshell;
    if (c->d_abi_cm.alloc(m) ||
        hipMemcpy(c->d_abi_cm.p, cutMatrix, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        set_err("cutMatrix upload failed");
        fatal("gpu_upload_cutoff");
    }
    c->abi_have_cutmatrix = true;
}
// the per-cycle uploads are consumed by the call that reads them
static void abi_consume(qk_ctx* c) {
    c->abi_have_dense = false;
    c->abi_have_cutmatrix = false;
}
extern "C" void gpu_get2e_(double* o) {
    qk_ctx* c = g_ctx;
    if (!c || !c->abi_have_dense) {
        set_err("gpu_get2e: density not uploaded (call gpu_upload_calculated first)");
        fatal("gpu_get2e");
    }
    qk_opts op;
    qk_default_opts(&op);
    op.tau = c->abi_tau;
    op.prim_limit = c->abi_prim_limit;
    op.hybrid = hybrid_of_method(g_method);
    const size_t nn = (size_t)c->nbasis * c->nbasis;
    const double* cmp = c->abi_have_cutmatrix ? c->d_abi_cm.p : nullptr;
    (void)hipSetDevice(c->device);
    if (fock_enqueue(c, c->d_D.p, cmp, c->d_G.p, &op, nullptr)) fatal("gpu_get2e");
    std::vector<double> G(nn);
    if (hipMemcpy(G.data(), c->d_G.p, nn * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        set_err("operator download failed");
        fatal("gpu_get2e");
    }
    if (check_range(c->d_counters.p + kRangeFlag, "gpu_get2e")) fatal("gpu_get2e");
    for (size_t k = 0; k < nn; ++k) o[k] += G[k];
    abi_consume(c);
}
// Keeps the gradient seed (the caller's partial gradient, 3 natom) and the
// gradient cutoff; the next gpu_grad_ writes seed + four-centre term
// (gpu.cu:1393-1443, 1481-1530).
extern "C" void gpu_upload_grad_(double* grad, double* gradCutoff) {
    if (!g_ctx) fatal("gpu_upload_grad (no context)");
    g_ctx->abi_grad_seed.assign(grad, grad + 3 * (size_t)g_ctx->natom);
    g_ctx->abi_grad_cutoff = *gradCutoff;
    g_ctx->abi_have_grad_seed = true;
}
// grad <- seed + the four-centre gradient of the uploaded density.  Screening
// follows the reference CPU path (gradient.f90:289-316, shell.f90:1941): the
// quartet test and the primitive test both use the gradient cutoff from
// gpu_upload_grad_.  (The reference GPU kernel screens with integralCutoff and
// primLimit instead; QUICK's CUTOFF= keyword sets integralCutoff = gradCutoff.)
extern "C" void gpu_grad_(double* grad) {
    qk_ctx* c = g_ctx;
    if (!c || !c->abi_have_dense) {
        set_err("gpu_grad: density not uploaded (call gpu_upload_calculated first)");
        fatal("gpu_grad");
    }
    if (!c->abi_have_grad_seed) {
        set_err("gpu_grad: gradient not uploaded (call gpu_upload_grad first)");
        fatal("gpu_grad");
    }
    qk_opts op;
    qk_default_opts(&op);
    op.tau = c->abi_grad_cutoff;
    op.hybrid = hybrid_of_method(g_method);
    const double* cmp = c->abi_have_cutmatrix ? c->d_abi_cm.p : nullptr;
    (void)hipSetDevice(c->device);
    const size_t m = 3 * (size_t)c->natom;
    if (c->d_grad.alloc(m) ||
        hipMemcpy(c->d_grad.p, c->abi_grad_seed.data(), m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        fatal("gpu_grad");
    if (grad_enqueue(c, c->d_D.p, cmp, c->d_grad.p, &op, nullptr)) fatal("gpu_grad");
    if (hipMemcpy(grad, c->d_grad.p, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) fatal("gpu_grad");
    if (check_range(c->d_gcounters.p + kGradRangeFlag, "gpu_grad")) fatal("gpu_grad");
    abi_consume(c);
    c->abi_have_grad_seed = false;
}
// B3LYP XC on the SG-1 grid for the uploaded density (gpu.cu:1866-1975):
// o += V_xc, Eelxc += E_xc, aelec and belec += N/2; sigrad2 gives the radii of
// significance of the basis functions (get_sigrad).  The grid is built on the
// device once per geometry.
extern "C" void gpu_getxc_(int* isg, double* sigrad2, double* Eelxc, double* aelec, double* belec, double* o) {
    qk_ctx* c = g_ctx;
    if (!c || !c->abi_have_dense) {
        set_err("gpu_getxc: density not uploaded (call gpu_upload_calculated first)");
        fatal("gpu_getxc");
    }
    if (*isg != 1) {
        set_err("gpu_getxc: only the SG-1 grid (isg = 1) is provided");
        fatal("gpu_getxc");
    }
    if ((int)g_atom_type.size() != g_natom || g_ncontract.empty()) {
        set_err("gpu_getxc: atoms or basis not uploaded");
        fatal("gpu_getxc");
    }
    if (!g_xc || g_xc_geometry != g_geometry) {
        if (!g_xc) g_xc = qk_xc_create(c->device);
        if (!g_xc ||
            qk_xc_set_system(g_xc, g_natom, g_xyz.data(), g_atom_type.data(), g_nbasis, g_maxcontract,
                             g_ncontract.data(), g_itype.data(), g_aexp.data(), g_dcoeff.data(), g_ncenter.data(),
                             sigrad2, 1e-10))
            fatal("gpu_getxc");
        qk_xc_opts xo;
        qk_xc_default_opts(&xo);
        if (qk_xc_build_grid(g_xc, &xo)) fatal("gpu_getxc");
        g_xc_geometry = g_geometry;
    }
    (void)hipSetDevice(c->device);
    double E = 0.0, N = 0.0;
    if (qk_xc_device(g_xc, c->d_D.p, c->d_G.p, &E, &N, nullptr)) fatal("gpu_getxc");
    const size_t nn = (size_t)c->nbasis * c->nbasis;
    std::vector<double> V(nn);
    if (hipMemcpy(V.data(), c->d_G.p, nn * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        set_err("V_xc download failed");
        fatal("gpu_getxc");
    }
    for (size_t k = 0; k < nn; ++k) o[k] += V[k];
    *Eelxc += E;
    *aelec += 0.5 * N;
    *belec += 0.5 * N;
    abi_consume(c);
}
extern "C" void gpu_aoint_(double*, double*, int*, char*) {
    set_err("gpu_aoint: stored-integral (NODIRECT) mode is not provided by this library");
}
extern "C" void gpu_addint_(double*, int*, char*) {
    set_err("gpu_addint: stored-integral (NODIRECT) mode is not provided by this library");
}
extern "C" void gpu_cleanup_(void) {
    if (g_ctx) {
        g_ctx->prepared = false;
        g_ctx->have_schwarz = false;
    }
}
extern "C" void gpu_shutdown_(void) {
    qk_xc_destroy(g_xc);
    g_xc = nullptr;
    g_xc_geometry = -1;
    qk_destroy(g_ctx);
    g_ctx = nullptr;
    g_started = false;
}
