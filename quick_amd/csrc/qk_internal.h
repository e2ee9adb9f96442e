// Internal types shared by the host code (qk_host.cpp) and the kernels
// (qk_fock.hip).  Device arrays are structure-of-arrays so that a wavefront
// whose lanes walk consecutive kets of one bra pair reads the bra data as a
// broadcast and the ket data as coalesced loads.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

// sets the text qk_last_error() returns (this thread)
void qk_set_error_text(const char* msg);

namespace qk {

constexpr int kNumPairClasses = 6;     // ss ps pp ds dp dd (la >= lb)
constexpr int kNumQuartetClasses = 21;  // bra pair class >= ket pair class
// device counters: per class (surviving quartets, primitive quartets) for the
// FP64 kernels at [2k, 2k+1] and for the FP32 kernels at [2K+2k, 2K+2k+1];
// the last one is the range flag of the fixed-point accumulator (set by the
// finalize kernel when a sum comes within a factor of two of wrapping)
constexpr int kNumCounters = 4 * kNumQuartetClasses + 3;
constexpr int kRangeFlag = kNumCounters - 1;
// gradient counters: per class (quartets, primitive quartets), FP64 then FP32,
// then the gradient accumulator's range flag
constexpr int kNumGradCounters = 4 * kNumQuartetClasses + 1;
constexpr int kGradRangeFlag = kNumGradCounters - 1;
// The operator accumulator X is n x n int64 (G = (X + X^T)/2): the exchange
// rows of a bra are added to X once per run as whole rows, which a full matrix
// keeps contiguous (in a packed triangle half of each row is strided by the
// column, and the heaviest thread-per-quartet class ran twice as long).  It is
// kept in up to kXCopies copies; block b adds into copy b % copies, and the
// copies are summed at the end.  Consecutive blocks run on different XCDs, so
// each copy's atomics come mostly from one XCD.
constexpr int kXCopies = 8;
__host__ __device__ inline long long xtri_size(int n) { return (long long)n * (n + 1) / 2; }
// slot of the unordered pair {i, j} in the packed lower triangle, in 32-bit
// unsigned arithmetic (exact for n <= kMaxBasis: hi (hi + 1) < 2^32); the
// 64-bit form doubled the spills of the heaviest thread-per-quartet class
constexpr int kMaxBasis = 65535;
__host__ __device__ __forceinline__ unsigned xtri(int i, int j) {
    const unsigned hi = (unsigned)(i > j ? i : j), lo = (unsigned)(i > j ? j : i);
    return ((hi * (hi + 1u)) >> 1) + lo;
}
// Scratch (private segment) backing of a dispatch: the runtime sizes it as the
// kernel's per-lane scratch x 64 lanes x the waves the grid can have in flight
// (at most 8192 on 256 CUs).  Kernels that spill get their grid bounded so that
// this stays under kScratchBudget; their grid-stride loops cover the rest.
constexpr double kScratchBudget = 256.0 * 1024 * 1024;
constexpr long long kMaxWavesInFlight = 256LL * 32;
// largest grid (blocks of waves_per_block wavefronts) whose scratch backing stays
// under kScratchBudget for a kernel with lane_bytes of scratch per lane
inline long long scratch_grid_cap(long long lane_bytes, int waves_per_block) {
    if (lane_bytes <= 0) return 1LL << 40;
    const long long c = (long long)(kScratchBudget / (double(lane_bytes) * 64.0 * waves_per_block));
    return c < 1 ? 1 : c;
}
// per-lane scratch of a kernel (0 when the query fails)
inline long long kernel_scratch(const void* k) {
    hipFuncAttributes at{};
    return hipFuncGetAttributes(&at, k) == hipSuccess ? (long long)at.localSizeBytes : 0;
}

// Scratch of the no-allocation Fock entry point (qk_fock_device_noalloc): one
// fixed device array of this many bytes, defined in qk_fock.hip; every GPU the
// library runs on holds its own copy until the process ends.  It carries one
// accumulator copy (n^2 int64), the pair-indexed Coulomb buffer (at most
// n(n+1)/2 int64), the counters, the scaled density (n^2 doubles, plus n^2
// floats in mixed precision) and the shell-block density maxima (nshell^2
// doubles): 62 MB at 1600 functions in mixed precision.
constexpr long long kNoAllocScratchBytes = 64LL << 20;
// Looks the array's address up on `device` (hipGetSymbolAddress, with that
// device current) the first time it is asked for that device; returns its
// HIP error code (0 on success; later calls return the first call's code).
int noalloc_scratch_init(int device);
// the array's address on `device`, nullptr before a successful init
void* noalloc_scratch_ptr(int device);

constexpr int kFixedShift = 48;         // Fock accumulator: int64 in units of 2^-48
constexpr int kGradShift = 44;          // gradient accumulator: int64 in units of 2^-44

// Per Qshell (one angular momentum of one shell).
struct DevQshells {
    int* atom;    // 0-based atom
    int* f0;      // first basis function (0-based)
    int* shell;   // parent shell (0-based)
    int* qfinal;  // Qfinal of the parent shell (QUICK's Boys-order convention)
    int* nc;      // Cartesian functions of the Qshell (1, 3, 6)
};

// Per significant Qshell pair (oriented: L(a) >= L(b); equal L: a <= b).
struct DevPairs {
    int* qa;
    int* qb;
    int* prim0;    // index of the pair's first primitive pair
    int* pstride;  // distance between its consecutive primitive pairs
    int* nprim;    // number of primitive pairs
    double* Y;     // Schwarz bound of the parent shell pair
};

// Per primitive pair.  In the kernels' copy the primitive pairs of a group of
// pairs with equal primitive count are stored primitive-major: entry j of the
// pairs q, q+1, ... of a group is contiguous, so the lanes of a wavefront (one
// ket pair each) load primitive j of their kets with coalesced accesses.
struct DevPrims {
    double* zeta;  // a + b
    double* oo2z;  // 0.5 / (a + b)
    double* Px;
    double* Py;
    double* Pz;
    double* K;     // sqrt(2) pi^(5/4) exp(-ab/(a+b)|AB|^2)/(a+b) c_a c_b
    double* cut;   // QUICK cutprim of the primitive pair
};

// A wavefront's unit of work: one bra pair p and up to 64 consecutive ket pairs
// q0 .. q0+len-1 of one ket group (one per lane).  Bra and ket groups are runs
// of pairs of one pair class with equal primitive-pair count, sorted by Schwarz
// bound descending, so the kets passing Y_p*Y_q > tau form a prefix of the ket
// group for every bra; those prefixes are cut into tiles (qk_lists.hip).
struct Tile {
    int p;
    int q0;
    int len;
    int pad;
};

// Work-list construction on the device (qk_lists.hip).  An entry is a block of
// items of one quartet class: bra pairs s1 .. s1+n1-1 of one bra group, each
// against the m ket groups kg[kg0 .. kg0+m) (item item0 + i*m + t: bra i, ket
// group t).  The Fock lists have one entry per bra group, the gradient lists
// one per (bra group, ket group) with m = 1.  bra0 numbers the entry's bras
// across all entries (the Fock runs).
struct ListEntry {
    int s1, n1, m, kg0;
    long long item0, bra0;
    int cls, pad;
};
struct KetGroup {
    int s2, n2, same, pad;   // first pair, count; same: the bra group itself (kets from the bra on)
};
struct ListClass {
    int chunk;          // kets per tile
    int cap64, cap32;   // candidates per run and per segment (FP64, FP32 lists)
    int split;          // mixed precision: kets with Y_ij*Y_kl >= threshold stay FP64
    int segs;           // compaction segments are listed
    int grad;           // gradient form: one segment per item, with its first tile
    int runs;           // runs of tiles and segments per bra (Fock)
    int pad;
};
// per class: where the lists go (FP64 / FP32 tiles, segments, the gradient's
// first tile per segment, runs of tiles and of segments)
struct ListDst {
    Tile *t64, *t32;
    int4 *s64, *s32;
    int *slot64, *slot32;
    int2 *r64, *r32, *sr64, *sr32;
};
constexpr int kListCounts = 6;   // per item: tiles, tiles32, segs, segs32, candidates, candidates32
constexpr int kRunCounts = 4;    // per bra: runs, runs32, segment runs, segment runs32
// the primitive-pair fields copied primitive-major per group: zeta, 1/2zeta,
// Px, Py, Pz, K, cut, and the two exponents
constexpr int kPrimArrays = 9;
struct PrimArrays {
    double* a[kPrimArrays];
};
struct PrimArraysF {
    float* a[kPrimArrays];   // a[0] == nullptr: no FP32 copy
};

struct ClassWork {
    const Tile* tiles = nullptr;
    long long ntiles = 0;
    // optional: runs of tiles sharing one bra pair (first tile, end tile); when
    // set, a block takes one run at a time and sums the exchange contributions
    // to the bra's rows of X in LDS (n columns per bra function) before adding
    // them to X once per run
    const int2* runs = nullptr;
    long long nruns = 0;
    // optional (with runs): the density-screened quartets compacted per
    // segment (one bra pair, one ket group): segs[s] = {bra pair, first ket,
    // end ket, offset}, the surviving kets of segment s at list[offset ..
    // offset + counts[s]); runs then index segments
    const int4* segs = nullptr;
    const int* list = nullptr;
    const int* counts = nullptr;
    // optional (without runs): the steps written by compact_kernel, one per
    // tile and in tile order: tiles whose kets are list[q0 .. q0 + len)
    // (len 0: nothing survived there)
    const Tile* steps = nullptr;
};

// Inputs of the per-build density compaction (compact_kernel): per pair its
// two shells and two atoms (packed), Schwarz bounds and the density screen.
struct CompactArgs {
    const int* psh;   // per pair: parent shells of its Qshells (bits 0-15, 16-31)
    const int* pat;   // per pair: atoms of its Qshells (bits 0-15, 16-31)
    const double* Y;
    const double* cutmat;
    int nshell;
    double tau;
    int grad;   // 1: the gradient's test (one-centre skip, Y Y > tau and Y Y DNmax > tau)
    // optional: work steps, each a bra pair and up to 64 surviving kets of one
    // segment ({bra, list offset, count}); segment s owns the ceil(len / 64)
    // slots from slot0[s] (the slots of its tiles), the unused ones get count 0
    Tile* steps;
    const int* slot0;
};

// FP32 copy of the kernels' primitive-pair data (same layout as DevPrims).
struct DevPrimsF {
    float* zeta;
    float* oo2z;
    float* Px;
    float* Py;
    float* Pz;
    float* K;
    float* cut;
};

// Boys table for the FP32 path: F_m(T_k), T_k = k / kBoysTabPerUnit, m < kBoysTabM
constexpr int kBoysTabPerUnit = 8;
constexpr int kBoysTabRows = 36 * kBoysTabPerUnit + 1;
constexpr int kBoysTabM = 16;
constexpr float kBoysTabTmax = 36.0f;
// the kernels keep the kBoysLdsCols orders they expand (F_LT .. F_LT+4 of every
// grid point) in LDS: kBoysTabRows x kBoysLdsCols floats (5.8 KB)
constexpr int kBoysLdsCols = 5;

// Nuclear-gradient build (four-centre part).  The per-class tables are one
// int32 blob (generated, grad_tables.inc) with this header, then per stage
// kGradStageFields ints: dir (0 A, 1 B, 2 C; bits 0-7) with the most terms of
// any bra-table row (bits 8-15) and ket-table row (bits 16-23), source set
// (0 S0, 1 SA, 2 SB, 3 SC), bra rows, source columns, ket rows, and the
// offsets of the bra HRR, ket HRR and contraction tables.
enum GradHdr {
    kGhNodes = 0, kGhLevels, kGhSlots, kGhLT, kGhN0, kGhNA, kGhNB, kGhNC, kGhF0, kGhFA, kGhFB, kGhFC,
    kGhStages, kGhNOut, kGhCA, kGhCB, kGhCC, kGhCD, kGhOffNodes, kGhOffLev, kGhOffTgt, kGhMaxStage,
    kGhSize = 24
};
constexpr int kGradStageFields = 8;
constexpr int kGradHrrTerms = 8;

struct GradArgs {
    DevQshells q;
    DevPairs pr;
    DevPrims pp;
    const double* ea;   // exponent of the first (A or C) primitive of each primitive pair
    const double* eb;   // exponent of the second (B or D) primitive
    const double* xyz;
    const double* Dp;   // cons-scaled density
    const double* cutmat;
    unsigned long long* G;   // 3 x natom, int64 fixed point (2^-kGradShift)
    unsigned long long* counters;
    int n, nshell;
    double cutoff;      // gradient cutoff: Schwarz x density and primitive screens
    double hybrid;
    int natom;
    int lds_grad;       // light-class kernels: accumulate per block in LDS (3 x natom) first
    // FP32 copies for the mixed-precision kernels (set when precision is mixed)
    DevPrimsF ppf;
    const float* eaf;
    const float* ebf;
    const float* Dpf;
    const float* boys_tab;   // kBoysTabRows x kBoysTabM
};

// One-electron integrals (qk_oneint.hip): per basis function its data, and the nuclei.
constexpr int k1eMaxPrim = 10;   // QUICK's MAXPRIM
struct OneFuncs {
    const int* atom;     // 0-based
    const int* lmn;      // l | m << 4 | n << 8
    const int* np;       // primitives
    const int* qf;       // Qfinal of the function's shell (Boys order)
    const double* a;     // k1eMaxPrim exponents per function
    const double* c;     // k1eMaxPrim coefficients per function: gccoeff(k, ksumtype + L) cons(f)
};
struct OneSys {
    const double* xyz;   // 3 x natom
    const double* chg;   // nuclear charges
    int natom;
    int n;
};

struct FockArgs {
    DevQshells q;
    DevPairs pr;
    DevPrims pp;    // primitive-major per group (see DevPrims)
    DevPrimsF ppf;  // FP32 copy for the mixed-precision kernels
    const float* Dpf;          // FP32 copy of the cons-scaled density
    const float* boys_tab;     // kBoysTabRows x kBoysTabM
    const double* xyz;        // 3 x natom
    const double* Dp;         // cons-scaled density, n x n
    const double* cutmat;     // nshell x nshell, max |D| over shell blocks
    unsigned long long* X;    // fixed-point accumulator: xcopies n x n copies
    unsigned long long* counters;  // [0] surviving quartets, [1] primitive quartets
    // Coulomb contributions per significant pair: element e of pair k's block
    // (e over its NA x NB functions) at J[jofs[k] + e * pstride[k]], so that the
    // lanes of a wavefront, which hold consecutive ket pairs of one group, add
    // to consecutive addresses; jpair_scatter adds J into X after the classes
    unsigned long long* J;
    const int* jofs;
    int xcopies;              // copies of the accumulator (1 .. kXCopies)
    int n;
    int nshell;
    double tau;
    double prim_limit;
    double hybrid;
    int boys_qshell_order;    // 1: Boys order = class L total (reference GPU)
    // FP64 run mode: the density rows of the run's bra functions ((NA + NB) x n)
    // staged in LDS after the bra slab, so the exchange contractions read them
    // there instead of gathering them from memory
    int stage_d;
};

}  // namespace qk
