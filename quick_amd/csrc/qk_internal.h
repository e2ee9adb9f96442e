// Internal types shared by the host code (qk_host.cpp) and the kernels
// (qk_fock.hip).  Device arrays are structure-of-arrays so that a wavefront
// whose lanes walk consecutive kets of one bra pair reads the bra data as a
// broadcast and the ket data as coalesced loads.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace qk {

constexpr int kNumPairClasses = 6;     // ss ps pp ds dp dd (la >= lb)
constexpr int kNumQuartetClasses = 21;  // bra pair class >= ket pair class
constexpr int kFixedShift = 48;         // Fock accumulator: int64 in units of 2^-48
constexpr int kGradShift = 44;          // gradient accumulator: int64 in units of 2^-44

// Per Qshell (one angular momentum of one shell).
struct DevQshells {
    int* atom;    // 0-based atom
    int* f0;      // first basis function (0-based)
    int* shell;   // parent shell (0-based)
    int* qfinal;  // Qfinal of the parent shell (QUICK's Boys-order convention)
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
// group for every bra; the host cuts those prefixes into tiles.
struct Tile {
    int p;
    int q0;
    int len;
    int pad;
};

struct ClassWork {
    const Tile* tiles;
    long long ntiles;
};

struct FockArgs {
    DevQshells q;
    DevPairs pr;
    DevPrims pp;    // primitive-major per group (see DevPrims)
    const double* xyz;        // 3 x natom
    const double* Dp;         // cons-scaled density, n x n
    const double* cutmat;     // nshell x nshell, max |D| over shell blocks
    unsigned long long* X;    // n x n fixed-point accumulator
    unsigned long long* counters;  // [0] surviving quartets, [1] primitive quartets
    int n;
    int nshell;
    double tau;
    double prim_limit;
    double hybrid;
    int boys_qshell_order;    // 1: Boys order = class L total (reference GPU)
};

}  // namespace qk
