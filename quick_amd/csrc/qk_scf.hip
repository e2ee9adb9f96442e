// qk_scf.hip -- the restricted Hartree-Fock SCF step beside the Fock build,
// with the density, the Fock matrix and the DIIS history resident in HBM.
//
// Replaces QUICK's electdiis loop (scf.f90:56-645): per cycle the operator
// F = H + G(D) from the library's own Fock build, the energy
// E = 1/2 sum D (H + F) + E_nn (getEnergy.f90:43-63), Pulay's DIIS on the
// commutator error F D S - S D F (scf.f90:249-523), the orthogonalised
// operator X F X diagonalised in FP64 (rocSOLVER dsyevd), C = X C' and
// D = 2 C_occ C_occ^T (scf.f90:528-537), until the RMS change of D falls below
// the convergence threshold (QUICK's DENSERMS).  X = S^{-1/2} (symmetric
// orthogonalisation); the guess is the core Hamiltonian.  The matrix products
// are rocBLAS DGEMMs; only the DIIS system (at most 11 x 11) is solved on the
// host.  All of it stays FP64; the Fock build takes the caller's options
// (FP64 or mixed precision).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/quick_mi355x.h"
#include "qk_internal.h"

// device index and number of basis functions of a context (qk_host.cpp)
int qk_context_info(qk_ctx* c, int* device, int* nbasis);

namespace {

void scf_err(const char* fmt, const char* a = "", int b = 0) {
    char buf[512];
    snprintf(buf, sizeof(buf), fmt, a, b);
    qk_set_error_text(buf);
}

struct Dev {
    double* p = nullptr;
    int alloc(size_t n) {
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) {
            p = nullptr;
            scf_err("qk_scf: device allocation of %s failed", "a matrix");
            return -1;
        }
        return 0;
    }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
};

// out = a - a^T (n x n)
__global__ void antisym(const double* __restrict__ a, double* __restrict__ out, int n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)n * n) return;
    const int i = int(k % n), j = int(k / n);
    out[k] = a[k] - a[(long long)i * n + j];
}

// column j of u scaled by w_j^{-1/4}
__global__ void scale_cols(double* __restrict__ u, const double* __restrict__ w, int n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)n * n) return;
    u[k] *= 1.0 / std::sqrt(std::sqrt(w[k / n]));
}

// Gaussian elimination with partial pivoting on the (m x m) system A x = b (host)
bool solve_small(std::vector<double> A, std::vector<double> b, int m, std::vector<double>& x) {
    for (int c = 0; c < m; ++c) {
        int piv = c;
        for (int r = c + 1; r < m; ++r)
            if (std::fabs(A[r * m + c]) > std::fabs(A[piv * m + c])) piv = r;
        if (std::fabs(A[piv * m + c]) < 1e-300) return false;
        if (piv != c) {
            for (int k = 0; k < m; ++k) std::swap(A[c * m + k], A[piv * m + k]);
            std::swap(b[c], b[piv]);
        }
        for (int r = c + 1; r < m; ++r) {
            const double f = A[r * m + c] / A[c * m + c];
            for (int k = c; k < m; ++k) A[r * m + k] -= f * A[c * m + k];
            b[r] -= f * b[c];
        }
    }
    x.assign(m, 0.0);
    for (int r = m - 1; r >= 0; --r) {
        double s = b[r];
        for (int k = r + 1; k < m; ++k) s -= A[r * m + k] * x[k];
        x[r] = s / A[r * m + r];
    }
    return true;
}

#define RB(call)                                                         \
    do {                                                                 \
        if ((call) != rocblas_status_success) {                          \
            scf_err("qk_scf: %s failed", #call);                         \
            return -1;                                                   \
        }                                                                \
    } while (0)
#define HC(call)                                                         \
    do {                                                                 \
        if ((call) != hipSuccess) {                                      \
            scf_err("qk_scf: %s failed", #call);                         \
            return -1;                                                   \
        }                                                                \
    } while (0)

struct Handle {
    rocblas_handle h = nullptr;
    ~Handle() {
        if (h) rocblas_destroy_handle(h);
    }
};

}  // namespace

extern "C" void qk_scf_default_opts(qk_scf_opts* o) {
    o->max_cycles = 100;
    o->max_diis = 10;
    o->dens_rms = 1e-7;
    qk_default_opts(&o->fock);
}

extern "C" int qk_scf(qk_ctx* c, int nocc, const qk_scf_opts* o, double* D_out, double* E_out, int* cycles_out,
                      double* rms_out) {
    qk_scf_opts d;
    if (!o) {
        qk_scf_default_opts(&d);
        o = &d;
    }
    if (!c) { scf_err("qk_scf: null context"); return -1; }
    int dev = 0, n = 0;
    if (qk_context_info(c, &dev, &n)) return -1;
    if (nocc < 1 || nocc > n) { scf_err("qk_scf: %s %d occupied orbitals", "invalid:", nocc); return -2; }
    if (o->max_diis < 1 || o->max_diis > 10) { scf_err("qk_scf: max_diis %s%d outside 1..10", "", o->max_diis); return -2; }
    (void)hipSetDevice(dev);
    const size_t nn = (size_t)n * n;
    const int m = o->max_diis;
    Dev S, H, X, F, G, D, Dn, T1, T2, W, Fh, Eh, w, e;
    if (S.alloc(nn) || H.alloc(nn) || X.alloc(nn) || F.alloc(nn) || G.alloc(nn) || D.alloc(nn) || Dn.alloc(nn) ||
        T1.alloc(nn) || T2.alloc(nn) || W.alloc(nn) || Fh.alloc(nn * m) || Eh.alloc(nn * m) || w.alloc(n) ||
        e.alloc(n))
        return -1;
    rocblas_int* info = nullptr;
    HC(hipMalloc(&info, sizeof(rocblas_int)));
    struct InfoFree { rocblas_int* p; ~InfoFree() { (void)hipFree(p); } } info_free{info};
    Handle hb;
    RB(rocblas_create_handle(&hb.h));
    RB(rocblas_set_pointer_mode(hb.h, rocblas_pointer_mode_host));
    const rocblas_handle h = hb.h;
    const double one = 1.0, zero = 0.0, two = 2.0, mone = -1.0;
    auto gemm = [&](rocblas_operation ta, rocblas_operation tb, int k, const double* A, const double* B, double* Cm,
                    double alpha) {
        return rocblas_dgemm(h, ta, tb, n, n, k, &alpha, A, n, B, n, &zero, Cm, n);
    };
    const int blocks = (int)((nn + 255) / 256);
    // one-electron operator, overlap, nuclear repulsion
    if (qk_oneint_device(c, S.p, H.p, nullptr)) return -1;
    double enn = 0.0;
    if (qk_nuclear_repulsion(c, &enn)) return -1;
    // X = S^{-1/2} = U w^{-1/2} U^T
    HC(hipMemcpy(T1.p, S.p, nn * sizeof(double), hipMemcpyDeviceToDevice));
    RB(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, n, T1.p, n, w.p, e.p, info));
    rocblas_int hinfo = 0;
    HC(hipMemcpy(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost));
    if (hinfo != 0) { scf_err("qk_scf: %s (info %d)", "overlap diagonalisation failed", hinfo); return -3; }
    hipLaunchKernelGGL(scale_cols, dim3(blocks), dim3(256), 0, 0, T1.p, w.p, n);
    RB(gemm(rocblas_operation_none, rocblas_operation_transpose, n, T1.p, T1.p, X.p, 1.0));
    // core-Hamiltonian guess: D = 0, F = H
    HC(hipMemset(D.p, 0, nn * sizeof(double)));
    std::vector<double> B;     // DIIS error inner products, m x m, ring slots
    std::vector<int> slots;    // ring slot of each stored cycle, oldest first
    int next = 0, cyc = 0;
    double E = 0.0, rms = 1.0;
    for (cyc = 1; cyc <= o->max_cycles; ++cyc) {
        // F = H + G(D)
        HC(hipMemcpy(F.p, H.p, nn * sizeof(double), hipMemcpyDeviceToDevice));
        if (cyc > 1) {
            if (qk_fock_device(c, D.p, G.p, &o->fock, nullptr)) return -1;
            RB(rocblas_daxpy(h, (rocblas_int)nn, &one, G.p, 1, F.p, 1));
        }
        // E = 1/2 sum D (H + F) + E_nn
        double dh = 0.0, df = 0.0;
        RB(rocblas_ddot(h, (rocblas_int)nn, D.p, 1, H.p, 1, &dh));
        RB(rocblas_ddot(h, (rocblas_int)nn, D.p, 1, F.p, 1, &df));
        E = 0.5 * (dh + df) + enn;
        // DIIS: error F D S - S D F, stored with F in the ring (not for the
        // first cycle: its density is the empty guess, its error zero, and DIIS
        // would return to the core Hamiltonian)
        if (cyc > 1) {
        const int sl = next;
        next = (next + 1) % m;
        if ((int)slots.size() == m) slots.erase(slots.begin());
        slots.push_back(sl);
        double* Fs = Fh.p + nn * sl;
        double* Es = Eh.p + nn * sl;
        HC(hipMemcpy(Fs, F.p, nn * sizeof(double), hipMemcpyDeviceToDevice));
        RB(gemm(rocblas_operation_none, rocblas_operation_none, n, F.p, D.p, T1.p, 1.0));
        RB(gemm(rocblas_operation_none, rocblas_operation_none, n, T1.p, S.p, T2.p, 1.0));
        hipLaunchKernelGGL(antisym, dim3(blocks), dim3(256), 0, 0, T2.p, Es, n);
        HC(hipGetLastError());
        B.resize((size_t)m * m, 0.0);
        for (int s2 : slots) {
            double v = 0.0;
            RB(rocblas_ddot(h, (rocblas_int)nn, Es, 1, Eh.p + nn * s2, 1, &v));
            B[(size_t)sl * m + s2] = B[(size_t)s2 * m + sl] = v;
        }
        // extrapolated F = sum c_i F_i with sum c_i = 1 minimising |sum c_i e_i|
        const int k = (int)slots.size();
        std::vector<double> coef(1, 1.0);
        if (k > 1) {
            std::vector<double> A((size_t)(k + 1) * (k + 1), 0.0), rhs(k + 1, 0.0), x;
            for (int i = 0; i < k; ++i) {
                for (int j = 0; j < k; ++j) A[(size_t)i * (k + 1) + j] = B[(size_t)slots[i] * m + slots[j]];
                A[(size_t)i * (k + 1) + k] = A[(size_t)k * (k + 1) + i] = -1.0;
            }
            rhs[k] = -1.0;
            if (solve_small(A, rhs, k + 1, x)) coef.assign(x.begin(), x.begin() + k);
            else coef.assign(k, 0.0), coef[k - 1] = 1.0;
            HC(hipMemset(F.p, 0, nn * sizeof(double)));
            for (int i = 0; i < k; ++i)
                RB(rocblas_daxpy(h, (rocblas_int)nn, &coef[i], Fh.p + nn * slots[i], 1, F.p, 1));
        }
        }
        // X F X, its eigenvectors, C = X C', D = 2 C_occ C_occ^T
        RB(gemm(rocblas_operation_none, rocblas_operation_none, n, F.p, X.p, T1.p, 1.0));
        RB(gemm(rocblas_operation_none, rocblas_operation_none, n, X.p, T1.p, T2.p, 1.0));
        RB(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, n, T2.p, n, w.p, e.p, info));
        HC(hipMemcpy(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost));
        if (hinfo != 0) { scf_err("qk_scf: %s (info %d)", "Fock diagonalisation failed", hinfo); return -3; }
        RB(gemm(rocblas_operation_none, rocblas_operation_none, n, X.p, T2.p, W.p, 1.0));
        RB(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, n, n, nocc, &two, W.p, n, W.p, n,
                         &zero, Dn.p, n));
        // RMS change of the density
        RB(rocblas_daxpy(h, (rocblas_int)nn, &mone, Dn.p, 1, D.p, 1));
        double nrm = 0.0;
        RB(rocblas_dnrm2(h, (rocblas_int)nn, D.p, 1, &nrm));
        rms = nrm / n;
        std::swap(D.p, Dn.p);
        if (rms < o->dens_rms && cyc > 1) break;
    }
    // the energy of the final density
    HC(hipMemcpy(F.p, H.p, nn * sizeof(double), hipMemcpyDeviceToDevice));
    if (qk_fock_device(c, D.p, G.p, &o->fock, nullptr)) return -1;
    RB(rocblas_daxpy(h, (rocblas_int)nn, &one, G.p, 1, F.p, 1));
    double dh = 0.0, df = 0.0;
    RB(rocblas_ddot(h, (rocblas_int)nn, D.p, 1, H.p, 1, &dh));
    RB(rocblas_ddot(h, (rocblas_int)nn, D.p, 1, F.p, 1, &df));
    E = 0.5 * (dh + df) + enn;
    if (D_out) HC(hipMemcpy(D_out, D.p, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (E_out) *E_out = E;
    if (cycles_out) *cycles_out = std::min(cyc, o->max_cycles);
    if (rms_out) *rms_out = rms;
    if (rms >= o->dens_rms) {
        scf_err("qk_scf: %s after the cycle limit (%d)", "no convergence", o->max_cycles);
        return -4;
    }
    return 0;
}
