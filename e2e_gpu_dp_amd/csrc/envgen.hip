// envgen.hip — input generation: the double-gyre environment of
// src/data_input/build_environment_DG_with_westward_clouds.py (realizations of the
// velocity field for a range of forcing frequencies, per-slice mean, the leading modes
// and coefficients of the centred realization matrix, the westward-moving cloud field
// and the obstacle mask), written in the layouts the builder reads
// (src/build_model.cu:380-398).
//
// The reference takes a full np.linalg.svd of each slice's n_wsamples x 2n matrix
// (:227-236) and keeps n_modes of it. Here only those modes are computed, from the
// Gram matrix of the smaller side: G = X X^T (or X^T X) by rocBLAS dgemm (a plain
// library GEMM), then a batched subspace iteration on G (GEMMs on the device; the
// l x l Cholesky and Rayleigh-Ritz steps of l = n_modes + 8 columns on the host),
// then C = U diag(s), M = Vh (:199-207). Realizations, means, centring, sign choice
// and the float32 field layouts are HIP kernels. Refinement to a finer grid
// (interpolate_degree > 1, scipy griddata) is not done: the generator works at the
// target resolution (degree 1, where griddata returns the samples themselves).
//
// Singular vectors carry an arbitrary sign in LAPACK as here; each mode is signed so
// that its entry of largest magnitude (the first such) is positive, and its coefficient
// column with it, which leaves every C M product unchanged.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

namespace e2e {

namespace {

struct RealArgs {
    double* R;     // [nb][nw][2n] velocity realizations (vx then vy), slices t0..t0+nb-1
    double* P;     // [nb][nw][n] phi realizations or NULL
    int t0, nb, g, nw;
    double dt, dxy, A, eps, wx, wy, w_lo, w_hi;
};

// generate_R (:149-171): vx_vy (:137-142), phi (:134-135), sample_w (:144-150),
// get_ij_from_xy (:152-155), xs = (0.5 + i) dxy (:131-132), t = tid * dt
__global__ void realizations(const RealArgs A) {
    const int64_t n = (int64_t)A.g * A.g;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)A.nb * A.nw * n) return;
    const int64_t col = q % n;
    const int k = (int)((q / n) % A.nw);
    const int b = (int)(q / (n * A.nw));
    const int i = (int)(col / A.g), j = (int)(col % A.g);
    const double t = (double)(A.t0 + b) * A.dt;
    const double r = (double)k / (double)(A.nw - 1);
    const double w = A.w_lo + r * (A.w_hi - A.w_lo);
    const double x = (0.5 + j) * A.dxy;
    const double y = (0.5 + (A.g - 1 - i)) * A.dxy;
    const double swt = sin(w * t);
    const double f = A.eps * swt * (x * x) + (1 - 2 * A.eps * swt) * x;
    const double fx = (2 * A.eps * swt * (x - 1)) + 1;
    const double vx = -(0.5 * A.A * M_PI * sin(A.wx * f) * cos(A.wy * y));
    const double vy = A.A * M_PI * sin(A.wy * y) * fx * cos(A.wx * f);
    const int64_t row = ((int64_t)b * A.nw + k) * 2 * n;
    A.R[row + col] = vx;
    A.R[row + n + col] = vy;
    if (A.P) A.P[((int64_t)b * A.nw + k) * n + col] = A.A * sin(M_PI * f) * sin(A.wy * y);
}

// mean over realizations (np.mean axis 1: added in realization order, then divided),
// then the realizations centred in place (:215-225)
__global__ void mean_center(double* X, double* mean, int nb, int nw, int64_t m) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nb * m) return;
    const int b = (int)(q / m);
    const int64_t c = q - (int64_t)b * m;
    double* x = X + (int64_t)b * nw * m + c;
    double s = 0.0;
    for (int k = 0; k < nw; k++) s += x[(int64_t)k * m];
    const double mu = s / nw;
    mean[q] = mu;
    for (int k = 0; k < nw; k++) x[(int64_t)k * m] -= mu;
}

// a column-major r x l start block per slice from a fixed hash (deterministic)
__global__ void start_block(double* Q, int64_t per, int nb) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= per * nb) return;
    uint64_t h = (uint64_t)(q % per) * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    Q[q] = (double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

// per slice and mode: s_c, the sign of the mode's first largest-magnitude entry, and the
// float32 outputs. Mc: column-major m x k (mode c = column c, entries in the reference's
// [vx | vy] or [phi] column order); Cc: column-major nw x k, already scaled by s.
__global__ void emit_modes(const double* Mc, const double* Cc, int nb, int64_t m, int nw, int k, int n_out,
                           int64_t slice_stride_mode, float* mode_a, float* mode_b, float* coef, int t0,
                           int nt_total) {
    // one block per (slice, mode): find the sign, then write
    const int b = blockIdx.x / k, c = blockIdx.x % k;
    const double* mv = Mc + ((int64_t)b * k + c) * m;
    const double* cv = Cc + ((int64_t)b * k + c) * nw;
    __shared__ double bv[256];
    __shared__ int64_t bi[256];
    double best = -1.0;
    int64_t besti = 0;
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) {
        const double a = fabs(mv[i]);
        if (a > best) {
            best = a;
            besti = i;
        }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = besti;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const double a = bv[threadIdx.x + o];
            const int64_t ai = bi[threadIdx.x + o];
            if (a > bv[threadIdx.x] || (a == bv[threadIdx.x] && ai < bi[threadIdx.x])) {
                bv[threadIdx.x] = a;
                bi[threadIdx.x] = ai;
            }
        }
        __syncthreads();
    }
    const double sg = mv[bi[0]] < 0 ? -1.0 : 1.0;
    const int t = t0 + b;
    // modes: mode_a[t][c][cell] = M[c][cell], mode_b (vy) = M[c][n_out + cell]
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) {
        const float v = (float)(sg * mv[i]);
        if (i < n_out)
            mode_a[(int64_t)t * slice_stride_mode + (int64_t)c * n_out + i] = v;
        else
            mode_b[(int64_t)t * slice_stride_mode + (int64_t)c * n_out + (i - n_out)] = v;
    }
    // coefficients: coef[t][rzn][c]
    for (int r = threadIdx.x; r < nw; r += blockDim.x)
        coef[((int64_t)t * nw + r) * k + c] = (float)(sg * cv[r]);
    (void)nt_total;
}

// means to the float32 fields: out_a[t][cell] = mean[b][cell], out_b = mean[b][n + cell]
__global__ void emit_mean(const double* mean, int nb, int64_t m, int64_t n, float* out_a, float* out_b, int t0) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)nb * m) return;
    const int b = (int)(q / m);
    const int64_t c = q - (int64_t)b * m;
    const float v = (float)mean[q];
    if (c < n)
        out_a[(int64_t)(t0 + b) * n + c] = v;
    else
        out_b[(int64_t)(t0 + b) * n + (c - n)] = v;
}

// DG_scalar_field.generate_phi (:59-99) at refinement degree 1: t = tid, T = t/2,
// l = xs[-1]/2, dom_size = xs[-1] + dxy; value A_sc * f(x, t), float32
__global__ void clouds(float* s, int nt, int g, double dxy, double A_sc) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)g * g;
    if (q >= (int64_t)nt * n) return;
    const int t = (int)(q / n);
    const int j = (int)(q % g);
    const double x = (0.5 + j) * dxy;
    const double xl = (0.5 + (g - 1)) * dxy;
    const double T = t * 1 / 2.0;
    const double l = xl / 2;
    const double dom = xl + dxy;
    const double xt = x + (T * dxy);
    double f;
    if (xt <= l) {
        f = 1;
    } else if (l < xt && xt < l + (0.2 * dom)) {
        const double slope = (1.0 - 0.0) / (-0.2 * dom);
        f = 1 + (slope * (x - (l - T * dxy)));
    } else {
        f = 0;
    }
    s[q] = (float)(A_sc * f);
}

// fill_obstacles (:470-505): 0 none, 1 static box rows [0.4g, 0.5g) x cols [0.5g, 0.6g),
// 2 a box of obs_width moving obs_speed cells per step, present while inside the interior
__global__ void obstacles(int32_t* mask, int nt, int g, int mode, int ru_i, int cl_i, int width, double vxs,
                          double vys) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)g * g;
    if (q >= (int64_t)nt * n) return;
    const int t = (int)(q / n);
    const int i = (int)((q % n) / g), j = (int)(q % g);
    int v = 0;
    if (mode == 1) {
        const int ru = (int)(0.4 * g), rd = (int)(0.5 * g), cl = (int)(0.5 * g), cr = (int)(0.6 * g);
        v = i >= ru && i < rd && j >= cl && j < cr;
    } else if (mode == 2) {
        const int cl = (int)floor(cl_i + (t * vxs));
        const int cr = cl + width;
        const int ru = (int)floor(ru_i + (t * vys));
        const int rd = ru + width;
        const bool in = (0 < cl && cl < g - 1) && (0 < cr && cr < g - 1) && (0 < ru && ru < g - 1) &&
                        (0 < rd && rd < g - 1);
        v = in && i >= ru && i < rd && j >= cl && j < cr;
    }
    mask[q] = v;
}

// --------------------------------------------------------------- host linear algebra
// Cholesky S = R^T R of an l x l column-major SPD block; returns R^-1 (upper), with the
// columns of (numerically) dependent directions zeroed
void chol_inv(const double* S, int l, double* Rinv) {
    std::vector<double> R(l * l, 0.0);
    double dmax = 0.0;
    for (int i = 0; i < l; i++) dmax = std::max(dmax, S[i * l + i]);
    std::vector<char> dead(l, 0);
    for (int j = 0; j < l; j++) {
        double d = S[j * l + j];
        for (int k = 0; k < j; k++) d -= R[j * l + k] * R[j * l + k];
        if (!(d > 1e-26 * dmax) || dead[j]) {
            dead[j] = 1;
            continue;
        }
        const double rjj = std::sqrt(d);
        R[j * l + j] = rjj;
        for (int i = j + 1; i < l; i++) {
            double v = S[i * l + j];
            for (int k = 0; k < j; k++) v -= R[i * l + k] * R[j * l + k];
            R[i * l + j] = v / rjj;  // R^T stored row i: lower factor L[i][j]
        }
    }
    // L = R^T lower (L[i][j] at R[i*l + j]); R^-1 = (L^-1)^T, column-major Rinv
    std::vector<double> Linv(l * l, 0.0);
    for (int j = 0; j < l; j++) {
        if (dead[j]) continue;
        Linv[j * l + j] = 1.0 / R[j * l + j];
        for (int i = j + 1; i < l; i++) {
            if (dead[i]) continue;
            double v = 0.0;
            for (int k = j; k < i; k++) v -= R[i * l + k] * Linv[k * l + j];
            Linv[i * l + j] = v / R[i * l + i];
        }
    }
    // Rinv (column-major) [row r][col c] = Linv[c][r]
    for (int c = 0; c < l; c++)
        for (int r = 0; r < l; r++) Rinv[c * l + r] = Linv[c * l + r];
}

// cyclic Jacobi eigen-decomposition of an l x l symmetric column-major T: eigenvalues
// descending in w, eigenvectors in the columns of V (column-major)
void jacobi_eig(const double* T, int l, double* w, double* V) {
    std::vector<double> a(T, T + l * l);
    std::vector<double> v(l * l, 0.0);
    for (int i = 0; i < l; i++) v[i * l + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < l; p++)
            for (int q = 0; q < l; q++) (p == q ? diag : off) += a[q * l + p] * a[q * l + p];
        if (off <= 1e-30 * diag) break;
        for (int p = 0; p < l; p++)
            for (int q = p + 1; q < l; q++) {
                const double apq = a[q * l + p];
                if (apq == 0.0) continue;
                const double app = a[p * l + p], aqq = a[q * l + q];
                const double th = (aqq - app) / (2 * apq);
                const double tt = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
                const double c = 1 / std::sqrt(tt * tt + 1), s = tt * c;
                for (int k = 0; k < l; k++) {  // columns p, q
                    const double akp = a[p * l + k], akq = a[q * l + k];
                    a[p * l + k] = c * akp - s * akq;
                    a[q * l + k] = s * akp + c * akq;
                }
                for (int k = 0; k < l; k++) {  // rows p, q
                    const double apk = a[k * l + p], aqk = a[k * l + q];
                    a[k * l + p] = c * apk - s * aqk;
                    a[k * l + q] = s * apk + c * aqk;
                }
                for (int k = 0; k < l; k++) {
                    const double vkp = v[p * l + k], vkq = v[q * l + k];
                    v[p * l + k] = c * vkp - s * vkq;
                    v[q * l + k] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(l);
    for (int i = 0; i < l; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a[x * l + x] > a[y * l + y]; });
    for (int c = 0; c < l; c++) {
        w[c] = a[idx[c] * l + idx[c]];
        for (int r = 0; r < l; r++) V[c * l + r] = v[idx[c] * l + r];
    }
}

struct Blas {
    rocblas_handle h = nullptr;
    ~Blas() {
        if (h) (void)rocblas_destroy_handle(h);
    }
};

#define E2E_BLAS(call)                                                               \
    do {                                                                             \
        rocblas_status _s = (call);                                                  \
        if (_s != rocblas_status_success) {                                          \
            set_error("%s failed: %s", #call, rocblas_status_to_string(_s));         \
            return E2E_ERR_HIP;                                                      \
        }                                                                            \
    } while (0)

// Leading k singular triplets of nb row-major nw x m matrices X (stride nw*m), as
// Mc (column-major m x k, the modes, = Vh[:k] row-major) and Cc (column-major nw x k,
// U[:, :k] diag(s)), singular values in sv[nb][k]. Returns the iterations used.
e2e_status top_modes(rocblas_handle h, hipStream_t s, const double* X, int nb, int nw, int64_t m, int k,
                     double* Mc, double* Cc, double* sv, int* iters_out) {
    const bool left = (int64_t)nw <= m;  // Gram of the smaller side
    const int r = left ? nw : (int)m;
    const int l = std::min(k + 8, r);
    const double one = 1.0, zero = 0.0;
    DevBuf dG, dQ, dY, dS, dRi;
    E2E_TRY(dmalloc(dG, sizeof(double) * (size_t)r * r * nb));
    E2E_TRY(dmalloc(dQ, sizeof(double) * (size_t)r * l * nb));
    E2E_TRY(dmalloc(dY, sizeof(double) * (size_t)r * l * nb));
    E2E_TRY(dmalloc(dS, sizeof(double) * (size_t)l * l * nb));
    E2E_TRY(dmalloc(dRi, sizeof(double) * (size_t)l * l * nb));
    double *G = dG.as<double>(), *Q = dQ.as<double>(), *Y = dY.as<double>(), *S = dS.as<double>(),
           *Ri = dRi.as<double>();
    const rocblas_stride sX = (rocblas_stride)nw * m, sG = (rocblas_stride)r * r, sQ = (rocblas_stride)r * l,
                         sS = (rocblas_stride)l * l;
    // X row-major nw x m is column-major m x nw (ld m): X X^T = A^T A, X^T X = A A^T
    if (left)
        E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, r, r, (int)m,
                                               &one, X, (int)m, sX, X, (int)m, sX, &zero, G, r, sG, nb));
    else
        E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_transpose, r, r, nw,
                                               &one, X, (int)m, sX, X, (int)m, sX, &zero, G, r, sG, nb));
    hipLaunchKernelGGL(start_block, dim3((unsigned)ceil_div((int64_t)r * l * nb, 256)), dim3(256), 0, s, Q,
                       (int64_t)r * l, nb);
    E2E_HIP(hipGetLastError());
    std::vector<double> hS((size_t)l * l * nb), hRi((size_t)l * l * nb, 0.0);
    // Q <- orthonormal basis of Y (CholeskyQR), applied twice for orthogonality
    auto orth = [&](double* Yin, double* Qout) -> e2e_status {
        for (int pass = 0; pass < 2; pass++) {
            double* src = pass == 0 ? Yin : Qout;
            E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, l, l, r,
                                                   &one, src, r, sQ, src, r, sQ, &zero, S, l, sS, nb));
            E2E_HIP(hipMemcpyAsync(hS.data(), S, sizeof(double) * hS.size(), hipMemcpyDeviceToHost, s));
            E2E_HIP(hipStreamSynchronize(s));
            for (int b = 0; b < nb; b++) chol_inv(hS.data() + (size_t)b * l * l, l, hRi.data() + (size_t)b * l * l);
            E2E_HIP(hipMemcpyAsync(Ri, hRi.data(), sizeof(double) * hRi.size(), hipMemcpyHostToDevice, s));
            // Qout = src Ri (src may be Qout: go through Y)
            double* dst = pass == 0 ? Qout : Yin;
            E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, r, l, l, &one,
                                                   src, r, sQ, Ri, l, sS, &zero, dst, r, sQ, nb));
            if (pass == 1)
                E2E_HIP(hipMemcpyAsync(Qout, Yin, sizeof(double) * (size_t)r * l * nb, hipMemcpyDeviceToDevice, s));
        }
        return E2E_OK;
    };
    E2E_HIP(hipMemcpyAsync(Y, Q, sizeof(double) * (size_t)r * l * nb, hipMemcpyDeviceToDevice, s));
    E2E_TRY(orth(Y, Q));
    std::vector<double> prev((size_t)k * nb, -1.0), w(l), V((size_t)l * l);
    int it = 0, stop = 0;
    const int kMaxIt = 400;
    for (it = 1; it <= kMaxIt; it++) {
        // Y = G Q; every 4th iteration the Ritz values (Q^T G Q) are compared with the
        // previous ones. Ritz values converge at twice the rate of the vectors, so once
        // they have settled the iteration runs as many steps again.
        E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, r, l, r, &one, G, r,
                                               sG, Q, r, sQ, &zero, Y, r, sQ, nb));
        if (stop && it >= stop) break;
        if (!stop && it % 4 == 0) {
            E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, l, l, r,
                                                   &one, Q, r, sQ, Y, r, sQ, &zero, S, l, sS, nb));
            E2E_HIP(hipMemcpyAsync(hS.data(), S, sizeof(double) * hS.size(), hipMemcpyDeviceToHost, s));
            E2E_HIP(hipStreamSynchronize(s));
            bool done = true;
            for (int b = 0; b < nb; b++) {
                jacobi_eig(hS.data() + (size_t)b * l * l, l, w.data(), V.data());
                for (int c = 0; c < k; c++) {
                    const double d = std::fabs(w[c] - prev[(size_t)b * k + c]);
                    if (!(d <= 1e-13 * std::fabs(w[c]) + 1e-15 * std::fabs(w[0]))) done = false;
                    prev[(size_t)b * k + c] = w[c];
                }
            }
            if (done) stop = 2 * it;
        }
        E2E_TRY(orth(Y, Q));
    }
    // Rayleigh-Ritz on the final basis: T = Q^T G Q, U = Q W
    E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, r, l, r, &one, G, r, sG,
                                           Q, r, sQ, &zero, Y, r, sQ, nb));
    E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, l, l, r, &one, Q, r,
                                           sQ, Y, r, sQ, &zero, S, l, sS, nb));
    E2E_HIP(hipMemcpyAsync(hS.data(), S, sizeof(double) * hS.size(), hipMemcpyDeviceToHost, s));
    E2E_HIP(hipStreamSynchronize(s));
    std::vector<double> hW((size_t)l * k * nb, 0.0), hScale((size_t)k * nb);
    for (int b = 0; b < nb; b++) {
        jacobi_eig(hS.data() + (size_t)b * l * l, l, w.data(), V.data());
        for (int c = 0; c < k; c++) {
            const double lam = std::max(w[c], 0.0);
            sv[(size_t)b * k + c] = std::sqrt(lam);
            for (int q = 0; q < l; q++) hW[((size_t)b * k + c) * l + q] = V[(size_t)c * l + q];
        }
    }
    // U (r x k) = Q W into S-sized scratch: reuse Y's first r*k entries per slice
    DevBuf dW, dU;
    E2E_TRY(dmalloc(dW, sizeof(double) * hW.size()));
    E2E_TRY(dmalloc(dU, sizeof(double) * (size_t)r * k * nb));
    E2E_HIP(hipMemcpyAsync(dW.p, hW.data(), sizeof(double) * hW.size(), hipMemcpyHostToDevice, s));
    E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, r, k, l, &one, Q, r, sQ,
                                           dW.as<double>(), l, (rocblas_stride)l * k, &zero, dU.as<double>(), r,
                                           (rocblas_stride)r * k, nb));
    // scale factors: left: C = U diag(s), M^T = X^T U diag(1/s); right: M^T = V, C = X V
    std::vector<double> hD((size_t)k * k * nb, 0.0);
    DevBuf dD;
    E2E_TRY(dmalloc(dD, sizeof(double) * hD.size()));
    const rocblas_stride sK = (rocblas_stride)k * k;
    if (left) {
        // Cc = U diag(s)
        for (int b = 0; b < nb; b++)
            for (int c = 0; c < k; c++) hD[(size_t)b * k * k + c * k + c] = sv[(size_t)b * k + c];
        E2E_HIP(hipMemcpyAsync(dD.p, hD.data(), sizeof(double) * hD.size(), hipMemcpyHostToDevice, s));
        E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, nw, k, k, &one,
                                               dU.as<double>(), nw, (rocblas_stride)nw * k, dD.as<double>(), k, sK,
                                               &zero, Cc, nw, (rocblas_stride)nw * k, nb));
        // Mc = (X^T U) diag(1/s) (column-major m x k), X column-major m x nw
        E2E_HIP(hipStreamSynchronize(s));
        for (int b = 0; b < nb; b++)
            for (int c = 0; c < k; c++) {
                const double sc = sv[(size_t)b * k + c];
                hD[(size_t)b * k * k + c * k + c] = sc > 0 ? 1.0 / sc : 0.0;
            }
        DevBuf dT;
        E2E_TRY(dmalloc(dT, sizeof(double) * (size_t)m * k * nb));
        E2E_HIP(hipMemcpyAsync(dD.p, hD.data(), sizeof(double) * hD.size(), hipMemcpyHostToDevice, s));
        E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, (int)m, k, nw, &one,
                                               X, (int)m, sX, dU.as<double>(), nw, (rocblas_stride)nw * k, &zero,
                                               dT.as<double>(), (int)m, (rocblas_stride)m * k, nb));
        E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, (int)m, k, k, &one,
                                               dT.as<double>(), (int)m, (rocblas_stride)m * k, dD.as<double>(), k, sK,
                                               &zero, Mc, (int)m, (rocblas_stride)m * k, nb));
        E2E_HIP(hipStreamSynchronize(s));
    } else {
        // Mc = V (column-major m x k); Cc = X V (column-major nw x k) = (A^T) V, A = X col-major m x nw
        E2E_HIP(hipMemcpyAsync(Mc, dU.p, sizeof(double) * (size_t)m * k * nb, hipMemcpyDeviceToDevice, s));
        E2E_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, nw, k, (int)m,
                                               &one, X, (int)m, sX, dU.as<double>(), (int)m, (rocblas_stride)m * k,
                                               &zero, Cc, nw, (rocblas_stride)nw * k, nb));
        E2E_HIP(hipStreamSynchronize(s));
    }
    if (iters_out) *iters_out = std::min(it, kMaxIt);
    return E2E_OK;
}

}  // namespace

e2e_status env_generate(const e2e_env_params* ep, const e2e_env_out* out, e2e_env_stats* stats) {
    E2E_CHECK(ep && out, E2E_ERR_INVALID, "env_generate: params and outputs are required");
    const int g = ep->gsize, nt = ep->nt, nw = ep->n_wsamples, k = ep->n_modes;
    E2E_CHECK(g >= 2 && nt >= 1 && nw >= 2, E2E_ERR_INVALID,
              "env_generate: gsize >= 2, nt >= 1 and n_wsamples >= 2 required (sample_w divides by n_wsamples-1)");
    const int64_t n = (int64_t)g * g, m = 2 * n;
    E2E_CHECK(k >= 1 && k <= std::min<int64_t>(nw, n), E2E_ERR_INVALID,
              "env_generate: n_modes=%d must be in 1..min(n_wsamples, gsize^2)", k);
    E2E_CHECK(out->u && out->v && out->ui && out->vi && out->yi, E2E_ERR_INVALID,
              "env_generate: u, v, ui, vi, yi outputs are required");
    E2E_CHECK(ep->obstacle_mode >= 0 && ep->obstacle_mode <= 2, E2E_ERR_INVALID, "obstacle_mode must be 0, 1 or 2");
    const bool phi = out->phi || out->phi_i || out->phi_yi;
    E2E_CHECK(!phi || (out->phi && out->phi_i && out->phi_yi), E2E_ERR_INVALID,
              "env_generate: phi, phi_i and phi_yi go together");
    const double t_start = now_ms();
    hipStream_t s = nullptr;
    E2E_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct SG {
        hipStream_t s;
        ~SG() { (void)hipStreamDestroy(s); }
    } sg{s};
    Blas blas;
    E2E_BLAS(rocblas_create_handle(&blas.h));
    E2E_BLAS(rocblas_set_stream(blas.h, s));

    // device outputs
    DevBuf dU, dV, dUi, dVi, dYi, dS, dM, dPhi, dPhiI, dPhiY;
    E2E_TRY(dmalloc(dU, sizeof(float) * nt * n));
    E2E_TRY(dmalloc(dV, sizeof(float) * nt * n));
    E2E_TRY(dmalloc(dUi, sizeof(float) * nt * k * n));
    E2E_TRY(dmalloc(dVi, sizeof(float) * nt * k * n));
    E2E_TRY(dmalloc(dYi, sizeof(float) * nt * nw * k));
    if (phi) {
        E2E_TRY(dmalloc(dPhi, sizeof(float) * nt * n));
        E2E_TRY(dmalloc(dPhiI, sizeof(float) * nt * k * n));
        E2E_TRY(dmalloc(dPhiY, sizeof(float) * nt * nw * k));
    }
    // slices per pass: the velocity realizations of a pass stay within 8 GB
    const int64_t per_slice = (int64_t)nw * m * 8 + (phi ? (int64_t)nw * n * 8 : 0);
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(nt, (8ll << 30) / std::max<int64_t>(per_slice, 1)));
    DevBuf dR, dP, dMean, dPMean, dMc, dCc;
    E2E_TRY(dmalloc(dR, sizeof(double) * (size_t)chunk * nw * m));
    E2E_TRY(dmalloc(dMean, sizeof(double) * (size_t)chunk * m));
    E2E_TRY(dmalloc(dMc, sizeof(double) * (size_t)chunk * m * k));
    E2E_TRY(dmalloc(dCc, sizeof(double) * (size_t)chunk * nw * k));
    if (phi) {
        E2E_TRY(dmalloc(dP, sizeof(double) * (size_t)chunk * nw * n));
        E2E_TRY(dmalloc(dPMean, sizeof(double) * (size_t)chunk * n));
    }
    std::vector<double> sv((size_t)chunk * k);
    double svd_ms = 0.0;
    int max_iters = 0;
    for (int t0 = 0; t0 < nt; t0 += chunk) {
        const int nb = std::min(chunk, nt - t0);
        RealArgs A{dR.as<double>(), phi ? dP.as<double>() : nullptr, t0, nb, g, nw, ep->dt, ep->dxy, ep->A,
                   ep->eps, ep->wx, ep->wy, ep->w_lo, ep->w_hi};
        const int64_t nel = (int64_t)nb * nw * n;
        hipLaunchKernelGGL(realizations, dim3((unsigned)ceil_div(nel, 256)), dim3(256), 0, s, A);
        hipLaunchKernelGGL(mean_center, dim3((unsigned)ceil_div((int64_t)nb * m, 256)), dim3(256), 0, s,
                           dR.as<double>(), dMean.as<double>(), nb, nw, m);
        hipLaunchKernelGGL(emit_mean, dim3((unsigned)ceil_div((int64_t)nb * m, 256)), dim3(256), 0, s,
                           dMean.as<double>(), nb, m, n, dU.as<float>(), dV.as<float>(), t0);
        E2E_HIP(hipGetLastError());
        const double t1 = now_ms();
        int iters = 0;
        E2E_TRY(top_modes(blas.h, s, dR.as<double>(), nb, nw, m, k, dMc.as<double>(), dCc.as<double>(), sv.data(),
                          &iters));
        max_iters = std::max(max_iters, iters);
        hipLaunchKernelGGL(emit_modes, dim3((unsigned)(nb * k)), dim3(256), 0, s, dMc.as<double>(), dCc.as<double>(),
                           nb, m, nw, k, (int)n, (int64_t)k * n, dUi.as<float>(), dVi.as<float>(), dYi.as<float>(),
                           t0, nt);
        E2E_HIP(hipGetLastError());
        if (out->singular_values)
            for (int b = 0; b < nb; b++)
                for (int c = 0; c < k; c++) out->singular_values[(size_t)(t0 + b) * k + c] = sv[(size_t)b * k + c];
        if (phi) {
            hipLaunchKernelGGL(mean_center, dim3((unsigned)ceil_div((int64_t)nb * n, 256)), dim3(256), 0, s,
                               dP.as<double>(), dPMean.as<double>(), nb, nw, n);
            hipLaunchKernelGGL(emit_mean, dim3((unsigned)ceil_div((int64_t)nb * n, 256)), dim3(256), 0, s,
                               dPMean.as<double>(), nb, n, n, dPhi.as<float>(), dPhi.as<float>(), t0);
            E2E_HIP(hipGetLastError());
            E2E_TRY(top_modes(blas.h, s, dP.as<double>(), nb, nw, n, k, dMc.as<double>(), dCc.as<double>(),
                              sv.data(), &iters));
            max_iters = std::max(max_iters, iters);
            hipLaunchKernelGGL(emit_modes, dim3((unsigned)(nb * k)), dim3(256), 0, s, dMc.as<double>(),
                               dCc.as<double>(), nb, n, nw, k, (int)n, (int64_t)k * n, dPhiI.as<float>(),
                               dPhiI.as<float>(), dPhiY.as<float>(), t0, nt);
            E2E_HIP(hipGetLastError());
        }
        E2E_HIP(hipStreamSynchronize(s));
        svd_ms += now_ms() - t1;
    }
    // clouds and obstacles (final grid = this grid at refinement degree 1)
    if (out->s) {
        E2E_TRY(dmalloc(dS, sizeof(float) * nt * n));
        hipLaunchKernelGGL(clouds, dim3((unsigned)ceil_div((int64_t)nt * n, 256)), dim3(256), 0, s, dS.as<float>(), nt,
                           g, ep->dxy, ep->A_sc);
    }
    if (out->mask) {
        E2E_TRY(dmalloc(dM, sizeof(int32_t) * nt * n));
        hipLaunchKernelGGL(obstacles, dim3((unsigned)ceil_div((int64_t)nt * n, 256)), dim3(256), 0, s,
                           dM.as<int32_t>(), nt, g, ep->obstacle_mode, ep->obs_upper_row, ep->obs_left_col,
                           ep->obs_width, ep->obs_speed_x, ep->obs_speed_y);
    }
    E2E_HIP(hipGetLastError());
    auto down = [&](void* h, const DevBuf& d, size_t bytes) -> e2e_status {
        if (h) E2E_HIP(hipMemcpyAsync(h, d.p, bytes, hipMemcpyDeviceToHost, s));
        return E2E_OK;
    };
    E2E_TRY(down(out->u, dU, sizeof(float) * nt * n));
    E2E_TRY(down(out->v, dV, sizeof(float) * nt * n));
    E2E_TRY(down(out->ui, dUi, sizeof(float) * nt * k * n));
    E2E_TRY(down(out->vi, dVi, sizeof(float) * nt * k * n));
    E2E_TRY(down(out->yi, dYi, sizeof(float) * nt * nw * k));
    if (out->s) E2E_TRY(down(out->s, dS, sizeof(float) * nt * n));
    if (out->mask) E2E_TRY(down(out->mask, dM, sizeof(int32_t) * nt * n));
    if (phi) {
        E2E_TRY(down(out->phi, dPhi, sizeof(float) * nt * n));
        E2E_TRY(down(out->phi_i, dPhiI, sizeof(float) * nt * k * n));
        E2E_TRY(down(out->phi_yi, dPhiY, sizeof(float) * nt * nw * k));
    }
    E2E_HIP(hipStreamSynchronize(s));
    if (stats) {
        *stats = e2e_env_stats{};
        stats->total_ms = now_ms() - t_start;
        stats->modes_ms = svd_ms;
        stats->slices_per_pass = chunk;
        stats->max_iterations = max_iters;
    }
    return E2E_OK;
}

}  // namespace e2e
