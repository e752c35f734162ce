// vi.hip — (d) sparse value iteration: one launch per sweep computes
//   Q(s,a) = R(s,a) + discount * sum_j P(s,a,j) V(j)          (spvi_2.cu:186,232-235)
//   V'(s)  = max_a Q(s,a), pi(s) = first argmax, strict '>' from (-1e6, -1)
//                                                              (spvi_2.cu:93-122)
//   residual = max_s |V'(s) - V(s)|                            (spvi_2.cu:129-171,263-356)
// fused, with Q never materialised and the convergence test done on the device.
//
// Mapping (CDNA wave64): a state's na rows are contiguous in the state-major CSR, so
// a group of NAP*LPR lanes (NAP = next power of two >= na) takes one state: LPR lanes
// per action row stride that row's nnz, their partial sums combine with xor shuffles,
// then an xor-butterfly over the action lanes gives max/argmax. 64/(NAP*LPR) states
// per wave, 4 waves per block.
//
// Accuracy: each row sum is accumulated with an error-free product (fma) and TwoSum
// (the compensated dot product of Ogita, Rump and Oishi, "Accurate sum and dot
// product", SIAM J. Sci. Comput. 26(6), 2005), so Q carries about one float rounding;
// values stay within 1e-5 relative of a float64 solve over 200-step horizons.
//
// Convergence without host round trips: the device state holds the sweep index k and
// `conv`, the first sweep whose residual fell below thresh. Each block stores its
// residual maximum (plain store); the last block of a sweep (ticket counter) reduces
// them, publishes the residual, sets conv, and advances k. Blocks are 1024 threads and
// at most 512 per grid (grid-stride over states), so one sweep costs at most 512
// same-address atomics. A launch
// that finds conv >= 0 returns at once, so a fixed chain of launches (eager or a
// captured hipGraph) stops doing work exactly where the reference loop stops
// (spvi_2.cu:669-698).
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace e2e {

namespace {

struct VIArgs {
    const uint32_t* rp;
    const int2* nz;
    const float* R;
    float* V0;
    float* V1;
    int32_t* pi;
    VIState* st;
    float* part;   // [gridDim.x] per-block residual maxima of the running sweep
    int64_t ns;
    int na;
    float thresh;
    float discount;
};

constexpr int kVIThreads = 1024;  // 16 waves per block
constexpr int kVIWaves = kVIThreads / 64;
constexpr int kVIMaxBlocks = 512;  // 2 blocks per CU

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

template <int LPR, int NAP>
__global__ __launch_bounds__(kVIThreads) void spmv_vi_sweep(const VIArgs A) {
    constexpr int G = LPR * NAP;
    constexpr int SPW = 64 / G;
    __shared__ float wmax[kVIWaves];
    __shared__ int is_last;
    VIState* st = A.st;
    if (st->conv >= 0) return;  // converged at an earlier sweep
    const uint32_t k = st->k;
    const float* __restrict__ Vin = (k & 1u) ? A.V0 : A.V1;
    float* __restrict__ Vout = (k & 1u) ? A.V1 : A.V0;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int a = (lane % G) / LPR;
    const int l = lane % LPR;
    float res = 0.0f;
    // grid-stride over state groups: a few thousand blocks, so the per-sweep residual
    // and ticket atomics stay few
    const int64_t stride = (int64_t)gridDim.x * kVIWaves * SPW;
    for (int64_t s = ((int64_t)blockIdx.x * kVIWaves + wv) * SPW + lane / G; s - lane / G < A.ns; s += stride) {
        const bool valid = (s < A.ns) && (a < A.na);
        float sh = 0.0f, sl = 0.0f, rr = 0.0f;
        if (valid) {
            const int64_t row = s * A.na + a;
            const uint32_t beg = A.rp[row], end = A.rp[row + 1];
            rr = A.R[row];
            uint32_t j = beg + l;
            // two independent loads in flight per lane
            for (; j + LPR < end; j += 2 * LPR) {
                const int2 e0 = A.nz[j];
                const int2 e1 = A.nz[j + LPR];
                const float v0 = __int_as_float(e0.y), v1 = __int_as_float(e1.y);
                const float x0 = Vin[e0.x], x1 = Vin[e1.x];
                float p, pe, t, err;
                p = v0 * x0;
                pe = fmaf(v0, x0, -p);
                two_sum(sh, p, t, err);
                sh = t;
                sl = sl + (err + pe);
                p = v1 * x1;
                pe = fmaf(v1, x1, -p);
                two_sum(sh, p, t, err);
                sh = t;
                sl = sl + (err + pe);
            }
            if (j < end) {
                const int2 e0 = A.nz[j];
                const float v0 = __int_as_float(e0.y);
                const float x0 = Vin[e0.x];
                float p = v0 * x0, pe = fmaf(v0, x0, -p), t, err;
                two_sum(sh, p, t, err);
                sh = t;
                sl = sl + (err + pe);
            }
        }
#pragma unroll
        for (int o = 1; o < LPR; o <<= 1) {
            const float h2 = __shfl_xor(sh, o, 64), l2 = __shfl_xor(sl, o, 64);
            float t, err;
            two_sum(sh, h2, t, err);
            sh = t;
            sl = sl + (l2 + err);
        }
        if (A.discount != 1.0f) {
            sh *= A.discount;
            sl *= A.discount;
        }
        float q;
        {
            float t, err;
            two_sum(rr, sh, t, err);
            q = t + (err + sl);
        }
        if (!valid || q != q) q = -INFINITY;
        int ba = valid ? a : 0x7fffffff;
#pragma unroll
        for (int o = LPR; o < G; o <<= 1) {
            const float q2 = __shfl_xor(q, o, 64);
            const int a2 = __shfl_xor(ba, o, 64);
            if (q2 > q || (q2 == q && a2 < ba)) {
                q = q2;
                ba = a2;
            }
        }
        if (s < A.ns && (lane % G) == 0) {
            float vnew;
            int32_t pol;
            if (q > -1e6f) {
                vnew = q;
                pol = ba;
            } else {
                vnew = -1e6f;
                pol = -1;
            }
            Vout[s] = vnew;
            A.pi[s] = pol;
            res = fmaxf(res, fabsf(vnew - Vin[s]));
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) res = fmaxf(res, __shfl_xor(res, o, 64));
    if (lane == 0) wmax[wv] = res;
    __syncthreads();
    if (tid == 0) {
        float bm = 0.0f;
#pragma unroll
        for (int w = 0; w < kVIWaves; w++) bm = fmaxf(bm, wmax[w]);
        A.part[blockIdx.x] = bm;
        __threadfence();
        is_last = atomicAdd(&st->done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (is_last) {  // the last block of the sweep: reduce the partials, publish
        __threadfence();
        float r = 0.0f;
        for (unsigned b = tid; b < gridDim.x; b += kVIThreads)
            r = fmaxf(r, __hip_atomic_load(A.part + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) r = fmaxf(r, __shfl_xor(r, o, 64));
        __syncthreads();
        if (lane == 0) wmax[wv] = r;
        __syncthreads();
        if (tid == 0) {
            float m = 0.0f;
#pragma unroll
            for (int w = 0; w < kVIWaves; w++) m = fmaxf(m, wmax[w]);
            st->last_res = m;
            if (m < A.thresh) st->conv = (int32_t)k;
            st->done = 0u;
            st->k = k + 1u;
        }
    }
}

__global__ void vi_init(float* V0, int64_t ns, VIState* st) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) V0[i] = 0.0f;  // PV starts at 0 (spvi_2.cu:433-434,470)
    if (i == 0) {
        st->k = 1u;
        st->conv = -1;
        st->done = 0u;
        st->last_res = 0.0f;
    }
}

using SweepFn = void (*)(const VIArgs);

template <int NAP>
SweepFn pick_lpr(int lpr) {
    constexpr int maxl = 64 / NAP;
    if (lpr >= 16 && maxl >= 16) return spmv_vi_sweep<(maxl >= 16 ? 16 : 1), NAP>;
    if (lpr >= 8 && maxl >= 8) return spmv_vi_sweep<(maxl >= 8 ? 8 : 1), NAP>;
    if (lpr >= 4 && maxl >= 4) return spmv_vi_sweep<(maxl >= 4 ? 4 : 1), NAP>;
    if (lpr >= 2 && maxl >= 2) return spmv_vi_sweep<(maxl >= 2 ? 2 : 1), NAP>;
    return spmv_vi_sweep<1, NAP>;
}

int next_pow2(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

SweepFn pick_sweep(int na, int lpr, int* lpr_used) {
    const int nap = next_pow2(na);
    int maxl = 64 / nap;
    int l = 1;
    while (l * 2 <= lpr && l * 2 <= maxl && l < 16) l *= 2;
    *lpr_used = l;
    switch (nap) {
        case 1: return pick_lpr<1>(l);
        case 2: return pick_lpr<2>(l);
        case 4: return pick_lpr<4>(l);
        case 8: return pick_lpr<8>(l);
        case 16: return pick_lpr<16>(l);
        case 32: return pick_lpr<32>(l);
        default: return pick_lpr<64>(l);
    }
}

int auto_lpr(double nnz_per_row) {
    if (nnz_per_row <= 1.5) return 1;
    if (nnz_per_row <= 3.0) return 2;
    if (nnz_per_row <= 12.0) return 4;
    if (nnz_per_row <= 24.0) return 8;
    return 16;
}

}  // namespace

e2e_status vi_solve(Model* m, const e2e_solve_opts* o, uint32_t* policy, float* value,
                    e2e_solve_stats* stats) {
    const double t_start = now_ms();
    e2e_solve_opts opts{};
    opts.thresh = 0.1f;  // spvi_2.cu:757
    opts.discount = 1.0f;  // spvi_2.cu:61
    if (o) opts = *o;
    const int64_t ns = m->ns;
    int max_sweeps = opts.max_sweeps > 0 ? opts.max_sweeps : (m->built ? (int)m->nt + 2 : 100000);
    const int spc = opts.sweeps_per_check > 0 ? opts.sweeps_per_check : 16;
    hipStream_t s = m->stream;
    if (!m->V[0]) {
        if (hipMalloc(&m->V[0], sizeof(float) * ns) != hipSuccess ||
            hipMalloc(&m->V[1], sizeof(float) * ns) != hipSuccess ||
            hipMalloc(&m->pi, sizeof(int32_t) * ns) != hipSuccess ||
            hipMalloc(&m->st, sizeof(VIState)) != hipSuccess ||
            hipMalloc(&m->vi_part, sizeof(float) * kVIMaxBlocks) != hipSuccess) {
            set_error("vi_solve: device allocation failed");
            return E2E_ERR_NOMEM;
        }
    }
    int lpr_req = opts.lanes_per_row > 0 ? opts.lanes_per_row
                                         : auto_lpr(m->rows ? (double)m->nnz / (double)m->rows : 1.0);
    int lpr = 1;
    SweepFn fn = pick_sweep((int)m->na, lpr_req, &lpr);
    const int G = next_pow2((int)m->na) * lpr;
    const int64_t states_per_block = kVIWaves * (64 / G);
    const int64_t nblocks = std::min<int64_t>(ceil_div(ns, states_per_block), kVIMaxBlocks);
    VIArgs A{m->row_ptr, m->nz, m->R, m->V[0], m->V[1], m->pi, m->st, m->vi_part, ns, (int)m->na,
             opts.thresh, opts.discount};

    hipLaunchKernelGGL(vi_init, dim3((unsigned)ceil_div(ns, 256)), dim3(256), 0, s, m->V[0], ns, m->st);
    E2E_HIP(hipGetLastError());

    if (opts.use_graph) {
        if (!m->graph || m->graph_chunk != spc || m->graph_lpr != lpr || m->graph_thresh != opts.thresh ||
            m->graph_discount != opts.discount) {
            if (m->graph) {
                (void)hipGraphExecDestroy(m->graph);
                m->graph = nullptr;
            }
            hipGraph_t gr;
            E2E_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < spc; i++)
                hipLaunchKernelGGL(fn, dim3((unsigned)nblocks), dim3(kVIThreads), 0, s, A);
            E2E_HIP(hipStreamEndCapture(s, &gr));
            hipError_t e = hipGraphInstantiate(&m->graph, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) {
                m->graph = nullptr;
                set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
                return E2E_ERR_HIP;
            }
            m->graph_chunk = spc;
            m->graph_lpr = lpr;
            m->graph_thresh = opts.thresh;
            m->graph_discount = opts.discount;
        }
    }

    hipEvent_t e0, e1;
    E2E_HIP(hipEventCreate(&e0));
    E2E_HIP(hipEventCreate(&e1));
    double kms = 0.0;
    int launched = 0;
    VIState hst{};
    e2e_status rc = E2E_OK;
    while (launched < max_sweeps) {
        int n = std::min(spc, max_sweeps - launched);
        E2E_HIP(hipEventRecord(e0, s));
        if (opts.use_graph && n == spc) {
            E2E_HIP(hipGraphLaunch(m->graph, s));
        } else {
            for (int i = 0; i < n; i++) hipLaunchKernelGGL(fn, dim3((unsigned)nblocks), dim3(kVIThreads), 0, s, A);
            E2E_HIP(hipGetLastError());
        }
        E2E_HIP(hipEventRecord(e1, s));
        E2E_HIP(hipMemcpyAsync(&hst, m->st, sizeof(VIState), hipMemcpyDeviceToHost, s));
        E2E_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        E2E_HIP(hipEventElapsedTime(&ms, e0, e1));
        kms += ms;
        launched += n;
        if (hst.conv >= 0) break;
        if (opts.timeout_ms > 0 && now_ms() - t_start > opts.timeout_ms) {
            rc = E2E_ERR_TIMEOUT;
            break;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int sweeps = hst.conv >= 0 ? hst.conv : (int)hst.k - 1;
    if (rc == E2E_OK && hst.conv < 0) rc = E2E_ERR_NOT_CONVERGED;
    if (rc == E2E_ERR_NOT_CONVERGED) set_error("not converged after %d sweeps (residual %g)", sweeps, hst.last_res);
    if (rc == E2E_ERR_TIMEOUT) set_error("timeout after %d sweeps", sweeps);
    const float* Vfin = (sweeps & 1) ? m->V[1] : m->V[0];
    if (value) E2E_HIP(hipMemcpyAsync(value, Vfin, sizeof(float) * ns, hipMemcpyDeviceToHost, s));
    if (policy) E2E_HIP(hipMemcpyAsync(policy, m->pi, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, s));
    E2E_HIP(hipStreamSynchronize(s));
    if (stats) {
        stats->sweeps = sweeps;
        stats->converged = hst.conv >= 0;
        stats->final_residual = hst.last_res;
        stats->lanes_per_row = lpr;
        stats->solve_ms = now_ms() - t_start;
        stats->sweep_kernel_ms = kms;
        stats->sweep_kernel_ms_per_sweep = sweeps > 0 ? kms / sweeps : 0.0;
        stats->bytes_per_sweep = 8.0 * m->nnz + 4.0 * (m->rows + 1) + 4.0 * m->rows + 12.0 * ns;
        stats->launches = launched;
    }
    return rc;
}

}  // namespace e2e
