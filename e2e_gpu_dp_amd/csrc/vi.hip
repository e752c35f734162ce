// vi.hip — (d) sparse value iteration: one launch per sweep computes
//   Q(s,a) = R(s,a) + discount * sum_j P(s,a,j) V(j)          (spvi_2.cu:186,232-235)
//   V'(s)  = max_a Q(s,a), pi(s) = first argmax, strict '>' from (-1e6, -1)
//                                                              (spvi_2.cu:93-122)
//   residual = max_s |V'(s) - V(s)|                            (spvi_2.cu:129-171,263-356)
// fused, with Q never materialised and the convergence test done on the device.
//
// Mapping (CDNA wave64): a state's na rows are contiguous in the state-major CSR, so a
// wave takes 64/NAP consecutive states (NAP = next power of two >= na) and streams their
// contiguous nnz range with independent coalesced loads (see spmv_vi_sweep).
//
// Accuracy: each row sum is accumulated with an error-free product (fma) and TwoSum
// (the compensated dot product of Ogita, Rump and Oishi, "Accurate sum and dot
// product", SIAM J. Sci. Comput. 26(6), 2005), so Q carries about one float rounding;
// values stay within 1e-5 relative of a float64 solve over 200-step horizons.
//
// Convergence without host round trips: the device state holds the sweep index k and
// `conv`, the first sweep whose residual fell below thresh. Each block stores its
// residual maximum (plain store); the last block of a sweep (ticket counter) reduces
// them, publishes the residual, sets conv, and advances k. Blocks are 512 threads and
// at most 1024 per grid (grid-stride over state chunks), so one sweep costs at most 1024
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

constexpr int kVIThreads = 512;   // 8 waves per block
constexpr int kVIWaves = kVIThreads / 64;
constexpr int kVIMaxBlocks = 1024;  // 4 blocks per CU

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

// One wave takes a chunk of SPC = 64/NAP consecutive states (NAP = next power of two
// >= na; lane = state_in_chunk * NAP + action). Per chunk:
//   1. each lane loads its row's two row pointers;
//   2. the chunk's nnz range is streamed in windows of 64*U entries: every lane issues U
//      independent coalesced 8-byte loads and U value gathers, forms the error-free
//      products (p = v*x, pe = fma(v, x, -p)) and parks them in its wave's LDS;
//   3. each lane sums its own row from LDS in ascending order (TwoSum, compensation
//      carried), adds R;
//   4. an xor butterfly over the NAP lanes of a state gives max / first argmax.
// All loads of a window are independent, so a wave keeps about 2*U*64 memory requests
// in flight instead of one state's dependent chain.
template <int NAP, int U>
__global__ __launch_bounds__(kVIThreads) void spmv_vi_sweep(const VIArgs A) {
    constexpr int SPC = 64 / NAP;
    constexpr int WIN = 64 * U;
    __shared__ float wmax[kVIWaves];
    __shared__ int is_last;
    __shared__ float2 stage[kVIWaves][WIN];
    VIState* st = A.st;
    if (st->conv >= 0) return;  // converged at an earlier sweep
    const uint32_t k = st->k;
    const float* __restrict__ Vin = (k & 1u) ? A.V0 : A.V1;
    float* __restrict__ Vout = (k & 1u) ? A.V1 : A.V0;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int a = lane % NAP, sl = lane / NAP;
    float2* wst = stage[wv];
    float res = 0.0f;
    const int64_t nchunks = (A.ns + SPC - 1) / SPC;
    for (int64_t ch = (int64_t)blockIdx.x * kVIWaves + wv; ch < nchunks; ch += (int64_t)gridDim.x * kVIWaves) {
        const int64_t s = ch * SPC + sl;
        const bool valid = (s < A.ns) && (a < A.na);
        const int64_t row = s * A.na + a;
        uint32_t rb = 0, re = 0;
        float rr = 0.0f;
        if (valid) {
            rb = A.rp[row];
            re = A.rp[row + 1];
            rr = A.R[row];
        }
        // chunk range: first valid lane's begin, last valid lane's end
        const uint64_t vm = __ballot(valid);
        const uint32_t cb = __shfl(rb, __ffsll((long long)vm) - 1, 64);
        const uint32_t ce = __shfl(re, 63 - __clzll((long long)vm), 64);
        float sh = 0.0f, slo = 0.0f;
        for (uint32_t w0 = cb; w0 < ce; w0 += WIN) {
            const uint32_t wn = min((uint32_t)WIN, ce - w0);
            int2 e[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t i = (uint32_t)(u * 64 + lane);
                e[u] = i < wn ? A.nz[w0 + i] : make_int2(0, 0);
            }
            float x[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t i = (uint32_t)(u * 64 + lane);
                x[u] = i < wn ? Vin[e[u].x] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t i = (uint32_t)(u * 64 + lane);
                if (i < wn) {
                    const float v = __int_as_float(e[u].y);
                    const float p = v * x[u];
                    wst[i] = make_float2(p, fmaf(v, x[u], -p));
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (valid) {
                const uint32_t jb = max(rb, w0), je = min(re, w0 + wn);
                for (uint32_t j = jb; j < je; j++) {
                    const float2 pp = wst[j - w0];
                    float t, err;
                    two_sum(sh, pp.x, t, err);
                    sh = t;
                    slo = slo + (err + pp.y);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (A.discount != 1.0f) {
            sh *= A.discount;
            slo *= A.discount;
        }
        float q;
        {
            float t, err;
            two_sum(rr, sh, t, err);
            q = t + (err + slo);
        }
        if (!valid || q != q) q = -INFINITY;
        int ba = valid ? a : 0x7fffffff;
#pragma unroll
        for (int o = 1; o < NAP; o <<= 1) {
            const float q2 = __shfl_xor(q, o, 64);
            const int a2 = __shfl_xor(ba, o, 64);
            if (q2 > q || (q2 == q && a2 < ba)) {
                q = q2;
                ba = a2;
            }
        }
        if (s < A.ns && a == 0) {
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

int next_pow2(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

// U: nnz loads in flight per lane per window (the window is 64*U entries).
template <int U>
SweepFn pick_nap(int nap) {
    switch (nap) {
        case 1: return spmv_vi_sweep<1, U>;
        case 2: return spmv_vi_sweep<2, U>;
        case 4: return spmv_vi_sweep<4, U>;
        case 8: return spmv_vi_sweep<8, U>;
        case 16: return spmv_vi_sweep<16, U>;
        case 32: return spmv_vi_sweep<32, U>;
        default: return spmv_vi_sweep<64, U>;
    }
}

SweepFn pick_sweep(int na, int u, int* u_used) {
    const int nap = next_pow2(na);
    if (u <= 2) {
        *u_used = 2;
        return pick_nap<2>(nap);
    }
    if (u <= 4) {
        *u_used = 4;
        return pick_nap<4>(nap);
    }
    *u_used = 8;
    return pick_nap<8>(nap);
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
    // loads in flight per lane: a chunk holds about 64/NAP states of nnz/row * na entries
    const int nap = next_pow2((int)m->na);
    const double per_chunk = (m->rows ? (double)m->nnz / (double)m->rows : 1.0) * m->na * (64 / nap);
    int u_req = opts.loads_in_flight > 0 ? opts.loads_in_flight : (per_chunk > 192 ? 8 : per_chunk > 96 ? 4 : 2);
    int lpr = 0;
    SweepFn fn = pick_sweep((int)m->na, u_req, &lpr);
    const int64_t states_per_block = kVIWaves * (64 / nap);
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
        stats->loads_in_flight = lpr;
        stats->solve_ms = now_ms() - t_start;
        stats->sweep_kernel_ms = kms;
        stats->sweep_kernel_ms_per_sweep = sweeps > 0 ? kms / sweeps : 0.0;
        stats->bytes_per_sweep = 8.0 * m->nnz + 4.0 * (m->rows + 1) + 4.0 * m->rows + 12.0 * ns;
        stats->launches = launched;
    }
    return rc;
}

}  // namespace e2e
