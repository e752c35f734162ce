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
    const uint32_t* wpart;  // [waves+1] state partition of the streaming kernel
};

constexpr int kVIThreads = 512;   // 8 waves per block
constexpr int kVIWaves = kVIThreads / 64;
constexpr int kVIMaxBlocks = 2048;  // 8 blocks of 512 threads per CU on 256 CUs

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

// One wave takes chunks of SPC = 64/NAP consecutive states (NAP = next power of two
// >= na; lane = state_in_chunk * NAP + action). Per chunk:
//   1. each lane loads its row's two row pointers, R, and (action-0 lanes) the state's
//      previous value;
//   2. the chunk's nnz range is streamed in windows of 64*U entries: every lane issues U
//      independent coalesced 8-byte loads and U value gathers, forms the error-free
//      products (p = v*x, pe = fma(v, x, -p)) and parks them in its wave's LDS;
//   3. each lane sums its own row from LDS in ascending order (TwoSum, compensation
//      carried), adds R;
//   4. an xor butterfly over the NAP lanes of a state gives max / first argmax.
// Software pipeline across the wave's chunks: the next chunk's row data is requested
// together with this chunk's gathers, and the next chunk's first nnz window is
// requested before this chunk's sums are formed, so about one memory latency per chunk
// is exposed instead of four.
template <int NAP, int U, int ACC>
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
    double* wsd = reinterpret_cast<double*>(wst);
    float res = 0.0f;
    const int64_t nchunks = (A.ns + SPC - 1) / SPC;
    const int64_t cstride = (int64_t)gridDim.x * kVIWaves;

    struct Meta {
        uint32_t rb, re;
        float rr, vold;
        bool valid;
    };
    auto load_meta = [&](int64_t ch, Meta& M) {
        const int64_t s = ch * SPC + sl;
        M.valid = (ch < nchunks) && (s < A.ns) && (a < A.na);
        M.rb = 0u;
        M.re = 0u;
        M.rr = 0.0f;
        M.vold = 0.0f;
        if (M.valid) {
            const int64_t row = s * A.na + a;
            M.rb = A.rp[row];
            M.re = A.rp[row + 1];
            M.rr = A.R[row];
            if (a == 0) M.vold = Vin[s];
        }
    };
    auto range_of = [&](const Meta& M, uint32_t& cb, uint32_t& ce) {
        const uint64_t vm = __ballot(M.valid);
        cb = vm ? (uint32_t)__shfl((int)M.rb, __ffsll((long long)vm) - 1, 64) : 0u;
        ce = vm ? (uint32_t)__shfl((int)M.re, 63 - __clzll((long long)vm), 64) : 0u;
    };
    auto issue_nz = [&](uint32_t cb, uint32_t ce, int2 (&e)[U]) {
        const uint32_t wn = min((uint32_t)WIN, ce - cb);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t i = (uint32_t)(u * 64 + lane);
            e[u] = i < wn ? A.nz[cb + i] : make_int2(0, 0);
        }
    };
    // products of one window into LDS, then each lane adds its row's part of the window.
    // ACC 0: error-free products (p, pe) and TwoSum (compensated); ACC 1: (prob, value)
    // pairs and one fma per entry.
    auto window = [&](const Meta& M, uint32_t w0, uint32_t wn, const int2 (&e)[U], float& sh, float& slo,
                      double& sd) {
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
                if (ACC == 0) {
                    const float p = v * x[u];
                    wst[i] = make_float2(p, fmaf(v, x[u], -p));
                } else if (ACC == 3) {
                    wsd[i] = (double)v * (double)x[u];  // exact: 24x24-bit product
                } else {
                    wst[i] = make_float2(v, x[u]);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (M.valid) {
            const uint32_t jb = max(M.rb, w0), je = min(M.re, w0 + wn);
            uint32_t j = jb;
            if (ACC == 3) {
                for (; j + 4 <= je; j += 4) {
                    double d4[4];
#pragma unroll
                    for (int t = 0; t < 4; t++) d4[t] = wsd[j + t - w0];
#pragma unroll
                    for (int t = 0; t < 4; t++) sd += d4[t];
                }
                for (; j < je; j++) sd += wsd[j - w0];
            }
            for (; ACC != 3 && j + 4 <= je; j += 4) {
                float2 q4[4];
#pragma unroll
                for (int t = 0; t < 4; t++) q4[t] = wst[j + t - w0];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    if (ACC == 0) {
                        float tt, err;
                        two_sum(sh, q4[t].x, tt, err);
                        sh = tt;
                        slo = slo + (err + q4[t].y);
                    } else {
                        sh = fmaf(q4[t].x, q4[t].y, sh);
                    }
                }
            }
            for (; ACC != 3 && j < je; j++) {
                const float2 pp = wst[j - w0];
                if (ACC == 0) {
                    float tt, err;
                    two_sum(sh, pp.x, tt, err);
                    sh = tt;
                    slo = slo + (err + pp.y);
                } else {
                    sh = fmaf(pp.x, pp.y, sh);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    };

    int64_t ch = (int64_t)blockIdx.x * kVIWaves + wv;
    Meta cur;
    uint32_t cb = 0, ce = 0;
    int2 ecur[U];
    load_meta(ch, cur);
    range_of(cur, cb, ce);
    issue_nz(cb, ce, ecur);
    for (; ch < nchunks; ch += cstride) {
        Meta nxt;
        load_meta(ch + cstride, nxt);  // in flight with this chunk's gathers
        float sh = 0.0f, slo = 0.0f;
        double sd = 0.0;
        const uint32_t wn0 = min((uint32_t)WIN, ce - cb);
        window(cur, cb, wn0, ecur, sh, slo, sd);
        // the next chunk's first window goes out before this chunk's remaining work
        uint32_t nb, ne;
        range_of(nxt, nb, ne);
        int2 enxt[U];
        issue_nz(nb, ne, enxt);
        for (uint32_t w0 = cb + wn0; w0 < ce; w0 += WIN) {  // chunks larger than a window
            issue_nz(w0, ce, ecur);  // ecur is consumed: reuse its registers
            window(cur, w0, min((uint32_t)WIN, ce - w0), ecur, sh, slo, sd);
        }
        float q;
        if (ACC == 3) {
            q = (float)((double)cur.rr + (double)A.discount * sd);
        } else {
            if (A.discount != 1.0f) {
                sh *= A.discount;
                slo *= A.discount;
            }
            float tt, err;
            two_sum(cur.rr, sh, tt, err);
            q = tt + (err + slo);
        }
        if (!cur.valid || q != q) q = -INFINITY;
        int ba = cur.valid ? a : 0x7fffffff;
#pragma unroll
        for (int o = 1; o < NAP; o <<= 1) {
            const float q2 = __shfl_xor(q, o, 64);
            const int a2 = __shfl_xor(ba, o, 64);
            if (q2 > q || (q2 == q && a2 < ba)) {
                q = q2;
                ba = a2;
            }
        }
        const int64_t s = ch * SPC + sl;
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
            res = fmaxf(res, fabsf(vnew - cur.vold));
        }
        cur = nxt;
        cb = nb;
        ce = ne;
#pragma unroll
        for (int u = 0; u < U; u++) ecur[u] = enxt[u];
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

// ACC 2: one lane per row, streaming that row's entries directly (no LDS): 4 loads
// per lane in flight, one fma per entry. Same chunk mapping and pipelined row data.
template <int NAP>
__global__ __launch_bounds__(kVIThreads) void spmv_vi_rows(const VIArgs A) {
    constexpr int SPC = 64 / NAP;
    __shared__ float wmax[kVIWaves];
    __shared__ int is_last;
    VIState* st = A.st;
    if (st->conv >= 0) return;
    const uint32_t k = st->k;
    const float* __restrict__ Vin = (k & 1u) ? A.V0 : A.V1;
    float* __restrict__ Vout = (k & 1u) ? A.V1 : A.V0;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int a = lane % NAP, sl = lane / NAP;
    float res = 0.0f;
    const int64_t nchunks = (A.ns + SPC - 1) / SPC;
    const int64_t cstride = (int64_t)gridDim.x * kVIWaves;
    for (int64_t ch = (int64_t)blockIdx.x * kVIWaves + wv; ch < nchunks; ch += cstride) {
        const int64_t s = ch * SPC + sl;
        const bool valid = (s < A.ns) && (a < A.na);
        float acc = 0.0f, rr = 0.0f, vold = 0.0f;
        if (valid) {
            const int64_t row = s * A.na + a;
            const uint32_t rb = A.rp[row], re = A.rp[row + 1];
            rr = A.R[row];
            if (a == 0) vold = Vin[s];
            uint32_t j = rb;
            for (; j + 4 <= re; j += 4) {
                int2 e[4];
#pragma unroll
                for (int t = 0; t < 4; t++) e[t] = A.nz[j + t];
                float x[4];
#pragma unroll
                for (int t = 0; t < 4; t++) x[t] = Vin[e[t].x];
#pragma unroll
                for (int t = 0; t < 4; t++) acc = fmaf(__int_as_float(e[t].y), x[t], acc);
            }
            for (; j < re; j++) {
                const int2 e = A.nz[j];
                acc = fmaf(__int_as_float(e.y), Vin[e.x], acc);
            }
        }
        float q = valid ? rr + A.discount * acc : -INFINITY;
        if (q != q) q = -INFINITY;
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
            res = fmaxf(res, fabsf(vnew - vold));
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
    if (is_last) {
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

// ---------------------------------------------------------------------------------
// Streaming sweep (the default). Each wavefront owns one contiguous range of states,
// [wpart[w], wpart[w+1]), cut so that every wave has about the same nnz + 2 * rows.
// Its rows are contiguous in the state-major CSR, so its entries form ONE contiguous
// stream [rp[s_begin*na], rp[s_end*na]) that is read in fixed windows of 64*U entries,
// with no dependence on the row pointers:
//   window k+2's entries are requested, window k+1's value gathers are requested, and
//   window k's products are formed and parked in LDS, all in the same step, so two
//   windows of entries and one of gathers are in flight while a window is reduced.
// Rows are consumed in groups of SPC = 64/NAP states (one row per lane, as above);
// a group's row data (row pointers, R, old value) is requested two groups ahead. When a
// window contains the end of a group, the group is finished: max / first argmax over
// the NAP lanes of each state, value + policy stores, residual.
// Windows may run past the wave's stream: the entry array carries kNzPad zero entries
// past nnz, and entries outside a lane's row range are never summed.
// XCD-aware: blocks are dealt to the 8 XCDs round robin, so block b takes partition
// (b % 8) * (nb / 8) + b / 8 and each XCD (and its L2) works on one eighth of the states.
// ACC 0: float products with their rounding errors (fma) and TwoSum; ACC 3: float64
// products (exact) and float64 sums; both round Q to float once.
constexpr int kStreamThreads = 256;
constexpr int kStreamWaves = kStreamThreads / 64;

template <int NAP, int U, int ACC>
__global__ __launch_bounds__(kStreamThreads) void spmv_vi_stream(const VIArgs A) {
    constexpr int SPC = 64 / NAP;
    constexpr int WIN = 64 * U;
    __shared__ float wmax[kStreamWaves];
    __shared__ int is_last;
    __shared__ double stage[kStreamWaves][WIN];
    VIState* st = A.st;
    if (st->conv >= 0) return;  // converged at an earlier sweep
    const uint32_t k = st->k;
    const float* __restrict__ Vin = (k & 1u) ? A.V0 : A.V1;
    float* __restrict__ Vout = (k & 1u) ? A.V1 : A.V0;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int a = lane % NAP, sl = lane / NAP;
    const int na = A.na;
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned pb = (nb % 8u == 0u) ? (b % 8u) * (nb / 8u) + b / 8u : b;
    const unsigned w = pb * kStreamWaves + wv;
    const int64_t s_begin = A.wpart[w], s_end = A.wpart[w + 1];
    double* wsd = stage[wv];
    float2* wsf = reinterpret_cast<float2*>(stage[wv]);
    float res = 0.0f;

    if (s_begin < s_end) {  // wave-uniform
        const int64_t ngroups = (s_end - s_begin + SPC - 1) / SPC;
        const uint32_t P0 = A.rp[s_begin * na];
        const uint32_t Pend = A.rp[s_end * na];
        const uint32_t nwin = (Pend - P0 + WIN - 1) / WIN;
        const int2* __restrict__ nzp = A.nz + P0 + lane;

        struct Meta {
            uint32_t rb, re, ce;
            float rr, vold;
            bool valid;
        };
        auto load_meta = [&](int64_t g, Meta& M) {
            const int64_t s0 = s_begin + g * SPC, s = s0 + sl;
            M.valid = (g < ngroups) && (s < s_end) && (a < na);
            M.rb = 0u;
            M.re = 0u;
            M.ce = 0u;
            M.rr = 0.0f;
            M.vold = 0.0f;
            if (g < ngroups) M.ce = A.rp[min(s0 + SPC, s_end) * na];  // same address in all lanes
            if (M.valid) {
                const int64_t row = s * na + a;
                M.rb = A.rp[row];
                M.re = A.rp[row + 1];
                M.rr = A.R[row];
                if (a == 0) M.vold = Vin[s];
            }
        };
        auto load_win = [&](uint32_t kw, int2(&e)[U]) {
            const int2* p = nzp + (size_t)kw * WIN;
#pragma unroll
            for (int u = 0; u < U; u++) e[u] = p[u * 64];
        };
        auto gather = [&](const int2(&e)[U], float(&x)[U]) {
#pragma unroll
            for (int u = 0; u < U; u++) x[u] = Vin[e[u].x];
        };

        Meta m0, m1, m2;
        int64_t g = 0;
        double sd = 0.0;
        float sh = 0.0f, slo = 0.0f;
        auto finish_group = [&]() {
            float q;
            if (ACC == 3) {
                q = (float)((double)m0.rr + (double)A.discount * sd);
            } else {
                float hs = sh, ls = slo;
                if (A.discount != 1.0f) {
                    hs *= A.discount;
                    ls *= A.discount;
                }
                float tt, err;
                two_sum(m0.rr, hs, tt, err);
                q = tt + (err + ls);
            }
            if (!m0.valid || q != q) q = -INFINITY;
            int ba = m0.valid ? a : 0x7fffffff;
#pragma unroll
            for (int o = 1; o < NAP; o <<= 1) {
                const float q2 = __shfl_xor(q, o, 64);
                const int a2 = __shfl_xor(ba, o, 64);
                if (q2 > q || (q2 == q && a2 < ba)) {
                    q = q2;
                    ba = a2;
                }
            }
            if (m0.valid && a == 0) {
                const int64_t s = s_begin + g * SPC + sl;
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
                res = fmaxf(res, fabsf(vnew - m0.vold));
            }
            sd = 0.0;
            sh = 0.0f;
            slo = 0.0f;
            g++;
            m0 = m1;
            m1 = m2;
            load_meta(g + 2, m2);
        };

        int2 e0[U], e1[U], e2[U];
        float x0[U], x1[U];
        if (nwin > 0) load_win(0, e0);
        if (nwin > 1) load_win(1, e1);
        load_meta(0, m0);
        load_meta(1, m1);
        load_meta(2, m2);
        if (nwin > 0) gather(e0, x0);
        for (uint32_t kw = 0; kw < nwin; kw++) {
            if (kw + 2 < nwin) load_win(kw + 2, e2);
            if (kw + 1 < nwin) gather(e1, x1);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const float v = __int_as_float(e0[u].y);
                if (ACC == 3) {
                    wsd[u * 64 + lane] = (double)v * (double)x0[u];  // exact: 24x24-bit product
                } else {
                    const float p = v * x0[u];
                    wsf[u * 64 + lane] = make_float2(p, fmaf(v, x0[u], -p));
                }
            }
            __builtin_amdgcn_wave_barrier();
            const uint32_t w0 = P0 + kw * WIN, wend = w0 + WIN;
            for (;;) {
                if (m0.valid) {  // this lane's row part inside the window, ascending
                    const uint32_t jb = max(m0.rb, w0), je = min(m0.re, wend);
                    if (ACC == 3) {
                        uint32_t j = jb;
                        for (; j + 2 <= je; j += 2) {
                            const double d0 = wsd[j - w0], d1 = wsd[j + 1 - w0];
                            sd += d0;
                            sd += d1;
                        }
                        if (j < je) sd += wsd[j - w0];
                    } else {
                        for (uint32_t j = jb; j < je; j++) {
                            const float2 pp = wsf[j - w0];
                            float tt, err;
                            two_sum(sh, pp.x, tt, err);
                            sh = tt;
                            slo = slo + (err + pp.y);
                        }
                    }
                }
                if (m0.ce > wend) break;  // the group continues in the next window
                finish_group();
                if (g >= ngroups) break;
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int u = 0; u < U; u++) {
                e0[u] = e1[u];
                e1[u] = e2[u];
                x0[u] = x1[u];
            }
        }
        while (g < ngroups) finish_group();  // trailing groups whose rows are all empty
    }

#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) res = fmaxf(res, __shfl_xor(res, o, 64));
    if (lane == 0) wmax[wv] = res;
    __syncthreads();
    if (tid == 0) {
        float bm = 0.0f;
#pragma unroll
        for (int i = 0; i < kStreamWaves; i++) bm = fmaxf(bm, wmax[i]);
        A.part[blockIdx.x] = bm;
        __threadfence();
        is_last = atomicAdd(&st->done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (is_last) {  // the last block of the sweep: reduce the partials, publish
        __threadfence();
        float r = 0.0f;
        for (unsigned i = tid; i < gridDim.x; i += kStreamThreads)
            r = fmaxf(r, __hip_atomic_load(A.part + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) r = fmaxf(r, __shfl_xor(r, o, 64));
        __syncthreads();
        if (lane == 0) wmax[wv] = r;
        __syncthreads();
        if (tid == 0) {
            float mx = 0.0f;
#pragma unroll
            for (int i = 0; i < kStreamWaves; i++) mx = fmaxf(mx, wmax[i]);
            st->last_res = mx;
            if (mx < A.thresh) st->conv = (int32_t)k;
            st->done = 0u;
            st->k = k + 1u;
        }
    }
}

// wpart[w] = first state of wave w: the smallest multiple of spc whose weight
// rp[s*na] + 2*s*na reaches w/W of the total (monotone in w), wpart[W] = ns.
__global__ void vi_partition(const uint32_t* __restrict__ rp, int64_t ns, int na, int64_t W, int spc,
                             uint32_t* __restrict__ wpart) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w > W) return;
    if (w == 0 || w == W) {
        wpart[w] = w == 0 ? 0u : (uint32_t)ns;
        return;
    }
    const double total = (double)rp[ns * na] + 2.0 * (double)(ns * na);
    const double target = total * (double)w / (double)W;
    int64_t lo = 0, hi = ns;  // smallest s with weight(s) >= target
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        const double wt = (double)rp[mid * na] + 2.0 * (double)(mid * na);
        if (wt >= target)
            hi = mid;
        else
            lo = mid + 1;
    }
    int64_t s = (lo + spc / 2) / spc * spc;
    wpart[w] = (uint32_t)(s < ns ? s : ns);
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
template <int U, int ACC>
SweepFn pick_nap(int nap) {
    switch (nap) {
        case 1: return spmv_vi_sweep<1, U, ACC>;
        case 2: return spmv_vi_sweep<2, U, ACC>;
        case 4: return spmv_vi_sweep<4, U, ACC>;
        case 8: return spmv_vi_sweep<8, U, ACC>;
        case 16: return spmv_vi_sweep<16, U, ACC>;
        case 32: return spmv_vi_sweep<32, U, ACC>;
        default: return spmv_vi_sweep<64, U, ACC>;
    }
}

SweepFn pick_rows(int nap) {
    switch (nap) {
        case 1: return spmv_vi_rows<1>;
        case 2: return spmv_vi_rows<2>;
        case 4: return spmv_vi_rows<4>;
        case 8: return spmv_vi_rows<8>;
        case 16: return spmv_vi_rows<16>;
        case 32: return spmv_vi_rows<32>;
        default: return spmv_vi_rows<64>;
    }
}

template <int ACC>
SweepFn pick_u(int nap, int u, int* u_used) {
    if (u <= 2) {
        *u_used = 2;
        return pick_nap<2, ACC>(nap);
    }
    if (u <= 4) {
        *u_used = 4;
        return pick_nap<4, ACC>(nap);
    }
    *u_used = 8;
    return pick_nap<8, ACC>(nap);
}

template <int U, int ACC>
SweepFn pick_stream_nap(int nap) {
    switch (nap) {
        case 1: return spmv_vi_stream<1, U, ACC>;
        case 2: return spmv_vi_stream<2, U, ACC>;
        case 4: return spmv_vi_stream<4, U, ACC>;
        case 8: return spmv_vi_stream<8, U, ACC>;
        case 16: return spmv_vi_stream<16, U, ACC>;
        case 32: return spmv_vi_stream<32, U, ACC>;
        default: return spmv_vi_stream<64, U, ACC>;
    }
}

template <int ACC>
SweepFn pick_stream(int nap, int u, int* u_used) {
    if (u <= 2) {
        *u_used = 2;
        return pick_stream_nap<2, ACC>(nap);
    }
    if (u <= 4) {
        *u_used = 4;
        return pick_stream_nap<4, ACC>(nap);
    }
    *u_used = 8;
    return pick_stream_nap<8, ACC>(nap);
}

// 0-based variants: 0 staged, compensated; 1 staged, fma; 2 row per lane, fma;
// 3 staged, float64 products and sums; 4 streaming, compensated; 5 streaming, float64
constexpr int kDefaultVariant = 5;
bool is_stream_variant(int variant) { return variant == 4 || variant == 5; }

SweepFn pick_sweep(int na, int variant, int u, int* u_used) {
    const int nap = next_pow2(na);
    switch (variant) {
        case 1: return pick_u<1>(nap, u, u_used);
        case 2: *u_used = 4; return pick_rows(nap);
        case 3: return pick_u<3>(nap, u, u_used);
        case 4: return pick_stream<0>(nap, u, u_used);
        case 5: return pick_stream<3>(nap, u, u_used);
        default: return pick_u<0>(nap, u, u_used);
    }
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
    const int nap = next_pow2((int)m->na);
    const int grp = 64 / nap;
    const int variant = opts.kernel_variant > 0 ? opts.kernel_variant - 1 : kDefaultVariant;
    const bool stream = is_stream_variant(variant);
    // loads in flight per lane. Chunked kernels: a chunk holds about 64/NAP states of
    // nnz/row * na entries. Streaming kernel: 4 (two windows of 256 entries in flight).
    const double per_chunk = (m->rows ? (double)m->nnz / (double)m->rows : 1.0) * m->na * grp;
    int u_req = opts.loads_in_flight > 0 ? opts.loads_in_flight
                : stream ? 4 : (per_chunk > 192 ? 8 : per_chunk > 96 ? 4 : 2);
    int lpr = 0;
    SweepFn fn = pick_sweep((int)m->na, variant, u_req, &lpr);
    const int threads = stream ? kStreamThreads : kVIThreads;
    // persistent grid: as many blocks as can be resident at once (grid-stride over chunks,
    // or one contiguous share per wave), so that no block starts after the others finished
    const int64_t states_per_block = (threads / 64) * grp;
    int bpc = opts.blocks_per_cu;
    if (bpc <= 0) E2E_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, (const void*)fn, threads, 0));
    int ncu = 0, dev = 0;
    E2E_HIP(hipGetDevice(&dev));
    E2E_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t resident = (int64_t)std::max(bpc, 1) * std::max(ncu, 1);
    const int64_t nblocks =
        std::min<int64_t>(std::min<int64_t>(ceil_div(ns, states_per_block), resident), kVIMaxBlocks);
    if (stream) {
        const int64_t W = nblocks * kStreamWaves;
        if (m->vi_wpart_waves != W || m->vi_wpart_spc != grp) {
            if (m->vi_wpart) (void)hipFree(m->vi_wpart);
            m->vi_wpart = nullptr;
            m->vi_wpart_waves = 0;
            if (hipMalloc(&m->vi_wpart, sizeof(uint32_t) * (W + 1)) != hipSuccess) {
                m->vi_wpart = nullptr;
                set_error("vi_solve: device allocation failed");
                return E2E_ERR_NOMEM;
            }
            hipLaunchKernelGGL(vi_partition, dim3((unsigned)ceil_div(W + 1, 256)), dim3(256), 0, s, m->row_ptr,
                               ns, (int)m->na, W, grp, m->vi_wpart);
            E2E_HIP(hipGetLastError());
            m->vi_wpart_waves = W;
            m->vi_wpart_spc = grp;
        }
    }
    VIArgs A{m->row_ptr, m->nz, m->R, m->V[0], m->V[1], m->pi, m->st, m->vi_part, ns, (int)m->na,
             opts.thresh, opts.discount, stream ? m->vi_wpart : nullptr};

    hipLaunchKernelGGL(vi_init, dim3((unsigned)ceil_div(ns, 256)), dim3(256), 0, s, m->V[0], ns, m->st);
    E2E_HIP(hipGetLastError());

    if (opts.use_graph) {
        if (!m->graph || m->graph_chunk != spc || m->graph_fn != (void*)fn || m->graph_blocks != nblocks ||
            m->graph_thresh != opts.thresh || m->graph_discount != opts.discount) {
            if (m->graph) {
                (void)hipGraphExecDestroy(m->graph);
                m->graph = nullptr;
            }
            hipGraph_t gr;
            E2E_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < spc; i++)
                hipLaunchKernelGGL(fn, dim3((unsigned)nblocks), dim3(threads), 0, s, A);
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
            m->graph_blocks = nblocks;
            m->graph_thresh = opts.thresh;
            m->graph_discount = opts.discount;
            m->graph_fn = (void*)fn;
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
            for (int i = 0; i < n; i++) hipLaunchKernelGGL(fn, dim3((unsigned)nblocks), dim3(threads), 0, s, A);
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
        stats->kernel_variant = variant + 1;
        stats->blocks = (int32_t)nblocks;
        stats->solve_ms = now_ms() - t_start;
        stats->sweep_kernel_ms = kms;
        stats->sweep_kernel_ms_per_sweep = sweeps > 0 ? kms / sweeps : 0.0;
        stats->bytes_per_sweep = 8.0 * m->nnz + 4.0 * (m->rows + 1) + 4.0 * m->rows + 12.0 * ns;
        stats->launches = launched;
    }
    return rc;
}

}  // namespace e2e
