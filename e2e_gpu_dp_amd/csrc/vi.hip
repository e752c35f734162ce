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
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "internal.h"
#include "lanes.h"

namespace e2e {

namespace {

struct VIArgs {
    const uint32_t* rp;
    const int2* nz;
    const float* R;
    float* V0;
    float* V1;
    double* W0;  // float64 value buffers of the float64 sweeps (lean, tiled, jagged)
    double* W1;
    int32_t* pi;
    VIState* st;
    float* part;   // [gridDim.x] per-block residual maxima of the running sweep
    int64_t ns;
    int na;
    float thresh;
    float discount;
    const uint32_t* wpart;  // [waves+1] state partition of the lean sweep
    const uint2* tdesc;     // [tiles] (first column, column span) of the tiled kernel
    uint32_t tile_states;   // states per tile (tiled kernel)
    uint32_t vwin;          // LDS value-window capacity in floats (tiled kernel)
    const int2* jnz;        // jagged sweep: entries in per-state jagged-diagonal order
    const uint32_t* jmeta;  // jagged sweep: length | (action << 16) per sorted row
    const float* jR;        // jagged sweep: R per sorted row
    const uint32_t* cz;     // compressed sweep: count << 16 | (column - window base) per entry
    const int32_t* cbase;   // compressed sweep: base column per 256-entry window, -1: kept 8-byte
    double inv_nr;          // compressed sweep: 1 / nrzns
    int ptab_n;             // compressed sweep: nrzns + 1 when the probability table sits in LDS, else 0
    int one_pass;           // lean sweep as one slice of the backward pass: no residual, no sweep state
    uint32_t pre_nt;        // lean sweep over a built model's active slices: nt (0: every state, every sweep)
    uint32_t pre_stride;    // active slices: waves + 1, the stride of wpart's per-prefix partitions
    uint32_t pool_parts;    // lean sweep: tail pieces per wave taken from a per-XCD pool (0: none)
    uint32_t head16;        // lean sweep with pool_parts: sixteenths of each range its wave sweeps itself
    uint32_t pool_min;      // lean sweep: states per wave from which a sweep uses the pool
};

// bytes of VIState the host reads back: k, conv, done, last_res (not the counters)
constexpr size_t kVIStateHead = offsetof(VIState, gdone);
constexpr int kVIThreads = 512;   // 8 waves per block
constexpr int kVIWaves = kVIThreads / 64;
constexpr int kVIMaxBlocks = 2048;  // 8 blocks of 512 threads per CU on 256 CUs
static_assert(kVIMaxBlocks <= kTicketGroup * kTicketGroups, "the lean sweep's ticket groups cover its grid");

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
    auto window = [&](const Meta& M, uint32_t w0, uint32_t wn, const int2 (&e)[U], float& sh, float& slo) {
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
                } else {
                    wst[i] = make_float2(v, x[u]);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (M.valid) {
            const uint32_t jb = max(M.rb, w0), je = min(M.re, w0 + wn);
            uint32_t j = jb;
            for (; j + 4 <= je; j += 4) {
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
            for (; j < je; j++) {
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
        const uint32_t wn0 = min((uint32_t)WIN, ce - cb);
        window(cur, cb, wn0, ecur, sh, slo);
        // the next chunk's first window goes out before this chunk's remaining work
        uint32_t nb, ne;
        range_of(nxt, nb, ne);
        int2 enxt[U];
        issue_nz(nb, ne, enxt);
        for (uint32_t w0 = cb + wn0; w0 < ce; w0 += WIN) {  // chunks larger than a window
            issue_nz(w0, ce, ecur);  // ecur is consumed: reuse its registers
            window(cur, w0, min((uint32_t)WIN, ce - w0), ecur, sh, slo);
        }
        float q;
        {
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

template <typename T>
using CPtr = const __attribute__((address_space(4))) T*;

// p[i] with a 32-bit byte offset (i * sizeof(T) < 2^32)
template <typename T>
__device__ __forceinline__ T ld_off(const T* __restrict__ p, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + (i * (uint32_t)sizeof(T)));
}

// float residual that is never below the double one r >= 0 (so res < thresh in float
// implies the double residual is below thresh too): round to nearest, then one ulp up
// if that fell below r (for a non-negative finite float the next one up is bits + 1)
__device__ __forceinline__ float res_up(double r) {
    const float f = (float)r;
    return (double)f < r ? __uint_as_float(__float_as_uint(f) + 1u) : f;
}

// A lane adds its row's products from LDS kRowReads at a time: all reads issued, one
// wait, then the in-range ones added in ascending order (the same order, and so the
// same bits, as one at a time). Reads past the row, up to kRowReads past the window,
// land in the window's padding or on other rows and are not added.
constexpr int kRowReads = 8;

__device__ __forceinline__ void row_part(const double* q, uint32_t jb, uint32_t je, double& sd) {
    while (jb < je) {
        double d[kRowReads];
#pragma unroll
        for (int i = 0; i < kRowReads; i++) d[i] = q[jb + i];
        const uint32_t c = je - jb;
#pragma unroll
        for (int i = 0; i < kRowReads; i++) sd = ((uint32_t)i < c) ? sd + d[i] : sd;
        jb += kRowReads;
    }
}

constexpr int kStreamThreads = 256;
constexpr int kStreamWaves = kStreamThreads / 64;

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

// Per-slice partitions, balanced as vi_partition over each range: for each slice t
// (blockIdx.y), wpart[t][w] = first state of wave w within [t*nc, (t+1)*nc) (the backward
// pass), or with prefix set within [0, (t+1)*nc) (the sweeps over the active slices);
// bnd (if given) receives the row pointer at every slice boundary, bnd[t] = rp[t*nc*na].
__global__ void vi_partition_slices(const uint32_t* __restrict__ rp, int64_t nc, int na, int64_t W, int spc,
                                    uint32_t* __restrict__ wpart, int prefix, uint32_t* __restrict__ bnd) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w > W) return;
    const int64_t hi = ((int64_t)blockIdx.y + 1) * nc, lo = prefix ? 0 : hi - nc;
    if (bnd && w == W) {
        bnd[blockIdx.y + 1] = rp[hi * na];
        if (blockIdx.y == 0) bnd[0] = rp[0];
    }
    uint32_t* out = wpart + (int64_t)blockIdx.y * (W + 1);
    if (w == 0 || w == W) {
        out[w] = (uint32_t)(w == 0 ? lo : hi);
        return;
    }
    auto weight = [&](int64_t s) { return (double)(rp[s * na] - rp[lo * na]) + 2.0 * (double)((s - lo) * na); };
    const double target = weight(hi) * (double)w / (double)W;
    int64_t a = lo, b = hi;
    while (a < b) {
        const int64_t mid = (a + b) / 2;
        if (weight(mid) >= target)
            b = mid;
        else
            a = mid + 1;
    }
    const int64_t s = lo + (a - lo + spc / 2) / spc * spc;
    out[w] = (uint32_t)(s < hi ? s : hi);
}

// ---------------------------------------------------------------------------------
// Lean streaming sweep (the default). One contiguous state range per wave, balanced by
// entries + rows, its entries read as one stream in windows of 64*U entries; part of each
// range goes to a per-XCD pool of pieces that the XCD's waves take as they finish (the
// kernel below). Built for occupancy: at most 64 VGPRs, so 7 or 8 waves per SIMD keep
// enough loads in flight without software pipelining (a gather-bound stream runs about
// 1.5x faster at 8 waves per SIMD than at 4). Wave-uniform values live in SGPRs
// (readfirstlane, scalar loads of the read-only partition and row pointers). The next
// group's row data is requested when a group starts. Products and sums in float64.
//
// One wave's part of a lean sweep: the states [s_begin, s_end) (s_begin < s_end), their
// entries read as one stream in windows of 64*U entries, NAP lanes per state and 64/NAP
// states per group. Writes Vout, pi of those states and raises res to their residual.
template <int NAP, int U, bool COMP>
__device__ __forceinline__ void lean_range(const VIArgs& A, const double* __restrict__ Vin,
                                           double* __restrict__ Vout, double* wsd, const float* ptabf,
                                           uint32_t s_begin, uint32_t s_end, float& res) {
    constexpr int SPC = 64 / NAP;
    constexpr int WIN = 64 * U;
    // lane is made opaque per call, so that what derives from it is recomputed per range
    // rather than held in registers across the ranges of a wave (64 VGPRs: 8 waves per SIMD)
    uint32_t tx = threadIdx.x;
    asm volatile("" : "+v"(tx));
    const int lane = (int)(tx & 63u);
    const int a = lane % NAP, sl = lane / NAP;
    const uint32_t na = (uint32_t)A.na;
    const CPtr<uint32_t> rps = (CPtr<uint32_t>)A.rp;  // uniform reads of the row pointers
    const uint32_t ngroups = (s_end - s_begin + SPC - 1) / SPC;
    const uint32_t Pb = rps[s_begin * na], Pend = rps[s_end * na];
    // COMP: windows are the stream's aligned 256-entry blocks (one base column each);
    // entries before Pb or from Pend on are staged but belong to no row of this wave
    const uint32_t P0 = COMP ? Pb / WIN * WIN : Pb;
    const uint32_t nwin = Pend > Pb ? (Pend - P0 + WIN - 1) / WIN : 0u;
    struct Meta {
        uint32_t rb, re;
        float rr;
        double vold;
    };
    // row data of group g (clamped to the wave's last group, so always in range)
    auto load_meta = [&](uint32_t g, Meta& M) {
        const uint32_t s = min(s_begin + g * SPC + (uint32_t)sl, s_end - 1u);
        const uint32_t row = s * na + min((uint32_t)a, na - 1u);
        M.rb = ld_off(A.rp, row);
        M.re = ld_off(A.rp, row + 1u);
        M.rr = ld_off(A.R, row);
        M.vold = ld_off(Vin, s);
    };
    Meta cur, nxt;
    uint32_t g = 0;
    load_meta(0, cur);
    load_meta(min(1u, ngroups - 1u), nxt);
    uint32_t ce = rps[min(s_begin + SPC, s_end) * na];
    bool valid = (s_begin + (uint32_t)sl < s_end) && ((uint32_t)a < na);
    double sd = 0.0;

    auto finish_group = [&]() {
        double q = valid ? (double)cur.rr + (double)A.discount * sd : -INFINITY;
        if (q != q) q = -INFINITY;
        int ba = valid ? a : 0x7fffffff;
        // max / first argmax over the NAP lanes of each state; the combine is
        // associative and commutative, so any reduction tree gives the same result
        if (NAP <= 16) {
            // DPP row shifts: after shifts 1, 2, .., NAP/2 the state's last lane
            // (a = NAP-1) has combined exactly its own NAP lanes
#pragma unroll
            for (int o = 1; o < NAP; o <<= 1) {
                const double q2 = dpp_row_shr(q, o, (double)-INFINITY);
                const int a2 = dpp_row_shr(ba, o, 0x7fffffff);
                const bool take = q2 > q || (q2 == q && a2 < ba);
                q = take ? q2 : q;
                ba = take ? a2 : ba;
            }
        } else {
#pragma unroll
            for (int o = 1; o < NAP; o <<= 1) {
                const double q2 = __shfl_xor(q, o, 64);
                const int a2 = __shfl_xor(ba, o, 64);
                if (q2 > q || (q2 == q && a2 < ba)) {
                    q = q2;
                    ba = a2;
                }
            }
        }
        const bool writer = (NAP <= 16) ? (a == NAP - 1) : (a == 0);
        if (writer && g < ngroups && s_begin + g * SPC + (uint32_t)sl < s_end) {
            const uint32_t s = s_begin + g * SPC + (uint32_t)sl;
            double vnew;
            int32_t pol;
            if (q > -1e6) {
                vnew = q;
                pol = ba;
            } else {
                vnew = -1e6;
                pol = -1;
            }
            Vout[s] = vnew;
            A.pi[s] = pol;
            res = fmaxf(res, res_up(fabs(vnew - cur.vold)));
        }
        sd = 0.0;
        g++;
        cur = nxt;
        load_meta(min(g + 1u, ngroups - 1u), nxt);
        const uint32_t s0 = s_begin + g * SPC;
        ce = rps[min(s0 + SPC, s_end) * na];
        valid = (g < ngroups) && (s0 + (uint32_t)sl < s_end) && ((uint32_t)a < na);
    };

    const int2* __restrict__ nzp = A.nz + P0 + lane;
    const CPtr<int32_t> cbs = (CPtr<int32_t>)A.cbase;
    // compressed windows: the next window's entries are requested as soon as this
    // window's value gathers are (ecur holds them when has_next)
    uint32_t ecur[U];
    bool has_next = false;
    auto load_comp = [&](uint32_t w, uint32_t* dst) {
#pragma unroll
        for (int u = 0; u < U; u++) dst[u] = A.cz[w + (uint32_t)(u * 64 + lane)];
    };
    if (COMP && nwin > 0 && cbs[P0 / WIN] >= 0) {
        load_comp(P0, ecur);
        has_next = true;
    }
    for (uint32_t kw = 0; kw < nwin; kw++) {
        const uint32_t w0 = P0 + kw * WIN, wend = w0 + WIN;
        double x[U], pv[U];
        const int32_t base = COMP ? cbs[w0 / WIN] : -1;
        const bool next_comp = COMP && kw + 1 < nwin && cbs[w0 / WIN + 1] >= 0;
        if (COMP && base >= 0) {
            // 4-byte entries w0 + 64 u + lane: the probability count / nrzns is
            // rounded to float from the double count * (1 / nrzns) (the same float as
            // the division: that product is within 2^-52 of count / nrzns, which for
            // count < 2^16 lies at least 2^-41 away from any rounding midpoint of
            // float); the window was kept compressed only if every entry decodes to
            // its stored bits
            uint32_t e[U];
            if (has_next) {
#pragma unroll
                for (int u = 0; u < U; u++) e[u] = ecur[u];
            } else {
                load_comp(w0, e);
            }
#pragma unroll
            for (int u = 0; u < U; u++) x[u] = ld_off(Vin, (uint32_t)base + (e[u] & 0xffffu));
            has_next = next_comp;
            if (next_comp) load_comp(wend, ecur);
            if (A.ptab_n) {
#pragma unroll
                for (int u = 0; u < U; u++) pv[u] = (double)ptabf[e[u] >> 16];
            } else {
#pragma unroll
                for (int u = 0; u < U; u++) pv[u] = (double)(float)((double)(e[u] >> 16) * A.inv_nr);
            }
#pragma unroll
            for (int u = 0; u < U; u++) wsd[u * 64 + lane] = pv[u] * x[u];
        } else {
            int2 e[U];
#pragma unroll
            for (int u = 0; u < U; u++) e[u] = nzp[kw * WIN + u * 64];
#pragma unroll
            for (int u = 0; u < U; u++) x[u] = ld_off(Vin, (uint32_t)e[u].x);
            has_next = next_comp;
            if (next_comp) load_comp(wend, ecur);
#pragma unroll
            for (int u = 0; u < U; u++) pv[u] = (double)__int_as_float(e[u].y);
#pragma unroll
            for (int u = 0; u < U; u++) wsd[u * 64 + lane] = pv[u] * x[u];
        }
        __builtin_amdgcn_wave_barrier();
        for (;;) {
            if (valid) row_part(wsd - w0, max(cur.rb, w0), min(cur.re, wend), sd);
            if (ce > wend) break;  // the group continues in the next window
            finish_group();
            if (g >= ngroups) break;
        }
        __builtin_amdgcn_wave_barrier();
    }
    while (g < ngroups) finish_group();  // trailing groups whose rows are all empty
}

// States [x, y) of one piece of a wave's range [sb, se) of the lean sweep, in whole groups
// of spc states: j < 0 the head, the first head16/16 of the groups; j in [0, P) the j-th
// of P pieces of the rest (the tail), each half the size of the one before (shares
// 2^(P-1-j) / (2^P - 1) of the tail). The head and the P pieces tile [sb, se).
__device__ __forceinline__ uint2 lean_piece(uint32_t sb, uint32_t se, uint32_t spc, uint32_t head16, uint32_t P,
                                            int j) {
    const uint32_t ng = se > sb ? (se - sb + spc - 1u) / spc : 0u;
    const uint32_t hg = ng * head16 / 16u, tg = ng - hg;
    const uint32_t F = (1u << P) - 1u;  // the tail's pieces start at tg * (2^P - 2^(P-j)) / F
    const uint32_t g0 = j < 0 ? 0u : hg + tg * (F + 1u - (1u << (P - (uint32_t)j))) / F;
    const uint32_t g1 = j < 0 ? hg : hg + tg * (F + 1u - (1u << (P - (uint32_t)j - 1u))) / F;
    return make_uint2(min(sb + g0 * spc, se), min(sb + g1 * spc, se));
}

template <int NAP, int U, bool COMP>
// At most 80 SGPRs (the rest spill to VGPR lanes): with more, fewer than 8 waves per SIMD
// are resident on the GPU, below what the occupancy calculation reports for the kernel
__global__ __launch_bounds__(kStreamThreads) __attribute__((amdgpu_num_sgpr(80))) void spmv_vi_lean(const VIArgs A) {
    constexpr int SPC = 64 / NAP;
    static_assert(!COMP || 64 * U == kCompWin, "compressed windows are 256 entries");
    __shared__ float wmax[kStreamWaves];
    __shared__ int is_last;
    __shared__ double stage[kStreamWaves][64 * U + kRowReads];
    VIState* st = A.st;
    if (!A.one_pass && st->conv >= 0) return;  // converged at an earlier sweep
    // float64 values between sweeps. In the backward pass W0 == W1: the launch writes the
    // values of one slice and reads only those of the next
    const uint32_t k = A.one_pass ? 1u : st->k;
    const double* __restrict__ Vin = (k & 1u) ? A.W0 : A.W1;
    double* __restrict__ Vout = (k & 1u) ? A.W1 : A.W0;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned pb = (nb % 8u == 0u) ? (b % 8u) * (nb / 8u) + b / 8u : b;
    const unsigned w = pb * kStreamWaves + wv;
    // a built model's sweep recomputes the active slices only (active_slices), over
    // that prefix's own partition
    const uint32_t* wpart = A.wpart;
    if (A.pre_nt) wpart += (size_t)(active_slices(k, (int)A.pre_nt) - 1) * A.pre_stride;
    const CPtr<uint32_t> wp = (CPtr<uint32_t>)wpart;
    const uint32_t s_begin = wp[w], s_end = wp[w + 1];
    double* wsd = stage[wv];
    float res = 0.0f;
    // compressed entries: the probability of each count, float(count * (1 / nrzns)), once
    // per block in LDS (the same expression as the inline decode, so the same bits)
    // (kept as float: each is a float rounding, widened to double exactly where it is used)
    extern __shared__ float ptabf[];
    if (COMP && A.ptab_n) {
        for (int c = tid; c < A.ptab_n; c += kStreamThreads) ptabf[c] = (float)((double)c * A.inv_nr);
        __syncthreads();
    }

    // Wave speeds differ by up to 1.4x on equal work, and a sweep lasts as long as its
    // slowest wave. So (pool_parts > 0) each wave sweeps only the head of its range; the
    // tails of the ranges of the waves of one XCD (blocks b = x mod 8; their ranges are
    // one contiguous run of states, which keeps a piece's value gathers in that XCD's L2)
    // form a pool of P pieces per wave, which the XCD's waves take from a counter as they
    // finish: every wave's first (largest) piece, then every wave's second, and so on, so
    // that the last pieces taken are the smallest. Every state is still swept once, by one
    // wave, with its entries added in the same order: the same bits whichever wave takes
    // a piece.
    // (only for sweeps of at least pool_min states per wave: a short sweep would spend more
    // on the pool's counters than it saves; every wave reads the same partition ends)
    const uint32_t nwv = nb * kStreamWaves;
    const uint32_t P = (A.one_pass || wp[nwv] - wp[0] < A.pool_min * nwv) ? 0u : A.pool_parts;
    const unsigned xcds = (nb % 8u == 0u) ? 8u : 1u;
    const unsigned x = b % xcds;
    const uint32_t wpx = nb / xcds * kStreamWaves, npool = wpx * P;
    uint2 r = P ? lean_piece(s_begin, s_end, SPC, A.head16, P, -1) : make_uint2(s_begin, s_end);
    for (;;) {
        if (r.x < r.y) lean_range<NAP, U, COMP>(A, Vin, Vout, wsd, ptabf, r.x, r.y, res);
        if (!P) break;
        uint32_t q = 0u;
        if (lane == 0) q = atomicAdd(&st->pool[x * kCounterStride], 1u);
        q = __builtin_amdgcn_readfirstlane(q);
        if (q >= npool) break;
        const uint32_t wq = x * wpx + q % wpx;
        r = lean_piece(wp[wq], wp[wq + 1u], SPC, A.head16, P, (int)(q / wpx));
    }
    if (A.one_pass) return;  // block-uniform

#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) res = fmaxf(res, __shfl_xor(res, o, 64));
    if (lane == 0) wmax[wv] = res;
    __syncthreads();
    if (tid == 0) {
        float bm = 0.0f;
#pragma unroll
        for (int i = 0; i < kStreamWaves; i++) bm = fmaxf(bm, wmax[i]);
        // The partial goes out by an atomic exchange and the last block takes the
        // partials by atomic exchanges (with 0): the same agent-scope atomics as the
        // ticket, whose counts every XCD sees. The exchange returns once performed, and
        // the ticket's increment depends on its returned value, so each block's partial
        // is in place before its ticket counts. (A __threadfence here writes back the
        // XCD's whole L2 in every block: 30 us of a 254 us sweep, 22 of the 41 us of a
        // sweep over two slices.)
        const uint32_t prev = __hip_atomic_exchange(reinterpret_cast<uint32_t*>(A.part) + blockIdx.x,
                                                    __float_as_uint(bm), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t one = 1u;
        asm volatile("" : "+v"(one) : "v"(prev));
        // two-level ticket: the last block of each group of kTicketGroup counts for its
        // group (same-address atomics per sweep: kTicketGroup per group counter and one
        // per group on the sweep's counter, instead of one per block); the second atomic
        // is issued after the first returned (it depends on its value)
        const unsigned grp = blockIdx.x / kTicketGroup, ngrp = (gridDim.x + kTicketGroup - 1) / kTicketGroup;
        const unsigned in_grp = min((unsigned)kTicketGroup, gridDim.x - grp * kTicketGroup);
        int last = 0;
        if (atomicAdd(&st->gdone[grp * kCounterStride], one) == in_grp - 1u) {
            // every block of the group has counted: ready for the next sweep
            __hip_atomic_exchange(&st->gdone[grp * kCounterStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = atomicAdd(&st->done, 1u) == ngrp - 1u;
        }
        is_last = last;
    }
    __syncthreads();
    if (is_last) {  // the last block of the sweep: reduce the partials, publish
        float r = 0.0f;
        for (unsigned i = tid; i < gridDim.x; i += kStreamThreads)
            r = fmaxf(r, __uint_as_float(__hip_atomic_exchange(reinterpret_cast<uint32_t*>(A.part) + i, 0u,
                                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
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
            // the counters the next sweep's atomics count on, reset by atomics too (every
            // wave has left its pool)
            __hip_atomic_exchange(&st->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int i = 0; i < 8; i++)
                __hip_atomic_exchange(&st->pool[i * kCounterStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            st->k = k + 1u;
        }
    }
}

// ---------------------------------------------------------------------------------
// Tiled sweep. One block per tile of tile_states consecutive states;
// each of its 4 waves streams the entries of a quarter of the tile exactly like
// spmv_vi_lean. The difference is where the value gathers go: a 64-lane gather from
// global memory costs the texture addresser about 5x the cycles of a coalesced load of
// the same instruction count, and the gathers of a state go to a narrow band of next
// states. So the block first copies the tile's band of V, V[cmin, cmin + span) (tdesc,
// computed once per model), into LDS with coalesced loads, and the per-entry gathers
// read LDS. A tile whose band is wider than the LDS window (vwin floats) gathers from
// global memory instead. Entries that a window loads past its wave's rows are never
// summed; their band index is clamped so that every LDS read stays in the window.
template <int NAP, int U>
__global__ __launch_bounds__(kStreamThreads) void spmv_vi_tile(const VIArgs A) {
    constexpr int SPC = 64 / NAP;
    constexpr int WIN = 64 * U;
    extern __shared__ double vband[];  // [vwin]
    __shared__ float wmax[kStreamWaves];
    __shared__ int is_last;
    __shared__ double stage[kStreamWaves][WIN + kRowReads];
    VIState* st = A.st;
    if (st->conv >= 0) return;  // converged at an earlier sweep
    const uint32_t k = st->k;
    const double* __restrict__ Vin = (k & 1u) ? A.W0 : A.W1;  // float64 values between sweeps
    double* __restrict__ Vout = (k & 1u) ? A.W1 : A.W0;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int a = lane % NAP, sl = lane / NAP;
    const uint32_t na = (uint32_t)A.na, ns = (uint32_t)A.ns;
    const CPtr<uint32_t> rps = (CPtr<uint32_t>)A.rp;  // uniform reads of the row pointers
    const CPtr<unsigned long long> tdq = (CPtr<unsigned long long>)A.tdesc;
    const uint32_t tile = blockIdx.x;
    const unsigned long long td = tdq[tile];
    const uint32_t cmin = (uint32_t)td, span = (uint32_t)(td >> 32);
    const bool in_lds = span > 0u && span <= A.vwin;  // block-uniform
    if (in_lds) {
        for (uint32_t i = tid; i < span; i += kStreamThreads) vband[i] = ld_off(Vin, cmin + i);
    }
    __syncthreads();
    const uint32_t per_wave = A.tile_states / kStreamWaves;
    const uint32_t s_begin = min(tile * A.tile_states + wv * per_wave, ns);
    const uint32_t s_end = min(s_begin + per_wave, ns);
    double* wsd = stage[wv];
    float res = 0.0f;

    if (s_begin < s_end) {
        const uint32_t ngroups = (s_end - s_begin + SPC - 1) / SPC;
        const uint32_t P0 = rps[s_begin * na], Pend = rps[s_end * na];
        const uint32_t nwin = (Pend - P0 + WIN - 1) / WIN;
        struct Meta {
            uint32_t rb, re;
            float rr;
            double vold;
        };
        auto load_meta = [&](uint32_t g, Meta& M) {
            const uint32_t s = min(s_begin + g * SPC + (uint32_t)sl, s_end - 1u);
            const uint32_t row = s * na + min((uint32_t)a, na - 1u);
            M.rb = ld_off(A.rp, row);
            M.re = ld_off(A.rp, row + 1u);
            M.rr = ld_off(A.R, row);
            M.vold = ld_off(Vin, s);
        };
        Meta cur, nxt;
        uint32_t g = 0;
        load_meta(0, cur);
        load_meta(min(1u, ngroups - 1u), nxt);
        uint32_t ce = rps[min(s_begin + SPC, s_end) * na];
        bool valid = (s_begin + (uint32_t)sl < s_end) && ((uint32_t)a < na);
        double sd = 0.0;

        auto finish_group = [&]() {
            double q = valid ? (double)cur.rr + (double)A.discount * sd : -INFINITY;
            if (q != q) q = -INFINITY;
            int ba = valid ? a : 0x7fffffff;
            if (NAP <= 16) {
#pragma unroll
                for (int o = 1; o < NAP; o <<= 1) {
                    const double q2 = dpp_row_shr(q, o, (double)-INFINITY);
                    const int a2 = dpp_row_shr(ba, o, 0x7fffffff);
                    const bool take = q2 > q || (q2 == q && a2 < ba);
                    q = take ? q2 : q;
                    ba = take ? a2 : ba;
                }
            } else {
#pragma unroll
                for (int o = 1; o < NAP; o <<= 1) {
                    const double q2 = __shfl_xor(q, o, 64);
                    const int a2 = __shfl_xor(ba, o, 64);
                    if (q2 > q || (q2 == q && a2 < ba)) {
                        q = q2;
                        ba = a2;
                    }
                }
            }
            const bool writer = (NAP <= 16) ? (a == NAP - 1) : (a == 0);
            if (writer && g < ngroups && s_begin + g * SPC + (uint32_t)sl < s_end) {
                const uint32_t s = s_begin + g * SPC + (uint32_t)sl;
                double vnew;
                int32_t pol;
                if (q > -1e6) {
                    vnew = q;
                    pol = ba;
                } else {
                    vnew = -1e6;
                    pol = -1;
                }
                Vout[s] = vnew;
                A.pi[s] = pol;
                res = fmaxf(res, res_up(fabs(vnew - cur.vold)));
            }
            sd = 0.0;
            g++;
            cur = nxt;
            load_meta(min(g + 1u, ngroups - 1u), nxt);
            const uint32_t s0 = s_begin + g * SPC;
            ce = rps[min(s0 + SPC, s_end) * na];
            valid = (g < ngroups) && (s0 + (uint32_t)sl < s_end) && ((uint32_t)a < na);
        };

        const int2* __restrict__ nzp = A.nz + P0 + lane;
        for (uint32_t kw = 0; kw < nwin; kw++) {
            const uint32_t w0 = P0 + kw * WIN, wend = w0 + WIN;
            int2 e[U];
#pragma unroll
            for (int u = 0; u < U; u++) e[u] = nzp[kw * WIN + u * 64];
            double x[U];
            if (in_lds) {
#pragma unroll
                for (int u = 0; u < U; u++) x[u] = vband[min((uint32_t)e[u].x - cmin, span - 1u)];
            } else {
#pragma unroll
                for (int u = 0; u < U; u++) x[u] = ld_off(Vin, (uint32_t)e[u].x);
            }
#pragma unroll
            for (int u = 0; u < U; u++) wsd[u * 64 + lane] = (double)__int_as_float(e[u].y) * x[u];
            __builtin_amdgcn_wave_barrier();
            for (;;) {
                if (valid) row_part(wsd - w0, max(cur.rb, w0), min(cur.re, wend), sd);
                if (ce > wend) break;  // the group continues in the next window
                finish_group();
                if (g >= ngroups) break;
            }
            __builtin_amdgcn_wave_barrier();
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

// tdesc[t] = (smallest column, largest - smallest + 1) over the entries of tile t
// (0 span for a tile without entries); *maxspan = the largest span.
__global__ __launch_bounds__(256) void tile_band(const uint32_t* __restrict__ rp, const int2* __restrict__ nz,
                                                 int64_t ns, int na, uint32_t ts, uint2* __restrict__ tdesc,
                                                 unsigned* __restrict__ maxspan) {
    __shared__ uint32_t smin[256], smax[256];
    const int64_t t = blockIdx.x;
    const int64_t s0 = t * ts, s1 = min(s0 + (int64_t)ts, ns);
    const uint32_t b = rp[s0 * na], e = rp[s1 * na];
    uint32_t mn = 0xffffffffu, mx = 0u;
    for (uint32_t j = b + threadIdx.x; j < e; j += 256) {
        const uint32_t c = (uint32_t)nz[j].x;
        mn = min(mn, c);
        mx = max(mx, c);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            smin[threadIdx.x] = min(smin[threadIdx.x], smin[threadIdx.x + o]);
            smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint32_t span = e > b ? smax[0] - smin[0] + 1u : 0u;
        tdesc[t] = make_uint2(e > b ? smin[0] : 0u, span);
        atomicMax(maxspan, span);
    }
}

// ---------------------------------------------------------------------------------
// Jagged sweep (the default). Layout (jds_convert, once per model): within each state
// the na rows are ranked by length, longest first (ties by action), and the entries are
// stored diagonal by diagonal: entry k of every row longer than k, in rank order, then
// entry k+1, ... (a jagged-diagonal / sliced-ELL layout with one slice per state and no
// padding). A state's entries keep their place [rp[s*na], rp[(s+1)*na]).
// Mapping: one lane per row, NAP lanes per state, 64/NAP states per wave, grid-stride
// over such chunks. At diagonal k the rows longer than k are ranks 0..c_k-1, so a lane
// finds its entry at base + sum_{k'<k} c_k' + rank, with c_k counted by a ballot: no
// row pointers per entry, no LDS, and each row's products are added in its own entry
// order in a float64 register (the same bits as the other float64 sweeps). Loads of
// rows shorter than the state's longest are masked, so the only waste is issue slots.
template <int NAP, int UNR>
__global__ __launch_bounds__(kStreamThreads) void spmv_vi_jagged(const VIArgs A) {
    constexpr int SPC = 64 / NAP;
    __shared__ float wmax[kStreamWaves];
    __shared__ int is_last;
    VIState* st = A.st;
    if (st->conv >= 0) return;  // converged at an earlier sweep
    const uint32_t k = st->k;
    const double* __restrict__ Vin = (k & 1u) ? A.W0 : A.W1;  // float64 values between sweeps
    double* __restrict__ Vout = (k & 1u) ? A.W1 : A.W0;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int rank = lane % NAP, sl = lane / NAP;
    const uint32_t na = (uint32_t)A.na, ns = (uint32_t)A.ns;
    const uint64_t gmask = (NAP == 64 ? ~0ull : ((1ull << NAP) - 1ull)) << (sl * NAP);
    const uint32_t nchunks = (ns + SPC - 1) / SPC;
    const uint32_t nwaves = gridDim.x * kStreamWaves;
    float res = 0.0f;
    for (uint32_t ch = blockIdx.x * kStreamWaves + wv; ch < nchunks; ch += nwaves) {
        const uint32_t s = ch * SPC + (uint32_t)sl;
        const bool vs = s < ns;
        const bool vr = vs && (uint32_t)rank < na;
        const uint32_t row = s * na + (uint32_t)rank;
        const uint32_t meta = vr ? ld_off(A.jmeta, row) : 0u;
        const uint32_t len = meta & 0xffffu, act = meta >> 16;
        const float rr = vr ? ld_off(A.jR, row) : 0.0f;
        const uint32_t base = vs ? ld_off(A.rp, s * na) : 0u;
        const double vold = vs ? ld_off(Vin, s) : 0.0;
        double acc = 0.0;
        uint32_t off = 0;
        for (uint32_t kd = 0;; kd += UNR) {
            bool on[UNR];
            uint32_t pos[UNR];
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                on[u] = kd + u < len;
                const uint64_t bm = __ballot(on[u]);
                pos[u] = base + off + (uint32_t)rank;
                off += (uint32_t)__popcll(bm & gmask);
            }
            if (__ballot(on[0]) == 0ull) break;  // wave-uniform: no row has entry kd
            int2 e[UNR];
#pragma unroll
            for (int u = 0; u < UNR; u++) e[u] = on[u] ? ld_off(A.jnz, pos[u]) : make_int2(0, 0);
            double x[UNR];
#pragma unroll
            for (int u = 0; u < UNR; u++) x[u] = on[u] ? ld_off(Vin, (uint32_t)e[u].x) : 0.0;
#pragma unroll
            for (int u = 0; u < UNR; u++)
                if (on[u]) acc += (double)__int_as_float(e[u].y) * x[u];
        }
        double q = vr ? (double)rr + (double)A.discount * acc : -INFINITY;
        if (q != q) q = -INFINITY;
        int ba = vr ? (int)act : 0x7fffffff;
        if (NAP <= 16) {
#pragma unroll
            for (int o = 1; o < NAP; o <<= 1) {
                const double q2 = dpp_row_shr(q, o, (double)-INFINITY);
                const int a2 = dpp_row_shr(ba, o, 0x7fffffff);
                const bool take = q2 > q || (q2 == q && a2 < ba);
                q = take ? q2 : q;
                ba = take ? a2 : ba;
            }
        } else {
#pragma unroll
            for (int o = 1; o < NAP; o <<= 1) {
                const double q2 = __shfl_xor(q, o, 64);
                const int a2 = __shfl_xor(ba, o, 64);
                if (q2 > q || (q2 == q && a2 < ba)) {
                    q = q2;
                    ba = a2;
                }
            }
        }
        const bool writer = (NAP <= 16) ? (rank == NAP - 1) : (rank == 0);
        if (writer && vs) {
            double vnew;
            int32_t pol;
            if (q > -1e6) {
                vnew = q;
                pol = ba;
            } else {
                vnew = -1e6;
                pol = -1;
            }
            Vout[s] = vnew;
            A.pi[s] = pol;
            res = fmaxf(res, res_up(fabs(vnew - vold)));
        }
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

// CSR -> per-state jagged-diagonal layout; one wave per state, one lane per row
// (na <= 64). Rank = place in (length descending, action ascending).
__global__ __launch_bounds__(256) void jds_convert(const uint32_t* __restrict__ rp, const int2* __restrict__ nz,
                                                   const float* __restrict__ R, int64_t ns, int na,
                                                   int2* __restrict__ jnz, uint32_t* __restrict__ jmeta,
                                                   float* __restrict__ jR) {
    const int lane = threadIdx.x & 63;
    const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= ns) return;  // wave-uniform
    const bool valid = lane < na;
    const int64_t row = s * na + lane;
    const uint32_t rb = valid ? rp[row] : 0u;
    const uint32_t len = valid ? rp[row + 1] - rb : 0u;
    uint32_t rank = 0, maxlen = len;
    for (int b = 0; b < na; b++) {
        const uint32_t lb = (uint32_t)__shfl((int)len, b, 64);
        if (b != lane && (lb > len || (lb == len && b < lane))) rank++;
        maxlen = max(maxlen, lb);
    }
    if (valid) {
        jmeta[s * na + rank] = len | ((uint32_t)lane << 16);
        jR[s * na + rank] = R[row];
    }
    const uint32_t base = rp[s * na];
    uint32_t off = 0;
    for (uint32_t kd = 0; kd < maxlen; kd++) {
        const bool on = valid && kd < len;
        const uint32_t c = (uint32_t)__popcll(__ballot(on));
        if (on) jnz[base + off + rank] = nz[rb + kd];
        off += c;
    }
}

// Compressed stream: wavefronts stride over the 256-entry windows of the CSR entries. A
// window is kept at 4 bytes per entry when its columns span less than `span` (at most
// 2^16) and every probability is count / nr for a count < 2^16 that the sweep's decode
// reproduces bit for bit; otherwise its base is -1 and the sweep reads the 8-byte
// entries. Each wavefront adds its count of compressed windows once.
__global__ __launch_bounds__(256) void comp_convert(const int2* __restrict__ nz, int64_t nnz, int64_t nwin,
                                                   int nr, double inv_nr, uint32_t* __restrict__ cz,
                                                   int32_t* __restrict__ cbase, int span,
                                                   unsigned long long* ncomp) {
    const int lane = threadIdx.x & 63;
    unsigned long long kept = 0;
    for (int64_t win = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); win < nwin; win += (int64_t)gridDim.x * 4) {
        int32_t col[4];
        uint32_t cnt[4];
        bool ok = true;
        int32_t cmin = 0x7fffffff, cmax = -1;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t j = win * kCompWin + u * 64 + lane;
            const int2 e = nz[j];  // padded past nnz with zeros
            const bool real = j < nnz;
            const float pf = __int_as_float(e.y);
            const double c = rint((double)pf * (double)nr);
            const bool in16 = c >= 0.0 && c < 65536.0;
            cnt[u] = real && in16 ? (uint32_t)c : 0u;
            ok = ok && (!real || (in16 && __float_as_int((float)((double)cnt[u] * inv_nr)) == e.y));
            col[u] = real ? e.x : -1;
            if (real) {
                cmin = min(cmin, e.x);
                cmax = max(cmax, e.x);
            }
        }
        for (int o = 32; o >= 1; o >>= 1) {
            cmin = min(cmin, __shfl_xor(cmin, o, 64));
            cmax = max(cmax, __shfl_xor(cmax, o, 64));
        }
        ok = __all(ok) && cmax >= 0 && (int64_t)cmax - cmin < span;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t j = win * kCompWin + u * 64 + lane;
            cz[j] = ok && col[u] >= 0 ? (cnt[u] << 16) | (uint32_t)(col[u] - cmin) : 0u;
        }
        if (lane == 0) cbase[win] = ok ? cmin : -1;
        kept += ok ? 1ull : 0ull;
    }
    if (lane == 0 && kept) atomicAdd(ncomp, kept);
}

__global__ void vi_to_f32(const double* __restrict__ W, float* __restrict__ V, int64_t ns) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) V[i] = (float)W[i];  // the f32 value function, rounded once
}

__global__ void vi_init(float* V0, double* W0, int64_t ns, VIState* st) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) {  // PV starts at 0 (spvi_2.cu:433-434,470)
        V0[i] = 0.0f;
        W0[i] = 0.0;
    }
    if (i < kTicketGroups) st->gdone[i * kCounterStride] = 0u;
    if (i < 8) st->pool[i * kCounterStride] = 0u;
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

template <int U, bool COMP = false>
SweepFn pick_lean_nap(int nap) {
    switch (nap) {
        case 1: return spmv_vi_lean<1, U, COMP>;
        case 2: return spmv_vi_lean<2, U, COMP>;
        case 4: return spmv_vi_lean<4, U, COMP>;
        case 8: return spmv_vi_lean<8, U, COMP>;
        case 16: return spmv_vi_lean<16, U, COMP>;
        case 32: return spmv_vi_lean<32, U, COMP>;
        default: return spmv_vi_lean<64, U, COMP>;
    }
}

SweepFn pick_lean(int nap, int u, int* u_used) {
    if (u <= 2) {
        *u_used = 2;
        return pick_lean_nap<2>(nap);
    }
    if (u <= 4) {
        *u_used = 4;
        return pick_lean_nap<4>(nap);
    }
    *u_used = 8;
    return pick_lean_nap<8>(nap);
}

template <int U>
SweepFn pick_tile_nap(int nap) {
    switch (nap) {
        case 1: return spmv_vi_tile<1, U>;
        case 2: return spmv_vi_tile<2, U>;
        case 4: return spmv_vi_tile<4, U>;
        case 8: return spmv_vi_tile<8, U>;
        case 16: return spmv_vi_tile<16, U>;
        case 32: return spmv_vi_tile<32, U>;
        default: return spmv_vi_tile<64, U>;
    }
}

SweepFn pick_tile(int nap, int u, int* u_used) {
    if (u <= 2) {
        *u_used = 2;
        return pick_tile_nap<2>(nap);
    }
    if (u <= 4) {
        *u_used = 4;
        return pick_tile_nap<4>(nap);
    }
    *u_used = 8;
    return pick_tile_nap<8>(nap);
}

template <int UNR>
SweepFn pick_jagged_nap(int nap) {
    switch (nap) {
        case 1: return spmv_vi_jagged<1, UNR>;
        case 2: return spmv_vi_jagged<2, UNR>;
        case 4: return spmv_vi_jagged<4, UNR>;
        case 8: return spmv_vi_jagged<8, UNR>;
        case 16: return spmv_vi_jagged<16, UNR>;
        case 32: return spmv_vi_jagged<32, UNR>;
        default: return spmv_vi_jagged<64, UNR>;
    }
}

SweepFn pick_jagged(int nap, int u, int* u_used) {
    if (u <= 2) {
        *u_used = 2;
        return pick_jagged_nap<2>(nap);
    }
    if (u <= 4) {
        *u_used = 4;
        return pick_jagged_nap<4>(nap);
    }
    *u_used = 8;
    return pick_jagged_nap<8>(nap);
}

// LDS value window of the tiled sweep: at most this many values
constexpr uint32_t kMaxBand = 4096;  // doubles: 32 KB

// Sweep kernels (0-based; the C ABI numbers them from 1):
//   0 chunked, compensated float   (spmv_vi_sweep<.., 0>)
//   1 chunked, fma float           (spmv_vi_sweep<.., 1>; does not meet 1e-5 over long horizons)
//   2 lean streaming, float64      (spmv_vi_lean; the default)
//   3 tiled with an LDS value band (spmv_vi_tile)
//   4 jagged-diagonal, float64     (spmv_vi_jagged)
//   5 lean streaming over the compressed 4-byte entries, float64 (spmv_vi_lean<.., true>;
//     the default for built models whose probabilities are counts over nrzns < 2^16)
//   6 backward pass (built models): one launch of the lean kernel per slice, from the last
//     slice to the first, each reading the values its successor slice wrote
constexpr int kDefaultVariant = 5;
constexpr int kBackward = 6;
constexpr int kDefaultImported = 2;  // imported models: their probabilities are arbitrary
// Lean sweeps: each wave sweeps the first kLeanHead16/16 of its range and leaves the rest,
// in kLeanPoolParts pieces of halving size, to its XCD's pool (spmv_vi_lean). Measured on
// full sweeps at 7 blocks per CU (100x100x120 model, 1.3e8 entries; 200x200x200, 1.0e9
// entries): 289 and 1855 us without the pool; with 4 equal pieces and a head of 6, 8, 10,
// 12 sixteenths 275 and 1691, 268 and 1661, 263 and 1631, 259 and 1682 us; with 4 halving
// pieces after a head of 10 (and at most 80 SGPRs) 255 and 1593 us
constexpr uint32_t kLeanPoolParts = 4;
constexpr uint32_t kLeanHead16 = 10;
// A sweep uses the pool only with at least kLeanPoolMin states per wave: the sweeps over a
// few active slices near the end of a solve otherwise spend more on the counters than
// the pool saves (measured on the 100x100x120 model at 7 blocks per CU: 132, 126 and
// 123.5 us per sweep of the whole solve from 32, 64 and 128 states per wave)
constexpr uint32_t kLeanPoolMin = 128;
constexpr int kLeanBlocksPerCU = 5;  // fastest resident grid of the lean sweep (100x100x120 model)
// Resident blocks per CU of the lean sweep over compressed entries. Measured with the pool
// and the fence-free ticket, full sweeps of the 100x100x120 model: 233.5, 222.6, 223.2 us
// at 6, 7, 8 blocks; of the 200x200x200 model 1629, 1557, 1538 us; the sweeps over the
// active slices of the smaller model 128.3, 126.1, 125.9 us per sweep, of the larger 850,
// 812, 806 us
constexpr int kCompBlocksPerCU = 8;

// Launch shape of the lean sweep over compressed entries: the probability table of the
// counts sits in LDS (float, nrzns + 1 entries) when the blocks per CU wanted (want, or
// kCompBlocksPerCU) still fit with it; else the probabilities are decoded inline (the
// same bits) and the freed LDS keeps the occupancy.
e2e_status comp_launch_shape(SweepFn fn, int nrzns, int want, int dflt, int* bpc, int* ptab_n) {
    const int target = want > 0 ? want : dflt;
    int n = (int64_t)(nrzns + 1) * 4 <= 32768 ? nrzns + 1 : 0;
    int occ = 0;
    if (n) {
        E2E_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)fn, kStreamThreads,
                                                             (size_t)n * sizeof(float)));
        if (occ < target) n = 0;
    }
    if (!n) E2E_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)fn, kStreamThreads, 0));
    *ptab_n = n;
    *bpc = want > 0 ? want : std::max(1, std::min(occ, dflt));
    return E2E_OK;
}
bool is_comp_variant(int variant) { return variant == 5; }
bool is_lean_variant(int variant) { return variant == 2 || variant == 5; }
bool is_tile_variant(int variant) { return variant == 3; }
bool is_jagged_variant(int variant) { return variant == 4; }

SweepFn pick_sweep(int na, int variant, int u, int* u_used) {
    const int nap = next_pow2(na);
    switch (variant) {
        case 0: return pick_u<0>(nap, u, u_used);
        case 1: return pick_u<1>(nap, u, u_used);
        case 2: return pick_lean(nap, u, u_used);
        case 3: return pick_tile(nap, u, u_used);
        case 4: return pick_jagged(nap, u, u_used);
        case 5:
            *u_used = 4;
            return pick_lean_nap<4, true>(nap);
        default: return nullptr;
    }
}

}  // namespace

e2e_status vi_solve(Model* m, const e2e_solve_opts* o, uint32_t* policy, float* value,
                    e2e_solve_stats* stats) {
    const double t_start = now_ms();
    // E2E_TRACE set: host time of each phase to stderr (it waits for the stream)
    const bool trace = std::getenv("E2E_TRACE") != nullptr;
    double t_mark = t_start;
    auto phase = [&](const char* what) {
        if (!trace) return;
        (void)hipStreamSynchronize(m->stream);
        const double t = now_ms();
        std::fprintf(stderr, "[e2e solve] %-28s %9.3f ms\n", what, t - t_mark);
        t_mark = t;
    };
    e2e_solve_opts opts{};
    opts.thresh = 0.1f;  // spvi_2.cu:757
    opts.discount = 1.0f;  // spvi_2.cu:61
    if (o) opts = *o;
    const int64_t ns = m->ns;
    int max_sweeps = opts.max_sweeps > 0 ? opts.max_sweeps : (m->built ? (int)m->nt + 2 : 100000);
    const int spc = opts.sweeps_per_check > 0 ? opts.sweeps_per_check : 16;
    hipStream_t s = m->stream;
    if (!m->V[0]) {
        if (pool_malloc(&m->V[0], sizeof(float) * ns) != hipSuccess ||
            pool_malloc(&m->W[0], sizeof(double) * ns) != hipSuccess ||
            pool_malloc(&m->W[1], sizeof(double) * ns) != hipSuccess ||
            pool_malloc(&m->V[1], sizeof(float) * ns) != hipSuccess ||
            pool_malloc(&m->pi, sizeof(int32_t) * ns) != hipSuccess ||
            pool_malloc(&m->st, sizeof(VIState)) != hipSuccess ||
            pool_malloc(&m->vi_part, sizeof(float) * kVIMaxBlocks) != hipSuccess) {
            void** bufs[] = {(void**)&m->V[0], (void**)&m->W[0], (void**)&m->W[1], (void**)&m->V[1],
                             (void**)&m->pi,   (void**)&m->st,   (void**)&m->vi_part};
            for (void** b : bufs) {  // all or nothing, so a later call allocates again
                if (*b) pool_free(*b);
                *b = nullptr;
            }
            set_error("vi_solve: device allocation failed");
            return E2E_ERR_NOMEM;
        }
        m->vi_part_cap = kVIMaxBlocks;
    }
    const int nap = next_pow2((int)m->na);
    const int grp = 64 / nap;
    const bool comp_ok = m->built && m->p.nrzns > 0 && m->p.nrzns < 65536;
    const int variant_req = opts.kernel_variant > 0 ? opts.kernel_variant - 1
                            : (comp_ok ? kDefaultVariant : kDefaultImported);
    // the backward pass runs the default lean kernel once per slice; its transitions must
    // go from each slice to the next, which holds for built models
    const bool bw = variant_req == kBackward;
    E2E_CHECK(!bw || (m->built && m->nc > 0 && m->nt > 0), E2E_ERR_INVALID,
              "vi_solve: the backward pass (kernel_variant 7) needs a model built by e2e_model_build");
    const int variant = bw ? (comp_ok ? kDefaultVariant : kDefaultImported) : variant_req;
    const bool comp = is_comp_variant(variant);
    E2E_CHECK(!comp || comp_ok, E2E_ERR_INVALID,
              "vi_solve: the compressed sweep needs a built model with nrzns below 65536");
    const bool lean = is_lean_variant(variant);
    const bool tiled = is_tile_variant(variant);
    const bool jagged = is_jagged_variant(variant);
    // loads in flight per lane. Chunked kernels: a chunk holds about 64/NAP states of
    // nnz/row * na entries; the other kernels: 4 (windows of 256 entries).
    const double per_chunk = (m->rows ? (double)m->nnz / (double)m->rows : 1.0) * m->na * grp;
    int u_req = opts.loads_in_flight > 0 ? opts.loads_in_flight
                : (lean || tiled || jagged) ? 4 : (per_chunk > 192 ? 8 : per_chunk > 96 ? 4 : 2);
    int lpr = 0;
    SweepFn fn = pick_sweep((int)m->na, variant, u_req, &lpr);
    E2E_CHECK(fn != nullptr, E2E_ERR_INVALID, "vi_solve: unknown kernel_variant %d", opts.kernel_variant);
    const int threads = (lean || tiled || jagged) ? kStreamThreads : kVIThreads;
    if (jagged) {
        E2E_CHECK(ns * (int64_t)sizeof(double) < (1ll << 32) && (m->rows + 1) * 4 < (1ll << 32) &&
                      (m->nnz + kNzPad) * 8 < (1ll << 32),
                  E2E_ERR_CAPACITY, "vi_solve: the jagged sweep addresses entries with 32-bit byte offsets");
        E2E_CHECK(m->na <= 64 && m->info.max_row_nnz < 65536, E2E_ERR_CAPACITY,
                  "vi_solve: the jagged sweep needs at most 64 actions and rows shorter than 65536");
        if (!m->vi_jnz) {
            if (pool_malloc(&m->vi_jnz, sizeof(int2) * (m->nnz + kNzPad)) != hipSuccess ||
                pool_malloc(&m->vi_jmeta, sizeof(uint32_t) * m->rows) != hipSuccess ||
                pool_malloc(&m->vi_jR, sizeof(float) * m->rows) != hipSuccess) {
                set_error("vi_solve: device allocation of the jagged layout failed");
                return E2E_ERR_NOMEM;
            }
            E2E_HIP(hipMemsetAsync(m->vi_jnz + m->nnz, 0, sizeof(int2) * kNzPad, s));
            hipLaunchKernelGGL(jds_convert, dim3((unsigned)ceil_div(ns, 4)), dim3(256), 0, s, m->row_ptr, m->nz,
                               m->R, ns, (int)m->na, m->vi_jnz, m->vi_jmeta, m->vi_jR);
            E2E_HIP(hipGetLastError());
        }
    }
    // tiled sweep: one block per tile of 16 state groups per wave, an LDS window sized to
    // the widest column band (once per model)
    uint32_t vwin = 0, tile_states = 0;
    int64_t ntiles = 0;
    if (tiled) {
        E2E_CHECK(ns * (int64_t)sizeof(double) < (1ll << 32) && (m->rows + 1) * 4 < (1ll << 32), E2E_ERR_CAPACITY,
                  "vi_solve: the tiled sweep addresses states and rows with 32-bit byte offsets");
        tile_states = (uint32_t)(16 * kStreamWaves * grp);
        ntiles = ceil_div(ns, (int64_t)tile_states);
        if (m->vi_tile_states != tile_states) {
            if (m->vi_tdesc) pool_free(m->vi_tdesc);
            m->vi_tdesc = nullptr;
            m->vi_tile_states = 0;
            if (pool_malloc(&m->vi_tdesc, sizeof(uint2) * ntiles) != hipSuccess) {
                m->vi_tdesc = nullptr;
                set_error("vi_solve: device allocation failed");
                return E2E_ERR_NOMEM;
            }
            DevBuf mx;
            E2E_TRY(dmalloc(mx, sizeof(unsigned)));
            E2E_HIP(hipMemsetAsync(mx.p, 0, sizeof(unsigned), s));
            hipLaunchKernelGGL(tile_band, dim3((unsigned)ntiles), dim3(256), 0, s, m->row_ptr, m->nz, ns, (int)m->na,
                               tile_states, m->vi_tdesc, mx.as<unsigned>());
            E2E_HIP(hipGetLastError());
            unsigned hmax = 0;
            E2E_HIP(hipMemcpyAsync(&hmax, mx.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
            E2E_HIP(hipStreamSynchronize(s));
            m->vi_maxspan = hmax;
            m->vi_tile_states = tile_states;
        }
        vwin = (uint32_t)std::min<uint64_t>((m->vi_maxspan + 63u) / 64u * 64u, kMaxBand);
        if (vwin == 0) vwin = 64;
        if (m->vi_part_cap < ntiles) {
            pool_free(m->vi_part);
            m->vi_part = nullptr;
            m->vi_part_cap = 0;
            if (pool_malloc(&m->vi_part, sizeof(float) * ntiles) != hipSuccess) {
                m->vi_part = nullptr;
                set_error("vi_solve: device allocation failed");
                return E2E_ERR_NOMEM;
            }
            m->vi_part_cap = ntiles;
        }
    }
    // persistent grid: as many blocks as can be resident at once (grid-stride over chunks,
    // or one contiguous share per wave), so that no block starts after the others finished;
    // the compressed sweep keeps its probability table in LDS if that leaves its occupancy
    int ptab_n = 0;
    int bpc = opts.blocks_per_cu;
    if (lean && comp) {
        E2E_TRY(comp_launch_shape(fn, m->p.nrzns, opts.blocks_per_cu, kCompBlocksPerCU, &bpc, &ptab_n));
    } else if (bpc <= 0) {
        E2E_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, (const void*)fn, threads,
                                                             tiled ? (size_t)vwin * sizeof(double) : 0));
        if (lean) bpc = std::min(bpc, kLeanBlocksPerCU);
    }
    const size_t shmem = tiled ? (size_t)vwin * sizeof(double) : (size_t)ptab_n * sizeof(float);
    const int64_t states_per_block = (threads / 64) * grp;
    int ncu = 0, dev = 0;
    E2E_HIP(hipGetDevice(&dev));
    E2E_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t resident = (int64_t)std::max(bpc, 1) * std::max(ncu, 1);
    const int64_t nblocks =
        tiled ? ntiles
              : std::min<int64_t>(std::min<int64_t>(ceil_div(ns, states_per_block), resident), kVIMaxBlocks);
    const int span = opts.comp_span > 0 ? opts.comp_span : kCompSpan;
    E2E_CHECK(span <= kCompSpan, E2E_ERR_INVALID, "vi_solve: comp_span must be at most 65536 (got %d)", opts.comp_span);
    if (comp && m->vi_cz && m->vi_comp_span != span) {  // written or converted with another span
        pool_free(m->vi_cz);
        pool_free(m->vi_cbase);
        m->vi_cz = nullptr;
        m->vi_cbase = nullptr;
    }
    if (comp && !m->vi_cz) {  // the build wrote none for this span: convert the CSR entries
        const int64_t nwin = ceil_div(m->nnz, (int64_t)kCompWin);
        if (pool_malloc(&m->vi_cz, sizeof(uint32_t) * (size_t)std::max<int64_t>(nwin, 1) * kCompWin) != hipSuccess ||
            pool_malloc(&m->vi_cbase, sizeof(int32_t) * (size_t)std::max<int64_t>(nwin, 1)) != hipSuccess) {
            if (m->vi_cz) pool_free(m->vi_cz);
            m->vi_cz = nullptr;
            m->vi_cbase = nullptr;
            set_error("vi_solve: device allocation of the compressed stream failed");
            return E2E_ERR_NOMEM;
        }
        m->vi_cwin = nwin;
        m->vi_cwin_comp = 0;
        m->vi_comp_span = span;
        if (nwin > 0) {
            DevBuf cnt;
            E2E_TRY(dmalloc(cnt, sizeof(unsigned long long)));
            E2E_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), s));
            const unsigned cblocks = (unsigned)std::min<int64_t>(ceil_div(nwin, 4), 4096);
            hipLaunchKernelGGL(comp_convert, dim3(cblocks), dim3(256), 0, s, m->nz, m->nnz, nwin,
                               m->p.nrzns, 1.0 / (double)m->p.nrzns, m->vi_cz, m->vi_cbase, span,
                               cnt.as<unsigned long long>());
            E2E_HIP(hipGetLastError());
            unsigned long long h = 0;
            E2E_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
            E2E_HIP(hipStreamSynchronize(s));
            m->vi_cwin_comp = (int64_t)h;
        }
    }
    // sweeps over the active slices only (spmv_vi_lean): built models, whose transitions
    // go from each slice to the next; not the backward pass, which has its own schedule
    const bool active = lean && !bw && !opts.full_sweeps && m->built && m->nt > 0 && m->nc > 0 &&
                        ns == m->nt * m->nc;
    if (lean) {  // (entries are addressed through 64-bit pointers)
        E2E_CHECK(ns * (int64_t)sizeof(double) < (1ll << 32) && (m->rows + 1) * 4 < (1ll << 32), E2E_ERR_CAPACITY,
                  "vi_solve: the lean sweep addresses states and rows with 32-bit byte offsets");
        const int64_t W = nblocks * kStreamWaves;
        if (m->vi_wpart_waves != W || m->vi_wpart_spc != grp) {
            if (m->vi_wpart) pool_free(m->vi_wpart);
            m->vi_wpart = nullptr;
            m->vi_wpart_waves = 0;
            if (pool_malloc(&m->vi_wpart, sizeof(uint32_t) * (W + 1)) != hipSuccess) {
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
        if (active) {  // the partitions of every prefix of slices, once per model and grid
            if (m->vi_pre_waves != W || m->vi_pre_spc != grp) {
                if (m->vi_pre_wpart) pool_free(m->vi_pre_wpart);
                m->vi_pre_wpart = nullptr;
                m->vi_pre_waves = 0;
                // + the row pointer at each slice boundary (nt + 1), for the bytes of each sweep
                if (pool_malloc(&m->vi_pre_wpart, sizeof(uint32_t) * ((W + 1) * m->nt + m->nt + 1)) != hipSuccess) {
                    m->vi_pre_wpart = nullptr;
                    set_error("vi_solve: device allocation failed");
                    return E2E_ERR_NOMEM;
                }
                hipLaunchKernelGGL(vi_partition_slices, dim3((unsigned)ceil_div(W + 1, 256), (unsigned)m->nt),
                                   dim3(256), 0, s, m->row_ptr, m->nc, (int)m->na, W, grp, m->vi_pre_wpart, 1,
                                   m->vi_pre_wpart + (W + 1) * m->nt);
                E2E_HIP(hipGetLastError());
                m->vi_pre_waves = W;
                m->vi_pre_spc = grp;
                m->slice_rp.clear();
            }
        }
    }
    phase("buffers, layouts, partitions");
    VIArgs A;
    std::memset(&A, 0, sizeof(A));  // padding too: the graph cache compares the bytes
    A.rp = m->row_ptr;
    A.nz = m->nz;
    A.R = m->R;
    A.V0 = m->V[0];
    A.V1 = m->V[1];
    A.W0 = m->W[0];
    A.W1 = m->W[1];
    A.pi = m->pi;
    A.st = m->st;
    A.part = m->vi_part;
    A.ns = ns;
    A.na = (int)m->na;
    A.thresh = opts.thresh;
    A.discount = opts.discount;
    A.wpart = lean ? (active ? m->vi_pre_wpart : m->vi_wpart) : nullptr;
    A.pre_nt = active ? (uint32_t)m->nt : 0u;
    A.pre_stride = active ? (uint32_t)(nblocks * kStreamWaves + 1) : 0u;
    A.pool_parts = lean ? kLeanPoolParts : 0u;
    A.head16 = kLeanHead16;
    A.pool_min = kLeanPoolMin;
    A.tdesc = tiled ? m->vi_tdesc : nullptr;
    A.tile_states = tile_states;
    A.vwin = vwin;
    A.jnz = jagged ? m->vi_jnz : nullptr;
    A.jmeta = jagged ? m->vi_jmeta : nullptr;
    A.jR = jagged ? m->vi_jR : nullptr;
    A.cz = comp ? m->vi_cz : nullptr;
    A.cbase = comp ? m->vi_cbase : nullptr;
    A.inv_nr = comp ? 1.0 / (double)m->p.nrzns : 0.0;
    A.ptab_n = ptab_n;
    A.one_pass = 0;

    if (bw) {
        // one launch per slice, last to first: slice t's rows read only slice t+1, which the
        // previous launch wrote; every slice takes its final values in this one pass
        const int64_t nc = m->nc, nts = m->nt;
        const int64_t groups = ceil_div(nc, (int64_t)grp);
        const int64_t nbs = std::max<int64_t>(1, std::min<int64_t>(ceil_div(groups, (int64_t)kStreamWaves), nblocks));
        const int64_t Ws = nbs * kStreamWaves;
        if (m->vi_bw_waves != Ws || m->vi_bw_spc != grp) {
            if (m->vi_bw_wpart) pool_free(m->vi_bw_wpart);
            m->vi_bw_wpart = nullptr;
            m->vi_bw_waves = 0;
            if (pool_malloc(&m->vi_bw_wpart, sizeof(uint32_t) * (Ws + 1) * nts) != hipSuccess) {
                m->vi_bw_wpart = nullptr;
                set_error("vi_solve: device allocation failed");
                return E2E_ERR_NOMEM;
            }
            hipLaunchKernelGGL(vi_partition_slices, dim3((unsigned)ceil_div(Ws + 1, 256), (unsigned)nts), dim3(256), 0,
                               s, m->row_ptr, nc, (int)m->na, Ws, grp, m->vi_bw_wpart, 0, nullptr);
            E2E_HIP(hipGetLastError());
            m->vi_bw_waves = Ws;
            m->vi_bw_spc = grp;
        }
        VIArgs B = A;
        B.one_pass = 1;
        B.W0 = B.W1 = m->W[0];
        hipEvent_t e0, e1;
        E2E_HIP(hipEventCreate(&e0));
        E2E_HIP(hipEventCreate(&e1));
        E2E_HIP(hipEventRecord(e0, s));
        for (int64_t t = nts - 1; t >= 0; t--) {
            B.wpart = m->vi_bw_wpart + t * (Ws + 1);
            hipLaunchKernelGGL(fn, dim3((unsigned)nbs), dim3(threads), shmem, s, B);
        }
        E2E_HIP(hipGetLastError());
        E2E_HIP(hipEventRecord(e1, s));
        hipLaunchKernelGGL(vi_to_f32, dim3((unsigned)ceil_div(ns, 256)), dim3(256), 0, s, m->W[0], m->V[0], ns);
        E2E_HIP(hipGetLastError());
        if (value) E2E_HIP(hipMemcpyAsync(value, m->V[0], sizeof(float) * ns, hipMemcpyDeviceToHost, s));
        if (policy) E2E_HIP(hipMemcpyAsync(policy, m->pi, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, s));
        E2E_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        E2E_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (stats) {
            *stats = e2e_solve_stats{};
            stats->sweeps = 1;
            stats->converged = 1;
            stats->final_residual = 0.0f;
            stats->loads_in_flight = lpr;
            stats->kernel_variant = kBackward + 1;
            stats->blocks = (int32_t)nbs;
            stats->solve_ms = now_ms() - t_start;
            stats->sweep_kernel_ms = ms;
            stats->sweep_kernel_ms_per_sweep = ms;
            const double meta = 4.0 * (m->rows + 1) + 4.0 * m->rows;
            stats->bytes_per_sweep = 8.0 * m->nnz + meta + 12.0 * ns;
            double entries = 8.0 * m->nnz;
            if (comp) {
                const double cw = (double)m->vi_cwin_comp, ww = (double)(m->vi_cwin - m->vi_cwin_comp);
                entries = cw * (4.0 * kCompWin + 4.0) + ww * (8.0 * kCompWin + 4.0);
            }
            stats->moved_bytes_per_sweep = entries + meta + 20.0 * ns;
            stats->compressed_windows = comp ? (int32_t)std::min<int64_t>(m->vi_cwin_comp, INT32_MAX) : 0;
            stats->launches = (int32_t)nts;
        }
        return E2E_OK;
    }

    hipLaunchKernelGGL(vi_init, dim3((unsigned)ceil_div(ns, 256)), dim3(256), 0, s, m->V[0], m->W[0], ns, m->st);
    E2E_HIP(hipGetLastError());

    if (opts.use_graph) {
        static_assert(sizeof(VIArgs) <= sizeof(m->graph_args), "graph key");
        if (!m->graph || m->graph_chunk != spc || m->graph_fn != (void*)fn || m->graph_blocks != nblocks ||
            std::memcmp(m->graph_args, &A, sizeof(A)) != 0) {
            if (m->graph) {
                (void)hipGraphExecDestroy(m->graph);
                m->graph = nullptr;
            }
            hipGraph_t gr;
            E2E_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < spc; i++)
                hipLaunchKernelGGL(fn, dim3((unsigned)nblocks), dim3(threads), shmem, s, A);
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
            std::memcpy(m->graph_args, &A, sizeof(A));
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
    // host checks after 16, 32, 64, 128, ... sweeps (chunks of 16, 16, 32, 64 and then 64)
    // unless sweeps_per_check fixes the chunk: fewer waits, and launches after convergence
    // return at once
    int chunk = spc;
    while (launched < max_sweeps) {
        int n = std::min(chunk, max_sweeps - launched);
        E2E_HIP(hipEventRecord(e0, s));
        if (opts.use_graph && n % spc == 0) {
            for (int r = 0; r < n / spc; r++) E2E_HIP(hipGraphLaunch(m->graph, s));
        } else {
            for (int i = 0; i < n; i++) hipLaunchKernelGGL(fn, dim3((unsigned)nblocks), dim3(threads), shmem, s, A);
            E2E_HIP(hipGetLastError());
        }
        E2E_HIP(hipEventRecord(e1, s));
        E2E_HIP(hipMemcpyAsync(&hst, m->st, kVIStateHead, hipMemcpyDeviceToHost, s));
        E2E_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        E2E_HIP(hipEventElapsedTime(&ms, e0, e1));
        kms += ms;
        launched += n;
        if (opts.sweeps_per_check <= 0 && launched >= 2 * spc) chunk = std::min(chunk * 2, 4 * spc);
        if (hst.conv >= 0) break;
        if (opts.timeout_ms > 0 && now_ms() - t_start > opts.timeout_ms) {
            rc = E2E_ERR_TIMEOUT;
            break;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    phase("sweeps");
    const int sweeps = hst.conv >= 0 ? hst.conv : (int)hst.k - 1;
    if (rc == E2E_OK && hst.conv < 0) rc = E2E_ERR_NOT_CONVERGED;
    if (rc == E2E_ERR_NOT_CONVERGED) set_error("not converged after %d sweeps (residual %g)", sweeps, hst.last_res);
    if (rc == E2E_ERR_TIMEOUT) set_error("timeout after %d sweeps", sweeps);
    const bool f64 = lean || tiled || jagged;  // values carried in float64 between sweeps
    const float* Vfin = (sweeps & 1) ? m->V[1] : m->V[0];
    if (f64) {
        hipLaunchKernelGGL(vi_to_f32, dim3((unsigned)ceil_div(ns, 256)), dim3(256), 0, s,
                           (sweeps & 1) ? m->W[1] : m->W[0], m->V[0], ns);
        E2E_HIP(hipGetLastError());
        Vfin = m->V[0];
    }
    if (value) E2E_HIP(hipMemcpyAsync(value, Vfin, sizeof(float) * ns, hipMemcpyDeviceToHost, s));
    if (policy) E2E_HIP(hipMemcpyAsync(policy, m->pi, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, s));
    if (stats && active && (int64_t)m->slice_rp.size() != m->nt + 1) {  // once per model
        m->slice_rp.assign((size_t)m->nt + 1, 0u);
        E2E_HIP(hipMemcpyAsync(m->slice_rp.data(), m->vi_pre_wpart + (m->vi_pre_waves + 1) * m->nt,
                               sizeof(uint32_t) * (m->nt + 1), hipMemcpyDeviceToHost, s));
    }
    E2E_HIP(hipStreamSynchronize(s));
    phase("results to the host");
    if (stats) {
        stats->sweeps = sweeps;
        stats->converged = hst.conv >= 0;
        // entries, row pointers, R, V read + V write + policy write
        const double meta = 4.0 * (m->rows + 1) + 4.0 * m->rows;
        double bytes = 8.0 * m->nnz + meta + 12.0 * ns;
        double entries = 8.0 * m->nnz;
        if (comp) {  // whole aligned windows: 4 bytes per entry + base, 8 for wide windows
            const double cw = (double)m->vi_cwin_comp, ww = (double)(m->vi_cwin - m->vi_cwin_comp);
            entries = cw * (4.0 * kCompWin + 4.0) + ww * (8.0 * kCompWin + 4.0);
        }
        double moved = entries + meta + (f64 ? 20.0 : 12.0) * ns;
        if (active && sweeps > 0) {  // the mean over the sweeps of the slices each one ran
            double b = 0.0, mv = 0.0;
            for (int k = 1; k <= sweeps; k++) {
                const int64_t T = active_slices((uint32_t)k, (int)m->nt);
                const double nnz_t = (double)m->slice_rp[(size_t)T];
                b += prefix_sweep_bytes(m, T, 12.0);
                mv += prefix_sweep_bytes(m, T, 20.0) - 8.0 * nnz_t + (m->nnz ? entries * nnz_t / (double)m->nnz : 0.0);
            }
            bytes = b / sweeps;
            moved = mv / sweeps;
        }
        stats->final_residual = hst.last_res;
        stats->loads_in_flight = lpr;
        stats->kernel_variant = variant + 1;
        stats->blocks = (int32_t)nblocks;
        stats->solve_ms = now_ms() - t_start;
        stats->sweep_kernel_ms = kms;
        stats->sweep_kernel_ms_per_sweep = sweeps > 0 ? kms / sweeps : 0.0;
        stats->bytes_per_sweep = bytes;
        stats->moved_bytes_per_sweep = moved;
        stats->compressed_windows = comp ? (int32_t)std::min<int64_t>(m->vi_cwin_comp, INT32_MAX) : 0;
        stats->launches = launched;
    }
    return rc;
}

// wpart[0..W] of the lean state partition (vi_partition) for W waves and groups of spc
// states, on stream s; shared with the batched streaming sweep
e2e_status fetch_slice_row_ptrs(Model* m, hipStream_t s) {
    if ((int64_t)m->slice_rp.size() == m->nt + 1) return E2E_OK;
    std::vector<uint32_t> h((size_t)m->nt + 1);
    const size_t pitch = sizeof(uint32_t) * (size_t)(m->nc * m->na);
    E2E_HIP(hipMemcpy2DAsync(h.data(), sizeof(uint32_t), m->row_ptr, pitch, sizeof(uint32_t), (size_t)m->nt + 1,
                             hipMemcpyDeviceToHost, s));
    E2E_HIP(hipStreamSynchronize(s));
    m->slice_rp = std::move(h);
    return E2E_OK;
}

double prefix_sweep_bytes(const Model* m, int64_t T, double per_state) {
    const double nnz_t = (double)m->slice_rp[(size_t)T], rows_t = (double)(T * m->nc * m->na);
    return 8.0 * nnz_t + 4.0 * (rows_t + 1) + 4.0 * rows_t + per_state * (double)(T * m->nc);
}

void launch_prefix_partitions(const uint32_t* rp, int64_t nc, int na, int64_t nt, int64_t W, int spc,
                              uint32_t* wpart, hipStream_t s) {
    hipLaunchKernelGGL(vi_partition_slices, dim3((unsigned)ceil_div(W + 1, 256), (unsigned)nt), dim3(256), 0, s, rp,
                       nc, na, W, spc, wpart, 1, nullptr);
}

void launch_wave_partition(const uint32_t* rp, int64_t ns, int na, int64_t W, int spc, uint32_t* wpart,
                           hipStream_t s) {
    hipLaunchKernelGGL(vi_partition, dim3((unsigned)ceil_div(W + 1, 256)), dim3(256), 0, s, rp, ns, na, W, spc,
                       wpart);
}


// ---------------------------------------------------------------------------------
// Solve without device allocation (e2e_vi_solve_noalloc). The sweeps of vi_solve keep
// their buffers in the model (allocated on its first solve). This variant keeps every
// device buffer beyond the model and the host outputs in one file-scope array per GPU,
// laid out as
//   VIState | residual partials [kVIMaxBlocks] | wave partitions | W0, W1 [ns] double |
//   pi [ns] int32 | V [ns] float
// The array lives as long as the process; each call reuses it and clears what must
// start at zero (vi_init: the first value buffer, the output vector, the sweep state).
// Calls that overlap on one GPU share it and race.
namespace {

constexpr int64_t kNaMaxStates = int64_t(1) << 21;  // 2,097,152 states
constexpr int64_t kNaMaxSlices = 256;               // slices of a built model (active-slice partitions)
constexpr int64_t kNaMaxWaves = (int64_t)kVIMaxBlocks * kStreamWaves;
constexpr size_t na_align(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t kNaOffPart = na_align(sizeof(VIState));
constexpr size_t kNaOffWpart = na_align(kNaOffPart + sizeof(float) * kVIMaxBlocks);
constexpr size_t kNaWpartWords = (size_t)(kNaMaxWaves + 1) * kNaMaxSlices + kNaMaxSlices + 1;
constexpr size_t kNaOffW0 = na_align(kNaOffWpart + sizeof(uint32_t) * kNaWpartWords);
constexpr size_t kNaOffW1 = kNaOffW0 + sizeof(double) * kNaMaxStates;
constexpr size_t kNaOffPi = kNaOffW1 + sizeof(double) * kNaMaxStates;
constexpr size_t kNaOffV = kNaOffPi + sizeof(int32_t) * kNaMaxStates;
constexpr size_t kNaScratchBytes = kNaOffV + sizeof(float) * kNaMaxStates;

// Host code never uses this name as an address: the device address of each GPU's copy
// comes from hipGetSymbolAddress (noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
unsigned char* g_na_scratch[kNaMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
e2e_status noalloc_scratch(unsigned char** out) {
    int dev = -1;
    E2E_HIP(hipGetDevice(&dev));
    E2E_CHECK(dev >= 0 && dev < kNaMaxDevices, E2E_ERR_CAPACITY, "e2e_vi_solve_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr) {
            set_error("e2e_vi_solve_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                      hipGetErrorString(e));
            return E2E_ERR_HIP;
        }
        g_na_scratch[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_na_scratch[dev];
    return E2E_OK;
}

}  // namespace

e2e_status vi_solve_noalloc(Model* m, const e2e_solve_opts* o, uint32_t* policy, float* value,
                            e2e_solve_stats* stats) {
    const double t_start = now_ms();
    e2e_solve_opts opts{};
    opts.thresh = 0.1f;    // spvi_2.cu:757
    opts.discount = 1.0f;  // spvi_2.cu:61
    if (o) opts = *o;
    const int64_t ns = m->ns;
    E2E_CHECK(ns >= 1 && ns <= kNaMaxStates, E2E_ERR_CAPACITY,
              "e2e_vi_solve_noalloc: %lld states; at most %lld fit its fixed device scratch", (long long)ns,
              (long long)kNaMaxStates);
    E2E_CHECK(!opts.use_graph, E2E_ERR_INVALID, "e2e_vi_solve_noalloc: use_graph must be 0 (a graph is allocated)");
    // the lean float64 sweep over the compressed entries when the model has them (a build
    // with nrzns < 65536 makes them), else over the 8-byte entries: the same bits
    const bool has_comp = m->vi_cz != nullptr && m->vi_cbase != nullptr;
    int variant;
    if (opts.kernel_variant == 0)
        variant = has_comp ? 5 : 2;
    else if (opts.kernel_variant == 3)
        variant = 2;
    else if (opts.kernel_variant == 6) {
        E2E_CHECK(has_comp, E2E_ERR_INVALID,
                  "e2e_vi_solve_noalloc: kernel_variant 6 needs the compressed entries of a built model with "
                  "nrzns below 65536");
        variant = 5;
    } else {
        set_error("e2e_vi_solve_noalloc: kernel_variant %d (0, 3 and 6 run without allocation)", opts.kernel_variant);
        return E2E_ERR_INVALID;
    }
    const bool comp = variant == 5;
    E2E_CHECK(ns * (int64_t)sizeof(double) < (1ll << 32) && (m->rows + 1) * 4 < (1ll << 32), E2E_ERR_CAPACITY,
              "e2e_vi_solve_noalloc: the lean sweep addresses states and rows with 32-bit byte offsets");
    const bool active = !opts.full_sweeps && m->built && m->nt > 0 && m->nc > 0 && ns == m->nt * m->nc;
    E2E_CHECK(!active || m->nt <= kNaMaxSlices, E2E_ERR_CAPACITY,
              "e2e_vi_solve_noalloc: %lld slices; at most %lld fit its fixed device scratch (full_sweeps = 1 "
              "lifts this limit)", (long long)m->nt, (long long)kNaMaxSlices);
    unsigned char* base = nullptr;
    E2E_TRY(noalloc_scratch(&base));

    const int max_sweeps = opts.max_sweeps > 0 ? opts.max_sweeps : (m->built ? (int)m->nt + 2 : 100000);
    const int spc = opts.sweeps_per_check > 0 ? opts.sweeps_per_check : 16;
    hipStream_t s = m->stream;
    const int nap = next_pow2((int)m->na);
    const int grp = 64 / nap;
    int lpr = 0;
    const SweepFn fn = comp ? pick_lean_nap<4, true>(nap) : pick_lean(nap, 4, &lpr);
    lpr = 4;
    int ptab_n = 0;
    int bpc = opts.blocks_per_cu;
    if (comp) {
        E2E_TRY(comp_launch_shape(fn, m->p.nrzns, opts.blocks_per_cu, kCompBlocksPerCU, &bpc, &ptab_n));
    } else if (bpc <= 0) {
        E2E_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, (const void*)fn, kStreamThreads, 0));
        bpc = std::min(bpc, kLeanBlocksPerCU);
    }
    const size_t shmem = (size_t)ptab_n * sizeof(float);
    int ncu = 0, dev = 0;
    E2E_HIP(hipGetDevice(&dev));
    E2E_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t resident = (int64_t)std::max(bpc, 1) * std::max(ncu, 1);
    const int64_t states_per_block = (int64_t)kStreamWaves * grp;
    const int64_t nblocks =
        std::min<int64_t>(std::min<int64_t>(ceil_div(ns, states_per_block), resident), kVIMaxBlocks);
    const int64_t W = nblocks * kStreamWaves;  // <= kNaMaxWaves

    VIState* st = reinterpret_cast<VIState*>(base);
    float* part = reinterpret_cast<float*>(base + kNaOffPart);
    uint32_t* wpart = reinterpret_cast<uint32_t*>(base + kNaOffWpart);
    double* W0 = reinterpret_cast<double*>(base + kNaOffW0);
    double* W1 = reinterpret_cast<double*>(base + kNaOffW1);
    int32_t* pi = reinterpret_cast<int32_t*>(base + kNaOffPi);
    float* V = reinterpret_cast<float*>(base + kNaOffV);

    // this model's partition (the array keeps the previous call's)
    if (active)
        launch_prefix_partitions(m->row_ptr, m->nc, (int)m->na, m->nt, W, grp, wpart, s);
    else
        launch_wave_partition(m->row_ptr, ns, (int)m->na, W, grp, wpart, s);
    E2E_HIP(hipGetLastError());

    VIArgs A;
    std::memset(&A, 0, sizeof(A));
    A.rp = m->row_ptr;
    A.nz = m->nz;
    A.R = m->R;
    A.V0 = V;
    A.V1 = V;
    A.W0 = W0;
    A.W1 = W1;
    A.pi = pi;
    A.st = st;
    A.part = part;
    A.ns = ns;
    A.na = (int)m->na;
    A.thresh = opts.thresh;
    A.discount = opts.discount;
    A.wpart = wpart;
    A.pre_nt = active ? (uint32_t)m->nt : 0u;
    A.pre_stride = active ? (uint32_t)(W + 1) : 0u;
    A.pool_parts = kLeanPoolParts;
    A.head16 = kLeanHead16;
    A.pool_min = kLeanPoolMin;
    A.cz = comp ? m->vi_cz : nullptr;
    A.cbase = comp ? m->vi_cbase : nullptr;
    A.inv_nr = comp ? 1.0 / (double)m->p.nrzns : 0.0;
    A.ptab_n = ptab_n;

    // W0, the output vector and the sweep state start at zero (a kernel, no memset node)
    hipLaunchKernelGGL(vi_init, dim3((unsigned)ceil_div(ns, 256)), dim3(256), 0, s, V, W0, ns, st);
    E2E_HIP(hipGetLastError());

    int launched = 0;
    VIState hst{};
    e2e_status rc = E2E_OK;
    int chunk = spc;
    while (launched < max_sweeps) {
        const int n = std::min(chunk, max_sweeps - launched);
        for (int i = 0; i < n; i++) hipLaunchKernelGGL(fn, dim3((unsigned)nblocks), dim3(kStreamThreads), shmem, s, A);
        E2E_HIP(hipGetLastError());
        E2E_HIP(hipMemcpyAsync(&hst, st, kVIStateHead, hipMemcpyDeviceToHost, s));
        E2E_HIP(hipStreamSynchronize(s));
        launched += n;
        if (opts.sweeps_per_check <= 0 && launched >= 2 * spc) chunk = std::min(chunk * 2, 4 * spc);
        if (hst.conv >= 0) break;
        if (opts.timeout_ms > 0 && now_ms() - t_start > opts.timeout_ms) {
            rc = E2E_ERR_TIMEOUT;
            break;
        }
    }
    const int sweeps = hst.conv >= 0 ? hst.conv : (int)hst.k - 1;
    if (rc == E2E_OK && hst.conv < 0) rc = E2E_ERR_NOT_CONVERGED;
    if (rc == E2E_ERR_NOT_CONVERGED) set_error("not converged after %d sweeps (residual %g)", sweeps, hst.last_res);
    if (rc == E2E_ERR_TIMEOUT) set_error("timeout after %d sweeps", sweeps);
    hipLaunchKernelGGL(vi_to_f32, dim3((unsigned)ceil_div(ns, 256)), dim3(256), 0, s, (sweeps & 1) ? W1 : W0, V, ns);
    E2E_HIP(hipGetLastError());
    if (value) E2E_HIP(hipMemcpyAsync(value, V, sizeof(float) * ns, hipMemcpyDeviceToHost, s));
    if (policy) E2E_HIP(hipMemcpyAsync(policy, pi, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, s));
    E2E_HIP(hipStreamSynchronize(s));
    if (stats) {  // no HIP events (they are created objects): kernel times and bytes are 0
        *stats = e2e_solve_stats{};
        stats->sweeps = sweeps;
        stats->converged = hst.conv >= 0;
        stats->final_residual = hst.last_res;
        stats->loads_in_flight = lpr;
        stats->kernel_variant = variant + 1;
        stats->blocks = (int32_t)nblocks;
        stats->solve_ms = now_ms() - t_start;
        stats->launches = launched;
        stats->compressed_windows = comp ? (int32_t)std::min<int64_t>(m->vi_cwin_comp, INT32_MAX) : 0;
    }
    return rc;
}

}  // namespace e2e
