// spmm_vi.hip — batched multi-goal value iteration: k goal columns that share one
// transition model, solved together as a CSR x dense-block SpMM per sweep.
//
// Column j is the model the reference would build with its terminal cell at goal g_j
// (term_subgrid_size 1), solved by the reference loop (src/spvi_2.cu:650-726). From one
// shared build with no terminal cell (term_subgrid_size 0, per-row reward sums kept):
//   * rows of a non-goal cell are identical in both models (the terminal test of
//     transition_calc, src/build_model.cu:523, only concerns S1), except for the reward of
//     lanes that land on the goal: move() adds r_terminal instead of the late-step
//     penalty (src/build_model.cu:343-354). With c the number of such lanes and
//     delta_t = r_terminal - (t > nt-20 ? 0.05*r_outbound : 0),
//         R_j(s,a) = (Rsum(s,a) + c * delta_t) / nr,
//     exact for integer-valued rewards (c = count is recovered exactly from
//     prob = count/nr);
//   * rows of the goal cell are self-loops with R = 0 in the per-goal model, so
//     V_j = 0 and pi_j = 0 there at every t; the kernel writes that directly.
//
// Layout: V[2][ns][KP] float64 (KP = padded column count, columns contiguous per
// state), pi [ns][KP]. grid.y walks blocks of KB columns. Two kernel families:
//   * spmm_vi_sweep<CPL, SPW>: a wave takes SPW states, 64/SPW lanes per state, CPL
//     adjacent columns per lane; each lane walks all na rows of its state one entry at
//     a time. KB = 64/SPW*CPL.
//   * spmm_vi_rows<LC, CPL, U>: a wave takes 64/(LC*NAP) states (NAP = na rounded up to
//     a power of two); lane (state, action, column lane) walks one row, U entries at a
//     time (all U entry loads, then all U gathers, then the in-range products added in
//     ascending order), so each lane keeps U gathers in flight and the LC lanes of a row
//     read each gathered line of V together. KB = LC*CPL.
//
// Arithmetic is that of the plain float64 sweeps: values carried in float64 between
// sweeps, each row's products (double)P * V added in ascending entry order, Q = R +
// discount * sum, first argmax, residual |V_new - V_old| rounded up to float. A column
// therefore reproduces its plain solve bit for bit.
//
// Each column freezes at the first sweep whose own residual is below thresh, exactly
// where its plain solve would stop; every column's arithmetic is independent of which
// other columns share the batch.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

#include "internal.h"
#include "lanes.h"

namespace e2e {

// bytes of BatchState the host reads back: k, done, all_conv (not the tile counters)
constexpr size_t kBatchStateHead = offsetof(BatchState, tnext);

struct BatchArgs {
    const uint32_t* rp;
    const int2* nz;
    const float* R;
    const float* Rsum;
    double* V0;
    double* V1;
    int32_t* pi;
    const int32_t* goal;   // [KP] goal cell per column (-1: none)
    int32_t* conv;         // [KP] sweep at which the column converged, -1 running
    float* part;           // [gridDim.x][KP] per-block column residual maxima
    float* last_res;       // [KP]
    BatchState* st;
    int64_t ns;
    int nc, na, nt, kp, nr;
    int nap;               // na rounded up to a power of two (spmm_vi_rows lane layout)
    float thresh, r_term, discount;
    double late_pen;       // 0.05 * r_outbound
    const float* Rcol;     // alpha columns: R of column j at Rcol[j * rows + row] (NULL: goal columns)
    int64_t rows;
    int rcol_n;            // columns Rcol holds (padding columns read the last one)
    const uint32_t* wpart; // spmm_vi_stream: [waves+1] first state of each wave
    uint32_t* colmax;      // spmm_vi_stream: [KP] running column residual maxima (float bits)
    int one_pass;          // spmm_vi_rows as one slice of a backward pass: states [s_lo, s_hi),
    uint32_t s_lo, s_hi;   // V0 == V1, no residual and no sweep state
    uint32_t pre_nt;       // sweeps of a solve whose columns all start at sweep 1, on a built model:
                           // nt, and each sweep recomputes the active slices only (0: every state)
    uint32_t pre_stride;   // spmm_vi_stream over the active slices: waves + 1, the stride of the
                           // per-prefix partitions in wpart
    int pre_live;          // live batches: a column's own sweep index is k - start + 1, and a block
    const int32_t* start;  // recomputes the slices its youngest running column needs ([KP])
    int g;                 // spmm_vi_halo: grid side (cells of slice t: i * g + j)
    uint32_t gmagic;       // spmm_vi_halo: ceil(2^32 / g), so that umulhi(c, gmagic) = c / g for c < g^2
};

namespace {

using BatchFn = void (*)(const BatchArgs);

// Slices [0, act) that sweep k recomputes for the columns c0 .. c0+kb-1 (act = the
// model's slices when every state is swept). A column admitted at sweep start[j] is at
// its own sweep k - start[j] + 1, and active_slices of that index covers every slice
// whose values are not yet final for it in both buffers; the youngest running column of
// the block needs the most. Wave-uniform (scalar loads of conv and start).
__device__ __forceinline__ int sweep_active(const BatchArgs& A, uint32_t k, int c0, int kb) {
    if (!A.pre_nt) return A.nt;
    uint32_t own = k;
    if (A.pre_live) {
        int young = 0;
        for (int j = c0; j < c0 + kb; j++)
            if (A.conv[j] < 0) young = max(young, A.start[j]);
        if (young <= 0) return 1;  // no running column in the block: nothing is written
        own = k - (uint32_t)young + 1u;  // start <= k for a running column
    }
    return active_slices(own, (int)A.pre_nt);
}

// float residual never below the double one r >= 0 (round to nearest, then one ulp up
// if that fell below r), the same rounding as the plain sweeps
__device__ __forceinline__ float res_up(double r) {
    const float f = (float)r;
    return (double)f < r ? __uint_as_float(__float_as_uint(f) + 1u) : f;
}

template <int CPL>
struct VecLoad;
template <>
struct VecLoad<1> {
    __device__ static void ld(const double* p, double* x) { x[0] = *p; }
};
template <>
struct VecLoad<2> {
    __device__ static void ld(const double* p, double* x) {
        const double2 v = *reinterpret_cast<const double2*>(p);
        x[0] = v.x;
        x[1] = v.y;
    }
};
template <>
struct VecLoad<4> {
    __device__ static void ld(const double* p, double* x) {
        const double2 v = reinterpret_cast<const double2*>(p)[0];
        const double2 w = reinterpret_cast<const double2*>(p)[1];
        x[0] = v.x;
        x[1] = v.y;
        x[2] = w.x;
        x[3] = w.y;
    }
};

template <>
struct VecLoad<8> {
    __device__ static void ld(const double* p, double* x) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const double2 v = reinterpret_cast<const double2*>(p)[i];
            x[2 * i] = v.x;
            x[2 * i + 1] = v.y;
        }
    }
};

constexpr int kBThreads = 1024;  // 16 waves per block
constexpr int kBWaves = kBThreads / 64;
constexpr int kBMaxBlocks = 512;

// End of a sweep, after each wave has put its per-column residual maxima in red[wave]:
// per-block column maxima to part[block x][KP]; the last block of the grid takes the
// maxima over the x-blocks, freezes the columns below thresh and publishes the sweep.
template <int KB, int NT = kBThreads>
__device__ __forceinline__ void sweep_tail(const BatchArgs& A, float (*red)[KB], float* fin, int& is_last,
                                           uint32_t k) {
    constexpr int NW = NT / 64;
    BatchState* st = A.st;
    const int tid = threadIdx.x;
    const int KP = A.kp;
    __syncthreads();
    // The block's partials go out by agent-scope atomic exchanges (wave 0: KB <= 64), and
    // the last block takes them by exchanges too: the same atomics as the ticket, whose
    // counts every XCD sees. The ticket's increment depends on the exchanges' returned
    // values, so it counts only once the partials are in place. No __threadfence: each
    // writes back the XCD's whole L2.
    uint32_t one = 1u;
    if (tid < KB) {
        float m = 0.0f;
#pragma unroll
        for (int w = 0; w < NW; w++) m = fmaxf(m, red[w][tid]);
        const uint32_t prev =
            __hip_atomic_exchange(reinterpret_cast<uint32_t*>(A.part) + (int64_t)blockIdx.x * KP + blockIdx.y * KB + tid,
                                  __float_as_uint(m), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        asm volatile("" : "+v"(one) : "v"(prev));
    }
    __syncthreads();
    if (tid == 0) is_last = atomicAdd(&st->done, one) == gridDim.x * gridDim.y - 1;
    __syncthreads();
    if (is_last) {
        // column j's residual = max over the x-blocks of part[bx][j]; thread t takes
        // column t % KP and every (NT/KP)-th x-block, then LDS combines
        const int groups = NT / KP;
        {
            const int j = tid % KP, g0 = tid / KP;
            float r = 0.0f;
#pragma unroll 8
            for (unsigned bx = g0; bx < gridDim.x; bx += groups)
                r = fmaxf(r, __uint_as_float(__hip_atomic_exchange(reinterpret_cast<uint32_t*>(A.part) + (int64_t)bx * KP + j,
                                                                   0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
            fin[tid] = r;
        }
        __syncthreads();
        int running = 0;
        for (int j = tid; j < KP; j += NT) {
            float r = 0.0f;
            for (int g2 = 0; g2 < groups; g2++) r = fmaxf(r, fin[g2 * KP + j]);
            if (A.conv[j] < 0) {
                A.last_res[j] = r;
                if (r < A.thresh) A.conv[j] = (int32_t)k;
                else running++;
            }
        }
        const int n_running = __syncthreads_count(running > 0);
        // every block has counted, so every tile sweep's block has left its counters; the
        // counters the next sweep's atomics count on are reset by atomics too
        for (int i = tid; i < kTileRows * 8; i += NT)
            __hip_atomic_exchange(&st->tnext[i * kCounterStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (tid == 0) {
            st->all_conv = n_running == 0;
            __hip_atomic_exchange(&st->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            st->k = k + 1u;
        }
    }
}

template <int CPL, int SPW>
__global__ __launch_bounds__(kBThreads) void spmm_vi_sweep(const BatchArgs A) {
    constexpr int LPS = 64 / SPW;
    constexpr int KB = LPS * CPL;
    __shared__ float red[kBWaves][KB];
    __shared__ float fin[kBThreads];
    __shared__ int is_last;
    BatchState* st = A.st;
    if (st->all_conv) return;
    const uint32_t k = st->k;
    const double* __restrict__ Vin = (k & 1u) ? A.V0 : A.V1;
    double* __restrict__ Vout = (k & 1u) ? A.V1 : A.V0;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int grp = lane / LPS, l = lane % LPS;
    const int j0 = blockIdx.y * KB + l * CPL;
    const int KP = A.kp;
    int goal[CPL];
    bool live[CPL];
    float rmax[CPL];
#pragma unroll
    for (int c = 0; c < CPL; c++) {
        goal[c] = A.goal[j0 + c];
        live[c] = A.conv[j0 + c] < 0;
        rmax[c] = 0.0f;
    }
    const int64_t stride = (int64_t)gridDim.x * kBWaves * SPW;
    const int64_t s_end = A.pre_nt ? (int64_t)sweep_active(A, k, blockIdx.y * KB, KB) * A.nc : A.ns;
    for (int64_t s = ((int64_t)blockIdx.x * kBWaves + wv) * SPW + grp; s < s_end; s += stride) {
        const int t = (int)(s / A.nc);
        const int cell = (int)(s - (int64_t)t * A.nc);
        const int colbase = (t + 1) * A.nc;
        const float delta = (float)((double)A.r_term - (t > A.nt - 20 ? A.late_pen : 0.0));
        double best[CPL];
        int bi[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            best[c] = -INFINITY;
            bi[c] = -1;
        }
        for (int a = 0; a < A.na; a++) {
            const int64_t row = s * A.na + a;
            const uint32_t beg = A.rp[row], end = A.rp[row + 1];
            double sd[CPL];
            int hit[CPL];
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                sd[c] = 0.0;
                hit[c] = -1;
            }
            for (uint32_t j = beg; j < end; j++) {
                const int2 e = A.nz[j];
                const float v = __int_as_float(e.y);
                double x[CPL];
                VecLoad<CPL>::ld(Vin + (int64_t)e.x * KP + j0, x);
                const int cc = e.x - colbase;
#pragma unroll
                for (int c = 0; c < CPL; c++) {
                    sd[c] = sd[c] + (double)v * x[c];
                    if (cc == goal[c]) hit[c] = (int)rintf(v * (float)A.nr);
                }
            }
            const float rr = A.R[row];
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                const float rj = A.Rcol ? A.Rcol[(int64_t)min(j0 + c, A.rcol_n - 1) * A.rows + row]
                                 : hit[c] >= 0 ? (A.Rsum[row] + (float)hit[c] * delta) / A.nr
                                               : rr;
                double q = (double)rj + (double)A.discount * sd[c];
                if (q != q) q = -INFINITY;
                if (q > best[c]) {  // strict '>' in action order (spvi_2.cu:112)
                    best[c] = q;
                    bi[c] = a;
                }
            }
        }
        double vo[CPL];
        int po[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            if (cell == goal[c]) {
                vo[c] = 0.0;
                po[c] = 0;
            } else if (best[c] > -1e6) {
                vo[c] = best[c];
                po[c] = bi[c];
            } else {
                vo[c] = -1e6;
                po[c] = -1;
            }
        }
        const int64_t o = s * KP + j0;
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            if (live[c]) {
                Vout[o + c] = vo[c];
                A.pi[o + c] = po[c];
                rmax[c] = fmaxf(rmax[c], res_up(fabs(vo[c] - Vin[o + c])));
            }
        }
    }
    // per-column residual: lanes of the same column across the SPW groups, then waves
#pragma unroll
    for (int c = 0; c < CPL; c++)
        for (int off = LPS; off < 64; off <<= 1) rmax[c] = fmaxf(rmax[c], __shfl_xor(rmax[c], off, 64));
    if (grp == 0) {
#pragma unroll
        for (int c = 0; c < CPL; c++) red[wv][l * CPL + c] = rmax[c];
    }
    sweep_tail<KB>(A, red, fin, is_last, k);
}

// Row-parallel sweep. Lane (sl, a, cl) = lane / (LC*NAP), (lane / LC) % NAP, lane % LC
// takes row a of state sl of the wave's current group and columns j0 .. j0+CPL-1. Each
// block takes a contiguous range of groups (blocks remapped so that the 8 XCDs each
// get a contiguous eighth of the states, keeping the gathers of V(t+1) in one XCD's L2);
// its waves interleave over that range.
template <int LC, int CPL, int U>
__global__ __launch_bounds__(kBThreads) void spmm_vi_rows(const BatchArgs A) {
    constexpr int KB = LC * CPL;
    __shared__ float red[kBWaves][KB];
    __shared__ float fin[kBThreads];
    __shared__ int is_last;
    BatchState* st = A.st;
    if (!A.one_pass && st->all_conv) return;
    const uint32_t k = A.one_pass ? 1u : st->k;
    const double* __restrict__ Vin = (k & 1u) ? A.V0 : A.V1;
    double* __restrict__ Vout = (k & 1u) ? A.V1 : A.V0;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nap = A.nap;
    const int cl = lane % LC;
    const int a = (lane / LC) & (nap - 1);
    const int sl = lane / (LC * nap);
    const int spw = 64 / (LC * nap);
    const int j0 = blockIdx.y * KB + cl * CPL;
    const int KP = A.kp;
    const uint32_t ns = (uint32_t)A.ns, na = (uint32_t)A.na, nc = (uint32_t)A.nc;
    int goal[CPL];
    bool live[CPL];
    float rmax[CPL];
#pragma unroll
    for (int c = 0; c < CPL; c++) {
        goal[c] = A.goal[j0 + c];
        live[c] = A.conv[j0 + c] < 0;
        rmax[c] = 0.0f;
    }
    // a sweep covers every state (a built model's: the active slices); one slice of a
    // backward pass covers [s_lo, s_hi)
    const uint32_t s_lo = A.one_pass ? A.s_lo : 0u;
    const uint32_t s_hi = A.one_pass ? A.s_hi
                          : A.pre_nt ? (uint32_t)sweep_active(A, k, blockIdx.y * KB, KB) * nc
                                     : ns;
    const uint32_t G = (s_hi - s_lo + spw - 1) / spw;
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned pb = (nb % 8u == 0u) ? (b % 8u) * (nb / 8u) + b / 8u : b;
    const uint32_t g_beg = (uint32_t)((uint64_t)G * pb / nb), g_end = (uint32_t)((uint64_t)G * (pb + 1) / nb);
    for (uint32_t g = g_beg + wv; g < g_end; g += kBWaves) {
        const uint32_t s = s_lo + g * spw + sl;
        const bool sval = s < s_hi;
        const bool valid = sval && (uint32_t)a < na;
        const uint32_t sc = sval ? s : ns - 1u;
        const uint32_t t = sc / nc;
        const int cell = (int)(sc - t * nc);
        const int colbase = (int)((t + 1) * nc);
        const float delta = (float)((double)A.r_term - ((int)t > A.nt - 20 ? A.late_pen : 0.0));
        const int64_t row = (int64_t)sc * na + min((uint32_t)a, na - 1u);
        uint32_t jb = 0, je = 0;
        if (valid) {
            jb = A.rp[row];
            je = A.rp[row + 1];
        }
        double sd[CPL];
        int hit[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            sd[c] = 0.0;
            hit[c] = -1;
        }
        // entries jb .. jb+U-1 may run past the row: they are loaded (the entry array
        // has a zeroed tail of kNzPad entries, and every column is a state < ns, so the
        // gather stays inside V) but not added
        while (jb < je) {
            int2 e[U];
#pragma unroll
            for (int u = 0; u < U; u++) e[u] = A.nz[jb + u];
            double x[U][CPL];
#pragma unroll
            for (int u = 0; u < U; u++) VecLoad<CPL>::ld(Vin + (int64_t)e[u].x * KP + j0, x[u]);
            const uint32_t cnt = je - jb;
#pragma unroll
            for (int u = 0; u < U; u++) {
                if ((uint32_t)u < cnt) {
                    const float v = __int_as_float(e[u].y);
                    const int cc = e[u].x - colbase;
#pragma unroll
                    for (int c = 0; c < CPL; c++) {
                        sd[c] = sd[c] + (double)v * x[u][c];
                        if (cc == goal[c]) hit[c] = (int)rintf(v * (float)A.nr);
                    }
                }
            }
            jb += U;
        }
        double q[CPL];
        int ba[CPL];
        const float rr = valid ? A.R[row] : 0.0f;
        const float rs = (valid && A.Rsum) ? A.Rsum[row] : 0.0f;
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            const float rj = A.Rcol ? (valid ? A.Rcol[(int64_t)min(j0 + c, A.rcol_n - 1) * A.rows + row] : 0.0f)
                             : hit[c] >= 0 ? (rs + (float)hit[c] * delta) / A.nr
                                           : rr;
            double qq = valid ? (double)rj + (double)A.discount * sd[c] : -INFINITY;
            if (qq != qq) qq = -INFINITY;
            q[c] = qq;
            ba[c] = valid ? a : 0x7fffffff;
        }
        // max / first argmax over the NAP lanes of the state (strict '>' in action
        // order, spvi_2.cu:112: the larger value, on a tie the lower action); the
        // combine is associative and commutative, so the butterfly gives every lane
        // the result of the sequential scan
        for (int o = LC; o < LC * nap; o <<= 1) {
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                const double q2 = __shfl_xor(q[c], o, 64);
                const int a2 = __shfl_xor(ba[c], o, 64);
                if (q2 > q[c] || (q2 == q[c] && a2 < ba[c])) {
                    q[c] = q2;
                    ba[c] = a2;
                }
            }
        }
        if (a == 0 && sval) {
            const int64_t o = (int64_t)s * KP + j0;
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                double vo;
                int po;
                if (cell == goal[c]) {
                    vo = 0.0;
                    po = 0;
                } else if (q[c] > -1e6) {
                    vo = q[c];
                    po = ba[c];
                } else {
                    vo = -1e6;
                    po = -1;
                }
                if (live[c]) {
                    Vout[o + c] = vo;
                    A.pi[o + c] = po;
                    rmax[c] = fmaxf(rmax[c], res_up(fabs(vo - Vin[o + c])));
                }
            }
        }
    }
    if (A.one_pass) return;  // block-uniform
    // per-column residual: the lanes of one column lane cl, then the waves
#pragma unroll
    for (int c = 0; c < CPL; c++)
        for (int off = LC; off < 64; off <<= 1) rmax[c] = fmaxf(rmax[c], __shfl_xor(rmax[c], off, 64));
    if (lane < LC) {
#pragma unroll
        for (int c = 0; c < CPL; c++) red[wv][cl * CPL + c] = rmax[c];
    }
    sweep_tail<KB>(A, red, fin, is_last, k);
}

// Halo sweep. Transitions of the built models go from a cell of slice t to cells near it
// in slice t+1 (the named workloads reach at most 7 cells, 97 % of entries at most 3,
// 99.5 % at most 4). A block takes a tile of kHTileH x kHTileW cells of one slice and
// stages the values of slice t+1 over the tile widened by H cells on every side (its KB
// columns, float64) in LDS with coalesced loads. Each wave then takes one tile row: a
// run of consecutive cells, so one contiguous range of the state-major entries, which it
// streams in windows of WIN entries (coalesced loads, staged in its own LDS window, the
// next window's loads issued before the current one is consumed). Lane (sl, a, cl) =
// state sl of the wave's current group of 64 / (LC * NAP) states, action a, column lane
// cl (CPL adjacent columns) walks its row's entries of each window from LDS and reads
// each entry's values from the halo (from global memory for an entry that lands outside
// it). Per row and column the products are added in ascending entry order, as in every
// other sweep, so the bits are theirs.
constexpr int kHThreads = 256;
constexpr int kHWaves = kHThreads / 64;
constexpr int kHTileH = 4, kHTileW = 16;
constexpr int kHMaxBlocks = 2048;

template <int LC, int CPL, int H>
__global__ __launch_bounds__(kHThreads) void spmm_vi_halo(const BatchArgs A) {
    constexpr int KB = LC * CPL;
    constexpr int HH = kHTileH + 2 * H, HW = kHTileW + 2 * H;
    constexpr int U = 4, WIN = 64 * U;
    constexpr int U2 = CPL <= 4 ? 4 : 2;  // entries per step of a lane's row walk
    static_assert(CPL % 2 == 0, "columns are staged and read in pairs");
    __shared__ double2 halo[HH * HW * KB / 2];
    __shared__ int2 win[kHWaves][WIN];
    __shared__ float red[kHWaves][KB];
    __shared__ float fin[kHThreads];
    __shared__ int is_last;
    BatchState* st = A.st;
    if (!A.one_pass && st->all_conv) return;
    const uint32_t k = A.one_pass ? 1u : st->k;
    const double* __restrict__ Vin = (k & 1u) ? A.V0 : A.V1;
    double* __restrict__ Vout = (k & 1u) ? A.V1 : A.V0;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nap = A.nap;
    const int cl = lane % LC;
    const int a = (lane / LC) & (nap - 1);
    const int sl = lane / (LC * nap);
    const int spw = 64 / (LC * nap);
    const int jb0 = blockIdx.y * KB;        // the block's first column
    const int j0 = jb0 + cl * CPL;          // this lane's first column
    const int KP = A.kp;
    const int g = A.g, nc = A.nc, na = A.na;
    int goal[CPL];
    bool live[CPL];
    float rmax[CPL];
#pragma unroll
    for (int c = 0; c < CPL; c++) {
        goal[c] = A.goal[j0 + c];
        live[c] = A.conv[j0 + c] < 0;
        rmax[c] = 0.0f;
    }
    const bool goals = A.Rcol == nullptr;
    int2* wb = win[wv];
    // tiles of the slices this launch covers: every active slice of the sweep, or the one
    // slice of a backward-pass step
    const int tr = (g + kHTileH - 1) / kHTileH, tcn = (g + kHTileW - 1) / kHTileW, tps = tr * tcn;
    const int t_lo = A.one_pass ? (int)(A.s_lo / (uint32_t)nc) : 0;
    const int t_hi = A.one_pass ? t_lo + 1 : A.pre_nt ? sweep_active(A, k, jb0, KB) : A.nt;
    const int ntiles = (t_hi - t_lo) * tps;
    // XCD x (blocks b = x mod 8) walks the band [tlo, thi) of the tiles; a block's first
    // tile is its own and it takes the band's next from the XCD's counter as it finishes,
    // as in spmm_vi_tile (the backward pass and column blocks beyond kTileRows stride)
    const unsigned nb = gridDim.x;
    const unsigned xcds = (nb % 8u == 0u) ? 8u : 1u;
    const unsigned xcd = blockIdx.x % xcds, nbx = nb / xcds, bi = blockIdx.x / xcds;
    const int tlo = (int)((int64_t)ntiles * xcd / xcds), thi = (int)((int64_t)ntiles * (xcd + 1) / xcds);
    const bool dyn = !A.one_pass && blockIdx.y < (unsigned)kTileRows;
    uint32_t* const tctr = dyn ? st->tnext + (blockIdx.y * 8u + xcd) * (unsigned)kCounterStride : nullptr;
    __shared__ int next_tile;
    auto advance = [&](int tile) {
        if (!dyn) return tile + (int)nbx;
        __syncthreads();  // every wave has read next_tile
        if (tid == 0) next_tile = tlo + (int)nbx + (int)atomicAdd(tctr, 1u);
        __syncthreads();
        return next_tile;
    };
    for (int tile = tlo + (int)bi; tile < thi; tile = advance(tile)) {
        const int t = t_lo + tile / tps;
        const int rem = tile - (t - t_lo) * tps;
        const int ti = rem / tcn, tj = rem - ti * tcn;
        const int i0 = ti * kHTileH, jc0 = tj * kHTileW;
        const int th = min(kHTileH, g - i0), tw = min(kHTileW, g - jc0);
        const int hi0 = i0 - H, hj0 = jc0 - H;  // grid cell of halo entry (0, 0)
        const int colbase = (t + 1) * nc;
        __syncthreads();  // the previous tile's reads of the halo are done
        if (t + 1 < A.nt) {
            // halo[(hi * HW + hj) * KB/2 + c2] = V(t+1)[cell (hi0 + hi, hj0 + hj)][columns
            // jb0 + 2 c2, +1]; cells outside the grid are never read (every entry lands
            // on a cell of the grid)
            for (int q = tid; q < HH * HW * (KB / 2); q += kHThreads) {
                const int c2 = q % (KB / 2), hc = q / (KB / 2);
                const int hi = hc / HW, hj = hc - hi * HW;
                const int i = hi0 + hi, j = hj0 + hj;
                if (i >= 0 && i < g && j >= 0 && j < g)
                    halo[q] = *reinterpret_cast<const double2*>(Vin + (int64_t)(colbase + i * g + j) * KP + jb0 +
                                                                2 * c2);
            }
        }
        __syncthreads();
        if (wv >= th) continue;  // block-uniform per wave; the next tile's barrier waits for all
        // does a goal of the block's columns lie in the halo (then entries landing in it
        // are compared with the goals; else only those landing outside it)
        bool goal_tile = false;
        if (goals) {
            for (int c = 0; c < KB; c++) {
                const int gc = A.goal[jb0 + c];
                if (gc >= 0) {
                    const int gi2 = gc / g, gj2 = gc - gi2 * g;
                    goal_tile = goal_tile || ((unsigned)(gi2 - hi0) < (unsigned)HH && (unsigned)(gj2 - hj0) < (unsigned)HW);
                }
            }
        }
        const float delta = (float)((double)A.r_term - (t > A.nt - 20 ? A.late_pen : 0.0));
        // this wave's run: tile row wv, cells jc0 .. jc0+tw-1 (states s_run .. s_run+tw-1)
        const int cell0 = (i0 + wv) * g + jc0;
        const int64_t s_run = (int64_t)t * nc + cell0;
        const int ngroups = (tw + spw - 1) / spw;
        const uint32_t P0 = A.rp[s_run * na], Pend = A.rp[(s_run + tw) * na];
        // row data of group gi (clamped to the run's last state, so always in range)
        struct Meta {
            uint32_t rb, re, ce;
            float rr, rs;
        };
        auto load_meta = [&](int gi, Meta& M) {
            const int q = min(gi * spw + sl, tw - 1);
            const int64_t row = (s_run + q) * na + min(a, na - 1);
            M.rb = A.rp[row];
            M.re = A.rp[row + 1];
            M.ce = A.rp[(s_run + min(gi * spw + spw, tw)) * na];
            M.rr = A.R[row];
            M.rs = A.Rsum ? A.Rsum[row] : 0.0f;
        };
        Meta cur, nxt;
        int gi = 0;
        load_meta(0, cur);
        load_meta(min(1, ngroups - 1), nxt);
        double sd[CPL];
        int hit[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            sd[c] = 0.0;
            hit[c] = -1;
        }
        auto finish_group = [&]() {
            const int q = gi * spw + sl;
            const bool sval = q < tw;
            const bool valid = sval && a < na;
            const int64_t s = s_run + min(q, tw - 1);
            const int64_t row = s * na + min(a, na - 1);
            double qv[CPL];
            int ba[CPL];
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                const float rj = !goals ? (valid ? A.Rcol[(int64_t)min(j0 + c, A.rcol_n - 1) * A.rows + row] : 0.0f)
                                 : hit[c] >= 0 ? (cur.rs + (float)hit[c] * delta) / A.nr
                                               : cur.rr;
                double qq = valid ? (double)rj + (double)A.discount * sd[c] : -INFINITY;
                if (qq != qq) qq = -INFINITY;
                qv[c] = qq;
                ba[c] = valid ? a : 0x7fffffff;
            }
            // max / first argmax over the NAP lanes of the state (strict '>' in action
            // order, spvi_2.cu:112), as in spmm_vi_rows
            for (int o = LC; o < LC * nap; o <<= 1) {
#pragma unroll
                for (int c = 0; c < CPL; c++) {
                    const double q2 = __shfl_xor(qv[c], o, 64);
                    const int a2 = __shfl_xor(ba[c], o, 64);
                    if (q2 > qv[c] || (q2 == qv[c] && a2 < ba[c])) {
                        qv[c] = q2;
                        ba[c] = a2;
                    }
                }
            }
            if (a == 0 && sval) {
                const int cell = cell0 + q;
                const int64_t o = s * KP + j0;
#pragma unroll
                for (int c = 0; c < CPL; c++) {
                    double vo;
                    int po;
                    if (cell == goal[c]) {
                        vo = 0.0;
                        po = 0;
                    } else if (qv[c] > -1e6) {
                        vo = qv[c];
                        po = ba[c];
                    } else {
                        vo = -1e6;
                        po = -1;
                    }
                    if (live[c]) {
                        Vout[o + c] = vo;
                        A.pi[o + c] = po;
                        rmax[c] = fmaxf(rmax[c], res_up(fabs(vo - Vin[o + c])));
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                sd[c] = 0.0;
                hit[c] = -1;
            }
            gi++;
            cur = nxt;
            load_meta(min(gi + 1, ngroups - 1), nxt);
        };
        // the run's entries, windows of WIN from P0; entries past Pend are loaded (the
        // entry array has a zeroed tail of kNzPad entries) but belong to no row of the run.
        // Each entry is decoded once, as it is staged: x = the halo index (double2 units)
        // of its landing cell's record, or 0x80000000 | column for an entry that lands
        // outside the halo; y = its probability bits
        auto decode = [&](int2 e) {
            const int cc = e.x - colbase;
            const int i2 = (int)__umulhi((uint32_t)cc, A.gmagic);
            const int hi = i2 - hi0, hj = cc - i2 * g - hj0;
            const bool in = (unsigned)hi < (unsigned)HH && (unsigned)hj < (unsigned)HW;
            return make_int2(in ? (hi * HW + hj) * (KB / 2) : (int)(0x80000000u | (uint32_t)e.x), e.y);
        };
        const uint32_t nwin = Pend > P0 ? (Pend - P0 + WIN - 1) / WIN : 0u;
        int2 e[U];
        if (nwin > 0) {
#pragma unroll
            for (int u = 0; u < U; u++) e[u] = A.nz[P0 + u * 64 + lane];
        }
        for (uint32_t kw = 0; kw < nwin; kw++) {
            const uint32_t w0 = P0 + kw * WIN, wend = w0 + WIN;
#pragma unroll
            for (int u = 0; u < U; u++) wb[u * 64 + lane] = decode(e[u]);
            if (kw + 1 < nwin) {  // the next window's loads, in flight while this one is used
#pragma unroll
                for (int u = 0; u < U; u++) e[u] = A.nz[wend + u * 64 + lane];
            }
            __builtin_amdgcn_wave_barrier();
            for (;;) {
                const bool valid = gi * spw + sl < tw && a < na;
                const uint32_t jlo = valid ? max(cur.rb, w0) : 0u, jhi = valid ? min(cur.re, wend) : 0u;
                // U2 entries per step: their window reads, then their halo reads, are issued
                // together (an entry outside the halo reads halo record 0 and is then taken
                // from global memory); the products are added in entry order
                for (uint32_t j = jlo; j < jhi; j += U2) {
                    int2 en[U2];
#pragma unroll
                    for (int u = 0; u < U2; u++) en[u] = wb[min(j + (uint32_t)u, jhi - 1u) - w0];
                    double2 hv[U2][CPL / 2];
#pragma unroll
                    for (int u = 0; u < U2; u++) {
                        const double2* hp = halo + (en[u].x >= 0 ? en[u].x : 0) + cl * (CPL / 2);
#pragma unroll
                        for (int c2 = 0; c2 < CPL / 2; c2++) hv[u][c2] = hp[c2];
                    }
#pragma unroll
                    for (int u = 0; u < U2; u++) {
                        if (j + (uint32_t)u >= jhi) break;
                        const float v = __int_as_float(en[u].y);
                        const double pv = (double)v;
                        if (en[u].x >= 0) {  // the landing cell's values from the halo
#pragma unroll
                            for (int c2 = 0; c2 < CPL / 2; c2++) {
                                sd[2 * c2] = sd[2 * c2] + pv * hv[u][c2].x;
                                sd[2 * c2 + 1] = sd[2 * c2 + 1] + pv * hv[u][c2].y;
                            }
                            if (goal_tile) {  // a goal of these columns lies in the halo
                                const int hc = en[u].x / (KB / 2), hi = hc / HW;
                                const int cc = (hi0 + hi) * g + hj0 + (hc - hi * HW);
                                const int n = (int)rintf(v * (float)A.nr);
#pragma unroll
                                for (int c = 0; c < CPL; c++)
                                    if (cc == goal[c]) hit[c] = n;
                            }
                        } else {  // outside the halo: from global memory
                            const int col = en[u].x & 0x7fffffff;
                            double x[CPL];
                            VecLoad<CPL>::ld(Vin + (int64_t)col * KP + j0, x);
                            const int cc = col - colbase;
                            const int n = (int)rintf(v * (float)A.nr);
#pragma unroll
                            for (int c = 0; c < CPL; c++) {
                                sd[c] = sd[c] + pv * x[c];
                                if (goals && cc == goal[c]) hit[c] = n;
                            }
                        }
                    }
                }
                if (cur.ce > wend) break;  // the group continues in the next window
                finish_group();
                if (gi >= ngroups) break;
            }
            __builtin_amdgcn_wave_barrier();
        }
        while (gi < ngroups) finish_group();  // trailing groups whose rows are all empty
    }
    if (A.one_pass) return;  // block-uniform
#pragma unroll
    for (int c = 0; c < CPL; c++)
        for (int off = LC; off < 64; off <<= 1) rmax[c] = fmaxf(rmax[c], __shfl_xor(rmax[c], off, 64));
    if (lane < LC) {
#pragma unroll
        for (int c = 0; c < CPL; c++) red[wv][cl * CPL + c] = rmax[c];
    }
    sweep_tail<KB, kHThreads>(A, red, fin, is_last, k);
}

// Tile sweep: the halo kernel's tiles and LDS halo, with one lane per row over all KB
// columns of the block and a row group's entries staged whole.
//   * The halo record of a cell holds its KB values of slice t+1 and, after them, the
//     mask of the block's columns whose goal is that cell (a stride of KB + 2 doubles, so
//     that lanes reading the same column pair of different cells fall on different banks).
//     Cells outside the grid hold zeros.
//   * A wave takes one tile row, in groups of 64 / NAP states: lane (state, action) walks
//     its row over every entry of the group's window (windows are cut at group bounds,
//     WIN = 64 * U entries; the next window's entries are loaded while this one is
//     walked). Staging decodes each entry to the LDS offset of its landing cell's record;
//     an entry landing outside the halo gets one of the wave's kTOvf overflow records,
//     filled from global memory as the window is staged; a window whose entries need more
//     ends at the first entry without one. The walk reads LDS only.
//   * The walk reads U2 entries per step, all their records together, and the next
//     step's entries ahead. A step slot past the row's end adds 0 * (a finite value),
//     which leaves the sum's bits unchanged (a sum that starts at +0 is never -0).
//   * Each row's Q (KB values) goes to LDS; lane (state, column) then takes the max and
//     first argmax over the state's actions in action order and writes V, pi and the
//     residual of one column, so the stores of a state's KB columns are contiguous.
// The per-row, per-column arithmetic is the other sweeps': products added in ascending
// entry order, the same Q, argmax and residual rounding, so the bits are theirs.
constexpr int kTOvf = 8;  // overflow records per wave

template <int KB, int U, int U2, int H, int QH>
__global__ __launch_bounds__(kHThreads) void spmm_vi_tile(const BatchArgs A) {
    constexpr int HH = kHTileH + 2 * H, HW = kHTileW + 2 * H;
    constexpr int CS = KB + 2;  // doubles per record: KB values, goal mask, pad
    constexpr int NREC = HH * HW;
    constexpr int WIN = 64 * U;
    constexpr int KQ = KB / QH;                          // columns per Q pass
    constexpr int WB = WIN > 64 * KQ ? WIN : 64 * KQ;    // doubles per wave: entries, then Q
    constexpr int QM = KQ / 2 - 1;                       // Q slots (double2) per row - 1
    static_assert(KB % 2 == 0 && 64 % KB == 0 && KB <= 32 && (KB & (KB - 1)) == 0,
                  "columns are staged in pairs, one column per lane");
    static_assert(QH >= 1 && KB % QH == 0 && KQ % 2 == 0, "Q passes of whole column pairs");
    static_assert(kHTileH == kHWaves, "one tile row per wave");
    __shared__ __align__(16) double halo[(NREC + kHWaves * kTOvf) * CS];
    __shared__ __align__(16) double wbuf[kHWaves][WB];
    __shared__ float red[kHWaves][KB];
    __shared__ float fin[kHThreads];
    __shared__ int is_last;
    // the block's goal cells and their grid rows and columns (-1: none), read from LDS so
    // that they take no scalar registers
    __shared__ int goal_s[KB], goal_i[KB], goal_j[KB];
    BatchState* st = A.st;
    if (!A.one_pass && st->all_conv) return;
    const uint32_t k = A.one_pass ? 1u : st->k;
    const double* __restrict__ Vin = (k & 1u) ? A.V0 : A.V1;
    double* __restrict__ Vout = (k & 1u) ? A.V1 : A.V0;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nap = A.nap;
    const int a = lane & (nap - 1);
    const int sl = lane / nap;
    const int spg = 64 / nap;  // states per group
    const int jb0 = blockIdx.y * KB;
    const int KP = A.kp;
    const int g = A.g, nc = A.nc, na = A.na;
    const bool goals = A.Rcol == nullptr;
    if (tid < KB) {  // read after the first tile's barrier
        const int gc = A.goal[jb0 + tid];
        goal_s[tid] = gc;
        goal_i[tid] = gc >= 0 ? gc / g : -1;
        goal_j[tid] = gc >= 0 ? gc - (gc / g) * g : -1;
    }
    // the columns of this lane in the argmax phase: cq + h * KQ in pass h; qoff[i]: the
    // offset of column cq in row i of a run of QM + 1 rows whose first row is a multiple of
    // QM + 1
    const int cq = lane % KQ;
    int qoff[QM + 1];
#pragma unroll
    for (int i = 0; i <= QM; i++) qoff[i] = i * KQ + 2 * ((cq >> 1) ^ i) + (cq & 1);
    int goal_q[QH];
    bool live_q[QH];
    float rmax[QH];
#pragma unroll
    for (int h = 0; h < QH; h++) {
        goal_q[h] = A.goal[jb0 + h * KQ + cq];
        live_q[h] = A.conv[jb0 + h * KQ + cq] < 0;
        rmax[h] = 0.0f;
    }
    int2* const wb = reinterpret_cast<int2*>(wbuf[wv]);
    double2* const qb = reinterpret_cast<double2*>(wbuf[wv]);
    const int ovf0 = (NREC + wv * kTOvf) * CS;  // this wave's first overflow record
    const uint64_t below = (1ull << lane) - 1ull;
    // XCD x (blocks b = x mod 8; consecutive blocks run on different XCDs) walks a
    // contiguous band [tlo, thi) of the sweep's tiles, so the halos it stages overlap in
    // its own L2. A block's first tile is its own; as it finishes one it takes the band's
    // next from the XCD's counter (tiles of equal work end at different times on different
    // CUs, and edge tiles are smaller), except in the backward pass (no sweep state to
    // reset) and for column blocks beyond kTileRows, which stride the band. Each tile is
    // still swept once, by one block: the same bits whichever block takes it.
    const unsigned nb = gridDim.x;
    const unsigned xcds = (nb % 8u == 0u) ? 8u : 1u;
    const unsigned xcd = blockIdx.x % xcds, nbx = nb / xcds, bi = blockIdx.x / xcds;
    const int tr = (g + kHTileH - 1) / kHTileH, tcn = (g + kHTileW - 1) / kHTileW, tps = tr * tcn;
    const int t_lo = A.one_pass ? (int)(A.s_lo / (uint32_t)nc) : 0;
    const int t_hi = A.one_pass ? t_lo + 1 : A.pre_nt ? sweep_active(A, k, jb0, KB) : A.nt;
    const int ntiles = (t_hi - t_lo) * tps;
    const int tlo = (int)((int64_t)ntiles * xcd / xcds), thi = (int)((int64_t)ntiles * (xcd + 1) / xcds);
    const bool dyn = !A.one_pass && blockIdx.y < (unsigned)kTileRows;
    uint32_t* const tctr = dyn ? st->tnext + (blockIdx.y * 8u + xcd) * (unsigned)kCounterStride : nullptr;
    __shared__ int next_tile;
    // the block's next tile (a loop increment, so the waves that skip a tile reach it too)
    auto advance = [&](int tile) {
        if (!dyn) return tile + (int)nbx;
        __syncthreads();  // every wave has read next_tile
        if (tid == 0) next_tile = tlo + (int)nbx + (int)atomicAdd(tctr, 1u);
        __syncthreads();
        return next_tile;
    };
    for (int tile = tlo + (int)bi; tile < thi; tile = advance(tile)) {
        const int t = t_lo + tile / tps;
        const int rem = tile - (t - t_lo) * tps;
        const int ti = rem / tcn, tj = rem - ti * tcn;
        const int i0 = ti * kHTileH, jc0 = tj * kHTileW;
        const int th = min(kHTileH, g - i0), tw = min(kHTileW, g - jc0);
        const int hi0 = i0 - H, hj0 = jc0 - H;
        const int colbase = (t + 1) * nc;
        // this wave's run: tile row wv, cells jc0 .. jc0+tw-1 (none when wv >= th)
        const bool act = wv < th;
        const int cell0 = (i0 + min(wv, th - 1)) * g + jc0;
        const int64_t s_run = (int64_t)t * nc + cell0;
        const int ngroups = act ? (tw + spg - 1) / spg : 0;
        // first entry of group gi (gi = ngroups: the run's end); lane i of gsv holds that
        // of group i, for the groups after the first three
        auto gstart = [&](int gi) { return A.rp[(s_run + min(gi * spg, tw)) * na]; };
        const uint32_t gsv = (act && lane <= ngroups) ? A.rp[(s_run + min(lane * spg, tw)) * na] : 0u;
        struct Meta {
            uint32_t rb, re;
            float rr, rs;
        };
        // the row data of this lane's row of group gi
        auto load_meta = [&](int gi) {
            Meta M;
            const int q = gi * spg + sl;
            const bool valid = q < tw && a < na;
            const int64_t row = (s_run + min(q, tw - 1)) * na + min(a, na - 1);
            M.rb = valid ? A.rp[row] : 0u;
            M.re = valid ? A.rp[row + 1] : 0u;
            M.rr = A.R[row];
            M.rs = (goals && A.Rsum) ? A.Rsum[row] : 0.0f;
            return M;
        };
        // group 0's row data and the run's first window, requested before the halo is staged
        Meta cur{0u, 0u, 0.0f, 0.0f};
        uint32_t G1 = 0u, G2 = 0u;   // ends of the current and the next group
        uint32_t nws = 0u, nwe = 0u;  // the window held in e[]
        int2 e[U];
        if (act) {
            cur = load_meta(0);
            nws = gstart(0);
            G1 = gstart(1);
            G2 = ngroups >= 2 ? gstart(2) : G1;
            nwe = min(nws + (uint32_t)WIN, G1);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t idx = nws + (uint32_t)(u * 64 + lane);
                e[u] = idx < nwe ? A.nz[idx] : make_int2(0, 0);
            }
        }
        __syncthreads();  // the previous tile's reads of the halo are done
        if (t + 1 < A.nt) {
            for (int q = tid; q < NREC * (KB / 2); q += kHThreads) {
                const int c2 = q % (KB / 2), hc = q / (KB / 2);
                const int hi = hc / HW, hj = hc - hi * HW;
                const int i = hi0 + hi, j = hj0 + hj;
                double2 v = make_double2(0.0, 0.0);
                if (i >= 0 && i < g && j >= 0 && j < g)
                    v = *reinterpret_cast<const double2*>(Vin + (int64_t)(colbase + i * g + j) * KP + jb0 + 2 * c2);
                *reinterpret_cast<double2*>(halo + hc * CS + 2 * c2) = v;
            }
            if (goals) {
                for (int hc = tid; hc < NREC; hc += kHThreads) {
                    const int hi = hc / HW, hj = hc - hi * HW;
                    const int cell = (hi0 + hi) * g + hj0 + hj;
                    uint32_t mask = 0u;
#pragma unroll
                    for (int c = 0; c < KB; c++) mask |= (goal_s[c] == cell ? 1u : 0u) << c;
                    reinterpret_cast<uint32_t*>(halo + hc * CS + KB)[0] = mask;
                }
            }
        }
        __syncthreads();
        if (!act) continue;  // block-uniform per wave; the next tile's barrier waits for all
        // does a goal of the block's columns lie in the halo (then the walk reads the masks)
        int gt = 0;
        if (goals) {
#pragma unroll
            for (int c = 0; c < KB; c++)
                gt |= (goal_i[c] >= 0 && (unsigned)(goal_i[c] - hi0) < (unsigned)HH &&
                       (unsigned)(goal_j[c] - hj0) < (unsigned)HW) ? 1 : 0;
        }
        const bool goal_tile = __builtin_amdgcn_readfirstlane(gt) != 0;
        const float delta = (float)((double)A.r_term - (t > A.nt - 20 ? A.late_pen : 0.0));
        auto decode = [&](int2 e) {
            const int cc = e.x - colbase;
            const int i2 = (int)__umulhi((uint32_t)cc, A.gmagic);
            const int hi = i2 - hi0, hj = cc - i2 * g - hj0;
            const bool in = (unsigned)hi < (unsigned)HH && (unsigned)hj < (unsigned)HW;
            return make_int2(in ? (hi * HW + hj) * CS : (int)(0x80000000u | (uint32_t)e.x), e.y);
        };
        for (int gi = 0; gi < ngroups; gi++) {
            // requested now, used after this group's walk: the next group's row data, the
            // end of the group after it, and the old values of this group's first (state,
            // column) pass
            const Meta nxt = gi + 1 < ngroups ? load_meta(gi + 1) : cur;
            const uint32_t G3 = gi + 3 <= ngroups ? (uint32_t)__builtin_amdgcn_readlane((int)gsv, gi + 3) : G2;
            const int qv = gi * spg + lane / KQ;
            const double vold = (live_q[0] && lane / KQ < spg && qv < tw) ? Vin[(s_run + qv) * KP + jb0 + cq] : 0.0;
            const int q = gi * spg + sl;
            const bool valid = q < tw && a < na;
            const int64_t row = (s_run + min(q, tw - 1)) * na + min(a, na - 1);
            double sd[KB];
            int hit[KB];
#pragma unroll
            for (int c = 0; c < KB; c++) {
                sd[c] = 0.0;
                hit[c] = -1;
            }
            for (;;) {
                const uint32_t w0 = nws, wend = nwe;
                // stage: decode; an entry landing outside the halo takes the next overflow
                // record; when none is left the window ends at that entry (the first entry
                // outside the halo always has one, so every window holds an entry)
                int novf = 0;           // wave-uniform
                uint32_t wcut = wend;   // wave-uniform: the window's end after staging
                bool ovf_goal = false;  // an overflow record holds a goal of the block's columns
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (w0 + (uint32_t)(u * 64) >= wcut) break;  // wave-uniform: no entry left
                    const uint32_t idx = w0 + (uint32_t)(u * 64 + lane);
                    int2 d = decode(e[u]);
                    const bool far = idx < wcut && d.x < 0;
                    const uint64_t m = __ballot(far);
                    if (m) {
                        const int slot = novf + (int)__popcll(m & below);
                        const uint64_t mo = __ballot(far && slot >= kTOvf);
                        if (mo) wcut = w0 + (uint32_t)(u * 64 + __builtin_ctzll(mo));
                        if (far && slot < kTOvf) {
                            const int col = d.x & 0x7fffffff;
                            const int r0 = ovf0 + slot * CS;
                            const double2* xp = reinterpret_cast<const double2*>(Vin + (int64_t)col * KP + jb0);
#pragma unroll
                            for (int c2 = 0; c2 < KB / 2; c2++)
                                *reinterpret_cast<double2*>(halo + r0 + 2 * c2) = xp[c2];
                            uint32_t mask = 0u;
                            if (goals) {
#pragma unroll
                                for (int c = 0; c < KB; c++) mask |= (goal_s[c] == col - colbase ? 1u : 0u) << c;
                            }
                            reinterpret_cast<uint32_t*>(halo + r0 + KB)[0] = mask;
                            ovf_goal = ovf_goal || mask != 0u;
                            d.x = r0;
                        }
                        novf += (int)__popcll(m);
                    }
                    if (idx < wcut) wb[u * 64 + lane] = d;
                }
                const bool chk = goal_tile || __ballot(ovf_goal) != 0ull;
                // the next window: the rest of this group, else the next group's first
                const bool more = wcut < G1;
                nws = wcut;
                nwe = more ? min(wcut + (uint32_t)WIN, G1) : min(wcut + (uint32_t)WIN, G2);
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const uint32_t idx = nws + (uint32_t)(u * 64 + lane);
                    if (idx < nwe) e[u] = A.nz[idx];
                }
                __builtin_amdgcn_wave_barrier();
                const uint32_t jlo = valid ? max(cur.rb, w0) : 0u, jhi = valid ? min(cur.re, wcut) : 0u;
                if (jlo < jhi) {
                    int2 en[U2];
#pragma unroll
                    for (int u = 0; u < U2; u++) en[u] = wb[min(jlo + (uint32_t)u, jhi - 1u) - w0];
                    for (uint32_t j = jlo; j < jhi; j += U2) {
                        double2 hv[U2][KB / 2];
#pragma unroll
                        for (int u = 0; u < U2; u++) {
                            const double2* hp = reinterpret_cast<const double2*>(halo + en[u].x);
#pragma unroll
                            for (int c2 = 0; c2 < KB / 2; c2++) hv[u][c2] = hp[c2];
                        }
                        uint32_t mk[U2];
#pragma unroll
                        for (int u = 0; u < U2; u++)
                            mk[u] = chk ? reinterpret_cast<const uint32_t*>(halo + en[u].x + KB)[0] : 0u;
                        float vf[U2];
                        double pv[U2];
#pragma unroll
                        for (int u = 0; u < U2; u++) {
                            const bool inr = j + (uint32_t)u < jhi;
                            vf[u] = __int_as_float(en[u].y);
                            pv[u] = inr ? (double)vf[u] : 0.0;
                            if (!inr) mk[u] = 0u;
                        }
#pragma unroll
                        for (int u = 0; u < U2; u++) en[u] = wb[min(j + (uint32_t)(U2 + u), jhi - 1u) - w0];
#pragma unroll
                        for (int u = 0; u < U2; u++) {
#pragma unroll
                            for (int c2 = 0; c2 < KB / 2; c2++) {
                                sd[2 * c2] = sd[2 * c2] + pv[u] * hv[u][c2].x;
                                sd[2 * c2 + 1] = sd[2 * c2 + 1] + pv[u] * hv[u][c2].y;
                            }
                            if (mk[u]) {
                                const int n = (int)rintf(vf[u] * (float)A.nr);
#pragma unroll
                                for (int c = 0; c < KB; c++)
                                    if ((mk[u] >> c) & 1u) hit[c] = n;
                            }
                        }
                    }
                }
                __builtin_amdgcn_wave_barrier();
                if (!more) break;
            }
            // the rows' Q to LDS (over the entries, which are no longer read), KQ columns
            // per pass: row r's column pair c2 at slot r * KQ/2 + (c2 ^ (r & QM)), so that
            // the stores of one pair by consecutive rows, and the loads of one row by
            // consecutive columns, fall on different banks. Then lane (state, column): max
            // / first argmax over the actions in action order (strict '>', spvi_2.cu:112)
            const double* const qd = reinterpret_cast<const double*>(qb);
            // a goal column reached by a lane of a row of the wave (rare): only then the
            // rewards of the goal columns are formed per column
            bool anyhit = false;
#pragma unroll
            for (int c = 0; c < KB; c++) anyhit = anyhit || hit[c] >= 0;
            const bool wave_hit = __ballot(anyhit) != 0ull;
#pragma unroll
            for (int h = 0; h < QH; h++) {
                if (valid) {
                    double qq[KQ];
                    if (!goals) {
#pragma unroll
                        for (int c = 0; c < KQ; c++) {
                            const float rj = A.Rcol[(int64_t)min(jb0 + h * KQ + c, A.rcol_n - 1) * A.rows + row];
                            qq[c] = (double)rj + (double)A.discount * sd[h * KQ + c];
                        }
                    } else if (wave_hit) {
#pragma unroll
                        for (int c = 0; c < KQ; c++) {
                            const int hc = hit[h * KQ + c];
                            const float rj = hc >= 0 ? (cur.rs + (float)hc * delta) / A.nr : cur.rr;
                            qq[c] = (double)rj + (double)A.discount * sd[h * KQ + c];
                        }
                    } else if (A.discount == 1.0f) {  // 1 * sum is the sum, bit for bit
#pragma unroll
                        for (int c = 0; c < KQ; c++) qq[c] = (double)cur.rr + sd[h * KQ + c];
                    } else {
#pragma unroll
                        for (int c = 0; c < KQ; c++) qq[c] = (double)cur.rr + (double)A.discount * sd[h * KQ + c];
                    }
#pragma unroll
                    for (int c2 = 0; c2 < KQ / 2; c2++) {
                        double x0 = qq[2 * c2], x1 = qq[2 * c2 + 1];
                        if (x0 != x0) x0 = -INFINITY;
                        if (x1 != x1) x1 = -INFINITY;
                        qb[lane * (KQ / 2) + (c2 ^ (lane & QM))] = make_double2(x0, x1);
                    }
                }
                __builtin_amdgcn_wave_barrier();
                for (int p = lane; p < spg * KQ; p += 64) {
                    const int s2 = p / KQ;
                    const int q2 = gi * spg + s2;
                    if (q2 >= tw) continue;
                    double best = -INFINITY;
                    int ba = 0x7fffffff;
                    if (nap > QM) {  // r & QM = a2 & QM: the swizzle offsets repeat every QM + 1 rows
                        for (int m = 0; m * (QM + 1) < na; m++) {
                            const double* qm = qd + (s2 * nap + m * (QM + 1)) * KQ;
#pragma unroll
                            for (int i = 0; i <= QM; i++) {
                                const int a2 = m * (QM + 1) + i;
                                if (a2 < na) {
                                    const double x = qm[qoff[i]];
                                    if (x > best) {
                                        best = x;
                                        ba = a2;
                                    }
                                }
                            }
                        }
                    } else {
                        for (int a2 = 0; a2 < na; a2++) {
                            const int r = s2 * nap + a2;
                            const double x = qd[2 * (r * (KQ / 2) + ((cq >> 1) ^ (r & QM))) + (cq & 1)];
                            if (x > best) {
                                best = x;
                                ba = a2;
                            }
                        }
                    }
                    const int cell = cell0 + q2;
                    const int64_t o = (s_run + q2) * KP + jb0 + h * KQ + cq;
                    double vo;
                    int po;
                    if (cell == goal_q[h]) {
                        vo = 0.0;
                        po = 0;
                    } else if (best > -1e6) {
                        vo = best;
                        po = ba;
                    } else {
                        vo = -1e6;
                        po = -1;
                    }
                    if (live_q[h]) {
                        const double vprev = (h == 0 && p == lane) ? vold : Vin[o];
                        Vout[o] = vo;
                        A.pi[o] = po;
                        rmax[h] = fmaxf(rmax[h], res_up(fabs(vo - vprev)));
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
            cur = nxt;
            G1 = G2;
            G2 = G3;
        }
    }
    if (A.one_pass) return;  // block-uniform
#pragma unroll
    for (int h = 0; h < QH; h++) {
        for (int off = KQ; off < 64; off <<= 1) rmax[h] = fmaxf(rmax[h], __shfl_xor(rmax[h], off, 64));
        if (lane < KQ) red[wv][h * KQ + lane] = rmax[h];
    }
    sweep_tail<KB, kHThreads>(A, red, fin, is_last, k);
}

// Streaming sweep with transposed gathers. The state partition and the entry stream
// are those of the plain lean sweep: each wave takes a contiguous state range (wpart,
// balanced by entries + rows), lanes (a, sl) = (action, state) of groups of SPC = 64/NAP
// states, and the entries are read as one stream in windows of WIN = 64*U. Per window:
//   * gathers: lane (q, c) = (lane / KB, lane % KB) takes entry q of each run of
//     E = 64/KB entries and column c, so the KB lanes of an entry read its KB adjacent
//     values of V (one line) together: one cache line per entry, as in the plain sweep,
//     whatever KB is. Products P*V(c) go to LDS [entry][column];
//   * rows: each lane adds its row's products of the window, per column in ascending
//     entry order (the order, and so the bits, of the plain sweep);
//   * groups end as in the plain sweep: max / first argmax over the NAP lanes of a state
//     per column, one writer lane per state and column.
// The residual of each column goes through a per-column atomic maximum (colmax, float
// bits, all residuals >= 0); the last block of the grid reads and clears it.
constexpr int kSThreads = 256;
constexpr int kSWaves = kSThreads / 64;
constexpr int kSRowReads = 2;  // entries a row lane reads per step (the LDS pad of a window)

template <int NAP, int KB, int U>
__global__ __launch_bounds__(kSThreads) void spmm_vi_stream(const BatchArgs A) {
    constexpr int SPC = 64 / NAP;
    constexpr int WIN = 64 * U;
    constexpr int E = 64 / KB;    // entries per gather instruction
    constexpr int NIT = WIN / E;  // gather instructions per window
    constexpr int BAT = NIT < 8 ? NIT : 8;
    static_assert(NIT % BAT == 0, "gather batches");
    __shared__ __align__(16) double stage[kSWaves][(WIN + kSRowReads) * KB];
    __shared__ int2 hitc[kSWaves][WIN + kSRowReads];  // goal columns: (column, count) of each entry
    __shared__ float red[kSWaves][KB];
    __shared__ int sgoal[KB];
    __shared__ int slive[KB];
    __shared__ int is_last;
    BatchState* st = A.st;
    if (st->all_conv) return;
    const uint32_t k = st->k;
    const double* __restrict__ Vin = (k & 1u) ? A.V0 : A.V1;
    double* __restrict__ Vout = (k & 1u) ? A.V1 : A.V0;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int a = lane % NAP, sl = lane / NAP;
    const int gq = lane / KB, gc = lane % KB;
    const uint32_t na = (uint32_t)A.na, nc = (uint32_t)A.nc;
    const int KP = A.kp;
    const int jb0 = blockIdx.y * KB;  // first column of this block
    const bool goals = A.Rcol == nullptr;
    if (tid < KB) {
        sgoal[tid] = A.goal[jb0 + tid];
        slive[tid] = A.conv[jb0 + tid] < 0;
    }
    __syncthreads();
    int goal[KB];
    float rmax[KB];
#pragma unroll
    for (int c = 0; c < KB; c++) {
        goal[c] = sgoal[c];
        rmax[c] = 0.0f;
    }
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned pb = (nb % 8u == 0u) ? (b % 8u) * (nb / 8u) + b / 8u : b;
    const unsigned w = pb * kSWaves + wv;
    const uint32_t* wpart = A.wpart;  // a built model's sweep: the active slices' partition
    if (A.pre_nt) wpart += (size_t)(sweep_active(A, k, blockIdx.y * KB, KB) - 1) * A.pre_stride;
    const uint32_t s_begin = wpart[w], s_end = wpart[w + 1];
    double* wst = stage[wv];
    int2* whc = hitc[wv];

    if (s_begin < s_end) {
        const uint32_t ngroups = (s_end - s_begin + SPC - 1) / SPC;
        const uint32_t Pb = A.rp[s_begin * na], Pend = A.rp[s_end * na];
        const uint32_t nwin = Pend > Pb ? (Pend - Pb + WIN - 1) / WIN : 0u;
        struct Meta {
            uint32_t rb, re;
            float rr, rs;
        };
        auto load_meta = [&](uint32_t g, Meta& M) {
            const uint32_t s = min(s_begin + g * SPC + (uint32_t)sl, s_end - 1u);
            const uint32_t row = s * na + min((uint32_t)a, na - 1u);
            M.rb = A.rp[row];
            M.re = A.rp[row + 1];
            M.rr = A.R[row];
            M.rs = goals ? A.Rsum[row] : 0.0f;
        };
        Meta cur, nxt;
        uint32_t g = 0;
        load_meta(0, cur);
        load_meta(min(1u, ngroups - 1u), nxt);
        uint32_t ce = A.rp[min(s_begin + SPC, s_end) * na];
        bool valid = (s_begin + (uint32_t)sl < s_end) && ((uint32_t)a < na);
        double sd[KB];
        int hit[KB];
#pragma unroll
        for (int c = 0; c < KB; c++) {
            sd[c] = 0.0;
            hit[c] = -1;
        }
        // the lane's state in group g, its slice's next-slice base, and per column the
        // goal column index of that base (hits are entries with column gbase + goal[c])
        auto group_base = [&](uint32_t gg) {
            const uint32_t s = min(s_begin + gg * SPC + (uint32_t)sl, s_end - 1u);
            return (s / nc + 1u) * nc;
        };
        uint32_t colbase = group_base(0);

        // products of entries [jb, je) of the window starting at entry w0, ascending
        // (absolute entry indices: a row that ended before the window gives je < jb)
        auto row_part = [&](uint32_t w0, uint32_t jb, uint32_t je) {
            while (jb < je) {
                const uint32_t r = jb - w0;
                double x[kSRowReads][KB];
                int2 h[kSRowReads];
#pragma unroll
                for (int i = 0; i < kSRowReads; i++) {
                    if (KB == 1) {
                        x[i][0] = wst[r + i];
                    } else {
#pragma unroll
                        for (int c = 0; c < KB; c += 2) {
                            const double2 v = *reinterpret_cast<const double2*>(wst + (r + i) * KB + c);
                            x[i][c] = v.x;
                            x[i][c + 1] = v.y;
                        }
                    }
                    if (goals) h[i] = whc[r + i];
                }
                const uint32_t cnt = je - jb;
#pragma unroll
                for (int i = 0; i < kSRowReads; i++) {
                    if ((uint32_t)i < cnt) {
#pragma unroll
                        for (int c = 0; c < KB; c++) sd[c] = sd[c] + x[i][c];
                        if (goals) {
                            const int cc = h[i].x - (int)colbase;
#pragma unroll
                            for (int c = 0; c < KB; c++)
                                if (cc == goal[c]) hit[c] = h[i].y;
                        }
                    }
                }
                jb += kSRowReads;
            }
        };

        auto finish_group = [&]() {
            const uint32_t s = s_begin + g * SPC + (uint32_t)sl;
            const bool own = g < ngroups && s < s_end;
            const uint32_t sc = min(s, s_end - 1u);
            const uint32_t t = sc / nc;
            const int cell = (int)(sc - t * nc);
            const float delta = (float)((double)A.r_term - ((int)t > A.nt - 20 ? A.late_pen : 0.0));
            const uint32_t row = sc * na + min((uint32_t)a, na - 1u);
            const bool writer = (NAP <= 16) ? (a == NAP - 1) : (a == 0);
#pragma unroll
            for (int c = 0; c < KB; c++) {
                const float rj = !goals ? A.Rcol[(int64_t)min(jb0 + c, A.rcol_n - 1) * A.rows + row]
                                 : hit[c] >= 0 ? (cur.rs + (float)hit[c] * delta) / A.nr
                                               : cur.rr;
                double q = valid ? (double)rj + (double)A.discount * sd[c] : -INFINITY;
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
                if (writer && own && slive[c]) {
                    double vo;
                    int po;
                    if (cell == goal[c]) {
                        vo = 0.0;
                        po = 0;
                    } else if (q > -1e6) {
                        vo = q;
                        po = ba;
                    } else {
                        vo = -1e6;
                        po = -1;
                    }
                    const int64_t o = (int64_t)s * KP + jb0 + c;
                    Vout[o] = vo;
                    A.pi[o] = po;
                    rmax[c] = fmaxf(rmax[c], res_up(fabs(vo - Vin[o])));
                }
                sd[c] = 0.0;
                hit[c] = -1;
            }
            g++;
            cur = nxt;
            load_meta(min(g + 1u, ngroups - 1u), nxt);
            const uint32_t s0 = s_begin + g * SPC;
            ce = A.rp[min(s0 + SPC, s_end) * na];
            valid = (g < ngroups) && (s0 + (uint32_t)sl < s_end) && ((uint32_t)a < na);
            colbase = group_base(min(g, ngroups - 1u));
        };

        for (uint32_t kw = 0; kw < nwin; kw++) {
            const uint32_t w0 = Pb + kw * WIN, wend = w0 + WIN;
            // entries w0 .. wend-1: those past the wave's range are loaded (the entry
            // array has a zeroed tail of kNzPad >= WIN entries and every column is a
            // state < ns) but no row of this wave reads them
#pragma unroll
            for (int i0 = 0; i0 < NIT; i0 += BAT) {
                int2 e[BAT];
                double x[BAT];
#pragma unroll
                for (int i = 0; i < BAT; i++) e[i] = A.nz[w0 + (uint32_t)((i0 + i) * E + gq)];
#pragma unroll
                for (int i = 0; i < BAT; i++) x[i] = Vin[(int64_t)e[i].x * KP + jb0 + gc];
#pragma unroll
                for (int i = 0; i < BAT; i++) {
                    const float v = __int_as_float(e[i].y);
                    wst[((i0 + i) * E + gq) * KB + gc] = (double)v * x[i];
                    if (goals && gc == 0) whc[(i0 + i) * E + gq] = make_int2(e[i].x, (int)rintf(v * (float)A.nr));
                }
            }
            __builtin_amdgcn_wave_barrier();
            for (;;) {
                if (valid) row_part(w0, max(cur.rb, w0), min(cur.re, wend));
                if (ce > wend) break;  // the group continues in the next window
                finish_group();
                if (g >= ngroups) break;
            }
            __builtin_amdgcn_wave_barrier();
        }
        while (g < ngroups) finish_group();  // trailing groups whose rows are all empty
    }

    // per-column residual: the wave's lanes, then the block's waves, then one atomic
    // maximum per column and block
#pragma unroll
    for (int c = 0; c < KB; c++) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) rmax[c] = fmaxf(rmax[c], __shfl_xor(rmax[c], o, 64));
    }
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < KB; c++) red[wv][c] = rmax[c];
    }
    __syncthreads();
    // per-column atomic maxima, then the ticket, ordered by the maxima's returned values
    // (as in sweep_tail; no __threadfence)
    uint32_t one = 1u;
    if (tid < KB) {
        float m = 0.0f;
#pragma unroll
        for (int i = 0; i < kSWaves; i++) m = fmaxf(m, red[i][tid]);
        const uint32_t prev = atomicMax(A.colmax + jb0 + tid, __float_as_uint(m));
        asm volatile("" : "+v"(one) : "v"(prev));
    }
    __syncthreads();
    if (tid == 0) is_last = atomicAdd(&st->done, one) == gridDim.x * gridDim.y - 1;
    __syncthreads();
    if (is_last) {
        int running = 0;
        for (int j = tid; j < KP; j += kSThreads) {
            const float r = __uint_as_float(atomicExch(A.colmax + j, 0u));
            if (A.conv[j] < 0) {
                A.last_res[j] = r;
                if (r < A.thresh) A.conv[j] = (int32_t)k;
                else running++;
            }
        }
        const int n_running = __syncthreads_count(running > 0);
        if (tid == 0) {
            st->all_conv = n_running == 0;
            __hip_atomic_exchange(&st->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            st->k = k + 1u;
        }
    }
}

// stream kernel for (NAP, KB); nullptr when not instantiated
template <int KB, int U>
BatchFn stream_fn_kb(int nap) {
    switch (nap) {
        case 1: return spmm_vi_stream<1, KB, U>;
        case 2: return spmm_vi_stream<2, KB, U>;
        case 4: return spmm_vi_stream<4, KB, U>;
        case 8: return spmm_vi_stream<8, KB, U>;
        case 16: return spmm_vi_stream<16, KB, U>;
        case 32: return spmm_vi_stream<32, KB, U>;
        case 64: return spmm_vi_stream<64, KB, U>;
        default: return nullptr;
    }
}

BatchFn stream_fn(int kb, int u, int nap) {
    if (kb == 1 && u == 4) return stream_fn_kb<1, 4>(nap);
    if (kb == 2 && u == 4) return stream_fn_kb<2, 4>(nap);
    if (kb == 4 && u == 2) return stream_fn_kb<4, 2>(nap);
    if (kb == 4 && u == 4) return stream_fn_kb<4, 4>(nap);
    if (kb == 8 && u == 1) return stream_fn_kb<8, 1>(nap);
    if (kb == 8 && u == 2) return stream_fn_kb<8, 2>(nap);
    return nullptr;
}

__global__ void batch_init(double* V0, int64_t n, const int32_t* goal, int32_t* conv, int32_t* start,
                           float* last_res, uint32_t* colmax, int kp, int k, BatchState* st) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) V0[i] = 0.0;
    if (i < kp) {
        colmax[i] = 0u;
        conv[i] = i < k ? -1 : 0;  // padding columns never run
        start[i] = 1;
        last_res[i] = 0.0f;
    }
    if (i < kTileRows * 8) st->tnext[i * kCounterStride] = 0u;
    if (i == 0) {
        st->k = 1u;
        st->done = 0u;
        st->all_conv = 0;
    }
}

// out[j][s] = V[conv_j parity][s][j] rounded to float once; pol[j][s] likewise from pi
// (written only while live)
__global__ void batch_gather(const double* V0, const double* V1, const int32_t* pi, const int32_t* conv,
                             int64_t ns, int kp, int k, int last_k, float* vout, int32_t* pout) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns * k) return;
    const int j = (int)(i / ns);
    const int64_t s = i - (int64_t)j * ns;
    const int cj = conv[j] >= 0 ? conv[j] : last_k;
    const double* V = (cj & 1) ? V1 : V0;
    vout[i] = (float)V[s * kp + j];
    pout[i] = pi[s * kp + j];
}

// kind 0: spmm_vi_sweep (lc = 64/spw lanes per state); kind 1: spmm_vi_rows (lc lanes
// per row); kind 2: spmm_vi_stream (cpl = KB columns per block, u = entries per lane
// and window); kind 3: spmm_vi_halo (lc lanes per row, built models only); kind 4:
// spmm_vi_tile (cpl = KB columns per row lane, u = window entries per lane, built models
// only). Kinds 1-4 take a number of states per wave that depends on na.
struct Variant {
    int cpl, lc;
    int kind;
    int u;
    BatchFn fn;  // kinds 0, 1, 3 and 4
};

const Variant kVariants[] = {
    {1, 64, 0, 0, spmm_vi_sweep<1, 1>},  {2, 32, 0, 0, spmm_vi_sweep<2, 2>},
    {4, 16, 0, 0, spmm_vi_sweep<4, 4>},  {1, 32, 0, 0, spmm_vi_sweep<1, 2>},
    {2, 16, 0, 0, spmm_vi_sweep<2, 4>},  {4, 8, 0, 0, spmm_vi_sweep<4, 8>},
    {1, 16, 0, 0, spmm_vi_sweep<1, 4>},  {2, 8, 0, 0, spmm_vi_sweep<2, 8>},
    {4, 4, 0, 0, spmm_vi_sweep<4, 16>},  {1, 4, 1, 0, spmm_vi_rows<4, 1, 4>},
    {1, 8, 1, 0, spmm_vi_rows<8, 1, 4>},  {2, 4, 1, 0, spmm_vi_rows<4, 2, 4>},
    {2, 8, 1, 0, spmm_vi_rows<8, 2, 4>},  {4, 4, 1, 0, spmm_vi_rows<4, 4, 4>},
    {1, 16, 1, 0, spmm_vi_rows<16, 1, 4>}, {4, 8, 1, 0, spmm_vi_rows<8, 4, 4>},
    {2, 16, 1, 0, spmm_vi_rows<16, 2, 4>}, {4, 1, 1, 0, spmm_vi_rows<1, 4, 4>},
    {8, 1, 1, 0, spmm_vi_rows<1, 8, 2>},  {4, 2, 1, 0, spmm_vi_rows<2, 4, 4>},
    {8, 2, 1, 0, spmm_vi_rows<2, 8, 2>},  {4, 1, 2, 2, nullptr},
    {4, 1, 2, 4, nullptr},                {8, 1, 2, 1, nullptr},
    {8, 1, 2, 2, nullptr},                {1, 1, 2, 4, nullptr},
    {2, 1, 2, 4, nullptr},
    {4, 4, 3, 0, spmm_vi_halo<4, 4, 4>},  {4, 2, 3, 0, spmm_vi_halo<2, 4, 4>},
    {2, 4, 3, 0, spmm_vi_halo<4, 2, 4>},  {4, 1, 3, 0, spmm_vi_halo<1, 4, 4>},
    {8, 2, 3, 0, spmm_vi_halo<2, 8, 4>},  {2, 2, 3, 0, spmm_vi_halo<2, 2, 4>},
    {8, 1, 4, 16, spmm_vi_tile<8, 16, 1, 4, 1>},  {16, 1, 4, 12, spmm_vi_tile<16, 12, 1, 4, 1>},
    {8, 1, 4, 8, spmm_vi_tile<8, 8, 2, 5, 1>},   {16, 1, 4, 16, spmm_vi_tile<16, 16, 1, 4, 1>},
    {16, 1, 4, 8, spmm_vi_tile<16, 8, 1, 4, 1>},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

int kb_of(const Variant& v) { return v.lc * v.cpl; }

int nap_of(int na) {
    int p = 1;
    while (p < na) p <<= 1;
    return p;
}

// states per wavefront; 0 when the variant cannot take this many actions
int spw_of(const Variant& v, int na) {
    if (v.kind == 0) return 64 / v.lc;
    if (v.kind == 2 || v.kind == 4) return 64 / nap_of(na);
    const int lanes = v.lc * nap_of(na);
    return lanes <= 64 ? 64 / lanes : 0;
}

bool variant_fits(const Variant& v, int na, int kp) { return spw_of(v, na) > 0 && kp % kb_of(v) == 0; }

// the halo kernel needs a built model (cells of slice t move to cells of slice t+1, on a
// square grid of at most 1024 cells a side, where umulhi by ceil(2^32 / g) divides exactly)
bool halo_model(const Model* m) {
    return m->built && m->nt > 0 && m->p.gsize >= 2 && m->p.gsize <= 1024 &&
           m->nc == (int64_t)m->p.gsize * m->p.gsize && m->ns == m->nt * m->nc;
}

bool variant_ok(const Model* m, const Variant& v, int kp) {
    return variant_fits(v, (int)m->na, kp) && ((v.kind != 3 && v.kind != 4) || halo_model(m));
}

// padded column count: 4 or 8 for small batches (the value block's stride per state is
// then 32 or 64 bytes, so the gathers of a few columns touch fewer lines), else a
// multiple of 16
int padded_cols(int k) { return k <= 4 ? 4 : k <= 8 ? 8 : (int)ceil_div(k, 16) * 16; }

// per-block residual rows of the partials buffer: the largest grid of any kernel
constexpr int kPartRows = kHMaxBlocks > kBMaxBlocks ? kHMaxBlocks : kBMaxBlocks;

struct BatchBufs {
    int kp = 0;
    int64_t ns = 0;
    double* V[2] = {nullptr, nullptr};
    int32_t* pi = nullptr;
    int32_t* goal = nullptr;
    int32_t* conv = nullptr;
    int32_t* start = nullptr;  // sweep index at which each column was admitted
    float* part = nullptr;     // [kPartRows][kp]
    float* last_res = nullptr;
    BatchState* st = nullptr;
    float* vout = nullptr;
    int32_t* pout = nullptr;
    uint32_t* colmax = nullptr;  // [kp] (spmm_vi_stream)
    uint32_t* wpart = nullptr;   // [wpart_W + 1] state partition of spmm_vi_stream
    int64_t wpart_W = 0;
    int wpart_spc = 0;
    uint32_t* wpre = nullptr;    // [nt][wpart_W + 1] the same partition of each prefix of slices
    int64_t wpre_W = 0;
    int wpre_spc = 0;
    void release() {
        void* ps[] = {V[0], V[1], pi, goal, conv, start, part, last_res, st, vout, pout, colmax, wpart, wpre};
        for (void* p : ps)
            if (p) pool_free(p);
        *this = BatchBufs();
    }
};

}  // namespace

struct BatchCache {
    BatchBufs b;
    int tuned_kp = -1;
    int tuned_variant = -1;
    hipGraphExec_t graph = nullptr;
    int graph_variant = -1, graph_chunk = 0, graph_grid = 0, graph_gy = 0;
    float graph_thresh = 0.f, graph_discount = 0.f;
    const float* graph_rcol = nullptr;  // goal columns (NULL) or the alpha rewards
    const uint32_t* graph_wpart = nullptr;
    uint32_t graph_pre = 0;
};

void batch_cache_free(BatchCache* c) {
    if (!c) return;
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    c->b.release();
    delete c;
}

static e2e_status alloc_bufs(const Model* m, BatchBufs& b, int kp) {
    const size_t n = (size_t)m->ns * kp;
    if (pool_malloc(&b.V[0], n * 8) != hipSuccess || pool_malloc(&b.V[1], n * 8) != hipSuccess ||
        pool_malloc(&b.pi, n * 4) != hipSuccess || pool_malloc(&b.goal, kp * 4) != hipSuccess ||
        pool_malloc(&b.conv, kp * 4) != hipSuccess || pool_malloc(&b.start, kp * 4) != hipSuccess ||
        pool_malloc(&b.part, (size_t)kPartRows * kp * 4) != hipSuccess ||
        pool_malloc(&b.last_res, kp * 4) != hipSuccess || pool_malloc(&b.st, sizeof(BatchState)) != hipSuccess ||
        pool_malloc(&b.vout, n * 4) != hipSuccess || pool_malloc(&b.pout, n * 4) != hipSuccess ||
        pool_malloc(&b.colmax, kp * 4) != hipSuccess) {
        b.release();
        set_error("batched solve: device allocation failed");
        return E2E_ERR_NOMEM;
    }
    b.kp = kp;
    b.ns = m->ns;
    return E2E_OK;
}

static e2e_status ensure_bufs(Model* m, BatchCache* c, int kp) {
    BatchBufs& b = c->b;
    if (b.kp == kp && b.ns == m->ns) return E2E_OK;
    b.release();
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    return alloc_bufs(m, b, kp);
}

static BatchArgs make_args(const Model* m, const BatchBufs& b, const e2e_solve_opts& o) {
    BatchArgs A{};
    A.rp = m->row_ptr;
    A.nz = m->nz;
    A.R = m->R;
    A.Rsum = m->Rsum;
    A.V0 = b.V[0];
    A.V1 = b.V[1];
    A.pi = b.pi;
    A.goal = b.goal;
    A.conv = b.conv;
    A.part = b.part;
    A.last_res = b.last_res;
    A.st = b.st;
    A.ns = m->ns;
    A.nc = (int)m->nc;
    A.na = (int)m->na;
    A.nt = (int)m->nt;
    A.kp = b.kp;
    A.nr = m->p.nrzns;
    A.nap = nap_of((int)m->na);
    A.thresh = o.thresh;
    A.r_term = m->p.r_terminal;
    A.discount = o.discount;
    A.late_pen = 0.05 * (double)m->p.r_outbound;
    A.Rcol = nullptr;
    A.rows = m->rows;
    A.rcol_n = 1;
    A.colmax = b.colmax;
    A.wpart = b.wpart;
    A.start = b.start;
    A.g = m->p.gsize;
    A.gmagic = m->p.gsize >= 2 ? (uint32_t)(((1ull << 32) + (uint64_t)m->p.gsize - 1) / (uint64_t)m->p.gsize) : 0u;
    return A;
}

// grid.y covers the column blocks that hold columns 0 .. ncols-1
static dim3 grid_of(const Model* m, const Variant& v, int kp, int ncols) {
    const int64_t states_per_block = (int64_t)kBWaves * spw_of(v, (int)m->na);
    int64_t gx = ceil_div(m->ns, states_per_block);
    gx = std::min<int64_t>(gx, kBMaxBlocks);  // grid-stride beyond this
    const int64_t gy = std::min<int64_t>(kp / kb_of(v), ceil_div((int64_t)std::max(ncols, 1), (int64_t)kb_of(v)));
    return dim3((unsigned)gx, (unsigned)gy);
}

// Kernel by the width a sweep has to cover (live batches: highest live slot + 1; large
// batched solves: the column count); all kernels share the layout and the sweep state, so
// the choice may change from sweep to sweep. Measured on a 150x150 x 125 x 8-heading model
// (1.8e8 entries; plain sweep 0.39 ms), ms per sweep over every state:
//   width   1-2: spmm_vi_stream<8, 2, 4> 0.95;
//   width     4: spmm_vi_halo<1, 4> 1.38 (spmm_vi_rows<1, 4, 4> 1.59);
//   width     8: spmm_vi_tile<8, 8, 2, 5, 1> 1.33 (spmm_vi_halo<2, 4> 1.98);
//   width 16/32: spmm_vi_tile<16, 16, 1, 4, 1> 1.99 / 4.03 (spmm_vi_halo<2, 8> 3.45 / 6.90).
// The halo and tile kernels need a built model; where one does not fit, the second table's
// kernel runs, and where that does not fit either, spmm_vi_sweep<1, 4> (fits every model).
constexpr int kLiveWidths[] = {2, 4, 8, 1 << 30};
constexpr int kLiveVariants[] = {26, 30, 35, 36};
constexpr int kLiveAlt[] = {26, 17, 20, 20};
constexpr int kLiveFallback = 6;  // spmm_vi_sweep<1, 4>: fits every model
constexpr int kLiveN = 4;

// index into kLiveWidths for a sweep over columns 0 .. ncols-1
int width_class(int ncols) {
    int i = 0;
    while (i < kLiveN - 1 && ncols > kLiveWidths[i]) i++;
    return i;
}

// the variant of width class i on this model
int width_variant(const Model* m, int kp, int i) {
    if (variant_ok(m, kVariants[kLiveVariants[i]], kp)) return kLiveVariants[i];
    if (variant_ok(m, kVariants[kLiveAlt[i]], kp)) return kLiveAlt[i];
    return kLiveFallback;
}

// batched solves on models of at least this many entries and at most 16 columns take the
// width choice instead of timing every variant (the timing would cost more than it saves)
constexpr int64_t kTuneMaxNnz = 16ll << 20;

// A variant's launch on a model: kernel, block size, grid (grid.y for all kp columns;
// launches may cut it to the blocks that hold live columns). spmm_vi_stream needs the
// state partition of its grid, computed here once per (waves, group size).
struct Plan {
    BatchFn fn = nullptr;
    dim3 grid, block;
    int kb = 0;
};

static e2e_status plan_of(const Model* m, BatchBufs& b, const Variant& v, hipStream_t s, Plan* p) {
    p->kb = kb_of(v);
    if (v.kind == 3 || v.kind == 4) {  // persistent blocks over the tiles of the swept slices
        E2E_CHECK(halo_model(m), E2E_ERR_INVALID,
                  "batched sweep: the halo kernel needs a model built by e2e_model_build (grid side <= 1024)");
        int bpc = 0, ncu = 0, dev = 0;
        E2E_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, (const void*)v.fn, kHThreads, 0));
        E2E_HIP(hipGetDevice(&dev));
        E2E_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const int64_t g = m->p.gsize;
        const int64_t tiles = ceil_div(g, (int64_t)kHTileH) * ceil_div(g, (int64_t)kHTileW) * m->nt;
        const int64_t nblocks = std::max<int64_t>(
            1, std::min<int64_t>({(int64_t)std::max(bpc, 1) * std::max(ncu, 1), tiles, (int64_t)kHMaxBlocks}));
        p->fn = v.fn;
        p->block = dim3(kHThreads);
        p->grid = dim3((unsigned)nblocks, (unsigned)(b.kp / p->kb));
        return E2E_OK;
    }
    if (v.kind != 2) {
        p->fn = v.fn;
        p->block = dim3(kBThreads);
        p->grid = grid_of(m, v, b.kp, b.kp);
        return E2E_OK;
    }
    const int nap = nap_of((int)m->na);
    p->fn = stream_fn(v.cpl, v.u, nap);
    E2E_CHECK(p->fn, E2E_ERR_INVALID, "batched sweep: no streaming kernel for %d columns", v.cpl);
    E2E_CHECK(m->ns * (int64_t)m->na < (1ll << 32), E2E_ERR_CAPACITY,
              "batched sweep: the streaming kernel indexes rows with 32 bits");
    int bpc = 0, ncu = 0, dev = 0;
    E2E_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, (const void*)p->fn, kSThreads, 0));
    E2E_HIP(hipGetDevice(&dev));
    E2E_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int spc = 64 / nap;
    int64_t nblocks = (int64_t)std::max(bpc, 1) * std::max(ncu, 1);
    nblocks = std::max<int64_t>(1, std::min<int64_t>(nblocks, ceil_div(m->ns, (int64_t)kSWaves * spc)));
    const int64_t W = nblocks * kSWaves;
    if (b.wpart_W != W || b.wpart_spc != spc) {
        if (b.wpart) pool_free(b.wpart);
        b.wpart = nullptr;
        b.wpart_W = 0;
        if (pool_malloc(&b.wpart, sizeof(uint32_t) * (W + 1)) != hipSuccess) {
            b.wpart = nullptr;
            set_error("batched sweep: device allocation failed");
            return E2E_ERR_NOMEM;
        }
        launch_wave_partition(m->row_ptr, m->ns, (int)m->na, W, spc, b.wpart, s);
        E2E_HIP(hipGetLastError());
        b.wpart_W = W;
        b.wpart_spc = spc;
    }
    p->block = dim3(kSThreads);
    p->grid = dim3((unsigned)nblocks, (unsigned)(b.kp / p->kb));
    return E2E_OK;
}

// p's grid cut to the column blocks that hold columns 0 .. ncols-1
static dim3 plan_grid(const Plan& p, int ncols) {
    dim3 g = p.grid;
    g.y = (unsigned)std::min<int64_t>(g.y, ceil_div((int64_t)std::max(ncols, 1), (int64_t)p.kb));
    return g;
}

// after a backward pass: every running column is done, its values in V0 (an even conv)
__global__ void live_pass_done(int32_t* conv, int kp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kp && conv[i] < 0) conv[i] = 2;
}


// One exact backward pass over columns 0 .. ncols-1 of b (kernel_variant 7's schedule with
// the batched columns' arithmetic): spmm_vi_rows once per slice, from the last to the
// first, each launch reading the values of the slice after it, which the previous launch
// wrote into V0. The running columns (conv < 0) then hold their fixed points in V0 and are
// marked done with an even conv.
static e2e_status backward_pass(const Model* m, BatchBufs& b, const BatchArgs& A, int ncols, hipStream_t s) {
    const int kp = b.kp;
    // the width table's kernel where it takes one slice per launch (the row and halo
    // kernels), else the row kernels
    int vi = width_variant(m, kp, width_class(ncols));
    if (kVariants[vi].kind != 1 && kVariants[vi].kind != 3 && kVariants[vi].kind != 4)
        vi = (ncols > 4 && variant_fits(kVariants[20], (int)m->na, kp)) ? 20 : 17;
    const Variant& v = kVariants[vi];
    Plan p;
    E2E_TRY(plan_of(m, b, v, s, &p));
    const int spw = spw_of(v, (int)m->na);
    const int64_t nc = m->nc, groups = ceil_div(nc, (int64_t)spw);
    const int64_t g = m->p.gsize;
    const int64_t gx = (v.kind == 3 || v.kind == 4) ? std::min<int64_t>(ceil_div(g, (int64_t)kHTileH) * ceil_div(g, (int64_t)kHTileW),
                                                       (int64_t)p.grid.x)
                                   : std::min<int64_t>(ceil_div(groups, (int64_t)kBWaves), kBMaxBlocks);
    const dim3 grid((unsigned)gx, plan_grid(p, ncols).y);
    BatchArgs B = A;
    B.one_pass = 1;
    B.V1 = B.V0;
    for (int64_t t = m->nt - 1; t >= 0; t--) {
        B.s_lo = (uint32_t)(t * nc);
        B.s_hi = (uint32_t)((t + 1) * nc);
        hipLaunchKernelGGL(p.fn, grid, p.block, 0, s, B);
    }
    hipLaunchKernelGGL(live_pass_done, dim3((unsigned)ceil_div(kp, 64)), dim3(64), 0, s, b.conv, kp);
    E2E_HIP(hipGetLastError());
    return E2E_OK;
}

// Goal columns (goals != NULL) or alpha columns (goals == NULL: the model's n_alpha
// reward vectors, the model's own terminal cells).
e2e_status vi_solve(Model* m, const e2e_solve_opts* o, uint32_t* policy, float* value, e2e_solve_stats* stats);

// Alpha columns one at a time: the plain solve (the default sweep kernel) with R set to
// each column's rewards. Every column is the plain solve of the model built with that
// alpha, bit for bit.
// A built model's transitions go from each slice to the next (the sweeps over the
// active slices need that); the state count is then nt * nc.
static bool active_ok(const Model* m) {
    return m->built && m->nt > 0 && m->nc > 0 && m->ns == m->nt * m->nc;
}

// The streaming kernel over the active slices: the partition of every prefix of slices
// (b.wpre, once per model and grid), in place of the full partition.
static e2e_status use_prefix_parts(const Model* m, BatchBufs& b, BatchArgs& A, hipStream_t s) {
    if (b.wpre_W != b.wpart_W || b.wpre_spc != b.wpart_spc) {
        if (b.wpre) pool_free(b.wpre);
        b.wpre = nullptr;
        b.wpre_W = 0;
        if (pool_malloc(&b.wpre, sizeof(uint32_t) * (b.wpart_W + 1) * m->nt) != hipSuccess) {
            b.wpre = nullptr;
            set_error("batched sweep: device allocation failed");
            return E2E_ERR_NOMEM;
        }
        launch_prefix_partitions(m->row_ptr, m->nc, (int)m->na, m->nt, b.wpart_W, b.wpart_spc, b.wpre, s);
        E2E_HIP(hipGetLastError());
        b.wpre_W = b.wpart_W;
        b.wpre_spc = b.wpart_spc;
    }
    A.wpart = b.wpre;
    A.pre_stride = (uint32_t)(b.wpart_W + 1);
    return E2E_OK;
}

static e2e_status vi_solve_alpha_columns(Model* m, const e2e_solve_opts* o, uint32_t* policy, float* value,
                                         int32_t* col_sweeps, e2e_batch_stats* stats, bool backward) {
    const double t_start = now_ms();
    float* const R_own = m->R;
    // backward: each column by the exact backward pass (kernel_variant 7)
    e2e_solve_opts ob{};
    ob.thresh = 0.1f;
    ob.discount = 1.0f;
    if (o) ob = *o;
    if (backward) ob.kernel_variant = 7;
    e2e_status rc = E2E_OK;
    int conv = 0, maxsw = 0, launches = 0;
    double kms = 0.0, bytes = 0.0;
    for (int j = 0; j < m->n_alpha; j++) {
        m->R = m->Ralpha + (size_t)j * m->rows;
        e2e_solve_stats st{};
        const e2e_status r = vi_solve(m, &ob, policy ? policy + (size_t)j * m->ns : nullptr,
                                      value ? value + (size_t)j * m->ns : nullptr, &st);
        m->R = R_own;
        if (r != E2E_OK && r != E2E_ERR_NOT_CONVERGED && r != E2E_ERR_TIMEOUT) return r;
        if (r != E2E_OK && rc == E2E_OK) rc = r;
        conv += st.converged ? 1 : 0;
        maxsw = std::max(maxsw, (int)st.sweeps);
        if (col_sweeps) col_sweeps[j] = st.converged ? st.sweeps : -st.sweeps;
        kms += st.sweep_kernel_ms;
        launches += st.launches;
        bytes = st.bytes_per_sweep;
    }
    if (stats) {
        *stats = e2e_batch_stats{};
        stats->sweeps = maxsw;
        stats->converged = conv;
        stats->variant = backward ? -3 : -2;
        stats->cols_per_lane = 1;
        stats->states_per_wave = 0;
        stats->col_block = 1;
        stats->solve_ms = now_ms() - t_start;
        stats->sweep_kernel_ms = kms;
        stats->bytes_per_sweep = bytes;
        stats->launches = launches;
    }
    if (rc == E2E_ERR_NOT_CONVERGED) set_error("%d of %d alpha columns did not converge", m->n_alpha - conv, m->n_alpha);
    return rc;
}

static e2e_status vi_solve_batch(Model* m, const int32_t* goals, int k, const e2e_solve_opts* o,
                                 const e2e_batch_opts* bo, uint32_t* policy, float* value, int32_t* col_sweeps,
                                 e2e_batch_stats* stats) {
    const double t_start = now_ms();
    if (goals) {
        E2E_CHECK(m->built, E2E_ERR_INVALID, "batched goals need a model built by e2e_model_build");
        E2E_CHECK(m->Rsum, E2E_ERR_INVALID, "batched goals need keep_rsum=1 at build time");
        E2E_CHECK(m->p.term_subgrid_size == 0, E2E_ERR_INVALID,
                  "batched goals need the shared model built without a terminal cell (term_subgrid_size 0)");
        E2E_CHECK(!m->has_obstacles, E2E_ERR_INVALID, "batched goals are not defined for models with obstacles");
        E2E_CHECK(k >= 1 && k <= 64, E2E_ERR_INVALID, "k=%d goals must be in 1..64", k);
        for (int j = 0; j < k; j++)
            E2E_CHECK(goals[j] >= 0 && goals[j] < m->nc, E2E_ERR_INVALID, "goal %d out of range", goals[j]);
    } else {
        E2E_CHECK(m->Ralpha && m->n_alpha >= 1 && k == m->n_alpha, E2E_ERR_INVALID,
                  "alpha columns need a model built with n_alpha > 0");
        if (bo && (bo->variant == -2 || bo->variant == -3))
            return vi_solve_alpha_columns(m, o, policy, value, col_sweeps, stats, bo->variant == -3);
    }
    e2e_solve_opts opts{};
    opts.thresh = 0.1f;
    opts.discount = 1.0f;
    if (o) opts = *o;
    e2e_batch_opts bopt{};
    bopt.variant = -1;
    if (bo) bopt = *bo;
    const int kp = padded_cols(k);
    if (!m->batch) m->batch = new BatchCache();
    BatchCache* c = m->batch;
    E2E_TRY(ensure_bufs(m, c, kp));
    BatchBufs& b = c->b;
    hipStream_t s = m->stream;
    std::vector<int32_t> hg(kp, -1);
    for (int j = 0; goals && j < k; j++) hg[j] = goals[j];
    E2E_HIP(hipMemcpyAsync(b.goal, hg.data(), kp * 4, hipMemcpyHostToDevice, s));
    BatchArgs A = make_args(m, b, opts);
    if (!goals) {
        A.Rcol = m->Ralpha;
        A.rcol_n = m->n_alpha;
    }
    const int64_t nV = m->ns * kp;
    auto init = [&]() {
        hipLaunchKernelGGL(batch_init, dim3((unsigned)ceil_div(nV, 256)), dim3(256), 0, s, b.V[0], nV, b.goal,
                           b.conv, b.start, b.last_res, b.colmax, kp, k, b.st);
    };
    const int max_sweeps = opts.max_sweeps > 0 ? opts.max_sweeps : (int)m->nt + 2;
    const int spc = opts.sweeps_per_check > 0 ? opts.sweeps_per_check : 8;

    if (bopt.variant == -3) {  // every column by one exact backward pass
        hipEvent_t e0, e1;
        E2E_HIP(hipEventCreate(&e0));
        E2E_HIP(hipEventCreate(&e1));
        init();
        E2E_HIP(hipEventRecord(e0, s));
        E2E_TRY(backward_pass(m, b, A, k, s));
        E2E_HIP(hipEventRecord(e1, s));
        hipLaunchKernelGGL(batch_gather, dim3((unsigned)ceil_div(m->ns * k, 256)), dim3(256), 0, s, b.V[0], b.V[1],
                           b.pi, b.conv, m->ns, kp, k, 0, b.vout, b.pout);
        E2E_HIP(hipGetLastError());
        if (value) E2E_HIP(hipMemcpyAsync(value, b.vout, sizeof(float) * m->ns * k, hipMemcpyDeviceToHost, s));
        if (policy) E2E_HIP(hipMemcpyAsync(policy, b.pout, sizeof(int32_t) * m->ns * k, hipMemcpyDeviceToHost, s));
        E2E_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        E2E_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        for (int j = 0; col_sweeps && j < k; j++) col_sweeps[j] = 1;
        if (stats) {
            *stats = e2e_batch_stats{};
            stats->sweeps = 1;
            stats->converged = k;
            stats->variant = -3;
            stats->solve_ms = now_ms() - t_start;
            stats->sweep_kernel_ms = ms;
            stats->bytes_per_sweep = 8.0 * m->nnz + 8.0 * m->rows + 4.0 + 20.0 * (double)k * m->ns;
            stats->launches = (int32_t)m->nt;
        }
        return E2E_OK;
    }

    // ---- autotune: time a few sweeps of every valid variant on this model
    double tune_ms = 0.0;
    int variant = bopt.variant;
    hipEvent_t e0, e1;
    E2E_HIP(hipEventCreate(&e0));
    E2E_HIP(hipEventCreate(&e1));
    if (variant < 0 || variant >= kNumVariants || !variant_ok(m, kVariants[variant], kp)) {
        if (c->tuned_kp == kp && c->tuned_variant >= 0) {
            variant = c->tuned_variant;
        } else if (m->nnz >= kTuneMaxNnz && kp <= 16) {
            variant = width_variant(m, kp, width_class(k));
        } else {
            const double t0 = now_ms();
            float best = 1e30f;
            const int trial = std::min(3, max_sweeps);
            for (int vi = 0; vi < kNumVariants; vi++) {
                const Variant& v = kVariants[vi];
                if (!variant_ok(m, v, kp)) continue;
                Plan pl;
                E2E_TRY(plan_of(m, b, v, s, &pl));
                A.wpart = b.wpart;
                init();
                E2E_HIP(hipEventRecord(e0, s));
                for (int i = 0; i < trial; i++) hipLaunchKernelGGL(pl.fn, plan_grid(pl, k), pl.block, 0, s, A);
                E2E_HIP(hipEventRecord(e1, s));
                E2E_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                E2E_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    variant = vi;
                }
            }
            c->tuned_kp = kp;
            c->tuned_variant = variant;
            tune_ms = now_ms() - t0;
        }
    }
    const Variant& V = kVariants[variant];
    Plan plan;
    E2E_TRY(plan_of(m, b, V, s, &plan));
    A.wpart = b.wpart;
    const dim3 grid = plan_grid(plan, k);
    // every column starts at sweep 1: on a built model each sweep recomputes the active
    // slices only (active_slices); the streaming kernel takes each prefix's own partition
    const bool pre = !opts.full_sweeps && active_ok(m);
    if (pre) {
        A.pre_nt = (uint32_t)m->nt;
        if (V.kind == 2) E2E_TRY(use_prefix_parts(m, b, A, s));
    }

    if (bopt.use_graph &&
        (!c->graph || c->graph_variant != variant || c->graph_chunk != spc || c->graph_grid != (int)grid.x || c->graph_gy != (int)grid.y ||
         c->graph_thresh != opts.thresh || c->graph_discount != opts.discount || c->graph_rcol != A.Rcol ||
         c->graph_wpart != A.wpart || c->graph_pre != A.pre_nt)) {
        if (c->graph) {
            (void)hipGraphExecDestroy(c->graph);
            c->graph = nullptr;
        }
        hipGraph_t gr;
        E2E_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < spc; i++) hipLaunchKernelGGL(plan.fn, grid, plan.block, 0, s, A);
        E2E_HIP(hipStreamEndCapture(s, &gr));
        hipError_t e = hipGraphInstantiate(&c->graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) {
            c->graph = nullptr;
            set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
            return E2E_ERR_HIP;
        }
        c->graph_variant = variant;
        c->graph_chunk = spc;
        c->graph_grid = (int)grid.x;
        c->graph_gy = (int)grid.y;
        c->graph_thresh = opts.thresh;
        c->graph_discount = opts.discount;
        c->graph_rcol = A.Rcol;
        c->graph_wpart = A.wpart;
        c->graph_pre = A.pre_nt;
    }

    init();
    E2E_HIP(hipGetLastError());
    double kms = 0.0;
    int launched = 0;
    BatchState hst{};
    // host checks after 8, 16, 32, 64 sweeps and then every 64 (chunks 8, 8, 16, 32, 64)
    // unless sweeps_per_check fixes the chunk; launches after convergence return at once
    int chunk = spc;
    while (launched < max_sweeps) {
        const int n = std::min(chunk, max_sweeps - launched);
        E2E_HIP(hipEventRecord(e0, s));
        if (bopt.use_graph && n % spc == 0) {
            for (int r = 0; r < n / spc; r++) E2E_HIP(hipGraphLaunch(c->graph, s));
        } else {
            for (int i = 0; i < n; i++) hipLaunchKernelGGL(plan.fn, grid, plan.block, 0, s, A);
            E2E_HIP(hipGetLastError());
        }
        E2E_HIP(hipEventRecord(e1, s));
        E2E_HIP(hipMemcpyAsync(&hst, b.st, kBatchStateHead, hipMemcpyDeviceToHost, s));
        E2E_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        E2E_HIP(hipEventElapsedTime(&ms, e0, e1));
        kms += ms;
        launched += n;
        if (opts.sweeps_per_check <= 0 && launched >= 2 * spc) chunk = std::min(chunk * 2, 8 * spc);
        if (hst.all_conv) break;
        if (opts.timeout_ms > 0 && now_ms() - t_start > opts.timeout_ms) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int last_k = (int)hst.k - 1;
    hipLaunchKernelGGL(batch_gather, dim3((unsigned)ceil_div(m->ns * k, 256)), dim3(256), 0, s, b.V[0], b.V[1],
                       b.pi, b.conv, m->ns, kp, k, last_k, b.vout, b.pout);
    E2E_HIP(hipGetLastError());
    std::vector<int32_t> hconv(kp);
    std::vector<float> hres(kp);
    E2E_HIP(hipMemcpyAsync(hconv.data(), b.conv, kp * 4, hipMemcpyDeviceToHost, s));
    E2E_HIP(hipMemcpyAsync(hres.data(), b.last_res, kp * 4, hipMemcpyDeviceToHost, s));
    if (value) E2E_HIP(hipMemcpyAsync(value, b.vout, sizeof(float) * m->ns * k, hipMemcpyDeviceToHost, s));
    if (policy) E2E_HIP(hipMemcpyAsync(policy, b.pout, sizeof(int32_t) * m->ns * k, hipMemcpyDeviceToHost, s));
    E2E_HIP(hipStreamSynchronize(s));
    int nconv = 0;
    for (int j = 0; j < k; j++) {
        if (hconv[j] >= 0) nconv++;
        if (col_sweeps) col_sweeps[j] = hconv[j] >= 0 ? hconv[j] : -last_k;
    }
    if (stats) {
        stats->sweeps = last_k;
        stats->converged = nconv;
        stats->variant = variant;
        stats->cols_per_lane = V.cpl;
        stats->states_per_wave = spw_of(V, (int)m->na);
        stats->col_block = kb_of(V);
        stats->solve_ms = now_ms() - t_start;
        stats->sweep_kernel_ms = kms;
        stats->tune_ms = tune_ms;
        // entries, row pointers + R, and per column: V read, V written (8 B each), policy;
        // over the active slices, the mean over the sweeps of the slices each one ran
        stats->bytes_per_sweep = 8.0 * m->nnz + 8.0 * m->rows + 4.0 + 20.0 * (double)k * m->ns;
        if (pre && last_k > 0) {
            E2E_TRY(fetch_slice_row_ptrs(m, s));
            double bsum = 0.0;
            for (int kk = 1; kk <= last_k; kk++)
                bsum += prefix_sweep_bytes(m, active_slices((uint32_t)kk, (int)m->nt), 20.0 * k);
            stats->bytes_per_sweep = bsum / last_k;
        }
        stats->launches = launched;
    }
    if (nconv < k) {
        set_error("%d of %d goal columns did not converge in %d sweeps", k - nconv, k, last_k);
        return E2E_ERR_NOT_CONVERGED;
    }
    return E2E_OK;
}

e2e_status vi_solve_goals(Model* m, const int32_t* goals, int k, const e2e_solve_opts* o,
                          const e2e_batch_opts* bo, uint32_t* policy, float* value, int32_t* col_sweeps,
                          e2e_batch_stats* stats) {
    E2E_CHECK(goals, E2E_ERR_INVALID, "goal_cells is NULL");
    return vi_solve_batch(m, goals, k, o, bo, policy, value, col_sweeps, stats);
}

e2e_status vi_solve_alphas(Model* m, const e2e_solve_opts* o, const e2e_batch_opts* bo, uint32_t* policy,
                           float* value, int32_t* col_sweeps, e2e_batch_stats* stats) {
    return vi_solve_batch(m, nullptr, m->n_alpha, o, bo, policy, value, col_sweeps, stats);
}

// ------------------------------------------------------------------ live batches
// A long-lived batch of goal columns on one model, for continuous batching: columns
// are admitted at sweep boundaries into free slots, every launch advances all live
// columns by one sweep, and a column leaves when its own residual falls below thresh.
// A column admitted before sweep k starts from V = 0 in the buffer sweep k reads, so
// after n of its sweeps it holds exactly the values of a solo n-sweep solve.

namespace {

// admit n columns: goal[col] = g, conv = -1, start = k, zero the column in the buffer
// the next sweep reads; clears all_conv so launches do work again.
__global__ void live_admit_kernel(double* V0, double* V1, int64_t ns, int kp, const int32_t* cols,
                                  const int32_t* goals, int n, int32_t* goal, int32_t* conv, int32_t* start,
                                  float* last_res, BatchState* st) {
    const uint32_t k = st->k;
    double* Vin = (k & 1u) ? V0 : V1;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns * n) {
        const int c = (int)(i / ns);
        const int64_t s = i - (int64_t)c * ns;
        Vin[s * kp + cols[c]] = 0.0;
    }
    if (i < n) {
        goal[cols[i]] = goals[i];
        conv[cols[i]] = -1;
        start[cols[i]] = (int32_t)k;
        last_res[cols[i]] = 0.0f;
    }
    if (i == 0) st->all_conv = 0;
}

__global__ void live_column(const double* V0, const double* V1, const int32_t* pi, const int32_t* conv,
                            int64_t ns, int kp, int col, float* vout, int32_t* pout) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int cj = conv[col];
    const double* V = (cj & 1) ? V1 : V0;
    vout[s] = (float)V[s * kp + col];
    pout[s] = pi[s * kp + col];
}

__global__ void live_init(double* V0, double* V1, int64_t n, int32_t* goal, int32_t* conv, int32_t* start,
                          float* last_res, uint32_t* colmax, int kp, BatchState* st) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        V0[i] = 0.0;
        V1[i] = 0.0;
    }
    if (i < kp) {
        goal[i] = -1;
        colmax[i] = 0u;
        conv[i] = 0;  // free slot
        start[i] = 1;
        last_res[i] = 0.0f;
    }
    if (i < kTileRows * 8) st->tnext[i * kCounterStride] = 0u;
    if (i == 0) {
        st->k = 1u;
        st->done = 0u;
        st->all_conv = 1;
    }
}

}  // namespace


struct LiveBatch {
    Model* m = nullptr;
    BatchBufs b;
    Plan plans[kLiveN];  // by kLiveWidths
    hipStream_t s = nullptr;
    BatchArgs A{};
    int32_t* d_list = nullptr;  // admission list scratch: [2][kp]
    int64_t bytes = 0;
};

int64_t live_batch_bytes(const Model* m, int kp) {
    // V[2] float64, pi, and the float value / policy staging of a harvest
    return (int64_t)m->ns * kp * (2 * 8 + 3 * 4) + (int64_t)kp * (44 + 4 * kPartRows) + 65536;
}

void live_free(LiveBatch* L);

e2e_status live_create(Model* m, int kp, float thresh, hipStream_t s, LiveBatch** out) {
    E2E_CHECK(m->built && m->Rsum && m->p.term_subgrid_size == 0 && !m->has_obstacles, E2E_ERR_INVALID,
              "live batch needs a shared goal model (keep_rsum, no terminal, no obstacles)");
    E2E_CHECK(kp >= 16 && kp <= 64 && kp % 16 == 0, E2E_ERR_INVALID, "kp=%d", kp);
    LiveBatch* L = new LiveBatch();
    L->m = m;
    L->s = s;
    e2e_status st = alloc_bufs(m, L->b, kp);
    if (st != E2E_OK || pool_malloc(&L->d_list, 2 * kp * sizeof(int32_t)) != hipSuccess) {
        L->b.release();
        delete L;
        if (st == E2E_OK) set_error("live batch: allocation failed");
        return st != E2E_OK ? st : E2E_ERR_NOMEM;
    }
    e2e_solve_opts o{};
    o.thresh = thresh;
    o.discount = 1.0f;
    for (int i = 0; i < kLiveN && st == E2E_OK; i++) st = plan_of(m, L->b, kVariants[width_variant(m, kp, i)], s, &L->plans[i]);
    if (st != E2E_OK) {
        live_free(L);
        return st;
    }
    L->A = make_args(m, L->b, o);  // after plan_of: the streaming kernel's partition
    // every block sweeps the slices its youngest running column needs (sweep_active)
    if (active_ok(m)) {
        L->A.pre_nt = (uint32_t)m->nt;
        L->A.pre_live = 1;
        bool stream = false;
        for (int i = 0; i < kLiveN; i++) stream = stream || kVariants[width_variant(m, kp, i)].kind == 2;
        if (stream) st = use_prefix_parts(m, L->b, L->A, s);
        if (st != E2E_OK) {
            live_free(L);
            return st;
        }
    }
    const int64_t n = m->ns * kp;
    hipLaunchKernelGGL(live_init, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, L->b.V[0], L->b.V[1], n,
                       L->b.goal, L->b.conv, L->b.start, L->b.last_res, L->b.colmax, kp, L->b.st);
    E2E_HIP(hipGetLastError());
    L->bytes = live_batch_bytes(m, kp);
    *out = L;
    return E2E_OK;
}

void live_free(LiveBatch* L) {
    if (!L) return;
    if (L->s) (void)hipStreamSynchronize(L->s);
    if (L->d_list) pool_free(L->d_list);
    L->b.release();
    delete L;
}

e2e_status live_admit(LiveBatch* L, const int32_t* cols, const int32_t* goals, int n) {
    if (n <= 0) return E2E_OK;
    const int kp = L->b.kp;
    E2E_CHECK(n <= kp, E2E_ERR_INVALID, "too many admissions");
    std::vector<int32_t> h(2 * kp);
    for (int i = 0; i < n; i++) {
        E2E_CHECK(cols[i] >= 0 && cols[i] < kp && goals[i] >= 0 && goals[i] < L->m->nc, E2E_ERR_INVALID,
                  "bad admission");
        h[i] = cols[i];
        h[kp + i] = goals[i];
    }
    E2E_HIP(hipMemcpyAsync(L->d_list, h.data(), 2 * kp * sizeof(int32_t), hipMemcpyHostToDevice, L->s));
    const int64_t tot = L->m->ns * n;
    hipLaunchKernelGGL(live_admit_kernel, dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, L->s, L->b.V[0],
                       L->b.V[1], L->m->ns, kp, L->d_list, L->d_list + kp, n, L->b.goal, L->b.conv, L->b.start,
                       L->b.last_res, L->b.st);
    E2E_HIP(hipGetLastError());
    // the host copy must outlive the async memcpy
    E2E_HIP(hipStreamSynchronize(L->s));
    return E2E_OK;
}

// Sweeps over the column blocks that hold slots 0 .. ncols-1 (the caller's highest live
// slot + 1): free slots beyond are not computed. The residual and convergence logic
// of the sweep only acts on running columns, all of which are in the launched blocks.
e2e_status live_step(LiveBatch* L, int nsweeps, int ncols) {
    const Plan& p = L->plans[width_class(ncols)];
    const dim3 grid = plan_grid(p, ncols);
    for (int n = 0; n < nsweeps; n++) hipLaunchKernelGGL(p.fn, grid, p.block, 0, L->s, L->A);
    E2E_HIP(hipGetLastError());
    return E2E_OK;
}

// One exact backward pass over the running columns 0 .. ncols-1 (kernel_variant 7's
// schedule with the goal columns' arithmetic): spmm_vi_rows once per slice, from the last
// to the first, each launch reading the values of the slice after it, which the previous
// launch wrote. Every running column then holds its fixed point in V0 and is marked done.
e2e_status live_pass(LiveBatch* L, int ncols) { return backward_pass(L->m, L->b, L->A, ncols, L->s); }

// Waits for the stream, then reports each slot's convergence sweep (conv) and
// admission sweep (start).
e2e_status live_state(LiveBatch* L, BatchState* st, int32_t* conv, int32_t* start) {
    const int kp = L->b.kp;
    E2E_HIP(hipMemcpyAsync(st, L->b.st, kBatchStateHead, hipMemcpyDeviceToHost, L->s));
    E2E_HIP(hipMemcpyAsync(conv, L->b.conv, kp * 4, hipMemcpyDeviceToHost, L->s));
    E2E_HIP(hipMemcpyAsync(start, L->b.start, kp * 4, hipMemcpyDeviceToHost, L->s));
    E2E_HIP(hipStreamSynchronize(L->s));
    return E2E_OK;
}

// Copies one converged column: V(start_state), pi(start_state), and optionally the full
// value/policy columns (host arrays of ns).
e2e_status live_harvest(LiveBatch* L, int col, int64_t start_state, float* v_start, int32_t* a_start,
                        float* vcol, uint32_t* pcol) {
    const int64_t ns = L->m->ns;
    hipLaunchKernelGGL(live_column, dim3((unsigned)ceil_div(ns, 256)), dim3(256), 0, L->s, L->b.V[0], L->b.V[1],
                       L->b.pi, L->b.conv, ns, L->b.kp, col, L->b.vout, L->b.pout);
    E2E_HIP(hipGetLastError());
    E2E_HIP(hipMemcpyAsync(v_start, L->b.vout + start_state, 4, hipMemcpyDeviceToHost, L->s));
    E2E_HIP(hipMemcpyAsync(a_start, L->b.pout + start_state, 4, hipMemcpyDeviceToHost, L->s));
    if (vcol) E2E_HIP(hipMemcpyAsync(vcol, L->b.vout, sizeof(float) * ns, hipMemcpyDeviceToHost, L->s));
    if (pcol) E2E_HIP(hipMemcpyAsync(pcol, L->b.pout, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, L->s));
    E2E_HIP(hipStreamSynchronize(L->s));
    return E2E_OK;
}

int live_kp(const LiveBatch* L) { return L->b.kp; }
int64_t live_bytes(const LiveBatch* L) { return L->bytes; }

}  // namespace e2e
