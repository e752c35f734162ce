// build.hip — Monte Carlo transition-model build on the GPU: (a) move, (b) LDS segmented
// radix sort + count-by-key, (c) fused CSR build; plus the reference-order COO
// export and the COO import used by the solver driver.
//
// Reference semantics: src/build_model.cu (transition_calc :483-559, move :268-356,
// extract_velocity :361-404, rewards :231-265, obstacles :419-477, compute_mean
// :563-570, COO emission :597-644, host loop :1142-1314), with the counting done as
// in the sort-based builder prel_tests/build_model.cu:410-487,705-706 (sort the S2 keys
// of each (s,a) segment, one entry per distinct key, count = run length), so every S2
// is counted (no 21x21 window drop; out-of-window lanes are reported).
//
// Layout in HBM
//   fields (read-only): u,v,s [nt][g][g], ui,vi [nt][nm][g][g], yi [nt][nr][nm], mask
//   staging  : uint2 [rows][K]  (sp2, count) per distinct S2 of a row; rows = ns*na,
//              row = (t*nc + cell)*na + a  (state-major, so the solver reads all
//              actions of a state from one contiguous nnz range)
//   arena    : uint2 overflow entries of rows with more than K distinct S2
//   row_nnz  : uint32 [rows]; R, Rsum : float [rows]
//   CSR      : row_ptr uint32 [rows+1], nz int2 [nnz] = (col, bits(val))
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "internal.h"

namespace e2e {

namespace {


constexpr int kStatSlots = 16;  // 64-bit build statistics (BuildArgs::stats)

struct BuildArgs {
    const float* u;
    const float* v;
    const float* ui;
    const float* vi;
    const float* yi;
    const float* s;
    const int32_t* mask;
    const int32_t* slice_obs;  // [nt] any obstacle in slice (NULL: no mask)
    const float* xs;
    const float* ys;
    float xs0, xsn, ys0, ysn, Dj, Di;
    double hDj, hDi;           // 0.5*Dj, 0.5*Di in double (move :318-330)
    double late_pen;           // 0.05*r_outbound in double (move :353)
    float cell_diag;           // fabsf(xs[1]-xs[0])*1.414213 rounded to float (:427)
    int g, nc, nt, nm, nr, na;
    int i_term, j_term, tsg, reward_type, m;
    float dt, r_out, r_term, alpha;
    float int_lim;             // integer reward-sum path: |reward| < int_lim, so nr of them fit int32
    int t0, K, radix_bits, R, tc;
    const uint32_t* seg_list;  // block-level kernel in list mode: one (s, a) row per block
    uint32_t* defer_list;      // wave kernel: rows handed to the block-level kernel
    unsigned* defer_count;
    int stage_y;
    uint2* staging;
    uint32_t* row_nnz;
    float* Rout;
    float* Rsum;
    uint2* arena;
    unsigned long long* arena_top;
    unsigned long long arena_cap;
    unsigned long long* stats;  // 0 oow lanes, 1 overflow rows, 2 max nuq, 3 nonfinite,
                                // 4 arena exhausted, 5 multipass rows, 6 lane off the grid,
                                // 7 alpha refusals, 8 counted cells (kStatSlots in all)
    int32_t pipe_rounds;  // per-row kernel: pair its rounds (few waves resident per SIMD)
    const uint32_t* cell_list;  // per-row kernel in list mode: the cells (t * nc + cell) it builds, one per block
    int tc_count;               // mc_count: cells per block
    uint32_t* left_cells;       // mc_count: the cells it leaves to the per-row kernel
    unsigned* left_count;
    int packed;                 // staging and arena entries as one u32 each, count << 16 | sp2 (counts and
                                // cells below 2^16), and an overflow row's arena offset as one u32
    int32_t n_alpha;      // alpha columns (reward method 4): rewards for each alpha too
    float* Ralpha;        // [n_alpha][rows_total]
    int64_t rows_total;
    float alphas[64];
    float act_speed[kMaxActions];
    float act_fcos[kMaxActions];
    float act_fsin[kMaxActions];
};

__device__ __forceinline__ bool is_terminal(const BuildArgs& A, int i, int j) {
    // unsigned compares: i - i_term in [0, tsg) and j - j_term in [0, tsg); no branches
    return ((unsigned)(i - A.i_term) < (unsigned)A.tsg) & ((unsigned)(j - A.j_term) < (unsigned)A.tsg);
}

// calculate_one_step_reward :231-265
__device__ __forceinline__ float step_reward(const BuildArgs& A, float ac_speed, float rad1,
                                             float rad2) {
    const int method = A.reward_type;
    const float Cr = 1.0f, Cf = 1.0f, Ct = (float)0.01;
    const float dt = A.dt;
    if (method == 0) return -dt;
    if (method == 1) return -(Cf * ac_speed * ac_speed + Ct) * dt;
    if (method == 2) return ((Cr * (rad2 + rad1) / 2) - (Cf * ac_speed * ac_speed) - Ct) * dt;
    if (method == 3) return ((Cr * (rad2 + rad1) / 2) - Ct) * dt;
    if (method == 4) {
        float energy_reward = -(ac_speed * ac_speed) * dt;
        float time_reward = -dt;
        return A.alpha * energy_reward + (1 - A.alpha) * time_reward;
    }
    return 0.0f;
}

// is_within_band :419-441
__device__ bool within_band(const BuildArgs& A, int i, int j, int i1, int j1, int i2, int j2) {
    if (i1 == i2 || j1 == j2) return true;
    const int g = A.g;
    float x = A.xs[j], y = A.ys[g - 1 - i];
    float x1 = A.xs[j1], y1 = A.ys[g - 1 - i1];
    float x2 = A.xs[j2], y2 = A.ys[g - 1 - i2];
    float Aa = (y2 - y1) / (x2 - x1);
    float B = -1.0f;
    float C = y1 - (Aa * x1);
    float dist = fabsf(Aa * x + B * y + C) / sqrtf((Aa * Aa) + (B * B));
    return dist < A.cell_diag;
}

// goes_through_obstacle :444-477
__device__ bool crosses_obstacle(const BuildArgs& A, int i1, int j1, int i2, int j2, int T) {
    const int g = A.g;
    const int imin = min(i1, i2), imax = max(i1, i2), jmin = min(j1, j2), jmax = max(j1, j2);
    const int32_t* m0 = A.mask + (size_t)T * A.nc;
    const int32_t* m1 = m0 + A.nc;
    for (int i = imin; i <= imax; i++)
        for (int j = jmin; j <= jmax; j++) {
            if (!(i == i1 && j == j1)) {
                int sp = j + g * i;
                if (m0[sp] == 1 || m1[sp] == 1)
                    if (within_band(A, i, j, i1, j1, i2, j2)) return true;
            }
        }
    return false;
}

// move :268-356 for a lane that starts at the centre (xc, yc) of cell (pi, pj).
// UNIT: cells of width exactly 1 (Dj == Di == 1.0f, the reference's dx = dy = 1). Then
// remainderf(x, 1) == x - rint(x) exactly (the IEEE remainder's quotient is
// round-half-even(x / 1), and x minus its nearest integer is exact), (x - rem) / 1 ==
// rint(x) exactly, and the double compare rem >= 0.5 * Dj equals rem >= 0.5f: the same
// bits as the general path at a fraction of its cost.
template <bool UNIT>
__device__ __forceinline__ void lane_move(const BuildArgs& A, float xc, float yc, float fc,
                                          float fs, float vx, float vy, bool late_slice, int& pi,
                                          int& pj, float& r, int* cls = nullptr) {
    // Written with selects instead of the reference's if/else-if ladders; every float
    // operation and its operands are the same, so the results are the same bits.
    const int n = A.g - 1;
    const float vnetx = fc + vx;
    const float vnety = fs + vy;
    float xnew = xc + (vnetx * A.dt);
    float ynew = yc + (vnety * A.dt);
    const bool xh = xnew > A.xsn, xl = xnew < A.xs0;  // :296-303
    xnew = xh ? A.xsn : (xl ? A.xs0 : xnew);
    r = (xh || xl) ? 0.0f + A.r_out : 0.0f;
    const bool yh = ynew > A.ysn, yl = ynew < A.ys0;  // :304-311
    ynew = yh ? A.ysn : (yl ? A.ys0 : ynew);
    r = (yh || yl) ? r + A.r_out : r;
    const float ax = xnew - A.xs0;
    const float ay = -(ynew - A.ysn);
    int xind, yind;
    if (UNIT) {
        const float nx = rintf(ax), ny = rintf(ay);
        const float remx = ax - nx, remy = ay - ny;
        xind = (int)nx + (remx >= 0.5f ? 1 : 0);
        yind = (int)ny + (remy >= 0.5f ? 1 : 0);
    } else {
        const float remx = remainderf(ax, A.Dj);
        const float remy = remainderf(ay, A.Di);
        xind = (int)((ax - remx) / A.Dj);
        yind = (int)((ay - remy) / A.Di);
        // the if/else-if ladder of :318-330 adds one to each index whose remainder is at
        // least half a cell (compared in double)
        xind += ((double)remx >= A.hDj) ? 1 : 0;
        yind += ((double)remy >= A.hDi) ? 1 : 0;
    }
    pi = yind;
    pj = xind;
    const bool edge = (pi == 0) | (pi == n) | (pj == 0) | (pj == n);  // :339-341
    r = edge ? r + A.r_out : r;
    const bool term = is_terminal(A, pi, pj);                          // :343-354
    // the reward class: how many r_outbound were added (x clamp, y clamp, edge), terminal
    if (cls) *cls = (((xh || xl) ? 1 : 0) + ((yh || yl) ? 1 : 0) + (edge ? 1 : 0)) * 2 + (term ? 1 : 0);
    if (late_slice) {  // T > nt - 20: uniform over the slice
        const float late = (float)((double)r + A.late_pen);
        r = term ? r + A.r_term : late;
    } else {
        r = term ? r + A.r_term : r;
    }
}

// Wave reductions with DPP: row shifts 1, 2, 4, 8 leave each row's total in its lane 15,
// row_bcast:15 / row_bcast:31 carry rows 0 and 2 into 1 and 3, so lane 63 holds the
// total; readlane makes it wave-uniform. Lanes without a source take the identity.
template <typename Op>
__device__ __forceinline__ int wave_reduce(int x, int id, Op op) {
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x111, 0xf, 0xf, false));
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x112, 0xf, 0xf, false));
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x114, 0xf, 0xf, false));
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x118, 0xf, 0xf, false));
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x142, 0xa, 0xf, false));
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(x, 63);
}
__device__ __forceinline__ int wave_min(int x) {
    return wave_reduce(x, 0x7fffffff, [](int a, int b) { return min(a, b); });
}
__device__ __forceinline__ int wave_max(int x) {
    return wave_reduce(x, (int)0x80000000, [](int a, int b) { return max(a, b); });
}
__device__ __forceinline__ int wave_sum(int x) {
    return wave_reduce(x, 0, [](int a, int b) { return a + b; });
}

// Exclusive scan over the NW waves of a block. lds: >= NW ints. Two barriers.
template <int NW>
__device__ __forceinline__ int block_excl_scan_int(int v, int* lds, int* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        int s = lds[w];
        if (w < wid) pre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

// Peers of this lane among `valid` lanes whose low `nbits` bits of `digit` agree.
__device__ __forceinline__ unsigned long long match_peers(uint32_t digit, int nbits,
                                                          bool valid) {
    unsigned long long peers = __ballot(valid);
    for (int b = 0; b < nbits; b++) {
        const bool bit = (digit >> b) & 1u;
        const unsigned long long bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

// Writes entry u of the current row (to the K staging slots or the overflow arena).
__device__ __forceinline__ void put_entry(const BuildArgs& A, size_t row, int u, int nuq,
                                          unsigned long long arena_off, uint32_t sp2,
                                          uint32_t count) {
    if (nuq <= A.K) {
        if (A.packed)
            reinterpret_cast<uint32_t*>(A.staging)[row * A.K + u] = count << 16 | sp2;
        else
            A.staging[row * A.K + u] = make_uint2(sp2, count);
    } else if (arena_off + (unsigned long long)nuq <= A.arena_cap) {
        if (A.packed)
            reinterpret_cast<uint32_t*>(A.arena)[arena_off + u] = count << 16 | sp2;
        else
            A.arena[arena_off + u] = make_uint2(sp2, count);
    }
}

// An overflow row's arena offset, in its first staging slot.
__device__ __forceinline__ void put_arena_offset(const BuildArgs& A, size_t row, unsigned long long arena_off) {
    if (A.packed)
        reinterpret_cast<uint32_t*>(A.staging)[row * A.K] = (uint32_t)arena_off;
    else
        A.staging[row * A.K] = make_uint2((uint32_t)arena_off, (uint32_t)(arena_off >> 32));
}

constexpr int kMiscInts = 256;  // wave range data [0,128), scan scratch [128,144), u64 at [160]

// LDS bytes of mc_move_count<WG> for R = ceil(nr / WG) rounds.
__host__ __device__ inline size_t build_lds_bytes(int WG, int R, int D, int nm, int nr, bool stage_y) {
    const int NW = WG / 64;
    const size_t NRP = (size_t)R * WG;
    return sizeof(double) * R * NW + sizeof(int) * kMiscInts + sizeof(uint32_t) * NW * (1 << D) +
           sizeof(float2) * NRP + 2 * sizeof(uint32_t) * (NRP + 4) +
           (stage_y ? sizeof(float) * (size_t)nr * nm : 0);
}

// One workgroup = TC consecutive cells of one time slice t; its threads own the
// realizations w = r*WG + tid (rounds r < R). For each cell the velocity sample of every
// realization is formed once (extract_velocity) and kept in LDS; then for each action:
//   (a) move every (s, a, w) lane, keep its relative S2 offset in LDS and its reward in
//       the canonical chunk tree;
//   (b) range-compress the segment's S2 keys to nbits bits and LSD-radix-sort them in
//       LDS with radix_bits-wide digits (ballot match ranking per wave, stable scatter
//       through per-(wave, digit) running offsets), then count runs (count-by-key);
//       when nbits <= radix_bits one counting pass orders the keys and its digit
//       histogram is the count-by-key, so the scatter is skipped;
//   write the row's distinct S2 with their counts and the row reward sum.
template <int WG, bool OBS, bool UNIT>
__global__ __launch_bounds__(WG) void mc_move_count(const BuildArgs A) {
    constexpr int NW = WG / 64;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int R = A.R;
    const int nr = A.nr, nm = A.nm, g = A.g, nc = A.nc;
    const int D = A.radix_bits, NBMAX = 1 << D;
    const int NRP = R * WG;
    double* csum = reinterpret_cast<double*>(smem_raw);          // [R*NW] chunk sums
    int* misc = reinterpret_cast<int*>(csum + R * NW);           // [kMiscInts]
    uint32_t* cnt = reinterpret_cast<uint32_t*>(misc + kMiscInts);  // [NW][NBMAX]
    float2* vxy = reinterpret_cast<float2*>(cnt + NW * NBMAX);   // [NRP]
    uint32_t* keysA = reinterpret_cast<uint32_t*>(vxy + NRP);    // [NRP + 4]
    uint32_t* keysB = keysA + NRP + 4;                           // [NRP + 4]
    float* rwl = reinterpret_cast<float*>(keysB);                // lane rewards (until the sum is formed)
    float* ylds = reinterpret_cast<float*>(keysB + NRP + 4);     // [nm][nr] when staged
    unsigned long long* arena_slot = reinterpret_cast<unsigned long long*>(misc + 160);
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    const int nchunks = (nr + 63) / 64;

    // grid mode: blockIdx.y = slice - t0, cells blockIdx.x*tc .. +tc, all actions;
    // list mode: the one row seg_list[blockIdx.x] = ((t*nc + c)*na + a)
    int t, c_begin, c_end, a_begin, a_end;
    if (A.seg_list) {
        const uint32_t lrow = A.seg_list[blockIdx.x];
        const uint32_t lmid = lrow / (uint32_t)A.na;
        a_begin = (int)(lrow - lmid * (uint32_t)A.na);
        a_end = a_begin + 1;
        t = (int)(lmid / (uint32_t)nc);
        c_begin = (int)(lmid - (uint32_t)t * (uint32_t)nc);
        c_end = c_begin + 1;
    } else {
        t = A.t0 + blockIdx.y;
        c_begin = blockIdx.x * A.tc;
        c_end = min(c_begin + A.tc, nc);
        a_begin = 0;
        a_end = A.na;
    }
    const size_t tslice = (size_t)t * nc;
    if (A.stage_y) {
        const float* src = A.yi + (size_t)t * nr * nm;
        for (int i = tid; i < nr * nm; i += WG) {
            int w = i / nm, k = i - w * nm;
            ylds[k * nr + w] = src[i];
        }
        __syncthreads();
    }
    const bool obs_slice = A.slice_obs && (A.slice_obs[t] | A.slice_obs[t + 1]);
    // the scalar field enters only the radiation rewards (methods 2 and 3, :249-255)
    const bool uses_rad = A.s && (A.reward_type == 2 || A.reward_type == 3);
    unsigned long long oow_acc = 0, multipass_acc = 0, overflow_acc = 0;
    uint32_t maxnuq_acc = 0;
    bool bad = false, bad_index = false;
    const int del = (A.m - 1) / 2;

    for (int c = c_begin; c < c_end; c++) {
        const int i1 = c / g, j1 = c - i1 * g;
        const size_t mid = tslice + c;
        // extract_velocity :361-404: mode sum from 0 in k order, then added to the mean
        const float u0 = A.u[mid], v0 = A.v[mid];
        for (int r = 0; r < R; r++) {
            const int w = r * WG + tid;
            if (w < nr) {
                float sx = 0.0f, sy = 0.0f;
                for (int k = 0; k < nm; k++) {
                    const size_t o = ((size_t)t * nm + k) * nc + c;
                    const float y = A.stage_y ? ylds[k * nr + w] : A.yi[((size_t)t * nr + w) * nm + k];
                    sx = sx + A.ui[o] * y;
                    sy = sy + A.vi[o] * y;
                }
                const float vx = u0 + sx, vy = v0 + sy;
                if (!(isfinite(vx) && isfinite(vy))) bad = true;
                vxy[w] = make_float2(vx, vy);
            }
        }
        const bool term1 = is_terminal(A, i1, j1);
        const bool obs1 = A.mask && A.mask[mid] == 1;
        const float rad1 = uses_rad ? A.s[mid] : 0.0f;
        const float xc = A.xs[j1], yc = A.ys[g - 1 - i1];

        for (int a = a_begin; a < a_end; a++) {
            const float speed = A.act_speed[a], fc = A.act_fcos[a], fs = A.act_fsin[a];
            int dimin = 0x7fffffff, dimax = -0x7fffffff, djmin = 0x7fffffff, djmax = -0x7fffffff;
            int oow = 0, ipart = 0;
            bool allint = true;
            // ---- (a) move: transition_calc :522-545
            for (int r = 0; r < R; r++) {
                const int w = r * WG + tid;
                const bool valid = w < nr;
                float rw = 0.0f;
                if (valid) {
                    int pi = i1, pj = j1;
                    if (!term1) {
                        if (!obs1) {
                            const float2 v = vxy[w];
                            lane_move<UNIT>(A, xc, yc, fc, fs, v.x, v.y, t > A.nt - 20, pi, pj, rw);
                            // impossible for dx, dy > 0; guards the gathers below
                            const int ci = min(max(pi, 0), g - 1), cj = min(max(pj, 0), g - 1);
                            bad_index |= (ci != pi) | (cj != pj);
                            pi = ci;
                            pj = cj;
                            const int sp2 = pj + g * pi;
                            const float rad2 = uses_rad ? A.s[tslice + nc + sp2] : 0.0f;
                            rw += step_reward(A, speed, rad1, rad2);
                            if (OBS && obs_slice && crosses_obstacle(A, i1, j1, pi, pj, t)) rw = A.r_out;
                        } else {
                            rw = A.r_out;
                        }
                    }
                    const int di = pi - i1, dj = pj - j1;
                    dimin = min(dimin, di);
                    dimax = max(dimax, di);
                    djmin = min(djmin, dj);
                    djmax = max(djmax, dj);
                    oow += ((unsigned)(di + del) >= (unsigned)A.m) | ((unsigned)(dj + del) >= (unsigned)A.m);
                    keysA[w] = ((uint32_t)(di + 32768) << 16) | (uint32_t)(dj + 32768);
                }
                if (valid) {
                    rwl[w] = rw;
                    // integer-valued rewards sum exactly in int in any order, which equals
                    // the canonical double sum
                    if (rw == rintf(rw) && fabsf(rw) < A.int_lim) ipart += (int)rw;
                    else allint = false;
                }
            }
            dimin = wave_min(dimin);
            dimax = wave_max(dimax);
            djmin = wave_min(djmin);
            djmax = wave_max(djmax);
            oow = wave_sum(oow);
            ipart = wave_sum(ipart);
            const bool wave_int = __all(allint);
            if (lane == 0) {
                misc[wv * 8 + 0] = dimin;
                misc[wv * 8 + 1] = dimax;
                misc[wv * 8 + 2] = djmin;
                misc[wv * 8 + 3] = djmax;
                misc[wv * 8 + 4] = oow;
                misc[wv * 8 + 5] = ipart;
                misc[wv * 8 + 6] = wave_int ? 1 : 0;
            }
            for (int i = tid; i < NW * NBMAX; i += WG) cnt[i] = 0u;
            __syncthreads();
            // block totals: lane l < NW reads wave l's entries, then wave reductions
            {
                const bool lw = lane < NW;
                dimin = wave_min(lw ? misc[lane * 8 + 0] : 0x7fffffff);
                dimax = wave_max(lw ? misc[lane * 8 + 1] : -0x7fffffff);
                djmin = wave_min(lw ? misc[lane * 8 + 2] : 0x7fffffff);
                djmax = wave_max(lw ? misc[lane * 8 + 3] : -0x7fffffff);
            }
            const int oow_blk = wave_sum(lane < NW ? misc[lane * 8 + 4] : 0);
            const long long isum = wave_sum(lane < NW ? misc[lane * 8 + 5] : 0);
            const bool blk_int = __all(lane >= NW || misc[lane * 8 + 6] != 0);
            double rsum = (double)isum;
            if (!blk_int) {
                for (int ch = wv; ch < nchunks; ch += NW) {
                    const int w = ch * 64 + lane;
                    double p = w < nr ? (double)rwl[w] : 0.0;
#pragma unroll
                    for (int o = 32; o >= 1; o >>= 1) {
                        const double q = __shfl_down(p, o, 64);
                        p = p + q;
                    }
                    if (lane == 0) csum[ch] = p;
                }
                __syncthreads();
                rsum = 0.0;
                if (tid == 0)
                    for (int ch = 0; ch < nchunks; ch++) rsum = rsum + csum[ch];
            }
            // ---- (b) range-compress: key = (di - dimin) * Wj + (dj - djmin) keeps the
            // (i2, j2) lexicographic order, i.e. ascending S2 = (t+1)*nc + j2 + g*i2
            const uint32_t Wj = (uint32_t)(djmax - djmin + 1);
            const uint64_t nkeys = (uint64_t)(dimax - dimin + 1) * Wj;
            const int nbits = nkeys <= 1 ? 0 : 64 - __clzll((long long)(nkeys - 1));
            const size_t row = mid * A.na + a;
            int nuq = 0;
            unsigned long long arena_off = 0;
            if (nbits <= D) {
                // single counting pass: the block's key histogram by LDS atomics (the
                // counts are integers, so the order of the adds does not matter); the
                // histogram of the compressed keys is the count-by-key in key order
                for (int r = 0; r < R; r++) {
                    const int w = r * WG + tid;
                    if (w < nr) {
                        const uint32_t k0 = keysA[w];
                        const int di = (int)(k0 >> 16) - 32768, dj = (int)(k0 & 0xffffu) - 32768;
                        atomicAdd(&cnt[(uint32_t)(di - dimin) * Wj + (uint32_t)(dj - djmin)], 1u);
                    }
                }
                __syncthreads();
                uint32_t tot = 0;
                if (tid < NBMAX) tot = cnt[tid];  // NBMAX <= 256 <= WG: one digit per thread
                int total;
                const int u = block_excl_scan_int<NW>(tot > 0 ? 1 : 0, misc + 128, &total);
                nuq = total;
                if (nuq > A.K) {
                    if (tid == 0) *arena_slot = atomicAdd(A.arena_top, (unsigned long long)nuq);
                    __syncthreads();
                    arena_off = *arena_slot;
                }
                if (tot > 0) {
                    const int di = (int)(tid / Wj) + dimin, dj = (int)(tid % Wj) + djmin;
                    put_entry(A, row, u, nuq, arena_off, (uint32_t)((j1 + dj) + g * (i1 + di)), tot);
                }
            } else {
                for (int r = 0; r < R; r++) {  // range-compress the keys in place
                    const int w = r * WG + tid;
                    if (w < nr) {
                        const uint32_t k0 = keysA[w];
                        const int di = (int)(k0 >> 16) - 32768, dj = (int)(k0 & 0xffffu) - 32768;
                        keysA[w] = (uint32_t)(di - dimin) * Wj + (uint32_t)(dj - djmin);
                    }
                }
                // multi-pass LSD radix sort of the nr keys in LDS; element (wave, round,
                // lane) of a pass is keysA[r*WG + tid] in pass 0 (any order is fine for the
                // first pass) and src[wv*R*64 + r*64 + lane] afterwards (the stable order)
                const int passes = (nbits + D - 1) / D;
                uint32_t* src = keysA;
                uint32_t* dst = keysB;
                for (int ps = 0; ps < passes; ps++) {
                    const int shift = ps * D;
                    const int dbits = min(D, nbits - shift);
                    const int nbp = 1 << dbits;
                    if (ps > 0) {
                        for (int i = tid; i < NW * NBMAX; i += WG) cnt[i] = 0u;
                        __syncthreads();
                    }
                    for (int r = 0; r < R; r++) {  // count
                        const int idx = ps == 0 ? r * WG + tid : wv * (R * 64) + r * 64 + lane;
                        const bool valid = idx < nr;
                        const uint32_t key = valid ? src[idx] : 0u;
                        const uint32_t d = (key >> shift) & (uint32_t)(nbp - 1);
                        const unsigned long long peers = match_peers(d, dbits, valid);
                        if (valid && lane == __ffsll((long long)peers) - 1)
                            cnt[wv * NBMAX + d] += (uint32_t)__popcll(peers);
                    }
                    __syncthreads();
                    // offsets: digit-major, then wave order (stable)
                    uint32_t tot = 0;
                    if (tid < nbp) {
#pragma unroll
                        for (int w = 0; w < NW; w++) tot += cnt[w * NBMAX + tid];
                    }
                    int total;
                    const int base = block_excl_scan_int<NW>((int)tot, misc + 128, &total);
                    if (tid < nbp) {
                        uint32_t run = (uint32_t)base;
#pragma unroll
                        for (int w = 0; w < NW; w++) {
                            const uint32_t cw = cnt[w * NBMAX + tid];
                            cnt[w * NBMAX + tid] = run;
                            run += cw;
                        }
                    }
                    __syncthreads();
                    for (int r = 0; r < R; r++) {  // scatter through running offsets
                        const int idx = ps == 0 ? r * WG + tid : wv * (R * 64) + r * 64 + lane;
                        const bool valid = idx < nr;
                        const uint32_t key = valid ? src[idx] : 0u;
                        const uint32_t d = (key >> shift) & (uint32_t)(nbp - 1);
                        const unsigned long long peers = match_peers(d, dbits, valid);
                        const uint32_t b0 = valid ? cnt[wv * NBMAX + d] : 0u;
                        if (valid) dst[b0 + (uint32_t)__popcll(peers & lt_mask)] = key;
                        if (valid && lane == __ffsll((long long)peers) - 1)
                            cnt[wv * NBMAX + d] = b0 + (uint32_t)__popcll(peers);
                    }
                    __syncthreads();
                    uint32_t* tmp = src;
                    src = dst;
                    dst = tmp;
                }
                // count-by-key: run heads of the sorted keys (thread-contiguous blocks)
                int heads = 0;
                const int p0 = tid * R;
                for (int k = 0; k < R; k++) {
                    const int p = p0 + k;
                    if (p < nr && (p == 0 || src[p] != src[p - 1])) heads++;
                }
                int total;
                int u = block_excl_scan_int<NW>(heads, misc + 128, &total);
                nuq = total;
                for (int k = 0; k < R; k++) {
                    const int p = p0 + k;
                    if (p < nr && (p == 0 || src[p] != src[p - 1])) dst[u++] = (uint32_t)p;
                }
                if (tid == 0) {
                    dst[nuq] = (uint32_t)nr;
                    if (nuq > A.K) *arena_slot = atomicAdd(A.arena_top, (unsigned long long)nuq);
                }
                __syncthreads();
                if (nuq > A.K) arena_off = *arena_slot;
                for (int q = tid; q < nuq; q += WG) {
                    const uint32_t p = dst[q];
                    const uint32_t kk = src[p];
                    const int di = (int)(kk / Wj) + dimin, dj = (int)(kk % Wj) + djmin;
                    put_entry(A, row, q, nuq, arena_off, (uint32_t)((j1 + dj) + g * (i1 + di)),
                              dst[q + 1] - p);
                }
                multipass_acc++;
            }
            if (tid == 0) {
                A.row_nnz[row] = (uint32_t)nuq;
                const float fs_ = (float)rsum;  // compute_mean :563-570
                A.Rout[row] = fs_ / nr;
                if (A.Rsum) A.Rsum[row] = fs_;
                oow_acc += (unsigned long long)oow_blk;
                if ((uint32_t)nuq > maxnuq_acc) maxnuq_acc = (uint32_t)nuq;
                if (nuq > A.K) {
                    overflow_acc++;
                    if (arena_off + (unsigned long long)nuq > A.arena_cap) atomicOr(A.stats + 4, 1ull);
                    else put_arena_offset(A, row, arena_off);
                }
            }
            __syncthreads();
        }
        __syncthreads();  // vxy of this cell is done before the next cell overwrites it
    }
    if (bad) atomicOr(A.stats + 3, 1ull);
    if (bad_index) atomicOr(A.stats + 6, 1ull);
    if (tid == 0) {
        if (oow_acc) atomicAdd(A.stats + 0, oow_acc);
        if (overflow_acc) atomicAdd(A.stats + 1, overflow_acc);
        if (maxnuq_acc) atomicMax(A.stats + 2, (unsigned long long)maxnuq_acc);
        if (multipass_acc) atomicAdd(A.stats + 5, multipass_acc);
    }
}

// floor(x) converted to int in one instruction (x finite, within int range)
__device__ __forceinline__ int cvt_flr_i32(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Wave-level DPP scans and reductions for the wave kernel (see wave_reduce).
__device__ __forceinline__ int wave_incl_scan(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);
    return x;
}

constexpr int kWaveBins = 256;  // key histogram bins per wave (nkeys above: deferred)
// Copies of each wave's histogram: lane l counts into copy l % kHistCopies (word
// bin * kHistCopies + copy, so a bin's copies sit in adjacent banks). Lanes of a wave
// mostly land in a few bins, and LDS atomics to one word are serialized; the copies cut
// those collisions by kHistCopies. The emit adds the copies.
#ifndef E2E_HIST_COPIES
#define E2E_HIST_COPIES 4
#endif
constexpr int kHistCopies = E2E_HIST_COPIES;
static_assert((kHistCopies & (kHistCopies - 1)) == 0, "power of two");
constexpr int kHistWords = kWaveBins * kHistCopies;
constexpr int kFixHalf = 7;     // fixed offset window +-7 cells, counted during the move
constexpr int kFixSide = 2 * kFixHalf + 1;
static_assert(kFixSide * kFixSide <= kWaveBins, "fixed window fits the histogram");

// Alpha columns (reward method 4): the row's reward for every alpha, from the counts of
// its lanes per reward class. A lane's reward is the class value of
// calculate_one_step_reward :256-260 and move :268-356: rk = k penalties of r_outbound
// added one at a time, then + r_terminal (terminal) or the late penalty (T > nt - 20,
// in double), then + the step reward alpha*(-F^2 dt) + (1-alpha)*(-dt); an obstacle
// crossing is r_outbound. The sum over classes, n_c * v_c in double, is exact, and so
// is every partial sum of the canonical chunk order, when the values' exponents span
// at most 53 - 24 - ceil(log2 nr) binades; then both give the same float, as the
// canonical sum of a build with that alpha. Otherwise stats[7] bit 0 is set (the build
// fails). Rows of a terminal or obstacle start cell have no step reward: every column
// takes the row's own reward.
__device__ __forceinline__ void alpha_rewards(const BuildArgs& A, size_t row, float r_own, bool fixed_start,
                                           unsigned long long acls, float speed, bool late_slice, int nr) {
    const int lane = threadIdx.x & 63;
    if (fixed_start) {
        for (int j = lane; j < A.n_alpha; j += 64) A.Ralpha[(size_t)j * A.rows_total + row] = r_own;
        return;
    }
    int n[9];
#pragma unroll
    for (int c = 0; c < 9; c++) n[c] = wave_sum((int)((acls >> (7 * c)) & 127ull));
    int lg = 0;
    while ((1 << lg) < nr) lg++;
    for (int j = lane; j < A.n_alpha; j += 64) {
        const float al = A.alphas[j];
        const float energy_reward = -(speed * speed) * A.dt, time_reward = -A.dt;
        const float rs = al * energy_reward + (1 - al) * time_reward;
        double sum = 0.0;
        int emin = 1 << 20, emax = -(1 << 20);
        auto add = [&](int cnt, float v) {
            if (cnt == 0 || v == 0.0f) return;
            const int e = ilogbf(v);
            emin = min(emin, e);
            emax = max(emax, e);
            sum += (double)cnt * (double)v;
        };
        float rk = 0.0f;
        for (int k = 0; k < 4; k++) {
            const float other = late_slice ? (float)((double)rk + A.late_pen) : rk;
            add(n[2 * k], other + rs);
            add(n[2 * k + 1], (rk + A.r_term) + rs);
            rk = rk + A.r_out;
        }
        add(n[8], A.r_out);
        if (emax >= emin && (emax - emin) + 24 + lg > 53) atomicOr(A.stats + 7, 1ull);
        A.Ralpha[(size_t)j * A.rows_total + row] = (float)sum / nr;  // compute_mean :563-570
    }
}

// The reward values a lane of a row can get, when the row's reward sum is an exact int32
// sum. Every reward is one of: rstep + r_k (+ r_terminal, or the late-slice penalty),
// r_k = k * r_outbound summed one at a time, k = 0..3 (x clamp, y clamp, edge), or
// r_outbound (obstacles), or 0. seg_int: all of them are integers below int_lim.
// seg_lin: also linear in k with slope r_outbound, so a lane's reward is the k = edge
// value (tt0 / tt1 landing on a terminal cell, tn0 / tn1 elsewhere, for k = 0 / 1) plus
// r_outbound times its clamp count.
struct RowRewards {
    bool seg_int, seg_lin;
    int tt0, tt1, tn0, tn1;
};

__device__ __forceinline__ RowRewards row_rewards(const BuildArgs& A, float rstep, bool late_slice, bool uses_rad) {
    RowRewards w{!uses_rad, !uses_rad, 0, 0, 0, 0};
    if (!w.seg_int) return w;
    auto ok = [&](float v) { return v == rintf(v) && fabsf(v) < A.int_lim; };
    float rk = 0.0f;
    w.seg_int = ok(A.r_out) && ok(0.0f);
    w.seg_lin = w.seg_int;
    for (int kk = 0; kk < 4; kk++) {
        const float other = late_slice ? (float)((double)rk + A.late_pen) : rk;
        const float rt = (rk + A.r_term) + rstep, ro = other + rstep;
        w.seg_int = w.seg_int && ok(rt) && ok(ro);
        if (w.seg_int) {
            if (kk == 0) {
                w.tt0 = (int)rt;
                w.tn0 = (int)ro;
            } else if (kk == 1) {
                w.tt1 = (int)rt;
                w.tn1 = (int)ro;
            }
            const int64_t slope = (int64_t)kk * (int64_t)A.r_out;
            w.seg_lin = w.seg_lin && (int64_t)(int)rt == (int64_t)w.tt0 + slope &&
                        (int64_t)(int)ro == (int64_t)w.tn0 + slope;
        }
        rk = rk + A.r_out;
    }
    w.seg_lin = w.seg_lin && w.seg_int;
    return w;
}

// ---------------------------------------------------------------------------------
// Counted cells: the rows of a cell from one pass over its realizations.
//
// On unit cells a moving lane's landing column is a non-decreasing function of its
// sample's vx alone, and its landing row (counted from the bottom) of vy alone:
//   ord_x(v) = 0 (clamped low) | 1 + xind (inside) | XH + 2 (clamped high),
//   xnew = xc + (fcos + v) * dt, xind = floor(ax) + (fract(ax) >= 0.5), ax = xnew - xs0,
// every operation rounded as in lane_move (each one is monotone, and so is the clamp),
// and ord_y the same with the row index. So for action a and each ord o the cell's
// samples reach (ord_lo(a) < o <= ord_hi(a)), ord_x(v) >= o exactly when
// key(v) >= tau(a, o), the smallest float key with ord_x >= o, found by bisection over
// the keys between the cell's smallest and largest sample (key: the float's bits mapped
// to an order-preserving uint32). A block then counts the cell once for all its actions:
//   1. the boundaries tau of every action, one lane each, ranked per axis;
//   2. every realization's position among them, I_x = #{tau_x <= key(vx)} and I_y, in a
//      2-D histogram in LDS, then its inclusive prefix;
//   3. per row, the landing cells of its box in ascending S2 order: a cell's count is a
//      rectangle of the prefix (its I ranges cover its clamped and inside ords), and the
//      reward is summed per landing cell and clamp as the per-lane fast pass sums it.
// Same counts, order, rewards and statistics as moving every lane. One wavefront counts
// one cell, with its own tables and grid in LDS (no block barrier), and recomputes the
// samples when it needs them. A group of actions holds at most kCntMax boundaries per
// axis and a grid of at most kCntArea bins; a cell that needs more is counted group by
// group (on the 100x100x120 fields 1.28 groups per cell, against 1.54 with at most 47
// boundaries per axis in a square grid of the same area).
constexpr int kCntMax = 63;                    // boundaries per axis per group (ranks fit 64 keys)
constexpr int kCntArea = 49 * 49;              // grid [I_y + 1][I_x + 1] bins, row pitch cnt_pitch(I_x)
// Row pitch of a group's grid for tx boundaries on x: tx + 2 bins, one more when that is
// a multiple of 4, so that the lanes walking the rows (one row each) spread over the LDS banks
__device__ __forceinline__ int cnt_pitch(int tx) { return tx + 2 + (((tx + 2) & 3) == 0 ? 1 : 0); }
constexpr int kCntWaves = 16;                  // waves of a counting block
constexpr int kCntModes = 8;                   // DO modes a counting wave holds in registers

// LDS of one counting wave. The block's dynamic LDS starts with the waves' tables, each a
// multiple of 256 bytes, so every u[ax] lies on a 256-byte boundary: a rank's bisection
// then carries the LDS address of its position itself (the base OR the offset).
struct alignas(256) CountWave {
    uint32_t u[2][64];                  // the group's boundary keys per axis, ascending; padding 0xffffffff
    uint32_t d[kCntMax + 1];            // boundary j of the axis being ranked: action << 16 | ord, then its
                                        // key (padding 0xffffffff), read four at a time
    uint32_t grid[(kCntArea + 1) / 2];  // 16-bit counts, two per word (a count is at most nrzns < 2^16)
    uint8_t c[2][kCntMax + 1];          // boundary j of the group: #{u <= its key}
};
static_assert(offsetof(CountWave, d) % 16 == 0, "the keys are read as uint4");
static_assert(sizeof(CountWave) % 256 == 0, "counting-wave tables keep u on 256-byte boundaries");

// 32-bit LDS address of a pointer into shared memory, and a load from such an address
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}
__device__ __forceinline__ uint32_t lds_load(uint32_t a) {
    return *(const __attribute__((address_space(3))) uint32_t*)(uintptr_t)a;
}
// a generic pointer to the shared-memory word at 32-bit LDS address a
__device__ __forceinline__ void* lds_ptr(uint32_t a) {
    return (void*)(__attribute__((address_space(3))) void*)(uintptr_t)a;
}
// LDS of a counting block: this table, the waves' tables, then Yi [nm][nr] when it fits
struct CountShared {
    float fcos[kMaxActions], fsin[kMaxActions], speed[kMaxActions];  // the action table, for per-lane reads
    RowRewards rew[kMaxActions];  // each action's row rewards on the block's slice (row_rewards)
    unsigned next_cell;           // the block's next cell for a free wave
};

__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t b = __float_as_uint(f);
    // negative: all bits flipped; else the sign bit set
    return b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u);
}
__device__ __forceinline__ float kfloat(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// the unit-cell index of lane_move<true> for a >= 0: floor(a) + (fract(a) >= 0.5)
__device__ __forceinline__ int unit_index(float a) {
    return cvt_flr_i32(a) + (__builtin_amdgcn_fractf(a) >= 0.5f ? 1 : 0);
}

// the scalars of the unit-cell move from one start cell
struct CountGeo {
    float xc, yc, xs0, xsn, ys0, ysn, dt;
    int XH, YL;  // the column of a lane clamped high, the row of a lane clamped low
};

__device__ __forceinline__ int ord_x(const CountGeo& G, float fc, float v) {
    const float xnew = G.xc + ((fc + v) * G.dt);
    const int inside = 1 + unit_index(xnew - G.xs0);
    return xnew < G.xs0 ? 0 : (xnew > G.xsn ? G.XH + 2 : inside);
}
__device__ __forceinline__ int ord_y(const CountGeo& G, float fs, float v) {
    const float ynew = G.yc + ((fs + v) * G.dt);
    const int inside = 1 + (G.YL - unit_index(G.ysn - ynew));
    return ynew < G.ys0 ? 0 : (ynew > G.ysn ? G.YL + 2 : inside);
}
template <int AX>
__device__ __forceinline__ int ord_of(const CountGeo& G, float f, float v) {
    return AX ? ord_y(G, f, v) : ord_x(G, f, v);
}

// tau: the smallest key whose float has ord >= o, given ord(k0) < o <= ord(k1). The
// bisection starts from a bracket around the real-number boundary (where the move's
// unrounded position reaches the ord's edge), which the rounding moves by a few ulps of
// the position; the bracket is used only if its ends verify, else [k0, k1]. Half-width
// 2^-23 of the magnitudes involved: on the 100x100x120 fields every bracket verifies, and
// a wave's slowest lane bisects 12.9 times (15.7 at 2^-20).
template <int AX>
__device__ __forceinline__ uint32_t boundary_key(const CountGeo& G, float f, int o, uint32_t k0, uint32_t k1) {
    float edge;
    if (AX == 0)
        edge = o == 1 ? G.xs0 : (o == G.XH + 2 ? G.xsn : G.xs0 + ((float)(o - 1) - 0.5f));
    else
        edge = o == 1 ? G.ys0 : (o == G.YL + 2 ? G.ysn : G.ysn - ((float)(G.YL + 1 - o) + 0.5f));
    const float c = AX ? G.yc : G.xc;
    const float vb = (edge - c) / G.dt - f;
    const float d = 0x1p-23f * (fabsf(c) + fabsf(edge) + (fabsf(f) + fabsf(vb)) * G.dt + 1.0f) / G.dt;
    uint32_t lo = max(fkey(vb - d), k0), hi = min(fkey(vb + d), k1);
    if (!(lo < hi && ord_of<AX>(G, f, kfloat(lo)) < o && ord_of<AX>(G, f, kfloat(hi)) >= o)) {
        lo = k0;
        hi = k1;
    }
    // every key between two finite floats is a finite float
    while (hi - lo > 1u) {
        const uint32_t m = lo + ((hi - lo) >> 1);
        if (ord_of<AX>(G, f, kfloat(m)) >= o)
            hi = m;
        else
            lo = m;
    }
    return hi;
}

// Sums over 16-lane segments (SW 16: the sum of each row of 16 lanes, in all its lanes)
// or over the wave (SW 64).
template <int SW>
__device__ __forceinline__ int seg_sum(int x) {
    if (SW == 64) return wave_sum(x);
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
    return __shfl(x, (int)(threadIdx.x & 63) | 15, 64);
}

// landing column / row of an ord
__device__ __forceinline__ int col_of(const CountGeo& G, int o) { return o == 0 ? 0 : (o == G.XH + 2 ? G.XH : o - 1); }
__device__ __forceinline__ int row_of(const CountGeo& G, int o) {
    return o == 0 ? G.YL : (o == G.YL + 2 ? 0 : G.YL - (o - 1));
}

// Wave reductions of floats (DPP, as wave_reduce): every lane gets the result.
__device__ __forceinline__ float wave_fmin(float x) {
    auto op = [](int a, int b) { return __float_as_int(fminf(__int_as_float(a), __int_as_float(b))); };
    return __int_as_float(wave_reduce(__float_as_int(x), __float_as_int(INFINITY), op));
}
__device__ __forceinline__ float wave_fmax(float x) {
    auto op = [](int a, int b) { return __float_as_int(fmaxf(__int_as_float(a), __int_as_float(b))); };
    return __int_as_float(wave_reduce(__float_as_int(x), __float_as_int(-INFINITY), op));
}

// Every realization's cell of a group's boundary grid, two rounds at a time: the rank
// #{u <= key} of both samples on both axes by bisection, the four lookups of a level
// issued together, over the first 2^LX - 1 keys of the x axis and 2^LY - 1 of the y axis
// (the group has at most that many; the padding keys 0xffffffff lie above every finite
// sample's key). A position is kept as the LDS address of u[ax][rank] (u[ax] is 256-byte
// aligned, so adding 4h is an OR). The count of bin (rank_y + 1) * P + rank_x + 1 is
// raised by one in the 16-bit grid.
template <int LX, int LY, typename SampleFn>
__device__ __forceinline__ void place_group(const SampleFn& sample, int lane, int nr, uint32_t ax0, uint32_t ay0,
                                            int P, uint32_t* grid) {
    // byte address of the 16-bit count of bin b = (rank_y + 1) * P + rank_x + 1: the grid's
    // address + 2 b = (oy * P + ox) / 2 + base2 (oy * P and ox are multiples of 4)
    const uint32_t base2 = lds_addr(grid) + 2u * (uint32_t)(P + 1) - ((ay0 * (uint32_t)P + ax0) >> 1);
    for (int w = lane; w < nr; w += 128) {
        const bool two = w + 64 < nr;
        const float2 v1 = sample(w), v2 = sample(two ? w + 64 : w);
        const uint32_t kx1 = fkey(v1.x), ky1 = fkey(v1.y), kx2 = fkey(v2.x), ky2 = fkey(v2.y);
        uint32_t ox1 = ax0, oy1 = ay0, ox2 = ax0, oy2 = ay0;
#pragma unroll
        for (int h = 32; h >= 1; h >>= 1) {
            const bool dx = h < (1 << LX), dy = h < (1 << LY);
            uint32_t ex1 = 0u, ex2 = 0u, ey1 = 0u, ey2 = 0u;
            if (dx) {
                ex1 = lds_load(ox1 + 4u * (h - 1));
                ex2 = lds_load(ox2 + 4u * (h - 1));
            }
            if (dy) {
                ey1 = lds_load(oy1 + 4u * (h - 1));
                ey2 = lds_load(oy2 + 4u * (h - 1));
            }
            if (dx) {
                ox1 |= ex1 <= kx1 ? 4u * h : 0u;
                ox2 |= ex2 <= kx2 ? 4u * h : 0u;
            }
            if (dy) {
                oy1 |= ey1 <= ky1 ? 4u * h : 0u;
                oy2 |= ey2 <= ky2 ? 4u * h : 0u;
            }
        }
        // (LDS addresses < 2^16, so the products fit 24 bits by 24 bits)
        const uint32_t a1 = ((__umul24(oy1, (uint32_t)P) + ox1) >> 1) + base2;
        const uint32_t a2 = ((__umul24(oy2, (uint32_t)P) + ox2) >> 1) + base2;
        // the word holding the count, and its half
        atomicAdd(reinterpret_cast<uint32_t*>(lds_ptr(a1 & ~3u)), 1u << ((a1 & 2u) << 3));
        if (two) atomicAdd(reinterpret_cast<uint32_t*>(lds_ptr(a2 & ~3u)), 1u << ((a2 & 2u) << 3));
    }
}

// One wavefront counts every row of a moving cell (the per-lane fast pass's conditions
// hold for the cell: unit cells, no scalar-field reward, neither terminal nor obstacle)
// and writes them as the per-lane pass would. Returns false, wave-uniformly and before
// writing anything, when the cell is left to that pass: a non-finite sample, an
// obstacle within the fixed window, a row without exact integer rewards, an action with
// more than kCntMax boundaries on an axis. Y: the slice's Yi, [nm][nr] (YLDS, in LDS)
// or [nr][nm] (global). NM: the number of DO modes when fixed at compile time (the mode
// sums fully unrolled), 0 for any number up to kCntModes. OBS: the model has obstacles.
template <bool YLDS, int NM>
__device__ __forceinline__ bool count_cell(const BuildArgs& A, const float* __restrict__ Y, CountWave& Wv,
                                           const CountShared& T, CountGeo& G, int t, int c, bool late_slice,
                                           bool obs_slice, unsigned long long& oow_acc,
                                           unsigned long long& overflow_acc, uint32_t& maxnuq_acc) {
    const int lane = threadIdx.x & 63;
    const int na = A.na, nr = A.nr, nm = NM ? NM : A.nm, g = A.g, nc = A.nc;
    const bool OBS = A.slice_obs != nullptr;
    const int i1 = c / g, j1 = c - i1 * g;
    const size_t tslice = (size_t)t * nc, mid = tslice + c;
    if (is_terminal(A, i1, j1) || (A.mask && A.mask[mid] == 1)) return false;
    if (OBS && obs_slice) {  // an obstacle within the fixed window (at t or t + 1): the per-lane pass
        bool near = false;
        const int32_t* m0 = A.mask + tslice;
        for (int k = lane; k < kFixSide * kFixSide; k += 64) {
            const int i = i1 + k / kFixSide - kFixHalf, j = j1 + k % kFixSide - kFixHalf;
            if (i >= 0 && i < g && j >= 0 && j < g) near = near || m0[j + g * i] == 1 || m0[nc + j + g * i] == 1;
        }
        if (__any(near)) return false;
    }
    G.xc = A.xs[j1];
    G.yc = A.ys[g - 1 - i1];
    const float u0 = A.u[mid], v0 = A.v[mid];
    // the cell's mode coefficients, once (scalar registers)
    float cu[kCntModes], cv[kCntModes];
#pragma unroll
    for (int k = 0; k < (NM ? NM : kCntModes); k++) {
        const size_t o = ((size_t)t * nm + (k < nm ? k : 0)) * nc + c;
        cu[k] = A.ui[o];
        cv[k] = A.vi[o];
    }
    // realization w's velocity sample, extract_velocity :361-404 (the mode sum in k
    // order, then added to the mean). The reference's sum starts from 0; starting from
    // the first product instead changes only the sign of a zero sum, and -0 and +0 have
    // the same ord, so no boundary key lies between them and every rank is the same
    auto sample = [&](int w) __attribute__((always_inline)) {
        float sx = 0.0f, sy = 0.0f;
#pragma unroll
        for (int k = 0; k < (NM ? NM : kCntModes); k++) {
            if (NM || k < nm) {
                const float y = YLDS ? Y[k * nr + w] : Y[w * nm + k];
                sx = k ? sx + cu[k] * y : cu[k] * y;
                sy = k ? sy + cv[k] * y : cv[k] * y;
            }
        }
        return make_float2(u0 + sx, v0 + sy);
    };
    // the cell's sample range (-0 and +0 have the same ords, so the float order serves).
    // A sample with a non-finite component has a non-finite x + y; so has a finite one
    // whose sum overflows, which only hands the cell to the per-lane pass (the same rows).
    // With Yi in LDS two rounds at a time; with Yi read from L2, one round at a time (two
    // were slower on the 200x200x200 fields).
    float x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    bool nf = false;
    if constexpr (YLDS) {
        for (int w = lane; w < nr; w += 128) {
            const float2 v1 = sample(w), v2 = sample(w + 64 < nr ? w + 64 : w);
            nf |= !(isfinite(v1.x + v1.y) && isfinite(v2.x + v2.y));
            x0 = fminf(fminf(x0, v1.x), v2.x);
            x1 = fmaxf(fmaxf(x1, v1.x), v2.x);
            y0 = fminf(fminf(y0, v1.y), v2.y);
            y1 = fmaxf(fmaxf(y1, v1.y), v2.y);
        }
    } else {
        for (int w = lane; w < nr; w += 64) {
            const float2 v = sample(w);
            nf |= !isfinite(v.x + v.y);
            x0 = fminf(x0, v.x);
            x1 = fmaxf(x1, v.x);
            y0 = fminf(y0, v.y);
            y1 = fmaxf(y1, v.y);
        }
    }
    if (__any(nf)) return false;
    x0 = wave_fmin(x0);
    x1 = wave_fmax(x1);
    y0 = wave_fmin(y0);
    y1 = wave_fmax(y1);
    const uint32_t kx0 = fkey(x0), kx1 = fkey(x1), ky0 = fkey(y0), ky1 = fkey(y1);

    // per action (lane a): the ords its samples reach, its landing box
    const bool act = lane < na;
    const int aa = act ? lane : 0;
    const float fca = T.fcos[aa], fsa = T.fsin[aa];
    const int olx = ord_x(G, fca, x0), ohx = ord_x(G, fca, x1);
    const int oly = ord_y(G, fsa, y0), ohy = ord_y(G, fsa, y1);
    const int nx = act ? ohx - olx : 0, ny = act ? ohy - oly : 0;
    const int bxlo = col_of(G, olx), bylo = row_of(G, ohy);
    const int bw = col_of(G, ohx) - bxlo + 1, ncell = bw * (row_of(G, oly) - bylo + 1);
    const RowRewards rr = T.rew[aa];
    bool ok = rr.seg_lin && nx >= 0 && ny >= 0 && nx <= kCntMax && ny <= kCntMax && cnt_pitch(nx) * (ny + 2) <= kCntArea;
    if (OBS && obs_slice)  // no obstacle within the fixed window: boxes must stay inside it
        ok = ok && bxlo >= j1 - kFixHalf && col_of(G, ohx) <= j1 + kFixHalf && bylo >= i1 - kFixHalf &&
             row_of(G, oly) <= i1 + kFixHalf;
    if (!__all(!act || ok)) return false;
    const int ix = wave_incl_scan(nx), iy = wave_incl_scan(ny);
    const int fx = ix - nx, fy = iy - ny;  // first boundary of each action in the raw lists
    const int del = (A.m - 1) / 2;
    const int i_term = A.i_term, j_term = A.j_term, tsg = A.tsg, mwin = A.m, K = A.K;
    // 16-bit bins: bin b is half b & 1 of word b >> 1
    auto bin = [&](int b) __attribute__((always_inline)) {
        return (int)reinterpret_cast<const uint16_t*>(Wv.grid)[b];
    };

    for (int start = 0; start < na;) {
        // the group: the longest run of actions from start with at most kCntMax
        // boundaries per axis and at most kCntArea grid bins (a prefix of the lanes from
        // start: ix, iy grow with a; every action fits alone, so the clamps below change
        // nothing)
        const int bx = start ? __builtin_amdgcn_readlane(ix, start - 1) : 0;
        const int by = start ? __builtin_amdgcn_readlane(iy, start - 1) : 0;
        const bool fits = act && lane >= start && ix - bx <= kCntMax && iy - by <= kCntMax &&
                          cnt_pitch(ix - bx) * (iy - by + 2) <= kCntArea;
        const int end = start + max(1, (int)__popcll(__ballot(fits)));
        const int Tx = min(kCntMax, __builtin_amdgcn_readlane(ix, end - 1) - bx);
        const int Ty = min(kCntMax, __builtin_amdgcn_readlane(iy, end - 1) - by);
        const int P = cnt_pitch(Tx);  // the grid's row pitch

        // the boundaries, x then y: one lane each, bisection, rank
#pragma unroll
        for (int ax = 0; ax < 2; ax++) {
            const int n_a = ax ? ny : nx, lo_a = ax ? oly : olx, first = ax ? fy - by : fx - bx;
            if (lane >= start && lane < end)
                for (int k = 0; k < n_a; k++) Wv.d[first + k] = ((uint32_t)lane << 16) | (uint32_t)(lo_a + 1 + k);
            __builtin_amdgcn_wave_barrier();
            const int nb = ax ? Ty : Tx;
            uint32_t key = 0xffffffffu;
            if (lane < nb) {
                const uint32_t dsc = Wv.d[lane];
                const int a = (int)(dsc >> 16), o = (int)(dsc & 0xffffu);
                key = ax ? boundary_key<1>(G, T.fsin[a], o, ky0, ky1) : boundary_key<0>(G, T.fcos[a], o, kx0, kx1);
            }
            // the descriptors are read: the keys take their place (0xffffffff past nb, above
            // every finite sample's key), and every lane counts #{keys <= mine} and
            // #{keys < mine} over them, four keys per LDS read
            __builtin_amdgcn_wave_barrier();
            Wv.d[lane] = key;
            __builtin_amdgcn_wave_barrier();
            uint32_t cnt = 0u, lt = 0u;
            const uint4* d4 = reinterpret_cast<const uint4*>(Wv.d);
            for (int i = 0; i < nb; i += 4) {
                const uint4 q = d4[i >> 2];
                cnt += (q.x <= key ? 1u : 0u) + (q.y <= key ? 1u : 0u) + (q.z <= key ? 1u : 0u) + (q.w <= key ? 1u : 0u);
                lt += (q.x < key ? 1u : 0u) + (q.y < key ? 1u : 0u) + (q.z < key ? 1u : 0u) + (q.w < key ? 1u : 0u);
            }
            __builtin_amdgcn_wave_barrier();
            // sorted: equal keys share the run [lt, cnt), each lane of the run writes all of it
            if (lane < nb) {
                Wv.c[ax][lane] = (uint8_t)cnt;
                for (uint32_t j = lt; j < cnt; j++) Wv.u[ax][j] = key;
            } else {
                Wv.u[ax][lane] = 0xffffffffu;
            }
        }
        for (int i = lane; i < ((Ty + 2) * P + 1) / 2; i += 64) Wv.grid[i] = 0u;
        __builtin_amdgcn_wave_barrier();

        // every realization's cell of the boundary grid (place_group), with as many
        // bisection levels per axis as the group's boundaries need (at least 4)
        const uint32_t ax0 = lds_addr(Wv.u[0]), ay0 = lds_addr(Wv.u[1]);
        const int lx = max(4, 32 - __clz(Tx)), ly = max(4, 32 - __clz(Ty));
        switch (lx * 8 + ly) {
            case 4 * 8 + 4: place_group<4, 4>(sample, lane, nr, ax0, ay0, P, Wv.grid); break;
            case 4 * 8 + 5: place_group<4, 5>(sample, lane, nr, ax0, ay0, P, Wv.grid); break;
            case 4 * 8 + 6: place_group<4, 6>(sample, lane, nr, ax0, ay0, P, Wv.grid); break;
            case 5 * 8 + 4: place_group<5, 4>(sample, lane, nr, ax0, ay0, P, Wv.grid); break;
            case 5 * 8 + 5: place_group<5, 5>(sample, lane, nr, ax0, ay0, P, Wv.grid); break;
            case 5 * 8 + 6: place_group<5, 6>(sample, lane, nr, ax0, ay0, P, Wv.grid); break;
            case 6 * 8 + 4: place_group<6, 4>(sample, lane, nr, ax0, ay0, P, Wv.grid); break;
            case 6 * 8 + 5: place_group<6, 5>(sample, lane, nr, ax0, ay0, P, Wv.grid); break;
            default: place_group<6, 6>(sample, lane, nr, ax0, ay0, P, Wv.grid); break;
        }
        __builtin_amdgcn_wave_barrier();
        // inclusive prefix: down the columns (lane x - 1 walks column x), then along the
        // rows (lane y - 1 walks row y)
        uint16_t* g16 = reinterpret_cast<uint16_t*>(Wv.grid);
        if (lane <= Tx) {
            int acc = 0;
            for (int y = 1; y <= Ty + 1; y++) {
                acc += g16[y * P + lane + 1];
                g16[y * P + lane + 1] = (uint16_t)acc;
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (lane <= Ty) {
            uint16_t* row = g16 + (lane + 1) * P + 1;
            int acc = 0;
            for (int x = 0; x <= Tx; x++) {
                acc += row[x];
                row[x] = (uint16_t)acc;
            }
        }
        __builtin_amdgcn_wave_barrier();

        // The group's rows. Segment width SW: 16 lanes per row, four rows per pass, when
        // their landing boxes have at most 16 cells; else 64, one row per pass (and
        // chunks of 64 cells). Lane kk of row ar's segment takes landing cells kk, kk + SW,
        // .. of the box in row-major order (ascending S2).
        auto emit = [&](auto sw_tag, int ar, bool arow, int kk) __attribute__((always_inline)) {
            constexpr int SW = decltype(sw_tag)::value;
            const int src = arow ? ar : start;
            const int rlx = __shfl(olx, src, 64), rhx = __shfl(ohx, src, 64);
            const int rly = __shfl(oly, src, 64), rhy = __shfl(ohy, src, 64);
            const int offx = __shfl(fx, src, 64) - bx, offy = __shfl(fy, src, 64) - by;
            const int xlo = __shfl(bxlo, src, 64), ylo = __shfl(bylo, src, 64), W = __shfl(bw, src, 64);
            const int nce = arow ? __shfl(ncell, src, 64) : 0;
            // the row's integer rewards, formed once per cell by its action's lane
            const int r_tn0 = __shfl(rr.tn0, src, 64), r_tn1 = __shfl(rr.tn1, src, 64);
            const int r_tt0 = __shfl(rr.tt0, src, 64), r_tt1 = __shfl(rr.tt1, src, 64);
            const float rW = __builtin_amdgcn_rcpf((float)W);
            // first grid column (row) of the realizations with ord >= o
            auto gx = [&](int o) __attribute__((always_inline)) {
                return o <= rlx ? 0 : (o > rhx ? Tx + 1 : (int)Wv.c[0][offx + o - rlx - 1]);
            };
            auto gy = [&](int o) __attribute__((always_inline)) {
                return o <= rly ? 0 : (o > rhy ? Ty + 1 : (int)Wv.c[1][offy + o - rly - 1]);
            };
            auto rect = [&](int lx, int hx, int ly, int hy) __attribute__((always_inline)) {
                return bin(hy * P + hx) - bin(ly * P + hx) - bin(hy * P + lx) + bin(ly * P + lx);
            };
            // landing cell k: its count, its lanes' clamps, its reward class (terminal * 2 +
            // edge: the reward before the clamp penalties is tn0, tn1, tt0 or tt1), its lanes
            // outside the neighb_gsize window, its S2 cell
            auto cell = [&](int k, int& cnt, int& kc, int& cls, int& oo, uint32_t& sp2) __attribute__((always_inline)) {
                // k / W from the reciprocal: (k + 1/2) / W is at least 1 / (2 W) from an integer
                const int q = (int)(((float)k + 0.5f) * rW);
                const int yi = ylo + q, xi = xlo + (k - q * W);
                const int oax = max(rlx, xi == 0 ? 0 : xi + 1), obx = min(rhx, xi == G.XH ? G.XH + 2 : xi + 1);
                const int oay = max(rly, yi == G.YL ? 0 : 1 + G.YL - yi), oby = min(rhy, yi == 0 ? G.YL + 2 : 1 + G.YL - yi);
                const int lx = gx(oax), hx = gx(obx + 1), ly = gy(oay), hy = gy(oby + 1);
                cnt = rect(lx, hx, ly, hy);
                kc = 0;
                if (oax == 0) kc += rect(lx, gx(1), ly, hy);                // x clamped low
                if (obx == G.XH + 2) kc += rect(gx(G.XH + 2), hx, ly, hy);  // x clamped high
                if (oay == 0) kc += rect(lx, hx, ly, gy(1));                // y clamped low
                if (oby == G.YL + 2) kc += rect(lx, hx, gy(G.YL + 2), hy);  // y clamped high
                const int n1 = g - 1;
                const bool edge = (yi == 0) | (yi == n1) | (xi == 0) | (xi == n1);  // :339-341
                const bool term = ((unsigned)(yi - i_term) < (unsigned)tsg) & ((unsigned)(xi - j_term) < (unsigned)tsg);  // :343-354
                cls = (term ? 2 : 0) + (edge ? 1 : 0);
                const int di = yi - i1, dj = xi - j1;
                oo = (((unsigned)(di + del) >= (unsigned)mwin) | ((unsigned)(dj + del) >= (unsigned)mwin)) ? cnt : 0;
                sp2 = (uint32_t)(xi + g * yi);
            };
            int nuq = 0, kcl = 0, oow = 0, cnt0 = 0, n0 = 0, n1c = 0, n2 = 0, n3 = 0;
            uint32_t sp0 = 0u;
            for (int k0 = 0; k0 < nce || (SW == 64 && k0 == 0); k0 += SW) {
                int cnt = 0, kc = 0, cls = 0, oo = 0;
                uint32_t sp2 = 0u;
                if (k0 + kk < nce) cell(k0 + kk, cnt, kc, cls, oo, sp2);
                if (k0 == 0) {
                    cnt0 = cnt;
                    sp0 = sp2;
                }
                nuq += cnt != 0 ? 1 : 0;
                kcl += kc;
                oow += oo;
                n0 += cls == 0 ? cnt : 0;
                n1c += cls == 1 ? cnt : 0;
                n2 += cls == 2 ? cnt : 0;
                n3 += cls == 3 ? cnt : 0;
                if (SW == 16) break;  // one chunk: boxes of at most 16 cells
            }
            nuq = seg_sum<SW>(nuq);
            kcl = seg_sum<SW>(kcl);
            oow = seg_sum<SW>(oow);
            n0 = seg_sum<SW>(n0);
            n1c = seg_sum<SW>(n1c);
            n2 = seg_sum<SW>(n2);
            n3 = seg_sum<SW>(n3);
            const size_t row = mid * na + ar;
            const int seg0 = lane & ~(SW - 1);  // first lane of this segment
            const bool lead = arow && lane == seg0 + SW - 1;
            unsigned long long slot = 0;
            if (lead && nuq > K) slot = atomicAdd(A.arena_top, (unsigned long long)nuq);
            const unsigned long long arena_off = (unsigned long long)__shfl((long long)slot, seg0 + SW - 1, 64);
            int u = 0;
            for (int k0 = 0; k0 < nce || (SW == 64 && k0 == 0); k0 += SW) {
                int cnt = cnt0;
                uint32_t sp2 = sp0;
                if (k0 > 0) {
                    int kc, cls, oo;
                    cnt = 0;
                    if (k0 + kk < nce) cell(k0 + kk, cnt, kc, cls, oo, sp2);
                }
                const unsigned long long bal = __ballot(cnt != 0);
                const unsigned long long segmask = SW == 64 ? ~0ull : (0xffffull << seg0);
                const int below = (int)__popcll(bal & segmask & ((1ull << lane) - 1ull));
                if (cnt != 0) put_entry(A, row, u + below, nuq, arena_off, sp2, (uint32_t)cnt);
                u += (int)__popcll(bal & segmask);
                if (SW == 16) break;
            }
            if (lead) {
                // the fast pass's sum: per landing cell, plus r_outbound per clamp
                const int64_t tot = (int64_t)n0 * r_tn0 + (int64_t)n1c * r_tn1 + (int64_t)n2 * r_tt0 +
                                    (int64_t)n3 * r_tt1 + (int64_t)A.r_out * (int64_t)kcl;
                const float fs_ = (float)(double)tot;  // compute_mean :563-570
                A.row_nnz[row] = (uint32_t)nuq;
                A.Rout[row] = fs_ / nr;
                if (A.Rsum) A.Rsum[row] = fs_;
                oow_acc += (unsigned long long)oow;
                maxnuq_acc = max(maxnuq_acc, (uint32_t)nuq);
                if (nuq > K) {
                    overflow_acc++;
                    if (arena_off + (unsigned long long)nuq > A.arena_cap) atomicOr(A.stats + 4, 1ull);
                    else put_arena_offset(A, row, arena_off);
                }
            }
        };
        for (int a0 = start; a0 < end; a0 += 4) {
            int mx = 0;  // the largest landing box of rows a0 .. a0 + 3
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (a0 + r < end) mx = max(mx, __builtin_amdgcn_readlane(ncell, a0 + r));
            if (mx <= 16) {
                const int ar = a0 + (lane >> 4);
                emit(std::integral_constant<int, 16>{}, ar, ar < end, lane & 15);
            } else {
                for (int r = 0; r < 4 && a0 + r < end; r++) emit(std::integral_constant<int, 64>{}, a0 + r, true, lane);
            }
        }
        __builtin_amdgcn_wave_barrier();  // the tables and grid are the next group's
        start = end;
    }
    return true;
}

// LDS bytes of mc_wave<NW>.
__host__ __device__ inline size_t wave_lds_bytes(int NW, int nr, int nm, bool stage_y) {
    return sizeof(float2) * (size_t)nr + NW * sizeof(uint32_t) * kHistWords +
           (stage_y ? sizeof(float) * (size_t)nr * nm : 0);
}



// Wave kernel: one wavefront per (cell, action) segment. A block of NW waves takes tc
// consecutive cells of one slice; for each cell its threads form the velocity sample of
// every realization once (extract_velocity, kept in LDS), then wave wv moves actions
// wv, wv + NW, ... over all realizations in rounds of 64 (one round = one reward chunk):
//   pass 1: move; S2 offset (di, dj) counted in the wave's LDS histogram over the fixed
//           window |di|, |dj| <= 7 (bin order = ascending S2; atomics, as the counts
//           are integers); wave min/max of di and dj; the row reward in the canonical
//           order (chunk tree sums added in ascending chunk order; while every chunk is
//           integer-valued the exact int32 sum stands in for them);
//   pass 2: only when the offsets leave the fixed window: the moves again, counted
//           by the range-compressed key (di - dimin) * Wj + (dj - djmin), which also
//           orders S2 ascending; and the lanes outside the neighb_gsize window;
//   emit:   nonzero bins in key order via a wave scan (the count-by-key), to the row's K
//           staging slots or the overflow arena.
// No block barrier separates the actions. A segment whose key range exceeds kWaveBins
// is appended to defer_list, and the block-level kernel (list mode) builds that row
// exactly as it builds every row in grid mode.
template <int NW, bool OBS, bool UNIT, bool ALPHA>
__global__ __launch_bounds__(NW * 64) void mc_wave(const BuildArgs A) {
    constexpr int WG = NW * 64;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nr = A.nr, nm = A.nm, g = A.g, nc = A.nc, na = A.na;
    float2* vxy = reinterpret_cast<float2*>(smem_raw);                                  // [nr]
    uint32_t* hist = reinterpret_cast<uint32_t*>(vxy + nr) + (size_t)wv * kHistWords;  // [kWaveBins][kHistCopies] per wave
    uint32_t* hcopy = hist + (lane & (kHistCopies - 1));  // this lane's copy: hcopy[bin * kHistCopies]
    float* ylds = reinterpret_cast<float*>(reinterpret_cast<uint32_t*>(vxy + nr) + (size_t)NW * kHistWords);  // [nm][nr]
    // grid mode: blockIdx.y is the slice, blockIdx.x its tc cells; list mode: one cell of
    // cell_list (the cells mc_count left to this kernel) per block
    int t, c_begin, c_end;
    if (A.cell_list) {
        const uint32_t sc = A.cell_list[blockIdx.x];
        t = (int)(sc / (uint32_t)nc);
        c_begin = (int)(sc - (uint32_t)t * (uint32_t)nc);
        c_end = c_begin + 1;
    } else {
        t = A.t0 + blockIdx.y;
        c_begin = blockIdx.x * A.tc;
        c_end = min(c_begin + A.tc, nc);
    }
    const size_t tslice = (size_t)t * nc;
    for (int i = lane; i < kHistWords; i += 64) hist[i] = 0u;
    if (A.stage_y) {
        const float* src = A.yi + (size_t)t * nr * nm;
        for (int i = tid; i < nr * nm; i += WG) {
            const int w = i / nm, k = i - w * nm;
            ylds[k * nr + w] = src[i];
        }
    }
    const bool obs_slice = A.slice_obs && (A.slice_obs[t] | A.slice_obs[t + 1]);
    const bool uses_rad = A.s && (A.reward_type == 2 || A.reward_type == 3);
    const bool late_slice = t > A.nt - 20;
    constexpr bool alpha_rows = ALPHA;  // alpha columns: separate instances carry their code
    const int del = (A.m - 1) / 2;
    const int nrounds = (nr + 63) / 64;
    unsigned long long oow_acc = 0, overflow_acc = 0;
    uint32_t maxnuq_acc = 0;
    bool bad = false, bad_index = false;

    for (int c = c_begin; c < c_end; c++) {
        const int i1 = c / g, j1 = c - i1 * g;
        const size_t mid = tslice + c;
        __syncthreads();  // the previous cell's actions are done with vxy (and ylds is staged)
        const float u0 = A.u[mid], v0 = A.v[mid];
        for (int w = tid; w < nr; w += WG) {  // extract_velocity :361-404
            float sx = 0.0f, sy = 0.0f;
            for (int k = 0; k < nm; k++) {
                const size_t o = ((size_t)t * nm + k) * nc + c;
                const float y = A.stage_y ? ylds[k * nr + w] : A.yi[((size_t)t * nr + w) * nm + k];
                sx = sx + A.ui[o] * y;
                sy = sy + A.vi[o] * y;
            }
            const float vx = u0 + sx, vy = v0 + sy;
            bad |= !(isfinite(vx) && isfinite(vy));
            vxy[w] = make_float2(vx, vy);
        }
        // obstacle slices: is any cell within +-kFixHalf of this one an obstacle at t or
        // t+1? If not, no lane whose S2 stays in that window can cross one
        // (goes_through_obstacle :444-477 only looks inside the S1-S2 bounding box)
        bool near_obs = false;
        if (OBS && obs_slice) {
            const int32_t* m0 = A.mask + tslice;
            for (int k = tid; k < kFixSide * kFixSide; k += WG) {
                const int i = i1 + k / kFixSide - kFixHalf, j = j1 + k % kFixSide - kFixHalf;
                if (i >= 0 && i < g && j >= 0 && j < g)
                    near_obs = near_obs || m0[j + g * i] == 1 || m0[nc + j + g * i] == 1;
            }
        }
        near_obs = __syncthreads_or(near_obs);
        const bool term1 = is_terminal(A, i1, j1);
        const bool obs1 = A.mask && A.mask[mid] == 1;
        const float rad1 = uses_rad ? A.s[mid] : 0.0f;
        const float xc = A.xs[j1], yc = A.ys[g - 1 - i1];

        for (int a = wv; a < na; a += NW) {
            const float speed = A.act_speed[a], fc = A.act_fcos[a], fs = A.act_fsin[a];
            // step reward of the methods that do not read the scalar field: one per action
            const float rstep = uses_rad ? 0.0f : step_reward(A, speed, 0.0f, 0.0f);
            int dimin = 0x7fffffff, dimax = -0x7fffffff, djmin = 0x7fffffff, djmax = -0x7fffffff;
            int oow = 0, ip = 0;
            bool mixed = false;
            double rsum = 0.0;
            // S2 offset and reward of realization w: transition_calc :522-545 (terminal
            // cells stay with reward 0, obstacle cells stay with r_outbound)
            int lcls = -1;  // alpha columns: the reward class of the lane just moved
            auto lane_offset = [&](int w, bool valid, int& di, int& dj, float& rw) {
                if (term1 || obs1) {
                    di = 0;
                    dj = 0;
                    rw = term1 ? 0.0f : A.r_out;
                    lcls = -1;
                    return;
                }
                const float2 v = vxy[valid ? w : 0];
                int pi, pj;
                lane_move<UNIT>(A, xc, yc, fc, fs, v.x, v.y, late_slice, pi, pj, rw,
                                alpha_rows ? &lcls : nullptr);
                if (!UNIT) {  // impossible for unit cells; guards the gathers below
                    const int ci = min(max(pi, 0), g - 1), cj = min(max(pj, 0), g - 1);
                    bad_index |= valid & ((ci != pi) | (cj != pj));
                    pi = ci;
                    pj = cj;
                }
                if (uses_rad) rw += step_reward(A, speed, rad1, A.s[tslice + nc + pj + g * pi]);
                else rw += rstep;
                if (OBS && obs_slice && crosses_obstacle(A, i1, j1, pi, pj, t)) {
                    rw = A.r_out;
                    lcls = 8;
                }
                di = pi - i1;
                dj = pj - j1;
            };
            // reward of the row: compute_mean's sum in the canonical order
            auto add_reward = [&](bool valid, float rw) {
                const bool okint = !valid || (rw == rintf(rw) && fabsf(rw) < A.int_lim);
                const bool round_int = __all(okint);
                if (!mixed && round_int) {
                    ip += valid ? (int)rw : 0;
                } else {
                    if (!mixed) {
                        rsum = (double)wave_sum(ip);  // exact: all earlier chunks are integers
                        mixed = true;
                    }
                    double cs;
                    if (round_int) {
                        cs = (double)wave_sum(valid ? (int)rw : 0);
                    } else {
                        double p = valid ? (double)rw : 0.0;
#pragma unroll
                        for (int o = 32; o >= 1; o >>= 1) {
                            const double q = __shfl_down(p, o, 64);
                            p = p + q;
                        }
                        cs = __shfl(p, 0, 64);
                    }
                    rsum = rsum + cs;
                }
            };
            // ---- pass 1, one round = one reward chunk of 64 realizations: bounding box
            // of the offsets, and the count in the fixed 15 x 15 offset histogram (bin
            // (di+7)*15 + (dj+7): ascending S2 order too) when within +-7 cells. Full
            // rounds carry no validity test.
            // Every reward a lane can get here is one of: rstep + r_k (+ r_terminal, or the
            // late-slice penalty), r_k = k * r_outbound summed one at a time, k = 0..3
            // (x clamp, y clamp, edge), or r_outbound (obstacles), or 0. If all of them are
            // integers within int_lim the row sum is an exact int32 sum: no per-lane test.
            const RowRewards rwd = row_rewards(A, rstep, late_slice, uses_rad);
            const bool seg_int = rwd.seg_int, seg_lin = rwd.seg_lin;
            const int tt0 = rwd.tt0, tt1 = rwd.tt1, tn0 = rwd.tn0, tn1 = rwd.tn1;
            // Fast pass 1: unit cells, a start cell that moves (neither terminal nor an
            // obstacle), no scalar-field reward, integer rewards linear in the penalty
            // count, and no obstacle test to make: not an obstacle slice, or no obstacle
            // within the fixed window (a row whose lanes then leave that window goes to
            // the block-level kernel). The edge and terminal terms of the reward depend
            // on the landing cell only, so they are added per histogram bin at the emit;
            // a lane only counts its clamps (x, y).
            const bool fast = UNIT && !uses_rad && !(OBS && obs_slice && near_obs) && !term1 && !obs1 && seg_lin &&
                              !alpha_rows;
            int kcl = 0;
            // alpha columns: per-lane lane counts of the 9 reward classes (class c = 2 *
            // (penalties added) + terminal, 8 = obstacle crossing), 7 bits each
            unsigned long long acls = 0ull;
            auto round1 = [&](int w, bool valid) {
                int di, dj;
                float rw;
                lane_offset(w, valid, di, dj, rw);
                if (valid) {
                    dimin = min(dimin, di);
                    dimax = max(dimax, di);
                    djmin = min(djmin, dj);
                    djmax = max(djmax, dj);
                    // bin of the fixed window; an offset outside it lands in some bin, but
                    // then the bounding box leaves the window and the counts are rebuilt
                    const unsigned bin = (unsigned)(di + kFixHalf) * kFixSide + (unsigned)(dj + kFixHalf);
                    atomicAdd(&hcopy[min(bin, (unsigned)kWaveBins - 1u) * kHistCopies], 1u);
                }
                if (seg_int) ip += valid ? (int)rw : 0;
                else add_reward(valid, rw);
                if (alpha_rows && valid && lcls >= 0) acls += 1ull << (7 * lcls);
            };
            // the move of lane_move<true> for a moving lane, cut to what the fast pass
            // needs: clamp with med3 (= the if/else-if clamp for ordered bounds; its
            // penalty is taken when the clamp changed the value; a non-finite velocity
            // fails the build anyway), then the index of lane_move's rint / remainder
            // rounding, written as floor(a) + (fract(a) >= 0.5): equal for every a >= 0
            // (a = k + f: f < 0.5 gives k, f > 0.5 and f == 0.5 give k + 1 in both)
            // The clamp bounds in VGPRs (med3 takes one scalar operand); the column offset
            // carries the fixed window's bias, so dj + bias is formed by the rounding's
            // carry add and bin = di * 15 + (dj + bias).
            constexpr int kBias = kFixHalf * kFixSide + kFixHalf;
            float vxs0 = A.xs0, vxsn = A.xsn, vys0 = A.ys0, vysn = A.ysn;
            int mi1 = -i1, mj1 = kBias - j1;
            float vxc = xc, vyc = yc;  // loaded before the loop, not waited for inside it
            asm volatile("" : "+v"(vxs0), "+v"(vxsn), "+v"(vys0), "+v"(vysn), "+v"(mi1), "+v"(mj1), "+v"(vxc),
                         "+v"(vyc));
            auto move_fast = [&](float2 v, bool valid, auto dt_one) {
                const float vnetx = fc + v.x;
                const float vnety = fs + v.y;
                float xnew, ynew;
                if constexpr (decltype(dt_one)::value) {  // vnet * 1.0f == vnet exactly
                    xnew = vxc + vnetx;
                    ynew = vyc + vnety;
                } else {
                    xnew = vxc + (vnetx * A.dt);
                    ynew = vyc + (vnety * A.dt);
                }
                const float xcl = __builtin_amdgcn_fmed3f(xnew, vxs0, vxsn);
                const float ycl = __builtin_amdgcn_fmed3f(ynew, vys0, vysn);
                const float ax = xcl - vxs0;
                // -(ycl - ysn) up to the sign of a zero, which neither floor nor fract
                // below can tell apart
                const float ay = vysn - ycl;
                const int djb = cvt_flr_i32(ax) + mj1 + (__builtin_amdgcn_fractf(ax) >= 0.5f ? 1 : 0);
                const int di = cvt_flr_i32(ay) + mi1 + (__builtin_amdgcn_fractf(ay) >= 0.5f ? 1 : 0);
                if (valid) {
                    kcl += xcl != xnew ? 1 : 0;
                    kcl += ycl != ynew ? 1 : 0;
                    dimin = min(dimin, di);
                    dimax = max(dimax, di);
                    djmin = min(djmin, djb);
                    djmax = max(djmax, djb);
                    const unsigned bin = (unsigned)(__mul24(di, kFixSide) + djb);
                    // copy (lane ^ (bin >> 3)) % 4 of the bin: any lane-to-copy map counts the
                    // same, and this one gives bins 8 apart (the same LDS bank for one copy)
                    // different copies (about 1% faster on the 100x100 x nt=120 build)
                    const unsigned b2 = min(bin, (unsigned)kWaveBins - 1u);
                    atomicAdd(&hist[b2 * kHistCopies + ((lane ^ (b2 >> 3)) & (kHistCopies - 1))], 1u);
                }
            };
            const int nfull = nr / 64;
            // rounds in pairs, each round's velocity read issued before the previous
            // round's histogram atomic (LDS operations complete in order, so a read
            // issued after an atomic would wait for it); va / vb alternate, no copies
            // Reads past realization nr - 1 (at most 63 entries, into the histograms that
            // follow vxy in LDS) only feed lanes of a round that is not run or not valid.
            auto rounds_fast = [&](auto dt_one) {
                float2 va = vxy[lane];
                int r = 0;
                for (; r + 2 <= nfull; r += 2) {
                    const float2 vb = vxy[(r + 1) * 64 + lane];
                    move_fast(va, true, dt_one);
                    va = vxy[(r + 2) * 64 + lane];
                    move_fast(vb, true, dt_one);
                }
                if (r < nfull) {
                    const float2 vb = vxy[(r + 1) * 64 + lane];
                    move_fast(va, true, dt_one);
                    va = vb;
                    r++;
                }
                if (r < nrounds) move_fast(va, r * 64 + lane < nr, dt_one);
            };
            // one round at a time (enough waves resident to hide the LDS latency)
            auto rounds_plain = [&](auto dt_one) {
                for (int r = 0; r < nfull; r++) move_fast(vxy[r * 64 + lane], true, dt_one);
                if (nfull < nrounds) {
                    const int w = nfull * 64 + lane;
                    move_fast(vxy[w < nr ? w : 0], w < nr, dt_one);
                }
            };
            if (fast) {
                if (A.pipe_rounds) {
                    if (A.dt == 1.0f)
                        rounds_fast(std::true_type{});
                    else
                        rounds_fast(std::false_type{});
                } else {
                    if (A.dt == 1.0f)
                        rounds_plain(std::true_type{});
                    else
                        rounds_plain(std::false_type{});
                }
            } else {
                for (int r = 0; r < nfull; r++) round1(r * 64 + lane, true);
                if (nfull < nrounds) round1(nfull * 64 + lane, nfull * 64 + lane < nr);
            }
            if (!mixed) rsum = (double)wave_sum(ip);
            dimin = wave_min(dimin);
            dimax = wave_max(dimax);
            djmin = wave_min(djmin);
            djmax = wave_max(djmax);
            if (fast) {  // every wave has a valid lane, so these are real offsets plus the bias
                djmin -= kBias;
                djmax -= kBias;
            }
            const bool wave_wide = (int64_t)dimax - dimin >= 65536 || (int64_t)djmax - djmin >= 65536;
            const bool fixed = dimin >= -kFixHalf && dimax <= kFixHalf && djmin >= -kFixHalf && djmax <= kFixHalf;
            // key decode: fixed window, or the range-compressed key of the bounding box
            const uint32_t Wj = fixed ? (uint32_t)kFixSide : (uint32_t)(djmax - djmin + 1);
            const int di0 = fixed ? -kFixHalf : dimin, dj0 = fixed ? -kFixHalf : djmin;
            const uint64_t nkeys = fixed ? (uint64_t)kFixSide * kFixSide : (uint64_t)(dimax - dimin + 1) * Wj;
            const size_t row = mid * na + a;
            if (wave_wide || nkeys > (uint64_t)kWaveBins || (fast && OBS && obs_slice && !fixed)) {
                if (alpha_rows && lane == 0) atomicOr(A.stats + 7, 2ull);  // no alpha rewards there
                for (int i = lane; i < kHistWords; i += 64) hist[i] = 0u;  // partial fixed counts
                if (lane == 0) A.defer_list[atomicAdd(A.defer_count, 1u)] = (uint32_t)row;
                continue;  // the block-level kernel writes this row
            }
            const bool window_miss = dimin + del < 0 || dimax + del >= A.m || djmin + del < 0 || djmax + del >= A.m;
            if (!fixed || window_miss) {
                // ---- pass 2 (spreads beyond +-7 cells): the moves again, for the histogram
                // of the range-compressed keys and for the lanes outside the window
                if (!fixed)
                    for (int i = lane; i < kHistWords; i += 64) hist[i] = 0u;
                for (int r = 0; r < nrounds; r++) {
                    const int w = r * 64 + lane;
                    if (w < nr) {
                        int di, dj;
                        float rw;
                        lane_offset(w, true, di, dj, rw);
                        if (!fixed)
                            atomicAdd(&hcopy[((uint32_t)(di - dimin) * Wj + (uint32_t)(dj - djmin)) * kHistCopies], 1u);
                        oow += ((unsigned)(di + del) >= (unsigned)A.m) | ((unsigned)(dj + del) >= (unsigned)A.m);
                    }
                }
                oow = wave_sum(oow);
            }
            // ---- emit the nonzero bins in key order; lane l owns bins 4l .. 4l+3
            // A key kk < 256 splits into (kk / Wj, kk % Wj), 1 <= Wj <= 256, without a
            // division per bin: q = (kk * ceil(2^16 / Wj)) >> 16 is exact, since the product
            // exceeds kk / Wj by less than 256 / 2^16 = 1/256, which for Wj < 256 stays below
            // the distance 1/Wj from kk / Wj to the next integer (Wj = 256 gives kk >> 8 = 0).
            const uint32_t wrec = (65536u + Wj - 1u) / Wj;
            auto key_split = [&](uint32_t kk, int& di, int& dj) {
                const uint32_t qq = __umul24(kk, wrec) >> 16;
                di = (int)qq + di0;
                dj = (int)(kk - __umul24(qq, Wj)) + dj0;
            };
            // fast rows whose landing box touches neither the grid's edge rows and columns
            // nor a terminal cell: every lane's reward before its clamp penalties is tn0
            const bool plain_box = fast && i1 + dimin > 0 && i1 + dimax < g - 1 && j1 + djmin > 0 &&
                                   j1 + djmax < g - 1 &&
                                   (i1 + dimax < A.i_term || i1 + dimin >= A.i_term + A.tsg ||
                                    j1 + djmax < A.j_term || j1 + djmin >= A.j_term + A.tsg);
            uint32_t cnt4[4];
            int mine = 0;
            int rbin = 0;  // fast rows: sum over the lane's bins of count * (the k = edge reward)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                uint32_t cq = 0u;
#pragma unroll
                for (int c = 0; c < kHistCopies; c++) cq += hist[(lane * 4 + q) * kHistCopies + c];
                cnt4[q] = cq;
                mine += cnt4[q] != 0u;
                if (fast && !plain_box && cnt4[q] != 0u) {
                    const uint32_t kk = (uint32_t)(lane * 4 + q);
                    int di, dj;
                    key_split(kk, di, dj);
                    const int pi = i1 + di, pj = j1 + dj;
                    const int n1 = g - 1;
                    const bool edge = (pi == 0) | (pi == n1) | (pj == 0) | (pj == n1);  // :339-341
                    const bool term = is_terminal(A, pi, pj);                             // :343-354
                    rbin += (int)cnt4[q] * (term ? (edge ? tt1 : tt0) : (edge ? tn1 : tn0));
                }
            }
            if (fast) {
                // every partial sum is a sum of lane rewards tk[t][edge] (each below
                // int_lim, nr of them), so rbin's wave sum fits int32; the clamp part
                // r_outbound * (clamp count) is summed in 64 bits
                const int64_t tot = (plain_box ? (int64_t)tn0 * nr : (int64_t)wave_sum(rbin)) +
                                    (int64_t)A.r_out * (int64_t)wave_sum(kcl);
                rsum = (double)tot;
            }
            const int incl = wave_incl_scan(mine);
            const int nuq = __builtin_amdgcn_readlane(incl, 63);
            int u = incl - mine;
            unsigned long long arena_off = 0;
            if (nuq > A.K) {
                unsigned long long slot = 0;
                if (lane == 0) slot = atomicAdd(A.arena_top, (unsigned long long)nuq);
                arena_off = __shfl(slot, 0, 64);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (cnt4[q] != 0u) {
                    const uint32_t kk = (uint32_t)(lane * 4 + q);
                    int di, dj;
                    key_split(kk, di, dj);
                    put_entry(A, row, u, nuq, arena_off, (uint32_t)((j1 + dj) + g * (i1 + di)), cnt4[q]);
                    u++;
#pragma unroll
                    for (int c = 0; c < kHistCopies; c++)  // all zero again for the next segment
                        hist[kk * kHistCopies + c] = 0u;
                }
            }
            if (alpha_rows) alpha_rewards(A, row, (float)rsum / nr, term1 || obs1, acls, speed, late_slice, nr);
            if (lane == 0) {
                A.row_nnz[row] = (uint32_t)nuq;
                const float fs_ = (float)rsum;  // compute_mean :563-570
                A.Rout[row] = fs_ / nr;
                if (A.Rsum) A.Rsum[row] = fs_;
                oow_acc += (unsigned long long)oow;
                if ((uint32_t)nuq > maxnuq_acc) maxnuq_acc = (uint32_t)nuq;
                if (nuq > A.K) {
                    overflow_acc++;
                    if (arena_off + (unsigned long long)nuq > A.arena_cap) atomicOr(A.stats + 4, 1ull);
                    else put_arena_offset(A, row, arena_off);
                }
            }
        }
    }
    if (__any(bad) && lane == 0) atomicOr(A.stats + 3, 1ull);
    if (__any(bad_index) && lane == 0) atomicOr(A.stats + 6, 1ull);
    if (lane == 0) {
        if (oow_acc) atomicAdd(A.stats + 0, oow_acc);
        if (overflow_acc) atomicAdd(A.stats + 1, overflow_acc);
        if (maxnuq_acc) atomicMax(A.stats + 2, (unsigned long long)maxnuq_acc);
    }
}

// LDS bytes of mc_count: the shared table, the waves' tables, Yi [nm][nr] (YLDS).
__host__ __device__ inline size_t count_lds_bytes(int nr, int nm, bool ylds) {
    return sizeof(CountWave) * kCntWaves + sizeof(CountShared) + (ylds ? sizeof(float) * (size_t)nr * nm : 0);
}

// Counting kernel: the cells count_cell can take, all their rows at once. A block of
// kCntWaves waves takes tc_count consecutive cells of one slice; each wave takes the
// block's next cell until none is left and counts it alone. A cell that count_cell
// declines goes to left_cells for the per-row kernel's list mode. Only for unit cells, no
// scalar-field reward, no alpha columns and dt > 0 (the host checks).
template <bool YLDS, int NM>
__global__ __launch_bounds__(kCntWaves * 64) void mc_count(const BuildArgs A) {
    constexpr int WG = kCntWaves * 64;
    extern __shared__ __align__(256) unsigned char smem_raw[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nr = A.nr, nm = A.nm, nc = A.nc, na = A.na;
    CountWave* waves = reinterpret_cast<CountWave*>(smem_raw);
    CountShared& T = *reinterpret_cast<CountShared*>(smem_raw + sizeof(CountWave) * kCntWaves);
    float* ylds = reinterpret_cast<float*>(smem_raw + sizeof(CountWave) * kCntWaves + sizeof(CountShared));  // [nm][nr]
    const int t = A.t0 + blockIdx.y;
    // the action table, one entry per lane (kernel-argument arrays take uniform indices only)
    for (int a = 0; a < na; a++)
        if (tid == a) {
            T.fcos[a] = A.act_fcos[a];
            T.fsin[a] = A.act_fsin[a];
            T.speed[a] = A.act_speed[a];
            // the rewards of a row depend on its action and the slice only: once per block
            T.rew[a] = row_rewards(A, step_reward(A, A.act_speed[a], 0.0f, 0.0f), A.t0 + (int)blockIdx.y > A.nt - 20, false);
        }
    const int c_begin = blockIdx.x * A.tc_count, c_end = min(c_begin + A.tc_count, nc);
    if (tid == 0) T.next_cell = (unsigned)c_begin;
    const float* src = A.yi + (size_t)t * nr * nm;
    if (YLDS)
        for (int i = tid; i < nr * nm; i += WG) {
            const int w = i / nm, k = i - w * nm;
            ylds[k * nr + w] = src[i];
        }
    __syncthreads();
    const bool obs_slice = A.slice_obs && (A.slice_obs[t] | A.slice_obs[t + 1]);
    const bool late_slice = t > A.nt - 20;
    unsigned long long oow_acc = 0, overflow_acc = 0, counted = 0;
    uint32_t maxnuq_acc = 0;
    CountGeo geo;
    geo.xs0 = A.xs0;
    geo.xsn = A.xsn;
    geo.ys0 = A.ys0;
    geo.ysn = A.ysn;
    geo.dt = A.dt;
    geo.XH = unit_index(A.xsn - A.xs0);
    geo.YL = unit_index(A.ysn - A.ys0);
    CountWave& Wv = waves[wv];
    // the bisection's address arithmetic needs the tables on 256-byte boundaries; were the
    // dynamic LDS placed otherwise, every cell would go to the per-row kernel instead
    const bool aligned = (lds_addr(smem_raw) & 255u) == 0u;
    for (;;) {
        unsigned cc = 0;
        if (lane == 0) cc = atomicAdd(&T.next_cell, 1u);
        const int c = __builtin_amdgcn_readfirstlane(__shfl((int)cc, 0, 64));  // wave-uniform
        if (c >= c_end) break;
        const bool done = aligned && count_cell<YLDS, NM>(A, YLDS ? ylds : src, Wv, T, geo, t, c, late_slice,
                                                           obs_slice, oow_acc, overflow_acc, maxnuq_acc);
        if (done)
            counted++;
        else if (lane == 0)
            A.left_cells[atomicAdd(A.left_count, 1u)] = (uint32_t)((size_t)t * nc + c);
    }
    // rows were written by the last lane of their segment: the wave's totals
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        oow_acc += __shfl_xor(oow_acc, o, 64);
        overflow_acc += __shfl_xor(overflow_acc, o, 64);
        maxnuq_acc = max(maxnuq_acc, (uint32_t)__shfl_xor((int)maxnuq_acc, o, 64));
    }
    if (lane == 0) {
        if (oow_acc) atomicAdd(A.stats + 0, oow_acc);
        if (overflow_acc) atomicAdd(A.stats + 1, overflow_acc);
        if (maxnuq_acc) atomicMax(A.stats + 2, (unsigned long long)maxnuq_acc);
        if (counted) atomicAdd(A.stats + 8, counted);
    }
}

// (c) The row holding the first entry of each 256-entry window (row_ptr[r] <= 256 w <
// row_ptr[r + 1]), and in wrow[nwin] the row holding the last entry; one thread per row.
__global__ __launch_bounds__(256) void window_rows(const uint32_t* __restrict__ row_ptr, int64_t rows, int64_t nnz,
                                                   int64_t nwin, uint32_t* __restrict__ wrow) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint32_t b = row_ptr[r], e = row_ptr[r + 1];
    for (uint32_t w = (b + kCompWin - 1) / kCompWin; w * (uint32_t)kCompWin < e; w++) wrow[w] = (uint32_t)r;
    if ((int64_t)b < nnz && (int64_t)e >= nnz) wrow[nwin] = (uint32_t)r;
}

// (c) CSR entries and the compressed stream, one wavefront per window of kCompWin
// entries. Each lane of the window's rows marks that row's entries in LDS; each entry
// then reads its staging slot (or its overflow-arena slot; PACKED: 4-byte slots, count
// << 16 | sp2) and writes
//   nz[j] = (col, val): col = (t+1)*nc + sp2, val = count/nr (reduce_kernel :635-637:
//   S2 = state1D_from_spid(t+1, sp_id2), prob = count/nrzns);
//   cz[j] (when cz is given): count << 16 | col - base, when the window keeps 4-byte
//   entries (the rule at kCompWin, with span for kCompSpan), else 0 and cbase[w] = -1.
// Entries past nnz in the last window get cz 0. ncomp counts the windows kept.
template <bool PACKED>
__global__ __launch_bounds__(256) void csr_emit(const uint2* __restrict__ staging, const uint2* __restrict__ arena,
                                                const uint32_t* __restrict__ row_ptr,
                                                const uint32_t* __restrict__ wrow, int2* __restrict__ nz,
                                                uint32_t* __restrict__ cz, int32_t* __restrict__ cbase,
                                                int64_t nnz, int64_t nwin, int K, uint32_t rows_per_slice,
                                                int64_t nc, int nr, double inv_nr, int span,
                                                unsigned long long* __restrict__ ncomp) {
    // per entry of the window: its row, and its slot (staging index, or arena index with
    // the top bit set)
    __shared__ uint32_t rowof[4][kCompWin];
    __shared__ unsigned long long slotof[4][kCompWin];
    constexpr unsigned long long kArenaBit = 1ull << 63;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t wstride = (int64_t)gridDim.x * 4;
    unsigned long long kept = 0;
    int64_t w = (int64_t)blockIdx.x * 4 + wv;
    // rows rlo .. rhi hold every entry of window w (rhi: the next window's first row, or
    // the last entry's row; the empty rows of the last slice lie past it)
    int64_t rlo = w < nwin ? (int64_t)wrow[w] : 0, rhi = w < nwin ? (int64_t)wrow[w + 1] : 0;
    for (; w < nwin; w += wstride) {
        const int64_t j0 = w * kCompWin, j1 = min(j0 + (int64_t)kCompWin, nnz);
        const int64_t wn = w + wstride;  // the next window's rows, loaded ahead
        const int64_t nlo = wn < nwin ? (int64_t)wrow[wn] : 0, nhi = wn < nwin ? (int64_t)wrow[wn + 1] : 0;
        for (int64_t r = rlo + lane; r <= rhi; r += 64) {
            const uint32_t rb = row_ptr[r], re = row_ptr[r + 1];
            const int64_t b = max((int64_t)rb, j0), e = min((int64_t)re, j1);
            if (b >= e) continue;
            // entry j's slot is slot + j (modulo 2^64), plus the arena bit
            unsigned long long slot = (unsigned long long)r * (unsigned)K - rb, arena_bit = 0ull;
            if (re - rb > (uint32_t)K) {  // an overflow row: its arena offset is in its first slot
                if (PACKED) {
                    slot = (unsigned long long)reinterpret_cast<const uint32_t*>(staging)[(int64_t)r * K] - rb;
                } else {
                    const uint2 h = staging[(int64_t)r * K];
                    slot = ((unsigned long long)h.x | ((unsigned long long)h.y << 32)) - rb;
                }
                arena_bit = kArenaBit;
            }
            for (int64_t j = b; j < e; j++) {
                rowof[wv][j - j0] = (uint32_t)r;
                slotof[wv][j - j0] = (slot + (unsigned long long)j) | arena_bit;
            }
        }
        __builtin_amdgcn_wave_barrier();
        int32_t col[4];
        uint32_t cnt[4];
        bool ok = true;
        int32_t cmin = 0x7fffffff, cmax = -1;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t j = j0 + u * 64 + lane;
            col[u] = -1;
            cnt[u] = 0u;
            if (j < j1) {
                const uint32_t r = rowof[wv][u * 64 + lane];
                const unsigned long long sl = slotof[wv][u * 64 + lane];
                uint2 e;
                if (PACKED) {
                    const uint32_t v = (sl & kArenaBit) ? reinterpret_cast<const uint32_t*>(arena)[sl & ~kArenaBit]
                                                        : reinterpret_cast<const uint32_t*>(staging)[sl];
                    e = make_uint2(v & 0xffffu, v >> 16);
                } else {
                    e = (sl & kArenaBit) ? arena[sl & ~kArenaBit] : staging[sl];
                }
                const int32_t c = (int32_t)((int64_t)(r / rows_per_slice + 1) * nc + e.x);
                const float p = (float)e.y / nr;
                nz[j] = make_int2(c, __float_as_int(p));
                col[u] = c;
                cnt[u] = e.y;
                ok = ok && e.y < 65536u && __float_as_int((float)((double)e.y * inv_nr)) == __float_as_int(p);
                cmin = min(cmin, c);
                cmax = max(cmax, c);
            }
        }
        __builtin_amdgcn_wave_barrier();  // rowof and slotof are the next window's
        rlo = nlo;
        rhi = nhi;
        if (!cz) continue;
        for (int o = 32; o >= 1; o >>= 1) {
            cmin = min(cmin, __shfl_xor(cmin, o, 64));
            cmax = max(cmax, __shfl_xor(cmax, o, 64));
        }
        ok = __all(ok) && cmax >= 0 && (int64_t)cmax - cmin < span;
#pragma unroll
        for (int u = 0; u < 4; u++)
            cz[j0 + u * 64 + lane] = ok && col[u] >= 0 ? (cnt[u] << 16) | (uint32_t)(col[u] - cmin) : 0u;
        if (lane == 0) cbase[w] = ok ? cmin : -1;
        kept += ok ? 1ull : 0ull;
    }
    if (lane == 0 && kept) atomicAdd(ncomp, kept);
}

__global__ void fill_tail(uint32_t* row_nnz, float* R, float* Rsum, int64_t from, int64_t to) {
    int64_t i = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < to) {
        row_nnz[i] = 0;
        R[i] = 0.0f;
        if (Rsum) Rsum[i] = 0.0f;
    }
}

__global__ void slice_any(const int32_t* __restrict__ mask, int64_t nc, int32_t* flags) {
    const int t = blockIdx.y;
    int any = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (int64_t)gridDim.x * blockDim.x)
        any |= (mask[t * nc + i] == 1);
    if (__any(any) && (threadIdx.x & 63) == 0) atomicOr(flags + t, 1);
}

// ---- export to the reference COO order (rows a*ns + s)
__global__ void counts_action_major(const uint32_t* __restrict__ row_ptr, uint32_t* __restrict__ out,
                                    int64_t ns, int64_t na) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // action-major row
    if (i >= ns * na) return;
    int64_t a = i / ns, s = i - a * ns;
    int64_t r = s * na + a;
    out[i] = row_ptr[r + 1] - row_ptr[r];
}

template <typename IT>
__global__ void export_rows(const uint32_t* __restrict__ row_ptr, const int2* __restrict__ nz,
                           const uint32_t* __restrict__ off_am, const float* __restrict__ R,
                           IT* __restrict__ S1, IT* __restrict__ S2, float* __restrict__ val,
                           float* __restrict__ Ram, int64_t ns, int64_t na) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // state-major row
    if (r >= ns * na) return;
    int64_t s = r / na, a = r - s * na;
    int64_t am = a * ns + s;
    uint32_t b = row_ptr[r], e = row_ptr[r + 1], d = off_am[am];
    for (uint32_t j = b; j < e; j++, d++) {
        int2 x = nz[j];
        if (S1) S1[d] = (IT)am;
        if (S2) S2[d] = (IT)x.x;
        if (val) val[d] = __int_as_float(x.y);
    }
    if (Ram) Ram[am] = R[r];
}

// ---- state-major R -> the reference's action-major order (rows a*ns + s)
__global__ void r_action_major(const float* __restrict__ R, float* __restrict__ out, int64_t ns, int64_t na) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // action-major row
    if (i >= ns * na) return;
    const int64_t a = i / ns, s = i - a * ns;
    out[i] = R[s * na + a];
}

// ---- import from the reference COO order
__global__ void counts_state_major(const uint32_t* __restrict__ rp_am, uint32_t* __restrict__ out,
                                   int64_t ns, int64_t na) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // state-major row
    if (r >= ns * na) return;
    int64_t s = r / na, a = r - s * na;
    int64_t am = a * ns + s;
    out[r] = rp_am[am + 1] - rp_am[am];
}

__global__ void import_rows(const uint32_t* __restrict__ rp_am, const uint32_t* __restrict__ row_ptr,
                            const int32_t* __restrict__ S2, const float* __restrict__ val,
                            const float* __restrict__ Ram, int2* __restrict__ nz,
                            float* __restrict__ R, int64_t ns, int64_t na) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ns * na) return;
    int64_t s = r / na, a = r - s * na;
    int64_t am = a * ns + s;
    uint32_t b = rp_am[am], e = rp_am[am + 1], d = row_ptr[r];
    for (uint32_t j = b; j < e; j++, d++) nz[d] = make_int2(S2[j], __float_as_int(val[j]));
    R[r] = Ram[am];
}

// ------------------------------------------------------------------ host helpers

// OBS: the goes_through_obstacle test (:444-477) is compiled in only for models whose
// mask has obstacles; UNIT: the unit-cell move. Both cost registers the common kernel
// does not need.
using BuildFn = void (*)(const BuildArgs);

template <int WG>
BuildFn pick_build(bool obs, bool unit) {
    if (obs) return unit ? mc_move_count<WG, true, true> : mc_move_count<WG, true, false>;
    return unit ? mc_move_count<WG, false, true> : mc_move_count<WG, false, false>;
}

BuildFn pick_build(int wg, bool obs, bool unit) {
    if (wg == 256) return pick_build<256>(obs, unit);
    if (wg == 512) return pick_build<512>(obs, unit);
    return pick_build<1024>(obs, unit);
}

template <int NW>
BuildFn pick_wave(bool obs, bool unit, bool alpha) {
    if (alpha) {
        if (obs) return unit ? mc_wave<NW, true, true, true> : mc_wave<NW, true, false, true>;
        return unit ? mc_wave<NW, false, true, true> : mc_wave<NW, false, false, true>;
    }
    if (obs) return unit ? mc_wave<NW, true, true, false> : mc_wave<NW, true, false, false>;
    return unit ? mc_wave<NW, false, true, false> : mc_wave<NW, false, false, false>;
}

BuildFn pick_wave(int nw, bool obs, bool unit, bool alpha) {
    if (nw == 4) return pick_wave<4>(obs, unit, alpha);
    if (nw == 8) return pick_wave<8>(obs, unit, alpha);
    return pick_wave<16>(obs, unit, alpha);
}

// Wavefronts per block of the per-row kernel: 4. Per-cell barriers cost less with small
// blocks, and 4-wave blocks measured fastest on both the 25x25 and the 100x100 workloads
// (100x100 x nt=120: 108 ms against 121 ms with 8 waves and 254 ms with 16). 0 when even
// that does not fit the LDS (velocity samples of nrzns realizations).
// The counting kernel for Yi in LDS or not, its mode sums unrolled for 4, 5 or 6 modes.
BuildFn pick_count(bool ylds, int nm) {
    if (ylds) return nm == 4 ? mc_count<true, 4> : nm == 5 ? mc_count<true, 5> : nm == 6 ? mc_count<true, 6> : mc_count<true, 0>;
    return nm == 4 ? mc_count<false, 4> : nm == 5 ? mc_count<false, 5> : nm == 6 ? mc_count<false, 6> : mc_count<false, 0>;
}

int choose_nw(int nr, int nm, bool stage_y) { return wave_lds_bytes(4, nr, nm, stage_y) <= 160 * 1024 ? 4 : 0; }

e2e_status dispatch_build(const BuildArgs& A, int wg, bool obs, bool unit, dim3 grid, size_t lds,
                          hipStream_t s) {
    BuildFn fn = pick_build(wg, obs, unit);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            set_error("hipFuncSetAttribute(%zu B LDS): %s", lds, hipGetErrorString(e));
            return E2E_ERR_HIP;
        }
    }
    hipLaunchKernelGGL(fn, grid, dim3(wg), lds, s, A);
    return E2E_OK;
}

// Block size: the one that keeps the most waves resident per CU, given the 160 KiB of
// LDS and the kernel's VGPR count (512 VGPRs per SIMD lane, 4 SIMDs per CU, at most 8
// waves per SIMD); variants that spill to scratch are avoided; the smaller on ties.
int choose_wg(int nr, int D, int nm, bool stage_y, bool obs, bool unit, int forced) {
    if (forced == 256 || forced == 512 || forced == 1024) return forced;
    int best = 256, best_w = -1;
    for (int wg : {256, 512, 1024}) {
        const int R = (nr + wg - 1) / wg;
        const size_t lds = build_lds_bytes(wg, R, D, nm, nr, stage_y);
        int blocks = lds > 0 ? (int)((160 * 1024) / lds) : 0;
        hipFuncAttributes fa{};
        int per_simd = 8;
        bool spills = false;
        if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(pick_build(wg, obs, unit))) == hipSuccess) {
            if (fa.numRegs > 0) per_simd = std::min(8, 512 / fa.numRegs);
            spills = fa.localSizeBytes > 0;
        } else {
            (void)hipGetLastError();
        }
        blocks = std::min(blocks, (4 * per_simd) / (wg / 64));
        int waves = blocks * wg / 64;
        if (spills) waves /= 2;
        if (waves > best_w) {
            best_w = waves;
            best = wg;
        }
    }
    return best;
}

void host_actions(const e2e_params* p, float* speed, float* angle, float* fcos, float* fsin) {
    // populate_ac_angles :1403-1408 (double, stored float), populate_ac_speeds :1412-1434
    const int nsp = p->num_ac_speeds, nan_ = p->num_ac_angles;
    std::vector<float> sp(nsp), an(nan_);
    for (int i = 0; i < nan_; i++) an[i] = (float)(i * (2 * M_PI) / nan_);
    if (nsp == 1) {
        sp[0] = p->F;
    } else {
        const float delF = p->F / (nsp);
        for (int i = 0; i < nsp; i++) sp[i] = (i + 1) * delF;
    }
    for (int i = 0; i < nsp; i++)
        for (int j = 0; j < nan_; j++) {
            const int idx = j + nan_ * i;  // populate_actions :1449
            if (speed) speed[idx] = sp[i];
            if (angle) angle[idx] = an[j];
            // F*cosf(ac_angle), F*sinf(ac_angle) of move :284-285, taken once per action
            if (fcos) fcos[idx] = sp[i] * cosf(an[j]);
            if (fsin) fsin[idx] = sp[i] * sinf(an[j]);
        }
}

e2e_status validate_params(const e2e_params* p) {
    E2E_CHECK(p, E2E_ERR_INVALID, "params is NULL");
    E2E_CHECK(p->gsize >= 2 && p->gsize <= 32767, E2E_ERR_INVALID, "gsize=%d out of range [2, 32767]", p->gsize);
    E2E_CHECK(p->nt >= 2, E2E_ERR_INVALID, "nt=%d must be >= 2", p->nt);
    E2E_CHECK(p->nrzns >= 1, E2E_ERR_INVALID, "nrzns=%d must be >= 1", p->nrzns);
    E2E_CHECK(p->nmodes >= 0 && p->nmodes <= kMaxModes, E2E_ERR_CAPACITY, "nmodes=%d must be in [0, %d]", p->nmodes, kMaxModes);
    E2E_CHECK(p->num_ac_speeds >= 1 && p->num_ac_angles >= 1, E2E_ERR_INVALID, "need >= 1 speed and angle");
    E2E_CHECK(p->num_ac_speeds * p->num_ac_angles <= kMaxActions, E2E_ERR_CAPACITY, "at most %d actions", kMaxActions);
    E2E_CHECK(p->reward_type >= 0 && p->reward_type <= 4, E2E_ERR_INVALID, "reward_type=%d not in 0..4", p->reward_type);
    E2E_CHECK(p->neighb_gsize >= 1 && (p->neighb_gsize % 2) == 1, E2E_ERR_INVALID, "neighb_gsize must be odd (build_model.cu:1028)");
    E2E_CHECK(p->term_subgrid_size >= 0, E2E_ERR_INVALID, "term_subgrid_size < 0");
    E2E_CHECK(p->dx > 0.0f && p->dy > 0.0f && std::isfinite(p->x0) && std::isfinite(p->y0) &&
                  std::isfinite(p->dt) && std::isfinite(p->F),
              E2E_ERR_INVALID, "dx, dy must be > 0 and x0, y0, dt, F finite");
    return E2E_OK;
}

}  // namespace

// the build's parameter checks, for the other entry points that take e2e_params
e2e_status check_params(const e2e_params* p) { return validate_params(p); }

// ------------------------------------------------------------------ build

e2e_status model_build(const e2e_params* p, const e2e_fields* f, const e2e_build_opts* o,
                       Model** out) {
    const double t_start = now_ms();
    const bool trace = std::getenv("E2E_TRACE") != nullptr;
    double t_mark = t_start;
    auto phase = [&](const char* what) {
        if (!trace) return;
        const double t = now_ms();
        std::fprintf(stderr, "[e2e build] %-28s %9.3f ms\n", what, t - t_mark);
        t_mark = t;
    };
    E2E_TRY(validate_params(p));
    E2E_CHECK(f && f->u && f->v && f->yi && (p->nmodes == 0 || (f->ui && f->vi)), E2E_ERR_INVALID,
              "fields u, v, yi (and ui, vi when nmodes > 0) are required");
    E2E_CHECK(!(p->reward_type == 2 || p->reward_type == 3) || f->s, E2E_ERR_INVALID,
              "reward_type %d needs the scalar field s", p->reward_type);
    e2e_build_opts opts{};
    if (o) opts = *o;
    const int K = opts.slot_k > 0 ? opts.slot_k : 16;
    const int D = opts.radix_bits > 0 ? opts.radix_bits : 8;
    E2E_CHECK(D >= 1 && D <= 8, E2E_ERR_INVALID, "radix_bits=%d not in 1..8", D);
    E2E_CHECK(K >= 1 && K <= 4096, E2E_ERR_INVALID, "slot_k=%d not in 1..4096", K);

    Model* m = new Model();
    m->p = *p;
    m->built = true;
    const int64_t g = p->gsize, nt = p->nt, nc = g * g, na = (int64_t)p->num_ac_speeds * p->num_ac_angles;
    m->nc = nc;
    m->nt = nt;
    m->ns = nc * nt;
    m->na = na;
    m->rows = m->ns * na;
    e2e_status st = model_alloc_stream(m, opts.stream);
    if (st != E2E_OK) {
        delete m;
        return st;
    }
    auto fail = [&](e2e_status s) {
        free_model(m);
        return s;
    };
    hipStream_t s = m->stream;
    const int nm = p->nmodes, nr = p->nrzns;

    // fields on the device
    DevBuf dU, dV, dUi, dVi, dYi, dS, dM;
    const float *u = f->u, *v = f->v, *ui = f->ui, *vi = f->vi, *yi = f->yi, *sf = f->s;
    const int32_t* mask = f->mask;
    if (!opts.fields_on_device) {
        auto up = [&](DevBuf& b, const void* h, size_t bytes) -> e2e_status {
            if (!h) return E2E_OK;
            E2E_TRY(dmalloc(b, bytes));
            E2E_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s));
            return E2E_OK;
        };
        const size_t f1 = sizeof(float) * nt * nc, fm = f1 * (nm > 0 ? nm : 1);
        if ((st = up(dU, f->u, f1)) || (st = up(dV, f->v, f1)) ||
            (st = up(dUi, nm ? f->ui : nullptr, fm)) || (st = up(dVi, nm ? f->vi : nullptr, fm)) ||
            (st = up(dYi, f->yi, sizeof(float) * nt * nr * (nm > 0 ? nm : 1))) ||
            (st = up(dS, f->s, f1)) || (st = up(dM, f->mask, sizeof(int32_t) * nt * nc)))
            return fail(st);
        u = dU.as<float>();
        v = dV.as<float>();
        ui = nm ? dUi.as<float>() : nullptr;
        vi = nm ? dVi.as<float>() : nullptr;
        yi = dYi.as<float>();
        sf = f->s ? dS.as<float>() : nullptr;
        mask = f->mask ? dM.as<int32_t>() : nullptr;
    }

    // grid ticks (define_xs_or_ys :1395-1399) and actions
    std::vector<float> xs(g), ys(g);
    for (int i = 0; i < g; i++) {
        xs[i] = p->x0 + i * p->dx;
        ys[i] = p->y0 + i * p->dy;
    }
    BuildArgs A;
    std::memset(&A, 0, sizeof(A));
    host_actions(p, A.act_speed, nullptr, A.act_fcos, A.act_fsin);
    DevBuf dXY, dFlags, dStats;
    if ((st = dmalloc(dXY, sizeof(float) * 2 * g))) return fail(st);
    E2E_HIP(hipMemcpyAsync(dXY.p, xs.data(), sizeof(float) * g, hipMemcpyHostToDevice, s));
    E2E_HIP(hipMemcpyAsync(dXY.as<float>() + g, ys.data(), sizeof(float) * g, hipMemcpyHostToDevice, s));
    if ((st = dmalloc(dStats, sizeof(unsigned long long) * kStatSlots))) return fail(st);
    E2E_HIP(hipMemsetAsync(dStats.p, 0, sizeof(unsigned long long) * kStatSlots, s));
    if (mask) {
        if ((st = dmalloc(dFlags, sizeof(int32_t) * nt))) return fail(st);
        E2E_HIP(hipMemsetAsync(dFlags.p, 0, sizeof(int32_t) * nt, s));
        hipLaunchKernelGGL(slice_any, dim3((unsigned)std::min<int64_t>(ceil_div(nc, 256), 64), (unsigned)nt),
                           dim3(256), 0, s, mask, nc, dFlags.as<int32_t>());
    }

    phase("fields + setup");
    A.u = u; A.v = v; A.ui = ui; A.vi = vi; A.yi = yi; A.s = sf; A.mask = mask;
    A.slice_obs = mask ? dFlags.as<int32_t>() : nullptr;
    A.xs = dXY.as<float>();
    A.ys = dXY.as<float>() + g;
    const int n = (int)g - 1;
    A.xs0 = xs[0]; A.xsn = xs[n]; A.ys0 = ys[0]; A.ysn = ys[n];
    A.Dj = fabsf(xs[1] - xs[0]);
    A.Di = fabsf(ys[1] - ys[0]);
    A.hDj = 0.5 * A.Dj;
    A.hDi = 0.5 * A.Di;
    A.late_pen = 0.05 * (double)p->r_outbound;
    A.cell_diag = (float)(fabsf(xs[1] - xs[0]) * 1.414213);
    A.g = (int)g; A.nc = (int)nc; A.nt = (int)nt; A.nm = nm; A.nr = nr; A.na = (int)na;
    A.i_term = p->i_term; A.j_term = p->j_term; A.tsg = p->term_subgrid_size;
    A.reward_type = p->reward_type; A.m = p->neighb_gsize;
    A.dt = p->dt; A.r_out = p->r_outbound; A.r_term = p->r_terminal; A.alpha = p->alpha;
    A.int_lim = (float)std::min<int64_t>(1 << 20, 2147483647ll / std::max<int64_t>(p->nrzns, 1));
    A.K = K;
    A.radix_bits = D;
    A.stats = dStats.as<unsigned long long>();

    // row-indexed outputs
    const int64_t rows = m->rows;
    DevBuf dNnz, dStage, dArena, dTop;
    if ((st = dmalloc(dNnz, sizeof(uint32_t) * rows))) return fail(st);
    // 4-byte staging entries when every count and landing cell fits 16 bits
    A.packed = nr < 65536 && nc <= 65536 ? 1 : 0;
    const size_t slot_bytes = A.packed ? sizeof(uint32_t) : sizeof(uint2);
    if ((st = dmalloc(dStage, slot_bytes * rows * K))) return fail(st);
    if (pool_malloc(&m->R, sizeof(float) * rows) != hipSuccess) {
        set_error("hipMalloc R failed");
        return fail(E2E_ERR_NOMEM);
    }
    if (opts.keep_rsum && pool_malloc(&m->Rsum, sizeof(float) * rows) != hipSuccess) {
        set_error("hipMalloc Rsum failed");
        return fail(E2E_ERR_NOMEM);
    }
    // overflow arena for rows with more than K distinct S2: 2 entries per row on average
    // covers the named workloads (a retry with a larger arena re-runs the kernel)
    unsigned long long arena_cap = std::max<unsigned long long>(1ull << 20, (unsigned long long)rows * 2);
    if ((st = dmalloc(dTop, sizeof(unsigned long long)))) return fail(st);

    A.staging = dStage.as<uint2>();
    A.row_nnz = dNnz.as<uint32_t>();
    A.Rout = m->R;
    A.Rsum = m->Rsum;

    // last slice: no rows (build_model.cu:1142), R = 0 (:1173-1176)
    {
        const int64_t from = (nt - 1) * nc * na, to = rows;
        hipLaunchKernelGGL(fill_tail, dim3((unsigned)ceil_div(to - from, 256)), dim3(256), 0, s,
                           A.row_nnz, m->R, m->Rsum, from, to);
    }

    const size_t ybytes = sizeof(float) * (size_t)nr * nm;
    const bool stage = ybytes > 0 && ybytes <= 16 * 1024;
    A.stage_y = stage ? 1 : 0;
    A.tc = stage ? 4 : 1;
    const bool unit = A.Dj == 1.0f && A.Di == 1.0f;
    bool any_obs = false;
    if (mask) {
        std::vector<int32_t> flags(nt);
        E2E_HIP(hipMemcpyAsync(flags.data(), dFlags.p, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, s));
        E2E_HIP(hipStreamSynchronize(s));
        for (int32_t fl : flags) any_obs = any_obs || fl != 0;
    }
    m->has_obstacles = any_obs;
    const int wg = choose_wg(nr, D, nm, stage, any_obs, unit, opts.block_threads);
    int nw = opts.count_kernel == 1 ? 0 : choose_nw(nr, nm, stage);
    if (nw && opts.block_threads && wave_lds_bytes(opts.block_threads / 64, nr, nm, stage) <= 160 * 1024)
        nw = opts.block_threads / 64;  // a forced workgroup size applies to the per-row kernel too
    const size_t wave_lds = nw ? wave_lds_bytes(nw, nr, nm, stage) : 0;
    // waves of the per-row kernel resident per SIMD, as its LDS allows (4 SIMDs per CU);
    // below 4 its rounds are paired so that LDS latency overlaps the next round
    E2E_CHECK(opts.paired_rounds >= 0 && opts.paired_rounds <= 2, E2E_ERR_INVALID,
              "paired_rounds must be 0, 1 or 2 (got %d)", opts.paired_rounds);
    A.pipe_rounds = opts.paired_rounds == 2 ? 1
                    : opts.paired_rounds == 1 ? 0
                    : (nw && (int)((160 * 1024) / wave_lds) * nw / 4 < 4 ? 1 : 0);
    m->info.paired_rounds = nw ? A.pipe_rounds : 0;
    E2E_CHECK(opts.lane_moves == 0 || opts.lane_moves == 1, E2E_ERR_INVALID, "lane_moves must be 0 or 1 (got %d)",
              opts.lane_moves);
    E2E_CHECK(opts.comp_span >= 0 && opts.comp_span <= kCompSpan, E2E_ERR_INVALID,
              "comp_span must be 0..65536 (got %d)", opts.comp_span);
    const int comp_span = opts.comp_span > 0 ? opts.comp_span : kCompSpan;
    // alpha columns: rewards for n_alpha values of alpha from this one build
    E2E_CHECK(opts.n_alpha >= 0 && opts.n_alpha <= 64, E2E_ERR_INVALID, "n_alpha=%d not in 0..64", opts.n_alpha);
    A.n_alpha = 0;
    A.Ralpha = nullptr;
    A.rows_total = rows;
    if (opts.n_alpha > 0) {
        E2E_CHECK(opts.alphas, E2E_ERR_INVALID, "n_alpha > 0 needs the alphas array");
        E2E_CHECK(p->reward_type == 4, E2E_ERR_INVALID,
                  "alpha columns need reward_type 4 (the only reward that depends on alpha)");
        E2E_CHECK(nw > 0 && opts.count_kernel == 0, E2E_ERR_CAPACITY,
                  "alpha columns need the per-row kernel (count_kernel 0, velocity samples within its LDS)");
        E2E_CHECK(nr <= 127 * 64, E2E_ERR_CAPACITY, "alpha columns count at most 8128 realizations per row");
        if (pool_malloc(&m->Ralpha, sizeof(float) * (size_t)opts.n_alpha * rows) != hipSuccess) {
            m->Ralpha = nullptr;
            set_error("hipMalloc of the alpha rewards failed");
            return fail(E2E_ERR_NOMEM);
        }
        E2E_HIP(hipMemsetAsync(m->Ralpha, 0, sizeof(float) * (size_t)opts.n_alpha * rows, s));  // last slice: 0
        m->n_alpha = opts.n_alpha;
        A.n_alpha = opts.n_alpha;
        A.Ralpha = m->Ralpha;
        for (int j = 0; j < opts.n_alpha; j++) {
            m->alphas[j] = opts.alphas[j];
            A.alphas[j] = opts.alphas[j];
        }
    }
    m->info.block_threads = wg;
    m->info.wave_blocks = nw;
    m->info.unit_cells = unit ? 1 : 0;
    A.R = (int)ceil_div(nr, wg);
    const size_t lds = build_lds_bytes(wg, A.R, D, nm, nr, stage);
    E2E_CHECK(lds <= 160 * 1024, E2E_ERR_CAPACITY, "build kernel needs %zu B of LDS (nrzns too large)", lds);
    m->info.n_multipass_rows = 0;

    DevBuf dDefer, dDcnt;
    if (nw) {
        if ((st = dmalloc(dDefer, sizeof(uint32_t) * std::max<int64_t>(rows, 1))) ||
            (st = dmalloc(dDcnt, sizeof(unsigned))))
            return fail(st);
        A.defer_list = dDefer.as<uint32_t>();
        A.defer_count = dDcnt.as<unsigned>();
    }
    // counted cells (mc_count) where the fast unit-cell pass applies to the whole model:
    // unit cells, no scalar-field reward, no alpha columns, a forward time step
    const bool uses_rad = sf && (p->reward_type == 2 || p->reward_type == 3);
    // Yi of a slice in the counting block's LDS when it fits
    const bool count_ylds = count_lds_bytes(nr, nm, true) <= 160 * 1024;
    const size_t count_lds = count_lds_bytes(nr, nm, count_ylds);
    const bool counting = nw > 0 && unit && !uses_rad && A.n_alpha == 0 && p->dt > 0.0f && !opts.lane_moves &&
                          nr < 65536 && nm <= kCntModes;
    A.tc_count = 160;
    DevBuf dLeft, dLcnt;
    if (counting) {
        if ((st = dmalloc(dLeft, sizeof(uint32_t) * std::max<int64_t>((nt - 1) * nc, 1))) ||
            (st = dmalloc(dLcnt, sizeof(unsigned))))
            return fail(st);
        A.left_cells = dLeft.as<uint32_t>();
        A.left_count = dLcnt.as<unsigned>();
    }
    phase("row buffers allocated");
    const int64_t nslices = nt - 1;
    const int64_t spl = opts.slices_per_launch > 0 ? opts.slices_per_launch : nslices;
    hipEvent_t e0, e1;
    E2E_HIP(hipEventCreate(&e0));
    E2E_HIP(hipEventCreate(&e1));
    double move_ms = 0.0;
    int launches = 0;
    for (int attempt = 0; attempt < 4; attempt++) {
        if (A.packed && arena_cap >= (1ull << 32)) {  // packed arena offsets are 32-bit
            set_error("overflow arena of %llu entries exceeds 32-bit offsets", arena_cap);
            st = E2E_ERR_CAPACITY;
            break;
        }
        if ((st = dmalloc(dArena, slot_bytes * arena_cap))) break;
        E2E_HIP(hipMemsetAsync(dTop.p, 0, sizeof(unsigned long long), s));
        E2E_HIP(hipMemsetAsync(dStats.p, 0, sizeof(unsigned long long) * kStatSlots, s));
        A.arena = dArena.as<uint2>();
        A.arena_top = dTop.as<unsigned long long>();
        A.arena_cap = arena_cap;
        move_ms = 0.0;
        launches = 0;
        unsigned ndefer = 0;
        if (nw) E2E_HIP(hipMemsetAsync(dDcnt.p, 0, sizeof(unsigned), s));
        if (counting) E2E_HIP(hipMemsetAsync(dLcnt.p, 0, sizeof(unsigned), s));
        BuildFn wf = nw ? pick_wave(nw, any_obs, unit, A.n_alpha > 0) : nullptr;
        if (nw && wave_lds > 64 * 1024)
            E2E_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(wf), hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)wave_lds));
        // time one launch on the stream into move_ms
        auto timed = [&](auto&& launch) -> e2e_status {
            E2E_HIP(hipEventRecord(e0, s));
            E2E_TRY(launch());
            E2E_HIP(hipGetLastError());
            E2E_HIP(hipEventRecord(e1, s));
            E2E_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            E2E_HIP(hipEventElapsedTime(&ms, e0, e1));
            move_ms += ms;
            launches++;
            return E2E_OK;
        };
        for (int64_t t0 = 0; t0 < nslices; t0 += spl) {
            const int64_t cnt = std::min(spl, nslices - t0);
            A.t0 = (int)t0;
            dim3 grid((unsigned)ceil_div(nc, A.tc), (unsigned)cnt);
            if (counting) {
                BuildFn cf = pick_count(count_ylds, nm);
                if (count_lds > 64 * 1024)
                    E2E_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(cf),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)count_lds));
                const dim3 cgrid((unsigned)ceil_div(nc, A.tc_count), (unsigned)cnt);
                st = timed([&] {
                    hipLaunchKernelGGL(cf, cgrid, dim3(kCntWaves * 64), count_lds, s, A);
                    return E2E_OK;
                });
            } else if (nw) {
                st = timed([&] {
                    hipLaunchKernelGGL(wf, grid, dim3(nw * 64), wave_lds, s, A);
                    return E2E_OK;
                });
            } else {
                st = timed([&] { return dispatch_build(A, wg, any_obs, unit, grid, lds, s); });
            }
            if (st) break;
        }
        if (st) break;
        if (counting) {  // the cells mc_count left: the per-row kernel in list mode
            unsigned nleft = 0;
            E2E_HIP(hipMemcpyAsync(&nleft, dLcnt.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
            E2E_HIP(hipStreamSynchronize(s));
            if (nleft) {
                A.cell_list = dLeft.as<uint32_t>();
                st = timed([&] {
                    hipLaunchKernelGGL(wf, dim3(nleft), dim3(nw * 64), wave_lds, s, A);
                    return E2E_OK;
                });
                A.cell_list = nullptr;
                if (st) break;
            }
        }
        if (nw) {  // rows the per-row kernel handed over: the block-level kernel in list mode
            E2E_HIP(hipMemcpyAsync(&ndefer, dDcnt.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
            E2E_HIP(hipStreamSynchronize(s));
            if (ndefer) {
                A.seg_list = dDefer.as<uint32_t>();
                st = timed([&] { return dispatch_build(A, wg, any_obs, unit, dim3(ndefer), lds, s); });
                A.seg_list = nullptr;
                if (st) break;
            }
        }
        m->info.n_deferred_rows = (int32_t)ndefer;
        unsigned long long hs[kStatSlots];
        E2E_HIP(hipMemcpyAsync(hs, dStats.p, sizeof(hs), hipMemcpyDeviceToHost, s));
        E2E_HIP(hipStreamSynchronize(s));
        if (hs[3]) {
            set_error("non-finite velocity sample in the input fields");
            st = E2E_ERR_FIELD;
            break;
        }
        if (hs[6]) {
            set_error("a lane left the grid (check dx, dy, x0, y0)");
            st = E2E_ERR_INVALID;
            break;
        }
        if (hs[7] & 2ull) {
            set_error("alpha columns: a row's S2 spread exceeds the per-row kernel; build each alpha separately");
            st = E2E_ERR_CAPACITY;
            break;
        }
        if (hs[7] & 1ull) {
            set_error("alpha columns: reward values too far apart in magnitude for an exact sum over "
                      "reward classes; build each alpha separately");
            st = E2E_ERR_CAPACITY;
            break;
        }
        m->info.n_out_of_window = (int64_t)hs[0];
        m->info.n_overflow_rows = (int64_t)hs[1];
        m->info.max_row_nnz = (int64_t)hs[2];
        m->info.n_multipass_rows = (int64_t)hs[5];
        m->info.counted_cells = (int32_t)hs[8];
        if (!hs[4]) break;
        // arena exhausted: grow and rebuild
        unsigned long long used = 0;
        E2E_HIP(hipMemcpy(&used, dTop.p, sizeof(used), hipMemcpyDeviceToHost));
        dArena.reset();
        arena_cap = std::max(arena_cap * 4, used + 1024);
        if (attempt == 3) {
            set_error("overflow arena exhausted after 4 attempts");
            st = E2E_ERR_CAPACITY;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return fail(st);
    phase("move/sort/count kernel");
    m->info.move_kernel_ms = move_ms;
    m->info.move_launches = launches;

    // (c) scan of row counts -> row_ptr, then scatter into the CSR
    DevBuf dTmp, dTot;
    if ((st = dmalloc(dTmp, sizeof(unsigned long long) * scan_tmp_words(rows)))) return fail(st);
    if ((st = dmalloc(dTot, sizeof(unsigned long long)))) return fail(st);
    if (pool_malloc(&m->row_ptr, sizeof(uint32_t) * (rows + 1)) != hipSuccess) {
        set_error("hipMalloc row_ptr failed");
        return fail(E2E_ERR_NOMEM);
    }
    hipEvent_t c0, c1, c2;
    E2E_HIP(hipEventCreate(&c0));
    E2E_HIP(hipEventCreate(&c1));
    E2E_HIP(hipEventCreate(&c2));
    E2E_HIP(hipEventRecord(c0, s));
    if ((st = scan_u32(A.row_nnz, m->row_ptr, rows, dTmp.as<unsigned long long>(),
                       dTot.as<unsigned long long>(), s)))
        return fail(st);
    E2E_HIP(hipEventRecord(c1, s));
    unsigned long long total = 0;
    E2E_HIP(hipMemcpyAsync(&total, dTot.p, sizeof(total), hipMemcpyDeviceToHost, s));
    E2E_HIP(hipStreamSynchronize(s));
    E2E_CHECK(total < (1ull << 32) && total < (1ull << 31), E2E_ERR_CAPACITY,
              "nnz=%llu exceeds the 2^31 column-index range", total);
    m->nnz = (int64_t)total;
    phase("row-count scan");
    if ((st = alloc_nz(m, (int64_t)total))) return fail(st);
    // the CSR entries and, when every count fits 16 bits, the lean sweep's compressed
    // stream, in one pass over the staging slots
    E2E_CHECK(rows < (1ll << 32), E2E_ERR_CAPACITY, "build: %lld rows exceed 32-bit row indices", (long long)rows);
    const int64_t nwin = ceil_div(m->nnz, (int64_t)kCompWin);
    const bool comp = nr < 65536;
    DevBuf dWrow, dKept;
    if ((st = dmalloc(dWrow, sizeof(uint32_t) * (size_t)(nwin + 1)))) return fail(st);
    if ((st = dmalloc(dKept, sizeof(unsigned long long)))) return fail(st);
    if (comp && (pool_malloc(&m->vi_cz, sizeof(uint32_t) * (size_t)std::max<int64_t>(nwin, 1) * kCompWin) != hipSuccess ||
                 pool_malloc(&m->vi_cbase, sizeof(int32_t) * (size_t)std::max<int64_t>(nwin, 1)) != hipSuccess)) {
        if (m->vi_cz) pool_free(m->vi_cz);
        m->vi_cz = nullptr;
        m->vi_cbase = nullptr;
        set_error("hipMalloc of the compressed entry stream failed");
        return fail(E2E_ERR_NOMEM);
    }
    E2E_HIP(hipMemsetAsync(dKept.p, 0, sizeof(unsigned long long), s));
    unsigned long long kept = 0;
    if (nwin > 0) {
        hipLaunchKernelGGL(window_rows, dim3((unsigned)ceil_div(rows, 256)), dim3(256), 0, s, m->row_ptr, rows,
                           m->nnz, nwin, dWrow.as<uint32_t>());
        E2E_HIP(hipGetLastError());
        hipLaunchKernelGGL(A.packed ? csr_emit<true> : csr_emit<false>,
                           dim3((unsigned)std::min<int64_t>(ceil_div(nwin, 4), 16384)), dim3(256), 0, s,
                           dStage.as<uint2>(), dArena.as<uint2>(), m->row_ptr, dWrow.as<uint32_t>(), m->nz, m->vi_cz,
                           m->vi_cbase, m->nnz, nwin, K, (uint32_t)(nc * na), nc, nr, 1.0 / (double)nr,
                           comp_span, dKept.as<unsigned long long>());
        E2E_HIP(hipGetLastError());
    }
    E2E_HIP(hipEventRecord(c2, s));
    E2E_HIP(hipMemcpyAsync(&kept, dKept.p, sizeof(kept), hipMemcpyDeviceToHost, s));
    E2E_HIP(hipEventSynchronize(c2));
    E2E_HIP(hipStreamSynchronize(s));
    if (comp) {
        m->vi_cwin = nwin;
        m->vi_cwin_comp = (int64_t)kept;
        m->vi_comp_span = comp_span;
    }
    float ms1 = 0, ms2 = 0;
    E2E_HIP(hipEventElapsedTime(&ms1, c0, c1));
    E2E_HIP(hipEventElapsedTime(&ms2, c1, c2));
    (void)hipEventDestroy(c0);
    (void)hipEventDestroy(c1);
    (void)hipEventDestroy(c2);
    phase("CSR allocated + scattered");
    m->info.scan_ms = ms1;
    m->info.csr_ms = ms2;
    m->info.lanes = (nt - 1) * nc * na * (int64_t)nr;
    // release the staging buffers inside the timed build
    dStage.reset();
    dArena.reset();
    dNnz.reset();
    phase("staging freed");
    m->info.build_ms = now_ms() - t_start;
    *out = m;
    return E2E_OK;
}

e2e_status alloc_nz(Model* m, int64_t nnz) {
    if (pool_malloc(&m->nz, sizeof(int2) * (nnz + kNzPad)) != hipSuccess) {
        m->nz = nullptr;
        set_error("hipMalloc nz (%lld entries) failed", (long long)nnz);
        return E2E_ERR_NOMEM;
    }
    E2E_HIP(hipMemsetAsync(m->nz + nnz, 0, sizeof(int2) * kNzPad, m->stream));
    return E2E_OK;
}

e2e_status model_import(int64_t ns, int64_t na, int64_t nnz, const int32_t* S1, const int32_t* S2,
                        const float* val, const float* R, void* stream, Model** out) {
    E2E_CHECK(ns > 0 && na > 0 && nnz >= 0 && S1 && S2 && val && R, E2E_ERR_INVALID,
              "import: bad arguments");
    E2E_CHECK(na <= kMaxActions, E2E_ERR_CAPACITY, "import: at most %d actions", kMaxActions);
    E2E_CHECK(nnz < (1ll << 31) && ns * na < (1ll << 31), E2E_ERR_CAPACITY, "import: model too large");
    const int64_t rows = ns * na;
    // action-major row pointers; S1 must be sorted by row (the reference writes it so)
    std::vector<uint32_t> rp(rows + 1, 0);
    for (int64_t j = 0; j < nnz; j++) {
        E2E_CHECK(S1[j] >= 0 && S1[j] < rows, E2E_ERR_INVALID, "import: S1[%lld]=%d out of range",
                  (long long)j, S1[j]);
        E2E_CHECK(j == 0 || S1[j] >= S1[j - 1], E2E_ERR_INVALID, "import: S1 not sorted at %lld",
                  (long long)j);
        E2E_CHECK(S2[j] >= 0 && S2[j] < ns, E2E_ERR_INVALID, "import: S2[%lld]=%d out of range",
                  (long long)j, S2[j]);
        rp[S1[j] + 1]++;
    }
    uint32_t max_len = 0;
    for (int64_t r = 0; r < rows; r++) {
        max_len = std::max(max_len, rp[r + 1]);
        rp[r + 1] += rp[r];
    }
    Model* m = new Model();
    m->info.max_row_nnz = max_len;
    m->built = false;
    m->ns = ns;
    m->na = na;
    m->rows = rows;
    m->nnz = nnz;
    e2e_status st = model_alloc_stream(m, stream);
    if (st) {
        delete m;
        return st;
    }
    auto fail = [&](e2e_status s) {
        free_model(m);
        return s;
    };
    hipStream_t s = m->stream;
    DevBuf dRp, dS2, dVal, dR, dCnt, dTmp, dTot;
    if ((st = dmalloc(dRp, sizeof(uint32_t) * (rows + 1))) || (st = dmalloc(dS2, sizeof(int32_t) * nnz)) ||
        (st = dmalloc(dVal, sizeof(float) * nnz)) || (st = dmalloc(dR, sizeof(float) * rows)) ||
        (st = dmalloc(dCnt, sizeof(uint32_t) * rows)) ||
        (st = dmalloc(dTmp, sizeof(unsigned long long) * scan_tmp_words(rows))) ||
        (st = dmalloc(dTot, sizeof(unsigned long long))))
        return fail(st);
    if (pool_malloc(&m->row_ptr, sizeof(uint32_t) * (rows + 1)) != hipSuccess ||
        pool_malloc(&m->R, sizeof(float) * rows) != hipSuccess) {
        set_error("import: device allocation failed");
        return fail(E2E_ERR_NOMEM);
    }
    if ((st = alloc_nz(m, nnz))) return fail(st);
    E2E_HIP(hipMemcpyAsync(dRp.p, rp.data(), sizeof(uint32_t) * (rows + 1), hipMemcpyHostToDevice, s));
    E2E_HIP(hipMemcpyAsync(dS2.p, S2, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
    E2E_HIP(hipMemcpyAsync(dVal.p, val, sizeof(float) * nnz, hipMemcpyHostToDevice, s));
    E2E_HIP(hipMemcpyAsync(dR.p, R, sizeof(float) * rows, hipMemcpyHostToDevice, s));
    const unsigned nb = (unsigned)ceil_div(rows, 256);
    hipLaunchKernelGGL(counts_state_major, dim3(nb), dim3(256), 0, s, dRp.as<uint32_t>(),
                       dCnt.as<uint32_t>(), ns, na);
    if ((st = scan_u32(dCnt.as<uint32_t>(), m->row_ptr, rows, dTmp.as<unsigned long long>(),
                       dTot.as<unsigned long long>(), s)))
        return fail(st);
    hipLaunchKernelGGL(import_rows, dim3(nb), dim3(256), 0, s, dRp.as<uint32_t>(), m->row_ptr,
                       dS2.as<int32_t>(), dVal.as<float>(), dR.as<float>(), m->nz, m->R, ns, na);
    E2E_HIP(hipGetLastError());
    E2E_HIP(hipStreamSynchronize(s));
    *out = m;
    return E2E_OK;
}

template <typename IT>
static e2e_status export_coo_t(const Model* m, IT* S1, IT* S2, float* val, float* R) {
    hipStream_t s = m->stream;
    const int64_t rows = m->rows, nnz = m->nnz;
    DevBuf dCnt, dOff, dTmp, dTot, dS1, dS2, dVal, dR;
    e2e_status st;
    if ((st = dmalloc(dCnt, sizeof(uint32_t) * rows)) || (st = dmalloc(dOff, sizeof(uint32_t) * (rows + 1))) ||
        (st = dmalloc(dTmp, sizeof(unsigned long long) * scan_tmp_words(rows))) ||
        (st = dmalloc(dTot, sizeof(unsigned long long))))
        return st;
    if (S1 && (st = dmalloc(dS1, sizeof(IT) * nnz))) return st;
    if (S2 && (st = dmalloc(dS2, sizeof(IT) * nnz))) return st;
    if (val && (st = dmalloc(dVal, sizeof(float) * nnz))) return st;
    if (R && (st = dmalloc(dR, sizeof(float) * rows))) return st;
    const unsigned nb = (unsigned)ceil_div(rows, 256);
    hipLaunchKernelGGL(counts_action_major, dim3(nb), dim3(256), 0, s, m->row_ptr, dCnt.as<uint32_t>(),
                       m->ns, m->na);
    E2E_TRY(scan_u32(dCnt.as<uint32_t>(), dOff.as<uint32_t>(), rows, dTmp.as<unsigned long long>(),
                     dTot.as<unsigned long long>(), s));
    hipLaunchKernelGGL(export_rows<IT>, dim3(nb), dim3(256), 0, s, m->row_ptr, m->nz, dOff.as<uint32_t>(),
                       m->R, dS1.as<IT>(), dS2.as<IT>(), dVal.as<float>(), dR.as<float>(), m->ns, m->na);
    E2E_HIP(hipGetLastError());
    if (S1) E2E_HIP(hipMemcpyAsync(S1, dS1.p, sizeof(IT) * nnz, hipMemcpyDeviceToHost, s));
    if (S2) E2E_HIP(hipMemcpyAsync(S2, dS2.p, sizeof(IT) * nnz, hipMemcpyDeviceToHost, s));
    if (val) E2E_HIP(hipMemcpyAsync(val, dVal.p, sizeof(float) * nnz, hipMemcpyDeviceToHost, s));
    if (R) E2E_HIP(hipMemcpyAsync(R, dR.p, sizeof(float) * rows, hipMemcpyDeviceToHost, s));
    E2E_HIP(hipStreamSynchronize(s));
    return E2E_OK;
}

e2e_status model_export_coo64(const Model* m, int64_t* S1, int64_t* S2, float* val, float* R) {
    return export_coo_t<int64_t>(m, S1, S2, val, R);
}

e2e_status model_export_coo32(const Model* m, int32_t* S1, int32_t* S2, float* val, float* R) {
    E2E_CHECK(m->rows < (1ll << 31), E2E_ERR_CAPACITY, "row index a*Ns+s exceeds int32");
    return export_coo_t<int32_t>(m, S1, S2, val, R);
}

e2e_status model_export_csr(const Model* m, uint32_t* row_ptr, int32_t* col, float* val, float* R) {
    hipStream_t s = m->stream;
    if (row_ptr) E2E_HIP(hipMemcpyAsync(row_ptr, m->row_ptr, sizeof(uint32_t) * (m->rows + 1), hipMemcpyDeviceToHost, s));
    if (R) E2E_HIP(hipMemcpyAsync(R, m->R, sizeof(float) * m->rows, hipMemcpyDeviceToHost, s));
    if (col || val) {
        std::vector<int2> h(m->nnz);
        E2E_HIP(hipMemcpyAsync(h.data(), m->nz, sizeof(int2) * m->nnz, hipMemcpyDeviceToHost, s));
        E2E_HIP(hipStreamSynchronize(s));
        for (int64_t j = 0; j < m->nnz; j++) {
            if (col) col[j] = h[j].x;
            if (val) {
                int bits = h[j].y;
                std::memcpy(&val[j], &bits, sizeof(float));
            }
        }
    }
    E2E_HIP(hipStreamSynchronize(s));
    return E2E_OK;
}

e2e_status model_export_R_alpha(const Model* m, int32_t j, float* R) {
    E2E_CHECK(m->Ralpha && j >= 0 && j < m->n_alpha, E2E_ERR_INVALID,
              "alpha column %d: the model has %d alpha columns", j, m->n_alpha);
    E2E_CHECK(R, E2E_ERR_INVALID, "R is NULL");
    hipStream_t s = m->stream;
    DevBuf d;
    E2E_TRY(dmalloc(d, sizeof(float) * m->rows));
    hipLaunchKernelGGL(r_action_major, dim3((unsigned)ceil_div(m->rows, 256)), dim3(256), 0, s,
                       m->Ralpha + (size_t)j * m->rows, d.as<float>(), m->ns, m->na);
    E2E_HIP(hipGetLastError());
    E2E_HIP(hipMemcpyAsync(R, d.p, sizeof(float) * m->rows, hipMemcpyDeviceToHost, s));
    E2E_HIP(hipStreamSynchronize(s));
    return E2E_OK;
}

void actions_table(const e2e_params* p, float* speed, float* angle) {
    host_actions(p, speed, angle, nullptr, nullptr);
}

}  // namespace e2e
