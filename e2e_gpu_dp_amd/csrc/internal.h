// internal.h — shared declarations of libe2edp (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/e2edp.h"

namespace e2e {

void set_error(const char* fmt, ...);

#define E2E_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::e2e::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,               \
                             hipGetErrorString(e_));                                    \
            return E2E_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

#define E2E_CHECK(cond, code, ...)                                                      \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::e2e::set_error(__VA_ARGS__);                                              \
            return (code);                                                              \
        }                                                                               \
    } while (0)

#define E2E_TRY(expr)                                                                   \
    do {                                                                                \
        e2e_status s_ = (expr);                                                         \
        if (s_ != E2E_OK) return s_;                                                    \
    } while (0)

constexpr int kMaxActions = 64;
constexpr int kMaxModes = 16;
// The CSR entry array is allocated with this many zero entries (column 0, value 0) past
// nnz, so that streaming kernels may load whole windows of up to kNzPad entries without
// per-entry bounds checks.
constexpr int64_t kNzPad = 4096;
// Compressed entry stream of the lean sweep: windows of kCompWin consecutive CSR entries.
// A window whose columns span less than kCompSpan, and whose probabilities are all
// count / nr for a count < 2^16 that (float)(count * (1.0 / nr)) reproduces bit for bit,
// is kept at 4 bytes per entry (count << 16 | column - base, with base = the window's
// smallest column); any other window has base -1 and is read from the 8-byte entries.
constexpr int kCompWin = 256;
constexpr int kCompSpan = 65536;

// Slices a sweep over a built model recomputes. Transitions go from slice t to t+1 only
// and the last slice's rows have none, so the values of slice t are final from sweep
// nt - t on; at sweep k (1-based) >= 3 the slices t >= nt - k + 2 hold their final
// values in both value buffers (written at sweep k - 2 or before, unchanged since): the
// sweep recomputes slices [0, active_slices(k, nt)) and gives every value, policy and
// residual the full sweep gives.
// Always in [1, nt], for every k.
__host__ __device__ inline int active_slices(uint32_t k, int nt) {
    if (k <= 2u) return nt;
    if (k >= (uint32_t)nt + 1u) return 1;
    return nt - (int)k + 2;  // 3 <= k <= nt: [2, nt - 1]
}

// Blocks per group of the lean sweep's two-level ticket, and the groups of its largest grid
constexpr int kTicketGroup = 64;
constexpr int kTicketGroups = 32;

// Words between the lean sweep's counters: one 128-byte cache line each, since atomics on
// one line serialize
constexpr int kCounterStride = 32;

// Device-side solver state of one solve (one per model).
struct VIState {
    uint32_t k;          // 1-based index of the next sweep
    int32_t conv;        // sweep at which the residual first fell below thresh, -1 if none
    uint32_t done;       // blocks (lean sweep: block groups) of the current sweep that finished (ticket)
    float last_res;      // residual of the last completed sweep
    // lean sweep, at [i * kCounterStride]: blocks of group i of kTicketGroup that finished
    uint32_t gdone[kTicketGroups * kCounterStride];
    // lean sweep, at [x * kCounterStride]: next tail piece of XCD x's pool (0 between sweeps)
    uint32_t pool[8 * kCounterStride];
};

// Column blocks (grid.y) of a batched tile sweep that take their tiles from counters
constexpr int kTileRows = 8;

// Per-column state of a batched (multi-goal) solve.
struct BatchState {
    uint32_t k;          // 1-based index of the next sweep
    uint32_t done;       // blocks finished in the current sweep
    int32_t all_conv;    // 1 once every column converged
    uint32_t pad;
    // tile sweeps, at [(y * 8 + x) * kCounterStride]: tiles of XCD x's band that column
    // block y's blocks have taken beyond their first (0 between sweeps)
    uint32_t tnext[kTileRows * 8 * kCounterStride];
};

struct BatchCache;
void batch_cache_free(BatchCache* c);

// wpart[0..W]: first state of each of W waves, balanced by entries + rows, group-aligned
void launch_wave_partition(const uint32_t* rp, int64_t ns, int na, int64_t W, int spc, uint32_t* wpart,
                           hipStream_t s);
struct Model;
// m->slice_rp: the row pointer at every slice boundary of a built model (once per model)
e2e_status fetch_slice_row_ptrs(Model* m, hipStream_t s);
// algorithmic bytes of the rows of slices [0, T): entries, row pointers, R, plus
// per_state bytes per state
double prefix_sweep_bytes(const Model* m, int64_t T, double per_state);
// wpart[t][0..W] for t < nt: the same partition of the states [0, (t+1) nc)
void launch_prefix_partitions(const uint32_t* rp, int64_t nc, int na, int64_t nt, int64_t W, int spc,
                              uint32_t* wpart, hipStream_t s);

struct Model {
    e2e_params p{};
    bool built = false;       // false: imported from COO
    int64_t ns = 0, na = 0, nc = 0, nt = 0, nnz = 0, rows = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;

    // state-major CSR, row = s*na + a
    uint32_t* row_ptr = nullptr;  // [rows+1]
    int2* nz = nullptr;           // [nnz] (col, val bits)
    float* R = nullptr;           // [rows]
    float* Rsum = nullptr;        // [rows] (optional)

    // plain solve buffers
    float* V[2] = {nullptr, nullptr};
    double* W[2] = {nullptr, nullptr};  // float64 values of the float64 sweeps
    int32_t* pi = nullptr;
    VIState* st = nullptr;
    float* vi_part = nullptr;     // per-block residual partials of the running sweep
    hipGraphExec_t graph = nullptr;
    int graph_chunk = 0;
    int graph_lpr = 0;
    float graph_thresh = 0.f, graph_discount = 0.f;
    void* graph_fn = nullptr;
    int64_t graph_blocks = 0;
    unsigned char graph_args[256] = {};  // kernel arguments the graph was captured with
    uint32_t* vi_wpart = nullptr;  // [waves+1] first state of each wave's contiguous share
    int64_t vi_wpart_waves = 0;
    int vi_wpart_spc = 0;
    uint32_t* vi_pre_wpart = nullptr;  // active-slice sweeps: [nt][waves+1] partitions of [0, (t+1) nc)
    int64_t vi_pre_waves = 0;
    int vi_pre_spc = 0;
    std::vector<uint32_t> slice_rp;    // host copy of row_ptr at each slice boundary, [nt+1]
    uint32_t* vi_bw_wpart = nullptr;  // backward pass: [nt][waves+1] per-slice partitions
    int64_t vi_bw_waves = 0;
    int vi_bw_spc = 0;
    // jagged-diagonal copy of the CSR for the jagged sweep: per state, rows sorted by
    // length (longest first); entry k of the rows longer than k stored contiguously
    int2* vi_jnz = nullptr;       // [nnz + kNzPad]
    uint32_t* vi_jmeta = nullptr; // [rows] length | (action << 16), sorted order
    float* vi_jR = nullptr;       // [rows] R in sorted order
    uint2* vi_tdesc = nullptr;    // tiled sweep: per tile (first column, column span)
    uint32_t vi_tile_states = 0, vi_maxspan = 0;
    int64_t vi_part_cap = 0;      // entries of vi_part
    // compressed stream of the default sweep: per entry count << 16 | (column - base),
    // per 256-entry window its base column (-1: that window keeps its 8-byte entries)
    uint32_t* vi_cz = nullptr;    // [vi_cwin * 256]
    int32_t* vi_cbase = nullptr;  // [vi_cwin]
    int64_t vi_cwin = 0;
    int64_t vi_cwin_comp = 0;     // windows kept at 4 bytes per entry
    int vi_comp_span = 0;         // column span limit the stream was converted with

    bool has_obstacles = false;
    BatchCache* batch = nullptr;  // buffers + tuning of the batched multi-goal solve

    // alpha columns: R of each alpha, state-major like R
    float* Ralpha = nullptr;      // [n_alpha][rows]
    int32_t n_alpha = 0;
    float alphas[64] = {};

    e2e_model_info info{};
};

// Device allocations that outlive one model (devpool.cpp): blocks of 1 MiB or more freed
// with pool_free stay mapped for later pool_malloc calls of a similar size. pool_free
// takes any device pointer (others go to hipFree); a pooled block must not reach hipFree.
hipError_t pool_malloc(void** p, size_t bytes);
void pool_free(void* p);
void pool_release_all();
size_t pool_cached_bytes();
template <typename T>
inline hipError_t pool_malloc(T** p, size_t bytes) {
    return pool_malloc(reinterpret_cast<void**>(p), bytes);
}

// Scoped device allocation for temporaries.
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) pool_free(p);
        p = nullptr;
    }
    template <typename T>
    T* as() const { return reinterpret_cast<T*>(p); }
};

inline e2e_status dmalloc(DevBuf& b, size_t bytes) {
    hipError_t e = pool_malloc(&b.p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        return E2E_ERR_NOMEM;
    }
    return E2E_OK;
}

e2e_status check_params(const e2e_params* p);  // the build's parameter checks
e2e_status model_alloc_stream(Model* m, void* stream);
// allocates m->nz with nnz entries plus a zeroed tail of kNzPad entries (on m->stream)
e2e_status alloc_nz(Model* m, int64_t nnz);
e2e_status free_model(Model* m);

// scan.hip: exclusive scan of n uint32 counts into out[0..n] (out[n] = total).
// tmp must hold scan_tmp_words(n) uint64 words. Returns total via *total_dev (uint64 on device).
size_t scan_tmp_words(int64_t n);
e2e_status scan_u32(const uint32_t* in, uint32_t* out, int64_t n, unsigned long long* tmp,
                    unsigned long long* total_dev, hipStream_t s);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

double now_ms();

}  // namespace e2e
