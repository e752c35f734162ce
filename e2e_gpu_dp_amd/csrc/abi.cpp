// abi.cpp — the extern "C" surface of libe2edp: argument checks, error state, handle
// lifetime. Compute lives in build.hip (model build/import/export) and vi.hip (solve).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdarg>
#include <cstring>
#include <exception>
#include <new>

#include "internal.h"

namespace e2e {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

double now_ms() {
    using namespace std::chrono;
    return duration_cast<duration<double, std::milli>>(steady_clock::now().time_since_epoch()).count();
}

e2e_status model_alloc_stream(Model* m, void* stream) {
    if (stream) {
        m->stream = reinterpret_cast<hipStream_t>(stream);
        m->own_stream = false;
        return E2E_OK;
    }
    E2E_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    m->own_stream = true;
    return E2E_OK;
}

e2e_status free_model(Model* m) {
    if (!m) return E2E_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    if (m->graph) (void)hipGraphExecDestroy(m->graph);
    batch_cache_free(m->batch);
    void* ptrs[] = {m->row_ptr, m->nz,      m->R,        m->Rsum,     m->V[0],   m->V[1],     m->W[0], m->W[1],
                    m->pi,      m->st,      m->vi_part,  m->vi_wpart, m->vi_tdesc, m->vi_jnz, m->vi_jmeta, m->vi_jR,
                    m->vi_cz,   m->vi_cbase, m->Ralpha, m->vi_bw_wpart, m->vi_pre_wpart};
    for (void* p : ptrs) pool_free(p);
    if (m->own_stream && m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return E2E_OK;
}

// defined in build.hip / vi.hip
e2e_status model_build(const e2e_params* p, const e2e_fields* f, const e2e_build_opts* o, Model** out);
e2e_status model_import(int64_t ns, int64_t na, int64_t nnz, const int32_t* S1, const int32_t* S2,
                        const float* val, const float* R, void* stream, Model** out);
e2e_status model_export_coo64(const Model* m, int64_t* S1, int64_t* S2, float* val, float* R);
e2e_status model_export_coo32(const Model* m, int32_t* S1, int32_t* S2, float* val, float* R);
e2e_status model_export_csr(const Model* m, uint32_t* row_ptr, int32_t* col, float* val, float* R);
e2e_status vi_solve(Model* m, const e2e_solve_opts* o, uint32_t* policy, float* value,
                    e2e_solve_stats* stats);
e2e_status vi_solve_noalloc(Model* m, const e2e_solve_opts* o, uint32_t* policy, float* value,
                            e2e_solve_stats* stats);
void actions_table(const e2e_params* p, float* speed, float* angle);
e2e_status vi_solve_goals(Model* m, const int32_t* goals, int k, const e2e_solve_opts* o,
                          const e2e_batch_opts* bo, uint32_t* policy, float* value, int32_t* col_sweeps,
                          e2e_batch_stats* stats);
e2e_status vi_solve_alphas(Model* m, const e2e_solve_opts* o, const e2e_batch_opts* bo, uint32_t* policy,
                           float* value, int32_t* col_sweeps, e2e_batch_stats* stats);
e2e_status model_export_R_alpha(const Model* m, int32_t j, float* R);
e2e_status env_generate(const e2e_env_params* ep, const e2e_env_out* out, e2e_env_stats* stats);
e2e_status rollout(const e2e_params* p, const e2e_fields* f, const int32_t* policy, int32_t n_starts,
                   const int32_t* starts, const e2e_rollout_opts* o, int8_t* success, int32_t* travel,
                   double* energy, double* xy, e2e_rollout_stats* stats);

template <typename F>
static e2e_status guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return E2E_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_error("%s", e.what());
        return E2E_ERR_INVALID;
    } catch (...) {
        set_error("unknown exception");
        return E2E_ERR_INVALID;
    }
}

static int64_t model_device_bytes(const Model* m) {
    int64_t b = sizeof(uint32_t) * (m->rows + 1) + sizeof(int2) * m->nnz + sizeof(float) * m->rows;
    if (m->Rsum) b += sizeof(float) * m->rows;
    if (m->V[0]) b += (2 * sizeof(float) + 2 * sizeof(double) + sizeof(int32_t)) * m->ns + sizeof(VIState);
    if (m->vi_jnz) b += sizeof(int2) * (m->nnz + kNzPad) + (sizeof(uint32_t) + sizeof(float)) * m->rows;
    if (m->vi_cz) b += (int64_t)(sizeof(uint32_t) * 256 + sizeof(int32_t)) * m->vi_cwin;
    if (m->Ralpha) b += (int64_t)sizeof(float) * m->n_alpha * m->rows;
    return b;
}

}  // namespace e2e

using namespace e2e;

extern "C" {

const char* e2e_last_error(void) { return g_err.c_str(); }

const char* e2e_version(void) { return "e2edp 0.1 (gfx950)"; }

int32_t e2e_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int64_t e2e_cached_device_bytes(void) { return (int64_t)pool_cached_bytes(); }

void e2e_release_cached_memory(void) { pool_release_all(); }

int32_t e2e_struct_size(int32_t which) {
    switch (which) {
        case 0: return (int32_t)sizeof(e2e_params);
        case 1: return (int32_t)sizeof(e2e_fields);
        case 2: return (int32_t)sizeof(e2e_build_opts);
        case 3: return (int32_t)sizeof(e2e_model_info);
        case 4: return (int32_t)sizeof(e2e_solve_opts);
        case 5: return (int32_t)sizeof(e2e_solve_stats);
        case 6: return (int32_t)sizeof(e2e_batch_opts);
        case 7: return (int32_t)sizeof(e2e_batch_stats);
        case 8: return (int32_t)sizeof(e2e_server_opts);
        case 9: return (int32_t)sizeof(e2e_request_result);
        case 10: return (int32_t)sizeof(e2e_server_stats);
        case 11: return (int32_t)sizeof(e2e_rollout_opts);
        case 12: return (int32_t)sizeof(e2e_rollout_stats);
        case 13: return (int32_t)sizeof(e2e_env_params);
        case 14: return (int32_t)sizeof(e2e_env_out);
        case 15: return (int32_t)sizeof(e2e_env_stats);
        default: return -1;
    }
}

e2e_status e2e_params_pack(const e2e_params* p, float out32[32]) {
    if (!p || !out32) {
        set_error("e2e_params_pack: NULL argument");
        return E2E_ERR_INVALID;
    }
    // H_params, src/build_model.cu:1032-1055
    out32[0] = (float)p->gsize;
    out32[1] = (float)(p->num_ac_speeds * p->num_ac_angles);
    out32[2] = (float)p->nrzns;
    out32[3] = p->F;
    out32[4] = p->dt;
    out32[5] = p->r_outbound;
    out32[6] = p->r_terminal;
    out32[7] = (float)p->nmodes;
    out32[8] = (float)p->i_term;
    out32[9] = (float)p->j_term;
    out32[10] = (float)p->nt;
    out32[11] = (float)p->is_stationary;
    out32[12] = (float)p->term_subgrid_size;
    out32[13] = (float)p->reward_type;
    out32[14] = (float)p->num_ac_speeds;
    out32[15] = (float)p->num_ac_angles;
    out32[16] = p->dx;
    out32[17] = p->dy;
    out32[18] = (float)p->neighb_gsize;
    out32[19] = p->alpha;
    for (int i = 20; i < 32; i++) out32[i] = p->z;
    return E2E_OK;
}

e2e_status e2e_actions(const e2e_params* p, float* speed, float* angle) {
    if (!p || p->num_ac_speeds < 1 || p->num_ac_angles < 1) {
        set_error("e2e_actions: bad params");
        return E2E_ERR_INVALID;
    }
    return guarded([&] {
        actions_table(p, speed, angle);
        return E2E_OK;
    });
}

e2e_status e2e_model_build(const e2e_params* p, const e2e_fields* f, const e2e_build_opts* opts,
                           void** model_out) {
    if (!model_out) {
        set_error("e2e_model_build: model_out is NULL");
        return E2E_ERR_INVALID;
    }
    *model_out = nullptr;
    return guarded([&] {
        Model* m = nullptr;
        e2e_status st = model_build(p, f, opts, &m);
        if (st == E2E_OK) {
            m->info.device_bytes = model_device_bytes(m);
            *model_out = m;
        }
        return st;
    });
}

e2e_status e2e_model_import_coo(int64_t ns, int64_t na, int64_t nnz, const int32_t* S1,
                                const int32_t* S2, const float* val, const float* R, void* stream,
                                void** model_out) {
    if (!model_out) {
        set_error("e2e_model_import_coo: model_out is NULL");
        return E2E_ERR_INVALID;
    }
    *model_out = nullptr;
    return guarded([&] {
        Model* m = nullptr;
        e2e_status st = model_import(ns, na, nnz, S1, S2, val, R, stream, &m);
        if (st == E2E_OK) {
            m->info.device_bytes = model_device_bytes(m);
            *model_out = m;
        }
        return st;
    });
}

e2e_status e2e_model_info_get(const void* model, e2e_model_info* out) {
    if (!model || !out) {
        set_error("e2e_model_info_get: NULL argument");
        return E2E_ERR_INVALID;
    }
    const Model* m = static_cast<const Model*>(model);
    *out = m->info;
    out->ns = m->ns;
    out->na = m->na;
    out->nnz = m->nnz;
    out->nc = m->built ? m->nc : 0;
    out->nt = m->built ? m->nt : 0;
    out->has_rsum = m->Rsum != nullptr;
    out->device_bytes = model_device_bytes(m);
    return E2E_OK;
}

e2e_status e2e_model_destroy(void* model) {
    return free_model(static_cast<Model*>(model));
}

e2e_status e2e_model_export_coo(const void* model, int64_t* S1, int64_t* S2, float* val, float* R) {
    if (!model) {
        set_error("e2e_model_export_coo: model is NULL");
        return E2E_ERR_INVALID;
    }
    return guarded([&] { return model_export_coo64(static_cast<const Model*>(model), S1, S2, val, R); });
}

e2e_status e2e_model_export_coo_i32(const void* model, int32_t* S1, int32_t* S2, float* val, float* R) {
    if (!model) {
        set_error("e2e_model_export_coo_i32: model is NULL");
        return E2E_ERR_INVALID;
    }
    return guarded([&] { return model_export_coo32(static_cast<const Model*>(model), S1, S2, val, R); });
}

e2e_status e2e_model_export_csr(const void* model, uint32_t* row_ptr, int32_t* col, float* val, float* R) {
    if (!model) {
        set_error("e2e_model_export_csr: model is NULL");
        return E2E_ERR_INVALID;
    }
    return guarded([&] { return model_export_csr(static_cast<const Model*>(model), row_ptr, col, val, R); });
}

e2e_status e2e_model_export_R_alpha(const void* model, int32_t j, float* R) {
    if (!model) {
        set_error("e2e_model_export_R_alpha: model is NULL");
        return E2E_ERR_INVALID;
    }
    return guarded([&] { return model_export_R_alpha(static_cast<const Model*>(model), j, R); });
}

e2e_status e2e_vi_solve_alphas(void* model, const e2e_solve_opts* opts, const e2e_batch_opts* bopts,
                               uint32_t* policy, float* value, int32_t* col_sweeps, e2e_batch_stats* stats) {
    if (!model) {
        set_error("e2e_vi_solve_alphas: model is NULL");
        return E2E_ERR_INVALID;
    }
    return guarded([&] {
        return vi_solve_alphas(static_cast<Model*>(model), opts, bopts, policy, value, col_sweeps, stats);
    });
}

e2e_status e2e_vi_solve(void* model, const e2e_solve_opts* opts, uint32_t* policy, float* value,
                        e2e_solve_stats* stats) {
    if (!model) {
        set_error("e2e_vi_solve: model is NULL");
        return E2E_ERR_INVALID;
    }
    return guarded([&] { return vi_solve(static_cast<Model*>(model), opts, policy, value, stats); });
}

e2e_status e2e_vi_solve_noalloc(void* model, const e2e_solve_opts* opts, uint32_t* policy, float* value,
                                e2e_solve_stats* stats) {
    if (!model) {
        set_error("e2e_vi_solve_noalloc: model is NULL");
        return E2E_ERR_INVALID;
    }
    return guarded([&] { return vi_solve_noalloc(static_cast<Model*>(model), opts, policy, value, stats); });
}

e2e_status e2e_vi_solve_goals(void* model, const int32_t* goal_cells, int32_t k,
                              const e2e_solve_opts* opts, const e2e_batch_opts* bopts, uint32_t* policy,
                              float* value, int32_t* col_sweeps, e2e_batch_stats* stats) {
    if (!model) {
        set_error("e2e_vi_solve_goals: model is NULL");
        return E2E_ERR_INVALID;
    }
    return guarded([&] {
        return vi_solve_goals(static_cast<Model*>(model), goal_cells, k, opts, bopts, policy, value, col_sweeps,
                              stats);
    });
}

e2e_status e2e_rollout(const e2e_params* p, const e2e_fields* fields, const int32_t* policy, int32_t n_starts,
                       const int32_t* starts, const e2e_rollout_opts* opts, int8_t* success, int32_t* travel,
                       double* energy, double* xy, e2e_rollout_stats* stats) {
    return guarded([&] { return rollout(p, fields, policy, n_starts, starts, opts, success, travel, energy, xy, stats); });
}

e2e_status e2e_env_generate(const e2e_env_params* p, const e2e_env_out* out, e2e_env_stats* stats) {
    return guarded([&] { return env_generate(p, out, stats); });
}

}  // extern "C"
