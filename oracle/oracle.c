/*
 * oracle.c — CPU restatement of the reference planner (checker + CPU baseline).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py. The product never links or calls it.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -shared -fPIC (oracle/Makefile). With
 * SSE2 float arithmetic (FLT_EVAL_METHOD 0) and no contraction, each float operation
 * below rounds to binary32 exactly where the reference's CUDA source rounds; the
 * double sub-expressions of the reference (0.5*Dj, 0.05*r_outbound, 1.414213) are kept
 * double. Trig is taken once per action on the host (glibc cosf/sinf) and enters the
 * lane arithmetic as the float products F*cosf(angle), F*sinf(angle).
 *
 * Line references are to src/build_model.cu unless noted.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_mut = 0;
void orc_set_mutation(int code) { g_mut = code; }

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------- setup */

/* define_xs_or_ys :1395-1399 */
void orc_grid(const e2e_params* p, float* xs, float* ys) {
    for (int i = 0; i < p->gsize; i++) {
        xs[i] = p->x0 + i * p->dx;
        ys[i] = p->y0 + i * p->dy;
    }
}

/* populate_ac_angles :1403-1408, populate_ac_speeds :1412-1434,
 * populate_actions :1438-1457; the trig of move() :284-285 taken on the host. */
int orc_actions(const e2e_params* p, float* speed, float* angle, float* fcos, float* fsin) {
    int ns = p->num_ac_speeds, na = p->num_ac_angles;
    if (ns < 1 || na < 1) return -1;
    float* sp = (float*)malloc(sizeof(float) * ns);
    float* an = (float*)malloc(sizeof(float) * na);
    for (int i = 0; i < na; i++) an[i] = i * (2 * M_PI) / na;
    if (ns == 1) {
        sp[0] = p->F;
    } else {
        float delF = p->F / (ns);
        for (int i = 0; i < ns; i++) sp[i] = (i + 1) * delF;
    }
    for (int i = 0; i < ns; i++)
        for (int j = 0; j < na; j++) {
            int idx = j + na * i;
            if (speed) speed[idx] = sp[i];
            if (angle) angle[idx] = an[j];
            if (fcos) fcos[idx] = sp[i] * cosf(an[j]);
            if (fsin) fsin[idx] = sp[i] * sinf(an[j]);
        }
    free(sp);
    free(an);
    return 0;
}

/* ---------------------------------------------------------------- lane */

typedef struct lane_ctx {
    const e2e_params* p;
    const e2e_fields* f;
    const float* xs;
    const float* ys;
    int g, nc, n, nm, nr, nt;
} lane_ctx;

/* is_terminal :181-190 */
static int is_terminal(const e2e_params* p, int32_t i, int32_t j) {
    int32_t it = p->i_term, jt = p->j_term, ts = p->term_subgrid_size;
    return (i >= it && i < it + ts) && (j >= jt && j < jt + ts);
}

/* is_edge_state :162-169 (n = gridDim.x - 1 = gsize - 1) */
static int is_edge_state(int32_t i, int32_t j, int32_t n) {
    return i == 0 || i == n || j == 0 || j == n;
}

/* is_in_obstacle :172-178 */
static int is_in_obstacle(const lane_ctx* c, long long sp, int T) {
    if (!c->f->mask) return 0;
    return c->f->mask[(long long)T * c->nc + sp] == 1;
}

/* is_within_band :419-441 */
static int is_within_band(const lane_ctx* c, int i, int j, int i1, int j1, int i2, int j2) {
    if (i1 == i2 || j1 == j2) return 1;
    const float* xs = c->xs;
    const float* ys = c->ys;
    int g = c->g;
    float cell_diag = fabsf(xs[1] - xs[0]) * 1.414213;
    float x = xs[j], y = (g_mut == 11) ? ys[i] : ys[g - 1 - i];
    float x1 = xs[j1], y1 = ys[g - 1 - i1];
    float x2 = xs[j2], y2 = ys[g - 1 - i2];
    float A = (y2 - y1) / (x2 - x1);
    float B = -1;
    float C = y1 - (A * x1);
    float dist = fabsf(A * x + B * y + C) / sqrtf((A * A) + (B * B));
    return dist < cell_diag;
}

/* goes_through_obstacle :444-477 */
static int goes_through_obstacle(const lane_ctx* c, long long sp1, long long sp2, int T) {
    if (!c->f->mask) return 0;
    int g = c->g;
    int p1i = (int)(sp1 / g), p1j = (int)(sp1 % g);
    int p2i = (int)(sp2 / g), p2j = (int)(sp2 % g);
    int imin = p1i < p2i ? p1i : p2i, imax = p1i > p2i ? p1i : p2i;
    int jmin = p1j < p2j ? p1j : p2j, jmax = p1j > p2j ? p1j : p2j;
    for (int i = imin; i <= imax; i++)
        for (int j = jmin; j <= jmax; j++) {
            if (g_mut == 13 || !(i == p1i && j == p1j)) {
                long long sp = j + (long long)g * i;
                if (is_in_obstacle(c, sp, T) || (g_mut != 12 && is_in_obstacle(c, sp, T + 1)))
                    if (is_within_band(c, i, j, p1i, p1j, p2i, p2j)) return 1;
            }
        }
    return 0;
}

/* calculate_one_step_reward :231-265 */
static float step_reward(const e2e_params* p, float ac_speed, float rad1, float rad2) {
    int method = p->reward_type;
    float alpha = p->alpha;
    float Cr = 1;
    float Cf = 1;
    float Ct = 0.01;
    float dt = p->dt;
    float energy_reward, time_reward;
    if (method == 0) return -dt;
    if (method == 1) return -(Cf * ac_speed * ac_speed + Ct) * dt;
    if (method == 2) {
        if (g_mut == 9) return ((Cr * (rad2 + rad1)) - (Cf * ac_speed * ac_speed) - Ct) * dt;
        return ((Cr * (rad2 + rad1) / 2) - (Cf * ac_speed * ac_speed) - Ct) * dt;
    }
    if (method == 3) {
        if (g_mut == 15) return ((Cr * (rad1 + rad1) / 2) - Ct) * dt;
        return ((Cr * (rad2 + rad1) / 2) - Ct) * dt;
    }
    if (method == 4) {
        energy_reward = -(ac_speed * ac_speed) * dt;
        time_reward = -dt;
        if (g_mut == 10) return (1 - alpha) * energy_reward + alpha * time_reward;
        return alpha * energy_reward + (1 - alpha) * time_reward;
    }
    return 0;
}

/* move :268-356. fcos/fsin are F*cosf(ac_angle), F*sinf(ac_angle). */
static void move_ref(const lane_ctx* c, float fcos, float fsin, float vx, float vy, int32_t T,
                     int32_t* posids, float* r) {
    const e2e_params* p = c->p;
    const float* xs = c->xs;
    const float* ys = c->ys;
    int32_t gsize = c->g;
    int32_t n = gsize - 1;
    int32_t nt = p->nt;
    float dt = p->dt;
    float r_outbound = p->r_outbound;
    float r_terminal = p->r_terminal;
    float Dj = fabsf(xs[1] - xs[0]);
    float Di = fabsf(ys[1] - ys[0]);
    int32_t i0 = posids[0];
    int32_t j0 = posids[1];
    float vnetx = fcos + vx;
    float vnety = fsin + vy;
    float x = xs[j0];
    float y = ys[gsize - 1 - i0];
    float xnew = x + (vnetx * dt);
    float ynew = y + (vnety * dt);
    *r = 0;
    if (xnew > xs[n]) {
        xnew = xs[n];
        *r += r_outbound;
    } else if (xnew < xs[0]) {
        xnew = xs[0];
        *r += r_outbound;
    }
    if (ynew > ys[n]) {
        ynew = ys[n];
        *r += r_outbound;
    } else if (ynew < ys[0]) {
        ynew = ys[0];
        *r += r_outbound;
    }
    int32_t xind, yind;
    float remx = remainderf((xnew - xs[0]), Dj);
    float remy = (g_mut == 2) ? remainderf((ynew - ys[n]), Di) : remainderf(-(ynew - ys[n]), Di);
    xind = ((xnew - xs[0]) - remx) / Dj;
    yind = (-(ynew - ys[n]) - remy) / Di;
    if ((remx >= 0.5 * Dj) && (remy >= 0.5 * Di)) {
        xind += 1;
        yind += 1;
    } else if ((remx >= 0.5 * Dj && remy < 0.5 * Di)) {
        xind += 1;
    } else if ((remx < 0.5 * Dj && remy >= 0.5 * Di)) {
        yind += 1;
    }
    /* my_isnan(int) is always false (:193-196, :331): posids are always updated. */
    if (g_mut == 3) {
        posids[0] = xind;
        posids[1] = yind;
    } else {
        posids[0] = yind;
        posids[1] = xind;
    }
    if (g_mut != 4 && is_edge_state(posids[0], posids[1], n)) *r += r_outbound;
    if (is_terminal(p, posids[0], posids[1])) {
        *r += r_terminal;
    } else {
        if ((g_mut == 5) ? (T >= nt - 20) : (T > nt - 20)) *r += 0.05 * r_outbound;
    }
}

static int lane_eval(const lane_ctx* c, float speed, float fcos, float fsin, int t, int sp,
                     int rzn, int32_t* sp2_out, float* r_out) {
    const e2e_params* p = c->p;
    const e2e_fields* f = c->f;
    int g = c->g;
    long long nc = c->nc;
    int nm = c->nm;
    int32_t posids[2] = {sp / g, sp % g};
    /* extract_velocity :361-404 — the mode sum starts at 0, accumulates in k order,
     * and is added to the mean afterwards. */
    long long mean_id = (long long)t * nc + sp;
    long long str_uvi = nc;
    long long sp_uvi = ((long long)t * nm * str_uvi) + ((long long)g * posids[0]) + posids[1];
    long long sp_Yi = ((long long)t * c->nr * nm) + ((long long)rzn * nm);
    float sum_x = 0, sum_y = 0;
    int kmax = (g_mut == 1) ? nm - 1 : nm;
    float vx_mean = f->u[mean_id];
    for (int k = 0; k < kmax; k++) sum_x += f->ui[sp_uvi + k * str_uvi] * f->yi[sp_Yi + k];
    float vy_mean = f->v[mean_id];
    for (int k = 0; k < kmax; k++) sum_y += f->vi[sp_uvi + k * str_uvi] * f->yi[sp_Yi + k];
    float vx = vx_mean + sum_x;
    float vy = vy_mean + sum_y;
    if (!isfinite(vx) || !isfinite(vy)) return -1;
    /* extract_radiation :407-416 */
    float rad1 = f->s ? f->s[mean_id] : 0.0f;
    float r = 0;
    /* transition_calc :522-545 */
    if (!is_terminal(p, posids[0], posids[1])) {
        if (!is_in_obstacle(c, sp, t)) {
            move_ref(c, fcos, fsin, vx, vy, t, posids, &r);
            long long sp2 = posids[1] + (long long)g * posids[0];
            float rad2 = f->s ? f->s[(long long)(g_mut == 14 ? t : t + 1) * nc + sp2] : 0.0f;
            float r_step = step_reward(p, speed, rad1, rad2);
            r += r_step;
            if (goes_through_obstacle(c, sp, sp2, t)) r = p->r_outbound;
        } else {
            r = p->r_outbound;
        }
    }
    *sp2_out = posids[1] + g * posids[0];
    *r_out = r;
    return 0;
}

static void ctx_init(lane_ctx* c, const e2e_params* p, const e2e_fields* f, float* xs, float* ys) {
    c->p = p;
    c->f = f;
    c->g = p->gsize;
    c->nc = p->gsize * p->gsize;
    c->n = p->gsize - 1;
    c->nm = p->nmodes;
    c->nr = p->nrzns;
    c->nt = p->nt;
    orc_grid(p, xs, ys);
    c->xs = xs;
    c->ys = ys;
}

int orc_lane(const e2e_params* p, const e2e_fields* f, int t, int a, int sp, int rzn,
             int32_t* sp2, float* r) {
    int na = p->num_ac_speeds * p->num_ac_angles;
    float* xs = (float*)malloc(sizeof(float) * p->gsize * 2);
    float* sp_ = (float*)malloc(sizeof(float) * na * 3);
    lane_ctx c;
    ctx_init(&c, p, f, xs, xs + p->gsize);
    orc_actions(p, sp_, NULL, sp_ + na, sp_ + 2 * na);
    int rc = lane_eval(&c, sp_[a], sp_[na + a], sp_[2 * na + a], t, sp, rzn, sp2, r);
    free(xs);
    free(sp_);
    return rc;
}

int orc_lanes(const e2e_params* p, const e2e_fields* f, int t, int a, int32_t* sp2, float* r) {
    int na = p->num_ac_speeds * p->num_ac_angles;
    float* xs = (float*)malloc(sizeof(float) * p->gsize * 2);
    float* sp_ = (float*)malloc(sizeof(float) * na * 3);
    lane_ctx c;
    ctx_init(&c, p, f, xs, xs + p->gsize);
    orc_actions(p, sp_, NULL, sp_ + na, sp_ + 2 * na);
    int rc = 0;
    for (int sp = 0; sp < c.nc; sp++)
        for (int w = 0; w < c.nr; w++) {
            long long i = (long long)sp * c.nr + w;
            if (lane_eval(&c, sp_[a], sp_[na + a], sp_[2 * na + a], t, sp, w, &sp2[i], &r[i]))
                rc = -1;
        }
    free(xs);
    free(sp_);
    return rc;
}

/* ---------------------------------------------------------------- build */

/* One (t, a) block of nc segments. */
typedef struct blk {
    int32_t* nuq;    /* [nc] */
    int32_t* key;    /* unique sp2, ascending, concatenated */
    int32_t* cnt;
    float* R;        /* [nc] */
    float* Rsum;     /* [nc] */
    long long n;     /* total uniques */
    long long cap;
    long long oow;   /* lanes outside the neighb_gsize window */
    int max_nuq;
    int bad;
} blk;

static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

static void blk_push(blk* b, int32_t k, int32_t n) {
    if (b->n == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 1024;
        b->key = (int32_t*)realloc(b->key, sizeof(int32_t) * b->cap);
        b->cnt = (int32_t*)realloc(b->cnt, sizeof(int32_t) * b->cap);
    }
    b->key[b->n] = k;
    b->cnt[b->n] = n;
    b->n++;
}

/* The reward sum of a segment. The reference adds the per-lane rewards with float
 * atomicAdd in an unspecified order (:553). The canonical order fixed here (and used by
 * the GPU kernel), in double: realizations are taken in chunks of 64 consecutive w;
 * within a chunk, p[l] = r[64c + l] (0 past nr) is reduced by the tree
 * p[l] += p[l + o] for o = 32, 16, ..., 1 (result p[0]); the chunk sums are then added
 * in ascending chunk order starting from 0. For integer-valued rewards every order
 * gives the same exact sum. */
static double canonical_sum(const float* r, int nr) {
    double total = 0.0;
    for (int c = 0; c * 64 < nr; c++) {
        double p[64];
        for (int l = 0; l < 64; l++) p[l] = (c * 64 + l < nr) ? (double)r[c * 64 + l] : 0.0;
        for (int o = 32; o >= 1; o >>= 1)
            for (int l = 0; l < o; l++) p[l] = p[l] + p[l + o];
        total = total + p[0];
    }
    return total;
}

/* Count-by-key of one segment (prel_tests/build_model.cu:410-487 semantics: sort the
 * S2 keys of the segment, one output per distinct key, count = run length). */
static void segment_count(blk* b, int32_t* keys, int nr) {
    /* counting sort when the key range is small (the usual case: S2 lies near S1),
     * comparison sort otherwise; both give the ascending run-length counts */
    int32_t kmin = keys[0], kmax = keys[0];
    for (int w = 1; w < nr; w++) {
        if (keys[w] < kmin) kmin = keys[w];
        if (keys[w] > kmax) kmax = keys[w];
    }
    if ((int64_t)kmax - kmin < 65536) {
        int range = kmax - kmin + 1;
        int32_t* h = (int32_t*)calloc((size_t)range, sizeof(int32_t));
        for (int w = 0; w < nr; w++) h[keys[w] - kmin]++;
        int nu = 0;
        for (int k = 0; k < range; k++)
            if (h[k]) {
                int n = h[k];
                if (g_mut == 6 && nu == 0) n += 1;
                blk_push(b, kmin + k, n);
                nu++;
            }
        free(h);
        if (nu > b->max_nuq) b->max_nuq = nu;
        return;
    }
    qsort(keys, nr, sizeof(int32_t), cmp_i32);
    int start = 0;
    int nu = 0;
    for (int w = 1; w <= nr; w++) {
        if (w == nr || keys[w] != keys[start]) {
            int n = w - start;
            if (g_mut == 6 && nu == 0) n += 1;
            blk_push(b, keys[start], n);
            nu++;
            start = w;
        }
    }
    if (nu > b->max_nuq) b->max_nuq = nu;
}

struct orc_model {
    int64_t ns, na, nnz, nc, nt, t0, t1;
    int64_t oow;
    int64_t max_nuq;
    int64_t* S1;
    int64_t* S2;
    float* val;
    float* R;
    float* Rsum;
};

orc_model* orc_build(const e2e_params* p, const e2e_fields* f, int t_begin, int t_end,
                     int nthreads) {
    int g = p->gsize, nt = p->nt, nr = p->nrzns;
    int na = p->num_ac_speeds * p->num_ac_angles;
    long long nc = (long long)g * g;
    long long ns = nc * nt;
    if (g < 2 || nt < 1 || nr < 1 || na < 1 || t_begin < 0 || t_end > nt - 1 || t_begin > t_end)
        return NULL;
    int nslices = t_end - t_begin;
    int nblk = nslices * na;
    float* xs = (float*)malloc(sizeof(float) * g * 2);
    float* act = (float*)malloc(sizeof(float) * na * 3);
    lane_ctx c;
    ctx_init(&c, p, f, xs, xs + g);
    orc_actions(p, act, NULL, act + na, act + 2 * na);
    int m = p->neighb_gsize;
    int del = (m - 1) / 2;
    blk* blks = (blk*)calloc(nblk > 0 ? nblk : 1, sizeof(blk));
    int bad = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(| : bad)
#endif
    for (int bi = 0; bi < nblk; bi++) {
        int t = t_begin + bi / na;
        int a = bi % na;
        blk* b = &blks[bi];
        b->nuq = (int32_t*)malloc(sizeof(int32_t) * nc);
        b->R = (float*)malloc(sizeof(float) * nc);
        b->Rsum = (float*)malloc(sizeof(float) * nc);
        int32_t* keys = (int32_t*)malloc(sizeof(int32_t) * nr);
        float* rw = (float*)malloc(sizeof(float) * nr);
        for (long long sp = 0; sp < nc; sp++) {
            long long before = b->n;
            int i1 = (int)(sp / g), j1 = (int)(sp % g);
            for (int w = 0; w < nr; w++) {
                if (lane_eval(&c, act[a], act[na + a], act[2 * na + a], t, (int)sp, w, &keys[w], &rw[w]))
                    b->bad = 1;
                /* get_rel_sp_id2 :126-141 window test, both signs checked */
                int di = keys[w] / g - (i1 - del), dj = keys[w] % g - (j1 - del);
                if (di < 0 || di >= m || dj < 0 || dj >= m) b->oow++;
            }
            double s = canonical_sum(rw, nr);
            segment_count(b, keys, nr);
            b->nuq[sp] = (int32_t)(b->n - before);
            /* compute_mean :563-570 */
            b->Rsum[sp] = (float)s;
            b->R[sp] = b->Rsum[sp] / nr;
        }
        free(keys);
        free(rw);
        bad |= b->bad;
    }
    orc_model* M = (orc_model*)calloc(1, sizeof(orc_model));
    M->ns = ns;
    M->na = na;
    M->nc = nc;
    M->nt = nt;
    M->t0 = t_begin;
    M->t1 = t_end;
    long long nnz = 0;
    for (int bi = 0; bi < nblk; bi++) {
        nnz += blks[bi].n;
        M->oow += blks[bi].oow;
        if (blks[bi].max_nuq > M->max_nuq) M->max_nuq = blks[bi].max_nuq;
    }
    M->nnz = nnz;
    M->S1 = (int64_t*)malloc(sizeof(int64_t) * (nnz ? nnz : 1));
    M->S2 = (int64_t*)malloc(sizeof(int64_t) * (nnz ? nnz : 1));
    M->val = (float*)malloc(sizeof(float) * (nnz ? nnz : 1));
    M->R = (float*)calloc((size_t)(na * ns), sizeof(float));
    M->Rsum = (float*)calloc((size_t)(na * ns), sizeof(float));
    /* Concatenation of save_master_Coos_to_file :1246-1299: action-major, then the
     * time order of insertion (:857-860), S1 += a*Ns (:1272). */
    long long o = 0;
    for (int a = 0; a < na; a++)
        for (int t = t_begin; t < t_end; t++) {
            blk* b = &blks[(t - t_begin) * na + a];
            long long k = 0;
            for (long long sp = 0; sp < nc; sp++) {
                long long s1 = sp + (long long)t * nc;
                for (int u = 0; u < b->nuq[sp]; u++, k++, o++) {
                    M->S1[o] = s1 + (long long)a * ns;
                    M->S2[o] = (long long)(t + 1) * nc + b->key[k];
                    M->val[o] = (float)b->cnt[k] / nr;
                }
                M->R[(long long)a * ns + s1] = b->R[sp];
                M->Rsum[(long long)a * ns + s1] = b->Rsum[sp];
            }
        }
    for (int bi = 0; bi < nblk; bi++) {
        free(blks[bi].nuq);
        free(blks[bi].key);
        free(blks[bi].cnt);
        free(blks[bi].R);
        free(blks[bi].Rsum);
    }
    free(blks);
    free(xs);
    free(act);
    if (bad) {
        orc_model_free(M);
        return NULL;
    }
    return M;
}

int64_t orc_model_nnz(const orc_model* m) { return m->nnz; }
int64_t orc_model_out_of_window(const orc_model* m) { return m->oow; }
int64_t orc_model_max_row_nnz(const orc_model* m) { return m->max_nuq; }

void orc_model_copy(const orc_model* m, int64_t* S1, int64_t* S2, float* val, float* R,
                    float* Rsum) {
    if (S1) memcpy(S1, m->S1, sizeof(int64_t) * m->nnz);
    if (S2) memcpy(S2, m->S2, sizeof(int64_t) * m->nnz);
    if (val) memcpy(val, m->val, sizeof(float) * m->nnz);
    if (R) memcpy(R, m->R, sizeof(float) * m->na * m->ns);
    if (Rsum) memcpy(Rsum, m->Rsum, sizeof(float) * m->na * m->ns);
}

void orc_model_free(orc_model* m) {
    if (!m) return;
    free(m->S1);
    free(m->S2);
    free(m->val);
    free(m->R);
    free(m->Rsum);
    free(m);
}

/* ---------------------------------------------------------------- VI */

/* Row pointers of a COO matrix whose row index S1 is sorted ascending. */
static int64_t* coo_row_ptr(int64_t nrows, int64_t nnz, const int64_t* S1) {
    int64_t* rp = (int64_t*)calloc((size_t)(nrows + 1), sizeof(int64_t));
    for (int64_t j = 0; j < nnz; j++) {
        if (S1[j] < 0 || S1[j] >= nrows || (j > 0 && S1[j] < S1[j - 1])) {
            free(rp);
            return NULL;
        }
        rp[S1[j] + 1]++;
    }
    for (int64_t r = 0; r < nrows; r++) rp[r + 1] += rp[r];
    return rp;
}

/* select_best_action src/spvi_2.cu:93-122: strict '>' from (-1e6, -1). */
#define ARGMAX_STEP(q, a, best, bi)                                   \
    do {                                                              \
        if ((g_mut == 7) ? ((q) >= (best)) : ((q) > (best))) {        \
            best = (q);                                               \
            bi = (a);                                                 \
        }                                                             \
    } while (0)

int orc_vi_backward_f64(int64_t ns, int64_t na, int64_t nc, int64_t nnz, const int64_t* S1,
                        const int64_t* S2, const float* val, const float* R, double discount,
                        double* V, int32_t* pi) {
    if (nc <= 0 || ns % nc) return -1;
    int64_t nt = ns / nc;
    int64_t* rp = coo_row_ptr(na * ns, nnz, S1);
    if (!rp) return -1;
    int rc = 0;
    for (int64_t t = nt - 1; t >= 0 && rc == 0; t--) {
        for (int64_t sp = 0; sp < nc; sp++) {
            int64_t s = t * nc + sp;
            double best = -1e6;
            int32_t bi = -1;
            for (int64_t a = 0; a < na; a++) {
                int64_t row = a * ns + s;
                double acc = 0.0;
                for (int64_t j = rp[row]; j < rp[row + 1]; j++) {
                    int64_t c2 = S2[j];
                    if (c2 < (t + 1) * nc || c2 >= (t + 2) * nc) rc = -1;
                    else acc += (double)val[j] * V[c2];
                }
                double q = ((g_mut == 8) ? 0.0 : (double)R[row]) + discount * acc;
                ARGMAX_STEP(q, (int32_t)a, best, bi);
            }
            V[s] = best;
            pi[s] = bi;
        }
    }
    free(rp);
    return rc;
}

static void backup_f64(int64_t ns, int64_t na, const int64_t* rp, const int64_t* S2,
                       const float* val, const float* R, double discount, const double* Vin,
                       double* Vout, int32_t* pi) {
    for (int64_t s = 0; s < ns; s++) {
        double best = -1e6;
        int32_t bi = -1;
        for (int64_t a = 0; a < na; a++) {
            int64_t row = a * ns + s;
            double acc = 0.0;
            for (int64_t j = rp[row]; j < rp[row + 1]; j++) acc += (double)val[j] * Vin[S2[j]];
            double q = ((g_mut == 8) ? 0.0 : (double)R[row]) + discount * acc;
            ARGMAX_STEP(q, (int32_t)a, best, bi);
        }
        Vout[s] = best;
        if (pi) pi[s] = bi;
    }
}

int orc_vi_sweeps_f64(int64_t ns, int64_t na, int64_t nnz, const int64_t* S1, const int64_t* S2,
                      const float* val, const float* R, double discount, double thresh,
                      int max_sweeps, double* V, int32_t* pi, double* last_residual) {
    int64_t* rp = coo_row_ptr(na * ns, nnz, S1);
    if (!rp) return -1;
    double* PV = (double*)calloc((size_t)ns, sizeof(double));
    double* CV = (double*)malloc(sizeof(double) * ns);
    int it = 0;
    double res = 0;
    while (it < max_sweeps) {
        it++;
        backup_f64(ns, na, rp, S2, val, R, discount, PV, CV, pi);
        res = 0;
        for (int64_t s = 0; s < ns; s++) {
            double d = fabs(CV[s] - PV[s]);
            if (d > res) res = d;
        }
        memcpy(PV, CV, sizeof(double) * ns);
        if (res < thresh) break;
    }
    memcpy(V, CV, sizeof(double) * ns);
    if (last_residual) *last_residual = res;
    free(rp);
    free(PV);
    free(CV);
    return it;
}

int orc_vi_sweeps_f32(int64_t ns, int64_t na, int64_t nnz, const int64_t* S1, const int64_t* S2,
                      const float* val, const float* R, float discount, float thresh,
                      int max_sweeps, int nthreads, float* V, int32_t* pi, float* last_residual) {
    int64_t* rp = coo_row_ptr(na * ns, nnz, S1);
    if (!rp) return -1;
    float* PV = (float*)calloc((size_t)ns, sizeof(float));
    float* CV = (float*)malloc(sizeof(float) * ns);
    int it = 0;
    float res = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
    while (it < max_sweeps) {
        it++;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
        for (int64_t s = 0; s < ns; s++) {
            float best = -1e6f;
            int32_t bi = -1;
            for (int64_t a = 0; a < na; a++) {
                int64_t row = a * ns + s;
                float acc = 0.0f;
                for (int64_t j = rp[row]; j < rp[row + 1]; j++) acc += val[j] * PV[S2[j]];
                float q = discount * acc + R[row];
                if (q > best) {
                    best = q;
                    bi = (int32_t)a;
                }
            }
            CV[s] = best;
            pi[s] = bi;
        }
        res = 0;
        for (int64_t s = 0; s < ns; s++) {
            float d = fabsf(CV[s] - PV[s]);
            if (d > res) res = d;
        }
        memcpy(PV, CV, sizeof(float) * ns);
        if (res < thresh) break;
    }
    memcpy(V, CV, sizeof(float) * ns);
    if (last_residual) *last_residual = res;
    free(rp);
    free(PV);
    free(CV);
    return it;
}

double orc_bellman_residual_f64(int64_t ns, int64_t na, int64_t nnz, const int64_t* S1,
                                const int64_t* S2, const float* val, const float* R,
                                double discount, const float* V) {
    int64_t* rp = coo_row_ptr(na * ns, nnz, S1);
    if (!rp) return -1.0;
    double* Vin = (double*)malloc(sizeof(double) * ns);
    double* Vout = (double*)malloc(sizeof(double) * ns);
    for (int64_t s = 0; s < ns; s++) Vin[s] = V[s];
    backup_f64(ns, na, rp, S2, val, R, discount, Vin, Vout, NULL);
    double res = 0;
    for (int64_t s = 0; s < ns; s++) {
        double d = fabs(Vout[s] - Vin[s]);
        if (d > res) res = d;
    }
    free(rp);
    free(Vin);
    free(Vout);
    return res;
}

/* ---------------------------------------------------------------- rollouts */
/* The post-processing's trajectory simulation, Post_processing/plot_e2eGPU_policy.py
 * and Post_processing/grid_world.py. Python float arithmetic is float64; the float32
 * inputs (fields, speeds, grid ticks) are widened where they enter it. */

/* Python's a % b and a // b for floats (b > 0): the sign of the remainder follows
 * the divisor, and the quotient is the floor of (a - a%b)/b rounded to an integer. */
static double py_mod(double a, double b) {
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0) != (m < 0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

static double py_floordiv(double a, double b) {
    double m = fmod(a, b);
    double d = (a - m) / b;
    if (m != 0.0 && ((b < 0) != (m < 0))) d -= 1.0;
    if (d != 0.0) {
        double fd = floor(d);
        if (d - fd > 0.5) fd += 1.0;
        return fd;
    }
    return copysign(0.0, a / b);
}

/* round(x, 14) (grid_world.py:352): the double nearest the decimal string of x
 * correctly rounded to 14 places. */
static double py_round14(double x) {
    char buf[64];
    snprintf(buf, sizeof buf, "%.14f", x);
    return strtod(buf, NULL);
}

void orc_rollout_actions(const e2e_params* p, double* speed, double* angle, double* ax, double* ay) {
    int ns = p->num_ac_speeds, na = p->num_ac_angles;
    for (int i = 0; i < ns; i++) {
        /* populate_ac_speeds :322-337 (float32: Fmax is the float32 prob_params entry) */
        float sp = ns == 1 ? p->F : (float)(i + 1) * (p->F / (float)ns);
        for (int j = 0; j < na; j++) {
            double an = py_round14((double)(j * 2) * M_PI / (double)na); /* :351-352 */
            int idx = j + na * i; /* itertools.product(speed_list, angle_list) :356 */
            if (speed) speed[idx] = (double)sp;
            if (angle) angle[idx] = an;
            if (ax) ax[idx] = (double)sp * cos(an);
            if (ay) ay[idx] = (double)sp * sin(an);
        }
    }
}

typedef struct ro_ctx {
    const e2e_params* p;
    const e2e_fields* f;
    const int32_t* policy;
    int g, nc, nm, nr, nt, na;
    double xs[4096], ys[4096];
    double dj, di;
    const double *speed, *ax, *ay;
} ro_ctx;

/* is_terminal() :140-157 without a query: the terminal subgrid */
static int ro_terminal(const e2e_params* p, int i, int j) {
    int it = p->i_term, jt = p->j_term, ts = p->term_subgrid_size;
    return i >= it && i < it + ts && j >= jt && j < jt + ts;
}

/* g.actions[policy_1d[s1_id]] (plot_e2eGPU_policy.py:236-239, s1_id :27-29); a policy
 * of -1 indexes the last action, as a Python list does */
static int ro_action(const ro_ctx* c, int t, int i, int j) {
    int a = c->policy[(long long)t * c->nc + (long long)i * c->g + j];
    return a < 0 ? a + c->na : a;
}

/* extract_velocity in the build's float32 order (the mode sum, then the mean) */
static void ro_velocity(const ro_ctx* c, int t, int i, int j, int rzn, double* vx, double* vy) {
    const e2e_fields* f = c->f;
    long long nc = c->nc, mean_id = (long long)t * nc + (long long)i * c->g + j;
    long long uvi = (long long)t * c->nm * nc + (long long)i * c->g + j;
    long long yi = ((long long)t * c->nr + rzn) * c->nm;
    float sx = 0, sy = 0;
    for (int k = 0; k < c->nm; k++) sx += f->ui[uvi + k * nc] * f->yi[yi + k];
    for (int k = 0; k < c->nm; k++) sy += f->vi[uvi + k * nc] * f->yi[yi + k];
    *vx = (double)(f->u[mean_id] + sx);
    *vy = (double)(f->v[mean_id] + sy);
}

/* get_interpolated_action :174-198 + interpolate_action_angles :113-171 +
 * interpolate_as_per_term_status :84-110. Returns -1 for the support patterns the
 * reference rejects (:158-161). */
static int ro_interp(const ro_ctx* c, int t, int i, int j, double x, double y, double* v_out,
                     double* th_out) {
    const e2e_params* p = c->p;
    int g = c->g;
    double x0 = c->xs[j], y0 = c->ys[g - 1 - i];
    double delx = x - x0, dely = y - y0;
    int fx = delx < 0 ? -1 : 1, fy = dely < 0 ? -1 : 1; /* f() :31-35 */
    int si[4] = {i, i - fy, i, i - fy}, sj[4] = {j, j, j + fx, j + fx};
    int term[4];
    double vx[4], vy[4];
    for (int k = 0; k < 4; k++) {
        term[k] = si[k] == p->i_term && sj[k] == p->j_term; /* is_terminal(query_s) :154-156 */
        int a = ro_action(c, t, si[k], sj[k]);
        vx[k] = c->ax[a];
        vy[k] = c->ay[a];
    }
    double x1 = c->xs[sj[0]], y1 = c->ys[g - 1 - si[0]];
    double x2 = c->xs[sj[3]], y2 = c->ys[g - 1 - si[3]];
    int tri = 0;
    if (!term[0] && !term[1] && !term[2] && !term[3]) {
    } else if (!term[0] && !term[1] && term[2]) { /* [0,0,1,1], [0,0,1,0] */
        vx[2] = vx[0]; vx[3] = vx[1];
        vy[2] = vy[0]; vy[3] = vy[1];
    } else if (!term[0] && term[1] && !term[2]) { /* [0,1,0,1], [0,1,0,0] */
        vx[1] = vx[0]; vx[3] = vx[2];
        vy[1] = vy[0]; vy[3] = vy[2];
    } else if (!term[0] && !term[1] && !term[2] && term[3]) { /* [0,0,0,1] */
        tri = 1;
    } else {
        return -1;
    }
    double ivx, ivy;
    if (!tri) {
        /* bilinear_interpolate :44-63: denom * x_vec @ (val_mat @ y_vec) */
        double denom = 1.0 / ((x2 - x1) * (y2 - y1));
        double r0x = vx[0] * (y2 - y) + vx[1] * (y - y1), r1x = vx[2] * (y2 - y) + vx[3] * (y - y1);
        double r0y = vy[0] * (y2 - y) + vy[1] * (y - y1), r1y = vy[2] * (y2 - y) + vy[3] * (y - y1);
        ivx = denom * ((x2 - x) * r0x + (x - x1) * r1x);
        ivy = denom * ((x2 - x) * r0y + (x - x1) * r1y);
    } else {
        /* traingle_interpolate :67-79 on (x1,y1), (x1,y2), (x2,y1) */
        double X1 = x1, Y1 = y1, X2 = x1, Y2 = y2, X3 = x2, Y3 = y1;
        double den = (Y2 - Y3) * (X1 - X3) + (X3 - X2) * (Y1 - Y3);
        double W0 = ((Y2 - Y3) * (x - X3) + (X3 - X2) * (y - Y3)) / den;
        double W1 = ((Y3 - Y1) * (x - X3) + (X1 - X3) * (y - Y3)) / den;
        double W2 = 1 - W0 - W1;
        ivx = vx[0] * W0 + vx[1] * W1 + vx[2] * W2;
        ivy = vy[0] * W0 + vy[1] * W1 + vy[2] * W2;
    }
    *th_out = py_mod(atan2(ivy, ivx) + 2 * M_PI, 2 * M_PI); /* get_angle_in_0_2pi */
    *v_out = sqrt(ivx * ivx + ivy * ivy);
    return 0;
}

/* One trajectory: plot_exact_trajectory_set_DP's realization loop body :436-516 */
static int ro_one(const ro_ctx* c, int si0, int sj0, int rzn, int interp, double Cf, double Cr,
                  int8_t* succ_out, int32_t* travel_out, double* en, double* xy) {
    const e2e_params* p = c->p;
    const e2e_fields* f = c->f;
    int g = c->g, nt = c->nt;
    int t = 0, i = si0, j = sj0;
    double x = c->xs[j], y = c->ys[g - 1 - i]; /* set_state :119-129 */
    int a = ro_action(c, t, i, j);
    double thrust = c->speed[a], cth = c->ax[a], sth = c->ay[a]; /* cos/sin products of the table */
    int table = 1;
    double ang = 0.0;
    int succ = 2; /* None */
    int travel = 0, bad = 0;
    double e_cons = 0, e_col = 0, e_net = 0;
    if (xy) {
        xy[0] = x;
        xy[1] = y;
    }
    for (int it = 0; it < nt; it++) {
        if (succ == 2) {
            double Vx, Vy;
            ro_velocity(c, t, i, j, rzn, &Vx, &Vy);
            double rad1 = f->s ? (double)f->s[(long long)t * c->nc + (long long)i * g + j] : 0.0;
            /* move_exact :160-230 (acts only while not terminal and t < nt-1) */
            if (!ro_terminal(p, i, j) && t < nt - 1) {
                double vnetx = (table ? cth : thrust * cos(ang)) + Vx;
                double vnety = (table ? sth : thrust * sin(ang)) + Vy;
                double xnew = x + (vnetx * (double)p->dt), ynew = y + (vnety * (double)p->dt);
                if (xnew > c->xs[g - 1]) xnew = c->xs[g - 1];
                else if (xnew < c->xs[0]) xnew = c->xs[0];
                if (ynew > c->ys[g - 1]) ynew = c->ys[g - 1];
                else if (ynew < c->ys[0]) ynew = c->ys[0];
                x = xnew;
                y = ynew;
                double remx = py_mod(xnew - c->xs[0], c->dj), remy = py_mod(-(ynew - c->ys[g - 1]), c->di);
                double xind = py_floordiv(xnew - c->xs[0], c->dj), yind = py_floordiv(-(ynew - c->ys[g - 1]), c->di);
                if (remx >= 0.5 * c->dj && remy >= 0.5 * c->di) {
                    xind += 1;
                    yind += 1;
                } else if (remx >= 0.5 * c->dj && remy < 0.5 * c->di) {
                    xind += 1;
                } else if (remx < 0.5 * c->dj && remy >= 0.5 * c->di) {
                    yind += 1;
                }
                if (!(isnan(xind) || isnan(yind))) {
                    /* the snap stays on the ticks; the clamp only guards the indexing */
                    i = (int)yind < 0 ? 0 : (int)yind > g - 1 ? g - 1 : (int)yind;
                    j = (int)xind < 0 ? 0 : (int)xind > g - 1 ? g - 1 : (int)xind;
                }
                t += 1;
            }
            double rad2 = f->s ? (double)f->s[(long long)t * c->nc + (long long)i * g + j] : 0.0;
            travel += 1;
            double rr = rad2 + rad1;
            e_cons += (Cf * thrust) * (Cf * thrust);
            e_col += (Cr * rr) / 2;
            e_net += (Cf * (thrust * thrust)) - ((Cr * rr) / 2);
        }
        if (xy) {
            xy[2 * (it + 1)] = x;
            xy[2 * (it + 1) + 1] = y;
        }
        int in_obs = f->mask && f->mask[(long long)t * c->nc + (long long)i * g + j] == 1;
        if (i == 0 || i == g - 1 || j == 0 || j == g - 1 || in_obs) succ = 0; /* :473-474 */
        if (!ro_terminal(p, i, j) && t < nt - 1 && succ == 2) {
            if (interp) {
                if (ro_interp(c, t, i, j, x, y, &thrust, &ang) < 0) {
                    bad = 1;
                    succ = -1;
                    for (int k = it + 1; xy && k < nt; k++) {
                        xy[2 * (k + 1)] = x;
                        xy[2 * (k + 1) + 1] = y;
                    }
                    break;
                }
                table = 0;
            } else {
                a = ro_action(c, t, i, j);
                thrust = c->speed[a];
                cth = c->ax[a];
                sth = c->ay[a];
                table = 1;
            }
        } else if (ro_terminal(p, i, j)) {
            succ = 1;
        } else if (succ != -1) {
            succ = 0;
        }
    }
    *succ_out = (int8_t)succ;
    *travel_out = travel;
    en[0] = e_cons;
    en[1] = e_col;
    en[2] = e_net;
    return bad;
}

int orc_rollout(const e2e_params* p, const e2e_fields* f, const int32_t* policy, int n_starts,
                const int32_t* starts, int interp, double Cf, double Cr, int nthreads,
                int8_t* success, int32_t* travel, double* energy, double* xy) {
    int g = p->gsize;
    if (g < 2 || g > 4096) return -1;
    ro_ctx* c = (ro_ctx*)malloc(sizeof(ro_ctx));
    int na = p->num_ac_speeds * p->num_ac_angles;
    double* tab = (double*)malloc(sizeof(double) * 3 * na);
    orc_rollout_actions(p, tab, NULL, tab + na, tab + 2 * na);
    c->p = p;
    c->f = f;
    c->policy = policy;
    c->g = g;
    c->nc = g * g;
    c->nm = p->nmodes;
    c->nr = p->nrzns;
    c->nt = p->nt;
    c->na = na;
    c->speed = tab;
    c->ax = tab + na;
    c->ay = tab + 2 * na;
    /* Test_grid ticks (Post_processing/utils/setup_grid.py:107-108): (dx/2) + (i*dx) in
     * float32, dj = |xs[1] - xs[0]| (grid_world.py:16-17) */
    for (int k = 0; k < g; k++) {
        c->xs[k] = (double)(p->dx / 2 + (float)k * p->dx);
        c->ys[k] = (double)(p->dy / 2 + (float)k * p->dy);
    }
    c->dj = (double)fabsf((float)c->xs[1] - (float)c->xs[0]);
    c->di = (double)fabsf((float)c->ys[1] - (float)c->ys[0]);
    long long n = (long long)n_starts * p->nrzns;
    int bad = 0;
    if (nthreads <= 0) nthreads = orc_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64) reduction(+ : bad)
    for (long long q = 0; q < n; q++) {
        int k = (int)(q / p->nrzns), r = (int)(q % p->nrzns);
        bad += ro_one(c, starts[2 * k], starts[2 * k + 1], r, interp, Cf, Cr, success + q, travel + q,
                      energy + 3 * q, xy ? xy + q * 2 * (p->nt + 1) : NULL);
    }
    free(tab);
    free(c);
    return bad;
}
