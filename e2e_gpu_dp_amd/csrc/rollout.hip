// rollout.hip — policy rollouts: the post-processing's trajectory simulation
// (Post_processing/plot_e2eGPU_policy.py:376-520 plot_exact_trajectory_set_DP, the
// vehicle step Post_processing/grid_world.py:160-230 Grid.move_exact, the interpolated
// action :174-198 with :44-171), one lane per (start cell, realization) trajectory.
//
// Arithmetic is float64, as the reference's Python floats, on the float32 inputs
// widened where they enter it: the velocity sample in the build's float32 order (mode
// sum, then the mean), the float32 grid ticks of the Test_grid set-up
// (Post_processing/utils/setup_grid.py:107-108) and the float32 speeds. The action
// table's speed*cos(angle), speed*sin(angle) products come from the host (glibc), so
// moves under a table action round as the reference's; the interpolated action's
// atan2 / cos / sin run on the device.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "internal.h"

namespace e2e {

namespace {

struct RolloutArgs {
    const float *u, *v, *ui, *vi, *yi, *s;
    const int32_t* mask;
    const int32_t* policy;   // [nt][g][g]
    const int32_t* starts;   // [n_starts][2] (i, j)
    const double* xs;        // [g] float32 ticks widened
    const double* ys;
    const double* speed;     // [na]
    const double* ax;        // [na] speed*cos(angle)
    const double* ay;        // [na] speed*sin(angle)
    int8_t* success;
    int32_t* travel;
    double* energy;          // [n][3]
    double* xy;              // [n][nt+1][2] or NULL
    unsigned long long* moves;  // total moves (one atomic per wave)
    int64_t n;
    int g, nt, nm, nr, na;
    int i_term, j_term, tsg;
    int interp;
    double dt, dj, di, cf, cr;
};

// Python's a % b and a // b for floats (b > 0)
__device__ __forceinline__ double py_mod(double a, double b) {
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0) != (m < 0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

__device__ __forceinline__ double py_floordiv(double a, double b) {
    const double m = fmod(a, b);
    double d = (a - m) / b;
    if (m != 0.0 && ((b < 0) != (m < 0))) d -= 1.0;
    if (d != 0.0) {
        double fd = floor(d);
        if (d - fd > 0.5) fd += 1.0;
        return fd;
    }
    return copysign(0.0, a / b);
}

__device__ __forceinline__ bool terminal(const RolloutArgs& A, int i, int j) {
    return i >= A.i_term && i < A.i_term + A.tsg && j >= A.j_term && j < A.j_term + A.tsg;
}

// g.actions[policy_1d[s1_id]]: a policy of -1 indexes the last action, as a Python list
__device__ __forceinline__ int action_at(const RolloutArgs& A, int t, int i, int j) {
    const int a = A.policy[((int64_t)t * A.g + i) * A.g + j];
    return a < 0 ? a + A.na : a;
}

// get_interpolated_action (plot_e2eGPU_policy.py:174-198, interpolate_action_angles
// :113-171, interpolate_as_per_term_status :84-110); false for the support patterns
// the reference rejects (:158-161)
__device__ bool interp_action(const RolloutArgs& A, int t, int i, int j, double x, double y, double& v_out,
                              double& th_out) {
    const int g = A.g;
    const double delx = x - A.xs[j], dely = y - A.ys[g - 1 - i];
    const int fx = delx < 0 ? -1 : 1, fy = dely < 0 ? -1 : 1;
    const int si[4] = {i, i - fy, i, i - fy}, sj[4] = {j, j, j + fx, j + fx};
    bool term[4];
    double vx[4], vy[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        term[k] = si[k] == A.i_term && sj[k] == A.j_term;  // is_terminal(query_s)
        const int a = action_at(A, t, si[k], sj[k]);
        vx[k] = A.ax[a];
        vy[k] = A.ay[a];
    }
    const double x1 = A.xs[sj[0]], y1 = A.ys[g - 1 - si[0]];
    const double x2 = A.xs[sj[3]], y2 = A.ys[g - 1 - si[3]];
    bool tri = false;
    if (!term[0] && !term[1] && !term[2] && !term[3]) {
    } else if (!term[0] && !term[1] && term[2]) {
        vx[2] = vx[0];
        vx[3] = vx[1];
        vy[2] = vy[0];
        vy[3] = vy[1];
    } else if (!term[0] && term[1] && !term[2]) {
        vx[1] = vx[0];
        vx[3] = vx[2];
        vy[1] = vy[0];
        vy[3] = vy[2];
    } else if (!term[0] && !term[1] && !term[2] && term[3]) {
        tri = true;
    } else {
        return false;
    }
    double ivx, ivy;
    if (!tri) {  // bilinear_interpolate :44-63
        const double denom = 1.0 / ((x2 - x1) * (y2 - y1));
        const double r0x = vx[0] * (y2 - y) + vx[1] * (y - y1), r1x = vx[2] * (y2 - y) + vx[3] * (y - y1);
        const double r0y = vy[0] * (y2 - y) + vy[1] * (y - y1), r1y = vy[2] * (y2 - y) + vy[3] * (y - y1);
        ivx = denom * ((x2 - x) * r0x + (x - x1) * r1x);
        ivy = denom * ((x2 - x) * r0y + (x - x1) * r1y);
    } else {  // traingle_interpolate :67-79 on (x1,y1), (x1,y2), (x2,y1)
        const double X1 = x1, Y1 = y1, X2 = x1, Y2 = y2, X3 = x2, Y3 = y1;
        const double den = (Y2 - Y3) * (X1 - X3) + (X3 - X2) * (Y1 - Y3);
        const double W0 = ((Y2 - Y3) * (x - X3) + (X3 - X2) * (y - Y3)) / den;
        const double W1 = ((Y3 - Y1) * (x - X3) + (X1 - X3) * (y - Y3)) / den;
        const double W2 = 1 - W0 - W1;
        ivx = vx[0] * W0 + vx[1] * W1 + vx[2] * W2;
        ivy = vy[0] * W0 + vy[1] * W1 + vy[2] * W2;
    }
    th_out = py_mod(atan2(ivy, ivx) + 2 * M_PI, 2 * M_PI);  // get_angle_in_0_2pi
    v_out = sqrt(ivx * ivx + ivy * ivy);
    return true;
}

__global__ __launch_bounds__(256) void rollout_kernel(const RolloutArgs A) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long my_moves = 0;
    if (q < A.n) {
        const int g = A.g, nt = A.nt;
        const int64_t nc = (int64_t)g * g;
        const int k = (int)(q / A.nr), rzn = (int)(q - (int64_t)k * A.nr);
        int t = 0, i = A.starts[2 * k], j = A.starts[2 * k + 1];
        double x = A.xs[j], y = A.ys[g - 1 - i];  // set_state (grid_world.py:119-129)
        int a = action_at(A, t, i, j);
        double thrust = A.speed[a], cth = A.ax[a], sth = A.ay[a], ang = 0.0;
        bool table = true;
        int succ = 2;  // None
        int trav = 0;
        double e_cons = 0, e_col = 0, e_net = 0;
        double* xy = A.xy ? A.xy + q * 2 * (int64_t)(nt + 1) : nullptr;
        if (xy) {
            xy[0] = x;
            xy[1] = y;
        }
        int it = 0;
        for (; it < nt; it++) {
            if (succ == 2) {
                // extract_velocity in the build's float32 order
                const int64_t cell = (int64_t)i * g + j;
                const int64_t mid = (int64_t)t * nc + cell;
                const int64_t uvi = (int64_t)t * A.nm * nc + cell;
                const int64_t yb = ((int64_t)t * A.nr + rzn) * A.nm;
                float sx = 0.0f, sy = 0.0f;
                for (int m = 0; m < A.nm; m++) sx += A.ui[uvi + m * nc] * A.yi[yb + m];
                for (int m = 0; m < A.nm; m++) sy += A.vi[uvi + m * nc] * A.yi[yb + m];
                const double Vx = (double)(A.u[mid] + sx), Vy = (double)(A.v[mid] + sy);
                const double rad1 = A.s ? (double)A.s[mid] : 0.0;
                if (!terminal(A, i, j) && t < nt - 1) {  // move_exact :160-230
                    const double vnetx = (table ? cth : thrust * cos(ang)) + Vx;
                    const double vnety = (table ? sth : thrust * sin(ang)) + Vy;
                    double xnew = x + (vnetx * A.dt), ynew = y + (vnety * A.dt);
                    if (xnew > A.xs[g - 1]) xnew = A.xs[g - 1];
                    else if (xnew < A.xs[0]) xnew = A.xs[0];
                    if (ynew > A.ys[g - 1]) ynew = A.ys[g - 1];
                    else if (ynew < A.ys[0]) ynew = A.ys[0];
                    x = xnew;
                    y = ynew;
                    const double remx = py_mod(xnew - A.xs[0], A.dj), remy = py_mod(-(ynew - A.ys[g - 1]), A.di);
                    double xind = py_floordiv(xnew - A.xs[0], A.dj), yind = py_floordiv(-(ynew - A.ys[g - 1]), A.di);
                    if (remx >= 0.5 * A.dj && remy >= 0.5 * A.di) {
                        xind += 1;
                        yind += 1;
                    } else if (remx >= 0.5 * A.dj && remy < 0.5 * A.di) {
                        xind += 1;
                    } else if (remx < 0.5 * A.dj && remy >= 0.5 * A.di) {
                        yind += 1;
                    }
                    if (!(isnan(xind) || isnan(yind))) {
                        // the snap stays on the ticks; the clamp only guards the indexing
                        i = min(max((int)yind, 0), g - 1);
                        j = min(max((int)xind, 0), g - 1);
                    }
                    t += 1;
                }
                const double rad2 = A.s ? (double)A.s[(int64_t)t * nc + (int64_t)i * g + j] : 0.0;
                trav += 1;
                const double rr = rad2 + rad1;
                e_cons += (A.cf * thrust) * (A.cf * thrust);
                e_col += (A.cr * rr) / 2;
                e_net += (A.cf * (thrust * thrust)) - ((A.cr * rr) / 2);
            }
            if (xy) {
                xy[2 * (it + 1)] = x;
                xy[2 * (it + 1) + 1] = y;
            }
            const bool obs = A.mask && A.mask[(int64_t)t * nc + (int64_t)i * g + j] == 1;
            if (i == 0 || i == g - 1 || j == 0 || j == g - 1 || obs) succ = 0;  // :473-474
            if (!terminal(A, i, j) && t < nt - 1 && succ == 2) {
                if (A.interp) {
                    if (!interp_action(A, t, i, j, x, y, thrust, ang)) {
                        succ = -1;
                        break;
                    }
                    table = false;
                } else {
                    a = action_at(A, t, i, j);
                    thrust = A.speed[a];
                    cth = A.ax[a];
                    sth = A.ay[a];
                    table = true;
                }
            } else if (terminal(A, i, j)) {
                succ = 1;
            } else {
                succ = 0;
            }
            if (succ != 2 && !xy) break;  // decided: the state no longer changes
        }
        for (it = it + 1; xy && it < nt; it++) {  // positions after a rejected interpolation
            xy[2 * (it + 1)] = x;
            xy[2 * (it + 1) + 1] = y;
        }
        A.success[q] = (int8_t)succ;
        A.travel[q] = trav;
        A.energy[3 * q] = e_cons;
        A.energy[3 * q + 1] = e_col;
        A.energy[3 * q + 2] = e_net;
        my_moves = (unsigned long long)trav;
    }
    // one atomic per wave
    for (int o = 32; o >= 1; o >>= 1) my_moves += __shfl_xor(my_moves, o, 64);
    if ((threadIdx.x & 63) == 0 && my_moves) atomicAdd(A.moves, my_moves);
}

// round(x, 14) (grid_world.py:352): the double nearest the decimal string of x correctly
// rounded to 14 places
double py_round14(double x) {
    char buf[64];
    snprintf(buf, sizeof buf, "%.14f", x);
    return strtod(buf, nullptr);
}

}  // namespace

void rollout_actions(const e2e_params* p, double* speed, double* angle, double* ax, double* ay) {
    const int ns = p->num_ac_speeds, na = p->num_ac_angles;
    for (int i = 0; i < ns; i++) {
        // populate_ac_speeds (grid_world.py:322-337), float32 as the prob_params entry
        const float sp = ns == 1 ? p->F : (float)(i + 1) * (p->F / (float)ns);
        for (int j = 0; j < na; j++) {
            const double an = py_round14((double)(j * 2) * M_PI / (double)na);  // :351-352
            const int idx = j + na * i;  // itertools.product(speed_list, angle_list) :356
            if (speed) speed[idx] = (double)sp;
            if (angle) angle[idx] = an;
            if (ax) ax[idx] = (double)sp * std::cos(an);
            if (ay) ay[idx] = (double)sp * std::sin(an);
        }
    }
}

e2e_status rollout(const e2e_params* p, const e2e_fields* f, const int32_t* policy, int32_t n_starts,
                   const int32_t* starts, const e2e_rollout_opts* o, int8_t* success, int32_t* travel,
                   double* energy, double* xy, e2e_rollout_stats* stats) {
    E2E_TRY(check_params(p));
    E2E_CHECK(f && f->u && f->v && f->yi && (p->nmodes == 0 || (f->ui && f->vi)), E2E_ERR_INVALID,
              "rollout: fields u, v, yi (and ui, vi with modes) are required");
    E2E_CHECK(policy && starts && n_starts >= 1, E2E_ERR_INVALID, "rollout: policy and >= 1 start cell required");
    E2E_CHECK(success && travel && energy, E2E_ERR_INVALID, "rollout: success, travel and energy are required");
    e2e_rollout_opts opts{};
    opts.interp = 1;
    if (o) opts = *o;
    const int g = p->gsize, nt = p->nt, nm = p->nmodes, nr = p->nrzns;
    const int na = p->num_ac_speeds * p->num_ac_angles;
    for (int k = 0; k < n_starts; k++)
        E2E_CHECK(starts[2 * k] >= 0 && starts[2 * k] < g && starts[2 * k + 1] >= 0 && starts[2 * k + 1] < g,
                  E2E_ERR_INVALID, "rollout: start cell %d (%d, %d) outside the grid", k, starts[2 * k],
                  starts[2 * k + 1]);
    const int64_t nc = (int64_t)g * g, ns = nc * nt;
    for (int64_t s = 0; s < ns; s++)
        E2E_CHECK(policy[s] >= -na && policy[s] < na, E2E_ERR_INVALID, "rollout: policy[%lld] = %d is not an action",
                  (long long)s, policy[s]);
    const int64_t n = (int64_t)n_starts * nr;

    hipStream_t s = (hipStream_t)opts.stream;
    const bool own_stream = s == nullptr;
    if (own_stream) E2E_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        bool own;
        ~StreamGuard() {
            if (own && s) (void)hipStreamDestroy(s);
        }
    } sg{s, own_stream};

    // fields
    DevBuf dU, dV, dUi, dVi, dYi, dS, dM;
    RolloutArgs A{};
    A.u = f->u;
    A.v = f->v;
    A.ui = f->ui;
    A.vi = f->vi;
    A.yi = f->yi;
    A.s = f->s;
    A.mask = f->mask;
    if (!opts.fields_on_device) {
        auto up = [&](DevBuf& b, const void* h, size_t bytes, const void** dst) -> e2e_status {
            if (!h) return E2E_OK;
            E2E_TRY(dmalloc(b, bytes));
            E2E_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s));
            *dst = b.p;
            return E2E_OK;
        };
        const size_t f1 = sizeof(float) * ns, fm = f1 * (nm > 0 ? nm : 1);
        E2E_TRY(up(dU, f->u, f1, (const void**)&A.u));
        E2E_TRY(up(dV, f->v, f1, (const void**)&A.v));
        if (nm) {
            E2E_TRY(up(dUi, f->ui, fm, (const void**)&A.ui));
            E2E_TRY(up(dVi, f->vi, fm, (const void**)&A.vi));
        }
        E2E_TRY(up(dYi, f->yi, sizeof(float) * nt * nr * (nm > 0 ? nm : 1), (const void**)&A.yi));
        E2E_TRY(up(dS, f->s, f1, (const void**)&A.s));
        E2E_TRY(up(dM, f->mask, sizeof(int32_t) * ns, (const void**)&A.mask));
    }
    // ticks, action table, policy, starts
    std::vector<double> tab(2 * g + 3 * na);
    for (int k = 0; k < g; k++) {
        tab[k] = (double)(p->dx / 2 + (float)k * p->dx);
        tab[g + k] = (double)(p->dy / 2 + (float)k * p->dy);
    }
    rollout_actions(p, tab.data() + 2 * g, nullptr, tab.data() + 2 * g + na, tab.data() + 2 * g + 2 * na);
    DevBuf dTab, dPol, dSt, dSucc, dTrav, dEn, dXY, dMoves;
    E2E_TRY(dmalloc(dTab, sizeof(double) * tab.size()));
    E2E_TRY(dmalloc(dPol, sizeof(int32_t) * ns));
    E2E_TRY(dmalloc(dSt, sizeof(int32_t) * 2 * n_starts));
    E2E_TRY(dmalloc(dSucc, n));
    E2E_TRY(dmalloc(dTrav, sizeof(int32_t) * n));
    E2E_TRY(dmalloc(dEn, sizeof(double) * 3 * n));
    E2E_TRY(dmalloc(dMoves, sizeof(unsigned long long)));
    if (xy) E2E_TRY(dmalloc(dXY, sizeof(double) * 2 * (size_t)(nt + 1) * n));
    E2E_HIP(hipMemcpyAsync(dTab.p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, s));
    E2E_HIP(hipMemcpyAsync(dPol.p, policy, sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
    E2E_HIP(hipMemcpyAsync(dSt.p, starts, sizeof(int32_t) * 2 * n_starts, hipMemcpyHostToDevice, s));
    E2E_HIP(hipMemsetAsync(dMoves.p, 0, sizeof(unsigned long long), s));
    A.policy = dPol.as<int32_t>();
    A.starts = dSt.as<int32_t>();
    A.xs = dTab.as<double>();
    A.ys = A.xs + g;
    A.speed = A.xs + 2 * g;
    A.ax = A.speed + na;
    A.ay = A.speed + 2 * na;
    A.success = dSucc.as<int8_t>();
    A.travel = dTrav.as<int32_t>();
    A.energy = dEn.as<double>();
    A.xy = xy ? dXY.as<double>() : nullptr;
    A.moves = dMoves.as<unsigned long long>();
    A.n = n;
    A.g = g;
    A.nt = nt;
    A.nm = nm;
    A.nr = nr;
    A.na = na;
    A.i_term = p->i_term;
    A.j_term = p->j_term;
    A.tsg = p->term_subgrid_size;
    A.interp = opts.interp ? 1 : 0;
    A.dt = (double)p->dt;
    // dj = |xs[1] - xs[0]| in float32 (grid_world.py:16-17)
    A.dj = (double)std::fabs((float)tab[1] - (float)tab[0]);
    A.di = (double)std::fabs((float)tab[g + 1] - (float)tab[g]);
    A.cf = opts.cf != 0.0 ? opts.cf : 1.0;
    A.cr = opts.cr != 0.0 ? opts.cr : 1.0;

    hipEvent_t e0, e1;
    E2E_HIP(hipEventCreate(&e0));
    E2E_HIP(hipEventCreate(&e1));
    E2E_HIP(hipEventRecord(e0, s));
    hipLaunchKernelGGL(rollout_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, A);
    E2E_HIP(hipGetLastError());
    E2E_HIP(hipEventRecord(e1, s));
    unsigned long long moves = 0;
    E2E_HIP(hipMemcpyAsync(success, dSucc.p, n, hipMemcpyDeviceToHost, s));
    E2E_HIP(hipMemcpyAsync(travel, dTrav.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    E2E_HIP(hipMemcpyAsync(energy, dEn.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
    if (xy) E2E_HIP(hipMemcpyAsync(xy, dXY.p, sizeof(double) * 2 * (size_t)(nt + 1) * n, hipMemcpyDeviceToHost, s));
    E2E_HIP(hipMemcpyAsync(&moves, dMoves.p, sizeof(moves), hipMemcpyDeviceToHost, s));
    E2E_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    E2E_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int64_t rejected = 0, reached = 0;
    for (int64_t q = 0; q < n; q++) {
        rejected += success[q] < 0;
        reached += success[q] == 1;
    }
    if (stats) {
        *stats = e2e_rollout_stats{};
        stats->trajectories = n;
        stats->reached = reached;
        stats->rejected = rejected;
        stats->moves = (int64_t)moves;
        stats->kernel_ms = ms;
    }
    if (rejected) {
        set_error("rollout: %lld trajectories met an interpolation support pattern the reference rejects",
                  (long long)rejected);
        return E2E_ERR_INVALID;
    }
    return E2E_OK;
}

}  // namespace e2e
