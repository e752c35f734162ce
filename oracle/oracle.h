/*
 * oracle.h — CPU restatement of the reference planner's arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker and the CPU baseline: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (libe2edp) never links, loads or calls it.
 *
 * It follows src/build_model.cu (the only builder the reference's Makefile compiles,
 * src/Makefile:23) and src/spvi_2.cu, compiled with -O2 -ffp-contract=off on SSE2 so
 * every float operation rounds to binary32 exactly where the reference's source does.
 */
#ifndef E2E_ORACLE_H
#define E2E_ORACLE_H

#include <stdint.h>
#include "../include/e2edp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model orc_model;

/* Action table (src/build_model.cu:1403-1457) and the host-side trig products
 * F*cosf(angle), F*sinf(angle) of move() (:284-285). Any output may be NULL. */
int orc_actions(const e2e_params* p, float* speed, float* angle, float* fcos, float* fsin);

/* xs / ys of define_xs_or_ys (src/build_model.cu:1395-1399). */
void orc_grid(const e2e_params* p, float* xs, float* ys);

/* One Monte Carlo lane: (t, action, spatial cell, realization) -> S2 cell and reward,
 * transition_calc body src/build_model.cu:506-545. Returns 0, or -1 when the velocity
 * sample is not finite. */
int orc_lane(const e2e_params* p, const e2e_fields* f, int t, int a, int sp, int rzn,
             int32_t* sp2, float* r);

/* All lanes of one (t, a): sp2[sp*nrzns + rzn], r[sp*nrzns + rzn]. */
int orc_lanes(const e2e_params* p, const e2e_fields* f, int t, int a,
              int32_t* sp2, float* r);

/* Builds the model for slices [t_begin, t_end) (t_end <= nt-1) with nthreads OpenMP
 * threads (0: all). Rows of S1 in the reference order (a, t, sp), S2 ascending. */
orc_model* orc_build(const e2e_params* p, const e2e_fields* f, int t_begin, int t_end,
                     int nthreads);
int64_t orc_model_nnz(const orc_model* m);
int64_t orc_model_out_of_window(const orc_model* m);
int64_t orc_model_max_row_nnz(const orc_model* m);
/* S1,S2,val: nnz entries; R, Rsum: Na*Ns entries (rows a*Ns+s; slices outside the
 * built range and the last slice are 0). Any pointer may be NULL. */
void orc_model_copy(const orc_model* m, int64_t* S1, int64_t* S2, float* val, float* R,
                    float* Rsum);
void orc_model_free(orc_model* m);

/* Value iteration oracles on a model in the reference COO layout (rows a*ns + s).
 * FP64 backward induction over time slices (exact fixed point of the finite-horizon
 * DAG). Returns 0, or -1 if some transition does not go from slice t to slice t+1. */
int orc_vi_backward_f64(int64_t ns, int64_t na, int64_t nc, int64_t nnz,
                        const int64_t* S1, const int64_t* S2, const float* val,
                        const float* R, double discount, double* V, int32_t* pi);

/* FP64 sweep-form value iteration: stop at the first sweep whose sup-norm residual
 * is < thresh (src/spvi_2.cu:669-698). Returns the number of sweeps (or -1). */
int orc_vi_sweeps_f64(int64_t ns, int64_t na, int64_t nnz, const int64_t* S1,
                      const int64_t* S2, const float* val, const float* R,
                      double discount, double thresh, int max_sweeps,
                      double* V, int32_t* pi, double* last_residual);

/* FP32 sweep-form value iteration in the reference's order (Q = R + discount*P V with
 * the row sum accumulated sequentially, strict '>' argmax from -1e6/-1), OpenMP over
 * states: the CPU baseline of the solver. Returns the number of sweeps (or -1). */
int orc_vi_sweeps_f32(int64_t ns, int64_t na, int64_t nnz, const int64_t* S1,
                      const int64_t* S2, const float* val, const float* R,
                      float discount, float thresh, int max_sweeps, int nthreads,
                      float* V, int32_t* pi, float* last_residual);

/* Policy rollouts: the post-processing's trajectory simulation
 * (Post_processing/plot_e2eGPU_policy.py:376-520 plot_exact_trajectory_set_DP, with
 * Grid.move_exact Post_processing/grid_world.py:160-230). Trajectory n = start k x
 * realization r (n = k*nrzns + r) starts at starts[2k], starts[2k+1] (row i, column j).
 * interp 1: actions after the first bilinearly interpolated (get_interpolated_action
 * :174-198, the logged run :949-952); 0: the grid policy. Outputs per trajectory:
 * success (1 reached the terminal subgrid, 0 not), travel (moves), energy[3n..3n+2]
 * (energy_cons, energy_col, net_energy_cons), xy[(n*(nt+1) + step)*2 + {0,1}] (positions,
 * NULL to skip; a position repeats once the trajectory has ended). Arithmetic in float64
 * on the float32 inputs widened. Returns the number of trajectories whose interpolation
 * met a support pattern the reference rejects (:158-161; success then -1), or -1. */
int orc_rollout(const e2e_params* p, const e2e_fields* f, const int32_t* policy, int n_starts,
                const int32_t* starts, int interp, double Cf, double Cr, int nthreads,
                int8_t* success, int32_t* travel, double* energy, double* xy);

/* The rollout's action table (Post_processing/grid_world.py:322-365): speed (float32
 * widened), angle round(i*2*pi/na, 14), and speed*cos(angle), speed*sin(angle) (glibc). */
void orc_rollout_actions(const e2e_params* p, double* speed, double* angle, double* ax, double* ay);

/* max_s | (T V)(s) - V(s) | in FP64 for a float value vector V (the final residual). */
double orc_bellman_residual_f64(int64_t ns, int64_t na, int64_t nnz, const int64_t* S1,
                                const int64_t* S2, const float* val, const float* R,
                                double discount, const float* V);

/* Test hook: plants a deliberate mistake in the oracle so tests can show that the
 * pins catch it. 0 = none (default). */
void orc_set_mutation(int code);

int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
