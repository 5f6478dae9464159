// Internal interface between the kernels (pyfwi_kernels.hip) and the plan /
// C-ABI layer (pyfwi_plan.hip).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pyfwi_hip.h"

namespace pyfwi {

constexpr int HALO = 8;      // zero rows above and below each padded field (>= 2*SDO)
constexpr int MARCH_ROWS_MIN = 8;   // strip height bounds (rows walked by a thread)
constexpr int MARCH_ROWS_MAX = 64;
// Wavefronts with work per CU that one launch aims for; the strip height is
// chosen to match.  Measured on MI355X (540 x 1040 cells, 10 shots a launch,
// all four stencil kernels): 20 per CU (18-row strips) runs 5-10 % faster
// than 22-30 per CU (12-17-row strips), although the forward kernels'
// registers allow 32; 13-17 per CU loses 2-6 % again.
constexpr int MARCH_WAVES_PER_CU = 20;
// When one round of MARCH_WAVES_PER_CU waves would need strips taller than
// MARCH_ROWS_ONE_ROUND rows, the launch takes MARCH_ROWS_MANY-row strips
// instead (several rounds of waves): on MI355X at 1040 x 2040 cells with 8
// shots a launch, 12-row strips ran all four stencil kernels 3-13 % faster
// than the 55 rows one round needs (8- and 12-row strips within 1 %, 24-row
// 2-4 % and 48-row 6-13 % slower than 12).
constexpr int MARCH_ROWS_ONE_ROUND = 18;
constexpr int MARCH_ROWS_MANY = 12;
// One-shot launches (one working wave per block): 10 working waves per CU,
// strips of 4 to 12 rows.  On MI355X, 4-row strips ran a 540 x 1040 one-shot
// gradient 21 % faster than 8-row strips; 12-row strips ran a 1040 x 2040
// one 6 % faster than 4-row strips, and a 2040 x 6040 one 12 % faster than
// 39-row strips.
constexpr int MARCH_WAVES_PER_CU_ONE_SHOT = 10;
constexpr int MARCH_ROWS_MIN_ONE_SHOT = 4;
constexpr int MARCH_ROWS_MAX_ONE_SHOT = 12;
constexpr int NUM_XCD = 8;          // MI355X: 8 XCDs, each with its own L2
constexpr int FUSED_WAVES = 4;      // fused two-stage kernels: wavefronts per block
#ifndef MARCH_UNROLL
#define MARCH_UNROLL 2              // rows per unrolled iteration: two rows of loads in flight per wave
#endif
#ifndef FWD_STEP_UNROLL
#define FWD_STEP_UNROLL 2           // rows per unrolled iteration of the one-launch forward step
#endif
#ifndef FWD_TWO_ROWS
#define FWD_TWO_ROWS 36             // strip height of the two-step forward (k_fwd_two)
#endif
#ifndef ADJ_VEL_UNROLL
#define ADJ_VEL_UNROLL 2            // rows per unrolled iteration of the adjoint velocity + imaging kernel
#endif
#ifndef ADJ_GROUP_ROWS
#define ADJ_GROUP_ROWS 2            // adjoint kernels: rows per block-wide sum of the imaging terms
#endif

// Shots per column-marching block (one wave each, over the same 64 columns)
// for a launch of `nshots` shots at half stencil width `sdo`: 4 or 5.
int march_spb(int sdo, int nshots);

// Gradient accumulator sets that the reference-mode adjoint kernels fill for
// `nshots` shots of one pass: one per block's group of shots (the shots of a
// block sum their imaging terms before they reach memory).
inline int adj_acc_sets(int sdo, int nshots)
{
    const int spb = march_spb(sdo, nshots);
    return (nshots + spb - 1) / spb;
}

// Accumulator sets for launches of any 1 .. nshots shots (a pass whose
// launches advance only the shots live at each step).
inline int adj_acc_sets_upto(int sdo, int nshots)
{
    int m = 1;
    for (int n = 1; n <= nshots; ++n) m = adj_acc_sets(sdo, n) > m ? adj_acc_sets(sdo, n) : m;
    return m;
}

// Constants every kernel takes by value.
struct DevGrid {
    int nz, nx;          // tnz, tnx
    int pitch;           // padded row pitch in floats
    int sdo;             // half stencil width
    int npml;
    int64_t fstride;     // floats per padded field
    int64_t origin;      // offset of cell (0, 0) inside a padded field
    float cx[4], cz[4];  // c_k / dx, c_k / dz
    float dt;
    int ncu = 256;       // compute units of the device
    // set per launch by the column-marching launchers
    int mrows = 0, ncb = 0, nstrips = 0, nbatch = 0, nshot = 0;
    int rows_override = 0;  // strip height chosen by the caller (0: automatic)
};

// Velocity half step: state read from `in`, velocities written to `out`
// (per-shot strides in floats; in == out with stride 5*fstride updates in
// place, the only form the time-reversed direction takes).  Forward with
// `src`: the step's source term is added to the values loaded at the source
// cell (memory is not changed); src == nullptr when the caller injected it.
void launch_vel(const DevGrid &g, int dir, int nshots, const pyfwi_shot *shots, int t,
                const float *in, int64_t in_stride, float *out, int64_t out_stride,
                const float *bu, const float *dxp, const float *dzp, hipStream_t st,
                const float *src = nullptr);
// Velocity + stress step in one launch (forward, dir = +1) or stress +
// velocity step (time reversed, dir = -1): reads `in`, writes `out`
// (distinct buffers; per-shot strides in_stride / out_stride, 5*fstride for
// the time-reversed step).  Forward with `src` / `seis`: the source term on
// the loaded values and the record sample of step t+1, as launch_vel and
// launch_str.
void launch_fused_step(const DevGrid &g, int dir, int nshots, const pyfwi_shot *shots, int t,
                       const float *in, int64_t in_stride, float *out, int64_t out_stride,
                       const float *bu, const float *lam, const float *mu, const float *dxp,
                       const float *dzp, hipStream_t st, const float *src = nullptr,
                       float *seis = nullptr);
// Stress half step: forward reads this step's velocities from `out` and the
// previous stresses from `in`, writes the stresses to `out`.  Forward with
// `src`, the source term on txx, tzz at the source cell; with `seis`, the
// record sample of step t+1 (this step's result) at the receivers.
void launch_str(const DevGrid &g, int dir, int nshots, const pyfwi_shot *shots, int t,
                const float *in, int64_t in_stride, float *out, int64_t out_stride,
                const float *lam, const float *mu, const float *dxp, const float *dzp,
                hipStream_t st, const float *src = nullptr, float *seis = nullptr);
// Forward steps t and t+1 in one launch (order 4; false, launching
// nothing, at order 8): state t-1 read from `in`, the states after steps t
// and t+1 written to `out0` and `out1` (per-shot stride out_stride); source
// terms and record samples as launch_fused_step.
bool launch_fwd_two(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                    const float *in, int64_t in_stride, float *out0, float *out1,
                    int64_t out_stride, const float *bu, const float *lam, const float *mu,
                    const float *dxp, const float *dzp, hipStream_t st, const float *src,
                    float *seis);
// Record the state before step t's source term, then add the term in memory
// (the fused steps, the exact mode's recomputation and sources off the
// active cells); with src == nullptr only record (receiver tables, whose
// gather the stencil kernels do not do).  `fields`: per-shot stride
// fields_stride.  rtab: the plan's receiver table (device), read by
// PYFWI_REC_TABLE shots.
void launch_inject_record(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                          float *fields, int64_t fields_stride, const float *src, float *seis,
                          const int32_t *rtab, hipStream_t st);
// The records of steps t0..t1-1 from a forward that kept every state (shot
// s's state after step u at kept + s*shot_stride + u*slot_stride): the same
// rows launch_inject_record(..., src = nullptr, ...) writes step by step.
void launch_record_kept(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t0, int t1,
                        const float *kept, int64_t shot_stride, int64_t slot_stride, float *seis,
                        const int32_t *rtab, hipStream_t st);
// Residual injection in memory at backward step t, components f0..f1-1 (all
// five by default); with first_only, only for shots whose first backward
// step t is.
void launch_adj_inject(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                       float *adj, const float *res, const int32_t *rtab, hipStream_t st,
                       int f0 = 0, int f1 = 5, bool first_only = false);
// Adjoint stresses + the lambda / mu imaging terms; adjoint velocities + the
// rho imaging term.  `fields`: the forward state of the step (per-shot
// stride fields_stride), gacc: [adj_acc_sets(nshots)][6][fstride].  With `res`, the stress
// update adds step t's residual to the adjoint stresses at the receivers and
// the velocity update writes its adjoint velocities with step t-1's residual
// added (launch_adj_inject(..., 0, 2, true) covers each shot's first step);
// res == nullptr when the caller injects every step in memory.
void launch_adj_str(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                    const float *fields, int64_t fields_stride, float *adj, const float *bu,
                    const float *dxp, const float *dzp, float *gacc, bool precond,
                    hipStream_t st, const float *res = nullptr);
void launch_adj_vel_grad(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                         const float *fields, int64_t fields_stride, float *adj,
                         const float *lam, const float *mu, const float *bu,
                         const float *dxp, const float *dzp, float *gacc, bool precond,
                         hipStream_t st, const float *res = nullptr);
// Both of the above in one launch (order 4 only; false, launching nothing,
// at order 8): the adjoint state is read from `ain` and written to `aout`
// (distinct buffers, [nshots][5][fstride] each, swapped by the caller every
// step); the same values, residual terms and accumulators as launch_adj_str
// followed by launch_adj_vel_grad.
bool launch_adj_fused(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                      const float *fields, int64_t fields_stride, const float *ain, float *aout,
                      const float *lam, const float *mu, const float *bu, const float *dxp,
                      const float *dzp, float *gacc, bool precond, hipStream_t st,
                      const float *res = nullptr);
void launch_l2_residual(int nshots, const pyfwi_shot *shots, const float *seis, const float *obs,
                        float *res, int comp_mask, double *misfit, hipStream_t st);
// Records `in` -> `out` (same layout, distinct buffers) filtered along time
// by the symmetric circular FIR k[0..L] on traces padded to n_samples + pad
// (the multiscale lowpass); components of comp_mask (0: txx and tzz), the
// n_rec columns of each shot; h_shots: host copy of the shot table.
// Needs 2 * (L + 128) + 1 + 4096 floats of LDS: L <= 15000.
void launch_fir(int nshots, const pyfwi_shot *shots, const pyfwi_shot *h_shots, const float *in,
                float *out, int comp_mask, const float *taps, int L, int pad, hipStream_t st);
void launch_exact_str(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                      const float *vel, int64_t vel_stride, float *adj, float *av,
                      const float *lam, const float *mu, const float *dxp,
                      const float *dzp, float *gacc, hipStream_t st);
// The exact adjoint's velocity kernel reads the stresses step t updates from
// `tau` (per-shot stride tau_stride); with `src`, a kept state without step
// t's source term, which the kernel adds to txx and tzz at the source cell.
void launch_exact_vel(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                      const float *tau, int64_t tau_stride, float *adj, const float *av_in,
                      float *av_out, const float *bu, const float *dxp, const float *dzp,
                      float *gacc, hipStream_t st, const float *src = nullptr);
void launch_exact_inject(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                         float *adj, float *av, const float *res, const int32_t *rtab,
                         hipStream_t st);
// Zero n floats from p on stream st.
void launch_zero(float *p, int64_t n, hipStream_t st);
void launch_copy_shots(float *dst, int64_t dst_stride, const float *src, int64_t src_stride,
                       int64_t count, int nshots, hipStream_t st);
void launch_grad_sum(const DevGrid &g, int nsets, int nslot, const float *gacc, float *out,
                     const float *bu, int b_to_rho, int accumulate, hipStream_t st);
void launch_pad(const DevGrid &g, const float *src, float *dst, hipStream_t st);
void launch_unpad(const DevGrid &g, const float *src, float *dst, int nf, hipStream_t st);
// out[0]: bits of the largest (lam+2mu)*b of the padded model (the squared
// largest P velocity); out[1]: 1 if lam, mu or b holds NaN or Inf.  The
// caller zeroes the two ints first.
void launch_model_check(const DevGrid &g, const float *lam, const float *mu, const float *bu,
                        int *out, hipStream_t st);

} // namespace pyfwi
