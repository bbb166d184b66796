/*
 * tqsv.h -- C ABI of libtqsv.so, the MI355X (gfx950) statevector kernels
 * behind torchquantum_amd.
 *
 * The reference (mit-han-lab/torchquantum 0.1.8) is pure Python over torch;
 * its hot path has no native interface.  Each entry point below replaces the
 * torch computation named in its comment (file:line in the reference tree).
 *
 * Conventions (all entry points):
 *   - state: device pointer to bsz rows of 2^n complex64 amplitudes
 *     (interleaved re, im float32), row b at [b << n, (b + 1) << n).
 *     Wire w of the reference API is bit position (n - 1 - w) of the in-row
 *     index (torchquantum/device/devices.py:64-71); every argument named
 *     *bit* / *mask* below is in bit positions.
 *   - matrices: device pointers to row-major complex64; a 2-qubit matrix is
 *     indexed 2*b(bit_hi) + b(bit_lo) (gate_wrapper.py:120-127: wires[0] is
 *     the most significant bit).  mat_bstride: complex elements between the
 *     matrices of consecutive batch rows, 0 = one matrix shared by all rows.
 *   - controls: cmask selects control bits, cval their required values.
 *   - stream: hipStream_t (passed as void*); every call is asynchronous on it.
 *   - return value: 0 on success, negative on error (-1 invalid argument,
 *     -2 HIP launch/runtime error, -3 a documented kernel limit exceeded);
 *     tqsv_last_error() describes the last error of the calling thread.
 *   - no entry point allocates device memory; workspaces come from the caller.
 */
#ifndef TQSV_H
#define TQSV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* tqsv_stream_t; /* hipStream_t */

int tqsv_version(void);
const char* tqsv_last_error(void);
int tqsv_num_cu(void);

/* |index> in every row.  Replaces QuantumDevice.__init__ / reset_states
 * (torchquantum/device/devices.py:64-71, 89-92). */
int tqsv_init_basis(tqsv_stream_t s, void* state, int64_t bsz, int n, int64_t index);

/* ---- per-gate kernels: replace apply_unitary_bmm / apply_unitary_einsum
 *      (torchquantum/functional/gate_wrapper.py:18-140) ------------------- */

/* (controlled) 2x2 on bit tbit. */
int tqsv_apply_1q(tqsv_stream_t s, void* state, int64_t bsz, int n, int tbit, uint64_t cmask, uint64_t cval,
                  const void* mat, int64_t mat_bstride);
/* (controlled) 4x4 on bits (bit_hi, bit_lo). */
int tqsv_apply_2q(tqsv_stream_t s, void* state, int64_t bsz, int n, int bit_hi, int bit_lo, uint64_t cmask,
                  uint64_t cval, const void* mat, int64_t mat_bstride);
/* diagonal: amplitudes j with (j & cmask) == cval are multiplied by
 * phases[popcount(j & pmask) & 1] (phases: 2 complex; ph_bstride as above).
 * Covers RZ, RZZ, MultiRZ, CRZ, Z, S, T, PhaseShift, CZ, CCZ, CU1 ... */
int tqsv_apply_diag(tqsv_stream_t s, void* state, int64_t bsz, int n, uint64_t cmask, uint64_t cval,
                    uint64_t pmask, const void* phases, int64_t ph_bstride);
/* (controlled) exchange of bits a and b: SWAP, CSWAP. */
int tqsv_apply_swap(tqsv_stream_t s, void* state, int64_t bsz, int n, int bit_a, int bit_b, uint64_t cmask,
                    uint64_t cval);
/* dense 2^k x 2^k (k <= 10) on host-listed bits (bits[0] = matrix MSB):
 * qubitunitary / qubitunitaryfast (functional/qubit_unitary.py:19-69), QFT, RCCX,
 * fused blocks (graph/graphs.py:836-860).  k = 5, 6 with one matrix for all
 * rows (mat_bstride = 0) and >= 64 groups per row run on the matrix cores. */
int tqsv_apply_kq(tqsv_stream_t s, void* state, int64_t bsz, int n, int k, const int* bits, uint64_t cmask,
                  uint64_t cval, const void* mat, int64_t mat_bstride);
/* A run of the calls above from one host array of gate records
 * (tqsv_gate_words() int64 words each: kind 1 = 1q, 2 = 2q, 3 = swap,
 * 4 = diagonal, 5 = k-qubit; k; bits[10]; cmask; cval; matrix offset in
 * complex elements from mats; matrix row stride).  Replaces the reference's
 * Python loop of gate_wrapper calls over a circuit's gates
 * (torchquantum/functional/gate_wrapper.py:324-446, called once per gate from
 * QuantumModule.forward).  Stops at the first invalid gate (its index is in
 * the message); gates before it have been enqueued. */
int tqsv_gate_words(void);
int tqsv_run_gates(tqsv_stream_t s, void* state, int64_t bsz, int n, const int64_t* gates, int n_gates,
                   const void* mats);

/* ---- fused stages: replace the static-mode block application
 *      (torchquantum/graph/graphs.py:384-532, 836-860) and runs of per-gate
 *      calls.  prog_host/prog_dev: the same stage program (int32 words, format
 *      in tqsv_stage.hip) on host and device; mats: matrix pool.  state_in may
 *      equal state_out, or be NULL with init_zero = 1 (start from |0..0>).
 *      grid_cap: workgroups per launch, 0 = automatic.  n <= 32. ---------- */
int tqsv_stage_forward(tqsv_stream_t s, void* state_out, const void* state_in, int64_t bsz, int n,
                       const int* prog_host, const void* prog_dev, const void* mats, int init_zero, int grid_cap);
/* Reverse sweep of one stage for the adjoint method: psi <- U^dag psi,
 * lam <- U^dag lam, grad[row * grad_stride + slot] += Re<lam|dU/dtheta psi>.
 * Replaces torch autograd through gate_wrapper.py:127-136.  work: device
 * buffer of at least tqsv_stage_adjoint_work_floats() floats for the
 * per-workgroup partials, which a second launch adds into grad in a fixed
 * order (gradients are bitwise reproducible; no float atomics). */
int tqsv_stage_adjoint(tqsv_stream_t s, void* psi, void* lam, int64_t bsz, int n, const int* prog_host,
                       const void* prog_dev, const void* mats, float* grad, int64_t grad_stride, float* work,
                       int64_t work_floats, int grid_cap);
/* Partial-buffer floats tqsv_stage_adjoint / tqsv_jit_stage_adjoint need for
 * this stage and launch shape (< 0 on a bad program). */
int64_t tqsv_stage_adjoint_work_floats(const int* prog_host, int n, int64_t bsz, int grid_cap);

/* ---- run-time compiled stages: the same stage program turned by the host
 *      planner into straight-line HIP source (every op unrolled with its bits,
 *      controls and matrix offsets as constants), compiled here with hipRTC
 *      for gfx950, and launched with the launch shape and argument meaning of
 *      tqsv_stage_forward / tqsv_stage_adjoint above (same reference
 *      counterparts).  Needs 2^T >= 64 threads per group (m >= 10). ------- */
/* source -> code object (no GPU needed).  opts (optional): extra compiler
 * options separated by spaces.  code may be NULL to query *size; log
 * (optional) receives the compiler log. */
int tqsv_jit_compile(const char* src, const char* opts, void* code, int64_t cap, int64_t* size, char* log,
                     int64_t log_cap);
/* load a code object on the current device; funcs[i] = kernel names[i]. */
int tqsv_jit_load(const void* code, int64_t size, int n, const char* const* names, void** module, void** funcs);
int tqsv_jit_unload(void* module);
/* func: kernel compiled from prog_host's program (parameters: out, in, mats,
 * init_zero, n_groups, groups_per_wg). */
int tqsv_jit_stage_forward(tqsv_stream_t s, const void* func, void* state_out, const void* state_in, int64_t bsz,
                           int n, const int* prog_host, const void* mats, int init_zero, int grid_cap);
/* func: adjoint kernel compiled from prog_host's program (parameters: psi,
 * lam, psi_in, lam_in, mats, partials, rows per workgroup, n_groups,
 * groups_per_wg).  psi_in / lam_in: where psi and lam are read (NULL = in
 * place), so the first stage of a sweep can read the forward output and the
 * incoming gradient without copying them.  prog_dev, work: as
 * tqsv_stage_adjoint (the fixed-order reduction reads the gradient slot
 * table from prog_dev). */
int tqsv_jit_stage_adjoint(tqsv_stream_t s, const void* func, void* psi, void* lam, const void* psi_in,
                           const void* lam_in, int64_t bsz, int n, const int* prog_host, const void* prog_dev,
                           const void* mats, float* grad, int64_t grad_stride, float* work, int64_t work_floats,
                           int grid_cap);

/* ---- per-gate adjoint steps (the unfused gradient path); mat holds U then
 *      P derivative matrices dU/dtheta_p (P <= 4); grad slots gslot..gslot+P-1.
 *      Replace autograd through apply_unitary_bmm (gate_wrapper.py:127-136).
 *      work: at least tqsv_adjoint_work_floats(bsz, n) floats of device
 *      memory for the per-workgroup partials (fixed-order sum, no float
 *      atomics). */
int tqsv_adjoint_1q(tqsv_stream_t s, void* psi, void* lam, int64_t bsz, int n, int tbit, uint64_t cmask,
                    uint64_t cval, const void* mat, int64_t mat_bstride, int P, float* grad, int64_t grad_stride,
                    int gslot, float* work, int64_t work_floats);
int tqsv_adjoint_2q(tqsv_stream_t s, void* psi, void* lam, int64_t bsz, int n, int bit_hi, int bit_lo,
                    uint64_t cmask, uint64_t cval, const void* mat, int64_t mat_bstride, int P, float* grad,
                    int64_t grad_stride, int gslot, float* work, int64_t work_floats);
int tqsv_adjoint_diag(tqsv_stream_t s, void* psi, void* lam, int64_t bsz, int n, uint64_t cmask, uint64_t cval,
                      uint64_t pmask, const void* phases, int64_t ph_bstride, int P, float* grad,
                      int64_t grad_stride, int gslot, float* work, int64_t work_floats);
/* Partial-buffer floats any per-gate adjoint step on a [bsz, 2^n] state needs. */
int64_t tqsv_adjoint_work_floats(int64_t bsz, int n);

/* ---- gate matrices from device parameters ------------------------------- */
/* sizeof(DevOp) | sizeof(DevMember) << 16 (layouts checked by the host). */
int tqsv_dev_gate_bytes(void);
/* Write the matrix-pool blocks (U, dU/dtheta_p, U^dagger) of single-wire ops
 * -- one gate or a merged run of single-wire gates (RX, RY, RZ,
 * PhaseShift/U1, U2, U3, Rot with shared [1, P] or per-row [bsz, P]
 * parameters in device memory; fixed H, SH, X, Y, SX, SXDG) -- from their
 * parameters on the GPU.  ops_dev / members_dev: n_ops DevOp records and
 * their DevMember records in device memory; max_rows: the largest row
 * count.  Replaces the per-call matrix build of the reference's functional
 * gates (rx.py:19-46, ry.py:50-67, rz.py:150-171, u3.py:52-77 via
 * gate_wrapper.py:361-425), for trainable Operator parameters and for
 * per-row encoder inputs (encoding/encodings.py:91-102). */
int tqsv_build_dev_gates(tqsv_stream_t s, void* pool, const void* ops_dev, const void* members_dev, int n_ops,
                         int64_t max_rows);

/* ---- expectations ------------------------------------------------------ */
/* bytes of float workspace needed by the calls below (-1: unsupported n). */
int64_t tqsv_expval_workspace_bytes(int64_t bsz, int n, int nterms);
/* <Z_w> for every wire: out[bsz, n].  Replaces expval / MeasureAll with
 * PauliZ (torchquantum/measurement/measurements.py:282-343). */
int tqsv_expval_z_all(tqsv_stream_t s, const void* state, int64_t bsz, int n, float* out, void* workspace,
                      int64_t ws_bytes);
/* <Z_w> for every wire, as tqsv_expval_z_all, allocating nothing and taking
 * no workspace: the chunk partials go to a fixed device array of this
 * library (4 MiB per GPU, held until the process ends; shapes needing more
 * return -3).  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Calls that overlap on
 * one GPU share the array and race: serialise them.  The stream must belong
 * to the current device. */
int tqsv_expval_z_all_noalloc(tqsv_stream_t s, const void* state, int64_t bsz, int n, float* out);
/* Re<psi|P_t|psi> for Pauli strings given by host masks (x: X/Y positions,
 * z: Z/Y positions): out[row * out_stride + t].  Replaces
 * expval_joint_analytical (measurements.py:232-279). */
int tqsv_expval_pauli(tqsv_stream_t s, const void* state, int64_t bsz, int n, int nterms, const uint64_t* xmask,
                      const uint64_t* zmask, float* out, int64_t out_stride, void* workspace, int64_t ws_bytes);
/* out (+)= sum_t coef[row, t] X^x_t Z^z_t psi  (coef: device complex64
 * [bsz, nterms]); the backward of the expectations (torch autograd through
 * measurements.py:271-279 and :299-312). */
int tqsv_pauli_apply(tqsv_stream_t s, void* out, const void* in, int64_t bsz, int n, int nterms,
                     const uint64_t* xmask, const uint64_t* zmask, const void* coef, int accumulate);
/* A Pauli sum with its state gradient in the same passes: out = sum_t
 * coef[t] X^x_t Z^z_t psi (coef: device complex64 [nterms], shared by all
 * rows) and dot[row] = Re <psi_row|out_row> (device float [bsz]).  Replaces
 * the forward and the autograd backward of a weighted sum of
 * expval_joint_analytical values (measurements.py:232-279, their backward
 * :271-279 and :299-312) when the loss is differentiated.  workspace:
 * tqsv_pauli_sum_grad_workspace_floats(bsz, n) floats. */
int tqsv_pauli_sum_grad(tqsv_stream_t s, void* out, const void* in, int64_t bsz, int n, int nterms,
                        const uint64_t* xmask, const uint64_t* zmask, const void* coef, float* dot,
                        float* workspace, int64_t workspace_floats);
int64_t tqsv_pauli_sum_grad_workspace_floats(int64_t bsz, int n);

/* ---- shot sampling: replaces measure()'s host-side random.choices over
 *      |psi|^2 (torchquantum/measurement/measurements.py:37-74), also behind
 *      expval_joint_sampling (:168-229).  out[row * n_shots + k]: in-row index
 *      of shot k (wire 0 = most significant bit); uniforms come from a
 *      counter-based generator keyed by (seed, row, shot).  n <= 32,
 *      bsz * n_shots <= 2^26.  workspace: tqsv_sample_workspace_bytes(). */
int64_t tqsv_sample_workspace_bytes(int64_t bsz, int n);
int tqsv_sample(tqsv_stream_t s, const void* state, int64_t bsz, int n, int64_t n_shots, uint64_t seed,
                int64_t* out, void* workspace, int64_t ws_bytes);

/* Fold matrix gradients (ops flagged F_MG in a stage program: 8 floats
 * sum(lam conj(psi)^T) per op) into parameter gradients:
 * grad[row, slot_p] += Re sum_ab D_p[a][b] conj(Mg[a][b]); items: MgItem array
 * (layout: tqsv_mg_item_bytes()).  Part of the adjoint sweep above. */
int tqsv_mg_item_bytes(void);
int tqsv_mg_to_grad(tqsv_stream_t s, float* grad, int64_t grad_stride, int64_t bsz, const void* items_dev,
                    int n_items, const void* mats);

/* ---- many circuits (bsz 1 each, different widths) in one launch: the work
 *      list holds (circuit, stage program, group range) items (BatchItem,
 *      tqsv_batch_item_bytes()); all items' programs have m <= max_m <= 12.
 *      Replaces a Python loop of per-circuit gate_wrapper calls. ---------- */
int tqsv_batch_item_bytes(void);
int tqsv_batch_forward(tqsv_stream_t s, void* states, const void* items_dev, int n_items, const int* progs_dev,
                       const void* mats, int max_m, int grid);
/* Every level of every circuit in ONE persistent launch: items (the layout
 * of tqsv_batch_queue_item_bytes: a batch item plus its circuit index and
 * the number of that circuit's items in earlier levels) listed level by
 * level; workgroups take them in order from a device counter and wait, at
 * most spin_limit polls, for the circuit's earlier levels.  ctrl: 2 device
 * ints (ctrl[1] != 0 afterwards: a wait ran out and the results are not
 * valid), done: n_circ device ints; both are reset by the call's first
 * kernel.  Replaces, like tqsv_batch_forward, a Python loop of per-circuit
 * gate_wrapper calls (gate_wrapper.py:324-446 per gate of every circuit),
 * without a launch per level. */
int tqsv_batch_queue(tqsv_stream_t s, void* states, const void* items, int n_items, const int* progs,
                     const void* mats, int max_m, int* ctrl, int* done, int n_circ, int64_t spin_limit, int grid);
int tqsv_batch_queue_item_bytes(void);
/* per-circuit Z-string expectation (ZCirc array, tqsv_batch_zcirc_bytes()). */
int tqsv_batch_zcirc_bytes(void);
int tqsv_batch_zexpval(tqsv_stream_t s, const void* states, const void* circ_dev, int n_circ,
                       const int64_t* chunk_circ_dev, int64_t n_chunks, float* part, float* out);

/* ---- learned plan cost model (no reference counterpart: it serves the
 *      fusion planner).  Row-major fp32.
 *      C = alpha op(A) op(B) + beta C (+ bias[N]) (ReLU if relu). ---------- */
int tqsv_gemm(tqsv_stream_t s, int M, int N, int K, float alpha, const float* A, int lda, int ta, const float* B,
              int ldb, int tb, float beta, float* C, int ldc, const float* bias, int relu);
int tqsv_relu_backward(tqsv_stream_t s, float* dY, const float* Y, int64_t n);
int tqsv_colsum(tqsv_stream_t s, const float* X, int M, int N, float* out);
int tqsv_mse_grad(tqsv_stream_t s, const float* pred, const float* y, int M, float* grad, float* loss);
int tqsv_adam_step(tqsv_stream_t s, float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                   float b2, float eps, float c1, float c2);

#ifdef __cplusplus
}
#endif

#endif /* TQSV_H */
