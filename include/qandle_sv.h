/*
 * qandle_sv.h -- C-ABI of the MI355X (gfx950) state-vector engine behind the
 * qandle_amd drop-in.  One shared library, libqandle_sv.so, exports these.
 *
 * Conventions for every entry point
 *   - Amplitudes are complex64 stored as interleaved float pairs (re, im).
 *     A state batch is a dense (B, 2^n) row-major array: sample b, amplitude
 *     index i at element b * 2^n + i.  qandle qubit q is amplitude bit n-1-q;
 *     every entry point here speaks of BIT POSITIONS, the Python host converts.
 *   - All pointers are device pointers unless a parameter says "host".
 *   - `stream` is a hipStream_t passed as void*; work is enqueued on it and the
 *     call returns without synchronising.
 *   - Return value: QS_OK (0) or a QS_E* code; no exception crosses the ABI.
 *     qs_last_hip_error() gives the HIP error behind QS_EHIP.
 *   - No entry point allocates device memory.  Entry points that need scratch
 *     take a caller-provided `work` buffer whose size is given by the matching
 *     *_work_bytes query (qs_measure_prob_noalloc uses a static device array).
 *
 * Reference citations are file:line into gstenzel/qandle v0.1.8.
 */
#ifndef QANDLE_SV_H
#define QANDLE_SV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  QS_OK = 0,
  QS_EINVAL = 1,       /* bad argument (shape, bit position, null pointer) */
  QS_EHIP = 2,         /* a HIP runtime call failed, see qs_last_hip_error() */
  QS_EWORKSPACE = 3,   /* caller workspace too small */
  QS_EUNSUPPORTED = 4  /* valid request this build does not handle */
};

/* Rotation kinds and Pauli generators: R_P(phi) = exp(-i phi P / 2). */
enum { QS_RX = 0, QS_RY = 1, QS_RZ = 2 };

int qs_abi_version(void);
int qs_last_hip_error(void);
const char *qs_status_string(int status);

/* ---------------------------------------------------------------------------
 * Gate coefficients.  cs[2k] = cos(phi[k]/2), cs[2k+1] = sin(phi[k]/2).
 * Replaces the per-call matrix rebuild `_a*a_op(t) + _b*b_op(t)`, t = phi/2
 * (src/qandle/operators.py:265-275): only the two scalars are ever needed.
 * ------------------------------------------------------------------------- */
int qs_half_angle_coeffs(float *cs, const float *phi, int64_t count, void *stream);

/* ---------------------------------------------------------------------------
 * Plain per-gate path (one HBM pass per gate).
 * ------------------------------------------------------------------------- */

/* AngleEmbedding: out[b] = (x)_j R_rot(x[b, j]) |0..0> over the embedded bit
 * positions bits[0..d) (host array); unembedded bits stay |0>.
 * x is (B, d) float32 with row stride x_stride (0 = one row shared by all B).
 * Replaces AngleEmbeddingBuilt.forward (src/qandle/embeddings.py:148-164). */
int qs_embed_angle(float *out, const float *x, int64_t B, int n,
                   const int32_t *bits, int d, int rot, int64_t x_stride,
                   void *stream);

/* One rotation on bit position `bit`: out = U psi (out may equal in).
 * Coefficients for sample b are cs[2*(b*cs_stride)], cs_stride 0 = shared,
 * 1 = one (cos, sin) pair per sample.  inverse != 0 applies U^dagger.
 * Replaces BuiltParametrizedOperator.forward `state @ M` / the batched bmm
 * (src/qandle/operators.py:289-298). */
int qs_rot(float *out, const float *in, int64_t B, int n, int bit, int kind,
           const float *cs, int64_t cs_stride, int inverse, void *stream);

/* CNOT as an exact permutation: amplitudes i and i ^ (1<<tbit) exchange when
 * bit cbit of i is set.  out may equal in (in-place swap).
 * Replaces BuiltCNOT.forward `state @ _M` (src/qandle/operators.py:547-558). */
int qs_cnot(float *out, const float *in, int64_t B, int n, int cbit, int tbit,
            void *stream);

/* MeasureProbability: out[b, q] = sum_{i: bit n-1-q of i = 0} |psi_b[i]|^2,
 * float32 (B, n).  Replaces MeasureProbabilityBuilt.forward
 * (src/qandle/measurements.py:120-123).  Deterministic two-stage reduction. */
int64_t qs_measure_prob_work_bytes(int64_t B, int n);
int qs_measure_prob(float *out, const float *psi, int64_t B, int n, void *work,
                    int64_t work_bytes, void *stream);

/* qs_measure_prob without a workspace argument, for callers that cannot
 * allocate (code under HIP graph capture, serving loops and harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The stage-1 partials
 * live in a file-scope device array of QS_NOALLOC_SCRATCH_BYTES, one per
 * device for the life of the process; the call allocates nothing.  Requests
 * whose partials exceed it (qs_measure_prob_work_bytes(B, n) > 16 MiB, i.e.
 * more than about 2^30 amplitudes) return QS_EWORKSPACE.  Every call on one
 * device shares that array: calls that overlap on one device (other streams,
 * other host threads) race and must be ordered by the caller. */
#define QS_NOALLOC_SCRATCH_BYTES (16 << 20)
int qs_measure_prob_noalloc(float *out, const float *psi, int64_t B, int n,
                            void *stream);

/* Backward of MeasureProbability: lam[b, i] = 2 psi[b, i] * sum_q w[b, q] [bit n-1-q of i = 0]
 * (torch's dL/dRe + i dL/dIm for L = sum w * P).  w is (B, n) float32 with
 * row stride w_stride (0 = shared). */
int qs_measure_prob_backward(float *lam, const float *psi, const float *w,
                             int64_t w_stride, int64_t B, int n, void *stream);

/* out[b * d + j] = Im <lam_b | P_j | psi_b> for the Pauli `pauli` on bit
 * position bits[j] (host array, d <= 64).  Deterministic; used for the
 * rotation and embedding gradients of the plain path, where
 * dL/dphi = 1/2 Im <lam | P | psi> at the gate output. */
int64_t qs_pauli_im_work_bytes(int64_t B, int n, int d);
int qs_pauli_im(float *out, const float *lam, const float *psi, int64_t B,
                int n, const int32_t *bits, int d, int pauli, void *work,
                int64_t work_bytes, void *stream);

/* qs_pauli_im restricted to the amplitudes whose bits cmask are all 1 (the
 * generator P (x) |1..1><1..1| of a controlled rotation; cmask must not hold
 * a bit of `bits`).  Gradient of a Controlled rotation
 * (src/qandle/operators.py:488-506) in the plain path. */
int qs_pauli_im_controlled(float *out, const float *lam, const float *psi,
                           int64_t B, int n, const int32_t *bits, int d,
                           int pauli, int64_t cmask, void *work,
                           int64_t work_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Remaining operators, plain per-gate path.  `cmask` is a set of bit
 * positions (bit p of the mask = position p) that must all be 1 for the gate
 * to act; 0 = uncontrolled.  A 2x2 unitary `u` is 8 floats, row-major complex
 * (u00, u01, u10, u11) in the column convention psi' = U psi, i.e. the
 * transpose of the reference's `state @ M` block.
 * ------------------------------------------------------------------------- */

/* psi' = U psi on bit position `bit` where every bit of cmask is set (out may
 * equal in).  u_stride 0: one U for the batch; 1: one U per sample (8 floats
 * each).  dagger != 0 applies U^dagger.  Replaces BuiltU.forward
 * (src/qandle/operators.py:125-126), BuiltControlled.forward (:488-506) and
 * BuiltInvert.forward (:445-447) for one-qubit targets. */
int qs_unitary_1q(float *out, const float *in, int64_t B, int n, int bit,
                  const float *u, int64_t u_stride, int64_t cmask, int dagger,
                  void *stream);

/* u[8k..8k+8) = R_kind(phi_k) from cs (cos, sin of phi_k / 2), the matrices of
 * src/qandle/operators.py:368-395 in the column convention. */
int qs_rot_unitaries(float *u, const float *cs, int64_t count, int kind,
                     void *stream);

/* Multi-controlled X: out[i] = in[i ^ (1 << tbit)] where every bit of cmask is
 * set in i, else in[i] (out != in).  CNOT under a Controlled
 * (src/qandle/operators.py:488-506, 557-558); exact permutation. */
int qs_mcx(float *out, const float *in, int64_t B, int n, int tbit,
           int64_t cmask, void *stream);

/* out[i] = -in[i] where every bit of mask is set in i (out may equal in).
 * CZ is mask = {control, target}; replaces BuiltCZ.forward
 * (src/qandle/operators.py:581-591). */
int qs_phase_flip(float *out, const float *in, int64_t B, int n, int64_t mask,
                  void *stream);

/* out[i] = in[i with bits abit and bbit exchanged] where every bit of cmask is
 * set in i (out != in).  Replaces BuiltSWAP.forward
 * (src/qandle/operators.py:662-674). */
int qs_swap(float *out, const float *in, int64_t B, int n, int abit, int bbit,
            int64_t cmask, void *stream);

/* Reset of bit position `bit`, per amplitude pair (a0: bit 0, a1: bit 1):
 * a0' = a0 * sqrt(|a0|^2 + |a1|^2) / (|a0| + 1e-7), a1' = 0 (out may equal
 * in).  Replaces BuiltReset.forward (src/qandle/operators.py:612-623). */
int qs_reset(float *out, const float *in, int64_t B, int n, int bit,
             void *stream);

/* Gradient of qs_reset for a real loss: gin from g = dL/dRe + i dL/dIm of
 * the output and the input `in` of the forward (gin != in). */
int qs_reset_backward(float *gin, const float *g, const float *in, int64_t B,
                      int n, int bit, void *stream);

/* AmplitudeEmbedding: out[b] (2^n complex64) = the first 2^n entries of
 * (x[b, 0..d), pad, pad, ...) as real amplitudes, divided by
 * max(||.||_2, 1e-12) when normalize != 0.  x is (B, d) float32 with row
 * stride x_stride.  Replaces AmplitudeEmbeddingBuilt.forward
 * (src/qandle/embeddings.py:50-61: F.pad with `pad_with`, F.normalize,
 * torch.complex).  The backward writes gx (B, d) from g = dL/dRe + i dL/dIm;
 * entries cropped away (d > 2^n) get 0. */
int qs_amplitude_embed(float *out, const float *x, int64_t B, int d, int n,
                       int64_t x_stride, float pad, int normalize,
                       void *stream);
int qs_amplitude_embed_backward(float *gx, const float *g, const float *x,
                                int64_t B, int d, int n, int64_t x_stride,
                                float pad, int normalize, void *stream);

/* MeasureJointProbability: out[i] = |psi[i]|^2 over `count` amplitudes, and
 * its gradient gpsi = 2 g psi.  Replaces MeasureAllAbsolute.forward
 * (src/qandle/measurements.py:78-79). */
int qs_abs2(float *out, const float *psi, int64_t count, void *stream);
int qs_abs2_backward(float *gpsi, const float *g, const float *psi,
                     int64_t count, void *stream);

/* ---------------------------------------------------------------------------
 * Fused pass path.  A pass loads a tile of 2^K amplitudes (K = R + TB local
 * bit positions, lpos[] ascending with lpos[0] = 0) into LDS once, applies a
 * run of gates in sub-passes that each hold R positions in registers, and
 * stores it back: a whole run of gates costs one HBM round trip.  The int32
 * `words` stream (device memory) encodes the sub-passes; its layout is in
 * qandle_amd/csrc/qs_program.h.  Together these replace the executor loop
 * UnsplittedCircuit.forward (src/qandle/qcircuit.py:163-174), the split
 * executor (src/qandle/qcircuit.py:246-295) and the per-gate forwards they
 * call (src/qandle/operators.py:289-298, 557-558).
 * ------------------------------------------------------------------------- */

/* 1 when a (register bits R, thread bits TB) kernel variant is compiled in. */
int qs_fused_supported(int R, int TB);

/* 1 when the measurement and gradient partial layouts of a (R, TB) tile are
 * defined (the run-time specialised kernels also use shapes that no compiled
 * variant has): 1 <= R <= 5, 6 <= TB <= 10, R + TB <= 16.  The work-bytes and
 * finish entry points below accept these shapes. */
int qs_layout_supported(int R, int TB);

/* Product-state factors efac[b, pos, v] (float2) of an AngleEmbedding over bit
 * positions bits[0..d) (host) with angles x (B, d), row stride x_stride;
 * positions that are not embedded get (1, 0) / (0, 0) (|0>).  d = 0 gives
 * |0..0>.  Replaces the embedding's einsum + matmul chain
 * (src/qandle/embeddings.py:148-164): the first pass multiplies these out. */
int qs_product_factors(float *efac, const float *x, int64_t B, int n,
                       const int32_t *bits, int d, int rot, int64_t x_stride,
                       void *stream);

/* Product-state factors with single-qubit unitaries applied:
 * out[b, pos] = U efac[b, pos] for every position with unit[pos] >= 0 (host
 * arrays of n entries), U the 2x2 utab entry unit[pos] (+ b when span[pos] is
 * B; span 1 is one entry for the batch), other positions copied.  utab has nu
 * entries (qs_compose_tables).  A circuit's rotations that precede every CNOT
 * on their qubit act on an unentangled qubit, so they are applied here to its
 * factor instead of to the 2^n amplitudes (the rotations of
 * src/qandle/operators.py:215-305 that follow src/qandle/embeddings.py:148-164
 * directly); no counterpart in the reference.  Returns QS_EINVAL when an entry
 * lies outside the table. */
int qs_factor_apply(float *out, const float *efac, const float *utab, int64_t nu,
                    const int32_t *unit, const int32_t *span, int64_t B, int n,
                    void *stream);

/* Fused per-bit unitaries and gradient generators for a program's words:
 * `recipes` (device int32, record r at offsets[r]) lists, per table entry,
 * u_index, span (batch size of a per-sample entry, else 1), nconst, q_index
 * and per fused rotation (kind, cs index, cs per-sample flag) in application
 * order.  For b < span writes utab[u_index + b] = R_m ... R_1 and
 * qtab[q_index + k*span + b] = V_k P_k V_k^dagger with
 * V_k = R_m ... R_{k+1} (8 floats each; qtab = NULL skips the generators,
 * which only the adjoint kernels read).  Replaces the per-call rebuild of
 * every gate's dense matrix (src/qandle/operators.py:265-275). */
int qs_compose_tables(float *utab, float *qtab, const int32_t *recipes,
                      const int32_t *offsets, int nrec, const float *cs,
                      void *stream);

/* Matrix-core operands of composed unitaries (utab, 8 floats each): entry e
 * (16 floats, ltab[16 e + 4 r + k]) is what lane row r = lane & 3 of a 4-lane
 * block supplies to v_mfma_f32_4x4x1f32 for the real 4x4 form of U = utab
 * entry idx[e] (e when idx = NULL; idx is device int32), or of U^dagger when
 * dagger != 0.  The run-time specialised pass kernels copy their pass's
 * entries into LDS and read each unitary's operands with one LDS load.  Part
 * of the same per-call table build (src/qandle/operators.py:265-275). */
int qs_lane_coeffs(float *ltab, const float *utab, const int32_t *idx,
                   int64_t count, int dagger, void *stream);

/* One pass.  src = NULL generates the product state of efac in registers;
 * otherwise tiles are read from src (src may equal state).  cs is the float2
 * (cos, sin) table of the phase entries and utab the fused unitaries that the
 * words index.  mpart != NULL folds
 * MeasureProbability partials into the pass (size from
 * qs_fused_measure_work_bytes).  store = 0 skips writing the state back. */
int qs_fused_pass(float *state, const float *src, int64_t B, int n, int R,
                  int TB, const int32_t *lpos, const int32_t *words, int nsub,
                  const float *cs, const float *utab, const float *efac,
                  float *mpart, int store, void *stream);

/* MeasureProbability from a measuring pass's partials: out (B, n) float32,
 * lpos (host) = that pass's local positions.
 * Replaces src/qandle/measurements.py:120-123. */
int64_t qs_fused_measure_work_bytes(int64_t B, int n, int R, int TB);
int qs_fused_measure_finish(float *out, const float *mpart, int64_t B, int n,
                            int R, int TB, const int32_t *lpos, void *stream);

/* MeasureProbability from readout partials: every tile (or, for tiles that
 * hold several samples, every sample of a tile) wrote n + 1 floats, entry p =
 * its share of P(measured bit at position p = 0) and entry n its total; out
 * (B, n) float32 sums the shares of each sample's tiles.  This is the form of
 * the run-time specialised kernels, whose measured bit may be a parity of
 * state bits (CNOTs that end a circuit are folded into its readout).
 * Replaces src/qandle/measurements.py:120-123 together with the trailing
 * CNOT layers of src/qandle/operators.py:557-558. */
int64_t qs_readout_work_bytes(int64_t B, int n, int R, int TB);
int qs_readout_finish(float *out, const float *mpart, int64_t B, int n, int R,
                      int TB, void *stream);

/* ---------------------------------------------------------------------------
 * Adjoint backward.  Replaces torch autograd through every dense gate of the
 * reference (implicit; the tape of src/qandle/operators.py:272-298 and the
 * grad_M = state^H @ g outer products it implies).  One call un-applies one
 * forward pass: `words` is that pass re-encoded in reverse order with each
 * rotation's gradient row; psi and lam (the co-state dL/dpsi) are updated in
 * place, and row r of `gpart` receives Im<lam|P|psi> partials for the gate
 * at its output.  seed_w != NULL (the first call of a sweep) builds the
 * MeasureProbability co-state from psi while loading instead of reading lam:
 * lam = 2 psi * sum_q seed_w[b, q] [bit n-1-q = 0], seed_w (B, n) float32.
 * qs_adjoint_finish turns the partials into
 * out[r, b] = dL/dphi_r for sample b and tot[r] = sum over b.
 * ------------------------------------------------------------------------- */
int64_t qs_adjoint_work_bytes(int64_t B, int n, int R, int TB, int rows);
int qs_adjoint_pass(float *psi, float *lam, int64_t B, int n, int R, int TB,
                    const int32_t *lpos, const int32_t *words, int nsub,
                    const float *cs, const float *utab, const float *qtab,
                    float *gpart, int rows, const float *seed_w, void *stream);
int qs_adjoint_finish(float *out, float *tot, const float *gpart, int rows,
                      int64_t B, int n, int R, int TB, void *stream);

/* ---------------------------------------------------------------------------
 * Work-item form for batches of uneven circuits.  A descriptor is an opaque
 * record (size from the *_desc_bytes query) filled on the host from exactly
 * the arguments of the direct call; the caller copies an array of them to
 * device memory.  `items` (device) holds int32 pairs (descriptor index, tile
 * index); one launch runs every item with a grid-stride loop over `grid`
 * workgroups, so circuits of different widths, depths and programs share one
 * launch and can be captured into a single hipGraph.  The item launches take
 * the tile shapes (R, TB) = (4, 7), (4, 8), (5, 6) and (5, 8) and return
 * QS_EUNSUPPORTED for any other.  Replaces running the reference's
 * per-circuit Python executor (src/qandle/qcircuit.py:163-174) once per
 * circuit.
 * ------------------------------------------------------------------------- */
/* Product-state factors of many circuits in one launch: entry e writes
 * efac[e] = (f(0), f(1)) (two float2) from x[entries[2e]] with rotation
 * entries[2e+1], or |0> when entries[2e] is -1.  The per-circuit equivalent
 * is qs_product_factors. */
int qs_factor_entries(float *efac, const float *x, const int32_t *entries,
                      int64_t count, void *stream);

int64_t qs_pass_desc_bytes(void);
int qs_pass_desc(void *desc, int64_t *ntiles, float *state, const float *src,
                 int64_t B, int n, int R, int TB, const int32_t *lpos,
                 const int32_t *words, int nsub, const float *cs,
                 const float *utab, const float *efac, float *mpart, int store);
int qs_items_pass(const void *descs, const int32_t *items, int nitems, int R,
                  int TB, int grid, void *stream);

int64_t qs_measure_desc_bytes(void);
int qs_measure_desc(void *desc, float *out, const float *mpart, int64_t B,
                    int n, int R, int TB, const int32_t *lpos);
int qs_items_measure_finish(const void *descs, int ndesc, int max_n,
                            void *stream);

int64_t qs_adjoint_desc_bytes(void);
int qs_adjoint_desc(void *desc, int64_t *ntiles, float *psi, float *lam,
                    int64_t B, int n, int R, int TB, const int32_t *lpos,
                    const int32_t *words, int nsub, const float *cs,
                    const float *utab, const float *qtab, float *gpart,
                    int rows, const float *seed_w);
int qs_items_adjoint(const void *descs, const int32_t *items, int nitems,
                     int R, int TB, int grid, void *stream);

int64_t qs_adjoint_finish_desc_bytes(void);
int qs_adjoint_finish_desc(void *desc, float *out, float *tot,
                           const float *gpart, int rows, int64_t B, int n,
                           int R, int TB);
int qs_items_adjoint_finish(const void *descs, int ndesc, int row_blocks,
                            void *stream);

/* ---- run-time specialised pass kernels -----------------------------------
 * Same role as qs_fused_pass (the executor loop of UnsplittedCircuit.forward,
 * src/qandle/qcircuit.py:163-174, and the per-gate `state @ M` it runs,
 * src/qandle/operators.py:289-298, 557-558), with the pass program compiled
 * into the kernel: the host emits the HIP source of one kernel per pass, these
 * calls compile it with hipRTC for the current device, load the code object
 * and launch its kernels.
 *
 * qs_jit_compile compiles `src` with the given compiler options; *handle
 * receives the result (code object and log; release with
 * qs_jit_release_code) even when compilation fails (QS_EHIP: read the log).
 * qs_jit_load loads a code object (from qs_jit_code_copy or a cache) into the
 * current device and qs_jit_function looks up an extern "C" kernel.
 * qs_jit_resident_blocks gives the workgroups of `block` threads that can be
 * resident at once on the current device (the grid of a grid-stride kernel).
 * qs_jit_launch launches `function` with the by-value argument block `args`
 * (the kernel's parameter layout, `args_size` bytes). */
int qs_jit_compile(const char *src, const char *name, const char *const *options,
                   int noptions, void **handle);
int64_t qs_jit_code_size(const void *handle);
int qs_jit_code_copy(const void *handle, void *out, int64_t size);
int64_t qs_jit_log_size(const void *handle);
int qs_jit_log_copy(const void *handle, char *out, int64_t size);
void qs_jit_release_code(void *handle);
int qs_jit_load(const void *code, int64_t size, void **module);
int qs_jit_function(void *module, const char *name, void **function);
int64_t qs_jit_resident_blocks(void *function, int block);
int qs_jit_launch(void *function, unsigned grid, unsigned block, const void *args,
                  int64_t args_size, void *stream);
int qs_jit_unload(void *module);

#ifdef __cplusplus
}
#endif
#endif /* QANDLE_SV_H */
