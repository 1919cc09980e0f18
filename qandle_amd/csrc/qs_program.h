// Word format of a fused pass program (shared by the host planner and the kernels).
//
// v2 layout (qs_fused_pass / qs_adjoint_pass):
//
//   for each sub-pass:
//     header (56 words): [0] number of segments, [6] number of phase entries,
//       [1, 1+R) register local bits, [8, 8+R) swizzled LDS strides,
//       [16, 16+R) global positions of the register bits, [24, 24+TB) thread-bit
//       local bits, [40, 40+TB) their global positions
//     phase entries, 4 words each: global position, cs index, flags
//       (QS_SEG_PER_SAMPLE; QS_SEG_NOUNAPPLY = gradient probe only), gradient row
//     segments:
//       word: mask of register bits with a unitary | (number of CNOTs << 8)
//       per masked bit, ascending: utab index, flags, nconst, qtab index, then
//         nconst gradient rows (one per fused rotation, application order)
//       per CNOT, 4 words: opcode (QS_OP_CNOT / QS_OP_CNOT_UNIFORM), target
//         register bit, control register bit or global position, 0
//   The forward runs each segment's unitaries then its CNOTs, and the phase
//   entries last; the adjoint encoding lists sub-passes, segments and CNOTs in
//   reverse, and the kernel un-applies phases first, then per segment the CNOTs
//   and the unitaries.
//
// v1 layout (kept for reference in this comment only):
// A pass is a stream of int32 words in device memory:
//
//   for each sub-pass:
//     word 0            : number of ops in this sub-pass
//     words 1 .. R      : the R tile-local bit indices held in registers
//                         (R is the kernel's compile-time register-bit count)
//     then 4 words per op: opcode, a, b, coef
//
// Opcodes (low 8 bits of word 0 of an op; bit 8 = per-sample coefficient,
// i.e. the coefficient index is coef + sample instead of coef):
//   QS_OP_RX, QS_OP_RY, QS_OP_RZ  rotation on register bit a
//   QS_OP_RZ_UNIFORM              RZ on global bit position a that is not a
//                                 register bit: a per-thread phase
//   QS_OP_CNOT                    target register bit a, control register bit b
//   QS_OP_CNOT_UNIFORM            target register bit a, control = global bit
//                                 position b that is not a register bit
//   QS_OP_PAULI_GRAD              adjoint kernel only: reduce Im<lam|P|psi> into
//                                 gradient row b without changing the state;
//                                 the coef word holds the Pauli (0 X, 1 Y, 2 Z)
//                                 plus QS_PAULI_UNIFORM when a is a global
//                                 position that is not a register bit (Z only)
// Coefficients are float2 (cos(phi/2), sin(phi/2)) entries of the `cs` table.
#pragma once

#define QS_OP_RX 0
#define QS_OP_RY 1
#define QS_OP_RZ 2
#define QS_OP_RZ_UNIFORM 3
#define QS_OP_CNOT 4
#define QS_OP_CNOT_UNIFORM 5
#define QS_OP_PAULI_GRAD 6
#define QS_OP_PER_SAMPLE 0x100
#define QS_PAULI_UNIFORM 0x10
#define QS_SEG_PER_SAMPLE 1
#define QS_SEG_NOUNAPPLY 2

#define QS_MAX_LOCAL_BITS 16
#define QS_MAX_QUBITS 40
