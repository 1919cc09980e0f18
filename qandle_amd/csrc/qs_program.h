// Word format of a fused pass program (shared by the host planner and the kernels).
//
// A pass is a stream of int32 words in device memory (qs_fused_pass,
// qs_adjoint_pass and their descriptor forms):
//
//   for each sub-pass:
//     header (56 words): [0] number of segments, [6] number of phase entries,
//       [1, 1+R) register local bits, [8, 8+R) swizzled LDS strides,
//       [16, 16+R) global positions of the register bits, [24, 24+TB) thread-bit
//       local bits, [40, 40+TB) their global positions
//     phase entries, 4 words each: global position, cs index, flags
//       (QS_SEG_PER_SAMPLE; QS_SEG_NOUNAPPLY = gradient probe only), gradient row
//     segments, each starting with a head word:
//       bits 0-7   mask of register bits that carry a fused unitary
//       bits 8-15  number of CNOTs
//       bits 16-23 mask of those unitaries that vary per sample
//       bits 24-30 mask of those that are gradient probes only (adjoint)
//     forward segment: R words, the utab index of every register bit (0, a
//       valid entry, for bits outside the mask), then the CNOT words
//     adjoint segment: the CNOT words, then per masked bit, ascending: utab
//       index, nconst, qtab index, nconst gradient rows (one per fused
//       rotation, in application order)
//     CNOT word: bit 0 = QS_CNOT_UNIFORM (control outside the registers),
//       bits 8-15 target register bit, bits 16+ control register bit or, when
//       uniform, the control's global bit position
//
// The forward runs each segment's unitaries then its CNOTs, and the phase
// entries last.  The adjoint encoding lists sub-passes, segments and CNOTs in
// reverse; the kernel un-applies the phases first, then per segment the
// CNOTs, then reads the gradients of and un-applies the unitaries.
//
// The fixed-size forward record lets a kernel issue every table read of a
// segment as soon as the head word arrives; the unitaries themselves are
// 2x2 complex matrices (8 floats) composed per call by qs_compose_tables.
#pragma once

#define QS_CNOT_UNIFORM 1
#define QS_SEG_PER_SAMPLE 1
#define QS_SEG_NOUNAPPLY 2

#define QS_MAX_LOCAL_BITS 16
#define QS_MAX_QUBITS 40
