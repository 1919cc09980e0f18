"""Module-level configuration (``src/qandle/config.py:1-75``).

``DEFAULT_MAPPING`` is read when an unnamed rotation is constructed
(``src/qandle/operators.py:168``), exactly as in the reference.
"""
from . import qw_map

DEFAULT_MAPPING = qw_map.tanh

# Text drawings (``src/qandle/config.py:9-75``, read by ``qandle_amd.drawer``).
DRAW_SPLITTED_PAD = 7          # blanks after each sub-circuit diagram of a split circuit
DRAW_DASH = "─"                # the wire character
DRAW_SHOW_VALUES = True        # False drops digits, dots, underscores, blanks and brackets from labels
DRAW_SHIFT_LEFT = False        # True draws gates as far left as their wires allow
DRAW_CROSS_BETWEEN_CNOT = True  # a crossing on the wires between a CNOT's control and target

# Splitter: True carries a rotation's remapping into its sub-circuit, so split and unsplit circuits
# compute the same angles; False rebuilds the gate with DEFAULT_MAPPING as the reference's splitter
# does (src/qandle/splitter/main.py:143-153).
SPLIT_KEEP_REMAPPING = True

# Engine knobs (no counterpart in the reference).
FUSED = True            # Circuit.forward uses the fused pass engine
CHECKPOINT_LAYERS = 5   # adjoint backward restores psi from a checkpoint every k layers
CHECKPOINT_GATES = None  # explicit checkpoint spacing in gates (overrides CHECKPOINT_LAYERS)
AUTOTUNE = False         # time the kernel shape / fusion budget candidates on first use per (circuit, batch)
AUTOTUNE_CACHE = None    # JSON file that keeps tuned choices across processes
SPECIALISE = True        # passes run as kernels compiled at run time for the pass program (hipRTC);
                         # False runs the interpreting pass kernels of the library
SPECIALISE_OPTIONS = {}  # forward kernel options (jit.ProgramKernels): waves, lanes, split_pairs,
                         # interleave, wave_sync, prefetch, grid, pipeline, stagger
ABSORB_LEADING_ROTATIONS = True  # forward-only runs from a product state apply the rotations that precede
                                 # every CNOT on their qubit to that qubit's factor, not to the state
SPECIALISE_ADJOINT_OPTIONS = {}  # adjoint sweep kernel options (jit.AdjointKernels): waves, lanes,
                                 # split_pairs, interleave, wave_sync, prefetch, grid
