"""A circuit stays an ordinary nn.Module after the engine has lowered it:
deepcopy, pickle and torch.save work (best-model and EMA snapshots), and a
copy lowers its own program bound to its own parameters."""
import copy
import ctypes
import io
import pickle

import torch

import qandle_amd as qandle
from qandle_amd import workloads
from qandle_amd._engine import fused


def _lowered_circuit():
    w = workloads.rxryrz_ladder(n=6, B=3, depth=2, seed=0)
    circ = qandle.Circuit(w.layers, num_qubits=6)
    prog = fused.lower_circuit(circ.circuit)
    assert prog is not None
    # what a GPU run leaves in the program's cache: compiled-kernel handles are ctypes pointers
    prog.words_cache[("jit", "planted")] = ctypes.c_void_p(0x1234)
    return circ, prog


def test_deepcopy_and_pickle_after_lowering():
    circ, prog = _lowered_circuit()
    dup = copy.deepcopy(circ)
    back = pickle.loads(pickle.dumps(circ))
    buf = io.BytesIO()
    torch.save(circ, buf)
    buf.seek(0)
    loaded = torch.load(buf, weights_only=False)  # a file this test wrote itself
    for other in (dup, back, loaded):
        sd0, sd1 = circ.state_dict(), other.state_dict()
        assert sd0.keys() == sd1.keys()
        assert all(torch.equal(sd0[k], sd1[k]) for k in sd0)
    # the original keeps its program; a copy lowers a new one over its own modules
    assert fused.lower_circuit(circ.circuit) is prog
    p2 = fused.lower_circuit(dup.circuit)
    assert p2 is not prog and ("jit", "planted") not in p2.words_cache
    own = {id(m) for m in dup.modules()}
    assert all(prog.slots[i].module is not None and id(p2.slots[i].module) in own for i in p2.param_slots)


def test_program_cache_does_not_keep_circuits_alive():
    import gc
    import weakref

    circ, _ = _lowered_circuit()
    ref = weakref.ref(circ.circuit)
    before = len(fused._PROGRAMS)
    del circ
    gc.collect()
    assert ref() is None and len(fused._PROGRAMS) == before - 1
