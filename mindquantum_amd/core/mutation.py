"""Global mutation epoch of circuits, gates and parameter resolvers.

Every change that can alter how a circuit lowers to the native gate list - a list mutation of a
Circuit, an attribute write on a gate, a flag or constant change of a ParameterResolver - advances
one counter.  Native circuits compiled from a Circuit object are reused without looking at the
circuit again while the counter is unchanged; after any change the circuit is lowered again and
its content compared with what the native circuit was built from (mindquantum_amd.ir.NativeCache).
In-place edits of a gate's numpy matrix are not tracked (replace the gate instead).
"""
_EPOCH = [0]


def bump():
    _EPOCH[0] += 1


def epoch():
    return _EPOCH[0]
