"""Example grammars (dpx.examples) plus the oracle helpers of the tests."""
from dpx.examples import *  # noqa: F401,F403
from dpx.examples import _wrap  # noqa: F401


def oracle(g):
    import adp_oracle
    return adp_oracle.Oracle(g.spec(), {n: f.py for n, f in g.functions().items()})


def run_oracle(g, in1, in2=None, trace=True):
    o = oracle(g)
    g._in1 = g.py_input(in1)
    g._in2 = g.py_input(in2) if in2 is not None else None
    return o, o.run(g._in1, g._in2, trace=trace)
