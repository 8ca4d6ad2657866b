"""fastfermion_amd -- MI355X implementation of fastfermion's truncated
Heisenberg-picture propagation (Pauli and Majorana), with the same `propagate`
call and operator types as fastfermion (fastfermion/__init__.py:9-11):

    import fastfermion_amd as ff
    obs = ff.propagate(circuit, ff.PauliString("Z10"), maxdegree=6)

Propagation runs on the GPU through libffamd.so (include/fastfermion_amd.h);
there is no CPU fallback.
"""
from ._ffcore import *  # noqa: F401,F403
from ._ffcore import __version__  # noqa: F401
from ._parse import poly  # noqa: F401


def coefficient(p, op):
    """Coefficient of monomial `op` (string, tuple or string object) in `p`."""
    return p.coefficient(op)


def degree(p, v=None):
    """Degree of a polynomial, optionally in one Pauli letter ('X', 'Y', 'Z')."""
    return p.degree() if v is None else p.degree(v)


def dagger(p):
    return p.dagger()


def norm(p, k=1):
    """(sum_i |c_i|^k)^(1/k) for k in {0, 1, 2}; k='inf' gives max |c_i|."""
    return p.norm(k)


def extent(p):
    return p.extent()


def truncate(p, maxdegree):
    return p.truncate(maxdegree)


def compress(p, threshold=0):
    return p.compress(threshold)


def commutes(a, b):
    return a.commutes(b)


def commutator(a, b):
    return a.commutator(b)
