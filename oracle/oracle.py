"""ctypes binding of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker and never as the thing
measured or shipped.  The C sources next to this file restate the reference's
mathematics; see oracle.h for the reference file:line map.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Potential(C.Structure):
    _fields_ = [
        ("potential_type", C.c_int),
        ("num_neurons", C.c_int),
        ("num_variables", C.c_int),
        ("cutoff", C.c_float),
        ("weight_force", C.c_float),
        ("weight_energy", C.c_float),
        ("weight_stress", C.c_float),
        ("pmin", C.POINTER(C.c_float)),
        ("pmax", C.POINTER(C.c_float)),
    ]


class _Train(C.Structure):
    _fields_ = [
        ("Nc", C.c_int), ("Nc_force", C.c_int), ("N", C.c_int), ("N_force", C.c_int),
        ("max_Na", C.c_int),
        ("Na", C.POINTER(C.c_int)), ("Na_sum", C.POINTER(C.c_int)), ("type", C.POINTER(C.c_int)),
        ("h", C.POINTER(C.c_float)), ("r", C.POINTER(C.c_float)),
        ("pe_ref", C.POINTER(C.c_float)), ("virial_ref", C.POINTER(C.c_float)),
        ("force_ref", C.POINTER(C.c_float)),
        ("energy_minimum", C.c_double), ("force_square_sum", C.c_double),
        ("potential_square_sum", C.c_double), ("virial_square_sum", C.c_double),
    ]


class _Neighbor(C.Structure):
    _fields_ = [("N", C.c_int), ("NN", C.POINTER(C.c_int)), ("start", C.POINTER(C.c_int)),
                ("NL", C.POINTER(C.c_int)), ("total", C.c_int)]


class _GAParams(C.Structure):
    _fields_ = [("maximum_generation", C.c_int), ("population_size", C.c_int),
                ("parent_number", C.c_int), ("mutation_rate", C.c_float)]


class _MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("idx", C.c_int)]


class _GA(C.Structure):
    _fields_ = [("p", _GAParams), ("V", C.c_int), ("child_number", C.c_int),
                ("population", C.POINTER(C.c_float)), ("population_copy", C.POINTER(C.c_float)),
                ("fitness", C.POINTER(C.c_float)), ("index", C.POINTER(C.c_int)),
                ("cumulative_probabilities", C.POINTER(C.c_float)), ("rng", _MT)]


def build():
    """Compile liboracle.so from the C sources next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER
        L.orc_last_error.restype = C.c_char_p
        L.orc_read_potential.argtypes = [C.c_char_p, P(_Potential)]
        L.orc_read_train.argtypes = [C.c_char_p, P(_Train)]
        L.orc_read_ga.argtypes = [C.c_char_p, P(_GAParams)]
        L.orc_make_train.argtypes = [C.c_int, C.c_int, P(C.c_int), P(C.c_int), P(C.c_float),
                                     P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float),
                                     P(_Train)]
        L.orc_build_neighbor.argtypes = [P(_Train), C.c_float, P(_Neighbor)]
        L.orc_find_force.argtypes = [P(_Potential), P(C.c_float), P(_Train), P(_Neighbor),
                                     P(C.c_double), P(C.c_double), P(C.c_double)]
        L.orc_error_terms.argtypes = [P(_Potential), P(_Train), P(C.c_double), P(C.c_double),
                                      P(C.c_double), P(C.c_double)]
        L.orc_error_terms.restype = C.c_double
        L.orc_fitness.argtypes = [P(_Potential), P(_Train), P(_Neighbor), C.c_int, P(C.c_float),
                                  P(C.c_double), P(C.c_double), C.c_int]
        L.orc_denormalise.argtypes = [P(_Potential), P(C.c_float), P(C.c_float)]
        L.orc_box_inverse.argtypes = [P(C.c_float)]
        L.orc_mt_seed.argtypes = [P(_MT), C.c_uint32]
        L.orc_mt_next.argtypes = [P(_MT)]
        L.orc_mt_next.restype = C.c_uint32
        L.orc_canonical_float.argtypes = [P(_MT)]
        L.orc_canonical_float.restype = C.c_float
        L.orc_uniform_int.argtypes = [P(_MT), C.c_int, C.c_int]
        L.orc_ga_init.argtypes = [P(_GA), P(_GAParams), C.c_int, C.c_uint32]
        for name in ("orc_ga_free", "orc_ga_sort", "orc_ga_crossover"):
            getattr(L, name).argtypes = [P(_GA)]
        L.orc_ga_mutation.argtypes = [P(_GA), C.c_int]
        L.orc_ga_step.argtypes = [P(_GA), C.c_int]
        L.orc_free_potential.argtypes = [P(_Potential)]
        L.orc_free_train.argtypes = [P(_Train)]
        L.orc_free_neighbor.argtypes = [P(_Neighbor)]
        _LIB = L
    return _LIB


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _err():
    return lib().orc_last_error().decode()


class Case:
    """A potential.in + train.in pair held by the oracle, with its neighbour list."""

    def __init__(self, directory=None, *, potential=None, arrays=None, cutoff=None):
        L = lib()
        self.pot = _Potential()
        self.ts = _Train()
        self.nb = _Neighbor()
        self._owned = []
        if directory is not None:
            if L.orc_read_potential(directory.encode(), C.byref(self.pot)):
                raise RuntimeError(_err())
            if L.orc_read_train(directory.encode(), C.byref(self.ts)):
                raise RuntimeError(_err())
        else:
            self._set_potential(**potential)
            a = arrays
            Na = np.ascontiguousarray(a["Na"], np.int32)
            typ = np.ascontiguousarray(a["type"], np.int32)
            box = np.ascontiguousarray(a["box"], np.float32)
            r = np.ascontiguousarray(a["r"], np.float32)
            fref = np.ascontiguousarray(a.get("force_ref", np.zeros_like(r)), np.float32)
            pref = np.ascontiguousarray(a.get("pe_ref", np.zeros(len(Na))), np.float32)
            vref = np.ascontiguousarray(a.get("virial_ref", np.zeros(6 * len(Na))), np.float32)
            if L.orc_make_train(len(Na), int(a["Nc_force"]), _ptr(Na, C.c_int),
                                _ptr(typ, C.c_int), _ptr(box, C.c_float), _ptr(r, C.c_float),
                                _ptr(pref, C.c_float), _ptr(vref, C.c_float),
                                _ptr(fref, C.c_float), C.byref(self.ts)):
                raise RuntimeError(_err())
        rc = self.pot.cutoff if cutoff is None else cutoff
        L.orc_build_neighbor(C.byref(self.ts), C.c_float(rc), C.byref(self.nb))

    def _set_potential(self, potential_type, cutoff, weights, pmin, pmax, num_neurons=0):
        self.pot.potential_type = potential_type
        self.pot.num_neurons = num_neurons
        self.pot.num_variables = len(pmin)
        self.pot.cutoff = cutoff
        self.pot.weight_force, self.pot.weight_energy, self.pot.weight_stress = weights
        self._pmin = np.ascontiguousarray(pmin, np.float32)
        self._pmax = np.ascontiguousarray(pmax, np.float32)
        self.pot.pmin = _ptr(self._pmin, C.c_float)
        self.pot.pmax = _ptr(self._pmax, C.c_float)

    # --------------------------------------------------------------- views
    @property
    def N(self):
        return self.ts.N

    @property
    def Nc(self):
        return self.ts.Nc

    @property
    def V(self):
        return self.pot.num_variables

    def array(self, name, n, ct=np.float32):
        p = getattr(self.ts, name)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def neighbors(self):
        N = self.ts.N
        NN = np.ctypeslib.as_array(self.nb.NN, shape=(N,)).copy()
        start = np.ctypeslib.as_array(self.nb.start, shape=(N,)).copy()
        NL = np.ctypeslib.as_array(self.nb.NL, shape=(max(self.nb.total, 1),)).copy()
        return NN, start, NL[: self.nb.total]

    def bounds(self):
        V = self.V
        return (np.ctypeslib.as_array(self.pot.pmin, shape=(V,)).copy(),
                np.ctypeslib.as_array(self.pot.pmax, shape=(V,)).copy())

    # ------------------------------------------------------------- compute
    def find_force(self, params):
        N = self.ts.N
        p = np.ascontiguousarray(params, np.float32)
        f = np.zeros(3 * N)
        w = np.zeros(6 * N)
        e = np.zeros(N)
        if lib().orc_find_force(C.byref(self.pot), _ptr(p, C.c_float), C.byref(self.ts),
                                C.byref(self.nb), _ptr(f, C.c_double), _ptr(w, C.c_double),
                                _ptr(e, C.c_double)):
            raise RuntimeError("bad potential type")
        return f, w, e

    def error_terms(self, f, w, e):
        t = np.zeros(3)
        fit = lib().orc_error_terms(C.byref(self.pot), C.byref(self.ts), _ptr(f, C.c_double),
                                    _ptr(w, C.c_double), _ptr(e, C.c_double), _ptr(t, C.c_double))
        return fit, t

    def denormalise(self, genes):
        g = np.ascontiguousarray(genes, np.float32)
        out = np.zeros(self.V, np.float32)
        lib().orc_denormalise(C.byref(self.pot), _ptr(g, C.c_float), _ptr(out, C.c_float))
        return out

    def fitness(self, genes, threads=1):
        g = np.ascontiguousarray(genes, np.float32).reshape(-1, self.V)
        P = g.shape[0]
        fit = np.zeros(P)
        terms = np.zeros(3 * P)
        lib().orc_fitness(C.byref(self.pot), C.byref(self.ts), C.byref(self.nb), P,
                          _ptr(g, C.c_float), _ptr(fit, C.c_double), _ptr(terms, C.c_double),
                          threads)
        return fit, terms.reshape(P, 3)

    def fitness_of_params(self, params):
        f, w, e = self.find_force(params)
        return self.error_terms(f, w, e)

    def close(self):
        L = lib()
        L.orc_free_neighbor(C.byref(self.nb))
        L.orc_free_train(C.byref(self.ts))
        if not hasattr(self, "_pmin"):
            L.orc_free_potential(C.byref(self.pot))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MT19937:
    def __init__(self, seed):
        self.s = _MT()
        lib().orc_mt_seed(C.byref(self.s), seed)

    def next(self):
        return lib().orc_mt_next(C.byref(self.s))

    def canonical(self):
        return lib().orc_canonical_float(C.byref(self.s))

    def uniform_int(self, a, b):
        return lib().orc_uniform_int(C.byref(self.s), a, b)


class GA:
    """Oracle GA state (population, fitness, rng) with the reference's steps."""

    def __init__(self, P, parents, V, seed, generations=1000, mutation_rate=0.2):
        self.g = _GA()
        p = _GAParams(generations, P, parents, mutation_rate)
        if lib().orc_ga_init(C.byref(self.g), C.byref(p), V, seed):
            raise ValueError("bad GA parameters")
        self.P, self.V = P, V

    @property
    def population(self):
        return np.ctypeslib.as_array(self.g.population, shape=(self.P * self.V,)).reshape(
            self.P, self.V)

    @property
    def fitness(self):
        return np.ctypeslib.as_array(self.g.fitness, shape=(self.P,))

    def step(self, generation, fitness):
        self.fitness[:] = np.asarray(fitness, np.float32)
        lib().orc_ga_step(C.byref(self.g), generation)

    def __del__(self):
        try:
            lib().orc_ga_free(C.byref(self.g))
        except Exception:
            pass
