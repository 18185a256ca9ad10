"""gpuga_amd: GPUGA's fitness-evaluation path on AMD Instinct MI355X (gfx950).

Thin ctypes layer over libgpuga.so (the C ABI declared in include/gpuga.h),
mirroring the reference's host interface:

    Fitness(input_dir)            class Fitness, src/fitness.cuh:30-40
      .compute(genes) -> fitness  Fitness::compute, src/fitness.cu:316-333
      .predict(dir, elite)        Fitness::predict, src/fitness.cu:347-390
      .number_of_variables, .parameters_min, .parameters_max
      .find_force(params)         Potential::find_force, src/potential.cuh:29-40
    GA(...)                       class GA, src/ga.cu:27-212
    run(input_dir)                GPUGA::GPUGA, src/gpuga.cu:24-28
    Server(...)                   concurrent fitting jobs (no reference counterpart)

The library is built in-tree (gpuga_amd/libgpuga.so) by `make -C gpuga_amd`.
There is no CPU fallback: if the library is missing or no HIP device is
present, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpuga.so")
_LIB = None

POTENTIAL_RI, POTENTIAL_TERSOFF1, POTENTIAL_TERSOFF2, POTENTIAL_NN2B = 0, 1, 2, 3


class GpugaError(RuntimeError):
    pass


class Options(C.Structure):
    _fields_ = [("backend", C.c_int), ("chunk_cost", C.c_int), ("use_graph", C.c_int),
                ("nn2b_bf16", C.c_int), ("nn2b_grid", C.c_int), ("verbose", C.c_int),
                ("blocks_per_cu", C.c_int), ("pack_lanes", C.c_int), ("queues", C.c_int),
                ("ga_batch", C.c_int), ("tail_split", C.c_int),
                ("serve_share", C.c_int)]


class ConvertOptions(C.Structure):
    _fields_ = [("cells", C.c_int * 3), ("delta_e", C.c_float), ("stress_in_kbar", C.c_int)]


def build(jobs: int = 8) -> None:
    """Compile libgpuga.so and the gpuga CLI for gfx950 (hipcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE, f"-j{jobs}"], check=True)


def lib():
    """Load libgpuga.so; raises GpugaError when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise GpugaError(f"{LIB_PATH} is missing: build it with `make -C gpuga_amd`")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        fp, ip, dp = P(C.c_float), P(C.c_int), P(C.c_double)
        vp = C.c_void_p
        sig = {
            "gpuga_last_error": (C.c_char_p, []),
            "gpuga_default_options": (None, [P(Options)]),
            "gpuga_fitness_create": (C.c_int, [C.c_char_p, P(Options), P(vp)]),
            "gpuga_fitness_create_arrays": (C.c_int, [C.c_int, C.c_int, C.c_float, C.c_float,
                                                      C.c_float, C.c_float, fp, fp, C.c_int,
                                                      C.c_int, ip, ip, fp, fp, fp, fp, fp,
                                                      P(Options), P(vp)]),
            "gpuga_fitness_destroy": (None, [vp]),
            "gpuga_fitness_num_variables": (C.c_int, [vp]),
            "gpuga_fitness_bounds": (C.c_int, [vp, fp, fp]),
            "gpuga_fitness_info": (C.c_int, [vp, ip]),
            "gpuga_fitness_compute": (C.c_int, [vp, C.c_int, fp, fp]),
            "gpuga_fitness_compute_terms": (C.c_int, [vp, C.c_int, fp, fp, dp]),
            "gpuga_fitness_compute_plain": (C.c_int, [vp, C.c_int, fp, fp, dp]),
            "gpuga_fitness_compute_plain_noalloc": (C.c_int, [vp, C.c_int, fp, fp, dp]),
            "gpuga_fitness_predict": (C.c_int, [vp, C.c_char_p, fp]),
            "gpuga_fitness_find_force": (C.c_int, [vp, fp, fp, fp, fp]),
            "gpuga_fitness_neighbors": (C.c_int, [vp, ip, ip, ip]),
            "gpuga_fitness_autotune": (C.c_int, [vp, C.c_int, ip]),
            "gpuga_fitness_launch_shape": (C.c_int, [vp, C.c_int, ip, ip, ip, ip]),
            "gpuga_inspect_neighbors": (C.c_int, [C.c_char_p, ip, ip, ip, ip]),
            "gpuga_fitness_set_timing": (C.c_int, [vp, C.c_int]),
            "gpuga_fitness_timing": (C.c_int, [vp, dp, P(C.c_long), C.c_int]),
            "gpuga_fitness_device_bytes": (C.c_int, [vp, dp]),
            "gpuga_fitness_set_wait_limit": (C.c_int, [vp, C.c_int]),
            "gpuga_fitness_trace": (C.c_int, [vp, C.c_int, fp, dp, ip]),
            "gpuga_ga_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                          C.c_uint, P(vp)]),
            "gpuga_ga_destroy": (None, [vp]),
            "gpuga_ga_get_population": (C.c_int, [vp, fp]),
            "gpuga_ga_set_population": (C.c_int, [vp, fp]),
            "gpuga_ga_step": (C.c_int, [vp, C.c_int, fp, fp]),
            "gpuga_ga_step_device": (C.c_int, [vp, C.c_int, fp, fp]),
            "gpuga_fitness_run_ga": (C.c_int, [vp, vp, C.c_int, C.c_int, fp, fp]),
            "gpuga_run": (C.c_int, [C.c_char_p, P(Options), dp]),
            "gpuga_read_ga_in": (C.c_int, [C.c_char_p, ip, ip, ip, fp, P(C.c_uint)]),
            "gpuga_server_create": (C.c_int, [C.c_double, C.c_int, P(Options), P(vp)]),
            "gpuga_server_destroy": (None, [vp]),
            "gpuga_server_submit": (C.c_int, [vp, C.c_char_p, C.c_int, ip]),
            "gpuga_server_drain": (C.c_int, [vp]),
            "gpuga_server_poll": (C.c_int, [vp, ip, ip, ip]),
            "gpuga_server_job": (C.c_int, [vp, C.c_int, ip, ip, fp, fp]),
            "gpuga_server_latencies": (C.c_int, [vp, dp, ip]),
            "gpuga_server_admissions": (C.c_int, [vp, dp, ip]),
            "gpuga_fitness_geometry": (C.c_int, [vp, ip, fp, fp, fp]),
            "gpuga_inspect_geometry": (C.c_int, [C.c_char_p, ip, fp, fp, fp]),
            "gpuga_default_convert_options": (None, [P(ConvertOptions)]),
            "gpuga_construct_train_in": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p,
                                                   P(ConvertOptions)]),
            "gpuga_expand_train_in": (C.c_int, [C.c_char_p, C.c_char_p, ip]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    """Names of the C-ABI entry points this binding expects."""
    lib()
    return [n for n in dir(_LIB) if n.startswith("gpuga_")]


def _check(rc):
    if rc != 0:
        raise GpugaError(lib().gpuga_last_error().decode())


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, ct=C.c_float):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def options(**kw) -> Options:
    o = Options()
    lib().gpuga_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, int(v))
    return o


class Fitness:
    """Resident training set + potential; the reference's class Fitness."""

    def __init__(self, input_dir=None, *, arrays=None, **opts):
        self._h = C.c_void_p()
        o = options(**opts)
        if input_dir is not None:
            _check(lib().gpuga_fitness_create(str(input_dir).encode(), C.byref(o),
                                              C.byref(self._h)))
        else:
            a = arrays
            Na = np.ascontiguousarray(a["Na"], np.int32)
            typ = np.ascontiguousarray(a["type"], np.int32)
            box = _f32(a["box"])
            r = _f32(a["r"])
            V = a.get("V")
            pmin = _f32(a.get("pmin", np.zeros(1)))
            pmax = _f32(a.get("pmax", np.zeros(1)))
            pe = _f32(a.get("pe_ref", np.zeros(len(Na))))
            vir = _f32(a.get("virial_ref", np.zeros(6 * len(Na))))
            fr = _f32(a.get("force_ref", np.zeros(r.size)))
            w = a.get("weights", (0.05, 0.15, 0.8))
            _check(lib().gpuga_fitness_create_arrays(
                int(a["potential_type"]), int(a.get("num_neurons", 0)), float(a["cutoff"]),
                float(w[0]), float(w[1]), float(w[2]), _ptr(pmin), _ptr(pmax), len(Na),
                int(a["Nc_force"]), _ptr(Na, C.c_int), _ptr(typ, C.c_int), _ptr(box), _ptr(r),
                _ptr(pe), _ptr(vir), _ptr(fr), C.byref(o), C.byref(self._h)))
            del V
        info = np.zeros(10, np.int32)
        _check(lib().gpuga_fitness_info(self._h, _ptr(info, C.c_int)))
        (self.potential_type, self.Nc, self.Nc_force, self.N, self.N_force, self.max_Na,
         self.bonds, self.max_neighbors, self.number_of_variables, self.num_neurons) = map(
            int, info)
        self.parameters_min = np.zeros(self.number_of_variables, np.float32)
        self.parameters_max = np.zeros(self.number_of_variables, np.float32)
        _check(lib().gpuga_fitness_bounds(self._h, _ptr(self.parameters_min),
                                          _ptr(self.parameters_max)))

    @property
    def V(self):
        return self.number_of_variables

    def compute(self, genes, *, terms=False, plain=False):
        g = _f32(genes).reshape(-1, self.V)
        P = g.shape[0]
        fit = np.zeros(P, np.float32)
        t = np.zeros((P, 3), np.float64)
        if plain:
            _check(lib().gpuga_fitness_compute_plain(self._h, P, _ptr(g), _ptr(fit),
                                                     _ptr(t, C.c_double)))
        else:
            _check(lib().gpuga_fitness_compute_terms(self._h, P, _ptr(g), _ptr(fit),
                                                     _ptr(t, C.c_double)))
        return (fit, t) if terms else fit

    def denormalise(self, genes):
        """genes in [0,1] -> physical parameters, float arithmetic as the reference."""
        g = _f32(genes)
        b = (self.parameters_max - self.parameters_min).astype(np.float32)
        return (self.parameters_min + (b * g).astype(np.float32)).astype(np.float32)

    def find_force(self, params):
        p = _f32(params)
        f = np.zeros(3 * self.N, np.float32)
        w = np.zeros(6 * self.N, np.float32)
        e = np.zeros(self.N, np.float32)
        _check(lib().gpuga_fitness_find_force(self._h, _ptr(p), _ptr(f), _ptr(w), _ptr(e)))
        return f, w, e

    def predict(self, directory, elite_genes):
        _check(lib().gpuga_fitness_predict(self._h, str(directory).encode(),
                                           _ptr(_f32(elite_genes))))

    def neighbors(self):
        NN = np.zeros(self.N, np.int32)
        start = np.zeros(self.N, np.int32)
        NL = np.zeros(max(self.bonds, 1), np.int32)
        _check(lib().gpuga_fitness_neighbors(self._h, _ptr(NN, C.c_int), _ptr(start, C.c_int),
                                             _ptr(NL, C.c_int)))
        return NN, start, NL[: self.bonds]

    def geometry(self):
        """The geometry tables built on the device: (bond, tri, rec), float4 rows."""
        return _tables(lambda *a: lib().gpuga_fitness_geometry(self._h, *a))

    def autotune(self, P):
        c = C.c_int()
        _check(lib().gpuga_fitness_autotune(self._h, int(P), C.byref(c)))
        return c.value

    def launch_shape(self, P):
        """The population launch's shape for P: chunk_cost, queues,
        blocks_per_cu, packed (tuned first when the options ask for it)."""
        v = [C.c_int() for _ in range(4)]
        _check(lib().gpuga_fitness_launch_shape(self._h, int(P), *[C.byref(x) for x in v]))
        return dict(zip(("chunk_cost", "queues", "blocks_per_cu", "packed"), (x.value for x in v)))

    def run_ga(self, ga, first_generation, generations):
        """GA generations with population and GA on the device; returns the
        per-generation best fitness [G] and elite genes [G][V]."""
        best = np.zeros(generations, np.float32)
        elite = np.zeros((generations, self.V), np.float32)
        _check(lib().gpuga_fitness_run_ga(self._h, ga._h, int(first_generation), int(generations),
                                          _ptr(best), _ptr(elite)))
        return best, elite

    def set_timing(self, on=True):
        _check(lib().gpuga_fitness_set_timing(self._h, int(bool(on))))

    def timing(self, reset=False):
        ms = C.c_double()
        n = C.c_long()
        _check(lib().gpuga_fitness_timing(self._h, C.byref(ms), C.byref(n), int(reset)))
        return ms.value, n.value

    def trace(self, genes):
        """Work-item timeline of one population generation: array [n_items, 12]
        of (config, lane block, chunk, atom begin, atom end, phase, wave slot,
        start ns, end ns, wave entry ns on the wave's first item else NaN,
        compute begin ns, compute end ns)."""
        g = _f32(genes).reshape(-1, self.V)
        n = C.c_int(0)
        _check(lib().gpuga_fitness_trace(self._h, g.shape[0], _ptr(g), None, C.byref(n)))
        rows = np.zeros((n.value, 12))
        _check(lib().gpuga_fitness_trace(self._h, g.shape[0], _ptr(g), _ptr(rows, C.c_double),
                                         C.byref(n)))
        return rows

    def set_wait_limit(self, spins):
        """Polls of the population kernel's dependency wait before a timeout
        (< 0: time out at once; a test hook for the failure path)."""
        _check(lib().gpuga_fitness_set_wait_limit(self._h, int(spins)))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def device_bytes(self):
        b = C.c_double()
        _check(lib().gpuga_fitness_device_bytes(self._h, C.byref(b)))
        return b.value

    def close(self):
        if getattr(self, "_h", None) and _LIB is not None and self._h.value:
            _LIB.gpuga_fitness_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GA:
    """The reference's GA bookkeeping (initial draw, sort, crossover, mutation)."""

    def __init__(self, population_size, parent_number, V, maximum_generation=1000,
                 mutation_rate=0.2, seed=12345678):
        self._h = C.c_void_p()
        _check(lib().gpuga_ga_create(population_size, parent_number, V, maximum_generation,
                                     mutation_rate, seed, C.byref(self._h)))
        self.P, self.V = population_size, V
        self.parent_number = parent_number

    @property
    def population(self):
        g = np.zeros((self.P, self.V), np.float32)
        _check(lib().gpuga_ga_get_population(self._h, _ptr(g)))
        return g

    @population.setter
    def population(self, genes):
        _check(lib().gpuga_ga_set_population(self._h, _ptr(_f32(genes))))

    def step(self, generation, fitness):
        f = _f32(fitness)
        out = np.zeros(self.P, np.float32)
        _check(lib().gpuga_ga_step(self._h, generation, _ptr(f), _ptr(out)))
        return out

    def step_device(self, generation, fitness):
        """The same step by the device GA kernel (identical results)."""
        f = _f32(fitness)
        out = np.zeros(self.P, np.float32)
        _check(lib().gpuga_ga_step_device(self._h, generation, _ptr(f), _ptr(out)))
        return out

    def __del__(self):
        try:
            if self._h.value:
                lib().gpuga_ga_destroy(self._h)
        except Exception:
            pass


def inspect_neighbors(input_dir):
    """Host-only neighbour list of a job directory (no HIP device needed)."""
    sizes = np.zeros(4, np.int32)
    d = str(input_dir).encode()
    _check(lib().gpuga_inspect_neighbors(d, _ptr(sizes, C.c_int), None, None, None))
    N, B = int(sizes[0]), int(sizes[1])
    NN = np.zeros(N, np.int32)
    start = np.zeros(N, np.int32)
    NL = np.zeros(max(B, 1), np.int32)
    _check(lib().gpuga_inspect_neighbors(d, _ptr(sizes, C.c_int), _ptr(NN, C.c_int),
                                         _ptr(start, C.c_int), _ptr(NL, C.c_int)))
    return NN, start, NL[:B]


def _tables(call):
    sizes = np.zeros(3, np.int32)
    _check(call(_ptr(sizes, C.c_int), None, None, None))
    t = [np.zeros((max(int(n), 1), 4), np.float32) for n in sizes]
    _check(call(_ptr(sizes, C.c_int), *[_ptr(a) for a in t]))
    return tuple(a[:int(n)] for a, n in zip(t, sizes))


def inspect_geometry(input_dir):
    """Host-built geometry tables of a job directory: (bond, tri, rec), float4 rows."""
    d = str(input_dir).encode()
    return _tables(lambda *a: lib().gpuga_inspect_geometry(d, *a))


def read_ga_in(input_dir):
    """ga.in of a job directory (host only)."""
    G, P, par, seed = C.c_int(), C.c_int(), C.c_int(), C.c_uint()
    rate = C.c_float()
    _check(lib().gpuga_read_ga_in(str(input_dir).encode(), C.byref(G), C.byref(P), C.byref(par),
                                  C.byref(rate), C.byref(seed)))
    return {"maximum_generation": G.value, "population_size": P.value,
            "parent_number": par.value, "mutation_rate": rate.value, "seed": seed.value}


def construct_train_in(list_path, out_path, cells=(4, 4, 3), delta_e=1.31684,
                       stress_in_kbar=True):
    """tools/construct_train.in/construct_train_in.cpp: DFT data files -> train.in.

    list_path holds what that tool reads on standard input (Nc_force
    Nc_energy, then the file names, relative to list_path's directory)."""
    o = ConvertOptions()
    lib().gpuga_default_convert_options(C.byref(o))
    o.cells[:] = [int(c) for c in cells]
    o.delta_e = float(delta_e)
    o.stress_in_kbar = int(bool(stress_in_kbar))
    with open(list_path, "rb") as f:
        text = f.read()
    d = os.path.dirname(os.path.abspath(list_path)).encode()
    _check(lib().gpuga_construct_train_in(text, d, str(out_path).encode(), C.byref(o)))


def expand_train_in(in_path, out_path, cells=(2, 2, 2)):
    """tools/construct_train.in/expand222.cpp: every cell replicated."""
    c = np.array(cells, np.int32)
    _check(lib().gpuga_expand_train_in(str(in_path).encode(), str(out_path).encode(),
                                       _ptr(c, C.c_int)))


def run(input_dir, **opts):
    """GPUGA::GPUGA(dir): GA fit with ga.out / force.out / energy.out / virial.out."""
    spg = C.c_double()
    o = options(**opts)
    _check(lib().gpuga_run(str(input_dir).encode(), C.byref(o), C.byref(spg)))
    return spg.value


class Server:
    """Concurrent fitting jobs with resident training sets (continuous batching)."""

    def __init__(self, memory_budget_bytes=64e9, max_active=16, **opts):
        self._h = C.c_void_p()
        o = options(**opts)
        _check(lib().gpuga_server_create(float(memory_budget_bytes), int(max_active),
                                         C.byref(o), C.byref(self._h)))

    def submit(self, input_dir, write_outputs=False):
        jid = C.c_int()
        _check(lib().gpuga_server_submit(self._h, str(input_dir).encode(), int(write_outputs),
                                         C.byref(jid)))
        return jid.value

    def poll(self):
        a, q, f = C.c_int(), C.c_int(), C.c_int()
        _check(lib().gpuga_server_poll(self._h, C.byref(a), C.byref(q), C.byref(f)))
        return a.value, q.value, f.value

    def drain(self):
        _check(lib().gpuga_server_drain(self._h))

    def job(self, job_id, V):
        st, gens, best = C.c_int(), C.c_int(), C.c_float()
        elite = np.zeros(V, np.float32)
        _check(lib().gpuga_server_job(self._h, job_id, C.byref(st), C.byref(gens),
                                      C.byref(best), _ptr(elite)))
        return {"state": st.value, "generations": gens.value, "best": best.value,
                "elite": elite}

    def latencies(self):
        """Per-generation latencies (ms), all jobs, completion order."""
        return self._times(lib().gpuga_server_latencies)

    def admissions(self):
        """Per admitted job: submission to ready (ms)."""
        return self._times(lib().gpuga_server_admissions)

    def _times(self, fn):
        n = C.c_int(0)
        _check(fn(self._h, None, C.byref(n)))
        out = np.zeros(max(n.value, 1))
        _check(fn(self._h, _ptr(out, C.c_double), C.byref(n)))
        return out[: n.value]

    def close(self):
        if getattr(self, "_h", None) and _LIB is not None and self._h.value:
            _LIB.gpuga_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
