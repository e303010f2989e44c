"""ctypes binding of the CPU oracle (libcd_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP engine is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcd_oracle.so")
_lib = None

i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(path):
    L = C.CDLL(path)
    _bind(L)
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _load(LIB)
    return _lib


@contextlib.contextmanager
def use_library(path):
    """Route every call (and every Job created) inside the block to another
    build of the oracle, e.g. one with a deliberately planted mistake that a
    test must detect; the default library is restored afterwards."""
    global _lib
    saved = lib()
    _lib = _load(path)
    try:
        yield _lib
    finally:
        _lib = saved


def build_variant(out_path, patches):
    """Compile the oracle with source substitutions into out_path (test use:
    planting a known mistake).  patches: (function signature, old, new)
    triples; `old` must occur exactly once inside that function's body."""
    src = open(os.path.join(HERE, "cd_oracle.cpp")).read()
    for sig, old, new in patches:
        a = src.index(sig)
        b = src.index("\n}\n", a)
        body = src[a:b]
        assert body.count(old) == 1, (sig, old)
        src = src[:a] + body.replace(old, new) + src[b:]
    cpp = out_path + ".cpp"
    with open(cpp, "w") as f:
        f.write(src)
    return subprocess.Popen(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", "-I", HERE,
                             "-o", out_path, cpp])


def build_native(out_path):
    """The oracle compiled for the host it runs on (-O3 -march=native, no
    fast-math; the same results): the CPU baseline's build.  Returns the
    library path, or None when the compiler fails."""
    r = subprocess.run(["g++", "-O3", "-march=native", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
                        "-shared", "-I", HERE, "-o", out_path, os.path.join(HERE, "cd_oracle.cpp")],
                       capture_output=True, text=True)
    return out_path if r.returncode == 0 else None


def _bind(L):
    L.or_philox4x32_10.argtypes = [u32p, u32p, u32p]
    L.or_det_exp.argtypes = [C.c_double]
    L.or_det_exp.restype = C.c_double
    L.or_dasdennis.argtypes = [C.c_int, C.c_int, f32p]
    L.or_sl_cds.argtypes = [u8p, C.c_int, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                            C.POINTER(C.c_int32)]
    L.or_eval.argtypes = [C.c_int, C.c_int, i8p, C.c_char_p, f32p, f32p, C.c_float, i8p, C.c_int,
                          f32p, i8p, i32p]
    L.or_cai_cpb_ref.argtypes = [C.c_int, C.c_int, i8p, C.c_char_p, f32p, f32p, i8p, C.c_int, C.c_int,
                                 f32p, f64p]
    L.or_mutate.argtypes = [C.c_int, C.c_int, i8p, C.c_char_p, f32p, f32p, C.c_float, C.c_float, C.c_uint64,
                            C.c_uint32, C.c_uint32, i8p, f32p, i8p, i32p, i8p]
    L.or_nds.argtypes = [f32p, C.c_int, C.c_int, i32p, C.POINTER(C.c_int32)]
    L.or_select.argtypes = [f32p, C.c_int, C.c_int, C.c_uint32, C.c_uint64, f32p, i32p, i32p, f32p, i32p]
    L.or_job_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64, i8p, C.c_char_p,
                                f32p, f32p, C.c_float]
    L.or_job_create.restype = C.c_void_p
    L.or_job_init.argtypes = [C.c_void_p]
    L.or_job_step.argtypes = [C.c_void_p, C.c_int]
    L.or_job_read.argtypes = [C.c_void_p, i8p, f32p, i8p, i32p, i32p, i32p, f32p]
    L.or_job_generation.argtypes = [C.c_void_p]
    L.or_job_destroy.argtypes = [C.c_void_p]
    L.or_set_threads.argtypes = [C.c_int]


def set_threads(n):
    return lib().or_set_threads(int(n))


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def det_exp(x):
    return lib().or_det_exp(float(x))


def dasdennis(M, H):
    from math import comb
    out = np.zeros(comb(M + H - 1, H) * M, np.float32)
    n = lib().or_dasdennis(M, H, out)
    return out.reshape(n, M)


NT = {"A": 0, "C": 1, "G": 2, "U": 3}


def codes(seq):
    if isinstance(seq, str):
        return np.array([NT[c] for c in seq], np.uint8)
    return np.ascontiguousarray(seq, np.uint8)


def sl(seq, mode=0):
    c = codes(seq)
    l, p, q = C.c_int32(), C.c_int32(), C.c_int32()
    rc = lib().or_sl_cds(c, len(c), mode, C.byref(l), C.byref(p), C.byref(q))
    assert rc == 0
    return (l.value, p.value, q.value) if mode == 0 else l.value


def evaluate(aa, tables, syn):
    """syn [count][m][L] int8 -> obj [count][6], idx [count][12], pql [count][3]."""
    syn = np.ascontiguousarray(syn, np.int8)
    count, m, L = syn.shape
    obj = np.zeros((count, 6), np.float32)
    idx = np.zeros((count, 12), np.int8)
    pql = np.zeros((count, 3), np.int32)
    rc = lib().or_eval(L, m, np.ascontiguousarray(aa, np.int8), tables.codon_bytes(), tables.weights,
                       tables.cps, tables.gc3_ref, syn, count, obj, idx, pql)
    assert rc == 0
    return obj, idx, pql


def cai_cpb_reference(aa, tables, syn, threads=256):
    syn = np.ascontiguousarray(syn, np.int8)
    count, m, L = syn.shape
    o32 = np.zeros((count, m, 2), np.float32)
    o64 = np.zeros((count, m, 2), np.float64)
    rc = lib().or_cai_cpb_ref(L, m, np.ascontiguousarray(aa, np.int8), tables.codon_bytes(), tables.weights,
                              tables.cps, syn, count, threads, o32, o64)
    assert rc == 0
    return o32, o64


def mutate(aa, tables, Pm, seed, gen, child, parent, pobj, pidx, ppql):
    """The child genome [m][L] that child `child` of generation `gen` gets
    from its parent's genome [m][L] and record."""
    parent = np.ascontiguousarray(parent, np.int8)
    m, L = parent.shape
    out = np.zeros_like(parent)
    rc = lib().or_mutate(L, m, np.ascontiguousarray(aa, np.int8), tables.codon_bytes(), tables.weights,
                         tables.cps, tables.gc3_ref, float(Pm), int(seed), int(gen), int(child), parent,
                         np.ascontiguousarray(pobj, np.float32), np.ascontiguousarray(pidx, np.int8),
                         np.ascontiguousarray(ppql, np.int32), out)
    assert rc == 0
    return out


def nds(obj, N):
    obj = np.ascontiguousarray(obj, np.float32)
    P = obj.shape[0]
    fo = np.zeros(P, np.int32)
    nf = C.c_int32()
    lib().or_nds(obj, P, N, fo, C.byref(nf))
    return fo, nf.value


def select(obj, N, H, gen, seed, extremes):
    obj = np.ascontiguousarray(obj, np.float32)
    ext = np.ascontiguousarray(extremes, np.float32).copy()
    sorted_out = np.zeros(N, np.int32)
    fo = np.zeros(2 * N, np.int32)
    norm = np.zeros(18, np.float32)
    flags = np.zeros(2, np.int32)
    lib().or_select(obj, N, H, gen, seed, ext, sorted_out, fo, norm, flags)
    return dict(sorted=sorted_out, front_of=fo, extremes=ext.reshape(6, 6), ideal=norm[:6], worst=norm[6:12],
                nadir=norm[12:], n_cut=int(flags[0]), hyp_exception=int(flags[1]))


class Job:
    """Whole-loop oracle job (main.cu:467-787)."""

    def __init__(self, aa, tables, m, N, H, Pm, seed):
        self.aa = np.ascontiguousarray(aa, np.int8)
        self.L, self.m, self.N = len(aa), m, N
        self._L = lib()  # the library that owns the handle
        self.h = self._L.or_job_create(self.L, m, N, H, Pm, seed, self.aa, tables.codon_bytes(), tables.weights,
                                       tables.cps, tables.gc3_ref)
        if not self.h:
            raise ValueError("invalid job")

    def init(self):
        self._L.or_job_init(self.h)

    def step(self, gens=1):
        self._L.or_job_step(self.h, gens)

    @property
    def generation(self):
        return self._L.or_job_generation(self.h)

    def read(self):
        P = 2 * self.N
        g = np.zeros((P, self.m, self.L), np.int8)
        obj = np.zeros((P, 6), np.float32)
        idx = np.zeros((P, 12), np.int8)
        pql = np.zeros((P, 3), np.int32)
        srt = np.zeros(self.N, np.int32)
        fo = np.zeros(P, np.int32)
        st = np.zeros(54, np.float32)
        self._L.or_job_read(self.h, g, obj, idx, pql, srt, fo, st)
        return dict(genome=g, obj=obj, idx=idx, pql=pql, sorted=srt, front_of=fo, extremes=st[:36].reshape(6, 6),
                    ideal=st[36:42], worst=st[42:48], nadir=st[48:54])

    def __del__(self):
        if getattr(self, "h", None):
            self._L.or_job_destroy(self.h)
            self.h = None
