"""MI355X-native NSGA-III codon-design engine.

Python binding (ctypes) of libcd_engine.so, the HIP/gfx950 engine behind the
C ABI in include/codon_nsga3.h.  It mirrors the reference's host loop
(src/main.cu:467-787): create a job from a protein and a codon table,
initialise the population, advance generations, read the pool back.
The engine has no CPU fallback: loading fails if the HIP library is absent.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from . import tables  # noqa: F401
from . import workloads  # noqa: F401
from .tables import Tables

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcd_engine.so")
CSRC = os.path.join(HERE, "csrc")

CD_OK = 0
# kernels timed by cd_job_profile: offspring, stem, dominance (with the stem
# objective), fronts, associate, niche
PROFILE_KERNELS = 6
ERRORS = {-1: "CD_E_INVALID", -2: "CD_E_NOMEM", -3: "CD_E_DEVICE", -4: "CD_E_STATE", -5: "CD_E_BUDGET"}

# Every entry point of include/codon_nsga3.h
SYMBOLS = ("cd_job_create", "cd_job_init", "cd_job_step", "cd_job_step_timed", "cd_job_sync", "cd_job_generation",
           "cd_job_read", "cd_job_profile", "cd_job_cds_scanned", "cd_job_destroy", "cd_eval_batch", "cd_nds", "cd_select",
           "cd_batch_create", "cd_batch_step", "cd_batch_step_timed", "cd_batch_sync", "cd_batch_profile",
           "cd_batch_destroy",
           "cd_job_device_bytes", "cd_last_error", "cd_hypervolume", "cd_normalize", "cd_min_distance_spread",
           "cd_server_create", "cd_server_submit", "cd_server_wait", "cd_server_result", "cd_server_stats",
           "cd_server_destroy", "cd_nds_noalloc", "cd_job_tune", "cd_batch_tune", "cd_job_shapes",
           "cd_batch_shapes")
SERVER_QUEUED, SERVER_RUNNING, SERVER_DONE, SERVER_FAILED = 0, 1, 2, 3


class CdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class _Tables(C.Structure):
    _fields_ = [("codons", C.c_char_p), ("weights", C.POINTER(C.c_float)), ("cps", C.POINTER(C.c_float)),
                ("gc3_ref", C.c_float)]


# launch shapes (cd_shapes): every shape gives the same results
SHAPE_FIELDS = ("design_split", "design_threads", "stem_waves", "dom_rows", "fronts_threads", "assoc_designs",
                "niche_threads")


class _Shapes(C.Structure):
    _fields_ = [(f, C.c_int32) for f in SHAPE_FIELDS]


def _shapes_dict(s):
    return {f: int(getattr(s, f)) for f in SHAPE_FIELDS}


def _shapes_struct(d):
    return _Shapes(*[int(d[f]) for f in SHAPE_FIELDS])


class _Desc(C.Structure):
    _fields_ = [("L", C.c_int32), ("m", C.c_int32), ("N", C.c_int32), ("H", C.c_int32), ("Pm", C.c_float),
                ("seed", C.c_uint64), ("aa", C.POINTER(C.c_int8)), ("tables", _Tables)]


def build(force=False):
    """Compile libcd_engine.so for gfx950 (hipcc, see csrc/Makefile)."""
    args = ["make", "-s", "-C", CSRC]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        P = np.ctypeslib.ndpointer
        i8 = P(np.int8, flags="C_CONTIGUOUS")
        f32 = P(np.float32, flags="C_CONTIGUOUS")
        i32 = P(np.int32, flags="C_CONTIGUOUS")
        L.cd_job_create.argtypes = [C.POINTER(_Desc), C.POINTER(vp)]
        L.cd_job_init.argtypes = [vp]
        L.cd_job_step.argtypes = [vp, C.c_int32]
        L.cd_job_step_timed.argtypes = [vp, C.c_int32, C.POINTER(C.c_float)]
        L.cd_job_sync.argtypes = [vp]
        L.cd_job_generation.argtypes = [vp, C.POINTER(C.c_int32)]
        L.cd_job_read.argtypes = [vp, i8, f32, i8, i32, i32, i32, f32]
        L.cd_job_profile.argtypes = [vp, C.c_int32, f32]
        L.cd_job_cds_scanned.argtypes = [vp, C.POINTER(C.c_int64)]
        L.cd_job_destroy.argtypes = [vp]
        L.cd_eval_batch.argtypes = [C.POINTER(_Desc), i8, C.c_int32, f32, i8, i32]
        L.cd_nds.argtypes = [f32, C.c_int32, C.c_int32, i32, C.POINTER(C.c_int32)]
        L.cd_nds_noalloc.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp]
        SP = C.POINTER(_Shapes)
        L.cd_job_tune.argtypes = [vp, C.c_int32, SP, C.POINTER(C.c_int32)]
        L.cd_batch_tune.argtypes = [vp, C.c_int32, SP, C.POINTER(C.c_int32)]
        L.cd_job_shapes.argtypes = [vp, SP, SP]
        L.cd_batch_shapes.argtypes = [vp, SP, SP]
        L.cd_select.argtypes = [f32, C.c_int32, C.c_int32, C.c_uint32, C.c_uint64, f32, i32, i32, f32, i32]
        L.cd_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp)]
        L.cd_batch_step.argtypes = [vp, C.c_int32]
        L.cd_batch_step_timed.argtypes = [vp, C.c_int32, C.POINTER(C.c_float)]
        L.cd_batch_sync.argtypes = [vp]
        L.cd_batch_profile.argtypes = [vp, C.c_int32, f32]
        L.cd_batch_destroy.argtypes = [vp]
        L.cd_job_device_bytes.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.cd_job_device_bytes.restype = C.c_int64
        L.cd_last_error.restype = C.c_char_p
        f64 = P(np.float64, flags="C_CONTIGUOUS")
        L.cd_hypervolume.argtypes = [f64, C.c_int32, C.c_int32, f64, C.POINTER(C.c_double)]
        L.cd_normalize.argtypes = [f32, C.c_int32, f32]
        L.cd_min_distance_spread.argtypes = [f32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.cd_server_create.argtypes = [C.c_int64, C.c_int32, C.POINTER(vp)]
        L.cd_server_submit.argtypes = [vp, C.POINTER(_Desc), C.c_int32, C.POINTER(C.c_int64)]
        L.cd_server_wait.argtypes = [vp, C.c_int64, C.c_double, C.POINTER(C.c_int32)]
        L.cd_server_result.argtypes = [vp, C.c_int64, i8, f32, i8, i32, i32, i32, f64]
        L.cd_server_stats.argtypes = [vp, P(np.int64, flags="C_CONTIGUOUS")]
        L.cd_server_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _check(rc):
    if rc != CD_OK:
        raise CdError(rc, lib().cd_last_error().decode())


def _desc(aa, tab: Tables, m, N, H, Pm, seed):
    aa = np.ascontiguousarray(aa, np.int8)
    w = np.ascontiguousarray(tab.weights, np.float32)
    cps = np.ascontiguousarray(tab.cps, np.float32)
    cod = tab.codon_bytes()
    t = _Tables(cod, w.ctypes.data_as(C.POINTER(C.c_float)), cps.ctypes.data_as(C.POINTER(C.c_float)),
                float(tab.gc3_ref))
    d = _Desc(len(aa), m, N, H, float(Pm), int(seed) & 0xFFFFFFFFFFFFFFFF, aa.ctypes.data_as(C.POINTER(C.c_int8)), t)
    return d, (aa, w, cps, cod)  # keep the arrays alive


def device_bytes(L, m, N, H):
    return int(lib().cd_job_device_bytes(L, m, N, H))


class Job:
    """One codon-design job resident on the GPU (cd_job_*)."""

    def __init__(self, aa, tab: Tables, m, N, H, Pm, seed):
        self.L, self.m, self.N, self.H = len(aa), m, N, H
        d, self._keep = _desc(aa, tab, m, N, H, Pm, seed)
        h = C.c_void_p()
        _check(lib().cd_job_create(C.byref(d), C.byref(h)))
        self.h = h

    def init(self):
        _check(lib().cd_job_init(self.h))
        return self

    def step(self, generations=1):
        _check(lib().cd_job_step(self.h, generations))

    def step_timed(self, generations):
        ms = C.c_float()
        _check(lib().cd_job_step_timed(self.h, generations, C.byref(ms)))
        return ms.value

    def sync(self):
        _check(lib().cd_job_sync(self.h))

    @property
    def generation(self):
        g = C.c_int32()
        _check(lib().cd_job_generation(self.h, C.byref(g)))
        return g.value

    def profile(self, generations):
        out = np.zeros(PROFILE_KERNELS, np.float32)
        _check(lib().cd_job_profile(self.h, generations, out))
        return out

    def tune(self, gens_per_candidate=3):
        """Autotune the launch shapes on this job's own generations (cd_job_tune);
        returns (chosen shapes, generations run)."""
        sh, n = _Shapes(), C.c_int32()
        _check(lib().cd_job_tune(self.h, gens_per_candidate, C.byref(sh), C.byref(n)))
        return _shapes_dict(sh), n.value

    def shapes(self, set=None):
        """The job's launch shapes; with `set` (a dict), install those first."""
        sh = _Shapes()
        st = C.byref(_shapes_struct(set)) if set is not None else None
        _check(lib().cd_job_shapes(self.h, C.byref(sh), st))
        return _shapes_dict(sh)

    def cds_scanned(self):
        n = C.c_int64()
        _check(lib().cd_job_cds_scanned(self.h, C.byref(n)))
        return n.value

    def read(self):
        P = 2 * self.N
        g = np.zeros((P, self.m, self.L), np.int8)
        obj = np.zeros((P, 6), np.float32)
        idx = np.zeros((P, 12), np.int8)
        pql = np.zeros((P, 3), np.int32)
        srt = np.zeros(self.N, np.int32)
        fo = np.zeros(P, np.int32)
        st = np.zeros(54, np.float32)
        _check(lib().cd_job_read(self.h, g, obj, idx, pql, srt, fo, st))
        return dict(genome=g, obj=obj, idx=idx, pql=pql, sorted=srt, front_of=fo, extremes=st[:36].reshape(6, 6),
                    ideal=st[36:42], worst=st[42:48], nadir=st[48:54])

    def close(self):
        if getattr(self, "h", None):
            # CD_E_STATE: a live batch still holds the job (only possible at
            # interpreter exit, since a Batch keeps its jobs referenced)
            if lib().cd_job_destroy(self.h) == CD_OK:
                self.h = None

    def __del__(self):
        self.close()


class Batch:
    """Initialised jobs advanced together (cd_batch_*)."""

    def __init__(self, jobs):
        self.jobs = list(jobs)
        arr = (C.c_void_p * len(self.jobs))(*[j.h.value for j in self.jobs])
        h = C.c_void_p()
        _check(lib().cd_batch_create(arr, len(self.jobs), C.byref(h)))
        self.h = h

    def step(self, generations=1):
        _check(lib().cd_batch_step(self.h, generations))

    def step_timed(self, generations):
        ms = C.c_float()
        _check(lib().cd_batch_step_timed(self.h, generations, C.byref(ms)))
        return ms.value

    def sync(self):
        _check(lib().cd_batch_sync(self.h))

    def tune(self, gens_per_candidate=3):
        """Autotune the launch shapes on the batch's own generations (cd_batch_tune);
        returns (chosen shapes, generations run)."""
        sh, n = _Shapes(), C.c_int32()
        _check(lib().cd_batch_tune(self.h, gens_per_candidate, C.byref(sh), C.byref(n)))
        return _shapes_dict(sh), n.value

    def shapes(self, set=None):
        """The batch's launch shapes; with `set` (a dict), install those first."""
        sh = _Shapes()
        st = C.byref(_shapes_struct(set)) if set is not None else None
        _check(lib().cd_batch_shapes(self.h, C.byref(sh), st))
        return _shapes_dict(sh)

    def profile(self, generations):
        """Mean milliseconds of each phase of a batched generation (cd_batch_profile)."""
        out = np.zeros(PROFILE_KERNELS, np.float32)
        _check(lib().cd_batch_profile(self.h, generations, out))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().cd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Server:
    """Continuous batching of jobs under a device-memory budget (cd_server_*)."""

    def __init__(self, hbm_budget_bytes, max_jobs=256):
        h = C.c_void_p()
        _check(lib().cd_server_create(int(hbm_budget_bytes), int(max_jobs), C.byref(h)))
        self.h = h
        self._shapes = {}

    def submit(self, aa, tab: Tables, m, N, H, Pm, seed, generations):
        d, keep = _desc(aa, tab, m, N, H, Pm, seed)
        t = C.c_int64()
        _check(lib().cd_server_submit(self.h, C.byref(d), int(generations), C.byref(t)))
        del keep  # the server copied the description
        self._shapes[t.value] = (len(aa), m, N)
        return t.value

    def wait(self, ticket, timeout=-1.0):
        st = C.c_int32()
        _check(lib().cd_server_wait(self.h, ticket, float(timeout), C.byref(st)))
        return st.value

    def result(self, ticket):
        """Final pool of a finished request (as Job.read) and its times
        (queued, resident, time to front) in seconds; releases the request."""
        if ticket not in self._shapes:
            raise CdError(-1, "unknown ticket")
        L, m, N = self._shapes.pop(ticket)
        P = 2 * N
        out = {"genome": np.zeros((P, m, L), np.int8), "obj": np.zeros((P, 6), np.float32),
               "idx": np.zeros((P, 12), np.int8), "pql": np.zeros((P, 3), np.int32),
               "sorted": np.zeros(N, np.int32), "front_of": np.zeros(P, np.int32)}
        times = np.zeros(3, np.float64)
        _check(lib().cd_server_result(self.h, ticket, out["genome"], out["obj"], out["idx"], out["pql"],
                                      out["sorted"], out["front_of"], times))
        out["times"] = {"queued": times[0], "resident": times[1], "time_to_front": times[2]}
        return out

    def stats(self):
        s = np.zeros(11, np.int64)
        _check(lib().cd_server_stats(self.h, s))
        return {"queued": int(s[0]), "resident": int(s[1]), "bytes": int(s[2]), "generations": int(s[3]),
                "designs": int(s[4]), "tick_gpu_s": s[5] / 1e6, "worker_busy_s": s[6] / 1e6,
                "rebuilds": int(s[7]), "rebuild_s": s[8] / 1e6, "launch_s": s[9] / 1e6, "wait_s": s[10] / 1e6}

    def close(self):
        if getattr(self, "h", None):
            lib().cd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def evaluate(aa, tab: Tables, genome):
    """cd_eval_batch: genome [count][m][L] int8 -> obj [count][6], idx [count][12], pql [count][3]."""
    genome = np.ascontiguousarray(genome, np.int8)
    count, m, L = genome.shape
    d, keep = _desc(aa, tab, m, 1, 1, 0.0, 0)
    obj = np.zeros((count, 6), np.float32)
    idx = np.zeros((count, 12), np.int8)
    pql = np.zeros((count, 3), np.int32)
    _check(lib().cd_eval_batch(C.byref(d), genome, count, obj, idx, pql))
    return obj, idx, pql


def nds(obj, N):
    """cd_nds: fronts of P objective vectors until >= N are ranked."""
    obj = np.ascontiguousarray(obj, np.float32)
    fo = np.zeros(len(obj), np.int32)
    nf = C.c_int32()
    _check(lib().cd_nds(obj, len(obj), N, fo, C.byref(nf)))
    return fo, nf.value


def select(obj, N, H, gen, seed, extremes):
    """cd_select: one NSGA-III selection of N survivors from 2N vectors."""
    obj = np.ascontiguousarray(obj, np.float32)
    ext = np.ascontiguousarray(extremes, np.float32).reshape(36).copy()
    srt = np.zeros(N, np.int32)
    fo = np.zeros(2 * N, np.int32)
    norm = np.zeros(18, np.float32)
    flags = np.zeros(2, np.int32)
    _check(lib().cd_select(obj, N, H, gen, seed, ext, srt, fo, norm, flags))
    return dict(sorted=srt, front_of=fo, extremes=ext.reshape(6, 6), ideal=norm[:6], worst=norm[6:12],
                nadir=norm[12:], n_cut=int(flags[0]), hyp_exception=int(flags[1]))


def hypervolume(points, ref=None):
    """Exact hypervolume (minimisation) of points [n][d] against ref (default all ones)."""
    pts = np.ascontiguousarray(points, np.float64)
    n, d = pts.shape
    r = np.ones(d) if ref is None else np.ascontiguousarray(ref, np.float64)
    out = C.c_double()
    _check(lib().cd_hypervolume(pts, n, d, r, C.byref(out)))
    return out.value


def normalize(obj):
    """main.cu:789-835 normalisation of raw objective records [n][6]."""
    obj = np.ascontiguousarray(obj, np.float32)
    out = np.zeros_like(obj)
    _check(lib().cd_normalize(obj, len(obj), out))
    return out


def quality(obj):
    """(hypervolume, min distance to ideal, spread) of a pool of raw records,
    as the reference reports them (normalised, reference point 1^6)."""
    z = normalize(obj)
    md, sp = C.c_float(), C.c_float()
    _check(lib().cd_min_distance_spread(z, len(z), C.byref(md), C.byref(sp)))
    return hypervolume(z.astype(np.float64)), md.value, sp.value
