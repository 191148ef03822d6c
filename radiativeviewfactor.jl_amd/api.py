"""Reference-shaped entry points over the C ABI.

``compute_view_factors`` mirrors src/Assembly.jl:64-96 for the GPU branch
(the hook ``compute_view_factors_gpu``, src/GPUAssembly.jl:32-105) with an
``MI355XBackend`` carrying the knobs the reference hook drops (precision,
Duffy, Monte Carlo seed, device; src/Assembly.jl:77-78,243-252).
``ViewFactorResult``, ``aggregate_by_group``, ``check_reciprocity`` and
``check_closure`` mirror src/Results.jl:35-127.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import RVF_AUDIT, RVF_TIMING, RVF_WANT_F, RVFError, check, ptr
from .mesh import MeshData

# Occluded-quadrature traversal variants (rvf_opts.variant, include/rvf.h)
VARIANT_DEFAULT = 0
VARIANT_LANE = 1


def VARIANT_QUEUE(slots: int) -> int:
    """Wave-shared ray queue with `slots` rays in flight per lane."""
    return 10 + slots


def VARIANT_SPLIT(slots: int) -> int:
    """Ray slots walk nodes only; triangle tests run wave-wide from a task ring."""
    return 20 + slots


FP64_VARIANTS = (VARIANT_LANE, VARIANT_QUEUE(1), VARIANT_QUEUE(2), VARIANT_QUEUE(3), VARIANT_SPLIT(1),
                 VARIANT_SPLIT(2), VARIANT_SPLIT(3))
MIXED_VARIANTS = (VARIANT_LANE, VARIANT_QUEUE(2), VARIANT_QUEUE(3), VARIANT_QUEUE(4), VARIANT_QUEUE(6),
                  VARIANT_SPLIT(1), VARIANT_SPLIT(2), VARIANT_SPLIT(3))
# variants whose FP64 traversal stops at the first blocking triangle, so their
# box / triangle test counters equal the canonical stackless counts
EXACT_COUNT_VARIANTS = (VARIANT_LANE, VARIANT_QUEUE(1), VARIANT_QUEUE(2), VARIANT_QUEUE(3))

SOUP_READ = {"reference": 0, ("geometric", "xyz_vertex"): 1, ("geometric", "vertex_xyz"): 2}
SOUP_SEGMENTS = 3


@dataclass
class MI355XBackend:
    """Backend object passed through the reference's GPU hook unchanged.

    precision: "fp64" (reference arithmetic) or "mixed" (FP32 bulk with FP64
    accumulation and FP64 re-tests of undecided ray-triangle decisions).
    parity: "reference" reproduces the reference's soup reading (quirk Q1:
    boxes from ``[xyz, vertex]``, Moller-Trumbore from ``[vertex, xyz]``);
    "geometric" reads the soup consistently by ``mesh.soup_layout``."""
    precision: str = "fp64"
    use_duffy: bool = False
    seed: int = 0
    device: int = 0
    parity: str = "reference"
    audit: bool = False
    # autotuner table (JSON path): occluded solves take the traversal
    # configuration stored for their problem class, measuring and storing it
    # on a miss; None keeps the default configuration
    tune_table: Optional[str] = None


@dataclass
class ViewFactorResult:
    F_elem: np.ndarray
    A_elem: np.ndarray
    F_group: np.ndarray
    A_group: np.ndarray
    group_tags: List[int]
    group_names: List[str]
    row_sums: Optional[np.ndarray] = None
    stats: Dict = field(default_factory=dict)
    audit_count: Optional[np.ndarray] = None
    audit_hash: Optional[np.ndarray] = None


# ----------------------------------------------------------------------------
# device objects
# ----------------------------------------------------------------------------
class Context:
    def __init__(self, device: int = 0):
        L = _lib.lib()
        h = C.c_void_p()
        check(L.rvf_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            _lib.lib().rvf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def record(self, slot):
        check(_lib.lib().rvf_ctx_event_record(self.h, slot))

    def elapsed(self, a, b):
        ms = C.c_float()
        check(_lib.lib().rvf_ctx_event_elapsed(self.h, a, b, C.byref(ms)))
        return ms.value

    def copy_probe(self, nbytes: int, reps: int = 1) -> float:
        """Device-to-device copy of nbytes with the library's copy kernel
        (a kernel of known traffic, for calibrating HBM counters); avg ms."""
        ms = C.c_float()
        check(_lib.lib().rvf_copy_probe(self.h, nbytes, reps, C.byref(ms)))
        return ms.value

    def sync(self):
        check(_lib.lib().rvf_ctx_sync(self.h))


def family_matrices(mesh: MeshData):
    """Per-family node matrices and 1-based elem_node_idx, as build_gpu_arrays
    (src/GPUKernels.jl:344-385) lays them out; curve meshes: Line3, Line2."""
    nn = (8, 6, 4, 3, 3, 2)
    fams = (4, 5) if mesh.mesh_dim == 1 else (0, 1, 2, 3)
    mats, counts = [], np.zeros(len(fams), np.int64)
    idx = np.zeros(mesh.n_elems, np.int32)
    for k, f in enumerate(fams):
        sel = np.nonzero(mesh.family == f)[0]
        counts[k] = sel.size
        idx[sel] = np.arange(1, sel.size + 1, dtype=np.int32)
        mats.append(np.ascontiguousarray(mesh.nodes[sel, :nn[f]], dtype=np.int32))
    return mats, counts, idx


class DeviceMesh:
    """A mesh in device memory: surface elements, or a curve mesh (mesh_dim
    1, Line3/Line2, which the reference computes on the CPU only)."""

    def __init__(self, ctx: Context, mesh: MeshData):
        self.ctx = ctx
        self.mesh = mesh
        mats, counts, idx = family_matrices(mesh)
        self._keep = (mats, counts, idx)
        h = C.c_void_p()
        if mesh.mesh_dim == 1:
            check(_lib.lib().rvf_curve_mesh_upload(ctx.h, ptr(mesh.coords), mesh.n_nodes, mesh.n_elems,
                                                   ptr(mesh.family), ptr(idx), ptr(mesh.group), ptr(mats[0]),
                                                   int(counts[0]), ptr(mats[1]), int(counts[1]), C.byref(h)))
        else:
            arr = (C.c_void_p * 4)(*[m.ctypes.data if m.size else None for m in mats])
            check(_lib.lib().rvf_mesh_upload(ctx.h, ptr(mesh.coords), mesh.n_nodes, mesh.n_elems, ptr(mesh.family),
                                             ptr(idx), ptr(mesh.group), arr, ptr(counts), C.byref(h)))
        self.h = h
        G = C.c_int32()
        _lib.lib().rvf_mesh_info(self.h, None, C.byref(G), None)
        tags = np.zeros(max(G.value, 1), np.int32)
        _lib.lib().rvf_mesh_info(self.h, None, None, ptr(tags))
        self.tags = [int(t) for t in tags[:G.value]]

    def close(self):
        if self.h:
            _lib.lib().rvf_mesh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceObstruction:
    def __init__(self, ctx: Context, dmesh: DeviceMesh, obstruction_groups: Sequence[int], parity: str):
        mesh = dmesh.mesh
        tags = sorted(set(int(g) for g in obstruction_groups if int(g) in mesh.group_tri_soup))
        if tags:
            soup = np.ascontiguousarray(np.concatenate([mesh.group_tri_soup[t] for t in tags]), np.float64)
            tg = np.concatenate([np.full(mesh.group_tri_soup[t].shape[0], t, np.int32) for t in tags])
        else:
            soup = np.zeros((0, mesh.soup_vertices, 3)); tg = np.zeros(0, np.int32)
        obs = np.ascontiguousarray(list(obstruction_groups), np.int32)
        if mesh.mesh_dim == 1:  # segment soups read one way only (src/BVH.jl:81-109, 307)
            mode = SOUP_SEGMENTS
        else:
            mode = SOUP_READ["reference"] if parity == "reference" else SOUP_READ[("geometric", mesh.soup_layout)]
        self._keep = (soup, tg, obs)
        h = C.c_void_p()
        check(_lib.lib().rvf_obs_build(ctx.h, dmesh.h, ptr(soup), ptr(tg), soup.shape[0], ptr(obs), len(obs), mode,
                                       C.byref(h)))
        self.h = h
        nt = C.c_int32(); nn = C.c_int64(); ntri = C.c_int64()
        _lib.lib().rvf_obs_info(self.h, C.byref(nt), C.byref(nn), C.byref(ntri))
        self.n_trees, self.n_nodes, self.n_tris = nt.value, nn.value, ntri.value

    def export(self, tree=0):
        L = _lib.lib()
        cap = max(self.n_nodes, self.n_tris, 1)
        box = np.zeros(6 * cap); meta = np.zeros(5 * cap, np.int32); tri = np.zeros(cap, np.int32)
        nn = C.c_int64(); nt = C.c_int64()
        check(L.rvf_obs_export(self.h, tree, cap, ptr(box), ptr(meta), ptr(tri), C.byref(nn), C.byref(nt)))
        return box[:6 * nn.value].reshape(-1, 6), meta[:5 * nn.value].reshape(-1, 5), tri[:nt.value]

    def close(self):
        if self.h:
            _lib.lib().rvf_obs_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TuneTable:
    """Persisted autotuner choices: a JSON file mapping a problem class
    (precision, method, nquad, element count and occluder triangle count
    rounded to powers of two) to the fastest (variant, refill_min, columns,
    waves) measured for it on this machine."""

    def __init__(self, path: str):
        self.path = path
        self.entries = {}
        if os.path.exists(path):
            with open(path) as f:
                self.entries = {k: tuple(v) + (0,) * (4 - len(v)) for k, v in json.load(f).items()}

    @staticmethod
    def key(plan: "Plan") -> Optional[str]:
        if plan.dobs is None or plan.dobs.n_tris == 0:
            return None

        def b2(x):
            return 1 << max(0, int(x) - 1).bit_length()
        o = plan.opts
        prec = [k for k, v in _lib.PRECISIONS.items() if v == o.precision][0]
        meth = [k for k, v in _lib.METHODS.items() if v == o.method][0]
        return f"v{_lib.lib().rvf_version()}|{prec}|{meth}|nq{o.nquad}|N{b2(plan.dmesh.mesh.n_elems)}|T{b2(plan.dobs.n_tris)}"

    def __contains__(self, key):
        return key in self.entries

    def __getitem__(self, key):
        return self.entries[key]

    def __setitem__(self, key, value):
        # (variant, refill_min, columns per tile, waves per SIMD; 0 = the kernel's default)
        value = tuple(value) + (0,) * (4 - len(value))
        self.entries[key] = (int(value[0]), int(value[1]), int(value[2]), int(value[3]))

    def save(self):
        d = os.path.dirname(os.path.abspath(self.path))
        os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump({k: list(v) for k, v in sorted(self.entries.items())}, f, indent=1)
        os.replace(tmp, self.path)


REQ_PENDING, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = range(5)


class Server:
    """Request server (include/rvf.h rvf_server_*): solves submitted from any
    thread are admitted by device-memory footprint and batched with the other
    running solves; each result equals the standalone solve's."""

    def __init__(self, ctx: Context, hbm_budget: int = 0, max_active: int = 0, round_tiles: int = 0,
                 direct_tiles: int = 0):
        o = _lib.ServerOpts(hbm_budget, max_active, round_tiles, direct_tiles, 0)
        h = C.c_void_p()
        check(_lib.lib().rvf_server_create(ctx.h, C.byref(o), C.byref(h)))
        self.ctx, self.h = ctx, h
        self._meta = {}

    def submit(self, mesh: MeshData, nquad: int, obstruction_groups: Sequence[int] = (), parity: str = "reference",
               precision: str = "fp64", want_F: bool = False, variant: int = 0, self_vf: bool = False) -> int:
        if mesh.mesh_dim == 1:
            raise ValueError("the server takes surface meshes; solve curve meshes with Plan or compute_view_factors")
        mats, counts, idx = family_matrices(mesh)
        arr = (C.c_void_p * 4)(*[m.ctypes.data if m.size else None for m in mats])
        tags = sorted(set(int(g) for g in obstruction_groups if int(g) in mesh.group_tri_soup))
        soup = tg = None
        T = 0
        if tags:
            soup = np.ascontiguousarray(np.concatenate([mesh.group_tri_soup[t] for t in tags]), np.float64)
            tg = np.concatenate([np.full(mesh.group_tri_soup[t].shape[0], t, np.int32) for t in tags])
            T = soup.shape[0]
        obs = np.ascontiguousarray(list(obstruction_groups), np.int32)
        mode = SOUP_READ["reference"] if parity == "reference" else SOUP_READ[("geometric", mesh.soup_layout)]
        o = _lib.Opts()
        o.nquad = nquad
        o.method = _lib.METHODS["quad"]
        o.precision = _lib.PRECISIONS[precision]
        o.flags = (RVF_WANT_F if want_F else 0) | (_lib.RVF_SELF_VF if self_vf else 0)
        o.variant = variant
        t = C.c_int64()
        check(_lib.lib().rvf_server_submit(self.h, ptr(mesh.coords), mesh.n_nodes, mesh.n_elems, ptr(mesh.family),
                                           ptr(idx), ptr(mesh.group), arr, ptr(counts), ptr(soup), ptr(tg), T,
                                           ptr(obs), len(obs), mode, C.byref(o), C.byref(t)))
        self._meta[t.value] = (mesh.n_elems, len(set(int(g) for g in mesh.group)), want_F)
        return t.value

    def poll(self, ticket: int) -> int:
        return check(_lib.lib().rvf_server_poll(self.h, ticket))

    def wait(self, ticket: int, timeout_ms: int = -1) -> int:
        return check(_lib.lib().rvf_server_wait(self.h, ticket, timeout_ms))

    def result(self, ticket: int):
        """(F or None, A, rowsum, F_group, A_group, tags, request stats dict)."""
        N, G, want_F = self._meta[ticket]
        F = np.empty((N, N), order="F") if want_F else None
        A = np.empty(N); rs = np.empty(N)
        Fg = np.empty((G, G), order="F"); Ag = np.empty(G); tags = np.empty(G, np.int32)
        st = _lib.RequestStats()
        check(_lib.lib().rvf_server_result(self.h, ticket, ptr(F), ptr(A), ptr(rs), ptr(Fg), ptr(Ag), ptr(tags),
                                           C.byref(st)))
        return F, A, rs, Fg, Ag, [int(x) for x in tags], st.as_dict()

    def request_stats(self, ticket: int):
        st = _lib.RequestStats()
        _lib.lib().rvf_server_result(self.h, ticket, None, None, None, None, None, None, C.byref(st))
        return st.as_dict()

    def cancel(self, ticket: int):
        check(_lib.lib().rvf_server_cancel(self.h, ticket))

    def release(self, ticket: int):
        check(_lib.lib().rvf_server_release(self.h, ticket))
        self._meta.pop(ticket, None)

    def stats(self):
        c = _lib.ServerCounters()
        check(_lib.lib().rvf_server_stats(self.h, C.byref(c)))
        return c.as_dict()

    def close(self):
        if self.h:
            _lib.lib().rvf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """Device buffers + launch sequence for one (mesh, obstruction, options)."""

    def __init__(self, ctx: Context, dmesh: DeviceMesh, dobs: Optional[DeviceObstruction], nquad: int,
                 method: str = "quad", precision: str = "fp64", n_samples: int = 10000, seed: int = 0,
                 want_F: bool = True, audit: bool = False, timing: bool = True, graph: bool = False,
                 variant: int = 0, self_vf: bool = False):
        o = _lib.Opts()
        o.nquad = nquad
        o.method = _lib.METHODS[method]
        o.precision = _lib.PRECISIONS[precision]
        o.n_samples = n_samples
        o.seed = seed & ((1 << 64) - 1)
        o.flags = (RVF_WANT_F if want_F else 0) | (RVF_AUDIT if audit else 0) | (RVF_TIMING if timing else 0) | \
            (_lib.RVF_GRAPH if graph else 0) | (_lib.RVF_SELF_VF if self_vf else 0)
        o.variant = variant
        self.opts = o
        self.ctx, self.dmesh, self.dobs = ctx, dmesh, dobs
        h = C.c_void_p()
        check(_lib.lib().rvf_plan_create(ctx.h, dmesh.h, dobs.h if dobs is not None else None, C.byref(o),
                                         C.byref(h)))
        self.h = h
        self.want_F, self.audit = want_F, audit and dobs is not None

    def tune(self, variant: int = 0, refill_min: int = 0):
        """Switch the traversal variant / wave-queue refill threshold (F, A and
        visibility decisions unchanged)."""
        check(_lib.lib().rvf_plan_tune(self.h, variant, refill_min))
        self.opts.variant = variant

    def tune_tiles(self, columns: int = 0):
        """Columns per pair tile, 1..16 (0: the kernel's default).  F, A and
        visibility decisions are unchanged; row and group sums are added in
        another order."""
        check(_lib.lib().rvf_plan_tune_tiles(self.h, columns))

    @property
    def columns(self) -> int:
        """Columns per pair tile of the plan's current layout."""
        st = _lib.Stats()
        check(_lib.lib().rvf_plan_stats(self.h, C.byref(st)))
        return int(st.tile_columns)

    def tune_waves(self, waves: int = 0):
        """Occupancy of the pair kernel: 0 = its default build; the mixed split
        traversal kernels also have a build for 3 waves per SIMD (4 = their
        default).  F, A and visibility decisions are unchanged."""
        check(_lib.lib().rvf_plan_tune_waves(self.h, waves))

    @property
    def waves(self) -> int:
        """The pair-kernel build in use (3: the three-wave build; 0: default)."""
        st = _lib.Stats()
        check(_lib.lib().rvf_plan_stats(self.h, C.byref(st)))
        return int(st.waves)

    def autotune(self, sample_tiles: int = 0, table: Optional["TuneTable"] = None):
        """Pick the fastest traversal configuration for this plan.

        With a table holding an entry for this plan's problem class, that
        entry is applied without measuring; otherwise the native autotuner
        times every variant over a tile sample (rvf_plan_autotune) and the
        winner is stored in the table.  Returns (variant, refill_min,
        measurements), measurements being [] for a table hit; the tile width
        and build chosen are `plan.columns` and `plan.waves`.  F, A and visibility decisions of later runs
        do not depend on the choice."""
        key = table.key(self) if table is not None else None
        if key is not None and key in table:
            v, r, c, w = table[key]
            self.tune(v, r)
            self.tune_tiles(c)
            self.tune_waves(w)
            return v, r, []
        ents = (_lib.TuneEntry * 32)()
        n, v, r = C.c_int32(), C.c_int32(), C.c_int32()
        check(_lib.lib().rvf_plan_autotune(self.h, sample_tiles, ents, 32, C.byref(n), C.byref(v), C.byref(r)))
        meas = [{"variant": e.variant, "refill_min": e.refill_min, "tiles": e.tiles, "ms": e.ms,
                 "columns": e.columns, "waves": e.waves} for e in ents[:n.value]]
        if not meas:
            return v.value, r.value, []
        self.opts.variant = v.value
        if table is not None and key is not None:
            table[key] = (v.value, r.value, self.columns, self.waves)
            table.save()
        return v.value, r.value, meas

    def kernel_name(self) -> str:
        """The pair kernel this plan launches, as profilers name it."""
        buf = C.create_string_buffer(128)
        check(_lib.lib().rvf_plan_kernel(self.h, buf, 128))
        return buf.value.decode()

    def run(self):
        check(_lib.lib().rvf_plan_run(self.h))

    def sync(self):
        check(_lib.lib().rvf_plan_sync(self.h))

    def stats(self):
        s = _lib.Stats()
        check(_lib.lib().rvf_plan_stats(self.h, C.byref(s)))
        return s.as_dict()

    def fetch(self, F=True, F_out: Optional[np.ndarray] = None):
        """Host copies of the results.  F_out: an N x N float64 Fortran-order
        array to fill instead of a new one (e.g. from pinned_empty, which makes
        the copy one DMA transfer)."""
        mesh = self.dmesh.mesh
        N = mesh.n_elems
        G = len(self.dmesh.tags)
        Fm = None
        if F and self.want_F:
            if F_out is not None:
                if F_out.shape != (N, N) or F_out.dtype != np.float64 or not F_out.flags.f_contiguous:
                    raise ValueError("F_out must be an N x N float64 Fortran-order array")
                Fm = F_out
            else:
                Fm = np.empty((N, N), order="F")
        A = np.empty(N); rs = np.empty(N)
        Fg = np.empty((G, G), order="F"); Ag = np.empty(G); tags = np.empty(G, np.int32)
        ac = np.empty((N, N), np.uint32, order="F") if self.audit else None
        ah = np.empty((N, N), np.uint32, order="F") if self.audit else None
        check(_lib.lib().rvf_plan_fetch(self.h, ptr(Fm), ptr(A), ptr(rs), ptr(Fg), ptr(Ag), ptr(tags), ptr(ac),
                                        ptr(ah)))
        return Fm, A, rs, Fg, Ag, [int(t) for t in tags], ac, ah

    def close(self):
        if self.h:
            _lib.lib().rvf_plan_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _PinnedBlock:
    """Owns page-locked host memory from rvf_host_alloc; freed with the
    last array that views it."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        check(_lib.lib().rvf_host_alloc(nbytes, C.byref(p)))
        self.ptr, self.nbytes = p.value, nbytes

    def __del__(self):
        if getattr(self, "ptr", None):
            _lib.lib().rvf_host_free(C.c_void_p(self.ptr))
            self.ptr = None


def pinned_empty(shape, dtype=np.float64, order: str = "F") -> np.ndarray:
    """An uninitialised array in page-locked host memory (no reference
    counterpart): Plan.fetch(F_out=...) into it copies F by DMA."""
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    blk = _PinnedBlock(max(n, 1))
    buf = (C.c_char * max(n, 1)).from_address(blk.ptr)
    buf._rvf_owner = blk  # the ctypes buffer keeps the block alive; the array keeps the buffer
    return np.ndarray(shape, dtype=dt, buffer=buf, order=order)


_CTX: Dict[int, Context] = {}


def context(device: int = 0) -> Context:
    if device not in _CTX:
        _CTX[device] = Context(device)
    return _CTX[device]


# ----------------------------------------------------------------------------
# reference-shaped API
# ----------------------------------------------------------------------------
def compute_view_factors(mesh: MeshData, nquad: int = 4, obstruction_groups: Sequence[int] = (),
                         backend: Optional[MI355XBackend] = None, self_vf: bool = False,
                         monte_carlo: bool = False, n_samples: int = 10000, rng=None, use_duffy: bool = False,
                         verbose: bool = True) -> ViewFactorResult:
    """GPU branch of compute_view_factors (src/Assembly.jl:64-96) on MI355X.  ``self_vf`` adds the
    pairs (i, i) (src/Assembly.jl:173,193), which the reference computes on the CPU only."""
    if backend is None:
        backend = MI355XBackend()
    if isinstance(backend, type):
        backend = backend()
    if not isinstance(backend, MI355XBackend):
        raise TypeError("only the MI355X backend is provided here; the CPU path is the reference's own")
    duffy = use_duffy or backend.use_duffy
    if duffy and monte_carlo:
        raise ValueError("use_duffy and monte_carlo cannot both be true.")
    method = "mc" if monte_carlo else ("duffy" if duffy else "quad")
    ctx = context(backend.device)
    N = mesh.n_elems
    if verbose:
        what = f"n_samples={n_samples} (Monte Carlo)" if monte_carlo else f"nquad={nquad}"
        print(f"GPU compute_view_factors: {N} elements, {what}, precision={backend.precision}, backend=MI355X")
    dmesh = DeviceMesh(ctx, mesh)
    dobs = None
    if len(obstruction_groups) > 0:
        dobs = DeviceObstruction(ctx, dmesh, obstruction_groups, backend.parity)
        if dobs.n_trees == 0:
            print("Warning: obstruction_groups specified but no triangle geometry found for those groups; "
                  "proceeding without obstruction checking.")
            dobs = None
    plan = Plan(ctx, dmesh, dobs, nquad, method, backend.precision, n_samples, backend.seed, want_F=True,
                audit=backend.audit, self_vf=self_vf)
    if backend.tune_table and dobs is not None and method != "mc":
        plan.autotune(table=TuneTable(backend.tune_table))
    plan.run()
    plan.sync()
    F, A, rs, Fg, Ag, tags, ac, ah = plan.fetch()
    names = [mesh.group_tags.get(t, str(t)) for t in tags]
    res = ViewFactorResult(F, A, Fg, Ag, tags, names, rs, plan.stats(), ac, ah)
    if verbose:
        print("  Row-sum check (element level) — max |Σⱼ Fᵢⱼ - 1| : ", float(np.max(np.abs(rs - 1.0))))
        print("  Row-sum check (group level)   — max |Σⱼ Fᵢⱼ - 1| : ", float(np.max(np.abs(Fg.sum(1) - 1.0))))
    plan.close(); dmesh.close()
    if dobs is not None:
        dobs.close()
    return res


def aggregate_by_group(result: ViewFactorResult, mesh: MeshData):
    """Re-aggregate element results to group level (src/Results.jl:56-93)."""
    tags = sorted(mesh.group_tags)
    ge = mesh.group_elems
    G = len(tags)
    A = result.A_elem
    Ag = np.array([A[ge[t]].sum() for t in tags])
    Fg = np.zeros((G, G))
    for a, ta in enumerate(tags):
        ia = ge[ta]
        for b, tb in enumerate(tags):
            Fg[a, b] = (A[ia] * result.F_elem[np.ix_(ia, ge[tb])].sum(1)).sum() / Ag[a]
    return Fg, Ag, tags, [mesh.group_tags[t] for t in tags]


def check_reciprocity(result: ViewFactorResult, tol: float = 1e-4) -> bool:
    """A_i F_ij ~ A_j F_ji (src/Results.jl:105-114)."""
    F, A = result.F_elem, result.A_elem
    AF = A[:, None] * F
    iu = np.triu_indices(F.shape[0], 1)
    err = np.abs(AF[iu] - AF.T[iu]) / np.maximum(AF[iu], 1e-30)
    m = float(err.max()) if err.size else 0.0
    print(f"Reciprocity max relative error: {m}")
    return m < tol


def check_closure(result: ViewFactorResult, tol: float = 1e-3) -> bool:
    """No row of F sums to more than 1 + tol (src/Results.jl:122-127)."""
    rs = result.F_elem.sum(1)
    print(f"Row sums: min={round(float(rs.min()), 6)}, max={round(float(rs.max()), 6)}")
    return float(rs.max()) <= 1.0 + tol


def quad_rule(nquad: int, tri: bool = False):
    """The quadrature table the device uses (host-computed, no GPU needed)."""
    pts = np.zeros(2 * 64); w = np.zeros(64)
    n = check(_lib.lib().rvf_quad_rule(nquad, 1 if tri else 0, ptr(pts), ptr(w)))
    return pts[:2 * n].reshape(n, 2).copy(), w[:n].copy()


def bvh_build_host(soup: np.ndarray, soup_read: int):
    soup = np.ascontiguousarray(soup, np.float64).reshape(-1, 3, 3)
    T = soup.shape[0]
    cap = 2 * T + 4
    box = np.zeros(6 * cap); meta = np.zeros(5 * cap, np.int32); tri = np.zeros(T, np.int32)
    nn = C.c_int64()
    check(_lib.lib().rvf_bvh_build_host(ptr(soup), T, soup_read, cap, ptr(box), ptr(meta), ptr(tri), C.byref(nn)))
    return box[:6 * nn.value].reshape(-1, 6), meta[:5 * nn.value].reshape(-1, 5), tri


__all__ = ["MI355XBackend", "ViewFactorResult", "compute_view_factors", "aggregate_by_group", "check_reciprocity",
           "check_closure", "Context", "DeviceMesh", "DeviceObstruction", "Plan", "RVFError", "quad_rule"]
