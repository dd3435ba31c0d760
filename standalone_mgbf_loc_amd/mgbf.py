"""ctypes binding of libmgbf (the C ABI in include/mgbf.h).

Mirrors the reference's operator surface for the localization path:
`Config` carries every key of the namelist &PARAMETERS_MGBETA
(src.mgbf/mg_parameter.f90:190-208); `MGBF(config)` is mg_initialize
(src.mgbf/mg_entrymod.f90:36-219) and `MGBF.close()` is mg_finalize
(:222-273); `MGBF.filtering_procedure(x)` applies W A W^T as
filtering_procedure(mgbf_proc) does (src.mgbf/mg_filtering.f90:67-120).
Fields are torch tensors on the HIP device, float32, shape (ny, nx, K) in C
order (k fastest), the memory order of the reference's VALL(k, i, j).

There is no CPU fallback: if libmgbf.so is missing or no HIP device is
present, the calls raise.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Optional

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libmgbf.so")

MGBF_OK = 0
(OP_WAWT, OP_WT, OP_W, OP_A, OP_ASQRT, OP_BSQRT, OP_BSQRTT, OP_A2F, OP_F2A, OP_A2F_ALL,
 OP_F2A_ALL) = range(11)
STAGE_VF, STAGE_VFT, STAGE_F, STAGE_FT, STAGE_R, STAGE_P, STAGE_A = range(7)

EXPORTED_SYMBOLS = (
    "mgbf_config_default", "mgbf_derive", "mgbf_init", "mgbf_finalize", "mgbf_get_dims",
    "mgbf_set_weights", "mgbf_apply", "mgbf_apply_stage", "mgbf_dot", "mgbf_last_error",
    "mgbf_version", "mgbf_timers_enable", "mgbf_timers_read", "mgbf_autotune", "mgbf_get_tile",
    "mgbf_set_tile", "mgbf_group_create", "mgbf_group_apply", "mgbf_group_destroy",
    "mgbf_apply_noalloc", "mgbf_set_aspect", "mgbf_timers_read_gen",
)
# floats of the static device scratch of mgbf_apply_noalloc (MGBF_NOALLOC_SCRATCH_FLOATS)
NOALLOC_SCRATCH_FLOATS = 32 << 20
TIMER_NAMES = ("vfiltT", "upsend", "hfiltT", "hfilt", "dnsend", "vfilt")


class MGBFError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"mgbf status {status}: {msg}")
        self.status = status


_DOUBLE_KEYS = ("mg_ampl01", "mg_ampl02", "mg_ampl03",
                "mg_weig1", "mg_weig2", "mg_weig3", "mg_weig4", "mg_weig5")
_INT_KEYS = ("hx", "hy", "hz", "p", "mgbf_line", "mgbf_proc", "n_ens", "lm_a", "lm", "km2", "km3",
             "l_loc", "l_loc_vertical", "l_filt_g1", "l_lin_vertical", "l_lin_horizontal",
             "l_quad_horizontal", "l_new_map", "l_vertical_filter", "ldelta", "lquart", "lhelm",
             "gm_max", "nm0", "mm0", "nxPE", "nyPE", "im_filt", "jm_filt")


class _CConfig(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in _DOUBLE_KEYS] + [(k, ctypes.c_int) for k in _INT_KEYS]


class _CDims(ctypes.Structure):
    _fields_ = [("gm", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("km", ctypes.c_int), ("km_all", ctypes.c_int), ("km3_all", ctypes.c_int),
                ("ntaps2", ctypes.c_int), ("ntaps1", ctypes.c_int),
                ("field_size", ctypes.c_longlong), ("mg_size", ctypes.c_longlong),
                ("na", ctypes.c_int), ("ma", ctypes.c_int), ("km_a_all", ctypes.c_int),
                ("anal_size", ctypes.c_longlong), ("ntaps3", ctypes.c_int)]


@dataclasses.dataclass
class Config:
    """Every key of &PARAMETERS_MGBETA (defaults: a 4-generation ensemble run)."""
    mg_ampl01: float = 16.0 / 7.0
    mg_ampl02: float = 49.0 / 8.0
    mg_ampl03: float = 1.0
    mg_weig1: float = 1.0
    mg_weig2: float = 0.5
    mg_weig3: float = 0.25
    mg_weig4: float = 0.125
    mg_weig5: float = 0.0625
    hx: int = 6
    hy: int = 6
    hz: int = 3
    p: int = 2
    mgbf_line: bool = False
    mgbf_proc: int = 6
    n_ens: int = 1
    lm_a: int = 32
    lm: int = 32
    km2: int = 0
    km3: int = 1
    l_loc: bool = False
    l_loc_vertical: bool = False
    l_filt_g1: bool = True
    l_lin_vertical: bool = False
    l_lin_horizontal: bool = False
    l_quad_horizontal: bool = False
    l_new_map: bool = True
    l_vertical_filter: bool = True
    ldelta: bool = False
    lquart: bool = False
    lhelm: bool = False
    gm_max: int = 4
    nm0: int = 192
    mm0: int = 192
    nxPE: int = 8
    nyPE: int = 8
    im_filt: int = 24
    jm_filt: int = 24

    @classmethod
    def for_grid(cls, nx: int, ny: int, lm: int, n_ens: int = 1, gm: int = 4, h: int = 6,
                 hz: int = 3, p: int = 2, weights=None, **kw) -> "Config":
        """A configuration whose global grid is nx x ny with gm generations.

        The reference derives the grid from nxPE*im_filt and caps gm by the PE
        grid (def_maxgen); this picks the smallest power-of-two PE grid that
        allows gm generations, and sets the aspect amplitudes so that the
        filter reach equals the halo (mg_ampl02 = (h+1)^2/(2p+4),
        mg_ampl01 = (hz+1)^2/(2p+3), and for the 3-D filter every
        configuration normalizes, mg_ampl03 = (min(h,hz)+1)^2/(2p+5)).
        """
        pe = 1
        while _def_maxgen(pe, pe) < gm:
            pe *= 2
        if nx % pe or ny % pe:
            raise ValueError(f"grid {nx}x{ny} not divisible by the {pe}x{pe} PE grid needed for gm={gm}")
        w = list(weights) if weights is not None else [2.0 ** -g for g in range(5)]
        w = (w + [w[-1]] * 5)[:5]
        c = cls(mg_ampl01=(hz + 1) ** 2 / (2 * p + 3), mg_ampl02=(h + 1) ** 2 / (2 * p + 4),
                mg_ampl03=(min(h, hz) + 1) ** 2 / (2 * p + 5),
                mg_weig1=w[0], mg_weig2=w[1], mg_weig3=w[2], mg_weig4=w[3], mg_weig5=w[4],
                hx=h, hy=h, hz=hz, p=p, n_ens=n_ens, lm_a=lm, lm=lm, gm_max=gm,
                nm0=nx, mm0=ny, nxPE=pe, nyPE=pe, im_filt=nx // pe, jm_filt=ny // pe)
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    def to_c(self) -> _CConfig:
        c = _CConfig()
        for k in _DOUBLE_KEYS:
            setattr(c, k, float(getattr(self, k)))
        for k in _INT_KEYS:
            setattr(c, k, int(getattr(self, k)))
        return c

    def weights(self):
        return [self.mg_weig1, self.mg_weig2, self.mg_weig3, self.mg_weig4, self.mg_weig5]


def _def_maxgen(nxm: int, nym: int) -> int:
    """def_maxgen (src.mgbf/mg_parameter.f90:578-611)."""
    def one(n):
        g = 1
        while True:
            n = (n + 1) // 2
            g += 1
            if n == 1:
                return g
    return min(one(nxm), one(nym))


_lib = None


def load_library(path: Optional[str] = None):
    """Loads libmgbf.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is missing: build it with `python -m standalone_mgbf_loc_amd.build`")
    L = ctypes.CDLL(p)
    cfgp = ctypes.POINTER(_CConfig)
    L.mgbf_config_default.argtypes = [cfgp]
    L.mgbf_derive.argtypes = [cfgp, ctypes.POINTER(_CDims)]
    L.mgbf_init.argtypes = [cfgp, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.mgbf_finalize.argtypes = [ctypes.c_void_p]
    L.mgbf_get_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(_CDims)]
    L.mgbf_set_weights.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_double]
    L.mgbf_apply.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p]
    L.mgbf_apply_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]
    if hasattr(L, "mgbf_set_aspect"):
        L.mgbf_set_aspect.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.mgbf_apply_stage.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p]
    L.mgbf_dot.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                           ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
    L.mgbf_timers_enable.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.mgbf_timers_read.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_double)]
    if hasattr(L, "mgbf_timers_read_gen"):
        L.mgbf_timers_read_gen.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_double)]
    L.mgbf_autotune.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.mgbf_get_tile.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.mgbf_set_tile.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int]
    L.mgbf_group_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_void_p)]
    L.mgbf_group_apply.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                   ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]
    L.mgbf_group_destroy.argtypes = [ctypes.c_void_p]
    L.mgbf_last_error.restype = ctypes.c_char_p
    L.mgbf_version.restype = ctypes.c_char_p
    if path is None:
        _lib = L
    return L


def _check(status: int):
    if status != MGBF_OK:
        raise MGBFError(status, load_library().mgbf_last_error().decode())


@dataclasses.dataclass
class Dims:
    gm: int
    nx: int
    ny: int
    km: int
    km_all: int
    km3_all: int
    ntaps2: int
    ntaps1: int
    field_size: int
    mg_size: int
    na: int
    ma: int
    km_a_all: int
    anal_size: int
    ntaps3: int

    def gen_shape(self, g: int):
        """(ny_g, nx_g, K) of generation g (1-based)."""
        return (self.ny >> (g - 1), self.nx >> (g - 1), self.km_all)

    def anal_shape(self):
        """(mm0, nm0, km_a_all): the analysis field WORKA in C order."""
        return (self.ma, self.na, self.km_a_all)


def derive(cfg: Config) -> Dims:
    """Validated derived sizes of a configuration (no device use)."""
    L = load_library()
    d = _CDims()
    c = cfg.to_c()
    _check(L.mgbf_derive(ctypes.byref(c), ctypes.byref(d)))
    return Dims(*[getattr(d, f) for f, _ in _CDims._fields_])


def op_shapes(dims: "Dims", op: int):
    """(input shape, output shape) of an operation (MGBF_OP_*); a multigrid
    vector is flat (mg_size,). The C ABI takes bare pointers, so the wrapper
    checks every buffer against these before a call."""
    fld, mg, an = dims.gen_shape(1), (dims.mg_size,), dims.anal_shape()
    table = {OP_WAWT: (fld, fld), OP_WT: (fld, mg), OP_W: (mg, fld), OP_A: (mg, mg),
             OP_ASQRT: (mg, mg), OP_BSQRT: (mg, fld), OP_BSQRTT: (fld, mg),
             OP_A2F: (an, fld), OP_A2F_ALL: (an, fld), OP_F2A: (fld, an), OP_F2A_ALL: (fld, an)}
    if op not in table:
        raise ValueError(f"unknown operation {op}")
    return table[op]


def stage_shapes(dims: "Dims", stage: int, gen: int):
    """(input shape, output shape) of a single stage (MGBF_STAGE_*) on
    generation gen (1-based)."""
    if not 1 <= gen <= dims.gm:
        raise ValueError(f"generation {gen} out of range 1..{dims.gm}")
    g = dims.gen_shape
    if stage in (STAGE_R, STAGE_P):
        if gen >= dims.gm:
            raise ValueError("no coarser generation")
        return (g(gen), g(gen + 1)) if stage == STAGE_R else (g(gen + 1), g(gen))
    if stage in (STAGE_VF, STAGE_VFT) and gen != 1:
        raise ValueError("the vertical filter acts on generation 1")
    if stage in (STAGE_VF, STAGE_VFT, STAGE_F, STAGE_FT, STAGE_A):
        return g(gen), g(gen)
    raise ValueError(f"unknown stage {stage}")


def _numel(shape) -> int:
    n = 1
    for s in shape:
        n *= int(s)
    return n


def check_sizes(what: str, x, out, ishape, oshape):
    """Raises ValueError unless x (and out, when given) hold exactly the
    elements of the expected shapes."""
    if x.numel() != _numel(ishape):
        raise ValueError(f"{what}: input has {x.numel()} elements, expected {_numel(ishape)} "
                         f"(shape {tuple(ishape)})")
    if out is not None and out.numel() != _numel(oshape):
        raise ValueError(f"{what}: output has {out.numel()} elements, expected {_numel(oshape)} "
                         f"(shape {tuple(oshape)})")


def _ptr(t) -> int:
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor")
    if t.device.type != "cuda" or t.dtype != torch.float32 or not t.is_contiguous():
        raise ValueError("fields must be contiguous float32 tensors on the HIP device")
    return t.data_ptr()


def _stream(stream) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class MGBF:
    """One operator context (mg_initialize ... mg_finalize) on one device."""

    def __init__(self, cfg: Config, device: int = 0):
        self.cfg = cfg
        self.device = device
        self._L = load_library()
        c = cfg.to_c()
        h = ctypes.c_void_p()
        _check(self._L.mgbf_init(ctypes.byref(c), int(device), ctypes.byref(h)))
        self._h = h
        d = _CDims()
        _check(self._L.mgbf_get_dims(self._h, ctypes.byref(d)))
        self.dims = Dims(*[getattr(d, f) for f, _ in _CDims._fields_])
        self._wkeep = {}

    def close(self):
        if getattr(self, "_h", None):
            self._L.mgbf_finalize(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- shapes ---
    def field_shape(self):
        return self.dims.gen_shape(1)

    def mg_numel(self) -> int:
        return self.dims.mg_size

    def split_mg(self, v):
        """Views of the generation fields of a flat multigrid vector."""
        out, off = [], 0
        for g in range(1, self.dims.gm + 1):
            s = self.dims.gen_shape(g)
            n = s[0] * s[1] * s[2]
            out.append(v[off:off + n].view(s))
            off += n
        return out

    # --- weights ---
    def set_weights(self, gen: int, field=None, value: Optional[float] = None):
        """Per-point weight field (a_diff) or constant for generation gen (1-based)."""
        if field is not None:
            if tuple(field.shape) != self.dims.gen_shape(gen):
                raise ValueError("weight field shape mismatch")
            self._wkeep[gen] = field
            _check(self._L.mgbf_set_weights(self._h, gen, _ptr(field), 0.0))
        else:
            self._wkeep.pop(gen, None)
            v = value if value is not None else self.cfg.weights()[min(gen, 5) - 1]
            _check(self._L.mgbf_set_weights(self._h, gen, None, float(v)))

    def set_aspect(self, gen: int, pasp2=None):
        """Per-point aspect tensors of generation gen (1-based): pasp2 is an
        array (m_g, n_g, 3) of (A11, A22, A12) per point, the tensor
        def_mg_weights builds from p_del and p_eps (mg_intstate.f90:383-390),
        or None for the namelist's mg_ampl02 * I."""
        if pasp2 is None:
            _check(self._L.mgbf_set_aspect(self._h, gen, None))
            return
        import numpy as np
        ny, nx, _ = self.dims.gen_shape(gen)
        a = np.ascontiguousarray(pasp2, dtype=np.float64)
        if a.shape != (ny, nx, 3):
            raise ValueError(f"aspect field must have shape {(ny, nx, 3)}")
        _check(self._L.mgbf_set_aspect(self._h, gen, a.ctypes.data_as(ctypes.c_void_p)))

    # --- sizes: the C ABI takes bare pointers, so every call checks here ---
    def op_shapes(self, op: int):
        return op_shapes(self.dims, op)

    def stage_shapes(self, stage: int, gen: int):
        return stage_shapes(self.dims, stage, gen)

    # --- operators ---
    def apply(self, op: int, x, out=None, stream=None):
        import torch
        ishape, oshape = op_shapes(self.dims, op)
        check_sizes(f"operation {op}", x, out, ishape, oshape)
        if out is None:
            out = torch.empty(oshape, dtype=torch.float32, device=x.device)
        _check(self._L.mgbf_apply(self._h, op, _ptr(x), _ptr(out), _stream(stream)))
        return out

    def apply_noalloc(self, op: int, x, out, stream=None):
        """apply() through mgbf_apply_noalloc: the workspaces live in the
        library's static device scratch, so the call allocates nothing (out
        must be given). Calls that overlap on one device race."""
        if out is None:
            raise ValueError("apply_noalloc writes into a caller-given output")
        ishape, oshape = op_shapes(self.dims, op)
        check_sizes(f"operation {op}", x, out, ishape, oshape)
        _check(self._L.mgbf_apply_noalloc(self._h, op, _ptr(x), _ptr(out), _stream(stream)))
        return out

    def filtering_procedure(self, x, out=None, stream=None):
        """y = W A W^T x (filtering_rad2_ens / filtering_rad2_bckg)."""
        return self.apply(OP_WAWT, x, out, stream)

    def anal_to_filt(self, w, out=None, stream=None):
        """Analysis field (mm0, nm0, km_a_all) -> generation-1 field
        (anal_to_filt_all2, src.mgbf/mg_transfer.f90:154-183)."""
        return self.apply(OP_A2F, w, out, stream)

    def filt_to_anal(self, v, out=None, stream=None):
        """Generation-1 field -> analysis field (filt_to_anal_all2,
        src.mgbf/mg_transfer.f90:186-203)."""
        return self.apply(OP_F2A, v, out, stream)

    def anal_to_filt_all(self, w, out=None, stream=None):
        """anal_to_filt_all (src.mgbf/mg_transfer.f90:58-110): the vertical
        map of l_lin_vertical (linear) or the lwq one; needs km2 = 0."""
        return self.apply(OP_A2F_ALL, w, out, stream)

    def filt_to_anal_all(self, v, out=None, stream=None):
        """filt_to_anal_all (src.mgbf/mg_transfer.f90:113-151)."""
        return self.apply(OP_F2A_ALL, v, out, stream)

    def localize_analysis(self, w, stream=None):
        """The driver's sequence on an analysis field (RBETA_TEST.f90:34-126):
        anal_to_filt_all2, filtering_procedure, filt_to_anal_all2 with
        l_new_map, the _all routines without."""
        new = self.cfg.l_new_map
        v = self.apply(OP_A2F if new else OP_A2F_ALL, w, stream=stream)
        self.apply(OP_WAWT, v, v, stream)
        return self.apply(OP_F2A if new else OP_F2A_ALL, v, stream=stream)

    def stage(self, stage: int, gen: int, x, out=None, stream=None):
        import torch
        ishape, oshape = stage_shapes(self.dims, stage, gen)
        check_sizes(f"stage {stage} on generation {gen}", x, out, ishape, oshape)
        if out is None:
            out = torch.empty(oshape, dtype=torch.float32, device=x.device)
        _check(self._L.mgbf_apply_stage(self._h, stage, gen, _ptr(x), _ptr(out), _stream(stream)))
        return out

    def autotune(self, stream=None) -> dict:
        """Times the candidate radial tiles on this context's shapes and keeps
        the fastest; returns {(gen, kind): (TI, TJ)}. TI = 0 for (1, 2): the
        vertical band runs as a pass of its own after the kind-1 pass, which
        was faster than the band fused into it."""
        _check(self._L.mgbf_autotune(self._h, _stream(stream)))
        return self.tiles()

    def tiles(self) -> dict:
        out = {}
        ti, tj = ctypes.c_int(), ctypes.c_int()
        for g in range(1, self.dims.gm + 1):
            for kind in range(4):
                _check(self._L.mgbf_get_tile(self._h, g, kind, ctypes.byref(ti), ctypes.byref(tj)))
                out[(g, kind)] = (ti.value, tj.value)
        return out

    def set_tiles(self, tiles: dict):
        """Pins tiles {(gen, kind): (TI, TJ)} as returned by autotune()/tiles()
        (TI = 16: the matrix cores; TI = 0 for (1, 2): the vertical band as a
        pass of its own)."""
        for (g, kind), (ti, tj) in tiles.items():
            _check(self._L.mgbf_set_tile(self._h, int(g), int(kind), int(ti), int(tj)))

    def timers(self, enable: bool = True):
        """Turns the per-phase device timers on or off (and clears them)."""
        _check(self._L.mgbf_timers_enable(self._h, int(enable)))

    def read_timers(self) -> dict:
        """{phase: (ms, launches, algorithmic bytes)} accumulated since enabling."""
        n = len(TIMER_NAMES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_longlong * n)()
        by = (ctypes.c_double * n)()
        _check(self._L.mgbf_timers_read(self._h, ms, cnt, by))
        return {name: (ms[i], cnt[i], by[i]) for i, name in enumerate(TIMER_NAMES)}

    def read_timers_gen(self, gen: int) -> dict:
        """read_timers() restricted to the launches on generation `gen` (1..gm)."""
        n = len(TIMER_NAMES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_longlong * n)()
        by = (ctypes.c_double * n)()
        _check(self._L.mgbf_timers_read_gen(self._h, int(gen), ms, cnt, by))
        return {name: (ms[i], cnt[i], by[i]) for i, name in enumerate(TIMER_NAMES)}

    def dot(self, a, b, stream=None) -> float:
        """FP64-accumulated <a, b> on the device."""
        if a.numel() != b.numel():
            raise ValueError("size mismatch")
        r = ctypes.c_double()
        _check(self._L.mgbf_dot(self._h, _ptr(a), _ptr(b), a.numel(), ctypes.byref(r),
                                _stream(stream)))
        return r.value


class Group:
    """Several contexts (regional domains of uneven size) applied as one
    hipGraph, captured on first use with a given set of buffers."""

    def __init__(self, ops):
        self.ops = list(ops)
        self._L = load_library()
        arr = (ctypes.c_void_p * len(self.ops))(*[o._h.value for o in self.ops])
        h = ctypes.c_void_p()
        _check(self._L.mgbf_group_create(arr, len(self.ops), ctypes.byref(h)))
        self._h = h

    def apply(self, op: int, xs, outs, stream=None):
        n = len(self.ops)
        if len(xs) != n or len(outs) != n:
            raise ValueError("one input and one output per domain")
        for o, x, y in zip(self.ops, xs, outs):
            ishape, oshape = op_shapes(o.dims, op)
            check_sizes(f"operation {op}", x, y, ishape, oshape)
        ins = (ctypes.c_void_p * n)(*[_ptr(x) for x in xs])
        ous = (ctypes.c_void_p * n)(*[_ptr(y) for y in outs])
        _check(self._L.mgbf_group_apply(self._h, op, ins, ous, _stream(stream)))
        return outs

    def close(self):
        if getattr(self, "_h", None):
            self._L.mgbf_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def version() -> str:
    return load_library().mgbf_version().decode()
