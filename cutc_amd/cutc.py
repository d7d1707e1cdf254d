"""Host-side mirror of the reference's cutc_int interface over libcutcint.so.

``cutc_int`` takes the same arguments with the same meaning as the reference's
Fortran binding (``src/cutc_module.f90:10-37``, with the 6-integer launch
geometry of the C definition ``src/cutc_int.c:15-20``) and returns the same
two arrays; the device-pointer stage functions mirror
``get_int2_grad1_u12_ao`` / ``get_int_2e_ao`` (``include/cutc_int.h:8-14``).
Errors raise :class:`CutcError` carrying the library's negative status code.

Everything here calls the HIP library; there is no CPU fallback.  torch is
used only to own device memory and to hand over a stream.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _lib

CUTC_OK = 0
CUTC_ERR_GEOMETRY = -1
CUTC_ERR_ARG = -2
CUTC_ERR_NOMEM = -3
CUTC_ERR_HIP = -4
CUTC_ERR_UNSUPPORTED = -5

PREC_FP64 = 0
PREC_BF16 = 1
PREC_FP16 = 2
PREC_INT8 = 3
PREC_BF16_CRT = 4
PREC_FP16_CRT = 5
PRECISIONS = {"fp64": PREC_FP64, "bf16": PREC_BF16, "fp16": PREC_FP16, "int8": PREC_INT8, "bf16crt": PREC_BF16_CRT,
              "fp16crt": PREC_FP16_CRT}
# what cutc_options.slices = 0 resolves to in the library (cutc_default_slices)
DEFAULT_SLICES = {PREC_FP64: 0, PREC_BF16: 4, PREC_FP16: 4, PREC_INT8: 11, PREC_BF16_CRT: 10, PREC_FP16_CRT: 10}


class CutcError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__(f"{where}: status {code} ({_error_string(code)})")


class Dim3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint), ("y", ctypes.c_uint), ("z", ctypes.c_uint)]


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("slices", ctypes.c_int), ("verbose", ctypes.c_int),
                ("n1_chunk", ctypes.c_int), ("mem_limit", ctypes.c_size_t), ("host_copy", ctypes.c_int),
                ("streams", ctypes.c_int), ("dense", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_jastrow", ctypes.c_double),
                ("ms_integrands", ctypes.c_double), ("ms_gemm1", ctypes.c_double), ("ms_gemm2", ctypes.c_double),
                ("ms_symm", ctypes.c_double), ("ms_split", ctypes.c_double), ("n_gemm1", ctypes.c_int),
                ("n_gemm2", ctypes.c_int), ("n_jastrow", ctypes.c_int), ("n1_chunk", ctypes.c_int64),
                ("n1_chunks", ctypes.c_int), ("flops_exec", ctypes.c_double), ("workspace_bytes", ctypes.c_size_t),
                ("ms_host_in", ctypes.c_double), ("ms_host_out", ctypes.c_double),
                ("ms_long_range", ctypes.c_double), ("ms_short_range", ctypes.c_double),
                ("ms_stage1", ctypes.c_double), ("ms_stage2", ctypes.c_double),
                ("k_density", ctypes.c_double), ("k_density_2e", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class System(ctypes.Structure):
    """one system of a batch (cutc_system): device pointers"""
    _fields_ = [("n_grid1", ctypes.c_int), ("n_grid2", ctypes.c_int), ("n_ao", ctypes.c_int), ("n_nuc", ctypes.c_int),
                ("size_bh", ctypes.c_int)] + [(n, ctypes.c_void_p) for n in (
                    "r1", "wr1", "r2", "wr2", "rn", "aos_data1", "aos_data2", "c_bh", "m_bh", "n_bh", "o_bh",
                    "int_2e_ao")]


class BatchStats(ctypes.Structure):
    _fields_ = [("ms_setup", ctypes.c_double), ("ms_tune", ctypes.c_double), ("ms_launch", ctypes.c_double),
                ("n_nodes", ctypes.c_int64), ("n_systems", ctypes.c_int), ("n_split", ctypes.c_int),
                ("n_branches", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class NoStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("ms_total", "ms_point", "ms_pair", "ms_gemm2e", "ms_1e", "ms_symm",
                                              "flops_gemm2e")] + [
        ("chunk", ctypes.c_int64), ("n_chunks", ctypes.c_int64), ("workspace_bytes", ctypes.c_size_t)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class MoStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("ms_total", "ms_left", "ms_right", "flops")] + [
        ("mo_chunk", ctypes.c_int64), ("n_chunks", ctypes.c_int64), ("workspace_bytes", ctypes.c_size_t)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Plan(ctypes.Structure):
    _fields_ = [("n1_chunk", ctypes.c_int64), ("ldk", ctypes.c_int64), ("n_pairs", ctypes.c_int64),
                ("n1_chunk_2e", ctypes.c_int64), ("bytes_stage1", ctypes.c_size_t),
                ("bytes_stage2", ctypes.c_size_t), ("k_chunk", ctypes.c_int64)]


# every C symbol the public header declares
HEADER_SYMBOLS = [
    "cutc_int", "get_int2_grad1_u12_ao", "get_int_2e_ao", "checkCudaErrors", "checkCublasErrors", "cutc_int_bh",
    "int_long_range", "int_short_range_herm", "int_short_range_nonherm", "trans_inplace", "cutc_get_last_error",
    "cutc_error_string", "cutc_default_options", "cutc_int_device", "cutc_int_ex", "cutc_make_plan",
    "cutc_release_workspace", "cutc_int_streamed_device", "cutc_int_batch_device", "cutc_no", "cutc_no_device",
    "cutc_ao_to_mo", "cutc_ao_to_mo_device", "cutc_aos", "cutc_aos_device", "cutc_int_device_noalloc",
]
NOALLOC_SCRATCH_BYTES = 256 << 20

_vp = ctypes.c_void_p
_ci = ctypes.c_int
_L = None


def lib():
    """The loaded libcutcint.so (raises if it is not built)."""
    global _L
    if _L is None:
        # One HIP runtime per process.  PyTorch-ROCm ships its own
        # libamdhip64 and asks for it as "libamdhip64.so"; the library asks
        # for "libamdhip64.so.7".  With torch loaded first the library binds
        # to torch's copy (same soname); loaded the other way round, torch
        # brings a second runtime that sees no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(_lib.lib_path("libcutcint.so"))
        L.cutc_int.restype = _ci
        L.cutc_int.argtypes = [_ci] * 11 + [_vp] * 13
        L.cutc_int_ex.restype = _ci
        L.cutc_int_ex.argtypes = [ctypes.POINTER(Options)] + [_ci] * 5 + [_vp] * 13 + [ctypes.POINTER(Stats)]
        L.cutc_int_batch_device.restype = _ci
        L.cutc_int_batch_device.argtypes = [ctypes.POINTER(Options), _ci, ctypes.POINTER(System), _vp,
                                            ctypes.POINTER(BatchStats)]
        L.cutc_int_streamed_device.restype = _ci
        L.cutc_int_streamed_device.argtypes = ([ctypes.POINTER(Options)] + [_ci] * 5 + [_vp] * 12 + [_vp]
                                               + [ctypes.POINTER(Stats)])
        L.cutc_int_device.restype = _ci
        L.cutc_int_device.argtypes = ([ctypes.POINTER(Options)] + [_ci] * 5 + [_vp] * 13 + [_vp]
                                      + [ctypes.POINTER(Stats)])
        L.cutc_int_device_noalloc.restype = _ci
        L.cutc_int_device_noalloc.argtypes = [ctypes.POINTER(Options)] + [_ci] * 5 + [_vp] * 13 + [_vp]
        L.get_int2_grad1_u12_ao.restype = None
        L.get_int2_grad1_u12_ao.argtypes = [Dim3, Dim3] + [_ci] * 5 + [_vp] * 10
        L.get_int_2e_ao.restype = None
        L.get_int_2e_ao.argtypes = [_ci, _ci] + [_vp] * 4
        L.cutc_int_bh.restype = None
        L.cutc_int_bh.argtypes = [Dim3, Dim3] + [_ci] * 8 + [_vp] * 8
        L.int_long_range.restype = None
        L.int_long_range.argtypes = [_ci] * 5 + [_vp] * 3
        L.int_short_range_herm.restype = None
        L.int_short_range_herm.argtypes = [_ci, _ci, _vp, _vp, _vp]
        L.int_short_range_nonherm.restype = None
        L.int_short_range_nonherm.argtypes = [_ci, _ci, _vp, _vp, _vp]
        L.trans_inplace.restype = None
        L.trans_inplace.argtypes = [_vp, _ci]
        L.cutc_get_last_error.restype = _ci
        L.cutc_error_string.restype = ctypes.c_char_p
        L.cutc_error_string.argtypes = [_ci]
        L.cutc_default_options.restype = None
        L.cutc_default_options.argtypes = [ctypes.POINTER(Options)]
        L.cutc_release_workspace.restype = None
        L.cutc_release_workspace.argtypes = []
        L.cutc_no.restype = _ci
        L.cutc_no.argtypes = [_ci] * 4 + [_vp] * 7
        L.cutc_no_device.restype = _ci
        L.cutc_no_device.argtypes = [_ci] * 4 + [_vp] * 7 + [ctypes.c_size_t, ctypes.c_int64, _vp,
                                                             ctypes.POINTER(NoStats)]
        L.cutc_ao_to_mo.restype = _ci
        L.cutc_ao_to_mo.argtypes = [_ci] * 3 + [_vp] * 4
        L.cutc_ao_to_mo_device.restype = _ci
        L.cutc_ao_to_mo_device.argtypes = [_ci] * 3 + [_vp] * 4 + [ctypes.c_size_t, ctypes.c_int64, _vp,
                                                                   ctypes.POINTER(MoStats)]
        L.cutc_aos.restype = _ci
        L.cutc_aos.argtypes = [_ci] * 4 + [_vp] * 8
        L.cutc_aos_device.restype = _ci
        L.cutc_aos_device.argtypes = [_ci] * 4 + [_vp] * 9
        L.cutc_make_plan.restype = _ci
        L.cutc_make_plan.argtypes = [ctypes.c_int64, ctypes.c_int64, _ci, _ci, _ci, _ci, ctypes.c_size_t,
                                     ctypes.c_int64, ctypes.POINTER(Plan)]
        _L = L
    return _L


def _error_string(code):
    try:
        return lib().cutc_error_string(code).decode()
    except Exception:  # library not loadable: keep the message usable
        return "unknown"


def _check(rc, where):
    if rc != CUTC_OK:
        raise CutcError(rc, where)


def options(precision="fp64", slices=None, verbose=False, n1_chunk=0, mem_limit=0, host_copy=0, streams=0,
            dense=0) -> Options:
    """slices: bf16/fp16 digit slices per operand, or the number of moduli
    for int8 / bf16crt / fp16crt (None: the precision's default,
    DEFAULT_SLICES).  host_copy (host entry points): 0 = pipelined download
    during the build, 1 = blocking copies after it.  dense (modular
    precisions): 1 = the stage-1 contraction runs every K step instead of
    skipping the all-zero tiles of the pair operand (same results)."""
    o = Options()
    lib().cutc_default_options(ctypes.byref(o))
    o.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    o.slices = DEFAULT_SLICES[o.precision] if slices is None else slices
    o.verbose = int(verbose)
    o.n1_chunk = n1_chunk
    o.mem_limit = mem_limit
    o.host_copy = int(host_copy)
    o.streams = int(streams)
    o.dense = int(dense)
    return o


def make_plan(n1, n2, n_ao, n_nuc, budget, precision="fp64", slices=None, n1_chunk=0) -> Plan:
    p = Plan()
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    slices = DEFAULT_SLICES[prec] if slices is None else slices
    _check(lib().cutc_make_plan(n1, n2, n_ao, n_nuc, prec, slices, budget, n1_chunk, ctypes.byref(p)),
           "cutc_make_plan")
    return p


def _np_ptr(a, dtype):
    assert a.dtype == dtype and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(ctypes.c_void_p)


def _host_inputs(sys):
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    i = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    return dict(r1=f(sys.r1), wr1=f(sys.wr1), r2=f(sys.r2), wr2=f(sys.wr2), rn=f(sys.rn), aos1=f(sys.aos1),
                aos2=f(sys.aos2), c=f(sys.c_bh), m=i(sys.m_bh), n=i(sys.n_bh), o=i(sys.o_bh))


def cutc_int(sys, geometry=(100, 1, 1, 32, 1, 1)):
    """Reference entry point on host arrays.  Returns (X, int_2e_ao):
    X numpy (3, n1, n_ao, n_ao) = Fortran int2_grad1_u12_ao(n_ao, n_ao, n1, 3),
    int_2e_ao numpy (n_ao,)*4 = Fortran int_2e_ao(n_ao, n_ao, n_ao, n_ao)."""
    h = _host_inputs(sys)
    X = np.zeros((3, sys.n1, sys.n_ao, sys.n_ao))
    I = np.zeros((sys.n_ao,) * 4)
    rc = lib().cutc_int(*geometry, sys.n1, sys.n2, sys.n_ao, sys.n_nuc, sys.size_bh,
                        *(_np_ptr(h[k], np.float64) for k in ("r1", "wr1", "r2", "wr2", "rn", "aos1", "aos2", "c")),
                        *(_np_ptr(h[k], np.int32) for k in ("m", "n", "o")),
                        _np_ptr(X, np.float64), _np_ptr(I, np.float64))
    _check(rc, "cutc_int")
    return X, I


def cutc_int_ex(sys, opt: Options | None = None):
    """cutc_int with options; returns (X[0:3], int_2e_ao, stats dict)."""
    h = _host_inputs(sys)
    X = np.zeros((3, sys.n1, sys.n_ao, sys.n_ao))
    I = np.zeros((sys.n_ao,) * 4)
    st = Stats()
    rc = lib().cutc_int_ex(ctypes.byref(opt or options()), sys.n1, sys.n2, sys.n_ao, sys.n_nuc, sys.size_bh,
                           *(_np_ptr(h[k], np.float64) for k in ("r1", "wr1", "r2", "wr2", "rn", "aos1", "aos2", "c")),
                           *(_np_ptr(h[k], np.int32) for k in ("m", "n", "o")),
                           _np_ptr(X, np.float64), _np_ptr(I, np.float64), ctypes.byref(st))
    _check(rc, "cutc_int_ex")
    return X, I, st.as_dict()


# ---------------------------------------------------------------- device side

def _tp(t):
    """Device pointer of a contiguous torch tensor (None -> NULL)."""
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous()
    return ctypes.c_void_p(t.data_ptr())


@dataclasses.dataclass
class DeviceSystem:
    """Inputs resident in HBM (torch tensors on cuda), same layouts as the host arrays."""
    r1: object
    wr1: object
    r2: object
    wr2: object
    rn: object
    aos1: object
    aos2: object
    c_bh: object
    m_bh: object
    n_bh: object
    o_bh: object
    n1: int
    n2: int
    n_ao: int
    n_nuc: int
    size_bh: int

    @classmethod
    def upload(cls, sys, device="cuda"):
        import torch
        h = _host_inputs(sys)
        d = {k: torch.from_numpy(v).to(device) for k, v in h.items()}
        return cls(d["r1"], d["wr1"], d["r2"], d["wr2"], d["rn"], d["aos1"], d["aos2"], d["c"], d["m"], d["n"],
                   d["o"], sys.n1, sys.n2, sys.n_ao, sys.n_nuc, sys.size_bh)

    @classmethod
    def upload_with_device_aos(cls, sys, device="cuda"):
        """Upload everything but the AO values, which cutc_aos_device
        evaluates in HBM from ``sys.basis`` (no aos_data transfer)."""
        import torch
        h = _host_inputs(sys)
        d = {k: torch.from_numpy(v).to(device) for k, v in h.items() if k not in ("aos1", "aos2")}
        db = upload_basis(sys.basis, device)
        for k, r in (("aos1", d["r1"]), ("aos2", d["r2"])):
            d[k] = torch.empty((4, sys.n_ao, r.shape[0]), dtype=torch.float64, device=device)
            cutc_aos_device(r, d["rn"], db, d[k])
        return cls(d["r1"], d["wr1"], d["r2"], d["wr2"], d["rn"], d["aos1"], d["aos2"], d["c"], d["m"], d["n"],
                   d["o"], sys.n1, sys.n2, sys.n_ao, sys.n_nuc, sys.size_bh)


def alloc_outputs(ds, device="cuda"):
    """(X, int_2e_ao) device buffers: X (4, n1, n_ao, n_ao), int_2e_ao (n_ao,)*4."""
    import torch
    X = torch.empty((4, ds.n1, ds.n_ao, ds.n_ao), dtype=torch.float64, device=device)
    I = torch.empty((ds.n_ao,) * 4, dtype=torch.float64, device=device)
    return X, I


def cutc_int_device(ds, X, I, opt: Options | None = None, stream=None):
    """Whole path on resident inputs, on ``stream`` (a torch stream or None
    for torch's current stream).  Returns the library's stats dict."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    st = Stats()
    rc = lib().cutc_int_device(ctypes.byref(opt or options()), ds.n1, ds.n2, ds.n_ao, ds.n_nuc, ds.size_bh,
                               _tp(ds.r1), _tp(ds.wr1), _tp(ds.r2), _tp(ds.wr2), _tp(ds.rn), _tp(ds.aos1),
                               _tp(ds.aos2), _tp(ds.c_bh), _tp(ds.m_bh), _tp(ds.n_bh), _tp(ds.o_bh), _tp(X), _tp(I),
                               ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st))
    _check(rc, "cutc_int_device")
    return st.as_dict()


def cutc_int_device_noalloc(ds, X, I, opt: Options | None = None, stream=None):
    """cutc_int_device with every scratch buffer in the library's fixed
    device array (no allocation); no statistics."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    rc = lib().cutc_int_device_noalloc(ctypes.byref(opt or options()), ds.n1, ds.n2, ds.n_ao, ds.n_nuc, ds.size_bh,
                                       _tp(ds.r1), _tp(ds.wr1), _tp(ds.r2), _tp(ds.wr2), _tp(ds.rn), _tp(ds.aos1),
                                       _tp(ds.aos2), _tp(ds.c_bh), _tp(ds.m_bh), _tp(ds.n_bh), _tp(ds.o_bh), _tp(X),
                                       _tp(I), ctypes.c_void_p(stream.cuda_stream))
    _check(rc, "cutc_int_device_noalloc")


def cutc_int_streamed_device(ds, I, opt: Options | None = None, stream=None):
    """Streamed whole path: int_2e_ao only, X built and contracted r1 chunk by
    r1 chunk (opt.n1_chunk, when set, caps the chunk).  Returns the stats dict."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    st = Stats()
    rc = lib().cutc_int_streamed_device(ctypes.byref(opt or options()), ds.n1, ds.n2, ds.n_ao, ds.n_nuc, ds.size_bh,
                                        _tp(ds.r1), _tp(ds.wr1), _tp(ds.r2), _tp(ds.wr2), _tp(ds.rn), _tp(ds.aos1),
                                        _tp(ds.aos2), _tp(ds.c_bh), _tp(ds.m_bh), _tp(ds.n_bh), _tp(ds.o_bh),
                                        _tp(I), ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st))
    _check(rc, "cutc_int_streamed_device")
    return st.as_dict()


def cutc_int_batch_device(dss, Is, opt: Options | None = None, stream=None):
    """Batched build of the DeviceSystems ``dss`` into the int_2e_ao tensors
    ``Is``: one hipGraph launch for the whole batch.  Returns the batch
    statistics dict."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    arr = (System * max(1, len(dss)))()
    for y, ds, I in zip(arr, dss, Is):
        y.n_grid1, y.n_grid2, y.n_ao, y.n_nuc, y.size_bh = ds.n1, ds.n2, ds.n_ao, ds.n_nuc, ds.size_bh
        for f, t in (("r1", ds.r1), ("wr1", ds.wr1), ("r2", ds.r2), ("wr2", ds.wr2), ("rn", ds.rn),
                     ("aos_data1", ds.aos1), ("aos_data2", ds.aos2), ("c_bh", ds.c_bh), ("m_bh", ds.m_bh),
                     ("n_bh", ds.n_bh), ("o_bh", ds.o_bh), ("int_2e_ao", I)):
            setattr(y, f, None if t is None else t.data_ptr())
    st = BatchStats()
    rc = lib().cutc_int_batch_device(ctypes.byref(opt or options()), len(dss), arr,
                                     ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st))
    _check(rc, "cutc_int_batch_device")
    return st.as_dict()


def cutc_no(ne_a, ne_b, wr1, mos_l, mos_r, G):
    """Normal ordering on host arrays (the reference's cutc_no).  Inputs in
    numpy C order of the Fortran layouts: wr1 (n_grid1,), mos_l / mos_r
    (n_mo, n_grid1), G (n_mo, n_mo, 3, n_grid1) = int2_grad1_u12(n_grid1, 3,
    n_mo, n_mo).  Returns (no_2e [t, s, q, p], no_1e [q, p], no_0e)."""
    G = np.ascontiguousarray(G, dtype=np.float64)
    nm, ng = G.shape[0], G.shape[3]
    w = np.ascontiguousarray(wr1, dtype=np.float64)
    ml = np.ascontiguousarray(mos_l, dtype=np.float64)
    mr = np.ascontiguousarray(mos_r, dtype=np.float64)
    o2 = np.zeros((nm,) * 4)
    o1 = np.zeros((nm, nm))
    o0 = np.zeros(1)
    rc = lib().cutc_no(ng, nm, int(ne_a), int(ne_b), *(a.ctypes.data for a in (w, ml, mr, G, o2, o1, o0)))
    _check(rc, "cutc_no")
    return o2, o1, float(o0[0])


def cutc_no_device(ne_a, ne_b, wr1, mos_l, mos_r, G, no_2e, no_1e, no_0e, mem_limit=0, grid_chunk=0, stream=None):
    """cutc_no_device on torch device tensors (layouts as cutc_no); returns the
    statistics dict."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    nm, ng = G.shape[0], G.shape[3]
    st = NoStats()
    rc = lib().cutc_no_device(ng, nm, int(ne_a), int(ne_b), *(_tp(t) for t in (wr1, mos_l, mos_r, G, no_2e, no_1e,
                                                                              no_0e)),
                              mem_limit, grid_chunk, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st))
    _check(rc, "cutc_no_device")
    return st.as_dict()


def cutc_ao_to_mo(mo_l_coef, mo_r_coef, X):
    """AO -> MO transform on host arrays.  numpy C order of the Fortran
    layouts: mo_l_coef / mo_r_coef (n_mo, n_ao), X (3, n_grid1, n_ao, n_ao) =
    int2_grad1_u12_ao(n_ao, n_ao, n_grid1, 3).  Returns int2_grad1_u12 as
    (n_mo, n_mo, 3, n_grid1), the array cutc_no takes."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    cl = np.ascontiguousarray(mo_l_coef, dtype=np.float64)
    cr = np.ascontiguousarray(mo_r_coef, dtype=np.float64)
    n1, na = X.shape[1], X.shape[2]
    nm = cl.shape[0]
    if X.shape != (3, n1, na, na) or cl.shape != (nm, na) or cr.shape != (nm, na):
        raise ValueError("cutc_ao_to_mo: shapes do not match")
    out = np.zeros((nm, nm, 3, n1))
    rc = lib().cutc_ao_to_mo(n1, na, nm, cl.ctypes.data, cr.ctypes.data, X.ctypes.data, out.ctypes.data)
    _check(rc, "cutc_ao_to_mo")
    return out


def cutc_ao_to_mo_device(mo_l_coef, mo_r_coef, X, out, mem_limit=0, mo_chunk=0, stream=None):
    """cutc_ao_to_mo_device on torch device tensors (shapes as cutc_ao_to_mo,
    X may hold more than 3 components: the first 3 are read); returns the
    statistics dict."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    n1, na = X.shape[1], X.shape[2]
    nm = mo_l_coef.shape[0]
    if (tuple(X.shape[1:]) != (n1, na, na) or X.shape[0] < 3 or tuple(mo_l_coef.shape) != (nm, na)
            or tuple(mo_r_coef.shape) != (nm, na) or tuple(out.shape) != (nm, nm, 3, n1)):
        raise ValueError("cutc_ao_to_mo_device: shapes do not match")
    st = MoStats()
    rc = lib().cutc_ao_to_mo_device(n1, na, nm, *(_tp(t) for t in (mo_l_coef, mo_r_coef, X, out)), mem_limit,
                                    mo_chunk, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st))
    _check(rc, "cutc_ao_to_mo_device")
    return st.as_dict()


def _basis_arrays(basis):
    nucl = np.ascontiguousarray(basis["ao_nucl"], dtype=np.int32)
    n_ao = nucl.shape[0]
    power = np.ascontiguousarray(basis["ao_power"], dtype=np.int32).reshape(3, n_ao)
    nprim = np.ascontiguousarray(basis["ao_prim_num"], dtype=np.int32)
    expo = np.ascontiguousarray(basis["ao_expo"], dtype=np.float64).reshape(n_ao, -1)
    coef = np.ascontiguousarray(basis["ao_coef"], dtype=np.float64).reshape(n_ao, -1)
    return nucl, power, nprim, expo, coef


def cutc_aos(r, rn, basis):
    """AO values and gradients on host arrays: r (n, 3), rn (n_nuc, 3), basis
    a dict of the cutc_aos tables in numpy C order of the Fortran layouts
    (ao_nucl (n_ao,) 1-based, ao_power (3, n_ao), ao_prim_num (n_ao,),
    ao_expo / ao_coef (n_ao, prim_max)).  Returns aos_data as (4, n_ao, n)."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    rn = np.ascontiguousarray(rn, dtype=np.float64)
    nucl, power, nprim, expo, coef = _basis_arrays(basis)
    n, n_ao, pmax = r.shape[0], nucl.shape[0], expo.shape[1]
    out = np.zeros((4, n_ao, n))
    rc = lib().cutc_aos(n, rn.shape[0], n_ao, pmax, *(a.ctypes.data for a in (r, rn, nucl, power, nprim, expo, coef,
                                                                              out)))
    _check(rc, "cutc_aos")
    return out


def upload_basis(basis, device="cuda"):
    """The cutc_aos tables as torch device tensors (dict, same keys)."""
    import torch
    nucl, power, nprim, expo, coef = _basis_arrays(basis)
    return {k: torch.from_numpy(v).to(device) for k, v in
            (("ao_nucl", nucl), ("ao_power", power), ("ao_prim_num", nprim), ("ao_expo", expo), ("ao_coef", coef))}


def cutc_aos_device(r, rn, dbasis, out, stream=None):
    """cutc_aos_device on torch device tensors: r (n, 3), rn (n_nuc, 3), the
    tables of upload_basis, out (4, n_ao, n) (or a view with that layout,
    e.g. DeviceSystem.aos1)."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    n, n_ao = r.shape[0], dbasis["ao_nucl"].shape[0]
    pmax = dbasis["ao_expo"].shape[1]
    if tuple(out.shape) != (4, n_ao, n):
        raise ValueError("cutc_aos_device: out must be (4, n_ao, n)")
    rc = lib().cutc_aos_device(n, rn.shape[0], n_ao, pmax, _tp(r), _tp(rn),
                               *(_tp(dbasis[k]) for k in ("ao_nucl", "ao_power", "ao_prim_num", "ao_expo", "ao_coef")),
                               _tp(out), ctypes.c_void_p(stream.cuda_stream))
    _check(rc, "cutc_aos_device")


def release_workspace():
    """Free the device workspace cached by cutc_int_device / cutc_int_ex."""
    lib().cutc_release_workspace()


def get_int2_grad1_u12_ao(ds, X, grid=(1, 1, 1), block=(32, 1, 1)):
    """Stage 1 into the device buffer X (4, n1, n_ao, n_ao)."""
    lib().get_int2_grad1_u12_ao(Dim3(*grid), Dim3(*block), ds.n1, ds.n2, ds.n_ao, ds.n_nuc, ds.size_bh, _tp(ds.r1),
                                _tp(ds.r2), _tp(ds.wr2), _tp(ds.rn), _tp(ds.aos2), _tp(ds.c_bh), _tp(ds.m_bh),
                                _tp(ds.n_bh), _tp(ds.o_bh), _tp(X))
    _check(lib().cutc_get_last_error(), "get_int2_grad1_u12_ao")


def get_int_2e_ao(ds, X, I):
    """Stage 2: int_2e_ao (n_ao,)*4 from X (4, n1, n_ao, n_ao)."""
    lib().get_int_2e_ao(ds.n1, ds.n_ao, _tp(ds.wr1), _tp(ds.aos1), _tp(X), _tp(I))
    _check(lib().cutc_get_last_error(), "get_int_2e_ao")


def cutc_int_bh(ds, G, ii0, n1_eff, n1_tot, jj0, n2_eff, n2_tot):
    """Reference-layout Jastrow block: G[m * n1_tot * n2_tot + i1 * n2_tot + i2]."""
    lib().cutc_int_bh(Dim3(1, 1, 1), Dim3(32, 1, 1), ii0, n1_eff, n1_tot, jj0, n2_eff, n2_tot, ds.n_nuc,
                      ds.size_bh, _tp(ds.r1), _tp(ds.r2), _tp(ds.rn), _tp(ds.c_bh), _tp(ds.m_bh), _tp(ds.n_bh),
                      _tp(ds.o_bh), _tp(G))
    _check(lib().cutc_get_last_error(), "cutc_int_bh")


def int_long_range(ds, out, jj0, n2_eff, n2_tot):
    lib().int_long_range(jj0, n2_eff, n2_tot, ds.n2, ds.n_ao, _tp(ds.wr2), _tp(ds.aos2), _tp(out))
    _check(lib().cutc_get_last_error(), "int_long_range")


def int_short_range_herm(ds, out):
    lib().int_short_range_herm(ds.n1, ds.n_ao, _tp(ds.wr1), _tp(ds.aos1), _tp(out))
    _check(lib().cutc_get_last_error(), "int_short_range_herm")


def int_short_range_nonherm(ds, out):
    lib().int_short_range_nonherm(ds.n1, ds.n_ao, _tp(ds.wr1), _tp(ds.aos1), _tp(out))
    _check(lib().cutc_get_last_error(), "int_short_range_nonherm")


def trans_inplace(M):
    lib().trans_inplace(_tp(M), M.shape[0])
    _check(lib().cutc_get_last_error(), "trans_inplace")
