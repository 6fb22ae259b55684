"""ctypes binding of libssmhip.so (the C-ABI declared in include/ssmhip.h).

There is no CPU fallback: if the library is missing, or no HIP device is present when a
context is created, the calls raise.  The library is loaded from this package directory
(built in-tree by ``__graft_entry__.build()`` / ``csrc/Makefile``).
"""
from __future__ import annotations

import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libssmhip.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "ssmhip.h")

SSM_OK = 0
SSM_ERR_ARG = -1
SSM_ERR_NOT_PD = -2
SSM_ERR_OOM = -3
SSM_ERR_ADMISSION = -4
SSM_ERR_HIP = -5
SSM_ERR_STATE = -6
SSM_ERR_NO_DEVICE = -7

SSM_SYSTEMATIC, SSM_SYSTEMATIC_MURRAY, SSM_STRATIFIED, SSM_MULTINOMIAL = 0, 1, 2, 3
SSM_CSMC_PLAIN, SSM_CSMC_ANCESTOR = 0, 1
SSM_APF_NONE, SSM_APF_MEAN = 0, 1
SSM_F32, SSM_F64 = 0, 1


class SSMError(RuntimeError):
    """A non-zero status from the library (maps the reference's Julia exceptions)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ArgumentError(SSMError, ValueError):
    pass


class PosDefException(SSMError):
    pass


class AdmissionError(SSMError):
    pass


_EXC = {SSM_ERR_ARG: ArgumentError, SSM_ERR_NOT_PD: PosDefException,
        SSM_ERR_ADMISSION: AdmissionError}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p


class LGModelC(C.Structure):
    _fields_ = [("Dx", C.c_int32), ("Dy", C.c_int32)] + [
        (n, _dp) for n in ("mu0", "Sigma0", "A", "b", "Q", "H", "c", "R")]


class RequestC(C.Structure):
    """ssm_request."""
    _fields_ = [("kind", C.c_int32), ("model", LGModelC), ("N", C.c_int32), ("T", C.c_int32),
                ("ys", _dp), ("threshold", C.c_double), ("resampler", C.c_int32),
                ("seed", C.c_uint64)]


class RBModelC(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("om0", "os0", "oa", "ob", "oq")] + [
        ("D", C.c_int32), ("Dy", C.c_int32), ("rs", C.c_double)] + [
        (n, _dp) for n in ("mu0", "Sigma0", "A", "b", "C", "Q", "H", "c", "R0")]


_lib = None


def load():
    """Loads libssmhip.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def header_symbols():
    """Names of the functions declared in include/ssmhip.h."""
    txt = open(HEADER_PATH).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(ssm_\w+)\s*\(",
                                 txt, flags=re.M)))


def _declare(L):
    i32, u64, u32, dbl = C.c_int32, C.c_uint64, C.c_uint32, C.c_double
    sig = {
        "ssm_version": ([], i32),
        "ssm_ctx_create": ([i32, u64, C.POINTER(_vp)], i32),
        "ssm_ctx_destroy": ([_vp], i32),
        "ssm_last_error": ([_vp], C.c_char_p),
        "ssm_ctx_synchronize": ([_vp], i32),
        "ssm_ctx_device_info": ([_vp, C.c_char_p, i32, _ip], i32),
        "ssm_dev_alloc": ([_vp, u64, C.POINTER(_vp)], i32),
        "ssm_dev_free": ([_vp, _vp], i32),
        "ssm_memcpy_h2d": ([_vp, _vp, _vp, u64], i32),
        "ssm_memcpy_d2h": ([_vp, _vp, _vp, u64], i32),
        "ssm_sample_ancestors": ([_vp, i32, i32, _vp, i32, _vp, u64, u32, _vp], i32),
        "ssm_sample_ancestors_noalloc": ([_vp, i32, i32, _vp, i32, _vp, u64, u32, _vp], i32),
        "ssm_sample_ancestors_host": ([_vp, i32, i32, _dp, _dp, _ip], i32),
        "ssm_offspring_to_ancestors": ([_vp, i32, _vp, _vp], i32),
        "ssm_ancestors_to_offspring": ([_vp, i32, _vp, _vp], i32),
        "ssm_bf_create": ([_vp, C.POINTER(LGModelC), i32, i32, u64, C.POINTER(_vp)], i32),
        "ssm_bf_destroy": ([_vp], i32),
        "ssm_bf_set_model": ([_vp, C.POINTER(LGModelC)], i32),
        "ssm_bf_initialise": ([_vp, _dp], i32),
        "ssm_bf_step": ([_vp, i32, _dp, dbl, i32, _dp, _dp, _dp, _ip], i32),
        "ssm_bf_run": ([_vp, i32, i32, _dp, dbl, i32, _dp, _dp], i32),
        "ssm_bf_export": ([_vp, _dp, _dp, _ip, _dp], i32),
        "ssm_bf_import": ([_vp, _dp, _dp], i32),
        "ssm_bf_last_timing": ([_vp, _dp, _dp], i32),
        "ssm_bf_profile": ([_vp, i32, i32, _dp, dbl, i32, i32, _dp, _dp], i32),
        "ssm_bf_set_reference": ([_vp, i32, _dp, i32], i32),
        "ssm_bf_sample_index": ([_vp, dbl, _ip], i32),
        "ssm_bf_set_seed": ([_vp, u64], i32),
        "ssm_bf_set_auxiliary": ([_vp, i32], i32),
        "ssm_kf_batch_create": ([_vp, i32, _ip, _ip, _dp, C.POINTER(C.c_int64), i32, _dp,
                                 C.POINTER(C.c_int64), C.POINTER(_vp)], i32),
        "ssm_kf_batch_run": ([_vp, _dp, _ip, _dp, _dp], i32),
        "ssm_kf_batch_last_timing": ([_vp, _dp], i32),
        "ssm_kf_batch_tune": ([_vp, i32, _ip, _dp], i32),
        "ssm_kf_batch_destroy": ([_vp], i32),
        "ssm_kf_loglikelihood": ([_vp, i32, i32, i32, i32, i32, _dp, _dp, _dp, _dp, _dp, _dp,
                                  _dp, _dp, _dp, dbl, _dp, _ip], i32),
        "ssm_kf_loglikelihood_device": ([_vp, i32, i32, i32, i32, i32] + [_vp] * 9 +
                                        [dbl, _dp, _ip], i32),
        "ssm_kf_loglikelihood_last_timing": ([_vp, _dp], i32),
        "ssm_bf_set_proposal": ([_vp, _dp, _dp, _dp, _dp], i32),
        "ssm_srkf_filter": ([_vp, i32, i32, i32, i32, i32] + [_dp] * 9 + [dbl, _dp, _ip, _dp, _dp],
                            i32),
        "ssm_kf_loglikelihood_grad": ([_vp, i32, i32, i32, i32, i32] + [_dp] * 9 +
                                      [dbl, _dp, _ip] + [_dp] * 8, i32),
        "ssm_rbpf_create": ([_vp, C.POINTER(RBModelC), i32, u64, C.POINTER(_vp)], i32),
        "ssm_rbpf_destroy": ([_vp], i32),
        "ssm_rbpf_initialise": ([_vp, _dp], i32),
        "ssm_rbpf_step": ([_vp, i32, _dp, dbl, i32, _dp, _dp, _dp, _ip], i32),
        "ssm_rbpf_run": ([_vp, i32, i32, _dp, dbl, i32, _dp, _dp], i32),
        "ssm_rbpf_export": ([_vp, _dp, _dp, _dp, _dp, _ip, _dp], i32),
        "ssm_rbpf_last_timing": ([_vp, _dp], i32),
        "ssm_bf_tree_create": ([_vp, C.c_int64, C.POINTER(_vp)], i32),
        "ssm_bf_tree_insert": ([_vp], i32),
        "ssm_bf_tree_info": ([_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _ip], i32),
        "ssm_bf_tree_ancestry": ([_vp, _dp], i32),
        "ssm_bf_tree_path": ([_vp, C.c_int32, _dp], i32),
        "ssm_bf_tree_export": ([_vp, _ip, _ip, _ip], i32),
        "ssm_bf_history_create": ([_vp, i32, C.POINTER(_vp)], i32),
        "ssm_bf_history_push": ([_vp], i32),
        "ssm_bf_history_reset": ([_vp], i32),
        "ssm_bf_history_record": ([_vp, i32], i32),
        "ssm_bf_history_backward_sample": ([_vp, _dp, _ip, _dp], i32),
        "ssm_bf_history_backward_sample_tv": ([_vp, C.POINTER(LGModelC), _dp, _ip, _dp], i32),
        "ssm_bf_history_info": ([_vp, _ip, _ip], i32),
        "ssm_bf_history_ancestry": ([_vp, i32, i32, _dp], i32),
        "ssm_bf_history_generation": ([_vp, i32, _dp, _dp, _ip], i32),
        "ssm_bf_history_destroy": ([_vp], i32),
        "ssm_bf_tree_destroy": ([_vp], i32),
        "ssm_serve_create": ([_vp, u64, i32, C.POINTER(_vp)], i32),
        "ssm_serve_submit": ([_vp, C.POINTER(RequestC), C.POINTER(C.c_int64)], i32),
        "ssm_serve_poll": ([_vp, C.c_int64, _ip, _dp, _dp, _ip], i32),
        "ssm_serve_wait": ([_vp, C.c_int64, dbl, _ip, _dp, _dp, _ip], i32),
        "ssm_serve_cancel": ([_vp, C.c_int64], i32),
        "ssm_serve_release": ([_vp, C.c_int64], i32),
        "ssm_serve_stats": ([_vp, _dp, i32], i32),
        "ssm_serve_destroy": ([_vp], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


class Context:
    """An ssm_ctx: one HIP device, one stream (not thread-safe, one per host thread)."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        L = load()
        p = _vp()
        rc = L.ssm_ctx_create(device, hbm_budget_bytes, C.byref(p))
        if rc == SSM_ERR_NO_DEVICE:
            raise SSMError(rc, "no HIP device available (the library has no CPU fallback)")
        if rc:
            raise SSMError(rc, "ssm_ctx_create failed")
        self.ptr = p
        self.lib = L

    def check(self, rc):
        if rc != SSM_OK:
            msg = self.lib.ssm_last_error(self.ptr)
            raise _EXC.get(rc, SSMError)(rc, msg.decode() if msg else "")
        return rc

    def synchronize(self):
        self.check(self.lib.ssm_ctx_synchronize(self.ptr))

    def device_info(self):
        buf = C.create_string_buffer(256)
        cu = C.c_int32()
        self.check(self.lib.ssm_ctx_device_info(self.ptr, buf, 256, C.byref(cu)))
        return buf.value.decode(), cu.value

    def close(self):
        if self.ptr:
            self.lib.ssm_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


class DeviceBuffer:
    """A raw device allocation owned by a Context (for hosts that stage device data)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = nbytes
        p = _vp()
        ctx.check(ctx.lib.ssm_dev_alloc(ctx.ptr, nbytes, C.byref(p)))
        self.ptr = p

    def upload(self, arr):
        import numpy as np
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        self.ctx.check(self.ctx.lib.ssm_memcpy_h2d(self.ctx.ptr, self.ptr,
                                                   a.ctypes.data_as(_vp), a.nbytes))

    def download(self, out):
        assert out.nbytes <= self.nbytes and out.flags.c_contiguous
        self.ctx.check(self.ctx.lib.ssm_memcpy_d2h(self.ctx.ptr, out.ctypes.data_as(_vp),
                                                   self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            self.ctx.lib.ssm_dev_free(self.ctx.ptr, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
