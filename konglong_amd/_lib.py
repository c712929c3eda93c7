"""ctypes binding of libkonglong_rtm.so (the C ABI in include/konglong_rtm.h).

The library is built in-tree by ``konglong_amd/build.py``; importing a wrapper
without it raises immediately -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkonglong_rtm.so")

KLRTM_OK = 0
ENGINE_CFD, ENGINE_FD = 4, 7
ISO, VTI, TTI = 1, 2, 3
SNAP_FP32, SNAP_C16, SNAP_FP16, SNAP_BF16, SNAP_LEARNED, SNAP_CONVAE = 0, 1, 16, 17, 32, 33
IC_IMAGE1, IC_INVERSION, IC_INVERSION2, IC_ALLIN3 = 0, 1, 2, 3

_vp, _i, _f, _d, _sz, _l = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_size_t, C.c_long
_pi, _pf, _pd, _pl = C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_long)


class ShotConfig(C.Structure):
    """klrtm_shot_config."""
    _fields_ = [
        ("nx", _i), ("ny", _i), ("nz", _i),
        ("dx", _f), ("dy", _f), ("dz", _f), ("dt", _f),
        ("nt", _i),
        ("engine", _i), ("order", _i), ("model", _i),
        ("nbnd", _i * 6),
        ("vmax", _f), ("pml_eps", _f), ("pml3d_scaler", _f),
        ("pml3d", _i), ("top_pml", _i),
        ("src_xyz", _f * 3),
        ("f_peak", _f),
        ("src_spread", _i), ("rcv_spread", _i),
        ("khpass", _f), ("khcut", _f),
        ("nrcv", _i),
        ("snap_every", _i), ("snap_format", _i), ("rec_through_codec", _i),
        ("tmute_frac", _f), ("ntaper_z", _i), ("fuse_step", _i),
        ("snap_slots", _i), ("ic_fp32_only", _i), ("free_surface_iz", _i), ("surface_taper", _i),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "klrtm_last_error": (C.c_char_p, []),
    "klrtm_version": (_i, [_pi, _pi]),
    "klrtm_device_count": (_i, [_pi]),
    "klrtm_set_device": (_i, [_i]),
    "klrtm_set_stream": (_i, [_vp]),
    "klrtm_malloc": (_i, [C.POINTER(_vp), _sz]),
    "klrtm_free": (_i, [_vp]),
    "klrtm_memcpy_h2d": (_i, [_vp, _vp, _sz]),
    "klrtm_memcpy_d2h": (_i, [_vp, _vp, _sz]),
    "klrtm_memcpy_d2d": (_i, [_vp, _vp, _sz]),
    "klrtm_memset": (_i, [_vp, _i, _sz]),
    "klrtm_synchronize": (_i, []),
    "klrtm_timer_enable": (_i, [_i]),
    "klrtm_timer_read": (_i, [C.c_char_p, _pd, _pl]),
    "klrtm_timer_reset": (_i, []),
    "klrtm_kernel_laplacian": (_i, [_vp, _vp, _vp]),
    "klrtm_kernel_deriv1": (_i, [_vp, _i, _vp, _vp, _i, _i]),
    "klrtm_pml_profile": (_i, [_i, _f, _f, _f, _f, _pf, _pf]),
    "klrtm_pml_apply_lines": (_i, [_vp, _i, _i, _l, _f, _vp, _vp, _vp, _i]),
    "klrtm_pml_apply_trans": (_i, [_f, _vp, _vp, _vp, _i]),
    "klrtm_vel2": (_i, [_vp, _sz, _f]),
    "klrtm_update2nd": (_i, [_vp, _vp, _sz]),
    "klrtm_vel2_q": (_i, [_vp, _vp, _sz, _f, _f]),
    "klrtm_fft_size": (_i, [_i]),
    "klrtm_timeshift": (_i, [_vp, _i, _i, _i, _pf, _vp]),
    "klrtm_update2nd_q": (_i, [_vp, _vp, _vp, _vp, _vp, _sz, _i, _pf]),
    "klrtm_kernel_create": (_i, [C.POINTER(_vp), _i, _i, _i, _f, _f, _f, _i, _i]),
    "klrtm_kernel_destroy": (_i, [_vp]),
    "klrtm_kernel_set_model": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "klrtm_kernel_set_pml_options": (_i, [_vp, _f, _i, _f, _i]),
    "klrtm_kernel_set_boundary": (_i, [_vp, _i, _i, _i, _i, _i, _i, _f, _f]),
    "klrtm_kernel_clean_pml": (_i, [_vp]),
    "klrtm_kernel_iso": (_i, [_vp, _vp, _vp, _vp]),
    "klrtm_kernel_vti": (_i, [_vp, _vp, _vp, _vp]),
    "klrtm_kernel_tti": (_i, [_vp, _vp, _vp, _vp]),
    "klrtm_kernel_step": (_i, [_vp, _i, _vp, _vp, _vp]),
    "klrtm_kernel_set_tile": (_i, [_vp, _i, _i, _i]),
    "klrtm_kernel_set_overlap": (_i, [_vp, _i]),
    "klrtm_kernel_set_zgrid": (_i, [_vp, _vp]),
    "klrtm_propagator_set_zgrid": (_i, [_vp, _vp]),
    "klrtm_kernel_get_tile": (_i, [_vp, _pi, _pi, _pi]),
    "klrtm_cfd_plan": (_i, [_i, _f, _pf, _pf, _pf]),
    "klrtm_spread_coeff": (_i, [_i, _pf, _f, _f, _f, _pi, _pi]),
    "klrtm_receiver_create": (_i, [C.POINTER(_vp), _i, _i, _i, _i, _pf, _i, _f, _f, _i, _f]),
    "klrtm_receiver_destroy": (_i, [_vp]),
    "klrtm_receiver_set_traces": (_i, [_vp, _pf]),
    "klrtm_receiver_apply": (_i, [_vp, _vp, _i, _vp, _f]),
    "klrtm_receiver_extract": (_i, [_vp, _vp, _i]),
    "klrtm_receiver_apply_ghost": (_i, [_vp, _vp, _i, _vp, _f, _i]),
    "klrtm_receiver_extract_ghost": (_i, [_vp, _vp, _i, _i]),
    "klrtm_receiver_get_record": (_i, [_vp, _pf]),
    "klrtm_receiver_get_tables": (_i, [_vp, _pi, _pi, _pi, _pi, _pi, _pi, _pf, _pf, _pf, _pi]),
    "klrtm_nshort_volume": (_sz, [_i, _i, _i]),
    "klrtm_compress16": (_i, [_vp, _vp, _i, _i, _i]),
    "klrtm_compress16_noalloc": (_i, [_vp, _vp, _i, _i, _i]),
    "klrtm_nshort_pack": (_sz, [_sz, _i]),
    "klrtm_compress_pair": (_i, [_vp, _vp, _i, _i, _i, _i]),
    "klrtm_uncompress_pair": (_i, [_vp, _vp, _i, _i, _i, _i]),
    "klrtm_compress_none": (_i, [_vp, _vp, _i, _i, _i]),
    "klrtm_uncompress_none": (_i, [_vp, _vp, _i, _i, _i]),
    "klrtm_uncompress16": (_i, [_vp, _vp, _i, _i, _i]),
    "klrtm_blockq_nscales": (_sz, [_i, _i, _i]),
    "klrtm_blockq_encode": (_i, [_vp, _vp, _vp, _i, _i, _i, _i]),
    "klrtm_blockq_decode": (_i, [_vp, _vp, _vp, _i, _i, _i, _i]),
    "klrtm_ic_correlate": (_i, [_vp, _vp, _i, _i, _i, _f, _vp, _i, _vp, _vp, _vp, _vp]),
    "klrtm_apply_tape": (_i, [_vp, _i, _i, _i, _f, _vp, _i]),
    "klrtm_record_write": (_i, [C.c_char_p, _i, _i, _f, _f, _vp, _vp]),
    "klrtm_record_read_header": (_i, [C.c_char_p, _pi, _pi, _pf, _pf]),
    "klrtm_record_read": (_i, [C.c_char_p, _vp, _vp]),
    "klrtm_phasecorr_create": (_i, [C.POINTER(_vp), _i, _f, _f, _f, _f, _i, _f, _f, _i, _i, _f, _f]),
    "klrtm_phasecorr_destroy": (_i, [_vp]),
    "klrtm_phasecorr_info": (_i, [_vp, _pi, _pi]),
    "klrtm_phasecorr_backward": (_i, [_vp, _vp, _i, _vp, _f]),
    "klrtm_correlator_create": (_i, [C.POINTER(_vp), _i, _i, _i, _f, _f, _f, _f]),
    "klrtm_correlator_destroy": (_i, [_vp]),
    "klrtm_correlator_run": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "klrtm_correlator_costerm": (_i, [_vp, _vp]),
    "klrtm_correlator_sinterm": (_i, [_vp, _vp, _vp]),
    "klrtm_correlator_unit_derivative": (_i, [_vp, _vp, _i, _vp]),
    "klrtm_correlator_kfilter": (_i, [_vp, _vp]),
    "klrtm_correlator_set_double": (_i, [_vp, _i, _i, _i, C.c_float, C.c_float]),
    "klrtm_correlator_nloop": (_i, [_vp, _pi]),
    "klrtm_correlator_interleave": (_i, [_vp, _vp, _vp]),
    "klrtm_correlator_interp": (_i, [_vp, _vp]),
    "klrtm_correlator_run_copies": (_i, [_vp, _vp, _vp, _vp, _vp, _i]),
    "klrtm_correlator_hightape": (_i, [_vp, _vp]),
    "klrtm_correlator_interp_illum": (_i, [_vp, _vp, _vp]),
    "klrtm_propagator_create": (_i, [C.POINTER(_vp), C.POINTER(ShotConfig)]),
    "klrtm_propagator_destroy": (_i, [_vp]),
    "klrtm_propagator_set_model": (_i, [_vp, _pf, _pf, _pf, _pf, _pf]),
    "klrtm_propagator_set_q": (_i, [_vp, _pf, _i, _pf]),
    "klrtm_propagator_set_rho": (_i, [_vp, _pf]),
    "klrtm_propagator_set_source_array": (_i, [_vp, _i, _pf, _pf]),
    "klrtm_propagator_set_ghost": (_i, [_vp, _i, _i, _i]),
    "klrtm_propagator_set_snapshot_cache": (_i, [_vp, C.c_char_p, C.c_double, _i]),
    "klrtm_propagator_snapshot_cache_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_sz), C.POINTER(_i)]),
    "klrtm_propagator_set_receivers": (_i, [_vp, _pf]),
    "klrtm_propagator_set_source_wavelet": (_i, [_vp, _pf]),
    "klrtm_propagator_model": (_i, [_vp, _pf]),
    "klrtm_propagator_rtm": (_i, [_vp, _pf, _pf, _pf, _pd, _pd]),
    "klrtm_propagator_timings": (_i, [_vp, _pd, _pd, _pd]),
    "klrtm_propagator_wavefields": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "klrtm_propagator_forward_steps": (_i, [_vp, _i, _i, _i]),
    "klrtm_propagator_reset": (_i, [_vp]),
    "klrtm_propagator_backward_steps": (_i, [_vp, _i, _i, _i]),
    "klrtm_propagator_set_record": (_i, [_vp, _pf]),
    "klrtm_propagator_set_lcodec": (_i, [_vp, _vp]),
    "klrtm_propagator_set_graph": (_i, [_vp, _i]),
    "klrtm_propagator_set_convae": (_i, [_vp, _vp, C.c_float]),
    "klrtm_convae_create": (_i, [C.POINTER(_vp), C.c_uint]),
    "klrtm_convae_destroy": (_i, [_vp]),
    "klrtm_convae_param_count": (_sz, []),
    "klrtm_convae_get_params": (_i, [_vp, _vp]),
    "klrtm_convae_set_params": (_i, [_vp, _vp]),
    "klrtm_convae_forward": (_i, [_vp, _vp, _i, _vp, _vp]),
    "klrtm_convae_loss_grad": (_i, [_vp, _vp, _i, _vp, _pf]),
    "klrtm_convae_adam_step": (_i, [_vp, C.c_float]),
    "klrtm_convae_train": (_i, [_vp, _vp, _i, _i, _i, _i, C.c_float, _i, _i, C.c_float, C.c_uint, _pf, _pf]),
    "klrtm_convae_code_bytes": (_sz, [_i, _i, _i]),
    "klrtm_convae_encode": (_i, [_vp, _vp, _i, _i, _i, C.c_float, _vp]),
    "klrtm_convae_decode": (_i, [_vp, _vp, _vp, _i, _i, _i]),
    "klrtm_propagator_graph_stats": (_i, [_vp, _pi, _pl]),
    "klrtm_propagator_kernel": (_i, [_vp, C.POINTER(_vp)]),
    "klrtm_propagator_images": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "klrtm_propagator_clear_images": (_i, [_vp]),
    "klrtm_propagator_set_imaging": (_i, [_vp, _i, _f]),
    "klrtm_propagator_set_fk": (_i, [_vp, _i, _f, _i, _f, _i, _f, _f, _f]),
    "klrtm_receiver_fk_spread": (_i, [_vp, _f, _f, _f, _i, _f, _i, _f, _i, _i, _f, _f, _f]),
    "klrtm_receiver_fk_data": (_i, [_vp, C.POINTER(_vp)]),
    "klrtm_propagator_ic_images": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "klrtm_propagator_finish_images": (_i, [_vp]),
    "klrtm_propagator_set_ic_double": (_i, [_vp, _i, _i, _i, _f, _f]),
    "klrtm_propagator_ic_double_images": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), _pi]),
    "klrtm_stack_image": (_i, [_vp, _i, _i, _i, _i, _vp, _i, _i, _i, _i, _i, _i]),
    "klrtm_apply_symmetry": (_i, [_vp, _i, _i, _i, _i, _i, _i, _i]),
    "klrtm_illum_damping": (_i, [_vp, _i, _i, _i, _i, _i, _i, _f, _f, _pd]),
    "klrtm_apply_damping": (_i, [_vp, _vp, _sz]),
    "klrtm_avg_volume3d": (_i, [_vp, _i, _i, _i, _i, _i, _i]),
    "klrtm_cone_mute": (_i, [_vp, _vp, _i, _i, _i] + [_f] * 11 + [_pf, _i]),
    "klrtm_save_image": (_i, [_vp, _i, _i, _i, _i, _i, _i, _i, _vp, _i, _i, _i, _i, _i]),
    "klrtm_lagrange_resample": (_i, [_vp, _i, _i, _i] + [_f] * 6 + [_vp, _i, _i, _i] + [_f] * 6 + [_i]),
    "klrtm_stream_create": (_i, [C.POINTER(_vp)]),
    "klrtm_stream_destroy": (_i, [_vp]),
    "klrtm_stream_synchronize": (_i, [_vp]),
    "klrtm_lcodec_create": (_i, [C.POINTER(_vp), _i, _pf]),
    "klrtm_lcodec_destroy": (_i, [_vp]),
    "klrtm_lcodec_code_bytes": (_sz, [_vp, _i, _i, _i]),
    "klrtm_lcodec_encode": (_i, [_vp, _vp, _vp, _i, _i, _i]),
    "klrtm_lcodec_decode": (_i, [_vp, _vp, _vp, _i, _i, _i]),
    "klrtm_lcodec_gram": (_i, [_vp, _i, _i, _i, _f, _vp]),
    "klrtm_lcodec_correlate": (_i, [_vp, _vp, _vp, _i, _i, _i, _f, _vp, _i, _vp, _vp, _vp, _vp]),
}

_lib = None


class KlrtmError(RuntimeError):
    pass


def lib():
    """Load libkonglong_rtm.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python konglong_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return list(_SIGS)


def require(cond, what):
    """Argument check that stays on under python -O: the C ABI reads and writes through the pointers it gets."""
    if not cond:
        raise ValueError(what)


def check(rc):
    if rc != KLRTM_OK:
        msg = lib().klrtm_last_error()
        raise KlrtmError(f"klrtm error {rc}: {msg.decode() if msg else ''}")


def fptr(a):
    """float* of a contiguous float32 numpy array (or None)."""
    if a is None:
        return None
    require(a.dtype == np.float32 and a.flags.c_contiguous, "needs a contiguous float32 array")
    return a.ctypes.data_as(_pf)


def dptr(a):
    if a is None:
        return None
    require(a.dtype == np.float64 and a.flags.c_contiguous, "needs a contiguous float64 array")
    return a.ctypes.data_as(_pd)


def iptr(a):
    require(a.dtype == np.int32 and a.flags.c_contiguous, "needs a contiguous int32 array")
    return a.ctypes.data_as(_pi)


def device_count():
    n = C.c_int(0)
    check(lib().klrtm_device_count(C.byref(n)))
    return n.value


class DeviceArray:
    """A device allocation made through klrtm_malloc, with host transfers."""

    def __init__(self, shape, dtype=np.float32, zero=True):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        check(lib().klrtm_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value
        if zero:
            self.zero()

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, zero=False)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        require(a.nbytes == self.nbytes, f"upload of {a.nbytes} bytes into a device array of {self.nbytes}")
        check(lib().klrtm_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes))

    def download(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().klrtm_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self):
        check(lib().klrtm_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if getattr(self, "ptr", None):
            lib().klrtm_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib().klrtm_synchronize())
