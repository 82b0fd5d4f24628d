"""Loader for ``libghx.so`` (the HIP/gfx950 implementation) via ctypes.

There is no CPU fallback: if the library is missing, or no GPU is visible,
calls raise instead of computing anything elsewhere.
"""
import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libghx.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "ghx.h")

_lib = None


class GhxError(RuntimeError):
    pass


def declared_symbols(header=HEADER_PATH):
    """Names of the functions that include/ghx.h declares."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ghx_[a-z0-9_]+)\s*\(", text)))


def load():
    """Load the library (raises GhxError if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GhxError(f"{LIB_PATH} not found: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        lib = C.CDLL(LIB_PATH)
        _declare(lib)
        _lib = lib
    return _lib


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_I = C.c_int
_D = C.c_double
_DP = C.POINTER(C.c_double)


def _declare(L):
    sig = {
        "ghx_last_error": (C.c_char_p, []),
        "ghx_device_ok": (_I, []),
        "ghx_malloc": (_I, [_PP, C.c_size_t]),
        "ghx_free": (_I, [_P]),
        "ghx_memcpy_h2d": (_I, [_P, _P, C.c_size_t]),
        "ghx_memcpy_d2h": (_I, [_P, _P, C.c_size_t]),
        "ghx_memcpy_d2d": (_I, [_P, _P, C.c_size_t]),
        "ghx_memset_zero": (_I, [_P, C.c_size_t]),
        "ghx_sync": (_I, []),
        "ghx_plan_create": (_I, [_PP, _I, _I, _I, _I, _I]),
        "ghx_plan_destroy": (_I, [_P]),
        "ghx_plan_set_stream": (_I, [_P, _P]),
        "ghx_r2c": (_I, [_P, _P, _P]),
        "ghx_c2r": (_I, [_P, _P, _P]),
        "ghx_r2c_many": (_I, [_P, _I, _PP, _PP]),
        "ghx_c2r_many": (_I, [_P, _I, _PP, _PP]),
        "ghx_r2c_host": (_I, [_P, _P, _P]),
        "ghx_c2r_host": (_I, [_P, _P, _P]),
        "ghx_plan_timers": (_I, [_P, _DP, _DP, _DP, _DP]),
        "ghx_plan_reset_timers": (_I, [_P]),
        "ghx_grid_create": (_I, [_PP, _I, _I, _I, _D, _D, _D, _I, _D]),
        "ghx_grid_destroy": (_I, [_P]),
        "ghx_grid_set_stream": (_I, [_P, _P]),
        "ghx_grid_info": (_I, [_P, _DP, _DP, C.POINTER(_I), _DP, C.POINTER(_I)]),
        "ghx_grid_r2c": (_I, [_P, _P, _P]),
        "ghx_grid_c2r": (_I, [_P, _P, _P]),
        "ghx_derivk3": (_I, [_P, _P, _P, _I]),
        "ghx_laplak3": (_I, [_P, _P, _P]),
        "ghx_rotor3": (_I, [_P, _P, _P, _P, _I]),
        "ghx_prodre3": (_I, [_P] + [_P] * 6),
        "ghx_nonlhd3": (_I, [_P, _P, _P, _P, _P, _I]),
        "ghx_vector3": (_I, [_P] + [_P] * 9),
        "ghx_nonlin3": (_I, [_P] + [_P] * 7 + [_I]),
        "ghx_gauge3": (_I, [_P, _P, _P, _P, _P, _I]),
        "ghx_energy": (_I, [_P, _P, _P, _P, _I, _DP]),
        "ghx_helicity": (_I, [_P, _P, _P, _P, _DP]),
        "ghx_cross": (_I, [_P] + [_P] * 6 + [_I, _DP]),
        "ghx_maxabs": (_I, [_P, _P, _P, _P, _I, _DP]),
        "ghx_divergence": (_I, [_P, _P, _P, _P, _DP]),
        "ghx_spectrumc": (_I, [_P, _P, _P, _P, _I, _I, _DP, _DP]),
        "ghx_spectrumc_noalloc": (_I, [_P, _P, _P, _P, _DP]),
        "ghx_phaseshift": (_I, [_P, _P, _D, _D]),
        "ghx_normalize": (_I, [_P, _P, _P, _P, _D, _I]),
        "ghx_sharp": (_I, [_P, _P, _P, _D]),
        "ghx_randu": (_D, [C.POINTER(C.c_int32)]),
        "ghx_initialv_tg": (_I, [_P, _D, _D, _D, _P, _P, _P]),
        "ghx_initialfv_random": (_I, [_P, _D, _D, _D, C.POINTER(C.c_int32), _P, _P, _P]),
        "ghx_initialv_k4exp": (_I, [_P, _D, _D, C.POINTER(C.c_int32), _P, _P, _P]),
        "ghx_hd_rkstep2": (_I, [_P, _I, _D, _D, _PP, _PP, _PP, _PP]),
        "ghx_roth_rkstep2": (_I, [_P, _I, _D, _D, _D, _PP, _PP, _PP, _PP]),
        "ghx_mhd_rkstep2": (_I, [_P, _I, _D, _D, _D, _PP, _PP, _PP, _PP, _PP]),
        "ghx_bouss_rkstep2": (_I, [_P, _I, _D, _D, _D, _D, _D, _PP, _PP, _PP, _PP]),
        "ghx_rotbouss_rkstep2": (_I, [_P, _I, _D, _D, _D, _D, _D, _D, _PP, _PP, _PP, _PP]),
        "ghx_bouss_ord3": (_I, [_P, _D, _D, _PP, _PP]),
        "ghx_hd_create": (_I, [_PP, _P, _I, _D, _D]),
        "ghx_hd_destroy": (_I, [_P]),
        "ghx_hd_fields": (_I, [_P, _PP, _PP]),
        "ghx_hd_set_forcing": (_I, [_P, _I]),
        "ghx_hd_set_rotation": (_I, [_P, _D]),
        "ghx_hd_step": (_I, [_P, _I]),
        "ghx_hd_touch": (_I, [_P]),
        "ghx_mhd_touch": (_I, [_P]),
        "ghx_mhd_create": (_I, [_PP, _P, _I, _D, _D, _D]),
        "ghx_mhd_destroy": (_I, [_P]),
        "ghx_mhd_fields": (_I, [_P, _PP, _PP, _PP, _PP]),
        "ghx_mhd_set_forcing": (_I, [_P, _I]),
        "ghx_mhd_step": (_I, [_P, _I]),
        "ghx_advect3": (_I, [_P, _P, _P, _P, _P, _P]),
        "ghx_truncate": (_I, [_P, _P, _P, _P]),
        "ghx_bouss_create": (_I, [_PP, _P, _I, _D, _D, _D, _D, _D]),
        "ghx_bouss_destroy": (_I, [_P]),
        "ghx_bouss_fields": (_I, [_P, _PP, _PP, _PP, _PP]),
        "ghx_bouss_touch": (_I, [_P]),
        "ghx_bouss_set_forcing": (_I, [_P, _I]),
        "ghx_bouss_set_rotation": (_I, [_P, _D]),
        "ghx_bouss_step": (_I, [_P, _I]),
        "ghx_bouss_set_tracers": (_I, [_P, _P]),
        "ghx_ensemble_create": (_I, [_PP, _I, C.POINTER(_I), _DP, _D, _I]),
        "ghx_ensemble_destroy": (_I, [_P]),
        "ghx_ensemble_fields": (_I, [_P, _I, _PP]),
        "ghx_ensemble_grid": (_I, [_P, _I, _PP]),
        "ghx_ensemble_touch": (_I, [_P]),
        "ghx_ensemble_step": (_I, [_P, _I]),
        "ghx_ensemble_set_graph": (_I, [_P, _I]),
        "ghx_ensemble_autotune": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I)]),
        "ghx_ensemble_groups": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
        "ghx_ensemble_schedules": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
        "ghx_ensemble_set_schedule": (_I, [_P, _I, _I, _I, _I]),
        "ghx_ensemble_sync": (_I, [_P]),
        "ghx_tracers_create": (_I, [_PP, _P, C.c_longlong, _I]),
        "ghx_tracers_destroy": (_I, [_P]),
        "ghx_tracers_positions": (_I, [_P, _PP, _PP, _PP]),
        "ghx_tracers_velocities": (_I, [_P, _PP, _PP, _PP]),
        "ghx_tracers_set_step": (_I, [_P]),
        "ghx_tracers_step": (_I, [_P, _PP, _D, _D]),
        "ghx_tracers_euler_to_lag": (_I, [_P, _P, _P]),
        "ghx_hd_set_tracers": (_I, [_P, _P]),
        "ghx_mhd_set_tracers": (_I, [_P, _P]),
        "ghx_io_init": (_I, [_PP, _I, C.POINTER(_I), _I, _I]),
        "ghx_io_initc": (_I, [_PP, _I, C.POINTER(_I), _I, _I]),
        "ghx_io_destroy": (_I, [_P]),
        "ghx_io_set_options": (_I, [_P, _I, _I, C.c_longlong]),
        "ghx_io_read": (_I, [_P, C.c_char_p, C.c_char_p, C.c_char_p, _P]),
        "ghx_io_write": (_I, [_P, C.c_char_p, C.c_char_p, C.c_char_p, _P]),
        "ghx_io_readc": (_I, [_P, C.c_char_p, C.c_char_p, C.c_char_p, _P]),
        "ghx_io_writec": (_I, [_P, C.c_char_p, C.c_char_p, C.c_char_p, _P]),
        "ghx_hd_output": (_I, [_P, _P, C.c_char_p, C.c_char_p, _I, _P, _P, _P]),
        "ghx_hd_restart": (_I, [_P, _P, C.c_char_p, C.c_char_p, _P, _P, _P]),
        "ghx_mlp_nparams": (_I, []),
        "ghx_mlp_create": (_I, [_PP, C.c_ulonglong]),
        "ghx_mlp_destroy": (_I, [_P]),
        "ghx_mlp_set_stream": (_I, [_P, _P]),
        "ghx_mlp_params": (_I, [_P, _PP]),
        "ghx_mlp_set_norm": (_I, [_P, C.POINTER(C.c_float)]),
        "ghx_mlp_get_norm": (_I, [_P, C.POINTER(C.c_float)]),
        "ghx_mlp_fit_norm": (_I, [_P, C.c_longlong, _PP, _D, _PP]),
        "ghx_mlp_pack": (_I, [_P, C.c_longlong, _PP, _D, _PP, _P, _P]),
        "ghx_mlp_forward": (_I, [_P, C.c_longlong, _P, _P]),
        "ghx_mlp_grad": (_I, [_P, C.c_longlong, _P, _P, _P, _DP]),
        "ghx_mlp_train": (_I, [_P, C.c_longlong, _P, _P, C.c_longlong, _I, C.c_float, C.c_ulonglong,
                               C.POINTER(C.c_float)]),
        "ghx_mlp_stress": (_I, [_P, C.c_longlong, _PP, _D, _PP]),
        "ghx_velocity_gradients": (_I, [_P, _PP, _PP]),
        "ghx_sgs_targets": (_I, [_P, _P, _PP, _D, _PP, _PP]),
        "ghx_hd_set_sgs": (_I, [_P, _P, _D]),
        "ghx_hd_set_sgs_spectral": (_I, [_P, _I, _DP]),
        "ghx_entrans": (_I, [_P] * 7 + [_I, _DP]),
        "ghx_heltrans": (_I, [_P] * 7 + [_DP]),
        "ghx_crosspecc": (_I, [_P] * 7 + [_D, _DP]),
        "ghx_spectr1d": (_I, [_P, _P, _I, _DP]),
        "ghx_variance": (_I, [_P, _P, _I, _DP]),
        "ghx_product": (_I, [_P, _P, _P, _DP]),
        "ghx_pscheck": (_I, [_P, _P, _P, _DP]),
        "ghx_spectrscc": (_I, [_P, _P, _D, _DP]),
        "ghx_sctrans": (_I, [_P, _P, _P, _DP]),
        "ghx_difucx": (_I, [_P, _P, _P, _DP]),
        "ghx_difucz": (_I, [_P, _P, _P, _DP]),
        "ghx_havgcomp": (_I, [_P] * 5 + [_D, _D, _I, _DP]),
        "ghx_tbouss": (_I, [_P] * 5 + [_D, _D, _DP, _DP, _DP]),
        "ghx_spectpv": (_I, [_P] * 5 + [_DP]),
        "ghx_axis_bins": (_I, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "ghx_specpara": (_I, [_P] * 4 + [_I, _I, _DP, _DP]),
        "ghx_specperp": (_I, [_P] * 4 + [_I, _I, _DP, _DP]),
        "ghx_entpara": (_I, [_P] * 7 + [_I, _DP]),
        "ghx_entperp": (_I, [_P] * 7 + [_I, _DP]),
        "ghx_heltpara": (_I, [_P] * 7 + [_DP]),
        "ghx_heltperp": (_I, [_P] * 7 + [_DP]),
        "ghx_spec2d": (_I, [_P] * 4 + [_I, _I, _DP, _DP]),
        "ghx_hdcheck": (_I, [_P] * 7 + [_I, _I, _DP]),
        "ghx_mhdcheck": (_I, [_P] * 7 + [_I, _I, _I, _DP]),
        "ghx_prof_enable": (_I, [_I]),
        "ghx_prof_reset": (_I, []),
        "ghx_prof_read": (_I, [_I, C.POINTER(C.c_char_p), C.POINTER(_I), _DP, _DP]),
        "ghx_stream_probe": (_I, [_I, _P, _P, C.c_size_t, _I, _P]),
        "ghx_pass_shape": (_I, [_I, _I, _I, _I, C.c_longlong, C.POINTER(_I), C.POINTER(_I),
                                C.POINTER(C.c_longlong), C.POINTER(_I)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def check(rc):
    if rc != 0:
        msg = load().ghx_last_error().decode(errors="replace")
        raise GhxError(f"ghx error {rc}: {msg}")
    return rc


def ptr_array(ptrs):
    arr = (C.c_void_p * len(ptrs))(*[C.c_void_p(p) for p in ptrs])
    return C.cast(arr, _PP), arr


def prof_enable(on=True):
    check(load().ghx_prof_enable(1 if on else 0))


def prof_reset():
    check(load().ghx_prof_reset())


def prof_read():
    """{pass name: (launches, total ms, total algorithmic bytes)} from the pass profiler."""
    L = load()
    out = {}
    slot = 0
    while True:
        name, cnt, ms, by = C.c_char_p(), C.c_int(), C.c_double(), C.c_double()
        if L.ghx_prof_read(slot, C.byref(name), C.byref(cnt), C.byref(ms), C.byref(by)) != 0:
            break
        if cnt.value:
            out[name.value.decode()] = (cnt.value, ms.value, by.value)
        slot += 1
    return out
