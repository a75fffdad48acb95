"""ctypes driver of libwannier.so -- the reference's pySWE interface.

Same classes, constructor arguments, call order and return values as
pySWE/pyswe.py:9-98 of HHG-Solids: ``Settings(param)`` fills a native
Settings block from a dict, ``SWE(settings)`` creates and initialises a solve,
``run_simulation()`` propagates, ``get_current()`` returns
``np.array([t, jx, jy, jz])`` and ``delete()`` frees the solve.

Differences, all additive: ``param["path_lib"]`` may be omitted (the in-tree
library is used); a failed native call raises ``RuntimeError`` with the
library's message instead of terminating the process; ``param["precision"]``
("fp64" or "fp32") selects the stage arithmetic.
"""
import ctypes as ct
import os

import numpy as np

_F64 = np.ctypeslib.ndpointer(dtype=np.float64, ndim=1, flags="C")
_C128 = np.ctypeslib.ndpointer(dtype=np.complex128, ndim=1, flags="C")

DEFAULT_LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "libwannier.so")
_PRECISION = {"fp64": 0, "fp32": 1}


def _load(path_lib):
    if not os.path.exists(path_lib):
        raise RuntimeError(f"libwannier.so not found at {path_lib}: build it with `make -C hhg-solids_amd`")
    lib = ct.CDLL(path_lib)
    lib.SWESim_get_status.argtypes = [ct.c_void_p]
    lib.SWESim_get_status.restype = ct.c_int
    lib.SWESim_get_error.argtypes = [ct.c_void_p]
    lib.SWESim_get_error.restype = ct.c_char_p
    lib.SWESim_set_precision.argtypes = [ct.c_void_p, ct.c_int]
    lib.SWESim_set_precision.restype = None
    lib.SWESim_last_run_ms.argtypes = [ct.c_void_p]
    lib.SWESim_last_run_ms.restype = ct.c_double
    lib.SWESim_get_nt.argtypes = [ct.c_void_p]
    lib.SWESim_get_nt.restype = ct.c_int
    lib.Settings_delete.argtypes = [ct.c_void_p]
    lib.Settings_delete.restype = None
    return lib


class SWE:
    def __init__(self, settings):
        self.settings = settings
        self.swe_cpp = _load(settings.path_lib)
        self.swe_cpp.SWESim_new.argtypes = []
        self.swe_cpp.SWESim_new.restype = ct.c_void_p
        self.swe_cpp.SWESim_run_simulation.argtypes = [ct.c_void_p]
        self.swe_cpp.SWESim_run_simulation.restype = None
        self.swe_cpp.SWESim_test_files.argtypes = [ct.c_void_p]
        self.swe_cpp.SWESim_test_files.restype = None
        self.swe_cpp.SWESim_set_path_tb.argtypes = [ct.c_void_p, ct.c_char_p]
        self.swe_cpp.SWESim_set_path_tb.restype = None
        self.swe_cpp.SWESim_set_settings.argtypes = [ct.c_void_p, ct.c_void_p]
        self.swe_cpp.SWESim_set_settings.restype = None
        self.swe_cpp.SWESim_init.argtypes = [ct.c_void_p]
        self.swe_cpp.SWESim_init.restype = None
        self.swe_cpp.SWESim_restart.argtypes = [ct.c_void_p]
        self.swe_cpp.SWESim_restart.restype = None
        self.swe_cpp.SWESim_get_current.argtypes = [ct.c_void_p, _F64, _C128, _C128, _C128]
        self.swe_cpp.SWESim_get_current.restype = None
        self.swe_cpp.SWESim_delete.argtypes = [ct.c_void_p]
        self.swe_cpp.SWESim_delete.restype = None

        self.ptr = self.swe_cpp.SWESim_new()
        if not self.ptr:
            raise RuntimeError("SWESim_new failed")
        if settings.precision:
            self.swe_cpp.SWESim_set_precision(self.ptr, settings.precision)
            self._check("set_precision")
        self.set_path_tb(settings.path_tb.encode("UTF-8"))
        self.set_settings(settings.ptr)
        self.init()

    def _check(self, what):
        if self.swe_cpp.SWESim_get_status(self.ptr) != 0:
            raise RuntimeError(f"SWESim_{what}: {self.swe_cpp.SWESim_get_error(self.ptr).decode()}")

    def run_simulation(self):
        self.swe_cpp.SWESim_run_simulation(self.ptr)
        self._check("run_simulation")

    def test_files(self):
        self.swe_cpp.SWESim_test_files(self.ptr)
        self._check("test_files")

    def set_path_tb(self, val):
        self.swe_cpp.SWESim_set_path_tb(self.ptr, val)
        self._check("set_path_tb")

    def set_settings(self, val):
        self.swe_cpp.SWESim_set_settings(self.ptr, val)
        self._check("set_settings")

    def init(self):
        self.swe_cpp.SWESim_init(self.ptr)
        self._check("init")

    def restart(self):
        self.swe_cpp.SWESim_restart(self.ptr)
        self._check("restart")

    def delete(self):
        if self.ptr:
            self.swe_cpp.SWESim_delete(self.ptr)
            self.ptr = None

    def last_run_ms(self):
        return self.swe_cpp.SWESim_last_run_ms(self.ptr)

    def get_current(self):
        t = np.zeros(self.settings.nt, dtype=np.float64)
        jx = np.zeros(self.settings.nt, dtype=np.complex128)
        jy = np.zeros(self.settings.nt, dtype=np.complex128)
        jz = np.zeros(self.settings.nt, dtype=np.complex128)
        self.swe_cpp.SWESim_get_current(self.ptr, t, jx, jy, jz)
        self._check("get_current")
        return np.array([t, jx, jy, jz])


class Settings:
    def __init__(self, param):
        self.path_lib = param.get("path_lib") or DEFAULT_LIB
        self.path_tb = param["path_tb"]
        self.nt = int(param["tmax"] / param["dt"])
        self.precision = _PRECISION[param.get("precision", "fp64")]
        self.settings_swe = _load(self.path_lib)
        self.settings_swe.Settings_new.argtypes = []
        self.settings_swe.Settings_new.restype = ct.c_void_p
        for name in ("nr1", "nr2", "nr3"):
            getattr(self.settings_swe, "Settings_set_" + name).argtypes = [ct.c_void_p, ct.c_int]
            getattr(self.settings_swe, "Settings_set_" + name).restype = None
        for name in ("tmax", "dt", "intensity", "lambda", "tmax_field"):
            getattr(self.settings_swe, "Settings_set_" + name).argtypes = [ct.c_void_p, ct.c_double]
            getattr(self.settings_swe, "Settings_set_" + name).restype = None
        for name in ("pol_vec", "phi_vec"):
            getattr(self.settings_swe, "Settings_set_" + name).argtypes = [ct.c_void_p, _F64]
            getattr(self.settings_swe, "Settings_set_" + name).restype = None

        self.ptr = self.settings_swe.Settings_new()
        self.settings_swe.Settings_set_nr1(self.ptr, param["nr1"])
        self.settings_swe.Settings_set_nr2(self.ptr, param["nr2"])
        self.settings_swe.Settings_set_nr3(self.ptr, param["nr3"])
        self.settings_swe.Settings_set_tmax(self.ptr, param["tmax"])
        self.settings_swe.Settings_set_dt(self.ptr, param["dt"])
        self.settings_swe.Settings_set_intensity(self.ptr, param["intensity"])
        self.settings_swe.Settings_set_lambda(self.ptr, param["lambda"])
        self.settings_swe.Settings_set_tmax_field(self.ptr, param["tmax_field"])
        self.settings_swe.Settings_set_pol_vec(self.ptr, np.ascontiguousarray(param["pol_vec"], dtype=np.float64))
        self.settings_swe.Settings_set_phi_vec(self.ptr, np.ascontiguousarray(param["phi_vec"], dtype=np.float64))

    def delete(self):
        if self.ptr:
            self.settings_swe.Settings_delete(self.ptr)
            self.ptr = None
