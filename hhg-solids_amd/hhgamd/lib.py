"""Loading libwannier.so and the list of symbols its C ABI exports."""
import ctypes as ct
import os
import re
import subprocess

import numpy as np

from . import LIB_PATH, PKG_DIR, REPO_DIR

HEADER = os.path.join(REPO_DIR, "include", "hhg_swe.h")


def build(jobs=8):
    """Compile libwannier.so in-tree for gfx950 (hipcc cross-compiles without a GPU)."""
    subprocess.run(["make", "-s", "-C", PKG_DIR, f"-j{jobs}"], check=True)


def load(path=LIB_PATH):
    """Load the native library; fails loudly when it has not been built."""
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `make -C hhg-solids_amd`")
    return ct.CDLL(path)


def declared_symbols(header=HEADER):
    """Function names declared in include/hhg_swe.h."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)) - {"if", "defined"})


_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C")
_C = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C")


def bind(lib):
    """Declare argtypes/restypes of every C-ABI function of include/hhg_swe.h."""
    V, I, Dd, P = ct.c_void_p, ct.c_int, ct.c_double, ct.c_char_p
    sig = {
        "SWESim_new": ([], V), "SWESim_run_simulation": ([V], None), "SWESim_test_files": ([V], None),
        "SWESim_set_path_tb": ([V, P], None), "SWESim_set_settings": ([V, V], None),
        "SWESim_init": ([V], None), "SWESim_restart": ([V], None), "SWESim_init_host": ([V], None),
        "SWESim_get_current": ([V, _D, _C, _C, _C], None), "SWESim_delete": ([V], None),
        "SWESim_get_current_run": ([V, I, _D, _C, _C, _C], I),
        "Settings_new": ([], V), "Settings_delete": ([V], None),
        "Settings_set_nr1": ([V, I], None), "Settings_set_nr2": ([V, I], None), "Settings_set_nr3": ([V, I], None),
        "Settings_set_tmax": ([V, Dd], None), "Settings_set_dt": ([V, Dd], None),
        "Settings_set_intensity": ([V, Dd], None), "Settings_set_lambda": ([V, Dd], None),
        "Settings_set_tmax_field": ([V, Dd], None), "Settings_set_pol_vec": ([V, _D], None),
        "Settings_set_phi_vec": ([V, _D], None),
        "SWESim_get_status": ([V], I), "SWESim_get_error": ([V], P), "hhg_last_error": ([], P),
        "SWESim_get_nt": ([V], I), "SWESim_get_num_orb": ([V], I), "SWESim_set_precision": ([V, I], None),
        "SWESim_get_fields": ([V, _D, _D, _D], I), "SWESim_get_rho_k": ([V, _C], I),
        "SWESim_last_run_ms": ([V], Dd), "SWESim_last_kernel_ms": ([V], Dd),
        "SWESim_set_kernel_variant": ([V, I], None), "SWESim_get_kernel": ([V], I),
        "SWESim_flops_per_kstep": ([V], Dd), "SWESim_min_flops_per_kstep": ([V], Dd),
        "SWESim_autotune": ([V, I], I),
        "SWESim_get_launch": ([V, np.ctypeslib.ndpointer(dtype=np.int32, flags="C")], I),
        "SWESim_get_kpoints_per_wave": ([V], I),
        "SWESim_get_batch_launch": ([V, np.ctypeslib.ndpointer(dtype=np.int32, flags="C")], I),
        "hhg_set_autotune": ([I], None), "hhg_get_autotune": ([], I),
        "hhg_fft2": ([np.ctypeslib.ndpointer(dtype=np.complex128, flags="C"), I, I, I, I], I),
        "hhg_fft2_plan_create": ([I, I, I], V), "hhg_fft2_plan_exec": ([V, V, I, I], Dd),
        "hhg_fft2_plan_destroy": ([V], None),
        "hhg_batch_new": ([], V), "hhg_batch_delete": ([V], None), "hhg_batch_add": ([V, V], I),
        "hhg_batch_size": ([V], I), "hhg_batch_run": ([V], I), "hhg_batch_last_ms": ([V], Dd),
        "hhg_batch_get_error": ([V], P),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


def fft2(a, sign=-1, lib=None):
    """Unnormalised 2-D FFT over the last two axes on the GPU (the FFT form's
    LDS-resident radix transform); returns a new complex128 array."""
    L = bind(lib or load())
    x = np.ascontiguousarray(a, dtype=np.complex128).copy()
    n1, n2 = x.shape[-2:]
    nch = int(np.prod(x.shape[:-2])) if x.ndim > 2 else 1
    if L.hhg_fft2(x.reshape(-1), nch, n1, n2, int(sign)) != 0:
        raise RuntimeError("hhg_fft2: " + L.hhg_last_error().decode())
    return x


class DeviceArray:
    """A complex128 array in device memory, allocated through the HIP runtime
    (libamdhip64, the runtime libwannier.so uses): for the entry points that
    take device addresses. upload/download copy synchronously."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            h = ct.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
            h.hipFree.argtypes = [ct.c_void_p]
            h.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
            for f in (h.hipMalloc, h.hipFree, h.hipMemcpy):
                f.restype = ct.c_int
            cls._hip = h
        return cls._hip

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 16
        self.ptr = ct.c_void_p()
        if self.hip().hipMalloc(ct.byref(self.ptr), self.nbytes) != 0:
            raise RuntimeError(f"hipMalloc of {self.nbytes} bytes failed")

    @property
    def address(self):
        return self.ptr.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        if self.hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1) != 0:  # host to device
            raise RuntimeError("hipMemcpy to the device failed")

    def download(self):
        out = np.empty(self.shape, dtype=np.complex128)
        if self.hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) != 0:  # device to host
            raise RuntimeError("hipMemcpy from the device failed")
        return out

    def close(self):
        if self.ptr.value:
            self.hip().hipFree(self.ptr)
            self.ptr = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the runtime may be gone already
            pass


class FFT2Plan:
    """The LDS-resident 2-D FFT on device memory: nch row-major n1 x n2
    complex128 fields, transformed in place. `exec(address, sign, repeats)`
    takes a device address (DeviceArray.address, or any device allocation whose
    producers have finished) and returns the device time in milliseconds."""

    def __init__(self, nch, n1, n2, lib=None):
        self.lib = bind(lib or load())
        self.shape = (nch, n1, n2)
        self.ptr = self.lib.hhg_fft2_plan_create(nch, n1, n2)
        if not self.ptr:
            raise RuntimeError("hhg_fft2_plan_create: " + self.lib.hhg_last_error().decode())

    def exec(self, device_ptr, sign=-1, repeats=1):
        ms = self.lib.hhg_fft2_plan_exec(self.ptr, ct.c_void_p(int(device_ptr)), int(sign), int(repeats))
        if ms < 0:
            raise RuntimeError("hhg_fft2_plan_exec: " + self.lib.hhg_last_error().decode())
        return ms

    def close(self):
        if self.ptr:
            self.lib.hhg_fft2_plan_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the library may be gone already
            pass


def set_autotune(on, lib=None):
    """Process-wide: tune every solve's launch configuration on its first run."""
    bind(lib or load()).hhg_set_autotune(1 if on else 0)


def get_autotune(lib=None):
    """The process-wide tuning switch."""
    return bool(bind(lib or load()).hhg_get_autotune())


class Solve:
    """One solve behind the C ABI, with the extension calls.

    ``param`` uses the keys of the reference's pyswe dictionary (nr1, nr2,
    nr3, tmax, dt, intensity, lambda, tmax_field, pol_vec, phi_vec) plus
    path_tb and optionally precision ("fp64" | "fp32").
    """

    def __init__(self, param, device=True, lib=None):
        self.lib = bind(lib or load())
        L = self.lib
        self.st = L.Settings_new()
        for k in ("nr1", "nr2", "nr3"):
            getattr(L, "Settings_set_" + k)(self.st, int(param.get(k, 1)))
        for k, f in (("tmax", "tmax"), ("dt", "dt"), ("intensity", "intensity"), ("lambda", "lambda"),
                     ("tmax_field", "tmax_field")):
            getattr(L, "Settings_set_" + f)(self.st, float(param[k]))
        L.Settings_set_pol_vec(self.st, np.ascontiguousarray(param["pol_vec"], dtype=np.float64))
        L.Settings_set_phi_vec(self.st, np.ascontiguousarray(param["phi_vec"], dtype=np.float64))
        self.ptr = L.SWESim_new()
        L.SWESim_set_path_tb(self.ptr, param["path_tb"].encode())
        L.SWESim_set_settings(self.ptr, self.st)
        prec = {"fp64": 0, "fp32": 1}.get(param.get("precision", "fp64"), -1)  # the library rejects -1
        L.SWESim_set_precision(self.ptr, prec)
        self._check("set_precision")
        L.SWESim_set_kernel_variant(self.ptr, {"auto": 0, "generic": 1, "fftform": 2}.get(param.get("kernel", "auto"), -1))
        self._check("set_kernel_variant")
        (L.SWESim_init if device else L.SWESim_init_host)(self.ptr)
        self._check("init")
        self.nt = L.SWESim_get_nt(self.ptr)
        self.nb = L.SWESim_get_num_orb(self.ptr)
        self.N = int(param["nr1"]) * int(param["nr2"])

    def _check(self, what):
        if self.lib.SWESim_get_status(self.ptr) != 0:
            raise RuntimeError(f"SWESim_{what}: {self.lib.SWESim_get_error(self.ptr).decode()}")

    def run(self):
        self.lib.SWESim_run_simulation(self.ptr)
        self._check("run_simulation")
        return self.lib.SWESim_last_run_ms(self.ptr)

    def restart(self):
        """Re-initialise: equilibrium state, current history cleared (SWESim_restart)."""
        self.lib.SWESim_restart(self.ptr)
        self._check("restart")

    def set_precision(self, precision):
        """Stage arithmetic of the next runs: "fp64" or "fp32"."""
        self.lib.SWESim_set_precision(self.ptr, {"fp64": 0, "fp32": 1}.get(precision, -1))
        self._check("set_precision")

    def kernel(self):
        return self.lib.SWESim_get_kernel(self.ptr)

    def flops_per_kstep(self):
        """Flops per k-point and RK4 step of the kernel the next run launches."""
        return self.lib.SWESim_flops_per_kstep(self.ptr)

    def min_flops_per_kstep(self):
        """Flops per k-point and RK4 step that the model fixes, independent of the kernel."""
        return self.lib.SWESim_min_flops_per_kstep(self.ptr)

    def autotune(self, steps=32):
        """Time the launch configurations that cover this solve on its first
        `steps` steps (state untouched) and keep the fastest; returns the
        chosen kernel id."""
        k = self.lib.SWESim_autotune(self.ptr, int(steps))
        self._check("autotune")
        return k

    def launch(self):
        """Launch configuration of the next run: dict(kernel, a, b, c, fp32)."""
        out = np.zeros(5, dtype=np.int32)
        self.lib.SWESim_get_launch(self.ptr, out)
        self._check("get_launch")
        return dict(zip(("kernel", "a", "b", "c", "fp32"), out.tolist()))

    def batch_launch(self):
        """Launch configuration as a batch member or served request: dict(kernel,
        a, b, c, fp32, kpw)."""
        out = np.zeros(6, dtype=np.int32)
        self.lib.SWESim_get_batch_launch(self.ptr, out)
        self._check("get_batch_launch")
        return dict(zip(("kernel", "a", "b", "c", "fp32", "kpw"), out.tolist()))

    def kpoints_per_wave(self):
        """k-points per lane group of the next run on the row-separable kernel (1, 2 or 4; 1 elsewhere)."""
        k = self.lib.SWESim_get_kpoints_per_wave(self.ptr)
        self._check("get_kpoints_per_wave")
        return k

    def kernel_ms(self):
        return self.lib.SWESim_last_kernel_ms(self.ptr)

    def current(self):
        t = np.zeros(self.nt)
        J = [np.zeros(self.nt, dtype=np.complex128) for _ in range(3)]
        self.lib.SWESim_get_current(self.ptr, t, *J)
        self._check("get_current")
        return t, np.stack(J, axis=1)

    def current_run(self, run=-1):
        """J(t) of one run (0 = first, -1 = last): shape (nt, 3)."""
        t = np.zeros(self.nt)
        J = [np.zeros(self.nt, dtype=np.complex128) for _ in range(3)]
        self.lib.SWESim_get_current_run(self.ptr, int(run), t, *J)
        self._check("get_current_run")
        return t, np.stack(J, axis=1)

    def fields(self):
        t = np.zeros(self.nt); E = np.zeros(3 * self.nt); A = np.zeros(3 * self.nt)
        self.lib.SWESim_get_fields(self.ptr, t, E, A)
        self._check("get_fields")
        return t, E.reshape(3, self.nt), A.reshape(3, self.nt)

    def rho_k(self):
        out = np.zeros(self.N * self.nb * self.nb, dtype=np.complex128)
        self.lib.SWESim_get_rho_k(self.ptr, out)
        self._check("get_rho_k")
        return out.reshape(self.N, self.nb, self.nb)

    def close(self):
        if self.ptr:
            self.lib.SWESim_delete(self.ptr)
            self.ptr = None
        if self.st:
            self.lib.Settings_delete(self.st)
            self.st = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many solves propagated together (hhg_batch_* of include/hhg_swe.h)."""

    def __init__(self, solves, lib=None):
        self.lib = bind(lib or load())
        self.solves = list(solves)
        self.ptr = self.lib.hhg_batch_new()
        if not self.ptr:
            raise RuntimeError("hhg_batch_new: " + self.lib.hhg_last_error().decode())
        for s in self.solves:
            self._check(self.lib.hhg_batch_add(self.ptr, s.ptr), "add")

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"hhg_batch_{what}: {self.lib.hhg_batch_get_error(self.ptr).decode()}")

    def run(self):
        self._check(self.lib.hhg_batch_run(self.ptr), "run")
        return self.lib.hhg_batch_last_ms(self.ptr)

    def close(self):
        if self.ptr:
            self.lib.hhg_batch_delete(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
