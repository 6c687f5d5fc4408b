"""ctypes view of include/spac_fk.h (plumbing for tests and bench.py).

The C ABI is the product boundary; this module only mirrors its structs and
loads the in-tree library.  Loading fails loudly when the library is missing:
there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = Path(__file__).resolve().parent / "lib" / "libspac_fk.so"
HEADER = ROOT / "include" / "spac_fk.h"

SPAC_OK = 0
SPAC_E_INVALID_ARG = -1
SPAC_E_UNREGISTERED_PLUGIN = -2
SPAC_E_MAX_STEPS = -3
SPAC_E_NAN = -4
SPAC_E_HIP = -5
SPAC_E_PATTERN = -6
SPAC_E_BAD_INTERFACE = -7
SPAC_E_NO_DEVICE = -8
SPAC_E_OUT_OF_MEMORY = -9
SPAC_NOALLOC_MAX_PATHS = 1 << 19
SPAC_NOALLOC_MAX_SOLVES = 1 << 15

SPAC_COEF_FRANKE_POISSON = 1
SPAC_COEF_VARCOEF_FRANKE = 2
SPAC_COEF_FRANKE_ROBIN = 3
SPAC_REFLECT_XMIN = 1
SPAC_REFLECT_YMIN = 2
SPAC_REFLECT_XMAX = 4
SPAC_REFLECT_YMAX = 8
SPAC_BOUNDARY_NONE = 0
SPAC_BOUNDARY_RECTANGLE = 1
SPAC_BOUNDARY_CIRCLE = 2
SPAC_FP64 = 0
SPAC_FP32 = 1
SPAC_OUT_INSIDE = 0
SPAC_OUT_DOMAIN = 1
SPAC_OUT_SUBDOMAIN = 2
SPAC_OUT_REFLECTED = 3
SPAC_OUT_MAX_STEPS = 4
SPAC_RBF_IMQ_GLOBAL = 0
SPAC_RBF_GAUSS_STENCIL = 1
SPAC_ORDER_NODE = 0
SPAC_ORDER_COST = 1
SPAC_ORDER_TAIL = 2
SPAC_ORDER_AUTO = 3
SPAC_CV_ANALYTIC = 0
SPAC_CV_NONE = 1
SPAC_CV_TABLE = 2
SPAC_CV_NETWORK = 3
SPAC_TAB_SIGMA = 1
SPAC_TAB_DRIFT = 2
SPAC_TAB_KILLING = 4
SPAC_TAB_SOURCE = 8
SPAC_TAB_DIRICHLET = 16
SPAC_TAB_SOLUTION = 32
SPAC_TAB_ALL = 63
MLP_N_PARAMS = 8707

GOBET_CONSTANT = -0.5826  # Solvers/FeynmanKacSolver.cpp:268,283

dp = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
i8p = C.POINTER(C.c_int8)


class ProblemDesc(C.Structure):
    _fields_ = [
        ("coef_kind", C.c_int32),
        ("domain_kind", C.c_int32),
        ("domain_params", C.c_double * 4),
        ("subdomain_kind", C.c_int32),
        ("precision", C.c_int32),
        ("subdomain_params", C.c_double * 4),
        ("gobet_constant", C.c_double),
        ("max_steps", C.c_int32),
        ("reflecting", C.c_int32),
        ("reflect_rho", C.c_double),
    ]


class SolveBatch(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("reserved", C.c_int32),
        ("x0", dp),
        ("y0", dp),
        ("paths", u32p),
        ("h", dp),
        ("stream", u32p),
        ("circle", dp),
    ]


class PathRecords(C.Structure):
    _fields_ = [
        ("x_tau", dp),
        ("y_tau", dp),
        ("Y", dp),
        ("Z", dp),
        ("xi", dp),
        ("tau", dp),
        ("gval", dp),
        ("steps", i32p),
        ("outcome", i8p),
    ]


class PointStats(C.Structure):
    _fields_ = [
        ("phi", C.c_double),
        ("phiphi", C.c_double),
        ("phi_VR", C.c_double),
        ("phiphi_VR", C.c_double),
        ("xi", C.c_double),
        ("xixi", C.c_double),
        ("xiphi", C.c_double),
        ("APL", C.c_double),
        ("RNGC", C.c_double),
        ("B", C.c_double),
        ("paths", C.c_int64),
        ("path_steps", C.c_int64),
        ("n_max_steps", C.c_int64),
        ("status", C.c_int32),
        ("n_domain_exits", C.c_int32),
        ("cv_fallback", C.c_int32),
        ("reserved", C.c_int32),
    ]


class InterfaceTable(C.Structure):
    _fields_ = [
        ("n_interfaces", C.c_int32),
        ("rbf_kind", C.c_int32),
        ("cx", dp),
        ("cy", dp),
        ("radius", dp),
        ("code", i32p),
        ("c2", dp),
        ("node_offset", i32p),
        ("node_theta", dp),
        ("node_global", i32p),
        ("ipsi_offset", i64p),
        ("ipsi", dp),
        ("gauss_eps", C.c_double),
        ("stencil_size", C.c_int32),
        ("reserved", C.c_int32),
    ]


class NodeBatch(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("reserved", C.c_int32),
        ("x", dp),
        ("y", dp),
        ("iface", i32p),
        ("frontera", i32p),
        ("paths", u32p),
        ("h", dp),
        ("stream", u32p),
    ]


class RowBlock(C.Structure):
    _fields_ = [
        ("B", dp),
        ("G", dp),
        ("G_col", i32p),
        ("G_offset", i64p),
        ("stats", C.POINTER(PointStats)),
    ]


class ResumeReport(C.Structure):
    _fields_ = [("batches", C.c_int32), ("computed", C.c_int32), ("loaded", C.c_int32), ("rejected", C.c_int32)]


class CoefTables(C.Structure):
    _fields_ = [("mask", C.c_int32), ("table_size", C.c_int32), ("table_domain", C.c_double * 4),
                ("values", C.POINTER(C.c_double))]


class PsReport(C.Structure):
    _fields_ = [("circles", C.c_int32), ("rows", C.c_int32), ("factorizations", C.c_int32), ("solves", C.c_int32),
                ("seconds", C.c_double)]


class Timing(C.Structure):
    _fields_ = [
        ("path_kernel_ms", C.c_double),
        ("reduce_kernel_ms", C.c_double),
        ("total_ms", C.c_double),
        ("path_steps", C.c_int64),
        ("path_kernel_launches", C.c_int32),
        ("grid_blocks", C.c_int32),
    ]


class LaunchConfig(C.Structure):
    _fields_ = [
        ("block_size", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("chunk_paths", C.c_int32),
        ("item_order", C.c_int32),
    ]


class AutotuneReport(C.Structure):
    _fields_ = [
        ("n_candidates", C.c_int32),
        ("reserved", C.c_int32),
        ("chosen", LaunchConfig),
        ("best_ms", C.c_double),
        ("default_ms", C.c_double),
    ]


class CvConfig(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("fallback", C.c_int32),
        ("max_ratio", C.c_double),
        ("table_size", C.c_int32),
        ("reserved", C.c_int32),
        ("table_domain", C.c_double * 4),
        ("table", C.POINTER(C.c_float)),
    ]


class MlpDesc(C.Structure):
    _fields_ = [("domain", C.c_double * 4), ("seed", C.c_uint64), ("first_layer_scale", C.c_double)]


class MlpTrainCfg(C.Structure):
    _fields_ = [
        ("steps", C.c_int32),
        ("eager", C.c_int32),
        ("lr", C.c_double),
        ("lr_final", C.c_double),
        ("grad_weight", C.c_double),
        ("beta2", C.c_double),
    ]


class MlpTrainReport(C.Structure):
    _fields_ = [
        ("final_loss", C.c_double),
        ("train_ms", C.c_double),
        ("steps", C.c_int32),
        ("reserved", C.c_int32),
    ]


_LIB = None


def declared_symbols(header: Path = HEADER) -> list[str]:
    """Function names declared in a public header (for the export check)."""
    import re

    text = header.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|spac_status|int32_t)\s+(spac_\w+)\s*\(", text, re.M)))


def load(path: Path | None = None) -> C.CDLL:
    """Load the in-tree library; raise if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(str(p))
    lib.spac_last_error.restype = C.c_char_p
    lib.spac_abi_version.restype = C.c_int32
    lib.spac_device_count.restype = C.c_int32
    lib.spac_fk_problem_create.argtypes = [C.POINTER(ProblemDesc), C.POINTER(C.c_void_p)]
    lib.spac_fk_problem_destroy.argtypes = [C.c_void_p]
    lib.spac_fk_solve_points.argtypes = [C.c_void_p, dp, dp, C.c_int32, C.c_uint32, C.c_double, C.c_uint64,
                                         C.POINTER(PointStats)]
    lib.spac_fk_solve_points_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32,
                                                C.c_double, C.c_uint64, C.c_void_p]
    lib.spac_fk_solve_points_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32,
                                                        C.c_double, C.c_uint64, C.c_void_p]
    lib.spac_debug_device_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
    lib.spac_debug_device_free.argtypes = [C.c_void_p]
    lib.spac_debug_copy_to_device.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    lib.spac_debug_copy_to_host.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    lib.spac_debug_copy_device.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    lib.spac_fk_simulate.argtypes = [C.c_void_p, C.POINTER(SolveBatch), C.c_uint64, C.POINTER(PathRecords),
                                     C.POINTER(PointStats)]
    lib.spac_fk_assemble_rows.argtypes = [C.c_void_p, C.POINTER(InterfaceTable), C.POINTER(NodeBatch), C.c_uint64,
                                          C.POINTER(RowBlock)]
    lib.spac_fk_assemble_system.argtypes = [C.c_void_p, C.POINTER(InterfaceTable), C.POINTER(NodeBatch),
                                            C.c_uint64, C.c_void_p, C.POINTER(PointStats)]
    lib.spac_fk_assemble_system_async.argtypes = lib.spac_fk_assemble_system.argtypes
    lib.spac_fk_assemble_system_uniform.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_double, C.c_uint64,
                                                    C.POINTER(PointStats)]
    lib.spac_fk_assemble_system_uniform_async.argtypes = lib.spac_fk_assemble_system_uniform.argtypes
    lib.spac_fk_solve_points_device_async.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32,
                                                      C.c_double, C.c_uint64, C.c_void_p]
    lib.spac_fk_wait.argtypes = [C.c_void_p]
    lib.spac_fk_query.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    lib.spac_fk_last_timing.argtypes = [C.c_void_p, C.POINTER(Timing)]
    lib.spac_fk_assemble_rows_resumable.argtypes = [C.c_void_p, C.POINTER(InterfaceTable), C.POINTER(NodeBatch),
                                                    C.c_uint64, C.c_int32, C.c_char_p, C.POINTER(RowBlock),
                                                    C.POINTER(ResumeReport)]
    lib.spac_ps_assemble_rows.argtypes = [C.c_int32, C.POINTER(InterfaceTable), C.POINTER(NodeBatch), C.c_int32,
                                          C.POINTER(RowBlock), C.POINTER(PsReport)]
    lib.spac_rows_file_write.argtypes = [C.c_char_p, C.c_int64, C.c_uint64, C.c_uint64, C.c_int32,
                                         C.POINTER(RowBlock)]
    lib.spac_rows_file_read.argtypes = [C.c_char_p, C.c_int64, C.c_uint64, C.c_uint64, C.c_int32, C.c_int64,
                                        C.POINTER(RowBlock), C.POINTER(C.c_int32)]
    lib.spac_fk_set_launch.argtypes = [C.c_void_p, C.POINTER(LaunchConfig)]
    lib.spac_fk_set_control_variate.argtypes = [C.c_void_p, C.POINTER(CvConfig)]
    lib.spac_fk_set_coef_tables.argtypes = [C.c_void_p, C.POINTER(CoefTables)]
    lib.spac_fk_plan_create.argtypes = [C.c_void_p, C.POINTER(InterfaceTable), C.POINTER(NodeBatch),
                                        C.POINTER(C.c_void_p)]
    lib.spac_fk_plan_execute.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(RowBlock), C.POINTER(C.c_double)]
    lib.spac_fk_plan_destroy.argtypes = [C.c_void_p]
    lib.spac_fk_autotune_rows.argtypes = [C.c_void_p, C.POINTER(InterfaceTable), C.POINTER(NodeBatch), C.c_double,
                                          C.POINTER(LaunchConfig), C.POINTER(C.c_double), C.c_int32,
                                          C.POINTER(AutotuneReport)]
    fp = C.POINTER(C.c_float)
    lib.spac_mlp_create.argtypes = [C.POINTER(MlpDesc), C.POINTER(C.c_void_p)]
    lib.spac_mlp_destroy.argtypes = [C.c_void_p]
    lib.spac_mlp_get_params.argtypes = [C.c_void_p, fp]
    lib.spac_mlp_set_params.argtypes = [C.c_void_p, fp]
    lib.spac_mlp_loss_grad.argtypes = [C.c_void_p, dp, dp, dp, C.c_int32, C.c_double, fp, dp]
    lib.spac_mlp_train.argtypes = [C.c_void_p, dp, dp, dp, C.c_int32, C.POINTER(MlpTrainCfg),
                                   C.POINTER(MlpTrainReport)]
    lib.spac_mlp_forward_fp32.argtypes = [C.c_void_p, dp, dp, C.c_int32, fp]
    lib.spac_mlp_infer.argtypes = [C.c_void_p, dp, dp, C.c_int32, fp]
    lib.spac_mlp_build_table.argtypes = [C.c_void_p, C.c_int32, dp]
    lib.spac_mlp_read_table.argtypes = [C.c_void_p, fp]
    lib.spac_mlp_attach.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_double]
    lib.spac_mlp_predict_system.argtypes = [C.c_void_p, C.c_void_p, dp]
    lib.spac_mlp_attach_network.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_double]
    for header in sorted(HEADER.parent.glob("*.h")):
        for name in declared_symbols(header):
            getattr(lib, name).restype = C.c_int32 if name != "spac_last_error" else C.c_char_p
    for name in declared_symbols():
        getattr(lib, name).restype = C.c_int32 if name != "spac_last_error" else C.c_char_p
    if path is None:
        _LIB = lib
    return lib


def last_error(lib: C.CDLL | None = None) -> str:
    lib = lib or load()
    msg = lib.spac_last_error()
    return msg.decode() if msg else ""


class SpacError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"spac status {code}: {msg}")
        self.code = code


def check(code: int, lib: C.CDLL | None = None) -> int:
    if code != SPAC_OK:
        raise SpacError(code, last_error(lib))
    return code


def ptr(arr, ctype):
    """ctypes pointer to a contiguous numpy array (or None)."""
    if arr is None:
        return C.cast(None, C.POINTER(ctype))
    return arr.ctypes.data_as(C.POINTER(ctype))


class DeviceArray:
    """A device allocation made through the library's test hooks (include/spac_fk_testing.h),
    so that a test needs no second HIP runtime in the process."""

    def __init__(self, nbytes: int, lib: C.CDLL | None = None):
        self.lib = lib or load()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(self.lib.spac_debug_device_alloc(self.nbytes, C.byref(p)), self.lib)
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr, lib: C.CDLL | None = None):
        import numpy as np

        a = np.ascontiguousarray(arr)
        d = cls(max(a.nbytes, 1), lib)
        check(d.lib.spac_debug_copy_to_device(d.ptr, a.ctypes.data, a.nbytes), d.lib)
        return d

    def to_bytes(self) -> bytes:
        buf = (C.c_char * self.nbytes)()
        check(self.lib.spac_debug_copy_to_host(C.addressof(buf), self.ptr, self.nbytes), self.lib)
        return bytes(buf)

    def zero(self):
        z = bytes(self.nbytes)
        check(self.lib.spac_debug_copy_to_device(self.ptr, z, self.nbytes), self.lib)

    def free(self):
        if self.ptr:
            self.lib.spac_debug_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def env_flag(name: str) -> bool:
    return os.environ.get(name, "") not in ("", "0")
