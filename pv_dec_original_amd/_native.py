"""ctypes bindings of the in-tree native libraries (include/pvlbm.h, include/pvnet.h, include/pvgen.h).

The simulation path has no fallback: if libpvlbm.so is missing or a call fails, an exception is
raised.  Build with ``python -m pv_dec_original_amd.build`` or ``__graft_entry__.build()``.
"""
import ctypes as C
import os

# One HIP runtime per process.  PyTorch (device memory and streams of the warm-start network)
# ships its own libamdhip64/libhsa-runtime64; the libraries here link the system ones under the
# same soname.  Loaded first, torch's runtime serves every library of the process; loaded after
# ours, a second runtime would come up beside the first and torch would see no device.
try:  # noqa: SIM105
    import torch  # noqa: F401
except ImportError:
    pass

HERE = os.path.dirname(os.path.abspath(__file__))

PVLBM_OK = 0
FP64, FP16S, BF16S = 0, 1, 2
INLET, OUTLET1, OUTLET2, OUTLET3 = 0, 1, 2, 3
INLET_XMIN, OUTLET_YMAX, OUTLET_ZMIN, OUTLET_XMAX = 0, 1, 2, 3
SOLID, FLUID, NTW, KINLET, KOUTLET, WALL = 0, 1, 2, 3, 4, 5
LIST_FLUID, LIST_INLET, LIST_OUTLET1, LIST_OUTLET2, LIST_OUTLET3, LIST_WALL, LIST_NTW, LIST_TILES = range(8)
WALL_REFERENCE, WALL_INTERPOLATED = 0, 1
STORES_AUTO, STORES_REGULAR, STORES_STREAMING = 0, 1, 2
STREAMING_AUTO, STREAMING_TWO_LATTICES, STREAMING_IN_PLACE = 0, 1, 2

_d = C.POINTER(C.c_double)
_u8 = C.POINTER(C.c_uint8)
_i64 = C.POINTER(C.c_int64)


class Face(C.Structure):
    """pvlbm_face: {role, closure, axis, reserved, plane, tol, lo[2], hi[2], rho, plane_lo}."""
    _fields_ = [("role", C.c_int32), ("closure", C.c_int32), ("axis", C.c_int32),
                ("reserved", C.c_int32), ("plane", C.c_double), ("tol", C.c_double),
                ("lo", C.c_double * 2), ("hi", C.c_double * 2), ("rho", C.c_double),
                ("plane_lo", C.c_double)]


class Grid(C.Structure):
    """pvlbm_grid: node (i,j,k) at origin + (i,j,k)*h."""
    _fields_ = [("origin", C.c_double * 3), ("dims", C.c_int32 * 3), ("reserved", C.c_int32)]


class NativeError(RuntimeError):
    pass


class NetShape(C.Structure):
    """pvnet_shape: a batch of nb boxes of d x h x w voxels."""
    _fields_ = [("nb", C.c_int32), ("d", C.c_int32), ("h", C.c_int32), ("w", C.c_int32)]


class NetConv(C.Structure):
    """pvnet_conv (forward / input-gradient convolution)."""
    _fields_ = [("s", NetShape), ("ntaps", C.c_int32), ("cin", C.c_int32), ("cout", C.c_int32),
                ("kpad", C.c_int32), ("x", C.c_void_p), ("ldx", C.c_int32), ("flags", C.c_int32),
                ("w", C.c_void_p), ("bias", C.c_void_p), ("y", C.c_void_p), ("ldy", C.c_int32),
                ("ldr", C.c_int32), ("relu_of", C.c_void_p)]


class NetWgrad(C.Structure):
    """pvnet_wgrad (weight and bias gradient)."""
    _fields_ = [("s", NetShape), ("ntaps", C.c_int32), ("cin", C.c_int32), ("cout", C.c_int32),
                ("x", C.c_void_p), ("ldx", C.c_int32), ("lddy", C.c_int32), ("dy", C.c_void_p),
                ("ws", C.c_void_p), ("dw", C.c_void_p), ("db", C.c_void_p), ("flags", C.c_int32)]


NET_RELU, NET_F32OUT, NET_GATHER = 1, 2, 4

_lib = None
_gen = None
_net = None

# (name, restype, argtypes) of every entry point of include/pvlbm.h
PVLBM_API = [
    ("pvlbm_last_error", C.c_char_p, [C.c_void_p]),
    ("pvlbm_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("pvlbm_read_stl_ascii", C.c_int, [C.c_char_p, _d, C.c_int64, _i64]),
    ("pvlbm_read_stl_binary", C.c_int, [C.c_char_p, _d, C.c_int64, _i64]),
    ("pvlbm_bbox", C.c_int, [_d, C.c_int64, _d]),
    ("pvlbm_grid_from_bbox", C.c_int, [_d, C.c_double, C.POINTER(Grid)]),
    ("pvlbm_voxelize", C.c_int, [_d, C.c_int64, C.c_double, C.POINTER(Grid), C.POINTER(Grid), _u8]),
    ("pvlbm_voxelize_stats", C.c_int, [_i64, _d]),
    ("pvlbm_create", C.c_int, [_u8, C.POINTER(Grid), C.c_double, C.c_int, C.c_int, C.POINTER(Face),
                               C.c_int, _d, C.c_int64, C.c_double, C.c_double, C.POINTER(C.c_void_p)]),
    ("pvlbm_set_tau", C.c_int, [C.c_void_p, C.c_double]),
    ("pvlbm_set_dt", C.c_int, [C.c_void_p, C.c_double]),
    ("pvlbm_count", C.c_int, [C.c_void_p, C.c_int, _i64]),
    ("pvlbm_list", C.c_int, [C.c_void_p, C.c_int, _i64]),
    ("pvlbm_get_kinds", C.c_int, [C.c_void_p, _u8]),
    ("pvlbm_sample_kinds", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                                     C.POINTER(C.c_int32), C.c_int32, _u8]),
    ("pvlbm_sample_kinds_noalloc", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                                     C.POINTER(C.c_int32), C.c_int32, _u8]),
    ("pvlbm_set_inlet_profile", C.c_int, [C.c_void_p, _d]),
    ("pvlbm_init_fields", C.c_int, [C.c_void_p, _d, _d]),
    ("pvlbm_init_fields_coarse", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32),
                                           C.c_double, C.c_double, C.c_double, C.c_double, _d, _d, _i64]),
    ("pvlbm_set_wss_every", C.c_int, [C.c_void_p, C.c_int]),
    ("pvlbm_link_fractions", C.c_int, [C.c_void_p, C.c_int, _d]),
    ("pvlbm_set_wall_closure", C.c_int, [C.c_void_p, C.c_int]),
    ("pvlbm_set_store_policy", C.c_int, [C.c_void_p, C.c_int]),
    ("pvlbm_get_store_policy", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("pvlbm_get_load_policy", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("pvlbm_set_streaming", C.c_int, [C.c_void_p, C.c_int]),
    ("pvlbm_get_streaming", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("pvlbm_step", C.c_int, [C.c_void_p, C.c_int, _d]),
    ("pvlbm_step_timed", C.c_int, [C.c_void_p, C.c_int, _d, _d]),
    ("pvlbm_iteration", C.c_int, [C.c_void_p, _i64]),
    ("pvlbm_get_macros", C.c_int, [C.c_void_p, _d, _d]),
    ("pvlbm_get_macros_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pvlbm_get_wss", C.c_int, [C.c_void_p, _d]),
    ("pvlbm_get_normals", C.c_int, [C.c_void_p, _d]),
    ("pvlbm_get_wss_magnitude", C.c_int, [C.c_void_p, _d]),
    ("pvlbm_get_wss_raw", C.c_int, [C.c_void_p, _d, _d]),
    ("pvlbm_reduce", C.c_int, [C.c_void_p, _d]),
    ("pvlbm_gate_events", C.c_int, [C.c_void_p, _i64]),
    ("pvlbm_pulsatile_velocity", C.c_double, [C.c_double, C.c_double]),
    ("pvlbm_destroy", C.c_int, [C.c_void_p]),
    ("pvlbm_copy_bench", C.c_int, [C.c_int64, C.c_int, _d]),
]

_vp = C.c_void_p
_sp = C.POINTER(NetShape)
PVNET_API = [
    ("pvnet_conv_fwd", C.c_int, [C.POINTER(NetConv), _vp]),
    ("pvnet_wgrad_ws_floats", C.c_int64, [_sp, C.c_int32, C.c_int32, C.c_int32]),
    ("pvnet_conv_wgrad", C.c_int, [C.POINTER(NetWgrad), _vp]),
    ("pvnet_pack_weights", C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, _vp, _vp, _vp]),
    ("pvnet_pool2_fwd", C.c_int, [_sp, C.c_int32, _vp, C.c_int32, _vp, C.c_int32, _vp]),
    ("pvnet_pool2_bwd", C.c_int, [_sp, C.c_int32, _vp, C.c_int32, _vp, C.c_int32, C.c_int32, _vp]),
    ("pvnet_pool2_bwd_relu", C.c_int, [_sp, C.c_int32, _vp, C.c_int32, _vp, C.c_int32, _vp, C.c_int32, _vp]),
    ("pvnet_up2_fwd", C.c_int, [_sp, C.c_int32, _vp, C.c_int32, _vp, C.c_int32, _vp]),
    ("pvnet_up2_bwd", C.c_int, [_sp, C.c_int32, _vp, C.c_int32, _vp, C.c_int32, _vp]),
    ("pvnet_relu_bwd", C.c_int, [C.c_int64, C.c_int32, _vp, C.c_int32, _vp, C.c_int32, _vp]),
    ("pvnet_wall_distance", C.c_int, [_sp, _vp, _vp, _vp]),
    ("pvnet_features", C.c_int, [_sp, _vp, _vp, _vp, _vp, _vp]),
    ("pvnet_loss_ws_doubles", C.c_int64, [C.c_int64]),
    ("pvnet_mse_loss", C.c_int, [C.c_int64, C.c_int32, _vp, C.c_int32, _vp, _vp, _vp, _vp, C.c_int32,
                                 _vp, _vp, _vp]),
    ("pvnet_adam", C.c_int, [C.c_int64, _vp, _vp, _vp, _vp, C.c_float, C.c_float, C.c_float,
                             C.c_float, C.c_int32, _vp]),
    ("pvnet_last_error", C.c_char_p, []),
]

PVGEN_API = [
    ("pvgen_tube", C.c_int, [C.c_double, C.c_double, C.c_int, C.POINTER(C.c_void_p)]),
    ("pvgen_ybifurcation", C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double, _d, _d, C.c_double,
                                     C.POINTER(C.c_void_p)]),
    ("pvgen_vessel", C.c_int, [C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double,
                               C.c_double, _d, _d, C.c_double, C.POINTER(C.c_void_p)]),
    ("pvgen_aneurysm", C.c_int, [C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_double, _d, _d, C.c_double,
                                 C.POINTER(C.c_void_p)]),
    ("pvgen_tree", C.c_int, [C.c_int, _d, _d, _d, _d, _d, C.c_double, C.POINTER(C.c_void_p)]),
    ("pvgen_mesh_size", C.c_int64, [C.c_void_p]),
    ("pvgen_mesh_copy", None, [C.c_void_p, _d]),
    ("pvgen_mesh_free", None, [C.c_void_p]),
    ("pvgen_write_stl_ascii", C.c_int, [C.c_char_p, _d, C.c_int64, C.c_char_p]),
    ("pvgen_write_stl_binary", C.c_int, [C.c_char_p, _d, C.c_int64, C.c_char_p]),
    ("pvgen_field", C.c_int, [C.c_int, _d, C.c_int64, _d, _d]),
]


def _bind(lib, api):
    for name, res, args in api:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    """The loaded libpvlbm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libpvlbm.so")
        if not os.path.exists(path):
            raise NativeError(f"{path} is missing: build it with __graft_entry__.build()")
        _lib = _bind(C.CDLL(path), PVLBM_API)
    return _lib


def net():
    """The loaded libpvnet.so (warm-start network kernels); raises if it has not been built."""
    global _net
    if _net is None:
        path = os.path.join(HERE, "libpvnet.so")
        if not os.path.exists(path):
            raise NativeError(f"{path} is missing: build it with __graft_entry__.build()")
        _net = _bind(C.CDLL(path), PVNET_API)
    return _net


def net_check(rc):
    if rc != PVLBM_OK:
        msg = net().pvnet_last_error()
        raise NativeError(f"pvnet error {rc}: {msg.decode() if msg else ''}")


def gen():
    global _gen
    if _gen is None:
        path = os.path.join(HERE, "libpvgen.so")
        if not os.path.exists(path):
            raise NativeError(f"{path} is missing: build it with __graft_entry__.build()")
        _gen = _bind(C.CDLL(path), PVGEN_API)
    return _gen


def check(rc, handle=None):
    if rc != PVLBM_OK:
        msg = lib().pvlbm_last_error(handle)
        raise NativeError(f"pvlbm error {rc}: {msg.decode() if msg else ''}")


def dptr(a):
    return a.ctypes.data_as(_d) if a is not None else None


def u8ptr(a):
    return a.ctypes.data_as(_u8) if a is not None else None


def i64ptr(a):
    return a.ctypes.data_as(_i64) if a is not None else None


def device_count():
    c = C.c_int(0)
    check(lib().pvlbm_device_count(C.byref(c)))
    return c.value
