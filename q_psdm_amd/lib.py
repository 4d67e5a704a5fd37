"""ctypes binding of libqpsdm.so (the C ABI declared in include/qpsdm.h,
include/segy.h and include/psdmpkg.h).

The library is built in-tree by ``make -C q_psdm_amd/csrc`` (see
__graft_entry__.build()).  Loading fails loudly when it is missing: there is
no CPU fallback for any entry point.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqpsdm.so")

QPSDM_F32, QPSDM_BF16, QPSDM_F16B, QPSDM_U16B, QPSDM_U16X, QPSDM_F16, QPSDM_U16Q = 0, 1, 2, 3, 4, 5, 6
QPSDM_F32P, QPSDM_F16P = 7, 8


class Window(C.Structure):
    _fields_ = [("nsb", C.c_int), ("itb1", C.c_int), ("itb2", C.c_int), ("nf1", C.c_int),
                ("nf4", C.c_int), ("nww", C.c_int), ("ncoef1", C.c_int), ("nf3", C.c_int),
                ("tdstart1", C.c_float), ("tdstart2", C.c_float), ("tdmid", C.c_float),
                ("df", C.c_float), ("dw", C.c_float)]


class GeomParams(C.Structure):
    _fields_ = [("scaled", C.c_double), ("fxv1", C.c_double), ("exv1", C.c_double),
                ("fyv1", C.c_double), ("eyv1", C.c_double), ("fxi", C.c_double),
                ("cyi", C.c_double), ("ox", C.c_double), ("oy", C.c_double),
                ("dxi", C.c_float), ("dyi", C.c_float)]


class TGeom(C.Structure):
    _fields_ = [("nnx", C.c_int), ("nny", C.c_int), ("n0x", C.c_float), ("n0y", C.c_float),
                ("ndx", C.c_float), ("ndy", C.c_float), ("nu", C.c_int), ("nv", C.c_int),
                ("nw", C.c_int), ("u0", C.c_float), ("v0", C.c_float), ("du", C.c_float),
                ("dv", C.c_float), ("dw", C.c_float)]


class Model(C.Structure):
    _fields_ = [("kind", C.c_int), ("v0", C.c_double), ("kgrad", C.c_double), ("q0", C.c_double),
                ("nlayer", C.c_int), ("layer_dz", C.c_double), ("vl", C.c_double * 16),
                ("ql", C.c_double * 16), ("ax", C.c_double), ("ay", C.c_double),
                ("az", C.c_double), ("asig", C.c_double), ("av", C.c_double), ("aq", C.c_double)]


class TidxHeader(C.Structure):
    _fields_ = [("cdp1", C.c_int32), ("cdp2", C.c_int32), ("dx", C.c_float), ("dz", C.c_float),
                ("nzg", C.c_int32), ("nxg", C.c_int32), ("nzi0", C.c_int32), ("nxi0", C.c_int32)]


class TttDesc(C.Structure):
    _fields_ = [("idx", C.c_int64), ("nys", C.c_uint16), ("nxs", C.c_uint16),
                ("dys", C.c_uint16), ("dxs", C.c_uint16)]


class RefLayout(C.Structure):
    _fields_ = [("line1", C.c_int), ("nlines", C.c_int), ("cdp1", C.c_int), ("ncol", C.c_int), ("nz", C.c_int),
                ("nxi0", C.c_int), ("nzi0", C.c_int), ("x0", C.c_float), ("y0", C.c_float), ("dx", C.c_float),
                ("dy", C.c_float), ("dz", C.c_float)]


class BankParams(C.Structure):
    _fields_ = [("nt", C.c_int), ("dt", C.c_float), ("win", Window), ("K", C.c_int),
                ("dtstar", C.c_float), ("fc", C.c_float), ("threshold", C.c_float)]


class ImageParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_float),
                ("dy", C.c_float), ("dz", C.c_float), ("nt", C.c_int), ("dt", C.c_float),
                ("K", C.c_int), ("dtstar", C.c_float), ("tanx1", C.c_float),
                ("tanx2", C.c_float), ("tany1", C.c_float), ("tany2", C.c_float),
                ("taperzone", C.c_float), ("tmin", C.c_float), ("tg", TGeom),
                ("weight_nodes", C.c_int), ("z_exact", C.c_float), ("span_n", C.c_int),
                ("span_z", C.c_float * 4), ("span_s", C.c_int * 4)]


class Launch(C.Structure):
    _fields_ = [("tile_x", C.c_int), ("tile_y", C.c_int), ("tile_z", C.c_int)]


class GatherParams(C.Structure):
    _fields_ = [("dipx", C.c_float), ("dipy", C.c_float), ("ddipx", C.c_float), ("ddipy", C.c_float),
                ("aspect", C.c_int)]


class CtxParams(C.Structure):
    _fields_ = [("image", ImageParams), ("bank", BankParams), ("geom", GeomParams),
                ("bank_dtype", C.c_int), ("table_dtype", C.c_int), ("batch_traces", C.c_int),
                ("shape", Launch)]


class TrHdr(C.Structure):
    _fields_ = [("sx", C.c_int32), ("sy", C.c_int32), ("gx", C.c_int32), ("gy", C.c_int32),
                ("fldr", C.c_int32), ("cdp", C.c_int32), ("offset", C.c_int32),
                ("tracl", C.c_int32), ("scalco", C.c_int16), ("pad", C.c_int16),
                ("ns", C.c_uint16), ("dt", C.c_uint16)]


class Shot(C.Structure):
    _fields_ = [("tr0", C.c_int), ("ntr", C.c_int), ("tanx1", C.c_float), ("tanx2", C.c_float),
                ("tany1", C.c_float), ("tany2", C.c_float)]


class CoorT(C.Structure):
    _fields_ = [("sx", C.c_int), ("sy", C.c_int)]


# Every symbol the headers in include/ declare with C linkage.
EXPORTS = [
    "qpsdm_window_params", "qpsdm_clamp_band", "qpsdm_npfar", "qpsdm_gain_coefs", "qpsdm_iaztab",
    "qpsdm_trace_geometry", "qpsdm_tables_build", "qpsdm_tables_f16b_bytes",
    "qpsdm_tables_quantize", "qpsdm_tables_blend", "qpsdm_read_tidx", "qpsdm_write_tidx",
    "qpsdm_load_line_table", "qpsdm_trace_spectra", "qpsdm_bank_row_bytes", "qpsdm_qbank", "qpsdm_qbank_noalloc",
    "qpsdm_image_shot",
    "qpsdm_count_shot", "qpsdm_image_scale", "qpsdm_last_error", "qpsdm_version",
    "qpsdm_su_fgettr", "qpsdm_su_fputtr", "qpsdm_su_sgettr", "qpsdm_su_sputtr", "qpsdm_su_decode",
    "getscale", "get_vel_coor", "time_offset_mute",
    "qpsdm_pkg_fetch_and_taper", "qpsdm_pkg_parseint", "qpsdm_pkg_parsetuple",
    "qpsdm_pkg_readparfile_i", "qpsdm_pkg_readparfile_f", "qpsdm_pkg_readparfile_s",
    "qpsdm_pkg_readparfile_ii", "qpsdm_pkg_initargs", "qpsdm_pkg_getparint",
    "qpsdm_pkg_getparfloat", "qpsdm_pkg_getpardouble", "qpsdm_pkg_getparstr",
    "qpsdm_pkg_time_offset_mute_santity", "qpsdm_pkg_parsepath", "qpsdm_pkg_strrtrim",
    "qpsdm_pkg_free", "qpsdm_ctx_create", "qpsdm_ctx_destroy", "qpsdm_ctx_set_tables",
    "qpsdm_ctx_build_tables", "qpsdm_ctx_push_su", "qpsdm_ctx_push_traces", "qpsdm_ctx_psdm_kernel",
    "qpsdm_ctx_flush", "qpsdm_ctx_image", "qpsdm_ctx_reset_image", "qpsdm_ctx_traces_pushed",
    "qpsdm_write_image_su", "qpsdm_image_batch", "qpsdm_count_batch", "qpsdm_batch_graph_create",
    "qpsdm_graph_launch", "qpsdm_graph_destroy", "qpsdm_image_batch_src",
    "qpsdm_surrogate_nparams", "qpsdm_surrogate_create", "qpsdm_surrogate_destroy",
    "qpsdm_surrogate_get_params", "qpsdm_surrogate_set_params", "qpsdm_surrogate_samples",
    "qpsdm_surrogate_eval", "qpsdm_surrogate_grad", "qpsdm_surrogate_train", "qpsdm_surrogate_mask_bricks",
    "qpsdm_surrogate_mask", "qpsdm_surrogate_src_bytes", "qpsdm_surrogate_src_tables",
    "qpsdm_src_tables_nodes", "qpsdm_copy_f32", "qpsdm_tables_write_ref", "qpsdm_tables_read_ref",
    "qpsdm_rays",
    "qpsdm_tables_bytes", "qpsdm_tables_lookup", "qpsdm_image_batch_grid",
    "qpsdm_segy_fgethd", "qpsdm_segy_initkeyindex", "qpsdm_segy_setkeyindex", "qpsdm_segy_fgettr",
    "qpsdm_segy_sgettr", "qpsdm_segy_decode", "qpsdm_segy_decode_dev", "qpsdm_ctx_push_segy",
    "qpsdm_gather_bins", "qpsdm_gathers_shot", "qpsdm_ctx_set_gathers", "qpsdm_ctx_gathers",
]

SEGY_TOTAL_KEY_NUM = 75


class KeyIndex(C.Structure):
    """struct keyindex of segy.h: 1-based start byte and width of a key."""
    _fields_ = [("start", C.c_int), ("bytes", C.c_int)]


KeyIndexArray = KeyIndex * SEGY_TOTAL_KEY_NUM

_lib = None


class QpsdmError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QpsdmError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                             "(make -C q_psdm_amd/csrc); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        vp, fp, ip = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int)
        L.qpsdm_last_error.restype = C.c_char_p
        L.qpsdm_version.restype = C.c_char_p
        L.qpsdm_window_params.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                          C.POINTER(Window)]
        L.qpsdm_npfar.argtypes = [C.c_int]
        L.qpsdm_clamp_band.argtypes = [C.c_float, fp, fp]
        L.qpsdm_gain_coefs.argtypes = [C.c_float, fp]
        L.qpsdm_iaztab.argtypes = [ip, C.c_int]
        L.qpsdm_trace_geometry.argtypes = [C.POINTER(GeomParams), vp, C.c_int, vp, vp, vp]
        L.qpsdm_tables_build.argtypes = [C.POINTER(Model), C.POINTER(TGeom), vp, vp]
        L.qpsdm_tables_f16b_bytes.restype = C.c_size_t
        L.qpsdm_tables_f16b_bytes.argtypes = [C.POINTER(TGeom)]
        L.qpsdm_tables_quantize.argtypes = [C.POINTER(TGeom), vp, vp, C.c_int, vp]
        L.qpsdm_tables_blend.argtypes = [vp, vp, C.c_size_t, C.c_float, vp, vp]
        L.qpsdm_read_tidx.argtypes = [C.c_char_p, C.POINTER(TidxHeader), vp, C.c_int]
        L.qpsdm_write_tidx.argtypes = [C.c_char_p, C.POINTER(TidxHeader), vp]
        L.qpsdm_load_line_table.restype = C.c_long
        L.qpsdm_load_line_table.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int,
                                            vp, C.c_size_t, ip, ip]
        L.qpsdm_trace_spectra.argtypes = [C.POINTER(BankParams), vp, C.c_int, vp, vp, vp]
        L.qpsdm_qbank.argtypes = [C.POINTER(BankParams), vp, C.c_int, vp, C.c_int, vp]
        L.qpsdm_qbank_noalloc.argtypes = [C.POINTER(BankParams), vp, C.c_int, vp, C.c_int, vp]
        L.qpsdm_bank_row_bytes.restype = C.c_size_t
        L.qpsdm_bank_row_bytes.argtypes = [C.c_int, C.c_int, C.c_int]
        L.qpsdm_image_shot.argtypes = [C.POINTER(ImageParams), C.POINTER(Launch), vp, vp, C.c_int,
                                       C.c_int, vp, C.c_int, vp, C.c_int, vp]
        L.qpsdm_count_shot.argtypes = [C.POINTER(ImageParams), vp, C.c_int, C.c_int, vp, vp]
        L.qpsdm_image_scale.argtypes = [vp, C.c_size_t, C.c_float, vp]
        L.qpsdm_su_decode.restype = C.c_long
        L.qpsdm_su_decode.argtypes = [vp, C.c_size_t, C.c_int, vp, vp]
        L.qpsdm_su_sgettr.argtypes = [vp, vp]
        L.qpsdm_segy_fgethd.argtypes = [vp, vp]
        L.qpsdm_segy_initkeyindex.argtypes = [vp]
        L.qpsdm_segy_setkeyindex.argtypes = [vp, C.c_char_p, C.c_int, C.c_int]
        L.qpsdm_segy_sgettr.argtypes = [vp, vp, C.c_int, vp]
        L.qpsdm_segy_decode.restype = C.c_long
        L.qpsdm_segy_decode.argtypes = [vp, C.c_size_t, C.c_int, C.c_int, vp, vp, vp]
        L.qpsdm_segy_decode_dev.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
        L.qpsdm_ctx_push_segy.argtypes = [vp, vp, C.c_size_t, C.c_int, vp]
        L.qpsdm_gather_bins.argtypes = [C.POINTER(GatherParams), ip, ip]
        L.qpsdm_ctx_set_gathers.argtypes = [vp, C.POINTER(GatherParams), vp, C.c_int]
        L.qpsdm_ctx_gathers.argtypes = [vp, vp, vp, C.c_int]
        L.qpsdm_gathers_shot.argtypes = [C.POINTER(ImageParams), C.POINTER(GatherParams), vp, C.c_int, vp, vp, vp,
                                         C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp]
        L.qpsdm_su_sputtr.argtypes = [vp, vp]
        L.getscale.restype = C.c_float
        L.getscale.argtypes = [C.c_short]
        L.time_offset_mute.restype = C.c_float
        L.time_offset_mute.argtypes = [fp, C.c_int, C.c_float]
        L.get_vel_coor.argtypes = [C.c_char_p, C.c_int, vp, C.POINTER(C.c_short), C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_int, C.c_int]
        L.qpsdm_pkg_fetch_and_taper.argtypes = [fp, C.c_int, C.c_int, C.c_int, fp]
        L.qpsdm_pkg_parseint.restype = C.POINTER(C.c_int)
        L.qpsdm_pkg_parseint.argtypes = [C.c_char_p, ip]
        L.qpsdm_pkg_parsetuple.restype = C.POINTER(C.c_float)
        L.qpsdm_pkg_parsetuple.argtypes = [C.c_char_p, ip, ip]
        L.qpsdm_pkg_strrtrim.restype = C.c_char_p
        L.qpsdm_pkg_free.argtypes = [vp]
        L.qpsdm_ctx_create.argtypes = [C.POINTER(CtxParams), C.POINTER(vp)]
        L.qpsdm_ctx_destroy.argtypes = [vp]
        L.qpsdm_ctx_set_tables.argtypes = [vp, vp, C.c_int]
        L.qpsdm_ctx_build_tables.argtypes = [vp, C.POINTER(Model)]
        L.qpsdm_ctx_push_su.argtypes = [vp, vp, C.c_size_t]
        L.qpsdm_ctx_push_traces.argtypes = [vp, vp, vp, C.c_int]
        L.qpsdm_ctx_psdm_kernel.argtypes = [vp, vp]
        L.qpsdm_ctx_flush.argtypes = [vp]
        L.qpsdm_ctx_image.argtypes = [vp, vp, C.c_int]
        L.qpsdm_ctx_reset_image.argtypes = [vp]
        L.qpsdm_ctx_traces_pushed.restype = C.c_longlong
        L.qpsdm_ctx_traces_pushed.argtypes = [vp]
        L.qpsdm_write_image_su.argtypes = [C.c_char_p, C.c_int, vp, C.c_int, C.c_int, C.c_float, C.c_int,
                                           C.c_float, C.c_int, vp, C.c_int, vp, C.c_short]
        L.qpsdm_image_batch.argtypes = [C.POINTER(ImageParams), C.POINTER(Launch), vp, vp, vp, C.c_int,
                                        vp, C.c_int, vp, C.c_int, vp, C.c_int, vp]
        L.qpsdm_count_batch.argtypes = [C.POINTER(ImageParams), vp, vp, C.c_int, vp, vp]
        L.qpsdm_batch_graph_create.argtypes = [C.POINTER(BankParams), vp, C.c_int, vp, C.c_int,
                                               C.POINTER(ImageParams), C.POINTER(Launch), vp, vp, vp,
                                               C.c_int, vp, C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_int,
                                               C.POINTER(vp)]
        L.qpsdm_copy_f32.argtypes = [vp, vp, C.c_size_t, vp]
        L.qpsdm_image_batch_grid.argtypes = [C.POINTER(ImageParams), C.POINTER(Launch), C.c_int, C.c_int,
                                             C.POINTER(C.c_int)]
        L.qpsdm_tables_bytes.restype = C.c_size_t
        L.qpsdm_tables_bytes.argtypes = [C.POINTER(TGeom), C.c_int]
        L.qpsdm_tables_lookup.argtypes = [C.POINTER(ImageParams), vp, C.c_int, C.c_float, C.c_float, vp, C.c_int,
                                          vp, vp]
        L.qpsdm_image_batch_src.argtypes = [C.POINTER(ImageParams), C.POINTER(Launch), vp, vp, vp, C.c_int,
                                            vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, vp]
        f5 = C.c_float * 5
        L.qpsdm_surrogate_create.argtypes = [C.POINTER(Model), f5, f5, C.c_float * 3, C.c_ulonglong,
                                             C.POINTER(vp)]
        L.qpsdm_surrogate_destroy.argtypes = [vp]
        L.qpsdm_surrogate_get_params.argtypes = [vp, vp]
        L.qpsdm_surrogate_set_params.argtypes = [vp, vp]
        L.qpsdm_surrogate_samples.argtypes = [vp, C.c_ulonglong, C.c_longlong, C.c_int, vp, vp, vp]
        L.qpsdm_surrogate_eval.argtypes = [vp, vp, C.c_int, vp, vp]
        L.qpsdm_surrogate_grad.argtypes = [vp, vp, vp, C.c_int, vp, C.POINTER(C.c_float), vp]
        L.qpsdm_surrogate_train.argtypes = [vp, vp, vp, C.c_longlong, C.c_int, C.c_int, C.c_float, C.c_float,
                                            vp, vp]
        L.qpsdm_surrogate_mask_bricks.restype = C.c_longlong
        L.qpsdm_surrogate_mask_bricks.argtypes = [C.POINTER(TGeom)]
        L.qpsdm_surrogate_mask.argtypes = [vp, C.POINTER(TGeom), vp, C.c_float, vp, vp, vp]
        L.qpsdm_surrogate_src_bytes.restype = C.c_size_t
        L.qpsdm_surrogate_src_bytes.argtypes = [C.POINTER(ImageParams), C.c_int]
        L.qpsdm_surrogate_src_tables.argtypes = [vp, C.POINTER(ImageParams), vp, vp, C.c_int, vp, vp, vp, vp]
        L.qpsdm_rays.argtypes = [C.POINTER(Model), vp, C.c_int, vp, vp]
        L.qpsdm_tables_write_ref.argtypes = [C.POINTER(TGeom), vp, C.POINTER(RefLayout), C.c_char_p]
        L.qpsdm_tables_read_ref.argtypes = [C.POINTER(TGeom), C.POINTER(RefLayout), C.c_char_p, vp]
        L.qpsdm_src_tables_nodes.argtypes = [C.POINTER(ImageParams), vp, vp, C.c_int, vp, vp, vp]
        L.qpsdm_graph_launch.argtypes = [vp, vp]
        L.qpsdm_graph_destroy.argtypes = [vp]
        _lib = L
    return _lib


def check(rc):
    """Raise QpsdmError with the library's message when rc < 0."""
    if rc < 0:
        raise QpsdmError(lib().qpsdm_last_error().decode())
    return rc


def exported_symbols():
    """Names of EXPORTS that the loaded library resolves."""
    L = lib()
    return [n for n in EXPORTS if hasattr(L, n)]
