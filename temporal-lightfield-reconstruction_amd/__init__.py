"""Python binding of libtlr (ctypes over the C ABI in include/tlr.h).

The product is the native library (lib/libtlr.so: HIP kernels for gfx950 +
C ABI); this module only marshals arguments.  Every compute call goes to the
device; there is no host fallback: on a machine without a usable GPU, Ctx()
raises.

Classes mirroring the reference interface (TreeGather,
src/reconstruction_lib/reconstruction/Reconstruction.hpp:42-48):
    UVTSampleBuffer  sample records + header (SampleBuffer.hpp:138-183)
    CameraParams     reconstruction flags used by the path (CameraParams.hpp:92-107)
    TreeGather       build in the constructor, reconstruct_dof_motion(image)
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
TLR_LIB = os.environ.get("TLR_LIBRARY") or os.path.join(LIB_DIR, "libtlr.so")
SCENE_LIB = os.path.join(LIB_DIR, "libtlr_scene.so")

STATUS = {0: "TLR_OK", 1: "TLR_E_INVALID", 2: "TLR_E_IO", 3: "TLR_E_ALL_DISCARDED",
          4: "TLR_E_UNSUPPORTED_SPP", 5: "TLR_E_IMAGE_TOO_SMALL", 6: "TLR_E_HIP",
          7: "TLR_E_CAPACITY", 8: "TLR_E_STATE"}

EXPORTED = [
    "tlr_create", "tlr_destroy", "tlr_last_error", "tlr_load_sample_file", "tlr_save_sample_file",
    "tlr_free", "tlr_set_samples", "tlr_set_samples_soa", "tlr_build", "tlr_build_ex", "tlr_reconstruct", "tlr_reconstruct_noalloc", "tlr_sync",
    "tlr_get_build_info", "tlr_copy_reprojected", "tlr_copy_layout", "tlr_copy_initial_leaves",
    "tlr_copy_leaves", "tlr_copy_hierarchy", "tlr_copy_query_colors", "tlr_copy_query_stages",
    "tlr_copy_debug_image", "tlr_copy_query_gather",
    "tlr_output_pattern", "tlr_calibrate_traffic",
    "tlr_gate_set_params", "tlr_gate_get_params", "tlr_gate_set_threshold", "tlr_gate_forward",
    "tlr_gate_grad", "tlr_gate_collect", "tlr_gate_append_training", "tlr_gate_clear_training",
    "tlr_gate_training_size", "tlr_gate_train",
    "tlr_batch_set_samples", "tlr_reconstruct_batch", "tlr_batch_retune", "tlr_batch_set_mode", "tlr_use_slot",
    "tlr_set_shadow_inputs", "tlr_reconstruct_queries", "tlr_shadow_setup", "tlr_reconstruct_dof_motion_shadows",
]
GATE_FEATURES = 24
GATE_PARAMS = 5825


class TlrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS.get(code, code), msg))
        self.code = code


class Header(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("spp", C.c_int),
                ("coc_c0", C.c_float), ("coc_c1", C.c_float)]


class Soa(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("x", "y", "w", "u", "v", "t", "r", "g", "b",
                                          "mv_x", "mv_y", "mv_w")]


class ReconParams(C.Structure):
    _fields_ = [("out_spp", C.c_int), ("override_uvt", C.c_int), ("uvt", C.c_float * 3),
                ("lens_gaussian", C.c_int), ("time_gaussian", C.c_int),
                ("collect_stats", C.c_int), ("gate", C.c_int), ("path", C.c_int),
                ("list_cap", C.c_int), ("debug_image", C.c_int), ("reserved", C.c_int)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "leaf_nodes", "samples_in_leaf_nodes", "samples_within_1r", "samples_fetched",
        "surfaces", "num_2r", "surfaces_merged", "at_least_one", "count", "overflow",
        "build_ms", "query_ms", "combine_ms", "total_ms", "num_leaves", "num_samples",
        "fast_ms", "general_ms", "general_queries", "gate_candidates", "gate_ms", "gate_rejected",
        "fast_leaf_nodes", "fast_samples_in_leaf_nodes", "fast_samples_within_1r",
        "wide_ms", "wide_queries", "wide_leaf_nodes", "wide_samples_in_leaf_nodes", "wide_samples_within_1r",
        "light_queries")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class BatchItem(C.Structure):
    _fields_ = [("slot", C.c_int), ("aperture_adjust", C.c_float), ("focal_adjust", C.c_float),
                ("out_spp", C.c_int), ("rgba_out", C.c_void_p), ("out_on_device", C.c_int),
                ("build_ms", C.POINTER(C.c_float)), ("gate", C.c_int), ("done_ms", C.POINTER(C.c_float))]


class BatchStats(C.Structure):
    _fields_ = [("build_ms", C.c_double), ("query_ms", C.c_double), ("total_ms", C.c_double),
                ("output_samples", C.c_double), ("general_queries", C.c_double),
                ("fast_leaf_cap", C.c_int), ("fast_block", C.c_int), ("general_blocks", C.c_int),
                ("graph_reused", C.c_int), ("autotuned", C.c_int), ("chained", C.c_int),
                ("gate_accepted", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ShadowParams(C.Structure):
    """tlr_shadow_params: matrices (row-major 4x4) and camera data of
    reconstructDofMotionShadows."""
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("camera_coc", C.c_float * 2),
                ("light_coc", C.c_float * 2), ("camera_far", C.c_float)] + [
        (n, C.c_float * 16) for n in ("c2l", "zfromw_ndc", "inv_projection", "shadow2_inv_proj",
                                      "shadow2_inv_t", "world_inv_t")]


class BuildInfo(C.Structure):
    _fields_ = [("num_input", C.c_int), ("num_kept", C.c_int), ("num_bucketed", C.c_int),
                ("spp_pow2", C.c_int), ("coc_adj", C.c_float * 2), ("bbmin", C.c_float * 2),
                ("bbmax", C.c_float * 2), ("scale", C.c_float), ("bbmin_int", C.c_int * 2),
                ("grid_w", C.c_int), ("grid_h", C.c_int), ("max_hits", C.c_int),
                ("num_initial_leaves", C.c_int), ("num_leaves", C.c_int),
                ("num_valid_leaves", C.c_int)]


class SceneDesc(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("spp", C.c_int),
                ("coc_c0", C.c_float), ("coc_c1", C.c_float),
                ("num_quads", C.c_int), ("quad_min_frac", C.c_float), ("quad_max_frac", C.c_float),
                ("num_occluders", C.c_int), ("occluder_min_px", C.c_float),
                ("occluder_max_px", C.c_float), ("w_min", C.c_float), ("w_max", C.c_float),
                ("w_log_uniform", C.c_int), ("mv_xy_px", C.c_float), ("mv_w_frac", C.c_float),
                ("static_frac", C.c_float), ("checker_period_px", C.c_float),
                ("seed", C.c_uint64), ("threads", C.c_int)]


_tlr = None
_scene = None


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib():
    """Loads lib/libtlr.so (raises if it has not been built)."""
    global _tlr
    if _tlr is None:
        if not os.path.exists(TLR_LIB):
            raise RuntimeError("libtlr.so is not built (run __graft_entry__.build())")
        L = C.CDLL(TLR_LIB)
        V = C.c_void_p
        L.tlr_create.argtypes = [C.POINTER(V), C.c_int]
        L.tlr_destroy.argtypes = [V]
        L.tlr_last_error.restype = C.c_char_p
        L.tlr_last_error.argtypes = [V]
        L.tlr_load_sample_file.argtypes = [C.c_char_p, C.POINTER(Header), C.POINTER(C.POINTER(C.c_float)),
                                           C.POINTER(C.c_int64)]
        L.tlr_save_sample_file.argtypes = [C.c_char_p, C.POINTER(Header), V, C.c_int, C.c_int]
        L.tlr_free.argtypes = [V]
        L.tlr_set_samples.argtypes = [V, C.POINTER(Header), V, C.c_int64]
        L.tlr_set_samples_soa.argtypes = [V, C.POINTER(Header), C.POINTER(Soa), C.c_int]
        L.tlr_build.argtypes = [V, C.c_float, C.c_float]
        L.tlr_build_ex.argtypes = [V, C.c_float, C.c_float, V]
        L.tlr_copy_debug_image.argtypes = [V, V]
        L.tlr_copy_query_gather.argtypes = [V, V, C.c_int, C.c_float, C.c_int, V, V, C.c_int, V, V, V]
        L.tlr_reconstruct.argtypes = [V, C.POINTER(ReconParams), V, C.c_int, C.POINTER(Stats)]
        L.tlr_reconstruct_noalloc.argtypes = [V, C.POINTER(ReconParams), V, C.POINTER(Stats)]
        L.tlr_sync.argtypes = [V]
        L.tlr_get_build_info.argtypes = [V, C.POINTER(BuildInfo)]
        L.tlr_copy_reprojected.argtypes = [V, V, V, V, V]
        L.tlr_copy_layout.argtypes = [V, V]
        L.tlr_copy_initial_leaves.argtypes = [V, V]
        L.tlr_copy_leaves.argtypes = [V, V, V]
        L.tlr_copy_hierarchy.argtypes = [V, V, V, V]
        L.tlr_batch_set_samples.argtypes = [V, C.c_int, C.POINTER(Header), V, C.c_int64]
        L.tlr_reconstruct_batch.argtypes = [V, C.POINTER(BatchItem), C.c_int, C.POINTER(BatchStats)]
        L.tlr_batch_retune.argtypes = [V]
        L.tlr_batch_set_mode.argtypes = [V, C.c_int]
        L.tlr_use_slot.argtypes = [V, C.c_int]
        L.tlr_copy_query_colors.argtypes = [V, V]
        L.tlr_copy_query_stages.argtypes = [V, V]
        L.tlr_output_pattern.restype = C.c_int
        L.tlr_output_pattern.argtypes = [C.c_int, C.c_int, V, V]
        L.tlr_calibrate_traffic.argtypes = [V, C.c_int, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.tlr_gate_set_params.argtypes = [V, V]
        L.tlr_gate_get_params.argtypes = [V, V]
        L.tlr_gate_set_threshold.argtypes = [V, C.c_float]
        L.tlr_gate_forward.argtypes = [V, V, C.c_int64, V]
        L.tlr_gate_grad.argtypes = [V, V, V, C.c_int64, C.c_float, V, V]
        L.tlr_gate_collect.argtypes = [V, C.POINTER(ReconParams), C.c_int, C.c_int64, V, V, V, V, V,
                                       C.POINTER(C.c_int64)]
        L.tlr_gate_append_training.argtypes = [V, C.POINTER(ReconParams), C.c_int]
        L.tlr_gate_clear_training.argtypes = [V]
        L.tlr_gate_training_size.restype = C.c_int64
        L.tlr_gate_training_size.argtypes = [V]
        L.tlr_gate_train.argtypes = [V, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint64, V]
        L.tlr_set_shadow_inputs.argtypes = [V, V, V, C.c_int64]
        L.tlr_reconstruct_queries.argtypes = [V, C.POINTER(ReconParams), V, C.c_int64, C.c_int, V, V,
                                              C.POINTER(Stats)]
        L.tlr_shadow_setup.argtypes = [V, V, V, V, C.c_int, C.c_int, V, V, C.c_float, C.POINTER(ShadowParams)]
        L.tlr_reconstruct_dof_motion_shadows.argtypes = [V, V, C.POINTER(ReconParams), C.POINTER(ShadowParams), V,
                                                         C.c_int, C.POINTER(Stats)]
        _tlr = L
    return _tlr


def scene_lib():
    global _scene
    if _scene is None:
        if not os.path.exists(SCENE_LIB):
            raise RuntimeError("libtlr_scene.so is not built (run __graft_entry__.build())")
        L = C.CDLL(SCENE_LIB)
        L.tlr_scene_preset.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(SceneDesc)]
        L.tlr_scene_generate.argtypes = [C.POINTER(SceneDesc), C.c_void_p]
        _scene = L
    return _scene


def scene_desc(preset, width, height, spp, seed, threads=0, **overrides):
    d = SceneDesc()
    if scene_lib().tlr_scene_preset(preset, width, height, spp, C.c_uint64(seed), C.byref(d)) != 0:
        raise ValueError("unknown scene preset %r" % preset)
    d.threads = threads
    for k, v in overrides.items():
        setattr(d, k, v)
    return d


def make_scene(preset, width, height, spp, seed, threads=0, **overrides):
    """Synthetic sample buffer: returns (UVTSampleBuffer)."""
    d = scene_desc(preset, width, height, spp, seed, threads, **overrides)
    rec = np.empty((height, width, spp, 16), np.float32)
    if scene_lib().tlr_scene_generate(C.byref(d), _p(rec)) != 0:
        raise RuntimeError("scene generation failed")
    return UVTSampleBuffer(width, height, spp, d.coc_c0, d.coc_c1, rec)


def shadow_setup(cam_proj, cam_to_world, light_proj, light_world_to_cam, width, height, cam_coc, light_coc,
                 cam_far):
    """tlr_shadow_setup: ShadowParams from the cameras' row-major 4x4 matrices."""
    m = [np.ascontiguousarray(a, np.float32).reshape(16) for a in (cam_proj, cam_to_world, light_proj,
                                                                   light_world_to_cam)]
    cc = np.asarray(cam_coc, np.float32)
    lc = np.asarray(light_coc, np.float32)
    sp = ShadowParams()
    rc = lib().tlr_shadow_setup(_p(m[0]), _p(m[1]), _p(m[2]), _p(m[3]), width, height, _p(cc), _p(lc),
                                float(cam_far), C.byref(sp))
    if rc != 0:
        raise TlrError(rc, "tlr_shadow_setup")
    return sp


def output_pattern(out_spp, override_uvt=None):
    uvt = np.zeros(3, np.float32)
    if override_uvt is not None:
        uvt[:] = override_uvt
    flag = 1 if override_uvt is not None else 0
    M = lib().tlr_output_pattern(out_spp, flag, _p(uvt), None)
    out = np.zeros((64 * M, 5), np.float32)
    lib().tlr_output_pattern(out_spp, flag, _p(uvt), _p(out))
    return out


class UVTSampleBuffer:
    """Sample records (v1.3 16-float layout) plus header; indexed (y*W+x)*spp+i."""

    def __init__(self, width, height, spp, coc_c0, coc_c1, records):
        self.width, self.height, self.spp = int(width), int(height), int(spp)
        self.coc = (float(coc_c0), float(coc_c1))
        self.records = np.ascontiguousarray(records, np.float32).reshape(-1, 16)
        # spp 0: an irregular buffer, any number of records
        if self.spp and self.records.shape[0] != self.width * self.height * self.spp:
            raise ValueError("record count does not match the header")

    @classmethod
    def load(cls, path):
        h = Header()
        ptr = C.POINTER(C.c_float)()
        n = C.c_int64(0)
        rc = lib().tlr_load_sample_file(path.encode(), C.byref(h), C.byref(ptr), C.byref(n))
        if rc != 0:
            raise TlrError(rc, "cannot load %s" % path)
        try:
            arr = np.ctypeslib.as_array(ptr, shape=(n.value * 16,)).copy()
        finally:
            lib().tlr_free(C.cast(ptr, C.c_void_p))
        return cls(h.width, h.height, h.spp, h.coc_c0, h.coc_c1, arr)

    def save(self, path, binary=False, separate_header=False):
        rc = lib().tlr_save_sample_file(path.encode(), C.byref(self.header()), _p(self.records),
                                        int(binary), int(separate_header))
        if rc != 0:
            raise TlrError(rc, "cannot save %s" % path)

    def header(self):
        return Header(self.width, self.height, self.spp, self.coc[0], self.coc[1])


def coc_coeffs(projection, window_width, aperture_radius, focus_distance):
    """CameraParams::getCocCoeffs (CameraParams.hpp:97-118) in FP32 with the
    framework's operation order: m = projection * getDofMatrix((0.5, 0)) -
    projection (products summed over k = 0..3 from 0, Math.hpp:1024-1035),
    coefficients (-m03, m02) * window width."""
    f = np.float32
    P = np.asarray(projection, np.float32).reshape(4, 4)
    A, fd = f(aperture_radius), f(focus_distance)
    dof = np.eye(4, dtype=np.float32)
    dof[0, 2] = f(f(f(0.5) * A) / fd)     # uv * apertureRadius / focusDistance, uv = (1,0) / 2
    dof[1, 2] = f(f(f(0.0) * A) / fd)
    dof[0, 3] = f(f(0.5) * A)              # uv * apertureRadius
    dof[1, 3] = f(f(0.0) * A)
    m = np.zeros((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            rr = f(0.0)
            for k in range(4):
                rr = f(rr + f(P[i, k] * dof[k, j]))
            m[i, j] = f(rr - P[i, j])
    ww = f(window_width)
    return np.array([f(-m[0, 3] * ww), f(m[0, 2] * ww)], np.float32)


class CameraParams:
    """The CameraParams fields the reconstruction reads (CameraParams.hpp:92-107).
    override_refocus_distance (with projection, window_width and
    aperture_radius, which it needs) refocuses through new CoC coefficients
    (Reconstruction.cpp:88-95)."""

    def __init__(self, out_spp=128, override_uvt=None, lens_gaussian=False, time_gaussian=False,
                 override_refocus_distance=None, projection=None, window_width=None, aperture_radius=None):
        self.out_spp = out_spp
        self.override_uvt = override_uvt
        self.lens_gaussian = lens_gaussian
        self.time_gaussian = time_gaussian
        self.override_refocus_distance = override_refocus_distance
        self.projection = projection
        self.window_width = window_width
        self.aperture_radius = aperture_radius

    def refocus_coc(self):
        if self.override_refocus_distance is None:
            return None
        if self.projection is None or self.window_width is None or self.aperture_radius is None:
            raise ValueError("overrideRefocusDistance needs projection, window_width and aperture_radius")
        return coc_coeffs(self.projection, self.window_width, self.aperture_radius, self.override_refocus_distance)


class Ctx:
    """One libtlr context (device buffers + stream) on one GPU."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib().tlr_create(C.byref(self.h), device)
        if rc != 0:
            raise TlrError(rc, "tlr_create failed (no usable GPU?)")

    def close(self):
        if self.h:
            lib().tlr_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise TlrError(rc, lib().tlr_last_error(self.h).decode())

    def set_samples(self, sbuf):
        self.sbuf = sbuf
        self._chk(lib().tlr_set_samples(self.h, C.byref(sbuf.header()), _p(sbuf.records),
                                        sbuf.records.shape[0]))

    def build(self, aperture_adjust=1.0, focal_adjust=1.0, coc_override=None):
        if coc_override is None:
            self._chk(lib().tlr_build(self.h, aperture_adjust, focal_adjust))
        else:
            cc = np.ascontiguousarray(coc_override, np.float32)
            self._chk(lib().tlr_build_ex(self.h, aperture_adjust, focal_adjust, _p(cc)))

    def debug_image(self):
        """Debug image of the last reconstruct(debug_image=True)."""
        out = np.empty((self.sbuf.height, self.sbuf.width, 4), np.float32)
        self._chk(lib().tlr_copy_debug_image(self.h, _p(out)))
        return out

    def query_gather(self, o5, radius, leaf_cap=4096, sample_cap=1 << 15):
        """The plain method's gather of output samples o5 ((n, 5): x,y,u,v,t)
        at `radius`: per output sample (leaf ids front to back, gathered
        sample indices in tree layout, their surfaces); None past capacity."""
        o5 = np.ascontiguousarray(o5, np.float32).reshape(-1, 5)
        n = o5.shape[0]
        lv = np.zeros((n, leaf_cap), np.int32)
        sm = np.zeros((n, sample_cap), np.int32)
        sf = np.zeros((n, sample_cap), np.int32)
        nl = np.zeros(n, np.int32)
        ns = np.zeros(n, np.int32)
        self._chk(lib().tlr_copy_query_gather(self.h, _p(o5), n, float(radius), leaf_cap, _p(lv), _p(nl), sample_cap,
                                              _p(sm), _p(sf), _p(ns)))
        out = []
        for j in range(n):
            if nl[j] < 0:
                out.append(None)
            else:
                out.append((lv[j, :min(nl[j], leaf_cap)], sm[j, :min(ns[j], sample_cap)],
                            sf[j, :min(ns[j], sample_cap)]))
        return out

    def reconstruct(self, out_spp=16, override_uvt=None, lens_gaussian=False, time_gaussian=False,
                    collect_stats=False, gate=False, path=0, out=None, list_cap=0, device_only=False,
                    debug_image=False):
        """Reconstructs into `out` (host, H x W x 4 float32; allocated when
        None).  device_only: the image stays in the context's device buffer
        and (None, stats) is returned."""
        p = ReconParams()
        p.out_spp = out_spp
        p.override_uvt = 1 if override_uvt is not None else 0
        if override_uvt is not None:
            for k in range(3):
                p.uvt[k] = override_uvt[k]
        p.lens_gaussian = int(lens_gaussian)
        p.time_gaussian = int(time_gaussian)
        p.collect_stats = int(collect_stats)
        p.gate = int(gate)
        p.path = int(path)
        p.list_cap = int(list_cap)
        p.debug_image = int(debug_image)
        if device_only:
            img = None
        else:
            img = out if out is not None else np.empty((self.sbuf.height, self.sbuf.width, 4), np.float32)
        st = Stats()
        self._chk(lib().tlr_reconstruct(self.h, C.byref(p), _p(img) if img is not None else None, 0, C.byref(st)))
        return img, st.as_dict()

    def set_shadow_inputs(self, wg=None, density=None):
        """Per input sample w-gradients ((N, 2), the records' floats 14-15)
        and light-space densities (N) for shadow reconstruction; None: 0, 1."""
        n = self.sbuf.records.shape[0]
        wg = None if wg is None else np.ascontiguousarray(wg, np.float32).reshape(n, 2)
        density = None if density is None else np.ascontiguousarray(density, np.float32).reshape(n)
        self._chk(lib().tlr_set_shadow_inputs(self.h, _p(wg), _p(density), n))

    def reconstruct_queries(self, queries, shadow=False, lens_gaussian=False, time_gaussian=False):
        """Explicit output samples (n x 8: x, y, u, v, t, w, density, 0):
        (colours n x 4, extended results n x 8, stats)."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 8)
        p = ReconParams()
        p.lens_gaussian = int(lens_gaussian)
        p.time_gaussian = int(time_gaussian)
        rgba = np.empty((q.shape[0], 4), np.float32)
        ext = np.empty((q.shape[0], 8), np.float32)
        st = Stats()
        self._chk(lib().tlr_reconstruct_queries(self.h, C.byref(p), _p(q), q.shape[0], 1 if shadow else 0,
                                                _p(rgba), _p(ext), C.byref(st)))
        return rgba, ext, st.as_dict()

    def reconstruct_dof_motion_shadows(self, light, sp, out_spp=16, override_uvt=None, lens_gaussian=False,
                                       time_gaussian=False, collect_stats=False):
        """reconstructDofMotionShadows with `light` (a built Ctx of the same
        size) as the shadow TreeGather: (image, stats)."""
        p = ReconParams()
        p.out_spp = out_spp
        p.collect_stats = int(collect_stats)
        p.override_uvt = 1 if override_uvt is not None else 0
        if override_uvt is not None:
            for k in range(3):
                p.uvt[k] = override_uvt[k]
        p.lens_gaussian = int(lens_gaussian)
        p.time_gaussian = int(time_gaussian)
        img = np.empty((self.sbuf.height, self.sbuf.width, 4), np.float32)
        st = Stats()
        self._chk(lib().tlr_reconstruct_dof_motion_shadows(self.h, light.h, C.byref(p), C.byref(sp), _p(img), 0,
                                                           C.byref(st)))
        return img, st.as_dict()

    def reconstruct_noalloc(self, device_out, out_spp=16, override_uvt=None, lens_gaussian=False,
                            time_gaussian=False, collect_stats=False, path=0):
        """tlr_reconstruct_noalloc into device memory at address device_out
        (W*H*4 floats); returns the stats."""
        p = self._params(out_spp=out_spp, override_uvt=override_uvt, lens_gaussian=lens_gaussian,
                         time_gaussian=time_gaussian, collect_stats=collect_stats, path=path)
        st = Stats()
        self._chk(lib().tlr_reconstruct_noalloc(self.h, C.byref(p), C.c_void_p(device_out), C.byref(st)))
        return st.as_dict()

    def batch_set_samples(self, slot, sbuf):
        """Uploads a sample buffer into batch slot `slot` (device-resident)."""
        self._chk(lib().tlr_batch_set_samples(self.h, slot, C.byref(sbuf.header()), _p(sbuf.records),
                                              sbuf.records.shape[0]))
        if not hasattr(self, "_slot_shapes"):
            self._slot_shapes = {}
        self._slot_shapes[slot] = (sbuf.height, sbuf.width)

    def reconstruct_batch(self, jobs, outs=None, gate=False, done=None):
        """jobs: list of (slot, out_spp[, aperture_adjust, focal_adjust]);
        outs: optional list of (H, W, 4) float32 arrays (host); gate: the
        learned gate for every item (or a list of flags); done: optional list
        that receives each item's completion time (ms from the batch's start).
        Returns (images, per-item build ms, batch stats)."""
        n = len(jobs)
        items = (BatchItem * n)()
        bms = (C.c_float * n)()
        dms = (C.c_float * n)()
        gates = list(gate) if isinstance(gate, (list, tuple)) else [bool(gate)] * n
        imgs = []
        for k, job in enumerate(jobs):
            slot, out_spp = job[0], job[1]
            items[k].slot = slot
            items[k].out_spp = out_spp
            items[k].aperture_adjust = job[2] if len(job) > 2 else 1.0
            items[k].focal_adjust = job[3] if len(job) > 3 else 1.0
            img = outs[k] if outs is not None else None
            if img is None and getattr(self, "_slot_shapes", {}).get(slot) is not None:
                h, w = self._slot_shapes[slot]
                img = np.empty((h, w, 4), np.float32)
            imgs.append(img)
            items[k].rgba_out = img.ctypes.data if img is not None else None
            items[k].out_on_device = 0
            items[k].build_ms = C.cast(C.addressof(bms) + 4 * k, C.POINTER(C.c_float))
            items[k].gate = 1 if gates[k] else 0
            items[k].done_ms = C.cast(C.addressof(dms) + 4 * k, C.POINTER(C.c_float))
        st = BatchStats()
        self._chk(lib().tlr_reconstruct_batch(self.h, items, n, C.byref(st)))
        if done is not None:
            done[:] = [dms[k] for k in range(n)]
        return imgs, [bms[k] for k in range(n)], st.as_dict()

    def calibrate_traffic(self, kind, nbytes):
        """One kernel moving a known number of bytes (profiler counter
        scaling): kind 0 streaming copy, 1 / 2 gathers of 32 / 64-byte
        records.  Returns (bytes read, bytes written)."""
        r, w = C.c_int64(0), C.c_int64(0)
        self._chk(lib().tlr_calibrate_traffic(self.h, kind, nbytes, C.byref(r), C.byref(w)))
        return r.value, w.value

    def use_slot(self, slot):
        """The context's input becomes batch slot `slot` (for build/reconstruct)."""
        self._chk(lib().tlr_use_slot(self.h, slot))
        h, w = self._slot_shapes[slot]

        class _Shape:
            pass
        sb = _Shape()
        sb.height, sb.width = h, w
        self.sbuf = sb

    def batch_retune(self):
        self._chk(lib().tlr_batch_retune(self.h))

    def batch_set_mode(self, mode):
        """Query-phase form of later batches: -1 autotuned, 0 merged, 1 chained."""
        self._chk(lib().tlr_batch_set_mode(self.h, mode))

    @staticmethod
    def _params(out_spp=16, override_uvt=None, lens_gaussian=False, time_gaussian=False,
                collect_stats=False, gate=False, path=0):
        p = ReconParams()
        p.out_spp = out_spp
        p.override_uvt = 1 if override_uvt is not None else 0
        if override_uvt is not None:
            for k in range(3):
                p.uvt[k] = override_uvt[k]
        p.lens_gaussian = int(lens_gaussian)
        p.time_gaussian = int(time_gaussian)
        p.collect_stats = int(collect_stats)
        p.gate = int(gate)
        p.path = int(path)
        return p

    # ---- learned gate ----
    def gate_set_params(self, params):
        params = np.ascontiguousarray(params, np.float32)
        assert params.size == GATE_PARAMS
        self._chk(lib().tlr_gate_set_params(self.h, _p(params)))

    def gate_get_params(self):
        out = np.empty(GATE_PARAMS, np.float32)
        self._chk(lib().tlr_gate_get_params(self.h, _p(out)))
        return out

    def gate_set_threshold(self, thr):
        self._chk(lib().tlr_gate_set_threshold(self.h, float(thr)))

    def gate_forward(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, GATE_FEATURES)
        out = np.empty(X.shape[0], np.float32)
        self._chk(lib().tlr_gate_forward(self.h, _p(X), X.shape[0], _p(out)))
        return out

    def gate_grad(self, X, y, pos_weight=1.0):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, GATE_FEATURES)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(GATE_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(lib().tlr_gate_grad(self.h, _p(X), _p(y), X.shape[0], pos_weight, _p(g), _p(loss)))
        return g, float(loss[0])

    def gate_collect(self, out_spp=16, stride=1, **kw):
        """Every stride-th hand-off entry of the current build: (features,
        labels, provisional colours, plain colours, output-sample indices)."""
        p = self._params(out_spp=out_spp, **kw)
        M = 1 if p.override_uvt else out_spp
        cap = (self.sbuf.width * self.sbuf.height * M + stride - 1) // stride
        X = np.empty((cap, GATE_FEATURES), np.float32)
        y = np.empty(cap, np.float32)
        pv = np.empty((cap, 4), np.float32)
        pl = np.empty((cap, 4), np.float32)
        qi = np.empty(cap, np.uint32)
        n = C.c_int64(0)
        self._chk(lib().tlr_gate_collect(self.h, C.byref(p), stride, cap, _p(X), _p(y), _p(pv), _p(pl), _p(qi),
                                         C.byref(n)))
        k = min(n.value, cap)
        return X[:k], y[:k], pv[:k], pl[:k], qi[:k]

    def gate_append_training(self, out_spp=16, stride=1, **kw):
        p = self._params(out_spp=out_spp, **kw)
        self._chk(lib().tlr_gate_append_training(self.h, C.byref(p), stride))
        return lib().tlr_gate_training_size(self.h)

    def gate_clear_training(self):
        self._chk(lib().tlr_gate_clear_training(self.h))

    def gate_train(self, epochs, batch=4096, lr=1e-3, pos_weight=1.0, seed=0):
        losses = np.zeros(max(epochs, 1), np.float32)
        self._chk(lib().tlr_gate_train(self.h, epochs, batch, lr, pos_weight, C.c_uint64(seed), _p(losses)))
        return losses[:epochs]

    def build_info(self):
        bi = BuildInfo()
        self._chk(lib().tlr_get_build_info(self.h, C.byref(bi)))
        return bi

    def reprojected(self):
        n = self.sbuf.records.shape[0]
        out = [np.empty(n, np.float32) for _ in range(4)]
        self._chk(lib().tlr_copy_reprojected(self.h, *[_p(a) for a in out]))
        return out

    def layout(self):
        out = np.empty(max(self.build_info().num_bucketed, 1), np.int32)
        self._chk(lib().tlr_copy_layout(self.h, _p(out)))
        return out[:self.build_info().num_bucketed]

    def initial_leaves(self):
        n = self.build_info().num_initial_leaves
        out = np.empty((max(n, 1), 6), np.int32)
        self._chk(lib().tlr_copy_initial_leaves(self.h, _p(out)))
        return out[:n]

    def leaves(self):
        n = self.build_info().num_leaves
        s01 = np.empty((max(n, 1), 2), np.int32)
        b = np.empty((max(n, 1), 14), np.float32)
        self._chk(lib().tlr_copy_leaves(self.h, _p(s01), _p(b)))
        return s01[:n], b[:n]

    def hierarchy(self):
        """(leaf order, internal children (n-1, 2), internal bounds (n-1, 14));
        leaf children carry bit 31."""
        n = self.build_info().num_valid_leaves
        order = np.empty(max(n, 1), np.int32)
        ch = np.empty((max(n - 1, 1), 2), np.int32)
        b = np.empty((max(n - 1, 1), 14), np.float32)
        self._chk(lib().tlr_copy_hierarchy(self.h, _p(order), _p(ch), _p(b)))
        return order[:n], ch[:max(n - 1, 0)], b[:max(n - 1, 0)]

    def query_colors(self, out_spp):
        out = np.empty((self.sbuf.height, self.sbuf.width, out_spp, 4), np.float32)
        self._chk(lib().tlr_copy_query_colors(self.h, _p(out)))
        return out

    def query_stages(self, out_spp):
        """After reconstruct(path=2): per output sample, the stage that gave
        its colour (0: 1R first surface, 1: 2R surfaces, 2: nearest)."""
        out = np.empty((self.sbuf.height, self.sbuf.width, out_spp), np.int32)
        self._chk(lib().tlr_copy_query_stages(self.h, _p(out)))
        return out


class TreeGather:
    """Mirror of FW::TreeGather (Reconstruction.hpp:42-48): the constructor
    builds the hierarchy, reconstruct_dof_motion fills an RGBA image."""

    def __init__(self, sbuf, params, aperture_adjust=1.0, focal_adjust=1.0, device=0):
        self.sbuf = sbuf
        self.params = params
        self.ctx = Ctx(device)
        self.ctx.set_samples(sbuf)
        self.ctx.build(aperture_adjust, focal_adjust, coc_override=params.refocus_coc())

    def reconstruct_dof_motion(self, image=None, debug_image=None):
        """Fills image (or returns a new one); debug_image, if given, receives
        (#surfaces, #samples within 1R, 0, 0) per pixel (Reconstruction.hpp:46)."""
        if image is not None and (image.shape[0] < self.sbuf.height or image.shape[1] < self.sbuf.width):
            raise TlrError(5, "TreeGather::reconstructDofMotion image smaller than sample buffer")
        p = self.params
        img, st = self.ctx.reconstruct(out_spp=p.out_spp, override_uvt=p.override_uvt,
                                       lens_gaussian=p.lens_gaussian, time_gaussian=p.time_gaussian,
                                       debug_image=debug_image is not None)
        if debug_image is not None:
            debug_image[:self.sbuf.height, :self.sbuf.width] = self.ctx.debug_image()
        if image is not None:
            image[:self.sbuf.height, :self.sbuf.width] = img
            return image
        return img
