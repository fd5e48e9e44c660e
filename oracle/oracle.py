"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TLR_ORACLE_LIB selects another build of the same source (the pin tests'
# mutants, tests/test_oracle_mutants.py)
LIB_PATH = os.environ.get("TLR_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

_lib = None


class BuildInfo(C.Structure):
    _fields_ = [
        ("num_input", C.c_int), ("num_kept", C.c_int), ("num_bucketed", C.c_int),
        ("spp_pow2", C.c_int), ("coc_adj", C.c_float * 2),
        ("bbmin", C.c_float * 2), ("bbmax", C.c_float * 2), ("scale", C.c_float),
        ("bbmin_int", C.c_int * 2), ("grid_w", C.c_int), ("grid_h", C.c_int),
        ("max_hits", C.c_int), ("num_initial_nodes", C.c_int),
        ("num_initial_leaves", C.c_int), ("num_leaves", C.c_int),
        ("num_internal", C.c_int), ("root", C.c_int),
    ]


class ReconParams(C.Structure):
    _fields_ = [
        ("out_spp", C.c_int), ("override_uvt", C.c_int), ("uvt", C.c_float * 3),
        ("lens_gaussian", C.c_int), ("time_gaussian", C.c_int),
        ("threads", C.c_int), ("row_stride", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "leaf_nodes", "samples_in_leaf_nodes", "samples_within_1r", "samples_fetched",
        "surfaces", "num_2r", "surfaces_merged", "at_least_one", "count")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ShadowParams(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("camera_coc", C.c_float * 2),
                ("light_coc", C.c_float * 2), ("camera_far", C.c_float)] + [
        (n, C.c_float * 16) for n in ("c2l", "zfromw_ndc", "inv_projection", "shadow2_inv_proj",
                                      "shadow2_inv_t", "world_inv_t")]


class QueryInfo(C.Structure):
    _fields_ = [
        ("fetch2r", C.c_int), ("num_surfaces_1r", C.c_int), ("num_samples_1r", C.c_int),
        ("num_samples_final", C.c_int), ("emergency_passes", C.c_int),
        ("found_surface", C.c_int), ("used_nearest", C.c_int),
        ("dispersion", C.c_float), ("color", C.c_float * 4),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def refgpu_triangles(xy, disp):
    """checkTriangles of the GPU-variant method on points relative to the
    output sample."""
    p = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    return bool(lib().orc_refgpu_triangles(_ptr(p), p.shape[0], float(disp)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, fp]
        L.orc_destroy.argtypes = [P]
        L.orc_error.restype = C.c_char_p
        L.orc_error.argtypes = [P]
        L.orc_build.argtypes = [P, C.c_float, C.c_float, C.c_int]
        L.orc_get_build_info.argtypes = [P, C.POINTER(BuildInfo)]
        L.orc_get_reprojected.argtypes = [P, fp, ip]
        L.orc_get_cells.argtypes = [P, ip]
        L.orc_get_initial_leaves.argtypes = [P, ip]
        L.orc_get_layout.argtypes = [P, ip]
        L.orc_get_leaves.argtypes = [P, ip, fp]
        L.orc_get_hierarchy.argtypes = [P, ip, ip, fp]
        L.orc_output_pattern.argtypes = [C.c_int, C.c_int, fp, C.c_void_p]
        L.orc_reconstruct.argtypes = [P, C.POINTER(ReconParams), C.c_void_p, C.c_void_p,
                                      C.POINTER(Stats)]
        L.orc_reconstruct_debug.argtypes = [P, C.POINTER(ReconParams), C.c_void_p, C.c_void_p,
                                            C.POINTER(Stats), C.c_void_p]
        L.orc_set_coc_override.argtypes = [P, C.c_void_p]
        L.orc_reconstruct_refgpu.argtypes = [P, C.POINTER(ReconParams), C.c_void_p, C.c_void_p,
                                             C.c_void_p]
        L.orc_refgpu_query.argtypes = [P, fp, C.c_void_p, C.POINTER(C.c_int32)]
        L.orc_refgpu_nearest.argtypes = [P, fp, C.c_void_p]
        L.orc_refgpu_failed.argtypes = [P, fp, C.c_void_p]
        L.orc_refgpu_triangles.argtypes = [C.c_void_p, C.c_int, C.c_float]
        L.orc_refgpu_reproject.argtypes = [P, fp, C.c_void_p]
        L.orc_gather.argtypes = [P, fp, C.c_float, C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                 C.c_void_p, C.c_void_p, C.c_int]
        L.orc_query.argtypes = [P, C.POINTER(ReconParams), C.c_int, C.c_int, C.c_int,
                                C.POINTER(QueryInfo)]
        L.orc_leaves_bruteforce.argtypes = [P, fp, C.c_float, C.c_void_p, C.c_int]
        L.orc_hash_bits.restype = C.c_uint32
        L.orc_hash_bits.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_ranrot.argtypes = [C.c_uint32, C.c_int, up]
        L.orc_sobol.restype = C.c_float
        L.orc_sobol.argtypes = [C.c_int, C.c_int]
        L.orc_sobol2d.argtypes = [C.c_int, fp]
        L.orc_halton.restype = C.c_float
        L.orc_halton.argtypes = [C.c_int, C.c_int]
        L.orc_lp.restype = C.c_float
        L.orc_lp.argtypes = [C.c_int]
        L.orc_to_unit_disk.argtypes = [C.c_float, C.c_float, fp]
        L.orc_tux_planes.argtypes = [fp, fp, C.c_float, C.c_float, fp, fp]
        L.orc_merge_bounds.argtypes = [fp, fp, fp]
        L.orc_fw_sort.argtypes = [C.c_int, fp, ip]
        L.orc_triangle_covers.argtypes = [fp, fp, fp, fp]
        L.orc_same_surface.argtypes = [fp, fp, fp]
        L.orc_surface_scan.argtypes = [C.c_int, fp, fp, fp, fp, fp, ip]
        L.orc_set_density.argtypes = [P, C.c_void_p]
        L.orc_create_irregular.restype = P
        L.orc_create_irregular.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, fp, C.c_int]
        L.orc_reconstruct_queries.argtypes = [P, C.POINTER(ReconParams), fp, C.c_int, C.c_int, C.c_void_p,
                                              C.c_void_p, C.POINTER(Stats)]
        L.orc_shadow_setup.argtypes = [fp, fp, fp, fp, C.c_int, C.c_int, fp, fp, C.c_float,
                                       C.POINTER(ShadowParams)]
        L.orc_mat4_inverse.argtypes = [fp, fp]
        L.orc_light_space_dot.restype = C.c_float
        L.orc_light_space_dot.argtypes = [fp, fp, C.POINTER(ShadowParams), fp]
        L.orc_reconstruct_dof_motion_shadows.argtypes = [P, P, C.POINTER(ReconParams), C.POINTER(ShadowParams),
                                                         C.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One reconstruction context: records -> build -> reconstruct."""

    def __init__(self, W, H, spp, C0, C1, records):
        L = lib()
        rec = np.ascontiguousarray(records, dtype=np.float32).reshape(-1)
        assert rec.size == W * H * spp * 16
        self.W, self.H, self.spp = W, H, spp
        self.h = L.orc_create(W, H, spp, C0, C1, rec)

    @classmethod
    def irregular(cls, W, H, C0, C1, records):
        """An irregular buffer (spp 0): n records in any order."""
        self = cls.__new__(cls)
        rec = np.ascontiguousarray(records, dtype=np.float32).reshape(-1)
        self.W, self.H, self.spp = W, H, 0
        self.n = rec.size // 16
        self.h = lib().orc_create_irregular(W, H, C0, C1, rec, self.n)
        return self

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self, aperture_adjust=1.0, focal_adjust=1.0, threads=0, coc_override=None):
        cc = None if coc_override is None else np.ascontiguousarray(coc_override, np.float32)
        lib().orc_set_coc_override(self.h, _ptr(cc))
        rc = lib().orc_build(self.h, aperture_adjust, focal_adjust, threads)
        if rc != 0:
            raise RuntimeError("oracle build failed (%d): %s" % (rc, lib().orc_error(self.h).decode()))
        return self.info()

    def info(self):
        bi = BuildInfo()
        lib().orc_get_build_info(self.h, C.byref(bi))
        return bi

    def reprojected(self):
        bi = self.info()
        out = np.zeros((bi.num_kept, 10), np.float32)
        src = np.zeros(bi.num_kept, np.int32)
        lib().orc_get_reprojected(self.h, out, src)
        return out, src

    def cells(self):
        out = np.zeros(self.info().num_kept, np.int32)
        lib().orc_get_cells(self.h, out)
        return out

    def initial_leaves(self):
        out = np.zeros((self.info().num_initial_leaves, 6), np.int32)
        lib().orc_get_initial_leaves(self.h, out)
        return out

    def layout(self):
        out = np.zeros(self.info().num_bucketed, np.int32)
        lib().orc_get_layout(self.h, out)
        return out

    def leaves(self):
        n = self.info().num_leaves
        s01 = np.zeros((n, 2), np.int32)
        b = np.zeros((n, 14), np.float32)
        lib().orc_get_leaves(self.h, s01, b)
        return s01, b

    def hierarchy(self):
        bi = self.info()
        nvalid = bi.num_internal + 1 if bi.root != -1 else 0
        order = np.zeros(max(nvalid, 1), np.int32)
        ch = np.zeros((max(bi.num_internal, 1), 2), np.int32)
        b = np.zeros((max(bi.num_internal, 1), 14), np.float32)
        lib().orc_get_hierarchy(self.h, order, ch, b)
        return order[:nvalid], ch[:bi.num_internal], b[:bi.num_internal]

    @staticmethod
    def params(out_spp=16, override_uvt=None, lens_gaussian=0, time_gaussian=0, threads=0,
               row_stride=1):
        p = ReconParams()
        p.out_spp = out_spp
        p.override_uvt = 1 if override_uvt is not None else 0
        if override_uvt is not None:
            for k in range(3):
                p.uvt[k] = override_uvt[k]
        p.lens_gaussian = lens_gaussian
        p.time_gaussian = time_gaussian
        p.threads = threads
        p.row_stride = row_stride
        return p

    def reconstruct(self, out_spp=16, want_queries=False, want_debug=False, **kw):
        """(image, per-output-sample colours or None, stats[, debug image])."""
        p = self.params(out_spp=out_spp, **kw)
        M = 1 if p.override_uvt else out_spp
        img = np.zeros((self.H, self.W, 4), np.float32)
        q = np.zeros((self.H, self.W, M, 4), np.float32) if want_queries else None
        dbg = np.zeros((self.H, self.W, 4), np.float32) if want_debug else None
        st = Stats()
        rc = lib().orc_reconstruct_debug(self.h, C.byref(p), _ptr(img), _ptr(q), C.byref(st), _ptr(dbg))
        if rc != 0:
            raise RuntimeError("oracle reconstruct failed (%d)" % rc)
        if want_debug:
            return img, q, st.as_dict(), dbg
        return img, q, st.as_dict()

    def set_density(self, density):
        """Per-sample LIGHTSPACE_DENSITY (input order) or None for 1; rebuild after."""
        d = None if density is None else np.ascontiguousarray(density, np.float32).reshape(-1)
        assert d is None or d.size == (self.W * self.H * self.spp if self.spp else self.n)
        lib().orc_set_density(self.h, _ptr(d))

    def reconstruct_queries(self, queries, shadow=False, **kw):
        """Explicit output samples (n x 8: x, y, u, v, t, w, density, 0):
        (colours n x 4, extended results n x 8, stats)."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 8)
        p = self.params(**kw)
        rgba = np.zeros((q.shape[0], 4), np.float32)
        ext = np.zeros((q.shape[0], 8), np.float32)
        st = Stats()
        rc = lib().orc_reconstruct_queries(self.h, C.byref(p), q, q.shape[0], 1 if shadow else 0, _ptr(rgba),
                                           _ptr(ext), C.byref(st))
        if rc != 0:
            raise RuntimeError("oracle queries failed (%d)" % rc)
        return rgba, ext, st.as_dict()

    def reconstruct_dof_motion_shadows(self, light, sp, out_spp=16, **kw):
        """FilterTask::process2 over the image with `light` (another built
        Oracle of the same size) as the shadow TreeGather."""
        p = self.params(out_spp=out_spp, **kw)
        img = np.zeros((self.H, self.W, 4), np.float32)
        rc = lib().orc_reconstruct_dof_motion_shadows(self.h, light.h, C.byref(p), C.byref(sp), _ptr(img))
        if rc != 0:
            raise RuntimeError("oracle shadows failed (%d)" % rc)
        return img

    def reconstruct_refgpu(self, out_spp=16, want_queries=False, **kw):
        """The reference's GPU-variant method (recKernel / failKernel /
        nearestKernel): image, per-output-sample colours and stages (0: 1R,
        1: 2R, 2: nearest) when want_queries."""
        p = self.params(out_spp=out_spp, **kw)
        M = 1 if p.override_uvt else out_spp
        img = np.zeros((self.H, self.W, 4), np.float32)
        q = np.zeros((self.H, self.W, M, 4), np.float32) if want_queries else None
        stg = np.zeros((self.H, self.W, M), np.int32) if want_queries else None
        rc = lib().orc_reconstruct_refgpu(self.h, C.byref(p), _ptr(img), _ptr(q), _ptr(stg))
        if rc != 0:
            raise RuntimeError("oracle reconstruct_refgpu failed (%d)" % rc)
        return img, q, stg

    def refgpu_query(self, o5):
        """One output sample through the GPU-variant method: (colour, stage)."""
        o = np.asarray(o5, np.float32)
        col = np.zeros(4, np.float32)
        st = C.c_int32(0)
        if lib().orc_refgpu_query(self.h, o, _ptr(col), C.byref(st)) != 0:
            raise RuntimeError("not built")
        return col, st.value

    def refgpu_nearest(self, o5):
        """nearestKernel alone: (tree-layout index or -1, colour)."""
        o = np.asarray(o5, np.float32)
        col = np.zeros(4, np.float32)
        k = lib().orc_refgpu_nearest(self.h, o, _ptr(col))
        return k, col

    def refgpu_failed(self, o5):
        """failKernel alone (2R surfaces): (found, colour)."""
        o = np.asarray(o5, np.float32)
        col = np.zeros(4, np.float32)
        return lib().orc_refgpu_failed(self.h, o, _ptr(col)) == 1, col

    def refgpu_reproject(self, o5):
        """dx, dy, dist^2 of every tree sample (layout order) at o5."""
        o = np.asarray(o5, np.float32)
        out = np.zeros((self.info().num_bucketed, 3), np.float32)
        lib().orc_refgpu_reproject(self.h, o, _ptr(out))
        return out

    def gather(self, o5, R, max_leaves=4096, max_samples=1 << 16):
        o = np.asarray(o5, np.float32)
        lv = np.zeros(max_leaves, np.int32)
        sm = np.zeros(max_samples, np.int32)
        sf = np.zeros(max_samples, np.int32)
        nl = C.c_int(0)
        ns = lib().orc_gather(self.h, o, R, _ptr(lv), max_leaves, C.byref(nl), _ptr(sm), _ptr(sf),
                              max_samples)
        return lv[:nl.value], sm[:ns], sf[:ns]

    def query(self, x, y, i, out_spp=16, **kw):
        p = self.params(out_spp=out_spp, **kw)
        qi = QueryInfo()
        lib().orc_query(self.h, C.byref(p), x, y, i, C.byref(qi))
        return qi

    def leaves_bruteforce(self, o5, R, max_leaves=1 << 16):
        o = np.asarray(o5, np.float32)
        lv = np.zeros(max_leaves, np.int32)
        n = lib().orc_leaves_bruteforce(self.h, o, R, _ptr(lv), max_leaves)
        return lv[:n]


def surface_scan(t, ba, bb, ba1, bb1):
    """Leaf surface scan (ReconstructionTreeBuilder.cpp:222-340) of candidates
    with times t and screen boxes ba, bb (t=0) and ba1, bb1 (t=1), (n, 2)
    each: surface id per candidate."""
    t = np.ascontiguousarray(t, np.float32)
    n = t.size
    arrs = [np.ascontiguousarray(a, np.float32).reshape(n, 2) for a in (ba, bb, ba1, bb1)]
    out = np.zeros(n, np.int32)
    lib().orc_surface_scan(n, t, *arrs, out)
    return out


def same_surface(b0, b1, o5):
    """sameSurface (Reconstruction.hpp:457-525) of two 14-float bounds (12
    plane coefficients: xmin, xmax, ymin, ymax planes as (t, uv, 1)
    coefficients; then the time span) around output sample o5 = (x,y,u,v,t)."""
    return bool(lib().orc_same_surface(np.ascontiguousarray(b0, np.float32), np.ascontiguousarray(b1, np.float32),
                                       np.ascontiguousarray(o5, np.float32)))


def output_pattern(out_spp, override_uvt=None):
    uvt = np.zeros(3, np.float32)
    if override_uvt is not None:
        uvt[:] = override_uvt
    M = lib().orc_output_pattern(out_spp, 1 if override_uvt is not None else 0, uvt, None)
    out = np.zeros((64 * M, 6), np.float32)
    lib().orc_output_pattern(out_spp, 1 if override_uvt is not None else 0, uvt, _ptr(out))
    return out


def _m16(m):
    return np.ascontiguousarray(m, np.float32).reshape(16)


def shadow_setup(cam_proj, cam_to_world, light_proj, light_world_to_cam, w, h, cam_coc, light_coc, cam_far):
    """ShadowParams from the cameras' row-major 4x4 matrices."""
    sp = ShadowParams()
    lib().orc_shadow_setup(_m16(cam_proj), _m16(cam_to_world), _m16(light_proj), _m16(light_world_to_cam), w, h,
                           np.asarray(cam_coc, np.float32), np.asarray(light_coc, np.float32), cam_far,
                           C.byref(sp))
    return sp


def mat4_inverse(m):
    out = np.zeros(16, np.float32)
    lib().orc_mat4_inverse(_m16(m), out)
    return out.reshape(4, 4)


def light_space_dot(xyw, wg, sp):
    n = np.zeros(3, np.float32)
    d = lib().orc_light_space_dot(np.asarray(xyw, np.float32), np.asarray(wg, np.float32), C.byref(sp), n)
    return d, n
