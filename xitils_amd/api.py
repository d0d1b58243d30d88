"""ctypes binding of include/xitils_amd.h (the C ABI of libxitils_amd.so).

This is the Python face of the drop-in: it drives the same entry points a
C++ caller (include/xitils/xitils.hpp) or any other FFI would bind.  There is
no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import scenes as _sc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("XITILS_AMD_LIB") or os.path.join(HERE, "_lib", "libxitils_amd.so")

XT_OK = 0
STATUS_NAMES = {0: "XT_OK", 1: "XT_ERR_INVALID_ARGUMENT", 2: "XT_ERR_UNSUPPORTED_MATERIAL", 3: "XT_ERR_OUT_OF_MEMORY",
                4: "XT_ERR_DEVICE", 5: "XT_ERR_NO_DEVICE", 6: "XT_ERR_TOO_LARGE"}

EXPORTED = ["xt_version", "xt_last_error", "xt_device_open", "xt_device_close", "xt_device_valu_rate",
            "xt_scene_create", "xt_scene_destroy",
            "xt_scene_info", "xt_trace_closest", "xt_trace_any", "xt_render", "xt_render_get_stats",
            "xt_scene_set_counting", "xt_guide_config_default", "xt_guide_init", "xt_guide_get_params",
            "xt_guide_set_params", "xt_guide_get_bounds", "xt_guide_forward", "xt_guide_gradient",
            "xt_guide_train_step", "xt_vmf_eval", "xt_vmf_sample", "xt_vmf_fit", "xt_render_batch",
            "xt_scene_autotune", "xt_render_static"]

XT_FLAG_UNIFORM_BARY = 1
XT_FLAG_ENV_NEE = 2
XT_FLAG_GUIDE = 4
XT_FLAG_GUIDE_TRAIN = 8
XT_FLAG_SORT_MATERIALS = 16
GUIDE_NUM_PARAMS = 17648


class XtError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class xt_material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("p", C.c_float * 7)]


class xt_object_desc(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("normals", C.c_void_p), ("indices", C.c_void_p),
                ("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32),
                ("object_to_world", C.c_void_p), ("world_to_object", C.c_void_p),
                ("material", C.c_uint32), ("tri_materials", C.c_void_p),
                ("texcoords", C.c_void_p), ("tangents", C.c_void_p), ("bitangents", C.c_void_p)]


class xt_texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.c_void_p)]


class xt_env(C.Structure):
    _fields_ = [("has_sky", C.c_int32), ("sky", C.c_float * 3), ("sun_dir", C.c_float * 3),
                ("sun_one_minus_cos", C.c_float), ("sun_sin2", C.c_float), ("sun_radiance", C.c_float * 3),
                ("map_kind", C.c_uint32), ("map_width", C.c_uint32), ("map_height", C.c_uint32),
                ("map_rgb", C.c_void_p)]


class xt_scene_desc(C.Structure):
    _fields_ = [("objects", C.c_void_p), ("num_objects", C.c_uint32), ("materials", C.c_void_p),
                ("num_materials", C.c_uint32), ("env", xt_env), ("textures", C.c_void_p),
                ("num_textures", C.c_uint32), ("material_normal_maps", C.c_void_p)]


class xt_camera(C.Structure):
    _fields_ = [("cam_to_world", C.c_float * 16), ("tan_x", C.c_float), ("tan_y", C.c_float)]


class xt_render_params(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("spp_offset", C.c_uint32),
                ("max_depth", C.c_int32), ("seed", C.c_uint32), ("stream", C.c_uint32), ("integrator", C.c_uint32),
                ("flags", C.c_uint32), ("rr_start", C.c_int32), ("rr_prob", C.c_float), ("ray_eps", C.c_float),
                ("shadow_margin", C.c_float)]


class xt_render_stats(C.Structure):
    _fields_ = [("camera_rays", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("ms_total", C.c_double), ("ms_extend", C.c_double), ("ms_shade", C.c_double),
                ("ms_shadow", C.c_double), ("ms_other", C.c_double), ("extend_launches", C.c_uint32),
                ("ext_node_fetches", C.c_uint64), ("ext_tri_fetches", C.c_uint64),
                ("shadow_node_fetches", C.c_uint64), ("shadow_tri_fetches", C.c_uint64),
                ("ms_guide_infer", C.c_double), ("ms_guide_train", C.c_double), ("guide_steps", C.c_uint64),
                ("guide_samples", C.c_uint64), ("guide_queries", C.c_uint64), ("guide_grad_samples", C.c_uint64)]


class xt_batch_item(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", xt_camera), ("params", xt_render_params), ("film", C.c_void_p),
                ("film_is_f64", C.c_int32), ("film_on_device", C.c_int32), ("ms", C.c_float)]


class xt_guide_config(C.Structure):
    _fields_ = [("guide_prob", C.c_float), ("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("adam_eps", C.c_float), ("minibatch", C.c_uint32), ("max_steps_per_batch", C.c_uint32),
                ("weight_clip", C.c_float), ("seed", C.c_uint32), ("max_records", C.c_uint32),
                ("train_target", C.c_uint32), ("epochs", C.c_uint32), ("max_guided_vertex", C.c_uint32)]


_lib = None


def lib():
    """Loads libxitils_amd.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libxitils_amd.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.xt_version.restype = C.c_char_p
        L.xt_last_error.restype = C.c_char_p
        L.xt_device_open.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.xt_device_close.argtypes = [C.c_void_p]
        L.xt_device_valu_rate.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.xt_scene_create.argtypes = [C.c_void_p, C.POINTER(xt_scene_desc), C.POINTER(C.c_void_p)]
        L.xt_scene_destroy.argtypes = [C.c_void_p]
        L.xt_scene_info.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.xt_trace_closest.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int]
        L.xt_trace_any.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int]
        L.xt_render.argtypes = [C.c_void_p, C.POINTER(xt_camera), C.POINTER(xt_render_params), C.c_void_p, C.c_int,
                                C.c_int]
        L.xt_render_get_stats.argtypes = [C.c_void_p, C.POINTER(xt_render_stats)]
        L.xt_scene_set_counting.argtypes = [C.c_void_p, C.c_int]
        L.xt_guide_config_default.argtypes = [C.POINTER(xt_guide_config)]
        L.xt_guide_config_default.restype = None
        L.xt_guide_init.argtypes = [C.c_void_p, C.POINTER(xt_guide_config)]
        L.xt_guide_get_params.argtypes = [C.c_void_p, C.c_void_p]
        L.xt_guide_set_params.argtypes = [C.c_void_p, C.c_void_p]
        L.xt_guide_get_bounds.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.xt_guide_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        L.xt_guide_gradient.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float,
                                        C.c_float, C.c_void_p]
        L.xt_guide_train_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]
        L.xt_vmf_eval.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                  C.c_void_p, C.c_int]
        L.xt_vmf_sample.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                    C.c_uint32, C.c_void_p, C.c_int]
        L.xt_vmf_fit.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_int]
        L.xt_render_batch.argtypes = [C.c_void_p, C.c_uint32]
        L.xt_render_static.argtypes = [C.c_void_p, C.POINTER(xt_camera), C.POINTER(xt_render_params), C.c_void_p,
                                       C.c_int]
        L.xt_scene_autotune.argtypes = [C.c_void_p, C.POINTER(xt_camera), C.POINTER(xt_render_params), C.c_void_p]
        for f in ("xt_vmf_eval", "xt_vmf_sample", "xt_vmf_fit", "xt_render_batch", "xt_scene_autotune",
                  "xt_render_static"):
            getattr(L, f).restype = C.c_int
        for f in ("xt_device_open", "xt_device_close", "xt_scene_create", "xt_scene_destroy", "xt_scene_info",
                  "xt_trace_closest", "xt_trace_any", "xt_render", "xt_render_get_stats", "xt_scene_set_counting",
                  "xt_guide_init", "xt_guide_get_params", "xt_guide_set_params", "xt_guide_get_bounds",
                  "xt_guide_forward", "xt_guide_gradient", "xt_guide_train_step"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def check(status):
    if status != XT_OK:
        raise XtError(status, lib().xt_last_error().decode())


class Device:
    def __init__(self, ordinal=0):
        h = C.c_void_p()
        check(lib().xt_device_open(ordinal, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().xt_device_close(self.h)
            self.h = None

    def valu_rate(self):
        """Wave-level VALU instructions per second that a kernel of 16
        independent fp32 FMA chains per lane issues on this device."""
        r = C.c_double()
        check(lib().xt_device_valu_rate(self.h, C.byref(r)))
        return r.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def vmf_eval(device, encoded, dirs, dist_of_dir=None):
    """pdf of each direction under its encoded mixture (VonMisesFisherDistribution::eval)."""
    enc = np.ascontiguousarray(encoded, np.float32)
    enc2 = enc.reshape(-1, enc.shape[-1])
    nl = (enc2.shape[1] + 1) // 5
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    idx = None if dist_of_dir is None else np.ascontiguousarray(dist_of_dir, np.uint32)
    out = np.zeros(d.shape[0], np.float32)
    check(lib().xt_vmf_eval(device.h, nl, enc2.ctypes.data, enc2.shape[0], None if idx is None else idx.ctypes.data,
                            d.ctypes.data, d.shape[0], out.ctypes.data, 0))
    return out


def vmf_sample(device, encoded, uniforms, dist_of_sample=None):
    """Directions from the reference's vMF sampler, 4 uniforms per sample."""
    enc = np.ascontiguousarray(encoded, np.float32)
    enc2 = enc.reshape(-1, enc.shape[-1])
    nl = (enc2.shape[1] + 1) // 5
    u = np.ascontiguousarray(uniforms, np.float32).reshape(-1, 4)
    idx = None if dist_of_sample is None else np.ascontiguousarray(dist_of_sample, np.uint32)
    out = np.zeros((u.shape[0], 3), np.float32)
    check(lib().xt_vmf_sample(device.h, nl, enc2.ctypes.data, enc2.shape[0], None if idx is None else idx.ctypes.data,
                              u.ctypes.data, u.shape[0], out.ctypes.data, 0))
    return out


def vmf_fit(device, num_lobes, samples, offsets, init_uniforms):
    """approximateBySEM per sample set; returns (encoded (F, 5N-1), steps (F,))."""
    s = np.ascontiguousarray(samples, np.float32).reshape(-1, 3)
    off = np.ascontiguousarray(offsets, np.uint32)
    nf = off.shape[0] - 1
    iu = np.ascontiguousarray(init_uniforms, np.float32).reshape(nf, 2 * num_lobes)
    enc = np.zeros((nf, 5 * num_lobes - 1), np.float32)
    steps = np.zeros(nf, np.int32)
    check(lib().xt_vmf_fit(device.h, num_lobes, s.ctypes.data, s.shape[0], off.ctypes.data, nf, iu.ctypes.data,
                           enc.ctypes.data, steps.ctypes.data, 0))
    return enc, steps


def render_batch(jobs):
    """jobs: list of (scene, width, height, spp, film or None, kwargs) -> the
    films (host numpy, accumulated like Scene.render) and per-item GPU ms.
    Films may also be (device pointer, f64) tuples."""
    items = (xt_batch_item * max(1, len(jobs)))()
    films = []
    for k, (sc, w, h, spp, film, kw) in enumerate(jobs):
        kw = dict(kw)
        cam = kw.pop("camera", None)
        items[k].scene = sc.h
        items[k].camera = camera_struct(cam or sc.desc.camera)
        items[k].params = Scene.params(w, h, spp, **kw)
        if isinstance(film, tuple):
            items[k].film = film[0]
            items[k].film_is_f64 = int(film[1])
            items[k].film_on_device = 1
            films.append(None)
        else:
            if film is None:
                film = np.zeros((h, w, 3), np.float32)
            items[k].film = film.ctypes.data
            items[k].film_is_f64 = int(film.dtype == np.float64)
            items[k].film_on_device = 0
            films.append(film)
    check(lib().xt_render_batch(items, len(jobs)))
    return films, [items[k].ms for k in range(len(jobs))]


def env_struct(e: "_sc.Env", keep=None) -> xt_env:
    omc, sin2 = e.sun_terms()
    s = xt_env()
    if e.map_kind and e.map is not None:
        m = np.ascontiguousarray(e.map, np.float32)
        if keep is not None:
            keep.append(m)
        s.map_kind = int(e.map_kind)
        s.map_height, s.map_width = int(m.shape[0]), int(m.shape[1])
        s.map_rgb = m.ctypes.data
    s.has_sky = int(e.has_sky)
    s.sky[:] = [float(x) for x in e.sky]
    s.sun_dir[:] = [float(x) for x in e.sun_dir]
    s.sun_one_minus_cos = float(omc)
    s.sun_sin2 = float(sin2)
    s.sun_radiance[:] = [float(x) for x in e.sun_radiance]
    return s


def camera_struct(cam: "_sc.Camera") -> xt_camera:
    c = xt_camera()
    packed = cam.packed()
    c.cam_to_world[:] = [float(x) for x in packed[:16]]
    c.tan_x = float(packed[16])
    c.tan_y = float(packed[17])
    return c


class Scene:
    """A scene uploaded to HBM (Scene.h:11-61 after buildAccelerationStructure)."""

    def __init__(self, device: Device, desc: "_sc.SceneDesc"):
        self.device = device
        self.desc = desc
        keep = []
        objs = (xt_object_desc * max(1, len(desc.objects)))()
        for k, o in enumerate(desc.objects):
            pos = np.ascontiguousarray(o.positions, np.float32)
            idx = np.ascontiguousarray(o.indices, np.int32)
            keep += [pos, idx]
            objs[k].positions = pos.ctypes.data
            objs[k].indices = idx.ctypes.data
            objs[k].num_vertices = pos.shape[0]
            objs[k].num_triangles = idx.shape[0]
            if o.normals is not None:
                n = np.ascontiguousarray(o.normals, np.float32)
                keep.append(n)
                objs[k].normals = n.ctypes.data
            if o.xf is not None:
                m = np.ascontiguousarray(o.xf, np.float32)
                mi = np.ascontiguousarray(o.xf_inv, np.float32)
                keep += [m, mi]
                objs[k].object_to_world = m.ctypes.data
                objs[k].world_to_object = mi.ctypes.data
            objs[k].material = o.material
            if o.tri_materials is not None:
                tm = np.ascontiguousarray(o.tri_materials, np.uint32)
                keep.append(tm)
                objs[k].tri_materials = tm.ctypes.data
            for name in ("texcoords", "tangents", "bitangents"):
                a = getattr(o, name, None)
                if a is not None:
                    a = np.ascontiguousarray(a, np.float32)
                    keep.append(a)
                    setattr(objs[k], name, a.ctypes.data)
        mats = _sc.pack_materials(desc.materials)
        keep.append(mats)
        sd = xt_scene_desc()
        sd.objects = C.cast(objs, C.c_void_p)
        sd.num_objects = len(desc.objects)
        sd.materials = mats.ctypes.data
        sd.num_materials = mats.shape[0]
        sd.env = env_struct(desc.env, keep)
        texs = getattr(desc, "textures", None) or []
        if texs:
            tarr = (xt_texture * len(texs))()
            for k, t in enumerate(texs):
                t = np.ascontiguousarray(t, np.float32)
                keep.append(t)
                tarr[k].width, tarr[k].height, tarr[k].rgb = t.shape[1], t.shape[0], t.ctypes.data
            keep.append(tarr)
            sd.textures = C.cast(tarr, C.c_void_p)
            sd.num_textures = len(texs)
        nmaps = _sc.material_normal_maps(desc.materials)
        if nmaps is not None:
            keep.append(nmaps)
            sd.material_normal_maps = nmaps.ctypes.data
        h = C.c_void_p()
        check(lib().xt_scene_create(device.h, C.byref(sd), C.byref(h)))
        self.h = h
        del keep

    def destroy(self):
        if getattr(self, "h", None):
            lib().xt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def info(self):
        out = (C.c_uint64 * 6)()
        check(lib().xt_scene_info(self.h, out))
        build_ms = np.array([out[5]], np.uint64).view(np.float64)[0]
        return {"num_prims": out[0], "bvh_nodes": out[1], "bvh_depth": out[2], "bytes_nodes": out[3],
                "bytes_tris": out[4], "build_ms": float(build_ms)}

    # ---- ray queries -----------------------------------------------------
    def trace_closest(self, rays, on_device=False):
        """rays: (n, 8) float32 {o, tmax, d, pad} (host array or device pointer
        with on_device and an explicit n).  Returns (n, 4) float32 {t,
        prim bits, b0, b1}."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), np.float32)
        check(lib().xt_trace_closest(self.h, rays.ctypes.data, rays.shape[0], hits.ctypes.data, 0))
        return hits

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        occ = np.zeros(rays.shape[0], np.uint8)
        check(lib().xt_trace_any(self.h, rays.ctypes.data, rays.shape[0], occ.ctypes.data, 0))
        return occ

    def trace_closest_dev(self, rays_ptr, n, hits_ptr):
        check(lib().xt_trace_closest(self.h, C.c_void_p(rays_ptr), n, C.c_void_p(hits_ptr), 1))

    def trace_any_dev(self, rays_ptr, n, occ_ptr):
        check(lib().xt_trace_any(self.h, C.c_void_p(rays_ptr), n, C.c_void_p(occ_ptr), 1))

    # ---- rendering -------------------------------------------------------
    @staticmethod
    def params(width, height, spp, seed=1, max_depth=0, integrator=0, flags=0, spp_offset=0, stream=0,
               rr_start=5, rr_prob=0.9, eps=1e-5, shadow_margin=1e-4):
        p = xt_render_params()
        p.width, p.height, p.spp, p.spp_offset = width, height, spp, spp_offset
        p.max_depth, p.seed, p.stream, p.integrator, p.flags = max_depth, seed, stream, integrator, flags
        p.rr_start, p.rr_prob, p.ray_eps, p.shadow_margin = rr_start, rr_prob, eps, shadow_margin
        return p

    def render(self, width, height, spp, film=None, f64=False, camera=None, **kw):
        """Host-film render: returns (height, width, 3) float32/float64 film +=
        radiance sums."""
        cam = camera_struct(camera or self.desc.camera)
        p = self.params(width, height, spp, **kw)
        if film is None:
            film = np.zeros((height, width, 3), np.float64 if f64 else np.float32)
        check(lib().xt_render(self.h, C.byref(cam), C.byref(p), film.ctypes.data, int(film.dtype == np.float64), 0))
        return film

    def render_static(self, film_ptr, width, height, spp, f64=False, camera=None, **kw):
        """xt_render_static: renders into a device film with no device
        allocation (the library's static scratch; needs max_depth >= 1)."""
        cam = camera_struct(camera or self.desc.camera)
        p = self.params(width, height, spp, **kw)
        check(lib().xt_render_static(self.h, C.byref(cam), C.byref(p), C.c_void_p(film_ptr), int(f64)))

    def render_device(self, film_ptr, width, height, spp, f64=False, camera=None, **kw):
        """Renders into a device film (e.g. a torch tensor's data_ptr())."""
        cam = camera_struct(camera or self.desc.camera)
        p = self.params(width, height, spp, **kw)
        check(lib().xt_render(self.h, C.byref(cam), C.byref(p), C.c_void_p(film_ptr), int(f64), 1))

    def autotune(self, width, height, spp, camera=None, **kw):
        """Picks this scene's traversal / shade grid sizes for graph renders;
        returns (traversal blocks per CU, shade blocks per CU)."""
        cam = camera_struct(camera or self.desc.camera)
        p = self.params(width, height, spp, **kw)
        out = (C.c_uint32 * 2)()
        check(lib().xt_scene_autotune(self.h, C.byref(cam), C.byref(p), out))
        return int(out[0]), int(out[1])

    def set_counting(self, on):
        check(lib().xt_scene_set_counting(self.h, int(on)))

    def stats(self):
        s = xt_render_stats()
        check(lib().xt_render_get_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in xt_render_stats._fields_}

    # ---- learned guiding -------------------------------------------------
    @staticmethod
    def guide_config(**kw):
        """xt_guide_config with the library defaults, fields overridden by kw."""
        c = xt_guide_config()
        lib().xt_guide_config_default(C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    def guide_init(self, **kw):
        check(lib().xt_guide_init(self.h, C.byref(self.guide_config(**kw))))

    def guide_params(self):
        out = np.zeros(GUIDE_NUM_PARAMS, np.float32)
        check(lib().xt_guide_get_params(self.h, out.ctypes.data))
        return out

    def guide_set_params(self, params):
        p = np.ascontiguousarray(params, np.float32).reshape(GUIDE_NUM_PARAMS)
        check(lib().xt_guide_set_params(self.h, p.ctypes.data))

    def guide_bounds(self):
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        check(lib().xt_guide_get_bounds(self.h, lo.ctypes.data, hi.ctypes.data))
        return lo, hi

    def guide_forward(self, queries):
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 9)
        out = np.zeros((q.shape[0], 40), np.float32)
        check(lib().xt_guide_forward(self.h, q.ctypes.data, q.shape[0], out.ctypes.data))
        return out

    def guide_gradient(self, queries, dirs, weights, weight_scale=1.0, weight_clip=float("inf")):
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 9)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        w = np.ascontiguousarray(weights, np.float32).reshape(-1)
        g = np.zeros(GUIDE_NUM_PARAMS, np.float32)
        check(lib().xt_guide_gradient(self.h, q.ctypes.data, d.ctypes.data, w.ctypes.data, q.shape[0], weight_scale,
                                      weight_clip, g.ctypes.data))
        return g

    def guide_train_step(self, queries, dirs, weights):
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 9)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        w = np.ascontiguousarray(weights, np.float32).reshape(-1)
        check(lib().xt_guide_train_step(self.h, q.ctypes.data, d.ctypes.data, w.ctypes.data, q.shape[0]))
