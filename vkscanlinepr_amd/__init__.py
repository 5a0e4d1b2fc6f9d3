"""vkscanlinepr_amd -- MI355X-native scanline path rasteriser.

Python mirror of the C ABI in include/spr.h (libspr.so, hand-written gfx950 HIP kernels).  The
classes follow the reference's plugin interface VGRasterizer (VkScanlinePR/src/core/rasterizer.h:
9-20): ``Rasterizer(w, h)`` ~ initialize, ``load_scene``/``load_paths``/``load_rvg`` ~ loadVG,
``set_mvp`` ~ setMVP, ``render`` ~ render.  Errors raise ``SprError`` (the reference throws
std::runtime_error).  There is no CPU fallback: if libspr.so is missing, loading fails.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

__all__ = ["SprError", "Scene", "Rasterizer", "lib", "STAGES", "COMPOSE_OVERWRITE",
           "COMPOSE_SRC_OVER", "MODE_PLAIN", "MODE_LEARNED"]

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libspr.so"

COMPOSE_OVERWRITE, COMPOSE_SRC_OVER = 0, 1
MODE_PLAIN, MODE_LEARNED = 0, 1
STAGES = {
    "tpos": 0, "visible": 1, "cuts": 2, "counts": 3, "inter": 4, "key": 5, "wn": 6, "seg": 7,
    "skey": 8, "sidx": 9, "winding": 10, "flags": 11, "records": 12, "image": 13, "features": 14,
    "pred": 15, "path_counts": 16, "coverage": 17,
}
_STAGE_DTYPE = {
    "tpos": np.float32, "visible": np.int32, "cuts": np.float32, "counts": np.int32,
    "inter": np.int32, "key": np.int32, "wn": np.int32, "seg": np.int32, "skey": np.int32,
    "sidx": np.int32, "winding": np.int32, "flags": np.int32, "records": np.int32,
    "image": np.uint8, "features": np.float32, "pred": np.float32, "path_counts": np.int32,
    "coverage": np.uint64,
}
AA_SAMPLES = 16  # samples per pixel of the anti-aliased mode
MLP_PARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2  # W1, b1, W2, b2, W3, b3
MLP_FLOATS = MLP_PARAMS + 32                          # + feature mean[16], 1/std[16]


class SprError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"spr error {code}: {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [("compose", C.c_uint32), ("mode", C.c_uint32), ("device", C.c_int32),
                ("debug", C.c_uint32), ("aa", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class Stats(C.Structure):
    _fields_ = [("n_points", C.c_int64), ("n_curves", C.c_int64), ("n_paths", C.c_int64),
                ("n_intersections", C.c_int64), ("n_cells", C.c_int64), ("n_spans", C.c_int64),
                ("n_overflow_paths", C.c_int64), ("alloc_intersections", C.c_int64),
                ("frame_ms", C.c_double), ("stage_ms", C.c_double * 16),
                ("learned_fallbacks", C.c_int64), ("n_wave_chains", C.c_int64),
                ("n_heavy_pieces", C.c_int64), ("n_long_bands", C.c_int64), ("n_big_bands", C.c_int64),
                ("graph_captures", C.c_int64)]


class TrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float),
                ("chunk_paths", C.c_uint32), ("seed", C.c_uint64), ("overflow_quantile", C.c_float),
                ("reserved", C.c_uint32 * 3)]


class TuneReport(C.Structure):
    _fields_ = [("chain_blocks_per_cu", C.c_int32), ("merge_blocks_per_cu", C.c_int32),
                ("n_tried", C.c_int32), ("reserved", C.c_int32), ("default_ms", C.c_double),
                ("best_ms", C.c_double), ("scan_shape", C.c_int32), ("band_window_log", C.c_int32)]


class TrainReport(C.Structure):
    _fields_ = [("samples", C.c_int64), ("steps", C.c_int64), ("first_loss", C.c_double),
                ("train_loss", C.c_double), ("val_loss", C.c_double), ("log_margin", C.c_double),
                ("collect_ms", C.c_double), ("train_ms", C.c_double)]


class SceneView(C.Structure):
    _fields_ = [("pos_xy", C.POINTER(C.c_float)), ("pos_path_idx", C.POINTER(C.c_uint32)),
                ("n_points", C.c_uint32),
                ("curve_pos_map", C.POINTER(C.c_uint32)), ("curve_type", C.POINTER(C.c_uint32)),
                ("curve_path_idx", C.POINTER(C.c_uint32)), ("n_curves", C.c_uint32),
                ("fill_rule", C.POINTER(C.c_uint32)), ("fill_info", C.POINTER(C.c_uint32)),
                ("n_paths", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("curve_weight", C.POINTER(C.c_float)), ("path_paint", C.POINTER(C.c_uint32)),
                ("paints", C.c_void_p), ("n_paints", C.c_uint32),
                ("stops", C.c_void_p), ("n_stops", C.c_uint32)]


# spr_paint (64 bytes) and spr_stop (5 floats) as numpy records
PAINT_DTYPE = np.dtype([("kind", "<u4"), ("spread", "<u4"), ("first_stop", "<u4"), ("n_stops", "<u4"),
                        ("opacity", "<f4"), ("xform", "<f4", (6,)), ("geom", "<f4", (5,))])
assert PAINT_DTYPE.itemsize == 64
PAINT_LINEAR, PAINT_RADIAL = 1, 2
SPREAD_PAD, SPREAD_REPEAT, SPREAD_REFLECT, SPREAD_TRANSPARENT = 0, 1, 2, 3
SOLID_PAINT = 0xFFFFFFFF
CURVE_LINE, CURVE_QUAD, CURVE_CUBIC, CURVE_ARC = 2, 3, 4, 0x13


_lib = None
EXPORTS = [
    "spr_create", "spr_destroy", "spr_last_error", "spr_global_error", "spr_load_paths",
    "spr_load_rvg", "spr_load_scene", "spr_set_mvp", "spr_render_frame", "spr_image_device_ptr",
    "spr_synchronize", "spr_set_stage_timing", "spr_get_stats", "spr_debug_export",
    "spr_debug_inject", "spr_scene_load_rvg", "spr_scene_generate", "spr_scene_subset",
    "spr_scene_view_get", "spr_scene_write_rvg", "spr_scene_free", "spr_scene_create", "spr_learned_train",
    "spr_learned_save", "spr_learned_load", "spr_learned_set_margin", "spr_mlp_get_params",
    "spr_mlp_set_params", "spr_mlp_predict", "spr_mlp_gradients", "spr_load_batch", "spr_batch_image",
    "spr_autotune", "spr_set_tuning", "spr_scan_i32", "spr_scan_i32_noalloc", "spr_write_png",
    "spr_write_png_rgba",
]


def lib() -> C.CDLL:
    """Loads libspr.so (raises if it was not built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FileNotFoundError(f"{LIB_PATH} is missing; run `python -m vkscanlinepr_amd.build`")
    L = C.CDLL(str(LIB_PATH))
    vp, u32p, f32p = C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_float)
    L.spr_create.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(Opts), C.POINTER(vp)]
    L.spr_destroy.argtypes = [vp]
    L.spr_last_error.argtypes = [vp]
    L.spr_last_error.restype = C.c_char_p
    L.spr_global_error.restype = C.c_char_p
    L.spr_load_paths.argtypes = [vp, f32p, u32p, C.c_uint32, u32p, u32p, u32p, C.c_uint32, u32p,
                                 u32p, C.c_uint32]
    L.spr_load_rvg.argtypes = [vp, C.c_char_p]
    L.spr_load_scene.argtypes = [vp, vp]
    L.spr_set_mvp.argtypes = [vp, f32p]
    L.spr_render_frame.argtypes = [vp, C.c_void_p]
    L.spr_image_device_ptr.argtypes = [vp, C.POINTER(vp)]
    L.spr_synchronize.argtypes = [vp]
    L.spr_set_stage_timing.argtypes = [vp, C.c_int]
    L.spr_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.spr_debug_export.argtypes = [vp, C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.spr_debug_inject.argtypes = [vp, C.c_int, C.c_void_p, C.c_size_t]
    L.spr_scene_load_rvg.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.spr_scene_generate.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(vp)]
    L.spr_scene_subset.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(vp)]
    L.spr_scene_view_get.argtypes = [vp, C.POINTER(SceneView)]
    L.spr_scene_write_rvg.argtypes = [vp, C.c_char_p]
    L.spr_scene_free.argtypes = [vp]
    L.spr_scene_create.argtypes = [C.POINTER(SceneView), C.POINTER(vp)]
    L.spr_scene_free.restype = None
    L.spr_learned_train.argtypes = [vp, vp, C.POINTER(TrainOpts), C.POINTER(TrainReport)]
    L.spr_learned_save.argtypes = [vp, C.c_char_p]
    L.spr_learned_load.argtypes = [vp, C.c_char_p]
    L.spr_learned_set_margin.argtypes = [vp, C.c_float, C.c_int32]
    L.spr_mlp_get_params.argtypes = [vp, f32p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.spr_mlp_set_params.argtypes = [vp, f32p, C.c_size_t]
    L.spr_mlp_predict.argtypes = [vp, f32p, C.c_uint32, f32p]
    L.spr_mlp_gradients.argtypes = [vp, f32p, f32p, C.c_uint32, f32p, f32p]
    L.spr_load_batch.argtypes = [vp, C.POINTER(vp), u32p, u32p, C.c_uint32]
    L.spr_batch_image.argtypes = [vp, C.c_uint32, C.c_void_p, C.c_size_t]
    L.spr_autotune.argtypes = [vp, C.c_uint32, C.POINTER(TuneReport)]
    L.spr_set_tuning.argtypes = [vp, C.POINTER(TuneReport)]
    L.spr_write_png.argtypes = [vp, C.c_char_p]
    L.spr_write_png_rgba.argtypes = [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]
    L.spr_scan_i32.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
    L.spr_scan_i32_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
    _lib = L
    return L


def _check(rc: int, handle=None) -> None:
    if rc != 0:
        L = lib()
        msg = (L.spr_last_error(handle) if handle else L.spr_global_error()) or b""
        raise SprError(rc, msg.decode(errors="replace"))


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def write_png(path, rgba: np.ndarray) -> None:
    """An (h, w, 4) uint8 image (top row first) to an 8-bit RGBA PNG file."""
    a = np.ascontiguousarray(rgba, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("expected an (h, w, 4) uint8 image")
    _check(lib().spr_write_png_rgba(str(path).encode(), a.ctypes.data_as(C.c_void_p), a.shape[1], a.shape[0]))


def scan_i32_noalloc(d_in: int, d_out: int, n: int, stream: int | None = None) -> None:
    """spr_scan_i32_noalloc: the exclusive scan with a static device workspace (no allocation),
    on the current device and `stream` (a hipStream_t handle; None = the null stream)."""
    _check(lib().spr_scan_i32_noalloc(C.c_void_p(d_in), C.c_void_p(d_out), int(n),
                                      C.c_void_p(stream) if stream else None))


class Scene:
    """A flattened vector scene (the SoA that loadVG uploads) held by the library."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        self._arrays = None

    @classmethod
    def generate(cls, kind: int, seed: int, n_paths: int = 0, canvas: int = 0) -> "Scene":
        out = C.c_void_p()
        _check(lib().spr_scene_generate(kind, seed, n_paths, canvas, C.byref(out)))
        return cls(out.value)

    @classmethod
    def load_rvg(cls, path: str | Path) -> "Scene":
        out = C.c_void_p()
        _check(lib().spr_scene_load_rvg(str(path).encode(), C.byref(out)))
        return cls(out.value)

    def subset(self, first: int, count: int) -> "Scene":
        out = C.c_void_p()
        _check(lib().spr_scene_subset(self._h, first, count, C.byref(out)))
        return Scene(out.value)

    def write_rvg(self, path: str | Path) -> None:
        _check(lib().spr_scene_write_rvg(self._h, str(path).encode()))

    def arrays(self) -> dict:
        """Copies of the SoA arrays (numpy), plus width/height."""
        if self._arrays is None:
            v = SceneView()
            _check(lib().spr_scene_view_get(self._h, C.byref(v)))

            def cp(p, n, dt):
                if n == 0:
                    return np.zeros(0, dtype=dt)
                return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)

            self._arrays = dict(
                pos=cp(v.pos_xy, 2 * v.n_points, np.float32).reshape(-1, 2),
                pos_path=cp(v.pos_path_idx, v.n_points, np.uint32),
                curve_pos=cp(v.curve_pos_map, v.n_curves, np.uint32),
                curve_type=cp(v.curve_type, v.n_curves, np.uint32),
                curve_path=cp(v.curve_path_idx, v.n_curves, np.uint32),
                rule=cp(v.fill_rule, v.n_paths, np.uint32),
                fill=cp(v.fill_info, v.n_paths, np.uint32),
                width=int(v.width), height=int(v.height),
                curve_weight=cp(v.curve_weight, v.n_curves, np.float32) if v.curve_weight else None,
                path_paint=cp(v.path_paint, v.n_paths, np.uint32) if v.path_paint else None,
                paints=(np.frombuffer(C.string_at(v.paints, 64 * v.n_paints), dtype=PAINT_DTYPE).copy()
                        if v.n_paints else np.zeros(0, dtype=PAINT_DTYPE)),
                stops=(np.frombuffer(C.string_at(v.stops, 20 * v.n_stops), dtype=np.float32).reshape(-1, 5).copy()
                       if v.n_stops else np.zeros((0, 5), dtype=np.float32)),
            )
        return self._arrays

    @classmethod
    def create(cls, pos, pos_path, curve_pos, curve_type, curve_path, rule, fill, width: int = 0,
               height: int = 0, curve_weight=None, path_paint=None, paints=None, stops=None) -> "Scene":
        """A scene from arrays (spr_scene_create): copied and validated by the library."""
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        v = SceneView()
        pos = arr(pos, np.float32).reshape(-1)
        pp, cpos, ctype, cpath = (arr(pos_path, np.uint32), arr(curve_pos, np.uint32),
                                  arr(curve_type, np.uint32), arr(curve_path, np.uint32))
        rl, fl = arr(rule, np.uint32), arr(fill, np.uint32)
        u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
        v.pos_xy = pos.ctypes.data_as(C.POINTER(C.c_float))
        v.pos_path_idx, v.n_points = u32(pp), len(pp)
        v.curve_pos_map, v.curve_type, v.curve_path_idx, v.n_curves = u32(cpos), u32(ctype), u32(cpath), len(ctype)
        v.fill_rule, v.fill_info, v.n_paths = u32(rl), u32(fl), len(rl)
        v.width, v.height = int(width), int(height)
        if curve_weight is not None:
            cw = arr(curve_weight, np.float32)
            v.curve_weight = cw.ctypes.data_as(C.POINTER(C.c_float))
        if path_paint is not None:
            v.path_paint = u32(arr(path_paint, np.uint32))
            g = arr(paints, PAINT_DTYPE)
            st = arr(stops, np.float32).reshape(-1, 5)
            v.paints, v.n_paints = g.ctypes.data, len(g)
            v.stops, v.n_stops = st.ctypes.data, len(st)
        out = C.c_void_p()
        _check(lib().spr_scene_create(C.byref(v), C.byref(out)))
        return cls(out.value)

    @property
    def n_paths(self) -> int:
        return len(self.arrays()["rule"])

    @property
    def n_curves(self) -> int:
        return len(self.arrays()["curve_type"])

    @property
    def n_points(self) -> int:
        return len(self.arrays()["pos_path"])

    def __del__(self):
        try:
            if self._h and self._h.value:
                lib().spr_scene_free(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


class Rasterizer:
    """One rasteriser handle: a w x h RGBA8 target on one GPU, with its own HIP stream."""

    def __init__(self, width: int, height: int, compose: int = COMPOSE_OVERWRITE,
                 mode: int = MODE_PLAIN, device: int = -1, debug: bool = False, aa: int = 0):
        self.width, self.height = int(width), int(height)
        o = Opts(compose=compose, mode=mode, device=device, debug=1 if debug else 0, aa=int(aa))
        out = C.c_void_p()
        _check(lib().spr_create(self.width, self.height, C.byref(o), C.byref(out)))
        self._h = out

    def _c(self, rc):
        _check(rc, self._h)

    def close(self):
        if self._h and self._h.value:
            lib().spr_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_scene(self, scene: Scene) -> None:
        self._c(lib().spr_load_scene(self._h, scene._h))

    def load_rvg(self, path: str | Path) -> None:
        self._c(lib().spr_load_rvg(self._h, str(path).encode()))

    def load_paths(self, pos, pos_path, curve_pos, curve_type, curve_path, rule, fill) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
        arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in
                (pos_path, curve_pos, curve_type, curve_path, rule, fill)]
        pp, cp_, ct, cpa, ru, fi = arrs
        self._c(lib().spr_load_paths(self._h, _ptr(pos, C.c_float), _ptr(pp, C.c_uint32), len(pp),
                                     _ptr(cp_, C.c_uint32), _ptr(ct, C.c_uint32),
                                     _ptr(cpa, C.c_uint32), len(ct), _ptr(ru, C.c_uint32),
                                     _ptr(fi, C.c_uint32), len(ru)))

    def set_mvp(self, m_rows) -> None:
        m = np.ascontiguousarray(m_rows, dtype=np.float32).reshape(16)
        self._c(lib().spr_set_mvp(self._h, _ptr(m, C.c_float)))

    def render(self, readback: bool = True):
        if readback:
            img = np.empty((self.height, self.width, 4), dtype=np.uint8)
            self._c(lib().spr_render_frame(self._h, img.ctypes.data_as(C.c_void_p)))
            return img
        self._c(lib().spr_render_frame(self._h, None))
        return None

    def synchronize(self) -> None:
        self._c(lib().spr_synchronize(self._h))

    def set_stage_timing(self, on: bool) -> None:
        self._c(lib().spr_set_stage_timing(self._h, 1 if on else 0))

    def stats(self) -> Stats:
        s = Stats()
        self._c(lib().spr_get_stats(self._h, C.byref(s)))
        return s

    def image_device_ptr(self) -> int:
        p = C.c_void_p()
        self._c(lib().spr_image_device_ptr(self._h, C.byref(p)))
        return p.value

    def export(self, stage: str) -> np.ndarray:
        sid = STAGES[stage]
        n = C.c_size_t()
        self._c(lib().spr_debug_export(self._h, sid, None, 0, C.byref(n)))
        dt = _STAGE_DTYPE[stage]
        out = np.empty(n.value // np.dtype(dt).itemsize, dtype=dt)
        self._c(lib().spr_debug_export(self._h, sid, out.ctypes.data_as(C.c_void_p), n.value,
                                       C.byref(n)))
        return out

    def scan_i32(self, d_in: int, d_out: int, n: int, stream: int | None = None) -> None:
        """Exclusive prefix sum of n int32 at device address d_in into d_out (n + 1 values, the
        total last), enqueued on `stream` (a hipStream_t handle; None = the handle's stream)."""
        self._c(lib().spr_scan_i32(self._h, C.c_void_p(d_in), C.c_void_p(d_out), int(n),
                                   C.c_void_p(stream) if stream else None))

    def write_png(self, path) -> None:
        """The last frame as an 8-bit RGBA PNG file."""
        self._c(lib().spr_write_png(self._h, str(path).encode()))

    def inject_records(self, records) -> None:
        """Following frames draw these (R, 4) int32 records (yx, width, fill, frag_index) instead
        of rendering the scene (record replay); None clears them."""
        if records is None:
            self._c(lib().spr_debug_inject(self._h, STAGES["records"], None, 0))
            return
        a = np.ascontiguousarray(records, dtype=np.int32).reshape(-1)
        self._c(lib().spr_debug_inject(self._h, STAGES["records"], a.ctypes.data_as(C.c_void_p), a.nbytes))

    def inject_intersections(self, inter) -> None:
        if inter is None:
            self._c(lib().spr_debug_inject(self._h, STAGES["inter"], None, 0))
            return
        a = np.ascontiguousarray(inter, dtype=np.int32).reshape(-1)
        self._c(lib().spr_debug_inject(self._h, STAGES["inter"], a.ctypes.data_as(C.c_void_p),
                                       a.nbytes))


def _rasterizer_learned_methods():
    """Learned single-pass allocation: training, model I/O and the predictor's kernels."""

    def learned_train(self, scene: Scene, epochs: int = 0, batch: int = 0, lr: float = 0.0,
                      chunk_paths: int = 0, seed: int = 4, overflow_quantile: float = 0.0) -> dict:
        o = TrainOpts(epochs=epochs, batch=batch, lr=lr, chunk_paths=chunk_paths, seed=seed,
                      overflow_quantile=overflow_quantile)
        r = TrainReport()
        self._c(lib().spr_learned_train(self._h, scene._h, C.byref(o), C.byref(r)))
        return {k: getattr(r, k) for k, _ in TrainReport._fields_}

    def learned_save(self, path) -> None:
        self._c(lib().spr_learned_save(self._h, str(path).encode()))

    def learned_load(self, path) -> None:
        self._c(lib().spr_learned_load(self._h, str(path).encode()))

    def learned_set_margin(self, log2_margin: float, slack: int) -> None:
        self._c(lib().spr_learned_set_margin(self._h, float(log2_margin), int(slack)))

    def mlp_params(self) -> np.ndarray:
        out = np.empty(MLP_FLOATS, dtype=np.float32)
        n = C.c_size_t()
        self._c(lib().spr_mlp_get_params(self._h, _ptr(out, C.c_float), out.size, C.byref(n)))
        return out

    def mlp_set_params(self, params) -> None:
        a = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        self._c(lib().spr_mlp_set_params(self._h, _ptr(a, C.c_float), a.size))

    def mlp_predict(self, features) -> np.ndarray:
        x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, 16)
        out = np.empty((len(x), 2), dtype=np.float32)
        self._c(lib().spr_mlp_predict(self._h, _ptr(x, C.c_float), len(x), _ptr(out, C.c_float)))
        return out

    def mlp_gradients(self, features, targets):
        x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 2)
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        loss = C.c_float()
        self._c(lib().spr_mlp_gradients(self._h, _ptr(x, C.c_float), _ptr(t, C.c_float), len(x),
                                        _ptr(g, C.c_float), C.byref(loss)))
        return g, float(loss.value)

    for f in (learned_train, learned_save, learned_load, learned_set_margin, mlp_params,
              mlp_set_params, mlp_predict, mlp_gradients):
        setattr(Rasterizer, f.__name__, f)


_rasterizer_learned_methods()


def _rasterizer_batch_methods():
    """Batches of documents rendered by one frame (one captured graph once unchanged)."""

    def load_batch(self, scenes, canvases) -> None:
        """scenes: list of Scene; canvases: list of (width, height)."""
        n = len(scenes)
        arr = (C.c_void_p * n)(*[sc._h.value for sc in scenes])
        w = np.ascontiguousarray([c[0] for c in canvases], dtype=np.uint32)
        h = np.ascontiguousarray([c[1] for c in canvases], dtype=np.uint32)
        self._c(lib().spr_load_batch(self._h, arr, _ptr(w, C.c_uint32), _ptr(h, C.c_uint32), n))
        self._batch = [(int(a), int(b)) for a, b in zip(w, h)]
        self._batch_scenes = list(scenes)  # keep the scene handles alive with the batch

    def batch_image(self, doc: int) -> np.ndarray:
        w, h = self._batch[doc]
        out = np.empty((h, w, 4), dtype=np.uint8)
        self._c(lib().spr_batch_image(self._h, int(doc), out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def autotune(self, reps: int = 3) -> dict:
        r = TuneReport()
        self._c(lib().spr_autotune(self._h, int(reps), C.byref(r)))
        return {k: getattr(r, k) for k, _ in TuneReport._fields_ if k != "reserved"}

    def set_tuning(self, chain_blocks_per_cu=8, merge_blocks_per_cu=8, scan_shape=0, band_window_log=9) -> None:
        r = TuneReport(chain_blocks_per_cu=chain_blocks_per_cu, merge_blocks_per_cu=merge_blocks_per_cu,
                       scan_shape=scan_shape, band_window_log=band_window_log)
        self._c(lib().spr_set_tuning(self._h, C.byref(r)))

    for f in (load_batch, batch_image, autotune, set_tuning):
        setattr(Rasterizer, f.__name__, f)


_rasterizer_batch_methods()
