"""Oracle (test infrastructure only): ctypes access to the C restatement in
oracle_trace.c. Imported by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg, never by the product.

Build with `make -C oracle` (the library lands in oracle/_build/).
"""
import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

NODE_DTYPE = np.dtype([("lo", "<f4", 3), ("first", "<u4"), ("hi", "<f4", 3), ("count", "<u4")])

_P = ctypes.c_void_p
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        sig = {
            "or_ray_frame": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, _P]),
            "or_look_frame": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.c_int, _P]),
            "or_pack_frame": (None, [_P, ctypes.c_uint32, ctypes.c_uint32, _P]),
            "or_bvh_build": (ctypes.c_size_t, [_P, _P, ctypes.c_size_t, _P, _P, _P]),
            "or_reorder_u32": (None, [_P, _P, _P, ctypes.c_size_t, ctypes.c_size_t]),
            "or_check_bvh": (ctypes.c_int, [_P, ctypes.c_size_t, _P, _P, ctypes.c_size_t]),
            "or_probe": (None, [_P, _P, _P, _P, _P, ctypes.c_uint32, ctypes.c_uint32, _P, _P, _P, _P]),
            "or_trace_rays": (None, [_P, _P, _P, _P, _P, ctypes.c_size_t, ctypes.c_float, _P, _P, _P]),
            "or_render": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, _P, _P, _P, _P,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _P, ctypes.c_size_t,
                                         ctypes.c_int, _P, _P, _P, ctypes.c_size_t, _P, _P, ctypes.c_size_t]),
            "or_render_guided": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, _P, _P, _P,
                                                _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _P,
                                                ctypes.c_size_t, ctypes.c_int, _P, _P, _P, ctypes.c_size_t, _P, _P,
                                                ctypes.c_size_t, _P, _P, _P]),
            "or_plant": (None, [ctypes.c_int]),
            "or_vmf_pdf": (ctypes.c_double, [_P, ctypes.c_double, _P]),
            "or_vmf_sample": (None, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _P]),
            "or_guide_fit_cell": (None, [_P, _P]),
            "or_guide_key": (ctypes.c_uint32, [_P, _P, _P, _P]),
            "or_run_program": (None, [_P, _P, ctypes.c_size_t, _P, _P, ctypes.c_size_t, _P, _P, _P, ctypes.c_uint32,
                                      _P, _P, _P, _P, _P, _P, _P]),
            "or_gradient_noise": (ctypes.c_double, [ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "or_sample4": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _P, _P]),
            "or_make_tables": (None, [_P]),
            "or_brute_force": (None, [_P, _P, _P, ctypes.c_size_t, _P, ctypes.c_uint32, ctypes.c_uint32, _P,
                                      ctypes.c_size_t, ctypes.c_int, _P, _P, _P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def ray_frame(matrix, orthographic, horizontal_aperture, focal_length, width, height, f_stop=0.0, focus=1.0):
    """camera.lucb:57-76 in f64: a dict of the frame's parts."""
    m = np.ascontiguousarray(np.asarray(matrix, np.float64)[:4, :3].reshape(12))
    out = np.zeros(18)
    rc = lib().or_ray_frame(_p(m), int(orthographic), horizontal_aperture, focal_length, f_stop, focus, width,
                            height, _p(out))
    if rc:
        raise ValueError(rc)
    return out


def look_frame(position, forward, up, tan_half_height, width, height, orthographic=False):
    pos = np.asarray(position, np.float64)
    fw = np.asarray(forward, np.float64)
    u = np.asarray(up, np.float64)
    out = np.zeros(18)
    rc = lib().or_look_frame(_p(pos), _p(fw), _p(u), tan_half_height, width, height, int(orthographic), _p(out))
    if rc:
        raise ValueError(rc)
    return out


def frame_ray(frame, x, y):
    """camera.lucb:51-55 on an f64 frame from ray_frame/look_frame."""
    film = frame[3:6] + frame[6:9] * x + frame[9:12] * y
    if frame[17] > 0.5:
        return frame[0:3] + film, frame[12:15].copy()
    return frame[0:3].copy(), film


def pack_frame(frame, width, height):
    out = np.zeros(16, np.float32)
    lib().or_pack_frame(_p(np.ascontiguousarray(frame)), width, height, _p(out))
    return out


def bvh_build(positions, indices):
    """bvh.lucb:66-129: (nodes, order, max |f32 - f64| SAH cost gap)."""
    positions = np.ascontiguousarray(positions, np.float32)
    indices = np.ascontiguousarray(indices, np.uint32)
    n = len(indices)
    nodes = np.zeros(max(1, 2 * n), NODE_DTYPE)
    order = np.zeros(max(1, n), np.uint32)
    gap = ctypes.c_double(0.0)
    count = lib().or_bvh_build(_p(positions), _p(indices), n, _p(nodes), _p(order), ctypes.byref(gap))
    return nodes[:count].copy(), order[:n].copy(), gap.value


def check_bvh(nodes, positions, indices):
    """tests.lucb:9-28: 0 when every invariant holds."""
    return lib().or_check_bvh(_p(np.ascontiguousarray(nodes)), len(nodes), _p(np.ascontiguousarray(positions, np.float32)),
                              _p(np.ascontiguousarray(indices, np.uint32)), len(indices))


def probe(nodes, positions, indices, faces, frame16, width, height):
    """probe.comp over the f32 mirror: faces, t, slots, counters
    (interior visits, leaf visits, triangle tests, dropped stack entries)."""
    n = width * height
    f = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    s = np.zeros(n, np.uint32)
    c = np.zeros(4, np.uint64)
    lib().or_probe(_p(np.ascontiguousarray(nodes)), _p(np.ascontiguousarray(positions, np.float32)),
                   _p(np.ascontiguousarray(indices, np.uint32)), _p(np.ascontiguousarray(faces, np.uint32)),
                   _p(np.ascontiguousarray(frame16, np.float32)), width, height, _p(f), _p(t), _p(s), _p(c))
    return f, t, s, c


def trace_rays(nodes, positions, indices, faces, rays, t_max=3.0e38):
    """trace_closest over rays (n, 6) f32 (origin, direction)."""
    rays = np.ascontiguousarray(rays, np.float32)
    n = len(rays)
    f = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    c = np.zeros(4, np.uint64)
    lib().or_trace_rays(_p(np.ascontiguousarray(nodes)), _p(np.ascontiguousarray(positions, np.float32)),
                        _p(np.ascontiguousarray(indices, np.uint32)), _p(np.ascontiguousarray(faces, np.uint32)),
                        _p(rays), n, t_max, _p(f), _p(t), _p(c))
    return f, t, c


def brute_force(positions, indices, faces, frame=None, width=0, height=0, rays=None, threads=1):
    """tests/gpu/main.lucb:82-98 in f64: (face or -1, distance, margin)."""
    n = len(rays) if rays is not None else width * height
    face = np.zeros(n, np.int64)
    dist = np.zeros(n, np.float64)
    margin = np.zeros(n, np.float64)
    r = None if rays is None else np.ascontiguousarray(rays, np.float64)
    fr = None if frame is None else np.ascontiguousarray(frame, np.float64)
    lib().or_brute_force(_p(np.ascontiguousarray(positions, np.float32)), _p(np.ascontiguousarray(indices, np.uint32)),
                         _p(np.ascontiguousarray(faces, np.uint32)), len(indices), _p(fr), width, height, _p(r),
                         n if r is not None else 0, threads, _p(face), _p(dist), _p(margin))
    return face, dist, margin


def agrees(got_face, got_t, face, dist, margin):
    """The reference's comparison (tests/gpu/main.lucb:61-78): same polygon
    (or both missed) and |dt| <= 1e-4 t; disagreements count only where the
    next polygon is more than 1e-3 farther (relative)."""
    got_face = np.asarray(got_face, np.int64)
    hit = face >= 0
    missed = got_face == 0xFFFFFFFF
    same = np.where(hit, (~missed) & (got_face == face) & (np.abs(got_t.astype(np.float64) - dist) <= 1e-4 * dist),
                    missed)
    differ = (~same) & (margin > 1e-3)
    return same, differ


def sample4(seed, pixel, index, group):
    """sampler.glsl:47-55: the four 24-bit integers and their values."""
    bits = np.zeros(4, np.uint32)
    vals = np.zeros(4, np.float64)
    lib().or_sample4(seed, pixel, index, group, _p(bits), _p(vals))
    return bits, vals


def tables():
    """tables.comp's albedo tables in f64 (17,440 values)."""
    t = np.zeros(32 * 32 + 32 + 32 * 32 * 16, np.float64)
    lib().or_make_tables(_p(t))
    return t


def render(scene, loaded, frame16, width, height, samples, max_bounces, spectral, seed, pixels, threads=8,
           guided=False, guide_state=None):
    """The f64 camera-path integrator over the listed pixels: ACEScg (n, 3).
    `loaded` carries the packed lights/constants/materials/emitters. With
    `guided` the hash-grid guide learns sample by sample over every pixel
    (the listed ones are returned)."""
    pixels = np.ascontiguousarray(pixels, np.uint32)
    out = np.zeros((len(pixels), 3), np.float64)
    extra = ()
    if guided:
        cells = 524288
        if guide_state is not None:
            guide_state["keys"] = np.zeros(cells, np.uint32)
            guide_state["lobes"] = np.zeros((cells, 5, 4), np.float32)
            guide_state["stats"] = np.zeros((cells, 4, 5), np.float32)
            extra = (_p(guide_state["keys"]), _p(guide_state["lobes"]), _p(guide_state["stats"]))
        else:
            extra = (None, None, None)
    fn = lib().or_render_guided if guided else lib().or_render
    fn(_p(np.ascontiguousarray(scene.positions, np.float32)), _p(np.ascontiguousarray(scene.indices, np.uint32)),
                    _p(np.ascontiguousarray(scene.shading, np.uint32)), scene.triangle_count(),
                    _p(np.ascontiguousarray(scene.nodes)), len(scene.nodes),
                    _p(loaded.lights), _p(loaded.constants), _p(loaded.materials), _p(loaded.emitters),
                    _p(np.ascontiguousarray(frame16, np.float32)), width, height, samples, max_bounces,
                    1 if spectral else 0, seed, loaded.through, _p(pixels), len(pixels), threads, _p(out),
                    _p(np.ascontiguousarray(scene.uvs, np.float32)), _p(loaded.programs), len(loaded.programs),
                    _p(loaded.textures), _p(loaded.texture_table), loaded.texture_count, *extra)
    return out


def run_program(materials, programs, material, x, n, view, uv, corners=None, uvs=None, frame=None, textures=None,
                texture_table=None, texture_count=0):
    """One node-graph program run (shader_program.glsl:205-349) at a hit on
    one triangle: (mask, viewed, values[23])."""
    materials = np.ascontiguousarray(materials, np.float32)
    programs = np.ascontiguousarray(programs, np.uint32)
    corners = np.ascontiguousarray(np.zeros(9) if corners is None else corners, np.float32).reshape(-1)
    uvs = np.ascontiguousarray(np.zeros(6) if uvs is None else uvs, np.float32).reshape(-1)
    frame = np.ascontiguousarray(np.zeros(16) if frame is None else frame, np.float32)
    textures = np.ascontiguousarray(np.zeros(4) if textures is None else textures, np.float32)
    texture_table = np.ascontiguousarray(np.zeros(4) if texture_table is None else texture_table, np.uint32)
    d = [np.ascontiguousarray(v, np.float64) for v in (x, n, view, uv)]
    mask, viewed = ctypes.c_uint32(0), ctypes.c_int(0)
    values = np.zeros(23, np.float64)
    lib().or_run_program(_p(materials), _p(programs), len(programs), _p(textures), _p(texture_table), texture_count,
                         _p(corners), _p(uvs), _p(frame), material, *[_p(v) for v in d], ctypes.byref(mask),
                         ctypes.byref(viewed), _p(values))
    return mask.value, bool(viewed.value), values


def gradient_noise(x, y, z):
    return lib().or_gradient_noise(x, y, z)


def adaptive_error(film_sum, film_half):
    """adaptive.comp:22-28, phase 0: each pixel's error, |I - A| summed over
    the colour channels over the weight, relative to 1e-4 + sqrt(brightness),
    in f32 as the kernel computes it. film_sum / film_half: (..., 4) f32."""
    s = np.asarray(film_sum, np.float32)
    h = np.asarray(film_half, np.float32)
    scale = np.float32(1.0) / np.maximum(s[..., 3], np.float32(1.0))
    diff = (np.abs(s[..., 0] - h[..., 0]) + np.abs(s[..., 1] - h[..., 1]) + np.abs(s[..., 2] - h[..., 2])) * scale
    bright = np.maximum(np.float32(0.0), (s[..., 0] + s[..., 1] + s[..., 2]) * scale)
    return (diff / (np.float32(1e-4) + np.sqrt(bright))).astype(np.float32)


def adaptive_verdict(film_sum, film_half, threshold):
    """adaptive.comp:16-42: (own, converged) per pixel of (H, W, 4) planes. A
    pixel's own verdict is error < threshold (phase 0); it is converged when
    its whole 3 x 3 neighbourhood, clipped at the image's edges, is (phase 1)."""
    own = adaptive_error(film_sum, film_half) < np.float32(threshold)
    h, w = own.shape
    padded = np.ones((h + 2, w + 2), bool)
    padded[1:-1, 1:-1] = own
    conv = np.ones((h, w), bool)
    for dy in (0, 1, 2):
        for dx in (0, 1, 2):
            conv &= padded[dy:dy + h, dx:dx + w]
    return own, conv


def vmf_pdf(mean, kappa, d):
    """guide.glsl:84-88: a vMF lobe's density at unit direction d."""
    return lib().or_vmf_pdf(_p(np.ascontiguousarray(mean, np.float64)), float(kappa),
                            _p(np.ascontiguousarray(d, np.float64)))


def vmf_sample(mean, kappa, ux, uy):
    """guide.glsl:90-98: the direction drawn from a lobe by (ux, uy)."""
    out = np.zeros(3, np.float64)
    lib().or_vmf_sample(_p(np.ascontiguousarray(mean, np.float64)), float(kappa), float(ux), float(uy), _p(out))
    return out


def guide_fit_cell(stats, lobes=None):
    """guide_fit.comp:16-35 on one cell: stats (4, 5) f32 (weight, weighted
    direction, count per lobe) -> (lobes (4, 4): mean, kappa; weights (4,);
    the stats after the decay)."""
    st = np.ascontiguousarray(np.asarray(stats, np.float32).reshape(20)).copy()
    lb = np.zeros(20, np.float32) if lobes is None else np.ascontiguousarray(np.asarray(lobes, np.float32).reshape(20)).copy()
    lib().or_guide_fit_cell(_p(st), _p(lb))
    return lb[:16].reshape(4, 4), lb[16:20].copy(), st.reshape(4, 5)


def guide_key(constants, frame16, x, n):
    """guide.glsl:39-61's key of point x with normal n (the cell before its slot)."""
    return lib().or_guide_key(_p(np.ascontiguousarray(constants, np.float32)),
                              _p(np.ascontiguousarray(frame16, np.float32)),
                              _p(np.ascontiguousarray(x, np.float64)), _p(np.ascontiguousarray(n, np.float64)))


def plant(mistake):
    """Plants a known mistake in the restatement (0: none; 1: the light
    sample's MIS weight without the light tracer's term), for tests that
    check that a comparison catches it."""
    lib().or_plant(int(mistake))
