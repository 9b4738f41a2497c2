"""Gaussian splat clouds in the kernels' layout (src/render/splats.lucb).

Every splat cloud of a set and its visible instances, placed in world space
and packed for ray tracing as 3D Gaussian Ray Tracing does (Moenne-Loccoz et
al. 2024): a ray meets each Gaussian once, at its point of maximum response,
where it covers alpha = min(0.99, opacity * exp(-d^2 / 2)), d the ray's
Mahalanobis distance from the centre; splats whose alpha stays under the cull
(1/255) or past 3 sigma never count.

Per splat, 16 u32 words (splats.lucb:15-25):
  0..3   the centre (f32 x 3) and the opacity
  4..7   the turn from world to the splat's axes (the conjugate orient, xyzw)
  8..11  one over the scales (f32 x 3), and the splat's first harmonics word
  12..15 the DC colour in the cloud's encoding (f32 x 3), and flags: the SH
         items (0, 3, 8 or 15) in bits 0..4, bit 5 when the encoding is sRGB.
`harmonics` holds, per splat, its SH frame turn (world to the frame the SH
were authored in, half x 4 in two words) and its SH items (RGB halves, padded
to a whole word). `boxes` reach out to where alpha falls to the cull (at most
3 sigma); the splats' own binary BVH is built over them (build_tree,
bvh.lucb:85) and the records put in its leaf order.
"""
import math

import numpy as np

from . import _lib
from ._lib import INVALID, LIMIT, LuceError

SPLAT_CULL = 0.00392156862745098  # 1/255: 3DGS's alpha cull (splats.lucb:33)
RECORD_WORDS = 16
FLAG_SRGB = 32
MAX_SPLATS = 1 << 26


def qmul(a, b):
    """Hamilton products of (..., 4) xyzw quaternions."""
    ax, ay, az, aw = np.moveaxis(a, -1, 0)
    bx, by, bz, bw = np.moveaxis(b, -1, 0)
    return np.stack([aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw,
                     aw * bw - ax * bx - ay * by - az * bz], axis=-1)


def qconj(q):
    return q * np.array([-1.0, -1.0, -1.0, 1.0])


def qmatrix(q):
    """Rotation matrices (..., 3, 3) of unit xyzw quaternions, acting on
    column vectors (local to world for an orient)."""
    x, y, z, w = np.moveaxis(q, -1, 0)
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
                     np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
                     np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)


def quaternion_of(r):
    """The unit xyzw quaternion of a 3x3 rotation matrix (Shepperd's method)."""
    r = np.asarray(r, np.float64)
    t = np.trace(r)
    if t > 0.0:
        s = math.sqrt(t + 1.0) * 2.0
        q = [(r[2, 1] - r[1, 2]) / s, (r[0, 2] - r[2, 0]) / s, (r[1, 0] - r[0, 1]) / s, 0.25 * s]
    elif r[0, 0] > r[1, 1] and r[0, 0] > r[2, 2]:
        s = math.sqrt(1.0 + r[0, 0] - r[1, 1] - r[2, 2]) * 2.0
        q = [0.25 * s, (r[0, 1] + r[1, 0]) / s, (r[0, 2] + r[2, 0]) / s, (r[2, 1] - r[1, 2]) / s]
    elif r[1, 1] > r[2, 2]:
        s = math.sqrt(1.0 + r[1, 1] - r[0, 0] - r[2, 2]) * 2.0
        q = [(r[0, 1] + r[1, 0]) / s, 0.25 * s, (r[1, 2] + r[2, 1]) / s, (r[0, 2] - r[2, 0]) / s]
    else:
        s = math.sqrt(1.0 + r[2, 2] - r[0, 0] - r[1, 1]) * 2.0
        q = [(r[0, 2] + r[2, 0]) / s, (r[1, 2] + r[2, 1]) / s, 0.25 * s, (r[1, 0] - r[0, 1]) / s]
    q = np.array(q)
    return q / np.linalg.norm(q)


def unit_or(q, fallback):
    """Each quaternion normalized, or `fallback`'s where its squared norm is
    at most 1e-30 (splats.lucb:131)."""
    n2 = (q * q).sum(-1, keepdims=True)
    return np.where(n2 > 1e-30, q / np.sqrt(np.where(n2 > 1e-30, n2, 1.0)), fallback)


def srgb_encode(c):
    """Linear to the sRGB encoding (IEC 61966-2-1)."""
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(np.maximum(c, 0.0031308), 1.0 / 2.4) - 0.055)


def _halves(values):
    return np.asarray(np.clip(values, -60000.0, 60000.0), np.float64).astype(np.float16).view(np.uint16).astype(np.uint32)


class CompiledSplats:
    """The splats of a compiled scene (splats.lucb:45-70): records (n, 16)
    u32 in leaf order once built, harmonics u32, boxes (n, 6) f32, the
    splats' BVH nodes, and the count of plain point clouds skipped."""

    def __init__(self):
        self.records = np.zeros((0, RECORD_WORDS), np.uint32)
        self.harmonics = np.zeros(0, np.uint32)
        self.boxes = np.zeros((0, 6), np.float32)
        self.nodes = None
        self.order = np.zeros(0, np.uint32)
        self.plain_clouds = 0
        self._records, self._harmonics, self._boxes = [], [], []

    def count(self):
        return len(self.records) + sum(len(r) for r in self._records)

    def add_cloud(self, cloud, matrix):
        """add_cloud (splats.lucb:71-108): `cloud` placed by `matrix` (row
        vectors: rows 0..2 axes, row 3 translation). A cloud without orient,
        scale and opacity is a plain point cloud: counted, not rendered."""
        if cloud.orient is None or cloud.scale is None or cloud.opacity is None:
            self.plain_clouds += 1
            return
        count = len(cloud.points)
        if count == 0:
            return
        if self.count() + count > MAX_SPLATS:
            raise LuceError(LIMIT, "a render holds at most 2^26 splats")
        m = np.asarray(matrix, np.float64)
        a = m[:3, :3].T  # the placement's linear part acting on column vectors
        scale_of = abs(np.linalg.det(a)) ** (1.0 / 3.0)
        identity = np.array_equal(m, np.eye(4))
        if not identity:
            if not (np.linalg.det(a) > 0.0) or not np.allclose(a.T @ a, np.eye(3) * scale_of * scale_of,
                                                               rtol=1e-6, atol=1e-9 * scale_of * scale_of):
                raise LuceError(INVALID, "a splat cloud is placed by rotations, translations and uniform scales only")
        p_rot = np.array([0.0, 0.0, 0.0, 1.0]) if identity else quaternion_of(a / scale_of)
        pts = np.asarray(cloud.points, np.float64).reshape(-1, 3) + np.asarray(cloud.origin, np.float64)
        center = pts @ m[:3, :3] + m[3, :3]
        ident = np.array([0.0, 0.0, 0.0, 1.0])
        local = unit_or(np.asarray(cloud.orient, np.float64).reshape(-1, 4), ident)
        spin = local if identity else unit_or(qmul(np.broadcast_to(p_rot, local.shape), local), ident)
        scale = np.maximum(np.abs(np.asarray(cloud.scale, np.float64).reshape(-1, 3)) * (1.0 if identity else scale_of),
                           1e-12)
        opacity = np.clip(np.asarray(cloud.opacity, np.float64).reshape(-1), 0.0, 1.0)
        srgb = cloud.color_space == "srgb"
        linear = (np.asarray(cloud.color, np.float64).reshape(-1, 3) if cloud.color is not None
                  else np.full((count, 3), 0.5))
        encoded = srgb_encode(linear) if srgb else linear
        items = 0
        sh = None
        if cloud.sh is not None:
            sh = np.asarray(cloud.sh, np.float64).reshape(count, -1, 3)
            items = sh.shape[1]
            if items not in (0, 3, 8, 15):
                raise LuceError(INVALID, "splat SH hold 0, 3, 8 or 15 items (degrees 0 to 3)")
        stride = 2 + (items * 3 + 1) // 2
        first = sum(len(h) for h in self._harmonics) + len(self.harmonics)
        records = np.zeros((count, RECORD_WORDS), np.uint32)
        turn = qconj(spin)
        f32 = np.column_stack([center, opacity, turn, 1.0 / scale]).astype(np.float32)
        records[:, 0:11] = f32.view(np.uint32)
        records[:, 11] = first + np.arange(count, dtype=np.uint64) * stride
        records[:, 12:15] = encoded.astype(np.float32).view(np.uint32)
        records[:, 15] = items | (FLAG_SRGB if srgb else 0)
        # The SH frame: world to the frame the SH were authored in (the
        # splat's turn since its rest orientation, in the cloud's frame).
        rest = local if cloud.restorient is None else unit_or(np.asarray(cloud.restorient, np.float64).reshape(-1, 4),
                                                              local)
        frame = qconj(qmul(spin, qconj(rest)))
        harm = np.zeros((count, stride), np.uint32)
        fh = _halves(frame)
        harm[:, 0] = fh[:, 0] | (fh[:, 1] << 16)
        harm[:, 1] = fh[:, 2] | (fh[:, 3] << 16)
        if items:
            h = _halves(sh.reshape(count, items * 3))
            if (items * 3) % 2:
                h = np.concatenate([h, np.zeros((count, 1), np.uint32)], axis=1)
            harm[:, 2:] = h[:, 0::2] | (h[:, 1::2] << 16)
        # Boxes out to where alpha falls to the cull, at most 3 sigma.
        reach = np.where(opacity > SPLAT_CULL, 2.0 * np.log(np.maximum(opacity, SPLAT_CULL) / SPLAT_CULL), 0.0)
        k = np.minimum(3.0, np.sqrt(np.maximum(reach, 0.0)))
        rot = qmatrix(spin)
        extent = k[:, None] * np.sqrt(((rot * scale[:, None, :]) ** 2).sum(-1))
        boxes = np.column_stack([center - extent, center + extent]).astype(np.float32)
        self._records.append(records)
        self._harmonics.append(harm.reshape(-1))
        self._boxes.append(boxes)

    def finish(self, threads=0):
        """build_splat_bvh (splats.lucb:193-213): the tree over the boxes and
        the records and boxes in its leaf order."""
        if self._records:
            self.records = np.ascontiguousarray(np.concatenate([self.records] + self._records))
            self.harmonics = np.ascontiguousarray(np.concatenate([self.harmonics] + self._harmonics))
            self.boxes = np.ascontiguousarray(np.concatenate([self.boxes] + self._boxes))
            self._records, self._harmonics, self._boxes = [], [], []
        from .scene import NODE_DTYPE
        n = len(self.records)
        nodes = np.zeros(max(1, 2 * n), NODE_DTYPE)
        order = np.zeros(max(1, n), np.uint32)
        got = _lib.ctypes.c_size_t(0)
        _lib.check(_lib.lib().lr_bvh_build_boxes(_lib.ptr(self.boxes), n, int(threads), _lib.ptr(nodes),
                                                 _lib.ctypes.byref(got), _lib.ptr(order)))
        self.nodes = nodes[:got.value].copy()
        self.order = order[:n].copy()
        if n:
            self.records = np.ascontiguousarray(self.records[self.order])
            self.boxes = np.ascontiguousarray(self.boxes[self.order])
        return self


def decode_records(records):
    """Records back to (centre, opacity, turn, inverse scales, first word,
    encoded colour, items, sRGB flag), f64 from the stored f32 bits."""
    r = np.asarray(records, np.uint32).reshape(-1, RECORD_WORDS)
    f = r.view(np.float32).astype(np.float64)
    return (f[:, 0:3], f[:, 3], f[:, 4:8], f[:, 8:11], r[:, 11].astype(np.int64), f[:, 12:15], r[:, 15] & 31,
            (r[:, 15] & FLAG_SRGB) != 0)


def harmonics_of(words, first, items):
    """A splat's SH frame (xyzw) and items (items, 3) from its harmonics words."""
    w = np.asarray(words, np.uint32)
    stride = 2 + (items * 3 + 1) // 2
    block = w[first:first + stride]
    halves = np.stack([block & 0xFFFF, block >> 16], 1).reshape(-1).astype(np.uint16).view(np.float16)
    frame = halves[0:4].astype(np.float64)
    return frame, halves[4:4 + items * 3].astype(np.float64).reshape(items, 3)
