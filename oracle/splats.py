"""Oracle (test infrastructure only): Gaussian splats along a ray in numpy f64,
the model of shaders/splats.glsl and of the reference's CPU check
(tests/gpu/splats.lucb:88-137), evaluated from the splats' description (not
from the packed records): each Gaussian at its point of maximum response
along the ray, alpha = min(0.99, opacity exp(-d^2/2)) when d <= 3 and alpha
>= 1/255, composited front to back. Nothing in the product imports this file.

Pinning: the maximum-response formula against a dense search along the ray;
the SH basis against its orthonormality over the sphere (a product Gauss
quadrature); the composite against the closed form of one splat's shadow
(1 - alpha of the light) in the GPU tests.
"""
import numpy as np

CULL = 0.00392156862745098
C1 = 0.4886025119029199
C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396)
C3 = (-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
      1.445305721320277, -0.5900435899266435)
REC709_TO_ACESCG = np.array([[0.613097, 0.339523, 0.047379],
                             [0.070194, 0.916354, 0.013452],
                             [0.020616, 0.109570, 0.869815]])


def quat_matrix(q):
    """Rotation matrices (n, 3, 3), local to world, of xyzw quaternions
    (normalized here)."""
    q = np.asarray(q, np.float64).reshape(-1, 4)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    x, y, z, w = q.T
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
                     np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
                     np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)


def about(axis, angle):
    """The xyzw quaternion turning `angle` radians about `axis`."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    return np.array([*(a * np.sin(0.5 * angle)), np.cos(0.5 * angle)])


def alpha(splats, o, d):
    """Each splat's (alpha, t) along the ray from o along unit d
    (splats.glsl:66-82 in f64)."""
    r = quat_matrix(splats["orient"]).transpose(0, 2, 1)  # world to the splat's axes
    inv = 1.0 / np.asarray(splats["scale"], np.float64)
    a = np.einsum("nij,nj->ni", r, np.asarray(o, np.float64) - splats["center"]) * inv
    b = np.einsum("nij,j->ni", r, np.asarray(d, np.float64)) * inv
    dd = (b * b).sum(1)
    t = -(a * b).sum(1) / dd
    c = np.cross(a, b)
    m2 = (c * c).sum(1) / dd
    al = np.minimum(0.99, np.asarray(splats["opacity"], np.float64) * np.exp(-0.5 * m2))
    al = np.where((m2 <= 9.0) & (al >= CULL), al, 0.0)
    return al, t


def sh_basis(d, items):
    """The real SH basis toward unit d (3DGS's signs and constants), items
    0, 3, 8 or 15 functions of bands 1..3."""
    x, y, z = d
    out = [-C1 * y, C1 * z, -C1 * x]
    if items >= 8:
        xx, yy, zz = x * x, y * y, z * z
        out += [C2[0] * x * y, C2[1] * y * z, C2[2] * (2 * zz - xx - yy), C2[3] * x * z, C2[4] * (xx - yy)]
    if items >= 15:
        out += [C3[0] * y * (3 * xx - yy), C3[1] * x * y * z, C3[2] * y * (4 * zz - xx - yy),
                C3[3] * z * (2 * zz - 3 * xx - 3 * yy), C3[4] * x * (4 * zz - xx - yy), C3[5] * z * (xx - yy),
                C3[6] * x * (xx - 3 * yy)]
    return np.array(out[:items]) if items else np.zeros(0)


def srgb_encode(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(np.maximum(c, 0.0031308), 1.0 / 2.4) - 0.055)


def srgb_decode(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, np.power((np.maximum(c, 0.04045) + 0.055) / 1.055, 2.4))


def display_color(color, sh, d, srgb=False):
    """A splat's colour seen along unit d (linear Rec.709): its linear colour
    in the cloud's encoding plus the SH bands toward d, clamped at 0, then
    decoded (SH authored in world axes)."""
    encoded = srgb_encode(color) if srgb else np.asarray(color, np.float64)
    if sh is not None and len(sh):
        encoded = encoded + sh_basis(d, len(sh)) @ np.asarray(sh, np.float64)
    encoded = np.maximum(encoded, 0.0)
    return srgb_decode(encoded) if srgb else encoded


def composite(splats, o, d, t_end, behind=(0.0, 0.0, 0.0), srgb=False):
    """The radiance (ACEScg) along the ray: every splat with its maximum in
    (0, t_end), front to back, then `behind` with what they pass on."""
    al, t = alpha(splats, o, d)
    keep = (al > 0.0) & (t > 0.0) & (t < t_end)
    idx = np.nonzero(keep)[0]
    idx = idx[np.argsort(t[idx], kind="stable")]
    a = al[idx]
    encoded = np.asarray(splats["color"], np.float64)[idx]
    if srgb:
        encoded = srgb_encode(encoded)
    sh = splats.get("sh")
    if sh is not None and sh.shape[1]:
        encoded = encoded + np.einsum("k,nkc->nc", sh_basis(d, sh.shape[1]), np.asarray(sh, np.float64)[idx])
    encoded = np.maximum(encoded, 0.0)
    colors = (srgb_decode(encoded) if srgb else encoded) @ REC709_TO_ACESCG.T
    before = np.concatenate([[1.0], np.cumprod(1.0 - a)])  # transmittance in front of each
    total = (colors * (before[:-1] * a)[:, None]).sum(0)
    kept = before[-1]
    return total + np.asarray(behind, np.float64) * kept, kept


def transmittance(splats, o, d, t_lo, t_hi):
    al, t = alpha(splats, o, d)
    keep = (al > 0.0) & (t > t_lo) & (t < t_hi)
    return float(np.prod(1.0 - al[keep]))
