"""Host colour science (src/render/spectrum.lucb:26-187): the XYZ_E <-> ACEScg
film matrix (Bradford), the Wyman–Sloan–Shirley colour-matching fit and its
integrals, the Jakob–Hanika sigmoid fit of an ACEScg colour by damped
Gauss–Newton, and the peak-normalized blackbody. f64 throughout."""
import math

import numpy as np

LAMBDA_MIN = 360.0
LAMBDA_MAX = 830.0

ACESCG_TO_XYZ = np.array([[0.6624541811, 0.1340042065, 0.1561876870],
                          [0.2722287168, 0.6740817658, 0.0536895174],
                          [-0.0055746495, 0.0040607335, 1.0103391003]])
BRADFORD = np.array([[0.8951, 0.2664, -0.1614], [-0.7502, 1.7135, 0.0367], [0.0389, -0.0685, 1.0296]])


def _inverse3(a):
    """Matrix3.inverse (spectrum.lucb:47-55): cofactors over the determinant."""
    a = np.asarray(a, np.float64).reshape(9)
    c0 = a[4] * a[8] - a[5] * a[7]
    c1 = a[5] * a[6] - a[3] * a[8]
    c2 = a[3] * a[7] - a[4] * a[6]
    d = 1.0 / (a[0] * c0 + a[1] * c1 + a[2] * c2)
    return np.array([[c0 * d, (a[2] * a[7] - a[1] * a[8]) * d, (a[1] * a[5] - a[2] * a[4]) * d],
                     [c1 * d, (a[0] * a[8] - a[2] * a[6]) * d, (a[2] * a[3] - a[0] * a[5]) * d],
                     [c2 * d, (a[1] * a[6] - a[0] * a[7]) * d, (a[0] * a[4] - a[1] * a[3]) * d]])


def bradford(source, target):
    s = BRADFORD @ np.asarray(source, np.float64)
    t = BRADFORD @ np.asarray(target, np.float64)
    return _inverse3(BRADFORD) @ np.diag(t / s) @ BRADFORD


def acescg_to_xyz_e():
    """ACEScg to XYZ relative to white E: what spectral values mean."""
    return bradford(ACESCG_TO_XYZ @ np.ones(3), np.ones(3)) @ ACESCG_TO_XYZ


def film_matrix():
    """XYZ under E to ACEScg."""
    return _inverse3(acescg_to_xyz_e())


def _lobe(x, mean, below, above):
    t = (x - mean) / np.where(x < mean, below, above)
    return np.exp(-0.5 * t * t)


def cmf(lam):
    """The CIE 1931 2-degree observer by Wyman et al.'s fit (unnormalized)."""
    lam = np.asarray(lam, np.float64)
    x = 1.056 * _lobe(lam, 599.8, 37.9, 31.0) + 0.362 * _lobe(lam, 442.0, 16.0, 26.7) - 0.065 * _lobe(lam, 501.1, 20.4, 26.2)
    y = 0.821 * _lobe(lam, 568.8, 46.9, 40.5) + 0.286 * _lobe(lam, 530.9, 16.3, 31.1)
    z = 1.217 * _lobe(lam, 437.0, 11.8, 36.0) + 0.681 * _lobe(lam, 459.0, 26.0, 13.8)
    return np.stack([x, y, z], axis=-1)


def cmf_integrals():
    """Each channel summed over 360..830 nm at 1 nm (spectrum.lucb:94-100)."""
    return cmf(np.arange(LAMBDA_MIN, LAMBDA_MAX + 0.5, 1.0)).sum(axis=0)


def sigmoid(z):
    return 0.5 + z / (2.0 * np.sqrt(1.0 + z * z))


class SpectrumFit:
    def __init__(self, c0, c1, c2, scale):
        self.c0, self.c1, self.c2, self.scale = c0, c1, c2, scale

    def at(self, lam):
        x = (np.asarray(lam, np.float64) - LAMBDA_MIN) / (LAMBDA_MAX - LAMBDA_MIN)
        return self.scale * sigmoid((self.c0 * x + self.c1) * x + self.c2)

    def row(self):
        return [self.c0, self.c1, self.c2, self.scale]


_FIT_LAMBDA = LAMBDA_MIN + 5.0 * np.arange(95)
_FIT_X = (_FIT_LAMBDA - LAMBDA_MIN) / (LAMBDA_MAX - LAMBDA_MIN)
_FIT_OBSERVER = None
_FIT_CACHE = {}


def _fit_xyz(k, integrals):
    s = sigmoid((k[0] * _FIT_X + k[1]) * _FIT_X + k[2])
    return (_FIT_OBSERVER * (s * 5.0)[:, None]).sum(axis=0) / integrals


def fit_spectrum(rgb):
    """fit_spectrum (spectrum.lucb:136-168): greys exact, others fitted to
    their XYZ under E at half their largest component."""
    global _FIT_OBSERVER
    rgb = tuple(float(c) for c in rgb)
    if rgb in _FIT_CACHE:
        return _FIT_CACHE[rgb]
    largest = max(rgb)
    if largest <= 0.0:
        fit = SpectrumFit(0.0, 0.0, 0.0, 0.0)
    elif abs(rgb[0] - rgb[1]) < 1e-6 * largest and abs(rgb[1] - rgb[2]) < 1e-6 * largest:
        fit = SpectrumFit(0.0, 0.0, 0.0, 2.0 * largest)
    else:
        if _FIT_OBSERVER is None:
            _FIT_OBSERVER = cmf(_FIT_LAMBDA)
        target = acescg_to_xyz_e() @ (np.array(rgb) * (0.5 / largest))
        integrals = cmf_integrals()
        k = np.zeros(3)
        for _ in range(40):
            r = _fit_xyz(k, integrals) - target
            if np.linalg.norm(r) < 1e-7:
                break
            h = 1e-4
            cols = []
            for i in range(3):
                kk = k.copy()
                kk[i] += h
                cols.append((_fit_xyz(kk, integrals) - target - r) / h)
            j = np.stack(cols, axis=1)
            step = _inverse3(j) @ r
            if not np.all(np.isfinite(step)):
                break
            size = np.linalg.norm(step)
            k = k - step * (20.0 / size if size > 20.0 else 1.0)
        fit = SpectrumFit(k[0], k[1], k[2], 2.0 * largest)
    _FIT_CACHE[rgb] = fit
    return fit


def blackbody(lam, kelvin):
    peak = 2.8977721e6 / kelvin
    x = 1.4387769e7 / (np.asarray(lam, np.float64) * kelvin)
    r = peak / lam
    return r ** 5 * ((math.exp(4.965114) - 1.0) / (np.exp(x) - 1.0))


def blackbody_xyz(kelvin):
    lam = np.arange(LAMBDA_MIN, LAMBDA_MAX + 0.5, 1.0)
    return (cmf(lam) * blackbody(lam, kelvin)[:, None]).sum(axis=0) / cmf_integrals()


def luminance(rgb):
    """ACEScg luminance (the Y row of ACEScg to XYZ)."""
    return 0.2722287168 * rgb[0] + 0.6740817658 * rgb[1] + 0.0536895174 * rgb[2]


def basis_weights():
    """Environment texels' basis (environment.lucb): ACEScg to the weights of
    three smooth spectra (blue, green, red logistic bands) that sum to 1."""
    lam = np.arange(LAMBDA_MIN, LAMBDA_MAX + 0.5, 1.0)
    blue = 1.0 - 1.0 / (1.0 + np.exp(-(lam - 490.0) / 12.0))
    red = 1.0 / (1.0 + np.exp(-(lam - 595.0) / 12.0))
    green = 1.0 - blue - red
    c = cmf(lam) / cmf_integrals()
    # XYZ_E of each basis spectrum, then ACEScg; weights are the inverse
    to_acescg = film_matrix()
    cols = [to_acescg @ (c * s[:, None]).sum(axis=0) for s in (blue, green, red)]
    return np.linalg.inv(np.stack(cols, axis=1))
