/*
 * Oracle (test infrastructure only): an f64 CPU restatement of luce-render's
 * camera-path integrator, one path at a time, drawing the same Owen-scrambled
 * Sobol numbers as the kernels (integer-exact), so a GPU render and this one
 * agree pixel by pixel up to f32 rounding and the rare decision it flips.
 * Nothing in the product links or calls this file.
 *
 * Follows, in f64:
 *   sampler          shaders/sampler.glsl:8-55, sobol_table.glsl
 *   wavelengths/film shaders/spectrum.glsl:10-118, finish.comp:17-27,
 *                    film_convert.comp:36-53
 *   camera           shaders/camera.comp:14-50
 *   closest + lights shaders/intersect_closest.comp:45-87, lights.glsl
 *   surface shading  shaders/shade_surface.comp:48-312, with the camera
 *                    side of light tracing's MIS and the guide's mixture
 *   light tracing    light_emit.comp:21-137, light_paths.glsl:15-189,
 *                    connect.glsl:30-151 (every light path of every sample,
 *                    splatted in path order)
 *   guide            guide.glsl:39-139, guide_fit.comp:16-35 (sample by
 *                    sample over every pixel, trained in pixel order)
 *   light tree       shaders/lighttree.glsl:18-137
 *   OpenPBR / GGX    shaders/openpbr.glsl:100-556, microfacet.glsl:9-84;
 *                    albedo tables by tables.comp's estimator, in f64
 *   miss             shaders/shade_miss.comp:12-51
 *   shadow           shaders/intersect_shadow.comp:28-97
 *   node graphs      shaders/shader_program.glsl:1-349 (the interpreter),
 *                    openpbr.glsl:41-70, 264-287, 391-401 (uvs, textures,
 *                    the parameters a program drives)
 * Ray casts use the scene's binary BVH with an f64 slab test and f64
 * Moller-Trumbore (the brute force's arithmetic, tests/gpu/main.lucb:105-119).
 *
 * The packed arrays (lights, constants, materials, emitters) are the ones the
 * renderer receives; the oracle reads them as the kernels do.
 *
 * Pinning: the furnace closed form (rho L), the rect-light irradiance closed
 * form (rho/pi L x the corner-irradiance of a rectangle, also reached by light
 * paths), f64 brute-force hit agreement of the caster; light tracing on and
 * off giving one image mean (the three techniques' weights sum to one; a
 * planted mistake in them, or_plant, breaks it); the vMF lobe's normalisation,
 * its sampler's mean resultant length coth k - 1/k and the Banerjee fit of
 * samples of a known lobe; and the GPU images themselves in the tests.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define PI 3.14159265358979323846
#define INF 3.0e38
#define NONE 0xffffffffu

typedef struct { double x, y, z; } v3;
static v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 mul(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
static double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 cross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static double len(v3 a) { return sqrt(dot(a, a)); }
static v3 norm(v3 a) { return mul(a, 1.0 / len(a)); }
static double clampd(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }
static double mixd(double a, double b, double t) { return a * (1.0 - t) + b * t; }
static double fract(double x) { return x - floor(x); }

typedef struct { double v[4]; } spec;
static spec S1(double s) { spec r = {{s, s, s, s}}; return r; }
static spec sadd(spec a, spec b) { spec r; for (int i = 0; i < 4; i++) r.v[i] = a.v[i] + b.v[i]; return r; }
static spec smul(spec a, spec b) { spec r; for (int i = 0; i < 4; i++) r.v[i] = a.v[i] * b.v[i]; return r; }
static spec sscale(spec a, double s) { spec r; for (int i = 0; i < 4; i++) r.v[i] = a.v[i] * s; return r; }

/* ------------------------------------------------------------- scene */

typedef struct { float lo[3]; uint32_t first; float hi[3]; uint32_t count; } onode;

typedef struct {
    const float *positions; const uint32_t *indices; const uint32_t *shading; size_t triangles;
    const onode *nodes; size_t node_count;
    const float *lights; const float *constants; const float *materials; const uint32_t *emitters;
    const double *frame;     /* 16 floats of the packed frame, as f64 */
    uint32_t width, height, samples, max_bounces, spectral, seed, through;
    double *tables;          /* f64 albedo tables */
    /* node graphs: leaf-order uvs (6 a triangle), program words, RGBA texels
     * and per texture (first texel, width, height, 0) */
    const float *uvs; const uint32_t *programs; size_t program_words;
    const float *textures; const uint32_t *texture_table; size_t texture_count;
    struct oguide *guide;    /* the hash grid of vMF mixtures (guiding 1), or NULL */
} oscene;

static v3 point_of(const oscene *s, uint32_t i) {
    return V(s->positions[(size_t)i * 3], s->positions[(size_t)i * 3 + 1], s->positions[(size_t)i * 3 + 2]);
}
static double K(const oscene *s, int row, int c) { return s->constants[row * 4 + c]; }
static v3 Kv(const oscene *s, int row) { return V(K(s, row, 0), K(s, row, 1), K(s, row, 2)); }
static double L(const oscene *s, uint32_t light, int row, int c) { return s->lights[(light * 7 + row) * 4 + c]; }
static v3 Lv(const oscene *s, uint32_t light, int row) { return V(L(s, light, row, 0), L(s, light, row, 1), L(s, light, row, 2)); }
static double M(const oscene *s, uint32_t m, int row, int c) { return s->materials[(m * 16 + row) * 4 + c]; }

/* ------------------------------------------------------------ sampler */

static const uint32_t SOBOL[128] = {
    0x80000000u, 0x40000000u, 0x20000000u, 0x10000000u, 0x08000000u, 0x04000000u, 0x02000000u, 0x01000000u,
    0x00800000u, 0x00400000u, 0x00200000u, 0x00100000u, 0x00080000u, 0x00040000u, 0x00020000u, 0x00010000u,
    0x00008000u, 0x00004000u, 0x00002000u, 0x00001000u, 0x00000800u, 0x00000400u, 0x00000200u, 0x00000100u,
    0x00000080u, 0x00000040u, 0x00000020u, 0x00000010u, 0x00000008u, 0x00000004u, 0x00000002u, 0x00000001u,
    0x80000000u, 0xc0000000u, 0xa0000000u, 0xf0000000u, 0x88000000u, 0xcc000000u, 0xaa000000u, 0xff000000u,
    0x80800000u, 0xc0c00000u, 0xa0a00000u, 0xf0f00000u, 0x88880000u, 0xcccc0000u, 0xaaaa0000u, 0xffff0000u,
    0x80008000u, 0xc000c000u, 0xa000a000u, 0xf000f000u, 0x88008800u, 0xcc00cc00u, 0xaa00aa00u, 0xff00ff00u,
    0x80808080u, 0xc0c0c0c0u, 0xa0a0a0a0u, 0xf0f0f0f0u, 0x88888888u, 0xccccccccu, 0xaaaaaaaau, 0xffffffffu,
    0x80000000u, 0xc0000000u, 0x60000000u, 0x90000000u, 0xe8000000u, 0x5c000000u, 0x8e000000u, 0xc5000000u,
    0x68800000u, 0x9cc00000u, 0xee600000u, 0x55900000u, 0x80680000u, 0xc09c0000u, 0x60ee0000u, 0x90550000u,
    0xe8808000u, 0x5cc0c000u, 0x8e606000u, 0xc5909000u, 0x6868e800u, 0x9c9c5c00u, 0xeeee8e00u, 0x5555c500u,
    0x8000e880u, 0xc0005cc0u, 0x60008e60u, 0x9000c590u, 0xe8006868u, 0x5c009c9cu, 0x8e00eeeeu, 0xc5005555u,
    0x80000000u, 0xc0000000u, 0x20000000u, 0x50000000u, 0xf8000000u, 0x74000000u, 0xa2000000u, 0x93000000u,
    0xd8800000u, 0x25400000u, 0x59e00000u, 0xe6d00000u, 0x78080000u, 0xb40c0000u, 0x82020000u, 0xc3050000u,
    0x208f8000u, 0x51474000u, 0xfbea2000u, 0x75d93000u, 0xa0858800u, 0x914e5400u, 0xdbe79e00u, 0x25db6d00u,
    0x58800080u, 0xe54000c0u, 0x79e00020u, 0xb6d00050u, 0x800800f8u, 0xc00c0074u, 0x200200a2u, 0x50050093u};

uint32_t or_hash_u32(uint32_t x) {
    x ^= x >> 16; x *= 0x21f0aaadu; x ^= x >> 15; x *= 0x735a2d97u; x ^= x >> 15;
    return x;
}
static uint32_t hash_combine(uint32_t seed, uint32_t v) { return seed ^ (or_hash_u32(v) + 0x9e3779b9u + (seed << 6) + (seed >> 2)); }
static uint32_t rev(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
}
static uint32_t lk(uint32_t x, uint32_t seed) {
    x += seed; x ^= x * 0x6c50b47cu; x ^= x * 0xb82f1e52u; x ^= x * 0xc7afe638u; x ^= x * 0x8d22f6e6u;
    return x;
}
static uint32_t nus(uint32_t x, uint32_t seed) { return rev(lk(rev(x), seed)); }
static uint32_t sobol(uint32_t index, uint32_t d) {
    uint32_t x = 0;
    for (uint32_t bit = 0; index; bit++, index >>= 1) if (index & 1u) x ^= SOBOL[d * 32 + bit];
    return x;
}
/* sampler.glsl:47-55; the four 24-bit integers and their values. */
void or_sample4(uint32_t seed0, uint32_t pixel, uint32_t index, uint32_t group, uint32_t bits[4], double out[4]) {
    uint32_t seed = hash_combine(hash_combine(seed0, pixel), group);
    uint32_t shuffled = nus(index, or_hash_u32(seed));
    for (uint32_t d = 0; d < 4; d++) {
        uint32_t b = nus(sobol(shuffled, d), hash_combine(seed, d + 1)) >> 8;
        if (bits) bits[d] = b;
        if (out) out[d] = (double)b * (1.0 / 16777216.0);
    }
}
static void sample4(const oscene *s, uint32_t pixel, uint32_t index, uint32_t group, double u[4]) {
    or_sample4(s->seed, pixel, index, group, NULL, u);
}

/* ------------------------------------------------------------ spectrum */

static spec sample_wavelengths(double u) {
    spec l;
    for (int i = 0; i < 4; i++) {
        double v = fract(u + i * 0.25);
        l.v[i] = 538.0 - 138.888889 * atanh(0.85691062 - 1.82750197 * v);
    }
    return l;
}
static double wavelength_pdf(double l) {
    if (l < 360.0 || l > 830.0) return 0.0;
    double c = cosh(0.0072 * (l - 538.0));
    return 0.0039398042 / (c * c);
}
static double lobe(double x, double m, double b, double a) { double t = (x - m) / (x < m ? b : a); return exp(-0.5 * t * t); }
static v3 cmf(double l) {
    return V(1.056 * lobe(l, 599.8, 37.9, 31.0) + 0.362 * lobe(l, 442.0, 16.0, 26.7) - 0.065 * lobe(l, 501.1, 20.4, 26.2),
             0.821 * lobe(l, 568.8, 46.9, 40.5) + 0.286 * lobe(l, 530.9, 16.3, 31.1),
             1.217 * lobe(l, 437.0, 11.8, 36.0) + 0.681 * lobe(l, 459.0, 26.0, 13.8));
}
static spec spec_of(const oscene *s, const float *fit, const float *rgb, spec lambda) {
    spec r = S1(0.0);
    if (s->spectral) {
        for (int i = 0; i < 4; i++) {
            double x = (lambda.v[i] - 360.0) / 470.0;
            double z = ((double)fit[0] * x + fit[1]) * x + fit[2];
            r.v[i] = fit[3] * (0.5 + z / (2.0 * sqrt(1.0 + z * z)));
        }
    } else {
        r.v[0] = rgb[0]; r.v[1] = rgb[1]; r.v[2] = rgb[2]; r.v[3] = 0.0;
    }
    return r;
}
static double blackbody(double l, double k) {
    double peak = 2.8977721e6 / k, x = 1.4387769e7 / (l * k), r = peak / l;
    return r * r * r * r * r * (142.32492 / (exp(x) - 1.0));
}
static double spec_max(const oscene *s, spec a) {
    if (s->spectral) return fmax(fmax(a.v[0], a.v[1]), fmax(a.v[2], a.v[3]));
    return fmax(a.v[0], fmax(a.v[1], a.v[2]));
}
static v3 spec_to_film(const oscene *s, spec r, spec lambda) {
    if (!s->spectral) return V(r.v[0], r.v[1], r.v[2]);
    v3 sum = V(0, 0, 0);
    for (int i = 0; i < 4; i++) {
        double pdf = wavelength_pdf(lambda.v[i]);
        if (pdf > 0.0) sum = add(sum, mul(cmf(lambda.v[i]), r.v[i] / pdf));
    }
    return V(sum.x * 0.25 / K(s, 6, 0), sum.y * 0.25 / K(s, 6, 1), sum.z * 0.25 / K(s, 6, 2));
}

/* ------------------------------------------------------------ ray casts */

typedef struct { double t; uint32_t tri; double u, v; } ohit;

static int box64(const onode *n, v3 o, v3 inv, double tmax) {
    double t0 = 0.0, t1 = tmax;
    double lo[3] = {n->lo[0], n->lo[1], n->lo[2]}, hi[3] = {n->hi[0], n->hi[1], n->hi[2]};
    double oo[3] = {o.x, o.y, o.z}, ii[3] = {inv.x, inv.y, inv.z};
    for (int k = 0; k < 3; k++) {
        double a = (lo[k] - oo[k]) * ii[k], b = (hi[k] - oo[k]) * ii[k];
        double nn = a < b ? a : b, ff = a < b ? b : a;
        if (nn > t0) t0 = nn;
        if (ff < t1) t1 = ff;
    }
    /* a little slack: f64 boxes of f32 geometry must not cull a true hit */
    return t0 <= t1 * (1.0 + 1e-9) + 1e-12;
}

static int tri64(const oscene *s, uint32_t t, v3 o, v3 d, double tmax, ohit *h) {
    v3 a = point_of(s, s->indices[t * 3]);
    v3 e1 = sub(point_of(s, s->indices[t * 3 + 1]), a), e2 = sub(point_of(s, s->indices[t * 3 + 2]), a);
    v3 q = cross(d, e2);
    double det = dot(e1, q);
    if (fabs(det) < 1e-30) return 0;
    v3 sv = sub(o, a);
    double u = dot(sv, q) / det;
    if (u < 0.0 || u > 1.0) return 0;
    v3 r = cross(sv, e1);
    double v = dot(d, r) / det;
    if (v < 0.0 || u + v > 1.0) return 0;
    double tt = dot(e2, r) / det;
    if (tt <= 0.0 || tt >= tmax || tt >= h->t) return 0;
    h->t = tt; h->tri = t; h->u = u; h->v = v;
    return 1;
}

static ohit closest(const oscene *s, v3 o, v3 d, double tmax) {
    ohit h = {tmax, NONE, 0, 0};
    if (s->triangles == 0 || s->nodes[0].lo[0] > s->nodes[0].hi[0]) return h;
    v3 inv = V(1.0 / (fabs(d.x) < 1e-20 ? (d.x < 0 ? -1e-20 : 1e-20) : d.x),
               1.0 / (fabs(d.y) < 1e-20 ? (d.y < 0 ? -1e-20 : 1e-20) : d.y),
               1.0 / (fabs(d.z) < 1e-20 ? (d.z < 0 ? -1e-20 : 1e-20) : d.z));
    uint32_t stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const onode *n = &s->nodes[stack[--sp]];
        if (!box64(n, o, inv, h.t)) continue;
        if (n->count) {
            for (uint32_t i = 0; i < n->count; i++) tri64(s, n->first + i, o, d, tmax, &h);
        } else if (sp + 2 <= 256) {
            stack[sp++] = n->first + 1;
            stack[sp++] = n->first;
        }
    }
    return h;
}

/* ------------------------------------------------------------ lights */

static uint32_t light_count(const oscene *s) { return (uint32_t)K(s, 0, 0); }
static uint32_t light_type(const oscene *s, uint32_t l) { return (uint32_t)L(s, l, 0, 3); }
static int light_visible(const oscene *s, uint32_t l) { return ((uint32_t)L(s, l, 5, 3) & 1u) != 0; }
static double light_select_pdf(const oscene *s, uint32_t l) { return K(s, 8, 2) * L(s, l, 6, 1); }

static spec light_radiance(const oscene *s, uint32_t l, spec lambda) {
    spec r = spec_of(s, &s->lights[(l * 7 + 4) * 4], &s->lights[(l * 7 + 5) * 4], lambda);
    if (s->spectral && ((uint32_t)L(s, l, 5, 3) & 2u)) {
        double k = L(s, l, 6, 0);
        for (int i = 0; i < 4; i++) r.v[i] *= blackbody(lambda.v[i], k);
    }
    return r;
}

/* Dome lights (lights.glsl:89-175, environment.lucb): the one textured dome
 * (constants row 16: light, texels' first float4, width, height; row 17: the
 * cumulative weights' first float, the weights' sum) looks its radiance up in
 * a lat-long image and samples it by luminance x sin theta; others are even. */
static int dome_textured(const oscene *s, uint32_t l) { return (double)l == K(s, 16, 0); }
static uint32_t dome_texel(const oscene *s, uint32_t l, v3 d, double *sin_theta) {
    v3 lv = V(dot(d, Lv(s, l, 1)), dot(d, Lv(s, l, 2)), dot(d, Lv(s, l, 3)));
    uint32_t w = (uint32_t)K(s, 16, 2), h = (uint32_t)K(s, 16, 3);
    double u = 0.5 + atan2(lv.x, -lv.z) * (0.5 / PI);
    double v = acos(fmin(1.0, fmax(-1.0, lv.y))) / PI;
    *sin_theta = sqrt(fmax(0.0, 1.0 - lv.y * lv.y));
    uint32_t x = (uint32_t)(u * w), y = (uint32_t)(v * h);
    if (x > w - 1) x = w - 1;
    if (y > h - 1) y = h - 1;
    return y * w + x;
}
static const float *dome_texel_value(const oscene *s, uint32_t texel) {
    return &s->lights[((uint32_t)K(s, 16, 1) + texel) * 4];
}
static double dome_pdf(const oscene *s, uint32_t l, v3 d) {
    if (!dome_textured(s, l)) return 1.0 / (4.0 * PI);
    double st;
    uint32_t texel = dome_texel(s, l, d, &st);
    if (st <= 0.0) return 0.0;
    return dome_texel_value(s, texel)[3] / K(s, 17, 1) * K(s, 16, 2) * K(s, 16, 3) / (2.0 * PI * PI * st);
}
static double logistic(double x) { return 1.0 / (1.0 + exp(-x)); }
/* spectrum.glsl:67-79: a texel's ACEScg as the three basis spectra's sum. */
static spec spec_of_rgb(const oscene *s, const double *rgb, spec lambda) {
    spec r = S1(0.0);
    if (s->spectral) {
        double w[3];
        for (int k = 0; k < 3; k++) w[k] = K(s, 18 + k, 0) * rgb[0] + K(s, 18 + k, 1) * rgb[1] + K(s, 18 + k, 2) * rgb[2];
        for (int i = 0; i < 4; i++) {
            double blue = 1.0 - logistic((lambda.v[i] - 490.0) / 12.0), red = logistic((lambda.v[i] - 595.0) / 12.0);
            r.v[i] = w[0] * blue + w[1] * (1.0 - blue - red) + w[2] * red;
        }
    } else {
        r.v[0] = rgb[0]; r.v[1] = rgb[1]; r.v[2] = rgb[2];
    }
    return r;
}
static spec spec_of_texel(const oscene *s, const float *rgb, spec lambda) {
    double d[3] = {rgb[0], rgb[1], rgb[2]};
    return spec_of_rgb(s, d, lambda);
}
static spec dome_radiance(const oscene *s, uint32_t l, v3 d, spec lambda) {
    spec r = light_radiance(s, l, lambda);
    if (!dome_textured(s, l)) return r;
    double st;
    spec t = spec_of_texel(s, dome_texel_value(s, dome_texel(s, l, d, &st)), lambda);
    for (int i = 0; i < 4; i++) r.v[i] *= t.v[i];
    return r;
}
static double dome_cdf(const oscene *s, uint32_t at) { return s->lights[(uint32_t)K(s, 17, 0) + at]; }
static uint32_t dome_search(const oscene *s, uint32_t first, uint32_t count, double u, double *within) {
    uint32_t low = 0, high = count - 1;
    while (low < high) {
        uint32_t middle = (low + high) / 2;
        if (dome_cdf(s, first + middle) > u) high = middle;
        else low = middle + 1;
    }
    double below = low > 0 ? dome_cdf(s, first + low - 1) : 0.0, above = dome_cdf(s, first + low);
    *within = above > below ? fmin(fmax((u - below) / (above - below), 0.0), 0.99999994) : 0.5;
    return low;
}

static uint32_t pick_distant(const oscene *s, double u, double *pdf) {
    uint32_t n = light_count(s), chosen = NONE;
    for (uint32_t l = 0; l < n; l++) {
        if (light_type(s, l) < 3) continue;
        chosen = l;
        if (u < L(s, l, 6, 2)) break;
    }
    *pdf = chosen == NONE ? 0.0 : L(s, chosen, 6, 1);
    return chosen;
}

typedef struct { v3 direction; double distance, pdf, cos_light; int delta; } lsample;

static void basis(v3 n, v3 *t, v3 *b) {
    double sg = n.z >= 0.0 ? 1.0 : -1.0;
    double a = -1.0 / (sg + n.z), c = n.x * n.y * a;
    *t = V(1.0 + sg * n.x * n.x * a, sg * c, -sg * n.x);
    *b = V(c, sg + n.y * n.y * a, -n.y);
}
static void concentric(double ux, double uy, double *x, double *y) {
    double ox = 2 * ux - 1, oy = 2 * uy - 1;
    if (ox == 0 && oy == 0) { *x = 0; *y = 0; return; }
    double r, th;
    if (fabs(ox) > fabs(oy)) { r = ox; th = (PI / 4) * (oy / ox); }
    else { r = oy; th = (PI / 2) - (PI / 4) * (ox / oy); }
    *x = r * cos(th); *y = r * sin(th);
}
static v3 sample_cone(v3 axis, double cos_max, double ux, double uy) {
    double c = 1.0 - ux * (1.0 - cos_max), sn = sqrt(fmax(0.0, 1.0 - c * c)), phi = 2 * PI * uy;
    v3 t, b;
    basis(axis, &t, &b);
    return norm(add(add(mul(t, sn * cos(phi)), mul(b, sn * sin(phi))), mul(axis, c)));
}
static double sphere_hit(v3 c, double r, v3 o, v3 d) {
    v3 oc = sub(o, c);
    double b = dot(oc, d), cc = dot(oc, oc) - r * r, h = b * b - cc;
    if (h < 0) return INF;
    h = sqrt(h);
    double t = -b - h;
    if (t > 0) return t;
    t = -b + h;
    return t > 0 ? t : INF;
}

static int sample_light(const oscene *s, uint32_t l, v3 x, double ux, double uy, lsample *o) {
    uint32_t type = light_type(s, l);
    o->delta = 0;
    o->cos_light = 1.0;
    if (type == 4) {
        v3 lv;
        if (!dome_textured(s, l)) {
            double z = 1 - 2 * ux, r = sqrt(fmax(0, 1 - z * z));
            lv = V(r * cos(2 * PI * uy), z, r * sin(2 * PI * uy));
        } else {
            uint32_t w = (uint32_t)K(s, 16, 2), h = (uint32_t)K(s, 16, 3);
            double fy, fx;
            uint32_t y = dome_search(s, w * h, h, uy, &fy);
            uint32_t x = dome_search(s, y * w, w, ux, &fx);
            double theta = PI * (y + fy) / h, phi = 2 * PI * ((x + fx) / w - 0.5);
            lv = V(sin(theta) * sin(phi), cos(theta), -sin(theta) * cos(phi));
        }
        o->direction = norm(add(add(mul(Lv(s, l, 1), lv.x), mul(Lv(s, l, 2), lv.y)), mul(Lv(s, l, 3), lv.z)));
        o->distance = INF;
        o->pdf = dome_pdf(s, l, o->direction);
        return o->pdf > 0.0;
    }
    if (type == 0 || type == 1) {
        double ax, ay;
        if (type == 0) { ax = 2 * ux - 1; ay = 2 * uy - 1; } else concentric(ux, uy, &ax, &ay);
        v3 q = add(add(Lv(s, l, 0), mul(Lv(s, l, 1), ax)), mul(Lv(s, l, 2), ay));
        v3 d = sub(q, x);
        double d2 = dot(d, d);
        o->distance = sqrt(d2);
        o->direction = mul(d, 1.0 / o->distance);
        double cl = -dot(o->direction, Lv(s, l, 3));
        if (cl <= 0 || L(s, l, 3, 3) <= 0) return 0;
        o->pdf = d2 / (cl * L(s, l, 3, 3));
        o->cos_light = cl;
        return 1;
    }
    if (type == 2) {
        v3 tc = sub(Lv(s, l, 0), x);
        double d2 = dot(tc, tc), r = L(s, l, 1, 3);
        if (d2 <= r * r) return 0;
        double cm = sqrt(fmax(0, 1 - r * r / d2));
        o->direction = sample_cone(mul(tc, 1.0 / sqrt(d2)), cm, ux, uy);
        o->distance = sphere_hit(Lv(s, l, 0), r, x, o->direction);
        if (o->distance >= INF) return 0;
        o->pdf = 1.0 / (2 * PI * (1 - cm));
        o->cos_light = fmax(0, -dot(o->direction, norm(sub(add(x, mul(o->direction, o->distance)), Lv(s, l, 0)))));
        return 1;
    }
    v3 axis = neg(Lv(s, l, 3));
    double ch = L(s, l, 2, 3);
    o->distance = INF;
    if (ch >= 1.0 - 1e-7) { o->direction = axis; o->pdf = 1; o->delta = 1; return 1; }
    o->direction = sample_cone(axis, ch, ux, uy);
    o->pdf = 1.0 / (2 * PI * (1 - ch));
    return 1;
}

static double intersect_light(const oscene *s, uint32_t l, v3 o, v3 d, double tmax, int flow) {
    uint32_t type = light_type(s, l);
    if (type == 2) { double t = sphere_hit(Lv(s, l, 0), L(s, l, 1, 3), o, d); return t < tmax ? t : INF; }
    if (type >= 3) return INF;
    v3 u = Lv(s, l, 1), v = Lv(s, l, 2), n = Lv(s, l, 3);
    double f = dot(d, n);
    if (flow ? f <= 0 : f >= 0) return INF;
    double t = dot(sub(Lv(s, l, 0), o), n) / f;
    if (t <= 0 || t >= tmax) return INF;
    v3 q = sub(add(o, mul(d, t)), Lv(s, l, 0));
    double a = dot(q, u) / dot(u, u), b = dot(q, v) / dot(v, v);
    int in = type == 0 ? (fabs(a) <= 1 && fabs(b) <= 1) : (a * a + b * b <= 1);
    return in ? t : INF;
}

static double light_hit_pdf(const oscene *s, uint32_t l, v3 o, v3 d, double t) {
    uint32_t type = light_type(s, l);
    if (type == 2) {
        v3 tc = sub(Lv(s, l, 0), o);
        double r = L(s, l, 1, 3), cm = sqrt(fmax(0, 1 - r * r / dot(tc, tc)));
        return 1.0 / (2 * PI * fmax(1e-12, 1 - cm));
    }
    if (type == 3) return 1.0 / (2 * PI * fmax(1e-12, 1 - L(s, l, 2, 3)));
    if (type == 4) return dome_pdf(s, l, d);
    double cl = fmax(1e-12, -dot(d, Lv(s, l, 3)));
    return t * t / (cl * L(s, l, 3, 3));
}

/* light tree (lighttree.glsl:18-95) */
static double tree_importance(const oscene *s, uint32_t node, v3 x, v3 n) {
    uint32_t base = (uint32_t)K(s, 8, 1) + node * 3;
    const float *a = &s->lights[base * 4], *b = &s->lights[(base + 1) * 4], *c = &s->lights[(base + 2) * 4];
    if (a[3] <= 0) return 0;
    v3 center = mul(add(V(a[0], a[1], a[2]), V(b[0], b[1], b[2])), 0.5);
    v3 off = sub(x, center);
    double radius = 0.5 * len(sub(V(b[0], b[1], b[2]), V(a[0], a[1], a[2])));
    double d2o = dot(off, off), dist2 = fmax(d2o, radius);
    v3 cv = V(c[0], c[1], c[2]);
    v3 wi = d2o > 0 ? norm(off) : cv;
    double cw = dot(cv, wi), sw = sqrt(fmax(0, 1 - cw * cw));
    double cb = d2o < radius * radius ? -1.0 : sqrt(fmax(0, 1 - radius * radius / d2o)), sb = sqrt(fmax(0, 1 - cb * cb));
    double co = b[3], so = sqrt(fmax(0, 1 - co * co));
    double cx = cw > co ? 1.0 : cw * co + sw * so, sx = cw > co ? 0.0 : sw * co - cw * so;
    double cp = cx > cb ? 1.0 : cx * cb + sx * sb;
    if (cp <= c[3]) return 0;
    double imp = a[3] * cp / dist2;
    double ci = fabs(dot(wi, n)), si = sqrt(fmax(0, 1 - ci * ci));
    double cpi = ci > cb ? 1.0 : ci * cb + si * sb;
    return fmax(imp * cpi, 0.0);
}
static int tree_sample(const oscene *s, double u, v3 x, v3 n, uint32_t *emitter, double *pdf) {
    *pdf = 1.0;
    *emitter = 0;
    if (K(s, 8, 0) < 1) return 0;
    uint32_t node = 0;
    for (int depth = 0; depth < 64; depth++) {
        uint32_t link = s->emitters[node * 2];
        if (s->emitters[node * 2 + 1] & 0x80000000u) { *emitter = link; return *pdf > 0; }
        double l = tree_importance(s, link, x, n), r = tree_importance(s, link + 1, x, n);
        if (l + r <= 0) return 0;
        double ch = l / (l + r);
        if (u < ch) { u = fmin(u / ch, 0.99999994); *pdf *= ch; node = link; }
        else { u = fmin((u - ch) / (1 - ch), 0.99999994); *pdf *= 1 - ch; node = link + 1; }
    }
    return 0;
}
static double tree_pdf(const oscene *s, uint32_t emitter, v3 x, v3 n) {
    uint32_t rec = (uint32_t)K(s, 10, 0) + emitter * 4;
    uint32_t trail = s->emitters[rec + 1], depth = s->emitters[rec + 2], node = 0;
    double pdf = 1.0;
    for (uint32_t lv = 0; lv < depth; lv++) {
        uint32_t link = s->emitters[node * 2];
        double l = tree_importance(s, link, x, n), r = tree_importance(s, link + 1, x, n);
        if (l + r <= 0) return 0;
        uint32_t bit = (trail >> lv) & 1u;
        pdf *= bit == 0 ? l / (l + r) : r / (l + r);
        node = link + bit;
    }
    return pdf;
}
static uint32_t emitter_code(const oscene *s, uint32_t e) { return s->emitters[(uint32_t)K(s, 10, 0) + e * 4]; }
static uint32_t triangle_emitter(const oscene *s, uint32_t t) { return s->emitters[(uint32_t)K(s, 10, 1) + t]; }
static spec material_color(const oscene *s, uint32_t m, int row, spec lambda) {
    return spec_of(s, &s->materials[(m * 16 + row) * 4], &s->materials[(m * 16 + row + 1) * 4], lambda);
}
static int sample_triangle(const oscene *s, uint32_t t, v3 x, double ux, double uy, spec lambda, lsample *o, spec *rad) {
    v3 a = point_of(s, s->indices[t * 3]), b = point_of(s, s->indices[t * 3 + 1]), c = point_of(s, s->indices[t * 3 + 2]);
    double r = sqrt(ux);
    v3 q = add(add(mul(a, 1 - r), mul(b, r * (1 - uy))), mul(c, r * uy));
    v3 cab = cross(sub(b, a), sub(c, a));
    double area = 0.5 * len(cab);
    v3 d = sub(q, x);
    double d2 = dot(d, d);
    if (area <= 0 || d2 <= 0) return 0;
    o->distance = sqrt(d2);
    o->direction = mul(d, 1.0 / o->distance);
    double cl = fabs(dot(o->direction, cab)) / (2 * area);
    if (cl <= 0) return 0;
    o->pdf = d2 / (cl * area);
    o->cos_light = cl;
    o->delta = 0;
    *rad = material_color(s, s->shading[t * 4 + 3], 6, lambda);
    return 1;
}
static double triangle_hit_pdf(const oscene *s, uint32_t t, v3 d, double tt) {
    v3 a = point_of(s, s->indices[t * 3]);
    v3 cab = cross(sub(point_of(s, s->indices[t * 3 + 1]), a), sub(point_of(s, s->indices[t * 3 + 2]), a));
    double area = 0.5 * len(cab);
    double cl = fabs(dot(d, cab)) / fmax(1e-20, 2 * area);
    return tt * tt / fmax(1e-20, cl * area);
}

/* ------------------------------------------------------------ GGX / OpenPBR */

typedef struct { double x, y; } d2v;
static d2v ggx_alpha(double r, double an) {
    double r2 = r * r, t = r2 * sqrt(2.0 / (1.0 + (1.0 - an) * (1.0 - an)));
    d2v a = {fmax(t, 1e-4), fmax((1.0 - an) * t, 1e-4)};
    return a;
}
static double ggx_d(v3 m, d2v a) {
    if (m.z <= 0) return 0;
    v3 sv = V(m.x / a.x, m.y / a.y, m.z);
    double k = dot(sv, sv);
    return 1.0 / (PI * a.x * a.y * k * k);
}
static double ggx_lambda(v3 v, d2v a) {
    double z2 = v.z * v.z;
    if (z2 <= 0) return INF;
    double a2 = (a.x * a.x * v.x * v.x + a.y * a.y * v.y * v.y) / z2;
    return 0.5 * (-1.0 + sqrt(1.0 + a2));
}
static double ggx_g1(v3 v, d2v a) { return 1.0 / (1.0 + ggx_lambda(v, a)); }
static double ggx_g2(v3 i, v3 o, d2v a) { return 1.0 / (1.0 + ggx_lambda(i, a) + ggx_lambda(o, a)); }
static v3 ggx_sample(v3 wo, d2v a, double ux, double uy) {
    v3 v = norm(V(wo.x * a.x, wo.y * a.y, wo.z));
    double phi = 2 * PI * ux, z = (1 - uy) * (1 + v.z) - v.z, sn = sqrt(clampd(1 - z * z, 0, 1));
    v3 h = add(V(sn * cos(phi), sn * sin(phi), z), v);
    return norm(V(h.x * a.x, h.y * a.y, fmax(h.z, 1e-6)));
}
static double ggx_reflection_pdf(v3 wo, v3 h, d2v a) { return ggx_g1(wo, a) * ggx_d(h, a) / (4 * wo.z); }
static double fresnel_dielectric(double ci, double eta) {
    double c = fabs(ci), g2 = eta * eta - 1 + c * c;
    if (g2 <= 0) return 1;
    double g = sqrt(g2), a = (g - c) / (g + c), b = (c * (g + c) - 1) / (c * (g - c) + 1);
    return 0.5 * a * a * (1 + b * b);
}
static double fresnel_avg(double eta) { return log((10893.0 * eta - 1438.2) / (-774.4 * eta * eta + 10212.0 * eta + 1.0)); }
static v3 reflectv(v3 i, v3 n) { return sub(i, mul(n, 2 * dot(n, i))); }
static v3 refractv(v3 i, v3 n, double eta) {
    double d = dot(n, i), k = 1 - eta * eta * (1 - d * d);
    if (k < 0) return V(0, 0, 0);
    return sub(mul(i, eta), mul(n, eta * d + sqrt(k)));
}

#define TMU 32
#define TR 32
#define TETA 16
#define T_AVG (TMU * TR)
#define T_DIE (T_AVG + TR)
#define T_SIZE (T_DIE + TMU * TR * TETA)
static double table_mu(int i) { return mixd(0.02, 1.0, (double)i / (TMU - 1)); }
static double table_albedo(double mu, double r, double eta) {
    v3 wo = V(sqrt(1 - mu * mu), 0, mu);
    d2v a = ggx_alpha(r, 0);
    double sum = 0;
    for (int i = 0; i < 32; i++)
        for (int j = 0; j < 32; j++) {
            v3 h = ggx_sample(wo, a, (i + 0.5) / 32, (j + 0.5) / 32);
            v3 wi = reflectv(neg(wo), h);
            if (wi.z <= 0) continue;
            double f = eta > 0 ? fresnel_dielectric(dot(wo, h), eta) : 1.0;
            sum += f * ggx_g2(wi, wo, a) / ggx_g1(wo, a);
        }
    return sum / 1024.0;
}
/* tables.comp in f64 */
void or_make_tables(double *t) {
    int plane = TMU * TR;
    for (int cell = 0; cell < plane * (1 + TETA); cell++) {
        int layer = cell / plane, at = cell % plane;
        double mu = table_mu(at % TMU), r = (double)(at / TMU) / (TR - 1);
        if (layer == 0) t[at] = table_albedo(mu, r, -1);
        else {
            double z = 0.8 * (layer - 1) / (TETA - 1);
            t[T_DIE + (layer - 1) * plane + at] = table_albedo(mu, r, (1 + z * z) / (1 - z * z));
        }
    }
    for (int c = 0; c < TR; c++) {
        double sum = 0;
        for (int m = 0; m + 1 < TMU; m++) {
            double a = table_mu(m), b = table_mu(m + 1);
            sum += 0.5 * (t[c * TMU + m] * a + t[c * TMU + m + 1] * b) * (b - a);
        }
        t[T_AVG + c] = 2 * sum;
    }
}
static double table_row(const double *t, int base, double r, double mu) {
    double x = clampd((mu - 0.02) / 0.98 * (TMU - 1), 0, TMU - 1), y = clampd(r * (TR - 1), 0, TR - 1);
    int x0 = (int)x, y0 = (int)y, x1 = x0 + 1 < TMU ? x0 + 1 : TMU - 1, y1 = y0 + 1 < TR ? y0 + 1 : TR - 1;
    double fx = x - x0, fy = y - y0;
    double a = mixd(t[base + y0 * TMU + x0], t[base + y0 * TMU + x1], fx);
    double b = mixd(t[base + y1 * TMU + x0], t[base + y1 * TMU + x1], fx);
    return mixd(a, b, fy);
}
static double table_ggx(const double *t, double mu, double r) { return table_row(t, 0, r, mu); }
static double table_dielectric(const double *t, double mu, double r, double eta) {
    double z = clampd(sqrt(fabs(eta - 1) / (eta + 1)) / 0.8 * (TETA - 1), 0, TETA - 1);
    int z0 = (int)z, z1 = z0 + 1 < TETA ? z0 + 1 : TETA - 1, plane = TMU * TR;
    return mixd(table_row(t, T_DIE + z0 * plane, r, mu), table_row(t, T_DIE + z1 * plane, r, mu), z - z0);
}

#define FON_C1 (0.5 - 2.0 / (3.0 * PI))
#define FON_C2 (2.0 / 3.0 - 28.0 / (15.0 * PI))
static double fon_albedo(double mu, double r) {
    double m = 1 - mu, g = m * (0.0571085289 + m * (0.491881867 + m * (-0.332181442 + m * 0.0714429953)));
    return (1 + r * g) / (1 + FON_C1 * r);
}
static spec eon(spec rho, double r, v3 wi, v3 wo) {
    double mi = wi.z, mo = wo.z, sv = dot(wi, wo) - mi * mo;
    double sot = sv > 0 ? sv / fmax(mi, mo) : sv, a = 1.0 / (1 + FON_C1 * r);
    double eo = fon_albedo(mo, r), ei = fon_albedo(mi, r), avg = a * (1 + FON_C2 * r);
    spec out;
    for (int k = 0; k < 4; k++) {
        double single = rho.v[k] * (a * (1 + r * sot) / PI);
        double ms = rho.v[k] * rho.v[k] * avg / (1 - rho.v[k] * (1 - avg));
        out.v[k] = single + ms * (fmax(1e-7, 1 - eo) * fmax(1e-7, 1 - ei) / (PI * fmax(1e-7, 1 - avg)));
    }
    return out;
}
static spec f82(spec f0, spec tint, double mu) {
    const double bar = 1.0 / 7.0;
    double m = 1 - mu, m5 = m * m * m * m * m, b5 = pow(1 - bar, 5);
    spec o;
    for (int k = 0; k < 4; k++) {
        double sch = f0.v[k] + (1 - f0.v[k]) * m5, sb = f0.v[k] + (1 - f0.v[k]) * b5;
        double a = sb * (1 - tint.v[k]) / (bar * b5 * (1 - bar));
        o.v[k] = clampd(sch - a * (mu * m5 * m), 0, 1);
    }
    return o;
}
static double lum(double r, double g, double b) { return r * 0.2722287168 + g * 0.6740817658 + b * 0.0536895174; }

/* ------------------------------------------------------- node graphs
 * shader_program.glsl: a material's program runs once a hit and drives some
 * OpenPBR parameters (openpbr.glsl:52-60). f64 throughout; word reads stop at
 * the words' end and stack indices stay in the 64-value stack. */

#define PROGRAM_STACK 64u
#define SLOT_TAG 0x7FC10000u
typedef struct { uint32_t mask; int viewed; double values[23]; } oprog;

static int driven(const oprog *g, uint32_t bit) { return g && ((g->mask >> bit) & 1u); }
static v3 driven3(const oprog *g, uint32_t at) { return V(g->values[at], g->values[at + 1], g->values[at + 2]); }

/* openpbr.glsl:41-50 */
static void triangle_uv(const oscene *s, uint32_t tri, double b1, double b2, double uv[2]) {
    const float *t = s->uvs + (size_t)tri * 6;
    double w = 1.0 - b1 - b2;
    uv[0] = t[0] * w + t[2] * b1 + t[4] * b2;
    uv[1] = t[1] * w + t[3] * b1 + t[5] * b2;
}

/* openpbr.glsl:66-68: linear filtering, repeat addressing, over fract(uv);
 * texel centres at (i + 0.5) / size, v = 0 the first row. */
static void texture_at(const oscene *s, uint32_t slot, double u, double v, double out[4]) {
    out[0] = out[1] = out[2] = out[3] = 1.0;
    if (slot >= s->texture_count) return;
    const uint32_t *t = s->texture_table + (size_t)slot * 4;
    long w = t[1], h = t[2];
    if (!w || !h) return;
    double x = fract(u) * w - 0.5, y = fract(v) * h - 0.5;
    double fx = floor(x), fy = floor(y);
    if (!(fx >= -1.0 && fx <= (double)w)) fx = x = 0.0;
    if (!(fy >= -1.0 && fy <= (double)h)) fy = y = 0.0;
    double ax = x - fx, ay = y - fy;
    long xs[2] = {(long)fx, (long)fx + 1}, ys[2] = {(long)fy, (long)fy + 1};
    for (int k = 0; k < 2; k++) {
        xs[k] = ((xs[k] % w) + w) % w;
        ys[k] = ((ys[k] % h) + h) % h;
    }
    const float *base = s->textures + (size_t)t[0] * 4;
    for (int c = 0; c < 4; c++) {
        double a = base[((size_t)ys[0] * w + xs[0]) * 4 + c], b = base[((size_t)ys[0] * w + xs[1]) * 4 + c];
        double d = base[((size_t)ys[1] * w + xs[0]) * 4 + c], e = base[((size_t)ys[1] * w + xs[1]) * 4 + c];
        out[c] = mixd(mixd(a, b, ax), mixd(d, e, ax), ay);
    }
}

/* shader_program.glsl:27-51: Perlin's improved noise, gradients by a hash of
 * the lattice point. */
static v3 gradient_at(int32_t cx, int32_t cy, int32_t cz) {
    uint32_t h = or_hash_u32((uint32_t)cx * 73856093u ^
                             or_hash_u32((uint32_t)cy * 19349663u ^ or_hash_u32((uint32_t)cz * 83492791u))) % 12u;
    double sx = (h & 1u) ? -1.0 : 1.0, sy = (h & 2u) ? -1.0 : 1.0;
    if (h < 4u) return V(sx, sy, 0);
    if (h < 8u) return V(sx, 0, sy);
    return V(0, sx, sy);
}
static double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
double or_gradient_noise(double px, double py, double pz) {
    v3 i = V(floor(px), floor(py), floor(pz)), f = V(px - i.x, py - i.y, pz - i.z);
    double ux = fade(f.x), uy = fade(f.y), uz = fade(f.z);
    int32_t cx = (int32_t)i.x, cy = (int32_t)i.y, cz = (int32_t)i.z;
    double n[8];
    for (int k = 0; k < 8; k++) {
        int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        n[k] = dot(gradient_at(cx + dx, cy + dy, cz + dz), V(f.x - dx, f.y - dy, f.z - dz));
    }
    return mixd(mixd(mixd(n[0], n[1], ux), mixd(n[2], n[3], ux), uy), mixd(mixd(n[4], n[5], ux), mixd(n[6], n[7], ux), uy), uz);
}
/* shader_program.glsl:53-72 */
static double fractal_noise(v3 q, double detail, double roughness) {
    double octaves = clampd(detail, 0.0, 15.0);
    int whole = (int)floor(octaves);
    double part = octaves - whole, sum = 0.0, total = 0.0, amplitude = 1.0;
    for (int octave = 0; octave <= whole; octave++) {
        sum += or_gradient_noise(q.x, q.y, q.z) * amplitude;
        total += amplitude;
        amplitude *= clampd(roughness, 0.0, 1.0);
        q = mul(q, 2.0);
    }
    if (part > 0.0) {
        sum += or_gradient_noise(q.x, q.y, q.z) * amplitude * part;
        total += amplitude * part;
    }
    return clampd(0.5 + 0.5 * sum / total, 0.0, 1.0);
}
/* shader_program.glsl:74-89 */
static double math_of(uint32_t op, double a, double b) {
    switch (op) {
    case 0: return a + b;
    case 1: return a - b;
    case 2: return a * b;
    case 3: return b != 0.0 ? a / b : 0.0;
    case 4:
        if (a >= 0.0) return pow(a, b);
        if (b != floor(b)) return 0.0;
        return pow(-a, b) * (b - 2.0 * floor(b / 2.0) == 1.0 ? -1.0 : 1.0);
    case 5: return fmin(a, b);
    case 6: return fmax(a, b);
    case 7: return fabs(a);
    case 8: return sin(a);
    case 9: return cos(a);
    case 10: return floor(a);
    default: return fract(a);
    }
}
/* shader_program.glsl:142-149: Duff et al.'s frame */
static void bump_tangents(v3 n, v3 *t, v3 *b) {
    double sg = n.z >= 0.0 ? 1.0 : -1.0, a = -1.0 / (sg + n.z), c = n.x * n.y * a;
    *t = V(1.0 + sg * n.x * n.x * a, sg * c, -sg * n.x);
    *b = V(c, sg + n.y * n.y * a, -n.y);
}
/* shader_program.glsl:121-130 */
static double bump_width(const oscene *s, v3 x, double pixels) {
    const double *f = s->frame;
    double pixel = len(V(f[8], f[9], f[10])) * (f[3] > 0.5 ? 1.0 : len(sub(x, V(f[0], f[1], f[2]))));
    return fmax(pixel * fmax(pixels, 0.0), 4e-6 * (1.0 + fmax(fabs(x.x), fmax(fabs(x.y), fabs(x.z)))));
}
static void tri_edges(const oscene *s, uint32_t tri, v3 *e1, v3 *e2) {
    v3 p0 = point_of(s, s->indices[tri * 3]);
    *e1 = sub(point_of(s, s->indices[tri * 3 + 1]), p0);
    *e2 = sub(point_of(s, s->indices[tri * 3 + 2]), p0);
}
/* shader_program.glsl:160-174 */
static void uv_step(const oscene *s, uint32_t tri, v3 step, double out[2]) {
    v3 e1, e2;
    tri_edges(s, tri, &e1, &e2);
    double g11 = dot(e1, e1), g12 = dot(e1, e2), g22 = dot(e2, e2), det = g11 * g22 - g12 * g12;
    out[0] = out[1] = 0.0;
    if (fabs(det) < 1e-24) return;
    double a = (g22 * dot(e1, step) - g12 * dot(e2, step)) / det, b = (g11 * dot(e2, step) - g12 * dot(e1, step)) / det;
    const float *t = s->uvs + (size_t)tri * 6;
    out[0] = (t[2] - t[0]) * a + (t[4] - t[0]) * b;
    out[1] = (t[3] - t[1]) * a + (t[5] - t[1]) * b;
}
/* shader_program.glsl:91-115 */
static v3 normal_from_map(const oscene *s, uint32_t tri, v3 n, v3 color, double strength) {
    v3 e1, e2;
    tri_edges(s, tri, &e1, &e2);
    const float *q = s->uvs + (size_t)tri * 6;
    double d1x = q[2] - q[0], d1y = q[3] - q[1], d2x = q[4] - q[0], d2y = q[5] - q[1];
    double det = d1x * d2y - d2x * d1y;
    if (fabs(det) < 1e-12) return n;
    v3 t = mul(sub(mul(e1, d2y), mul(e2, d1y)), 1.0 / det), b = mul(sub(mul(e2, d1x), mul(e1, d2x)), 1.0 / det);
    t = sub(t, mul(n, dot(n, t)));
    if (dot(t, t) < 1e-20) return n;
    t = norm(t);
    v3 bt = cross(n, t);
    if (dot(bt, b) < 0.0) bt = neg(bt);
    v3 c = V(color.x * 2 - 1, color.y * 2 - 1, color.z * 2 - 1);
    v3 bent = add(add(mul(t, c.x), mul(bt, c.y)), mul(n, c.z));
    if (dot(bent, bent) < 1e-12) return n;
    double k = clampd(strength, 0.0, 1.0);
    return norm(add(mul(n, 1 - k), mul(norm(bent), k)));
}
/* shader_program.glsl:176-190 (Cycles' svm_node_set_bump) */
static v3 bumped(const oscene *s, v3 x, v3 n, v3 normal, double strength, double dist, double pixels, double hc, double hx,
                 double hy) {
    v3 base = dot(normal, normal) > 1e-12 ? norm(normal) : n, t, b;
    bump_tangents(n, &t, &b);
    double w = bump_width(s, x, pixels);
    v3 dx = mul(t, w), dy = mul(b, w), rx = cross(dy, base), ry = cross(base, dx);
    double det = dot(dx, rx), sg = det > 0 ? 1.0 : (det < 0 ? -1.0 : 0.0);
    v3 gradient = add(mul(rx, hx - hc), mul(ry, hy - hc));
    v3 out = sub(mul(base, fabs(det)), mul(gradient, dist * sg));
    if (dot(out, out) < 1e-30) return base;
    double k = clampd(strength, 0.0, 1.0);
    return norm(add(mul(base, 1 - k), mul(norm(out), k)));
}
/* shader_program.glsl:192-203 */
static double facing_cos(v3 n, v3 normal, v3 view) { return dot(view, dot(normal, normal) > 1e-12 ? norm(normal) : n); }

/* shader_program.glsl:205-349 */
static void run_program(const oscene *s, oprog *g, uint32_t m, uint32_t tri, v3 x, v3 n, v3 view, const double uv[2]) {
    g->mask = 0;
    g->viewed = 0;
    double entry = M(s, m, 13, 0);
    if (!(entry >= 0.0) || !(entry < (double)s->program_words)) return;
    double stack[PROGRAM_STACK];
    memset(stack, 0, sizeof(stack));
    const uint32_t *code = s->programs;
    size_t words = s->program_words, pc = (size_t)entry;
#define WORD() (pc < words ? code[pc++] : 0u)
#define OPERAND(dst) do { uint32_t w_ = WORD(); float f_; memcpy(&f_, &w_, 4); \
        dst = (w_ & 0xFFFF0000u) == SLOT_TAG ? stack[w_ & (PROGRAM_STACK - 1u)] : (double)f_; } while (0)
#define OPERAND3(dst) do { OPERAND(dst.x); OPERAND(dst.y); OPERAND(dst.z); } while (0)
#define PUT(v) do { double v_ = (v); stack[WORD() & (PROGRAM_STACK - 1u)] = v_; } while (0)
#define PUT3(v) do { v3 p_ = (v); PUT(p_.x); PUT(p_.y); PUT(p_.z); } while (0)
    for (uint32_t step = 0; step < 4096u; step++) {
        if (pc >= words) return;
        uint32_t op = code[pc++];
        if (op == 0) return;
        if (op == 1 || op == 2) {
            uint32_t axis = WORD();
            double pixels;
            OPERAND(pixels);
            v3 offset = V(0, 0, 0);
            if (axis != 0) {
                v3 t, b;
                bump_tangents(n, &t, &b);
                offset = mul(axis == 1 ? t : b, bump_width(s, x, pixels));
            }
            if (op == 1) {
                double d[2] = {0, 0};
                if (axis != 0) uv_step(s, tri, offset, d);
                PUT3(V(uv[0] + d[0], uv[1] + d[1], 0));
            } else {
                PUT3(add(x, offset));
            }
        } else if (op == 3) {
            double v;
            OPERAND(v);
            PUT(v);
        } else if (op == 4 || op == 12 || op == 13) {
            v3 v;
            OPERAND3(v);
            PUT3(v);
        } else if (op == 5) {
            uint32_t slot = WORD();
            int color = WORD() != 0;
            v3 at;
            OPERAND3(at);
            double t[4];
            texture_at(s, slot, at.x, at.y, t);
            v3 rgb = V(t[0], t[1], t[2]);
            if (color)
                rgb = V(0.613097 * t[0] + 0.339523 * t[1] + 0.047379 * t[2], 0.070194 * t[0] + 0.916354 * t[1] + 0.013452 * t[2],
                        0.020616 * t[0] + 0.109570 * t[1] + 0.869815 * t[2]);
            PUT3(rgb);
            PUT(t[3]);
        } else if (op == 6) {
            v3 at, a, b;
            double scale;
            OPERAND3(at); OPERAND3(a); OPERAND3(b); OPERAND(scale);
            long cx = (long)floor(at.x * scale + 1e-5), cy = (long)floor(at.y * scale + 1e-5), cz = (long)floor(at.z * scale + 1e-5);
            double fac = (double)((cx + cy + cz) & 1);
            PUT3(fac > 0.5 ? a : b);
            PUT(fac);
        } else if (op == 7) {
            v3 at;
            double scale, detail, rough;
            OPERAND3(at); OPERAND(scale); OPERAND(detail); OPERAND(rough);
            v3 q = mul(at, scale);
            double fac = fractal_noise(q, detail, rough);
            PUT(fac);
            PUT3(V(fac, fractal_noise(add(q, V(19.1, 33.4, 47.2)), detail, rough),
                   fractal_noise(add(q, V(73.7, 11.3, 29.9)), detail, rough)));
        } else if (op == 8) {
            uint32_t operation = WORD();
            double a, b;
            OPERAND(a); OPERAND(b);
            PUT(math_of(operation, a, b));
        } else if (op == 9) {
            double fac;
            v3 a, b;
            OPERAND(fac); OPERAND3(a); OPERAND3(b);
            fac = clampd(fac, 0, 1);
            PUT3(V(mixd(a.x, b.x, fac), mixd(a.y, b.y, fac), mixd(a.z, b.z, fac)));
        } else if (op == 10) {
            double v, f0, f1, t0, t1;
            OPERAND(v); OPERAND(f0); OPERAND(f1); OPERAND(t0); OPERAND(t1);
            double t = f1 != f0 ? (v - f0) / (f1 - f0) : 0.0;
            PUT(t0 + t * (t1 - t0));
        } else if (op == 11) {
            double v, lo, hi;
            OPERAND(v); OPERAND(lo); OPERAND(hi);
            PUT(clampd(v, fmin(lo, hi), fmax(lo, hi)));
        } else if (op == 14) {
            double fac, first, last;
            v3 a, b;
            OPERAND(fac); OPERAND3(a); OPERAND3(b); OPERAND(first); OPERAND(last);
            double t = last != first ? clampd((fac - first) / (last - first), 0, 1) : (fac < first ? 0.0 : 1.0);
            PUT3(V(mixd(a.x, b.x, t), mixd(a.y, b.y, t), mixd(a.z, b.z, t)));
        } else if (op == 15) {
            v3 color;
            double strength;
            OPERAND3(color); OPERAND(strength);
            PUT3(normal_from_map(s, tri, n, color, strength));
        } else if (op == 16) {
            uint32_t first = WORD(), width = WORD(), bit = WORD();
            for (uint32_t k = 0; k < width && k < 23u; k++) {
                double v;
                OPERAND(v);
                if (first + k < 23u) g->values[first + k] = v;
            }
            if (bit < 32u) g->mask |= 1u << bit;
        } else if (op == 17) {
            double strength, dist, pixels, hc, hx, hy;
            v3 normal;
            OPERAND(strength); OPERAND(dist); OPERAND(pixels); OPERAND(hc); OPERAND(hx); OPERAND(hy); OPERAND3(normal);
            PUT3(bumped(s, x, n, normal, strength, dist, pixels, hc, hx, hy));
        } else if (op == 18 || op == 19) {
            double parameter;
            v3 normal;
            OPERAND(parameter); OPERAND3(normal);
            g->viewed = 1;
            double c = facing_cos(n, normal, view);
            if (op == 18) {
                double eta = fmax(parameter, 1e-5);
                PUT(fresnel_dielectric(c, c < 0.0 ? 1.0 / eta : eta));
            } else {
                double eta = fmax(1.0 - parameter, 1e-5), facing = fabs(c);
                double fr = fresnel_dielectric(c, c < 0.0 ? eta : 1.0 / eta);
                if (parameter != 0.5) {
                    double b = clampd(parameter, 0.0, 1.0 - 1e-5);
                    facing = pow(facing, b < 0.5 ? 2.0 * b : 0.5 / (1.0 - b));
                }
                PUT(fr);
                PUT(1.0 - facing);
            }
        } else {
            return;
        }
    }
#undef WORD
#undef OPERAND
#undef OPERAND3
#undef PUT
#undef PUT3
}

/* One program run at a given hit, for tests: the material records, program
 * words and textures as the renderer receives them, the triangle's corners and
 * uvs, the frame; out: mask, viewed, values[23]. */
void or_run_program(const float *materials, const uint32_t *programs, size_t program_words, const float *textures,
                    const uint32_t *texture_table, size_t texture_count, const float *corners, const float *uvs,
                    const float *frame, uint32_t material, const double *x, const double *n, const double *view,
                    const double *uv, uint32_t *mask, int *viewed, double *values) {
    oscene s;
    memset(&s, 0, sizeof(s));
    static const uint32_t idx[3] = {0, 1, 2};
    double fr[16];
    for (int k = 0; k < 16; k++) fr[k] = frame[k];
    s.positions = corners; s.indices = idx; s.materials = materials; s.frame = fr; s.uvs = uvs;
    s.programs = programs; s.program_words = program_words; s.textures = textures; s.texture_table = texture_table;
    s.texture_count = texture_count;
    oprog g;
    memset(&g, 0, sizeof(g));
    run_program(&s, &g, material, 0, V(x[0], x[1], x[2]), V(n[0], n[1], n[2]), V(view[0], view[1], view[2]), uv);
    *mask = g.mask;
    *viewed = g.viewed;
    for (int k = 0; k < 23; k++) values[k] = g.values[k];
}

typedef struct {
    v3 n, t, b, g;
    uint32_t m;
    int back, inside;
    double droughness, metal, sw, trans, rough, croughness, ior, dark, lbase, lspec, ltrans, lcoat;
    d2v alpha, calpha;
    double eta, coat, ceta;
    double dcomp, missing, salb, ccomp, calb, ck;
    double pc, pm, ps, pt, pd;
    spec base, mtint, stint, ttint, ctint, substrate;
} surf;

static v3 to_local(const surf *s, v3 v) { return V(dot(v, s->t), dot(v, s->b), dot(v, s->n)); }
static v3 to_world(const surf *s, v3 v) { return add(add(mul(s->t, v.x), mul(s->b, v.y)), mul(s->n, v.z)); }

/* A material row as read_material sees it: the record's, with what the
 * program drives in its place (openpbr.glsl:264-287). */
static double Mg(const oscene *sc, const oprog *g, uint32_t m, int row, int c) {
    static const struct { int row, c, bit, at; } drive[] = {
        {1, 0, 0, 0}, {1, 1, 0, 1}, {1, 2, 0, 2}, {1, 3, 1, 3}, {10, 0, 2, 4}, {10, 1, 3, 5}, {3, 3, 4, 6},
        {10, 2, 5, 7}, {5, 3, 6, 8}, {11, 1, 7, 9}, {9, 3, 8, 10}, {3, 0, 10, 14}, {3, 1, 10, 15}, {3, 2, 10, 16},
        {9, 0, 11, 17}, {9, 1, 11, 18}, {9, 2, 11, 19}, {5, 0, 12, 20}, {5, 1, 12, 21}, {5, 2, 12, 22}};
    if (g && g->mask)
        for (size_t k = 0; k < sizeof(drive) / sizeof(drive[0]); k++)
            if (drive[k].row == row && drive[k].c == c && driven(g, (uint32_t)drive[k].bit)) return g->values[drive[k].at];
    return M(sc, m, row, c);
}

static void read_material(const oscene *sc, surf *s, double hero, const oprog *g) {
    uint32_t m = s->m;
#define M(sc, m, row, c) Mg(sc, g, m, row, c)
    s->droughness = M(sc, m, 10, 1);
    s->metal = clampd(M(sc, m, 10, 0), 0, 1);
    s->sw = fmax(M(sc, m, 3, 3), 0);
    s->trans = clampd(M(sc, m, 9, 3), 0, 1);
    s->inside = s->back && s->trans > 0;
    s->coat = !s->inside ? clampd(M(sc, m, 5, 3), 0, 1) : 0;
    s->croughness = clampd(M(sc, m, 11, 1), 0, 1);
    s->ceta = fmax(M(sc, m, 11, 2), 1.0001);
    s->calpha.x = s->calpha.y = 1;
    if (s->coat > 0) s->calpha = ggx_alpha(s->croughness, clampd(M(sc, m, 12, 2), 0, 1));
    s->dark = clampd(M(sc, m, 11, 3), 0, 1);
    double rough = clampd(M(sc, m, 10, 2), 0, 1);
    double r4 = rough * rough * rough * rough, c4 = pow(s->croughness, 4);
    s->rough = mixd(rough, fmin(1, pow(r4 + 2 * c4, 0.25)), s->coat);
    s->alpha = ggx_alpha(s->rough, clampd(M(sc, m, 10, 3), 0, 1));
    double eta = fmax(M(sc, m, 11, 0), 1.0001);
    if (sc->spectral && M(sc, m, 12, 3) != 0 && hero > 0)
        eta = fmax(eta + M(sc, m, 12, 3) * (1 / (hero * hero) - 1 / (587.6 * 587.6)), 1.0001);
    double rel = eta / s->ceta;
    if (rel < 1) rel = 1 / rel;
    s->ior = fmax(mixd(eta, rel, s->coat), 1.0001);
    s->eta = s->inside ? 1 / s->ior : s->ior;
    double bw = M(sc, m, 1, 3);
    s->lbase = fmax(lum(M(sc, m, 1, 0) * bw, M(sc, m, 1, 1) * bw, M(sc, m, 1, 2) * bw), 0);
    s->lspec = fmax(lum(M(sc, m, 3, 0), M(sc, m, 3, 1), M(sc, m, 3, 2)), 0);
    s->ltrans = fmax(lum(M(sc, m, 9, 0), M(sc, m, 9, 1), M(sc, m, 9, 2)), 0.05);
    s->lcoat = lum(M(sc, m, 5, 0), M(sc, m, 5, 1), M(sc, m, 5, 2));
#undef M
}
static void choose(surf *s, double wc, double wm, double ws, double wt, double wd) {
    if (s->inside) { wm = 0; wd = 0; ws = fmax(ws, 0.05); wt = fmax(wt, s->trans > 0 ? 0.05 : 0); }
    double tot = wc + wm + ws + wt + wd, sc = tot > 0 ? 1 / tot : 0;
    s->pc = wc * sc; s->pm = wm * sc; s->ps = ws * sc; s->pt = wt * sc; s->pd = wd * sc;
}
static void orient(const oscene *sc, surf *s, v3 wo) {
    double mu = clampd(wo.z, 0, 1);
    const double *t = sc->tables;
    s->ccomp = 1; s->calb = 0;
    if (s->coat > 0) {
        double e = fmax(table_ggx(t, mu, s->croughness), 1e-4);
        s->ccomp = 1 + fresnel_avg(s->ceta) * (1 - e) / e;
        s->calb = clampd(s->ccomp * table_dielectric(t, mu, s->croughness, s->ceta), 0, 1);
    }
    double e = fmax(table_ggx(t, mu, s->rough), 1e-4), miss = (1 - e) / e;
    s->missing = miss;
    s->dcomp = 1 + fresnel_avg(s->ior) * miss;
    s->salb = clampd(s->sw * s->dcomp * table_dielectric(t, mu, s->rough, s->ior), 0, 1);
    double sub_rgb = 1;
    s->ck = 0;
    if (s->coat > 0) {
        double ks = fresnel_dielectric(mu, s->ceta), kr = 1 - (1 - fresnel_avg(s->ceta)) / (s->ceta * s->ceta);
        s->ck = mixd(ks, kr, mixd(1, s->rough, s->metal));
        double drgb = (1 - s->ck) / fmax(1e-4, 1 - s->lbase * s->ck);
        sub_rgb = mixd(1, s->lcoat * (1 - s->calb) * mixd(1, drgb, s->dark), s->coat);
    }
    double di = 1 - s->metal;
    choose(s, s->coat * s->calb, sub_rgb * s->metal, sub_rgb * di * s->salb * s->lspec,
           sub_rgb * di * s->trans * (1 - s->salb) * s->ltrans, sub_rgb * di * (1 - s->trans) * (1 - s->salb) * s->lbase);
}
/* openpbr.glsl:378-387: what reaches the layers under the coat. */
static void update_substrate(surf *s) {
    s->substrate = S1(1);
    if (s->coat > 0)
        for (int k = 0; k < 4; k++) {
            double delta = (1 - s->ck) / (1 - s->base.v[k] * s->ck);
            double under = s->ctint.v[k] * (1 - s->calb) * mixd(1, delta, s->dark);
            s->substrate.v[k] = mixd(1, under, s->coat);
        }
}
/* A color a program drives: a reflectance through the environment basis,
 * at most one (openpbr.glsl:391-401). */
static spec driven_color(const oscene *sc, const oprog *g, uint32_t at, spec lambda) {
    double rgb[3] = {g->values[at], g->values[at + 1], g->values[at + 2]};
    spec c = spec_of_rgb(sc, rgb, lambda);
    for (int k = 0; k < 4; k++) c.v[k] = clampd(c.v[k], 0, 1);
    return c;
}
static void colors(const oscene *sc, surf *s, spec lambda, const oprog *g) {
    double weight = driven(g, 1) ? g->values[3] : M(sc, s->m, 1, 3);
    s->base = sscale(driven(g, 0) ? driven_color(sc, g, 0, lambda) : material_color(sc, s->m, 0, lambda), weight);
    s->mtint = driven(g, 10) ? driven_color(sc, g, 14, lambda) : material_color(sc, s->m, 2, lambda);
    s->stint = sscale(s->mtint, s->sw);
    s->ttint = driven(g, 11) ? driven_color(sc, g, 17, lambda) : material_color(sc, s->m, 8, lambda);
    s->ctint = s->coat > 0 ? (driven(g, 12) ? driven_color(sc, g, 20, lambda) : material_color(sc, s->m, 4, lambda)) : S1(1);
    update_substrate(s);
}
static int refraction_half(v3 wi, v3 wo, double eta, v3 *h) {
    *h = norm(add(wo, mul(wi, eta)));
    if (h->z < 0) *h = neg(*h);
    return dot(wo, *h) > 0 && dot(wi, *h) < 0;
}
static double surface_pdf(const surf *s, v3 wi, v3 wo) {
    if (wo.z <= 0 || wi.z == 0) return 0;
    if (wi.z < 0) {
        if (s->pt <= 0) return 0;
        v3 h;
        if (!refraction_half(wi, wo, s->eta, &h)) return 0;
        double oh = dot(wo, h), ih = dot(wi, h), den = oh + s->eta * ih, jac = s->eta * s->eta * fabs(ih) / (den * den);
        return s->pt * ggx_g1(wo, s->alpha) * ggx_d(h, s->alpha) * oh / wo.z * jac;
    }
    v3 h = norm(add(wi, wo));
    double pdf = 0;
    if (s->pc > 0) pdf += s->pc * ggx_reflection_pdf(wo, h, s->calpha);
    if (!s->inside) pdf += s->pd * wi.z / PI;
    pdf += (s->ps + s->pm) * ggx_reflection_pdf(wo, h, s->alpha);
    return pdf;
}
static spec evaluate(const surf *s, v3 wi, v3 wo, double *pdf) {
    *pdf = 0;
    if (wo.z <= 0 || wi.z == 0) return S1(0);
    *pdf = surface_pdf(s, wi, wo);
    double di = 1 - s->metal;
    if (wi.z < 0) {
        if (s->pt <= 0) return S1(0);
        v3 h;
        if (!refraction_half(wi, wo, s->eta, &h)) return S1(0);
        double oh = dot(wo, h), ih = dot(wi, h), den = oh + s->eta * ih;
        double f = (1 - fresnel_dielectric(oh, s->eta)) * ggx_d(h, s->alpha) * ggx_g2(neg(wi), wo, s->alpha) * fabs(ih) * oh /
                   (fabs(wi.z) * wo.z * den * den);
        return sscale(smul(s->substrate, s->ttint), di * s->trans * f);
    }
    v3 h = norm(add(wi, wo));
    double ch = dot(wo, h);
    spec f = S1(0), below = S1(0);
    if (s->pc > 0) {
        double coat = ggx_d(h, s->calpha) * ggx_g2(wi, wo, s->calpha) / (4 * wi.z * wo.z);
        f = S1(s->coat * fresnel_dielectric(ch, s->ceta) * s->ccomp * coat);
    }
    if (!s->inside) below = sadd(below, sscale(eon(s->base, s->droughness, wi, wo), di * (1 - s->trans) * (1 - s->salb)));
    double mf = ggx_d(h, s->alpha) * ggx_g2(wi, wo, s->alpha) / (4 * wi.z * wo.z);
    if (di > 0) below = sadd(below, sscale(s->stint, di * fresnel_dielectric(ch, s->eta) * s->dcomp * mf));
    if (s->metal > 0 && !s->inside) {
        spec F = f82(s->base, s->mtint, ch);
        for (int k = 0; k < 4; k++) {
            double avg = (20 * s->base.v[k] + 1) / 21, comp = 1 + avg * s->missing;
            below.v[k] += fmin(1.0, F.v[k] * s->sw) * comp * (s->metal * mf);
        }
    }
    return sadd(f, smul(s->substrate, below));
}
static int sample_dir(const surf *s, v3 wo, double ux, double uy, double uz, v3 *wi, double *pdf, uint32_t *lobe) {
    *pdf = 0;
    if (wo.z <= 0) return 0;
    double pick = uz;
    if (pick < s->pd) {
        double dx, dy;
        concentric(ux, uy, &dx, &dy);
        *wi = V(dx, dy, sqrt(fmax(0, 1 - (dx * dx + dy * dy))));
        *lobe = 0;
    } else if ((pick -= s->pd) < s->pc) {
        *wi = reflectv(neg(wo), ggx_sample(wo, s->calpha, ux, uy));
        *lobe = 1;
    } else if ((pick -= s->pc) < s->ps + s->pm) {
        *wi = reflectv(neg(wo), ggx_sample(wo, s->alpha, ux, uy));
        *lobe = 1;
    } else if (s->pt > 0) {
        v3 h = ggx_sample(wo, s->alpha, ux, uy);
        *wi = refractv(neg(wo), h, 1 / s->eta);
        if (dot(*wi, *wi) == 0) return 0;
        *lobe = 2;
    } else return 0;
    if (wi->z == 0) return 0;
    *pdf = surface_pdf(s, *wi, wo);
    return *pdf > 0;
}

/* ------------------------------------------------------------ the guide
 * The hash grid of vMF mixtures (shaders/guide.glsl:39-139,
 * guide_fit.comp:16-35): cells keyed by a quantised position (power-of-two
 * levels sized by the distance from the camera) and the normal's octant, in
 * an open-addressed table of GUIDE_CELLS slots probed linearly; per cell four
 * lobes (mean direction, kappa), their weights, and five statistics a lobe.
 * The state is f32 as the reference's buffers are; keys are computed from the
 * hit point rounded to f32, as the kernels see it. Cells are made when a
 * training record names them (a cell made earlier in a pass has no samples
 * yet, so it cannot guide in that pass either). */
#define GUIDE_CELLS 524288u
#define GUIDE_LOBES 4u
#define GUIDE_RING 3u
#define GUIDE_TRAINING 1u
#define GUIDE_PROBES 8u

typedef struct oguide {
    uint32_t *keys;   /* GUIDE_CELLS */
    float *lobes;     /* (GUIDE_LOBES + 1) x 4 a cell: lobes, then the weights */
    float *stats;     /* GUIDE_LOBES x 5 a cell */
} oguide;

/* One training record of a path vertex (shade_surface.comp:288-295): the
 * luminance the path held once this vertex's light sample was in (`base`)
 * and what that sample would have added had it got through (`blocked`). */
typedef struct { uint32_t key; v3 dir; double pdf, through, base, blocked, cosine; } orecord;

static uint32_t guide_key(const oscene *s, v3 xd, v3 nd) {
    float minimum = (float)K(s, 13, 0);
    float x = (float)xd.x, y = (float)xd.y, z = (float)xd.z;
    float cx = (float)s->frame[0], cy = (float)s->frame[1], cz = (float)s->frame[2];
    float dx = x - cx, dy = y - cy, dz = z - cz;
    float dist = sqrtf(dx * dx + dy * dy + dz * dz);
    float wanted = fmaxf(minimum, dist * (float)K(s, 12, 1));
    float lv = ceilf(log2f(wanted / minimum));
    lv = lv < 0.0f ? 0.0f : (lv > 31.0f ? 31.0f : lv);
    uint32_t level = (uint32_t)lv;
    float size = minimum * exp2f((float)level);
    int32_t qx = (int32_t)floorf(x / size), qy = (int32_t)floorf(y / size), qz = (int32_t)floorf(z / size);
    uint32_t octant = (nd.x >= 0 ? 1u : 0u) | (nd.y >= 0 ? 2u : 0u) | (nd.z >= 0 ? 4u : 0u);
    uint32_t key = or_hash_u32((uint32_t)qx * 73856093u ^
                               or_hash_u32((uint32_t)qy * 19349663u ^ or_hash_u32((uint32_t)qz * 83492791u ^ (octant | (level << 3u)))));
    return key > 1u ? key : 1u;
}
/* The cell holding `key`, NONE if none (make: claim an empty slot). */
static uint32_t guide_find(oguide *g, uint32_t key, int make) {
    uint32_t slot = key % GUIDE_CELLS;
    for (uint32_t probe = 0; probe < GUIDE_PROBES; probe++) {
        uint32_t at = (slot + probe) % GUIDE_CELLS;
        if (g->keys[at] == key) return at;
        if (g->keys[at] != 0u) continue;
        if (!make) return NONE;
        g->keys[at] = key;
        return at;
    }
    return NONE;
}
static double guide_count(const oguide *g, uint32_t cell) {
    double c = 0;
    for (uint32_t k = 0; k < GUIDE_LOBES; k++) c += g->stats[(cell * GUIDE_LOBES + k) * 5 + 4];
    return c;
}
/* guide.glsl:71-76: a lobe; before its first fit the tetrahedron's direction, kappa 1. */
static void guide_lobe(const oguide *g, uint32_t cell, uint32_t k, v3 *mean, double *kappa) {
    const float *v = &g->lobes[((size_t)cell * (GUIDE_LOBES + 1) + k) * 4];
    if (v[3] > 0.0f) { *mean = V(v[0], v[1], v[2]); *kappa = v[3]; return; }
    double c = 1.0 / sqrt(3.0);
    *mean = V((k == 0 || k == 1) ? c : -c, (k == 0 || k == 2) ? c : -c, (k == 0 || k == 3) ? c : -c);
    *kappa = 1.0;
}
static void guide_weights(const oguide *g, uint32_t cell, double w[4]) {
    const float *v = &g->lobes[((size_t)cell * (GUIDE_LOBES + 1) + GUIDE_LOBES) * 4];
    double t = (double)v[0] + v[1] + v[2] + v[3];
    for (int k = 0; k < 4; k++) w[k] = t > 0 ? v[k] / t : 0.25;
}
/* guide.glsl:84-98 */
static double vmf_pdf(v3 mean, double kappa, v3 d) {
    if (kappa < 1e-3) return 1.0 / (4.0 * PI);
    return kappa / (2.0 * PI * (1.0 - exp(-2.0 * kappa))) * exp(kappa * (dot(mean, d) - 1.0));
}
static v3 vmf_sample(v3 mean, double kappa, double ux, double uy) {
    kappa = fmax(kappa, 1e-3);
    double w = 1.0 + log(ux + (1.0 - ux) * exp(-2.0 * kappa)) / kappa;
    double sn = sqrt(fmax(0.0, 1.0 - w * w)), phi = 2.0 * PI * uy;
    v3 t, b;
    basis(mean, &t, &b);
    return norm(add(add(mul(t, sn * cos(phi)), mul(b, sn * sin(phi))), mul(mean, w)));
}
/* guide.glsl:100-119 */
static double guide_pdf(const oguide *g, uint32_t cell, v3 d) {
    double w[4], pdf = 0, kappa;
    v3 mean;
    guide_weights(g, cell, w);
    for (uint32_t k = 0; k < GUIDE_LOBES; k++) { guide_lobe(g, cell, k, &mean, &kappa); pdf += w[k] * vmf_pdf(mean, kappa, d); }
    return pdf;
}
static v3 guide_sample(const oguide *g, uint32_t cell, double ux, double uy, double uz) {
    double w[4], kappa;
    v3 mean;
    guide_weights(g, cell, w);
    uint32_t lobe = GUIDE_LOBES - 1;
    double pick = uz;
    for (uint32_t k = 0; k < GUIDE_LOBES; k++) {
        if (pick < w[k]) { lobe = k; break; }
        pick -= w[k];
    }
    guide_lobe(g, cell, lobe, &mean, &kappa);
    return vmf_sample(mean, kappa, ux, uy);
}
/* guide.glsl:121-139: the lobe that explains d best takes the sample. */
static void guide_train(oguide *g, uint32_t cell, v3 d, double value, double pdf) {
    if (cell == NONE || !(value > 0) || !(pdf > 0) || isinf(value)) return;
    double w[4], best_score = -1, kappa;
    v3 mean;
    uint32_t best = 0;
    guide_weights(g, cell, w);
    for (uint32_t k = 0; k < GUIDE_LOBES; k++) {
        guide_lobe(g, cell, k, &mean, &kappa);
        double score = (w[k] + 0.05) * vmf_pdf(mean, kappa, d);
        if (score > best_score) { best_score = score; best = k; }
    }
    float x = (float)fmin(value / pdf, 1e4);
    float *st = &g->stats[((size_t)cell * GUIDE_LOBES + best) * 5];
    st[0] += x;
    st[1] += x * (float)d.x;
    st[2] += x * (float)d.y;
    st[3] += x * (float)d.z;
    st[4] += 1.0f;
}
/* guide_fit.comp:16-35, in f32 as the buffers hold it: mean and Banerjee's
 * kappa from each lobe's statistics, weights from their shares; then the
 * statistics decay by 0.9. */
static void guide_fit(oguide *g) {
    for (uint32_t cell = 0; cell < GUIDE_CELLS; cell++) {
        if (g->keys[cell] == 0u) continue;
        float w4[4];
        for (uint32_t k = 0; k < GUIDE_LOBES; k++) {
            float *st = &g->stats[((size_t)cell * GUIDE_LOBES + k) * 5];
            float w = st[0], sx = st[1], sy = st[2], sz = st[3];
            float ss = sx * sx + sy * sy + sz * sz;
            if (w > 0.0f && ss > 0.0f) {
                float l = sqrtf(ss);
                float r = l / w;
                r = r < 0.0f ? 0.0f : (r > 0.999f ? 0.999f : r);
                float kappa = r * (3.0f - r * r) / (1.0f - r * r);
                kappa = kappa < 0.1f ? 0.1f : (kappa > 1000.0f ? 1000.0f : kappa);
                float *lb = &g->lobes[((size_t)cell * (GUIDE_LOBES + 1) + k) * 4];
                lb[0] = sx / l; lb[1] = sy / l; lb[2] = sz / l; lb[3] = kappa;
            }
            w4[k] = w;
            for (int c = 0; c < 5; c++) st[c] *= 0.9f;
        }
        if (w4[0] + w4[1] + w4[2] + w4[3] > 0.0f) {
            float *wt = &g->lobes[((size_t)cell * (GUIDE_LOBES + 1) + GUIDE_LOBES) * 4];
            for (int k = 0; k < 4; k++) wt[k] = w4[k];
        }
    }
}
/* The guide's pieces alone, for tests: a lobe's density and a direction
 * drawn from it, the fit of one cell's statistics (20 floats: 4 lobes x
 * weight, weighted direction, count) into its lobes and weights (20 floats),
 * the statistics decayed in place, and the key of a point. */
double or_vmf_pdf(const double mean[3], double kappa, const double d[3]) {
    return vmf_pdf(V(mean[0], mean[1], mean[2]), kappa, V(d[0], d[1], d[2]));
}
void or_vmf_sample(const double mean[3], double kappa, double ux, double uy, double out[3]) {
    v3 r = vmf_sample(V(mean[0], mean[1], mean[2]), kappa, ux, uy);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void or_guide_fit_cell(float stats[20], float lobes[20]) {
    static uint32_t key[GUIDE_CELLS];
    static float lb[(size_t)GUIDE_CELLS * (GUIDE_LOBES + 1) * 4], st[(size_t)GUIDE_CELLS * GUIDE_LOBES * 5];
    static pthread_mutex_t lock = PTHREAD_MUTEX_INITIALIZER;
    pthread_mutex_lock(&lock);
    oguide g = {key, lb, st};
    memset(key, 0, sizeof(key));
    key[7] = 1u;
    memcpy(&st[7 * GUIDE_LOBES * 5], stats, 20 * 4);
    memcpy(&lb[7 * (GUIDE_LOBES + 1) * 4], lobes, 20 * 4);
    guide_fit(&g);
    memcpy(stats, &st[7 * GUIDE_LOBES * 5], 20 * 4);
    memcpy(lobes, &lb[7 * (GUIDE_LOBES + 1) * 4], 20 * 4);
    pthread_mutex_unlock(&lock);
}
uint32_t or_guide_key(const float *constants, const float *frame16, const double x[3], const double n[3]) {
    oscene s;
    double fr[16];
    memset(&s, 0, sizeof(s));
    for (int k = 0; k < 16; k++) fr[k] = frame16[k];
    s.constants = constants;
    s.frame = fr;
    return guide_key(&s, V(x[0], x[1], x[2]), V(n[0], n[1], n[2]));
}

/* spec_luminance (spectrum.glsl): Y of the estimate. */
static double spec_luminance(const oscene *s, spec r, spec lambda) {
    if (!s->spectral) return r.v[0] * 0.2722287168 + r.v[1] * 0.6740817658 + r.v[2] * 0.0536895174;
    double sum = 0;
    for (int i = 0; i < 4; i++) {
        double pdf = wavelength_pdf(lambda.v[i]);
        if (pdf > 0) sum += cmf(lambda.v[i]).y * (r.v[i] / pdf);
    }
    return sum * 0.25 / K(s, 6, 1);
}

/* ------------------------------------------------------------ light tracing
 * Light paths and the three-technique MIS (shaders/connect.glsl:30-151,
 * light_paths.glsl:15-189, light_emit.comp:21-137, common.glsl:113-143,
 * 238-242): paths start at emitters chosen by power (or from a dome or wide
 * distant light, on a disk facing it), connect to the camera at every surface
 * they reach and splat what gets through onto the pixel it projects to;
 * camera paths weigh their BSDF hits and light samples against them through
 * the density ratios they carry (SV_MIS). */
#define AFAR_EMITTER 0x40000000u

/* A mistake planted for tests that check the checks: 1 drops the light
 * tracer's term from the light sample's MIS weight. */
static int planted = 0;
void or_plant(int mistake) { planted = mistake; }

static int tracing_on(const oscene *s) { return K(s, 14, 0) > 0; }
static double paths_per_sample(const oscene *s) { return K(s, 14, 0); }
static v3 frame_pos(const oscene *s) { return V(s->frame[0], s->frame[1], s->frame[2]); }
static v3 frame_du(const oscene *s) { return V(s->frame[8], s->frame[9], s->frame[10]); }
static v3 frame_dv(const oscene *s) { return V(s->frame[12], s->frame[13], s->frame[14]); }

/* camera.comp:40-48's lens point */
static v3 lens_point(const oscene *s, double ux, double uy) {
    double radius = K(s, 14, 2);
    if (radius <= 0) return frame_pos(s);
    double lx, ly;
    concentric(ux, uy, &lx, &ly);
    return add(add(frame_pos(s), mul(norm(frame_du(s)), lx * radius)), mul(norm(frame_dv(s)), ly * radius));
}
/* common.glsl:125-131: the camera's density of direction d (solid angle). */
static double camera_pdf(const oscene *s, v3 d) {
    double c = dot(d, Kv(s, 15));
    return c > 0 ? K(s, 14, 1) / (c * c * c) : 0;
}
/* connect.glsl:30-45 */
static int camera_pixel(const oscene *s, v3 x, v3 lens, uint32_t *pixel, v3 *d, double *dist) {
    v3 offset = sub(x, lens);
    *dist = len(offset);
    *d = mul(offset, 1.0 / fmax(*dist, 1e-20));
    double c = dot(*d, Kv(s, 15));
    if (c <= 1e-6 || *dist <= 0) return 0;
    double focus = K(s, 14, 3);
    v3 corner = V(s->frame[4], s->frame[5], s->frame[6]);
    v3 film = sub(K(s, 14, 2) > 0 ? mul(sub(add(lens, mul(*d, focus / c)), frame_pos(s)), 1.0 / focus) : mul(*d, 1.0 / c),
                  corner);
    v3 du = frame_du(s), dv = frame_dv(s);
    double fx = dot(film, du) / dot(du, du), fy = dot(film, dv) / dot(dv, dv);
    if (!(fx >= 0 && fy >= 0 && fx < s->frame[11] && fy < s->frame[15])) return 0;
    *pixel = (uint32_t)fy * (uint32_t)s->frame[11] + (uint32_t)fx;
    return 1;
}
/* connect.glsl:66-110: an emitter's chance by power, and the choice. */
static double emitter_power_pdf(const oscene *s, uint32_t e) {
    uint32_t record = (uint32_t)K(s, 10, 0) + e * 4;
    uint32_t trail = s->emitters[record + 1], depth = s->emitters[record + 2];
    uint32_t base = (uint32_t)K(s, 8, 1), node = 0;
    double pdf = 1;
    for (uint32_t level = 0; level < depth; level++) {
        uint32_t link = s->emitters[node * 2];
        double left = fmax(s->lights[(base + link * 3) * 4 + 3], 0), right = fmax(s->lights[(base + (link + 1) * 3) * 4 + 3], 0);
        if (left + right <= 0) return 0;
        uint32_t bit = (trail >> level) & 1u;
        pdf *= (bit == 0 ? left : right) / (left + right);
        node = link + bit;
    }
    return pdf;
}
static int emitter_by_power(const oscene *s, double u, uint32_t *e, double *pdf) {
    *pdf = 1;
    *e = 0;
    if (K(s, 8, 0) < 1) return 0;
    uint32_t base = (uint32_t)K(s, 8, 1), node = 0;
    for (uint32_t depth = 0; depth < 64; depth++) {
        uint32_t link = s->emitters[node * 2];
        if (s->emitters[node * 2 + 1] & 0x80000000u) { *e = link; return *pdf > 0; }
        double left = fmax(s->lights[(base + link * 3) * 4 + 3], 0), right = fmax(s->lights[(base + (link + 1) * 3) * 4 + 3], 0);
        if (left + right <= 0) return 0;
        double chance = left / (left + right);
        if (u < chance) { u = fmin(u / chance, 0.99999994); *pdf *= chance; node = link; }
        else { u = fmin((u - chance) / (1 - chance), 0.99999994); *pdf *= 1 - chance; node = link + 1; }
    }
    return 0;
}
static double emitter_area(const oscene *s, uint32_t code) {
    if (code & 0x80000000u) {
        uint32_t t = code & ~0x80000000u;
        v3 a = point_of(s, s->indices[t * 3]);
        return 0.5 * len(cross(sub(point_of(s, s->indices[t * 3 + 1]), a), sub(point_of(s, s->indices[t * 3 + 2]), a)));
    }
    if (light_type(s, code) == 2) { double r = L(s, code, 1, 3); return 4 * PI * r * r; }
    return L(s, code, 3, 3);
}
/* connect.glsl:112-133 */
static double emitter_area_pdf(const oscene *s, uint32_t e) {
    double finite = 1 - K(s, 21, 0), area = emitter_area(s, emitter_code(s, e));
    return area > 0 ? finite * emitter_power_pdf(s, e) / area : 0;
}
static double emission_direction_pdf(const oscene *s, uint32_t e, double c) {
    return fabs(c) / ((emitter_code(s, e) & 0x80000000u) ? 2 * PI : PI);
}
/* connect.glsl:140-151 */
static double camera_hit_ratio(const oscene *s, const double mis[4], uint32_t e, double cos_light, double dist,
                               double bsdf_pdf) {
    if (e == NONE || mis[1] <= 0 || bsdf_pdf <= 0) return 0;
    double d2 = dist * dist;
    double at_vertex = emission_direction_pdf(s, e, cos_light) * mis[2] / d2 / mis[1];
    double at_light = emitter_area_pdf(s, e) / (bsdf_pdf * fabs(cos_light) / d2);
    return mis[0] * at_vertex * at_light;
}
/* lights.glsl:290-308: light paths from afar */
static int emits_from_afar(const oscene *s, uint32_t l) {
    uint32_t type = light_type(s, l);
    return type == 4 || (type == 3 && L(s, l, 2, 3) < 1 - 1e-7);
}
static double afar_emission_pdf(const oscene *s, uint32_t l, v3 w) {
    return K(s, 21, 0) * L(s, l, 6, 1) * light_hit_pdf(s, l, V(0, 0, 0), w, INF);
}
static double afar_first_density(const oscene *s, v3 x, v3 w, double c) {
    double radius = K(s, 21, 1);
    v3 offset = sub(x, Kv(s, 22));
    v3 across = sub(offset, mul(w, dot(offset, w)));
    if (dot(across, across) > radius * radius) return 0;
    return c / (PI * radius * radius);
}
/* lights.glsl's cosine at a light toward the ray's origin */
static double light_cos(const oscene *s, uint32_t l, v3 o, v3 d, double t) {
    if (light_type(s, l) == 2) return fmax(0, -dot(d, norm(sub(add(o, mul(d, t)), Lv(s, l, 0)))));
    return fmax(0, -dot(d, Lv(s, l, 3)));
}
/* intersect_shadow.comp:51-57: a light the camera sees stands in the way */
static int camera_sees_light(const oscene *s, v3 o, v3 d, double reach) {
    for (uint32_t l = 0; l < light_count(s); l++)
        if (light_visible(s, l) && intersect_light(s, l, o, d, reach, 1) < reach) return 1;
    return 0;
}
/* openpbr.glsl:248-261: a surface at a hit, and the one seen from its other side */
static void make_surf(surf *sf, v3 normal, v3 geo, uint32_t m, int back) {
    memset(sf, 0, sizeof(*sf));
    sf->n = normal;
    basis(normal, &sf->t, &sf->b);
    sf->g = geo;
    sf->m = m;
    sf->back = back;
}
static void turned(const surf *a, surf *out) { make_surf(out, neg(a->n), neg(a->g), a->m, !a->back); }
/* openpbr.glsl:417-449: light paths' lobe choice, Fresnel at normal incidence */
static void light_lobes(surf *s) {
    double f0 = (s->ior - 1) / (s->ior + 1), spec_alb = clampd(s->sw * f0 * f0, 0, 1), coat_alb = 0;
    if (s->coat > 0) { double c0 = (s->ceta - 1) / (s->ceta + 1); coat_alb = c0 * c0; }
    double sub_rgb = mixd(1, s->lcoat * (1 - coat_alb), s->coat), di = 1 - s->metal;
    choose(s, s->coat * coat_alb, sub_rgb * s->metal, sub_rgb * di * spec_alb * s->lspec,
           sub_rgb * di * s->trans * (1 - spec_alb) * s->ltrans, sub_rgb * di * (1 - s->trans) * (1 - spec_alb) * s->lbase);
}
/* connect.glsl:47-63: the light sampler's pdf of `to` having arrived from `from` */
static double light_sampler_pdf(const oscene *sc, const surf *s, v3 from, v3 to, double hero, const oprog *g) {
    int same = dot(from, s->g) >= 0;
    surf v;
    if (same) v = *s; else turned(s, &v);
    v3 wo = to_local(&v, from), wi = to_local(&v, to);
    if ((dot(to, v.g) > 0) != (wi.z > 0)) return 0;
    if (!same) read_material(sc, &v, hero, g);
    light_lobes(&v);
    return surface_pdf(&v, wi, wo);
}

/* ------------------------------------------------------------ the path */

static double mis_weight(double a, double b) { a = fmin(a, 1e18); b = fmin(b, 1e18); return 1.0 / (1 + a * a + b * b); }
static v3 offset_origin(v3 p, v3 n) {
    double sc = fmax(1.0, fmax(fabs(p.x), fmax(fabs(p.y), fabs(p.z))));
    return add(p, mul(n, 1e-4 * sc));
}
static v3 oct_decode(uint32_t p) {
    double x = fmax(-1.0, (double)(int16_t)(p & 0xffff) / 32767.0), y = fmax(-1.0, (double)(int16_t)(p >> 16) / 32767.0);
    v3 n = V(x, y, 1 - fabs(x) - fabs(y));
    if (n.z < 0) {
        double nx = (1 - fabs(n.y)) * (n.x >= 0 ? 1 : -1), ny = (1 - fabs(n.x)) * (n.y >= 0 ? 1 : -1);
        n.x = nx; n.y = ny;
    }
    return norm(n);
}
/* octahedral_encode (scene.lucb:400-412) in f32, as the kernels store a
 * guide record's direction. */
static uint32_t oct_encode(v3 nd) {
    float nx = (float)nd.x, ny = (float)nd.y, nz = (float)nd.z;
    float l = fabsf(nx) + fabsf(ny) + fabsf(nz);
    float px = nx / l, py = ny / l;
    if (nz < 0.0f) {
        float qx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        float qy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = qx; py = qy;
    }
    px = px < -1.0f ? -1.0f : (px > 1.0f ? 1.0f : px);
    py = py < -1.0f ? -1.0f : (py > 1.0f ? 1.0f : py);
    int ix = (int)rintf(px * 32767.0f), iy = (int)rintf(py * 32767.0f);
    return ((uint32_t)ix & 0xffffu) | (((uint32_t)iy & 0xffffu) << 16);
}
static spec clamp_ind(const oscene *s, spec c, uint32_t total) {
    double cv = K(s, 0, 2);
    if (total == 0 || cv <= 0) return c;
    double m = spec_max(s, c);
    return m > cv ? sscale(c, cv / m) : c;
}
static uint32_t next_bounce(const oscene *s, uint32_t b, uint32_t lobe) {
    uint32_t shift = 8 * (lobe + 1), kind = ((b >> shift) & 0xff) + 1;
    if ((double)kind > K(s, 7, (int)lobe) || (b & 0xff) >= 255) return NONE;
    return (b & ~(0xffu << shift)) + (kind << shift) + 1;
}

/* Whether the shadow ray gets through; `kept` what the cut-outs pass
 * (intersect_shadow.comp:56-97). */
static int shadow(const oscene *s, v3 o, v3 d, double reach, double *kept) {
    *kept = 1.0;
    ohit h = closest(s, o, d, reach);
    if (h.tri == NONE) return 1;
    if (K(s, 0, 3) < 0.5) return 0;
    v3 from = o;
    double remaining = reach;
    for (int hop = 0; hop < 16; hop++) {
        ohit g = closest(s, from, d, remaining);
        if (g.tri == NONE) return 1;
        double op = clampd(M(s, s->shading[g.tri * 4 + 3], 12, 0), 0, 1);
        *kept *= 1 - op;
        if (*kept < 1e-4) return 0;
        v3 q = add(from, mul(d, g.t));
        double margin = 1e-4 * fmax(1.0, fmax(fabs(q.x), fmax(fabs(q.y), fabs(q.z))));
        from = add(q, mul(d, margin));
        remaining -= g.t + margin;
        if (remaining <= 0) return 1;
    }
    return 0;
}

/* shade_miss.comp:12-61; `mis` (light tracing) the path's density ratios and
 * `vertex` its last vertex. */
static spec miss_radiance(const oscene *s, uint32_t bounce, v3 rd, double ray_pdf, spec lambda, const double *mis,
                          v3 vertex) {
    spec rad = spec_of(s, &s->constants[4], &s->constants[8], lambda);
    for (uint32_t l = 0; l < light_count(s); l++) {
        uint32_t type = light_type(s, l);
        if (type < 3) continue;
        if (type == 3) {
            double ch = L(s, l, 2, 3);
            if (ch >= 1.0 - 1e-7 || dot(rd, neg(Lv(s, l, 3))) < ch) continue;
        }
        double w = 1.0;
        if (bounce == 0) { if (!light_visible(s, l)) continue; }
        else if (ray_pdf > 0) {
            double traced = 0;
            if (mis && emits_from_afar(s, l) && mis[1] > 0)
                traced = mis[0] * (afar_first_density(s, vertex, rd, mis[2]) / mis[1]) * (afar_emission_pdf(s, l, rd) / ray_pdf);
            w = mis_weight(light_hit_pdf(s, l, V(0, 0, 0), rd, INF) * light_select_pdf(s, l) / ray_pdf, traced);
        }
        rad = sadd(rad, sscale(type == 4 ? dome_radiance(s, l, rd, lambda) : light_radiance(s, l, lambda), w));
    }
    return rad;
}

/* One camera path of pixel `pixel`, sample `index`: its film value. With a
 * guide, `rec` receives the training records of its first GUIDE_RING
 * vertices (key 0 ends them) and `found` the luminance it carried home. */
static v3 trace_path(const oscene *s, uint32_t pixel, uint32_t index, int *nan, orecord *rec, double *found_out) {
    double u[4];
    sample4(s, pixel, index, 0, u);
    const double *f = s->frame;
    double fx = (double)(pixel % s->width) + u[0], fy = (double)(pixel / s->width) + u[1];
    v3 film = add(add(V(f[4], f[5], f[6]), mul(V(f[8], f[9], f[10]), fx)), mul(V(f[12], f[13], f[14]), fy));
    int ortho = f[3] > 0.5;
    v3 origin = ortho ? add(V(f[0], f[1], f[2]), film) : V(f[0], f[1], f[2]);
    v3 rd = norm(ortho ? cross(V(f[8], f[9], f[10]), V(f[12], f[13], f[14])) : film);
    if (!ortho && K(s, 14, 2) > 0) {
        v3 focused = add(V(f[0], f[1], f[2]), mul(rd, K(s, 14, 3) / dot(rd, Kv(s, 15))));
        double lx, ly;
        concentric(u[2], u[3], &lx, &ly);
        v3 du = V(f[8], f[9], f[10]), dv = V(f[12], f[13], f[14]);
        origin = add(add(V(f[0], f[1], f[2]), mul(norm(du), lx * K(s, 14, 2))), mul(norm(dv), ly * K(s, 14, 2)));
        rd = norm(sub(focused, origin));
    }
    spec lambda = S1(0);
    if (s->spectral) {
        double w[4];
        sample4(s, pixel, index, 1, w);
        lambda = sample_wavelengths(w[0]);
    }
    spec throughput = S1(1), radiance = S1(0);
    uint32_t bounces = 0;
    double ray_pdf = 0.0, survival_scale = 1.0;
    /* light tracing: the density ratios the path carries (camera.comp:47-49) */
    int tracing = tracing_on(s);
    double mis[4] = {0, 0, 0, 0};
    if (tracing) { mis[0] = paths_per_sample(s); mis[3] = camera_pdf(s, rd); }
    v3 vertex = origin;
    v3 prev_normal = V(0, 0, 1);
    uint32_t rounds = s->max_bounces + 1 + s->through;
    int guide_on = s->guide && K(s, 12, 0) > 0.5;
    if (rec) for (uint32_t k = 0; k < GUIDE_RING; k++) rec[k].key = 0;
    for (uint32_t round = 0; round < rounds; round++) {
        uint32_t total = bounces & 0xff;
        ohit h = closest(s, origin, rd, INF);
        uint32_t what = h.tri;
        double ht = h.t;
        for (uint32_t l = 0; l < light_count(s); l++) {
            if (total == 0 && !light_visible(s, l)) continue;
            double t = intersect_light(s, l, origin, rd, ht, 0);
            if (t < ht) { ht = t; what = 0x80000000u | l; }
        }
        if (what == NONE) {
            spec c = smul(throughput, miss_radiance(s, total, rd, ray_pdf, lambda, tracing ? mis : NULL, vertex));
            double cv = K(s, 0, 2), m = spec_max(s, c);
            if (total > 0 && cv > 0 && m > cv) c = sscale(c, cv / m);
            radiance = sadd(radiance, c);
            break;
        }
        if (what & 0x80000000u) {
            uint32_t light = what & ~0x80000000u;
            double w = 1.0;
            if (total == 0) { if (!light_visible(s, light)) break; }
            else if (ray_pdf > 0) {
                uint32_t em = (uint32_t)L(s, light, 6, 3);
                double reach = len(sub(vertex, add(origin, mul(rd, ht))));
                double chance = (1 - K(s, 8, 2)) * tree_pdf(s, em, vertex, prev_normal);
                double traced = tracing ? camera_hit_ratio(s, mis, em, light_cos(s, light, vertex, rd, reach), reach, ray_pdf) : 0;
                w = mis_weight(light_hit_pdf(s, light, vertex, rd, reach) * chance / ray_pdf, traced);
            }
            radiance = sadd(radiance, clamp_ind(s, sscale(smul(throughput, light_radiance(s, light, lambda)), w), total));
            break;
        }
        uint32_t tri = what;
        v3 p0 = point_of(s, s->indices[tri * 3]);
        v3 geo = norm(cross(sub(point_of(s, s->indices[tri * 3 + 1]), p0), sub(point_of(s, s->indices[tri * 3 + 2]), p0)));
        int inside = dot(geo, rd) > 0;
        if (inside) geo = neg(geo);
        const uint32_t *attr = &s->shading[tri * 4];
        v3 x = add(origin, mul(rd, ht));
        double reach = len(sub(vertex, x));
        double density = (total == 0 ? mis[3] : ray_pdf) * fabs(dot(rd, geo)) / (reach * reach);
        double back_reach = mis[2] / (reach * reach);
        double uo[4];
        sample4(s, pixel, index, 4 + 3 * total, uo);
        double op = clampd(M(s, attr[3], 12, 0), 0, 1);
        if (op < 1) {
            float htf = (float)ht;
            uint32_t bits;
            memcpy(&bits, &htf, 4);
            if (fract(uo[0] + (double)or_hash_u32(bits) * (1.0 / 4294967296.0)) >= op) {
                origin = offset_origin(x, neg(geo));
                continue;
            }
        }
        if (s->spectral && M(s, attr[3], 12, 3) != 0 && lambda.v[1] >= 0) {
            throughput = (spec){{throughput.v[0] * 4, 0, 0, 0}};
            radiance = (spec){{radiance.v[0] * 4, 0, 0, 0}};
            lambda = (spec){{lambda.v[0], -lambda.v[1], -lambda.v[2], -lambda.v[3]}};
        }
        double hero = lambda.v[1] < 0 ? lambda.v[0] : 0;
        if (M(s, attr[3], 7, 3) > 0) {
            double w = 1;
            uint32_t em = triangle_emitter(s, tri);
            if (total > 0 && ray_pdf > 0 && em != NONE) {
                double chance = (1 - K(s, 8, 2)) * tree_pdf(s, em, vertex, prev_normal);
                double traced = tracing ? camera_hit_ratio(s, mis, em, fabs(dot(rd, geo)), reach, ray_pdf) : 0;
                w = mis_weight(triangle_hit_pdf(s, tri, rd, reach) * chance / ray_pdf, traced);
            }
            radiance = sadd(radiance, clamp_ind(s, sscale(smul(throughput, material_color(s, attr[3], 6, lambda)), w), total));
        }
        if (total >= (uint32_t)K(s, 0, 1)) break;
        double b1 = h.u, b2 = h.v;
        v3 normal = norm(add(add(mul(oct_decode(attr[0]), 1 - b1 - b2), mul(oct_decode(attr[1]), b1)), mul(oct_decode(attr[2]), b2)));
        if (dot(normal, geo) < 0) normal = neg(normal);
        /* the material's program, on the outward normal (shade_surface.comp:134-145) */
        oprog prog;
        memset(&prog, 0, sizeof(prog));
        if (s->program_words) {
            double uv[2];
            triangle_uv(s, tri, b1, b2, uv);
            run_program(s, &prog, attr[3], tri, x, inside ? neg(normal) : normal, neg(rd), uv);
            if (driven(&prog, 9)) {
                normal = norm(driven3(&prog, 11));
                if (dot(normal, geo) < 0) normal = neg(normal);
            }
        }
        surf sf;
        memset(&sf, 0, sizeof(sf));
        sf.n = normal;
        basis(normal, &sf.t, &sf.b);
        sf.g = geo;
        sf.m = attr[3];
        sf.back = inside;
        v3 wo = to_local(&sf, neg(rd));
        read_material(s, &sf, hero, &prog);
        orient(s, &sf, wo);
        colors(s, &sf, lambda, &prog);
        v3 from = offset_origin(x, geo);
        /* Path guiding (shade_surface.comp:155-164): in a cell with enough
         * samples, BSDF sampling becomes a one-sample mix of the guide at
         * share K_GUIDE.z and the BSDF. */
        uint32_t key = 0, cell = NONE;
        double guided = 0;
        if (guide_on && total < GUIDE_RING) {
            key = guide_key(s, x, geo);
            cell = guide_find(s->guide, key, 0);
            int suited = !sf.inside && sf.pt <= 0 && (sf.pd > 0.05 || fmin(sf.alpha.x, sf.alpha.y) > 0.1);
            float count = 0.0f;
            if (cell != NONE)
                for (uint32_t k = 0; k < GUIDE_LOBES; k++) count += s->guide->stats[((size_t)cell * GUIDE_LOBES + k) * 5 + 4];
            if (cell != NONE && suited && count >= (float)K(s, 12, 3)) guided = K(s, 12, 2);
        }
        int pushed = 0;
        spec nee = S1(0);
        double ul[4];
        sample4(s, pixel, index, 2 + 3 * total, ul);
        double p_distant = K(s, 8, 2);
        lsample ls;
        spec emitted = S1(0);
        double chance = 0;
        int sampled = 0;
        uint32_t emitter = NONE, afar = NONE;
        if (ul[0] < p_distant) {
            double share;
            uint32_t light = pick_distant(s, ul[0] / p_distant, &share);
            if (light != NONE && share > 0 && sample_light(s, light, x, ul[1], ul[2], &ls)) {
                afar = light;
                chance = p_distant * share;
                emitted = light_type(s, light) == 4 ? dome_radiance(s, light, ls.direction, lambda)
                                                    : light_radiance(s, light, lambda);
                sampled = 1;
            }
        } else {
            double tc;
            if (tree_sample(s, (ul[0] - p_distant) / fmax(1e-7, 1 - p_distant), from, normal, &emitter, &tc)) {
                chance = (1 - p_distant) * tc;
                uint32_t code = emitter_code(s, emitter);
                if (code & 0x80000000u) sampled = sample_triangle(s, code & ~0x80000000u, x, ul[1], ul[2], lambda, &ls, &emitted);
                else if (sample_light(s, code, x, ul[1], ul[2], &ls)) { emitted = light_radiance(s, code, lambda); sampled = 1; }
            }
        }
        if (sampled && chance > 0) {
            double side = dot(ls.direction, geo);
            v3 wi = to_local(&sf, ls.direction);
            if ((side > 0) == (wi.z > 0) && (side > 0 || sf.pt > 0)) {
                double bp;
                spec fv = evaluate(&sf, wi, wo, &bp);
                if (guided > 0 && !tracing) bp = guided * guide_pdf(s->guide, cell, ls.direction) + (1 - guided) * bp;
                if (bp > 0) {
                    double lp = ls.pdf * chance;
                    /* the light tracer's density of this path over the light
                     * sample's (shade_surface.comp:206-228) */
                    double traced = 0;
                    int viewed = prog.viewed;
                    if (tracing && !ls.delta && emitter != NONE && density > 0) {
                        double d2 = ls.distance * ls.distance;
                        double at_last = total == 0 ? 1.0 : (viewed ? 0.0 :
                            light_sampler_pdf(s, &sf, ls.direction, neg(rd), hero, &prog) * back_reach / mis[1]);
                        double at_vertex = emission_direction_pdf(s, emitter, ls.cos_light) * fabs(side) / d2 / density;
                        double at_light = emitter_area_pdf(s, emitter) / (bp * ls.cos_light / d2);
                        traced = mis[0] * at_last * at_vertex * at_light;
                    }
                    if (tracing && afar != NONE && emits_from_afar(s, afar) && density > 0) {
                        double at_last = total == 0 ? 1.0 : (viewed ? 0.0 :
                            light_sampler_pdf(s, &sf, ls.direction, neg(rd), hero, &prog) * back_reach / mis[1]);
                        double at_vertex = afar_first_density(s, x, ls.direction, fabs(side)) / density;
                        double at_light = afar_emission_pdf(s, afar, ls.direction) / bp;
                        traced = mis[0] * at_last * at_vertex * at_light;
                    }
                    double w = ls.delta ? 1.0 : mis_weight(bp / lp, (planted == 1 ? 0.0 : traced) * bp / lp);
                    spec c = clamp_ind(s, sscale(smul(smul(throughput, fv), emitted), fabs(wi.z) * (w / lp)), total);
                    pushed = 1;
                    nee = c;
                    v3 target = add(x, mul(ls.direction, ls.distance));
                    double margin = 4e-4 * fmax(1.0, fmax(fabs(target.x), fmax(fabs(target.y), fabs(target.z))));
                    double rs = ls.distance >= INF ? INF : fmax(0, ls.distance - margin);
                    double kept;
                    if (shadow(s, offset_origin(x, side > 0 ? geo : neg(geo)), ls.direction, rs, &kept)) {
                        radiance = sadd(radiance, sscale(c, kept));
                        nee = sscale(c, 1.0 - kept);
                    }
                }
            }
        }
        double ub[4];
        sample4(s, pixel, index, 3 + 3 * total, ub);
        v3 wi;
        double pdf;
        uint32_t lobe;
        spec weight;
        double unguided = 1.0, bsdf_only;
        if (uo[1] < guided) {
            /* a guide sample (shade_surface.comp:252-262) */
            wi = to_local(&sf, guide_sample(s->guide, cell, uo[2], uo[3], ub[3]));
            double bp;
            spec fv = evaluate(&sf, wi, wo, &bp);
            pdf = guided * guide_pdf(s->guide, cell, to_world(&sf, wi)) + (1 - guided) * bp;
            bsdf_only = bp;
            if (pdf <= 0 || spec_max(s, fv) <= 0) break;
            unguided = pdf / fmax(bp, 1e-6 * pdf);
            weight = sscale(fv, fabs(wi.z) / pdf);
            lobe = wi.z < 0 ? 2u : (sf.pd > 0.5 ? 0u : 1u);
        } else {
            if (!sample_dir(&sf, wo, ub[0], ub[1], ub[3], &wi, &pdf, &lobe)) break;
            double again;
            spec fv = evaluate(&sf, wi, wo, &again);
            weight = sscale(fv, fabs(wi.z) / pdf);
            bsdf_only = pdf;
            if (guided > 0) {
                double mixed = guided * guide_pdf(s->guide, cell, to_world(&sf, wi)) + (1 - guided) * pdf;
                unguided = mixed / pdf;
                weight = sscale(weight, pdf / mixed);
                pdf = mixed;
            }
        }
        v3 dir = to_world(&sf, wi);
        double side = dot(dir, geo);
        if ((side > 0) != (wi.z > 0)) break;
        bounces = next_bounce(s, bounces, lobe);
        if (bounces == NONE) break;
        throughput = smul(throughput, weight);
        survival_scale = fmin(survival_scale * unguided, 1e6);
        if (total + 1 >= 3) {
            double survive = fmin(0.95, spec_max(s, throughput) * survival_scale);
            if (ub[2] >= survive) break;
            throughput = sscale(throughput, 1.0 / survive);
        }
        /* the vertex's training record (shade_surface.comp:288-295) */
        if (guide_on && total < GUIDE_RING && rec) {
            orecord *r = &rec[total];
            r->key = key ? key : guide_key(s, x, geo);
            r->dir = oct_decode(oct_encode(dir));
            r->pdf = (float)pdf;
            r->through = (float)spec_luminance(s, throughput, lambda);
            r->base = spec_luminance(s, radiance, lambda);
            r->blocked = pushed ? spec_luminance(s, nee, lambda) : 0.0;
            r->cosine = (float)fabs(dot(dir, normal));
            if (total + 1 < GUIDE_RING) rec[total + 1].key = 0;
        }
        prev_normal = normal;
        origin = side > 0 ? from : offset_origin(x, neg(geo));
        vertex = origin;
        if (tracing) {
            /* settle this vertex's factor (shade_surface.comp:302-309) */
            double at_last = total == 0 ? 1.0 : (prog.viewed ? 0.0 :
                light_sampler_pdf(s, &sf, dir, neg(rd), hero, &prog) * back_reach / mis[1]);
            mis[0] *= at_last;
            mis[1] = density;
            mis[2] = fabs(side);
            mis[3] = 0;
            pdf = bsdf_only;
        }
        rd = dir;
        ray_pdf = pdf;
    }
    if (found_out) *found_out = spec_luminance(s, radiance, lambda);
    v3 value = spec_to_film(s, radiance, lambda);
    if (isnan(value.x) || isnan(value.y) || isnan(value.z) || isinf(value.x) || isinf(value.y) || isinf(value.z)) {
        value = V(0, 0, 0);
        if (nan) (*nan)++;
    }
    return value;
}

/* One light path (light_emit.comp:21-137, then light_paths.glsl:31-189 at each
 * surface it reaches): what its camera connections deliver, MIS-weighted, is
 * added to `splat` (3 doubles a pixel, the film's terms). `slot` is the path's
 * slot after the band's camera paths; `index` the sample. */
static void trace_light(const oscene *s, uint32_t slot, uint32_t index, double *splat) {
    uint32_t id = 0x80000000u | slot;
    double u_point[4], u_pick[4];
    sample4(s, id, index, 0, u_point);
    sample4(s, id, index, 1, u_pick);
    spec lambda = s->spectral ? sample_wavelengths(u_pick[0]) : S1(0);
    double afar = K(s, 21, 0);
    v3 origin, dir, vertex;
    double ray_pdf, survival, mis[4];
    spec throughput;
    uint32_t emitter;
    if (u_pick[1] < afar) {
        /* from afar: a direction by the light's distribution, a point on the
         * disk facing it */
        double share;
        uint32_t light = pick_distant(s, u_pick[1] / afar, &share);
        if (light == NONE || share <= 0 || !emits_from_afar(s, light)) return;
        lsample ls;
        if (!sample_light(s, light, V(0, 0, 0), u_point[0], u_point[1], &ls)) return;
        v3 w = ls.direction, t, b;
        double radius = K(s, 21, 1), dx, dy;
        basis(w, &t, &b);
        concentric(u_point[2], u_point[3], &dx, &dy);
        v3 start = add(add(add(Kv(s, 22), mul(w, 2 * radius)), mul(t, dx * radius)), mul(b, dy * radius));
        double chance = afar * share * ls.pdf;
        spec emitted = light_type(s, light) == 4 ? dome_radiance(s, light, w, lambda) : light_radiance(s, light, lambda);
        throughput = sscale(emitted, PI * radius * radius / chance);
        double most = spec_max(s, throughput);
        if (!(most > 0)) return;
        emitter = AFAR_EMITTER | light;
        origin = start; dir = neg(w); ray_pdf = 0; vertex = start; survival = 1 / most;
        mis[0] = 1; mis[1] = chance; mis[2] = 1; mis[3] = -1;
    } else {
        double chance;
        if (!emitter_by_power(s, (u_pick[1] - afar) / fmax(1e-7, 1 - afar), &emitter, &chance)) return;
        chance *= 1 - afar;
        uint32_t code = emitter_code(s, emitter);
        v3 q, n;
        spec radiance;
        double area;
        int two_sided = 0;
        if (code & 0x80000000u) {
            uint32_t tri = code & ~0x80000000u;
            v3 a = point_of(s, s->indices[tri * 3]), b = point_of(s, s->indices[tri * 3 + 1]), c = point_of(s, s->indices[tri * 3 + 2]);
            double r = sqrt(u_point[0]);
            q = add(add(mul(a, 1 - r), mul(b, r * (1 - u_point[1]))), mul(c, r * u_point[1]));
            v3 cab = cross(sub(b, a), sub(c, a));
            area = 0.5 * len(cab);
            n = mul(cab, 1.0 / fmax(2 * area, 1e-30));
            if (u_pick[2] < 0.5) n = neg(n);
            two_sided = 1;
            radiance = material_color(s, s->shading[tri * 4 + 3], 6, lambda);
        } else {
            uint32_t type = light_type(s, code);
            if (type == 2) {
                double r = L(s, code, 1, 3), z = 1 - 2 * u_point[0], sn = sqrt(fmax(0, 1 - z * z)), phi = 2 * PI * u_point[1];
                n = V(sn * cos(phi), sn * sin(phi), z);
                q = add(Lv(s, code, 0), mul(n, r));
                area = 4 * PI * r * r;
            } else {
                double ax, ay;
                if (type == 0) { ax = 2 * u_point[0] - 1; ay = 2 * u_point[1] - 1; }
                else concentric(u_point[0], u_point[1], &ax, &ay);
                q = add(add(Lv(s, code, 0), mul(Lv(s, code, 1), ax)), mul(Lv(s, code, 2), ay));
                n = Lv(s, code, 3);
                area = L(s, code, 3, 3);
            }
            radiance = light_radiance(s, code, lambda);
        }
        if (area <= 0) return;
        double dx, dy;
        concentric(u_point[2], u_point[3], &dx, &dy);
        double cos_n = sqrt(fmax(0, 1 - (dx * dx + dy * dy)));
        if (cos_n <= 0) return;
        v3 t, b;
        basis(n, &t, &b);
        dir = norm(add(add(mul(t, dx), mul(b, dy)), mul(n, cos_n)));
        double point_pdf = chance / area, direction_pdf = cos_n / (two_sided ? 2 * PI : PI);
        throughput = sscale(radiance, cos_n / (point_pdf * direction_pdf));
        double largest = spec_max(s, throughput);
        if (!(largest > 0)) return;
        origin = offset_origin(q, n); ray_pdf = direction_pdf; vertex = q; survival = 1 / largest;
        mis[0] = 1; mis[1] = point_pdf; mis[2] = cos_n; mis[3] = -1;
    }
    double lpps = paths_per_sample(s);
    uint32_t bounces = 0;
    uint32_t rounds = s->max_bounces + 1 + s->through;
    for (uint32_t round = 0; round < rounds; round++) {
        uint32_t total = bounces & 0xff;
        ohit h = closest(s, origin, dir, INF);
        uint32_t what = h.tri;
        double ht = h.t;
        for (uint32_t l = 0; l < light_count(s); l++) {
            if (total == 0 && !light_visible(s, l)) continue;
            double tl = intersect_light(s, l, origin, dir, ht, 1);
            if (tl < ht) { ht = tl; what = 0x80000000u | l; }
        }
        if (what == NONE || (what & 0x80000000u)) return;  /* light paths that leave, or reach a light, end */
        uint32_t tri = what;
        v3 p0 = point_of(s, s->indices[tri * 3]);
        v3 geo = norm(cross(sub(point_of(s, s->indices[tri * 3 + 1]), p0), sub(point_of(s, s->indices[tri * 3 + 2]), p0)));
        int back = dot(geo, dir) > 0;
        if (back) geo = neg(geo);
        const uint32_t *attr = &s->shading[tri * 4];
        v3 x = add(origin, mul(dir, ht));
        double uo[4];
        sample4(s, id, index, 4 + 3 * total, uo);
        double op = clampd(M(s, attr[3], 12, 0), 0, 1);
        if (op < 1) {
            float htf = (float)ht;
            uint32_t bits;
            memcpy(&bits, &htf, 4);
            if (fract(uo[0] + (double)or_hash_u32(bits) * (1.0 / 4294967296.0)) >= op) {
                origin = offset_origin(x, neg(geo));
                continue;
            }
        }
        if (total >= (uint32_t)K(s, 0, 1)) return;
        if (s->spectral && M(s, attr[3], 12, 3) != 0 && lambda.v[1] >= 0) {
            throughput = (spec){{throughput.v[0] * 4, 0, 0, 0}};
            lambda = (spec){{lambda.v[0], -lambda.v[1], -lambda.v[2], -lambda.v[3]}};
        }
        double hero = lambda.v[1] < 0 ? lambda.v[0] : 0;
        double b1 = h.u, b2 = h.v;
        v3 normal = norm(add(add(mul(oct_decode(attr[0]), 1 - b1 - b2), mul(oct_decode(attr[1]), b1)), mul(oct_decode(attr[2]), b2)));
        if (dot(normal, geo) < 0) normal = neg(normal);
        double ul[4];
        sample4(s, id, index, 2 + 3 * total, ul);
        v3 lens = lens_point(s, ul[0], ul[1]), to_point;
        uint32_t pixel = 0;
        double cam_dist;
        int on_film = camera_pixel(s, x, lens, &pixel, &to_point, &cam_dist);
        oprog prog;
        memset(&prog, 0, sizeof(prog));
        int viewed = 0;
        if (s->program_words) {
            double uv[2];
            triangle_uv(s, tri, b1, b2, uv);
            run_program(s, &prog, attr[3], tri, x, back ? neg(normal) : normal, on_film ? neg(to_point) : neg(dir), uv);
            viewed = prog.viewed;
            if (viewed && !on_film) return;
            if (driven(&prog, 9)) {
                normal = norm(driven3(&prog, 11));
                if (dot(normal, geo) < 0) normal = neg(normal);
            }
        }
        surf lit;
        make_surf(&lit, normal, geo, attr[3], back);
        v3 toward_light = neg(dir);
        read_material(s, &lit, hero, &prog);
        colors(s, &lit, lambda, &prog);
        v3 x_last = vertex;
        double reach = len(sub(x_last, x)), cos_in = fabs(dot(dir, geo));
        double density = ray_pdf * cos_in / (reach * reach), back_reach = mis[2] / (reach * reach);
        if (total == 0 && (emitter & AFAR_EMITTER)) {
            density = afar_first_density(s, x, neg(dir), cos_in);
            back_reach = 1;
            if (!(density > 0)) return;
        }
        double cos_shading_in = dot(toward_light, normal);
        if (cos_shading_in <= 0 || cos_in < 1e-6) return;
        double shading_in = cos_shading_in / cos_in;
        if (on_film) {
            /* connect to the camera */
            v3 to_camera = neg(to_point);
            double side = dot(to_camera, geo);
            surf seen;
            if (side > 0) seen = lit; else turned(&lit, &seen);
            v3 wo = to_local(&seen, to_camera), wi = to_local(&seen, toward_light);
            if ((dot(toward_light, seen.g) > 0) == (wi.z > 0) && wo.z > 0) {
                if (side > 0) { orient(s, &seen, wo); update_substrate(&seen); }
                else { read_material(s, &seen, hero, &prog); orient(s, &seen, wo); colors(s, &seen, lambda, &prog); }
                double cbp;
                spec f = evaluate(&seen, wi, wo, &cbp);
                double camera_density = camera_pdf(s, to_point) * fabs(side) / (cam_dist * cam_dist);
                spec contribution = sscale(smul(throughput, f), shading_in * camera_density / lpps);
                if (spec_max(s, contribution) > 0) {
                    double last = mis[1] > 0 ? cbp * back_reach / mis[1] : 0;
                    double ratio = mis[0] * last * camera_density / density;
                    double sample_ratio = total == 0 ? 0 : fmax(mis[3], 0);
                    if (total == 0) {
                        /* the light sample a camera path here would take toward x0 (light_paths.glsl:15-29) */
                        if (cbp > 0) {
                            v3 from = offset_origin(x, seen.g);
                            if (emitter & AFAR_EMITTER) {
                                uint32_t light = emitter & ~AFAR_EMITTER;
                                sample_ratio = light_select_pdf(s, light) * light_hit_pdf(s, light, from, toward_light, INF) / cbp;
                            } else {
                                v3 off = sub(x_last, from);
                                double rr = len(off);
                                v3 dd = mul(off, 1.0 / rr);
                                double ch = (1 - K(s, 8, 2)) * tree_pdf(s, emitter, from, seen.n);
                                uint32_t code = emitter_code(s, emitter);
                                double lpdf = (code & 0x80000000u) ? triangle_hit_pdf(s, code & ~0x80000000u, dd, rr)
                                                                   : light_hit_pdf(s, code, from, dd, rr);
                                sample_ratio = ch * lpdf / cbp;
                            }
                        }
                    }
                    double weight = mis_weight(ratio / lpps, ratio * sample_ratio / lpps);
                    double margin = 4e-4 * fmax(1.0, fmax(fabs(x.x), fmax(fabs(x.y), fabs(x.z))));
                    v3 so = offset_origin(x, seen.g);
                    double rs = fmax(0, cam_dist - margin);
                    v3 value = spec_to_film(s, sscale(contribution, weight), lambda);
                    double kept;
                    if (!camera_sees_light(s, so, to_camera, rs) && shadow(s, so, to_camera, rs, &kept) &&
                        !isnan(value.x) && !isnan(value.y) && !isnan(value.z) && !isinf(value.x) && !isinf(value.y) &&
                        !isinf(value.z)) {
                        splat[(size_t)pixel * 3] += value.x * kept;
                        splat[(size_t)pixel * 3 + 1] += value.y * kept;
                        splat[(size_t)pixel * 3 + 2] += value.z * kept;
                    }
                }
            }
        }
        /* scatter on */
        if (viewed || total + 1 >= (uint32_t)K(s, 0, 1)) return;
        double u[4];
        sample4(s, id, index, 3 + 3 * total, u);
        surf choosing = lit;
        light_lobes(&choosing);
        v3 wo_light = to_local(&lit, toward_light), wi_light;
        double light_pdf;
        uint32_t lobe;
        if (!sample_dir(&choosing, wo_light, u[0], u[1], u[3], &wi_light, &light_pdf, &lobe)) return;
        v3 direction = to_world(&lit, wi_light);
        double side = dot(direction, geo);
        if ((side > 0) != (wi_light.z > 0)) return;
        surf seen;
        if (side > 0) seen = lit; else turned(&lit, &seen);
        v3 wo = to_local(&seen, direction), wi = to_local(&seen, toward_light);
        if ((dot(toward_light, seen.g) > 0) != (wi.z > 0)) return;
        if (side > 0) { orient(s, &seen, wo); update_substrate(&seen); }
        else { read_material(s, &seen, hero, &prog); orient(s, &seen, wo); colors(s, &seen, lambda, &prog); }
        double cbp;
        spec f = evaluate(&seen, wi, wo, &cbp);
        bounces = next_bounce(s, bounces, lobe);
        if (bounces == NONE) return;
        throughput = sscale(smul(throughput, f), shading_in * fabs(side) / light_pdf);
        if (total + 1 >= 3) {
            double survive = fmin(0.95, spec_max(s, throughput) * survival);
            if (u[2] >= survive) return;
            throughput = sscale(throughput, 1.0 / survive);
        }
        if (!(spec_max(s, throughput) > 0)) return;
        double last = mis[1] > 0 ? cbp * back_reach / mis[1] : 0;
        double sample_ratio = mis[3];
        if (total == 0) {
            sample_ratio = 0;
            if (cbp > 0) {
                v3 from = offset_origin(x, seen.g);
                if (emitter & AFAR_EMITTER) {
                    uint32_t light = emitter & ~AFAR_EMITTER;
                    sample_ratio = light_select_pdf(s, light) * light_hit_pdf(s, light, from, toward_light, INF) / cbp;
                } else {
                    v3 off = sub(x_last, from);
                    double rr = len(off);
                    v3 dd = mul(off, 1.0 / rr);
                    double ch = (1 - K(s, 8, 2)) * tree_pdf(s, emitter, from, seen.n);
                    uint32_t code = emitter_code(s, emitter);
                    double lpdf = (code & 0x80000000u) ? triangle_hit_pdf(s, code & ~0x80000000u, dd, rr)
                                                       : light_hit_pdf(s, code, from, dd, rr);
                    sample_ratio = ch * lpdf / cbp;
                }
            }
        }
        origin = offset_origin(x, side > 0 ? geo : neg(geo));
        vertex = x;
        dir = direction;
        ray_pdf = light_pdf;
        mis[0] *= last; mis[1] = density; mis[2] = fabs(side); mis[3] = sample_ratio;
    }
}

typedef struct { const oscene *s; uint32_t sample, first; size_t begin, end; double *splat; } ljob;

static void *light_run(void *arg) {
    ljob *j = (ljob *)arg;
    for (size_t i = j->begin; i < j->end; i++) trace_light(j->s, j->first + (uint32_t)i, j->sample, j->splat);
    return NULL;
}

/* Every sample's light paths (light_count = K_CAMERA.x a sample, in the slots
 * after the one band's camera paths), splatted into `splat` (3 doubles a
 * pixel); returns the samples that traced them. */
static uint32_t splat_light_paths(const oscene *s, int threads, double *splat) {
    uint32_t count = (uint32_t)K(s, 14, 0);
    if (!tracing_on(s) || count == 0) return 0;
    size_t pixels = (size_t)s->width * s->height;
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    double *films = (double *)calloc((size_t)threads * pixels * 3, sizeof(double));
    if (!films) return 0;
    pthread_t th[64];
    ljob jobs[64];
    for (uint32_t sample = 0; sample < s->samples; sample++) {
        for (int k = 0; k < threads; k++) {
            jobs[k] = (ljob){s, sample, (uint32_t)pixels, (size_t)count * k / threads, (size_t)count * (k + 1) / threads,
                             films + (size_t)k * pixels * 3};
            pthread_create(&th[k], NULL, light_run, &jobs[k]);
        }
        for (int k = 0; k < threads; k++) pthread_join(th[k], NULL);
    }
    for (int k = 0; k < threads; k++)
        for (size_t i = 0; i < pixels * 3; i++) splat[i] += films[(size_t)k * pixels * 3 + i];
    free(films);
    return s->samples;
}

typedef struct { const oscene *s; const uint32_t *pixels; size_t n; double *out; size_t begin, end; } rjob;

static void *render_run(void *arg) {
    rjob *j = (rjob *)arg;
    for (size_t i = j->begin; i < j->end; i++) {
        uint32_t pixel = j->pixels[i];
        v3 sum = V(0, 0, 0);
        int nan = 0;
        for (uint32_t k = 0; k < j->s->samples; k++) sum = add(sum, trace_path(j->s, pixel, k, &nan, NULL, NULL));
        v3 mean = mul(sum, 1.0 / fmax((double)j->s->samples, 1.0));
        j->out[i * 3] = mean.x; j->out[i * 3 + 1] = mean.y; j->out[i * 3 + 2] = mean.z;
    }
    return NULL;
}

/* film_convert.comp:36-53 for the listed pixels whose camera means are in
 * `out`: the light paths' splats over the samples that traced them added,
 * then (spectral) XYZ under E to ACEScg. */
static void finish_pixels(const oscene *s, int threads, const uint32_t *pixels, size_t n, double *out) {
    size_t count = (size_t)s->width * s->height;
    double *splat = tracing_on(s) ? (double *)calloc(count * 3, sizeof(double)) : NULL;
    uint32_t light_samples = splat ? splat_light_paths(s, threads, splat) : 0;
    for (size_t i = 0; i < n; i++) {
        v3 mean = V(out[i * 3], out[i * 3 + 1], out[i * 3 + 2]);
        if (light_samples) {
            size_t px = pixels[i];
            mean = add(mean, mul(V(splat[px * 3], splat[px * 3 + 1], splat[px * 3 + 2]), 1.0 / light_samples));
        }
        if (s->spectral) {
            v3 r0 = Kv(s, 3), r1 = Kv(s, 4), r2 = Kv(s, 5);
            mean = V(dot(r0, mean), dot(r1, mean), dot(r2, mean));
        }
        out[i * 3] = mean.x; out[i * 3 + 1] = mean.y; out[i * 3 + 2] = mean.z;
    }
    free(splat);
}

/* A guided render's pass over every pixel for one sample: film values and
 * each path's training records, the guide read only. */
typedef struct { const oscene *s; uint32_t sample; size_t begin, end; double *film; orecord *recs; double *found; } gjob;

static void *guided_run(void *arg) {
    gjob *j = (gjob *)arg;
    for (size_t px = j->begin; px < j->end; px++) {
        int nan = 0;
        v3 v = trace_path(j->s, (uint32_t)px, j->sample, &nan, &j->recs[px * GUIDE_RING], &j->found[px]);
        j->film[px * 3] += v.x; j->film[px * 3 + 1] += v.y; j->film[px * 3 + 2] += v.z;
    }
    return NULL;
}

static void setup_scene(oscene *s, double fr[16], const float *positions, const uint32_t *indices,
                        const uint32_t *shading, size_t triangles, const void *nodes, size_t node_count,
                        const float *lights, const float *constants, const float *materials, const uint32_t *emitters,
                        const float *frame, uint32_t width, uint32_t height, uint32_t samples, uint32_t max_bounces,
                        uint32_t spectral, uint32_t seed, uint32_t through, const float *uvs, const uint32_t *programs,
                        size_t program_words, const float *textures, const uint32_t *texture_table,
                        size_t texture_count);

/* Renders a guided image (guiding 1, camera paths): sample by sample over
 * every pixel, each sample's paths then training the grid in pixel order
 * (finish.comp; the kernels' float atomics add in another order) and the fit
 * running after it (guide_fit.comp, integrator.lucb:815-817), as one band a
 * sample does. Returns the listed pixels as or_render does. */
int or_render_guided(const float *positions, const uint32_t *indices, const uint32_t *shading, size_t triangles,
                     const void *nodes, size_t node_count, const float *lights, const float *constants,
                     const float *materials, const uint32_t *emitters, const float *frame, uint32_t width,
                     uint32_t height, uint32_t samples, uint32_t max_bounces, uint32_t spectral, uint32_t seed,
                     uint32_t through, const uint32_t *pixels, size_t n, int threads, double *out, const float *uvs,
                     const uint32_t *programs, size_t program_words, const float *textures,
                     const uint32_t *texture_table, size_t texture_count, uint32_t *keys_out, float *lobes_out,
                     float *stats_out) {
    oscene s;
    double fr[16];
    setup_scene(&s, fr, positions, indices, shading, triangles, nodes, node_count, lights, constants, materials,
                emitters, frame, width, height, samples, max_bounces, spectral, seed, through, uvs, programs,
                program_words, textures, texture_table, texture_count);
    oguide g;
    g.keys = (uint32_t *)calloc(GUIDE_CELLS, 4);
    g.lobes = (float *)calloc((size_t)GUIDE_CELLS * (GUIDE_LOBES + 1) * 4, 4);
    g.stats = (float *)calloc((size_t)GUIDE_CELLS * GUIDE_LOBES * 5, 4);
    size_t pixel_count = (size_t)width * height;
    double *film = (double *)calloc(pixel_count * 3, sizeof(double));
    orecord *recs = (orecord *)calloc(pixel_count * GUIDE_RING, sizeof(orecord));
    double *found = (double *)calloc(pixel_count, sizeof(double));
    if (!g.keys || !g.lobes || !g.stats || !film || !recs || !found) {
        free(g.keys); free(g.lobes); free(g.stats); free(film); free(recs); free(found);
        return -1;
    }
    s.guide = &g;
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    gjob jobs[64];
    for (uint32_t sample = 0; sample < samples; sample++) {
        for (int k = 0; k < threads; k++) {
            jobs[k] = (gjob){&s, sample, pixel_count * (size_t)k / (size_t)threads,
                             pixel_count * (size_t)(k + 1) / (size_t)threads, film, recs, found};
            pthread_create(&th[k], NULL, guided_run, &jobs[k]);
        }
        for (int k = 0; k < threads; k++) pthread_join(th[k], NULL);
        for (size_t px = 0; px < pixel_count; px++) {
            if (or_hash_u32((uint32_t)px * 9781u + sample) % GUIDE_TRAINING != 0u) continue;
            for (uint32_t k = 0; k < GUIDE_RING; k++) {
                const orecord *r = &recs[px * GUIDE_RING + k];
                if (r->key == 0) break;
                uint32_t cell = guide_find(&g, r->key, 1);
                if (cell == NONE) break;
                if (r->through <= 0) continue;
                /* finish.comp's found - v.z: the light past the vertex less
                 * what its blocked light sample would have added; nothing
                 * when no light came past it (as the kernels count it) */
                double gain = (found[px] - r->base) - r->blocked;
                double at_vertex = found[px] - gain;
                if (gain <= 1e-5 * fmax(found[px], at_vertex)) gain = 0.0;
                guide_train(&g, cell, r->dir, (double)(float)(gain / r->through * r->cosine), r->pdf);
            }
        }
        guide_fit(&g);
    }
    for (size_t i = 0; i < n; i++) {
        uint32_t px = pixels[i];
        v3 mean = mul(V(film[px * 3], film[px * 3 + 1], film[px * 3 + 2]), 1.0 / fmax((double)samples, 1.0));
        out[i * 3] = mean.x; out[i * 3 + 1] = mean.y; out[i * 3 + 2] = mean.z;
    }
    finish_pixels(&s, threads, pixels, n, out);
    if (keys_out) memcpy(keys_out, g.keys, (size_t)GUIDE_CELLS * 4);
    if (lobes_out) memcpy(lobes_out, g.lobes, (size_t)GUIDE_CELLS * (GUIDE_LOBES + 1) * 16);
    if (stats_out) memcpy(stats_out, g.stats, (size_t)GUIDE_CELLS * GUIDE_LOBES * 20);
    free(g.keys); free(g.lobes); free(g.stats); free(film); free(recs); free(found);
    return 0;
}

/* Renders the listed pixels (camera paths, light tracing and guiding off):
 * ACEScg, 3 doubles each. `frame`: the 16 packed floats. */
int or_render(const float *positions, const uint32_t *indices, const uint32_t *shading, size_t triangles,
              const void *nodes, size_t node_count, const float *lights, const float *constants,
              const float *materials, const uint32_t *emitters, const float *frame, uint32_t width, uint32_t height,
              uint32_t samples, uint32_t max_bounces, uint32_t spectral, uint32_t seed, uint32_t through,
              const uint32_t *pixels, size_t n, int threads, double *out, const float *uvs, const uint32_t *programs,
              size_t program_words, const float *textures, const uint32_t *texture_table, size_t texture_count) {
    oscene s;
    double fr[16];
    setup_scene(&s, fr, positions, indices, shading, triangles, nodes, node_count, lights, constants, materials,
                emitters, frame, width, height, samples, max_bounces, spectral, seed, through, uvs, programs,
                program_words, textures, texture_table, texture_count);
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    rjob jobs[64];
    for (int k = 0; k < threads; k++) {
        jobs[k] = (rjob){&s, pixels, n, out, n * (size_t)k / (size_t)threads, n * (size_t)(k + 1) / (size_t)threads};
        pthread_create(&th[k], NULL, render_run, &jobs[k]);
    }
    for (int k = 0; k < threads; k++) pthread_join(th[k], NULL);
    finish_pixels(&s, threads, pixels, n, out);
    return 0;
}

static void setup_scene(oscene *s, double fr[16], const float *positions, const uint32_t *indices,
                        const uint32_t *shading, size_t triangles, const void *nodes, size_t node_count,
                        const float *lights, const float *constants, const float *materials, const uint32_t *emitters,
                        const float *frame, uint32_t width, uint32_t height, uint32_t samples, uint32_t max_bounces,
                        uint32_t spectral, uint32_t seed, uint32_t through, const float *uvs, const uint32_t *programs,
                        size_t program_words, const float *textures, const uint32_t *texture_table,
                        size_t texture_count) {
    static double tables[T_SIZE];
    static int tables_ready = 0;
    static pthread_mutex_t lock = PTHREAD_MUTEX_INITIALIZER;
    pthread_mutex_lock(&lock);
    if (!tables_ready) { or_make_tables(tables); tables_ready = 1; }
    pthread_mutex_unlock(&lock);
    for (int k = 0; k < 16; k++) fr[k] = frame[k];
    memset(s, 0, sizeof(*s));
    s->positions = positions; s->indices = indices; s->shading = shading; s->triangles = triangles;
    s->nodes = (const onode *)nodes; s->node_count = node_count;
    s->lights = lights; s->constants = constants; s->materials = materials; s->emitters = emitters;
    s->frame = fr; s->width = width; s->height = height; s->samples = samples; s->max_bounces = max_bounces;
    s->spectral = spectral; s->seed = seed; s->through = through; s->tables = tables;
    s->uvs = uvs; s->programs = programs; s->program_words = uvs ? program_words : 0;
    s->textures = textures; s->texture_table = texture_table; s->texture_count = texture_count;
    s->guide = NULL;
}
