/*
 * pmx_scene.cpp -- seeded synthetic multi-view scenes for PatchMatch.
 *
 * The reference ships no data (its inputs are NVM/Middlebury files read in
 * src/image.cpp:64-245), so every workload here is rendered: a
 * piecewise-planar world of textured rectangles and spheres in front of a
 * textured background, seen by one reference camera (K = [f 0 W/2; 0 f H/2;
 * 0 0 1], R = I, C = 0) and V source cameras placed on a random baseline and
 * turned towards the scene.  The texture is 3D value noise evaluated at the
 * surface point, so it is consistent across views; per-view Gaussian pixel
 * noise is added after rendering.  Pixel (x, y) is rendered along the ray
 * through the continuous image point (x + 0.5, y + 0.5).
 *
 * Host-only code (OpenMP over image rows); deterministic for any thread
 * count.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

typedef struct pmx_scene_spec {
    int W, H;            /* image size of every view */
    int V;               /* number of source views */
    int n_planes;        /* textured rectangles */
    int n_spheres;       /* textured spheres */
    double depth_min;    /* scene depth range along the reference axis */
    double depth_max;
    double focal;        /* f in pixels, principal point at the centre */
    double base_min;     /* source baseline length range */
    double base_max;
    double rot_max_deg;  /* bound on source yaw / pitch */
    double noise_sigma;  /* per-pixel Gaussian noise */
    uint32_t scene_seed;
    uint32_t noise_seed;
} pmx_scene_spec;

}

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double range(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        double u1 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 normalize(V3 a) { return (1.0 / std::sqrt(dot(a, a))) * a; }

struct Prim {
    int kind;      /* 0 rectangle, 1 sphere, 2 infinite plane */
    V3 c;          /* centre */
    V3 n, e1, e2;  /* plane normal and in-plane axes */
    double a, b;   /* rectangle half sizes or sphere radius (a) */
    V3 tex_off;    /* texture offset */
};

inline uint32_t hash3(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    uint64_t h = (uint64_t)(uint32_t)x * 0x8DA6B343ULL ^ (uint64_t)(uint32_t)y * 0xD8163841ULL ^
                 (uint64_t)(uint32_t)z * 0xCB1AB31FULL ^ ((uint64_t)seed << 32);
    h ^= h >> 33;
    h *= 0xFF51AFD7ED558CCDULL;
    h ^= h >> 33;
    h *= 0xC4CEB9FE1A85EC53ULL;
    h ^= h >> 33;
    return (uint32_t)h;
}

inline double lattice(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    return (double)(hash3(x, y, z, seed) >> 8) * (1.0 / 16777216.0);
}

inline double fade(double t) { return t * t * (3.0 - 2.0 * t); }

double value_noise(V3 p, uint32_t seed) {
    const double fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
    const int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    const double tx = fade(p.x - fx), ty = fade(p.y - fy), tz = fade(p.z - fz);
    double acc = 0.0;
    for (int c = 0; c < 8; c++) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        const double w = (dx ? tx : 1.0 - tx) * (dy ? ty : 1.0 - ty) * (dz ? tz : 1.0 - tz);
        acc += w * lattice(ix + dx, iy + dy, iz + dz, seed);
    }
    return acc;
}

/* Returns the nearest hit distance along the ray (origin o, unit dir d);
 * sets *hit_id. */
double cast(const std::vector<Prim>& prims, V3 o, V3 d, int* hit_id) {
    double best = INFINITY;
    int bid = -1;
    for (size_t i = 0; i < prims.size(); i++) {
        const Prim& P = prims[i];
        double t = INFINITY;
        if (P.kind == 1) {
            const V3 oc = o - P.c;
            const double b = dot(oc, d), cc = dot(oc, oc) - P.a * P.a;
            const double disc = b * b - cc;
            if (disc >= 0.0) {
                const double sq = std::sqrt(disc);
                double t0 = -b - sq;
                if (t0 <= 1e-9) t0 = -b + sq;
                if (t0 > 1e-9) t = t0;
            }
        } else {
            const double den = dot(P.n, d);
            if (std::fabs(den) > 1e-12) {
                const double t0 = dot(P.n, P.c - o) / den;
                if (t0 > 1e-9) {
                    if (P.kind == 2) {
                        t = t0;
                    } else {
                        const V3 q = (o + t0 * d) - P.c;
                        if (std::fabs(dot(q, P.e1)) <= P.a && std::fabs(dot(q, P.e2)) <= P.b) t = t0;
                    }
                }
            }
        }
        if (t < best) {
            best = t;
            bid = (int)i;
        }
    }
    *hit_id = bid;
    return best;
}

void rot_xyz(double yaw, double pitch, double roll, double* R) {
    /* Rz(roll) * Rx(pitch) * Ry(yaw) */
    const double cy = std::cos(yaw), sy = std::sin(yaw);
    const double cp = std::cos(pitch), sp = std::sin(pitch);
    const double cr = std::cos(roll), sr = std::sin(roll);
    const double Ry[9] = {cy, 0, sy, 0, 1, 0, -sy, 0, cy};
    const double Rx[9] = {1, 0, 0, 0, cp, -sp, 0, sp, cp};
    const double Rz[9] = {cr, -sr, 0, sr, cr, 0, 0, 0, 1};
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += Rx[i * 3 + k] * Ry[k * 3 + j];
            t[i * 3 + j] = s;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += Rz[i * 3 + k] * t[k * 3 + j];
            R[i * 3 + j] = s;
        }
}

} // namespace

extern "C" {

/* images: (V+1) layers of H*W float32, layer 0 = reference.
 * gt_depth: H*W z-depth of the reference view (may be NULL).
 * cams: (V+1) * 21 doubles, K (9, row-major), R (9, row-major), C (3). */
int pmx_scene_render(const pmx_scene_spec* sp, float* images, float* gt_depth, double* cams) {
    if (!sp || !images || !cams || sp->W <= 0 || sp->H <= 0 || sp->V < 1) return -1;
    const int W = sp->W, H = sp->H, NV = sp->V + 1;
    const double f = sp->focal, cx = 0.5 * W, cy = 0.5 * H;
    const double dmin = sp->depth_min, dmax = sp->depth_max;
    const double dmid = 0.5 * (dmin + dmax);
    Rng rng{(uint64_t)sp->scene_seed * 0x100000001B3ULL + 17};

    std::vector<Prim> prims;
    /* Background: a slightly slanted plane near the far end of the range. */
    {
        Prim bg{};
        bg.kind = 2;
        const double tilt = rng.range(-0.15, 0.15);
        bg.n = normalize(V3{std::sin(tilt), rng.range(-0.1, 0.1), -1.0});
        bg.c = V3{0, 0, 0.92 * dmax + 0.08 * dmin};
        bg.tex_off = V3{rng.range(0, 100), rng.range(0, 100), rng.range(0, 100)};
        prims.push_back(bg);
    }
    const double half_w_ndc = 0.5 * W / f; /* image half-width at unit depth */
    for (int i = 0; i < sp->n_planes; i++) {
        Prim P{};
        P.kind = 0;
        for (int attempt = 0; attempt < 64; attempt++) {
            const double u = rng.range(0.1 * W, 0.9 * W), v = rng.range(0.1 * H, 0.9 * H);
            const double d = rng.range(dmin, dmax);
            P.c = V3{(u - cx) / f * d, (v - cy) / f * d, d};
            const double ax = rng.range(-0.7, 0.7), ay = rng.range(-0.7, 0.7);
            P.n = normalize(V3{std::sin(ay), std::sin(ax), -1.0});
            const V3 up = std::fabs(P.n.y) < 0.9 ? V3{0, 1, 0} : V3{1, 0, 0};
            V3 e1 = normalize(cross(up, P.n));
            V3 e2 = cross(P.n, e1);
            const double th = rng.range(0, 3.141592653589793);
            P.e1 = std::cos(th) * e1 + std::sin(th) * e2;
            P.e2 = cross(P.n, P.e1);
            P.a = rng.range(0.15, 0.45) * half_w_ndc * d;
            P.b = rng.range(0.15, 0.45) * half_w_ndc * d;
            /* keep the corners inside the depth range */
            bool ok = true;
            for (int k = 0; k < 4; k++) {
                const double sa = (k & 1) ? 1.0 : -1.0, sb = (k & 2) ? 1.0 : -1.0;
                const V3 q = P.c + (sa * P.a) * P.e1 + (sb * P.b) * P.e2;
                if (q.z < dmin || q.z > dmax) ok = false;
            }
            if (ok) break;
        }
        P.tex_off = V3{rng.range(0, 100), rng.range(0, 100), rng.range(0, 100)};
        prims.push_back(P);
    }
    for (int i = 0; i < sp->n_spheres; i++) {
        Prim P{};
        P.kind = 1;
        const double u = rng.range(0.15 * W, 0.85 * W), v = rng.range(0.15 * H, 0.85 * H);
        const double d = rng.range(dmin, dmax);
        P.a = std::fmin(rng.range(0.06, 0.14) * half_w_ndc * d, 0.45 * (d - dmin) + 1e-3);
        const double dc = d + P.a;
        P.c = V3{(u - cx) / f * dc, (v - cy) / f * dc, dc};
        P.tex_off = V3{rng.range(0, 100), rng.range(0, 100), rng.range(0, 100)};
        prims.push_back(P);
    }

    /* Cameras. */
    std::vector<double> K(9), Rall((size_t)NV * 9), Call((size_t)NV * 3);
    const double Kv[9] = {f, 0, cx, 0, f, cy, 0, 0, 1};
    const double rmax = sp->rot_max_deg * 3.141592653589793 / 180.0;
    for (int c = 0; c < NV; c++) {
        double* R = &Rall[(size_t)c * 9];
        double* C = &Call[(size_t)c * 3];
        if (c == 0) {
            const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            std::memcpy(R, I, sizeof(I));
            C[0] = C[1] = C[2] = 0.0;
        } else {
            const double th = rng.range(0, 6.283185307179586);
            const double b = rng.range(sp->base_min, sp->base_max);
            V3 dir = normalize(V3{std::cos(th), 0.5 * std::sin(th), rng.range(-0.2, 0.2)});
            C[0] = b * dir.x;
            C[1] = b * dir.y;
            C[2] = b * dir.z;
            /* turn towards the scene centre, then perturb */
            const V3 to = normalize(V3{-C[0], -C[1], dmid - C[2]});
            double yaw = std::atan2(to.x, to.z) + rng.range(-rmax / 3, rmax / 3);
            double pitch = -std::asin(to.y) + rng.range(-rmax / 3, rmax / 3);
            const double roll = rng.range(-rmax / 4, rmax / 4);
            yaw = std::fmax(-rmax, std::fmin(rmax, yaw));
            pitch = std::fmax(-rmax, std::fmin(rmax, pitch));
            /* camera forward (world) = R^T e_z; choose R = (Rz Rx Ry)^T-style
             * so that yaw turns the forward axis towards +x. */
            double Rw[9];
            rot_xyz(yaw, pitch, roll, Rw);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) R[i * 3 + j] = Rw[j * 3 + i];
        }
        double* out = cams + (size_t)c * 21;
        std::memcpy(out, Kv, sizeof(Kv));
        std::memcpy(out + 9, R, 9 * sizeof(double));
        std::memcpy(out + 18, C, 3 * sizeof(double));
    }

    const double cell = 3.0 * dmid / f; /* texture lattice ~3 px at mid depth */
    const uint32_t tseed = sp->scene_seed * 2654435761u + 99u;

#pragma omp parallel for schedule(dynamic, 4)
    for (int row = 0; row < NV * H; row++) {
        const int c = row / H, y = row % H;
        const double* R = &Rall[(size_t)c * 9];
        const double* C = &Call[(size_t)c * 3];
        const V3 o{C[0], C[1], C[2]};
        Rng nrng{((uint64_t)sp->noise_seed << 32) ^ ((uint64_t)c << 24) ^ (uint64_t)y};
        for (int x = 0; x < W; x++) {
            /* camera ray through (x+0.5, y+0.5), rotated to world: R^T k */
            const V3 kc{(x + 0.5 - cx) / f, (y + 0.5 - cy) / f, 1.0};
            const V3 dw{R[0] * kc.x + R[3] * kc.y + R[6] * kc.z,
                        R[1] * kc.x + R[4] * kc.y + R[7] * kc.z,
                        R[2] * kc.x + R[5] * kc.y + R[8] * kc.z};
            const V3 d = normalize(dw);
            int id = -1;
            const double t = cast(prims, o, d, &id);
            double val = 0.5;
            if (id >= 0) {
                const V3 p = o + t * d;
                const V3 q = p + prims[id].tex_off;
                val = 0.40 * value_noise((1.0 / (0.5 * cell)) * q, tseed) +
                      0.35 * value_noise((1.0 / cell) * q, tseed + 1) +
                      0.25 * value_noise((1.0 / (3.0 * cell)) * q, tseed + 2);
                if (c == 0 && gt_depth) gt_depth[(size_t)y * W + x] = (float)p.z;
            } else if (c == 0 && gt_depth) {
                gt_depth[(size_t)y * W + x] = (float)dmax;
            }
            val += sp->noise_sigma * nrng.normal();
            images[((size_t)c * H + y) * W + x] = (float)val;
        }
    }
    return 0;
}

} // extern "C"
