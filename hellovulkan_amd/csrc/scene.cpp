// scene.cpp — host-side camera/UBO construction and seeded synthetic scenes
// (include/hvx.h hvx_camera_build, include/hvx_scene.h).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/hvx.h"
#include "../../include/hvx_scene.h"

namespace {

struct D3 {
    double x, y, z;
};
D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
D3 normalize(D3 a) {
    const double l = std::sqrt(dot(a, a));
    return {a.x / l, a.y / l, a.z / l};
}

// column-major 4x4 in double: m[c*4 + r]
void matmul(const double* a, const double* b, double* o) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + r] * b[c * 4 + k];
            o[c * 4 + r] = s;
        }
}

// Gauss-Jordan inverse with partial pivoting; returns false if singular.
bool invert(const double* m, double* inv) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = m[c * 4 + r];
            a[r][4 + c] = (r == c) ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        if (a[piv][col] == 0.0) return false;
        if (piv != col)
            for (int c = 0; c < 8; ++c) std::swap(a[piv][c], a[col][c]);
        const double d = a[col][col];
        for (int c = 0; c < 8; ++c) a[col][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            const double f = a[r][col];
            if (f == 0.0) continue;
            for (int c = 0; c < 8; ++c) a[r][c] -= f * a[col][c];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[c * 4 + r] = a[r][4 + c];
    return true;
}

void to_float(const double* d, float* f, int n) {
    for (int i = 0; i < n; ++i) f[i] = (float)d[i];
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014).
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
// uniform double in [0, 1) of element i of stream k
inline double u01(uint64_t seed, uint64_t stream, uint64_t i) {
    return (double)(splitmix64(seed + (stream << 40) + i) >> 11) * (1.0 / 9007199254740992.0);
}

template <class F> void parallel_rows(uint32_t rows, int threads, F&& fn) {
    threads = std::max(1, std::min(threads, 64));
    if (threads == 1 || rows < 2) {
        for (uint32_t r = 0; r < rows; ++r) fn(r);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t]() {
            for (uint32_t r = (uint32_t)t; r < rows; r += (uint32_t)threads) fn(r);
        });
    for (auto& th : pool) th.join();
}

constexpr double kPi = 3.14159265358979323846;

}  // namespace

extern "C" {

int hvx_camera_build(const hvx_camera_desc* d, hvx_camera_ubo* camera, hvx_cf_ubo* cf, float inv_view_proj[16]) {
    if (!d) return HVX_ERR_NULL;
    if (!(d->znear > 0.0f) || !(d->zfar > d->znear) || !(d->fovy > 0.0f) || d->width == 0 || d->height == 0)
        return HVX_ERR_INVALID_ARG;
    const D3 eye = {d->eye[0], d->eye[1], d->eye[2]};
    const D3 tgt = {d->target[0], d->target[1], d->target[2]};
    // Camera::SetPositionAndTarget + ConstrainPitch
    const D3 dir = normalize(sub(tgt, eye));
    double pitch = std::asin(dir.y);
    const double yaw = std::atan2(dir.z, dir.x);
    pitch = std::min(std::max(pitch, -1.55334), 1.55334);
    // Camera::UpdateInternal (worldUp = +Y)
    const D3 front = normalize({std::cos(yaw) * std::cos(pitch), std::sin(pitch), std::sin(yaw) * std::cos(pitch)});
    const D3 right = normalize(cross(front, D3{0.0, 1.0, 0.0}));
    const D3 up = normalize(cross(right, front));
    // perspective, right-handed, depth [0, 1], then P[1][1] *= -1
    const double n = d->znear, f = d->zfar;
    const double aspect = (double)d->width / (double)d->height;
    const double th = std::tan((double)d->fovy / 2.0);
    double P[16] = {0};
    P[0 * 4 + 0] = 1.0 / (aspect * th);
    P[1 * 4 + 1] = -(1.0 / th);
    P[2 * 4 + 2] = f / (n - f);
    P[2 * 4 + 3] = -1.0;
    P[3 * 4 + 2] = -(f * n) / (f - n);
    // lookAt(eye, eye + front, up), right-handed
    const D3 s = normalize(cross(front, up));
    const D3 u = cross(s, front);
    double V[16] = {0};
    V[0 * 4 + 0] = s.x; V[1 * 4 + 0] = s.y; V[2 * 4 + 0] = s.z;
    V[0 * 4 + 1] = u.x; V[1 * 4 + 1] = u.y; V[2 * 4 + 1] = u.z;
    V[0 * 4 + 2] = -front.x; V[1 * 4 + 2] = -front.y; V[2 * 4 + 2] = -front.z;
    V[3 * 4 + 0] = -dot(s, eye);
    V[3 * 4 + 1] = -dot(u, eye);
    V[3 * 4 + 2] = dot(front, eye);
    V[3 * 4 + 3] = 1.0;
    double invP[16], PV[16], invPV[16];
    if (!invert(P, invP)) return HVX_ERR_INVALID_ARG;
    matmul(P, V, PV);
    if (!invert(PV, invPV)) return HVX_ERR_INVALID_ARG;
    if (camera) {
        std::memset(camera, 0, sizeof(*camera));
        to_float(P, camera->projection, 16);
        to_float(V, camera->view, 16);
        camera->position[0] = d->eye[0];
        camera->position[1] = d->eye[1];
        camera->position[2] = d->eye[2];
        camera->position[3] = 1.0f;
        camera->camera_near = d->znear;
        camera->camera_far = d->zfar;
    }
    if (cf) {
        if (d->slice_count_x == 0 || d->slice_count_y == 0 || d->slice_count_z == 0) return HVX_ERR_INVALID_ARG;
        std::memset(cf, 0, sizeof(*cf));
        to_float(invP, cf->camera_inverse_projection, 16);
        to_float(V, cf->camera_view, 16);
        cf->screen_size[0] = (float)d->width;
        cf->screen_size[1] = (float)d->height;
        // Camera::GetClusterForwardUBO (Camera.cpp:178-197)
        const double Z = d->slice_count_z;
        const double l2fn = std::log2(f / n), l2n = std::log2(n);
        cf->slice_scaling = (float)(Z / l2fn);
        cf->slice_bias = (float)(-(Z * l2n / l2fn));
        cf->camera_near = d->znear;
        cf->camera_far = d->zfar;
        cf->slice_count_x = d->slice_count_x;
        cf->slice_count_y = d->slice_count_y;
        cf->slice_count_z = d->slice_count_z;
    }
    if (inv_view_proj) to_float(invPV, inv_view_proj, 16);
    return HVX_OK;
}

int hvx_scene_lights(uint64_t seed, uint32_t n, hvx_light* out) {
    if (!out && n) return HVX_ERR_NULL;
    for (uint32_t i = 0; i < n; ++i) {
        hvx_light& l = out[i];
        std::memset(&l, 0, sizeof(l));
        l.position[0] = (float)(-15.0 + 30.0 * u01(seed, 0, i));
        l.position[1] = (float)(0.0 + 12.0 * u01(seed, 1, i));
        l.position[2] = (float)(-7.0 + 14.0 * u01(seed, 2, i));
        l.position[3] = 1.0f;
        l.color[0] = (float)u01(seed, 3, i);
        l.color[1] = (float)u01(seed, 4, i);
        l.color[2] = (float)u01(seed, 5, i);
        l.color[3] = 1.0f;
        l.radius = (float)(0.5 + 2.5 * u01(seed, 6, i));
    }
    return HVX_OK;
}

int hvx_scene_fragments(uint64_t seed, const hvx_camera_desc* cam, const float ivp[16], float* depth, float* nx,
                        float* ny, float* nz, float* ar, float* ag, float* ab, float* roughness, uint8_t* metallic,
                        int threads) {
    if (!cam || !ivp || !depth || !nx || !ny || !nz || !ar || !ag || !ab || !roughness || !metallic)
        return HVX_ERR_NULL;
    const uint32_t W = cam->width, H = cam->height;
    if (W == 0 || H == 0) return HVX_ERR_INVALID_ARG;
    const double n = cam->znear, f = cam->zfar;
    const D3 eye = {cam->eye[0], cam->eye[1], cam->eye[2]};
    parallel_rows(H, threads, [&](uint32_t y) {
        for (uint32_t x = 0; x < W; ++x) {
            const uint64_t i = (uint64_t)y * W + x;
            // linear depth log-uniform in [near, far) -> NDC z (inverse of
            // LinearDepth, Scene.frag:55-58)
            const double lin = n * std::pow(f / n, u01(seed, 10, i));
            double z = f * (lin - n) / (lin * (f - n));
            float zf = (float)z;
            if (!(zf < 1.0f)) zf = std::nextafter(1.0f, 0.0f);
            if (zf < 0.0f) zf = 0.0f;
            depth[i] = zf;
            // world position of the fragment (for the facing test only)
            const double cx = ((x + 0.5) / W) * 2.0 - 1.0, cy = ((y + 0.5) / H) * 2.0 - 1.0;
            double wh[4];
            for (int r = 0; r < 4; ++r)
                wh[r] = ivp[r] * cx + ivp[4 + r] * cy + ivp[8 + r] * (double)zf + ivp[12 + r];
            const D3 wp = {wh[0] / wh[3], wh[1] / wh[3], wh[2] / wh[3]};
            const D3 v = sub(eye, wp);
            // uniform direction on the sphere, flipped into the hemisphere facing the eye
            const double cz = 2.0 * u01(seed, 11, i) - 1.0;
            const double ph = 2.0 * kPi * u01(seed, 12, i);
            const double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
            D3 nn = {sz * std::cos(ph), sz * std::sin(ph), cz};
            if (dot(nn, v) < 0.0) nn = {-nn.x, -nn.y, -nn.z};
            nx[i] = (float)nn.x;
            ny[i] = (float)nn.y;
            nz[i] = (float)nn.z;
            ar[i] = (float)u01(seed, 13, i);
            ag[i] = (float)u01(seed, 14, i);
            ab[i] = (float)u01(seed, 15, i);
            roughness[i] = (float)(0.05 + 0.95 * u01(seed, 16, i));
            metallic[i] = u01(seed, 17, i) < 0.3 ? 255 : 0;
        }
    });
    return HVX_OK;
}

int hvx_scene_equirect(uint64_t seed, uint32_t w, uint32_t h, float* rgba, int threads) {
    if (!rgba) return HVX_ERR_NULL;
    if (w == 0 || h == 0) return HVX_ERR_INVALID_ARG;
    struct Spot {
        D3 d;
        double peak, width, c[3];
    };
    std::vector<Spot> spots(32);
    for (uint32_t k = 0; k < 32; ++k) {
        const double cz = 2.0 * u01(seed, 100, k) - 1.0, ph = 2.0 * kPi * u01(seed, 101, k);
        const double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
        // log-normal peak: exp(mu + sigma * N(0,1)), Box-Muller normal
        const double u1 = std::max(u01(seed, 102, k), 1e-300), u2 = u01(seed, 103, k);
        const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
        spots[k].d = {sz * std::cos(ph), cz, sz * std::sin(ph)};
        spots[k].peak = std::exp(3.0 + 1.0 * g);
        spots[k].width = 0.03 + 0.12 * u01(seed, 104, k);
        for (int c = 0; c < 3; ++c) spots[k].c[c] = 0.5 + 0.5 * u01(seed, 105 + c, k);
    }
    const double sel = 35.0 * kPi / 180.0, saz = 40.0 * kPi / 180.0;
    const D3 sun = {std::cos(sel) * std::cos(saz), std::sin(sel), std::cos(sel) * std::sin(saz)};
    const double sun_r = 0.02;  // angular radius (rad)
    parallel_rows(h, threads, [&](uint32_t y) {
        for (uint32_t x = 0; x < w; ++x) {
            // inverse of Equirect2Cube.frag:87-94: s = 0.5 + atan2(z, x)/(2 pi), t = 1 - acos(y)/pi
            const double s = (x + 0.5) / w, t = (y + 0.5) / h;
            const double phi = (s - 0.5) * 2.0 * kPi, theta = (1.0 - t) * kPi;
            const D3 d = {std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi)};
            double c[3];
            if (d.y >= 0.0) {
                const double k = std::sqrt(d.y);
                const double hz[3] = {1.2, 1.1, 0.9}, zn[3] = {0.15, 0.3, 0.8};
                for (int i = 0; i < 3; ++i) c[i] = hz[i] * (1.0 - k) + zn[i] * k;
            } else {
                const double gr[3] = {0.12, 0.10, 0.08};
                for (int i = 0; i < 3; ++i) c[i] = gr[i];
            }
            const double cs = std::min(1.0, std::max(-1.0, dot(d, sun)));
            if (std::acos(cs) < sun_r)
                for (int i = 0; i < 3; ++i) c[i] += 1.0e4;
            for (const Spot& sp : spots) {
                const double ca = std::min(1.0, std::max(-1.0, dot(d, sp.d)));
                const double a = std::acos(ca) / sp.width;
                if (a < 6.0) {
                    const double e = sp.peak * std::exp(-0.5 * a * a);
                    for (int i = 0; i < 3; ++i) c[i] += e * sp.c[i];
                }
            }
            float* o = rgba + ((size_t)y * w + x) * 4;
            o[0] = (float)c[0];
            o[1] = (float)c[1];
            o[2] = (float)c[2];
            o[3] = 1.0f;
        }
    });
    return HVX_OK;
}

}  // extern "C"
