// Synthetic depth input for the KinectFusion frame loop (host-only, libazu_scene.so).
//
// Scene: a sphere (r = 0.3 m, centre (0, 0, 0.8)) in front of the plane z = 1.2 m
// (solid behind it), optionally with axis-aligned boxes (edges U[0.05, 0.20] m,
// centres x, y ~ U[-0.45, 0.45], z ~ U[0.85, 1.15], from std::mt19937_64(seed + 1000003)).
// Depth is the Z-depth t * dir_z of the nearest hit along the pixel-centre ray through
// pinhole intrinsics, plus N(0, (0.0012 z^2)^2) m noise drawn per pixel in row-major
// order from std::mt19937_64(noise_seed), rounded to integer millimetres; 0 = no hit.
//
// Poses: the camera orbits the sphere centre about the world y axis at radius 0.8 m,
// 2 degrees per frame, ping-pong within +-60 degrees (0 -> 60 -> -60 -> 0, period 120
// frames) so it never passes behind the plane. The camera looks at the centre; at
// angle 0 it sits at the world origin looking down +z. Optional jitter composes a
// camera-frame perturbation [R(w) | t] with t ~ N(0, 5 mm) and w ~ N(0, 0.5 deg) per
// axis, from std::mt19937_64(seed * 7919 + frame).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#define AZU_SCENE_API extern "C" __attribute__((visibility("default")))

namespace {
struct Box {
    double lo[3], hi[3];
};
struct Scene {
    std::vector<Box> boxes;
};

constexpr double kSphereC[3] = {0.0, 0.0, 0.8};
constexpr double kSphereR = 0.3;
constexpr double kPlaneZ = 1.2;
constexpr double kPi = 3.14159265358979323846;

Scene makeScene(int nboxes, uint64_t seed) {
    Scene s;
    std::mt19937_64 rng(seed + 1000003ull);
    std::uniform_real_distribution<double> edge(0.05, 0.20), cxy(-0.45, 0.45), cz(0.85, 1.15);
    for (int i = 0; i < nboxes; ++i) {
        Box b;
        double e[3], c[3];
        for (double& v : e) v = edge(rng);
        c[0] = cxy(rng);
        c[1] = cxy(rng);
        c[2] = cz(rng);
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = c[a] - 0.5 * e[a];
            b.hi[a] = c[a] + 0.5 * e[a];
        }
        s.boxes.push_back(b);
    }
    return s;
}

// Nearest positive ray parameter of o + t d against the scene, or -1.
double hit(const Scene& s, const double o[3], const double d[3]) {
    double best = -1.0;
    auto take = [&](double t) {
        if (t > 1e-9 && (best < 0 || t < best)) best = t;
    };
    // sphere
    {
        const double oc[3] = {o[0] - kSphereC[0], o[1] - kSphereC[1], o[2] - kSphereC[2]};
        const double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double b = 2 * (oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2]);
        const double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - kSphereR * kSphereR;
        const double disc = b * b - 4 * a * c;
        if (disc >= 0) {
            const double sq = std::sqrt(disc);
            const double t0 = (-b - sq) / (2 * a), t1 = (-b + sq) / (2 * a);
            if (t0 > 1e-9) take(t0);
            else if (t1 > 1e-9) take(t1);
        }
    }
    // plane z = kPlaneZ, solid behind: a camera in front sees its front face
    if (d[2] != 0) {
        const double t = (kPlaneZ - o[2]) / d[2];
        if (o[2] < kPlaneZ) take(t);
        else take(1e-8);  // inside the solid half-space: immediate hit
    }
    // boxes (slab test)
    for (const Box& b : s.boxes) {
        double t0 = -1e300, t1 = 1e300;
        bool miss = false;
        for (int a = 0; a < 3; ++a) {
            if (d[a] == 0) {
                if (o[a] < b.lo[a] || o[a] > b.hi[a]) miss = true;
                continue;
            }
            double ta = (b.lo[a] - o[a]) / d[a], tb = (b.hi[a] - o[a]) / d[a];
            if (ta > tb) std::swap(ta, tb);
            t0 = std::max(t0, ta);
            t1 = std::min(t1, tb);
        }
        if (miss || t0 > t1) continue;
        if (t0 > 1e-9) take(t0);
        else if (t1 > 1e-9) take(1e-8);
    }
    return best;
}

double orbitAngleDeg(int frame) {
    const int u = ((frame % 120) + 120) % 120;
    const int a = (u <= 30) ? u : (u <= 90 ? 60 - u : u - 120);
    return 2.0 * a;
}

void rodrigues(const double w[3], double R[9]) {
    const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (th < 1e-15) {
        const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        std::memcpy(R, I, sizeof(I));
        return;
    }
    const double k[3] = {w[0] / th, w[1] / th, w[2] / th};
    const double c = std::cos(th), s = std::sin(th), oc = 1 - c;
    R[0] = c + oc * k[0] * k[0];
    R[1] = oc * k[0] * k[1] - s * k[2];
    R[2] = oc * k[0] * k[2] + s * k[1];
    R[3] = oc * k[1] * k[0] + s * k[2];
    R[4] = c + oc * k[1] * k[1];
    R[5] = oc * k[1] * k[2] - s * k[0];
    R[6] = oc * k[2] * k[0] - s * k[1];
    R[7] = oc * k[2] * k[1] + s * k[0];
    R[8] = c + oc * k[2] * k[2];
}
}  // namespace

// Camera pose of `frame` as a column-major 4x4 world-from-camera matrix, in double.
AZU_SCENE_API void azu_scene_pose_d(int frame, int jitter, uint64_t seed, double* out) {
    const double th = orbitAngleDeg(frame) * kPi / 180.0;
    const double c = std::cos(th), s = std::sin(th);
    double R[9] = {c, 0, s, 0, 1, 0, -s, 0, c};  // row-major R_y(th)
    double t[3] = {kSphereC[0] - 0.8 * s, kSphereC[1], kSphereC[2] - 0.8 * c};
    if (jitter) {
        std::mt19937_64 rng(seed * 7919ull + static_cast<uint64_t>(frame));
        std::normal_distribution<double> n01(0.0, 1.0);
        double dt[3], w[3];
        for (double& v : dt) v = 0.005 * n01(rng);
        for (double& v : w) v = (0.5 * kPi / 180.0) * n01(rng);
        double Rj[9];
        rodrigues(w, Rj);
        double Rn[9], tn[3];
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) Rn[i * 3 + j] = R[i * 3 + 0] * Rj[0 + j] + R[i * 3 + 1] * Rj[3 + j] + R[i * 3 + 2] * Rj[6 + j];
            tn[i] = R[i * 3 + 0] * dt[0] + R[i * 3 + 1] * dt[1] + R[i * 3 + 2] * dt[2] + t[i];
        }
        std::memcpy(R, Rn, sizeof(R));
        std::memcpy(t, tn, sizeof(t));
    }
    for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) out[cc * 4 + r] = R[r * 3 + cc];
    out[12] = t[0];
    out[13] = t[1];
    out[14] = t[2];
    out[3] = out[7] = out[11] = 0;
    out[15] = 1;
}

AZU_SCENE_API void azu_scene_pose(int frame, int jitter, uint64_t seed, float* out) {
    double d[16];
    azu_scene_pose_d(frame, jitter, seed, d);
    for (int i = 0; i < 16; ++i) out[i] = static_cast<float>(d[i]);
}

// Render one frame: pose (float, world-from-camera, column-major), K = {fx, fy, cx, cy}.
// out_mm: W*H uint16 millimetres. true_depth (optional): noise-free Z-depth in metres.
AZU_SCENE_API void azu_scene_render(int nboxes, uint64_t box_seed, const float* pose, const float* K, int W,
                                    int H, uint64_t noise_seed, int noise_on, uint16_t* out_mm,
                                    double* true_depth) {
    const Scene sc = makeScene(nboxes, box_seed);
    double R[9], o[3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) R[r * 3 + c] = pose[c * 4 + r];
    o[0] = pose[12];
    o[1] = pose[13];
    o[2] = pose[14];
    std::mt19937_64 rng(noise_seed);
    std::normal_distribution<double> n01(0.0, 1.0);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const double dc[3] = {(x - static_cast<double>(K[2])) / K[0], (y - static_cast<double>(K[3])) / K[1], 1.0};
            const double d[3] = {R[0] * dc[0] + R[1] * dc[1] + R[2] * dc[2], R[3] * dc[0] + R[4] * dc[1] + R[5] * dc[2],
                                 R[6] * dc[0] + R[7] * dc[1] + R[8] * dc[2]};
            const double t = hit(sc, o, d);
            const double g = n01(rng);  // drawn for every pixel: the stream never shifts
            const size_t i = static_cast<size_t>(y) * W + x;
            if (true_depth) true_depth[i] = t > 0 ? t : 0.0;
            if (t <= 0) {
                out_mm[i] = 0;
                continue;
            }
            double z = t;  // dc.z == 1, so the Z-depth equals the ray parameter
            if (noise_on) z += 0.0012 * z * z * g;
            const double mm = std::round(z * 1000.0);
            out_mm[i] = (mm <= 0 || mm > 65535) ? 0 : static_cast<uint16_t>(mm);
        }
}

// 2x bicubic upsampling of a millimetre frame (Keys kernel, a = -0.5, pixel centres
// aligned: output (X, Y) samples source ((X + 0.5) / 2 - 0.5, (Y + 0.5) / 2 - 0.5),
// borders clamped). A window touching an invalid (0) sample copies the nearest
// source sample instead, so holes never blend into depth.
AZU_SCENE_API void azu_scene_upsample2x(const uint16_t* src, int W, int H, uint16_t* dst) {
    auto keys = [](double x) {
        x = std::fabs(x);
        const double a = -0.5;
        if (x < 1) return (a + 2) * x * x * x - (a + 3) * x * x + 1;
        if (x < 2) return a * x * x * x - 5 * a * x * x + 8 * a * x - 4 * a;
        return 0.0;
    };
    const int W2 = 2 * W, H2 = 2 * H;
    for (int Y = 0; Y < H2; ++Y)
        for (int X = 0; X < W2; ++X) {
            const double sx = (X + 0.5) / 2 - 0.5, sy = (Y + 0.5) / 2 - 0.5;
            const int ix = static_cast<int>(std::floor(sx)), iy = static_cast<int>(std::floor(sy));
            const double fx = sx - ix, fy = sy - iy;
            double acc = 0;
            bool hole = false;
            for (int j = -1; j <= 2 && !hole; ++j)
                for (int i = -1; i <= 2; ++i) {
                    const int px = std::min(std::max(ix + i, 0), W - 1), py = std::min(std::max(iy + j, 0), H - 1);
                    const uint16_t v = src[static_cast<size_t>(py) * W + px];
                    if (v == 0) { hole = true; break; }
                    acc += keys(i - fx) * keys(j - fy) * v;
                }
            const int nx = std::min(std::max(static_cast<int>(std::lround(sx)), 0), W - 1);
            const int ny = std::min(std::max(static_cast<int>(std::lround(sy)), 0), H - 1);
            if (hole) {
                dst[static_cast<size_t>(Y) * W2 + X] = src[static_cast<size_t>(ny) * W + nx];
            } else {
                const double r = std::round(acc);
                dst[static_cast<size_t>(Y) * W2 + X] = (r <= 0 || r > 65535) ? 0 : static_cast<uint16_t>(r);
            }
        }
}
