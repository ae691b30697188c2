// Environment-map sky (SkyLight built from an HDR image, light.cpp:19-134): equirectangular radiance
// looked up bilinearly (lerp_helper, maths.h:760-780, u repeating, v clamped, texel centres at
// (i + 0.5) / res), importance-sampled from DistribTable2D (distrib.cpp:9-107) over the thresholded
// luminance sin(theta) lum - mean. The same functions run on the device (shading, NEE) and on the host
// (the light's power for the light sampler).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace ksd {

#define KS_HD __host__ __device__ inline

struct SkyMap {
    const float4 *rgb;  // w * h texels, row 0 at theta = 0
    const float *cond;  // h rows of w + 1 CDF entries
    const float *marg;  // h + 1 CDF entries
    int32_t w, h;
    float l2w[9], w2l[9]; // row-major direction transforms (map frame <-> world)
    int32_t yup;          // to_spherical_yup (1) or to_spherical (0)
    float strength;
};

constexpr float kSkyPi = 3.14159265359f;     // maths.h:58
constexpr float kSkyTwoPi = 6.28318530718f;  // maths.h:59
constexpr float kSkyInvPi = 1.0f / kSkyPi;   // maths.h:63

KS_HD void sky_mat(const float m[9], float x, float y, float z, float &ox, float &oy, float &oz) {
    ox = m[0] * x + m[1] * y + m[2] * z;
    oy = m[3] * x + m[4] * y + m[5] * z;
    oz = m[6] * x + m[7] * y + m[8] * z;
}

// world direction -> (u, v) = (phi / 2 pi, theta / pi) of the map frame, and sin(theta)
KS_HD void sky_uv(const SkyMap &s, float wx, float wy, float wz, float &u, float &v, float &sin_theta) {
    float x, y, z;
    sky_mat(s.w2l, wx, wy, wz, x, y, z);
    float theta, phi;
    if (s.yup) { // to_spherical_yup (maths.h:477-484)
        theta = acosf(fminf(fmaxf(y, -1.0f), 1.0f));
        phi = atan2f(z, x);
    } else { // to_spherical (maths.h:458-465)
        theta = acosf(fminf(fmaxf(z, -1.0f), 1.0f));
        phi = atan2f(y, x);
    }
    if (phi < 0.0f) phi += kSkyTwoPi;
    u = phi * kSkyInvPi * 0.5f;
    v = theta * kSkyInvPi;
    sin_theta = sinf(theta);
}

KS_HD void sky_lerp_index(float u, int res, bool repeat, int &i0, int &i1, float &t) {
    float us = u * (float)res - 0.5f; // TickMode::Middle
    i0 = (int)floorf(us);
    i1 = i0 + 1;
    t = us - (float)i0;
    if (repeat) {
        i0 = (i0 % res + res) % res;
        i1 = (i1 % res + res) % res;
    } else {
        i0 = i0 < 0 ? 0 : (i0 > res - 1 ? res - 1 : i0);
        i1 = i1 < 0 ? 0 : (i1 > res - 1 ? res - 1 : i1);
    }
}

// SkyLight::eval (light.cpp:59-86)
KS_HD void sky_eval(const SkyMap &s, float wx, float wy, float wz, float L[3]) {
    float u, v, st;
    sky_uv(s, wx, wy, wz, u, v, st);
    int u0, u1, v0, v1;
    float tu, tv;
    sky_lerp_index(u, s.w, true, u0, u1, tu);
    sky_lerp_index(v, s.h, false, v0, v1, tv);
    const float4 a = s.rgb[v0 * s.w + u0], b = s.rgb[v1 * s.w + u0], c = s.rgb[v0 * s.w + u1],
                 d = s.rgb[v1 * s.w + u1];
    const float wa = (1.0f - tu) * (1.0f - tv), wb = (1.0f - tu) * tv, wc = tu * (1.0f - tv), wd = tu * tv;
    const float ax[3] = {a.x, a.y, a.z}, bx[3] = {b.x, b.y, b.z}, cx[3] = {c.x, c.y, c.z}, dx[3] = {d.x, d.y, d.z};
    for (int k = 0; k < 3; ++k) {
        float l = 0.0f;
        l += wa * ax[k];
        l += wb * bx[k];
        l += wc * cx[k];
        l += wd * dx[k];
        L[k] = l * s.strength;
    }
}

// DistribTable::sample_linear (distrib.cpp:43-55) over cdf[0..n]
KS_HD float sky_sample_linear(const float *cdf, int n, float u, float &pdf, int &index) {
    const float before_one = 0x1.fffffep-1f;
    u = fminf(fmaxf(u, 0.0f), before_one);
    int lo = 0, hi = n + 1; // upper_bound: first entry > u
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    index = lo - 1;
    pdf = (cdf[index + 1] - cdf[index]) * (float)n;
    float du = (u - cdf[index]) / (cdf[index + 1] - cdf[index]);
    return ((float)index + du) / (float)n;
}

// DistribTable2D::pdf(vec2) (distrib.cpp:96-107)
KS_HD float sky_pdf2(const SkyMap &s, float u, float v) {
    int x = (int)(u * (float)s.w), y = (int)(v * (float)s.h);
    x = x < 0 ? 0 : (x > s.w - 1 ? s.w - 1 : x);
    y = y < 0 ? 0 : (y > s.h - 1 ? s.h - 1 : y);
    const float *c = s.cond + (size_t)y * (s.w + 1);
    return ((s.marg[y + 1] - s.marg[y]) * (float)s.h) * ((c[x + 1] - c[x]) * (float)s.w);
}

// SkyLight::pdf (light.cpp:113-134)
KS_HD float sky_map_pdf(const SkyMap &s, float wx, float wy, float wz) {
    float u, v, st;
    sky_uv(s, wx, wy, wz, u, v, st);
    if (st == 0.0f) return 0.0f;
    float pdf = sky_pdf2(s, u, v);
    return pdf / (2.0f * kSkyPi * kSkyPi * st);
}

// SkyLight::sample (light.cpp:88-111): returns Le / pdf in L, the direction and the pdf
KS_HD void sky_map_sample(const SkyMap &s, float u0, float u1, float &wx, float &wy, float &wz, float &pdf,
                          float L[3]) {
    L[0] = L[1] = L[2] = 0.0f;
    float pm, pc;
    int iv, iu;
    float sv = sky_sample_linear(s.marg, s.h, u1, pm, iv);
    float su = sky_sample_linear(s.cond + (size_t)iv * (s.w + 1), s.w, u0, pc, iu);
    pdf = pm * pc;
    if (pdf == 0.0f) return;
    float phi = su * 2.0f * kSkyPi, theta = sv * kSkyPi;
    float ct = cosf(theta), st = sinf(theta), cp = cosf(phi), sp = sinf(phi);
    float lx, ly, lz;
    if (s.yup) { // to_cartesian_yup (maths.h:486-493)
        lx = cp * st;
        ly = ct;
        lz = sp * st;
    } else { // to_cartesian (maths.h:467-475)
        lx = cp * st;
        ly = sp * st;
        lz = ct;
    }
    pdf /= (2.0f * kSkyPi * kSkyPi * st);
    if (st == 0.0f) {
        pdf = 0.0f;
        return;
    }
    sky_mat(s.l2w, lx, ly, lz, wx, wy, wz);
    float Le[3];
    sky_eval(s, wx, wy, wz, Le);
    for (int k = 0; k < 3; ++k) L[k] = Le[k] / pdf;
}

} // namespace ksd
