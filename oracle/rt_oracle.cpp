/*
 * rt_oracle.cpp -- CPU restatement of RenderToy's path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  It is loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg and by
 * nothing else.  Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).
 *
 * Operation order follows the reference line by line (file:line cited per
 * function) so that, built without FP contraction, every float result is
 * bit-identical to the reference compiled the same way.
 *
 * How each part is pinned (tests/test_oracle.py):
 *   - ray generation, triangle cache, Moller-Trumbore, octree traversal, AOV
 *     passes, BSDF eval/sample and the whole mt19937 path tracer are compared
 *     bit-for-bit with the reference itself, compiled from its own sources
 *     into oracle/_ref (fixtures under tests/golden/);
 *   - Moller-Trumbore against closed-form hits on axis-aligned dyadic
 *     triangles; brute force == octree away from exact-t ties;
 *   - BSDF against closed forms (Lambert at roughness 0.5 lobes, Fresnel F0,
 *     GTR2 normalisation) and the integrator against the analytic furnace
 *     E[L] = Le * sum_k rho^k;
 *   - Philox4x32-10 against the published Random123 known-answer vectors.
 */
#include "rt_oracle.h"

#include <cmath>
#include <cstring>
#include <cstdint>
#include <algorithm>
#include <queue>
#include <random>
#include <vector>
#include <memory>

namespace {

constexpr float kPi = 3.14159265358979323846f;          /* rtmath.h:17-18 */
constexpr float kFloatEpsilon = 1e-7f;                   /* rtmath.h:21 */
constexpr float kFloatInfinity = 1e7f;                   /* rtmath.h:23 */
constexpr float kE = 2.71828182845904523536f;            /* M_Ef32 */
constexpr uint32_t kNone = 0xFFFFFFFFu;

/* ---------------- Vector3f semantics (rtmath.h:74-534) ---------------- */
struct V3 { float x, y, z; };
inline V3 v3(float a) { return {a, a, a}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 mulv(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 muls(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 divv(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
/* Dot: ret(0); ret += a[i]*b[i] (rtmath.h:320-329) */
inline float dot(V3 a, V3 b) {
    float r = 0.0f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
/* Cross (rtmath.h:399-418) */
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
/* Length2 / Length / Normalize (rtmath.h:126-150) */
inline float length2(V3 a) {
    float r = 0.0f;
    r += a.x * a.x;
    r += a.y * a.y;
    r += a.z * a.z;
    return r;
}
inline float length(V3 a) { return std::sqrt(length2(a)); }
inline V3 normalized(V3 a) { float l = length(a); return {a.x / l, a.y / l, a.z / l}; }
/* operator== : |a-b| > 1e-7 on any component -> false (rtmath.h:296-308) */
inline bool approx_eq(V3 a, V3 b) {
    if (std::abs(a.x - b.x) > kFloatEpsilon) return false;
    if (std::abs(a.y - b.y) > kFloatEpsilon) return false;
    if (std::abs(a.z - b.z) > kFloatEpsilon) return false;
    return true;
}
inline V3 vpow(V3 a, V3 b) { return {std::pow(a.x, b.x), std::pow(a.y, b.y), std::pow(a.z, b.z)}; }
inline V3 vlog(V3 a) { return {std::log(a.x), std::log(a.y), std::log(a.z)}; }
/* Lerp x*(1-a)+y*a (rtmath.h:1245-1249) */
inline float lerp(float x, float y, float a) { return x * (1.0f - a) + y * a; }
inline V3 lerpv(V3 x, V3 y, V3 a) { return add(mulv(x, sub(v3(1.0f), a)), mulv(y, a)); }
/* Luma BT.709 (rtmath.h:1140-1149) */
inline float luma(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
inline float fmax0(float x) { return (0.0f < x) ? x : 0.0f; }          /* std::max(0.0f, x) */
inline float clamp01(float v) { return (v < 0.0f) ? 0.0f : ((1.0f < v) ? 1.0f : v); }

/* Matrix4x4f * Vector4f: row dots (rtmath.h:724-733) */
inline void mat4_mul(const float *m, const float v[4], float out[4]) {
    for (int i = 0; i < 4; ++i) {
        float r = 0.0f;
        for (int k = 0; k < 4; ++k) r += m[i * 4 + k] * v[k];
        out[i] = r;
    }
}

/* ---------------- PBR helpers (rtmath.cpp:169-353) ---------------- */
float GTR1(float NdotH, float a) {                                   /* :169-176 */
    if (a >= 1.0f) return (1.0f / kPi);
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * NdotH * NdotH;
    return (a2 - 1.0f) / (kPi * std::log(a2) * t);
}
V3 SampleGTR1(float rgh, float r1, float r2) {                       /* :178-191 */
    (void)r2;
    float a = std::max(0.001f, rgh);
    float a2 = a * a;
    float phi = r1 * 2.0f * kPi;
    float cos_theta = std::sqrt((1.0f - std::pow(a2, 1.0f - r1)) / (1.0f - a2));
    float st = std::sqrt(1.0f - (cos_theta * cos_theta));
    float sin_theta = (st < 0.0f) ? 0.0f : ((1.0f < st) ? 1.0f : st);
    float sin_phi = std::sin(phi);
    float cos_phi = std::cos(phi);
    return {sin_theta * cos_phi, sin_theta * sin_phi, cos_theta};
}
float GTR2(float NdotH, float a) {                                   /* :193-198 */
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * NdotH * NdotH;
    return a2 / (kPi * t * t);
}
V3 SampleGGXVNDF(V3 V, float rgh, float r1, float r2) {              /* :214-232 */
    V3 Vh = normalized(V3{rgh * V.x, rgh * V.y, V.z});
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    V3 T1 = lensq > 0.0f ? muls(V3{-Vh.y, Vh.x, 0.0f}, 1.0f / std::sqrt(lensq)) : V3{1.0f, 0.0f, 0.0f};
    V3 T2 = cross(Vh, T1);
    float r = std::sqrt(r1);
    float phi = 2.0f * kPi * r2;
    float t1 = r * std::cos(phi);
    float t2 = r * std::sin(phi);
    float s = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - s) * std::sqrt(1.0f - t1 * t1) + s * t2;
    V3 Nh = add(add(muls(T1, t1), muls(T2, t2)), muls(Vh, std::sqrt(fmax0(1.0f - t1 * t1 - t2 * t2))));
    return normalized(V3{rgh * Nh.x, rgh * Nh.y, fmax0(Nh.z)});
}
float SmithG(float NdotV, float alpha_g) {                           /* :253-258 */
    float a = alpha_g * alpha_g;
    float b = NdotV * NdotV;
    return (2.0f * NdotV) / (NdotV + std::sqrt(a + b - a * b));
}
float SchlickFresnel(float u) {                                      /* :268-273 */
    float m = 1.0f - u;
    m = (m < 0.0f) ? 0.0f : ((1.0f < m) ? 1.0f : m);
    float m2 = m * m;
    return m2 * m2 * m;
}
float DielectricFresnel(float cos_theta_i, float eta) {              /* :275-289 */
    float sin_theta_t_sq = eta * eta * (1.0f - cos_theta_i * cos_theta_i);
    if (sin_theta_t_sq > 1.0f) return 1.0f;
    float x = 1.0f - sin_theta_t_sq;
    float cos_theta_t = std::sqrt((x < 0.0f) ? 0.0f : x);
    float rs = (eta * cos_theta_t - cos_theta_i) / (eta * cos_theta_t + cos_theta_i);
    float rp = (eta * cos_theta_i - cos_theta_t) / (eta * cos_theta_i + cos_theta_t);
    return 0.5f * (rs * rs + rp * rp);
}
V3 CosineSampleHemisphere(float r1, float r2) {                      /* :291-300 */
    float r = std::sqrt(r1);
    float phi = 2.0f * kPi * r2;
    V3 d;
    d.x = r * std::cos(phi);
    d.y = r * std::sin(phi);
    d.z = std::sqrt(fmax0(1.0f - d.x * d.x - d.y * d.y));
    return d;
}
inline float PowerHeuristic(float a, float b) { float t = a * a; return t / (b * b + t); }  /* :317-321 */
void Onb(V3 N, V3 &T, V3 &B) {                                       /* :323-328 */
    V3 up = std::abs(N.z) < 0.999f ? V3{0.0f, 0.0f, 1.0f} : V3{1.0f, 0.0f, 0.0f};
    T = normalized(cross(up, N));
    B = cross(N, T);
}
inline V3 ToWorld(V3 X, V3 Y, V3 Z, V3 V) { return add(add(muls(X, V.x), muls(Y, V.y)), muls(Z, V.z)); }
inline V3 ToLocal(V3 X, V3 Y, V3 Z, V3 V) { return {dot(V, X), dot(V, Y), dot(V, Z)}; }
inline V3 Reflect(V3 I, V3 N) { return sub(I, muls(N, 2.0f * dot(I, N))); }     /* :340-343 */
V3 Refract(V3 I, V3 N, float eta) {                                  /* :345-353 */
    float NdotI = dot(N, I);
    float k = 1.0f - eta * eta * (1.0f - NdotI * NdotI);
    if (k < 0.0f) return {0.0f, 0.0f, 0.0f};
    return sub(muls(I, eta), muls(N, eta * NdotI + std::sqrt(k)));
}

/* ---------------- Disney BSDF (material.cpp:35-295) ---------------- */
struct Mat {
    V3 base, emission;
    float roughness, metallic, spec_trans, ior, anisotropic, subsurface, specular_tint,
        sheen, sheen_tint, clearcoat, clearcoat_roughness, at_distance;
    V3 extinction;
};
Mat load_mat(const float *m) {
    Mat r;
    r.base = {m[0], m[1], m[2]};
    r.emission = {m[3], m[4], m[5]};
    r.roughness = m[6]; r.metallic = m[7]; r.spec_trans = m[8]; r.ior = m[9];
    r.anisotropic = m[10]; r.subsurface = m[11]; r.specular_tint = m[12]; r.sheen = m[13];
    r.sheen_tint = m[14]; r.clearcoat = m[15]; r.clearcoat_roughness = m[16]; r.at_distance = m[17];
    r.extinction = {m[18], m[19], m[20]};
    return r;
}
float FresnelMix(const Mat &m, float eta, float VdotH) {             /* :35-40 */
    float metallic_fresnel = SchlickFresnel(VdotH);
    float dielectric_fresnel = DielectricFresnel(VdotH, eta);
    return lerp(dielectric_fresnel, metallic_fresnel, m.metallic);
}
V3 EvalDiffuse(const Mat &m, V3 Csheen, V3 V, V3 L, V3 H, float &pdf) {   /* :42-65 */
    pdf = 0.0f;
    if (L.z <= 0.0f) return v3(0.0f);
    float FL = SchlickFresnel(L.z);
    float FV = SchlickFresnel(V.z);
    float FH = SchlickFresnel(dot(L, H));
    float Fd90 = 0.5f + 2.0f * dot(L, H) * dot(L, H) * m.roughness;
    float Fd = lerp(1.0f, Fd90, FL) * lerp(1.0f, Fd90, FV);
    float Fss90 = dot(L, H) * dot(L, H) * m.roughness;
    float Fss = lerp(1.0f, Fss90, FL) * lerp(1.0f, Fss90, FV);
    float ss = 1.25f * (Fss * (1.0f / (L.z + V.z) - 0.5f) + 0.5f);
    V3 Fsheen = muls(Csheen, FH * m.sheen);
    pdf = L.z * (1.0f / kPi);
    V3 r = add(muls(m.base, (1.0f / kPi) * lerp(Fd, ss, m.subsurface)), Fsheen);
    return muls(muls(r, 1.0f - m.metallic), 1.0f - m.spec_trans);
}
V3 EvalSpecReflection(const Mat &m, float eta, V3 specCol, V3 V, V3 L, V3 H, float &pdf) {  /* :67-82 */
    pdf = 0.0f;
    if (L.z <= 0.0f) return v3(0.0f);
    float FM = FresnelMix(m, eta, dot(L, H));
    V3 F = lerpv(specCol, v3(1.0f), v3(FM));
    float D = GTR2(H.z, m.roughness);
    float G1 = SmithG(std::abs(V.z), m.roughness);
    float G2 = G1 * SmithG(std::abs(L.z), m.roughness);
    float jacobian = 1.0f / (4.0f * dot(V, H));
    pdf = G1 * fmax0(dot(V, H)) * D * jacobian / V.z;
    return divs(muls(muls(F, D), G2), 4.0f * L.z * V.z);
}
V3 EvalSpecRefraction(const Mat &m, float eta, V3 V, V3 L, V3 H, float &pdf) {   /* :84-102 */
    pdf = 0.0f;
    if (L.z >= 0.0f) return v3(0.0f);
    float F = DielectricFresnel(std::abs(dot(V, H)), eta);
    float D = GTR2(H.z, m.roughness);
    float denom = dot(L, H) + dot(V, H) * eta;
    denom *= denom;
    float G1 = SmithG(std::abs(V.z), m.roughness);
    float G2 = G1 * SmithG(std::abs(L.z), m.roughness);
    float jacobian = std::abs(dot(L, H)) / denom;
    pdf = G1 * fmax0(dot(V, H)) * D * jacobian / V.z;
    V3 spec_color = vpow(m.base, v3(0.5f));
    V3 r = muls(spec_color, (float)(1.0 - (double)m.metallic));
    r = muls(r, m.spec_trans);
    r = muls(r, (float)(1.0 - (double)F));
    r = muls(r, D);
    r = muls(r, G2);
    r = muls(r, std::abs(dot(V, H)));
    r = muls(r, std::abs(dot(L, H)));
    r = muls(r, eta);
    r = muls(r, eta);
    return divs(r, denom * std::abs(L.z) * std::abs(V.z));
}
V3 EvalClearcoat(const Mat &m, V3 V, V3 L, V3 H, float &pdf) {      /* :104-118 */
    pdf = 0.0f;
    if (L.z <= 0.0f) return v3(0.0f);
    float FH = DielectricFresnel(dot(V, H), 1.0f / 1.5f);
    float F = lerp(0.04f, 1.0f, FH);
    float D = GTR1(H.z, m.clearcoat_roughness);
    float G = SmithG(L.z, 0.25f) * SmithG(V.z, 0.25f);
    float jacobian = 1.0f / (4.0f * dot(V, H));
    pdf = D * H.z * jacobian;
    V3 r = muls(muls(muls(muls(v3(0.25f), m.clearcoat), F), D), G);
    return divs(r, 4.0f * L.z * V.z);
}
void GetSpecColor(const Mat &m, float eta, V3 &spec, V3 &sheen) {   /* :120-127 */
    float lum = luma(m.base);
    V3 ctint = lum > 0.0f ? divs(m.base, lum) : v3(1.0f);
    float F0 = (1.0f - eta) / (1.0f + eta);
    spec = lerpv(muls(lerpv(v3(1.0f), ctint, v3(m.specular_tint)), F0 * F0), m.base, v3(m.metallic));
    sheen = lerpv(v3(1.0f), ctint, v3(m.sheen_tint));
}
void GetLobeProbabilities(const Mat &m, V3 spec, float approx_fresnel,   /* :129-141 */
                          float &dw, float &srw, float &stw, float &cw) {
    dw = luma(m.base) * (1.0f - m.metallic) * (1.0f - m.spec_trans);
    srw = luma(lerpv(spec, v3(1.0f), v3(approx_fresnel)));
    stw = (1.0f - approx_fresnel) * (1.0f - m.metallic) * m.spec_trans * luma(m.base);
    cw = m.clearcoat * (1.0f - m.metallic);
    float total = dw + srw + stw + cw;
    dw /= total;
    srw /= total;
    stw /= total;
    cw /= total;
}
V3 BsdfEval(const Mat &m, float eta, V3 V, V3 N, V3 L, float &bsdf_pdf) {   /* :143-204 */
    bsdf_pdf = 0.0f;
    V3 f = v3(0.0f);
    V3 T, B;
    Onb(N, T, B);
    V = ToLocal(T, B, N, V);
    L = ToLocal(T, B, N, L);
    V3 H;
    if (L.z > 0.0f) H = normalized(add(L, V));
    else H = normalized(add(L, muls(V, eta)));
    if (H.z < 0.0f) H = neg(H);
    V3 spec, sheen;
    GetSpecColor(m, eta, spec, sheen);
    float dw, srw, stw, cw;
    float fresnel = FresnelMix(m, eta, dot(V, H));
    GetLobeProbabilities(m, spec, fresnel, dw, srw, stw, cw);
    float pdf;
    if (dw > 0.0f && L.z > 0.0f) { f = add(f, EvalDiffuse(m, sheen, V, L, H, pdf)); bsdf_pdf += pdf * dw; }
    if (srw > 0.0f && L.z > 0.0f && V.z > 0.0f) { f = add(f, EvalSpecReflection(m, eta, spec, V, L, H, pdf)); bsdf_pdf += pdf * srw; }
    if (stw > 0.0f && L.z < 0.0f) { f = add(f, EvalSpecRefraction(m, eta, V, L, H, pdf)); bsdf_pdf += pdf * stw; }
    if (cw > 0.0f && L.z > 0.0f && V.z > 0.0f) { f = add(f, EvalClearcoat(m, V, L, H, pdf)); bsdf_pdf += pdf * cw; }
    return muls(f, std::abs(L.z));
}
bool BsdfSample(const Mat &m, float eta, V3 N, V3 in_dir, float r1, float r2,   /* :206-295 */
                V3 &out_dir, V3 &color, float &pdf) {
    pdf = 0.0f;
    V3 f = v3(0.0f);
    V3 V = in_dir;
    V3 T, B;
    Onb(N, T, B);
    V = ToLocal(T, B, N, V);
    V3 spec, sheen;
    GetSpecColor(m, eta, spec, sheen);
    float dw, srw, stw, cw;
    float approx = FresnelMix(m, eta, V.z);
    GetLobeProbabilities(m, spec, approx, dw, srw, stw, cw);
    float cdf0 = dw, cdf1 = cdf0 + srw, cdf2 = cdf1 + stw;
    if (r1 < cdf0) {
        r1 /= cdf0;
        out_dir = CosineSampleHemisphere(r1, r2);
        V3 H = normalized(add(out_dir, V));
        f = EvalDiffuse(m, sheen, V, out_dir, H, pdf);
        pdf *= dw;
    } else if (r1 < cdf1) {
        r1 = (r1 - cdf0) / (cdf1 - cdf0);
        V3 H = SampleGGXVNDF(V, m.roughness, r1, r2);
        if (H.z < 0.0f) H = neg(H);
        out_dir = normalized(Reflect(neg(V), H));
        f = EvalSpecReflection(m, eta, spec, V, out_dir, H, pdf);
        pdf *= srw;
    } else if (r1 < cdf2) {
        r1 = (r1 - cdf1) / (cdf2 - cdf1);
        V3 H = SampleGGXVNDF(V, m.roughness, r1, r2);
        if (H.z < 0.0f) H = neg(H);
        out_dir = normalized(Refract(neg(V), H, eta));
        f = EvalSpecRefraction(m, eta, V, out_dir, H, pdf);
        pdf *= stw;
    } else {
        r1 = (float)((double)(r1 - cdf2) / (1.0 - (double)cdf2));
        V3 H = SampleGTR1(m.clearcoat_roughness, r1, r2);
        if (H.z < 0.0f) H = neg(H);
        out_dir = normalized(Reflect(neg(V), H));
        f = EvalClearcoat(m, V, out_dir, H, pdf);
        pdf *= cw;
    }
    out_dir = ToWorld(T, B, N, out_dir);
    color = muls(f, std::abs(dot(N, out_dir)));
    return !approx_eq(out_dir, v3(0.0f));
}

/* ---------------- Philox4x32-10 ---------------- */
inline void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
    uint32_t c[4] = {c_in[0], c_in[1], c_in[2], c_in[3]};
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c[0] = hi1 ^ c[1] ^ k0;
        c[1] = lo1;
        c[2] = hi0 ^ c[3] ^ k1;
        c[3] = lo0;
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}
inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

/* RNG with per-purpose draws.  MT mode draws sequentially in the reference's
 * order (world.cpp:9, object.cpp:171-172, material.cpp:214-215); Philox mode
 * addresses each draw by (pixel, sample, depth, purpose). */
struct Rng {
    int mode = 0;
    std::mt19937 *mt = nullptr;
    uint32_t key[2] = {0, 0};
    uint32_t pixel = 0, sample = 0;
    uint32_t depth = 0;
    uint32_t a[4], b[4];
    void vertex(uint32_t d) {
        depth = d;
        if (mode == 1) {
            uint32_t c0[4] = {pixel, sample, d, 0};
            uint32_t c1[4] = {pixel, sample, d, 1};
            philox(c0, key, a);
            philox(c1, key, b);
        }
    }
    int pick(int n) {
        if (mode == 0) { std::uniform_int_distribution<> d(0, n - 1); return d(*mt); }
        return (int)(uint32_t)(((uint64_t)a[0] * (uint64_t)(uint32_t)n) >> 32);
    }
    float uni(int slot) {
        if (mode == 0) { std::uniform_real_distribution<float> d(0.0f, 1.0f); return d(*mt); }
        switch (slot) {
        case 1: return u01(a[1]);
        case 2: return u01(a[2]);
        case 3: return u01(a[3]);
        default: return u01(b[0]);
        }
    }
};

/* ---------------- scene ---------------- */
struct Tri {
    V3 v[3];         /* vert_w */
    V3 e1, e2;       /* v0v1_w, v0v2_w */
    V3 normal;       /* object.cpp:185-190 */
    float area;      /* object.cpp:192-195 */
    V3 n[3];         /* per-vertex normals */
    uint32_t mesh;
};
struct Box { V3 vmin, vmax; };

struct OctNode {
    int child[8];
    std::vector<int> items;   /* indices into triangle list (bbox_list order == triangle order) */
    Box bbox;
    bool is_leaf;
};

}  // namespace

struct orc_scene {
    std::vector<Tri> tris;
    std::vector<uint32_t> mesh_mat;
    std::vector<uint8_t> mesh_smooth;
    std::vector<Mat> mats;
    std::vector<uint32_t> emitters;
    V3 sky, ground;
    /* octree */
    bool oct_built = false;
    std::vector<OctNode> nodes;
    std::vector<Box> tri_box;
    Box tree_box;
};

namespace {

/* Triangle::UpdateCache (object.cpp:224-243) on already transformed vertices. */
void update_cache(Tri &t) {
    t.e1 = sub(t.v[1], t.v[0]);
    t.e2 = sub(t.v[2], t.v[0]);
    t.area = length(cross(t.e1, t.e2)) * 0.5f;
    V3 tangent = normalized(t.e1);
    t.normal = normalized(cross(tangent, sub(t.v[2], t.v[1])));
}

/* Triangle::Intersect, two-sided Moller-Trumbore (object.cpp:107-167). */
inline bool tri_intersect(const Tri &tr, V3 src, V3 dir, float &t, float &u, float &v) {
    V3 pvec = cross(dir, tr.e2);
    float det = dot(tr.e1, pvec);
    if (det > 0.0f) {
        V3 tvec = sub(src, tr.v[0]);
        float alpha = dot(pvec, tvec);
        if (alpha < 0.0f || alpha > det) return false;
        V3 qvec = cross(tvec, tr.e1);
        float beta = dot(qvec, dir);
        if (beta < 0.0f || alpha + beta > det) return false;
        float inv_det = 1.0f / det;
        t = dot(tr.e2, qvec) * inv_det;
        if (t < 0.0f) return false;
        u = alpha * inv_det;
        v = beta * inv_det;
        return true;
    } else if (det < 0.0f) {
        V3 tvec = sub(src, tr.v[0]);
        float alpha = dot(pvec, tvec);
        if (alpha > 0.0f || alpha < det) return false;
        V3 qvec = cross(tvec, tr.e1);
        float beta = dot(qvec, dir);
        if (beta > 0.0f || alpha + beta < det) return false;
        float inv_det = 1.0f / det;
        t = dot(tr.e2, qvec) * inv_det;
        if (t < 0.0f) return false;
        u = alpha * inv_det;
        v = beta * inv_det;
        return true;
    }
    return false;
}

/* Triangle::BBox (object.cpp:83-99) */
Box tri_bbox(const Tri &t) {
    auto mx = [](float a, float b) { return (a < b) ? b : a; };
    auto mn = [](float a, float b) { return (b < a) ? b : a; };
    Box b;
    b.vmax = {mx(t.v[0].x, mx(t.v[1].x, t.v[2].x)), mx(t.v[0].y, mx(t.v[1].y, t.v[2].y)),
              mx(t.v[0].z, mx(t.v[1].z, t.v[2].z))};
    b.vmin = {mn(t.v[0].x, mn(t.v[1].x, t.v[2].x)), mn(t.v[0].y, mn(t.v[1].y, t.v[2].y)),
              mn(t.v[0].z, mn(t.v[1].z, t.v[2].z))};
    return b;
}
/* BoundingBox::ExtendBy (bvh.cpp:76-84) */
void extend(Box &a, const Box &b) {
    auto mx = [](float p, float q) { return (p < q) ? q : p; };
    auto mn = [](float p, float q) { return (q < p) ? q : p; };
    a.vmin = {mn(a.vmin.x, b.vmin.x), mn(a.vmin.y, b.vmin.y), mn(a.vmin.z, b.vmin.z)};
    a.vmax = {mx(a.vmax.x, b.vmax.x), mx(a.vmax.y, b.vmax.y), mx(a.vmax.z, b.vmax.z)};
}
inline V3 centroid(const Box &b) { return muls(add(b.vmin, b.vmax), 0.5f); }   /* bvh.cpp:86-89 */

/* BoundingBox::Intersect: division slabs, t_max never written (bvh.cpp:19-74). */
inline bool box_intersect(const Box &b, V3 src, V3 dir, float &t_min) {
    float tmin = (b.vmin.x - src.x) / dir.x;
    float tmax = (b.vmax.x - src.x) / dir.x;
    if (tmin > tmax) std::swap(tmin, tmax);
    float tymin = (b.vmin.y - src.y) / dir.y;
    float tymax = (b.vmax.y - src.y) / dir.y;
    if (tymin > tymax) std::swap(tymin, tymax);
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = (b.vmin.z - src.z) / dir.z;
    float tzmax = (b.vmax.z - src.z) / dir.z;
    if (tzmin > tzmax) std::swap(tzmin, tzmax);
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    t_min = tmin;
    return true;
}

int new_node(orc_scene *s) {
    OctNode n;
    for (int i = 0; i < 8; ++i) n.child[i] = -1;
    n.bbox = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};     /* BoundingBox() seeds at origin (bvh.cpp:9-12) */
    n.is_leaf = true;
    s->nodes.push_back(std::move(n));
    return (int)s->nodes.size() - 1;
}

/* Octree::Insert (bvh.cpp:127-195) */
void oct_insert(orc_scene *s, int node, int item, const Box &cell, int depth) {
    if (s->nodes[node].is_leaf) {
        if (s->nodes[node].items.empty() || depth == 8) {
            s->nodes[node].items.push_back(item);
        } else {
            s->nodes[node].is_leaf = false;
            while (!s->nodes[node].items.empty()) {
                oct_insert(s, node, s->nodes[node].items.back(), cell, depth);
                s->nodes[node].items.pop_back();
            }
            oct_insert(s, node, item, cell, depth);
        }
    } else {
        V3 ic = centroid(s->tri_box[item]);
        V3 nc = centroid(cell);
        Box cb = {{0, 0, 0}, {0, 0, 0}};
        int ci = 0;
        if (ic.x > nc.x) { ci = 4; cb.vmin.x = nc.x; cb.vmax.x = cell.vmax.x; }
        else { cb.vmin.x = cell.vmin.x; cb.vmax.x = nc.x; }
        if (ic.y > nc.y) { ci += 2; cb.vmin.y = nc.y; cb.vmax.y = cell.vmax.y; }
        else { cb.vmin.y = cell.vmin.y; cb.vmax.y = nc.y; }
        if (ic.z > nc.z) { ci += 1; cb.vmin.z = nc.z; cb.vmax.z = cell.vmax.z; }
        else { cb.vmin.z = cell.vmin.z; cb.vmax.z = nc.z; }
        if (s->nodes[node].child[ci] < 0) {
            int c = new_node(s);
            s->nodes[node].child[ci] = c;
        }
        oct_insert(s, s->nodes[node].child[ci], item, cb, depth + 1);
    }
}

/* Octree::Build (bvh.cpp:197-228) */
void oct_build(orc_scene *s, int node, const Box &cell) {
    if (s->nodes[node].is_leaf) {
        for (int e : s->nodes[node].items) extend(s->nodes[node].bbox, s->tri_box[e]);
    } else {
        for (int i = 0; i < 8; ++i) {
            int c = s->nodes[node].child[i];
            if (c >= 0) {
                V3 cen = centroid(cell);
                Box cb;
                cb.vmin.x = (i & 4) ? cen.x : cell.vmin.x;
                cb.vmax.x = (i & 4) ? cell.vmax.x : cen.x;
                cb.vmin.y = (i & 2) ? cen.y : cell.vmin.y;
                cb.vmax.y = (i & 2) ? cell.vmax.y : cen.y;
                cb.vmin.z = (i & 1) ? cen.z : cell.vmin.z;
                cb.vmax.z = (i & 1) ? cell.vmax.z : cen.z;
                oct_build(s, c, cb);
                Box cbb = s->nodes[c].bbox;
                extend(s->nodes[node].bbox, cbb);
            }
        }
    }
}

/* BVH::BVH (bvh.cpp:242-264) */
void ensure_octree(orc_scene *s) {
    if (s->oct_built) return;
    size_t n = s->tris.size();
    s->tri_box.resize(n);
    s->tree_box = {{0, 0, 0}, {0, 0, 0}};
    for (size_t i = 0; i < n; ++i) {
        s->tri_box[i] = tri_bbox(s->tris[i]);
        extend(s->tree_box, s->tri_box[i]);
    }
    s->nodes.clear();
    s->nodes.reserve(n * 2 + 16);
    new_node(s);
    for (size_t i = 0; i < n; ++i) oct_insert(s, 0, (int)i, s->tree_box, 0);
    oct_build(s, 0, s->tree_box);
    s->oct_built = true;
}

struct QE {
    int node;
    float t;
    friend bool operator<(const QE &a, const QE &b) { return a.t > b.t; }   /* bvh.h:54-63 */
};

struct Hit { uint32_t id; float t, u, v; V3 pos; };

/* BVH::Intersect best-first (bvh.cpp:288-339) */
Hit intersect_octree(orc_scene *s, V3 src, V3 dir, uint32_t exclude) {
    Hit h;
    h.id = kNone;
    float t = kFloatInfinity;
    float u = 0.0f, v = 0.0f;
    float t_min;
    h.t = t;
    if (!box_intersect(s->nodes[0].bbox, src, dir, t_min)) return h;
    std::priority_queue<QE> q;
    q.push(QE{0, 0.0f});
    while (!q.empty() && q.top().t < t) {
        int node = q.top().node;
        q.pop();
        const OctNode &nd = s->nodes[node];
        if (nd.is_leaf) {
            for (int e : nd.items) {
                float it, iu, iv;
                if ((uint32_t)e != exclude && tri_intersect(s->tris[e], src, dir, it, iu, iv)) {
                    if (it < t) { t = it; u = iu; v = iv; h.id = (uint32_t)e; }
                }
            }
        } else {
            for (int i = 0; i < 8; ++i) {
                int c = nd.child[i];
                if (c >= 0) {
                    float tc;
                    if (box_intersect(s->nodes[c].bbox, src, dir, tc)) q.push(QE{c, tc});
                }
            }
        }
    }
    h.t = t; h.u = u; h.v = v;
    h.pos = add(src, muls(dir, t));
    return h;
}

/* DISABLE_BVH brute force (bvh.cpp:268-286): lowest index wins exact ties. */
Hit intersect_brute(orc_scene *s, V3 src, V3 dir, uint32_t exclude) {
    Hit h;
    h.id = kNone;
    float t = kFloatInfinity, u = 0.0f, v = 0.0f;
    const size_t n = s->tris.size();
    for (size_t i = 0; i < n; ++i) {
        float it, iu, iv;
        if ((uint32_t)i != exclude && tri_intersect(s->tris[i], src, dir, it, iu, iv)) {
            if (it < t) { t = it; u = iu; v = iv; h.id = (uint32_t)i; }
        }
    }
    h.t = t; h.u = u; h.v = v;
    h.pos = add(src, muls(dir, t));
    return h;
}

inline Hit intersect(orc_scene *s, int mode, V3 src, V3 dir, uint32_t exclude) {
    return mode == 1 ? intersect_octree(s, src, dir, exclude) : intersect_brute(s, src, dir, exclude);
}

/* Triangle::NormalC (object.cpp:207-217) */
inline V3 normal_c(const orc_scene *s, const Tri &t, float u, float v) {
    if (s->mesh_smooth[t.mesh]) {
        float w = 1.0f - u - v;
        return normalized(add(add(muls(t.n[0], w), muls(t.n[1], u)), muls(t.n[2], v)));
    }
    return t.normal;
}
inline const Mat &mat_of(const orc_scene *s, const Tri &t) { return s->mats[s->mesh_mat[t.mesh]]; }

/* SurfacePoint::GetEmission<true> (surfacepoint.h:20-45) */
inline V3 get_emission(const orc_scene *s, const Tri &t, float u, float v, V3 position, V3 to_pos,
                       V3 out_dir, float &pdf) {
    V3 ray = sub(to_pos, position);
    float distance2 = dot(ray, ray);
    float cos_area = dot(out_dir, normal_c(s, t, u, v)) * t.area;
    pdf = std::abs(distance2 / cos_area);
    return mat_of(s, t).emission;
}

/* Camera ray (renderer.cpp:253-259, object.cpp:42-64) */
inline void camera_ray(const float *M, float top, float right, float divW, float divH, uint32_t x,
                       uint32_t y, V3 &src, V3 &dir) {
    float ndc_x = float(x) * divW;
    float ndc_y = float(y) * divH;
    float sx = 2.0f * right * ndc_x - right;
    float sy = 2.0f * top * ndc_y - top;
    float o4[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    float d4[4] = {sx, sy, -1.0f, 1.0f};
    float so[4], dd[4];
    mat4_mul(M, o4, so);
    mat4_mul(M, d4, dd);
    src = {so[0], so[1], so[2]};
    dir = normalized(sub(V3{dd[0], dd[1], dd[2]}, src));
}

struct RayState { float eta; V3 absorption; V3 ffnormal; };   /* ray.h:25-30 */

struct PT {
    orc_scene *s;
    int mode;
    uint32_t max_depth;
    int64_t rays = 0;          /* extension rays traced */
    int64_t shadow = 0;        /* shadow rays whose contribution is not identically zero */

    /* DirectLight (renderer.cpp:335-384), World::SampleEmitter (world.cpp:5-17) */
    V3 direct_light(const RayState &st, V3 ray_dir, uint32_t tri, V3 pos, Rng &rng) {
        V3 ret = v3(0.0f);
        if (s->emitters.empty()) return ret;
        int k = rng.pick((int)s->emitters.size());
        uint32_t emit = s->emitters[k];
        const Tri &et = s->tris[emit];
        /* Triangle::GetSamplePoint (object.cpp:169-178) */
        float sqr1 = std::sqrt(rng.uni(1));
        float r2 = rng.uni(2);
        float a = 1.0f - sqr1;
        float b = (1.0f - r2) * sqr1;
        V3 emit_pos = add(add(muls(et.e1, a), muls(et.e2, b)), et.v[0]);
        V3 dir_to = normalized(sub(emit_pos, pos));
        Hit sh = intersect(s, mode, pos, dir_to, tri);
        {
            /* counting only: Eval has no side effects */
            float pdf_c, lp_c;
            V3 f_c = BsdfEval(mat_of(s, s->tris[tri]), st.eta, neg(ray_dir), st.ffnormal, dir_to, pdf_c);
            V3 e_c = get_emission(s, et, 0.0f, 0.0f, emit_pos, pos, neg(dir_to), lp_c);
            V3 c_c = muls(mulv(muls(f_c, PowerHeuristic(lp_c, pdf_c)), e_c), 1.0f);
            if (pdf_c > 0.0f && !(c_c.x == 0.0f && c_c.y == 0.0f && c_c.z == 0.0f)) ++shadow;
        }
        if (sh.id == kNone || sh.id == emit) {
            float su = sh.id == kNone ? 0.0f : sh.u, sv = sh.id == kNone ? 0.0f : sh.v;
            float lightpdf;
            V3 emission_in = get_emission(s, et, su, sv, emit_pos, pos, neg(dir_to), lightpdf);
            float bsdfpdf;
            V3 f = BsdfEval(mat_of(s, s->tris[tri]), st.eta, neg(ray_dir), st.ffnormal, dir_to, bsdfpdf);
            float weight = PowerHeuristic(lightpdf, bsdfpdf);
            if (bsdfpdf > 0.0f)
                ret = add(ret, divs(muls(mulv(muls(f, weight), emission_in), (float)s->emitters.size()), lightpdf));
        }
        return ret;
    }

    /* PathTracingRenderer::Radiance (renderer.cpp:272-333) */
    V3 radiance(V3 src, V3 dir, uint32_t last_hit, RayState &st, uint32_t depth, float last_bsdfpdf, Rng &rng) {
        if (depth > max_depth - 1) return v3(0.0f);
        Hit h = intersect(s, mode, src, dir, last_hit);
        ++rays;
        V3 rad = v3(0.0f);
        if (h.id != kNone) {
            rng.vertex(depth);
            const Tri &tr = s->tris[h.id];
            const Mat &m = mat_of(s, tr);
            st.ffnormal = normal_c(s, tr, h.u, h.v);
            if (dot(dir, st.ffnormal) > 0.0f) {
                st.ffnormal = neg(st.ffnormal);
                st.eta = m.ior;
                st.absorption = v3(0.0f);
            } else {
                st.eta = 1.0f / m.ior;
            }
            float pdf_e;
            V3 emis = get_emission(s, tr, h.u, h.v, h.pos, src, neg(dir), pdf_e);
            if (depth == 0) rad = add(rad, emis);
            else rad = add(rad, muls(emis, PowerHeuristic(last_bsdfpdf, pdf_e)));
            V3 bounce_ratio = vpow(v3(kE), muls(neg(st.absorption), length(sub(h.pos, src))));
            rad = add(rad, mulv(direct_light(st, dir, h.id, h.pos, rng), bounce_ratio));
            float r1 = rng.uni(3);
            float r2 = rng.uni(4);
            V3 nd, color;
            float bsdfpdf;
            if (BsdfSample(m, st.eta, st.ffnormal, neg(dir), r1, r2, nd, color, bsdfpdf)) {
                st.absorption = divs(neg(vlog(m.extinction)), m.at_distance);
                if (bsdfpdf > 0.0f) {
                    V3 w = divs(mulv(bounce_ratio, color), bsdfpdf);
                    rad = add(rad, mulv(w, radiance(h.pos, nd, h.id, st, depth + 1, bsdfpdf, rng)));
                }
            }
        } else {
            rad = (-dir.z < 0.0f) ? s->sky : s->ground;   /* World::GetDefaultEmission (world.cpp:39-42) */
        }
        return rad;
    }
};

}  // namespace

extern "C" {

orc_scene *orc_scene_create(uint32_t n, const float *verts, const float *norms, const uint32_t *tri_mesh,
                            uint32_t n_meshes, const uint32_t *mesh_mat, const uint8_t *mesh_smooth,
                            uint32_t n_mats, const float *mats, uint32_t n_emit, const uint32_t *emitters,
                            const float *sky, const float *ground) {
    orc_scene *s = new orc_scene();
    s->tris.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        Tri &t = s->tris[i];
        for (int k = 0; k < 3; ++k) {
            t.v[k] = {verts[i * 9 + k * 3 + 0], verts[i * 9 + k * 3 + 1], verts[i * 9 + k * 3 + 2]};
            if (norms) t.n[k] = {norms[i * 9 + k * 3 + 0], norms[i * 9 + k * 3 + 1], norms[i * 9 + k * 3 + 2]};
            else t.n[k] = v3(0.0f);
        }
        t.mesh = tri_mesh[i];
        update_cache(t);
    }
    s->mesh_mat.assign(mesh_mat, mesh_mat + n_meshes);
    s->mesh_smooth.assign(mesh_smooth, mesh_smooth + n_meshes);
    for (uint32_t i = 0; i < n_mats; ++i) s->mats.push_back(load_mat(mats + i * ORC_MAT_FLOATS));
    s->emitters.assign(emitters, emitters + n_emit);
    s->sky = sky ? V3{sky[0], sky[1], sky[2]} : v3(0.0f);
    s->ground = ground ? V3{ground[0], ground[1], ground[2]} : v3(0.0f);
    return s;
}

void orc_scene_destroy(orc_scene *s) { delete s; }

void orc_tri_cache(const orc_scene *s, float *out) {
    for (size_t i = 0; i < s->tris.size(); ++i) {
        const Tri &t = s->tris[i];
        float *o = out + i * 13;
        o[0] = t.v[0].x; o[1] = t.v[0].y; o[2] = t.v[0].z;
        o[3] = t.e1.x; o[4] = t.e1.y; o[5] = t.e1.z;
        o[6] = t.e2.x; o[7] = t.e2.y; o[8] = t.e2.z;
        o[9] = t.normal.x; o[10] = t.normal.y; o[11] = t.normal.z;
        o[12] = t.area;
    }
}

void orc_octree_stats(orc_scene *s, int64_t *nodes, int64_t *leaves, int64_t *max_leaf, float *root_box) {
    ensure_octree(s);
    int64_t nl = 0, ml = 0;
    for (auto &n : s->nodes)
        if (n.is_leaf) { ++nl; ml = std::max<int64_t>(ml, (int64_t)n.items.size()); }
    *nodes = (int64_t)s->nodes.size();
    *leaves = nl;
    *max_leaf = ml;
    const Box &b = s->nodes[0].bbox;
    root_box[0] = b.vmin.x; root_box[1] = b.vmin.y; root_box[2] = b.vmin.z;
    root_box[3] = b.vmax.x; root_box[4] = b.vmax.y; root_box[5] = b.vmax.z;
}

/* Renderer::PrepareScreenSpace (renderer.cpp:103-124) */
void orc_prepare_screen_space(float gate_x, float gate_y, float focal, uint32_t W, uint32_t H, float *top,
                              float *right) {
    float t = (gate_y * 25.4f / 2.0f) / focal;
    float r = (gate_x * 25.4f / 2.0f) / focal;
    float gate_aspr = gate_x / gate_y;
    float film_aspr = float(W) / float(H);
    float xscale = 1.0f, yscale = 1.0f;
    if (gate_aspr >= film_aspr) xscale = film_aspr / gate_aspr;
    else yscale = gate_aspr / film_aspr;
    *right = r * xscale;
    *top = t * yscale;
}

void orc_camera_rays(const float *o2w, float top, float right, uint32_t W, uint32_t H, float *out_src,
                     float *out_dir) {
    float divW = 1.0f / float(W), divH = 1.0f / float(H);
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            V3 s, d;
            camera_ray(o2w, top, right, divW, divH, x, y, s, d);
            size_t p = (size_t)y * W + x;
            out_src[p * 3 + 0] = s.x; out_src[p * 3 + 1] = s.y; out_src[p * 3 + 2] = s.z;
            out_dir[p * 3 + 0] = d.x; out_dir[p * 3 + 1] = d.y; out_dir[p * 3 + 2] = d.z;
        }
}

void orc_intersect(orc_scene *s, int mode, uint64_t nrays, const float *src, const float *dir,
                   const uint32_t *exclude, uint32_t *out_id, float *out_t, float *out_u, float *out_v,
                   int nthreads) {
    if (mode == 1) ensure_octree(s);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < (int64_t)nrays; ++i) {
        V3 o = {src[i * 3], src[i * 3 + 1], src[i * 3 + 2]};
        V3 d = {dir[i * 3], dir[i * 3 + 1], dir[i * 3 + 2]};
        Hit h = intersect(s, mode, o, d, exclude ? exclude[i] : kNone);
        out_id[i] = h.id;
        if (h.id != kNone) {
            if (out_t) out_t[i] = h.t;
            if (out_u) out_u[i] = h.u;
            if (out_v) out_v[i] = h.v;
        }
    }
}

/* NormalRenderer (renderer.cpp:186-229), DepthBufferRenderer (131-174),
 * AlbedoRenderer (391-427), IntersectTestRenderer (20-54). */
void orc_render_aov(orc_scene *s, int mode, const float *o2w, float top, float right, uint32_t W, uint32_t H,
                    float near_clip, float far_clip, uint64_t npix, const uint32_t *pix, uint32_t *out_id,
                    float *out_t, float *out_normal, float *out_depth, float *out_albedo, float *out_mask,
                    int nthreads) {
    if (mode == 1) ensure_octree(s);
    float divW = 1.0f / float(W), divH = 1.0f / float(H);
    float div_fmn = 1.0f / (far_clip - near_clip);
    uint64_t n = pix ? npix : (uint64_t)W * H;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        uint32_t p = pix ? pix[i] : (uint32_t)i;
        uint32_t x = p % W, y = p / W;
        V3 o, d;
        camera_ray(o2w, top, right, divW, divH, x, y, o, d);
        Hit h = intersect(s, mode, o, d, kNone);
        if (out_id) out_id[i] = h.id;
        if (out_t) out_t[i] = h.id != kNone ? h.t : kFloatInfinity;
        if (h.id != kNone) {
            const Tri &tr = s->tris[h.id];
            V3 nrm = normal_c(s, tr, h.u, h.v);
            if (dot(tr.normal, neg(d)) < 0.0f) nrm = neg(nrm);
            V3 nout = divs(add(nrm, v3(1.0f)), 2.0f);
            float dv = (h.t - near_clip) * div_fmn;
            dv = (dv < 0.0f) ? 0.0f : ((1.0f < dv) ? 1.0f : dv);
            V3 alb = mat_of(s, tr).base;
            if (out_normal) { out_normal[i * 3] = nout.x; out_normal[i * 3 + 1] = nout.y; out_normal[i * 3 + 2] = nout.z; }
            if (out_depth) out_depth[i] = dv;
            if (out_albedo) { out_albedo[i * 3] = alb.x; out_albedo[i * 3 + 1] = alb.y; out_albedo[i * 3 + 2] = alb.z; }
            if (out_mask) out_mask[i] = 1.0f;
        } else {
            if (out_normal) { out_normal[i * 3] = 0.0f; out_normal[i * 3 + 1] = 0.0f; out_normal[i * 3 + 2] = 0.0f; }
            if (out_depth) out_depth[i] = 1.0f;
            if (out_albedo) { out_albedo[i * 3] = 0.0f; out_albedo[i * 3 + 1] = 0.0f; out_albedo[i * 3 + 2] = 0.0f; }
            if (out_mask) out_mask[i] = 0.0f;
        }
    }
}

void orc_bsdf_eval(const float *mat, float eta, const float *V, const float *N, const float *L, float *f_out,
                   float *pdf_out) {
    Mat m = load_mat(mat);
    float pdf;
    V3 f = BsdfEval(m, eta, V3{V[0], V[1], V[2]}, V3{N[0], N[1], N[2]}, V3{L[0], L[1], L[2]}, pdf);
    f_out[0] = f.x; f_out[1] = f.y; f_out[2] = f.z;
    *pdf_out = pdf;
}

int orc_bsdf_sample(const float *mat, float eta, const float *ffnormal, const float *in_dir, float r1, float r2,
                    float *out_dir, float *color, float *pdf) {
    Mat m = load_mat(mat);
    V3 od = v3(0.0f), c;
    float p;
    bool ok = BsdfSample(m, eta, V3{ffnormal[0], ffnormal[1], ffnormal[2]}, V3{in_dir[0], in_dir[1], in_dir[2]},
                         r1, r2, od, c, p);
    out_dir[0] = od.x; out_dir[1] = od.y; out_dir[2] = od.z;
    color[0] = c.x; color[1] = c.y; color[2] = c.z;
    *pdf = p;
    return ok ? 1 : 0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }

/* PathTracingRenderer::Render (renderer.cpp:236-270) */
int64_t orc_render_pt(orc_scene *s, int mode, const float *o2w, float top, float right, uint32_t W, uint32_t H,
                      uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, uint32_t spp, uint32_t max_depth,
                      int rng_mode, uint64_t seed, uint64_t warmup, float *out, int nthreads, int64_t *counts) {
    if (mode == 1) ensure_octree(s);
    int64_t total_shadow = 0;
    float divW = 1.0f / float(W), divH = 1.0f / float(H);
    uint32_t RW = x1 - x0, RH = y1 - y0;
    std::memset(out, 0, sizeof(float) * 3 * (size_t)RW * RH);
    int64_t total_rays = 0;
    if (rng_mode == 0) {
        /* one shared engine, reference loop order: sample, row, column */
        std::mt19937 eng;
        for (uint64_t i = 0; i < warmup; ++i) { std::uniform_real_distribution<float> d(0.0f, 1.0f); (void)d(eng); }
        PT pt{s, mode, max_depth};
        Rng rng;
        rng.mode = 0;
        rng.mt = &eng;
        for (uint32_t i = 0; i < spp; ++i)
            for (uint32_t y = y0; y < y1; ++y)
                for (uint32_t x = x0; x < x1; ++x) {
                    V3 o, d;
                    camera_ray(o2w, top, right, divW, divH, x, y, o, d);
                    RayState st{0.0f, v3(0.0f), v3(0.0f)};
                    V3 L = pt.radiance(o, d, kNone, st, 0, 0.0f, rng);
                    size_t q = (size_t)(y - y0) * RW + (x - x0);
                    V3 cur = {out[q * 3], out[q * 3 + 1], out[q * 3 + 2]};
                    cur = add(cur, divs(L, (float)spp));
                    out[q * 3] = cur.x; out[q * 3 + 1] = cur.y; out[q * 3 + 2] = cur.z;
                }
        if (counts) { counts[0] = pt.rays; counts[1] = pt.shadow; }
        return pt.rays + pt.shadow;
    }
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total_rays, total_shadow) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t q = 0; q < (int64_t)RW * RH; ++q) {
        uint32_t x = x0 + (uint32_t)(q % RW), y = y0 + (uint32_t)(q / RW);
        PT pt{s, mode, max_depth};
        Rng rng;
        rng.mode = 1;
        rng.key[0] = key[0]; rng.key[1] = key[1];
        rng.pixel = y * W + x;
        V3 acc = v3(0.0f);
        V3 o, d;
        camera_ray(o2w, top, right, divW, divH, x, y, o, d);
        for (uint32_t i = 0; i < spp; ++i) {
            rng.sample = i;
            RayState st{0.0f, v3(0.0f), v3(0.0f)};
            V3 L = pt.radiance(o, d, kNone, st, 0, 0.0f, rng);
            acc = add(acc, divs(L, (float)spp));
        }
        out[q * 3] = acc.x; out[q * 3 + 1] = acc.y; out[q * 3 + 2] = acc.z;
        total_rays += pt.rays;
        total_shadow += pt.shadow;
    }
    if (counts) { counts[0] = total_rays; counts[1] = total_shadow; }
    return total_rays + total_shadow;
}

void orc_pt_samples(orc_scene *s, int mode, const float *o2w, float top, float right, uint32_t W, uint32_t H,
                    uint64_t n, const uint32_t *pixel, const uint32_t *sample, uint32_t max_depth, uint64_t seed,
                    float *out, int nthreads) {
    if (mode == 1) ensure_octree(s);
    float divW = 1.0f / float(W), divH = 1.0f / float(H);
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        uint32_t p = pixel[i];
        PT pt{s, mode, max_depth};
        Rng rng;
        rng.mode = 1;
        rng.key[0] = key[0]; rng.key[1] = key[1];
        rng.pixel = p;
        rng.sample = sample[i];
        V3 o, d;
        camera_ray(o2w, top, right, divW, divH, p % W, p / W, o, d);
        RayState st{0.0f, v3(0.0f), v3(0.0f)};
        V3 L = pt.radiance(o, d, kNone, st, 0, 0.0f, rng);
        out[i * 3] = L.x; out[i * 3 + 1] = L.y; out[i * 3 + 2] = L.z;
    }
}

}  // extern "C"
