// Procedural materials and lighting of the ARBOR render path (Arbor.cu:256-289),
// evaluated per pixel from a loaded primitive record.
#pragma once
#include "arbor_trace.hpp"

namespace arbor {

AD v3 sky(v3 rd, v3 sun) {
  float h = sat(rd.y);
  v3 c = mix(mk(0.75f, 0.79f, 0.77f), mk(0.19f, 0.35f, 0.54f), powf(h, 0.48f));
  float warm = powf(sat(dot(rd, sun)), 32.0f);
  c = c + mk(0.48f, 0.26f, 0.075f) * warm;
  float disc = smooth(0.99996f, 0.999984f, dot(rd, sun));
  c = c + mk(8.0f, 6.7f, 4.9f) * disc;
  float clouds = noise(rd.x * 5.0f / (0.2f + h), rd.z * 5.0f / (0.2f + h), 231);
  float veil = smooth(0.62f, 0.81f, clouds) * smooth(0.02f, 0.18f, h) * 0.38f;
  return mix(c, mk(1.1f, 1.10f, 1.05f), veil);
}
__device__ inline v3 barkColor(const Prim& pr, v3 p, float footprint) {
  int seed = (int)pr.seed();
  v2 uv = woodUV(pr, p);
  v3 coord = mk(cosf(uv.x) * 9.0f, sinf(uv.x) * 9.0f, uv.y * 7.0f);
  float mottle = noise3(p * 3.0f, 714);
  float rough = noise3(coord, seed);
  v3 d = pr.b() - pr.a();
  float t = sat(dot(p - pr.a(), d) / fmaxf(0.000001f, dot(d, d)));
  float amp = fminf(0.021f, lerp(pr.ra(), pr.rb(), t) * 0.055f);
  float cut = barkCut(pr, p, footprint) / fmaxf(0.00001f, amp);
  v3 c = mix(mk(0.10f, 0.075f, 0.050f), mk(0.25f, 0.21f, 0.16f), 0.25f + rough * 0.62f);
  c = c * (0.75f + 0.35f * mottle);
  c = c * (1.0f - 0.62f * smooth(0.28f, 0.85f, cut));
  float moss = smooth(0.52f, 0.72f, noise3(p * 5.0f, 189)) * (1.0f - smooth(0.15f, 1.4f, p.y));
  c = mix(c, mk(0.045f, 0.064f, 0.018f), moss * 0.62f);
  float lichen = smooth(0.65f, 0.77f, noise3(p * 23.0f, 27)) * smooth(0.25f, 0.65f, noise3(p * 90.0f, 712)) *
                 bandWeight(0.018f, footprint);
  c = mix(c, mk(0.31f, 0.33f, 0.23f), lichen * 0.5f);
  float grain = (noise3(coord * 15.0f, seed) - 0.5f) * bandWeight(0.0025f, footprint);
  return c * (1.0f + grain * 0.38f);
}
AD float leafVein(const Prim& pr, v3 u, v3 p, float footprint) {
  v3 q = p - pr.a();
  float x = dot(q, u) - pr.f16();
  float y = dot(q, pr.c());
  float len = pr.ra();
  float mid = expf(-fabsf(y) / fmaxf(0.00065f, footprint * 0.55f));
  float fy = frac((x - fabsf(y) * 0.70f) / len * 7.0f);
  float dd = fminf(fy, 1.0f - fy) * len / 7.0f;
  float lateral = expf(-dd / fmaxf(0.00033f, footprint * 0.5f));
  return sat(mid * 0.82f + lateral * 0.48f) * bandWeight(0.003f, footprint);
}
AD v3 leafColor(const Prim& pr, v3 u, v3 p, float fp) {
  float h = rnd((int)pr.seed(), 887);
  v3 q = p - pr.a();
  float x = dot(q, u) - pr.f16();
  if (x < 0.0f) return mk(0.17f, 0.15f, 0.048f);
  float y = dot(q, pr.c());
  float v = leafVein(pr, u, p, fp);
  v3 c = mix(mk(0.035f, 0.11f, 0.008f), mk(0.15f, 0.255f, 0.025f), h);
  c = mix(c, mk(0.22f, 0.26f, 0.055f), v * 0.57f);
  float speck = (noise(x * 700.0f, y * 700.0f, (int)pr.seed()) - 0.5f) * bandWeight(0.0017f, fp);
  float edge = smooth(0.78f, 1.0f, fabsf(y) / fmaxf(0.001f, leafWidth(x, pr.ra(), pr.rb(), fp)));
  c = c * (1.0f + speck * 0.27f);
  c = mix(c, mk(0.23f, 0.20f, 0.05f), edge * 0.22f);
  return c;
}
AD v3 groundColor(v3 p, float fp) {
  float a = noise(p.x * 3.0f, p.z * 3.0f, 819);
  float b = noise(p.x * 15.0f, p.z * 15.0f, 28);
  float r = sqrtf(p.x * p.x + p.z * p.z);
  float soil = 1.0f - smooth(4.1f, 6.8f, r + noise(p.x, p.z, 13));
  v3 grass = mix(mk(0.085f, 0.115f, 0.035f), mk(0.23f, 0.24f, 0.092f), a);
  v3 dirt = mix(mk(0.075f, 0.055f, 0.035f), mk(0.16f, 0.13f, 0.080f), b);
  float moss = smooth(0.53f, 0.73f, a);
  dirt = mix(dirt, mk(0.049f, 0.079f, 0.022f), moss * 0.72f);
  float grit = (noise(p.x * 200.0f, p.z * 200.0f, 127) - 0.5f) * bandWeight(0.007f, fp);
  return mix(grass, dirt, soil) * (1.0f + grit * 0.40f);
}
AD float specGGX(v3 n, v3 v, v3 l, float rough) {
  v3 h = unit(v + l);
  float nh = sat(dot(n, h));
  float nv = fmaxf(0.002f, sat(dot(n, v)));
  float nl = sat(dot(n, l));
  float vh = sat(dot(v, h));
  float a = rough * rough, aa = a * a;
  float den = nh * nh * (aa - 1.0f) + 1.0f;
  float D = aa / (kPi * den * den);
  float k = (rough + 1.0f) * (rough + 1.0f) * 0.125f;
  float G = nv / (nv * (1.0f - k) + k) * nl / (nl * (1.0f - k) + k);
  float F = 0.035f + 0.965f * powf(1.0f - vh, 5.0f);
  return D * G * F / (4.0f * nv + 0.0001f);
}
// Up to four shadow traversals through leaves (Arbor.cu:271-274). The
// sun-visibility pass inlines this (its kernel then needs 121 VGPRs, 4 waves
// per SIMD); the fused shading paths call the out-of-line lightVisibility
// below, so their own register budget is not added to the traversal's.
template <class W, class Trees>
__device__ __forceinline__ v3 lightVisibilityInline(const float* P, const float4* Nodes4, const float4* top, const uint32_t* Order,
                                     Trees trees, v3 p, v3 direction, float time, float wind) {
  v3 transmission = mk(1.0f, 1.0f, 1.0f);
  v3 ro = p;
  for (int j = 0; j < 4; j++) {
    int tree = -1;
    v2 h = traceForest<W>(P, Nodes4, top, Order, trees, ro, direction, 0.0001f, time, wind, 36.0f, &tree, true);
    int id = (int)h.y;
    if (id < 0) break;
    if (P[id * kPS + 3] != 1.0f) return mk(0.0f, 0.0f, 0.0f);
    transmission = transmission * mk(0.18f, 0.29f, 0.065f);
    ro = ro + direction * (h.x + 0.0018f);
  }
  return transmission;
}
template <class W>
__device__ __noinline__ v3 lightVisibility(const float* P, const float4* Nodes4, const float4* top, const uint32_t* Order,
                                           const TreeInst* trees, v3 p, v3 direction, float time, float wind) {
  return lightVisibilityInline<W, const TreeInst*>(P, Nodes4, top, Order, trees, p, direction, time, wind);
}


// Shading state shared by every pixel of one view (per view, not per pixel:
// the same values shadePixels recomputes per pixel).
struct ShadeCtx {
  View view;
  v3 sun, su, sv;
  float time, wind, mode;
  bool shadows;
  int sample, width, height;
};
AD ShadeCtx makeShadeCtx(const float* C, int width, int height) {
  ShadeCtx sc;
  sc.view = makeView(C);
  sc.sun = unit(mk(cos_cr(C[12]) * cos_cr(C[13]), sin_cr(C[13]), sin_cr(C[12]) * cos_cr(C[13])));
  sc.su = basisU(sc.sun);
  sc.sv = cross(sc.sun, sc.su);
  sc.time = C[8];
  sc.wind = C[11];
  sc.mode = C[15];
  sc.shadows = C[22] > 0.5f;
  sc.sample = (int)C[9];
  sc.width = width;
  sc.height = height;
  return sc;
}

// The jittered sun direction of pixel `pixel` for the current sample and the
// shadow ray's origin off hit point p (Arbor.cu:282-284): shading and the
// separate sun-visibility pass both take them from here.
struct SunRay { v3 light, origin; };
AD SunRay sunRayOf(const ShadeCtx& sc, int pixel, v3 p, int id) {
  float r1 = rnd(pixel + sc.sample * 1823, 991), r2 = rnd(pixel + sc.sample * 971, 731);
  float angle = r1 * 6.28318f, disk = sqrtf(r2) * 0.012f;
  SunRay s;
  s.light = unit(sc.sun + sc.su * (cosf(angle) * disk) + sc.sv * (sinf(angle) * disk));
  float offset = id >= 0 ? 0.0014f : 0.004f;
  s.origin = p + s.light * offset;
  return s;
}

// Sun visibility of one pixel from its Hit record: lightVisibility along its
// sun ray, or 1 for sky or with tree-to-tree visibility off.
template <class W, class Trees>
__device__ inline v3 sunVisibilityOne(const float* P, const float4* Nodes4, const float4* top, const uint32_t* Order,
                                      Trees trees, const ShadeCtx& sc, int x, int y, float4 hit) {
  if (!sc.shadows || !(hit.x < kFar)) return mk(1.0f, 1.0f, 1.0f);
  v3 rd = viewRay(sc.view, x, y, sc.width, sc.height);
  v3 p = sc.view.eye + rd * hit.x;
  SunRay s = sunRayOf(sc, y * sc.width + x, p, (int)hit.y);
  return lightVisibilityInline<W>(P, Nodes4, top, Order, trees, s.origin, s.light, sc.time, sc.wind);
}

// One pixel of shadePixels (Arbor.cu:277-289): Linear and Guide of pixel
// (x, y) from its Hit record; with outNormal, also the shading normal (facing
// the camera; zero for sky). With vis, the sun visibility comes from the
// separate pass (sunVisibilityOne: same rays, same value) instead of being
// traced here; kTraceVis = false compiles the tracing out (callers then pass
// vis whenever sun visibility is on).
template <class W, bool kTraceVis = true>
__device__ inline void shadeOne(const float* P, const float4* Nodes4, const uint32_t* Order, const TreeInst* trees,
                                const ShadeCtx& sc, int x, int y, float4 hit, float4& outLinear, float4& outGuide,
                                v3* outNormal = nullptr, const float4* vis = nullptr) {
  const int pixel = y * sc.width + x;
  float tt = hit.x;
  int id = (int)hit.y, tree = (int)hit.z;
  v3 ro = sc.view.eye, rd = viewRay(sc.view, x, y, sc.width, sc.height);
  v3 col = sky(rd, sc.sun);
  v3 guide = mk(1.0f, 1.0f, 1.0f);
  if (outNormal) *outNormal = mk(0.0f, 0.0f, 0.0f);
  if (tt < kFar) {
    v3 p = ro + rd * tt;
    float fp = fmaxf(0.000005f, hit.w);
    v3 localP = p, n = mk(0.0f, 1.0f, 0.0f);
    float localFp = fp, phase = 0.0f;
    Prim pr;
    bool leaf = false;
    if (id >= 0) {
      pr = loadPrim(P, id);
      leaf = pr.type() == 1.0f;
    }
    v3 leafUdir = mk(0.0f, 0.0f, 0.0f);
    if (id >= 0 && tree >= 0) {
      const TreeInst& ti = trees[tree];
      localP = toLocalPoint(p, ti);
      localFp = fp / ti.sc;
      phase = ti.phase;
      n = toWorldNormal(objectNormal(pr, localP, localFp, sc.time + phase, sc.wind), ti);
    }
    if (leaf) leafUdir = leafFrame(pr, sc.time + phase, sc.wind).u;
    if (dot(n, rd) > 0.0f) n = n * -1.0f;
    if (id < 0) {
      float e = fmaxf(0.001f, fp);
      float nx = (noise((p.x + e) * 18.0f, p.z * 18.0f, 133) - noise((p.x - e) * 18.0f, p.z * 18.0f, 133)) * 0.18f;
      float nz = (noise(p.x * 18.0f, (p.z + e) * 18.0f, 133) - noise(p.x * 18.0f, (p.z - e) * 18.0f, 133)) * 0.18f;
      n = unit(mk(nx, 1.0f, nz));
    }
    if (outNormal) *outNormal = n;
    v3 albedo;
    if (id < 0) albedo = groundColor(p, localFp);
    else {
      int type = (int)pr.type();
      if (type == 0) albedo = barkColor(pr, localP, localFp);
      else if (type == 1) albedo = leafColor(pr, leafUdir, localP, localFp);
      else albedo = mix(mk(0.09f, 0.10f, 0.086f), mk(0.31f, 0.30f, 0.24f), noise3(localP * 42.0f, (int)pr.seed()));
    }
    if (tree >= 0) {
      int ts = trees[tree].seed;
      float hue = rnd(8, ts), age = rnd(9, ts), season = rnd(10, ts);
      float dark = 0.76f + age * 0.38f;
      v3 tint = mk(0.88f + 0.22f * hue, 0.84f + 0.28f * (1.0f - fabsf(hue - 0.55f)), 0.72f + 0.25f * (1.0f - hue));
      if (leaf) tint = mix(tint, mk(1.12f, 0.88f, 0.48f), smooth(0.78f, 0.98f, season) * 0.38f);
      else tint = mix(tint, mk(0.78f, 0.66f, 0.53f), season * 0.18f);
      albedo = albedo * tint * dark;
    }
    guide = albedo;
    const SunRay sray = sunRayOf(sc, pixel, p, id);
    const v3 light = sray.light;
    v3 visibility = mk(1.0f, 1.0f, 1.0f);
    if (sc.shadows) {
      if (!kTraceVis || vis) {
        const float4 v = vis[pixel];
        visibility = mk(v.x, v.y, v.z);
      } else {
        visibility = lightVisibility<W>(P, Nodes4, nullptr, Order, trees, sray.origin, light, sc.time, sc.wind);
      }
    }
    float nl = sat(dot(n, light)), back = sat(-dot(n, light));
    v3 direct = albedo * (nl / kPi);
    if (leaf) {
      float vein = leafVein(pr, leafUdir, localP, localFp);
      direct = direct + albedo * (back * (0.75f - 0.25f * vein) + powf(sat(dot(rd, light)), 9.0f) * 0.12f);
    }
    float spec = specGGX(n, rd * -1.0f, light, leaf ? 0.39f : 0.82f);
    direct = direct + mk(spec, spec, spec);
    const float ao = 1.0f;
    v3 ambient = mk(0.30f, 0.37f, 0.45f) * (0.45f + 0.55f * sat(n.y)) * ao + mk(0.08f, 0.068f, 0.035f) * sat(-n.y);
    if (leaf) ambient = ambient + mk(0.04f, 0.055f, 0.022f) * ao;
    col = direct * visibility * mk(4.8f, 4.2f, 3.1f) + albedo * ambient;
    float haze = 1.0f - expf(-tt * 0.0018f);
    col = mix(col, sky(rd, sc.sun), haze);
    if (sc.mode == 1.0f) {
      float w0 = bandWeight(0.065f, localFp), w1 = bandWeight(0.005f, localFp), w2 = bandWeight(0.001f, localFp);
      col = mk(w0 * 0.80f, w1 * 0.85f, w2 * 0.9f);
    }
    if (sc.mode == 2.0f) {
      float hh = rnd(id + 177 + (tree + 1) * 71, 32);
      col = id < 0 ? mk(0.10f, 0.10f, 0.10f)
                   : (leaf ? mk(0.1f, 0.3f + hh * 0.5f, 0.17f) : mk(0.47f + hh * 0.3f, 0.25f, 0.10f));
    }
    if (sc.mode == 3.0f) col = n * 0.5f + mk(0.5f, 0.5f, 0.5f);
  }
  outLinear = make_float4(col.x, col.y, col.z, 1.0f);
  outGuide = make_float4(guide.x, guide.y, guide.z, 1.0f);
}

}  // namespace arbor
