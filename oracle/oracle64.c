/*
 * oracle64.c -- FP64 forward, analytic adjoint and central finite differences
 * of the deferred-lighting direct + ambient terms (test infrastructure only;
 * see oracle.h).
 *
 * Forward per pixel (deferred_lighting.wgsl:991-1259 with the fallbacks of the
 * FP32 restatement, all lights evaluated, F0 resolved from albedo/metallic as
 * gbuffer.wgsl:324-328 does):
 *   out_k = sum_i (kD_k a_k / pi + D G F_k / (4 NdV NdL + 1e-4)) rad_k NdL
 *           + mix(0.15 sky_k, sky_k, 0.5 N.y + 0.5) a_k ao + emissive_k
 * pbr_direct_light (deferred_lighting.wgsl:699-820) supplies L, atten, rad and
 * the cut-offs; pbr_eval.wgsl:6-66 supplies D, G, F.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double PI64 = 3.14159265359;
static const double EPS = 0.0001;

static inline double d_dot(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void d_normalize(double v[3]) {
    double l = sqrt(d_dot(v, v));
    v[0] /= l;
    v[1] /= l;
    v[2] /= l;
}

static double d_half(uint16_t h) {
    int e = (h >> 10) & 0x1f, m = h & 0x3ff;
    double s = (h >> 15) ? -1.0 : 1.0;
    if (e == 0) return s * ldexp((double)m, -24);
    if (e == 31) return m ? NAN : s * INFINITY;
    return s * ldexp((double)(m | 0x400), e - 25);
}

static void d_load(const hl_texture* t, uint32_t x, uint32_t y, double out[4]) {
    const uint8_t* row = (const uint8_t*)t->data_d + (size_t)y * t->pitch_bytes;
    for (int k = 0; k < 4; ++k) {
        switch (t->format) {
            case HL_FMT_RGBA8_UNORM: out[k] = (double)row[4 * (size_t)x + k] / 255.0; break;
            case HL_FMT_RGBA16_FLOAT: {
                uint16_t h;
                memcpy(&h, row + 8 * (size_t)x + 2 * k, 2);
                out[k] = d_half(h);
                break;
            }
            case HL_FMT_RGBA32_FLOAT: {
                float f;
                memcpy(&f, row + 16 * (size_t)x + 4 * k, 4);
                out[k] = f;
                break;
            }
            default: out[k] = 0.0;
        }
    }
}

void or64_decode(const GpuCameraUniforms* cam, const hl_gbuffer* gb, or64_pixel* px) {
#pragma omp parallel for schedule(static)
    for (int64_t yy = 0; yy < (int64_t)gb->height; ++yy) {
        uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < gb->width; ++x) {
            or64_pixel* p = &px[(size_t)y * gb->width + x];
            float depth;
            memcpy(&depth, (const uint8_t*)gb->depth.data_d + (size_t)y * gb->depth.pitch_bytes + 4 * (size_t)x, 4);
            memset(p, 0, sizeof(*p));
            p->sky = depth >= 1.0f;
            if (p->sky) continue;
            double a[4], n[4], o[4], e[4];
            d_load(&gb->albedo, x, y, a);
            d_load(&gb->normal, x, y, n);
            d_load(&gb->orm, x, y, o);
            d_load(&gb->emissive, x, y, e);
            for (int k = 0; k < 3; ++k) {
                p->albedo[k] = a[k];
                p->N[k] = n[k];
                p->emissive[k] = e[k];
            }
            p->alpha = a[3];
            d_normalize(p->N);
            p->ao = o[0];
            p->rough = o[1];
            p->metal = o[2];
            /* The world position is the FP32 reconstruction of the shader (its
             * rounding near depth 1 is a property of the input, not of the
             * lighting arithmetic this path evaluates in FP64). */
            float wp32[3];
            or_world_pos(cam, gb->width, gb->height, x, y, depth, wp32);
            for (int k = 0; k < 3; ++k) {
                p->wp[k] = wp32[k];
                p->V[k] = (double)cam->position_near[k] - p->wp[k];
            }
            d_normalize(p->V);
        }
    }
}

/* Everything one (pixel, light) term needs for forward and adjoint. */
typedef struct term {
    int active;
    double L[3], d, att0, fall, s, atten, t_spot, dsdcos;
    double rad[3], NdL, h[3], hlen, H[3], HdV, u, P;
    double F0[3], F[3], kD[3];
    int f0_free[3];
    double NdV, alpha2, NdH, den, Q, D, k, EV, EL, gV, gLt, G, den2, B[3];
} term;

static void d_f0(const or64_pixel* p, double F0[3], int free_[3]) {
    for (int c = 0; c < 3; ++c) {
        double v = 0.04 * (1.0 - p->metal) + p->albedo[c] * p->metal;
        free_[c] = (v > 0.0 && v < 0.999);
        F0[c] = v < 0.0 ? 0.0 : (v > 0.999 ? 0.999 : v);
    }
}

static void d_term(const or64_pixel* p, const GpuLight* l, term* t) {
    memset(t, 0, sizeof(*t));
    const double ci[3] = {l->color_intensity[0], l->color_intensity[1], l->color_intensity[2]};
    const double I = l->color_intensity[3];
    if (l->light_type == 0u) {
        double L[3] = {-(double)l->direction_outer[0], -(double)l->direction_outer[1],
                       -(double)l->direction_outer[2]};
        d_normalize(L);
        memcpy(t->L, L, sizeof L);
        t->atten = 1.0;
        for (int c = 0; c < 3; ++c) t->rad[c] = ci[c] * I;
    } else {
        double tl[3] = {l->position_range[0] - p->wp[0], l->position_range[1] - p->wp[1],
                        l->position_range[2] - p->wp[2]};
        double d = sqrt(d_dot(tl, tl));
        const double R = l->position_range[3];
        if (d > R) return;
        t->d = d;
        for (int c = 0; c < 3; ++c) t->L[c] = tl[c] / d;
        t->att0 = 1.0 / (d * d + EPS);
        double nd = d / R;
        t->fall = 1.0 - nd * nd * nd * nd;
        if (t->fall < 0.0) t->fall = 0.0;
        t->s = 1.0;
        if (l->light_type == 2u) {
            double cos_a = -(t->L[0] * l->direction_outer[0] + t->L[1] * l->direction_outer[1] +
                             t->L[2] * l->direction_outer[2]);
            double e0 = l->direction_outer[3], e1 = l->inner_angle;
            double tt = (cos_a - e0) / (e1 - e0);
            double tc = tt < 0.0 ? 0.0 : (tt > 1.0 ? 1.0 : tt);
            t->t_spot = tc;
            t->s = tc * tc * (3.0 - 2.0 * tc);
            t->dsdcos = (tt > 0.0 && tt < 1.0) ? 6.0 * tc * (1.0 - tc) / (e1 - e0) : 0.0;
        }
        t->atten = t->att0 * t->fall * t->s;
        for (int c = 0; c < 3; ++c) t->rad[c] = ci[c] * I * t->atten;
    }
    double ndl = d_dot(p->N, t->L);
    t->NdL = ndl > 0.0 ? ndl : 0.0;
    if (t->NdL == 0.0) return;
    if (t->rad[0] < 0.002 && t->rad[1] < 0.002 && t->rad[2] < 0.002) return;
    t->active = 1;
    for (int c = 0; c < 3; ++c) t->h[c] = p->V[c] + t->L[c];
    t->hlen = sqrt(d_dot(t->h, t->h));
    for (int c = 0; c < 3; ++c) t->H[c] = t->h[c] / t->hlen;
    double hdv = d_dot(t->H, p->V);
    t->HdV = hdv > 0.0 ? hdv : 0.0;
    t->u = 1.0 - t->HdV;
    if (t->u < 0.0) t->u = 0.0;
    if (t->u > 1.0) t->u = 1.0;
    t->P = t->u * t->u * t->u * t->u * t->u;
    d_f0(p, t->F0, t->f0_free);
    for (int c = 0; c < 3; ++c) {
        t->F[c] = t->F0[c] + (1.0 - t->F0[c]) * t->P;
        t->kD[c] = (1.0 - t->F[c]) * (1.0 - p->metal);
    }
    double ndv = d_dot(p->N, p->V);
    t->NdV = ndv > 0.0 ? ndv : 0.0;
    double a = p->rough * p->rough;
    t->alpha2 = a * a;
    double ndh = d_dot(p->N, t->H);
    t->NdH = ndh > 0.0 ? ndh : 0.0;
    t->den = t->NdH * t->NdH * (t->alpha2 - 1.0) + 1.0;
    t->Q = PI64 * t->den * t->den + EPS;
    t->D = t->alpha2 / t->Q;
    double r1 = p->rough + 1.0;
    t->k = r1 * r1 / 8.0;
    t->EV = t->NdV * (1.0 - t->k) + t->k + EPS;
    t->EL = t->NdL * (1.0 - t->k) + t->k + EPS;
    t->gV = t->NdV / t->EV;
    t->gLt = t->NdL / t->EL;
    t->G = t->gV * t->gLt;
    t->den2 = 4.0 * t->NdV * t->NdL + EPS;
    for (int c = 0; c < 3; ++c)
        t->B[c] = t->kD[c] * p->albedo[c] / PI64 + t->D * t->G * t->F[c] / t->den2;
}

static inline uint64_t mix_hash(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h;
}

/* Forward of one pixel with light `mod_i` replaced by *mod (mod may be NULL). */
static uint64_t d_pixel(const or64_pixel* p, const DeferredGlobals* g, const GpuLight* lights,
                        uint32_t n, uint32_t mod_i, const GpuLight* mod, double out[3]) {
    uint64_t h = 1469598103934665603ull;
    double Lo[3] = {0, 0, 0};
    for (uint32_t i = 0; i < n; ++i) {
        const GpuLight* l = (mod && i == mod_i) ? mod : &lights[i];
        term t;
        d_term(p, l, &t);
        if (!t.active) continue;
        h = mix_hash(h, i);
        h = mix_hash(h, (uint64_t)(t.fall > 0.0) | ((uint64_t)(t.t_spot > 0.0 && t.t_spot < 1.0) << 1));
        for (int c = 0; c < 3; ++c) Lo[c] += t.B[c] * t.rad[c] * t.NdL;
    }
    int f0f[3];
    double F0[3];
    d_f0(p, F0, f0f);
    h = mix_hash(h, (uint64_t)(f0f[0] | (f0f[1] << 1) | (f0f[2] << 2)));
    double ht = p->N[1] * 0.5 + 0.5;
    for (int c = 0; c < 3; ++c) {
        double sky = (double)g->ambient_color[c] * (double)g->ambient_intensity;
        double hemi = (0.15 * sky * (1.0 - ht) + sky * ht) * p->albedo[c];
        out[c] = Lo[c] + hemi * p->ao + p->emissive[c];
    }
    return h;
}

uint64_t or64_shade_pixel(const or64_pixel* p, const DeferredGlobals* g, const GpuLight* lights,
                          uint32_t num_lights, double out[3]) {
    return d_pixel(p, g, lights, num_lights, 0, NULL, out);
}

static inline const float* tex_px(const float* base, uint32_t pitch, uint32_t x, uint32_t y) {
    return (const float*)((const uint8_t*)base + (size_t)y * pitch) + 4 * (size_t)x;
}

double or64_forward(const or64_pixel* px, uint32_t width, uint32_t height,
                    const DeferredGlobals* g, const GpuLight* lights, uint32_t num_lights,
                    double* out, const float* target, uint32_t target_pitch) {
    double loss = 0.0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : loss)
    for (int64_t yy = 0; yy < (int64_t)height; ++yy) {
        uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < width; ++x) {
            const or64_pixel* p = &px[(size_t)y * width + x];
            double o[3] = {0, 0, 0};
            if (!p->sky) d_pixel(p, g, lights, num_lights, 0, NULL, o);
            if (out) memcpy(out + 3 * ((size_t)y * width + x), o, sizeof o);
            if (target && !p->sky) {
                const float* tg = tex_px(target, target_pitch, x, y);
                for (int c = 0; c < 3; ++c) loss += (o[c] - tg[c]) * (o[c] - tg[c]);
            }
        }
    }
    return loss;
}

/* Adjoint of one pixel given dL/dout (gl).  Per-pixel grads into dpx[5];
 * per-light grads accumulated into dpos/dint (caller serializes). */
static void d_pixel_adjoint(const or64_pixel* p, const DeferredGlobals* g, const GpuLight* lights,
                            uint32_t n, const double gl[3], double dpx[5], double* dpos,
                            double* dint) {
    double ht = p->N[1] * 0.5 + 0.5;
    for (int c = 0; c < 3; ++c) {
        double sky = (double)g->ambient_color[c] * (double)g->ambient_intensity;
        double mixv = 0.15 * sky * (1.0 - ht) + sky * ht;
        dpx[c] += gl[c] * mixv * p->ao;
    }
    const double m = p->metal, r = p->rough;
    for (uint32_t i = 0; i < n; ++i) {
        const GpuLight* l = &lights[i];
        term t;
        d_term(p, l, &t);
        if (!t.active) continue;
        const double I = l->color_intensity[3];
        double W[3], dBdF[3];
        for (int c = 0; c < 3; ++c) {
            W[c] = t.rad[c] * t.NdL;
            dBdF[c] = -(1.0 - m) * p->albedo[c] / PI64 + t.D * t.G / t.den2;
        }
        /* per-pixel material */
        double dDdr = (1.0 / t.Q - t.alpha2 * 2.0 * PI64 * t.den * t.NdH * t.NdH / (t.Q * t.Q)) *
                      4.0 * r * r * r;
        double dgdk_v = -t.NdV * (1.0 - t.NdV) / (t.EV * t.EV);
        double dgdk_l = -t.NdL * (1.0 - t.NdL) / (t.EL * t.EL);
        double dGdr = (dgdk_v * t.gLt + t.gV * dgdk_l) * (r + 1.0) / 4.0;
        for (int c = 0; c < 3; ++c) {
            double gw = gl[c] * W[c];
            double f0a = t.f0_free[c] ? m * (1.0 - t.P) * dBdF[c] : 0.0;
            double f0m = t.f0_free[c] ? (p->albedo[c] - 0.04) * (1.0 - t.P) * dBdF[c] : 0.0;
            dpx[c] += gw * (t.kD[c] / PI64 + f0a);
            dpx[4] += gw * (-(1.0 - t.F[c]) * p->albedo[c] / PI64 + f0m);
            dpx[3] += gw * (dDdr * t.G + t.D * dGdr) * t.F[c] / t.den2;
        }
        /* per-light intensity */
        double di = 0.0;
        for (int c = 0; c < 3; ++c) di += gl[c] * t.B[c] * (double)l->color_intensity[c] * t.atten * t.NdL;
        dint[i] += di;
        if (l->light_type == 0u) continue;
        /* per-light position: reverse mode through L, d */
        double adj_NdL = 0.0, adj_atten = 0.0, adj_D = 0.0, adj_G = 0.0, adj_den2 = 0.0, adj_P = 0.0;
        for (int c = 0; c < 3; ++c) {
            double adj_B = gl[c] * W[c];
            double adj_rad = gl[c] * t.B[c] * t.NdL;
            adj_NdL += gl[c] * t.B[c] * t.rad[c];
            adj_atten += adj_rad * (double)l->color_intensity[c] * I;
            double adj_F = adj_B * dBdF[c];
            adj_D += adj_B * t.G * t.F[c] / t.den2;
            adj_G += adj_B * t.D * t.F[c] / t.den2;
            adj_den2 -= adj_B * t.D * t.G * t.F[c] / (t.den2 * t.den2);
            adj_P += adj_F * (1.0 - t.F0[c]);
        }
        double adj_H[3] = {0, 0, 0};
        if (t.HdV > 0.0 && t.u > 0.0) {
            double adj_HdV = adj_P * (-5.0 * t.u * t.u * t.u * t.u);
            for (int c = 0; c < 3; ++c) adj_H[c] += adj_HdV * p->V[c];
        }
        if (t.NdH > 0.0) {
            double dDdNdH = -t.alpha2 / (t.Q * t.Q) * (2.0 * PI64 * t.den) * (2.0 * t.NdH * (t.alpha2 - 1.0));
            double adj_NdH = adj_D * dDdNdH;
            for (int c = 0; c < 3; ++c) adj_H[c] += adj_NdH * p->N[c];
        }
        adj_NdL += adj_G * t.gV * (t.k + EPS) / (t.EL * t.EL);
        adj_NdL += adj_den2 * 4.0 * t.NdV;
        double Hd = d_dot(t.H, adj_H);
        double adj_L[3];
        for (int c = 0; c < 3; ++c) adj_L[c] = (adj_H[c] - t.H[c] * Hd) / t.hlen + adj_NdL * p->N[c];
        double adj_att0 = adj_atten * t.fall * t.s;
        double adj_fall = adj_atten * t.att0 * t.s;
        double adj_s = adj_atten * t.att0 * t.fall;
        const double R = l->position_range[3];
        double adj_d = adj_att0 * (-2.0 * t.d * t.att0 * t.att0);
        if (t.fall > 0.0) adj_d += adj_fall * (-4.0 * t.d * t.d * t.d / (R * R * R * R));
        if (l->light_type == 2u && t.dsdcos != 0.0) {
            double adj_cos = adj_s * t.dsdcos;
            for (int c = 0; c < 3; ++c) adj_L[c] += adj_cos * (-(double)l->direction_outer[c]);
        }
        double Ld = d_dot(t.L, adj_L);
        for (int c = 0; c < 3; ++c) dpos[3 * i + c] += (adj_L[c] - t.L[c] * Ld) / t.d + adj_d * t.L[c];
    }
}

static void d_backward(const or64_pixel* px, uint32_t width, uint32_t height,
                       const DeferredGlobals* g, const GpuLight* lights, uint32_t n,
                       const float* buf, uint32_t pitch, int is_target, double* d_px_out,
                       double* d_pos, double* d_int) {
    const size_t np = (size_t)width * height;
    memset(d_px_out, 0, sizeof(double) * 5 * np);
    memset(d_pos, 0, sizeof(double) * 3 * n);
    memset(d_int, 0, sizeof(double) * n);
#pragma omp parallel
    {
        double* lpos = (double*)calloc(3 * (size_t)(n ? n : 1), sizeof(double));
        double* lint = (double*)calloc((size_t)(n ? n : 1), sizeof(double));
#pragma omp for schedule(dynamic, 4)
        for (int64_t yy = 0; yy < (int64_t)height; ++yy) {
            uint32_t y = (uint32_t)yy;
            for (uint32_t x = 0; x < width; ++x) {
                const or64_pixel* p = &px[(size_t)y * width + x];
                if (p->sky) continue;
                const float* b = tex_px(buf, pitch, x, y);
                double gl[3];
                if (is_target) {
                    double o[3];
                    d_pixel(p, g, lights, n, 0, NULL, o);
                    for (int c = 0; c < 3; ++c) gl[c] = 2.0 * (o[c] - (double)b[c]);
                } else {
                    for (int c = 0; c < 3; ++c) gl[c] = b[c];
                }
                double dpx[5] = {0, 0, 0, 0, 0};
                d_pixel_adjoint(p, g, lights, n, gl, dpx, lpos, lint);
                size_t idx = (size_t)y * width + x;
                for (int k = 0; k < 5; ++k) d_px_out[k * np + idx] = dpx[k];
            }
        }
#pragma omp critical
        {
            for (uint32_t i = 0; i < 3 * n; ++i) d_pos[i] += lpos[i];
            for (uint32_t i = 0; i < n; ++i) d_int[i] += lint[i];
        }
        free(lpos);
        free(lint);
    }
}

void or64_backward_l2(const or64_pixel* px, uint32_t width, uint32_t height,
                      const DeferredGlobals* g, const GpuLight* lights, uint32_t num_lights,
                      const float* target, uint32_t target_pitch, double* d_pixel,
                      double* d_pos, double* d_int) {
    d_backward(px, width, height, g, lights, num_lights, target, target_pitch, 1, d_pixel, d_pos,
               d_int);
}

void or64_backward(const or64_pixel* px, uint32_t width, uint32_t height,
                   const DeferredGlobals* g, const GpuLight* lights, uint32_t num_lights,
                   const float* dout, uint32_t dout_pitch, double* d_pixel, double* d_pos,
                   double* d_int) {
    d_backward(px, width, height, g, lights, num_lights, dout, dout_pitch, 0, d_pixel, d_pos,
               d_int);
}

static double d_pixel_loss(const or64_pixel* p, const DeferredGlobals* g, const GpuLight* lights,
                           uint32_t n, uint32_t mod_i, const GpuLight* mod, const float* tg,
                           uint64_t* hash) {
    double o[3];
    *hash = d_pixel(p, g, lights, n, mod_i, mod, o);
    double s = 0.0;
    for (int c = 0; c < 3; ++c) s += (o[c] - tg[c]) * (o[c] - tg[c]);
    return s;
}

double or64_fd(const or64_pixel* px, uint32_t width, uint32_t height, const DeferredGlobals* g,
               const GpuLight* lights, uint32_t num_lights, const float* target,
               uint32_t target_pitch, int kind, uint32_t index, double h, int* crosses) {
    *crosses = 0;
    if (kind < 5) {
        const or64_pixel* p0 = &px[index];
        if (p0->sky) return 0.0;
        or64_pixel pp = *p0, pm = *p0;
        double* fp = kind < 3 ? &pp.albedo[kind] : (kind == 3 ? &pp.rough : &pp.metal);
        double* fm = kind < 3 ? &pm.albedo[kind] : (kind == 3 ? &pm.rough : &pm.metal);
        *fp += h;
        *fm -= h;
        uint32_t x = index % width, y = index / width;
        const float* tg = tex_px(target, target_pitch, x, y);
        uint64_t hp, hm, h0;
        double lp = d_pixel_loss(&pp, g, lights, num_lights, 0, NULL, tg, &hp);
        double lm = d_pixel_loss(&pm, g, lights, num_lights, 0, NULL, tg, &hm);
        d_pixel_loss(p0, g, lights, num_lights, 0, NULL, tg, &h0);
        *crosses = (hp != hm) || (hp != h0);
        return (lp - lm) / (2.0 * h);
    }
    GpuLight lp_ = lights[index], lm_ = lights[index];
    /* FP64 perturbation of an FP32 record: the step is applied in double and
     * rounded once; the effective step is recomputed from the stored values. */
    double step;
    if (kind <= 7) {
        int c = kind - 5;
        lp_.position_range[c] = (float)((double)lights[index].position_range[c] + h);
        lm_.position_range[c] = (float)((double)lights[index].position_range[c] - h);
        step = (double)lp_.position_range[c] - (double)lm_.position_range[c];
    } else {
        lp_.color_intensity[3] = (float)((double)lights[index].color_intensity[3] + h);
        lm_.color_intensity[3] = (float)((double)lights[index].color_intensity[3] - h);
        step = (double)lp_.color_intensity[3] - (double)lm_.color_intensity[3];
    }
    const GpuLight* l = &lights[index];
    const double R = l->position_range[3] + 2.0 * h + 1e-3;
    double acc = 0.0;
    int cross = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : acc) reduction(| : cross)
    for (int64_t yy = 0; yy < (int64_t)height; ++yy) {
        uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < width; ++x) {
            const or64_pixel* p = &px[(size_t)y * width + x];
            if (p->sky) continue;
            if (l->light_type != 0u) {
                double dx = p->wp[0] - l->position_range[0], dy = p->wp[1] - l->position_range[1],
                       dz = p->wp[2] - l->position_range[2];
                if (dx * dx + dy * dy + dz * dz > R * R) continue;
            }
            const float* tg = tex_px(target, target_pitch, x, y);
            uint64_t hp, hm;
            double a = d_pixel_loss(p, g, lights, num_lights, index, &lp_, tg, &hp);
            double b = d_pixel_loss(p, g, lights, num_lights, index, &lm_, tg, &hm);
            acc += a - b;
            cross |= (hp != hm);
        }
    }
    *crosses = cross;
    return acc / step;
}
