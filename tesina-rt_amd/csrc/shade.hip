// Ray shading and framebuffer.
//
// Reference: RayShader::shade (src/rt/ray-shader.cpp:33-154) with shade_sample
// (src/kernel/ray-shader.cl:116-333), FrameBuffer clear/copy
// (src/rt/framebuffer.cpp:118-165; src/kernel/framebuffer.cl:21-54).
// Accumulation is integer (floor(c*1e6) per channel), so atomics make it
// exact and order-free; shade_secondary's run-head serial loop
// (src/kernel/ray-shader.cl:415-457) is not needed for that.
#include "devmath.hpp"
#include "rtx_impl.hpp"
#include "wavefront.hpp"

namespace {

// Images of the shading stage: the material textures, then (cube != 0) the
// cubemap's six faces +x, -x, +y, -y, +z, -z.
struct Images {
    const uchar4* texel;
    const TexEnt* tab;
    int count;
    int cube;
};

// UNORM_INT8 texel to float: c / 255.
__device__ __forceinline__ v3 texel(const Images& im, const TexEnt& e, int x, int y) {
    const uchar4 c = im.texel[e.off + (uint64_t)x + (uint64_t)y * (uint64_t)e.w];
    return mk3((float)c.x / 255.f, (float)c.y / 255.f, (float)c.z / 255.f);
}

// read_imagef with normalized coordinates and CLK_FILTER_LINEAR
// (src/kernel/ray-shader.cl:145-153), REPEAT or CLAMP_TO_EDGE, by the
// OpenCL 1.2 filter formulas in float32 (the weights' precision is left to
// the implementation there; this is this library's definition):
//   repeat: u = (s - floor(s)) * w, i0 = floor(u - 0.5), i1 = i0 + 1, wrapped;
//   clamp:  u = s * w, i0 and i1 clamped to [0, w);  a = frac(u - 0.5);
//   T = (1-a)(1-b) T00 + a(1-b) T10 + (1-a)b T01 + ab T11, in that order.
// A non-finite coordinate samples as 0.
__device__ v3 sample_linear(const Images& im, int entry, float s, float t, bool repeat) {
    const TexEnt e = im.tab[entry];
    const int w = e.w, h = e.h;
    if (!__builtin_isfinite(s)) s = 0.f;
    if (!__builtin_isfinite(t)) t = 0.f;
    float u, v;
    if (repeat) {
        u = (s - floorf(s)) * (float)w;
        v = (t - floorf(t)) * (float)h;
    } else {
        u = s * (float)w;
        v = t * (float)h;
    }
    const float fu = floorf(u - 0.5f), fv = floorf(v - 0.5f);
    const float a = (u - 0.5f) - fu, b = (v - 0.5f) - fv;
    int i0, i1, j0, j1;
    if (repeat) {
        i0 = (int)fu;
        j0 = (int)fv;
        i1 = i0 + 1;
        j1 = j0 + 1;
        if (i0 < 0) i0 += w;
        if (j0 < 0) j0 += h;
        if (i1 > w - 1) i1 -= w;
        if (j1 > h - 1) j1 -= h;
    } else {
        const int iu = (int)fminf(fmaxf(fu, -1.f), (float)w);
        const int iv = (int)fminf(fmaxf(fv, -1.f), (float)h);
        i0 = min(max(iu, 0), w - 1);
        i1 = min(max(iu + 1, 0), w - 1);
        j0 = min(max(iv, 0), h - 1);
        j1 = min(max(iv + 1, 0), h - 1);
    }
    const float w00 = (1.f - a) * (1.f - b), w10 = a * (1.f - b), w01 = (1.f - a) * b, w11 = a * b;
    return ((w00 * texel(im, e, i0, j0) + w10 * texel(im, e, i1, j0)) + w01 * texel(im, e, i0, j1)) +
           w11 * texel(im, e, i1, j1);
}

__device__ __forceinline__ bool in_range(float f1, float f2) {
    return f1 <= 1.f && f1 >= -1.f && f2 <= 1.f && f2 >= -1.f;
}

// The miss branch of shade_sample (src/kernel/ray-shader.cl:275-307): the
// first face, in the order +x, -x, +y, -y, +z, -z, whose projection of the
// direction lies in [-1, 1]^2; none (a NaN direction) gives black.
__device__ v3 cubemap_rgb(const Images& im, v3 d) {
    const v3 inv = recip(d);
    const v3 xp = d * inv.x, yp = d * inv.y, zp = d * inv.z;
    const int c = im.count;
    if (d.x > 0.f && in_range(xp.y, xp.z)) return sample_linear(im, c + 0, (-xp.z + 1.f) * 0.5f, (xp.y + 1.f) * 0.5f, false);
    if (d.x < 0.f && in_range(xp.y, xp.z)) return sample_linear(im, c + 1, (-xp.z + 1.f) * 0.5f, (-xp.y + 1.f) * 0.5f, false);
    if (d.y > 0.f && in_range(yp.x, yp.z)) return sample_linear(im, c + 2, (yp.x + 1.f) * 0.5f, (-yp.z + 1.f) * 0.5f, false);
    if (d.y < 0.f && in_range(yp.x, yp.z)) return sample_linear(im, c + 3, (-yp.x + 1.f) * 0.5f, (-yp.z + 1.f) * 0.5f, false);
    if (d.z > 0.f && in_range(zp.x, zp.y)) return sample_linear(im, c + 4, (zp.x + 1.f) * 0.5f, (zp.y + 1.f) * 0.5f, false);
    if (d.z < 0.f && in_range(zp.x, zp.y)) return sample_linear(im, c + 5, (zp.x + 1.f) * 0.5f, (-zp.y + 1.f) * 0.5f, false);
    return mk3(0.f, 0.f, 0.f);
}

// unique: no other ray of this launch adds to this pixel (the primary rays
// of a frame, one per pixel), so a plain read-add-write gives what the
// atomic gives (the reference's shade_primary adds with +=,
// src/kernel/ray-shader.cl:318-330) without a device-scope atomic, which
// on this chip goes past the XCD's L2 to memory.
__device__ __forceinline__ void add_pixel(uint32_t* __restrict__ accum, uint32_t px, const float ch[3], bool unique) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        uint32_t u = (uint32_t)floorf(clamp01(ch[q]) * 1e6f);
        if (!u) continue;
        if (unique)
            accum[3 * (size_t)px + q] += u;
        else
            atomicAdd(&accum[3 * (size_t)px + q], u);
    }
}

// shade_sample for ray r: its colour times its contribution added to its
// pixel's accumulators.
__device__ __forceinline__ void shade_ray(uint64_t r, const float4* __restrict__ dir4, const int32_t* __restrict__ pixel,
                                          const float* __restrict__ contrib, const uint32_t* __restrict__ flags,
                                          const int32_t* __restrict__ hid, const float4* __restrict__ hn,
                                          const float4* __restrict__ hp, const float2* __restrict__ huv,
                                          const rtx_material* __restrict__ mats, const int32_t* __restrict__ matmap,
                                          const rtx_lights* __restrict__ lights, int nl, v3 amb, const Images& im,
                                          uint32_t* __restrict__ accum, bool unique) {
    uint32_t fl = flags[r];
    if (!(fl & 1u)) {  // a miss adds the cubemap's colour, or nothing without one
        if (!im.cube) return;
        const v3 rgb = cubemap_rgb(im, xyz(dir4[r]));
        const float c = contrib[r];
        const float ch[3] = {c * rgb.x, c * rgb.y, c * rgb.z};
        add_pixel(accum, (uint32_t)pixel[r], ch, unique);
        return;
    }
    const rtx_material m = mats[matmap[hid[r]]];
    v3 diffuse = mk3(m.diffuse[0], m.diffuse[1], m.diffuse[2]);
    if (m.texture >= 0) {
        const float2 uv = huv[r];
        diffuse = m.texture < im.count ? sample_linear(im, m.texture, uv.x, uv.y, true) : mk3(0.f, 0.f, 0.f);
    }
    v3 d = xyz(dir4[r]);
    v3 nn = xyz(hn[r]);
    v3 p = xyz(hp[r]);
    v3 val = amb * diffuse;
    for (int l = 0; l < nl; ++l) {
        const rtx_light& L = lights->l[l];
        v3 Ld;
        float LD;
        if (L.type == RTX_LIGHT_DIR) {
            Ld = mk3(L.dir[0], L.dir[1], L.dir[2]);
            LD = 1.f;
        } else if (L.type == RTX_LIGHT_SPOT) {
            Ld = normalize(p - mk3(L.pos[0], L.pos[1], L.pos[2]));
            LD = dot(Ld, mk3(L.dir[0], L.dir[1], L.dir[2])) - L.angle;
            LD /= 1.f - L.angle;
            LD = clamp01(LD);
        } else {
            Ld = normalize(p - mk3(L.pos[0], L.pos[1], L.pos[2]));
            LD = 1.f;
        }
        v3 rgb = mk3(L.color[0], L.color[1], L.color[2]) * LD;
        if (fl & (1u << (8 + l))) continue;
        v3 rr = d - 2.f * nn * dot(d, nn);
        float cosL = clamp01(dot(nn, -Ld));
        cosL = cosL * cosL;
        float spec = clamp01(dot(rr, -Ld));
        float s2 = spec * spec, s4 = s2 * s2;
        spec = s4 * s4;
        spec *= cosL * m.shininess;
        val = val + rgb * diffuse * cosL;
        val = val + rgb * spec;
    }
    val = val * (1.f - m.reflectiveness);
    const float c = contrib[r];
    const float ch[3] = {val.x * c, val.y * c, val.z * c};
    add_pixel(accum, (uint32_t)pixel[r], ch, unique);
}

__global__ __launch_bounds__(256) void k_shade(const float4* __restrict__ dir4, const int32_t* __restrict__ pixel,
                                               const float* __restrict__ contrib, const uint32_t* __restrict__ flags,
                                               const int32_t* __restrict__ hid, const float4* __restrict__ hn,
                                               const float4* __restrict__ hp, const float2* __restrict__ huv,
                                               const rtx_material* __restrict__ mats,
                                               const int32_t* __restrict__ matmap,
                                               const rtx_lights* __restrict__ lights, const Images im,
                                               const uint32_t* dcount, uint64_t hcount,
                                               uint32_t* __restrict__ accum, int unique) {
    const uint64_t n = hcount == RTX_DEVICE_COUNT ? (uint64_t)*dcount : hcount;
    const int nl = lights->count;
    const v3 amb = mk3(lights->ambient[0], lights->ambient[1], lights->ambient[2]);
    for (uint64_t r = blockIdx.x * 256ull + threadIdx.x; r < n; r += (uint64_t)gridDim.x * 256)
        shade_ray(r, dir4, pixel, contrib, flags, hid, hn, hp, huv, mats, matmap, lights, nl, amb, im, accum,
                  unique != 0);
}

__global__ void k_resolve(const uint32_t* __restrict__ accum, uchar4* __restrict__ rgba, uint32_t P) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float inv = 1.f / 1e6f;
    uint8_t c[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        float f = (float)accum[3 * (size_t)i + q] * inv;
        f = clamp01(f) * 255.f;
        c[q] = (uint8_t)rintf(f);
    }
    rgba[i] = make_uchar4(c[0], c[1], c[2], 255);
}

// Shading of every frame's rays of this bounce: blocks take 256-ray chunks
// in a fixed stride.
__global__ __launch_bounds__(256) void k_wf_shade(const WfSeg* __restrict__ segs, uint32_t nseg) {
    __shared__ uint32_t s_end[RTX_WF_MAX_SEGS];
    const uint32_t total = wf_chunks<WF_RAYS>(segs, nseg, 256u, s_end);
    for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
        const uint32_t sg = wf_find(s_end, nseg, c);
        const WfSeg& S = segs[sg];
        const uint64_t r = (uint64_t)(c - wf_first(s_end, sg)) * 256u + threadIdx.x;
        if (r >= (uint64_t)S.count[0]) continue;
        const Images im{S.tex, S.tex_tab, S.tex ? S.ntex : 0, S.tex && S.cube ? 1 : 0};
        const v3 amb = mk3(S.lights->ambient[0], S.lights->ambient[1], S.lights->ambient[2]);
        // a frame's primary rays (one per pixel) are shaded in the step that
        // generated them
        shade_ray(r, S.dir4, S.pixel, S.contrib, S.flags, S.id, S.n4, S.p4, S.uv2, S.mats, S.matmap, S.lights,
                  S.lights->count, amb, im, S.accum, S.primary != 0);
    }
}

// RGBA8 of the frames whose last bounce this step shaded.
__global__ __launch_bounds__(256) void k_wf_resolve(const WfSeg* __restrict__ segs, uint32_t nseg,
                                                    unsigned long long* __restrict__ fin_out) {
    __shared__ uint32_t s_end[RTX_WF_MAX_SEGS];
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < nseg; i += 256)
            if (segs[i].resolve)
                for (int q = 0; q < 3; ++q) fin_out[4 * (size_t)segs[i].fin + q] = segs[i].stats[q];
    const uint32_t total = wf_chunks<WF_PIXELS_DONE>(segs, nseg, 256u, s_end);
    const float inv = 1.f / 1e6f;
    for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
        const uint32_t sg = wf_find(s_end, nseg, c);
        const WfSeg& S = segs[sg];
        const uint32_t i = (c - wf_first(s_end, sg)) * 256u + threadIdx.x;
        if (i >= (uint32_t)S.w * (uint32_t)S.h) continue;
        uint8_t ch[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            float f = (float)S.accum[3 * (size_t)i + q] * inv;
            f = clamp01(f) * 255.f;
            ch[q] = (uint8_t)rintf(f);
        }
        S.rgba[i] = make_uchar4(ch[0], ch[1], ch[2], 255);
    }
}

}  // namespace

int32_t rtx_fb_create(rtx_ctx* ctx, int32_t width, int32_t height, rtx_fb** out) {
    RTX_REQUIRE(ctx && out && width > 0 && height > 0, RTX_EINVAL, "rtx_fb_create: bad argument");
    rtx_fb* f = new rtx_fb();
    f->ctx = ctx;
    f->w = width;
    f->h = height;
    size_t P = (size_t)width * height;
    int32_t e = 0;
    if ((e = f->accum.reserve(ctx, P * 12)) || (e = f->rgba.reserve(ctx, P * 4))) {
        rtx_fb_destroy(f);
        return e;
    }
    *out = f;
    return rtx_fb_clear(f, nullptr);
}

int32_t rtx_fb_destroy(rtx_fb* fb) {
    RTX_REQUIRE(fb, RTX_EINVAL, "rtx_fb_destroy: null");
    (void)hipStreamSynchronize(fb->ctx->stream);
    fb->accum.release();
    fb->rgba.release();
    delete fb;
    return RTX_OK;
}

int32_t rtx_fb_clear(rtx_fb* fb, void* stream) {
    RTX_REQUIRE(fb, RTX_EINVAL, "rtx_fb_clear: null");
    return rtx_fill_u32(fb->ctx, fb->accum.p, 0u, 3ull * fb->w * fb->h, fb->ctx->pick(stream));
}

int32_t rtx_fb_resolve(rtx_fb* fb, void* stream) {
    RTX_REQUIRE(fb, RTX_EINVAL, "rtx_fb_resolve: null");
    hipStream_t st = fb->ctx->pick(stream);
    uint32_t P = (uint32_t)fb->w * fb->h;
    ProfScope ps(fb->ctx, st, "fb_resolve");
    hipLaunchKernelGGL(k_resolve, dim3(rtx_div_up(P, 256)), dim3(256), 0, st, fb->accum.as<uint32_t>(),
                       fb->rgba.as<uchar4>(), P);
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

int32_t rtx_fb_read_accum(rtx_fb* fb, uint32_t* rgb3) {
    RTX_REQUIRE(fb && rgb3, RTX_EINVAL, "rtx_fb_read_accum: null");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    RTX_CHECK_HIP(hipMemcpy(rgb3, fb->accum.p, (size_t)fb->w * fb->h * 12, hipMemcpyDeviceToHost));
    return RTX_OK;
}

int32_t rtx_fb_read_rgba8(rtx_fb* fb, uint8_t* rgba) {
    RTX_REQUIRE(fb && rgba, RTX_EINVAL, "rtx_fb_read_rgba8: null");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    RTX_CHECK_HIP(hipMemcpy(rgba, fb->rgba.p, (size_t)fb->w * fb->h * 4, hipMemcpyDeviceToHost));
    return RTX_OK;
}

int32_t rtx_shade(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, rtx_fb* fb, uint64_t count, int32_t primary,
                  void* stream) {
    // the caller's primary rays may hold several samples per pixel
    return rtx_shade_impl(s, rays, hits, fb, count, primary, false, stream);
}

// unique_pixels: every ray of the bundle has its own pixel (a renderer's
// primary rays, one sample per pixel).
int32_t rtx_shade_impl(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, rtx_fb* fb, uint64_t count, int32_t primary,
                       bool unique_pixels, void* stream) {
    (void)primary;
    RTX_REQUIRE(s && rays && hits && fb, RTX_EINVAL, "rtx_shade: null argument");
    RTX_REQUIRE(s->mats.p && s->lights.p, RTX_ESTATE, "rtx_shade: scene has no materials or lights");
    RTX_REQUIRE(count == RTX_DEVICE_COUNT || (count <= rays->cap && count <= hits->cap), RTX_EINVAL,
                "rtx_shade: count exceeds bundle capacity");
    hipStream_t st = s->ctx->pick(stream);
    if (count == 0) return RTX_OK;
    uint64_t n = count == RTX_DEVICE_COUNT ? rays->cap : count;
    uint64_t g = (n + 255) / 256;
    if (g > (uint64_t)s->ctx->cus * 32) g = (uint64_t)s->ctx->cus * 32;
    ProfScope ps(s->ctx, st, "shade");
    hipLaunchKernelGGL(k_shade, dim3((uint32_t)g), dim3(256), 0, st, rays->dir4.as<float4>(), rays->pixel.as<int32_t>(),
                       rays->contrib.as<float>(), hits->flags.as<uint32_t>(), hits->id.as<int32_t>(),
                       hits->n4.as<float4>(), hits->p4.as<float4>(), hits->uv2.as<float2>(),
                       s->mats.as<rtx_material>(), scene_view(s).matmap, s->lights.as<rtx_lights>(),
                       Images{s->tex.as<uchar4>(), s->tex_tab.as<TexEnt>(), s->tex.p ? s->ntex : 0,
                              s->tex.p && s->cube ? 1 : 0},
                       rays->count.as<uint32_t>(), count, fb->accum.as<uint32_t>(), unique_pixels ? 1 : 0);
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

// ------------------------------------------------- wavefront launchers --
void rtx_wf_shade(const WfSeg* segs, uint32_t nseg, int cus, hipStream_t st) {
    hipLaunchKernelGGL(k_wf_shade, dim3((uint32_t)cus * 16), dim3(256), 0, st, segs, nseg);
}

void rtx_wf_resolve(const WfSeg* segs, uint32_t nseg, uint64_t pixels, unsigned long long* fin_out, int cus,
                    hipStream_t st) {
    uint64_t g = (pixels + 255) / 256;
    if (g > (uint64_t)cus * 16) g = (uint64_t)cus * 16;
    hipLaunchKernelGGL(k_wf_resolve, dim3((uint32_t)(g ? g : 1)), dim3(256), 0, st, segs, nseg, fin_out);
}
