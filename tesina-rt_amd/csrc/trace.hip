// Ray bundles, primary ray generation, closest-hit and shadow traversal.
//
// Reference: PrimaryRayGenerator (src/rt/primary-ray-generator.cpp:58-165,
// src/kernel/primary-ray-generator.cl:28-103), Tracer::trace / shadow_trace
// (src/rt/tracer.cpp:140-352; src/kernel/trace-bvh.cl:190-481,
// src/kernel/shadow-trace-bvh.cl:160-418).
//
// Traversal keeps the reference's visiting order exactly (test a node's box
// when it is visited, near child first by dot(dir, right.lo - left.lo) > 0,
// pop only while 0 < level < 32, tMax shrinks to the best hit) so that hit
// ids, distances and tie winners are those of the oracle.  The MI355X layout
// is what differs: 32-byte nodes plus a 16-byte delta fetched only at
// internal nodes, triangles stored contiguously in leaf order, the stack in
// LDS ([level][thread], conflict-free), SoA rays and hits.
#include "devmath.hpp"
#include "rtx_impl.hpp"

namespace {

constexpr int TB = 128;        // traversal block: 2 waves, 16 KB LDS stack
constexpr int STACK = 32;      // MAX_LEVELS of src/kernel/trace-bvh.cl:330

struct RaySoA {
    const float4* ori4;
    const float4* dir4;
};

__device__ __forceinline__ uint64_t ray_count(const uint32_t* dcount, uint64_t hcount) {
    return hcount == RTX_DEVICE_COUNT ? (uint64_t)*dcount : hcount;
}

// bbox_hit (src/kernel/trace-bvh.cl:247-273).
__device__ __forceinline__ bool slab(v3 lo, v3 hi, v3 ori, v3 inv, float tmin, float tmax) {
    v3 tl = (lo - ori) * inv;
    v3 th = (hi - ori) * inv;
    v3 tM = vmax(tl, th);
    v3 tm = vmin(tl, th);
    float t0 = tmin, t1 = tmax;
    if (fabsf(inv.x) > 1e-6f) { t0 = fmaxf(t0, tm.x); t1 = fminf(t1, tM.x); }
    if (fabsf(inv.y) > 1e-6f) { t0 = fmaxf(t0, tm.y); t1 = fminf(t1, tM.y); }
    if (fabsf(inv.z) > 1e-6f) { t0 = fmaxf(t0, tm.z); t1 = fminf(t1, tM.z); }
    return t0 <= t1;
}

// Möller–Trumbore of leaf_hit (src/kernel/trace-bvh.cl:282-327).
__device__ __forceinline__ bool tri_hit(const float4* __restrict__ tri, uint32_t t, v3 o, v3 d, float* tt, float* uu,
                                        float* vv) {
    v3 v0 = xyz(tri[3 * (size_t)t]);
    v3 v1 = xyz(tri[3 * (size_t)t + 1]);
    v3 v2 = xyz(tri[3 * (size_t)t + 2]);
    v3 e1 = v1 - v0, e2 = v2 - v0;
    v3 h = cross(d, e2);
    float a = dot(e1, h);
    if (a > -1e-26f && a < 1e-26f) return false;
    float f = 1.f / a;
    v3 s = o - v0;
    float u = f * dot(s, h);
    if (u < 0.f || u > 1.f) return false;
    v3 q = cross(s, e1);
    float v = f * dot(d, q);
    if (v < 0.f || u + v > 1.f) return false;
    *tt = f * dot(e2, q);
    *uu = u;
    *vv = v;
    return true;
}

// ------------------------------------------------------------- raygen ----
__global__ __launch_bounds__(256) void k_raygen_primary(float4* __restrict__ ori4, float4* __restrict__ dir4,
                                                        int32_t* __restrict__ pixel, float* __restrict__ contrib,
                                                        uint32_t* dcount, rtx_camera cam, int width, int height,
                                                        int quad, const float* __restrict__ ps, uint64_t offset,
                                                        uint64_t count) {
    uint64_t k = blockIdx.x * 256ull + threadIdx.x;
    if (k == 0) *dcount = (uint32_t)count;
    if (k >= count) return;
    int gid = (int)(offset + k);
    int pixels = width * height;
    int i = gid % pixels;
    const float* smp = ps + 3 * (gid / pixels);
    int bs = quad, ba = quad * quad;
    int xb = width / bs, yb = height / bs;
    int block_rays = xb * yb * ba;
    int right_rays = (width - xb * bs) * height;
    int x, y;
    if (i < block_rays) {
        int block = i / ba;
        int y_block = block / xb, x_block = block % xb;
        int bp = i % ba;
        x = x_block * bs + bp % bs;
        y = y_block * bs + bp / bs;
    } else if (i < block_rays + right_rays) {
        int sx = xb * bs, slab_w = width - sx, ri = i - block_rays;
        x = sx + ri % slab_w;
        y = ri / slab_w;
    } else {
        int sy = yb * bs, slab_h = height - sy, ti = i - (block_rays + right_rays);
        x = ti / slab_h;
        y = sy + ti % slab_h;
    }
    float xn = (0.5f + (float)x + smp[0]) / (float)width;
    float yn = (0.5f + (float)y + smp[1]) / (float)height;
    v3 dir = mk3(cam.dir[0], cam.dir[1], cam.dir[2]);
    v3 right = mk3(cam.right[0], cam.right[1], cam.right[2]);
    v3 up = mk3(cam.up[0], cam.up[1], cam.up[2]);
    v3 d = normalize(dir + right * (xn * 2.f - 1.f) + up * (yn * 2.f - 1.f));
    ori4[k] = make_float4(cam.pos[0], cam.pos[1], cam.pos[2], 0.f);
    dir4[k] = make_float4(d.x, d.y, d.z, 1e36f);
    pixel[k] = width * y + x;
    contrib[k] = smp[2];
}

// -------------------------------------------------------------- trace ----
__global__ __launch_bounds__(TB) void k_trace(const float4* __restrict__ node, const float4* __restrict__ delta,
                                              const float4* __restrict__ tri, const float4* __restrict__ nrm,
                                              const float2* __restrict__ uvs, const float4* __restrict__ ori4,
                                              const float4* __restrict__ dir4, const uint32_t* __restrict__ dcount,
                                              uint64_t hcount, uint32_t* __restrict__ flags, float* __restrict__ tout,
                                              int32_t* __restrict__ idout, float2* __restrict__ bary,
                                              float4* __restrict__ n4, float4* __restrict__ p4,
                                              float2* __restrict__ uv2) {
    __shared__ uint32_t stack[STACK][TB];
    const uint64_t n = ray_count(dcount, hcount);
    const uint32_t tid = threadIdx.x;
    for (uint64_t r = blockIdx.x * (uint64_t)TB + tid; r < n; r += (uint64_t)gridDim.x * TB) {
        float4 o = ori4[r], dd = dir4[r];
        v3 ori = xyz(o), dir = xyz(dd);
        v3 inv = recip(dir);
        const float tmin = o.w;
        float best_t = dd.w, bu = 0.f, bv = 0.f;
        int32_t best = -1;
        uint32_t level = 0, curr = 0;
        for (;;) {
            float4 f0 = node[2 * (size_t)curr], f1 = node[2 * (size_t)curr + 1];
            uint32_t a = __float_as_uint(f0.w), cnt = __float_as_uint(f1.w);
            bool pop = true;
            if (slab(xyz(f0), xyz(f1), ori, inv, tmin, best_t)) {
                if (cnt) {
                    for (uint32_t t = a; t < a + cnt; ++t) {
                        float tt, u, v;
                        if (!tri_hit(tri, t, ori, dir, &tt, &u, &v)) continue;
                        if (tt <= best_t && tt >= tmin) {
                            best_t = tt;
                            best = (int32_t)t;
                            bu = u;
                            bv = v;
                        }
                    }
                } else {
                    v3 dl = xyz(delta[curr]);
                    bool near_left = dot(dir, dl) > 0.f;
                    stack[level & (STACK - 1)][tid] = near_left ? a + 1 : a;
                    level++;
                    curr = near_left ? a : a + 1;
                    pop = false;
                }
            }
            if (pop) {
                if (level > 0 && level < STACK) {
                    level--;
                    curr = stack[level][tid];
                } else {
                    break;
                }
            }
        }
        // complete_trace_info (src/kernel/trace-bvh.cl:190-245)
        if (best < 0) {
            flags[r] = 0u;
            tout[r] = 0.f;
            idout[r] = -1;
            bary[r] = make_float2(0.f, 0.f);
            n4[r] = make_float4(0.f, 0.f, 0.f, 0.f);
            p4[r] = make_float4(0.f, 0.f, 0.f, 0.f);
            uv2[r] = make_float2(0.f, 0.f);
            continue;
        }
        v3 hp = ori + dir * best_t;
        uint32_t i0 = __float_as_uint(tri[3 * (size_t)best].w);
        uint32_t i1 = __float_as_uint(tri[3 * (size_t)best + 1].w);
        uint32_t i2 = __float_as_uint(tri[3 * (size_t)best + 2].w);
        float w = 1.f - (bu + bv);
        v3 n0 = xyz(nrm[i0]), n1 = xyz(nrm[i1]), n2 = xyz(nrm[i2]);
        v3 nn = normalize(w * n0 + bv * n2 + bu * n1);
        uint32_t fl = 1u;
        if (dot(nn, dir) > 0.f) {
            fl |= 2u;
            nn = -nn;
        }
        float2 t0 = uvs[i0], t1 = uvs[i1], t2 = uvs[i2];
        flags[r] = fl;
        tout[r] = best_t;
        idout[r] = best;
        bary[r] = make_float2(bu, bv);
        n4[r] = make_float4(nn.x, nn.y, nn.z, 0.f);
        p4[r] = make_float4(hp.x, hp.y, hp.z, 0.f);
        uv2[r] = make_float2(w * t0.x + bv * t2.x + bu * t1.x, w * t0.y + bv * t2.y + bu * t1.y);
    }
}

// ------------------------------------------------------------- shadow ----
// shadow_trace_single (src/kernel/shadow-trace-bvh.cl:375-418) for every
// light: DIR toward -dir, SPOT toward pos (outside the cone = shadowed),
// POINT toward pos with tMax = distance; any hit with tMin < t < tMax.
__global__ __launch_bounds__(TB) void k_shadow(const float4* __restrict__ node, const float4* __restrict__ delta,
                                               const float4* __restrict__ tri, const rtx_lights* __restrict__ lights,
                                               const uint32_t* __restrict__ dcount, uint64_t hcount,
                                               uint32_t* __restrict__ flags, const float4* __restrict__ p4,
                                               unsigned long long* shadow_count) {
    __shared__ uint32_t stack[STACK][TB];
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint64_t n = ray_count(dcount, hcount);
    const uint32_t tid = threadIdx.x;
    const int nl = lights->count;
    uint32_t my_shadow_rays = 0;
    for (uint64_t r = blockIdx.x * (uint64_t)TB + tid; r < n; r += (uint64_t)gridDim.x * TB) {
        uint32_t fl = flags[r] & 0xffu;
        if (!(fl & 1u)) {
            flags[r] = fl;
            continue;
        }
        v3 p = xyz(p4[r]);
        for (int l = 0; l < nl; ++l) {
            const rtx_light& L = lights->l[l];
            v3 dir;
            float tmax = 1e37f;
            bool shadowed = false, trace = true;
            if (L.type == RTX_LIGHT_DIR) {
                dir = -mk3(L.dir[0], L.dir[1], L.dir[2]);
            } else if (L.type == RTX_LIGHT_SPOT) {
                dir = normalize(mk3(L.pos[0], L.pos[1], L.pos[2]) - p);
                if (dot(-dir, mk3(L.dir[0], L.dir[1], L.dir[2])) < L.angle) {
                    shadowed = true;
                    trace = false;
                }
            } else if (L.type == RTX_LIGHT_POINT) {
                v3 toL = mk3(L.pos[0], L.pos[1], L.pos[2]) - p;
                tmax = __builtin_sqrtf(dot(toL, toL));
                dir = normalize(toL);
            } else {
                shadowed = true;
                trace = false;
            }
            if (trace) {
                my_shadow_rays++;
                v3 inv = recip(dir);
                const float tmin = 0.01f;
                uint32_t level = 0, curr = 0;
                for (;;) {
                    float4 f0 = node[2 * (size_t)curr], f1 = node[2 * (size_t)curr + 1];
                    uint32_t a = __float_as_uint(f0.w), cnt = __float_as_uint(f1.w);
                    bool pop = true;
                    if (slab(xyz(f0), xyz(f1), p, inv, tmin, tmax)) {
                        if (cnt) {
                            for (uint32_t t = a; t < a + cnt; ++t) {
                                float tt, u, v;
                                if (tri_hit(tri, t, p, dir, &tt, &u, &v) && tt < tmax && tt > tmin) {
                                    shadowed = true;
                                    break;
                                }
                            }
                            if (shadowed) break;
                        } else {
                            v3 dl = xyz(delta[curr]);
                            bool near_left = dot(dir, dl) > 0.f;
                            stack[level & (STACK - 1)][tid] = near_left ? a + 1 : a;
                            level++;
                            curr = near_left ? a : a + 1;
                            pop = false;
                        }
                    }
                    if (pop) {
                        if (level > 0 && level < STACK) {
                            level--;
                            curr = stack[level][tid];
                        } else {
                            break;
                        }
                    }
                }
            }
            if (shadowed) fl |= 1u << (8 + l);
        }
        flags[r] = fl;
    }
    if (shadow_count) {
        atomicAdd(&s_cnt, my_shadow_rays);
        __syncthreads();
        if (threadIdx.x == 0 && s_cnt) atomicAdd(shadow_count, (unsigned long long)s_cnt);
    }
}

uint32_t trace_grid(rtx_ctx* c, uint64_t count, uint64_t cap) {
    uint64_t n = count == RTX_DEVICE_COUNT ? cap : count;
    uint64_t g = (n + TB - 1) / TB;
    uint64_t gmax = (uint64_t)c->cus * 64;  // grid-stride beyond 64 blocks per CU
    if (g > gmax) g = gmax;
    return (uint32_t)(g ? g : 1);
}

}  // namespace

// ----------------------------------------------------------- bundles -----
int32_t rtx_rays_create(rtx_ctx* ctx, uint64_t capacity, rtx_rays** out) {
    RTX_REQUIRE(ctx && out && capacity, RTX_EINVAL, "rtx_rays_create: null argument or capacity 0");
    RTX_REQUIRE(capacity < (1ull << 30), RTX_EUNSUPPORTED, "rtx_rays_create: capacity >= 2^30");
    rtx_rays* r = new rtx_rays();
    r->ctx = ctx;
    r->cap = capacity;
    int32_t e = 0;
    if ((e = r->ori4.reserve(ctx, capacity * 16)) || (e = r->dir4.reserve(ctx, capacity * 16)) ||
        (e = r->pixel.reserve(ctx, capacity * 4)) || (e = r->contrib.reserve(ctx, capacity * 4)) ||
        (e = r->count.reserve(ctx, 16)) || (e = r->scan_status.reserve(ctx, ((capacity + 255) / 256 + 1) * 4))) {
        rtx_rays_destroy(r);
        return e;
    }
    uint32_t init[4] = {0, 0, 0xffffffffu, 0};
    if (hipMemcpy(r->count.p, init, 16, hipMemcpyHostToDevice) != hipSuccess) {
        rtx_rays_destroy(r);
        rtx_set_error("rtx_rays_create: count init failed");
        return RTX_EDEVICE;
    }
    *out = r;
    return RTX_OK;
}

int32_t rtx_rays_destroy(rtx_rays* r) {
    RTX_REQUIRE(r, RTX_EINVAL, "rtx_rays_destroy: null");
    (void)hipStreamSynchronize(r->ctx->stream);
    r->ori4.release();
    r->dir4.release();
    r->pixel.release();
    r->contrib.release();
    r->count.release();
    r->scan_status.release();
    r->psamp.release();
    delete r;
    return RTX_OK;
}

uint64_t rtx_rays_capacity(rtx_rays* r) { return r ? r->cap : 0; }

int32_t rtx_rays_count(rtx_rays* r, uint64_t* count) {
    RTX_REQUIRE(r && count, RTX_EINVAL, "rtx_rays_count: null argument");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    uint32_t c = 0;
    RTX_CHECK_HIP(hipMemcpy(&c, r->count.p, 4, hipMemcpyDeviceToHost));
    *count = c;
    return RTX_OK;
}

int32_t rtx_rays_read(rtx_rays* r, float* ori4, float* dir4, int32_t* pixel, float* contrib, uint64_t n) {
    RTX_REQUIRE(r && n <= r->cap, RTX_EINVAL, "rtx_rays_read: bad argument");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    if (ori4) RTX_CHECK_HIP(hipMemcpy(ori4, r->ori4.p, n * 16, hipMemcpyDeviceToHost));
    if (dir4) RTX_CHECK_HIP(hipMemcpy(dir4, r->dir4.p, n * 16, hipMemcpyDeviceToHost));
    if (pixel) RTX_CHECK_HIP(hipMemcpy(pixel, r->pixel.p, n * 4, hipMemcpyDeviceToHost));
    if (contrib) RTX_CHECK_HIP(hipMemcpy(contrib, r->contrib.p, n * 4, hipMemcpyDeviceToHost));
    return RTX_OK;
}

int32_t rtx_rays_write(rtx_rays* r, const float* ori4, const float* dir4, const int32_t* pixel, const float* contrib,
                       uint64_t n) {
    RTX_REQUIRE(r && ori4 && dir4 && n <= r->cap, RTX_EINVAL, "rtx_rays_write: bad argument");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    RTX_CHECK_HIP(hipMemcpy(r->ori4.p, ori4, n * 16, hipMemcpyHostToDevice));
    RTX_CHECK_HIP(hipMemcpy(r->dir4.p, dir4, n * 16, hipMemcpyHostToDevice));
    std::vector<int32_t> px(n, 0);
    std::vector<float> cb(n, 1.f);
    RTX_CHECK_HIP(hipMemcpy(r->pixel.p, pixel ? pixel : px.data(), n * 4, hipMemcpyHostToDevice));
    RTX_CHECK_HIP(hipMemcpy(r->contrib.p, contrib ? contrib : cb.data(), n * 4, hipMemcpyHostToDevice));
    uint32_t c = (uint32_t)n;
    RTX_CHECK_HIP(hipMemcpy(r->count.p, &c, 4, hipMemcpyHostToDevice));
    return RTX_OK;
}

int32_t rtx_hits_create(rtx_ctx* ctx, uint64_t capacity, rtx_hits** out) {
    RTX_REQUIRE(ctx && out && capacity, RTX_EINVAL, "rtx_hits_create: null argument or capacity 0");
    rtx_hits* h = new rtx_hits();
    h->ctx = ctx;
    h->cap = capacity;
    int32_t e = 0;
    if ((e = h->flags.reserve(ctx, capacity * 4)) || (e = h->t.reserve(ctx, capacity * 4)) ||
        (e = h->id.reserve(ctx, capacity * 4)) || (e = h->bary.reserve(ctx, capacity * 8)) ||
        (e = h->n4.reserve(ctx, capacity * 16)) || (e = h->p4.reserve(ctx, capacity * 16)) ||
        (e = h->uv2.reserve(ctx, capacity * 8))) {
        rtx_hits_destroy(h);
        return e;
    }
    *out = h;
    return RTX_OK;
}

int32_t rtx_hits_destroy(rtx_hits* h) {
    RTX_REQUIRE(h, RTX_EINVAL, "rtx_hits_destroy: null");
    (void)hipStreamSynchronize(h->ctx->stream);
    DevBuf* all[] = {&h->flags, &h->t, &h->id, &h->bary, &h->n4, &h->p4, &h->uv2};
    for (DevBuf* b : all) b->release();
    delete h;
    return RTX_OK;
}

int32_t rtx_hits_read(rtx_hits* h, uint32_t* flags, float* t, int32_t* id, float* bary2, float* n4, float* p4,
                      float* uv2, uint64_t n) {
    RTX_REQUIRE(h && n <= h->cap, RTX_EINVAL, "rtx_hits_read: bad argument");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    if (flags) RTX_CHECK_HIP(hipMemcpy(flags, h->flags.p, n * 4, hipMemcpyDeviceToHost));
    if (t) RTX_CHECK_HIP(hipMemcpy(t, h->t.p, n * 4, hipMemcpyDeviceToHost));
    if (id) RTX_CHECK_HIP(hipMemcpy(id, h->id.p, n * 4, hipMemcpyDeviceToHost));
    if (bary2) RTX_CHECK_HIP(hipMemcpy(bary2, h->bary.p, n * 8, hipMemcpyDeviceToHost));
    if (n4) RTX_CHECK_HIP(hipMemcpy(n4, h->n4.p, n * 16, hipMemcpyDeviceToHost));
    if (p4) RTX_CHECK_HIP(hipMemcpy(p4, h->p4.p, n * 16, hipMemcpyDeviceToHost));
    if (uv2) RTX_CHECK_HIP(hipMemcpy(uv2, h->uv2.p, n * 8, hipMemcpyDeviceToHost));
    return RTX_OK;
}

int32_t rtx_hits_write(rtx_hits* h, const uint32_t* flags, const float* t, const int32_t* id, const float* n4,
                       const float* p4, uint64_t n) {
    RTX_REQUIRE(h && n <= h->cap, RTX_EINVAL, "rtx_hits_write: bad argument");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    if (flags) RTX_CHECK_HIP(hipMemcpy(h->flags.p, flags, n * 4, hipMemcpyHostToDevice));
    if (t) RTX_CHECK_HIP(hipMemcpy(h->t.p, t, n * 4, hipMemcpyHostToDevice));
    if (id) RTX_CHECK_HIP(hipMemcpy(h->id.p, id, n * 4, hipMemcpyHostToDevice));
    if (n4) RTX_CHECK_HIP(hipMemcpy(h->n4.p, n4, n * 16, hipMemcpyHostToDevice));
    if (p4) RTX_CHECK_HIP(hipMemcpy(h->p4.p, p4, n * 16, hipMemcpyHostToDevice));
    return RTX_OK;
}

// ------------------------------------------------------------ entry pts --
int32_t rtx_raygen_primary(rtx_rays* rays, const rtx_camera* cam, int32_t width, int32_t height, int32_t quad_size,
                           int32_t spp, const float* psamples, uint64_t offset, uint64_t count, void* stream) {
    RTX_REQUIRE(rays && cam, RTX_EINVAL, "rtx_raygen_primary: null argument");
    RTX_REQUIRE(width > 0 && height > 0 && quad_size > 0 && spp > 0, RTX_EINVAL, "rtx_raygen_primary: bad size");
    RTX_REQUIRE(count <= rays->cap, RTX_EINVAL, "rtx_raygen_primary: count exceeds bundle capacity");
    RTX_REQUIRE(offset + count <= (uint64_t)width * height * spp, RTX_EINVAL,
                "rtx_raygen_primary: samples beyond width*height*spp");
    RTX_REQUIRE((uint64_t)width * height * spp < (1ull << 31), RTX_EUNSUPPORTED, "rtx_raygen_primary: too many samples");
    hipStream_t st = rays->ctx->pick(stream);
    float one[3] = {0.f, 0.f, 1.f};
    RTX_REQUIRE(psamples || spp == 1, RTX_EINVAL, "rtx_raygen_primary: spp > 1 needs pixel samples");
    const float* hs = psamples ? psamples : one;
    const size_t n_s = psamples ? (size_t)spp : 1;
    int32_t e = rays->psamp.reserve(rays->ctx, 12 * n_s);
    if (e) return e;
    float* dtab = rays->psamp.as<float>();
    RTX_CHECK_HIP(hipMemcpyAsync(dtab, hs, 12 * n_s, hipMemcpyHostToDevice, st));
    ProfScope ps(rays->ctx, st, "raygen_primary");
    uint32_t grid = rtx_div_up(count ? count : 1, 256);
    hipLaunchKernelGGL(k_raygen_primary, dim3(grid), dim3(256), 0, st, rays->ori4.as<float4>(),
                       rays->dir4.as<float4>(), rays->pixel.as<int32_t>(), rays->contrib.as<float>(),
                       rays->count.as<uint32_t>(), *cam, width, height, quad_size, dtab, offset, count);
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

int32_t rtx_trace(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream) {
    RTX_REQUIRE(s && rays && hits, RTX_EINVAL, "rtx_trace: null argument");
    RTX_REQUIRE(s->built, RTX_ESTATE, "rtx_trace: scene has no BVH (call rtx_lbvh_build)");
    RTX_REQUIRE(count == RTX_DEVICE_COUNT || (count <= rays->cap && count <= hits->cap), RTX_EINVAL,
                "rtx_trace: count exceeds bundle capacity");
    RTX_REQUIRE(hits->cap >= rays->cap || count != RTX_DEVICE_COUNT, RTX_EINVAL,
                "rtx_trace: hit bundle smaller than ray bundle");
    hipStream_t st = s->ctx->pick(stream);
    if (count == 0) return RTX_OK;
    ProfScope ps(s->ctx, st, "trace");
    hipLaunchKernelGGL(k_trace, dim3(trace_grid(s->ctx, count, rays->cap)), dim3(TB), 0, st, s->node.as<float4>(),
                       s->delta.as<float4>(), s->tri.as<float4>(), s->nrm.as<float4>(), s->uv.as<float2>(),
                       rays->ori4.as<float4>(), rays->dir4.as<float4>(), rays->count.as<uint32_t>(), count,
                       hits->flags.as<uint32_t>(), hits->t.as<float>(), hits->id.as<int32_t>(),
                       hits->bary.as<float2>(), hits->n4.as<float4>(), hits->p4.as<float4>(), hits->uv2.as<float2>());
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

int32_t rtx_trace_shadow_impl(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, uint64_t count, uint64_t* dev_shadow_count,
                              hipStream_t st) {
    RTX_REQUIRE(s && rays && hits, RTX_EINVAL, "rtx_trace_shadow: null argument");
    RTX_REQUIRE(s->built, RTX_ESTATE, "rtx_trace_shadow: scene has no BVH");
    RTX_REQUIRE(s->lights.p, RTX_ESTATE, "rtx_trace_shadow: scene has no lights");
    RTX_REQUIRE(count == RTX_DEVICE_COUNT || count <= hits->cap, RTX_EINVAL, "rtx_trace_shadow: count too large");
    if (count == 0) return RTX_OK;
    ProfScope ps(s->ctx, st, "trace_shadow");
    hipLaunchKernelGGL(k_shadow, dim3(trace_grid(s->ctx, count, rays->cap)), dim3(TB), 0, st, s->node.as<float4>(),
                       s->delta.as<float4>(), s->tri.as<float4>(), s->lights.as<rtx_lights>(),
                       rays->count.as<uint32_t>(), count, hits->flags.as<uint32_t>(), hits->p4.as<float4>(),
                       (unsigned long long*)dev_shadow_count);
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

int32_t rtx_trace_shadow(rtx_scene* s, rtx_rays* rays, rtx_hits* hits, uint64_t count, void* stream) {
    RTX_REQUIRE(s, RTX_EINVAL, "rtx_trace_shadow: null scene");
    return rtx_trace_shadow_impl(s, rays, hits, count, nullptr, s->ctx->pick(stream));
}
