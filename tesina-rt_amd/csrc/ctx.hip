// Context, errors, device buffers, profiling, scene uploads and read-backs.
//
// Replaces CLInfo / DeviceInterface / DeviceMemory (src/cl-gl/opencl-init.cpp,
// src/gpu/interface.cpp, src/gpu/memory.cpp) and the Scene device-buffer
// management of src/rt/scene.cpp.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "devmath.hpp"
#include "rtx_impl.hpp"

// ------------------------------------------------------------- errors ----
static thread_local std::string g_err;
void rtx_set_error(const std::string& msg) { g_err = msg; }
const char* rtx_last_error(void) { return g_err.c_str(); }
int32_t rtx_abi_version(void) { return 1; }

// ------------------------------------------------------------- buffers ---
int32_t DevBuf::reserve(rtx_ctx* c, size_t n) {
    if (n <= bytes && p) return RTX_OK;
    release();
    ctx = c;
    size_t want = n ? n : 16;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        rtx_set_error(std::string("hipMalloc(") + std::to_string(want) + "): " + hipGetErrorString(e));
        (void)hipGetLastError();
        return RTX_ENOMEM;
    }
    bytes = want;
    if (ctx) ctx->bytes_in_use += want;
    return RTX_OK;
}

void DevBuf::release() {
    if (p) {
        (void)hipFree(p);
        if (ctx) ctx->bytes_in_use -= bytes;
    }
    p = nullptr;
    bytes = 0;
}

namespace {
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Vertex upload: positions to float4, normals normalized once
// (complete_trace_info normalizes each vertex normal at every hit,
// src/kernel/trace-bvh.cl:222-224; same operation, done here once).
__global__ void k_vertices(const float* __restrict__ pos, const float* __restrict__ nrm, const float* __restrict__ uv,
                           float4* __restrict__ pos4, float4* __restrict__ nrm4, float4* __restrict__ vnrm4,
                           float2* __restrict__ uv2, uint32_t V, uint32_t stride_f) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const float* p = pos + (size_t)i * stride_f;
    pos4[i] = make_float4(p[0], p[1], p[2], 0.f);
    if (nrm4) {
        v3 raw = mk3(0.f, 0.f, 0.f), n = raw;
        if (nrm) {
            const float* q = nrm + (size_t)i * stride_f;
            raw = mk3(q[0], q[1], q[2]);
            n = normalize(raw);
        }
        vnrm4[i] = make_float4(raw.x, raw.y, raw.z, 0.f);
        nrm4[i] = make_float4(n.x, n.y, n.z, 0.f);
    }
    if (uv2) {
        if (uv) {
            const float* t = uv + (size_t)i * (stride_f == 3 ? 2 : stride_f);
            uv2[i] = make_float2(t[0], t[1]);
        } else {
            uv2[i] = make_float2(0.f, 0.f);
        }
    }
}

// Heightfield displacement: y = sum_k a sin(kappa (d . (x,z)) - omega t + phi),
// vertex normal = normalize(-dy/dx, 1, -dy/dz).  Reads x,z from pos4, writes
// y, the vertex normal (vnrm4, what a vertex buffer holds) and its
// normalization (nrm4, what the hit completion uses).
__device__ __forceinline__ void displace_vertex(float4* __restrict__ pos4, float4* __restrict__ nrm4,
                                                float4* __restrict__ vnrm4, const float* __restrict__ waves, int nw,
                                                float t, uint32_t i) {
    float4 p = pos4[i];
    float y = 0.f, gx = 0.f, gz = 0.f;
    for (int k = 0; k < nw; ++k) {
        const float* w = waves + 6 * k;
        float ph = w[1] * (w[2] * p.x + w[3] * p.z) - w[4] * t + w[5];
        float s, c;
        sincosf(ph, &s, &c);
        y += w[0] * s;
        gx += w[0] * c * w[1] * w[2];
        gz += w[0] * c * w[1] * w[3];
    }
    pos4[i] = make_float4(p.x, y, p.z, 0.f);
    const v3 vn = normalize(mk3(-gx, 1.f, -gz));
    vnrm4[i] = make_float4(vn.x, vn.y, vn.z, 0.f);
    const v3 n = normalize(vn);
    nrm4[i] = make_float4(n.x, n.y, n.z, 0.f);
}

__global__ void k_waves(float4* __restrict__ pos4, float4* __restrict__ nrm4, float4* __restrict__ vnrm4,
                        const float* __restrict__ waves, int nw, float t, uint32_t V) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    displace_vertex(pos4, nrm4, vnrm4, waves, nw, t, i);
}

// Same at t = t0 + dt * frame, with the frame number read from the device
// clock (so that a captured frame advances the animation on every replay).
__global__ void k_waves_clock(float4* __restrict__ pos4, float4* __restrict__ nrm4, float4* __restrict__ vnrm4,
                              const float* __restrict__ waves, int nw, float t0, float dt,
                              const uint32_t* __restrict__ clock, uint32_t V) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const float t = t0 + dt * (float)(*clock);
    displace_vertex(pos4, nrm4, vnrm4, waves, nw, t, i);
}

__global__ void k_tick(uint32_t* clock) { *clock += 1u; }
}  // namespace

int32_t rtx_fill_u32(rtx_ctx* ctx, void* p, uint32_t value, uint64_t n, hipStream_t st) {
    if (!n) return RTX_OK;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    ProfScope ps(ctx, st, "fill");
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, st, (uint32_t*)p, value, n);
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

// ------------------------------------------------------------ profiling --
ProfScope::ProfScope(rtx_ctx* c, hipStream_t s, const char* name) : ctx(c), st(s) {
    if (!ctx || !ctx->profiling) return;
    hipEvent_t a, b;
    if (ctx->event_pool.size() >= 2) {
        a = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        b = ctx->event_pool.back();
        ctx->event_pool.pop_back();
    } else {
        if (hipEventCreate(&a) != hipSuccess) return;
        if (hipEventCreate(&b) != hipSuccess) return;
    }
    (void)hipEventRecord(a, st);
    ctx->pending.push_back(KernelTiming{name, a, b});
    slot = (int)ctx->pending.size() - 1;
}

ProfScope::~ProfScope() {
    if (slot >= 0) (void)hipEventRecord(ctx->pending[slot].stop, st);
}

static void drain_profile(rtx_ctx* c) {
    for (auto& k : c->pending) {
        (void)hipEventSynchronize(k.stop);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, k.start, k.stop);
        bool found = false;
        for (auto& a : c->acc)
            if (a.name == k.name) {
                a.count++;
                a.ms += ms;
                found = true;
                break;
            }
        if (!found) c->acc.push_back(rtx_ctx::Acc{k.name, 1, (double)ms});
        c->event_pool.push_back(k.start);
        c->event_pool.push_back(k.stop);
    }
    c->pending.clear();
}

int32_t rtx_profile_enable(rtx_ctx* ctx, int32_t on) {
    RTX_REQUIRE(ctx, RTX_EINVAL, "rtx_profile_enable: null context");
    ctx->profiling = on != 0;
    return RTX_OK;
}

int32_t rtx_profile_enabled(rtx_ctx* ctx) { return ctx && ctx->profiling ? 1 : 0; }

int32_t rtx_profile_reset(rtx_ctx* ctx) {
    RTX_REQUIRE(ctx, RTX_EINVAL, "rtx_profile_reset: null context");
    drain_profile(ctx);
    ctx->acc.clear();
    return RTX_OK;
}

int64_t rtx_profile_report(rtx_ctx* ctx, char* buf, uint64_t cap) {
    RTX_REQUIRE(ctx, RTX_EINVAL, "rtx_profile_report: null context");
    drain_profile(ctx);
    std::string out;
    char line[256];
    for (auto& a : ctx->acc) {
        snprintf(line, sizeof line, "%s %llu %.6f\n", a.name.c_str(), (unsigned long long)a.count, a.ms);
        out += line;
    }
    if (buf && cap) {
        size_t n = out.size() < cap - 1 ? out.size() : (size_t)cap - 1;
        memcpy(buf, out.data(), n);
        buf[n] = 0;
    }
    return (int64_t)out.size() + 1;
}

// -------------------------------------------------------------- context --
int32_t rtx_init(int32_t device, rtx_ctx** out) {
    RTX_REQUIRE(out, RTX_EINVAL, "rtx_init: null out");
    *out = nullptr;
    int n = 0;
    RTX_CHECK_HIP(hipGetDeviceCount(&n));
    RTX_REQUIRE(device >= 0 && device < n, RTX_EINVAL, "rtx_init: no such device");
    RTX_CHECK_HIP(hipSetDevice(device));
    rtx_ctx* c = new rtx_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        rtx_set_error("rtx_init: hipStreamCreate failed");
        return RTX_EDEVICE;
    }
    if (int32_t e = rtx_lbvh_scratch_init(c)) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return e;
    }
    *out = c;
    return RTX_OK;
}

int32_t rtx_shutdown(rtx_ctx* ctx) {
    RTX_REQUIRE(ctx, RTX_EINVAL, "rtx_shutdown: null context");
    (void)hipStreamSynchronize(ctx->stream);
    drain_profile(ctx);
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RTX_OK;
}

void* rtx_stream(rtx_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int32_t rtx_sync(rtx_ctx* ctx, void* stream) {
    RTX_REQUIRE(ctx, RTX_EINVAL, "rtx_sync: null context");
    RTX_CHECK_HIP(hipStreamSynchronize(ctx->pick(stream)));
    return RTX_OK;
}

uint64_t rtx_mem_in_use(rtx_ctx* ctx) { return ctx ? ctx->bytes_in_use : 0; }

int32_t rtx_tune_set(rtx_ctx* ctx, const char* key, int64_t value) {
    RTX_REQUIRE(ctx && key, RTX_EINVAL, "rtx_tune_set: null argument");
    std::string k(key);
    if (k == "trace_variant") {
        RTX_REQUIRE(value >= 0 && value <= 4, RTX_EINVAL, "rtx_tune_set: trace_variant must be 0 to 4");
        ctx->tune.trace_variant = value;
    } else if (k == "trace_block") {
        RTX_REQUIRE(value == 64 || value == 128 || value == 256, RTX_EINVAL,
                    "rtx_tune_set: trace_block must be 64, 128 or 256");
        ctx->tune.trace_block = value;
    } else if (k == "xcds") {
        RTX_REQUIRE(value >= 1 && value <= 64, RTX_EINVAL, "rtx_tune_set: xcds must be in [1, 64]");
        ctx->tune.xcds = value;
    } else if (k == "refill") {
        RTX_REQUIRE(value >= 1 && value <= 64, RTX_EINVAL, "rtx_tune_set: refill must be in [1, 64]");
        ctx->tune.refill = value;
    } else if (k == "coop_leaf") {
        RTX_REQUIRE(value >= 0 && value <= (1 << 20), RTX_EINVAL, "rtx_tune_set: coop_leaf must be in [0, 2^20]");
        ctx->tune.coop_leaf = value;
    } else if (k == "side_stream") {
        RTX_REQUIRE(value == 0 || value == 1, RTX_EINVAL, "rtx_tune_set: side_stream must be 0 or 1");
        ctx->tune.side_stream = value;
    } else if (k == "inst_lane_roots") {
        RTX_REQUIRE(value >= 0 && value <= 2, RTX_EINVAL, "rtx_tune_set: inst_lane_roots must be 0, 1 or 2");
        ctx->tune.inst_lane_roots = value;
    } else if (k == "wf_sched") {
        RTX_REQUIRE(value == 0 || value == 1, RTX_EINVAL, "rtx_tune_set: wf_sched must be 0 or 1");
        ctx->tune.wf_sched = value;
    } else if (k == "sort_bits") {
        RTX_REQUIRE(value == 4 || value == 8, RTX_EINVAL, "rtx_tune_set: sort_bits must be 4 or 8");
        ctx->tune.sort_bits = value;
    } else if (k == "sort_items") {
        RTX_REQUIRE(value == 16 || value == 32, RTX_EINVAL, "rtx_tune_set: sort_items must be 16 or 32");
        ctx->tune.sort_items = value;
    } else if (k == "shadow_order") {
        RTX_REQUIRE(value == 0 || value == 1, RTX_EINVAL, "rtx_tune_set: shadow_order must be 0 or 1");
        ctx->tune.shadow_order = value;
    } else if (k == "build_grid") {
        RTX_REQUIRE(value >= 1 && value <= 16, RTX_EINVAL, "rtx_tune_set: build_grid must be in [1, 16]");
        ctx->tune.build_grid = value;
    } else if (k == "srv_step_pixels") {
        RTX_REQUIRE(value >= 0, RTX_EINVAL, "rtx_tune_set: srv_step_pixels must be >= 0");
        ctx->tune.srv_step_pixels = value;
    } else if (k == "wf_share") {
        RTX_REQUIRE(value >= 0 && value <= 100, RTX_EINVAL, "rtx_tune_set: wf_share must be in [0, 100]");
        ctx->tune.wf_share = value;
    } else if (k == "stack_lds") {
        RTX_REQUIRE(value == 8 || value == 16 || value == 32, RTX_EINVAL, "rtx_tune_set: stack_lds must be 8, 16 or 32");
        ctx->tune.stack_lds = value;
    } else if (k == "wf_grain") {
        RTX_REQUIRE(value >= 1 && value <= 64, RTX_EINVAL, "rtx_tune_set: wf_grain must be in [1, 64]");
        ctx->tune.wf_grain = value;
    } else {
        rtx_set_error("rtx_tune_set: unknown key " + k);
        return RTX_EINVAL;
    }
    return RTX_OK;
}

int64_t rtx_tune_get(rtx_ctx* ctx, const char* key) {
    if (!ctx || !key) return -1;
    std::string k(key);
    if (k == "trace_variant") return ctx->tune.trace_variant;
    if (k == "trace_block") return ctx->tune.trace_block;
    if (k == "xcds") return ctx->tune.xcds;
    if (k == "coop_leaf") return ctx->tune.coop_leaf;
    if (k == "refill") return ctx->tune.refill;
    if (k == "side_stream") return ctx->tune.side_stream;
    if (k == "inst_lane_roots") return ctx->tune.inst_lane_roots;
    if (k == "wf_grain") return ctx->tune.wf_grain;
    if (k == "wf_sched") return ctx->tune.wf_sched;
    if (k == "stack_lds") return ctx->tune.stack_lds;
    if (k == "shadow_order") return ctx->tune.shadow_order;
    if (k == "wf_share") return ctx->tune.wf_share;
    if (k == "srv_step_pixels") return ctx->tune.srv_step_pixels;
    if (k == "build_grid") return ctx->tune.build_grid;
    if (k == "sort_bits") return ctx->tune.sort_bits;
    if (k == "sort_items") return ctx->tune.sort_items;
    return -1;
}

// ---------------------------------------------------------------- scene --
int32_t rtx_scene_create(rtx_ctx* ctx, rtx_scene** out) {
    RTX_REQUIRE(ctx && out, RTX_EINVAL, "rtx_scene_create: null argument");
    rtx_scene* s = new rtx_scene();
    s->ctx = ctx;
    *out = s;
    return RTX_OK;
}

// Device bytes held by a scene (buffers and BVH, build scratch included).
uint64_t rtx_scene_device_bytes(const rtx_scene* sc) {
    rtx_scene* s = const_cast<rtx_scene*>(sc);
    const DevBuf* all[] = {&s->pos,        &s->nrm,  &s->vnrm,     &s->uv,   &s->idx,    &s->idx_alt, &s->matmap,
                           &s->matmap_alt, &s->orig, &s->orig_alt, &s->mats, &s->lights, &s->waves,   &s->stage,
                           &s->clock,      &s->tex,  &s->tex_tab};
    uint64_t b = 0;
    for (const DevBuf* d : all) b += d->p ? d->bytes : 0;
    for (auto& set : s->sets) set.each([&](DevBuf& d) { b += d.p ? d.bytes : 0; });
    for (const auto& c : s->rc) b += c.node.bytes + c.delta.bytes + c.tri.bytes + c.nrm.bytes + c.matmap.bytes;
    return b;
}

int32_t rtx_scene_destroy(rtx_scene* s) {
    RTX_REQUIRE(s, RTX_EINVAL, "rtx_scene_destroy: null scene");
    (void)hipStreamSynchronize(s->ctx->stream);
    DevBuf* all[] = {&s->pos,        &s->nrm,  &s->vnrm,     &s->uv,   &s->idx,    &s->idx_alt, &s->matmap,
                     &s->matmap_alt, &s->orig, &s->orig_alt, &s->mats, &s->lights, &s->waves,   &s->stage,
                     &s->clock,      &s->tex,  &s->tex_tab};
    for (auto& set : s->sets) set.each([](DevBuf& d) { d.release(); });
    for (auto& c : s->rc) {
        c.node.release();
        c.delta.release();
        c.tri.release();
        c.nrm.release();
        c.matmap.release();
    }
    for (DevBuf* b : all) b->release();
    delete s;
    return RTX_OK;
}

static int32_t upload_vertices(rtx_scene* s, const float* pos, const float* nrm, const float* uv, uint32_t V,
                               uint32_t stride_f) {
    rtx_ctx* c = s->ctx;
    int32_t e;
    if ((e = s->pos.reserve(c, (size_t)V * 16))) return e;
    if ((e = s->nrm.reserve(c, (size_t)V * 16))) return e;
    if ((e = s->vnrm.reserve(c, (size_t)V * 16))) return e;
    if ((e = s->uv.reserve(c, (size_t)V * 8))) return e;
    // staging: raw host arrays copied as-is, converted on the device
    size_t fpv = stride_f;
    DevBuf sp, sn, su;
    if ((e = sp.reserve(nullptr, (size_t)V * fpv * 4))) return e;
    RTX_CHECK_HIP(hipMemcpy(sp.p, pos, (size_t)V * fpv * 4, hipMemcpyHostToDevice));
    if (nrm) {
        if ((e = sn.reserve(nullptr, (size_t)V * fpv * 4))) return e;
        RTX_CHECK_HIP(hipMemcpy(sn.p, nrm, (size_t)V * fpv * 4, hipMemcpyHostToDevice));
    }
    size_t uvf = stride_f == 3 ? 2 : stride_f;
    if (uv) {
        if ((e = su.reserve(nullptr, (size_t)V * uvf * 4))) return e;
        RTX_CHECK_HIP(hipMemcpy(su.p, uv, (size_t)V * uvf * 4, hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(k_vertices, dim3(rtx_div_up(V, 256)), dim3(256), 0, c->stream, sp.as<float>(),
                       nrm ? sn.as<float>() : nullptr, uv ? su.as<float>() : nullptr, s->pos.as<float4>(),
                       s->nrm.as<float4>(), s->vnrm.as<float4>(), s->uv.as<float2>(), V, stride_f);
    RTX_CHECK_HIP(hipGetLastError());
    RTX_CHECK_HIP(hipStreamSynchronize(c->stream));
    sp.release();
    sn.release();
    su.release();
    s->V = V;
    s->built = false;
    return RTX_OK;
}

int32_t rtx_scene_set_vertices(rtx_scene* s, const float* pos, const float* nrm, const float* uv, uint32_t V) {
    RTX_REQUIRE(s && pos && V, RTX_EINVAL, "rtx_scene_set_vertices: null argument or V == 0");
    return upload_vertices(s, pos, nrm, uv, V, 3);
}

int32_t rtx_scene_set_vertices_ref(rtx_scene* s, const void* vertex80, uint32_t V) {
    RTX_REQUIRE(s && vertex80 && V, RTX_EINVAL, "rtx_scene_set_vertices_ref: null argument or V == 0");
    // Vertex: position@0, normal@16, tangent@32, bitangent@48, texCoord@64 (80 B).
    const float* f = static_cast<const float*>(vertex80);
    std::vector<float> p(3 * (size_t)V), n(3 * (size_t)V), t(2 * (size_t)V);
    for (uint32_t i = 0; i < V; ++i) {
        for (int k = 0; k < 3; ++k) {
            p[3 * (size_t)i + k] = f[20 * (size_t)i + k];
            n[3 * (size_t)i + k] = f[20 * (size_t)i + 4 + k];
        }
        t[2 * (size_t)i] = f[20 * (size_t)i + 16];
        t[2 * (size_t)i + 1] = f[20 * (size_t)i + 17];
    }
    return upload_vertices(s, p.data(), n.data(), t.data(), V, 3);
}

int32_t rtx_scene_update_positions(rtx_scene* s, const float* host_pos, void* stream) {
    RTX_REQUIRE(s && host_pos && s->V, RTX_EINVAL, "rtx_scene_update_positions: no vertices");
    hipStream_t st = s->ctx->pick(stream);
    int32_t e = s->stage.reserve(s->ctx, (size_t)s->V * 12);
    if (e) return e;
    RTX_CHECK_HIP(hipMemcpyAsync(s->stage.p, host_pos, (size_t)s->V * 12, hipMemcpyHostToDevice, st));
    ProfScope ps(s->ctx, st, "vertex_update");
    hipLaunchKernelGGL(k_vertices, dim3(rtx_div_up(s->V, 256)), dim3(256), 0, st, s->stage.as<float>(),
                       (const float*)nullptr, (const float*)nullptr, s->pos.as<float4>(), (float4*)nullptr,
                       (float4*)nullptr, (float2*)nullptr, s->V, 3u);
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

int32_t rtx_scene_displace_waves(rtx_scene* s, const float* waves, int32_t n, float t, void* stream) {
    RTX_REQUIRE(s && s->V && waves && n > 0 && n <= 64, RTX_EINVAL, "rtx_scene_displace_waves: bad argument");
    hipStream_t st = s->ctx->pick(stream);
    int32_t e = s->waves.reserve(s->ctx, 64 * 6 * 4);
    if (e) return e;
    float* dev_w = s->waves.as<float>();
    RTX_CHECK_HIP(hipMemcpyAsync(dev_w, waves, (size_t)n * 6 * 4, hipMemcpyHostToDevice, st));
    ProfScope ps(s->ctx, st, "waves");
    hipLaunchKernelGGL(k_waves, dim3(rtx_div_up(s->V, 256)), dim3(256), 0, st, s->pos.as<float4>(),
                       s->nrm.as<float4>(), s->vnrm.as<float4>(), dev_w, n, t, s->V);
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

int32_t rtx_scene_set_waves(rtx_scene* s, const float* waves, int32_t n) {
    RTX_REQUIRE(s && waves && n > 0 && n <= 64, RTX_EINVAL, "rtx_scene_set_waves: bad argument");
    int32_t e = s->waves.reserve(s->ctx, 64 * 6 * 4);
    if (e) return e;
    if ((e = s->clock.reserve(s->ctx, 4))) return e;
    RTX_CHECK_HIP(hipDeviceSynchronize());
    RTX_CHECK_HIP(hipMemcpy(s->waves.p, waves, (size_t)n * 6 * 4, hipMemcpyHostToDevice));
    const uint32_t zero = 0;
    RTX_CHECK_HIP(hipMemcpy(s->clock.p, &zero, 4, hipMemcpyHostToDevice));
    s->nwaves = n;
    return RTX_OK;
}

int32_t rtx_scene_set_clock(rtx_scene* s, uint32_t frame) {
    RTX_REQUIRE(s && s->nwaves > 0, RTX_ESTATE, "rtx_scene_set_clock: no waves set");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    RTX_CHECK_HIP(hipMemcpy(s->clock.p, &frame, 4, hipMemcpyHostToDevice));
    return RTX_OK;
}

int32_t rtx_scene_get_clock(rtx_scene* s, uint32_t* frame) {
    RTX_REQUIRE(s && frame && s->nwaves > 0, RTX_ESTATE, "rtx_scene_get_clock: no waves set");
    RTX_CHECK_HIP(hipDeviceSynchronize());
    RTX_CHECK_HIP(hipMemcpy(frame, s->clock.p, 4, hipMemcpyDeviceToHost));
    return RTX_OK;
}

int32_t rtx_scene_animate(rtx_scene* s, float t0, float dt, void* stream) {
    RTX_REQUIRE(s && s->V, RTX_EINVAL, "rtx_scene_animate: no vertices");
    RTX_REQUIRE(s->nwaves > 0, RTX_ESTATE, "rtx_scene_animate: no waves set (rtx_scene_set_waves)");
    hipStream_t st = s->ctx->pick(stream);
    ProfScope ps(s->ctx, st, "waves");
    hipLaunchKernelGGL(k_waves_clock, dim3(rtx_div_up(s->V, 256)), dim3(256), 0, st, s->pos.as<float4>(),
                       s->nrm.as<float4>(), s->vnrm.as<float4>(), s->waves.as<float>(), s->nwaves, t0, dt,
                       s->clock.as<uint32_t>(), s->V);
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, st, s->clock.as<uint32_t>());
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

int32_t rtx_scene_set_triangles(rtx_scene* s, const uint32_t* idx, const int32_t* matmap, uint32_t T) {
    RTX_REQUIRE(s && idx && T, RTX_EINVAL, "rtx_scene_set_triangles: null argument or T == 0");
    RTX_REQUIRE(s->V, RTX_ESTATE, "rtx_scene_set_triangles: set vertices first");
    for (uint64_t i = 0; i < 3ull * T; ++i)
        RTX_REQUIRE(idx[i] < s->V, RTX_EINVAL, "rtx_scene_set_triangles: vertex index out of range");
    rtx_ctx* c = s->ctx;
    int32_t e;
    if ((e = s->idx.reserve(c, (size_t)T * 12))) return e;
    if ((e = s->idx_alt.reserve(c, (size_t)T * 12))) return e;
    if ((e = s->matmap.reserve(c, (size_t)T * 4))) return e;
    if ((e = s->matmap_alt.reserve(c, (size_t)T * 4))) return e;
    if ((e = s->orig.reserve(c, (size_t)T * 4))) return e;
    if ((e = s->orig_alt.reserve(c, (size_t)T * 4))) return e;
    RTX_CHECK_HIP(hipMemcpy(s->idx.p, idx, (size_t)T * 12, hipMemcpyHostToDevice));
    std::vector<int32_t> mm(T, 0);
    if (matmap) memcpy(mm.data(), matmap, (size_t)T * 4);
    RTX_CHECK_HIP(hipMemcpy(s->matmap.p, mm.data(), (size_t)T * 4, hipMemcpyHostToDevice));
    std::vector<uint32_t> o(T);
    for (uint32_t i = 0; i < T; ++i) o[i] = i;
    RTX_CHECK_HIP(hipMemcpy(s->orig.p, o.data(), (size_t)T * 4, hipMemcpyHostToDevice));
    s->T = T;
    s->built = false;
    return RTX_OK;
}

int32_t rtx_scene_set_materials(rtx_scene* s, const rtx_material* mats, uint32_t M) {
    RTX_REQUIRE(s && mats && M, RTX_EINVAL, "rtx_scene_set_materials: null argument or M == 0");
    int32_t e = s->mats.reserve(s->ctx, (size_t)M * sizeof(rtx_material));
    if (e) return e;
    RTX_CHECK_HIP(hipMemcpy(s->mats.p, mats, (size_t)M * sizeof(rtx_material), hipMemcpyHostToDevice));
    s->M = M;
    return RTX_OK;
}

int32_t rtx_scene_set_lights(rtx_scene* s, const rtx_lights* lights) {
    RTX_REQUIRE(s && lights, RTX_EINVAL, "rtx_scene_set_lights: null argument");
    RTX_REQUIRE(lights->count >= 0 && lights->count <= RTX_MAX_LIGHTS, RTX_EINVAL,
                "rtx_scene_set_lights: count out of range");
    int32_t e = s->lights.reserve(s->ctx, sizeof(rtx_lights));
    if (e) return e;
    RTX_CHECK_HIP(hipMemcpy(s->lights.p, lights, sizeof(rtx_lights), hipMemcpyHostToDevice));
    s->host_lights = *lights;
    return RTX_OK;
}

uint32_t rtx_scene_triangle_count(rtx_scene* s) { return s ? s->T : 0; }

// The atlas holds the material images, then the cubemap's faces; it is
// rebuilt whenever either changes (uploads synchronize: shading in flight may
// read the old atlas).
static int32_t upload_atlas(rtx_scene* s) {
    std::vector<TexEnt> tab;
    uint64_t off = 0;
    auto add = [&](const std::vector<int32_t>& dims) {
        for (size_t k = 0; k + 1 < dims.size(); k += 2) {
            tab.push_back(TexEnt{off, dims[k], dims[k + 1]});
            off += (uint64_t)dims[k] * (uint64_t)dims[k + 1];
        }
    };
    add(s->host_tex_dims);
    if (s->cube) add(s->host_cube_dims);
    RTX_CHECK_HIP(hipDeviceSynchronize());
    if (tab.empty()) {
        s->tex.release();
        s->tex_tab.release();
        return RTX_OK;
    }
    int32_t e = 0;
    if ((e = s->tex.reserve(s->ctx, off * 4)) || (e = s->tex_tab.reserve(s->ctx, tab.size() * sizeof(TexEnt))))
        return e;
    const size_t nt = s->host_tex.size();
    if (nt) RTX_CHECK_HIP(hipMemcpy(s->tex.p, s->host_tex.data(), nt, hipMemcpyHostToDevice));
    if (s->cube)
        RTX_CHECK_HIP(hipMemcpy(static_cast<uint8_t*>(s->tex.p) + nt, s->host_cube.data(), s->host_cube.size(),
                                hipMemcpyHostToDevice));
    RTX_CHECK_HIP(hipMemcpy(s->tex_tab.p, tab.data(), tab.size() * sizeof(TexEnt), hipMemcpyHostToDevice));
    return RTX_OK;
}

static int32_t image_bytes(const int32_t* dims, int32_t n, const char* what, uint64_t* bytes) {
    uint64_t b = 0;
    for (int32_t k = 0; k < n; ++k) {
        RTX_REQUIRE(dims[2 * k] > 0 && dims[2 * k + 1] > 0 && dims[2 * k] <= 65536 && dims[2 * k + 1] <= 65536,
                    RTX_EINVAL, std::string(what) + ": image size out of range");
        b += 4ull * (uint64_t)dims[2 * k] * (uint64_t)dims[2 * k + 1];
    }
    *bytes = b;
    return RTX_OK;
}

int32_t rtx_scene_set_textures(rtx_scene* s, const uint8_t* rgba, const int32_t* dims, int32_t count) {
    RTX_REQUIRE(s && count >= 0 && count <= RTX_MAX_TEXTURES, RTX_EINVAL,
                "rtx_scene_set_textures: null scene or count out of range");
    RTX_REQUIRE(count == 0 || (rgba && dims), RTX_EINVAL, "rtx_scene_set_textures: null images");
    uint64_t bytes = 0;
    if (count) {
        int32_t e = image_bytes(dims, count, "rtx_scene_set_textures", &bytes);
        if (e) return e;
    }
    s->host_tex.assign(rgba, rgba + bytes);
    s->host_tex_dims.assign(dims, dims + 2 * (size_t)count);
    s->ntex = count;
    return upload_atlas(s);
}

int32_t rtx_scene_set_cubemap(rtx_scene* s, const uint8_t* rgba, const int32_t* dims) {
    RTX_REQUIRE(s, RTX_EINVAL, "rtx_scene_set_cubemap: null scene");
    if (!rgba) {
        s->cube = false;
        s->host_cube.clear();
        s->host_cube_dims.clear();
        return upload_atlas(s);
    }
    RTX_REQUIRE(dims, RTX_EINVAL, "rtx_scene_set_cubemap: null dims");
    uint64_t bytes = 0;
    int32_t e = image_bytes(dims, 6, "rtx_scene_set_cubemap", &bytes);
    if (e) return e;
    s->host_cube.assign(rgba, rgba + bytes);
    s->host_cube_dims.assign(dims, dims + 12);
    s->cube = true;
    return upload_atlas(s);
}

// ----------------------------------------------------------- read-backs --
static int32_t readback(rtx_scene* s, const DevBuf& b, void* dst, size_t bytes, const char* what) {
    RTX_REQUIRE(s && dst, RTX_EINVAL, std::string(what) + ": null argument");
    RTX_REQUIRE(b.p && b.bytes >= bytes, RTX_ESTATE, std::string(what) + ": buffer not available");
    RTX_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    RTX_CHECK_HIP(hipDeviceSynchronize());
    RTX_CHECK_HIP(hipMemcpy(dst, b.p, bytes, hipMemcpyDeviceToHost));
    return RTX_OK;
}

int32_t rtx_read_codes(rtx_scene* s, uint32_t* codes, uint32_t T) {
    RTX_REQUIRE(s && s->built && T == s->T, RTX_EINVAL, "rtx_read_codes: no BVH or wrong T");
    return readback(s, s->b().codes, codes, (size_t)T * 4, "rtx_read_codes");
}
int32_t rtx_read_perm(rtx_scene* s, uint32_t* perm, uint32_t T) {
    RTX_REQUIRE(s && s->built && T == s->T, RTX_EINVAL, "rtx_read_perm: no BVH or wrong T");
    return readback(s, s->b().perm, perm, (size_t)T * 4, "rtx_read_perm");
}
int32_t rtx_read_indices(rtx_scene* s, uint32_t* idx3, uint32_t T) {
    RTX_REQUIRE(s && T == s->T, RTX_EINVAL, "rtx_read_indices: wrong T");
    return readback(s, s->idx, idx3, (size_t)T * 12, "rtx_read_indices");
}
int32_t rtx_read_matmap(rtx_scene* s, int32_t* matmap, uint32_t T) {
    RTX_REQUIRE(s && T == s->T, RTX_EINVAL, "rtx_read_matmap: wrong T");
    return readback(s, s->matmap, matmap, (size_t)T * 4, "rtx_read_matmap");
}
int32_t rtx_read_orig_ids(rtx_scene* s, uint32_t* orig, uint32_t T) {
    RTX_REQUIRE(s && T == s->T, RTX_EINVAL, "rtx_read_orig_ids: wrong T");
    return readback(s, s->orig, orig, (size_t)T * 4, "rtx_read_orig_ids");
}
int32_t rtx_read_vertices(rtx_scene* s, float* pos3, float* nrm3, uint32_t V) {
    RTX_REQUIRE(s && V == s->V, RTX_EINVAL, "rtx_read_vertices: wrong vertex count");
    std::vector<float> tmp((size_t)V * 4);
    for (int which = 0; which < 2; ++which) {
        float* dst = which ? nrm3 : pos3;
        if (!dst) continue;
        int32_t e = readback(s, which ? s->vnrm : s->pos, tmp.data(), tmp.size() * 4, "rtx_read_vertices");
        if (e) return e;
        for (size_t i = 0; i < V; ++i)
            for (int k = 0; k < 3; ++k) dst[3 * i + k] = tmp[4 * i + k];
    }
    return RTX_OK;
}

int32_t rtx_read_global_bbox(rtx_scene* s, float* lo3, float* hi3) {
    RTX_REQUIRE(s && s->built && lo3 && hi3, RTX_EINVAL, "rtx_read_global_bbox: no BVH");
    float g[6];
    int32_t e = readback(s, s->b().gbox, g, sizeof g, "rtx_read_global_bbox");
    if (e) return e;
    memcpy(lo3, g, 12);
    memcpy(hi3, g + 3, 12);
    return RTX_OK;
}

// ------------------------------------------------------------ host math --
int32_t rtx_camera_set(rtx_camera* cam, const float pos[3], const float dir[3], const float up[3], float fov,
                       float aspect, int32_t right_handed) {
    RTX_REQUIRE(cam && pos && dir && up, RTX_EINVAL, "rtx_camera_set: null argument");
    v3 d = mk3(dir[0], dir[1], dir[2]);
    float dn = std::sqrt(dot(d, d));
    d = d * (1.0f / dn);
    v3 U = mk3(up[0], up[1], up[2]);
    v3 right = right_handed ? cross(d, U) : cross(U, d);
    v3 upv = right_handed ? cross(right, d) : cross(d, right);
    float plane_h = (float)std::tan((double)(fov * 0.5f));
    float rn = std::sqrt(dot(right, right));
    float un = std::sqrt(dot(upv, upv));
    right = right * ((plane_h * aspect) / rn);
    upv = upv * (plane_h / un);
    for (int k = 0; k < 3; ++k) cam->pos[k] = pos[k];
    cam->dir[0] = d.x; cam->dir[1] = d.y; cam->dir[2] = d.z;
    cam->right[0] = right.x; cam->right[1] = right.y; cam->right[2] = right.z;
    cam->up[0] = upv.x; cam->up[1] = upv.y; cam->up[2] = upv.z;
    return RTX_OK;
}
