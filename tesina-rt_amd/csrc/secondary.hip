// Whitted secondary rays: one fused mark + scan + generate pass.
//
// Reference: SecondaryRayGenerator::gen_scan (src/rt/secondary-ray-generator.cpp:248-340)
// with mark_secondary_rays / generate_secondary_rays
// (src/kernel/secondary-ray-generator.cl:43-181) and gpu_scan_uint
// (src/gpu/scan.cpp:17-157).  The reference marks, scans in several
// launches, reads the total back to the host and generates; here each
// 256-ray tile computes its children, takes its output offset from a
// decoupled look-back over the tiles before it (tiles are taken in order),
// and writes.  Output order equals the reference's: parents in order,
// refraction before reflection.  The count stays on the device.
#include "devmath.hpp"
#include "lookback.hpp"
#include "rtx_impl.hpp"
#include "wavefront.hpp"

namespace {

constexpr int SB = 256;

__global__ void k_sec_init(uint32_t* status, uint32_t nwords, uint32_t* in_cnt, uint32_t* out_cnt) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += gridDim.x * blockDim.x) status[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        in_cnt[3] = 0;             // tile ticket
        out_cnt[1] = 0;            // scan total
        out_cnt[2] = 0xffffffffu;  // first dropped offset
    }
}

// Children of parent k (mark_secondary_rays / generate_secondary_rays,
// src/kernel/secondary-ray-generator.cl:43-181): bit 0 a refracted ray
// (refractive material and k > 0), bit 1 a reflected ray (reflective
// material), with the Schlick R of the refraction (Q22).
struct Kids {
    uint32_t c;
    float R, reflectiveness;
    v3 refr, refl;
};

__device__ __forceinline__ Kids sec_children(uint64_t k, uint64_t n, const float4* __restrict__ dir_in,
                                             const uint32_t* __restrict__ flags, const int32_t* __restrict__ hid,
                                             const float4* __restrict__ hn, const rtx_material* __restrict__ mats,
                                             const int32_t* __restrict__ matmap) {
    Kids q{0u, 1.f, 0.f, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
    if (k < n && (flags[k] & 1u)) {
        const rtx_material m = mats[matmap[hid[k]]];
        q.reflectiveness = m.reflectiveness;
        v3 d = -normalize(xyz(dir_in[k]));
        v3 nn = normalize(xyz(hn[k]));
        if (m.refractive_index > 0.f) {
            float etai, etat;
            if (flags[k] & 2u) {
                etat = 1.f;
                etai = m.refractive_index;
            } else {
                etai = 1.f;
                etat = m.refractive_index;
            }
            float cosi = dot(d, nn);
            float rel = etai / etat;
            float wv = rel * cosi;
            float kk = 1.f + (wv - rel) * (wv + rel);
            if (kk > 0.f) {
                float R0 = (etat - 1.f) / (etat + 1.f);
                R0 *= R0;
                float x = 1.f - cosi;
                float x2 = x * x;
                float cos5 = x2 * x2 * x;
                q.R = R0 + (1.f - R0) * cos5;
                kk = __builtin_sqrtf(kk);
                q.refr = normalize((wv - kk) * nn - etai * d);
                q.c |= 1u;
            }
        }
        if (m.reflectiveness > 0.f) {
            v3 dd = -d;
            q.refl = normalize(dd - 2.f * nn * dot(dd, nn));
            q.c |= 2u;
        }
    }
    return q;
}

__device__ __forceinline__ uint32_t kid_count(const Kids& q) { return (q.c & 1u) + (q.c >> 1); }

// The children of parent k at output offset nid: refraction, then reflection.
__device__ __forceinline__ void sec_write(const Kids& q, uint32_t nid, uint64_t k, const float4* __restrict__ hp,
                                          const int32_t* __restrict__ pix_in, const float* __restrict__ con_in,
                                          float4* __restrict__ ori_out, float4* __restrict__ dir_out,
                                          int32_t* __restrict__ pix_out, float* __restrict__ con_out) {
    float4 hpk = hp[k];
    const int32_t px = pix_in[k];
    const float ck = con_in[k];
    if (q.c & 1u) {
        ori_out[nid] = make_float4(hpk.x, hpk.y, hpk.z, 0.0001f);
        dir_out[nid] = make_float4(q.refr.x, q.refr.y, q.refr.z, 1e37f);
        pix_out[nid] = px;
        con_out[nid] = ck * q.reflectiveness * (1.f - q.R);
        ++nid;
    }
    if (q.c & 2u) {
        ori_out[nid] = make_float4(hpk.x, hpk.y, hpk.z, 0.0001f);
        dir_out[nid] = make_float4(q.refl.x, q.refl.y, q.refl.z, 1e37f);
        pix_out[nid] = px;
        con_out[nid] = ck * q.reflectiveness * q.R;
    }
}

// One tile of SB * IPT input rays (tiles taken in order), thread t holding
// rays IPT t .. IPT t + IPT - 1: children per parent, the tile's output
// offset by look-back, then the children written, parents in order.  A
// parent whose children would reach the capacity (nid + 2 >= cap,
// src/kernel/secondary-ray-generator.cl:61-63) writes none and records its
// offset.  Every thread of the block calls it (two block barriers inside);
// s_w and s_prefix are the block's shared scratch.
template <int IPT>
__device__ __forceinline__ void sec_tile(
    uint32_t tile, uint32_t ntiles, uint64_t n, const float4* __restrict__ dir_in,
    const int32_t* __restrict__ pix_in, const float* __restrict__ con_in, const uint32_t* __restrict__ flags,
    const int32_t* __restrict__ hid, const float4* __restrict__ hn, const float4* __restrict__ hp,
    const rtx_material* __restrict__ mats, const int32_t* __restrict__ matmap, float4* __restrict__ ori_out,
    float4* __restrict__ dir_out, int32_t* __restrict__ pix_out, float* __restrict__ con_out, uint32_t* out_cnt,
    uint32_t* status, uint32_t cap, uint32_t* s_w, uint32_t* s_prefix_p) {
    uint32_t& s_prefix = *s_prefix_p;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint64_t k0 = ((uint64_t)tile * SB + tid) * IPT;
    Kids q[IPT];
    uint32_t cnt = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        q[j] = sec_children(k0 + j, n, dir_in, flags, hid, hn, mats, matmap);
        cnt += kid_count(q[j]);
    }
    // inclusive scan of the per-thread counts over the wave
    uint32_t v = cnt;
#pragma unroll
    for (int off = 1; off < RTX_WAVE; off <<= 1) {
        const uint32_t u = __shfl_up(v, off, RTX_WAVE);
        if (lane >= (uint32_t)off) v += u;
    }
    const uint32_t excl = v - cnt;
    if (lane == RTX_WAVE - 1) s_w[w] = v;
    __syncthreads();
    uint32_t wofs = 0, agg = 0;
#pragma unroll
    for (int i = 0; i < SB / RTX_WAVE; ++i) {
        wofs += (i < (int)w) ? s_w[i] : 0u;
        agg += s_w[i];
    }
    if (w == 0) {
        const uint32_t prefix = lookback_wave(status, tile, agg, 1u);
        if (lane == 0) {
            s_prefix = prefix;
            if (tile == ntiles - 1) out_cnt[1] = prefix + agg;
        }
    }
    __syncthreads();
    if (!cnt) return;
    uint32_t nid = s_prefix + wofs + excl;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const uint32_t cj = kid_count(q[j]);
        if (!cj) continue;
        if ((uint64_t)nid + 2u >= (uint64_t)cap)
            atomicMin(&out_cnt[2], nid);
        else
            sec_write(q[j], nid, k0 + j, hp, pix_in, con_in, ori_out, dir_out, pix_out, con_out);
        nid += cj;
    }
}

__global__ __launch_bounds__(SB) void k_secondary(
    const float4* __restrict__ dir_in, const int32_t* __restrict__ pix_in, const float* __restrict__ con_in,
    const uint32_t* __restrict__ flags, const int32_t* __restrict__ hid, const float4* __restrict__ hn,
    const float4* __restrict__ hp, const rtx_material* __restrict__ mats, const int32_t* __restrict__ matmap,
    uint32_t* in_cnt, uint64_t hcount, float4* __restrict__ ori_out, float4* __restrict__ dir_out,
    int32_t* __restrict__ pix_out, float* __restrict__ con_out, uint32_t* out_cnt, uint32_t* status, uint32_t cap) {
    __shared__ uint32_t s_tile, s_prefix;
    __shared__ uint32_t s_w[SB / RTX_WAVE];
    const uint64_t n = hcount == RTX_DEVICE_COUNT ? (uint64_t)in_cnt[0] : hcount;
    const uint32_t ntiles = (uint32_t)((n + SB - 1) / SB);
    // blocks past the last tile leave before taking a ticket, so the first
    // ntiles blocks to arrive take exactly the tickets [0, ntiles)
    if (blockIdx.x >= ntiles) return;  // (n == 0: out_cnt[1] stays 0 from k_sec_init)
    if (threadIdx.x == 0) s_tile = atomicAdd(&in_cnt[3], 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= ntiles) return;
    sec_tile<1>(tile, ntiles, n, dir_in, pix_in, con_in, flags, hid, hn, hp, mats, matmap, ori_out, dir_out, pix_out,
                con_out, out_cnt, status, cap, s_w, &s_prefix);
}

// Secondary rays of every frame that generates this step: persistent blocks
// take tickets over the concatenation of the frames' tiles, so each frame's
// tiles are taken in order and a tile's look-back only waits on tiles
// already taken (by running blocks).  Look-back words were cleared by this
// step's trace kernel; scan totals and drop offsets by its first kernel.
__global__ __launch_bounds__(SB) void k_wf_secondary(const WfSeg* __restrict__ segs, uint32_t nseg,
                                                     uint32_t* __restrict__ ticket) {
    __shared__ uint32_t s_end[RTX_WF_MAX_SEGS];
    __shared__ uint32_t s_tile, s_prefix;
    __shared__ uint32_t s_w[SB / RTX_WAVE];
    const uint32_t total = wf_chunks<WF_GEN_TILES>(segs, nseg, 1u, s_end);
    // only the first `total` blocks take tickets (a ticket is a same-address
    // atomic; idle blocks would only queue behind the working ones)
    if (blockIdx.x >= total) return;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t g = s_tile;
        if (g >= total) break;
        const uint32_t sg = wf_find(s_end, nseg, g);
        const WfSeg& S = segs[sg];
        const uint64_t n = S.count[0];
        sec_tile<WF_GEN_TILE / SB>(g - wf_first(s_end, sg), (uint32_t)((n + WF_GEN_TILE - 1) / WF_GEN_TILE), n,
                                   S.dir4, S.pixel, S.contrib, S.flags, S.id, S.n4, S.p4, S.mats, S.matmap, S.o_ori4,
                                   S.o_dir4, S.o_pixel, S.o_contrib, S.o_count, S.status, S.o_cap, s_w, &s_prefix);
    }
}

// Output count = rays actually written (a prefix): min(scan total, first
// dropped offset).  Optionally accumulates the frame's secondary-ray count.
__global__ void k_sec_finalize(uint32_t* out_cnt, unsigned long long* acc) {
    uint32_t total = out_cnt[1], drop = out_cnt[2];
    uint32_t n = total < drop ? total : drop;
    out_cnt[0] = n;
    if (acc) *acc += n;
}

}  // namespace

int32_t rtx_raygen_secondary_impl(rtx_scene* s, rtx_rays* in, rtx_hits* hits, rtx_rays* out, uint64_t count_in,
                                  uint64_t* dev_total_acc, hipStream_t st) {
    RTX_REQUIRE(s && in && hits && out, RTX_EINVAL, "rtx_raygen_secondary: null argument");
    RTX_REQUIRE(in != out, RTX_EINVAL, "rtx_raygen_secondary: input and output bundles must differ");
    RTX_REQUIRE(s->mats.p && s->matmap.p, RTX_ESTATE, "rtx_raygen_secondary: scene has no materials");
    RTX_REQUIRE(count_in == RTX_DEVICE_COUNT || (count_in <= in->cap && count_in <= hits->cap), RTX_EINVAL,
                "rtx_raygen_secondary: count exceeds bundle capacity");
    uint32_t max_tiles = rtx_div_up(in->cap, SB) + 1;
    {
        ProfScope ps(s->ctx, st, "sec_init");
        uint32_t g = rtx_div_up(max_tiles, 256);
        hipLaunchKernelGGL(k_sec_init, dim3(g > 256 ? 256 : g), dim3(256), 0, st, in->scan_status.as<uint32_t>(),
                           max_tiles, in->count.as<uint32_t>(), out->count.as<uint32_t>());
    }
    {
        ProfScope ps(s->ctx, st, "sec_generate");
        uint64_t n = count_in == RTX_DEVICE_COUNT ? in->cap : count_in;
        uint32_t grid = rtx_div_up(n ? n : 1, SB);
        hipLaunchKernelGGL(k_secondary, dim3(grid), dim3(SB), 0, st, in->dir4.as<float4>(), in->pixel.as<int32_t>(),
                           in->contrib.as<float>(), hits->flags.as<uint32_t>(), hits->id.as<int32_t>(),
                           hits->n4.as<float4>(), hits->p4.as<float4>(), s->mats.as<rtx_material>(),
                           scene_view(s).matmap, in->count.as<uint32_t>(), count_in, out->ori4.as<float4>(),
                           out->dir4.as<float4>(), out->pixel.as<int32_t>(), out->contrib.as<float>(),
                           out->count.as<uint32_t>(), in->scan_status.as<uint32_t>(), (uint32_t)out->cap);
    }
    {
        ProfScope ps(s->ctx, st, "sec_finalize");
        hipLaunchKernelGGL(k_sec_finalize, dim3(1), dim3(1), 0, st, out->count.as<uint32_t>(),
                           (unsigned long long*)dev_total_acc);
    }
    RTX_CHECK_HIP(hipGetLastError());
    return RTX_OK;
}

int32_t rtx_raygen_secondary(rtx_scene* s, rtx_rays* rays_in, rtx_hits* hits, rtx_rays* rays_out, uint64_t count_in,
                             uint64_t* count_out, void* stream) {
    RTX_REQUIRE(s, RTX_EINVAL, "rtx_raygen_secondary: null scene");
    hipStream_t st = s->ctx->pick(stream);
    int32_t e = rtx_raygen_secondary_impl(s, rays_in, hits, rays_out, count_in, nullptr, st);
    if (e) return e;
    if (count_out) {
        uint32_t c = 0;
        RTX_CHECK_HIP(hipMemcpyAsync(&c, rays_out->count.p, 4, hipMemcpyDeviceToHost, st));
        RTX_CHECK_HIP(hipStreamSynchronize(st));
        *count_out = c;
    }
    return RTX_OK;
}

void rtx_wf_secondary(const WfSeg* segs, uint32_t nseg, uint32_t* ctr, int cus, hipStream_t st) {
    hipLaunchKernelGGL(k_wf_secondary, dim3((uint32_t)cus * 8), dim3(SB), 0, st, segs, nseg, ctr + WF_Q_SEC);
}
